"""Independent mathematics used to pin the oracle and the device path.

Nothing here restates the oracle's code: the Dirac basis is taken from the
reference's gamma-matrix definitions (lib/Dirac_ops/wilsonLikeUtils.cpp:9-49)
as 4x4 matrices and checked against the Clifford algebra; the Wilson
operator is then built densely from the textbook formula
  D = 1 - kappa sum_mu [(1 - g_mu) U_mu(x) d_{x+mu,y} + (1 + g_mu) U_mu(x-mu)^dag d_{x-mu,y}]
with numpy, independently of the projection tables.
"""
import numpy as np

# Dirac basis of the reference (wilsonLikeUtils.cpp:9-49), acting on spin
GX = np.array([[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]])
GY = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex)
GZ = np.array([[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]])
GT = np.diag([1, 1, -1, -1]).astype(complex)
G5 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)
GAMMAS = [GX, GY, GZ, GT]


def to_cplx_spinor(f, nsites):
    a = np.asarray(f).reshape(nsites, 4, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def from_cplx_spinor(c):
    out = np.empty(c.shape + (2,))
    out[..., 0] = c.real
    out[..., 1] = c.imag
    return out.ravel()


def to_cplx_gauge(u, V):
    a = np.asarray(u).reshape(4, V, 3, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def from_cplx_gauge(c):
    out = np.empty(c.shape + (2,))
    out[..., 0] = c.real
    out[..., 1] = c.imag
    return out.ravel()


def coords(dims):
    Lx, Ly, Lz, Lt = dims
    s = np.arange(Lx * Ly * Lz * Lt)
    return np.stack([s % Lx, (s // Lx) % Ly, (s // (Lx * Ly)) % Lz, s // (Lx * Ly * Lz)], axis=1)


def shift_index(dims, mu, step):
    """Lexicographic index of x + step*mu (periodic) for every site x."""
    c = coords(dims).copy()
    c[:, mu] = (c[:, mu] + step) % dims[mu]
    Lx, Ly, Lz, _ = dims
    return c[:, 0] + Lx * (c[:, 1] + Ly * (c[:, 2] + Lz * c[:, 3]))


def random_su3(n, rng):
    z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    q = q / (det ** (1.0 / 3.0))[:, None, None]
    return q


def dense_wilson(dims, u, kappa):
    """Dense full Wilson matrix from the textbook formula, index
    (site, spin, colour) -> 12*site + 3*spin + colour."""
    V = int(np.prod(dims))
    U = to_cplx_gauge(u, V)
    D = np.eye(12 * V, dtype=complex)
    one = np.eye(4)
    for mu in range(4):
        fw = shift_index(dims, mu, +1)
        bw = shift_index(dims, mu, -1)
        Pm = one - GAMMAS[mu]
        Pp = one + GAMMAS[mu]
        for x in range(V):
            D[12 * x:12 * x + 12, 12 * fw[x]:12 * fw[x] + 12] -= kappa * np.kron(Pm, U[mu, x])
            D[12 * x:12 * x + 12, 12 * bw[x]:12 * bw[x] + 12] -= kappa * np.kron(
                Pp, U[mu, bw[x]].conj().T)
    return D


def spinor_vec(f, V):
    """Format_F real array -> complex vector in (site, spin, colour) order."""
    return to_cplx_spinor(f, V).reshape(-1)


def vec_spinor(v, V):
    return from_cplx_spinor(v.reshape(V, 4, 3))
