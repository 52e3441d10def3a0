"""Restatement of MesonCorrelator (lib/Measurements/FermionicM/
meson_correlator.hpp) and the gamma-matrix tables it uses
(lib/Tools/gammaMatrices.cpp) in numpy.

TEST INFRASTRUCTURE ONLY.  Importable from tests/ (the checker of the
library's ii_meson_correlators); the product never imports it.

A propagator q is 12 full fields back to back, column c1 + 3*s1 (source
colour, spin), each in Format_F: q[col][24*site + 2*(c2 + 3*s2) + re/im].
"""
import numpy as np

# gammaMatrices.cpp:13-170: (spn index per row, factor (re, im) per row);
# Gamma(s) maps spin s to spn[s] with factor fac[s]
GAMMAS = {
    "Unit": ([0, 1, 2, 3], [(1, 0), (1, 0), (1, 0), (1, 0)]),
    "Gamma1": ([3, 2, 1, 0], [(0, -1), (0, -1), (0, 1), (0, 1)]),
    "Gamma2": ([3, 2, 1, 0], [(-1, 0), (1, 0), (1, 0), (-1, 0)]),
    "Gamma3": ([2, 3, 0, 1], [(0, -1), (0, 1), (0, 1), (0, -1)]),
    "Gamma4": ([0, 1, 2, 3], [(1, 0), (1, 0), (-1, 0), (-1, 0)]),
    "Gamma5": ([2, 3, 0, 1], [(1, 0), (1, 0), (1, 0), (1, 0)]),
    "Gamma1_5": ([1, 0, 3, 2], [(0, -1), (0, -1), (0, 1), (0, 1)]),
    "Gamma2_5": ([1, 0, 3, 2], [(-1, 0), (1, 0), (1, 0), (-1, 0)]),
    "Gamma3_5": ([0, 1, 2, 3], [(0, -1), (0, 1), (0, 1), (0, -1)]),
    "Gamma4_5": ([2, 3, 0, 1], [(1, 0), (1, 0), (-1, 0), (-1, 0)]),
    "Sigma12_5": ([2, 3, 0, 1], [(0, 1), (0, -1), (0, 1), (0, -1)]),
    "Sigma13_5": ([3, 2, 1, 0], [(1, 0), (-1, 0), (1, 0), (-1, 0)]),
    # the reference's table as it stands (rows 2 and 3 are not swapped)
    "Sigma14_5": ([1, 0, 2, 3], [(1, 0), (1, 0), (1, 0), (1, 0)]),
    "Sigma23_5": ([3, 2, 1, 0], [(0, 1), (0, 1), (0, 1), (0, 1)]),
    "Sigma24_5": ([1, 0, 3, 2], [(1, 0), (-1, 0), (1, 0), (-1, 0)]),
    "Sigma34_5": ([0, 1, 2, 3], [(0, 1), (0, -1), (0, 1), (0, -1)]),
}

# MesonType (meson_correlator.hpp:20-24) -> (G1, G2) (:37-120)
TYPES = [
    ("Pion", "Unit", "Unit"), ("Scalar", "Gamma5", "Gamma5"),
    ("Vector1", "Gamma1_5", "Gamma1_5"), ("Vector2", "Gamma2_5", "Gamma2_5"),
    ("Vector3", "Gamma3_5", "Gamma3_5"), ("Vector4", "Gamma4_5", "Gamma4_5"),
    ("AVector1", "Gamma1", "Gamma1"), ("AVector2", "Gamma2", "Gamma2"),
    ("AVector3", "Gamma3", "Gamma3"), ("AVector4", "Gamma4", "Gamma4"),
    ("Tensor12", "Sigma12_5", "Sigma12_5"), ("Tensor13", "Sigma13_5", "Sigma13_5"),
    ("Tensor14", "Sigma14_5", "Sigma14_5"), ("Tensor23", "Sigma23_5", "Sigma23_5"),
    ("Tensor24", "Sigma24_5", "Sigma24_5"), ("Tensor34", "Sigma34_5", "Sigma34_5"),
    ("AV4_PS", "Gamma4", "Unit"), ("PS_AV4", "Unit", "Gamma4"),
]


def _prop(q, V):
    a = np.asarray(q, dtype=np.float64).reshape(12, V, 4, 3, 2)
    return a[..., 0] + 1j * a[..., 1]          # [col][site][spin][colour]


def _site_values(q1, q2, V, type_idx):
    """Per site: sum_{s4,s2} gf(s4,s2) sum_{c1} sum_{c2} conj(q1[c1+3 s1][s2][c2])
    q2[c1+3 s4][s3][c2] with s1 = G2(s4).spn, s3 = G1(s2).spn and
    gf = Re(G2(s4).fac G1(s2).fac) (meson_correlator.hpp:156-176); complex
    (the real part is calculate's temp_corr, the imaginary part
    calculate_mom's im_prod)."""
    _, n1, n2 = TYPES[type_idx]
    i1, f1 = GAMMAS[n1]
    i2, f2 = GAMMAS[n2]
    a = _prop(q1, V).reshape(4, 3, V, 4, 3)   # [s1][c1][site][s][c2]
    b = _prop(q2, V).reshape(4, 3, V, 4, 3)
    out = np.zeros(V, dtype=np.complex128)
    for s4 in range(4):
        s1 = i2[s4]
        for s2 in range(4):
            s3 = i1[s2]
            gf = f2[s4][0] * f1[s2][0] - f2[s4][1] * f1[s2][1]
            if gf == 0.0:
                continue
            out += gf * np.sum(np.conj(a[s1, :, :, s2, :]) * b[s4, :, :, s3, :], axis=(0, 2))
    return out


def coords(dims):
    V = int(np.prod(dims))
    s = np.arange(V)
    x = s % dims[0]
    y = (s // dims[0]) % dims[1]
    z = (s // (dims[0] * dims[1])) % dims[2]
    t = s // (dims[0] * dims[1] * dims[2])
    return np.stack([x, y, z, t], axis=1)


def calculate(q1, q2, dims, type_idx, direction=3):
    """MesonCorrelator::calculate (meson_correlator.hpp:138-190): sum of the
    real part over each slice along `direction`."""
    V = int(np.prod(dims))
    v = _site_values(q1, q2, V, type_idx).real
    c = coords(dims)[:, direction]
    return np.bincount(c, weights=v, minlength=dims[direction])


def calculate_mom(q1, q2, dims, type_idx, p):
    """calculate_mom (:192-247): time slices of sum z e^{i p.x}, (re, im)
    interleaved."""
    V = int(np.prod(dims))
    v = _site_values(q1, q2, V, type_idx)
    c = coords(dims)
    ph = c[:, 0] * p[0] + c[:, 1] * p[1] + c[:, 2] * p[2]
    w = v * (np.cos(ph) + 1j * np.sin(ph))
    re = np.bincount(c[:, 3], weights=w.real, minlength=dims[3])
    im = np.bincount(c[:, 3], weights=w.imag, minlength=dims[3])
    out = np.empty(2 * dims[3])
    out[0::2], out[1::2] = re, im
    return out


def calculate_PSmom(q1, q2, dims, type_idx, p):
    """calculate_PSmom (:249-296): time slices of Re z cos(px x) cos(py y)
    cos(pz z)."""
    V = int(np.prod(dims))
    v = _site_values(q1, q2, V, type_idx).real
    c = coords(dims)
    w = v * np.cos(p[0] * c[:, 0]) * np.cos(p[1] * c[:, 1]) * np.cos(p[2] * c[:, 2])
    return np.bincount(c[:, 3], weights=w, minlength=dims[3])


def gamma_matrix(name):
    """The 4x4 matrix of a table entry: row s has fac[s] in column spn[s]."""
    idx, fac = GAMMAS[name]
    m = np.zeros((4, 4), dtype=np.complex128)
    for s in range(4):
        m[s, idx[s]] = complex(*fac[s])
    return m

