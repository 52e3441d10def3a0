"""A lattice beyond one 32-bit buffer descriptor per gauge field.

48^3 x 72 has V/2 = 3,981,312 half sites: its FP64 gauge field (4.6 GB)
spans more than the 4 GiB a buffer descriptor's 32-bit offsets reach, so
the kernels address it in (parity, direction) blocks whose bases are formed
in 64 bits.  The checks run the operators that read the gauge field through
those blocks (FP64 hop and Dhat, the fused 12-rhs FP32 and FP16 hops with
reconstruct-12 links) against the oracle on the same field, the plaquette,
and the size limits of the network paths: the single-rhs layers refuse the
lattice (their 8-channel activation array would exceed one descriptor), the
batched network takes it.
"""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle

pytestmark = pytest.mark.gpu

DIMS = (48, 48, 48, 72)
KAPPA = 0.1355


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def big():
    lat = oracle.Lattice(DIMS)
    assert lat.Vh * 72 * 16 > 2 ** 32  # the FP64 gauge field exceeds one descriptor
    ctx = ii.Context(DIMS)
    ctx.unit_gauge()
    ctx.heatbath(6.0, 2026, 1)
    ctx.apply_apbc()
    u = ctx.download_gauge()
    D = ii.Dirac_Wilson_EvenOdd(0.5 / KAPPA - 4.0, u, ctx)
    yield lat, u, ctx, D
    ctx.close()


def test_fp64_operators_and_plaquette(big):
    lat, u, ctx, D = big
    assert abs(ctx.plaquette() - lat.plaquette(u)) < 1e-12
    rng = np.random.default_rng(3)
    fo = rng.normal(size=lat.half_size)
    fe = rng.normal(size=lat.half_size)
    k = D.getKappa()
    assert rel(D.mult_eo(fo), lat.hop_eo(u, fo, 1, 0, k)) < 1e-13
    assert rel(D.mult(fe), lat.dirac_eo(u, fe, 0, k)) < 1e-13


@pytest.mark.parametrize("precision,tol", [("single", 6 * 2.0 ** -24), ("half", 6 * 2.0 ** -11)])
def test_fused_12rhs_hop(big, precision, tol):
    lat, u, ctx, D = big
    f = np.random.default_rng(5).normal(size=(12, lat.half_size))
    got = D.mult_eo_prec(f.ravel(), precision, 0, False).reshape(12, -1)
    for k in (0, 11):
        assert rel(got[k], lat.hop_eo(u, f[k], 1, 0, KAPPA)) < tol, (precision, k)


def test_network_paths_size_limits(big):
    lat, u, ctx, D = big
    net = ii.LearnedPreconditioner(ctx)
    net.init_neumann(0.13, noise=0.3, seed=5)
    b = np.random.default_rng(7).normal(size=lat.half_size)
    with pytest.raises(ii.IIError):
        net.apply(b)
    # the batched network: every rounding step scales with a power of two, so
    # z(2 b) = 2 z(b) exactly, and the rhs of a batch do not mix
    z = net.apply_batched(np.stack([b, 2.0 * b]))
    assert np.all(np.isfinite(z))
    assert np.array_equal(z[1], 2.0 * z[0])
    z1 = net.apply_batched(b[None, :])
    assert np.array_equal(z1[0], z[0])
    assert 0.1 < np.linalg.norm(z[0]) / np.linalg.norm(b) < 10.0
