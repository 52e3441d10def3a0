"""Meson correlators on the device (ii_meson_correlators) against the
reference's own MesonCorrelator outputs (golden fixture) and, on a real
point-source propagator solved on the device at 8^3 x 16, against the
oracle's contraction of the downloaded columns.  Tolerance 1e-12 relative
to the largest entry (the device sums in a different order and with
fused multiply-adds)."""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import meson
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b))


@pytest.fixture(scope="module")
def mz(golden):
    z = golden("meson_2x4x2x6")
    dims = tuple(int(d) for d in z["dims"])
    ctx = ii.Context(dims)
    V = int(np.prod(dims))
    q1 = [z["q1"][k * 24 * V:(k + 1) * 24 * V] for k in range(12)]
    q2 = [z["q2"][k * 24 * V:(k + 1) * 24 * V] for k in range(12)]
    f1, f2 = ii.DeviceField(ctx, "full", 12), ii.DeviceField(ctx, "full", 12)
    for k in range(12):
        f1.upload(k, q1[k])
        f2.upload(k, q2[k])
    yield dims, z, ctx, f1, f2
    ctx.close()


def test_all_types_and_directions_vs_reference(mz):
    dims, z, ctx, f1, f2 = mz
    calc = np.concatenate([ii.meson_correlators(ctx, f1, f2, 0, d) for d in range(4)], axis=1)
    for t in range(18):
        assert rel(calc[t], z["calc"][t]) < 1e-12, t
    mom = ii.meson_correlators(ctx, f1, f2, 1, 3, z["mom"])
    ps = ii.meson_correlators(ctx, f1, f2, 2, 3, z["mom"])
    for t in range(18):
        assert rel(mom[t], z["mom_out"][t]) < 1e-12, t
        assert rel(ps[t], z["psmom_out"][t]) < 1e-12, t


def test_same_propagator_and_host_columns(mz):
    dims, z, ctx, f1, f2 = mz
    V = int(np.prod(dims))
    got = ii.meson_correlators(ctx, f1, f1, 0, 3)
    for t in range(18):
        want = meson.calculate(z["q1"], z["q1"], dims, t, 3)
        assert rel(got[t], want) < 1e-12, t
    # MesonCorrelator with a host prop_t (list of 12 fields)
    cols = [z["q1"][k * 24 * V:(k + 1) * 24 * V] for k in range(12)]
    mc = ii.MesonCorrelator(ctx, "Vector2")
    assert rel(mc.calculate(cols, cols), meson.calculate(z["q1"], z["q1"], dims, 3, 3)) < 1e-12


def test_point_source_propagator_8c16():
    dims = (8, 8, 8, 16)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(6.0, 31, 20)
    ctx.apply_apbc()
    src = ii.DeviceField(ctx, "full", 12)
    sol = ii.DeviceField(ctx, "full", 12)
    src.point_sources((0, 0, 0, 0))
    ii.PropagatorSolver(ctx, 0.13, tol=1e-12, precision="double", batched=True).solve(src, sol)
    got = ii.meson_correlators(ctx, sol, sol, 0, 3)
    V = ctx.V
    q = np.concatenate([sol.download(k) for k in range(12)])
    for t in (0, 1, 2, 5, 9, 16, 17):
        want = meson.calculate(q, q, dims, t, 3)
        assert rel(got[t], want) < 1e-12, t
    pion = got[0]
    assert np.all(pion > 0)
    # a bound state: decays away from the source towards the middle
    assert pion[4] < pion[1] and pion[8] < pion[4]
    assert pion[12] < pion[15]
    ctx.close()
