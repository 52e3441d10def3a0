"""ii_fopr_ddagd_mult_noalloc, the variant of Fopr_DdagD::mult (lib/Fopr/fopr.h:146-153)
that allocates nothing (its buffers live in one static device array per GPU), called
through the C ABI: equal bit for bit to ii_fopr_ddagd_mult on two calls in a row, the
second on a smaller lattice than the first, and a lattice beyond its scratch rejected."""
import ctypes

import numpy as np
import pytest

import iroiro_amd as ii
from iroiro_amd import _lib
from oracle import oracle

pytestmark = pytest.mark.gpu


def _apply(ctx, fn, kappa, x):
    out = np.empty_like(x)
    st = getattr(_lib.load(), fn)(ctx.handle, ctypes.c_double(kappa), _lib.dptr(out),
                                  _lib.dptr(x))
    return st, out


def _case(dims, seed, kappa):
    print("no-alloc case dims=%s seed=%d" % (dims, seed), flush=True)
    lat = oracle.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.3, seed))
    x = np.random.default_rng(seed).standard_normal(lat.half_size)
    ctx = ii.Context(dims)
    ctx.upload_gauge(u)
    st_a, ref = _apply(ctx, "ii_fopr_ddagd_mult", kappa, x)
    st_b, got = _apply(ctx, "ii_fopr_ddagd_mult_noalloc", kappa, x)
    assert st_a == _lib.II_OK and st_b == _lib.II_OK, (st_a, st_b)
    # same kernels, same traversal-independent site arithmetic: bit-identical
    assert np.array_equal(got, ref)
    # and it is Fopr_DdagD::mult: the oracle's operator to rounding
    want = lat.ddagd(u, x, kappa)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13
    ctx.close()


def test_noalloc_equals_ddagd_mult_twice_second_smaller():
    _case((8, 8, 8, 16), 11, 0.1355)   # V/2 = 4096
    _case((4, 4, 4, 8), 12, 0.12)      # smaller, on the same device scratch


def test_noalloc_rejects_lattice_beyond_scratch():
    ctx = ii.Context((16, 16, 16, 32))  # V/2 = 65536: needs ~63 MB of scratch
    x = np.zeros(24 * ctx.Vh)
    st, _ = _apply(ctx, "ii_fopr_ddagd_mult_noalloc", 0.13, x)
    assert st == _lib.II_ERR_BAD_ARG
    msg = _lib.load().ii_last_error(ctx.handle).decode()
    assert "scratch" in msg and "maximum" in msg, msg
    ctx.close()
