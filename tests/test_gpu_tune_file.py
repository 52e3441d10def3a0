"""The autotuner's choices saved to a text file and loaded into a fresh
context on the same lattice (ii_tune_save / ii_tune_load): the loaded
context writes the same file back, applies the network and the multi-rhs
hops without re-timing them, with results bitwise equal to the tuned
context's; a file for another lattice is refused."""
import numpy as np
import pytest

import iroiro_amd as ii
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_tune_file_round_trip(tmp_path):
    dims = (8, 8, 8, 16)
    lat = orc.Lattice(dims)
    u = lat.apbc(lat.gauge_expH(0.4, 91))
    b = np.random.default_rng(5).normal(size=(12, lat.half_size))
    outs = []
    files = [tmp_path / "a.tune", tmp_path / "b.tune"]
    for i in range(2):
        ctx = ii.Context(dims)
        try:
            ctx.upload_gauge(u)
            net = ii.LearnedPreconditioner(ctx)
            net.init_neumann(0.13, noise=0.3, seed=5)
            if i == 0:
                print(ctx.autotune())
                net.tune(12)
                ctx.tune_save(files[0])
            else:
                ctx.tune_load(files[0])
                ctx.tune_save(files[1])
            outs.append(net.apply_batched(b))
        finally:
            ctx.close()
    text = files[0].read_text()
    print(text)
    assert text == files[1].read_text()
    assert "hop " in text and "mrhs " in text and "net 12 " in text
    assert np.array_equal(outs[0], outs[1])
    other = ii.Context((8, 8, 8, 8))
    try:
        with pytest.raises(ii.IIError):
            other.tune_load(files[0])
    finally:
        other.close()
