"""Gauge-configuration files through the device (ii_gauge_read /
ii_gauge_write) against the reference's own files (golden fixture), the
oracle's restatement of the formats, and at 16^3 x 32 by write-read round
trips.  Byte and index work: bit-exact; the plaquette / link trace the
device computes for a NERSC header agree with the reference's to 1e-14."""
import os

import numpy as np
import pytest

import iroiro_amd as ii
from oracle import gauge_io as gio
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def io(golden):
    base = golden("wilson_aniso_4x6x2x4")
    z = golden("wilson_aniso_4x6x2x4_io")
    dims = tuple(int(d) for d in base["dims"])
    ctx = ii.Context(dims)
    yield dims, base["u_apbc"], z, ctx
    ctx.close()


def test_read_reference_nersc_file(io, tmp_path):
    dims, u, z, ctx = io
    p = tmp_path / "ref.nersc"
    p.write_bytes(z["written_NERSC"].tobytes())
    info = ctx.read_gauge(str(p), "NERSC", check=True)
    assert np.array_equal(ctx.download_gauge(), u)
    kv, _ = gio.nersc_parse_header(z["written_NERSC"].tobytes())
    assert info["checksum"] == int(kv["CHECKSUM"], 16) == info["header_checksum"]
    assert abs(info["plaquette"] - float(kv["PLAQUETTE"])) < 1e-15
    assert abs(info["link_trace"] - float(kv["LINK_TRACE"])) < 1e-15
    assert info["bits"] == 64 and info["is3x2"] == 0


def test_read_binary_big_endian(io, tmp_path):
    dims, u, z, ctx = io
    p = tmp_path / "in.be"
    p.write_bytes(z["binary_input_be"].tobytes())
    ctx.read_gauge(str(p), "Binary")
    assert np.array_equal(ctx.download_gauge(), z["binary_read"])


def test_write_files_match_reference_and_oracle(io, tmp_path):
    dims, u, z, ctx = io
    V = int(np.prod(dims))
    ctx.upload_gauge(u)
    pb, pn, pi = tmp_path / "w.bin", tmp_path / "w.nersc", tmp_path / "w.ildg"
    ctx.write_gauge(str(pb), "Binary")
    assert pb.read_bytes() == z["written_Binary"].tobytes()
    info = ctx.write_gauge(str(pn), "NERSC")
    got = pn.read_bytes()
    ref = z["written_NERSC"].tobytes()
    kv, off = gio.nersc_parse_header(got)
    kv_ref, off_ref = gio.nersc_parse_header(ref)
    assert got[off:] == ref[off_ref:]
    for k in kv_ref:
        if k in ("PLAQUETTE", "LINK_TRACE"):
            assert abs(float(kv[k]) - float(kv_ref[k])) < 1e-14, k
        else:
            assert kv[k] == kv_ref[k], k
    ctx.write_gauge(str(pi), "ILDG")
    assert pi.read_bytes() == gio.ildg_write(u, V)


@pytest.mark.parametrize("is3x2,bits", [(True, 64), (False, 32), (True, 32)])
def test_read_nersc_variants_vs_oracle(io, tmp_path, is3x2, bits):
    dims, u, z, ctx = io
    lat = oracle.Lattice(dims)
    su3 = lat.gauge_expH(0.3, 91)
    blob = gio.nersc_write_variant(su3, dims, is3x2, bits, lat.plaquette(su3))
    p = tmp_path / "v.nersc"
    p.write_bytes(blob)
    want, kv, chk = gio.nersc_read(blob, dims)
    info = ctx.read_gauge(str(p), "NERSC", check=(bits == 64))
    got = ctx.download_gauge()
    # same arithmetic (float -> double, the reference's third-row expression)
    assert np.max(np.abs(got - want)) <= (1e-15 if bits == 64 else 1e-15)
    assert info["checksum"] == chk
    assert info["is3x2"] == int(is3x2) and info["bits"] == bits


def test_plaquette_check_rejects_a_wrong_header(io, tmp_path):
    dims, u, z, ctx = io
    blob = z["written_NERSC"].tobytes().replace(b"PLAQUETTE = 0.4", b"PLAQUETTE = 0.5")
    p = tmp_path / "bad.nersc"
    p.write_bytes(blob)
    with pytest.raises(ii.IIError):
        ctx.read_gauge(str(p), "NERSC", check=True)
    bad = tmp_path / "dims.nersc"
    bad.write_bytes(z["written_NERSC"].tobytes().replace(b"DIMENSION_2 = 6", b"DIMENSION_2 = 8"))
    with pytest.raises(ii.IIError):
        ctx.read_gauge(str(bad), "NERSC")


def test_round_trips_16c32(tmp_path):
    dims = (16, 16, 16, 32)
    ctx = ii.Context(dims)
    ctx.unit_gauge()
    ctx.heatbath(6.0, 99, 3)
    ctx.apply_apbc()
    u = ctx.download_gauge()
    plaq = ctx.plaquette()
    for fmt in ("NERSC", "ILDG"):
        p = tmp_path / ("rt." + fmt)
        w = ctx.write_gauge(str(p), fmt)
        ctx.unit_gauge()
        r = ctx.read_gauge(str(p), fmt, check=(fmt == "NERSC"))
        assert np.array_equal(ctx.download_gauge(), u), fmt
        assert r["plaquette"] == plaq
        if fmt == "NERSC":
            assert r["checksum"] == w["checksum"] == r["header_checksum"]
        os.remove(p)
    # the oracle reads the device-written NERSC file to the same field
    p = tmp_path / "o.nersc"
    ctx.write_gauge(str(p), "NERSC")
    got, kv, chk = gio.nersc_read(p.read_bytes(), dims)
    assert np.array_equal(got, u) and chk == int(kv["CHECKSUM"], 16)
    ctx.close()
