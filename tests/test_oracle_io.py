"""Pins of the oracle's gauge-file restatement (oracle/gauge_io.py).

 * Compiled reference: tests/golden/wilson_aniso_4x6x2x4_io.npz holds the
   files the reference's BinaryWriter and NERSCWriter wrote for the golden
   gauge field, and what its BinaryReader decoded from a big-endian file.
   The NERSC reader needs a configure-generated header (include/config.h,
   via IO/generic_header.hpp) and the ILDG classes need the LIME library;
   neither is in this image, so those two are pinned by the files the
   reference writes and by the format's own invariants below.
 * Mathematics: a 3x2 (two-row) SU(3) link is completed by conj(row0 x
   row1); the plaquette and link trace of the unit field are 1; a planted
   byte flip changes the checksum by the word difference.
"""
import struct

import numpy as np
import pytest

from oracle import gauge_io as gio
from oracle import oracle


@pytest.fixture(scope="module")
def io(golden):
    base = golden("wilson_aniso_4x6x2x4")
    z = golden("wilson_aniso_4x6x2x4_io")
    dims = tuple(int(d) for d in base["dims"])
    return dims, base["u_apbc"], z


def test_binary_writer_and_reader_bit_exact(io):
    dims, u, z = io
    V = int(np.prod(dims))
    assert gio.binary_write(u, V) == z["written_Binary"].tobytes()
    assert np.array_equal(gio.binary_read(z["binary_input_be"], V), z["binary_read"])
    assert np.array_equal(z["binary_read"], u)
    # the reference's Binary files are host-order: reading one back with
    # BinaryReader (big-endian) does not give the field
    assert not np.array_equal(gio.binary_read(z["written_Binary"], V), u)


def test_nersc_writer_bit_exact(io):
    dims, u, z = io
    lat = oracle.Lattice(dims)
    blob = gio.nersc_write(u, dims, lat.plaquette(u))
    ref = z["written_NERSC"].tobytes()
    kv_ref, off_ref = gio.nersc_parse_header(ref)
    kv, off = gio.nersc_parse_header(blob)
    assert blob[off:] == ref[off_ref:]                      # data, bit for bit
    assert kv == kv_ref                                     # every header line
    assert blob == ref


def test_nersc_reader_on_reference_file(io):
    dims, u, z = io
    got, kv, chk = gio.nersc_read(z["written_NERSC"], dims)
    assert np.array_equal(got, u)
    assert chk == int(kv["CHECKSUM"], 16)
    lat = oracle.Lattice(dims)
    assert abs(lat.plaquette(got) - float(kv["PLAQUETTE"])) < 1e-15
    assert abs(gio.link_trace(got, lat.V) - float(kv["LINK_TRACE"])) < 1e-15


@pytest.mark.parametrize("is3x2,bits", [(True, 64), (False, 32), (True, 32)])
def test_nersc_compressed_and_single_precision(io, is3x2, bits):
    """4D_SU3_GAUGE (two rows) and IEEE32BIG: the third row of an SU(3) link
    is conj(row0 x row1), so the two-row file reproduces the field to
    rounding; single precision to float rounding."""
    dims, u, z = io
    lat = oracle.Lattice(dims)
    su3 = lat.gauge_expH(0.3, 77)           # no APBC: det = +1 everywhere
    blob = gio.nersc_write_variant(su3, dims, is3x2, bits, lat.plaquette(su3))
    got, kv, chk = gio.nersc_read(blob, dims)
    tol = 1e-14 if bits == 64 else 3e-7
    assert np.max(np.abs(got - su3)) < tol
    assert chk == int(kv["CHECKSUM"], 16)
    # a planted sign error in the reconstruction would show: the third row
    # of the unit matrix is (0, 0, 1)
    if is3x2:
        unit = np.zeros((1, 18))
        unit[0, [0, 8, 16]] = 1.0
        assert np.array_equal(gio.reconstruct_third_row(unit[:, :12])[0], unit[0])


def test_checksum_is_the_word_sum():
    data = struct.pack("<4I", 1, 2, 0xFFFFFFFF, 7)
    assert gio.checksum(data) == (1 + 2 + 0xFFFFFFFF + 7) & 0xFFFFFFFF
    flipped = bytearray(data)
    flipped[4] ^= 0x10
    assert gio.checksum(bytes(flipped)) == (gio.checksum(data) + 0x10) & 0xFFFFFFFF


def test_ildg_lime_layout_and_round_trip(io):
    """LIME record format: 144-byte header with magic 0x456789ab, version 1,
    MB/ME flags, data length and type; data padded to 8 bytes."""
    dims, u, z = io
    V = int(np.prod(dims))
    blob = gio.ildg_write(u, V)
    magic, ver, flags, n = struct.unpack(">IHHQ", blob[:16])
    assert (magic, ver, flags, n) == (0x456789AB, 1, 0x8000, 17)
    assert blob[16:32].rstrip(b"\0") == b"lime-test-text"
    assert blob[144:144 + 17] == b"LIME test message" and blob[161:168] == b"\0" * 7
    magic, ver, flags, n = struct.unpack(">IHHQ", blob[168:184])
    assert (magic, ver, flags, n) == (0x456789AB, 1, 0x4000, V * 72 * 8)
    assert blob[184:200].rstrip(b"\0") == b"ildg-binary-data"
    assert len(blob) == 168 + 144 + V * 72 * 8
    assert np.array_equal(gio.ildg_read(blob, V), u)
    # the data record is the big-endian (in, ex, site) block
    assert blob[312:] == z["binary_input_be"].tobytes()


def test_unit_field_observables():
    dims = (2, 2, 2, 2)
    lat = oracle.Lattice(dims)
    u = np.zeros(lat.gauge_size)
    m = u.reshape(4, lat.V, 18)
    m[:, :, [0, 8, 16]] = 1.0
    assert gio.link_trace(u, lat.V) == 1.0
    assert abs(lat.plaquette(u) - 1.0) < 1e-15
