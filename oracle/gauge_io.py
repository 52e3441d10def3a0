"""Restatement of the reference's gauge-configuration file formats in numpy
(byte work): CCIO::BinaryReader / BinaryWriter, NERSCReader / NERSCWriter
and ILDGReader / ILDGWriter (lib/IO/*.cpp) as driven by ReadFromDisk /
SaveOnDisk (lib/IO/fields_io.hpp:47-138, 250-360) on one process.

TEST INFRASTRUCTURE ONLY.  Importable from tests/ (the checker of the
library's ii_gauge_read / ii_gauge_write); the product never imports it.

Arrays: u is a Format_G float64 array, u[18*(site + V*mu) + 2*(3a+b) + re/im]
(include/format_G.h:34-40); files are bytes.  The file order of every format
is format(gsite, in, ex) = in + 18*(ex + 4*gsite) (binary_reader.cpp:20-22).
"""
import struct

import numpy as np


def file_order(u, V):
    """Format_G [mu][site][18] -> file [site][mu][18]."""
    return np.asarray(u, dtype=np.float64).reshape(4, V, 18).transpose(1, 0, 2)


def from_file_order(a, V):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(V, 4, 18)
                                .transpose(1, 0, 2)).ravel()


def checksum(data):
    """NERSC checksum as the reference computes it: the sum, modulo 2^32, of
    the data's 32-bit words as stored in the file, read on a little-endian
    host (NERSC_reader.cpp:126-166; NERSC_writer.cpp:18-46 sums the byte-
    swapped, i.e. stored, words)."""
    w = np.frombuffer(bytes(data), dtype="<u4")
    return int(np.sum(w, dtype=np.uint64) & 0xFFFFFFFF)


def link_trace(u, V):
    """Staples::link_trace (staples.cpp:24-33) with ReTr (sunMatUtils.cpp:
    15-19): tr = ((0 + U00) + U11) + U22 per link, summed sequentially over
    mu (outer) and site, divided by V * 3 * 4.  np.cumsum adds in order."""
    m = np.asarray(u).reshape(4, V, 18)
    tr = ((0.0 + m[:, :, 0]) + m[:, :, 8]) + m[:, :, 16]
    return float(np.cumsum(tr.ravel())[-1] / (V * 3 * 4))


# ---------------------------------------------------------------- Binary
def binary_read(data, V):
    """BinaryReader::read (binary_reader.cpp:24-35): big-endian doubles."""
    a = np.frombuffer(bytes(data), dtype=">f8", count=V * 72)
    return from_file_order(a.astype(np.float64), V)


def binary_write(u, V):
    """BinaryWriter::write (binary_writer.cpp:19-22): host-order doubles
    (little-endian here), no byte swap."""
    return file_order(u, V).astype("<f8").tobytes()


# ---------------------------------------------------------------- NERSC
def nersc_header_text(dims, link_tr, plaq, chk):
    """NERSCWriter::header (NERSC_writer.cpp:55-76), C printf formats."""
    lines = ["BEGIN_HEADER", "HDR_VERSION = 1.0", "DATATYPE = 4D_SU3_GAUGE_3x3",
             "STORAGE_FORMAT = 1.0"]
    lines += ["DIMENSION_%d = %d" % (m + 1, dims[m]) for m in range(4)]
    lines += ["LINK_TRACE = %1.15g" % link_tr, "PLAQUETTE = %1.15g" % plaq,
              "CHECKSUM = %x" % chk, "ENSEMBLE_ID = JLQCD_null", "SEQUENCE_NUMBER = 0",
              "FLOATING_POINT = IEEE64BIG", "END_HEADER"]
    return ("\n".join(lines) + "\n").encode()


def nersc_write(u, dims, plaq):
    """NERSCWriter: header, then 3x3 big-endian doubles."""
    V = int(np.prod(dims))
    data = file_order(u, V).astype(">f8").tobytes()
    return nersc_header_text(dims, link_trace(u, V), plaq, checksum(data)) + data


def nersc_parse_header(blob):
    """NERSCReader::header (NERSC_reader.cpp:48-72): 'KEY = VALUE' lines
    between BEGIN_HEADER and END_HEADER; returns (dict, data offset)."""
    text_end = blob.index(b"END_HEADER\n") + len(b"END_HEADER\n")
    lines = blob[:text_end].decode().split("\n")
    if lines[0] != "BEGIN_HEADER":
        raise ValueError("NERSC type header not found")
    kv = {}
    for line in lines[1:]:
        if line == "END_HEADER":
            break
        key, _, rest = line.partition(" ")
        kv[key] = rest[2:] if rest.startswith("= ") else rest
    return kv, text_end


def reconstruct_third_row(m):
    """NERSCReader::reconstruct3x3 (NERSC_reader.cpp:218-240): the third row
    from the first two, in the reference's expression order (double)."""
    p = np.asarray(m, dtype=np.float64)
    out = np.empty(p.shape[:-1] + (18,))
    out[..., :12] = p
    out[..., 12] = p[..., 2] * p[..., 10] - p[..., 4] * p[..., 8] - p[..., 3] * p[..., 11] + p[..., 5] * p[..., 9]
    out[..., 13] = p[..., 4] * p[..., 9] - p[..., 2] * p[..., 11] + p[..., 5] * p[..., 8] - p[..., 3] * p[..., 10]
    out[..., 14] = p[..., 4] * p[..., 6] - p[..., 0] * p[..., 10] - p[..., 5] * p[..., 7] + p[..., 1] * p[..., 11]
    out[..., 15] = p[..., 0] * p[..., 11] - p[..., 4] * p[..., 7] + p[..., 1] * p[..., 10] - p[..., 5] * p[..., 6]
    out[..., 16] = p[..., 0] * p[..., 8] - p[..., 2] * p[..., 6] - p[..., 1] * p[..., 9] + p[..., 3] * p[..., 7]
    out[..., 17] = p[..., 2] * p[..., 7] - p[..., 0] * p[..., 9] + p[..., 3] * p[..., 6] - p[..., 1] * p[..., 8]
    return out


def nersc_read(blob, dims):
    """NERSCReader (header, read): returns (u, header dict, checksum of the
    data as stored)."""
    blob = bytes(blob)
    kv, off = nersc_parse_header(blob)
    V = int(np.prod(dims))
    for m in range(4):
        if int(kv["DIMENSION_%d" % (m + 1)]) != dims[m]:
            raise ValueError("NERSC dimension mismatch")
    is3x2 = kv["DATATYPE"] == "4D_SU3_GAUGE"
    bits = 32 if kv["FLOATING_POINT"] in ("IEEE32", "IEEE32BIG") else 64
    nreal = 12 if is3x2 else 18
    dt = ">f4" if bits == 32 else ">f8"
    nbytes = V * 4 * nreal * (bits // 8)
    data = blob[off:off + nbytes]
    a = np.frombuffer(data, dtype=dt).astype(np.float64).reshape(V, 4, nreal)
    if is3x2:
        a = reconstruct_third_row(a)
    return from_file_order(a, V), kv, checksum(data)


def nersc_write_variant(u, dims, is3x2, bits, plaq):
    """A NERSC file in the other accepted layouts (4D_SU3_GAUGE with two
    rows, IEEE32BIG), for the reader tests; the header follows the
    writer's lines with DATATYPE / FLOATING_POINT changed."""
    V = int(np.prod(dims))
    a = file_order(u, V)
    if is3x2:
        a = a[..., :12]
    data = a.astype(">f4" if bits == 32 else ">f8").tobytes()
    hdr = nersc_header_text(dims, link_trace(u, V), plaq, checksum(data)).decode()
    if is3x2:
        hdr = hdr.replace("DATATYPE = 4D_SU3_GAUGE_3x3", "DATATYPE = 4D_SU3_GAUGE")
    if bits == 32:
        hdr = hdr.replace("FLOATING_POINT = IEEE64BIG", "FLOATING_POINT = IEEE32BIG")
    return hdr.encode() + data


# ---------------------------------------------------------------- ILDG / LIME
LIME_MAGIC = 0x456789AB


def lime_record(rtype, data, mb, me):
    """One LIME record: 144-byte big-endian header (magic, version 1, MB/ME
    flags in bits 15/14, data length, 128-byte type) and the data padded to
    8 bytes (the c-lime record format the reference links)."""
    flags = (0x8000 if mb else 0) | (0x4000 if me else 0)
    hdr = struct.pack(">IHHQ", LIME_MAGIC, 1, flags, len(data)) + rtype.encode().ljust(128, b"\0")
    pad = (8 - len(data) % 8) % 8
    return hdr + bytes(data) + b"\0" * pad


def ildg_write(u, V):
    """ILDGWriter (ILDG_writer.cpp:22-74): a 'lime-test-text' record (MB) and
    the 'ildg-binary-data' record (ME) of big-endian doubles."""
    data = file_order(u, V).astype(">f8").tobytes()
    return (lime_record("lime-test-text", b"LIME test message", True, False)
            + lime_record("ildg-binary-data", data, False, True))


def lime_records(blob):
    """[(type, flags, data)] of a LIME file."""
    blob = bytes(blob)
    out, pos = [], 0
    while pos + 144 <= len(blob):
        magic, ver, flags, n = struct.unpack(">IHHQ", blob[pos:pos + 16])
        if magic != LIME_MAGIC:
            raise ValueError("bad LIME magic")
        rtype = blob[pos + 16:pos + 144].split(b"\0")[0].decode()
        out.append((rtype, flags, blob[pos + 144:pos + 144 + n]))
        pos += 144 + n + (8 - n % 8) % 8
    return out


def ildg_read(blob, V):
    """ILDGReader (ILDG_reader.cpp:22-54): the first 'ildg-binary-data'
    record as big-endian doubles."""
    for rtype, _, data in lime_records(blob):
        if rtype.startswith("ildg-binary-data"):
            return binary_read(data, V)
    raise ValueError("no ildg-binary-data record")
