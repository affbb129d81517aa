"""Output formats (mmx.output): #dump002 with CRC-64/ISO, OVF 2.0, OMF 1.0,
table files.  CPU only (the CRC runs in libmmx's host code).

Pins: the CRC-64/GO-ISO check value of the CRC catalogue ("123456789" ->
0xb90956c775a41001); the frame of the reference's own dump test
(dump/dump_test.go: 1 component, 2x4x9 cells, data 0..71, t = 1e-15, steps
1e-9, 2e-9, 3e-9) against the field layout of dump/reader.go; the OVF 2.0
(little-endian) and OMF 1.0 (big-endian) control number 1234567.0; the cell
and component order of ovf2_output.go:139-160 (user x fastest, components
x, y, z); Go's %v float printing on documented cases; round trips.
"""
import struct

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O():
    from mmx import output
    return output


def test_crc64_iso_check_value(O):
    assert O.crc64_iso(b"123456789") == 0xB90956C775A41001
    # continuation = one pass
    d = bytes(range(256)) * 3
    assert O.crc64_iso(d[100:], O.crc64_iso(d[:100])) == O.crc64_iso(d)


def test_dump_frame_of_the_reference_test(O):
    data = np.arange(2 * 4 * 9, dtype=np.float32).reshape(1, 2, 4, 9)
    b = O.dump_bytes(data, (1e-9, 2e-9, 3e-9), 1e-15, "", "")
    assert len(b) == 112 + 4 * 72 + 8
    assert b[:8] == b"#dump002"
    assert struct.unpack_from("<Q", b, 8)[0] == 1
    assert struct.unpack_from("<3Q", b, 16) == (2, 4, 9)
    assert struct.unpack_from("<3d", b, 40) == (1e-9, 2e-9, 3e-9)
    assert b[64:72] == b"m" + b"\0" * 7
    assert struct.unpack_from("<d", b, 72)[0] == 1e-15
    assert b[80:88] == b"s" + b"\0" * 7
    assert struct.unpack_from("<Q", b, 104)[0] == 4
    assert np.array_equal(np.frombuffer(b[112:112 + 288], "<f4"), np.arange(72, dtype=np.float32))
    assert struct.unpack_from("<Q", b, 400)[0] == O.crc64_iso(b[:400])
    hdr, back, nxt = O.read_dump(b)
    assert nxt == len(b) and hdr["meshsize"] == (2, 4, 9) and np.array_equal(back, data)
    bad = bytearray(b)
    bad[200] ^= 1
    with pytest.raises(ValueError):
        O.read_dump(bytes(bad))


def test_dump_vector_and_labels(O):
    rng = np.random.default_rng(0)
    a = rng.normal(size=(3, 2, 3, 5)).astype(np.float32)
    b = O.dump_bytes(a, (3e-9, 4e-9, 5e-9), 2.5e-12, "H_eff_long_label", "A/m")
    hdr, back, _ = O.read_dump(b)
    assert hdr["components"] == 3 and hdr["label"] == "H_eff_lo" and hdr["unit"] == "A/m"
    assert np.array_equal(back, a)       # internal component order, no swap
    # frames append (AutoSaveSingleFile)
    two = b + O.dump_bytes(2 * a, (3e-9, 4e-9, 5e-9), 5e-12, "m", "")
    _, _, o = O.read_dump(two)
    h2, d2, _ = O.read_dump(two, o)
    assert h2["time"] == 5e-12 and np.array_equal(d2, 2 * a)


def test_ovf2_binary_and_text(O):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 2, 3, 4)).astype(np.float32)     # internal (z, y, x)
    cell = (5e-9, 4e-9, 3e-9)                                  # internal
    b = O.ovf2_text(a, cell, 1e-9, "m", "")
    head = b[:b.index(b"# Begin: Data Binary 4\n")].decode()
    lines = head.splitlines()
    assert lines[0] == "# OOMMF OVF 2.0" and "# Segment count: 1" in lines
    for want in ("# xnodes: 4", "# ynodes: 3", "# znodes: 2", "# xstepsize: 3e-09",
                 "# zstepsize: 5e-09", "# xmax: 1.2e-08", "# valuedim: 3",
                 "# valuelabels: m_x m_y m_z", "# valueunits: 1 1 1",
                 "# Desc: Total simulation time:  1e-09  s", "# xbase: 1.5e-09"):
        assert want in lines, want
    o = b.index(b"# Begin: Data Binary 4\n") + len(b"# Begin: Data Binary 4\n")
    assert struct.unpack_from("<f", b, o)[0] == 1234567.0
    vals = O.read_ovf_binary(b)
    assert vals.shape == (24, 3)
    # cell (i, j, k) = internal indices, k (user x) fastest; components x, y, z
    for (i, j, k) in [(0, 0, 0), (1, 2, 3), (0, 1, 2)]:
        row = vals[(i * 3 + j) * 4 + k]
        assert np.array_equal(row, [a[2, i, j, k], a[1, i, j, k], a[0, i, j, k]])
    assert b.endswith(b"# End: Data Binary 4\n# End: Segment\n")
    t = O.ovf2_text(a, cell, 1e-9, "m", "", "Text")
    body = t[t.index(b"# Begin: Data Text\n") + 19:t.index(b"# End: Data Text")].decode()
    rows = [list(map(float, r.split())) for r in body.splitlines()]
    assert np.array_equal(np.array(rows, np.float32), vals)
    assert body.splitlines()[0].endswith(" ")
    with pytest.raises(ValueError):
        O.ovf2_text(a, cell, 0, "m", "", "Binary 8")


def test_omf_big_endian(O):
    a = np.arange(3 * 1 * 2 * 2, dtype=np.float32).reshape(3, 1, 2, 2)
    b = O.omf_text(a, (3e-9, 4e-9, 4e-9), 2e-12, "m", "")
    assert b.startswith(b"# OOMMF: rectangular mesh v1.0\n")
    assert b"# Desc: Time :  2e-12\n" in b
    o = b.index(b"# Begin: Data Binary 4\n") + len(b"# Begin: Data Binary 4\n")
    assert struct.unpack_from(">f", b, o)[0] == 1234567.0
    vals = O.read_ovf_binary(b)
    assert np.array_equal(vals[3], [a[2, 0, 1, 1], a[1, 0, 1, 1], a[0, 0, 1, 1]])


def test_go_float_printing(O):
    cases = {0.0: "0", 1.0: "1", 4e-9: "4e-09", 2.5e-9: "2.5e-09", 1e6: "1e+06",
             123456.0: "123456", 1234567.0: "1.234567e+06", 0.0001: "0.0001", 1e-5: "1e-05",
             -3.25: "-3.25", 1e21: "1e+21", 5.12e-7: "5.12e-07", 0.5: "0.5", 100.0: "100"}
    for v, want in cases.items():
        assert O.go_float(v) == want, (v, O.go_float(v))
    assert O.go_float(0.1, 32) == "0.1"
    assert O.go_float(1 / 3, 32) == "0.33333334"


def test_table_format(O):
    h = O.table_header([("t", 1, "s"), ("<m>", 3, "")])
    assert h == "#t (s)\t<m>_x ()\t<m>_y ()\t<m>_z ()\t\n"
    assert O.table_row([1e-12, 0.5, -0.25, 1.0]) == "1e-12\t0.5\t-0.25\t1\t\n"
