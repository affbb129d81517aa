"""Output formats of mumax/2 for this path's quantities: #dump002 (with its
CRC-64/ISO checksum), OVF 2.0 and OMF 1.0 (Text and Binary 4), and the table
file of tabulate/autotabulate.

Arrays are in the engine's internal layout, (ncomp, n0, n1, n2) with
component 0 = user z, cell sizes in internal order; each writer does the
reference's own reordering (the formats are defined in user x, y, z).

  dump: dump/writer.go:28-97, dump/frame.go:8-45, engine/format_dump.go:34-58
  ovf:  ovf/ovf2_output.go:28-161
  omf:  ovf/omf_output.go:36-206
  table: engine/table.go:26-71
"""
import decimal
import struct

import numpy as np

from ._lib import lib

MAGIC = b"#dump002"
OMF_CONTROL = 1234567.0


def crc64_iso(data, crc=0):
    """Go's crc64.Update(crc, crc64.MakeTable(crc64.ISO), data)."""
    b = bytes(data)
    return int(lib().mmx_crc64_iso(b, len(b), crc))


def _fixed8(s):
    b = s.encode()[:8]
    return b + b"\0" * (8 - len(b))


# ------------------------------------------------------------------ dump

def dump_bytes(data, cell, t, label, unit):
    """One #dump002 frame: header, float32 data in the internal component and
    cell order, CRC-64/ISO of everything before it (format_dump.go:40-57)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 3:
        a = a[None]
    ncomp, n0, n1, n2 = a.shape
    head = (MAGIC + struct.pack("<Q", ncomp) + struct.pack("<3Q", n0, n1, n2) +
            struct.pack("<3d", *[float(c) for c in cell]) + _fixed8("m") +
            struct.pack("<d", float(t)) + _fixed8("s") + _fixed8(label) + _fixed8(unit) +
            struct.pack("<Q", 4))
    body = a.tobytes()
    crc = int(lib().mmx_crc64_iso(head, len(head), 0))
    crc = int(lib().mmx_crc64_iso(body, len(body), crc))
    return head + body + struct.pack("<Q", crc)


def write_dump(f, data, cell, t, label, unit):
    f.write(dump_bytes(data, cell, t, label, unit))


def read_dump(buf, offset=0):
    """Parses one frame; returns (header dict, data (ncomp, n0, n1, n2), next
    offset).  Raises ValueError on a wrong magic or a CRC mismatch."""
    o = offset
    if buf[o:o + 8] != MAGIC:
        raise ValueError("not a #dump002 frame")
    ncomp, = struct.unpack_from("<Q", buf, o + 8)
    size = struct.unpack_from("<3Q", buf, o + 16)
    step = struct.unpack_from("<3d", buf, o + 40)
    meshunit = bytes(buf[o + 64:o + 72]).rstrip(b"\0").decode()
    t, = struct.unpack_from("<d", buf, o + 72)
    timeunit = bytes(buf[o + 80:o + 88]).rstrip(b"\0").decode()
    label = bytes(buf[o + 88:o + 96]).rstrip(b"\0").decode()
    unit = bytes(buf[o + 96:o + 104]).rstrip(b"\0").decode()
    prec, = struct.unpack_from("<Q", buf, o + 104)
    n = int(ncomp) * int(np.prod(size))
    end = o + 112 + 4 * n
    data = np.frombuffer(bytes(buf[o + 112:end]), dtype="<f4").reshape((int(ncomp),) + tuple(size))
    crc, = struct.unpack_from("<Q", buf, end)
    want = int(lib().mmx_crc64_iso(bytes(buf[o:end]), end - o, 0))
    if crc != want:
        raise ValueError("dump frame CRC mismatch: %x != %x" % (crc, want))
    hdr = dict(components=int(ncomp), meshsize=tuple(int(s) for s in size), meshstep=step,
               meshunit=meshunit, time=t, timeunit=timeunit, label=label, unit=unit,
               precision=int(prec), crc=crc)
    return hdr, data, end + 8


# ------------------------------------------------------- Go number printing

def go_float(x, bits=64):
    """A number as Go's fmt %v prints it: strconv's 'g' format with the
    shortest digits that round-trip (float64, or float32 for bits = 32), in
    exponent form when the decimal exponent is below -4 or at least 6, with a
    two-digit minimum exponent ("1e-09", "1.234567e+06", "123456")."""
    x = float(x)
    if bits == 32:
        s = np.format_float_scientific(np.float32(x), unique=True, trim="-")
    else:
        s = repr(x)
    if x == 0:
        return "0" if not str(s).startswith("-") else "-0"
    d = decimal.Decimal(s)
    sign, digits, exp = d.as_tuple()
    digits = "".join(str(v) for v in digits).rstrip("0") or "0"
    # decimal exponent of the leading digit
    e10 = len("".join(str(v) for v in d.as_tuple()[1])) + exp - 1
    neg = "-" if sign else ""
    if e10 < -4 or e10 >= 6:
        mant = digits[0] + ("." + digits[1:] if len(digits) > 1 else "")
        return "%s%se%s%02d" % (neg, mant, "-" if e10 < 0 else "+", abs(e10))
    if e10 >= len(digits) - 1:
        return neg + digits + "0" * (e10 - len(digits) + 1)
    if e10 >= 0:
        return neg + digits[:e10 + 1] + "." + digits[e10 + 1:]
    return neg + "0." + "0" * (-e10 - 1) + digits


def _hdr(key, *values):
    """# Key: v1 v2 ... (ovf/omf_output.go:199-202: Fprint then Fprintln)."""
    vs = []
    for v in values:
        vs.append(v if isinstance(v, str) else
                  (str(v) if isinstance(v, (int, np.integer)) else go_float(v)))
    return "# %s: %s\n" % (key, " ".join(vs))


def _user_cells(a):
    """(ncomp, n0, n1, n2) internal -> values per cell in the formats' order:
    internal axis 0 outermost, axis 2 innermost (user x fastest), components
    swapped to user x, y, z (SwapIndex, common/index.go:99-110)."""
    ncomp = a.shape[0]
    comps = a[::-1] if ncomp == 3 else a
    return np.moveaxis(comps, 0, -1)   # (n0, n1, n2, ncomp)


def _mesh_header(cell, size):
    # internal X = axis 0 = user z: xmax uses cell[Z] (internal 2)
    return cell[2], cell[1], cell[0], size[2], size[1], size[0]


# ------------------------------------------------------------------- OVF 2

def ovf2_text(data, cell, t, name, unit, dataformat="Binary 4"):
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 3:
        a = a[None]
    ncomp = a.shape[0]
    cx, cy, cz, nx, ny, nz = _mesh_header(cell, a.shape[1:])
    out = ["# OOMMF OVF 2.0\n", "#\n", _hdr("Segment count", "1"), "#\n", _hdr("Begin", "Segment"),
           _hdr("Begin", "Header"), "#\n", _hdr("Title", "mumax data"), _hdr("meshtype", "rectangular"),
           _hdr("meshunit", "m"), _hdr("xmin", 0), _hdr("ymin", 0), _hdr("zmin", 0),
           _hdr("xmax", cx * nx), _hdr("ymax", cy * ny), _hdr("zmax", cz * nz)]
    labels = [name] if ncomp == 1 else [name + "_" + "xyz"[i] for i in range(ncomp)]
    out.append(_hdr("valuedim", ncomp))
    out.append(_hdr("valuelabels", *labels))
    u = unit or "1"
    out.append(_hdr("valueunits", *([u] if ncomp == 1 else [u, u, u])))
    out.append("# Desc: Stage simulation time:  %s  s\n" % go_float(t))
    out.append("# Desc: Total simulation time:  %s  s\n" % go_float(t))
    out += [_hdr("xbase", cx / 2), _hdr("ybase", cy / 2), _hdr("zbase", cz / 2),
            _hdr("xnodes", nx), _hdr("ynodes", ny), _hdr("znodes", nz),
            _hdr("xstepsize", cx), _hdr("ystepsize", cy), _hdr("zstepsize", cz), "#\n",
            _hdr("End", "Header"), "#\n", _hdr("Begin", "Data " + dataformat)]
    head = "".join(out).encode()
    vals = _user_cells(a)
    fmt = dataformat.lower()
    if fmt == "binary 4":
        body = struct.pack("<f", OMF_CONTROL) + np.ascontiguousarray(vals, dtype="<f4").tobytes()
    elif fmt == "text":
        body = _omf_text_body(vals)
    else:
        raise ValueError("Illegal OVF data format %s. Options are: Text, Binary 4" % dataformat)
    tail = (_hdr("End", "Data " + dataformat) + _hdr("End", "Segment")).encode()
    return head + body + tail


def _omf_text_body(vals):
    rows = vals.reshape(-1, vals.shape[-1])
    lines = ["".join(go_float(v, 32) + " " for v in r) + "\n" for r in rows]
    return "".join(lines).encode()


# ------------------------------------------------------------------- OMF 1

def omf_text(data, cell, t, name, unit, dataformat="Binary 4"):
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 3:
        a = a[None]
    cx, cy, cz, nx, ny, nz = _mesh_header(cell, a.shape[1:])
    out = [_hdr("OOMMF", "rectangular mesh v1.0"), _hdr("Segment count", "1"),
           _hdr("Begin", "Segment"), _hdr("Begin", "Header"),
           "# Desc: Time :  %s\n" % go_float(t), _hdr("Title", name),
           _hdr("meshtype", "rectangular"), _hdr("meshunit", "m"),
           _hdr("xbase", cx / 2), _hdr("ybase", cy / 2), _hdr("zbase", cz / 2),
           _hdr("xstepsize", cx), _hdr("ystepsize", cy), _hdr("zstepsize", cz),
           _hdr("xmin", 0), _hdr("ymin", 0), _hdr("zmin", 0),
           _hdr("xmax", cx * nx), _hdr("ymax", cy * ny), _hdr("zmax", cz * nz),
           _hdr("xnodes", nx), _hdr("ynodes", ny), _hdr("znodes", nz),
           _hdr("ValueRangeMinMag", 1e-08), _hdr("ValueRangeMaxMag", 1),
           _hdr("valueunit", unit), _hdr("valuemultiplier", 1), _hdr("End", "Header"),
           _hdr("Begin", "Data " + dataformat)]
    head = "".join(out).encode()
    vals = _user_cells(a)
    fmt = dataformat.lower()
    if fmt == "binary 4":
        # big-endian, control number first (omf_output.go:112-140)
        body = struct.pack(">f", OMF_CONTROL) + np.ascontiguousarray(vals, dtype=">f4").tobytes()
    elif fmt == "text":
        body = _omf_text_body(vals)
    else:
        raise ValueError("Illegal OMF data format %s. Options are: Text, Binary 4" % dataformat)
    tail = (_hdr("End", "Data " + dataformat) + _hdr("End", "Segment")).encode()
    return head + body + tail


def read_ovf_binary(buf):
    """Values (cells in file order, ncomp) of an OVF 2 / OMF 1 Binary 4 file."""
    key = b"# Begin: Data Binary 4\n"
    o = buf.index(key) + len(key)
    end = buf.index(b"# End: Data Binary 4")
    le = struct.unpack_from("<f", buf, o)[0] == OMF_CONTROL
    if not le and struct.unpack_from(">f", buf, o)[0] != OMF_CONTROL:
        raise ValueError("bad OVF/OMF control number")
    vals = np.frombuffer(bytes(buf[o + 4:end]), dtype="<f4" if le else ">f4")
    dim = 3
    for line in bytes(buf[:o]).decode().splitlines():
        if line.startswith("# valuedim:"):
            dim = int(line.split(":")[1])
    return vals.reshape(-1, dim)


# ------------------------------------------------------------------- table

def table_header(quants):
    """#q_x (unit)\t... with vector components in user x, y, z order
    (table.go:51-71)."""
    out = "#"
    for name, ncomp, unit in quants:
        for i in range(ncomp - 1, -1, -1):
            comp = ("_" + chr(ord("z") - i)) if ncomp > 1 else ""
            out += name + comp + " (" + unit + ")\t"
    return out + "\n"


def table_row(values):
    """One row: each value followed by a tab (table.go:33-49); `values` are
    already in user order."""
    return "".join(go_float(v) + "\t" for v in values) + "\n"
