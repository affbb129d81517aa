"""User-facing script API of mumax/2 for this path, in user (x, y, z) order.

Mirrors the functions a mumax/2 script calls (src/python/mumax2.py, generated
from src/mumax/engine/api.go): setgridsize, setcellsize, setperiodic, load,
setv, getv, gets, setarray, getarray, makearray, setmask, run, step, steps,
run_until_smaller, autotabulate / tabulate, save / saveas / autosave / remove,
outputdir, filenumberformat (output formats in mmx.output).  The reference's API layer swaps
user XYZ into the engine's internal ZYX order (engine/api.go:43-45,
frontend/jsonrpc.go:297-312); so does this module, over mmx.Sim.

Unlike the reference there is no separate process: one Engine object holds
one simulation (module-level functions act on a default engine).
"""
import numpy as np

from .sim import Sim

MU0 = 4.0 * 3.14159265358979323846264338327950288 * 1e-7


def _swap3(v):
    return (v[2], v[1], v[0])


def makearray(ncomp, nx, ny, nz):
    """[ncomp][nx][ny][nz] nested lists (mumax2.py.tail makearray)."""
    return np.zeros((ncomp, nx, ny, nz)).tolist()


def resample(arr, size):
    """engine/resample.go:22-42: out[c][i][j][k] = in[c][i*n1/n2][j*..][k*..]
    (integer division) for an array [c][x][y][z] onto size (nx, ny, nz)."""
    a = np.asarray(arr, dtype=np.float64)
    if a.shape[1:] == tuple(size):
        return a
    ix, iy, iz = (np.arange(n2) * n1 // n2 for n1, n2 in zip(a.shape[1:], size))
    return a[:, ix][:, :, iy][:, :, :, iz]


class Engine:
    """One mumax/2 engine restricted to this path."""

    VECTORS = ("B_ext", "anisU", "j", "<m>", "<H_therm>")

    UNITS = {"m": "", "H_eff": "A/m", "B": "T", "torque": "/s"}   # modules/*.go AddNewQuant

    def __init__(self):
        self.size = None
        self.cell = None
        self.periodic = (0, 0, 0)
        self.sim = None
        self._pending = []
        self._tables = []
        self._autosaves = {}
        self._handles = 0
        self.out_dir = ""
        self.filename_format = "%06d"      # engine.go:84
        self._last_output_t = None
        self._outid = 0

    # -- geometry ---------------------------------------------------------
    def setgridsize(self, nx, ny, nz):
        self.size = (int(nx), int(ny), int(nz))

    def setcellsize(self, cx, cy, cz):
        self.cell = (float(cx), float(cy), float(cz))

    def setperiodic(self, px, py, pz):
        self.periodic = (int(px), int(py), int(pz))

    def _need(self):
        if self.sim is None:
            if self.size is None or self.cell is None:
                raise ValueError("setgridsize and setcellsize first")
            self.sim = Sim(_swap3(self.size), _swap3(self.cell), _swap3(self.periodic))
            for f in self._pending:
                f()
            self._pending = []
        return self.sim

    # -- modules / quantities -------------------------------------------
    def load(self, module):
        if module in ("micromagnetism", "demag", "exchange6", "zeeman", "llg", "maxtorque"):
            return
        self._need().load(module)

    def setv(self, name, value):
        s = self._need()
        if name in self.VECTORS:
            s.setv(name, _swap3(value))
        else:
            s.setv(name, value)

    def getv(self, name):
        s = self._need()
        if name in self.VECTORS:
            return list(_swap3(s.getv(name, 3)))
        return s.getv(name, 1)

    def gets(self, name):
        return self.getv(name)

    def _resampled(self, arr):
        """[c][x][y][z] onto the grid size by nearest (floored) cells, as
        API.SetArray/SetMask resample an array of another size
        (engine/api.go:234-264, engine/resample.go:22-42)."""
        return resample(arr, self.size)

    def _to_internal(self, arr):
        """[3][nx][ny][nz] (user) -> (3, nz, ny, nx) internal order, comps swapped."""
        a = self._resampled(arr)
        a = a[::-1].transpose(0, 3, 2, 1)
        return np.ascontiguousarray(a, dtype=np.float32)

    def _to_user(self, a):
        return np.ascontiguousarray(np.asarray(a)[::-1].transpose(0, 3, 2, 1))

    def setarray(self, name, arr):
        self._need().setarray(name, self._to_internal(arr))

    def getarray(self, name):
        return self._to_user(self._need().getarray(name))

    def setmask(self, name, arr):
        s = self._need()
        a = np.asarray(arr, dtype=np.float64)
        if a.ndim == 3:
            a = a[None]
        if a.shape[0] == 3:
            s.setmask(name, self._to_internal(a), 3)
        else:
            a = self._resampled(a)[0]
            s.setmask(name, np.ascontiguousarray(a.astype(np.float32).transpose(2, 1, 0)), 1)

    # -- output (engine.go:98-105,554-582, notify_autosave.go:25-31) -----
    def outputdir(self, d):
        self.out_dir = d

    def filenumberformat(self, fmt):
        self.filename_format = fmt

    def _output_id(self):
        t = self._need().getv("t")
        if t != self._last_output_t:
            self._last_output_t = t
            self._outid += 1
        return self._outid

    def autofilename(self, quantity, format):
        name = quantity + (self.filename_format % self._output_id()) + "." + format
        return (self.out_dir + "/" if self.out_dir else "") + name

    def _field(self, quantity):
        s = self._need()
        a = np.asarray(s.getarray(quantity), dtype=np.float64)
        if quantity == "torque":          # Quant.Buffer applies the multiplier gamma
            a = a * s.getv("gamma")
        return a.astype(np.float32)

    def saveas(self, quantity, format, options, filename):
        """Engine.SaveAs: formats dump, ovf (OVF 2.0), omf (OMF 1.0); options
        ["Text"] or ["Binary 4"] (the default) for ovf/omf."""
        from . import output as O
        s = self._need()
        data = self._field(quantity)
        t = s.getv("t")
        unit = self.UNITS.get(quantity, "")
        opts = list(options or [])
        if format == "dump":
            if opts:
                raise ValueError("dump output format does not take options")
            blob = O.dump_bytes(data, s.cell, t, quantity, unit)
        elif format in ("ovf", "omf"):
            if len(opts) > 1:
                raise ValueError("Illegal %s options: %s" % (format.upper(), opts))
            df = opts[0] if opts else "Binary 4"
            blob = (O.ovf2_text if format == "ovf" else O.omf_text)(data, s.cell, t, quantity,
                                                                     unit, df)
        else:
            raise ValueError("Unknown output format: %s. Options are: dump ovf omf" % format)
        with open(filename, "wb") as f:
            f.write(blob)
        return filename

    def save(self, quantity, format, options=()):
        return self.saveas(quantity, format, options, self.autofilename(quantity, format))

    def autosave(self, quantity, format, options=(), period=0.0):
        self._handles += 1
        self._autosaves[self._handles] = dict(q=quantity, f=format, o=list(options or []),
                                              period=float(period), start=self._need().getv("t"),
                                              count=0)
        return self._handles

    def remove(self, handle):
        if handle in self._autosaves:
            del self._autosaves[handle]
        else:
            self._tables = [t for t in self._tables if getattr(t, "handle", None) != handle]

    # -- time stepping --------------------------------------------------
    def _notify(self):
        for tab in self._tables:
            tab.notify(self)
        for a in self._autosaves.values():
            t = self._need().getv("t") - a["start"]
            if t - a["count"] * a["period"] >= a["period"]:
                self.save(a["q"], a["f"], a["o"])
                a["count"] += 1

    def step(self):
        self._need().step()
        self._notify()

    def steps(self, n):
        for _ in range(int(n)):
            self.step()

    def run(self, duration):
        s = self._need()
        start = s.getv("t")
        if not self._tables and not self._autosaves:
            s.run(duration)
            return
        while s.getv("t") < start + duration:
            self.step()

    def run_until_smaller(self, name, value):
        """engine/api.go:122-131"""
        while self.getv(name) >= value:
            self.step()

    def autotabulate(self, quants, filename=None, period=0.0):
        """notify_tabulate.go:23-28: a row at the first step past each period;
        with a file name the rows are also appended to that table file
        (table.go:26-71)."""
        tab = AutoTable(quants, period, filename)
        self._handles += 1
        tab.handle = self._handles
        self._tables.append(tab)
        return tab

    def tabulate(self, quants, filename=None):
        """Engine.Tabulate: the values (user order); with a file name, one
        row appended to that table file."""
        vals = [self.getv(q) for q in quants]
        if filename:
            _append_row(filename, self, quants, vals)
        return vals


_UNIT_OF = {"t": "s", "dt": "s", "<m>": "", "maxtorque": "/s", "B_ext": "T", "j": "A/m2",
            "E": "J", "E_demag": "J", "E_ex": "J", "E_zeeman": "J", "E_anis": "J",
            "Msat": "A/m", "Aex": "J/m", "alpha": "", "step": "", "m_error": "",
            "m_peakerror": "", "badsteps": ""}


def _append_row(filename, e, quants, vals):
    """table.go:26-49: the header at the first row, then values + tabs."""
    from . import output as O
    import os
    fresh = not os.path.exists(filename) or os.path.getsize(filename) == 0
    with open(filename, "a") as f:
        if fresh:
            spec = [(q, len(v) if isinstance(v, (list, tuple)) else 1, _UNIT_OF.get(q, ""))
                    for q, v in zip(quants, vals)]
            f.write(O.table_header(spec))
        flat = []
        for v in vals:
            flat.extend(v if isinstance(v, (list, tuple)) else [v])
        f.write(O.table_row(flat))


class AutoTable:
    def __init__(self, quants, period, filename=None):
        self.quants = list(quants)
        self.period = float(period)
        self.filename = filename
        self.count = 0
        self.rows = []

    def notify(self, e):
        t = e.getv("t")
        if t - self.count * self.period >= self.period:
            vals = [e.getv(q) for q in self.quants]
            row = []
            for v in vals:
                row.extend(v if isinstance(v, (list, tuple)) else [v])
            self.rows.append(row)
            if self.filename:
                _append_row(self.filename, e, self.quants, vals)
            self.count += 1


_default = Engine()


def engine():
    return _default


def reset():
    global _default
    _default = Engine()
    return _default


for _name in ("setgridsize", "setcellsize", "setperiodic", "load", "setv", "getv", "gets",
              "setarray", "getarray", "setmask", "step", "steps", "run", "run_until_smaller",
              "autotabulate", "tabulate", "save", "saveas", "autosave", "remove", "outputdir",
              "filenumberformat"):
    globals()[_name] = (lambda n: lambda *a, **k: getattr(_default, n)(*a, **k))(_name)
