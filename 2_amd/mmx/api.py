"""User-facing script API of mumax/2 for this path, in user (x, y, z) order.

Mirrors the functions a mumax/2 script calls (src/python/mumax2.py, generated
from src/mumax/engine/api.go): setgridsize, setcellsize, setperiodic, load,
setv, getv, gets, setarray, getarray, makearray, setmask, run, step, steps,
run_until_smaller, autotabulate / tabulate.  The reference's API layer swaps
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


class Engine:
    """One mumax/2 engine restricted to this path."""

    VECTORS = ("B_ext", "anisU", "j", "<m>")

    def __init__(self):
        self.size = None
        self.cell = None
        self.periodic = (0, 0, 0)
        self.sim = None
        self._pending = []
        self._tables = []

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
        if module in ("micromagnetism", "demag", "exchange6", "zeeman", "llg", "solver/rk12",
                      "maxtorque"):
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

    def _to_internal(self, arr):
        """[3][nx][ny][nz] (user) -> (3, nz, ny, nx) internal order, comps swapped."""
        a = np.asarray(arr, dtype=np.float64)
        nx, ny, nz = self.size
        if a.shape[1:] == (1, 1, 1):
            a = np.broadcast_to(a, (a.shape[0], nx, ny, nz))
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
        if a.ndim == 4 and a.shape[0] == 3:
            s.setmask(name, self._to_internal(a), 3)
        else:
            s.setmask(name, np.ascontiguousarray(np.asarray(a, np.float32).transpose(2, 1, 0)), 1)

    # -- time stepping --------------------------------------------------
    def _notify(self):
        for tab in self._tables:
            tab.notify(self)

    def step(self):
        self._need().step()
        self._notify()

    def steps(self, n):
        for _ in range(int(n)):
            self.step()

    def run(self, duration):
        s = self._need()
        start = s.getv("t")
        if not self._tables:
            s.run(duration)
            return
        while s.getv("t") < start + duration:
            self.step()

    def run_until_smaller(self, name, value):
        """engine/api.go:122-131"""
        while self.getv(name) >= value:
            self.step()

    def autotabulate(self, quants, filename=None, period=0.0):
        """notify_tabulate.go:23-28: a row at the first step past each period."""
        tab = AutoTable(quants, period)
        self._tables.append(tab)
        return tab

    def tabulate(self, quants):
        return [self.getv(q) for q in quants]


class AutoTable:
    def __init__(self, quants, period):
        self.quants = list(quants)
        self.period = float(period)
        self.count = 0
        self.rows = []

    def notify(self, e):
        t = e.getv("t")
        if t - self.count * self.period >= self.period:
            row = []
            for q in self.quants:
                v = e.getv(q)
                row.extend(v if isinstance(v, (list, tuple)) else [v])
            self.rows.append(row)
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
              "autotabulate", "tabulate"):
    globals()[_name] = (lambda n: lambda *a, **k: getattr(_default, n)(*a, **k))(_name)
