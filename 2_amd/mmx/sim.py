"""Object wrappers of the engine-level C entry points (internal z,y,x order)."""
import ctypes as C

import numpy as np

from ._lib import lib, check, MMX_KEEP_REAL_KERNEL, MMX_KERNEL_ARNE


def _f32(a, ncomp, N):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.size != ncomp * N:
        raise ValueError("array has %d values, expected %d" % (a.size, ncomp * N))
    return a


class Sim:
    """One micromagnetic simulation (mmx_sim): load('micromagnetism') semantics.

    n, cell, pbc are in the reference's internal order (z, y, x).  Quantity
    names are the reference's (Msat, Aex, alpha, gamma, B_ext, Ku, anisU, j,
    polarisation, xi, t, dt, mindt, maxdt, m_maxerror, <m>, maxtorque, ...).
    """

    def __init__(self, n, cell, pbc=(0, 0, 0)):
        self.n = tuple(int(x) for x in n)
        self.N = int(np.prod(self.n))
        self.cell = tuple(float(c) for c in cell)
        self.pbc = tuple(int(p) for p in pbc)
        h = C.c_void_p()
        check(lib().mmx_sim_create(C.byref(h), *self.n, *self.cell, *self.pbc))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mmx_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, module):
        check(lib().mmx_sim_load(self._h, module.encode()))

    def set_surrogate(self, sur, tol=1e-3, check_every=8):
        check(lib().mmx_sim_set_surrogate(self._h, None if sur is None else sur._h, float(tol),
                                          int(check_every)))

    def set_kernel(self, kernel_type=MMX_KERNEL_ARNE, accuracy=6):
        check(lib().mmx_sim_set_kernel(self._h, int(kernel_type), int(accuracy)))

    def setv(self, name, value):
        v = np.atleast_1d(np.asarray(value, dtype=np.float64))
        buf = (C.c_double * len(v))(*v)
        check(lib().mmx_sim_setv(self._h, name.encode(), buf, len(v)))

    def getv(self, name, ncomp=None):
        if ncomp is None:
            ncomp = 3 if name in ("<m>", "<H_therm>", "B_ext", "anisU", "j") else 1
        buf = (C.c_double * ncomp)()
        check(lib().mmx_sim_getv(self._h, name.encode(), buf, ncomp))
        return list(buf) if ncomp > 1 else buf[0]

    def setmask(self, name, array, ncomp=1):
        if array is None:
            check(lib().mmx_sim_setmask(self._h, name.encode(), None, ncomp))
            return
        a = _f32(array, ncomp, self.N)
        check(lib().mmx_sim_setmask(self._h, name.encode(), a.ctypes.data, ncomp))

    def setarray(self, name, array):
        a = _f32(array, 3, self.N)
        check(lib().mmx_sim_setarray(self._h, name.encode(), a.ctypes.data, 3))

    def getarray(self, name):
        out = np.empty((3,) + self.n, dtype=np.float32)
        check(lib().mmx_sim_getarray(self._h, name.encode(), out.ctypes.data, 3))
        return out

    def device_array(self, name):
        p = C.c_void_p()
        check(lib().mmx_sim_device_array(self._h, name.encode(), C.byref(p)))
        return p.value

    def stream(self):
        p = C.c_void_p()
        check(lib().mmx_sim_stream(self._h, C.byref(p)))
        return p.value

    def update(self, name="H_eff"):
        check(lib().mmx_sim_update(self._h, name.encode()))

    def step(self):
        check(lib().mmx_sim_step(self._h))

    def steps(self, n):
        check(lib().mmx_sim_steps(self._h, int(n)))

    def run(self, duration):
        check(lib().mmx_sim_run(self._h, float(duration)))

    PASSES_3D = ("xf", "yf", "z", "yi", "xi")
    PASSES_2D = ("xf", "y2d", "xi")

    def time_passes(self, reps=10):
        """{pass name: average ms} over `reps` H_eff evaluations (HIP events)."""
        ms = (C.c_double * 8)()
        npass = C.c_int(0)
        check(lib().mmx_sim_time_passes(self._h, int(reps), ms, C.byref(npass)))
        names = self.PASSES_2D if npass.value == 3 else self.PASSES_3D
        return {names[i]: ms[i] for i in range(npass.value)}

    def try_step(self, dt):
        err = C.c_double(0)
        check(lib().mmx_sim_try_step(self._h, float(dt), C.byref(err)))
        return err.value


class Surrogate:
    """Learned demag correction for one simulation's grid (mmx_surrogate).

    Parameters are a flat float32 vector: per layer l = 0..3 the weights
    (3, 3, Cin, Cout) then the biases (Cout,), channels 32, 32, 32, 32, 16."""

    CH = (32, 32, 32, 32, 16)

    def __init__(self, sim, seed=0):
        self.sim = sim
        h = C.c_void_p()
        check(lib().mmx_surrogate_create(C.byref(h), sim._h, int(seed)))
        self._h = h
        n = C.c_int()
        check(lib().mmx_surrogate_num_params(h, C.byref(n)))
        self.nparams = n.value
        self.n = sim.n

    def close(self):
        if getattr(self, "_h", None):
            lib().mmx_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.empty(self.nparams, np.float32)
        check(lib().mmx_surrogate_get_params(self._h, out.ctypes.data, self.nparams))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.nparams
        check(lib().mmx_surrogate_set_params(self._h, flat.ctypes.data, self.nparams))

    def get_grads(self):
        out = np.empty(self.nparams, np.float32)
        check(lib().mmx_surrogate_get_grads(self._h, out.ctypes.data, self.nparams))
        return out

    @classmethod
    def unflatten(cls, flat):
        Ws, bs, o = [], [], 0
        for l in range(4):
            ci, co = cls.CH[l], cls.CH[l + 1]
            Ws.append(np.asarray(flat[o:o + 9 * ci * co], np.float64).reshape(3, 3, ci, co))
            o += 9 * ci * co
            bs.append(np.asarray(flat[o:o + co], np.float64))
            o += co
        return Ws, bs

    @classmethod
    def flatten(cls, Ws, bs):
        return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in zip(Ws, bs)]
                              ).astype(np.float32)

    def net(self, x, tgt=None):
        """x: (n1, n2, 32) -> out (n1, n2, 16); with tgt (n1, n2, 3 n0) also the loss."""
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty(x.shape[:2] + (16,), np.float32)
        loss = C.c_double(0)
        t = None if tgt is None else np.ascontiguousarray(tgt, np.float32)
        check(lib().mmx_surrogate_net(self._h, x.ctypes.data, None if t is None else t.ctypes.data,
                                      out.ctypes.data, C.byref(loss)))
        return out if tgt is None else (out, loss.value)

    def train_step(self, m_ptr, lr):
        loss = C.c_double(0)
        check(lib().mmx_surrogate_train_step(self._h, m_ptr, float(lr), C.byref(loss)))
        return loss.value

    def predict_b(self, m_ptr, b_ptr, coarse_only=False):
        check(lib().mmx_surrogate_predict_b(self._h, m_ptr, b_ptr, 1 if coarse_only else 0))


class Batch:
    """Many simulations stepped together (mmx_batch): one launch per demag
    pass for all of them, RK12 step control on the device, tries replayed
    from a hipGraph.  Borrows the Sim objects (keeps references to them)."""

    PASSES = ("xf", "y", "z", "yi", "xi")

    def __init__(self, sims, tries_per_graph=8):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        check(lib().mmx_batch_create(C.byref(h), arr, len(self.sims)))
        self._h = h
        self.set_graph(tries_per_graph)

    def close(self):
        if getattr(self, "_h", None):
            lib().mmx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, tries_per_graph):
        check(lib().mmx_batch_set_graph(self._h, int(tries_per_graph)))

    def set_persistent(self, tries_per_launch):
        """One persistent launch per `tries_per_launch` tries (0: off)."""
        check(lib().mmx_batch_set_persistent(self._h, int(tries_per_launch)))

    def run(self, duration):
        check(lib().mmx_batch_run(self._h, float(duration)))

    def steps(self, n):
        check(lib().mmx_batch_steps(self._h, int(n)))

    def tune(self, reps=3):
        out = (C.c_int * 5)()
        check(lib().mmx_batch_tune(self._h, int(reps), out))
        return dict(zip(self.PASSES, list(out)))

    def set_lds(self, kb):
        arr = (C.c_int * 5)(*[int(kb[p]) for p in self.PASSES])
        check(lib().mmx_batch_set_lds(self._h, arr))

    def info(self):
        nb = (C.c_int * 5)()
        lds = (C.c_int * 5)()
        check(lib().mmx_batch_info(self._h, nb, lds))
        return {p: {"blocks": nb[i], "lds_bytes": lds[i]} for i, p in enumerate(self.PASSES)}


class DemagPlan:
    """MaxwellPlan for the demag source (mmx_demag_plan)."""

    def __init__(self, n, cell, pbc=(0, 0, 0), kernel_type=MMX_KERNEL_ARNE, accuracy=6,
                 keep_real_kernel=False):
        self.n = tuple(int(x) for x in n)
        h = C.c_void_p()
        kt = int(kernel_type) | (MMX_KEEP_REAL_KERNEL if keep_real_kernel else 0)
        check(lib().mmx_demag_plan_create(C.byref(h), *self.n, *[float(c) for c in cell],
                                          *[int(p) for p in pbc], kt, int(accuracy)))
        self._h = h
        p = (C.c_int * 3)()
        ko = (C.c_int * 3)()
        bad = C.c_double(0)
        check(lib().mmx_demag_plan_info(h, p, ko, C.byref(bad)))
        self.p = tuple(p)
        self.ko = tuple(ko)
        self.bad_over_good = bad.value

    def close(self):
        if getattr(self, "_h", None):
            lib().mmx_demag_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def kernel(self):
        """Octant kernel, float32 (6, ko0, ko1, ko2), XX,YY,ZZ,YZ,XZ,XY, incl. 1/Np."""
        out = np.empty((6,) + self.ko, dtype=np.float32)
        check(lib().mmx_demag_plan_kernel(self._h, out.ctypes.data))
        return out

    def real_kernel(self):
        """Real-space kernel (6, p0, p1, p2) as Kernel_Arne stores it."""
        out = np.empty((6,) + self.p, dtype=np.float32)
        check(lib().mmx_demag_plan_real_kernel(self._h, out.ctypes.data))
        return out

    def update_b(self, b_ptr, m_ptr, msat_mul, msat_map_ptr=None, stream=None):
        """B = mu0 Ms K*m on device buffers (base addresses of 3N floats)."""
        N = int(np.prod(self.n))
        check(lib().mmx_demag_update_b(self._h, b_ptr, b_ptr + 4 * N, b_ptr + 8 * N, m_ptr,
                                       m_ptr + 4 * N, m_ptr + 8 * N, msat_map_ptr,
                                       float(msat_mul), stream))


class Reductor:
    """Reductor (mmx_reductor) over device float arrays."""

    def __init__(self, n):
        h = C.c_void_p()
        check(lib().mmx_reductor_create(C.byref(h), int(n)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mmx_reductor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sum(self, a, n, stream=None):
        out = C.c_double(0)
        check(lib().mmx_reductor_sum(self._h, a, int(n), stream, C.byref(out)))
        return out.value

    def dot(self, a, b, n, stream=None):
        out = C.c_double(0)
        check(lib().mmx_reductor_dot(self._h, a, b, int(n), stream, C.byref(out)))
        return out.value

    def _f(self, fn, *args):
        out = C.c_float(0)
        check(fn(self._h, *args, C.byref(out)))
        return out.value

    def max(self, a, n, stream=None):
        return self._f(lib().mmx_reductor_max, a, int(n), stream)

    def min(self, a, n, stream=None):
        return self._f(lib().mmx_reductor_min, a, int(n), stream)

    def maxabs(self, a, n, stream=None):
        return self._f(lib().mmx_reductor_maxabs, a, int(n), stream)

    def maxdiff(self, a, b, n, stream=None):
        return self._f(lib().mmx_reductor_maxdiff, a, b, int(n), stream)

    def maxnorm(self, ax, ay, az, n, stream=None):
        return self._f(lib().mmx_reductor_maxnorm, ax, ay, az, int(n), stream)
