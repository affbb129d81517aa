"""ctypes front end of the FP64 CPU oracle (oracle/libmmx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, as the checker.
The product (2_amd/) never imports it.

All arrays are numpy float64 in the reference's internal (z,y,x) order,
vector fields shaped (3, n0, n1, n2) with component 0 = user z.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmmx_oracle.so")

MU0 = 4.0 * 3.14159265358979323846264338327950288 * 1e-7
GAMMA0 = 2.211e5
MUB = 9.2740091523e-24
QE = 1.60217646e-19

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def build():
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    lib.orc_pad_size.argtypes = [_ip, _ip, _ip]
    lib.orc_kernel_arne.argtypes = [_ip, _dp, _ip, C.c_int, _fp]
    lib.orc_fft_line.argtypes = [_dp, C.c_int, C.c_int]
    lib.orc_demag_kernel_fft.argtypes = [_ip, _fp, _dp, _dp, _dp]
    lib.orc_demag_conv.argtypes = [_ip, _ip, _dp, _dp, _dp, _dp]
    lib.orc_demag_brute.argtypes = [_ip, _ip, _fp, _dp, _dp, _dp]
    lib.orc_exchange6.argtypes = [_ip, _ip, _dp, C.c_double, _dp, _dp, _dp, _dp]
    lib.orc_exchange6.restype = None
    lib.orc_uniaxial.argtypes = [C.c_long, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_uniaxial.restype = None
    lib.orc_torque.argtypes = [C.c_long, _dp, _dp, _dp, C.c_double, _dp]
    lib.orc_torque.restype = None
    lib.orc_normalize.argtypes = [C.c_long, _dp, _dp]
    lib.orc_normalize.restype = None
    lib.orc_zhangli.argtypes = [_ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                C.c_double, C.c_double, C.c_double, _dp]
    lib.orc_zhangli.restype = None
    lib.orc_heff.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.orc_torque_full.argtypes = [C.c_void_p, _dp, _dp, _dp]
    lib.orc_rk12_init.argtypes = [C.c_void_p]
    lib.orc_rk12_init.restype = None
    lib.orc_rk12_step.argtypes = [C.c_void_p, C.c_void_p, _dp]
    lib.orc_rk12_try.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_double, _dp]
    lib.orc_average.argtypes = [C.c_long, C.c_long, _dp, _dp]
    lib.orc_average.restype = None
    lib.orc_maxnorm.argtypes = [C.c_long, _dp]
    lib.orc_maxnorm.restype = C.c_double
    lib.orc_maxdiff.argtypes = [C.c_long, _dp, _dp]
    lib.orc_maxdiff.restype = C.c_double
    lib.orc_newell_point.argtypes = [_dp, _dp, C.c_int, _dp]
    lib.orc_newell_point.restype = None
    lib.orc_newell_rule.argtypes = [C.c_int, _dp, _dp]
    lib.orc_kernel_newell.argtypes = [_ip, _dp, _ip, _dp]
    lib.orc_newell_khat.argtypes = [_ip, _dp, _ip, _dp, _dp]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def pad_size(n, pbc=(0, 0, 0)):
    p = _i3([0, 0, 0])
    lib().orc_pad_size(_i3(n), _i3(pbc), p)
    return tuple(p)


def kernel_arne(p, cell, pbc=(0, 0, 0), accuracy=6):
    """Kernel_Arne on the padded grid p: float32 array (9, p0, p1, p2)."""
    K = np.zeros((9,) + tuple(p), dtype=np.float32)
    rc = lib().orc_kernel_arne(_i3(p), _d3(cell), _i3(pbc), int(accuracy), _ptr(K, C.c_float))
    if rc:
        raise ValueError("orc_kernel_arne rejected size %s" % (p,))
    return K


def kernel_newell(p, cell, pbc=(0, 0, 0)):
    """FP64 Newell kernel on the padded grid p: float64 array (9, p0, p1, p2)."""
    K = np.zeros((9,) + tuple(p), dtype=np.float64)
    rc = lib().orc_kernel_newell(_i3(p), _d3(cell), _i3(pbc), _ptr(K))
    if rc:
        raise ValueError("orc_kernel_newell rejected %s pbc %s" % (p, pbc))
    return K


def newell_khat(p, cell, pbc=(0, 0, 0)):
    """Transformed Newell kernel (as kernel_fft) without the real-space kernel."""
    h2 = p[2] // 2 + 1
    kh = np.zeros((6, p[0], p[1], h2))
    bad = C.c_double(0)
    rc = lib().orc_newell_khat(_i3(p), _d3(cell), _i3(pbc), _ptr(kh), C.byref(bad))
    if rc:
        raise ValueError("orc_newell_khat rejected %s pbc %s" % (p, pbc))
    return kh, bad.value


def newell_point(R, d, method=0):
    """K = -N at displacement R for cell d (6 stored components)."""
    out = np.zeros(6)
    lib().orc_newell_point(_d3(R), _d3(d), int(method), _ptr(out))
    return out


def newell_rule(n):
    t, w = np.zeros(n), np.zeros(n)
    if lib().orc_newell_rule(int(n), _ptr(t), _ptr(w)):
        raise ValueError("rule order %d" % n)
    return t, w


def fft_line(z, sign=-1):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().orc_fft_line(z.view(np.float64).ctypes.data_as(_dp), len(z), int(sign))
    return z


def kernel_fft(p, K9=None, K9d=None):
    """Returns (khat6 float64 (6, p0, p1, p2//2+1), bad_over_good)."""
    h2 = p[2] // 2 + 1
    kh = np.zeros((6, p[0], p[1], h2))
    bad = C.c_double(0)
    K9 = None if K9 is None else np.ascontiguousarray(K9, dtype=np.float32)
    K9d = _f64(K9d)
    rc = lib().orc_demag_kernel_fft(_i3(p), _ptr(K9, C.c_float), _ptr(K9d), _ptr(kh), C.byref(bad))
    if rc:
        raise RuntimeError("orc_demag_kernel_fft failed %d" % rc)
    return kh, bad.value


def demag_conv(n, p, khat6, m, msat_map=None):
    m = _f64(m)
    out = np.zeros((3,) + tuple(n))
    rc = lib().orc_demag_conv(_i3(n), _i3(p), _ptr(_f64(khat6)), _ptr(m), _ptr(_f64(msat_map)), _ptr(out))
    if rc:
        raise RuntimeError("orc_demag_conv failed")
    return out


def demag_brute(n, p, K9, m, msat_map=None):
    m = _f64(m)
    out = np.zeros((3,) + tuple(n))
    K9 = np.ascontiguousarray(K9, dtype=np.float32)
    lib().orc_demag_brute(_i3(n), _i3(p), _ptr(K9, C.c_float), _ptr(m), _ptr(_f64(msat_map)), _ptr(out))
    return out


def exchange6(n, pbc, cell, pre, m, msat_map=None, aex_map=None):
    h = np.zeros((3,) + tuple(n))
    lib().orc_exchange6(_i3(n), _i3(pbc), _d3(cell), float(pre), _ptr(_f64(msat_map)),
                        _ptr(_f64(aex_map)), _ptr(_f64(m)), _ptr(h))
    return h


def uniaxial(m, ku2, u, ku_map=None, msat_map=None, u_map=None):
    m = _f64(m)
    N = m.size // 3
    h = np.zeros_like(m)
    lib().orc_uniaxial(N, float(ku2), _ptr(_f64(ku_map)), _ptr(_f64(msat_map)), _d3(u),
                       _ptr(_f64(u_map)), _ptr(m), _ptr(h))
    return h


def torque(m, h, alpha, alpha_map=None):
    m = _f64(m)
    h = _f64(h)
    N = m.size // 3
    t = np.zeros_like(m)
    lib().orc_torque(N, _ptr(m), _ptr(h), _ptr(_f64(alpha_map)), float(alpha), _ptr(t))
    return t


def normalize(m, norm_map=None):
    m = _f64(m).copy()
    lib().orc_normalize(m.size // 3, _ptr(m), _ptr(_f64(norm_map)))
    return m


def zhangli(n, pbc, cell, m, jmul, pred, xi, alpha, jmask=None, msat_map=None,
            pol_map=None, xi_map=None, alpha_map=None):
    out = np.zeros((3,) + tuple(n))
    lib().orc_zhangli(_i3(n), _i3(pbc), _d3(cell), _ptr(_f64(m)), _ptr(_f64(jmask)), _d3(jmul),
                      _ptr(_f64(msat_map)), _ptr(_f64(pol_map)), _ptr(_f64(xi_map)),
                      _ptr(_f64(alpha_map)), float(pred), float(xi), float(alpha), _ptr(out))
    return out


class _Problem(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3), ("pbc", C.c_int * 3), ("cell", C.c_double * 3),
        ("msat", C.c_double), ("msat_map", _dp),
        ("aex", C.c_double), ("aex_map", _dp),
        ("alpha", C.c_double), ("alpha_map", _dp),
        ("gamma", C.c_double),
        ("bext", C.c_double * 3),
        ("demag", C.c_int), ("p", C.c_int * 3), ("khat6", _dp),
        ("anis", C.c_int), ("ku", C.c_double), ("ku_map", _dp), ("anis_u", C.c_double * 3),
        ("zhangli", C.c_int), ("jmul", C.c_double * 3), ("jmask", _dp), ("pol", C.c_double),
        ("xi", C.c_double),
        ("bext_map", _dp), ("u_map", _dp), ("pol_map", _dp), ("xi_map", _dp),
    ]


class _RK12(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("mindt", C.c_double), ("maxdt", C.c_double),
        ("maxerr", C.c_double), ("err", C.c_double), ("peakerr", C.c_double),
        ("step", C.c_long), ("badsteps", C.c_long), ("m_normalized", C.c_int), ("evals", C.c_int),
    ]


class Problem:
    """Micromagnetic problem on the oracle side (internal z,y,x order).

    Mirrors what load('micromagnetism') (+ 'anisotropy/uniaxial', 'zhang-li')
    sets up in the reference: demag (Kernel_Arne, accuracy 6), exchange6,
    zeeman, llg and solver/rk12 with the micromagnetism defaults.
    """

    def __init__(self, n, cell, pbc=(0, 0, 0), msat=8e5, aex=1.3e-11, alpha=0.02,
                 gamma=GAMMA0, bext=(0, 0, 0), demag=True, ku=None, anis_u=(1, 0, 0),
                 j=None, pol=1.0, xi=0.05, msat_map=None, aex_map=None, alpha_map=None,
                 ku_map=None, jmask=None, kernel=None, accuracy=6, bext_map=None,
                 anis_u_map=None, pol_map=None, xi_map=None):
        self.n = tuple(int(x) for x in n)
        self.N = int(np.prod(self.n))
        self.cell = tuple(float(c) for c in cell)
        self.pbc = tuple(int(x) for x in pbc)
        self.p = pad_size(self.n, self.pbc)
        self._keep = {}
        P = _Problem()
        P.n = _i3(self.n)
        P.pbc = _i3(self.pbc)
        P.cell = _d3(self.cell)
        P.msat = msat
        P.aex = aex
        P.alpha = alpha
        P.gamma = gamma
        P.bext = _d3(bext)
        P.msat_map = self._map("msat_map", msat_map)
        P.aex_map = self._map("aex_map", aex_map)
        P.alpha_map = self._map("alpha_map", alpha_map)
        P.ku_map = self._map("ku_map", ku_map)
        P.jmask = self._map("jmask", jmask)
        P.bext_map = self._map("bext_map", bext_map)
        P.u_map = self._map("u_map", anis_u_map)
        P.pol_map = self._map("pol_map", pol_map)
        P.xi_map = self._map("xi_map", xi_map)
        P.demag = 1 if demag else 0
        P.p = _i3(self.p)
        if demag:
            if kernel is None:
                K9 = kernel_arne(self.p, self.cell, self.pbc, accuracy)
                kernel, _ = kernel_fft(self.p, K9=K9)
            elif isinstance(kernel, str) and kernel == "newell":
                kernel, _ = newell_khat(self.p, self.cell, self.pbc)
            self.khat6 = np.ascontiguousarray(kernel, dtype=np.float64)
            P.khat6 = _ptr(self.khat6)
        if ku is not None:
            P.anis = 1
            P.ku = ku
            P.anis_u = _d3(anis_u)
        if j is not None:
            P.zhangli = 1
            P.jmul = _d3(j)
            P.pol = pol
            P.xi = xi
        self.P = P
        self.S = _RK12()
        lib().orc_rk12_init(C.byref(self.S))

    def _map(self, name, a):
        if a is None:
            return None
        a = _f64(a)
        self._keep[name] = a
        return _ptr(a)

    def set_param(self, name, value):
        if name in ("bext", "jmul", "anis_u"):
            setattr(self.P, name, _d3(value))
        else:
            setattr(self.P, name, value)

    def heff(self, m, terms=False):
        m = _f64(m)
        h = np.zeros_like(m)
        if terms:
            hd, he, ha = np.zeros_like(m), np.zeros_like(m), np.zeros_like(m)
            rc = lib().orc_heff(C.byref(self.P), _ptr(m), _ptr(h), _ptr(hd), _ptr(he), _ptr(ha))
            if rc:
                raise RuntimeError("orc_heff failed")
            return h, hd, he, ha
        rc = lib().orc_heff(C.byref(self.P), _ptr(m), _ptr(h), None, None, None)
        if rc:
            raise RuntimeError("orc_heff failed")
        return h

    def torque(self, m, h):
        t = np.zeros_like(_f64(m))
        lib().orc_torque_full(C.byref(self.P), _ptr(_f64(m)), _ptr(_f64(h)), _ptr(t))
        return t

    # solver/rk12 state, engine quantity names
    def get(self, name):
        return getattr(self.S, name)

    def set(self, name, value):
        setattr(self.S, name, value)

    def step(self, m):
        """One accepted RK12 step, m (3,n0,n1,n2) float64 updated in place."""
        assert m.dtype == np.float64 and m.flags.c_contiguous
        rc = lib().orc_rk12_step(C.byref(self.P), C.byref(self.S), _ptr(m))
        if rc:
            raise RuntimeError("rk12: cannot converge (rc=%d)" % rc)

    def try_step(self, m, dt):
        assert m.dtype == np.float64 and m.flags.c_contiguous
        err = C.c_double(0)
        rc = lib().orc_rk12_try(C.byref(self.P), C.byref(self.S), _ptr(m), float(dt), C.byref(err))
        if rc:
            raise RuntimeError("rk12 try failed")
        return err.value

    def normalized(self, m):
        """m as the reference's quantity m reads after Update()."""
        if not self.S.m_normalized:
            lib().orc_normalize(self.N, _ptr(m), self._keep.get("msat_map") is not None
                                and _ptr(self._keep["msat_map"]) or None)
            self.S.m_normalized = 1
        return m

    def average(self, a):
        a = _f64(a)
        out = (C.c_double * 3)()
        lib().orc_average(a.size // 3, self.N, _ptr(a), out)
        return np.array(list(out))
