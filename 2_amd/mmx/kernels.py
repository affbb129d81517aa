"""Drop-in kernel entry points on torch CUDA (HIP) tensors.

Each function calls the C entry point that replaces the named libmumax2
function; tensors are float32, contiguous, on the GPU, in internal (z,y,x)
order, vector fields shaped (3, n0, n1, n2).  A map argument of None is the
reference's NULL map (uniform).  The work is queued on torch's current stream.
"""
import torch

from ._lib import lib, check


def _p(t):
    if t is None:
        return None
    if t.dtype != torch.float32 or not t.is_contiguous() or t.device.type != "cuda":
        raise ValueError("expected a contiguous float32 GPU tensor")
    return t.data_ptr()


def _c(t, c):
    return None if t is None else _p(t[c])


def _stream():
    return torch.cuda.current_stream().cuda_stream


def normalize(m, norm_map=None):
    """normalizeAsync (normalize.cu:45-60), in place."""
    N = m[0].numel()
    check(lib().mmx_normalize(_c(m, 0), _c(m, 1), _c(m, 2), _p(norm_map), N, _stream()))
    return m


def torque(m, h, alpha, alpha_map=None, out=None):
    """torqueAsync (torque.cu:59-78): tau/gamma."""
    out = torch.empty_like(m) if out is None else out
    N = m[0].numel()
    check(lib().mmx_torque(_c(out, 0), _c(out, 1), _c(out, 2), _c(m, 0), _c(m, 1), _c(m, 2),
                           _c(h, 0), _c(h, 1), _c(h, 2), _p(alpha_map), float(alpha), N,
                           _stream()))
    return out


def exchange6(m, pre, cell, pbc=(0, 0, 0), msat_map=None, aex_map=None, out=None):
    """exchange6Async (exchange6.cu:145-163); pre = 2 Aex/(mu0 Ms)."""
    out = torch.empty_like(m) if out is None else out
    n = m.shape[1:]
    check(lib().mmx_exchange6(_c(out, 0), _c(out, 1), _c(out, 2), _c(m, 0), _c(m, 1), _c(m, 2),
                              _p(msat_map), _p(aex_map), float(pre), *[int(x) for x in n],
                              *[int(x) for x in pbc], *[float(c) for c in cell], _stream()))
    return out


def uniaxial_anisotropy(m, ku2, u, ku_map=None, msat_map=None, u_map=None, out=None):
    """uniaxialAnisotropyAsync (uniaxialanisotropy.cu:234-265); ku2 = 2Ku/(mu0 Ms)."""
    out = torch.empty_like(m) if out is None else out
    N = m[0].numel()
    check(lib().mmx_uniaxial_anisotropy(_c(out, 0), _c(out, 1), _c(out, 2), _c(m, 0), _c(m, 1),
                                        _c(m, 2), _p(ku_map), _p(msat_map), float(ku2),
                                        _c(u_map, 0), float(u[0]), _c(u_map, 1), float(u[1]),
                                        _c(u_map, 2), float(u[2]), N, _stream()))
    return out


def zhangli(m, j, pred, xi, alpha, cell, pbc=(0, 0, 0), j_map=None, msat_map=None,
            pol_map=None, xi_map=None, alpha_map=None, out=None):
    """zhangli_async (zhang-li_torque.cu:307-429); pred = P muB/(e Ms)."""
    out = torch.empty_like(m) if out is None else out
    n = m.shape[1:]
    check(lib().mmx_zhangli(_c(out, 0), _c(out, 1), _c(out, 2), _c(m, 0), _c(m, 1), _c(m, 2),
                            _c(j_map, 0), _c(j_map, 1), _c(j_map, 2), _p(msat_map), _p(pol_map),
                            _p(xi_map), _p(alpha_map), float(j[0]), float(j[1]), float(j[2]),
                            float(pred), float(xi), float(alpha), *[int(x) for x in n],
                            *[float(c) for c in cell], *[int(x) for x in pbc], _stream()))
    return out


def vec_madd(a, b, mul, out=None):
    """vecMaddAsync (add.cu:119-150): out_c = a_c + mul_c b_c (b None: + mul_c)."""
    out = torch.empty_like(a) if out is None else out
    N = a[0].numel()
    check(lib().mmx_vec_madd(_c(out, 0), _c(out, 1), _c(out, 2), _c(a, 0), _c(a, 1), _c(a, 2),
                             _c(b, 0), _c(b, 1), _c(b, 2), float(mul[0]), float(mul[1]),
                             float(mul[2]), N, _stream()))
    return out


def madd(a, b, mul, out=None):
    """maddAsync (add.cu:102-118): out = a + mul*b (b None: a + mul)."""
    out = torch.empty_like(a) if out is None else out
    check(lib().mmx_madd(_p(out), _p(a), _p(b), float(mul), a.numel(), _stream()))
    return out


def madd1(a, b, mulb):
    """madd1Async (add.cu:169-177): a += mulb*b, in place."""
    check(lib().mmx_madd1(_p(a), _p(b), float(mulb), a.numel(), _stream()))
    return a


def madd2(a, b, mulb, c, mulc):
    """madd2Async (add.cu:191-199): a += mulb*b + mulc*c, in place."""
    check(lib().mmx_madd2(_p(a), _p(b), float(mulb), _p(c), float(mulc), a.numel(), _stream()))
    return a


def mul(a, b, out=None):
    """mulAsync: out = a*b."""
    out = torch.empty_like(a) if out is None else out
    check(lib().mmx_mul(_p(out), _p(a), _p(b), a.numel(), _stream()))
    return out
