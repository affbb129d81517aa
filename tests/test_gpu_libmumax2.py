"""The libmumax2-shaped entry points (include/mmx_libmumax2.h) and the ConvPlan
seam (mmx_convplan_*), called through the C ABI as the cgo wrappers would:
every array a float** (an array of one device pointer), streams a
hipStream_t*, the reference's argument order.

* the entry points that forward to the mmx_* kernels give the same bits as
  those kernels, with maps given as float** and as NULL in both forms;
* the FFT-side helpers (kernelMulMicromag3D, cmadd, copyPad3D, copyUnPad3D,
  zeroArray) and the Reductor's float partials against numpy;
* ConvPlan: a caller-built kernel (Kernel_Arne, the demagexch kernel) against
  the FP64 oracle's convolution with the same kernel, the identity kernel
  round trip (convplan.go:195-231 checks 1e-3), and the refusals.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from oracle import demagexch as OD  # noqa: E402

TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    ids = (C.c_int * 1)(0)
    mmx.lib().mmx_setUsedGPUs(ids, 1)
    return torch


class Args:
    """float** / hipStream_t* arguments (arrays of one element), kept alive."""

    def __init__(self, torch):
        self.torch = torch
        self.keep = []

    def pp(self, t):
        """float** of a tensor (None: a NULL float**)."""
        if t is None:
            return None
        a = (C.c_void_p * 1)(t.data_ptr())
        self.keep.append(a)
        return C.addressof(a)

    def null_part(self):
        """a float** whose element [0] is NULL (the Go side's empty Array)."""
        a = (C.c_void_p * 1)(None)
        self.keep.append(a)
        return C.addressof(a)

    def stream(self):
        a = (C.c_void_p * 1)(self.torch.cuda.current_stream().cuda_stream)
        self.keep.append(a)
        return C.addressof(a)


def _rand(torch, shape, seed, lo=-1.0, hi=1.0):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.uniform(lo, hi, size=shape).astype(np.float32)).cuda()


def _same(a, b):
    return np.array_equal(a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32))


def test_forwarding_entry_points_same_bits(torch_gpu):
    import mmx
    from mmx import kernels as K
    torch = torch_gpu
    L = mmx.lib()
    n = (3, 10, 17)
    N = int(np.prod(n))
    m = _rand(torch, (3,) + n, 1)
    m = m / torch.linalg.norm(m, dim=0, keepdim=True)
    h = _rand(torch, (3,) + n, 2, -1e5, 1e5)
    amap = _rand(torch, n, 3, 0.5, 1.5)
    A = Args(torch)
    st = A.stream()
    c = lambda t, i: A.pp(t[i])  # noqa: E731
    # torqueAsync, with the alpha map as a float** and as NULL (both forms)
    for amp, am in ((amap, A.pp(amap)), (None, None), (None, A.null_part())):
        out = torch.empty_like(m)
        L.mmx_torqueAsync(c(out, 0), c(out, 1), c(out, 2), c(m, 0), c(m, 1), c(m, 2), c(h, 0),
                          c(h, 1), c(h, 2), am, 0.1, st, N)
        want = K.torque(m, h, 0.1, alpha_map=amp)
        torch.cuda.synchronize()
        assert _same(out, want)
    # normalizeAsync
    x = m * 1.7
    y = x.clone()
    L.mmx_normalizeAsync(c(x, 0), c(x, 1), c(x, 2), A.null_part(), st, N)
    K.normalize(y)
    torch.cuda.synchronize()
    assert _same(x, y)
    # exchange6Async with maps and periodic y
    ms = _rand(torch, n, 4, 0.5, 1.0)
    ax = _rand(torch, n, 5, 0.5, 1.0)
    out = torch.empty_like(m)
    cell = (3e-9, 4e-9, 5e-9)
    L.mmx_exchange6Async(c(out, 0), c(out, 1), c(out, 2), c(m, 0), c(m, 1), c(m, 2), A.pp(ms),
                         A.pp(ax), 2.5e-11, n[0], n[1], n[2], 0, 1, 0, *cell, st)
    want = K.exchange6(m, 2.5e-11, cell, (0, 1, 0), msat_map=ms, aex_map=ax)
    torch.cuda.synchronize()
    assert _same(out, want)
    # uniaxialAnisotropyAsync
    out = torch.empty_like(m)
    L.mmx_uniaxialAnisotropyAsync(c(out, 0), c(out, 1), c(out, 2), c(m, 0), c(m, 1), c(m, 2),
                                  A.pp(amap), None, 1e5, None, 0.6, A.null_part(), 0.8, None, 0.0,
                                  st, N)
    want = K.uniaxial_anisotropy(m, 1e5, (0.6, 0.8, 0.0), ku_map=amap)
    torch.cuda.synchronize()
    assert _same(out, want)
    # zhangli_async
    out = torch.empty_like(m)
    L.mmx_zhangli_async(c(out, 0), c(out, 1), c(out, 2), c(m, 0), c(m, 1), c(m, 2), None, None,
                        None, A.pp(ms), None, None, None, 0.0, 0.0, 1e12, 3e-11, 0.05, 0.02,
                        n[0], n[1], n[2], *cell, 0, 0, 1, st)
    want = K.zhangli(m, (0.0, 0.0, 1e12), 3e-11, 0.05, 0.02, cell, (0, 0, 1), msat_map=ms)
    torch.cuda.synchronize()
    assert _same(out, want)
    # the madd family and mul
    a, b, d = _rand(torch, (N,), 6), _rand(torch, (N,), 7), _rand(torch, (N,), 8)
    out = torch.empty_like(a)
    L.mmx_maddAsync(A.pp(out), A.pp(a), A.pp(b), 0.3, st, N)
    torch.cuda.synchronize()
    assert _same(out, K.madd(a, b, 0.3))
    x, y = a.clone(), a.clone()
    L.mmx_madd1Async(A.pp(x), A.pp(b), -0.7, st, N)
    K.madd1(y, b, -0.7)
    torch.cuda.synchronize()
    assert _same(x, y)
    x, y = a.clone(), a.clone()
    L.mmx_madd2Async(A.pp(x), A.pp(b), 0.25, A.pp(d), -0.5, st, N)
    K.madd2(y, b, 0.25, d, -0.5)
    torch.cuda.synchronize()
    assert _same(x, y)
    out = torch.empty_like(a)
    L.mmx_mulAsync(A.pp(out), A.pp(a), A.pp(b), st, N)
    torch.cuda.synchronize()
    assert _same(out, K.mul(a, b))
    out = torch.empty_like(m)
    L.mmx_vecMaddAsync(c(out, 0), c(out, 1), c(out, 2), c(m, 0), c(m, 1), c(m, 2), c(h, 0),
                       c(h, 1), c(h, 2), 1.0, 2.0, 3.0, st, N)
    torch.cuda.synchronize()
    assert _same(out, K.vec_madd(m, h, (1.0, 2.0, 3.0)))


def test_fft_side_helpers(torch_gpu):
    import mmx
    torch = torch_gpu
    L = mmx.lib()
    A = Args(torch)
    st = A.stream()
    rng = np.random.default_rng(11)
    nc = 4099
    fm = [rng.normal(size=2 * nc).astype(np.float32) for _ in range(3)]
    kk = [rng.normal(size=nc).astype(np.float32) for _ in range(6)]
    dfm = [torch.from_numpy(f.copy()).cuda() for f in fm]
    dk = [torch.from_numpy(k).cuda() for k in kk]
    L.mmx_kernelMulMicromag3DAsync(*[A.pp(t) for t in dfm], *[A.pp(t) for t in dk], st, 2 * nc)
    torch.cuda.synchronize()
    M = np.stack([f.astype(np.float64).reshape(nc, 2) for f in fm])   # (3, nc, 2)
    xx, yy, zz, yz, xz, xy = [k.astype(np.float64) for k in kk]
    T = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])           # (3, 3, nc)
    want = np.einsum("abn,bnr->anr", T, M)
    got = np.stack([t.cpu().numpy().astype(np.float64).reshape(nc, 2) for t in dfm])
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # cmadd: dst += (a + b i) k src
    dst0 = rng.normal(size=2 * nc).astype(np.float32)
    src = rng.normal(size=2 * nc).astype(np.float32)
    ddst, dsrc = torch.from_numpy(dst0.copy()).cuda(), torch.from_numpy(src).cuda()
    L.mmx_cmaddAsync(A.pp(ddst), 0.5, -2.0, A.pp(dk[0]), A.pp(dsrc), st, nc)
    torch.cuda.synchronize()
    z = (dst0[0::2] + 1j * dst0[1::2]) + (0.5 - 2.0j) * xx * (src[0::2] + 1j * src[1::2])
    got = ddst.cpu().numpy()
    assert np.abs(got[0::2] - z.real).max() <= 1e-5 and np.abs(got[1::2] - z.imag).max() <= 1e-5
    # copyPad3D / copyUnPad3D / zeroArray (2 components)
    S, D = (3, 5, 7), (6, 10, 14)
    a = rng.normal(size=(2,) + S).astype(np.float32)
    da = torch.from_numpy(a).cuda()
    dp = torch.full((2,) + D, 7.0, device="cuda")
    L.mmx_copyPad3DAsync(A.pp(dp), *D, A.pp(da), *S, 2, st)
    back = torch.empty_like(da)
    L.mmx_copyUnPad3DAsync(A.pp(back), *S, A.pp(dp), *D, 2, st)
    torch.cuda.synchronize()
    want = np.full((2,) + D, 7.0, np.float32)
    want[:, :S[0], :S[1], :S[2]] = a
    assert np.array_equal(dp.cpu().numpy(), want)       # padding region left as it was
    assert np.array_equal(back.cpu().numpy(), a)
    L.mmx_zeroArrayAsync(A.pp(dp), dp.numel(), st)
    torch.cuda.synchronize()
    assert not dp.abs().sum().item()


@pytest.mark.parametrize("n", [1, 1000, 262145, 3 * 1024 * 1024 + 17])
def test_partial_reductions(torch_gpu, n):
    import mmx
    torch = torch_gpu
    L = mmx.lib()
    A = Args(torch)
    st = A.stream()
    rng = np.random.default_rng(n % 997)
    h = [rng.normal(size=n).astype(np.float32) for _ in range(6)]
    d = [torch.from_numpy(x).cuda() for x in h]
    blocks = max(1, min(1024, (n + 1023) // 1024))
    part = torch.empty(blocks, device="cuda")
    threads = 512

    def run(fn, *ins):
        getattr(L, fn)(*[A.pp(t) for t in ins], A.pp(part), blocks, threads, n, st)
        torch.cuda.synchronize()
        return part.cpu().numpy().astype(np.float64)

    h64 = [x.astype(np.float64) for x in h]
    s = run("mmx_partialSumAsync", d[0]).sum()
    assert abs(s - h64[0].sum()) <= 1e-6 * max(1.0, np.abs(h64[0]).sum())
    s = run("mmx_partialSDotAsync", d[0], d[1]).sum()
    assert abs(s - (h64[0] * h64[1]).sum()) <= 1e-6 * max(1.0, np.abs(h64[0] * h64[1]).sum())
    assert run("mmx_partialMaxAsync", d[0]).max() == h[0].max()
    assert run("mmx_partialMinAsync", d[0]).min() == h[0].min()
    assert run("mmx_partialMaxAbsAsync", d[0]).max() == np.abs(h[0]).max()
    assert run("mmx_partialMaxDiffAsync", d[0], d[1]).max() == np.abs(h[0] - h[1]).max()
    assert run("mmx_partialMaxSumAsync", d[0], d[1]).max() == np.abs(h[0] + h[1]).max()
    # sums of three squares: the GPU contracts them into fused multiply-adds,
    # so they may differ from numpy's separately rounded float32 by an ulp
    sq = h[0] * h[0] + h[1] * h[1] + h[2] * h[2]
    got = run("mmx_partialMaxNorm3SqAsync", d[0], d[1], d[2]).max()
    assert abs(got - sq.max()) <= 4e-7 * sq.max()
    dx, dy, dz = h[0] - h[3], h[1] - h[4], h[2] - h[5]
    sq = dx * dx + dy * dy + dz * dz
    got = run("mmx_partialMaxNorm3SqDiffAsync", *d).max()
    assert abs(got - sq.max()) <= 4e-7 * sq.max()


# ----------------------------------------------------------------- ConvPlan

class ConvPlan:
    def __init__(self, n, p, K6):
        import mmx
        self.lib = mmx.lib()
        self.K = [None if k is None else np.ascontiguousarray(k, np.float32) for k in K6]
        ptrs = (C.c_void_p * 6)(*[None if k is None else k.ctypes.data for k in self.K])
        self.h = C.c_void_p()
        self.rc = self.lib.mmx_convplan_create(C.byref(self.h), (C.c_int * 3)(*n),
                                               (C.c_int * 3)(*p), ptrs)

    def convolve(self, torch, m):
        out = torch.empty_like(m)
        mmx_check(self.lib.mmx_convplan_convolve(self.h, m.data_ptr(), out.data_ptr(),
                                                 torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def close(self):
        if self.h:
            self.lib.mmx_convplan_destroy(self.h)
            self.h = C.c_void_p()


def mmx_check(rc):
    import mmx
    mmx.check(rc)


CONV_CASES = [((4, 16, 32), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9), (0, 0, 0)),
              ((6, 10, 12), (2e-9, 3e-9, 4e-9), (0, 1, 0)),
              ((8, 8, 24), (3e-9, 3e-9, 3e-9), (1, 0, 0))]


@pytest.mark.parametrize("n,cell,pbc", CONV_CASES)
def test_convplan_demag_and_demagexch_kernels(torch_gpu, n, cell, pbc):
    torch = torch_gpu
    p = O.pad_size(n, pbc)
    rng = np.random.default_rng(sum(n))
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    for name, K9 in (("Kernel_Arne", O.kernel_arne(p, cell, pbc)),
                     ("demagexch", OD.dex_kernel(p, cell, pbc, 8e5, 1.3e-11))):
        _say("convplan", name, n, pbc)
        cp = ConvPlan(n, p, K9[:6])
        mmx_check(cp.rc)
        got = cp.convolve(torch, dm)
        kh, _ = O.kernel_fft(p, K9=K9)
        want = O.demag_conv(n, p, kh, m.astype(np.float64))
        err = float(np.abs(got - want).max() / np.abs(want).max())
        _say("convplan", name, "rel err %.2e" % err)
        assert err < TOL
        bad = C.c_double(-1)
        mmx_check(cp.lib.mmx_convplan_info(cp.h, C.byref(bad)))
        assert 0.0 <= bad.value < 1e-5
        cp.close()


def test_convplan_identity_round_trip_and_refusals(torch_gpu):
    import mmx
    torch = torch_gpu
    n = (4, 12, 20)
    p = O.pad_size(n)
    delta = np.zeros(p, np.float32)
    delta[0, 0, 0] = 1.0
    cp = ConvPlan(n, p, [delta, delta, delta, None, None, None])
    mmx_check(cp.rc)
    rng = np.random.default_rng(3)
    a = (2 * rng.uniform(size=(3,) + n) - 1).astype(np.float32)
    got = cp.convolve(torch, torch.from_numpy(a).cuda())
    assert np.abs(got - a).max() < 1e-6   # convplan.go:195-231 accepts 1e-3
    cp.close()
    # a kernel without the mirror symmetry of a demag kernel
    skew = np.zeros(p, np.float32)
    skew[0, 1, 1] = 1.0   # not even in y at fixed kx
    bad = ConvPlan(n, p, [skew, delta, delta, None, None, None])
    assert bad.rc == 1 and b"mirror symmetry" in mmx.lib().mmx_last_error()
    # a kernel size that is neither the data size nor twice it
    odd = ConvPlan(n, (p[0], p[1], p[2] + 2), [np.zeros((p[0], p[1], p[2] + 2), np.float32)] * 3
                   + [None] * 3)
    assert odd.rc == 2
