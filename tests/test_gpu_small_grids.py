"""Small grids on the default pass choice: the x passes of grids with fewer
than 256 register-FFT workgroups run the generic LDS kernels (faster there),
with the register y/z passes where their lengths allow.  H_eff and the torque
against the FP64 oracle (1e-5 of the field's maximum), same-dt RK12 tries
(m to 1e-5), Zhang-Li on a film, and the two x-pass choices against each
other on one grid."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def mmx_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return mmx


def relerr(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def _pair(mmx, n, cell, zl=False, seed=1):
    m = O.normalize(np.random.default_rng(seed).normal(size=(3,) + n)).astype(np.float32)
    s = mmx.Sim(n, cell)
    s.setv("Msat", 8e5)
    s.setv("Aex", 1.3e-11)
    s.setv("alpha", 0.02)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    kw = dict(msat=8e5, aex=1.3e-11, alpha=0.02, bext=(0.0, 0.004, 0.009))
    if zl:
        s.load("zhang-li")
        s.setv("j", (0.0, 0.0, 1e12))
        s.setv("polarisation", 0.56)
        kw.update(j=(0.0, 0.0, 1e12), pol=0.56, xi=0.05)
    s.setarray("m", m)
    return s, O.Problem(n, cell, **kw), m.astype(np.float64)


def test_default_threshold(mmx_gpu):
    prev = C.c_int()
    mmx_gpu.check(mmx_gpu.lib().mmx_set_fast_x_min_blocks(256, C.byref(prev)))
    assert prev.value == 256
    with pytest.raises(mmx_gpu.MMXError):
        mmx_gpu.check(mmx_gpu.lib().mmx_set_fast_x_min_blocks(-1, None))


@pytest.mark.parametrize("n,zl", [((1, 32, 128), False), ((1, 32, 128), True),
                                  ((4, 64, 64), False), ((2, 96, 96), True),
                                  ((4, 128, 256), False)])
def test_heff_torque_small(mmx_gpu, n, zl):
    _say("small grid H_eff/torque", n, zl)
    cell = (3e-9, 3.90625e-9, 3.90625e-9)
    s, prob, m = _pair(mmx_gpu, n, cell, zl)
    mo = prob.normalized(m.copy())
    want = prob.heff(mo)
    eh = relerr(s.getarray("H_eff"), want)
    et = relerr(s.getarray("torque"), prob.torque(mo, want))
    _say("  H_eff %.2e torque %.2e" % (eh, et))
    assert eh < TOL and et < TOL
    s.close()


@pytest.mark.parametrize("n", [(1, 32, 128), (4, 64, 64)])
def test_rk12_tries_small(mmx_gpu, n):
    _say("small grid rk12 tries", n)
    cell = (3e-9, 3.90625e-9, 3.90625e-9)
    s, prob, m = _pair(mmx_gpu, n, cell, seed=4)
    mo = m.copy()
    for k in range(10):
        dt = 1e-14 * (1 + 0.1 * k)
        eg = s.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12, (k, eg, eo)
    err = np.abs(s.getarray("m") - prob.normalized(mo)).max()
    _say("  max |dm| %.2e" % err)
    assert err < TOL
    s.close()


def test_both_x_pass_choices_agree(mmx_gpu):
    """One grid through the generic x passes (default) and the register x
    passes (threshold 0): the same H_eff to FP32 rounding."""
    n, cell = (2, 64, 128), (3e-9, 4e-9, 4e-9)
    _say("x-pass choices", n)
    s, prob, m = _pair(mmx_gpu, n, cell, zl=True, seed=9)
    h_generic = s.getarray("H_eff").astype(np.float64)
    prev = C.c_int()
    mmx_gpu.check(mmx_gpu.lib().mmx_set_fast_x_min_blocks(0, C.byref(prev)))
    try:
        s.setv("alpha", 0.02)          # invalidates the cached field
        h_register = s.getarray("H_eff").astype(np.float64)
    finally:
        mmx_gpu.check(mmx_gpu.lib().mmx_set_fast_x_min_blocks(prev.value, None))
    e = relerr(h_register, h_generic)
    _say("  rel diff %.2e" % e)
    assert e < 1e-6
    s.close()
