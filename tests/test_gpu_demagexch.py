"""demagexch on the GPU, through the C ABI, against oracle/demagexch.py.

* H_dex (getarray) against the FP64 oracle conv(m, dex) on 3-D, film, periodic
  (in y, in z) and non-power-of-two grids, to 1e-5 of max |H_dex|; with Aex = 0 the field
  is the demag field of the accuracy-8 kernel alone, checked to the same bar;
* H_eff with demagexch loaded (micromagnetism's terms plus H_dex, evaluated as
  one convolution with the folded kernel) against the oracle Problem given the
  folded transform, and the torque with it;
* rk12 tries at fixed dt replayed by the oracle: m to 1e-5;
* a new Msat, Aex or demag_acc rebuilds the kernel; an Msat mask is refused;
* tests/rk12.py of the reference (standard problem 4 on 128x32x1 with
  micromagnetism + demagexch + solver/rk12) through the user API.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import demagexch as D  # noqa: E402
from oracle import oracle as O  # noqa: E402

TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


def relerr(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def mmx_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return mmx


def _sim(mmx, n, cell, pbc=(0, 0, 0), msat=8e5, aex=1.3e-11, seed=3):
    s = mmx.Sim(n, cell, pbc)
    s.setv("Msat", msat)
    s.setv("Aex", aex)
    s.setv("alpha", 0.05)
    s.load("demagexch")
    m = O.normalize(np.random.default_rng(seed).normal(size=(3,) + tuple(n))).astype(np.float32)
    s.setarray("m", m)
    return s, m.astype(np.float64)


HDEX_CASES = [((4, 16, 32), (2e-9, 3e-9, 5e-9), (0, 0, 0)),
              ((1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9), (0, 0, 0)),
              ((4, 16, 32), (3e-9, 3e-9, 3e-9), (0, 1, 0)),
              ((3, 10, 12), (2e-9, 3e-9, 4e-9), (0, 0, 0)),
              ((8, 8, 16), (3e-9, 2e-9, 2.5e-9), (1, 0, 0))]


@pytest.mark.parametrize("aex", [1.3e-11, 0.0])
@pytest.mark.parametrize("n,cell,pbc", HDEX_CASES)
def test_hdex_against_oracle(mmx_gpu, n, cell, pbc, aex):
    _say("H_dex", n, pbc, aex)
    s, m = _sim(mmx_gpu, n, cell, pbc, aex=aex)
    got = s.getarray("H_dex")
    want = D.h_dex(n, pbc, cell, 8e5, aex, m)
    e = relerr(got, want)
    _say("  rel err %.2e" % e)
    assert e < TOL
    s.close()


@pytest.mark.parametrize("n,cell,pbc", HDEX_CASES[:3])
def test_heff_with_demagexch(mmx_gpu, n, cell, pbc):
    _say("H_eff + H_dex", n, pbc)
    s, m = _sim(mmx_gpu, n, cell, pbc)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    prob = O.Problem(n, cell, pbc=pbc, msat=8e5, aex=1.3e-11, alpha=0.05,
                     bext=(0.0, 0.004, 0.009),
                     kernel=D.folded_khat(n, pbc, cell, 8e5, 1.3e-11))
    want = prob.heff(m)
    assert relerr(s.getarray("H_eff"), want) < TOL
    assert relerr(s.getarray("torque"), prob.torque(m, want)) < TOL
    # B is still the demag field alone
    p = O.pad_size(n, pbc)
    kd, _ = O.kernel_fft(p, K9=O.kernel_arne(p, cell, pbc, 6))
    assert relerr(s.getarray("B"), O.MU0 * 8e5 * O.demag_conv(n, p, kd, m)) < TOL
    s.close()


def test_rk12_replay_with_demagexch(mmx_gpu):
    n, cell = (1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    _say("rk12 replay + demagexch", n)
    s, m = _sim(mmx_gpu, n, cell)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    prob = O.Problem(n, cell, msat=8e5, aex=1.3e-11, alpha=0.05, bext=(0.0, 0.004, 0.009),
                     kernel=D.folded_khat(n, (0, 0, 0), cell, 8e5, 1.3e-11))
    mo = m.copy()
    for k in range(20):
        dt = 1e-14 * (1 + 0.1 * k)
        eg = s.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12
    err = np.abs(s.getarray("m") - prob.normalized(mo)).max()
    _say("  max |dm| %.2e" % err)
    assert err < 1e-5
    s.close()


def test_parameters_rebuild_the_kernel(mmx_gpu):
    n, cell, pbc = (2, 16, 32), (3e-9, 3e-9, 3e-9), (0, 0, 0)
    _say("demagexch rebuild")
    s, m = _sim(mmx_gpu, n, cell)
    assert s.getv("demag_acc") == 8
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 8e5, 1.3e-11, m)) < TOL
    s.setv("Msat", 6e5)
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 6e5, 1.3e-11, m)) < TOL
    s.setv("Aex", 2e-11)
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 6e5, 2e-11, m)) < TOL
    s.setv("demag_acc", 3)
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 6e5, 2e-11, m, accuracy=3)) < TOL
    with pytest.raises(Exception):
        s.setv("demag_acc", 2.5)
    # with an Msat mask H_dex takes its own convolution (the kernel keeps the
    # multiplier); a mask of ones gives the folded convolution's field
    h_folded = s.getarray("H_eff")
    s.setmask("Msat", np.ones(n, np.float32))
    h_masked = s.getarray("H_eff")
    assert relerr(h_masked, h_folded.astype(np.float64)) < TOL
    s.close()


def test_reference_rk12_script(mmx_gpu):
    """tests/rk12.py of the reference: 128x32x1 cells of 500/128 x 125/32 x 3
    nm, micromagnetism + demagexch + solver/rk12, Msat 800e3, Aex 1.3e-11,
    alpha 1, m_maxerror 1e-3, m = (1, 1, 0), relaxed for 2 ns, then the
    standard problem 4 field (-24.6, 4.3, 0) mT / mu0 at alpha 0.02 for 0.1 ns
    (the script runs 1 ns).  H_eff at the end against the oracle's field of
    the same m with the folded kernel."""
    from mmx import api
    _say("reference rk12.py")
    e = api.reset()
    e.setgridsize(128, 32, 1)
    e.setcellsize(500e-9 / 128, 125e-9 / 32, 3e-9)
    e.load("micromagnetism")
    e.load("demagexch")
    e.load("solver/rk12")
    e.setv("Msat", 800e3)
    e.setv("Aex", 1.3e-11)
    e.setv("alpha", 1)
    e.setv("dt", 1e-12)
    e.setv("m_maxerror", 1.0 / 1000)
    e.setarray("m", [[[[1]]], [[[1]]], [[[0]]]])
    tab = e.autotabulate(["t", "<m>", "m_error", "dt"], period=1e-12)
    e.run(2e-9)
    relaxed = np.array(e.getv("<m>"))
    _say("relaxed <m> %s after %d steps" % (np.round(relaxed, 4), e.getv("step")))
    assert len(tab.rows) >= 1900
    assert relaxed[0] > 0.9 and abs(relaxed[2]) < 1e-3
    e.setv("B_ext", (-24.6e-3, 4.3e-3, 0.0))
    e.setv("alpha", 0.02)
    e.setv("dt", 0.1e-12)
    e.run(0.1e-9)
    after = np.array(e.getv("<m>"))
    _say("<m> after 0.1 ns in the field %s" % np.round(after, 4))
    assert np.all(np.isfinite(after)) and after[0] < relaxed[0]
    # the field at the final state (internal order: z, y, x)
    s = e.sim
    n = (1, 32, 128)
    cell = (3e-9, 125e-9 / 32, 500e-9 / 128)
    m = s.getarray("m").astype(np.float64)
    prob = O.Problem(n, cell, msat=800e3, aex=1.3e-11, alpha=0.02, bext=(0.0, 4.3e-3, -24.6e-3),
                     kernel=D.folded_khat(n, (0, 0, 0), cell, 800e3, 1.3e-11))
    assert relerr(s.getarray("H_eff"), prob.heff(m)) < TOL


@pytest.mark.parametrize("register_x", [False, True])
def test_heff_with_demagexch_and_msat_mask(mmx_gpu, register_x):
    """With an Msat mask the reference keeps building the dex kernel from the
    Msat multiplier and convolves m itself (mod_demagexchange.go:74-79,143),
    while B takes Msat_map m: H_eff = H_demag(Msat_map m) + H_ex + H_dex(m).
    register_x: the register-FFT x passes at this size (their two-cell
    epilogue adds H_dex); otherwise the generic passes this grid gets by
    default."""
    import ctypes as C
    n, cell, pbc = (2, 16, 32), (3e-9, 4e-9, 4e-9), (0, 0, 0)
    _say("H_eff + H_dex with an Msat mask", n, "register x" if register_x else "default")
    lib = mmx_gpu.lib()
    prev = C.c_int()
    mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(0 if register_x else 256, C.byref(prev)))
    try:
        _heff_dex_msat_mask(mmx_gpu, n, cell, pbc)
    finally:
        mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(prev.value, None))


def _heff_dex_msat_mask(mmx_gpu, n, cell, pbc):
    s, m = _sim(mmx_gpu, n, cell, pbc, seed=11)
    ms = np.random.default_rng(12).uniform(0.5, 1.0, size=n).astype(np.float32)
    s.setmask("Msat", ms, 1)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    mo = O.normalize(m, ms)
    prob = O.Problem(n, cell, pbc=pbc, msat=8e5, aex=1.3e-11, alpha=0.05,
                     bext=(0.0, 0.004, 0.009), msat_map=ms)
    want = prob.heff(mo) + D.h_dex(n, pbc, cell, 8e5, 1.3e-11, mo)
    assert relerr(s.getarray("H_eff"), want) < TOL
    assert relerr(s.getarray("torque"), prob.torque(mo, want)) < TOL
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 8e5, 1.3e-11, mo)) < TOL
    # one rk12 try: the predictor's H_dex comes from its own convolution
    e = s.try_step(1e-14)
    assert np.isfinite(e) and e > 0
    s.close()


def test_demag_acc_zero(mmx_gpu):
    n, cell, pbc = (4, 16, 32), (2e-9, 3e-9, 5e-9), (0, 0, 0)
    s, m = _sim(mmx_gpu, n, cell, pbc)
    s.setv("demag_acc", 0)
    assert relerr(s.getarray("H_dex"), D.h_dex(n, pbc, cell, 8e5, 1.3e-11, m, accuracy=0)) < TOL
    s.close()
