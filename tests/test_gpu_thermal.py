"""temperature/brown on the GPU, through the C ABI, against oracle/thermal.py.

* the library's Philox4x32-10 block function against the Random123
  known-answer vectors;
* H_therm of a draw (Temp, alpha and Msat maps, a zero-Msat cell) equals the
  oracle's field of the same (seed, draw), to float32 rounding of the FP64
  Box-Muller step;
* solver/heun with the thermal field replayed by the oracle (draw k in step
  k, the field held over the step): m after every step to 1e-5;
* a new Therm_seed restarts the draws; Temp = 0 removes the field;
* the same Heun replay on the register-FFT x pass (its two-cell epilogue
  adds the field), and with maps against the generic passes.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from oracle import thermal as T  # noqa: E402

CELL = (3e-9, 4e-9, 4e-9)
N3 = (2, 16, 32)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def mmx_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return mmx


# Random123 kat_vectors for philox4x32 with 10 rounds
KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


def test_library_philox_known_answers(mmx_gpu):
    f = mmx_gpu.lib().mmx_philox4x32_10
    for ctr, key, want in KAT:
        c = (C.c_uint * 4)(*ctr)
        k = (C.c_uint * 2)(*key)
        out = (C.c_uint * 4)()
        assert f(c, k, out) == 0
        assert tuple(out) == want


def _maps():
    rng = np.random.default_rng(2)
    tm = rng.uniform(0.5, 1.5, N3).astype(np.float32)
    am = rng.uniform(0.5, 2.0, N3).astype(np.float32)
    mm = np.ones(N3, np.float32)
    mm[0, 3, 5] = 0.0
    mm[1, 7, :4] = 0.5
    return tm, am, mm


def _sim(mmx, maps=True, temp=300.0, seed=7, alpha=0.1):
    s = mmx.Sim(N3, CELL)
    s.setv("Msat", 8e5)
    s.setv("Aex", 1.3e-11)
    s.setv("alpha", alpha)
    s.load("temperature/brown")
    s.setv("Temp", temp)
    s.setv("Therm_seed", seed)
    if maps:
        tm, am, mm = _maps()
        s.setmask("Temp", tm)
        s.setmask("alpha", am)
        s.setmask("Msat", mm)
    return s


def test_field_matches_oracle(mmx_gpu):
    _say("H_therm draw 0", N3)
    s = _sim(mmx_gpu)
    s.setv("dt", 2e-14)
    rng = np.random.default_rng(5)
    s.setarray("m", O.normalize(rng.normal(size=(3,) + N3)).astype(np.float32))
    got = s.getarray("H_therm")
    tm, am, mm = _maps()
    want = T.htherm(N3, CELL, 7, 0, 300.0, 0.1, 8e5, O.GAMMA0, 2e-14, temp_map=tm,
                    alpha_map=am, msat_map=mm)
    scale = np.abs(want).max()
    _say("max |dH| / max|H| = %.2e, std %.4e" % (np.abs(got - want).max() / scale, want.std()))
    assert np.abs(got - want).max() <= 1e-6 * scale
    assert np.all(got[:, 0, 3, 5] == 0)
    # no redraw within the step
    assert np.array_equal(s.getarray("H_therm"), got)
    # a new seed restarts the draws
    s.setv("Therm_seed", 9)
    want9 = T.htherm(N3, CELL, 9, 0, 300.0, 0.1, 8e5, O.GAMMA0, 2e-14, temp_map=tm,
                     alpha_map=am, msat_map=mm)
    assert np.abs(s.getarray("H_therm") - want9).max() <= 1e-6 * scale
    # Temp = 0: no field
    s.setv("Temp", 0.0)
    assert not np.any(s.getarray("H_therm"))
    s.close()


def test_heun_with_thermal_field_against_oracle(mmx_gpu):
    _say("heun + H_therm", N3)
    dt = 1e-14
    s = _sim(mmx_gpu, maps=False, seed=3)
    s.load("solver/heun")
    s.setv("dt", dt)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    rng = np.random.default_rng(6)
    m = O.normalize(rng.normal(size=(3,) + N3))
    s.setarray("m", m.astype(np.float32))
    prob = O.Problem(N3, CELL, msat=8e5, aex=1.3e-11, alpha=0.1, bext=(0.0, 0.004, 0.009))
    mo = m.astype(np.float32).astype(np.float64)
    for k in range(6):
        s.step()
        ht = T.htherm(N3, CELL, 3, k, 300.0, 0.1, 8e5, O.GAMMA0, dt).astype(np.float64)
        mo = T.heun_step(prob, mo, dt, ht, O.GAMMA0)
        got = s.getarray("m")
        err = np.abs(got - mo).max()
        _say("step %d max |dm| %.2e" % (k, err))
        assert err < 1e-5
    # the field of the next step is draw 6
    want = T.htherm(N3, CELL, 3, 6, 300.0, 0.1, 8e5, O.GAMMA0, dt)
    assert np.abs(s.getarray("H_therm") - want).max() <= 1e-6 * np.abs(want).max()
    s.close()


@pytest.fixture
def register_x_passes(mmx_gpu):
    """The register-FFT x passes at this grid size (by default grids of fewer
    than 256 x-pass workgroups take the generic passes)."""
    lib = mmx_gpu.lib()
    prev = C.c_int()
    mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(0, C.byref(prev)))
    yield
    mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(prev.value, None))


def test_heun_thermal_on_register_x_pass(mmx_gpu, register_x_passes):
    """The thermal field in the register x pass's two-cell epilogue: heun
    against the oracle as above, and with Temp/alpha/Msat maps against the
    generic passes' per-cell epilogue on the same GPU."""
    _say("heun + H_therm, register x pass", N3)
    dt = 1e-14
    s = _sim(mmx_gpu, maps=False, seed=3)
    s.load("solver/heun")
    s.setv("dt", dt)
    s.setv("B_ext", (0.0, 0.004, 0.009))
    rng = np.random.default_rng(6)
    m = O.normalize(rng.normal(size=(3,) + N3))
    s.setarray("m", m.astype(np.float32))
    prob = O.Problem(N3, CELL, msat=8e5, aex=1.3e-11, alpha=0.1, bext=(0.0, 0.004, 0.009))
    mo = m.astype(np.float32).astype(np.float64)
    for k in range(6):
        s.step()
        ht = T.htherm(N3, CELL, 3, k, 300.0, 0.1, 8e5, O.GAMMA0, dt).astype(np.float64)
        mo = T.heun_step(prob, mo, dt, ht, O.GAMMA0)
        err = np.abs(s.getarray("m") - mo).max()
        _say("step %d max |dm| %.2e" % (k, err))
        assert err < 1e-5
    s.close()
    # with maps: register x pass against the generic passes
    lib = mmx_gpu.lib()
    out = []
    for threshold in (0, 1 << 30):
        mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(threshold, None))
        s = _sim(mmx_gpu, maps=True, seed=4)
        s.load("solver/heun")
        s.setv("dt", dt)
        s.setv("B_ext", (0.0, 0.004, 0.009))
        s.setarray("m", m.astype(np.float32))
        s.steps(6)
        out.append((s.getarray("m"), s.getarray("H_eff")))
        s.close()
    mmx_gpu.check(lib.mmx_set_fast_x_min_blocks(0, None))
    dm = np.abs(out[0][0] - out[1][0]).max()
    dh = np.abs(out[0][1] - out[1][1]).max() / np.abs(out[1][1]).max()
    _say("maps: register vs generic |dm| %.2e, H_eff %.2e" % (dm, dh))
    assert dm < 1e-5 and dh < 1e-5


def test_temperature_changes_the_trajectory_only_through_the_field(mmx_gpu):
    """Temp = 0 with the module loaded follows the athermal solver exactly."""
    _say("Temp 0")
    a = _sim(mmx_gpu, maps=False, temp=0.0)
    b = mmx_gpu.Sim(N3, CELL)
    b.setv("Msat", 8e5)
    b.setv("Aex", 1.3e-11)
    b.setv("alpha", 0.1)
    m = O.normalize(np.random.default_rng(8).normal(size=(3,) + N3)).astype(np.float32)
    for s in (a, b):
        s.setarray("m", m)
        s.steps(5)
    assert np.array_equal(a.getarray("m"), b.getarray("m"))
    a.close()
    b.close()


def test_reference_temperature_script(mmx_gpu):
    """tests/temperature.py of the reference through the user API: 32x32x1,
    micromagnetism + solver/rk12 + temperature/brown, Temp 100 K masked to
    the upper half in y (a 1x2x1 mask, resampled), 1 ns.  The field is zero
    where the mask is zero, and elsewhere its spread is the Brown variance at
    the current dt."""
    from mmx import api
    _say("reference temperature.py")
    e = api.reset()
    e.setgridsize(32, 32, 1)
    e.setcellsize(5e-9, 5e-9, 5e-9)
    e.load("micromagnetism")
    e.load("solver/rk12")
    e.load("temperature/brown")
    e.setv("Msat", 800e3)
    e.setv("Aex", 1.3e-11)
    e.setv("alpha", 0.01)
    e.setv("dt", 1e-12)
    e.setv("m_maxerror", 1.0 / 100)
    e.setv("temp", 100)
    e.setmask("temp", [[[[0], [1]]]])
    e.setarray("m", [[[[1]]], [[[1]]], [[[0]]]])
    tab = e.autotabulate(["t", "<m>"], period=10e-12)
    e.run(1e-9)
    assert len(tab.rows) >= 100
    h = e.getarray("H_therm")                      # user order [c][x][y][z]
    assert not np.any(h[:, :, :16, :])
    hot = h[:, :, 16:, :].ravel()
    dt = e.getv("dt")
    sigma = np.sqrt(2 * 0.01 * T.KB * 100 / (T.MU0 * 125e-27 * O.GAMMA0 * dt * 800e3))
    _say("steps %d, dt %.3e, std %.4e vs %.4e, <m> %s" % (e.getv("step"), dt, hot.std(), sigma,
                                                        np.round(e.getv("<m>"), 4)))
    assert abs(hot.std() / sigma - 1) < 0.05
    assert abs(hot.mean()) < 0.1 * sigma
