"""Parity of the HIP path with the FP64 oracle, through the C ABI.

Tolerances: integer/index work and the Kernel_Arne construction are exact;
FP32 fields are judged by ||dH||_inf / ||H||_inf <= 1e-5 (the plain-path
tolerance); m(t) by same-dt replays and free-running <m> traces.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = 1e-5


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return torch


@pytest.fixture(scope="module", autouse=True)
def register_x_passes_everywhere():
    """The register-FFT x passes at every grid size (by default grids of
    fewer than 256 x-pass workgroups take the generic passes): these tests pin
    those kernels at small sizes; the default choice is covered by
    tests/test_gpu_small_grids.py and the other modules."""
    import ctypes as C
    import mmx
    prev = C.c_int()
    mmx.check(mmx.lib().mmx_set_fast_x_min_blocks(0, C.byref(prev)))
    yield
    mmx.check(mmx.lib().mmx_set_fast_x_min_blocks(prev.value, None))


def relerr(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------ kernel build

ARNE_CASES = [((4, 6, 8), (3e-9, 4e-9, 5e-9), (0, 0, 0)),
              ((1, 16, 32), (3e-9, 3.90625e-9, 3.90625e-9), (0, 0, 0)),
              ((6, 8, 10), (2e-9, 2e-9, 2e-9), (0, 0, 0)),
              ((2, 12, 6), (4e-9, 3e-9, 4e-9), (0, 0, 0))]

# periodic images: several displacements wrap onto one element and are summed
# in float32 in the reference's loop order (kernel_magnetostatic.go:168-172)
ARNE_PBC_CASES = [((4, 6, 8), (3e-9, 4e-9, 5e-9), (1, 0, 0)),
                  ((3, 8, 8), (2e-9, 2e-9, 2e-9), (0, 1, 0)),
                  ((2, 12, 6), (4e-9, 3e-9, 4e-9), (0, 0, 1)),
                  ((1, 16, 8), (3e-9, 3.90625e-9, 3.90625e-9), (0, 1, 1)),
                  ((2, 6, 10), (3e-9, 4e-9, 2e-9), (0, 2, 0)),
                  ((4, 4, 6), (3e-9, 3e-9, 3e-9), (1, 1, 1))]


@pytest.mark.parametrize("n,cell,pbc", ARNE_CASES + ARNE_PBC_CASES)
def test_kernel_arne_bit_exact(torch_gpu, n, cell, pbc):
    import mmx
    _say("arne", n, cell)
    plan = mmx.DemagPlan(n, cell, pbc, keep_real_kernel=True)
    Kg = plan.real_kernel()
    Ko = O.kernel_arne(plan.p, cell, pbc)[:6]
    assert np.array_equal(Kg, Ko), np.abs(Kg - Ko).max()
    if any(pbc):
        # the same bits on a second build (no order-dependent accumulation)
        again = mmx.DemagPlan(n, cell, pbc, keep_real_kernel=True).real_kernel()
        assert np.array_equal(again.view(np.uint32), Kg.view(np.uint32))


@pytest.mark.parametrize("n,cell,pbc", ARNE_CASES + [((3, 8, 8), (2e-9,) * 3, (0, 1, 0))])
def test_kernel_fft_octant(torch_gpu, n, cell, pbc):
    import mmx
    _say("kernel fft", n, pbc)
    plan = mmx.DemagPlan(n, cell, pbc, keep_real_kernel=True)
    p = plan.p
    K = plan.real_kernel()
    kh, bad = O.kernel_fft(p, K9=np.concatenate([K, np.zeros((3,) + p, np.float32)]))
    ko = plan.ko
    want = kh[:, :ko[0], :ko[1], :ko[2]]
    got = plan.kernel()
    assert relerr(got, want) < 1e-6
    assert plan.bad_over_good < 1e-5


# ----------------------------------------------------------- demag field

CONV_CASES = [((1, 32, 128), (3e-9, 3.90625e-9, 3.90625e-9), (0, 0, 0)),
              ((4, 16, 16), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((8, 16, 24), (4e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((3, 10, 14), (2e-9, 3e-9, 5e-9), (0, 0, 0)),
              ((2, 12, 20), (3e-9, 3e-9, 3e-9), (0, 1, 0)),
              ((4, 8, 16), (4e-9, 4e-9, 4e-9), (0, 0, 1)),
              ((1, 7, 9), (5e-9, 3e-9, 3e-9), (0, 0, 0)),
              # power-of-two transforms (register-FFT passes), incl. mixed paths
              ((16, 64, 64), (3e-9, 3e-9, 3e-9), (0, 0, 0)),
              ((1, 256, 512), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((32, 128, 256), (4e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((4, 48, 64), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((6, 32, 40), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((2, 16, 512), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((1, 1024, 8), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              # the longest transforms: y 2048 in 3-D, x half-lengths 1024 and 2048
              ((2, 1024, 24), (2e-9, 2e-9, 2e-9), (0, 0, 0)),
              # y 2048 on the lane-pair split: under a register z pass, and
              # periodic (every row holds data)
              ((8, 1024, 4), (2e-9, 2e-9, 3e-9), (0, 0, 0)),
              ((2, 2048, 4), (3e-9, 2e-9, 2e-9), (0, 1, 0)),
              ((2, 12, 1024), (2e-9, 2e-9, 2e-9), (0, 0, 0)),
              ((1, 8, 2048), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              # z transforms of length 128, 512, 2048 (zero-padded half), and
              # periodic z (every plane holds data) at lengths 8 and 64
              ((64, 8, 16), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((256, 4, 8), (2e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((1024, 2, 4), (2e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((8, 16, 16), (3e-9, 4e-9, 4e-9), (1, 0, 0)),
              ((64, 8, 8), (3e-9, 3e-9, 3e-9), (1, 0, 0)),
              # periodic y at register-FFT lengths: film (y2d) and 3-D (yf, yi)
              ((1, 64, 32), (3e-9, 4e-9, 4e-9), (0, 1, 0)),
              ((4, 32, 16), (3e-9, 4e-9, 4e-9), (0, 1, 0)),
              # periodic y of length 6 under a register-FFT z pass (tile pairs
              # ky, p1 - ky around p1/2 = 3)
              ((4, 6, 8), (3e-9, 4e-9, 4e-9), (0, 1, 0)),
              # prime factors above 7 (the direct-DFT Stockham stage): x half
              # lengths 13 and 31, y 22 and 34, z 6 with y 22, periodic y 34
              ((1, 11, 13), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((3, 11, 26), (3e-9, 4e-9, 4e-9), (0, 0, 0)),
              ((2, 17, 62), (3e-9, 3e-9, 3e-9), (0, 0, 0)),
              ((2, 34, 12), (3e-9, 3e-9, 3e-9), (0, 1, 0))]


@pytest.mark.parametrize("n,cell,pbc", CONV_CASES)
def test_demag_update_b(torch_gpu, n, cell, pbc):
    import mmx
    torch = torch_gpu
    _say("demag B", n, pbc)
    rng = np.random.default_rng(sum(n))
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    ms_map = rng.uniform(0.5, 1.0, size=n).astype(np.float32)
    plan = mmx.DemagPlan(n, cell, pbc)
    p = plan.p
    K9 = O.kernel_arne(p, cell, pbc)
    kh, _ = O.kernel_fft(p, K9=K9)
    for msat_map in (None, ms_map):
        want = O.MU0 * 8e5 * O.demag_conv(n, p, kh, m.astype(np.float64), msat_map)
        dm = torch.from_numpy(m).cuda()
        db = torch.empty_like(dm)
        dmap = None if msat_map is None else torch.from_numpy(msat_map).cuda()
        plan.update_b(db.data_ptr(), dm.data_ptr(), 8e5,
                      None if dmap is None else dmap.data_ptr())
        torch.cuda.synchronize()
        assert relerr(db.cpu().numpy(), want) < TOL


def test_demag_linearity_full_size(torch_gpu):
    """256x256x64 (the bulk configuration's grid): B(2a - 3b) = 2B(a) - 3B(b)."""
    import mmx
    torch = torch_gpu
    n = (64, 256, 256)
    _say("linearity", n)
    plan = mmx.DemagPlan(n, (4e-9,) * 3)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    a = torch.randn((3,) + n, device="cuda", generator=g)
    b = torch.randn((3,) + n, device="cuda", generator=g)
    c = 2 * a - 3 * b
    outs = []
    for x in (a, b, c):
        o = torch.empty_like(x)
        plan.update_b(o.data_ptr(), x.data_ptr(), 8e5)
        outs.append(o)
    torch.cuda.synchronize()
    lhs = outs[2]
    rhs = 2 * outs[0] - 3 * outs[1]
    err = (lhs - rhs).abs().max().item() / lhs.abs().max().item()
    assert err < 1e-5


# -------------------------------------------------------- drop-in kernels

def _t(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("pbc", [(0, 0, 0), (1, 0, 1)])
@pytest.mark.parametrize("maps", [False, True])
def test_exchange6(torch_gpu, pbc, maps):
    from mmx import kernels as K
    torch = torch_gpu
    n = (5, 9, 17)
    _say("exchange6", pbc, maps)
    rng = np.random.default_rng(11)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    cell = (3e-9, 4e-9, 5e-9)
    ms = rng.uniform(0.5, 1.5, n).astype(np.float32) if maps else None
    ax = rng.uniform(0.5, 1.5, n).astype(np.float32) if maps else None
    pre = 2 * 1.3e-11 / (O.MU0 * 8e5)
    cell_f = tuple(float(np.float32(c)) for c in cell)
    want = O.exchange6(n, pbc, cell_f, float(np.float32(pre)), m, ms, ax)
    got = K.exchange6(_t(torch, m), pre, cell, pbc,
                      None if ms is None else _t(torch, ms),
                      None if ax is None else _t(torch, ax)).cpu().numpy()
    assert relerr(got, want) < TOL


@pytest.mark.parametrize("n,pbc", [((5, 9, 17), (0, 0, 0)), ((3, 8, 128), (1, 1, 1)),
                                   ((1, 6, 70), (0, 1, 0)), ((2, 1, 65), (1, 0, 0)),
                                   ((4, 7, 1), (0, 0, 1)), ((2, 5, 33), (0, 0, 1))])
def test_exchange6_lds_path_is_bit_identical(torch_gpu, n, pbc):
    """Without maps mmx_exchange6 runs the LDS-staged stencil; with an Aex map
    of ones it runs the per-cell kernel whose weights are then exactly 1: the
    two must agree bit for bit (partial tiles, wrapped and clamped edges,
    length-1 axes)."""
    from mmx import kernels as K
    torch = torch_gpu
    _say("exchange6 LDS vs per-cell", n, pbc)
    m = O.normalize(np.random.default_rng(13).normal(size=(3,) + n)).astype(np.float32)
    cell, pre = (3e-9, 4e-9, 5e-9), 2 * 1.3e-11 / (O.MU0 * 8e5)
    a = K.exchange6(_t(torch, m), pre, cell, pbc).cpu().numpy()
    b = K.exchange6(_t(torch, m), pre, cell, pbc, None,
                    _t(torch, np.ones(n, np.float32))).cpu().numpy()
    assert np.array_equal(a, b)


def test_uniaxial_torque_normalize(torch_gpu):
    from mmx import kernels as K
    torch = torch_gpu
    n = (4, 8, 8)
    _say("uniaxial/torque/normalize")
    rng = np.random.default_rng(12)
    m = rng.normal(size=(3,) + n).astype(np.float32)
    mask = rng.uniform(0, 1, n).astype(np.float32)
    mask[0, 0, :3] = 0
    dm = _t(torch, m)
    K.normalize(dm, _t(torch, mask))
    want = O.normalize(m.astype(np.float64), mask)
    assert np.abs(dm.cpu().numpy() - want).max() < 1e-6
    mn = want.astype(np.float32)
    u = np.array([0.6, 0.0, 0.8])
    ku2 = 2 * 5e5 / (O.MU0 * 8e5)
    got = K.uniaxial_anisotropy(_t(torch, mn), ku2, u).cpu().numpy()
    assert relerr(got, O.uniaxial(mn, float(np.float32(ku2)), u.astype(np.float32))) < TOL
    h = rng.normal(size=(3,) + n).astype(np.float32) * 1e5
    got = K.torque(_t(torch, mn), _t(torch, h), 0.05, _t(torch, mask)).cpu().numpy()
    want = O.torque(mn, h, float(np.float32(0.05)), mask)
    assert relerr(got, want) < TOL


def test_zhangli_kernel(torch_gpu):
    from mmx import kernels as K
    torch = torch_gpu
    n = (2, 6, 20)
    _say("zhang-li")
    rng = np.random.default_rng(13)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    cell = (3e-9, 3.125e-9, 3.125e-9)
    pred = 0.56 * O.MUB / (O.QE * 8e5)
    j = (0.0, 0.3e12, 1e12)
    got = K.zhangli(_t(torch, m), j, pred, 0.05, 0.1, cell).cpu().numpy()
    want = O.zhangli(n, (0, 0, 0), tuple(float(np.float32(c)) for c in cell), m, j,
                     float(np.float32(pred)), float(np.float32(0.05)), float(np.float32(0.1)))
    assert relerr(got, want) < TOL


def test_madd_family_exact(torch_gpu):
    from mmx import kernels as K
    torch = torch_gpu
    rng = np.random.default_rng(14)
    a, b, c = (rng.normal(size=1000).astype(np.float32) for _ in range(3))
    ta, tb, tc = _t(torch, a), _t(torch, b), _t(torch, c)
    got = K.madd(ta, tb, 0.25).cpu().numpy()
    assert np.array_equal(got, a + np.float32(0.25) * b)
    x = ta.clone()
    K.madd2(x, tb, 0.5, tc, -0.5)
    want = a + (np.float32(0.5) * b + np.float32(-0.5) * c)
    assert np.abs(x.cpu().numpy() - want).max() <= 2e-7 * np.abs(want).max()
    v = _t(torch, rng.normal(size=(3, 100)))
    w = _t(torch, rng.normal(size=(3, 100)))
    got = K.vec_madd(v, w, (1.0, 2.0, 3.0)).cpu().numpy()
    vv, ww = v.cpu().numpy(), w.cpu().numpy()
    # a fused multiply-add rounds once: within 1 ulp of the two-rounding value
    want = vv.astype(np.float64) + np.array([1, 2, 3])[:, None] * ww
    assert np.abs(got - want).max() <= 1.2e-7 * np.abs(want).max()
    got = K.vec_madd(v, None, (1.0, 2.0, 3.0)).cpu().numpy()
    assert np.array_equal(got, vv + np.float32([1, 2, 3])[:, None])


def test_reductor(torch_gpu):
    import mmx
    torch = torch_gpu
    rng = np.random.default_rng(15)
    for n in (1, 7, 1000, 300001):
        a = rng.normal(size=n).astype(np.float32)
        b = rng.normal(size=n).astype(np.float32)
        ta, tb = _t(torch, a), _t(torch, b)
        r = mmx.Reductor(n)
        assert abs(r.sum(ta.data_ptr(), n) - a.astype(np.float64).sum()) < 1e-9 * max(1, np.abs(a).sum())
        assert abs(r.dot(ta.data_ptr(), tb.data_ptr(), n) - (a.astype(np.float64) * b).sum()) < 1e-9 * n
        assert r.max(ta.data_ptr(), n) == a.max()
        assert r.min(ta.data_ptr(), n) == a.min()
        assert r.maxabs(ta.data_ptr(), n) == np.abs(a).max()
        assert r.maxdiff(ta.data_ptr(), tb.data_ptr(), n) == np.abs(a - b).max()


# ------------------------------------------------------------ H_eff, RK12

def _problem_pair(n, cell, anis=False, zl=False, bext=(0.0, 0.0, 0.0), seed=21, alpha=0.02):
    import mmx
    rng = np.random.default_rng(seed)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    sim = mmx.Sim(n, cell)
    sim.setv("Msat", 8e5)
    sim.setv("Aex", 1.3e-11)
    sim.setv("alpha", alpha)
    sim.setv("B_ext", bext)
    kw = {}
    if anis:
        sim.load("anisotropy/uniaxial")
        sim.setv("Ku", 5e5)
        sim.setv("anisU", (1.0, 0.0, 0.0))
        kw.update(ku=5e5, anis_u=(1.0, 0.0, 0.0))
    if zl:
        sim.load("zhang-li")
        sim.setv("j", (0.0, 0.0, 1e12))
        sim.setv("polarisation", 0.56)
        sim.setv("xi", 0.05)
        kw.update(j=(0.0, 0.0, 1e12), pol=0.56, xi=0.05)
    sim.setarray("m", m)
    prob = O.Problem(n, cell, msat=8e5, aex=1.3e-11, alpha=alpha, bext=bext, **kw)
    return sim, prob, m


HEFF_CASES = [((1, 32, 128), (3e-9, 3.90625e-9, 3.90625e-9), False, False),
              ((8, 32, 32), (4e-9, 4e-9, 4e-9), True, False),
              ((2, 16, 1024), (2e-9, 2e-9, 2e-9), True, False),
              ((2, 16, 48), (3e-9, 4e-9, 4e-9), False, True)]


@pytest.mark.parametrize("n,cell,anis,zl", HEFF_CASES)
def test_heff_and_torque(torch_gpu, n, cell, anis, zl):
    _say("heff", n, anis, zl)
    sim, prob, m = _problem_pair(n, cell, anis, zl, bext=(0.0, 0.004, 0.009))
    mo = prob.normalized(m.astype(np.float64).copy())
    want_h = prob.heff(mo)
    got_h = sim.getarray("H_eff")
    assert relerr(got_h, want_h) < TOL
    want_t = prob.torque(mo, want_h)
    got_t = sim.getarray("torque")
    assert relerr(got_t, want_t) < TOL


def test_rk12_same_dt_replay(torch_gpu):
    n, cell = (1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    _say("rk12 replay", n)
    sim, prob, m = _problem_pair(n, cell, bext=(0.0, 0.004, 0.009))
    mo = m.astype(np.float64).copy()
    for k in range(20):
        dt = 2e-14 * (1 + 0.1 * k)
        eg = sim.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12
    got = sim.getarray("m")
    want = prob.normalized(mo)
    assert np.abs(got - want).max() < 1e-5


def test_rk12_free_running_cfg1(torch_gpu):
    """The film configuration (128x32x1, seeded random m, 10 mT in-plane field),
    adaptive rk12 at m_maxerror 1e-5 (dt ~ 2 fs): <m>(t) against the oracle
    for 3 ps (~1700 steps)."""
    import mmx
    from mmx import synth
    cfg = synth.config(1)
    _say("rk12 free run cfg1")
    sim = mmx.Sim(cfg["n"], cfg["cell"])
    synth.setup_sim(sim, cfg)
    prob = O.Problem(cfg["n"], cfg["cell"], msat=cfg["msat"], aex=cfg["aex"], alpha=cfg["alpha"],
                     bext=cfg["bext"])
    prob.set("maxerr", cfg["maxerr"])
    mo = cfg["m"].astype(np.float64).copy()
    for t_check in (1e-12, 2e-12, 3e-12):
        while sim.getv("t") < t_check:
            sim.step()
        while prob.get("t") < t_check:
            prob.step(mo)
        g = np.array(sim.getv("<m>"))
        o = prob.average(prob.normalized(mo))
        # both stop at the first step past t_check (dt ~ 2 fs, <m> moves
        # ~1e-4 per fs at most here)
        assert np.abs(g - o).max() < 2e-4, (t_check, g, o)


def test_maxtorque_is_the_max_norm(torch_gpu):
    """tests/reduce.py: maxtorque (MaxNormUpdater of the torque) equals the
    largest |torque| over the cells, and is at least the largest component."""
    import mmx
    n, cell = (1, 32, 32), (3e-9, 500e-9 / 32 / 4, 500e-9 / 32 / 4)
    _say("maxtorque")
    sim = mmx.Sim(n, cell)
    sim.setv("Msat", 800e3)
    sim.setv("Aex", 1.3e-11)
    sim.setv("alpha", 1.0)
    sim.setarray("m", O.normalize(np.random.default_rng(9).normal(size=(3,) + n)).astype(np.float32))
    mt = sim.getv("maxtorque")
    tq = sim.getarray("torque").astype(np.float64) * sim.getv("gamma")
    norm = np.sqrt((tq ** 2).sum(axis=0)).max()
    assert abs(mt - norm) <= 1e-6 * norm
    assert mt >= np.abs(tq).max()


def test_rk12_look_ahead_changes_nothing(torch_gpu):
    """solver/rk12 computes the next step's first stage at the corrected m while
    the host decides on the step; a run where that look-ahead is discarded
    before every step (device_array invalidates it) must be bit-identical."""
    import mmx
    from mmx import synth
    cfg = synth.config(1)
    _say("rk12 look-ahead")
    runs = []
    for defeat in (False, True):
        sim = mmx.Sim(cfg["n"], cfg["cell"])
        synth.setup_sim(sim, cfg)
        for _ in range(300):
            if defeat:
                sim.device_array("H_eff")   # hands out a buffer: no look-ahead
            sim.step()
        runs.append((sim.getarray("m"), sim.getv("t"), sim.getv("dt"), sim.getv("badsteps"),
                     sim.getv("evals")))
        sim.close()
    (ma, ta, da, ba, ea), (mb, tb, db, bb, eb) = runs
    assert np.array_equal(ma, mb) and ta == tb and da == db and ba == bb
    _say("evals with look-ahead %d, without %d, badsteps %d" % (ea, eb, ba))
    assert ea <= eb + ba + 1


def test_step_counts_and_errors(torch_gpu):
    import mmx
    sim = mmx.Sim((1, 8, 8), (3e-9, 5e-9, 5e-9))
    with pytest.raises(mmx.MMXError):
        sim.step()                      # Msat not set
    with pytest.raises(mmx.MMXError):
        sim.setv("nonsense", 1.0)
    with pytest.raises(mmx.MMXError):
        sim.load("nonsense")
    with pytest.raises(mmx.MMXError):
        mmx.DemagPlan((1, 8, 0), (1e-9,) * 3)    # empty axis



@pytest.mark.parametrize("solver,dt,nstep", [("solver/bdf-euler", 1e-14, 8),
                                             ("solver/am12", 1e-13, 6)])
def test_implicit_solvers_against_oracle(torch_gpu, solver, dt, nstep):
    """solver/bdf-euler (solve_euler_backward.go:29-91) and solver/am12
    (solve_am01.go:45-288) replayed by the oracle's restatement
    (oracle/solvers.py) on the oracle's H_eff and torque: m after every step
    (1e-5), t and dt, and the exact iteration and bad-step counts."""
    from oracle import solvers as OS
    n, cell = (2, 16, 32), (3e-9, 4e-9, 4e-9)
    _say("implicit", solver, n, dt)
    sim, prob, _ = _problem_pair(n, cell, zl=True, bext=(0.0, 0.004, 0.009))
    # a smooth state: from a random one the fixed-point iterations of
    # BDF-Euler need more than m_maxIterations passes at any useful dt
    i, j, k = np.meshgrid(*(np.arange(x) for x in n), indexing="ij")
    m = O.normalize(np.stack([0.3 * np.cos(2 * np.pi * k / n[2]) + 0.1 * i,
                              0.4 * np.sin(2 * np.pi * j / n[1]), np.ones(n)])).astype(np.float32)
    sim.setarray("m", m)
    sim.load(solver)
    sim.setv("dt", dt)
    rhs = OS.llg_rhs(prob)
    o = OS.BDFEuler(dt) if solver == "solver/bdf-euler" else OS.AM12(dt)
    mo = m.astype(np.float64)
    for k in range(nstep):
        sim.step()
        mo = o.step(mo, rhs, O.GAMMA0)
        got_t, got_dt = sim.getv("t"), sim.getv("dt")
        _say("step %d t %.6e/%.6e dt %.6e/%.6e iters %d/%d" % (
            k, got_t, o.t, got_dt, o.dt, sim.getv("bdf_iterations"), o.iterations))
        assert sim.getv("bdf_iterations") == o.iterations
        # AM12's step sizes come from ratios of FP32 maxima of |dm| ~ 1e-5
        # (relative precision ~ 1e-3 with m ~ 1 in float32), as in the reference
        rel = 1e-9 if solver == "solver/bdf-euler" else 2e-3
        assert abs(got_t - o.t) <= rel * o.t
        assert abs(got_dt - o.dt) <= rel * o.dt
        if solver == "solver/am12":
            assert sim.getv("badsteps") == o.badsteps
        assert sim.getv("step") == k + 1
        sim.getv("<m>")              # normalises m on both sides, as the next update does
        got = sim.getarray("m")
        assert np.abs(got - O.normalize(mo)).max() < 1e-5

@pytest.mark.parametrize("solver", ["solver/heun", "solver/euler"])
def test_fixed_step_solvers(torch_gpu, solver):
    """solver/heun (solve_heun.go:27-79) = one predictor/corrector try at the
    fixed dt; solver/euler (solve_euler.go:22-50) = m + dt gamma tau(m) with
    the lazy normalisation; 12 steps against the oracle's stages."""
    n, cell = (2, 16, 32), (3e-9, 4e-9, 4e-9)
    _say("fixed-step", solver, n)
    sim, prob, m = _problem_pair(n, cell, zl=True, bext=(0.0, 0.004, 0.009))
    sim.load(solver)
    dt = 2e-14
    sim.setv("dt", dt)
    mo = m.astype(np.float64).copy()
    for _ in range(12):
        sim.step()
        if solver == "solver/heun":
            prob.try_step(mo, dt)
        else:
            # the normUpdater before each step's first stage (every step)
            mn = O.normalize(mo)
            tau = prob.torque(mn, prob.heff(mn))
            mo = mn + (O.GAMMA0 * dt) * tau
    assert abs(sim.getv("t") - 12 * dt) < 1e-6 * dt and sim.getv("step") == 12
    got = sim.getarray("m") if solver == "solver/heun" else None
    if solver == "solver/euler":
        # compare the normalised state, as the next update sees it
        sim.getv("<m>")
        got = sim.getarray("m")
        want = O.normalize(mo)
    else:
        want = prob.normalized(mo)
    assert np.abs(got - want).max() < 1e-5


# Transform lengths 2^a * {3, 5, 7} on the register path (mixed-radix
# in-register FFTs): every x length (zero-padded and periodic), film y, 3-D y
# up to 448 and z up to 56, against the oracle's convolution.
RADIX_CASES = [((2, 12, 48), (0, 0, 0)), ((2, 40, 24), (0, 0, 0)), ((1, 28, 96), (0, 0, 0)),
               ((1, 48, 40), (0, 0, 0)), ((12, 8, 16), (0, 0, 0)), ((20, 4, 8), (0, 0, 0)),
               ((28, 4, 8), (0, 0, 0)), ((4, 96, 56), (0, 0, 0)), ((2, 192, 16), (0, 0, 0)),
               ((2, 160, 16), (0, 0, 0)), ((2, 112, 16), (0, 0, 0)), ((2, 224, 16), (0, 0, 0)),
               ((2, 80, 112), (0, 0, 0)), ((2, 20, 192), (0, 0, 0)), ((2, 56, 80), (0, 0, 0)),
               ((4, 24, 96), (0, 1, 0)), ((4, 8, 48), (0, 0, 1)), ((24, 8, 8), (1, 0, 0)),
               ((1, 40, 56), (0, 1, 0))]


@pytest.mark.parametrize("n,pbc", RADIX_CASES)
def test_demag_odd_lengths_register_path(torch_gpu, n, pbc):
    import mmx
    torch = torch_gpu
    cell = (3e-9, 4e-9, 4e-9)
    _say("odd-length demag B", n, pbc)
    rng = np.random.default_rng(sum(n) + 3)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    ms_map = rng.uniform(0.5, 1.0, size=n).astype(np.float32)
    plan = mmx.DemagPlan(n, cell, pbc)
    p = plan.p
    kh, _ = O.kernel_fft(p, K9=O.kernel_arne(p, cell, pbc))
    for msat_map in (None, ms_map):
        want = O.MU0 * 8e5 * O.demag_conv(n, p, kh, m.astype(np.float64), msat_map)
        dm = torch.from_numpy(m).cuda()
        db = torch.empty_like(dm)
        dmap = None if msat_map is None else torch.from_numpy(msat_map).cuda()
        plan.update_b(db.data_ptr(), dm.data_ptr(), 8e5, None if dmap is None else dmap.data_ptr())
        torch.cuda.synchronize()
        assert relerr(db.cpu().numpy(), want) < TOL


@pytest.mark.parametrize("n", [(2, 24, 96), (1, 40, 80), (4, 56, 112)])
def test_heff_odd_lengths_general_epilogue(torch_gpu, n):
    """Zhang-Li with an Msat map at lengths with factors 3, 5, 7: the register
    x pass with the general per-cell epilogue, H_eff, torque and RK tries."""
    cell = (3e-9, 4e-9, 4e-9)
    _say("odd-length H_eff/torque general epilogue", n)
    sim, prob, m = _problem_pair(n, cell, anis=True, zl=True, bext=(0.0, 0.004, 0.009))
    mo = prob.normalized(m.astype(np.float64).copy())
    want_h = prob.heff(mo)
    assert relerr(sim.getarray("H_eff"), want_h) < TOL
    assert relerr(sim.getarray("torque"), prob.torque(mo, want_h)) < TOL
    mo = m.astype(np.float64).copy()
    for k in range(4):
        dt = 1e-14 * (1 + 0.1 * k)
        eg = sim.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12
    assert np.abs(sim.getarray("m") - prob.normalized(mo)).max() < TOL
