"""Batched stepping of uneven simulations (mmx_batch), through the C ABI.

* same-dt replays: every simulation of a batch (films and 3-D grids, FFT
  lengths with factors 2, 3, 5, 7, Zhang-Li on and off) against the FP64
  oracle's RK12 tries (orc_rk12_try), 1e-5 after 6 steps;
* the hipGraph replays against plain launches of the same kernels: bit-exact
  after every replay, with eager work (reductions) between the replays;
* the persistent launch (every pass and the controller in one cooperative
  launch, grid-wide barriers) against plain launches: bit-exact;
* tile shapes (the autotuner's LDS budgets) do not change the arithmetic:
  bit-exact across budgets;
* free runs against each simulation's own engine (mmx_sim_run) within the
  float32 rounding of the different FFT passes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

CELL = (3e-9, 4e-9, 4e-9)
# internal (n0, n1, n2), Zhang-Li on
CASES = [((1, 32, 96), True),      # film, x half-length 96 = 3 * 32
         ((2, 64, 32), True),      # 3-D, z length 4
         ((3, 32, 160), False),    # z length 6, x half-length 160 = 5 * 32
         ((4, 96, 64), True),      # z length 8, y length 192
         ((1, 128, 224), True)]    # film, x half-length 224 = 7 * 32


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def mmx_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return mmx


def _params(k):
    rng = np.random.default_rng(100 + k)
    return dict(msat=float(rng.uniform(6e5, 1.2e6)), alpha=float(rng.uniform(0.005, 0.1)))


def _make(mmx, cases=CASES, seed=7, dt=None, maxerr=1e-4):
    sims, ms = [], []
    for k, (n, zl) in enumerate(cases):
        p = _params(k)
        s = mmx.Sim(n, CELL)
        s.setv("Msat", p["msat"])
        s.setv("Aex", 1.3e-11)
        s.setv("alpha", p["alpha"])
        s.setv("m_maxerror", maxerr)
        if zl:
            s.load("zhang-li")
            s.setv("j", (0.0, 0.0, 1e12))
            s.setv("polarisation", 0.56)
            s.setv("xi", 0.05)
        m = O.normalize(np.random.default_rng(seed + k).normal(size=(3,) + n)).astype(np.float32)
        s.setarray("m", m)
        if dt is not None:
            s.setv("dt", dt)
            s.setv("mindt", dt)
            s.setv("maxdt", dt)
        sims.append(s)
        ms.append(m)
    return sims, ms


def test_same_dt_replay_against_oracle(mmx_gpu):
    mmx = mmx_gpu
    dt = 1.5e-14
    sims, ms = _make(mmx, dt=dt)
    b = mmx.Batch(sims, tries_per_graph=4)
    b.steps(6)
    for k, ((n, zl), s, m) in enumerate(zip(CASES, sims, ms)):
        p = _params(k)
        kw = dict(j=(0.0, 0.0, 1e12), pol=0.56, xi=0.05) if zl else {}
        prob = O.Problem(n, CELL, msat=p["msat"], aex=1.3e-11, alpha=p["alpha"], **kw)
        mo = m.astype(np.float64).copy()
        for _ in range(6):
            prob.try_step(mo, dt)
        got = s.getarray("m")
        want = prob.normalized(mo)
        err = np.abs(got - want).max()
        _say("batch replay", n, zl, err, s.getv("step"), s.getv("t"))
        assert s.getv("step") == 6
        assert abs(s.getv("t") - 6 * dt) < 1e-9 * dt
        assert err < 1e-5


def _snapshot(sims):
    return [(s.getarray("m"), s.getv("t"), s.getv("dt"), s.getv("step")) for s in sims]


def test_graph_replays_bit_exact(mmx_gpu):
    mmx = mmx_gpu
    sa, _ = _make(mmx)
    sb, _ = _make(mmx)
    ba = mmx.Batch(sa, tries_per_graph=0)       # plain launches
    bb = mmx.Batch(sb, tries_per_graph=3)       # graph of 3 tries per replay
    for rep in range(4):
        ba.steps(5)
        bb.steps(5)
        for x, y in zip(_snapshot(sa), _snapshot(sb)):
            assert np.array_equal(x[0], y[0]) and x[1:] == y[1:], rep
        # eager work between the replays
        for s in sb:
            s.getv("<m>")
        _say("graph replay", rep, sb[0].getv("t"))


@pytest.mark.parametrize("tries", [1, 3, 64])
def test_persistent_launch_bit_exact(mmx_gpu, tries):
    """One cooperative launch per `tries` tries (passes and controller behind
    grid barriers) against plain launches: same bits, step counts and times,
    across steps() and run() stops that fall inside a launch."""
    mmx = mmx_gpu
    sa, _ = _make(mmx)
    sb, _ = _make(mmx)
    ba = mmx.Batch(sa, tries_per_graph=0)
    bb = mmx.Batch(sb)
    bb.set_persistent(tries)
    for rep, (kind, amount) in enumerate((("steps", 5), ("run", 2e-13), ("steps", 7))):
        getattr(ba, kind)(amount)
        getattr(bb, kind)(amount)
        for x, y in zip(_snapshot(sa), _snapshot(sb)):
            assert np.array_equal(x[0], y[0]) and x[1:] == y[1:], (rep, kind)
        _say("persistent", tries, kind, sb[0].getv("step"), sb[0].getv("t"))


def test_tile_budgets_do_not_change_results(mmx_gpu):
    mmx = mmx_gpu
    sa, _ = _make(mmx)
    sb, _ = _make(mmx)
    ba = mmx.Batch(sa)
    bb = mmx.Batch(sb)
    ba.set_lds({"xf": 8, "y": 8, "z": 8, "yi": 8, "xi": 8})
    chosen = bb.tune(reps=2)
    _say("tuned", chosen, bb.info())
    ba.steps(4)
    bb.steps(4)
    for x, y in zip(_snapshot(sa), _snapshot(sb)):
        assert np.array_equal(x[0], y[0]) and x[1:] == y[1:]


def test_free_run_against_single_engines(mmx_gpu):
    mmx = mmx_gpu
    sa, _ = _make(mmx, maxerr=1e-5)
    sb, _ = _make(mmx, maxerr=1e-5)
    mmx.Batch(sa).run(3e-12)
    for s in sb:
        s.run(3e-12)
    for (n, _), x, y in zip(CASES, sa, sb):
        ga, gb = np.array(x.getv("<m>")), np.array(y.getv("<m>"))
        _say("free run", n, x.getv("step"), y.getv("step"), np.abs(ga - gb).max())
        assert x.getv("t") >= 3e-12 and y.getv("t") >= 3e-12
        assert np.abs(ga - gb).max() < 2e-4


def test_batch_errors(mmx_gpu):
    mmx = mmx_gpu
    with pytest.raises(mmx.MMXError):
        mmx.Batch([])
    sims, _ = _make(mmx, cases=CASES[:1])
    b = mmx.Batch(sims)
    with pytest.raises(mmx.MMXError):
        b.set_lds({"xf": 0, "y": 8, "z": 8, "yi": 8, "xi": 8})
    with pytest.raises(mmx.MMXError):
        b.set_graph(-1)
    with pytest.raises(mmx.MMXError):
        b.set_persistent(-1)
