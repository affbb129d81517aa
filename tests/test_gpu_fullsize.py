"""Every benchmark configuration at its stated size, against the FP64 oracle.

* bulk 256x256x64 (internal 64x256x256, padded 128x512x512), Ku = 5e5 along
  z, seeded random m: the Kernel_Arne real-space kernel bit for bit, H_eff
  and the torque to 1e-5 of max |field|, and 10 same-dt RK12 tries
  (solve_rk12.go:124-214) with m to 1e-5;
* precision film 1024x1024x32 of 2 nm cells, FP64 Newell kernel: B_demag and
  H_eff to 1e-5 of max |field| (padded 64x2048x2048);
* the batch of 64 uneven films with Zhang-Li (synth.batch_configs(64)):
  same-dt tries of several simulations replayed by the oracle, and all 64
  against their own engines;
* the learned-correction film 256x256x4: the coarse pipeline and the full
  prediction (coarse field + net) against the oracle's, and the net's
  gradients at that size.

The oracle runs on the host cores (OpenMP); each case prints what it is
doing before every long step.
"""
import time

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


# ------------------------------------------------------------------- bulk

@pytest.fixture(scope="module")
def bulk(mmx_gpu):
    """The bulk configuration on both sides, sharing nothing but the inputs."""
    from mmx import synth
    mmx = mmx_gpu
    cfg = synth.config(2)
    n, cell = cfg["n"], cfg["cell"]
    p = O.pad_size(n)
    t0 = time.perf_counter()
    _say("bulk: oracle Kernel_Arne", p)
    K9 = O.kernel_arne(p, cell)
    _say("bulk: oracle kernel FFT (%.1f s so far)" % (time.perf_counter() - t0))
    kh, _ = O.kernel_fft(p, K9=K9)
    prob = O.Problem(n, cell, msat=cfg["msat"], aex=cfg["aex"], alpha=cfg["alpha"],
                     bext=cfg["bext"], ku=cfg["ku"], anis_u=cfg["anis_u"], kernel=kh)
    _say("bulk: oracle ready (%.1f s)" % (time.perf_counter() - t0))
    return cfg, K9, prob


def test_bulk_kernel_arne_bit_exact(mmx_gpu, bulk):
    mmx = mmx_gpu
    cfg, K9, _ = bulk
    _say("bulk: GPU Kernel_Arne 128x512x512")
    plan = mmx.DemagPlan(cfg["n"], cfg["cell"], (0, 0, 0), keep_real_kernel=True)
    Kg = plan.real_kernel()
    assert Kg.shape == (6,) + tuple(plan.p)
    same = np.array_equal(Kg, K9[:6])
    _say("bulk: kernel bit-exact", same)
    assert same, np.abs(Kg - K9[:6]).max()


def test_bulk_heff_and_torque(mmx_gpu, bulk):
    from mmx import synth
    mmx = mmx_gpu
    cfg, _, prob = bulk
    sim = mmx.Sim(cfg["n"], cfg["cell"])
    synth.setup_sim(sim, cfg)
    got_h = sim.getarray("H_eff")
    got_t = sim.getarray("torque")
    _say("bulk: oracle H_eff")
    m = prob.normalized(cfg["m"].astype(np.float64).copy())
    want_h = prob.heff(m)
    want_t = prob.torque(m, want_h)
    eh, et = relerr(got_h, want_h), relerr(got_t, want_t)
    _say("bulk: H_eff %.2e torque %.2e" % (eh, et))
    assert eh < TOL and et < TOL
    sim.close()


def test_bulk_rk12_same_dt_tries(mmx_gpu, bulk):
    from mmx import synth
    mmx = mmx_gpu
    cfg, _, prob = bulk
    sim = mmx.Sim(cfg["n"], cfg["cell"])
    synth.setup_sim(sim, cfg)
    # a fresh oracle controller state (the Problem is shared by the module)
    prob.S = O._RK12()
    O.lib().orc_rk12_init(O.C.byref(prob.S))
    mo = cfg["m"].astype(np.float64).copy()
    for k in range(10):
        dt = 1e-15 * (1 + 0.2 * k)
        eg = sim.try_step(dt)
        eo = prob.try_step(mo, dt)
        _say("bulk: try %d dt %.2e err gpu %.6e oracle %.6e" % (k, dt, eg, eo))
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12
    got = sim.getarray("m")
    want = prob.normalized(mo)
    dm = np.abs(got - want).max()
    _say("bulk: |dm| after 10 tries %.2e" % dm)
    assert dm < TOL
    sim.close()


# -------------------------------------------------------------- precision

def test_precision_film_newell(mmx_gpu):
    from mmx import synth
    mmx = mmx_gpu
    cfg = synth.config(3)
    n, cell = cfg["n"], cfg["cell"]
    _say("precision: GPU sim", n)
    sim = mmx.Sim(n, cell)
    sim.set_kernel(mmx.MMX_KERNEL_NEWELL)
    synth.setup_sim(sim, cfg)
    got_h = sim.getarray("H_eff").astype(np.float64)
    got_b = sim.getarray("B").astype(np.float64)
    sim.close()
    t0 = time.perf_counter()
    _say("precision: oracle Newell kernel")
    prob = O.Problem(n, cell, msat=cfg["msat"], aex=cfg["aex"], alpha=cfg["alpha"],
                     bext=cfg["bext"], kernel="newell")
    _say("precision: oracle H_eff (%.1f s so far)" % (time.perf_counter() - t0))
    m = prob.normalized(cfg["m"].astype(np.float64).copy())
    h, hd, _, _ = prob.heff(m, terms=True)
    eb, eh = relerr(got_b, O.MU0 * hd), relerr(got_h, h)
    _say("precision: B %.2e H_eff %.2e (%.1f s)" % (eb, eh, time.perf_counter() - t0))
    assert eb < TOL and eh < TOL


# ------------------------------------------------------------------ batch

def _fixed_dt(sims, dt):
    for s in sims:
        s.setv("dt", dt)
        s.setv("mindt", dt)
        s.setv("maxdt", dt)


def test_batch_of_64_against_oracle_and_engines(mmx_gpu):
    from mmx import synth
    mmx = mmx_gpu
    cfgs = synth.batch_configs(64)

    def make():
        out = []
        for c in cfgs:
            s = mmx.Sim(c["n"], c["cell"])
            synth.setup_sim(s, c)
            out.append(s)
        return out

    dt, tries = 1.5e-14, 4
    sa, sb = make(), make()
    _fixed_dt(sa, dt)
    _fixed_dt(sb, dt)
    _say("batch: 64 simulations, %d tries at dt %.1e" % (tries, dt))
    mmx.Batch(sa, tries_per_graph=4).steps(tries)
    for s in sb:
        s.steps(tries)
    worst = 0.0
    for k, (x, y) in enumerate(zip(sa, sb)):
        assert x.getv("step") == tries and y.getv("step") == tries
        d = float(np.abs(x.getarray("m") - y.getarray("m")).max())
        worst = max(worst, d)
    _say("batch: max |dm| batch vs own engines over 64 sims %.2e" % worst)
    assert worst < TOL
    # oracle replays: the smallest, the largest and three others
    sizes = [int(np.prod(c["n"])) for c in cfgs]
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 7, 31, 50})
    for k in pick:
        c = cfgs[k]
        prob = O.Problem(c["n"], c["cell"], msat=c["msat"], aex=c["aex"], alpha=c["alpha"],
                         j=c["j"], pol=c["pol"], xi=c["xi"])
        mo = c["m"].astype(np.float64).copy()
        for _ in range(tries):
            prob.try_step(mo, dt)
        err = float(np.abs(sa[k].getarray("m") - prob.normalized(mo)).max())
        _say("batch: sim %d %s oracle |dm| %.2e" % (k, c["n"], err))
        assert err < TOL
    for s in sa + sb:
        s.close()


# -------------------------------------------------------------- surrogate

def test_surrogate_film_256x256x4(mmx_gpu):
    import torch
    from mmx import synth
    from oracle import surrogate as OS
    mmx = mmx_gpu
    cfg = synth.config(5)
    n, cell, ms = cfg["n"], cfg["cell"], cfg["msat"]
    sim = mmx.Sim(n, cell)
    synth.setup_sim(sim, cfg)
    sur = mmx.Surrogate(sim, seed=7)
    Ws, bs = OS.init_params(seed=8, n0=n[0])
    Ws[3] *= 10.0
    rng = np.random.default_rng(9)
    bs = [rng.normal(size=b.shape) * 0.05 for b in bs]
    sur.set_params(mmx.Surrogate.flatten(Ws, bs))
    Ws, bs = mmx.Surrogate.unflatten(sur.get_params())
    m = synth.vortex_m(12, n, cell)
    dm = torch.from_numpy(m).cuda()
    bco, bnet = torch.empty_like(dm), torch.empty_like(dm)
    sur.predict_b(dm.data_ptr(), bco.data_ptr(), coarse_only=True)
    sur.predict_b(dm.data_ptr(), bnet.data_ptr())
    torch.cuda.synchronize()
    bco, bnet = bco.cpu().numpy().astype(np.float64), bnet.cpu().numpy().astype(np.float64)
    _say("surrogate: oracle coarse convolution")
    nc = OS.coarse_shape(n)
    cc = (cell[0], cell[1] * 2, cell[2] * 2)
    pc = O.pad_size(nc)
    kh, _ = O.kernel_fft(pc, K9=O.kernel_arne(pc, cc))
    bc = O.MU0 * ms * O.demag_conv(nc, pc, kh, OS.coarsen(m.astype(np.float64), n))
    up = OS.upsample(bc, n)
    e_co = relerr(bco, up)
    _say("surrogate: coarse pipeline %.2e" % e_co)
    assert e_co < TOL
    # the full prediction: the net on the packed (m, coarse B) input
    _say("surrogate: oracle net forward 256x256")
    x0 = OS.pack_input(m.astype(np.float64), up, ms)
    out, _ = OS.forward(Ws, bs, x0)
    want = OS.predict_b(out, up, ms)
    # the net's FP32 sums and bf16 roundings (flipped by FP32-vs-FP64 sums)
    # differ from the oracle's at the 1e-3 level of the correction itself
    corr_g, corr_o = bnet - bco, want - up
    e_net = float(np.linalg.norm(corr_g - corr_o) / np.linalg.norm(corr_o))
    _say("surrogate: net correction rel L2 %.2e, |correction| / |B| %.2e"
         % (e_net, np.linalg.norm(corr_o) / np.linalg.norm(up)))
    assert e_net < 5e-3
    # gradients at full size on a random input / target
    xr = np.zeros((n[1], n[2], 32))
    xr[:, :, :6 * n[0] + 2] = rng.normal(size=(n[1], n[2], 6 * n[0] + 2))
    tgt = rng.normal(size=(n[1], n[2], 3 * n[0])) * 0.3
    out_g, loss_g = sur.net(xr, tgt)
    out_o, acts = OS.forward(Ws, bs, OS.bf16(xr))
    loss_o, gout = OS.loss_and_grad_out(out_o, tgt, 3 * n[0])
    dWs, _ = OS.backward(Ws, acts, gout)
    gW, _ = mmx.Surrogate.unflatten(sur.get_grads())
    _say("surrogate: loss %.6e vs %.6e" % (loss_g, loss_o))
    assert abs(loss_g - loss_o) < 1e-2 * loss_o
    for l in range(4):
        used = slice(None) if l else slice(0, 6 * n[0] + 2)
        cos = float(np.sum(gW[l][:, :, used] * dWs[l][:, :, used])
                    / (np.linalg.norm(gW[l][:, :, used]) * np.linalg.norm(dWs[l][:, :, used])))
        _say("surrogate: layer %d gradient cosine %.8f" % (l, cos))
        assert cos > 0.9998
    sim.close()
