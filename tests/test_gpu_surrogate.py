"""The learned demag correction (mmx_surrogate) on the GPU, through the C ABI.

* the BF16-MFMA network forward, loss and gradients (every weight and bias
  of the 4 layers) against the FP64 oracle (oracle/surrogate.py) that rounds
  to bf16 at the same points: the remaining differences are FP32 vs FP64
  accumulation order and the bf16 roundings it flips (tolerances below);
* the coarse pipeline (in-plane block means, coarse exact demag, upsampling) against
  the oracle's coarse convolution, 1e-5;
* training on exact-solver trajectories lowers the error of the fine field
  below the coarse-only method's (the gain of the network);
* in RK12 the surrogate drives the corrector, the judge measures the
  residual, and a tolerance it cannot meet switches the solver to the exact
  field.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from oracle import surrogate as OS  # noqa: E402

CELL = (3e-9, 4e-9, 4e-9)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def mmx_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import mmx
    mmx.check(mmx.lib().mmx_init(0))
    return mmx


def _sim(mmx, n, msat=8e5):
    s = mmx.Sim(n, CELL)
    s.setv("Msat", msat)
    s.setv("Aex", 1.3e-11)
    s.setv("alpha", 0.02)
    return s


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [(4, 32, 32), (2, 16, 48), (1, 32, 16)])
def test_net_forward_loss_gradients_against_oracle(mmx_gpu, n):
    mmx = mmx_gpu
    _say("net vs oracle", n)
    n0 = n[0]
    sim = _sim(mmx, n)
    sur = mmx.Surrogate(sim, seed=1)
    Ws, bs = OS.init_params(seed=5, n0=n0)
    Ws[3] *= 10.0
    rng = np.random.default_rng(6)
    bs = [rng.normal(size=b.shape) * 0.1 for b in bs]
    sur.set_params(mmx.Surrogate.flatten(Ws, bs))
    # parameters as the GPU holds them (FP32)
    Ws, bs = mmx.Surrogate.unflatten(sur.get_params())
    x = np.zeros((n[1], n[2], 32))
    x[:, :, :6 * n0 + 2] = rng.normal(size=(n[1], n[2], 6 * n0 + 2))
    tgt = rng.normal(size=(n[1], n[2], 3 * n0)) * 0.3
    out_g, loss_g = sur.net(x, tgt)
    xo = OS.bf16(x)
    out_o, acts = OS.forward(Ws, bs, xo)
    loss_o, gout = OS.loss_and_grad_out(out_o, tgt, 3 * n0)
    dWs, dbs = OS.backward(Ws, acts, gout)
    e_out = _rel(out_g, out_o)
    _say("forward rel L2 %.2e, loss %.6e vs %.6e" % (e_out, loss_g, loss_o))
    assert e_out < 5e-3
    assert abs(loss_g - loss_o) < 1e-2 * loss_o
    gW, gb = mmx.Surrogate.unflatten(sur.get_grads())
    for l in range(4):
        used = slice(None) if l else slice(0, 6 * n0 + 2)
        ew = _rel(gW[l][:, :, used], dWs[l][:, :, used])
        eb = _rel(gb[l][:3 * n0] if l == 3 else gb[l], dbs[l][:3 * n0] if l == 3 else dbs[l])
        cos = float(np.sum(gW[l] * dWs[l]) / (np.linalg.norm(gW[l]) * np.linalg.norm(dWs[l])))
        _say("layer %d dW rel %.2e db rel %.2e cos %.8f" % (l, ew, eb, cos))
        assert ew < 2e-2 and eb < 2e-2 and cos > 0.9998


def test_coarse_pipeline_against_oracle(mmx_gpu):
    """coarse_only prediction = upsampled exact demag of the block-mean grid."""
    import torch
    mmx = mmx_gpu
    n = (4, 32, 48)
    sim = _sim(mmx, n)
    sur = mmx.Surrogate(sim, seed=2)
    m = O.normalize(np.random.default_rng(3).normal(size=(3,) + n)).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    db = torch.empty_like(dm)
    sur.predict_b(dm.data_ptr(), db.data_ptr(), coarse_only=True)
    torch.cuda.synchronize()
    nc = OS.coarse_shape(n)
    cc = (CELL[0], CELL[1] * 2, CELL[2] * 2)
    pc = O.pad_size(nc)
    kh, _ = O.kernel_fft(pc, K9=O.kernel_arne(pc, cc))
    mc = OS.coarsen(m.astype(np.float64), n)
    bc = O.MU0 * 8e5 * O.demag_conv(nc, pc, kh, mc)
    want = OS.upsample(bc, n)
    assert np.abs(db.cpu().numpy() - want).max() < 1e-5 * np.abs(want).max()


def test_training_beats_coarse_only(mmx_gpu):
    """After training on exact trajectories, the fine field's error on unseen
    textures is below the coarse-only method's."""
    import torch
    import mmx.train as T
    from mmx import synth
    mmx = mmx_gpu
    n = (4, 64, 64)
    sim = _sim(mmx, n)
    sur = mmx.Surrogate(sim, seed=3)
    plan = mmx.DemagPlan(n, CELL)
    # validation states: textures not used in training, relaxed a little like
    # the states the solver meets
    val = []
    for k in range(4):
        sim.setarray("m", synth.training_texture(1001 + 7 * k, n, CELL))
        sim.setv("alpha", 0.2)
        sim.steps(20)
        val.append(sim.getarray("m"))

    def errors():
        out = []
        for m in val:
            dm = torch.from_numpy(m).cuda()
            exact, bn, bco = torch.empty_like(dm), torch.empty_like(dm), torch.empty_like(dm)
            plan.update_b(exact.data_ptr(), dm.data_ptr(), 8e5)
            sur.predict_b(dm.data_ptr(), bn.data_ptr())
            sur.predict_b(dm.data_ptr(), bco.data_ptr(), coarse_only=True)
            torch.cuda.synchronize()
            e = exact.cpu().numpy().astype(np.float64)
            out.append((_rel(bco.cpu().numpy(), e), _rel(bn.cpu().numpy(), e)))
        return out

    before = errors()
    T.train(sur, sim, nsamples=4000, steps_per_sample=2, samples_per_trajectory=50,
            lr=5e-4, lr_final=2e-4, alpha=0.2)
    after = errors()
    for (co, n0_), (_, n1_) in zip(before, after):
        _say("coarse-only %.3e, untrained net %.3e, trained net %.3e" % (co, n0_, n1_))
    # the trained net beats the coarse field on every unseen state, by 1.3x on
    # the geometric mean (4000 pairs; the benchmark trains on 20000)
    assert all(a[1] < a[0] for a in after), after
    gain = float(np.exp(np.mean([np.log(a[0] / a[1]) for a in after])))
    _say("geometric-mean gain %.2f" % gain)
    assert gain > 1.3, after


def test_surrogate_in_rk12_and_fallback(mmx_gpu):
    mmx = mmx_gpu
    from mmx import synth
    n = (4, 32, 32)
    m = synth.vortex_m(11, n, CELL)
    sa, sb = _sim(mmx, n), _sim(mmx, n)
    for s in (sa, sb):
        s.setarray("m", m)
        s.setv("m_maxerror", 1e-4)
    sur = mmx.Surrogate(sa, seed=4)
    # an untrained net with a tolerance it cannot meet: the first check falls back
    sa.set_surrogate(sur, tol=1e-9, check_every=1)
    sa.steps(5)
    sb.steps(5)
    assert sa.getv("surrogate_fallbacks") == 1 and sa.getv("surrogate_on") == 0
    assert sa.getv("surrogate_checks") == 1
    # after the fallback the exact path runs: same trajectory as the exact solver
    assert np.array_equal(sa.getarray("m"), sb.getarray("m"))
    # a tolerance the untrained net meets: the corrector runs on the surrogate
    sc = _sim(mmx, n)
    sc.setarray("m", m)
    sc.setv("m_maxerror", 1e-4)
    sur2 = mmx.Surrogate(sc, seed=4)
    sc.set_surrogate(sur2, tol=10.0, check_every=4)
    sc.steps(8)
    assert sc.getv("surrogate_on") == 1 and sc.getv("surrogate_evals") >= 8
    assert sc.getv("surrogate_checks") == 2
    _say("residual of the untrained net", sc.getv("surrogate_residual"))
    assert 0 < sc.getv("surrogate_residual") < 10.0
