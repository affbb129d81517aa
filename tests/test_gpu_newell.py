"""FP64 Newell kernel on the GPU and the compensated energy reductions,
against the FP64 oracle, through the C ABI.

* real-space kernel: the GPU's FP64 table rounded to float32 against the
  oracle's FP64 table rounded the same way (both sides run the same
  formulas in FP64; the oracle's are pinned by tests/test_oracle_newell.py);
* transformed kernel: FP64 FFT on both sides, float32 store, 1e-6;
* fields and H_eff with the Newell kernel: 1e-5 (the FP32 path's bar);
* energies (micromag/energy, modules/mod_energy.go:24-100): FP32 fields
  with FP64 compensated dot products against FP64 everywhere, 1e-5.
"""
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


def relerr(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


NEWELL_CASES = [((4, 6, 8), (3e-9, 4e-9, 5e-9)),
                ((1, 16, 32), (3e-9, 3.90625e-9, 3.90625e-9)),
                ((6, 8, 10), (2e-9, 2e-9, 2e-9)),
                ((8, 32, 32), (2e-9, 2e-9, 2e-9)),
                ((2, 24, 40), (3e-9, 4e-9, 4e-9))]


@pytest.mark.parametrize("n,cell", NEWELL_CASES)
def test_newell_real_and_fft_kernel(torch_gpu, n, cell):
    import mmx
    _say("newell kernel", n, cell)
    plan = mmx.DemagPlan(n, cell, kernel_type=mmx.MMX_KERNEL_NEWELL, keep_real_kernel=True)
    p = plan.p
    K9d = O.kernel_newell(p, cell)
    Kg = plan.real_kernel().astype(np.float64)
    Ko = K9d[:6]
    # FP64 on both sides, then float32 rounding: equal to ~1 float ulp
    assert np.abs(Kg - Ko).max() <= 1.2e-7 * np.abs(Ko).max()
    kh, bad = O.kernel_fft(p, K9d=K9d)
    ko = plan.ko
    assert relerr(plan.kernel(), kh[:, :ko[0], :ko[1], :ko[2]]) < 1e-6
    assert plan.bad_over_good < 1e-10


CONV_CASES = [((1, 32, 128), (3e-9, 3.90625e-9, 3.90625e-9)),
              ((8, 64, 64), (2e-9, 2e-9, 2e-9)),
              ((4, 16, 24), (3e-9, 4e-9, 4e-9)),
              ((1, 256, 256), (3e-9, 4e-9, 4e-9))]


@pytest.mark.parametrize("n,cell", CONV_CASES)
def test_newell_demag_field(torch_gpu, n, cell):
    import mmx
    torch = torch_gpu
    _say("newell B", n)
    rng = np.random.default_rng(sum(n) + 3)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    plan = mmx.DemagPlan(n, cell, kernel_type=mmx.MMX_KERNEL_NEWELL)
    p = plan.p
    kh, _ = O.kernel_fft(p, K9d=O.kernel_newell(p, cell))
    want = O.MU0 * 1e6 * O.demag_conv(n, p, kh, m.astype(np.float64))
    dm = torch.from_numpy(m).cuda()
    db = torch.empty_like(dm)
    plan.update_b(db.data_ptr(), dm.data_ptr(), 1e6)
    torch.cuda.synchronize()
    assert relerr(db.cpu().numpy(), want) < TOL


def test_newell_uniform_cube_mean_field(torch_gpu):
    """A uniformly magnetised cube of 16^3 cells: <B> = -mu0 Ms / 3 (exact with
    Newell; tests/demag3D.py's known answer)."""
    import mmx
    torch = torch_gpu
    n = (16, 16, 16)
    plan = mmx.DemagPlan(n, (2e-9,) * 3, kernel_type=mmx.MMX_KERNEL_NEWELL)
    for c in range(3):
        m = np.zeros((3,) + n, np.float32)
        m[c] = 1.0
        dm = torch.from_numpy(m).cuda()
        db = torch.empty_like(dm)
        plan.update_b(db.data_ptr(), dm.data_ptr(), 8e5)
        torch.cuda.synchronize()
        mean = db.cpu().numpy().astype(np.float64).mean(axis=(1, 2, 3))
        want = np.zeros(3)
        want[c] = -O.MU0 * 8e5 / 3
        assert np.abs(mean - want).max() < 1e-6 * O.MU0 * 8e5


def _pair(n, cell, kernel, anis, seed, smooth=False):
    import mmx
    from mmx import synth
    rng = np.random.default_rng(seed)
    if smooth:
        m = synth.smooth_texture_m(seed, n, cell).astype(np.float32)
    else:
        m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    sim = mmx.Sim(n, cell)
    if kernel == "newell":
        sim.set_kernel(mmx.MMX_KERNEL_NEWELL)
    sim.setv("Msat", 1e6)
    sim.setv("Aex", 1.5e-11)
    sim.setv("alpha", 0.05)
    sim.setv("B_ext", (0.002, -0.003, 0.01))
    kw = {}
    if anis:
        sim.load("anisotropy/uniaxial")
        sim.setv("Ku", 5e5)
        sim.setv("anisU", (1.0, 0.0, 0.0))
        kw.update(ku=5e5, anis_u=(1.0, 0.0, 0.0))
    sim.setarray("m", m)
    prob = O.Problem(n, cell, msat=1e6, aex=1.5e-11, alpha=0.05, bext=(0.002, -0.003, 0.01),
                     kernel=("newell" if kernel == "newell" else None), **kw)
    return sim, prob, m


@pytest.mark.parametrize("n,cell,kernel,anis", [
    ((8, 64, 64), (2e-9,) * 3, "newell", False),
    ((1, 64, 128), (3e-9, 3.90625e-9, 3.90625e-9), "newell", True),
    ((4, 32, 48), (3e-9, 4e-9, 4e-9), "arne", True)])
def test_heff_and_energies(torch_gpu, n, cell, kernel, anis):
    _say("heff+energies", n, kernel, anis)
    sim, prob, m = _pair(n, cell, kernel, anis, seed=31, smooth=(n[0] == 8))
    mo = prob.normalized(m.astype(np.float64).copy())
    h, hd, he, ha = prob.heff(mo, terms=True)
    assert relerr(sim.getarray("H_eff"), h) < TOL
    V = float(np.prod(cell))
    msat = 1e6
    want = {
        "E_zeeman": -V * msat * sum(mo[c].sum() * b for c, b in enumerate((0.002, -0.003, 0.01))),
        "E_ex": -0.5 * V * O.MU0 * msat * float((mo * he).sum()),
        "E_demag": -0.5 * V * msat * float((mo * (O.MU0 * hd)).sum()),
    }
    if anis:
        want["E_anis"] = -0.5 * V * O.MU0 * msat * float((mo * ha).sum())
    want["E"] = sum(want.values())
    scale = max(abs(v) for v in want.values())
    for name, w in want.items():
        got = sim.getv(name)
        _say(name, got, w)
        assert abs(got - w) <= TOL * scale, name


def test_newell_rejects_periodic(torch_gpu):
    import mmx
    with pytest.raises(mmx.MMXError):
        mmx.DemagPlan((2, 8, 8), (2e-9,) * 3, (0, 0, 1), kernel_type=mmx.MMX_KERNEL_NEWELL)
