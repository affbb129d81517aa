"""MASK quantities in the fused engine, and maps on the register-FFT path.

* B_ext (mod_zeeman.go:26), anisU (mod_uniaxialanisotropy.go:30), and the
  Zhang-Li polarisation and xi masks (zhang-li_torque.cu:167,171) set with
  mmx_sim_setmask: H_eff and the torque against the FP64 oracle (1e-5 of the
  field's maximum), same-dt RK12 tries, and the Zeeman / anisotropy energies;
* the same with Msat, Aex, alpha and Ku maps on top; power-of-two x sizes
  run the register-FFT x pass with the general per-cell epilogue, the others
  the generic LDS passes: both against the oracle;
* the maps (Msat, Aex, alpha, Ku) and the B_ext / anisU masks without the
  spin-transfer term (with Msat and Aex holes, open and periodic): the
  register x pass's two-cell epilogue with exchange weights, against the
  oracle, H_eff, torque and RK12 tries;
* the standalone uniaxialAnisotropyAsync with its u maps.
"""
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


def _pair(mmx, n, cell, seed, anis=True, zl=True, msat_map=False, other_maps=False):
    rng = np.random.default_rng(seed)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    bext = (0.004, -0.006, 0.009)
    bmap = rng.uniform(-1.0, 2.0, size=(3,) + n).astype(np.float32)
    umap = rng.normal(size=(3,) + n).astype(np.float32)
    pmap = rng.uniform(0.2, 1.0, size=n).astype(np.float32)
    xmap = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    sim = mmx.Sim(n, cell)
    sim.setv("Msat", 8e5)
    sim.setv("Aex", 1.3e-11)
    sim.setv("alpha", 0.05)
    sim.setv("B_ext", bext)
    sim.setmask("B_ext", bmap, 3)
    kw = dict(bext=bext, bext_map=bmap)
    if anis:
        sim.load("anisotropy/uniaxial")
        sim.setv("Ku", 5e5)
        sim.setv("anisU", (0.6, 0.0, 0.8))
        sim.setmask("anisU", umap, 3)
        kw.update(ku=5e5, anis_u=(0.6, 0.0, 0.8), anis_u_map=umap)
    if zl:
        sim.load("zhang-li")
        sim.setv("j", (2e11, 0.0, 1e12))
        sim.setv("polarisation", 0.56)
        sim.setv("xi", 0.05)
        sim.setmask("polarisation", pmap, 1)
        sim.setmask("xi", xmap, 1)
        kw.update(j=(2e11, 0.0, 1e12), pol=0.56, xi=0.05, pol_map=pmap, xi_map=xmap)
    if msat_map:
        ms = rng.uniform(0.6, 1.0, size=n).astype(np.float32)
        sim.setmask("Msat", ms, 1)
        kw.update(msat_map=ms)
    if other_maps:
        for name, key in (("Aex", "aex_map"), ("alpha", "alpha_map"), ("Ku", "ku_map")):
            a = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
            sim.setmask(name, a, 1)
            kw[key] = a
    sim.setarray("m", m)
    prob = O.Problem(n, cell, msat=8e5, aex=1.3e-11, alpha=0.05, **kw)
    return sim, prob, m


# power-of-two x sizes run the register-FFT x pass; 48 and 96 the generic one
CASES = [((1, 32, 64), False, False), ((4, 16, 32), True, False), ((8, 16, 16), True, True),
         ((2, 24, 48), False, True), ((1, 64, 96), True, True)]


@pytest.mark.parametrize("n,msat_map,other_maps", CASES)
def test_masks_heff_torque(mmx_gpu, n, msat_map, other_maps):
    mmx = mmx_gpu
    cell = (3e-9, 4e-9, 4e-9)
    _say("masks H_eff/torque", n, msat_map, other_maps)
    sim, prob, m = _pair(mmx, n, cell, sum(n), msat_map=msat_map, other_maps=other_maps)
    mo = prob.normalized(m.astype(np.float64).copy())
    want_h = prob.heff(mo)
    eh = relerr(sim.getarray("H_eff"), want_h)
    et = relerr(sim.getarray("torque"), prob.torque(mo, want_h))
    _say("masks", n, "H_eff %.2e torque %.2e" % (eh, et))
    assert eh < TOL and et < TOL
    sim.close()


def test_masks_rk12_tries_and_energies(mmx_gpu):
    mmx = mmx_gpu
    n, cell = (2, 32, 64), (3e-9, 4e-9, 4e-9)
    sim, prob, m = _pair(mmx, n, cell, 3)
    mo = m.astype(np.float64).copy()
    for k in range(8):
        dt = 1e-14 * (1 + 0.1 * k)
        eg = sim.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12, (k, eg, eo)
    got = sim.getarray("m")
    want = prob.normalized(mo)
    _say("masks rk12 |dm| %.2e" % np.abs(got - want).max())
    assert np.abs(got - want).max() < TOL
    # energies at the current m: E_zeeman = -V Ms sum m . B_ext_eff,
    # E_anis = -1/2 V mu0 Ms sum m . H_anis (mod_energy.go:24-100)
    V = cell[0] * cell[1] * cell[2]
    mg = sim.getarray("m").astype(np.float64)
    beff = np.array(prob.P.bext[:])[:, None, None, None] * prob._keep["bext_map"]
    ez = -V * 8e5 * float((mg * beff).sum())
    assert abs(sim.getv("E_zeeman") - ez) <= 1e-5 * abs(ez)
    ku2 = 2 * 5e5 / (O.MU0 * 8e5)
    ha = O.uniaxial(mg, float(np.float32(ku2)), (0.6, 0.0, 0.8), u_map=prob._keep["u_map"])
    ea = -0.5 * V * O.MU0 * 8e5 * float((mg * ha).sum())
    assert abs(sim.getv("E_anis") - ea) <= 1e-5 * abs(ea)
    sim.close()


def test_mask_removal_restores_uniform(mmx_gpu):
    mmx = mmx_gpu
    n, cell = (1, 16, 32), (3e-9, 4e-9, 4e-9)
    sim, prob, m = _pair(mmx, n, cell, 9, zl=False)
    for name in ("B_ext", "anisU"):
        sim.setmask(name, None, 3)
    prob2 = O.Problem(n, cell, msat=8e5, aex=1.3e-11, alpha=0.05, bext=(0.004, -0.006, 0.009),
                      ku=5e5, anis_u=(0.6, 0.0, 0.8))
    mo = prob2.normalized(m.astype(np.float64).copy())
    assert relerr(sim.getarray("H_eff"), prob2.heff(mo)) < TOL
    sim.close()


def test_uniaxial_kernel_u_maps(mmx_gpu):
    import torch
    from mmx import kernels as K
    n = (3, 8, 16)
    rng = np.random.default_rng(21)
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    umap = rng.normal(size=(3,) + n).astype(np.float32)
    kmap = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    ms = rng.uniform(0.5, 1.0, size=n).astype(np.float32)
    ms[0, 0, :2] = 0.0
    ku2 = float(np.float32(2 * 5e5 / (O.MU0 * 8e5)))
    u = (0.6, -0.3, 0.8)
    got = K.uniaxial_anisotropy(torch.from_numpy(m).cuda(), ku2, u, ku_map=torch.from_numpy(kmap).cuda(),
                                msat_map=torch.from_numpy(ms).cuda(),
                                u_map=torch.from_numpy(umap).cuda()).cpu().numpy()
    want = O.uniaxial(m, ku2, np.float32(u).astype(np.float64), ku_map=kmap, msat_map=ms,
                      u_map=umap)
    assert relerr(got, want) < TOL


# Maps (Msat, Aex, alpha, Ku) and the B_ext / anisU masks without the
# spin-transfer term: the register x pass runs the two-cell epilogue with
# exchange weights; Msat holes (zero cells) check the masked renormalisation
# and the anisotropy constant's zero guard.
# maps: "msat", "msat+aex", "all" (Msat, Aex, alpha, Ku), "alpha+ku" (no Msat),
# "masks" (the B_ext and anisU masks alone), "all+masks"
MAP_CASES = [((4, 16, 32), (0, 0, 0), "msat"), ((2, 32, 64), (0, 0, 0), "msat+aex"),
             ((2, 16, 64), (0, 0, 1), "all"), ((4, 16, 32), (0, 1, 0), "msat+aex"),
             ((4, 8, 16), (1, 0, 0), "msat"), ((2, 32, 32), (0, 0, 0), "all"),
             ((4, 16, 32), (0, 0, 0), "alpha+ku"), ((4, 16, 32), (0, 0, 0), "masks"),
             ((2, 16, 64), (0, 0, 1), "all+masks"), ((4, 16, 32), (0, 1, 0), "all+masks"),
             # x lengths with a factor 3, 5, 7 (the map variant of the mixed-radix passes)
             ((2, 16, 48), (0, 0, 0), "all+masks"), ((2, 8, 40), (0, 0, 0), "msat"),
             ((2, 8, 56), (0, 1, 0), "msat+aex")]


@pytest.mark.parametrize("n,pbc,maps", MAP_CASES)
def test_maps_without_spin_transfer(mmx_gpu, n, pbc, maps):
    mmx = mmx_gpu
    cell = (3e-9, 4e-9, 4e-9)
    _say("scalar maps only", n, pbc, maps)
    rng = np.random.default_rng(sum(n) + 7 * sum(pbc) + len(maps))
    m = O.normalize(rng.normal(size=(3,) + n)).astype(np.float32)
    bext = (0.004, -0.006, 0.009)
    sim = mmx.Sim(n, cell, pbc)
    sim.setv("Msat", 8e5)
    sim.setv("Aex", 1.3e-11)
    sim.setv("alpha", 0.05)
    sim.setv("B_ext", bext)
    sim.load("anisotropy/uniaxial")
    sim.setv("Ku", 5e5)
    sim.setv("anisU", (0.6, 0.0, 0.8))
    kw = dict(bext=bext)
    if maps in ("masks", "all+masks"):
        bmap = rng.uniform(-1.0, 2.0, size=(3,) + n).astype(np.float32)
        umap = rng.normal(size=(3,) + n).astype(np.float32)
        sim.setmask("B_ext", bmap, 3)
        sim.setmask("anisU", umap, 3)
        kw.update(bext_map=bmap, anis_u_map=umap)
    if maps not in ("alpha+ku", "masks"):
        ms = rng.uniform(0.6, 1.0, size=n).astype(np.float32)
        ms[rng.uniform(size=n) < 0.1] = 0.0
        sim.setmask("Msat", ms, 1)
        kw.update(msat_map=ms)
    if maps in ("msat+aex", "all", "all+masks"):
        ax = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
        ax[rng.uniform(size=n) < 0.05] = 0.0
        sim.setmask("Aex", ax, 1)
        kw.update(aex_map=ax)
    if maps in ("all", "alpha+ku", "all+masks"):
        for name, key in (("alpha", "alpha_map"), ("Ku", "ku_map")):
            a = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
            sim.setmask(name, a, 1)
            kw[key] = a
    sim.setarray("m", m)
    prob = O.Problem(n, cell, pbc=pbc, msat=8e5, aex=1.3e-11, alpha=0.05, ku=5e5,
                     anis_u=(0.6, 0.0, 0.8), **kw)
    mo = prob.normalized(m.astype(np.float64).copy())
    want_h = prob.heff(mo)
    eh = relerr(sim.getarray("H_eff"), want_h)
    et = relerr(sim.getarray("torque"), prob.torque(mo, want_h))
    # the tries start from m as normalised with the Msat map (holes are 0)
    for k in range(6):
        dt = 1e-14 * (1 + 0.1 * k)
        eg = sim.try_step(dt)
        eo = prob.try_step(mo, dt)
        assert abs(eg - eo) <= 1e-3 * eo + 1e-12, (k, eg, eo)
    dm = np.abs(sim.getarray("m") - prob.normalized(mo)).max()
    _say("maps only", n, pbc, maps, "H_eff %.2e torque %.2e rk12 |dm| %.2e" % (eh, et, dm))
    assert eh < TOL and et < TOL and dm < TOL
    sim.close()
