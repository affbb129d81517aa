"""Pins the FP64 CPU oracle before it is trusted as the checker.

Every oracle function is checked against something it does not share code
with: closed forms, numpy, an independent O(N^2) real-space sum, the
reference's own known-answer tests (tests/demag3D.py, tests/dipole.py) and
data fixtures (tests/golden/).  A planted mistake (sign, index, transposed
operand, dropped term) fails at least one of these.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

MU0 = O.MU0


# ----------------------------------------------------------------- FFT

@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 7, 8, 12, 60, 64, 105, 121, 256, 448])
def test_fft_line_matches_numpy(L):
    rng = np.random.default_rng(L)
    z = rng.normal(size=L) + 1j * rng.normal(size=L)
    ref = np.fft.fft(z)
    assert np.abs(O.fft_line(z, -1) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(O.fft_line(z, +1) - np.fft.ifft(z) * L).max() <= 1e-12 * max(1.0, np.abs(ref).max())


# --------------------------------------------------------- Kernel_Arne

def test_kernel_arne_cube_self_term():
    K = O.kernel_arne((8, 8, 8), (5e-9,) * 3)
    diag = [K[c, 0, 0, 0] for c in range(3)]
    for d in diag:
        assert abs(d + 1.0 / 3.0) < 2e-4           # cube demag factor -1/3
    assert abs(sum(diag) + 1.0) < 5e-4             # trace of N is 1
    for c in (3, 4, 5):                            # off-diagonal self terms vanish
        assert abs(K[c, 0, 0, 0]) < 1e-6


def test_kernel_arne_converges_with_accuracy():
    # the midpoint rule's error falls roughly like 1/accuracy^2 until the
    # staggered face grid's own ~2e-4 bias is reached
    errs = []
    for acc in (1, 2, 4, 6, 10):
        K = O.kernel_arne((4, 4, 4), (1e-9, 1e-9, 1e-9), accuracy=acc)
        errs.append(abs(K[0, 0, 0, 0] + 1.0 / 3.0))
    assert errs[0] > 3 * errs[1] > 9 * errs[2]
    assert max(errs[3:]) < 3e-4


def test_kernel_arne_far_field_is_point_dipole():
    """Far away the cell-averaged kernel is the dipole tensor
    K_sd(r) = V (3 r_s r_d / r^2 - delta_sd) / (4 pi r^3): pins signs, the
    source/destination order and the axis order of the storage."""
    p = (32, 32, 32)
    cell = (2e-9, 3e-9, 4e-9)
    K = O.kernel_arne(p, cell)
    V = cell[0] * cell[1] * cell[2]
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (1, 2): 3, (0, 2): 4, (0, 1): 5,
           (2, 1): 6, (2, 0): 7, (1, 0): 8}
    for (a, b, c) in [(9, 4, -7), (-11, 6, 3), (5, -12, 10)]:
        r = np.array([a * cell[0], b * cell[1], c * cell[2]])
        rr = np.linalg.norm(r)
        for (s, d), ci in idx.items():
            want = V * (3 * r[s] * r[d] / rr ** 2 - (s == d)) / (4 * np.pi * rr ** 3)
            got = K[ci, a % p[0], b % p[1], c % p[2]]
            # quadrupole corrections are O((cell/r)^2) ~ 1 %; a wrong sign,
            # index or axis order is O(1)
            assert abs(got - want) < 2e-2 * V / (4 * np.pi * rr ** 3), (a, b, c, s, d, got, want)


def test_kernel_arne_symmetry():
    K = O.kernel_arne((8, 12, 10), (2e-9, 3e-9, 5e-9)).astype(np.float64)
    kmax = np.abs(K).max()
    # K_sd ~ K_ds (computed by different integrations in the reference)
    for a, b in [(5, 8), (4, 7), (3, 6)]:
        assert np.abs(K[a] - K[b]).max() < 1e-3 * kmax
    # diagonal components even, off-diagonal odd in their two axes
    flip = lambda A, ax: np.roll(np.flip(A, axis=ax), 1, axis=ax)
    for c in range(3):
        for ax in range(3):
            assert np.abs(flip(K[c], ax) - K[c]).max() < 1e-6 * kmax
    for c, odd in [(3, (1, 2)), (4, (0, 2)), (5, (0, 1))]:
        for ax in range(3):
            sgn = -1 if ax in odd else 1
            assert np.abs(flip(K[c], ax) - sgn * K[c]).max() < 1e-6 * kmax


def test_kernel_fft_real_and_matches_numpy():
    p = (8, 12, 10)
    K = O.kernel_arne(p, (2e-9, 3e-9, 5e-9))
    kh, bad = O.kernel_fft(p, K9=K)
    assert bad < 1e-5                                 # maxwell_convolution.go:608-610
    for c in range(6):
        ref = np.fft.rfftn(K[c].astype(np.float64)).real / np.prod(p)
        assert np.abs(kh[c] - ref).max() < 1e-12 * np.abs(ref).max()


# ------------------------------------------------------- convolution

@pytest.mark.parametrize("n,pbc", [((2, 4, 6), (0, 0, 0)), ((1, 6, 5), (0, 0, 0)),
                                   ((3, 4, 4), (0, 1, 0)), ((1, 3, 8), (0, 0, 1))])
def test_conv_equals_bruteforce(n, pbc):
    rng = np.random.default_rng(sum(n))
    cell = (3e-9, 4e-9, 5e-9)
    p = O.pad_size(n, pbc)
    K = O.kernel_arne(p, cell, pbc)
    kh, _ = O.kernel_fft(p, K9=K)
    m = rng.normal(size=(3,) + n)
    ms = rng.uniform(0.5, 1.5, size=n)
    for msat_map in (None, ms):
        a = O.demag_conv(n, p, kh, m, msat_map)
        b = O.demag_brute(n, p, K, m, msat_map)
        assert np.abs(a - b).max() < 1e-10 * np.abs(b).max()


def test_conv_linearity():
    n = (2, 8, 8)
    p = O.pad_size(n)
    K = O.kernel_arne(p, (2e-9, 2e-9, 2e-9))
    kh, _ = O.kernel_fft(p, K9=K)
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=(3,) + n), rng.normal(size=(3,) + n)
    lhs = O.demag_conv(n, p, kh, 2.0 * a - 3.0 * b)
    rhs = 2.0 * O.demag_conv(n, p, kh, a) - 3.0 * O.demag_conv(n, p, kh, b)
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(lhs).max()


def test_demag3d_uniform_cube():
    """tests/demag3D.py:11-30: 32^3 cube of 5 nm cells, m = x: <H> = -Ms/3.

    The script's +-10 A/m predates the fixed accuracy 6 of loadDipoleKernel
    (maxwell_convolution.go:188, the script also reads a quantity name that no
    longer exists); Kernel_Arne at accuracy 6 is ~114 A/m off, and converges
    to -Ms/3 as the accuracy grows, which pins it to the closed form."""
    n = (32, 32, 32)
    prob = O.Problem(n, (5e-9,) * 3, msat=800e3, aex=0.0, alpha=0.02)
    m = np.zeros((3,) + n)
    m[2] = 1.0                                        # user x = internal component 2
    h, hd, _, _ = prob.heff(m, terms=True)
    avg = hd.reshape(3, -1).mean(axis=1)
    assert abs(avg[0]) < 10 and abs(avg[1]) < 10
    assert abs(avg[2] - (-800e3 / 3)) < 150
    errs = []
    for acc in (6, 12, 20):
        n = (16, 16, 16)
        p = O.pad_size(n)
        kh, _ = O.kernel_fft(p, K9=O.kernel_arne(p, (5e-9,) * 3, accuracy=acc))
        m = np.zeros((3,) + n)
        m[2] = 1.0
        errs.append(abs(800e3 * O.demag_conv(n, p, kh, m)[2].mean() + 800e3 / 3))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 30


def test_dipole_thin_film():
    """tests/dipole.py:9-33: 32x32x1 film, 5x5x2 nm, m = z, B_ext = (2,3,4) T:
    <B_z> within 1 % of -790000 mu0 + 4."""
    n = (1, 32, 32)
    cell = (2e-9, 5e-9, 5e-9)
    p = O.pad_size(n)
    K = O.kernel_arne(p, cell)
    kh, _ = O.kernel_fft(p, K9=K)
    m = np.zeros((3,) + n)
    m[0] = 1.0                                        # user z = internal component 0
    B = MU0 * 800e3 * O.demag_conv(n, p, kh, m)
    bext_user = (2.0, 3.0, 4.0)
    bz = B[0].mean() + bext_user[2]
    by = B[1].mean() + bext_user[1]
    bx = B[2].mean() + bext_user[0]
    good = -790000 * MU0 + 4
    assert abs((bz - good) / good) < 0.01
    assert abs(bx - 2.0) < 1e-9 and abs(by - 3.0) < 1e-9


# ------------------------------------------------------- local terms

def test_exchange_uniform_is_zero():
    n = (3, 5, 7)
    m = np.zeros((3,) + n)
    m[0], m[1], m[2] = 0.3, -0.4, np.sqrt(1 - 0.25)
    h = O.exchange6(n, (0, 0, 0), (1e-9, 2e-9, 3e-9), 1.0, m)
    assert np.abs(h).max() == 0.0


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_exchange_spin_wave(dim):
    """Periodic spin wave m = (cos kx, sin kx, 0) along one axis:
    H = -pre (2 - 2 cos k d)/d^2 m."""
    n = [2, 3, 4]
    n[dim] = 16
    n = tuple(n)
    cell = (2e-9, 3e-9, 4e-9)
    pbc = [0, 0, 0]
    pbc[dim] = 1
    k = 2 * np.pi * 3 / (n[dim] * cell[dim])
    x = (np.arange(n[dim]) * cell[dim])
    shape = [1, 1, 1]
    shape[dim] = n[dim]
    ph = np.broadcast_to((k * x).reshape(shape), n)
    m = np.stack([np.cos(ph), np.sin(ph), np.zeros(n)])
    pre = 2 * 1.3e-11 / (MU0 * 8e5)
    h = O.exchange6(n, pbc, cell, pre, m)
    want = -pre * (2 - 2 * np.cos(k * cell[dim])) / cell[dim] ** 2 * m
    assert np.abs(h - want).max() < 1e-9 * np.abs(want).max()


def test_exchange_neumann_edge_and_weights():
    # 2 cells along x: clamped neighbour of each cell is itself or the other
    n = (1, 1, 2)
    m = np.zeros((3,) + n)
    m[:, 0, 0, 0] = (1, 0, 0)
    m[:, 0, 0, 1] = (0, 1, 0)
    cell = (1e-9, 1e-9, 2e-9)
    h = O.exchange6(n, (0, 0, 0), cell, 1.0, m)
    d = (m[:, 0, 0, 1] - m[:, 0, 0, 0]) / cell[2] ** 2
    assert np.allclose(h[:, 0, 0, 0], d, rtol=1e-14)
    assert np.allclose(h[:, 0, 0, 1], -d, rtol=1e-14)
    # harmonic mean of Aex/Msat between the two cells
    aex = np.array([1.0, 3.0]).reshape(n)
    msat = np.array([2.0, 1.0]).reshape(n)
    h = O.exchange6(n, (0, 0, 0), cell, 1.0, m, msat_map=msat, aex_map=aex)
    l0, l1 = 0.5, 3.0
    w = 2 * l0 * l1 / (l0 + l1)
    assert np.allclose(h[:, 0, 0, 0], w * d, rtol=1e-14)


def test_uniaxial_closed_form():
    rng = np.random.default_rng(0)
    m = rng.normal(size=(3, 2, 3, 4))
    u = np.array([0.2, -0.5, 0.7])
    u /= np.linalg.norm(u)
    h = O.uniaxial(m, 2.5, u)
    mu = np.einsum("c...,c->...", m, u)
    assert np.allclose(h, 2.5 * mu[None] * u[:, None, None, None], rtol=1e-14)
    ms = rng.uniform(1, 2, size=(2, 3, 4))
    ms[0, 0, 0] = 0.0                               # zero Msat: divide by 1
    h2 = O.uniaxial(m, 2.5, u, msat_map=ms)
    msd = np.where(ms == 0, 1.0, ms)
    assert np.allclose(h2, h / msd[None], rtol=1e-14)


def test_mask_quantities_in_the_problem():
    """The MASK paths of B_ext, anisU, polarisation and xi: B_ext's map scales
    its uniform term cell by cell (vecMadd mul * map); anisU's map gives each
    cell its own axis u_c = mul_c * map_c (uniaxialanisotropy.cu:51-79);
    polarisation and xi enter the Zhang-Li term only through the cell's own
    prefactor, so a map equals, cell by cell, the uniform value it holds there."""
    rng = np.random.default_rng(5)
    n, cell = (2, 5, 6), (3e-9, 4e-9, 5e-9)
    m = O.normalize(rng.normal(size=(3,) + n))
    bext = (0.01, -0.02, 0.03)
    bmap = rng.uniform(-1, 2, size=(3,) + n)
    base = O.Problem(n, cell, demag=False)
    with_b = O.Problem(n, cell, demag=False, bext=bext, bext_map=bmap)
    want = np.array(bext)[:, None, None, None] / O.MU0 * bmap
    assert np.allclose(with_b.heff(m) - base.heff(m), want, rtol=1e-12, atol=1e-9)
    # anisU map against the closed form with each cell's axis
    umul = np.array([0.5, 2.0, -1.0])
    umap = rng.normal(size=(3,) + n)
    h = O.uniaxial(m, 2.5, umul, u_map=umap)
    uc = umul[:, None, None, None] * umap
    mu = (m * uc).sum(axis=0)
    assert np.allclose(h, 2.5 * mu[None] * uc, rtol=1e-13)
    pa = O.Problem(n, cell, demag=False, ku=5e5, anis_u=tuple(umul), anis_u_map=umap)
    pb = O.Problem(n, cell, demag=False, ku=5e5, anis_u=tuple(umul))
    ku2 = 2 * 5e5 / (O.MU0 * 8e5)
    da = pa.heff(m) - pb.heff(m)
    assert np.allclose(da, O.uniaxial(m, ku2, umul, u_map=umap) - O.uniaxial(m, ku2, umul),
                       rtol=1e-12, atol=1e-6)
    # polarisation / xi maps: cell by cell the uniform value
    j = (0.0, 3e11, 1e12)
    pol_map = rng.choice([0.3, 0.9], size=n)
    xi_map = rng.choice([0.5, 2.0], size=n)
    tau_map = O.Problem(n, cell, demag=False, j=j, pol=0.56, xi=0.05, pol_map=pol_map,
                        xi_map=xi_map).torque(m, base.heff(m))
    for pv in (0.3, 0.9):
        for xv in (0.5, 2.0):
            t = O.Problem(n, cell, demag=False, j=j, pol=0.56 * pv,
                          xi=0.05 * xv).torque(m, base.heff(m))
            sel = (pol_map == pv) & (xi_map == xv)
            assert np.allclose(tau_map[:, sel], t[:, sel], rtol=1e-12, atol=1e-12)


def test_torque_physics():
    rng = np.random.default_rng(1)
    m = O.normalize(rng.normal(size=(3, 4, 4, 4)))
    h = rng.normal(size=m.shape)
    t = O.torque(m, h, 0.1)
    assert np.abs((t * m).sum(axis=0)).max() < 1e-14
    t0 = O.torque(m, 3.0 * m, 0.1)
    assert np.abs(t0).max() < 1e-14
    # user space: m = x, H = y, alpha = 0 -> dm/dt = -gamma m x H = -gamma z.
    # Internal order (z, y, x): m = (0,0,1), H = (0,1,0), tau_internal[0] = user z.
    mi = np.array([0.0, 0.0, 1.0]).reshape(3, 1, 1, 1)
    hi = np.array([0.0, 1.0, 0.0]).reshape(3, 1, 1, 1)
    ti = O.torque(mi, hi, 0.0)[:, 0, 0, 0]
    assert np.allclose(ti, [-1.0, 0.0, 0.0])
    # damping pulls m towards H: m = x, H = y, alpha = 1 -> tau . H > 0
    ti = O.torque(mi, hi, 1.0)[:, 0, 0, 0]
    assert ti[1] > 0 and abs(ti[1] - 0.5) < 1e-15 and abs(ti[0] + 0.5) < 1e-15


def test_normalize():
    rng = np.random.default_rng(2)
    m = rng.normal(size=(3, 5, 5, 5))
    mask = np.ones((5, 5, 5))
    mask[1, 2, 3] = 0
    m[:, 4, 4, 4] = 0
    mn = O.normalize(m, mask)
    nrm = np.sqrt((mn ** 2).sum(axis=0))
    assert nrm[1, 2, 3] == 0 and nrm[4, 4, 4] == 0
    sel = np.ones_like(nrm, bool)
    sel[1, 2, 3] = sel[4, 4, 4] = False
    assert np.abs(nrm[sel] - 1).max() < 1e-15


def _zhangli_user_numpy(m_user, cell_user, j_user, pred, xi, alpha):
    """Independent rendition in *user* (x,y,z) coordinates with np.cross and
    the textbook Zhang-Li form m x ((u.grad)m x m), (u.grad)m x m."""
    nx = m_user.shape[1]
    u = np.asarray(j_user) / np.linalg.norm(j_user)
    # only x-current used here; 5-point derivative, out-of-range -> centre
    D = np.zeros_like(m_user)
    for i in range(nx):
        idx = [i - 2, i - 1, i + 1, i + 2]
        idx = [q if 0 <= q < nx else i for q in idx]
        D[:, i] = (m_user[:, idx[0]] - 8 * m_user[:, idx[1]] + 8 * m_user[:, idx[2]]
                   - m_user[:, idx[3]]) / (12 * cell_user[0])
    D *= u[0]
    b = np.linalg.norm(j_user) * pred / ((1 + xi * xi) * (1 + alpha * alpha))
    Dxm = np.cross(D, m_user, axis=0)
    return b * (1 + xi * alpha) * np.cross(m_user, Dxm, axis=0) + b * (xi - alpha) * Dxm


def test_zhangli_against_user_space_rendition():
    rng = np.random.default_rng(5)
    nx, ny, nz = 12, 3, 2
    mu = O.normalize(rng.normal(size=(3, nx, ny, nz)))          # user (c, x, y, z)
    cell_user = (3e-9, 4e-9, 5e-9)
    want = _zhangli_user_numpy(mu, cell_user, (2e12, 0, 0), 1.7e-11, 0.05, 0.1)
    # to internal: comps (z,y,x) and axes (z,y,x)
    mi = np.ascontiguousarray(mu[::-1].transpose(0, 3, 2, 1))
    got = O.zhangli((nz, ny, nx), (0, 0, 0), cell_user[::-1], mi, (0, 0, 2e12), 1.7e-11, 0.05, 0.1)
    got_user = got[::-1].transpose(0, 3, 2, 1)
    assert np.abs(got_user - want).max() < 1e-10 * np.abs(want).max()


def test_zhangli_uniform_is_zero():
    n = (2, 3, 8)
    m = np.zeros((3,) + n)
    m[1] = 1.0
    s = O.zhangli(n, (0, 0, 0), (1e-9,) * 3, m, (0, 0, 1e12), 1e-11, 0.05, 0.1)
    assert np.abs(s).max() == 0.0


# -------------------------------------------------------------- RK12

def _macrospin(alpha, b=0.05, T=2e-10, maxerr=1e-5):
    prob = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=alpha, demag=False,
                     bext=(b, 0.0, 0.0))                         # B along user z
    prob.set("maxerr", maxerr)
    th0 = 0.6
    m = np.zeros((3, 1, 1, 2))
    m[0], m[2] = np.cos(th0), np.sin(th0)                       # polar angle th0 from z
    while prob.get("t") < T:
        prob.step(m)
    return prob, m, th0, b


@pytest.mark.parametrize("alpha", [0.0, 0.1, 1.0])
def test_rk12_macrospin_matches_analytic_llg(alpha):
    prob, m, th0, b = _macrospin(alpha)
    t = prob.get("t")
    H = b / MU0
    gp = O.GAMMA0 / (1 + alpha * alpha)
    th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * gp * H * t))
    # dm/dt = -gamma' m x H with H = H z: d(mx)/dt = -gamma' H my, d(my)/dt = +gamma' H mx
    phi = gp * H * t
    mz_user, mx_user, my_user = m[0, 0, 0, 0], m[2, 0, 0, 0], m[1, 0, 0, 0]
    assert abs(mz_user - np.cos(th)) < 2e-3
    assert abs(mx_user - np.sin(th) * np.cos(phi)) < 2e-3
    assert abs(my_user - np.sin(th) * np.sin(phi)) < 2e-3


def test_rk12_controller_rules():
    # a huge first step at a tiny tolerance must be rejected and cut by the
    # cube-root rule, clamped to a factor 0.1 (solve_rk12.go:161-203)
    prob = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=0.1, demag=False,
                     bext=(0.1, 0, 0))
    m = np.zeros((3, 1, 1, 2))
    m[2] = 1.0
    prob.set("dt", 1e-10)
    prob.set("mindt", 1e-30)
    prob.set("maxerr", 1e-7)
    prob.step(m)
    assert prob.get("badsteps") >= 1
    assert prob.get("err") <= 1e-7
    # every rejection cuts dt by 0.8 (maxerr/err)^(1/3) < 0.8, never below 0.1
    b = prob.get("badsteps")
    assert 1e-10 * 0.1 ** b * 0.9999999 <= prob.get("t") <= 1e-10 * 0.8 ** b
    # accepted at mindt even when bad (solve_rk12.go:202)
    prob2 = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=0.1, demag=False,
                      bext=(0.1, 0, 0))
    m2 = np.zeros((3, 1, 1, 2))
    m2[2] = 1.0
    prob2.set("dt", 1e-11)
    prob2.set("mindt", 1e-11)
    prob2.set("maxerr", 1e-12)
    prob2.step(m2)
    assert prob2.get("step") == 1 and prob2.get("badsteps") == 1
    assert prob2.get("t") == 1e-11 and prob2.get("dt") == 1e-11
    # good step: growth factor 0.8 sqrt(maxerr/err) clamped to 1.5
    prob3 = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=0.1, demag=False,
                      bext=(0.01, 0, 0))
    m3 = np.zeros((3, 1, 1, 2))
    m3[2] = 1.0
    prob3.set("dt", 1e-15)
    prob3.step(m3)
    assert prob3.get("dt") == 1e-15 * 1.5
    # cannot converge -> error (the reference panics)
    prob4 = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=0.1, demag=False,
                      bext=(0.1, 0, 0))
    m4 = np.zeros((3, 1, 1, 2))
    m4[2] = 1.0
    prob4.set("dt", 1e-9)
    prob4.set("mindt", 1e-300)
    prob4.set("maxerr", 1e-300)
    with pytest.raises(RuntimeError):
        prob4.step(m4)


# ----------------------------------------------------------- fixtures

def _load_sp5():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "sp5_m.txt")
    return np.loadtxt(path, comments="#")


@pytest.mark.slow
def test_sp5_prefix_matches_reference_trace():
    """certification/stdproblem5/m.txt (mumax2's own output, recipe
    examples/stdproblem5.py): relax a vortex, then Zhang-Li current 1e12 A/m^2
    along x; the oracle's <m>(t) must follow the reference's trace."""
    ref = _load_sp5()
    prob, m, rows = sp5_run(t_end=0.3e-9)
    for t, avg in rows:
        i = np.argmin(np.abs(ref[:, 0] - t))
        assert abs(ref[i, 0] - t) < 2e-12
        # vortex displacement <mx>, <my> (amplitude ~0.2; the trace was made
        # by an earlier mumax2 build whose relaxed core <mz> is 1.2 % lower,
        # so the orbit differs by a few %) and core <mz>; a sign error in the
        # spin-transfer term sends the core the other way (error ~0.4)
        assert np.abs(avg[:2] - ref[i, 1:3]).max() < 1.5e-2, (t, avg, ref[i])
        assert abs(avg[2] - ref[i, 3]) < 1e-3, (t, avg, ref[i])


def sp5_run(t_end, Problem=None):
    """examples/stdproblem5.py in internal order on the oracle."""
    nx, ny, nz = 32, 32, 1
    cs = (100e-9 / nx, 100e-9 / ny, 10e-9 / nz)
    n = (nz, ny, nx)
    prob = O.Problem(n, cs[::-1], msat=800e3, aex=1.3e-11, alpha=1.0, gamma=2.211e5,
                     j=(0.0, 0.0, 0.0), pol=1.0, xi=0.05)
    prob.P.zhangli = 0
    prob.set("dt", 1e-15)
    prob.set("maxdt", 1e-12)
    prob.set("maxerr", 1e-4)
    m = np.zeros((3,) + n)
    for a in range(nx):
        for b in range(ny):
            xx = a * cs[0] - 50e-9
            yy = 50e-9 - b * cs[1]
            m[2, 0, b, a] = yy       # user x
            m[1, 0, b, a] = xx       # user y
            m[0, 0, b, a] = 40e-9    # user z
    # examples/stdproblem5.py relaxes with run_until_smaller('maxtorque',
    # 1e-3 gamma Ms); the reference trace starts from a fully centred vortex
    # (<mx>, <my> ~ 1e-8), which that criterion does not guarantee for an
    # FP64 run (its torque noise drops below the threshold early), so relax
    # for a fixed 3 ns instead
    while prob.get("t") < 3e-9:
        prob.step(m)
    prob.P.zhangli = 1
    prob.set_param("alpha", 0.1)
    prob.set_param("jmul", (0.0, 0.0, 1e12))
    prob.set_param("xi", 0.05)
    prob.set_param("pol", 1.0)
    prob.set("dt", 1e-15)
    prob.set("t", 0.0)
    period = 50e-12
    count = 0
    rows = [(0.0, _user_avg(prob, m))]
    while prob.get("t") < t_end:
        prob.step(m)
        if prob.get("t") - count * period >= period:            # notify_tabulate.go:23-28
            rows.append((prob.get("t"), _user_avg(prob, m)))
            count += 1
    return prob, m, rows


def _user_avg(prob, m):
    a = prob.average(prob.normalized(m))
    return np.array([a[2], a[1], a[0]])
