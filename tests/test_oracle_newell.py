"""Pins of the oracle's FP64 Newell kernel (oracle/newell.c).

The Newell tensor has no counterpart in the reference (its dipole kernel is
Kernel_Arne, modules/kernel_magnetostatic.go:30-181).  What fixes it here:

* the same analytic formulas evaluated in __float128 (method 2), which do not
  suffer the sixth-order-difference cancellation at these distances: the
  production method (analytic FP64 near, Gauss rule of the cell-pair offset
  density far) must agree to 1e-9;
* the trace identity K_xx + K_yy + K_zz = -1 at R = 0 and 0 elsewhere
  (a dropped or mis-permuted f term breaks it), the cube self term -1/3;
* the point-dipole limit far away (normalisation and sign of the far rule);
* the mean field of a uniformly magnetised box: for a cube of n^3 cells the
  average of the Newell-convolved field is exactly -Ms/3 (tests/demag3D.py's
  known answer, which Kernel_Arne reaches only approximately);
* the reference's own Kernel_Arne converging to it with the integration
  accuracy (pins sign, axis order and component order against the reference).
"""
import numpy as np
import pytest

from oracle import oracle as O

CUBE = (2e-9, 2e-9, 2e-9)
FILM = (3e-9, 3.90625e-9, 3.90625e-9)      # internal (z, y, x) of the film cell
SURR = (3e-9, 4e-9, 4e-9)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [4, 5])
def test_rule_moments(n):
    t, w = O.newell_rule(n)
    assert abs(w.sum() - 1) < 1e-15
    for k in range(1, 2 * n):
        want = 0.0 if k % 2 else 2.0 / ((k + 1) * (k + 2))
        assert abs((w * t ** k).sum() - want) < 1e-15, k
    # and not beyond: an n-point rule cannot integrate t^(2n) (order pinned)
    assert abs((w * t ** (2 * n)).sum() - 2.0 / ((2 * n + 1) * (2 * n + 2))) > 1e-6
    assert np.all(w > 0) and np.all(np.abs(t) < 1)


def test_cube_self_term_and_trace():
    for method in (0, 1, 2):
        k = O.newell_point((0, 0, 0), CUBE, method)
        assert np.allclose(k[:3], -1 / 3, rtol=0, atol=1e-14)
        assert np.allclose(k[3:], 0, atol=1e-15)
    rng = np.random.default_rng(11)
    for d in (CUBE, FILM, SURR):
        k0 = O.newell_point((0, 0, 0), d, 0)
        assert abs(k0[:3].sum() + 1) < 1e-13
        for _ in range(40):
            R = rng.integers(-12, 13, size=3) * np.array(d)
            if not R.any():
                continue
            k = O.newell_point(R, d, 0)
            # zero to the method's accuracy (a structural mistake is O(1))
            assert abs(k[:3].sum()) < 1e-9 * np.abs(k).max() + 1e-16


@pytest.mark.parametrize("d", [CUBE, FILM, SURR])
def test_production_method_against_quad_precision(d):
    """Analytic near field and the far rule, on both sides of the switch
    (8 largest cell sizes), against the analytic formulas in __float128."""
    dm = max(d)
    rng = np.random.default_rng(5)
    pts = []
    for r in (0.0, 1.0, 1.5, 3.0, 5.0, 7.5, 7.99, 8.01, 9.0, 12.0, 20.0, 40.0, 80.0):
        for _ in range(6):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            pts.append(np.round(v * r * dm / np.array(d)) * np.array(d))
    worst = 0.0
    for R in pts:
        q = O.newell_point(R, d, 2)
        worst = max(worst, _rel(O.newell_point(R, d, 0), q))
    assert worst < 1e-9, worst


def test_fp64_analytic_alone_fails_far_away():
    """The reason for the switch: the plain FP64 formulas lose ~all digits at
    100 cells, where the far rule is exact to rounding."""
    R = np.array([100.0, 37.0, 61.0]) * 2e-9
    q = O.newell_point(R, CUBE, 2)
    assert _rel(O.newell_point(R, CUBE, 1), q) > 1e-6
    assert _rel(O.newell_point(R, CUBE, 0), q) < 1e-13


def test_point_dipole_limit():
    """Far away K -> (V/4pi)(3 R R^T - |R|^2 I)/|R|^5 (B = mu0 Ms K m)."""
    d = SURR
    V = np.prod(d)
    for R in (np.array([300, 0, 0]) * 3e-9, np.array([0, 170, 200]) * 4e-9,
              np.array([100, 120, -90]) * 4e-9):
        r = np.linalg.norm(R)
        T = V / (4 * np.pi) * (3 * np.outer(R, R) - r * r * np.eye(3)) / r ** 5
        want = np.array([T[0, 0], T[1, 1], T[2, 2], T[1, 2], T[0, 2], T[0, 1]])
        assert _rel(O.newell_point(R, d, 0), want) < 1e-4


def test_kernel_layout_and_symmetry():
    n = (3, 4, 6)
    p = O.pad_size(n)
    K = O.kernel_newell(p, SURR)
    assert K.shape == (9,) + tuple(p)
    np.testing.assert_array_equal(K[6], K[3])
    np.testing.assert_array_equal(K[7], K[4])
    np.testing.assert_array_equal(K[8], K[5])
    # the p/2 planes are left zero (displacement range of Kernel_Arne)
    assert not K[:, p[0] // 2].any() and not K[:, :, p[1] // 2].any()
    assert not K[:, :, :, p[2] // 2].any()
    for (i, j, k) in [(1, 2, 3), (p[0] - 1, 1, p[2] - 2), (2, p[1] - 3, 0)]:
        disp = [q if q < pp // 2 else q - pp for q, pp in zip((i, j, k), p)]
        want = O.newell_point(np.array(disp) * np.array(SURR), SURR, 0)
        # the table is built at |R| with signs applied afterwards: same value up
        # to the summation order of the 27-point difference
        np.testing.assert_allclose(K[:6, i, j, k], want, rtol=1e-11, atol=1e-15 * np.abs(want).max())
    # even in every axis for the diagonal, odd in both axes of an off-diagonal pair
    def mirror(a, axis):
        return np.roll(np.flip(a, axis=axis), 1, axis=axis)
    pairs = {3: (1, 2), 4: (0, 2), 5: (0, 1)}
    for c in range(6):
        for ax in range(3):
            sgn = -1 if c >= 3 and ax in pairs[c] else 1
            np.testing.assert_allclose(mirror(K[c], ax), sgn * K[c], rtol=0,
                                       atol=1e-15 * np.abs(K[c]).max())


@pytest.mark.parametrize("n", [(4, 4, 4), (6, 6, 6)])
def test_uniform_cube_mean_field_is_minus_one_third(n):
    """tests/demag3D.py: a uniformly magnetised cube has <B> = -mu0 Ms / 3.
    With the Newell kernel the cell-averaged fields sum to the exact box
    average, so this holds to rounding (Kernel_Arne: to ~1e-3 at accuracy 6)."""
    p = O.pad_size(n)
    kh, bad = O.kernel_fft(p, K9d=O.kernel_newell(p, CUBE))
    assert bad < 1e-10
    for c in range(3):
        m = np.zeros((3,) + n)
        m[c] = 1.0
        out = O.demag_conv(n, p, kh, m)
        assert abs(out[c].mean() + 1 / 3) < 1e-12
        for o in range(3):
            if o != c:
                assert abs(out[o].mean()) < 1e-12


def test_uniform_prism_mean_field_is_box_demag_factor():
    """For an a x b x c prism of cells, <K> over the prism equals the self
    term of one cell of size (a dx, b dy, c dz) (the box demag tensor,
    Aharoni 1998), evaluated in __float128."""
    n = (2, 5, 7)
    d = SURR
    p = O.pad_size(n)
    kh, _ = O.kernel_fft(p, K9d=O.kernel_newell(p, d))
    box = O.newell_point((0, 0, 0), tuple(ni * di for ni, di in zip(n, d)), 2)
    for c in range(3):
        m = np.zeros((3,) + n)
        m[c] = 1.0
        out = O.demag_conv(n, p, kh, m)
        assert abs(out[c].mean() - box[c]) < 1e-12


def test_kernel_arne_converges_to_newell():
    """The reference's Kernel_Arne integrates the same cell-pair average
    numerically; with rising accuracy it must approach the Newell kernel on
    every stored component (pins sign, axis and component order)."""
    n = (2, 3, 5)
    d = FILM
    p = O.pad_size(n)
    Kn = O.kernel_newell(p, d)[:6]
    errs, worst = [], []
    for acc in (3, 6, 12, 24):
        Ka = O.kernel_arne(p, d, accuracy=acc)[:6].astype(np.float64)
        errs.append(np.sqrt(((Ka - Kn) ** 2).mean()) / np.abs(Kn).max())
        worst.append(np.abs(Ka - Kn).max() / np.abs(Kn).max())
    # rms over the kernel (relative to its largest entry) falls steadily
    # (2.1e-4, 5.9e-5, 3.6e-5, 2.5e-5); the max sits in the self term, which
    # Kernel_Arne's charge integration resolves slowly (~5e-4 at accuracy 6..24)
    assert errs[0] > errs[1] > errs[2] > errs[3], errs
    assert errs[-1] < 5e-5 and max(worst[1:]) < 1e-3, (errs, worst)


def test_newell_khat_matches_kernel_fft():
    """The memory-lean transform (octant table -> one component at a time)
    equals the transform of the full real-space kernel."""
    for n, d in (((3, 4, 6), SURR), ((1, 8, 12), FILM)):
        p = O.pad_size(n)
        a, bad_a = O.kernel_fft(p, K9d=O.kernel_newell(p, d))
        b, bad_b = O.newell_khat(p, d)
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-14 * np.abs(a).max())
        assert bad_b < 1e-12
