"""The implicit-solver oracle (oracle/solvers.py) against closed forms.

* linear system y' = A y (A acting on the component axis): the BDF-Euler
  fixed point is the implicit-Euler solution (I - hA)^-1 y0, the AM12
  predictor and corrector fixed points are implicit Euler and the trapezoidal
  rule (I - hA/2)^-1 (I + hA/2) y0, and the next AM12 step is
  dt sqrt(0.8 / err) with err the scaled max-norm of their difference
  (solve_am01.go:252-262) -- a dropped factor 1/2, a sign error or another
  step rule fails one of these;
* a non-contracting step (|hA| > 1) is cut by the alpha rule and retried;
* a macrospin follows the analytic LLG solution (first order for BDF-Euler,
  second order with step control for AM12);
* BDF-Euler that cannot converge in m_maxIterations raises, as the reference
  panics.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import solvers as S

A = np.array([[-0.3, 0.4, 0.0], [-0.4, -0.3, 0.1], [0.0, -0.1, -0.2]]) * 1e10


def _lin_rhs(y):
    return y, np.einsum("ij,j...->i...", A, y)


def _y0():
    rng = np.random.default_rng(4)
    return rng.normal(size=(3, 2, 3, 4))


def _apply(M, y):
    return np.einsum("ij,j...->i...", M, y)


def _solve(M, y):
    return np.einsum("ij,j...->i...", np.linalg.inv(M), y)


def test_bdf_euler_is_implicit_euler_on_a_linear_system():
    y0 = _y0()
    dt = 5e-12                                  # |h A| ~ 0.03
    s = S.BDFEuler(dt)
    y = s.step(y0.copy(), _lin_rhs, 1.0)
    h = S.f32(dt)
    want = _solve(np.eye(3) - h * A, y0)
    assert np.abs(y - want).max() < 2e-5 * np.abs(want).max()
    # explicit Euler would be off by O(h^2 A^2 y0) ~ 1e-3
    expl = y0 + h * _apply(A, y0)
    assert np.abs(expl - want).max() > 50 * np.abs(y - want).max()
    assert s.t == dt and s.step_count == 1 and 2 <= s.iterations <= 4


def test_am12_step_is_trapezoid_and_step_rule():
    y0 = _y0()
    # small enough that the first predictor iterate is within 0.1 (otherwise
    # alpha = err / 0.1 > 1 flags the iterator and the step is cut first)
    dt = 1e-12
    s = S.AM12(dt)
    y = s.step(y0.copy(), _lin_rhs, 1.0)
    h = S.f32(dt)
    ie = _solve(np.eye(3) - h * A, y0)
    tr = _solve(np.eye(3) - 0.5 * h * A, _apply(np.eye(3) + 0.5 * h * A, y0))
    scale = s.abs_err + np.abs(y0).max() * s.rel_err
    assert np.abs(y - tr).max() < 0.2 * scale
    assert np.abs(ie - tr).max() > 20 * np.abs(y - tr).max()   # the corrector moved
    assert s.t == dt and s.badsteps == 0
    err = s.norm(tr, ie)
    assert abs(s.err - err) < 0.05 * err
    assert abs(s.dt - dt * math.sqrt(0.8 / err)) < 0.03 * s.dt


def test_am12_non_contracting_step_is_cut_by_alpha_rule():
    y0 = _y0()
    dt = 4e-10                                  # |h A| ~ 2: fixed point diverges
    s = S.AM12(dt)
    s.step(y0.copy(), _lin_rhs, 1.0)
    assert s.t < dt and s.badsteps == 0         # accepted with a smaller step
    # every cut is 0.5 dt or dt sqrt(0.6 / alpha) < dt
    assert s.t <= dt * math.sqrt(0.6)


def _macrospin(alpha=0.1, b=0.05):
    prob = O.Problem((1, 1, 2), (5e-9,) * 3, msat=8e5, aex=0.0, alpha=alpha, demag=False,
                     bext=(b, 0.0, 0.0))                         # B along user z
    th0 = 0.6
    m = np.zeros((3, 1, 1, 2))
    m[0], m[2] = np.cos(th0), np.sin(th0)
    return prob, m, th0, b


def _analytic(t, th0, b, alpha):
    H = b / O.MU0
    gp = O.GAMMA0 / (1 + alpha * alpha)
    th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * gp * H * t))
    phi = gp * H * t
    return np.cos(th), np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi)


@pytest.mark.parametrize("solver", ["bdf", "am12"])
def test_macrospin_matches_analytic_llg(solver):
    alpha = 0.1
    prob, m, th0, b = _macrospin(alpha)
    rhs = S.llg_rhs(prob)
    if solver == "bdf":
        s = S.BDFEuler(2e-14)
        T = 5e-11
    else:
        s = S.AM12(1e-13, max_abs_err=1e-6, max_rel_err=1e-5)
        T = 2e-10
    nstep = 0
    while s.t < T:
        m = s.step(m, rhs, O.GAMMA0)
        nstep += 1
        assert nstep <= 5000, "step size collapsed (t = %g)" % s.t
    m = O.normalize(m)
    mz, mx, my = _analytic(s.t, th0, b, alpha)
    err = max(abs(m[0, 0, 0, 0] - mz), abs(m[2, 0, 0, 0] - mx), abs(m[1, 0, 0, 0] - my))
    # first order: error ~ dt gamma H * (phase) ; second order with control
    assert err < (3e-3 if solver == "bdf" else 1e-3), err


def test_macrospin_bdf_euler_is_first_order():
    alpha = 0.1
    errs = []
    for dt in (4e-14, 2e-14):
        prob, m, th0, b = _macrospin(alpha)
        rhs = S.llg_rhs(prob)
        s = S.BDFEuler(dt)
        for _ in range(int(round(4e-11 / dt))):
            m = s.step(m, rhs, O.GAMMA0)
        m = O.normalize(m)
        mz, mx, my = _analytic(s.t, th0, b, alpha)
        errs.append(math.hypot(m[2, 0, 0, 0] - mx, m[1, 0, 0, 0] - my))
    assert 1.6 < errs[0] / errs[1] < 2.4, errs


def test_bdf_euler_gives_up():
    y0 = _y0()
    s = S.BDFEuler(2e-10, max_iter=3)           # |h A| ~ 1: no convergence in 3
    with pytest.raises(S.ConvergenceError):
        s.step(y0.copy(), _lin_rhs, 1.0)
