"""FP64 restatement of the implicit solvers (test infrastructure only).

* solver/bdf-euler: BDFEuler.Step, engine/solve_euler_backward.go:29-91
* solver/am12: BDFAM12.Step, engine/solve_am01.go:45-288

The equation dy/dt = mul * f(y) is given as a function ``rhs(y) -> (y', f)``:
the engine updates y's own dependencies before it evaluates the right-hand
side (for m the normUpdater normalises it in place, update_norm.go:20-22),
so ``rhs`` returns the y it used together with f.  For the LLG equation y = m,
f = the torque array (tau / gamma) and mul = gamma (mod_llg.go:59-70).

The float32 conversions the reference makes on the scalar multipliers
(gpu.Madd takes float64 and converts, libmumax2.go:167-176; the AM12
corrector multiplies in float32, solve_am01.go:200-201) are reproduced; the
arrays stay FP64.

Pinned in tests/test_oracle_solvers.py by closed forms: on a linear system the
fixed points are the implicit-Euler and trapezoidal solutions, the AM12 step
size follows sqrt(0.8 / err) of their difference, and a macrospin follows the
analytic LLG solution.
"""
import math

import numpy as np


def f32(x):
    return float(np.float32(x))


class ConvergenceError(RuntimeError):
    pass


def _maxdiff(a, b):
    return float(np.abs(a - b).max())


def _maxsum(a, b):
    return float(np.abs(a + b).max())


class BDFEuler:
    """solve_euler_backward.go: state t, dt, step, bdf_iterations."""

    def __init__(self, dt, max_iter=3, max_iter_err=1e-5):
        self.t = 0.0
        self.dt = dt
        self.step_count = 0
        self.iterations = 0
        self.max_iter = max_iter
        self.max_iter_err = max_iter_err

    def step(self, y, rhs, mul):
        y, dy = rhs(y)                       # equation input Update()
        dt = self.dt
        self.t += dt
        self.iterations = 0
        h = f32(dt * mul)
        y0 = y.copy()
        y = y0 + h * dy                      # explicit Euler (zero order)
        y, dy = rhs(y)
        self.iterations += 1
        err, it = 1.0e38, 0
        while err > self.max_iter_err:
            yb = y0 + h * dy
            err = _maxdiff(y, yb)
            s = _maxsum(y, yb)
            if s > 0.0:
                err = err / s
            it += 1
            self.iterations += 1
            y = yb
            y, dy = rhs(y)
            if it > self.max_iter:
                raise ConvergenceError("The BDF iterator cannot converge")
        self.step_count += 1
        return y


class AM12:
    """solve_am01.go: adaptive predictor (implicit Euler) / corrector
    (trapezoidal), both by fixed-point iteration."""

    MAX_TRY, HEADROOM, MAX_ITER_ERR, MAX_ITER, ALPHA_REF = 10, 0.8, 0.1, 5, 0.6

    def __init__(self, dt, max_abs_err=1e-4, max_rel_err=1e-3, mindt=1e-38, maxdt=1e38):
        self.t = 0.0
        self.dt = dt
        self.step_count = 0
        self.iterations = 0
        self.badsteps = 0
        self.err = 0.0
        self.alpha = 0.0
        self.abs_err = max_abs_err
        self.rel_err = max_rel_err
        self.mindt = mindt
        self.maxdt = maxdt

    def norm(self, y, yb):
        """srCOMP * sqrt(sum_p (MaxDiff_p / (absErr + MaxAbs_p(yb) relErr))^2)"""
        e = 0.0
        ncomp = y.shape[0]
        for p in range(ncomp):
            d = float(np.abs(y[p] - yb[p]).max())
            mx = float(np.abs(yb[p]).max())
            e += (d / (self.abs_err + mx * self.rel_err)) ** 2
        return math.sqrt(e) / math.sqrt(ncomp)

    def step(self, y, rhs, mul):
        t0 = self.t
        self.badsteps = 0
        self.iterations = 0
        y, dy0 = rhs(y)
        y0 = y.copy()
        tries, restrict, guard = 0, False, 0
        while True:
            # the reference retries without bound; 1000 cuts shrink dt by
            # more than 1e100, so a step that still fails raises here
            guard += 1
            if guard > 1000:
                raise ConvergenceError("am12: no step size converges")
            dt = self.dt
            bad_step = bad_iter = False
            y = y0 + f32(dt * mul) * dy0     # explicit Euler start
            self.iterations += 1
            self.t = t0 + dt
            y, dy = rhs(y)
            dy_euler = dy.copy()
            er, it, alpha = self.MAX_ITER_ERR, 0, 0.0
            while True:                      # predictor: implicit Euler
                yb = y0 + f32(dt * mul) * dy
                terr = self.norm(y, yb)
                alpha = terr / er
                self.alpha = alpha
                er = terr
                y = yb
                it += 1
                self.iterations += 1
                if terr <= self.MAX_ITER_ERR:
                    break
                if alpha >= 1.0 or it > self.MAX_ITER:
                    bad_iter = True
                    break
                y, dy = rhs(y)
            if bad_iter and tries == self.MAX_TRY - 1:
                raise ConvergenceError("The BDF Euler iterator cannot converge")
            if bad_iter:
                h_alpha = 0.5 * dt
                if alpha > self.ALPHA_REF:
                    h_alpha = dt * (self.ALPHA_REF / alpha) ** 0.5
                self.dt = h_alpha
                restrict = True
                continue
            y1 = y.copy()
            dy = dy_euler.copy()             # restore dy of the explicit Euler point
            er, it = self.MAX_ITER_ERR, 0
            while True:                      # corrector: trapezoidal
                hf = f32(dt * mul)
                yb = y0 + f32(0.5 * hf) * (dy + dy0)
                terr = self.norm(y, yb)
                alpha = terr / er
                er = terr
                y = yb
                it += 1
                self.iterations += 1
                if terr <= self.MAX_ITER_ERR:
                    break
                if alpha >= 1.0 or it > self.MAX_ITER:
                    bad_iter = True
                    break
                y, dy = rhs(y)
            if bad_iter and tries == self.MAX_TRY - 1:
                raise ConvergenceError("The BDF Trapezoidal iterator cannot converge")
            if bad_iter:
                h_alpha = 0.5 * dt
                if alpha > self.ALPHA_REF:
                    h_alpha = dt * (self.ALPHA_REF / alpha) ** 0.5
                self.dt = h_alpha
                continue
            terr = self.norm(y, y1)
            if terr > 1.0:
                self.badsteps += 1
                bad_step = True
            self.err = terr
            h_r = dt * (self.HEADROOM / terr) ** 0.5
            new_dt = h_r
            if restrict:
                new_dt = min(h_r, dt * (self.ALPHA_REF / self.alpha) ** 0.5)
                restrict = False
            new_dt = min(max(new_dt, self.mindt), self.maxdt)
            self.dt = new_dt
            if not bad_step or new_dt == self.mindt:
                break
            if tries > self.MAX_TRY:
                raise ConvergenceError("The solver cannot converge after %d badsteps" % self.MAX_TRY)
            tries += 1
        self.step_count += 1
        return y


def llg_rhs(prob):
    """rhs for the LLG equation of an oracle Problem: normalise m, torque."""
    from oracle import oracle as O

    def rhs(m):
        mn = O.normalize(m)
        return mn, prob.torque(mn, prob.heff(mn))
    return rhs
