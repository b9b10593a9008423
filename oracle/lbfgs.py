"""ORACLE — test infrastructure only (only tests/ import this, as the checker).

numpy (FP64) restatement of the reference's l-BFGS solver with the regular-Wolfe line
search, the optimiser of the elastic mixed-precision FWI configuration:
  lbfgs::run / updateSY / computeHg ..... src/cpp/nlsolver.cpp:1158-1359
  regular_wolfe::lineSearch / zoom ...... src/cpp/nlsolver.cpp:185-423
  quadratic / cubic step interpolation .. src/cpp/nlsolver.cpp:10-36 (ZERO 1e-10, STOL 1e-3)
  defaults .............................. src/include/param.hpp:55-62, nlsolver.hpp:99
The problem is a callable m -> (f, g): like nlls_fwi::res(), which computes the
residual and the gradient together (src/cpp/optimization.cpp:5-197).
"""
from __future__ import annotations

import math

import numpy as np

ZERO = 1e-10
STOL = 1e-3


def quadratic(f, df, stp0, f0):
    if f0 <= f + df * stp0:
        ans = stp0 * 1.1
    else:
        ans = -df * stp0 * stp0 / (2 * (f0 - f - df * stp0))
    if ans / stp0 < STOL or abs(ans - stp0) / stp0 < STOL:
        ans = stp0 / 2.0
    return ans


def cubic(f, df, stp0, f0, stp1, f1):
    x1 = f1 - f - df * stp1
    x2 = f0 - f - df * stp0
    x3 = stp0 * stp0 * stp1 * stp1 * (stp1 - stp0)
    a = (stp0 * stp0 * x1 - stp1 * stp1 * x2) / x3
    b = (-stp0 * stp0 * stp0 * x1 + stp1 * stp1 * stp1 * x2) / x3
    delta = b * b - 3 * a * df
    if f1 < f0:
        step, func = stp1, f1
    else:
        step, func = stp0, f0
    if abs(a) < ZERO or delta < 0:
        ans = quadratic(f, df, step, func)
    else:
        ans = (-b + math.sqrt(delta)) / (3 * a)
    if ans / step < STOL or abs(step - ans) / step < STOL:
        ans = step / 2.0
    elif abs(stp0 + stp1 - step - ans) / (stp0 + stp1 - step) < STOL:
        ans = (stp0 + stp1 - step) / 2.0
    return ans


class Problem:
    """Holds the current model, its objective and gradient (the optimization object)."""

    def __init__(self, fn, m):
        self.fn = fn
        self.m = np.array(m, dtype=np.float64)
        self.f = None
        self.g = None
        self.evals = 0

    def res(self):
        self.f, g = self.fn(self.m)
        self.g = np.array(g, dtype=np.float64)
        self.evals += 1


class RegularWolfe:
    def __init__(self, c1=1e-4, c2=0.9, a0=1.0, a1=0.0, step_max=1e6, version=0):
        assert 0 < c1 < c2 < 1
        self.c1, self.c2, self.a0, self.a1, self.stp_max, self.version = c1, c2, a0, a1, step_max, version
        self.flag_g = False
        self.reset()

    def reset(self):
        self.stp0 = self.stp1 = self.stp = self.f = self.df = self.f0 = self.f1 = 0.0

    def line_search(self, prob, m0, p, it, max_trial, cnt):
        df0 = self.df
        self.f = prob.f
        self.df = float(np.dot(prob.g, p))
        trial = 0
        self.stp0 = 0.0
        if it == 0:
            self.f0 = self.f
            self.stp1 = self.a0 / cnt["gnorm"]
            if self.a1 > 0:
                self.stp1 = max(self.stp1, self.a1 * float(np.linalg.norm(prob.m)) / cnt["gnorm"])
        else:
            self.stp1 = self.stp * df0 / self.df
            if self.version == 1:
                self.stp1 = min(1.0, 1.00001 * 2 * (self.f - self.f0) / self.df)
        alpha = self.stp1
        while True:
            if trial >= max_trial:
                return False
            prob.m = m0 + alpha * p
            prob.res()
            cnt["feval"] += 1
            trial += 1
            f = prob.f
            self.f1 = f
            if f != f:
                raise RuntimeError("the objective function is NaN")
            if f <= ZERO:
                self.stp, self.f0, self.f = alpha, self.f, f
                return False
            if self.f1 > self.f + self.c1 * alpha * self.df or (self.f1 >= self.f0 and trial > 1):
                return self._zoom(False, prob, m0, p, trial, max_trial, cnt)
            cnt["geval"] += 1
            cnt["gnorm"] = float(np.linalg.norm(prob.g))
            df1 = float(np.dot(prob.g, p))
            if df1 >= self.c2 * self.df:
                self.stp, self.f0, self.f = alpha, self.f, f
                return True
            if df1 >= 0:
                return self._zoom(True, prob, m0, p, trial, max_trial, cnt)
            self.stp0 = self.stp1
            self.stp1 = min(quadratic(self.f, self.df, self.stp1, self.f1), self.stp_max)
            alpha = self.stp1
            self.f0 = self.f1

    def _zoom(self, reverse, prob, m0, p, trial, max_trial, cnt):
        alo, ahi, flo, fhi = self.stp0, self.stp1, self.f0, self.f1
        if reverse:
            alo, ahi, flo, fhi = self.stp1, self.stp0, self.f1, self.f0
        while True:
            if trial >= max_trial:
                return False
            if alo == 0:
                alpha = quadratic(self.f, self.df, ahi, fhi)
            elif ahi == 0:
                alpha = quadratic(self.f, self.df, alo, flo)
            else:
                alpha = cubic(self.f, self.df, alo, flo, ahi, fhi)
            prob.m = m0 + alpha * p
            prob.res()
            cnt["feval"] += 1
            trial += 1
            f = prob.f
            self.f0, self.f1 = self.f1, f
            if f != f:
                raise RuntimeError("the objective function is NaN")
            if f <= ZERO:
                self.stp, self.f0, self.f = alpha, self.f, f
                return False
            if self.f1 > self.f + self.c1 * alpha * self.df or self.f1 >= flo:
                ahi, fhi = alpha, self.f1
            else:
                cnt["geval"] += 1
                cnt["gnorm"] = float(np.linalg.norm(prob.g))
                df1 = float(np.dot(prob.g, p))
                if df1 >= self.c2 * self.df:
                    self.stp, self.f0, self.f = alpha, self.f, f
                    return True
                if df1 * (ahi - alo) >= 0:
                    ahi, fhi = alo, flo
                alo, flo = alpha, self.f1


def compute_hg(g, s, y, rho, k, mmax, h0):
    """Two-loop recursion (nlsolver.cpp:1327-1359): returns H g with rho = s.y."""
    q = g.copy()
    m = min(k, mmax)
    a = [0.0] * m
    gamma = 1.0
    if m > 0:
        for i in range(m - 1, -1, -1):
            a[i] = float(np.dot(s[i], q)) / rho[i]
            q = q - a[i] * y[i]
        gamma = rho[m - 1] / float(np.dot(y[m - 1], y[m - 1]))
    p = gamma * h0 * q
    for i in range(m):
        b = float(np.dot(y[i], p)) / rho[i]
        p = p + (a[i] - b) * s[i]
    return p


def lbfgs(fn, m_init, niter, mmax=5, max_trial=10, threshold=0.0, h0=None, ls=None):
    """Run the solver; returns (model, objective history, counters)."""
    ls = ls or RegularWolfe()
    ls.reset()
    prob = Problem(fn, m_init)
    m0 = prob.m.copy()
    cnt = {"feval": 0, "geval": 0, "gnorm": 999.0}
    prob.res()
    cnt["feval"] += 1
    f = prob.f
    func = [f]
    if func[0] <= ZERO or func[0] != func[0]:
        return prob.m, func, cnt
    n = prob.m.size
    h0 = np.ones(n) if h0 is None else np.asarray(h0, np.float64)
    s = [np.zeros(n) for _ in range(mmax)]
    y = [np.zeros(n) for _ in range(mmax)]
    rho = [0.0] * mmax
    g0 = np.zeros(n)
    k = 0
    rate = 1.0
    success = True
    while k < niter and rate > threshold and cnt["gnorm"] > ZERO and f > ZERO and success:
        if ls.flag_g or k == 0:
            cnt["geval"] += 1
            cnt["gnorm"] = float(np.linalg.norm(prob.g))
        # updateSY (nlsolver.cpp:1290-1325)
        if 0 < k <= mmax:
            s[k - 1] = prob.m - m0
            y[k - 1] = prob.g - g0
            rho[k - 1] = float(np.dot(s[k - 1], y[k - 1]))
        elif k > mmax:
            s = s[1:] + [prob.m - m0]
            y = y[1:] + [prob.g - g0]
            rho = rho[1:] + [float(np.dot(s[-1], y[-1]))]
        p = -compute_hg(prob.g, s, y, rho, k, mmax, h0)
        if float(np.dot(p, prob.g)) >= 0:
            break
        if k > 0:
            m0 = prob.m.copy()
        g0 = prob.g.copy()
        success = ls.line_search(prob, m0, p, k, max_trial, cnt)
        if not success:
            break
        f = prob.f
        func.append(f)
        if k > 0:
            rate = (func[k] - func[k + 1]) / func[k]
        k += 1
    cnt["iterations"] = k
    return prob.m, func, cnt
