"""Pin the oracle's l-BFGS / regular-Wolfe restatement (oracle/lbfgs.py; reference
src/cpp/nlsolver.cpp:10-36, 185-423, 1158-1359) by what the mathematics fixes:
the step interpolants are exact on parabolas and cubics, the two-loop recursion equals
the explicit L-BFGS inverse-Hessian matrix, accepted steps satisfy the Wolfe
conditions, and the solver converges on convex quadratics and on Rosenbrock's
function."""
import numpy as np
import pytest

from oracle import lbfgs as L


def test_quadratic_step_is_parabola_minimiser():
    f, df, c = 2.0, -3.0, 4.0
    phi = lambda a: f + df * a + c * a * a
    s = 1.0
    assert L.quadratic(f, df, s, phi(s)) == pytest.approx(-df / (2 * c), rel=1e-14)
    # f0 below the tangent line: extrapolate by 10 %
    assert L.quadratic(f, df, s, f + df * s - 1.0) == pytest.approx(1.1)


def test_cubic_step_is_cubic_minimiser():
    f, df, b, a = 1.0, -2.0, 0.5, 0.3
    phi = lambda x: f + df * x + b * x * x + a * x ** 3
    s0, s1 = 0.7, 2.5
    want = (-b + np.sqrt(b * b - 3 * a * df)) / (3 * a)
    assert L.cubic(f, df, s0, phi(s0), s1, phi(s1)) == pytest.approx(want, rel=1e-12)


def _explicit_lbfgs_matrix(s, y, h0):
    """H from gamma*H0 through the BFGS inverse update H <- (I - r s y^T) H (I - r y s^T) + r s s^T."""
    n = len(h0)
    m = len(s)
    gamma = float(s[-1] @ y[-1]) / float(y[-1] @ y[-1])
    H = gamma * np.diag(h0)
    I = np.eye(n)
    for i in range(m):
        r = 1.0 / float(s[i] @ y[i])
        H = (I - r * np.outer(s[i], y[i])) @ H @ (I - r * np.outer(y[i], s[i])) + r * np.outer(s[i], s[i])
    return H


@pytest.mark.parametrize("k", [1, 3, 5, 8])
def test_two_loop_recursion_is_the_lbfgs_matrix(k, rng):
    n, mmax = 12, 5
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    s = [rng.standard_normal(n) for _ in range(mmax)]
    y = [A @ v for v in s]  # s.y > 0
    rho = [float(a @ b) for a, b in zip(s, y)]
    h0 = rng.uniform(0.5, 2.0, n)
    g = rng.standard_normal(n)
    m = min(k, mmax)
    got = L.compute_hg(g, s, y, rho, k, mmax, h0)
    want = _explicit_lbfgs_matrix(s[:m], y[:m], h0) @ g
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12)


def test_converges_on_spd_quadratic(rng):
    n = 30
    A = rng.standard_normal((n, n))
    A = A @ A.T / n + np.diag(np.linspace(0.5, 5, n))
    x_star = rng.standard_normal(n)
    fn = lambda x: (0.5 * float((x - x_star) @ A @ (x - x_star)) + 1.0, A @ (x - x_star))
    x, hist, cnt = L.lbfgs(fn, np.zeros(n), niter=80)
    assert all(b <= a for a, b in zip(hist, hist[1:]))  # non-increasing (sufficient decrease)
    assert np.linalg.norm(x - x_star) < 1e-5 * np.linalg.norm(x_star)
    assert cnt["iterations"] <= 80 and cnt["feval"] >= cnt["iterations"]


def test_rosenbrock():
    def fn(x):
        a, b = x
        return (1 - a) ** 2 + 100 * (b - a * a) ** 2 + 1e-3, np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])
    x, hist, cnt = L.lbfgs(fn, np.array([-1.2, 1.0]), niter=200)
    assert np.allclose(x, [1.0, 1.0], atol=1e-4)


def test_accepted_steps_satisfy_wolfe(rng):
    """Every accepted step of the search satisfies sufficient decrease and the (regular)
    curvature condition, f(a) <= f + c1 a f' and f'(a) >= c2 f'."""
    n = 20
    A = rng.standard_normal((n, n))
    A = A @ A.T / n + np.eye(n)
    fn = lambda x: (0.5 * float(x @ A @ x) + float(np.sum(np.cos(x))) + 30.0, A @ x - np.sin(x))
    ls = L.RegularWolfe()
    prob = L.Problem(fn, rng.standard_normal(n) * 2)
    prob.res()
    cnt = {"feval": 1, "geval": 0, "gnorm": float(np.linalg.norm(prob.g))}
    for it in range(5):
        m0 = prob.m.copy()
        f0, g0 = prob.f, prob.g.copy()
        p = -g0
        assert ls.line_search(prob, m0, p, it, 20, cnt)
        a = ls.stp
        assert np.allclose(prob.m, m0 + a * p)
        assert prob.f <= f0 + ls.c1 * a * float(g0 @ p) + 1e-12
        assert float(prob.g @ p) >= ls.c2 * float(g0 @ p) - 1e-12
