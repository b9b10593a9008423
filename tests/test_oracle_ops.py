"""Pin the oracle's SBP/SAT/taper operators against the reference's own compiled
operators (oracle/_ref, built from src/cpp/spatial_operators.cpp and
src/include/spatial_operators.hpp) in FP32 and FP64."""
import numpy as np
import pytest

from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available("f") or not O.ref_available("d"),
                               reason="reference operator library not built (no reference checkout)")

SHAPE = (14, 13, 16)  # ny, nx, nz: every axis >= 12 so all D2 closure rows are distinct
D = (0.011, 0.009, 0.01)
TOL = {"f": 2e-6, "d": 1e-13}


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def field(rng, s, n=1, lo=-1.0, hi=1.0, shape=SHAPE):
    T = np.float32 if s == "f" else np.float64
    return np.ascontiguousarray(rng.uniform(lo, hi, size=(n,) + shape).astype(T).squeeze(0) if n == 1
                                else rng.uniform(lo, hi, size=(n,) + shape).astype(T))


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("mult", [False, True])
@pytest.mark.parametrize("add", [False, True])
def test_d1_matches_reference(rng, s, axis, mult, add):
    u = field(rng, s)
    par = field(rng, s, lo=0.5, hi=2.0) if mult else None
    base = field(rng, s)
    a, b = base.copy(), base.copy()
    O.Ops("orc", s).d1(axis, u, a, D[axis], par=par, add=add)
    O.Ops("ref", s).d1(axis, u, b, D[axis], par=par, add=add)
    assert rel(a, b) < TOL[s]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("a", [1.0, 2.0])
def test_d2_matches_reference(rng, s, axis, a):
    u = field(rng, s)
    bb = field(rng, s, lo=0.5, hi=3.0)
    base = field(rng, s)
    x, y = base.copy(), base.copy()
    O.Ops("orc", s).d2(axis, u, x, bb, D[axis], a=a, add=True)
    O.Ops("ref", s).d2(axis, u, y, bb, D[axis], a=a, add=True)
    assert rel(x, y) < 4 * TOL[s]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("inv", [False, True])
def test_h_matches_reference(rng, s, axis, inv):
    u = field(rng, s)
    x, y = np.zeros_like(u), np.zeros_like(u)
    O.Ops("orc", s).apply_h(axis, u, x, D[axis], inv=inv)
    O.Ops("ref", s).apply_h(axis, u, y, D[axis], inv=inv)
    assert rel(x, y) < TOL[s]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("tp", [(3, 4, 2, 5, 4, 3), (0, 5, 5, 5, 5, 5), (6, 6, 6, 6, 7, 7)])
def test_taper_matches_reference(rng, s, tp):
    u = field(rng, s, n=3)
    x, y = u.copy(), u.copy()
    O.Ops("orc", s).taper(x, tp, 0.3)
    O.Ops("ref", s).taper(y, tp, 0.3)
    assert rel(x, y) < TOL[s]
    assert not np.allclose(x, u)


def acoustic_model(rng, s, shape=SHAPE):
    T = np.float32 if s == "f" else np.float64
    vp = rng.uniform(1.5, 3.0, size=shape)
    rho = rng.uniform(1.8, 2.6, size=shape)
    return np.ascontiguousarray(np.stack([rho * vp * vp, 1.0 / rho]).astype(T))


def elastic_model(rng, s, vti, shape=SHAPE):
    T = np.float32 if s == "f" else np.float64
    vp = rng.uniform(2.0, 3.5, size=shape)
    vs = vp / rng.uniform(1.7, 2.0, size=shape)
    rho = rng.uniform(1.8, 2.6, size=shape)
    lam = rho * (vp ** 2 - 2 * vs ** 2)
    mu = rho * vs ** 2
    if not vti:
        return np.ascontiguousarray(np.stack([lam, mu, rho]).astype(T))
    eps = rng.uniform(0.0, 0.2, size=shape)
    dlt = rng.uniform(0.0, 0.1, size=shape)
    gam = rng.uniform(0.0, 0.1, size=shape)
    c13 = np.sqrt(2 * (lam + 2 * mu) * (lam + mu) * dlt + (lam + mu) ** 2) - mu
    c66 = (1 + 2 * gam) * mu
    return np.ascontiguousarray(np.stack([lam, mu, rho, c13, eps, c66]).astype(T))


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("bc", [1, 2, 3])
def test_acoustic_sat_matches_reference(rng, s, face, bc):
    m = acoustic_model(rng, s)
    curr, prev, base = field(rng, s), field(rng, s), field(rng, s)
    x, y = base.copy(), base.copy()
    alpha = 0.2508560249 / 1.05
    O.Ops("orc", s).asat(face, bc, curr, prev, x, m, D, 1e-3, alpha)
    O.Ops("ref", s).asat(face, bc, curr, prev, y, m, D, 1e-3, alpha)
    assert rel(x - base, y - base) < 10 * TOL[s]
    assert np.abs(x - base).max() > 0


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("bc", [(2, 2, 2, 2, 2, 2), (1, 2, 2, 2, 2, 2), (2, 1, 1, 2, 1, 2), (1, 1, 1, 1, 1, 1)])
def test_acoustic_scale_matches_reference(rng, s, bc):
    m = acoustic_model(rng, s)
    u = field(rng, s)
    x, y = u.copy(), u.copy()
    O.Ops("orc", s).ascale(x, m, D, 1e-3, bc)
    O.Ops("ref", s).ascale(y, m, D, 1e-3, bc)
    assert rel(x, y) < 5 * TOL[s]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("vti", [False, True])
@pytest.mark.parametrize("face", range(6))
@pytest.mark.parametrize("bc", [1, 2])
def test_elastic_sat_matches_reference(rng, s, vti, face, bc):
    m = elastic_model(rng, s, vti)
    curr, prev, base = field(rng, s, n=3), field(rng, s, n=3), field(rng, s, n=3)
    x, y = base.copy(), base.copy()
    O.Ops("orc", s).esat(face, bc, vti, curr, prev, x, m, D, 1e-3)
    O.Ops("ref", s).esat(face, bc, vti, curr, prev, y, m, D, 1e-3)
    assert np.abs(y - base).max() > 0
    assert rel(x - base, y - base) < 10 * TOL[s]


@needs_ref
@pytest.mark.parametrize("s", ["f", "d"])
@pytest.mark.parametrize("vti", [False, True])
@pytest.mark.parametrize("bc", [(2, 2, 2, 2, 2, 2), (1, 2, 2, 2, 2, 2), (2, 1, 2, 1, 1, 2), (1, 1, 1, 1, 1, 1)])
def test_elastic_scale_matches_reference(rng, s, vti, bc):
    m = elastic_model(rng, s, vti)
    u = field(rng, s, n=3)
    x, y = u.copy(), u.copy()
    O.Ops("orc", s).escale(vti, x, m, D, 1e-3, bc)
    O.Ops("ref", s).escale(vti, y, m, D, 1e-3, bc)
    assert rel(x, y) < 5 * TOL[s]
