"""The C-ABI library loads and exports every function include/fwi3d_hip.h declares;
host-side entry points that need no GPU behave like the reference."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fwi3d_amd as F
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "fwi3d_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fwi3d_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 8
    L = F.lib()
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/fwi3d_hip.h but not exported"
    assert set(F.EXPORTS) == set(names)


def test_engine_fails_loudly_without_device():
    # No GPU in the CPU suite: creating a context must raise, never fall back.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        F.Engine(0)


@pytest.mark.parametrize("kind", [0, 1])
def test_injector_tables_match_oracle(kind):
    shape = (40, 37, 33)  # ny, nx, nz
    d = (0.01, 0.012, 0.011)
    o = (0.1, -0.05, 0.0)
    rng = np.random.default_rng(7)
    loc = np.stack([rng.uniform(o[0], o[0] + 36 * d[0], 50), rng.uniform(o[1], o[1] + 39 * d[1], 50),
                    rng.uniform(o[2], o[2] + 32 * d[2], 50)], axis=1).astype(np.float32)
    # include the exact edges and half-cell ties
    loc[0] = (o[0], o[1], o[2])
    loc[1] = (o[0] + 36 * d[0], o[1] + 39 * d[1], o[2] + 32 * d[2])
    loc[2] = (o[0] + 10.5 * d[0], o[1] + 20.5 * d[1], o[2] + 5.5 * d[2])
    a = F.Injector(kind, loc, shape, d, o)
    b = O.Injector(kind, loc, shape, d, o, s="f")
    assert np.array_equal(a.xind, b.xi) and np.array_equal(a.yind, b.yi) and np.array_equal(a.zind, b.zi)
    for wa, wb in ((a.xw, b.xw), (a.yw, b.yw), (a.zw, b.zw)):
        np.testing.assert_allclose(wa, wb, rtol=2e-6, atol=1e-6 * np.abs(wb).max())


def test_injector_rejects_points_outside():
    with pytest.raises(ValueError):
        F.Injector(0, [[5.0, 0.1, 0.1]], (20, 20, 20), (0.01, 0.01, 0.01))


def test_injector_moment_conditions():
    # Injection weights reproduce the H-weighted moments of a delta: sum h H_i w_i (x_i-x)^k = delta_k0
    n, h = 30, 0.01
    for x in (0.0, 0.004, 0.013, 0.1234, 0.29):
        inj = F.Injector(0, [[x, 0.1, 0.1]], (n, n, n), (h, h, h))
        idx, w = inj.xind[0], inj.xw[0, 0].astype(np.float64)
        H = np.array([O.hweight(i, n) for i in idx]) if hasattr(O, "hweight") else None
        Hw = np.array([[17 / 48, 59 / 48, 43 / 48, 49 / 48][i] if i < 4 else
                       ([17 / 48, 59 / 48, 43 / 48, 49 / 48][n - 1 - i] if i >= n - 4 else 1.0) for i in idx])
        xi = idx * h
        for k in range(3):
            m = np.sum(h * Hw * w * (xi - x) ** k)
            assert abs(m - (1.0 if k == 0 else 0.0)) < 1e-5 * (h ** k if k else 1) + 1e-9, (x, k, m)
        # extraction weights reproduce quadratics exactly
        we = inj.xw[0, 1].astype(np.float64)
        for k in range(3):
            assert abs(np.sum(we * xi ** k) - x ** k) < 1e-6


@pytest.mark.parametrize("nmodels", [2, 3, 6])
@pytest.mark.parametrize("dt,sub", [(-1.0, -1), (0.0, 0), (0.0015, -1), (0.01, 3)])
def test_host_analyze_model_matches_oracle(nmodels, dt, sub):
    """fwi3d_analyze_model (host code of the library) against the oracle's analyzeModel:
    the same clipped model bit for bit and the same time axis."""
    from oracle import oracle as O
    rng = np.random.default_rng(nmodels * 10 + sub)
    shape = (13, 14, 15)
    vp = rng.uniform(0.1, 9.0, shape).astype(np.float32)      # some outside [0.2, 8]
    vs = (vp / rng.uniform(1.2, 2.2, shape)).astype(np.float32)  # some above vp/sqrt(2)
    rho = rng.uniform(0.05, 3.0, shape).astype(np.float32)
    if nmodels == 2:
        m = np.stack([vp, rho])
    elif nmodels == 3:
        m = np.stack([vp, vs, rho])
    else:
        m = np.stack([vp, vs, rho, rng.uniform(-0.6, 1.1, shape), rng.uniform(-0.1, 1.1, shape),
                      rng.uniform(-0.1, 1.1, shape)]).astype(np.float32)
    kw = dict(nmodels=nmodels, dt=dt, sub=sub, courant=0.45)
    p = F.op_param(**kw)
    shot = F.Shot(shape, (0.01, 0.012, 0.011), (0, 0, 0), 301, 0.002, (0.05, 0.05, 0.05), [(0.06, 0.06, 0.02)])
    got = F.analyze_model(p, m, shot)
    q = O.op_par(**kw)
    g = O.ShotGeom(shape, (0.01, 0.012, 0.011), (0, 0, 0), 301, 0.002, (0.05, 0.05, 0.05), [(0.06, 0.06, 0.02)], s="f")
    want = O.analyze(q, m, g)
    assert np.array_equal(got, want)
    assert (p.nt, p.sub) == (q.nt, q.sub)
    assert np.float32(p.dt) == np.float32(q.dt) and np.float32(p.courant) == np.float32(q.courant)
    assert np.float32(p.vmax) == np.float32(q.vmax) and np.float32(p.vmin) == np.float32(q.vmin)


def test_dipole_tables_match_oracle():
    """fwi3d_dipole_tables (host code) against the oracle's dipole_m3 restatement: same
    indices and weights (pinned themselves in test_oracle_das.py)."""
    from oracle import oracle as O
    rng = np.random.default_rng(8)
    shape = (30, 34, 38)
    ny, nx, nz = shape
    d = (0.01, 0.012, 0.009)
    L = np.array([(nx - 1) * d[0], (ny - 1) * d[1], (nz - 1) * d[2]])
    loc5 = np.array([list(rng.uniform(0.2, 0.8, 3) * L) + [rng.uniform(-1.5, 1.5), rng.uniform(-3, 3)]
                     for _ in range(40)], np.float32)
    gl = 0.03
    a = F.Injector.dipole(loc5, gl, shape, d, (0.0, 0.0, 0.0))
    b = O.Injector.dipole(loc5, gl, shape, d, (0, 0, 0), s="f")
    for k in ("xind", "yind", "zind"):
        assert np.array_equal(getattr(a, k), getattr(b, k[0] + "i")), k
    for k in ("xw", "yw", "zw"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    with pytest.raises(ValueError):
        F.Injector.dipole([[0.0, 0.0, 0.0, 0.0, 0.0]], gl, shape, d)  # an end falls outside the grid
