"""DAS channels in the oracle: dipole_m3 tables (src/cpp/injector.cpp:313-492),
dipole_to_strain (src/cpp/misc.cpp:11-42) and the DAS branches of apply_forward /
apply_jacobianT (src/cpp/we_op.cpp:918-930, 1013-1023), pinned by what the
mathematics fixes: the end positions along the fiber, exactness of the extraction on
quadratics, the H-weighted moment conditions of the injection, the strain of a
constant-strain field, adjointness of injection / extraction and of the strain
conversion, and a finite-difference check of the gradient with DAS data."""
import numpy as np
import pytest

from oracle import oracle as O

H = 0.01
D = (H, H, H)
SHAPE = (16, 18, 20)  # ny, nx, nz


def hw(i, n):
    w = [17 / 48, 59 / 48, 43 / 48, 49 / 48]
    return w[i] if i < 4 else (w[n - 1 - i] if i >= n - 4 else 1.0)


def channels(rng, n=7, gl=0.025):
    ny, nx, nz = SHAPE
    L = np.array([(nx - 1) * H, (ny - 1) * H, (nz - 1) * H])
    loc = []
    for _ in range(n):
        p = rng.uniform(0.3, 0.7, 3) * L
        loc.append([p[0], p[1], p[2], rng.uniform(-1.2, 1.2), rng.uniform(-3, 3)])
    return np.array(loc), gl


def direction(dip, az):
    return np.array([np.cos(dip) * np.cos(az), np.cos(dip) * np.sin(az), np.sin(dip)])


def test_dipole_tables(rng):
    loc, gl = channels(rng)
    I = O.Injector.dipole(loc, gl, SHAPE, D, (0, 0, 0), s="d")
    ny, nx, nz = SHAPE
    lens = {0: nx, 1: ny, 2: nz}
    for r in range(len(loc)):
        n = direction(loc[r, 3], loc[r, 4])
        for e, sg in ((0, -1.0), (1, 1.0)):
            end = loc[r, :3] + sg * gl / 2 * n
            for ax, (ind, w) in enumerate(((I.xi, I.xw), (I.yi, I.yw), (I.zi, I.zw))):
                idx = ind[r, 3 * e:3 * e + 3]
                x = idx * H
                wx = w[r, 1, 3 * e:3 * e + 3]   # extraction: Lagrange weights at the end
                wi = w[r, 0, 3 * e:3 * e + 3]   # injection: H-weighted discrete delta
                assert list(idx) == [idx[1] - 1, idx[1], idx[1] + 1]
                assert abs(idx[1] * H - end[ax]) <= H / 2 + 1e-12 or idx[1] in (1, lens[ax] - 2)
                for k in range(3):
                    assert np.dot(wx, (x - end[ax]) ** k) == pytest.approx(1.0 if k == 0 else 0.0, abs=1e-12)
                    hh = np.array([H * hw(i, lens[ax]) for i in idx])
                    assert np.dot(hh * wi, (x - end[ax]) ** k) == pytest.approx(1.0 if k == 0 else 0.0, abs=1e-12)


def test_constant_strain_is_recovered(rng):
    """u = G x: each component's dipole difference is gl (G n)_c, and dipole_to_strain
    gives n . G n (the axial strain along the fiber)."""
    loc, gl = channels(rng)
    I = O.Injector.dipole(loc, gl, SHAPE, D, (0, 0, 0), s="d")
    ny, nx, nz = SHAPE
    G = rng.standard_normal((3, 3))
    yy, xx, zz = np.meshgrid(np.arange(ny) * H, np.arange(nx) * H, np.arange(nz) * H, indexing="ij")
    X = np.stack([xx, yy, zz])
    nt = 1
    rcv = np.zeros((3, len(loc), nt))
    for c in range(3):
        u = np.ascontiguousarray(np.tensordot(G[c], X, axes=1))
        O.extract(I, u, nt, c, 0, rcv, s="d")
    for r in range(len(loc)):
        n = direction(loc[r, 3], loc[r, 4])
        assert np.allclose(rcv[:, r, 0], gl * (G @ n), atol=1e-12)
    O.dipole_to_strain(False, rcv, loc[:, 3:], gl, s="d")
    for r in range(len(loc)):
        n = direction(loc[r, 3], loc[r, 4])
        assert rcv[0, r, 0] == pytest.approx(n @ G @ n, abs=1e-10)
        assert rcv[1, r, 0] == 0 and rcv[2, r, 0] == 0


def test_injection_is_h_adjoint_of_extraction(rng):
    loc, gl = channels(rng)
    I = O.Injector.dipole(loc, gl, SHAPE, D, (0, 0, 0), s="d")
    ny, nx, nz = SHAPE
    Hq = (np.array([hw(i, ny) for i in range(ny)])[:, None, None] * np.array([hw(i, nx) for i in range(nx)])[None, :, None]
          * np.array([hw(i, nz) for i in range(nz)])[None, None, :]) * H ** 3
    u = rng.standard_normal(SHAPE)
    s = rng.standard_normal((1, len(loc), 1))
    out = np.zeros(SHAPE)
    O.inject(I, s, 1, 0, 0, out, s="d")
    rcv = np.zeros((1, len(loc), 1))
    O.extract(I, u, 1, 0, 0, rcv, s="d")
    a, b = float(np.sum(Hq * out * u)), float(np.sum(s * rcv))
    assert a == pytest.approx(b, rel=1e-12)


def test_strain_conversion_adjoint(rng):
    loc, gl = channels(rng)
    p = rng.standard_normal((3, len(loc), 5))
    q = np.zeros_like(p)
    q[0] = rng.standard_normal((len(loc), 5))
    Dp = O.dipole_to_strain(False, p.copy(), loc[:, 3:], gl, s="d")
    Dtq = O.dipole_to_strain(True, q.copy(), loc[:, 3:], gl, s="d")
    assert float(np.sum(Dp * q)) == pytest.approx(float(np.sum(p * Dtq)), rel=1e-12)


def das_case(nt=300, dt=0.001, f=10.0):
    ny, nx, nz = 24, 26, 28
    shape = (ny, nx, nz)
    z = np.arange(nz) * H
    yy, xx, zz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    vp = 2.0 + 0.6 * z[None, None, :] + 0.1 * np.sin(xx / 5.0) * np.cos(yy / 6.0)
    model = np.stack([vp, vp / 1.8, 0.31 * (1000 * vp) ** 0.25])
    t = np.arange(nt) * dt
    tau = t - 1.2 / f
    w = (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)
    src = np.stack([w, 0.5 * w, 0.8 * w])
    rx = [(0.05 + 0.03 * i, 0.06 + 0.02 * i, 0.04) for i in range(6)]
    da = [(0.3 * i - 0.6, 0.5 * i) for i in range(6)]
    g = O.ShotGeom(shape, D, (0, 0, 0), nt, dt, (0.12, 0.12, 0.12), rx, s="d", rdipaz=da)
    kw = dict(nmodels=3, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4), sub=1, dt=dt, gl=0.02, lag=0)
    return model, src, g, kw


def test_das_forward_is_strain_of_displacement_traces():
    """apply_forward with gl > 0 equals the strain conversion of the dipole traces that
    a plain propagate extracts with the same tables."""
    model, src, g, kw = das_case(nt=120)
    das, _ = O.apply_forward(O.op_par(**kw), g, model, src)
    assert das.shape == (1, 6, 120)
    conv = O.convert(model, 3, s="d").reshape(model.shape)
    loc5 = np.concatenate([g.rxyz, g.rdipaz], axis=1)
    ext = O.Injector.dipole(loc5, kw["gl"], g.shape, D, (0, 0, 0), s="d")
    inj = O.Injector(0, g.sxyz[None], g.shape, D, (0, 0, 0), s="d")
    par = O.make_par(120, 0, kw["dt"], bc=kw["bc"], taper=kw["taper"])
    rcv, _ = O.propagate("elastic", conv, src[:, None, :], inj, ext, par, D, s="d")
    O.dipole_to_strain(False, rcv, g.rdipaz, kw["gl"], s="d")
    assert np.abs(rcv[0]).max() > 0
    assert np.allclose(das[0], rcv[0], rtol=1e-9, atol=1e-12 * np.abs(rcv[0]).max())


@pytest.mark.slow
def test_das_gradient_finite_difference():
    """FWI objective on DAS data: <gradient, dm> against the central difference in the vp
    direction (lag-free imaging, 10 Hz; measured 8e-5, tolerance 1e-3). The data path of
    the strain channels (forward conversion, adjoint injection through the dipoles) is
    what this pins; the other directions' imaging terms are pinned with point receivers."""
    model, src, g, kw = das_case()
    obsm = model.copy()
    obsm[0] *= 1.03
    obs, _ = O.apply_forward(O.op_par(**kw), g, obsm, src)
    _, _, grad = O.fwi_shot(O.op_par(**kw), g, model, src, obs)
    ny, nx, nz = g.shape
    yy, xx, zz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    bump = np.exp(-((yy - ny / 2) ** 2 + (xx - nx / 2) ** 2 + (zz - nz / 2) ** 2) / 20.0)
    for k in (0,):
        dm = np.zeros_like(model)
        dm[k] = 0.05 * bump * model[k]
        eps = 0.01
        mp, _, _ = O.fwi_shot(O.op_par(**kw), g, model + eps * dm, src, obs)
        mm, _, _ = O.fwi_shot(O.op_par(**kw), g, model - eps * dm, src, obs)
        fd = (mp - mm) / (2 * eps)
        an = float(np.sum(grad[k] * dm[k]))
        print(f"DAS parameter {k}: <g, dm> {an:.6e}, central difference {fd:.6e}")
        assert abs(an - fd) / abs(fd) < 1e-3
