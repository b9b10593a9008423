"""Pin the oracle's shot-operator restatement (oracle/pipeline.hpp) — the chain around
propagate in apply_forward / apply_jacobianT (src/cpp/we_op.cpp:872-1078), the time
resamplers (src/cpp/operator.cpp:419-520), Ht and Dt (src/cpp/misc.cpp:44-94), the
model analysis (we_op.cpp:234-446) and the FWI shot (src/cpp/optimization.cpp:35-188).

The reference's operator.cpp and misc.cpp need fftw3.h, absent here, so these parts
cannot be compiled from the reference; they are pinned by what the mathematics fixes:
adjointness, exactness on polynomials, band-limited interpolation, closed-form time
axes, conversion identities, and a finite-difference check of the whole gradient
chain (which fails on a wrong sign, a missing time reversal, a wrong alpha or a wrong
chain-rule factor)."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# ------------------------------------------------------------------ resamplers
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("ratio", [1, 2, 3, 0.5])
def test_resampler_adjoint(method, ratio, rng):
    """<R m, d> = <m, R^T d> in FP64 (the adjoint branch is the transpose)."""
    nm, dm = 40, 0.004
    dd = dm / ratio
    nd = int(round((nm - 1) * dm / dd)) + 1
    m = rng.standard_normal((3, nm))
    d = rng.standard_normal((3, nd))
    Rm = np.zeros((3, nd))
    O.resample(method, False, False, m.copy(), Rm, dm, dd, s="d")
    Rtd = np.zeros((3, nm))
    O.resample(method, True, False, Rtd, d.copy(), dm, dd, s="d")
    assert abs(np.vdot(Rm, d) - np.vdot(m, Rtd)) < 1e-12 * max(1.0, abs(np.vdot(Rm, d)))


def test_resampler_add_flag(rng):
    nm, dm, dd = 30, 0.004, 0.002
    nd = 59
    m = rng.standard_normal((2, nm))
    base = rng.standard_normal((2, nd))
    for method in (0, 1):
        out0 = np.zeros((2, nd))
        O.resample(method, False, False, m.copy(), out0, dm, dd, s="d")
        out = base.copy()
        O.resample(method, False, True, m.copy(), out, dm, dd, s="d")
        assert np.allclose(out, base + out0, atol=1e-13)
        adj0 = np.zeros((2, nm))
        O.resample(method, True, False, adj0, base.copy(), dm, dd, s="d")
        adj = m.copy()
        O.resample(method, True, True, adj, base.copy(), dm, dd, s="d")
        assert np.allclose(adj, m + adj0, atol=1e-13)


def test_sinc_same_rate_is_identity():
    """With equal steps the taps sit on sinc zeros: identity up to the FP32 rounding of
    pi*k (the reference runs the resampler even at alpha=1)."""
    x = np.random.default_rng(3).standard_normal((2, 50)).astype(np.float32)
    y = np.zeros_like(x)
    O.resample(0, False, False, x.copy(), y, 0.002, 0.002, s="f")
    assert rel(y, x) < 5e-6


def test_sinc_interpolates_bandlimited():
    """Windowed sinc (hl=11) reproduces a band-limited signal on a 3x finer axis away
    from the ends; a wrong argument, sign or window breaks this."""
    dm, nm = 0.004, 200
    dd = dm / 3
    nd = (nm - 1) * 3 + 1
    tm = np.arange(nm) * dm
    td = np.arange(nd) * dd
    f = lambda t: np.sin(2 * np.pi * 15 * t) + 0.5 * np.cos(2 * np.pi * 31 * t)  # <= 0.25 of Nyquist
    out = np.zeros((1, nd))
    O.resample(0, False, False, f(tm)[None].copy(), out, dm, dd, s="d")
    core = slice(3 * 15, nd - 3 * 15)
    assert np.max(np.abs(out[0, core] - f(td)[core])) < 5e-3


def test_linear_exact_on_lines():
    dm, nm, dd = 0.004, 30, 0.0013
    nd = int((nm - 1) * dm / dd) + 1
    tm = np.arange(nm) * dm
    td = np.arange(nd) * dd
    out = np.zeros((1, nd))
    O.resample(1, False, False, (2.0 * tm + 0.3)[None].copy(), out, dm, dd, s="d")
    assert np.allclose(out[0], 2.0 * td + 0.3, atol=1e-12)


# ------------------------------------------------------------------ Ht, Dt
def test_ht_trapezoid_and_inverse(rng):
    nt, dt = 17, 0.003
    x = np.ones((2, nt))
    O.ht(False, x, dt, s="d")
    assert np.isclose(x[0].sum(), (nt - 1) * dt)  # trapezoid integral of 1 over [0, (nt-1)dt]
    assert np.isclose(x[0, 0], 0.5 * dt) and np.isclose(x[0, 1], dt)
    y = rng.standard_normal((3, nt))
    z = O.ht(True, O.ht(False, y.copy(), dt, s="d"), dt, s="d")
    assert np.allclose(z, y, rtol=1e-14)


def test_dt_forward_exact_on_quadratics():
    nt, dt = 21, 0.01
    t = np.arange(nt) * dt
    q = (3 * t * t - t + 1)[None]
    d = O.time_derivative(False, q.copy(), dt, s="d")
    assert np.allclose(d[0, 1:-1], 6 * t[1:-1] - 1, atol=1e-10)
    lin = (2 * t + 1)[None]
    d2 = O.time_derivative(False, lin.copy(), dt, s="d")
    assert np.allclose(d2[0], 2.0, atol=1e-10)  # one-sided ends are exact on lines


def test_dt_adjoint_acts_on_time_reversed_data(rng):
    """The reference's adjoint Dt is applied after the time reversal (we_op.cpp:1002-1013):
    reverse(Dt_adj(reverse(y))) equals the transpose of Dt on interior samples."""
    nt, dt = 25, 0.01
    D = np.zeros((nt, nt))
    for j in range(nt):
        e = np.zeros((1, nt))
        e[0, j] = 1
        D[:, j] = O.time_derivative(False, e, dt, s="d")[0]
    y = rng.standard_normal((1, nt))
    a = O.time_derivative(True, y[:, ::-1].copy(), dt, s="d")[:, ::-1]
    assert np.allclose(a[0, 2:-2], (D.T @ y[0])[2:-2], atol=1e-10)


# ------------------------------------------------------------------ analysis, conversion
def _geom(shape, nt_data, dt_data, s="d"):
    ny, nx, nz = shape
    return O.ShotGeom(shape, (0.01, 0.01, 0.01), (0, 0, 0), nt_data, dt_data,
                      (nx * 0.005, ny * 0.005, 0.03), [(nx * 0.004, ny * 0.004, 0.02)], s=s)


def test_analyze_time_axis_rules():
    shape = (12, 12, 12)
    g = _geom(shape, 101, 0.004)
    m = np.stack([np.full(shape, 3.0), np.full(shape, 1.5), np.full(shape, 2.0)])
    # dt = -1: CFL step 0.5*0.01/3 rounded to divide the data step: 0.004/3
    p = O.op_par(nmodels=3, dt=-1, sub=-1)
    O.analyze(p, m, g)
    assert np.isclose(p.dt, 0.004 / 3) and p.nt == 301 and p.sub == 3
    assert np.isclose(p.courant, p.dt * 3.0 / 0.01)
    # dt = 0: the CFL step itself; nt = round(tmax/dt) + 1
    p = O.op_par(nmodels=3, dt=0)
    O.analyze(p, m, g)
    assert np.isclose(p.dt, 0.5 * 0.01 / 3.0) and p.nt == int(round(0.4 / (0.5 * 0.01 / 3.0))) + 1
    # dt > 0: capped by the data step
    p = O.op_par(nmodels=3, dt=0.01, sub=-1)
    O.analyze(p, m, g)
    assert np.isclose(p.dt, 0.004) and p.nt == 101 and p.sub == 1


def test_analyze_clips_and_fixes():
    shape = (12, 12, 12)
    g = _geom(shape, 11, 0.004)
    vp = np.full(shape, 2.0)
    vp[0, 0, 0] = 9.0      # above vpmax=8
    vs = np.full(shape, 1.0)
    vs[1, 1, 1] = 1.9      # vp < sqrt(2) vs -> vs = vp/sqrt(2.0001)
    rho = np.full(shape, 2.0)
    rho[2, 2, 2] = 0.01    # below rhomin=0.1
    p = O.op_par(nmodels=3)
    c = O.analyze(p, np.stack([vp, vs, rho]), g)
    assert c[0, 0, 0, 0] == 8.0 and c[2, 2, 2, 2] == 0.1
    assert np.isclose(c[1, 1, 1, 1], 2.0 / np.sqrt(2.0001))
    assert p.vmax == 8.0 and np.isclose(p.vmin, 1.0)
    p = O.op_par(nmodels=3, soft_clip=1)
    c = O.analyze(p, np.stack([vp, vs, rho]), g)
    assert c[0, 0, 0, 0] == 9.0 and c[1, 1, 1, 1] == 1.9


def test_analyze_vti_delta_scan_covers_first_plane_only():
    """Thomsen-delta admissibility -0.5(1-vs^2/vp^2) < delta < 2/(vp^2/vs^2-1) is
    enforced on the first nx*nz points only (we_op.cpp:273)."""
    shape = (12, 12, 12)
    g = _geom(shape, 11, 0.004)
    vp, vs = np.full(shape, 2.0), np.full(shape, 1.0)
    upper = 2.0 / (4.0 - 1.0)
    delta = np.full(shape, 0.9)  # above the bound 2/3 everywhere
    m = np.stack([vp, vs, np.full(shape, 2.0), delta, np.full(shape, 0.1), np.zeros(shape)])
    p = O.op_par(nmodels=6)
    c = O.analyze(p, m, g)
    assert np.allclose(c[3, 0], 0.9999 * upper)   # plane iy=0 (nx*nz points)
    assert np.allclose(c[3, 1:], 0.9)             # untouched elsewhere
    assert np.isclose(p.vmax, 2.0 * np.sqrt(1.2))  # horizontal P speed vp*sqrt(1+2 eps)


def test_convert_identities(rng):
    n = 50
    vp = rng.uniform(2, 4, n)
    vs = vp / rng.uniform(1.6, 2.2, n)
    rho = rng.uniform(1.8, 2.6, n)
    lam_mu_rho = O.convert(np.concatenate([vp, vs, rho]), 3, s="d").reshape(3, n)
    assert np.allclose(lam_mu_rho[0] + 2 * lam_mu_rho[1], rho * vp ** 2)  # lambda + 2 mu = rho vp^2
    assert np.allclose(lam_mu_rho[1], rho * vs ** 2)
    # VTI: delta = 0 gives c13 = lambda; gamma gives c66 = (1 + 2 gamma) mu
    eps, gam = rng.uniform(0, 0.2, n), rng.uniform(0, 0.1, n)
    v = O.convert(np.concatenate([vp, vs, rho, np.zeros(n), eps, gam]), 6, s="d").reshape(6, n)
    assert np.allclose(v[3], v[0]) and np.allclose(v[5], (1 + 2 * gam) * v[1]) and np.allclose(v[4], eps)
    # Thomsen: delta = ((c13+c44)^2 - (c33-c44)^2) / (2 c33 (c33-c44)) with c33 = lambda+2mu, c44 = mu
    dl = rng.uniform(-0.05, 0.1, n)
    v = O.convert(np.concatenate([vp, vs, rho, dl, eps, gam]), 6, s="d").reshape(6, n)
    c33, c44 = v[0] + 2 * v[1], v[1]
    assert np.allclose(((v[3] + c44) ** 2 - (c33 - c44) ** 2) / (2 * c33 * (c33 - c44)), dl)
    a = O.convert(np.concatenate([vp, rho]), 2, s="d").reshape(2, n)
    assert np.allclose(a[0], rho * vp ** 2) and np.allclose(a[1], 1 / rho)


# ------------------------------------------------------------------ full chain
def small_case(physics, n=20, nt_data=90, dt_data=0.002, s="d"):
    shape = (n, n + 2, n + 4)
    ny, nx, nz = shape
    z = np.arange(nz) * 0.01
    vp = np.broadcast_to(1.8 + 6.0 * z, shape).copy()
    r = np.random.default_rng(7)
    vp *= 1 + 0.02 * r.standard_normal(shape).clip(-2, 2)
    rho = 0.31 * (1000 * vp) ** 0.25
    if physics == "acoustic":
        model = np.stack([vp, rho])
        nmodels, nsc = 2, 1
    else:
        vs = vp / 1.8
        model = [vp, vs, rho]
        if physics == "vti":
            model += [np.full(shape, 0.05), np.full(shape, 0.1), np.full(shape, 0.02)]
        model = np.stack(model)
        nmodels, nsc = (3 if physics == "elastic" else 6), 3
    t = np.arange(nt_data) * dt_data
    f = 25.0
    tau = t - 1.2 / f
    w = (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)
    src = np.stack([w * (k + 1) for k in range(nsc)])
    rx = [(0.05 + 0.04 * i, 0.06 + 0.03 * j, 0.03) for i in range(3) for j in range(3)]
    g = O.ShotGeom(shape, (0.01, 0.01, 0.01), (0, 0, 0), nt_data, dt_data, (0.1, 0.1, 0.08), rx, s=s)
    bc = (1, 2, 2, 2, 2, 2)
    taper = (0, 4, 4, 4, 4, 4)
    return model, src, g, dict(nmodels=nmodels, bc=bc, taper=taper, sub=-1, dt=-1, courant=0.4)


@pytest.mark.parametrize("physics", ["acoustic", "elastic"])
def test_forward_linear_in_source(physics):
    model, src, g, kw = small_case(physics)
    d1, _ = O.apply_forward(O.op_par(**kw), g, model, src)
    d2, _ = O.apply_forward(O.op_par(**kw), g, model, 2.0 * src)
    assert rel(d2, 2 * d1) < 1e-12


def test_forward_decimation_matches_point_sampling():
    """Data at dt_data = 2 dt (sinc decimation alpha*R^T) agrees with the data computed
    at dt_data = dt, sampled every second step (pins alpha and the decimation)."""
    model, src, g, kw = small_case("acoustic", nt_data=160, dt_data=0.001)
    kw.update(courant=0.4, dt=0.001)
    fine, _ = O.apply_forward(O.op_par(**kw), g, model, src)
    t = np.arange(80) * 0.002
    f = 25.0
    tau = t - 1.2 / f
    w = (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)
    g2 = O.ShotGeom(g.shape, (0.01,) * 3, (0, 0, 0), 80, 0.002, g.sxyz, g.rxyz, s="d")
    p2 = O.op_par(**dict(kw, dt=0.001))
    coarse, _ = O.apply_forward(p2, g2, model, w[None])
    assert p2.dt == pytest.approx(0.001) and p2.nt == 159
    assert rel(coarse[..., 5:-5], fine[..., ::2][..., 5:80 - 5]) < 2e-2


def _fd_check(physics, k, obs_factors, eps=0.02):
    model, src, g, kw = small_case(physics)
    fac = np.ones(model.shape[0])
    fac[:len(obs_factors)] = obs_factors
    obs, _ = O.apply_forward(O.op_par(**kw), g, model * fac[:, None, None, None], src)
    _, _, grad = O.fwi_shot(O.op_par(**kw), g, model, src, obs)
    ny, nx, nz = g.shape
    yy, xx, zz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    bump = np.exp(-((yy - ny / 2) ** 2 + (xx - nx / 2) ** 2 + (zz - nz / 2) ** 2) / 20.0)
    dm = np.zeros_like(model)
    dm[k] = 0.05 * bump * model[k]
    mp, _, _ = O.fwi_shot(O.op_par(**kw), g, model + eps * dm, src, obs)
    mm, _, _ = O.fwi_shot(O.op_par(**kw), g, model - eps * dm, src, obs)
    fd = (mp - mm) / (2 * eps)
    an = float(np.sum(grad[k] * dm[k]))
    assert abs(fd) > 0
    return abs(an - fd) / abs(fd)


@pytest.mark.parametrize("physics,k,obs_factors,tol", [
    ("acoustic", 0, (1.03, 1.05), 2e-2),   # vp
    ("acoustic", 1, (1.03, 1.05), 1e-1),   # rho (the 1/rho imaging condition uses D1*D1 against the narrow D2(b))
    ("elastic", 0, (1.03,), 1e-2),         # vp (lambda term only)
    ("vti", 0, (1.03,), 2e-2),
])
def test_gradient_finite_difference(physics, k, obs_factors, tol):
    """Directional derivative of the FWI objective (central difference, FP64 oracle)
    against <grad, dm> for a smooth bump in one raw parameter. Fails on a wrong sign,
    a missing time reversal, a wrong alpha or resampling, or a wrong vp chain factor.
    Elastic/VTI use a vp-only residual: the reference's lagged adjoint derivatives
    (we_op.cpp:637-688) bias the lambda term by ~17% once converted waves dominate."""
    assert _fd_check(physics, k, obs_factors) < tol


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
def test_chain_rule_is_conversion_jacobian_transpose(physics):
    """apply_jacobianT = J^T (H G): G = gradient w.r.t. the propagation parameters from
    the oracle's adjoint propagate, H = SBP quadrature, J = d(convert)/d(raw) taken
    numerically from convert_model. With dt_data = dt the adjoint source is
    -reverse(Ht^-1 data) up to the sinc rounding (~1e-15 in FP64). Pins the chain-rule
    lines of we_op.cpp:1054-1065 (a transposed operand or wrong factor fails)."""
    model, src, g, kw = small_case(physics)
    kw = dict(kw, dt=0.001, sub=2)
    g = O.ShotGeom(g.shape, (0.01,) * 3, (0, 0, 0), 150, 0.001, g.sxyz, g.rxyz, s="d")
    t = np.arange(150) * 0.001
    src = np.stack([np.exp(-((t - 0.04) / 0.01) ** 2) * (i + 1) for i in range(src.shape[0])])
    p = O.op_par(**kw)
    data, snaps = O.apply_forward(p, g, model, src)
    assert p.nt == 150 and p.sub == 2 and p.dt == pytest.approx(0.001)
    resid = np.random.default_rng(5).standard_normal(data.shape) * np.abs(data).max()
    graw = O.apply_jacobianT(p, g, model, resid, snaps)
    # independent composition
    clipped = O.analyze(O.op_par(**kw), model, g)
    nm = model.shape[0]
    conv = O.convert(clipped, nm, s="d").reshape(model.shape)
    ny, nx, nz = g.shape
    ncomp = 1 if physics == "acoustic" else 3
    hts = resid.reshape(-1, 150).copy()
    O.ht(True, hts, 0.001, s="d")
    asrc = -hts[:, ::-1].reshape(ncomp, g.nrcv, 150).copy()
    rinj = O.Injector(0, g.rxyz, g.shape, (0.01,) * 3, (0, 0, 0), s="d")
    sinj = O.Injector(1 if kw.get("mt") else 0, g.sxyz[None], g.shape, (0.01,) * 3, (0, 0, 0), s="d")
    par = O.make_par(150, 2, 0.001, bc=kw["bc"], taper=kw["taper"], taper_strength=0.05,
                     free_surface_stiffness=1.05 if physics != "acoustic" else max(1.05, 1.0))
    ng = 2 if physics == "acoustic" else 3
    G = np.zeros((ng, ny, nx, nz))
    O.propagate(physics, conv, asrc, rinj, sinj, par, (0.01,) * 3, adj=True, grad=G, snaps=snaps.copy(), s="d")
    ops = O.Ops("orc", "d")
    for k in range(ng):
        for ax in (0, 1, 2):
            ops.apply_h(ax, G[k].copy(), G[k], 0.01)
    # J^T by central differences of convert (pointwise, so one perturbation per raw parameter)
    expect = np.zeros((ng, ny, nx, nz))
    for k in range(ng):
        h = 1e-6 * np.abs(clipped[k]).max()
        mp, mm = clipped.copy(), clipped.copy()
        mp[k] += h
        mm[k] -= h
        dconv = (O.convert(mp, nm, s="d").reshape(model.shape) - O.convert(mm, nm, s="d").reshape(model.shape)) / (2 * h)
        expect[k] = sum(dconv[j] * G[j] for j in range(ng))
    assert rel(graw, expect) < 1e-7


def test_fwi_shot_misfit_and_residual():
    model, src, g, kw = small_case("acoustic")
    obs = np.zeros((1, g.nrcv, g.nt_data))
    mis, res, _ = O.fwi_shot(O.op_par(**kw), g, model, src, obs)
    syn, _ = O.apply_forward(O.op_par(**kw), g, model, src)
    assert np.allclose(res, syn)
    w = np.ones(g.nt_data)
    w[0] = w[-1] = 0.5
    assert mis == pytest.approx(0.5 * 0.002 * float(np.sum(syn ** 2 * w)), rel=1e-12)


def test_jacobianT_add_flag():
    model, src, g, kw = small_case("elastic")
    p = O.op_par(**kw)
    d, snaps = O.apply_forward(p, g, model, src)
    g0 = O.apply_jacobianT(p, g, model, d, snaps)
    base = np.random.default_rng(1).standard_normal(g0.shape)
    g1 = O.apply_jacobianT(p, g, model, d, snaps, add=True, grad=base.copy())
    assert np.allclose(g1, base + g0, rtol=1e-12, atol=1e-12 * np.abs(g0).max())
