"""GPU parity of the shot operator (fwi3d_apply_forward / fwi3d_apply_jacobianT /
fwi3d_fwi_shot_device) against the oracle's restatement of apply_forward,
apply_jacobianT (src/cpp/we_op.cpp:872-1078) and one shot of compute_res_and_grad
(src/cpp/optimization.cpp:35-188), in FP32 on the same inputs.

Tolerances: traces and misfit relative L2 <= 1e-4, gradients <= 1e-3."""
import numpy as np
import pytest

import fwi3d_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-4
GRAD_TOL = 1e-3


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def eng():
    e = F.Engine(0)
    yield e
    e.close()


def case(physics, n=22, nt_data=70, dt_data=0.002, mt=0, seed=7):
    """Raw model (vp[, vs], rho[, delta, eps, gamma]), wavelets at the data rate, and
    the same geometry for the product and the oracle. With courant 0.4 the CFL step
    rounds to dt_data/2, so the resamplers and a derived sub = 2 are exercised."""
    shape = (n, n + 2, n + 4)
    ny, nx, nz = shape
    z = np.arange(nz) * 0.01
    r = np.random.default_rng(seed)
    vp = np.broadcast_to(1.8 + 6.0 * z, shape) * (1 + 0.02 * r.standard_normal(shape).clip(-2, 2))
    rho = 0.31 * (1000 * vp) ** 0.25
    if physics == "acoustic":
        model, nmodels, nsc = np.stack([vp, rho]), 2, 1
    else:
        parts = [vp, vp / 1.8, rho]
        if physics == "vti":
            parts += [np.full(shape, 0.05), np.full(shape, 0.1), np.full(shape, 0.02)]
        model, nmodels = np.stack(parts), (3 if physics == "elastic" else 6)
        nsc = 6 if mt else 3
    model = model.astype(np.float32)
    t = np.arange(nt_data) * dt_data
    f = 25.0
    tau = t - 1.2 / f
    w = (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)
    src = np.stack([w * (1 + 0.3 * k) for k in range(nsc)]).astype(np.float32)
    rx = [(0.05 + 0.04 * i, 0.06 + 0.03 * j, 0.03) for i in range(3) for j in range(3)]
    sx = (0.1, 0.1, 0.08)
    d, o = (0.01, 0.01, 0.01), (0.0, 0.0, 0.0)
    kw = dict(nmodels=nmodels, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4), sub=-1, dt=-1.0, courant=0.4, mt=mt)
    shot = F.Shot(shape, d, o, nt_data, dt_data, sx, rx)
    geom = O.ShotGeom(shape, d, o, nt_data, dt_data, sx, rx, s="f")
    return model, src, shot, geom, kw


VARIANTS = [
    ("acoustic", dict()),
    ("acoustic", dict(resampling=1)),
    ("acoustic", dict(seismotype=1)),
    ("elastic", dict()),
    ("elastic", dict(mt=1)),
    ("elastic", dict(resampling=1, seismotype=1)),
    ("vti", dict()),
    ("vti", dict(mt=1, seismotype=1)),
]


def _ids(v):
    return v[0] + "".join(f"-{k}{x}" for k, x in v[1].items())


@pytest.mark.parametrize("physics,extra", VARIANTS, ids=[_ids(v) for v in VARIANTS])
def test_forward_and_jacobianT_match_oracle(eng, physics, extra):
    mt = extra.pop("mt", 0) if "mt" in extra else 0
    model, src, shot, geom, kw = case(physics, mt=mt)
    kw.update(extra)
    p = F.op_param(**kw)
    op = F.WeOp(eng, shot, src, p)
    data = op.apply_forward(False, model)
    q = O.op_par(**kw)
    want, snaps = O.apply_forward(q, geom, model, src)
    print(f"{physics} {extra} mt={mt}: nt={p.nt} sub={p.sub} dt={p.dt:.6g} data rel L2 {rel(data, want):.2e}", flush=True)
    assert (p.nt, p.sub) == (q.nt, q.sub) and p.sub == 2
    assert rel(data, want) <= TRACE_TOL
    # adjoint of a residual-like trace set
    resid = (0.5 * want[..., ::-1] + 0.1 * want).astype(np.float32).copy()
    grad = np.zeros((min(3, kw["nmodels"]),) + shot.shape, np.float32)
    op.apply_jacobianT(False, grad, model, resid)
    gwant = O.apply_jacobianT(q, geom, model, resid, snaps)
    for k in range(grad.shape[0]):
        e = rel(grad[k], gwant[k])
        print(f"  grad[{k}] rel L2 {e:.2e}", flush=True)
        assert e <= GRAD_TOL


def test_add_semantics(eng):
    model, src, shot, geom, kw = case("elastic")
    p = F.op_param(**kw)
    op = F.WeOp(eng, shot, src, p)
    d0 = op.apply_forward(False, model)
    base = np.random.default_rng(2).standard_normal(d0.shape).astype(np.float32)
    d1 = op.apply_forward(True, model, base.copy())
    assert np.allclose(d1, base + d0, rtol=1e-5, atol=1e-6 * np.abs(d0).max())
    g0 = np.zeros((3,) + shot.shape, np.float32)
    op.apply_jacobianT(False, g0, model, d0)
    gb = np.random.default_rng(3).standard_normal(g0.shape).astype(np.float32) * np.abs(g0).max()
    g1 = gb.copy()
    op.apply_jacobianT(True, g1, model, d0)
    assert np.allclose(g1, gb + g0, rtol=1e-5, atol=1e-5 * np.abs(g0).max())


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
def test_fwi_shot_device_matches_oracle(eng, physics):
    import torch
    model, src, shot, geom, kw = case(physics)
    fac = np.ones(model.shape[0], np.float32)
    fac[0] = 1.03
    obs, _ = O.apply_forward(O.op_par(**kw), geom, model * fac[:, None, None, None], src)
    q = O.op_par(**kw)
    mis_w, res_w, grad_w = O.fwi_shot(q, geom, model, src, obs)
    dev = torch.device("cuda:0")
    dm = torch.from_numpy(model).to(dev)
    ds = torch.from_numpy(src).to(dev)
    do = torch.from_numpy(obs).to(dev)
    dr = torch.zeros_like(do)
    ng = min(3, kw["nmodels"])
    g64 = torch.zeros((ng,) + shot.shape, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    p = F.op_param(**kw)
    mis = F.fwi_shot_device(eng, p, shot, dm.data_ptr(), ds.data_ptr(), do.data_ptr(), g64.data_ptr(), dr.data_ptr())
    torch.cuda.synchronize()
    print(f"{physics}: misfit {mis:.6e} vs {mis_w:.6e}", flush=True)
    assert abs(mis - mis_w) / abs(mis_w) <= TRACE_TOL
    assert rel(dr.cpu().numpy(), res_w) <= TRACE_TOL
    g = g64.cpu().numpy()
    for k in range(ng):
        e = rel(g[k], grad_w[k])
        print(f"  grad[{k}] rel L2 {e:.2e}", flush=True)
        assert e <= GRAD_TOL
    # a second shot accumulates into the same FP64 buffer
    p2 = F.op_param(**kw)
    F.fwi_shot_device(eng, p2, shot, dm.data_ptr(), ds.data_ptr(), do.data_ptr(), g64.data_ptr())
    torch.cuda.synchronize()
    assert rel(g64.cpu().numpy(), 2 * g) < 1e-6


def test_jacobianT_without_forward_is_an_error(eng):
    """The adjoint needs the snapshots of a matching forward on the same context."""
    model, src, shot, geom, kw = case("acoustic")
    fresh = F.Engine(0)
    try:
        p = F.op_param(**kw)
        op = F.WeOp(fresh, shot, src, p)
        grad = np.zeros((2,) + shot.shape, np.float32)
        with pytest.raises(RuntimeError, match="snapshots"):
            op.apply_jacobianT(False, grad, model, np.zeros((1, shot.nrcv, shot.nt_data), np.float32))
    finally:
        fresh.close()
