"""Serving loop (fwi3d_server_*): shot-modelling and shot-gradient requests of the three
physics on uneven grids arrive while others run; the server admits them by HBM
footprint into one continuously refilled batch. Each result in the mix equals, bit for
bit, what the one-shot entry points (fwi3d_apply_forward_device, fwi3d_fwi_shot_device)
compute for that request alone; the FP64 misfit, summed with atomics in any order,
agrees to 1e-12 relative."""
import time

import numpy as np
import pytest

import fwi3d_amd as F
from fwi3d_amd import workloads as W

pytestmark = pytest.mark.gpu

H = 0.01


def raw_model(physics, shape, seed):
    vp = (2.0 + 1.2 * W.smooth_random(shape, 4, seed)).astype(np.float32)
    rho = W.gardner(vp)
    if physics == "acoustic":
        return np.stack([vp, rho])
    vs = (vp / np.float32(1.8)).astype(np.float32)
    if physics == "elastic":
        return np.stack([vp, vs, rho])
    dlt = (0.1 * W.smooth_random(shape, 3, seed + 1)).astype(np.float32)
    eps = (0.2 * W.smooth_random(shape, 3, seed + 2)).astype(np.float32)
    return np.stack([vp, vs, rho, dlt, eps, 0 * eps])


# kind, physics, shape (ny, nx, nz), nt_data, dt_data, op-param overrides
REQS = [
    ("model", "acoustic", (40, 44, 48), 150, 0.0008, dict(dt=0.0008)),
    ("gradient", "elastic", (30, 34, 38), 120, 0.0016, dict(dt=-1.0, sub=-1, mt=1)),       # CFL dt, sub = data/dt
    ("model", "vti", (28, 26, 40), 110, 0.0008, dict(dt=0.0008, mt=1)),
    ("gradient", "acoustic", (36, 30, 70), 130, 0.0008, dict(dt=0.0008, sub=4, resampling="linear")),
    ("model", "elastic", (26, 32, 30), 100, 0.0008, dict(dt=0.0008, seismotype=1)),
    ("gradient", "vti", (28, 30, 32), 90, 0.0008, dict(dt=0.0008, sub=3, mt=1, snap_precision="bf16")),
    ("gradient", "elastic", (24, 28, 40), 100, 0.0008, dict(dt=0.0008, sub=4, snap_precision="f16")),
]


def build_request(k, kind, physics, shape, nt, dt, over, dev):
    import torch
    rng = np.random.default_rng(500 + k)
    ny, nx, nz = shape
    raw = raw_model(physics, shape, 40 + k)
    nm = raw.shape[0]
    ns = 1 if nm == 2 else (6 if over.get("mt") else 3)
    w = W.ricker(20.0, nt, dt)
    src = np.stack([w * (1.0 + 0.2 * q) for q in range(ns)]).astype(np.float32)
    L = np.array([(nx - 1) * H, (ny - 1) * H, (nz - 1) * H])
    sxyz = rng.uniform(0.35, 0.65, 3) * L
    rxyz = np.stack([rng.uniform(0.05, 0.95, 3) * L for _ in range(12)])
    geo = F.Shot(shape, (H, H, H), (0.0, 0.0, 0.0), nt, dt, sxyz, rxyz)
    kw = dict(nmodels=nm, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4))
    kw.update(over)
    nrc = 1 if nm == 2 else 3
    ng = min(3, nm)
    req = dict(kind=kind, physics=physics, geo=geo, kw=kw, nrc=nrc, ng=ng,
               model=torch.from_numpy(raw).to(dev), src=torch.from_numpy(src).to(dev),
               data=torch.zeros((nrc, geo.nrcv, nt), dtype=torch.float32, device=dev),
               obs=None, grad=None)
    if kind == "gradient":
        obs = (0.01 * rng.standard_normal((nrc, geo.nrcv, nt))).astype(np.float32)
        req["obs"] = torch.from_numpy(obs).to(dev)
        req["grad"] = torch.zeros((ng,) + tuple(shape), dtype=torch.float64, device=dev)
    return req


def alone(eng, q):
    """The one-shot entry point for request q on its own context."""
    import torch
    par = F.op_param(**q["kw"])
    data = torch.zeros_like(q["data"])
    if q["kind"] == "model":
        F.apply_forward_device(eng, par, q["geo"], False, q["model"].data_ptr(), q["src"].data_ptr(), data.data_ptr())
        torch.cuda.synchronize()
        return data.cpu().numpy(), None, None
    grad = torch.zeros_like(q["grad"])
    mis = F.fwi_shot_device(eng, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(), q["obs"].data_ptr(),
                            grad.data_ptr(), data.data_ptr())
    torch.cuda.synchronize()
    return data.cpu().numpy(), grad.cpu().numpy(), mis


def submit(srv, q):
    par = F.op_param(**q["kw"])
    if q["kind"] == "model":
        return srv.submit(F.REQ_MODEL, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(),
                          data_ptr=q["data"].data_ptr())
    return srv.submit(F.REQ_GRADIENT, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(),
                      data_ptr=q["data"].data_ptr(), obs_ptr=q["obs"].data_ptr(), grad64_ptr=q["grad"].data_ptr())


def test_serve_mix_equals_alone():
    import torch
    dev = torch.device("cuda:0")
    reqs = [build_request(k, *r, dev) for k, r in enumerate(REQS)]
    torch.cuda.synchronize()
    ref_eng = F.Engine(0)
    try:
        want = []
        for k, q in enumerate(reqs):
            print(f"  alone: request {k} {q['kind']} {q['physics']} {q['geo'].shape}", flush=True)
            want.append(alone(ref_eng, q))
    finally:
        ref_eng.close()
    eng = F.Engine(0)
    srv = F.Server(eng, hbm_budget=8 << 30, max_active=4)
    try:
        print("  serving: 3 requests, then the rest while they run", flush=True)
        ids = [submit(srv, q) for q in reqs[:3]]
        t0 = time.time()
        while srv.poll(ids[0])[0] == 0 and time.time() - t0 < 60:
            time.sleep(0.001)
        for q in reqs[3:]:
            ids.append(submit(srv, q))
            time.sleep(0.002)
        srv.wait()
        st = srv.stats()
        print(f"  counters: {st}", flush=True)
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["queued"] == 0 and st["active"] == 0
        assert st["peak_active"] >= 2 and st["hbm_in_use"] == 0 and st["steps"] > 0
        for k, (q, rid) in enumerate(zip(reqs, ids)):
            state, mis, lat, qs = srv.poll(rid)
            assert state == 2, srv.error(rid)
            d_want, g_want, m_want = want[k]
            d = q["data"].cpu().numpy()
            msg = f"  request {k} {q['kind']} {q['physics']}: latency {lat * 1e3:.1f} ms (queued {qs * 1e3:.1f} ms)"
            assert np.abs(d_want).max() > 0
            assert np.array_equal(d, d_want), k  # data, or residual of a gradient request
            if q["kind"] == "gradient":
                g = q["grad"].cpu().numpy()
                msg += f", misfit {mis:.6e} vs {m_want:.6e}, gradient max|d| {np.abs(g - g_want).max():.1e}"
                assert np.abs(g_want).max() > 0
                assert np.array_equal(g, g_want), k
                assert abs(mis - m_want) <= 1e-12 * abs(m_want), k
                p = srv.result_param(rid)
                assert p.nt >= 2 and p.sub > 0
            print(msg, flush=True)
    finally:
        srv.close()
        eng.close()


def test_serve_budget_admission_and_errors():
    """A request larger than the whole budget fails with a clear error while the others
    complete; the budget keeps the reserved bytes at or under it."""
    import torch
    dev = torch.device("cuda:0")
    small = [build_request(k, *REQS[k], dev) for k in (0, 2)]
    big = build_request(9, "model", "elastic", (100, 100, 100), 80, 0.0008, dict(dt=0.0008), dev)
    torch.cuda.synchronize()
    eng = F.Engine(0)
    budget = 40 << 20  # fits one small request at a time, never the 100^3 elastic one
    srv = F.Server(eng, hbm_budget=budget, max_active=8)
    try:
        ids = [submit(srv, small[0]), submit(srv, big), submit(srv, small[1])]
        srv.wait()
        st = srv.stats()
        print(f"  counters: {st}", flush=True)
        assert srv.poll(ids[0])[0] == 2 and srv.poll(ids[2])[0] == 2
        assert srv.poll(ids[1])[0] == -1 and "budget" in srv.error(ids[1])
        assert st["failed"] == 1 and st["completed"] == 2 and 0 < st["hbm_peak"] <= budget
        with pytest.raises(RuntimeError, match="unknown request"):
            srv.poll(12345)
    finally:
        srv.close()
        eng.close()
