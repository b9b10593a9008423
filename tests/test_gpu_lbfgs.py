"""GPU l-BFGS (fwi3d_lbfgs_run: device vectors, FP64 dots and curvature pairs) against
the oracle's restatement of lbfgs::run + regular_wolfe (oracle/lbfgs.py) on the same
objective: the same iteration path (iterations, evaluations) and misfit history."""
import numpy as np
import pytest

import fwi3d_amd as F
from oracle import lbfgs as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = F.Engine(0)
    yield e
    e.close()


def test_quadratic_matches_oracle(eng):
    import torch
    dev = torch.device("cuda:0")
    n = 4096
    r = np.random.default_rng(11)
    d = r.uniform(0.2, 5.0, n)
    w = r.standard_normal(n) / np.sqrt(n)
    t = r.standard_normal(n)

    def f_np(x):
        e = x - t
        we = float(w @ e)
        return 0.5 * float(np.sum(d * e * e)) + 0.5 * we * we + 1.0, d * e + w * we

    x = torch.zeros(n, dtype=torch.float32, device=dev)
    g = torch.zeros_like(x)
    D, W, Tt = (torch.from_numpy(a).to(dev) for a in (d, w, t))

    def objective(mp, gp):
        assert mp == x.data_ptr() and gp == g.data_ptr()
        e = x.double() - Tt
        we = torch.dot(W, e)
        g.copy_((D * e + W * we).float())
        torch.cuda.synchronize()
        return float(0.5 * torch.sum(D * e * e) + 0.5 * we * we + 1.0)

    hist, (it, fe, ge) = F.lbfgs_run(eng, n, x.data_ptr(), g.data_ptr(), objective, niter=25)
    xo, hist_o, cnt = L.lbfgs(f_np, np.zeros(n), niter=25)
    print(f"gpu: it={it} feval={fe} geval={ge} f={hist[-1]:.9g}; oracle: it={cnt['iterations']} "
          f"feval={cnt['feval']} f={hist_o[-1]:.9g}", flush=True)
    assert it == cnt["iterations"] and fe == cnt["feval"]
    k = min(len(hist), 12)
    assert np.allclose(hist[:k], hist_o[:k], rtol=1e-5)
    assert np.allclose(x.cpu().numpy(), xo, atol=2e-4)


def test_acoustic_fwi_matches_oracle(eng):
    """Three l-BFGS iterations of a one-shot acoustic FWI: objective = fwi3d_fwi_shot_device
    (FP64 misfit, FP64 gradient accumulation) vs the oracle's FP32 fwi_shot."""
    import torch
    dev = torch.device("cuda:0")
    shape = (20, 22, 24)
    ny, nx, nz = shape
    zz = np.arange(nz) * 0.01
    vp = np.broadcast_to(1.8 + 4.0 * zz, shape).astype(np.float32).copy()
    rho = (0.31 * (1000 * vp) ** 0.25).astype(np.float32)
    true = np.stack([vp * np.float32(1.04), rho])
    start = np.stack([vp, rho]).astype(np.float32)
    nt, dtd = 80, 0.002
    tt = np.arange(nt) * dtd
    tau = tt - 1.2 / 25.0
    src = ((1 - 2 * (np.pi * 25 * tau) ** 2) * np.exp(-(np.pi * 25 * tau) ** 2))[None].astype(np.float32)
    rx = [(0.04 + 0.04 * i, 0.04 + 0.04 * j, 0.02) for i in range(4) for j in range(4)]
    kw = dict(nmodels=2, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4), sub=-1, dt=-1.0, courant=0.4)
    shot = F.Shot(shape, (0.01,) * 3, (0, 0, 0), nt, dtd, (0.1, 0.1, 0.06), rx)
    geom = O.ShotGeom(shape, (0.01,) * 3, (0, 0, 0), nt, dtd, (0.1, 0.1, 0.06), rx, s="f")
    obs, _ = O.apply_forward(O.op_par(**kw), geom, true, src)

    def f_oracle(m):
        mis, _, grad = O.fwi_shot(O.op_par(**kw), geom, m.astype(np.float32).reshape(start.shape), src, obs)
        return mis, grad.astype(np.float64).ravel()

    xo, hist_o, cnt = L.lbfgs(f_oracle, start.ravel().astype(np.float64), niter=3)

    m = torch.from_numpy(start.copy()).to(dev)
    g = torch.zeros_like(m)
    g64 = torch.zeros(m.shape, dtype=torch.float64, device=dev)
    ds = torch.from_numpy(src).to(dev)
    do = torch.from_numpy(obs).to(dev)

    def objective(mp, gp):
        g64.zero_()
        torch.cuda.synchronize()
        mis = F.fwi_shot_device(eng, F.op_param(**kw), shot, mp, ds.data_ptr(), do.data_ptr(), g64.data_ptr())
        g.copy_(g64.float())
        torch.cuda.synchronize()
        return mis

    hist, (it, fe, ge) = F.lbfgs_run(eng, m.numel(), m.data_ptr(), g.data_ptr(), objective, niter=3)
    print(f"gpu f: {hist} it={it} feval={fe}; oracle f: {hist_o} it={cnt['iterations']} feval={cnt['feval']}",
          flush=True)
    assert it == cnt["iterations"] and fe == cnt["feval"]
    assert np.allclose(hist, hist_o, rtol=1e-3)
    assert hist[-1] < hist[0]
    mo = xo.reshape(start.shape)
    assert np.linalg.norm(m.cpu().numpy() - mo) / np.linalg.norm(mo - start) < 1e-2
