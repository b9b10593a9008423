"""GPU parity of the elastic isotropic and VTI paths (nl_we_op_e / nl_we_op_vti
::propagate and ::compute_gradients, src/cpp/we_op.cpp:488-870, 1108-1531) through
the C ABI, against the oracle on the same inputs.

Tolerances (north_star): 3-C traces relative L2 <= 1e-4 against the single-precision
oracle; gradients relative L2 <= 1e-3."""
import numpy as np
import pytest

import fwi3d_amd as F
from fwi3d_amd import workloads as W
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


def small_model(physics, shape, seed=5):
    vp = (2.0 + 1.3 * W.smooth_random(shape, 4, seed)).astype(np.float32)
    vs = (vp / (1.75 + 0.2 * W.smooth_random(shape, 3, seed + 1))).astype(np.float32)
    rho = W.gardner(vp)
    if physics == "elastic":
        return W.elastic_convert(vp, vs, rho)
    eps = (0.2 * W.smooth_random(shape, 3, seed + 2)).astype(np.float32)
    dlt = (0.1 * W.smooth_random(shape, 3, seed + 3)).astype(np.float32)
    gam = (0.1 * W.smooth_random(shape, 3, seed + 4)).astype(np.float32)
    return W.vti_convert(vp, vs, rho, dlt, eps, gam)


def small_shot(physics, kind, shape=(30, 34, 38), nt=200, dt=0.0008, h=0.01, seed=5):
    ny, nx, nz = shape
    rng = np.random.default_rng(seed)
    model = small_model(physics, shape, seed)
    w = W.ricker(30.0, nt, dt)
    if kind == 1:  # moment tensor (mxx, myy, mzz, mxy, mxz, myz)
        src = np.stack([w, 0.8 * w, 1.2 * w, 0.3 * w, -0.2 * w, 0.1 * w])[:, None, :].astype(np.float32)
    else:          # vector force (fx, fy, fz)
        src = np.stack([0.4 * w, -0.3 * w, w])[:, None, :].astype(np.float32)
    sloc = np.array([[0.45 * (nx - 1) * h, 0.55 * (ny - 1) * h, 0.4 * (nz - 1) * h]], np.float32)
    L = np.array([(nx - 1) * h, (ny - 1) * h, (nz - 1) * h])
    pts = [rng.uniform(0, 1, 3) * L for _ in range(30)]
    for ax in range(3):
        for end in (0.0, 1.0):
            p = rng.uniform(0.2, 0.8, 3) * L
            p[ax] = end * L[ax]
            pts.append(p)
    return model, src, sloc, np.array(pts, np.float32), (h, h, h), dt


CASES = {
    "free+absorb+taper": dict(bc=(1, 2, 2, 2, 2, 2), taper=(0, 6, 6, 6, 6, 6)),
    "absorb-all": dict(bc=(2, 2, 2, 2, 2, 2), taper=(0,) * 6),
    "free-all": dict(bc=(1, 1, 1, 1, 1, 1), taper=(0,) * 6),
    "mixed": dict(bc=(2, 1, 0, 2, 1, 2), taper=(3, 0, 4, 0, 0, 5)),
}


def run_pair(eng, physics, kind, cfg, nt=200, sub=0, shape=(30, 34, 38)):
    model, src, sloc, rloc, d, dt = small_shot(physics, kind, shape=shape, nt=nt)
    shape = model.shape[-3:]
    inj, ext = F.Injector(kind, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, sub=sub, bc=cfg["bc"], taper=cfg["taper"])
    got = eng.propagate(physics, model, src, inj, ext, par, d)
    oinj, oext = O.Injector(kind, sloc, shape, d, (0, 0, 0)), O.Injector(0, rloc, shape, d, (0, 0, 0))
    opar = O.make_par(nt, sub, dt, bc=cfg["bc"], taper=cfg["taper"])
    want, snaps = O.propagate(physics, model, src, oinj, oext, opar, d, s="f")
    return got, want, snaps, (model, src, sloc, rloc, d, dt, inj, ext, par, oinj, oext, opar)


@pytest.mark.parametrize("physics", ["elastic", "vti"])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kind", [1, 0])
def test_elastic_forward_matches_oracle(eng, physics, case, kind):
    print(f"[{physics} fwd] case={case} src_kind={kind}", flush=True)
    got, want, _, _ = run_pair(eng, physics, kind, CASES[case])
    e = rel(got, want)
    print(f"  rel L2 vs oracle f32 {e:.2e}")
    assert np.isfinite(got).all() and np.abs(want).max() > 0
    assert e <= TRACE_TOL


@pytest.mark.parametrize("physics", ["elastic", "vti"])
def test_elastic_tile_edges_near_faces(eng, physics):
    """nz = 66 and nx = 27: the last z tile (z 64-65) and the last x tiles end within 3
    points of the high faces, so those D1 closure rows of the flux pass read global
    memory while the rows of the other tiles read the staged plane; both against the
    oracle."""
    print(f"[{physics} tile edges] shape=(20, 27, 66)", flush=True)
    got, want, _, _ = run_pair(eng, physics, 1, CASES["mixed"], nt=120, shape=(20, 27, 66))
    e = rel(got, want)
    print(f"  rel L2 vs oracle f32 {e:.2e}")
    assert np.isfinite(got).all() and np.abs(want).max() > 0
    assert e <= TRACE_TOL


@pytest.mark.parametrize("physics", ["elastic", "vti"])
def test_elastic_forward_fp64(eng, physics):
    """Error against the FP64 oracle stays at the level of the FP32 oracle's."""
    cfg = CASES["free+absorb+taper"]
    got, want, _, a = run_pair(eng, physics, 1, cfg)
    model, src, sloc, rloc, d, dt, inj, ext, par, oinj, oext, opar = a
    want64, _ = O.propagate(physics, model.astype(np.float64), src.astype(np.float64), oinj.as_double(),
                            oext.as_double(), opar, d, s="d")
    e64, o64 = rel(got, want64), rel(want, want64)
    print(f"[{physics}] vs f64 {e64:.2e} (oracle f32 vs f64 {o64:.2e})")
    assert e64 <= max(TRACE_TOL, 3 * o64)


@pytest.mark.parametrize("physics", ["elastic", "vti"])
@pytest.mark.parametrize("sub", [4, 3])
def test_elastic_gradient_matches_oracle(eng, physics, sub):
    """Forward with snapshots, adjoint with a residual at the receivers, imaging
    condition (g_lambda, g_mu, g_rho) including the lagged adjoint derivatives."""
    nt = 201
    cfg = CASES["free+absorb+taper"]
    print(f"[{physics} grad] sub={sub} nt={nt}", flush=True)
    got, want, snaps, a = run_pair(eng, physics, 1, cfg, nt=nt, sub=sub)
    model, src, sloc, rloc, d, dt, inj, ext, par, oinj, oext, opar = a
    assert rel(got, want) <= TRACE_TOL
    shape = model.shape[-3:]
    res = np.ascontiguousarray(got[:, :, ::-1] * 0.5)
    grad = np.zeros((3,) + shape, np.float32)
    eng.propagate(physics, model, res, ext, None, par, d, adj=True, grad=grad)
    ograd = np.zeros((3,) + shape, np.float32)
    O.propagate(physics, model, res, oext, oinj, opar, d, adj=True, grad=ograd, snaps=snaps, s="f")
    for k in range(3):
        e = rel(grad[k], ograd[k])
        print(f"  grad[{k}] rel L2 {e:.2e}")
        assert np.abs(ograd[k]).max() > 0
        assert e <= GRAD_TOL


@pytest.mark.parametrize("physics", ["elastic", "vti"])
@pytest.mark.parametrize("shape,case", [((19, 17, 45), "mixed"), ((13, 26, 14), "free-all"), ((21, 14, 40), "absorb-all"),
                                        # an axis of exactly 12 points: no interior tiles, and no slab
                                        # tiles at all for (40, 12, 12) (the edge kernel covers the grid)
                                        ((40, 12, 30), "mixed"), ((40, 12, 12), "absorb-all"),
                                        ((12, 20, 12), "free-all")])
def test_elastic_every_point_matches_oracle(eng, physics, shape, case):
    """Receivers on every grid node (quadratic extraction weights reduce to the node
    value), so the traces are the whole displacement field at every step: each point of
    the interior, the z slabs, the x / y side slabs (several 32-point z tiles and march
    chunks with remainders) and the shell's edges and corners is checked against the
    oracle, not only the points that reach a few receivers."""
    nt = 40
    print(f"[{physics} every point] shape={shape} case={case} nt={nt}", flush=True)
    model, src, sloc, _, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    ny, nx, nz = shape
    h = d[0]
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    rloc = np.stack([ix.ravel() * h, iy.ravel() * h, iz.ravel() * h], axis=1).astype(np.float32)
    cfg = CASES[case]
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=cfg["bc"], taper=cfg["taper"])
    got = eng.propagate(physics, model, src, inj, ext, par, d)
    oinj, oext = O.Injector(1, sloc, shape, d, (0, 0, 0)), O.Injector(0, rloc, shape, d, (0, 0, 0))
    opar = O.make_par(nt, 0, dt, bc=cfg["bc"], taper=cfg["taper"])
    want, _ = O.propagate(physics, model, src, oinj, oext, opar, d, s="f")
    e = rel(got, want)
    # the worst point of the last sample, relative to the field's largest value
    last = np.abs(got[:, :, -1] - want[:, :, -1]).max() / np.abs(want[:, :, -1]).max()
    print(f"  rel L2 vs oracle f32 {e:.2e}, worst point of the last sample {last:.2e}")
    assert np.isfinite(got).all() and np.abs(want[:, :, -1]).max() > 0
    assert e <= TRACE_TOL
    assert last <= 10 * TRACE_TOL
