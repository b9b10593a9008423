"""The tiling of the time-step kernels (fwi3d_tune: x rows per workgroup, y planes per
workgroup) changes the schedule, not the arithmetic of any point: traces and
gradients are bit-identical across tilings."""
import numpy as np
import pytest

import fwi3d_amd as F

pytestmark = pytest.mark.gpu


def run(physics, tx, ychunk, autotune=False):
    from fwi3d_amd import workloads as W
    shape = (21, 27, 70)  # nz > 64: two z tiles
    ny, nx, nz = shape
    vp = np.broadcast_to(1.8 + 3.0 * np.arange(nz, dtype=np.float32) * 0.01, shape).astype(np.float32).copy()
    rho = W.gardner(vp)
    if physics == "acoustic":
        model, ns, kind = W.acoustic_convert(vp, rho), 1, 0
    elif physics == "elastic":
        model, ns, kind = W.elastic_convert(vp, vp / 1.8, rho), 6, 1
    else:
        z = np.full(shape, 0.1, np.float32)
        model, ns, kind = W.vti_convert(vp, vp / 2, rho, 0.5 * z, z, 0 * z), 6, 1
    nt = 60
    src = np.zeros((ns, 1, nt), np.float32)
    src[:3 if ns == 6 else 1, 0] = W.ricker(25, nt, 0.001)
    d = (0.01, 0.01, 0.01)
    inj = F.Injector(kind, [[0.13, 0.1, 0.3]], shape, d)
    ext = F.Injector(0, [[0.05 + 0.04 * i, 0.05 + 0.03 * j, 0.03] for i in range(4) for j in range(4)], shape, d)
    eng = F.Engine(0)
    try:
        eng.tune(ychunk=ychunk, tx=tx)
        if autotune:
            t = eng.autotune(physics, shape, steps=4)
            print(f"[{physics}] autotuned tiles (tx, ychunk) = {t}", flush=True)
            assert t[0] in (4, 8) and t[1] in (4, 8, 16, 32)
        par = F.make_param(nt, 0.001, sub=3, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4))
        rcv = eng.propagate(physics, model, src, inj, ext, par, d)
        g = np.zeros((2 if physics == "acoustic" else 3,) + shape, np.float32)
        eng.propagate(physics, model, np.ascontiguousarray(rcv[..., ::-1]), ext, None, par, d, adj=True, grad=g)
    finally:
        eng.close()
    return rcv, g


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
def test_tilings_agree_bitwise(physics):
    r0, g0 = run(physics, 4, 0)
    for tx, yc in ((8, 7), (4, 64), (8, 32)):
        r, g = run(physics, tx, yc)
        assert np.array_equal(r, r0) and np.array_equal(g, g0), (tx, yc)


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
def test_autotune_keeps_results(physics):
    """The autotuner's choice (measured on a synthetic model) then drives propagate on
    that grid; traces and gradients are unchanged."""
    r0, g0 = run(physics, 4, 0)
    r, g = run(physics, 4, 0, autotune=True)
    assert np.array_equal(r, r0) and np.array_equal(g, g0)


def test_tune_rejects_bad_values():
    eng = F.Engine(0)
    try:
        with pytest.raises(RuntimeError):
            eng.tune(tx=6)
        with pytest.raises(RuntimeError):
            eng.tune(ychunk=-1)
    finally:
        eng.close()


def fwd_traces(physics, nt, graphs):
    from fwi3d_amd import workloads as W
    shape = (21, 27, 70)
    ny, nx, nz = shape
    vp = np.broadcast_to(1.8 + 3.0 * np.arange(nz, dtype=np.float32) * 0.01, shape).astype(np.float32).copy()
    rho = W.gardner(vp)
    if physics == "acoustic":
        model, ns, kind = W.acoustic_convert(vp, rho), 1, 0
    elif physics == "elastic":
        model, ns, kind = W.elastic_convert(vp, vp / 1.8, rho), 6, 1
    else:
        z = np.full(shape, 0.1, np.float32)
        model, ns, kind = W.vti_convert(vp, vp / 2, rho, 0.5 * z, z, 0 * z), 6, 1
    src = np.zeros((ns, 1, nt), np.float32)
    src[:3 if ns == 6 else 1, 0] = W.ricker(25, nt, 0.001)
    d = (0.01, 0.01, 0.01)
    inj = F.Injector(kind, [[0.13, 0.1, 0.3]], shape, d)
    ext = F.Injector(0, [[0.05 + 0.04 * i, 0.05 + 0.03 * j, 0.03] for i in range(4) for j in range(4)], shape, d)
    eng = F.Engine(0)
    try:
        eng.tune(graphs=graphs)
        par = F.make_param(nt, 0.001, sub=0, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4))
        return eng.propagate(physics, model, src, inj, ext, par, d)
    finally:
        eng.close()


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
@pytest.mark.parametrize("nt", [20, 33, 60])
def test_graph_replay_equals_eager(physics, nt):
    """Forward propagation without snapshots replays a captured graph of 16 steps whose
    kernels read the step index from a device counter (nt=20: too short, eager; 33: two
    windows exactly; 60: three windows and an eager remainder): bit-identical traces."""
    print(f"  {physics} nt={nt}", flush=True)
    r_graph = fwd_traces(physics, nt, True)
    r_eager = fwd_traces(physics, nt, False)
    assert np.abs(r_eager).max() > 0
    assert np.array_equal(r_graph, r_eager)
