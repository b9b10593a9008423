"""16-bit snapshot store (BF16 / FP16 codes with a per-snapshot max-abs scale) against
the FP32 store: the forward run is unchanged, the adjoint gradient moves by the code
precision only, and an adjoint must use the forward's format.

Tolerances (relative L2 of the gradient against the FP32-snapshot gradient, which is
itself within 1e-6 of the oracle): FP16 (11-bit significand) <= 1e-3, BF16 (8-bit)
<= 1e-2."""
import numpy as np
import pytest

import fwi3d_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="module")
def eng():
    e = F.Engine(0)
    yield e
    e.close()


def shot(physics, n=28, nt=160):
    from fwi3d_amd import workloads as W
    shape = (n, n + 2, n + 4)
    ny, nx, nz = shape
    h = 0.01
    z = np.arange(nz, dtype=np.float32) * h
    vp = np.broadcast_to(1.8 + 5.0 * z, shape).astype(np.float32).copy()
    rho = W.gardner(vp)
    if physics == "acoustic":
        model, ns, kind = W.acoustic_convert(vp, rho), 1, 0
    else:
        model, ns, kind = W.elastic_convert(vp, vp / 1.8, rho), 6, 1
    src = np.zeros((ns, 1, nt), np.float32)
    src[:3 if ns == 6 else 1, 0] = W.ricker(20, nt, 0.001)
    d = (h, h, h)
    inj = F.Injector(kind, [[0.12, 0.12, 0.08]], shape, d)
    rx = [[0.05 + 0.05 * i, 0.06 + 0.04 * j, 0.03] for i in range(4) for j in range(4)]
    ext = F.Injector(0, rx, shape, d)
    return model, src, inj, ext, d, nt


@pytest.mark.parametrize("physics", ["acoustic", "elastic"])
def test_16bit_snapshots(eng, physics):
    model, src, inj, ext, d, nt = shot(physics)
    ng = 2 if physics == "acoustic" else 3
    grads, traces = {}, {}
    for fmt in ("f32", "f16", "bf16"):
        par = F.make_param(nt, 0.001, sub=3, bc=(1, 2, 2, 2, 2, 2), taper=(0, 5, 5, 5, 5, 5), snap_precision=fmt)
        traces[fmt] = eng.propagate(physics, model, src, inj, ext, par, d)
        res = np.ascontiguousarray(traces[fmt][..., ::-1]) * np.float32(0.3)
        g = np.zeros((ng,) + model.shape[1:], np.float32)
        eng.propagate(physics, model, res, ext, None, par, d, adj=True, grad=g)
        grads[fmt] = g
    assert np.array_equal(traces["f16"], traces["f32"]) and np.array_equal(traces["bf16"], traces["f32"])
    for fmt, tol in (("f16", 1e-3), ("bf16", 1e-2)):
        for k in range(ng):
            e = rel(grads[fmt][k], grads["f32"][k])
            print(f"{physics} {fmt} grad[{k}] rel L2 vs FP32 snapshots {e:.2e}", flush=True)
            assert e <= tol


def test_adjoint_needs_forward_format(eng):
    model, src, inj, ext, d, nt = shot("acoustic", n=24, nt=40)
    par = F.make_param(nt, 0.001, sub=2, snap_precision="f16")
    eng.propagate("acoustic", model, src, inj, ext, par, d)
    par2 = F.make_param(nt, 0.001, sub=2, snap_precision="f32")
    g = np.zeros((2,) + model.shape[1:], np.float32)
    res = np.zeros((1, ext.npts, nt), np.float32)
    with pytest.raises(RuntimeError, match="snap_precision"):
        eng.propagate("acoustic", model, res, ext, None, par2, d, adj=True, grad=g)
