"""The allocation-free propagate entry point (fwi3d_propagate_device_noalloc) against
the allocating one (fwi3d_propagate_device) through the C ABI: the same kernels over the
library's static device workspace instead of context buffers, so the traces must agree
bit for bit. Two calls in a row on different grids (the second smaller) check that the
workspace's previous contents do not leak into a call; an oversized shot is rejected."""
import numpy as np
import pytest

import fwi3d_amd as F

from test_gpu_elastic import small_shot

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = F.Engine(0)
    yield e
    e.close()


def _run(eng, physics, shape, adj=False, noalloc=False, nt=60):
    import torch
    dev = torch.device("cuda:0")
    kind = 1
    if physics == "acoustic":
        from fwi3d_amd import workloads as W
        ny, nx, nz = shape
        h, dt = 0.01, 0.001
        vp = (1.6 + 0.8 * W.smooth_random(shape, 4, 3)).astype(np.float32)
        model = W.acoustic_convert(vp, W.gardner(vp))
        src = W.ricker(25.0, nt, dt)[None, None, :].astype(np.float32)
        sloc = np.array([[0.5 * (nx - 1) * h, 0.5 * (ny - 1) * h, 0.3 * (nz - 1) * h]], np.float32)
        rloc = np.array([[0.1 * (i + 1) * (nx - 1) * h, 0.4 * (ny - 1) * h, 0.1 * (nz - 1) * h] for i in range(8)],
                        np.float32)
        d, kind = (h, h, h), 0
    else:
        model, src, sloc, rloc, d, dt = small_shot(physics, kind, shape=shape, nt=nt)
        src = src[:, :, :nt].copy()
    shape = model.shape[-3:]
    inj, ext = F.Injector(kind, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4))
    ncomp = 1 if physics == "acoustic" else 3
    m = torch.from_numpy(np.ascontiguousarray(model)).to(dev)
    if adj:
        # adjoint: receiver traces injected as the source
        s = torch.from_numpy(np.ascontiguousarray(np.random.default_rng(7).standard_normal(
            (ncomp, len(rloc), nt)).astype(np.float32))).to(dev)
        r = None
        a, b = ext, inj
    else:
        s = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
        r = torch.zeros((ncomp, len(rloc), nt), dtype=torch.float32, device=dev)
        a, b = inj, ext
    torch.cuda.synchronize()
    f = eng.propagate_device_noalloc if noalloc else eng.propagate_device
    f(physics, m.data_ptr(), s.data_ptr(), r.data_ptr() if r is not None else 0, a, b, par, shape, d, adj=adj)
    torch.cuda.synchronize()
    return r.cpu().numpy() if r is not None else None


@pytest.mark.parametrize("physics", ["acoustic", "elastic", "vti"])
def test_noalloc_equals_propagate_device(eng, physics):
    for shape in ((28, 40, 70), (17, 20, 33)):  # the second call smaller than the first
        print(f"[noalloc {physics}] shape={shape}", flush=True)
        want = _run(eng, physics, shape)
        got = _run(eng, physics, shape, noalloc=True)
        assert np.isfinite(got).all() and np.abs(want).max() > 0
        assert np.array_equal(got, want)


def test_noalloc_adjoint_and_limits(eng):
    import torch
    print("[noalloc adjoint]", flush=True)
    _run(eng, "elastic", (24, 30, 40), adj=True, noalloc=True)
    # a shot whose buffers exceed the workspace is rejected before any device work
    from fwi3d_amd import workloads as W
    shape = (160, 160, 160)
    d = (0.01, 0.01, 0.01)
    inj = F.Injector(1, np.array([[0.8, 0.8, 0.5]], np.float32), shape, d)
    ext = F.Injector(0, np.array([[0.5, 0.5, 0.2]], np.float32), shape, d)
    par = F.make_param(10, 0.0008)
    m = torch.zeros((3,) + shape, dtype=torch.float32, device="cuda:0")
    s = torch.zeros((6, 1, 10), dtype=torch.float32, device="cuda:0")
    r = torch.zeros((3, 1, 10), dtype=torch.float32, device="cuda:0")
    with pytest.raises(RuntimeError, match="workspace"):
        eng.propagate_device_noalloc("elastic", m.data_ptr(), s.data_ptr(), r.data_ptr(), inj, ext, par, shape, d)
    # and a snapshot store is refused
    par = F.make_param(10, 0.0008, sub=2)
    small = (20, 20, 20)
    with pytest.raises(RuntimeError, match="sub must be 0"):
        eng.propagate_device_noalloc("elastic", m.data_ptr(), s.data_ptr(), r.data_ptr(),
                                     F.Injector(1, np.array([[0.1, 0.1, 0.1]], np.float32), small, d),
                                     F.Injector(0, np.array([[0.1, 0.1, 0.05]], np.float32), small, d), par, small, d)
    del W
