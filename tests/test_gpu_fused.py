"""GPU parity against the oracle, point by point, of the fused elastic / VTI time step
(fwi3d_amd/csrc/fused.hpp, the default form) and of the flux / interior / slab / edge
form. The fused launch updates the planes 6 cells inside the y faces with the stresses
and fluxes of nl_we_op_e / nl_we_op_vti::propagate (src/cpp/we_op.cpp:646-836,
1290-1497) formed on chip, including the x / z closure rows, the traction SATs of the x
and z faces and their absorbing scaling.

The shapes cover each tiling case of an axis (tile widths 16 along x, 32 along z): one
tile, two tiles where the axis is shorter than a tile plus 8 points, and several tiles
with a short middle tile and a last tile shifted to the face; the y chunk of the fused
kernel is forced small so that chunks with remainders occur. Receivers on every grid
node make the traces the whole displacement field at every step (as in
test_gpu_elastic.py). Tolerance: relative L2 <= 1e-4 over all samples and the worst
point of the last sample <= 1e-3 of the field's maximum."""
import numpy as np
import pytest

import fwi3d_amd as F
from oracle import oracle as O

from test_gpu_elastic import CASES, rel, small_shot

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-4


@pytest.fixture(scope="module")
def eng():
    e = F.Engine(0)
    yield e
    e.close()


def every_point(eng, physics, shape, case, nt=40, fchunk=5):
    model, src, sloc, _, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    ny, nx, nz = shape
    h = d[0]
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    rloc = np.stack([ix.ravel() * h, iy.ravel() * h, iz.ravel() * h], axis=1).astype(np.float32)
    cfg = CASES[case]
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=cfg["bc"], taper=cfg["taper"])
    eng.tune(fused=True, fchunk=fchunk)
    try:
        got = eng.propagate(physics, model, src, inj, ext, par, d)
    finally:
        eng.tune(fchunk=0)
    return got, (model, src, sloc, rloc, d, dt, cfg)


@pytest.mark.parametrize("physics", ["elastic", "vti"])
@pytest.mark.parametrize("shape,case", [((30, 45, 70), "mixed"), ((24, 20, 36), "free+absorb+taper"),
                                        ((17, 16, 33), "absorb-all"), ((22, 24, 40), "free-all"),
                                        ((13, 40, 31), "mixed")])
def test_fused_every_point_matches_oracle(eng, physics, shape, case):
    nt = 40
    print(f"[fused {physics} every point] shape={shape} case={case} nt={nt}", flush=True)
    got, (model, src, sloc, rloc, d, dt, cfg) = every_point(eng, physics, shape, case, nt=nt)
    oinj, oext = O.Injector(1, sloc, shape, d, (0, 0, 0)), O.Injector(0, rloc, shape, d, (0, 0, 0))
    opar = O.make_par(nt, 0, dt, bc=cfg["bc"], taper=cfg["taper"])
    want, _ = O.propagate(physics, model, src, oinj, oext, opar, d, s="f")
    e = rel(got, want)
    last = np.abs(got[:, :, -1] - want[:, :, -1]).max() / np.abs(want[:, :, -1]).max()
    # where the largest difference of the last sample sits (component, y, x, z)
    k = int(np.argmax(np.abs(got[:, :, -1] - want[:, :, -1]).max(axis=0)))
    print(f"  rel L2 vs oracle f32 {e:.2e}, worst point of the last sample {last:.2e} at "
          f"(y,x,z)={np.unravel_index(k, shape)}")
    assert np.isfinite(got).all() and np.abs(want[:, :, -1]).max() > 0
    assert e <= TRACE_TOL
    assert last <= 10 * TRACE_TOL


@pytest.mark.parametrize("physics", ["elastic", "vti"])
@pytest.mark.parametrize("shape,case", [((30, 45, 70), "mixed"), ((17, 16, 33), "absorb-all")])
def test_unfused_every_point_matches_oracle(eng, physics, shape, case):
    """The flux / interior / slab / edge form (fwi3d_tune FWI3D_TUNE_FUSED = 0) at every
    point against the oracle."""
    nt = 40
    print(f"[unfused {physics} every point] shape={shape} case={case} nt={nt}", flush=True)
    model, src, sloc, _, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    ny, nx, nz = shape
    h = d[0]
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    rloc = np.stack([ix.ravel() * h, iy.ravel() * h, iz.ravel() * h], axis=1).astype(np.float32)
    cfg = CASES[case]
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=cfg["bc"], taper=cfg["taper"])
    eng.tune(fused=False)
    try:
        got = eng.propagate(physics, model, src, inj, ext, par, d)
    finally:
        eng.tune(fused=True)
    oinj, oext = O.Injector(1, sloc, shape, d, (0, 0, 0)), O.Injector(0, rloc, shape, d, (0, 0, 0))
    opar = O.make_par(nt, 0, dt, bc=cfg["bc"], taper=cfg["taper"])
    want, _ = O.propagate(physics, model, src, oinj, oext, opar, d, s="f")
    e = rel(got, want)
    last = np.abs(got[:, :, -1] - want[:, :, -1]).max() / np.abs(want[:, :, -1]).max()
    print(f"  rel L2 vs oracle f32 {e:.2e}, worst point of the last sample {last:.2e}")
    assert np.isfinite(got).all() and e <= TRACE_TOL and last <= 10 * TRACE_TOL


@pytest.mark.parametrize("physics", ["elastic", "vti"])
def test_fused_equals_unfused(eng, physics):
    """The fused step and the flux / interior / slab / edge kernels compute the same
    field to rounding (a larger grid than the every-point cases, default y chunks)."""
    shape, nt = (48, 66, 100), 60
    print(f"[fused vs unfused {physics}] shape={shape} nt={nt}", flush=True)
    model, src, sloc, rloc, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    out = {}
    for fused in (True, False):
        for case in ("mixed", "free+absorb+taper"):
            cfg = CASES[case]
            par = F.make_param(nt, dt, bc=cfg["bc"], taper=cfg["taper"])
            eng.tune(fused=fused)
            try:
                out[fused, case] = eng.propagate(physics, model, src, inj, ext, par, d)
            finally:
                eng.tune(fused=True)
    for case in ("mixed", "free+absorb+taper"):
        e = rel(out[True, case], out[False, case])
        print(f"  {case}: rel L2 fused vs unfused {e:.2e}")
        assert np.isfinite(out[True, case]).all()
        assert e <= 1e-5


@pytest.mark.parametrize("physics", ["elastic", "vti"])
def test_overlap_modes_identical(eng, physics):
    """The fused step's edge tiles (fwi3d_tune FWI3D_TUNE_OVERLAP 3) or y-face launches
    (1, 2) on the second stream, or all on one stream (0), with and without graph replay:
    the launches write disjoint points of the next field, so the traces are bitwise
    equal."""
    shape, nt = (40, 50, 70), 50
    print(f"[overlap modes {physics}] shape={shape} nt={nt}", flush=True)
    model, src, sloc, rloc, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=CASES["mixed"]["bc"], taper=CASES["mixed"]["taper"])
    out = {}
    try:
        for ov in (0, 1, 2, 3):
            for graphs in (True, False):
                eng.tune(overlap=ov, graphs=graphs)
                out[ov, graphs] = eng.propagate(physics, model, src, inj, ext, par, d)
    finally:
        eng.tune(overlap=3, graphs=True)
    ref = out[0, False]
    assert np.isfinite(ref).all() and np.abs(ref).max() > 0
    for k, v in out.items():
        assert np.array_equal(v, ref), k


@pytest.mark.parametrize("physics", ["elastic", "vti"])
def test_edge_chunks_identical(eng, physics):
    """The y chunk of the fused kernel's side tiles (fwi3d_tune FWI3D_TUNE_ECHUNK; the
    corner tiles take half of it) only changes which workgroup updates a plane: traces
    are bitwise equal for every chunk, including one past the planes' count and the
    autotuner's choice."""
    shape, nt = (40, 50, 70), 40
    print(f"[edge chunks {physics}] shape={shape} nt={nt}", flush=True)
    model, src, sloc, rloc, d, dt = small_shot(physics, 1, shape=shape, nt=nt)
    inj, ext = F.Injector(1, sloc, shape, d), F.Injector(0, rloc, shape, d)
    par = F.make_param(nt, dt, bc=CASES["mixed"]["bc"], taper=CASES["mixed"]["taper"])
    out = {}
    try:
        for ec in (0, 1, 5, 8, 13, 100):
            eng.tune(echunk=ec)
            out[ec] = eng.propagate(physics, model, src, inj, ext, par, d)
        eng.tune(echunk=0)
        eng.autotune(physics, shape, steps=2)
        out["auto"] = eng.propagate(physics, model, src, inj, ext, par, d)
    finally:
        eng.tune(echunk=0)
    ref = out[0]
    assert np.isfinite(ref).all() and np.abs(ref).max() > 0
    for k, v in out.items():
        assert np.array_equal(v, ref), k
