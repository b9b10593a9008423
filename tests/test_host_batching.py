"""Host-side logic of batching and serving that needs no GPU: the device footprints the
batch and the server admit by, and the seeded request stream of the serving benchmark."""
import numpy as np
import pytest

import fwi3d_amd as F
from fwi3d_amd import workloads as W


def padded(n):
    return (n + 63) // 64 * 64


@pytest.mark.parametrize("physics,ncomp", [("acoustic", 1), ("elastic", 3), ("vti", 3)])
@pytest.mark.parametrize("nt,sub", [(100, 4), (101, 4), (90, 3), (2000, 4)])
def test_snapshot_store_footprint(physics, ncomp, nt, sub):
    shape = (30, 34, 70)  # ny, nx, nz
    ny, nx, nz = shape
    slots = 1 + nt // sub
    cells = ny * nx * padded(nz)
    al = lambda b: (b + 255) // 256 * 256
    for fmt, esz in (("f32", 4), ("bf16", 2), ("f16", 2)):
        want = al(slots * ncomp * cells * esz) + al(slots * 4)
        assert F.Batch.snap_footprint(physics, shape, nt, sub, fmt) == want


def test_footprints_reject_invalid_arguments():
    assert F.Batch.snap_footprint("elastic", (11, 30, 30), 100, 4) == -1   # an axis below 12 points
    assert F.Batch.snap_footprint("elastic", (30, 30, 30), 100, 0) == -1   # no imaging interval
    assert F.Batch.snap_footprint("elastic", (30, 30, 30), 1, 4) == -1     # nt < 2
    assert F.Batch.snap_footprint(7, (30, 30, 30), 100, 4) == -1           # unknown physics
    assert F.Batch.snap_footprint("elastic", (2000, 2000, 2000), 100, 4) == -1  # 32-bit field offsets


def test_batch_shot_footprint_covers_fields_model_and_fluxes():
    shape = (24, 26, 40)
    ny, nx, nz = shape
    cells = ny * nx * padded(nz)
    inj = F.Injector(1, [[0.1, 0.1, 0.1]], shape, (0.01, 0.01, 0.01))
    ext = F.Injector(0, [[0.05, 0.06, 0.02], [0.15, 0.1, 0.02]], shape, (0.01, 0.01, 0.01))
    # three wavefield sets, model fields, flux buffer (elastic 7, VTI 9) at the padded size
    assert F.Batch.footprint("acoustic", shape, inj, ext) >= (3 * 1 + 2) * cells * 4
    assert F.Batch.footprint("elastic", shape, inj, ext) >= (3 * 3 + 3 + 7) * cells * 4
    assert F.Batch.footprint("vti", shape, inj, ext) >= (3 * 3 + 7 + 9) * cells * 4
    assert F.Batch.footprint("vti", shape, inj, ext) > F.Batch.footprint("elastic", shape, inj, ext)


def test_serve_mix_is_seeded_and_follows_the_stream_spec():
    a, b = W.serve_mix(500, seed=42), W.serve_mix(500, seed=42)
    assert a == b and W.serve_mix(500, seed=43) != a
    kinds = [(r.kind, r.physics) for r in a]
    n_ac, n_el, n_vti = (kinds.count(("model", "acoustic")), kinds.count(("gradient", "elastic")),
                         kinds.count(("model", "vti")))
    assert n_ac + n_el + n_vti == 500
    # 60 / 30 / 10 % within binomial noise (3 sigma)
    for n, p in ((n_ac, 0.6), (n_el, 0.3), (n_vti, 0.1)):
        assert abs(n - 500 * p) <= 3 * np.sqrt(500 * p * (1 - p))
    for r in a:
        assert 1000 <= r.nt <= 3000
        if r.physics == "acoustic":
            assert 100 <= r.n <= 200 and r.dt == 0.001
        elif r.physics == "elastic":
            assert 128 <= r.n <= 256 and r.dt == 0.0008
        else:
            assert r.n == 160 and r.dt == 0.0008
    assert W.cell_updates(W.Request(0, "gradient", "elastic", 10, 11, 0.0008, 1)) == 2 * 1000 * 10


def test_poisson_arrivals():
    t = W.poisson_arrivals(20000, rate=2.5, seed=42)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 0.4) < 0.02          # mean inter-arrival 1/rate
    assert abs(gaps.std() - 0.4) < 0.03           # exponential: std = mean
    assert np.array_equal(t, W.poisson_arrivals(20000, rate=2.5, seed=42))
