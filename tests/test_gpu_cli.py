"""The drop-in executables on the GPU: WE_MODELING_3D.x and FWI3D.x run from a parfile
and a srcoord file on hypercube files, as the reference's executables do
(src/main/we_modeling.cpp:15-119, src/main/fwi3d.cpp:16-417), and their outputs are
checked against the oracle's apply_forward (src/cpp/we_op.cpp:872-946) and FWI objective
(src/cpp/optimization.cpp:5-197, src/include/optimization.hpp:410-424)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fwi3d_amd", "bin")
H = 0.01


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def write_cube(path, data, axes):
    """Binary + header (format=1), the reference's binWrite layout."""
    data = np.ascontiguousarray(data, np.float32)
    with open(path, "w") as f:
        for i, (n, o, d) in enumerate(axes):
            f.write(f"n{i + 1}={n}\no{i + 1}={o}\nd{i + 1}={d}\n")
        f.write(f"esize=4\nin={path}@\n")
    data.tofile(str(path) + "@")


def read_cube(path):
    kv = {}
    for line in open(path):
        for w in line.split():
            if "=" in w:
                k, v = w.split("=", 1)
                kv[k] = v
    n = [int(kv.get(f"n{i}", 1)) for i in range(1, 7)]
    data = np.fromfile(kv["in"].strip('"'), dtype="<f4")
    assert data.size == int(np.prod(n))
    return data, n, kv


def run(exe, args, cwd):
    cmd = [os.path.join(BIN, exe)] + args
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


def model_raw(physics, shape, seed=2):
    ny, nx, nz = shape
    z = np.arange(nz) * H
    r = np.random.default_rng(seed)
    vp = (np.broadcast_to(2.0 + 5.0 * z, shape) * (1 + 0.02 * r.standard_normal(shape).clip(-2, 2))).astype(np.float32)
    rho = (0.31 * (1000 * vp.astype(np.float64)) ** 0.25).astype(np.float32)
    if physics == "acoustic":
        return np.stack([vp, rho])
    parts = [vp, (vp / 1.8).astype(np.float32), rho]
    if physics == "vti":
        parts += [np.full(shape, 0.05, np.float32), np.full(shape, 0.1, np.float32), np.full(shape, 0.02, np.float32)]
    return np.stack(parts)


def ricker(nt, dt, f=25.0):
    t = np.arange(nt) * dt
    tau = t - 1.2 / f
    return ((1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)).astype(np.float32)


SHOTS = [  # shot id, source x y z, receivers (x y z dip az)
    (0, (0.10, 0.12, 0.06), [(0.05 + 0.03 * i, 0.10, 0.02, 0, 0) for i in range(5)]),
    (1, (0.16, 0.10, 0.05), [(0.06, 0.05 + 0.03 * i, 0.03, 0, 0) for i in range(4)]),
]


def write_srcoord(path):
    with open(path, "w") as f:
        f.write("# shot sx sy sz rx ry rz dip az\n")
        for sid, s, rs in SHOTS:
            for r in rs:
                f.write(f"{sid} {s[0]} {s[1]} {s[2]} {r[0]} {r[1]} {r[2]} {r[3]} {r[4]}\n")


@pytest.mark.parametrize("physics", ["elastic", "acoustic", "vti"])
def test_we_modeling_matches_oracle(tmp_path, physics):
    shape = (24, 26, 22)  # ny, nx, nz
    ny, nx, nz = shape
    raw = model_raw(physics, shape)
    nm = raw.shape[0]
    nt, dt = 120, 0.002
    w = ricker(nt, dt)
    write_cube(tmp_path / "model.H", raw, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (nm, 0, 1)])
    write_cube(tmp_path / "wavelet.H", w, [(nt, 0, dt)])
    write_srcoord(tmp_path / "geom.txt")
    mt = 1 if physics == "vti" else 0
    (tmp_path / "run.p").write_text(
        "# modelling parameters\n"
        f"srcoord={tmp_path}/geom.txt\nbc_top=1\nbc_bottom=2\nbc_left=2\nbc_right=2\nbc_front=2\nbc_back=2\n"
        "taper_bottom=4\ntaper_left=4\ntaper_right=4\ntaper_front=4\ntaper_back=4\ncourant=0.4\n"
        f"mt={mt}\nmxx=1.0\nmyy=0.8\nmzz=1.2\nmxy=0.1\nmxz=0.0\nmyz=-0.1\nfangle=0.3\nfazimuth=0.4\nformat=1\nverbose=1\n")
    run("WE_MODELING_3D.x", [f"parfile={tmp_path}/run.p", "source=wavelet.H", "model=model.H", "output=data.H",
                             f"datapath={tmp_path}"], tmp_path)
    got, n, _ = read_cube(tmp_path / "data.H")
    nrc = 1 if physics == "acoustic" else 3
    ntr = sum(len(s[2]) for s in SHOTS) * nrc
    assert n[:2] == [nt, ntr]
    got = got.reshape(ntr, nt)
    row = 0
    for sid, s, rs in SHOTS:
        if physics == "acoustic":
            src = w[None]
        elif mt:
            src = np.stack([w * m for m in (1.0, 0.8, 1.2, 0.1, 0.0, -0.1)])
        else:
            src = np.stack([w * np.float32(np.cos(0.3) * np.cos(0.4)), w * np.float32(np.cos(0.3) * np.sin(0.4)),
                            w * np.float32(np.sin(0.3))])
        g = O.ShotGeom(shape, (H,) * 3, (0, 0, 0), nt, dt, s, [r[:3] for r in rs], s="f")
        kw = dict(nmodels=nm, mt=mt, bc=(1, 2, 2, 2, 2, 2), taper=(0, 4, 4, 4, 4, 4), courant=0.4, dt=-1, sub=0)
        want, _ = O.apply_forward(O.op_par(**kw), g, raw, src.astype(np.float32))
        k = len(rs) * nrc
        e = rel(got[row:row + k], want.reshape(k, nt))
        print(f"{physics} shot {sid}: traces vs oracle {e:.2e}", flush=True)
        assert e <= 1e-4
        row += k


def test_we_modeling_wavefield(tmp_path):
    """One shot with wavefield= and sub=2: the snapshot store as a 5-axis hypercube."""
    shape = (20, 22, 24)
    ny, nx, nz = shape
    raw = model_raw("elastic", shape)
    nt, dt = 80, 0.002
    w = ricker(nt, dt)
    write_cube(tmp_path / "model.H", raw, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (3, 0, 1)])
    write_cube(tmp_path / "wavelet.H", w, [(nt, 0, dt)])
    args = ["source=wavelet.H", "model=model.H", "output=data.H", "wavefield=wfld.H", f"datapath={tmp_path}",
            "format=1", "ns=1", "sx0=0.1", "sy0=0.1", "sz0=0.1", "nr=3", "rx0=0.05", "rxinc=0.04", "ry0=0.1",
            "rz0=0.02", "sub=2", "dt=0.001", "bc_top=1"]
    run("WE_MODELING_3D.x", args, tmp_path)
    wf, n, kv = read_cube(tmp_path / "wfld.H")
    g = O.ShotGeom(shape, (H,) * 3, (0, 0, 0), nt, dt, (0.1, 0.1, 0.1), [(0.05 + 0.04 * i, 0.1, 0.02) for i in range(3)],
                   s="f")
    p = O.op_par(nmodels=3, dt=0.001, sub=2)
    src = np.stack([w, w * 0, w * 0]).astype(np.float32)  # fangle = fazimuth = 0: a force along x
    _, snaps = O.apply_forward(p, g, raw, src)
    assert n[:5] == [nz, nx, ny, 3, snaps.shape[0]]
    assert float(kv["d5"]) == pytest.approx(0.002)
    e = rel(wf, snaps.ravel())
    print(f"wavefield vs oracle snapshots {e:.2e}", flush=True)
    assert e <= 1e-4


def test_fwi3d_objective_and_descent(tmp_path):
    """FWI3D.x on two shots (niter=2): the first objective value equals the oracle's FWI
    objective (sum of shot misfits / dnorm) at the starting model, and it decreases."""
    shape = (22, 24, 20)
    ny, nx, nz = shape
    true = model_raw("elastic", shape, seed=4)
    start = true.copy()
    start[0] *= np.float32(0.98)
    start[1] *= np.float32(0.98)
    nt, dt = 100, 0.002
    w = ricker(nt, dt)
    write_cube(tmp_path / "wavelet.H", w, [(nt, 0, dt)])
    write_cube(tmp_path / "true.H", true, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (3, 0, 1)])
    write_cube(tmp_path / "start.H", start, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (3, 0, 1)])
    write_srcoord(tmp_path / "geom.txt")
    common = [f"srcoord={tmp_path}/geom.txt", "format=1", f"datapath={tmp_path}", "bc_top=1", "bc_bottom=2",
              "bc_left=2", "bc_right=2", "bc_front=2", "bc_back=2", "taper_bottom=3", "taper_left=3",
              "taper_right=3", "taper_front=3", "taper_back=3", "courant=0.4", "source=wavelet.H"]
    run("WE_MODELING_3D.x", common + ["model=true.H", "output=obs.H"], tmp_path)
    run("FWI3D.x", common + ["model=start.H", "data=obs.H", "output=inv.H", "obj_func=f.H", "niter=2",
                             "normalize_obj_func=1", "verbose=2"], tmp_path)
    f, n, _ = read_cube(tmp_path / "f.H")
    obs, _, _ = read_cube(tmp_path / "obs.H")
    obs = obs.reshape(-1, nt)
    # oracle objective at the start model
    kw = dict(nmodels=3, bc=(1, 2, 2, 2, 2, 2), taper=(0, 3, 3, 3, 3, 3), courant=0.4, dt=-1, sub=-1)
    src = np.stack([w * np.float32(np.cos(0.0)), w * 0, w * 0]).astype(np.float32)
    total, row = 0.0, 0
    for sid, s, rs in SHOTS:
        g = O.ShotGeom(shape, (H,) * 3, (0, 0, 0), nt, dt, s, [r[:3] for r in rs], s="f")
        k = 3 * len(rs)
        m, _, _ = O.fwi_shot(O.op_par(**kw), g, start, src, obs[row:row + k].reshape(3, len(rs), nt))
        total += m
        row += k
    dnorm = float(np.sum(obs.astype(np.float64) ** 2)) * dt
    want = total / dnorm
    print(f"objective history {f}, oracle start objective {want:.6e}", flush=True)
    assert abs(f[0] - want) / want < 1e-4
    assert len(f) >= 2 and f[-1] < f[0]
    inv, n, _ = read_cube(tmp_path / "inv.H")
    assert n[:4] == [nz, nx, ny, 3] and np.isfinite(inv).all()
    # with weights, trace normalisation and a gradient mask the driver still descends
    wts = np.ones_like(obs)
    wts[:, : nt // 4] = 0.5
    write_cube(tmp_path / "w.H", wts, [(nt, 0, dt), (obs.shape[0], 0, 1)])
    mask = np.ones_like(start)
    mask[:, :, :, :3] = 0
    write_cube(tmp_path / "mask.H", mask, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (3, 0, 1)])
    run("FWI3D.x", common + ["model=start.H", "data=obs.H", "output=inv2.H", "obj_func=f2.H", "niter=2",
                             "weights=w.H", "normalize=1", "mask=mask.H"], tmp_path)
    f2, _, _ = read_cube(tmp_path / "f2.H")
    inv2, _, _ = read_cube(tmp_path / "inv2.H")
    inv2 = inv2.reshape(3, ny, nx, nz)
    assert f2[-1] < f2[0]
    assert np.array_equal(inv2[:, :, :, :3], start[:, :, :, :3])  # masked cells keep their values


def test_we_modeling_das(tmp_path):
    """DAS channels (gl > 0): the srcoord dip / azimuth columns orient the fibers; one
    strain trace per channel, against the oracle's apply_forward."""
    shape = (24, 26, 22)
    ny, nx, nz = shape
    raw = model_raw("elastic", shape)
    nt, dt = 100, 0.002
    w = ricker(nt, dt)
    write_cube(tmp_path / "model.H", raw, [(nz, 0, H), (nx, 0, H), (ny, 0, H), (3, 0, 1)])
    write_cube(tmp_path / "wavelet.H", w, [(nt, 0, dt)])
    shots = [(0, (0.12, 0.12, 0.08), [(0.06 + 0.03 * i, 0.10, 0.05, 0.3 * i - 0.4, 0.5 * i) for i in range(5)]),
             (1, (0.15, 0.10, 0.09), [(0.08, 0.06 + 0.03 * i, 0.06, 0.2, 1.0 - 0.3 * i) for i in range(4)])]
    with open(tmp_path / "geom.txt", "w") as f:
        for sid, s, rs in shots:
            for r in rs:
                f.write(f"{sid} {s[0]} {s[1]} {s[2]} {r[0]} {r[1]} {r[2]} {r[3]} {r[4]}\n")
    args = [f"srcoord={tmp_path}/geom.txt", "gl=0.02", "format=1", f"datapath={tmp_path}", "courant=0.4",
            "bc_top=1", "bc_bottom=2", "bc_left=2", "bc_right=2", "bc_front=2", "bc_back=2",
            "source=wavelet.H", "model=model.H", "output=das.H"]
    run("WE_MODELING_3D.x", args, tmp_path)
    got, n, _ = read_cube(tmp_path / "das.H")
    assert n[:2] == [nt, 9]
    got = got.reshape(9, nt)
    row = 0
    src = np.stack([w, w * 0, w * 0]).astype(np.float32)
    for sid, s, rs in shots:
        g = O.ShotGeom(shape, (H,) * 3, (0, 0, 0), nt, dt, s, [r[:3] for r in rs], s="f", rdipaz=[r[3:] for r in rs])
        kw = dict(nmodels=3, gl=0.02, bc=(1, 2, 2, 2, 2, 2), courant=0.4, dt=-1, sub=0)
        want, _ = O.apply_forward(O.op_par(**kw), g, raw, src)
        e = rel(got[row:row + len(rs)], want.reshape(len(rs), nt))
        print(f"DAS shot {sid}: strain traces vs oracle {e:.2e}", flush=True)
        assert e <= 1e-4
        row += len(rs)
