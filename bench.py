#!/usr/bin/env python3
"""Benchmark of the MI355X wave-propagation / adjoint-gradient engine.

Default workload (the configuration the headline metric is quoted on): elastic
isotropic forward + adjoint gradient, 1 shot, 200^3 cells, h = 10 m, nt = 2000,
dt = 0.8 ms, FP32 snapshots every 4 steps, 5-layer model with a +5% Gaussian vp
anomaly, explosive 12 Hz Ricker at z = 30 m, 40x40 3-C receivers at z = 20 m,
residual against data modelled without the anomaly.

One step = one shot gradient: forward propagation with snapshots, residual, adjoint
propagation with the imaging condition (all inputs resident in HBM before timing).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 200] [--nt 2000]

Prints one JSON line (see the keys at the end of main()).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & HBM GB/s vs roofline; shot-gradients/s, p50/p99 latency"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
ELASTIC_BYTES_PER_CELL = 48.0  # algorithmic bytes per elastic cell-update (u, u_prev, u_next x3 + lambda, mu, rho)
ACOUSTIC_BYTES_PER_CELL = 20.0  # u, u_prev, u_next + K, 1/rho


def cpu_model():
    """The host CPU's model name (lscpu / /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(shot, seconds=6.0, gradient=True):
    """Time the oracle (FP32 CPU restatement of the reference's propagate and
    compute_gradients, OpenMP) on a bounded sample of the same workload: a forward run of
    nsteps time steps with snapshots every `sub` steps, then the adjoint run with the
    imaging condition (forward + adjoint cell-updates, the GPU figure's unit); with
    gradient=False the forward run only (modelling workloads). One warm-up, then the
    median of 3."""
    from oracle import oracle as O
    shape = shot.shape
    inj = O.Injector(shot.src_kind, shot.src_loc, shape, shot.d, (0, 0, 0))
    ext = O.Injector(0, shot.rcv_loc, shape, shot.d, (0, 0, 0))

    def run(nsteps):
        nt = nsteps + 1
        par = O.make_par(nt, shot.sub, shot.dt, bc=shot.bc, taper=shot.taper, taper_strength=shot.taper_strength)
        src = np.ascontiguousarray(shot.src[..., :nt])
        t = time.perf_counter()
        rcv, snaps = O.propagate(shot.physics, shot.model, src, inj, ext, par, shot.d)
        if gradient:
            grad = np.zeros((3,) + tuple(shape), np.float32)
            O.propagate(shot.physics, shot.model, np.ascontiguousarray(rcv[:, :, ::-1]), ext, inj, par, shot.d,
                        adj=True, grad=grad, snaps=snaps)
        return time.perf_counter() - t

    t4 = run(4)  # warm-up, and the per-step cost
    nsteps = int(max(4, min(400, seconds / max(t4 / 4, 1e-3))))
    if shot.sub > 0:
        nsteps -= nsteps % shot.sub
    times = sorted(run(nsteps) for _ in range(3))
    t = times[1]
    cells = float(np.prod(shape))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = (f"{nsteps} forward time steps with snapshots every {shot.sub} steps, then {nsteps} adjoint steps with the "
            f"imaging condition" if gradient else f"{nsteps} forward time steps")
    return {"value": round((2.0 if gradient else 1.0) * cells * nsteps / t / 1e9, 6), "unit": "Gcell-updates/s",
            "cores": cores, "kind": "port", "cpu": cpu_model(), "runs_s": [round(x, 2) for x in times],
            "sample": f"{what}, of the same {shape[2]}x{shape[1]}x{shape[0]} {shot.physics} shot "
                      f"(oracle FP32 C++/OpenMP); median of 3 runs after one warm-up"}


# fwi3d_tune settings of every engine the workloads create (--unfused, --overlap)
ENGINE_TUNE = {}

# profiled kernel classes of the elastic time step (fwi3d_profile_read), by step form
KERNEL_CLASSES = {
    True: ((0, "elastic_fused"), (1, "elastic_flux_yfaces"), (3, "elastic_yslabs")),
    False: ((0, "elastic_interior"), (1, "elastic_flux"), (2, "elastic_boundary"), (3, "elastic_zslab")),
}
STEP_KERNEL = {
    True: "elastic time step = elastic_fused_batch (middle and edge tiles, planes [6, ny-6)) + elastic_flux_batch "
          "(the 8 planes at each y face) + elastic_yslabs_ext (y-face slabs, edges, receivers) launches, the edge "
          "tiles on a second stream beside the middle tiles and the y-face launches after the middle tiles (HIP "
          "events around each step on the engine stream; the per-kernel averages come from a run with one stream)",
    False: "elastic time step = elastic_flux + elastic_interior + elastic_boundary + elastic_zslab launches "
           "(HIP events around each step on the engine stream)",
}


def new_engine(F):
    eng = F.Engine(0)
    if ENGINE_TUNE:
        eng.tune(**ENGINE_TUNE)
    return eng


def traffic_profile(n):
    """The committed rocprofv3 --pmc summary of the elastic time step at n^3 for the
    step form in use (fused by default, --unfused)."""
    form = "unfused" if ENGINE_TUNE.get("fused") is False else "fused"
    return os.path.join("profiles", f"elastic_{form}_step_{n}cube_traffic_pmc.md")


def kernel_rooflines(fused, nz, nx, ny, avg_us, n):
    """Per kernel class of the elastic step: the cells it handles, its algorithmic bytes
    (48 B per cell it updates: u, prev, lambda, mu, rho read, next written; the y-face
    flux pass reads u, lambda, mu and writes T and 6 fluxes, 48 B per cell too), its
    one-stream launch time, achieved GB/s and fraction of 8 TB/s, and its measured HBM
    bytes per launch from the committed PMC summary when there is one."""
    nxz = nx * nz
    if fused:
        cells = {"elastic_fused": nxz * max(0, ny - 12), "elastic_flux_yfaces": nxz * min(ny, 16),
                 "elastic_yslabs": nxz * min(ny, 12)}
        prof_keys = {"elastic_fused": ("elastic_fused (middle tiles)", "elastic_fused (edge tiles)"),
                     "elastic_flux_yfaces": ("elastic_flux_batch (y faces)",), "elastic_yslabs": ("elastic_yslabs_ext",)}
    else:
        inner = max(0, nx - 12) * max(0, ny - 12) * max(0, nz - 12)
        cells = {"elastic_interior": inner, "elastic_flux": nxz * ny, "elastic_boundary": 0,
                 "elastic_zslab": nxz * ny - inner}
        prof_keys = {k: (k,) for k in cells}
    path = os.path.join(ROOT, traffic_profile(n))
    per_launch = {}
    if os.path.exists(path):
        try:
            block = open(path).read().split("```json", 1)[1].split("```", 1)[0]
            per_launch = json.loads(block)["traffic_bytes_per_launch"]
        except (IndexError, KeyError, ValueError):
            per_launch = {}
    out = {}
    for k, c in cells.items():
        us = avg_us.get(k, 0.0)
        if c <= 0 or us <= 0:
            continue
        b = ELASTIC_BYTES_PER_CELL * c
        gbs = b / (us * 1e-6) / 1e9
        t = [per_launch[q] for q in prof_keys[k] if q in per_launch]
        out[k] = {"cells": int(c), "algorithmic_bytes": int(b), "avg_us": round(us, 2), "achieved_GBs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic_bytes": int(sum(t)) if t else None}
    return out


def measured_traffic(n):
    """HBM bytes per elastic time step (scaled FETCH_SIZE + WRITE_SIZE of the step kernels)
    from the committed rocprofv3 --pmc summary of the same workload, or None."""
    path = os.path.join(ROOT, traffic_profile(n))
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)["traffic_bytes_per_step"])
    except (IndexError, KeyError, ValueError):
        return None


def copy_probe(torch, mib):
    """A plain elementwise device copy of `mib` MiB, 5 launches: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_launch_read": n * 4, "copy_probe_bytes_per_launch_write": n * 4}), flush=True)


def run_lbfgs(args, torch, F, W):
    """Mixed-precision elastic FWI: 256^3 (or --n), nt=3000 (or --nt), layered model with
    the anomaly as the target, anomaly-free starting model, snapshots every 4 steps in
    --snap precision, --iters l-BFGS iterations (m=5, regular Wolfe) with FP64 dot
    products; the gradient at the starting model is compared with the FP32-snapshot one."""
    n, nt = args.n if args.n != 200 else 256, args.nt if args.nt != 2000 else 3000
    true = W.config_b(n=n, nt=nt)
    start = W.config_b(n=n, nt=nt, anomaly=False)
    shape = true.shape
    dev = torch.device("cuda:0")
    eng = new_engine(F)
    geo = F.Shot(shape, true.d, (0.0, 0.0, 0.0), nt, true.dt, true.src_loc[0], true.rcv_loc)
    opkw = dict(nmodels=3, dt=true.dt, sub=true.sub, mt=1, bc=true.bc, taper=true.taper,
                taper_strength=true.taper_strength)
    src = torch.from_numpy(np.ascontiguousarray(true.src[:, 0, :])).to(dev)
    m_true = torch.from_numpy(np.ascontiguousarray(true.raw)).to(dev)
    dobs = torch.zeros((3, geo.nrcv, nt), dtype=torch.float32, device=dev)
    F.apply_forward_device(eng, F.op_param(**dict(opkw, sub=0)), geo, False, m_true.data_ptr(), src.data_ptr(),
                           dobs.data_ptr())
    del m_true
    m = torch.from_numpy(np.ascontiguousarray(start.raw)).to(dev)
    g = torch.zeros_like(m)
    g64 = torch.zeros(m.shape, dtype=torch.float64, device=dev)

    def gradient(fmt):
        g64.zero_()
        f = F.fwi_shot_device(eng, F.op_param(**dict(opkw, snap_precision=fmt)), geo, m.data_ptr(), src.data_ptr(),
                              dobs.data_ptr(), g64.data_ptr())
        return f, g64.clone()

    t0 = time.perf_counter()
    _, g_ref = gradient("f32")
    _, g_mp = gradient(args.snap)
    t_two = time.perf_counter() - t0
    err = [float(torch.linalg.vector_norm(g_mp[k] - g_ref[k]) / torch.linalg.vector_norm(g_ref[k])) for k in range(3)]
    del g_ref, g_mp
    evals = [0]

    def objective(mp, gp):
        g64.zero_()
        f = F.fwi_shot_device(eng, F.op_param(**dict(opkw, snap_precision=args.snap)), geo, mp, src.data_ptr(),
                              dobs.data_ptr(), g64.data_ptr())
        g.copy_(g64.float())
        torch.cuda.synchronize()
        evals[0] += 1
        print(f"# evaluation {evals[0]}: misfit {f:.6e}", file=sys.stderr, flush=True)
        return f

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist, (it, fe, ge) = F.lbfgs_run(eng, m.numel(), m.data_ptr(), g.data_ptr(), objective, niter=args.iters)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    N = float(np.prod(shape))
    out = {
        "metric": METRIC,
        "value": round(2.0 * N * (nt - 1) * fe / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": fe,
        "warmup": 0,
        "ms_per_step": round(elapsed / max(fe, 1) * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision elastic FWI, l-BFGS (m=5, regular Wolfe, FP64 dots), 1 shot, "
                               f"{shape[2]}x{shape[1]}x{shape[0]}, nt={nt}, dt={true.dt}, sub={true.sub}, "
                               f"{args.snap.upper()} snapshots with per-snapshot max-abs scale",
                   "iterations": it, "function_evaluations": fe, "gradient_evaluations": ge},
        "shot_gradients_per_s": round(fe / elapsed, 4),
        "misfit_history": hist,
        "misfit_decrease": round(1.0 - hist[-1] / hist[0], 6) if hist else None,
        "snapshot_gradient_rel_l2_vs_fp32": {"vp": err[0], "vs": err[1], "rho": err[2]},
        "two_gradients_s": round(t_two, 2),
    }
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
    eng.close()


def run_batch(args, torch, F, W, emit=True):
    """Batched VTI forward modelling: 32 shots split 11/11/10 over the grids (nz,nx,ny) =
    (128,128,128), (200,160,160), (96,192,192), nt=1500 (or --nt), dt=0.8 ms. The same
    shots run one at a time through the plain path (fwi3d_propagate_device) and then
    together through the batch (fwi3d_batch_*: every shot admitted at step 0, one launch
    per kernel class and step); traces of both are compared bit for bit."""
    nt = args.nt if args.nt != 2000 else 1500
    split = (11, 11, 10)
    shots = [W.config_c_shot(g, k % 12, nt=nt) for g in range(3) for k in range(split[g])]
    dev = torch.device("cuda:0")
    eng = new_engine(F)
    bufs = []
    for s in shots:
        inj = F.Injector(1, s.src_loc, s.shape, s.d)
        ext = F.Injector(0, s.rcv_loc, s.shape, s.d)
        par = F.make_param(nt, s.dt, sub=0, bc=s.bc, taper=s.taper, taper_strength=s.taper_strength)
        m = torch.from_numpy(np.ascontiguousarray(s.model)).to(dev)
        src = torch.from_numpy(np.ascontiguousarray(s.src)).to(dev)
        r1 = torch.zeros((3, ext.npts, nt), dtype=torch.float32, device=dev)
        r2 = torch.zeros_like(r1)
        bufs.append((s, inj, ext, par, m, src, r1, r2))
    torch.cuda.synchronize()
    cells = sum(float(np.prod(s.shape)) * (nt - 1) for s in shots)
    # tile shapes per grid from the autotuner (used by both paths below)
    tiles = {}
    for g in range(3):
        shp = shots[sum(split[:g])].shape
        tiles[f"{shp[2]}x{shp[1]}x{shp[0]}"] = eng.autotune("vti", shp, steps=10)
    print(f"# autotuned (tx, ychunk): {tiles}", file=sys.stderr, flush=True)

    # plain path, one shot after another (first shot once untimed, to warm up)
    s, inj, ext, par, m, src, r1, _ = bufs[0]
    eng.propagate_device("vti", m.data_ptr(), src.data_ptr(), r1.data_ptr(), inj, ext, par, s.shape, s.d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat_plain = []
    for s, inj, ext, par, m, src, r1, _ in bufs:
        ts = time.perf_counter()
        eng.propagate_device("vti", m.data_ptr(), src.data_ptr(), r1.data_ptr(), inj, ext, par, s.shape, s.d)
        lat_plain.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    print(f"# plain: {len(shots)} shots in {t_plain:.2f} s", file=sys.stderr, flush=True)

    # batched: all shots admitted at step 0
    b = F.Batch(eng)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, inj, ext, par, m, src, _, r2 in bufs:
        b.add("vti", m.data_ptr(), src.data_ptr(), r2.data_ptr(), inj, ext, par, s.shape, s.d)
    resident = b.status()[1]
    torch.cuda.synchronize()
    t_admit = time.perf_counter() - t0
    b.autotune(window=10)  # online: the first 40 steps time y chunks 4/8/16/32
    b.run()
    batch_tiles = b.tiles()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    b.close()
    print(f"# batched: {len(shots)} shots in {t_batch:.2f} s", file=sys.stderr, flush=True)
    same = all(torch.equal(r1, r2) for *_, r1, r2 in bufs)
    out = {
        "metric": METRIC,
        "value": round(cells / t_batch / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(t_batch * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched VTI forward modelling, 32 shots over (nz,nx,ny) = (128,128,128) x11, "
                               f"(200,160,160) x11, (96,192,192) x10, nt={nt}, dt=0.0008, one launch per kernel "
                               f"class and time step, replayed from a captured graph of 16 steps between "
                               f"departures (fwi3d_batch)", "shots": len(shots)},
        "plain_gcell_s": round(cells / t_plain / 1e9, 3),
        "shots_per_s": {"batched": round(len(shots) / t_batch, 4), "plain": round(len(shots) / t_plain, 4)},
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "plain_latency_s": {"p50": round(float(np.percentile(lat_plain, 50)), 4),
                            "p99": round(float(np.percentile(lat_plain, 99)), 4)},
        "batch_resident_gib": round(resident / 2**30, 2),
        "batch_phase_s": {"admit": round(t_admit, 3), "steps": round(t_batch - t_admit, 3)},
        "autotuned_tiles_tx_ychunk": tiles,
        "batch_online_ychunk": batch_tiles[0],
        "batch_candidate_ms_per_step": {str(y): round(t, 3) for y, t in zip((4, 8, 16, 32), batch_tiles[1])},
        "traces_identical_to_plain": bool(same),
    }
    eng.close()
    if not emit:
        return out
    print(json.dumps(out), flush=True)


def run_acoustic(args, torch, F, W):
    """Acoustic variable-density forward modelling (the acoustic configuration): 100^3
    (or --n), nt=1000 (or --nt), dt=1 ms, vp 1.5 -> 2.5 km/s with depth, Gardner rho,
    15 Hz Ricker pressure source at z=50 m, 10x10 receivers at z=20 m. One step = one
    apply_forward (raw model and wavelet in HBM -> 100 traces at the data rate)."""
    n = args.n if args.n != 200 else 100
    nt = args.nt if args.nt != 2000 else 1000
    shot = W.config_a(n=n, nt=nt)
    shape = shot.shape
    N = float(np.prod(shape))
    dev = torch.device("cuda:0")
    eng = new_engine(F)
    tiles = eng.autotune("acoustic", shape, steps=20)
    geo = F.Shot(shape, shot.d, (0.0, 0.0, 0.0), nt, shot.dt, shot.src_loc[0], shot.rcv_loc)
    par = lambda: F.op_param(nmodels=2, dt=shot.dt, sub=0, bc=shot.bc, taper=shot.taper,
                             taper_strength=shot.taper_strength)
    model = torch.from_numpy(np.ascontiguousarray(shot.raw)).to(dev)
    src = torch.from_numpy(np.ascontiguousarray(shot.src[:, 0, :])).to(dev)
    data = torch.zeros((1, geo.nrcv, nt), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    # per-kernel launch times: one untimed shot launched kernel by kernel with HIP events
    eng.tune(graphs=False)
    eng.profile(True)
    F.apply_forward_device(eng, par(), geo, False, model.data_ptr(), src.data_ptr(), data.data_ptr())
    kern = {name: eng.profile_read(k) for k, name in ((0, "acoustic_interior"), (2, "acoustic_boundary"),
                                                      (4, "acoustic_time_step_eager"))}
    eng.profile(False)
    eng.tune(graphs=True)
    for _ in range(args.warmup):
        F.apply_forward_device(eng, par(), geo, False, model.data_ptr(), src.data_ptr(), data.data_ptr())
    torch.cuda.synchronize()
    # timed region: the time loop replays captured graphs of 16 steps; HIP events around
    # each replay
    eng.profile(True, steps_only=True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        F.apply_forward_device(eng, par(), geo, False, model.data_ptr(), src.data_ptr(), data.data_ptr())
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    elapsed = time.perf_counter() - t0
    prof = dict(kern, acoustic_time_step=eng.profile_read(4))
    eng.profile(False)
    avg_us = {k: ms / max(c, 1) * 1e3 for k, (ms, c) in prof.items()}
    t_step = avg_us.pop("acoustic_time_step") / 1e6
    achieved = ACOUSTIC_BYTES_PER_CELL * N / t_step / 1e9 if t_step > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(N * (nt - 1) * args.steps / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"acoustic variable-density forward modelling, 1 shot, {shape[2]}x{shape[1]}x{shape[0]}, "
                               f"nt={nt}, dt={shot.dt}, 100 pressure receivers (fwi3d_apply_forward_device)",
                   "autotuned_tiles_tx_ychunk": list(tiles)},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
        "trace_max_abs": float(data.abs().max()),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "acoustic time step = acoustic_interior + acoustic_boundary (+ extraction, injection) "
                               "replayed from a captured graph (HIP events around each 16-step replay)",
                     "step_us": round(t_step * 1e6, 2), "bytes_per_cell": ACOUSTIC_BYTES_PER_CELL,
                     "cells_per_step": N, "avg_us": {k: round(v, 2) for k, v in avg_us.items()},
                     "note": "a 100^3 step's 20 MB working set is served from the 256 MB Infinity Cache"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(shot, gradient=False)
    print(json.dumps(out), flush=True)
    eng.close()


def _smooth(torch, x, sigma):
    """Gaussian filter (sigma in cells, truncated at 4 sigma, edge values repeated) along
    the three axes of x, on the device."""
    r = int(4 * sigma + 0.5)
    k = torch.exp(-0.5 * (torch.arange(-r, r + 1, device=x.device, dtype=torch.float32) / sigma) ** 2)
    k = (k / k.sum()).view(1, 1, -1)
    for axis in range(3):
        y = x.movedim(axis, -1)
        shp = y.shape
        y = torch.nn.functional.pad(y.reshape(-1, 1, shp[-1]), (r, r), mode="replicate")
        x = torch.nn.functional.conv1d(y, k).reshape(shp).movedim(-1, axis).contiguous()
    return x


def serve_inputs(torch, F, W, req, dev):
    """Device inputs of one serving request, generated from its seed: the acoustic
    configuration's linear-gradient model, the elastic one's five layers with a +5%
    Gaussian anomaly (seeded centre within +-100 m), the VTI one's linear vp with smooth
    random epsilon/delta; Gardner density; sources and receivers as in those configs."""
    n, nt, dt, h = req.n, req.nt, req.dt, 0.01
    gen = torch.Generator(device=dev)
    gen.manual_seed(req.seed)
    z = torch.arange(n, device=dev, dtype=torch.float64)
    c = (n - 1) * h / 2
    gardner = lambda vp: (0.31 * (1000.0 * vp.double()) ** 0.25).float()
    if req.physics == "acoustic":
        vp = (1.5 + 1.0 * z / (n - 1)).float().expand(n, n, n).contiguous()
        raw = torch.stack([vp, gardner(vp)])
        w = W.ricker(15.0, nt, dt)[None, :]
        sxyz, rxyz, nm, mt, ns = (c, c, 0.05), W._surface_grid(c, c, 0.02, 10, 0.05), 2, 0, 1
    elif req.physics == "elastic":
        vals = torch.linspace(2.0, 3.5, 5, device=dev, dtype=torch.float64)
        bounds = [round(k * n / 5) for k in range(6)]
        vz = torch.empty(n, device=dev, dtype=torch.float64)
        for k in range(5):
            vz[bounds[k]:bounds[k + 1]] = vals[k]
        off = (torch.rand(3, generator=gen, device=dev, dtype=torch.float64) - 0.5) * 0.2
        x = torch.arange(n, device=dev, dtype=torch.float64) * h
        r2 = ((x[:, None, None] - c - off[0]) ** 2 + (x[None, :, None] - c - off[1]) ** 2 +
              (x[None, None, :] - c - off[2]) ** 2)
        vp = (vz[None, None, :] * (1 + 0.05 * torch.exp(-r2 / 0.3 ** 2))).float()
        raw = torch.stack([vp, (vp / 1.8).float(), gardner(vp)])
        w = W.ricker(12.0, nt, dt)
        nrec = min(40, int(((n - 1) * h - 0.04) / 0.05) + 1)
        sxyz, rxyz, nm, mt, ns = (c, c, 0.03), W._surface_grid(c, c, 0.02, nrec, 0.05), 3, 1, 6
        w = np.stack([w, w, w] + [0 * w] * 3)
    else:
        vp = (1.8 + 2.2 * z / (n - 1)).float().expand(n, n, n).contiguous()

        def field(scale):
            f = _smooth(torch, torch.randn((n, n, n), generator=gen, device=dev, dtype=torch.float32), 5.0)
            return (scale * (f - f.min()) / (f.max() - f.min()).clamp_min(1e-30)).float()

        eps, dlt = field(0.2), field(0.1)
        raw = torch.stack([vp, (vp / 2).float(), gardner(vp), dlt, eps, torch.zeros_like(vp)])
        w = W.ricker(15.0, nt, dt)
        sxyz, rxyz, nm, mt, ns = (c, c, 0.02), W._surface_grid(c, c, 0.02, 16, (n - 40) * h / 16), 6, 1, 6
        w = np.stack([w, w, w] + [0 * w] * 3)
    geo = F.Shot((n, n, n), (h, h, h), (0.0, 0.0, 0.0), nt, dt, sxyz, rxyz)
    grad = req.kind == "gradient"
    par = dict(nmodels=nm, dt=dt, sub=4 if grad else 0, mt=mt, bc=(1, 2, 2, 2, 2, 2), taper=(0, 20, 20, 20, 20, 20),
               snap_precision="bf16")
    nrc = 1 if nm == 2 else 3
    out = dict(req=req, geo=geo, par=par, model=raw.contiguous(),
               src=torch.from_numpy(np.ascontiguousarray(np.asarray(w, np.float32).reshape(ns, nt))).to(dev),
               data=torch.zeros((nrc, geo.nrcv, nt), dtype=torch.float32, device=dev))
    if grad:
        out["obs"] = torch.zeros_like(out["data"])  # residual = synthetic data
        out["grad"] = torch.zeros((min(3, nm), n, n, n), dtype=torch.float64, device=dev)
    return out


def run_serve(args, torch, F, W, emit=True):
    """Serving stream: --requests seeded requests (60% acoustic modelling n^3, n in
    100..200; 30% elastic shot gradients, n in 128..256, BF16 snapshots every 4 steps;
    10% VTI modelling 160^3; nt in 1000..3000). (1) plain: one request after another
    through the one-shot entry points; (2) served at saturation: every request queued at
    once into the serving loop (capacity); (3) served under a Poisson stream at 0.7 x that
    capacity: p50/p99 latency. Results of (2) are compared with (1) bit for bit."""
    dev = torch.device("cuda:0")
    reqs = W.serve_mix(args.requests)
    t0 = time.perf_counter()
    inputs = [serve_inputs(torch, F, W, r, dev) for r in reqs]
    torch.cuda.synchronize()
    print(f"# {len(reqs)} requests generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cells = sum(W.cell_updates(r) for r in reqs)
    ngrad = sum(r.kind == "gradient" for r in reqs)

    def plain_one(eng, q, data, grad):
        par = F.op_param(**q["par"])
        if q["req"].kind == "model":
            F.apply_forward_device(eng, par, q["geo"], False, q["model"].data_ptr(), q["src"].data_ptr(), data.data_ptr())
        else:
            F.fwi_shot_device(eng, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(), q["obs"].data_ptr(),
                              grad.data_ptr(), data.data_ptr())

    # (1) plain, one request at a time (a small warm-up request first, untimed)
    eng = new_engine(F)
    warm = serve_inputs(torch, F, W, W.Request(-1, "gradient", "elastic", 40, 50, 0.0008, 1), dev)
    plain_one(eng, warm, warm["data"], warm["grad"])
    plain_out, lat_plain = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_note = t0
    for k, q in enumerate(inputs):
        ts = time.perf_counter()
        d = torch.zeros_like(q["data"])
        g = torch.zeros_like(q["grad"]) if "grad" in q else None
        plain_one(eng, q, d, g)
        plain_out.append((d, g))
        lat_plain.append(time.perf_counter() - ts)
        if time.perf_counter() - t_note > 20:
            t_note = time.perf_counter()
            print(f"# plain: {k + 1}/{len(inputs)} requests, {t_note - t0:.0f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    eng.close()
    torch.cuda.empty_cache()
    print(f"# plain: {t_plain:.2f} s", file=sys.stderr, flush=True)

    free, _ = torch.cuda.mem_get_info()
    budget = int(min(args.hbm_budget_gib * 2**30, free - 16 * 2**30))

    def serve(arrivals):
        eng = new_engine(F)
        srv = F.Server(eng, budget, args.max_active)
        for q in inputs:
            if "grad" in q:
                q["grad"].zero_()
        torch.cuda.synchronize()
        ids = []
        t0 = time.perf_counter()
        t_note = [t0]

        def note(tag):
            if time.perf_counter() - t_note[0] > 20:
                t_note[0] = time.perf_counter()
                st = srv.stats()
                print(f"# {tag}: {st['completed']}/{len(inputs)} done, {st['active']} running, {st['queued']} queued, "
                      f"{t_note[0] - t0:.0f} s", file=sys.stderr, flush=True)

        for k, q in enumerate(inputs):
            if arrivals is not None:
                while True:  # sleep until the arrival, reporting progress meanwhile
                    wait = t0 + arrivals[k] - time.perf_counter()
                    if wait <= 0:
                        break
                    time.sleep(min(wait, 1.0))
                    note("serving")
            par = F.op_param(**q["par"])
            if q["req"].kind == "model":
                ids.append(srv.submit(F.REQ_MODEL, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(),
                                      data_ptr=q["data"].data_ptr()))
            else:
                ids.append(srv.submit(F.REQ_GRADIENT, par, q["geo"], q["model"].data_ptr(), q["src"].data_ptr(),
                                      data_ptr=q["data"].data_ptr(), obs_ptr=q["obs"].data_ptr(),
                                      grad64_ptr=q["grad"].data_ptr()))
        while True:
            st = srv.stats()
            if st["completed"] + st["failed"] == st["submitted"]:
                break
            time.sleep(0.05)
            note("serving")
        srv.wait()
        elapsed = time.perf_counter() - t0
        polls = [srv.poll(i) for i in ids]
        stats = srv.stats()
        errors = [srv.error(i) for i, p in zip(ids, polls) if p[0] != 2]
        srv.close()
        eng.close()
        if errors:
            raise RuntimeError(f"serving requests failed: {errors[:3]}")
        return elapsed, [p[2] for p in polls], [p[3] for p in polls], stats

    # (2) saturation: every request queued at t = 0
    t_srv, lat_sat, _, st_sat = serve(None)
    print(f"# served (saturation): {t_srv:.2f} s, counters {st_sat}", file=sys.stderr, flush=True)
    same = all(torch.equal(q["data"], d) and (g is None or torch.equal(q["grad"], g))
               for q, (d, g) in zip(inputs, plain_out))
    del plain_out
    # (3) Poisson stream at 0.7 x the measured capacity
    rate = 0.7 * len(reqs) / t_srv
    arr = W.poisson_arrivals(len(reqs), rate)
    t_poi, lat, qwait, st_poi = serve(arr)
    print(f"# served (Poisson, {rate:.3f} req/s): {t_poi:.2f} s", file=sys.stderr, flush=True)
    pct = lambda v, q: round(float(np.percentile(v, q)), 4)
    kinds = {}
    for r, l in zip(reqs, lat):
        kinds.setdefault(f"{r.kind}_{r.physics}", []).append(l)
    out = {
        "metric": METRIC,
        "value": round(cells / t_srv / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(t_srv * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (models generated per request from its seed; zero observed data, so the residual is the "
                "synthetic data)",
        "config": {"workload": f"serving stream: {len(reqs)} seeded requests (60% acoustic modelling n^3 n~U{{100..200}}, "
                               f"30% elastic shot gradients n~U{{128..256}} with BF16 snapshots every 4 steps, 10% VTI "
                               f"modelling 160^3; nt~U{{1000..3000}}), continuous batching (fwi3d_server)",
                   "requests": len(reqs), "gradients": ngrad, "hbm_budget_gib": round(budget / 2**30, 1),
                   "max_active": args.max_active},
        "shot_gradients_per_s": {"served": round(ngrad / t_srv, 4), "plain": round(ngrad / t_plain, 4)},
        "requests_per_s": {"served": round(len(reqs) / t_srv, 4), "plain": round(len(reqs) / t_plain, 4)},
        "speedup_vs_plain": round(t_plain / t_srv, 3),
        "plain_gcell_s": round(cells / t_plain / 1e9, 3),
        "poisson": {"rate_req_per_s": round(rate, 4), "makespan_s": round(t_poi, 2),
                    "latency_s": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                    "queue_s": {"p50": pct(qwait, 50), "p99": pct(qwait, 99)},
                    "latency_p50_by_kind_s": {k: pct(v, 50) for k, v in kinds.items()},
                    "hbm_peak_gib": round(st_poi["hbm_peak"] / 2**30, 2), "peak_active": st_poi["peak_active"]},
        "saturation": {"hbm_peak_gib": round(st_sat["hbm_peak"] / 2**30, 2), "peak_active": st_sat["peak_active"],
                       "latency_s": {"p50": pct(lat_sat, 50), "p99": pct(lat_sat, 99)}},
        "plain_latency_s": {"p50": pct(lat_plain, 50), "p99": pct(lat_plain, 99)},
        "results_identical_to_plain": bool(same),
    }
    if not emit:
        return out
    print(json.dumps(out), flush=True)


def config_d_gradients(torch, F, W, n=256, nt=3000):
    """The mixed-precision gradient configuration's shot gradient at full size (256^3,
    nt=3000, sub=4; layered model with the anomaly, anomaly-free observed data): one shot
    gradient with FP32 snapshots and one with BF16 snapshots, each timed, and the BF16
    gradient's relative L2 error against the FP32 one."""
    true = W.config_b(n=n, nt=nt)
    obs_model = W.config_b(n=n, nt=nt, anomaly=False)
    shape = true.shape
    dev = torch.device("cuda:0")
    eng = new_engine(F)
    tiles = eng.autotune("elastic", shape, steps=10)  # as the headline workload does
    geo = F.Shot(shape, true.d, (0.0, 0.0, 0.0), nt, true.dt, true.src_loc[0], true.rcv_loc)
    opkw = dict(nmodels=3, dt=true.dt, sub=true.sub, mt=1, bc=true.bc, taper=true.taper,
                taper_strength=true.taper_strength)
    src = torch.from_numpy(np.ascontiguousarray(true.src[:, 0, :])).to(dev)
    m = torch.from_numpy(np.ascontiguousarray(obs_model.raw)).to(dev)
    dobs = torch.zeros((3, geo.nrcv, nt), dtype=torch.float32, device=dev)
    F.apply_forward_device(eng, F.op_param(**dict(opkw, sub=0)), geo, False, m.data_ptr(), src.data_ptr(),
                           dobs.data_ptr())
    m = torch.from_numpy(np.ascontiguousarray(true.raw)).to(dev)
    g64 = torch.zeros(m.shape, dtype=torch.float64, device=dev)
    N = float(np.prod(shape))
    res, grads = {}, {}
    for fmt in ("f32", "bf16"):
        g64.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = F.fwi_shot_device(eng, F.op_param(**dict(opkw, snap_precision=fmt)), geo, m.data_ptr(), src.data_ptr(),
                              dobs.data_ptr(), g64.data_ptr())
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res[fmt] = {"shot_gradient_s": round(t, 3), "gcell_updates_per_s": round(2 * N * (nt - 1) / t / 1e9, 3),
                    "misfit": f}
        grads[fmt] = g64.clone()
        print(f"# config D {fmt}: {t:.2f} s", file=sys.stderr, flush=True)
    err = [float(torch.linalg.vector_norm(grads["bf16"][k] - grads["f32"][k]) / torch.linalg.vector_norm(grads["f32"][k]))
           for k in range(3)]
    eng.close()
    return {"workload": f"elastic shot gradient {shape[2]}x{shape[1]}x{shape[0]}, nt={nt}, sub={true.sub}, FP32 and "
                        f"BF16 (per-snapshot max-abs scale) snapshots, FP64 gradient accumulation",
            **res, "bf16_vs_f32_gradient_rel_l2": {"vp": err[0], "vs": err[1], "rho": err[2]},
            "autotuned_tiles_tx_ychunk": list(tiles)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-precision 256^3 gradients, the batched VTI comparison and the reduced "
                         "serving stream that the default run adds after the headline")
    ap.add_argument("--snap", default="f32", choices=("f32", "bf16", "f16"), help="snapshot precision")
    ap.add_argument("--unfused", action="store_true",
                    help="elastic / VTI steps with the flux, interior, slab and edge kernels (fwi3d_tune "
                         "FWI3D_TUNE_FUSED = 0) instead of the fused kernel")
    ap.add_argument("--overlap", type=int, default=None, choices=(0, 1, 2, 3),
                    help="streams of the fused step (fwi3d_tune FWI3D_TUNE_OVERLAP; default: the library's, 3)")
    ap.add_argument("--workload", default="gradient", choices=("gradient", "acoustic", "lbfgs", "batch", "serve"),
                    help="gradient: elastic shot gradient (default); acoustic: 100^3 acoustic forward modelling; "
                         "lbfgs: 256^3 mixed-precision l-BFGS FWI; batch: 32 VTI shots on 3 grids, batched vs one at "
                         "a time; serve: a seeded stream of mixed requests through the serving loop")
    ap.add_argument("--requests", type=int, default=48, help="requests of the serving stream (--workload serve)")
    ap.add_argument("--max-active", type=int, default=12, help="serving slots (--workload serve)")
    ap.add_argument("--hbm-budget-gib", type=float, default=200.0, help="serving HBM budget (--workload serve)")
    ap.add_argument("--iters", type=int, default=10, help="l-BFGS iterations (--workload lbfgs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates the HBM traffic counters)")
    args = ap.parse_args()
    if args.unfused:
        ENGINE_TUNE["fused"] = False
    if args.overlap is not None:
        ENGINE_TUNE["overlap"] = args.overlap
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (shots would shard one per GPU); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_probe:
        copy_probe(torch, args.copy_probe)
        return
    import fwi3d_amd as F
    from fwi3d_amd import workloads as W

    if args.workload == "lbfgs":
        run_lbfgs(args, torch, F, W)
        return
    if args.workload == "batch":
        run_batch(args, torch, F, W)
        return
    if args.workload == "serve":
        run_serve(args, torch, F, W)
        return
    if args.workload == "acoustic":
        run_acoustic(args, torch, F, W)
        return

    shot = W.config_b(n=args.n, nt=args.nt)
    obs = W.config_b(n=args.n, nt=args.nt, anomaly=False)
    shape = shot.shape
    ny, nx, nz = shape
    N = nx * ny * nz
    nt = shot.nt
    dev = torch.device("cuda:0")
    eng = new_engine(F)

    geo = F.Shot(shape, shot.d, (0.0, 0.0, 0.0), nt, shot.dt, shot.src_loc[0], shot.rcv_loc)
    # the reference's param for this shot: raw (vp, vs, rho), explosive moment tensor,
    # propagation dt = data dt (sinc resampling still runs), imaging every 4 steps
    opkw = dict(nmodels=3, dt=shot.dt, sub=shot.sub, mt=1, bc=shot.bc, taper=shot.taper,
                taper_strength=shot.taper_strength, snap_precision=args.snap)

    model = torch.from_numpy(np.ascontiguousarray(shot.raw)).to(dev)
    model_obs = torch.from_numpy(np.ascontiguousarray(obs.raw)).to(dev)
    src = torch.from_numpy(np.ascontiguousarray(shot.src[:, 0, :])).to(dev)  # [6][nt] wavelets
    dobs = torch.zeros((3, geo.nrcv, nt), dtype=torch.float32, device=dev)
    grad = torch.zeros((3, ny, nx, nz), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    # tile shapes of the time-step kernels for this physics and grid (fwi3d_autotune: every
    # TX / y-chunk pair timed on this grid; results do not depend on the choice)
    tiles = eng.autotune("elastic", shape, steps=10)
    # observed data: the anomaly-free model (not timed)
    F.apply_forward_device(eng, F.op_param(**dict(opkw, sub=0)), geo, False, model_obs.data_ptr(), src.data_ptr(),
                           dobs.data_ptr())

    misfits = []

    def shot_gradient():
        # one shot of nlls_fwi::compute_res_and_grad: forward with snapshots, residual and
        # FP64 misfit, adjoint with the imaging condition, quadrature and chain rule
        grad.zero_()
        misfits.append(F.fwi_shot_device(eng, F.op_param(**opkw), geo, model.data_ptr(), src.data_ptr(),
                                         dobs.data_ptr(), grad.data_ptr()))

    # per-kernel launch times: one untimed shot gradient with every kernel bracketed by
    # HIP events (also the warm-up)
    eng.profile(True)
    shot_gradient()
    torch.cuda.synchronize()
    kern = {name: eng.profile_read(k)
            for k, name in KERNEL_CLASSES[not args.unfused]}
    eng.profile(False)
    for _ in range(max(0, args.warmup - 1)):
        shot_gradient()
    torch.cuda.synchronize()
    # timed region: HIP events around each time step only
    eng.profile(True, steps_only=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        shot_gradient()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = dict(kern, elastic_time_step=eng.profile_read(4))
    eng.profile(False)

    g = grad.cpu().numpy()
    finite = bool(np.isfinite(g).all() and np.abs(g).max() > 0)

    updates = 2.0 * N * (nt - 1) * args.steps  # forward + adjoint cell-updates
    gcell = updates / elapsed / 1e9
    # one elastic time step = the launches of its step form (STEP_KERNEL), timed as a whole
    # (HIP events around the step on the engine stream, after both streams join)
    avg_us = {k: ms / max(n, 1) * 1e3 for k, (ms, n) in prof.items()}
    t_pair = avg_us.pop("elastic_time_step") / 1e6
    achieved = ELASTIC_BYTES_PER_CELL * N / t_pair / 1e9 if t_pair > 0 else 0.0
    traffic = measured_traffic(args.n)
    per_kernel = kernel_rooflines(not args.unfused, nz, nx, ny, avg_us, args.n)

    out = {
        "metric": METRIC,
        "value": round(gcell, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"elastic isotropic forward + adjoint gradient, 1 shot, {nz}x{nx}x{ny}, nt={nt}, "
                               f"dt={shot.dt}, sub={shot.sub}, {args.snap.upper()} snapshots, FP64 misfit and gradient "
                               f"accumulation (fwi3d_fwi_shot_device)",
                   "nz": nz, "nx": nx, "ny": ny, "nt": nt, "receivers": int(geo.nrcv) * 3,
                   "autotuned_tiles_tx_ychunk": list(tiles)},
        "ms_per_shot_gradient": round(elapsed / args.steps * 1e3, 2),
        "time_steps_per_iteration": 2 * (nt - 1),
        "step_definition": "one bench step is one shot gradient: forward propagation with snapshots, residual, "
                           "adjoint propagation with the imaging condition (2 x (nt-1) time steps)",
        "shot_gradients_per_s": round(args.steps / elapsed, 4),
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4)},
        "misfit": misfits[-1] if misfits else None,
        "gradient_finite": finite,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_profile(args.n) if traffic else None,
            "kernel": STEP_KERNEL[not args.unfused],
            "step_us": round(t_pair * 1e6, 2),
            "bytes_per_cell": ELASTIC_BYTES_PER_CELL,
            "cells_per_step": N,
            "avg_us": {k: round(v, 2) for k, v in avg_us.items()},
            "per_kernel": per_kernel,
            "launches": {k: n for k, (ms, n) in prof.items() if k != "elastic_time_step"},
            "step_timing": "timed region: HIP events around each time step; per-kernel averages from one untimed shot "
                           "gradient with every kernel bracketed by events",
        },
    }
    eng.close()
    del model, model_obs, src, dobs, grad
    torch.cuda.empty_cache()
    if not args.no_extras and args.n == 200 and args.nt == 2000:
        # the other configurations, untimed by the headline: mixed-precision 256^3
        # gradients, 32 batched VTI shots against one at a time, a reduced serving stream
        print("# extras: 256^3 gradients (FP32, BF16 snapshots)", file=sys.stderr, flush=True)
        out["config_mixed_precision_256"] = config_d_gradients(torch, F, W)
        torch.cuda.empty_cache()
        print("# extras: batched VTI shots", file=sys.stderr, flush=True)
        out["config_batched_vti"] = run_batch(argparse.Namespace(nt=1500), torch, F, W, emit=False)
        torch.cuda.empty_cache()
        print("# extras: serving stream (16 requests)", file=sys.stderr, flush=True)
        sv = run_serve(argparse.Namespace(requests=16, max_active=12, hbm_budget_gib=200.0), torch, F, W, emit=False)
        out["config_serving"] = {k: sv[k] for k in ("config", "shot_gradients_per_s", "requests_per_s",
                                                    "speedup_vs_plain", "poisson", "plain_latency_s",
                                                    "results_identical_to_plain")}
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(shot)
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, 'item') else str(o)), flush=True)


if __name__ == "__main__":
    main()
