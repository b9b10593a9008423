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


def cpu_baseline(shot, seconds=15.0):
    """Time the oracle (FP32 CPU restatement, OpenMP) on a bounded sample of the same
    workload: forward time steps of the same model and geometry."""
    from oracle import oracle as O
    shape = shot.shape
    inj = O.Injector(shot.src_kind, shot.src_loc, shape, shot.d, (0, 0, 0))
    ext = O.Injector(0, shot.rcv_loc, shape, shot.d, (0, 0, 0))

    def run(nsteps):
        par = O.make_par(nsteps + 1, 0, shot.dt, bc=shot.bc, taper=shot.taper, taper_strength=shot.taper_strength)
        src = np.ascontiguousarray(shot.src[..., :nsteps + 1])
        t = time.perf_counter()
        O.propagate(shot.physics, shot.model, src, inj, ext, par, shot.d)
        return time.perf_counter() - t

    t2 = run(2)
    nsteps = int(max(2, min(200, seconds / max(t2 / 2, 1e-3))))
    t = run(nsteps)
    cells = float(np.prod(shape))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(cells * nsteps / t / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{nsteps} forward time steps of the same {shape[2]}x{shape[1]}x{shape[0]} {shot.physics} shot "
                      f"(oracle FP32 C++/OpenMP, {t:.1f} s)"}


def measured_traffic(n):
    """HBM bytes per elastic time step (scaled FETCH_SIZE + WRITE_SIZE of the step kernels)
    from the committed rocprofv3 --pmc summary of the same workload, or None."""
    path = os.path.join(ROOT, "profiles", f"elastic_{n}cube_step_traffic_pmc.md")
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
    eng = F.Engine(0)
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


def run_batch(args, torch, F, W):
    """Batched VTI forward modelling: 32 shots split 11/11/10 over the grids (nz,nx,ny) =
    (128,128,128), (200,160,160), (96,192,192), nt=1500 (or --nt), dt=0.8 ms. The same
    shots run one at a time through the plain path (fwi3d_propagate_device) and then
    together through the batch (fwi3d_batch_*: every shot admitted at step 0, one launch
    per kernel class and step); traces of both are compared bit for bit."""
    nt = args.nt if args.nt != 2000 else 1500
    split = (11, 11, 10)
    shots = [W.config_c_shot(g, k % 12, nt=nt) for g in range(3) for k in range(split[g])]
    dev = torch.device("cuda:0")
    eng = F.Engine(0)
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
                               f"class and time step (fwi3d_batch)", "shots": len(shots)},
        "plain_gcell_s": round(cells / t_plain / 1e9, 3),
        "shots_per_s": {"batched": round(len(shots) / t_batch, 4), "plain": round(len(shots) / t_plain, 4)},
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "plain_latency_s": {"p50": round(float(np.percentile(lat_plain, 50)), 4),
                            "p99": round(float(np.percentile(lat_plain, 99)), 4)},
        "batch_resident_gib": round(resident / 2**30, 2),
        "autotuned_tiles_tx_ychunk": tiles,
        "batch_online_ychunk": batch_tiles[0],
        "batch_candidate_ms_per_step": {str(y): round(t, 3) for y, t in zip((4, 8, 16, 32), batch_tiles[1])},
        "traces_identical_to_plain": bool(same),
    }
    print(json.dumps(out), flush=True)
    eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=200)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--snap", default="f32", choices=("f32", "bf16", "f16"), help="snapshot precision")
    ap.add_argument("--workload", default="gradient", choices=("gradient", "lbfgs", "batch"),
                    help="gradient: elastic shot gradient (default); lbfgs: 256^3 mixed-precision l-BFGS FWI; "
                         "batch: 32 VTI shots on 3 grids, batched vs one at a time")
    ap.add_argument("--iters", type=int, default=10, help="l-BFGS iterations (--workload lbfgs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates the HBM traffic counters)")
    args = ap.parse_args()
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

    shot = W.config_b(n=args.n, nt=args.nt)
    obs = W.config_b(n=args.n, nt=args.nt, anomaly=False)
    shape = shot.shape
    ny, nx, nz = shape
    N = nx * ny * nz
    nt = shot.nt
    dev = torch.device("cuda:0")
    eng = F.Engine(0)
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

    for _ in range(args.warmup):
        shot_gradient()
    torch.cuda.synchronize()
    eng.profile(True)
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
    prof = {name: eng.profile_read(k)
            for k, name in ((0, "elastic_interior"), (1, "elastic_flux"), (2, "elastic_boundary"))}
    eng.profile(False)

    g = grad.cpu().numpy()
    finite = bool(np.isfinite(g).all() and np.abs(g).max() > 0)

    updates = 2.0 * N * (nt - 1) * args.steps  # forward + adjoint cell-updates
    gcell = updates / elapsed / 1e9
    # one elastic time step = flux pre-pass + interior y-march + boundary shell launches
    avg_us = {k: ms / max(n, 1) * 1e3 for k, (ms, n) in prof.items()}
    t_pair = sum(avg_us.values()) / 1e6
    achieved = ELASTIC_BYTES_PER_CELL * N / t_pair / 1e9 if t_pair > 0 else 0.0
    traffic = measured_traffic(args.n)

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
                   "nz": nz, "nx": nx, "ny": ny, "nt": nt, "receivers": int(geo.nrcv) * 3},
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
            "traffic_source": f"profiles/elastic_{args.n}cube_step_traffic_pmc.md" if traffic else None,
            "kernel": "elastic time step = elastic_flux + elastic_interior + elastic_boundary launches "
                      "(HIP events on the engine stream)",
            "bytes_per_cell": ELASTIC_BYTES_PER_CELL,
            "cells_per_step": N,
            "avg_us": {k: round(v, 2) for k, v in avg_us.items()},
            "launches": {k: n for k, (ms, n) in prof.items()},
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(shot)
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, 'item') else str(o)), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
