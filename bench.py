"""Benchmark of the Joseph forward + matched adjoint on one MI355X.

Workload (the 3D cone-beam drop-in case): 256^3 FP32 modified Shepp-Logan volume, 360 poses
over 360 degrees (LinSpaced, both ends), 384x384 detector, source-centre 25600, centre-
detector 1280. One step = one forward projection plus one back projection of the whole
volume/sinogram, inputs resident in HBM.

Prints ONE JSON line:
  value       whole-job Joseph samples per second, (S_fwd + S_adj) * steps / wall time, Gsamples/s
  roofline    dominant kernel (the longest launch), bound by its vector ALU: VALU lane operations
              per second (SQ_INSTS_VALU x 64 per launch, from the committed counters in
              profiles/, over its mean duration from HIP events on the launch stream) against
              the VALU's published 78.6 T lane operations/s (157.3 TFLOP/s FP32); beside it the
              counter-measured HBM bytes against 8 TB/s and the algorithmic tap bytes (16 B per
              3D sample, served from L2 / LDS). The back projection runs one launch per driving
              axis; each is given the share of the pass's time that its samples have
  cpu_baseline  the CPU restatement (oracle/, OpenMP over rays) on a bounded sample of the same
              workload (a subset of the poses spread over the orbit), on this host's cores: median
              of 5 runs with every thread and with one thread
  cg          CGLS (elsa's CGLS.cpp semantics) from b = A x_true on the device, eagerly and with
              one iteration captured in a hipGraph: wall time to reach a relative error of 1e-2
              against x* = the FP64 CGLS iterate after 100 iterations on the same b, and the
              residual and errors after 20 iterations

--projector siddon | blob | bspline times those projectors on the same workload instead
(voxel crossings or footprint taps per second), without the CG and CPU parts.

--workload batch2d_64 is the batched form: 64 independent 2D fan-beam problems of sizes drawn
from {128,192,256,384,512}^2 with 90-360 poses each from a random start angle, random-ellipse
phantoms with 1% Gaussian noise. value = Joseph samples per second of one fused forward +
back projection pass over all problems; "solve" times 50 CGLS + 50 FISTA (L1, lambda 1e-3)
iterations in hipGraphs on every problem, once problem by problem (the plain operator) and
once as one batch, and checks both reach the same reconstructions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Joseph fwd+adj Gsamples/s & %HBM roofline; CG wall time to fixed rel. error"
HBM_PEAK_GBS = 8000.0
# MI355X FP32 vector peak 157.3 TFLOP/s = 78.6 T lane operations per second (FMA = 2 FLOP)
VALU_PEAK_TOPS = 78.64
WORKLOADS = {
    "cone3d_256": dict(dim=3, vol=(256, 256, 256), views=360, det=(384, 384), s2c=25600.0, c2d=1280.0,
                       text="3D cone-beam Joseph fwd+adj: 256^3 f32 Shepp-Logan, 360 poses over 360 deg, "
                            "384x384 detector, s2c 25600, c2d 1280"),
    "fan2d_256": dict(dim=2, vol=(256, 256), views=180, det=(384,), s2c=25600.0, c2d=1280.0,
                      text="2D fan-beam Joseph fwd+adj: 256^2 f32 Shepp-Logan, 180 poses over 360 deg, "
                           "384 detector, s2c 25600, c2d 1280"),
    "mixed512": dict(dim=3, vol=(512, 512, 512), views=720, det=(768, 768), s2c=51200.0, c2d=2560.0, sub=64,
                     text="mixed precision: Joseph and blob voxel (r 2, alpha 10.83, m 2) fwd+adj at 512^3 f32 / bf16, "
                          "720 poses, 768x768 detector, s2c 51200, c2d 2560; CGLS FP32 vs bf16 with iterative "
                          "refinement on a 64^3 subproblem"),
    "autograd128": dict(dim=3, vol=(128, 128, 128), views=180, det=(192, 192), s2c=12800.0, c2d=640.0, batch=4,
                        text="differentiable Joseph: 4 x 128^3 f32 volumes of 20 random ellipsoids, 180 poses, "
                             "192x192 detector; one step = loss 0.5||Ax-b||^2 + 0.01 TV(x), backward (hand-written "
                             "adjoint), gradient step"),
    "batch2d_64": dict(dim=2, n=64, sizes=(128, 192, 256, 384, 512), views=(90, 360), seed=2, cg=50, fista=50,
                       lam=1e-3,
                       text="64 independent 2D fan-beam Joseph problems: N in {128,192,256,384,512}, 90-360 poses "
                            "over 360 deg from a random start, detector floor(N sqrt 2), s2c 100 N, c2d 5 N, "
                            "random ellipses + 1% noise; one fused fwd+adj pass per step"),
}


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="cone3d_256", choices=sorted(WORKLOADS))
    p.add_argument("--cpu-poses", type=int, nargs=2, default=[40, 2], metavar=("ALL", "ONE"),
                   help="poses of the bounded CPU samples (evenly spread over the trajectory) timed with every "
                        "host thread and with one thread")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--kernels", default="fast", choices=["fast", "simple"],
                   help="Joseph kernels: plane-major forward + LDS-tiled back projection (default) or one "
                        "thread per ray")
    p.add_argument("--tile", default="", help="LA,TB,TC of the tiled Joseph / Siddon back projection, or W,H "
                                              "pixels of the tiled voxel forward (default: built-in), or "
                                              "'auto' to time a set of candidates first (untimed)")
    p.add_argument("--pose-batch", type=int, default=0,
                   help="poses per step of the tiled back projection (default: built-in)")
    p.add_argument("--projector", default="joseph", choices=["joseph", "siddon", "blob", "bspline"],
                   help="projector to time (default: Joseph, the headline metric)")
    p.add_argument("--cg-iters", type=int, default=20,
                   help="CGLS iterations of the fixed-count reconstruction part (0 skips the CG part)")
    p.add_argument("--cg-target", type=float, default=1e-2,
                   help="relative error to the FP64 reference iterate that the timed CGLS must reach")
    p.add_argument("--cg-ref-iters", type=int, default=100,
                   help="FP64 CGLS iterations of the reference solution x*")
    p.add_argument("--dtype", default="f32", choices=["f32", "f64"],
                   help="data type of the operator (f64: elsa's data_t = double classes; the CG leg and the "
                        "CPU baseline are measured for f32 only)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB (3 launches), the reference kernel for "
                        "scaling FETCH_SIZE / WRITE_SIZE counters, and exit")
    return p.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: counters read under it convert to bytes. Also times
    it, which gives the device-to-device copy rate (read + write bytes per second)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / 3
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 4,
                      "copy_GBps": round(2 * n * 4 / t / 1e9, 1)}), flush=True)


def _cpu_run(wl, poses, threads, reps):
    """One oracle timing (oracle/cpu_time.py) in a child process: OMP_NUM_THREADS and
    OMP_PROC_BIND=close take effect there, not in this process whose OpenMP runtime is loaded."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, "-m", "oracle.cpu_time", "--dim", str(wl["dim"]), "--vol", *map(str, wl["vol"]),
           "--views", str(wl["views"]), "--det", *map(str, wl["det"]), "--s2c", str(wl["s2c"]),
           "--c2d", str(wl["c2d"]), "--poses", str(poses), "--threads", str(threads), "--reps", str(reps)]
    out = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, check=True).stdout
    return json.loads(out.strip().splitlines()[-1])


def cpu_baseline(wl, poses_all, poses_one, reps=5):
    """The oracle (FP32, elsa's own position rule, OpenMP over rays) on a subset of the poses:
    median of `reps` runs after one warm-up, once with every host thread this process may use
    and once with one thread, threads pinned to cores."""
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
    allt = _cpu_run(wl, poses_all, threads, reps)
    one = _cpu_run(wl, poses_one, 1, reps)
    return {"value": allt["gsamples_per_s"], "unit": "Gsamples/s", "cores": threads, "kind": "port",
            "sample": f"{poses_all} of {wl['views']} poses of the same workload spread over the orbit (Joseph fwd+adj, "
                      f"{allt['samples']} samples per run), median of {reps} runs ({allt['median_s']} s) after a "
                      f"warm-up; oracle/elsa_oracle.cpp, OpenMP over rays, OMP_PROC_BIND=close",
            "cpu_model": allt["cpu_model"], "runs_s": allt["times_s"],
            "one_thread": {"value": one["gsamples_per_s"], "unit": "Gsamples/s", "cores": 1,
                           "sample": f"{poses_one} of {wl['views']} poses ({one['samples']} samples per run), "
                                     f"median of {reps} runs ({one['median_s']} s)", "runs_s": one["times_s"]}}


def cg_run(E, op, x, iters, vd, dd, target, ref_iters):
    """CGLS (elsa's CGLS.cpp semantics, tol 0) from b = A x_true, whole iterations on the
    device.

    Wall time to a fixed relative error: x* is the FP64 CGLS iterate after ref_iters
    iterations on the same b (FP64 operator and reductions; an iteration-count proxy for the
    FP64 solution of the normal equations). The FP32 iteration is deterministic (fixed-point
    back projection), so solve(k) from scratch gives its k-th iterate: the smallest k with
    ||x_k - x*|| / ||x*|| <= target is found by bisection, untimed, then solve(k) is timed eagerly
    and with one iteration captured in a hipGraph. Also the error and residual after `iters`
    iterations."""
    import torch

    b = op.apply(x)
    res = {"iterations": iters}
    op64 = E.JosephsMethodCUDA(vd, dd, dtype=np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs = E.CGLS(op64, b.double(), tol=0.0, graph=True).solve(ref_iters)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    nrm = float(torch.linalg.norm(xs))
    del op64

    def err(k):
        xr = E.CGLS(op, b, tol=0.0, graph=True).solve(k)
        return float(torch.linalg.norm(xr.double() - xs)) / nrm

    errs = {}
    lo, hi = 0, ref_iters
    errs[hi] = err(hi)
    if errs[hi] <= target:
        while hi - lo > 1:
            mid = (lo + hi) // 2
            errs[mid] = err(mid)
            if errs[mid] <= target:
                hi = mid
            else:
                lo = mid
        k = hi
    else:
        k = None
    res.update({"target_rel_error": target,
                "reference": f"FP64 CGLS iterate after {ref_iters} iterations on the same b (FP64 operator), "
                             f"{t_ref:.2f} s",
                "iterations_to_target": k, "rel_error_at_target": errs.get(k) if k else None,
                "rel_error_trace": {str(kk): round(v, 6) for kk, v in sorted(errs.items())}})
    for graph in (False, True):
        for n, key in ((k, "to_target"), (iters, "fixed")):
            if n is None:
                continue
            solver = E.CGLS(op, b, tol=0.0, graph=graph)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xr = solver.solve(n)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            mode = "graph" if graph else "eager"
            res[f"wall_s_{key}_{mode}"] = round(dt, 4)
            res[f"ms_per_iter_{key}_{mode}"] = round(dt / n * 1e3, 3)
            if key == "fixed":
                res["rel_residual_fixed"] = float(solver.residuals[solver.iterations_done] / solver.residuals[0])
                res["rel_error_vs_phantom_fixed"] = float(torch.linalg.norm(xr - x) / torch.linalg.norm(x))
                res["rel_error_vs_reference_fixed"] = float(torch.linalg.norm(xr.double() - xs)) / nrm
    return res


def batch_problems(E, wl):
    """The operators, phantoms and noisy sinograms of the batched workload (seeded)."""
    import torch

    from elsa_amd.phantoms import random_ellipsoids

    rng = np.random.default_rng(wl["seed"])
    ops, xs, bs = [], [], []
    for j in range(wl["n"]):
        N = int(rng.choice(wl["sizes"]))
        V = int(rng.integers(wl["views"][0], wl["views"][1] + 1))
        start = float(rng.uniform(0, 360))
        vd = E.VolumeDescriptor([N, N])
        det = int(np.float32(N) * np.float32(np.sqrt(np.float32(2))))
        deg = np.float32(start) + np.linspace(0, 360, V, dtype=np.float32)
        rad = (deg * np.float32(np.pi) / np.float32(180)).astype(np.float32)
        geoms = [E.Geometry(100.0 * N, 5.0 * N, vd, [det], float(r)) for r in rad]
        dd = E.PlanarDetectorDescriptor([det, V], geoms)
        op = E.JosephsMethodCUDA(vd, dd)
        x = torch.from_numpy(random_ellipsoids((N, N), n=10, seed=1000 + j)).cuda()
        y = op.apply(x)
        sigma = 0.01 * float(y.max())
        g = torch.Generator(device="cuda").manual_seed(2000 + j)
        y = y + sigma * torch.randn(y.shape, device="cuda", generator=g)
        ops.append(op)
        xs.append(x)
        bs.append(y.contiguous())
    return ops, xs, bs


def batch_main(a, E, wl):
    """Batched workload: fused passes over all problems, and the solves problem by problem
    against the batch."""
    import torch

    ops, xs, bs = batch_problems(E, wl)
    tuned = None
    if a.tile in ("", "auto"):  # the batched workload's tiles are chosen by the autotuner
        B, tile, pose_batch, times = E.autotune_batch(ops)
        tuned = {"tile": tile, "pose_batch": pose_batch}
        print(json.dumps({"autotune_batch": {str(k): round(v, 3) for k, v in times.items()}}), file=sys.stderr)
    else:
        for op in ops:
            if a.tile:
                op.tile(*[int(v) for v in a.tile.split(",")])
            if a.pose_batch:
                op.pose_batch(a.pose_batch)
        B = E.ProjectorBatch(ops)
    S = sum(op.sample_count() for op in ops)
    R = int(B.sino_offsets[-1])
    x = B.pack(xs)
    sino = torch.empty(B._n_sino, dtype=torch.float32, device="cuda")
    back = torch.empty(B._n_vol, dtype=torch.float32, device="cuda")
    for _ in range(a.warmup):
        B.apply(x, sino)
        B.applyAdjoint(sino, back)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for e0, e1, e2 in evs:
        e0.record()
        B.apply(x, sino)
        e1.record()
        B.applyAdjoint(sino, back)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_fwd = float(np.mean([e0.elapsed_time(e1) for e0, e1, _ in evs])) * 1e-3
    t_adj = float(np.mean([e1.elapsed_time(e2) for _, e1, e2 in evs])) * 1e-3
    # the same pass operator by operator
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        for op, xj in zip(ops, xs):
            op.applyAdjoint(op.apply(xj))
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t1

    solve = {"cgls_iterations": wl["cg"], "fista_iterations": wl["fista"], "fista_lambda": wl["lam"]}
    b = B.pack(bs, adjoint_space=True)

    def run_batch():
        cg = E.CGLS(B, b, tol=0.0, graph=True)
        xc = cg.solve(wl["cg"])
        ap = E.APGD(B, b, prox=E.ProximalL1(wl["lam"]), graph=True)
        xa = ap.solve(wl["fista"])
        return B.unpack(xc), B.unpack(xa)

    def run_plain():
        rc, ra = [], []
        for op, bj in zip(ops, bs):
            rc.append(E.CGLS(op, bj, tol=0.0, graph=True).solve(wl["cg"]))
            ra.append(E.APGD(op, bj, prox=E.ProximalL1(wl["lam"]), graph=True).solve(wl["fista"]))
        return rc, ra

    for name, fn in (("plain", run_plain), ("batched", run_batch)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        xc, xa = fn()
        torch.cuda.synchronize()
        solve[f"wall_s_{name}"] = round(time.perf_counter() - t, 3)
        solve[f"{name}_rel_error_cgls"] = [round(float(torch.linalg.norm(u - v) / torch.linalg.norm(v)), 6)
                                           for u, v in zip(xc, xs)]
        solve[f"{name}_rel_error_fista"] = [round(float(torch.linalg.norm(u - v) / torch.linalg.norm(v)), 6)
                                            for u, v in zip(xa, xs)]
        if name == "plain":
            ref = (xc, xa)
        else:
            solve["max_abs_diff_batched_vs_plain"] = max(float((u - v).abs().max()) for u, v in
                                                         zip(ref[0] + ref[1], xc + xa))
    solve["speedup"] = round(solve["wall_s_plain"] / solve["wall_s_batched"], 2)
    for name in ("plain", "batched"):
        for k in ("cgls", "fista"):
            errs = solve.pop(f"{name}_rel_error_{k}")
            solve[f"{name}_rel_error_{k}_mean"] = round(float(np.mean(errs)), 6)
    out = {
        "metric": METRIC,
        "value": round(2 * S * a.steps / wall / 1e9, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (random ellipses, 1% Gaussian noise)",
        "config": {"workload": wl["text"], "problems": len(ops), "fused": B.fused, "samples_per_pass": S,
                   "rays": R, "voxels": int(B.vol_offsets[-1]), "fwd_ms": round(t_fwd * 1e3, 3),
                   "adj_ms": round(t_adj * 1e3, 3), "ms_per_step_operator_by_operator":
                       round(wall_plain / a.steps * 1e3, 3), "tile": list(ops[0].tile()),
                   "pose_batch": ops[0].pose_batch(), "autotuned": tuned},
        "solve": solve,
    }
    print(json.dumps(out), flush=True)


def _line(a, wl, value, ms, dtype, data, config, **extra):
    out = {"metric": METRIC, "value": value, "unit": "Gsamples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": data, "config": {"workload": wl["text"], **config}}
    out.update(extra)
    print(json.dumps(out), flush=True)


def mixed_main(a, E, wl):
    """Joseph and the blob voxel projector at 512^3 in FP32 and bfloat16 storage, then CGLS in
    FP32 against bfloat16 CGLS with iterative refinement on a 64^3 subproblem, both measured
    against an FP64 GPU CGLS of 400 iterations (the reference iterate) at equal iteration
    budgets."""
    import torch

    from elsa_amd.phantoms import modified_shepp_logan

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    n = wl["vol"][0]
    vd = E.VolumeDescriptor(list(wl["vol"]))
    dd = E.CircleTrajectoryGenerator.createTrajectory(wl["views"], vd, 360, wl["s2c"], wl["c2d"],
                                                      detectorSize=list(wl["det"]))
    passes = {}
    S = None
    for name, dt in (("f32", np.float32), ("bf16", "bfloat16")):
        op = E.JosephsMethodCUDA(vd, dd, dtype=dt)
        S = op.sample_count()
        tdt = torch.float32 if name == "f32" else torch.bfloat16
        x = torch.from_numpy(modified_shepp_logan((n, n, n))).cuda().to(tdt)
        sino = torch.empty(dd.shape, dtype=tdt, device="cuda")
        back = torch.empty(vd.shape, dtype=tdt, device="cuda")
        tf = timed(lambda: op.apply(x, sino), a.steps)
        ta = timed(lambda: op.applyAdjoint(sino, back), a.steps)
        passes[name] = {"fwd_ms": round(tf, 2), "adj_ms": round(ta, 2), "Gsamples_per_s": round(2 * S / (tf + ta) / 1e6, 1)}
        print(f"joseph {name}: {passes[name]}", file=sys.stderr, flush=True)
        del op, x, sino, back
        torch.cuda.empty_cache()
    # the blob voxel projector (Kaiser-Bessel, radius 2, alpha 10.83, order 2) on the same problem
    blob = {}
    for name, dt in (("f32", np.float32), ("bf16", "bfloat16")):
        op = E.BlobVoxelProjectorCUDA(vd, dd, radius=2.0, alpha=10.83, order=2, dtype=dt)
        taps = op.sample_count()
        tdt = torch.float32 if name == "f32" else torch.bfloat16
        x = torch.from_numpy(modified_shepp_logan((n, n, n))).cuda().to(tdt)
        sino = torch.empty(dd.shape, dtype=tdt, device="cuda")
        back = torch.empty(vd.shape, dtype=tdt, device="cuda")
        tf = timed(lambda: op.apply(x, sino), 1)
        ta = timed(lambda: op.applyAdjoint(sino, back), 1)
        blob[name] = {"fwd_ms": round(tf, 1), "adj_ms": round(ta, 1), "taps_per_pass": taps,
                      "Gtaps_per_s": round(2 * taps / (tf + ta) / 1e6, 1)}
        print(f"blob {name}: {blob[name]}", file=sys.stderr, flush=True)
        del op, x, sino, back
        torch.cuda.empty_cache()
    m = wl["sub"]
    svd = E.VolumeDescriptor([m, m, m])
    sdd = E.CircleTrajectoryGenerator.createTrajectory(wl["views"] // 6, svd, 360, 100.0 * m, 5.0 * m,
                                                       detectorSize=[m * 3 // 2, m * 3 // 2])
    A64 = E.JosephsMethodCUDA(svd, sdd, dtype=np.float64)
    A32 = E.JosephsMethodCUDA(svd, sdd)
    A16 = E.JosephsMethodCUDA(svd, sdd, dtype="bfloat16")
    xt = torch.from_numpy(modified_shepp_logan((m, m, m))).cuda()
    b = A32.apply(xt)
    x_ref = E.CGLS(A64, b.double(), tol=0.0, graph=True).solve(400)
    err = lambda x: round(float(torch.linalg.norm(x.double() - x_ref) / torch.linalg.norm(x_ref)), 5)
    cg = {}
    for it in (40, 80, 160):
        torch.cuda.synchronize()
        t = time.perf_counter()
        x = E.CGLS(A32, b, tol=0.0, graph=True).solve(it)
        torch.cuda.synchronize()
        cg[f"fp32_cgls_{it}"] = {"rel_error": err(x), "wall_s": round(time.perf_counter() - t, 4)}
    for inner, outer in ((20, 2), (20, 4), (40, 4)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = E.CGLSRefined(A32, A16, b, inner=inner, graph=True)
        x = r.solve(outer)
        torch.cuda.synchronize()
        cg[f"bf16_refined_{outer}x{inner}"] = {"rel_error": err(x), "wall_s": round(time.perf_counter() - t, 4)}
    f32 = passes["f32"]
    _line(a, wl, f32["Gsamples_per_s"], round(f32["fwd_ms"] + f32["adj_ms"], 2), "f32 and bf16 (storage)",
          "synthetic (modified Shepp-Logan phantom)",
          {"volume": list(wl["vol"]), "poses": wl["views"], "detector": list(wl["det"]), "samples_per_pass": S,
           "passes": passes, "blob_passes": blob, "subproblem": {"volume": [m] * 3, "poses": wl["views"] // 6,
                                            "reference": "FP64 CGLS, 400 iterations, on the GPU"}},
          solve=cg)


def autograd_main(a, E, wl):
    """Training steps through the autograd wrapper (backward = the hand-written adjoint), plus
    the dot-product test and FP64 central finite differences (h = 1e-3) of the same loss on 64
    random directions."""
    import torch

    from elsa_amd.autograd import project
    from elsa_amd.phantoms import random_ellipsoids

    n, B = wl["vol"][0], wl["batch"]
    vd = E.VolumeDescriptor(list(wl["vol"]))
    dd = E.CircleTrajectoryGenerator.createTrajectory(wl["views"], vd, 360, wl["s2c"], wl["c2d"],
                                                      detectorSize=list(wl["det"]))
    op = E.JosephsMethodCUDA(vd, dd)
    S = op.sample_count()
    truth = torch.stack([torch.from_numpy(random_ellipsoids((n, n, n), 20, seed=1 + i)) for i in range(B)]).cuda()
    with torch.no_grad():
        b = project(op, truth)

    def tv(v):
        dz = v[:, 1:, :-1, :-1] - v[:, :-1, :-1, :-1]
        dy = v[:, :-1, 1:, :-1] - v[:, :-1, :-1, :-1]
        dx = v[:, :-1, :-1, 1:] - v[:, :-1, :-1, :-1]
        return torch.sqrt(dx * dx + dy * dy + dz * dz + 1e-8).sum()

    def loss_fn(v, A=op, bb=b):
        r = project(A, v) - bb
        return 0.5 * (r * r).sum() + 0.01 * tv(v)

    x = torch.zeros_like(truth, requires_grad=True)
    lr = 1e-5

    def step():
        loss = loss_fn(x)
        loss.backward()
        with torch.no_grad():
            x.sub_(lr * x.grad)
            x.grad = None
        return loss

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        last = step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # checks (untimed): dot-product test in FP32 and FP64 finite differences on one volume
    g = torch.Generator(device="cuda").manual_seed(5)
    u = torch.rand(vd.shape, device="cuda", generator=g)
    w = torch.rand(dd.shape, device="cuda", generator=g)
    Au, Atw = op.apply(u), op.applyAdjoint(w)
    dot_rel = abs(float((Au.double() * w.double()).sum() - (u.double() * Atw.double()).sum())) / \
        float(torch.linalg.norm(Au.double()) * torch.linalg.norm(w.double()))
    op64 = E.JosephsMethodCUDA(vd, dd, dtype=np.float64)
    b64 = b[0].double()
    x64 = (truth[0].double() * 0.5).requires_grad_(True)
    l64 = loss_fn(x64[None], op64, b64[None])
    grad, = torch.autograd.grad(l64, x64)
    h, worst = 1e-3, 0.0
    with torch.no_grad():
        for _ in range(64):
            d = torch.randn(x64.shape, dtype=torch.float64, device="cuda", generator=g)
            fd = (loss_fn((x64 + h * d)[None], op64, b64[None]) - loss_fn((x64 - h * d)[None], op64, b64[None])) / (2 * h)
            an = (grad * d).sum()
            worst = max(worst, abs(float(fd - an)) / max(abs(float(an)), 1e-300))
    passes = 2 * B * a.steps
    _line(a, wl, round(passes * S / wall / 1e9, 3), round(wall / a.steps * 1e3, 3), "f32",
          "synthetic (random ellipsoids)",
          {"batch": B, "volume": list(wl["vol"]), "poses": wl["views"], "detector": list(wl["det"]),
           "samples_per_pass": S, "final_loss": float(last)},
          checks={"dot_product_rel": dot_rel, "fd_fp64_h": h, "fd_directions": 64, "fd_max_rel_error": worst})


def roofline(a, wl, name, launch_bytes, t_dom):
    """Roofline of the dominant kernel (the longest launch), from its live HIP-event time and the
    per-launch hardware counters committed under profiles/ for this projector and workload.

    Its bound is the vector ALU: it issues a VALU instruction nearly every cycle it can
    (VALUBusy ~100 %) while HBM is far from busy, because the taps are served from L2 (forward)
    and LDS (back projection). achieved = VALU lane operations per second, SQ_INSTS_VALU x 64
    lanes per launch over the launch's mean time; peak = the VALU's published capability,
    MI355X's 157.3 TFLOP/s FP32 counted as 78.6 T lane operations (an FMA is 2 FLOP). Beside it:
    the counter-measured HBM bytes per launch against 8 TB/s, and the algorithmic tap bytes of
    the launch (16 B per 3D sample), which exceed any HBM rate since they never reach HBM.
    Without a counter file for the launch, the HBM form of the algorithmic bytes is reported."""
    out = {"kernel": name, "ms_per_launch": round(t_dom * 1e3, 3), "algorithmic_bytes_per_launch": launch_bytes,
           "algorithmic_GBps": round(launch_bytes / t_dom / 1e9, 1)}
    key = a.projector + ("_f64" if a.dtype == "f64" else "")
    cfile = os.path.join(ROOT, "profiles", f"{key}_{a.workload}_counters.json")
    k = None
    if a.kernels == "fast" and os.path.exists(cfile):
        k = json.load(open(cfile))["kernels"].get(name)
    if k is None:
        achieved = launch_bytes / t_dom / 1e9
        out.update({"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "note": "algorithmic bytes (no counter file for this launch)"})
        return out
    achieved = k["valu_insts"] * 64 / t_dom / 1e12
    hbm = k["hbm_bytes"] / t_dom / 1e9
    out.update({"bound": "valu", "achieved": round(achieved, 2), "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
                "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": k["hbm_bytes"],
                "valu_busy_pct": k["valu_busy_pct"],
                "hbm": {"achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(hbm / HBM_PEAK_GBS, 4), "l2_hit_rate": k["l2_hit_rate"]},
                "counters": cfile[len(ROOT) + 1:]})
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {a.gpus}); the projectors are single-GPU here",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    import torch

    import elsa_amd as E
    from elsa_amd.phantoms import modified_shepp_logan

    wl = WORKLOADS[a.workload]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.workload == "batch2d_64":
        batch_main(a, E, wl)
        return
    if a.workload == "mixed512":
        mixed_main(a, E, wl)
        return
    if a.workload == "autograd128":
        autograd_main(a, E, wl)
        return
    vd = E.VolumeDescriptor(list(wl["vol"]))
    dd = E.CircleTrajectoryGenerator.createTrajectory(wl["views"], vd, 360, wl["s2c"], wl["c2d"],
                                                      detectorSize=list(wl["det"]))
    ndt = np.float64 if a.dtype == "f64" else np.float32
    if a.projector == "joseph":
        op = E.JosephsMethodCUDA(vd, dd, kernels=a.kernels, dtype=ndt)
    elif a.projector == "siddon":
        op = E.SiddonsMethodCUDA(vd, dd, kernels=a.kernels, dtype=ndt)
    elif a.projector == "blob":
        op = E.BlobVoxelProjectorCUDA(vd, dd, kernels=a.kernels, dtype=ndt)
    else:
        op = E.BSplineVoxelProjectorCUDA(vd, dd, kernels=a.kernels, dtype=ndt)
    tile = None
    S = op.sample_count()
    R = int(np.prod(wl["det"])) * wl["views"]
    x_np = modified_shepp_logan(wl["vol"][::-1]).astype(ndt)
    x = torch.from_numpy(x_np).to(dev)
    tdt = torch.float64 if a.dtype == "f64" else torch.float32
    if a.kernels == "fast":
        if a.tile == "auto":
            # tune on the workload's own data: the phantom for the voxel forward, its sinogram
            # for the back projections
            voxel = a.projector in ("blob", "bspline")
            tile, tile_times = E.autotune(op, x=x if voxel else op.apply(x))
            print(json.dumps({"autotune": {str(k): round(v, 3) for k, v in tile_times.items()}}), file=sys.stderr)
        else:
            tile = op.tile(*[int(v) for v in a.tile.split(",")]) if a.tile else op.tile()
            if a.pose_batch and a.projector in ("joseph", "siddon"):
                op.pose_batch(a.pose_batch)
    sino = torch.empty(dd.shape, dtype=tdt, device=dev)
    back = torch.empty(vd.shape, dtype=tdt, device=dev)

    for _ in range(a.warmup):
        op.apply(x, sino)
        op.applyAdjoint(sino, back)
    torch.cuda.synchronize()

    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in evs:
        e0.record()
        op.apply(x, sino)
        e1.record()
        op.applyAdjoint(sino, back)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    t_fwd = float(np.mean([e0.elapsed_time(e1) for e0, e1, _ in evs])) * 1e-3
    t_adj = float(np.mean([e1.elapsed_time(e2) for _, e1, e2 in evs])) * 1e-3
    nvox = int(np.prod(wl["vol"]))
    eb = 8 if a.dtype == "f64" else 4  # bytes per value
    if a.projector == "joseph":
        # 4 (2D: 2) taps per sample, one value per ray
        alg_bytes = (4 if wl["dim"] == 3 else 2) * eb * S + eb * R
        adj_name = "joseph_adj_tiled_kernel" if a.kernels == "fast" else "joseph_adj_kernel"
        fwd_name = "joseph_fwd_planes_kernel" if a.kernels == "fast" else "joseph_fwd_kernel"
        unit, metric = "Gsamples/s", METRIC
    elif a.projector == "siddon":
        # one voxel per crossing, one value per ray
        alg_bytes = eb * S + eb * R
        fwd_name = "siddon_fwd_planes_kernel" if a.kernels == "fast" else "siddon_kernel"
        adj_name = "siddon_adj_tiled_kernel" if a.kernels == "fast" else "siddon_kernel"
        unit, metric = "Gcrossings/s", "Siddon fwd+adj voxel crossings/s & %HBM roofline"
    else:
        # one detector value per footprint tap, one voxel per voxel and pose
        alg_bytes = eb * S + eb * nvox * wl["views"]
        # FP64: the tiled forward and the simple kernels' gather
        adj_name = "voxel_adj_fast_kernel" if a.kernels == "fast" and a.dtype == "f32" else "voxel_adj_kernel"
        fwd_name = "voxel_fwd_tiled_kernel" if a.kernels == "fast" else "voxel_fwd_kernel"
        unit, metric = "Gtaps/s", f"{a.projector} voxel projector fwd+adj footprint taps/s & %HBM roofline"
    # The tiled back projections run one launch per driving axis that occurs (a "pass"), each
    # a kernel of its own in rocprof's statistics. Their durations are not timed one by one, so
    # each launch gets the share of the back projection's time that its samples have; the
    # dominant kernel is the one with the longest launch.
    adj_launches = [(adj_name, alg_bytes, t_adj)]  # one pass, timed as a whole
    if a.kernels == "fast" and a.projector == "joseph":
        axis, nsteps = op.ray_info()
        per_axis = [(ax, int(nsteps[axis == ax].sum()), int(np.count_nonzero((axis == ax) & (nsteps > 0))))
                    for ax in range(wl["dim"])]
        per_axis = [(ax, s_a, r_a) for ax, s_a, r_a in per_axis if s_a > 0]
        s_tot = sum(s_a for _, s_a, _ in per_axis)
        b_tap = (4 if wl["dim"] == 3 else 2) * eb
        adj_launches = [(f"{adj_name}<{wl['dim']},{ax}>", b_tap * s_a + eb * r_a, t_adj * s_a / s_tot)
                        for ax, s_a, r_a in per_axis]
    adj_launch = max(adj_launches, key=lambda x: x[2])
    if adj_launch[2] >= t_fwd:
        name, launch_bytes, t_dom = adj_launch
    else:
        name, launch_bytes, t_dom = fwd_name, alg_bytes, t_fwd
    roof = roofline(a, wl, name, launch_bytes, t_dom)
    out = {
        "metric": metric,
        "value": round(2 * S * a.steps / wall / 1e9, 3),
        "unit": unit,
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": "synthetic (modified Shepp-Logan phantom)",
        "config": {"workload": wl["text"], "projector": a.projector, "kernels": a.kernels, "tile": tile,
                   "pose_batch": op.pose_batch() if a.kernels == "fast" and a.projector in ("joseph", "siddon")
                   else None,
                   "volume": list(wl["vol"]), "poses": wl["views"],
                   "detector": list(wl["det"]), "samples_per_pass": S, "rays": R,
                   "fwd_ms": round(t_fwd * 1e3, 3), "adj_ms": round(t_adj * 1e3, 3)},
        "roofline": dict(roof, launches={"forward": 1, "back_projection": len(adj_launches) if len(adj_launches) > 1
                                                                             else "pass"}),
    }
    if a.cg_iters > 0 and a.projector == "joseph" and a.dtype == "f32":
        out["cg"] = cg_run(E, op, x, a.cg_iters, vd, dd, a.cg_target, a.cg_ref_iters)
    if not a.no_cpu and a.projector == "joseph" and a.dtype == "f32":
        out["cpu_baseline"] = cpu_baseline(wl, *a.cpu_poses)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
