"""Benchmark: seconds per 256^3 reconstruction (rel-L2 <= 1e-3 vs FP64) on one MI355X.

Workload: swirl-only Burgers vortex pressure on a 256^3 grid (L = 1, core
radius 0.1 L, axis z through the centre, Gamma = 1) with 2 % Gaussian noise
on the gradient; 162 icosahedral directions (twice subdivided), line spacing
1 voxel; plain method; ray sums in FP32 with Kahan compensation, averaging,
source update and residual in FP64; stop at a relative change < 1e-6 of the
source voxels or 200 iterations.  The field is resident on the GPU before
the timed region; a step is one reconstruction of that snapshot with the
geometry plan (rays, visit counts, source CSR) of the grid already built,
as in the reference's batch loop over snapshots of one grid.  The plan build
time and the cold total are reported beside it.

Also measured in the same run: the FP64 GPU path (rel-L2 of the FP32 result
against it, speed ratio), the per-phase times from the library's HIP events
on its stream, the roofline of the dominant kernel (the final traversal), and
the oracle's single-thread CPU time on a bounded sample (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-3d-pressure-reconstruction_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--divide", type=int, default=3)
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-dirs", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(direction sharding is described in DESIGN.md); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import omni3d
    from omni3d import fields as F

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    shape = (n, n, n)
    V = n ** 3
    log(f"bench: generating Burgers vortex {n}^3")
    f = F.burgers_vortex(n, noise=0.02, seed=0)
    dirs = omni3d.generate_directions(args.divide)
    D = dirs.shape[0]

    def make_ctx(precision):
        ctx = omni3d.Context(shape, f["delta"], density=f["rho"], line_spacing=1.0,
                             precision=precision, max_iterations=args.max_iter,
                             tolerance=args.tol, seed_pn_with_p=1)
        ctx.set_directions(dirs)
        ctx.set_field(f["acc"])
        return ctx

    p_dev = torch.zeros(V, dtype=torch.float64, device="cuda")

    def step(ctx):
        p_dev.zero_()
        return ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)

    # FP32-Kahan: cold (plan build) then warm steps
    log("bench: FP32-Kahan context")
    ctx32 = make_ctx(omni3d.FP32_KAHAN)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep_cold = step(ctx32)
    torch.cuda.synchronize()
    cold_s = time.perf_counter() - t0
    log(f"bench: cold {cold_s:.3f}s plan {rep_cold.ms_plan:.1f}ms rays {rep_cold.rays} "
        f"steps {rep_cold.steps} sources {rep_cold.sources} iters {rep_cold.iterations}")
    for _ in range(max(0, args.warmup - 1)):
        step(ctx32)
    torch.cuda.synchronize()
    reps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reps.append(step(ctx32))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s_per = (t1 - t0) / args.steps
    p32 = p_dev.clone()
    r = reps[-1]
    log(f"bench: FP32-Kahan {s_per * 1e3:.2f} ms/recon (A {r.ms_pass_a:.2f} it {r.ms_iterate:.2f} "
        f"B {r.ms_pass_b:.2f} fin {r.ms_finalize:.2f}) iters {r.iterations} rel {r.rel_change:.3e}")

    # FP64 GPU path on the same input
    fp64 = {}
    if not args.no_fp64:
        log("bench: FP64 context")
        ctx64 = make_ctx(omni3d.FP64)
        step(ctx64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k64 = max(1, min(args.steps, 2))
        for _ in range(k64):
            r64 = step(ctx64)
        torch.cuda.synchronize()
        s64 = (time.perf_counter() - t0) / k64
        p64 = p_dev.clone()
        rel = float(torch.linalg.norm(p32 - p64) / torch.linalg.norm(p64))
        maxerr = float((p32 - p64).abs().max() / p64.abs().max())
        fp64 = {"s_per_recon_fp64": round(s64, 6), "rel_l2_vs_fp64": rel,
                "max_err_rel_vs_fp64": maxerr, "speedup_vs_fp64": round(s64 / s_per, 3),
                "fp64_iterations": r64.iterations, "ms_pass_b_fp64": round(r64.ms_pass_b, 4)}
        log(f"bench: FP64 {s64 * 1e3:.2f} ms/recon, FP32 rel-L2 vs FP64 {rel:.3e}")
        assert rel <= 1e-3, f"FP32-Kahan result misses rel-L2 <= 1e-3 vs FP64 ({rel})"
        ctx64.close()

    pa = np.asarray(f["p"]).reshape(-1)
    p32h = p32.cpu().numpy()
    err_analytic = F.rel_l2(F.demean(p32h), F.demean(pa))

    # roofline of the final traversal: D*V*3*4 B of FP32 gradient + V*20 B
    # (p read+write in FP64, weight read)
    ms_b = float(np.mean([x.ms_pass_b for x in reps]))
    bytes_pass = D * V * 3 * 4 + V * 20
    achieved = bytes_pass / (ms_b * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_walk_plain<float,true> (final traversal)", "ms": round(ms_b, 4)}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sub = dirs[:: max(1, D // args.cpu_dirs)][: args.cpu_dirs]
        log(f"bench: CPU oracle sample, {len(sub)} of {D} directions, one pass at {n}^3")
        t0 = time.perf_counter()
        O.one_pass(shape, sub, 1.0, f["delta"], f["rho"], f["acc"], np.zeros(shape))
        ts = time.perf_counter() - t0
        per_pass = ts * D / len(sub)
        iters = reps[-1].iterations
        cpu = {"value": round(per_pass * iters, 2), "unit": "s per reconstruction", "cores": 1,
               "kind": "port",
               "sample": f"oracle FP64 serial pass over {len(sub)} of {D} directions at {n}^3 "
                         f"took {ts:.2f}s; scaled x{D / len(sub):.1f} to one pass "
                         f"({per_pass:.1f}s) and x{iters} passes (the reference's CPU loop "
                         f"runs one full pass per iteration)"}

    out = {
        "metric": "s per 256^3 reconstruction at rel-L2<=1e-3 vs FP64 oracle; % HBM roofline",
        "value": round(s_per, 6), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(s_per * 1e3, 4),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32-kahan ray sums, fp64 averaging", "data": "synthetic",
        "config": {"workload": f"burgers-vortex {n}^3, 2% noise, {D} icosahedral directions, "
                               f"line spacing 1, plain method, affine solver, tol {args.tol:g} / "
                               f"{args.max_iter} iterations",
                   "grid": [n, n, n], "directions": D, "iterations": reps[-1].iterations,
                   "rel_change": reps[-1].rel_change, "rays": reps[-1].rays,
                   "steps_per_pass": reps[-1].steps, "sources": reps[-1].sources,
                   "source_visits": reps[-1].source_visits, "segments": reps[-1].segments,
                   "sell_cells": reps[-1].sell_cells,
                   "plan_s": round(rep_cold.ms_plan * 1e-3, 4), "cold_s": round(cold_s, 4),
                   "ms_pass_a": round(float(np.mean([x.ms_pass_a for x in reps])), 4),
                   "ms_iterate": round(float(np.mean([x.ms_iterate for x in reps])), 4),
                   "ms_pass_b": round(ms_b, 4),
                   "ms_finalize": round(float(np.mean([x.ms_finalize for x in reps])), 4),
                   "rel_l2_vs_analytic_demeaned": err_analytic, **fp64},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
