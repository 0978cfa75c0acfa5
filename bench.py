"""Benchmark: seconds per 256^3 reconstruction (rel-L2 <= 1e-3 vs FP64) on one MI355X.

Workload: swirl-only Burgers vortex pressure on a 256^3 grid (L = 1, core
radius 0.1 L, axis z through the centre, Gamma = 1) with 2 % Gaussian noise
on the gradient; 162 icosahedral directions (twice subdivided), line spacing
1 voxel; plain method; gradient and trapezoid terms in FP32, added into FP64
accumulators by the tile engine, a ray's partial sums stored rounded once to
FP32 (the compensated FP32 mode: OMNI3D_FP32_KAHAN),
averaging, source update and residual in FP64; stop at a relative change
< 1e-6 of the source voxels or 200 iterations.  The field is resident on the GPU before
the timed region; a step is one reconstruction of that snapshot with the
geometry plan (rays, visit counts, source CSR, tile segments) of the grid
already built, as in the reference's loop over the snapshots of one grid.
The plan build time and the cold total are reported beside it.

Also measured in the same run: the FP64 GPU path on the same input (rel-L2
of the FP32 result against it, speed ratio), the cost of one reference-
schedule FP64 pass (a full traversal per iteration, as the reference
iterates), the per-phase times from the library's HIP events on its stream,
the roofline of the dominant kernel, the source iteration (with the two
ray traversals beside it: their counter traffic over their time), the
Poisson path (GPU red-black SOR, the CPU solvers), the trilinear traversal's accuracy against the
analytic pressure beside the staircase's, and the oracle's single-thread CPU
time on a bounded sample (cpu_baseline; --cpu-full times a whole pass).

Other modes (not the headline): --copy-calibration runs a plain 1 GiB device
copy for scaling the FETCH_SIZE / WRITE_SIZE counters; --batch times the
64-snapshot uneven-grid series (one hipGraph vs one launch sequence per
snapshot); --adjoint times forward + backward of the 128^3 Fourier-sphere
reconstruction with the plain method (source-restricted forward and
backward), and once with the plain-mask method (a traversal per iteration);
--methods times one reference-schedule pass (a full traversal) of each
integration method on the 128^3 masked Fourier-sphere field; --sor times the
GPU red-black SOR (omega 1.8) on the 256^3 Neumann right-hand side of the
Burgers-vortex field, beside the oracle's serial red-black SOR; --pairs times
the README's boundary-pair pipeline at 256^3 (pairs, 4 boundary iterations,
the inner pass); --threshold times the curl-histogram threshold at 256^3.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-3d-pressure-reconstruction_amd"))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_tile_passes_burgers256.json")


def host_info():
    """CPU model, threads visible and the compiler the oracle is built with."""
    import platform
    import subprocess
    model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                            timeout=20).stdout.splitlines()[0]
    except Exception:
        cc = "unknown"
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return {"cpu_model": model, "nproc": ncpu, "compiler": f"{cc} -O2 -ffp-contract=off"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration():
    """A plain device copy of known size (1 GiB read + 1 GiB written) whose
    FETCH_SIZE / WRITE_SIZE readings scale the counters of the kernels."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def batch_mode(args):
    import torch
    import omni3d
    from omni3d import fields as F
    series = F.snapshot_series(64, seed=0)
    shapes = [s["shape"] for s in series]
    dirs = omni3d.generate_directions(3)
    h = 2 * np.pi / 64
    with omni3d.Batch(shapes, (h, h, h), precision=omni3d.FP32_KAHAN,
                      max_iterations=args.max_iter, tolerance=args.tol) as b:
        b.set_directions(dirs)
        for i, s in enumerate(series):
            f = F.taylor_green(s["shape"][2], noise=s["noise"], seed=s["seed"], shape=s["shape"],
                               phase=s["phase"])
            b.set_field(i, f["acc"], dtype=np.float32)
        t0 = time.perf_counter()
        picks = b.tune()
        tune_s = time.perf_counter() - t0
        log(f"batch: tuned {picks} in {tune_s:.2f}s")
        b.run(use_graph=False)
        seq = [b.run(use_graph=False) for _ in range(args.steps)]
        b.run(use_graph=True)
        gr = [b.run(use_graph=True) for _ in range(args.steps)]
        its = [b.result(i)[1].iterations for i in range(len(shapes))]
        vox = int(sum(np.prod(s) for s in shapes))
    ms_seq, ms_graph = float(np.mean(seq)), float(np.mean(gr))
    print(json.dumps({
        "metric": "snapshots per second, 64 uneven snapshots (one hipGraph)",
        "value": round(64 / (ms_graph * 1e-3), 3), "unit": "snapshots/s", "n_gpus": 1,
        "steps": args.steps, "ms_per_step": round(ms_graph, 3), "higher_is_better": True,
        "dtype": "f32 gradient, trapezoid terms and stored ray partial sums; f64 accumulation, iteration and averaging",
        "data": "synthetic",
        "config": {"workload": "taylor-green series, 64 snapshots from {64^3, 96^3, 128x96x64, "
                               "160x128x96}, 0-5% noise, 162 directions, plain affine",
                   "voxels": vox, "tile_choices": picks, "tune_s": round(tune_s, 3),
                   "ms_sequential": round(ms_seq, 3), "speedup_graph_vs_sequential":
                   round(ms_seq / ms_graph, 3), "iterations_min_max": [min(its), max(its)]}}))


def adjoint_mode(args):
    import torch
    import omni3d
    from omni3d import fields as F
    n = 128
    f = F.fourier_sphere(n, seed=0)
    V = n ** 3
    dirs = omni3d.generate_directions(3)
    g = torch.from_numpy(np.concatenate([c.reshape(-1) for c in f["grad"]]).astype(np.float32)).cuda()
    p = torch.empty(V, dtype=torch.float32, device="cuda")
    y = torch.randn(V, dtype=torch.float32, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(0))
    dg = torch.empty(3 * V, dtype=torch.float32, device="cuda")

    def timed(method, steps):
        ctx = omni3d.Context((n, n, n), f["delta"], method=method,
                             precision=omni3d.FP32_KAHAN, max_iterations=args.max_iter)
        ctx.set_directions(dirs)
        ctx.set_field([np.zeros((n, n, n))] * 3, mask=f["mask"])
        ctx.forward_device(g.data_ptr(), p.data_ptr())
        ctx.backward_device(y.data_ptr(), dg.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.forward_device(g.data_ptr(), p.data_ptr())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            ctx.backward_device(y.data_ptr(), dg.data_ptr())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ctx.close()
        return (t1 - t0) / steps, (t2 - t1) / steps

    fw, bw = timed(omni3d.PLAIN, args.steps)
    mfw, mbw = timed(omni3d.PLAIN_MASK, 1)
    print(json.dumps({
        "metric": "s per forward+backward, 128^3 Fourier sphere, plain method",
        "value": round(fw + bw, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": False, "dtype": "fp32-kahan forward, fp64 adjoint", "data": "synthetic",
        "config": {"workload": f"fourier-sphere 128^3, 162 directions, {args.max_iter} iterations",
                   "forward_s": round(fw, 5), "backward_s": round(bw, 5),
                   "plain_mask_forward_s": round(mfw, 5), "plain_mask_backward_s": round(mbw, 5)}}))


def methods_mode(args):
    """One full traversal per iteration (the reference's schedule) for each
    integration method, 128^3 masked Fourier sphere, 162 directions, FP32-Kahan."""
    import torch
    import omni3d
    from omni3d import fields as F
    n = 128
    shape = (n, n, n)
    V = n ** 3
    f = F.fourier_sphere(n, seed=0)
    dirs = omni3d.generate_directions(3)
    D = dirs.shape[0]
    bytes_pass = D * V * 3 * 4 + V * 20
    p_dev = torch.zeros(V, dtype=torch.float64, device="cuda")
    passes = args.naive_passes
    out = {}
    thr = None
    # the two reference variants walk one lane per ray (their step rules
    # leave the plan's staircase)
    for name, method in (("plain", omni3d.PLAIN), ("plain-mask", omni3d.PLAIN_MASK),
                         ("select", omni3d.SELECT), ("weighted", omni3d.WEIGHTED),
                         ("select-fixedbc", omni3d.SELECT_FIXEDBC),
                         ("minicurl", omni3d.MINICURL)):
        kw = {}
        if method in (omni3d.SELECT, omni3d.SELECT_FIXEDBC):
            kw["threshold"] = thr
        ctx = omni3d.Context(shape, f["delta"], density=f["rho"], precision=omni3d.FP32_KAHAN,
                             method=method, solver=omni3d.NAIVE, max_iterations=passes,
                             tolerance=0.0, **kw)
        ctx.set_directions(dirs)
        ctx.set_field(f["acc"], mask=f["mask"])
        if thr is None:
            # Select threshold: the 80th percentile of |curl a| over the fluid
            c = ctx.field_curl()
            thr = float(np.quantile(c[f["mask"] > 0], 0.8))
        ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)
        ms, setup = [], []
        for _ in range(max(1, args.steps)):
            r = ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)
            ms.append(r.ms_iterate / passes)
            setup.append(r.ms_segment_sums)
        torch.cuda.synchronize()
        ctx.close()
        m = float(np.mean(ms))
        out[name] = {"ms_per_pass": round(m, 3),
                     "frac": round(bytes_pass / (m * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        # the per-field replay of the path-break rule (k_mask_walk and the
        # setup tile pass), once per reconstruction
        out[name]["ms_rule_replay_per_field"] = round(float(np.mean(setup)), 3)
        log(f"bench: {name} {m:.2f} ms per pass")
    # whole reconstructions of the masked methods: the affine schedule (the
    # iteration on the run-start voxels, then one full pass) against the
    # reference schedule (a full traversal per iteration), same field
    # (the reference's variants whose walks leave the staircase too: their
    # affine schedule walks one lane per ray)
    recon = {}
    for name, method in (("plain-mask", omni3d.PLAIN_MASK), ("select", omni3d.SELECT),
                         ("weighted", omni3d.WEIGHTED), ("select-fixedbc", omni3d.SELECT_FIXEDBC),
                         ("minicurl", omni3d.MINICURL),
                         ("weighted-fixedbc", omni3d.WEIGHTED_FIXEDBC),
                         ("weighted-minicurl", omni3d.WEIGHTED_MINICURL)):
        kw = {"threshold": thr} if method in (omni3d.SELECT, omni3d.SELECT_FIXEDBC) else {}
        row = {}
        for iters in (10, 200):
            for solver, sname in ((omni3d.AFFINE, "affine"), (omni3d.NAIVE, "reference")):
                ctx = omni3d.Context(shape, f["delta"], density=f["rho"],
                                     precision=omni3d.FP32_KAHAN, method=method, solver=solver,
                                     max_iterations=iters, tolerance=0.0, **kw)
                ctx.set_directions(dirs)
                ctx.set_field(f["acc"], mask=f["mask"])
                ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)
                torch.cuda.synchronize()
                wall = time.perf_counter() - t0
                ctx.close()
                e = {"ms_solve": round(r.ms_solve, 3), "ms_wall": round(wall * 1e3, 3)}
                if solver == omni3d.AFFINE:
                    e.update({"ms_rows_and_b": round(r.ms_pass_a, 3),
                              "ms_iterate": round(r.ms_iterate, 3),
                              "ms_final_pass": round(r.ms_pass_b, 3),
                              "run_starts": int(r.sources), "row_entries": int(r.source_visits)})
                row[f"{sname}_K{iters}"] = e
                log(f"bench: {name} {sname} K={iters}: {r.ms_solve:.2f} ms")
        row["speedup_K10"] = round(row["reference_K10"]["ms_solve"] / row["affine_K10"]["ms_solve"], 2)
        row["speedup_K200"] = round(row["reference_K200"]["ms_solve"] /
                                    row["affine_K200"]["ms_solve"], 2)
        recon[name] = row
    first = out["select"]
    cpu = None
    if not args.no_cpu:
        # the oracle's serial Select pass (the reference's shipped method) over
        # a sample of the directions, scaled to one pass
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        with omni3d.Context(shape, f["delta"], density=f["rho"]) as cc:
            cc.set_field(f["acc"], mask=f["mask"])
            curl = cc.field_curl()
        sub = dirs[:: max(1, D // args.cpu_dirs)][: args.cpu_dirs]
        t0 = time.perf_counter()
        O.one_pass(shape, sub, 1.0, f["delta"], f["rho"], f["acc"], np.zeros(shape), curl=curl,
                   threshold=thr, method=O.SELECT, faithful=1)
        ts = time.perf_counter() - t0
        cpu = {"value": round(ts * D / len(sub) * 1e3, 1), "unit": "ms per pass", "cores": 1,
               "kind": "port",
               "sample": f"oracle FP64 serial Select pass over {len(sub)} of {D} directions at "
                         f"{n}^3 took {ts:.2f}s, scaled x{D / len(sub):.1f} to one pass"}
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_masked_fourier128.json")
    if os.path.exists(tf):
        t = json.load(open(tf))
        if t.get("workload_n") == n and t.get("directions") == D:
            traffic = t.get("traffic_bytes", {}).get("select")
    print(json.dumps({
        "metric": "ms per reference-schedule pass, 128^3 masked Fourier sphere",
        "value": first["ms_per_pass"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "higher_is_better": False, "dtype": "fp32-kahan ray sums, fp64 averaging",
        "data": "synthetic",
        "roofline": {"bound": "hbm", "achieved": round(bytes_pass / (first["ms_per_pass"] * 1e-3)
                                                       / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": first["frac"], "traffic": traffic,
                     "kernel": "k_mask_carry + k_tile_masked<float, select, 16^3, 1024> "
                               "(one iteration)", "algorithmic_bytes": bytes_pass},
        "cpu_baseline": cpu,
        "config": {"workload": f"fourier-sphere 128^3 with the sphere masked, {D} directions, "
                               f"one full traversal per iteration, {passes} iterations timed; "
                               f"tile engine (rule replay once per field, carry chain + tile pass per iteration)",
                   "select_threshold": thr, "methods": out,
                   "reconstructions": recon}}))


SOR_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_sor256.json")


def sor_mode(args):
    """GPU red-black SOR on device buffers: 256^3, omega 1.8, fixed sweeps."""
    print(json.dumps(sor_measure(args)))


def sor_measure(args):
    """The Poisson path's numbers (the --sor line; a summary rides on the
    default line as "poisson"): GPU red-black SOR per iteration with its
    roofline, the oracle's serial red-black SOR and its restatement of the
    reference's CPU solver sor3D (damped Jacobi, weight 0.8) per sweep on
    the same right-hand side, and sor3D's drop-in on host buffers."""
    import torch
    import omni3d
    from omni3d import fields as F
    n = 256
    shape = (n, n, n)
    V = n ** 3
    f = F.burgers_vortex(n, noise=0.02, seed=0)
    rhs = omni3d.calrhs(shape, f["delta"], f["rho"], f["acc"])
    rhs -= rhs.mean()
    rd = torch.from_numpy(rhs).cuda()
    pd = torch.zeros(V, dtype=torch.float64, device="cuda")
    iters = args.sor_iters
    omni3d.sor_solve_device(shape, rd.data_ptr(), pd.data_ptr(), omega=1.8, max_iter=5)
    pd.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it, res = omni3d.sor_solve_device(shape, rd.data_ptr(), pd.data_ptr(), omega=1.8,
                                      max_iter=iters)
    s_total = time.perf_counter() - t0
    ms_it = s_total * 1e3 / iters
    # one iteration (red then black half sweep) must read p and rhs and write
    # p once: 3 x 8 B per voxel
    bytes_it = 24 * V
    achieved = bytes_it / (ms_it * 1e-3) / 1e9
    traffic = None
    if os.path.exists(SOR_TRAFFIC_FILE):
        t = json.load(open(SOR_TRAFFIC_FILE))
        if t.get("workload_n") == n:
            traffic = t.get("traffic_bytes_per_iteration")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        k = 4
        t0 = time.perf_counter()
        O.sor_redblack(np.zeros(shape), rhs, k, 1.8)
        ts = time.perf_counter() - t0
        cpu = {"value": round(ts / k * 1e3, 2), "unit": "ms per iteration", "cores": 1,
               "kind": "port",
               "sample": f"oracle serial red-black SOR, {k} iterations at {n}^3 took {ts:.2f}s",
               **host_info()}
        # the reference's own CPU Poisson solver, sor3D.cpp: damped Jacobi
        # (weight 0.8, sor3D.cpp:7,739) with its 26 boundary stencils
        kj = 4
        t0 = time.perf_counter()
        O.sor3d(np.zeros(shape), rhs, kj, 0.0)
        tj = time.perf_counter() - t0
        cpu["reference_jacobi_w0.8"] = {
            "value": round(tj / kj * 1e3, 2), "unit": "ms per sweep", "cores": 1,
            "sample": f"oracle restatement of sor3D (damped Jacobi, weight 0.8), {kj} sweeps "
                      f"at {n}^3 took {tj:.2f}s"}
    # sor3D's drop-in on the GPU (same signature, host buffers: the PCIe copies
    # of p, pn and the right-hand side are included)
    kg = 40
    t0 = time.perf_counter()
    _, _, sw, _ = omni3d.sor3d(np.zeros(shape), rhs, kg, 0.0)
    tg = time.perf_counter() - t0
    jacobi_gpu = {"sweeps": sw, "s_total_host_buffers": round(tg, 4),
                  "note": "sor3D signature on the GPU: damped Jacobi, weight 0.8, host buffers"}
    return {
        "metric": "ms per red-black SOR iteration, 256^3", "value": round(ms_it, 4), "unit": "ms",
        "n_gpus": 1, "steps": iters, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic",
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_redblack_multi<64, 16, 64, 2, 1024> (two red-black iterations per pass)",
                     "algorithmic_bytes": bytes_it},
        "cpu_baseline": cpu, "gpu_sor3D_dropin": jacobi_gpu,
        "config": {"workload": f"Neumann right-hand side of the burgers-vortex {n}^3 field, "
                               f"omega 1.8, {iters} iterations from p = 0",
                   "residual": res, "iterations": it}}


def pairs_mode(args):
    """Boundary-pair pipeline (README:18-28) at 256^3, 162 directions, device
    buffers: per-ray entry-to-exit sums, pair list, 4 boundary iterations,
    then the inner pass (one plain iteration from that boundary)."""
    import torch
    import omni3d
    from omni3d import fields as F
    n = args.n
    shape = (n, n, n)
    V = n ** 3
    f = F.burgers_vortex(n, noise=0.02, seed=0)
    dirs = omni3d.generate_directions(args.divide)
    out = {}
    p_dev = torch.zeros(V, dtype=torch.float64, device="cuda")
    for name, prec in (("fp32_kahan", omni3d.FP32_KAHAN), ("fp64", omni3d.FP64)):
        with omni3d.Context(shape, f["delta"], density=f["rho"], line_spacing=1.0,
                            precision=prec, max_iterations=1, seed_pn_with_p=0) as ctx:
            ctx.set_directions(dirs)
            ctx.set_field(f["acc"])
            npairs = C.c_longlong(0)

            def pairs():
                p_dev.zero_()
                omni3d.check(omni3d.lib().omni3d_boundary_pairs(
                    ctx._h, 4, C.c_void_p(p_dev.data_ptr()), omni3d.F64, omni3d.DEVICE,
                    C.byref(npairs)))
            pairs()  # plan build
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                pairs()
            torch.cuda.synchronize()
            ms_pairs = (time.perf_counter() - t0) * 1e3 / args.steps
            t0 = time.perf_counter()
            for _ in range(args.steps):
                pairs()
                ctx.reconstruct_device(p_dev.data_ptr(), None, omni3d.F64)
            torch.cuda.synchronize()
            ms_all = (time.perf_counter() - t0) * 1e3 / args.steps
            p = p_dev.view(shape).cpu().numpy()
            a = f["p"] - f["p"].mean()
            err = float(np.linalg.norm((p - p.mean()) - a) / np.linalg.norm(a))
            out[name] = {"ms_pairs_and_4_boundary_iterations": round(ms_pairs, 3),
                         "ms_readme_pipeline": round(ms_all, 3), "pairs": npairs.value,
                         "rel_l2_vs_analytic_demeaned": err}
            log(f"pairs {name}: {out[name]}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        k = 2
        t0 = time.perf_counter()
        O.boundary_pairs(shape, dirs[:k], 1.0, f["delta"], f["rho"], f["acc"], np.zeros(shape), 4)
        ts = time.perf_counter() - t0
        cpu = {"value": round(ts * dirs.shape[0] / k, 1), "unit": "s", "cores": 1, "kind": "port",
               "sample": f"oracle boundary pairs + 4 iterations over {k} of {dirs.shape[0]} "
                         f"directions at {n}^3 took {ts:.2f}s, scaled by directions"}
    print(json.dumps({
        "metric": "ms per boundary-pair reconstruction (README pipeline), 256^3",
        "value": out["fp32_kahan"]["ms_readme_pipeline"], "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "higher_is_better": False, "dtype": "fp32-kahan ray sums, f64 pairs",
        "data": "synthetic", "results": out, "cpu_baseline": cpu,
        "config": {"workload": f"burgers-vortex {n}^3, 2% noise, {dirs.shape[0]} directions, "
                               "4 boundary iterations + one plain inner pass"}}))


def threshold_mode(args):
    """Curl-histogram threshold (kernel.cu:8380-8418) at 256^3 on a device
    FP32 curl field: min/max pass, histogram pass, top-down walk."""
    import torch
    import omni3d
    from omni3d import fields as F
    n = args.n
    shape = (n, n, n)
    V = n ** 3
    f = F.burgers_vortex(n, noise=0.02, seed=0)
    with omni3d.Context(shape, f["delta"], density=f["rho"]) as ctx:
        ctx.set_field(f["acc"])
        curl = ctx.field_curl().astype(np.float32)
    cd = torch.from_numpy(curl).cuda()
    out = torch.empty(1, dtype=torch.float32, device="cuda")
    L = omni3d.lib()

    def call():
        omni3d.check(L.omni3d_threshold_hist(n, n, n, C.c_void_p(cd.data_ptr()), 0.05,
                                             C.c_void_p(out.data_ptr()), omni3d.DEVICE, None))
    call()
    torch.cuda.synchronize()
    k = max(args.steps, 20)
    t0 = time.perf_counter()
    for _ in range(k):
        call()
    us = (time.perf_counter() - t0) * 1e6 / k
    thr = float(out.item())
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        want = O.threshold_hist(curl, 0.05)
        ts = time.perf_counter() - t0
        assert want == thr, (want, thr)
        cpu = {"value": round(ts * 1e6, 1), "unit": "us per call", "cores": 1, "kind": "port",
               "sample": f"oracle threshold histogram at {n}^3 (one call, {ts:.3f}s)"}
    bytes_call = 8 * V  # the curl read twice (min/max, then the histogram)
    achieved = bytes_call / (us * 1e-6) / 1e9
    print(json.dumps({
        "metric": "us per curl-histogram threshold, 256^3", "value": round(us, 1), "unit": "us",
        "n_gpus": 1, "steps": k, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "threshold": thr,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole call (k_thr_minmax + k_thr_hist + k_thr_pick, host sync)",
                     "algorithmic_bytes": bytes_call},
        "cpu_baseline": cpu,
        "config": {"workload": f"|curl a| of the burgers-vortex {n}^3 field, percentage 0.05"}}))


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
    ap.add_argument("--cpu-dirs", type=int, default=18)
    ap.add_argument("--cpu-full", action="store_true",
                    help="time the oracle's whole pass (every direction) instead of a sample")
    ap.add_argument("--no-trilinear", action="store_true")
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--naive-passes", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--batch", action="store_true")
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--methods", action="store_true")
    ap.add_argument("--sor", action="store_true")
    ap.add_argument("--sor-iters", type=int, default=200)
    ap.add_argument("--pairs", action="store_true")
    ap.add_argument("--threshold", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(directions would shard across GPUs; not built); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration:
        return copy_calibration()
    if args.batch:
        return batch_mode(args)
    if args.adjoint:
        return adjoint_mode(args)
    if args.methods:
        return methods_mode(args)
    if args.sor:
        return sor_mode(args)
    if args.pairs:
        return pairs_mode(args)
    if args.threshold:
        return threshold_mode(args)

    import omni3d
    from omni3d import fields as F

    n = args.n
    shape = (n, n, n)
    V = n ** 3
    log(f"bench: generating Burgers vortex {n}^3")
    f = F.burgers_vortex(n, noise=0.02, seed=0)
    dirs = omni3d.generate_directions(args.divide)
    D = dirs.shape[0]

    def make_ctx(precision, **kw):
        ctx = omni3d.Context(shape, f["delta"], density=f["rho"], line_spacing=1.0,
                             precision=precision, max_iterations=args.max_iter,
                             tolerance=args.tol, seed_pn_with_p=1, **kw)
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
    ctx32.close()

    # FP64 GPU paths on the same input
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
        ctx64.close()
        # the reference's schedule in FP64: a full traversal per iteration;
        # a few passes are timed and the per-pass cost is reported
        ctxn = omni3d.Context(shape, f["delta"], density=f["rho"], precision=omni3d.FP64,
                              solver=omni3d.NAIVE, max_iterations=args.naive_passes)
        ctxn.set_directions(dirs)
        ctxn.set_field(f["acc"])
        step(ctxn)
        rn = step(ctxn)
        torch.cuda.synchronize()
        ms_naive = rn.ms_iterate / args.naive_passes
        # the same schedule walked one lane per ray with global atomics (the
        # reference kernels' organisation)
        os.environ["OMNI3D_MASKED"] = "naive"
        rl = step(ctxn)
        torch.cuda.synchronize()
        del os.environ["OMNI3D_MASKED"]
        ms_lanes = rl.ms_iterate / args.naive_passes
        ctxn.close()
        fp64 = {"s_per_recon_fp64": round(s64, 6), "rel_l2_vs_fp64": rel,
                "max_err_rel_vs_fp64": maxerr, "speedup_vs_fp64": round(s64 / s_per, 3),
                "fp64_iterations": r64.iterations, "ms_pass_b_fp64": round(r64.ms_pass_b, 4),
                "fp64_reference_schedule_ms_per_pass": round(ms_naive, 3),
                "fp64_reference_schedule_lane_walk_ms_per_pass": round(ms_lanes, 3),
                "fp64_reference_schedule_s_for_same_iterations":
                    round(ms_naive * 1e-3 * r.iterations, 3),
                "speedup_vs_fp64_reference_schedule":
                    round(ms_naive * 1e-3 * r.iterations / s_per, 2)}
        log(f"bench: FP64 {s64 * 1e3:.2f} ms/recon, FP32 rel-L2 vs FP64 {rel:.3e}; "
            f"FP64 full-traversal pass {ms_naive:.2f} ms")
        assert rel <= 1e-3, f"FP32-Kahan result misses rel-L2 <= 1e-3 vs FP64 ({rel})"

    pa = np.asarray(f["p"]).reshape(-1)
    p32h = p32.cpu().numpy()
    err_analytic = F.rel_l2(F.demean(p32h), F.demean(pa))

    # the trilinear traversal (straight-ray integral, the accuracy mode) on the
    # same input and schedule, beside the staircase's error
    tri = {}
    if not args.no_trilinear:
        log("bench: trilinear traversal")
        ctxt = make_ctx(omni3d.FP32_KAHAN, traversal=omni3d.TRILINEAR)
        step(ctxt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rt = step(ctxt)
        torch.cuda.synchronize()
        st_tri = time.perf_counter() - t0
        ptri = p_dev.cpu().numpy()
        ctxt.close()
        tri = {"s_per_recon": round(st_tri, 4), "iterations": rt.iterations,
               "ms_pass_a": round(rt.ms_pass_a, 3), "ms_pass_b": round(rt.ms_pass_b, 3),
               "rel_l2_vs_analytic_demeaned": F.rel_l2(F.demean(ptri), F.demean(pa)),
               "staircase_rel_l2_vs_analytic_demeaned": err_analytic}
        log(f"bench: trilinear {st_tri * 1e3:.1f} ms, rel-L2 vs analytic "
            f"{tri['rel_l2_vs_analytic_demeaned']:.3e} (staircase {err_analytic:.3e})")

    # roofline of the dominant traversal kernel: the slower of the segment-sum
    # traversal of pass A (k_tile_pass, local-sum mode) and the final full
    # traversal (k_tile_pass, full mode), each one ray pass.  A ray pass must
    # stream the gradient once per direction, D*V*3*4 B in FP32, and read and
    # write the FP64 pressure and FP32 weight once, V*20 B.  The kernels'
    # durations are the library's HIP events around them on its stream.
    ms_b = float(np.mean([x.ms_pass_b for x in reps]))
    ms_s = float(np.mean([x.ms_segment_sums for x in reps]))
    bytes_pass = D * V * 3 * 4 + V * 20
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        t = json.load(open(TRAFFIC_FILE))
        if t.get("workload_n") == n and t.get("directions") == D:
            traffic = t.get("traffic_bytes", {})
    kern = {"local_sum": ("k_tile_pass<float, local-sum, 16^3 tiles + halo, 1024 threads> "
                          "(segment sums of pass A)", ms_s),
            "full": ("k_tile_pass<float, full, 16^3 tiles, 1024 threads> (final traversal)", ms_b)}
    # the source iteration is the dominant kernel by time (200 launches, half
    # of a reconstruction): one launch must read the iteration matrix once,
    # 4 B per word (equal entries of a row merged), and per source row its
    # vectors: row sum, p_S, constant, diagonal, source index and visit count
    # read, row sum, p_S and FP32 increment written, the increment gathered
    # (60 B).  Its duration is the iteration phase's HIP events / launches.
    it_ms = float(np.mean([x.ms_iterate for x in reps])) / max(1, reps[-1].iterations)
    it_bytes = reps[-1].matrix_words * 4 + reps[-1].sources * 60
    it_ach = it_bytes / (it_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(it_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(it_ach / HBM_PEAK_GBS, 4), "traffic": traffic.get("iteration"),
            "kernel": "k_iterate_grp<8, increments> (one source iteration)",
            "algorithmic_bytes": it_bytes, "ms": round(it_ms, 5),
            "matrix_words": reps[-1].matrix_words, "source_visits": reps[-1].source_visits,
            # the two ray traversals: the bytes they move (counter traffic per
            # launch) over their time against the HBM peak, and beside it the
            # per-direction model D*V*12 + 20V (above 1: the tile engine
            # streams each tile's gradient once for every direction crossing it)
            "traversals": {
                k: {"kernel": kern[k][0], "ms": round(kern[k][1], 4),
                    "traffic": traffic.get(k),
                    "achieved": (round(traffic[k] / (kern[k][1] * 1e-3) / 1e9, 1)
                                 if traffic.get(k) else None),
                    "frac": (round(traffic[k] / (kern[k][1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             if traffic.get(k) else None),
                    "model_bytes": bytes_pass,
                    "model_frac": round(bytes_pass / (kern[k][1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                for k in kern}}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sub = dirs if args.cpu_full else dirs[:: max(1, D // args.cpu_dirs)][: args.cpu_dirs]
        log(f"bench: CPU oracle, {len(sub)} of {D} directions, one pass at {n}^3")
        t0 = time.perf_counter()
        O.one_pass(shape, sub, 1.0, f["delta"], f["rho"], f["acc"], np.zeros(shape))
        ts = time.perf_counter() - t0
        per_pass = ts * D / len(sub)
        iters = reps[-1].iterations
        what = (f"oracle FP64 serial pass over all {D} directions at {n}^3 took {ts:.2f}s"
                if len(sub) == D else
                f"oracle FP64 serial pass over {len(sub)} of {D} directions (every "
                f"{D // len(sub)}th) at {n}^3 took {ts:.2f}s; scaled x{D / len(sub):.1f} to one "
                f"pass ({per_pass:.1f}s)")
        cpu = {"value": round(per_pass * iters, 2), "unit": "s per reconstruction", "cores": 1,
               "kind": "port",
               "sample": what + f", x{iters} passes (the reference's CPU loop runs one full pass "
                                f"per iteration)",
               "measured_s": round(ts, 3), "s_per_pass": round(per_pass, 3),
               "directions_timed": len(sub), **host_info()}

    # the Poisson path beside it: GPU red-black SOR and the CPU solvers
    poisson = None
    if not args.no_poisson:
        log("bench: Poisson path (red-black SOR, CPU red-black SOR and sor3D)")
        sm = sor_measure(args)
        c = sm.get("cpu_baseline") or {}
        poisson = {
            "workload": sm["config"]["workload"],
            "gpu_ms_per_sor_iteration": sm["value"],
            "roofline_frac": sm["roofline"]["frac"],
            "cpu_red_black_sor_ms_per_iteration": c.get("value"),
            "cpu_reference_jacobi_w0.8_ms_per_sweep":
                (c.get("reference_jacobi_w0.8") or {}).get("value"),
            "cpu_cores": c.get("cores"), "cpu_kind": c.get("kind"),
            "gpu_sor3D_dropin": sm["gpu_sor3D_dropin"],
        }

    out = {
        "metric": "s per 256^3 reconstruction at rel-L2<=1e-3 vs FP64 oracle; % HBM roofline",
        "value": round(s_per, 6), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(s_per * 1e3, 4),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32 gradient, trapezoid terms and stored ray partial sums; f64 accumulation, iteration and averaging",
        "data": "synthetic",
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
                   "ms_segment_sums": round(ms_s, 4),
                   "ms_iterate": round(float(np.mean([x.ms_iterate for x in reps])), 4),
                   "ms_pass_b": round(ms_b, 4),
                   "ms_finalize": round(float(np.mean([x.ms_finalize for x in reps])), 4),
                   "rel_l2_vs_analytic_demeaned": err_analytic, **fp64},
        # one snapshot on a fresh context: plan build + first reconstruction
        "single_snapshot_s": round(cold_s, 4),
        "fp32_vs_fp64_target": {
            "target": 5.0, "measured": fp64.get("speedup_vs_fp64"),
            "met": bool(fp64) and fp64.get("speedup_vs_fp64", 0) >= 5.0,
            "note": "both precisions run the same affine schedule: the FP64 source iteration "
                    "(fixed-point row sums, the dominant kernel) is shared, and the tile "
                    "engine's traversals are bound by LDS and instruction issue, not by FP64 "
                    "arithmetic; against the reference's schedule in FP64 (a full traversal "
                    "per iteration) the ratio is speedup_vs_fp64_reference_schedule"},
        "trilinear": tri,
        "poisson": poisson,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
