"""The masked methods (plain-mask, Select, Weighted) on their affine schedule
(the default solver): the iteration on the run-start voxels S only, then one
full pass.  Against the reference schedule (a full traversal per iteration,
kernel.cu:8840-8861, the NAIVE solver here) and the FP64 oracle on the same
seeded inputs, through the C-ABI.

Why the two agree: the runs, their start voxels and the weights depend on the
curl / mask alone, so each pass maps p on S affinely (masked.hip).  Only the
rounding differs: the reference schedule rounds p[start] + the run's terms per
visit, the affine one rounds the terms (b) and w p[start] apart.

Tolerances: FP64 rel-L2 <= 1e-12 for plain-mask (fixed-point sums, unit
weights), <= 1e-10 for Select / Weighted (FP64 atomics, curl weights), over
~80 passes <= 1e-11 (per-pass rounding differences add up); FP32 ray sums:
<= 1e-7 between the schedules (plain-mask iterates in increment form, its
p_S changes rounded to FP32), <= 1e-5 against the FP64 oracle; iteration
counts equal.
"""
import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu

CASES = [
    (omni3d.PLAIN_MASK, 0, omni3d.FP64),
    (omni3d.PLAIN_MASK, 0, omni3d.FP32_KAHAN),
    (omni3d.SELECT, 1, omni3d.FP64),
    (omni3d.SELECT, 0, omni3d.FP64),
    (omni3d.SELECT, 1, omni3d.FP32_KAHAN),
    (omni3d.WEIGHTED, 0, omni3d.FP64),
    (omni3d.WEIGHTED, 0, omni3d.FP32_KAHAN),
]


def _threshold(shape, f, dtype, precision):
    with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision) as ctx:
        ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], mask=f["mask"], dtype=dtype)
        c = ctx.field_curl()
    return float(np.quantile(c[f["mask"] > 0], 0.7))


def _run(shape, f, dirs, method, faithful, precision, solver, p0, iters, thr, tol=0.0, bits=4):
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision, method=method,
                        solver=solver, max_iterations=iters, tolerance=tol, threshold=thr,
                        faithful_select=faithful) as ctx:
        ctx.set_tile_shape(bits, 1024)
        ctx.set_directions(dirs)
        ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], mask=f["mask"], dtype=dtype)
        return ctx.reconstruct(p0)


@pytest.mark.parametrize("method,faithful,precision", CASES)
def test_affine_masked_matches_reference_schedule(gpu_available, ref_dirs3, method, faithful,
                                                  precision):
    n = 40
    shape = (n, n, n)
    print(f"case affine masked 40^3 method={method} faithful={faithful} precision={precision}",
          flush=True)
    f = F.fourier_sphere(n, seed=6)
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    thr = _threshold(shape, f, dtype, precision) if method == omni3d.SELECT else 0.0
    p0 = np.random.default_rng(7).normal(size=shape)
    pa, wa, ra = _run(shape, f, ref_dirs3, method, faithful, precision, omni3d.AFFINE, p0, 6, thr)
    pn, wn, rn = _run(shape, f, ref_dirs3, method, faithful, precision, omni3d.NAIVE, p0, 6, thr)
    assert ra.iterations == rn.iterations == 6
    # the affine schedule ran: S and its rows are reported
    assert 0 < ra.sources < n ** 3 and ra.source_visits > 0
    if method == omni3d.PLAIN_MASK:
        np.testing.assert_array_equal(wa, wn)
    else:
        assert F.rel_l2(wa, wn) < 1e-12
    err = F.rel_l2(pa, pn)
    print("affine masked vs reference schedule rel-L2", err)
    if precision == omni3d.FP64:
        assert err < (1e-12 if method == omni3d.PLAIN_MASK else 1e-10)
    else:
        # plain-mask in FP32 iterates in increment form (p_S changes rounded
        # to FP32 between iterations), as the plain method does
        assert err < 1e-7


@pytest.mark.parametrize("method,faithful", [(omni3d.PLAIN_MASK, 0), (omni3d.SELECT, 1),
                                             (omni3d.WEIGHTED, 0)])
def test_affine_masked_vs_oracle_seeded(gpu_available, ref_dirs2, method, faithful):
    """Small grid from a random initial pressure, FP64, against the oracle
    (including the first pass's seed term, kernel.cu:8813), 8x8 tiles."""
    shape = (18, 20, 19)
    print(f"case affine masked vs oracle method={method}", flush=True)
    f = F.taylor_green(20, noise=0.05, seed=4, shape=shape)
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    m = (np.sqrt((i - 9) ** 2 + (j - 9.5) ** 2 + (k - 8.5) ** 2) > 4.0).astype(np.float64)
    f = dict(f, mask=m)
    curl = O.curl(shape, f["delta"], f["acc"], m)
    thr = float(np.quantile(curl[m > 0], 0.8)) if method == omni3d.SELECT else 0.0
    p0 = np.random.default_rng(8).normal(size=shape)
    p, w, rep = _run(shape, f, ref_dirs2, method, faithful, omni3d.FP64, omni3d.AFFINE, p0, 7, thr,
                     bits=3)
    ocurl = m if method == omni3d.PLAIN_MASK else curl
    om = {omni3d.PLAIN_MASK: O.PLAIN_MASK, omni3d.SELECT: O.SELECT,
          omni3d.WEIGHTED: O.WEIGHTED}[method]
    po, co, it, rel = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                    curl=ocurl, threshold=thr, method=om, faithful=faithful,
                                    p0=p0, max_iter=7)
    assert rep.iterations == it == 7
    assert rep.sources > 0
    assert F.rel_l2(w, co) < 1e-12
    err = F.rel_l2(p, po)
    print("affine masked vs oracle rel-L2", err)
    assert err < (1e-12 if method == omni3d.PLAIN_MASK else 1e-10)
    assert abs(rep.rel_change - rel) <= 1e-8 * rel


def test_affine_masked_tolerance_stop(gpu_available, ref_dirs2):
    """The on-device stop rule over S stops both schedules at the same pass."""
    n = 24
    shape = (n, n, n)
    print("case affine masked tolerance stop", flush=True)
    f = F.fourier_sphere(n, seed=2)
    p0 = np.random.default_rng(3).normal(size=shape)
    pa, wa, ra = _run(shape, f, ref_dirs2, omni3d.PLAIN_MASK, 0, omni3d.FP64, omni3d.AFFINE, p0,
                      200, 0.0, tol=1e-2)
    pn, wn, rn = _run(shape, f, ref_dirs2, omni3d.PLAIN_MASK, 0, omni3d.FP64, omni3d.NAIVE, p0,
                      200, 0.0, tol=1e-2)
    print("stopped after", ra.iterations, rn.iterations, "rel", ra.rel_change, rn.rel_change)
    assert ra.iterations == rn.iterations < 200
    assert abs(ra.rel_change - rn.rel_change) <= 1e-9 * rn.rel_change
    # per-pass rounding differences add up over the ~80 passes
    assert F.rel_l2(pa, pn) < 1e-11


def test_affine_masked_twice_and_after_field_change(gpu_available, ref_dirs2):
    """A context reused: the second reconstruction (same field) equals the
    first bit for bit; a new mask rebuilds S and the rows."""
    n = 24
    shape = (n, n, n)
    print("case affine masked reuse", flush=True)
    f = F.fourier_sphere(n, seed=4)
    with omni3d.Context(shape, f["delta"], density=f["rho"], method=omni3d.PLAIN_MASK,
                        max_iterations=9) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"], mask=f["mask"])
        p1, w1, r1 = ctx.reconstruct()
        p2, w2, r2 = ctx.reconstruct()
        np.testing.assert_array_equal(p1, p2)
        m2 = f["mask"].copy()
        m2[n // 2:] = 1.0  # the upper half of the solid sphere becomes fluid
        ctx.set_field(f["acc"], mask=m2)
        p3, w3, r3 = ctx.reconstruct()
    with omni3d.Context(shape, f["delta"], density=f["rho"], method=omni3d.PLAIN_MASK,
                        solver=omni3d.NAIVE, max_iterations=9) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"], mask=m2)
        p4, w4, r4 = ctx.reconstruct()
    assert r3.sources != r1.sources
    np.testing.assert_array_equal(w3, w4)
    assert F.rel_l2(p3, p4) < 1e-12


@pytest.mark.parametrize("precision", [omni3d.FP64, omni3d.FP32_KAHAN])
def test_plain_mask_series_reuses_mask_structures(gpu_available, ref_dirs2, precision):
    """Snapshots with one mask: the second field reuses the first's rule
    records, run starts and rows and equals a fresh context bit for bit; a
    new mask, a removed mask and a context switched to the reference
    schedule in between all rebuild them."""
    n = 28
    shape = (n, n, n)
    print(f"case plain-mask series precision={precision}", flush=True)
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    fa, fb = F.fourier_sphere(n, seed=11), F.fourier_sphere(n, seed=12)
    m2 = fa["mask"].copy()
    m2[: n // 3] = 1.0

    def fresh(acc, mask, solver=omni3d.AFFINE):
        with omni3d.Context(shape, fa["delta"], density=fa["rho"], method=omni3d.PLAIN_MASK,
                            precision=precision, solver=solver, max_iterations=12) as c:
            c.set_directions(ref_dirs2)
            c.set_field([np.asarray(a, dtype=dtype) for a in acc], mask=mask, dtype=dtype)
            return c.reconstruct()

    with omni3d.Context(shape, fa["delta"], density=fa["rho"], method=omni3d.PLAIN_MASK,
                        precision=precision, max_iterations=12) as ctx:
        ctx.set_directions(ref_dirs2)
        run = []
        for acc, mask in ((fa["acc"], fa["mask"]), (fb["acc"], fa["mask"]),
                          (fb["acc"], m2), (fa["acc"], None), (fb["acc"], fa["mask"])):
            ctx.set_field([np.asarray(a, dtype=dtype) for a in acc], mask=mask, dtype=dtype)
            run.append((ctx.reconstruct(), acc, mask))
    for (p, w, r), acc, mask in run:
        pf, wf, rf = fresh(acc, mask)
        np.testing.assert_array_equal(w, wf)
        np.testing.assert_array_equal(p, pf)
        assert r.sources == rf.sources and r.source_visits == rf.source_visits
    # the second snapshot's result also matches the reference schedule
    pn, wn, rn = fresh(fb["acc"], fa["mask"], solver=omni3d.NAIVE)
    np.testing.assert_array_equal(run[1][0][1], wn)
    assert F.rel_l2(run[1][0][0], pn) < (1e-12 if precision == omni3d.FP64 else 1e-7)

