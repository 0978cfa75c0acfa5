"""The reference's integrators whose walks leave the staircase of the plan
(SelectFixedBC kernel.cu:2538-2796, MiniCurl kernel.cu:2797-2936,
WeightedFixedBC kernel.cu:2136-2279, WeightedMiniCurl kernel.cu:2937-3085)
on the affine schedule (the default solver): one pass from p = 0 for the
constant, the weights and the run starts, the rows of the visits landing on
the run starts from two walks, the iteration on the run starts, one full
pass.  Against the reference schedule (a full walk of every ray per
iteration, the NAIVE solver) through the C-ABI; the oracle comparisons of
these methods (tests/test_gpu_methods.py) run on the default solver too.

Why they agree: the runs, their start voxels and the weights depend on the
curl field alone, so each pass maps p on the run starts affinely.  Only the
rounding differs (w * p[start] and the run's terms rounded apart instead of
(p[start] + terms) * w): rel-L2 <= 1e-10 in FP64 and FP32 ray sums (the
terms are the same FP32 sums on both schedules); MiniCurl's unit weights
equal.
"""
import numpy as np
import pytest

import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    f = F.taylor_green(max(shape), noise=0.05, seed=seed, shape=shape)
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    m = (np.sqrt((i - X / 2) ** 2 + (j - Y / 2) ** 2 + (k - Z / 2) ** 2) > 4.0).astype(np.float64)
    return dict(f, mask=m)


def _run(shape, f, dirs, method, precision, solver, p0, iters, thr=0.0, faithful=1, tol=0.0,
         seed_pn=1):
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision, method=method,
                        solver=solver, max_iterations=iters, tolerance=tol, threshold=thr,
                        faithful_select=faithful, seed_pn_with_p=seed_pn) as ctx:
        ctx.set_directions(dirs)
        ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], mask=f["mask"], dtype=dtype)
        return ctx.reconstruct(p0)


def _threshold(shape, f):
    with omni3d.Context(shape, f["delta"], density=f["rho"]) as ctx:
        ctx.set_field(f["acc"], mask=f["mask"])
        c = ctx.field_curl()
    return float(np.quantile(c[f["mask"] > 0], 0.75))


CASES = [
    # method, faithful, precision, seed_pn
    (omni3d.SELECT_FIXEDBC, 1, omni3d.FP64, 1),
    (omni3d.SELECT_FIXEDBC, 0, omni3d.FP64, 0),
    (omni3d.SELECT_FIXEDBC, 1, omni3d.FP32_KAHAN, 0),
    (omni3d.MINICURL, 1, omni3d.FP64, 0),
    (omni3d.MINICURL, 1, omni3d.FP32_KAHAN, 1),
    (omni3d.WEIGHTED_FIXEDBC, 1, omni3d.FP64, 1),
    (omni3d.WEIGHTED_FIXEDBC, 1, omni3d.FP32_KAHAN, 0),
    (omni3d.WEIGHTED_MINICURL, 1, omni3d.FP64, 0),
    (omni3d.WEIGHTED_MINICURL, 1, omni3d.FP32_KAHAN, 1),
]


@pytest.mark.parametrize("method,faithful,precision,seed_pn", CASES)
def test_lane_affine_matches_reference_schedule(gpu_available, ref_dirs2, method, faithful,
                                                precision, seed_pn):
    shape = (26, 22, 24)
    print(f"case lane affine method={method} faithful={faithful} precision={precision} "
          f"seed_pn={seed_pn}", flush=True)
    f = _field(shape, 12)
    thr = _threshold(shape, f) if method == omni3d.SELECT_FIXEDBC else 0.0
    p0 = np.random.default_rng(13).normal(size=shape)
    pa, wa, ra = _run(shape, f, ref_dirs2, method, precision, omni3d.AFFINE, p0, 6, thr,
                      faithful, seed_pn=seed_pn)
    pn, wn, rn = _run(shape, f, ref_dirs2, method, precision, omni3d.NAIVE, p0, 6, thr,
                      faithful, seed_pn=seed_pn)
    assert ra.iterations == rn.iterations == 6
    # the affine schedule ran: the run starts and their rows are reported
    assert 0 < ra.sources < np.prod(shape) and ra.source_visits > 0
    if method == omni3d.MINICURL:
        np.testing.assert_array_equal(wa, wn)
    else:
        assert F.rel_l2(wa, wn) < 1e-12
    err = F.rel_l2(pa, pn)
    print("lane affine vs reference schedule rel-L2", err)
    assert err < 1e-10
    if method in (omni3d.SELECT_FIXEDBC,):
        # SelectFixedBC keeps the j = 0 plane at its starting pressure
        np.testing.assert_array_equal(pa[:, 0, :], p0[:, 0, :])


@pytest.mark.parametrize("method", [omni3d.MINICURL, omni3d.WEIGHTED_MINICURL])
def test_lane_affine_tolerance_stop(gpu_available, ref_dirs2, method):
    """The stop rule over the run starts stops both schedules at the same pass."""
    shape = (20, 20, 20)
    print(f"case lane affine tolerance stop method={method}", flush=True)
    f = _field(shape, 14)
    p0 = np.random.default_rng(15).normal(size=shape)
    # a tolerance the iteration reaches within about 30 passes: 1.5 x the
    # relative change of pass 30
    _, _, r30 = _run(shape, f, ref_dirs2, method, omni3d.FP64, omni3d.NAIVE, p0, 30, seed_pn=0)
    tol = 1.5 * r30.rel_change
    pa, wa, ra = _run(shape, f, ref_dirs2, method, omni3d.FP64, omni3d.AFFINE, p0, 200,
                      tol=tol, seed_pn=0)
    pn, wn, rn = _run(shape, f, ref_dirs2, method, omni3d.FP64, omni3d.NAIVE, p0, 200,
                      tol=tol, seed_pn=0)
    print("stopped after", ra.iterations, rn.iterations, "rel", ra.rel_change, rn.rel_change)
    assert ra.iterations == rn.iterations < 200
    assert abs(ra.rel_change - rn.rel_change) <= 1e-8 * rn.rel_change
    assert F.rel_l2(pa, pn) < 1e-10


def test_lane_affine_context_reused(gpu_available, ref_dirs2):
    """A context reconstructing twice gives the same result (to the order of
    the one-lane walks' FP64 atomic adds of the weights and the constant),
    and a changed field rebuilds the run starts and the rows."""
    shape = (20, 18, 16)
    print("case lane affine context reused", flush=True)
    f = _field(shape, 16)
    g = _field(shape, 17)
    p0 = np.random.default_rng(18).normal(size=shape)
    with omni3d.Context(shape, f["delta"], density=f["rho"], method=omni3d.WEIGHTED_MINICURL,
                        max_iterations=5, tolerance=0.0) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"], mask=f["mask"])
        p1, w1, r1 = ctx.reconstruct(p0)
        p2, w2, r2 = ctx.reconstruct(p0)
        ctx.set_field(g["acc"], mask=g["mask"])
        p3, w3, r3 = ctx.reconstruct(p0)
    assert r1.sources == r2.sources and r1.source_visits == r2.source_visits
    assert F.rel_l2(w1, w2) < 1e-14
    assert F.rel_l2(p1, p2) < 1e-13
    assert r3.sources != r1.sources or r3.source_visits != r1.source_visits
    pn, wn, rn = _run(shape, g, ref_dirs2, omni3d.WEIGHTED_MINICURL, omni3d.FP64, omni3d.NAIVE,
                      p0, 5)
    assert F.rel_l2(p3, pn) < 1e-10
