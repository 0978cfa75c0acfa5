"""The masked methods' rule replay (the per-field records: how each segment's
first visit is taken, its last restart, the curl weights, the run starts) on
the tile engine against the one-lane-per-ray replay (OMNI3D_MASK_WALK=lane),
through the C-ABI, on both schedules (the reference's full traversal per
iteration, kernel.cu:8840-8861, and the affine one).

The two replays take every step of the rule in the same order along a ray;
the tile replay runs each segment from each entry state and a per-ray scan
picks the state the ray actually enters with.  So the flags, restarts, run
starts and the unit weights agree exactly: plain and plain-mask results are
bit-identical.  The curl weights (Select, Weighted) are the same terms
summed in another order (the weights at a segment's start plus the
segment's own sum): rel-L2 <= 1e-13 on the weights and <= 1e-12 on p.
"""
import os

import numpy as np
import pytest

import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def _threshold(shape, f, dtype, precision):
    with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision) as ctx:
        ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], mask=f["mask"], dtype=dtype)
        c = ctx.field_curl()
    return float(np.quantile(c[f["mask"] > 0], 0.7))


def _run(shape, f, dirs, method, faithful, precision, solver, p0, iters, thr, bits, lane):
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    old = os.environ.pop("OMNI3D_MASK_WALK", None)
    if lane:
        os.environ["OMNI3D_MASK_WALK"] = "lane"
    try:
        with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision,
                            method=method, solver=solver, max_iterations=iters, tolerance=0.0,
                            threshold=thr, faithful_select=faithful) as ctx:
            ctx.set_tile_shape(bits, 1024)
            ctx.set_directions(dirs)
            ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], mask=f["mask"],
                          dtype=dtype)
            return ctx.reconstruct(p0)
    finally:
        os.environ.pop("OMNI3D_MASK_WALK", None)
        if old is not None:
            os.environ["OMNI3D_MASK_WALK"] = old


def _field(shape):
    # a cube's field cut to the (ragged) shape: the sphere mask off centre
    f = F.fourier_sphere(max(shape), seed=11)
    cut = tuple(slice(0, s) for s in shape)
    out = dict(f)
    out["acc"] = [np.ascontiguousarray(a[cut]) for a in f["acc"]]
    out["mask"] = np.ascontiguousarray(f["mask"][cut])
    return out


CASES = [
    # method, faithful, precision, solver, tile bits, shape
    (omni3d.PLAIN, 0, omni3d.FP64, omni3d.NAIVE, 4, (40, 40, 40)),
    (omni3d.PLAIN_MASK, 0, omni3d.FP64, omni3d.NAIVE, 4, (40, 40, 40)),
    (omni3d.PLAIN_MASK, 0, omni3d.FP32_KAHAN, omni3d.AFFINE, 3, (37, 29, 23)),
    (omni3d.SELECT, 1, omni3d.FP64, omni3d.NAIVE, 4, (40, 40, 40)),
    (omni3d.SELECT, 0, omni3d.FP64, omni3d.AFFINE, 4, (40, 40, 40)),
    (omni3d.SELECT, 1, omni3d.FP32_KAHAN, omni3d.AFFINE, 3, (37, 29, 23)),
    (omni3d.SELECT, 0, omni3d.FP64, omni3d.NAIVE, 3, (33, 41, 26)),
    (omni3d.WEIGHTED, 0, omni3d.FP64, omni3d.NAIVE, 4, (40, 40, 40)),
    (omni3d.WEIGHTED, 0, omni3d.FP32_KAHAN, omni3d.AFFINE, 3, (37, 29, 23)),
]


@pytest.mark.parametrize("method,faithful,precision,solver,bits,shape", CASES)
def test_tile_rule_replay_matches_lane_replay(gpu_available, ref_dirs3, method, faithful,
                                              precision, solver, bits, shape):
    print(f"case rule replay method={method} faithful={faithful} precision={precision} "
          f"solver={solver} bits={bits} shape={shape}", flush=True)
    f = _field(shape)
    dtype = np.float64 if precision == omni3d.FP64 else np.float32
    thr = _threshold(shape, f, dtype, precision) if method == omni3d.SELECT else 0.0
    p0 = np.random.default_rng(5).normal(size=shape)
    pt, wt, rt = _run(shape, f, ref_dirs3, method, faithful, precision, solver, p0, 5, thr, bits,
                      lane=False)
    pl, wl, rl = _run(shape, f, ref_dirs3, method, faithful, precision, solver, p0, 5, thr, bits,
                      lane=True)
    assert rt.iterations == rl.iterations == 5
    if method in (omni3d.PLAIN, omni3d.PLAIN_MASK):
        np.testing.assert_array_equal(wt, wl)
        np.testing.assert_array_equal(pt, pl)
    else:
        assert F.rel_l2(wt, wl) < 1e-13
        err = F.rel_l2(pt, pl)
        print("tile vs lane rule replay rel-L2", err)
        assert err < 1e-12
    if solver == omni3d.AFFINE:
        assert rt.sources == rl.sources and rt.source_visits == rl.source_visits
