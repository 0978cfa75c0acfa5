"""GPU parity of the trilinear traversal (OMNI3D_TRAVERSAL_TRILINEAR: the plain
method's staircase visits with the line integral taken along the straight
ray through trilinear samples of the gradient), through the C-ABI:

  * FP64 against the oracle's trilinear walk (orc method PLAIN_TRILINEAR,
    the naive schedule) on the affine solver: rel-L2 <= 1e-12, visit counts
    and iteration counts equal, with and without the stop rule;
  * FP32 ray sums (Kahan-compensated on the one-lane walk) <= 1e-5;
  * a quadratic pressure is reproduced from its exact seed at 48^3 with 162
    directions (the traversal is exact for a linear gradient): <= 1e-9;
  * the staircase and trilinear results differ on a noisy field (the
    traversal parameter is honoured).
"""
import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def run(shape, f, dirs, **kw):
    with omni3d.Context(shape, f["delta"], density=f["rho"], **kw) as ctx:
        ctx.set_directions(dirs)
        ctx.set_field(f["acc"])
        return ctx.reconstruct(f.get("p0"))


@pytest.mark.parametrize("precision,tol", [(omni3d.FP64, 0.0), (omni3d.FP64, 1e-6),
                                           (omni3d.FP32_KAHAN, 0.0)])
def test_trilinear_vs_oracle(gpu_available, ref_dirs2, precision, tol):
    shape = (18, 20, 22)
    print(f"case trilinear precision={precision} tol={tol}", flush=True)
    f = F.taylor_green(22, noise=0.05, seed=21, shape=shape)
    if precision == omni3d.FP32_KAHAN:
        f = dict(f, acc=[np.asarray(a, np.float32).astype(np.float64) for a in f["acc"]])
    K = 40 if tol else 12
    p, w, rep = run(shape, f, ref_dirs2, traversal=omni3d.TRILINEAR, precision=precision,
                    max_iterations=K, tolerance=tol)
    po, co, it, rel = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                    method=O.PLAIN_TRILINEAR, max_iter=K, tol=tol)
    np.testing.assert_array_equal(w, co)
    assert rep.iterations == it
    err = F.rel_l2(p, po)
    print("trilinear rel-L2 vs oracle", err, "iterations", it)
    if precision == omni3d.FP64:
        # per-pass rounding differences (~1e-15) add up over the passes: the
        # averaging keeps a constant shift of p and damps smooth modes slowly
        assert err < (1e-12 if K <= 12 else 1e-11)
    else:
        assert err < 1e-5


def test_trilinear_quadratic_exact(gpu_available, ref_dirs3):
    shape = (48, 48, 48)
    print("case trilinear quadratic 48^3", flush=True)
    f = F.quadratic(shape, seed=5, delta=(0.02, 0.025, 0.03), rho=1.3)
    f = dict(f, p0=f["p"])
    p, w, rep = run(shape, f, ref_dirs3, traversal=omni3d.TRILINEAR, max_iterations=3,
                    seed_pn_with_p=0)
    m = w > 0
    err = np.abs(p[m] - f["p"][m]).max() / np.abs(f["p"]).max()
    print("trilinear quadratic max rel error", err)
    assert err < 1e-9


def test_traversal_changes_result(gpu_available, ref_dirs2):
    shape = (16, 16, 16)
    print("case trilinear vs staircase", flush=True)
    f = F.taylor_green(16, noise=0.1, seed=3, shape=shape)
    ps, ws, _ = run(shape, f, ref_dirs2, max_iterations=8)
    pt, wt, _ = run(shape, f, ref_dirs2, max_iterations=8, traversal=omni3d.TRILINEAR)
    np.testing.assert_array_equal(ws, wt)
    assert F.rel_l2(ps, pt) > 1e-4
