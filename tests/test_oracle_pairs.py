"""Pins the oracle's boundary-pair (PINT) pipeline (CPU only):
omni3dparallellinesEqualSpacing (kernel.cu:3414-3489) with bodyIntegral
(kernel.cu:1353-1467), devidecount (kernel.cu:3896-3909) and BCiteration
(kernel.cu:5572-5712), kept there as a dense n x n matrix and restated in the
oracle as a sorted pair list.

  * a dense-matrix restatement in Python, written from the reference's
    loops (pint[in, out] += integral of the staircase path, pcount[in, out]++,
    divide, then pn[out] = mean over in with pint != 0 of p[in] + pint),
    with the paths from the oracle's staircase and the trapezoid summed here:
    equal to the sorted-list form after 1 and 3 iterations (a swapped in/out,
    a missing divide or a Gauss-Seidel update fails it);
  * a quadratic pressure (linear gradient, trapezoid exact): every pair's
    value is p[out] - p[in], so the exact boundary is a fixed point, and from
    a zero boundary the iteration converges to p + constant (the averaging
    of boundary errors), the behaviour the README describes (README:25);
  * a zero field makes every pair zero: all are skipped and p is unchanged;
  * only exit voxels on the boundary change.
"""
import numpy as np
import pytest

import oracle as O
from omni3d import fields as F


def boundary(shape):
    b = np.zeros(shape, bool)
    b[0] = b[-1] = True
    b[:, 0] = b[:, -1] = True
    b[:, :, 0] = b[:, :, -1] = True
    return b


def dense_pipeline(shape, dirs, ls, delta, rho, acc, p0, iters):
    Z, Y, X = shape
    V = X * Y * Z
    a = [np.asarray(c, np.float64).ravel() for c in acc]
    pint = np.zeros((V, V))
    pcount = np.zeros((V, V))
    for d, point, iin, jin, kin, iout, jout, kout in O.enumerate_rays(shape, dirs, ls):
        path = O.ray_path(shape, dirs[d], ls, point)
        s = 0.0
        for vl, vn in zip(path[:-1], path[1:]):
            step = vn - vl
            ax = {1: 0, X: 1, X * Y: 2}[abs(step)]
            sg = 1 if step > 0 else -1
            s += -rho * sg * delta[ax] * 0.5 * (a[ax][vn] + a[ax][vl])
        nin = iin + jin * X + kin * X * Y
        nout = iout + jout * X + kout * X * Y
        pint[nin, nout] += s
        pcount[nin, nout] += 1
    nz = pcount > 0
    pint[nz] /= pcount[nz]
    p = np.asarray(p0, np.float64).ravel().copy()
    for _ in range(iters):
        pn = p.copy()
        for nout in range(V):
            col = np.nonzero(pint[:, nout] != 0)[0]
            if col.size:
                pn[nout] = np.sum(p[col] + pint[col, nout]) / col.size
        p = pn
    return p.reshape(shape)


@pytest.mark.parametrize("iters", [1, 3])
def test_dense_matrix_restatement(ref_dirs2, iters):
    shape = (5, 6, 7)
    f = F.taylor_green(7, noise=0.05, seed=2, shape=shape)
    rng = np.random.default_rng(1)
    p0 = rng.normal(size=shape)
    want = dense_pipeline(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0, iters)
    got, npairs = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0,
                                   iters)
    assert npairs > 0
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # only boundary voxels change
    np.testing.assert_array_equal(got[~boundary(shape)], p0[~boundary(shape)])


def test_quadratic_exact_boundary_is_fixed_point(ref_dirs2):
    shape = (12, 13, 14)
    f = F.quadratic(shape, seed=1)
    p, _ = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], f["p"], 3)
    assert np.abs(p - f["p"]).max() <= 1e-12 * np.abs(f["p"]).max()


def test_quadratic_converges_to_pressure_plus_constant(ref_dirs2):
    shape = (12, 13, 14)
    f = F.quadratic(shape, seed=1)
    b = boundary(shape)
    spread = []
    for it in (1, 3, 30):
        p, _ = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                np.zeros(shape), it)
        e = (p - f["p"])[b]
        spread.append(np.std(e) / np.std(f["p"][b]))
    assert spread[0] > spread[1] > spread[2]
    assert spread[2] < 1e-6


def test_zero_field_pairs_skipped(ref_dirs2):
    shape = (6, 7, 8)
    p0 = np.random.default_rng(3).normal(size=shape)
    z = [np.zeros(shape)] * 3
    p, npairs = O.boundary_pairs(shape, ref_dirs2, 1.0, (1.0, 1.0, 1.0), 1.0, z, p0, 4)
    assert npairs > 0
    np.testing.assert_array_equal(p, p0)
