"""Pins the oracle's integration, averaging and iteration (CPU only).

  * one plain pass against a brute-force restatement written from the
    reference's description (kernel.cu:1743-1863: entry self-add, trapezoid
    -rho*(next-last)*delta*0.5*(a[next]+a[last]), scatter p_in + pint);
  * closed-form KAT: a quadratic pressure has a linear gradient, which the
    trapezoid integrates exactly, so from the exact seed every method
    (plain, plain-mask with slab and sphere masks, select, weighted)
    reproduces p on every visited voxel -- this fixes the sign of rho, the
    delta of each axis, the 1/2 and the restart rule (p0 read at the re-entry
    voxel, pint reset);
  * count identity: sum(cnt) = sum over rays of (L1 + 1) (plain method);
  * affine structure: recon(alpha g) = alpha recon(g) from a zero seed, and
    recon(g, p0 + c) = recon(g, p0) + c without the seed term;
  * Taylor-Green (32^3, 42 directions, 5 % noise): the converged oracle
    approaches the analytic pressure.
"""
import numpy as np
import pytest

import oracle as O
from omni3d import fields as F


def brute_pass(shape, dirs, ls, delta, rho, acc, p):
    Z, Y, X = shape
    pn = np.zeros(X * Y * Z)
    cnt = np.zeros(X * Y * Z)
    a = [c.reshape(-1) for c in acc]
    pf = p.reshape(-1)
    for r in O.enumerate_rays(shape, dirs, ls):
        path = O.ray_path(shape, dirs[int(r[0])], ls, int(r[1]))
        pin = pf[path[0]]
        pn[path[0]] += pin
        cnt[path[0]] += 1
        pint = 0.0
        for last, nxt in zip(path[:-1], path[1:]):
            step = int(nxt) - int(last)
            ax = {1: 0, X: 1, X * Y: 2}[abs(step)]
            s = 1 if step > 0 else -1
            pint += -rho * s * delta[ax] * 0.5 * (a[ax][nxt] + a[ax][last])
            pn[nxt] += pin + pint
            cnt[nxt] += 1
    return pn.reshape(shape), cnt.reshape(shape)


def test_one_pass_matches_brute_force(ref_dirs2):
    shape = (6, 7, 5)
    f = F.taylor_green(7, noise=0.1, seed=11, shape=shape)
    rng = np.random.default_rng(2)
    p = rng.normal(size=shape)
    delta = (0.3, 0.5, 0.7)
    pn, cnt, _ = O.one_pass(shape, ref_dirs2, 1.0, delta, 1.3, f["acc"], p)
    bpn, bcnt = brute_pass(shape, ref_dirs2, 1.0, delta, 1.3, f["acc"], p)
    np.testing.assert_array_equal(cnt, bcnt)
    np.testing.assert_allclose(pn, bpn, rtol=1e-12, atol=1e-12)


def test_count_identity(ref_dirs3):
    shape = (12, 10, 14)
    f = F.quadratic(shape, seed=1)
    rays = O.enumerate_rays(shape, ref_dirs3, 1.0)
    _, cnt, src = O.one_pass(shape, ref_dirs3, 1.0, f["delta"], 1.0, f["acc"], np.zeros(shape))
    assert cnt.sum() == np.sum(np.abs(rays[:, 5:8] - rays[:, 2:5]).sum(axis=1) + 1)
    # the plain method reads p only at entry voxels, all on the surface
    Z, Y, X = shape
    k, j, i = np.nonzero(src)
    assert np.all((i == 0) | (i == X - 1) | (j == 0) | (j == Y - 1) | (k == 0) | (k == Z - 1))


def sphere_mask(shape, r):
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    c = [(X - 1) / 2, (Y - 1) / 2, (Z - 1) / 2]
    return (np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2) > r).astype(float)


@pytest.mark.parametrize("method,maskkind,faithful", [
    (O.PLAIN, None, 1), (O.PLAIN_MASK, "slab", 0), (O.PLAIN_MASK, "sphere", 0),
    (O.SELECT, "curl", 0), (O.SELECT, "curlzero", 0), (O.SELECT, "curl", 1),
    (O.WEIGHTED, "curl", 1), (O.WEIGHTED, "curlzero", 1)])
def test_quadratic_exact_every_method(ref_dirs2, method, maskkind, faithful):
    """With faithful=1 the select walk's y/z steps skip zero-curl voxels
    without integrating across them (kernel.cu:2442,2488), which is not
    exact; that case is covered by test_faithful_select_skips_zero_curl."""
    shape = (11, 9, 13)
    f = F.quadratic(shape, seed=7, delta=(0.4, 0.6, 0.5), rho=1.7)
    Z, Y, X = shape
    aux, thr = None, 0.0
    if maskkind == "slab":
        aux = np.ones(shape)
        aux[:, :, 4:6] = 0.0
    elif maskkind == "sphere":
        aux = sphere_mask(shape, 3.2)
    elif maskkind in ("curl", "curlzero"):
        # a hand-built positive "curl" field with a high-curl blob above the
        # threshold and (curlzero) a masked, zero-curl slab
        rng = np.random.default_rng(3)
        aux = rng.uniform(0.5, 1.0, shape)
        aux[3:6, 2:5, 5:9] = 5.0
        if maskkind == "curlzero":
            aux[:, 6, :] = 0.0
        thr = 2.0
    p, cnt, it, rel = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                    p0=f["p"], curl=aux, threshold=thr, method=method,
                                    faithful=faithful, max_iter=3, seed_pn=0)
    vis = cnt > 0
    assert vis.sum() > 0.5 * vis.size
    scale = np.abs(f["p"]).max()
    assert np.abs(p - f["p"]).max() <= 1e-11 * scale
    if maskkind in ("slab", "sphere"):
        assert np.all(cnt[aux == 0] == 0)    # masked voxels are never visited
    if method == O.SELECT:
        assert np.all(cnt[aux == 0] == 0)    # curl == 0 means masked
        assert np.all(cnt[aux >= thr] == 0)  # high-curl voxels are skipped
    if method == O.WEIGHTED:
        # no path break: every voxel after the first nonzero curl of a ray counts
        assert np.all(cnt[aux >= thr] > 0)


def test_faithful_select_skips_zero_curl(ref_dirs2):
    """The faithful select walk does not integrate across a zero-curl voxel
    reached by a y or z step (kernel.cu:2442,2488), so with such voxels it
    departs from the symmetric rule and from the exact quadratic; without
    zero-curl voxels the two rules agree wherever the threshold test agrees."""
    shape = (11, 9, 13)
    f = F.quadratic(shape, seed=7, delta=(0.4, 0.6, 0.5), rho=1.7)
    rng = np.random.default_rng(3)
    aux = rng.uniform(0.5, 1.0, shape)
    aux[:, 6, :] = 0.0
    args = (shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"])
    pf = O.reconstruct(*args, p0=f["p"], curl=aux, threshold=2.0, method=O.SELECT, faithful=1,
                       max_iter=1, seed_pn=0)[0]
    ps = O.reconstruct(*args, p0=f["p"], curl=aux, threshold=2.0, method=O.SELECT, faithful=0,
                       max_iter=1, seed_pn=0)[0]
    assert np.abs(pf - f["p"]).max() > 1e-6 * np.abs(f["p"]).max()
    assert np.abs(ps - f["p"]).max() <= 1e-11 * np.abs(f["p"]).max()


def test_linear_in_gradient_and_shift_invariant(ref_dirs2):
    shape = (8, 9, 10)
    f = F.taylor_green(10, noise=0.05, seed=4, shape=shape)
    args = (shape, ref_dirs2, 1.0, f["delta"], f["rho"])
    p1 = O.reconstruct(*args, f["acc"], max_iter=4, seed_pn=0)[0]
    p2 = O.reconstruct(*args, [2.5 * a for a in f["acc"]], max_iter=4, seed_pn=0)[0]
    np.testing.assert_allclose(p2, 2.5 * p1, rtol=1e-12, atol=1e-12 * np.abs(p1).max())
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=shape)
    q1 = O.reconstruct(*args, f["acc"], p0=p0, max_iter=4, seed_pn=0)[0]
    q2 = O.reconstruct(*args, f["acc"], p0=p0 + 3.0, max_iter=4, seed_pn=0)[0]
    np.testing.assert_allclose(q2, q1 + 3.0, rtol=0, atol=1e-12)


def test_seed_term_first_pass_only(ref_dirs2):
    """kernel.cu:8812-8813: pn starts as a copy of p, so the first average has
    one extra term p_init[v]/cnt[v]; later passes start from zero."""
    shape = (7, 8, 6)
    f = F.taylor_green(8, noise=0.0, seed=0, shape=shape)
    rng = np.random.default_rng(1)
    p0 = rng.normal(size=shape)
    args = (shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"])
    a = O.reconstruct(*args, p0=p0, max_iter=1, seed_pn=1)
    b = O.reconstruct(*args, p0=p0, max_iter=1, seed_pn=0)
    cnt = a[1]
    vis = cnt > 0
    np.testing.assert_allclose(a[0][vis], b[0][vis] + p0[vis] / cnt[vis], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(a[0][~vis], p0[~vis])


def test_taylor_green_converges_to_analytic(ref_dirs2):
    f = F.taylor_green(32, noise=0.05, seed=0)
    p, cnt, it, rel = O.reconstruct((32, 32, 32), ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                    max_iter=200, tol=1e-6)
    err = F.rel_l2(F.demean(p), F.demean(f["p"]))
    print("taylor-green 32^3: iterations", it, "rel change", rel, "rel-L2 vs analytic", err)
    assert err < 0.05
    assert it <= 200
