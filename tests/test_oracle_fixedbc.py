"""Pins the oracle's SelectFixedBC variant (bodyIntegralInnerSelectFixedBC,
kernel.cu:2538-2796, launcher omni3dparallellinesESInnerSelectFixedBC,
kernel.cu:6481-6549) on the CPU.

The variant is the Select integrator with another step rule: an x / y
distance tie goes to y (kernel.cu:2628 `d1 < d2`, 2676 `d2 <= d1`), no x- or
z-step is taken while on the j = 0 plane (`jnext1 != 0`, `jnext3 != 0`) and
no y-step onto it (`jnext2 > 0`); a refused step ends the walk, as any
refused step does.  So j = 0 voxels are only ever entry voxels: the plane
keeps its prescribed (Bernoulli) pressure.

  * paths: a brute-force walker written from those lines (Euclidean
    distances of the candidate centres to the line) against the oracle's
    staircase with the rule on, differing only at exact distance ties;
  * the j = 0 plane: never visited after the entry, and kept exactly by a
    reconstruction from any starting pressure, which plain Select changes;
  * quadratic pressure (linear gradient, trapezoid exact): from the exact
    seed every visited voxel is reproduced, so the result is p.
"""
import numpy as np
import pytest

import oracle as O
from omni3d import fields as F


def brute_fixedbc(shape, kv, ls, point):
    Z, Y, X = shape
    k = O.make_dir(kv)[0]
    ng = O.nogrid(shape, ls)
    o = O.origin(kv, ng, ls, point)
    ok, a, b = O.clip(shape, o, k)
    if not ok:
        return None, None
    line = o + np.array([(X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0])
    c = a.astype(int).copy()
    path = [c.copy()]
    dists = []
    while not np.array_equal(c, b):
        d = [np.inf, np.inf, np.inf]
        s = [0, 0, 0]
        for ax in range(3):
            if c[ax] == b[ax]:
                continue
            s[ax] = 1 if b[ax] > c[ax] else -1
            cand = c.copy()
            cand[ax] += s[ax]
            d[ax] = np.linalg.norm(np.cross(cand - line, k))
        if d[0] < d[1] and d[0] <= d[2]:
            ax = 0
        elif d[1] <= d[0] and d[1] <= d[2]:
            ax = 1
        elif d[2] < d[0] and d[2] < d[1]:
            ax = 2
        else:
            break
        if ax != 1 and c[1] == 0:
            break
        if ax == 1 and c[1] + s[1] == 0:
            break
        c[ax] += s[ax]
        path.append(c.copy())
        dists.append(d)
    return path, dists


def to_index(shape, ijk):
    Z, Y, X = shape
    return int(ijk[0] + ijk[1] * X + ijk[2] * X * Y)


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 8, 11)])
def test_fixedbc_paths_match_brute_force(shape, ref_dirs2):
    Z, Y, X = shape
    rays = O.enumerate_rays(shape, ref_dirs2, 1.0)
    ties = 0
    differ_from_plain = 0
    for r in rays:
        d, pt = int(r[0]), int(r[1])
        got = list(O.ray_path(shape, ref_dirs2[d], 1.0, pt, fixedbc=1))
        bp, dists = brute_fixedbc(shape, ref_dirs2[d], 1.0, pt)
        want = [to_index(shape, c) for c in bp]
        if got != list(O.ray_path(shape, ref_dirs2[d], 1.0, pt)):
            differ_from_plain += 1
        if got == want:
            continue
        j = next((i for i in range(min(len(got), len(want))) if got[i] != want[i]),
                 min(len(got), len(want)))
        dd = [v for v in dists[j - 1] if np.isfinite(v)]
        assert len(dd) > 1 and np.ptp(sorted(dd)[:2]) <= 1e-9, (shape, d, pt, j, dists[j - 1])
        ties += 1
    assert ties <= 0.01 * len(rays)
    # the rule matters: rays along or onto the j = 0 plane change
    assert differ_from_plain > 0


def test_fixedbc_never_steps_on_plane_j0(ref_dirs3):
    shape = (8, 9, 10)
    Z, Y, X = shape
    rays = O.enumerate_rays(shape, ref_dirs3, 1.0)
    for r in rays[::3]:
        path = O.ray_path(shape, ref_dirs3[int(r[0])], 1.0, int(r[1]), fixedbc=1)
        j = (path // X) % Y
        assert np.all(j[1:] != 0)


def test_fixedbc_keeps_plane_j0_and_reproduces_quadratic(ref_dirs2):
    shape = (9, 10, 11)
    f = F.taylor_green(11, noise=0.05, seed=3, shape=shape)
    curl = O.curl(shape, f["delta"], f["acc"])
    thr = float(np.quantile(curl, 0.9))
    p0 = np.random.default_rng(4).normal(size=shape)
    p, cnt, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0=p0,
                                 curl=curl, threshold=thr, method=O.SELECT_FIXEDBC,
                                 max_iter=3, seed_pn=0)
    np.testing.assert_array_equal(p[:, 0, :], p0[:, 0, :])
    np.testing.assert_array_equal(cnt[:, 0, :], 0.0)
    ps, _, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0=p0,
                                curl=curl, threshold=thr, method=O.SELECT, max_iter=3, seed_pn=0)
    assert np.abs(ps[:, 0, :] - p0[:, 0, :]).max() > 0
    q = F.quadratic(shape, seed=2)
    cq = O.curl(shape, q["delta"], q["acc"]) + 1.0  # curl of a gradient is 0: keep every voxel
    pq, _, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, q["delta"], q["rho"], q["acc"],
                                p0=q["p"], curl=cq, threshold=1e30, method=O.SELECT_FIXEDBC,
                                max_iter=3, seed_pn=0)
    assert np.abs(pq - q["p"]).max() <= 1e-9 * np.abs(q["p"]).max()


# ---------------------------------------------------------------------------
# MiniCurl (bodyIntegralInnerMiniCurl, kernel.cu:2797-2936): the path steps
# toward the exit through the candidate with the smallest curl, ties x, then
# y, then z; an axis at the exit offers a sentinel that counts as 1e10 when
# its aliased index falls outside the volume (always, below 60000 voxels).

def runs_path(ijk_in, ijk_out, order):
    """Voxels of the path that runs axis by axis in the given order."""
    c = list(ijk_in)
    path = [tuple(c)]
    for ax in order:
        s = 1 if ijk_out[ax] > c[ax] else -1
        while c[ax] != ijk_out[ax]:
            c[ax] += s
            path.append(tuple(c))
    return path


def pass_along(shape, rays, order_of, p, acc, delta, rho):
    Z, Y, X = shape
    a = [np.asarray(x, np.float64).ravel() for x in acc]
    pf = np.asarray(p, np.float64).ravel()
    pn = np.zeros(X * Y * Z)
    cnt = np.zeros(X * Y * Z)
    for r in rays:
        i_in, i_out = tuple(int(v) for v in r[2:5]), tuple(int(v) for v in r[5:8])
        path = runs_path(i_in, i_out, order_of(i_in, i_out))
        lin = [c[0] + c[1] * X + c[2] * X * Y for c in path]
        pin, pint = pf[lin[0]], 0.0
        for q in range(1, len(path)):
            ax = next(k for k in range(3) if path[q][k] != path[q - 1][k])
            s = path[q][ax] - path[q - 1][ax]
            pint += -rho * s * delta[ax] * 0.5 * (a[ax][lin[q]] + a[ax][lin[q - 1]])
            pn[lin[q]] += pin + pint
            cnt[lin[q]] += 1
    return pn.reshape(shape), cnt.reshape(shape)


@pytest.mark.parametrize("curl_kind", ["uniform", "rising_x"])
def test_minicurl_paths_closed_form(ref_dirs2, curl_kind):
    shape = (8, 9, 10)
    Z, Y, X = shape
    f = F.taylor_green(10, noise=0.05, seed=5, shape=shape)
    i = np.broadcast_to(np.arange(X, dtype=np.float64), shape)
    if curl_kind == "uniform":
        curl = np.ones(shape)
        order_of = lambda a, b: (0, 1, 2)  # noqa: E731  (ties: x, then y, then z)
    else:
        # curl 1 + i: a +x candidate is the dearest, a -x candidate the cheapest
        curl = 1.0 + i
        order_of = lambda a, b: (1, 2, 0) if b[0] > a[0] else (0, 1, 2)  # noqa: E731
    p = np.random.default_rng(6).normal(size=shape)
    rays = O.enumerate_rays(shape, ref_dirs2, 1.0)
    want_pn, want_cnt = pass_along(shape, rays, order_of, p, f["acc"], f["delta"], f["rho"])
    pn, cnt, _ = O.one_pass(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p,
                            curl=np.ascontiguousarray(curl), method=O.MINICURL)
    np.testing.assert_array_equal(cnt, want_cnt)
    np.testing.assert_allclose(pn, want_pn, rtol=1e-12, atol=1e-12 * np.abs(want_pn).max())


def test_minicurl_reproduces_quadratic(ref_dirs2):
    shape = (9, 10, 11)
    q = F.quadratic(shape, seed=3)
    curl = np.random.default_rng(7).random(shape) + 0.1  # any path: the trapezoid is exact
    p, cnt, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, q["delta"], q["rho"], q["acc"],
                                 p0=q["p"], curl=curl, method=O.MINICURL, max_iter=3, seed_pn=0)
    assert np.abs(p - q["p"]).max() <= 1e-9 * np.abs(q["p"]).max()
    assert cnt.sum() > 0
