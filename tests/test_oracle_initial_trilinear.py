"""Pins two oracle functions (CPU only).

initialIntegration (initialIntegration.cpp:9-408):
  * closed form: a quadratic pressure has a linear gradient, which one
    trapezoid step integrates exactly, so from its exact value at the centre
    voxel every voxel is reproduced exactly (this fixes the signs, the 0.5,
    the spacing and density factors and the neighbour toward the centre);
  * hand-worked grids with a rotational acceleration, where the one to three
    predictions a voxel averages differ: 3x3x1 with a_y = c i (two-way means)
    and 3x3x3 with a_z = c i (three-way means), every value derived by hand
    from the file's loops;
  * the centre keeps the caller's pn, p equals pn everywhere, pref only
    lands on voxel 0 when voxel 0 is the centre;
  * one-line mutants of the restatement must fail one of the above.

Trilinear plain walk (the accuracy mode beside the staircase of
kernel.cu:1743-1863):
  * closed form: trilinear sampling of a linear gradient is exact (with the
    linear extrapolation at the faces) and every trapezoid of the path is
    exact, so a quadratic pressure is reproduced exactly from the exact seed,
    for generic, pole and axis directions and anisotropic spacing;
  * brute force on 6^3 grids: an independent restatement in Python that
    computes each voxel's projection onto the ray from the ray's origin and
    direction (not the walker's incremental offset) and samples with its own
    trilinear interpolation;
  * it visits exactly the plain staircase's voxels (equal counts and
    sources) and is linear in the field.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from omni3d import fields as F


# --- initialIntegration ---------------------------------------------------

@pytest.mark.parametrize("shape", [(5, 6, 7), (4, 4, 4), (3, 8, 2), (1, 5, 4)])
def test_initial_quadratic_exact(shape):
    delta, rho = (0.3, 0.7, 0.45), 1.3
    f = F.quadratic(shape, seed=4, delta=delta, rho=rho)
    Z, Y, X = shape
    pn0 = np.zeros(shape)
    pn0[Z // 2, Y // 2, X // 2] = f["p"][Z // 2, Y // 2, X // 2]
    p, pn = O.initial_integration(shape, delta, rho, f["acc"], pref=0.0, pn=pn0)
    scale = np.abs(f["p"]).max()
    np.testing.assert_allclose(p, f["p"], rtol=0, atol=1e-12 * scale)
    np.testing.assert_array_equal(p, pn)


def _hand_3x3x1(L=None):
    # a = (0, c i, 0), rho * dy = 1; centre (1, 1) holds P0
    c, P0 = 0.8, 2.5
    shape = (1, 3, 3)
    ay = np.zeros(shape)
    ay[0] = c * np.arange(3)[None, :]
    acc = [np.zeros(shape), ay, np.zeros(shape)]
    pn0 = np.zeros(shape)
    pn0[0, 1, 1] = P0
    p, _ = O.initial_integration(shape, (0.9, 0.5, 1.0), 2.0, acc, pref=0.0, pn=pn0, L=L)
    want = np.array([[P0 + c / 2, P0 + c, P0 + 1.5 * c],
                     [P0, P0, P0],
                     [P0 - c / 2, P0 - c, P0 - 1.5 * c]])
    np.testing.assert_allclose(p[0], want, rtol=1e-14, atol=1e-14)


def _hand_3x3x3(L=None):
    # a = (0, 0, c i), rho * dz = 1; centre (1, 1, 1) holds P0.  With
    # c' = c: corner (0,0,0) = P0 + c/2 (mean of P0 + c, P0 + c/2, P0),
    # (2,0,0) = P0 + 1.5 c (mean of P0 + c, P0 + 1.5 c, P0 + 2 c).
    c, P0 = 0.6, -1.25
    shape = (3, 3, 3)
    az = np.zeros(shape)
    az[:] = c * np.arange(3)[None, None, :]
    acc = [np.zeros(shape), np.zeros(shape), az]
    pn0 = np.zeros(shape)
    pn0[1, 1, 1] = P0
    p, _ = O.initial_integration(shape, (0.7, 0.9, 0.25), 4.0, acc, pref=0.0, pn=pn0, L=L)
    want = {(0, 0, 0): P0 + c / 2, (1, 0, 0): P0 + c, (2, 0, 0): P0 + 1.5 * c,
            (0, 1, 0): P0 + c / 2, (1, 1, 0): P0 + c, (2, 1, 0): P0 + 1.5 * c,
            (0, 0, 1): P0, (2, 0, 1): P0, (1, 1, 1): P0}
    for (i, j, k), w in want.items():
        assert p[k, j, i] == pytest.approx(w, rel=1e-14, abs=1e-14), (i, j, k)


def test_initial_hand_3x3x1():
    _hand_3x3x1()


def test_initial_hand_3x3x3():
    _hand_3x3x3()


def test_initial_centre_pref_and_shift():
    shape = (4, 5, 6)
    f = F.taylor_green(6, noise=0.2, seed=5, shape=shape)
    pa, _ = O.initial_integration(shape, f["delta"], 1.0, f["acc"], pref=7.0)
    pn0 = np.zeros(shape)
    pn0[2, 2, 3] = 3.5
    pb, _ = O.initial_integration(shape, f["delta"], 1.0, f["acc"], pref=-2.0, pn=pn0)
    # voxel 0 is not the centre: pref is overwritten; the result shifts by
    # the centre value
    np.testing.assert_allclose(pb - pa, 3.5, rtol=0, atol=1e-12)
    assert pa[2, 2, 3] == 0.0
    # a single voxel is its own centre: it keeps pref
    p1, pn1 = O.initial_integration((1, 1, 1), (1, 1, 1), 1.0, [np.zeros((1, 1, 1))] * 3,
                                    pref=4.25)
    assert p1[0, 0, 0] == 4.25 and pn1[0, 0, 0] == 4.25


INITIAL_MUTANTS = [
    ("three-way mean not divided", "if (n == 3) val = val / 3.0;", ""),
    ("two-way mean not divided", "if (n == 2) val = val / 2.0;", ""),
    ("signs of the two sides swapped", "sg[a] < 0 ? pn[nb] + t : pn[nb] - t",
     "sg[a] < 0 ? pn[nb] - t : pn[nb] + t"),
    ("trapezoid without the half", "0.5 * (acc[a][v] + acc[a][nb])", "(acc[a][v] + acc[a][nb])"),
    ("one-sided difference", "0.5 * (acc[a][v] + acc[a][nb])", "acc[a][nb]"),
    ("spacing of the x axis used for all", "* del[a] * rho", "* del[0] * rho"),
    ("last prediction replaces the mean", "val = n == 0 ? pr : val + pr;",
     "val = n == 0 ? pr : pr + pr;"),
]


def _mutant(tmp, src, old, new):
    assert src.count(old) == 1, old
    d = os.path.join(tmp, f"m{abs(hash(old + new)) % 10**8}")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "omni3d_oracle.c"), "w") as fh:
        fh.write(src.replace(old, new))
    shutil.copy(os.path.join(O.HERE, "omni3d_oracle.h"), d)
    so = os.path.join(d, "liboracle_mutant.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(d, "omni3d_oracle.c"), "-lm"], check=True)
    return O.load(so)


def _quadratic_check(L):
    shape, delta, rho = (5, 6, 7), (0.3, 0.7, 0.45), 1.3
    f = F.quadratic(shape, seed=4, delta=delta, rho=rho)
    pn0 = np.zeros(shape)
    pn0[2, 3, 3] = f["p"][2, 3, 3]
    p, _ = O.initial_integration(shape, delta, rho, f["acc"], pref=0.0, pn=pn0, L=L)
    np.testing.assert_allclose(p, f["p"], rtol=0, atol=1e-12 * np.abs(f["p"]).max())


def test_initial_mutants_fail(tmp_path):
    with open(os.path.join(O.HERE, "omni3d_oracle.c")) as fh:
        src = fh.read()
    survivors = []
    for name, old, new in INITIAL_MUTANTS:
        L = _mutant(str(tmp_path), src, old, new)
        failed = False
        for check in (_quadratic_check, _hand_3x3x1, _hand_3x3x3):
            try:
                check(L)
            except AssertionError:
                failed = True
                break
        if not failed:
            survivors.append(name)
    assert not survivors, survivors


# --- trilinear plain walk --------------------------------------------------

def _dirs_mix(ref_dirs2):
    rng = np.random.default_rng(8)
    g = rng.normal(size=(5, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    axes = np.array([[1, 0, 0], [0, -1, 0], [0, 0, 1]], dtype=np.float64)
    return np.concatenate([ref_dirs2, g, axes]).astype(np.float32)


@pytest.mark.parametrize("shape", [(9, 8, 10), (6, 6, 6), (12, 5, 7)])
def test_trilinear_quadratic_exact(ref_dirs2, shape):
    delta, rho = (0.4, 0.25, 0.55), 1.7
    f = F.quadratic(shape, seed=11, delta=delta, rho=rho)
    dirs = _dirs_mix(ref_dirs2)
    # one pass from the exact pressure: every visit carries p exactly, so
    # the average is p wherever a ray passes
    pn, cnt, _ = O.one_pass(shape, dirs, 1.0, delta, rho, f["acc"], f["p"],
                            method=O.PLAIN_TRILINEAR)
    m = cnt > 0
    assert m.mean() > 0.9
    scale = np.abs(f["p"]).max()
    np.testing.assert_allclose(pn[m] / cnt[m], f["p"][m], rtol=0, atol=1e-12 * scale)
    # and the iteration from the exact seed stays there
    p, _, _, _ = O.reconstruct(shape, dirs, 1.0, delta, rho, f["acc"], p0=f["p"],
                               method=O.PLAIN_TRILINEAR, max_iter=3, seed_pn=0)
    np.testing.assert_allclose(p, f["p"], rtol=0, atol=1e-12 * scale)


def _tri_np(acc, x):
    """Trilinear sample (own implementation): the cell's lower corner clamped
    into the grid, unclamped weights (linear extrapolation)."""
    Z, Y, X = acc[0].shape
    size = (X, Y, Z)
    lo, fr = [], []
    for a in range(3):
        fl = min(max(int(np.floor(x[a])), 0), size[a] - 2)
        lo.append(fl)
        fr.append(x[a] - fl)
    out = np.zeros(3)
    for c in range(3):
        s = 0.0
        for dk in (0, 1):
            for dj in (0, 1):
                for di in (0, 1):
                    w = ((fr[0] if di else 1 - fr[0]) * (fr[1] if dj else 1 - fr[1]) *
                         (fr[2] if dk else 1 - fr[2]))
                    s += w * acc[c][lo[2] + dk, lo[1] + dj, lo[0] + di]
        out[c] = s
    return out


def _brute_trilinear_pass(shape, dirs, delta, rho, acc, p):
    Z, Y, X = shape
    pn = np.zeros(X * Y * Z)
    cnt = np.zeros(X * Y * Z)
    pf = p.reshape(-1)
    centre = np.array([(X - 1) / 2, (Y - 1) / 2, (Z - 1) / 2])
    dl = np.array(delta)
    for r in O.enumerate_rays(shape, dirs, 1.0):
        kvec = dirs[int(r[0])]
        k, _, _, _ = O.make_dir(kvec)
        o = O.origin(kvec, O.nogrid(shape, 1.0), 1.0, int(r[1])) + centre
        path = O.ray_path(shape, kvec, 1.0, int(r[1]))
        ijk = [np.array([v % X, (v // X) % Y, v // (X * Y)], dtype=float) for v in path]
        # projection of each visited centre onto the line and its gradient
        proj = [o + np.dot(c - o, k) * k for c in ijk]
        gl = [-rho * _tri_np(acc, x) for x in proj]
        gv = [-rho * np.array([acc[a].reshape(-1)[v] for a in range(3)]) for v in path]
        pin = pf[path[0]]
        pn[path[0]] += pin
        cnt[path[0]] += 1
        line = 0.5 * np.dot(gv[0] + gl[0], (proj[0] - ijk[0]) * dl)
        for n in range(1, len(path)):
            line += 0.5 * np.dot(gl[n - 1] + gl[n], (proj[n] - proj[n - 1]) * dl)
            leg = 0.5 * np.dot(gl[n] + gv[n], (ijk[n] - proj[n]) * dl)
            pn[path[n]] += pin + line + leg
            cnt[path[n]] += 1
    return pn.reshape(shape), cnt.reshape(shape)


@pytest.mark.parametrize("seed", [0, 1])
def test_trilinear_brute_force(ref_dirs2, seed):
    shape = (6, 6, 6)
    rng = np.random.default_rng(seed)
    acc = [rng.normal(size=shape) for _ in range(3)]
    p = rng.normal(size=shape)
    delta, rho = (0.5, 0.8, 0.3), 1.2
    dirs = _dirs_mix(ref_dirs2)
    pn, cnt, _ = O.one_pass(shape, dirs, 1.0, delta, rho, acc, p, method=O.PLAIN_TRILINEAR)
    bpn, bcnt = _brute_trilinear_pass(shape, dirs, delta, rho, acc, p)
    np.testing.assert_array_equal(cnt, bcnt)
    np.testing.assert_allclose(pn, bpn, rtol=1e-11, atol=1e-11 * np.abs(bpn).max())


def test_trilinear_visits_and_linearity(ref_dirs3):
    shape = (10, 9, 11)
    f = F.taylor_green(11, noise=0.1, seed=2, shape=shape)
    p = np.random.default_rng(3).normal(size=shape)
    pn0, c0, s0 = O.one_pass(shape, ref_dirs3, 1.0, f["delta"], 1.0, f["acc"], p)
    pn1, c1, s1 = O.one_pass(shape, ref_dirs3, 1.0, f["delta"], 1.0, f["acc"], p,
                             method=O.PLAIN_TRILINEAR)
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_array_equal(s0, s1)
    # linear in (field, p) jointly
    a2 = [2.5 * a for a in f["acc"]]
    pn2, _, _ = O.one_pass(shape, ref_dirs3, 1.0, f["delta"], 1.0, a2, 2.5 * p,
                           method=O.PLAIN_TRILINEAR)
    np.testing.assert_allclose(pn2, 2.5 * pn1, rtol=1e-12, atol=1e-12 * np.abs(pn1).max())
    # a different discretisation from the staircase on a noisy field
    assert not np.allclose(pn0, pn1)
