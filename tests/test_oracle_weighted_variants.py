"""Pins the oracle's two integrators that no reference launcher calls (CPU only).

WeightedFixedBC (bodyIntegralInnerWeightedFixedBC, kernel.cu:2136-2279) is
bodyIntegralInnerWeighted with one changed bound: the y-branch needs
jnext2 > 0 (kernel.cu:2226) where Weighted has jnext2 >= 0, so a step onto the
j = 0 plane runs no branch and ends the walk; x- and z-steps keep their
bounds (unlike SelectFixedBC) and ties keep the plain order x, y, z.

WeightedMiniCurl (bodyIntegralInnerWeightedMiniCurl, kernel.cu:2937-3085)
walks MiniCurl's path (the candidate with the smallest curl, ties x, y, z)
with curl weights that differ from Weighted's: curltmp starts at the entry's
curl and counttmp at 1, and the entry adds its own p with weight
1 / curltmp * counttmp (kernel.cu:2956-2960); each step adds the next curl
to curltmp (no / 1e4) and 1 to counttmp, and adds (p_entry + pint) *
1 / curltmp * counttmp, with no test of curltmp (kernel.cu:3029-3060).

Cases: single rays along a grid axis (the setup of test_oracle_weights: one
accepted ray, the interior row from voxel 0 to 7, a constant gradient G
along it so every step's term is -T, T = rho * delta * G per unit step)
with the sums and weights worked out by hand; the WeightedFixedBC paths
against a brute-force walker written from the branch conditions; a quadratic
pressure reproduced from its exact seed by both (any positive weights
average exact values).  test_mutants_fail compiles one-line mutations of the
two walks and requires each to fail a case.
"""
import os

import numpy as np
import pytest

import oracle as O
from omni3d import fields as F
from test_oracle_weights import (AXES, DELTA, G, N, P, RHO, T, _build_mutant, _line, _row,
                                 expect)
from test_oracle_fixedbc import to_index


def run_dir(axis, direction, curl_line, method, L=None):
    """One pass of the single ray of AXES[axis] walked along `direction`
    (+1 or -1 along the axis); returns (pn, cnt, src) along the row."""
    shape, d = AXES[axis]
    kv = np.array([[direction * c for c in d]], dtype=np.float32)
    acc = [np.zeros(shape) for _ in range(3)]
    acc[axis][:] = G
    pn, cnt, src = O.one_pass(shape, kv, 1.0, (DELTA,) * 3, RHO, acc, _line(axis, P),
                              curl=_line(axis, curl_line), method=method, L=L)
    rows = [_row(axis, a).copy() for a in (pn, cnt, src)]
    for a in (pn, cnt, src):
        _row(axis, a)[:] = 0
        assert not a.any(), "a visit off the row"
    return rows


# --- WeightedFixedBC ------------------------------------------------------

def case_wfbc_down_y(L=None):
    # walked from j = 7 toward j = 0: each step s = -1 adds +T; the step onto
    # j = 0 is refused, so visits are j = 6..1 (Weighted reaches j = 0)
    c = 2.0
    pn, cnt, src = run_dir(1, -1, [c] * N, O.WEIGHTED_FIXEDBC, L)
    expect(pn, cnt, src, {j: (P[N - 1] + (N - 1 - j) * T, 1e4 / c) for j in range(1, N - 1)},
           [N - 1])
    pn, cnt, src = run_dir(1, -1, [c] * N, O.WEIGHTED, L)
    expect(pn, cnt, src, {j: (P[N - 1] + (N - 1 - j) * T, 1e4 / c) for j in range(0, N - 1)},
           [N - 1])


def case_wfbc_up_y_and_x_z(L=None):
    # from j = 0 upward nothing is refused; x and z rays are Weighted's walks
    ramp = 1.0 + np.arange(N)
    for axis in range(3):
        pn, cnt, src = run_dir(axis, +1, ramp, O.WEIGHTED_FIXEDBC, L)
        expect(pn, cnt, src, {v: (P[0] - v * T, 2e4 / (v + 3)) for v in range(1, N)}, [0])


# --- WeightedMiniCurl -----------------------------------------------------

def case_wmc_ramp(L=None):
    # curl 1 + v: curltmp at voxel n = sum_{m=0..n} (1 + m) = (n + 1)(n + 2) / 2,
    # counttmp n + 1, weight 2 / (n + 2); the entry (n = 0) adds P[0] with 1
    for axis in range(3):
        pn, cnt, src = run_dir(axis, +1, 1.0 + np.arange(N), O.WEIGHTED_MINICURL, L)
        expect(pn, cnt, src, {v: (P[0] - v * T, 2.0 / (v + 2)) for v in range(N)}, [0])


def case_wmc_steps(L=None):
    # curl 4 at the entry, then 1: curltmp 4 + n, counttmp n + 1
    curl = [4.0] + [1.0] * (N - 1)
    pn, cnt, src = run_dir(0, +1, curl, O.WEIGHTED_MINICURL, L)
    expect(pn, cnt, src, {v: (P[0] - v * T, (v + 1) / (4.0 + v)) for v in range(N)}, [0])


def case_wfbc_counts_vs_paths(L=None):
    # a uniform curl c weighs every visit after the entry 1e4 / c: one pass's
    # counts are the visits of brute_wfbc's paths (every accepted ray of the
    # Divide2 directions on a 5 x 6 x 7 box, which has no exact distance tie)
    from conftest import read_direction_fixture
    shape = (5, 6, 7)
    dirs = read_direction_fixture("GridIcosahedronDivide2.dat")
    c = 2.0
    acc = [np.zeros(shape) for _ in range(3)]
    _, cnt, _ = O.one_pass(shape, dirs, 1.0, (DELTA,) * 3, RHO, acc, np.zeros(shape),
                           curl=np.full(shape, c), method=O.WEIGHTED_FIXEDBC, L=L)
    want = np.zeros(shape).reshape(-1)
    for r in O.enumerate_rays(shape, dirs, 1.0):
        path, _ = brute_wfbc(shape, dirs[int(r[0])], 1.0, int(r[1]))
        for ijk in path[1:]:
            want[to_index(shape, ijk)] += 1e4 / c
    np.testing.assert_allclose(cnt.reshape(-1), want, rtol=1e-12, atol=0)


CASES = [case_wfbc_down_y, case_wfbc_up_y_and_x_z, case_wfbc_counts_vs_paths, case_wmc_ramp,
         case_wmc_steps]


@pytest.mark.parametrize("case", CASES, ids=[c.__name__ for c in CASES])
def test_case(case):
    case()


def brute_wfbc(shape, kv, ls, point):
    """The branch conditions of kernel.cu:2211-2242 with Euclidean distances
    of the candidate centres to the line: x when d1 <= d2, d3; y when d2 < d1
    and d2 <= d3 and the next j > 0; z when d3 < d1, d2; a refused step ends
    the walk."""
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
        if d[0] <= d[1] and d[0] <= d[2]:
            ax = 0
        elif d[1] < d[0] and d[1] <= d[2]:
            ax = 1
            if c[1] + s[1] <= 0:
                break
        elif d[2] < d[0] and d[2] < d[1]:
            ax = 2
        else:
            break
        c[ax] += s[ax]
        path.append(c.copy())
        dists.append(d)
    return path, dists


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 8, 11)])
def test_wfbc_paths_match_brute_force(shape, ref_dirs2):
    rays = O.enumerate_rays(shape, ref_dirs2, 1.0)
    ties = differ = 0
    for r in rays:
        d, pt = int(r[0]), int(r[1])
        got = list(O.ray_path(shape, ref_dirs2[d], 1.0, pt, fixedbc=2))
        bp, dists = brute_wfbc(shape, ref_dirs2[d], 1.0, pt)
        want = [to_index(shape, c) for c in bp]
        if got != list(O.ray_path(shape, ref_dirs2[d], 1.0, pt)):
            differ += 1
        if got == want:
            continue
        j = next((i for i in range(min(len(got), len(want))) if got[i] != want[i]),
                 min(len(got), len(want)))
        dd = [v for v in dists[j - 1] if np.isfinite(v)]
        assert len(dd) > 1 and np.ptp(sorted(dd)[:2]) <= 1e-9, (shape, d, pt, j)
        ties += 1
    assert ties <= 0.01 * len(rays)
    assert differ > 0  # rays stepping onto j = 0 are cut


def test_quadratic_exact_both(ref_dirs2):
    shape = (9, 10, 11)
    q = F.quadratic(shape, seed=2)
    cq = O.curl(shape, q["delta"], q["acc"]) + 1.0  # positive everywhere
    for method in (O.WEIGHTED_FIXEDBC, O.WEIGHTED_MINICURL):
        p, cnt, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, q["delta"], q["rho"], q["acc"],
                                     p0=q["p"], curl=cq, method=method, max_iter=3, seed_pn=0)
        m = cnt > 0
        assert m.mean() > 0.5
        assert np.abs(p[m] - q["p"][m]).max() <= 1e-9 * np.abs(q["p"]).max()


MUTANTS = [
    ("wfbc: the y bound of Weighted", "if (w->fixedbc == 2 && a == 1 && w->c[1] + st[1] == 0) return -1;",
     "if (w->fixedbc == 2 && a == 1 && w->c[1] + st[1] < 0) return -1;"),
    ("wfbc: SelectFixedBC's rule", "          w.fixedbc = 2;\n          walk_weighted(&P, &w);",
     "          w.fixedbc = 1;\n          walk_weighted(&P, &w);"),
    ("wmc: no entry self-add", "  P->pn[vin] += pin * 1 / curltmp * counttmp;\n  P->cnt[vin] += 1 / curltmp * counttmp;\n  double pint = 0.0;\n  int steps = 0;\n  while (!walker_done(w)) {\n    const int64_t v = vidx(X, Y, w->c[0], w->c[1], w->c[2]);\n    int st[3];\n    double d[3];\n    for (int a = 0; a < 3; a++) {\n      st[a] = w->c[a] < w->e[a] ? 1 : (w->c[a] > w->e[a] ? -1 : -60000);\n      const int64_t idx = v + (int64_t)st[a] * stride[a];\n      d[a] = (idx >= 0 && idx < V) ? P->curl[idx] : 1e10;\n    }\n    int a;\n    if (d[0] <= d[1] && d[0] <= d[2]) a = 0;\n    else if (d[1] < d[0] && d[1] <= d[2]) a = 1;\n    else if (d[2] < d[0] && d[2] < d[1]) a = 2;\n    else break;\n    const int n = w->c[a] + st[a];\n    if (n < 0 || n >= size[a]) break;\n    w->c[a] = n;\n    const int64_t vn = v + (int64_t)st[a] * stride[a];\n    pint += trap(P->rho, st[a], P->del[a], P->acc[a][vn], P->acc[a][v]);\n    curltmp += P->curl[vn];",
     "  double pint = 0.0;\n  int steps = 0;\n  while (!walker_done(w)) {\n    const int64_t v = vidx(X, Y, w->c[0], w->c[1], w->c[2]);\n    int st[3];\n    double d[3];\n    for (int a = 0; a < 3; a++) {\n      st[a] = w->c[a] < w->e[a] ? 1 : (w->c[a] > w->e[a] ? -1 : -60000);\n      const int64_t idx = v + (int64_t)st[a] * stride[a];\n      d[a] = (idx >= 0 && idx < V) ? P->curl[idx] : 1e10;\n    }\n    int a;\n    if (d[0] <= d[1] && d[0] <= d[2]) a = 0;\n    else if (d[1] < d[0] && d[1] <= d[2]) a = 1;\n    else if (d[2] < d[0] && d[2] < d[1]) a = 2;\n    else break;\n    const int n = w->c[a] + st[a];\n    if (n < 0 || n >= size[a]) break;\n    w->c[a] = n;\n    const int64_t vn = v + (int64_t)st[a] * stride[a];\n    pint += trap(P->rho, st[a], P->del[a], P->acc[a][vn], P->acc[a][v]);\n    curltmp += P->curl[vn];"),
    ("wmc: curl / 1e4 as Weighted", "    curltmp += P->curl[vn];\n    counttmp += 1.0;\n    P->pn[vn] += (pin + pint) * 1 / curltmp * counttmp;",
     "    curltmp += P->curl[vn] / 10000;\n    counttmp += 1.0;\n    P->pn[vn] += (pin + pint) * 1 / curltmp * counttmp;"),
    ("wmc: counttmp from 0", "  double curltmp = P->curl[vin], counttmp = 1.0;",
     "  double curltmp = P->curl[vin], counttmp = 0.0;"),
    ("wmc: curltmp from 0", "  double curltmp = P->curl[vin], counttmp = 1.0;",
     "  double curltmp = 0.0 * P->curl[vin] + 1.0, counttmp = 1.0;"),
]


def test_mutants_fail(tmp_path):
    with open(os.path.join(O.HERE, "omni3d_oracle.c")) as f:
        src = f.read()
    survivors = []
    for name, old, new in MUTANTS:
        L = _build_mutant(str(tmp_path), src, old, new)
        failed = False
        for case in CASES:
            try:
                with np.errstate(all="ignore"):
                    case(L)
            except AssertionError:
                failed = True
                break
        if not failed:
            survivors.append(name)
    assert not survivors, f"mutations no case detects: {survivors}"
