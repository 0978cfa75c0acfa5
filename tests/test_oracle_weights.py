"""Pins the Select / Weighted weight rule and branch structure of the oracle
with single-ray cases whose expected sums and weights are worked out by hand
(CPU only).

Setup.  One direction along a grid axis, in a box 3 voxels wide across that
axis and N = 8 long along it, gives exactly one accepted ray: the interior
row (the reference's axis-aligned cross2point cases take only lines strictly
inside the box), walked from voxel 0 to voxel 7.  The acceleration component
along the axis is a constant G, so every step's trapezoid term
-rho*(+1)*delta*0.5*(G + G) is -T with T = rho*delta*G, and a run that
(re)started at voxel s carries P[s] - (v - s)*T at voxel v.  P is a known
pressure along the row, the curl (|curl a|*mask) is given along the row.

Expected weights, from the reference's arithmetic: every visit that adds
increments counttmp by 1 and curltmp by curl/1e4 and adds (p0 + pint)/curltmp
* counttmp to pn and 1/curltmp * counttmp to the count (kernel.cu:2071-2079
Weighted, 2385-2392 Select).  So a uniform curl c weighs every visit 1e4/c;
a curl 1 + v along the row weighs the n-th visit of a run from voxel 0
n / (sum_{m=1..n} (1+m) / 1e4) = 2e4 / (n + 3).

Branch rules pinned (Select, kernel.cu:2280-2537):
  * the entry is a start when curl != 0 and curl <= thr (kernel.cu:2302);
    later voxels are valid when curl != 0 and curl < thr;
  * a step onto a voxel with curl >= thr enters the high-error zone, a later
    valid voxel restarts the run there (p0 = p at that voxel, pint = 0,
    nothing added at the restart voxel);
  * faithful rule: an x-step onto curl == 0 enters the high-error zone
    (kernel.cu:2394), a y- or z-step onto it does nothing -- no trapezoid
    term, no break (kernel.cu:2442, 2488); the first start after an invalid
    entry needs only curl != 0 on a y- or z-step (kernel.cu:2422, 2468);
    a z-restart keeps counttmp and curltmp (kernel.cu:2497-2506) while x-
    and y-restarts reset them (kernel.cu:2404-2410, 2450-2459);
  * symmetric rule (faithful = 0): every axis as the x-branch, entry as the
    later voxels.
Weighted (kernel.cu:1993-2135): no breaks, no entry self-add, and no scatter
while curltmp == 0 (counttmp still counts those visits).

test_mutants_fail compiles one-line mutations of the oracle's Select /
Weighted walks on the CPU and requires each to fail at least one case.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

N = 8
RHO, DELTA, G = 1.3, 0.5, 0.3
T = RHO * DELTA * G
P = 1.0 + 0.25 * np.arange(N) ** 1.5  # pressure along the row (any distinct values)
AXES = {0: ((3, 3, N), (1.0, 0.0, 0.0)),
        1: ((3, N, 3), (0.0, 1.0, 0.0)),
        2: ((N, 3, 3), (0.0, 0.0, 1.0))}


def _line(axis, values):
    shape, _ = AXES[axis]
    v = np.asarray(values, dtype=float)
    if axis == 0:
        return np.broadcast_to(v[None, None, :], shape).copy()
    if axis == 1:
        return np.broadcast_to(v[None, :, None], shape).copy()
    return np.broadcast_to(v[:, None, None], shape).copy()


def _row(axis, a):
    return a[1, 1, :] if axis == 0 else (a[1, :, 1] if axis == 1 else a[:, 1, 1])


def run(axis, curl_line, method, faithful, thr, L=None):
    """One pass of the single ray; returns (pn, cnt, src) along the row and
    checks that nothing off the row was touched."""
    shape, d = AXES[axis]
    acc = [np.zeros(shape) for _ in range(3)]
    acc[axis][:] = G
    pn, cnt, src = O.one_pass(shape, np.array([d], dtype=np.float32), 1.0, (DELTA,) * 3, RHO,
                              acc, _line(axis, P), curl=_line(axis, curl_line), threshold=thr,
                              method=method, faithful=faithful, L=L)
    rows = [_row(axis, a).copy() for a in (pn, cnt, src)]
    for a in (pn, cnt, src):
        _row(axis, a)[:] = 0
        assert not a.any(), "a visit off the row"
    return rows


def expect(pn, cnt, src, visits, sources):
    """visits: {v: (carried value, weight)}; every other voxel gets nothing."""
    epn, ecnt = np.zeros(N), np.zeros(N)
    for v, (val, w) in visits.items():
        epn[v] = val * w
        ecnt[v] = w
    np.testing.assert_allclose(cnt, ecnt, rtol=1e-12, atol=0)
    np.testing.assert_allclose(pn, epn, rtol=1e-12, atol=1e-9)
    assert sorted(np.nonzero(src)[0].tolist()) == sorted(sources)


# --- Weighted ------------------------------------------------------------

def case_weighted_uniform(L=None):
    c = 2.5
    for axis in range(3):
        pn, cnt, src = run(axis, [c] * N, O.WEIGHTED, 1, 0.0, L)
        expect(pn, cnt, src, {v: (P[0] - v * T, 1e4 / c) for v in range(1, N)}, [0])


def case_weighted_ramp(L=None):
    for axis in range(3):
        pn, cnt, src = run(axis, 1.0 + np.arange(N), O.WEIGHTED, 1, 0.0, L)
        expect(pn, cnt, src, {v: (P[0] - v * T, 2e4 / (v + 3)) for v in range(1, N)}, [0])


def case_weighted_zero_start(L=None):
    # curl 0 at voxel 1: curltmp == 0 there, no scatter; counttmp still counts
    # it, so voxel v >= 2 has counttmp v and curltmp (v - 1)/1e4
    curl = [1.0] * N
    curl[1] = 0.0
    pn, cnt, src = run(0, curl, O.WEIGHTED, 1, 0.0, L)
    expect(pn, cnt, src, {v: (P[0] - v * T, 1e4 * v / (v - 1)) for v in range(2, N)}, [0])


# --- Select --------------------------------------------------------------

def case_select_uniform(L=None):
    c = 2.5
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, [c] * N, O.SELECT, faithful, 10.0, L)
            expect(pn, cnt, src, {v: (P[0] - v * T, 1e4 / c) for v in range(1, N)}, [0])


def case_select_restart_weights(L=None):
    # curl 1 + v with a high-error voxel at 3 (curl 20 >= thr 10): voxels 1, 2
    # add with weights 1/2e-4 and 2/5e-4; 3 adds nothing and enters the zone;
    # 4 restarts (p0 = P[4], nothing added); 5..7 carry P[4] - (v - 4) T.
    # A z-restart keeps counttmp = 2, curltmp = 5e-4; x, y (and the symmetric
    # rule on z) start again from 0.
    curl = 1.0 + np.arange(N)
    curl[3] = 20.0
    keep = {5: 3 / 11e-4, 6: 4 / 18e-4, 7: 5 / 26e-4}
    reset = {5: 1 / 6e-4, 6: 2 / 13e-4, 7: 3 / 21e-4}
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, curl, O.SELECT, faithful, 10.0, L)
            w = keep if (faithful and axis == 2) else reset
            visits = {1: (P[0] - T, 5000.0), 2: (P[0] - 2 * T, 4000.0)}
            visits.update({v: (P[4] - (v - 4) * T, w[v]) for v in (5, 6, 7)})
            expect(pn, cnt, src, visits, [0, 4])


def case_select_zero_curl_step(L=None):
    # curl 1 except 0 at voxel 3.  Faithful y/z: the step onto 3 adds nothing
    # and loses its trapezoid term, the run goes on (counttmp 3 at voxel 4):
    # voxel v >= 4 carries P[0] - (v - 1) T with weight 1e4.  x-steps and the
    # symmetric rule: 3 enters the zone, 4 restarts, v >= 5 carry
    # P[4] - (v - 4) T.
    curl = [1.0] * N
    curl[3] = 0.0
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, curl, O.SELECT, faithful, 10.0, L)
            visits = {1: (P[0] - T, 1e4), 2: (P[0] - 2 * T, 1e4)}
            if faithful and axis != 0:
                visits.update({v: (P[0] - (v - 1) * T, 1e4) for v in range(4, N)})
                expect(pn, cnt, src, visits, [0])
            else:
                visits.update({v: (P[4] - (v - 4) * T, 1e4) for v in range(5, N)})
                expect(pn, cnt, src, visits, [0, 4])


def case_select_later_at_threshold(L=None):
    # curl 1 except curl == thr at voxel 3: not valid (curl < thr fails) and
    # entering the zone (curl >= thr) on every axis and under both rules;
    # 4 restarts
    curl = [1.0] * N
    curl[3] = 10.0
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, curl, O.SELECT, faithful, 10.0, L)
            visits = {1: (P[0] - T, 1e4), 2: (P[0] - 2 * T, 1e4)}
            visits.update({v: (P[4] - (v - 4) * T, 1e4) for v in range(5, N)})
            expect(pn, cnt, src, visits, [0, 4])


def case_select_entry_at_threshold(L=None):
    # entry curl == thr: a start under the faithful entry test (curl > thr
    # rejects), not under the symmetric one (curl < thr accepts)
    curl = [1.0] * N
    curl[0] = 10.0
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, curl, O.SELECT, faithful, 10.0, L)
            if faithful:
                expect(pn, cnt, src, {v: (P[0] - v * T, 1e4) for v in range(1, N)}, [0])
            else:
                expect(pn, cnt, src, {v: (P[1] - (v - 1) * T, 1e4) for v in range(2, N)}, [1])


def case_select_first_start_ignores_threshold(L=None):
    # entry curl 0 (no start), voxel 1 curl 20 >= thr, then curl 1.  Faithful
    # y/z start at voxel 1 (only curl != 0 is tested there); x and the
    # symmetric rule start at voxel 2.
    curl = [1.0] * N
    curl[0], curl[1] = 0.0, 20.0
    for faithful in (0, 1):
        for axis in range(3):
            pn, cnt, src = run(axis, curl, O.SELECT, faithful, 10.0, L)
            s = 1 if (faithful and axis != 0) else 2
            expect(pn, cnt, src, {v: (P[s] - (v - s) * T, 1e4) for v in range(s + 1, N)}, [s])


def case_plain_mask_break(L=None):
    # plain-mask: the Select walk on the mask with unit weights and an
    # infinite threshold; a masked voxel breaks the run, the next valid
    # voxel restarts it, no entry self-add
    mask = [1.0] * N
    mask[2] = 0.0
    for axis in range(3):
        pn, cnt, src = run(axis, mask, O.PLAIN_MASK, 0, 0.0, L)
        visits = {1: (P[0] - T, 1.0)}
        visits.update({v: (P[3] - (v - 3) * T, 1.0) for v in range(4, N)})
        expect(pn, cnt, src, visits, [0, 3])


CASES = [case_weighted_uniform, case_weighted_ramp, case_weighted_zero_start,
         case_select_uniform, case_select_restart_weights, case_select_zero_curl_step,
         case_select_later_at_threshold, case_select_entry_at_threshold, case_select_first_start_ignores_threshold,
         case_plain_mask_break]


@pytest.mark.parametrize("case", CASES, ids=[c.__name__ for c in CASES])
def test_hand_worked(case):
    case()


# One-line mutations of omni3d_oracle.c's Select / Weighted walks (text to
# find, replacement).  Each must fail at least one case above.
MUTANTS = [
    ("weighted: curl not scaled by 1e4",
     "curltmp += P->curl[vn] / 10000;", "curltmp += P->curl[vn];"),
    ("weighted: weight without counttmp",
     "      P->pn[vn] += (pin + pint) * 1 / curltmp * counttmp;\n      P->cnt[vn]",
     "      P->pn[vn] += (pin + pint) / curltmp;\n      P->cnt[vn]"),
    ("weighted: count without counttmp",
     "counttmp;\n      P->cnt[vn] += 1 / curltmp * counttmp;",
     "counttmp;\n      P->cnt[vn] += 1 / curltmp;"),
    ("select: count without counttmp",
     "counttmp;\n          P->cnt[vn] += 1 / curltmp * counttmp;",
     "counttmp;\n          P->cnt[vn] += 1 / curltmp;"),
    ("weighted: scatter while curltmp == 0 guarded by counttmp",
     "if (curltmp != 0) {", "if (counttmp != 0) {"),
    ("weighted: entry self-add",
     "double pint = 0.0, curltmp = 0.0, counttmp = 0.0;\n  int steps = 0;\n  while",
     "double pint = 0.0, curltmp = 0.0, counttmp = 0.0;\n  P->pn[vin] += pin;"
     " P->cnt[vin] += 1.0;\n  int steps = 0;\n  while"),
    ("select: curl not scaled by 1e4", "curltmp += cn / 10000;", "curltmp += cn;"),
    ("select: weight without counttmp",
     "P->pn[vn] += (p0 + pint) * 1 / curltmp * counttmp;", "P->pn[vn] += (p0 + pint) / curltmp;"),
    ("select: z-restart resets the weights",
     "if (!(faithful && a == 2)) { counttmp = 0; curltmp = 0; }",
     "{ counttmp = 0; curltmp = 0; }"),
    ("select: the y-restart keeps the weights",
     "if (!(faithful && a == 2)) { counttmp = 0; curltmp = 0; }",
     "if (!(faithful && a != 0)) { counttmp = 0; curltmp = 0; }"),
    ("select: restart keeps pint", "pint = 0;\n        outthr = 0;", "outthr = 0;"),
    ("select: faithful y/z break on zero curl",
     "? (!outthr && cn >= thr)", "? ((!outthr && cn >= thr) || cn == 0)"),
    ("select: faithful entry uses the later-voxel test",
     "!(ci == 0 || ci > thr)", "(ci != 0 && ci < thr)"),
    ("select: symmetric entry accepts curl == thr",
     ": (ci != 0 && ci < thr);", ": (ci != 0 && ci <= thr);"),
    ("select: faithful y/z first start tests the threshold",
     "? (cn != 0) : valid;", "? valid : valid;"),
    ("select: adds inside the high-error zone", "if (!outthr && valid) {", "if (valid) {"),
    ("select: the high-error zone starts at curl > thr",
     ": ((!outthr && cn >= thr) || cn == 0);", ": ((!outthr && cn > thr) || cn == 0);"),
]


def _build_mutant(tmp, src, old, new):
    assert src.count(old) == 1, f"mutation anchor not unique: {old!r}"
    d = os.path.join(tmp, f"m{abs(hash(old + new)) % 10**8}")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "omni3d_oracle.c"), "w") as f:
        f.write(src.replace(old, new))
    shutil.copy(os.path.join(O.HERE, "omni3d_oracle.h"), d)
    so = os.path.join(d, "liboracle_mutant.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(d, "omni3d_oracle.c"), "-lm"], check=True)
    return O.load(so)


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
