"""GPU parity of the plain-method reconstruction against the FP64 oracle.

Every test calls libomni3d.so through its C-ABI on the GPU and compares with
the oracle (oracle/omni3d_oracle.c) on the same seeded input.  Tolerances:
FP64 path rel-L2 <= 1e-12 (the only difference is the fixed-point vs FP64
summation order); FP32-Kahan path rel-L2 <= 1e-5 against the FP64 oracle;
visit counts bit-exact.
"""
import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def run_gpu(f, dirs, shape, precision=omni3d.FP64, max_iter=5, tol=0.0, p0=None, seed_pn=1,
            ls=1.0, dtype=np.float64):
    with omni3d.Context(shape, f["delta"], density=f["rho"], line_spacing=ls,
                        precision=precision, max_iterations=max_iter, tolerance=tol,
                        seed_pn_with_p=seed_pn) as ctx:
        ctx.set_directions(dirs)
        ctx.set_field(f["acc"], dtype=dtype)
        return ctx.reconstruct(p0)


def run_oracle(f, dirs, shape, max_iter=5, tol=0.0, p0=None, seed_pn=1, ls=1.0):
    return O.reconstruct(shape, dirs, ls, f["delta"], f["rho"], f["acc"], p0=p0,
                         max_iter=max_iter, tol=tol, seed_pn=seed_pn)


@pytest.mark.parametrize("shape", [(9, 10, 11), (16, 12, 7)])
def test_quadratic_exact_seed(gpu_available, ref_dirs2, shape):
    print(f"case quadratic {shape}", flush=True)
    f = F.quadratic(shape, seed=3)
    p, w, rep = run_gpu(f, ref_dirs2, shape, max_iter=3, p0=f["p"], seed_pn=0)
    po, co, it, rel = run_oracle(f, ref_dirs2, shape, max_iter=3, p0=f["p"], seed_pn=0)
    assert rep.iterations == it == 3
    np.testing.assert_array_equal(w, co)
    assert F.rel_l2(p, po) < 1e-12
    # trapezoid integrates a linear gradient exactly: p is reproduced
    assert np.abs(p - f["p"]).max() < 1e-9 * max(1.0, np.abs(f["p"]).max())


@pytest.mark.parametrize("dirs_name", ["ref2", "gen2", "gen3"])
def test_taylor_green_fp64_matches_oracle(gpu_available, ref_dirs2, dirs_name):
    shape = (20, 20, 20)
    print(f"case taylor-green fp64 {dirs_name}", flush=True)
    dirs = ref_dirs2 if dirs_name == "ref2" else omni3d.generate_directions(int(dirs_name[-1]))
    f = F.taylor_green(20, noise=0.05, seed=1)
    p, w, rep = run_gpu(f, dirs, shape, max_iter=12)
    po, co, it, rel = run_oracle(f, dirs, shape, max_iter=12)
    assert rep.iterations == it
    np.testing.assert_array_equal(w, co)
    assert F.rel_l2(p, po) < 1e-12
    assert abs(rep.rel_change - rel) <= 1e-9 * max(rel, 1e-30)


def test_seeded_pn_and_tolerance_stop(gpu_available, ref_dirs2):
    shape = (14, 15, 16)
    print("case seed+tolerance", flush=True)
    f = F.taylor_green(16, noise=0.02, seed=2, shape=shape)
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=shape)
    p, w, rep = run_gpu(f, ref_dirs2, shape, max_iter=60, tol=1e-4, p0=p0, seed_pn=1)
    po, co, it, rel = run_oracle(f, ref_dirs2, shape, max_iter=60, tol=1e-4, p0=p0, seed_pn=1)
    assert rep.iterations == it
    assert F.rel_l2(p, po) < 1e-11


def test_fp32_kahan_close_to_fp64_oracle(gpu_available, ref_dirs3):
    shape = (24, 24, 24)
    print("case fp32 kahan", flush=True)
    f = F.taylor_green(24, noise=0.05, seed=4)
    p, w, rep = run_gpu(f, ref_dirs3, shape, precision=omni3d.FP32_KAHAN, max_iter=10)
    po, co, it, rel = run_oracle(f, ref_dirs3, shape, max_iter=10)
    np.testing.assert_array_equal(w, co)
    err = F.rel_l2(p, po)
    print("fp32-kahan rel-L2 vs FP64 oracle", err)
    assert err < 1e-5


def test_repeatable_bitwise(gpu_available, ref_dirs3):
    shape = (18, 17, 19)
    print("case determinism", flush=True)
    f = F.taylor_green(18, noise=0.05, seed=6, shape=shape)
    a = run_gpu(f, ref_dirs3, shape, max_iter=8)[0]
    b = run_gpu(f, ref_dirs3, shape, max_iter=8)[0]
    np.testing.assert_array_equal(a, b)


def test_ray_engine_matches_oracle(gpu_available, ref_dirs2, monkeypatch):
    """OMNI3D_ENGINE=rays: one lane per ray with global fixed-point atomics
    instead of the tile engine, same results."""
    shape = (15, 14, 13)
    print("case ray engine", flush=True)
    monkeypatch.setenv("OMNI3D_ENGINE", "rays")
    f = F.taylor_green(15, noise=0.05, seed=9, shape=shape)
    p, w, rep = run_gpu(f, ref_dirs2, shape, max_iter=7)
    po, co, it, rel = run_oracle(f, ref_dirs2, shape, max_iter=7)
    np.testing.assert_array_equal(w, co)
    assert F.rel_l2(p, po) < 1e-12


# the fallback layouts of the default build: dense words one chunk per
# group (24-bit sources, for more than 2^22 sources) and the masked
# sliced-ELL slots (rows of 4096 or more entries, e.g. thousands of
# directions); and p_S gathers instead of FP32 increments
LAYOUTS = [
    {"OMNI3D_DENSE_GROUP": "1"},
    {"OMNI3D_ITER": "slots"},
    {"OMNI3D_NO_INCREMENTS": "1"},
]


@pytest.mark.parametrize("env", LAYOUTS, ids=lambda e: ",".join(f"{k[7:]}={v}" for k, v in e.items()))
@pytest.mark.parametrize("precision", [omni3d.FP64, omni3d.FP32_KAHAN])
def test_iteration_layouts_agree(gpu_available, ref_dirs3, monkeypatch, env, precision):
    """Every fallback layout of the source iteration (dense words per chunk,
    masked sliced-ELL slots, p_S gathers instead of increments) against the
    default (merged words in groups of eight chunks) from a random initial
    pressure: FP64 rel-L2 <= 1e-12, FP32 <= 1e-6, same iteration count."""
    shape = (32, 34, 36)
    print(f"case iteration layout {env} precision={precision}", flush=True)
    f = F.taylor_green(36, noise=0.05, seed=13, shape=shape)
    p0 = np.random.default_rng(14).normal(size=shape)
    dtype = np.float64 if precision == omni3d.FP64 else np.float32

    def go():
        with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision,
                            max_iterations=25, tolerance=0.0) as ctx:
            ctx.set_directions(ref_dirs3)
            ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], dtype=dtype)
            return ctx.reconstruct(p0)

    pd, wd, rd = go()
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    pv, wv, rv = go()
    assert rv.iterations == rd.iterations == 25
    np.testing.assert_array_equal(wv, wd)
    err = F.rel_l2(pv, pd)
    print("layout vs default rel-L2", err)
    assert err < (1e-12 if precision == omni3d.FP64 else 1e-6)


@pytest.mark.parametrize("precision", [omni3d.FP64, omni3d.FP32_KAHAN])
def test_word_groups_of_four_and_eight_chunks_agree_bitwise(gpu_available, ref_dirs3,
                                                            monkeypatch, precision):
    """The merged words in groups of 4 chunks (256 rows, 8-bit rows) against
    the default groups of 8 chunks (512 rows, 9-bit rows).  A row's sum is an
    exact 64-bit fixed-point sum whatever group holds it, so p is bit-equal
    (only the stop rule's partial sums are added in another order: tolerance
    0 here).  86 k sources, so that the 256- and 512-row caps split the rows
    differently (smaller grids give both the same CU-filling groups)."""
    shape = (120, 112, 128)
    print(f"case word groups 4 vs 8 precision={precision}", flush=True)
    f = F.taylor_green(128, noise=0.05, seed=21, shape=shape)
    p0 = np.random.default_rng(22).normal(size=shape)
    dtype = np.float64 if precision == omni3d.FP64 else np.float32

    def go():
        with omni3d.Context(shape, f["delta"], density=f["rho"], precision=precision,
                            max_iterations=12, tolerance=0.0) as ctx:
            ctx.set_directions(ref_dirs3)
            ctx.set_field([np.asarray(a, dtype=dtype) for a in f["acc"]], dtype=dtype)
            return ctx.reconstruct(p0)

    p8, w8, r8 = go()
    monkeypatch.setenv("OMNI3D_DENSE_GROUP", "4")
    p4, w4, r4 = go()
    assert r4.iterations == r8.iterations == 12
    assert r4.sources == r8.sources > 65536
    np.testing.assert_array_equal(w4, w8)
    np.testing.assert_array_equal(p4, p8)


@pytest.mark.parametrize("shape", [(2, 9, 7), (12, 2, 5), (11, 9, 3)])
def test_thin_grids_match_oracle(gpu_available, ref_dirs2, shape):
    """Grids two or three voxels thick, where voxels lie on several box faces
    at once: the source numbering (a voxel belongs to the first of its faces),
    the tiles' partial edges and the visit counts against the oracle (counts
    bit-exact, FP64 rel-L2 <= 1e-12)."""
    print(f"case thin grid {shape}", flush=True)
    f = F.taylor_green(max(shape), noise=0.05, seed=9, shape=shape)
    p, w, rep = run_gpu(f, ref_dirs2, shape, max_iter=4)
    po, co, it, rel = run_oracle(f, ref_dirs2, shape, max_iter=4)
    assert rep.iterations == it == 4
    np.testing.assert_array_equal(w, co)
    assert w.max() > 0
    assert F.rel_l2(p, po) < 1e-12
