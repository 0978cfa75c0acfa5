"""GPU parity of the kernel-level and CPU-level drop-ins, called through the
C-ABI with the reference's argument lists:

  * omni3d_es_inner / _weighted / _select (kernel.cu:6151/6281/6414): one
    pass over every line x direction pair added into FP32 pn / pcountinner;
    against one oracle pass on the same FP32 inputs, pn within FP32 rounding
    (rel <= 1e-6 per voxel), pcountinner exact for the plain method;
  * omni3d_devide_count_inner (kernel.cu:6746-6759), omni3d_curl
    (kernel.cu:8352-8379, and without a mask kernel.cu:6863-6888);
  * calRHS, sor3D (same signatures as calRHS.h:4 and sor3D.h:5) and the
    red-black omni3d_sor_solve: bit-identical to the oracle's serial
    evaluation (every voxel update has the same expression order);
  * thredholdHistMaterialAccCPU (kernel.cu:8380-8418) and
    omni3d_threshold_hist: the curl threshold bit-identical to the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def es_call(which, shape, dirs, f, p, curl=None, thr=0.0, pn=None, pc=None):
    import torch
    Z, Y, X = shape
    k = [dev(dirs[:, i]) for i in range(3)]
    a = [dev(c) for c in f["acc"]]
    pd = dev(p)
    pnd = dev(np.zeros(shape) if pn is None else pn)
    pcd = dev(np.zeros(shape) if pc is None else pc)
    L = omni3d.lib()
    args = [X, Y, Z, len(dirs), 1.0, k[0].data_ptr(), k[1].data_ptr(), k[2].data_ptr(), None,
            f["delta"][0], f["delta"][1], f["delta"][2], f["rho"],
            a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr()]
    if which == "plain":
        rc = L.omni3d_es_inner(*args, pd.data_ptr(), pnd.data_ptr(), pcd.data_ptr(), None)
    elif which == "weighted":
        cd = dev(curl)
        rc = L.omni3d_es_inner_weighted(*args, cd.data_ptr(), pd.data_ptr(), pnd.data_ptr(),
                                        pcd.data_ptr(), None)
    else:
        cd = dev(curl)
        rc = L.omni3d_es_inner_select(*args, cd.data_ptr(), pd.data_ptr(), pnd.data_ptr(),
                                      pcd.data_ptr(), thr, None)
    omni3d.check(rc)
    torch.cuda.synchronize()
    return pnd.cpu().numpy().astype(np.float64), pcd.cpu().numpy().astype(np.float64)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


@pytest.mark.parametrize("which", ["plain", "weighted", "select"])
def test_es_inner_one_pass(gpu_available, ref_dirs2, which):
    shape = (12, 13, 14)
    print(f"case es_inner {which}", flush=True)
    f = F.taylor_green(14, noise=0.05, seed=7, shape=shape)
    f = dict(f, acc=[f32(a) for a in f["acc"]])
    rng = np.random.default_rng(1)
    p = f32(rng.normal(size=shape))
    curl = f32(O.curl(shape, f["delta"], f["acc"])) if which != "plain" else None
    thr = float(np.quantile(curl, 0.8)) if which == "select" else 0.0
    thr = float(np.float32(thr))
    method = {"plain": O.PLAIN, "weighted": O.WEIGHTED, "select": O.SELECT}[which]
    pn_o, cnt_o, _ = O.one_pass(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p,
                                curl=curl, threshold=thr, method=method, faithful=1)
    pn, pc = es_call(which, shape, ref_dirs2, f, p, curl, thr)
    if which == "plain":
        np.testing.assert_array_equal(pc, cnt_o)
    else:
        np.testing.assert_allclose(pc, cnt_o, rtol=1e-6, atol=1e-30)
    scale = np.abs(pn_o).max()
    np.testing.assert_allclose(pn, pn_o, rtol=1e-6, atol=1e-6 * scale)
    # a second call accumulates, as the reference's kernels do
    pn2, pc2 = es_call(which, shape, ref_dirs2, f, p, curl, thr, pn=pn, pc=pc)
    np.testing.assert_allclose(pn2, 2 * pn_o, rtol=2e-6, atol=2e-6 * scale)


def test_devide_count_inner(gpu_available):
    import torch
    shape = (5, 6, 7)
    rng = np.random.default_rng(2)
    p = rng.normal(size=shape).astype(np.float32)
    pn = rng.normal(size=shape).astype(np.float32)
    pc = rng.integers(0, 3, size=shape).astype(np.float32)
    pd, pnd, pcd = dev(p), dev(pn), dev(pc)
    omni3d.check(omni3d.lib().omni3d_devide_count_inner(7, 6, 5, pd.data_ptr(), pnd.data_ptr(),
                                                        pcd.data_ptr(), None))
    torch.cuda.synchronize()
    m = pc > 0
    want_p = np.where(m, pn / np.where(m, pc, 1), p)
    np.testing.assert_array_equal(pd.cpu().numpy(), want_p.astype(np.float32))
    np.testing.assert_array_equal(pnd.cpu().numpy(), np.where(m, 0, pn).astype(np.float32))


def test_curl(gpu_available):
    import torch
    shape = (9, 10, 11)
    f = F.taylor_green(11, noise=0.1, seed=3, shape=shape)
    acc = [f32(a) for a in f["acc"]]
    mask = (np.random.default_rng(0).uniform(size=shape) > 0.2).astype(np.float64)
    a = [dev(c) for c in acc]
    md = dev(mask)
    out = dev(np.zeros(shape))
    Z, Y, X = shape
    omni3d.check(omni3d.lib().omni3d_curl(X, Y, Z, f["delta"][0], f["delta"][1], f["delta"][2],
                                          a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(),
                                          md.data_ptr(), out.data_ptr(), None))
    torch.cuda.synchronize()
    want = O.curl(shape, tuple(float(np.float32(d)) for d in f["delta"]), acc, mask)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-6, atol=1e-7 * np.abs(want).max())
    # no mask: the reference's GPU kernel calCurlofMaterialAcc (kernel.cu:6863-6888)
    omni3d.check(omni3d.lib().omni3d_curl(X, Y, Z, f["delta"][0], f["delta"][1], f["delta"][2],
                                          a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(),
                                          None, out.data_ptr(), None))
    torch.cuda.synchronize()
    want = O.curl(shape, tuple(float(np.float32(d)) for d in f["delta"]), acc, np.ones(shape))
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-6, atol=1e-7 * np.abs(want).max())


def test_calrhs_bit_exact(gpu_available):
    shape = (7, 9, 8)
    print("case calRHS", flush=True)
    f = F.taylor_green(9, noise=0.1, seed=4, shape=shape)
    got = omni3d.calrhs(shape, (0.3, 0.4, 0.5), 1.2, f["acc"])
    want = O.calrhs(shape, (0.3, 0.4, 0.5), 1.2, f["acc"])
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("bc", [(0, 0, 0, 0, 0, 0), (1, 1, 1, 1, 1, 1), (0, 2, 1, 0, 0, 1)])
def test_sor3d_bit_exact(gpu_available, bc):
    shape = (7, 8, 9)
    print(f"case sor3D bc={bc}", flush=True)
    rng = np.random.default_rng(5)
    p0 = rng.normal(size=shape)
    pn0 = rng.normal(size=shape)
    rhs = O.calrhs(shape, (1, 1, 1), 1.0, F.taylor_green(9, noise=0.1, seed=5, shape=shape)["acc"])
    p, pn, sweeps, res = omni3d.sor3d(p0, rhs, 7, 0.0, bcscheme=bc, pn=pn0)
    po, n_o, res_o = O.sor3d(p0, rhs, 7, 0.0, bcscheme=bc, pn=pn0)
    assert sweeps == n_o == 8
    np.testing.assert_array_equal(p, po)
    assert abs(res - res_o) <= 1e-6 * res_o


def test_sor3d_eps_stop(gpu_available):
    shape = (6, 7, 8)
    print("case sor3D eps stop", flush=True)
    Z, Y, X = shape
    k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    v = np.cos(np.pi * (i + 0.5) / X) * np.cos(np.pi * 2 * (j + 0.5) / Y) * np.cos(np.pi * (k + 0.5) / Z)
    lam = -sum(2 - 2 * np.cos(np.pi * m / n) for m, n in ((1, X), (2, Y), (1, Z)))
    rhs = -lam * v
    # eps just above sweep 12's residue (sweeps shrink it by ~2 % each), far
    # from any other sweep's, so both sides stop at sweep 12 whatever the
    # residue's summation order
    _, _, r12 = O.sor3d(np.zeros(shape), rhs, 12, 0.0)
    eps = r12 * 1.005
    p, pn, sweeps, res = omni3d.sor3d(np.zeros(shape), rhs, 100, eps)
    po, n_o, res_o = O.sor3d(np.zeros(shape), rhs, 100, eps)
    assert sweeps == n_o == 12
    np.testing.assert_array_equal(p, po)


def test_sor_redblack_bit_exact_and_converges(gpu_available):
    shape = (9, 8, 7)
    print("case red-black", flush=True)
    f = F.taylor_green(9, noise=0.0, seed=0, shape=shape)
    rhs = O.calrhs(shape, f["delta"], 1.0, f["acc"])
    rhs -= rhs.mean()
    p, it, res = omni3d.sor_solve(rhs, omega=1.8, max_iter=25)
    po = O.sor_redblack(np.zeros(shape), rhs, 25, 1.8)
    assert it == 25
    np.testing.assert_array_equal(p, po)
    assert abs(res - O.residual(po, rhs)) <= 1e-12 * max(res, 1e-30)
    p2, it2, res2 = omni3d.sor_solve(rhs, omega=1.8, max_iter=5000, eps=1e-10)
    assert it2 < 5000 and res2 < 1e-10


def test_sor_redblack_rows_bit_exact_device_buffers(gpu_available):
    """Row-parallel half sweeps on an odd-sized grid, through the host and
    the device-buffer entry points: bit-identical to the serial oracle."""
    import torch
    shape = (33, 45, 67)
    print("case red-black rows 67x45x33", flush=True)
    rng = np.random.default_rng(11)
    rhs = rng.normal(size=shape)
    rhs -= rhs.mean()
    po = O.sor_redblack(np.zeros(shape), rhs, 12, 1.7)
    p, it, res = omni3d.sor_solve(rhs, omega=1.7, max_iter=12)
    assert it == 12
    np.testing.assert_array_equal(p, po)
    pd = torch.zeros(shape, dtype=torch.float64, device="cuda")
    rd = torch.from_numpy(rhs).cuda()
    it2, res2 = omni3d.sor_solve_device(shape, rd.data_ptr(), pd.data_ptr(), omega=1.7,
                                        max_iter=12)
    assert it2 == 12 and res2 == res
    np.testing.assert_array_equal(pd.cpu().numpy(), po)


@pytest.mark.parametrize("iters", [1, 2, 7, 16])
def test_sor_paired_passes_equal_single_passes(gpu_available, monkeypatch, iters):
    """With no stop test between iterations, pairs of red-black iterations
    run as one temporally blocked pass: the same voxel updates on the same
    values, so the result equals one pass per iteration (OMNI3D_SOR=single)
    and the serial oracle bit for bit, on a grid that is no multiple of the
    column (64 x 8 x 32) in any axis; an odd count ends with a single pass."""
    shape = (37, 29, 71)
    print(f"case red-black pairs {iters} iterations", flush=True)
    rng = np.random.default_rng(21)
    rhs = rng.normal(size=shape)
    rhs -= rhs.mean()
    p2, it2, res2 = omni3d.sor_solve(rhs, omega=1.8, max_iter=iters)
    monkeypatch.setenv("OMNI3D_SOR", "single")
    p1, it1, res1 = omni3d.sor_solve(rhs, omega=1.8, max_iter=iters)
    monkeypatch.delenv("OMNI3D_SOR")
    assert it1 == it2 == iters and res1 == res2
    np.testing.assert_array_equal(p2, p1)
    if iters <= 7:
        np.testing.assert_array_equal(p2, O.sor_redblack(np.zeros(shape), rhs, iters, 1.8))


@pytest.mark.parametrize("shape,kind", [((1, 1, 1), "uniform"), ((19, 23, 37), "uniform"),
                                        ((64, 64, 64), "curl"), ((128, 128, 128), "normal"),
                                        ((40, 40, 40), "constant"),
                                        ((257, 256, 256), "normal")])
def test_threshold_hist_bit_exact(gpu_available, shape, kind):
    """thredholdHistMaterialAccCPU (kernel.cu:8380-8418): min / max, the
    10000-bin histogram and the FP32 top-down walk, bit-identical to the
    oracle through both entry points (host signature of the reference and
    the device-buffer form); the 257 x 256 x 256 case (above 2^24 voxels)
    takes the serial FP32 walk, the others the integer suffix sums."""
    import torch
    print(f"case threshold {kind} {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    if kind == "uniform":
        c = rng.random(shape).astype(np.float32)
    elif kind == "normal":
        c = np.abs(rng.normal(size=shape)).astype(np.float32) * np.float32(40.0)
    elif kind == "constant":
        c = np.full(shape, 3.5, np.float32)
    else:
        f = F.fourier_sphere(shape[0], seed=2)
        a = [dev(x) for x in f["acc"]]
        cd = torch.empty(shape, dtype=torch.float32, device="cuda")
        Z, Y, X = shape
        assert omni3d.lib().omni3d_curl(X, Y, Z, f["delta"][0], f["delta"][1], f["delta"][2],
                                        a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(),
                                        None, cd.data_ptr(), None) == 0
        c = cd.cpu().numpy()
    L = omni3d.lib()
    for p in (0.0, 0.003, 0.05, 0.5, 0.97, 1.0, 2.0):
        want = O.threshold_hist(c, p)
        got = C.c_float(-1.0)
        Z, Y, X = shape
        L.thredholdHistMaterialAccCPU(X, Y, Z, np.ascontiguousarray(c).ctypes.data_as(
            C.POINTER(C.c_float)), p, C.byref(got))
        assert got.value == want, (p, got.value, want)
        assert omni3d.threshold_hist(dev(c), p) == want
        assert omni3d.threshold_hist(c, p) == want
