"""GPU checks of the drop-in boundary, called through the C-ABI:

  * initialIntegration (initialIntegration.h:3), both the void reference
    signature and the status form, bit-identical to the oracle's serial
    octant sweeps (initialIntegration.cpp:9-408) on odd, even, flat and
    single-voxel-wide grids;
  * the reference kernels' own names (kernel.cu:6414, 6746) with the
    reference's argument lists, including long dims and long* index, against
    one oracle pass;
  * the drop-ins' input cache: repeated calls with new p, a field or curl
    rewritten in place under the same pointers, a changed threshold and a
    second stream each give the oracle's pass for the inputs of that call; a
    non-finite p is reported before anything is queued;
  * the Bernoulli seed of the j = 0 face (kernel.cu:8806-8811) through
    omni3d_set_velocity, against the oracle started from that face, for the
    plain and the faithful Select methods; removing it restores the unseeded
    result;
  * a snapshot batch captured without masks and replayed after masked fields
    are set (and the reverse) equals lone contexts of the new fields.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(5, 6, 7), (8, 8, 8), (3, 9, 2), (1, 4, 5), (33, 31, 30)])
def test_initial_integration_bit_exact(gpu_available, shape):
    print(f"case initialIntegration {shape}", flush=True)
    f = F.taylor_green(max(shape), noise=0.3, seed=4, shape=shape)
    rng = np.random.default_rng(5)
    pn0 = rng.normal(size=shape)
    p, pn = omni3d.initial_integration(shape, (0.3, 0.45, 0.6), 1.7, f["acc"], 2.5, pn=pn0)
    po, pno = O.initial_integration(shape, (0.3, 0.45, 0.6), 1.7, f["acc"], 2.5, pn=pn0)
    np.testing.assert_array_equal(pn, pno)
    np.testing.assert_array_equal(p, po)


def test_initial_integration_reference_signature(gpu_available):
    shape = (6, 7, 9)
    print("case initialIntegration reference signature", flush=True)
    Z, Y, X = shape
    f = F.quadratic(shape, seed=6, delta=(0.2, 0.3, 0.4), rho=1.1)
    a = [np.ascontiguousarray(c, dtype=np.float64) for c in f["acc"]]
    p = np.zeros(shape)
    pn = np.zeros(shape)
    pn[Z // 2, Y // 2, X // 2] = f["p"][Z // 2, Y // 2, X // 2]
    pbc = np.zeros(4)
    dp = C.POINTER(C.c_double)
    omni3d.lib().initialIntegration(X, Y, Z, 0.2, 0.3, 0.4, 1.1, a[0].ctypes.data_as(dp),
                                    a[1].ctypes.data_as(dp), a[2].ctypes.data_as(dp), 0.0,
                                    p.ctypes.data_as(dp), pn.ctypes.data_as(dp),
                                    pbc.ctypes.data_as(dp))
    # a quadratic pressure is marched exactly from its centre value
    np.testing.assert_allclose(p, f["p"], rtol=0, atol=1e-12 * np.abs(f["p"]).max())
    np.testing.assert_array_equal(p, pn)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def select_inputs(shape, seed):
    f = F.taylor_green(shape[2], noise=0.05, seed=seed, shape=shape)
    f = dict(f, acc=[f32(a) for a in f["acc"]])
    curl = f32(O.curl(shape, f["delta"], f["acc"]))
    thr = float(np.float32(np.quantile(curl, 0.8)))
    return f, curl, thr


def oracle_select(shape, dirs, f, p, curl, thr):
    pn, cnt, _ = O.one_pass(shape, dirs, 1.0, f["delta"], f["rho"], f["acc"], p, curl=curl,
                            threshold=thr, method=O.SELECT, faithful=1)
    return pn, cnt


def check_pass(pn, pc, pn_o, cnt_o):
    np.testing.assert_allclose(pc, cnt_o, rtol=1e-6, atol=1e-30)
    np.testing.assert_allclose(pn, pn_o, rtol=1e-6, atol=1e-6 * np.abs(pn_o).max())


def test_reference_named_kernels(gpu_available, ref_dirs2):
    import torch
    shape = (12, 13, 14)
    print("case reference-named drop-ins", flush=True)
    Z, Y, X = shape
    f, curl, thr = select_inputs(shape, 7)
    p = f32(np.random.default_rng(1).normal(size=shape))
    k = [dev(ref_dirs2[:, i]) for i in range(3)]
    a = [dev(c) for c in f["acc"]]
    cd, pd = dev(curl), dev(p)
    pnd = torch.zeros(shape, dtype=torch.float32, device="cuda")
    pcd = torch.zeros(shape, dtype=torch.float32, device="cuda")
    index = torch.zeros(X * Y * Z, dtype=torch.int64, device="cuda")  # not read
    L = omni3d.lib()
    # the call site of kernel.cu:8849 without its <<<grid, block>>>
    L.omni3dparallellinesESInnerSelect(X, Y, Z, len(ref_dirs2), 1.0, k[0].data_ptr(),
                                       k[1].data_ptr(), k[2].data_ptr(), index.data_ptr(),
                                       f["delta"][0], f["delta"][1], f["delta"][2], f["rho"],
                                       a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(),
                                       cd.data_ptr(), pd.data_ptr(), pnd.data_ptr(),
                                       pcd.data_ptr(), thr)
    torch.cuda.synchronize()
    pn_o, cnt_o = oracle_select(shape, ref_dirs2, f, p, curl, thr)
    check_pass(pnd.cpu().numpy().astype(np.float64), pcd.cpu().numpy().astype(np.float64),
               pn_o, cnt_o)
    # kernel.cu:8854: devidecountInner with the same argument list
    pd2 = dev(p)
    L.devidecountInner(X, Y, Z, pd2.data_ptr(), pnd.data_ptr(), pcd.data_ptr())
    torch.cuda.synchronize()
    pc = pcd.cpu().numpy()
    m = pc > 0
    got = pd2.cpu().numpy()
    np.testing.assert_allclose(got[m], (pn_o[m] / cnt_o[m]), rtol=1e-5)
    np.testing.assert_array_equal(got[~m], np.float32(p)[~m])


def test_dropin_cache_follows_inputs(gpu_available, ref_dirs2):
    import torch
    shape = (12, 13, 14)
    print("case drop-in cache", flush=True)
    Z, Y, X = shape
    L = omni3d.lib()
    f, curl, thr = select_inputs(shape, 8)
    k = [dev(ref_dirs2[:, i]) for i in range(3)]
    a = [dev(c) for c in f["acc"]]
    cd = dev(curl)

    def call(p, thr, stream=None):
        pd = dev(p)
        pnd = torch.zeros(shape, dtype=torch.float32, device="cuda")
        pcd = torch.zeros(shape, dtype=torch.float32, device="cuda")
        rc = L.omni3d_es_inner_select(X, Y, Z, len(ref_dirs2), 1.0, k[0].data_ptr(),
                                      k[1].data_ptr(), k[2].data_ptr(), None, f["delta"][0],
                                      f["delta"][1], f["delta"][2], f["rho"], a[0].data_ptr(),
                                      a[1].data_ptr(), a[2].data_ptr(), cd.data_ptr(),
                                      pd.data_ptr(), pnd.data_ptr(), pcd.data_ptr(), thr,
                                      stream)
        omni3d.check(rc)
        torch.cuda.synchronize()
        return (pnd.cpu().numpy().astype(np.float64), pcd.cpu().numpy().astype(np.float64))

    rng = np.random.default_rng(2)
    # the same field, a new p each call
    for _ in range(3):
        p = f32(rng.normal(size=shape))
        check_pass(*call(p, thr), *oracle_select(shape, ref_dirs2, f, p, curl, thr))
    # the field rewritten in place (same device pointers)
    f2, curl2, _ = select_inputs(shape, 9)
    for c in range(3):
        a[c].copy_(dev(f2["acc"][c]))
    check_pass(*call(p, thr), *oracle_select(shape, ref_dirs2, dict(f, acc=f2["acc"]), p, curl,
                                             thr))
    # the curl rewritten in place, then a new threshold
    cd.copy_(dev(curl2))
    check_pass(*call(p, thr), *oracle_select(shape, ref_dirs2, dict(f, acc=f2["acc"]), p, curl2,
                                             thr))
    thr2 = float(np.float32(np.quantile(curl2, 0.6)))
    check_pass(*call(p, thr2), *oracle_select(shape, ref_dirs2, dict(f, acc=f2["acc"]), p,
                                              curl2, thr2))
    # on a second stream (default priority)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        got = call(p, thr2, stream=s.cuda_stream)
    check_pass(*got, *oracle_select(shape, ref_dirs2, dict(f, acc=f2["acc"]), p, curl2, thr2))
    # a non-finite p is refused before anything is queued
    pbad = p.copy()
    pbad[3, 4, 5] = np.nan
    with pytest.raises(omni3d.Omni3DError) as e:
        call(pbad, thr2)
    assert e.value.code == omni3d.ERR_NONFINITE
    L.omni3d_dropin_release()


@pytest.mark.parametrize("method", [omni3d.PLAIN, omni3d.SELECT])
def test_bernoulli_seed(gpu_available, ref_dirs2, method):
    shape = (14, 15, 16)
    print(f"case bernoulli seed method={method}", flush=True)
    f = F.taylor_green(16, noise=0.05, seed=12, shape=shape)
    rng = np.random.default_rng(13)
    uvw = [rng.normal(size=shape) for _ in range(3)]
    pref, rho = 0.7, 1.3
    face = pref - 0.5 * rho * (uvw[0] ** 2 + uvw[1] ** 2 + uvw[2] ** 2)
    p0 = np.zeros(shape)
    p0[:, 0, :] = face[:, 0, :]
    kw = dict(max_iterations=6, method=method, density=rho)
    curl = None
    thr = 0.0
    if method == omni3d.SELECT:
        with omni3d.Context(shape, f["delta"], **kw) as ctx:
            ctx.set_directions(ref_dirs2)
            ctx.set_field(f["acc"])
            curl = ctx.field_curl()
        thr = float(np.quantile(curl, 0.85))
        kw["threshold"] = thr
    with omni3d.Context(shape, f["delta"], **kw) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"])
        ctx.set_velocity(uvw, pref)
        p, w, rep = ctx.reconstruct()
        ctx.set_velocity(None, 0.0)
        pu, _, _ = ctx.reconstruct()
    po, co, it, _ = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], rho, f["acc"], p0=p0,
                                  curl=curl, threshold=thr,
                                  method=O.PLAIN if method == omni3d.PLAIN else O.SELECT,
                                  faithful=1, max_iter=6, seed_pn=1)
    assert rep.iterations == it
    assert F.rel_l2(p, po) < 1e-12
    pz, _, _, _ = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], rho, f["acc"], curl=curl,
                                threshold=thr,
                                method=O.PLAIN if method == omni3d.PLAIN else O.SELECT,
                                faithful=1, max_iter=6, seed_pn=1)
    assert F.rel_l2(pu, pz) < 1e-12
    assert F.rel_l2(p, pu) > 1e-3


@pytest.mark.parametrize("method", [omni3d.PLAIN_MASK, omni3d.SELECT])
def test_batch_graph_mask_added_then_removed(gpu_available, ref_dirs2, method):
    print(f"case batch mask change method={method}", flush=True)
    shapes = [(16, 18, 20), (20, 16, 18), (16, 18, 20)]
    h = 2 * np.pi / 20
    # Select: curl * mask, with no threshold break (only the mask's zeros)
    kw = dict(max_iterations=6, method=method, threshold=1e9)
    fs = [F.taylor_green(s[2], noise=0.04, seed=30 + i, shape=s) for i, s in enumerate(shapes)]
    masks = []
    for s in shapes:
        Z, Y, X = s
        k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        masks.append(((i - X / 2) ** 2 + (j - Y / 2) ** 2 + (k - Z / 2) ** 2 > 12).astype(float))

    def lone(i, mask):
        with omni3d.Context(shapes[i], (h, h, h), **kw) as ctx:
            ctx.set_directions(ref_dirs2)
            ctx.set_field(fs[i]["acc"], mask=mask)
            return ctx.reconstruct()[0]

    with omni3d.Batch(shapes, (h, h, h), **kw) as b:
        b.set_directions(ref_dirs2)
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"])
        b.run(use_graph=True)  # captured without masks
        for i in range(len(shapes)):
            assert F.rel_l2(b.result(i)[0], lone(i, None)) < 1e-12
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"], mask=masks[i])
        b.run(use_graph=True)
        for i in range(len(shapes)):
            assert F.rel_l2(b.result(i)[0], lone(i, masks[i])) < 1e-12
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"])
        b.run(use_graph=True)
        for i in range(len(shapes)):
            assert F.rel_l2(b.result(i)[0], lone(i, None)) < 1e-12
