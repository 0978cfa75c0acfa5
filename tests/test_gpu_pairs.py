"""GPU parity of the boundary-pair (PINT) pipeline, omni3d_boundary_pairs
(kernel.cu:3414-3489, 1353-1467, 3896-3909, 5572-5712; README:18-28),
through the C-ABI, against the oracle's sorted-pair restatement:

  * FP64 ray sums: rel <= 1e-12 against the oracle after 1, 3 and 8 Jacobi
    iterations, the same number of pairs, interior voxels untouched;
    FP32-Kahan ray sums within 1e-6;
  * a quadratic pressure: the exact boundary is a fixed point; from zero the
    boundary converges to p + constant (README:25);
  * the README's inner pass: omni3d_reconstruct with the plain method and one
    iteration from the boundary-pair result, against the oracle's same chain.
"""
import numpy as np
import pytest

import oracle as O
import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu


def boundary(shape):
    b = np.zeros(shape, bool)
    b[0] = b[-1] = True
    b[:, 0] = b[:, -1] = True
    b[:, :, 0] = b[:, :, -1] = True
    return b


def gpu_pairs(f, dirs, shape, p0, iters, precision=omni3d.FP64, ls=1.0, engine=None):
    import os
    old = os.environ.pop("OMNI3D_ENGINE", None)
    if engine:
        os.environ["OMNI3D_ENGINE"] = engine  # read when the context is created
    try:
        with omni3d.Context(shape, f["delta"], density=f["rho"], line_spacing=ls,
                            precision=precision) as ctx:
            ctx.set_directions(dirs)
            ctx.set_field(f["acc"])
            first = ctx.boundary_pairs(p0, iters)
            # a second call on the same context reuses its work buffers
            second = ctx.boundary_pairs(p0, iters)
            np.testing.assert_array_equal(first[0], second[0])
            return first
    finally:
        os.environ.pop("OMNI3D_ENGINE", None)
        if old is not None:
            os.environ["OMNI3D_ENGINE"] = old


@pytest.mark.parametrize("shape,iters,engine", [((9, 10, 11), 1, None), ((16, 12, 7), 3, None),
                                                ((20, 20, 20), 8, None),
                                                ((16, 12, 7), 3, "rays")])
def test_fp64_matches_oracle(gpu_available, ref_dirs2, shape, iters, engine):
    """Ray sums from the tile engine's segment sums (default) or one lane per
    ray (engine "rays")."""
    print(f"case boundary pairs {shape} x{iters} {engine}", flush=True)
    f = F.taylor_green(max(shape), noise=0.05, seed=5, shape=shape)
    p0 = np.random.default_rng(7).normal(size=shape)
    want, nw = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0, iters)
    got, ng = gpu_pairs(f, ref_dirs2, shape, p0, iters, engine=engine)
    assert ng == nw
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= 1e-12, err
    inner = ~boundary(shape)
    np.testing.assert_array_equal(got[inner], p0[inner])


def test_fp32_kahan_matches_oracle(gpu_available, ref_dirs2):
    shape = (14, 15, 16)
    print("case boundary pairs fp32", flush=True)
    f = F.taylor_green(16, noise=0.05, seed=6, shape=shape)
    p0 = np.zeros(shape)
    want, nw = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], p0, 4)
    got, ng = gpu_pairs(f, ref_dirs2, shape, p0, 4, precision=omni3d.FP32_KAHAN)
    assert ng == nw
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-6


def test_quadratic_fixed_point_and_convergence(gpu_available, ref_dirs2):
    shape = (12, 13, 14)
    print("case boundary pairs quadratic", flush=True)
    f = F.quadratic(shape, seed=1)
    p, _ = gpu_pairs(f, ref_dirs2, shape, f["p"], 3)
    assert np.abs(p - f["p"]).max() <= 1e-12 * np.abs(f["p"]).max()
    b = boundary(shape)
    p, _ = gpu_pairs(f, ref_dirs2, shape, np.zeros(shape), 30)
    e = (p - f["p"])[b]
    assert np.std(e) <= 1e-6 * np.std(f["p"][b])


def test_readme_pipeline_inner_pass(gpu_available, ref_dirs2):
    shape = (16, 16, 16)
    print("case boundary pairs + inner pass", flush=True)
    f = F.taylor_green(16, noise=0.03, seed=8, shape=shape)
    z = np.zeros(shape)
    pb, _ = gpu_pairs(f, ref_dirs2, shape, z, 4)
    with omni3d.Context(shape, f["delta"], density=f["rho"], max_iterations=1,
                        seed_pn_with_p=0) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"])
        p, w, rep = ctx.reconstruct(pb)
    ob, _ = O.boundary_pairs(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"], z, 4)
    po, co, it, rel = O.reconstruct(shape, ref_dirs2, 1.0, f["delta"], f["rho"], f["acc"],
                                    p0=ob, max_iter=1, seed_pn=0)
    assert np.linalg.norm(p - po) / np.linalg.norm(po) <= 1e-12
    np.testing.assert_array_equal(w, co)
    # against the analytic pressure (mean removed): the README pipeline
    # reconstructs the field
    d = (p - p.mean()) - (f["p"] - f["p"].mean())
    assert np.linalg.norm(d) / np.linalg.norm(f["p"] - f["p"].mean()) < 0.2


def test_device_f32_buffers(gpu_available, ref_dirs2):
    """The device-buffer, FP32 form of the call gives the host FP64 result
    rounded to FP32."""
    import ctypes as C
    import torch
    shape = (12, 11, 10)
    print("case boundary pairs device f32", flush=True)
    f = F.taylor_green(12, noise=0.05, seed=13, shape=shape)
    p0 = np.random.default_rng(14).normal(size=shape).astype(np.float32)
    with omni3d.Context(shape, f["delta"], density=f["rho"]) as ctx:
        ctx.set_directions(ref_dirs2)
        ctx.set_field(f["acc"])
        want, nw = ctx.boundary_pairs(p0.astype(np.float64), 3)
        pd = torch.from_numpy(p0).cuda()
        n = C.c_longlong(0)
        omni3d.check(omni3d.lib().omni3d_boundary_pairs(ctx._h, 3, C.c_void_p(pd.data_ptr()),
                                                        omni3d.F32, omni3d.DEVICE, C.byref(n)))
        got = pd.cpu().numpy()
    assert n.value == nw
    np.testing.assert_array_equal(got, want.astype(np.float32))
