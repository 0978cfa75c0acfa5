"""The allocation-free threshold entry point (omni3d_threshold_hist_noalloc):
through the C-ABI, on two calls in a row (the second on a smaller volume, so
the shared scratch holds the first call's bins when it starts), its result
is bit-identical to omni3d_threshold_hist's on device buffers and to the
oracle's (thredholdHistMaterialAccCPU, kernel.cu:8380-8418)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import omni3d

pytestmark = pytest.mark.gpu


def test_threshold_noalloc_matches_allocating_entry(gpu_available):
    import torch
    L = omni3d.lib()
    L.omni3d_threshold_hist_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                C.c_float, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(17)
    for shape, pct in (((40, 48, 56), 0.2), ((17, 19, 23), 0.05)):
        print(f"case threshold no-alloc {shape} {pct}", flush=True)
        Z, Y, X = shape
        curl = rng.lognormal(0.0, 1.0, size=shape).astype(np.float32)
        curl.reshape(-1)[::97] = 0.0  # masked voxels
        cd = torch.from_numpy(curl).cuda()
        ref = omni3d.threshold_hist(cd, pct)
        out = torch.full((1,), -1.0, dtype=torch.float32, device="cuda")
        omni3d.check(L.omni3d_threshold_hist_noalloc(X, Y, Z, cd.data_ptr(), pct,
                                                     out.data_ptr(), None))
        torch.cuda.synchronize()
        got = float(out.item())
        assert np.float32(got).tobytes() == np.float32(ref).tobytes(), (got, ref)
        want = O.threshold_hist(curl, pct)
        assert np.float32(got).tobytes() == np.float32(want).tobytes(), (got, want)


def test_threshold_noalloc_rejects_too_large(gpu_available):
    L = omni3d.lib()
    L.omni3d_threshold_hist_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                C.c_float, C.c_void_p, C.c_void_p]
    assert L.omni3d_threshold_hist_noalloc(1 << 16, 1 << 16, 1, 1, 0.5, 1, None) == omni3d.ERR_ARG
