"""Python harness binding of libomni3d.so (ctypes over include/omni3d.h).

The library is C++/HIP; this module only loads it and wraps the C-ABI for
tests and the benchmark.  It never falls back to anything else: if the
shared library is missing or no GPU is visible, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libomni3d.so")

OK = 0
ERR_ARG, ERR_STATE, ERR_HIP, ERR_IO, ERR_NONFINITE, ERR_NODEVICE = 1, 2, 3, 4, 5, 6
PLAIN, PLAIN_MASK, SELECT, WEIGHTED, SELECT_FIXEDBC, MINICURL = 0, 1, 2, 3, 4, 5
# the integrators no reference launcher calls
WEIGHTED_FIXEDBC, WEIGHTED_MINICURL = 6, 7
FP64, FP32_KAHAN = 0, 1
FP32 = FP32_KAHAN
STAIRCASE, TRILINEAR = 0, 1
AFFINE, NAIVE = 0, 1
F32, F64 = 0, 1
HOST, DEVICE = 0, 1

EXPORTS = [
    "omni3d_default_params", "omni3d_create", "omni3d_destroy", "omni3d_last_error",
    "omni3d_last_global_error", "omni3d_set_stream", "omni3d_set_directions",
    "omni3d_load_directions_file", "omni3d_generate_directions", "omni3d_set_field",
    "omni3d_reconstruct", "omni3d_es_inner", "omni3d_es_inner_weighted",
    "omni3d_es_inner_select", "omni3d_devide_count_inner", "omni3d_curl", "omni3d_forward",
    "omni3d_backward", "omni3d_sor_solve", "omni3d_sor_solve_device", "sor3D", "calRHS", "omni3d_sor3d_ex",
    "omni3d_calrhs", "omni3d_run_parameter_file", "omni3d_set_tile_shape",
    "omni3d_batch_create", "omni3d_batch_destroy", "omni3d_batch_last_error",
    "omni3d_batch_set_directions", "omni3d_batch_set_field", "omni3d_batch_tune",
    "omni3d_batch_run", "omni3d_batch_result", "omni3d_field_curl",
    "omni3d_read_parameter_file", "omni3d_run_parameter_file_ex",
    "thredholdHistMaterialAccCPU", "omni3d_threshold_hist", "omni3d_boundary_pairs",
    "omni3d_es_inner_select_fixedbc", "omni3d_es_inner_weighted_minicurl",
    "omni3d_set_velocity", "omni3d_dropin_release", "initialIntegration",
    "omni3d_initial_integration", "omni3dparallellinesESInner",
    "omni3dparallellinesESInnerWeighted", "omni3dparallellinesESInnerSelect",
    "omni3dparallellinesESInnerSelectFixedBC", "omni3dparallellinesESInnerWeightedMiniCurl",
    "devidecountInner",
]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("density", C.c_double), ("line_spacing", C.c_double),
        ("method", C.c_int), ("threshold", C.c_double), ("faithful_select", C.c_int),
        ("precision", C.c_int), ("solver", C.c_int), ("max_iterations", C.c_int),
        ("tolerance", C.c_double), ("seed_pn_with_p", C.c_int), ("traversal", C.c_int),
    ]


class ParameterFile(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("density", C.c_double), ("scale", C.c_int), ("line_spacing", C.c_double),
        ("n_angles", C.c_int), ("grid_file", C.c_char * 1024), ("acc_dir", C.c_char * 1024),
        ("pressure_dir", C.c_char * 1024), ("n_iterations", C.c_int), ("threshold", C.c_double),
        ("pref", C.c_double), ("roi", C.c_int * 6),
    ]


class Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("rel_change", C.c_double), ("rays", C.c_int64),
        ("steps", C.c_int64), ("sources", C.c_int64), ("status", C.c_int),
        ("ms_plan", C.c_double), ("ms_solve", C.c_double), ("ms_pass_a", C.c_double),
        ("ms_iterate", C.c_double), ("ms_pass_b", C.c_double), ("ms_finalize", C.c_double),
        ("source_visits", C.c_int64), ("segments", C.c_int64), ("sell_cells", C.c_int64),
        ("ms_segment_sums", C.c_double), ("matrix_words", C.c_int64),
    ]


_lib = None


class Omni3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"omni3d error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    L.omni3d_default_params.argtypes = [C.POINTER(Params)]
    L.omni3d_default_params.restype = None
    L.omni3d_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.omni3d_destroy.argtypes = [vp]
    L.omni3d_destroy.restype = None
    L.omni3d_last_error.argtypes = [vp]
    L.omni3d_last_error.restype = C.c_char_p
    L.omni3d_last_global_error.argtypes = []
    L.omni3d_last_global_error.restype = C.c_char_p
    L.omni3d_set_stream.argtypes = [vp, vp]
    L.omni3d_set_directions.argtypes = [vp, fp, fp, fp, i]
    L.omni3d_load_directions_file.argtypes = [vp, C.c_char_p, i]
    L.omni3d_generate_directions.argtypes = [i, fp, fp, fp, i]
    L.omni3d_set_field.argtypes = [vp, vp, vp, vp, vp, i, i]
    L.omni3d_reconstruct.argtypes = [vp, vp, vp, i, i, C.POINTER(Report)]
    lg = C.c_long
    # the reference kernels' argument lists (kernel.cu:6151...): long dims,
    # direction arrays, long* index, spacing, density, field
    es = [lg, lg, lg, i, f, vp, vp, vp, vp, f, f, f, f, vp, vp, vp]
    L.omni3d_es_inner.argtypes = es + [vp, vp, vp, vp]
    L.omni3d_es_inner_weighted.argtypes = es + [vp, vp, vp, vp, vp]
    L.omni3d_es_inner_select.argtypes = es + [vp, vp, vp, vp, f, vp]
    L.omni3d_es_inner_select_fixedbc.argtypes = es + [vp, vp, vp, vp, f, vp]
    L.omni3d_es_inner_weighted_minicurl.argtypes = es + [vp, vp, vp, vp, vp]
    L.omni3d_devide_count_inner.argtypes = [lg, lg, lg, vp, vp, vp, vp]
    L.omni3d_dropin_release.argtypes = []
    L.omni3d_dropin_release.restype = None
    for name, extra in (("omni3dparallellinesESInner", [vp, vp, vp]),
                        ("omni3dparallellinesESInnerWeighted", [vp, vp, vp, vp]),
                        ("omni3dparallellinesESInnerSelect", [vp, vp, vp, vp, f]),
                        ("omni3dparallellinesESInnerSelectFixedBC", [vp, vp, vp, vp, f]),
                        ("omni3dparallellinesESInnerWeightedMiniCurl", [vp, vp, vp, vp])):
        getattr(L, name).argtypes = es + extra
        getattr(L, name).restype = None
    L.devidecountInner.argtypes = [lg, lg, lg, vp, vp, vp]
    L.devidecountInner.restype = None
    L.omni3d_set_velocity.argtypes = [vp, vp, vp, vp, d, i, i]
    L.initialIntegration.argtypes = [lg, lg, lg, d, d, d, d, dp, dp, dp, d, dp, dp, dp]
    L.initialIntegration.restype = None
    L.omni3d_initial_integration.argtypes = [lg, lg, lg, d, d, d, d, dp, dp, dp, d, dp, dp]
    L.omni3d_curl.argtypes = [i, i, i, f, f, f, vp, vp, vp, vp, vp, vp]
    L.thredholdHistMaterialAccCPU.argtypes = [i, i, i, C.POINTER(C.c_float), f,
                                              C.POINTER(C.c_float)]
    L.thredholdHistMaterialAccCPU.restype = None
    L.omni3d_threshold_hist.argtypes = [i, i, i, vp, f, vp, i, vp]
    L.omni3d_boundary_pairs.argtypes = [vp, i, vp, i, i, C.POINTER(C.c_longlong)]
    L.omni3d_forward.argtypes = [vp, vp, vp]
    L.omni3d_backward.argtypes = [vp, vp, vp]
    L.omni3d_sor_solve.argtypes = [i, i, i, d, i, d, dp, dp, C.POINTER(C.c_int), dp]
    L.omni3d_sor_solve_device.argtypes = [i, i, i, d, i, d, vp, vp, vp, C.POINTER(C.c_int), dp]
    L.sor3D.argtypes = [i, i, i, i, C.POINTER(C.c_int), dp, dp, dp, dp, dp, dp, d]
    L.sor3D.restype = None
    L.calRHS.argtypes = [i, i, i, d, d, d, d, dp, dp, dp, dp]
    L.calRHS.restype = None
    L.omni3d_sor3d_ex.argtypes = [i, i, i, i, C.POINTER(C.c_int), dp, dp, dp, d,
                                  C.POINTER(C.c_int), dp]
    L.omni3d_calrhs.argtypes = [i, i, i, d, d, d, d, dp, dp, dp, dp]
    L.omni3d_run_parameter_file.argtypes = [C.c_char_p]
    L.omni3d_set_tile_shape.argtypes = [vp, i, i]
    L.omni3d_field_curl.argtypes = [vp, vp, i, i]
    L.omni3d_read_parameter_file.argtypes = [C.c_char_p, C.POINTER(ParameterFile)]
    L.omni3d_run_parameter_file_ex.argtypes = [C.c_char_p, i]
    ip = C.POINTER(C.c_int)
    L.omni3d_batch_create.argtypes = [C.POINTER(Params), i, ip, ip, ip, C.POINTER(vp)]
    L.omni3d_batch_destroy.argtypes = [vp]
    L.omni3d_batch_destroy.restype = None
    L.omni3d_batch_last_error.argtypes = [vp]
    L.omni3d_batch_last_error.restype = C.c_char_p
    L.omni3d_batch_set_directions.argtypes = [vp, fp, fp, fp, i]
    L.omni3d_batch_set_field.argtypes = [vp, i, vp, vp, vp, vp, i, i]
    L.omni3d_batch_tune.argtypes = [vp, ip, ip]
    L.omni3d_batch_run.argtypes = [vp, i, dp]
    L.omni3d_batch_result.argtypes = [vp, i, vp, i, i, C.POINTER(Report)]


def check(rc, ctx=None):
    if rc != OK:
        L = lib()
        msg = (L.omni3d_last_error(ctx) if ctx else L.omni3d_last_global_error()) or b""
        raise Omni3DError(rc, msg.decode(errors="replace"))


def generate_directions(divide):
    """Icosahedral unit vectors, shape (D, 3) float32."""
    L = lib()
    n = L.omni3d_generate_directions(divide, None, None, None, 0)
    if n < 0:
        check(-n)
    k = np.zeros((3, n), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    L.omni3d_generate_directions(divide, k[0].ctypes.data_as(fp), k[1].ctypes.data_as(fp),
                                 k[2].ctypes.data_as(fp), n)
    return np.ascontiguousarray(k.T)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def calrhs(shape, delta, density, acc):
    """calRHS on the GPU (host arrays, FP64); returns the RHS volume."""
    Z, Y, X = shape
    a = [np.ascontiguousarray(c, dtype=np.float64) for c in acc]
    out = np.zeros(shape)
    check(lib().omni3d_calrhs(X, Y, Z, delta[0], delta[1], delta[2], density, _dptr(a[0]),
                              _dptr(a[1]), _dptr(a[2]), _dptr(out)))
    return out


def initial_integration(shape, delta, density, acc, pref, pn=None):
    """initialIntegration on the GPU (host arrays, FP64); returns (p, pn).
    pn: the starting pn (its centre voxel is kept), zeros if None."""
    Z, Y, X = shape
    a = [np.ascontiguousarray(c, dtype=np.float64) for c in acc]
    pn = np.zeros(shape) if pn is None else np.ascontiguousarray(pn, dtype=np.float64).copy()
    p = np.zeros(shape)
    check(lib().omni3d_initial_integration(X, Y, Z, delta[0], delta[1], delta[2], density,
                                           _dptr(a[0]), _dptr(a[1]), _dptr(a[2]), pref,
                                           _dptr(p), _dptr(pn)))
    return p, pn


def sor3d(p, rhs, noitr, eps, bcscheme=(0, 0, 0, 0, 0, 0), pn=None):
    """sor3D on the GPU; returns (p, pn, sweeps, residue)."""
    Z, Y, X = p.shape
    p = np.ascontiguousarray(p, dtype=np.float64).copy()
    pn = np.zeros_like(p) if pn is None else np.ascontiguousarray(pn, dtype=np.float64).copy()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    bc = (C.c_int * 6)(*bcscheme)
    sw = C.c_int(0)
    res = C.c_double(0)
    check(lib().omni3d_sor3d_ex(X, Y, Z, noitr, bc, _dptr(p), _dptr(pn), _dptr(rhs), eps,
                                C.byref(sw), C.byref(res)))
    return p, pn, sw.value, res.value


def sor_solve(rhs, p0=None, omega=1.8, max_iter=1000, eps=0.0):
    """Red-black SOR on the GPU; returns (p, iterations, residual)."""
    Z, Y, X = rhs.shape
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    p = np.zeros_like(rhs) if p0 is None else np.ascontiguousarray(p0, dtype=np.float64).copy()
    it = C.c_int(0)
    res = C.c_double(0)
    check(lib().omni3d_sor_solve(X, Y, Z, omega, max_iter, eps, _dptr(rhs), _dptr(p),
                                 C.byref(it), C.byref(res)))
    return p, it.value, res.value


def sor_solve_device(shape, rhs_ptr, p_ptr, omega=1.8, max_iter=1000, eps=0.0, stream=None):
    """Red-black SOR on device buffers (double, z-major shape); returns
    (iterations, residual)."""
    Z, Y, X = shape
    it = C.c_int(0)
    res = C.c_double(0)
    check(lib().omni3d_sor_solve_device(X, Y, Z, omega, max_iter, eps, rhs_ptr, p_ptr, stream,
                                        C.byref(it), C.byref(res)))
    return it.value, res.value


def threshold_hist(curl, percentage, stream=None):
    """Curl threshold above which `percentage` of the voxels lie (10000-bin
    histogram, FP32).  curl: host numpy float32 (z-major) or a device torch
    float32 tensor."""
    Z, Y, X = curl.shape
    out = C.c_float(0)
    if hasattr(curl, "is_cuda") and curl.is_cuda:
        import torch
        c = curl.contiguous()
        t = torch.empty(1, dtype=torch.float32, device=c.device)
        check(lib().omni3d_threshold_hist(X, Y, Z, C.c_void_p(c.data_ptr()), percentage,
                                          C.c_void_p(t.data_ptr()), DEVICE, stream))
        return float(t.item())
    c = np.ascontiguousarray(curl, dtype=np.float32)
    check(lib().omni3d_threshold_hist(X, Y, Z, C.c_void_p(c.ctypes.data), percentage,
                                      C.c_void_p(C.addressof(out)), HOST, stream))
    return out.value


def read_parameter_file(path):
    """Parameter_Omni3D.dat as a ParameterFile structure (host only)."""
    pf = ParameterFile()
    check(lib().omni3d_read_parameter_file(str(path).encode(), C.byref(pf)))
    return pf


def run_parameter_file(path, precision=FP64):
    """The Omni3D_8.0.exe drop-in; returns the number of files reconstructed."""
    n = lib().omni3d_run_parameter_file_ex(str(path).encode(), precision)
    if n < 0:
        check(-n)
    return n


def default_params(**kw):
    p = Params()
    lib().omni3d_default_params(C.byref(p))
    for key, val in kw.items():
        setattr(p, key, val)
    return p


class Context:
    """One omni3d_ctx: directions + field -> reconstruct."""

    def __init__(self, shape, delta=(1.0, 1.0, 1.0), **kw):
        Z, Y, X = shape
        self.shape = tuple(shape)
        self.params = default_params(nx=X, ny=Y, nz=Z, dx=delta[0], dy=delta[1], dz=delta[2],
                                     **kw)
        self._h = C.c_void_p()
        check(lib().omni3d_create(C.byref(self.params), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().omni3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        check(rc, self._h)

    def set_directions(self, k):
        k = np.ascontiguousarray(np.asarray(k, dtype=np.float32).reshape(-1, 3).T)
        fp = C.POINTER(C.c_float)
        self._chk(lib().omni3d_set_directions(self._h, k[0].ctypes.data_as(fp),
                                              k[1].ctypes.data_as(fp), k[2].ctypes.data_as(fp),
                                              k.shape[1]))

    def load_directions_file(self, path, count=0):
        self._chk(lib().omni3d_load_directions_file(self._h, path.encode(), count))

    def field_curl(self):
        """|curl a| * mask of the context's field (FP64 host array)."""
        out = np.zeros(self.shape)
        self._chk(lib().omni3d_field_curl(self._h, out.ctypes.data, F64, HOST))
        return out

    def set_tile_shape(self, tile_bits, threads):
        self._chk(lib().omni3d_set_tile_shape(self._h, tile_bits, threads))

    def set_field(self, acc, mask=None, dtype=np.float64):
        self._keep = [np.ascontiguousarray(a, dtype=dtype) for a in acc]
        m = None if mask is None else np.ascontiguousarray(mask, dtype=dtype)
        self._keep.append(m)
        dt = F64 if dtype == np.float64 else F32
        self._chk(lib().omni3d_set_field(self._h, self._keep[0].ctypes.data,
                                         self._keep[1].ctypes.data, self._keep[2].ctypes.data,
                                         None if m is None else m.ctypes.data, dt, HOST))

    def set_velocity(self, uvw, pref, dtype=np.float64):
        """Bernoulli seed of the j = 0 face from host velocity arrays (uvw =
        None removes it)."""
        if uvw is None:
            self._chk(lib().omni3d_set_velocity(self._h, None, None, None, 0.0, F64, HOST))
            return
        vel = [np.ascontiguousarray(a, dtype=dtype) for a in uvw]
        dt = F64 if dtype == np.float64 else F32
        self._chk(lib().omni3d_set_velocity(self._h, vel[0].ctypes.data, vel[1].ctypes.data,
                                            vel[2].ctypes.data, pref, dt, HOST))

    def set_field_device(self, ptrs, mask_ptr=None, dtype=F32):
        self._chk(lib().omni3d_set_field(self._h, ptrs[0], ptrs[1], ptrs[2], mask_ptr, dtype,
                                         DEVICE))

    def reconstruct(self, p0=None, dtype=np.float64, weights=True):
        """Host buffers; p0 = initial pressure (zeros if None).  Returns (p, w, report)."""
        p = (np.zeros(self.shape, dtype=dtype) if p0 is None
             else np.ascontiguousarray(p0, dtype=dtype).copy())
        w = np.zeros(self.shape, dtype=dtype) if weights else None
        rep = Report()
        dt = F64 if dtype == np.float64 else F32
        self._chk(lib().omni3d_reconstruct(self._h, p.ctypes.data,
                                           w.ctypes.data if w is not None else None, dt, HOST,
                                           C.byref(rep)))
        return p, w, rep

    def boundary_pairs(self, p0, iterations, dtype=np.float64):
        """Boundary-pair pipeline (README:18-28): host buffers; returns (p with
        the exit voxels updated, number of pairs)."""
        p = np.ascontiguousarray(p0, dtype=dtype).copy()
        n = C.c_longlong(0)
        self._chk(lib().omni3d_boundary_pairs(self._h, iterations, p.ctypes.data,
                                              F64 if dtype == np.float64 else F32, HOST,
                                              C.byref(n)))
        return p, n.value

    def forward_device(self, grad_ptr, p_ptr):
        """p = F(g) for g = grad p (3 components, c*V + v), device buffers in
        the context's precision (double for FP64, float for FP32_KAHAN)."""
        self._chk(lib().omni3d_forward(self._h, grad_ptr, p_ptr))

    def backward_device(self, dldp_ptr, dldg_ptr):
        """dL/dg = F^T dL/dp, device buffers as for forward_device."""
        self._chk(lib().omni3d_backward(self._h, dldp_ptr, dldg_ptr))

    def reconstruct_device(self, p_ptr, w_ptr=None, dtype=F64, report=True):
        rep = Report()
        self._chk(lib().omni3d_reconstruct(self._h, p_ptr, w_ptr, dtype, DEVICE,
                                           C.byref(rep) if report else None))
        return rep


class Batch:
    """omni3d_batch: snapshots of uneven grids reconstructed in one hipGraph."""

    def __init__(self, shapes, delta=(1.0, 1.0, 1.0), **kw):
        self.shapes = [tuple(s) for s in shapes]
        n = len(self.shapes)
        nx = (C.c_int * n)(*[s[2] for s in self.shapes])
        ny = (C.c_int * n)(*[s[1] for s in self.shapes])
        nz = (C.c_int * n)(*[s[0] for s in self.shapes])
        self.params = default_params(dx=delta[0], dy=delta[1], dz=delta[2], **kw)
        self._h = C.c_void_p()
        check(lib().omni3d_batch_create(C.byref(self.params), n, nx, ny, nz, C.byref(self._h)))
        self._keep = {}

    def _chk(self, rc):
        if rc != OK:
            msg = lib().omni3d_batch_last_error(self._h) or b""
            raise Omni3DError(rc, msg.decode(errors="replace"))

    def close(self):
        if self._h:
            lib().omni3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_directions(self, k):
        k = np.ascontiguousarray(np.asarray(k, dtype=np.float32).reshape(-1, 3).T)
        fp = C.POINTER(C.c_float)
        self._chk(lib().omni3d_batch_set_directions(self._h, k[0].ctypes.data_as(fp),
                                                    k[1].ctypes.data_as(fp),
                                                    k[2].ctypes.data_as(fp), k.shape[1]))

    def set_field(self, i, acc, mask=None, dtype=np.float64):
        keep = [np.ascontiguousarray(a, dtype=dtype) for a in acc]
        m = None if mask is None else np.ascontiguousarray(mask, dtype=dtype)
        self._keep[i] = keep + [m]
        dt = F64 if dtype == np.float64 else F32
        self._chk(lib().omni3d_batch_set_field(self._h, i, keep[0].ctypes.data,
                                               keep[1].ctypes.data, keep[2].ctypes.data,
                                               None if m is None else m.ctypes.data, dt, HOST))

    def tune(self):
        n = len(self.shapes)
        bits = (C.c_int * n)()
        thr = (C.c_int * n)()
        r = lib().omni3d_batch_tune(self._h, bits, thr)
        if r < 0:
            self._chk(-r)
        return [(bits[j], thr[j]) for j in range(r)]

    def run(self, use_graph=True):
        ms = C.c_double(0)
        self._chk(lib().omni3d_batch_run(self._h, 1 if use_graph else 0, C.byref(ms)))
        return ms.value

    def result(self, i):
        p = np.zeros(self.shapes[i])
        rep = Report()
        self._chk(lib().omni3d_batch_result(self._h, i, p.ctypes.data, F64, HOST, C.byref(rep)))
        return p, rep
