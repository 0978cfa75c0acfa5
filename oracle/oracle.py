"""ctypes binding of the FP64 oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  Every function mirrors one function of
omni3d_oracle.c, which cites the reference line it restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PLAIN, PLAIN_MASK, SELECT, WEIGHTED, SELECT_FIXEDBC, MINICURL = 0, 1, 2, 3, 4, 5
# the plain method integrated along the straight ray (trilinear samples)
PLAIN_TRILINEAR = 6
# the two integrators no reference launcher calls
WEIGHTED_FIXEDBC, WEIGHTED_MINICURL = 7, 8

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def load(path):
    """A separately built copy of the oracle (e.g. a mutant the tests compile
    to show that a check fails on a one-line change), bound like lib()."""
    L = C.CDLL(path)
    _declare(L)
    return L


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_change", C.c_double)]


class Dir(C.Structure):
    _fields_ = [("k", C.c_double * 3), ("e1", C.c_double * 3), ("e2", C.c_double * 3),
                ("M", (C.c_double * 3) * 3)]


def _declare(L):
    i, d, i64 = C.c_int, C.c_double, C.c_int64
    L.orc_make_dir.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(Dir)]
    L.orc_make_dir.restype = None
    L.orc_origin.argtypes = [C.POINTER(Dir), i, d, i64, _dp]
    L.orc_origin.restype = None
    L.orc_clip.argtypes = [i, i, i, _dp, _dp, _ip, _ip]
    L.orc_clip.restype = i
    L.orc_nogrid.argtypes = [i, i, i, d]
    L.orc_nogrid.restype = i
    L.orc_enumerate_rays.argtypes = [i, i, i, i, _fp, _fp, _fp, d, _i32p, i64]
    L.orc_enumerate_rays.restype = i64
    L.orc_ray_path.argtypes = [i, i, i, C.c_float, C.c_float, C.c_float, d, i64, _i64p, i, i]
    L.orc_ray_path.restype = i
    L.orc_pass.argtypes = [i, i, i, i, _fp, _fp, _fp, d, d, d, d, d, _dp, _dp, _dp, _dp, d, i,
                           i, _dp, _dp, _dp, _u8p]
    L.orc_pass.restype = None
    L.orc_divide.argtypes = [i64, _dp, _dp, _dp]
    L.orc_divide.restype = None
    L.orc_reconstruct.argtypes = [i, i, i, i, _fp, _fp, _fp, d, d, d, d, d, _dp, _dp, _dp, _dp,
                                  d, i, i, i, d, i, _dp, _dp, C.POINTER(Report)]
    L.orc_reconstruct.restype = None
    L.orc_curl.argtypes = [i, i, i, d, d, d, _dp, _dp, _dp, _dp, _dp]
    L.orc_curl.restype = None
    L.orc_calrhs.argtypes = [i, i, i, d, d, d, d, _dp, _dp, _dp, _dp]
    L.orc_calrhs.restype = None
    L.orc_sor3d.argtypes = [i, i, i, i, _ip, _dp, _dp, _dp, d, _dp]
    L.orc_sor3d.restype = i
    L.orc_sor_redblack.argtypes = [i, i, i, i, d, _dp, _dp]
    L.orc_threshold_hist.argtypes = [i, i, i, _fp, C.c_float]
    L.orc_threshold_hist.restype = C.c_float
    L.orc_sor_redblack.restype = None
    L.orc_boundary_pairs.argtypes = [i, i, i, i, _fp, _fp, _fp, d, d, d, d, d, _dp, _dp, _dp,
                                     i, _dp]
    L.orc_boundary_pairs.restype = i64
    L.orc_residual.argtypes = [i, i, i, _dp, _dp]
    L.orc_residual.restype = d
    L.orc_initial_integration.argtypes = [i, i, i, d, d, d, d, _dp, _dp, _dp, d, _dp, _dp]
    L.orc_initial_integration.restype = None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _dirs(k):
    k = np.ascontiguousarray(k, dtype=np.float32).reshape(-1, 3)
    return (np.ascontiguousarray(k[:, 0]), np.ascontiguousarray(k[:, 1]),
            np.ascontiguousarray(k[:, 2]), k.shape[0])


def make_dir(kvec):
    """(k, e1, e2, M) of one direction as float64 arrays (kernel.cu:6441-6457)."""
    d = Dir()
    lib().orc_make_dir(float(kvec[0]), float(kvec[1]), float(kvec[2]), C.byref(d))
    return (np.array(d.k[:]), np.array(d.e1[:]), np.array(d.e2[:]),
            np.array([d.M[r][:] for r in range(3)]))


def origin(kvec, ng, ls, point):
    """Centred ray origin of lattice point `point` (kernel.cu:6172-6186)."""
    d = Dir()
    lib().orc_make_dir(float(kvec[0]), float(kvec[1]), float(kvec[2]), C.byref(d))
    o = np.zeros(3)
    lib().orc_origin(C.byref(d), ng, ls, int(point), _ptr(o, _dp))
    return o


def clip(shape, o, kvec):
    """cross2point + launcher acceptance: (accepted, in_ijk, out_ijk)."""
    Z, Y, X = shape
    o = _f64(o)
    k = _f64(kvec)
    a = np.zeros(3, dtype=np.int32)
    b = np.zeros(3, dtype=np.int32)
    ok = lib().orc_clip(X, Y, Z, _ptr(o, _dp), _ptr(k, _dp), _ptr(a, _ip), _ptr(b, _ip))
    return bool(ok), a.copy(), b.copy()


def nogrid(shape, ls):
    Z, Y, X = shape
    return lib().orc_nogrid(X, Y, Z, ls)


def enumerate_rays(shape, dirs, ls):
    """All accepted rays as an int64 array of rows (d, point, iin, jin, kin, iout, jout, kout)."""
    Z, Y, X = shape
    k1, k2, k3, D = _dirs(dirs)
    n = lib().orc_enumerate_rays(X, Y, Z, D, _ptr(k1, _fp), _ptr(k2, _fp), _ptr(k3, _fp), ls,
                                 None, 0)
    buf = np.zeros((max(n, 1), 9), dtype=np.int32)
    lib().orc_enumerate_rays(X, Y, Z, D, _ptr(k1, _fp), _ptr(k2, _fp), _ptr(k3, _fp), ls,
                             _ptr(buf, _i32p), n)
    buf = buf[:n].astype(np.int64)
    pt = buf[:, 1] | (buf[:, 2] << 31)
    return np.concatenate([buf[:, :1], pt[:, None], buf[:, 3:]], axis=1)


def ray_path(shape, kvec, ls, point, capacity=1 << 16, fixedbc=0):
    Z, Y, X = shape
    out = np.zeros(capacity, dtype=np.int64)
    n = lib().orc_ray_path(X, Y, Z, float(kvec[0]), float(kvec[1]), float(kvec[2]), ls, point,
                           _ptr(out, _i64p), capacity, fixedbc)
    return out[:n]


def one_pass(shape, dirs, ls, delta, rho, acc, p, curl=None, threshold=0.0, method=PLAIN,
             faithful=1, pn=None, cnt=None, L=None):
    """One full pass; returns (pn, cnt, src_mark) with pn/cnt accumulated in place.
    L: a library from load() instead of the default build."""
    Z, Y, X = shape
    k1, k2, k3, D = _dirs(dirs)
    ax, ay, az = (_f64(a) for a in acc)
    p = _f64(p)
    pn = np.zeros(shape) if pn is None else pn
    cnt = np.zeros(shape) if cnt is None else cnt
    src = np.zeros(shape, dtype=np.uint8)
    cu = _f64(curl) if curl is not None else None
    (L or lib()).orc_pass(X, Y, Z, D, _ptr(k1, _fp), _ptr(k2, _fp), _ptr(k3, _fp), ls,
                          delta[0], delta[1], delta[2], rho, _ptr(ax, _dp), _ptr(ay, _dp),
                          _ptr(az, _dp), _ptr(cu, _dp), threshold, method, faithful,
                          _ptr(p, _dp), _ptr(pn, _dp), _ptr(cnt, _dp), _ptr(src, _u8p))
    return pn, cnt, src


def reconstruct(shape, dirs, ls, delta, rho, acc, p0=None, curl=None, threshold=0.0,
                method=PLAIN, faithful=1, max_iter=10, tol=0.0, seed_pn=1):
    """kernel.cu:8805-8861 with full coverage; returns (p, cnt, iterations, rel_change)."""
    Z, Y, X = shape
    k1, k2, k3, D = _dirs(dirs)
    ax, ay, az = (_f64(a) for a in acc)
    p = np.zeros(shape) if p0 is None else _f64(p0).copy()
    cnt = np.zeros(shape)
    cu = _f64(curl) if curl is not None else None
    rep = Report()
    lib().orc_reconstruct(X, Y, Z, D, _ptr(k1, _fp), _ptr(k2, _fp), _ptr(k3, _fp), ls, delta[0],
                          delta[1], delta[2], rho, _ptr(ax, _dp), _ptr(ay, _dp), _ptr(az, _dp),
                          _ptr(cu, _dp), threshold, method, faithful, max_iter, tol, seed_pn,
                          _ptr(p, _dp), _ptr(cnt, _dp), C.byref(rep))
    return p, cnt, rep.iterations, rep.rel_change


def boundary_pairs(shape, dirs, ls, delta, rho, acc, p0, iters):
    """Boundary-pair pipeline (kernel.cu:3414-3489, 3896-3909, 5572-5712),
    sparse pairs, Jacobi boundary iterations; returns (p, number of pairs)."""
    Z, Y, X = shape
    k1, k2, k3, D = _dirs(dirs)
    ax, ay, az = (_f64(a) for a in acc)
    p = _f64(p0).copy()
    n = lib().orc_boundary_pairs(X, Y, Z, D, _ptr(k1, _fp), _ptr(k2, _fp), _ptr(k3, _fp), ls,
                                 delta[0], delta[1], delta[2], rho, _ptr(ax, _dp),
                                 _ptr(ay, _dp), _ptr(az, _dp), iters, _ptr(p, _dp))
    return p, int(n)


def curl(shape, delta, acc, mask=None):
    Z, Y, X = shape
    ax, ay, az = (_f64(a) for a in acc)
    m = np.ones(shape) if mask is None else _f64(mask)
    out = np.zeros(shape)
    lib().orc_curl(X, Y, Z, delta[0], delta[1], delta[2], _ptr(ax, _dp), _ptr(ay, _dp),
                   _ptr(az, _dp), _ptr(m, _dp), _ptr(out, _dp))
    return out


def calrhs(shape, delta, rho, acc):
    Z, Y, X = shape
    ax, ay, az = (_f64(a) for a in acc)
    out = np.zeros(shape)
    lib().orc_calrhs(X, Y, Z, delta[0], delta[1], delta[2], rho, _ptr(ax, _dp), _ptr(ay, _dp),
                     _ptr(az, _dp), _ptr(out, _dp))
    return out


def sor3d(p, rhs, noitr, eps, bcscheme=(0, 0, 0, 0, 0, 0), pn=None):
    """sor3D.cpp:9-1589; returns (p, sweeps, residue)."""
    Z, Y, X = p.shape
    p = _f64(p).copy()
    pn = np.zeros_like(p) if pn is None else _f64(pn).copy()
    rhs = _f64(rhs)
    bc = np.ascontiguousarray(bcscheme, dtype=np.int32)
    res = C.c_double(0)
    n = lib().orc_sor3d(X, Y, Z, noitr, _ptr(bc, _ip), _ptr(p, _dp), _ptr(pn, _dp),
                        _ptr(rhs, _dp), eps, C.byref(res))
    return p, n, res.value


def sor_redblack(p, rhs, iters, omega):
    Z, Y, X = p.shape
    p = _f64(p).copy()
    rhs = _f64(rhs)
    lib().orc_sor_redblack(X, Y, Z, iters, omega, _ptr(p, _dp), _ptr(rhs, _dp))
    return p


def threshold_hist(curl, percentage):
    """kernel.cu:8380-8418 with the intended bin walk; FP32 in and out."""
    Z, Y, X = curl.shape
    c = np.ascontiguousarray(curl, dtype=np.float32)
    return float(lib().orc_threshold_hist(X, Y, Z, _ptr(c, _fp), percentage))


def residual(p, rhs):
    Z, Y, X = p.shape
    return lib().orc_residual(X, Y, Z, _ptr(_f64(p), _dp), _ptr(_f64(rhs), _dp))


def initial_integration(shape, delta, rho, acc, pref, pn=None, L=None):
    """initialIntegration.cpp:9-408; returns (p, pn).  pn: the caller's
    starting pn (its centre voxel is kept), zeros if None."""
    Z, Y, X = shape
    ax, ay, az = (_f64(a) for a in acc)
    pn = np.zeros(shape) if pn is None else _f64(pn).copy()
    p = np.zeros(shape)
    (L or lib()).orc_initial_integration(X, Y, Z, delta[0], delta[1], delta[2], rho,
                                         _ptr(ax, _dp), _ptr(ay, _dp), _ptr(az, _dp), pref,
                                         _ptr(p, _dp), _ptr(pn, _dp))
    return p, pn
