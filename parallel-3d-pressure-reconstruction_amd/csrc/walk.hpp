// Device-side staircase walker (kernel.cu:1743-1863 restated for one lane).
#pragma once

#include <hip/hip_runtime.h>

#include "geometry.hpp"
#include "kernels.hpp"

namespace omni3d {

// dist^2 change of a step s along an axis with perpendicular offset u and
// q = 1 - k_a^2: (2s) u + s^2 q, evaluated exactly as the host/oracle do.
__device__ __forceinline__ double step_metric(int s, double u, double q) {
  const double t = u + u;
  return s == 1 ? q + t : (s == -1 ? q - t : (-120000.0 * u + 3.6e9 * q));
}

__device__ __forceinline__ int sgn_step(int c, int e) {
  return c < e ? 1 : (c > e ? -1 : kSentinelStep);
}

// Walker registers for one ray.
struct Walker {
  int c0, c1, c2;
  int e0, e1, e2;
  double u0, u1, u2;
  double m00, m01, m02, m11, m12, m22;
  int64_t v;
  int64_t sy, sz;
  int X, Y, Z;

  __device__ __forceinline__ void init(const RayRec& rr, const DirConst* __restrict__ dirs,
                                       const GridDesc& g) {
    const DirConst& d = dirs[rr.dir_len & 0xffffu];
    double o[3];
    lattice_origin(d, g.ng, g.ls, (int64_t)rr.point, o);
    int c[3], e[3];
    unpack_voxel(rr.in, c);
    unpack_voxel(rr.out, e);
    double u[3];
    perp_offset(d, c, o, g.cc, u);
    c0 = c[0]; c1 = c[1]; c2 = c[2];
    e0 = e[0]; e1 = e[1]; e2 = e[2];
    u0 = u[0]; u1 = u[1]; u2 = u[2];
    m00 = d.m[0][0]; m01 = d.m[0][1]; m02 = d.m[0][2];
    m11 = d.m[1][1]; m12 = d.m[1][2]; m22 = d.m[2][2];
    X = g.X; Y = g.Y; Z = g.Z;
    sy = g.X;
    sz = (int64_t)g.X * g.Y;
    v = (int64_t)c0 + (int64_t)c1 * sy + (int64_t)c2 * sz;
  }

  __device__ __forceinline__ bool done() const { return c0 == e0 && c1 == e1 && c2 == e2; }

  // Chooses and takes one step.  Returns the axis (0..2) and the sign in *s,
  // or -1 if the walk ends here (kernel.cu:1841-1861).
  __device__ __forceinline__ int advance(int* s_out) {
    const int s0 = sgn_step(c0, e0), s1 = sgn_step(c1, e1), s2 = sgn_step(c2, e2);
    const double f0 = step_metric(s0, u0, m00);
    const double f1 = step_metric(s1, u1, m11);
    const double f2 = step_metric(s2, u2, m22);
    int a, s, nc, lim;
    if (f0 <= f1 && f0 <= f2) { a = 0; s = s0; nc = c0 + s0; lim = X; }
    else if (f1 < f0 && f1 <= f2) { a = 1; s = s1; nc = c1 + s1; lim = Y; }
    else if (f2 < f0 && f2 < f1) { a = 2; s = s2; nc = c2 + s2; lim = Z; }
    else return -1;
    if (nc < 0 || nc >= lim) return -1;
    const double r0 = a == 0 ? m00 : (a == 1 ? m01 : m02);
    const double r1 = a == 0 ? m01 : (a == 1 ? m11 : m12);
    const double r2 = a == 0 ? m02 : (a == 1 ? m12 : m22);
    const double sd = (double)s;
    u0 = u0 + sd * r0;
    u1 = u1 + sd * r1;
    u2 = u2 + sd * r2;
    if (a == 0) { c0 = nc; v += s; }
    else if (a == 1) { c1 = nc; v += s * sy; }
    else { c2 = nc; v += s * sz; }
    *s_out = s;
    return a;
  }

  // The step rule of bodyIntegralInnerSelectFixedBC (kernel.cu:2628, 2676,
  // 2722): an x / y tie goes to y, no x- or z-step while on the j = 0 plane
  // and no y-step onto it; a refused step ends the walk.
  __device__ __forceinline__ int advance_fixedbc(int* s_out) {
    const int s0 = sgn_step(c0, e0), s1 = sgn_step(c1, e1), s2 = sgn_step(c2, e2);
    const double f0 = step_metric(s0, u0, m00);
    const double f1 = step_metric(s1, u1, m11);
    const double f2 = step_metric(s2, u2, m22);
    int a, s, nc, lim;
    if (f0 < f1 && f0 <= f2) { a = 0; s = s0; nc = c0 + s0; lim = X; }
    else if (f1 <= f0 && f1 <= f2) { a = 1; s = s1; nc = c1 + s1; lim = Y; }
    else if (f2 < f0 && f2 < f1) { a = 2; s = s2; nc = c2 + s2; lim = Z; }
    else return -1;
    if (a != 1 && c1 == 0) return -1;
    if (a == 1 && nc == 0) return -1;
    if (nc < 0 || nc >= lim) return -1;
    const double r0 = a == 0 ? m00 : (a == 1 ? m01 : m02);
    const double r1 = a == 0 ? m01 : (a == 1 ? m11 : m12);
    const double r2 = a == 0 ? m02 : (a == 1 ? m12 : m22);
    const double sd = (double)s;
    u0 = u0 + sd * r0;
    u1 = u1 + sd * r1;
    u2 = u2 + sd * r2;
    if (a == 0) { c0 = nc; v += s; }
    else if (a == 1) { c1 = nc; v += s * sy; }
    else { c2 = nc; v += s * sz; }
    *s_out = s;
    return a;
  }

  // The plain staircase step (ties x, y, z) refusing a y-step onto the
  // j = 0 plane (bodyIntegralInnerWeightedFixedBC, kernel.cu:2226: jnext2 > 0):
  // such a step ends the walk, as in the reference.
  __device__ __forceinline__ int advance_noj0(int* s_out) {
    const int s1 = sgn_step(c1, e1);
    if (c1 + s1 == 0) {
      const int s0 = sgn_step(c0, e0), s2 = sgn_step(c2, e2);
      const double f0 = step_metric(s0, u0, m00);
      const double f1 = step_metric(s1, u1, m11);
      const double f2 = step_metric(s2, u2, m22);
      if (f1 < f0 && f1 <= f2) return -1;
    }
    return advance(s_out);
  }

  __device__ __forceinline__ bool on_surface() const {
    return c0 == 0 || c0 == X - 1 || c1 == 0 || c1 == Y - 1 || c2 == 0 || c2 == Z - 1;
  }
};

// Geometry-only walk.  vis.entry(v) at the entry voxel, vis.visit(v, surface)
// after every step.  Returns the number of steps.
template <typename Visitor>
__device__ __forceinline__ int walk_geometry(const RayRec& rr, const DirConst* __restrict__ dirs,
                                             const GridDesc& g, Visitor& vis) {
  Walker w;
  w.init(rr, dirs, g);
  vis.entry(w.v);
  int steps = 0;
  while (!w.done()) {
    int s;
    if (w.advance(&s) < 0) break;
    vis.visit(w.v, w.on_surface());
    steps++;
  }
  return steps;
}

// round(x * scale) to an integer for |x * scale| < 2^51 (the scale is chosen
// so): adding 1.5 * 2^52 leaves the rounded integer in the low mantissa bits.
__device__ __forceinline__ long long to_fixed(double x, double scale) {
  const double y = x * scale + 6755399441055744.0;
  return __double_as_longlong(y) - 0x4338000000000000LL;
}

__device__ __forceinline__ void fx_add(long long* p, long long v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// Ray sum in the arithmetic of T: FP64 plain accumulation, or FP32 with
// Kahan compensation.  value() is the FP64 view of the running sum.
template <typename T>
struct RaySum;

template <>
struct RaySum<double> {
  double s = 0.0;
  __device__ __forceinline__ void add(double t) { s = s + t; }
  __device__ __forceinline__ double value() const { return s; }
  __device__ __forceinline__ void reset() { s = 0.0; }
};

template <>
struct RaySum<float> {
  float s = 0.0f, comp = 0.0f;
  __device__ __forceinline__ void add(float t) {
    const float y = t - comp;
    const float n = s + y;
    comp = (n - s) - y;
    s = n;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)comp; }
  __device__ __forceinline__ void reset() { s = 0.0f; comp = 0.0f; }
};

template <typename T>
__device__ __forceinline__ T trapezoid(const WalkParams& wp, int a, int s, T an, T al);

// The step factor is picked with constant indices (selects, no indexed
// load), so the six factors stay in scalar registers across the walk.
template <>
__device__ __forceinline__ double trapezoid<double>(const WalkParams& wp, int a, int s, double an,
                                                    double al) {
  const bool up = s > 0;
  const double h0 = up ? wp.hs[0][1] : wp.hs[0][0];
  const double h1 = up ? wp.hs[1][1] : wp.hs[1][0];
  const double h2 = up ? wp.hs[2][1] : wp.hs[2][0];
  return (a == 0 ? h0 : (a == 1 ? h1 : h2)) * (an + al);
}

template <>
__device__ __forceinline__ float trapezoid<float>(const WalkParams& wp, int a, int s, float an,
                                                  float al) {
  const bool up = s > 0;
  const float h0 = up ? wp.hsf[0][1] : wp.hsf[0][0];
  const float h1 = up ? wp.hsf[1][1] : wp.hsf[1][0];
  const float h2 = up ? wp.hsf[2][1] : wp.hsf[2][0];
  return (a == 0 ? h0 : (a == 1 ? h1 : h2)) * (an + al);
}

// Field staged as one 4-vector per voxel (a_x, a_y, a_z, 0): a step loads the
// next voxel once and the trapezoid's "last" value comes from registers.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

template <typename V4>
__device__ __forceinline__ auto component(const V4& v, int a) -> decltype(v.x) {
  return a == 0 ? v.x : (a == 1 ? v.y : v.z);
}

// Gradient walk of the plain method (kernel.cu:1743-1863).  vis.entry(v) for
// the entry self-add, vis.visit(v, surface, pint) after every step with the
// running trapezoid sum (kernel.cu:1816-1838).
template <typename T, typename Visitor>
__device__ __forceinline__ int walk_gradient(const RayRec& rr, const DirConst* __restrict__ dirs,
                                             const GridDesc& g, const WalkParams& wp,
                                             const typename Vec4<T>::type* __restrict__ field,
                                             Visitor& vis) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  vis.entry(w.v);
  V4 cur = field[w.v];
  RaySum<T> pint;
  int steps = 0;
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    const V4 nxt = field[w.v];
    pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
    vis.visit(w.v, w.on_surface(), pint.value());
    cur = nxt;
    steps++;
  }
  return steps;
}

// Trilinear sample of the packed field (a_x, a_y, a_z) at the index-space
// point x (0-based voxel centres).  The cell's lower corner is floor(x)
// clamped to [0, size - 2] and the weights are not clamped, so points within
// a voxel of the centres of the faces are extrapolated linearly and a linear
// field is sampled exactly.  Interpolation order x, then y, then z.
template <typename T>
__device__ __forceinline__ void tri_sample(const typename Vec4<T>::type* __restrict__ field,
                                           const Walker& w, const double x[3], T out[3]) {
  using V4 = typename Vec4<T>::type;
  const int size[3] = {w.X, w.Y, w.Z};
  int lo[3];
  T f[3], gw[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int fl = (int)floor(x[a]);
    fl = fl > size[a] - 2 ? size[a] - 2 : fl;
    fl = fl < 0 ? 0 : fl;
    lo[a] = fl;
    const double fd = x[a] - (double)fl;
    f[a] = (T)fd;
    gw[a] = (T)(1.0 - fd);
  }
  const int64_t b = lo[0] + lo[1] * w.sy + lo[2] * w.sz;
  const V4 q000 = field[b], q100 = field[b + 1];
  const V4 q010 = field[b + w.sy], q110 = field[b + w.sy + 1];
  const V4 q001 = field[b + w.sz], q101 = field[b + w.sz + 1];
  const V4 q011 = field[b + w.sy + w.sz], q111 = field[b + w.sy + w.sz + 1];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T c00 = component(q000, c) * gw[0] + component(q100, c) * f[0];
    const T c10 = component(q010, c) * gw[0] + component(q110, c) * f[0];
    const T c01 = component(q001, c) * gw[0] + component(q101, c) * f[0];
    const T c11 = component(q011, c) * gw[0] + component(q111, c) * f[0];
    const T c0 = c00 * gw[1] + c10 * f[1];
    const T c1 = c01 * gw[1] + c11 * f[1];
    out[c] = c0 * gw[2] + c1 * f[2];
  }
}

// Trapezoid of grad p = -rho a over a straight piece of path with index-space
// displacement disp: -rho/2 (ga + gb) . (disp * delta).
template <typename T>
__device__ __forceinline__ T tri_term(const WalkParams& wp, const T ga[3], const T gb[3],
                                      const T disp[3]) {
  const T hr = (T)(-wp.rho * 0.5);
  const T d0 = (T)wp.del[0], d1 = (T)wp.del[1], d2 = (T)wp.del[2];
  return hr * (((ga[0] + gb[0]) * (disp[0] * d0) + (ga[1] + gb[1]) * (disp[1] * d1)) +
               (ga[2] + gb[2]) * (disp[2] * d2));
}

// Gradient walk of the plain method along the straight ray (the accuracy
// mode beside kernel.cu:1743-1863).  The visited voxels are the staircase's;
// the integral to a visited voxel v runs in -> x_in (leg), along the line
// from x_in to x_v with one trapezoid per staircase step (the projection
// x = c - u of a voxel centre moves by s k_a along k), then x_v -> v (leg),
// with the gradient sampled trilinearly on the line.  vis.visit receives the
// carried line integral plus v's own leg.
template <typename T, typename Visitor>
__device__ __forceinline__ int walk_gradient_tri(const RayRec& rr,
                                                 const DirConst* __restrict__ dirs,
                                                 const GridDesc& g, const WalkParams& wp,
                                                 const typename Vec4<T>::type* __restrict__ field,
                                                 Visitor& vis) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  const DirConst& d = dirs[rr.dir_len & 0xffffu];
  const double k0 = d.k[0], k1 = d.k[1], k2 = d.k[2];
  vis.entry(w.v);
  double x[3] = {(double)w.c0 - w.u0, (double)w.c1 - w.u1, (double)w.c2 - w.u2};
  T sl[3], sn[3], ac[3], disp[3];
  tri_sample<T>(field, w, x, sl);
  V4 av = field[w.v];
  ac[0] = av.x;
  ac[1] = av.y;
  ac[2] = av.z;
  disp[0] = (T)(-w.u0);
  disp[1] = (T)(-w.u1);
  disp[2] = (T)(-w.u2);
  RaySum<T> pint;
  pint.add(tri_term<T>(wp, ac, sl, disp));  // in -> x_in
  int steps = 0;
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    x[0] = (double)w.c0 - w.u0;
    x[1] = (double)w.c1 - w.u1;
    x[2] = (double)w.c2 - w.u2;
    tri_sample<T>(field, w, x, sn);
    const double ka = a == 0 ? k0 : (a == 1 ? k1 : k2);
    const double sd = (double)s;
    disp[0] = (T)(sd * (ka * k0));
    disp[1] = (T)(sd * (ka * k1));
    disp[2] = (T)(sd * (ka * k2));
    pint.add(tri_term<T>(wp, sl, sn, disp));  // x_last -> x_next along the line
    av = field[w.v];
    ac[0] = av.x;
    ac[1] = av.y;
    ac[2] = av.z;
    disp[0] = (T)w.u0;
    disp[1] = (T)w.u1;
    disp[2] = (T)w.u2;
    const T leg = tri_term<T>(wp, sn, ac, disp);  // x_next -> v
    vis.visit(w.v, w.on_surface(), pint.value() + (double)leg);
    sl[0] = sn[0];
    sl[1] = sn[1];
    sl[2] = sn[2];
    steps++;
  }
  return steps;
}

// Logical block index that gives each XCD a contiguous range of blocks
// (hardware dispatches workgroups round-robin over the 8 XCDs), so that
// neighbouring ray patches share an L2.
__device__ __forceinline__ unsigned int xcd_block() {
  const unsigned int nb = gridDim.x, b = blockIdx.x;
  const unsigned int x = b & 7u, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Visitor of the plain reconstruction walks.  kFull: every visit adds
// fix(p_src + pint) into acc[v] (and fix(p_src) at the entry).  !kFull: visits
// to source voxels add fix(pint) into acc[sid[v]].
template <bool kFull>
struct PlainAccum {
  long long* acc;
  const int* sid;
  double scale;
  double pin;
  __device__ __forceinline__ void entry(int64_t v) {
    if (kFull) fx_add(&acc[v], to_fixed(pin, scale));
  }
  __device__ __forceinline__ void visit(int64_t v, bool surface, double pint) {
    if (kFull) {
      fx_add(&acc[v], to_fixed(pin + pint, scale));
    } else if (surface) {
      const int s = sid[v];
      if (s >= 0) fx_add(&acc[s], to_fixed(pint, scale));
    }
  }
};

}  // namespace omni3d
