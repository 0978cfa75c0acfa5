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
