// Reference-schedule iteration: one full traversal of every ray per
// iteration (kernel.cu:8840-8861) for all four integration methods, the
// curl field of the Select / Weighted methods, and the per-iteration
// averaging with the on-device stop rule.
//
// The weights (visit counts, curl weights) and the source flags depend on the
// field only: the first iteration accumulates them and later ones pass null
// pointers and add the pressure sums alone.
//
// One lane walks one ray.  Sums whose weights are whole visit counts (plain,
// plain-mask) are rounded to 64-bit fixed point and added with integer
// atomics, so they do not depend on the order lanes arrive in; the
// curl-weighted methods carry real-valued weights 1/curltmp * counttmp with
// no useful bound, and add value and weight with FP64 atomics instead.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"
#include "launch.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {

__device__ __forceinline__ void add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }

// Plain walk (bodyIntegralInner, kernel.cu:1743-1863) reading p over the
// whole volume: entry self-add, then p_in + pint at every visit.
template <typename T>
__device__ __forceinline__ void naive_plain(const RayRec& rr, const DirConst* __restrict__ dirs,
                                            const GridDesc& g, const WalkParams& wp,
                                            const typename Vec4<T>::type* __restrict__ field,
                                            const double* __restrict__ p, double scale,
                                            long long* __restrict__ sum) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  const double pin = p[w.v];
  fx_add(&sum[w.v], to_fixed(pin, scale));
  V4 cur = field[w.v];
  RaySum<T> pint;
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    const V4 nxt = field[w.v];
    pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
    fx_add(&sum[w.v], to_fixed(pin + pint.value(), scale));
    cur = nxt;
  }
}

// What a walk does where a run (re)starts and where a visit adds.  The
// walks below take the rule once; the sink decides what a visit becomes:
//   PassSink   the pass itself: p is read where a run starts (and the voxel
//              flagged as a run start), each visit adds its value and weight
//              to the sums in the reference's expressions;
//   EntrySink  the visits landing on a run-start voxel as entries (row =
//              that voxel's id, source = the id of the voxel the run
//              started from, weight) of the affine map a pass applies to p
//              on the run starts: counted per row (kFill false) or written
//              at the row's offset plus a slot from its cursor, with the row
//              relative to its tile's first row (the tile layout of
//              k_mtile_accum).  No gradient is read.
struct PassSink {
  using Start = double;
  static constexpr bool kValues = true;
  const double* __restrict__ p;
  double scale;
  long long* __restrict__ sumfx;
  unsigned int* __restrict__ wcnt;
  double* __restrict__ sumd;
  double* __restrict__ wsum;
  uint8_t* __restrict__ src;
  __device__ __forceinline__ double start(int64_t v) const {
    if (src) src[v] = 1;
    return p[v];
  }
  // a visit with unit weight: p0 + pint in 64-bit fixed point, a count
  __device__ __forceinline__ void unit(int64_t v, double p0, double pv) const {
    fx_add(&sumfx[v], to_fixed(p0 + pv, scale));
    if (wcnt) atomicAdd(&wcnt[v], 1u);
  }
  // a visit with curl weight counttmp / curltmp
  __device__ __forceinline__ void curl(int64_t v, double p0, double pv, double c, double n) const {
    add_f64(&sumd[v], (p0 + pv) * 1 / c * n);
    if (wsum) add_f64(&wsum[v], 1 / c * n);
  }
  // the entry voxel's own value (pint = 0) with curl weight
  __device__ __forceinline__ void curl_entry(int64_t v, double p0, double c, double n) const {
    add_f64(&sumd[v], p0 * 1 / c * n);
    if (wsum) add_f64(&wsum[v], 1 / c * n);
  }
};

template <bool kFill>
struct EntrySink {
  using Start = int;
  static constexpr bool kValues = false;
  const int* __restrict__ sid;
  GridDesc g;
  TileGrid tg;
  const int* __restrict__ tfirst;
  unsigned int* __restrict__ cnt;  // entries per row (count) or the rows' cursors (fill)
  const unsigned long long* __restrict__ rowptr;
  unsigned int* __restrict__ esrc;
  unsigned short* __restrict__ erow;
  double* __restrict__ ew;
  __device__ __forceinline__ int start(int64_t v) const { return sid[v]; }
  __device__ __forceinline__ void put(int64_t v, int s0, double w) const {
    const int row = sid[v];
    if (row < 0 || s0 < 0) return;
    if constexpr (!kFill) {
      atomicAdd(&cnt[row], 1u);
    } else {
      const unsigned long long at = rowptr[row] + atomicAdd(&cnt[row], 1u);
      const int i = (int)(v % g.X), j = (int)((v / g.X) % g.Y);
      const int k = (int)(v / ((int64_t)g.X * g.Y));
      esrc[at] = (unsigned int)s0;
      erow[at] = (unsigned short)(row - tfirst[tile_of(i, j, k, tg)]);
      if (ew) ew[at] = w;
    }
  }
  __device__ __forceinline__ void unit(int64_t v, int s0, double) const { put(v, s0, 1.0); }
  __device__ __forceinline__ void curl(int64_t v, int s0, double, double c, double n) const {
    put(v, s0, 1 / c * n);
  }
  __device__ __forceinline__ void curl_entry(int64_t v, int s0, double c, double n) const {
    put(v, s0, 1 / c * n);
  }
};

// Select walk (bodyIntegralInnerSelect, kernel.cu:2280-2537).  aux is the
// |curl a| * mask field (Select) or the mask (plain-mask, thr = +inf, unit
// weights, symmetric rule).  With faithful the per-axis asymmetries of the
// reference are kept: y/z steps test only curl != 0 for the first valid
// voxel and do not treat curl == 0 as leaving the valid zone
// (kernel.cu:2422,2442,2468,2488), and a z-step restart keeps the weight
// accumulators (kernel.cu:2497-2506).  kFixedBC: the step rule of
// bodyIntegralInnerSelectFixedBC (kernel.cu:2538-2796).
template <typename T, bool kUnit, bool kFixedBC, typename Sink>
__device__ __forceinline__ void walk_select(const RayRec& rr, const DirConst* __restrict__ dirs,
                                            const GridDesc& g, const WalkParams& wp,
                                            const typename Vec4<T>::type* __restrict__ field,
                                            const double* __restrict__ aux, const Sink& sk) {
  using V4 = typename Vec4<T>::type;
  const double thr = wp.thr;
  const bool faithful = wp.faithful != 0;
  Walker w;
  w.init(rr, dirs, g);
  typename Sink::Start p0{};
  double curltmp = 0.0, counttmp = 0.0;
  bool p0set = false, outthr = false;
  RaySum<T> pint;
  {
    const double ci = aux[w.v];
    const bool entry_valid = faithful ? !(ci == 0.0 || ci > thr) : (ci != 0.0 && ci < thr);
    if (entry_valid) {
      p0 = sk.start(w.v);
      p0set = true;
    }
  }
  V4 cur{};
  if constexpr (Sink::kValues) cur = field[w.v];
  while (!w.done()) {
    int s;
    const int a = kFixedBC ? w.advance_fixedbc(&s) : w.advance(&s);
    if (a < 0) break;
    V4 nxt{};
    if constexpr (Sink::kValues) nxt = field[w.v];
    const double cn = aux[w.v];
    const bool valid = cn < thr && cn != 0.0;
    if (!p0set) {
      const bool ok = (faithful && a != 0) ? (cn != 0.0) : valid;
      if (ok) {
        p0 = sk.start(w.v);
        p0set = true;
        pint.reset();
        counttmp = 0.0;
        curltmp = 0.0;
      }
    } else {
      if (!outthr && valid) {
        if constexpr (Sink::kValues)
          pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
        const double pv = Sink::kValues ? pint.value() : 0.0;
        if (kUnit) {
          sk.unit(w.v, p0, pv);
        } else {
          curltmp += cn / 10000;
          counttmp += 1.0;
          sk.curl(w.v, p0, pv, curltmp, counttmp);
        }
      }
      const bool enter_high = (faithful && a != 0) ? (!outthr && cn >= thr)
                                                   : ((!outthr && cn >= thr) || cn == 0.0);
      if (enter_high) outthr = true;
      if (outthr && valid) {
        p0 = sk.start(w.v);
        pint.reset();
        outthr = false;
        if (!(faithful && a == 2)) {
          counttmp = 0.0;
          curltmp = 0.0;
        }
      }
    }
    if constexpr (Sink::kValues) cur = nxt;
  }
}

// Weighted walk (bodyIntegralInnerWeighted, kernel.cu:1993-2135): no entry
// self-add, no path break; weight counttmp / curltmp with curltmp the running
// sum of curl / 10^4.  kNoJ0: bodyIntegralInnerWeightedFixedBC
// (kernel.cu:2136-2279), the same walk ended by a step onto the j = 0 plane.
template <typename T, bool kNoJ0, typename Sink>
__device__ __forceinline__ void walk_weighted(const RayRec& rr, const DirConst* __restrict__ dirs,
                                              const GridDesc& g, const WalkParams& wp,
                                              const typename Vec4<T>::type* __restrict__ field,
                                              const double* __restrict__ aux, const Sink& sk) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  const typename Sink::Start pin = sk.start(w.v);
  double curltmp = 0.0, counttmp = 0.0;
  RaySum<T> pint;
  V4 cur{};
  if constexpr (Sink::kValues) cur = field[w.v];
  while (!w.done()) {
    int s;
    const int a = kNoJ0 ? w.advance_noj0(&s) : w.advance(&s);
    if (a < 0) break;
    if constexpr (Sink::kValues) {
      const V4 nxt = field[w.v];
      pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
      cur = nxt;
    }
    curltmp += aux[w.v] / 10000;
    counttmp += 1.0;
    if (curltmp != 0.0) sk.curl(w.v, pin, Sink::kValues ? pint.value() : 0.0, curltmp, counttmp);
  }
}

// One step of the MiniCurl path (kernel.cu:2797-2936): toward the exit
// through the candidate voxel with the smallest curl (ties x, then y, then
// z).  An axis already at the exit offers the reference's sentinel
// (coordinate - 60000): its curl is read at the aliased linear index when
// that lies in the volume, else it counts as 1e10; choosing it ends the
// walk.  Returns the axis (sign in *s_out) or -1.
__device__ __forceinline__ int minicurl_step(Walker& w, const GridDesc& g, int64_t V,
                                             const double* __restrict__ aux, int* s_out) {
  const int s0 = w.c0 < w.e0 ? 1 : (w.c0 > w.e0 ? -1 : -60000);
  const int s1 = w.c1 < w.e1 ? 1 : (w.c1 > w.e1 ? -1 : -60000);
  const int s2 = w.c2 < w.e2 ? 1 : (w.c2 > w.e2 ? -1 : -60000);
  const int64_t i0 = w.v + s0, i1 = w.v + (int64_t)s1 * w.sy, i2 = w.v + (int64_t)s2 * w.sz;
  const double d0 = (i0 >= 0 && i0 < V) ? aux[i0] : 1e10;
  const double d1 = (i1 >= 0 && i1 < V) ? aux[i1] : 1e10;
  const double d2 = (i2 >= 0 && i2 < V) ? aux[i2] : 1e10;
  int a, s, nc, lim;
  int64_t vn;
  if (d0 <= d1 && d0 <= d2) { a = 0; s = s0; nc = w.c0 + s0; lim = g.X; vn = i0; }
  else if (d1 < d0 && d1 <= d2) { a = 1; s = s1; nc = w.c1 + s1; lim = g.Y; vn = i1; }
  else if (d2 < d0 && d2 < d1) { a = 2; s = s2; nc = w.c2 + s2; lim = g.Z; vn = i2; }
  else return -1;
  if (nc < 0 || nc >= lim) return -1;
  if (a == 0) w.c0 = nc;
  else if (a == 1) w.c1 = nc;
  else w.c2 = nc;
  w.v = vn;
  *s_out = s;
  return a;
}

// MiniCurl walk (bodyIntegralInnerMiniCurl, kernel.cu:2797-2936, the pass of
// omni3dparallellinesESInnerWeightedMiniCurl): plain sums and unit counts,
// no entry self-add, along the MiniCurl path.
template <typename T, typename Sink>
__device__ __forceinline__ void walk_minicurl(const RayRec& rr, const DirConst* __restrict__ dirs,
                                              const GridDesc& g, const WalkParams& wp,
                                              const typename Vec4<T>::type* __restrict__ field,
                                              const double* __restrict__ aux, const Sink& sk) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  const typename Sink::Start pin = sk.start(w.v);
  V4 cur{};
  if constexpr (Sink::kValues) cur = field[w.v];
  RaySum<T> pint;
  while (!w.done()) {
    int s;
    const int a = minicurl_step(w, g, V, aux, &s);
    if (a < 0) break;
    if constexpr (Sink::kValues) {
      const V4 nxt = field[w.v];
      pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
      cur = nxt;
    }
    sk.unit(w.v, pin, Sink::kValues ? pint.value() : 0.0);
  }
}

// WeightedMiniCurl walk (bodyIntegralInnerWeightedMiniCurl,
// kernel.cu:2937-3085): the MiniCurl path; the entry adds its own p with
// weight 1 / curl[entry] (curltmp = curl there, counttmp = 1, kernel.cu:2956-
// 2960), every step adds (p_in + pint) * counttmp / curltmp with curltmp the
// running sum of curl (no / 10^4) and no test of curltmp against zero.
template <typename T, typename Sink>
__device__ __forceinline__ void walk_weighted_minicurl(
    const RayRec& rr, const DirConst* __restrict__ dirs, const GridDesc& g, const WalkParams& wp,
    const typename Vec4<T>::type* __restrict__ field, const double* __restrict__ aux,
    const Sink& sk) {
  using V4 = typename Vec4<T>::type;
  Walker w;
  w.init(rr, dirs, g);
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  const typename Sink::Start pin = sk.start(w.v);
  double curltmp = aux[w.v], counttmp = 1.0;
  sk.curl_entry(w.v, pin, curltmp, counttmp);
  V4 cur{};
  if constexpr (Sink::kValues) cur = field[w.v];
  RaySum<T> pint;
  while (!w.done()) {
    int s;
    const int a = minicurl_step(w, g, V, aux, &s);
    if (a < 0) break;
    if constexpr (Sink::kValues) {
      const V4 nxt = field[w.v];
      pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
      cur = nxt;
    }
    curltmp += aux[w.v];
    counttmp += 1.0;
    sk.curl(w.v, pin, Sink::kValues ? pint.value() : 0.0, curltmp, counttmp);
  }
}

template <typename T, int kMethod>
__global__ void __launch_bounds__(256)
k_walk_naive(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, WalkParams wp, const typename Vec4<T>::type* __restrict__ field,
             const double* __restrict__ aux, const double* __restrict__ p0,
             const double* __restrict__ p1, int k, const DevScalars* __restrict__ sc,
             long long* __restrict__ sumfx, unsigned int* __restrict__ wcnt,
             double* __restrict__ sumd, double* __restrict__ wsum, uint8_t* __restrict__ src) {
  if (sc->converged) return;
  // iteration k reads the pressure of iteration k-1
  const double* __restrict__ p = ((k - 1) & 1) ? p1 : p0;
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const RayRec rr = rays[r];
  const PassSink sk{p, sc->scale, sumfx, wcnt, sumd, wsum, src};
  if (kMethod == 0)
    naive_plain<T>(rr, dirs, g, wp, field, p, sc->scale, sumfx);
  else if (kMethod == 1)
    walk_select<T, true, false>(rr, dirs, g, wp, field, aux, sk);
  else if (kMethod == 2)
    walk_select<T, false, false>(rr, dirs, g, wp, field, aux, sk);
  else if (kMethod == 4)
    walk_select<T, false, true>(rr, dirs, g, wp, field, aux, sk);
  else if (kMethod == 5)
    walk_minicurl<T>(rr, dirs, g, wp, field, aux, sk);
  else if (kMethod == 6)
    walk_weighted<T, true>(rr, dirs, g, wp, field, aux, sk);
  else if (kMethod == 7)
    walk_weighted_minicurl<T>(rr, dirs, g, wp, field, aux, sk);
  else
    walk_weighted<T, false>(rr, dirs, g, wp, field, aux, sk);
}

// The entries of the affine map of one pass on the run starts, one lane per
// ray, for the methods whose walks leave the plan's staircase (SelectFixedBC,
// MiniCurl, WeightedFixedBC, WeightedMiniCurl): the walk of the pass with an
// EntrySink (no gradient read).
template <int kMethod, bool kFill>
__global__ void __launch_bounds__(256)
k_lane_rows(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, TileGrid tg, WalkParams wp, const double* __restrict__ aux,
            const int* __restrict__ sid, const int* __restrict__ tfirst,
            unsigned int* __restrict__ cnt, const unsigned long long* __restrict__ rowptr,
            unsigned int* __restrict__ esrc, unsigned short* __restrict__ erow,
            double* __restrict__ ew) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const RayRec rr = rays[r];
  const EntrySink<kFill> sk{sid, g, tg, tfirst, cnt, rowptr, esrc, erow, ew};
  using V4 = Vec4<float>::type;
  const V4* nofield = nullptr;
  if (kMethod == 4)
    walk_select<float, false, true>(rr, dirs, g, wp, nofield, aux, sk);
  else if (kMethod == 5)
    walk_minicurl<float>(rr, dirs, g, wp, nofield, aux, sk);
  else if (kMethod == 6)
    walk_weighted<float, true>(rr, dirs, g, wp, nofield, aux, sk);
  else
    walk_weighted_minicurl<float>(rr, dirs, g, wp, nofield, aux, sk);
}

// Averaging of one iteration (devidecountInner, kernel.cu:6746-6759: where
// the weight is > 0, p = pn / weight, elsewhere p keeps its value) and the
// relative change over the voxels whose p the rays read.  The first
// iteration adds the initial pressure to the sums when pn starts as a copy
// of p (kernel.cu:8813).  The last block reduces and sets the stop flag.
constexpr int kUpdBlock = 256;

__global__ void __launch_bounds__(kUpdBlock)
k_naive_update(int k, int64_t V, int fixed, const long long* __restrict__ sumfx,
               const unsigned int* __restrict__ cnt, const double* __restrict__ sumd,
               const double* __restrict__ wsum, const double* __restrict__ pinit, int seed,
               const int* __restrict__ sid, const uint8_t* __restrict__ src,
               double* __restrict__ p0, double* __restrict__ p1, DevScalars* __restrict__ sc,
               double* __restrict__ partials, int max_iter, double tol) {
  if (sc->converged) return;
  constexpr int kWaves = kUpdBlock / 64;
  __shared__ double s_d2[kWaves], s_n2[kWaves];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double* __restrict__ pold = ((k - 1) & 1) ? p1 : p0;
  double* __restrict__ pnew = (k & 1) ? p1 : p0;
  const double inv = sc->inv_scale;
  const bool add_seed = seed && pinit && k == 1;
  double d2 = 0.0, n2 = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double po = pold[v];
    double pn = po;
    if (fixed) {
      const unsigned int c = cnt[v];
      if (c > 0) {
        double tot = (double)sumfx[v] * inv;
        if (add_seed) tot += pinit[v];
        pn = tot / (double)c;
      }
    } else {
      const double w = wsum[v];
      if (w > 0.0) {
        double tot = sumd[v];
        if (add_seed) tot += pinit[v];
        pn = tot / w;
      }
    }
    pnew[v] = pn;
    const bool is_src = sid ? sid[v] >= 0 : src[v] != 0;
    if (is_src) {
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    d2 += __shfl_xor(d2, off);
    n2 += __shfl_xor(n2, off);
  }
  if (lane == 0) {
    s_d2[wave] = d2;
    s_n2[wave] = n2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kWaves; w++) {
      a += s_d2[w];
      b += s_n2[w];
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
    __threadfence();
    const unsigned int t =
        __hip_atomic_fetch_add(&sc->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double a = 0.0, b = 0.0;
  for (unsigned int i = threadIdx.x; i < gridDim.x; i += kUpdBlock) {
    a += __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b += __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  __syncthreads();
  if (lane == 0) {
    s_d2[wave] = a;
    s_n2[wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < kWaves; w++) {
      a += s_d2[w];
      b += s_n2[w];
    }
    const double num = sqrt(a), den = sqrt(b);
    const double rel = den > 0.0 ? num / den : (num > 0.0 ? INFINITY : 0.0);
    sc->rel = rel;
    sc->dnum = num;
    sc->dden = den;
    sc->iter_done = k;
    if ((tol > 0.0 && rel < tol) || k >= max_iter) {
      sc->converged = 1;
      sc->kstar = k;
    }
    sc->ticket = 0;
  }
}

// Output of a naive reconstruction: the pressure of iteration k* and the
// weights of its pass.
__global__ void __launch_bounds__(256)
k_naive_output(int64_t V, const double* __restrict__ p0, const double* __restrict__ p1,
               const DevScalars* __restrict__ sc, int fixed, const unsigned int* __restrict__ cnt,
               const double* __restrict__ wsum, double* __restrict__ pd, float* __restrict__ pf,
               double* __restrict__ wd, float* __restrict__ wf) {
  const double* __restrict__ p = (sc->kstar & 1) ? p1 : p0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double x = p[v];
    const double w = fixed ? (double)cnt[v] : wsum[v];
    if (pd) pd[v] = x;
    if (pf) pf[v] = (float)x;
    if (wd) wd[v] = w;
    if (wf) wf[v] = (float)w;
  }
}

// |curl a| * mask (calCurlofMaterialAccCPU, kernel.cu:8352-8379): central
// differences inside, one-sided at the walls, evaluated in FP64 in the
// reference's operation order.
template <typename Load>
__device__ __forceinline__ double curl_at(int X, int Y, int Z, int i, int j, int k, double dx,
                                          double dy, double dz, const Load& ld) {
  const int i0 = i - 1 >= 0 ? i - 1 : i, j0 = j - 1 >= 0 ? j - 1 : j, k0 = k - 1 >= 0 ? k - 1 : k;
  const int ie = i + 1 <= X - 1 ? i + 1 : i, je = j + 1 <= Y - 1 ? j + 1 : j,
            ke = k + 1 <= Z - 1 ? k + 1 : k;
  auto at = [&](int c, int a, int b, int d) {
    return ld(c, (int64_t)a + (int64_t)b * X + (int64_t)d * X * Y);
  };
  double cx = (at(2, i, je, k) - at(2, i, j0, k)) / (double)(je - j0) / dy;
  cx += -(at(1, i, j, ke) - at(1, i, j, k0)) / (double)(ke - k0) / dz;
  double cy = -(at(2, ie, j, k) - at(2, i0, j, k)) / (double)(ie - i0) / dx;
  cy += (at(0, i, j, ke) - at(0, i, j, k0)) / (double)(ke - k0) / dz;
  double cz = (at(1, ie, j, k) - at(1, i0, j, k)) / (double)(ie - i0) / dx;
  cz += -(at(0, i, je, k) - at(0, i, j0, k)) / (double)(je - j0) / dy;
  return sqrt(cx * cx + cy * cy + cz * cz);
}

template <typename V4>
__global__ void __launch_bounds__(256)
k_curl_packed(int X, int Y, int Z, double dx, double dy, double dz, const V4* __restrict__ f,
              const double* __restrict__ mask, double* __restrict__ out) {
  const int64_t V = (int64_t)X * Y * Z;
  auto ld = [&](int c, int64_t v) -> double { return (double)component(f[v], c); };
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % X), j = (int)((v / X) % Y), k = (int)(v / ((int64_t)X * Y));
    const double c = curl_at(X, Y, Z, i, j, k, dx, dy, dz, ld);
    out[v] = c * (mask ? mask[v] : 1.0);
  }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_curl_arrays(int X, int Y, int Z, double dx, double dy, double dz, const S* __restrict__ ax,
              const S* __restrict__ ay, const S* __restrict__ az, const S* __restrict__ mask,
              S* __restrict__ out) {
  const int64_t V = (int64_t)X * Y * Z;
  const S* comp[3] = {ax, ay, az};
  auto ld = [&](int c, int64_t v) -> double { return (double)comp[c][v]; };
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % X), j = (int)((v / X) % Y), k = (int)(v / ((int64_t)X * Y));
    const double c = curl_at(X, Y, Z, i, j, k, dx, dy, dz, ld);
    out[v] = (S)(c * (mask ? (double)mask[v] : 1.0));
  }
}

// Fills n 32-bit words (a kernel, not a memset: it may be captured in a graph).
__global__ void __launch_bounds__(256) k_fill_u32(unsigned int* __restrict__ x, int64_t n,
                                                  unsigned int v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// Adds one pass's sums and weights into the caller's FP32 pn / pcountinner
// (the kernel-level drop-ins accumulate like the reference kernels do).
__global__ void __launch_bounds__(256)
k_accumulate_out(int64_t V, int fixed, const long long* __restrict__ sumfx,
                 const unsigned int* __restrict__ cnt, const double* __restrict__ sumd,
                 const double* __restrict__ wsum, const DevScalars* __restrict__ sc,
                 float* __restrict__ pn, float* __restrict__ pc) {
  const double inv = sc->inv_scale;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double s = fixed ? (double)sumfx[v] * inv : sumd[v];
    const double w = fixed ? (double)cnt[v] : wsum[v];
    if (w != 0.0 || s != 0.0) {
      pn[v] = (float)((double)pn[v] + s);
      pc[v] = (float)((double)pc[v] + w);
    }
  }
}

// devidecountInner (kernel.cu:6746-6759): where pcountinner > 0,
// p = pn / pcountinner and pn = 0.
__global__ void __launch_bounds__(256)
k_divide_count(int64_t V, float* __restrict__ p, float* __restrict__ pn,
               const float* __restrict__ pc) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const float c = pc[v];
    if (c > 0.0f) {
      p[v] = pn[v] / c;
      pn[v] = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256) k_fill_f64(double* __restrict__ x, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// ---------------------------------------------------------------------------
// Launchers.

template <typename T>
hipError_t launch_walk_naive(int method, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                             GridDesc g, const WalkParams& wp, const void* field,
                             const double* aux, const double* p0, const double* p1, int k,
                             const DevScalars* sc, long long* sumfx, unsigned int* wcnt,
                             double* sumd, double* wsum, uint8_t* src, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const unsigned int grid = grid_for(nrays, 1u << 30);
  const V4* f = static_cast<const V4*>(field);
#define OMNI_NAIVE(M)                                                                          \
  hipLaunchKernelGGL((k_walk_naive<T, M>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g, \
                     wp, f, aux, p0, p1, k, sc, sumfx, wcnt, sumd, wsum, src)
  switch (method) {
    case 0: OMNI_NAIVE(0); break;
    case 1: OMNI_NAIVE(1); break;
    case 2: OMNI_NAIVE(2); break;
    case 4: OMNI_NAIVE(4); break;
    case 5: OMNI_NAIVE(5); break;
    case 6: OMNI_NAIVE(6); break;
    case 7: OMNI_NAIVE(7); break;
    default: OMNI_NAIVE(3); break;
  }
#undef OMNI_NAIVE
  return hipGetLastError();
}
template hipError_t launch_walk_naive<float>(int, const RayRec*, int64_t, const DirConst*,
                                             GridDesc, const WalkParams&, const void*,
                                             const double*, const double*, const double*, int,
                                             const DevScalars*, long long*, unsigned int*,
                                             double*, double*, uint8_t*, hipStream_t);
template hipError_t launch_walk_naive<double>(int, const RayRec*, int64_t, const DirConst*,
                                              GridDesc, const WalkParams&, const void*,
                                              const double*, const double*, const double*, int,
                                              const DevScalars*, long long*, unsigned int*,
                                              double*, double*, uint8_t*, hipStream_t);

hipError_t launch_lane_rows(bool fill, int method, const RayRec* rays, int64_t nrays,
                            const DirConst* dirs, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const double* aux, const int* sid, const int* tfirst,
                            unsigned int* cnt, const unsigned long long* rowptr,
                            unsigned int* esrc, unsigned short* erow, double* ew, hipStream_t st) {
  const unsigned int grid = grid_for(nrays, 1u << 30);
#define OMNI_LR(M, FL)                                                                         \
  hipLaunchKernelGGL((k_lane_rows<M, FL>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g, \
                     tg, wp, aux, sid, tfirst, cnt, rowptr, esrc, erow, ew)
#define OMNI_LR_M(FL)                      \
  switch (method) {                        \
    case 4: OMNI_LR(4, FL); break;         \
    case 5: OMNI_LR(5, FL); break;         \
    case 6: OMNI_LR(6, FL); break;         \
    case 7: OMNI_LR(7, FL); break;         \
    default: return hipErrorInvalidValue;  \
  }
  if (fill) {
    OMNI_LR_M(true)
  } else {
    OMNI_LR_M(false)
  }
#undef OMNI_LR_M
#undef OMNI_LR
  return hipGetLastError();
}

// tbase[t] = rowptr[tfirst[t]], t = 0..nt: the first entry of each tile's rows
__global__ void __launch_bounds__(256)
k_tile_bases(int nt, const int* __restrict__ tfirst, const unsigned long long* __restrict__ rowptr,
             unsigned long long* __restrict__ tbase) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= nt) tbase[t] = rowptr[tfirst[t]];
}

hipError_t launch_tile_bases(int nt, const int* tfirst, const unsigned long long* rowptr,
                             unsigned long long* tbase, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_bases, dim3((unsigned int)((nt + 1 + 255) / 256)), dim3(256), 0, st, nt,
                     tfirst, rowptr, tbase);
  return hipGetLastError();
}

unsigned int naive_update_grid(int64_t V) { return grid_for(V, 2048); }

hipError_t launch_naive_update(int k, int64_t V, int fixed, const long long* sumfx,
                               const unsigned int* cnt, const double* sumd, const double* wsum,
                               const double* pinit, int seed, const int* sid, const uint8_t* src,
                               double* p0, double* p1, DevScalars* sc, double* partials,
                               int max_iter, double tol, hipStream_t st) {
  hipLaunchKernelGGL(k_naive_update, dim3(naive_update_grid(V)), dim3(kUpdBlock), 0, st, k, V,
                     fixed, sumfx, cnt, sumd, wsum, pinit, seed, sid, src, p0, p1, sc, partials,
                     max_iter, tol);
  return hipGetLastError();
}

hipError_t launch_naive_output(int64_t V, const double* p0, const double* p1,
                               const DevScalars* sc, int fixed, const unsigned int* cnt,
                               const double* wsum, double* pd, float* pf, double* wd, float* wf,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_naive_output, dim3(grid_for(V)), dim3(256), 0, st, V, p0, p1, sc, fixed,
                     cnt, wsum, pd, pf, wd, wf);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_curl_packed(int X, int Y, int Z, double dx, double dy, double dz,
                              const void* field, const double* mask, double* out,
                              hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const int64_t V = (int64_t)X * Y * Z;
  hipLaunchKernelGGL((k_curl_packed<V4>), dim3(grid_for(V)), dim3(256), 0, st, X, Y, Z, dx, dy,
                     dz, static_cast<const V4*>(field), mask, out);
  return hipGetLastError();
}
template hipError_t launch_curl_packed<float>(int, int, int, double, double, double, const void*,
                                              const double*, double*, hipStream_t);
template hipError_t launch_curl_packed<double>(int, int, int, double, double, double,
                                               const void*, const double*, double*, hipStream_t);

hipError_t launch_curl_f32(int X, int Y, int Z, double dx, double dy, double dz, const float* ax,
                           const float* ay, const float* az, const float* mask, float* out,
                           hipStream_t st) {
  const int64_t V = (int64_t)X * Y * Z;
  hipLaunchKernelGGL((k_curl_arrays<float>), dim3(grid_for(V)), dim3(256), 0, st, X, Y, Z, dx,
                     dy, dz, ax, ay, az, mask, out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(void* x, int64_t words, unsigned int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(words)), dim3(256), 0, st,
                     static_cast<unsigned int*>(x), words, v);
  return hipGetLastError();
}

hipError_t launch_accumulate_out(int64_t V, int fixed, const long long* sumfx,
                                 const unsigned int* cnt, const double* sumd, const double* wsum,
                                 const DevScalars* sc, float* pn, float* pc, hipStream_t st) {
  hipLaunchKernelGGL(k_accumulate_out, dim3(grid_for(V)), dim3(256), 0, st, V, fixed, sumfx, cnt,
                     sumd, wsum, sc, pn, pc);
  return hipGetLastError();
}

hipError_t launch_divide_count(int64_t V, float* p, float* pn, const float* pc, hipStream_t st) {
  hipLaunchKernelGGL(k_divide_count, dim3(grid_for(V)), dim3(256), 0, st, V, p, pn, pc);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
k_mask_diff(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
            unsigned int* __restrict__ diff) {
  bool d = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d |= __double_as_longlong(a[i]) != __double_as_longlong(b[i]);
  if (__any(d) && (threadIdx.x & 63) == 0) *diff = 1u;
}

hipError_t launch_mask_diff(const double* a, const double* b, int64_t n, unsigned int* diff,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_mask_diff, dim3(grid_for(n, 4096)), dim3(256), 0, st, a, b, n, diff);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double* x, int64_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(256), 0, st, x, n, v);
  return hipGetLastError();
}

}  // namespace omni3d
