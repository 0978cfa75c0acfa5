// Tile engine for the reference's schedule -- a full traversal per
// iteration (kernel.cu:8840-8861) -- of every integration method: the
// masked ones (plain-mask, Select, Weighted) and plain (the rule that never
// breaks a path).
//
// The path-break rule of these methods (bodyIntegralInnerSelect,
// kernel.cu:2280-2537; bodyIntegralInnerWeighted, kernel.cu:1993-2135)
// depends on the curl / mask field alone, never on p: where a ray's runs
// start and stop, which visits add, and their curl weights are fixed for a
// field.  Only the value a run carries depends on p, through the pressure at
// the voxel the run (re)starts from.  So
//   once per field:  k_mask_walk   one lane per ray replays the rule on the
//                    curl / mask alone and stores, per segment (ray x tile,
//                    in its value slot), how its first visit is taken (with
//                    the term of the step into the tile, the only gradient
//                    it reads) and the voxel of its last restart, and sets
//                    the source flags; then k_tile_masked<setup> replays the
//                    rule per tile from those records with the gradient in
//                    LDS: each segment's sum of added terms after its last
//                    restart, and the weights (visit counts or curl weights);
//   per iteration:   k_mask_carry  one lane per ray chains its segments: the
//                    value a run enters each segment with;
//                    k_tile_masked one workgroup per tile stages the tile's
//                    gradient and curl / mask in LDS, replays the rule over
//                    the tile's segments and adds every visit's value in LDS
//                    (64-bit fixed point for unit weights, FP64 otherwise),
//                    then stores the tile's sums -- no global atomics.
// The first visit of a segment (the step into the tile) is taken by
// k_mask_walk, which reads the voxel outside the tile from memory; the tile
// pass walks the visits after it.
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"
#include "launch.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {

// method ids as in k_walk_naive: 0 plain (bodyIntegralInner: every step
// adds, the entry adds its own p, weights = the plan's visit counts),
// 1 plain-mask (unit weights), 2 Select, 3 Weighted
enum { kMaskPlain = 0, kMaskUnit = 1, kMaskSelect = 2, kMaskWeighted = 3 };
// integer (fixed-point) sums for whole-count weights, FP64 for curl weights
template <int kMethod>
constexpr bool kCurlW = kMethod == kMaskSelect || kMethod == kMaskWeighted;

// flags of a segment's first visit
constexpr uint32_t kModeNone = 0u;     // no run: nothing carried
constexpr uint32_t kModeCont = 1u;     // continues the run: carry + E
constexpr uint32_t kModeRestart = 2u;  // (re)starts at its first voxel: p there
constexpr uint32_t kFlagAdd = 4u;      // the first visit adds
constexpr uint32_t kFlagP0 = 8u;       // run started (p0set) after the first visit
constexpr uint32_t kFlagOut = 16u;     // inside a high / masked stretch (outthr)
constexpr uint32_t kFlagRestart = 32u; // some visit of the segment restarts the run

// The rule of one step onto a voxel with curl / mask value cn along axis a
// (naive_select: the reference's asymmetries when faithful).  add: the visit
// adds p0 + pint (the step's term joins pint first); restart: p0 = p here,
// pint = 0, and the curl weights restart unless keep_w.
struct StepRule {
  bool add, restart, keep_w;
};

template <int kMethod>
__device__ __forceinline__ StepRule rule_step(const WalkParams& wp, int a, double cn, bool& p0set,
                                              bool& outthr) {
  StepRule r{false, false, false};
  if (kMethod == kMaskWeighted || kMethod == kMaskPlain) {
    r.add = true;  // every step adds (Weighted: when the curl sum is nonzero)
    return r;
  }
  const double thr = wp.thr;
  const bool faithful = wp.faithful != 0;
  const bool valid = cn < thr && cn != 0.0;
  if (!p0set) {
    const bool ok = (faithful && a != 0) ? (cn != 0.0) : valid;
    if (ok) {
      p0set = true;
      r.restart = true;
    }
  } else {
    r.add = !outthr && valid;
    const bool enter_high = (faithful && a != 0) ? (!outthr && cn >= thr)
                                                 : ((!outthr && cn >= thr) || cn == 0.0);
    if (enter_high) outthr = true;
    if (outthr && valid) {
      r.restart = true;
      outthr = false;
      r.keep_w = faithful && a == 2;
    }
  }
  return r;
}

// The entry voxel of a ray: a run starts there (p0 = p, no add) when valid.
template <int kMethod>
__device__ __forceinline__ bool rule_entry(const WalkParams& wp, double ci) {
  if (kMethod == kMaskWeighted || kMethod == kMaskPlain) return true;
  const double thr = wp.thr;
  return wp.faithful ? !(ci == 0.0 || ci > thr) : (ci != 0.0 && ci < thr);
}

// The visit's weight (curl methods): 1 / curltmp * counttmp as the
// reference writes it, and the visit's value times it.
__device__ __forceinline__ double curl_weight(double curl, double cnt) { return 1 / curl * cnt; }
__device__ __forceinline__ double curl_value(double c, double curl, double cnt) {
  return c * 1 / curl * cnt;
}

// Per-segment record of a field (value-slot indexed); S comes from the
// setup tile pass, the rest from k_mask_walk.
struct MaskSegs {
  double* E;       // term of the step into the tile when it joins the run
  uint32_t* R;     // voxel of the last restart
  uint32_t* F;     // first-visit flags
  double* wc;      // curl weights after the first visit (curl methods):
  double* wn;      // curltmp, counttmp
};

template <typename T, int kMethod>
__global__ void __launch_bounds__(256)
k_mask_walk(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, TileGrid tg, WalkParams wp, const typename Vec4<T>::type* __restrict__ field,
            const double* __restrict__ aux, const unsigned long long* __restrict__ val_off,
            MaskSegs ms, uint8_t* __restrict__ src) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  Walker w;
  w.init(rays[r], dirs, g);
  unsigned long long gs = val_off[r >> 6] + (unsigned long long)(r & 63);
  int tile = tile_of(w.c0, w.c1, w.c2, tg);
  bool p0set = false, outthr = false;
  double curl = 0.0, cnt = 0.0;
  double E = 0.0;
  uint32_t R = 0, F;
  bool hasR = false;
  {
    const double ci = kMethod == kMaskPlain ? 0.0 : aux[w.v];
    if (rule_entry<kMethod>(wp, ci)) {
      p0set = true;
      hasR = true;
      R = (uint32_t)w.v;
      src[w.v] = 1;
      // the plain entry adds its own p (pint = 0)
      F = kModeRestart | (kMethod == kMaskPlain ? kFlagAdd : 0u);
    } else {
      F = kModeNone;
    }
    F |= (p0set ? kFlagP0 : 0u) | (outthr ? kFlagOut : 0u);
  }
  double c0 = curl, n0 = cnt;
  auto store = [&]() {
    ms.E[gs] = E;
    ms.R[gs] = R;
    ms.F[gs] = F | (hasR ? kFlagRestart : 0u);
    if (kCurlW<kMethod>) {
      ms.wc[gs] = c0;
      ms.wn[gs] = n0;
    }
    gs += 64;
  };
  while (!w.done()) {
    const int64_t vprev = w.v;
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    const double cn = kMethod == kMaskPlain ? 0.0 : aux[w.v];
    const int t_now = tile_of(w.c0, w.c1, w.c2, tg);
    const bool enter = t_now != tile;
    if (enter) {
      store();
      tile = t_now;
      E = 0.0;
      hasR = false;
    }
    const StepRule d = rule_step<kMethod>(wp, a, cn, p0set, outthr);
    bool added = false;
    if (d.add) {
      if (kCurlW<kMethod>) {
        curl += cn / 10000;
        cnt += 1.0;
      }
      added = kMethod != kMaskWeighted || curl != 0.0;
      // the step into a tile: its term (the in-tile terms are summed by the
      // setup tile pass from the gradient staged in LDS)
      if (enter)
        E = (double)trapezoid<T>(wp, a, s, component(field[w.v], a), component(field[vprev], a));
    }
    if (d.restart) {
      hasR = true;
      R = (uint32_t)w.v;
      src[w.v] = 1;
      if (!d.keep_w) {
        curl = 0.0;
        cnt = 0.0;
      }
    }
    if (enter) {
      F = (d.restart ? kModeRestart : (p0set ? kModeCont : kModeNone)) | (added ? kFlagAdd : 0u) |
          (p0set ? kFlagP0 : 0u) | (outthr ? kFlagOut : 0u);
      c0 = curl;
      n0 = cnt;
    }
  }
  store();
}

// Carry chain of one iteration: the value each segment's first visit
// continues from (mode "continue": the run's value at the previous
// segment's exit plus the step into the tile).  A segment that restarts
// leaves with p at its last restart plus the terms after it.
__global__ void __launch_bounds__(256)
k_mask_carry(int64_t nrays, const unsigned int* __restrict__ nseg,
             const unsigned long long* __restrict__ val_off, const double* __restrict__ E,
             const double* __restrict__ S, const uint32_t* __restrict__ R,
             const uint32_t* __restrict__ F, const double* __restrict__ p0,
             const double* __restrict__ p1, int k, const DevScalars* __restrict__ sc,
             double* __restrict__ cval) {
  if (sc->converged) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const double* __restrict__ p = ((k - 1) & 1) ? p1 : p0;
  unsigned long long gs = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  double prev = 0.0;
  for (unsigned int j = 0; j < n; j++, gs += 64) {
    const uint32_t f = F[gs];
    const double cf = (f & 3u) == kModeCont ? prev + E[gs] : 0.0;
    cval[gs] = cf;
    prev = (f & kFlagRestart) ? p[R[gs]] + S[gs] : cf + S[gs];
  }
}

template <typename T, int kMethod, int kVox>
struct MaskLds {
  using Acc = typename std::conditional<kCurlW<kMethod>, double, long long>::type;
  T g[3][kVox];
  Acc acc[kVox];
  // curl / mask values and the previous iteration's p (read where a run
  // restarts) in LDS where they fit (FP32 gradient)
  double aux[sizeof(T) == 4 && kMethod != kMaskPlain ? kVox : 1];
  double pl[sizeof(T) == 4 ? kVox : 1];
};

// kSetup (once per field): instead of the visits' values, the visits'
// weights (visit counts or curl weights, into wcnt / wsum) and each
// segment's sum of added terms after its last restart (into segS).
template <typename T, int kMethod, int kBits, int kThreads, bool kSetup>
__global__ void __launch_bounds__(kThreads)
k_tile_masked(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
              GridDesc g, TileGrid tg, WalkParams wp,
              const typename Vec4<T>::type* __restrict__ field, const double* __restrict__ aux,
              const uint32_t* __restrict__ F, const double* __restrict__ wc,
              const double* __restrict__ wn, const double* __restrict__ cval,
              const double* __restrict__ p0, const double* __restrict__ p1, int k,
              const DevScalars* __restrict__ sc, long long* __restrict__ sumfx,
              double* __restrict__ sumd, double* __restrict__ segS,
              unsigned int* __restrict__ wcnt, double* __restrict__ wsum) {
  using V4 = typename Vec4<T>::type;
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kVox = kTile * kTile * kTile;
  constexpr bool kAuxLds = sizeof(T) == 4;
  constexpr bool kAux = kMethod != kMaskPlain;  // the plain rule reads no curl / mask
  __shared__ MaskLds<T, kMethod, kVox> L;
  if (sc->converged) return;
  const double* __restrict__ p = ((k - 1) & 1) ? p1 : p0;
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), kk = oz + (l >> (2 * kBits));
    const bool in = i < g.X && j < g.Y && kk < g.Z;
    V4 v;
    double cv = 0.0, pv = 0.0;
    if (in) {
      const int64_t gv = i + j * sy + kk * sz;
      v = field[gv];
      if (kAuxLds) {
        if (kAux) cv = aux[gv];
        if (!kSetup) pv = p[gv];
      }
    } else {
      v.x = v.y = v.z = v.w = 0;
    }
    const int sl = lds_slot<kBits>(l);
    L.g[0][sl] = v.x;
    L.g[1][sl] = v.y;
    L.g[2][sl] = v.z;
    L.acc[sl] = 0;
    if (kAuxLds) {
      if (kAux) L.aux[sl] = cv;
      L.pl[sl] = pv;
    }
  }
  __syncthreads();
  const double scale = sc->scale;
  // global voxel of in-tile index l
  auto gvox = [&](int l) -> int64_t {
    return (int64_t)(ox + (l & kTileMask)) + (int64_t)(oy + ((l >> kBits) & kTileMask)) * sy +
           (int64_t)(oz + (l >> (2 * kBits))) * sz;
  };
  auto add = [&](int sl, double c, double curl, double cnt) {
    if constexpr (kSetup) {
      // the plain method's weights are the plan's visit counts
      if constexpr (kMethod == kMaskUnit)
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[sl]), 1ull);
      else if constexpr (kCurlW<kMethod>)
        atomicAdd(&L.acc[sl], curl_weight(curl, cnt));
    } else if constexpr (!kCurlW<kMethod>) {
      atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[sl]),
                (unsigned long long)to_fixed(c, scale));
    } else {
      atomicAdd(&L.acc[sl], curl_value(c, curl, cnt));
    }
  };
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const uint32_t f = F[sr.g];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool up0 = sg & 1u, up1 = (sg >> 1) & 1u, up2 = (sg >> 2) & 1u;
    const int d0 = up0 ? 1 : -1;
    const int d1 = up1 ? kTile : -kTile;
    const int d2 = up2 ? kTile * kTile : -kTile * kTile;
    const T h0 = step_factor<T>(wp, 0, up0);
    const T h1 = step_factor<T>(wp, 1, up1);
    const T h2 = step_factor<T>(wp, 2, up2);
    bool p0set = (f & kFlagP0) != 0, outthr = (f & kFlagOut) != 0;
    const uint32_t mode = f & 3u;
    double c = 0.0, S = 0.0;
    if constexpr (!kSetup) {
      if (mode == kModeCont) c = cval[sr.g];
      else if (mode == kModeRestart) c = kAuxLds ? L.pl[lds_slot<kBits>(lv)] : p[gvox(lv)];
    }
    double curl = 0.0, cnt = 0.0;
    if (kCurlW<kMethod>) {
      curl = wc[sr.g];
      cnt = wn[sr.g];
    }
    int slv = lds_slot<kBits>(lv);
    if (f & kFlagAdd) add(slv, c, curl, cnt);
    uint32_t cw = sr.code[0];
    int rem = (int)n - 1;
#pragma unroll 1
    for (int wi = 0; wi < 3 && rem > 0; wi++) {
      const int m = rem < 16 ? rem : 16;
      for (int q = 0; q < m; q++) {
        const int a = (int)(cw & 3u);
        cw >>= 2;
        const int ln = lv + (a == 0 ? d0 : (a == 1 ? d1 : d2));
        const int sln = lds_slot<kBits>(ln);
        double cn = 0.0;
        if constexpr (!kAux) cn = 0.0;
        else if constexpr (kAuxLds) cn = L.aux[sln];
        else cn = aux[gvox(ln)];
        const StepRule d = rule_step<kMethod>(wp, a, cn, p0set, outthr);
        if (d.add) {
          const T h = a == 0 ? h0 : (a == 1 ? h1 : h2);
          const T t = h * (L.g[a][sln] + L.g[a][slv]);
          if constexpr (kSetup) S += (double)t;
          else c += (double)t;
          bool added = true;
          if (kCurlW<kMethod>) {
            curl += cn / 10000;
            cnt += 1.0;
            added = kMethod != kMaskWeighted || curl != 0.0;
          }
          if (added) add(sln, c, curl, cnt);
        }
        if (d.restart) {
          if constexpr (kSetup) S = 0.0;
          else if constexpr (kAuxLds) c = L.pl[sln];
          else c = p[gvox(ln)];
          if (!d.keep_w) {
            curl = 0.0;
            cnt = 0.0;
          }
        }
        lv = ln;
        slv = sln;
      }
      rem -= m;
      cw = wi == 0 ? sr.code[1] : sr.code[2];
    }
    if constexpr (kSetup) segS[sr.g] = S;
  }
  __syncthreads();
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), kk = oz + (l >> (2 * kBits));
    if (i >= g.X || j >= g.Y || kk >= g.Z) continue;
    const int64_t gv = i + j * sy + kk * sz;
    const int sl = lds_slot<kBits>(l);
    if constexpr (kSetup) {
      if constexpr (kMethod == kMaskUnit) wcnt[gv] = (unsigned int)L.acc[sl];
      else if constexpr (kCurlW<kMethod>) wsum[gv] = L.acc[sl];
    } else if constexpr (!kCurlW<kMethod>) {
      sumfx[gv] = L.acc[sl];
    } else {
      sumd[gv] = L.acc[sl];
    }
  }
}

template <typename T, bool kSetup>
hipError_t launch_tile_pass_masked(int method, int ntiles, const unsigned long long* tile_ptr,
                                   const SegRec* segs, GridDesc g, TileGrid tg,
                                   const WalkParams& wp, const void* field, const double* aux,
                                   const uint32_t* F, const double* wc, const double* wn,
                                   const double* cval, const double* p0, const double* p1, int k,
                                   const DevScalars* sc, long long* sumfx, double* sumd,
                                   double* segS, unsigned int* wcnt, double* wsum,
                                   hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  constexpr int kThreads = 1024;
#define OMNI_TM(M, B)                                                                           \
  hipLaunchKernelGGL((k_tile_masked<T, M, B, kThreads, kSetup>), dim3(ntiles), dim3(kThreads), 0, \
                     st, tile_ptr, segs, g, tg, wp, f, aux, F, wc, wn, cval, p0, p1, k, sc, sumfx, \
                     sumd, segS, wcnt, wsum)
  if (tg.bits == 3) {
    switch (method) {
      case kMaskPlain: OMNI_TM(kMaskPlain, 3); break;
      case kMaskUnit: OMNI_TM(kMaskUnit, 3); break;
      case kMaskSelect: OMNI_TM(kMaskSelect, 3); break;
      default: OMNI_TM(kMaskWeighted, 3); break;
    }
  } else {
    switch (method) {
      case kMaskPlain: OMNI_TM(kMaskPlain, 4); break;
      case kMaskUnit: OMNI_TM(kMaskUnit, 4); break;
      case kMaskSelect: OMNI_TM(kMaskSelect, 4); break;
      default: OMNI_TM(kMaskWeighted, 4); break;
    }
  }
#undef OMNI_TM
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mask_walk(int method, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                            GridDesc g, TileGrid tg, const WalkParams& wp, const void* field,
                            const double* aux, const unsigned long long* val_off, int ntiles,
                            const unsigned long long* tile_ptr, const SegRec* segs,
                            const DevScalars* sc, double* E, double* S, uint32_t* R, uint32_t* F,
                            double* wc, double* wn, unsigned int* wcnt, double* wsum,
                            uint8_t* src, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  const MaskSegs ms{E, R, F, wc, wn};
  const unsigned int grid = grid_for(nrays, 1u << 30);
#define OMNI_MW(M)                                                                              \
  hipLaunchKernelGGL((k_mask_walk<T, M>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g, tg, \
                     wp, f, aux, val_off, ms, src)
  switch (method) {
    case kMaskPlain: OMNI_MW(kMaskPlain); break;
    case kMaskUnit: OMNI_MW(kMaskUnit); break;
    case kMaskSelect: OMNI_MW(kMaskSelect); break;
    default: OMNI_MW(kMaskWeighted); break;
  }
#undef OMNI_MW
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // setup tile pass: in-tile sums after the last restart, and the weights
  return launch_tile_pass_masked<T, true>(method, ntiles, tile_ptr, segs, g, tg, wp, field, aux, F,
                                          wc, wn, nullptr, nullptr, nullptr, 1, sc, nullptr,
                                          nullptr, S, wcnt, wsum, st);
}
#define OMNI_MW_INST(T)                                                                        \
  template hipError_t launch_mask_walk<T>(                                                      \
      int, const RayRec*, int64_t, const DirConst*, GridDesc, TileGrid, const WalkParams&,       \
      const void*, const double*, const unsigned long long*, int, const unsigned long long*,    \
      const SegRec*, const DevScalars*, double*, double*, uint32_t*, uint32_t*, double*, double*, \
      unsigned int*, double*, uint8_t*, hipStream_t);
OMNI_MW_INST(float)
OMNI_MW_INST(double)
#undef OMNI_MW_INST

hipError_t launch_mask_carry(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, const double* E, const double* S,
                             const uint32_t* R, const uint32_t* F, const double* p0,
                             const double* p1, int k, const DevScalars* sc, double* cval,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_mask_carry, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, nrays, nseg,
                     val_off, E, S, R, F, p0, p1, k, sc, cval);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tile_masked(int method, int ntiles, const unsigned long long* tile_ptr,
                              const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                              const void* field, const double* aux, const uint32_t* F,
                              const double* wc, const double* wn, const double* cval,
                              const double* p0, const double* p1, int k, const DevScalars* sc,
                              long long* sumfx, double* sumd, hipStream_t st) {
  return launch_tile_pass_masked<T, false>(method, ntiles, tile_ptr, segs, g, tg, wp, field, aux,
                                           F, wc, wn, cval, p0, p1, k, sc, sumfx, sumd, nullptr,
                                           nullptr, nullptr, st);
}
template hipError_t launch_tile_masked<float>(int, int, const unsigned long long*, const SegRec*,
                                              GridDesc, TileGrid, const WalkParams&, const void*,
                                              const double*, const uint32_t*, const double*,
                                              const double*, const double*, const double*,
                                              const double*, int, const DevScalars*, long long*,
                                              double*, hipStream_t);
template hipError_t launch_tile_masked<double>(int, int, const unsigned long long*, const SegRec*,
                                               GridDesc, TileGrid, const WalkParams&, const void*,
                                               const double*, const uint32_t*, const double*,
                                               const double*, const double*, const double*,
                                               const double*, int, const DevScalars*, long long*,
                                               double*, hipStream_t);

}  // namespace omni3d
