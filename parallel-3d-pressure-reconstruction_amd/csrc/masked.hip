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
//   once per field:  k_rule_tile   one workgroup per tile replays the rule on
//                    the curl / mask (in LDS) for each segment from each
//                    state a ray can enter it in, and k_rule_scan (one lane
//                    per ray) follows the states the rays actually enter
//                    with: per segment (ray x tile, in its value slot), how
//                    its first visit is taken (with the term of the step
//                    into the tile, the only gradient it reads), the voxel
//                    of its last restart and the curl weights at its first
//                    visit (k_mask_walk, one lane per ray, writes the same
//                    records: OMNI3D_MASK_WALK=lane); then k_tile_masked<setup>
//                    replays the rule per tile from those records with the
//                    gradient in LDS: each segment's sum of added terms after
//                    its last restart, the weights (visit counts or curl
//                    weights) and the run starts (source flags);
//   per iteration:   k_mask_carry  one lane per ray chains its segments: the
//                    value a run enters each segment with;
//                    k_tile_masked one workgroup per tile stages the tile's
//                    gradient and curl / mask in LDS, replays the rule over
//                    the tile's segments and adds every visit's value in LDS
//                    (64-bit fixed point for unit weights, FP64 otherwise),
//                    then stores the tile's sums -- no global atomics.
// The first visit of a segment (the step into the tile) is decided by the
// replay, which reads the voxel outside the tile from memory; the tile pass
// walks the visits after it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "kernels.hpp"
#include "launch.hpp"
#include "rule.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {

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

// ---------------------------------------------------------------------------
// The rule replay on the tile engine (the records of k_mask_walk without a
// lane per ray).  The rule's state between two visits is (p0set, outthr) --
// N (no run), L (in a run), H (in a high / masked stretch) -- plus, for the
// curl weights, the running (curltmp, counttmp).  Per segment, a tile pass
// (k_rule_tile: the curl / mask in LDS) replays the segment from each entry
// state and stores what it does: the first visit's flags, the state it
// leaves in, its last restart, and how it changes the weights (the first
// visit adds cn / 1e4, resets them or keeps them; later visits add to sums
// kept since the first visit or since the segment's last reset).  A scan per
// ray (k_rule_scan) then follows the actual states along the ray's segments
// and writes the records k_mask_walk writes.  The source flags are set by the
// setup tile pass, which replays every restart.  The weights come out as
// (weights at the segment start) + (the segment's sums): the same terms as
// the lane walk's running sum, added in another order.

// per entry state: F flags (bits 0-5), exit state (6-7), the first visit's
// weight change (8-9: 0 keep, 1 add, 2 reset), a reset after the first
// visit (10), the visits added after the first or the last reset (11-16)
constexpr int kRuleStates = 3;
struct RuleSegs {
  double* E;      // the trapezoid term of the step into the tile, in place in MaskSegs::E
  double* x;      // curl / 1e4 at the first visit, in place in MaskSegs::wc (curl weights)
  uint32_t* f;    // [planes][nval] per entry state, bits as above
  uint32_t* r;    // [planes][nval] voxel of the last restart per entry state
  double* cs;     // [planes][nval] curl sum since the first visit or the last reset
  uint64_t nval;  // value slots per entry-state plane
};

__device__ __forceinline__ uint32_t rule_state(bool p0set, bool outthr) {
  return p0set ? (outthr ? 2u : 1u) : 0u;
}

// Select and plain-mask enter a segment in one of three states; the plain
// and Weighted walks are always in a run after the entry voxel.
template <int kMethod>
constexpr int kRulePlanes = (kMethod == kMaskUnit || kMethod == kMaskSelect) ? kRuleStates : 1;

bool mask_walk_lane() {
  const char* e = std::getenv("OMNI3D_MASK_WALK");
  return e && std::strcmp(e, "lane") == 0;
}

size_t mask_rule_bytes(int method, int64_t nval) {
  const int planes = (method == kMaskUnit || method == kMaskSelect) ? kRuleStates : 1;
  const bool curl = method == kMaskSelect || method == kMaskWeighted;
  return (size_t)planes * (size_t)std::max<int64_t>(nval, 1) * (curl ? 16 : 8);
}

template <typename T, int kMethod, int kBits>
__global__ void __launch_bounds__(1024)
k_rule_tile(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
            GridDesc g, TileGrid tg, WalkParams wp,
            const typename Vec4<T>::type* __restrict__ field, const double* __restrict__ aux,
            RuleSegs rs) {
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kVox = kTile * kTile * kTile;
  constexpr bool kAux = kMethod != kMaskPlain;
  __shared__ double s_aux[kAux ? kVox : 1];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  if constexpr (kAux) {
    for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
      const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), kk = oz + (l >> (2 * kBits));
      const bool in = i < g.X && j < g.Y && kk < g.Z;
      s_aux[lds_slot<kBits>(l)] = in ? aux[i + j * sy + kk * sz] : 0.0;
    }
    __syncthreads();
  }
  auto gvox = [&](int l) -> int64_t {
    return (int64_t)(ox + (l & kTileMask)) + (int64_t)(oy + ((l >> kBits) & kTileMask)) * sy +
           (int64_t)(oz + (l >> (2 * kBits))) * sz;
  };
  auto cval = [&](int l) -> double {
    if constexpr (kAux) return s_aux[lds_slot<kBits>(l)];
    else return 0.0;
  };
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const int lv0 = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool first_seg = (sr.info >> 21) & 1u;
    const int ae = (int)((sr.info >> 23) & 3u);  // axis of the step into the tile
    const int d0 = (sg & 1u) ? 1 : -1;
    const int d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    const double c_first = cval(lv0);
    // the step into the tile: its term (whether it joins the run depends on
    // the entry state; the scan decides)
    double E = 0.0;
    if (!first_seg) {
      const int se = ((sg >> ae) & 1u) ? 1 : -1;
      const int64_t vf = gvox(lv0);
      const int64_t vp = vf - (ae == 0 ? (int64_t)se : (ae == 1 ? se * sy : se * sz));
      E = (double)trapezoid<T>(wp, ae, se, component(field[vf], ae), component(field[vp], ae));
    }
    rs.E[sr.g] = E;
    if constexpr (kCurlW<kMethod>) rs.x[sr.g] = c_first / 10000;
    const int nplanes = first_seg ? 1 : kRulePlanes<kMethod>;
    for (int st = 0; st < nplanes; st++) {
      // entry state: N, L, H; the single state of the plain / Weighted walks is L
      const int es = kRulePlanes<kMethod> == 1 ? 1 : st;
      bool p0set = es >= 1, outthr = es == 2;
      uint32_t F, wop = 0u, rst = 0u;
      bool hasR = false;
      uint32_t R = 0u;
      if (first_seg) {
        // the ray's entry voxel (kernel.cu:2302 Select; every entry for the
        // plain and Weighted walks)
        p0set = outthr = false;
        if (rule_entry<kMethod>(wp, c_first)) {
          p0set = true;
          hasR = true;
          R = (uint32_t)gvox(lv0);
          F = kModeRestart | (kMethod == kMaskPlain ? kFlagAdd : 0u);
        } else {
          F = kModeNone;
        }
      } else {
        const StepRule d = rule_step<kMethod>(wp, ae, c_first, p0set, outthr);
        if (d.restart) {
          hasR = true;
          R = (uint32_t)gvox(lv0);
          if (!d.keep_w) wop = 2u;
        }
        if (d.add) wop = 1u;  // a step adds or restarts, never both
        F = (d.restart ? kModeRestart : (p0set ? kModeCont : kModeNone)) | (d.add ? kFlagAdd : 0u);
      }
      F |= (p0set ? kFlagP0 : 0u) | (outthr ? kFlagOut : 0u);
      // the visits after the first
      double csum = 0.0;
      uint32_t ccnt = 0u;
      int lv = lv0;
      uint32_t cw = sr.code[0];
      int rem = (int)n - 1;
#pragma unroll 1
      for (int wi = 0; wi < 3 && rem > 0; wi++) {
        const int m = rem < 16 ? rem : 16;
        for (int q = 0; q < m; q++) {
          const int a = (int)(cw & 3u);
          cw >>= 2;
          const int ln = lv + (a == 0 ? d0 : (a == 1 ? d1 : d2));
          const double cn = cval(ln);
          const StepRule d = rule_step<kMethod>(wp, a, cn, p0set, outthr);
          if (d.add && kCurlW<kMethod>) {
            csum += cn / 10000;
            ccnt += 1u;
          }
          if (d.restart) {
            hasR = true;
            R = (uint32_t)gvox(ln);
            if (!d.keep_w) {
              csum = 0.0;
              ccnt = 0u;
              rst = 1u;
            }
          }
          lv = ln;
        }
        rem -= m;
        cw = wi == 0 ? sr.code[1] : sr.code[2];
      }
      const uint64_t o = (uint64_t)st * rs.nval + sr.g;
      rs.f[o] = F | (hasR ? kFlagRestart : 0u) | (rule_state(p0set, outthr) << 6) |
                (wop << 8) | (rst << 10) | (ccnt << 11);
      rs.r[o] = R;
      if constexpr (kCurlW<kMethod>) rs.cs[o] = csum;
    }
  }
}

// Per ray: the segments' records from the actual entry states (one lane per
// ray, its slots 64 apart: every load and store is coalesced).
template <int kMethod>
__global__ void __launch_bounds__(256)
k_rule_scan(int64_t nrays, const unsigned int* __restrict__ nseg,
            const unsigned long long* __restrict__ val_off, RuleSegs rs, MaskSegs ms) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  unsigned long long gs = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int ns = nseg[r];
  int state = 0;
  double curl = 0.0, cnt = 0.0;
  for (unsigned int j = 0; j < ns; j++, gs += 64) {
    const int st = (j == 0 || kRulePlanes<kMethod> == 1) ? 0 : state;
    const uint64_t o = (uint64_t)st * rs.nval + gs;
    const uint32_t f = rs.f[o];
    const uint32_t wop = (f >> 8) & 3u;
    uint32_t F = f & 63u;
    if constexpr (kCurlW<kMethod>) {
      double c0 = curl, n0 = cnt;
      if (j > 0 && kMethod == kMaskWeighted) {
        // every step adds; it counts as added once the curl sum is nonzero
        c0 = curl + rs.x[gs];
        n0 = cnt + 1.0;
        if (c0 == 0.0) F &= ~kFlagAdd;
      } else if (wop == 1u) {
        c0 = curl + rs.x[gs];
        n0 = cnt + 1.0;
      } else if (wop == 2u) {
        c0 = 0.0;
        n0 = 0.0;
      }
      const bool reset = (f >> 10) & 1u;
      curl = (reset ? 0.0 : c0) + rs.cs[o];
      cnt = (reset ? 0.0 : n0) + (double)((f >> 11) & 63u);
      ms.wc[gs] = c0;
      ms.wn[gs] = n0;
    }
    // the step into the tile joins the run where it adds
    const bool step_adds =
        j > 0 && (kMethod == kMaskWeighted || kMethod == kMaskPlain || wop == 1u);
    if (!step_adds) ms.E[gs] = 0.0;
    ms.F[gs] = F;
    ms.R[gs] = rs.r[o];
    state = (int)((f >> 6) & 3u);
  }
}

// The terms of the steps into the tiles of a field, for records F kept from
// an earlier field with the same mask (plain-mask: the runs, and so F, depend
// on the mask alone): E = the step's trapezoid where the segment's first
// visit adds (not for a ray's first segment), else 0 -- what the rule replay
// writes for E.
template <typename T, int kBits>
__global__ void __launch_bounds__(256)
k_mask_entry_terms(const unsigned long long* __restrict__ tile_ptr,
                   const SegRec* __restrict__ segs, GridDesc g, TileGrid tg, WalkParams wp,
                   const typename Vec4<T>::type* __restrict__ field,
                   const uint32_t* __restrict__ F, double* __restrict__ E) {
  constexpr int kTileMask = (1 << kBits) - 1;
  const int tile = (int)blockIdx.y;
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + blockIdx.x * 256ull + threadIdx.x; jj < e;
       jj += (unsigned long long)gridDim.x * 256ull) {
    const SegRec sr = segs[jj];
    const bool first_seg = (sr.info >> 21) & 1u;
    double x = 0.0;
    if (!first_seg && (F[sr.g] & kFlagAdd)) {
      const int lv0 = (int)(sr.info & ((1u << kSegLvBits) - 1u));
      const unsigned int sg = (sr.info >> 18) & 7u;
      const int ae = (int)((sr.info >> 23) & 3u);
      const int se = ((sg >> ae) & 1u) ? 1 : -1;
      const int64_t vf = (int64_t)(ox + (lv0 & kTileMask)) +
                         (int64_t)(oy + ((lv0 >> kBits) & kTileMask)) * sy +
                         (int64_t)(oz + (lv0 >> (2 * kBits))) * sz;
      const int64_t vp = vf - (ae == 0 ? (int64_t)se : (ae == 1 ? se * sy : se * sz));
      x = (double)trapezoid<T>(wp, ae, se, component(field[vf], ae), component(field[vp], ae));
    }
    E[sr.g] = x;
  }
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
              unsigned int* __restrict__ wcnt, double* __restrict__ wsum,
              uint8_t* __restrict__ src) {
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
    // the run starts (source flags): every voxel a run (re)starts at
    if constexpr (kSetup)
      if (mode == kModeRestart) src[gvox(lv)] = 1;
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
          if constexpr (kSetup) {
            S = 0.0;
            src[gvox(ln)] = 1;
          } else if constexpr (kAuxLds) c = L.pl[sln];
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
                                   uint8_t* src, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  constexpr int kThreads = 1024;
#define OMNI_TM(M, B)                                                                           \
  hipLaunchKernelGGL((k_tile_masked<T, M, B, kThreads, kSetup>), dim3(ntiles), dim3(kThreads), 0, \
                     st, tile_ptr, segs, g, tg, wp, f, aux, F, wc, wn, cval, p0, p1, k, sc, sumfx, \
                     sumd, segS, wcnt, wsum, src)
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

// The rule replay on the tile engine alone (k_rule_tile + k_rule_scan): the
// records E, R, F (and the curl weights wc, wn) of every segment.
template <typename T>
hipError_t launch_rule_replay(int method, int64_t nrays, const unsigned int* nseg,
                              const unsigned long long* val_off, int64_t nval, int ntiles,
                              const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                              TileGrid tg, const WalkParams& wp, const void* field,
                              const double* aux, double* E, uint32_t* R, uint32_t* F, double* wc,
                              double* wn, void* rule, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  const MaskSegs ms{E, R, F, wc, wn};
  const int planes = (method == kMaskUnit || method == kMaskSelect) ? kRuleStates : 1;
  const bool curl = method == kMaskSelect || method == kMaskWeighted;
  const uint64_t nv = (uint64_t)std::max<int64_t>(nval, 1);
  char* base = static_cast<char*>(rule);
  RuleSegs rs;
  rs.E = E;
  rs.x = wc;
  rs.cs = curl ? reinterpret_cast<double*>(base) : nullptr;
  if (curl) base += sizeof(double) * planes * nv;
  rs.f = reinterpret_cast<uint32_t*>(base);
  rs.r = rs.f + planes * nv;
  rs.nval = nv;
#define OMNI_RT(M, B)                                                                         \
  hipLaunchKernelGGL((k_rule_tile<T, M, B>), dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, g, \
                     tg, wp, f, aux, rs)
#define OMNI_RT_M(B)                                  \
  switch (method) {                                   \
    case kMaskPlain: OMNI_RT(kMaskPlain, B); break;   \
    case kMaskUnit: OMNI_RT(kMaskUnit, B); break;     \
    case kMaskSelect: OMNI_RT(kMaskSelect, B); break; \
    default: OMNI_RT(kMaskWeighted, B); break;        \
  }
  if (tg.bits == 3) {
    OMNI_RT_M(3)
  } else {
    OMNI_RT_M(4)
  }
#undef OMNI_RT_M
#undef OMNI_RT
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned int sgrid = grid_for(nrays, 1u << 30);
#define OMNI_RS(M) \
  hipLaunchKernelGGL((k_rule_scan<M>), dim3(sgrid), dim3(256), 0, st, nrays, nseg, val_off, rs, ms)
  switch (method) {
    case kMaskPlain: OMNI_RS(kMaskPlain); break;
    case kMaskUnit: OMNI_RS(kMaskUnit); break;
    case kMaskSelect: OMNI_RS(kMaskSelect); break;
    default: OMNI_RS(kMaskWeighted); break;
  }
#undef OMNI_RS
  return hipGetLastError();
}
template hipError_t launch_rule_replay<float>(int, int64_t, const unsigned int*,
                                              const unsigned long long*, int64_t, int,
                                              const unsigned long long*, const SegRec*, GridDesc,
                                              TileGrid, const WalkParams&, const void*,
                                              const double*, double*, uint32_t*, uint32_t*,
                                              double*, double*, void*, hipStream_t);
template hipError_t launch_rule_replay<double>(int, int64_t, const unsigned int*,
                                               const unsigned long long*, int64_t, int,
                                               const unsigned long long*, const SegRec*, GridDesc,
                                               TileGrid, const WalkParams&, const void*,
                                               const double*, double*, uint32_t*, uint32_t*,
                                               double*, double*, void*, hipStream_t);

template <typename T>
hipError_t launch_mask_walk(int method, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                            GridDesc g, TileGrid tg, const WalkParams& wp, const void* field,
                            const double* aux, const unsigned int* nseg,
                            const unsigned long long* val_off, int64_t nval, int ntiles,
                            const unsigned long long* tile_ptr, const SegRec* segs,
                            const DevScalars* sc, double* E, double* S, uint32_t* R, uint32_t* F,
                            double* wc, double* wn, unsigned int* wcnt, double* wsum,
                            uint8_t* src, void* rule, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  const MaskSegs ms{E, R, F, wc, wn};
  hipError_t e;
  if (rule) {
    // the rule replay on the tile engine: per segment and entry state, then
    // per ray
    e = launch_rule_replay<T>(method, nrays, nseg, val_off, nval, ntiles, tile_ptr, segs, g, tg,
                              wp, field, aux, E, R, F, wc, wn, rule, st);
    if (e != hipSuccess) return e;
  } else {
    // one lane per ray (the records without the tile engine)
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
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // setup tile pass: in-tile sums after the last restart, the weights and
  // the run starts
  return launch_tile_pass_masked<T, true>(method, ntiles, tile_ptr, segs, g, tg, wp, field, aux, F,
                                          wc, wn, nullptr, nullptr, nullptr, 1, sc, nullptr,
                                          nullptr, S, wcnt, wsum, src, st);
}
#define OMNI_MW_INST(T)                                                                        \
  template hipError_t launch_mask_walk<T>(                                                      \
      int, const RayRec*, int64_t, const DirConst*, GridDesc, TileGrid, const WalkParams&,       \
      const void*, const double*, const unsigned int*, const unsigned long long*, int64_t, int,  \
      const unsigned long long*, const SegRec*, const DevScalars*, double*, double*, uint32_t*,  \
      uint32_t*, double*, double*, unsigned int*, double*, uint8_t*, void*, hipStream_t);
OMNI_MW_INST(float)
OMNI_MW_INST(double)
#undef OMNI_MW_INST

template <typename T>
hipError_t launch_mask_reuse(int method, int ntiles, const unsigned long long* tile_ptr,
                             const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                             const void* field, const double* aux, const DevScalars* sc,
                             double* E, double* S, const uint32_t* F, unsigned int* wcnt,
                             uint8_t* src, hipStream_t st) {
  if (method != kMaskUnit) return hipErrorInvalidValue;
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  // a few workgroups per tile: the tiles' segment lists average hundreds
  const dim3 grid(4, (unsigned int)ntiles);
  if (tg.bits == 3)
    hipLaunchKernelGGL((k_mask_entry_terms<T, 3>), grid, dim3(256), 0, st, tile_ptr, segs, g, tg,
                       wp, f, F, E);
  else
    hipLaunchKernelGGL((k_mask_entry_terms<T, 4>), grid, dim3(256), 0, st, tile_ptr, segs, g, tg,
                       wp, f, F, E);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_tile_pass_masked<T, true>(method, ntiles, tile_ptr, segs, g, tg, wp, field, aux, F,
                                          nullptr, nullptr, nullptr, nullptr, nullptr, 1, sc,
                                          nullptr, nullptr, S, wcnt, nullptr, src, st);
}
template hipError_t launch_mask_reuse<float>(int, int, const unsigned long long*, const SegRec*,
                                             GridDesc, TileGrid, const WalkParams&, const void*,
                                             const double*, const DevScalars*, double*, double*,
                                             const uint32_t*, unsigned int*, uint8_t*,
                                             hipStream_t);
template hipError_t launch_mask_reuse<double>(int, int, const unsigned long long*, const SegRec*,
                                              GridDesc, TileGrid, const WalkParams&, const void*,
                                              const double*, const DevScalars*, double*, double*,
                                              const uint32_t*, unsigned int*, uint8_t*,
                                              hipStream_t);

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
                                           nullptr, nullptr, nullptr, st);
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

// ---------------------------------------------------------------------------
// The masked methods' affine schedule.  A run's value at a visit is
// p[start] + (the terms since the start), and the weights, the runs and their
// start voxels depend on the curl / mask alone.  So with S = the run-start
// voxels (the source flags of k_mask_walk) every pass maps p on S affinely:
//   p_new[s] = (b[s] + sum over the visits of s: w p[start of the run]) / W[s]
// where W[s] > 0 (else p[s] stays), b = the pass's sums from p = 0 and W the
// weights (kernel.cu:2385-2392 Select, 2071-2079 Weighted; unit weights for
// plain-mask).  Per field the visits landing on S are listed once as rows:
// k_mask_runsrc gives each segment the start voxel of the run it continues,
// then a tile pass replays the rule and counts, then fills, the entries --
// as CSR rows for the unit weights (k_tile_rows; the merged-word iteration
// of the plain method takes them from there) or in tile layout for the curl
// weights (k_tile_entries; iterated by k_mtile_accum + k_mtile_finish).  The
// iteration runs on S alone, and one full pass from p_S^{k*-1} gives every
// voxel (k_masked_final).

constexpr uint32_t kNoRun = 0xffffffffu;

// Per segment (value slot): the start voxel of the run a segment in the
// "continue" mode carries -- the last restart before it along its ray.
__global__ void __launch_bounds__(256)
k_mask_runsrc(int64_t nrays, const unsigned int* __restrict__ nseg,
              const unsigned long long* __restrict__ val_off, const uint32_t* __restrict__ R,
              const uint32_t* __restrict__ F, uint32_t* __restrict__ runsrc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  unsigned long long gs = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  uint32_t cur = kNoRun;
  for (unsigned int j = 0; j < n; j++, gs += 64) {
    const uint32_t f = F[gs];
    runsrc[gs] = (f & 3u) == kModeCont ? cur : kNoRun;
    if (f & kFlagRestart) cur = R[gs];
  }
}

// The rows: per tile, the rule replayed over the tile's segments (the
// curl / mask and the source ids staged in LDS); every visit that adds and
// lands on a source voxel v is an entry (row sid[v], source sid[start],
// weight).  Every visit of a voxel is walked by the tile that owns it, so the
// rows of the tile's voxels are counted and filled in LDS: kFill == false
// stores each row's entry count into cnt; kFill takes a slot per entry from
// an LDS cursor and writes the entry's source at rowptr[row] + slot.  Entries
// of a row are in no fixed order: the iteration's fixed-point sums do not
// depend on it.  Launched for unit weights (plain-mask).
template <int kMethod, int kBits, bool kFill>
__global__ void __launch_bounds__(1024)
k_tile_rows(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
            GridDesc g, TileGrid tg, WalkParams wp, const double* __restrict__ aux,
            const uint32_t* __restrict__ F, const uint32_t* __restrict__ runsrc,
            const int* __restrict__ sid, unsigned int* __restrict__ cnt,
            const unsigned long long* __restrict__ rowptr, unsigned int* __restrict__ esrc) {
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kVox = kTile * kTile * kTile;
  __shared__ double s_aux[kVox];
  __shared__ int s_sid[kVox];
  __shared__ unsigned int s_n[kVox];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), kk = oz + (l >> (2 * kBits));
    const bool in = i < g.X && j < g.Y && kk < g.Z;
    const int64_t gv = i + j * sy + kk * sz;
    const int sl = lds_slot<kBits>(l);
    s_aux[sl] = in ? aux[gv] : 0.0;
    s_sid[sl] = in ? sid[gv] : -1;
    s_n[sl] = 0u;
  }
  __syncthreads();
  auto gvox = [&](int l) -> int64_t {
    return (int64_t)(ox + (l & kTileMask)) + (int64_t)(oy + ((l >> kBits) & kTileMask)) * sy +
           (int64_t)(oz + (l >> (2 * kBits))) * sz;
  };
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const uint32_t f = F[sr.g];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const int d0 = (sg & 1u) ? 1 : -1;
    const int d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    bool p0set = (f & kFlagP0) != 0, outthr = (f & kFlagOut) != 0;
    const uint32_t mode = f & 3u;
    uint32_t t = mode == kModeCont ? runsrc[sr.g]
                                   : (mode == kModeRestart ? (uint32_t)gvox(lv) : kNoRun);
    // the source id of the run's start voxel (every run start is a source:
    // k_mask_walk flags them all)
    int src = t == kNoRun ? -1 : sid[t];
    auto emit = [&](int l) {
      const int sl = lds_slot<kBits>(l);
      const int row = s_sid[sl];
      if (row < 0 || src < 0) return;
      const unsigned int slot = atomicAdd(&s_n[sl], 1u);
      if constexpr (kFill) esrc[rowptr[row] + slot] = (unsigned int)src;
    };
    if (f & kFlagAdd) emit(lv);
    uint32_t cw = sr.code[0];
    int rem = (int)n - 1;
#pragma unroll 1
    for (int wi = 0; wi < 3 && rem > 0; wi++) {
      const int m = rem < 16 ? rem : 16;
      for (int q = 0; q < m; q++) {
        const int a = (int)(cw & 3u);
        cw >>= 2;
        const int ln = lv + (a == 0 ? d0 : (a == 1 ? d1 : d2));
        const double cn = s_aux[lds_slot<kBits>(ln)];
        const StepRule d = rule_step<kMethod>(wp, a, cn, p0set, outthr);
        if (d.add) emit(ln);
        if (d.restart) src = s_sid[lds_slot<kBits>(ln)];
        lv = ln;
      }
      rem -= m;
      cw = wi == 0 ? sr.code[1] : sr.code[2];
    }
  }
  if constexpr (!kFill) {
    __syncthreads();
    for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
      const int sl = lds_slot<kBits>(l);
      const int row = s_sid[sl];
      if (row >= 0) cnt[row] = s_n[sl];
    }
  }
}

// The rows in tile layout (curl weights, and unit weights past the merged-
// word format): the entries whose row is a voxel of tile t sit in
// [tbase[t], tbase[t + 1]) in the order the tile pass emits them (source id,
// the row's index among the tile's rows, weight), so the pass writes them
// coalesced (a wavefront takes consecutive slots from one LDS counter) and
// the iteration streams a tile's entries with neighbouring rays' sources side
// by side.  kFill == false stores each tile's entry count into tcnt.
template <int kMethod, int kBits, bool kFill>
__global__ void __launch_bounds__(1024)
k_tile_entries(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
               GridDesc g, TileGrid tg, WalkParams wp, const double* __restrict__ aux,
               const uint32_t* __restrict__ F, const double* __restrict__ wc,
               const double* __restrict__ wn, const uint32_t* __restrict__ runsrc,
               const int* __restrict__ sid, const int* __restrict__ tfirst,
               unsigned int* __restrict__ tcnt, const unsigned long long* __restrict__ tbase,
               unsigned int* __restrict__ esrc, unsigned short* __restrict__ erow,
               double* __restrict__ ew) {
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kVox = kTile * kTile * kTile;
  __shared__ double s_aux[kVox];
  __shared__ int s_sid[kVox];
  __shared__ unsigned int s_total;
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  if (threadIdx.x == 0) s_total = 0u;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), kk = oz + (l >> (2 * kBits));
    const bool in = i < g.X && j < g.Y && kk < g.Z;
    const int64_t gv = i + j * sy + kk * sz;
    const int sl = lds_slot<kBits>(l);
    s_aux[sl] = in ? aux[gv] : 0.0;
    s_sid[sl] = in ? sid[gv] : -1;
  }
  __syncthreads();
  const int first = kFill ? tfirst[tile] : 0;
  const unsigned long long base = kFill ? tbase[tile] : 0ull;
  const int lane = threadIdx.x & 63;
  auto gvox = [&](int l) -> int64_t {
    return (int64_t)(ox + (l & kTileMask)) + (int64_t)(oy + ((l >> kBits) & kTileMask)) * sy +
           (int64_t)(oz + (l >> (2 * kBits))) * sz;
  };
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const uint32_t f = F[sr.g];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const int d0 = (sg & 1u) ? 1 : -1;
    const int d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    bool p0set = (f & kFlagP0) != 0, outthr = (f & kFlagOut) != 0;
    const uint32_t mode = f & 3u;
    const uint32_t t = mode == kModeCont ? runsrc[sr.g]
                                         : (mode == kModeRestart ? (uint32_t)gvox(lv) : kNoRun);
    int src = t == kNoRun ? -1 : sid[t];
    double curl = 0.0, num = 0.0;
    if (kCurlW<kMethod>) {
      curl = wc[sr.g];
      num = wn[sr.g];
    }
    auto emit = [&](int l) {
      const int row = s_sid[lds_slot<kBits>(l)];
      const bool want = row >= 0 && src >= 0;
      // the lanes emitting now take consecutive slots
      const unsigned long long act = __ballot(want);
      if (!want) return;
      const int leader = __ffsll((unsigned long long)act) - 1;
      const unsigned int rank = (unsigned int)__popcll(act & ((1ull << lane) - 1ull));
      unsigned int b0 = 0u;
      if (lane == leader) b0 = atomicAdd(&s_total, (unsigned int)__popcll(act));
      b0 = __shfl(b0, leader);
      if constexpr (kFill) {
        const unsigned long long at = base + b0 + rank;
        esrc[at] = (unsigned int)src;
        erow[at] = (unsigned short)(row - first);
        if constexpr (kCurlW<kMethod>) ew[at] = curl_weight(curl, num);
      }
    };
    if (f & kFlagAdd) emit(lv);
    uint32_t cw = sr.code[0];
    int rem = (int)n - 1;
#pragma unroll 1
    for (int wi = 0; wi < 3 && rem > 0; wi++) {
      const int m = rem < 16 ? rem : 16;
      for (int q = 0; q < m; q++) {
        const int a = (int)(cw & 3u);
        cw >>= 2;
        const int ln = lv + (a == 0 ? d0 : (a == 1 ? d1 : d2));
        const double cn = s_aux[lds_slot<kBits>(ln)];
        const StepRule d = rule_step<kMethod>(wp, a, cn, p0set, outthr);
        if (d.add) {
          bool added = true;
          if (kCurlW<kMethod>) {
            curl += cn / 10000;
            num += 1.0;
            added = kMethod != kMaskWeighted || curl != 0.0;
          }
          if (added) emit(ln);
        }
        if (d.restart) {
          src = s_sid[lds_slot<kBits>(ln)];
          if (!d.keep_w) {
            curl = 0.0;
            num = 0.0;
          }
        }
        lv = ln;
      }
      rem -= m;
      cw = wi == 0 ? sr.code[1] : sr.code[2];
    }
  }
  if constexpr (!kFill) {
    __syncthreads();
    if (threadIdx.x == 0) tcnt[tile] = s_total;
  }
}

// One iteration on S in tile layout, part 1: a workgroup per chunk of one
// tile's entries adds w p[src] (64-bit fixed point, scale 2^(51 - e),
// 2^e > max w * max |p_S^{k-1}|) into LDS sums of the tile's rows, then adds
// each nonzero sum into the row's global fixed-point sum (a row's total over
// all its entries stays below 2^63: fewer than 4096 entries per row).
constexpr int kChunkThreads = 1024;
constexpr int kChunkUnroll = 8;
constexpr int kTileRowsMax = 4096;

__device__ __forceinline__ bool masked_scale(const DevScalars* sc, int k,
                                             const unsigned long long* wmax_bits, double* scale,
                                             double* iscale, double* B) {
  const double b = __longlong_as_double((long long)sc->itmax_bits[(k - 1) & 1]) *
                   __longlong_as_double((long long)*wmax_bits);
  const bool finite = b <= DBL_MAX;
  int ex = 0;
  (void)frexp(finite ? b : 1.0, &ex);
  ex = ex < -960 ? -960 : ex;
  *scale = ldexp(1.0, 51 - ex);
  *iscale = ldexp(1.0, ex - 51);
  *B = b;
  return finite;
}

template <bool kUnit>
__global__ void __launch_bounds__(kChunkThreads)
k_mtile_accum(int k, const MaskChunk* __restrict__ chunks, const int* __restrict__ tfirst,
              const unsigned int* __restrict__ esrc, const unsigned short* __restrict__ erow,
              const double* __restrict__ ew, const double* __restrict__ p0,
              const double* __restrict__ p1, const DevScalars* __restrict__ sc,
              const unsigned long long* __restrict__ wmax_bits,
              unsigned long long* __restrict__ rowsum) {
  if (sc->converged) return;
  __shared__ unsigned long long s_acc[kTileRowsMax];
  const MaskChunk c = chunks[xcd_block()];
  const int r0 = tfirst[c.tile], nr = tfirst[c.tile + 1] - r0;
  for (int i = threadIdx.x; i < nr; i += kChunkThreads) s_acc[i] = 0ull;
  __syncthreads();
  const double* __restrict__ pold = ((k - 1) & 1) ? p1 : p0;
  double scale, iscale, B;
  (void)masked_scale(sc, k, wmax_bits, &scale, &iscale, &B);
  // kChunkUnroll entries per thread and round: their loads, then their
  // gathers, then their adds (the integer sums do not depend on the order)
  constexpr unsigned long long kSpan = (unsigned long long)kChunkThreads * kChunkUnroll;
  for (unsigned long long e0 = c.begin + threadIdx.x; e0 < c.end; e0 += kSpan) {
    unsigned int src[kChunkUnroll];
    unsigned short row[kChunkUnroll];
    double w[kChunkUnroll], x[kChunkUnroll];
#pragma unroll
    for (int q = 0; q < kChunkUnroll; q++) {
      const unsigned long long e = e0 + (unsigned long long)q * kChunkThreads;
      src[q] = 0u;
      row[q] = 0;
      w[q] = 0.0;
      if (e < c.end) {
        src[q] = esrc[e];
        row[q] = erow[e];
        w[q] = kUnit ? 1.0 : ew[e];
      }
    }
#pragma unroll
    for (int q = 0; q < kChunkUnroll; q++) {
      const unsigned long long e = e0 + (unsigned long long)q * kChunkThreads;
      x[q] = e < c.end ? pold[src[q]] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kChunkUnroll; q++) {
      const unsigned long long e = e0 + (unsigned long long)q * kChunkThreads;
      if (e < c.end) atomicAdd(&s_acc[row[q]], (unsigned long long)to_fixed(w[q] * x[q], scale));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += kChunkThreads) {
    const unsigned long long a = s_acc[i];
    if (a) atomicAdd(&rowsum[r0 + i], a);
  }
}

// Part 2, per row: p_new = (b + sum [+ the starting p on the first pass,
// kernel.cu:8813]) / W where W > 0 (else p stays), the sum cleared for the
// next iteration; the relative-change partials per workgroup and max |p_new|
// (slot k & 1); k_iter_finish (two slots) applies the stop rule.
__global__ void __launch_bounds__(256)
k_mtile_finish(int k, int nS, const double* __restrict__ bvec, const double* __restrict__ wvec,
               const int* __restrict__ slist, const double* __restrict__ pinit, int seed,
               double* __restrict__ p0, double* __restrict__ p1,
               unsigned long long* __restrict__ rowsum, DevScalars* __restrict__ sc,
               const unsigned long long* __restrict__ wmax_bits, double* __restrict__ partials) {
  if (sc->converged) return;
  __shared__ double s_d2[4], s_n2[4];
  __shared__ unsigned long long s_pm[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double* __restrict__ pold = ((k - 1) & 1) ? p1 : p0;
  double* __restrict__ pnew = (k & 1) ? p1 : p0;
  double scale, iscale, B;
  const bool finite = masked_scale(sc, k, wmax_bits, &scale, &iscale, &B);
  double d2 = 0.0, n2 = 0.0;
  unsigned long long pm = 0ull;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < nS; r += gridDim.x * 256) {
    const unsigned long long a = rowsum[r];
    rowsum[r] = 0ull;
    double tot = bvec[r] + (finite ? (double)(long long)a * iscale : B - B);
    if (k == 1 && seed && pinit) tot += pinit[slist[r]];
    const double po = pold[r];
    const double w = wvec[r];
    const double pn = w > 0.0 ? tot / w : po;
    pnew[r] = pn;
    const double df = pn - po;
    d2 += df * df;
    n2 += pn * pn;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(pn));
    pm = bits > pm ? bits : pm;
  }
  for (int off = 32; off > 0; off >>= 1) {
    d2 += __shfl_xor(d2, off);
    n2 += __shfl_xor(n2, off);
    const unsigned long long o = __shfl_xor(pm, off);
    pm = o > pm ? o : pm;
  }
  if (lane == 0) {
    s_d2[wave] = d2;
    s_n2[wave] = n2;
    s_pm[wave] = pm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    unsigned long long m = 0ull;
    for (int q = 0; q < 4; q++) {
      a += s_d2[q];
      c += s_n2[q];
      m = s_pm[q] > m ? s_pm[q] : m;
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = c;
    atomicMax(&sc->itmax_bits[k & 1], m);
  }
}

// Per row of S: the constant b (the sums of the pass from p = 0) and the
// weight W, from the full-volume sums and weights.
__global__ void __launch_bounds__(256)
k_masked_rowvec(int nS, const int* __restrict__ slist, int fixed, const long long* __restrict__ sumfx,
                const unsigned int* __restrict__ wcnt, const double* __restrict__ sumd,
                const double* __restrict__ wsum, const DevScalars* __restrict__ sc,
                double* __restrict__ bvec, double* __restrict__ wvec) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nS) return;
  const int v = slist[r];
  bvec[r] = fixed ? (double)sumfx[v] * sc->inv_scale : sumd[v];
  wvec[r] = fixed ? (double)wcnt[v] : wsum[v];
}

// Compact ids of the flagged voxels (the run starts) in tile order (the
// voxels of tile 0 in in-tile order, then tile 1, ...): the rows a tile pass
// builds are then one contiguous id range, tile t's [tfirst[t],
// tfirst[t + 1]), and consecutive rows are nearby voxels.  Over many workgroups: per 1024 positions the flag count, an
// exclusive scan of the counts in one workgroup, then each block's ids from
// its base and ballots.  Positions past the grid (partial tiles) are unflagged.
constexpr int kIdBlock = 1024;

__device__ __forceinline__ int64_t tile_order_voxel(int64_t q, GridDesc g, TileGrid tg) {
  const int b = tg.bits;
  const int64_t tv = 1ll << (3 * b);
  const int64_t t = q >> (3 * b);
  const int l = (int)(q & (tv - 1));
  const int tx = (int)(t % tg.ntx), ty = (int)((t / tg.ntx) % tg.nty), tz = (int)(t / ((int64_t)tg.ntx * tg.nty));
  const int m = (1 << b) - 1;
  const int i = (tx << b) + (l & m), j = (ty << b) + ((l >> b) & m), k = (tz << b) + (l >> (2 * b));
  if (i >= g.X || j >= g.Y || k >= g.Z) return -1;
  return (int64_t)i + (int64_t)j * g.X + (int64_t)k * g.X * g.Y;
}

__global__ void __launch_bounds__(kIdBlock)
k_srcid_partial(const uint8_t* __restrict__ flag, int64_t nq, GridDesc g, TileGrid tg,
                unsigned int* __restrict__ bcnt) {
  __shared__ unsigned int s_w[kIdBlock / 64];
  const int64_t q = (int64_t)blockIdx.x * kIdBlock + threadIdx.x;
  const int64_t v = q < nq ? tile_order_voxel(q, g, tg) : -1;
  const unsigned long long m = __ballot(v >= 0 && flag[v]);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = (unsigned int)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = 0;
    for (int w = 0; w < kIdBlock / 64; w++) t += s_w[w];
    bcnt[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kIdBlock)
k_srcid_scan(unsigned int* __restrict__ bcnt, int64_t nb, int* __restrict__ nS,
             int* __restrict__ tfirst, int ntiles) {
  __shared__ unsigned int s_w[kIdBlock / 64];
  __shared__ unsigned int s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += kIdBlock) {
    const int64_t i = b0 + threadIdx.x;
    const unsigned int v = i < nb ? bcnt[i] : 0u;
    unsigned int x = v;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    unsigned int wb = 0, tot = 0;
    for (int w = 0; w < kIdBlock / 64; w++) {
      if (w < wave) wb += s_w[w];
      tot += s_w[w];
    }
    const unsigned int base = s_base;
    if (i < nb) bcnt[i] = base + wb + x - v;
    __syncthreads();
    if (threadIdx.x == 0) s_base = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *nS = (int)s_base;
    tfirst[ntiles] = (int)s_base;
  }
}

__global__ void __launch_bounds__(kIdBlock)
k_srcid_final(const uint8_t* __restrict__ flag, int64_t nq, GridDesc g, TileGrid tg,
              const unsigned int* __restrict__ bbase, int* __restrict__ sid,
              int* __restrict__ slist, int* __restrict__ tfirst) {
  __shared__ unsigned int s_w[kIdBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t q = (int64_t)blockIdx.x * kIdBlock + threadIdx.x;
  const int64_t v = q < nq ? tile_order_voxel(q, g, tg) : -1;
  const bool f = v >= 0 && flag[v];
  const unsigned long long m = __ballot(f);
  if (lane == 0) s_w[wave] = (unsigned int)__popcll(m);
  __syncthreads();
  unsigned int wb = 0;
  for (int w = 0; w < wave; w++) wb += s_w[w];
  const unsigned int before = bbase[blockIdx.x] + wb + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
  // a tile's first position: the first id of its rows
  if (q < nq && (q & ((1ll << (3 * tg.bits)) - 1)) == 0) tfirst[q >> (3 * tg.bits)] = (int)before;
  if (v >= 0) {
    if (f) {
      const int id = (int)before;
      sid[v] = id;
      slist[id] = (int)v;
    } else {
      sid[v] = -1;
    }
  }
}

// Unit weights: the rows' constant in fixed point (the grouped iteration's
// form, bs * inv_scale) and the rows' entries as 64-bit sort keys (source in
// the low bits) for the merged-word layout.
__global__ void __launch_bounds__(256)
k_masked_bs(int nS, const int* __restrict__ slist, const long long* __restrict__ sumfx,
            long long* __restrict__ bs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nS) bs[r] = sumfx[slist[r]];
}

__global__ void __launch_bounds__(256)
k_widen_keys(int64_t n, const unsigned int* __restrict__ esrc, unsigned long long* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = esrc[i];
}

// p_S^{k*-1} (the buffer iteration k* read) into the full-volume pressure
// the final pass reads at the run starts.
__global__ void __launch_bounds__(256)
k_masked_scatter(int nS, const int* __restrict__ slist, const DevScalars* __restrict__ sc,
                 const double* __restrict__ p0, const double* __restrict__ p1,
                 double* __restrict__ pfull) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nS) return;
  const double* __restrict__ p = ((sc->kstar - 1) & 1) ? p1 : p0;
  pfull[slist[r]] = p[r];
}

// The scalars of the final pass: the iteration's scale, the stop flag
// cleared (the pass kernels return at once on a stopped iteration).
__global__ void k_masked_final_scalars(const DevScalars* __restrict__ sc,
                                       DevScalars* __restrict__ out) {
  *out = *sc;
  out->converged = 0;
}

// Averaging of the final pass (devidecountInner, kernel.cu:6746-6759):
// where the weight is positive p = sum / weight (plus the starting p when the
// output is the first pass and pn started as p), elsewhere p keeps
// p^{k*-1} (pprev).  Outputs in FP64 or FP32, the weights beside.
__global__ void __launch_bounds__(256)
k_masked_final(int64_t V, int fixed, const long long* __restrict__ sumfx,
               const unsigned int* __restrict__ wcnt, const double* __restrict__ sumd,
               const double* __restrict__ wsum, const double* __restrict__ pinit, int seed,
               const DevScalars* __restrict__ sc, const double* __restrict__ pprev,
               double* __restrict__ pd, float* __restrict__ pf, double* __restrict__ wd,
               float* __restrict__ wf) {
  const double inv = sc->inv_scale;
  const bool add_seed = seed && pinit && sc->kstar == 1;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const double w = fixed ? (double)wcnt[v] : wsum[v];
    double p = pprev[v];
    if (w > 0.0) {
      double tot = fixed ? (double)sumfx[v] * inv : sumd[v];
      if (add_seed) tot += pinit[v];
      p = tot / w;
    }
    if (pd) pd[v] = p;
    if (pf) pf[v] = (float)p;
    if (wd) wd[v] = w;
    if (wf) wf[v] = (float)w;
  }
}

hipError_t launch_mask_runsrc(int64_t nrays, const unsigned int* nseg,
                              const unsigned long long* val_off, const uint32_t* R,
                              const uint32_t* F, uint32_t* runsrc, hipStream_t st) {
  hipLaunchKernelGGL(k_mask_runsrc, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, nrays, nseg,
                     val_off, R, F, runsrc);
  return hipGetLastError();
}

hipError_t launch_tile_rows(bool fill, int method, int ntiles, const unsigned long long* tile_ptr,
                            const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const double* aux, const uint32_t* F, const uint32_t* runsrc,
                            const int* sid, unsigned int* cnt, const unsigned long long* rowptr,
                            unsigned int* esrc, hipStream_t st) {
#define OMNI_TR(M, B, FL)                                                                       \
  hipLaunchKernelGGL((k_tile_rows<M, B, FL>), dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, g, \
                     tg, wp, aux, F, runsrc, sid, cnt, rowptr, esrc)
#define OMNI_TR_B(M, FL)                  \
  do {                                    \
    if (tg.bits == 3) OMNI_TR(M, 3, FL);  \
    else OMNI_TR(M, 4, FL);               \
  } while (0)
  // (unit weights only: the curl-weighted rows use the tile layout)
  if (method != kMaskUnit) return hipErrorInvalidValue;
  if (fill) OMNI_TR_B(kMaskUnit, true);
  else OMNI_TR_B(kMaskUnit, false);
#undef OMNI_TR_B
#undef OMNI_TR
  return hipGetLastError();
}

static int64_t tile_positions(const TileGrid& tg) {
  return (int64_t)tg.ntx * tg.nty * tg.ntz << (3 * tg.bits);
}

hipError_t launch_source_ids_tiled(const uint8_t* flag, GridDesc g, TileGrid tg,
                                  unsigned int* bcnt, int* sid, int* slist, int* nS,
                                  int* tfirst, hipStream_t st) {
  const int64_t nq = tile_positions(tg);
  const int64_t nb = (nq + kIdBlock - 1) / kIdBlock;
  const int ntiles = tg.ntx * tg.nty * tg.ntz;
  hipLaunchKernelGGL(k_srcid_partial, dim3((unsigned int)nb), dim3(kIdBlock), 0, st, flag, nq, g,
                     tg, bcnt);
  hipLaunchKernelGGL(k_srcid_scan, dim3(1), dim3(kIdBlock), 0, st, bcnt, nb, nS, tfirst, ntiles);
  hipLaunchKernelGGL(k_srcid_final, dim3((unsigned int)nb), dim3(kIdBlock), 0, st, flag, nq, g,
                     tg, bcnt, sid, slist, tfirst);
  return hipGetLastError();
}

int64_t source_ids_tiled_blocks(const TileGrid& tg) {
  return (tile_positions(tg) + kIdBlock - 1) / kIdBlock;
}

hipError_t launch_masked_bs(int nS, const int* slist, const long long* sumfx, long long* bs,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_masked_bs, dim3(grid_for(nS, 1u << 30)), dim3(256), 0, st, nS, slist, sumfx,
                     bs);
  return hipGetLastError();
}

hipError_t launch_widen_keys(int64_t n, const unsigned int* esrc, unsigned long long* key,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_widen_keys, dim3(grid_for(n, 1u << 30)), dim3(256), 0, st, n, esrc, key);
  return hipGetLastError();
}

hipError_t launch_tile_entries(bool fill, int method, int ntiles,
                               const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                               TileGrid tg, const WalkParams& wp, const double* aux,
                               const uint32_t* F, const double* wc, const double* wn,
                               const uint32_t* runsrc, const int* sid, const int* tfirst,
                               unsigned int* tcnt, const unsigned long long* tbase,
                               unsigned int* esrc, unsigned short* erow, double* ew,
                               hipStream_t st) {
#define OMNI_TE(M, B, FL)                                                                       \
  hipLaunchKernelGGL((k_tile_entries<M, B, FL>), dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, \
                     g, tg, wp, aux, F, wc, wn, runsrc, sid, tfirst, tcnt, tbase, esrc, erow, ew)
#define OMNI_TE_B(M, FL)                  \
  do {                                    \
    if (tg.bits == 3) OMNI_TE(M, 3, FL);  \
    else OMNI_TE(M, 4, FL);               \
  } while (0)
#define OMNI_TE_M(FL)                                      \
  do {                                                     \
    switch (method) {                                      \
      case kMaskUnit: OMNI_TE_B(kMaskUnit, FL); break;     \
      case kMaskSelect: OMNI_TE_B(kMaskSelect, FL); break; \
      default: OMNI_TE_B(kMaskWeighted, FL); break;        \
    }                                                      \
  } while (0)
  if (fill) OMNI_TE_M(true);
  else OMNI_TE_M(false);
#undef OMNI_TE_M
#undef OMNI_TE_B
#undef OMNI_TE
  return hipGetLastError();
}

hipError_t launch_mtile_accum(int k, int nchunks, const MaskChunk* chunks, const int* tfirst,
                              const unsigned int* esrc, const unsigned short* erow,
                              const double* ew, const double* p0, const double* p1,
                              const DevScalars* sc, const unsigned long long* wmax_bits,
                              unsigned long long* rowsum, hipStream_t st) {
  if (ew)
    hipLaunchKernelGGL(k_mtile_accum<false>, dim3(nchunks), dim3(kChunkThreads), 0, st, k, chunks,
                       tfirst, esrc, erow, ew, p0, p1, sc, wmax_bits, rowsum);
  else
    hipLaunchKernelGGL(k_mtile_accum<true>, dim3(nchunks), dim3(kChunkThreads), 0, st, k, chunks,
                       tfirst, esrc, erow, ew, p0, p1, sc, wmax_bits, rowsum);
  return hipGetLastError();
}

unsigned int mtile_finish_grid(int nS) {
  const int b = (nS + 255) / 256;
  return (unsigned int)(b < 1 ? 1 : (b > 1024 ? 1024 : b));
}

hipError_t launch_mtile_finish(int k, int nS, const double* bvec, const double* wvec,
                               const int* slist, const double* pinit, int seed, double* p0,
                               double* p1, unsigned long long* rowsum, DevScalars* sc,
                               const unsigned long long* wmax_bits, double* partials,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_mtile_finish, dim3(mtile_finish_grid(nS)), dim3(256), 0, st, k, nS, bvec,
                     wvec, slist, pinit, seed, p0, p1, rowsum, sc, wmax_bits, partials);
  return hipGetLastError();
}

hipError_t launch_masked_rowvec(int nS, const int* slist, int fixed, const long long* sumfx,
                                const unsigned int* wcnt, const double* sumd, const double* wsum,
                                const DevScalars* sc, double* bvec, double* wvec,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_masked_rowvec, dim3(grid_for(nS, 1u << 30)), dim3(256), 0, st, nS, slist,
                     fixed, sumfx, wcnt, sumd, wsum, sc, bvec, wvec);
  return hipGetLastError();
}

hipError_t launch_masked_scatter(int nS, const int* slist, const DevScalars* sc, const double* p0,
                                 const double* p1, double* pfull, hipStream_t st) {
  hipLaunchKernelGGL(k_masked_scatter, dim3(grid_for(nS, 1u << 30)), dim3(256), 0, st, nS, slist,
                     sc, p0, p1, pfull);
  return hipGetLastError();
}

hipError_t launch_masked_final_scalars(const DevScalars* sc, DevScalars* out, hipStream_t st) {
  hipLaunchKernelGGL(k_masked_final_scalars, dim3(1), dim3(1), 0, st, sc, out);
  return hipGetLastError();
}

hipError_t launch_masked_final(int64_t V, int fixed, const long long* sumfx,
                               const unsigned int* wcnt, const double* sumd, const double* wsum,
                               const double* pinit, int seed, const DevScalars* sc,
                               const double* pprev, double* pd, float* pf, double* wd, float* wf,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_masked_final, dim3(grid_for(V)), dim3(256), 0, st, V, fixed, sumfx, wcnt,
                     sumd, wsum, pinit, seed, sc, pprev, pd, pf, wd, wf);
  return hipGetLastError();
}

}  // namespace omni3d
