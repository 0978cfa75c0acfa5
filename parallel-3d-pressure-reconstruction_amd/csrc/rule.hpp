// The path-break rule of the masked methods (plain-mask, Select, Weighted)
// and plain, step by step, as the tile engine replays it (masked.hip) and
// the plain-mask adjoint transposes it (adjoint.hip); the flags of a
// segment's first visit that the per-field rule replay stores.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

}  // namespace omni3d
