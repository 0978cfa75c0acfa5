// Tile engine: the traversals of a reconstruction reorganised by space.
//
// The plan compiles every ray's staircase once (geometry only) into 2-bit
// step codes and cuts it into segments, one per 16^3 tile the ray crosses.
// A workgroup then owns one tile: it stages the tile's gradient in LDS, walks
// every segment of every direction that crosses the tile, and accumulates
// into 64-bit fixed-point LDS counters, which it finally stores to its own
// voxels -- no global atomics.  The prefix sum a segment continues from (its
// "carry") comes from a first pass of per-segment sums and a per-ray scan.
//
//   plan:   k_tile_count -> scans -> k_tile_fill            (once per grid)
//   solve:  k_tile_pass<kLocalSum> -> k_carry_scan ->
//           k_tile_pass<kSourceSums> (surface tiles) -> source iteration ->
//           k_tile_pass<kFull> -> averaging
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.hpp"
#include "launch.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {

// ---------------------------------------------------------------------------
// Plan kernels.  Visitor-style geometry walks (the same step rule as every
// other walk) that count steps and segments, then write codes and records.

// Segments of a tile are grouped by length, longest first, so that the lanes
// of a wavefront walk segments of similar length: slot list index
// tile * kLenBuckets + bucket.
__device__ __forceinline__ int len_slot(int tile, unsigned int nvis, int lshift) {
  const unsigned int b = nvis >> lshift;
  return tile * kLenBuckets + (kLenBuckets - 1) - (int)(b < kLenBuckets ? b : kLenBuckets - 1);
}

__global__ void __launch_bounds__(256)
k_tile_count(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, TileGrid tg, unsigned int* __restrict__ nseg,
             unsigned int* __restrict__ tile_cnt, SrcRows rows) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  Walker w;
  w.init(rays[r], dirs, g);
  // the entry visit reads the ray's own source: the row's diagonal
  if (rows.rowcnt) atomicAdd(&rows.diag[rows.sid[w.v]], 1u);
  int tile = tile_of(w.c0, w.c1, w.c2, tg);
  unsigned int segs = 1, nvis = 1;
  while (!w.done()) {
    int s;
    if (w.advance(&s) < 0) break;
    if (rows.rowcnt && w.on_surface()) {
      const int q = rows.sid[w.v];
      if (q >= 0) atomicAdd(&rows.rowcnt[q], 1u);
    }
    const int t = tile_of(w.c0, w.c1, w.c2, tg);
    if (t != tile) {
      atomicAdd(&tile_cnt[len_slot(tile, nvis, tg.lshift)], 1u);
      tile = t;
      segs++;
      nvis = 0;
    }
    nvis++;
  }
  atomicAdd(&tile_cnt[len_slot(tile, nvis, tg.lshift)], 1u);
  nseg[r] = segs;
}

__global__ void __launch_bounds__(256)
k_tile_fill(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, TileGrid tg, const int* __restrict__ sid,
            const unsigned long long* __restrict__ val_off,
            const unsigned long long* __restrict__ tile_ptr, unsigned int* __restrict__ cursor,
            SegRec* __restrict__ segs, SrcRows rows) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const RayRec rr = rays[r];
  Walker w;
  w.init(rr, dirs, g);
  const unsigned long long rkey =
      ((unsigned long long)(rr.dir_len & 0xffffu) << 32) | (unsigned long long)(uint32_t)sid[w.v];
  int ci[3], co[3];
  unpack_voxel(rr.in, ci);
  unpack_voxel(rr.out, co);
  const uint32_t signs = (co[0] > ci[0] ? 1u : 0u) | (co[1] > ci[1] ? 2u : 0u) |
                         (co[2] > ci[2] ? 4u : 0u);
  const uint32_t src = (uint32_t)sid[w.v];
  // value slot of the k-th segment: blocked transposed layout (k * 64 + lane
  // within the ray's group of 64), so that a wavefront's rays touch
  // consecutive slots in the carry scan and, within a tile, neighbouring
  // parallel rays at the same ordinal read neighbouring carries
  unsigned long long gseg = val_off[r >> 6] + (unsigned long long)(r & 63);
  int tile = tile_of(w.c0, w.c1, w.c2, tg);
  SegRec sr;
  uint32_t nvis = 1, ncode = 0, first = 1;
  const int tb = tg.bits, tmask = (1 << tb) - 1;
  auto start = [&](int i, int j, int k) {
    sr.g = (uint32_t)gseg;
    sr.src = src;
    const uint32_t lv = (uint32_t)((i & tmask) | ((j & tmask) << tb) | ((k & tmask) << (2 * tb)));
    sr.info = lv | (signs << 18) | (first << 21);
    sr.code[0] = sr.code[1] = sr.code[2] = 0;
  };
  auto emit = [&](bool last) {
    sr.info |= (nvis << kSegLvBits) | ((last ? 1u : 0u) << 22);
    const int ls = len_slot(tile, nvis, tg.lshift);
    const unsigned int slot = atomicAdd(&cursor[ls], 1u);
    segs[tile_ptr[ls] + slot] = sr;
    gseg += 64;
  };
  auto put_code = [&](int a) {
    const uint32_t sh = 2 * (ncode & 15);
    const uint32_t word = ncode >> 4;
    if (word == 0) sr.code[0] |= (uint32_t)a << sh;
    else if (word == 1) sr.code[1] |= (uint32_t)a << sh;
    else sr.code[2] |= (uint32_t)a << sh;
    ncode++;
  };
  start(w.c0, w.c1, w.c2);
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    if (rows.key && w.on_surface()) {
      const int q = rows.sid[w.v];
      if (q >= 0) rows.key[rows.rowptr[q] + atomicAdd(&rows.cursor[q], 1u)] = rkey;
    }
    const int t = tile_of(w.c0, w.c1, w.c2, tg);
    // the step out of a tile ends the code list of the segment it leaves
    put_code(a);
    if (t != tile) {
      emit(false);
      tile = t;
      first = 0;
      nvis = 1;
      ncode = 0;
      start(w.c0, w.c1, w.c2);
      sr.info |= (uint32_t)a << 23;  // the axis of the step into the tile
    } else {
      nvis++;
    }
  }
  emit(true);
}

// ---------------------------------------------------------------------------
// Tile passes.

// LDS of a tile pass: the tile's gradient; the fixed-point counters (full
// and source-sum passes); the source ids of the tile's voxels (source-sum
// pass).  Exact sizes matter: two FP32 16^3 full-pass workgroups fill the
// 160 KB of a CU.
// The segment-sum pass keeps a one-voxel halo ((edge + 2)^3 voxels, x
// fastest) instead: the step into the tile then reads the voxel outside it
// from LDS.
template <typename T, int kVox, int kHaloVox, int kMode>
struct TileLds {
  T g[3][kVox];
  long long acc[kVox];
};
template <typename T, int kVox, int kHaloVox>
struct TileLds<T, kVox, kHaloVox, kLocalSum> {
  T g[3][kHaloVox];
};
template <typename T, int kVox, int kHaloVox>
struct TileLds<T, kVox, kHaloVox, kSourceSums> {
  T g[3][kVox];
  long long acc[kVox];
  int sidl[kVox];
};

template <typename T, int kMode, int kBits, int kThreads>
__global__ void __launch_bounds__(kThreads)
k_tile_pass(const int* __restrict__ tile_list, const unsigned long long* __restrict__ tile_ptr,
            const SegRec* __restrict__ segs, GridDesc g, TileGrid tg, WalkParams wp,
            const typename Vec4<T>::type* __restrict__ field, const int* __restrict__ sid,
            const unsigned int* __restrict__ vcnt,
            const double* __restrict__ pbuf0, const double* __restrict__ pbuf1,
            const DevScalars* __restrict__ sc, T* __restrict__ segval,
            long long* __restrict__ out) {
  using V4 = typename Vec4<T>::type;
  constexpr int kTileBits = kBits;
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kTileVox = kTile * kTile * kTile;
  constexpr bool kHalo = kMode == kLocalSum;
  constexpr int kW = kTile + 2;  // halo edge
  constexpr int kHaloVox = kW * kW * kW;
  constexpr int kPlane = kHalo ? kHaloVox : kTileVox;
  __shared__ TileLds<T, kTileVox, kHaloVox, kMode> L;
  // contiguous tile ranges per XCD: neighbouring tiles share an L2
  const int tb = (int)xcd_block();
  const int tile = tile_list ? tile_list[tb] : tb;
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kTileBits, oy = ty << kTileBits, oz = tz << kTileBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  // stage the tile's gradient (zero outside the grid) and clear the counters
  if constexpr (kHalo) {
    for (int l = threadIdx.x; l < kHaloVox; l += blockDim.x) {
      const int i = ox + l % kW - 1, j = oy + (l / kW) % kW - 1, k = oz + l / (kW * kW) - 1;
      V4 v;
      const bool in = i >= 0 && j >= 0 && k >= 0 && i < g.X && j < g.Y && k < g.Z;
      if (in) v = field[i + j * sy + k * sz];
      else v.x = v.y = v.z = v.w = 0;
      L.g[0][l] = v.x;
      L.g[1][l] = v.y;
      L.g[2][l] = v.z;
    }
  }
  for (int l = threadIdx.x; l < (kHalo ? 0 : kTileVox); l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kTileBits) & kTileMask),
              k = oz + (l >> (2 * kTileBits));
    V4 v;
    const bool in = i < g.X && j < g.Y && k < g.Z;
    if (in) v = field[i + j * sy + k * sz];
    else v.x = v.y = v.z = v.w = 0;
    if constexpr (!kHalo) {
      const int sl = lds_slot<kBits>(l);
      L.g[0][sl] = v.x;
      L.g[1][sl] = v.y;
      L.g[2][sl] = v.z;
      L.acc[sl] = 0;
      // source voxels (all on the box surface) by their compact id
      if constexpr (kMode == kSourceSums) L.sidl[sl] = in ? sid[i + j * sy + k * sz] : -1;
    }
  }
  __syncthreads();
  const double scale = sc->scale;
  const double* __restrict__ ps = ((sc->kstar - 1) & 1) ? pbuf1 : pbuf0;
  // kSourceSums walks the compact list of the tile's segments that reach a
  // source voxel and kLocalSum the tile's segments in value-slot order (one
  // range per tile: coalesced slot stores); kFull the length groups
  constexpr int kStride = kMode == kFull ? kLenBuckets : 1;
  const unsigned long long b = tile_ptr[tile * kStride], e = tile_ptr[(tile + 1) * kStride];
  // software pipeline over a thread's segments: the record two segments
  // ahead and the values the next segment starts from (its carry, entry
  // term and source pressure) are in flight while this segment is walked
  struct Pre {
    double pin = 0.0, cv = 0.0;
  };
  const unsigned int bd = blockDim.x;
  auto fetch = [&](const SegRec& r, unsigned long long jj) {
    Pre q;
    const bool fs = (r.info >> 21) & 1u;
    if constexpr (kMode == kFull) q.pin = ps[r.src];
    if constexpr (kMode != kLocalSum) {
      if (!fs) q.cv = (double)segval[r.g];
    }
    (void)jj;
    return q;
  };
  unsigned long long j = b + threadIdx.x;
  SegRec r1, r2;
  Pre p1;
  if (j < e) r1 = segs[j];
  if (j + bd < e) r2 = segs[j + bd];
  if (j < e) p1 = fetch(r1, j);
  for (; j < e; j += bd) {
    const SegRec sr = r1;
    const Pre pr = p1;
    r1 = r2;
    if (j + 2ull * bd < e) r2 = segs[j + 2ull * bd];
    if (j + bd < e) p1 = fetch(r1, j + bd);
    // carry: the ray's sum before this segment plus the step into the tile
    // (kLocalSum computes that step, the only read outside the tile, and
    // stores it per segment in tile order)
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool first = (sr.info >> 21) & 1u;
    // the segment's step signs are fixed: per axis, the LDS step and the
    // trapezoid factor of a step along it
    const bool up0 = sg & 1u, up1 = (sg >> 1) & 1u, up2 = (sg >> 2) & 1u;
    constexpr int kSy = kHalo ? kW : kTile;  // LDS strides of the walked index
    const int d0 = up0 ? 1 : -1;
    const int d1 = up1 ? kSy : -kSy;
    const int d2 = up2 ? kSy * kSy : -kSy * kSy;
    // the walked index: the halo index (kLocalSum) or the in-tile index,
    // whose LDS slot is swizzled
    if constexpr (kHalo)
      lv = (lv & kTileMask) + 1 + (((lv >> kTileBits) & kTileMask) + 1) * kW +
           ((lv >> (2 * kTileBits)) + 1) * kW * kW;
    auto slot = [&](int l) {
      if constexpr (kHalo) return l;
      else return lds_slot<kBits>(l);
    };
    const T h0 = step_factor<T>(wp, 0, up0);
    const T h1 = step_factor<T>(wp, 1, up1);
    const T h2 = step_factor<T>(wp, 2, up2);
    // in-tile sums accumulate in FP64: kLocalSum the segment's sum; the
    // other passes (base + sum so far) * scale + 1.5 * 2^52, whose low
    // mantissa bits are the 64-bit fixed-point value (|value| < 2^51, the
    // scale is a power of two: each step adds its term rounded to a unit)
    double loc = 0.0, accs = 6755399441055744.0;
    auto add_fixed = [&](int l) {
      if constexpr (kMode == kFull) {
        // the biased bits; the flush removes visits x the bias
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[l]),
                  (unsigned long long)__double_as_longlong(accs));
      } else if constexpr (kMode == kSourceSums) {
        const long long f = __double_as_longlong(accs) - 0x4338000000000000LL;
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[l]), (unsigned long long)f);
      }
    };
    auto on_source = [&](int l) {
      if constexpr (kMode == kSourceSums) return L.sidl[l] >= 0;
      else return false;
    };
    uint32_t cw = sr.code[0];
    int inw = 16;  // codes left in the current word
    if (first) {
      // the entry visit: self-add p_src (pint = 0)
      if constexpr (kMode == kFull) {
        accs = pr.pin * scale + 6755399441055744.0;
        add_fixed(slot(lv));
      }
    } else if constexpr (kMode != kLocalSum) {
      // the carry already holds the step into the tile (it ends the previous
      // segment's sum)
      accs = (kMode == kFull ? pr.pin + pr.cv : pr.cv) * scale + 6755399441055744.0;
      const int slv = slot(lv);
      if (kMode == kFull || on_source(slv)) add_fixed(slv);
    }
    // every later visit adds base + (in-tile sum so far)
    int rem = (int)n - 1;
    int slv = slot(lv);
#pragma unroll 1
    for (int wi = 0; wi < 3 && rem > 0; wi++) {
      const int m = rem < inw ? rem : inw;
      for (int q = 0; q < m; q++) {
        const int a = (int)(cw & 3u);
        cw >>= 2;
        const bool a0 = a == 0, a1 = a == 1;
        const int ln = lv + (a0 ? d0 : (a1 ? d1 : d2));
        const T h = a0 ? h0 : (a1 ? h1 : h2);
        const T* __restrict__ gp = &L.g[0][0] + a * kPlane;
        const int sln = slot(ln);
        const T t = h * (gp[sln] + gp[slv]);
        lv = ln;
        slv = sln;
        if constexpr (kMode == kLocalSum) {
          loc += (double)t;
        } else {
          accs = fma((double)t, scale, accs);
          if (kMode == kFull || on_source(slv)) add_fixed(slv);
        }
      }
      rem -= m;
      cw = wi == 0 ? sr.code[1] : sr.code[2];
      inw = 16;
    }
    if constexpr (kMode == kLocalSum) {
      if (!((sr.info >> 22) & 1u)) {
        // the step out of the tile (the next voxel lies in the halo) ends
        // the segment: its code follows the n - 1 in-tile codes
        const unsigned int q = n - 1;
        const uint32_t wq = q < 16 ? sr.code[0] : (q < 32 ? sr.code[1] : sr.code[2]);
        const int a = (int)((wq >> (2 * (q & 15))) & 3u);
        const int sgn = ((sg >> a) & 1u) ? 1 : -1;
        const int dn = a == 0 ? d0 : (a == 1 ? d1 : d2);
        loc += (double)trapezoid<T>(wp, a, sgn, L.g[a][lv + dn], L.g[a][lv]);
      }
      segval[sr.g] = (T)loc;
    }
  }
  if (kMode == kLocalSum) return;
  __syncthreads();
  for (int l = threadIdx.x; l < kTileVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kTileBits) & kTileMask),
              k = oz + (l >> (2 * kTileBits));
    if (i >= g.X || j >= g.Y || k >= g.Z) continue;
    const int64_t v = i + j * sy + k * sz;
    const int sl = lds_slot<kBits>(l);
    // (every visit of v is walked by this tile: cnt[v] biased adds)
    if constexpr (kMode == kFull)
      out[v] = (long long)((unsigned long long)L.acc[sl] -
                           (unsigned long long)vcnt[v] * 0x4338000000000000ull);
    if constexpr (kMode == kSourceSums) {
      const int s = L.sidl[sl];
      if (s >= 0) out[s] = L.acc[sl];
    }
  }
}

// Visit counts of every voxel (plan build): a workgroup per tile counts its
// segments' visits in LDS and stores the tile's counts (the entry visit of a
// ray is its first segment's first visit).
template <int kBits>
__global__ void __launch_bounds__(1024)
k_tile_counts(const unsigned long long* __restrict__ tile_ptr1, const SegRec* __restrict__ segs,
              GridDesc g, TileGrid tg, unsigned int* __restrict__ cnt) {
  constexpr int kTile = 1 << kBits, kTileMask = kTile - 1, kVox = kTile * kTile * kTile;
  __shared__ unsigned int C[kVox];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) C[l] = 0u;
  __syncthreads();
  const unsigned long long b = tile_ptr1[tile], e = tile_ptr1[tile + 1];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const int d0 = (sg & 1u) ? 1 : -1, d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    atomicAdd(&C[lv], 1u);
    for (unsigned int q = 1; q < n; q++) {
      const unsigned int c = q - 1;
      const uint32_t w = c < 16 ? sr.code[0] : (c < 32 ? sr.code[1] : sr.code[2]);
      const int a = (int)((w >> (2 * (c & 15))) & 3u);
      lv += a == 0 ? d0 : (a == 1 ? d1 : d2);
      atomicAdd(&C[lv], 1u);
    }
  }
  __syncthreads();
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    if (i < g.X && j < g.Y && k < g.Z) cnt[i + j * sy + k * sz] = C[l];
  }
}

hipError_t launch_tile_counts(int ntiles, const unsigned long long* tile_ptr1, const SegRec* segs,
                              GridDesc g, TileGrid tg, unsigned int* cnt, hipStream_t st) {
  if (tg.bits == 3)
    hipLaunchKernelGGL(k_tile_counts<3>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs, g, tg,
                       cnt);
  else
    hipLaunchKernelGGL(k_tile_counts<4>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs, g, tg,
                       cnt);
  return hipGetLastError();
}

// Compact list of the segments of the source-sum pass (plan build).  One
// workgroup per surface tile replays the step codes of the tile's segments
// and keeps those whose visits reach a source voxel after the ray's entry
// (for a later segment the first visit counts too), with the visit count
// cut after the last such visit and the position in the full list (where
// the pass reads the segment's step term) in the source field.  pass 0
// counts per tile, pass 1 writes at sptr[tile] (order inside a tile
// immaterial: the pass sums exactly).
template <int kBits>
__global__ void __launch_bounds__(256)
k_src_segs(int pass, const int* __restrict__ tile_list, const unsigned long long* __restrict__ tile_ptr,
           const SegRec* __restrict__ segs, GridDesc g, TileGrid tg, const int* __restrict__ sid,
           unsigned int* __restrict__ cnt, const unsigned long long* __restrict__ sptr,
           SegRec* __restrict__ out) {
  constexpr int kTile = 1 << kBits;
  constexpr int kTileMask = kTile - 1;
  constexpr int kTileVox = kTile * kTile * kTile;
  __shared__ unsigned char s_src[kTileVox];
  __shared__ unsigned int s_n;
  const int tile = tile_list[blockIdx.x];
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kTileVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    s_src[l] = (i < g.X && j < g.Y && k < g.Z && sid[i + j * sy + k * sz] >= 0) ? 1 : 0;
  }
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long j = b + threadIdx.x; j < e; j += blockDim.x) {
    const SegRec sr = segs[j];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool first = (sr.info >> 21) & 1u;
    uint32_t c0 = sr.code[0], c1 = sr.code[1], c2 = sr.code[2];
    auto next_code = [&]() {
      const int a = (int)(c0 & 3u);
      c0 = (c0 >> 2) | (c1 << 30);
      c1 = (c1 >> 2) | (c2 << 30);
      c2 >>= 2;
      return a;
    };
    int last = -1;
    if (!first && s_src[lv]) last = 0;
    for (unsigned int vis = 1; vis < n; vis++) {
      const int a = next_code();
      const int up = (sg >> a) & 1u;
      const int d = a == 0 ? 1 : (a == 1 ? kTile : kTile * kTile);
      lv += up ? d : -d;
      if (s_src[lv]) last = (int)vis;
    }
    if (last < 0) continue;
    const unsigned int pos = atomicAdd(&s_n, 1u);
    if (pass == 1) {
      SegRec r = sr;
      r.info = (sr.info & ~(63u << kSegLvBits)) | ((unsigned int)(last + 1) << kSegLvBits);
      r.src = (uint32_t)j;
      out[sptr[tile] + pos] = r;
    }
  }
  __syncthreads();
  if (pass == 0 && threadIdx.x == 0) cnt[tile] = s_n;
}

hipError_t launch_src_segs(int pass, int nsurf, const int* tile_list,
                           const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                           TileGrid tg, const int* sid, unsigned int* cnt,
                           const unsigned long long* sptr, SegRec* out, hipStream_t st) {
  if (nsurf <= 0) return hipSuccess;
  if (tg.bits == 3)
    hipLaunchKernelGGL(k_src_segs<3>, dim3(nsurf), dim3(256), 0, st, pass, tile_list, tile_ptr, segs,
                       g, tg, sid, cnt, sptr, out);
  else
    hipLaunchKernelGGL(k_src_segs<4>, dim3(nsurf), dim3(256), 0, st, pass, tile_list, tile_ptr, segs,
                       g, tg, sid, cnt, sptr, out);
  return hipGetLastError();
}

// Iteration matrix layout (plan build).  A chunk is 64 consecutive rows
// (neighbouring surface voxels under the blocked source numbering).  Its
// entries are visited in slots (d, k): the k-th entry of direction d of each
// row, for k below the largest count of d among the chunk's rows.  A slot
// stores a 64-bit mask of the rows that have such an entry, and the entries
// themselves are packed without padding in slot order, each row's at its
// rank among the set bits.  The lanes of one slot then gather sources of one
// direction for neighbouring targets -- neighbouring sources -- and a row's
// entries are summed in (direction, source) order.
constexpr int kSellDirCap = 4096;

__device__ __forceinline__ void sell_count_dirs(const unsigned long long* __restrict__ rowptr,
                                                const unsigned long long* __restrict__ key,
                                                int nS, int row, unsigned int* K) {
  if (row >= nS) return;
  const unsigned long long b = rowptr[row], e = rowptr[row + 1];
  unsigned int run = 0, d = 0xffffffffu;
  for (unsigned long long j = b; j < e; j++) {
    const unsigned int dj = (unsigned int)(key[j] >> 32);
    if (dj != d) {
      if (run) atomicMax(&K[d], run);
      d = dj;
      run = 0;
    }
    run++;
  }
  if (run) atomicMax(&K[d], run);
}

__global__ void __launch_bounds__(64)
k_sell_layout(int nS, int D, const unsigned long long* __restrict__ rowptr,
              const unsigned long long* __restrict__ key, unsigned int* __restrict__ nslot) {
  __shared__ unsigned int K[kSellDirCap];
  const int chunk = blockIdx.x, lane = threadIdx.x, row = chunk * 64 + lane;
  for (int d = lane; d < D; d += 64) K[d] = 0;
  __syncthreads();
  sell_count_dirs(rowptr, key, nS, row, K);
  __syncthreads();
  unsigned int tot = 0;
  for (int d = lane; d < D; d += 64) tot += K[d];
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off);
  if (lane == 0) nslot[chunk] = tot;
}

__global__ void __launch_bounds__(64)
k_sell_fill(int nS, int D, const unsigned long long* __restrict__ rowptr,
            const unsigned long long* __restrict__ key, const unsigned long long* __restrict__ moff,
            unsigned long long* __restrict__ masks, unsigned int* __restrict__ cell_eb,
            unsigned short* __restrict__ sdir, int* __restrict__ ent) {
  const int chunk = blockIdx.x, lane = threadIdx.x, row = chunk * 64 + lane;
  const unsigned long long rb = row < nS ? rowptr[row] : 0ull;
  const unsigned long long re = row < nS ? rowptr[row + 1] : 0ull;
  unsigned long long mp = moff[chunk];
  unsigned long long eb = rowptr[chunk * 64];  // the chunk's entries start where its rows do
  unsigned long long j = rb;
  for (int d = 0; d < D; d++) {
    // entries of direction d in this lane's row (rows are sorted by (d, source))
    unsigned int c = 0;
    while (j + c < re && (unsigned int)(key[j + c] >> 32) == (unsigned int)d) c++;
    unsigned int kmax = c;
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned int o = __shfl_xor(kmax, off);
      kmax = o > kmax ? o : kmax;
    }
    for (unsigned int k = 0; k < kmax; k++) {
      const bool has = c > k;
      const unsigned long long m = __ballot(has);
      if (lane == 0) {
        masks[mp] = m;
        cell_eb[mp] = (unsigned int)eb;
        sdir[mp] = (unsigned short)d;
      }
      mp++;
      if (has) {
        const unsigned int rank = __popcll(m & ((1ull << lane) - 1ull));
        ent[eb + rank] = (int)(unsigned int)(key[j + k] & 0xffffffffull);
      }
      eb += __popcll(m);
    }
    j += c;
  }
}

// Value slots of the segments: rays in groups of 64 (one wavefront of the
// carry scan); group q holds 64 * (its largest segment count) slots, the k-th
// segment of its ray l at val_off[q] + 64 k + l.
__global__ void __launch_bounds__(64)
k_seg_groups(int64_t nrays, const unsigned int* __restrict__ nseg,
             unsigned int* __restrict__ gcnt) {
  const int64_t r = (int64_t)blockIdx.x * 64 + threadIdx.x;
  unsigned int m = r < nrays ? nseg[r] : 0u;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if (threadIdx.x == 0) gcnt[blockIdx.x] = 64u * m;
}

// Slot order inside every (tile, length bucket) list: the passes then read
// the carries of neighbouring lanes from neighbouring value slots, and the
// plan no longer depends on the order the fill kernel's atomics gave.
__global__ void __launch_bounds__(256)
k_seg_keys(const SegRec* __restrict__ segs, int64_t n, unsigned int* __restrict__ key,
           unsigned int* __restrict__ idx) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    key[i] = segs[i].g;
    idx[i] = (unsigned int)i;
  }
}

__global__ void __launch_bounds__(256)
k_seg_gather(const SegRec* __restrict__ in, const unsigned int* __restrict__ idx, int64_t n,
             SegRec* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[idx[i]];
}

__global__ void __launch_bounds__(256)
k_u64_to_u32(const unsigned long long* __restrict__ in, int64_t n, unsigned int* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (unsigned int)in[i];
}

// Per-ray exclusive scan of the segment sums into carries, in place; the
// lanes of a wavefront are the 64 rays of one group, so every step reads 64
// consecutive slots.  The running sum is FP64; the sums and carries are
// stored in the compute precision C (FP32 ray sums: each carry is the FP64
// prefix rounded once to FP32).
template <typename C>
__global__ void __launch_bounds__(256)
k_carry_scan(int64_t nrays, const unsigned int* __restrict__ nseg,
             const unsigned long long* __restrict__ val_off, C* __restrict__ segval) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const unsigned long long b = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  double run = 0.0;
  unsigned int k = 0;
  // sixteen loads in flight per lane; the sum order is the segment order
  for (; k + 16 <= n; k += 16) {
    double x[16];
#pragma unroll
    for (int u = 0; u < 16; u++) x[u] = (double)segval[b + 64ull * (k + u)];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      segval[b + 64ull * (k + u)] = (C)run;
      run = run + x[u];
    }
  }
  for (; k + 4 <= n; k += 4) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) x[u] = (double)segval[b + 64ull * (k + u)];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      segval[b + 64ull * (k + u)] = (C)run;
      run = run + x[u];
    }
  }
  for (; k < n; k++) {
    const double x = (double)segval[b + 64ull * k];
    segval[b + 64ull * k] = (C)run;
    run = run + x;
  }
}

// Multi-block exclusive scan of uint32 counts into uint64 offsets:
// per-block totals, a single-block scan of those, then the block-local scans.
constexpr int kScanBlock = 1024;

__device__ __forceinline__ unsigned long long scan_item(const unsigned int* cnt, int64_t i,
                                                        int64_t n, int words) {
  if (i >= n) return 0ull;
  const unsigned int c = cnt[i];
  return words ? (unsigned long long)((c + 15u) >> 4) : (unsigned long long)c;
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_partial(const unsigned int* __restrict__ cnt, int64_t n, int words,
               unsigned long long* __restrict__ bsum) {
  __shared__ unsigned long long s_w[kScanBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  unsigned long long x = scan_item(cnt, i, n, words);
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kScanBlock / 64; w++) t += s_w[w];
    bsum[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_blocksums(unsigned long long* __restrict__ bsum, int64_t nb) {
  __shared__ unsigned long long s_w[kScanBlock / 64];
  __shared__ unsigned long long s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += kScanBlock) {
    const int64_t i = b0 + threadIdx.x;
    const unsigned long long v = i < nb ? bsum[i] : 0ull;
    unsigned long long x = v;
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned long long y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    unsigned long long wb = 0, tot = 0;
    for (int w = 0; w < kScanBlock / 64; w++) {
      if (w < wave) wb += s_w[w];
      tot += s_w[w];
    }
    const unsigned long long base = s_base;
    if (i < nb) bsum[i] = base + wb + x - v;
    __syncthreads();
    if (threadIdx.x == 0) s_base = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = s_base;
}

__global__ void __launch_bounds__(kScanBlock)
k_scan_final(const unsigned int* __restrict__ cnt, int64_t n, int words,
             const unsigned long long* __restrict__ bsum, int64_t nb,
             unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_w[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  const unsigned long long v = scan_item(cnt, i, n, words);
  unsigned long long x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  unsigned long long wb = 0;
  for (int w = 0; w < wave; w++) wb += s_w[w];
  if (i < n) out[i] = bsum[blockIdx.x] + wb + x - v;
  if (i == n - 1) out[n] = bsum[nb];
}

// ---------------------------------------------------------------------------
// Launchers.

hipError_t launch_tile_count(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                             TileGrid tg, unsigned int* nseg, unsigned int* tile_cnt,
                             const SrcRows& rows, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_count, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, tg, nseg, tile_cnt, rows);
  return hipGetLastError();
}

hipError_t launch_tile_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            TileGrid tg, const int* sid, const unsigned long long* val_off,
                            const unsigned long long* tile_ptr, unsigned int* cursor,
                            SegRec* segs, const SrcRows& rows, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_fill, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, tg, sid, val_off, tile_ptr, cursor, segs, rows);
  return hipGetLastError();
}

size_t sort_segments_tmp_bytes(int64_t nsegs, int nslots) {
  size_t cub = 0;
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, cub, (const unsigned int*)nullptr,
                                              (unsigned int*)nullptr, (const unsigned int*)nullptr,
                                              (unsigned int*)nullptr, (int)nsegs, nslots,
                                              (const unsigned int*)nullptr,
                                              (const unsigned int*)nullptr, 0, 32, nullptr);
  const size_t arrays = 4 * sizeof(unsigned int) * (size_t)nsegs + sizeof(unsigned int) * (nslots + 1);
  return ((cub + 255) & ~(size_t)255) + ((arrays + 255) & ~(size_t)255) + 256 * 8;
}

hipError_t launch_sort_segments(SegRec* segs, SegRec* out, int64_t nsegs, int nslots,
                                const unsigned long long* slot_ptr, int key_bits, void* tmp,
                                size_t tmp_bytes, hipStream_t st) {
  if (nsegs <= 0) return hipSuccess;
  if (nsegs >= (1ll << 31) || tmp_bytes < sort_segments_tmp_bytes(nsegs, nslots))
    return hipErrorInvalidValue;
  char* const tmp_end = static_cast<char*>(tmp) + tmp_bytes;
  auto carve = [&](size_t bytes) {
    void* p = tmp;
    tmp = static_cast<char*>(tmp) + ((bytes + 255) & ~(size_t)255);
    return p;
  };
  unsigned int* kin = static_cast<unsigned int*>(carve(sizeof(unsigned int) * nsegs));
  unsigned int* kout = static_cast<unsigned int*>(carve(sizeof(unsigned int) * nsegs));
  unsigned int* vin = static_cast<unsigned int*>(carve(sizeof(unsigned int) * nsegs));
  unsigned int* vout = static_cast<unsigned int*>(carve(sizeof(unsigned int) * nsegs));
  unsigned int* off = static_cast<unsigned int*>(carve(sizeof(unsigned int) * (nslots + 1)));
  size_t cub = 0;
  hipError_t e = hipcub::DeviceSegmentedRadixSort::SortPairs(
      nullptr, cub, kin, kout, vin, vout, (int)nsegs, nslots, off, off + 1, 0, key_bits, st);
  if (e != hipSuccess) return e;
  void* cub_tmp = carve(cub);
  if (static_cast<char*>(tmp) > tmp_end) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_seg_keys, dim3(grid_for(nsegs)), dim3(256), 0, st, segs, nsegs, kin, vin);
  hipLaunchKernelGGL(k_u64_to_u32, dim3(grid_for(nslots + 1)), dim3(256), 0, st, slot_ptr,
                     (int64_t)nslots + 1, off);
  e = hipcub::DeviceSegmentedRadixSort::SortPairs(cub_tmp, cub, kin, kout, vin, vout, (int)nsegs,
                                                  nslots, off, off + 1, 0, key_bits, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_seg_gather, dim3(grid_for(nsegs)), dim3(256), 0, st, segs, vout, nsegs,
                     out);
  return hipGetLastError();
}

hipError_t launch_seg_groups(int64_t nrays, const unsigned int* nseg, unsigned int* gcnt,
                             hipStream_t st) {
  const unsigned int ng = (unsigned int)((nrays + 63) / 64);
  hipLaunchKernelGGL(k_seg_groups, dim3(ng), dim3(64), 0, st, nrays, nseg, gcnt);
  return hipGetLastError();
}

hipError_t launch_sell_layout(int nS, int D, const unsigned long long* rowptr,
                              const unsigned long long* key, unsigned int* nslot, hipStream_t st) {
  if (D > kSellDirCap) return hipErrorInvalidValue;
  const unsigned int nch = (unsigned int)((nS + 63) / 64);
  hipLaunchKernelGGL(k_sell_layout, dim3(nch), dim3(64), 0, st, nS, D, rowptr, key, nslot);
  return hipGetLastError();
}

hipError_t launch_sell_fill(int nS, int D, const unsigned long long* rowptr,
                            const unsigned long long* key, const unsigned long long* moff,
                            unsigned long long* masks, unsigned int* cell_eb,
                            unsigned short* sdir, int* ent, hipStream_t st) {
  const unsigned int nch = (unsigned int)((nS + 63) / 64);
  hipLaunchKernelGGL(k_sell_fill, dim3(nch), dim3(64), 0, st, nS, D, rowptr, key, moff, masks,
                     cell_eb, sdir, ent);
  return hipGetLastError();
}

template <typename T, int kMode, int kBits>
static void tile_pass_threads(int threads, int ntiles, const int* tile_list,
                              const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                              TileGrid tg, const WalkParams& wp,
                              const typename Vec4<T>::type* f, const int* sid,
                              const unsigned int* vcnt, const double* p0,
                              const double* p1, const DevScalars* sc, T* segval,
                              long long* out, hipStream_t st) {
#define OMNI_TILE(NT)                                                                        \
  hipLaunchKernelGGL((k_tile_pass<T, kMode, kBits, NT>), dim3(ntiles), dim3(NT), 0, st,       \
                     tile_list, tile_ptr, segs, g, tg, wp, f, sid, vcnt, p0, p1, sc, segval, out)
  if (threads <= 256) OMNI_TILE(256);
  else if (threads <= 512) OMNI_TILE(512);
  else OMNI_TILE(1024);
#undef OMNI_TILE
}

template <typename T, int kMode>
static void tile_pass_bits(int threads, int ntiles, const int* tile_list,
                           const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                           TileGrid tg, const WalkParams& wp, const typename Vec4<T>::type* f,
                           const int* sid, const unsigned int* vcnt, const double* p0,
                           const double* p1, const DevScalars* sc, T* segval,
                           long long* out, hipStream_t st) {
  if (tg.bits == 3)
    tile_pass_threads<T, kMode, 3>(threads, ntiles, tile_list, tile_ptr, segs, g, tg, wp, f, sid,
                                   vcnt, p0, p1, sc, segval, out, st);
  else
    tile_pass_threads<T, kMode, 4>(threads, ntiles, tile_list, tile_ptr, segs, g, tg, wp, f, sid,
                                   vcnt, p0, p1, sc, segval, out, st);
}

template <typename T>
hipError_t launch_tile_pass(int mode, int threads, int ntiles, const int* tile_list,
                            const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                            TileGrid tg, const WalkParams& wp, const void* field, const int* sid,
                            const unsigned int* vcnt, const double* p0, const double* p1, const DevScalars* sc,
                            T* segval, long long* out, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  if (ntiles <= 0) return hipSuccess;
  if (tg.bits < kTileBitsMin || tg.bits > kTileBitsMax) return hipErrorInvalidValue;
  if (mode == kLocalSum)
    tile_pass_bits<T, kLocalSum>(threads, ntiles, tile_list, tile_ptr, segs, g, tg, wp, f, sid,
                                 vcnt, p0, p1, sc, segval, out, st);
  else if (mode == kSourceSums)
    tile_pass_bits<T, kSourceSums>(threads, ntiles, tile_list, tile_ptr, segs, g, tg, wp, f, sid,
                                   vcnt, p0, p1, sc, segval, out, st);
  else
    tile_pass_bits<T, kFull>(threads, ntiles, tile_list, tile_ptr, segs, g, tg, wp, f, sid,
                             vcnt, p0, p1, sc, segval, out, st);
  return hipGetLastError();
}

template hipError_t launch_tile_pass<float>(int, int, int, const int*, const unsigned long long*,
                                            const SegRec*, GridDesc, TileGrid, const WalkParams&,
                                            const void*, const int*, const unsigned int*,
                                            const double*, const double*,
                                            const DevScalars*, float*, long long*,
                                            hipStream_t);
template hipError_t launch_tile_pass<double>(int, int, int, const int*, const unsigned long long*,
                                             const SegRec*, GridDesc, TileGrid,
                                             const WalkParams&, const void*, const int*,
                                             const unsigned int*, const double*, const double*,
                                             const DevScalars*,
                                             double*, long long*, hipStream_t);

size_t scan_u32_tmp_bytes(int64_t n) {
  return sizeof(unsigned long long) * (size_t)((n + kScanBlock - 1) / kScanBlock + 1);
}

hipError_t launch_scan_u32(const unsigned int* cnt, int64_t n, unsigned long long* out,
                           int words, void* tmp, size_t tmp_bytes, hipStream_t st) {
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (n <= 0 || tmp_bytes < scan_u32_tmp_bytes(n)) return hipErrorInvalidValue;
  unsigned long long* bsum = static_cast<unsigned long long*>(tmp);
  hipLaunchKernelGGL(k_scan_partial, dim3(nb), dim3(kScanBlock), 0, st, cnt, n, words, bsum);
  hipLaunchKernelGGL(k_scan_blocksums, dim3(1), dim3(kScanBlock), 0, st, bsum, nb);
  hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(kScanBlock), 0, st, cnt, n, words, bsum, nb, out);
  return hipGetLastError();
}

hipError_t launch_carry_scan(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, double* segval, hipStream_t st) {
  hipLaunchKernelGGL(k_carry_scan<double>, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st,
                     nrays, nseg, val_off, segval);
  return hipGetLastError();
}

hipError_t launch_carry_scan(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, float* segval, hipStream_t st) {
  hipLaunchKernelGGL(k_carry_scan<float>, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st,
                     nrays, nseg, val_off, segval);
  return hipGetLastError();
}

}  // namespace omni3d
