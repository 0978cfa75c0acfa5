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

#include "kernels.hpp"
#include "launch.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {

__device__ __forceinline__ int tile_of(int i, int j, int k, const TileGrid& tg) {
  return (i >> kTileBits) + (j >> kTileBits) * tg.ntx + (k >> kTileBits) * tg.ntx * tg.nty;
}

// ---------------------------------------------------------------------------
// Plan kernels.  Visitor-style geometry walks (the same step rule as every
// other walk) that count steps and segments, then write codes and records.

__global__ void __launch_bounds__(256)
k_tile_count(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, TileGrid tg, unsigned int* __restrict__ nsteps,
             unsigned int* __restrict__ nseg, unsigned int* __restrict__ tile_cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  Walker w;
  w.init(rays[r], dirs, g);
  int tile = tile_of(w.c0, w.c1, w.c2, tg);
  unsigned int segs = 1, steps = 0;
  atomicAdd(&tile_cnt[tile], 1u);
  while (!w.done()) {
    int s;
    if (w.advance(&s) < 0) break;
    steps++;
    const int t = tile_of(w.c0, w.c1, w.c2, tg);
    if (t != tile) {
      tile = t;
      segs++;
      atomicAdd(&tile_cnt[tile], 1u);
    }
  }
  nsteps[r] = steps;
  nseg[r] = segs;
}

__global__ void __launch_bounds__(256)
k_tile_fill(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, TileGrid tg, const unsigned long long* __restrict__ word_off,
            const unsigned long long* __restrict__ seg_base,
            const unsigned long long* __restrict__ tile_ptr, unsigned int* __restrict__ cursor,
            uint32_t* __restrict__ words, SegRec* __restrict__ segs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  Walker w;
  w.init(rays[r], dirs, g);
  uint32_t* wp = words + word_off[r];
  unsigned long long gseg = seg_base[r];
  int tile = tile_of(w.c0, w.c1, w.c2, tg);
  uint32_t cur = 0;
  unsigned int step = 0, seg_start = 0;
  uint32_t seg_vox = pack_voxel(w.c0, w.c1, w.c2);
  auto emit = [&](unsigned int end_visit) {
    // visits seg_start .. end_visit-1 lie in `tile`
    const unsigned int slot = atomicAdd(&cursor[tile], 1u);
    SegRec sr;
    sr.ray = (uint32_t)r;
    sr.g = (uint32_t)gseg;
    sr.s0n = seg_start | ((end_visit - seg_start) << 16);
    sr.vox = seg_vox;
    segs[tile_ptr[tile] + slot] = sr;
    gseg++;
  };
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;
    cur |= (uint32_t)a << (2 * (step & 15));
    step++;
    if ((step & 15) == 0) {
      wp[(step >> 4) - 1] = cur;
      cur = 0;
    }
    const int t = tile_of(w.c0, w.c1, w.c2, tg);
    if (t != tile) {
      emit(step);
      tile = t;
      seg_start = step;
      seg_vox = pack_voxel(w.c0, w.c1, w.c2);
    }
  }
  if (step & 15) wp[step >> 4] = cur;
  emit(step + 1);
}

// ---------------------------------------------------------------------------
// Tile passes.

template <typename T>
struct TileLds {
  T g[3][kTileVox];
  long long acc[kTileVox];
};

// Step code of step s (1-based) of a ray whose codes start at `words`.
struct CodeReader {
  const uint32_t* __restrict__ words;
  uint32_t cur;
  unsigned int idx;  // index of the next code
  __device__ __forceinline__ void seek(const uint32_t* w, unsigned int first) {
    words = w;
    idx = first;
    cur = w[first >> 4] >> (2 * (first & 15));
  }
  __device__ __forceinline__ int next() {
    const int a = (int)(cur & 3u);
    idx++;
    cur = (idx & 15) ? (cur >> 2) : words[idx >> 4];
    return a;
  }
};

template <typename T, int kMode>
__global__ void __launch_bounds__(256)
k_tile_pass(const int* __restrict__ tile_list, const unsigned long long* __restrict__ tile_ptr,
            const SegRec* __restrict__ segs, const RayRec* __restrict__ rays,
            const unsigned long long* __restrict__ word_off, const uint32_t* __restrict__ words,
            GridDesc g, TileGrid tg, WalkParams wp,
            const typename Vec4<T>::type* __restrict__ field, const int* __restrict__ sid,
            const double* __restrict__ pbuf0, const double* __restrict__ pbuf1,
            const DevScalars* __restrict__ sc, double* __restrict__ segval,
            long long* __restrict__ out) {
  using V4 = typename Vec4<T>::type;
  __shared__ TileLds<T> L;
  const int tile = tile_list ? tile_list[blockIdx.x] : (int)blockIdx.x;
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kTileBits, oy = ty << kTileBits, oz = tz << kTileBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  // stage the tile's gradient (zero outside the grid) and clear the counters
  for (int l = threadIdx.x; l < kTileVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kTileBits) & kTileMask),
              k = oz + (l >> (2 * kTileBits));
    V4 v;
    if (i < g.X && j < g.Y && k < g.Z) v = field[i + j * sy + k * sz];
    else v.x = v.y = v.z = v.w = 0;
    L.g[0][l] = v.x;
    L.g[1][l] = v.y;
    L.g[2][l] = v.z;
    if (kMode != kLocalSum) L.acc[l] = 0;
  }
  __syncthreads();
  const double scale = sc->scale;
  const double* __restrict__ ps = ((sc->kstar - 1) & 1) ? pbuf1 : pbuf0;
  const unsigned long long b = tile_ptr[tile], e = tile_ptr[tile + 1];
  for (unsigned long long j = b + threadIdx.x; j < e; j += blockDim.x) {
    const SegRec sr = segs[j];
    const RayRec rr = rays[sr.ray];
    int ci[3], co[3], cv[3];
    unpack_voxel(rr.in, ci);
    unpack_voxel(rr.out, co);
    unpack_voxel(sr.vox, cv);
    const int sg0 = co[0] > ci[0] ? 1 : -1, sg1 = co[1] > ci[1] ? 1 : -1,
              sg2 = co[2] > ci[2] ? 1 : -1;
    const unsigned int s0 = sr.s0n & 0xffffu, n = sr.s0n >> 16;
    int lv = (cv[0] - ox) + ((cv[1] - oy) << kTileBits) + ((cv[2] - oz) << (2 * kTileBits));
    double pin = 0.0, carry = 0.0;
    if (kMode == kFull) {
      const int64_t vin = (int64_t)ci[0] + ci[1] * sy + ci[2] * sz;
      pin = ps[sid[vin]];
    }
    if (kMode != kLocalSum) carry = segval[sr.g];
    RaySum<T> loc;
    CodeReader cr;
    const uint32_t* w = words + word_off[sr.ray];
    unsigned int vis = 0;
    if (s0 == 0) {
      // the entry visit: self-add p_src (pint = 0)
      if (kMode == kFull) atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[lv]),
                                    (unsigned long long)to_fixed(pin, scale));
      cr.seek(w, 0);
      vis = 1;
    } else {
      // crossing step into the tile: the previous voxel lies outside it
      cr.seek(w, s0 - 1);
      const int a = cr.next();
      const int sgn = a == 0 ? sg0 : (a == 1 ? sg1 : sg2);
      const int64_t vg = (int64_t)cv[0] + cv[1] * sy + cv[2] * sz;
      const int64_t vp = vg - sgn * (a == 0 ? 1 : (a == 1 ? sy : sz));
      const T prev = component(field[vp], a);
      loc.add(trapezoid<T>(wp, a, sgn, L.g[a][lv], prev));
      if (kMode == kFull)
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[lv]),
                  (unsigned long long)to_fixed(pin + (carry + loc.value()), scale));
      if (kMode == kSourceSums && sid[vg] >= 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[lv]),
                  (unsigned long long)to_fixed(carry + loc.value(), scale));
      vis = 1;
    }
    for (; vis < n; vis++) {
      const int a = cr.next();
      const int sgn = a == 0 ? sg0 : (a == 1 ? sg1 : sg2);
      const int ln = lv + sgn * (a == 0 ? 1 : (a == 1 ? kTile : kTile * kTile));
      loc.add(trapezoid<T>(wp, a, sgn, L.g[a][ln], L.g[a][lv]));
      lv = ln;
      if (kMode == kFull)
        atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[lv]),
                  (unsigned long long)to_fixed(pin + (carry + loc.value()), scale));
      if (kMode == kSourceSums) {
        const int i = ox + (lv & kTileMask), jj = oy + ((lv >> kTileBits) & kTileMask),
                  k = oz + (lv >> (2 * kTileBits));
        const bool surface =
            i == 0 || i == g.X - 1 || jj == 0 || jj == g.Y - 1 || k == 0 || k == g.Z - 1;
        if (surface && sid[i + jj * sy + k * sz] >= 0)
          atomicAdd(reinterpret_cast<unsigned long long*>(&L.acc[lv]),
                    (unsigned long long)to_fixed(carry + loc.value(), scale));
      }
    }
    if (kMode == kLocalSum) segval[sr.g] = loc.value();
  }
  if (kMode == kLocalSum) return;
  __syncthreads();
  for (int l = threadIdx.x; l < kTileVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kTileBits) & kTileMask),
              k = oz + (l >> (2 * kTileBits));
    if (i >= g.X || j >= g.Y || k >= g.Z) continue;
    const int64_t v = i + j * sy + k * sz;
    if (kMode == kFull) out[v] = L.acc[l];
    if (kMode == kSourceSums) {
      const int s = sid[v];
      if (s >= 0) out[s] = L.acc[l];
    }
  }
}

// Per-ray exclusive scan of the segment sums into carries, in place.
__global__ void __launch_bounds__(256)
k_carry_scan(int64_t nrays, const unsigned long long* __restrict__ seg_base,
             double* __restrict__ segval) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const unsigned long long b = seg_base[r], e = seg_base[r + 1];
  double run = 0.0;
  for (unsigned long long j = b; j < e; j++) {
    const double x = segval[j];
    segval[j] = run;
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
                             TileGrid tg, unsigned int* nsteps, unsigned int* nseg,
                             unsigned int* tile_cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_count, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, tg, nsteps, nseg, tile_cnt);
  return hipGetLastError();
}

hipError_t launch_tile_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            TileGrid tg, const unsigned long long* word_off,
                            const unsigned long long* seg_base, const unsigned long long* tile_ptr,
                            unsigned int* cursor, uint32_t* words, SegRec* segs, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_fill, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, tg, word_off, seg_base, tile_ptr, cursor, words, segs);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tile_pass(int mode, int ntiles, const int* tile_list,
                            const unsigned long long* tile_ptr, const SegRec* segs,
                            const RayRec* rays, const unsigned long long* word_off,
                            const uint32_t* words, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const void* field, const int* sid, const double* p0,
                            const double* p1, const DevScalars* sc, double* segval,
                            long long* out, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const V4* f = static_cast<const V4*>(field);
  if (ntiles <= 0) return hipSuccess;
  if (mode == kLocalSum)
    hipLaunchKernelGGL((k_tile_pass<T, kLocalSum>), dim3(ntiles), dim3(256), 0, st, tile_list,
                       tile_ptr, segs, rays, word_off, words, g, tg, wp, f, sid, p0, p1, sc,
                       segval, out);
  else if (mode == kSourceSums)
    hipLaunchKernelGGL((k_tile_pass<T, kSourceSums>), dim3(ntiles), dim3(256), 0, st, tile_list,
                       tile_ptr, segs, rays, word_off, words, g, tg, wp, f, sid, p0, p1, sc,
                       segval, out);
  else
    hipLaunchKernelGGL((k_tile_pass<T, kFull>), dim3(ntiles), dim3(256), 0, st, tile_list,
                       tile_ptr, segs, rays, word_off, words, g, tg, wp, f, sid, p0, p1, sc,
                       segval, out);
  return hipGetLastError();
}

template hipError_t launch_tile_pass<float>(int, int, const int*, const unsigned long long*,
                                            const SegRec*, const RayRec*,
                                            const unsigned long long*, const uint32_t*, GridDesc,
                                            TileGrid, const WalkParams&, const void*, const int*,
                                            const double*, const double*, const DevScalars*,
                                            double*, long long*, hipStream_t);
template hipError_t launch_tile_pass<double>(int, int, const int*, const unsigned long long*,
                                             const SegRec*, const RayRec*,
                                             const unsigned long long*, const uint32_t*, GridDesc,
                                             TileGrid, const WalkParams&, const void*, const int*,
                                             const double*, const double*, const DevScalars*,
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

hipError_t launch_carry_scan(int64_t nrays, const unsigned long long* seg_base, double* segval,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_carry_scan, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, nrays,
                     seg_base, segval);
  return hipGetLastError();
}

}  // namespace omni3d
