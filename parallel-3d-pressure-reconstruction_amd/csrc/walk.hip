// Staircase traversal kernels for gfx950.
//
// One lane walks one ray (line x direction pair) of the precomputed ray list;
// consecutive lanes hold neighbouring parallel lines of one direction, so the
// gradient gathers of a wavefront hit neighbouring voxels.  Accumulation is
// deterministic: every contribution is rounded once to a 64-bit fixed-point
// value (scale 2^e chosen from bounds of the data) and added with integer
// atomics, which are associative, so the result does not depend on the order
// in which lanes arrive.  Averaging then happens in FP64.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "launch.hpp"
#include "walk.hpp"

namespace omni3d {

// ---------------------------------------------------------------------------
// Plan: enumerate accepted rays (kernel.cu:6153-6206 for every line x
// direction pair), in candidate order, with a deterministic compaction.

constexpr int kEnumBlock = 256;
constexpr int kEnumPerThread = 4;

__device__ __forceinline__ bool candidate_ray(const DirConst* __restrict__ dirs, GridDesc g,
                                              int64_t cand, int64_t npts, uint32_t* in_p,
                                              uint32_t* out_p, uint32_t* len) {
  const int dir = (int)(cand / npts);
  const int64_t pt = cand - (int64_t)dir * npts;
  const DirConst d = dirs[dir];
  double o[3];
  lattice_origin(d, g.ng, g.ls, pt, o);
  int in[3], out[3];
  if (!clip_line(g.X, g.Y, g.Z, o, d.k, in, out)) return false;
  *in_p = pack_voxel(in[0], in[1], in[2]);
  *out_p = pack_voxel(out[0], out[1], out[2]);
  *len = (uint32_t)(abs(out[0] - in[0]) + abs(out[1] - in[1]) + abs(out[2] - in[2]));
  return true;
}

__global__ void __launch_bounds__(kEnumBlock)
k_enum_count(const DirConst* __restrict__ dirs, GridDesc g, int64_t npts, int64_t ncand,
             unsigned int* __restrict__ block_counts) {
  __shared__ unsigned int s_cnt;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kEnumBlock * kEnumPerThread;
  unsigned int mine = 0;
  for (int r = 0; r < kEnumPerThread; r++) {
    const int64_t c = base + (int64_t)r * kEnumBlock + threadIdx.x;
    uint32_t a, b, l;
    if (c < ncand && candidate_ray(dirs, g, c, npts, &a, &b, &l)) mine++;
  }
  atomicAdd(&s_cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = s_cnt;
}

// Writes the accepted rays of each block at its scanned offset, keeping the
// candidate order inside the block (wave ballots + LDS prefix of wave counts).
__global__ void __launch_bounds__(kEnumBlock)
k_enum_write(const DirConst* __restrict__ dirs, GridDesc g, int64_t npts, int64_t ncand,
             const unsigned long long* __restrict__ block_offsets, RayRec* __restrict__ rays) {
  constexpr int kWaves = kEnumBlock / 64;
  __shared__ unsigned int s_wave[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kEnumBlock * kEnumPerThread;
  unsigned long long offset = block_offsets[blockIdx.x];
  for (int r = 0; r < kEnumPerThread; r++) {
    const int64_t c = base + (int64_t)r * kEnumBlock + threadIdx.x;
    uint32_t a = 0, b = 0, l = 0;
    const bool ok = c < ncand && candidate_ray(dirs, g, c, npts, &a, &b, &l);
    const unsigned long long m = __ballot(ok);
    const unsigned int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    unsigned int wbase = 0, total = 0;
    for (int w = 0; w < kWaves; w++) {
      if (w < wave) wbase += s_wave[w];
      total += s_wave[w];
    }
    if (ok) {
      const int dir = (int)(c / npts);
      RayRec rr;
      rr.in = a;
      rr.out = b;
      rr.point = (uint32_t)(c - (int64_t)dir * npts);
      rr.dir_len = (uint32_t)dir | (l << 16);
      rays[offset + wbase + below] = rr;
    }
    offset += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Geometry-only walks of the plan: visit counts, source voxels, CSR of the
// visits that land on source voxels.

struct CountVisitor {
  unsigned int* cnt;
  uint8_t* src;
  __device__ __forceinline__ void entry(int64_t v) {
    atomicAdd(&cnt[v], 1u);
    src[v] = 1;
  }
  __device__ __forceinline__ void visit(int64_t v, bool /*surface*/) { atomicAdd(&cnt[v], 1u); }
};

__global__ void __launch_bounds__(256)
k_plan_counts(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
              GridDesc g, unsigned int* __restrict__ cnt, uint8_t* __restrict__ src,
              unsigned long long* __restrict__ steps_total) {
  unsigned long long steps = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrays;
       r += (int64_t)gridDim.x * blockDim.x) {
    CountVisitor vis{cnt, src};
    steps += (unsigned long long)walk_geometry(rays[r], dirs, g, vis);
  }
  // wave-level sum then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(steps_total, steps);
}

struct RowCountVisitor {
  const int* sid;
  unsigned int* rowcnt;
  __device__ __forceinline__ void entry(int64_t v) { atomicAdd(&rowcnt[sid[v]], 1u); }
  __device__ __forceinline__ void visit(int64_t v, bool surface) {
    if (!surface) return;
    const int s = sid[v];
    if (s >= 0) atomicAdd(&rowcnt[s], 1u);
  }
};

__global__ void __launch_bounds__(256)
k_plan_rowcount(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
                GridDesc g, const int* __restrict__ sid, unsigned int* __restrict__ rowcnt) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrays;
       r += (int64_t)gridDim.x * blockDim.x) {
    RowCountVisitor vis{sid, rowcnt};
    walk_geometry(rays[r], dirs, g, vis);
  }
}

struct RowFillVisitor {
  const int* sid;
  const unsigned long long* rowptr;
  unsigned int* cursor;
  int* col;
  int src_id;
  __device__ __forceinline__ void put(int s) {
    const unsigned int pos = atomicAdd(&cursor[s], 1u);
    col[rowptr[s] + pos] = src_id;
  }
  __device__ __forceinline__ void entry(int64_t v) { put(sid[v]); }
  __device__ __forceinline__ void visit(int64_t v, bool surface) {
    if (!surface) return;
    const int s = sid[v];
    if (s >= 0) put(s);
  }
};

__global__ void __launch_bounds__(256)
k_plan_rowfill(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
               GridDesc g, const int* __restrict__ sid, const unsigned long long* __restrict__ rowptr,
               unsigned int* __restrict__ cursor, int* __restrict__ col) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrays;
       r += (int64_t)gridDim.x * blockDim.x) {
    int c[3];
    unpack_voxel(rays[r].in, c);
    const int64_t vin = (int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y;
    RowFillVisitor vis{sid, rowptr, cursor, col, sid[vin]};
    walk_geometry(rays[r], dirs, g, vis);
  }
}

// ---------------------------------------------------------------------------
// Gradient walks of a reconstruction (plain method).
//
// Pass A: the pint part of every visit that lands on a source voxel
//   (bs[sid(v)] += fix(pint)).
// Pass B: every visit, p_src + pint (sum[v] += fix(p_src + pint)).

template <typename T, bool kFull>
__global__ void __launch_bounds__(256)
k_walk_plain(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, WalkParams wp, const T* __restrict__ ax, const T* __restrict__ ay,
             const T* __restrict__ az, const int* __restrict__ sid,
             const double* __restrict__ pbuf0, const double* __restrict__ pbuf1,
             const DevScalars* __restrict__ sc, long long* __restrict__ acc) {
  const double scale = sc->scale;
  // Pass B reads p^{k*-1} from the buffer the iteration left it in.
  const double* __restrict__ ps = ((sc->kstar - 1) & 1) ? pbuf1 : pbuf0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrays;
       r += (int64_t)gridDim.x * blockDim.x) {
    const RayRec rr = rays[r];
    int c[3];
    unpack_voxel(rr.in, c);
    const int64_t vin = (int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y;
    const double pin = kFull ? ps[sid[vin]] : 0.0;
    PlainAccum<T, kFull> vis{acc, sid, scale, pin};
    walk_gradient<T>(rr, dirs, g, wp, ax, ay, az, vis);
  }
}

// ---------------------------------------------------------------------------
// Launchers.

static_assert(kEnumBlock * kEnumPerThread == kEnumItemsPerBlock, "enumeration tiling");

static unsigned int enum_blocks(int64_t ncand) {
  return (unsigned int)((ncand + kEnumItemsPerBlock - 1) / kEnumItemsPerBlock);
}

hipError_t launch_enum_count(const DirConst* dirs, GridDesc g, int64_t npts, int64_t ncand,
                             unsigned int* block_counts, hipStream_t st) {
  hipLaunchKernelGGL(k_enum_count, dim3(enum_blocks(ncand)), dim3(kEnumBlock), 0, st, dirs, g,
                     npts, ncand, block_counts);
  return hipGetLastError();
}

hipError_t launch_enum_write(const DirConst* dirs, GridDesc g, int64_t npts, int64_t ncand,
                             const unsigned long long* block_offsets, RayRec* rays,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_enum_write, dim3(enum_blocks(ncand)), dim3(kEnumBlock), 0, st, dirs, g,
                     npts, ncand, block_offsets, rays);
  return hipGetLastError();
}

hipError_t launch_plan_counts(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                              GridDesc g, unsigned int* cnt, uint8_t* src,
                              unsigned long long* steps, hipStream_t st) {
  hipLaunchKernelGGL(k_plan_counts, dim3(grid_for(nrays)), dim3(256), 0, st, rays, nrays, dirs,
                     g, cnt, src, steps);
  return hipGetLastError();
}

hipError_t launch_plan_rowcount(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                GridDesc g, const int* sid, unsigned int* rowcnt,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_plan_rowcount, dim3(grid_for(nrays)), dim3(256), 0, st, rays, nrays,
                     dirs, g, sid, rowcnt);
  return hipGetLastError();
}

hipError_t launch_plan_rowfill(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                               GridDesc g, const int* sid, const unsigned long long* rowptr,
                               unsigned int* cursor, int* col, hipStream_t st) {
  hipLaunchKernelGGL(k_plan_rowfill, dim3(grid_for(nrays)), dim3(256), 0, st, rays, nrays, dirs,
                     g, sid, rowptr, cursor, col);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_walk_plain(bool full, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                             GridDesc g, const WalkParams& wp, const T* ax, const T* ay,
                             const T* az, const int* sid, const double* p0, const double* p1,
                             const DevScalars* sc, long long* acc, hipStream_t st) {
  const unsigned int grid = grid_for(nrays, 1u << 30);
  if (full)
    hipLaunchKernelGGL((k_walk_plain<T, true>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs,
                       g, wp, ax, ay, az, sid, p0, p1, sc, acc);
  else
    hipLaunchKernelGGL((k_walk_plain<T, false>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs,
                       g, wp, ax, ay, az, sid, p0, p1, sc, acc);
  return hipGetLastError();
}

template hipError_t launch_walk_plain<float>(bool, const RayRec*, int64_t, const DirConst*,
                                             GridDesc, const WalkParams&, const float*,
                                             const float*, const float*, const int*,
                                             const double*, const double*, const DevScalars*,
                                             long long*, hipStream_t);
template hipError_t launch_walk_plain<double>(bool, const RayRec*, int64_t, const DirConst*,
                                              GridDesc, const WalkParams&, const double*,
                                              const double*, const double*, const int*,
                                              const double*, const double*, const DevScalars*,
                                              long long*, hipStream_t);

}  // namespace omni3d
