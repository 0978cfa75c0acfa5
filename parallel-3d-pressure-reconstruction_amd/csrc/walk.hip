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

// Candidate order: direction, then 16x16 patches of the lattice, then 8x8
// sub-patches (one wavefront each), row-major inside.  Neighbouring lanes then
// walk neighbouring parallel lines, which share cache lines of the field.
__device__ __forceinline__ bool candidate_point(int64_t cand, int ng, int64_t per_dir, int* dir,
                                                int64_t* pt) {
  const int side = (ng + 15) >> 4;
  *dir = (int)(cand / per_dir);
  const int64_t r = cand - (int64_t)(*dir) * per_dir;
  const int64_t patch = r >> 8;
  const int l = (int)(r & 255);
  const int pm = (int)(patch / side), pn = (int)(patch % side);
  const int sub = l >> 6, q = l & 63;
  const int m = pm * 16 + (sub >> 1) * 8 + (q >> 3);
  const int n = pn * 16 + (sub & 1) * 8 + (q & 7);
  if (m >= ng || n >= ng) return false;
  *pt = (int64_t)m * ng + n;
  return true;
}

__device__ __forceinline__ bool candidate_ray(const DirConst* __restrict__ dirs, GridDesc g,
                                              int64_t cand, int64_t per_dir, uint32_t* in_p,
                                              uint32_t* out_p, uint32_t* len, int* dir_out,
                                              int64_t* pt_out) {
  int dir;
  int64_t pt;
  if (!candidate_point(cand, g.ng, per_dir, &dir, &pt)) return false;
  const DirConst d = dirs[dir];
  double o[3];
  lattice_origin(d, g.ng, g.ls, pt, o);
  int in[3], out[3];
  if (!clip_line(g.X, g.Y, g.Z, o, d.k, in, out)) return false;
  *in_p = pack_voxel(in[0], in[1], in[2]);
  *out_p = pack_voxel(out[0], out[1], out[2]);
  *len = (uint32_t)(abs(out[0] - in[0]) + abs(out[1] - in[1]) + abs(out[2] - in[2]));
  *dir_out = dir;
  *pt_out = pt;
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
    int dir;
    int64_t pt;
    if (c < ncand && candidate_ray(dirs, g, c, npts, &a, &b, &l, &dir, &pt)) mine++;
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
    int dir = 0;
    int64_t pt = 0;
    const bool ok = c < ncand && candidate_ray(dirs, g, c, npts, &a, &b, &l, &dir, &pt);
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
      RayRec rr;
      rr.in = a;
      rr.out = b;
      rr.point = (uint32_t)pt;
      rr.dir_len = (uint32_t)dir | (l << 16);
      rays[offset + wbase + below] = rr;
    }
    offset += total;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Plan kernels over the rays: source voxels and step totals, and their
// numbering (the CSR of the visits that land on source voxels is counted
// and filled by the tile engine's plan walks, k_tile_count / k_tile_fill).

// Entry voxels (the plain method's sources) and the total staircase steps
// (the L1 length of every ray) straight from the ray records; the visit
// counts come from the tile engine (k_tile_counts).
__global__ void __launch_bounds__(256)
k_plan_flags(const RayRec* __restrict__ rays, int64_t nrays, GridDesc g, uint8_t* __restrict__ src,
             unsigned long long* __restrict__ steps_total) {
  unsigned long long steps = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrays;
       r += (int64_t)gridDim.x * blockDim.x) {
    const RayRec rr = rays[r];
    int c[3];
    unpack_voxel(rr.in, c);
    src[(int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y] = 1;
    steps += rr.dir_len >> 16;
  }
  for (int off = 32; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
  __shared__ unsigned long long s_st[4];
  if ((threadIdx.x & 63) == 0) s_st[threadIdx.x >> 6] = steps;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) steps += s_st[w];
    atomicAdd(steps_total, steps);
  }
}

hipError_t launch_plan_flags(const RayRec* rays, int64_t nrays, GridDesc g, uint8_t* src,
                             unsigned long long* steps, hipStream_t st) {
  hipLaunchKernelGGL(k_plan_flags, dim3(grid_for(nrays, 2048)), dim3(256), 0, st, rays, nrays, g,
                     src, steps);
  return hipGetLastError();
}

// Source numbering on the device.  The flagged (entry) voxels of the box
// surface are numbered face by face (k = 0, k = Z - 1, j = 0, j = Y - 1,
// i = 0, i = X - 1; a voxel on several faces belongs to the first), each face
// in 8x8 blocks of its (u, w) coordinates, blocks row by row, inside a block
// w-major: slot q of face f is block q / 64 at (u, w) = (8 (b % nbu) + q % 8,
// 8 (b / nbu) + q % 64 / 8).  Consecutive ids are thus neighbouring voxels.
__device__ __forceinline__ int surface_face(int i, int j, int k, int X, int Y, int Z) {
  if (k == 0) return 0;
  if (k == Z - 1) return 1;
  if (j == 0) return 2;
  if (j == Y - 1) return 3;
  if (i == 0) return 4;
  if (i == X - 1) return 5;
  return 6;
}

// The voxel of slot q, or -1 where the slot lies outside its face or its
// voxel belongs to an earlier face.
__device__ __forceinline__ int64_t face_slot_voxel(const FaceSlots& fs, GridDesc g,
                                                    unsigned long long q) {
  int f = 0;
  while (f < 5 && q >= fs.off[f + 1]) f++;
  const unsigned long long pos = q - fs.off[f];
  const unsigned long long blk = pos >> 6;
  const int in = (int)(pos & 63ull);
  const int nbu = (fs.nu[f] + 7) >> 3;
  const int u = (int)(blk % (unsigned long long)nbu) * 8 + (in & 7);
  const int w = (int)(blk / (unsigned long long)nbu) * 8 + (in >> 3);
  if (u >= fs.nu[f] || w >= fs.nw[f]) return -1;
  int i, j, k;
  switch (f) {
    case 0: i = u; j = w; k = 0; break;
    case 1: i = u; j = w; k = g.Z - 1; break;
    case 2: i = u; j = 0; k = w; break;
    case 3: i = u; j = g.Y - 1; k = w; break;
    case 4: i = 0; j = u; k = w; break;
    default: i = g.X - 1; j = u; k = w; break;
  }
  if (surface_face(i, j, k, g.X, g.Y, g.Z) != f) return -1;
  return (int64_t)i + (int64_t)j * g.X + (int64_t)k * g.X * g.Y;
}

__global__ void __launch_bounds__(256)
k_face_slot_flags(FaceSlots fs, GridDesc g, const uint8_t* __restrict__ flag,
                  unsigned int* __restrict__ valid) {
  const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= fs.off[6]) return;
  const int64_t v = face_slot_voxel(fs, g, q);
  valid[q] = v >= 0 && flag[v] ? 1u : 0u;
}

__global__ void __launch_bounds__(256)
k_face_slot_ids(FaceSlots fs, GridDesc g, const unsigned int* __restrict__ valid,
                const unsigned long long* __restrict__ off, int* __restrict__ sid,
                int* __restrict__ slist) {
  const unsigned long long q = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= fs.off[6] || !valid[q]) return;
  const int64_t v = face_slot_voxel(fs, g, q);
  const int s = (int)off[q];
  sid[v] = s;
  slist[s] = (int)v;
}

// Flagged voxels off the box surface (none for rays that enter at the
// surface; the host numbering handles them if any).
__global__ void __launch_bounds__(256)
k_interior_flags(GridDesc g, const uint8_t* __restrict__ flag, unsigned int* __restrict__ count) {
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  unsigned int c = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[v]) continue;
    const int i = (int)(v % g.X), j = (int)((v / g.X) % g.Y), k = (int)(v / ((int64_t)g.X * g.Y));
    if (surface_face(i, j, k, g.X, g.Y, g.Z) == 6) c++;
  }
  if (c) atomicAdd(count, c);
}

FaceSlots face_slots(int X, int Y, int Z) {
  FaceSlots fs{};
  const int nu[6] = {X, X, X, X, Y, Y}, nw[6] = {Y, Y, Z, Z, Z, Z};
  fs.off[0] = 0;
  for (int f = 0; f < 6; f++) {
    fs.nu[f] = nu[f];
    fs.nw[f] = nw[f];
    fs.off[f + 1] = fs.off[f] + (unsigned long long)((nu[f] + 7) / 8) *
                                    (unsigned long long)((nw[f] + 7) / 8) * 64ull;
  }
  return fs;
}

hipError_t launch_face_slot_flags(const FaceSlots& fs, GridDesc g, const uint8_t* flag,
                                  unsigned int* valid, unsigned int* interior, hipStream_t st) {
  const unsigned long long n = fs.off[6];
  hipLaunchKernelGGL(k_face_slot_flags, dim3((unsigned int)((n + 255) / 256)), dim3(256), 0, st,
                     fs, g, flag, valid);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  hipLaunchKernelGGL(k_interior_flags, dim3(grid_for(V, 4096)), dim3(256), 0, st, g, flag,
                     interior);
  return hipGetLastError();
}

hipError_t launch_face_slot_ids(const FaceSlots& fs, GridDesc g, const unsigned int* valid,
                                const unsigned long long* off, int* sid, int* slist,
                                hipStream_t st) {
  const unsigned long long n = fs.off[6];
  hipLaunchKernelGGL(k_face_slot_ids, dim3((unsigned int)((n + 255) / 256)), dim3(256), 0, st,
                     fs, g, valid, off, sid, slist);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient walks of a reconstruction (plain method).
//
// Pass A: the pint part of every visit that lands on a source voxel
//   (bs[sid(v)] += fix(pint)).
// Pass B: every visit, p_src + pint (sum[v] += fix(p_src + pint)).

template <typename T, bool kFull, bool kTri>
__global__ void __launch_bounds__(256)
k_walk_plain(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, WalkParams wp, const typename Vec4<T>::type* __restrict__ field,
             const int* __restrict__ sid, const double* __restrict__ pbuf0,
             const double* __restrict__ pbuf1, const DevScalars* __restrict__ sc,
             long long* __restrict__ acc, int64_t copy_stride, int ncopies) {
  const double scale = sc->scale;
  // Pass B reads p^{k*-1} from the buffer the iteration left it in.
  const double* __restrict__ ps = ((sc->kstar - 1) & 1) ? pbuf1 : pbuf0;
  // private accumulator copy of this block's XCD (summed by the averaging kernel)
  long long* __restrict__ mine = acc + (int64_t)(blockIdx.x % (unsigned)ncopies) * copy_stride;
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const RayRec rr = rays[r];
  int c[3];
  unpack_voxel(rr.in, c);
  const int64_t vin = (int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y;
  const double pin = kFull ? ps[sid[vin]] : 0.0;
  PlainAccum<kFull> vis{mine, sid, scale, pin};
  if constexpr (kTri) walk_gradient_tri<T>(rr, dirs, g, wp, field, vis);
  else walk_gradient<T>(rr, dirs, g, wp, field, vis);
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



template <typename T>
hipError_t launch_walk_plain(bool full, bool trilinear, const RayRec* rays, int64_t nrays,
                             const DirConst* dirs, GridDesc g, const WalkParams& wp,
                             const void* field, const int* sid, const double* p0,
                             const double* p1, const DevScalars* sc, long long* acc,
                             int64_t copy_stride, int ncopies, hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  const unsigned int grid = grid_for(nrays, 1u << 30);
  const V4* f = static_cast<const V4*>(field);
#define OMNI_WP(F, R)                                                                          \
  hipLaunchKernelGGL((k_walk_plain<T, F, R>), dim3(grid), dim3(256), 0, st, rays, nrays, dirs, \
                     g, wp, f, sid, p0, p1, sc, acc, copy_stride, ncopies)
  if (full) {
    if (trilinear) OMNI_WP(true, true);
    else OMNI_WP(true, false);
  } else {
    if (trilinear) OMNI_WP(false, true);
    else OMNI_WP(false, false);
  }
#undef OMNI_WP
  return hipGetLastError();
}

#define OMNI_WP_INST(T)                                                                        \
  template hipError_t launch_walk_plain<T>(bool, bool, const RayRec*, int64_t, const DirConst*, \
                                           GridDesc, const WalkParams&, const void*,           \
                                           const int*, const double*, const double*,           \
                                           const DevScalars*, long long*, int64_t, int,        \
                                           hipStream_t);
OMNI_WP_INST(float)
OMNI_WP_INST(double)
#undef OMNI_WP_INST

}  // namespace omni3d
