// Boundary-pair (PINT) pipeline of the reference's README (README:18-28):
// the pressure difference between the entry and exit voxels of every ray,
// then Jacobi iterations on the boundary values alone.
//
//   omni3dparallellinesEqualSpacing (kernel.cu:3414-3489) + bodyIntegral
//   (kernel.cu:1353-1467): per accepted ray the staircase trapezoid sum from
//   entry to exit, added into pint[in, out] with pcount[in, out]++;
//   devidecount (kernel.cu:3896-3909): pint /= pcount;
//   BCiteration (kernel.cu:5572-5712): per exit node, pn = mean over the
//   pairs with pint != 0 of p[in] + pint, then p[out] = pn.
//
// The reference keeps pint and pcount as dense n x n matrices over the n
// boundary nodes (n = 6 N^2 voxels: 1.5e11 entries at 256^3); here every ray
// emits one (out, in) key and its integral, a radix sort groups the rays of
// a pair (stable, so a pair's rays add in ray order), and a pair list sorted
// by exit then entry voxel replaces the matrix: per exit node its pairs are
// contiguous and the iteration is one lane per exit node.  The ray sums come
// from the tile engine's segment sums (or one lane per ray with the ray
// engine).  The iteration is
// a Jacobi sweep (the reference's only separation between iterations is a
// __syncthreads(), an inter-block race); an exit node whose pairs are all
// zero keeps its value (the reference divides 0 by 0).  Ray integrals use
// the context's precision (FP64, or FP32 gradient and trapezoid terms,
// added into FP64 segment sums by the tile engine and into
// Kahan-compensated FP32 sums by the ray engine).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "../../include/omni3d.h"
#include "driver.hpp"
#include "launch.hpp"
#include "tiles.hpp"
#include "walk.hpp"

namespace omni3d {
namespace {

constexpr unsigned long long kVoxMask = (1ull << 30) - 1;

__device__ __forceinline__ int64_t lin_of(uint32_t packed, int X, int Y) {
  return (int64_t)(packed & 1023u) + (int64_t)((packed >> 10) & 1023u) * X +
         (int64_t)(packed >> 20) * X * Y;
}

// One lane per ray: key (exit << 30 | entry), value the entry-to-exit sum.
template <typename T>
__global__ void __launch_bounds__(256)
k_bp_integral(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
              GridDesc g, WalkParams wp, const typename Vec4<T>::type* __restrict__ field,
              unsigned long long* __restrict__ key, double* __restrict__ val) {
  using V4 = typename Vec4<T>::type;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const RayRec rr = rays[r];
  Walker w;
  w.init(rr, dirs, g);
  V4 cur = field[w.v];
  RaySum<T> pint;
  while (!w.done()) {
    int s;
    const int a = w.advance(&s);
    if (a < 0) break;  // the reference's walk ends there too (flag == 0)
    const V4 nxt = field[w.v];
    pint.add(trapezoid<T>(wp, a, s, component(nxt, a), component(cur, a)));
    cur = nxt;
  }
  key[r] = ((unsigned long long)rr.out << 30) | (unsigned long long)rr.in;
  val[r] = pint.value();
}

// Tile engine: the ray's sum is the sum of its segment sums (each step of
// the ray is in exactly one segment's sum; stored in the compute precision
// C, added in FP64), read in segment order from the value slots of its
// group of 64 rays.
template <typename C>
__global__ void __launch_bounds__(256)
k_bp_total(const RayRec* __restrict__ rays, int64_t nrays, const unsigned int* __restrict__ nseg,
           const unsigned long long* __restrict__ val_off, const C* __restrict__ segval,
           unsigned long long* __restrict__ key, double* __restrict__ val) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const unsigned long long b = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  double s = 0.0;
  unsigned int k = 0;
  for (; k + 4 <= n; k += 4) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) x[u] = (double)segval[b + 64ull * (k + u)];
#pragma unroll
    for (int u = 0; u < 4; u++) s = s + x[u];
  }
  for (; k < n; k++) s = s + (double)segval[b + 64ull * k];
  const RayRec rr = rays[r];
  key[r] = ((unsigned long long)rr.out << 30) | (unsigned long long)rr.in;
  val[r] = s;
}

// 1 where a run of equal keys starts
__global__ void k_bp_heads(const unsigned long long* __restrict__ key, int64_t n, int shift,
                           unsigned int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || (key[i] >> shift) != (key[i - 1] >> shift)) ? 1u : 0u;
}

// start index of every run (runs numbered by the inclusive scan of the heads)
__global__ void k_bp_starts(const unsigned int* __restrict__ flag,
                            const unsigned int* __restrict__ incl, int64_t n,
                            unsigned int* __restrict__ start) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) start[incl[i] - 1] = (unsigned int)i;
  if (i == n - 1) start[incl[i]] = (unsigned int)n;
}

// pair p: rays start[p] .. start[p+1]-1; mean of their integrals
// (devidecount), entry voxel and exit key
__global__ void k_bp_pairs(const unsigned long long* __restrict__ key,
                           const double* __restrict__ val, const unsigned int* __restrict__ start,
                           int64_t npairs, unsigned long long* __restrict__ pkey,
                           double* __restrict__ pmean) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const unsigned int b = start[p], e = start[p + 1];
  double s = 0.0;
  for (unsigned int i = b; i < e; i++) s += val[i];
  pkey[p] = key[b];
  pmean[p] = s / (double)(e - b);
}

// BCiteration for exit node o: pairs ostart[o] .. ostart[o+1]-1, one
// wavefront per node (coalesced pair reads, a fixed-order tree sum)
constexpr int kBpNodesPerBlock = 4;
__global__ void __launch_bounds__(64 * kBpNodesPerBlock)
k_bp_iterate(int64_t nout, const unsigned int* __restrict__ ostart,
             const unsigned long long* __restrict__ pkey, const double* __restrict__ pmean,
             int X, int Y, const double* __restrict__ p, double* __restrict__ pn) {
  const int64_t o = (int64_t)blockIdx.x * kBpNodesPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (o >= nout) return;  // whole wavefronts leave together
  const unsigned int b = ostart[o], e = ostart[o + 1];
  double s = 0.0;
  int c = 0;
  for (unsigned int i = b + lane; i < e; i += 64) {
    const double m = pmean[i];
    if (m != 0.0) {  // kernel.cu:5671: pairs with pint == 0 are skipped
      s += p[lin_of((uint32_t)(pkey[i] & kVoxMask), X, Y)] + m;
      c++;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    c += __shfl_xor(c, off);
  }
  if (lane == 0) pn[o] = c ? s / (double)c : p[lin_of((uint32_t)(pkey[b] >> 30), X, Y)];
}

__global__ void k_bp_commit(int64_t nout, const unsigned int* __restrict__ ostart,
                            const unsigned long long* __restrict__ pkey, int X, int Y,
                            const double* __restrict__ pn, double* __restrict__ p) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  p[lin_of((uint32_t)(pkey[ostart[o]] >> 30), X, Y)] = pn[o];
}

unsigned int blocks_for(int64_t n) { return (unsigned int)((n + 255) / 256); }

template <typename T>
int boundary_pairs(omni3d_ctx* ctx, int iterations, double* p, int64_t* npairs_out) {
  const omni3d_params& P = ctx->prm;
  Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  const int64_t n = pl.nrays;
  // the sort and scans take an int item count
  if (n > (int64_t)INT32_MAX - 1) return fail(ctx, OMNI3D_ERR_ARG, "too many rays");
  *npairs_out = 0;
  if (n == 0) return OMNI3D_OK;
  // work buffers stay with the context (freeing them costs more than the
  // pipeline); sized for the largest call so far
  DevBuf& k0 = ctx->bp[0];
  DevBuf& k1 = ctx->bp[1];
  DevBuf& v0 = ctx->bp[2];
  DevBuf& v1 = ctx->bp[3];
  DevBuf& flag = ctx->bp[4];
  DevBuf& incl = ctx->bp[5];
  DevBuf& start = ctx->bp[6];
  DevBuf& tmp = ctx->bp[7];
  OMNI_TRY_HIP(ctx, k0.alloc(sizeof(unsigned long long) * n));
  OMNI_TRY_HIP(ctx, k1.alloc(sizeof(unsigned long long) * n));
  OMNI_TRY_HIP(ctx, v0.alloc(sizeof(double) * n));
  OMNI_TRY_HIP(ctx, v1.alloc(sizeof(double) * n));
  OMNI_TRY_HIP(ctx, flag.alloc(sizeof(unsigned int) * n));
  OMNI_TRY_HIP(ctx, incl.alloc(sizeof(unsigned int) * n));
  OMNI_TRY_HIP(ctx, start.alloc(sizeof(unsigned int) * (n + 1)));
  using V4 = typename Vec4<T>::type;
  if (ctx->engine == 0 && pl.ntiles > 0 && pl.nval > 0) {
    // segment sums on the tile engine (pass A's first traversal), then per
    // ray the sum of its segments
    OMNI_TRY_HIP(ctx, ctx->segval.alloc(sizeof(T) * pl.nval));
    OMNI_TRY_HIP(ctx, launch_tile_pass<T>(kLocalSum, ctx->tile_threads, pl.ntiles, nullptr,
                                          pl.tile_ptr1.as<unsigned long long>(),
                                          pl.segs_slot.as<SegRec>(), pl.g, pl.tg, walk_params(P),
                                          ctx->field.as<void>(), pl.sid.as<int>(),
                                          pl.cnt.as<unsigned int>(), nullptr, nullptr,
                                          ctx->scal.as<DevScalars>(), ctx->segval.as<T>(),
                                          nullptr, st));
    hipLaunchKernelGGL(k_bp_total<T>, dim3(blocks_for(n)), dim3(256), 0, st, pl.rays.as<RayRec>(),
                       n, pl.nseg.as<unsigned int>(), pl.val_off.as<unsigned long long>(),
                       ctx->segval.as<T>(), k0.as<unsigned long long>(), v0.as<double>());
  } else {
    hipLaunchKernelGGL(k_bp_integral<T>, dim3(blocks_for(n)), dim3(256), 0, st,
                       pl.rays.as<RayRec>(), n, pl.dirs.as<DirConst>(), pl.g, walk_params(P),
                       ctx->field.as<V4>(), k0.as<unsigned long long>(), v0.as<double>());
  }
  OMNI_TRY_HIP(ctx, hipGetLastError());
  // group the rays of each (exit, entry) pair; stable, so a pair's rays keep
  // the ray order
  size_t sort_bytes = 0, scan_bytes = 0;
  OMNI_TRY_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(
                        nullptr, sort_bytes, k0.as<unsigned long long>(),
                        k1.as<unsigned long long>(), v0.as<double>(), v1.as<double>(), (int)n, 0,
                        60, st));
  OMNI_TRY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, flag.as<unsigned int>(),
                                                     incl.as<unsigned int>(), (int)n, st));
  OMNI_TRY_HIP(ctx, tmp.alloc(std::max(sort_bytes, scan_bytes)));
  OMNI_TRY_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(
                        tmp.as<void>(), sort_bytes, k0.as<unsigned long long>(),
                        k1.as<unsigned long long>(), v0.as<double>(), v1.as<double>(), (int)n, 0,
                        60, st));
  const unsigned long long* key = k1.as<unsigned long long>();
  // pairs: runs of equal keys
  hipLaunchKernelGGL(k_bp_heads, dim3(blocks_for(n)), dim3(256), 0, st, key, n, 0,
                     flag.as<unsigned int>());
  OMNI_TRY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp.as<void>(), scan_bytes,
                                                     flag.as<unsigned int>(),
                                                     incl.as<unsigned int>(), (int)n, st));
  hipLaunchKernelGGL(k_bp_starts, dim3(blocks_for(n)), dim3(256), 0, st, flag.as<unsigned int>(),
                     incl.as<unsigned int>(), n, start.as<unsigned int>());
  unsigned int np32 = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&np32, incl.as<unsigned int>() + (n - 1), sizeof(np32),
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  const int64_t np = np32;
  DevBuf& pkey = ctx->bp[8];
  DevBuf& pmean = ctx->bp[9];
  DevBuf& ostart = ctx->bp[10];
  DevBuf& pn = ctx->bp[11];
  OMNI_TRY_HIP(ctx, pkey.alloc(sizeof(unsigned long long) * np));
  OMNI_TRY_HIP(ctx, pmean.alloc(sizeof(double) * np));
  OMNI_TRY_HIP(ctx, ostart.alloc(sizeof(unsigned int) * (np + 1)));
  hipLaunchKernelGGL(k_bp_pairs, dim3(blocks_for(np)), dim3(256), 0, st, key, v1.as<double>(),
                     start.as<unsigned int>(), np, pkey.as<unsigned long long>(),
                     pmean.as<double>());
  // exit nodes: runs of equal exit voxels in the pair list
  size_t scan2 = 0;
  OMNI_TRY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, scan2, flag.as<unsigned int>(),
                                                     incl.as<unsigned int>(), (int)np, st));
  if (scan2 > scan_bytes) {
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));  // tmp is still in use by the sort
    OMNI_TRY_HIP(ctx, tmp.alloc(std::max(sort_bytes, scan2)));
    scan_bytes = scan2;
  }
  hipLaunchKernelGGL(k_bp_heads, dim3(blocks_for(np)), dim3(256), 0, st,
                     pkey.as<unsigned long long>(), np, 30, flag.as<unsigned int>());
  OMNI_TRY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp.as<void>(), scan_bytes,
                                                     flag.as<unsigned int>(),
                                                     incl.as<unsigned int>(), (int)np, st));
  hipLaunchKernelGGL(k_bp_starts, dim3(blocks_for(np)), dim3(256), 0, st, flag.as<unsigned int>(),
                     incl.as<unsigned int>(), np, ostart.as<unsigned int>());
  unsigned int no32 = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&no32, incl.as<unsigned int>() + (np - 1), sizeof(no32),
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  const int64_t nout = no32;
  OMNI_TRY_HIP(ctx, pn.alloc(sizeof(double) * nout));
  for (int it = 0; it < iterations; it++) {
    hipLaunchKernelGGL(k_bp_iterate,
                       dim3((unsigned int)((nout + kBpNodesPerBlock - 1) / kBpNodesPerBlock)),
                       dim3(64 * kBpNodesPerBlock), 0, st, nout,
                       ostart.as<unsigned int>(), pkey.as<unsigned long long>(),
                       pmean.as<double>(), P.nx, P.ny, p, pn.as<double>());
    hipLaunchKernelGGL(k_bp_commit, dim3(blocks_for(nout)), dim3(256), 0, st, nout,
                       ostart.as<unsigned int>(), pkey.as<unsigned long long>(), P.nx, P.ny,
                       pn.as<double>(), p);
  }
  OMNI_TRY_HIP(ctx, hipGetLastError());
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  *npairs_out = np;
  return OMNI3D_OK;
}

}  // namespace
}  // namespace omni3d

extern "C" int omni3d_boundary_pairs(omni3d_ctx* ctx, int iterations, void* p, int dtype,
                                     int mem_kind, long long* npairs_out) {
  using namespace omni3d;
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (!p) return fail(ctx, OMNI3D_ERR_ARG, "null pressure buffer");
  if (iterations < 0) return fail(ctx, OMNI3D_ERR_ARG, "iterations must be >= 0");
  if (ctx->prm.traversal != OMNI3D_TRAVERSAL_STAIRCASE)
    return fail(ctx, OMNI3D_ERR_ARG, "the boundary pairs are built for the staircase traversal");
  if (dtype != OMNI3D_F32 && dtype != OMNI3D_F64) return fail(ctx, OMNI3D_ERR_ARG, "bad dtype");
  if (mem_kind != OMNI3D_MEM_HOST && mem_kind != OMNI3D_MEM_DEVICE)
    return fail(ctx, OMNI3D_ERR_ARG, "bad mem_kind");
  if (ctx->k.empty()) return fail(ctx, OMNI3D_ERR_STATE, "no directions set");
  if (!ctx->have_field) return fail(ctx, OMNI3D_ERR_STATE, "no field set");
  const omni3d_params& P = ctx->prm;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  int rc = build_plan(ctx);
  if (rc) return rc;
  const size_t esz = elem_size(dtype);
  DevBuf& pd = ctx->bp[12];
  DevBuf stage;
  OMNI_TRY_HIP(ctx, pd.alloc(sizeof(double) * V));
  const void* src = p;
  if (mem_kind == OMNI3D_MEM_HOST) {
    OMNI_TRY_HIP(ctx, stage.alloc(esz * V));
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(stage.as<void>(), p, esz * V, hipMemcpyHostToDevice, st));
    src = stage.as<void>();
  }
  OMNI_TRY_HIP(ctx, dtype == OMNI3D_F64
                        ? (launch_convert<double, double>((const double*)src, pd.as<double>(), V, st))
                        : (launch_convert<float, double>((const float*)src, pd.as<double>(), V, st)));
  int64_t np = 0;
  rc = P.precision == OMNI3D_FP32_KAHAN ? boundary_pairs<float>(ctx, iterations, pd.as<double>(), &np)
                                        : boundary_pairs<double>(ctx, iterations, pd.as<double>(), &np);
  if (rc) return rc;
  void* dst = mem_kind == OMNI3D_MEM_HOST ? stage.as<void>() : p;
  OMNI_TRY_HIP(ctx, dtype == OMNI3D_F64
                        ? (launch_convert<double, double>(pd.as<double>(), (double*)dst, V, st))
                        : (launch_convert<double, float>(pd.as<double>(), (float*)dst, V, st)));
  if (mem_kind == OMNI3D_MEM_HOST)
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(p, stage.as<void>(), esz * V, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (npairs_out) *npairs_out = np;
  return OMNI3D_OK;
}
