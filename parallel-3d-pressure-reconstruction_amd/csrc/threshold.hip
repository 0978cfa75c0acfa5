// Curl threshold from a histogram: thredholdHistMaterialAccCPU
// (kernel.cu:8380-8418), the helper that picks the Select method's curl
// threshold as the value above which a given fraction of the voxels lie.
//
// Same arithmetic as the reference, in FP32: min / max seeded with 1e4 and
// -1e4; bin = (int)((c - min) / (max - min) * 10000); from the top bin down,
// a FP32 running count; the first bin j where it reaches (X*Y*Z) * percentage
// gives j / 10000 * (max - min) + min, else max.  Two differences, both where
// the reference leaves its arrays: the reference's bin walk counts j upwards
// from N-1 (kernel.cu:8406) and the maximum lands in bin N; here the walk
// goes down and the maximum goes into the last bin (NaN quotients, from
// max == min or NaN values, into bin 0).  min / max treat -0 as +0.
//
// Kernels: k_thr_minmax (one pass, one atomic pair per workgroup on an
// order-preserving integer image of the floats), k_thr_hist (per-workgroup
// LDS histogram of the 10000 bins, non-empty bins added to the global one),
// k_thr_pick (one workgroup: the top-down walk, as integer suffix sums over
// the lanes where FP32 counts are exact, else serial in FP32).  HBM traffic: the volume
// read twice (8 B per voxel) plus 40 KB of bins.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/omni3d.h"
#include "driver.hpp"
#include "launch.hpp"

namespace omni3d {
namespace {

constexpr int kBins = 10000;
constexpr int kThrBlock = 1024;

// order-preserving map float -> uint32 (-0 folded onto +0)
__device__ __forceinline__ unsigned int ord_of(float x) {
  unsigned int u = __float_as_uint(x + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of(unsigned int o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__global__ void k_thr_init(unsigned int* mm, unsigned int* hist) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    mm[0] = ord_of(1e4f);
    mm[1] = ord_of(-1e4f);
  }
  for (int b = t; b < kBins; b += gridDim.x * blockDim.x) hist[b] = 0u;
}

__global__ void __launch_bounds__(kThrBlock) k_thr_minmax(const float* __restrict__ c, int64_t V,
                                                         unsigned int* mm) {
  float lo = 1e4f, hi = -1e4f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
    const float x = c[v];
    if (x < lo) lo = x;  // NaN compares false and is skipped, as in the reference
    if (x > hi) hi = x;
  }
  unsigned int olo = ord_of(lo), ohi = ord_of(hi);
  for (int off = 32; off > 0; off >>= 1) {
    olo = min(olo, (unsigned int)__shfl_xor((int)olo, off));
    ohi = max(ohi, (unsigned int)__shfl_xor((int)ohi, off));
  }
  __shared__ unsigned int slo[kThrBlock / 64], shi[kThrBlock / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    slo[w] = olo;
    shi[w] = ohi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
      olo = min(olo, slo[i]);
      ohi = max(ohi, shi[i]);
    }
    atomicMin(&mm[0], olo);
    atomicMax(&mm[1], ohi);
  }
}

__device__ __forceinline__ int bin_of(float x, float lo, float range) {
  const float q = __fmul_rn(__fdiv_rn(__fsub_rn(x, lo), range), (float)kBins);
  return q >= (float)(kBins - 1) ? kBins - 1 : (q > 0.0f ? (int)q : 0);
}

__global__ void __launch_bounds__(kThrBlock) k_thr_hist(const float* __restrict__ c, int64_t V,
                                                       const unsigned int* mm,
                                                       unsigned int* hist) {
  __shared__ unsigned int h[kBins];
  for (int b = threadIdx.x; b < kBins; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  const float lo = float_of(mm[0]), hi = float_of(mm[1]);
  const float range = __fsub_rn(hi, lo);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride)
    atomicAdd(&h[bin_of(c[v], lo, range)], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < kBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// The top-down walk.  Up to 2^24 voxels every running count is an integer
// that FP32 holds exactly, so the count above bin j is an integer suffix sum
// and the walk's first hit is the largest j whose suffix sum reaches need:
// kPickThreads lanes take kBins / kPickThreads bins each, a scan of their
// totals gives each lane the count above its bins.  Larger volumes walk
// serially in FP32 as the reference does.
constexpr int kPickThreads = 1000;
constexpr int kPickBins = kBins / kPickThreads;
static_assert(kPickBins * kPickThreads == kBins, "bins split evenly over the lanes");

__global__ void __launch_bounds__(1024) k_thr_pick(const unsigned int* mm,
                                                   const unsigned int* hist, int64_t V,
                                                   float need, float* out) {
  __shared__ unsigned int h[kBins];
  __shared__ unsigned int sfx[2][1024];
  __shared__ int best;
  const int t = threadIdx.x;
  for (int b = t; b < kBins; b += blockDim.x) h[b] = hist[b];
  if (t == 0) best = -1;
  __syncthreads();
  const float lo = float_of(mm[0]), hi = float_of(mm[1]);
  auto edge = [&](int j) {
    return __fadd_rn(__fmul_rn(__fdiv_rn((float)j, (float)kBins), __fsub_rn(hi, lo)), lo);
  };
  if (V > (1ll << 24)) {
    if (t != 0) return;
    float tot = 0.0f, r = hi;
    for (int j = kBins - 1; j >= 0; j--) {
      tot = __fadd_rn(tot, (float)h[j]);
      if (tot >= need) {
        r = edge(j);
        break;
      }
    }
    out[0] = r;
    return;
  }
  // lane t owns bins [t * kPickBins, (t + 1) * kPickBins); inclusive suffix
  // sums of the lane totals (Hillis-Steele over the 1024 slots, ping-pong)
  unsigned int mine = 0u;
  if (t < kPickThreads)
    for (int q = 0; q < kPickBins; q++) mine += h[t * kPickBins + q];
  sfx[0][t] = mine;
  __syncthreads();
  int cur = 0;
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int v = sfx[cur][t] + (t + off < 1024 ? sfx[cur][t + off] : 0u);
    sfx[cur ^ 1][t] = v;
    cur ^= 1;
    __syncthreads();
  }
  if (t < kPickThreads) {
    unsigned int run = sfx[cur][t] - mine;  // voxels in the bins above this lane's
    for (int q = kPickBins - 1; q >= 0; q--) {
      const int j = t * kPickBins + q;
      run += h[j];
      if ((float)run >= need) {
        atomicMax(&best, j);
        break;
      }
    }
  }
  __syncthreads();
  if (t == 0) out[0] = best >= 0 ? edge(best) : hi;
}

// Scratch of the allocation-free entry point: the min / max pair and the
// kBins bins, then one spare word (40,012 bytes).  Each GPU the library runs
// on holds its own copy until the process ends.  Host code uses only the
// device address hipGetSymbolAddress returns for the current device (the
// array's name in host code is a stand-in, not a device address).
constexpr int kScratchWords = 2 + kBins + 1;
static __device__ __align__(256) unsigned int scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned int* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address, looked up once per device.
int scratch_address(unsigned int** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, OMNI3D_ERR_NODEVICE, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, OMNI3D_ERR_ARG, "device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(nullptr, OMNI3D_ERR_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned int*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return OMNI3D_OK;
}

#define THR_TRY(expr)                                                           \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      return fail(nullptr, OMNI3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                           \
  } while (0)

int threshold_impl(int X, int Y, int Z, const float* curl, float percentage, float* threshold,
                   int mem_kind, hipStream_t st) {
  if (X < 1 || Y < 1 || Z < 1 || (int64_t)X * Y * Z > INT32_MAX)
    return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!curl || !threshold) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  if (mem_kind != OMNI3D_MEM_HOST && mem_kind != OMNI3D_MEM_DEVICE)
    return fail(nullptr, OMNI3D_ERR_ARG, "mem_kind must be OMNI3D_MEM_HOST or OMNI3D_MEM_DEVICE");
  const int64_t V = (int64_t)X * Y * Z;
  // kernel.cu:8408: Imax*Jmax*Kmax*percentage, an int product times a float
  const float need = (float)(int)V * percentage;
  // the 40 KB of bins stay allocated per host thread (the call synchronises
  // before returning, so the next call on this thread may reuse them); a
  // call on another device than the last one allocates them there
  thread_local DevBuf scratch;
  thread_local int scratch_dev = -1;
  int dev = 0;
  THR_TRY(hipGetDevice(&dev));
  if (dev != scratch_dev) {
    scratch.release();
    scratch_dev = dev;
  }
  DevBuf dcurl;
  THR_TRY(scratch.alloc(sizeof(unsigned int) * (2 + kBins) + sizeof(float)));
  unsigned int* mm = scratch.as<unsigned int>();
  unsigned int* hist = mm + 2;
  float* dout = reinterpret_cast<float*>(hist + kBins);
  const float* src = curl;
  if (mem_kind == OMNI3D_MEM_HOST) {
    THR_TRY(dcurl.alloc(sizeof(float) * V));
    THR_TRY(hipMemcpyAsync(dcurl.as<void>(), curl, sizeof(float) * V, hipMemcpyHostToDevice, st));
    src = dcurl.as<float>();
  }
  const unsigned int nb = grid_for((V + 15) / 16, 512);  // >= 16 voxels per lane
  hipLaunchKernelGGL(k_thr_init, dim3(10), dim3(1024), 0, st, mm, hist);
  hipLaunchKernelGGL(k_thr_minmax, dim3(nb), dim3(kThrBlock), 0, st, src, V, mm);
  hipLaunchKernelGGL(k_thr_hist, dim3(nb), dim3(kThrBlock), 0, st, src, V, mm, hist);
  hipLaunchKernelGGL(k_thr_pick, dim3(1), dim3(1024), 0, st, mm, hist, V, need,
                     mem_kind == OMNI3D_MEM_HOST ? dout : threshold);
  THR_TRY(hipGetLastError());
  if (mem_kind == OMNI3D_MEM_HOST)
    THR_TRY(hipMemcpyAsync(threshold, dout, sizeof(float), hipMemcpyDeviceToHost, st));
  THR_TRY(hipStreamSynchronize(st));
  return OMNI3D_OK;
}

// The four kernels on device buffers with the given bins / min-max / result
// scratch (mm: 2 words, then kBins bins); the result goes to out (device).
void launch_threshold(const float* src, int64_t V, float need, unsigned int* mm, float* out,
                      hipStream_t st) {
  unsigned int* hist = mm + 2;
  const unsigned int nb = grid_for((V + 15) / 16, 512);  // >= 16 voxels per lane
  hipLaunchKernelGGL(k_thr_init, dim3(10), dim3(1024), 0, st, mm, hist);
  hipLaunchKernelGGL(k_thr_minmax, dim3(nb), dim3(kThrBlock), 0, st, src, V, mm);
  hipLaunchKernelGGL(k_thr_hist, dim3(nb), dim3(kThrBlock), 0, st, src, V, mm, hist);
  hipLaunchKernelGGL(k_thr_pick, dim3(1), dim3(1024), 0, st, mm, hist, V, need, out);
}

int threshold_noalloc(int X, int Y, int Z, const float* curl, float percentage, float* threshold,
                      hipStream_t st) {
  if (X < 1 || Y < 1 || Z < 1 || (int64_t)X * Y * Z > INT32_MAX)
    return fail(nullptr, OMNI3D_ERR_ARG,
                "bad grid dimensions (at most 2^31 - 1 voxels: the bins count in 32 bits)");
  if (!curl || !threshold) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  unsigned int* mm = nullptr;
  const int rc = scratch_address(&mm);
  if (rc) return rc;
  const int64_t V = (int64_t)X * Y * Z;
  launch_threshold(curl, V, (float)(int)V * percentage, mm, threshold, st);
  THR_TRY(hipGetLastError());
  return OMNI3D_OK;
}

}  // namespace
}  // namespace omni3d

extern "C" {

int omni3d_threshold_hist_noalloc(int X, int Y, int Z, const float* curl_d, float percentage,
                                  float* threshold_d, void* hip_stream) {
  return omni3d::threshold_noalloc(X, Y, Z, curl_d, percentage, threshold_d,
                                   static_cast<hipStream_t>(hip_stream));
}

void thredholdHistMaterialAccCPU(int Imax, int Jmax, int Kmax, float* curl, float percentage,
                                 float* threshold) {
  (void)omni3d::threshold_impl(Imax, Jmax, Kmax, curl, percentage, threshold, OMNI3D_MEM_HOST,
                               nullptr);
}

int omni3d_threshold_hist(int X, int Y, int Z, const float* curl, float percentage,
                          float* threshold, int mem_kind, void* hip_stream) {
  return omni3d::threshold_impl(X, Y, Z, curl, percentage, threshold, mem_kind,
                                static_cast<hipStream_t>(hip_stream));
}

}  // extern "C"
