// Reductions, fixed-point scale, source-voxel iteration and averaging
// kernels (kernel.cu:6746-6759 and the boundary update of the iteration
// kernel.cu:8840-8861).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.hpp"
#include "launch.hpp"
#include "walk.hpp"

namespace omni3d {

// max |x| over n values into *bits (non-negative doubles order like their
// bit patterns; NaN compares above +inf and is caught by the scale kernel).
template <typename T>
__global__ void __launch_bounds__(256)
k_absmax(const T* __restrict__ x, int64_t n, unsigned long long* __restrict__ bits) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = fabs((double)x[i]);
    m = (a > m || a != a) ? a : m;
  }
  unsigned long long b = (unsigned long long)__double_as_longlong(m);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(bits, b);
}


// Fixed-point scale: every accumulator holds at most cmax contributions of
// magnitude <= pmax + (iters + 1) * gstep * lmax (plus one seed term), so the
// scale 2^e keeps the sums below 2^62.
__global__ void k_compute_scale(DevScalars* sc, double gfac, double lmax, double cmax,
                                double iters) {
  const double amax = __longlong_as_double((long long)sc->amax_bits);
  const double pmax = __longlong_as_double((long long)sc->pmax_bits);
  const double bound = (cmax + 1.0) * (pmax + (iters + 1.0) * gfac * amax * lmax);
  if (!(bound <= DBL_MAX)) {
    sc->status = 5;  // OMNI3D_ERR_NONFINITE
    sc->scale = 1.0;
    sc->inv_scale = 1.0;
    return;
  }
  int e = 40;
  if (bound > 0.0) {
    int ex;
    frexp(bound, &ex);  // bound < 2^ex
    e = 62 - ex;
  }
  if (e > 1000) e = 1000;
  sc->scale = ldexp(1.0, e);
  sc->inv_scale = ldexp(1.0, -e);
}

__global__ void __launch_bounds__(256)
k_gather_sources(const int* __restrict__ slist, int nS, const double* __restrict__ pfull,
                 double* __restrict__ ps) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nS; r += gridDim.x * blockDim.x)
    ps[r] = pfull ? pfull[slist[r]] : 0.0;
}

// One iteration on the source voxels S (affine form of kernel.cu:8840-8861):
//   p_S^k[v] = (sum_{visits of v} fix(p_S^{k-1}[src]) + bs[v] + seed) / C[v]
// where C[v] > 0, else the old value.  One wavefront per row of the visit
// CSR.  The last block reduces the relative change and sets the stop flag;
// once it is set the remaining launches return at once.
constexpr int kIterBlock = 256;

__global__ void __launch_bounds__(kIterBlock)
k_iterate(int k, int nS, const unsigned long long* __restrict__ rowptr,
          const int* __restrict__ col, const int* __restrict__ slist,
          const unsigned int* __restrict__ cnt, const long long* __restrict__ bs,
          const double* __restrict__ pinit, int seed, const double* __restrict__ pold,
          double* __restrict__ pnew, DevScalars* __restrict__ sc, double* __restrict__ partials,
          int max_iter, double tol) {
  if (sc->converged) return;
  constexpr int kWaves = kIterBlock / 64;
  __shared__ double s_d2[kWaves], s_n2[kWaves];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gw = blockIdx.x * kWaves + wave, nw = gridDim.x * kWaves;
  const double scale = sc->scale, inv = sc->inv_scale;
  double d2 = 0.0, n2 = 0.0;
  for (int row = gw; row < nS; row += nw) {
    long long acc = 0;
    const unsigned long long b = rowptr[row], e = rowptr[row + 1];
    for (unsigned long long j = b + lane; j < e; j += 64) acc += to_fixed(pold[col[j]], scale);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
      const int v = slist[row];
      const unsigned int c = cnt[v];
      long long tot = acc + bs[row];
      if (k == 1 && seed && pinit) tot += to_fixed(pinit[v], scale);
      const double po = pold[row];
      const double pn = c > 0 ? (double)tot * inv / (double)c : po;
      pnew[row] = pn;
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
    }
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
  if (s_last && threadIdx.x == 0) {
    __threadfence();
    double a = 0.0, b = 0.0;
    for (unsigned int i = 0; i < gridDim.x; i++) {
      a += __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b += __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Averaging of the final pass (devidecountInner, kernel.cu:6746-6759): where
// C > 0, p = sum / C; elsewhere p keeps its initial value.
__global__ void __launch_bounds__(256)
k_finalize(int64_t V, const long long* __restrict__ sum, const unsigned int* __restrict__ cnt,
           const double* __restrict__ pinit, int seed, const DevScalars* __restrict__ sc,
           double* __restrict__ pd, float* __restrict__ pf, double* __restrict__ wd,
           float* __restrict__ wf) {
  const double inv = sc->inv_scale, scale = sc->scale;
  const bool add_seed = seed && pinit && sc->kstar == 1;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int c = cnt[v];
    const double p0 = pinit ? pinit[v] : 0.0;
    double p = p0;
    if (c > 0) {
      long long tot = sum[v];
      if (add_seed) tot += to_fixed(p0, scale);
      p = (double)tot * inv / (double)c;
    }
    if (pd) pd[v] = p;
    if (pf) pf[v] = (float)p;
    if (wd) wd[v] = (double)c;
    if (wf) wf[v] = (float)c;
  }
}

// Conversions used to stage the input field in the compute precision.
template <typename S, typename D>
__global__ void __launch_bounds__(256) k_convert(const S* __restrict__ s, D* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (D)s[i];
}


// Exclusive scan of 0/1 source flags into compact ids (single block; the
// flag array is only scanned once per plan).
__global__ void __launch_bounds__(1024)
k_source_ids(const uint8_t* __restrict__ flag, int64_t V, int* __restrict__ sid,
             int* __restrict__ slist, int* __restrict__ nS) {
  __shared__ int s_wave[16];
  __shared__ int s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < V; b0 += 1024) {
    const int64_t v = b0 + threadIdx.x;
    const bool f = v < V && flag[v];
    const unsigned long long m = __ballot(f);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int wb = 0, tot = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wave) wb += s_wave[w];
      tot += s_wave[w];
    }
    const int base = s_base;
    if (v < V) {
      if (f) {
        sid[v] = base + wb + below;
        slist[base + wb + below] = (int)v;
      } else {
        sid[v] = -1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *nS = s_base;
}

// Exclusive scan of counts into 64-bit pointers (one block of 1024 threads);
// ptr[n] gets the total.
__device__ void scan_counts(const unsigned int* __restrict__ cnt, int n,
                            unsigned long long* __restrict__ ptr) {
  __shared__ unsigned long long s_wave[16];
  __shared__ unsigned long long s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int b0 = 0; b0 < n; b0 += 1024) {
    const int i = b0 + threadIdx.x;
    unsigned long long x = i < n ? cnt[i] : 0ull;
    // inclusive wave scan
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned long long y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    unsigned long long wb = 0, tot = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wave) wb += s_wave[w];
      tot += s_wave[w];
    }
    const unsigned long long base = s_base;
    if (i < n) ptr[i] = base + wb + x - (unsigned long long)cnt[i];
    __syncthreads();
    if (threadIdx.x == 0) s_base = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) ptr[n] = s_base;
}

__global__ void __launch_bounds__(1024)
k_scan_counts(const unsigned int* __restrict__ cnt, int n, unsigned long long* __restrict__ ptr) {
  scan_counts(cnt, n, ptr);
}

__global__ void __launch_bounds__(256)
k_max_u32(const unsigned int* __restrict__ x, int64_t n, unsigned int* __restrict__ out) {
  unsigned int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = x[i] > m ? x[i] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void __launch_bounds__(256)
k_max_len(const RayRec* __restrict__ rays, int64_t n, unsigned int* __restrict__ out) {
  unsigned int m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int l = rays[i].dir_len >> 16;
    m = l > m ? l : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned int o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ---------------------------------------------------------------------------
// Launchers.

template <typename T>
hipError_t launch_absmax(const T* x, int64_t n, unsigned long long* bits, hipStream_t st) {
  hipLaunchKernelGGL((k_absmax<T>), dim3(grid_for(n, 4096)), dim3(256), 0, st, x, n, bits);
  return hipGetLastError();
}
template hipError_t launch_absmax<float>(const float*, int64_t, unsigned long long*, hipStream_t);
template hipError_t launch_absmax<double>(const double*, int64_t, unsigned long long*,
                                          hipStream_t);

hipError_t launch_compute_scale(DevScalars* sc, double gfac, double lmax, double cmax,
                                double iters, hipStream_t st) {
  hipLaunchKernelGGL(k_compute_scale, dim3(1), dim3(1), 0, st, sc, gfac, lmax, cmax, iters);
  return hipGetLastError();
}

hipError_t launch_gather_sources(const int* slist, int nS, const double* pfull, double* ps,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_gather_sources, dim3(grid_for(nS)), dim3(256), 0, st, slist, nS, pfull,
                     ps);
  return hipGetLastError();
}

unsigned int iterate_grid(int nS) {
  // one wavefront per row, 4 rows in flight per block
  int64_t b = ((int64_t)nS + 3) / 4;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned int)b;
}

hipError_t launch_iterate(int k, int nS, const unsigned long long* rowptr, const int* col,
                          const int* slist, const unsigned int* cnt, const long long* bs,
                          const double* pinit, int seed, const double* pold, double* pnew,
                          DevScalars* sc, double* partials, int max_iter, double tol,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_iterate, dim3(iterate_grid(nS)), dim3(kIterBlock), 0, st, k, nS, rowptr,
                     col, slist, cnt, bs, pinit, seed, pold, pnew, sc, partials, max_iter, tol);
  return hipGetLastError();
}

hipError_t launch_finalize(int64_t V, const long long* sum, const unsigned int* cnt,
                           const double* pinit, int seed, const DevScalars* sc, double* pd,
                           float* pf, double* wd, float* wf, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(V)), dim3(256), 0, st, V, sum, cnt, pinit, seed,
                     sc, pd, pf, wd, wf);
  return hipGetLastError();
}

template <typename S, typename D>
hipError_t launch_convert(const S* s, D* d, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<S, D>), dim3(grid_for(n)), dim3(256), 0, st, s, d, n);
  return hipGetLastError();
}
template hipError_t launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);
template hipError_t launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);
template hipError_t launch_convert<float, float>(const float*, float*, int64_t, hipStream_t);
template hipError_t launch_convert<double, double>(const double*, double*, int64_t, hipStream_t);

hipError_t launch_source_ids(const uint8_t* flag, int64_t V, int* sid, int* slist, int* nS,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_source_ids, dim3(1), dim3(1024), 0, st, flag, V, sid, slist, nS);
  return hipGetLastError();
}

hipError_t launch_scan_counts(const unsigned int* cnt, int n, unsigned long long* ptr,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, cnt, n, ptr);
  return hipGetLastError();
}

hipError_t launch_max_u32(const unsigned int* x, int64_t n, unsigned int* out, hipStream_t st) {
  hipLaunchKernelGGL(k_max_u32, dim3(grid_for(n, 4096)), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

hipError_t launch_max_len(const RayRec* rays, int64_t n, unsigned int* out, hipStream_t st) {
  hipLaunchKernelGGL(k_max_len, dim3(grid_for(n, 4096)), dim3(256), 0, st, rays, n, out);
  return hipGetLastError();
}

}  // namespace omni3d
