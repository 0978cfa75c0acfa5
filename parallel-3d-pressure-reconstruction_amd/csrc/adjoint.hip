// Adjoint of the linear map g = grad p -> p (plain and plain-mask methods,
// zero initial pressure, K iterations, fixed unit weights).
//
// Forward: p_k = A p_{k-1} + b(g), p_0 = 0, so p_K = sum_{j<K} A^j b with
//   (A p)[v] = (1/C[v]) sum_{visits (r,t) -> v} p[src(r,t)]
//   b[v]     = (1/C[v]) sum_{visits (r,t) -> v} pint_r(t),
//   pint_r(t) = sum_{contributing steps tau <= t of the segment} h_tau (g_a[n_tau] + g_a[l_tau]),
//   h = s * delta_a / 2 (the trapezoid of kernel.cu:1816 with a = -g / rho).
// Backward, for y = dL/dp_K:
//   z = sum_{j<K} (A^T)^j y  (Horner: z <- y + A^T z, K-1 times),
//   A^T w = P^T (w / C)     : every segment gathers u = w / C along its
//                             visits into its source voxel,
//   dL/dg = B^T (z / C)     : step tau of a segment scatters h_tau * S_tau into
//                             g_a at its two voxels, S_tau the suffix sum of u
//                             over the segment's visits from tau on.
// The suffix sums need the segment total first: each segment is walked
// twice from a saved walker state (no per-ray storage).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "launch.hpp"
#include "walk.hpp"

namespace omni3d {

// Segment walker of the plain (one segment from the entry voxel, which is
// also a visit of its own) and plain-mask rules (kernel.cu:2280-2537 with
// unit weights, the symmetric rule and no threshold: a segment starts at a
// valid voxel and collects valid voxels until the first masked one).
template <bool kMask>
struct SegWalker {
  Walker w;
  const double* __restrict__ mask;
  bool done = false;

  // Moves to the next segment source; false at the end of the ray.
  __device__ __forceinline__ bool next_source(bool first) {
    if (!kMask) return first;
    if (first && mask[w.v] != 0.0) return true;
    while (!w.done()) {
      int s;
      if (w.advance(&s) < 0) return false;
      if (mask[w.v] != 0.0) return true;
    }
    return false;
  }
  // Takes one contributing step of the current segment: false when the
  // segment ends (masked voxel or end of the ray).  On a masked voxel the
  // walker is left there; next_source(false) continues from it.
  __device__ __forceinline__ bool step(int* a_out, int* s_out, int64_t* last) {
    if (w.done()) return false;
    const int64_t l = w.v;
    int s;
    const int a = w.advance(&s);
    if (a < 0) {
      done = true;
      return false;
    }
    if (kMask && mask[w.v] == 0.0) return false;
    *a_out = a;
    *s_out = s;
    *last = l;
    return true;
  }
};

template <bool kMask>
__global__ void __launch_bounds__(256)
k_adj_gather(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, const double* __restrict__ mask, const double* __restrict__ u,
             double* __restrict__ t) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  SegWalker<kMask> sw;
  sw.w.init(rays[r], dirs, g);
  sw.mask = mask;
  bool first = true;
  while (!sw.done && sw.next_source(first)) {
    first = false;
    const int64_t src = sw.w.v;
    double T = kMask ? 0.0 : u[src];  // the plain entry visits its own source
    int a, s;
    int64_t l;
    while (sw.step(&a, &s, &l)) T += u[sw.w.v];
    if (T != 0.0) unsafeAtomicAdd(&t[src], T);
    if (!kMask) break;
  }
}

template <bool kMask>
__global__ void __launch_bounds__(256)
k_adj_scatter(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
              GridDesc g, WalkParams wp, const double* __restrict__ mask,
              const double* __restrict__ u, double* __restrict__ dg, int64_t V) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  SegWalker<kMask> sw;
  sw.w.init(rays[r], dirs, g);
  sw.mask = mask;
  bool first = true;
  while (!sw.done && sw.next_source(first)) {
    first = false;
    const Walker w0 = sw.w;
    double T = 0.0;
    int m = 0, a, s;
    int64_t l;
    while (sw.step(&a, &s, &l)) {
      T += u[sw.w.v];
      m++;
    }
    const Walker wend = sw.w;
    const bool end_done = sw.done;
    sw.w = w0;
    sw.done = false;
    double Q = 0.0;
    for (int i = 0; i < m; i++) {
      sw.step(&a, &s, &l);
      const double hS = wp.hs[a][s > 0] * (T - Q);
      unsafeAtomicAdd(&dg[(int64_t)a * V + sw.w.v], hS);
      unsafeAtomicAdd(&dg[(int64_t)a * V + l], hS);
      Q += u[sw.w.v];
    }
    sw.w = wend;
    sw.done = end_done;
    if (!kMask) break;
  }
}

// Unit weights of the plain-mask walk (its visit counts C).
__global__ void __launch_bounds__(256)
k_adj_count(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, const double* __restrict__ mask, unsigned int* __restrict__ cnt) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  SegWalker<true> sw;
  sw.w.init(rays[r], dirs, g);
  sw.mask = mask;
  bool first = true;
  while (!sw.done && sw.next_source(first)) {
    first = false;
    int a, s;
    int64_t l;
    while (sw.step(&a, &s, &l)) atomicAdd(&cnt[sw.w.v], 1u);
  }
}

// u = z / C where C > 0, else 0.
__global__ void __launch_bounds__(256)
k_adj_scale(int64_t V, const double* __restrict__ z, const unsigned int* __restrict__ cnt,
            double* __restrict__ u) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int c = cnt[v];
    u[v] = c > 0 ? z[v] / (double)c : 0.0;
  }
}

// z = y + t where C > 0 (voxels no ray updates keep p = 0 and pass nothing on).
__global__ void __launch_bounds__(256)
k_adj_horner(int64_t V, const double* __restrict__ y, const double* __restrict__ t,
             const unsigned int* __restrict__ cnt, double* __restrict__ z) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x)
    z[v] = cnt[v] > 0 ? y[v] + t[v] : y[v];
}

hipError_t launch_adj_gather(bool mask_rule, const RayRec* rays, int64_t nrays,
                             const DirConst* dirs, GridDesc g, const double* mask,
                             const double* u, double* t, hipStream_t st) {
  const unsigned int grid = grid_for(nrays, 1u << 30);
  if (mask_rule)
    hipLaunchKernelGGL(k_adj_gather<true>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       mask, u, t);
  else
    hipLaunchKernelGGL(k_adj_gather<false>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       mask, u, t);
  return hipGetLastError();
}

hipError_t launch_adj_scatter(bool mask_rule, const RayRec* rays, int64_t nrays,
                              const DirConst* dirs, GridDesc g, const WalkParams& wp,
                              const double* mask, const double* u, double* dg, int64_t V,
                              hipStream_t st) {
  const unsigned int grid = grid_for(nrays, 1u << 30);
  if (mask_rule)
    hipLaunchKernelGGL(k_adj_scatter<true>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       wp, mask, u, dg, V);
  else
    hipLaunchKernelGGL(k_adj_scatter<false>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       wp, mask, u, dg, V);
  return hipGetLastError();
}

hipError_t launch_adj_count(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            const double* mask, unsigned int* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_count, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, mask, cnt);
  return hipGetLastError();
}

hipError_t launch_adj_scale(int64_t V, const double* z, const unsigned int* cnt, double* u,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_adj_scale, dim3(grid_for(V)), dim3(256), 0, st, V, z, cnt, u);
  return hipGetLastError();
}

hipError_t launch_adj_horner(int64_t V, const double* y, const double* t,
                             const unsigned int* cnt, double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_horner, dim3(grid_for(V)), dim3(256), 0, st, V, y, t, cnt, z);
  return hipGetLastError();
}

}  // namespace omni3d
