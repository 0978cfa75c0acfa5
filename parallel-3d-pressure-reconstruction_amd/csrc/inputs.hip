// Kernels over the caller's input buffers.
//
//  * k_fingerprint: the kernel-level drop-ins' change detector.  One pass over
//    the directions, the field, the curl and p: an order-free 64-bit hash of
//    the directions and one of the field and curl (sum of a 64-bit mix of
//    (value bits, element index, array) over the elements, added with integer
//    atomics, so the result does not depend on scheduling), and a flag for a
//    non-finite value in any of them.
//  * k_bernoulli_face / k_apply_face: the Bernoulli seed of the j = 0 face,
//    p = pref - rho/2 (u^2 + v^2 + w^2) (kernel.cu:8806-8811), computed once
//    from the caller's velocity and written into the starting pressure.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "launch.hpp"

namespace omni3d {

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long key(float x, int64_t idx, int arr) {
  // the value's bits plus (element, array) times an odd constant
  return (unsigned long long)__float_as_uint(x) +
         ((unsigned long long)idx * 8ull + (unsigned long long)arr) * 0xD1B54A32D192ED03ull;
}

__device__ __forceinline__ bool finite(float x) {
  return (__float_as_uint(x) & 0x7f800000u) != 0x7f800000u;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(256)
k_fingerprint(int D, const float* __restrict__ k1, const float* __restrict__ k2,
              const float* __restrict__ k3, int64_t V, const float* __restrict__ ax,
              const float* __restrict__ ay, const float* __restrict__ az,
              const float* __restrict__ curl, const float* __restrict__ p,
              InputPrint* __restrict__ out) {
  unsigned long long hd = 0, hf = 0;
  unsigned int bad = 0;
  const int64_t n = V > D ? V : D;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < D) {
      const float a = k1[t], b = k2[t], c = k3[t];
      hd += mix64(key(a, t, 0)) + mix64(key(b, t, 1)) + mix64(key(c, t, 2));
      bad |= !(finite(a) && finite(b) && finite(c));
    }
    if (t < V) {
      const float a = ax[t], b = ay[t], c = az[t];
      hf += mix64(key(a, t, 3)) + mix64(key(b, t, 4)) + mix64(key(c, t, 5));
      bool ok = finite(a) && finite(b) && finite(c);
      if (curl) {
        const float q = curl[t];
        hf += mix64(key(q, t, 6));
        ok = ok && finite(q);
      }
      if (p) ok = ok && finite(p[t]);
      bad |= !ok;
    }
  }
  hd = wave_sum(hd);
  hf = wave_sum(hf);
  const unsigned long long anybad = __ballot(bad != 0);
  if ((threadIdx.x & 63) == 0) {
    if (hd) atomicAdd(&out->dirs, hd);
    if (hf) atomicAdd(&out->field, hf);
    if (anybad) atomicOr(&out->nonfinite, 1u);
  }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_bernoulli_face(int X, int Y, int Z, const S* __restrict__ u, const S* __restrict__ v,
                 const S* __restrict__ w, double pref, double rho, double* __restrict__ face) {
  const int64_t n = (int64_t)X * Z;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % X), k = (int)(t / X);
    const int64_t ind = i + (int64_t)k * X * Y;  // j = 0
    const double a = (double)u[ind], b = (double)v[ind], c = (double)w[ind];
    face[t] = pref - 0.5 * rho * (a * a + b * b + c * c);
  }
}

__global__ void __launch_bounds__(256)
k_apply_face(int X, int Y, int Z, const double* __restrict__ face, double* __restrict__ p) {
  const int64_t n = (int64_t)X * Z;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % X), k = (int)(t / X);
    p[i + (int64_t)k * X * Y] = face[t];
  }
}

}  // namespace

hipError_t launch_fingerprint(int D, const float* k1, const float* k2, const float* k3, int64_t V,
                              const float* ax, const float* ay, const float* az,
                              const float* curl, const float* p, InputPrint* out,
                              hipStream_t st) {
  const int64_t n = V > D ? V : D;
  hipLaunchKernelGGL(k_fingerprint, dim3(grid_for(n, 4096)), dim3(256), 0, st, D, k1, k2, k3, V,
                     ax, ay, az, curl, p, out);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_bernoulli_face(int X, int Y, int Z, const S* u, const S* v, const S* w,
                                 double pref, double rho, double* face, hipStream_t st) {
  hipLaunchKernelGGL(k_bernoulli_face<S>, dim3(grid_for((int64_t)X * Z)), dim3(256), 0, st, X, Y,
                     Z, u, v, w, pref, rho, face);
  return hipGetLastError();
}
template hipError_t launch_bernoulli_face<float>(int, int, int, const float*, const float*,
                                                 const float*, double, double, double*,
                                                 hipStream_t);
template hipError_t launch_bernoulli_face<double>(int, int, int, const double*, const double*,
                                                  const double*, double, double, double*,
                                                  hipStream_t);

hipError_t launch_apply_face(int X, int Y, int Z, const double* face, double* p, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_face, dim3(grid_for((int64_t)X * Z)), dim3(256), 0, st, X, Y, Z,
                     face, p);
  return hipGetLastError();
}

}  // namespace omni3d
