// Reductions, fixed-point scale, source-voxel iteration and averaging
// kernels (kernel.cu:6746-6759 and the boundary update of the iteration
// kernel.cu:8840-8861).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "kernels.hpp"
#include "launch.hpp"
#include "stage.hpp"
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
  // one atomic per workgroup (atomics on one address serialise)
  __shared__ unsigned long long s_b[4];
  if ((threadIdx.x & 63) == 0) s_b[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = s_b[w] > b ? s_b[w] : b;
    atomicMax(bits, b);
  }
}

// The same over n4 4-vectors (x aligned to a 4-vector, e.g. the packed
// field): one 16-byte (FP32) or 32-byte (FP64) load per thread and step.
template <typename T>
__global__ void __launch_bounds__(256)
k_absmax4(const typename Vec4<T>::type* __restrict__ x, int64_t n4,
          unsigned long long* __restrict__ bits) {
  double m = 0.0;
  auto take = [&](T v) {
    const double a = fabs((double)v);
    m = (a > m || a != a) ? a : m;
  };
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x) {
    const typename Vec4<T>::type v = x[i];
    take(v.x);
    take(v.y);
    take(v.z);
    take(v.w);
  }
  unsigned long long b = (unsigned long long)__double_as_longlong(m);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  __shared__ unsigned long long s_b[4];
  if ((threadIdx.x & 63) == 0) s_b[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = s_b[w] > b ? s_b[w] : b;
    atomicMax(bits, b);
  }
}


// Fixed-point scale: every accumulator holds at most cmax contributions of
// magnitude <= pmax + (iters + 1) * gstep * lmax (plus one seed term), so the
// scale 2^e keeps every term below 2^51 (the range of the rounding in
// to_fixed) and the sums below 2^62.
__global__ void k_compute_scale(DevScalars* sc, double gfac, double lmax, double cmax,
                                double iters) {
  const double amax = __longlong_as_double((long long)sc->amax_bits);
  const double pmax = __longlong_as_double((long long)sc->pmax_bits);
  const double term = pmax + (iters + 1.0) * gfac * amax * lmax;
  const double bound = (cmax + 1.0) * term;
  if (!(bound <= DBL_MAX)) {
    sc->status = 5;  // OMNI3D_ERR_NONFINITE
    sc->scale = 1.0;
    sc->inv_scale = 1.0;
    return;
  }
  int e = 40;
  if (bound > 0.0) {
    int ex, et;
    frexp(bound, &ex);  // bound < 2^ex
    frexp(term, &et);   // term < 2^et
    e = 62 - ex < 51 - et ? 62 - ex : 51 - et;
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
//   p_S^k[v] = (bs[v] + diag[v] p_S^{k-1}[v] + sum_{visits of v} p_S^{k-1}[src]
//               + seed) / C[v]
// where C[v] > 0, else the old value.  One workgroup per 64-row chunk of the
// masked sliced-ELL layout: its kWaves wavefronts take consecutive shares of
// the chunk's slots (one lane per row, sixteen gathers in flight), and the
// partial row sums are added in wavefront order, so the FP64 sum has a fixed
// order.  Each workgroup leaves its share of the relative-change sums in
// partials; k_iter_finish reduces them and sets the stop flag, after which
// the remaining launches return at once.
template <int kWaves, bool kMasked>
__global__ void __launch_bounds__(64 * kWaves)
k_iterate(int k, int nS, const unsigned long long* __restrict__ moff,
          const unsigned long long* __restrict__ masks, const unsigned int* __restrict__ cell_eb,
          const int* __restrict__ ent, const unsigned int* __restrict__ diag,
          const int* __restrict__ slist, const unsigned int* __restrict__ cnt,
          const long long* __restrict__ bs, const double* __restrict__ pinit, int seed,
          const double* __restrict__ pold, double* __restrict__ pnew,
          const DevScalars* __restrict__ sc, double* __restrict__ partials) {
  if (sc->converged) return;
  constexpr int kU = 16;
  // per wavefront: the packed entries of one round of kU slots (at most
  // 64 each) plus the alignment slack of the wide loads
  constexpr int kStage = kU * 64 + 8;
  __shared__ double s_acc[kWaves][64];
  __shared__ __attribute__((aligned(16))) int s_ent[kWaves][kStage];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int* __restrict__ se = s_ent[wave];
  const double inv = sc->inv_scale;
  const unsigned long long below = (1ull << lane) - 1ull;
  double d2 = 0.0, n2 = 0.0;
  const int nchunks = (nS + 63) >> 6;
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const unsigned long long m0 = moff[chunk], m1 = moff[chunk + 1];
    const unsigned long long q = (m1 - m0 + kWaves - 1) / kWaves;
    const unsigned long long ms = m0 + (unsigned long long)wave * q;
    const unsigned long long me = ms + q < m1 ? ms + q : m1;
    unsigned long long eb = ms < me ? (unsigned long long)cell_eb[ms] : 0ull;
    double acc = 0.0;
    unsigned long long s = ms;
    for (; s + kU <= me; s += kU) {
      // the slot masks (scalar loads: the slot is uniform across the
      // wavefront) give the round's entry range [eb, e_end)
      unsigned long long m[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) m[u] = masks[s + u];
      unsigned int off[kU];
      unsigned int n = 0;
#pragma unroll
      for (int u = 0; u < kU; u++) {
        off[u] = n;
        n += (unsigned int)__popcll(m[u]);
      }
      // stage the range with 16-byte loads from the aligned-down start (a
      // few address batches instead of one per slot), then read each
      // lane's entries from LDS
      const unsigned long long a0 = eb & ~3ull;
      const unsigned int sh = (unsigned int)(eb - a0);
      const unsigned int n4 = (sh + n + 3) >> 2;
      const int4* __restrict__ src4 = reinterpret_cast<const int4*>(ent + a0);
      for (unsigned int t = lane; t < n4; t += 64)
        reinterpret_cast<int4*>(se)[t] = src4[t];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bool has[kU];
      int idx[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        has[u] = (m[u] >> lane) & 1ull;
        const unsigned int r = has[u] ? (unsigned int)__popcll(m[u] & below) : 0u;
        idx[u] = se[sh + off[u] + r];
      }
      double x[kU];
      if (kMasked) {
        // only the lanes with an entry issue their gather
#pragma unroll
        for (int u = 0; u < kU; u++) {
          x[u] = 0.0;
          if (has[u]) x[u] = pold[idx[u]];
        }
#pragma unroll
        for (int u = 0; u < kU; u++) acc += x[u];
      } else {
#pragma unroll
        for (int u = 0; u < kU; u++) x[u] = pold[idx[u]];
#pragma unroll
        for (int u = 0; u < kU; u++) acc += has[u] ? x[u] : 0.0;
      }
      eb += n;
      __builtin_amdgcn_wave_barrier();
    }
    for (; s < me; s++) {
      const unsigned long long mm = masks[s];
      if ((mm >> lane) & 1ull) acc += pold[ent[eb + __popcll(mm & below)]];
      eb += __popcll(mm);
    }
    if (kWaves > 1) {
      s_acc[wave][lane] = acc;
      __syncthreads();
    }
    const int row = chunk * 64 + lane;
    if (wave == 0 && row < nS) {
      double tacc = acc;
#pragma unroll
      for (int w = 1; w < kWaves; w++) tacc += s_acc[w][lane];
      const int v = slist[row];
      const unsigned int c = cnt[v];
      const double po = pold[row];
      double tot = (double)bs[row] * inv + (double)diag[row] * po + tacc;
      if (k == 1 && seed && pinit) tot += pinit[v];
      const double pn = c > 0 ? tot / (double)c : po;
      pnew[row] = pn;
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
    }
    if (kWaves > 1) __syncthreads();
  }
  if (wave == 0) {
    for (int off = 32; off > 0; off >>= 1) {
      d2 += __shfl_xor(d2, off);
      n2 += __shfl_xor(n2, off);
    }
    if (lane == 0) {
      partials[2 * blockIdx.x] = d2;
      partials[2 * blockIdx.x + 1] = n2;
    }
  }
}

#ifdef OMNI3D_LAYOUT_STUDY
// The same iteration over the LDS-staged layout: a workgroup of kG
// wavefronts owns kG consecutive chunks.  Per stage, the workgroup copies the
// stage's distinct source lines (16 doubles each) into LDS with 16-byte
// loads, the next stage's lines already in flight in registers, and every
// wavefront adds its chunk's entries of the stage's directions from LDS, in
// slot order (the same fixed order on every run).
template <int kG, int kCap>
__global__ void __launch_bounds__(64 * kG)
k_iterate_staged(StagedIter si, int k, const unsigned int* __restrict__ diag,
                 const int* __restrict__ slist, const unsigned int* __restrict__ cnt,
                 const long long* __restrict__ bs, const double* __restrict__ pinit, int seed,
                 const double* __restrict__ pold, double* __restrict__ pnew,
                 const DevScalars* __restrict__ sc, double* __restrict__ partials) {
  if (sc->converged) return;
  constexpr int kU = 16;
  constexpr int kOffStage = kU * 64 + 16;
  constexpr int kThreads = 64 * kG;
  constexpr int kPer = (kCap * 8 + kThreads - 1) / kThreads;
  static_assert((kCap & (kCap - 1)) == 0, "stage capacity must be a power of two");
  __shared__ __attribute__((aligned(16))) double s_line[kCap * 16];
  __shared__ __attribute__((aligned(16))) unsigned short s_off[kG][kOffStage];
  __shared__ double s_red[2][kG];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  unsigned short* __restrict__ so = s_off[wave];
  int4* __restrict__ sl4 = reinterpret_cast<int4*>(s_line);
  const int4* __restrict__ p4 = reinterpret_cast<const int4*>(pold);
  const double inv = sc->inv_scale;
  const unsigned long long below = (1ull << lane) - 1ull;
  double d2 = 0.0, n2 = 0.0;
  const int nch = (si.nS + 63) >> 6;
  const int ngroups = (nch + kG - 1) / kG;
  for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const int chunk = g * kG + wave;
    const bool valid = chunk < nch;
    unsigned long long s = valid ? si.moff[chunk] : 0ull;
    const unsigned long long send = valid ? si.moff[chunk + 1] : 0ull;
    unsigned long long eb = s < send ? (unsigned long long)si.cell_eb[s] : 0ull;
    double acc = 0.0;
    const size_t b = (size_t)g * si.D;
    const unsigned int ns = si.nst[g];
    int4 r[kPer];
    unsigned int n4 = (si.st_hi[b] - si.st_lo[b]) * 8u;
    {
      const unsigned int lo = si.st_lo[b];
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const unsigned int i = tid + q * kThreads;
        r[q] = i < n4 ? p4[(size_t)si.ulines[lo + (i >> 3)] * 8 + (i & 7)] : int4{0, 0, 0, 0};
      }
    }
    for (unsigned int t = 0; t < ns; t++) {
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const unsigned int i = tid + q * kThreads;
        if (i < n4) sl4[i] = r[q];
      }
      __syncthreads();
      if (t + 1 < ns) {
        // the next stage's lines travel while this stage is summed
        const unsigned int lo = si.st_lo[b + t + 1];
        n4 = (si.st_hi[b + t + 1] - lo) * 8u;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
          const unsigned int i = tid + q * kThreads;
          if (i < n4) r[q] = p4[(size_t)si.ulines[lo + (i >> 3)] * 8 + (i & 7)];
        }
      }
      const unsigned long long se = valid ? (unsigned long long)si.cse[(b + t) * kG + wave] : 0ull;
      for (; s < se; s += kU) {
        unsigned long long m[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) m[u] = s + u < se ? si.masks[s + u] : 0ull;
        unsigned int off[kU];
        unsigned int n = 0;
#pragma unroll
        for (int u = 0; u < kU; u++) {
          off[u] = n;
          n += (unsigned int)__popcll(m[u]);
        }
        // the round's LDS positions, staged with 16-byte loads
        const unsigned long long a0 = eb & ~7ull;
        const unsigned int sh = (unsigned int)(eb - a0);
        const unsigned int n8 = (sh + n + 7) >> 3;
        const int4* __restrict__ src4 = reinterpret_cast<const int4*>(si.off + a0);
        for (unsigned int q = lane; q < n8; q += 64) reinterpret_cast<int4*>(so)[q] = src4[q];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double x[kU];
        bool has[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
          has[u] = (m[u] >> lane) & 1ull;
          const unsigned int rk = has[u] ? (unsigned int)__popcll(m[u] & below) : 0u;
          // (lanes without an entry may read a stale position: kept in range)
          x[u] = s_line[so[sh + off[u] + rk] & (kCap * 16 - 1)];
        }
#pragma unroll
        for (int u = 0; u < kU; u++) acc += has[u] ? x[u] : 0.0;
        eb += n;
        __builtin_amdgcn_wave_barrier();
      }
      s = se;
      __syncthreads();
    }
    const int row = chunk * 64 + lane;
    if (valid && row < si.nS) {
      const int v = slist[row];
      const unsigned int c = cnt[v];
      const double po = pold[row];
      double tot = (double)bs[row] * inv + (double)diag[row] * po + acc;
      if (k == 1 && seed && pinit) tot += pinit[v];
      const double pn = c > 0 ? tot / (double)c : po;
      pnew[row] = pn;
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    d2 += __shfl_xor(d2, off);
    n2 += __shfl_xor(n2, off);
  }
  if (lane == 0) {
    s_red[0][wave] = d2;
    s_red[1][wave] = n2;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, c = 0.0;
    for (int w = 0; w < kG; w++) {
      a += s_red[0][w];
      c += s_red[1][w];
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = c;
  }
}

#endif  // OMNI3D_LAYOUT_STUDY

// The same iteration over the dense layout: each 64-row chunk's entries
// sorted by source, one word (row in the chunk << 24 | source) each.  A
// wavefront owns a chunk (kWaves chunks per workgroup) and its lanes take
// rounds of kU * 64 consecutive entries, so one gather touches a few 128-byte
// lines.  Row sums are int64 fixed-point sums in the wavefront's LDS with the
// scale 2^(51 - e), 2^(e-1) <= max|p_S| < 2^e (every term below 2^51 in
// magnitude, exact sums in any order: the result is the same on every run).
// The iteration's own max |p_S| feeds the scale of the next one.
template <int kWaves, int kU>
__global__ void __launch_bounds__(64 * kWaves)
k_iterate_dense(int k, int nS, const unsigned long long* __restrict__ rowptr,
                const unsigned int* __restrict__ dent, const unsigned int* __restrict__ diag,
                const int* __restrict__ slist, const unsigned int* __restrict__ cnt,
                const long long* __restrict__ bs, const double* __restrict__ pinit, int seed,
                const double* __restrict__ pold, double* __restrict__ pnew,
                DevScalars* __restrict__ sc, double* __restrict__ partials) {
  if (sc->converged) return;
  static_assert(kU % 4 == 0, "rounds are whole 16-byte pieces per lane");
  __shared__ unsigned long long s_acc[kWaves][64];
  __shared__ __attribute__((aligned(16))) unsigned int s_ent[kWaves][kU * 64];
  __shared__ double s_red[2][kWaves];
  __shared__ unsigned long long s_max[kWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned int* __restrict__ se = s_ent[wave];
  unsigned long long* __restrict__ sa = s_acc[wave];
  const double inv = sc->inv_scale;
  const double B = __longlong_as_double((long long)sc->itmax_bits[(k - 1) & 1]);
  const bool finite = B <= DBL_MAX;
  int ex = 0;
  (void)frexp(finite ? B : 1.0, &ex);
  ex = ex < -960 ? -960 : ex;  // (a tiny bound: keep the scale finite)
  const double scale = ldexp(1.0, 51 - ex), iscale = ldexp(1.0, ex - 51);
  double d2 = 0.0, n2 = 0.0;
  unsigned long long pm = 0ull;
  const int nch = (nS + 63) >> 6;
  for (int chunk = blockIdx.x * kWaves + wave; chunk < nch; chunk += gridDim.x * kWaves) {
    sa[lane] = 0ull;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int r1 = chunk * 64 + 64 < nS ? chunk * 64 + 64 : nS;
    const unsigned long long e0 = rowptr[chunk * 64], e1 = rowptr[r1];
    // rounds of kU * 64 words from 4-aligned starts; each lane holds kU / 4
    // 16-byte pieces of the next round in registers while the current round
    // gathers
    const int4* __restrict__ d4 = reinterpret_cast<const int4*>(dent);
    unsigned long long base = e0 & ~3ull;
    int4 pre[kU / 4];
    {
      const unsigned long long q0 = base >> 2, q1 = (e1 + 3) >> 2;
#pragma unroll
      for (int j = 0; j < kU / 4; j++) {
        const unsigned long long q = q0 + (unsigned long long)(j * 64 + lane);
        pre[j] = q < q1 ? d4[q] : int4{0, 0, 0, 0};
      }
    }
    for (; base < e1; base += kU * 64) {
#pragma unroll
      for (int j = 0; j < kU / 4; j++) reinterpret_cast<int4*>(se)[j * 64 + lane] = pre[j];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const unsigned long long nb = base + kU * 64;
      if (nb < e1) {
        const unsigned long long q0 = nb >> 2, q1 = (e1 + 3) >> 2;
#pragma unroll
        for (int j = 0; j < kU / 4; j++) {
          const unsigned long long q = q0 + (unsigned long long)(j * 64 + lane);
          if (q < q1) pre[j] = d4[q];
        }
      }
      // words base + i with e0 <= base + i < e1 belong to this chunk
      const unsigned long long lo = e0 > base ? e0 - base : 0ull;
      const unsigned long long hi = e1 - base;
      unsigned int w[kU];
      double x[kU];
      bool ok[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const unsigned int i = u * 64 + lane;
        ok[u] = i >= lo && i < hi;
        w[u] = ok[u] ? se[i] : se[lo];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) x[u] = pold[w[u] & 0xffffffu];
#pragma unroll
      for (int u = 0; u < kU; u++)
        if (ok[u]) atomicAdd(&sa[w[u] >> 24], (unsigned long long)to_fixed(x[u], scale));
      __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int row = chunk * 64 + lane;
    if (row < nS) {
      double acc = (double)(long long)sa[lane] * iscale;
      if (!finite) acc = B - B;
      const int v = slist[row];
      const unsigned int c = cnt[v];
      const double po = pold[row];
      double tot = (double)bs[row] * inv + (double)diag[row] * po + acc;
      if (k == 1 && seed && pinit) tot += pinit[v];
      const double pn = c > 0 ? tot / (double)c : po;
      pnew[row] = pn;
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
      const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(pn));
      pm = b > pm ? b : pm;
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int off = 32; off > 0; off >>= 1) {
    d2 += __shfl_xor(d2, off);
    n2 += __shfl_xor(n2, off);
    const unsigned long long o = __shfl_xor(pm, off);
    pm = o > pm ? o : pm;
  }
  if (lane == 0) {
    s_red[0][wave] = d2;
    s_red[1][wave] = n2;
    s_max[wave] = pm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    unsigned long long m = 0ull;
    for (int q = 0; q < kWaves; q++) {
      a += s_red[0][q];
      c += s_red[1][q];
      m = s_max[q] > m ? s_max[q] : m;
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = c;
    atomicMax(&sc->itmax_bits[k & 1], m);
  }
}

constexpr int kFinishBlock = 1024;

// Relative change of iteration k and the stop rule, by one workgroup of
// kThreads (the caller has checked that no earlier iteration stopped).
template <int kThreads>
__device__ void finish_iteration(int k, int nparts, const double* __restrict__ partials,
                                 DevScalars* __restrict__ sc, int max_iter, double tol) {
  __shared__ double s_a[kThreads / 64], s_b[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kThreads) {
    a += partials[2 * i];
    b += partials[2 * i + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if (lane == 0) {
    s_a[wave] = a;
    s_b[wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < kThreads / 64; w++) {
      a += s_a[w];
      b += s_b[w];
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
  }
}

// The dense iteration over groups of kG chunks: a workgroup of kG
// wavefronts owns a group (rows [grow[g], grow[g + 1]), at most kG * 64 =
// 256 or 512, the group's words sorted by source, word (row in group << (32
// - b)) | (multiplicity - 1) << (30 - b) | source with b = 8 or 9 row bits,
// the group's words from wptr[g]), its wavefronts take the
// group's rounds of 512 words in turn and add into the group's LDS row sums;
// each wavefront then finishes the 64 rows of its chunk.  Sources shared by
// neighbouring chunks touch the same 128-byte lines in one gather.
//
// kInc selects the increment form (FP32 ray-sum precision): a row keeps
// S = sum of its entries' p_S in FP64 across iterations.  kInc == 1 (the
// first iteration) gathers p_S, stores S and the FP32 increment dnew = p_new
// - p_old; kInc == 2 gathers the previous increments din (4 bytes per entry)
// and adds their sum to S.  The fixed-point scale then follows max |increment|
// instead of max |p_S|, so the FP32 rounding applies to increments only.
// The body of one workgroup: block lb of the nb workgroups of this
// iteration (one snapshot's, in a multi-snapshot launch).
template <int kG, int kInc>
__device__ __forceinline__ void iterate_grp_body(
    const unsigned int lb, const unsigned int nb, int k, int nS,
    const unsigned int* __restrict__ wptr, const unsigned int* __restrict__ dent,
    const int* __restrict__ grow, int ngr, const unsigned int* __restrict__ diag, const int* __restrict__ slist,
    const unsigned int* __restrict__ cnt, const long long* __restrict__ bs,
    const double* __restrict__ pinit, int seed, const double* __restrict__ pold,
    double* __restrict__ pnew, double* __restrict__ srow, const float* __restrict__ din,
    float* __restrict__ dnew, DevScalars* __restrict__ sc, double* __restrict__ partials,
    int max_iter, double tol) {
  // the first group's word range and rows ([grow[g], grow[g + 1]), at most
  // kRows, chosen so that groups hold about equal words) are loaded with the
  // stop flag: independent loads in flight together
  unsigned int fw0 = 0, fw1 = 0;
  int fr0 = 0, fr1 = 0;
  if ((int)lb < ngr) {
    fw0 = wptr[lb];
    fw1 = wptr[lb + 1];
    fr0 = grow[lb];
    fr1 = grow[lb + 1];
  }
  // the stop rule of iteration k - 1 runs in workgroup 0 of this launch: the
  // launch computes iteration k before knowing it (the buffers keep p_S of
  // k - 1), and launch k + 1 returns at once if k - 1 stopped
  if (sc->converged) return;
  constexpr int kU = 8;
  constexpr int kRows = 64 * kG;
  static_assert(kRows == 256 || kRows == 512, "merged words hold 8- or 9-bit rows");
  constexpr int kRowShift = kRows == 256 ? 24 : 23;
  constexpr int kMulShift = kRowShift - 2;
  constexpr unsigned int kSrcMask = (1u << kMulShift) - 1u;
  __shared__ unsigned long long s_acc[kRows];
  __shared__ __attribute__((aligned(16))) unsigned int s_ent[kG][kU * 64];
  __shared__ double s_red[2][kG];
  __shared__ unsigned long long s_max[kG];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned int* __restrict__ se = s_ent[wave];
  const int4* __restrict__ d4 = reinterpret_cast<const int4*>(dent);
  // the first group's first words, issued before the loads of the scale
  // (and before workgroup 0's stop rule) so that their latencies overlap
  int4 pre[kU / 4];
  {
    const unsigned long long q1 = ((unsigned long long)fw1 + 3) >> 2;
    const unsigned long long base = (fw0 & ~3ull) + (unsigned long long)wave * (kU * 64);
#pragma unroll
    for (int j = 0; j < kU / 4; j++) {
      const unsigned long long q = (base >> 2) + (unsigned long long)(j * 64 + lane);
      pre[j] = q < q1 ? d4[q] : int4{0, 0, 0, 0};
    }
  }
  const double* __restrict__ prev_part = partials + (size_t)((k - 1) & 1) * 2 * nb;
  partials += (size_t)(k & 1) * 2 * nb;
  if (lb == 0) {
    if (k > 1) finish_iteration<64 * kG>(k - 1, nb, prev_part, sc, max_iter, tol);
    // the max slot launch k + 1 fills (launch k - 1 read it)
    if (threadIdx.x == 0) sc->itmax_bits[(k + 1) % 3] = 0ull;
  }
  const double inv = sc->inv_scale;
  const double B = __longlong_as_double((long long)sc->itmax_bits[(k - 1) % 3]);
  const bool finite = B <= DBL_MAX;
  int ex = 0;
  (void)frexp(finite ? B : 1.0, &ex);
  ex = ex < -960 ? -960 : ex;  // (a tiny bound: keep the scale finite)
  const double scale = ldexp(1.0, 51 - ex), iscale = ldexp(1.0, ex - 51);
  double d2 = 0.0, n2 = 0.0;
  unsigned long long pm = 0ull;
  for (int g = (int)lb; g < ngr; g += (int)nb) {
    const bool first = g == (int)lb;
    const int rbeg = first ? fr0 : grow[g];
    const int nrows = (first ? fr1 : grow[g + 1]) - rbeg;
    const unsigned long long e0 = first ? fw0 : wptr[g], e1 = first ? fw1 : wptr[g + 1];
    const unsigned long long q1 = (e1 + 3) >> 2;
    unsigned long long base = (e0 & ~3ull) + (unsigned long long)wave * (kU * 64);
    if (!first) {
#pragma unroll
      for (int j = 0; j < kU / 4; j++) {
        const unsigned long long q = (base >> 2) + (unsigned long long)(j * 64 + lane);
        pre[j] = q < q1 ? d4[q] : int4{0, 0, 0, 0};
      }
    }
    // the row sums are cleared while the first words are in flight (the
    // previous group's rows were read before the barrier that ended it)
    s_acc[threadIdx.x] = 0ull;
    __syncthreads();
    for (; base < e1; base += (unsigned long long)kG * (kU * 64)) {
#pragma unroll
      for (int j = 0; j < kU / 4; j++) reinterpret_cast<int4*>(se)[j * 64 + lane] = pre[j];
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const unsigned long long nb = base + (unsigned long long)kG * (kU * 64);
      if (nb < e1) {
#pragma unroll
        for (int j = 0; j < kU / 4; j++) {
          const unsigned long long q = (nb >> 2) + (unsigned long long)(j * 64 + lane);
          if (q < q1) pre[j] = d4[q];
        }
      }
      // this round's words i with lo <= i < hi belong to the group (lo <= 3:
      // the group's first round may start up to 3 words early; both bounds
      // fit 32 bits)
      const unsigned int lo = (unsigned int)(e0 > base ? e0 - base : 0ull);
      const unsigned long long hl = e1 - base;
      const unsigned int hi = hl < (unsigned long long)(kU * 64) ? (unsigned int)hl : kU * 64;
      unsigned int w[kU];
      double x[kU];
      bool ok[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const unsigned int i = u * 64 + lane;
        ok[u] = i >= lo && i < hi;
        w[u] = ok[u] ? se[i] : se[lo];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        if constexpr (kInc == 2) x[u] = (double)din[w[u] & kSrcMask];
        else x[u] = pold[w[u] & kSrcMask];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        // a word stands for 1..4 equal entries (the same source in one row):
        // the fixed-point value times its multiplicity (an integer product,
        // exact; scaling x first would leave the bias rounding's 2^51 range);
        // lanes outside the group's words add 0 (no branch around the add)
        const unsigned int m = ok[u] ? ((w[u] >> kMulShift) & 3u) + 1u : 0u;
        atomicAdd(&s_acc[w[u] >> kRowShift], (unsigned long long)to_fixed(x[u], scale) * m);
      }
      __builtin_amdgcn_wave_barrier();
    }
    // the row's own values do not depend on the row sums: loaded before the
    // barrier, they arrive while the group's last wavefronts finish
    const int row = rbeg + wave * 64 + lane;
    const bool own = wave * 64 + lane < nrows;
    int v = 0;
    unsigned int c = 0, dg = 0;
    double po = 0.0, sr = 0.0;
    long long bv = 0;
    if (own) {
      v = slist[row];
      c = cnt[v];
      po = pold[row];
      bv = bs[row];
      dg = diag[row];
      if constexpr (kInc == 2) sr = srow[row];
    }
    __syncthreads();
    if (own) {
      double acc = (double)(long long)s_acc[wave * 64 + lane] * iscale;
      if (!finite) acc = B - B;
      if constexpr (kInc == 2) acc += sr;
      if constexpr (kInc != 0) srow[row] = acc;
      double tot = (double)bv * inv + (double)dg * po + acc;
      if (k == 1 && seed && pinit) tot += pinit[v];
      const double pn = c > 0 ? tot / (double)c : po;
      pnew[row] = pn;
      const double df = pn - po;
      d2 += df * df;
      n2 += pn * pn;
      double mv = pn;
      if constexpr (kInc != 0) {
        const float dq = (float)df;
        dnew[row] = dq;
        mv = (double)dq;
      }
      const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(mv));
      pm = b > pm ? b : pm;
    }
    __syncthreads();
  }
  for (int off = 32; off > 0; off >>= 1) {
    d2 += __shfl_xor(d2, off);
    n2 += __shfl_xor(n2, off);
    const unsigned long long o = __shfl_xor(pm, off);
    pm = o > pm ? o : pm;
  }
  if (lane == 0) {
    s_red[0][wave] = d2;
    s_red[1][wave] = n2;
    s_max[wave] = pm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    unsigned long long m = 0ull;
    for (int q = 0; q < kG; q++) {
      a += s_red[0][q];
      c += s_red[1][q];
      m = s_max[q] > m ? s_max[q] : m;
    }
    partials[2 * lb] = a;
    partials[2 * lb + 1] = c;
    atomicMax(&sc->itmax_bits[k % 3], m);
  }
}

template <int kG, int kInc>
__global__ void __launch_bounds__(64 * kG)
k_iterate_grp(int k, int nS, const unsigned int* __restrict__ wptr,
              const unsigned int* __restrict__ dent, const int* __restrict__ grow, int ngr,
              const unsigned int* __restrict__ diag,
              const int* __restrict__ slist, const unsigned int* __restrict__ cnt,
              const long long* __restrict__ bs, const double* __restrict__ pinit, int seed,
              const double* __restrict__ pold, double* __restrict__ pnew,
              double* __restrict__ srow, const float* __restrict__ din, float* __restrict__ dnew,
              DevScalars* __restrict__ sc, double* __restrict__ partials, int max_iter,
              double tol) {
  iterate_grp_body<kG, kInc>(blockIdx.x, gridDim.x, k, nS, wptr, dent, grow, ngr, diag, slist,
                             cnt, bs, pinit, seed, pold, pnew, srow, din, dnew, sc, partials,
                             max_iter, tol);
}

// Iteration k of many snapshots in one launch (a snapshot batch): workgroup
// b works for snapshot blk_snap[b] as its block b - first of nb, on that
// snapshot's buffers (p_S rotating over three, increments over two, as in
// the one-snapshot launches), so every snapshot's result is the one its own
// launches give; a snapshot that stopped returns at once.
template <int kG, int kInc>
__global__ void __launch_bounds__(64 * kG)
k_iterate_multi(int k, const IterDesc* __restrict__ desc, const unsigned int* __restrict__ blk_snap,
                int max_iter, double tol) {
  const IterDesc& q = desc[blk_snap[blockIdx.x]];
  const unsigned int lb = blockIdx.x - (unsigned int)q.first;
  iterate_grp_body<kG, kInc>(lb, (unsigned int)q.nb, k, q.nS, q.wptr, q.dent, q.grow, q.ngr,
                             q.diag, q.slist,
                             q.cnt, q.bs, q.pinit, q.seed, q.pb[(k - 1) % 3], q.pb[k % 3],
                             q.srow, kInc == 2 ? q.db[(k - 1) & 1] : nullptr,
                             kInc != 0 ? q.db[k & 1] : nullptr, q.sc, q.partials, max_iter, tol);
}

// Relative change of iteration k from the workgroups' partial sums (fixed
// order: a strided share per thread, then a tree), and the stop rule.
__global__ void __launch_bounds__(kFinishBlock)
k_iter_finish(int k, int nparts, const double* __restrict__ partials, DevScalars* __restrict__ sc,
              int max_iter, double tol, int slots) {
  if (sc->converged) return;
  // the max |p_S| slot iteration k + 1 fills (iteration k read it)
  if (threadIdx.x == 0 && slots == 2) sc->itmax_bits[(k + 1) & 1] = 0ull;
  finish_iteration<kFinishBlock>(k, nparts, partials, sc, max_iter, tol);
}

// Sorts the (direction, source) keys of every CSR row (plan build), so that
// the iteration's FP64 row sums do not depend on the order the fill kernel's
// atomics gave.  One block per row: bitonic sort in LDS for rows up to 2048
// entries, a serial insertion sort beyond.
constexpr int kSortCap = 2048;

__global__ void __launch_bounds__(256)
k_sort_rows(int nS, const unsigned long long* __restrict__ rowptr,
            unsigned long long* __restrict__ key) {
  __shared__ unsigned long long s[kSortCap];
  for (int row = blockIdx.x; row < nS; row += gridDim.x) {
    const unsigned long long b = rowptr[row];
    const int n = (int)(rowptr[row + 1] - b);
    if (n <= 1) continue;
    if (n > kSortCap) {
      if (threadIdx.x == 0) {
        for (int i = 1; i < n; i++) {
          const unsigned long long x = key[b + i];
          int j = i - 1;
          while (j >= 0 && key[b + j] > x) {
            key[b + j + 1] = key[b + j];
            j--;
          }
          key[b + j + 1] = x;
        }
      }
      __syncthreads();
      continue;
    }
    int p = 1;
    while (p < n) p <<= 1;
    for (int i = threadIdx.x; i < p; i += blockDim.x) s[i] = i < n ? key[b + i] : ~0ull;
    __syncthreads();
    for (int kk = 2; kk <= p; kk <<= 1) {
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int i = threadIdx.x; i < p; i += blockDim.x) {
          const int l = i ^ jj;
          if (l > i) {
            const unsigned long long x = s[i], y = s[l];
            const bool up = (i & kk) == 0;
            if ((x > y) == up) {
              s[i] = y;
              s[l] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) key[b + i] = s[i];
    __syncthreads();
  }
}

// Stages three component arrays as one 4-vector per voxel.
template <typename S, typename V4>
__global__ void __launch_bounds__(256)
k_pack_field(const S* __restrict__ ax, const S* __restrict__ ay, const S* __restrict__ az,
             V4* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    V4 v;
    v.x = ax[i];
    v.y = ay[i];
    v.z = az[i];
    v.w = 0;
    out[i] = v;
  }
}

// Averaging of the final pass (devidecountInner, kernel.cu:6746-6759): where
// C > 0, p = sum / C; elsewhere p keeps its initial value.
__global__ void __launch_bounds__(256)
k_finalize(int64_t V, const long long* __restrict__ sum, int ncopies,
           const unsigned int* __restrict__ cnt, const double* __restrict__ pinit, int seed,
           const DevScalars* __restrict__ sc, double* __restrict__ pd, float* __restrict__ pf,
           double* __restrict__ wd, float* __restrict__ wf) {
  const double inv = sc->inv_scale, scale = sc->scale;
  const bool add_seed = seed && pinit && sc->kstar == 1;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const unsigned int c = cnt[v];
    const double p0 = pinit ? pinit[v] : 0.0;
    double p = p0;
    if (c > 0) {
      long long tot = 0;
      for (int q = 0; q < ncopies; q++) tot += sum[(int64_t)q * V + v];
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
  using V4 = typename Vec4<T>::type;
  if (n % 4 == 0 && reinterpret_cast<uintptr_t>(x) % alignof(V4) == 0) {
    const int64_t n4 = n / 4;
    hipLaunchKernelGGL((k_absmax4<T>), dim3(grid_for(n4, 4096)), dim3(256), 0, st,
                       reinterpret_cast<const V4*>(x), n4, bits);
  } else {
    hipLaunchKernelGGL((k_absmax<T>), dim3(grid_for(n, 1024)), dim3(256), 0, st, x, n, bits);
  }
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
  const int64_t nchunks = ((int64_t)nS + 63) / 64;
  return (unsigned int)std::max<int64_t>(1, std::min<int64_t>(nchunks, 65536));
}

hipError_t launch_iterate(int waves, int k, int nS, const unsigned long long* moff,
                          const unsigned long long* masks, const unsigned int* cell_eb,
                          const int* ent, const unsigned int* diag, const int* slist,
                          const unsigned int* cnt, const long long* bs, const double* pinit,
                          int seed, const double* pold, double* pnew, DevScalars* sc,
                          double* partials, int max_iter, double tol, hipStream_t st) {
  const unsigned int grid = iterate_grid(nS);
#define OMNI_ITER(W, M)                                                                        \
  hipLaunchKernelGGL((k_iterate<W, M>), dim3(grid), dim3(64 * W), 0, st, k, nS, moff, masks,     \
                     cell_eb, ent, diag, slist, cnt, bs, pinit, seed, pold, pnew, sc, partials)
#ifdef OMNI3D_LAYOUT_STUDY
  const bool masked = waves < 0;
  if (waves < 0) waves = -waves;
  if (masked) {
    if (waves <= 1) OMNI_ITER(1, true);
    else if (waves == 2) OMNI_ITER(2, true);
    else OMNI_ITER(4, true);
  } else {
    if (waves <= 1) OMNI_ITER(1, false);
    else if (waves == 2) OMNI_ITER(2, false);
    else OMNI_ITER(4, false);
  }
#else
  (void)waves;
  OMNI_ITER(2, false);
#endif
#undef OMNI_ITER
  hipLaunchKernelGGL(k_iter_finish, dim3(1), dim3(kFinishBlock), 0, st, k, (int)grid, partials, sc,
                     max_iter, tol, 2);
  return hipGetLastError();
}

// p_S of iteration k* - 1 out of the three rotating buffers of the grouped
// iteration (iteration k writes buffer k % 3).
__global__ void k_select_ps(int nS, const double* __restrict__ b0, const double* __restrict__ b1,
                            const double* __restrict__ b2, const DevScalars* __restrict__ sc,
                            double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nS) return;
  const int q = (sc->kstar - 1) % 3;
  out[i] = q == 0 ? b0[i] : (q == 1 ? b1[i] : b2[i]);
}

hipError_t launch_select_ps(int nS, const double* b0, const double* b1, const double* b2,
                            const DevScalars* sc, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_select_ps, dim3(grid_for(nS)), dim3(256), 0, st, nS, b0, b1, b2, sc, out);
  return hipGetLastError();
}

hipError_t launch_iter_finish2(int k, int nparts, const double* partials, DevScalars* sc,
                               int max_iter, double tol, hipStream_t st) {
  hipLaunchKernelGGL(k_iter_finish, dim3(1), dim3(kFinishBlock), 0, st, k, nparts, partials, sc,
                     max_iter, tol, 2);
  return hipGetLastError();
}

hipError_t launch_iter_finish(int k, int nparts, const double* partials, DevScalars* sc,
                              int max_iter, double tol, hipStream_t st) {
  hipLaunchKernelGGL(k_iter_finish, dim3(1), dim3(kFinishBlock), 0, st, k, nparts, partials, sc,
                     max_iter, tol, 3);
  return hipGetLastError();
}

unsigned int dense_grid(int nS) {
  // one chunk per wavefront, up to 4 wavefronts per workgroup
  const int nch = (nS + 63) / 64;
  return (unsigned int)std::max(1, std::min(nch, 65536));
}

hipError_t launch_iterate_dense(int gchunks, int variant, int k, int nS,
                                const unsigned long long* rowptr, const unsigned int* wptr,
                                const unsigned int* dent, const unsigned int* diag,
                                const int* slist, const unsigned int* cnt, const long long* bs,
                                const double* pinit, int seed, const double* pold, double* pnew,
                                DevScalars* sc, double* partials, int max_iter, double tol,
                                hipStream_t st, double* srow, const float* din, float* dnew,
                                const int* grow, int ngroups) {
  unsigned int grid = dense_grid(nS);
  if (gchunks > 1) {
    // groups carry their first rows (grow[ngroups + 1])
    if (!grow || ngroups < 1) return hipErrorInvalidValue;
    const int ngr = ngroups;
    grid = grouped_grid_n(ngr);
    const int inc = srow ? (k == 1 ? 1 : 2) : 0;
#define OMNI_GRP(G, I)                                                                         \
  hipLaunchKernelGGL((k_iterate_grp<G, I>), dim3(grid), dim3(64 * G), 0, st, k, nS, wptr, dent,  \
                     grow, ngr, diag, slist, cnt, bs, pinit, seed, pold, pnew, srow, din, dnew,   \
                     sc, partials, max_iter, tol)
    if (gchunks == 8) {
      if (inc == 1) OMNI_GRP(8, 1);
      else if (inc == 2) OMNI_GRP(8, 2);
      else OMNI_GRP(8, 0);
    } else {
      if (inc == 1) OMNI_GRP(4, 1);
      else if (inc == 2) OMNI_GRP(4, 2);
      else OMNI_GRP(4, 0);
    }
#undef OMNI_GRP
    return hipGetLastError();
  }
#define OMNI_DENSE(W, U)                                                                       \
  hipLaunchKernelGGL((k_iterate_dense<W, U>), dim3(grid), dim3(64 * W), 0, st, k, nS, rowptr,    \
                     dent, diag, slist, cnt, bs, pinit, seed, pold, pnew, sc, partials)
  const int nch = (nS + 63) / 64;
#ifdef OMNI3D_LAYOUT_STUDY
  const int w = variant == 1 || variant == 2 ? 8 : 4;
  grid = (unsigned int)std::max(1, std::min((nch + w - 1) / w, 65536));
  switch (variant) {
    case 1: OMNI_DENSE(8, 8); break;
    case 2: OMNI_DENSE(8, 4); break;
    case 3: OMNI_DENSE(4, 4); break;
    case 4: OMNI_DENSE(4, 16); break;
    default: OMNI_DENSE(4, 8); break;
  }
#else
  (void)variant;
  grid = (unsigned int)std::max(1, std::min((nch + 3) / 4, 65536));
  OMNI_DENSE(4, 8);
#endif
#undef OMNI_DENSE
  hipLaunchKernelGGL(k_iter_finish, dim3(1), dim3(kFinishBlock), 0, st, k, (int)grid, partials, sc,
                     max_iter, tol, 2);
  return hipGetLastError();
}

hipError_t launch_iterate_multi(int gchunks, int inc, int k, const IterDesc* desc,
                                const unsigned int* blk_snap, unsigned int nblocks, int max_iter,
                                double tol, hipStream_t st) {
  const int v = inc ? (k == 1 ? 1 : 2) : 0;
#define OMNI_MULTI(G, I)                                                                      \
  hipLaunchKernelGGL((k_iterate_multi<G, I>), dim3(nblocks), dim3(64 * G), 0, st, k, desc,       \
                     blk_snap, max_iter, tol)
  if (gchunks == 8) {
    if (v == 1) OMNI_MULTI(8, 1);
    else if (v == 2) OMNI_MULTI(8, 2);
    else OMNI_MULTI(8, 0);
  } else {
    if (v == 1) OMNI_MULTI(4, 1);
    else if (v == 2) OMNI_MULTI(4, 2);
    else OMNI_MULTI(4, 0);
  }
#undef OMNI_MULTI
  return hipGetLastError();
}

#ifdef OMNI3D_LAYOUT_STUDY
unsigned int staged_grid(int nS) {
  const int nch = (nS + 63) / 64;
  const int ng = (nch + kStageGroup - 1) / kStageGroup;
  return (unsigned int)std::max(1, std::min(ng, 65536));
}

hipError_t launch_iterate_staged(const StagedIter& si, int k, const unsigned int* diag,
                                 const int* slist, const unsigned int* cnt, const long long* bs,
                                 const double* pinit, int seed, const double* pold, double* pnew,
                                 DevScalars* sc, double* partials, int max_iter, double tol,
                                 hipStream_t st) {
  const unsigned int grid = staged_grid(si.nS);
  hipLaunchKernelGGL((k_iterate_staged<kStageGroup, kStageCap>), dim3(grid), dim3(64 * kStageGroup),
                     0, st, si, k, diag, slist, cnt, bs, pinit, seed, pold, pnew, sc, partials);
  hipLaunchKernelGGL(k_iter_finish, dim3(1), dim3(kFinishBlock), 0, st, k, (int)grid, partials, sc,
                     max_iter, tol, 2);
  return hipGetLastError();
}
#endif  // OMNI3D_LAYOUT_STUDY

hipError_t launch_sort_rows(int nS, const unsigned long long* rowptr, unsigned long long* key,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_sort_rows, dim3(grid_for((int64_t)nS * 256, 16384)), dim3(256), 0, st, nS,
                     rowptr, key);
  return hipGetLastError();
}

template <typename S, typename T>
hipError_t launch_pack_field(const S* ax, const S* ay, const S* az, void* out, int64_t n,
                             hipStream_t st) {
  using V4 = typename Vec4<T>::type;
  hipLaunchKernelGGL((k_pack_field<S, V4>), dim3(grid_for(n)), dim3(256), 0, st, ax, ay, az,
                     static_cast<V4*>(out), n);
  return hipGetLastError();
}
template hipError_t launch_pack_field<float, float>(const float*, const float*, const float*,
                                                    void*, int64_t, hipStream_t);
template hipError_t launch_pack_field<double, float>(const double*, const double*, const double*,
                                                     void*, int64_t, hipStream_t);
template hipError_t launch_pack_field<float, double>(const float*, const float*, const float*,
                                                     void*, int64_t, hipStream_t);
template hipError_t launch_pack_field<double, double>(const double*, const double*,
                                                      const double*, void*, int64_t, hipStream_t);

hipError_t launch_finalize(int64_t V, const long long* sum, int ncopies, const unsigned int* cnt,
                           const double* pinit, int seed, const DevScalars* sc, double* pd,
                           float* pf, double* wd, float* wf, hipStream_t st) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(V)), dim3(256), 0, st, V, sum, ncopies, cnt, pinit,
                     seed, sc, pd, pf, wd, wf);
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
