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

#include <cfloat>

#include "kernels.hpp"
#include "launch.hpp"
#include "tiles.hpp"
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

// ---------------------------------------------------------------------------
// Source-restricted backward of the plain method.  A^T w is nonzero only on
// the source voxels S after its first application, so with
//   t = P^T (y / C) restricted to S  (one gather pass),
//   h_1 = t,  h_m = t + A_S^T h_{m-1}  (m = 2 .. K-1, vectors over S),
// the Horner sum is z = y + E^T h_{K-1} (E^T places h at its voxels).  A_S^T
// is the transpose of the source rows of the forward iteration: row s' lists,
// for every ray entering at s', the source voxels its later visits land on,
// and (A_S^T h)[s'] = diag[s'] w[s'] + sum_{s in row s'} w[s], w = h / C.

// the ray's entry voxel, whose source id names its transposed row
__device__ __forceinline__ int ray_source(const RayRec& rr, const GridDesc& g,
                                          const int* __restrict__ sid) {
  int c[3];
  unpack_voxel(rr.in, c);
  return sid[(int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y];
}

struct TCountVisitor {
  const int* sid;
  unsigned int n;
  __device__ __forceinline__ void entry(int64_t) {}
  __device__ __forceinline__ void visit(int64_t v, bool surface) {
    if (surface && sid[v] >= 0) n++;
  }
};

__global__ void __launch_bounds__(256)
k_adj_tcount(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
             GridDesc g, const int* __restrict__ sid, unsigned int* __restrict__ tcnt) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  TCountVisitor vis{sid, 0u};
  walk_geometry(rays[r], dirs, g, vis);
  if (vis.n) atomicAdd(&tcnt[ray_source(rays[r], g, sid)], vis.n);
}

struct TFillVisitor {
  const int* sid;
  unsigned long long* key;
  unsigned long long base;
  unsigned int i;
  __device__ __forceinline__ void entry(int64_t) {}
  __device__ __forceinline__ void visit(int64_t v, bool surface) {
    if (!surface) return;
    const int s = sid[v];
    if (s >= 0) key[base + i++] = (unsigned long long)(unsigned int)s;
  }
};

// Each ray reserves its visits' run of its row (rows are sorted afterwards,
// so the order the rays reserve in does not matter).
__global__ void __launch_bounds__(256)
k_adj_tfill(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
            GridDesc g, const int* __restrict__ sid, const unsigned long long* __restrict__ tptr,
            unsigned int* __restrict__ cursor, unsigned long long* __restrict__ key) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  TCountVisitor cv{sid, 0u};
  walk_geometry(rays[r], dirs, g, cv);
  if (cv.n == 0) return;
  const int row = ray_source(rays[r], g, sid);
  const unsigned int pos = atomicAdd(&cursor[row], cv.n);
  TFillVisitor fv{sid, key, tptr[row] + pos, 0u};
  walk_geometry(rays[r], dirs, g, fv);
}

// Voxel-order compaction of 0/1 flags: the counts to scan, then sid / slist
// from the scanned offsets (sid = -1 where unflagged).
__global__ void __launch_bounds__(256)
k_flag_counts(const uint8_t* __restrict__ flag, int64_t V, unsigned int* __restrict__ cnt) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x)
    cnt[v] = flag[v] ? 1u : 0u;
}

__global__ void __launch_bounds__(256)
k_flag_ids(const uint8_t* __restrict__ flag, int64_t V, const unsigned long long* __restrict__ off,
           int* __restrict__ sid, int* __restrict__ slist) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (flag[v]) {
      const int s = (int)off[v];
      sid[v] = s;
      slist[s] = (int)v;
    } else {
      sid[v] = -1;
    }
  }
}

__global__ void __launch_bounds__(256)
k_adj_narrow(const unsigned long long* __restrict__ key, int64_t n, unsigned int* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (unsigned int)key[i];
}

// t_S = t at the source voxels; h_1 = t_S, w_1 = h_1 / C (0 where C = 0), and
// C per source.
__global__ void __launch_bounds__(256)
k_adj_src_init(int nS, const int* __restrict__ slist, const double* __restrict__ t,
               const unsigned int* __restrict__ cnt, double* __restrict__ ts,
               double* __restrict__ h, double* __restrict__ w, double* __restrict__ cs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nS) return;
  const int64_t v = slist[s];
  const double c = (double)cnt[v];
  const double x = t[v];
  ts[s] = x;
  h[s] = x;
  // a source no visit updates keeps p = 0 and passes nothing on
  w[s] = c > 0.0 ? x / c : 0.0;
  cs[s] = c;
}

// One wavefront per transposed row: h' = t + diag w + sum of w over the row,
// w' = h' / C.  The lane split and the reduction tree are fixed, so the sums
// are bitwise repeatable.
__global__ void __launch_bounds__(256)
k_adj_src_iter(int nS, const unsigned long long* __restrict__ tptr,
               const unsigned int* __restrict__ tent, const unsigned int* __restrict__ diag,
               const double* __restrict__ cs, const double* __restrict__ ts,
               const double* __restrict__ w, double* __restrict__ hn, double* __restrict__ wn) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= nS) return;
  const unsigned long long e1 = tptr[r + 1];
  double acc = 0.0;
  for (unsigned long long e = tptr[r] + lane; e < e1; e += 64) acc += w[tent[e]];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) {
    const double h = ts[r] + (double)diag[r] * w[r] + acc;
    hn[r] = h;
    wn[r] = cs[r] > 0.0 ? h / cs[r] : 0.0;
  }
}

// z = y + E^T h (source voxels are distinct, so plain stores); with cs, the
// vector given is w = h / C and h = w C.
__global__ void __launch_bounds__(256)
k_adj_src_out(int nS, const int* __restrict__ slist, const double* __restrict__ h,
              const double* __restrict__ cs, double* __restrict__ z) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nS) return;
  z[slist[s]] += cs ? h[s] * cs[s] : h[s];
}

// t_S -> the fixed-point source term of the grouped iteration: bs = round(t s)
// with s = 2^(51 - e), |t| < 2^e from max |t| (in pmax_bits), inv_scale = 1 / s.
__global__ void __launch_bounds__(256)
k_adj_fixed(int nS, const double* __restrict__ t, DevScalars* __restrict__ sc,
            long long* __restrict__ bs) {
  const double B = __longlong_as_double((long long)sc->pmax_bits);
  int ex = 0;
  (void)frexp(B > 0.0 && B <= DBL_MAX ? B : 1.0, &ex);
  ex = ex < -960 ? -960 : ex;
  const double scale = ldexp(1.0, 51 - ex);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) sc->inv_scale = ldexp(1.0, ex - 51);
  if (s < nS) bs[s] = to_fixed(t[s], scale);
}

// The plain-mask rule's version: its runs start at valid voxels after a
// masked one (anywhere in the volume, not only at ray entries), have no
// entry self-visit, and the sources S are the run-start voxels.

// marks every run's first voxel
__global__ void __launch_bounds__(256)
k_adj_mask_starts(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
                  GridDesc g, const double* __restrict__ mask, uint8_t* __restrict__ flag) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  SegWalker<true> sw;
  sw.w.init(rays[r], dirs, g);
  sw.mask = mask;
  bool first = true;
  while (!sw.done && sw.next_source(first)) {
    first = false;
    flag[sw.w.v] = 1;
    int a, s;
    int64_t l;
    while (sw.step(&a, &s, &l)) {
    }
  }
}

// pass 0: counts each run's visits to run-start voxels into its start's row;
// pass 1: writes them (the run is walked twice from a saved walker state)
template <int kPass>
__global__ void __launch_bounds__(256)
k_adj_mask_rows(const RayRec* __restrict__ rays, int64_t nrays, const DirConst* __restrict__ dirs,
                GridDesc g, const double* __restrict__ mask, const int* __restrict__ sid,
                const unsigned long long* __restrict__ tptr, unsigned int* __restrict__ cnt,
                unsigned long long* __restrict__ key) {
  const int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  SegWalker<true> sw;
  sw.w.init(rays[r], dirs, g);
  sw.mask = mask;
  bool first = true;
  while (!sw.done && sw.next_source(first)) {
    first = false;
    const int row = sid[sw.w.v];
    const Walker w0 = sw.w;
    unsigned int n = 0;
    int a, s;
    int64_t l;
    while (sw.step(&a, &s, &l))
      if (sid[sw.w.v] >= 0) n++;
    if (n == 0) continue;
    if (kPass == 0) {
      atomicAdd(&cnt[row], n);
      continue;
    }
    const Walker wend = sw.w;
    const bool end_done = sw.done;
    const unsigned long long base = tptr[row] + atomicAdd(&cnt[row], n);
    sw.w = w0;
    sw.done = false;
    unsigned int i = 0;
    while (i < n && sw.step(&a, &s, &l)) {
      const int q = sid[sw.w.v];
      if (q >= 0) key[base + i++] = (unsigned long long)(unsigned int)q;
    }
    sw.w = wend;
    sw.done = end_done;
  }
}

hipError_t launch_adj_mask_starts(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                  GridDesc g, const double* mask, uint8_t* flag, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_mask_starts, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays,
                     nrays, dirs, g, mask, flag);
  return hipGetLastError();
}

hipError_t launch_adj_mask_rows(int pass, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                GridDesc g, const double* mask, const int* sid,
                                const unsigned long long* tptr, unsigned int* cnt,
                                unsigned long long* key, hipStream_t st) {
  const unsigned int grid = grid_for(nrays, 1u << 30);
  if (pass == 0)
    hipLaunchKernelGGL(k_adj_mask_rows<0>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       mask, sid, tptr, cnt, key);
  else
    hipLaunchKernelGGL(k_adj_mask_rows<1>, dim3(grid), dim3(256), 0, st, rays, nrays, dirs, g,
                       mask, sid, tptr, cnt, key);
  return hipGetLastError();
}

hipError_t launch_adj_transpose_rows(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                     GridDesc g, const int* sid, unsigned int* tcnt,
                                     hipStream_t st) {
  hipLaunchKernelGGL(k_adj_tcount, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, sid, tcnt);
  return hipGetLastError();
}

hipError_t launch_adj_transpose_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                     GridDesc g, const int* sid, const unsigned long long* tptr,
                                     unsigned int* cursor, unsigned long long* key,
                                     hipStream_t st) {
  hipLaunchKernelGGL(k_adj_tfill, dim3(grid_for(nrays, 1u << 30)), dim3(256), 0, st, rays, nrays,
                     dirs, g, sid, tptr, cursor, key);
  return hipGetLastError();
}

hipError_t launch_adj_narrow(const unsigned long long* key, int64_t n, unsigned int* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_adj_narrow, dim3(grid_for(n)), dim3(256), 0, st, key, n, out);
  return hipGetLastError();
}

hipError_t launch_flag_counts(const uint8_t* flag, int64_t V, unsigned int* cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_flag_counts, dim3(grid_for(V)), dim3(256), 0, st, flag, V, cnt);
  return hipGetLastError();
}

hipError_t launch_flag_ids(const uint8_t* flag, int64_t V, const unsigned long long* off, int* sid,
                           int* slist, hipStream_t st) {
  hipLaunchKernelGGL(k_flag_ids, dim3(grid_for(V)), dim3(256), 0, st, flag, V, off, sid, slist);
  return hipGetLastError();
}

hipError_t launch_adj_src_init(int nS, const int* slist, const double* t, const unsigned int* cnt,
                               double* ts, double* h, double* w, double* cs, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_src_init, dim3((nS + 255) / 256), dim3(256), 0, st, nS, slist, t, cnt,
                     ts, h, w, cs);
  return hipGetLastError();
}

hipError_t launch_adj_src_iter(int nS, const unsigned long long* tptr, const unsigned int* tent,
                               const unsigned int* diag, const double* cs, const double* ts,
                               const double* w, double* hn, double* wn, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_src_iter, dim3((nS + 3) / 4), dim3(256), 0, st, nS, tptr, tent, diag,
                     cs, ts, w, hn, wn);
  return hipGetLastError();
}

hipError_t launch_adj_src_out(int nS, const int* slist, const double* h, const double* cs,
                              double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_adj_src_out, dim3((nS + 255) / 256), dim3(256), 0, st, nS, slist, h, cs, z);
  return hipGetLastError();
}

hipError_t launch_adj_fixed(int nS, const double* t, DevScalars* sc, long long* bs,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_adj_fixed, dim3((nS + 255) / 256), dim3(256), 0, st, nS, t, sc, bs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The plain method's transposed traversals on the tile engine.  With u a
// volume (z / C), step tau of a ray (from voxel l to n along axis a)
// scatters h_a * S_tau into dL/dg_a at l and n, S_tau = the sum of u over the
// ray's visits from n on; the gather adds the sum of u over all of a ray's
// visits (its entry included) into its source.  Per segment (ray x tile):
//   k_adj_seg_u      the sum U of u over the segment's visits (u in LDS);
//   k_adj_rev_scan   per ray, T = the sum of U over the later segments (in
//                    place), and the ray's total into its source voxel;
//   k_adj_tile_scatter  per tile, S of every step from U, T and the running
//                    prefix, added into LDS accumulators of the three
//                    components (the step into a tile counts at its first
//                    voxel here, at its last voxel in the tile it leaves).

template <int kBits>
__global__ void __launch_bounds__(1024)
k_adj_seg_u(const unsigned long long* __restrict__ tile_ptr1, const SegRec* __restrict__ segs,
            GridDesc g, TileGrid tg, const double* __restrict__ u, double* __restrict__ useg) {
  constexpr int kTile = 1 << kBits, kTileMask = kTile - 1, kVox = kTile * kTile * kTile;
  __shared__ double U[kVox];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    U[lds_slot<kBits>(l)] = (i < g.X && j < g.Y && k < g.Z) ? u[i + j * sy + k * sz] : 0.0;
  }
  __syncthreads();
  const unsigned long long b = tile_ptr1[tile], e = tile_ptr1[tile + 1];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const int d0 = (sg & 1u) ? 1 : -1, d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    double sum = U[lds_slot<kBits>(lv)];
    for (unsigned int q = 1; q < n; q++) {
      const unsigned int c = q - 1;
      const uint32_t w = c < 16 ? sr.code[0] : (c < 32 ? sr.code[1] : sr.code[2]);
      const int a = (int)((w >> (2 * (c & 15))) & 3u);
      lv += a == 0 ? d0 : (a == 1 ? d1 : d2);
      sum += U[lds_slot<kBits>(lv)];
    }
    useg[sr.g] = sum;
  }
}

__global__ void __launch_bounds__(256)
k_adj_rev_scan(const RayRec* __restrict__ rays, int64_t nrays, GridDesc g,
               const unsigned int* __restrict__ nseg, const unsigned long long* __restrict__ val_off,
               double* __restrict__ useg, double* __restrict__ tvol) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const unsigned long long b = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  double t = 0.0;
  for (unsigned int j = n; j-- > 0;) {
    const double x = useg[b + 64ull * j];
    useg[b + 64ull * j] = t;
    t += x;
  }
  if (tvol) {
    int c[3];
    unpack_voxel(rays[r].in, c);
    unsafeAtomicAdd(&tvol[(int64_t)c[0] + (int64_t)c[1] * g.X + (int64_t)c[2] * g.X * g.Y], t);
  }
}

template <int kBits>
__global__ void __launch_bounds__(1024)
k_adj_tile_scatter(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
                   GridDesc g, TileGrid tg, WalkParams wp, const double* __restrict__ u,
                   const double* __restrict__ tafter, double* __restrict__ dg) {
  constexpr int kTile = 1 << kBits, kTileMask = kTile - 1, kVox = kTile * kTile * kTile;
  __shared__ double U[kVox];
  __shared__ double A[3][kVox];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    const int sl = lds_slot<kBits>(l);
    U[sl] = (i < g.X && j < g.Y && k < g.Z) ? u[i + j * sy + k * sz] : 0.0;
    A[0][sl] = A[1][sl] = A[2][sl] = 0.0;
  }
  __syncthreads();
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const int lv0 = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool first = (sr.info >> 21) & 1u, last = (sr.info >> 22) & 1u;
    const bool up[3] = {(sg & 1u) != 0, (sg & 2u) != 0, (sg & 4u) != 0};
    const int d0 = up[0] ? 1 : -1, d1 = up[1] ? kTile : -kTile;
    const int d2 = up[2] ? kTile * kTile : -kTile * kTile;
    auto code = [&](unsigned int c) {
      const uint32_t w = c < 16 ? sr.code[0] : (c < 32 ? sr.code[1] : sr.code[2]);
      return (int)((w >> (2 * (c & 15))) & 3u);
    };
    auto hfac = [&](int a) {
      return a == 0 ? step_factor<double>(wp, 0, up[0])
                    : (a == 1 ? step_factor<double>(wp, 1, up[1]) : step_factor<double>(wp, 2, up[2]));
    };
    const double T = tafter[sr.g];
    // the segment's sum of u
    double Useg = 0.0;
    {
      int lv = lv0;
      Useg = U[lds_slot<kBits>(lv)];
      for (unsigned int q = 1; q < n; q++) {
        const int a = code(q - 1);
        lv += a == 0 ? d0 : (a == 1 ? d1 : d2);
        Useg += U[lds_slot<kBits>(lv)];
      }
    }
    int lv = lv0;
    int sl = lds_slot<kBits>(lv);
    if (!first) {
      // the step into the tile: S = every visit from this segment's first on
      const int ae = (int)((sr.info >> 23) & 3u);
      atomicAdd(&A[ae][sl], hfac(ae) * (Useg + T));
    }
    double P = U[sl];  // u of the visits before the next step's target
    for (unsigned int q = 1; q < n; q++) {
      const int a = code(q - 1);
      const int ln = lv + (a == 0 ? d0 : (a == 1 ? d1 : d2));
      const int sn = lds_slot<kBits>(ln);
      const double hS = hfac(a) * ((Useg - P) + T);
      atomicAdd(&A[a][sl], hS);
      atomicAdd(&A[a][sn], hS);
      P += U[sn];
      lv = ln;
      sl = sn;
    }
    if (!last) {
      // the step out of the tile: S = the later segments' visits
      const int ax = code(n - 1);
      atomicAdd(&A[ax][sl], hfac(ax) * T);
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    if (i >= g.X || j >= g.Y || k >= g.Z) continue;
    const int64_t v = i + j * sy + k * sz;
    const int sl = lds_slot<kBits>(l);
    dg[v] = A[0][sl];
    dg[V + v] = A[1][sl];
    dg[2 * V + v] = A[2][sl];
  }
}

hipError_t launch_adj_tiles(int what, int ntiles, const unsigned long long* tile_ptr1,
                            const SegRec* segs_slot, const unsigned long long* tile_ptr,
                            const SegRec* segs, const RayRec* rays, int64_t nrays,
                            const unsigned int* nseg, const unsigned long long* val_off,
                            GridDesc g, TileGrid tg, const WalkParams& wp, const double* u,
                            double* useg, double* tvol, double* dg, hipStream_t st) {
  // what 0: segment sums of u; 1: reverse scan (+ ray totals into tvol);
  // 2: the scatter into dg
  if (what == 1) {
    hipLaunchKernelGGL(k_adj_rev_scan, dim3((unsigned int)((nrays + 255) / 256)), dim3(256), 0, st,
                       rays, nrays, g, nseg, val_off, useg, tvol);
    return hipGetLastError();
  }
  if (what == 0) {
    if (tg.bits == 3)
      hipLaunchKernelGGL(k_adj_seg_u<3>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs_slot, g,
                         tg, u, useg);
    else
      hipLaunchKernelGGL(k_adj_seg_u<4>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs_slot, g,
                         tg, u, useg);
    return hipGetLastError();
  }
  if (tg.bits == 3)
    hipLaunchKernelGGL(k_adj_tile_scatter<3>, dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, g,
                       tg, wp, u, useg, dg);
  else
    hipLaunchKernelGGL(k_adj_tile_scatter<4>, dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, g,
                       tg, wp, u, useg, dg);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The plain-mask rule's transposed traversals on the tile engine (the rule of
// SegWalker<true>: a run starts at a valid voxel -- mask != 0 -- that is the
// ray's entry or follows a masked voxel, and its later valid voxels are its
// adding visits, up to the next masked one).  A visit v adds iff v and the
// voxel before it are valid, so a step into an adding visit scatters h_a S
// with S = the sum of u over the run's adding visits from its target on, and
// a run's start gathers the sum of u over all of the run's adding visits.
// Per segment (the voxel before its first visit read from the mask):
//   k_adjm_seg_u     Upre = the sum of u over the leading adding visits (the
//                    run the segment enters with), and whether that run
//                    spans the whole segment;
//   k_adjm_rev_scan  per ray, T = what the later segments add to the run the
//                    segment ends in (Upre of the next segment, plus its T if
//                    the run spans it), in place;
//   k_adjm_tile      per tile, each segment walked backwards from T: S of
//                    every adding step into LDS accumulators of the three
//                    components (kScatter), or each run's total into its
//                    start (the gather); complete values stored per voxel.

// validity of the voxel before a segment's first visit (none for a ray's
// first segment): the first voxel minus the step into the tile
template <int kBits>
__device__ __forceinline__ bool adjm_prev_valid(const SegRec& sr, int ox, int oy, int oz, int64_t sy,
                                                int64_t sz, const double* __restrict__ mask) {
  constexpr int kTileMask = (1 << kBits) - 1;
  if ((sr.info >> 21) & 1u) return false;
  const int lv0 = (int)(sr.info & ((1u << kSegLvBits) - 1u));
  const unsigned int sg = (sr.info >> 18) & 7u;
  const int ae = (int)((sr.info >> 23) & 3u);
  const int se = ((sg >> ae) & 1u) ? 1 : -1;
  const int64_t vf = (int64_t)(ox + (lv0 & kTileMask)) + (int64_t)(oy + ((lv0 >> kBits) & kTileMask)) * sy +
                     (int64_t)(oz + (lv0 >> (2 * kBits))) * sz;
  const int64_t vp = vf - (ae == 0 ? (int64_t)se : (ae == 1 ? se * sy : se * sz));
  return mask[vp] != 0.0;
}

template <int kBits>
__global__ void __launch_bounds__(1024)
k_adjm_seg_u(const unsigned long long* __restrict__ tile_ptr1, const SegRec* __restrict__ segs,
             GridDesc g, TileGrid tg, const double* __restrict__ u,
             const double* __restrict__ mask, double* __restrict__ useg,
             uint8_t* __restrict__ span) {
  constexpr int kTile = 1 << kBits, kTileMask = kTile - 1, kVox = kTile * kTile * kTile;
  __shared__ double U[kVox];
  __shared__ uint8_t M[kVox];
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    const bool in = i < g.X && j < g.Y && k < g.Z;
    const int64_t v = i + j * sy + k * sz;
    U[lds_slot<kBits>(l)] = in ? u[v] : 0.0;
    M[lds_slot<kBits>(l)] = in && mask[v] != 0.0 ? 1 : 0;
  }
  __syncthreads();
  const unsigned long long b = tile_ptr1[tile], e = tile_ptr1[tile + 1];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    int lv = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const int d0 = (sg & 1u) ? 1 : -1, d1 = (sg & 2u) ? kTile : -kTile;
    const int d2 = (sg & 4u) ? kTile * kTile : -kTile * kTile;
    // the entering run continues through a visit while the visits are valid
    bool alive = M[lds_slot<kBits>(lv)] && adjm_prev_valid<kBits>(sr, ox, oy, oz, sy, sz, mask);
    double pre = alive ? U[lds_slot<kBits>(lv)] : 0.0;
    for (unsigned int q = 1; q < n && alive; q++) {
      const unsigned int c = q - 1;
      const uint32_t w = c < 16 ? sr.code[0] : (c < 32 ? sr.code[1] : sr.code[2]);
      const int a = (int)((w >> (2 * (c & 15))) & 3u);
      lv += a == 0 ? d0 : (a == 1 ? d1 : d2);
      const int sl = lds_slot<kBits>(lv);
      if (M[sl]) pre += U[sl];
      else alive = false;
    }
    useg[sr.g] = pre;
    span[sr.g] = alive ? 1 : 0;
  }
}

__global__ void __launch_bounds__(256)
k_adjm_rev_scan(int64_t nrays, const unsigned int* __restrict__ nseg,
                const unsigned long long* __restrict__ val_off, double* __restrict__ useg,
                const uint8_t* __restrict__ span) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  const unsigned long long b = val_off[r >> 6] + (unsigned long long)(r & 63);
  const unsigned int n = nseg[r];
  double t = 0.0;
  for (unsigned int j = n; j-- > 0;) {
    const unsigned long long o = b + 64ull * j;
    const double x = useg[o];
    const bool whole = span[o] != 0;
    useg[o] = t;
    t = x + (whole ? t : 0.0);
  }
}

template <int kBits, bool kScatter>
__global__ void __launch_bounds__(1024)
k_adjm_tile(const unsigned long long* __restrict__ tile_ptr, const SegRec* __restrict__ segs,
            GridDesc g, TileGrid tg, WalkParams wp, const double* __restrict__ u,
            const double* __restrict__ mask, const double* __restrict__ tafter,
            double* __restrict__ out) {
  constexpr int kTile = 1 << kBits, kTileMask = kTile - 1, kVox = kTile * kTile * kTile;
  __shared__ double U[kVox];
  __shared__ uint8_t M[kVox];
  __shared__ double A[kScatter ? 3 : 1][kVox];  // dg components, or the gathered totals
  const int tile = (int)xcd_block();
  const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
  const int ox = tx << kBits, oy = ty << kBits, oz = tz << kBits;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  const int64_t V = (int64_t)g.X * g.Y * g.Z;
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    const bool in = i < g.X && j < g.Y && k < g.Z;
    const int64_t v = i + j * sy + k * sz;
    const int sl = lds_slot<kBits>(l);
    U[sl] = in ? u[v] : 0.0;
    M[sl] = in && mask[v] != 0.0 ? 1 : 0;
#pragma unroll
    for (int c = 0; c < (kScatter ? 3 : 1); c++) A[c][sl] = 0.0;
  }
  __syncthreads();
  const unsigned long long b = tile_ptr[tile * kLenBuckets], e = tile_ptr[(tile + 1) * kLenBuckets];
  for (unsigned long long jj = b + threadIdx.x; jj < e; jj += blockDim.x) {
    const SegRec sr = segs[jj];
    const int lv0 = (int)(sr.info & ((1u << kSegLvBits) - 1u));
    const unsigned int n = (sr.info >> kSegLvBits) & 63u;
    const unsigned int sg = (sr.info >> 18) & 7u;
    const bool last = (sr.info >> 22) & 1u;
    const bool up[3] = {(sg & 1u) != 0, (sg & 2u) != 0, (sg & 4u) != 0};
    const int d0 = up[0] ? 1 : -1, d1 = up[1] ? kTile : -kTile;
    const int d2 = up[2] ? kTile * kTile : -kTile * kTile;
    auto code = [&](unsigned int c) {
      const uint32_t w = c < 16 ? sr.code[0] : (c < 32 ? sr.code[1] : sr.code[2]);
      return (int)((w >> (2 * (c & 15))) & 3u);
    };
    auto delta = [&](int a) { return a == 0 ? d0 : (a == 1 ? d1 : d2); };
    auto hfac = [&](int a) {
      return a == 0 ? step_factor<double>(wp, 0, up[0])
                    : (a == 1 ? step_factor<double>(wp, 1, up[1]) : step_factor<double>(wp, 2, up[2]));
    };
    // forward: which visits add and which start a run (at most 46 visits)
    unsigned long long addb = 0ull, startb = 0ull;
    bool prev = adjm_prev_valid<kBits>(sr, ox, oy, oz, sy, sz, mask);
    int lv = lv0;
    for (unsigned int q = 0; q < n; q++) {
      if (q > 0) lv += delta(code(q - 1));
      const bool val = M[lds_slot<kBits>(lv)] != 0;
      if (val && prev) addb |= 1ull << q;
      if (val && !prev) startb |= 1ull << q;
      prev = val;
    }
    // backward from what the later segments add to the run the segment ends in
    const double T = tafter[sr.g];
    double S = T;
    if (kScatter && !last) {
      // the step out of the tile (into an adding visit whenever T != 0)
      const int ax = code(n - 1);
      atomicAdd(&A[ax][lds_slot<kBits>(lv)], hfac(ax) * T);
    }
    for (unsigned int q = n; q-- > 0;) {
      const int sl = lds_slot<kBits>(lv);
      const int a = q > 0 ? code(q - 1) : (int)((sr.info >> 23) & 3u);
      const int lp = q > 0 ? lv - delta(a) : lv;
      if ((addb >> q) & 1ull) {
        S += U[sl];
        if constexpr (kScatter) {
          // the step into visit q (for q = 0 the step into the tile; its
          // other end is counted by the segment it leaves)
          const double hS = hfac(a) * S;
          atomicAdd(&A[a][sl], hS);
          if (q > 0) atomicAdd(&A[a][lds_slot<kBits>(lp)], hS);
        }
      }
      if ((startb >> q) & 1ull) {
        if constexpr (!kScatter) atomicAdd(&A[0][sl], S);
        S = 0.0;
      }
      lv = lp;
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < kVox; l += blockDim.x) {
    const int i = ox + (l & kTileMask), j = oy + ((l >> kBits) & kTileMask), k = oz + (l >> (2 * kBits));
    if (i >= g.X || j >= g.Y || k >= g.Z) continue;
    const int64_t v = i + j * sy + k * sz;
    const int sl = lds_slot<kBits>(l);
    if constexpr (kScatter) {
      out[v] = A[0][sl];
      out[V + v] = A[1][sl];
      out[2 * V + v] = A[2][sl];
    } else {
      out[v] = A[0][sl];
    }
  }
}

hipError_t launch_adjm_tiles(int what, int ntiles, const unsigned long long* tile_ptr1,
                             const SegRec* segs_slot, const unsigned long long* tile_ptr,
                             const SegRec* segs, int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, GridDesc g, TileGrid tg,
                             const WalkParams& wp, const double* u, const double* mask,
                             double* useg, uint8_t* span, double* out, hipStream_t st) {
  // what 0: segment sums of the entering runs; 1: reverse scan; 2: the
  // scatter into out (dg, 3 V); 3: the gather into out (V)
  if (what == 1) {
    hipLaunchKernelGGL(k_adjm_rev_scan, dim3((unsigned int)((nrays + 255) / 256)), dim3(256), 0, st,
                       nrays, nseg, val_off, useg, span);
    return hipGetLastError();
  }
  if (what == 0) {
    if (tg.bits == 3)
      hipLaunchKernelGGL(k_adjm_seg_u<3>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs_slot, g,
                         tg, u, mask, useg, span);
    else
      hipLaunchKernelGGL(k_adjm_seg_u<4>, dim3(ntiles), dim3(1024), 0, st, tile_ptr1, segs_slot, g,
                         tg, u, mask, useg, span);
    return hipGetLastError();
  }
#define OMNI_AM(B, SC)                                                                       \
  hipLaunchKernelGGL((k_adjm_tile<B, SC>), dim3(ntiles), dim3(1024), 0, st, tile_ptr, segs, g, tg, \
                     wp, u, mask, useg, out)
  if (tg.bits == 3) {
    if (what == 2) OMNI_AM(3, true);
    else OMNI_AM(3, false);
  } else {
    if (what == 2) OMNI_AM(4, true);
    else OMNI_AM(4, false);
  }
#undef OMNI_AM
  return hipGetLastError();
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
