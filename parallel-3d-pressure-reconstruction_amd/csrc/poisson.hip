// Poisson path on the GPU: the reference's CPU entry points calRHS
// (calRHS.cpp:7-309) and sor3D (sor3D.cpp:9-1589) with their signatures, and
// a red-black SOR solver (omni3d_sor_solve) on the same 7-point operator.
//
// Stencil: in-box face neighbours only (the reference's 26 face / edge /
// corner cases are exactly "drop the neighbours outside the box"), summed
// in the order x-, x+, y-, y+, z-, z+ and divided by their count.  Every
// voxel update uses the reference's expression order, so Jacobi and
// red-black sweeps are bit-identical to a serial evaluation; only the
// residual's reduction order differs (FP64 tree instead of a serial FP32
// sum, which moves the eps stop only when the residual ties with eps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/omni3d.h"
#include "driver.hpp"
#include "launch.hpp"

namespace omni3d {

struct SorState {
  int stop_at;              // sweep (Jacobi) / iteration (red-black) that met eps, 0 = none
  int done;                 // sweeps / iterations executed
  double residual;          // last residual
};

__device__ __forceinline__ double nb_sum(int X, int Y, int Z, const double* __restrict__ p,
                                         int i, int j, int k, int64_t v, int* n_out) {
  const int64_t sy = X, sz = (int64_t)X * Y;
  double s = 0.0;
  int n = 0;
  if (i > 0) { s = n ? s + p[v - 1] : p[v - 1]; n++; }
  if (i < X - 1) { s = n ? s + p[v + 1] : p[v + 1]; n++; }
  if (j > 0) { s = n ? s + p[v - sy] : p[v - sy]; n++; }
  if (j < Y - 1) { s = n ? s + p[v + sy] : p[v + sy]; n++; }
  if (k > 0) { s = n ? s + p[v - sz] : p[v - sz]; n++; }
  if (k < Z - 1) { s = n ? s + p[v + sz] : p[v + sz]; n++; }
  *n_out = n;
  return s;
}

// calRHS: interior (calRHS.cpp:224-231) per axis
// (-d a[i] - d a[i-1]) / 2 + (d a[i] + d a[i+1]) / 2; a voxel on the box
// surface keeps only the inward half-term of each axis whose face it lies on
// (low face (d a[i] + d a[i+1]) / 2, high face (-d a[i] - d a[i-1]) / 2) and
// no term for its other axes (calRHS.cpp:19-222); times density.
__global__ void __launch_bounds__(256)
k_calrhs(int X, int Y, int Z, double dx, double dy, double dz, double rho,
         const double* __restrict__ ax, const double* __restrict__ ay,
         const double* __restrict__ az, double* __restrict__ rhs) {
  const int64_t V = (int64_t)X * Y * Z;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int c[3] = {(int)(v % X), (int)((v / X) % Y), (int)(v / ((int64_t)X * Y))};
    const int size[3] = {X, Y, Z};
    const double del[3] = {dx, dy, dz};
    const double* acc[3] = {ax, ay, az};
    const int64_t stride[3] = {1, X, (int64_t)X * Y};
    bool onb = false;
    for (int a = 0; a < 3; a++) onb |= (c[a] == 0 || c[a] == size[a] - 1);
    double r = 0.0;
    for (int a = 0; a < 3; a++) {
      const int64_t st = stride[a];
      const bool lo = c[a] == 0, hi = c[a] == size[a] - 1;
      if (!onb) {
        r += (-del[a] * acc[a][v] - del[a] * acc[a][v - st]) * 0.5;
        r += (del[a] * acc[a][v] + del[a] * acc[a][v + st]) * 0.5;
      } else if (lo) {
        r += (del[a] * acc[a][v] + del[a] * acc[a][v + st]) * 0.5;
      } else if (hi) {
        r += (-del[a] * acc[a][v] - del[a] * acc[a][v - st]) * 0.5;
      }
    }
    rhs[v] = r * rho;
  }
}

// One damped-Jacobi sweep (weight 0.8) over the BCScheme range.
__global__ void __launch_bounds__(256)
k_jacobi(int X, int Y, int Z, int i0, int ie, int j0, int je, int k0, int ke,
         const double* __restrict__ p, double* __restrict__ pn, const double* __restrict__ rhs,
         const SorState* __restrict__ st) {
  if (st->stop_at) return;
  const int nx = ie - i0, ny = je - j0;
  const int64_t n = (int64_t)nx * ny * (ke - k0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = i0 + (int)(t % nx), j = j0 + (int)((t / nx) % ny), k = k0 + (int)(t / ((int64_t)nx * ny));
    const int64_t v = (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
    int cn;
    double s = nb_sum(X, Y, Z, p, i, j, k, v, &cn);
    s += rhs[v];
    s = s / cn;
    pn[v] = p[v] + 0.8 * (s - p[v]);
  }
}

// Red-black SOR half sweep: voxels with (i + j + k) % 2 == color.
__global__ void __launch_bounds__(256)
k_redblack(int X, int Y, int Z, int color, double omega, double* __restrict__ p,
           const double* __restrict__ rhs, const SorState* __restrict__ st) {
  if (st->stop_at) return;
  const int64_t V = (int64_t)X * Y * Z;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % X), j = (int)((v / X) % Y), k = (int)(v / ((int64_t)X * Y));
    if (((i + j + k) & 1) != color) continue;
    int cn;
    double s = nb_sum(X, Y, Z, p, i, j, k, v, &cn);
    s += rhs[v];
    s = s / cn;
    p[v] = p[v] + omega * (s - p[v]);
  }
}

// The same half sweep with one thread per voxel of the colour: thread
// (h, j, k) updates i = 2h + ((j + k + color) & 1), so no lane idles and no
// index is divided.  Blocks of 64 x 4 threads: a wavefront covers 128
// consecutive voxels of one row.
__global__ void __launch_bounds__(256)
k_redblack_rows(int X, int Y, int Z, int color, double omega, double* __restrict__ p,
                const double* __restrict__ rhs, const SorState* __restrict__ st) {
  if (st->stop_at) return;
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int k = (int)blockIdx.z;
  const int i = 2 * (int)(blockIdx.x * 64 + threadIdx.x) + ((j + k + color) & 1);
  if (i >= X || j >= Y) return;
  const int64_t v = (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
  int cn;
  double s = nb_sum(X, Y, Z, p, i, j, k, v, &cn);
  s += rhs[v];
  s = s / cn;
  p[v] = p[v] + omega * (s - p[v]);
}

// One whole red-black iteration in one pass: a workgroup owns a TX x TY
// column of the grid over a chunk of z planes and marches up it with a ring
// of planes of p (the column plus a two-voxel halo) in LDS.  At plane k it
// takes the red half sweep of plane k (over the column plus a one-voxel
// halo: red needs only old black values, which the halo of two holds), and
// the black half sweep of plane k - 1 (its red neighbours on planes k - 2 ..
// k are new by then); plane k - 1 is stored one step later.  The plane two
// steps ahead and the next step's right-hand sides load into registers while
// a step computes.  Reads come from p_in and writes go to p_out, so
// workgroups never see each other's new values; each voxel's update is the
// same expression on the same values as in the half-sweep kernels.
template <int kRbTX, int kRbTY, int kRbZ, int kRbThreads>
__global__ void __launch_bounds__(kRbThreads)
k_redblack_fused(int X, int Y, int Z, double omega, const double* __restrict__ pin,
                 const double* __restrict__ rhs, double* __restrict__ pout,
                 const SorState* __restrict__ st) {
  if (st->stop_at) return;
  constexpr int kRbW = kRbTX + 4, kRbH = kRbTY + 4;
  constexpr int kHalfX = kRbTX / 2 + 1;  // red voxels per row of the column plus halo
  constexpr int kLoads = (kRbW * kRbH + kRbThreads - 1) / kRbThreads;  // plane values per thread
  constexpr int kRed = (kHalfX * (kRbTY + 2) + kRbThreads - 1) / kRbThreads;
  constexpr int kBlack = (kRbTX / 2 * kRbTY + kRbThreads - 1) / kRbThreads;
  __shared__ double P[5][kRbH][kRbW];
  const int x0 = (int)blockIdx.x * kRbTX, y0 = (int)blockIdx.y * kRbTY;
  const int z0 = (int)blockIdx.z * kRbZ, z1 = min(Z, z0 + kRbZ);
  const int64_t sy = X, sz = (int64_t)X * Y;
  const int tid = (int)threadIdx.x;
  auto slot = [](int k) { return (k + 5) % 5; };  // k >= -2
  // the plane two steps ahead and the right-hand sides of the next step are
  // loaded into registers while the current step computes
  double pf[kLoads], rr[kRed], rb[kBlack];
  auto issue = [&](int k) {
    if (k < 0 || k >= Z) return;
#pragma unroll
    for (int m = 0; m < kLoads; m++) {
      const int l = tid + m * kRbThreads;
      const int i = x0 - 2 + l % kRbW, j = y0 - 2 + l / kRbW;
      if (l < kRbW * kRbH && i >= 0 && i < X && j >= 0 && j < Y) pf[m] = pin[i + j * sy + k * sz];
    }
  };
  auto commit = [&](int k) {
    if (k < 0 || k >= Z) return;
#pragma unroll
    for (int m = 0; m < kLoads; m++) {
      const int l = tid + m * kRbThreads;
      const int i = x0 - 2 + l % kRbW, j = y0 - 2 + l / kRbW;
      if (l < kRbW * kRbH && i >= 0 && i < X && j >= 0 && j < Y) P[slot(k)][l / kRbW][l % kRbW] = pf[m];
    }
  };
  // red voxels ((i + j + k) even) of the column plus a one-voxel halo:
  // TX / 2 + 1 per row of TX + 2; black voxels of the column: TX / 2 per row
  auto red_at = [&](int l, int k, int* i, int* j) {
    *j = y0 - 1 + l / kHalfX;
    *i = x0 - 1 + ((x0 - 1 + *j + k) & 1) + 2 * (l % kHalfX);
    return l < kHalfX * (kRbTY + 2) && *i >= 0 && *i < X && *j >= 0 && *j < Y;
  };
  auto black_at = [&](int l, int k, int* i, int* j) {
    *j = y0 + l / (kRbTX / 2);
    *i = x0 + ((x0 + *j + k + 1) & 1) + 2 * (l % (kRbTX / 2));
    return l < kRbTX / 2 * kRbTY && *i < X && *j < Y;
  };
  auto issue_rhs = [&](int kr, int kb) {
#pragma unroll
    for (int m = 0; m < kRed; m++) {
      int i, j;
      if (kr >= 0 && kr < Z && red_at(tid + m * kRbThreads, kr, &i, &j)) rr[m] = rhs[i + j * sy + kr * sz];
    }
#pragma unroll
    for (int m = 0; m < kBlack; m++) {
      int i, j;
      if (kb >= z0 && kb < z1 && black_at(tid + m * kRbThreads, kb, &i, &j))
        rb[m] = rhs[i + j * sy + kb * sz];
    }
  };
  // in-box neighbour sum in the reference's order (x-, x+, y-, y+, z-, z+)
  auto update = [&](int i, int j, int k, double r) {
    const int xx = i - x0 + 2, yy = j - y0 + 2;
    const double(*Pc)[kRbW] = P[slot(k)];
    double s = 0.0;
    int n = 0;
    if (i > 0) { s = n ? s + Pc[yy][xx - 1] : Pc[yy][xx - 1]; n++; }
    if (i < X - 1) { s = n ? s + Pc[yy][xx + 1] : Pc[yy][xx + 1]; n++; }
    if (j > 0) { s = n ? s + Pc[yy - 1][xx] : Pc[yy - 1][xx]; n++; }
    if (j < Y - 1) { s = n ? s + Pc[yy + 1][xx] : Pc[yy + 1][xx]; n++; }
    if (k > 0) { const double q = P[slot(k - 1)][yy][xx]; s = n ? s + q : q; n++; }
    if (k < Z - 1) { const double q = P[slot(k + 1)][yy][xx]; s = n ? s + q : q; n++; }
    s += r;
    s = s / n;
    const double old = Pc[yy][xx];
    P[slot(k)][yy][xx] = old + omega * (s - old);
  };
  issue(z0 - 2);
  commit(z0 - 2);
  issue(z0 - 1);
  commit(z0 - 1);
  issue(z0);
  issue_rhs(z0 - 1, z0 - 2);
  // step k: red half sweep of plane k, store of plane k - 2 (finished by the
  // previous step), black half sweep of plane k - 1; a ring of five planes
  // keeps every plane a step reads or stores clear of the commit two steps on
  for (int k = z0 - 1; k <= z1 + 1; k++) {
    if (k + 1 <= z1 + 1) commit(k + 1);
    if (k + 2 <= z1 + 1) issue(k + 2);
    double r0[kRed], b0[kBlack];
#pragma unroll
    for (int m = 0; m < kRed; m++) r0[m] = rr[m];
#pragma unroll
    for (int m = 0; m < kBlack; m++) b0[m] = rb[m];
    issue_rhs(k + 1, k);
    __syncthreads();
    if (k >= 0 && k < Z && k <= z1) {
#pragma unroll
      for (int m = 0; m < kRed; m++) {
        int i, j;
        if (red_at(tid + m * kRbThreads, k, &i, &j)) update(i, j, k, r0[m]);
      }
    }
    const int ks = k - 2;
    if (ks >= z0 && ks < z1) {
      const double(*Pb)[kRbW] = P[slot(ks)];
      for (int l = tid; l < kRbTX * kRbTY; l += kRbThreads) {
        const int i = x0 + (l % kRbTX), j = y0 + l / kRbTX;
        if (i < X && j < Y) pout[i + j * sy + ks * sz] = Pb[j - y0 + 2][i - x0 + 2];
      }
    }
    __syncthreads();
    const int kb = k - 1;
    if (kb >= z0 && kb < z1) {
#pragma unroll
      for (int m = 0; m < kBlack; m++) {
        int i, j;
        if (black_at(tid + m * kRbThreads, kb, &i, &j)) update(i, j, kb, b0[m]);
      }
    }
  }
}

// kK whole red-black iterations in one pass (temporal blocking): the same
// z-march with 2 kK half sweeps ("stages") in flight.  Stage s (red for even
// s, black for odd s, iteration s / 2 + 1) takes plane k - s at step k over
// the column plus a halo of h = 2 kK - 1 - s voxels in x, y and z: each
// stage reads the previous stage's colour on the planes around it, which
// that stage finished one step earlier or earlier in this step, so the
// stages run in order within a step, a barrier apart.  The column's planes
// load with a halo of 2 kK (a ring of 2 kK + 3 planes; the plane two steps
// ahead loads into registers during a step); a plane is stored one step
// after the last stage finished it.  Reads come from p_in and writes go to
// p_out; every voxel update is the same expression on the same values as in
// the half-sweep kernels, so kK iterations here equal kK single ones bit for
// bit.  The halo voxels are updated redundantly by each column that needs
// them (the price of one read and one write of p per kK iterations).
template <int kTX, int kTY, int kZ, int kK, int kThreads>
__global__ void __launch_bounds__(kThreads)
k_redblack_multi(int X, int Y, int Z, double omega, const double* __restrict__ pin,
                 const double* __restrict__ rhs, double* __restrict__ pout,
                 const SorState* __restrict__ st) {
  if (st->stop_at) return;
  constexpr int kStages = 2 * kK;
  constexpr int kHalo = 2 * kK;
  constexpr int kW = kTX + 2 * kHalo, kH = kTY + 2 * kHalo;
  constexpr int kRing = 2 * kK + 3;
  constexpr int kLoads = (kW * kH + kThreads - 1) / kThreads;
  static_assert(kTX % 2 == 0, "even column width");
  __shared__ double P[kRing][kH][kW];
  const int x0 = (int)blockIdx.x * kTX, y0 = (int)blockIdx.y * kTY;
  const int z0 = (int)blockIdx.z * kZ, z1 = min(Z, z0 + kZ);
  const int64_t sy = X, sz = (int64_t)X * Y;
  const int tid = (int)threadIdx.x;
  auto slot = [](int k) { return (k + 8 * kRing) % kRing; };  // k >= -2 kHalo
  double pf[kLoads];
  auto issue = [&](int k) {
    if (k < 0 || k >= Z) return;
#pragma unroll
    for (int m = 0; m < kLoads; m++) {
      const int l = tid + m * kThreads;
      const int i = x0 - kHalo + l % kW, j = y0 - kHalo + l / kW;
      if (l < kW * kH && i >= 0 && i < X && j >= 0 && j < Y) pf[m] = pin[i + j * sy + k * sz];
    }
  };
  auto commit = [&](int k) {
    if (k < 0 || k >= Z) return;
#pragma unroll
    for (int m = 0; m < kLoads; m++) {
      const int l = tid + m * kThreads;
      const int i = x0 - kHalo + l % kW, j = y0 - kHalo + l / kW;
      if (l < kW * kH && i >= 0 && i < X && j >= 0 && j < Y) P[slot(k)][l / kW][l % kW] = pf[m];
    }
  };
  // in-box neighbour sum in the reference's order (x-, x+, y-, y+, z-, z+)
  auto update = [&](int i, int j, int k, double r) {
    const int xx = i - x0 + kHalo, yy = j - y0 + kHalo;
    double(*Pc)[kW] = P[slot(k)];
    double s = 0.0;
    int n = 0;
    if (i > 0) { s = n ? s + Pc[yy][xx - 1] : Pc[yy][xx - 1]; n++; }
    if (i < X - 1) { s = n ? s + Pc[yy][xx + 1] : Pc[yy][xx + 1]; n++; }
    if (j > 0) { s = n ? s + Pc[yy - 1][xx] : Pc[yy - 1][xx]; n++; }
    if (j < Y - 1) { s = n ? s + Pc[yy + 1][xx] : Pc[yy + 1][xx]; n++; }
    if (k > 0) { const double q = P[slot(k - 1)][yy][xx]; s = n ? s + q : q; n++; }
    if (k < Z - 1) { const double q = P[slot(k + 1)][yy][xx]; s = n ? s + q : q; n++; }
    s += r;
    s = s / n;
    const double old = Pc[yy][xx];
    Pc[yy][xx] = old + omega * (s - old);
  };
  // stage s on plane m: the voxels of its colour in the column plus halo h;
  // thread tid takes voxels tid + q kThreads, q < kPer
  constexpr int kPer = ((kTX / 2 + kStages - 1) * (kTY + 2 * (kStages - 1)) + kThreads - 1) /
                       kThreads;
  auto voxel = [&](int s, int m, int q, int* i, int* j) {
    const int h = kStages - 1 - s;
    const int half = kTX / 2 + h, rows = kTY + 2 * h;
    const int l = tid + q * kThreads;
    *j = y0 - h + l / half;
    *i = x0 - h + ((x0 - h + *j + m + (s & 1)) & 1) + 2 * (l % half);
    return l < half * rows && m >= 0 && m < Z && m >= z0 - h && m < z1 + h && *i >= 0 &&
           *i < X && *j >= 0 && *j < Y;
  };
  // the right-hand sides of a step's stages load into registers one step
  // ahead
  double rr[kStages][kPer];
  auto issue_rhs = [&](int k) {
#pragma unroll
    for (int s = 0; s < kStages; s++) {
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        int i, j;
        if (voxel(s, k - s, q, &i, &j)) rr[s][q] = rhs[i + j * sy + (int64_t)(k - s) * sz];
      }
    }
  };
  auto stage = [&](int s, int m, const double (&r)[kPer]) {
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      int i, j;
      if (voxel(s, m, q, &i, &j)) update(i, j, m, r[q]);
    }
  };
  const int kb = z0 - kStages + 1;  // the first step: stage 0 on its lowest plane
  issue(kb - 1);
  commit(kb - 1);
  issue(kb);
  commit(kb);
  issue(kb + 1);
  issue_rhs(kb);
  for (int k = kb; k <= z1 + kStages - 1; k++) {
    commit(k + 1);
    issue(k + 2);
    double r0[kStages][kPer];
#pragma unroll
    for (int s = 0; s < kStages; s++)
#pragma unroll
      for (int q = 0; q < kPer; q++) r0[s][q] = rr[s][q];
    issue_rhs(k + 1);
    __syncthreads();
    stage(0, k, r0[0]);
    // the plane the last stage finished at the previous step
    const int ms = k - kStages;
    if (ms >= z0 && ms < z1) {
      const double(*Pb)[kW] = P[slot(ms)];
      for (int l = tid; l < kTX * kTY; l += kThreads) {
        const int i = x0 + (l % kTX), j = y0 + l / kTX;
        if (i < X && j < Y) pout[i + j * sy + ms * sz] = Pb[j - y0 + kHalo][i - x0 + kHalo];
      }
    }
#pragma unroll
    for (int s = 1; s < kStages; s++) {
      __syncthreads();
      stage(s, k - s, r0[s]);
    }
  }
}

// Interior residual sum |sum_nb - 6 p + rhs| (sor3D.cpp:741-781) reduced
// over the grid; the last block applies the stop rule of sweep / iteration
// `it`: residue = sum / X / Y / Z, stop when residue < eps.  The sum is FP64
// in a fixed tree order; fp32 rounds it to float once, where the reference
// adds each |temp| into a float (sor3D.cpp:774): the residues can differ in
// their last float bits, which moves the stop sweep only when the residue
// ties with eps.
constexpr int kResBlock = 256;
constexpr int kResPlanes = 4;

// Blocks of 64 x 4 threads over the interior (i, j), each thread over
// kResPlanes consecutive planes; each block writes its partial sum.  k_residual_finish then
// adds the partials in a fixed order (each thread a fixed strided subset,
// then a fixed tree) and applies the stop rule.
__global__ void __launch_bounds__(kResBlock)
k_residual(int X, int Y, int Z, const double* __restrict__ p, const double* __restrict__ rhs,
           double* __restrict__ partials, const SorState* __restrict__ st) {
  if (st->stop_at) return;
  __shared__ double s_part[kResBlock / 64];
  const int i = 1 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * 4 + threadIdx.y);
  const int tid = (int)(threadIdx.y * 64 + threadIdx.x);
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kResPlanes; q++) {
    const int k = 1 + (int)blockIdx.z * kResPlanes + q;
    if (i < X - 1 && j < Y - 1 && k < Z - 1) {
      const int64_t v = (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
      int cn;
      double tmp = nb_sum(X, Y, Z, p, i, j, k, v, &cn);
      tmp += -6.0 * p[v];
      tmp += rhs[v];
      acc += fabs(tmp);
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) s_part[wave] = acc;
  __syncthreads();
  if (tid == 0) {
    double a = 0.0;
    for (int w = 0; w < kResBlock / 64; w++) a += s_part[w];
    partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = a;
  }
}

// One workgroup: residue = sum / X / Y / Z (fp32: the FP64 sum rounded to
// float once, then divided in float), stop when residue < eps.
__global__ void __launch_bounds__(kResBlock)
k_residual_finish(int X, int Y, int Z, int64_t nblk, const double* __restrict__ partials, int it,
                  double eps, int fp32, SorState* __restrict__ st) {
  if (st->stop_at) return;
  __shared__ double s_part[kResBlock / 64];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double a = 0.0;
  for (int64_t q = tid; q < nblk; q += kResBlock) a += partials[q];
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
  if (lane == 0) s_part[wave] = a;
  __syncthreads();
  if (tid != 0) return;
  a = 0.0;
  for (int w = 0; w < kResBlock / 64; w++) a += s_part[w];
  double res;
  if (fp32) {
    float r = (float)a;
    r = r / X / Y / Z;
    res = r;
  } else {
    res = a / X / Y / Z;
  }
  st->residual = res;
  st->done = it;
  if (res < eps) st->stop_at = it;
}

// After Jacobi sweep `it`: if it met eps, p = pn over the whole grid
// (sor3D.cpp:785-789); otherwise p = pn and pn = 0 over the BCScheme range.
__global__ void __launch_bounds__(256)
k_jacobi_commit(int X, int Y, int Z, int i0, int ie, int j0, int je, int k0, int ke, int it,
                double* __restrict__ p, double* __restrict__ pn, const SorState* __restrict__ st) {
  const int stop = st->stop_at;
  if (stop && stop != it) return;
  if (stop == it) {
    const int64_t V = (int64_t)X * Y * Z;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x)
      p[v] = pn[v];
    return;
  }
  const int nx = ie - i0, ny = je - j0;
  const int64_t n = (int64_t)nx * ny * (ke - k0);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = i0 + (int)(t % nx), j = j0 + (int)((t / nx) % ny), k = k0 + (int)(t / ((int64_t)nx * ny));
    const int64_t v = (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
    p[v] = pn[v];
    pn[v] = 0.0;
  }
}

}  // namespace omni3d

using namespace omni3d;

namespace {

// Device copies of the host arrays of one Poisson call.
struct PoissonBufs {
  DevBuf a[6];
  DevBuf state, partials;
};

bool grid_ok(int X, int Y, int Z) { return X >= 1 && Y >= 1 && Z >= 1 && (int64_t)X * Y * Z < (1ll << 40); }

// residual launch: 64 x 4 blocks over the interior (i, j), kResPlanes planes per block
dim3 res_grid(int X, int Y, int Z) {
  return dim3((unsigned int)((std::max(X - 2, 1) + 63) / 64),
              (unsigned int)((std::max(Y - 2, 1) + 3) / 4),
              (unsigned int)((std::max(Z - 2, 1) + kResPlanes - 1) / kResPlanes));
}
size_t res_blocks(int X, int Y, int Z) {
  const dim3 g = res_grid(X, Y, Z);
  return (size_t)g.x * g.y * g.z;
}

void launch_residual(int X, int Y, int Z, const double* p, const double* rhs, int it, double eps,
                     int fp32, double* partials, SorState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_residual, res_grid(X, Y, Z), dim3(64, 4), 0, s, X, Y, Z, p, rhs, partials,
                     st);
  hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(kResBlock), 0, s, X, Y, Z,
                     (int64_t)res_blocks(X, Y, Z), partials, it, eps, fp32, st);
}

#define POIS_TRY(expr)                                                          \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      return fail(nullptr, OMNI3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                           \
  } while (0)

int calrhs_impl(int X, int Y, int Z, double dx, double dy, double dz, double rho,
                const double* ax, const double* ay, const double* az, double* rhs) {
  if (!grid_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!ax || !ay || !az || !rhs) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  const int64_t V = (int64_t)X * Y * Z;
  const size_t b = sizeof(double) * V;
  PoissonBufs B;
  for (int i = 0; i < 4; i++) POIS_TRY(B.a[i].alloc(b));
  POIS_TRY(hipMemcpy(B.a[0].as<void>(), ax, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemcpy(B.a[1].as<void>(), ay, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemcpy(B.a[2].as<void>(), az, b, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_calrhs, dim3(grid_for(V)), dim3(256), 0, nullptr, X, Y, Z, dx, dy, dz, rho,
                     B.a[0].as<double>(), B.a[1].as<double>(), B.a[2].as<double>(),
                     B.a[3].as<double>());
  POIS_TRY(hipGetLastError());
  POIS_TRY(hipMemcpy(rhs, B.a[3].as<void>(), b, hipMemcpyDeviceToHost));
  return OMNI3D_OK;
}

int sor3d_impl(int X, int Y, int Z, int NoItr, const int* bc, double* p, double* pn,
               const double* rhs, double eps, int* sweeps_out, double* res_out) {
  if (!grid_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!bc || !p || !pn || !rhs) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  const int i0 = bc[0], ie = X - bc[1], j0 = bc[2], je = Y - bc[3], k0 = bc[4], ke = Z - bc[5];
  if (i0 < 0 || j0 < 0 || k0 < 0 || ie > X || je > Y || ke > Z)
    return fail(nullptr, OMNI3D_ERR_ARG, "BCScheme outside the grid");
  const int64_t V = (int64_t)X * Y * Z;
  const size_t b = sizeof(double) * V;
  PoissonBufs B;
  for (int i = 0; i < 3; i++) POIS_TRY(B.a[i].alloc(b));
  POIS_TRY(B.state.alloc(sizeof(SorState)));
  POIS_TRY(B.partials.alloc(sizeof(double) * res_blocks(X, Y, Z)));
  double* dp = B.a[0].as<double>();
  double* dpn = B.a[1].as<double>();
  double* drhs = B.a[2].as<double>();
  SorState* st = B.state.as<SorState>();
  POIS_TRY(hipMemcpy(dp, p, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemcpy(dpn, pn, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemcpy(drhs, rhs, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemset(st, 0, sizeof(SorState)));
  const bool empty = ie <= i0 || je <= j0 || ke <= k0;
  const int64_t nr = empty ? 0 : (int64_t)(ie - i0) * (je - j0) * (ke - k0);
  // sor3D.cpp:21-1587: each pass of the while loop runs two sweeps (forward
  // and reversed order; Jacobi makes the order irrelevant), each followed by
  // the residue test and the commit.
  int iteration = 0, sweep = 0;
  while (iteration < NoItr) {
    for (int half = 0; half < 2; half++) {
      sweep++;
      if (!empty)
        hipLaunchKernelGGL(k_jacobi, dim3(grid_for(nr)), dim3(256), 0, nullptr, X, Y, Z, i0, ie,
                           j0, je, k0, ke, dp, dpn, drhs, st);
      launch_residual(X, Y, Z, dpn, drhs, sweep, eps, 1, B.partials.as<double>(), st, nullptr);
      hipLaunchKernelGGL(k_jacobi_commit, dim3(grid_for(std::max<int64_t>(nr, V))), dim3(256), 0,
                         nullptr, X, Y, Z, i0, empty ? i0 : ie, j0, je, k0, ke, sweep, dp, dpn, st);
      POIS_TRY(hipGetLastError());
      iteration++;
    }
  }
  SorState hs;
  POIS_TRY(hipMemcpy(&hs, st, sizeof(hs), hipMemcpyDeviceToHost));
  POIS_TRY(hipMemcpy(p, dp, b, hipMemcpyDeviceToHost));
  POIS_TRY(hipMemcpy(pn, dpn, b, hipMemcpyDeviceToHost));
  if (sweeps_out) *sweeps_out = hs.stop_at ? hs.stop_at : hs.done;
  if (res_out) *res_out = hs.residual;
  return OMNI3D_OK;
}

}  // namespace

extern "C" {

void calRHS(int Xsize, int Ysize, int Zsize, double deltx, double delty, double deltz,
            double density, double* DuDt, double* DvDt, double* DwDt, double* RHS) {
  (void)calrhs_impl(Xsize, Ysize, Zsize, deltx, delty, deltz, density, DuDt, DvDt, DwDt, RHS);
}

void sor3D(int Xsize, int Ysize, int Zsize, int NoItr, int* BCScheme, double* x, double* y,
           double* z, double* p, double* pn, double* RHS, double eps) {
  (void)x;
  (void)y;
  (void)z;
  (void)sor3d_impl(Xsize, Ysize, Zsize, NoItr, BCScheme, p, pn, RHS, eps, nullptr, nullptr);
}

int omni3d_sor3d_ex(int X, int Y, int Z, int NoItr, const int* BCScheme, double* p, double* pn,
                    const double* rhs, double eps, int* sweeps_out, double* residue_out) {
  return sor3d_impl(X, Y, Z, NoItr, BCScheme, p, pn, rhs, eps, sweeps_out, residue_out);
}

int omni3d_calrhs(int X, int Y, int Z, double dx, double dy, double dz, double density,
                  const double* DuDt, const double* DvDt, const double* DwDt, double* rhs) {
  return calrhs_impl(X, Y, Z, dx, dy, dz, density, DuDt, DvDt, DwDt, rhs);
}

int omni3d_sor_solve(int X, int Y, int Z, double omega, int max_iter, double eps,
                     const double* rhs, double* p, int* iterations_out, double* residual_out) {
  if (!grid_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!rhs || !p) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  const int64_t V = (int64_t)X * Y * Z;
  const size_t b = sizeof(double) * V;
  DevBuf dp, drhs;
  POIS_TRY(dp.alloc(b));
  POIS_TRY(drhs.alloc(b));
  POIS_TRY(hipMemcpy(dp.as<void>(), p, b, hipMemcpyHostToDevice));
  POIS_TRY(hipMemcpy(drhs.as<void>(), rhs, b, hipMemcpyHostToDevice));
  const int rc = omni3d_sor_solve_device(X, Y, Z, omega, max_iter, eps, drhs.as<double>(),
                                         dp.as<double>(), nullptr, iterations_out, residual_out);
  if (rc) return rc;
  POIS_TRY(hipMemcpy(p, dp.as<void>(), b, hipMemcpyDeviceToHost));
  return OMNI3D_OK;
}

int omni3d_sor_solve_device(int X, int Y, int Z, double omega, int max_iter, double eps,
                            const double* rhs_d, double* p_d, void* hip_stream,
                            int* iterations_out, double* residual_out) {
  if (!grid_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!rhs_d || !p_d) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  if (!(omega > 0.0 && omega < 2.0)) return fail(nullptr, OMNI3D_ERR_ARG, "omega must be in (0, 2)");
  if (max_iter < 1) return fail(nullptr, OMNI3D_ERR_ARG, "max_iter must be >= 1");
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  const int64_t V = (int64_t)X * Y * Z;
  PoissonBufs B;
  POIS_TRY(B.state.alloc(sizeof(SorState)));
  POIS_TRY(B.partials.alloc(sizeof(double) * res_blocks(X, Y, Z)));
  SorState* st = B.state.as<SorState>();
  POIS_TRY(hipMemsetAsync(st, 0, sizeof(SorState), s));
  // default: one fused pass per iteration (p_d and a second buffer in
  // turn); OMNI3D_SOR=halves: two row-parallel half sweeps in place
  const char* env = std::getenv("OMNI3D_SOR");
  // column 64 x 8 over 32 planes, 256 threads (fastest of 64 x {8, 16} x
  // {16, 32} planes x {256, 512} threads at 256^3)
  constexpr int tx = 64, ty = 8, tz = 32;
  const bool fused = !(env && std::strcmp(env, "halves") == 0) &&
                     (Z + tz - 1) / tz <= 65535 && (Y + ty - 1) / ty <= 65535;
  double* buf[2] = {p_d, nullptr};
  if (fused) {
    POIS_TRY(B.a[0].alloc(sizeof(double) * V));
    buf[1] = B.a[0].as<double>();
  }
  const dim3 fgrid((unsigned int)((X + tx - 1) / tx), (unsigned int)((Y + ty - 1) / ty),
                   (unsigned int)((Z + tz - 1) / tz));
  const unsigned int nh = (unsigned int)((X + 1) / 2);
  const bool rows = (Y + 3) / 4 <= 65535 && Z <= 65535;
  const dim3 rgrid((nh + 63) / 64, (unsigned int)((Y + 3) / 4), (unsigned int)Z), rblock(64, 4);
  const unsigned int g = grid_for(V);
  // with no stop test between iterations (eps <= 0) pairs of iterations run
  // as one temporally blocked pass (k_redblack_multi, two iterations, same
  // columns); OMNI3D_SOR=single keeps one pass per iteration
  // column 64 x 16 over 64 planes, 1024 threads, one workgroup per CU (97 KB
  // of LDS): the fastest of 32-128 x 8-16 columns, 16-64 planes, 256-1024
  // threads and two or three iterations per pass at 256^3
  constexpr int mtx = 64, mty = 16, mtz = 64, mthreads = 1024;
  const dim3 mgrid((unsigned int)((X + mtx - 1) / mtx), (unsigned int)((Y + mty - 1) / mty),
                   (unsigned int)((Z + mtz - 1) / mtz));
  const bool pairs = fused && eps <= 0.0 && !(env && std::strcmp(env, "single") == 0) &&
                     mgrid.y <= 65535 && mgrid.z <= 65535;
  double* src = p_d;  // the buffer holding the latest iterate (fused passes)
  for (int it = 1; it <= max_iter; it++) {
    double* cur = p_d;
    if (fused) {
      double* dst = src == buf[0] ? buf[1] : buf[0];
      if (pairs && it + 1 <= max_iter) {
        hipLaunchKernelGGL((k_redblack_multi<mtx, mty, mtz, 2, mthreads>), mgrid, dim3(mthreads), 0,
                           s, X, Y, Z, omega, src, rhs_d, dst, st);
        it++;
      } else {
        hipLaunchKernelGGL((k_redblack_fused<tx, ty, tz, 256>), fgrid, dim3(256), 0, s, X, Y, Z,
                           omega, src, rhs_d, dst, st);
      }
      src = dst;
      cur = dst;
    } else {
      for (int color = 0; color < 2; color++) {
        if (rows)
          hipLaunchKernelGGL(k_redblack_rows, rgrid, rblock, 0, s, X, Y, Z, color, omega, p_d,
                             rhs_d, st);
        else
          hipLaunchKernelGGL(k_redblack, dim3(g), dim3(256), 0, s, X, Y, Z, color, omega, p_d,
                             rhs_d, st);
      }
    }
    if (eps > 0.0 || it == max_iter)
      launch_residual(X, Y, Z, cur, rhs_d, it, eps, 0, B.partials.as<double>(), st, s);
    POIS_TRY(hipGetLastError());
  }
  SorState hs;
  POIS_TRY(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
  POIS_TRY(hipStreamSynchronize(s));
  const int last = hs.stop_at ? hs.stop_at : max_iter;
  // the buffer the last executed iteration wrote: with single passes
  // iteration i writes buf[i & 1] (later launches return at once after a
  // stop); with pairs (no stop test) the latest one
  const double* res = pairs ? src : buf[last & 1];
  if (fused && res != p_d) {
    POIS_TRY(hipMemcpyAsync(p_d, res, sizeof(double) * V, hipMemcpyDeviceToDevice, s));
    POIS_TRY(hipStreamSynchronize(s));
  }
  if (iterations_out) *iterations_out = last;
  if (residual_out) *residual_out = hs.residual;
  return OMNI3D_OK;
}

}  // extern "C"
