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

#include <cmath>
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
  unsigned int ticket;      // last-block counter
  unsigned int pad;
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

// Interior residual sum |sum_nb - 6 p + rhs| (sor3D.cpp:741-781) reduced
// over the grid; the last block applies the stop rule of sweep / iteration
// `it`: residue = sum / X / Y / Z, stop when residue < eps.  fp32 rounds the
// residue to float as the reference's accumulator does.
constexpr int kResBlock = 256;

__global__ void __launch_bounds__(kResBlock)
k_residual(int X, int Y, int Z, const double* __restrict__ p, const double* __restrict__ rhs,
           int it, double eps, int fp32, double* __restrict__ partials, SorState* __restrict__ st) {
  if (st->stop_at) return;
  __shared__ double s_part[kResBlock / 64];
  __shared__ int s_last;
  const int nx = X - 2, ny = Y - 2, nz = Z - 2;
  const int64_t n = (nx > 0 && ny > 0 && nz > 0) ? (int64_t)nx * ny * nz : 0;
  double acc = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(t % nx), j = 1 + (int)((t / nx) % ny), k = 1 + (int)(t / ((int64_t)nx * ny));
    const int64_t v = (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
    int cn;
    double tmp = nb_sum(X, Y, Z, p, i, j, k, v, &cn);
    tmp += -6.0 * p[v];
    tmp += rhs[v];
    acc += fabs(tmp);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_part[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int w = 0; w < kResBlock / 64; w++) a += s_part[w];
    partials[blockIdx.x] = a;
    __threadfence();
    const unsigned int t =
        __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last || threadIdx.x != 0) return;
  __threadfence();
  double a = 0.0;
  for (unsigned int b = 0; b < gridDim.x; b++)
    a += __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  st->ticket = 0;
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

unsigned int res_grid(int X, int Y, int Z) { return grid_for((int64_t)X * Y * Z, 1024); }

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
  POIS_TRY(B.partials.alloc(sizeof(double) * res_grid(X, Y, Z)));
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
      hipLaunchKernelGGL(k_residual, dim3(res_grid(X, Y, Z)), dim3(kResBlock), 0, nullptr, X, Y,
                         Z, dpn, drhs, sweep, eps, 1, B.partials.as<double>(), st);
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
  POIS_TRY(B.partials.alloc(sizeof(double) * res_grid(X, Y, Z)));
  SorState* st = B.state.as<SorState>();
  POIS_TRY(hipMemsetAsync(st, 0, sizeof(SorState), s));
  // row-parallel half sweeps where the grid fits the launch limits
  const unsigned int nh = (unsigned int)((X + 1) / 2);
  const bool rows = (Y + 3) / 4 <= 65535 && Z <= 65535;
  const dim3 rgrid((nh + 63) / 64, (unsigned int)((Y + 3) / 4), (unsigned int)Z), rblock(64, 4);
  const unsigned int g = grid_for(V);
  for (int it = 1; it <= max_iter; it++) {
    for (int color = 0; color < 2; color++) {
      if (rows)
        hipLaunchKernelGGL(k_redblack_rows, rgrid, rblock, 0, s, X, Y, Z, color, omega, p_d, rhs_d,
                           st);
      else
        hipLaunchKernelGGL(k_redblack, dim3(g), dim3(256), 0, s, X, Y, Z, color, omega, p_d,
                           rhs_d, st);
    }
    if (eps > 0.0 || it == max_iter)
      hipLaunchKernelGGL(k_residual, dim3(res_grid(X, Y, Z)), dim3(kResBlock), 0, s, X, Y, Z,
                         p_d, rhs_d, it, eps, 0, B.partials.as<double>(), st);
    POIS_TRY(hipGetLastError());
  }
  SorState hs;
  POIS_TRY(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
  POIS_TRY(hipStreamSynchronize(s));
  if (iterations_out) *iterations_out = hs.stop_at ? hs.stop_at : max_iter;
  if (residual_out) *residual_out = hs.residual;
  return OMNI3D_OK;
}

}  // extern "C"
