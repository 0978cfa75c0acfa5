// initialIntegration (initialIntegration.cpp:9-408) on the GPU.
//
// The reference marches from the centre voxel (X/2, Y/2, Z/2) outward in
// eight serial octant sweeps.  A voxel's value is the mean of the one-step
// trapezoid predictions from its neighbours toward the centre, along the
// axes on which it is off the centre (x, then y, then z):
//     neighbour on the high side:  pn[nb] + 0.5 * (a[v] + a[nb]) * delta * rho
//     neighbour on the low side:   pn[nb] - 0.5 * (a[v] + a[nb]) * delta * rho
// so it depends only on voxels one L1 step closer to the centre.  Here the
// march is level-synchronous: launch L computes every voxel at L1 distance L
// from the centre, all independently, with the reference's expression order
// (the first prediction assigned, the others added, then / 2.0 or / 3.0), so
// each voxel equals the serial sweep's value bit for bit.  A voxel on a
// centre plane, which the reference recomputes in every octant it lies in
// with the same expression, is computed once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <string>

#include "../../include/omni3d.h"
#include "driver.hpp"
#include "host.hpp"

namespace omni3d {
namespace {

struct InitGrid {
  int X, Y, Z;
  int cx, cy, cz;
  double del[3];
  double rho;
};

// One launch per level: a thread per (j, k) column takes the voxels of its
// column at distance lev (i = cx -/+ di).
__global__ void __launch_bounds__(256)
k_initial_level(InitGrid g, int lev, const double* __restrict__ ax,
                const double* __restrict__ ay, const double* __restrict__ az,
                double* __restrict__ pn) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (int64_t)g.Y * g.Z) return;
  const int j = (int)(col % g.Y), k = (int)(col / g.Y);
  const int di = lev - abs(j - g.cy) - abs(k - g.cz);
  if (di < 0) return;
  const int64_t sy = g.X, sz = (int64_t)g.X * g.Y;
  const double* acc[3] = {ax, ay, az};
  const int ctr[3] = {g.cx, g.cy, g.cz};
  const int64_t stride[3] = {1, sy, sz};
  for (int side = -1; side <= 1; side += 2) {
    if (di == 0 && side > 0) break;
    const int i = g.cx + side * di;
    if (i < 0 || i >= g.X) continue;
    const int c[3] = {i, j, k};
    const int64_t v = i + j * sy + k * sz;
    int n = 0;
    double val = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      if (c[a] == ctr[a]) continue;
      const bool low = c[a] < ctr[a];  // octant toward the low face
      const int64_t nb = low ? v + stride[a] : v - stride[a];
      const double t = 0.5 * (acc[a][v] + acc[a][nb]) * g.del[a] * g.rho;
      const double pr = low ? pn[nb] + t : pn[nb] - t;
      val = n == 0 ? pr : val + pr;
      n++;
    }
    if (n == 2) val = val / 2.0;
    if (n == 3) val = val / 3.0;
    if (n > 0) pn[v] = val;
  }
}

__global__ void k_set_first(double* __restrict__ pn, double pref) { pn[0] = pref; }

#define INIT_TRY(expr)                                                                \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return fail(nullptr, OMNI3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

int initial_impl(long X, long Y, long Z, double dx, double dy, double dz, double rho,
                 const double* ax, const double* ay, const double* az, double pref, double* p,
                 double* pn) {
  if (X < 1 || Y < 1 || Z < 1 || X > (1l << 20) || Y > (1l << 20) || Z > (1l << 20) ||
      (int64_t)X * Y * Z >= (1ll << 40))
    return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!ax || !ay || !az || !p || !pn) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, OMNI3D_ERR_NODEVICE, "no HIP device visible");
  const int64_t V = (int64_t)X * Y * Z;
  const size_t b = sizeof(double) * V;
  DevBuf da[3], dpn;
  for (int c = 0; c < 3; c++) INIT_TRY(da[c].alloc(b));
  INIT_TRY(dpn.alloc(b));
  INIT_TRY(hipMemcpy(da[0].as<void>(), ax, b, hipMemcpyHostToDevice));
  INIT_TRY(hipMemcpy(da[1].as<void>(), ay, b, hipMemcpyHostToDevice));
  INIT_TRY(hipMemcpy(da[2].as<void>(), az, b, hipMemcpyHostToDevice));
  INIT_TRY(hipMemcpy(dpn.as<void>(), pn, b, hipMemcpyHostToDevice));
  InitGrid g;
  g.X = (int)X;
  g.Y = (int)Y;
  g.Z = (int)Z;
  g.cx = (int)(X / 2);
  g.cy = (int)(Y / 2);
  g.cz = (int)(Z / 2);
  g.del[0] = dx;
  g.del[1] = dy;
  g.del[2] = dz;
  g.rho = rho;
  // p[0] = pn[0] = pref (initialIntegration.cpp:11-12), before the march:
  // it survives only when voxel 0 is the centre
  hipLaunchKernelGGL(k_set_first, dim3(1), dim3(1), 0, nullptr, dpn.as<double>(), pref);
  INIT_TRY(hipGetLastError());
  const int levels = std::max(g.cx, g.X - 1 - g.cx) + std::max(g.cy, g.Y - 1 - g.cy) +
                     std::max(g.cz, g.Z - 1 - g.cz);
  const int64_t cols = (int64_t)g.Y * g.Z;
  const unsigned int grid = (unsigned int)((cols + 255) / 256);
  for (int lev = 1; lev <= levels; lev++) {
    hipLaunchKernelGGL(k_initial_level, dim3(grid), dim3(256), 0, nullptr, g, lev,
                       da[0].as<double>(), da[1].as<double>(), da[2].as<double>(),
                       dpn.as<double>());
    INIT_TRY(hipGetLastError());
  }
  // p = pn on every voxel (each voxel ends its sweep with p[ind] = pn[ind])
  INIT_TRY(hipMemcpy(pn, dpn.as<void>(), b, hipMemcpyDeviceToHost));
  if (p != pn) std::copy(pn, pn + V, p);
  return OMNI3D_OK;
}

}  // namespace
}  // namespace omni3d

extern "C" {

void initialIntegration(long Xsize, long Ysize, long Zsize, double deltx, double delty,
                        double deltz, double density, double* DuDt, double* DvDt, double* DwDt,
                        double pref, double* p, double* pn, double* pBC) {
  (void)pBC;  // unused by the reference too
  (void)omni3d::initial_impl(Xsize, Ysize, Zsize, deltx, delty, deltz, density, DuDt, DvDt,
                             DwDt, pref, p, pn);
}

int omni3d_initial_integration(long X, long Y, long Z, double dx, double dy, double dz,
                               double density, const double* DuDt, const double* DvDt,
                               const double* DwDt, double pref, double* p, double* pn) {
  return omni3d::initial_impl(X, Y, Z, dx, dy, dz, density, DuDt, DvDt, DwDt, pref, p, pn);
}

}  // extern "C"
