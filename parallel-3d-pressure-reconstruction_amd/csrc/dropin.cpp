// Kernel-level drop-ins for the reference's device entry points
// (kernel.cu:6151 omni3dparallellinesESInner, 6281 ...Weighted, 6414
// ...Select, 6746 devidecountInner, 8352 calCurlofMaterialAccCPU): the same
// arguments as device pointers, one pass over every line x direction pair
// accumulated into the caller's pn / pcountinner.
//
// Each host thread keeps one hidden context whose geometry plan (rays, visit
// counts, tile segments) is reused while the grid, line spacing and
// direction set stay the same, as they do across the reference's iteration
// loop.  A pass runs on the tile engine of the reference's schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "driver.hpp"
#include "host.hpp"
#include "launch.hpp"
#include "tiles.hpp"

using namespace omni3d;

namespace {

struct KernelCache {
  omni3d_ctx* ctx = nullptr;
  ~KernelCache() {
    if (ctx) omni3d_destroy(ctx);
  }
};
thread_local KernelCache g_kc;

bool dims_ok(int X, int Y, int Z) {
  return X >= 2 && Y >= 2 && Z >= 2 && X <= kMaxDim && Y <= kMaxDim && Z <= kMaxDim;
}

// Returns the cached context configured for this call (plan rebuilt only if
// the grid, line spacing or directions changed).
int kernel_context(int X, int Y, int Z, int D, float ls, const float* k1_d, const float* k2_d,
                   const float* k3_d, float dx, float dy, float dz, float rho, int method,
                   float thr, hipStream_t st, omni3d_ctx** out) {
  if (!dims_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "grid dimensions must be in [2, 1023]");
  if (D < 1 || D > 4096) return fail(nullptr, OMNI3D_ERR_ARG, "NoAngles must be in [1, 4096]");
  if (!k1_d || !k2_d || !k3_d) return fail(nullptr, OMNI3D_ERR_ARG, "null direction array");
  std::vector<float> a(D), b(D), c(D);
  hipError_t e = hipMemcpyAsync(a.data(), k1_d, sizeof(float) * D, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(b.data(), k2_d, sizeof(float) * D, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(c.data(), k3_d, sizeof(float) * D, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(nullptr, OMNI3D_ERR_HIP, std::string("direction copy: ") + hipGetErrorString(e));
  omni3d_ctx*& ctx = g_kc.ctx;
  if (ctx && (ctx->prm.nx != X || ctx->prm.ny != Y || ctx->prm.nz != Z ||
              ctx->prm.line_spacing != (double)ls)) {
    omni3d_destroy(ctx);
    ctx = nullptr;
  }
  if (!ctx) {
    omni3d_params P;
    omni3d_default_params(&P);
    P.nx = X;
    P.ny = Y;
    P.nz = Z;
    P.line_spacing = ls;
    P.solver = OMNI3D_SOLVER_NAIVE;
    P.max_iterations = 1;
    int rc = omni3d_create(&P, &ctx);
    if (rc) return rc;
  }
  omni3d_params& P = ctx->prm;
  P.dx = dx;
  P.dy = dy;
  P.dz = dz;
  P.density = rho;
  P.method = method;
  P.threshold = thr;
  P.faithful_select = 1;
  P.precision = OMNI3D_FP64;
  if (!(P.dx > 0 && P.dy > 0 && P.dz > 0)) return fail(nullptr, OMNI3D_ERR_ARG, "grid spacing must be positive");
  ctx->stream = st;
  bool same = (int)(ctx->k.size() / 3) == D;
  for (int i = 0; same && i < D; i++)
    same = ctx->k[3 * i] == a[i] && ctx->k[3 * i + 1] == b[i] && ctx->k[3 * i + 2] == c[i];
  if (!same) {
    int rc = omni3d_set_directions(ctx, a.data(), b.data(), c.data(), D);
    if (rc) return fail(nullptr, rc, ctx->err);
  }
  *out = ctx;
  return OMNI3D_OK;
}

// One pass of `method` over every ray, added into pn / pcountinner.
int es_pass(int method, int X, int Y, int Z, int D, float ls, const float* k1_d, const float* k2_d,
            const float* k3_d, float dx, float dy, float dz, float rho, const float* ax,
            const float* ay, const float* az, const float* curl, float thr, const float* p,
            float* pn, float* pc, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!ax || !ay || !az || !p || !pn || !pc)
    return fail(nullptr, OMNI3D_ERR_ARG, "null field / pressure / count pointer");
  if (method != OMNI3D_METHOD_PLAIN && !curl) return fail(nullptr, OMNI3D_ERR_ARG, "null curl");
  omni3d_ctx* ctx = nullptr;
  int rc = kernel_context(X, Y, Z, D, ls, k1_d, k2_d, k3_d, dx, dy, dz, rho, method, thr, st, &ctx);
  if (rc) return rc;
  rc = omni3d_set_field(ctx, ax, ay, az, nullptr, OMNI3D_F32, OMNI3D_MEM_DEVICE);
  if (rc) return fail(nullptr, rc, ctx->err);
  rc = build_plan(ctx);
  if (rc) return fail(nullptr, rc, ctx->err);
  const Plan& pl = *ctx->plan;
  const int64_t V = (int64_t)X * Y * Z;
  DevScalars* sc = ctx->scal.as<DevScalars>();
  // fixed-point sums with unit weights: plain (counts of the plan) and
  // MiniCurl (counts of its own paths, counted by the walk)
  const bool minicurl = method == OMNI3D_METHOD_MINICURL;
  const bool fixed = method == OMNI3D_METHOD_PLAIN || minicurl;
  OMNI_TRY_HIP(ctx, ctx->pa.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->srcf.alloc((size_t)((V + 3) / 4) * 4));
  OMNI_TRY_HIP(ctx, (launch_convert<float, double>(p, ctx->pa.as<double>(), V, st)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(sc, sizeof(DevScalars) / 4, 0u, st));
  OMNI_TRY_HIP(ctx, launch_absmax<double>(ctx->field.as<double>(), 4 * V, &sc->amax_bits, st));
  OMNI_TRY_HIP(ctx, launch_absmax<double>(ctx->pa.as<double>(), V, &sc->pmax_bits, st));
  const double gfac = (double)rho * std::max((double)dx, std::max((double)dy, (double)dz));
  OMNI_TRY_HIP(ctx, launch_compute_scale(sc, std::fabs(gfac), (double)pl.lmax, (double)pl.cmax,
                                         1.0, st));
  const double* aux = nullptr;
  if (method != OMNI3D_METHOD_PLAIN) {
    OMNI_TRY_HIP(ctx, ctx->aux.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, (launch_convert<float, double>(curl, ctx->aux.as<double>(), V, st)));
    aux = ctx->aux.as<double>();
  }
  if (minicurl) {
    OMNI_TRY_HIP(ctx, ctx->wcnt.alloc(sizeof(unsigned int) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wcnt.as<void>(), V, 0u, st));
  }
  if (!fixed) {
    OMNI_TRY_HIP(ctx, ctx->sumd.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, ctx->wsum.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sumd.as<void>(), 2 * V, 0u, st));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wsum.as<void>(), 2 * V, 0u, st));
  } else {
    OMNI_TRY_HIP(ctx, ctx->sum.alloc(sizeof(long long) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sum.as<void>(), 2 * V, 0u, st));
  }
  WalkParams wp = walk_params(ctx->prm);
  // the pass runs on the tile engine (rule replay, carry chain, tile pass);
  // OMNI3D_MASKED=naive walks one lane per ray with global atomics
  const char* menv = std::getenv("OMNI3D_MASKED");
  // SelectFixedBC's step rule differs from the plan's staircase: one lane per ray
  if (!(menv && std::strcmp(menv, "naive") == 0) && method != OMNI3D_METHOD_SELECT_FIXEDBC &&
      !minicurl) {
    const size_t nv = (size_t)std::max<int64_t>(pl.nval, 1);
    OMNI_TRY_HIP(ctx, ctx->mk_e.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_s.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_c.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_r.alloc(sizeof(uint32_t) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_f.alloc(sizeof(uint32_t) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_wc.alloc(sizeof(double) * (fixed ? 1 : nv)));
    OMNI_TRY_HIP(ctx, ctx->mk_wn.alloc(sizeof(double) * (fixed ? 1 : nv)));
    const double* pa = ctx->pa.as<double>();
    OMNI_TRY_HIP(ctx, launch_mask_walk<double>(method, pl.rays.as<RayRec>(), pl.nrays,
                                               pl.dirs.as<DirConst>(), pl.g, pl.tg, wp,
                                               ctx->field.as<void>(), aux,
                                               pl.val_off.as<unsigned long long>(), pl.ntiles,
                                               pl.tile_ptr.as<unsigned long long>(),
                                               pl.segs.as<SegRec>(), sc, ctx->mk_e.as<double>(),
                                               ctx->mk_s.as<double>(), ctx->mk_r.as<uint32_t>(),
                                               ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                               ctx->mk_wn.as<double>(), nullptr,
                                               fixed ? nullptr : ctx->wsum.as<double>(),
                                               ctx->srcf.as<uint8_t>(), st));
    OMNI_TRY_HIP(ctx, launch_mask_carry(pl.nrays, pl.nseg.as<unsigned int>(),
                                        pl.val_off.as<unsigned long long>(),
                                        ctx->mk_e.as<double>(), ctx->mk_s.as<double>(),
                                        ctx->mk_r.as<uint32_t>(), ctx->mk_f.as<uint32_t>(), pa, pa,
                                        1, sc, ctx->mk_c.as<double>(), st));
    OMNI_TRY_HIP(ctx, launch_tile_masked<double>(method, pl.ntiles,
                                                 pl.tile_ptr.as<unsigned long long>(),
                                                 pl.segs.as<SegRec>(), pl.g, pl.tg, wp,
                                                 ctx->field.as<void>(), aux,
                                                 ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                                 ctx->mk_wn.as<double>(), ctx->mk_c.as<double>(),
                                                 pa, pa, 1, sc,
                                                 fixed ? ctx->sum.as<long long>() : nullptr,
                                                 fixed ? nullptr : ctx->sumd.as<double>(), st));
  } else {
    OMNI_TRY_HIP(ctx, launch_walk_naive<double>(method, pl.rays.as<RayRec>(), pl.nrays,
                                                pl.dirs.as<DirConst>(), pl.g, wp,
                                                ctx->field.as<void>(), aux, ctx->pa.as<double>(),
                                                ctx->pa.as<double>(), 1, sc,
                                                fixed ? ctx->sum.as<long long>() : nullptr,
                                                minicurl ? ctx->wcnt.as<unsigned int>() : nullptr,
                                                fixed ? nullptr : ctx->sumd.as<double>(),
                                                fixed ? nullptr : ctx->wsum.as<double>(),
                                                ctx->srcf.as<uint8_t>(), st));
  }
  OMNI_TRY_HIP(ctx, launch_accumulate_out(V, fixed ? 1 : 0,
                                          fixed ? ctx->sum.as<long long>() : nullptr,
                                          minicurl ? ctx->wcnt.as<unsigned int>()
                                                   : pl.cnt.as<unsigned int>(),
                                          fixed ? nullptr : ctx->sumd.as<double>(),
                                          fixed ? nullptr : ctx->wsum.as<double>(), sc, pn, pc, st));
  // the pass reads the context's work buffers: finish before another call reuses them
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  DevScalars hs;
  OMNI_TRY_HIP(ctx, hipMemcpy(&hs, sc, sizeof(hs), hipMemcpyDeviceToHost));
  if (hs.status) return fail(nullptr, hs.status, "non-finite value in the field or pressure");
  return OMNI3D_OK;
}

}  // namespace

extern "C" {

int omni3d_es_inner(int X, int Y, int Z, int NoAngles, float linespacing, const float* k1_d,
                    const float* k2_d, const float* k3_d, float deltx, float delty, float deltz,
                    float density, const float* DuDt, const float* DvDt, const float* DwDt,
                    const float* p, float* pn, float* pcountinner, void* hip_stream) {
  return es_pass(OMNI3D_METHOD_PLAIN, X, Y, Z, NoAngles, linespacing, k1_d, k2_d, k3_d, deltx,
                 delty, deltz, density, DuDt, DvDt, DwDt, nullptr, 0.0f, p, pn, pcountinner,
                 hip_stream);
}

int omni3d_es_inner_weighted(int X, int Y, int Z, int NoAngles, float linespacing,
                             const float* k1_d, const float* k2_d, const float* k3_d,
                             float deltx, float delty, float deltz, float density,
                             const float* DuDt, const float* DvDt, const float* DwDt,
                             const float* curl, const float* p, float* pn, float* pcountinner,
                             void* hip_stream) {
  return es_pass(OMNI3D_METHOD_WEIGHTED, X, Y, Z, NoAngles, linespacing, k1_d, k2_d, k3_d, deltx,
                 delty, deltz, density, DuDt, DvDt, DwDt, curl, 0.0f, p, pn, pcountinner,
                 hip_stream);
}

int omni3d_es_inner_select(int X, int Y, int Z, int NoAngles, float linespacing,
                           const float* k1_d, const float* k2_d, const float* k3_d, float deltx,
                           float delty, float deltz, float density, const float* DuDt,
                           const float* DvDt, const float* DwDt, const float* curl, const float* p,
                           float* pn, float* pcountinner, float threshold, void* hip_stream) {
  return es_pass(OMNI3D_METHOD_SELECT, X, Y, Z, NoAngles, linespacing, k1_d, k2_d, k3_d, deltx,
                 delty, deltz, density, DuDt, DvDt, DwDt, curl, threshold, p, pn, pcountinner,
                 hip_stream);
}

int omni3d_es_inner_select_fixedbc(int X, int Y, int Z, int NoAngles, float linespacing,
                                   const float* k1_d, const float* k2_d, const float* k3_d,
                                   float deltx, float delty, float deltz, float density,
                                   const float* DuDt, const float* DvDt, const float* DwDt,
                                   const float* curl, const float* p, float* pn,
                                   float* pcountinner, float threshold, void* hip_stream) {
  return es_pass(OMNI3D_METHOD_SELECT_FIXEDBC, X, Y, Z, NoAngles, linespacing, k1_d, k2_d, k3_d,
                 deltx, delty, deltz, density, DuDt, DvDt, DwDt, curl, threshold, p, pn,
                 pcountinner, hip_stream);
}

int omni3d_es_inner_weighted_minicurl(int X, int Y, int Z, int NoAngles, float linespacing,
                                      const float* k1_d, const float* k2_d, const float* k3_d,
                                      float deltx, float delty, float deltz, float density,
                                      const float* DuDt, const float* DvDt, const float* DwDt,
                                      const float* curl, const float* p, float* pn,
                                      float* pcountinner, void* hip_stream) {
  return es_pass(OMNI3D_METHOD_MINICURL, X, Y, Z, NoAngles, linespacing, k1_d, k2_d, k3_d, deltx,
                 delty, deltz, density, DuDt, DvDt, DwDt, curl, 0.0f, p, pn, pcountinner,
                 hip_stream);
}

int omni3d_devide_count_inner(int X, int Y, int Z, float* p, float* pn, const float* pcountinner,
                              void* hip_stream) {
  if (X < 1 || Y < 1 || Z < 1) return fail(nullptr, OMNI3D_ERR_ARG, "bad grid dimensions");
  if (!p || !pn || !pcountinner) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  hipError_t e = launch_divide_count((int64_t)X * Y * Z, p, pn, pcountinner,
                                     static_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail(nullptr, OMNI3D_ERR_HIP, hipGetErrorString(e));
  return OMNI3D_OK;
}

int omni3d_curl(int X, int Y, int Z, float deltx, float delty, float deltz, const float* DuDt,
                const float* DvDt, const float* DwDt, const float* mask, float* curl,
                void* hip_stream) {
  if (!dims_ok(X, Y, Z)) return fail(nullptr, OMNI3D_ERR_ARG, "grid dimensions must be in [2, 1023]");
  if (!DuDt || !DvDt || !DwDt || !curl) return fail(nullptr, OMNI3D_ERR_ARG, "null pointer");
  hipError_t e = launch_curl_f32(X, Y, Z, deltx, delty, deltz, DuDt, DvDt, DwDt, mask, curl,
                                 static_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return fail(nullptr, OMNI3D_ERR_HIP, hipGetErrorString(e));
  return OMNI3D_OK;
}

}  // extern "C"
