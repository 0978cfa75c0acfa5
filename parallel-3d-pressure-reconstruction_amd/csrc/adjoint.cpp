// C-ABI of the adjoint pair: omni3d_forward (g = grad p -> p) and
// omni3d_backward (dL/dp -> dL/dg) of the plain / plain-mask reconstruction
// with zero initial pressure and the context's iteration count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "driver.hpp"
#include "host.hpp"
#include "launch.hpp"
#include "tiles.hpp"

using namespace omni3d;

namespace {

int adjoint_checks(omni3d_ctx* ctx) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  const omni3d_params& P = ctx->prm;
  if (P.method != OMNI3D_METHOD_PLAIN && P.method != OMNI3D_METHOD_PLAIN_MASK)
    return fail(ctx, OMNI3D_ERR_ARG,
                "the adjoint exists for the linear methods (plain, plain-mask) only");
  if (P.traversal != OMNI3D_TRAVERSAL_STAIRCASE)
    return fail(ctx, OMNI3D_ERR_ARG, "the adjoint is built for the staircase traversal");
  if (ctx->k.empty()) return fail(ctx, OMNI3D_ERR_STATE, "no directions set");
  return OMNI3D_OK;
}

int io_dtype(const omni3d_ctx* ctx) {
  return ctx->prm.precision == OMNI3D_FP64 ? OMNI3D_F64 : OMNI3D_F32;
}

// Merged words of sorted CSR rows in the forward iteration's layout (groups
// of 8 chunks, 4 from 2^21 sources on, sized to about equal entries and as
// many as fill the CUs evenly) when they fit it -- sources below 2^22, rows
// shorter than 4096 entries (int64 row sums); *dense tells whether they
// were built, *gchunks / *ngroups / grow the groups.
int build_row_words(omni3d_ctx* ctx, int nS, unsigned long long n, const DevBuf& tptr,
                    const DevBuf& key, DevBuf* tdent, DevBuf* twptr, DevBuf* grow,
                    int* gchunks, int* ngroups, bool* dense) {
  hipStream_t st = ctx->stream;
  *dense = false;
  std::vector<unsigned long long> hp((size_t)nS + 1);
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(hp.data(), tptr.as<void>(),
                                   sizeof(unsigned long long) * (nS + 1), hipMemcpyDeviceToHost,
                                   st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  unsigned long long mx = 0;
  for (int r = 0; r < nS; r++) mx = std::max(mx, hp[r + 1] - hp[r]);
  if (!(nS < (1 << 22) && mx < 4096 && n > 0 && n < (1ull << 31))) return OMNI3D_OK;
  const int grp = nS < (1 << 21) ? 8 : 4;
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  std::vector<int> hgrow, hrg((size_t)nS);
  balance_groups(hp, 64 * grp, std::max(ncu, 1), &hgrow, &hrg);
  const int ng = (int)hgrow.size() - 1;
  DevBuf rgroup;
  OMNI_TRY_HIP(ctx, grow->alloc(sizeof(int) * hgrow.size()));
  OMNI_TRY_HIP(ctx, rgroup.alloc(sizeof(int) * (size_t)nS));
  OMNI_TRY_HIP(ctx, hipMemcpy(grow->as<void>(), hgrow.data(), sizeof(int) * hgrow.size(),
                              hipMemcpyHostToDevice));
  OMNI_TRY_HIP(ctx, hipMemcpy(rgroup.as<void>(), hrg.data(), sizeof(int) * (size_t)nS,
                              hipMemcpyHostToDevice));
  DevBuf ka, kb, tmp, flag, excl;
  OMNI_TRY_HIP(ctx, ka.alloc(sizeof(unsigned long long) * n));
  OMNI_TRY_HIP(ctx, kb.alloc(sizeof(unsigned long long) * n));
  OMNI_TRY_HIP(ctx, tmp.alloc(dense_tmp_bytes((int64_t)n)));
  OMNI_TRY_HIP(ctx, tdent->alloc(sizeof(unsigned int) * (n + 8)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(tdent->as<void>(), (int64_t)n + 8, 0u, st));
  OMNI_TRY_HIP(ctx, build_dense_entries(nS, (int64_t)n, grp, tptr.as<unsigned long long>(),
                                        key.as<unsigned long long>(), ka.as<unsigned long long>(),
                                        kb.as<unsigned long long>(), tmp.as<void>(), tmp.size(),
                                        tdent->as<unsigned int>(), st, rgroup.as<int>(),
                                        grow->as<int>(), ng));
  OMNI_TRY_HIP(ctx, flag.alloc(sizeof(unsigned int) * n));
  OMNI_TRY_HIP(ctx, excl.alloc(sizeof(unsigned int) * n));
  OMNI_TRY_HIP(ctx, twptr->alloc(sizeof(unsigned int) * (ng + 1)));
  unsigned int nw = 0;
  OMNI_TRY_HIP(ctx, merge_dense_entries(nS, (int64_t)n, grp, tptr.as<unsigned long long>(),
                                        kb.as<unsigned long long>(), flag.as<unsigned int>(),
                                        excl.as<unsigned int>(), tmp.as<void>(), tmp.size(),
                                        tdent->as<unsigned int>(), twptr->as<unsigned int>(), &nw,
                                        st, grow->as<int>(), ng));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  *gchunks = grp;
  *ngroups = ng;
  *dense = true;
  return OMNI3D_OK;
}

// Transposed source rows of the plain method, kept with the plan: row s'
// lists the source voxels that rays entering at s' visit after their entry
// (the entries of the forward source rows, regrouped by their source).
int build_transposed_rows(omni3d_ctx* ctx) {
  Plan& pl = *ctx->plan;
  if (pl.trows) return OMNI3D_OK;
  hipStream_t st = ctx->stream;
  const int nS = pl.nS;
  DevBuf tcnt, key;
  OMNI_TRY_HIP(ctx, tcnt.alloc(sizeof(unsigned int) * (nS + 1)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(tcnt.as<void>(), nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_adj_transpose_rows(pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                              pl.g, pl.sid.as<int>(), tcnt.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, pl.tptr.alloc(sizeof(unsigned long long) * (nS + 1)));
  OMNI_TRY_HIP(ctx, launch_scan_counts(tcnt.as<unsigned int>(), nS,
                                       pl.tptr.as<unsigned long long>(), st));
  unsigned long long n = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&n, pl.tptr.as<unsigned long long>() + nS, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if ((int64_t)n != pl.nnz)
    return fail(ctx, OMNI3D_ERR_STATE, "transposed source rows disagree with the source rows");
  OMNI_TRY_HIP(ctx, key.alloc(sizeof(unsigned long long) * (n + 1)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(tcnt.as<void>(), nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_adj_transpose_fill(pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                              pl.g, pl.sid.as<int>(),
                                              pl.tptr.as<unsigned long long>(),
                                              tcnt.as<unsigned int>(),
                                              key.as<unsigned long long>(), st));
  OMNI_TRY_HIP(ctx, launch_sort_rows(nS, pl.tptr.as<unsigned long long>(),
                                     key.as<unsigned long long>(), st));
  {
    bool dense = false;
    const int rc = build_row_words(ctx, nS, n, pl.tptr, key, &pl.tdent, &pl.twptr, &pl.tgrow,
                                   &pl.tgroup, &pl.tngroups, &dense);
    if (rc) return rc;
    pl.tdense = dense;
  }
  OMNI_TRY_HIP(ctx, pl.tent.alloc(sizeof(unsigned int) * (n + 1)));
  OMNI_TRY_HIP(ctx, launch_adj_narrow(key.as<unsigned long long>(), (int64_t)n,
                                      pl.tent.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  pl.trows = true;
  return OMNI3D_OK;
}

// Sources and transposed rows of the plain-mask rule for the current mask:
// S = the run-start voxels (numbered in voxel order), row s' = the run-start
// voxels that runs starting at s' visit (sorted).  No self visits: diag = 0.
using MaskRows = omni3d_ctx::AdjMask;

int build_mask_rows(omni3d_ctx* ctx, const double* mask, MaskRows* mr) {
  const Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(ctx->prm);
  DevBuf flag;
  OMNI_TRY_HIP(ctx, flag.alloc((size_t)((V + 3) / 4) * 4));
  OMNI_TRY_HIP(ctx, launch_fill_u32(flag.as<void>(), (V + 3) / 4, 0u, st));
  OMNI_TRY_HIP(ctx, launch_adj_mask_starts(pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                           pl.g, mask, flag.as<uint8_t>(), st));
  // S numbered in voxel order on the device: flags -> counts -> scan -> ids
  DevBuf fcnt, foff, ftmp;
  OMNI_TRY_HIP(ctx, fcnt.alloc(sizeof(unsigned int) * (size_t)V));
  OMNI_TRY_HIP(ctx, foff.alloc(sizeof(unsigned long long) * ((size_t)V + 1)));
  OMNI_TRY_HIP(ctx, ftmp.alloc(scan_u32_tmp_bytes(V)));
  OMNI_TRY_HIP(ctx, launch_flag_counts(flag.as<uint8_t>(), V, fcnt.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, launch_scan_u32(fcnt.as<unsigned int>(), V, foff.as<unsigned long long>(), 0,
                                    ftmp.as<void>(), ftmp.size(), st));
  unsigned long long nsrc = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nsrc, foff.as<unsigned long long>() + V, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (nsrc >= (1ull << 31)) return fail(ctx, OMNI3D_ERR_ARG, "too many run starts");
  const int nS = (int)nsrc;
  mr->nS = nS;
  mr->tdense = false;
  if (nS == 0) return OMNI3D_OK;
  OMNI_TRY_HIP(ctx, mr->sid.alloc(sizeof(int) * V));
  OMNI_TRY_HIP(ctx, mr->slist.alloc(sizeof(int) * nS));
  OMNI_TRY_HIP(ctx, launch_flag_ids(flag.as<uint8_t>(), V, foff.as<unsigned long long>(),
                                    mr->sid.as<int>(), mr->slist.as<int>(), st));
  DevBuf cnt, key;
  OMNI_TRY_HIP(ctx, cnt.alloc(sizeof(unsigned int) * (nS + 1)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(cnt.as<void>(), nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_adj_mask_rows(0, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                         pl.g, mask, mr->sid.as<int>(), nullptr,
                                         cnt.as<unsigned int>(), nullptr, st));
  OMNI_TRY_HIP(ctx, mr->tptr.alloc(sizeof(unsigned long long) * (nS + 1)));
  OMNI_TRY_HIP(ctx, launch_scan_counts(cnt.as<unsigned int>(), nS,
                                       mr->tptr.as<unsigned long long>(), st));
  unsigned long long n = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&n, mr->tptr.as<unsigned long long>() + nS, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  OMNI_TRY_HIP(ctx, key.alloc(sizeof(unsigned long long) * (n + 1)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(cnt.as<void>(), nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_adj_mask_rows(1, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                         pl.g, mask, mr->sid.as<int>(),
                                         mr->tptr.as<unsigned long long>(), cnt.as<unsigned int>(),
                                         key.as<unsigned long long>(), st));
  OMNI_TRY_HIP(ctx, launch_sort_rows(nS, mr->tptr.as<unsigned long long>(),
                                     key.as<unsigned long long>(), st));
  {
    bool dense = false;
    const int rc = build_row_words(ctx, nS, n, mr->tptr, key, &mr->tdent, &mr->twptr, &mr->tgrow,
                                   &mr->tgroup, &mr->tngroups, &dense);
    if (rc) return rc;
    mr->tdense = dense;
  }
  OMNI_TRY_HIP(ctx, mr->tent.alloc(sizeof(unsigned int) * (n + 1)));
  OMNI_TRY_HIP(ctx, launch_adj_narrow(key.as<unsigned long long>(), (int64_t)n,
                                      mr->tent.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, mr->diag.alloc(sizeof(unsigned int) * nS));
  OMNI_TRY_HIP(ctx, launch_fill_u32(mr->diag.as<void>(), nS, 0u, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  return OMNI3D_OK;
}

}  // namespace

extern "C" {

int omni3d_forward(omni3d_ctx* ctx, const void* grad_d, void* p_d) {
  int rc = adjoint_checks(ctx);
  if (rc) return rc;
  if (!grad_d || !p_d) return fail(ctx, OMNI3D_ERR_ARG, "null pointer");
  const int64_t V = voxels(ctx->prm);
  const int dt = io_dtype(ctx);
  const size_t esz = elem_size(dt);
  const char* g = static_cast<const char*>(grad_d);
  // the gradient becomes the context's field, integrated with density -1
  // (-rho * a = g); the mask set before stays
  const bool had_mask = ctx->have_mask;
  const uint64_t epoch = ctx->mask_epoch;
  rc = omni3d_set_field(ctx, g, g + esz * V, g + 2 * esz * V, nullptr, dt, OMNI3D_MEM_DEVICE);
  ctx->have_mask = had_mask;
  ctx->mask_epoch = epoch;
  if (rc) return rc;
  const omni3d_params saved = ctx->prm;
  ctx->prm.density = -1.0;
  ctx->prm.tolerance = 0.0;
  ctx->prm.seed_pn_with_p = 0;
  // zero initial pressure, written into the output buffer
  hipError_t e = dt == OMNI3D_F64 ? launch_fill_f64(static_cast<double*>(p_d), V, 0.0, ctx->stream)
                                  : launch_fill_u32(p_d, V, 0u, ctx->stream);
  if (e != hipSuccess) {
    ctx->prm = saved;
    return fail(ctx, OMNI3D_ERR_HIP, hipGetErrorString(e));
  }
  rc = omni3d_reconstruct(ctx, p_d, nullptr, dt, OMNI3D_MEM_DEVICE, nullptr);
  ctx->prm = saved;
  return rc;
}

int omni3d_backward(omni3d_ctx* ctx, const void* dldp_d, void* dldgrad_d) {
  int rc = adjoint_checks(ctx);
  if (rc) return rc;
  if (!dldp_d || !dldgrad_d) return fail(ctx, OMNI3D_ERR_ARG, "null pointer");
  rc = build_plan(ctx);
  if (rc) return rc;
  const omni3d_params& P = ctx->prm;
  const Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  const int dt = io_dtype(ctx);
  const bool mask_rule = P.method == OMNI3D_METHOD_PLAIN_MASK;
  const double* mask = nullptr;
  if (mask_rule) {
    OMNI_TRY_HIP(ctx, ctx->aux.alloc(sizeof(double) * V));
    if (ctx->have_mask)
      OMNI_TRY_HIP(ctx, (launch_convert<double, double>(ctx->mask.as<double>(), ctx->aux.as<double>(), V, st)));
    else
      OMNI_TRY_HIP(ctx, launch_fill_f64(ctx->aux.as<double>(), V, 1.0, st));
    mask = ctx->aux.as<double>();
  }
  // visit counts C of the method (the plain-mask rule's: built once per
  // mask and plan, with its run-start rows below)
  const unsigned int* cnt = pl.cnt.as<unsigned int>();
  MaskRows& mr = ctx->adj_mask;
  if (mask_rule) {
    uint64_t mepoch = 0;
    rc = mask_epoch_now(ctx, &mepoch);
    if (rc) return rc;
    if (mr.mask_epoch != mepoch || mr.plan_epoch != pl.epoch) {
      mr.counts = mr.rows = false;
      mr.mask_epoch = mepoch;
      mr.plan_epoch = pl.epoch;
    }
    if (!mr.counts) {
      OMNI_TRY_HIP(ctx, mr.cnt.alloc(sizeof(unsigned int) * V));
      OMNI_TRY_HIP(ctx, launch_fill_u32(mr.cnt.as<void>(), V, 0u, st));
      OMNI_TRY_HIP(ctx, launch_adj_count(pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                         pl.g, mask, mr.cnt.as<unsigned int>(), st));
      mr.counts = true;
    }
    cnt = mr.cnt.as<unsigned int>();
  }
  for (DevBuf* b : {&ctx->adj_y, &ctx->adj_z, &ctx->adj_u, &ctx->adj_t})
    OMNI_TRY_HIP(ctx, b->alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->adj_g.alloc(sizeof(double) * 3 * V));
  double* y = ctx->adj_y.as<double>();
  double* z = ctx->adj_z.as<double>();
  double* u = ctx->adj_u.as<double>();
  double* t = ctx->adj_t.as<double>();
  double* dg = ctx->adj_g.as<double>();
  if (dt == OMNI3D_F64)
    OMNI_TRY_HIP(ctx, (launch_convert<double, double>(static_cast<const double*>(dldp_d), y, V, st)));
  else
    OMNI_TRY_HIP(ctx, (launch_convert<float, double>(static_cast<const float*>(dldp_d), y, V, st)));
  OMNI_TRY_HIP(ctx, (launch_convert<double, double>(y, z, V, st)));
  // the transposed traversals (plain and plain-mask) run on the tile engine
  // with the affine solver (the NAIVE solver keeps one lane per ray: the
  // cross-check)
  omni3d_params pg = P;
  pg.density = -1.0;
  const WalkParams wp = walk_params(pg);
  const bool tiles = P.solver == OMNI3D_SOLVER_AFFINE && pl.ntiles > 0;
  if (tiles) OMNI_TRY_HIP(ctx, ctx->adj_seg.alloc(sizeof(double) * std::max<int64_t>(pl.nval, 1)));
  if (tiles && mask_rule)
    OMNI_TRY_HIP(ctx, ctx->adj_span.alloc((size_t)std::max<int64_t>(pl.nval, 1)));
  // the plain-mask rule's transposed traversals on the tile engine: the
  // entering-run sums per segment, the reverse scan per ray, then the tiles'
  // gather (what 3) or scatter (what 2)
  auto mask_tiles = [&](const double* uu, int what, double* out) -> hipError_t {
    hipError_t e = launch_adjm_tiles(0, pl.ntiles, pl.tile_ptr1.as<unsigned long long>(),
                                     pl.segs_slot.as<SegRec>(), nullptr, nullptr, 0, nullptr,
                                     nullptr, pl.g, pl.tg, wp, uu, mask,
                                     ctx->adj_seg.as<double>(), ctx->adj_span.as<uint8_t>(),
                                     nullptr, st);
    if (e == hipSuccess)
      e = launch_adjm_tiles(1, pl.ntiles, nullptr, nullptr, nullptr, nullptr, pl.nrays,
                            pl.nseg.as<unsigned int>(), pl.val_off.as<unsigned long long>(), pl.g,
                            pl.tg, wp, nullptr, nullptr, ctx->adj_seg.as<double>(),
                            ctx->adj_span.as<uint8_t>(), nullptr, st);
    if (e == hipSuccess)
      e = launch_adjm_tiles(what, pl.ntiles, nullptr, nullptr, pl.tile_ptr.as<unsigned long long>(),
                            pl.segs.as<SegRec>(), 0, nullptr, nullptr, pl.g, pl.tg, wp, uu, mask,
                            ctx->adj_seg.as<double>(), nullptr, out, st);
    return e;
  };
  // t = the ray sums of u gathered into the rays' sources
  auto gather = [&](const double* uu, double* tt) -> hipError_t {
    if (!tiles)
      return launch_adj_gather(mask_rule, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                               pl.g, mask, uu, tt, st);
    if (mask_rule) return mask_tiles(uu, 3, tt);
    hipError_t e = launch_adj_tiles(0, pl.ntiles, pl.tile_ptr1.as<unsigned long long>(),
                                    pl.segs_slot.as<SegRec>(), nullptr, nullptr, nullptr, 0,
                                    nullptr, nullptr, pl.g, pl.tg, wp, uu,
                                    ctx->adj_seg.as<double>(), nullptr, nullptr, st);
    if (e == hipSuccess)
      e = launch_adj_tiles(1, pl.ntiles, nullptr, nullptr, nullptr, nullptr, pl.rays.as<RayRec>(),
                           pl.nrays, pl.nseg.as<unsigned int>(),
                           pl.val_off.as<unsigned long long>(), pl.g, pl.tg, wp, nullptr,
                           ctx->adj_seg.as<double>(), tt, nullptr, st);
    return e;
  };
  // z = sum_{j<K} (A^T)^j y
  if (P.solver == OMNI3D_SOLVER_AFFINE) {
    // after the first gather only the source voxels (ray entries; for the
    // plain-mask rule the run starts) carry values, so the Horner steps run
    // over the nS sources with the transposed source rows
    int nS = pl.nS;
    const int* slist = pl.slist.as<int>();
    const unsigned long long* tptr = pl.tptr.as<unsigned long long>();
    const unsigned int* tent = pl.tent.as<unsigned int>();
    const unsigned int* diagp = pl.diag.as<unsigned int>();
    if (P.max_iterations > 1) {
      if (mask_rule) {
        if (!mr.rows) {
          rc = build_mask_rows(ctx, mask, &mr);
          if (rc) return rc;
          mr.rows = true;
        }
        nS = mr.nS;
        slist = mr.slist.as<int>();
        tptr = mr.tptr.as<unsigned long long>();
        tent = mr.tent.as<unsigned int>();
        diagp = mr.diag.as<unsigned int>();
      } else {
        rc = build_transposed_rows(ctx);
        if (rc) return rc;
        tptr = pl.tptr.as<unsigned long long>();
        tent = pl.tent.as<unsigned int>();
      }
    }
    if (P.max_iterations > 1 && nS > 0) {
      OMNI_TRY_HIP(ctx, ctx->adj_s.alloc(sizeof(double) * 7 * (size_t)nS));
      double* ts = ctx->adj_s.as<double>();
      double* h[2] = {ts + nS, ts + 2 * (size_t)nS};
      double* w[2] = {ts + 3 * (size_t)nS, ts + 4 * (size_t)nS};
      double* cs = ts + 5 * (size_t)nS;
      double* w3 = ts + 6 * (size_t)nS;
      OMNI_TRY_HIP(ctx, launch_adj_scale(V, y, cnt, u, st));
      OMNI_TRY_HIP(ctx, launch_fill_f64(t, V, 0.0, st));
      OMNI_TRY_HIP(ctx, gather(u, t));
      OMNI_TRY_HIP(ctx, launch_adj_src_init(nS, slist, t, cnt, ts, h[0], w[0], cs, st));
      const int M = P.max_iterations - 2;
      // the transposed rows as merged words: the plain method's (with the
      // plan) or the plain-mask rule's (with the mask)
      const bool tdense = mask_rule ? mr.tdense : pl.tdense;
      const int tgroup = mask_rule ? mr.tgroup : pl.tgroup;
      const int tngroups = mask_rule ? mr.tngroups : pl.tngroups;
      const DevBuf& twptr = mask_rule ? mr.twptr : pl.twptr;
      const DevBuf& tdent = mask_rule ? mr.tdent : pl.tdent;
      const DevBuf& tgrow = mask_rule ? mr.tgrow : pl.tgrow;
      if (tdense && M > 0) {
        // w <- (t + diag w + sum w) / C is the forward iteration's form: the
        // grouped kernel on the transposed merged words (exact fixed-point
        // row sums), t as its source term, w over three buffers (a source
        // with C = 0 keeps w = 0, as in the one-wavefront-per-row steps)
        const unsigned int gg = grouped_grid_n(tngroups);
        OMNI_TRY_HIP(ctx, ctx->adj_sc.alloc(sizeof(DevScalars)));
        OMNI_TRY_HIP(ctx, ctx->adj_part.alloc(sizeof(double) * 4 * (size_t)gg));
        OMNI_TRY_HIP(ctx, ctx->adj_bs.alloc(sizeof(long long) * (size_t)nS));
        DevScalars* asc = ctx->adj_sc.as<DevScalars>();
        OMNI_TRY_HIP(ctx, launch_fill_u32(asc, sizeof(DevScalars) / 4, 0u, st));
        OMNI_TRY_HIP(ctx, launch_absmax<double>(ts, nS, &asc->pmax_bits, st));
        OMNI_TRY_HIP(ctx, launch_adj_fixed(nS, ts, asc, ctx->adj_bs.as<long long>(), st));
        OMNI_TRY_HIP(ctx, launch_absmax<double>(w[0], nS, &asc->itmax_bits[0], st));
        double* pb3[3] = {w[0], w[1], w3};
        for (int k = 1; k <= M; k++)
          OMNI_TRY_HIP(ctx, launch_iterate_dense(tgroup, 0, k, nS, nullptr,
                                                 twptr.as<unsigned int>(),
                                                 tdent.as<unsigned int>(), diagp, slist, cnt,
                                                 ctx->adj_bs.as<long long>(), nullptr, 0,
                                                 pb3[(k - 1) % 3], pb3[k % 3], asc,
                                                 ctx->adj_part.as<double>(), M, 0.0, st, nullptr,
                                                 nullptr, nullptr, tgrow.as<int>(), tngroups));
        OMNI_TRY_HIP(ctx, launch_adj_src_out(nS, slist, pb3[M % 3], cs, z, st));
      } else {
        int cur = 0;
        for (int m = 2; m < P.max_iterations; m++) {
          OMNI_TRY_HIP(ctx, launch_adj_src_iter(nS, tptr, tent, diagp, cs, ts, w[cur], h[cur ^ 1],
                                                w[cur ^ 1], st));
          cur ^= 1;
        }
        OMNI_TRY_HIP(ctx, launch_adj_src_out(nS, slist, h[cur], nullptr, z, st));
      }
    }
  } else {
    for (int m = 1; m < P.max_iterations; m++) {
      OMNI_TRY_HIP(ctx, launch_adj_scale(V, z, cnt, u, st));
      OMNI_TRY_HIP(ctx, launch_fill_f64(t, V, 0.0, st));
      OMNI_TRY_HIP(ctx, launch_adj_gather(mask_rule, pl.rays.as<RayRec>(), pl.nrays,
                                          pl.dirs.as<DirConst>(), pl.g, mask, u, t, st));
      OMNI_TRY_HIP(ctx, launch_adj_horner(V, y, t, cnt, z, st));
    }
  }
  // dL/dg = B^T (z / C)
  OMNI_TRY_HIP(ctx, launch_adj_scale(V, z, cnt, u, st));
  if (tiles && mask_rule) {
    OMNI_TRY_HIP(ctx, mask_tiles(u, 2, dg));
  } else if (tiles) {
    // per-segment sums, per-ray suffix sums, then the tiles' scatter (every
    // voxel of dg stored)
    OMNI_TRY_HIP(ctx, launch_adj_tiles(0, pl.ntiles, pl.tile_ptr1.as<unsigned long long>(),
                                       pl.segs_slot.as<SegRec>(), nullptr, nullptr, nullptr, 0,
                                       nullptr, nullptr, pl.g, pl.tg, wp, u,
                                       ctx->adj_seg.as<double>(), nullptr, nullptr, st));
    OMNI_TRY_HIP(ctx, launch_adj_tiles(1, pl.ntiles, nullptr, nullptr, nullptr, nullptr,
                                       pl.rays.as<RayRec>(), pl.nrays, pl.nseg.as<unsigned int>(),
                                       pl.val_off.as<unsigned long long>(), pl.g, pl.tg, wp,
                                       nullptr, ctx->adj_seg.as<double>(), nullptr, nullptr, st));
    OMNI_TRY_HIP(ctx, launch_adj_tiles(2, pl.ntiles, nullptr, nullptr,
                                       pl.tile_ptr.as<unsigned long long>(), pl.segs.as<SegRec>(),
                                       nullptr, 0, nullptr, nullptr, pl.g, pl.tg, wp, u,
                                       ctx->adj_seg.as<double>(), nullptr, dg, st));
  } else {
    OMNI_TRY_HIP(ctx, launch_fill_f64(dg, 3 * V, 0.0, st));
    OMNI_TRY_HIP(ctx, launch_adj_scatter(mask_rule, pl.rays.as<RayRec>(), pl.nrays,
                                         pl.dirs.as<DirConst>(), pl.g, wp, mask, u, dg, V, st));
  }
  if (dt == OMNI3D_F64)
    OMNI_TRY_HIP(ctx, (launch_convert<double, double>(dg, static_cast<double*>(dldgrad_d), 3 * V, st)));
  else
    OMNI_TRY_HIP(ctx, (launch_convert<double, float>(dg, static_cast<float*>(dldgrad_d), 3 * V, st)));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  return OMNI3D_OK;
}

}  // extern "C"
