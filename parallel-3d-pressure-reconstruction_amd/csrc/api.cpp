// C-ABI of libomni3d: context, plan build and the reconstruction driver
// (kernel.cu:8805-8861 restated as: plan -> pass A -> source iteration ->
// pass B -> averaging, all queued on the context's stream).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "driver.hpp"
#include "host.hpp"
#include "launch.hpp"
#include "stage.hpp"

using namespace omni3d;

namespace omni3d {

thread_local std::string g_last_error;

int fail(omni3d_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  g_last_error = msg;
  return code;
}

int64_t voxels(const omni3d_params& p) { return (int64_t)p.nx * p.ny * p.nz; }

bool params_ok(const omni3d_params& p, std::string* why) {
  if (p.nx < 2 || p.ny < 2 || p.nz < 2 || p.nx > kMaxDim || p.ny > kMaxDim || p.nz > kMaxDim) {
    *why = "grid dimensions must be in [2, 1023]";
    return false;
  }
  if (!(p.dx > 0 && p.dy > 0 && p.dz > 0)) {
    *why = "grid spacing must be positive";
    return false;
  }
  if (!(p.line_spacing > 0)) {
    *why = "line spacing must be positive";
    return false;
  }
  if (p.method < OMNI3D_METHOD_PLAIN || p.method > OMNI3D_METHOD_WEIGHTED_MINICURL) {
    *why = "unknown method";
    return false;
  }
  if (p.precision != OMNI3D_FP64 && p.precision != OMNI3D_FP32_KAHAN) {
    *why = "unknown precision";
    return false;
  }
  if (p.solver != OMNI3D_SOLVER_AFFINE && p.solver != OMNI3D_SOLVER_NAIVE) {
    *why = "unknown solver";
    return false;
  }
  if (p.max_iterations < 1) {
    *why = "max_iterations must be >= 1";
    return false;
  }
  if (p.traversal != OMNI3D_TRAVERSAL_STAIRCASE && p.traversal != OMNI3D_TRAVERSAL_TRILINEAR) {
    *why = "unknown traversal";
    return false;
  }
  if (p.traversal == OMNI3D_TRAVERSAL_TRILINEAR &&
      (p.method != OMNI3D_METHOD_PLAIN || p.solver != OMNI3D_SOLVER_AFFINE)) {
    *why = "the trilinear traversal runs the plain method with the affine solver";
    return false;
  }
  return true;
}

size_t elem_size(int dtype) { return dtype == OMNI3D_F64 ? 8 : 4; }

// Source numbering for the iteration's gathers: the voxels flagged as ray
// sources, grouped by box face (the first face a voxel lies on, in the order
// z=0, z=max, y=0, y=max, x=0, x=max), then by 8x8 block of the face, then
// row-major inside the block; sources off the surface (none for the plain
// method) follow in voxel order.
void number_sources(int X, int Y, int Z, const std::vector<uint8_t>& flag, std::vector<int>* sid,
                    std::vector<int>* slist) {
  const int64_t V = (int64_t)X * Y * Z;
  std::vector<std::pair<unsigned long long, int>> keys;
  for (int64_t v = 0; v < V; v++) {
    if (!flag[v]) continue;
    const int i = (int)(v % X), j = (int)((v / X) % Y), k = (int)(v / ((int64_t)X * Y));
    int face, u, w, nu;
    if (k == 0) { face = 0; u = i; w = j; nu = X; }
    else if (k == Z - 1) { face = 1; u = i; w = j; nu = X; }
    else if (j == 0) { face = 2; u = i; w = k; nu = X; }
    else if (j == Y - 1) { face = 3; u = i; w = k; nu = X; }
    else if (i == 0) { face = 4; u = j; w = k; nu = Y; }
    else if (i == X - 1) { face = 5; u = j; w = k; nu = Y; }
    else { face = 6; u = 0; w = 0; nu = 1; }
    unsigned long long key;
    if (face < 6) {
      const unsigned long long blk = (unsigned long long)(w >> 3) * (unsigned long long)((nu + 7) >> 3) +
                                     (unsigned long long)(u >> 3);
      key = ((unsigned long long)face << 56) | (blk << 6) | (unsigned long long)(((w & 7) << 3) | (u & 7));
    } else {
      key = (6ull << 56) | (unsigned long long)v;
    }
    keys.emplace_back(key, (int)v);
  }
  std::sort(keys.begin(), keys.end());
  sid->assign(V, -1);
  slist->resize(keys.size());
  for (size_t s = 0; s < keys.size(); s++) {
    (*slist)[s] = keys[s].second;
    (*sid)[keys[s].second] = (int)s;
  }
}

#ifdef OMNI3D_LAYOUT_STUDY
// LDS-staged layout of the iteration (stage.hip).  Left off (pl.staged
// false, the slot-ordered gathers are used) when a (group, direction) needs
// more lines than a stage holds or the sizes exceed the 32-bit indices.
int build_staged(omni3d_ctx* ctx, int nS, int D, int64_t nnz) {
  Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  pl.staged = false;
  const int nch = (nS + 63) / 64;
  const int ngroups = (nch + kStageGroup - 1) / kStageGroup;
  const int64_t ngd = (int64_t)ngroups * D;
  if (nnz <= 0 || nnz >= (1ll << 31) || pl.ncell >= (1ll << 32) || D > 65535 ||
      stage_key_bits(nS, D) > 64)
    return OMNI3D_OK;
  DevBuf ka, kb, va, vb, flag, incl, tmp, gdcnt, gdptr, st_dend, sbase, ovf;
  OMNI_TRY_HIP(ctx, ka.alloc(sizeof(unsigned long long) * nnz));
  OMNI_TRY_HIP(ctx, kb.alloc(sizeof(unsigned long long) * nnz));
  OMNI_TRY_HIP(ctx, va.alloc(sizeof(unsigned int) * nnz));
  OMNI_TRY_HIP(ctx, vb.alloc(sizeof(unsigned int) * nnz));
  OMNI_TRY_HIP(ctx, flag.alloc(sizeof(unsigned int) * nnz));
  OMNI_TRY_HIP(ctx, incl.alloc(sizeof(unsigned int) * nnz));
  OMNI_TRY_HIP(ctx, tmp.alloc(stage_tmp_bytes(nnz, ngd)));
  StageBuild sb;
  sb.nS = nS;
  sb.D = D;
  sb.nch = nch;
  sb.nnz = nnz;
  sb.moff = pl.moff.as<unsigned long long>();
  sb.masks = pl.masks.as<unsigned long long>();
  sb.sdir = pl.sdir.as<unsigned short>();
  sb.cell_eb = pl.cell_eb.as<unsigned int>();
  sb.ent = pl.ent.as<int>();
  sb.keys_a = ka.as<unsigned long long>();
  sb.keys_b = kb.as<unsigned long long>();
  sb.vals_a = va.as<unsigned int>();
  sb.vals_b = vb.as<unsigned int>();
  sb.flag = flag.as<unsigned int>();
  sb.incl = incl.as<unsigned int>();
  sb.tmp = tmp.as<void>();
  sb.tmp_bytes = tmp.size();
  OMNI_TRY_HIP(ctx, stage_keys_sorted(sb, st));
  unsigned int nl = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nl, sb.incl + (nnz - 1), 4, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  OMNI_TRY_HIP(ctx, pl.ulines.alloc(sizeof(unsigned int) * ((size_t)nl + 1)));
  OMNI_TRY_HIP(ctx, gdcnt.alloc(sizeof(unsigned int) * (ngd + 1)));
  OMNI_TRY_HIP(ctx, gdptr.alloc(sizeof(unsigned int) * (ngd + 1)));
  OMNI_TRY_HIP(ctx, pl.st_n.alloc(sizeof(unsigned int) * ngroups));
  OMNI_TRY_HIP(ctx, pl.st_lo.alloc(sizeof(unsigned int) * ngd));
  OMNI_TRY_HIP(ctx, pl.st_hi.alloc(sizeof(unsigned int) * ngd));
  OMNI_TRY_HIP(ctx, st_dend.alloc(sizeof(unsigned short) * ngd));
  OMNI_TRY_HIP(ctx, sbase.alloc(sizeof(unsigned int) * ngd));
  OMNI_TRY_HIP(ctx, pl.st_cse.alloc(sizeof(unsigned int) * ngd * kStageGroup));
  // slack after the positions: the iteration stages them with 16-byte loads
  OMNI_TRY_HIP(ctx, pl.soff.alloc(sizeof(unsigned short) * (nnz + 16)));
  OMNI_TRY_HIP(ctx, ovf.alloc(sizeof(unsigned int)));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(ovf.as<void>(), 0, sizeof(unsigned int), st));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(pl.soff.as<void>(), 0, sizeof(unsigned short) * (nnz + 16), st));
  OMNI_TRY_HIP(ctx, stage_finish(sb, ngroups, pl.ulines.as<unsigned int>(), gdcnt.as<unsigned int>(),
                                 gdptr.as<unsigned int>(), pl.st_n.as<unsigned int>(),
                                 pl.st_lo.as<unsigned int>(), pl.st_hi.as<unsigned int>(),
                                 st_dend.as<unsigned short>(), sbase.as<unsigned int>(),
                                 pl.st_cse.as<unsigned int>(), pl.soff.as<unsigned short>(),
                                 ovf.as<unsigned int>(), st));
  unsigned int overflow = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&overflow, ovf.as<void>(), 4, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  pl.nlines = nl;
  pl.staged = overflow == 0;
  if (std::getenv("OMNI3D_VERBOSE"))
    std::fprintf(stderr, "omni3d plan: %d sources, %lld entries, %lld slots, %u staged lines%s\n",
                 nS, (long long)nnz, (long long)pl.ncell, nl,
                 pl.staged ? "" : " (stage overflow: slot-ordered gathers)");
  return OMNI3D_OK;
}

#endif  // OMNI3D_LAYOUT_STUDY

// Tile engine part of the plan: compiled step codes and per-tile segments.
// The source CSR built by the tile walks: the count walk counts the visits
// landing on source voxels (and the diagonal), the fill walk writes their
// keys (pl.rowptr / pl.diag are filled; key holds the unsorted rows).
struct SrcRowBuild {
  int nS = 0;
  DevBuf rowcnt;  // counts, then the fill walk's cursors
  DevBuf key;     // uint64[nnz + 1]
  unsigned long long nnz = 0;
};

int build_tiles(omni3d_ctx* ctx, const GridDesc& g, int64_t nrays, SrcRowBuild* rb) {
  Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  TileGrid tg;
  tg.bits = ctx->tile_bits;
  tg.lshift = 3;  // segment-length groups of 8 visits
  if (const char* e = std::getenv("OMNI3D_LEN_SHIFT")) tg.lshift = std::max(0, std::min(6, std::atoi(e)));
  const int kTile = 1 << tg.bits;
  tg.ntx = (g.X + kTile - 1) / kTile;
  tg.nty = (g.Y + kTile - 1) / kTile;
  tg.ntz = (g.Z + kTile - 1) / kTile;
  const int nt = tg.ntx * tg.nty * tg.ntz;
  const int nslots = nt * kLenBuckets;
  const int64_t ngroups = (nrays + 63) / 64;
  DevBuf tcnt, gcnt, tmp;
  OMNI_TRY_HIP(ctx, pl.nseg.alloc(sizeof(unsigned int) * nrays));
  OMNI_TRY_HIP(ctx, tcnt.alloc(sizeof(unsigned int) * nslots));
  OMNI_TRY_HIP(ctx, gcnt.alloc(sizeof(unsigned int) * ngroups));
  OMNI_TRY_HIP(ctx, tmp.alloc(scan_u32_tmp_bytes(std::max<int64_t>(ngroups, nslots))));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(tcnt.as<void>(), 0, sizeof(unsigned int) * nslots, st));
  const int nS = rb->nS;
  OMNI_TRY_HIP(ctx, rb->rowcnt.alloc(sizeof(unsigned int) * (nS + 1)));
  OMNI_TRY_HIP(ctx, pl.diag.alloc(sizeof(unsigned int) * (nS + 1)));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(rb->rowcnt.as<void>(), 0, sizeof(unsigned int) * (nS + 1), st));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(pl.diag.as<void>(), 0, sizeof(unsigned int) * (nS + 1), st));
  SrcRows cnt_rows;
  cnt_rows.sid = pl.sid.as<int>();
  cnt_rows.rowcnt = rb->rowcnt.as<unsigned int>();
  cnt_rows.diag = pl.diag.as<unsigned int>();
  OMNI_TRY_HIP(ctx, launch_tile_count(pl.rays.as<RayRec>(), nrays, pl.dirs.as<DirConst>(), g, tg,
                                      pl.nseg.as<unsigned int>(), tcnt.as<unsigned int>(),
                                      cnt_rows, st));
  OMNI_TRY_HIP(ctx, pl.rowptr.alloc(sizeof(unsigned long long) * (nS + 1)));
  OMNI_TRY_HIP(ctx, launch_scan_counts(rb->rowcnt.as<unsigned int>(), nS,
                                       pl.rowptr.as<unsigned long long>(), st));
  // value slots: groups of 64 rays, segment-ordinal major inside a group
  OMNI_TRY_HIP(ctx, launch_seg_groups(nrays, pl.nseg.as<unsigned int>(), gcnt.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, pl.val_off.alloc(sizeof(unsigned long long) * (ngroups + 1)));
  OMNI_TRY_HIP(ctx, launch_scan_u32(gcnt.as<unsigned int>(), ngroups,
                                    pl.val_off.as<unsigned long long>(), 0, tmp.as<void>(),
                                    tmp.size(), st));
  OMNI_TRY_HIP(ctx, pl.tile_ptr.alloc(sizeof(unsigned long long) * (nslots + 1)));
  OMNI_TRY_HIP(ctx, launch_scan_u32(tcnt.as<unsigned int>(), nslots,
                                    pl.tile_ptr.as<unsigned long long>(), 0, tmp.as<void>(),
                                    tmp.size(), st));
  unsigned long long nsegs = 0, nval = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nsegs, pl.tile_ptr.as<unsigned long long>() + nslots, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nval, pl.val_off.as<unsigned long long>() + ngroups, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&rb->nnz, pl.rowptr.as<unsigned long long>() + nS, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (nsegs >= (1ull << 32) || nval >= (1ull << 32))
    return fail(ctx, OMNI3D_ERR_ARG, "too many ray segments");
  OMNI_TRY_HIP(ctx, pl.segs.alloc(sizeof(SegRec) * std::max<unsigned long long>(nsegs, 1)));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(tcnt.as<void>(), 0, sizeof(unsigned int) * nslots, st));
  OMNI_TRY_HIP(ctx, rb->key.alloc(sizeof(unsigned long long) * (rb->nnz + 1)));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(rb->rowcnt.as<void>(), 0, sizeof(unsigned int) * (nS + 1), st));
  SrcRows fill_rows;
  fill_rows.sid = pl.sid.as<int>();
  fill_rows.rowptr = pl.rowptr.as<unsigned long long>();
  fill_rows.cursor = rb->rowcnt.as<unsigned int>();
  fill_rows.key = rb->key.as<unsigned long long>();
  OMNI_TRY_HIP(ctx, launch_tile_fill(pl.rays.as<RayRec>(), nrays, pl.dirs.as<DirConst>(), g, tg,
                                     pl.sid.as<int>(), pl.val_off.as<unsigned long long>(),
                                     pl.tile_ptr.as<unsigned long long>(), tcnt.as<unsigned int>(),
                                     pl.segs.as<SegRec>(), fill_rows, st));
  {
    // every (tile, length bucket) list in value-slot order
    int key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) < nval) key_bits++;
    DevBuf sorted, stmp;
    OMNI_TRY_HIP(ctx, sorted.alloc(sizeof(SegRec) * std::max<unsigned long long>(nsegs, 1)));
    const size_t tb = sort_segments_tmp_bytes((int64_t)nsegs, nslots);
    OMNI_TRY_HIP(ctx, stmp.alloc(tb));
    OMNI_TRY_HIP(ctx, launch_sort_segments(pl.segs.as<SegRec>(), sorted.as<SegRec>(), (int64_t)nsegs,
                                           nslots, pl.tile_ptr.as<unsigned long long>(), key_bits,
                                           stmp.as<void>(), tb, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    pl.segs.swap(sorted);
  }
  {
    // each tile's segments in value-slot order (the segment-sum pass stores
    // one slot per segment: coalesced stores)
    std::vector<unsigned long long> hp((size_t)nslots + 1), h1((size_t)nt + 1);
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(hp.data(), pl.tile_ptr.as<void>(),
                                     sizeof(unsigned long long) * (nslots + 1),
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    for (int t = 0; t <= nt; t++) h1[t] = hp[(size_t)t * kLenBuckets];
    OMNI_TRY_HIP(ctx, pl.tile_ptr1.alloc(sizeof(unsigned long long) * (nt + 1)));
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(pl.tile_ptr1.as<void>(), h1.data(),
                                     sizeof(unsigned long long) * (nt + 1), hipMemcpyHostToDevice,
                                     st));
    int key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) < nval) key_bits++;
    DevBuf stmp;
    OMNI_TRY_HIP(ctx, pl.segs_slot.alloc(sizeof(SegRec) * std::max<unsigned long long>(nsegs, 1)));
    const size_t tb = sort_segments_tmp_bytes((int64_t)nsegs, nt);
    OMNI_TRY_HIP(ctx, stmp.alloc(tb));
    OMNI_TRY_HIP(ctx, launch_sort_segments(pl.segs.as<SegRec>(), pl.segs_slot.as<SegRec>(),
                                           (int64_t)nsegs, nt, pl.tile_ptr1.as<unsigned long long>(),
                                           key_bits, stmp.as<void>(), tb, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  }
  // tiles that touch the box surface hold every source voxel of the plain method
  std::vector<int> surf;
  for (int t = 0; t < nt; t++) {
    const int tx = t % tg.ntx, ty = (t / tg.ntx) % tg.nty, tz = t / (tg.ntx * tg.nty);
    if (tx == 0 || ty == 0 || tz == 0 || tx == tg.ntx - 1 || ty == tg.nty - 1 || tz == tg.ntz - 1)
      surf.push_back(t);
  }
  OMNI_TRY_HIP(ctx, pl.surf.alloc(sizeof(int) * std::max<size_t>(surf.size(), 1)));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(pl.surf.as<void>(), surf.data(), sizeof(int) * surf.size(),
                                   hipMemcpyHostToDevice, st));
  {
    // compact lists of the source-sum pass
    DevBuf scnt, stmp2;
    OMNI_TRY_HIP(ctx, scnt.alloc(sizeof(unsigned int) * nt));
    OMNI_TRY_HIP(ctx, hipMemsetAsync(scnt.as<void>(), 0, sizeof(unsigned int) * nt, st));
    OMNI_TRY_HIP(ctx, pl.sseg_ptr.alloc(sizeof(unsigned long long) * (nt + 1)));
    OMNI_TRY_HIP(ctx, stmp2.alloc(scan_u32_tmp_bytes(nt)));
    OMNI_TRY_HIP(ctx, launch_src_segs(0, (int)surf.size(), pl.surf.as<int>(),
                                      pl.tile_ptr.as<unsigned long long>(), pl.segs.as<SegRec>(), g,
                                      tg, pl.sid.as<int>(), scnt.as<unsigned int>(), nullptr,
                                      nullptr, st));
    OMNI_TRY_HIP(ctx, launch_scan_u32(scnt.as<unsigned int>(), nt,
                                      pl.sseg_ptr.as<unsigned long long>(), 0, stmp2.as<void>(),
                                      stmp2.size(), st));
    unsigned long long nss = 0;
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nss, pl.sseg_ptr.as<unsigned long long>() + nt, 8,
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    OMNI_TRY_HIP(ctx, pl.ssegs.alloc(sizeof(SegRec) * std::max<unsigned long long>(nss, 1)));
    OMNI_TRY_HIP(ctx, launch_src_segs(1, (int)surf.size(), pl.surf.as<int>(),
                                      pl.tile_ptr.as<unsigned long long>(), pl.segs.as<SegRec>(), g,
                                      tg, pl.sid.as<int>(), scnt.as<unsigned int>(),
                                      pl.sseg_ptr.as<unsigned long long>(), pl.ssegs.as<SegRec>(),
                                      st));
    pl.nssegs = (int64_t)nss;
    if (std::getenv("OMNI3D_VERBOSE"))
      std::fprintf(stderr, "omni3d plan: %llu segments, %llu reach a source voxel\n", nsegs, nss);
  }
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  pl.tg = tg;
  pl.ntiles = nt;
  pl.nsurf = (int)surf.size();
  pl.nsegs = (int64_t)nsegs;
  pl.nval = (int64_t)nval;
  return OMNI3D_OK;
}

// Builds the geometry plan: ray list, visit counts, source voxels and the CSR
// of visits that land on source voxels.
int build_plan(omni3d_ctx* ctx) {
  const omni3d_params& P = ctx->prm;
  const int D = (int)(ctx->k.size() / 3);
  {
    const Plan& cur = *ctx->plan;
    if (cur.valid && !ctx->dirs_dirty && cur.X == P.nx && cur.Y == P.ny && cur.Z == P.nz &&
        cur.ls == P.line_spacing && cur.D == D && cur.tg.bits == ctx->tile_bits)
      return OMNI3D_OK;
  }
  // a plan shared with other contexts stays theirs: build a new one
  if (ctx->plan.use_count() > 1) ctx->plan = std::make_shared<Plan>();
  Plan& pl = *ctx->plan;
  pl.valid = false;
  hipStream_t st = ctx->stream;
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const int64_t V = (int64_t)X * Y * Z;
  std::vector<DirConst> dc(D);
  for (int i = 0; i < D; i++) dc[i] = make_dir_const(ctx->k[3 * i], ctx->k[3 * i + 1], ctx->k[3 * i + 2]);
  OMNI_TRY_HIP(ctx, pl.dirs.alloc(sizeof(DirConst) * D));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(pl.dirs.as<void>(), dc.data(), sizeof(DirConst) * D,
                                   hipMemcpyHostToDevice, st));
  GridDesc g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.ls = P.line_spacing;
  g.ng = lattice_size(X, Y, Z, P.line_spacing);
  g.cc[0] = (X - 1) / 2.0;
  g.cc[1] = (Y - 1) / 2.0;
  g.cc[2] = (Z - 1) / 2.0;
  // candidates per direction: the lattice padded to whole 16x16 patches
  const int64_t side = (g.ng + 15) / 16;
  const int64_t npts = side * side * 256;
  const int64_t ncand = npts * D;
  const int64_t nblk = (ncand + kEnumItemsPerBlock - 1) / kEnumItemsPerBlock;
  if (nblk > (int64_t)1 << 31) return fail(ctx, OMNI3D_ERR_ARG, "too many candidate rays");
  DevBuf bcnt, boff;
  OMNI_TRY_HIP(ctx, bcnt.alloc(sizeof(unsigned int) * nblk));
  OMNI_TRY_HIP(ctx, boff.alloc(sizeof(unsigned long long) * (nblk + 1)));
  OMNI_TRY_HIP(ctx, launch_enum_count(pl.dirs.as<DirConst>(), g, npts, ncand,
                                      bcnt.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, launch_scan_counts(bcnt.as<unsigned int>(), (int)nblk,
                                       boff.as<unsigned long long>(), st));
  unsigned long long nrays = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nrays, boff.as<unsigned long long>() + nblk, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (nrays == 0) return fail(ctx, OMNI3D_ERR_ARG, "no ray crosses the grid");
  OMNI_TRY_HIP(ctx, pl.rays.alloc(sizeof(RayRec) * nrays));
  OMNI_TRY_HIP(ctx, launch_enum_write(pl.dirs.as<DirConst>(), g, npts, ncand,
                                      boff.as<unsigned long long>(), pl.rays.as<RayRec>(), st));
  // visit counts and source flags
  DevBuf flag, steps, mx;
  OMNI_TRY_HIP(ctx, pl.cnt.alloc(sizeof(unsigned int) * V));
  OMNI_TRY_HIP(ctx, flag.alloc(V));
  OMNI_TRY_HIP(ctx, steps.alloc(16));
  OMNI_TRY_HIP(ctx, mx.alloc(16));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(pl.cnt.as<void>(), 0, sizeof(unsigned int) * V, st));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(flag.as<void>(), 0, V, st));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(steps.as<void>(), 0, 16, st));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(mx.as<void>(), 0, 16, st));
  // entry voxels and step totals from the ray records; the visit counts
  // follow from the tile segments below
  OMNI_TRY_HIP(ctx, launch_plan_flags(pl.rays.as<RayRec>(), (int64_t)nrays, g, flag.as<uint8_t>(),
                                      steps.as<unsigned long long>(), st));
  OMNI_TRY_HIP(ctx, launch_max_len(pl.rays.as<RayRec>(), (int64_t)nrays, mx.as<unsigned int>() + 1, st));
  unsigned int mxh[2] = {0, 0};
  unsigned long long nsteps = 0;
  // source numbering: 8x8 blocks of each box face, so that consecutive
  // source ids are neighbouring voxels in both face directions -- on the
  // device (face slots flagged, scanned, scattered); on the host, the same
  // order, only if a flagged voxel lies off the surface
  const FaceSlots fs = face_slots(X, Y, Z);
  const unsigned long long nslot = fs.off[6];
  DevBuf svalid, soff, stmp, sint;
  OMNI_TRY_HIP(ctx, svalid.alloc(sizeof(unsigned int) * nslot));
  OMNI_TRY_HIP(ctx, soff.alloc(sizeof(unsigned long long) * (nslot + 1)));
  OMNI_TRY_HIP(ctx, stmp.alloc(scan_u32_tmp_bytes((int64_t)nslot)));
  OMNI_TRY_HIP(ctx, sint.alloc(sizeof(unsigned int)));
  OMNI_TRY_HIP(ctx, hipMemsetAsync(sint.as<void>(), 0, sizeof(unsigned int), st));
  OMNI_TRY_HIP(ctx, launch_face_slot_flags(fs, g, flag.as<uint8_t>(), svalid.as<unsigned int>(),
                                           sint.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, launch_scan_u32(svalid.as<unsigned int>(), (int64_t)nslot,
                                    soff.as<unsigned long long>(), 0, stmp.as<void>(), stmp.size(),
                                    st));
  unsigned long long nsurf_src = 0;
  unsigned int ninterior = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nsurf_src, soff.as<unsigned long long>() + nslot, 8,
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&ninterior, sint.as<void>(), 4, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nsteps, steps.as<void>(), 8, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  int nS = 0;
  OMNI_TRY_HIP(ctx, pl.sid.alloc(sizeof(int) * V));
  if (ninterior == 0) {
    nS = (int)nsurf_src;
    OMNI_TRY_HIP(ctx, pl.slist.alloc(sizeof(int) * std::max(nS, 1)));
    OMNI_TRY_HIP(ctx, launch_fill_u32(pl.sid.as<void>(), V, 0xffffffffu, st));
    OMNI_TRY_HIP(ctx, launch_face_slot_ids(fs, g, svalid.as<unsigned int>(),
                                           soff.as<unsigned long long>(), pl.sid.as<int>(),
                                           pl.slist.as<int>(), st));
  } else {
    std::vector<uint8_t> hflag(V);
    OMNI_TRY_HIP(ctx, hipMemcpy(hflag.data(), flag.as<void>(), V, hipMemcpyDeviceToHost));
    std::vector<int> hsid, hslist;
    number_sources(X, Y, Z, hflag, &hsid, &hslist);
    nS = (int)hslist.size();
    OMNI_TRY_HIP(ctx, pl.slist.alloc(sizeof(int) * std::max(nS, 1)));
    OMNI_TRY_HIP(ctx, hipMemcpy(pl.sid.as<void>(), hsid.data(), sizeof(int) * V,
                                hipMemcpyHostToDevice));
    OMNI_TRY_HIP(ctx, hipMemcpy(pl.slist.as<void>(), hslist.data(), sizeof(int) * nS,
                                hipMemcpyHostToDevice));
  }
  // the tile engine's segments (and, by the same walks, the CSR of the
  // visits landing on source voxels), then every voxel's visit count from
  // the segments
  SrcRowBuild rows;
  rows.nS = nS;
  {
    const int rc = build_tiles(ctx, g, (int64_t)nrays, &rows);
    if (rc) return rc;
  }
  OMNI_TRY_HIP(ctx, launch_tile_counts(pl.ntiles, pl.tile_ptr1.as<unsigned long long>(),
                                       pl.segs_slot.as<SegRec>(), g, pl.tg,
                                       pl.cnt.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, launch_max_u32(pl.cnt.as<unsigned int>(), V, mx.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(mxh, mx.as<unsigned int>(), 8, hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  // CSR of visits landing on source voxels (counted and filled by the
  // tile walks): each row's keys sorted
  const unsigned long long nnz = rows.nnz;
  DevBuf& key = rows.key;
  OMNI_TRY_HIP(ctx, launch_sort_rows(nS, pl.rowptr.as<unsigned long long>(),
                                     key.as<unsigned long long>(), st));
  // the dense layout (default iteration) unless sizes exceed its word format:
  // 24-bit source ids, rows shorter than 4096 entries (int64 row sums of
  // terms below 2^51); otherwise (e.g. thousands of directions) the masked
  // sliced-ELL layout
  const char* iter_env = std::getenv("OMNI3D_ITER");
  // groups of 8 (4) chunks use merged words with 21-bit (22-bit) sources
  int grp = ctx->dense_group;
  if (grp == 8 && nS >= (1 << 21)) grp = 4;
  if (grp == 4 && nS >= (1 << 22)) grp = 1;
#ifdef OMNI3D_LAYOUT_STUDY
  const bool want_slots = iter_env && (std::strcmp(iter_env, "slots") == 0 ||
                                       std::strcmp(iter_env, "staged") == 0);
#else
  const bool want_slots = iter_env && std::strcmp(iter_env, "slots") == 0;
#endif
  const bool use_dense = !want_slots && nS < (1 << (32 - dense_row_bits(grp))) &&
                         mxh[0] < 4096u && nnz > 0 && nnz < (1ull << 31);
  pl.ncell = 0;
  if (!use_dense) {
    // masked sliced-ELL copy of the sorted rows for the iteration
    if (D > 4096) return fail(ctx, OMNI3D_ERR_ARG, "at most 4096 directions");
    const int nch = (nS + 63) / 64;
    DevBuf nslot, tmp;
    OMNI_TRY_HIP(ctx, nslot.alloc(sizeof(unsigned int) * (nch + 1)));
    OMNI_TRY_HIP(ctx, tmp.alloc(scan_u32_tmp_bytes(nch)));
    OMNI_TRY_HIP(ctx, pl.moff.alloc(sizeof(unsigned long long) * (nch + 1)));
    OMNI_TRY_HIP(ctx, launch_sell_layout(nS, D, pl.rowptr.as<unsigned long long>(),
                                         key.as<unsigned long long>(), nslot.as<unsigned int>(), st));
    OMNI_TRY_HIP(ctx, launch_scan_u32(nslot.as<unsigned int>(), nch, pl.moff.as<unsigned long long>(),
                                      0, tmp.as<void>(), tmp.size(), st));
    unsigned long long nm = 0;
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nm, pl.moff.as<unsigned long long>() + nch, 8,
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (nnz >= (1ull << 32)) return fail(ctx, OMNI3D_ERR_ARG, "too many source visits");
    OMNI_TRY_HIP(ctx, pl.masks.alloc(sizeof(unsigned long long) * (nm + 1)));
    OMNI_TRY_HIP(ctx, pl.cell_eb.alloc(sizeof(unsigned int) * (nm + 1)));
    OMNI_TRY_HIP(ctx, pl.sdir.alloc(sizeof(unsigned short) * (nm + 1)));
    // slack after the entries: the iteration stages them with 16-byte loads
    OMNI_TRY_HIP(ctx, pl.ent.alloc(sizeof(int) * (nnz + 8)));
    OMNI_TRY_HIP(ctx, launch_sell_fill(nS, D, pl.rowptr.as<unsigned long long>(),
                                       key.as<unsigned long long>(), pl.moff.as<unsigned long long>(),
                                       pl.masks.as<unsigned long long>(),
                                       pl.cell_eb.as<unsigned int>(),
                                       pl.sdir.as<unsigned short>(), pl.ent.as<int>(), st));
    pl.ncell = (int64_t)nm;
  }
  {
    pl.dense = false;
    if (use_dense) {
      DevBuf ka, kb, tmp;
      OMNI_TRY_HIP(ctx, ka.alloc(sizeof(unsigned long long) * nnz));
      OMNI_TRY_HIP(ctx, kb.alloc(sizeof(unsigned long long) * nnz));
      OMNI_TRY_HIP(ctx, tmp.alloc(dense_tmp_bytes((int64_t)nnz)));
      // slack after the words: the iteration stages them with 16-byte loads
      OMNI_TRY_HIP(ctx, pl.dent.alloc(sizeof(unsigned int) * (nnz + 8)));
      OMNI_TRY_HIP(ctx, hipMemsetAsync(pl.dent.as<void>(), 0, sizeof(unsigned int) * (nnz + 8), st));
      // groups of at most 64 * grp consecutive rows holding about equal entries:
      // one workgroup per group, as many groups as fill the CUs evenly, so
      // the launch does not wait on the largest uniform group
      DevBuf rgroup;
      if (grp > 1) {
        std::vector<unsigned long long> hrp((size_t)nS + 1);
        OMNI_TRY_HIP(ctx, hipMemcpy(hrp.data(), pl.rowptr.as<void>(),
                                    sizeof(unsigned long long) * ((size_t)nS + 1),
                                    hipMemcpyDeviceToHost));
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) == hipSuccess)
          (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        ncu = ctx->groups_fill_cus ? std::max(ncu, 1) : 1;
        std::vector<int> hgrow, hrg((size_t)nS);
        balance_groups(hrp, 64 * grp, ncu, &hgrow, &hrg);
        pl.dense_ngroups = (int)hgrow.size() - 1;
        OMNI_TRY_HIP(ctx, pl.dense_grow.alloc(sizeof(int) * hgrow.size()));
        OMNI_TRY_HIP(ctx, rgroup.alloc(sizeof(int) * (size_t)std::max(nS, 1)));
        OMNI_TRY_HIP(ctx, hipMemcpy(pl.dense_grow.as<void>(), hgrow.data(),
                                    sizeof(int) * hgrow.size(), hipMemcpyHostToDevice));
        OMNI_TRY_HIP(ctx, hipMemcpy(rgroup.as<void>(), hrg.data(), sizeof(int) * (size_t)nS,
                                    hipMemcpyHostToDevice));
      }
      OMNI_TRY_HIP(ctx, build_dense_entries(nS, (int64_t)nnz, grp, pl.rowptr.as<unsigned long long>(),
                                            key.as<unsigned long long>(),
                                            ka.as<unsigned long long>(), kb.as<unsigned long long>(),
                                            tmp.as<void>(), tmp.size(), pl.dent.as<unsigned int>(),
                                            st, grp > 1 ? rgroup.as<int>() : nullptr,
                                            grp > 1 ? pl.dense_grow.as<int>() : nullptr,
                                            pl.dense_ngroups));
      pl.nwords = (int64_t)nnz;
      if (grp > 1) {
        DevBuf flag, excl;
        OMNI_TRY_HIP(ctx, flag.alloc(sizeof(unsigned int) * nnz));
        OMNI_TRY_HIP(ctx, excl.alloc(sizeof(unsigned int) * nnz));
        const int ng = pl.dense_ngroups;
        OMNI_TRY_HIP(ctx, pl.dense_wptr.alloc(sizeof(unsigned int) * (ng + 1)));
        unsigned int nw = 0;
        OMNI_TRY_HIP(ctx, merge_dense_entries(nS, (int64_t)nnz, grp,
                                              pl.rowptr.as<unsigned long long>(),
                                              kb.as<unsigned long long>(), flag.as<unsigned int>(),
                                              excl.as<unsigned int>(), tmp.as<void>(), tmp.size(),
                                              pl.dent.as<unsigned int>(),
                                              pl.dense_wptr.as<unsigned int>(), &nw, st,
                                              pl.dense_grow.as<int>(), ng));
        pl.nwords = nw;
      }
      OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
      pl.dense = true;
      pl.dense_group = grp;
      if (std::getenv("OMNI3D_VERBOSE")) {
        std::fprintf(stderr, "omni3d plan: %lld entries in %lld dense words (groups of %d chunks)\n",
                     (long long)nnz, (long long)pl.nwords, grp);
        if (grp > 1) {
          const int ng = pl.dense_ngroups;
          std::vector<unsigned int> wp(ng + 1);
          if (hipMemcpy(wp.data(), pl.dense_wptr.as<void>(), sizeof(unsigned int) * (ng + 1),
                        hipMemcpyDeviceToHost) == hipSuccess) {
            std::vector<unsigned int> n(ng);
            for (int i = 0; i < ng; i++) n[i] = wp[i + 1] - wp[i];
            std::sort(n.begin(), n.end());
            std::fprintf(stderr, "omni3d plan: words per group min %u median %u p90 %u max %u\n",
                         n[0], n[ng / 2], n[(ng * 9) / 10], n[ng - 1]);
          }
        }
      }
    }
    pl.staged = false;
#ifdef OMNI3D_LAYOUT_STUDY
    if (iter_env && std::strcmp(iter_env, "staged") == 0) {
      const int rc = build_staged(ctx, nS, D, (int64_t)nnz);
      if (rc) return rc;
    }
#endif
  }
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  pl.X = X;
  pl.Y = Y;
  pl.Z = Z;
  pl.D = D;
  pl.ls = P.line_spacing;
  pl.g = g;
  pl.nrays = (int64_t)nrays;
  pl.steps = (int64_t)nsteps;
  pl.nS = nS;
  pl.nnz = (int64_t)nnz;
  pl.cmax = mxh[0];
  pl.lmax = mxh[1];
  static std::atomic<uint64_t> builds{0};
  pl.epoch = ++builds;
  pl.valid = true;
  ctx->dirs_dirty = false;
  return OMNI3D_OK;
}

void drop_iteration_graph(omni3d_ctx* ctx) {
  if (ctx->it_exec) (void)hipGraphExecDestroy(ctx->it_exec);
  if (ctx->it_graph) (void)hipGraphDestroy(ctx->it_graph);
  ctx->it_exec = nullptr;
  ctx->it_graph = nullptr;
  ctx->it_key.clear();
}

// Timing events of the phases (not recorded while a graph is being captured).
void record_event(omni3d_ctx* ctx, int i) {
  if (!ctx->capturing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

// partials of the measurement-only iteration layouts (none without them)
unsigned int study_grid(int nS) {
#ifdef OMNI3D_LAYOUT_STUDY
  return staged_grid(nS);
#else
  (void)nS;
  return 1;
#endif
}

bool trilinear(const omni3d_params& P) { return P.traversal == OMNI3D_TRAVERSAL_TRILINEAR; }

// The traversals of the plain method run one lane per ray (ray engine) on
// request, and always for the trilinear traversal, whose samples need the
// ray's line.
bool ray_engine(const omni3d_ctx* ctx) { return ctx->engine != 0 || trilinear(ctx->prm); }

// Bound on a visit's path, in units of rho * max(delta) * max|a|, for the
// fixed-point scale.  Staircase: the longest ray's steps.  Trilinear: the
// line (at most the steps) plus two legs below one voxel, each term a dot
// product of up to three components (sqrt 3), and trilinear samples that
// extrapolate up to a voxel beyond the face centres (at most 20 max|a| at a
// corner): 40 (L + 2).
double path_bound(const omni3d_params& P, const Plan& pl) {
  return trilinear(P) ? 40.0 * ((double)pl.lmax + 2.0) : (double)pl.lmax;
}

WalkParams walk_params(const omni3d_params& P) {
  WalkParams wp;
  wp.rho = P.density;
  const double del[3] = {P.dx, P.dy, P.dz};
  for (int a = 0; a < 3; a++) {
    // ((-rho * s) * delta) * 0.5, the factor of -rho*(next-last)*delta*0.5*(a_n+a_l)
    // (kernel.cu:1816) for s = -1 / +1
    wp.hs[a][0] = -P.density * -1.0 * del[a] * 0.5;
    wp.hs[a][1] = -P.density * 1.0 * del[a] * 0.5;
    wp.hsf[a][0] = (float)wp.hs[a][0];
    wp.hsf[a][1] = (float)wp.hs[a][1];
    wp.del[a] = del[a];
  }
  wp.thr = P.threshold;
  wp.method = P.method;
  wp.faithful = P.faithful_select;
  return wp;
}

// Groups of consecutive rows for the grouped iteration: at most maxrows rows
// each and G = the uniform group count rounded up to a multiple of ncu
// groups at most, with the smallest entry budget per group for which the
// greedy split fits in G.  The launch runs one workgroup per group and the
// CUs take G / ncu groups each, so equal groups make equal work per CU.
// grow: first row of each group and nS at the end; rgroup: the group of
// every row.
void balance_groups(const std::vector<unsigned long long>& rowptr, int maxrows, int ncu,
                    std::vector<int>* grow, std::vector<int>* rgroup) {
  const int nS = (int)rowptr.size() - 1;
  const unsigned long long nnz = rowptr[nS];
  long long G = std::max<long long>((nS + maxrows - 1) / maxrows, 1);
  G = (G + ncu - 1) / ncu * ncu;
  // greedy split: each group takes rows while its entries stay within the
  // budget (at least one row, at most maxrows), found by binary search on
  // the offsets, so a split costs O(groups log nS)
  auto split = [&](unsigned long long budget, bool store) -> long long {
    long long groups = 0;
    if (store) grow->clear();
    int r = 0;
    while (r < nS) {
      const int cap = std::min(nS, r + maxrows);
      const auto it = std::upper_bound(rowptr.begin() + r + 1, rowptr.begin() + cap + 1,
                                       rowptr[r] + budget);
      const int t = std::max(r + 1, (int)(it - rowptr.begin()) - 1);
      if (store) {
        grow->push_back(r);
        for (int q = r; q < t; q++) (*rgroup)[q] = (int)groups;
      }
      groups++;
      r = t;
    }
    if (store) grow->push_back(nS);
    return groups;
  };
  // the smallest budget whose split has at most G groups (a budget of nnz
  // gives the uniform count, which is at most G)
  unsigned long long lo = std::max<unsigned long long>(1ull, nnz / (unsigned long long)G), hi = std::max(nnz, 1ull);
  while (lo < hi) {
    const unsigned long long mid = lo + (hi - lo) / 2;
    if (split(mid, false) <= G) hi = mid;
    else lo = mid + 1;
  }
  (void)split(lo, true);
}

template <typename T>
int run_affine(omni3d_ctx* ctx, int seed) {
  const omni3d_params& P = ctx->prm;
  Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  DevScalars* sc = ctx->scal.as<DevScalars>();
  const void* field = ctx->field.as<void>();
  const double* pinit = seed >= 0 ? ctx->pinit.as<double>() : nullptr;
  const WalkParams wp = walk_params(P);
  // phase 1: up to the source iteration; phase 2: from its end (the batch
  // runs the iterations of all snapshots in between); 0: everything
  const int phase = ctx->affine_phase;
  const size_t ns = (size_t)std::max(pl.nS, 1);
  OMNI_TRY_HIP(ctx, ctx->bs.alloc(sizeof(long long) * ns));
  // segment sums, then carries, in the compute precision
  OMNI_TRY_HIP(ctx, ctx->segval.alloc(sizeof(T) * std::max<int64_t>(pl.nval, 1)));
  if (phase != 2) {
  OMNI_TRY_HIP(ctx, launch_fill_u32(sc, sizeof(DevScalars) / 4, 0u, st));
  // max |a| over the packed field (the fourth component is 0)
  OMNI_TRY_HIP(ctx, launch_absmax<T>(ctx->field.as<T>(), 4 * V, &sc->amax_bits, st));
  if (pinit) OMNI_TRY_HIP(ctx, launch_absmax<double>(pinit, V, &sc->pmax_bits, st));
  const double gfac = P.density * std::max(P.dx, std::max(P.dy, P.dz));
  OMNI_TRY_HIP(ctx, launch_compute_scale(sc, gfac, path_bound(P, pl), (double)pl.cmax,
                                         (double)P.max_iterations, st));
  // pass A: trapezoid sums landing on source voxels
  OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->bs.as<void>(), 2 * (int64_t)ns, 0u, st));
  record_event(ctx, 2);
  if (!ray_engine(ctx)) {
    // segment sums over all tiles, per-ray carries, then the source sums on
    // the surface tiles
    OMNI_TRY_HIP(ctx, launch_tile_pass<T>(kLocalSum, ctx->tile_threads, pl.ntiles, nullptr,
                                          pl.tile_ptr1.as<unsigned long long>(),
                                          pl.segs_slot.as<SegRec>(), pl.g, pl.tg, wp, field,
                                          pl.sid.as<int>(), pl.cnt.as<unsigned int>(), nullptr, nullptr, sc,
                                          ctx->segval.as<T>(),
                                          nullptr, st));
    record_event(ctx, 8);
    OMNI_TRY_HIP(ctx, launch_carry_scan(pl.nrays, pl.nseg.as<unsigned int>(),
                                        pl.val_off.as<unsigned long long>(),
                                        ctx->segval.as<T>(), st));
    OMNI_TRY_HIP(ctx, launch_tile_pass<T>(kSourceSums, ctx->tile_threads, pl.nsurf,
                                          pl.surf.as<int>(), pl.sseg_ptr.as<unsigned long long>(),
                                          pl.ssegs.as<SegRec>(), pl.g, pl.tg, wp, field,
                                          pl.sid.as<int>(), pl.cnt.as<unsigned int>(), nullptr, nullptr, sc,
                                          ctx->segval.as<T>(),
                                          ctx->bs.as<long long>(), st));
  } else {
    OMNI_TRY_HIP(ctx, launch_walk_plain<T>(false, trilinear(P), pl.rays.as<RayRec>(), pl.nrays,
                                           pl.dirs.as<DirConst>(), pl.g, wp, field, pl.sid.as<int>(),
                                           nullptr, nullptr, sc, ctx->bs.as<long long>(), 0, 1, st));
    record_event(ctx, 8);
  }
  record_event(ctx, 3);
  }
  // source iteration
  // p_S padded to whole 16-source lines (the staged iteration copies lines)
  const size_t ns16 = (ns + 15) & ~(size_t)15;
  OMNI_TRY_HIP(ctx, ctx->ps0.alloc(sizeof(double) * ns16));
  OMNI_TRY_HIP(ctx, ctx->ps1.alloc(sizeof(double) * ns16));
  OMNI_TRY_HIP(ctx, ctx->partials.alloc(
                        sizeof(double) * 2 *
                        std::max(iterate_grid(pl.nS),
                                 std::max(study_grid(pl.nS),
                                          std::max(dense_grid(pl.nS),
                                                   2 * grouped_grid_n(pl.dense_ngroups))))));
  if (phase != 2) {
    OMNI_TRY_HIP(ctx, launch_gather_sources(pl.slist.as<int>(), pl.nS, pinit, ctx->ps0.as<double>(), st));
    // max |p_S| of the starting values: the fixed-point scale of iteration 1
    OMNI_TRY_HIP(ctx, launch_absmax<double>(ctx->ps0.as<double>(), pl.nS, &sc->itmax_bits[0], st));
  }
  double* pbuf[2] = {ctx->ps0.as<double>(), ctx->ps1.as<double>()};
  // the grouped dense iteration runs one launch ahead of its stop rule, so
  // p_S rotates over three buffers and p_S^{k*-1} is copied out after the loop
  const bool rot3 = pl.dense && pl.dense_group > 1 && !pl.staged && P.max_iterations > 0;
  double* pb3[3] = {pbuf[0], pbuf[1], nullptr};
  if (rot3) {
    OMNI_TRY_HIP(ctx, ctx->ps2.alloc(sizeof(double) * ns16));
    OMNI_TRY_HIP(ctx, ctx->psel.alloc(sizeof(double) * ns16));
    pb3[2] = ctx->ps2.as<double>();
  }
  // FP32 ray sums: the grouped dense iteration runs in increment form
  const bool inc = std::is_same<T, float>::value && pl.dense && pl.dense_group > 1 &&
                   !std::getenv("OMNI3D_NO_INCREMENTS");
  if (inc) {
    OMNI_TRY_HIP(ctx, ctx->srow.alloc(sizeof(double) * ns));
    OMNI_TRY_HIP(ctx, ctx->dinc0.alloc(sizeof(float) * ns16));
    OMNI_TRY_HIP(ctx, ctx->dinc1.alloc(sizeof(float) * ns16));
  }
  float* dbuf[2] = {ctx->dinc0.as<float>(), ctx->dinc1.as<float>()};
  if (phase == 1) return OMNI3D_OK;
  auto iterations = [&](hipStream_t s) -> hipError_t {
    StagedIter si;
    si.nS = pl.nS;
    si.D = pl.D;
    si.nst = pl.st_n.as<unsigned int>();
    si.st_lo = pl.st_lo.as<unsigned int>();
    si.st_hi = pl.st_hi.as<unsigned int>();
    si.cse = pl.st_cse.as<unsigned int>();
    si.ulines = pl.ulines.as<unsigned int>();
    si.moff = pl.moff.as<unsigned long long>();
    si.masks = pl.masks.as<unsigned long long>();
    si.cell_eb = pl.cell_eb.as<unsigned int>();
    si.off = pl.soff.as<unsigned short>();
    for (int k = 1; k <= P.max_iterations; k++) {
      if (pl.dense && !pl.staged) {
        const hipError_t e = launch_iterate_dense(
            pl.dense_group, ctx->dense_variant, k, pl.nS, pl.rowptr.as<unsigned long long>(),
            pl.dense_wptr.as<unsigned int>(), pl.dent.as<unsigned int>(),
            pl.diag.as<unsigned int>(), pl.slist.as<int>(), pl.cnt.as<unsigned int>(),
            ctx->bs.as<long long>(), pinit, seed > 0 ? 1 : 0,
            rot3 ? pb3[(k - 1) % 3] : pbuf[(k - 1) & 1], rot3 ? pb3[k % 3] : pbuf[k & 1], sc,
            ctx->partials.as<double>(), P.max_iterations, P.tolerance, s,
            inc ? ctx->srow.as<double>() : nullptr, inc ? dbuf[(k - 1) & 1] : nullptr,
            inc ? dbuf[k & 1] : nullptr, pl.dense_group > 1 ? pl.dense_grow.as<int>() : nullptr,
            pl.dense_ngroups);
        if (e != hipSuccess) return e;
        if (rot3 && k == P.max_iterations) {
          // the stop rule of the last iteration (each grouped launch runs
          // the previous one's), then p_S^{k*-1} for pass B
          const unsigned int gg = grouped_grid_n(pl.dense_ngroups);
          hipError_t e2 = launch_iter_finish(
              k, (int)gg, ctx->partials.as<double>() + (size_t)(k & 1) * 2 * gg, sc,
              P.max_iterations, P.tolerance, s);
          if (e2 == hipSuccess)
            e2 = launch_select_ps(pl.nS, pb3[0], pb3[1], pb3[2], sc, ctx->psel.as<double>(), s);
          if (e2 != hipSuccess) return e2;
        }
        continue;
      }
#ifdef OMNI3D_LAYOUT_STUDY
      if (pl.staged) {
        const hipError_t e = launch_iterate_staged(
            si, k, pl.diag.as<unsigned int>(), pl.slist.as<int>(), pl.cnt.as<unsigned int>(),
            ctx->bs.as<long long>(), pinit, seed > 0 ? 1 : 0, pbuf[(k - 1) & 1], pbuf[k & 1], sc,
            ctx->partials.as<double>(), P.max_iterations, P.tolerance, s);
        if (e != hipSuccess) return e;
        continue;
      }
#endif
      const hipError_t e = launch_iterate(ctx->iter_waves, k, pl.nS, pl.moff.as<unsigned long long>(),
                                          pl.masks.as<unsigned long long>(),
                                          pl.cell_eb.as<unsigned int>(), pl.ent.as<int>(),
                                          pl.diag.as<unsigned int>(), pl.slist.as<int>(),
                                          pl.cnt.as<unsigned int>(), ctx->bs.as<long long>(), pinit,
                                          seed > 0 ? 1 : 0, pbuf[(k - 1) & 1], pbuf[k & 1], sc,
                                          ctx->partials.as<double>(), P.max_iterations,
                                          P.tolerance, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  };
  if (phase == 2) {
    // the iterations ran elsewhere: the stop rule of the last one, then
    // p_S^{k*-1} for pass B
    if (rot3) {
      const int k = P.max_iterations;
      const unsigned int gg = grouped_grid_n(pl.dense_ngroups);
      OMNI_TRY_HIP(ctx, launch_iter_finish(k, (int)gg,
                                           ctx->partials.as<double>() + (size_t)(k & 1) * 2 * gg,
                                           sc, P.max_iterations, P.tolerance, st));
      OMNI_TRY_HIP(ctx, launch_select_ps(pl.nS, pb3[0], pb3[1], pb3[2], sc,
                                         ctx->psel.as<double>(), st));
    }
  } else if (ctx->capturing) {
    OMNI_TRY_HIP(ctx, iterations(st));
  } else {
    // the iteration launches are replayed from a graph captured once per set
    // of arguments (no launch gaps between the short iterations)
    double tol = P.tolerance;
    uint64_t tolbits;
    std::memcpy(&tolbits, &tol, 8);
    const std::vector<uint64_t> key = {
        (uint64_t)P.max_iterations, (uint64_t)pl.nS, (uint64_t)(uintptr_t)pbuf[0],
        (uint64_t)(uintptr_t)pbuf[1], (uint64_t)(uintptr_t)sc,
        (uint64_t)(uintptr_t)ctx->partials.as<void>(), (uint64_t)(uintptr_t)ctx->bs.as<void>(),
        (uint64_t)(uintptr_t)pinit, (uint64_t)seed, tolbits,
        (uint64_t)(uintptr_t)pl.moff.as<void>(), (uint64_t)(uintptr_t)pl.masks.as<void>(),
        (uint64_t)(uintptr_t)pl.cell_eb.as<void>(), (uint64_t)(uintptr_t)pl.ent.as<void>(),
        (uint64_t)(uintptr_t)pl.diag.as<void>(), (uint64_t)(uintptr_t)pl.slist.as<void>(),
        (uint64_t)(uintptr_t)pl.cnt.as<void>(), (uint64_t)ctx->iter_waves, (uint64_t)pl.staged,
        (uint64_t)(uintptr_t)pl.soff.as<void>(), (uint64_t)(uintptr_t)pl.ulines.as<void>(),
        (uint64_t)(uintptr_t)pl.st_cse.as<void>(), (uint64_t)pl.dense,
        (uint64_t)(uintptr_t)pl.dent.as<void>(), (uint64_t)(uintptr_t)pl.rowptr.as<void>(), (uint64_t)ctx->dense_variant,
        (uint64_t)pl.dense_group, (uint64_t)inc, (uint64_t)(uintptr_t)ctx->srow.as<void>(),
        (uint64_t)rot3, (uint64_t)(uintptr_t)pb3[2], (uint64_t)(uintptr_t)ctx->psel.as<void>(),
        (uint64_t)(uintptr_t)dbuf[0], (uint64_t)(uintptr_t)dbuf[1]};
    if (!ctx->it_exec || key != ctx->it_key) {
      drop_iteration_graph(ctx);
      if (!ctx->cap_stream) OMNI_TRY_HIP(ctx, hipStreamCreate(&ctx->cap_stream));
      OMNI_TRY_HIP(ctx, hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
      const hipError_t e1 = iterations(ctx->cap_stream);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(ctx->cap_stream, &g);
      if (e1 != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        OMNI_TRY_HIP(ctx, e1 != hipSuccess ? e1 : e2);
      }
      ctx->it_graph = g;
      OMNI_TRY_HIP(ctx, hipGraphInstantiate(&ctx->it_exec, g, nullptr, nullptr, 0));
      ctx->it_key = key;
    }
    OMNI_TRY_HIP(ctx, hipGraphLaunch(ctx->it_exec, st));
  }
  record_event(ctx, 4);
  // pass B: every voxel from p_S^{k*-1}
  const int nc = ray_engine(ctx) ? ctx->acc_copies : 1;
  ctx->used_copies = nc;
  OMNI_TRY_HIP(ctx, ctx->sum.alloc(sizeof(long long) * V * nc));
  if (!ray_engine(ctx)) {
    // every voxel belongs to exactly one tile, which stores its sum
    record_event(ctx, 5);
    OMNI_TRY_HIP(ctx, launch_tile_pass<T>(kFull, ctx->tile_threads, pl.ntiles, nullptr,
                                          pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), pl.g, pl.tg, wp, field,
                                          pl.sid.as<int>(), pl.cnt.as<unsigned int>(),
                                          rot3 ? ctx->psel.as<double>() : pbuf[0],
                                          rot3 ? ctx->psel.as<double>() : pbuf[1], sc,
                                          ctx->segval.as<T>(),
                                          ctx->sum.as<long long>(), st));
  } else {
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sum.as<void>(), 2 * V * nc, 0u, st));
    record_event(ctx, 5);
    OMNI_TRY_HIP(ctx, launch_walk_plain<T>(true, trilinear(P), pl.rays.as<RayRec>(), pl.nrays,
                                           pl.dirs.as<DirConst>(), pl.g, wp, field, pl.sid.as<int>(),
                                           rot3 ? ctx->psel.as<double>() : pbuf[0],
                                           rot3 ? ctx->psel.as<double>() : pbuf[1], sc,
                                           ctx->sum.as<long long>(), V, nc, st));
  }
  record_event(ctx, 6);
  return OMNI3D_OK;
}


// Reference schedule (kernel.cu:8840-8861): every iteration walks every ray
// over the whole volume and averages.  Used by the NAIVE solver and by every
// method other than PLAIN.  Writes the result of iteration k* to pdst/wdst.
template <typename T>
int run_naive(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype) {
  ctx->am_cache.valid = false;  // the records and rows it keeps are rewritten here
  const omni3d_params& P = ctx->prm;
  Plan& pl = *ctx->plan;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  const int m = P.method;
  DevScalars* sc = ctx->scal.as<DevScalars>();
  const void* field = ctx->field.as<void>();
  const double* pinit = seed >= 0 ? ctx->pinit.as<double>() : nullptr;
  WalkParams wp = walk_params(P);
  OMNI_TRY_HIP(ctx, launch_fill_u32(sc, sizeof(DevScalars) / 4, 0u, st));
  OMNI_TRY_HIP(ctx, launch_absmax<T>(ctx->field.as<T>(), 4 * V, &sc->amax_bits, st));
  if (pinit) OMNI_TRY_HIP(ctx, launch_absmax<double>(pinit, V, &sc->pmax_bits, st));
  const double gfac = P.density * std::max(P.dx, std::max(P.dy, P.dz));
  OMNI_TRY_HIP(ctx, launch_compute_scale(sc, gfac, (double)pl.lmax, (double)pl.cmax,
                                         (double)P.max_iterations, st));
  // the field the path-break rules test
  const double* aux = nullptr;
  if (m != OMNI3D_METHOD_PLAIN) {
    OMNI_TRY_HIP(ctx, ctx->aux.alloc(sizeof(double) * V));
    const double* mask = ctx->have_mask ? ctx->mask.as<double>() : nullptr;
    if (m == OMNI3D_METHOD_PLAIN_MASK) {
      if (mask)
        OMNI_TRY_HIP(ctx, (launch_convert<double, double>(mask, ctx->aux.as<double>(), V, st)));
      else
        OMNI_TRY_HIP(ctx, launch_fill_f64(ctx->aux.as<double>(), V, 1.0, st));
      wp.thr = INFINITY;
      wp.faithful = 0;
    } else {
      OMNI_TRY_HIP(ctx, launch_curl_packed<T>(P.nx, P.ny, P.nz, P.dx, P.dy, P.dz, field, mask,
                                              ctx->aux.as<double>(), st));
    }
    aux = ctx->aux.as<double>();
  }
  // unit weights: plain, plain-mask and MiniCurl (its counts come from its own walk)
  const bool unitw = m == OMNI3D_METHOD_PLAIN_MASK || m == OMNI3D_METHOD_MINICURL;
  const bool fixed = m == OMNI3D_METHOD_PLAIN || unitw;
  // every method runs on the tile engine (OMNI3D_MASKED=naive: one lane per
  // ray with global atomics)
  const char* menv = std::getenv("OMNI3D_MASKED");
  // SelectFixedBC steps by its own rule, not along the plan's staircase
  const bool tiled = !(menv && std::strcmp(menv, "naive") == 0) &&
                     m != OMNI3D_METHOD_SELECT_FIXEDBC && m != OMNI3D_METHOD_MINICURL &&
                     m != OMNI3D_METHOD_WEIGHTED_FIXEDBC && m != OMNI3D_METHOD_WEIGHTED_MINICURL;
  if (tiled) {
    const size_t nv = (size_t)std::max<int64_t>(pl.nval, 1);
    const bool curlw = !fixed;
    OMNI_TRY_HIP(ctx, ctx->mk_e.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_s.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_c.alloc(sizeof(double) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_r.alloc(sizeof(uint32_t) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_f.alloc(sizeof(uint32_t) * nv));
    OMNI_TRY_HIP(ctx, ctx->mk_wc.alloc(sizeof(double) * (curlw ? nv : 1)));
    OMNI_TRY_HIP(ctx, ctx->mk_wn.alloc(sizeof(double) * (curlw ? nv : 1)));
  }
  OMNI_TRY_HIP(ctx, ctx->pa.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->pb.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->srcf.alloc((size_t)((V + 3) / 4) * 4));
  OMNI_TRY_HIP(ctx, ctx->partials.alloc(sizeof(double) * 2 * naive_update_grid(V)));
  if (fixed) {
    OMNI_TRY_HIP(ctx, ctx->sum.alloc(sizeof(long long) * V));
    if (unitw) OMNI_TRY_HIP(ctx, ctx->wcnt.alloc(sizeof(unsigned int) * V));
  } else {
    OMNI_TRY_HIP(ctx, ctx->sumd.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, ctx->wsum.alloc(sizeof(double) * V));
  }
  double* pa = ctx->pa.as<double>();
  double* pb = ctx->pb.as<double>();
  if (pinit)
    OMNI_TRY_HIP(ctx, (launch_convert<double, double>(pinit, pa, V, st)));
  else
    OMNI_TRY_HIP(ctx, launch_fill_f64(pa, V, 0.0, st));
  OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->srcf.as<void>(), (V + 3) / 4, 0u, st));
  // the weights depend on the field only: accumulated by the first iteration
  if (unitw) OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wcnt.as<void>(), V, 0u, st));
  if (!fixed) OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wsum.as<void>(), 2 * V, 0u, st));
  const unsigned int* cntw =
      m == OMNI3D_METHOD_PLAIN ? pl.cnt.as<unsigned int>() : ctx->wcnt.as<unsigned int>();
  record_event(ctx, 2);
  void* rule = nullptr;
  if (tiled && !mask_walk_lane()) {
    OMNI_TRY_HIP(ctx, ctx->mk_rule.alloc(mask_rule_bytes(m, pl.nval)));
    rule = ctx->mk_rule.as<void>();
  }
  if (tiled)
    OMNI_TRY_HIP(ctx, launch_mask_walk<T>(m, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                          pl.g, pl.tg, wp, field, aux, pl.nseg.as<unsigned int>(),
                                          pl.val_off.as<unsigned long long>(), pl.nval, pl.ntiles,
                                          pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), sc,
                                          ctx->mk_e.as<double>(), ctx->mk_s.as<double>(),
                                          ctx->mk_r.as<uint32_t>(), ctx->mk_f.as<uint32_t>(),
                                          ctx->mk_wc.as<double>(), ctx->mk_wn.as<double>(),
                                          m == OMNI3D_METHOD_PLAIN_MASK
                                              ? ctx->wcnt.as<unsigned int>() : nullptr,
                                          fixed ? nullptr : ctx->wsum.as<double>(),
                                          ctx->srcf.as<uint8_t>(), rule, st));
  record_event(ctx, 8);
  record_event(ctx, 3);
  for (int k = 1; k <= P.max_iterations; k++) {
    if (tiled) {
      OMNI_TRY_HIP(ctx, launch_mask_carry(pl.nrays, pl.nseg.as<unsigned int>(),
                                          pl.val_off.as<unsigned long long>(),
                                          ctx->mk_e.as<double>(), ctx->mk_s.as<double>(),
                                          ctx->mk_r.as<uint32_t>(), ctx->mk_f.as<uint32_t>(), pa,
                                          pb, k, sc, ctx->mk_c.as<double>(), st));
      OMNI_TRY_HIP(ctx, launch_tile_masked<T>(m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                              pl.segs.as<SegRec>(), pl.g, pl.tg, wp, field, aux,
                                              ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                              ctx->mk_wn.as<double>(), ctx->mk_c.as<double>(), pa,
                                              pb, k, sc,
                                              fixed ? ctx->sum.as<long long>() : nullptr,
                                              fixed ? nullptr : ctx->sumd.as<double>(), st));
      OMNI_TRY_HIP(ctx, launch_naive_update(k, V, fixed ? 1 : 0,
                                            fixed ? ctx->sum.as<long long>() : nullptr, cntw,
                                            fixed ? nullptr : ctx->sumd.as<double>(),
                                            fixed ? nullptr : ctx->wsum.as<double>(), pinit,
                                            seed > 0 ? 1 : 0,
                                            m == OMNI3D_METHOD_PLAIN ? pl.sid.as<int>() : nullptr,
                                            ctx->srcf.as<uint8_t>(), pa, pb, sc,
                                            ctx->partials.as<double>(), P.max_iterations,
                                            P.tolerance, st));
      continue;
    }
    if (fixed)
      OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sum.as<void>(), 2 * V, 0u, st));
    else
      OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sumd.as<void>(), 2 * V, 0u, st));
    const bool first = k == 1;
    OMNI_TRY_HIP(ctx, launch_walk_naive<T>(m, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                           pl.g, wp, field, aux, pa, pb, k, sc,
                                           fixed ? ctx->sum.as<long long>() : nullptr,
                                           (first && unitw) ? ctx->wcnt.as<unsigned int>()
                                                            : nullptr,
                                           fixed ? nullptr : ctx->sumd.as<double>(),
                                           (first && !fixed) ? ctx->wsum.as<double>() : nullptr,
                                           first ? ctx->srcf.as<uint8_t>() : nullptr, st));
    OMNI_TRY_HIP(ctx, launch_naive_update(k, V, fixed ? 1 : 0, fixed ? ctx->sum.as<long long>() : nullptr,
                                          cntw, fixed ? nullptr : ctx->sumd.as<double>(),
                                          fixed ? nullptr : ctx->wsum.as<double>(), pinit,
                                          seed > 0 ? 1 : 0,
                                          m == OMNI3D_METHOD_PLAIN ? pl.sid.as<int>() : nullptr,
                                          ctx->srcf.as<uint8_t>(), pa, pb, sc,
                                          ctx->partials.as<double>(), P.max_iterations,
                                          P.tolerance, st));
  }
  record_event(ctx, 4);
  record_event(ctx, 5);
  record_event(ctx, 6);
  const bool f64 = dtype == OMNI3D_F64;
  OMNI_TRY_HIP(ctx, launch_naive_output(V, pa, pb, sc, fixed ? 1 : 0, cntw,
                                        fixed ? nullptr : ctx->wsum.as<double>(),
                                        f64 ? (double*)pdst : nullptr, f64 ? nullptr : (float*)pdst,
                                        (wdst && f64) ? (double*)wdst : nullptr,
                                        (wdst && !f64) ? (float*)wdst : nullptr, st));
  return OMNI3D_OK;
}

// The masked methods (plain-mask, Select, Weighted) on their affine
// schedule (masked.hip): the rule replay of the field, the run starts S and
// the rows of the visits landing on them, the pass from p = 0 (the constant
// b), the iteration on S, and one full pass from p_S^{k*-1}.  Sets
// *fallback (nothing of the result queued) when the rows do not fit the
// iteration's fixed-point sums (a voxel visited 4096 times or more) or there
// is no run start, for the caller to run the reference schedule.
template <typename T>
int run_affine_masked(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype,
                      bool* fallback) {
  *fallback = false;
  const omni3d_params& P = ctx->prm;
  Plan& pl = *ctx->plan;
  if (pl.cmax >= 4096u || pl.nval <= 0 || P.max_iterations <= 0) {
    *fallback = true;
    return OMNI3D_OK;
  }
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  const int m = P.method;
  DevScalars* sc = ctx->scal.as<DevScalars>();
  const void* field = ctx->field.as<void>();
  const double* pinit = seed >= 0 ? ctx->pinit.as<double>() : nullptr;
  WalkParams wp = walk_params(P);
  OMNI_TRY_HIP(ctx, launch_fill_u32(sc, sizeof(DevScalars) / 4, 0u, st));
  OMNI_TRY_HIP(ctx, launch_absmax<T>(ctx->field.as<T>(), 4 * V, &sc->amax_bits, st));
  if (pinit) OMNI_TRY_HIP(ctx, launch_absmax<double>(pinit, V, &sc->pmax_bits, st));
  const double gfac = P.density * std::max(P.dx, std::max(P.dy, P.dz));
  OMNI_TRY_HIP(ctx, launch_compute_scale(sc, gfac, (double)pl.lmax, (double)pl.cmax,
                                         (double)P.max_iterations, st));
  // the field the path-break rule tests
  OMNI_TRY_HIP(ctx, ctx->aux.alloc(sizeof(double) * V));
  const double* mask = ctx->have_mask ? ctx->mask.as<double>() : nullptr;
  if (m == OMNI3D_METHOD_PLAIN_MASK) {
    if (mask)
      OMNI_TRY_HIP(ctx, (launch_convert<double, double>(mask, ctx->aux.as<double>(), V, st)));
    else
      OMNI_TRY_HIP(ctx, launch_fill_f64(ctx->aux.as<double>(), V, 1.0, st));
    wp.thr = INFINITY;
    wp.faithful = 0;
  } else {
    OMNI_TRY_HIP(ctx, launch_curl_packed<T>(P.nx, P.ny, P.nz, P.dx, P.dy, P.dz, field, mask,
                                            ctx->aux.as<double>(), st));
  }
  const double* aux = ctx->aux.as<double>();
  const bool fixed = m == OMNI3D_METHOD_PLAIN_MASK;  // unit weights
  const size_t nv = (size_t)pl.nval;
  OMNI_TRY_HIP(ctx, ctx->mk_e.alloc(sizeof(double) * nv));
  OMNI_TRY_HIP(ctx, ctx->mk_s.alloc(sizeof(double) * nv));
  OMNI_TRY_HIP(ctx, ctx->mk_c.alloc(sizeof(double) * nv));
  OMNI_TRY_HIP(ctx, ctx->mk_r.alloc(sizeof(uint32_t) * nv));
  OMNI_TRY_HIP(ctx, ctx->mk_f.alloc(sizeof(uint32_t) * nv));
  OMNI_TRY_HIP(ctx, ctx->mk_wc.alloc(sizeof(double) * (fixed ? 1 : nv)));
  OMNI_TRY_HIP(ctx, ctx->mk_wn.alloc(sizeof(double) * (fixed ? 1 : nv)));
  OMNI_TRY_HIP(ctx, ctx->am_run.alloc(sizeof(uint32_t) * nv));
  OMNI_TRY_HIP(ctx, ctx->pa.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->pb.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->srcf.alloc((size_t)((V + 3) / 4) * 4));
  if (fixed) {
    OMNI_TRY_HIP(ctx, ctx->sum.alloc(sizeof(long long) * V));
    OMNI_TRY_HIP(ctx, ctx->wcnt.alloc(sizeof(unsigned int) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wcnt.as<void>(), V, 0u, st));
  } else {
    OMNI_TRY_HIP(ctx, ctx->sumd.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, ctx->wsum.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wsum.as<void>(), 2 * V, 0u, st));
  }
  long long* sumfx = fixed ? ctx->sum.as<long long>() : nullptr;
  double* sumd = fixed ? nullptr : ctx->sumd.as<double>();
  unsigned int* wcnt = fixed ? ctx->wcnt.as<unsigned int>() : nullptr;
  double* wsum = fixed ? nullptr : ctx->wsum.as<double>();
  double* pa = ctx->pa.as<double>();
  double* pz = ctx->pb.as<double>();
  if (pinit)
    OMNI_TRY_HIP(ctx, (launch_convert<double, double>(pinit, pa, V, st)));
  else
    OMNI_TRY_HIP(ctx, launch_fill_f64(pa, V, 0.0, st));
  OMNI_TRY_HIP(ctx, launch_fill_f64(pz, V, 0.0, st));
  OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->srcf.as<void>(), (V + 3) / 4, 0u, st));
  record_event(ctx, 2);
  // plain-mask: the runs depend on the mask alone, so the mask-only
  // structures of the last field (the rule records F and R, the run-start
  // ids, the run sources, the rows as merged words) serve this one while the
  // mask and the plan are unchanged: then only E, the setup pass, b, the
  // iteration and the final pass run
  omni3d_ctx::AmCache& cache = ctx->am_cache;
  uint64_t mepoch = 0;
  if (m == OMNI3D_METHOD_PLAIN_MASK) {
    const int rc = mask_epoch_now(ctx, &mepoch);
    if (rc) return rc;
  }
  const bool reuse = m == OMNI3D_METHOD_PLAIN_MASK && cache.valid &&
                     cache.mask_epoch == mepoch && cache.plan_epoch == pl.epoch;
  cache.valid = false;
  if (reuse) {
    OMNI_TRY_HIP(ctx, launch_mask_reuse<T>(m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                           pl.segs.as<SegRec>(), pl.g, pl.tg, wp, field, aux, sc,
                                           ctx->mk_e.as<double>(), ctx->mk_s.as<double>(),
                                           ctx->mk_f.as<uint32_t>(), wcnt,
                                           ctx->srcf.as<uint8_t>(), st));
  } else {
    // the rule replay: first-visit records, weights, run starts
    void* rule = nullptr;
    if (!mask_walk_lane()) {
      OMNI_TRY_HIP(ctx, ctx->mk_rule.alloc(mask_rule_bytes(m, pl.nval)));
      rule = ctx->mk_rule.as<void>();
    }
    OMNI_TRY_HIP(ctx, launch_mask_walk<T>(m, pl.rays.as<RayRec>(), pl.nrays,
                                          pl.dirs.as<DirConst>(), pl.g, pl.tg, wp, field, aux,
                                          pl.nseg.as<unsigned int>(),
                                          pl.val_off.as<unsigned long long>(), pl.nval, pl.ntiles,
                                          pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), sc, ctx->mk_e.as<double>(),
                                          ctx->mk_s.as<double>(), ctx->mk_r.as<uint32_t>(),
                                          ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                          ctx->mk_wn.as<double>(), wcnt, wsum,
                                          ctx->srcf.as<uint8_t>(), rule, st));
  }
  record_event(ctx, 8);
  // S: compact ids of the run starts (the voxels a pass reads p at), in tile
  // order: tile t's rows are [tfirst[t], tfirst[t + 1])
  int nS = reuse ? cache.nS : 0;
  if (!reuse) {
    OMNI_TRY_HIP(ctx, ctx->am_sid.alloc(sizeof(int) * V));
    OMNI_TRY_HIP(ctx, ctx->am_slist.alloc(sizeof(int) * (V + 1)));
    OMNI_TRY_HIP(ctx, ctx->am_tfirst.alloc(sizeof(int) * ((size_t)pl.ntiles + 1)));
    DevBuf dns, bcnt;
    OMNI_TRY_HIP(ctx, dns.alloc(sizeof(int)));
    OMNI_TRY_HIP(ctx, bcnt.alloc(sizeof(unsigned int) * (size_t)source_ids_tiled_blocks(pl.tg)));
    OMNI_TRY_HIP(ctx, launch_source_ids_tiled(ctx->srcf.as<uint8_t>(), pl.g, pl.tg,
                                              bcnt.as<unsigned int>(), ctx->am_sid.as<int>(),
                                              ctx->am_slist.as<int>(), dns.as<int>(),
                                              ctx->am_tfirst.as<int>(), st));
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nS, dns.as<void>(), sizeof(int), hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (nS <= 0) {
      *fallback = true;
      return OMNI3D_OK;
    }
  }
  const int* sid = ctx->am_sid.as<int>();
  const int* slist = ctx->am_slist.as<int>();
  const int* tfirst = ctx->am_tfirst.as<int>();
  // the run each segment continues
  if (!reuse)
    OMNI_TRY_HIP(ctx, launch_mask_runsrc(pl.nrays, pl.nseg.as<unsigned int>(),
                                         pl.val_off.as<unsigned long long>(),
                                         ctx->mk_r.as<uint32_t>(), ctx->mk_f.as<uint32_t>(),
                                         ctx->am_run.as<uint32_t>(), st));
  // unit weights (plain-mask): the rows as the merged-word layout of the
  // plain method's source iteration (k_iterate_grp: groups of up to 512
  // rows sized to about equal entries, as many as fill the CUs evenly;
  // words sorted by source, equal entries merged), built from CSR rows;
  // otherwise the rows in tile layout (k_mtile_accum / k_mtile_finish)
  const bool grouped = fixed && nS < (1 << 22);
  unsigned long long nnz = 0;
  unsigned int nwords = 0;
  int nchunks = 0;
  DevBuf stmp;
  if (grouped && reuse) {
    nnz = cache.nnz;
    nwords = cache.nwords;
  } else if (grouped) {
    DevBuf rcnt;
    OMNI_TRY_HIP(ctx, rcnt.alloc(sizeof(unsigned int) * ((size_t)nS + 1)));
    OMNI_TRY_HIP(ctx, ctx->am_ptr.alloc(sizeof(unsigned long long) * ((size_t)nS + 1)));
    OMNI_TRY_HIP(ctx, launch_fill_u32(rcnt.as<void>(), (int64_t)nS + 1, 0u, st));
    OMNI_TRY_HIP(ctx, launch_tile_rows(false, m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                       pl.segs.as<SegRec>(), pl.g, pl.tg, wp, aux,
                                       ctx->mk_f.as<uint32_t>(), ctx->am_run.as<uint32_t>(), sid,
                                       rcnt.as<unsigned int>(), nullptr, nullptr, st));
    OMNI_TRY_HIP(ctx, stmp.alloc(scan_u32_tmp_bytes(nS)));
    OMNI_TRY_HIP(ctx, launch_scan_u32(rcnt.as<unsigned int>(), nS,
                                      ctx->am_ptr.as<unsigned long long>(), 0, stmp.as<void>(),
                                      stmp.size(), st));
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nnz, ctx->am_ptr.as<unsigned long long>() + nS, 8,
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (nnz == 0 || nnz >= (1ull << 31)) {
      *fallback = true;
      return OMNI3D_OK;
    }
    OMNI_TRY_HIP(ctx, ctx->am_src.alloc(sizeof(unsigned int) * (nnz + 1)));
    OMNI_TRY_HIP(ctx, launch_tile_rows(true, m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                       pl.segs.as<SegRec>(), pl.g, pl.tg, wp, aux,
                                       ctx->mk_f.as<uint32_t>(), ctx->am_run.as<uint32_t>(), sid,
                                       rcnt.as<unsigned int>(), ctx->am_ptr.as<unsigned long long>(),
                                       ctx->am_src.as<unsigned int>(), st));
    DevBuf keys, ka, kb, tmp, flag, excl;
    OMNI_TRY_HIP(ctx, keys.alloc(sizeof(unsigned long long) * nnz));
    OMNI_TRY_HIP(ctx, ka.alloc(sizeof(unsigned long long) * nnz));
    OMNI_TRY_HIP(ctx, kb.alloc(sizeof(unsigned long long) * nnz));
    OMNI_TRY_HIP(ctx, tmp.alloc(dense_tmp_bytes((int64_t)nnz)));
    OMNI_TRY_HIP(ctx, flag.alloc(sizeof(unsigned int) * nnz));
    OMNI_TRY_HIP(ctx, excl.alloc(sizeof(unsigned int) * nnz));
    // slack after the words: the iteration stages them with 16-byte loads
    OMNI_TRY_HIP(ctx, ctx->am_dent.alloc(sizeof(unsigned int) * (nnz + 8)));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_dent.as<void>(), (int64_t)nnz + 8, 0u, st));
    // groups of 8 chunks (4 from 2^21 run starts on), balanced as the plain
    // method's
    const int grp = nS < (1 << 21) ? 8 : 4;
    std::vector<unsigned long long> hrp((size_t)nS + 1);
    OMNI_TRY_HIP(ctx, hipMemcpy(hrp.data(), ctx->am_ptr.as<void>(),
                                sizeof(unsigned long long) * ((size_t)nS + 1),
                                hipMemcpyDeviceToHost));
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    std::vector<int> hgrow, hrg((size_t)nS);
    balance_groups(hrp, 64 * grp, std::max(ncu, 1), &hgrow, &hrg);
    const int ng = (int)hgrow.size() - 1;
    DevBuf rgroup;
    OMNI_TRY_HIP(ctx, ctx->am_grow.alloc(sizeof(int) * hgrow.size()));
    OMNI_TRY_HIP(ctx, rgroup.alloc(sizeof(int) * (size_t)nS));
    OMNI_TRY_HIP(ctx, hipMemcpy(ctx->am_grow.as<void>(), hgrow.data(), sizeof(int) * hgrow.size(),
                                hipMemcpyHostToDevice));
    OMNI_TRY_HIP(ctx, hipMemcpy(rgroup.as<void>(), hrg.data(), sizeof(int) * (size_t)nS,
                                hipMemcpyHostToDevice));
    ctx->am_group = grp;
    ctx->am_ngroups = ng;
    OMNI_TRY_HIP(ctx, ctx->am_wptr.alloc(sizeof(unsigned int) * ((size_t)ng + 1)));
    OMNI_TRY_HIP(ctx, launch_widen_keys((int64_t)nnz, ctx->am_src.as<unsigned int>(),
                                        keys.as<unsigned long long>(), st));
    OMNI_TRY_HIP(ctx, build_dense_entries(nS, (int64_t)nnz, grp,
                                          ctx->am_ptr.as<unsigned long long>(),
                                          keys.as<unsigned long long>(),
                                          ka.as<unsigned long long>(), kb.as<unsigned long long>(),
                                          tmp.as<void>(), tmp.size(),
                                          ctx->am_dent.as<unsigned int>(), st, rgroup.as<int>(),
                                          ctx->am_grow.as<int>(), ng));
    OMNI_TRY_HIP(ctx, merge_dense_entries(nS, (int64_t)nnz, grp,
                                          ctx->am_ptr.as<unsigned long long>(),
                                          kb.as<unsigned long long>(), flag.as<unsigned int>(),
                                          excl.as<unsigned int>(), tmp.as<void>(), tmp.size(),
                                          ctx->am_dent.as<unsigned int>(),
                                          ctx->am_wptr.as<unsigned int>(), &nwords, st,
                                          ctx->am_grow.as<int>(), ng));
    if (m == OMNI3D_METHOD_PLAIN_MASK) {
      cache.mask_epoch = mepoch;
      cache.plan_epoch = pl.epoch;
      cache.nS = nS;
      cache.nnz = nnz;
      cache.nwords = nwords;
    }
  } else {
    // entry counts per tile, their offsets, then the entries
    const int nt = pl.ntiles;
    DevBuf tcnt;
    OMNI_TRY_HIP(ctx, tcnt.alloc(sizeof(unsigned int) * ((size_t)nt + 1)));
    OMNI_TRY_HIP(ctx, ctx->am_tbase.alloc(sizeof(unsigned long long) * ((size_t)nt + 1)));
    OMNI_TRY_HIP(ctx, launch_tile_entries(false, m, nt, pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), pl.g, pl.tg, wp, aux,
                                          ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                          ctx->mk_wn.as<double>(), ctx->am_run.as<uint32_t>(),
                                          sid, tfirst, tcnt.as<unsigned int>(), nullptr, nullptr,
                                          nullptr, nullptr, st));
    OMNI_TRY_HIP(ctx, stmp.alloc(scan_u32_tmp_bytes(nt)));
    OMNI_TRY_HIP(ctx, launch_scan_u32(tcnt.as<unsigned int>(), nt,
                                      ctx->am_tbase.as<unsigned long long>(), 0, stmp.as<void>(),
                                      stmp.size(), st));
    std::vector<unsigned long long> hbase((size_t)nt + 1);
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(hbase.data(), ctx->am_tbase.as<void>(),
                                     sizeof(unsigned long long) * ((size_t)nt + 1),
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    nnz = hbase[nt];
    if (nnz == 0) {
      *fallback = true;
      return OMNI3D_OK;
    }
    OMNI_TRY_HIP(ctx, ctx->am_src.alloc(sizeof(unsigned int) * nnz));
    OMNI_TRY_HIP(ctx, ctx->am_erow.alloc(sizeof(unsigned short) * nnz));
    OMNI_TRY_HIP(ctx, ctx->am_w.alloc(sizeof(double) * (fixed ? 1 : nnz)));
    OMNI_TRY_HIP(ctx, launch_tile_entries(true, m, nt, pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), pl.g, pl.tg, wp, aux,
                                          ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                          ctx->mk_wn.as<double>(), ctx->am_run.as<uint32_t>(),
                                          sid, tfirst, nullptr,
                                          ctx->am_tbase.as<unsigned long long>(),
                                          ctx->am_src.as<unsigned int>(),
                                          ctx->am_erow.as<unsigned short>(),
                                          fixed ? nullptr : ctx->am_w.as<double>(), st));
    // chunks of at most csz entries of one tile, a workgroup each
    const unsigned long long csz = std::max<unsigned long long>(
        2048ull, std::min<unsigned long long>(65536ull, nnz / 4096ull));
    std::vector<MaskChunk> chunks;
    for (int t = 0; t < nt; t++)
      for (unsigned long long b = hbase[t]; b < hbase[t + 1]; b += csz)
        chunks.push_back(MaskChunk{b, std::min(hbase[t + 1], b + csz), t, 0});
    nchunks = (int)chunks.size();
    OMNI_TRY_HIP(ctx, ctx->am_chunks.alloc(sizeof(MaskChunk) * chunks.size()));
    OMNI_TRY_HIP(ctx, hipMemcpy(ctx->am_chunks.as<void>(), chunks.data(),
                                sizeof(MaskChunk) * chunks.size(), hipMemcpyHostToDevice));
    // max w of the fixed-point scale
    OMNI_TRY_HIP(ctx, ctx->am_wmax.alloc(sizeof(unsigned long long)));
    if (fixed) {
      const double one = 1.0;
      OMNI_TRY_HIP(ctx, hipMemcpy(ctx->am_wmax.as<void>(), &one, 8, hipMemcpyHostToDevice));
    } else {
      OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_wmax.as<void>(), 2, 0u, st));
      OMNI_TRY_HIP(ctx, launch_absmax<double>(ctx->am_w.as<double>(), (int64_t)nnz,
                                              ctx->am_wmax.as<unsigned long long>(), st));
    }
    OMNI_TRY_HIP(ctx, ctx->am_rowsum.alloc(sizeof(unsigned long long) * (size_t)nS));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_rowsum.as<void>(), 2 * (int64_t)nS, 0u, st));
  }
  // b: the pass from p = 0, read at the run starts, and the weights W there
  OMNI_TRY_HIP(ctx, launch_mask_carry(pl.nrays, pl.nseg.as<unsigned int>(),
                                      pl.val_off.as<unsigned long long>(), ctx->mk_e.as<double>(),
                                      ctx->mk_s.as<double>(), ctx->mk_r.as<uint32_t>(),
                                      ctx->mk_f.as<uint32_t>(), pz, pz, 1, sc,
                                      ctx->mk_c.as<double>(), st));
  OMNI_TRY_HIP(ctx, launch_tile_masked<T>(m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), pl.g, pl.tg, wp, field, aux,
                                          ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                          ctx->mk_wn.as<double>(), ctx->mk_c.as<double>(), pz, pz,
                                          1, sc, sumfx, sumd, st));
  const size_t ns = (size_t)nS;
  const size_t ns16 = (ns + 15) & ~(size_t)15;
  double* q0 = nullptr;
  double* q1 = nullptr;
  if (grouped) {
    // the plain method's iteration on these rows: p_new = (bs * inv + M p) / W
    // (no diagonal: a run start's own visits are entries), p_S over three
    // buffers (each launch runs the previous iteration's stop rule), the
    // increment form for FP32 ray sums
    OMNI_TRY_HIP(ctx, ctx->am_bs.alloc(sizeof(long long) * ns));
    OMNI_TRY_HIP(ctx, ctx->am_diag.alloc(sizeof(unsigned int) * ns));
    OMNI_TRY_HIP(ctx, launch_masked_bs(nS, slist, sumfx, ctx->am_bs.as<long long>(), st));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_diag.as<void>(), (int64_t)ns, 0u, st));
    record_event(ctx, 3);
    OMNI_TRY_HIP(ctx, ctx->ps0.alloc(sizeof(double) * ns16));
    OMNI_TRY_HIP(ctx, ctx->ps1.alloc(sizeof(double) * ns16));
    OMNI_TRY_HIP(ctx, ctx->ps2.alloc(sizeof(double) * ns16));
    OMNI_TRY_HIP(ctx, ctx->psel.alloc(sizeof(double) * ns16));
    double* pb3[3] = {ctx->ps0.as<double>(), ctx->ps1.as<double>(), ctx->ps2.as<double>()};
    const bool inc = std::is_same<T, float>::value && !std::getenv("OMNI3D_NO_INCREMENTS");
    if (inc) {
      OMNI_TRY_HIP(ctx, ctx->srow.alloc(sizeof(double) * ns));
      OMNI_TRY_HIP(ctx, ctx->dinc0.alloc(sizeof(float) * ns16));
      OMNI_TRY_HIP(ctx, ctx->dinc1.alloc(sizeof(float) * ns16));
    }
    float* dbuf[2] = {ctx->dinc0.as<float>(), ctx->dinc1.as<float>()};
    const unsigned int gg = grouped_grid_n(ctx->am_ngroups);
    OMNI_TRY_HIP(ctx, ctx->partials.alloc(sizeof(double) * 4 * gg));
    OMNI_TRY_HIP(ctx, launch_gather_sources(slist, nS, pinit, pb3[0], st));
    OMNI_TRY_HIP(ctx, launch_absmax<double>(pb3[0], nS, &sc->itmax_bits[0], st));
    for (int k = 1; k <= P.max_iterations; k++)
      OMNI_TRY_HIP(ctx, launch_iterate_dense(ctx->am_group, 0, k, nS,
                                             ctx->am_ptr.as<unsigned long long>(),
                                             ctx->am_wptr.as<unsigned int>(),
                                             ctx->am_dent.as<unsigned int>(),
                                             ctx->am_diag.as<unsigned int>(), slist, wcnt,
                                             ctx->am_bs.as<long long>(), pinit, seed > 0 ? 1 : 0,
                                             pb3[(k - 1) % 3], pb3[k % 3], sc,
                                             ctx->partials.as<double>(), P.max_iterations,
                                             P.tolerance, st, inc ? ctx->srow.as<double>() : nullptr,
                                             inc ? dbuf[(k - 1) & 1] : nullptr,
                                             inc ? dbuf[k & 1] : nullptr,
                                             ctx->am_grow.as<int>(), ctx->am_ngroups));
    const int K = P.max_iterations;
    OMNI_TRY_HIP(ctx, launch_iter_finish(K, (int)gg,
                                         ctx->partials.as<double>() + (size_t)(K & 1) * 2 * gg, sc,
                                         K, P.tolerance, st));
    OMNI_TRY_HIP(ctx, launch_select_ps(nS, pb3[0], pb3[1], pb3[2], sc, ctx->psel.as<double>(), st));
    q0 = q1 = ctx->psel.as<double>();
  } else {
    OMNI_TRY_HIP(ctx, ctx->am_b.alloc(sizeof(double) * ns));
    OMNI_TRY_HIP(ctx, ctx->am_wv.alloc(sizeof(double) * ns));
    OMNI_TRY_HIP(ctx, launch_masked_rowvec(nS, slist, fixed ? 1 : 0, sumfx, wcnt, sumd, wsum, sc,
                                           ctx->am_b.as<double>(), ctx->am_wv.as<double>(), st));
    record_event(ctx, 3);
    // the iteration on S from p_S^0
    OMNI_TRY_HIP(ctx, ctx->am_p0.alloc(sizeof(double) * ns));
    OMNI_TRY_HIP(ctx, ctx->am_p1.alloc(sizeof(double) * ns));
    q0 = ctx->am_p0.as<double>();
    q1 = ctx->am_p1.as<double>();
    OMNI_TRY_HIP(ctx, launch_gather_sources(slist, nS, pinit, q0, st));
    OMNI_TRY_HIP(ctx, launch_absmax<double>(q0, nS, &sc->itmax_bits[0], st));
    const unsigned int grid = mtile_finish_grid(nS);
    OMNI_TRY_HIP(ctx, ctx->partials.alloc(sizeof(double) * 2 * grid));
    for (int k = 1; k <= P.max_iterations; k++) {
      OMNI_TRY_HIP(ctx, launch_mtile_accum(k, nchunks, ctx->am_chunks.as<MaskChunk>(), tfirst,
                                           ctx->am_src.as<unsigned int>(),
                                           ctx->am_erow.as<unsigned short>(),
                                           fixed ? nullptr : ctx->am_w.as<double>(), q0, q1, sc,
                                           ctx->am_wmax.as<unsigned long long>(),
                                           ctx->am_rowsum.as<unsigned long long>(), st));
      OMNI_TRY_HIP(ctx, launch_mtile_finish(k, nS, ctx->am_b.as<double>(), ctx->am_wv.as<double>(),
                                            slist, pinit, seed > 0 ? 1 : 0, q0, q1,
                                            ctx->am_rowsum.as<unsigned long long>(), sc,
                                            ctx->am_wmax.as<unsigned long long>(),
                                            ctx->partials.as<double>(), st));
      OMNI_TRY_HIP(ctx, launch_iter_finish2(k, (int)grid, ctx->partials.as<double>(), sc,
                                            P.max_iterations, P.tolerance, st));
    }
  }
  record_event(ctx, 4);
  // the final pass (pass k*) from p_S^{k*-1}, on scalars whose stop flag is
  // clear; pa keeps the starting p off S (voxels no visit updates)
  OMNI_TRY_HIP(ctx, ctx->am_sc.alloc(sizeof(DevScalars)));
  DevScalars* scf = ctx->am_sc.as<DevScalars>();
  OMNI_TRY_HIP(ctx, launch_masked_final_scalars(sc, scf, st));
  OMNI_TRY_HIP(ctx, launch_masked_scatter(nS, slist, sc, q0, q1, pa, st));
  record_event(ctx, 5);
  OMNI_TRY_HIP(ctx, launch_mask_carry(pl.nrays, pl.nseg.as<unsigned int>(),
                                      pl.val_off.as<unsigned long long>(), ctx->mk_e.as<double>(),
                                      ctx->mk_s.as<double>(), ctx->mk_r.as<uint32_t>(),
                                      ctx->mk_f.as<uint32_t>(), pa, pa, 1, scf,
                                      ctx->mk_c.as<double>(), st));
  OMNI_TRY_HIP(ctx, launch_tile_masked<T>(m, pl.ntiles, pl.tile_ptr.as<unsigned long long>(),
                                          pl.segs.as<SegRec>(), pl.g, pl.tg, wp, field, aux,
                                          ctx->mk_f.as<uint32_t>(), ctx->mk_wc.as<double>(),
                                          ctx->mk_wn.as<double>(), ctx->mk_c.as<double>(), pa, pa,
                                          1, scf, sumfx, sumd, st));
  record_event(ctx, 6);
  const bool f64 = dtype == OMNI3D_F64;
  OMNI_TRY_HIP(ctx, launch_masked_final(V, fixed ? 1 : 0, sumfx, wcnt, sumd, wsum, pinit,
                                        seed > 0 ? 1 : 0, sc, pa, f64 ? (double*)pdst : nullptr,
                                        f64 ? nullptr : (float*)pdst,
                                        (wdst && f64) ? (double*)wdst : nullptr,
                                        (wdst && !f64) ? (float*)wdst : nullptr, st));
  ctx->am_nnz = (int64_t)nnz;
  ctx->am_words = grouped ? (int64_t)nwords : (int64_t)nnz;
  ctx->am_ns = nS;
  cache.valid = m == OMNI3D_METHOD_PLAIN_MASK && grouped;
  return OMNI3D_OK;
}

// The walks that leave the plan's staircase (SelectFixedBC, MiniCurl,
// WeightedFixedBC, WeightedMiniCurl) on the same affine schedule, one lane
// per ray: their runs, run starts and weights also depend on the curl field
// alone.  One pass from p = 0 gives b (the sums), W (the weights) and the
// run starts S; two walks with an entry sink count, then write, the visits
// landing on S as rows in tile layout (S numbered in tile order); the
// iteration on S is the curl-weighted masked methods' (k_mtile_accum +
// k_mtile_finish, unit weights for MiniCurl); one full pass from
// p_S^{k*-1} gives every voxel.  Sets *fallback (nothing of the result
// queued) where the rows do not fit the iteration (a row of 4096 entries or
// more) or there is no run start.
template <typename T>
int run_affine_lane(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype,
                    bool* fallback) {
  *fallback = false;
  ctx->am_cache.valid = false;  // its rows share the masked schedule's buffers
  const omni3d_params& P = ctx->prm;
  Plan& pl = *ctx->plan;
  if (!pl.valid || pl.ntiles <= 0 || pl.nrays <= 0 || P.max_iterations <= 0) {
    *fallback = true;
    return OMNI3D_OK;
  }
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  const int m = P.method;
  DevScalars* sc = ctx->scal.as<DevScalars>();
  const void* field = ctx->field.as<void>();
  const double* pinit = seed >= 0 ? ctx->pinit.as<double>() : nullptr;
  WalkParams wp = walk_params(P);
  OMNI_TRY_HIP(ctx, launch_fill_u32(sc, sizeof(DevScalars) / 4, 0u, st));
  OMNI_TRY_HIP(ctx, launch_absmax<T>(ctx->field.as<T>(), 4 * V, &sc->amax_bits, st));
  if (pinit) OMNI_TRY_HIP(ctx, launch_absmax<double>(pinit, V, &sc->pmax_bits, st));
  const double gfac = P.density * std::max(P.dx, std::max(P.dy, P.dz));
  OMNI_TRY_HIP(ctx, launch_compute_scale(sc, gfac, (double)pl.lmax, (double)pl.cmax,
                                         (double)P.max_iterations, st));
  OMNI_TRY_HIP(ctx, ctx->aux.alloc(sizeof(double) * V));
  const double* mask = ctx->have_mask ? ctx->mask.as<double>() : nullptr;
  OMNI_TRY_HIP(ctx, launch_curl_packed<T>(P.nx, P.ny, P.nz, P.dx, P.dy, P.dz, field, mask,
                                          ctx->aux.as<double>(), st));
  const double* aux = ctx->aux.as<double>();
  const bool fixed = m == OMNI3D_METHOD_MINICURL;  // unit weights
  OMNI_TRY_HIP(ctx, ctx->pa.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->pb.alloc(sizeof(double) * V));
  OMNI_TRY_HIP(ctx, ctx->srcf.alloc((size_t)((V + 3) / 4) * 4));
  if (fixed) {
    OMNI_TRY_HIP(ctx, ctx->sum.alloc(sizeof(long long) * V));
    OMNI_TRY_HIP(ctx, ctx->wcnt.alloc(sizeof(unsigned int) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sum.as<void>(), 2 * V, 0u, st));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wcnt.as<void>(), V, 0u, st));
  } else {
    OMNI_TRY_HIP(ctx, ctx->sumd.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, ctx->wsum.alloc(sizeof(double) * V));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sumd.as<void>(), 2 * V, 0u, st));
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->wsum.as<void>(), 2 * V, 0u, st));
  }
  long long* sumfx = fixed ? ctx->sum.as<long long>() : nullptr;
  double* sumd = fixed ? nullptr : ctx->sumd.as<double>();
  unsigned int* wcnt = fixed ? ctx->wcnt.as<unsigned int>() : nullptr;
  double* wsum = fixed ? nullptr : ctx->wsum.as<double>();
  double* pa = ctx->pa.as<double>();
  double* pz = ctx->pb.as<double>();
  if (pinit)
    OMNI_TRY_HIP(ctx, (launch_convert<double, double>(pinit, pa, V, st)));
  else
    OMNI_TRY_HIP(ctx, launch_fill_f64(pa, V, 0.0, st));
  OMNI_TRY_HIP(ctx, launch_fill_f64(pz, V, 0.0, st));
  OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->srcf.as<void>(), (V + 3) / 4, 0u, st));
  record_event(ctx, 2);
  // the pass from p = 0: b, W and the run starts
  OMNI_TRY_HIP(ctx, launch_walk_naive<T>(m, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                         pl.g, wp, field, aux, pz, pz, 1, sc, sumfx, wcnt, sumd,
                                         wsum, ctx->srcf.as<uint8_t>(), st));
  record_event(ctx, 8);
  // S in tile order: tile t's rows are [tfirst[t], tfirst[t + 1])
  OMNI_TRY_HIP(ctx, ctx->am_sid.alloc(sizeof(int) * V));
  OMNI_TRY_HIP(ctx, ctx->am_slist.alloc(sizeof(int) * (V + 1)));
  OMNI_TRY_HIP(ctx, ctx->am_tfirst.alloc(sizeof(int) * ((size_t)pl.ntiles + 1)));
  DevBuf dns, bcnt;
  OMNI_TRY_HIP(ctx, dns.alloc(sizeof(int)));
  OMNI_TRY_HIP(ctx, bcnt.alloc(sizeof(unsigned int) * (size_t)source_ids_tiled_blocks(pl.tg)));
  OMNI_TRY_HIP(ctx, launch_source_ids_tiled(ctx->srcf.as<uint8_t>(), pl.g, pl.tg,
                                            bcnt.as<unsigned int>(), ctx->am_sid.as<int>(),
                                            ctx->am_slist.as<int>(), dns.as<int>(),
                                            ctx->am_tfirst.as<int>(), st));
  int nS = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&nS, dns.as<void>(), sizeof(int), hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (nS <= 0) {
    *fallback = true;
    return OMNI3D_OK;
  }
  const int* sid = ctx->am_sid.as<int>();
  const int* slist = ctx->am_slist.as<int>();
  const int* tfirst = ctx->am_tfirst.as<int>();
  // the rows: counts, offsets (CSR over S, so each tile's entries are one
  // range), then the entries
  const int nt = pl.ntiles;
  DevBuf rcnt, dmax, stmp;
  OMNI_TRY_HIP(ctx, rcnt.alloc(sizeof(unsigned int) * ((size_t)nS + 1)));
  OMNI_TRY_HIP(ctx, dmax.alloc(sizeof(unsigned int)));
  OMNI_TRY_HIP(ctx, ctx->am_ptr.alloc(sizeof(unsigned long long) * ((size_t)nS + 1)));
  OMNI_TRY_HIP(ctx, ctx->am_tbase.alloc(sizeof(unsigned long long) * ((size_t)nt + 1)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(rcnt.as<void>(), (int64_t)nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_fill_u32(dmax.as<void>(), 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_lane_rows(false, m, pl.rays.as<RayRec>(), pl.nrays,
                                     pl.dirs.as<DirConst>(), pl.g, pl.tg, wp, aux, sid, tfirst,
                                     rcnt.as<unsigned int>(), nullptr, nullptr, nullptr, nullptr,
                                     st));
  OMNI_TRY_HIP(ctx, launch_max_u32(rcnt.as<unsigned int>(), nS, dmax.as<unsigned int>(), st));
  OMNI_TRY_HIP(ctx, stmp.alloc(scan_u32_tmp_bytes(nS)));
  OMNI_TRY_HIP(ctx, launch_scan_u32(rcnt.as<unsigned int>(), nS,
                                    ctx->am_ptr.as<unsigned long long>(), 0, stmp.as<void>(),
                                    stmp.size(), st));
  OMNI_TRY_HIP(ctx, launch_tile_bases(nt, tfirst, ctx->am_ptr.as<unsigned long long>(),
                                      ctx->am_tbase.as<unsigned long long>(), st));
  std::vector<unsigned long long> hbase((size_t)nt + 1);
  unsigned int rmax = 0;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(hbase.data(), ctx->am_tbase.as<void>(),
                                   sizeof(unsigned long long) * ((size_t)nt + 1),
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&rmax, dmax.as<void>(), sizeof(unsigned int),
                                   hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  const unsigned long long nnz = hbase[nt];
  if (nnz == 0 || rmax >= 4096u) {
    *fallback = true;
    return OMNI3D_OK;
  }
  OMNI_TRY_HIP(ctx, ctx->am_src.alloc(sizeof(unsigned int) * nnz));
  OMNI_TRY_HIP(ctx, ctx->am_erow.alloc(sizeof(unsigned short) * nnz));
  OMNI_TRY_HIP(ctx, ctx->am_w.alloc(sizeof(double) * (fixed ? 1 : nnz)));
  OMNI_TRY_HIP(ctx, launch_fill_u32(rcnt.as<void>(), (int64_t)nS + 1, 0u, st));
  OMNI_TRY_HIP(ctx, launch_lane_rows(true, m, pl.rays.as<RayRec>(), pl.nrays,
                                     pl.dirs.as<DirConst>(), pl.g, pl.tg, wp, aux, sid, tfirst,
                                     rcnt.as<unsigned int>(), ctx->am_ptr.as<unsigned long long>(),
                                     ctx->am_src.as<unsigned int>(),
                                     ctx->am_erow.as<unsigned short>(),
                                     fixed ? nullptr : ctx->am_w.as<double>(), st));
  // chunks of at most csz entries of one tile, a workgroup each
  const unsigned long long csz = std::max<unsigned long long>(
      2048ull, std::min<unsigned long long>(65536ull, nnz / 4096ull));
  std::vector<MaskChunk> chunks;
  for (int t = 0; t < nt; t++)
    for (unsigned long long b = hbase[t]; b < hbase[t + 1]; b += csz)
      chunks.push_back(MaskChunk{b, std::min(hbase[t + 1], b + csz), t, 0});
  const int nchunks = (int)chunks.size();
  OMNI_TRY_HIP(ctx, ctx->am_chunks.alloc(sizeof(MaskChunk) * chunks.size()));
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(ctx->am_chunks.as<void>(), chunks.data(),
                                   sizeof(MaskChunk) * chunks.size(), hipMemcpyHostToDevice, st));
  OMNI_TRY_HIP(ctx, ctx->am_wmax.alloc(sizeof(unsigned long long)));
  if (fixed) {
    OMNI_TRY_HIP(ctx, launch_fill_f64(reinterpret_cast<double*>(ctx->am_wmax.as<void>()), 1, 1.0,
                                      st));
  } else {
    // a weight 1 / curltmp * counttmp with curltmp = 0 (WeightedMiniCurl
    // tests no curltmp against zero) is not finite: such a field runs the
    // reference schedule, whose sums carry it as the reference's do
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_wmax.as<void>(), 2, 0u, st));
    OMNI_TRY_HIP(ctx, launch_absmax<double>(ctx->am_w.as<double>(), (int64_t)nnz,
                                            ctx->am_wmax.as<unsigned long long>(), st));
    double wmax = 0.0;
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&wmax, ctx->am_wmax.as<void>(), sizeof(double),
                                     hipMemcpyDeviceToHost, st));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (!std::isfinite(wmax)) {
      *fallback = true;
      return OMNI3D_OK;
    }
  }
  OMNI_TRY_HIP(ctx, ctx->am_rowsum.alloc(sizeof(unsigned long long) * (size_t)nS));
  OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->am_rowsum.as<void>(), 2 * (int64_t)nS, 0u, st));
  const size_t ns = (size_t)nS;
  OMNI_TRY_HIP(ctx, ctx->am_b.alloc(sizeof(double) * ns));
  OMNI_TRY_HIP(ctx, ctx->am_wv.alloc(sizeof(double) * ns));
  OMNI_TRY_HIP(ctx, launch_masked_rowvec(nS, slist, fixed ? 1 : 0, sumfx, wcnt, sumd, wsum, sc,
                                         ctx->am_b.as<double>(), ctx->am_wv.as<double>(), st));
  record_event(ctx, 3);
  // the iteration on S from p_S^0
  OMNI_TRY_HIP(ctx, ctx->am_p0.alloc(sizeof(double) * ns));
  OMNI_TRY_HIP(ctx, ctx->am_p1.alloc(sizeof(double) * ns));
  double* q0 = ctx->am_p0.as<double>();
  double* q1 = ctx->am_p1.as<double>();
  OMNI_TRY_HIP(ctx, launch_gather_sources(slist, nS, pinit, q0, st));
  OMNI_TRY_HIP(ctx, launch_absmax<double>(q0, nS, &sc->itmax_bits[0], st));
  const unsigned int grid = mtile_finish_grid(nS);
  OMNI_TRY_HIP(ctx, ctx->partials.alloc(sizeof(double) * 2 * grid));
  for (int k = 1; k <= P.max_iterations; k++) {
    OMNI_TRY_HIP(ctx, launch_mtile_accum(k, nchunks, ctx->am_chunks.as<MaskChunk>(), tfirst,
                                         ctx->am_src.as<unsigned int>(),
                                         ctx->am_erow.as<unsigned short>(),
                                         fixed ? nullptr : ctx->am_w.as<double>(), q0, q1, sc,
                                         ctx->am_wmax.as<unsigned long long>(),
                                         ctx->am_rowsum.as<unsigned long long>(), st));
    OMNI_TRY_HIP(ctx, launch_mtile_finish(k, nS, ctx->am_b.as<double>(), ctx->am_wv.as<double>(),
                                          slist, pinit, seed > 0 ? 1 : 0, q0, q1,
                                          ctx->am_rowsum.as<unsigned long long>(), sc,
                                          ctx->am_wmax.as<unsigned long long>(),
                                          ctx->partials.as<double>(), st));
    OMNI_TRY_HIP(ctx, launch_iter_finish2(k, (int)grid, ctx->partials.as<double>(), sc,
                                          P.max_iterations, P.tolerance, st));
  }
  record_event(ctx, 4);
  // the final pass (pass k*) from p_S^{k*-1} on scalars whose stop flag is
  // clear; pa keeps the starting p off S; the sums cleared, the weights kept
  OMNI_TRY_HIP(ctx, ctx->am_sc.alloc(sizeof(DevScalars)));
  DevScalars* scf = ctx->am_sc.as<DevScalars>();
  OMNI_TRY_HIP(ctx, launch_masked_final_scalars(sc, scf, st));
  OMNI_TRY_HIP(ctx, launch_masked_scatter(nS, slist, sc, q0, q1, pa, st));
  if (fixed)
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sum.as<void>(), 2 * V, 0u, st));
  else
    OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->sumd.as<void>(), 2 * V, 0u, st));
  record_event(ctx, 5);
  OMNI_TRY_HIP(ctx, launch_walk_naive<T>(m, pl.rays.as<RayRec>(), pl.nrays, pl.dirs.as<DirConst>(),
                                         pl.g, wp, field, aux, pa, pa, 1, scf, sumfx, nullptr, sumd,
                                         nullptr, nullptr, st));
  record_event(ctx, 6);
  const bool f64 = dtype == OMNI3D_F64;
  OMNI_TRY_HIP(ctx, launch_masked_final(V, fixed ? 1 : 0, sumfx, wcnt, sumd, wsum, pinit,
                                        seed > 0 ? 1 : 0, sc, pa, f64 ? (double*)pdst : nullptr,
                                        f64 ? nullptr : (float*)pdst,
                                        (wdst && f64) ? (double*)wdst : nullptr,
                                        (wdst && !f64) ? (float*)wdst : nullptr, st));
  ctx->am_nnz = (int64_t)nnz;
  ctx->am_words = (int64_t)nnz;
  ctx->am_ns = nS;
  return OMNI3D_OK;
}

// The mask epoch with a pending comparison resolved (one flag read, on the
// context's stream): caches keyed on the mask compare against this.
int mask_epoch_now(omni3d_ctx* ctx, uint64_t* epoch) {
  if (ctx->mask_cmp_pending) {
    unsigned int diff = 0;
    OMNI_TRY_HIP(ctx, hipMemcpyAsync(&diff, ctx->mask_diff.as<void>(), sizeof(diff),
                                     hipMemcpyDeviceToHost, ctx->stream));
    OMNI_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (diff) ctx->mask_epoch++;
    ctx->mask_cmp_pending = false;
  }
  *epoch = ctx->mask_epoch;
  return OMNI3D_OK;
}

// Whether a reconstruction's source iteration can run inside a
// multi-snapshot launch: the plain method on the affine solver with the
// grouped dense iteration (tile engine, p_S over three buffers).
bool affine_fusable(const omni3d_ctx* ctx) {
  const omni3d_params& P = ctx->prm;
  const Plan& pl = *ctx->plan;
  return P.solver == OMNI3D_SOLVER_AFFINE && P.method == OMNI3D_METHOD_PLAIN && pl.valid &&
         pl.dense && pl.dense_group > 1 && !pl.staged && !ray_engine(ctx) &&
         P.max_iterations > 0 && pl.nS > 0;
}

// The iteration descriptor of a prepared context (buffers sized by an
// earlier reconstruction), seedless (zero initial pressure).
IterDesc affine_iter_desc(omni3d_ctx* ctx, int first) {
  const Plan& pl = *ctx->plan;
  const bool inc = ctx->prm.precision == OMNI3D_FP32_KAHAN && !std::getenv("OMNI3D_NO_INCREMENTS");
  IterDesc d{};
  d.nS = pl.nS;
  d.nb = (int)grouped_grid_n(pl.dense_ngroups);
  d.grow = pl.dense_grow.as<int>();
  d.ngr = pl.dense_ngroups;
  d.first = first;
  d.seed = 0;
  d.wptr = pl.dense_wptr.as<unsigned int>();
  d.dent = pl.dent.as<unsigned int>();
  d.diag = pl.diag.as<unsigned int>();
  d.slist = pl.slist.as<int>();
  d.cnt = pl.cnt.as<unsigned int>();
  d.bs = ctx->bs.as<long long>();
  d.pinit = nullptr;
  d.pb[0] = ctx->ps0.as<double>();
  d.pb[1] = ctx->ps1.as<double>();
  d.pb[2] = ctx->ps2.as<double>();
  d.srow = inc ? ctx->srow.as<double>() : nullptr;
  d.db[0] = inc ? ctx->dinc0.as<float>() : nullptr;
  d.db[1] = inc ? ctx->dinc1.as<float>() : nullptr;
  d.sc = ctx->scal.as<DevScalars>();
  d.partials = ctx->partials.as<double>();
  return d;
}

// Queues a whole reconstruction of the context's field on its stream (no
// host synchronisation, except the masked methods' affine schedule, which
// sizes its rows on the host; every buffer must already have its size when
// this runs under graph capture).  The result goes to pdst / wdst (device,
// dtype).
int enqueue_solve(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype) {
  const omni3d_params& P = ctx->prm;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  int rc;
  const bool f64 = dtype == OMNI3D_F64;
  ctx->am_ns = -1;
  const bool fp32 = P.precision == OMNI3D_FP32_KAHAN;
  // the source-voxel (affine) schedule exists for the plain method; every
  // other method, and the NAIVE solver, walks the whole volume per iteration
  if (P.solver == OMNI3D_SOLVER_AFFINE && P.method == OMNI3D_METHOD_PLAIN) {
    rc = fp32 ? run_affine<float>(ctx, seed) : run_affine<double>(ctx, seed);
    if (rc) return rc;
    if (ctx->affine_phase == 1) return OMNI3D_OK;
    OMNI_TRY_HIP(ctx, launch_finalize(V, ctx->sum.as<long long>(), ctx->used_copies,
                                      ctx->plan->cnt.as<unsigned int>(),
                                      seed >= 0 ? ctx->pinit.as<double>() : nullptr, seed > 0 ? 1 : 0,
                                      ctx->scal.as<DevScalars>(), f64 ? (double*)pdst : nullptr,
                                      f64 ? nullptr : (float*)pdst,
                                      (wdst && f64) ? (double*)wdst : nullptr,
                                      (wdst && !f64) ? (float*)wdst : nullptr, st));
  } else {
    // the masked methods' affine schedule, where it applies (the tile engine;
    // not inside a captured graph: it sizes its rows on the host)
    const char* menv = std::getenv("OMNI3D_MASKED");
    bool fallback = true;
    if (P.solver == OMNI3D_SOLVER_AFFINE && ctx->affine_masked && !ctx->capturing &&
        !(menv && std::strcmp(menv, "naive") == 0) &&
        (P.method == OMNI3D_METHOD_PLAIN_MASK || P.method == OMNI3D_METHOD_SELECT ||
         P.method == OMNI3D_METHOD_WEIGHTED)) {
      rc = fp32 ? run_affine_masked<float>(ctx, seed, pdst, wdst, dtype, &fallback)
                : run_affine_masked<double>(ctx, seed, pdst, wdst, dtype, &fallback);
      if (rc) return rc;
    }
    // the walks off the plan's staircase: the same schedule, one lane per ray
    if (P.solver == OMNI3D_SOLVER_AFFINE && ctx->affine_masked && !ctx->capturing &&
        !(menv && std::strcmp(menv, "naive") == 0) &&
        (P.method == OMNI3D_METHOD_SELECT_FIXEDBC || P.method == OMNI3D_METHOD_MINICURL ||
         P.method == OMNI3D_METHOD_WEIGHTED_FIXEDBC ||
         P.method == OMNI3D_METHOD_WEIGHTED_MINICURL)) {
      rc = fp32 ? run_affine_lane<float>(ctx, seed, pdst, wdst, dtype, &fallback)
                : run_affine_lane<double>(ctx, seed, pdst, wdst, dtype, &fallback);
      if (rc) return rc;
    }
    if (fallback) {
      rc = fp32 ? run_naive<float>(ctx, seed, pdst, wdst, dtype)
                : run_naive<double>(ctx, seed, pdst, wdst, dtype);
      if (rc) return rc;
    }
  }
  return OMNI3D_OK;
}

template int run_naive<float>(omni3d_ctx*, int, void*, void*, int);
template int run_naive<double>(omni3d_ctx*, int, void*, void*, int);

}  // namespace omni3d

extern "C" {

void omni3d_default_params(omni3d_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->dx = p->dy = p->dz = 1.0;
  p->density = 1.0;
  p->line_spacing = 1.0;
  p->method = OMNI3D_METHOD_PLAIN;
  p->threshold = 0.0;
  p->faithful_select = 1;
  p->precision = OMNI3D_FP64;
  p->solver = OMNI3D_SOLVER_AFFINE;
  p->max_iterations = 10;
  p->tolerance = 0.0;
  p->seed_pn_with_p = 1;
}

const char* omni3d_last_error(const omni3d_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

const char* omni3d_last_global_error(void) { return g_last_error.c_str(); }

int omni3d_create(const omni3d_params* params, omni3d_ctx** out) {
  if (!params || !out) return fail(nullptr, OMNI3D_ERR_ARG, "null argument");
  std::string why;
  if (!params_ok(*params, &why)) return fail(nullptr, OMNI3D_ERR_ARG, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, OMNI3D_ERR_NODEVICE, "no HIP device visible");
  omni3d_ctx* ctx = new omni3d_ctx();
  ctx->prm = *params;
  if (const char* e = std::getenv("OMNI3D_ACC_COPIES")) {
    const int n = std::atoi(e);
    if (n >= 1 && n <= 64) ctx->acc_copies = n;
  }
  if (const char* e = std::getenv("OMNI3D_ENGINE")) ctx->engine = std::strcmp(e, "rays") == 0 ? 1 : 0;
  if (const char* e = std::getenv("OMNI3D_TILE_BITS")) {
    const int b = std::atoi(e);
    if (b >= kTileBitsMin && b <= kTileBitsMax) ctx->tile_bits = b;
  }
  if (const char* e = std::getenv("OMNI3D_TILE_THREADS")) {
    const int t = std::atoi(e);
    if (t == 256 || t == 512 || t == 1024) ctx->tile_threads = t;
  }
#ifdef OMNI3D_LAYOUT_STUDY
  if (const char* e = std::getenv("OMNI3D_ITER_WAVES")) {
    const int w = std::atoi(e);
    if (w == 1 || w == 2 || w == 4 || w == -1 || w == -2 || w == -4) ctx->iter_waves = w;
  }
#endif
  if (const char* e = std::getenv("OMNI3D_DENSE_GROUP")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 4 || v == 8) ctx->dense_group = v;
  }
#ifdef OMNI3D_LAYOUT_STUDY
  if (const char* e = std::getenv("OMNI3D_DENSE_VARIANT")) {
    const int v = std::atoi(e);
    if (v >= 0 && v <= 4) ctx->dense_variant = v;
  }
#endif
  for (auto& e : ctx->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return fail(nullptr, OMNI3D_ERR_HIP, "hipEventCreate failed");
    }
  }
  if (ctx->scal.alloc(sizeof(DevScalars)) != hipSuccess) {
    delete ctx;
    return fail(nullptr, OMNI3D_ERR_HIP, "hipMalloc failed");
  }
  *out = ctx;
  return OMNI3D_OK;
}

void omni3d_destroy(omni3d_ctx* ctx) {
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->stream);
  drop_iteration_graph(ctx);
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
}

int omni3d_set_stream(omni3d_ctx* ctx, void* s) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  ctx->stream = static_cast<hipStream_t>(s);
  return OMNI3D_OK;
}

int omni3d_set_tile_shape(omni3d_ctx* ctx, int tile_bits, int threads) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (tile_bits < kTileBitsMin || tile_bits > kTileBitsMax)
    return fail(ctx, OMNI3D_ERR_ARG, "tile_bits must be 3 (8^3 tiles) or 4 (16^3 tiles)");
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(ctx, OMNI3D_ERR_ARG, "threads must be 256, 512 or 1024");
  ctx->tile_bits = tile_bits;
  ctx->tile_threads = threads;
  return OMNI3D_OK;
}

int omni3d_set_directions(omni3d_ctx* ctx, const float* k1, const float* k2, const float* k3,
                          int count) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (!k1 || !k2 || !k3 || count < 1 || count > 65535)
    return fail(ctx, OMNI3D_ERR_ARG, "need 1..65535 direction vectors");
  std::vector<float> k(3 * (size_t)count);
  for (int i = 0; i < count; i++) {
    const double n = std::sqrt((double)k1[i] * k1[i] + (double)k2[i] * k2[i] + (double)k3[i] * k3[i]);
    if (!(n > 0) || !std::isfinite(n)) return fail(ctx, OMNI3D_ERR_ARG, "zero or non-finite direction");
    k[3 * i] = k1[i];
    k[3 * i + 1] = k2[i];
    k[3 * i + 2] = k3[i];
  }
  ctx->k.swap(k);
  ctx->dirs_dirty = true;
  return OMNI3D_OK;
}

int omni3d_load_directions_file(omni3d_ctx* ctx, const char* path, int count) {
  if (!ctx || !path) return fail(ctx, OMNI3D_ERR_ARG, "null argument");
  std::vector<float> k;
  std::string err;
  if (!read_direction_file(path, count, &k, &err)) return fail(ctx, OMNI3D_ERR_IO, err);
  const int n = (int)(k.size() / 3);
  std::vector<float> a(n), b(n), c(n);
  for (int i = 0; i < n; i++) {
    a[i] = k[3 * i];
    b[i] = k[3 * i + 1];
    c[i] = k[3 * i + 2];
  }
  return omni3d_set_directions(ctx, a.data(), b.data(), c.data(), n);
}

int omni3d_generate_directions(int divide, float* k1, float* k2, float* k3, int capacity) {
  if (divide < 1 || divide > 8) return -fail(nullptr, OMNI3D_ERR_ARG, "divide must be 1..8");
  const std::vector<double> v = icosahedral_directions(divide);
  const int n = (int)(v.size() / 3);
  for (int i = 0; i < n && i < capacity; i++) {
    if (k1) k1[i] = (float)v[3 * i];
    if (k2) k2[i] = (float)v[3 * i + 1];
    if (k3) k3[i] = (float)v[3 * i + 2];
  }
  return n;
}

int omni3d_set_field(omni3d_ctx* ctx, const void* dudt, const void* dvdt, const void* dwdt,
                     const void* mask, int dtype, int mem_kind) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (!dudt || !dvdt || !dwdt) return fail(ctx, OMNI3D_ERR_ARG, "null field component");
  if (dtype != OMNI3D_F32 && dtype != OMNI3D_F64) return fail(ctx, OMNI3D_ERR_ARG, "bad dtype");
  if (mem_kind != OMNI3D_MEM_HOST && mem_kind != OMNI3D_MEM_DEVICE)
    return fail(ctx, OMNI3D_ERR_ARG, "bad mem_kind");
  const int64_t V = voxels(ctx->prm);
  const bool f32 = ctx->prm.precision == OMNI3D_FP32_KAHAN;
  const size_t esz = elem_size(dtype);
  hipStream_t st = ctx->stream;
  DevBuf stage;
  const void* comp[4] = {dudt, dvdt, dwdt, mask};
  const void* dsrc[4] = {dudt, dvdt, dwdt, mask};
  if (mem_kind == OMNI3D_MEM_HOST) {
    OMNI_TRY_HIP(ctx, stage.alloc(esz * V * 4));
    for (int c = 0; c < 4; c++) {
      if (!comp[c]) continue;
      void* d = static_cast<char*>(stage.as<void>()) + esz * V * c;
      OMNI_TRY_HIP(ctx, hipMemcpyAsync(d, comp[c], esz * V, hipMemcpyHostToDevice, st));
      dsrc[c] = d;
    }
  }
  OMNI_TRY_HIP(ctx, ctx->field.alloc((f32 ? 4 : 8) * 4 * V));
  hipError_t e;
  if (dtype == OMNI3D_F64)
    e = f32 ? launch_pack_field<double, float>((const double*)dsrc[0], (const double*)dsrc[1],
                                               (const double*)dsrc[2], ctx->field.as<void>(), V, st)
            : launch_pack_field<double, double>((const double*)dsrc[0], (const double*)dsrc[1],
                                                (const double*)dsrc[2], ctx->field.as<void>(), V, st);
  else
    e = f32 ? launch_pack_field<float, float>((const float*)dsrc[0], (const float*)dsrc[1],
                                              (const float*)dsrc[2], ctx->field.as<void>(), V, st)
            : launch_pack_field<float, double>((const float*)dsrc[0], (const float*)dsrc[1],
                                               (const float*)dsrc[2], ctx->field.as<void>(), V, st);
  OMNI_TRY_HIP(ctx, e);
  if (mask) {
    // a mask over a mask of the same grid: converted beside it and compared
    // on the device, so that structures kept for the old mask stay valid when
    // the values are the same (a snapshot series with one mask)
    const bool compare = ctx->have_mask && ctx->mask.size() >= sizeof(double) * (size_t)V;
    DevBuf& dst = compare ? ctx->mask_next : ctx->mask;
    OMNI_TRY_HIP(ctx, dst.alloc(sizeof(double) * V));
    e = dtype == OMNI3D_F64
            ? launch_convert<double, double>((const double*)dsrc[3], dst.as<double>(), V, st)
            : launch_convert<float, double>((const float*)dsrc[3], dst.as<double>(), V, st);
    OMNI_TRY_HIP(ctx, e);
    if (compare) {
      if (!ctx->mask_cmp_pending) {
        OMNI_TRY_HIP(ctx, ctx->mask_diff.alloc(sizeof(unsigned int)));
        OMNI_TRY_HIP(ctx, launch_fill_u32(ctx->mask_diff.as<void>(), 1, 0u, st));
      }
      // (a comparison still pending accumulates into the same flag)
      OMNI_TRY_HIP(ctx, launch_mask_diff(ctx->mask.as<double>(), ctx->mask_next.as<double>(), V,
                                         ctx->mask_diff.as<unsigned int>(), st));
      // copied, not swapped: graphs captured over the mask keep its address
      OMNI_TRY_HIP(ctx, (launch_convert<double, double>(ctx->mask_next.as<double>(),
                                                        ctx->mask.as<double>(), V, st)));
      ctx->mask_cmp_pending = true;
    } else {
      ctx->mask_epoch++;
    }
  } else if (ctx->have_mask) {
    ctx->mask_epoch++;
  }
  if (mem_kind == OMNI3D_MEM_HOST) OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  ctx->have_field = true;
  ctx->have_mask = mask != nullptr;
  return OMNI3D_OK;
}

int omni3d_set_velocity(omni3d_ctx* ctx, const void* u, const void* v, const void* w,
                        double pref, int dtype, int mem_kind) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (!u) {
    ctx->have_seed = false;
    return OMNI3D_OK;
  }
  if (!v || !w) return fail(ctx, OMNI3D_ERR_ARG, "null velocity component");
  if (dtype != OMNI3D_F32 && dtype != OMNI3D_F64) return fail(ctx, OMNI3D_ERR_ARG, "bad dtype");
  if (mem_kind != OMNI3D_MEM_HOST && mem_kind != OMNI3D_MEM_DEVICE)
    return fail(ctx, OMNI3D_ERR_ARG, "bad mem_kind");
  if (!std::isfinite(pref)) return fail(ctx, OMNI3D_ERR_ARG, "pref must be finite");
  const omni3d_params& P = ctx->prm;
  const int64_t V = voxels(P);
  const size_t esz = elem_size(dtype);
  hipStream_t st = ctx->stream;
  DevBuf stage;
  const void* src[3] = {u, v, w};
  if (mem_kind == OMNI3D_MEM_HOST) {
    OMNI_TRY_HIP(ctx, stage.alloc(esz * V * 3));
    for (int c = 0; c < 3; c++) {
      void* d = static_cast<char*>(stage.as<void>()) + esz * V * c;
      OMNI_TRY_HIP(ctx, hipMemcpyAsync(d, src[c], esz * V, hipMemcpyHostToDevice, st));
      src[c] = d;
    }
  }
  OMNI_TRY_HIP(ctx, ctx->seed_face.alloc(sizeof(double) * (size_t)P.nx * P.nz));
  hipError_t e =
      dtype == OMNI3D_F64
          ? launch_bernoulli_face<double>(P.nx, P.ny, P.nz, (const double*)src[0],
                                          (const double*)src[1], (const double*)src[2], pref,
                                          P.density, ctx->seed_face.as<double>(), st)
          : launch_bernoulli_face<float>(P.nx, P.ny, P.nz, (const float*)src[0],
                                         (const float*)src[1], (const float*)src[2], pref,
                                         P.density, ctx->seed_face.as<double>(), st);
  OMNI_TRY_HIP(ctx, e);
  if (mem_kind == OMNI3D_MEM_HOST) OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  ctx->have_seed = true;
  return OMNI3D_OK;
}

int omni3d_field_curl(omni3d_ctx* ctx, void* out, int dtype, int mem_kind) {
  if (!ctx || !out) return fail(ctx, OMNI3D_ERR_ARG, "null argument");
  if (!ctx->have_field) return fail(ctx, OMNI3D_ERR_STATE, "no field set");
  if (dtype != OMNI3D_F32 && dtype != OMNI3D_F64) return fail(ctx, OMNI3D_ERR_ARG, "bad dtype");
  const omni3d_params& P = ctx->prm;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  DevBuf c, c32;
  OMNI_TRY_HIP(ctx, c.alloc(sizeof(double) * V));
  const double* mask = ctx->have_mask ? ctx->mask.as<double>() : nullptr;
  if (P.precision == OMNI3D_FP32_KAHAN)
    OMNI_TRY_HIP(ctx, launch_curl_packed<float>(P.nx, P.ny, P.nz, P.dx, P.dy, P.dz,
                                                ctx->field.as<void>(), mask, c.as<double>(), st));
  else
    OMNI_TRY_HIP(ctx, launch_curl_packed<double>(P.nx, P.ny, P.nz, P.dx, P.dy, P.dz,
                                                 ctx->field.as<void>(), mask, c.as<double>(), st));
  const void* src = c.as<void>();
  if (dtype == OMNI3D_F32) {
    OMNI_TRY_HIP(ctx, c32.alloc(sizeof(float) * V));
    OMNI_TRY_HIP(ctx, (launch_convert<double, float>(c.as<double>(), c32.as<float>(), V, st)));
    src = c32.as<void>();
  }
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(out, src, elem_size(dtype) * V,
                                   mem_kind == OMNI3D_MEM_HOST ? hipMemcpyDeviceToHost
                                                               : hipMemcpyDeviceToDevice, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  return OMNI3D_OK;
}

int omni3d_reconstruct(omni3d_ctx* ctx, void* p, void* weight_out, int dtype, int mem_kind,
                       omni3d_report* rep) {
  if (!ctx) return fail(nullptr, OMNI3D_ERR_ARG, "null context");
  if (dtype != OMNI3D_F32 && dtype != OMNI3D_F64) return fail(ctx, OMNI3D_ERR_ARG, "bad dtype");
  if (mem_kind != OMNI3D_MEM_HOST && mem_kind != OMNI3D_MEM_DEVICE)
    return fail(ctx, OMNI3D_ERR_ARG, "bad mem_kind");
  if (ctx->k.empty()) return fail(ctx, OMNI3D_ERR_STATE, "no directions set");
  if (!ctx->have_field) return fail(ctx, OMNI3D_ERR_STATE, "no field set");
  const omni3d_params& P = ctx->prm;
  hipStream_t st = ctx->stream;
  const int64_t V = voxels(P);
  const bool had_plan = ctx->plan->valid && !ctx->dirs_dirty;
  OMNI_TRY_HIP(ctx, hipEventRecord(ctx->ev[0], st));
  int rc = build_plan(ctx);
  if (rc) return rc;
  OMNI_TRY_HIP(ctx, hipEventRecord(ctx->ev[1], st));
  // initial pressure
  int seed = -1;
  if (p) {
    OMNI_TRY_HIP(ctx, ctx->pinit.alloc(sizeof(double) * V));
    DevBuf stage;
    const void* src = p;
    if (mem_kind == OMNI3D_MEM_HOST) {
      OMNI_TRY_HIP(ctx, stage.alloc(elem_size(dtype) * V));
      OMNI_TRY_HIP(ctx, hipMemcpyAsync(stage.as<void>(), p, elem_size(dtype) * V,
                                       hipMemcpyHostToDevice, st));
      src = stage.as<void>();
    }
    hipError_t e = dtype == OMNI3D_F64
                       ? launch_convert<double, double>((const double*)src, ctx->pinit.as<double>(), V, st)
                       : launch_convert<float, double>((const float*)src, ctx->pinit.as<double>(), V, st);
    OMNI_TRY_HIP(ctx, e);
    if (mem_kind == OMNI3D_MEM_HOST) OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
    seed = P.seed_pn_with_p ? 1 : 0;
  }
  if (ctx->have_seed) {
    // the j = 0 face gets the Bernoulli pressure (kernel.cu:8806-8811), over
    // the caller's p or zeros
    if (!p) {
      OMNI_TRY_HIP(ctx, ctx->pinit.alloc(sizeof(double) * V));
      OMNI_TRY_HIP(ctx, launch_fill_f64(ctx->pinit.as<double>(), V, 0.0, st));
    }
    OMNI_TRY_HIP(ctx, launch_apply_face(P.nx, P.ny, P.nz, ctx->seed_face.as<double>(),
                                        ctx->pinit.as<double>(), st));
    seed = P.seed_pn_with_p ? 1 : 0;
  }
  // averaging into the caller's buffers (device) or a staging buffer (host)
  const size_t esz = elem_size(dtype);
  void* pdst = p;
  void* wdst = weight_out;
  if (mem_kind == OMNI3D_MEM_HOST) {
    OMNI_TRY_HIP(ctx, ctx->pout.alloc(esz * V));
    pdst = ctx->pout.as<void>();
    if (weight_out) {
      OMNI_TRY_HIP(ctx, ctx->wout.alloc(esz * V));
      wdst = ctx->wout.as<void>();
    }
  }
  DevBuf ptmp;
  if (!pdst) {
    OMNI_TRY_HIP(ctx, ptmp.alloc(esz * V));
    pdst = ptmp.as<void>();
  }
  rc = enqueue_solve(ctx, seed, pdst, wdst, dtype);
  if (rc) return rc;
  OMNI_TRY_HIP(ctx, hipEventRecord(ctx->ev[7], st));
  if (mem_kind == OMNI3D_MEM_HOST) {
    if (p) OMNI_TRY_HIP(ctx, hipMemcpyAsync(p, pdst, esz * V, hipMemcpyDeviceToHost, st));
    if (weight_out)
      OMNI_TRY_HIP(ctx, hipMemcpyAsync(weight_out, wdst, esz * V, hipMemcpyDeviceToHost, st));
  }
  DevScalars hs;
  OMNI_TRY_HIP(ctx, hipMemcpyAsync(&hs, ctx->scal.as<void>(), sizeof(hs), hipMemcpyDeviceToHost, st));
  OMNI_TRY_HIP(ctx, hipStreamSynchronize(st));
  if (rep) {
    rep->iterations = hs.kstar;
    rep->rel_change = hs.rel;
    rep->rays = ctx->plan->nrays;
    rep->steps = ctx->plan->steps;
    rep->sources = ctx->plan->nS;
    rep->status = hs.status;
    auto ms = [&](int a, int b) {
      float t = 0;
      (void)hipEventElapsedTime(&t, ctx->ev[a], ctx->ev[b]);
      return (double)t;
    };
    rep->ms_plan = had_plan ? 0.0 : ms(0, 1);
    rep->ms_solve = ms(1, 7);
    rep->ms_pass_a = ms(2, 3);
    rep->ms_iterate = ms(3, 4);
    rep->ms_pass_b = ms(5, 6);
    rep->ms_finalize = ms(6, 7);
    rep->source_visits = ctx->plan->nnz;
    rep->segments = ctx->plan->nsegs;
    rep->sell_cells = ctx->plan->ncell;
    rep->matrix_words = ctx->plan->nwords;
    if (ctx->am_ns >= 0) {  // the masked methods' run starts and their rows
      rep->sources = ctx->am_ns;
      rep->source_visits = ctx->am_nnz;
      rep->matrix_words = ctx->am_words;
    }
    rep->ms_segment_sums = ms(2, 8);
  }
  if (hs.status) return fail(ctx, hs.status, "non-finite value in the input field or pressure");
  return OMNI3D_OK;
}

}  // extern "C"
