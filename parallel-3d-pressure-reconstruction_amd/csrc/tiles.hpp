// Tile engine data structures (see tiles.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace omni3d {

// Tile edge 2^bits voxels: 8^3 (bits 3) or 16^3 (bits 4), chosen per grid
// (the autotuner's "tile shape"); a segment then holds at most 3 * 2^bits - 2
// visits (22 or 46: its "segment length").
constexpr int kTileBitsMin = 3;
constexpr int kTileBitsMax = 4;
constexpr int kTileBitsDefault = 4;

enum { kLocalSum = 0, kSourceSums = 1, kFull = 2 };
constexpr int kLenBuckets = 12;  // segment-length groups per tile (2^lshift visits each)

struct TileGrid {
  int ntx, nty, ntz;
  int bits;
  int lshift = 2;  // segment-length group of a segment: visits >> lshift
};

// tile of voxel (i, j, k)
__device__ __forceinline__ int tile_of(int i, int j, int k, const TileGrid& tg) {
  const int b = tg.bits;
  return (i >> b) + (j >> b) * tg.ntx + (k >> b) * tg.ntx * tg.nty;
}

// LDS slot of in-tile voxel l = x + (y << b) + (z << 2b): x is XORed with
// bits of y and z, so that voxels one step apart along y or z -- the
// positions of neighbouring parallel rays at the same step -- fall in
// different LDS banks (a bijection within every (y, z) row).
template <int kBits>
__device__ __forceinline__ int lds_slot(int l) {
  const unsigned int u = (unsigned int)l;
  return (int)(u ^ __builtin_amdgcn_ubfe(u, kBits + 1, kBits) ^
               __builtin_amdgcn_ubfe(u, 2 * kBits, kBits));
}

// Trapezoid factor ((-rho * s) * delta_a) * 0.5 of a step along axis a with
// sign s (up: s = +1) in the arithmetic of T.
template <typename T>
__device__ __forceinline__ T step_factor(const WalkParams& wp, int a, bool up);
template <>
__device__ __forceinline__ float step_factor<float>(const WalkParams& wp, int a, bool up) {
  return up ? wp.hsf[a][1] : wp.hsf[a][0];
}
template <>
__device__ __forceinline__ double step_factor<double>(const WalkParams& wp, int a, bool up) {
  return up ? wp.hs[a][1] : wp.hs[a][0];
}

// One segment: the visits of one ray inside one tile, self-contained so that
// a tile pass starts it with one coalesced load.  A monotone staircase visits
// at most 16+15+15 = 46 voxels of a 16^3 tile, so its step codes fit in 96 bits.
struct SegRec {
  uint32_t g;        // value slot of the segment (its sum, then its carry)
  uint32_t src;      // source id of the ray (index of p_src)
  uint32_t info;     // local voxel of the first visit (12 bits) | visits (6) << 12 |
                     // step signs x,y,z (3) << 18 | first segment of the ray (1) << 21 |
                     // last segment of the ray (1) << 22 | axis of the step into
                     // the tile (2, not for the first segment) << 23
  uint32_t code[3];  // 2-bit axis codes: the steps between its visits, then the step
                     // out of the tile (not for the last segment), low bits first
};
constexpr int kSegLvBits = 12;

// The source rows (the visits that land on source voxels) built by the same
// walks: the count walk adds each ray's entry visit to diag[sid] and every
// later visit of a surface voxel with a source id to rowcnt[sid]; the fill
// walk puts the ray's key (direction << 32 | its source id) at rowptr[s] +
// cursor[s]++ for those visits.  Null pointers skip them.
struct SrcRows {
  const int* sid = nullptr;
  unsigned int* rowcnt = nullptr;  // count walk
  unsigned int* diag = nullptr;    // count walk
  const unsigned long long* rowptr = nullptr;  // fill walk
  unsigned int* cursor = nullptr;              // fill walk
  unsigned long long* key = nullptr;           // fill walk
};
hipError_t launch_tile_count(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                             TileGrid tg, unsigned int* nseg, unsigned int* tile_cnt,
                             const SrcRows& rows, hipStream_t st);
hipError_t launch_tile_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            TileGrid tg, const int* sid, const unsigned long long* val_off,
                            const unsigned long long* tile_ptr, unsigned int* cursor,
                            SegRec* segs, const SrcRows& rows, hipStream_t st);
// sorts every (tile, length bucket) list of segments by value slot into out
// (slot_ptr: the nslots + 1 list offsets; key_bits: bits of the largest slot)
size_t sort_segments_tmp_bytes(int64_t nsegs, int nslots);
hipError_t launch_sort_segments(SegRec* segs, SegRec* out, int64_t nsegs, int nslots,
                                const unsigned long long* slot_ptr, int key_bits, void* tmp,
                                size_t tmp_bytes, hipStream_t st);
// slots per group of 64 rays: 64 * the group's largest segment count
hipError_t launch_seg_groups(int64_t nrays, const unsigned int* nseg, unsigned int* gcnt,
                             hipStream_t st);
// threads: workgroup size of the tile passes, 256, 512 or 1024
// Compact per-tile lists of the segments that reach a source voxel (source-
// sum pass): pass 0 counts into cnt[tile], pass 1 writes at sptr[tile].
hipError_t launch_src_segs(int pass, int nsurf, const int* tile_list,
                           const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                           TileGrid tg, const int* sid, unsigned int* cnt,
                           const unsigned long long* sptr, SegRec* out, hipStream_t st);
// kSourceSums: tile_ptr / segs are the compact lists, kLocalSum the slot-
// ordered copy (one range per tile); kFull the length-grouped lists.
// segval[nval]: per value slot the segment sums (kLocalSum writes them),
// then the carries (kSourceSums, kFull read them), in the compute precision
template <typename T>
hipError_t launch_tile_pass(int mode, int threads, int ntiles, const int* tile_list,
                            const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                            TileGrid tg, const WalkParams& wp, const void* field, const int* sid,
                            const unsigned int* vcnt,
                            const double* p0, const double* p1, const DevScalars* sc,
                            T* segval, long long* out, hipStream_t st);
// masked sliced-ELL copy of the source CSR (64-row chunks, per-direction
// slots with lane masks, entries packed without padding); D <= 4096
hipError_t launch_sell_layout(int nS, int D, const unsigned long long* rowptr,
                              const unsigned long long* key, unsigned int* nslot, hipStream_t st);
hipError_t launch_sell_fill(int nS, int D, const unsigned long long* rowptr,
                            const unsigned long long* key, const unsigned long long* moff,
                            unsigned long long* masks, unsigned int* cell_eb,
                            unsigned short* sdir, int* ent, hipStream_t st);
// per-ray exclusive scan of the segment sums into carries, in place (FP64
// running sum; storage in the compute precision)
hipError_t launch_carry_scan(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, double* segval, hipStream_t st);
hipError_t launch_carry_scan(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, float* segval, hipStream_t st);
// exclusive scan of n uint32 counts into n+1 uint64 offsets (multi-block);
// `words` != 0 scans ceil(count/16) instead of count
hipError_t launch_scan_u32(const unsigned int* cnt, int64_t n, unsigned long long* out,
                           int words, void* tmp, size_t tmp_bytes, hipStream_t st);
size_t scan_u32_tmp_bytes(int64_t n);


// masked.hip: the masked methods' traversal per iteration (method 1
// plain-mask, 2 Select, 3 Weighted as in the naive walk)
// the per-field rule replay -- on the tile engine (k_rule_tile per segment
// and entry state, k_rule_scan per ray; `rule` = mask_rule_bytes of scratch)
// or, with rule == nullptr, one lane per ray (k_mask_walk) -- then the setup
// tile pass (segment sums S, the weights wcnt (plain-mask) or wsum (curl
// methods) and the run starts src)
template <typename T>
hipError_t launch_mask_walk(int method, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                            GridDesc g, TileGrid tg, const WalkParams& wp, const void* field,
                            const double* aux, const unsigned int* nseg,
                            const unsigned long long* val_off, int64_t nval, int ntiles,
                            const unsigned long long* tile_ptr, const SegRec* segs,
                            const DevScalars* sc, double* E, double* S, uint32_t* R, uint32_t* F,
                            double* wc, double* wn, unsigned int* wcnt, double* wsum,
                            uint8_t* src, void* rule, hipStream_t st);
size_t mask_rule_bytes(int method, int64_t nval);
// plain-mask records F kept from an earlier field with the same mask: the new
// field's entering-step terms E, then the setup pass (segment sums S, the
// weights wcnt, the run starts src)
template <typename T>
hipError_t launch_mask_reuse(int method, int ntiles, const unsigned long long* tile_ptr,
                             const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                             const void* field, const double* aux, const DevScalars* sc,
                             double* E, double* S, const uint32_t* F, unsigned int* wcnt,
                             uint8_t* src, hipStream_t st);
// the tile rule replay alone (no setup pass): E, R, F (wc, wn) per segment
template <typename T>
hipError_t launch_rule_replay(int method, int64_t nrays, const unsigned int* nseg,
                              const unsigned long long* val_off, int64_t nval, int ntiles,
                              const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                              TileGrid tg, const WalkParams& wp, const void* field,
                              const double* aux, double* E, uint32_t* R, uint32_t* F, double* wc,
                              double* wn, void* rule, hipStream_t st);
// naive.hip: the rows of the affine map of one pass for the walks that leave
// the plan's staircase (methods 4-7), one lane per ray: entries per run-start
// row (fill false: counted into cnt) or written (fill true: cnt = zeroed
// cursors; source id, row within its tile, weight unless ew is null)
hipError_t launch_lane_rows(bool fill, int method, const RayRec* rays, int64_t nrays,
                            const DirConst* dirs, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const double* aux, const int* sid, const int* tfirst,
                            unsigned int* cnt, const unsigned long long* rowptr,
                            unsigned int* esrc, unsigned short* erow, double* ew, hipStream_t st);
// tbase[t] = rowptr[tfirst[t]] for t = 0..nt
hipError_t launch_tile_bases(int nt, const int* tfirst, const unsigned long long* rowptr,
                             unsigned long long* tbase, hipStream_t st);
// OMNI3D_MASK_WALK=lane selects the one-lane-per-ray rule replay
bool mask_walk_lane();
hipError_t launch_mask_carry(int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, const double* E, const double* S,
                             const uint32_t* R, const uint32_t* F, const double* p0,
                             const double* p1, int k, const DevScalars* sc, double* cval,
                             hipStream_t st);
template <typename T>
hipError_t launch_tile_masked(int method, int ntiles, const unsigned long long* tile_ptr,
                              const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                              const void* field, const double* aux, const uint32_t* F,
                              const double* wc, const double* wn, const double* cval,
                              const double* p0, const double* p1, int k, const DevScalars* sc,
                              long long* sumfx, double* sumd, hipStream_t st);


// adjoint.hip: the plain method's transposed traversals on the tile engine
// (what 0: per-segment sums of u over the slot-ordered lists; 1: per-ray
// suffix sums in place, and each ray's total into tvol at its entry voxel
// when tvol is given; 2: the scatter of h * S into dg, every voxel stored)
hipError_t launch_adj_tiles(int what, int ntiles, const unsigned long long* tile_ptr1,
                            const SegRec* segs_slot, const unsigned long long* tile_ptr,
                            const SegRec* segs, const RayRec* rays, int64_t nrays,
                            const unsigned int* nseg, const unsigned long long* val_off,
                            GridDesc g, TileGrid tg, const WalkParams& wp, const double* u,
                            double* useg, double* tvol, double* dg, hipStream_t st);
// the plain-mask rule's transposed traversals on the tile engine (what 0:
// entering-run sums and span flags per segment; 1: per-ray reverse scan;
// 2: scatter into out = dg (3 V); 3: gather of run totals into out (V))
hipError_t launch_adjm_tiles(int what, int ntiles, const unsigned long long* tile_ptr1,
                             const SegRec* segs_slot, const unsigned long long* tile_ptr,
                             const SegRec* segs, int64_t nrays, const unsigned int* nseg,
                             const unsigned long long* val_off, GridDesc g, TileGrid tg,
                             const WalkParams& wp, const double* u, const double* mask,
                             double* useg, uint8_t* span, double* out, hipStream_t st);


// visit counts of every voxel from the tile segments (plan build)
hipError_t launch_tile_counts(int ntiles, const unsigned long long* tile_ptr1, const SegRec* segs,
                              GridDesc g, TileGrid tg, unsigned int* cnt, hipStream_t st);

// masked.hip: the masked methods' affine schedule ---------------------------
hipError_t launch_mask_runsrc(int64_t nrays, const unsigned int* nseg,
                              const unsigned long long* val_off, const uint32_t* R,
                              const uint32_t* F, uint32_t* runsrc, hipStream_t st);
// method 1 plain-mask, 2 Select, 3 Weighted; fill = false counts each row's
// entries into cnt, fill = true writes them (cnt zeroed: the rows' cursors)
hipError_t launch_tile_rows(bool fill, int method, int ntiles, const unsigned long long* tile_ptr,
                            const SegRec* segs, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const double* aux, const uint32_t* F, const uint32_t* runsrc,
                            const int* sid, unsigned int* cnt, const unsigned long long* rowptr,
                            unsigned int* esrc, hipStream_t st);
// compact ids of the flagged voxels in tile order, over many workgroups
// (bcnt: source_ids_tiled_blocks(tg) u32); tile t's rows are
// [tfirst[t], tfirst[t + 1]) (tfirst: ntiles + 1 ints)
hipError_t launch_source_ids_tiled(const uint8_t* flag, GridDesc g, TileGrid tg,
                                  unsigned int* bcnt, int* sid, int* slist, int* nS,
                                  int* tfirst, hipStream_t st);
// the masked rows in tile layout and their iteration
struct MaskChunk {
  unsigned long long begin, end;  // entries [begin, end) of one tile
  int tile, pad;
};
hipError_t launch_tile_entries(bool fill, int method, int ntiles,
                               const unsigned long long* tile_ptr, const SegRec* segs, GridDesc g,
                               TileGrid tg, const WalkParams& wp, const double* aux,
                               const uint32_t* F, const double* wc, const double* wn,
                               const uint32_t* runsrc, const int* sid, const int* tfirst,
                               unsigned int* tcnt, const unsigned long long* tbase,
                               unsigned int* esrc, unsigned short* erow, double* ew,
                               hipStream_t st);
hipError_t launch_mtile_accum(int k, int nchunks, const MaskChunk* chunks, const int* tfirst,
                              const unsigned int* esrc, const unsigned short* erow,
                              const double* ew, const double* p0, const double* p1,
                              const DevScalars* sc, const unsigned long long* wmax_bits,
                              unsigned long long* rowsum, hipStream_t st);
unsigned int mtile_finish_grid(int nS);
hipError_t launch_mtile_finish(int k, int nS, const double* bvec, const double* wvec,
                               const int* slist, const double* pinit, int seed, double* p0,
                               double* p1, unsigned long long* rowsum, DevScalars* sc,
                               const unsigned long long* wmax_bits, double* partials,
                               hipStream_t st);
int64_t source_ids_tiled_blocks(const TileGrid& tg);
hipError_t launch_masked_bs(int nS, const int* slist, const long long* sumfx, long long* bs,
                           hipStream_t st);
hipError_t launch_widen_keys(int64_t n, const unsigned int* esrc, unsigned long long* key,
                             hipStream_t st);
hipError_t launch_masked_rowvec(int nS, const int* slist, int fixed, const long long* sumfx,
                                const unsigned int* wcnt, const double* sumd, const double* wsum,
                                const DevScalars* sc, double* bvec, double* wvec,
                                hipStream_t st);
hipError_t launch_masked_scatter(int nS, const int* slist, const DevScalars* sc, const double* p0,
                                 const double* p1, double* pfull, hipStream_t st);
hipError_t launch_masked_final_scalars(const DevScalars* sc, DevScalars* out, hipStream_t st);
hipError_t launch_masked_final(int64_t V, int fixed, const long long* sumfx,
                               const unsigned int* wcnt, const double* sumd, const double* wsum,
                               const double* pinit, int seed, const DevScalars* sc,
                               const double* pprev, double* pd, float* pf, double* wd, float* wf,
                               hipStream_t st);
// the stop rule of iteration k for an iteration with two max |p| slots
hipError_t launch_iter_finish2(int k, int nparts, const double* partials, DevScalars* sc,
                               int max_iter, double tol, hipStream_t st);

}  // namespace omni3d
