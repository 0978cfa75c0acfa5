// Host-side internals of libomni3d: the context, the geometry plan and the
// device buffers they own.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/omni3d.h"
#include "geometry.hpp"
#include "kernels.hpp"
#include "tiles.hpp"

namespace omni3d {

std::vector<double> icosahedral_directions(int divide);
bool read_direction_file(const char* path, int count, std::vector<float>* k, std::string* err);
DirConst make_dir_const(float k1, float k2, float k3);

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t alloc(size_t bytes) {
    if (bytes <= size_) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&ptr_, bytes ? bytes : 1);
    if (e == hipSuccess) size_ = bytes;
    else ptr_ = nullptr;
    return e;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    size_ = 0;
  }
  void swap(DevBuf& o) {
    std::swap(ptr_, o.ptr_);
    std::swap(size_, o.size_);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t size() const { return size_; }

 private:
  void* ptr_ = nullptr;
  size_t size_ = 0;
};

// Geometry plan: everything that depends only on the grid, the direction set
// and the line spacing (not on the field), reused across snapshots.
struct Plan {
  bool valid = false;
  uint64_t epoch = 0;  // distinct per build: what a cache keyed on the plan compares
  int X = 0, Y = 0, Z = 0, D = 0;
  double ls = 0;
  GridDesc g{};
  int64_t nrays = 0;
  int64_t steps = 0;
  int nS = 0;
  int64_t nnz = 0;
  unsigned int cmax = 0;
  unsigned int lmax = 0;
  DevBuf dirs;     // DirConst[D]
  DevBuf rays;     // RayRec[nrays]
  DevBuf cnt;      // uint32[V] visits per voxel
  DevBuf sid;      // int32[V] compact source id or -1
  DevBuf slist;    // int32[nS] voxel of each source
  DevBuf rowptr;   // uint64[nS+1] row offsets of the visits landing on sources
  DevBuf diag;     // uint32[nS] rays entering at each source (their self visit)
  DevBuf moff;     // uint64[nchunks+1] first slot of each 64-row chunk
  DevBuf masks;    // uint64[ncell] lane mask of each (direction, k) slot
  DevBuf cell_eb;  // uint32[ncell] first packed entry of each slot
  DevBuf ent;      // int32[nnz] source ids packed in slot order
  DevBuf sdir;     // uint16[ncell] direction of each slot
  int64_t ncell = 0;
  // LDS-staged iteration (valid when staged): per (group, stage) line ranges,
  // distinct source lines, per-chunk stage ends and 16-bit LDS positions
  bool staged = false;
  int64_t nlines = 0;  // distinct (group, direction, line) triples
  DevBuf st_n;         // uint32[ngroups] stages of each group
  DevBuf st_lo, st_hi; // uint32[ngroups*D] distinct-line range of each stage
  DevBuf st_cse;       // uint32[ngroups*D*kStageGroup] end slot of each (stage, chunk)
  DevBuf ulines;       // uint32[nlines] source line (16 sources) of each distinct triple
  DevBuf soff;         // uint16[nnz+16] LDS position of each packed entry
  // dense iteration layout (valid when dense): per chunk, entries sorted by
  // source, word (row in chunk << 24 | source), in the CSR range of its rows
  bool dense = false;
  int dense_group = 1; // chunks per group of the words (1, 4 or 8)
  int64_t nwords = 0;  // dense words (merged equal entries for groups)
  DevBuf dent;         // uint32[nnz+8]
  DevBuf dense_wptr;   // uint32[ngroups+1] first word of each group (groups only)
  DevBuf dense_grow;   // int32[ngroups+1] first row of each group (groups only)
  int dense_ngroups = 0;
  // tile engine: compiled paths and per-tile segment lists
  TileGrid tg{};
  int ntiles = 0;
  int nsurf = 0;
  int64_t nsegs = 0;
  int64_t nval = 0;  // value slots (segments padded per group of 64 rays)
  DevBuf nseg;       // uint32[nrays] segments of each ray
  DevBuf val_off;    // uint64[ngroups+1] first value slot of each group of 64 rays
  DevBuf tile_ptr;  // uint64[ntiles+1]
  DevBuf segs;      // SegRec[nsegs], grouped by tile
  DevBuf surf;      // int32[nsurf] tiles that hold source voxels
  DevBuf segs_slot; // SegRec[nsegs] each tile's segments in value-slot order
  DevBuf tile_ptr1; // uint64[ntiles+1] their range per tile
  DevBuf ssegs;     // SegRec[...] segments reaching a source voxel, per surface tile
  DevBuf sseg_ptr;  // uint64[ntiles+1] their range per tile
  int64_t nssegs = 0;
  // transposed source rows of the plain method (built by the first
  // source-restricted backward): row s' = the sources that rays entering at
  // s' visit, sorted
  bool trows = false;
  DevBuf tptr;  // uint64[nS+1]
  DevBuf tent;  // uint32[nnz+1]
  // the same rows as merged words in groups of tgroup chunks sized to equal
  // entries (the layout of the forward iteration), when sizes allow
  bool tdense = false;
  int tgroup = 8;     // chunks per group (8, or 4 for 2^21 sources or more)
  int tngroups = 0;   // groups
  DevBuf tdent;   // uint32[nnz+8]
  DevBuf twptr;   // uint32[ngroups+1]
  DevBuf tgrow;   // int32[ngroups+1] first row of each group
};

}  // namespace omni3d

struct omni3d_ctx {
  omni3d_params prm{};
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<float> k;  // 3*D, row-major (k1,k2,k3)
  bool dirs_dirty = true;
  // geometry plan; contexts of one grid may share it (snapshot batches)
  std::shared_ptr<omni3d::Plan> plan = std::make_shared<omni3d::Plan>();
  int tile_bits = omni3d::kTileBitsDefault;  // tile edge 2^bits of the tile engine
  int tile_threads = 1024;                   // workgroup size of the tile passes
  int iter_waves = 2;                        // wavefronts per 64-row chunk of the iteration
  int dense_variant = 0;                     // dense iteration: wavefronts x round size
  int dense_group = 8;                       // dense iteration: chunks per row group (1, 4, 8)
  // the row groups of the grouped iteration are as many as fill the CUs
  // evenly (a lone reconstruction); a batch, whose snapshots share one
  // launch per iteration, keeps the uniform count
  bool groups_fill_cus = true;
  bool capturing = false;                    // work is being captured into a graph
  int affine_phase = 0;                      // 1: up to the source iteration, 2: after it
  // graph of the source-iteration launches, reused while its arguments stay
  hipStream_t cap_stream = nullptr;
  hipGraph_t it_graph = nullptr;
  hipGraphExec_t it_exec = nullptr;
  std::vector<uint64_t> it_key;
  // field in the compute precision, one 4-vector (a_x, a_y, a_z, 0) per voxel
  omni3d::DevBuf field, mask;
  bool have_field = false;
  bool have_mask = false;
  // Bernoulli seed of the j = 0 face (X*Z doubles), written into the
  // starting pressure of every reconstruction while set
  omni3d::DevBuf seed_face;
  bool have_seed = false;
  // per-reconstruction work buffers
  omni3d::DevBuf scal;      // DevScalars
  omni3d::DevBuf pinit;     // double[V]
  omni3d::DevBuf bs;        // int64[nS]
  omni3d::DevBuf ps0, ps1;  // double[nS]
  omni3d::DevBuf ps2, psel; // third p_S buffer of the grouped iteration, p_S^{k*-1}
  // increment form of the iteration (FP32 ray sums): row sums of p_S and the
  // FP32 increments of two consecutive iterations
  omni3d::DevBuf srow, dinc0, dinc1;
  omni3d::DevBuf partials;  // double[2*grid]
  omni3d::DevBuf sum;       // int64[ncopies*V]
  omni3d::DevBuf segval;    // T[nval] segment sums, then carries (T: compute precision)
  int acc_copies = 1;       // private accumulator copies of the ray-engine final pass
  int engine = 0;           // 0 = tile engine, 1 = ray engine (one lane per ray, atomics)
  omni3d::DevBuf bp[13];    // boundary-pair pipeline: keys, sums, runs, pairs, pressure
  int used_copies = 1;
  omni3d::DevBuf pout, wout;  // staging for host outputs
  // naive schedule: pressure double buffer, aux field (mask or |curl a|*mask),
  // curl-weighted sums, unit visit counts, flags of the voxels read
  omni3d::DevBuf pa, pb, aux, sumd, wsum, wcnt, srcf;
  // adjoint: dL/dp, Horner sum, scaled sum, gathered sums, dL/dg (3 V)
  omni3d::DevBuf adj_y, adj_z, adj_u, adj_t, adj_g;
  // source-restricted backward: t_S, h (2), w (2), C_S over the nS sources
  omni3d::DevBuf adj_s;
  // per-segment sums of the tile engine's transposed traversals
  omni3d::DevBuf adj_seg;
  omni3d::DevBuf adj_span;  // plain-mask: per segment, whether its entering run spans it
  // the plain-mask backward's per-mask structures, reused while the mask
  // (mask_epoch: bumped when set_field sets or removes a mask) and the plan
  // are unchanged: visit counts of the rule, and the run starts with their
  // transposed rows (sorted run-start ids per row, no self visits)
  uint64_t mask_epoch = 0;
  // a mask set over one of the same size is compared with it on the device
  // (mask_diff: nonzero if any value differs); the epoch moves only if so,
  // resolved by mask_epoch_now() when a cache asks
  bool mask_cmp_pending = false;
  omni3d::DevBuf mask_next, mask_diff;
  struct AdjMask {
    bool counts = false, rows = false;
    uint64_t mask_epoch = 0, plan_epoch = 0;
    int nS = 0;
    omni3d::DevBuf cnt, sid, slist, tptr, tent, diag;
    // the rows as merged words (the forward iteration's layout), when sizes
    // allow: chunks per group, groups, words, first word and row per group
    bool tdense = false;
    int tgroup = 8, tngroups = 0;
    omni3d::DevBuf tdent, twptr, tgrow;
  } adj_mask;
  // the transposed iteration on merged words: its scalars, partial sums and
  // fixed-point source term
  omni3d::DevBuf adj_sc, adj_part, adj_bs;
  // masked methods on the tile engine: per value slot the first-visit record
  // (E, S, R, F, curl weights) of the field and the carries of an iteration
  omni3d::DevBuf mk_e, mk_s, mk_r, mk_f, mk_wc, mk_wn, mk_c;
  omni3d::DevBuf mk_rule;  // the tile rule replay's per-entry-state records
  // the masked methods' affine schedule: per value slot the start voxel of
  // the run a segment continues; source ids and list of the run starts S;
  // the rows (pointers, source ids, weights); per row b, W and p_S (two);
  // max weight; the final pass's scalars
  omni3d::DevBuf am_run, am_sid, am_slist, am_ptr, am_src, am_w, am_b, am_wv, am_p0, am_p1;
  omni3d::DevBuf am_wmax, am_sc, am_dent, am_wptr, am_bs, am_diag;
  omni3d::DevBuf am_tfirst, am_tbase, am_erow, am_chunks, am_rowsum;
  // plain-mask rows as merged words: chunks per group, groups, first row of
  // each group (int32[ngroups+1])
  int am_group = 8, am_ngroups = 0;
  omni3d::DevBuf am_grow;
  bool affine_masked = true;  // off for snapshot batches (their graphs replay the reference schedule)
  // plain-mask on the affine schedule: the mask-only structures (rule
  // records F and R, run-start ids, run sources, the rows as merged words)
  // kept for the next field while the mask and the plan are unchanged
  struct AmCache {
    bool valid = false;
    uint64_t mask_epoch = 0, plan_epoch = 0;
    int nS = 0;
    unsigned long long nnz = 0;
    unsigned int nwords = 0;
  } am_cache;
  int am_ns = -1;             // run starts of the last reconstruction (-1: reference schedule)
  int64_t am_nnz = 0;         // its row entries
  int64_t am_words = 0;       // and the words the iteration reads
  // 0 start, 1 plan done, 2 pass A start, 3 pass A done, 4 iterations done,
  // 5 pass B start, 6 pass B done, 7 finalize done
  hipEvent_t ev[9] = {};
};
