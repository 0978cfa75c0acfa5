// LDS-staged source iteration: plan construction (stage.hip) and launcher
// (solve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

namespace omni3d {

constexpr int kStageGroup = 4;   // 64-row chunks per workgroup (one wavefront each)
constexpr int kStageCap = 256;   // 128-byte source lines per stage (32 KB of LDS)

// Inputs and temporaries of the stage construction (device pointers).
struct StageBuild {
  int nS = 0, D = 0, nch = 0;
  int64_t nnz = 0;
  const unsigned long long* moff = nullptr;
  const unsigned long long* masks = nullptr;
  const unsigned short* sdir = nullptr;
  const unsigned int* cell_eb = nullptr;
  const int* ent = nullptr;
  unsigned long long* keys_a = nullptr;
  unsigned long long* keys_b = nullptr;
  unsigned int* vals_a = nullptr;
  unsigned int* vals_b = nullptr;
  unsigned int* flag = nullptr;
  unsigned int* incl = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
};

#ifdef OMNI3D_LAYOUT_STUDY
int stage_key_bits(int nS, int D);
size_t stage_tmp_bytes(int64_t nnz, int64_t ngd);
// sorted (group, direction, line) keys and the inclusive count of distinct keys
hipError_t stage_keys_sorted(const StageBuild& sb, hipStream_t st);
hipError_t stage_finish(const StageBuild& sb, int ngroups, unsigned int* ulines, unsigned int* gdcnt,
                        unsigned int* gdptr, unsigned int* nst, unsigned int* st_lo,
                        unsigned int* st_hi, unsigned short* st_dend, unsigned int* sbase,
                        unsigned int* cse, unsigned short* off, unsigned int* overflow,
                        hipStream_t st);
#endif

// One source iteration k over the staged layout, then the relative change
// and stop rule (as launch_iterate).
struct StagedIter {
  int nS = 0, D = 0;
  const unsigned int* nst = nullptr;
  const unsigned int* st_lo = nullptr;
  const unsigned int* st_hi = nullptr;
  const unsigned int* cse = nullptr;
  const unsigned int* ulines = nullptr;
  const unsigned long long* moff = nullptr;
  const unsigned long long* masks = nullptr;
  const unsigned int* cell_eb = nullptr;
  const unsigned short* off = nullptr;
};
#ifdef OMNI3D_LAYOUT_STUDY
unsigned int staged_grid(int nS);
#endif

// Dense layout: per 64-row chunk, entries sorted by source, one word
// (row in chunk << 24 | source) each; the chunk's words occupy the CSR range
// of its rows.
unsigned int dense_grid(int nS);
// Stop rule of iteration k alone (after the last grouped dense launch, whose
// own stop rule the next launch would have run); partials of iteration k.
hipError_t launch_iter_finish(int k, int nparts, const double* partials, DevScalars* sc,
                              int max_iter, double tol, hipStream_t st);
// The grouped iteration rotates p_S over three buffers (its launches run
// one iteration ahead of the stop rule): out = p_S of iteration k* - 1.
hipError_t launch_select_ps(int nS, const double* b0, const double* b1, const double* b2,
                            const DevScalars* sc, double* out, hipStream_t st);
// grid of the grouped dense iteration (one group per workgroup)
inline unsigned int grouped_grid_n(int ngroups) {
  return (unsigned int)(ngroups < 1 ? 1 : (ngroups > 65536 ? 65536 : ngroups));
}
// Row bits of a dense word for groups of gchunks chunks: (row << (32 - bits)) | source.
constexpr int dense_row_bits(int gchunks) { return gchunks <= 4 ? 8 : (gchunks <= 8 ? 9 : 10); }
// gchunks: chunks per group sharing the LDS row sums (1, 4, 8 or 16); the
// group's words are sorted by source.  For gchunks > 1 this leaves the sorted
// keys in keys_b for merge_dense_entries.
// With rgroup / grow (groups > 1 only): row r belongs to group rgroup[r],
// whose rows are [grow[g], grow[g + 1]) (at most 64 * gchunks), ngroups in
// all; without them, groups of 64 * gchunks consecutive rows.
hipError_t build_dense_entries(int nS, int64_t nnz, int gchunks, const unsigned long long* rowptr,
                               const unsigned long long* csr_key, unsigned long long* keys_a,
                               unsigned long long* keys_b, void* tmp, size_t tmp_bytes,
                               unsigned int* dent, hipStream_t st, const int* rgroup = nullptr,
                               const int* grow = nullptr, int ngroups = 0);
size_t dense_tmp_bytes(int64_t nnz);
// Groups (gchunks > 1): equal (group, source, row) keys of the sorted keys
// merged into words (row << (32 - b)) | (multiplicity - 1) << (30 - b) |
// source, b = dense_row_bits(gchunks) (sources below 2^(30 - b)),
// wptr[ngroups+1] the first word of every group.
hipError_t merge_dense_entries(int nS, int64_t nnz, int gchunks, const unsigned long long* rowptr,
                               const unsigned long long* keys, unsigned int* flag,
                               unsigned int* excl, void* tmp, size_t tmp_bytes, unsigned int* dent,
                               unsigned int* wptr, unsigned int* nwords, hipStream_t st,
                               const int* grow = nullptr, int ngroups = 0);
// gchunks 1: one chunk per wavefront, variant = chunks per workgroup x
// entries per round (0: 4 x 512, 1: 8 x 512, 2: 8 x 256, 3: 4 x 256,
// 4: 4 x 1024); gchunks 4 or 8: one group per workgroup of 4 or 8
// wavefronts sharing the group's LDS row sums, merged words from wptr; partials then hold two
// halves (iterations alternate) and each launch runs the stop rule of the
// previous iteration, max |.| slots k % 3 -- launch_iter_finish ends the loop.  With srow (groups only) the
// increment form: srow[nS] FP64 row sums of p_S kept across iterations,
// din / dnew the FP32 increments of the previous / this iteration.
hipError_t launch_iterate_dense(int gchunks, int variant, int k, int nS,
                                const unsigned long long* rowptr, const unsigned int* wptr,
                                const unsigned int* dent, const unsigned int* diag,
                                const int* slist, const unsigned int* cnt, const long long* bs,
                                const double* pinit, int seed, const double* pold, double* pnew,
                                DevScalars* sc, double* partials, int max_iter, double tol,
                                hipStream_t st, double* srow = nullptr,
                                const float* din = nullptr, float* dnew = nullptr,
                                const int* grow = nullptr, int ngroups = 0);
// One snapshot's share of a multi-snapshot launch of the grouped iteration:
// its nb = grouped_grid_n(ngr) workgroups start at block `first`; p_S rotates
// over pb[0..2] and the increments over db[0..1] as in its own launches.
struct IterDesc {
  int nS, nb, first, seed;
  const unsigned int* wptr;
  const unsigned int* dent;
  const int* grow;  // first row of each group (ngr + 1 entries)
  int ngr;          // groups
  const unsigned int* diag;
  const int* slist;
  const unsigned int* cnt;
  const long long* bs;
  const double* pinit;
  double* pb[3];
  double* srow;
  float* db[2];
  DevScalars* sc;
  double* partials;
};
// iteration k of every snapshot described (gchunks: chunks per group, 4 or 8,
// the same for all; inc: increment form, FP32 ray sums); blk_snap[b] = the
// snapshot of workgroup b
hipError_t launch_iterate_multi(int gchunks, int inc, int k, const IterDesc* desc,
                                const unsigned int* blk_snap, unsigned int nblocks, int max_iter,
                                double tol, hipStream_t st);
#ifdef OMNI3D_LAYOUT_STUDY
hipError_t launch_iterate_staged(const StagedIter& si, int k, const unsigned int* diag,
                                 const int* slist, const unsigned int* cnt, const long long* bs,
                                 const double* pinit, int seed, const double* pold, double* pnew,
                                 DevScalars* sc, double* partials, int max_iter, double tol,
                                 hipStream_t st);
#endif

}  // namespace omni3d
