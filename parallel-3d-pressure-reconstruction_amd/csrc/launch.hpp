// Host-side launchers of the kernels (each defined next to its kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace omni3d {

// Grid size for a grid-stride loop over n items with 256-thread blocks.
inline unsigned int grid_for(int64_t n, unsigned int cap = 65536) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned int)b;
}

// walk.hip -----------------------------------------------------------------
constexpr int kEnumItemsPerBlock = 1024;  // candidates per enumeration block
hipError_t launch_enum_count(const DirConst* dirs, GridDesc g, int64_t npts, int64_t ncand,
                             unsigned int* block_counts, hipStream_t st);
hipError_t launch_enum_write(const DirConst* dirs, GridDesc g, int64_t npts, int64_t ncand,
                             const unsigned long long* block_offsets, RayRec* rays,
                             hipStream_t st);
hipError_t launch_plan_flags(const RayRec* rays, int64_t nrays, GridDesc g, uint8_t* src,
                             unsigned long long* steps, hipStream_t st);
// Source numbering of the flagged surface voxels: face f's slots are
// [off[f], off[f + 1]), 8x8 blocks of its (u, w) plane (nu x nw voxels).
struct FaceSlots {
  int nu[6], nw[6];
  unsigned long long off[7];
};
FaceSlots face_slots(int X, int Y, int Z);
// valid[q] = 1 where slot q holds a flagged voxel of its face; *interior +=
// the flagged voxels off the surface
hipError_t launch_face_slot_flags(const FaceSlots& fs, GridDesc g, const uint8_t* flag,
                                  unsigned int* valid, unsigned int* interior, hipStream_t st);
// sid[v] = off[q], slist[off[q]] = v for the valid slots (sid set to -1
// beforehand)
hipError_t launch_face_slot_ids(const FaceSlots& fs, GridDesc g, const unsigned int* valid,
                                const unsigned long long* off, int* sid, int* slist,
                                hipStream_t st);
// field: float4 / double4 per voxel (a_x, a_y, a_z, 0); acc: ncopies accumulator
// copies copy_stride apart (block b adds into copy b % ncopies).
// trilinear: the straight-ray integral of the trilinear traversal instead of
// the staircase's trapezoids (same visits).
template <typename T>
hipError_t launch_walk_plain(bool full, bool trilinear, const RayRec* rays, int64_t nrays,
                             const DirConst* dirs, GridDesc g, const WalkParams& wp,
                             const void* field, const int* sid, const double* p0,
                             const double* p1, const DevScalars* sc, long long* acc,
                             int64_t copy_stride, int ncopies, hipStream_t st);

// inputs.hip ---------------------------------------------------------------
// out must be zeroed before the launch (it accumulates).
hipError_t launch_fingerprint(int D, const float* k1, const float* k2, const float* k3, int64_t V,
                              const float* ax, const float* ay, const float* az,
                              const float* curl, const float* p, InputPrint* out,
                              hipStream_t st);
template <typename S>
hipError_t launch_bernoulli_face(int X, int Y, int Z, const S* u, const S* v, const S* w,
                                 double pref, double rho, double* face, hipStream_t st);
hipError_t launch_apply_face(int X, int Y, int Z, const double* face, double* p, hipStream_t st);

// naive.hip ----------------------------------------------------------------
// method: 0 plain, 1 plain-mask, 2 select, 3 weighted.  Iteration k reads
// p0/p1[(k-1)&1]; fixed-point sums go to sumfx (+ unit counts to wcnt for
// plain-mask), curl-weighted sums to sumd/wsum; src flags the voxels read.
template <typename T>
hipError_t launch_walk_naive(int method, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                             GridDesc g, const WalkParams& wp, const void* field,
                             const double* aux, const double* p0, const double* p1, int k,
                             const DevScalars* sc, long long* sumfx, unsigned int* wcnt,
                             double* sumd, double* wsum, uint8_t* src, hipStream_t st);
unsigned int naive_update_grid(int64_t V);
hipError_t launch_naive_update(int k, int64_t V, int fixed, const long long* sumfx,
                               const unsigned int* cnt, const double* sumd, const double* wsum,
                               const double* pinit, int seed, const int* sid, const uint8_t* src,
                               double* p0, double* p1, DevScalars* sc, double* partials,
                               int max_iter, double tol, hipStream_t st);
hipError_t launch_naive_output(int64_t V, const double* p0, const double* p1,
                               const DevScalars* sc, int fixed, const unsigned int* cnt,
                               const double* wsum, double* pd, float* pf, double* wd, float* wf,
                               hipStream_t st);
template <typename T>
hipError_t launch_curl_packed(int X, int Y, int Z, double dx, double dy, double dz,
                              const void* field, const double* mask, double* out, hipStream_t st);
hipError_t launch_curl_f32(int X, int Y, int Z, double dx, double dy, double dz, const float* ax,
                           const float* ay, const float* az, const float* mask, float* out,
                           hipStream_t st);
hipError_t launch_fill_u32(void* x, int64_t words, unsigned int v, hipStream_t st);
hipError_t launch_fill_f64(double* x, int64_t n, double v, hipStream_t st);
// *diff = 1 when a and b differ in any of their n values (bitwise); else untouched
hipError_t launch_mask_diff(const double* a, const double* b, int64_t n, unsigned int* diff,
                            hipStream_t st);
hipError_t launch_accumulate_out(int64_t V, int fixed, const long long* sumfx,
                                 const unsigned int* cnt, const double* sumd, const double* wsum,
                                 const DevScalars* sc, float* pn, float* pc, hipStream_t st);
hipError_t launch_divide_count(int64_t V, float* p, float* pn, const float* pc, hipStream_t st);

// adjoint.hip --------------------------------------------------------------
hipError_t launch_adj_gather(bool mask_rule, const RayRec* rays, int64_t nrays,
                             const DirConst* dirs, GridDesc g, const double* mask,
                             const double* u, double* t, hipStream_t st);
hipError_t launch_adj_scatter(bool mask_rule, const RayRec* rays, int64_t nrays,
                              const DirConst* dirs, GridDesc g, const WalkParams& wp,
                              const double* mask, const double* u, double* dg, int64_t V,
                              hipStream_t st);
hipError_t launch_adj_count(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            const double* mask, unsigned int* cnt, hipStream_t st);
hipError_t launch_adj_scale(int64_t V, const double* z, const unsigned int* cnt, double* u,
                            hipStream_t st);
hipError_t launch_adj_horner(int64_t V, const double* y, const double* t,
                             const unsigned int* cnt, double* z, hipStream_t st);
// source-restricted backward of the plain method
hipError_t launch_adj_transpose_rows(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                     GridDesc g, const int* sid, unsigned int* tcnt,
                                     hipStream_t st);
hipError_t launch_adj_transpose_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                     GridDesc g, const int* sid, const unsigned long long* tptr,
                                     unsigned int* cursor, unsigned long long* key,
                                     hipStream_t st);
hipError_t launch_adj_mask_starts(const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                  GridDesc g, const double* mask, uint8_t* flag, hipStream_t st);
hipError_t launch_adj_mask_rows(int pass, const RayRec* rays, int64_t nrays, const DirConst* dirs,
                                GridDesc g, const double* mask, const int* sid,
                                const unsigned long long* tptr, unsigned int* cnt,
                                unsigned long long* key, hipStream_t st);
hipError_t launch_adj_narrow(const unsigned long long* key, int64_t n, unsigned int* out,
                             hipStream_t st);
// voxel-order compaction of 0/1 flags: cnt[v] = flag[v], then (after an
// exclusive scan into off) sid[v] = off[v] or -1, slist[off[v]] = v
hipError_t launch_flag_counts(const uint8_t* flag, int64_t V, unsigned int* cnt, hipStream_t st);
hipError_t launch_flag_ids(const uint8_t* flag, int64_t V, const unsigned long long* off, int* sid,
                           int* slist, hipStream_t st);
hipError_t launch_adj_src_init(int nS, const int* slist, const double* t, const unsigned int* cnt,
                               double* ts, double* h, double* w, double* cs, hipStream_t st);
hipError_t launch_adj_src_iter(int nS, const unsigned long long* tptr, const unsigned int* tent,
                               const unsigned int* diag, const double* cs, const double* ts,
                               const double* w, double* hn, double* wn, hipStream_t st);
hipError_t launch_adj_src_out(int nS, const int* slist, const double* h, const double* cs,
                              double* z, hipStream_t st);
hipError_t launch_adj_fixed(int nS, const double* t, DevScalars* sc, long long* bs,
                            hipStream_t st);

// solve.hip ----------------------------------------------------------------
template <typename T>
hipError_t launch_absmax(const T* x, int64_t n, unsigned long long* bits, hipStream_t st);
hipError_t launch_compute_scale(DevScalars* sc, double gfac, double lmax, double cmax,
                                double iters, hipStream_t st);
hipError_t launch_gather_sources(const int* slist, int nS, const double* pfull, double* ps,
                                 hipStream_t st);
unsigned int iterate_grid(int nS);
// waves: wavefronts per 64-row chunk (1, 2 or 4)
hipError_t launch_iterate(int waves, int k, int nS, const unsigned long long* moff,
                          const unsigned long long* masks, const unsigned int* cell_eb,
                          const int* ent, const unsigned int* diag, const int* slist,
                          const unsigned int* cnt, const long long* bs, const double* pinit,
                          int seed, const double* pold, double* pnew, DevScalars* sc,
                          double* partials, int max_iter, double tol, hipStream_t st);
hipError_t launch_sort_rows(int nS, const unsigned long long* rowptr, unsigned long long* key,
                            hipStream_t st);
template <typename S, typename T>
hipError_t launch_pack_field(const S* ax, const S* ay, const S* az, void* out, int64_t n,
                             hipStream_t st);
hipError_t launch_finalize(int64_t V, const long long* sum, int ncopies, const unsigned int* cnt,
                           const double* pinit, int seed, const DevScalars* sc, double* pd,
                           float* pf, double* wd, float* wf, hipStream_t st);
template <typename S, typename D>
hipError_t launch_convert(const S* s, D* d, int64_t n, hipStream_t st);
hipError_t launch_scan_counts(const unsigned int* cnt, int n, unsigned long long* ptr,
                              hipStream_t st);
hipError_t launch_max_u32(const unsigned int* x, int64_t n, unsigned int* out, hipStream_t st);
hipError_t launch_max_len(const RayRec* rays, int64_t n, unsigned int* out, hipStream_t st);

}  // namespace omni3d
