// Plan construction of the source iteration's word layouts.
//
// Dense words (the default): per group of 64-row chunks, the CSR entries
// sorted by (source, row) and equal entries merged into one word
// (row << (32 - b)) | (multiplicity - 1) << (30 - b) | source, with b = 8 row
// bits for groups of 4 chunks and 9 for groups of 8.
//
// LDS-staged lines (measurement only, built with OMNI3D_LAYOUT_STUDY): rows
// (sources) are grouped kStageGroup chunks of 64 at a time.  For every
// (group, direction) the distinct 16-source lines (128 bytes of p) that the
// group's entries of that direction gather are listed once; consecutive
// directions are packed into stages of at most kStageCap lines, which the
// iteration copies into LDS with coalesced 16-byte loads before the rows
// gather from LDS.  Each entry is stored as its 16-bit position in its
// stage's LDS copy.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "launch.hpp"
#include "stage.hpp"

namespace omni3d {

namespace {

__host__ __device__ inline int bits_for(unsigned long long n) {
  int b = 0;
  while (b < 64 && (1ull << b) < n) b++;
  return b;
}

#ifdef OMNI3D_LAYOUT_STUDY
// One wavefront per chunk: key (group, direction, line) and the entry
// index of every packed entry.
__global__ void __launch_bounds__(64)
k_stage_keys(int nch, int lb, int db, const unsigned long long* __restrict__ moff,
             const unsigned long long* __restrict__ masks, const unsigned short* __restrict__ sdir,
             const unsigned int* __restrict__ cell_eb, const int* __restrict__ ent,
             unsigned long long* __restrict__ keys, unsigned int* __restrict__ vals) {
  const int chunk = blockIdx.x, lane = threadIdx.x;
  if (chunk >= nch) return;
  const unsigned long long g = (unsigned long long)(chunk / kStageGroup);
  const unsigned long long below = (1ull << lane) - 1ull;
  for (unsigned long long s = moff[chunk]; s < moff[chunk + 1]; s++) {
    const unsigned long long m = masks[s];
    if (!((m >> lane) & 1ull)) continue;
    const unsigned int e = cell_eb[s] + (unsigned int)__popcll(m & below);
    const unsigned long long line = (unsigned long long)(unsigned int)ent[e] >> 4;
    keys[e] = (g << (db + lb)) | ((unsigned long long)sdir[s] << lb) | line;
    vals[e] = e;
  }
}

__global__ void k_stage_flags(int64_t n, const unsigned long long* __restrict__ k,
                              unsigned int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

// Distinct lines (in (group, direction, line) order) and their count per
// (group, direction).
__global__ void k_stage_unique(int64_t n, int lb, int db, int D, const unsigned long long* __restrict__ k,
                               const unsigned int* __restrict__ incl, unsigned int* __restrict__ ulines,
                               unsigned int* __restrict__ gdcnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i != 0 && k[i] == k[i - 1]) return;
  const unsigned long long key = k[i];
  ulines[incl[i] - 1u] = (unsigned int)(key & ((1ull << lb) - 1ull));
  const unsigned long long d = (key >> lb) & ((1ull << db) - 1ull);
  const unsigned long long g = key >> (lb + db);
  atomicAdd(&gdcnt[g * (unsigned long long)D + d], 1u);
}

// One lane per group: pack consecutive directions into stages of at most
// kStageCap lines.  sbase[g*D+d] = first distinct line of d's stage.
__global__ void k_stage_pack(int ngroups, int D, const unsigned int* __restrict__ gdptr,
                             unsigned int* __restrict__ nst, unsigned int* __restrict__ st_lo,
                             unsigned int* __restrict__ st_hi, unsigned short* __restrict__ st_dend,
                             unsigned int* __restrict__ sbase, unsigned int* __restrict__ overflow) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const size_t b = (size_t)g * D;
  unsigned int t = 0, lo = gdptr[b], cur = 0;
  for (int d = 0; d < D; d++) {
    const unsigned int L = gdptr[b + d + 1] - gdptr[b + d];
    if (L > (unsigned int)kStageCap) atomicOr(overflow, 1u);
    if (cur > 0 && cur + L > (unsigned int)kStageCap) {
      st_lo[b + t] = lo;
      st_hi[b + t] = gdptr[b + d];
      st_dend[b + t] = (unsigned short)d;
      t++;
      lo = gdptr[b + d];
      cur = 0;
    }
    sbase[b + d] = lo;
    cur += L;
  }
  st_lo[b + t] = lo;
  st_hi[b + t] = gdptr[b + D];
  st_dend[b + t] = (unsigned short)D;
  nst[g] = t + 1;
}

// One lane per chunk: the end slot of each of its group's stages.
__global__ void k_stage_slot_ends(int nch, int D, const unsigned long long* __restrict__ moff,
                                  const unsigned short* __restrict__ sdir,
                                  const unsigned int* __restrict__ nst,
                                  const unsigned short* __restrict__ st_dend,
                                  unsigned int* __restrict__ cse) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nch) return;
  const int g = c / kStageGroup, w = c % kStageGroup;
  const size_t b = (size_t)g * D;
  const unsigned int ns = nst[g];
  unsigned long long s = moff[c];
  const unsigned long long e = moff[c + 1];
  for (unsigned int t = 0; t < ns; t++) {
    const unsigned int dend = st_dend[b + t];
    while (s < e && sdir[s] < dend) s++;
    cse[(b + t) * kStageGroup + w] = (unsigned int)s;
  }
}

// LDS position of every entry: (distinct-line index - stage base) * 16 +
// position of the source in its line.
__global__ void k_stage_offsets(int64_t n, int lb, int db, int D,
                                const unsigned long long* __restrict__ k,
                                const unsigned int* __restrict__ v,
                                const unsigned int* __restrict__ incl,
                                const unsigned int* __restrict__ sbase,
                                const int* __restrict__ ent, unsigned short* __restrict__ off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long key = k[i];
  const unsigned long long d = (key >> lb) & ((1ull << db) - 1ull);
  const unsigned long long g = key >> (lb + db);
  const unsigned int u = incl[i] - 1u;
  const unsigned int e = v[i];
  const unsigned int base = sbase[g * (unsigned long long)D + d];
  off[e] = (unsigned short)((u - base) * 16u + ((unsigned int)ent[e] & 15u));
}

#endif  // OMNI3D_LAYOUT_STUDY

unsigned int grid1(int64_t n) { return (unsigned int)((n + 255) / 256); }

// One lane per row: key (group, source, row in group) of each CSR entry
// (CSR keys are (direction << 32) | source); a group is 2^gshift rows.
__global__ void k_dense_keys(int nS, int gshift, int rb, const unsigned long long* __restrict__ rowptr,
                             const unsigned long long* __restrict__ ckey,
                             const int* __restrict__ rgroup, const int* __restrict__ grow,
                             unsigned long long* __restrict__ keys) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nS) return;
  const int g = rgroup ? rgroup[r] : r >> gshift;
  const int rin = rgroup ? r - grow[g] : r & ((1 << gshift) - 1);
  const unsigned long long hi = ((unsigned long long)g << 32) | (unsigned long long)rin;
  const unsigned long long smask = (1ull << (32 - rb)) - 1ull;
  for (unsigned long long j = rowptr[r]; j < rowptr[r + 1]; j++)
    keys[j] = hi | ((ckey[j] & smask) << rb);
}

__global__ void k_dense_words(int64_t n, int rb, const unsigned long long* __restrict__ keys,
                              unsigned int* __restrict__ dent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long k = keys[i];
  const unsigned int row = (unsigned int)(k & ((1ull << rb) - 1ull));
  const unsigned int src = (unsigned int)((k >> rb) & ((1ull << (32 - rb)) - 1ull));
  dent[i] = (row << (32 - rb)) | src;
}

// Merge of equal (group, source, row) keys into one word with a
// multiplicity of 1..4: flag the first key of every run of up to 4.
__global__ void k_merge_flags(int64_t n, const unsigned long long* __restrict__ k,
                              unsigned int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c = 0;
  while (i - c - 1 >= 0 && k[i - c - 1] == k[i]) c++;
  flag[i] = (c & 3) == 0 ? 1u : 0u;
}

// word (row << (32 - rb)) | (multiplicity - 1) << (30 - rb) | source at the
// run's slot
__global__ void k_merge_write(int64_t n, int rb, const unsigned long long* __restrict__ k,
                              const unsigned int* __restrict__ flag,
                              const unsigned int* __restrict__ excl, unsigned int* __restrict__ dent) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  unsigned int m = 1;
  while (m < 4 && i + m < n && k[i + m] == k[i]) m++;
  const unsigned long long key = k[i];
  const unsigned int row = (unsigned int)(key & ((1ull << rb) - 1ull));
  const unsigned int src = (unsigned int)((key >> rb) & ((1ull << (30 - rb)) - 1ull));
  dent[excl[i]] = (row << (32 - rb)) | ((m - 1u) << (30 - rb)) | src;
}

// first word of every group: the merged index of the group's first entry
__global__ void k_merge_gptr(int ng, int rows, int nS, int64_t n,
                             const unsigned long long* __restrict__ rowptr,
                             const unsigned int* __restrict__ excl, unsigned int nwords,
                             const int* __restrict__ grow, unsigned int* __restrict__ wptr) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g > ng) return;
  const long long r = grow ? (long long)grow[g] : (long long)g * rows;
  const unsigned long long pos = rowptr[r < nS ? r : nS];
  wptr[g] = pos < (unsigned long long)n ? excl[pos] : nwords;
}

}  // namespace

size_t dense_tmp_bytes(int64_t nnz) {
  size_t a = 0, b = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, a, (const unsigned long long*)nullptr,
                                          (unsigned long long*)nullptr, (int)nnz, 0, 64);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const unsigned int*)nullptr,
                                         (unsigned int*)nullptr, (int)nnz);
  return std::max(a, b) + 256;
}

hipError_t merge_dense_entries(int nS, int64_t nnz, int gchunks, const unsigned long long* rowptr,
                               const unsigned long long* keys, unsigned int* flag,
                               unsigned int* excl, void* tmp, size_t tmp_bytes, unsigned int* dent,
                               unsigned int* wptr, unsigned int* nwords, hipStream_t st,
                               const int* grow, int ngroups) {
  const int rb = dense_row_bits(gchunks);
  hipLaunchKernelGGL(k_merge_flags, dim3(grid1(nnz)), dim3(256), 0, st, nnz, keys, flag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = tmp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, flag, excl, (int)nnz, st);
  if (e != hipSuccess) return e;
  unsigned int last[2] = {0, 0};
  e = hipMemcpyAsync(&last[0], excl + (nnz - 1), 4, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(&last[1], flag + (nnz - 1), 4, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  *nwords = last[0] + last[1];
  hipLaunchKernelGGL(k_merge_write, dim3(grid1(nnz)), dim3(256), 0, st, nnz, rb, keys, flag, excl,
                     dent);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int rows = 64 * gchunks;
  const int ng = grow ? ngroups : (nS + rows - 1) / rows;
  hipLaunchKernelGGL(k_merge_gptr, dim3(grid1(ng + 1)), dim3(256), 0, st, ng, rows, nS, nnz, rowptr,
                     excl, *nwords, grow, wptr);
  return hipGetLastError();
}

hipError_t build_dense_entries(int nS, int64_t nnz, int gchunks, const unsigned long long* rowptr,
                               const unsigned long long* csr_key, unsigned long long* keys_a,
                               unsigned long long* keys_b, void* tmp, size_t tmp_bytes,
                               unsigned int* dent, hipStream_t st, const int* rgroup,
                               const int* grow, int ngroups) {
  const int gshift = 6 + bits_for((unsigned long long)gchunks);
  const int rb = dense_row_bits(gchunks);
  hipLaunchKernelGGL(k_dense_keys, dim3(grid1(nS)), dim3(256), 0, st, nS, gshift, rb, rowptr,
                     csr_key, rgroup, grow, keys_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int ng = rgroup ? ngroups : (nS + (1 << gshift) - 1) >> gshift;
  const int end_bit = 32 + bits_for((unsigned long long)ng);
  size_t bytes = tmp_bytes;
  e = hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys_a, keys_b, (int)nnz, 0, end_bit, st);
  if (e != hipSuccess) return e;
  if (gchunks > 1) return hipSuccess;  // merged words: merge_dense_entries
  hipLaunchKernelGGL(k_dense_words, dim3(grid1(nnz)), dim3(256), 0, st, nnz, rb, keys_b, dent);
  return hipGetLastError();
}

#ifdef OMNI3D_LAYOUT_STUDY
int stage_key_bits(int nS, int D) {
  const int nch = (nS + 63) / 64;
  const int ngroups = (nch + kStageGroup - 1) / kStageGroup;
  return bits_for((unsigned long long)ngroups) + bits_for((unsigned long long)D) +
         bits_for((unsigned long long)((nS + 15) / 16));
}

size_t stage_tmp_bytes(int64_t nnz, int64_t ngd) {
  size_t a = 0, b = 0, c = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr,
                                           (const unsigned int*)nullptr, (unsigned int*)nullptr,
                                           (int)nnz, 0, 64);
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, b, (const unsigned int*)nullptr,
                                         (unsigned int*)nullptr, (int)nnz);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const unsigned int*)nullptr,
                                         (unsigned int*)nullptr, (int)(ngd + 1));
  return std::max(a, std::max(b, c)) + 256;
}

hipError_t stage_keys_sorted(const StageBuild& sb, hipStream_t st) {
  const int lb = bits_for((unsigned long long)((sb.nS + 15) / 16));
  const int db = bits_for((unsigned long long)sb.D);
  const int end_bit = stage_key_bits(sb.nS, sb.D);
  hipLaunchKernelGGL(k_stage_keys, dim3(sb.nch), dim3(64), 0, st, sb.nch, lb, db, sb.moff, sb.masks,
                     sb.sdir, sb.cell_eb, sb.ent, sb.keys_a, sb.vals_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = sb.tmp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(sb.tmp, bytes, sb.keys_a, sb.keys_b, sb.vals_a, sb.vals_b,
                                         (int)sb.nnz, 0, end_bit, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stage_flags, dim3(grid1(sb.nnz)), dim3(256), 0, st, sb.nnz, sb.keys_b,
                     sb.flag);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  bytes = sb.tmp_bytes;
  return hipcub::DeviceScan::InclusiveSum(sb.tmp, bytes, sb.flag, sb.incl, (int)sb.nnz, st);
}

hipError_t stage_finish(const StageBuild& sb, int ngroups, unsigned int* ulines, unsigned int* gdcnt,
                        unsigned int* gdptr, unsigned int* nst, unsigned int* st_lo,
                        unsigned int* st_hi, unsigned short* st_dend, unsigned int* sbase,
                        unsigned int* cse, unsigned short* off, unsigned int* overflow,
                        hipStream_t st) {
  const int lb = bits_for((unsigned long long)((sb.nS + 15) / 16));
  const int db = bits_for((unsigned long long)sb.D);
  const int64_t ngd = (int64_t)ngroups * sb.D;
  hipError_t e = hipMemsetAsync(gdcnt, 0, sizeof(unsigned int) * (ngd + 1), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stage_unique, dim3(grid1(sb.nnz)), dim3(256), 0, st, sb.nnz, lb, db, sb.D,
                     sb.keys_b, sb.incl, ulines, gdcnt);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = sb.tmp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(sb.tmp, bytes, gdcnt, gdptr, (int)(ngd + 1), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stage_pack, dim3(grid1(ngroups)), dim3(256), 0, st, ngroups, sb.D, gdptr,
                     nst, st_lo, st_hi, st_dend, sbase, overflow);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stage_slot_ends, dim3(grid1(sb.nch)), dim3(256), 0, st, sb.nch, sb.D,
                     sb.moff, sb.sdir, nst, st_dend, cse);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stage_offsets, dim3(grid1(sb.nnz)), dim3(256), 0, st, sb.nnz, lb, db, sb.D,
                     sb.keys_b, sb.vals_b, sb.incl, sbase, sb.ent, off);
  return hipGetLastError();
}

#endif  // OMNI3D_LAYOUT_STUDY

}  // namespace omni3d
