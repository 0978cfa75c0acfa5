// Tile engine data structures (see tiles.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace omni3d {

constexpr int kTileBits = 4;
constexpr int kTile = 1 << kTileBits;  // 16^3 voxels per tile
constexpr int kTileMask = kTile - 1;
constexpr int kTileVox = kTile * kTile * kTile;

enum { kLocalSum = 0, kSourceSums = 1, kFull = 2 };

struct TileGrid {
  int ntx, nty, ntz;
};

// One segment: the visits of one ray inside one tile.
struct SegRec {
  uint32_t ray;  // index into the ray list
  uint32_t g;    // segment id in ray-major order (index of its carry)
  uint32_t s0n;  // first visit index (low 16 bits) | number of visits (high 16 bits)
  uint32_t vox;  // packed voxel of the first visit
};

hipError_t launch_tile_count(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                             TileGrid tg, unsigned int* nsteps, unsigned int* nseg,
                             unsigned int* tile_cnt, hipStream_t st);
hipError_t launch_tile_fill(const RayRec* rays, int64_t nrays, const DirConst* dirs, GridDesc g,
                            TileGrid tg, const unsigned long long* word_off,
                            const unsigned long long* seg_base, const unsigned long long* tile_ptr,
                            unsigned int* cursor, uint32_t* words, SegRec* segs, hipStream_t st);
template <typename T>
hipError_t launch_tile_pass(int mode, int ntiles, const int* tile_list,
                            const unsigned long long* tile_ptr, const SegRec* segs,
                            const RayRec* rays, const unsigned long long* word_off,
                            const uint32_t* words, GridDesc g, TileGrid tg, const WalkParams& wp,
                            const void* field, const int* sid, const double* p0,
                            const double* p1, const DevScalars* sc, double* segval,
                            long long* out, hipStream_t st);
hipError_t launch_carry_scan(int64_t nrays, const unsigned long long* seg_base, double* segval,
                             hipStream_t st);
// exclusive scan of n uint32 counts into n+1 uint64 offsets (multi-block);
// `words` != 0 scans ceil(count/16) instead of count
hipError_t launch_scan_u32(const unsigned int* cnt, int64_t n, unsigned long long* out,
                           int words, void* tmp, size_t tmp_bytes, hipStream_t st);
size_t scan_u32_tmp_bytes(int64_t n);

}  // namespace omni3d
