// Host-side internals of libomni3d: the context, the geometry plan and the
// device buffers they own.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/omni3d.h"
#include "geometry.hpp"
#include "kernels.hpp"

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
  DevBuf rowptr;   // uint64[nS+1]
  DevBuf col;      // int32[nnz] source id of each visit landing on a source
};

}  // namespace omni3d

struct omni3d_ctx {
  omni3d_params prm{};
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<float> k;  // 3*D, row-major (k1,k2,k3)
  bool dirs_dirty = true;
  omni3d::Plan plan;
  // field in the compute precision
  omni3d::DevBuf ax, ay, az, mask;
  bool have_field = false;
  bool have_mask = false;
  // per-reconstruction work buffers
  omni3d::DevBuf scal;      // DevScalars
  omni3d::DevBuf pinit;     // double[V]
  omni3d::DevBuf bs;        // int64[nS]
  omni3d::DevBuf ps0, ps1;  // double[nS]
  omni3d::DevBuf partials;  // double[2*grid]
  omni3d::DevBuf sum;       // int64[V]
  omni3d::DevBuf pout, wout;  // staging for host outputs
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};
