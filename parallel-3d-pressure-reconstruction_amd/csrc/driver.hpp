// Internal driver functions shared by the C-ABI files (api.cpp defines them).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "host.hpp"
#include "stage.hpp"

namespace omni3d {

extern thread_local std::string g_last_error;

// Records msg on the context (if any) and as the global last error.
int fail(omni3d_ctx* ctx, int code, const std::string& msg);
int64_t voxels(const omni3d_params& p);
size_t elem_size(int dtype);
bool params_ok(const omni3d_params& p, std::string* why);
// Geometry plan of the context's grid, direction set and line spacing
// (rebuilt only when one of them changed).
int build_plan(omni3d_ctx* ctx);
// The context's mask epoch with a pending device comparison of a newly set
// mask resolved (a flag read): what caches keyed on the mask compare.
int mask_epoch_now(omni3d_ctx* ctx, uint64_t* epoch);
WalkParams walk_params(const omni3d_params& P);
// traversal == OMNI3D_TRAVERSAL_TRILINEAR
bool trilinear(const omni3d_params& P);
// whether the plain method's traversals run one lane per ray
bool ray_engine(const omni3d_ctx* ctx);
// bound on a visit's path for the fixed-point scale (see launch_compute_scale)
double path_bound(const omni3d_params& P, const Plan& pl);
// Row groups of the grouped iteration with about equal entries (rowptr:
// nS + 1 CSR offsets on the host): at most maxrows rows each, grow the
// first row of every group (nS last), rgroup[nS] the group of every row.
void balance_groups(const std::vector<unsigned long long>& rowptr, int maxrows, int ncu,
                    std::vector<int>* grow, std::vector<int>* rgroup);
// Reference-schedule iteration; the result of iteration k* goes to
// pdst/wdst (device, dtype), seed: -1 no initial pressure, 0 p only,
// 1 p and the seed term.
template <typename T>
int run_naive(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype);
// A whole reconstruction queued on the context's stream without host
// synchronisation (capturable once the buffers are sized); pdst / wdst device.
int enqueue_solve(omni3d_ctx* ctx, int seed, void* pdst, void* wdst, int dtype);
// multi-snapshot source iteration (snapshot batches)
bool affine_fusable(const omni3d_ctx* ctx);
IterDesc affine_iter_desc(omni3d_ctx* ctx, int first);
// Releases the context's cached graph of the iteration launches.
void drop_iteration_graph(omni3d_ctx* ctx);
// Phase timing event i of the context (skipped during graph capture).
void record_event(omni3d_ctx* ctx, int i);

}  // namespace omni3d

#define OMNI_TRY_HIP(ctx, expr)                                                   \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      const std::string _m = std::string(#expr) + ": " + hipGetErrorString(_e);  \
      return omni3d::fail((ctx), OMNI3D_ERR_HIP, _m);                             \
    }                                                                             \
  } while (0)
