/*
 * omni3d.h -- C-ABI of the MI355X (gfx950) omni-directional 3D pressure
 * reconstruction library libomni3d.so.
 *
 * Plain C linkage: no C++ types, no exceptions, plain pointers and sizes.
 * Every entry point returns an int status (OMNI3D_OK = 0); the message of the
 * last failure on a context is omni3d_last_error(ctx) (or
 * omni3d_last_global_error() for calls without a context).
 *
 * Reference interfaces replaced (paths relative to the reference root,
 * Omni3D_8.0/Omni3D_8.0/):
 *   kernel.cu:8467-8983  main()                      -> omni3d_run_parameter_file
 *   kernel.cu:8622-8648  direction-file reader       -> omni3d_load_directions_file
 *   kernel.cu:8805-8861  iteration driver            -> omni3d_reconstruct
 *   kernel.cu:6151       omni3dparallellinesESInner  -> omni3d_es_inner, and the
 *                        same name as a host function (call site minus <<<...>>>)
 *   kernel.cu:6281       ...ESInnerWeighted          -> omni3d_es_inner_weighted (+ same name)
 *   kernel.cu:6414       ...ESInnerSelect            -> omni3d_es_inner_select (+ same name)
 *   kernel.cu:6481       ...ESInnerSelectFixedBC     -> omni3d_es_inner_select_fixedbc (+ same name)
 *   kernel.cu:6347       ...ESInnerWeightedMiniCurl  -> omni3d_es_inner_weighted_minicurl (+ same name)
 *   kernel.cu:6746       devidecountInner            -> omni3d_devide_count_inner (+ same name)
 *   kernel.cu:8806-8811  Bernoulli seed of the j = 0 face -> omni3d_set_velocity
 *   kernel.cu:8352       calCurlofMaterialAccCPU     -> omni3d_curl
 *   kernel.cu:6863       calCurlofMaterialAcc (GPU)  -> omni3d_curl, mask NULL
 *   kernel.cu:8380       thredholdHistMaterialAccCPU -> thredholdHistMaterialAccCPU
 *                        (same signature) / omni3d_threshold_hist
 *   sor3D.h:5            sor3D                       -> sor3D (same signature)
 *   calRHS.h:4           calRHS                      -> calRHS (same signature; the
 *                        result is written to the caller's RHS buffer, which
 *                        the reference loses, calRHS.cpp:10)
 *   initialIntegration.h:3 initialIntegration        -> initialIntegration (same signature)
 *   kernel.cu:3414,5572  EqualSpacing + BCiteration -> omni3d_boundary_pairs
 *                        (the README's boundary-pair pipeline, commented out
 *                        in main, kernel.cu:8830-8836)
 * Additions without a reference counterpart: the adjoint pair
 * (omni3d_forward / omni3d_backward), the GPU red-black SOR
 * (omni3d_sor_solve) and the snapshot batch (omni3d_batch_*).
 *
 * Threading: a context is not thread-safe; use one per host thread.  All GPU
 * work of a context is ordered on the stream given to it (NULL = the null
 * stream).  Layout: x fastest, index = i + j*nx + k*nx*ny.
 */
#ifndef OMNI3D_H
#define OMNI3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMNI3D_OK 0
#define OMNI3D_ERR_ARG 1       /* invalid argument */
#define OMNI3D_ERR_STATE 2     /* call out of order (no field / no directions) */
#define OMNI3D_ERR_HIP 3       /* HIP runtime error (message has the HIP string) */
#define OMNI3D_ERR_IO 4        /* file missing or malformed */
#define OMNI3D_ERR_NONFINITE 5 /* NaN / Inf in the input field */
#define OMNI3D_ERR_NODEVICE 6  /* no GPU visible */

/* Integration method (which reference integrator a pass restates). */
#define OMNI3D_METHOD_PLAIN 0      /* bodyIntegralInner, kernel.cu:1743 */
#define OMNI3D_METHOD_PLAIN_MASK 1 /* Select walk with mask, unit weights, threshold +inf */
#define OMNI3D_METHOD_SELECT 2     /* bodyIntegralInnerSelect, kernel.cu:2280 */
#define OMNI3D_METHOD_WEIGHTED 3   /* bodyIntegralInnerWeighted, kernel.cu:1993 */
/* bodyIntegralInnerSelectFixedBC, kernel.cu:2538: Select with its own step
 * rule (x / y ties to y, no step along or onto the j = 0 plane, which keeps
 * its prescribed pressure); one lane per ray, reference schedule */
#define OMNI3D_METHOD_SELECT_FIXEDBC 4
/* bodyIntegralInnerMiniCurl, kernel.cu:2797 (the pass of
 * ESInnerWeightedMiniCurl): plain sums and counts, no entry self-add, along
 * the path through the smallest curl toward the exit; one lane per ray,
 * reference schedule */
#define OMNI3D_METHOD_MINICURL 5
/* bodyIntegralInnerWeightedFixedBC, kernel.cu:2136-2279 (no launcher calls
 * it): the Weighted sums and weights along the plain staircase, except that
 * a step onto the j = 0 plane ends the ray; one lane per ray, reference
 * schedule */
#define OMNI3D_METHOD_WEIGHTED_FIXEDBC 6
/* bodyIntegralInnerWeightedMiniCurl, kernel.cu:2937-3085 (no launcher calls
 * it): the MiniCurl path with curl weights counttmp / curltmp, curltmp the
 * running sum of curl (not / 10^4) starting at the entry voxel, which adds
 * its own p with weight 1 / curl[entry]; a zero curl sum gives infinite
 * weights, as in the reference; one lane per ray, reference schedule */
#define OMNI3D_METHOD_WEIGHTED_MINICURL 7

/* Arithmetic of the ray sums. */
#define OMNI3D_FP64 0        /* gradient and ray sums in FP64 */
/* Gradient and trapezoid terms in FP32, ray sums compensated.  The tile
 * engine (the default traversals) adds the FP32 terms into FP64 accumulators
 * and stores a ray's partial sums (per segment, and the prefix sums carried
 * between tiles) rounded once to FP32; the one-lane-per-ray walks (reference
 * schedule, trilinear traversal, the cross-checks) add them into
 * Kahan-compensated FP32 sums.  Averaging, the source iteration's row sums,
 * the boundary update and the residual stay FP64. */
#define OMNI3D_FP32_KAHAN 1
#define OMNI3D_FP32 OMNI3D_FP32_KAHAN

/* How a ray's line integral is formed. */
#define OMNI3D_TRAVERSAL_STAIRCASE 0 /* the voxel staircase with one trapezoid per
                                        axis step (kernel.cu:1764-1840) */
/* The staircase's visited voxels (same counts, sources and averaging), with
 * the integral taken along the straight ray: one trapezoid of grad p . k per
 * step between the projections of consecutive voxel centres, the gradient
 * sampled trilinearly there, plus a trapezoid leg from each end voxel to its
 * projection.  Exact for a linear gradient.  No reference counterpart (the
 * accuracy mode beside the staircase); plain method, affine solver. */
#define OMNI3D_TRAVERSAL_TRILINEAR 1

/* Iteration schedule. */
#define OMNI3D_SOLVER_AFFINE 0 /* iterate only the voxels the rays read, then one full pass */
#define OMNI3D_SOLVER_NAIVE 1  /* a full pass per iteration (kernel.cu:8840-8861) */

/* Buffer dtypes and memory kinds. */
#define OMNI3D_F32 0
#define OMNI3D_F64 1
#define OMNI3D_MEM_HOST 0
#define OMNI3D_MEM_DEVICE 1

typedef struct omni3d_params {
  int nx, ny, nz;          /* grid (Parameter_Omni3D.dat lines 1-3, after the ROI crop) */
  double dx, dy, dz;       /* spacing (lines 4-6) */
  double density;          /* rho (line 7) */
  double line_spacing;     /* parallel-line spacing in grid units (line 9) */
  int method;              /* OMNI3D_METHOD_* */
  double threshold;        /* curl threshold of the Select method (line 15) */
  int faithful_select;     /* 1 = keep the Select per-axis asymmetries (kernel.cu:2394 vs 2442) */
  int precision;           /* OMNI3D_FP64 / OMNI3D_FP32_KAHAN */
  int solver;              /* OMNI3D_SOLVER_AFFINE / OMNI3D_SOLVER_NAIVE */
  int max_iterations;      /* NoItr (line 14) */
  double tolerance;        /* stop when ||dp_S|| / ||p_S|| < tolerance over the voxels the
                              rays read; 0 = always run max_iterations (reference) */
  int seed_pn_with_p;      /* 1 = the first pass's sums start from p (kernel.cu:8813) */
  int traversal;           /* OMNI3D_TRAVERSAL_* (default staircase) */
} omni3d_params;

typedef struct omni3d_report {
  int iterations;          /* passes whose result the output holds */
  double rel_change;       /* last relative change over the source voxels */
  int64_t rays;            /* accepted rays (all directions) */
  int64_t steps;           /* staircase steps per full pass */
  int64_t sources;         /* voxels whose p the rays read */
  int status;              /* OMNI3D_OK or OMNI3D_ERR_NONFINITE */
  double ms_plan;          /* GPU time of the geometry plan build (0 if reused) */
  double ms_solve;         /* GPU time of the reconstruction (everything after the plan) */
  double ms_pass_a;        /* first traversal: ray sums landing on source voxels */
  double ms_iterate;       /* source-voxel iterations (or the naive passes) */
  double ms_pass_b;        /* final traversal into every voxel */
  double ms_finalize;      /* averaging into the output */
  int64_t source_visits;   /* visits landing on source voxels (iteration matrix entries) */
  int64_t segments;        /* ray segments of the tile engine (one per ray and tile) */
  int64_t sell_cells;      /* stored cells of the iteration matrix incl. padding */
  double ms_segment_sums;  /* the segment-sum traversal of pass A alone (tile engine) */
  int64_t matrix_words;    /* 4-byte words of the iteration matrix one iteration reads
                              (equal entries of a row merged) */
} omni3d_report;

typedef struct omni3d_ctx omni3d_ctx;

/* Defaults: rho 1, line spacing 1, plain FP64 affine, 10 iterations (the
 * shipped Parameter_Omni3D.dat line 14), tolerance 0, seed_pn_with_p 1,
 * staircase traversal. */
void omni3d_default_params(omni3d_params* p);

int omni3d_create(const omni3d_params* params, omni3d_ctx** out);
void omni3d_destroy(omni3d_ctx* ctx);
const char* omni3d_last_error(const omni3d_ctx* ctx);
const char* omni3d_last_global_error(void);
int omni3d_set_stream(omni3d_ctx* ctx, void* hip_stream);

/* Tile engine shape of the plain reconstruction: tile edge 2^tile_bits voxels
 * (3: 8^3, 4: 16^3 -- a segment then holds at most 22 / 46 visits) and the
 * workgroup size of its passes (256, 512, 1024).  Default 4 / 1024; the
 * batch autotuner sets them per grid.  No reference counterpart. */
int omni3d_set_tile_shape(omni3d_ctx* ctx, int tile_bits, int threads);

/* Direction set (kernel.cu:8622-8672: float k1,k2,k3[NoAngles]); vectors are
 * normalised in FP64 on the host. */
int omni3d_set_directions(omni3d_ctx* ctx, const float* k1, const float* k2, const float* k3,
                          int count);
/* Reads the first `count` rows "k1 k2 k3" of a GridIcosahedronDivideN.dat file
 * (count <= 0: the count line at the end of the file). */
int omni3d_load_directions_file(omni3d_ctx* ctx, const char* path, int count);
/* Pole-aligned icosahedral directions, `divide` = N of GridIcosahedronDivideN
 * (N-1 midpoint subdivisions): 10*4^(N-1)+2 unit vectors in exact antipodal
 * pairs.  Writes min(count, capacity) rows; returns the count (<0 on error). */
int omni3d_generate_directions(int divide, float* k1, float* k2, float* k3, int capacity);

/* Input field: material acceleration Du/Dt, Dv/Dt, Dw/Dt (grad p = -rho a,
 * kernel.cu:1816), dtype OMNI3D_F32/F64, host or device memory.  mask may
 * be NULL (all valid); for SELECT/WEIGHTED the curl*mask field is derived
 * from it (kernel.cu:8824, 8352-8379).  The context keeps its own copy in
 * the precision it computes in. */
int omni3d_set_field(omni3d_ctx* ctx, const void* dudt, const void* dvdt, const void* dwdt,
                     const void* mask, int dtype, int mem_kind);

/* Velocity of the snapshot for the Bernoulli seed (kernel.cu:8806-8811): every
 * later omni3d_reconstruct starts from p with its j = 0 face replaced by
 * pref - rho/2 (u^2 + v^2 + w^2) (from zeros when p is NULL), as the
 * reference's driver does before its first pass.  u, v, w: X*Y*Z values
 * (only the j = 0 face is read), dtype / mem_kind as for the field; the
 * context keeps the face values.  u == NULL removes the seed. */
int omni3d_set_velocity(omni3d_ctx* ctx, const void* u, const void* v, const void* w,
                        double pref, int dtype, int mem_kind);

/* Iterate to convergence on the GPU (kernel.cu:8805-8861).  p: in = initial
 * pressure (NULL = zeros), out = result, dtype/mem_kind as given.
 * weight_out (may be NULL): the per-voxel weight sum of the last pass
 * ("Count" column, kernel.cu:8929), same dtype/mem_kind.  rep may be NULL.
 * Returns after the work is queued and, for host buffers, copied back. */
int omni3d_reconstruct(omni3d_ctx* ctx, void* p, void* weight_out, int dtype, int mem_kind,
                       omni3d_report* rep);

/* Kernel-level drop-ins: the reference kernels' argument lists in their order
 * (kernel.cu:6151, 6281, 6414, 6481, 6347), device pointers, float, plus the
 * HIP stream (NULL = the null stream) at the end.  One pass over every line x
 * direction pair with deterministic accumulation: pn += sums over visits,
 * pcountinner += weights.  index (the boundary numbering of calIndexGPU) is
 * accepted and not read; NULL is allowed.
 *
 * The work is queued on the stream and the call returns without waiting for
 * it.  Each host thread keeps a hidden context: the geometry plan is rebuilt
 * only when the grid, line spacing or directions change, and the staged field
 * with the Select / Weighted rule replay only when the field, curl, spacing,
 * density, method or threshold change.  Changes are found by a fingerprint of
 * the directions, field and curl computed on the GPU (one 16-byte read-back
 * per call, which also reports a non-finite input or p as
 * OMNI3D_ERR_NONFINITE before anything is queued), so buffers may be reused
 * or rewritten between calls.  omni3d_dropin_release frees the thread's
 * hidden context. */
int omni3d_es_inner(long X, long Y, long Z, int NoAngles, float linespacing, const float* k1_d,
                    const float* k2_d, const float* k3_d, const long* index, float deltx,
                    float delty, float deltz, float density, const float* DuDt,
                    const float* DvDt, const float* DwDt, const float* p, float* pn,
                    float* pcountinner, void* hip_stream);
int omni3d_es_inner_weighted(long X, long Y, long Z, int NoAngles, float linespacing,
                             const float* k1_d, const float* k2_d, const float* k3_d,
                             const long* index, float deltx, float delty, float deltz,
                             float density, const float* DuDt, const float* DvDt,
                             const float* DwDt, const float* curl, const float* p, float* pn,
                             float* pcountinner, void* hip_stream);
int omni3d_es_inner_select(long X, long Y, long Z, int NoAngles, float linespacing,
                           const float* k1_d, const float* k2_d, const float* k3_d,
                           const long* index, float deltx, float delty, float deltz,
                           float density, const float* DuDt, const float* DvDt,
                           const float* DwDt, const float* curl, const float* p, float* pn,
                           float* pcountinner, float threshold, void* hip_stream);
/* omni3dparallellinesESInnerSelectFixedBC, kernel.cu:6481-6549 (commented out
 * in the reference's main): the Select pass with the FixedBC step rule. */
int omni3d_es_inner_select_fixedbc(long X, long Y, long Z, int NoAngles, float linespacing,
                                   const float* k1_d, const float* k2_d, const float* k3_d,
                                   const long* index, float deltx, float delty, float deltz,
                                   float density, const float* DuDt, const float* DvDt,
                                   const float* DwDt, const float* curl, const float* p,
                                   float* pn, float* pcountinner, float threshold,
                                   void* hip_stream);
/* omni3dparallellinesESInnerWeightedMiniCurl, kernel.cu:6347-6413 (commented
 * out in the reference's main; it calls bodyIntegralInnerMiniCurl): one
 * MiniCurl pass, unit counts into pcountinner. */
int omni3d_es_inner_weighted_minicurl(long X, long Y, long Z, int NoAngles, float linespacing,
                                      const float* k1_d, const float* k2_d, const float* k3_d,
                                      const long* index, float deltx, float delty, float deltz,
                                      float density, const float* DuDt, const float* DvDt,
                                      const float* DwDt, const float* curl, const float* p,
                                      float* pn, float* pcountinner, void* hip_stream);
/* devidecountInner, kernel.cu:6746-6759: where pcountinner > 0,
 * p = pn / pcountinner and pn = 0. */
int omni3d_devide_count_inner(long X, long Y, long Z, float* p, float* pn,
                              const float* pcountinner, void* hip_stream);
/* Frees the calling thread's hidden drop-in context (plan, staged field). */
void omni3d_dropin_release(void);

/* The same six under the reference kernels' own names and argument lists
 * (kernel.cu:6151, 6281, 6414, 6481, 6347, 6746): a reference call site
 * binds unchanged once its <<<grid, block>>> is dropped.  Null stream; void,
 * as a kernel launch is: failures only through omni3d_last_global_error(). */
void omni3dparallellinesESInner(long Xsize, long Ysize, long Zsize, int NoAngles,
                                float linespacing, float* k1_d, float* k2_d, float* k3_d,
                                long* index, float deltx, float delty, float deltz,
                                float density, float* DuDt, float* DvDt, float* DwDt, float* p,
                                float* pn, float* pcountinner);
void omni3dparallellinesESInnerWeighted(long Xsize, long Ysize, long Zsize, int NoAngles,
                                        float linespacing, float* k1_d, float* k2_d,
                                        float* k3_d, long* index, float deltx, float delty,
                                        float deltz, float density, float* DuDt, float* DvDt,
                                        float* DwDt, float* curl, float* p, float* pn,
                                        float* pcountinner);
void omni3dparallellinesESInnerSelect(long Xsize, long Ysize, long Zsize, int NoAngles,
                                      float linespacing, float* k1_d, float* k2_d, float* k3_d,
                                      long* index, float deltx, float delty, float deltz,
                                      float density, float* DuDt, float* DvDt, float* DwDt,
                                      float* curl, float* p, float* pn, float* pcountinner,
                                      float threshold);
void omni3dparallellinesESInnerSelectFixedBC(long Xsize, long Ysize, long Zsize, int NoAngles,
                                             float linespacing, float* k1_d, float* k2_d,
                                             float* k3_d, long* index, float deltx, float delty,
                                             float deltz, float density, float* DuDt,
                                             float* DvDt, float* DwDt, float* curl, float* p,
                                             float* pn, float* pcountinner, float threshold);
void omni3dparallellinesESInnerWeightedMiniCurl(long Xsize, long Ysize, long Zsize,
                                                int NoAngles, float linespacing, float* k1_d,
                                                float* k2_d, float* k3_d, long* index,
                                                float deltx, float delty, float deltz,
                                                float density, float* DuDt, float* DvDt,
                                                float* DwDt, float* curl, float* p, float* pn,
                                                float* pcountinner);
void devidecountInner(long Xsize, long Ysize, long Zsize, float* p, float* pn,
                      float* pcountinner);

/* |curl a| * mask with the reference's one-sided wall differences
 * (kernel.cu:8352-8379).  Device pointers, float; mask may be NULL, which
 * is the reference's GPU kernel calCurlofMaterialAcc (kernel.cu:6863-6888,
 * the same differences without the mask). */
int omni3d_curl(int X, int Y, int Z, float deltx, float delty, float deltz, const float* DuDt,
                const float* DvDt, const float* DwDt, const float* mask, float* curl,
                void* hip_stream);

/* Select threshold from the curl histogram (thredholdHistMaterialAccCPU,
 * kernel.cu:8380-8418): the value above which `percentage` of the voxels
 * lie, on 10000 FP32 bins between min(curl, 1e4) and max(curl, -1e4), walked
 * from the top bin with a FP32 running count; the reference's arithmetic,
 * except that its bin walk (j++ from N-1, kernel.cu:8406) and the maximum's
 * bin (N) leave the table: here the walk goes down and the maximum falls in
 * the last bin.  The first form keeps the reference's signature (host
 * buffers, void: failures only through omni3d_last_global_error()); the
 * second takes curl and threshold (one float) as host or device buffers
 * (mem_kind) and returns once *threshold is written. */
void thredholdHistMaterialAccCPU(int Imax, int Jmax, int Kmax, float* curl, float percentage,
                                 float* threshold);
int omni3d_threshold_hist(int X, int Y, int Z, const float* curl, float percentage,
                          float* threshold, int mem_kind, void* hip_stream);
/* The same threshold with nothing allocated, for callers that cannot call
 * hipMalloc (code under HIP graph capture, serving loops); no counterpart in
 * the reference.  curl_d and threshold_d (one float) are device buffers; the
 * work is queued on hip_stream and the call returns without waiting.  The
 * min / max pair and the 10000 bins live in a fixed 40,012-byte scratch array
 * in device memory, one per GPU, looked up on the first call on that GPU and
 * cleared by every call; calls that overlap on one GPU (e.g. on two streams
 * at once) race on it.  Volumes above 2^31 - 1 voxels are rejected
 * (OMNI3D_ERR_ARG). */
int omni3d_threshold_hist_noalloc(int X, int Y, int Z, const float* curl_d, float percentage,
                                  float* threshold_d, void* hip_stream);

/* Boundary-pair pipeline of the README (README:18-28; kernels
 * omni3dparallellinesEqualSpacing kernel.cu:3414-3489 with bodyIntegral
 * kernel.cu:1353-1467, devidecount kernel.cu:3896-3909, BCiteration
 * kernel.cu:5572-5712), in sparse form: per ray the entry-to-exit integral,
 * per (entry, exit) pair the mean over its rays, then `iterations` Jacobi
 * sweeps over the exit voxels: p[out] = mean over the pairs with a non-zero
 * value of p[in] + pint.  p (X*Y*Z, dtype, host or device by mem_kind): in =
 * starting pressure, out = the same with the exit voxels updated (an exit
 * voxel whose pairs are all zero keeps its value; the reference divides 0 by
 * 0 there).  The README's inner pass follows as omni3d_reconstruct with the
 * plain method and max_iterations = 1 from this p.  *npairs_out (may be
 * NULL) gets the number of pairs.  Directions, field and the context's
 * precision (ray sums) as for omni3d_reconstruct. */
int omni3d_boundary_pairs(omni3d_ctx* ctx, int iterations, void* p, int dtype, int mem_kind,
                          long long* npairs_out);

/* Adjoint pair of the linear operator g -> p (plain or plain-mask method,
 * zero initial pressure, the context's iteration count, no seed term).
 * forward: grad p field (3 components, device F32 or F64 by the context's
 * precision) -> p; backward: dL/dp -> dL/d(grad p).  Component layout:
 * g[c*V + v], c = x,y,z.  With the affine solver the backward iterates over
 * the source voxels only: ray entries for plain (its first call on a plan
 * builds the transposed source rows, kept with the plan), run starts for
 * plain-mask (rebuilt per call from the mask). */
int omni3d_forward(omni3d_ctx* ctx, const void* grad_d, void* p_d);
int omni3d_backward(omni3d_ctx* ctx, const void* dldp_d, void* dldgrad_d);

/* GPU red-black SOR on the 7-point Neumann operator of sor3D (same 26
 * boundary stencils), omega e.g. 1.8; iterates until the mean |residual|
 * (sor3D.cpp:741-781) is below eps or max_iter red+black sweeps.  Host
 * buffers, double.  *iterations_out / *residual_out may be NULL. */
int omni3d_sor_solve(int X, int Y, int Z, double omega, int max_iter, double eps,
                     const double* rhs, double* p, int* iterations_out, double* residual_out);
/* The same solver on device buffers (double, resident in HBM), queued on
 * hip_stream (NULL = the null stream); returns after the last sweep.  No
 * reference counterpart: for callers that keep the field on the GPU. */
int omni3d_sor_solve_device(int X, int Y, int Z, double omega, int max_iter, double eps,
                            const double* rhs_d, double* p_d, void* hip_stream,
                            int* iterations_out, double* residual_out);

/* The reference's CPU entry points, same signatures (sor3D.h:5, calRHS.h:4),
 * executed on the GPU with host buffers.  sor3D: damped Jacobi (weight 0.8)
 * over the BCScheme range, two sweeps per NoItr pair, stop when the mean
 * |interior residual| < eps (then p = pn over the whole grid); x, y, z are
 * unused, as in the reference.  calRHS writes RHS (the reference computes it
 * into a local copy and loses it, calRHS.cpp:10).  Being void, they report
 * failures only through omni3d_last_global_error(). */
void sor3D(int Xsize, int Ysize, int Zsize, int NoItr, int* BCScheme, double* x, double* y,
           double* z, double* p, double* pn, double* RHS, double eps);
void calRHS(int Xsize, int Ysize, int Zsize, double deltx, double delty, double deltz,
            double density, double* DuDt, double* DvDt, double* DwDt, double* RHS);
/* The same two with a status return; omni3d_sor3d_ex also returns the
 * number of sweeps run and the last residue (may be NULL). */
int omni3d_sor3d_ex(int X, int Y, int Z, int NoItr, const int* BCScheme, double* p, double* pn,
                    const double* rhs, double eps, int* sweeps_out, double* residue_out);
int omni3d_calrhs(int X, int Y, int Z, double dx, double dy, double dz, double density,
                  const double* DuDt, const double* DvDt, const double* DwDt, double* rhs);

/* initialIntegration (initialIntegration.h:3, initialIntegration.cpp:9-408),
 * same signature, host buffers, run on the GPU: p[0] = pn[0] = pref, then
 * from the centre voxel (X/2, Y/2, Z/2), which keeps the caller's pn, every
 * voxel becomes the mean of the one-step trapezoid predictions (rho * delta *
 * (a[v] + a[nb]) / 2) from its neighbours toward the centre along the axes on
 * which it is off the centre; p = pn on return.  pBC is not read (nor is it
 * by the reference).  Void: failures only through
 * omni3d_last_global_error(); omni3d_initial_integration returns a status. */
void initialIntegration(long Xsize, long Ysize, long Zsize, double deltx, double delty,
                        double deltz, double density, double* DuDt, double* DvDt, double* DwDt,
                        double pref, double* p, double* pn, double* pBC);
int omni3d_initial_integration(long X, long Y, long Z, double dx, double dy, double dz,
                               double density, const double* DuDt, const double* DvDt,
                               const double* DwDt, double pref, double* p, double* pn);

/* Snapshot batches (no reference counterpart; the reference loops over the
 * files of a folder, kernel.cu:8650-8960).  `count` snapshots of grids
 * nx[i] x ny[i] x nz[i] sharing the other parameters of `base`; snapshots of
 * one grid share one geometry plan.  omni3d_batch_run reconstructs every
 * snapshot from a zero pressure: use_graph = 1 replays one hipGraph holding
 * every snapshot's reconstruction (captured on the first run over four
 * fork-join branches; convergence is per snapshot), use_graph = 0 queues
 * them one by one with a synchronisation each.  *ms_out gets the GPU time.
 * omni3d_batch_tune times the tile shapes (8^3 / 16^3 tiles x 256 / 512 /
 * 1024 threads) on the first snapshot of each distinct grid and keeps the
 * fastest; it returns the number of distinct grids (< 0: error) and writes
 * the choices in the order of the grids' first snapshots (arrays may be
 * NULL).  Results: omni3d_batch_result copies snapshot i's pressure
 * (dtype/mem_kind) and fills its report (iterations, rel_change, sizes). */
typedef struct omni3d_batch omni3d_batch;
int omni3d_batch_create(const omni3d_params* base, int count, const int* nx, const int* ny,
                        const int* nz, omni3d_batch** out);
void omni3d_batch_destroy(omni3d_batch* b);
const char* omni3d_batch_last_error(const omni3d_batch* b);
int omni3d_batch_set_directions(omni3d_batch* b, const float* k1, const float* k2,
                                const float* k3, int count);
int omni3d_batch_set_field(omni3d_batch* b, int index, const void* dudt, const void* dvdt,
                           const void* dwdt, const void* mask, int dtype, int mem_kind);
int omni3d_batch_tune(omni3d_batch* b, int* tile_bits_out, int* threads_out);
int omni3d_batch_run(omni3d_batch* b, int use_graph, double* ms_out);
int omni3d_batch_result(omni3d_batch* b, int index, void* p, int dtype, int mem_kind,
                        omni3d_report* rep);

/* |curl a| * mask of the context's field (calCurlofMaterialAccCPU,
 * kernel.cu:8352-8379, in FP64 from the stored field) into out. */
int omni3d_field_curl(omni3d_ctx* ctx, void* out, int dtype, int mem_kind);

/* Parameter_Omni3D.dat (kernel.cu:8540-8571): lines 1-15 as below, line 16
 * pref and lines 17-22 the region of interest (blank = 0 / whole grid). */
typedef struct omni3d_parameter_file {
  int nx, ny, nz;            /* lines 1-3 (the grid of the files, before the crop) */
  double dx, dy, dz;         /* lines 4-6 */
  double density;            /* line 7 */
  int scale;                 /* line 8 (read, unused, as in the reference) */
  double line_spacing;       /* line 9 */
  int n_angles;              /* line 10: rows of the direction file to read */
  char grid_file[1024];      /* line 11: direction file */
  char acc_dir[1024];        /* line 12: acceleration folder */
  char pressure_dir[1024];   /* line 13: output folder (created if missing) */
  int n_iterations;          /* line 14 */
  double threshold;          /* line 15: curl threshold of the Select method */
  double pref;               /* line 16: Bernoulli reference pressure */
  int roi[6];                /* lines 17-22: i0, i1, j0, j1, k0, k1 (inclusive) */
} omni3d_parameter_file;
int omni3d_read_parameter_file(const char* path, omni3d_parameter_file* out);

/* Process-level drop-in for Omni3D_8.0.exe (kernel.cu:8467-8983): reads the
 * Parameter_Omni3D.dat at `path` (relative paths in it are taken from the
 * file's folder, '\' read as '/'), and for every file of the acceleration
 * folder (sorted; Tecplot POINT, 3 header lines, columns x y z u v w DuDt
 * DvDt DwDt mask, i fastest) crops the region of interest, starts from the
 * previous file's result (zero before the first; kernel.cu:8863) with the j = 0
 * face set to pref - rho |u|^2 / 2, runs the Select method (faithful, curl
 * threshold of line 15) for the line-14 iterations and writes
 * PressureOmni3D_%05d.dat (X Y Z P Count, %15.9f); file 0 also writes
 * CurlofMaterialAcc_Sample.dat; log.dat lists the files.  Outputs and
 * log.dat go beside the parameter file.  Returns the number of files
 * reconstructed, or -(error code).  _ex chooses the precision
 * (OMNI3D_FP64 / OMNI3D_FP32_KAHAN); the plain call uses FP64. */
int omni3d_run_parameter_file(const char* path);
int omni3d_run_parameter_file_ex(const char* path, int precision);

#ifdef __cplusplus
}
#endif
#endif
