/*
 * omni3d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference's omni-directional integration
 * (Omni3D_8.0/Omni3D_8.0/kernel.cu) and of its CPU Poisson path
 * (sor3D.cpp, calRHS.cpp).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; it is the checker,
 * never the thing measured or shipped.
 *
 * Layout everywhere: x fastest, index = i + j*X + k*X*Y (kernel.cu:1760).
 */
#ifndef OMNI3D_ORACLE_H
#define OMNI3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PLAIN = 0, ORC_PLAIN_MASK = 1, ORC_SELECT = 2, ORC_WEIGHTED = 3,
       ORC_SELECT_FIXEDBC = 4, /* bodyIntegralInnerSelectFixedBC, kernel.cu:2538-2796 */
       ORC_MINICURL = 5, /* bodyIntegralInnerMiniCurl, kernel.cu:2797-2936 */
       /* The plain method's visits and sources with the line integral taken
        * along the straight ray through trilinear samples of the gradient
        * (no reference counterpart: the accuracy mode beside the staircase
        * of kernel.cu:1743-1863; see walk_plain_trilinear). */
       ORC_PLAIN_TRILINEAR = 6,
       ORC_WEIGHTED_FIXEDBC = 7, /* bodyIntegralInnerWeightedFixedBC, kernel.cu:2136-2279 */
       ORC_WEIGHTED_MINICURL = 8 /* bodyIntegralInnerWeightedMiniCurl, kernel.cu:2937-3085 */ };

/* Per-direction constants derived from one unit vector. */
typedef struct {
  double k[3];     /* normalised direction */
  double e1[3];    /* lattice frame, kernel.cu:6455-6457 */
  double e2[3];
  double M[3][3];  /* I - k k^T : change of the perpendicular offset per unit step */
} orc_dir;

/* Build the per-direction constants (normalise, frame).  kernel.cu:6441-6457 */
void orc_make_dir(float k1, float k2, float k3, orc_dir* out);

/* NoGrid = int(max(X,Y,Z) * 1.732 / ls), kernel.cu:6157-6167 */
int orc_nogrid(int X, int Y, int Z, double ls);

/* Ray origin of lattice point `point` (centred coordinates), kernel.cu:6172-6186 */
void orc_origin(const orc_dir* d, int nogrid, double ls, int64_t point, double o[3]);

/* Entry/exit voxels of a line, kernel.cu:663-1113 (case 3 uses y, not z; the
 * reference's z there, kernel.cu:854-855, is a typo).  Returns 1 if the
 * launcher accepts the ray (both voxels in the box and in != out). */
int orc_clip(int X, int Y, int Z, const double o[3], const double k[3], int in[3], int out[3]);

/* Enumerate every accepted ray of every direction (full coverage).
 * rays: 9 int32 per ray = {d, point(lo 31 bits), point(hi), iin, jin, kin, iout, jout, kout}
 * Returns the ray count (rows are written only while it is below capacity;
 * call with capacity 0 to size the buffer). */
int64_t orc_enumerate_rays(int X, int Y, int Z, int D, const float* k1, const float* k2,
                           const float* k3, double ls, int32_t* rays, int64_t capacity);

/* Staircase path of one ray as voxel indices (entry first); returns its
 * length.  fixedbc: the step rule of bodyIntegralInnerSelectFixedBC. */
int orc_ray_path(int X, int Y, int Z, float k1, float k2, float k3, double ls, int64_t point,
                 int64_t* path, int capacity, int fixedbc);

/* One full pass (kernel.cu:6151-6215 / 6281-6346 / 6414-6480 with full
 * (line x direction) coverage): pn += contributions, cnt += weights.
 * curl: for SELECT/WEIGHTED the |curl a| * mask field; for PLAIN_MASK the
 * mask (0 = masked).  src_mark (may be NULL) gets 1 at every voxel whose p is
 * read by the pass (entry / restart voxels). */
void orc_pass(int X, int Y, int Z, int D, const float* k1, const float* k2, const float* k3,
              double ls, double dx, double dy, double dz, double rho, const double* ax,
              const double* ay, const double* az, const double* curl, double threshold,
              int method, int faithful, const double* p, double* pn, double* cnt,
              uint8_t* src_mark);

/* devidecountInner, kernel.cu:6746-6759 (GPU semantics: cnt > 0). */
void orc_divide(int64_t V, double* p, double* pn, const double* cnt);

typedef struct {
  int iterations;
  double rel_change;  /* ||p_S^k - p_S^{k-1}|| / ||p_S^k|| over source voxels S */
} orc_report;

/* Iteration driver, kernel.cu:8805-8861, plus the relative-change stop.
 * p: in = initial p (kernel.cu:8806-8813), out = result.  cnt_out may be NULL.
 * seed_pn: 1 = pn starts as a copy of p (kernel.cu:8813), 0 = pn starts at 0. */
void orc_reconstruct(int X, int Y, int Z, int D, const float* k1, const float* k2,
                     const float* k3, double ls, double dx, double dy, double dz, double rho,
                     const double* ax, const double* ay, const double* az, const double* curl,
                     double threshold, int method, int faithful, int max_iter, double tol,
                     int seed_pn, double* p, double* cnt_out, orc_report* rep);

/* calCurlofMaterialAccCPU, kernel.cu:8352-8379 (in FP64). */
void orc_curl(int X, int Y, int Z, double dx, double dy, double dz, const double* ax,
              const double* ay, const double* az, const double* mask, double* curl);

/* calRHS, calRHS.cpp:7-309, writing the caller's buffer. */
void orc_calrhs(int X, int Y, int Z, double dx, double dy, double dz, double rho,
                const double* ax, const double* ay, const double* az, double* rhs);

/* sor3D, sor3D.cpp:9-1589: damped Jacobi (weight 0.8) with the 26 boundary
 * stencils, mean-|residual| stop.  Returns the number of sweeps done;
 * *residue_out gets the last residue (float accumulation as sor3D.cpp:18-20). */
int orc_sor3d(int X, int Y, int Z, int noitr, const int* bcscheme, double* p, double* pn,
              const double* rhs, double eps, double* residue_out);

/* Red-black SOR with the same stencils (omega given), for the GPU solver's
 * parity: one "iteration" = red half-sweep then black half-sweep. */
void orc_sor_redblack(int X, int Y, int Z, int iters, double omega, double* p, const double* rhs);

/* Interior residual of the 7-point operator, sum |sum_nb - 6p + rhs| / V (FP64). */
double orc_residual(int X, int Y, int Z, const double* p, const double* rhs);

/* Boundary-pair pipeline (kernel.cu:3414-3489, 1353-1467, 3896-3909,
 * 5572-5712) on a sorted pair list, Jacobi boundary iterations: p (X*Y*Z)
 * in/out, only exit voxels change.  Returns the number of (in, out) pairs. */
int64_t orc_boundary_pairs(int X, int Y, int Z, int D, const float* k1, const float* k2,
                           const float* k3, double ls, double dx, double dy, double dz,
                           double rho, const double* ax, const double* ay, const double* az,
                           int iters, double* p);

/* initialIntegration, initialIntegration.cpp:9-408 (FP64, host buffers;
 * pBC is unused there and omitted here). */
void orc_initial_integration(int X, int Y, int Z, double dx, double dy, double dz, double rho,
                             const double* ax, const double* ay, const double* az, double pref,
                             double* p, double* pn);

/* kernel.cu:8380-8418 (intended bin walk, see the definition) */
float orc_threshold_hist(int X, int Y, int Z, const float* curl, float percentage);

#ifdef __cplusplus
}
#endif
#endif
