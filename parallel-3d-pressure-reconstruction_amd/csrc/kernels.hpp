// Device data structures and kernel launch wrappers of libomni3d.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "geometry.hpp"

namespace omni3d {

// One accepted ray (line x direction pair), 16 bytes.
struct RayRec {
  uint32_t in;       // packed entry voxel (10 bits per axis)
  uint32_t out;      // packed exit voxel
  uint32_t point;    // lattice index
  uint32_t dir_len;  // direction (low 16 bits) | L1 length (high 16 bits)
};

struct GridDesc {
  int X, Y, Z;
  int ng;        // lattice size NoGrid
  double ls;     // line spacing
  double cc[3];  // (dims - 1) / 2
};

// Scalars shared between the kernels of one reconstruction, kept on the device
// so that the iteration needs no host round trip.
struct DevScalars {
  unsigned long long amax_bits;  // max |a| (bit pattern of a non-negative double)
  unsigned long long pmax_bits;  // max |p_init|
  double scale;                  // fixed-point scale, a power of two
  double inv_scale;
  double rel;                    // last relative change
  double dnum, dden;             // its numerator / denominator
  int status;                    // 0 ok, OMNI3D_ERR_NONFINITE
  int iter_done;                 // iterations whose p_S is computed
  int converged;                 // 1 once the stop rule fired
  int kstar;                     // iterations the output represents
  unsigned int ticket;           // last-block counter of the reductions
  unsigned int pad;
  unsigned long long steps;      // staircase steps of the last full pass
  unsigned long long itmax_bits[3];  // max |p_S| (or increment) after iteration k, slot k & 1
                                     // (k % 3 with the finish folded into the next launch)
};

// Fingerprint of a kernel-level drop-in's inputs (directions; field and
// curl) and a flag for a non-finite input or p.
struct InputPrint {
  unsigned long long dirs;
  unsigned long long field;
  unsigned int nonfinite;
  unsigned int pad;
};

// Method-specific walk parameters.
struct WalkParams {
  double rho;
  double hs[3][2];   // ((-rho * s) * delta_a) * 0.5 for s = -1 (index 0) / +1 (index 1)
  float hsf[3][2];   // the same factors rounded to FP32
  double del[3];     // grid spacing (the trilinear walk's displacements)
  double thr;        // Select threshold (+inf for plain-mask)
  int method;
  int faithful;
};

}  // namespace omni3d
