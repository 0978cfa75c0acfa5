// Ray geometry shared by the host plan code and the gfx950 kernels.
//
// Everything here is FP64 and evaluated without FMA contraction
// (-ffp-contract=off) so that the host and the device produce bit-identical
// entry/exit voxels and staircase paths.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define OMNI_HD __host__ __device__ __forceinline__

namespace omni3d {

constexpr double kZero = 1e-7;  // kernel.cu:22 "zero"
constexpr int kSentinelStep = -60000;  // kernel.cu:1769: "ilast - 60000"
constexpr int kMaxDim = 1023;   // voxel coordinates are packed in 10 bits

// Per-direction constants.  k is the normalised direction; e1, e2 span the
// lattice plane (kernel.cu:6455-6457); m = I - k k^T (symmetric) gives the
// change of the perpendicular offset per unit axis step.
struct DirConst {
  double k[3];
  double e1[3];
  double e2[3];
  double m[3][3];
};

OMNI_HD uint32_t pack_voxel(int i, int j, int k) {
  return (uint32_t)i | ((uint32_t)j << 10) | ((uint32_t)k << 20);
}
OMNI_HD void unpack_voxel(uint32_t w, int c[3]) {
  c[0] = (int)(w & 1023u);
  c[1] = (int)((w >> 10) & 1023u);
  c[2] = (int)((w >> 20) & 1023u);
}

// NoGrid = int(max(X,Y,Z) * 1.732 / linespacing), kernel.cu:6157-6167.
OMNI_HD int lattice_size(int X, int Y, int Z, double ls) {
  int m = X > Y ? X : Y;
  m = m > Z ? m : Z;
  return (int)((double)m * 1.732 / ls);
}

// Origin of lattice line `point` in box-centred coordinates, kernel.cu:6172-6186.
OMNI_HD void lattice_origin(const DirConst& d, int ng, double ls, int64_t point, double o[3]) {
  const double half = 0.5 * (double)(ng - 1);
  const double xp = ((double)(point / ng) - half) * ls;
  const double yp = ((double)(point % ng) - half) * ls;
  o[0] = xp * d.e1[0] + yp * d.e2[0];
  o[1] = xp * d.e1[1] + yp * d.e2[1];
  o[2] = xp * d.e1[2] + yp * d.e2[2];
}

OMNI_HD int nearest(double v, double c) { return (int)floor(v + c + 0.5); }

// Face-hit bookkeeping: the first accepted face is the entry, each later one
// replaces the exit (kernel.cu:683-735).
struct FaceHits {
  int in[3], out[3];
  bool any;
  OMNI_HD void add(int i, int j, int k) {
    int* t = any ? out : in;
    t[0] = i;
    t[1] = j;
    t[2] = k;
    any = true;
  }
};

// Line clipped against the grid box [-c, c] (c = (dims-1)/2): entry and exit
// voxels per the seven cases of kernel.cu:663-1113.  The planar cases (one
// exact zero component) test the four faces of the two free axes, lower axis
// first; the axis-aligned cases (two components below 1e-7) take the fixed
// voxel column; the general case tests x-,x+,y-,y+,z-,z+.  A planar line whose
// fixed coordinate is not strictly inside falls through to the later cases.
// Returns true when the launcher accepts the ray (kernel.cu:6201-6206): both
// voxels in the box and distinct.
OMNI_HD bool clip_line(int X, int Y, int Z, const double o[3], const double k[3], int in[3],
                       int out[3]) {
  const int n[3] = {X, Y, Z};
  const double c[3] = {(X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0};
  FaceHits h;
  h.any = false;
  for (int a = 0; a < 3; a++) {
    h.in[a] = n[a];
    h.out[a] = n[a];
  }
  bool done = false;
  // planar cases 1..3: exactly one zero component
  for (int z0 = 0; z0 < 3 && !done; z0++) {
    const int a = z0 == 0 ? 1 : 0;  // lower free axis
    const int b = z0 == 2 ? 1 : 2;  // upper free axis
    if (!(k[z0] == 0 && k[a] != 0 && k[b] != 0)) continue;
    if (!(o[z0] > -c[z0] && o[z0] < c[z0])) continue;
    const int fixed = nearest(o[z0], c[z0]);
    // faces of axis a (other coordinate b), then faces of axis b (other a)
    for (int f = 0; f < 4; f++) {
      const int ax = f < 2 ? a : b, ot = f < 2 ? b : a;
      const double side = (f & 1) ? c[ax] : -c[ax];
      const double r = (side - o[ax]) / k[ax];
      const double v = o[ot] + k[ot] * r;
      if (v <= c[ot] && v >= -c[ot]) {
        int t[3];
        t[z0] = fixed;
        t[ax] = (f & 1) ? n[ax] - 1 : 0;
        t[ot] = nearest(v, c[ot]);
        h.add(t[0], t[1], t[2]);
      }
    }
    if (h.any && (double)(h.out[a] - h.in[a]) * k[a] + (double)(h.out[b] - h.in[b]) * k[b] < 0) {
      int t = h.in[a]; h.in[a] = h.out[a]; h.out[a] = t;
      t = h.in[b]; h.in[b] = h.out[b]; h.out[b] = t;
    }
    done = true;
  }
  // axis-aligned cases 4 (z), 5 (y), 6 (x), tested in that order
  for (int q = 0; q < 3 && !done; q++) {
    const int ax = 2 - q;  // the non-negligible axis
    const int a = ax == 0 ? 1 : 0;
    const int b = ax == 2 ? 1 : 2;
    if (!(fabs(k[a]) < kZero && fabs(k[b]) < kZero && fabs(k[ax]) >= kZero)) continue;
    if (o[a] > -c[a] && o[a] < c[a] && o[b] > -c[b] && o[b] < c[b]) {
      h.in[a] = h.out[a] = nearest(o[a], c[a]);
      h.in[b] = h.out[b] = nearest(o[b], c[b]);
      h.in[ax] = k[ax] > 0 ? 0 : n[ax] - 1;
      h.out[ax] = k[ax] > 0 ? n[ax] - 1 : 0;
      done = true;
    } else if (ax == 0) {
      done = true;  // case 6 returns with the sentinels (kernel.cu:989)
    }
  }
  // general case 7
  if (!done && fabs(k[0]) >= kZero && fabs(k[1]) >= kZero && fabs(k[2]) >= kZero) {
    for (int f = 0; f < 6; f++) {
      const int ax = f >> 1;
      const int a = ax == 0 ? 1 : 0;
      const int b = ax == 2 ? 1 : 2;
      const double side = (f & 1) ? c[ax] : -c[ax];
      const double r = (side - o[ax]) / k[ax];
      const double va = o[a] + k[a] * r;
      const double vb = o[b] + k[b] * r;
      if (va <= c[a] && va >= -c[a] && vb <= c[b] && vb >= -c[b]) {
        int t[3];
        t[ax] = (f & 1) ? n[ax] - 1 : 0;
        t[a] = nearest(va, c[a]);
        t[b] = nearest(vb, c[b]);
        h.add(t[0], t[1], t[2]);
      }
    }
    if ((double)(h.out[0] - h.in[0]) * k[0] + (double)(h.out[1] - h.in[1]) * k[1] +
            (double)(h.out[2] - h.in[2]) * k[2] <
        0) {
      for (int a = 0; a < 3; a++) {
        int t = h.in[a]; h.in[a] = h.out[a]; h.out[a] = t;
      }
    }
  }
  for (int a = 0; a < 3; a++) {
    in[a] = h.in[a];
    out[a] = h.out[a];
  }
  for (int a = 0; a < 3; a++) {
    if (in[a] < 0 || in[a] >= n[a] || out[a] < 0 || out[a] >= n[a]) return false;
  }
  return !(in[0] == out[0] && in[1] == out[1] && in[2] == out[2]);
}

// Perpendicular offset of voxel centre `v` from the line through o + c with
// direction k: u = w - (w.k) k, w = v - (o + c).  The walk compares
//   dist^2(v + s e_a) - dist^2(v) = 2 s u_a + s^2 (1 - k_a^2)
// instead of the reference's three sqrt distances (kernel.cu:1800-1810):
// same argmin, same tie order.
OMNI_HD void perp_offset(const DirConst& d, const int v[3], const double o[3], const double cc[3],
                         double u[3]) {
  double w[3];
  for (int a = 0; a < 3; a++) w[a] = (double)v[a] - (o[a] + cc[a]);
  const double t = w[0] * d.k[0] + w[1] * d.k[1] + w[2] * d.k[2];
  for (int a = 0; a < 3; a++) u[a] = w[a] - t * d.k[a];
}

// Step choice of kernel.cu:1764-1840.  Returns the axis, or -1 when the
// chosen candidate leaves the box (the walk ends, kernel.cu:1841).
OMNI_HD int choose_step(const int cur[3], const int end[3], const int dims[3], const double u[3],
                        double q0, double q1, double q2, int* step) {
  int s[3];
  double f[3];
  const double q[3] = {q0, q1, q2};
  for (int a = 0; a < 3; a++) {
    s[a] = cur[a] < end[a] ? 1 : (cur[a] > end[a] ? -1 : kSentinelStep);
    const double dd = (double)s[a];
    f[a] = (2.0 * dd) * u[a] + (dd * dd) * q[a];
  }
  int a;
  if (f[0] <= f[1] && f[0] <= f[2]) a = 0;
  else if (f[1] < f[0] && f[1] <= f[2]) a = 1;
  else if (f[2] < f[0] && f[2] < f[1]) a = 2;
  else return -1;
  const int nxt = cur[a] + s[a];
  if (nxt < 0 || nxt >= dims[a]) return -1;
  *step = s[a];
  return a;
}

}  // namespace omni3d
