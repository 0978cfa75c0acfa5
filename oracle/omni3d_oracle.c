/*
 * omni3d_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker; see the header).
 *
 * A serial FP64 restatement of the reference's inner-iteration method with
 * full (line x direction) coverage and exact (non-racy) accumulation.
 * Citations are to Omni3D_8.0/Omni3D_8.0/kernel.cu unless stated.
 *
 * The step rule.  The reference picks, among the three axis neighbours of
 * the current voxel that move toward the exit voxel, the one whose centre is
 * closest to the ideal line (kernel.cu:1764-1810: three sqrt distances,
 * ties to x then y, kernel.cu:1815,1824,1833).  Here the comparison uses the
 * exact identity
 *     dist^2(c + s*e_a) - dist^2(c) = 2 s u_a + s^2 (1 - k_a^2),
 * where u = (c - o) - ((c - o).k) k is the perpendicular offset of the
 * current voxel centre, updated per step by u += s (e_a - k_a k).  sqrt is
 * monotone, so this is the same argmin with the same tie order; only the
 * rounding at exact ties can differ from the reference's FP32 evaluation.
 * The "-60000" sentinel for an axis already at the exit coordinate
 * (kernel.cu:1769,1781,1793) is kept as a step of s = -60000 whose
 * out-of-box target ends the walk (kernel.cu:1841-1861).
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction: the HIP library is
 * compiled the same way so both evaluate identical expressions).
 */
#include "omni3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_ZERO 1e-7 /* kernel.cu:22 */
#define ORC_PI 3.1415926535897932384626433832795

static inline int64_t vidx(int X, int Y, int i, int j, int k) {
  return (int64_t)i + (int64_t)j * X + (int64_t)k * X * Y;
}

/* kernel.cu:6441-6457: theta = acos(k3), phi = asin(k2/sin theta) flipped by
 * pi when k1/sin theta < 0; frame e1 = (cos t cos f, cos t sin f, -sin t),
 * e2 = (-sin f, cos f, 0).  The asin argument is clamped to [-1, 1]. */
void orc_make_dir(float k1, float k2, float k3, orc_dir* out) {
  double a = k1, b = k2, c = k3;
  double n = sqrt(a * a + b * b + c * c);
  double k[3] = {a / n, b / n, c / n};
  double theta = acos(k[2]);
  double st = sin(theta), ct = cos(theta);
  double phi = 0.0;
  if (st != 0.0) {
    double s = k[1] / st;
    if (s > 1.0) s = 1.0;
    if (s < -1.0) s = -1.0;
    phi = asin(s);
    if (k[0] / st < 0.0) phi = -phi + ORC_PI;
  }
  double sf = sin(phi), cf = cos(phi);
  for (int i = 0; i < 3; i++) out->k[i] = k[i];
  out->e1[0] = ct * cf;
  out->e1[1] = ct * sf;
  out->e1[2] = -st;
  out->e2[0] = -sf;
  out->e2[1] = cf;
  out->e2[2] = 0.0;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) out->M[r][s] = (r == s ? 1.0 : 0.0) - k[r] * k[s];
}

int orc_nogrid(int X, int Y, int Z, double ls) {
  int m = X;
  if (m < Y) m = Y;
  if (m < Z) m = Z;
  return (int)((double)m * 1.732 / ls);
}

void orc_origin(const orc_dir* d, int nogrid, double ls, int64_t point, double o[3]) {
  double half = 0.5 * (double)(nogrid - 1);
  double xp = ((double)(point / nogrid) - half) * ls;
  double yp = ((double)(point % nogrid) - half) * ls;
  for (int i = 0; i < 3; i++) o[i] = xp * d->e1[i] + yp * d->e2[i];
}

/* One candidate face hit of cross2point: the first accepted hit is the entry,
 * every later one overwrites the exit (kernel.cu:683-735). */
static inline void hit(int* flag, int in[3], int out[3], int i, int j, int k) {
  if (*flag == 0) {
    in[0] = i; in[1] = j; in[2] = k;
  } else {
    out[0] = i; out[1] = j; out[2] = k;
  }
  *flag = 1;
}

static inline int rnd(double v, double c) { return (int)floor(v + c + 0.5); }

/* kernel.cu:663-1113.  Returns the reference's bool; in/out get sentinels
 * (X,Y,Z) when no face pair is found. */
static int cross2point(int X, int Y, int Z, const double o[3], const double kk[3], int in[3],
                       int out[3]) {
  const double cx = (X - 1) / 2.0, cy = (Y - 1) / 2.0, cz = (Z - 1) / 2.0;
  const double xin = o[0], yin = o[1], zin = o[2];
  const double k1 = kk[0], k2 = kk[1], k3 = kk[2];
  in[0] = X; in[1] = Y; in[2] = Z;
  out[0] = X; out[1] = Y; out[2] = Z;
  /* case 1: line perpendicular to x (kernel.cu:672-756) */
  if (k1 == 0 && k2 != 0 && k3 != 0) {
    if (xin > -cx && xin < cx) {
      double r = (-cy - yin) / k2; double z1 = zin + k3 * r;
      r = (cy - yin) / k2; double z2 = zin + k3 * r;
      r = (-cz - zin) / k3; double y3 = yin + k2 * r;
      r = (cz - zin) / k3; double y4 = yin + k2 * r;
      int flag = 0, ii = rnd(xin, cx);
      if (z1 <= cz && z1 >= -cz) hit(&flag, in, out, ii, 0, rnd(z1, cz));
      if (z2 <= cz && z2 >= -cz) hit(&flag, in, out, ii, Y - 1, rnd(z2, cz));
      if (y3 <= cy && y3 >= -cy) hit(&flag, in, out, ii, rnd(y3, cy), 0);
      if (y4 <= cy && y4 >= -cy) hit(&flag, in, out, ii, rnd(y4, cy), Z - 1);
      if (flag != 0 && (double)(out[1] - in[1]) * k2 + (double)(out[2] - in[2]) * k3 < 0) {
        int t = in[1]; in[1] = out[1]; out[1] = t;
        t = in[2]; in[2] = out[2]; out[2] = t;
      }
      return 1;
    }
  }
  /* case 2: perpendicular to y (kernel.cu:758-845) */
  if (k1 != 0 && k2 == 0 && k3 != 0) {
    if (yin > -cy && yin < cy) {
      double r = (-cx - xin) / k1; double z1 = zin + k3 * r;
      r = (cx - xin) / k1; double z2 = zin + k3 * r;
      r = (-cz - zin) / k3; double x3 = xin + k1 * r;
      r = (cz - zin) / k3; double x4 = xin + k1 * r;
      int flag = 0, jj = rnd(yin, cy);
      if (z1 <= cz && z1 >= -cz) hit(&flag, in, out, 0, jj, rnd(z1, cz));
      if (z2 <= cz && z2 >= -cz) hit(&flag, in, out, X - 1, jj, rnd(z2, cz));
      if (x3 <= cx && x3 >= -cx) hit(&flag, in, out, rnd(x3, cx), jj, 0);
      if (x4 <= cx && x4 >= -cx) hit(&flag, in, out, rnd(x4, cx), jj, Z - 1);
      if (flag != 0 && (double)(out[0] - in[0]) * k1 + (double)(out[2] - in[2]) * k3 < 0) {
        int t = in[0]; in[0] = out[0]; out[0] = t;
        t = in[2]; in[2] = out[2]; out[2] = t;
      }
      return 1;
    }
  }
  /* case 3: perpendicular to z (kernel.cu:846-933); y-face intersections use
   * yin (the reference's zin at kernel.cu:854-855 is a typo). */
  if (k1 != 0 && k2 != 0 && k3 == 0) {
    if (zin > -cz && zin < cz) {
      double r = (-cx - xin) / k1; double y1 = yin + k2 * r;
      r = (cx - xin) / k1; double y2 = yin + k2 * r;
      r = (-cy - yin) / k2; double x3 = xin + k1 * r;
      r = (cy - yin) / k2; double x4 = xin + k1 * r;
      int flag = 0, kz = rnd(zin, cz);
      if (y1 <= cy && y1 >= -cy) hit(&flag, in, out, 0, rnd(y1, cy), kz);
      if (y2 <= cy && y2 >= -cy) hit(&flag, in, out, X - 1, rnd(y2, cy), kz);
      if (x3 <= cx && x3 >= -cx) hit(&flag, in, out, rnd(x3, cx), 0, kz);
      if (x4 <= cx && x4 >= -cx) hit(&flag, in, out, rnd(x4, cx), Y - 1, kz);
      if (flag != 0 && (double)(out[0] - in[0]) * k1 + (double)(out[1] - in[1]) * k2 < 0) {
        int t = in[0]; in[0] = out[0]; out[0] = t;
        t = in[1]; in[1] = out[1]; out[1] = t;
      }
      return 1;
    }
  }
  /* case 4: along z (kernel.cu:935-953) */
  if (fabs(k1) < ORC_ZERO && fabs(k2) < ORC_ZERO && fabs(k3) >= ORC_ZERO) {
    if (xin < cx && xin > -cx && yin < cy && yin > -cy) {
      in[0] = out[0] = rnd(xin, cx);
      in[1] = out[1] = rnd(yin, cy);
      if (k3 > 0) { in[2] = 0; out[2] = Z - 1; } else { in[2] = Z - 1; out[2] = 0; }
      return 1;
    }
  }
  /* case 5: along y (kernel.cu:955-972) */
  if (fabs(k1) < ORC_ZERO && fabs(k2) >= ORC_ZERO && fabs(k3) < ORC_ZERO) {
    if (xin > -cx && xin < cx && zin > -cz && zin < cz) {
      in[0] = out[0] = rnd(xin, cx);
      in[2] = out[2] = rnd(zin, cz);
      if (k2 > 0) { out[1] = Y - 1; in[1] = 0; } else { in[1] = Y - 1; out[1] = 0; }
      return 1;
    }
  }
  /* case 6: along x (kernel.cu:974-990); returns true even when the line
   * misses, leaving the sentinels. */
  if (fabs(k1) >= ORC_ZERO && fabs(k2) < ORC_ZERO && fabs(k3) < ORC_ZERO) {
    if (yin > -cy && yin < cy && zin > -cz && zin < cz) {
      in[1] = out[1] = rnd(yin, cy);
      in[2] = out[2] = rnd(zin, cz);
      if (k1 > 0) { out[0] = X - 1; in[0] = 0; } else { in[0] = X - 1; out[0] = 0; }
    }
    return 1;
  }
  /* case 7: general direction, faces in the order x-,x+,y-,y+,z-,z+
   * (kernel.cu:992-1111) */
  if (fabs(k1) >= ORC_ZERO && fabs(k2) >= ORC_ZERO && fabs(k3) >= ORC_ZERO) {
    double r;
    r = (-cx - xin) / k1; double y1 = yin + k2 * r, z1 = zin + k3 * r;
    r = (cx - xin) / k1; double y2 = yin + k2 * r, z2 = zin + k3 * r;
    r = (-cy - yin) / k2; double x3 = xin + k1 * r, z3 = zin + k3 * r;
    r = (cy - yin) / k2; double x4 = xin + k1 * r, z4 = zin + k3 * r;
    r = (-cz - zin) / k3; double x5 = xin + k1 * r, y5 = yin + k2 * r;
    r = (cz - zin) / k3; double x6 = xin + k1 * r, y6 = yin + k2 * r;
    int flag = 0;
    if (y1 <= cy && y1 >= -cy && z1 <= cz && z1 >= -cz) hit(&flag, in, out, 0, rnd(y1, cy), rnd(z1, cz));
    if (y2 <= cy && y2 >= -cy && z2 <= cz && z2 >= -cz) hit(&flag, in, out, X - 1, rnd(y2, cy), rnd(z2, cz));
    if (x3 <= cx && x3 >= -cx && z3 <= cz && z3 >= -cz) hit(&flag, in, out, rnd(x3, cx), 0, rnd(z3, cz));
    if (x4 <= cx && x4 >= -cx && z4 <= cz && z4 >= -cz) hit(&flag, in, out, rnd(x4, cx), Y - 1, rnd(z4, cz));
    if (x5 <= cx && x5 >= -cx && y5 <= cy && y5 >= -cy) hit(&flag, in, out, rnd(x5, cx), rnd(y5, cy), 0);
    if (x6 <= cx && x6 >= -cx && y6 <= cy && y6 >= -cy) hit(&flag, in, out, rnd(x6, cx), rnd(y6, cy), Z - 1);
    if ((double)(out[0] - in[0]) * k1 + (double)(out[1] - in[1]) * k2 + (double)(out[2] - in[2]) * k3 < 0) {
      for (int a = 0; a < 3; a++) { int t = in[a]; in[a] = out[a]; out[a] = t; }
    }
    return 1;
  }
  return 0;
}

int orc_clip(int X, int Y, int Z, const double o[3], const double k[3], int in[3], int out[3]) {
  cross2point(X, Y, Z, o, k, in, out);
  /* launcher acceptance, kernel.cu:6201-6206 (index[in] != index[out] is
   * in != out: the boundary numbering of calIndexGPU is one-to-one). */
  if (in[0] < 0 || in[0] >= X || in[1] < 0 || in[1] >= Y || in[2] < 0 || in[2] >= Z) return 0;
  if (out[0] < 0 || out[0] >= X || out[1] < 0 || out[1] >= Y || out[2] < 0 || out[2] >= Z) return 0;
  if (in[0] == out[0] && in[1] == out[1] && in[2] == out[2]) return 0;
  return 1;
}

/* Walker state shared by the three integrators. */
typedef struct {
  int X, Y, Z;
  int c[3], e[3];
  double u[3];
  const orc_dir* d;
  int fixedbc; /* 1: the step rule of bodyIntegralInnerSelectFixedBC; 2: the
                  plain rule without a y-step onto j = 0 (WeightedFixedBC) */
} walker;

static void walker_init(walker* w, int X, int Y, int Z, const orc_dir* d, const double o[3],
                        const int in[3], const int out[3]) {
  /* kernel.cu:1749-1751: the line passes through o + centre in index space */
  const double cc[3] = {(X - 1) / 2.0, (Y - 1) / 2.0, (Z - 1) / 2.0};
  w->X = X; w->Y = Y; w->Z = Z; w->d = d;
  w->fixedbc = 0;
  double wv[3];
  for (int a = 0; a < 3; a++) {
    w->c[a] = in[a];
    w->e[a] = out[a];
    wv[a] = (double)in[a] - (o[a] + cc[a]);
  }
  double t = wv[0] * d->k[0] + wv[1] * d->k[1] + wv[2] * d->k[2];
  for (int a = 0; a < 3; a++) w->u[a] = wv[a] - t * d->k[a];
}

/* One step decision (kernel.cu:1764-1840).  Returns the axis (0..2) and the
 * signed step in *s, or -1 when the chosen candidate leaves the box / no
 * candidate qualifies ("Wrong Point Found", kernel.cu:1841-1859). */
static int walker_choose(const walker* w, int* s) {
  const int size[3] = {w->X, w->Y, w->Z};
  double f[3];
  int st[3];
  for (int a = 0; a < 3; a++) {
    st[a] = w->c[a] < w->e[a] ? 1 : (w->c[a] > w->e[a] ? -1 : -60000);
    double dd = (double)st[a];
    f[a] = (2.0 * dd) * w->u[a] + (dd * dd) * w->d->M[a][a];
  }
  int a;
  if (w->fixedbc == 1) {
    /* kernel.cu:2628,2676,2722: an x / y tie goes to y; no x- or z-step
     * while on the j = 0 plane, no y-step onto it (the plane keeps its
     * prescribed pressure: only rays entering there read it) */
    if (f[0] < f[1] && f[0] <= f[2]) a = 0;
    else if (f[1] <= f[0] && f[1] <= f[2]) a = 1;
    else if (f[2] < f[0] && f[2] < f[1]) a = 2;
    else return -1;
    if (a != 1 && w->c[1] == 0) return -1;
    if (a == 1 && w->c[1] + st[1] == 0) return -1;
  } else {
    if (f[0] <= f[1] && f[0] <= f[2]) a = 0;
    else if (f[1] < f[0] && f[1] <= f[2]) a = 1;
    else if (f[2] < f[0] && f[2] < f[1]) a = 2;
    else return -1;
    /* kernel.cu:2226: bodyIntegralInnerWeightedFixedBC takes the y-step only
     * when jnext2 > 0; otherwise no branch runs and the walk ends */
    if (w->fixedbc == 2 && a == 1 && w->c[1] + st[1] == 0) return -1;
  }
  int n = w->c[a] + st[a];
  if (n < 0 || n >= size[a]) return -1;
  *s = st[a];
  return a;
}

static void walker_step(walker* w, int a, int s) {
  double dd = (double)s;
  w->c[a] += s;
  for (int b = 0; b < 3; b++) w->u[b] = w->u[b] + dd * w->d->M[a][b];
}

static inline int walker_done(const walker* w) {
  return w->c[0] == w->e[0] && w->c[1] == w->e[1] && w->c[2] == w->e[2];
}

/* trapezoid term -rho*(next-last)*delta*0.5*(a[next]+a[last]) (kernel.cu:1816) */
static inline double trap(double rho, int s, double delta, double an, double al) {
  return -rho * (double)s * delta * 0.5 * (an + al);
}

typedef struct {
  int X, Y, Z;
  double rho, del[3];
  const double* acc[3];
  const double* curl;
  double thr;
  int method, faithful;
  const double* p;
  double* pn;
  double* cnt;
  uint8_t* src;
} pass_ctx;

/* bodyIntegralInner, kernel.cu:1743-1863 */
static int walk_plain(const pass_ctx* P, walker* w) {
  const int X = P->X, Y = P->Y;
  int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  double pin = P->p[vin];
  if (P->src) P->src[vin] = 1;
  P->pn[vin] += pin; /* entry self-add, kernel.cu:1760-1761 */
  P->cnt[vin] += 1.0;
  double pint = 0.0;
  int steps = 0;
  while (!walker_done(w)) {
    int s, a = walker_choose(w, &s);
    if (a < 0) break;
    int64_t vl = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    walker_step(w, a, s);
    int64_t vn = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    pint += trap(P->rho, s, P->del[a], P->acc[a][vn], P->acc[a][vl]);
    P->pn[vn] += pin + pint;
    P->cnt[vn] += 1.0;
    steps++;
  }
  return steps;
}

/* Trilinear sample of the acceleration at the index-space point x (0-based
 * voxel centres).  The cell's lower corner is floor(x) clamped to
 * [0, size - 2] and the weights are not clamped, so a point up to one voxel
 * outside the centres of the faces is extrapolated linearly: the sample of a
 * linear field is exact everywhere.  Interpolation order x, then y, then z. */
static void tri_sample(const pass_ctx* P, const double x[3], double out[3]) {
  const int size[3] = {P->X, P->Y, P->Z};
  int i0[3];
  double f[3], g[3];
  for (int a = 0; a < 3; a++) {
    int fl = (int)floor(x[a]);
    if (fl > size[a] - 2) fl = size[a] - 2;
    if (fl < 0) fl = 0;
    i0[a] = fl;
    f[a] = x[a] - (double)fl;
    g[a] = 1.0 - f[a];
  }
  const int64_t b = vidx(P->X, P->Y, i0[0], i0[1], i0[2]);
  const int64_t sy = P->X, sz = (int64_t)P->X * P->Y;
  for (int c = 0; c < 3; c++) {
    const double* A = P->acc[c];
    const double c00 = A[b] * g[0] + A[b + 1] * f[0];
    const double c10 = A[b + sy] * g[0] + A[b + sy + 1] * f[0];
    const double c01 = A[b + sz] * g[0] + A[b + sz + 1] * f[0];
    const double c11 = A[b + sy + sz] * g[0] + A[b + sy + sz + 1] * f[0];
    const double c0 = c00 * g[1] + c10 * f[1];
    const double c1 = c01 * g[1] + c11 * f[1];
    out[c] = c0 * g[2] + c1 * f[2];
  }
}

/* Trapezoid of grad p = -rho a over a straight path: -rho/2 (a_start +
 * a_end) . (disp * delta), disp in index units. */
static inline double tri_term(const pass_ctx* P, const double ga[3], const double gb[3],
                              const double disp[3]) {
  const double hr = -P->rho * 0.5;
  return hr * (((ga[0] + gb[0]) * (disp[0] * P->del[0]) + (ga[1] + gb[1]) * (disp[1] * P->del[1])) +
               (ga[2] + gb[2]) * (disp[2] * P->del[2]));
}

/* The plain method (kernel.cu:1743-1863: entry self-add, the staircase's
 * visited voxels, p_in + integral at each visit) with the integral taken
 * along the straight ray instead of the staircase.  Each voxel centre c
 * projects onto the line at x = c - u (u: its perpendicular offset, which the
 * walker tracks for the step rule).  The integral from the entry voxel to a
 * visited voxel v follows in -> x_in (leg), x_in -> x_v along the line, one
 * trapezoid per staircase step (the projection advances by s k_a along k),
 * and x_v -> v (leg); the gradient at the points on the line is sampled
 * trilinearly, at voxel centres it is the grid value.  Every piece is exact
 * for a linear gradient, so a quadratic pressure is reproduced exactly. */
static int walk_plain_trilinear(const pass_ctx* P, walker* w) {
  const int X = P->X, Y = P->Y;
  const double* k = w->d->k;
  int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  double pin = P->p[vin];
  if (P->src) P->src[vin] = 1;
  P->pn[vin] += pin;
  P->cnt[vin] += 1.0;
  double x[3], sl[3], ac[3], disp[3];
  for (int a = 0; a < 3; a++) {
    x[a] = (double)w->c[a] - w->u[a];
    ac[a] = P->acc[a][vin];
    disp[a] = -w->u[a];
  }
  tri_sample(P, x, sl);
  double pint = tri_term(P, ac, sl, disp); /* leg in -> x_in */
  int steps = 0;
  while (!walker_done(w)) {
    int s, a = walker_choose(w, &s);
    if (a < 0) break;
    walker_step(w, a, s);
    int64_t vn = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    double sn[3];
    for (int b = 0; b < 3; b++) {
      x[b] = (double)w->c[b] - w->u[b];
      disp[b] = (double)s * (k[a] * k[b]);
    }
    tri_sample(P, x, sn);
    pint += tri_term(P, sl, sn, disp); /* x_l -> x_n along the line */
    for (int b = 0; b < 3; b++) {
      ac[b] = P->acc[b][vn];
      disp[b] = w->u[b];
    }
    const double leg = tri_term(P, sn, ac, disp); /* x_n -> v */
    P->pn[vn] += pin + (pint + leg);
    P->cnt[vn] += 1.0;
    for (int b = 0; b < 3; b++) sl[b] = sn[b];
    steps++;
  }
  return steps;
}

/* bodyIntegralInnerMiniCurl, kernel.cu:2797-2936 (the pass of
 * omni3dparallellinesESInnerWeightedMiniCurl, kernel.cu:6347-6413): the
 * plain sums and unit counts without the entry self-add, along a path that
 * steps toward the exit through the candidate with the smallest curl (ties x,
 * then y, then z, kernel.cu:2883,2893,2903).  A candidate on an axis already
 * at the exit is the reference's sentinel i - 60000 (or j, k): its curl is
 * read at the aliased linear index when that lies inside the volume, else
 * it counts as 1e10, and choosing it ends the walk (kernel.cu:2867-2879). */
static int walk_minicurl(const pass_ctx* P, walker* w) {
  const int X = P->X, Y = P->Y, Z = P->Z;
  const int64_t V = (int64_t)X * Y * Z;
  const int size[3] = {X, Y, Z};
  const int64_t stride[3] = {1, X, (int64_t)X * Y};
  const int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  const double pin = P->p[vin];
  if (P->src) P->src[vin] = 1;
  double pint = 0.0;
  int steps = 0;
  while (!walker_done(w)) {
    const int64_t v = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    int st[3];
    double d[3];
    for (int a = 0; a < 3; a++) {
      st[a] = w->c[a] < w->e[a] ? 1 : (w->c[a] > w->e[a] ? -1 : -60000);
      const int64_t idx = v + (int64_t)st[a] * stride[a];
      d[a] = (idx >= 0 && idx < V) ? P->curl[idx] : 1e10;
    }
    int a;
    if (d[0] <= d[1] && d[0] <= d[2]) a = 0;
    else if (d[1] < d[0] && d[1] <= d[2]) a = 1;
    else if (d[2] < d[0] && d[2] < d[1]) a = 2;
    else break;
    const int n = w->c[a] + st[a];
    if (n < 0 || n >= size[a]) break;
    w->c[a] = n;
    const int64_t vn = v + (int64_t)st[a] * stride[a];
    pint += trap(P->rho, st[a], P->del[a], P->acc[a][vn], P->acc[a][v]);
    P->pn[vn] += pin + pint;
    P->cnt[vn] += 1.0;
    steps++;
  }
  return steps;
}

/* bodyIntegralInnerWeightedMiniCurl, kernel.cu:2937-3085: the path of
 * walk_minicurl; curltmp = curl[entry] and counttmp = 1 at the entry, which
 * adds p[entry] * 1 / curltmp * counttmp (kernel.cu:2956-2960); each step adds
 * curl[next] (not / 1e4) to curltmp, 1 to counttmp, and (p[entry] + pint) *
 * 1 / curltmp * counttmp with no test of curltmp (kernel.cu:3029-3060). */
static int walk_weighted_minicurl(const pass_ctx* P, walker* w) {
  const int X = P->X, Y = P->Y, Z = P->Z;
  const int64_t V = (int64_t)X * Y * Z;
  const int size[3] = {X, Y, Z};
  const int64_t stride[3] = {1, X, (int64_t)X * Y};
  const int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  const double pin = P->p[vin];
  if (P->src) P->src[vin] = 1;
  double curltmp = P->curl[vin], counttmp = 1.0;
  P->pn[vin] += pin * 1 / curltmp * counttmp;
  P->cnt[vin] += 1 / curltmp * counttmp;
  double pint = 0.0;
  int steps = 0;
  while (!walker_done(w)) {
    const int64_t v = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    int st[3];
    double d[3];
    for (int a = 0; a < 3; a++) {
      st[a] = w->c[a] < w->e[a] ? 1 : (w->c[a] > w->e[a] ? -1 : -60000);
      const int64_t idx = v + (int64_t)st[a] * stride[a];
      d[a] = (idx >= 0 && idx < V) ? P->curl[idx] : 1e10;
    }
    int a;
    if (d[0] <= d[1] && d[0] <= d[2]) a = 0;
    else if (d[1] < d[0] && d[1] <= d[2]) a = 1;
    else if (d[2] < d[0] && d[2] < d[1]) a = 2;
    else break;
    const int n = w->c[a] + st[a];
    if (n < 0 || n >= size[a]) break;
    w->c[a] = n;
    const int64_t vn = v + (int64_t)st[a] * stride[a];
    pint += trap(P->rho, st[a], P->del[a], P->acc[a][vn], P->acc[a][v]);
    curltmp += P->curl[vn];
    counttmp += 1.0;
    P->pn[vn] += (pin + pint) * 1 / curltmp * counttmp;
    P->cnt[vn] += 1 / curltmp * counttmp;
    steps++;
  }
  return steps;
}

/* bodyIntegralInnerWeighted, kernel.cu:1993-2135: no entry self-add; weight
 * counttmp/curltmp with curltmp = sum(curl/1e4) over the path so far.  With
 * w->fixedbc == 2 it is bodyIntegralInnerWeightedFixedBC (kernel.cu:2136-2279):
 * the same walk ended by a y-step onto the j = 0 plane. */
static int walk_weighted(const pass_ctx* P, walker* w) {
  const int X = P->X, Y = P->Y;
  int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  double pin = P->p[vin];
  if (P->src) P->src[vin] = 1;
  double pint = 0.0, curltmp = 0.0, counttmp = 0.0;
  int steps = 0;
  while (!walker_done(w)) {
    int s, a = walker_choose(w, &s);
    if (a < 0) break;
    int64_t vl = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    walker_step(w, a, s);
    int64_t vn = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    pint += trap(P->rho, s, P->del[a], P->acc[a][vn], P->acc[a][vl]);
    curltmp += P->curl[vn] / 10000;
    counttmp += 1.0;
    if (curltmp != 0) {
      P->pn[vn] += (pin + pint) * 1 / curltmp * counttmp;
      P->cnt[vn] += 1 / curltmp * counttmp;
    }
    steps++;
  }
  return steps;
}

/* bodyIntegralInnerSelect, kernel.cu:2280-2537.  With faithful != 0 the
 * per-axis asymmetries are kept (x treats curl == 0 as leaving the
 * low-error zone while y/z silently skip such voxels, kernel.cu:2394 vs
 * 2442,2488; y/z first-valid tests ignore the threshold, kernel.cu:2422,2468;
 * the z restart keeps counttmp/curltmp, kernel.cu:2497-2506).  With
 * faithful == 0 every axis uses valid(v) = curl != 0 && curl < thr.
 * PLAIN_MASK is this walk with curl := mask, thr := +inf, unit weights
 * and the symmetric rule. */
static int walk_select(const pass_ctx* P, walker* w, int unit_weight) {
  const int X = P->X, Y = P->Y;
  const double* cu = P->curl;
  const double thr = P->thr;
  const int faithful = P->faithful;
  int64_t vin = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
  double pint = 0.0, curltmp = 0.0, counttmp = 0.0, p0 = 0.0;
  int outthr = 0, p0set = 0;
  double ci = cu[vin];
  int entry_valid = faithful ? !(ci == 0 || ci > thr) : (ci != 0 && ci < thr);
  if (entry_valid) {
    p0 = P->p[vin];
    if (P->src) P->src[vin] = 1;
    p0set = 1;
  }
  int steps = 0;
  while (!walker_done(w)) {
    int s, a = walker_choose(w, &s);
    if (a < 0) break;
    int64_t vl = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    walker_step(w, a, s);
    int64_t vn = vidx(X, Y, w->c[0], w->c[1], w->c[2]);
    double cn = cu[vn];
    int valid = (cn < thr && cn != 0);
    if (!p0set) {
      int ok = (faithful && a != 0) ? (cn != 0) : valid;
      if (ok) {
        p0 = P->p[vn];
        if (P->src) P->src[vn] = 1;
        p0set = 1;
        pint = 0; counttmp = 0; curltmp = 0;
      }
    } else {
      if (!outthr && valid) {
        pint += trap(P->rho, s, P->del[a], P->acc[a][vn], P->acc[a][vl]);
        if (unit_weight) {
          P->pn[vn] += p0 + pint;
          P->cnt[vn] += 1.0;
        } else {
          curltmp += cn / 10000;
          counttmp += 1.0;
          P->pn[vn] += (p0 + pint) * 1 / curltmp * counttmp;
          P->cnt[vn] += 1 / curltmp * counttmp;
        }
      }
      int enter_high = (faithful && a != 0) ? (!outthr && cn >= thr)
                                            : ((!outthr && cn >= thr) || cn == 0);
      if (enter_high) outthr = 1;
      if (outthr && valid) {
        p0 = P->p[vn];
        if (P->src) P->src[vn] = 1;
        pint = 0;
        outthr = 0;
        if (!(faithful && a == 2)) { counttmp = 0; curltmp = 0; }
      }
    }
    steps++;
  }
  return steps;
}

void orc_pass(int X, int Y, int Z, int D, const float* k1, const float* k2, const float* k3,
              double ls, double dx, double dy, double dz, double rho, const double* ax,
              const double* ay, const double* az, const double* curl, double threshold,
              int method, int faithful, const double* p, double* pn, double* cnt,
              uint8_t* src_mark) {
  pass_ctx P;
  P.X = X; P.Y = Y; P.Z = Z;
  P.rho = rho;
  P.del[0] = dx; P.del[1] = dy; P.del[2] = dz;
  P.acc[0] = ax; P.acc[1] = ay; P.acc[2] = az;
  P.curl = curl;
  P.thr = (method == ORC_PLAIN_MASK) ? INFINITY : threshold;
  P.method = method;
  P.faithful = (method == ORC_PLAIN_MASK) ? 0 : faithful;
  P.p = p; P.pn = pn; P.cnt = cnt; P.src = src_mark;
  const int ng = orc_nogrid(X, Y, Z, ls);
  const int64_t npts = (int64_t)ng * ng;
  for (int ang = 0; ang < D; ang++) {
    orc_dir d;
    orc_make_dir(k1[ang], k2[ang], k3[ang], &d);
    for (int64_t pt = 0; pt < npts; pt++) {
      double o[3];
      int in[3], out[3];
      orc_origin(&d, ng, ls, pt, o);
      if (!orc_clip(X, Y, Z, o, d.k, in, out)) continue;
      walker w;
      walker_init(&w, X, Y, Z, &d, o, in, out);
      switch (method) {
        case ORC_PLAIN: walk_plain(&P, &w); break;
        case ORC_WEIGHTED: walk_weighted(&P, &w); break;
        case ORC_SELECT: walk_select(&P, &w, 0); break;
        case ORC_PLAIN_MASK: walk_select(&P, &w, 1); break;
        case ORC_SELECT_FIXEDBC:
          w.fixedbc = 1;
          walk_select(&P, &w, 0);
          break;
        case ORC_MINICURL: walk_minicurl(&P, &w); break;
        case ORC_WEIGHTED_FIXEDBC:
          w.fixedbc = 2;
          walk_weighted(&P, &w);
          break;
        case ORC_WEIGHTED_MINICURL: walk_weighted_minicurl(&P, &w); break;
        case ORC_PLAIN_TRILINEAR: walk_plain_trilinear(&P, &w); break;
        default: break;
      }
    }
  }
}

int64_t orc_enumerate_rays(int X, int Y, int Z, int D, const float* k1, const float* k2,
                           const float* k3, double ls, int32_t* rays, int64_t capacity) {
  const int ng = orc_nogrid(X, Y, Z, ls);
  const int64_t npts = (int64_t)ng * ng;
  int64_t n = 0;
  for (int ang = 0; ang < D; ang++) {
    orc_dir d;
    orc_make_dir(k1[ang], k2[ang], k3[ang], &d);
    for (int64_t pt = 0; pt < npts; pt++) {
      double o[3];
      int in[3], out[3];
      orc_origin(&d, ng, ls, pt, o);
      if (!orc_clip(X, Y, Z, o, d.k, in, out)) continue;
      if (rays && n < capacity) {
        int32_t* r = rays + 9 * n;
        r[0] = ang;
        r[1] = (int32_t)(pt & 0x7fffffff);
        r[2] = (int32_t)(pt >> 31);
        for (int a = 0; a < 3; a++) { r[3 + a] = in[a]; r[6 + a] = out[a]; }
      }
      n++;
    }
  }
  return n;
}

int orc_ray_path(int X, int Y, int Z, float k1, float k2, float k3, double ls, int64_t point,
                 int64_t* path, int capacity, int fixedbc) {
  orc_dir d;
  orc_make_dir(k1, k2, k3, &d);
  const int ng = orc_nogrid(X, Y, Z, ls);
  double o[3];
  int in[3], out[3];
  orc_origin(&d, ng, ls, point, o);
  if (!orc_clip(X, Y, Z, o, d.k, in, out)) return 0;
  walker w;
  walker_init(&w, X, Y, Z, &d, o, in, out);
  w.fixedbc = fixedbc;
  int n = 0;
  if (n < capacity) path[n] = vidx(X, Y, w.c[0], w.c[1], w.c[2]);
  n++;
  while (!walker_done(&w)) {
    int s, a = walker_choose(&w, &s);
    if (a < 0) break;
    walker_step(&w, a, s);
    if (n < capacity) path[n] = vidx(X, Y, w.c[0], w.c[1], w.c[2]);
    n++;
  }
  return n;
}

void orc_divide(int64_t V, double* p, double* pn, const double* cnt) {
  for (int64_t t = 0; t < V; t++) {
    if (cnt[t] > 0) {
      pn[t] = pn[t] / cnt[t];
      p[t] = pn[t];
      pn[t] = 0;
    }
  }
}

void orc_reconstruct(int X, int Y, int Z, int D, const float* k1, const float* k2,
                     const float* k3, double ls, double dx, double dy, double dz, double rho,
                     const double* ax, const double* ay, const double* az, const double* curl,
                     double threshold, int method, int faithful, int max_iter, double tol,
                     int seed_pn, double* p, double* cnt_out, orc_report* rep) {
  const int64_t V = (int64_t)X * Y * Z;
  double* pn = (double*)malloc(sizeof(double) * V);
  double* cnt = (double*)malloc(sizeof(double) * V);
  double* pold = (double*)malloc(sizeof(double) * V);
  uint8_t* src = (uint8_t*)calloc((size_t)V, 1);
  /* kernel.cu:8812-8813: pn_d starts as a copy of p */
  for (int64_t t = 0; t < V; t++) pn[t] = seed_pn ? p[t] : 0.0;
  int it = 0;
  double rel = 0.0;
  for (it = 1; it <= max_iter; it++) {
    memcpy(pold, p, sizeof(double) * V);
    memset(cnt, 0, sizeof(double) * V); /* kernel.cu:8860 */
    orc_pass(X, Y, Z, D, k1, k2, k3, ls, dx, dy, dz, rho, ax, ay, az, curl, threshold, method,
             faithful, p, pn, cnt, src);
    orc_divide(V, p, pn, cnt);
    double num = 0.0, den = 0.0;
    for (int64_t t = 0; t < V; t++) {
      if (!src[t]) continue;
      double df = p[t] - pold[t];
      num += df * df;
      den += p[t] * p[t];
    }
    num = sqrt(num);
    den = sqrt(den);
    rel = den > 0 ? num / den : (num > 0 ? INFINITY : 0.0);
    if (tol > 0 && rel < tol) break;
  }
  if (it > max_iter) it = max_iter;
  if (cnt_out) memcpy(cnt_out, cnt, sizeof(double) * V);
  if (rep) {
    rep->iterations = it;
    rep->rel_change = rel;
  }
  free(pn);
  free(cnt);
  free(pold);
  free(src);
}

/* calCurlofMaterialAccCPU, kernel.cu:8352-8379 */
void orc_curl(int X, int Y, int Z, double dx, double dy, double dz, const double* ax,
              const double* ay, const double* az, const double* mask, double* curl) {
  for (int k = 0; k < Z; k++)
    for (int j = 0; j < Y; j++)
      for (int i = 0; i < X; i++) {
        int i0 = i - 1 >= 0 ? i - 1 : i, j0 = j - 1 >= 0 ? j - 1 : j, k0 = k - 1 >= 0 ? k - 1 : k;
        int ie = i + 1 <= X - 1 ? i + 1 : i, je = j + 1 <= Y - 1 ? j + 1 : j,
            ke = k + 1 <= Z - 1 ? k + 1 : k;
        double cx = (az[vidx(X, Y, i, je, k)] - az[vidx(X, Y, i, j0, k)]) / (je - j0) / dy;
        cx += -(ay[vidx(X, Y, i, j, ke)] - ay[vidx(X, Y, i, j, k0)]) / (ke - k0) / dz;
        double cy = -(az[vidx(X, Y, ie, j, k)] - az[vidx(X, Y, i0, j, k)]) / (ie - i0) / dx;
        cy += (ax[vidx(X, Y, i, j, ke)] - ax[vidx(X, Y, i, j, k0)]) / (ke - k0) / dz;
        double cz = (ay[vidx(X, Y, ie, j, k)] - ay[vidx(X, Y, i0, j, k)]) / (ie - i0) / dx;
        cz += -(ax[vidx(X, Y, i, je, k)] - ax[vidx(X, Y, i, j0, k)]) / (je - j0) / dy;
        curl[vidx(X, Y, i, j, k)] = sqrt(cx * cx + cy * cy + cz * cz) * mask[vidx(X, Y, i, j, k)];
      }
}

/* calRHS, calRHS.cpp:7-309.  Interior (calRHS.cpp:224-231): per axis
 * (-d*a[i]-d*a[i-1])*0.5 + (d*a[i]+d*a[i+1])*0.5.  A boundary voxel keeps
 * only the inward half-term of each axis whose face it lies on
 * (low face: (d*a[i]+d*a[i+1])*0.5, high face: (-d*a[i]-d*a[i-1])*0.5) and
 * no term for the other axes (calRHS.cpp:19-222).  Everything * density.
 * 1-voxel-thick axes are not covered by the reference and get no term. */
void orc_calrhs(int X, int Y, int Z, double dx, double dy, double dz, double rho,
                const double* ax, const double* ay, const double* az, double* rhs) {
  const double del[3] = {dx, dy, dz};
  const double* acc[3] = {ax, ay, az};
  const int size[3] = {X, Y, Z};
  for (int k = 0; k < Z; k++)
    for (int j = 0; j < Y; j++)
      for (int i = 0; i < X; i++) {
        const int c[3] = {i, j, k};
        int64_t v = vidx(X, Y, i, j, k);
        int onb = 0;
        for (int a = 0; a < 3; a++) onb |= (c[a] == 0 || c[a] == size[a] - 1);
        double r = 0.0;
        for (int a = 0; a < 3; a++) {
          int64_t st = a == 0 ? 1 : (a == 1 ? X : (int64_t)X * Y);
          int lo = c[a] == 0, hi = c[a] == size[a] - 1;
          if (size[a] < 2) continue;
          if (!onb) {
            r += (-del[a] * acc[a][v] - del[a] * acc[a][v - st]) * 0.5;
            r += (del[a] * acc[a][v] + del[a] * acc[a][v + st]) * 0.5;
          } else if (lo) {
            r += (del[a] * acc[a][v] + del[a] * acc[a][v + st]) * 0.5;
          } else if (hi) {
            r += (-del[a] * acc[a][v] - del[a] * acc[a][v - st]) * 0.5;
          }
        }
        rhs[v] = r * rho;
      }
}

/* Neighbour sum used by sor3D: in-box face neighbours in the order
 * x-, x+, y-, y+, z-, z+ (sor3D.cpp:36,62,...), and their count. */
static inline double nbsum(int X, int Y, int Z, const double* p, int i, int j, int k, int* cntn) {
  double s = 0.0;
  int n = 0;
  int64_t v = vidx(X, Y, i, j, k);
  int64_t sy = X, sz = (int64_t)X * Y;
  if (i > 0) { s = n ? s + p[v - 1] : p[v - 1]; n++; }
  if (i < X - 1) { s = n ? s + p[v + 1] : p[v + 1]; n++; }
  if (j > 0) { s = n ? s + p[v - sy] : p[v - sy]; n++; }
  if (j < Y - 1) { s = n ? s + p[v + sy] : p[v + sy]; n++; }
  if (k > 0) { s = n ? s + p[v - sz] : p[v - sz]; n++; }
  if (k < Z - 1) { s = n ? s + p[v + sz] : p[v + sz]; n++; }
  *cntn = n;
  return s;
}

double orc_residual(int X, int Y, int Z, const double* p, const double* rhs) {
  double r = 0.0;
  for (int k = 1; k < Z - 1; k++)
    for (int j = 1; j < Y - 1; j++)
      for (int i = 1; i < X - 1; i++) {
        int n;
        int64_t v = vidx(X, Y, i, j, k);
        double t = nbsum(X, Y, Z, p, i, j, k, &n);
        t += -6.0 * p[v];
        t += rhs[v];
        r += fabs(t);
      }
  return r / X / Y / Z;
}

/* One damped-Jacobi sweep over the BCScheme range (sor3D.cpp:27-740) and the
 * float-accumulated interior residue (sor3D.cpp:741-781). */
static float sor_sweep(int X, int Y, int Z, const int rng[6], const double* p, double* pn,
                       const double* rhs) {
  for (int k = rng[4]; k < rng[5]; k++)
    for (int j = rng[2]; j < rng[3]; j++)
      for (int i = rng[0]; i < rng[1]; i++) {
        int n;
        int64_t v = vidx(X, Y, i, j, k);
        double s = nbsum(X, Y, Z, p, i, j, k, &n);
        s += rhs[v];
        s = s / n;
        pn[v] = p[v] + 0.8 * (s - p[v]);
      }
  float residue = 0.0f;
  for (int k = 1; k < Z - 1; k++)
    for (int j = 1; j < Y - 1; j++)
      for (int i = 1; i < X - 1; i++) {
        int n;
        int64_t v = vidx(X, Y, i, j, k);
        double t = nbsum(X, Y, Z, pn, i, j, k, &n);
        t += -6.0 * pn[v];
        t += rhs[v];
        residue += fabs(t);
      }
  residue = residue / X / Y / Z;
  return residue;
}

int orc_sor3d(int X, int Y, int Z, int noitr, const int* bc, double* p, double* pn,
              const double* rhs, double eps, double* residue_out) {
  const int rng[6] = {bc[0], X - bc[1], bc[2], Y - bc[3], bc[4], Z - bc[5]};
  int iteration = 0, sweeps = 0;
  float residue = 0.0f;
  /* sor3D.cpp:21-1587: each pass of the while loop does a forward and a
   * reversed sweep, each followed by the residue test; the sweep order does
   * not matter for Jacobi (pn reads only p). */
  while (iteration < noitr) {
    for (int half = 0; half < 2; half++) {
      residue = sor_sweep(X, Y, Z, rng, p, pn, rhs);
      sweeps++;
      if (residue < eps) {
        memcpy(p, pn, sizeof(double) * (size_t)X * Y * Z);
        if (residue_out) *residue_out = residue;
        return sweeps;
      }
      for (int k = rng[4]; k < rng[5]; k++)
        for (int j = rng[2]; j < rng[3]; j++)
          for (int i = rng[0]; i < rng[1]; i++) {
            int64_t v = vidx(X, Y, i, j, k);
            p[v] = pn[v];
            pn[v] = 0;
          }
      iteration++;
    }
  }
  if (residue_out) *residue_out = residue;
  return sweeps;
}

void orc_sor_redblack(int X, int Y, int Z, int iters, double omega, double* p, const double* rhs) {
  for (int it = 0; it < iters; it++)
    for (int color = 0; color < 2; color++)
      for (int k = 0; k < Z; k++)
        for (int j = 0; j < Y; j++)
          for (int i = 0; i < X; i++) {
            if (((i + j + k) & 1) != color) continue;
            int n;
            int64_t v = vidx(X, Y, i, j, k);
            double s = nbsum(X, Y, Z, p, i, j, k, &n);
            s += rhs[v];
            s = s / n;
            p[v] = p[v] + omega * (s - p[v]);
          }
}

/* Threshold from the curl histogram (thredholdHistMaterialAccCPU,
 * kernel.cu:8380-8418): min / max of the values (seeded with 1e4 / -1e4 as
 * there), 10000 bins over [min, max] in FP32, then from the top bin down the
 * first bin where the running count (FP32) reaches count * percentage gives
 * bin / N * (max - min) + min; the maximum itself when none does.  The
 * reference walks the bins with j++ (kernel.cu:8406) and puts the maximum in
 * bin N, both out of bounds; this restates the intended walk (j--) with the
 * maximum in the last bin. */
/* initialIntegration, initialIntegration.cpp:9-408.  p[0] = pn[0] = pref
 * (lines 11-12); then eight sweeps, one per octant around the centre voxel
 * (X/2, Y/2, Z/2), in the file's order (No.1 .. No.8: x, then y, then z
 * toward the low side first), each with k outermost, then j, then i, every
 * index running from the centre toward its face.  A voxel gets the mean of
 * the one-step trapezoid predictions from its neighbours toward the centre
 * along the axes on which it is off the centre, added in the order x, y, z:
 *   neighbour on the high side:  pn[nb] + 0.5 * (a[v] + a[nb]) * delta * rho
 *   neighbour on the low side:   pn[nb] - 0.5 * (a[v] + a[nb]) * delta * rho
 * (the first assigned, the others added with +=, then / 2.0 or / 3.0), and
 * p[v] = pn[v].  The centre keeps the caller's pn.  Voxels on a centre plane
 * lie in two or four octants and are recomputed with the same expression. */
void orc_initial_integration(int X, int Y, int Z, double dx, double dy, double dz, double rho,
                             const double* ax, const double* ay, const double* az, double pref,
                             double* p, double* pn) {
  p[0] = pref;
  pn[0] = pref;
  const int ctr[3] = {X / 2, Y / 2, Z / 2};
  const int size[3] = {X, Y, Z};
  const int64_t stride[3] = {1, X, (int64_t)X * Y};
  const double del[3] = {dx, dy, dz};
  const double* acc[3] = {ax, ay, az};
  for (int oc = 0; oc < 8; oc++) {
    const int sg[3] = {(oc & 1) ? 1 : -1, (oc & 2) ? 1 : -1, (oc & 4) ? 1 : -1};
    for (int k = ctr[2]; k >= 0 && k < size[2]; k += sg[2])
      for (int j = ctr[1]; j >= 0 && j < size[1]; j += sg[1])
        for (int i = ctr[0]; i >= 0 && i < size[0]; i += sg[0]) {
          const int c[3] = {i, j, k};
          const int64_t v = vidx(X, Y, i, j, k);
          int n = 0;
          double val = 0.0;
          for (int a = 0; a < 3; a++) {
            if (c[a] == ctr[a]) continue;
            const int64_t nb = v - (int64_t)sg[a] * stride[a];
            const double t = 0.5 * (acc[a][v] + acc[a][nb]) * del[a] * rho;
            const double pr = sg[a] < 0 ? pn[nb] + t : pn[nb] - t;
            val = n == 0 ? pr : val + pr;
            n++;
          }
          if (n == 2) val = val / 2.0;
          if (n == 3) val = val / 3.0;
          if (n > 0) pn[v] = val;
          p[v] = pn[v];
        }
  }
}

float orc_threshold_hist(int X, int Y, int Z, const float* curl, float percentage) {
  const int64_t V = (int64_t)X * Y * Z;
  float minv = 1e4f, maxv = -1e4f;
  for (int64_t i = 0; i < V; i++) {
    if (minv > curl[i]) minv = curl[i];
    if (maxv < curl[i]) maxv = curl[i];
  }
  enum { N = 10000 };
  static int hist[N];
  memset(hist, 0, sizeof(hist));
  for (int64_t i = 0; i < V; i++) {
    const float q = (curl[i] - minv) / (maxv - minv) * N;
    /* (int)q as kernel.cu:8401; the maximum (q == N) and NaN quotients
     * (max == min, NaN values) are kept inside the table */
    hist[q >= N - 1 ? N - 1 : (q > 0 ? (int)q : 0)]++;
  }
  float totnum = 0;
  const float need = (float)(int)V * percentage;
  for (int j = N - 1; j >= 0; j--) {
    totnum += hist[j];
    if (totnum >= need) return (float)j / N * (maxv - minv) + minv;
  }
  return maxv;
}

/* ---- boundary-pair (PINT) pipeline in sparse per-ray form ----------------
 * omni3dparallellinesEqualSpacing (kernel.cu:3414-3489): per accepted ray,
 * bodyIntegral (kernel.cu:1353-1467, the plain staircase and trapezoid of
 * bodyIntegralInner without the scatter) from the entry to the exit voxel,
 * added into pint[in, out] with pcount[in, out]++; devidecount
 * (kernel.cu:3896-3909) makes each pair's value the mean over its rays.
 * BCiteration (kernel.cu:5572-5712): per exit node, the mean over the pairs
 * with a non-zero value of p[in] + pint[in, out], then p[out] = that mean.
 * Here the pairs are a sorted list instead of the n x n matrix, the
 * iteration is a Jacobi sweep (every new value from the previous p; the
 * reference has only __syncthreads() between its iterations, an inter-block
 * race), and an exit node whose pairs are all zero keeps its p (the
 * reference divides 0 by 0 there). */
typedef struct {
  int64_t out, in, ray;
  double v;
} bp_rec;

static int bp_cmp(const void* a, const void* b) {
  const bp_rec* x = (const bp_rec*)a;
  const bp_rec* y = (const bp_rec*)b;
  if (x->out != y->out) return x->out < y->out ? -1 : 1;
  if (x->in != y->in) return x->in < y->in ? -1 : 1;
  return x->ray < y->ray ? -1 : (x->ray > y->ray);
}

int64_t orc_boundary_pairs(int X, int Y, int Z, int D, const float* k1, const float* k2,
                           const float* k3, double ls, double dx, double dy, double dz,
                           double rho, const double* ax, const double* ay, const double* az,
                           int iters, double* p) {
  const int ng = orc_nogrid(X, Y, Z, ls);
  const int64_t npts = (int64_t)ng * ng;
  const double del[3] = {dx, dy, dz};
  const double* acc[3] = {ax, ay, az};
  int64_t cap = 1024, n = 0;
  bp_rec* r = (bp_rec*)malloc(sizeof(bp_rec) * cap);
  for (int ang = 0; ang < D; ang++) {
    orc_dir d;
    orc_make_dir(k1[ang], k2[ang], k3[ang], &d);
    for (int64_t pt = 0; pt < npts; pt++) {
      double o[3];
      int in[3], out[3];
      orc_origin(&d, ng, ls, pt, o);
      if (!orc_clip(X, Y, Z, o, d.k, in, out)) continue;
      walker w;
      walker_init(&w, X, Y, Z, &d, o, in, out);
      double pint = 0.0;
      while (!walker_done(&w)) {
        int s, a = walker_choose(&w, &s);
        if (a < 0) break;
        const int64_t vl = vidx(X, Y, w.c[0], w.c[1], w.c[2]);
        walker_step(&w, a, s);
        const int64_t vn = vidx(X, Y, w.c[0], w.c[1], w.c[2]);
        pint += trap(rho, s, del[a], acc[a][vn], acc[a][vl]);
      }
      if (n == cap) {
        cap *= 2;
        r = (bp_rec*)realloc(r, sizeof(bp_rec) * cap);
      }
      r[n].out = vidx(X, Y, out[0], out[1], out[2]);
      r[n].in = vidx(X, Y, in[0], in[1], in[2]);
      r[n].ray = n;
      r[n].v = pint;
      n++;
    }
  }
  qsort(r, (size_t)n, sizeof(bp_rec), bp_cmp);
  /* pairs: (out, in, mean) in (out, in) order */
  int64_t np = 0;
  for (int64_t i = 0; i < n;) {
    int64_t j = i;
    double sum = 0.0;
    while (j < n && r[j].out == r[i].out && r[j].in == r[i].in) sum += r[j++].v;
    r[np].out = r[i].out;
    r[np].in = r[i].in;
    r[np].v = sum / (double)(j - i);
    np++;
    i = j;
  }
  const int64_t V = (int64_t)X * Y * Z;
  double* pn = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  for (int it = 0; it < iters; it++) {
    int64_t nout = 0;
    for (int64_t i = 0; i < np;) {
      int64_t j = i;
      double sum = 0.0;
      int c = 0;
      for (; j < np && r[j].out == r[i].out; j++) {
        if (r[j].v != 0.0) {
          sum += p[r[j].in] + r[j].v;
          c++;
        }
      }
      pn[nout++] = c ? sum / (double)c : p[r[i].out];
      i = j;
    }
    nout = 0;
    for (int64_t i = 0; i < np;) {
      int64_t j = i;
      while (j < np && r[j].out == r[i].out) j++;
      p[r[i].out] = pn[nout++];
      i = j;
    }
  }
  (void)V;
  free(pn);
  free(r);
  return np;
}
