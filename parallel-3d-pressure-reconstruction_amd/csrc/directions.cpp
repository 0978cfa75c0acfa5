// Direction sets: the project's icosahedral generator and the reader of the
// reference's GridIcosahedronDivideN.dat files (kernel.cu:8622-8648).
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace omni3d {

namespace {

struct V3 {
  double x, y, z;
};

V3 normalized(V3 a) {
  const double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x / n, a.y / n, a.z / n};
}

}  // namespace

// Pole-aligned icosahedron (vertices at +-z and two rings of five at
// z = +-1/sqrt5, the lower ring rotated by 36 degrees), subdivided divide-1
// times by normalised edge midpoints.  The set is centrally symmetric; each
// antipodal partner is stored as the exact negation of the first of the
// pair, components below 1e-12 are set to exactly 0 (the reference's clipping
// treats exact zeros and |k| < 1e-7 differently, kernel.cu:672-992), and the
// rows are sorted by (k1, k2, k3).
std::vector<double> icosahedral_directions(int divide) {
  std::vector<V3> vs;
  const double h = 1.0 / std::sqrt(5.0), r = 2.0 / std::sqrt(5.0);
  const double pi = 3.14159265358979323846;
  vs.push_back({0.0, 0.0, 1.0});
  for (int i = 0; i < 5; i++) {
    const double a = 2.0 * pi * i / 5.0;
    vs.push_back({r * std::cos(a), r * std::sin(a), h});
  }
  for (int i = 0; i < 5; i++) {
    const double a = 2.0 * pi * i / 5.0 + pi / 5.0;
    vs.push_back({r * std::cos(a), r * std::sin(a), -h});
  }
  vs.push_back({0.0, 0.0, -1.0});
  std::vector<std::array<int, 3>> faces;
  for (int i = 0; i < 5; i++) {
    const int u0 = 1 + i, u1 = 1 + (i + 1) % 5, l0 = 6 + i, l1 = 6 + (i + 1) % 5;
    faces.push_back({0, u0, u1});
    faces.push_back({11, l1, l0});
    faces.push_back({u0, l0, u1});
    faces.push_back({u1, l0, l1});
  }
  for (int level = 1; level < divide; level++) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      const auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      const V3 m = normalized({vs[a].x + vs[b].x, vs[a].y + vs[b].y, vs[a].z + vs[b].z});
      vs.push_back(m);
      const int id = (int)vs.size() - 1;
      mid[key] = id;
      return id;
    };
    std::vector<std::array<int, 3>> nf;
    nf.reserve(faces.size() * 4);
    for (const auto& f : faces) {
      const int a = midpoint(f[0], f[1]), b = midpoint(f[1], f[2]), c = midpoint(f[2], f[0]);
      nf.push_back({f[0], a, c});
      nf.push_back({f[1], b, a});
      nf.push_back({f[2], c, b});
      nf.push_back({a, b, c});
    }
    faces.swap(nf);
  }
  // exact zeros, then exact antipodal partners
  for (auto& v : vs) {
    if (std::fabs(v.x) < 1e-12) v.x = 0.0;
    if (std::fabs(v.y) < 1e-12) v.y = 0.0;
    if (std::fabs(v.z) < 1e-12) v.z = 0.0;
  }
  const size_t n = vs.size();
  std::vector<char> fixed(n, 0);
  for (size_t i = 0; i < n; i++) {
    if (fixed[i]) continue;
    fixed[i] = 1;
    size_t best = n;
    double bd = 1e-6;
    for (size_t j = 0; j < n; j++) {
      if (fixed[j]) continue;
      const double d = std::fabs(vs[i].x + vs[j].x) + std::fabs(vs[i].y + vs[j].y) +
                       std::fabs(vs[i].z + vs[j].z);
      if (d < bd) {
        bd = d;
        best = j;
      }
    }
    if (best < n) {
      vs[best] = {-vs[i].x, -vs[i].y, -vs[i].z};
      fixed[best] = 1;
    }
  }
  std::sort(vs.begin(), vs.end(), [](const V3& a, const V3& b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    return a.z < b.z;
  });
  std::vector<double> out;
  out.reserve(3 * n);
  for (const auto& v : vs) {
    out.push_back(v.x == 0.0 ? 0.0 : v.x);
    out.push_back(v.y == 0.0 ? 0.0 : v.y);
    out.push_back(v.z == 0.0 ? 0.0 : v.z);
  }
  return out;
}

// Reader of the reference's direction files: `count` rows of "k1 k2 k3"
// (kernel.cu:8622-8648 parses each line right to left; whitespace-separated
// tokens give the same three numbers), followed by a line holding the count.
bool read_direction_file(const char* path, int count, std::vector<float>* k, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    *err = std::string("direction file open error: ") + path;
    return false;
  }
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(f, line)) {
    std::replace(line.begin(), line.end(), '\r', ' ');
    std::istringstream is(line);
    std::vector<double> vals;
    double x;
    while (is >> x) vals.push_back(x);
    if (!vals.empty()) rows.push_back(vals);
  }
  if (count <= 0) {
    if (rows.empty() || rows.back().size() != 1) {
      *err = "direction file has no trailing count line";
      return false;
    }
    count = (int)rows.back()[0];
  }
  if (count <= 0 || (size_t)count > rows.size()) {
    *err = "direction file holds fewer rows than requested";
    return false;
  }
  k->clear();
  // the reference takes the last three numbers of a row (it parses from the
  // right, kernel.cu:8626-8647)
  for (int i = 0; i < count; i++) {
    const size_t n = rows[i].size();
    if (n < 3) {
      *err = "direction file row " + std::to_string(i) + " has fewer than 3 values";
      return false;
    }
    for (size_t c = n - 3; c < n; c++) k->push_back((float)rows[i][c]);
  }
  return true;
}

// Per-direction constants in FP64: normalise; theta = acos(k3),
// phi = asin(k2 / sin theta) (+pi flip when k1 / sin theta < 0),
// kernel.cu:6441-6457; the asin argument is clamped to [-1, 1].
DirConst make_dir_const(float k1, float k2, float k3) {
  DirConst d;
  const double a = k1, b = k2, c = k3;
  const double n = std::sqrt(a * a + b * b + c * c);
  const double k[3] = {a / n, b / n, c / n};
  const double theta = std::acos(k[2]);
  const double st = std::sin(theta), ct = std::cos(theta);
  double phi = 0.0;
  if (st != 0.0) {
    double s = k[1] / st;
    s = s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s);
    phi = std::asin(s);
    if (k[0] / st < 0.0) phi = -phi + 3.1415926535897932384626433832795;
  }
  const double sf = std::sin(phi), cf = std::cos(phi);
  for (int i = 0; i < 3; i++) d.k[i] = k[i];
  d.e1[0] = ct * cf;
  d.e1[1] = ct * sf;
  d.e1[2] = -st;
  d.e2[0] = -sf;
  d.e2[1] = cf;
  d.e2[2] = 0.0;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) d.m[r][s] = (r == s ? 1.0 : 0.0) - k[r] * k[s];
  return d;
}

}  // namespace omni3d
