// Process-level drop-in for Omni3D_8.0.exe (kernel.cu:8467-8983): the
// 22-line Parameter_Omni3D.dat, the Tecplot acceleration files of a folder,
// and the PressureOmni3D_%05d.dat / CurlofMaterialAcc_Sample.dat / log.dat
// outputs, with the reference's method (Select, faithful branches, curl
// threshold of line 15), Bernoulli boundary seed and fixed iteration count.
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <dirent.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "driver.hpp"
#include "host.hpp"
#include "launch.hpp"

using namespace omni3d;

namespace {

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r' || s[b] == '\n')) b++;
  while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r' || s[e - 1] == '\n')) e--;
  return s.substr(b, e - b);
}

// Windows separators to '/', relative paths against the parameter file's folder.
std::string resolve(const std::string& base_dir, std::string p) {
  std::replace(p.begin(), p.end(), '\\', '/');
  if (p.size() >= 2 && p[0] == '.' && p[1] == '/') p = p.substr(2);
  if (!p.empty() && p[0] == '/') return p;
  return base_dir.empty() ? p : base_dir + "/" + p;
}

bool is_dir(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// Regular files under dir (recursively, as "dir /B /S" lists them,
// kernel.cu:90-110), sorted by path.
void list_files(const std::string& dir, std::vector<std::string>* out) {
  DIR* d = opendir(dir.c_str());
  if (!d) return;
  while (dirent* e = readdir(d)) {
    const std::string name = e->d_name;
    if (name == "." || name == "..") continue;
    const std::string full = dir + (dir.back() == '/' ? "" : "/") + name;
    struct stat st;
    if (stat(full.c_str(), &st) != 0) continue;
    if (S_ISDIR(st.st_mode)) list_files(full, out);
    else if (S_ISREG(st.st_mode)) out->push_back(full);
  }
  closedir(d);
  std::sort(out->begin(), out->end());
}

// The last `n` numbers of a line (the reference parses rows from the right).
bool last_numbers(const std::string& line, int n, double* v) {
  std::vector<double> all;
  const char* s = line.c_str();
  char* end;
  for (;;) {
    while (*s == ' ' || *s == '\t' || *s == ',' || *s == '\r') s++;
    if (!*s) break;
    const double x = std::strtod(s, &end);
    if (end == s) return false;
    all.push_back(x);
    s = end;
  }
  if ((int)all.size() < n) return false;
  for (int i = 0; i < n; i++) v[i] = all[all.size() - n + i];
  return true;
}

struct Snapshot {
  std::vector<double> x, y, z, u, v, w, ax, ay, az, mask;
};

// Tecplot POINT file with 3 header lines and, per line (i fastest), the
// columns x y z u v w DuDt DvDt DwDt mask; voxels outside the ROI are skipped
// (kernel.cu:8717-8797).
bool read_acceleration(const std::string& path, const omni3d_parameter_file& pf, Snapshot* s,
                       std::string* err) {
  std::ifstream f(path);
  if (!f) {
    *err = "failed to open";
    return false;
  }
  const int X = pf.roi[1] - pf.roi[0] + 1, Y = pf.roi[3] - pf.roi[2] + 1,
            Z = pf.roi[5] - pf.roi[4] + 1;
  const size_t V = (size_t)X * Y * Z;
  for (auto* a : {&s->x, &s->y, &s->z, &s->u, &s->v, &s->w, &s->ax, &s->ay, &s->az, &s->mask})
    a->assign(V, 0.0);
  std::string line;
  for (int h = 0; h < 3; h++) std::getline(f, line);
  for (int k = 0; k < pf.nz; k++)
    for (int j = 0; j < pf.ny; j++)
      for (int i = 0; i < pf.nx; i++) {
        if (!std::getline(f, line)) {
          *err = "file ends before " + std::to_string((long long)pf.nx * pf.ny * pf.nz) + " data lines";
          return false;
        }
        if (i < pf.roi[0] || i > pf.roi[1] || j < pf.roi[2] || j > pf.roi[3] || k < pf.roi[4] ||
            k > pf.roi[5])
          continue;
        double c[10];
        if (!last_numbers(line, 10, c)) {
          *err = "data line with fewer than 10 numbers";
          return false;
        }
        const size_t ind = (size_t)(i - pf.roi[0]) + (size_t)(j - pf.roi[2]) * X +
                           (size_t)(k - pf.roi[4]) * X * Y;
        s->x[ind] = c[0];
        s->y[ind] = c[1];
        s->z[ind] = c[2];
        s->u[ind] = c[3];
        s->v[ind] = c[4];
        s->w[ind] = c[5];
        s->ax[ind] = c[6];
        s->ay[ind] = c[7];
        s->az[ind] = c[8];
        s->mask[ind] = c[9];
      }
  return true;
}

bool write_tecplot(const std::string& path, const char* title, const char* vars, int X, int Y,
                   int Z, const std::vector<const std::vector<double>*>& cols) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "TITLE = \"%s\"\n", title);
  std::fprintf(f, "VARIABLES = %s\n", vars);
  std::fprintf(f, "ZONE I=%i, J=%i, K=%i,F=POINT\n", X, Y, Z);
  const size_t V = (size_t)X * Y * Z;
  for (size_t v = 0; v < V; v++) {
    for (size_t c = 0; c < cols.size(); c++)
      std::fprintf(f, c + 1 < cols.size() ? "%15.9f " : "%15.9f\n", (*cols[c])[v]);
  }
  return std::fclose(f) == 0;
}

}  // namespace

extern "C" {

int omni3d_read_parameter_file(const char* path, omni3d_parameter_file* pf) {
  if (!path || !pf) return fail(nullptr, OMNI3D_ERR_ARG, "null argument");
  std::ifstream f(path);
  if (!f) return fail(nullptr, OMNI3D_ERR_IO, std::string("Parameter input file: \"") + path + "\" open error");
  std::vector<std::string> L;
  std::string line;
  while ((int)L.size() < 22 && std::getline(f, line)) L.push_back(trim(line));
  while (L.size() < 22) L.push_back("");
  std::memset(pf, 0, sizeof(*pf));
  pf->nx = std::atoi(L[0].c_str());
  pf->ny = std::atoi(L[1].c_str());
  pf->nz = std::atoi(L[2].c_str());
  pf->dx = std::atof(L[3].c_str());
  pf->dy = std::atof(L[4].c_str());
  pf->dz = std::atof(L[5].c_str());
  pf->density = std::atof(L[6].c_str());
  pf->scale = std::atoi(L[7].c_str());
  pf->line_spacing = std::atof(L[8].c_str());
  pf->n_angles = std::atoi(L[9].c_str());
  std::snprintf(pf->grid_file, sizeof(pf->grid_file), "%s", L[10].c_str());
  std::snprintf(pf->acc_dir, sizeof(pf->acc_dir), "%s", L[11].c_str());
  std::snprintf(pf->pressure_dir, sizeof(pf->pressure_dir), "%s", L[12].c_str());
  pf->n_iterations = std::atoi(L[13].c_str());
  pf->threshold = std::atof(L[14].c_str());
  // hidden lines 16-22: pref and the region of interest (blank = defaults)
  pf->pref = L[15].empty() ? 0.0 : std::atof(L[15].c_str());
  const int dflt[6] = {0, pf->nx - 1, 0, pf->ny - 1, 0, pf->nz - 1};
  for (int i = 0; i < 6; i++) pf->roi[i] = L[16 + i].empty() ? dflt[i] : std::atoi(L[16 + i].c_str());
  if (pf->nx < 2 || pf->ny < 2 || pf->nz < 2) return fail(nullptr, OMNI3D_ERR_IO, "grid size lines 1-3 invalid");
  for (int a = 0; a < 3; a++) {
    const int n = a == 0 ? pf->nx : (a == 1 ? pf->ny : pf->nz);
    if (pf->roi[2 * a] < 0 || pf->roi[2 * a + 1] >= n || pf->roi[2 * a + 1] - pf->roi[2 * a] < 1)
      return fail(nullptr, OMNI3D_ERR_IO, "region of interest outside the grid");
  }
  if (pf->n_angles < 1) return fail(nullptr, OMNI3D_ERR_IO, "line 10: number of directions < 1");
  if (pf->n_iterations < 1) return fail(nullptr, OMNI3D_ERR_IO, "line 14: number of iterations < 1");
  return OMNI3D_OK;
}

int omni3d_run_parameter_file_ex(const char* path, int precision) {
  omni3d_parameter_file pf;
  int rc = omni3d_read_parameter_file(path, &pf);
  if (rc) return -rc;
  std::string base = path;
  const size_t slash = base.find_last_of('/');
  base = slash == std::string::npos ? "" : base.substr(0, slash);
  const std::string accdir = resolve(base, pf.acc_dir), pdir = resolve(base, pf.pressure_dir);
  const std::string logpath = base.empty() ? "log.dat" : base + "/log.dat";
  const std::string curlpath =
      base.empty() ? "CurlofMaterialAcc_Sample.dat" : base + "/CurlofMaterialAcc_Sample.dat";
  if (!is_dir(accdir)) return -fail(nullptr, OMNI3D_ERR_IO, "Acceleration folder does not exist: " + accdir);
  if (!is_dir(pdir) && mkdir(pdir.c_str(), 0755) != 0 && errno != EEXIST)
    return -fail(nullptr, OMNI3D_ERR_IO, "cannot create pressure folder " + pdir);
  std::vector<std::string> files;
  list_files(accdir, &files);
  omni3d_params P;
  omni3d_default_params(&P);
  P.nx = pf.roi[1] - pf.roi[0] + 1;
  P.ny = pf.roi[3] - pf.roi[2] + 1;
  P.nz = pf.roi[5] - pf.roi[4] + 1;
  P.dx = pf.dx;
  P.dy = pf.dy;
  P.dz = pf.dz;
  P.density = pf.density;
  P.line_spacing = pf.line_spacing;
  P.method = OMNI3D_METHOD_SELECT;  // the shipped driver runs ESInnerSelect (kernel.cu:8845)
  P.threshold = pf.threshold;
  P.faithful_select = 1;
  P.precision = precision;
  P.solver = OMNI3D_SOLVER_NAIVE;
  P.max_iterations = pf.n_iterations;
  P.tolerance = 0.0;
  P.seed_pn_with_p = 1;  // pn_d starts as a copy of p (kernel.cu:8813)
  omni3d_ctx* ctx = nullptr;
  rc = omni3d_create(&P, &ctx);
  if (rc) return -rc;
  struct Guard {
    omni3d_ctx* c;
    ~Guard() { omni3d_destroy(c); }
  } guard{ctx};
  rc = omni3d_load_directions_file(ctx, resolve(base, pf.grid_file).c_str(), pf.n_angles);
  if (rc) return -fail(nullptr, rc, ctx->err);
  std::ofstream log(logpath);
  log << "Processing Starts" << std::endl;
  const int X = P.nx, Y = P.ny, Z = P.nz;
  const size_t V = (size_t)X * Y * Z;
  std::vector<double> p(V, 0.0);
  int done = 0;
  for (size_t fi = 0; fi < files.size(); fi++) {
    const std::string& fpath = files[fi];
    const std::string name = fpath.substr(fpath.find_last_of('/') + 1);
    log << name;
    Snapshot s;
    std::string err;
    if (!read_acceleration(fpath, pf, &s, &err)) {
      log << " " << err << std::endl;
      continue;
    }
    rc = omni3d_set_field(ctx, s.ax.data(), s.ay.data(), s.az.data(), s.mask.data(), OMNI3D_F64,
                          OMNI3D_MEM_HOST);
    if (rc) return -fail(nullptr, rc, ctx->err);
    // the previous file's result stays the initial pressure (the host array
    // carries over, kernel.cu:8863 -> 8812), with the Bernoulli pressure
    // written on the j = 0 face (kernel.cu:8806-8811); zero before file 0
    std::vector<double> cnt(V, 0.0);
    for (int i = 0; i < X; i++)
      for (int k = 0; k < Z; k++) {
        const size_t ind = (size_t)i + (size_t)k * X * Y;
        p[ind] = pf.pref - 0.5 * pf.density * (s.u[ind] * s.u[ind] + s.v[ind] * s.v[ind] + s.w[ind] * s.w[ind]);
      }
    rc = omni3d_reconstruct(ctx, p.data(), cnt.data(), OMNI3D_F64, OMNI3D_MEM_HOST, nullptr);
    if (rc) return -fail(nullptr, rc, ctx->err);
    char fname[64];
    std::snprintf(fname, sizeof(fname), "PressureOmni3D_%05d.dat", (int)fi);
    if (!write_tecplot(pdir + "/" + fname, "Pressure Integrated From GPU Based Omni 3D Method",
                       "\"X\",\"Y\",\"Z\",\"P\",\"Count\"", X, Y, Z, {&s.x, &s.y, &s.z, &p, &cnt}))
      return -fail(nullptr, OMNI3D_ERR_IO, std::string("cannot write ") + fname);
    if (fi == 0) {
      std::vector<double> curl(V);
      rc = omni3d_field_curl(ctx, curl.data(), OMNI3D_F64, OMNI3D_MEM_HOST);
      if (rc) return -fail(nullptr, rc, ctx->err);
      if (!write_tecplot(curlpath, "Curl of Material Acceleration Multiplied by the Mask",
                         "\"X\",\"Y\",\"Z\",\"Curl of Acceleration\"", X, Y, Z,
                         {&s.x, &s.y, &s.z, &curl}))
        return -fail(nullptr, OMNI3D_ERR_IO, "cannot write CurlofMaterialAcc_Sample.dat");
    }
    done++;
  }
  return done;
}

int omni3d_run_parameter_file(const char* path) {
  return omni3d_run_parameter_file_ex(path, OMNI3D_FP64);
}

}  // extern "C"
