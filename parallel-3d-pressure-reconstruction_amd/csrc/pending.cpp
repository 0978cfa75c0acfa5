// Entry points whose implementation is not in this build yet: they fail
// loudly with OMNI3D_ERR_STATE.
#include "host.hpp"

extern "C" {
static int pending(const char* what) {
  (void)what;
  return OMNI3D_ERR_STATE;
}
int omni3d_es_inner(int, int, int, int, float, const float*, const float*, const float*, float,
                    float, float, float, const float*, const float*, const float*, const float*,
                    float*, float*, void*) { return pending("es_inner"); }
int omni3d_es_inner_weighted(int, int, int, int, float, const float*, const float*, const float*,
                             float, float, float, float, const float*, const float*, const float*,
                             const float*, const float*, float*, float*, void*) { return pending("w"); }
int omni3d_es_inner_select(int, int, int, int, float, const float*, const float*, const float*,
                           float, float, float, float, const float*, const float*, const float*,
                           const float*, const float*, float*, float*, float, void*) { return pending("s"); }
int omni3d_devide_count_inner(int, int, int, float*, float*, const float*, void*) { return pending("d"); }
int omni3d_curl(int, int, int, float, float, float, const float*, const float*, const float*,
                const float*, float*, void*) { return pending("curl"); }
int omni3d_forward(omni3d_ctx*, const void*, void*) { return pending("fwd"); }
int omni3d_backward(omni3d_ctx*, const void*, void*) { return pending("bwd"); }
int omni3d_sor_solve(int, int, int, double, int, double, const double*, double*, int*, double*) { return pending("sor"); }
void sor3D(int, int, int, int, int*, double*, double*, double*, double*, double*, double*, double) {}
void calRHS(int, int, int, double, double, double, double, double*, double*, double*, double*) {}
int omni3d_run_parameter_file(const char*) { return -1; }
}
