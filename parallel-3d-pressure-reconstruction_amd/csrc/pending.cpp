// Entry points whose implementation is not in this build yet: they fail
// loudly with OMNI3D_ERR_STATE.
#include "host.hpp"

extern "C" {
static int pending(const char* what) {
  (void)what;
  return OMNI3D_ERR_STATE;
}
int omni3d_run_parameter_file(const char*) { return -1; }
}
