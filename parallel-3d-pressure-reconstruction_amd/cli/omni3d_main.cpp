// omni3d: command-line drop-in for Omni3D_8.0.exe.
//
//   omni3d [--fp32] [Parameter_Omni3D.dat]
//
// Reads the parameter file (default: Parameter_Omni3D.dat in the current
// folder), reconstructs every acceleration file it names and writes the
// pressure files, like the reference executable (kernel.cu:8467-8983).
#include <cstdio>
#include <cstring>

#include "omni3d.h"

int main(int argc, char** argv) {
  const char* path = "Parameter_Omni3D.dat";
  int precision = OMNI3D_FP64;
  for (int i = 1; i < argc; i++) {
    if (std::strcmp(argv[i], "--fp32") == 0) {
      precision = OMNI3D_FP32_KAHAN;
    } else if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
      std::printf("usage: %s [--fp32] [Parameter_Omni3D.dat]\n", argv[0]);
      return 0;
    } else {
      path = argv[i];
    }
  }
  std::printf("Processing Starts\n");
  const int n = omni3d_run_parameter_file_ex(path, precision);
  if (n < 0) {
    std::fprintf(stderr, "omni3d: %s\n", omni3d_last_global_error());
    return 1;
  }
  std::printf("%d file(s) reconstructed\n", n);
  return 0;
}
