// `phi <paramfile> <method> [threads]` on one MI355X: the process-level
// drop-in for PHI's main (phi.cc:4-10).
#include "../../include/phi_hip.h"

int main(int argc, char* argv[]) {
  const int rc = phi_hip_main(argc, (const char* const*)argv);
  return rc == PHI_HIP_OK ? 0 : 1;
}
