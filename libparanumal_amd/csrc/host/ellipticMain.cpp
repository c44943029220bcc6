// ellipticMain: reads a settings (.rc) file and runs the C0 hex elliptic solve
// on one MI355X GPU. Mirrors solvers/elliptic/ellipticMain.cpp:29-80, including
// the "Solution norm = ..." last line parsed by the reference test harness.
#include <fstream>
#include <iostream>
#include <sstream>

#include "elliptic.hpp"

using namespace libp;

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "Usage: ./ellipticMain setupfile\n";
    return 1;
  }
  try {
    std::ifstream f(argv[1]);
    LIBP_ABORT("Failed to open: " << argv[1], !f.is_open());
    std::stringstream ss;
    ss << f.rdbuf();
    settings_t settings = ellipticDefaultSettings();
    settings.readSettingsFromText(ss.str());
    platform_t platform(settings);
    mesh_t mesh(platform, settings, comm_t());
    dfloat lambda = 0.0;
    settings.getSetting("LAMBDA", lambda);
    memory<int> BCType(3);
    BCType[0] = 0;
    BCType[1] = 1;
    BCType[2] = 2;
    elliptic_t elliptic(platform, mesh, settings, lambda, 3, BCType);
    ellipticRunResult_t r = elliptic.Run(1);
    (void)r;
  } catch (std::exception& e) {
    std::cerr << "libp error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
