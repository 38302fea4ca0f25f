// wpic_app3d -- command-line entry of the 3D Weibel driver, as the
// reference's app binary takes it (3d/proj/weibel/main.f90): an optional
// config file path (default config.json); an optional second argument
// "parity" selects the bit-reproducible arithmetic mode.
#include <cstring>

#include "../../include/wpic.h"
#include "../../include/wpic_app.h"

int main(int argc, char** argv) {
  const char* config = argc > 1 ? argv[1] : "config.json";
  const int mode = argc > 2 && std::strcmp(argv[2], "parity") == 0 ? WPIC_MODE_PARITY : WPIC_MODE_FAST;
  return wpic_app3d_main(config, mode) == 0 ? 0 : 1;
}
