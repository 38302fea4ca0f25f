// wpic_app2d -- command-line entry of the 2D problem drivers, as the
// reference's app binaries take it (2d/proj/*/main.f90): an optional config
// file path (default config.json); an optional second argument "parity"
// selects the bit-reproducible arithmetic mode.
#include <cstring>

#include "../../include/wpic.h"
#include "../../include/wpic_app.h"

int main(int argc, char** argv) {
  const char* config = argc > 1 ? argv[1] : "config.json";
  const int mode = argc > 2 && std::strcmp(argv[2], "parity") == 0 ? WPIC_MODE_PARITY : WPIC_MODE_FAST;
  return wpic_app2d_main(config, mode) == 0 ? 0 : 1;
}
