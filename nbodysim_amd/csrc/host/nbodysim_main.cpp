// nbodysim_main.cpp -- the command-line driver, `nbodysim [settings file]`
// (src/main.cpp:1359-1504; default file settings.cfg).
#include <cstdio>

#include "../../../include/nbs_host.h"

int main(int argc, char** argv) {
  if (argc > 2) {
    fprintf(stderr, "Usage: %s [settings file]\n", argv[0]);
    return 1;
  }
  return nbs_run_frames(argc == 2 ? argv[1] : "settings.cfg");
}
