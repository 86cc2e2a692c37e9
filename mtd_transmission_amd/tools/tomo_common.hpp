// Shared pieces of the drop-in executables: argv handling, the
// $HOME/.config/gui_tomo*.conf indirection, HIP error checks, stage timing.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "host_io.hpp"
#include "mtd.h"

namespace tomo {

struct Args {
  std::string conf = "gui_tomo.conf";
  std::string etat;
  bool polar = false;
  bool help = false;
  // -arrayfire (no reference counterpart): the windows of the ArrayFire
  // variant pretraitement_hors_axe_GPU (main.cpp:74-79): alpha 0.025 for the
  // hologram and a (1+i) tukey2D(C, 0.025) crop window
  bool arrayfire = false;
};

// Order-sensitive argv loop of pretraitement_hors_axe_CPU/main.cpp:37-61 and
// Reconstruction/main.cpp:63-86: every unmatched argument (argv[0] included)
// resets the configuration name to gui_tomo.conf.
inline Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 0; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "--help") {
      a.help = true;
      return a;
    } else if (arg == "-polar") {
      a.polar = true;
      a.conf = "gui_tomo_polar.conf";
    } else if (arg == "-etat" && i + 1 < argc) {
      a.etat = argv[++i];
      a.conf = "gui_tomo_polar.conf";
    } else if (arg == "-arrayfire") {
      a.arrayfire = true;
    } else {
      a.conf = "gui_tomo.conf";
    }
  }
  return a;
}

inline void print_help() {
  std::cout << "-polar : mode polar\n-etat : indiquez un des 4 états :\n"
               "/LC/polar0/\n/LC/polar90/\n/RC/polar0/\n/RC/polar90/\n"
               "-arrayfire : fenêtres de la variante ArrayFire (alpha 0.025, découpe (1+i))\n";
}

inline std::string gui_conf_path(const std::string& name) {
  const char* home = getenv("HOME");
  return std::string(home ? home : "") + "/.config/" + name;
}

inline void hip_or_die(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::cerr << "HIP error in " << what << ": " << hipGetErrorString(e) << "\n";
    exit(1);
  }
}

inline void mtd_or_die(mtd_ctx* ctx, int rc, const char* what) {
  if (rc != MTD_OK) {
    std::cerr << what << " failed: " << mtd_last_error(ctx) << "\n";
    exit(1);
  }
}

struct Clock {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double s() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

inline bool exists(const std::string& p) {
  FILE* f = fopen(p.c_str(), "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

}  // namespace tomo
