// tomo_pretraitement_GPU: drop-in for pretraitement_hors_axe_GPU, the
// ArrayFire variant (installed as tomo_pretraitement_GPU, INSTALL.sh:23).
// Its interface: argv[1] is the GPU number (main.cpp:88, default 0), the
// configuration is gui_tomo.conf (no polarisation mode: manip's default
// constructor, main.cpp:32). Its semantics: the tukey windows of alpha 0.025
// for the hologram and the (1+i) crop window (main.cpp:55-60,
// src/GPU_fonctions.cpp:122), and its two extra outputs, phasePI_atan2.raw
// and UBornAmpFinal.raw (main.cpp:214,319; see tomo::Pre). Its Born field
// carries no -1 on the imaginary part (main.cpp:286 writes
// |A| sin(phi) - 1, a slip the CPU program does not have), and
// UBornAmpFinal.raw holds |A| where the variant keeps the clamped A, whose
// sign the fields do not carry (A < 0 only where the fitted amplitude
// background is negative).
#include <cstdlib>

#include "pretraitement.hpp"

int main(int argc, char** argv) {
  tomo::Pre p;
  p.device = argc > 1 ? std::atoi(argv[1]) : 0;
  p.arrayfire = true;
  p.variant_outputs = true;
  return tomo::run_pretraitement(p);
}
