// tomo_pretraitement: drop-in for pretraitement_hors_axe_CPU (installed as
// tomo_pretraitement, INSTALL.sh:22). Same argv (pretraitement_hors_axe_CPU/
// main.cpp:37-61: -polar, -etat <state>, --help), same configuration files
// and outputs (see pretraitement.hpp). -arrayfire (no reference counterpart)
// selects the ArrayFire variant's windows.
#include "pretraitement.hpp"

int main(int argc, char** argv) {
  const tomo::Args a = tomo::parse_args(argc, argv);
  if (a.help) {
    tomo::print_help();
    return 0;
  }
  tomo::Pre p;
  p.conf = a.conf;
  p.polar = a.polar;
  p.etat = a.etat;
  p.arrayfire = a.arrayfire;
  return tomo::run_pretraitement(p);
}
