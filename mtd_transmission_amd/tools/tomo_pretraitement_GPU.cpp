// tomo_pretraitement_GPU: the installed name of the ArrayFire variant
// (INSTALL.sh:23). It runs the same gfx950 preprocessing with the CPU
// program's semantics (see tomo_pretraitement.cpp).
#include "tomo_pretraitement.cpp"
