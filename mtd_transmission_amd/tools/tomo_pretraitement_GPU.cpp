// tomo_pretraitement_GPU: the installed name of the ArrayFire variant
// (INSTALL.sh:23). It runs the same gfx950 preprocessing with the CPU
// program's semantics (see tomo_pretraitement.cpp): its Born field carries
// no -1 on the imaginary part (pretraitement_hors_axe_GPU/main.cpp:286), and
// the variant's windows (alpha 0.025, the (1+i) crop window) are selected by
// the -arrayfire argument, on either name.
#include "tomo_pretraitement.cpp"
