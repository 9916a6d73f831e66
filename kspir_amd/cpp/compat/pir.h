// pir.h — drop-in for the reference's umbrella header (src/pir.h): a driver
// such as tests/test-pir.cpp compiles unchanged against the MI355X library
// with -I kspir_amd/cpp/compat -I kspir_amd/cpp and links libkspir_amd.so.
// Everything test-pir calls is declared by kspir_pir.h; params.h adds the
// global names N, crtMod and bsMod.
#ifndef KSPIR_COMPAT_PIR_H
#define KSPIR_COMPAT_PIR_H

#include "kspir_pir.h"
#include "params.h"

#endif
