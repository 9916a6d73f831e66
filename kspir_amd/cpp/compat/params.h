// params.h — the reference's global names for drivers written against its
// headers (src/params.h:16 N, :94 bsgsp; src/crt.h:16-18 crtq1, crtq2,
// crtMod, :32 bsMod), as the same macros.  kspir_pir.h is included first, so
// its own declarations are never touched by these names.
#ifndef KSPIR_COMPAT_PARAMS_H
#define KSPIR_COMPAT_PARAMS_H

#include "kspir_pir.h"

#define N 4096
#define bsgsp 65537
#define crtq1 268369921
#define crtq2 249561089UL
#define crtMod (crtq1 * crtq2)
#define bsMod 16760833

#endif
