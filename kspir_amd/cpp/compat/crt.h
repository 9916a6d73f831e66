// crt.h — drop-in for the reference's src/crt.h on the test-pir path:
// database_tocrt (crt.h:133, declared by kspir_pir.h) and the CRT moduli
// crtq1, crtq2, crtMod, bsMod (params.h of this directory).
#ifndef KSPIR_COMPAT_CRT_H
#define KSPIR_COMPAT_CRT_H

#include "kspir_pir.h"
#include "params.h"

#endif
