/*
 * TEST INFRASTRUCTURE ONLY - the oracle (see oracle_kernels.h).  Builds the
 * float and double restatements of the reference kernels into one library.
 * Compile with -ffp-contract=off so no multiply-add is fused.
 */
#include <math.h>
#include <stdlib.h>

#define REAL float
#define SFX _f32
#define SQRT sqrtf
#include "oracle_kernels.h"
#undef REAL
#undef SFX
#undef SQRT

#define REAL double
#define SFX _f64
#define SQRT sqrt
#include "oracle_kernels.h"
#undef REAL
#undef SFX
#undef SQRT

int oracle_abi_version(void) { return 1; }
