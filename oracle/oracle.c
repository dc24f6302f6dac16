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
#define SIN sinf
#define COS cosf
#define ASIN asinf
#include "oracle_kernels.h"
#undef REAL
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef ASIN

#define REAL double
#define SFX _f64
#define SQRT sqrt
#define SIN sin
#define COS cos
#define ASIN asin
#include "oracle_kernels.h"
#undef REAL
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef ASIN

int oracle_abi_version(void) { return 1; }
