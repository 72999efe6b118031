/*
 * ORACLE — test infrastructure only (see oracle_fvm.h).
 * Instantiates oracle_kernels.inc for T = double and T = float.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, x86-64 baseline, no FMA;
 * the reference's Release flags are -O2 -g with no -march, CMakeLists.txt:83-90).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_fvm.h"

static double or_carry_init = 0.0;

void or_set_carry_init(double v) { or_carry_init = v; }

#define T double
#define S f64
#define SQRTW sqrt
#include "oracle_kernels.inc"
#include "oracle_adjoint.inc"
#undef T
#undef S
#undef SQRTW

#define T float
#define S f32
#define SQRTW sqrtf
#include "oracle_kernels.inc"
#undef T
#undef S
#undef SQRTW
