/*
 * sph_oracle.c -- instantiates the CPU oracle for Float64 and Float32
 * (TEST INFRASTRUCTURE: loaded only by tests, smoke() and bench.py's
 * cpu_baseline leg; never by the product path).
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include "sph_oracle.h"

#include <math.h>
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288

#define REAL double
#define SFX _f64
#define FLOORR floor
#define SQRTR sqrt
#define EXPR exp
#define EPSR DBL_EPSILON
#include "sph_oracle_t.h"
#undef REAL
#undef SFX
#undef FLOORR
#undef SQRTR
#undef EXPR
#undef EPSR

#define clampi64 clampi64_f32_
#define nl_slot0 nl_slot0_f32_
#define xcalloc xcalloc_f32_
#undef AGE_NEVER
#define REAL float
#define SFX _f32
#define FLOORR floorf
#define SQRTR sqrtf
#define EXPR expf
#define EPSR FLT_EPSILON
#include "sph_oracle_t.h"

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
