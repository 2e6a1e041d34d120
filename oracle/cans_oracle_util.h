/*
 * Shared helpers of the C restatement (test infrastructure only).
 *
 * Fortran intrinsic semantics as compiled by the reference's compiler:
 *   MAX(a,b) = (a > b) ? a : b      (NaN in b propagates, -0/+0 keeps b)
 *   MIN(a,b) = (a < b) ? a : b
 *   SIGN(a,b) = |a| with the sign bit of b (copysign)
 *   x**2     = x*x
 * REAL*4 literals that the reference mixes into double expressions are
 * promoted from float, e.g. 1.0e-4 -> (double)1.0e-4f.
 */
#ifndef CANS_ORACLE_UTIL_H
#define CANS_ORACLE_UTIL_H
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double fmaxF(double a, double b) { return (a > b) ? a : b; }
static inline double fminF(double a, double b) { return (a < b) ? a : b; }
static inline double signF(double a, double b) { return copysign(fabs(a), b); }
static inline double sq(double a) { return a * a; }

#define R4_1EM4 ((double)1.0e-4f)  /* REAL*4 literal 1.0e-4 */
#define R4_1E20 ((double)1.e20f)   /* REAL*4 literal 1.e20  */
#define R4_PI   ((double)3.14159274101257324) /* acos(-1.0), REAL*4 pi */

static inline double *dalloc(size_t n) {
    double *p = (double *)malloc(n * sizeof(double) + 64);
    /* temporaries the reference leaves uninitialised are NaN-filled here so
       that any use of them in the interior shows up in the parity tests */
    for (size_t i = 0; i < n; i++) p[i] = NAN;
    return p;
}

/* column-major index helpers with the reference's 1-based indices */
#define I1(i) ((i) - 1)
#define I2(i, j) ((size_t)((i) - 1) + (size_t)ix * (size_t)((j) - 1))
#define I3(i, j, k) ((size_t)((i) - 1) + (size_t)ix * ((size_t)((j) - 1) + (size_t)jx * (size_t)((k) - 1)))

#endif
