/* CPU oracle internals (test infrastructure only; never linked into the product). */
#ifndef ORACLE_INTERNAL_H
#define ORACLE_INTERNAL_H

#include <stdint.h>

typedef struct osfmt {
  uint32_t state[624];
  int idx;
} osfmt;

void osfmt_seed(osfmt* s, uint32_t seed);
uint32_t osfmt_u32(osfmt* s);

/* omath.c: Numerical-Recipes-derived helpers the reference uses to build its
   Gaussian lookup table (source/libSteve/math2.c, random2.c). */
double omath_gammaln(double x);
double omath_gammp(double a, double x);
double omath_erfn(double x);
double omath_erfcc(double x);
double omath_inversefn(double (*fn)(double), double y, double x1, double x2, int n);
void omath_randtable(double* a, int n);
void omath_dircos(double* c, double theta, double phi, double chi);
void omath_dircos2(double* c, double theta);
double omath_sphere_reflect_sphere(const double* a0, const double* a1, const double* b0, const double* b1, int dim,
                                   double radius2, double* a1p, double* b1p);

/* orxnparam.c: Andrews-Bray reaction parameter solver (source/libSteve/rxnparam.c). */
double orxn_modelrxnrate(double a, double b, double difc, double chi);
double orxn_numrxnrate(double step, double a, double b);
double orxn_numrxnrateprob(double step, double a, double b, double prob);
double orxn_actrxnrate(double step, double a, double prob);
double orxn_bindingradius(double rate, double dt, double difc, double b, int rel);
double orxn_bindingradiusprob(double rate, double dt, double difc, double b, int rel, double chi,
                              double* probptr);
double orxn_unbindingradius(double pgem, double dt, double difc, double a);

#endif
