/* CPU oracle (test infrastructure only; never linked into the product).
 *
 * Restatement of the numerical helpers behind Smoldyn's quantized Gaussian
 * table and product-placement rotations:
 *   gammaln    source/libSteve/math2.c:261-296 (Numerical Recipes + CRC)
 *   gammp      source/libSteve/math2.c:321-348
 *   erfn       source/libSteve/math2.c:387-388
 *   erfcc      source/libSteve/math2.c:408-414 (and rxnparam.c:719-725)
 *   inversefn  source/libSteve/math2.c:526-538 (bisection)
 *   randtable  source/libSteve/random2.c:339-353 (eq==1 branch)
 *   dircos     source/libSteve/Rn.c:83-101, 110-113
 * Operation order follows the reference so results are bit-identical; checked
 * against the compiled reference by tests/test_oracle_ref.py.
 */
#include <float.h>
#include <math.h>

#include "oracle_internal.h"

#define OM_SQRT2 1.41421356237

double omath_gammaln(double x) {
  static const double c[6] = {76.18009173, -86.50532033, 24.01409822,
                              -1.231739516, 0.120858003e-2, -0.536382e-5};
  double sum, t;
  int j;
  if (x == floor(x) && x <= 0) {
    sum = DBL_MAX;
  } else if (x == floor(x)) {
    sum = 0;
    for (t = 2; t < x - 0.1; t += 1.0) sum += log(t);
  } else if (x == 0.5) {
    sum = 0.572364942;
  } else if (2 * x == floor(2 * x) && x > 0) {
    sum = 0.572364942;
    for (t = 0.5; t < x - 0.1; t += 1) sum += log(t);
  } else if (2 * x == floor(2 * x)) {
    sum = 0.572364942;
    for (t = 0.5; t < -x + 0.1; t += 1) sum -= log(t);
  } else if (x < 0) {
    sum = omath_gammaln(x + 1) - log(-x);
  } else {
    x -= 1.0;
    t = x + 5.5;
    t -= (x + 0.5) * log(t);
    sum = 1.0;
    for (j = 0; j <= 5; j++) {
      x += 1.0;
      sum += c[j] / x;
    }
    sum = -t + log(2.50662827465 * sum);
  }
  return sum;
}

double omath_gammp(double a, double x) {
  double sum, del, ap, eps;
  double gold = 0, g = 1, fac = 1, b1 = 1, b0 = 0, anf, ana, an = 0, a1, a0 = 1;
  eps = 3e-7;
  if (x < 0 || a <= 0) return -1;
  if (x == 0) return 0;
  if (x < a + 1) {
    ap = a;
    del = sum = 1.0 / a;
    while (fabs(del) > fabs(sum) * eps && ap - a < 100) {
      ap += 1;
      del *= x / ap;
      sum += del;
    }
    return sum * exp(-x + a * log(x) - omath_gammaln(a));
  }
  a1 = x;
  for (an = 1; an < 100; an++) {
    ana = an - a;
    a0 = (a1 + a0 * ana) * fac;
    b0 = (b1 + b0 * ana) * fac;
    anf = an * fac;
    a1 = x * a0 + anf * a1;
    b1 = x * b0 + anf * b1;
    if (a1) {
      fac = 1.0 / a1;
      g = b1 * fac;
      if (fabs((g - gold) / g) < eps) return 1.0 - exp(-x + a * log(x) - omath_gammaln(a)) * g;
      gold = g;
    }
  }
  return -1;
}

double omath_erfn(double x) { return x < 0 ? -omath_gammp(0.5, x * x) : omath_gammp(0.5, x * x); }

double omath_erfcc(double x) {
  double t, z, ans;
  z = fabs(x);
  t = 1.0 / (1.0 + 0.5 * z);
  ans = t * exp(-z * z - 1.26551223 +
                t * (1.00002368 +
                     t * (0.37409196 +
                          t * (0.09678418 +
                               t * (-0.18628806 +
                                    t * (0.27886807 +
                                         t * (-1.13520398 +
                                              t * (1.48851587 +
                                                   t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}

double omath_inversefn(double (*fn)(double), double y, double x1, double x2, int n) {
  double dx, y2;
  if ((*fn)(x1) < (*fn)(x2)) {
    dx = x2 - x1;
  } else {
    dx = x1 - x2;
    x1 = x2;
  }
  for (; n > 0; n--) {
    dx *= 0.5;
    x2 = x1 + dx;
    y2 = (*fn)(x2);
    if (y2 < y) x1 = x2;
  }
  return x1 + 0.5 * dx;
}

/* Gaussian quantiles at (i+0.5)/n, antisymmetric; the unshuffled table. */
void omath_randtable(double* a, int n) {
  int i;
  double dy = 2.0 / n;
  for (i = 0; i < n / 2; i++) a[i] = OM_SQRT2 * omath_inversefn(omath_erfn, (i + 0.5) * dy - 1.0, -20, 20, 30);
  for (i = n / 2; i < n; i++) a[i] = -a[n - i - 1];
}

void omath_dircos(double* c, double theta, double phi, double chi) {
  double cp = cos(phi), ct = cos(theta), cc = cos(chi);
  double sp = sin(phi), st = sin(theta), sc = sin(chi);
  c[0] = cp * ct * cc - sp * sc;
  c[1] = sp * ct * cc + cp * sc;
  c[2] = -st * cc;
  c[3] = -cp * ct * sc - sp * cc;
  c[4] = -sp * ct * sc + cp * cc;
  c[5] = st * sc;
  c[6] = cp * st;
  c[7] = sp * st;
  c[8] = ct;
}

void omath_dircos2(double* c, double theta) {
  c[0] = c[3] = cos(theta);
  c[1] = sin(theta);
  c[2] = -c[1];
}

/* Geo_SphereReflectSphere (source/libSteve/Geometry.c:1264-1293): two
   molecules move on straight lines a0->a1 and b0->b1; p is the fraction of
   the step at which their separation first equals the binding radius
   (radius2 = radius squared, the smaller root of the quadratic in p).  At
   that moment each molecule's remaining motion is reflected off the plane
   normal to their separation (deltap); the reflected end points go to a1p and
   b1p.  Returns p (the caller uses the result only for 0 < p < 1). */
double omath_sphere_reflect_sphere(const double* a0, const double* a1, const double* b0, const double* b1, int dim,
                                   double radius2, double* a1p, double* b1p) {
  double qa = 0, qb = 0, qc = 0, p, f1, f2, k, deltap[3];
  int d;
  for (d = 0; d < dim; d++) {
    const double rel = b1[d] - a1[d] - b0[d] + a0[d]; /* change of the separation over the step */
    const double sep0 = b0[d] - a0[d];                 /* separation at the start */
    qa += rel * rel;
    qb += 2 * sep0 * rel;
    qc += sep0 * sep0;
  }
  qc -= radius2;
  p = (-qb - sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
  f1 = f2 = 0;
  for (d = 0; d < dim; d++) {
    deltap[d] = (1 - p) * (b0[d] - a0[d]) + p * (b1[d] - a1[d]);
    f1 += (a1[d] - a0[d]) * deltap[d];
    f2 += (b1[d] - b0[d]) * deltap[d];
  }
  k = 2 * (1 - p) / radius2;
  if (f1 >= 0)
    for (d = 0; d < dim; d++) a1p[d] = a1[d] - k * f1 * deltap[d];
  else
    for (d = 0; d < dim; d++) a1p[d] = 2 * (1 - p) * a0[d] + (2 * p - 1) * a1[d] + k * f1 * deltap[d];
  if (f2 <= 0)
    for (d = 0; d < dim; d++) b1p[d] = b1[d] - k * f2 * deltap[d];
  else
    for (d = 0; d < dim; d++) b1p[d] = 2 * (1 - p) * b0[d] + (2 * p - 1) * b1[d] + k * f2 * deltap[d];
  return p;
}
