/* Oracle (test infrastructure only): restatement of the reference's surface
 * parameter functions that turn interaction rates into per-collision
 * probabilities, source/libSteve/SurfaceParam.c (public domain):
 *   surfaceprob        :66-200
 *   surfacetransmit    :297-438, the branch that finds probabilities from
 *                       kappas (:311-352); its other branches search kappas
 *                       with random draws and are not restated
 *   interpolate1D      :452-468
 *   interpolate2D      :472-501
 *   lookupirrevadsorb  :504-524 (probability from kappa)
 *   lookuprevads       :566-662 (probabilities from kappas)
 * and experfcD / experfcm1D of source/libSteve/math2.c:430-447.
 * Tables are the reference's numbers (smoldyn_amd/csrc/surfparam_tables.inc,
 * data only).  Pinned against the compiled reference's surfaceprob,
 * surfacetransmit, lookupirrevadsorb, lookuprevads and experfcm1D on grids
 * that reach every branch, and by closed forms (tests/test_oracle_surfparam.py).
 * SPArevAdsND (bisection on lookuprevadsorbnd) is not restated: srfcalcprob
 * never asks for it. */
#include <math.h>

#include "oracle.h"
#include "oracle_internal.h"

#include "../smoldyn_amd/csrc/surfparam_tables.inc"

#define OSP_PI 3.14159265358979323846
#define OSP_SQRTPI 1.7724538509   /* math2.h:13, truncated as the reference has it */
#define OSP_SQRT2PI 2.50662827462 /* math2.h:14 */

enum { OSP_IRRTRANS, OSP_IRRTRANST, OSP_IRRTRANSQ, OSP_REVTRANS, OSP_IRRADS, OSP_IRRADST, OSP_IRRADSQ,
       OSP_IRRADSEC, OSP_REVADS, OSP_REVADSND, OSP_IRRDES, OSP_REVDES, OSP_IRRFLIP, OSP_REVFLIP };

/* experfcD (math2.c:430-439) */
static double osp_experfc(double x) {
  double ans, xxinv;
  if (fabs(x) < 20) ans = exp(x * x) * omath_erfcc(x);
  else {
    xxinv = 1.0 / (x * x);
    ans = 1.0 + xxinv * (-1.0 / 2.0 + xxinv * (3.0 / 4.0 + xxinv * (-15.0 / 8.0 + xxinv * (105.0 / 16.0 + xxinv * (-945.0 / 32.0)))));
    ans /= x * OSP_SQRTPI;
    if (x < 0) ans += 2.0 * exp(x * x);
  }
  return ans;
}

/* experfcm1D (math2.c:442-447) */
double osurf_experfcm1(double x) {
  if (fabs(x) > 0.05) return osp_experfc(x) - 1;
  return x * (-2.0 / OSP_SQRTPI +
              x * (1.0 + x * (-4.0 / 3.0 / OSP_SQRTPI +
                              x * (1.0 / 2.0 + x * (-8.0 / 15.0 / OSP_SQRTPI +
                                                    x * (1.0 / 6.0 + x * (-16.0 / 105.0 / OSP_SQRTPI + x * (1.0 / 24.0))))))));
}

/* interpolate1D (SurfaceParam.c:452-468) */
static double osp_interp1(const double* xdata, const double* ydata, int n, double x) {
  int i;
  const double* xd;
  double ans;
  for (i = 0; i < n && xdata[i] < x; i++) {
  }
  if (i < 2) i = 2;
  else if (i > n - 2) i = n - 2;
  xd = xdata + (i - 2);
  ans = ydata[i - 2] * (x - xd[1]) * (x - xd[2]) * (x - xd[3]) / ((xd[0] - xd[1]) * (xd[0] - xd[2]) * (xd[0] - xd[3]));
  ans += ydata[i - 1] * (x - xd[0]) * (x - xd[2]) * (x - xd[3]) / ((xd[1] - xd[0]) * (xd[1] - xd[2]) * (xd[1] - xd[3]));
  ans += ydata[i] * (x - xd[0]) * (x - xd[1]) * (x - xd[3]) / ((xd[2] - xd[0]) * (xd[2] - xd[1]) * (xd[2] - xd[3]));
  ans += ydata[i + 1] * (x - xd[0]) * (x - xd[1]) * (x - xd[2]) / ((xd[3] - xd[0]) * (xd[3] - xd[1]) * (xd[3] - xd[2]));
  return ans;
}

/* interpolate2D (SurfaceParam.c:472-501); y is the fast index of zdata */
static double osp_interp2(const double* xdata, const double* ydata, const double* zdata, int nx, int ny, double x,
                          double y) {
  int i, j, m;
  const double *xd, *yd;
  double xc[4], zv[4], ans;
  for (i = 0; i < nx && xdata[i] < x; i++) {
  }
  for (j = 0; j < ny && ydata[j] < y; j++) {
  }
  if (i < 2) i = 2;
  else if (i > nx - 2) i = nx - 2;
  if (j < 2) j = 2;
  else if (j > ny - 2) j = ny - 2;
  xd = xdata + (i - 2);
  xc[0] = (x - xd[1]) * (x - xd[2]) * (x - xd[3]) / ((xd[0] - xd[1]) * (xd[0] - xd[2]) * (xd[0] - xd[3]));
  xc[1] = (x - xd[0]) * (x - xd[2]) * (x - xd[3]) / ((xd[1] - xd[0]) * (xd[1] - xd[2]) * (xd[1] - xd[3]));
  xc[2] = (x - xd[0]) * (x - xd[1]) * (x - xd[3]) / ((xd[2] - xd[0]) * (xd[2] - xd[1]) * (xd[2] - xd[3]));
  xc[3] = (x - xd[0]) * (x - xd[1]) * (x - xd[2]) / ((xd[3] - xd[0]) * (xd[3] - xd[1]) * (xd[3] - xd[2]));
  for (m = 0; m < 4; m++) {
    int jj = j - 2 + m;
    zv[m] = xc[0] * zdata[ny * (i - 2) + jj] + xc[1] * zdata[ny * (i - 1) + jj] + xc[2] * zdata[ny * i + jj] +
            xc[3] * zdata[ny * (i + 1) + jj];
  }
  yd = ydata + (j - 2);
  ans = zv[0] * (y - yd[1]) * (y - yd[2]) * (y - yd[3]) / ((yd[0] - yd[1]) * (yd[0] - yd[2]) * (yd[0] - yd[3]));
  ans += zv[1] * (y - yd[0]) * (y - yd[2]) * (y - yd[3]) / ((yd[1] - yd[0]) * (yd[1] - yd[2]) * (yd[1] - yd[3]));
  ans += zv[2] * (y - yd[0]) * (y - yd[1]) * (y - yd[3]) / ((yd[2] - yd[0]) * (yd[2] - yd[1]) * (yd[2] - yd[3]));
  ans += zv[3] * (y - yd[0]) * (y - yd[1]) * (y - yd[2]) / ((yd[3] - yd[0]) * (yd[3] - yd[1]) * (yd[3] - yd[2]));
  return ans;
}

/* lookupirrevadsorb, probability from kappa (SurfaceParam.c:504-524) */
double osurf_lookupirrevadsorb(double value) {
  if (value <= 0) return 0;
  if (value >= SMOLSP_irr_kappa[20]) return 1.0;
  return osp_interp1(SMOLSP_irr_kappa, SMOLSP_irr_pon, 21, value);
}

/* lookuprevads, probabilities from kappas (SurfaceParam.c:639-648) */
double osurf_lookuprevads(double value1, double value2, double* ans2ptr) {
  double kappar, kr, ans1, ans2;
  if (value1 < 0) value1 = 0;
  if (value2 < 0) value2 = 0;
  kappar = value1 / (1.0 + value1);
  kr = value2 / (1.0 + value2);
  ans1 = osp_interp2(SMOLSP_rev_x, SMOLSP_rev_x, SMOLSP_rev_adsprob, 21, 21, kappar, kr);
  if (ans1 < 0) ans1 = 0;
  else if (ans1 > 1) ans1 = 1;
  ans2 = ans1 * value2 / value1 / OSP_SQRT2PI;
  if (ans2ptr) *ans2ptr = ans2;
  return ans1;
}

/* surfacetransmit with p1 = p2 = -2 on entry: probabilities from kappas
   (SurfaceParam.c:297-352).  Returns the reference's codes (1: a diffusion
   coefficient or dt not positive, 2: a kappa missing) and 3 where the
   reference would search with random draws (one side maximal, the other
   finite: :328-339 re-enter the function in its search branches). */
int osurf_transmit(double kap1, double kap2, double difc1, double difc2, double dt, double* p1ptr, double* p2ptr) {
  double newp1, newp2, c, cterm;
  if (difc1 <= 0 || difc2 <= 0 || dt <= 0) return 1;
  if (kap1 == -2 || kap2 == -2) return 2;
  else if (kap1 == 0 && kap2 == 0) {
    newp1 = 0;
    newp2 = 0;
  } else if (kap1 == 0 && kap2 == -1) {
    newp1 = 0;
    newp2 = 1;
  } else if (kap1 == -1 && kap2 == 0) {
    newp1 = 1;
    newp2 = 0;
  } else if (kap1 == -1 && kap2 == -1) {
    if (difc1 <= difc2) {
      newp1 = 1;
      newp2 = sqrt(difc1 / difc2);
    } else {
      newp1 = sqrt(difc2 / difc1);
      newp2 = 1;
    }
  } else if (kap1 == -1 || kap2 == -1) return 3;
  else {
    c = sqrt(dt) * (kap1 / sqrt(difc1) + kap2 / sqrt(difc2));
    cterm = (2 * c / sqrt(OSP_PI) + osurf_experfcm1(c)) / (c * c);
    newp1 = kap1 * sqrt(OSP_PI * dt) / sqrt(difc1) * cterm;
    newp2 = kap2 * sqrt(OSP_PI * dt) / sqrt(difc2) * cterm;
  }
  if (newp1 > 1) { /* capprob: 1 outside the search branches */
    newp1 = 1;
    newp2 = kap2 / kap1 * sqrt(difc1 / difc2);
  }
  if (newp2 > 1) {
    newp2 = 1;
    newp1 = kap1 / kap2 * sqrt(difc2 / difc1);
  }
  if (p1ptr) *p1ptr = newp1;
  if (p2ptr) *p2ptr = newp2;
  return 0;
}

/* surfaceprob (SurfaceParam.c:66-200) */
double osurf_surfaceprob(double k1, double k2, double dt, double difc, double* p2ptr, int algo) {
  double step, p1, p2, kapp, kp;
  step = sqrt(2.0 * difc * dt);
  p1 = p2 = 0;
  if (algo == OSP_IRRTRANS) return osurf_surfaceprob(k1, k2, dt, difc, p2ptr, OSP_IRRADS);
  else if (algo == OSP_IRRTRANST) return osurf_surfaceprob(k1, k2, dt, difc, p2ptr, OSP_IRRADST);
  else if (algo == OSP_IRRTRANSQ) return osurf_surfaceprob(k1, k2, dt, difc, p2ptr, OSP_IRRADSQ);
  else if (algo == OSP_REVTRANS) {
    p1 = p2 = -2;
    osurf_transmit(k1, k2, difc, difc, dt, &p1, &p2);
  } else if (algo == OSP_IRRADS || algo == OSP_IRRADST) {
    if (k1 == -1) p1 = 1;
    else {
      kapp = k1 * dt / step;
      p1 = osurf_lookupirrevadsorb(kapp);
    }
  } else if (algo == OSP_IRRADSQ) {
    if (k1 == -1) p1 = 1;
    else {
      kapp = k1 * dt / step;
      if (kapp <= 0) p1 = 0;
      else if (kapp >= 0.9) p1 = 1.0;
      else {
        p1 = kapp * (OSP_SQRT2PI + kapp * (-3.3332146 + kapp * (3.356688 + kapp * (-1.5209235))));
        if (p1 > 1.0) p1 = 1.0;
      }
    }
  } else if (algo == OSP_IRRADSEC) {
    if (k1 == -1) p1 = 1;
    else {
      kapp = k1 * dt / step;
      p1 = kapp * OSP_SQRT2PI;
      if (p1 < 0) p1 = 0;
      else if (p1 > 1) p1 = 1;
    }
  } else if (algo == OSP_REVADS) {
    if (k1 == -1 && k2 == -1) p1 = p2 = 1;
    else if (k1 == -1) {
      p1 = 1;
      p2 = 1.0 - exp(-k2 * dt);
    } else if (k2 == -1) {
      p2 = 1;
      p1 = osurf_lookupirrevadsorb(k1 * dt / step);
    } else {
      kapp = k1 * dt / step;
      kp = k2 * dt;
      p1 = osurf_lookuprevads(kapp, kp, &p2);
    }
  } else if (algo == OSP_IRRDES || algo == OSP_IRRFLIP) {
    if (k1 == -1) p1 = 1;
    else {
      if (k2 <= 0) k2 = k1;
      p1 = k1 / k2 * (1.0 - exp(-k2 * dt));
    }
  } else if (algo == OSP_REVDES) {
    if (k1 == -1 && k2 == -1) p1 = p2 = 1;
    else if (k1 == -1) {
      p1 = 1;
      p2 = osurf_lookupirrevadsorb(k2 * dt / step);
    } else if (k2 == -1) {
      p1 = 1.0 - exp(-k2 * dt);
      p2 = 1;
    } else p2 = osurf_surfaceprob(k2, k1, dt, difc, &p1, OSP_REVADS);
  } else if (algo == OSP_REVFLIP) {
    if (k1 == -1 && k2 == -1) p1 = p2 = 1;
    else if (k1 == -1) {
      p1 = 1;
      p2 = 1.0 - exp(-k2 * dt);
    } else if (k2 == -1) {
      p1 = 1.0 - exp(-k1 * dt);
      p2 = 1;
    } else {
      p1 = k1 / (k1 + k2) * (1.0 - exp(-(k1 + k2) * dt));
      p2 = k2 / (k1 + k2) * (1.0 - exp(-(k1 + k2) * dt));
    }
  } else p1 = p2 = -1; /* SPArevAdsND (not restated) and unrecognized */
  if (p2ptr) *p2ptr = p2;
  return p1;
}
