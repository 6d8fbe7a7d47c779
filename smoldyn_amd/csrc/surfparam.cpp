// Surface interaction rates -> per-collision probabilities for the host setup
// of the device engine, as the reference computes them when a surface action
// is given by a rate (source/libSteve/SurfaceParam.c, public domain:
// surfaceprob :66-200, surfacetransmit :297-438 (its kappa -> probability
// branch), interpolate1D :452-468, interpolate2D :472-501, lookupirrevadsorb
// :504-524, lookuprevads :566-662 (probability from kappa); experfcD and
// experfcm1D of source/libSteve/math2.c:430-447).  Floating-point operations
// keep the reference's order so probabilities agree bit for bit; the tables
// come from surfparam_tables.inc.
#include <math.h>

#include "rxnparam.h"
#include "surfparam.h"

namespace smolamd {
namespace surfparam {

#include "surfparam_tables.inc"

// the reference's truncated constants (math2.h:11-14)
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.7724538509;
constexpr double kSqrt2Pi = 2.50662827462;
constexpr int kN = 21;  // points per table axis

// exp(x^2) erfc(x): direct for |x| < 20, else the asymptotic series
// 1/(x sqrt(pi)) (1 - 1/2 x^-2 + 3/4 x^-4 - 15/8 x^-6 + 105/16 x^-8 - 945/32 x^-10)
static double experfc(double x) {
  if (fabs(x) < 20) return exp(x * x) * rxnparam::erfcc(x);
  double u = 1.0 / (x * x);
  double s = 1.0 + u * (-0.5 + u * (0.75 + u * (-1.875 + u * (6.5625 + u * (-29.53125)))));
  s /= x * kSqrtPi;
  if (x < 0) s += 2.0 * exp(x * x);
  return s;
}

// exp(x^2) erfc(x) - 1; near 0 its Taylor series to x^8, whose odd
// coefficients carry 1/sqrt(pi): -2, 4/3, 8/15 and 16/105 over sqrt(pi)
double experfcm1(double x) {
  if (fabs(x) > 0.05) return experfc(x) - 1;
  constexpr double c1 = -2.0 / kSqrtPi, c3 = -4.0 / 3.0 / kSqrtPi, c5 = -8.0 / 15.0 / kSqrtPi,
                   c7 = -16.0 / 105.0 / kSqrtPi;
  return x * (c1 + x * (1.0 + x * (c3 + x * (0.5 + x * (c5 + x * (1.0 / 6.0 + x * (c7 + x * (1.0 / 24.0))))))));
}

// one term of the cubic Lagrange interpolant through xd[0..3]: the value v
// at node k times the basis polynomial of node k evaluated at x, multiplied
// in node order and divided last
static double lagrange(double v, const double* xd, int k, double x) {
  double num = v, den = 1.0;
  for (int j = 0; j < 4; j++)
    if (j != k) {
      num *= x - xd[j];
      den *= xd[k] - xd[j];
    }
  return num / den;
}

// first node of the four-point stencil: the first grid point not below x,
// two to the left, kept inside the table
static int stencil(const double* grid, int n, double x) {
  int i = 0;
  while (i < n && grid[i] < x) i++;
  if (i < 2) i = 2;
  else if (i > n - 2) i = n - 2;
  return i - 2;
}

static double interp1(const double* xs, const double* ys, int n, double x) {
  int s = stencil(xs, n, x);
  double ans = lagrange(ys[s], xs + s, 0, x);
  for (int k = 1; k < 4; k++) ans += lagrange(ys[s + k], xs + s, k, x);
  return ans;
}

// z[ix * ny + iy]: y is the fast index
static double interp2(const double* xs, const double* ys, const double* z, int nx, int ny, double x, double y) {
  int sx = stencil(xs, nx, x), sy = stencil(ys, ny, y);
  double w[4], col[4];
  for (int k = 0; k < 4; k++) w[k] = lagrange(1.0, xs + sx, k, x);
  for (int m = 0; m < 4; m++) {
    const double* zc = z + (size_t)sx * ny + sy + m;
    col[m] = w[0] * zc[0] + w[1] * zc[ny] + w[2] * zc[2 * ny] + w[3] * zc[3 * ny];
  }
  double ans = lagrange(col[0], ys + sy, 0, y);
  for (int m = 1; m < 4; m++) ans += lagrange(col[m], ys + sy, m, y);
  return ans;
}

// adsorption probability of an irreversibly adsorbing surface from its
// reduced adsorption coefficient kappa = k dt / step
double lookup_irrev_adsorb(double kappa) {
  if (kappa <= 0) return 0;
  if (kappa >= SMOLSP_irr_kappa[kN - 1]) return 1.0;
  return interp1(SMOLSP_irr_kappa, SMOLSP_irr_pon, kN, kappa);
}

// adsorption and desorption probabilities of a reversibly adsorbing surface
// from kappa = ka dt / step and kr = kd dt, interpolated on the table's
// axes kappa/(1+kappa) and kr/(1+kr)
double lookup_rev_ads(double kappa, double kr, double* p2) {
  if (kappa < 0) kappa = 0;
  if (kr < 0) kr = 0;
  double p1 = interp2(SMOLSP_rev_x, SMOLSP_rev_x, SMOLSP_rev_adsprob, kN, kN, kappa / (1.0 + kappa), kr / (1.0 + kr));
  if (p1 < 0) p1 = 0;
  else if (p1 > 1) p1 = 1;
  if (p2) *p2 = p1 * kr / kappa / kSqrt2Pi;
  return p1;
}

int transmit_probs(double kap1, double kap2, double difc1, double difc2, double dt, double* p1, double* p2) {
  if (difc1 <= 0 || difc2 <= 0 || dt <= 0) return 1;
  if (kap1 == -2 || kap2 == -2) return 2;
  double q1, q2;
  if (kap1 == 0 && kap2 == 0) {
    q1 = q2 = 0;
  } else if (kap1 == 0 && kap2 == -1) {
    q1 = 0;
    q2 = 1;
  } else if (kap1 == -1 && kap2 == 0) {
    q1 = 1;
    q2 = 0;
  } else if (kap1 == -1 && kap2 == -1) {
    // both sides absorb every hit: the slower side's probability is scaled
    // by the ratio of the diffusive fluxes
    if (difc1 <= difc2) {
      q1 = 1;
      q2 = sqrt(difc1 / difc2);
    } else {
      q1 = sqrt(difc2 / difc1);
      q2 = 1;
    }
  } else if (kap1 == -1 || kap2 == -1) {
    // the reference finds the other side's probability by a randomized
    // greedy search (SurfaceParam.c:406-426), which is not restated
    return 3;
  } else {
    double c = sqrt(dt) * (kap1 / sqrt(difc1) + kap2 / sqrt(difc2));
    double cterm = (2 * c / sqrt(kPi) + experfcm1(c)) / (c * c);
    q1 = kap1 * sqrt(kPi * dt) / sqrt(difc1) * cterm;
    q2 = kap2 * sqrt(kPi * dt) / sqrt(difc2) * cterm;
  }
  // capped at 1, the other side keeping the ratio of the two probabilities
  if (q1 > 1) {
    q1 = 1;
    q2 = kap2 / kap1 * sqrt(difc1 / difc2);
  }
  if (q2 > 1) {
    q2 = 1;
    q1 = kap1 / kap2 * sqrt(difc2 / difc1);
  }
  if (p1) *p1 = q1;
  if (p2) *p2 = q2;
  return 0;
}

double surfaceprob(double k1, double k2, double dt, double difc, double* p2ptr, int algo) {
  double step = sqrt(2.0 * difc * dt);
  double p1 = 0, p2 = 0;
  switch (algo) {
    case kIrrTrans:
    case kIrrTransT:
      return surfaceprob(k1, k2, dt, difc, p2ptr, kIrrAds);
    case kIrrTransQ:
      return surfaceprob(k1, k2, dt, difc, p2ptr, kIrrAdsQ);
    case kRevTrans:
      p1 = p2 = -2;
      transmit_probs(k1, k2, difc, difc, dt, &p1, &p2);
      break;
    case kIrrAds:
    case kIrrAdsT:
      p1 = k1 == -1 ? 1 : lookup_irrev_adsorb(k1 * dt / step);
      break;
    case kIrrAdsQ:  // quartic fit of the irreversible adsorption curve
      if (k1 == -1) {
        p1 = 1;
      } else {
        double kap = k1 * dt / step;
        if (kap <= 0) p1 = 0;
        else if (kap >= 0.9) p1 = 1.0;
        else {
          p1 = kap * (kSqrt2Pi + kap * (-3.3332146 + kap * (3.356688 + kap * (-1.5209235))));
          if (p1 > 1.0) p1 = 1.0;
        }
      }
      break;
    case kIrrAdsEC:  // the linear (small-kappa) limit sqrt(2 pi) kappa
      if (k1 == -1) {
        p1 = 1;
      } else {
        p1 = k1 * dt / step * kSqrt2Pi;
        if (p1 < 0) p1 = 0;
        else if (p1 > 1) p1 = 1;
      }
      break;
    case kRevAds:
      if (k1 == -1 && k2 == -1) {
        p1 = p2 = 1;
      } else if (k1 == -1) {
        p1 = 1;
        p2 = 1.0 - exp(-k2 * dt);
      } else if (k2 == -1) {
        p2 = 1;
        p1 = lookup_irrev_adsorb(k1 * dt / step);
      } else {
        p1 = lookup_rev_ads(k1 * dt / step, k2 * dt, &p2);
      }
      break;
    case kIrrDes:
    case kIrrFlip:  // first-order loss through every exit channel of summed rate k2
      if (k1 == -1) {
        p1 = 1;
      } else {
        if (k2 <= 0) k2 = k1;
        p1 = k1 / k2 * (1.0 - exp(-k2 * dt));
      }
      break;
    case kRevDes:  // the reversible adsorption problem read backwards
      if (k1 == -1 && k2 == -1) {
        p1 = p2 = 1;
      } else if (k1 == -1) {
        p1 = 1;
        p2 = lookup_irrev_adsorb(k2 * dt / step);
      } else if (k2 == -1) {
        p1 = 1.0 - exp(-k2 * dt);  // as the reference writes it (k2 = -1 here)
        p2 = 1;
      } else {
        p2 = surfaceprob(k2, k1, dt, difc, &p1, kRevAds);
      }
      break;
    case kRevFlip:  // two-state relaxation toward equilibrium
      if (k1 == -1 && k2 == -1) {
        p1 = p2 = 1;
      } else if (k1 == -1) {
        p1 = 1;
        p2 = 1.0 - exp(-k2 * dt);
      } else if (k2 == -1) {
        p1 = 1.0 - exp(-k1 * dt);
        p2 = 1;
      } else {
        p1 = k1 / (k1 + k2) * (1.0 - exp(-(k1 + k2) * dt));
        p2 = k2 / (k1 + k2) * (1.0 - exp(-(k1 + k2) * dt));
      }
      break;
    default:  // kRevAdsND (not restated), kIrrDesC, kRevAdsC, unrecognized
      p1 = p2 = -1;
  }
  if (p2ptr) *p2ptr = p2;
  return p1;
}

}  // namespace surfparam
}  // namespace smolamd
