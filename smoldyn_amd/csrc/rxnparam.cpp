// Andrews-Bray reaction-parameter solver for the host setup of the device
// engine: converts rate constants to binding/unbinding radii exactly as the
// reference does (source/libSteve/rxnparam.c, public domain:
// numrxnrate :57-156, numrxnrateprob :161-528, actrxnrate :536-544,
// bindingradius :558-583, bindingradiusprob :587-680, unbindingradius
// :690-711).  The same floating-point operation order is kept so radii agree
// bit for bit; the tabulated data come from rxnparam_tables.inc.
#include <math.h>

#include "rxnparam.h"

namespace smolamd {
namespace rxnparam {

#include "rxnparam_tables.inc"

// complementary error function, Numerical Recipes form used by the reference
// (rxnparam.c:719-725; erfccD of math2.c:408-414 is the same expression)
double erfcc(double x) {
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
                                              t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2Pi = 2.50662827462;

double modelrxnrate(double a, double b, double difc, double chi) {
  if (b < 0) return 4 * kPi * difc * a * chi;
  if (b <= a) return -1;
  return 4 * kPi * difc * a * b * chi / (b - a);
}

/* Lagrange cubic weights on 4 equally spaced nodes x[0..3] with spacing h. */
static void cubicw(double t, const double* x, double h, double* z) {
  z[0] = (t - x[1]) * (t - x[2]) * (t - x[3]) / (-6.0 * h * h * h);
  z[1] = (t - x[0]) * (t - x[2]) * (t - x[3]) / (2.0 * h * h * h);
  z[2] = (t - x[0]) * (t - x[1]) * (t - x[3]) / (-2.0 * h * h * h);
  z[3] = (t - x[0]) * (t - x[1]) * (t - x[2]) / (6.0 * h * h * h);
}

double numrxnrate(double step, double a, double b) {
  const double *yi = SMOLTBL_yi, *yr = SMOLTBL_yr, *yb = SMOLTBL_yb;
  const double slo = 3, sinc = -0.2;
  const double blor = 0, bincr = 0.2;
  const double blob = 0, bincb = 0.1;
  const int snum = 31, bnumr = 16, bnumb = 11;
  double x[4], y[4], z[4], ans;
  int sindx, bindx, i, j;

  if (step < 0 || a < 0) return -1;
  if (step == 0 && b >= 0 && b < 1) return -1;
  if (step == 0) return 0;
  step = log(step / a);
  b /= a;
  sindx = (int)floor((step - slo) / sinc);
  for (i = 0; i < 4; i++) x[i] = slo + (sindx - 1 + i) * sinc;
  cubicw(step, x, sinc, z);
  if (b < 0) {
    ans = 0;
    for (i = 0; i < 4; i++) {
      if (sindx - 1 + i >= snum) ans += z[i] * 2.0 * kPi * exp(2.0 * x[i]);
      else if (sindx - 1 + i < 0) ans += z[i] * 4.0 * kPi / 3.0;
      else ans += z[i] * yi[sindx - 1 + i];
    }
  } else if (b < 1.0) {
    bindx = (int)floor((b - blob) / bincb);
    if (bindx < 1) bindx = 1;
    else if (bindx + 2 >= bnumb) bindx = bnumb - 3;
    while (sindx + 3 >= snum || (sindx > 0 && yb[(bindx - 1) * snum + sindx + 3] < 0)) sindx--;
    for (i = 0; i < 4; i++) x[i] = slo + (sindx - 1 + i) * sinc;
    cubicw(step, x, sinc, z);
    for (j = 0; j < 4; j++) {
      y[j] = 0;
      for (i = 0; i < 4; i++) {
        if (sindx - 1 + i >= snum) y[j] += z[i] * yb[(bindx - 1 + j) * snum];
        else if (sindx - 1 + i < 0) y[j] += z[i] * 4.0 * kPi / 3.0;
        else y[j] += z[i] * yb[(bindx - 1 + j) * snum + sindx - 1 + i];
      }
    }
    for (j = 0; j < 4; j++) x[j] = blob + (bindx - 1 + j) * bincb;
    cubicw(b, x, bincb, z);
    ans = z[0] * y[0] + z[1] * y[1] + z[2] * y[2] + z[3] * y[3];
  } else {
    b = log(b);
    bindx = (int)floor((b - blor) / bincr);
    if (bindx < 1) bindx = 1;
    else if (bindx + 2 >= bnumr) bindx = bnumr - 3;
    for (j = 0; j < 4; j++) {
      y[j] = 0;
      for (i = 0; i < 4; i++) {
        if (sindx - 1 + i >= snum && b == 0) y[j] += z[i] * 2.0 * kPi * exp(x[i]) * (1.0 + exp(x[i]));
        else if (sindx - 1 + i >= snum) y[j] += z[i] * 2.0 * kPi * exp(2.0 * x[i]) * exp(b) / (exp(b) - 1.0);
        else if (sindx - 1 + i < 0) y[j] += z[i] * 4.0 * kPi / 3.0;
        else y[j] += z[i] * yr[(bindx - 1 + j) * snum + sindx - 1 + i];
      }
    }
    for (j = 0; j < 4; j++) x[j] = blor + (bindx - 1 + j) * bincr;
    cubicw(b, x, bincr, z);
    ans = z[0] * y[0] + z[1] * y[1] + z[2] * y[2] + z[3] * y[3];
  }
  return ans * a * a * a;
}

double numrxnrateprob(double step, double a, double b, double prob) {
  const double *kirr = SMOLTBL_kirr, *krev = SMOLTBL_krev, *kb = SMOLTBL_kb;
  const double slo = -3, sinc = 0.2, shi = 3;
  const double blor = 0, bincr = 0.2;
  const double blob = 0, bincb = 0.1;
  const double pvect[] = {0, 0.01, 0.02, 0.05, 0.1, 0.2, 0.5, 1.0};
  const int snum = 31, bnumr = 16, bnumb = 11, pnum = 8;
  int sindx, bindx, pindx, sflag = 0;
  double logstep, steplo, stephi, problo, probhi, blo, bhi, step0, step1, logb;
  double denom, numll, numlh, numhl, numhh, kll, klh, khl, khh;
  double numlll, numllh, numlhl, numlhh, numhll, numhlh, numhhl, numhhh;
  double klll, kllh, klhl, klhh, khll, khlh, khhl, khhh;
  double lambdap, theory, ans, ansp1;

  if (prob == 1) return numrxnrate(step, a, b);
  if (step < 0 || a < 0) return -1;
  if (prob < 0 || prob > 1) return -1;
  if (prob == 0) return 0;
  if (a == 0) return 0;
  if (step == 0 && b >= 0 && b < a) return -1;
  if (step == 0) return 0;

  ansp1 = numrxnrate(step, a, b);
  step /= a;
  logstep = log(step);
  b /= a;

  if (b < 0 || b >= 1.0) {
    sindx = (int)floor((logstep - slo) / sinc);
    if (sindx < 0) {
      sindx = 0;
      sflag = -1;
    } else if (sindx > snum - 2) {
      sindx = snum - 2;
    } else {
      sflag = 0;
    }
    steplo = slo + sindx * sinc;
    stephi = slo + (sindx + 1) * sinc;
  } else {
    sindx = (int)floor((shi - logstep) / sinc);
    if (sindx < 0) {
      sindx = 0;
      sflag = -1;
    } else if (sindx > snum - 2) {
      sindx = snum - 2;
    } else {
      sflag = 0;
    }
    steplo = shi - sindx * sinc;
    stephi = shi - (sindx + 1) * sinc;
  }

  for (pindx = 0; pindx < pnum - 2 && prob > pvect[pindx + 1]; pindx++)
    ;
  problo = pvect[pindx];
  probhi = pvect[pindx + 1];

  if (b < 0) {
    denom = (probhi - problo) * (stephi - steplo);
    numll = (probhi - prob) * (stephi - logstep);
    numlh = (probhi - prob) * (logstep - steplo);
    numhl = (prob - problo) * (stephi - logstep);
    numhh = (prob - problo) * (logstep - steplo);
    kll = kirr[pindx * snum + sindx];
    klh = kirr[pindx * snum + sindx + 1];
    khl = kirr[(pindx + 1) * snum + sindx];
    khh = kirr[(pindx + 1) * snum + sindx + 1];
    ans = (numll * kll + numlh * klh + numhl * khl + numhh * khh) / denom;
    if (sindx == 0) {
      lambdap = sqrt(2 * prob) / step;
      theory = 2 * kPi * step * step * (1 - tanh(lambdap) / lambdap);
      if (sflag == -1) {
        ans = theory;
      } else {
        step0 = exp(steplo);
        step1 = exp(steplo + sinc);
        ans = ((step - step0) * ans + (step1 - step) * theory) / (step1 - step0);
      }
    } else if (ans > 4.0 * kPi / 3.0 * prob) {
      ans = 4.0 * kPi / 3.0 * prob;
    }
  } else if (b >= 1.0) {
    logb = log(b);
    bindx = (int)floor((logb - blor) / bincr);
    if (bindx < 0) bindx = 0;
    else if (bindx > bnumr - 2) bindx = bnumr - 2;
    blo = blor + bindx * bincr;
    bhi = blor + (bindx + 1) * bincr;
    denom = (bhi - blo) * (probhi - problo) * (stephi - steplo);
    numlll = (bhi - logb) * (probhi - prob) * (stephi - logstep);
    numllh = (bhi - logb) * (probhi - prob) * (logstep - steplo);
    numlhl = (bhi - logb) * (prob - problo) * (stephi - logstep);
    numlhh = (bhi - logb) * (prob - problo) * (logstep - steplo);
    numhll = (logb - blo) * (probhi - prob) * (stephi - logstep);
    numhlh = (logb - blo) * (probhi - prob) * (logstep - steplo);
    numhhl = (logb - blo) * (prob - problo) * (stephi - logstep);
    numhhh = (logb - blo) * (prob - problo) * (logstep - steplo);
    klll = krev[bindx * snum * pnum + pindx * snum + sindx];
    kllh = krev[bindx * snum * pnum + pindx * snum + sindx + 1];
    klhl = krev[bindx * snum * pnum + (pindx + 1) * snum + sindx];
    klhh = krev[bindx * snum * pnum + (pindx + 1) * snum + sindx + 1];
    khll = krev[(bindx + 1) * snum * pnum + pindx * snum + sindx];
    khlh = krev[(bindx + 1) * snum * pnum + pindx * snum + sindx + 1];
    khhl = krev[(bindx + 1) * snum * pnum + (pindx + 1) * snum + sindx];
    khhh = krev[(bindx + 1) * snum * pnum + (pindx + 1) * snum + sindx + 1];
    ans = (numlll * klll + numllh * kllh + numlhl * klhl + numlhh * klhh + numhll * khll +
           numhlh * khlh + numhhl * khhl + numhhh * khhh) /
          denom;
    if (sindx == 0) {
      lambdap = sqrt(2 * prob) / step;
      if (lambdap < 10)
        theory = 2 * kPi * step * step * b * (lambdap * cosh(lambdap) - sinh(lambdap)) /
                 (lambdap * (b - 1) * cosh(lambdap) + sinh(lambdap));
      else
        theory = 2 * kPi * step * step * b * (lambdap - 1) / (lambdap * (b - 1) + 1);
      if (sflag == -1) {
        ans = theory;
      } else {
        step0 = exp(steplo);
        step1 = exp(steplo + sinc);
        ans = ((step - step0) * ans + (step1 - step) * theory) / (step1 - step0);
      }
    } else if (ans > 4.0 * kPi / 3.0 * prob) {
      ans = 4.0 * kPi / 3.0 * prob;
    }
  } else {
    bindx = (int)floor((b - blob) / bincb);
    if (bindx < 0) bindx = 0;
    else if (bindx > bnumb - 2) bindx = bnumb - 2;
    blo = blob + bindx * bincb;
    bhi = blob + (bindx + 1) * bincb;
    while (kb[bindx * snum * pnum + (pindx + 1) * snum + sindx + 1] == -1) sindx--;
    steplo = shi - sindx * sinc;
    stephi = shi - (sindx + 1) * sinc;
    denom = (bhi - blo) * (probhi - problo) * (stephi - steplo);
    numlll = (bhi - b) * (probhi - prob) * (stephi - logstep);
    numllh = (bhi - b) * (probhi - prob) * (logstep - steplo);
    numlhl = (bhi - b) * (prob - problo) * (stephi - logstep);
    numlhh = (bhi - b) * (prob - problo) * (logstep - steplo);
    numhll = (b - blo) * (probhi - prob) * (stephi - logstep);
    numhlh = (b - blo) * (probhi - prob) * (logstep - steplo);
    numhhl = (b - blo) * (prob - problo) * (stephi - logstep);
    numhhh = (b - blo) * (prob - problo) * (logstep - steplo);
    klll = kb[bindx * snum * pnum + pindx * snum + sindx];
    kllh = kb[bindx * snum * pnum + pindx * snum + sindx + 1];
    klhl = kb[bindx * snum * pnum + (pindx + 1) * snum + sindx];
    klhh = kb[bindx * snum * pnum + (pindx + 1) * snum + sindx + 1];
    khll = kb[(bindx + 1) * snum * pnum + pindx * snum + sindx];
    khlh = kb[(bindx + 1) * snum * pnum + pindx * snum + sindx + 1];
    khhl = kb[(bindx + 1) * snum * pnum + (pindx + 1) * snum + sindx];
    khhh = kb[(bindx + 1) * snum * pnum + (pindx + 1) * snum + sindx + 1];
    ans = (numlll * klll + numllh * kllh + numlhl * klhl + numlhh * klhh + numhll * khll +
           numhlh * khlh + numhhl * khhl + numhhh * khhh) /
          denom;
  }
  ans *= a * a * a;
  if (ans > ansp1) ans = ansp1;
  return ans;
}

double actrxnrate(double step, double a, double prob) {
  double ka;
  if (step < 0 || a <= 0) return -1;
  if (step == 0) return 0;
  step /= a;
  ka = 4.0 * kPi / 3.0 * (erfcc(sqrt(2.0) / step) + step * sqrt(2.0 / kPi));
  ka += 2.0 * kSqrt2Pi / 3.0 * step * (step * step - 1.0) * (exp(-2.0 / step / step) - 1.0);
  return prob * ka * a * a * a;
}

double bindingradius(double rate, double dt, double difc, double b, int rel) {
  double a, lo, dif, step;
  int n;
  if (rate < 0 || dt < 0 || difc <= 0) return -1;
  if (rate == 0) return 0;
  if (dt == 0) {
    if (b < 0) return rate / (4 * kPi * difc);
    if (rel && b > 1) return rate * (1 - 1 / b) / (4 * kPi * difc);
    if (rel && b <= 1) return -1;
    if (b > 0) return rate / (4 * kPi * difc + rate / b);
    return -1;
  }
  step = sqrt(2.0 * difc * dt);
  lo = 0;
  a = step;
  while (numrxnrate(step, a, rel ? b * a : b) < rate * dt) {
    lo = a;
    a *= 2.0;
  }
  dif = a - lo;
  for (n = 0; n < 15; n++) {
    dif *= 0.5;
    a = lo + dif;
    if (numrxnrate(step, a, rel ? b * a : b) < rate * dt) lo = a;
  }
  a = lo + 0.5 * dif;
  return a;
}

static double chitry_of(double rate, double difc, double a, double unbindrad) {
  return (unbindrad >= 0) ? rate / (4 * kPi * difc * a * unbindrad / (unbindrad - a))
                          : rate / (4 * kPi * difc * a);
}

double bindingradiusprob(double rate, double dt, double difc, double b, int rel, double chi,
                              double* probptr) {
  double a, prob, step, lambdap, lambda;
  double xtry[2], xmult[2], xlo[2], xhi[2], errtry, errold, chitry, ratetry, errbest, abest, pbest,
      unbindrad;
  int it, indx;

  if (rate < 0 || dt < 0 || difc <= 0 || chi > 1) return -1;
  if (chi < 0 && (!probptr || *probptr == 1)) return bindingradius(rate, dt, difc, b, rel);
  if (chi < 0 && probptr && *probptr == 0) return 0;
  if (rate == 0) {
    if (probptr && *probptr < 0) *probptr = 0;
    return 0;
  }
  if (dt == 0) {
    if (chi < 0) return -1;
    if (b < 0) a = rate / (4 * kPi * difc * chi);
    else if (rel && b > 1) a = rate * (1 - 1 / b) / (4 * kPi * difc * chi);
    else if (rel && b <= 1) return -1;
    else if (b > 0) a = rate * b / (rate + 4 * kPi * difc * b * chi);
    else return -1;
    lambdap = (3.18243 * chi - 3.40864 * chi * chi + 0.984385 * chi * chi * chi) / pow(1 - chi, 2.08761);
    lambda = difc * lambdap / a * lambdap / a;
    if (probptr) *probptr = lambda;
    return a;
  }

  abest = pbest = 0;
  if (chi < 0) {
    prob = pbest = *probptr;
    step = sqrt(2 * difc * dt);
    xtry[0] = abest = step;
    xmult[0] = 0.1;
    xlo[0] = 1.0e-15;
    xhi[0] = 1.0e15;
    unbindrad = (b >= 0) ? ((rel) ? b * xtry[0] : b) : -1;
    ratetry = numrxnrateprob(step, xtry[0], unbindrad, prob) / dt;
    errtry = errbest = (ratetry / rate - 1.0) * (ratetry / rate - 1.0);
    for (it = 0; it < 200 && errtry > 1.0e-6; it++) {
      errold = errtry;
      xtry[0] = xtry[0] * (1.0 + xmult[0]);
      if (xtry[0] < xlo[0]) xtry[0] = xlo[0];
      else if (xtry[0] > xhi[0]) xtry[0] = xhi[0];
      unbindrad = (b >= 0) ? ((rel) ? b * xtry[0] : b) : -1;
      ratetry = numrxnrateprob(step, xtry[0], unbindrad, prob) / dt;
      errtry = (ratetry / rate - 1.0) * (ratetry / rate - 1.0);
      if (errtry < errold) xmult[0] *= 1.1;
      else xmult[0] *= -0.5;
      if (errtry < errbest) {
        errbest = errtry;
        abest = xtry[0];
      }
    }
  } else {
    step = sqrt(2 * difc * dt);
    xtry[0] = abest = bindingradius(rate, dt, difc, b, rel);
    xtry[1] = pbest = 1;
    ratetry = rate;
    unbindrad = (b >= 0) ? ((rel) ? b * xtry[0] : b) : -1;
    chitry = chitry_of(rate, difc, xtry[0], unbindrad);
    if (chitry > chi) {
      xmult[0] = xmult[1] = 0.1;
      xlo[0] = xlo[1] = 1.0e-15;
      xhi[0] = 1.0e15;
      xhi[1] = 1.0;
      xtry[0] = abest = step;
      xtry[1] = pbest = 0.5;
      ratetry = numrxnrateprob(step, xtry[0], rel ? b * xtry[0] : b, xtry[1]) / dt;
      unbindrad = (b >= 0) ? ((rel) ? b * xtry[0] : b) : -1;
      chitry = chitry_of(rate, difc, xtry[0], unbindrad);
      errtry = errbest = (ratetry / rate - 1.0) * (ratetry / rate - 1.0) + (chitry / chi - 1.0) * (chitry / chi - 1.0);
      for (it = 0; it < 300 && errtry > 1.0e-6; it++) {
        for (indx = 0; indx < 2; indx++) {
          errold = errtry;
          xtry[indx] = xtry[indx] * (1.0 + xmult[indx]);
          if (xtry[indx] < xlo[indx]) xtry[indx] = xlo[indx];
          else if (xtry[indx] > xhi[indx]) xtry[indx] = xhi[indx];
          ratetry = numrxnrateprob(step, xtry[0], rel ? b * xtry[0] : b, xtry[1]) / dt;
          unbindrad = (b >= 0) ? ((rel) ? b * xtry[0] : b) : -1;
          chitry = chitry_of(rate, difc, xtry[0], unbindrad);
          errtry = (ratetry / rate - 1.0) * (ratetry / rate - 1.0) + (chitry / chi - 1.0) * (chitry / chi - 1.0);
          if (errtry < errold) xmult[indx] *= 1.1;
          else xmult[indx] *= -0.5;
          if (errtry < errbest) {
            errbest = errtry;
            abest = xtry[0];
            pbest = xtry[1];
          }
        }
      }
    }
  }
  if (chi >= 0 && probptr) *probptr = pbest;
  return abest;
}

double unbindingradius(double pgem, double dt, double difc, double a) {
  double b, lo, dif, step, ki, kmax;
  int n;
  if (pgem <= 0 || pgem >= 1 || difc <= 0 || a <= 0 || dt < 0) return -2;
  if (dt == 0) return a / pgem;
  step = sqrt(2.0 * difc * dt);
  ki = numrxnrate(step, a, -1);
  kmax = numrxnrate(step, a, 0);
  if (1.0 - ki / kmax < pgem) return ki / kmax - 1.0;
  lo = 0;
  b = a;
  while (1.0 - ki / numrxnrate(step, a, b) > pgem) {
    lo = b;
    b *= 2.0;
  }
  dif = b - lo;
  for (n = 0; n < 15; n++) {
    dif *= 0.5;
    b = lo + dif;
    if (1.0 - ki / numrxnrate(step, a, b) > pgem) lo = b;
  }
  b = lo + 0.5 * dif;
  return b;
}

}  // namespace rxnparam
}  // namespace smolamd
