// Host side of the rate surrogate: training samples from the reaction-
// parameter solver, bulk inversion and rate checks, and the process-wide
// surrogate that smolUpdateSim consults for reversible reactions whose
// unbinding radius is a ratio of the binding radius (see surrogate.hip).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "rxnparam.h"

namespace smolamd {
namespace surrogate {

constexpr int kIn = 3, kH = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kNParam = oB3 + kOut;

struct State {
  bool on = false;
  std::vector<float> P;
  double in_mean[kIn], in_std[kIn], out_mean[kOut], out_std[kOut];
  double tol = 1e-3;
  long long accepted = 0, rejected = 0;
};
static State g_state;
static std::mutex g_mu;

// FP32 weights, double arithmetic: the standardized network output
static void forward(const std::vector<float>& P, const double* x, double* y) {
  double h1[kH];
  for (int i = 0; i < kH; i++) {
    double a = P[oB1 + i];
    for (int c = 0; c < kIn; c++) a += (double)P[oW1 + i * kIn + c] * x[c];
    h1[i] = tanh(a);
  }
  for (int k = 0; k < kOut; k++) y[k] = P[oB3 + k];
  for (int j = 0; j < kH; j++) {
    double a = P[oB2 + j];
    for (int i = 0; i < kH; i++) a += (double)P[oW2 + j * kH + i] * h1[i];
    double h2 = tanh(a);
    for (int k = 0; k < kOut; k++) y[k] += (double)P[oW3 + k * kH + j] * h2;
  }
}

// relative error of the reference's rate function at binding radius a
static double rate_error(double ratedt, double step, double ratio, double a) {
  double r = rxnparam::numrxnrate(step, a, ratio * a);
  return fabs(r - ratedt) / ratedt;
}

// One secant step on log rate versus log radius: the rate function is
// evaluated at a and at a(1+h), its local power law gives the correction.
static double polish(double ratedt, double step, double ratio, double a) {
  const double h = 1e-3;
  double r0 = rxnparam::numrxnrate(step, a, ratio * a);
  double a1 = a * (1.0 + h);
  double r1 = rxnparam::numrxnrate(step, a1, ratio * a1);
  if (!(r0 > 0) || !(r1 > r0)) return a;
  double p = log(r1 / r0) / log1p(h);
  return a * exp(log(ratedt / r0) / p);
}

// Secant steps from the network's radius: a second one only when the first
// leaves the rate error above tol (rate-function kinks between table points).
static double refine(double ratedt, double step, double ratio, double a, double tol) {
  a = polish(ratedt, step, ratio, a);
  if (!(rate_error(ratedt, step, ratio, a) <= tol)) a = polish(ratedt, step, ratio, a);
  return a;
}

// network features of a query: reduced rate, scale, ratio
static void features(double ratedt, double step, double ratio, double* x) {
  x[0] = log(ratedt) - 3.0 * log(step);
  x[1] = log(step);
  x[2] = ratio;
}

bool enabled() { return g_state.on; }

// Binding radius for rate (per unit time) over time step dt with summed
// diffusion coefficient difc and unbinding/binding ratio: the surrogate's
// value when the reference's rate function confirms it within the tolerance,
// else the numerical inversion (bindingradius, rxnparam.c:558-583).
double binding_radius(double rate, double dt, double difc, double ratio, bool* used) {
  double step = sqrt(2.0 * difc * dt), ratedt = rate * dt;
  if (used) *used = false;
  if (g_state.on && rate > 0 && dt > 0 && difc > 0 && ratio >= 1) {
    double x[kIn], xs[kIn], y[kOut];
    features(ratedt, step, ratio, x);
    for (int c = 0; c < kIn; c++) xs[c] = (x[c] - g_state.in_mean[c]) / g_state.in_std[c];
    forward(g_state.P, xs, y);
    double a = refine(ratedt, step, ratio, step * exp(y[0] * g_state.out_std[0] + g_state.out_mean[0]), g_state.tol);
    if (a > 0 && rate_error(ratedt, step, ratio, a) <= g_state.tol) {
      std::lock_guard<std::mutex> lk(g_mu);
      g_state.accepted++;
      if (used) *used = true;
      return a;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_state.rejected++;
  }
  return rxnparam::bindingradius(rate, dt, difc, ratio, 1);
}

// splitmix64 (Steele, Lea, Flood 2014): a small seeded generator for samples
static uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unif(uint64_t* s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace surrogate
}  // namespace smolamd

using namespace smolamd;

extern "C" {

// Install (params != NULL) or remove the process-wide rate surrogate.  norm:
// input means[3], input standard deviations[3], output means[2], output
// standard deviations[2]; tol: largest accepted relative rate error.
int smolamdSetRateSurrogate(const float* params, const double* norm, double tol) {
  std::lock_guard<std::mutex> lk(surrogate::g_mu);
  surrogate::State& st = surrogate::g_state;
  if (!params) {
    st.on = false;
    return 0;
  }
  st.P.assign(params, params + surrogate::kNParam);
  memcpy(st.in_mean, norm, sizeof(double) * 3);
  memcpy(st.in_std, norm + 3, sizeof(double) * 3);
  memcpy(st.out_mean, norm + 6, sizeof(double) * 2);
  memcpy(st.out_std, norm + 8, sizeof(double) * 2);
  st.tol = tol > 0 ? tol : 1e-3;
  st.accepted = st.rejected = 0;
  st.on = true;
  return 0;
}

int smolamdRateSurrogateStats(long long* accepted, long long* rejected) {
  std::lock_guard<std::mutex> lk(surrogate::g_mu);
  if (accepted) *accepted = surrogate::g_state.accepted;
  if (rejected) *rejected = surrogate::g_state.rejected;
  return 0;
}

// n training samples from the reaction-parameter solver: rms step / binding
// radius log-uniform in [s_lo, s_hi], unbinding/binding ratio uniform in
// [r_lo, r_hi], rms step log-uniform in [step_lo, step_hi]; rate*dt from the
// reference's rate function numrxnrate.  X[n][3] = (log(rate*dt/step^3),
// log step, ratio), Y[n][2] = (log(binding radius/step), log(unbinding
// radius/step)).  Draws where the rate function's table interpolation is not
// positive (unbinding radius within about 0.1% of the binding radius at small
// step) are drawn again; the return value is the number of such redraws.
long smolamdRxnparamSamples(long n, unsigned long long seed, const double* ranges, double* X, double* Y) {
  double s_lo = ranges[0], s_hi = ranges[1], r_lo = ranges[2], r_hi = ranges[3], st_lo = ranges[4],
         st_hi = ranges[5];
  uint64_t s = seed;
  long redraws = 0;
  for (long k = 0; k < n; k++) {
    double soa = exp(log(s_lo) + surrogate::unif(&s) * (log(s_hi) - log(s_lo)));
    double ratio = r_lo + surrogate::unif(&s) * (r_hi - r_lo);
    double step = exp(log(st_lo) + surrogate::unif(&s) * (log(st_hi) - log(st_lo)));
    double a = step / soa, b = ratio * a;
    double ratedt = rxnparam::numrxnrate(step, a, b);
    if (!(ratedt > 0)) {
      redraws++;
      k--;
      continue;
    }
    surrogate::features(ratedt, step, ratio, X + 3 * k);
    Y[2 * k] = log(a / step);
    Y[2 * k + 1] = log(b / step);
  }
  return redraws;
}

// Numerical inversion (the reference's bindingradius) for n queries.
int smolamdRxnparamInvert(long n, const double* ratedt, const double* step, const double* ratio, double* a) {
  for (long k = 0; k < n; k++) {
    double dt = 1.0, difc = step[k] * step[k] / 2.0;  // step = sqrt(2 difc dt)
    a[k] = rxnparam::bindingradius(ratedt[k] / dt, dt, difc, ratio[k], 1);
  }
  return 0;
}

// Network features of n queries (see smolamdRxnparamSamples).
int smolamdRxnparamFeatures(long n, const double* ratedt, const double* step, const double* ratio, double* X) {
  for (long k = 0; k < n; k++) surrogate::features(ratedt[k], step[k], ratio[k], X + 3 * k);
  return 0;
}

// Secant correction of binding radii a_in: one step (two rate-function
// evaluations), and a second one where the first leaves the relative rate
// error above tol.
int smolamdRxnparamPolish(long n, const double* ratedt, const double* step, const double* ratio, const double* a_in,
                          double tol, double* a_out) {
  for (long k = 0; k < n; k++) a_out[k] = surrogate::refine(ratedt[k], step[k], ratio[k], a_in[k], tol);
  return 0;
}

// Relative error of the reference's rate function at binding radii a[k].
int smolamdRxnparamRateError(long n, const double* ratedt, const double* step, const double* ratio, const double* a,
                             double* relerr) {
  for (long k = 0; k < n; k++) relerr[k] = surrogate::rate_error(ratedt[k], step[k], ratio[k], a[k]);
  return 0;
}

}  // extern "C"
