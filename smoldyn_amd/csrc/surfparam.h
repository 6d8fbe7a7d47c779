// Surface interaction rates -> per-collision probabilities (see surfparam.cpp).
#pragma once

namespace smolamd {
namespace surfparam {

// algorithm codes in the order of the reference's SurfParamAlgo
// (source/libSteve/SurfaceParam.h:11)
enum Algo {
  kIrrTrans, kIrrTransT, kIrrTransQ, kRevTrans, kIrrAds, kIrrAdsT, kIrrAdsQ, kIrrAdsEC,
  kRevAds, kRevAdsND, kIrrDes, kRevDes, kIrrFlip, kRevFlip, kIrrDesC, kRevAdsC
};

double experfcm1(double x);
double lookup_irrev_adsorb(double kappa);
double lookup_rev_ads(double kappa, double kr, double* p2);
// kappas -> probabilities of a two-sided transmitting surface; returns 0, or
// 1 / 2 for the reference's refusals (p1 and p2 then untouched)
int transmit_probs(double kap1, double kap2, double difc1, double difc2, double dt, double* p1, double* p2);
// rate(s) -> probability; -1 for an algorithm not restated (kRevAdsND) or
// unrecognized, as the reference returns for an unrecognized one
double surfaceprob(double k1, double k2, double dt, double difc, double* p2, int algo);

}  // namespace surfparam
}  // namespace smolamd
