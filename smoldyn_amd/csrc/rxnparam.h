// Reaction-parameter solver of the host setup (see rxnparam.cpp).
#pragma once

namespace smolamd {
namespace rxnparam {
double erfcc(double x);
double modelrxnrate(double a, double b, double difc, double chi);
double numrxnrate(double step, double a, double b);
double numrxnrateprob(double step, double a, double b, double prob);
double actrxnrate(double step, double a, double prob);
double bindingradius(double rate, double dt, double difc, double b, int rel);
double bindingradiusprob(double rate, double dt, double difc, double b, int rel, double chi, double* probptr);
double unbindingradius(double pgem, double dt, double difc, double a);
}  // namespace rxnparam
}  // namespace smolamd
