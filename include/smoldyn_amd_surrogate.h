/* Learned surrogate for the reaction-parameter inversion of libsmoldyn_amd.
 *
 * No counterpart in the reference: the reference computes the binding radius
 * of a reversible reaction by bisection on its numerical rate function
 * (source/libSteve/rxnparam.c:558-583, bindingradius).  These calls train a
 * 3-64-64-2 MLP on samples of that rate function with the library's own GPU
 * kernels, evaluate it in BF16 on the GPU, and install it as the first guess
 * of the setup, with the rate function as the judge: a surrogate radius is
 * used only when the rate function at that radius is within the tolerance of
 * the requested rate, else the numerical inversion runs.
 *
 * Network inputs (standardized by the caller): log(rate*dt/step^3) (the
 * reduced rate), log(rms step), unbinding/binding radius ratio.  Outputs
 * (standardized): log(binding radius/step), log(unbinding radius/step).  The
 * setup refines the network's radius by one secant step on the rate function
 * (two where the first leaves the error above the tolerance) before judging it.  Parameter vector: W1[64][3] b1[64] W2[64][64] b2[64]
 * W3[2][64] b3[2] (float32).
 */
#ifndef SMOLDYN_AMD_SURROGATE_H
#define SMOLDYN_AMD_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

int smolamdSurrogateParamCount(void);
void* smolamdSurrogateNew(const float* params);
void smolamdSurrogateFree(void* net);
int smolamdSurrogateGetParams(void* net, float* params);
/* gradient of the mean squared error over n samples X[n][3], T[n][2] */
int smolamdSurrogateGradient(void* net, const float* X, const float* T, int n, float* grad, double* loss);
/* Adam on the device; perm holds epochs*nsamples sample indices */
int smolamdSurrogateTrain(void* net, const float* X, const float* T, long nsamples, const int* perm, int epochs,
                          int batch, float lr, float lr_final, double* loss_per_epoch);
/* inference X[n][3] -> Y[n][2]; bf16 != 0: BF16 kernel; *ms: kernel time */
int smolamdSurrogatePredict(void* net, const float* X, int n, float* Y, int bf16, double* ms);

/* install (params != NULL) or remove the surrogate used by smolUpdateSim for
   reversible reactions with ratio product placement; norm = input means[3],
   input sds[3], output means[2], output sds[2]; tol = accepted relative rate
   error */
int smolamdSetRateSurrogate(const float* params, const double* norm, double tol);
int smolamdRateSurrogateStats(long long* accepted, long long* rejected);

/* rxnparam in bulk: training samples, numerical inversion, rate error check.
   ranges = rms step/binding radius [lo,hi], ratio [lo,hi], rms step [lo,hi];
   the sampler returns how many draws it replaced because the rate function
   was not positive there */
long smolamdRxnparamSamples(long n, unsigned long long seed, const double* ranges, double* X, double* Y);
int smolamdRxnparamInvert(long n, const double* ratedt, const double* step, const double* ratio, double* a);
int smolamdRxnparamFeatures(long n, const double* ratedt, const double* step, const double* ratio, double* X);
/* secant correction of radii a_in: one step, a second where the first leaves
   the relative rate error above tol */
int smolamdRxnparamPolish(long n, const double* ratedt, const double* step, const double* ratio, const double* a_in,
                          double tol, double* a_out);
int smolamdRxnparamRateError(long n, const double* ratedt, const double* step, const double* ratio, const double* a,
                             double* relerr);

#ifdef __cplusplus
}
#endif
#endif
