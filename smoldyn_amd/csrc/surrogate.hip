// Learned surrogate for the reaction-parameter inversion (gfx950).
//
// The reference turns a reversible reaction's rate into a binding radius by
// bisection on its numerical rate function (bindingradius, rxnparam.c:558-583,
// about 20-40 evaluations of numrxnrate per inversion; the unbinding radius is
// the given ratio times it).  The surrogate is a 3-64-64-2 MLP (tanh hidden
// layers) that maps standardized (log rate*dt, log rms step, unbinding/binding
// ratio) to standardized (log binding radius, log unbinding radius).  It is
// trained here, on the GPU, by the kernels below (FP32 forward, per-sample
// backward, deterministic weight-gradient reductions, Adam), and evaluated
// with BF16 weights and activations.  The numerical solver stays the judge:
// callers accept a surrogate radius only when the reference's rate function
// at that radius is within the requested relative error of the target rate.
//
// Parameter vector (float32): W1[64][3] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

namespace smolamd {
namespace mlp {

constexpr int kIn = 3, kH = 64, kOut = 2;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kOut * kH, kNParam = oB3 + kOut;  // 4546
constexpr int kThreads = 256;

// Forward pass, one thread per sample, weights staged in LDS.  Inputs and
// activations are stored feature-major ([feature][n]) so the weight-gradient
// reductions read contiguous rows.
__global__ __launch_bounds__(kThreads) void k_forward(const float* __restrict__ P, const float* __restrict__ XT, int n,
                                                      float* __restrict__ H1T, float* __restrict__ H2T,
                                                      float* __restrict__ YT) {
  __shared__ float w[kNParam];
  for (int k = threadIdx.x; k < kNParam; k += blockDim.x) w[k] = P[k];
  __syncthreads();
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn];
#pragma unroll
  for (int c = 0; c < kIn; c++) x[c] = XT[c * n + s];
  float h1[kH];
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float a = w[oB1 + i];
#pragma unroll
    for (int c = 0; c < kIn; c++) a += w[oW1 + i * kIn + c] * x[c];
    h1[i] = tanhf(a);
    if (H1T) H1T[i * n + s] = h1[i];
  }
  float y[kOut];
#pragma unroll
  for (int k = 0; k < kOut; k++) y[k] = w[oB3 + k];
  for (int j = 0; j < kH; j++) {
    float a = w[oB2 + j];
#pragma unroll
    for (int i = 0; i < kH; i++) a += w[oW2 + j * kH + i] * h1[i];
    float h2 = tanhf(a);
    if (H2T) H2T[j * n + s] = h2;
#pragma unroll
    for (int k = 0; k < kOut; k++) y[k] += w[oW3 + k * kH + j] * h2;
  }
#pragma unroll
  for (int k = 0; k < kOut; k++) YT[k * n + s] = y[k];
}

// Per-sample backward pass of the mean squared error L = sum_{s,k} (y-t)^2 / n:
// dY = 2 (y - t) / n, D2 = (W3^T dY) (1 - h2^2), D1 = (W2^T D2) (1 - h1^2).
__global__ __launch_bounds__(kThreads) void k_backward(const float* __restrict__ P, const float* __restrict__ H1T,
                                                       const float* __restrict__ H2T, const float* __restrict__ YT,
                                                       const float* __restrict__ TT, int n, float* __restrict__ DYT,
                                                       float* __restrict__ D2T, float* __restrict__ D1T,
                                                       float* __restrict__ sqerr) {
  __shared__ float w[kNParam];
  for (int k = threadIdx.x; k < kNParam; k += blockDim.x) w[k] = P[k];
  __syncthreads();
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float dy[kOut], e2 = 0;
#pragma unroll
  for (int k = 0; k < kOut; k++) {
    float d = YT[k * n + s] - TT[k * n + s];
    e2 += d * d;
    dy[k] = 2.0f * d / (float)n;
    DYT[k * n + s] = dy[k];
  }
  sqerr[s] = e2;
  float d2[kH];
#pragma unroll
  for (int j = 0; j < kH; j++) {
    float h2 = H2T[j * n + s];
    float g = 0;
#pragma unroll
    for (int k = 0; k < kOut; k++) g += w[oW3 + k * kH + j] * dy[k];
    d2[j] = g * (1.0f - h2 * h2);
    D2T[j * n + s] = d2[j];
  }
  for (int i = 0; i < kH; i++) {
    float h1 = H1T[i * n + s];
    float g = 0;
#pragma unroll
    for (int j = 0; j < kH; j++) g += w[oW2 + j * kH + i] * d2[j];
    D1T[i * n + s] = g * (1.0f - h1 * h1);
  }
}

// Weight gradients: one thread per parameter and sample chunk (blockIdx.y),
// summing its chunk's samples in order into GP[chunk][p]; k_wsum adds the
// chunks in order, so the result does not depend on scheduling.  Rows of the
// feature-major arrays are read contiguously.
constexpr int kMaxChunks = 32;
__global__ __launch_bounds__(kThreads) void k_wgrad(const float* __restrict__ XT, const float* __restrict__ H1T,
                                                    const float* __restrict__ H2T, const float* __restrict__ D1T,
                                                    const float* __restrict__ D2T, const float* __restrict__ DYT, int n,
                                                    int chunk, float* __restrict__ GP) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  const float *a, *b = nullptr;
  if (p < oB1) {
    a = D1T + (p / kIn) * n;
    b = XT + (p % kIn) * n;
  } else if (p < oW2) {
    a = D1T + (p - oB1) * n;
  } else if (p < oB2) {
    int q = p - oW2;
    a = D2T + (q / kH) * n;
    b = H1T + (q % kH) * n;
  } else if (p < oW3) {
    a = D2T + (p - oB2) * n;
  } else if (p < oB3) {
    int q = p - oW3;
    a = DYT + (q / kH) * n;
    b = H2T + (q % kH) * n;
  } else {
    a = DYT + (p - oB3) * n;
  }
  int s0 = blockIdx.y * chunk, s1 = min(n, s0 + chunk);
  float acc = 0;
  if (b) {
    for (int s = s0; s < s1; s++) acc += a[s] * b[s];
  } else {
    for (int s = s0; s < s1; s++) acc += a[s];
  }
  GP[blockIdx.y * kNParam + p] = acc;
}

__global__ __launch_bounds__(kThreads) void k_wsum(const float* __restrict__ GP, int nchunk, float* __restrict__ G) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float acc = 0;
  for (int c = 0; c < nchunk; c++) acc += GP[c * kNParam + p];
  G[p] = acc;
}

// Adam (Kingma & Ba, 2015), bias-corrected.
__global__ __launch_bounds__(kThreads) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                                   float* __restrict__ M, float* __restrict__ V, float lr, float b1,
                                                   float b2, float eps, float c1, float c2) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam) return;
  float g = G[p];
  float m = b1 * M[p] + (1.0f - b1) * g;
  float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Gather a minibatch into feature-major arrays: XT[c][i] = X[perm[off+i]][c].
__global__ __launch_bounds__(kThreads) void k_gather(const float* __restrict__ X, const float* __restrict__ T,
                                                     const int* __restrict__ perm, int off, int n,
                                                     float* __restrict__ XT, float* __restrict__ TT) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = perm[off + i];
#pragma unroll
  for (int c = 0; c < kIn; c++) XT[c * n + i] = X[s * kIn + c];
#pragma unroll
  for (int k = 0; k < kOut; k++) TT[k * n + i] = T[s * kOut + k];
}

// BF16: round-to-nearest-even of the FP32 bit pattern, and back.
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// BF16 inference: weights, inputs and activations in BF16, products
// accumulated in FP32.  Sample-major input X[n][3] and output Y[n][2].
__global__ __launch_bounds__(kThreads) void k_infer_bf16(const uint16_t* __restrict__ PB, const float* __restrict__ X,
                                                         int n, float* __restrict__ Y) {
  __shared__ float w[kNParam];
  for (int k = threadIdx.x; k < kNParam; k += blockDim.x) w[k] = bf2f(PB[k]);
  __syncthreads();
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn];
#pragma unroll
  for (int c = 0; c < kIn; c++) x[c] = bf2f(f2bf(X[s * kIn + c]));
  float h1[kH];
#pragma unroll
  for (int i = 0; i < kH; i++) {
    float a = w[oB1 + i];
#pragma unroll
    for (int c = 0; c < kIn; c++) a += w[oW1 + i * kIn + c] * x[c];
    h1[i] = bf2f(f2bf(tanhf(a)));
  }
  float y[kOut];
#pragma unroll
  for (int k = 0; k < kOut; k++) y[k] = w[oB3 + k];
  for (int j = 0; j < kH; j++) {
    float a = w[oB2 + j];
#pragma unroll
    for (int i = 0; i < kH; i++) a += w[oW2 + j * kH + i] * h1[i];
    float h2 = bf2f(f2bf(tanhf(a)));
#pragma unroll
    for (int k = 0; k < kOut; k++) y[k] += w[oW3 + k * kH + j] * h2;
  }
#pragma unroll
  for (int k = 0; k < kOut; k++) Y[s * kOut + k] = y[k];
}

struct Net {
  float *P = nullptr, *G = nullptr, *GP = nullptr, *M = nullptr, *V = nullptr;
  uint16_t* PB = nullptr;
  float *XT = nullptr, *TT = nullptr, *H1T = nullptr, *H2T = nullptr, *YT = nullptr, *DYT = nullptr, *D1T = nullptr,
        *D2T = nullptr, *sq = nullptr;
  int cap = 0;  // minibatch capacity
  long step = 0;
  hipStream_t stream = nullptr;
};

template <typename T>
static bool dal(T** p, size_t n) {
  return hipMalloc((void**)p, (n ? n : 1) * sizeof(T)) == hipSuccess;
}

static bool ensure_batch(Net* m, int n) {
  if (n <= m->cap) return true;
  float** bufs[] = {&m->XT, &m->TT, &m->H1T, &m->H2T, &m->YT, &m->DYT, &m->D1T, &m->D2T, &m->sq};
  for (float** b : bufs)
    if (*b) (void)hipFree(*b), *b = nullptr;
  bool ok = dal(&m->XT, (size_t)kIn * n) && dal(&m->TT, (size_t)kOut * n) && dal(&m->H1T, (size_t)kH * n) &&
            dal(&m->H2T, (size_t)kH * n) && dal(&m->YT, (size_t)kOut * n) && dal(&m->DYT, (size_t)kOut * n) &&
            dal(&m->D1T, (size_t)kH * n) && dal(&m->D2T, (size_t)kH * n) && dal(&m->sq, (size_t)n);
  m->cap = ok ? n : 0;
  return ok;
}

// forward + backward + weight gradients of one minibatch already in XT/TT
static void fwd_bwd(Net* m, int n) {
  int nb = (n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_forward, dim3(nb), dim3(kThreads), 0, m->stream, m->P, m->XT, n, m->H1T, m->H2T, m->YT);
  hipLaunchKernelGGL(k_backward, dim3(nb), dim3(kThreads), 0, m->stream, m->P, m->H1T, m->H2T, m->YT, m->TT, n, m->DYT,
                     m->D2T, m->D1T, m->sq);
  int nchunk = (n + 127) / 128;
  if (nchunk > kMaxChunks) nchunk = kMaxChunks;
  int chunk = (n + nchunk - 1) / nchunk;
  nchunk = (n + chunk - 1) / chunk;
  int pb = (kNParam + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_wgrad, dim3(pb, nchunk), dim3(kThreads), 0, m->stream, m->XT, m->H1T, m->H2T, m->D1T, m->D2T,
                     m->DYT, n, chunk, m->GP);
  hipLaunchKernelGGL(k_wsum, dim3(pb), dim3(kThreads), 0, m->stream, m->GP, nchunk, m->G);
}

}  // namespace mlp
}  // namespace smolamd

using namespace smolamd::mlp;

extern "C" {

int smolamdSurrogateParamCount(void) { return kNParam; }

void* smolamdSurrogateNew(const float* params) {
  Net* m = new Net();
  if (!dal(&m->P, kNParam) || !dal(&m->G, kNParam) || !dal(&m->GP, (size_t)kNParam * kMaxChunks) ||
      !dal(&m->M, kNParam) || !dal(&m->V, kNParam) ||
      !dal(&m->PB, kNParam) || hipStreamCreate(&m->stream) != hipSuccess) {
    delete m;
    return nullptr;
  }
  (void)hipMemcpy(m->P, params, sizeof(float) * kNParam, hipMemcpyHostToDevice);
  (void)hipMemset(m->M, 0, sizeof(float) * kNParam);
  (void)hipMemset(m->V, 0, sizeof(float) * kNParam);
  return m;
}

void smolamdSurrogateFree(void* h) {
  Net* m = (Net*)h;
  if (!m) return;
  void* ptrs[] = {m->P, m->G, m->GP, m->M, m->V, m->PB, m->XT, m->TT, m->H1T, m->H2T, m->YT, m->DYT, m->D1T, m->D2T, m->sq};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int smolamdSurrogateGetParams(void* h, float* params) {
  Net* m = (Net*)h;
  if (hipStreamSynchronize(m->stream) != hipSuccess) return -1;
  return hipMemcpy(params, m->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

// One forward/backward pass over n samples (X[n][3], T[n][2], standardized):
// the gradient of the mean squared error into grad[kNParam]; returns the loss
// through *loss.
int smolamdSurrogateGradient(void* h, const float* X, const float* T, int n, float* grad, double* loss) {
  Net* m = (Net*)h;
  if (n <= 0 || !ensure_batch(m, n)) return -1;
  std::vector<float> xt((size_t)kIn * n), tt((size_t)kOut * n);
  for (int s = 0; s < n; s++) {
    for (int c = 0; c < kIn; c++) xt[(size_t)c * n + s] = X[(size_t)s * kIn + c];
    for (int k = 0; k < kOut; k++) tt[(size_t)k * n + s] = T[(size_t)s * kOut + k];
  }
  (void)hipMemcpy(m->XT, xt.data(), sizeof(float) * xt.size(), hipMemcpyHostToDevice);
  (void)hipMemcpy(m->TT, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice);
  fwd_bwd(m, n);
  if (hipStreamSynchronize(m->stream) != hipSuccess) return -1;
  (void)hipMemcpy(grad, m->G, sizeof(float) * kNParam, hipMemcpyDeviceToHost);
  if (loss) {
    std::vector<float> sq(n);
    (void)hipMemcpy(sq.data(), m->sq, sizeof(float) * n, hipMemcpyDeviceToHost);
    double acc = 0;
    for (float v : sq) acc += v;
    *loss = acc / n;
  }
  return 0;
}

// Adam training on the device: nsamples standardized samples, minibatches of
// `batch` in a per-epoch order given by perm (epochs * nsamples indices).
// loss_per_epoch[e] receives the mean minibatch loss of epoch e.
int smolamdSurrogateTrain(void* h, const float* X, const float* T, long nsamples, const int* perm, int epochs,
                          int batch, float lr, float lr_final, double* loss_per_epoch) {
  Net* m = (Net*)h;
  if (nsamples <= 0 || batch <= 0 || batch > nsamples || !ensure_batch(m, batch)) return -1;
  float *dX = nullptr, *dT = nullptr;
  int* dperm = nullptr;
  if (!dal(&dX, (size_t)kIn * nsamples) || !dal(&dT, (size_t)kOut * nsamples) || !dal(&dperm, (size_t)nsamples)) {
    if (dX) (void)hipFree(dX);
    if (dT) (void)hipFree(dT);
    return -1;
  }
  (void)hipMemcpy(dX, X, sizeof(float) * kIn * nsamples, hipMemcpyHostToDevice);
  (void)hipMemcpy(dT, T, sizeof(float) * kOut * nsamples, hipMemcpyHostToDevice);
  long nbatch = nsamples / batch;
  long total = (long)epochs * nbatch;
  std::vector<float> sq(batch);
  int rc = 0;
  for (int e = 0; e < epochs && rc == 0; e++) {
    (void)hipMemcpyAsync(dperm, perm + (size_t)e * nsamples, sizeof(int) * nsamples, hipMemcpyHostToDevice, m->stream);
    double eloss = 0;
    for (long b = 0; b < nbatch; b++) {
      long t = (long)e * nbatch + b;
      // learning rate: geometric decay from lr to lr_final over the run
      float lrt = lr * powf(lr_final / lr, (float)t / (float)(total > 1 ? total - 1 : 1));
      hipLaunchKernelGGL(k_gather, dim3((batch + kThreads - 1) / kThreads), dim3(kThreads), 0, m->stream, dX, dT,
                         dperm, (int)(b * batch), batch, m->XT, m->TT);
      fwd_bwd(m, batch);
      m->step++;
      float c1 = 1.0f - powf(0.9f, (float)m->step), c2 = 1.0f - powf(0.999f, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3((kNParam + kThreads - 1) / kThreads), dim3(kThreads), 0, m->stream, m->P, m->G,
                         m->M, m->V, lrt, 0.9f, 0.999f, 1e-8f, c1, c2);
      if (b == nbatch - 1) {  // loss of the last minibatch of the epoch
        if (hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * batch, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
          rc = -1;
        if (hipStreamSynchronize(m->stream) != hipSuccess) rc = -1;
        for (float v : sq) eloss += v;
        eloss /= batch;
      }
    }
    if (loss_per_epoch) loss_per_epoch[e] = eloss;
  }
  if (hipStreamSynchronize(m->stream) != hipSuccess) rc = -1;
  (void)hipFree(dX);
  (void)hipFree(dT);
  (void)hipFree(dperm);
  return rc;
}

// Inference on n standardized inputs X[n][3] into Y[n][2]: bf16 != 0 uses the
// BF16 kernel, else the FP32 forward pass.  *ms receives the device time of
// the inference kernel.
int smolamdSurrogatePredict(void* h, const float* X, int n, float* Y, int bf16, double* ms) {
  Net* m = (Net*)h;
  if (n <= 0) return -1;
  float *dX = nullptr, *dY = nullptr;
  if (!dal(&dX, (size_t)kIn * n) || !dal(&dY, (size_t)kOut * n)) return -1;
  (void)hipMemcpy(dX, X, sizeof(float) * kIn * n, hipMemcpyHostToDevice);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int nb = (n + kThreads - 1) / kThreads;
  if (bf16) {
    std::vector<float> p(kNParam);
    std::vector<uint16_t> pb(kNParam);
    (void)hipMemcpy(p.data(), m->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost);
    for (int k = 0; k < kNParam; k++) pb[k] = f2bf(p[k]);
    (void)hipMemcpy(m->PB, pb.data(), sizeof(uint16_t) * kNParam, hipMemcpyHostToDevice);
    (void)hipEventRecord(e0, m->stream);
    hipLaunchKernelGGL(k_infer_bf16, dim3(nb), dim3(kThreads), 0, m->stream, m->PB, dX, n, dY);
    (void)hipEventRecord(e1, m->stream);
  } else {
    // feature-major through the training forward pass
    std::vector<float> xt((size_t)kIn * n);
    for (int s = 0; s < n; s++)
      for (int c = 0; c < kIn; c++) xt[(size_t)c * n + s] = X[(size_t)s * kIn + c];
    float *dXT = nullptr, *dYT = nullptr;
    if (!dal(&dXT, (size_t)kIn * n) || !dal(&dYT, (size_t)kOut * n)) {
      void* bufs[] = {dXT, dYT, dX, dY};
      for (void* b : bufs)
        if (b) (void)hipFree(b);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return -1;
    }
    (void)hipMemcpy(dXT, xt.data(), sizeof(float) * xt.size(), hipMemcpyHostToDevice);
    (void)hipEventRecord(e0, m->stream);
    hipLaunchKernelGGL(k_forward, dim3(nb), dim3(kThreads), 0, m->stream, m->P, dXT, n, nullptr, nullptr, dYT);
    (void)hipEventRecord(e1, m->stream);
    std::vector<float> yt((size_t)kOut * n);
    (void)hipStreamSynchronize(m->stream);
    (void)hipMemcpy(yt.data(), dYT, sizeof(float) * yt.size(), hipMemcpyDeviceToHost);
    for (int s = 0; s < n; s++)
      for (int k = 0; k < kOut; k++) Y[(size_t)s * kOut + k] = yt[(size_t)k * n + s];
    (void)hipFree(dXT);
    (void)hipFree(dYT);
  }
  int rc = hipStreamSynchronize(m->stream) == hipSuccess ? 0 : -1;
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  if (ms) *ms = t;
  if (bf16) (void)hipMemcpy(Y, dY, sizeof(float) * kOut * n, hipMemcpyDeviceToHost);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dX);
  (void)hipFree(dY);
  return rc;
}

}  // extern "C"
