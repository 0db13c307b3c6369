// Host side of the far-field surrogate: device buffers, normalisation,
// training loop (forward, backward, weight gradients, Adam) and inference,
// exported through the qgpu_mlp_* functions of include/qgpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qgpu.h"
#include "mlp.hpp"
#include "mlp_host.hpp"

namespace qgpu {

__global__ void k_mlp_fwd(MlpDev, const float*, const int*, int, int, float*, int);
__global__ void k_mlp_bwd(MlpDev, const float*, const int*, int, int, const float*, float*);
__global__ void k_mlp_wgrad(MlpDev, int, float*);
__global__ void k_mlp_adam(MlpDev, const float*, float, float, float);
__global__ void k_mlp_sync(MlpDev);

#define MCHECK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static T* dalloc(size_t n) {
  void* p = nullptr;
  MCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  MCHECK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

Mlp::Mlp(int seed) {
  MCHECK(hipGetDevice(&device));
  MCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  D.P = dalloc<float>(kMlpNParam);
  D.adam_m = dalloc<float>(kMlpNParam);
  D.adam_v = dalloc<float>(kMlpNParam);
  D.W1b = dalloc<__bf16>(kMlpH * kMlpIn);
  D.W2b = dalloc<__bf16>(kMlpH * kMlpH);
  D.W2Tb = dalloc<__bf16>(kMlpH * kMlpH);
  D.W3b = dalloc<__bf16>(kMlpOut * kMlpH);
  D.b1 = D.P + kMlpOffB1;
  D.b2 = D.P + kMlpOffB2;
  D.b3 = D.P + kMlpOffB3;
  norm = dalloc<float>(14);
  D.xmean = norm; D.xinv = norm + 4; D.ymean = norm + 8; D.yinv = norm + 11;
  ystd_d = dalloc<float>(3);
  grad = dalloc<float>(kMlpNParam);
  // He-style uniform initialisation from a seeded generator
  std::mt19937 rng((unsigned)seed);
  std::vector<float> P(kMlpNParam, 0.0f);
  auto fill = [&](int off, int n, int fan_in) {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    const float a = std::sqrt(3.0f / (float)fan_in);
    for (int k = 0; k < n; ++k) P[off + k] = a * u(rng);
  };
  fill(kMlpOffW1, kMlpH * kMlpIn, kMlpIn);
  fill(kMlpOffW2, kMlpH * kMlpH, kMlpH);
  fill(kMlpOffW3, kMlpOut * kMlpH, kMlpH);
  std::vector<float> nm = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1};
  set_norm(nm);
  set_params(P);
}

Mlp::~Mlp() {
  for (void* p : {(void*)D.P, (void*)D.adam_m, (void*)D.adam_v, (void*)D.W1b, (void*)D.W2b, (void*)D.W2Tb,
                  (void*)D.W3b, (void*)norm, (void*)ystd_d, (void*)grad})
    (void)hipFree(p);
  free_train();
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  if (stream) (void)hipStreamDestroy(stream);
}

void Mlp::free_train() {
  for (void* p : {(void*)D.xT, (void*)D.z1T, (void*)D.z2T, (void*)D.dyT, (void*)D.dz2T, (void*)D.dz1T,
                  (void*)D.h1T, (void*)D.dz2Tb, (void*)pred, (void*)loss_part, (void*)didx})
    if (p) (void)hipFree(p);
  D.xT = D.z1T = D.z2T = D.dyT = D.dz2T = D.dz1T = nullptr;
  D.h1T = D.dz2Tb = nullptr;
  pred = loss_part = nullptr;
  didx = nullptr;
  cap_pad = 0;
}

// Training buffers for up to `n` samples per pass (padded to 64).
void Mlp::ensure_train(int n) {
  const int npad = (n + 63) / 64 * 64;
  if (npad <= cap_pad) return;
  free_train();
  cap_pad = npad;
  D.xT = dalloc<float>((size_t)kMlpIn * npad);
  D.z1T = dalloc<float>((size_t)kMlpH * npad);
  D.z2T = dalloc<float>((size_t)kMlpH * npad);
  D.dyT = dalloc<float>((size_t)kMlpOut * npad);
  D.dz2T = dalloc<float>((size_t)kMlpH * npad);
  D.dz1T = dalloc<float>((size_t)kMlpH * npad);
  D.h1T = dalloc<__bf16>((size_t)kMlpH * npad);
  D.dz2Tb = dalloc<__bf16>((size_t)kMlpH * npad);
  pred = dalloc<float>((size_t)kMlpOut * npad);
  loss_part = dalloc<float>(npad / 16);
  didx = dalloc<int>(npad);
}

void Mlp::set_norm(const std::vector<float>& nm) {
  // nm: xmean[4], xstd[4], ymean[3], ystd[3]; stored as mean / 1/std
  std::vector<float> d(14);
  for (int i = 0; i < 4; ++i) { d[i] = nm[i]; d[4 + i] = 1.0f / nm[4 + i]; }
  for (int i = 0; i < 3; ++i) { d[8 + i] = nm[8 + i]; d[11 + i] = 1.0f / nm[11 + i]; }
  MCHECK(hipMemcpy(norm, d.data(), 14 * sizeof(float), hipMemcpyHostToDevice));
  MCHECK(hipMemcpy(ystd_d, nm.data() + 11, 3 * sizeof(float), hipMemcpyHostToDevice));
  norm_h = nm;
}

void Mlp::set_params(const std::vector<float>& P) {
  if ((int)P.size() != kMlpNParam) throw std::runtime_error("wrong parameter count");
  MCHECK(hipMemcpy(D.P, P.data(), kMlpNParam * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_sync, dim3((kMlpNParam + 255) / 256), dim3(256), 0, stream, D);
  MCHECK(hipGetLastError());
  MCHECK(hipStreamSynchronize(stream));
}

std::vector<float> Mlp::params() {
  std::vector<float> P(kMlpNParam);
  MCHECK(hipStreamSynchronize(stream));
  MCHECK(hipMemcpy(P.data(), D.P, kMlpNParam * sizeof(float), hipMemcpyDeviceToHost));
  return P;
}

// Data on the device; normalisation from its means and standard deviations.
void Mlp::set_data(const float* X, const float* Y, int n) {
  if (n <= 0) throw std::runtime_error("no training samples");
  if (dX) (void)hipFree(dX);
  if (dY) (void)hipFree(dY);
  dX = dalloc<float>((size_t)n * kMlpIn);
  dY = dalloc<float>((size_t)n * kMlpOut);
  MCHECK(hipMemcpy(dX, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice));
  MCHECK(hipMemcpy(dY, Y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice));
  ndata = n;
  std::vector<float> nm(14);
  auto stats = [&](const float* A, int w, int c, float& mean, float& sd) {
    double s = 0, s2 = 0;
    for (int k = 0; k < n; ++k) s += A[(size_t)k * w + c];
    const double m = s / n;
    for (int k = 0; k < n; ++k) { const double d = A[(size_t)k * w + c] - m; s2 += d * d; }
    mean = (float)m;
    sd = (float)std::max(std::sqrt(s2 / n), 1e-6);
  };
  for (int i = 0; i < 4; ++i) stats(X, kMlpIn, i, nm[i], nm[4 + i]);
  for (int i = 0; i < 3; ++i) stats(Y, kMlpOut, i, nm[8 + i], nm[11 + i]);
  set_norm(nm);
}

// One forward + backward pass over samples idx (device array) or 0..n-1 of
// (X, Y); gradient in `grad`; returns the mean squared error (normalised).
double Mlp::pass(const float* X, const float* Y, const int* idx, int n) {
  ensure_train(n);
  const int npad = (n + 63) / 64 * 64;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(npad / 64), dim3(256), 0, stream, D, X, idx, n, npad, pred, 1);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(npad / 64), dim3(256), 0, stream, D, Y, idx, n, npad, pred, loss_part);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(64 + (kMlpOther + 3) / 4), dim3(256), 0, stream, D, npad, grad);
  MCHECK(hipGetLastError());
  std::vector<float> lp(npad / 16);
  MCHECK(hipMemcpyAsync(lp.data(), loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
  MCHECK(hipStreamSynchronize(stream));
  double s = 0;
  for (float v : lp) s += v;
  return s / (3.0 * n);
}

double Mlp::train(int epochs, int batch, float lr, int seed) {
  if (!dX) throw std::runtime_error("no training data");
  if (batch < 64) throw std::runtime_error("batch must be at least 64");
  std::vector<int> perm(ndata);
  for (int k = 0; k < ndata; ++k) perm[k] = k;
  std::mt19937 rng((unsigned)seed);
  double last = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double tot = 0;
    int cnt = 0;
    for (int b0 = 0; b0 < ndata; b0 += batch) {
      const int nb = std::min(batch, ndata - b0);
      ensure_train(nb);
      MCHECK(hipMemcpyAsync(didx, perm.data() + b0, nb * sizeof(int), hipMemcpyHostToDevice, stream));
      tot += pass(dX, dY, didx, nb) * nb;
      cnt += nb;
      ++adam_t;
      const float b1t = (float)std::pow(0.9, (double)adam_t), b2t = (float)std::pow(0.999, (double)adam_t);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpNParam + 255) / 256), dim3(256), 0, stream, D, grad, lr, b1t, b2t);
      MCHECK(hipGetLastError());
    }
    last = tot / cnt;
  }
  MCHECK(hipStreamSynchronize(stream));
  return last;
}

// Normalised predictions for n samples (device arrays).
void Mlp::predict_dev(const float* X, int n, float* Y, hipStream_t s) {
  const int npad = (n + 63) / 64 * 64;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(npad / 64), dim3(256), 0, s, D, X, (const int*)nullptr, n, npad, Y, 0);
  MCHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Gradient pass without device allocation.  The training activations and the
// staging copies of X and Y live in `scratch`, a static device array of
// every GPU the library runs on (held until the process ends), for up to
// kNoAllocMaxSamples samples.  Host code never uses the array's name as an
// address: the address of each device's copy comes from hipGetSymbolAddress
// (scratch_ptr) and is what the kernels receive.  Concurrent calls on one GPU
// share the array and race; every buffer a pass reads is written by the
// same pass first (padded samples included), so nothing needs clearing.
constexpr int kNoAllocMaxSamples = 2048;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
// bytes for npad samples: X, Y staging; xT, z1T, z2T, dyT, dz2T, dz1T, pred,
// loss_part (FP32); h1T, dz2Tb (BF16)
constexpr size_t scratch_bytes(size_t npad) {
  return align_up(npad * kMlpIn * 4) + align_up(npad * kMlpOut * 4) + align_up(npad * kMlpIn * 4) +
         4 * align_up(npad * kMlpH * 4) + align_up(npad * kMlpOut * 4) + align_up(npad * kMlpOut * 4) +
         align_up(npad / 16 * 4) + 2 * align_up(npad * kMlpH * 2);
}
constexpr size_t kScratchBytes = scratch_bytes(kNoAllocMaxSamples);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of this GPU's copy of `scratch`, looked up on the first
// call on each device (with that device current).
static unsigned char* scratch_ptr(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static unsigned char* ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptr[device] = static_cast<unsigned char*>(p);
  }
  return ptr[device];
}

// Same kernels, launch shapes and sample order as Mlp::pass over X, Y
// (host arrays, n samples); gradient to G (host), returns the loss.
static double pass_no_alloc(Mlp& m, const float* X, const float* Y, int n, float* G) {
  if (n <= 0) throw std::runtime_error("no samples");
  if (n > kNoAllocMaxSamples)
    throw std::runtime_error("qgpu_mlp_gradient_noalloc: at most " + std::to_string(kNoAllocMaxSamples) +
                             " samples per call");
  int dev = -1;
  MCHECK(hipGetDevice(&dev));
  if (dev != m.device) throw std::runtime_error("the current device is not the model's device");
  unsigned char* base = scratch_ptr(dev);
  const size_t npad = (size_t)(n + 63) / 64 * 64;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += align_up(bytes);
    return p;
  };
  float* dx = reinterpret_cast<float*>(take(npad * kMlpIn * 4));
  float* dy = reinterpret_cast<float*>(take(npad * kMlpOut * 4));
  MlpDev D = m.D;
  D.xT = reinterpret_cast<float*>(take(npad * kMlpIn * 4));
  D.z1T = reinterpret_cast<float*>(take(npad * kMlpH * 4));
  D.z2T = reinterpret_cast<float*>(take(npad * kMlpH * 4));
  D.dz2T = reinterpret_cast<float*>(take(npad * kMlpH * 4));
  D.dz1T = reinterpret_cast<float*>(take(npad * kMlpH * 4));
  D.dyT = reinterpret_cast<float*>(take(npad * kMlpOut * 4));
  float* pred = reinterpret_cast<float*>(take(npad * kMlpOut * 4));
  float* loss_part = reinterpret_cast<float*>(take(npad / 16 * 4));
  D.h1T = reinterpret_cast<__bf16*>(take(npad * kMlpH * 2));
  D.dz2Tb = reinterpret_cast<__bf16*>(take(npad * kMlpH * 2));
  if (off > kScratchBytes) throw std::runtime_error("scratch layout exceeds the array");
  MCHECK(hipMemcpyAsync(dx, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m.stream));
  MCHECK(hipMemcpyAsync(dy, Y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice, m.stream));
  hipLaunchKernelGGL(k_mlp_fwd, dim3(npad / 64), dim3(256), 0, m.stream, D, dx, (const int*)nullptr, n, (int)npad,
                     pred, 1);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(npad / 64), dim3(256), 0, m.stream, D, dy, (const int*)nullptr, n, (int)npad,
                     pred, loss_part);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(64 + (kMlpOther + 3) / 4), dim3(256), 0, m.stream, D, (int)npad, m.grad);
  MCHECK(hipGetLastError());
  float lp[kNoAllocMaxSamples / 16];
  MCHECK(hipMemcpyAsync(lp, loss_part, npad / 16 * sizeof(float), hipMemcpyDeviceToHost, m.stream));
  MCHECK(hipMemcpyAsync(G, m.grad, kMlpNParam * sizeof(float), hipMemcpyDeviceToHost, m.stream));
  MCHECK(hipStreamSynchronize(m.stream));
  double s = 0;
  for (size_t k = 0; k < npad / 16; ++k) s += lp[k];
  return s / (3.0 * n);
}

}  // namespace qgpu

// ---------------------------------------------------------------------------
// C ABI
namespace {
thread_local std::string g_mlp_err;
template <typename F>
int mguard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_mlp_err = e.what();
  } catch (...) {
    g_mlp_err = "unknown error";
  }
  return 1;
}
}  // namespace

extern "C" {

int qgpu_mlp_create(int seed, qgpu_mlp** out) {
  return mguard([&] {
    *out = nullptr;
    auto* h = new qgpu_mlp{new qgpu::Mlp(seed)};
    *out = h;
  });
}

int qgpu_mlp_destroy(qgpu_mlp* h) {
  return mguard([&] {
    if (!h) return;
    delete h->m;
    delete h;
  });
}

int qgpu_mlp_nparams(int* n) {
  *n = qgpu::kMlpNParam;
  return 0;
}

int qgpu_mlp_get_params(qgpu_mlp* h, float* params, float* norm) {
  return mguard([&] {
    auto P = h->m->params();
    std::memcpy(params, P.data(), P.size() * sizeof(float));
    if (norm) std::memcpy(norm, h->m->norm_h.data(), 14 * sizeof(float));
  });
}

int qgpu_mlp_set_params(qgpu_mlp* h, const float* params, const float* norm) {
  return mguard([&] {
    h->m->set_params(std::vector<float>(params, params + qgpu::kMlpNParam));
    if (norm) h->m->set_norm(std::vector<float>(norm, norm + 14));
  });
}

int qgpu_mlp_set_data(qgpu_mlp* h, const float* X, const float* Y, int n) {
  return mguard([&] { h->m->set_data(X, Y, n); });
}

int qgpu_mlp_train(qgpu_mlp* h, int epochs, int batch, float lr, int seed, double* loss) {
  return mguard([&] { *loss = h->m->train(epochs, batch, lr, seed); });
}

int qgpu_mlp_gradient(qgpu_mlp* h, const float* X, const float* Y, int n, float* grad, double* loss) {
  return mguard([&] {
    float *dx = qgpu::dalloc<float>((size_t)n * qgpu::kMlpIn), *dy = qgpu::dalloc<float>((size_t)n * qgpu::kMlpOut);
    MCHECK(hipMemcpy(dx, X, (size_t)n * qgpu::kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(dy, Y, (size_t)n * qgpu::kMlpOut * sizeof(float), hipMemcpyHostToDevice));
    *loss = h->m->pass(dx, dy, nullptr, n);
    MCHECK(hipMemcpy(grad, h->m->grad, qgpu::kMlpNParam * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(dx);
    (void)hipFree(dy);
  });
}

int qgpu_mlp_gradient_noalloc(qgpu_mlp* h, const float* X, const float* Y, int n, float* grad, double* loss) {
  return mguard([&] {
    if (!h || !X || !Y || !grad || !loss) throw std::runtime_error("null argument");
    *loss = qgpu::pass_no_alloc(*h->m, X, Y, n, grad);
  });
}

int qgpu_mlp_predict(qgpu_mlp* h, const float* X, int n, float* Y) {
  return mguard([&] {
    float *dx = qgpu::dalloc<float>((size_t)n * qgpu::kMlpIn), *dy = qgpu::dalloc<float>((size_t)n * qgpu::kMlpOut);
    MCHECK(hipMemcpy(dx, X, (size_t)n * qgpu::kMlpIn * sizeof(float), hipMemcpyHostToDevice));
    h->m->predict_dev(dx, n, dy, h->m->stream);
    MCHECK(hipStreamSynchronize(h->m->stream));
    std::vector<float> y((size_t)n * qgpu::kMlpOut);
    MCHECK(hipMemcpy(y.data(), dy, y.size() * sizeof(float), hipMemcpyDeviceToHost));
    // de-normalise
    for (int k = 0; k < n; ++k)
      for (int o = 0; o < qgpu::kMlpOut; ++o)
        Y[(size_t)k * qgpu::kMlpOut + o] = y[(size_t)k * qgpu::kMlpOut + o] * h->m->norm_h[11 + o] + h->m->norm_h[8 + o];
    (void)hipFree(dx);
    (void)hipFree(dy);
  });
}

const char* qgpu_mlp_last_error(void) { return g_mlp_err.c_str(); }

}  // extern "C"
