// Host object of the far-field surrogate MLP (mlp_host.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mlp.hpp"

namespace qgpu {

class Mlp {
 public:
  explicit Mlp(int seed);
  ~Mlp();
  void set_data(const float* X, const float* Y, int n);
  // Adam epochs over shuffled mini-batches; returns the last epoch's mean
  // squared error of the normalised outputs
  double train(int epochs, int batch, float lr, int seed);
  // forward + backward over n samples (device arrays); gradient in `grad`
  double pass(const float* X, const float* Y, const int* idx, int n);
  void predict_dev(const float* X, int n, float* Y, hipStream_t s);
  void set_params(const std::vector<float>& P);
  std::vector<float> params();
  void set_norm(const std::vector<float>& nm);

  MlpDev D{};
  hipStream_t stream = nullptr;
  int device = 0;                   // device the model's buffers and stream live on
  std::vector<float> norm_h;        // xmean[4] xstd[4] ymean[3] ystd[3]
  float* norm = nullptr;            // device: xmean, 1/xstd, ymean, 1/ystd
  float* ystd_d = nullptr;          // device: ystd (for de-normalising)
  float* grad = nullptr;

 private:
  void ensure_train(int n);
  void free_train();
  float *dX = nullptr, *dY = nullptr;
  int ndata = 0;
  int cap_pad = 0;
  float *pred = nullptr, *loss_part = nullptr;
  int* didx = nullptr;
  long adam_t = 0;
};

}  // namespace qgpu

// The C ABI's opaque model handle.
struct qgpu_mlp {
  qgpu::Mlp* m;
};
