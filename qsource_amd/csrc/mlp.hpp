// Far-field surrogate MLP: sizes, parameter layout and the device view used
// by the inference and training kernels (kernels_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace qgpu {

constexpr int kMlpIn = 4;      // q, dipole radial, dipole tangential, r / R
constexpr int kMlpH = 128;     // two hidden layers of 128 SiLU units
constexpr int kMlpOut = 3;     // potential, radial field, tangential field
// FP32 master parameters, one flat array: W1 [H][In], b1 [H], W2 [H][H]
// (row = output unit), b2 [H], W3 [Out][H], b3 [Out]
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpH * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpOut * kMlpH;
constexpr int kMlpNParam = kMlpOffB3 + kMlpOut;
// parameters other than W2 (their gradients are per-wavefront sums)
constexpr int kMlpOther = kMlpH * kMlpIn + 2 * kMlpH + kMlpOut * kMlpH + kMlpOut;

struct MlpDev {
  float* P;                          // FP32 master parameters
  float *adam_m, *adam_v;
  __bf16 *W1b, *W2b, *W2Tb, *W3b;    // BF16 copies (W2Tb = W2 transposed)
  const float *b1, *b2, *b3;         // biases (FP32, inside P)
  const float *xmean, *xinv, *ymean, *yinv;   // input / output normalisation
  // training buffers, feature-major over npad samples
  float *xT, *z1T, *z2T, *dyT, *dz2T, *dz1T;
  __bf16 *h1T, *dz2Tb;
};

__device__ __forceinline__ void mlp_refresh_copy(const MlpDev& M, int p, float w) {
  if (p >= kMlpOffW1 && p < kMlpOffB1) {
    M.W1b[p - kMlpOffW1] = (__bf16)w;
  } else if (p >= kMlpOffW2 && p < kMlpOffB2) {
    const int q = p - kMlpOffW2, o = q / kMlpH, k = q % kMlpH;
    M.W2b[q] = (__bf16)w;
    M.W2Tb[k * kMlpH + o] = (__bf16)w;
  } else if (p >= kMlpOffW3 && p < kMlpOffB3) {
    M.W3b[p - kMlpOffW3] = (__bf16)w;
  }
}

}  // namespace qgpu
