// Learned far-field surrogate: a 4-128-128-3 SiLU MLP mapping per-charge-
// group multipole features (charge, radial and tangential dipole, radial
// position in the sphere) to the far-field potential and field at the group
// centre, with hand-written CDNA4 kernels for inference and training.
//
// The 128x128 layer runs on BF16 matrix cores (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation); the 4->128 and 128->3 layers are VALU.  Weights are
// kept as FP32 master copies for the optimizer and as BF16 copies for the
// forward and backward products.  Every reduction is done in a fixed order,
// so training is bitwise reproducible.
//
// MFMA 16x16x32 lane layout (lane l, k8 = 8 * (l / 16)):
//   A[i][k]: i = l % 16, k = k8 .. k8+7      (8 bf16 per lane)
//   B[k][j]: j = l % 16, k = k8 .. k8+7
//   D[i][j]: j = l % 16, i = 4 * (l / 16) + 0..3   (4 floats per lane)
#include "mlp.hpp"

namespace qgpu {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// ---------------------------------------------------------------------------
// Forward pass, one wavefront per 16 samples (sample idx[s] of X when idx is
// given, else s).  Writes normalised outputs Y (n x 3).  With T != nullptr
// (training) it also keeps, feature-major over the padded sample range, the
// normalised inputs xT (4 x npad), pre-activations z1T/z2T (128 x npad) and
// the BF16 layer-1 activations h1T (128 x npad).
__global__ __launch_bounds__(256) void k_mlp_fwd(MlpDev M, const float* __restrict__ X, const int* __restrict__ idx,
                                                 int n, int npad, float* __restrict__ Y, int train) {
  __shared__ __attribute__((aligned(16))) __bf16 h1s[4][16][kMlpH];
  __shared__ float h2s[4][16][kMlpH];
  // npad is a multiple of 64 and the grid has npad / 64 blocks, so every
  // wavefront of every block has 16 samples in range (no early exit before
  // the barriers)
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int s0 = (blockIdx.x * 4 + w) * 16;
  // layer 1: lane -> sample s0 + l%16, outputs l/16 + 4m
  {
    const int s = l % 16, gs = s0 + s;
    float xn[kMlpIn];
    if (gs < n) {
      const int src = idx ? idx[gs] : gs;
      for (int i = 0; i < kMlpIn; ++i) xn[i] = (X[(size_t)src * kMlpIn + i] - M.xmean[i]) * M.xinv[i];
    } else {
      for (int i = 0; i < kMlpIn; ++i) xn[i] = 0.0f;
    }
    if (train && l < 16)
      for (int i = 0; i < kMlpIn; ++i) M.xT[(size_t)i * npad + gs] = xn[i];
    for (int m = 0; m < kMlpH / 4; ++m) {
      const int o = l / 16 + 4 * m;
      float z = M.b1[o];
      for (int i = 0; i < kMlpIn; ++i) z += (float)M.W1b[o * kMlpIn + i] * xn[i];
      if (gs >= n) z = 0.0f;
      const __bf16 hb = (__bf16)silu(z);
      h1s[w][s][o] = gs < n ? hb : (__bf16)0.0f;
      if (train) {
        M.z1T[(size_t)o * npad + gs] = z;
        M.h1T[(size_t)o * npad + gs] = gs < n ? hb : (__bf16)0.0f;
      }
    }
  }
  __syncthreads();
  // layer 2 on matrix cores: z2[s][o] = sum_k h1[s][k] W2[o][k] + b2[o]
  const int k8 = 8 * (l / 16);
  for (int nt = 0; nt < kMlpH / 16; ++nt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int kb = 0; kb < kMlpH / 32; ++kb) {
      const bf16x8 a = load8(&h1s[w][l % 16][kb * 32 + k8]);
      const bf16x8 b = load8(M.W2b + (size_t)(nt * 16 + l % 16) * kMlpH + kb * 32 + k8);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int o = nt * 16 + l % 16;
    for (int ii = 0; ii < 4; ++ii) {
      const int s = 4 * (l / 16) + ii, gs = s0 + s;
      const float z = acc[ii] + M.b2[o];
      h2s[w][s][o] = gs < n ? silu(z) : 0.0f;
      if (train) M.z2T[(size_t)o * npad + gs] = gs < n ? z : 0.0f;
    }
  }
  __syncthreads();
  // layer 3: lane l < 48 -> sample l % 16, output l / 16
  if (l < 16 * kMlpOut) {
    const int s = l % 16, o = l / 16, gs = s0 + s;
    float y = M.b3[o];
    for (int k = 0; k < kMlpH; ++k) y += (float)M.W3b[o * kMlpH + k] * h2s[w][s][k];
    if (gs < n) Y[(size_t)gs * kMlpOut + o] = y;
  }
}

// ---------------------------------------------------------------------------
// Backward through the activations, one wavefront per 16 samples: with the
// loss L = sum over samples and outputs of (y - t)^2 / (3 n), writes dyT
// (3 x npad), dz2T (FP32 and BF16, 128 x npad), dz1T (128 x npad) and one
// partial squared error per wavefront.
__global__ __launch_bounds__(256) void k_mlp_bwd(MlpDev M, const float* __restrict__ T, const int* __restrict__ idx,
                                                 int n, int npad, const float* __restrict__ Y,
                                                 float* __restrict__ loss_part) {
  __shared__ __attribute__((aligned(16))) __bf16 dz2s[4][16][kMlpH];
  __shared__ float dys[4][16][kMlpOut];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + w;
  const int s0 = wave * 16;   // in range for every wavefront (grid npad / 64)
  // output error
  float se = 0.0f;
  if (l < 16 * kMlpOut) {
    const int s = l % 16, o = l / 16, gs = s0 + s;
    float dy = 0.0f;
    if (gs < n) {
      const int src = idx ? idx[gs] : gs;
      const float t = (T[(size_t)src * kMlpOut + o] - M.ymean[o]) * M.yinv[o];
      const float e = Y[(size_t)gs * kMlpOut + o] - t;
      se = e * e;
      dy = 2.0f * e / (3.0f * (float)n);
    }
    dys[w][s][o] = dy;
    M.dyT[(size_t)o * npad + gs] = dy;
  }
  for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
  if (l == 0) loss_part[wave] = se;
  __syncthreads();
  // dh2 = W3^T dy, dz2 = dh2 * silu'(z2): lane -> sample l%16, units l/16 + 4m
  {
    const int s = l % 16, gs = s0 + s;
    for (int m = 0; m < kMlpH / 4; ++m) {
      const int k = l / 16 + 4 * m;
      float dh = 0.0f;
      for (int o = 0; o < kMlpOut; ++o) dh += (float)M.W3b[o * kMlpH + k] * dys[w][s][o];
      const float dz = gs < n ? dh * dsilu(M.z2T[(size_t)k * npad + gs]) : 0.0f;
      dz2s[w][s][k] = (__bf16)dz;
      M.dz2T[(size_t)k * npad + gs] = dz;
      M.dz2Tb[(size_t)k * npad + gs] = (__bf16)dz;
    }
  }
  __syncthreads();
  // dh1[s][k] = sum_o dz2[s][o] W2[o][k] on matrix cores (B = W2 via W2^T rows)
  const int k8 = 8 * (l / 16);
  for (int kt = 0; kt < kMlpH / 16; ++kt) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ob = 0; ob < kMlpH / 32; ++ob) {
      const bf16x8 a = load8(&dz2s[w][l % 16][ob * 32 + k8]);
      const bf16x8 b = load8(M.W2Tb + (size_t)(kt * 16 + l % 16) * kMlpH + ob * 32 + k8);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int k = kt * 16 + l % 16;
    for (int ii = 0; ii < 4; ++ii) {
      const int s = 4 * (l / 16) + ii, gs = s0 + s;
      M.dz1T[(size_t)k * npad + gs] = gs < n ? acc[ii] * dsilu(M.z1T[(size_t)k * npad + gs]) : 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// Weight gradients.  Blocks 0..63: one 16x16 tile of dW2 per block (wave 0),
// dW2[o][k] = sum_s dz2[s][o] h1[s][k] on matrix cores over 32-sample chunks
// (A = dz2^T rows, B = h1 via h1^T rows).  Remaining blocks: one parameter
// of W1, b1, b2, W3, b3 per wavefront, its lanes summing interleaved
// samples, then a fixed butterfly over the lanes (so the sum order is fixed
// and training stays bitwise reproducible).
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpDev M, int npad, float* __restrict__ G) {
  if (blockIdx.x < 64) {
    if (threadIdx.x >= 64) return;
    const int l = threadIdx.x, ot = blockIdx.x / 8, kt = blockIdx.x % 8, k8 = 8 * (l / 16);
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = 0; c < npad; c += 32) {
      const bf16x8 a = load8(M.dz2Tb + (size_t)(ot * 16 + l % 16) * npad + c + k8);
      const bf16x8 b = load8(M.h1T + (size_t)(kt * 16 + l % 16) * npad + c + k8);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    for (int ii = 0; ii < 4; ++ii)
      G[kMlpOffW2 + (ot * 16 + 4 * (l / 16) + ii) * kMlpH + kt * 16 + l % 16] = acc[ii];
    return;
  }
  const int l = threadIdx.x & 63;
  const int p = (blockIdx.x - 64) * (blockDim.x >> 6) + (int)(threadIdx.x >> 6);
  if (p >= kMlpOther) return;   // the whole wavefront
  float g = 0.0f;
  int dst;
  if (p < kMlpH * kMlpIn) {                       // W1[o][i]
    const int o = p / kMlpIn, i = p % kMlpIn;
    for (int s = l; s < npad; s += 64) g += M.dz1T[(size_t)o * npad + s] * M.xT[(size_t)i * npad + s];
    dst = kMlpOffW1 + p;
  } else if (p < kMlpH * kMlpIn + kMlpH) {        // b1
    const int o = p - kMlpH * kMlpIn;
    for (int s = l; s < npad; s += 64) g += M.dz1T[(size_t)o * npad + s];
    dst = kMlpOffB1 + o;
  } else if (p < kMlpH * kMlpIn + 2 * kMlpH) {    // b2
    const int o = p - kMlpH * kMlpIn - kMlpH;
    for (int s = l; s < npad; s += 64) g += M.dz2T[(size_t)o * npad + s];
    dst = kMlpOffB2 + o;
  } else if (p < kMlpH * kMlpIn + 2 * kMlpH + kMlpOut * kMlpH) {   // W3[o][k]
    const int q = p - kMlpH * kMlpIn - 2 * kMlpH, o = q / kMlpH, k = q % kMlpH;
    for (int s = l; s < npad; s += 64) g += M.dyT[(size_t)o * npad + s] * silu(M.z2T[(size_t)k * npad + s]);
    dst = kMlpOffW3 + q;
  } else {                                        // b3
    const int o = p - kMlpH * kMlpIn - 2 * kMlpH - kMlpOut * kMlpH;
    for (int s = l; s < npad; s += 64) g += M.dyT[(size_t)o * npad + s];
    dst = kMlpOffB3 + o;
  }
  for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off, 64);
  if (l == 0) G[dst] = g;
}

// Adam step on the FP32 master parameters; refreshes the BF16 copies.
__global__ void k_mlp_adam(MlpDev M, const float* __restrict__ G, float lr, float b1t, float b2t) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpNParam) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float g = G[p];
  const float m = beta1 * M.adam_m[p] + (1.0f - beta1) * g;
  const float v = beta2 * M.adam_v[p] + (1.0f - beta2) * g * g;
  M.adam_m[p] = m;
  M.adam_v[p] = v;
  const float mh = m / (1.0f - b1t), vh = v / (1.0f - b2t);
  const float w = M.P[p] - lr * mh / (sqrtf(vh) + eps);
  M.P[p] = w;
  mlp_refresh_copy(M, p, w);
}

// BF16 copies of all parameters from the FP32 masters.
__global__ void k_mlp_sync(MlpDev M) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpNParam) mlp_refresh_copy(M, p, M.P[p]);
}

}  // namespace qgpu
