// Far-field kernels on a batch of instances: per-charge-group features of
// the learned surrogate, its training targets (the exact far-field potential
// and field at each group centre), the LRF Taylor estimate at the current
// centres, and the application of the surrogate's output to the atoms.
//
// Frame of a group: e_r = unit vector from the solvent centre to the group
// centre, t = unit vector of the dipole's component perpendicular to e_r.
// Features: net charge, dipole along e_r, dipole along t, |centre|/rwat.
// Outputs:  potential, field along e_r, field along t at the centre.
#include "kernels_common.hpp"

namespace qgpu {

// Centre (mean of the group's atoms, as cgp_centers; the whole group's for
// the slices of a split group), charge and dipole of the group's (slice's)
// non-Q atoms, and the group frame.
__device__ void far_group_frame(const DevSys& S, int g, double c[3], double& q, double er[3], double t[3],
                                double& pr, double& pt, double& r) {
  const int f = S.cgp_first[g], l = S.cgp_last[g];
  const int cf = S.cgp_cfirst[g], cl = S.cgp_clast[g];
  c[0] = c[1] = c[2] = 0.0;
  for (int a = cf; a <= cl; ++a) {
    const int i = S.cgpatom[a];
    c[0] = c[0] + S.x[3 * i]; c[1] = c[1] + S.x[3 * i + 1]; c[2] = c[2] + S.x[3 * i + 2];
  }
  const double nat = (double)(cl - cf + 1);
  c[0] = c[0] / nat; c[1] = c[1] / nat; c[2] = c[2] / nat;
  double p[3] = {0.0, 0.0, 0.0};
  q = 0.0;
  for (int a = f; a <= l; ++a) {
    const int i = S.cgpatom[a];
    if (S.flags[i] & F_Q) continue;
    q = q + S.crg[i];
    for (int k = 0; k < 3; ++k) p[k] = p[k] + S.crg[i] * (S.x[3 * i + k] - c[k]);
  }
  double d[3] = {c[0] - S.xwcent[0], c[1] - S.xwcent[1], c[2] - S.xwcent[2]};
  r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (r > 1e-12) { er[0] = d[0] / r; er[1] = d[1] / r; er[2] = d[2] / r; }
  else { er[0] = 0.0; er[1] = 0.0; er[2] = 1.0; }
  pr = p[0] * er[0] + p[1] * er[1] + p[2] * er[2];
  double v[3] = {p[0] - pr * er[0], p[1] - pr * er[1], p[2] - pr * er[2]};
  pt = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (pt < 1e-12) {
    // any unit vector perpendicular to e_r
    const double a[3] = {fabs(er[0]) < 0.9 ? 1.0 : 0.0, fabs(er[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    const double ad = a[0] * er[0] + a[1] * er[1] + a[2] * er[2];
    v[0] = a[0] - ad * er[0]; v[1] = a[1] - ad * er[1]; v[2] = a[2] - ad * er[2];
    const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    t[0] = v[0] / vn; t[1] = v[1] / vn; t[2] = v[2] / vn;
  } else {
    t[0] = v[0] / pt; t[1] = v[1] / pt; t[2] = v[2] / pt;
  }
}

// Features (4 per group) into X at the instance's offset, and the frame
// (e_r, t, centre: 9 per group) into the frame buffer.
__global__ void k_far_features(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  double c[3], er[3], t[3], q, pr, pt, r;
  far_group_frame(S, g, c, q, er, t, pr, pt, r);
  float* X = S.far_x + (size_t)(S.far_off + g) * 4;
  X[0] = (float)q; X[1] = (float)pr; X[2] = (float)pt; X[3] = (float)(r / S.rwat);
  double* F = S.far_frame + (size_t)(S.far_off + g) * 9;
  for (int k = 0; k < 3; ++k) { F[k] = er[k]; F[3 + k] = t[k]; F[6 + k] = c[k]; }
}

// Exact far-field at each group centre from the freshly built LRF
// multipoles (phi0 and phi1 are sums over every far atom at the centre):
// potential, field along e_r and along t into Y (3 per group).
__global__ void k_far_targets(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  const double* L = S.lrf + (size_t)g * kLrfStride;
  const double* F = S.far_frame + (size_t)(S.far_off + g) * 9;
  float* Y = S.far_y + (size_t)(S.far_off + g) * 3;
  Y[0] = (float)L[3];
  Y[1] = (float)(L[4] * F[0] + L[5] * F[1] + L[6] * F[2]);
  Y[2] = (float)(L[4] * F[3] + L[5] * F[4] + L[6] * F[5]);
}

// LRF estimate at the current centres from the multipoles of the last list
// build (the Taylor expansion lrf_taylor applies, src/md.f90:5843-5888):
// potential and field projected on the group frame, into Y.
__global__ void k_far_lrf_eval(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  const double* L = S.lrf + (size_t)g * kLrfStride;
  const double* F = S.far_frame + (size_t)(S.far_off + g) * 9;
  double df[3];
  const double V = lrf_taylor_atom(L, F + 6, df);
  float* Y = S.far_y + (size_t)(S.far_off + g) * 3;
  Y[0] = (float)V;
  Y[1] = (float)(df[0] * F[0] + df[1] * F[1] + df[2] * F[2]);
  Y[2] = (float)(df[0] * F[3] + df[1] * F[4] + df[2] * F[5]);
}

// Surrogate output (normalised; de-normalised with ymean, ystd) applied to
// the group's non-Q, non-excluded atoms to first order about the centre:
// V_a = phi - E.(x_a - c), E_far += q_a V_a / 2, d_a -= q_a E.
__global__ void k_far_apply(const DevSys* __restrict__ sys, const float* __restrict__ ymean,
                            const float* __restrict__ ystd) {
  const DevSys& S = sys[blockIdx.y];
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  const float* Y = S.far_y + (size_t)(S.far_off + g) * 3;
  const double* F = S.far_frame + (size_t)(S.far_off + g) * 9;
  const double phi = (double)(Y[0] * ystd[0] + ymean[0]);
  const double Er = (double)(Y[1] * ystd[1] + ymean[1]);
  const double Et = (double)(Y[2] * ystd[2] + ymean[2]);
  const double E[3] = {Er * F[0] + Et * F[3], Er * F[1] + Et * F[4], Er * F[2] + Et * F[5]};
  double e = 0.0;
  for (int a = S.cgp_first[g]; a <= S.cgp_last[g]; ++a) {
    const int i = S.cgpatom[a];
    if (S.flags[i] & (F_EXCL | F_Q)) continue;
    const double dx[3] = {S.x[3 * i] - F[6], S.x[3 * i + 1] - F[7], S.x[3 * i + 2] - F[8]};
    const double V = phi - (E[0] * dx[0] + E[1] * dx[1] + E[2] * dx[2]);
    e = e + 0.5 * S.crg[i] * V;
    for (int k = 0; k < 3; ++k) S.d[3 * i + k] = S.d[3 * i + k] - S.crg[i] * E[k];
  }
  S.e_far[g] = e;
}

}  // namespace qgpu
