// Device helpers shared by the engine's kernels: wavefront/workgroup
// reductions, the charge-group parity ownership rule, the LRF multipole
// accumulator and Taylor evaluation, and the water-dipole geometry of the
// polarisation restraint.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qgpu {

constexpr double kPiD = 3.14159265358979323846;

__device__ __forceinline__ double psqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float psqrt(float v) { return sqrtf(v); }

// Inverse distance powers of one atom pair from the squared distance r2sq:
// returns 1/r^2 and sets rinv = 1/r.  FP64 keeps the reference's
// operations (one/r2, then sqrt, src/md.f90:15074-15080); FP32 uses the
// hardware reciprocal square root (v_rsq_f32, 1 ulp).
__device__ __forceinline__ double inv_r2(double r2sq, double& rinv) {
  double r2 = 1.0 / r2sq;
  rinv = sqrt(r2);
  return r2;
}
__device__ __forceinline__ float inv_r2(float r2sq, float& rinv) {
  rinv = __builtin_amdgcn_rsqf(r2sq);
  return rinv * rinv;
}

// Coordinates read by the pair arithmetic: FP64 x, or its FP32 copy
// A device array read through the global address space: pointers held in
// DevSys are generic to the compiler, which then emits flat loads that wait
// on both the LDS and the vector-memory counters; global loads do not.
template <typename T>
using gconst = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gconst<T> gptr(const T* p) {
  return (gconst<T>)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gptr_w(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Instance and first item of this workgroup.  MAPPED = 0: the 2-D grid
// (blockIdx.y = instance, blockIdx.x * items_per_block = first item).
// MAPPED = 1: one work-map entry per workgroup of a 1-D grid (instance,
// first item; instance < 0 for an empty slot), which lets the host place
// each instance's workgroups on the XCDs it chooses and skip empty ones.
template <int MAPPED>
__device__ __forceinline__ bool block_work(const DevSys* sys, const int2* map, int items_per_block,
                                           const DevSys*& S, int& base) {
  if constexpr (MAPPED != 0) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return false;
    S = sys + m.x;
    base = m.y;
  } else {
    S = sys + blockIdx.y;
    base = (int)blockIdx.x * items_per_block;
  }
  return true;
}

// FP32 coordinates (O, H1, H2) of water w from its 64-byte record.
typedef float f32v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void water_xf(const DevSys& S, int w, float* o) {
  const gconst<f32v4> r = gptr(reinterpret_cast<const f32v4*>(S.xw)) + 4 * (size_t)w;
  const f32v4 a = r[0], b = r[1], c = r[2];
  o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
  o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
  o[8] = c.x;
}
// The record of water w from its FP64 coordinates x (O, H1, H2).
__device__ __forceinline__ void water_xf_store(const DevSys& S, int w, const double* x) {
  f32v4* r = reinterpret_cast<f32v4*>(S.xw) + 4 * (size_t)w;
  r[0] = f32v4{(float)x[0], (float)x[1], (float)x[2], (float)x[3]};
  r[1] = f32v4{(float)x[4], (float)x[5], (float)x[6], (float)x[7]};
  r[2] = f32v4{(float)x[8], 0.0f, 0.0f, 0.0f};
}

template <typename P>
__device__ __forceinline__ const P* pair_x(const DevSys& S);
template <>
__device__ __forceinline__ const double* pair_x<double>(const DevSys& S) { return S.x; }
template <>
__device__ __forceinline__ const float* pair_x<float>(const DevSys& S) { return S.xf; }

// Per-solute-atom LJ products with the water atoms, in the pair precision
template <typename P>
__device__ __forceinline__ const P* qw_rec(const DevSys& S);
template <>
__device__ __forceinline__ const double* qw_rec<double>(const DevSys& S) { return S.qw_rec_d; }
template <>
__device__ __forceinline__ const float* qw_rec<float>(const DevSys& S) { return S.qw_rec_f; }

template <typename P>
__device__ __forceinline__ const P* solute_water_lj(const DevSys& S);
template <>
__device__ __forceinline__ const double* solute_water_lj<double>(const DevSys& S) { return S.swp_d; }
template <>
__device__ __forceinline__ const float* solute_water_lj<float>(const DevSys& S) { return S.swp_f; }

// c - a*b and the squared norm: FP64 as separate operations in the
// reference's order (d1*d1 + d2*d2 + d3*d3), FP32 with fused multiply-add
__device__ __forceinline__ double pmsub(double a, double b, double c) { return c - a * b; }
__device__ __forceinline__ float pmsub(float a, float b, float c) { return __builtin_fmaf(-a, b, c); }
__device__ __forceinline__ double pnorm2(double a, double b, double c) { return a * a + b * b + c * c; }
__device__ __forceinline__ float pnorm2(float a, float b, float c) {
  return __builtin_fmaf(c, c, __builtin_fmaf(b, b, a * a));
}

// Sum over an aligned group of 8 lanes; every lane of the group returns the
// same total
__device__ __forceinline__ double seg8_sum(double v) {
  v += __shfl_xor(v, 1, 64);
  v += __shfl_xor(v, 2, 64);
  v += __shfl_xor(v, 4, 64);
  return v;
}
// the same over an aligned group of N lanes (N a power of two <= 64)
template <int N>
__device__ __forceinline__ double segn_sum(double v) {
  for (int o = 1; o < N; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Failure handling.  A failure found by a launch of phase p at step s is
// recorded as the key s * 16 + p (the smallest key wins).  A kernel that
// changes the simulation state returns at once when a launch before it
// recorded a failure, so the state stops where the failing step found it.
// Within one launch a key is never below that launch's own key, so all of
// its workgroups take the same decision.
__device__ __forceinline__ unsigned long long fail_key(const DevSys& S, int phase) {
  return (unsigned long long)(unsigned)*S.step_ctr * 16ull + (unsigned long long)phase;
}
__device__ __forceinline__ bool halted(const DevSys& S, int phase) { return *S.fail < fail_key(S, phase); }
__device__ __forceinline__ void record_failure(const DevSys& S, int bit, int phase) {
  atomicOr(S.status, bit);
  atomicCAS(S.status + 1, 0, *S.step_ctr + 1);
  atomicMin(S.fail, fail_key(S, phase));
}

// True on a step whose energy-file frame is recorded (step counter a positive
// multiple of iene since recording started) when the batch also keeps the
// FP64 re-evaluation of its Q energies.
__device__ __forceinline__ bool shadow_record_step(const DevSys& S) {
  if (!S.eseries64 || S.iene <= 0) return false;
  const int st = *S.step_ctr - S.eseries_base;
  return st > 0 && st % S.iene == 0;
}

// Soft-core alpha sc_lookup(iq, column, state); column 0-based over the
// topology types, then natyps + the Q atom.
__device__ __forceinline__ double sc_alpha(const DevSys& S, int iq, int col, int st) {
  return S.sc_lookup[((size_t)iq * S.sc_ncol + col) * S.nstates + st];
}

// One state of a Q-solute pair with Q-atom types (nonbond_qp_qvdw,
// src/md.f90:14150-14172): Q atom iq, partner of type tj and charge cj, LJ
// code lj; r2i = 1/r^2, ri = 1/r, r6h = r^6.  Returns dv (times lambda).
template <typename P>
__device__ __forceinline__ P qp_qvdw_state(const DevSys& S, int iq, int st, int tj, int lj, P cj, P r2i,
                                           P ri, P r6h, P& Vel, P& Va, P& Vb) {
  const int qt = S.qiac[iq * S.nstates + st];
  const int qlj = lj == 2 ? 1 : lj;
  const P aLJ = (P)(S.qavdw[3 * qt + qlj - 1] * S.avdw[tj * 3 + lj - 1]);
  const P bLJ = (P)(S.qbvdw[3 * qt + qlj - 1] * S.bvdw[tj * 3 + lj - 1]);
  Vel = (P)S.qcrg[iq * S.nstates + st] * cj * ri;
  if (lj == 3) Vel = Vel * (P)S.el14_scale;
  const P r6s = r6h + (P)sc_alpha(S, iq, tj, st);
  Va = aLJ / (r6s * r6s);
  Vb = bLJ / r6s;
  return r2i * (-Vel - ((P)12 * Va - (P)6 * Vb) * (r6h / r6s)) * (P)S.lambda[st];
}

// One state of the Q-atom / water-atom pair with Q-atom types
// (nonbond_qw_spc src/md.f90:14495-14512 for the O, nonbond_qw_3atom
// :14831-14838 for any atom ip of an all-atom solvent): LJ code 1 of the Q
// type with the water atom's aLJ_solv(ip, lo), soft-core column lo.
// Returns dv (times lambda).
template <typename P>
__device__ __forceinline__ P qw_qvdw_state(const DevSys& S, int iq, int st, int ip, int lo, P r2i, P ri, P r6h,
                                           P& Vel, P& Va, P& Vb) {
  const int qt = S.qiac[iq * S.nstates + st];
  const P r6s = r6h + (P)sc_alpha(S, iq, lo - 1, st);
  Va = (P)S.qavdw[3 * qt] * (P)S.alj_w[ip][lo - 1] / (r6s * r6s);
  Vb = (P)S.qbvdw[3 * qt] * (P)S.blj_w[ip][lo - 1] / r6s;
  Vel = (P)S.chg_solv[ip] * (P)S.qcrg[iq * S.nstates + st] * ri;
  return r2i * (-Vel - ((P)12 * Va - (P)6 * Vb) * (r6h / r6s)) * (P)S.lambda[st];
}

// 64-lane butterfly sum; every lane returns the total
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NW>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  if (ln == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NW; ++w) t += red[w];
  __syncthreads();
  return t;  // valid in thread 0
}
__device__ __forceinline__ double block_sum_256(double v, double* red) { return block_sum<4>(v, red); }
__device__ __forceinline__ double block_sum_1024(double v, double* red) { return block_sum<16>(v, red); }

// Periodic box: the shift boxlength * nint(d / boxlength) of one component
// of a difference vector (nint rounds half away from zero, as round does).
// FP64 in the reference's form; FP32 from the FP32 copies for the
// mixed-precision pair terms (a listed pair's switching atoms are well
// within half a box apart, so both pick the same image).
__device__ __forceinline__ double pbc_shift(const DevSys& S, double d, int k) {
  return S.boxl[k] * round(d * S.inv_boxl[k]);
}
__device__ __forceinline__ float pbc_shift(const DevSys& S, float d, int k) {
  return S.boxl_f[k] * roundf(d * S.inv_boxl_f[k]);
}
// the minimum image of one component (FP64, as the list builders form it)
__device__ __forceinline__ double pbc_image(const DevSys& S, double d, int k) {
  return d - S.boxl[k] * round(d * S.inv_boxl[k]);
}

// The reference visits a pair of charge groups / waters (a, b), 1-based,
// only when (a < b and a+b even) or (a > b and a+b odd)
// (src/md.f90:7700-7702, 10977-10980).  With 0-based indices the parity of
// a+b is unchanged.
__device__ __forceinline__ bool owns_pair(int a, int b) {
  return (a < b) == (((a + b) & 1) == 0);
}

// Unique multipole components of charges about a centre (lrf_update,
// src/md.f90:5927-5977): phi0 = sum q/r, phi1 = -sum q d/r^3,
// phi2_ab = sum (3 q d_a d_b/r^5 - q delta_ab/r^3),
// phi3_abc = sum -3 q (5 d_a d_b d_c - r^2 (delta_ab d_c + delta_ac d_b +
// delta_bc d_a))/r^7, with d = atom - centre; p2 holds (00 01 02 11 12 22),
// p3 holds (000 001 002 011 012 022 111 112 122 222).
template <typename T>
struct LrfTerms {
  T p0, p1[3], p2[6], p3[10];
  __device__ void zero() {
    p0 = 0; for (int k = 0; k < 3; ++k) p1[k] = 0;
    for (int k = 0; k < 6; ++k) p2[k] = 0;
    for (int k = 0; k < 10; ++k) p3[k] = 0;
  }
  // component-wise sum with another record
  __device__ void add(const LrfTerms& o) {
    p0 = p0 + o.p0;
    for (int k = 0; k < 3; ++k) p1[k] = p1[k] + o.p1[k];
    for (int k = 0; k < 6; ++k) p2[k] = p2[k] + o.p2[k];
    for (int k = 0; k < 10; ++k) p3[k] = p3[k] + o.p3[k];
  }
  __device__ void add(T q, T d0, T d1, T d2) {
#pragma clang fp contract(fast)
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    T ri;
    const T ri2 = inv_r2(r2, ri);
    const T f0 = q * ri * ri2;          // q / r^3
    const T f1 = (T)3 * f0 * ri2;       // 3 q / r^5
    const T t = (T)-5 * f1 * ri2;       // 5 * (-3 q / r^7); -3q/r^7 * r^2 = -f1
    p0 = p0 + q * ri;
    p1[0] = p1[0] - f0 * d0; p1[1] = p1[1] - f0 * d1; p1[2] = p1[2] - f0 * d2;
    const T d00 = d0 * d0, d01 = d0 * d1, d02 = d0 * d2, d11 = d1 * d1, d12 = d1 * d2, d22 = d2 * d2;
    p2[0] = p2[0] + (f1 * d00 - f0);
    p2[1] = p2[1] + f1 * d01;
    p2[2] = p2[2] + f1 * d02;
    p2[3] = p2[3] + (f1 * d11 - f0);
    p2[4] = p2[4] + f1 * d12;
    p2[5] = p2[5] + (f1 * d22 - f0);
    const T t00 = t * d00, t01 = t * d01, t11 = t * d11, t22 = t * d22;
    p3[0] = p3[0] + (t00 * d0 + (T)3 * f1 * d0);   // 000
    p3[1] = p3[1] + (t00 * d1 + f1 * d1);          // 001
    p3[2] = p3[2] + (t00 * d2 + f1 * d2);          // 002
    p3[3] = p3[3] + (t11 * d0 + f1 * d0);          // 011
    p3[4] = p3[4] + t01 * d2;                      // 012
    p3[5] = p3[5] + (t22 * d0 + f1 * d0);          // 022
    p3[6] = p3[6] + (t11 * d1 + (T)3 * f1 * d1);   // 111
    p3[7] = p3[7] + (t11 * d2 + f1 * d2);          // 112
    p3[8] = p3[8] + (t22 * d1 + f1 * d1);          // 122
    p3[9] = p3[9] + (t22 * d2 + (T)3 * f1 * d2);   // 222
  }
};

// FP64 accumulator of the 20 unique components
struct LrfAcc : LrfTerms<double> {
  // add a partner group's components computed in FP32
  __device__ void add(const LrfTerms<float>& g) {
    p0 = p0 + (double)g.p0;
    for (int k = 0; k < 3; ++k) p1[k] = p1[k] + (double)g.p1[k];
    for (int k = 0; k < 6; ++k) p2[k] = p2[k] + (double)g.p2[k];
    for (int k = 0; k < 10; ++k) p3[k] = p3[k] + (double)g.p3[k];
  }
  using LrfTerms<double>::add;
  // sum over an aligned group of 8 lanes
  __device__ void seg8_reduce() {
    p0 = seg8_sum(p0);
    for (int k = 0; k < 3; ++k) p1[k] = seg8_sum(p1[k]);
    for (int k = 0; k < 6; ++k) p2[k] = seg8_sum(p2[k]);
    for (int k = 0; k < 10; ++k) p3[k] = seg8_sum(p3[k]);
  }
  template <int N>
  __device__ void segn_reduce() {
    p0 = segn_sum<N>(p0);
    for (int k = 0; k < 3; ++k) p1[k] = segn_sum<N>(p1[k]);
    for (int k = 0; k < 6; ++k) p2[k] = segn_sum<N>(p2[k]);
    for (int k = 0; k < 10; ++k) p3[k] = segn_sum<N>(p3[k]);
  }
  __device__ void wave_reduce() {
    p0 = wave_sum(p0);
    for (int k = 0; k < 3; ++k) p1[k] = wave_sum(p1[k]);
    for (int k = 0; k < 6; ++k) p2[k] = wave_sum(p2[k]);
    for (int k = 0; k < 10; ++k) p3[k] = wave_sum(p3[k]);
  }
  // write phi0, phi1, full phi2(9) and phi3(27) after the centre (L[0..2])
  __device__ void store(double* L) const {
    L[3] = p0;
    L[4] = p1[0]; L[5] = p1[1]; L[6] = p1[2];
    const int i2[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; ++k) L[7 + k] = p2[i2[k]];
    // index of sorted (a<=b<=c) triple in p3
    const int i3[27] = {0, 1, 2, 1, 3, 4, 2, 4, 5,
                        1, 3, 4, 3, 6, 7, 4, 7, 8,
                        2, 4, 5, 4, 7, 8, 5, 8, 9};
    for (int k = 0; k < 27; ++k) L[16 + k] = p3[i3[k]];
  }
};

// Raw moments of charges about a centre, the form in which the
// mixed-precision list kernels gather each partner group in FP32: with a = q/r, f0 = q/r^3, g1 = q/r^5, t = q/r^7,
// p0 = sum a, p1 = sum f0 d, f0s = sum f0, s1 = sum g1 d, m2 = sum g1 d d
// (00 01 02 11 12 22), m3 = sum t d d d (000 001 002 011 012 022 111 112
// 122 222).  They are linear in the charges; the LRF components follow as
// phi1 = -p1, phi2_ab = 3 m2_ab - delta_ab f0s and
// phi3_abc = -15 m3_abc + 3 (delta_ab s1_c + delta_ac s1_b + delta_bc s1_a),
// the same sums as LrfTerms regrouped (fewer operations per atom).
template <typename T>
struct LrfRaw {
  T p0, p1[3], f0s, s1[3], m2[6], m3[10];
  __device__ void zero() {
    p0 = 0; f0s = 0;
    for (int k = 0; k < 3; ++k) { p1[k] = 0; s1[k] = 0; }
    for (int k = 0; k < 6; ++k) m2[k] = 0;
    for (int k = 0; k < 10; ++k) m3[k] = 0;
  }
  __device__ void add(T q, T d0, T d1, T d2) {
#pragma clang fp contract(fast)
    const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
    T ri;
    const T ri2 = inv_r2(r2, ri);
    const T a = q * ri, f0 = a * ri2, g1 = f0 * ri2, t = g1 * ri2;
    p0 = p0 + a;
    f0s = f0s + f0;
    p1[0] = p1[0] + f0 * d0; p1[1] = p1[1] + f0 * d1; p1[2] = p1[2] + f0 * d2;
    s1[0] = s1[0] + g1 * d0; s1[1] = s1[1] + g1 * d1; s1[2] = s1[2] + g1 * d2;
    const T d00 = d0 * d0, d01 = d0 * d1, d02 = d0 * d2, d11 = d1 * d1, d12 = d1 * d2, d22 = d2 * d2;
    m2[0] = m2[0] + g1 * d00; m2[1] = m2[1] + g1 * d01; m2[2] = m2[2] + g1 * d02;
    m2[3] = m2[3] + g1 * d11; m2[4] = m2[4] + g1 * d12; m2[5] = m2[5] + g1 * d22;
    const T t00 = t * d00, t01 = t * d01, t02 = t * d02, t11 = t * d11, t12 = t * d12, t22 = t * d22;
    m3[0] = m3[0] + t00 * d0; m3[1] = m3[1] + t00 * d1; m3[2] = m3[2] + t00 * d2;
    m3[3] = m3[3] + t11 * d0; m3[4] = m3[4] + t01 * d2; m3[5] = m3[5] + t22 * d0;
    m3[6] = m3[6] + t11 * d1; m3[7] = m3[7] + t11 * d2; m3[8] = m3[8] + t22 * d1;
    m3[9] = m3[9] + t22 * d2;
    (void)t02; (void)t12;
  }
  // the unique LRF components of these moments
  __device__ LrfTerms<T> terms() const {
    LrfTerms<T> o;
    o.p0 = p0;
    for (int k = 0; k < 3; ++k) o.p1[k] = -p1[k];
    o.p2[0] = (T)3 * m2[0] - f0s; o.p2[1] = (T)3 * m2[1]; o.p2[2] = (T)3 * m2[2];
    o.p2[3] = (T)3 * m2[3] - f0s; o.p2[4] = (T)3 * m2[4]; o.p2[5] = (T)3 * m2[5] - f0s;
    o.p3[0] = (T)-15 * m3[0] + (T)9 * s1[0];   // 000
    o.p3[1] = (T)-15 * m3[1] + (T)3 * s1[1];   // 001
    o.p3[2] = (T)-15 * m3[2] + (T)3 * s1[2];   // 002
    o.p3[3] = (T)-15 * m3[3] + (T)3 * s1[0];   // 011
    o.p3[4] = (T)-15 * m3[4];                  // 012
    o.p3[5] = (T)-15 * m3[5] + (T)3 * s1[0];   // 022
    o.p3[6] = (T)-15 * m3[6] + (T)9 * s1[1];   // 111
    o.p3[7] = (T)-15 * m3[7] + (T)3 * s1[2];   // 112
    o.p3[8] = (T)-15 * m3[8] + (T)3 * s1[1];   // 122
    o.p3[9] = (T)-15 * m3[9] + (T)9 * s1[2];   // 222
    return o;
  }
};

// Periodic box: the shift of the partner group whose switching atom is sw
// against centre c (lrf_update, src/md.f90:5904-5917), zero for a sphere.
// BOX (here and in the gathers below): the system is a periodic box, a
// compile-time flag so that the sphere's list builds carry no box arithmetic.
template <bool BOX>
__device__ __forceinline__ void lrf_box_shift(const DevSys& S, int sw, double c0, double c1, double c2,
                                              double* sh) {
  if constexpr (BOX) {
    sh[0] = pbc_shift(S, S.x[3 * sw] - c0, 0);
    sh[1] = pbc_shift(S, S.x[3 * sw + 1] - c1, 1);
    sh[2] = pbc_shift(S, S.x[3 * sw + 2] - c2, 2);
  } else {
    sh[0] = sh[1] = sh[2] = 0.0;
  }
}

// The FP32 raw moments of the non-Q atoms of charge group g (as seen from
// centre c), added to part.
template <bool BOX>
__device__ __forceinline__ void lrf_raw_group(const DevSys& S, int g, double c0, double c1, double c2,
                                              LrfRaw<float>& part) {
  double sh[3];
  lrf_box_shift<BOX>(S, S.cgp_sw[g], c0, c1, c2, sh);
  for (int a = S.cgp_first[g]; a <= S.cgp_last[g]; ++a) {
    const int i = S.cgpatom[a];
    if (S.flags[i] & F_Q) continue;
    part.add((float)S.crg[i], (float)(S.x[3 * i] - c0 - sh[0]), (float)(S.x[3 * i + 1] - c1 - sh[1]),
             (float)(S.x[3 * i + 2] - c2 - sh[2]));
  }
}

// Add the non-Q atoms of charge group g (as seen from centre c) to acc.
template <typename P, bool BOX>
__device__ __forceinline__ void lrf_gather_group(const DevSys& S, int g, double c0, double c1,
                                                 double c2, LrfAcc& acc) {
  if constexpr (sizeof(P) == 8) {
    double sh[3];
    lrf_box_shift<BOX>(S, S.cgp_sw[g], c0, c1, c2, sh);
    for (int a = S.cgp_first[g]; a <= S.cgp_last[g]; ++a) {
      const int i = S.cgpatom[a];
      if (S.flags[i] & F_Q) continue;
      if constexpr (BOX)
        acc.add(S.crg[i], S.x[3 * i] - c0 - sh[0], S.x[3 * i + 1] - c1 - sh[1], S.x[3 * i + 2] - c2 - sh[2]);
      else
        acc.add(S.crg[i], S.x[3 * i] - c0, S.x[3 * i + 1] - c1, S.x[3 * i + 2] - c2);
    }
  } else {
    // mixed precision: the group's moments in FP32 (raw form, then the
    // components), accumulated in FP64
    LrfRaw<float> part;
    part.zero();
    lrf_raw_group<BOX>(S, g, c0, c1, c2, part);
    acc.add(part.terms());
  }
}

// lrf_gather_group for a standard water (its charge group is its own three
// atoms jo, jo+1, jo+2 with the solvent charges, none of them a Q atom):
// the same terms in the same order, without the group/atom indirection.
// In mixed precision the partner's offsets from the centre come from the
// FP32 coordinate copy and the FP32 centre (cf).
template <typename P, bool BOX>
__device__ __forceinline__ void lrf_gather_water(const DevSys& S, int jo, int jw, double c0, double c1,
                                                 double c2, const float* cf, LrfAcc& acc) {
  double sh[3];
  lrf_box_shift<BOX>(S, jo, c0, c1, c2, sh);
  if constexpr (sizeof(P) == 8) {
    (void)jw;
    const gconst<double> xj = gptr(S.x) + 3 * jo;
    for (int k = 0; k < 3; ++k) {
      if constexpr (BOX)
        acc.add(S.chg_solv[k], xj[3 * k] - c0 - sh[0], xj[3 * k + 1] - c1 - sh[1], xj[3 * k + 2] - c2 - sh[2]);
      else
        acc.add(S.chg_solv[k], xj[3 * k] - c0, xj[3 * k + 1] - c1, xj[3 * k + 2] - c2);
    }
  } else {
    float xj[9];
    water_xf(S, jw, xj);
    const float sf[3] = {(float)sh[0], (float)sh[1], (float)sh[2]};
    LrfRaw<float> part;
    part.zero();
    for (int k = 0; k < 3; ++k)
      part.add((float)S.chg_solv[k], xj[3 * k] - cf[0] - sf[0], xj[3 * k + 1] - cf[1] - sf[1],
               xj[3 * k + 2] - cf[2] - sf[2]);
    acc.add(part.terms());
  }
}

// The raw moments of standard water jw's three atoms about the FP32 centre
// cf, added to part (FP32); in a periodic box moved by the shift of the
// water's O against the FP64 centre c.
constexpr int kLrfPend = 8;   // partner tiles whose far waters' FP32 raw moments are summed before conversion
template <bool BOX>
__device__ __forceinline__ void lrf_raw_water(const DevSys& S, int jw, const double* c, const float* cf,
                                              LrfRaw<float>& part) {
  float xj[9];
  water_xf(S, jw, xj);
  if constexpr (BOX) {
    double sh[3];
    lrf_box_shift<BOX>(S, S.nat_solute + 3 * jw, c[0], c[1], c[2], sh);
    const float sf[3] = {(float)sh[0], (float)sh[1], (float)sh[2]};
    for (int k = 0; k < 3; ++k)
      part.add((float)S.chg_solv[k], xj[3 * k] - cf[0] - sf[0], xj[3 * k + 1] - cf[1] - sf[1],
               xj[3 * k + 2] - cf[2] - sf[2]);
    return;
  }
  for (int k = 0; k < 3; ++k)
    part.add((float)S.chg_solv[k], xj[3 * k] - cf[0], xj[3 * k + 1] - cf[1], xj[3 * k + 2] - cf[2]);
}

// lrf_taylor for one atom at xa with its group's multipoles L: returns the
// potential Vij and the field df (src/md.f90:5843-5888).
__device__ __forceinline__ double lrf_taylor_atom(const double* L, const double* xa, double* df) {
  const double dr0 = L[0] - xa[0], dr1 = L[1] - xa[1], dr2 = L[2] - xa[2];
  const double* p1 = L + 4;
  const double* p2 = L + 7;
  const double* p3 = L + 16;
  double V = L[3] + p1[0] * dr0 + p1[1] * dr1 + p1[2] * dr2
           + 0.5 * (p2[0] * dr0 + p2[1] * dr1 + p2[2] * dr2) * dr0
           + 0.5 * (p2[3] * dr0 + p2[4] * dr1 + p2[5] * dr2) * dr1
           + 0.5 * (p2[6] * dr0 + p2[7] * dr1 + p2[8] * dr2) * dr2;
  for (int a = 0; a < 3; ++a) {
    const double* q = p3 + 9 * a;
    df[a] = p1[a] + p2[3 * a] * dr0 + p2[3 * a + 1] * dr1 + p2[3 * a + 2] * dr2
          + 0.5 * (q[0] * dr0 + q[1] * dr1 + q[2] * dr2) * dr0
          + 0.5 * (q[3] * dr0 + q[4] * dr1 + q[5] * dr2) * dr1
          + 0.5 * (q[6] * dr0 + q[7] * dr1 + q[8] * dr2) * dr2;
  }
  return V;
}

// Water dipole unit vector, radial unit vector and their cosine for the
// water whose O is atom i (src/md.f90:18392-18412).
__device__ __forceinline__ void watpol_geometry(const DevSys& S, int i, double* rmu, double* rcu,
                                                double& rm, double& rc, double& scp) {
  const double* x = S.x + 3 * i;
  for (int k = 0; k < 3; ++k) rmu[k] = x[3 + k] + x[6 + k] - 2.0 * x[k];
  rm = sqrt(rmu[0] * rmu[0] + rmu[1] * rmu[1] + rmu[2] * rmu[2]);
  for (int k = 0; k < 3; ++k) rmu[k] = rmu[k] / rm;
  for (int k = 0; k < 3; ++k) rcu[k] = x[k] - S.xwcent[k];
  rc = sqrt(rcu[0] * rcu[0] + rcu[1] * rcu[1] + rcu[2] * rcu[2]);
  for (int k = 0; k < 3; ++k) rcu[k] = rcu[k] / rc;
  scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
  if (scp > 1.0) scp = 1.0;
  if (scp < -1.0) scp = -1.0;
}

}  // namespace qgpu
