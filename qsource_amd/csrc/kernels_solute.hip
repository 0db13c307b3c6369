// CDNA4 kernels for the solute and Q-atom side of qdyn's MD step: solute
// charge-group lists with the LRF gather, solute nonbonded forces (solute-
// solute, solute-water, LRF Taylor, shell restraints and the solute side of
// the Q-solute terms), the Q-atom kernel (Q-solute and Q-water terms seen
// from the Q atom, per-state energies), and the term kernel for bonded terms
// and Q-Q / Q-neighbour pairs with its fixed-order gather.  Every atom's
// gradient is written by one lane or summed in a fixed order, so results
// are independent of scheduling.
#include "kernels_common.hpp"

namespace qgpu {

// Exclusion (bit 0) and 1-4 (bit 1) status of a solute atom pair: the bit
// rows for pairs at most kMaxNbrRange apart, else the lower atom's long-pair
// row (listex/list14 and listexlong/list14long, src/md.f90:7745-7760).
__device__ __forceinline__ int pair_status(const DevSys& S, int i, int j) {
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  if (hi - lo <= qsys::kMaxNbrRange)
    return (int)((S.listex[lo] >> (hi - lo - 1)) & 1u) | (int)(((S.list14[lo] >> (hi - lo - 1)) & 1u) << 1);
  for (int n = S.exl_first[lo]; n < S.exl_first[lo + 1]; ++n) {
    const int c = S.exl_code[n];
    if ((c >> 2) == hi) return c & 3;
  }
  return 0;
}
__device__ __forceinline__ bool zero_interaction(const DevSys& S, int i, int j, int lj) {
  const int ti = S.iac[i], tj = S.iac[j];
  return S.crg[i] * S.crg[j] == 0.0 && S.avdw[ti * 3 + lj] * S.avdw[tj * 3 + lj] == 0.0 &&
         S.bvdw[ti * 3 + lj] * S.bvdw[tj * 3 + lj] == 0.0;
}

// Skip and 1-4 masks of the atom pairs of solute groups (g, jg), with the
// skip rules of nbpplist (src/md.f90:8080-8120): the same atom, a Q atom on
// either side, a zero interaction, an excluded pair.
__device__ ulonglong2 pair_masks(const DevSys& S, int g, int jg) {
  unsigned long long skip = 0, m14 = 0;
  const int fi = S.cgp_first[g], ni = S.cgp_last[g] - fi + 1;
  const int fj = S.cgp_first[jg], nj = S.cgp_last[jg] - fj + 1;
  // slots past the end of either group are never evaluated
  const unsigned long long absent_cols = 0xffull & ~((1ull << nj) - 1ull);
  for (int a = 0; a < ni; ++a) {
    const int i = S.cgpatom[fi + a];
    const bool qi = (S.flags[i] & F_Q) != 0;
    skip |= absent_cols << (a * kMaxGroupAtoms);
    for (int b = 0; b < nj; ++b) {
      const int j = S.cgpatom[fj + b];
      const unsigned long long bit = 1ull << (a * kMaxGroupAtoms + b);
      if (qi || j == i || (S.flags[j] & F_Q)) { skip |= bit; continue; }
      const int lj = S.ljcod[S.iac[i] * S.natyps + S.iac[j]];
      if (zero_interaction(S, i, j, lj - 1)) { skip |= bit; continue; }
      const int st = pair_status(S, i, j);
      if (st & 1) { skip |= bit; continue; }
      if (st & 2) m14 |= bit;
    }
  }
  for (int a = ni; a < kMaxGroupAtoms; ++a) skip |= 0xffull << (a * kMaxGroupAtoms);
  return make_ulonglong2(skip, m14);
}

// ---------------------------------------------------------------------------
// Solute lists (nbpplist_lrf src/md.f90:7998-8188 and the solute side of
// nbpwlist_lrf :9300-9429): one wavefront per solute charge group; P as in
// k_water_lists (precision of the LRF multipole terms); BOX: periodic box
// (nbpplist_box_lrf and nbpwlist_box_lrf: minimum-image distances, shifted
// LRF partners).
template <typename P, bool BOX>
__device__ __forceinline__ void solute_lists_body(const DevSys& S, const int base) {
  const int lane = threadIdx.x & 63;
  const int g = base + (threadIdx.x >> 6);
  if (g >= S.ncgp_solute) return;
  const int sw = S.cgp_sw[g];
  if (S.flags[sw] & F_EXCL) {
    if (lane == 0) { S.pp_cnt[g] = 0; S.pw_cnt[g] = 0; }
    return;
  }
  const double xs0 = S.x[3 * sw], xs1 = S.x[3 * sw + 1], xs2 = S.x[3 * sw + 2];
  double* Lg = S.lrf + (size_t)g * kLrfStride;
  const double c0 = Lg[0], c1 = Lg[1], c2 = Lg[2];
  const float cf[3] = {(float)c0, (float)c1, (float)c2};
  LrfAcc acc;
  acc.zero();
  int n_pp = 0;
  int* prow = S.pp_nb + (size_t)g * S.pp_cap;
  for (int base = 0; base < S.ncgp_solute; base += 64) {
    int jg = base + lane;
    bool inside = false, far = false;
    if (jg < S.ncgp_solute) {
      int js = S.cgp_sw[jg];
      if (!(S.flags[js] & F_EXCL)) {
        double a = xs0 - S.x[3 * js], b = xs1 - S.x[3 * js + 1], c = xs2 - S.x[3 * js + 2];
        if constexpr (BOX) { a = pbc_image(S, a, 0); b = pbc_image(S, b, 1); c = pbc_image(S, c, 2); }
        double r2 = a * a + b * b + c * c;
        inside = r2 <= S.rcpp2;
        far = !inside && r2 <= S.rclrf2;
      }
    }
    unsigned long long m = __ballot(inside);
    if (inside) {
      int pos = n_pp + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < S.pp_cap) {
        prow[pos] = jg;
        S.pp_mask[(size_t)g * S.pp_cap + pos] = pair_masks(S, g, jg);
      }
    }
    n_pp += __popcll(m);
    if (far && S.use_lrf) lrf_gather_group<P, BOX>(S, jg, c0, c1, c2, acc);
  }
  int n_pw = 0;
  int* wrow = S.pw_nb + (size_t)g * S.pw_cap;
  for (int base = 0; base < S.nwat; base += 64) {
    int jw = base + lane;
    bool inside = false, far = false;
    int jo = S.nat_solute + 3 * jw;
    if (jw < S.nwat && !(S.flags[jo] & F_EXCL)) {
      double a = xs0 - S.x[3 * jo], b = xs1 - S.x[3 * jo + 1], c = xs2 - S.x[3 * jo + 2];
      if constexpr (BOX) { a = pbc_image(S, a, 0); b = pbc_image(S, b, 1); c = pbc_image(S, c, 2); }
      double r2 = a * a + b * b + c * c;
      inside = r2 <= S.rcpw2;
      far = !inside && r2 <= S.rclrf2;
    }
    unsigned long long m = __ballot(inside);
    if (inside) {
      int pos = n_pw + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < S.pw_cap) wrow[pos] = jw;
    }
    n_pw += __popcll(m);
    if (far && S.use_lrf) {
      if (S.std_water) lrf_gather_water<P, BOX>(S, jo, jw, c0, c1, c2, cf, acc);
      else lrf_gather_group<P, BOX>(S, S.iwhich_cgp[jo], c0, c1, c2, acc);
    }
  }
  acc.wave_reduce();
  if (lane == 0) {
    if (S.use_lrf) acc.store(Lg);
    S.pp_cnt[g] = n_pp < S.pp_cap ? n_pp : S.pp_cap;
    S.pw_cnt[g] = n_pw < S.pw_cap ? n_pw : S.pw_cap;
    if (n_pp > S.pp_cap || n_pw > S.pw_cap) {
      atomicMax(S.status + 4, n_pw);
      record_failure(S, ST_LIST_OVERFLOW, PH_LISTS);
    }
  }
}
template <typename P, int MAPPED>
__global__ __launch_bounds__(256) void k_solute_lists(const DevSys* __restrict__ sys, const int2* __restrict__ map) {
  const DevSys* Sp;
  int base;
  if (!block_work<MAPPED>(sys, map, blockDim.x >> 6, Sp, base)) return;
  if (Sp->pbc) solute_lists_body<P, true>(*Sp, base);
  else solute_lists_body<P, false>(*Sp, base);
}
template __global__ void k_solute_lists<double, 0>(const DevSys*, const int2*);
template __global__ void k_solute_lists<float, 0>(const DevSys*, const int2*);
template __global__ void k_solute_lists<double, 1>(const DevSys*, const int2*);
template __global__ void k_solute_lists<float, 1>(const DevSys*, const int2*);

// ---------------------------------------------------------------------------
// Solute forces (nonbond_pp :13069-13155, nonbond_pw :13282-13365,
// lrf_taylor, fix_shell :2816-2862, and the solute-atom side of nonbond_qp
// :13788-13921).  kSFLanes lanes per solute atom (16: the short rows of a
// solute atom leave latency to hide, so more, shorter lanes): the lanes of
// an atom take the neighbour groups of its charge group (pair status from
// the list-time masks), the listed waters and the Q atoms in turn; lane 1
// adds the LRF Taylor term and lane 2 the shell/fix restraint; the atom's
// gradient is summed over its lanes and written once.  A pair's energy is
// counted by the atom that is i of the reference pair: the owner group's
// atom, or the lower index within one group.
// FP32 pair arithmetic fits 6 wavefronts per SIMD without spilling
// en as in k_water_forces (energy partials e_solute only when set)
template <typename P, int MAPPED, int PF>
__global__ __launch_bounds__(256, sizeof(P) == 4 ? 6 : 1) void k_solute_forces(const DevSys* __restrict__ sys,
                                                                               const int2* __restrict__ map,
                                                                               int en) {
  const DevSys* Sp;
  int base;
  if (!block_work<MAPPED>(sys, map, blockDim.x / kSFLanes, Sp, base)) return;
  const DevSys& S = *Sp;
  const int t = threadIdx.x;
  if (base + (t >> 6) * (64 / kSFLanes) >= S.nat_solute) return;   // whole wavefronts past the last atom
  const int sub = t & (kSFLanes - 1);
  const int i = base + t / kSFLanes;
  const bool active = i < S.nat_solute;
  const int ia = active ? i : 0;
  const int g = S.iwhich_cgp[ia];
  const int a = S.cgp_slot[ia];
  const unsigned char fl = S.flags[ia];
  const gconst<P> X = gptr(pair_x<P>(S));
  const P xi0 = X[3 * ia], xi1 = X[3 * ia + 1], xi2 = X[3 * ia + 2];
  const int ti = S.iac[ia];
  const P qi = (P)S.crg[ia];
  double g0 = 0, g1 = 0, g2 = 0;
  double e_ppel = 0, e_ppvdw = 0, e_pwel = 0, e_pwvdw = 0, e_lrf = 0, e_fix = 0, e_shell = 0;
  if (active && !(fl & F_Q)) {
    // solute-solute
    const int npp = (S.term_mask & TM_PP) ? S.pp_cnt[g] : 0;
    const gconst<int> prow = gptr(S.pp_nb) + (size_t)g * S.pp_cap;
    // (skip, 1-4) mask pairs of the entries, read as two 64-bit words
    const gconst<unsigned long long> mrow =
        gptr(reinterpret_cast<const unsigned long long*>(S.pp_mask + (size_t)g * S.pp_cap));
    const gconst<int> gatom = gptr(S.cgpatom);
    const gconst<int> giac = gptr(S.iac);
    const gconst<double> gcrg = gptr(S.crg);
    for (int k = sub; k < npp; k += kSFLanes) {
      const int jg = prow[k];
      const unsigned long long mk_x = mrow[2 * k], mk_y = mrow[2 * k + 1];
      unsigned todo = (unsigned)(~mk_x >> (a * kMaxGroupAtoms)) & 0xffu;
      const unsigned is14 = (unsigned)(mk_y >> (a * kMaxGroupAtoms)) & 0xffu;
      const bool same = jg == g;
      const bool own = same || owns_pair(g, jg);
      const int fj = S.cgp_first[jg];
      // periodic box (nonbond_pp_box :13159-13280): the group pair's shift
      // from the two switching atoms, as seen from this atom's group
      P ps0 = 0, ps1 = 0, ps2 = 0;
      if (S.pbc) {
        const int swi = S.cgp_sw[g], swj = S.cgp_sw[jg];
        ps0 = pbc_shift(S, X[3 * swj] - X[3 * swi], 0);
        ps1 = pbc_shift(S, X[3 * swj + 1] - X[3 * swi + 1], 1);
        ps2 = pbc_shift(S, X[3 * swj + 2] - X[3 * swi + 2], 2);
      }
      while (todo) {
        const int b = __builtin_ctz(todo);
        todo &= todo - 1;
        const int j = gatom[fj + b];
        const int tj = giac[j];
        const int lj = ((is14 >> b) & 1) ? 3 : S.ljcod[ti * S.natyps + tj];
        const int l = lj - 1;
        const bool me_first = same ? (i < j) : own;
        // dx = x_j - x_i; the reference pair (ia, ja) has dx of the same sign
        // when i is first, and the opposite sign otherwise
        P dx1 = X[3 * j] - xi0, dx2 = X[3 * j + 1] - xi1, dx3 = X[3 * j + 2] - xi2;
        if (S.pbc) { dx1 = dx1 - ps0; dx2 = dx2 - ps1; dx3 = dx3 - ps2; }
        if (!me_first) { dx1 = -dx1; dx2 = -dx2; dx3 = -dx3; }
        P r;
        const P r2 = inv_r2(pnorm2(dx1, dx2, dx3), r);
        const P r6 = r2 * r2 * r2;
        P cj, ai, aj, bi, bj;
        if constexpr (sizeof(P) == 4) {
          // FP32 copies of the same values (the FP64 ones rounded once)
          cj = S.crg_f[j];
          ai = S.avdw_f[ti * 3 + l]; aj = S.avdw_f[tj * 3 + l];
          bi = S.bvdw_f[ti * 3 + l]; bj = S.bvdw_f[tj * 3 + l];
        } else {
          cj = gcrg[j];
          ai = S.avdw[ti * 3 + l]; aj = S.avdw[tj * 3 + l];
          bi = S.bvdw[ti * 3 + l]; bj = S.bvdw[tj * 3 + l];
        }
        P Vel = qi * cj * r;
        if (lj == 3) Vel = Vel * (P)S.el14_scale;
        const P Va = r6 * r6 * ai * aj;
        const P Vb = r6 * bi * bj;
        const P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb);
        if (me_first) {
          g0 = g0 - (double)(dv * dx1);
          g1 = g1 - (double)(dv * dx2);
          g2 = g2 - (double)(dv * dx3);
          if (en) {
            e_ppel = e_ppel + (double)Vel;
            e_ppvdw = e_ppvdw + (double)Va - (double)Vb;
          }
        } else {
          g0 = g0 + (double)(dv * dx1);
          g1 = g1 + (double)(dv * dx2);
          g2 = g2 + (double)(dv * dx3);
        }
      }
    }
    // solute-water (the solute atom is i of the reference pair)
    const int npw = (S.term_mask & TM_PW) ? S.pw_cnt[g] : 0;
    const gconst<int> wrow = gptr(S.pw_nb) + (size_t)g * S.pw_cap;
    // standard waters: LJ products and zero-interaction mask of this atom
    // with O, H1, H2 from the per-atom table
    const P* wlj = S.std_water ? solute_water_lj<P>(S) + 6 * (size_t)i : nullptr;
    const unsigned wzero = S.std_water ? S.swp_zero[i] : 0u;
    // standard waters carry the solvent charges (checked at upload)
    const P qw3[3] = {(P)S.chg_solv[0], (P)S.chg_solv[1], (P)S.chg_solv[2]};
    auto pw_water = [&](const int jo, const P* xw) {
      // periodic box (nonbond_pw_box): the shift of the water O against the
      // group's switching atom
      P ws0 = 0, ws1 = 0, ws2 = 0;
      if (S.pbc) {
        const int sw = S.cgp_sw[g];
        ws0 = pbc_shift(S, xw[0] - X[3 * sw], 0);
        ws1 = pbc_shift(S, xw[1] - X[3 * sw + 1], 1);
        ws2 = pbc_shift(S, xw[2] - X[3 * sw + 2], 2);
      }
      for (int ip = 0; ip < 3; ++ip) {
        const int j = jo + ip;
        P aLJ, bLJ;
        if (wlj) {
          if ((wzero >> ip) & 1u) continue;
          aLJ = wlj[ip];
          bLJ = wlj[3 + ip];
        } else {
          const int tj = S.iac[j];
          const int lj = S.ljcod[ti * S.natyps + tj] - 1;
          if (zero_interaction(S, ia, j, lj)) continue;
          aLJ = (P)(S.avdw[ti * 3 + lj] * S.avdw[tj * 3 + lj]);
          bLJ = (P)(S.bvdw[ti * 3 + lj] * S.bvdw[tj * 3 + lj]);
        }
        P dx1 = xw[3 * ip] - xi0, dx2 = xw[3 * ip + 1] - xi1, dx3 = xw[3 * ip + 2] - xi2;
        if (S.pbc) { dx1 = dx1 - ws0; dx2 = dx2 - ws1; dx3 = dx3 - ws2; }
        P r;
        const P r2 = inv_r2(pnorm2(dx1, dx2, dx3), r);
        const P r6 = r2 * r2 * r2;
        const P r12 = r6 * r6;
        const P Vel = qi * (wlj ? qw3[ip] : (P)S.crg[j]) * r;
        const P Va = aLJ * r12;
        const P Vb = bLJ * r6;
        const P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb);
        g0 = g0 - (double)(dv * dx1);
        g1 = g1 - (double)(dv * dx2);
        g2 = g2 - (double)(dv * dx3);
        if (en) {
          e_pwel = e_pwel + (double)Vel;
          e_pwvdw = e_pwvdw + (double)Va - (double)Vb;
        }
      }
    };
    // coordinates of water jw: FP64, or the FP32 water record
    auto load_water = [&](const int jw, P* o) {
      if constexpr (sizeof(P) == 4) {
        water_xf(S, jw, o);
      } else {
        const int jo = S.nat_solute + 3 * jw;
        for (int u = 0; u < 9; ++u) o[u] = X[3 * jo + u];
      }
    };
    if constexpr (PF != 0) {
      // software pipeline: the next water's coordinates and the index of
      // the one after are in flight during this water's pairs
      int k = sub;
      int jwn = 0, jwn2 = 0;
      P xn[9];
      if (k < npw) {
        jwn = wrow[k];
        load_water(jwn, xn);
      }
      if (k + kSFLanes < npw) jwn2 = wrow[k + kSFLanes];
      while (k < npw) {
        P xw[9];
        for (int u = 0; u < 9; ++u) xw[u] = xn[u];
        const int jo = S.nat_solute + 3 * jwn;
        k += kSFLanes;
        if (k < npw) {
          jwn = jwn2;
          load_water(jwn, xn);
          if (k + kSFLanes < npw) jwn2 = wrow[k + kSFLanes];
        }
        pw_water(jo, xw);
      }
    } else {
      for (int k = sub; k < npw; k += kSFLanes) {
        const int jw = wrow[k];
        P xw[9];
        load_water(jw, xw);
        pw_water(S.nat_solute + 3 * jw, xw);
      }
    }
    // Q-solute pairs seen from the solute atom (the Q atom is i of the pair)
    if ((fl & F_QP) && (S.term_mask & TM_QX)) {
      // periodic box (nonbond_qp_box :13925-14096): the group's shift atom
      // against the Q switching atom
      P qs0 = 0, qs1 = 0, qs2 = 0;
      if (S.pbc) {
        const int sa = S.qp_shatom[g], qw = S.qswitch;
        qs0 = pbc_shift(S, X[3 * sa] - X[3 * qw], 0);
        qs1 = pbc_shift(S, X[3 * sa + 1] - X[3 * qw + 1], 1);
        qs2 = pbc_shift(S, X[3 * sa + 2] - X[3 * qw + 2], 2);
      }
      for (int iq = sub; iq < S.nqat; iq += kSFLanes) {
        const int qa = S.iqseq[iq];
        const int lj = S.qp_lj[(size_t)i * S.nqat + iq];
        const int l = lj - 1;
        const int tq = S.iac[qa];
        P dx1 = xi0 - X[3 * qa], dx2 = xi1 - X[3 * qa + 1], dx3 = xi2 - X[3 * qa + 2];
        if (S.pbc) { dx1 = dx1 - qs0; dx2 = dx2 - qs1; dx3 = dx3 - qs2; }
        const P r2sq = pnorm2(dx1, dx2, dx3);
        P r;
        const P r2 = inv_r2(r2sq, r);
        const P r6 = r2 * r2 * r2;
        const P aLJ = (P)(S.avdw[tq * 3 + l] * S.avdw[ti * 3 + l]);
        const P bLJ = (P)(S.bvdw[tq * 3 + l] * S.bvdw[ti * 3 + l]);
        if (S.qvdw) {
          const P r6h = r2sq * r2sq * r2sq;
          for (int st = 0; st < S.nstates; ++st) {
            P Vel, Va, Vb;
            const P dv = qp_qvdw_state<P>(S, iq, st, ti, lj, qi, r2, r, r6h, Vel, Va, Vb);
            g0 = g0 + (double)(dv * dx1);
            g1 = g1 + (double)(dv * dx2);
            g2 = g2 + (double)(dv * dx3);
          }
          continue;
        }
        for (int st = 0; st < S.nstates; ++st) {
          P Vel = (P)S.qcrg[iq * S.nstates + st] * qi * r;
          if (lj == 3) Vel = Vel * (P)S.el14_scale;
          const P Va = aLJ * r6 * r6;
          const P Vb = bLJ * r6;
          const P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb) * (P)S.lambda[st];
          g0 = g0 + (double)(dv * dx1);
          g1 = g1 + (double)(dv * dx2);
          g2 = g2 + (double)(dv * dx3);
        }
      }
    }
    // LRF Taylor expansion about the group centre (FP64)
    if (sub == 1 && S.use_lrf && !(fl & F_EXCL) && (S.term_mask & TM_LRF)) {
      const double* L = S.lrf + (size_t)g * kLrfStride;
      const double xa[3] = {S.x[3 * i], S.x[3 * i + 1], S.x[3 * i + 2]};
      double df[3];
      const double V = lrf_taylor_atom(L, xa, df);
      e_lrf = 0.5 * S.crg[i] * V;
      g0 = g0 - S.crg[i] * df[0];
      g1 = g1 - S.crg[i] * df[1];
      g2 = g2 - S.crg[i] * df[2];
    }
  }
  // harmonic restraint of excluded and shell atoms to the topology position
  // (excluded_freeze: the atom is put back there instead, no restraint)
  if (active && sub == 2 && (fl & (F_EXCL | F_SHELL)) && (S.term_mask & TM_OTHER)) {
    const double fk = (fl & F_EXCL) ? (S.freeze ? 0.0 : S.fk_fix) : S.fk_pshell;
    const double dr0 = S.x[3 * i] - S.xtop[3 * i], dr1 = S.x[3 * i + 1] - S.xtop[3 * i + 1],
                 dr2 = S.x[3 * i + 2] - S.xtop[3 * i + 2];
    const double erst = 0.5 * fk * (dr0 * dr0 + dr1 * dr1 + dr2 * dr2);
    if (fl & F_EXCL) e_fix = erst;
    if (fl & F_SHELL) e_shell = erst;
    g0 = g0 + fk * dr0;
    g1 = g1 + fk * dr1;
    g2 = g2 + fk * dr2;
  }
  g0 = segn_sum<kSFLanes>(g0); g1 = segn_sum<kSFLanes>(g1); g2 = segn_sum<kSFLanes>(g2);
  if (en) {
    e_ppel = segn_sum<kSFLanes>(e_ppel); e_ppvdw = segn_sum<kSFLanes>(e_ppvdw);
    e_pwel = segn_sum<kSFLanes>(e_pwel); e_pwvdw = segn_sum<kSFLanes>(e_pwvdw);
    e_lrf = segn_sum<kSFLanes>(e_lrf); e_fix = segn_sum<kSFLanes>(e_fix); e_shell = segn_sum<kSFLanes>(e_shell);
  }
  if (active && sub == 0) {
    S.d[3 * i] = g0; S.d[3 * i + 1] = g1; S.d[3 * i + 2] = g2;
  }
  if (active && sub == 0 && en) {
    double* es = S.e_solute + (size_t)i * 8;
    es[0] = e_ppel; es[1] = e_ppvdw; es[2] = e_pwel; es[3] = e_pwvdw;
    es[4] = e_lrf; es[5] = e_fix; es[6] = e_shell; es[7] = 0.0;
  }
}
template __global__ void k_solute_forces<double, 0, 0>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<float, 0, 0>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<double, 1, 0>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<float, 1, 0>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<double, 0, 1>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<float, 0, 1>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<double, 1, 1>(const DevSys*, const int2*, int);
template __global__ void k_solute_forces<float, 1, 1>(const DevSys*, const int2*, int);

// ---------------------------------------------------------------------------
// Q atoms (nonbond_qp :13788-13921, nonbond_qw_spc :14411-14570 seen from
// the Q atom): kQChunks wavefronts per Q atom, each taking an interleaved
// share of the Q-solute partners and the Q-water waters (the Q-water list
// spans the whole sphere); each wavefront writes its gradient and per-state
// energies to q_part, and k_q_reduce sums the chunks in a fixed order.
// NS bounds the number of FEP states at compile time so that the per-state
// accumulators stay in registers.
template <typename P, int NS, bool BOX>
__device__ __forceinline__ void q_forces_body(const DevSys& S, const int shadow) {
  const int lane = threadIdx.x & 63;
  const int wv = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int iq = wv / kQChunks, chunk = wv % kQChunks;
  if (iq >= S.nqat) return;
  const int ns = S.nstates;
  const int qi = S.iqseq[iq];
  const double xq0 = S.x[3 * qi], xq1 = S.x[3 * qi + 1], xq2 = S.x[3 * qi + 2];
  const int ti = S.iac[qi];
  double g0 = 0, g1 = 0, g2 = 0;
  double e[NS][4];
  P qs[NS], lam[NS];
#pragma unroll
  for (int st = 0; st < NS; ++st) {
    e[st][0] = e[st][1] = e[st][2] = e[st][3] = 0.0;
    qs[st] = st < ns ? (P)S.qcrg[iq * ns + st] : (P)0;
    lam[st] = st < ns ? (P)S.lambda[st] : (P)0;
  }
  const int stride = 64 * kQChunks;
  const int nqp = S.q_counts[0], nqw = S.q_counts[1];
  for (int k = chunk * 64 + lane; k < nqp; k += stride) {
    const int j = S.qp_atoms[k];
    const int lj = S.qp_lj[(size_t)j * S.nqat + iq];
    const int l = lj - 1;
    const int tj = S.iac[j];
    P dx1, dx2, dx3;
    if constexpr (BOX) {
      // periodic box (nonbond_qp_box): the shift of the partner's group
      const int sa = S.qp_shatom[S.iwhich_cgp[j]], qw = S.qswitch;
      const double s0 = pbc_shift(S, S.x[3 * sa] - S.x[3 * qw], 0);
      const double s1 = pbc_shift(S, S.x[3 * sa + 1] - S.x[3 * qw + 1], 1);
      const double s2 = pbc_shift(S, S.x[3 * sa + 2] - S.x[3 * qw + 2], 2);
      dx1 = (P)(S.x[3 * j] - xq0 - s0); dx2 = (P)(S.x[3 * j + 1] - xq1 - s1); dx3 = (P)(S.x[3 * j + 2] - xq2 - s2);
    } else {
      dx1 = (P)(S.x[3 * j] - xq0); dx2 = (P)(S.x[3 * j + 1] - xq1); dx3 = (P)(S.x[3 * j + 2] - xq2);
    }
    const P r2sq = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    P r;
    P r2 = inv_r2(r2sq, r);
    P r6 = r2 * r2 * r2;
    const P aLJ = (P)(S.avdw[ti * 3 + l] * S.avdw[tj * 3 + l]);
    const P bLJ = (P)(S.bvdw[ti * 3 + l] * S.bvdw[tj * 3 + l]);
    const P cj = (P)S.crg[j];
    if (S.qvdw) {
      const P r6h = r2sq * r2sq * r2sq;
#pragma unroll
      for (int st = 0; st < NS; ++st) {
        if (st >= ns) continue;
        P Vel, Va, Vb;
        const P dv = qp_qvdw_state<P>(S, iq, st, tj, lj, cj, r2, r, r6h, Vel, Va, Vb);
        g0 = g0 - (double)(dv * dx1);
        g1 = g1 - (double)(dv * dx2);
        g2 = g2 - (double)(dv * dx3);
        e[st][0] = e[st][0] + (double)Vel;
        e[st][1] = e[st][1] + (double)Va - (double)Vb;
      }
      continue;
    }
#pragma unroll
    for (int st = 0; st < NS; ++st) {
      if (st >= ns) continue;
      P Vel = qs[st] * cj * r;
      if (lj == 3) Vel = Vel * (P)S.el14_scale;
      P Va = aLJ * r6 * r6;
      P Vb = bLJ * r6;
      P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb) * lam[st];
      g0 = g0 - (double)(dv * dx1);
      g1 = g1 - (double)(dv * dx2);
      g2 = g2 - (double)(dv * dx3);
      e[st][0] = e[st][0] + (double)Vel;
      e[st][1] = e[st][1] + (double)Va - (double)Vb;
    }
  }
  const int lo = S.qw_lj[3 * iq];  // LJ code of the (water O, Q atom) pair
  const P aO = (P)S.avdw[ti * 3 + lo - 1], bO = (P)S.bvdw[ti * 3 + lo - 1];
  const P aS = (P)S.alj_w[0][lo - 1], bS = (P)S.blj_w[0][lo - 1];
  const bool allatom = S.solv_allatom != 0;
  const P cO = (P)S.chg_solv[0];
  for (int k = chunk * 64 + lane; k < nqw; k += stride) {
    const int jo = S.nat_solute + 3 * S.qw_waters[k];
    // periodic box (nonbond_qw_spc_box): the water O's shift against the Q
    // switching atom, for all three water atoms
    double ws0 = 0.0, ws1 = 0.0, ws2 = 0.0;
    if constexpr (BOX) {
      const int qw = S.qswitch;
      ws0 = pbc_shift(S, S.x[3 * jo] - S.x[3 * qw], 0);
      ws1 = pbc_shift(S, S.x[3 * jo + 1] - S.x[3 * qw + 1], 1);
      ws2 = pbc_shift(S, S.x[3 * jo + 2] - S.x[3 * qw + 2], 2);
    }
    P dx1, dx2, dx3;
    if constexpr (BOX) {
      dx1 = (P)(S.x[3 * jo] - xq0 - ws0); dx2 = (P)(S.x[3 * jo + 1] - xq1 - ws1); dx3 = (P)(S.x[3 * jo + 2] - xq2 - ws2);
    } else {
      dx1 = (P)(S.x[3 * jo] - xq0); dx2 = (P)(S.x[3 * jo + 1] - xq1); dx3 = (P)(S.x[3 * jo + 2] - xq2);
    }
    P r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    P r6_hc = r2 * r2 * r2;
    P r;
    r2 = inv_r2(r2, r);
    P r6 = (P)1 / r6_hc;
    P Va = aO * aS * r6 * r6;
    P Vb = bO * bS * r6;
    P dv = 0;
    if (S.qvdw) {
#pragma unroll
      for (int st = 0; st < NS; ++st) {
        if (st >= ns) continue;
        P Vel;
        dv = dv + qw_qvdw_state<P>(S, iq, st, 0, lo, r2, r, r6_hc, Vel, Va, Vb);
        e[st][2] = e[st][2] + (double)Vel;
        e[st][3] = e[st][3] + (double)Va - (double)Vb;
      }
    } else {
#pragma unroll
      for (int st = 0; st < NS; ++st) {
        if (st >= ns) continue;
        P Vel = cO * qs[st] * r;
        dv = dv + r2 * (-Vel - ((P)12 * Va - (P)6 * Vb) * (r6_hc / r6_hc)) * lam[st];
        e[st][2] = e[st][2] + (double)Vel;
        e[st][3] = e[st][3] + (double)Va - (double)Vb;
      }
    }
    g0 = g0 - (double)(dv * dx1);
    g1 = g1 - (double)(dv * dx2);
    g2 = g2 - (double)(dv * dx3);
    for (int ip = 1; ip < 3; ++ip) {
      const int ja = jo + ip;
      P ex1, ex2, ex3;
      if constexpr (BOX) {
        ex1 = (P)(S.x[3 * ja] - xq0 - ws0); ex2 = (P)(S.x[3 * ja + 1] - xq1 - ws1); ex3 = (P)(S.x[3 * ja + 2] - xq2 - ws2);
      } else {
        ex1 = (P)(S.x[3 * ja] - xq0); ex2 = (P)(S.x[3 * ja + 1] - xq1); ex3 = (P)(S.x[3 * ja + 2] - xq2);
      }
      P rr;
      const P h2 = ex1 * ex1 + ex2 * ex2 + ex3 * ex3;
      P q2 = inv_r2(h2, rr);
      const P ch = (P)S.chg_solv[ip];
      P dvh = 0;
      if (allatom) {
        // all-atom solvent: the H atoms carry LJ too, with their own LJ code
        // (nonbond_qw_3atom, src/md.f90:14801-14858)
        const int lh = S.qw_lj[3 * iq + ip];
        const P h6_hc = h2 * h2 * h2;
        if (S.qvdw) {
#pragma unroll
          for (int st = 0; st < NS; ++st) {
            if (st >= ns) continue;
            P Vel;
            dvh = dvh + qw_qvdw_state<P>(S, iq, st, ip, lh, q2, rr, h6_hc, Vel, Va, Vb);
            e[st][2] = e[st][2] + (double)Vel;
            e[st][3] = e[st][3] + (double)Va - (double)Vb;
          }
        } else {
          const P h6 = (P)1 / h6_hc;
          const P Vah = (P)S.avdw[ti * 3 + lh - 1] * (P)S.alj_w[ip][lh - 1] * h6 * h6;
          const P Vbh = (P)S.bvdw[ti * 3 + lh - 1] * (P)S.blj_w[ip][lh - 1] * h6;
#pragma unroll
          for (int st = 0; st < NS; ++st) {
            if (st >= ns) continue;
            P Vel = ch * qs[st] * rr;
            dvh = dvh + q2 * (-Vel - ((P)12 * Vah - (P)6 * Vbh) * (h6_hc / h6_hc)) * lam[st];
            e[st][2] = e[st][2] + (double)Vel;
            e[st][3] = e[st][3] + (double)Vah - (double)Vbh;
          }
        }
      } else {
#pragma unroll
        for (int st = 0; st < NS; ++st) {
          if (st >= ns) continue;
          P Vel = ch * qs[st] * rr;
          e[st][2] = e[st][2] + (double)Vel;
          dvh = dvh - q2 * Vel * lam[st];
        }
      }
      g0 = g0 - (double)(dvh * ex1);
      g1 = g1 - (double)(dvh * ex2);
      g2 = g2 - (double)(dvh * ex3);
    }
  }
  g0 = wave_sum(g0); g1 = wave_sum(g1); g2 = wave_sum(g2);
  double* out = (shadow ? S.q_part64 : S.q_part) + (size_t)wv * kQPartStride;
  if (lane == 0) { out[0] = g0; out[1] = g1; out[2] = g2; }
#pragma unroll
  for (int st = 0; st < NS; ++st) {
    if (st >= ns) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double v = wave_sum(e[st][c]);
      if (lane == 0) out[3 + st * 4 + c] = v;
    }
  }
}
template <typename P, int NS>
__global__ __launch_bounds__(256) void k_q_forces(const DevSys* __restrict__ sys, int shadow) {
  const DevSys& S = sys[blockIdx.y];
  if (shadow && !shadow_record_step(S)) return;
  if (S.pbc) q_forces_body<P, NS, true>(S, shadow);
  else q_forces_body<P, NS, false>(S, shadow);
}
template __global__ void k_q_forces<double, 2>(const DevSys*, int);
template __global__ void k_q_forces<float, 2>(const DevSys*, int);
template __global__ void k_q_forces<double, 4>(const DevSys*, int);
template __global__ void k_q_forces<float, 4>(const DevSys*, int);
template __global__ void k_q_forces<double, kMaxQStates>(const DevSys*, int);
template __global__ void k_q_forces<float, kMaxQStates>(const DevSys*, int);

// Q-solute and Q-water lists rebuilt from the current coordinates when
// Rcq <= rexcl_o (nbqplist src/md.f90:10215-10322, nbqwlist :10464-10549):
// the non-connected atoms of each solute charge group whose switch atom lies
// within Rcq of the solute centre, and the waters whose O does, neither
// excluded.  One workgroup per instance compacts them in ascending order
// (wavefront ballots, then the waves in order) and sets the F_QP / F_QW
// flags the force kernels read.
__device__ __forceinline__ int ordered_compact(bool take, int* wave_cnt, int& total) {
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const unsigned long long m = __ballot(take);
  if (ln == 0) wave_cnt[wv] = __popcll(m);
  __syncthreads();
  int off = total;
  for (int w = 0; w < wv; ++w) off += wave_cnt[w];
  int all = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) all += wave_cnt[w];
  const int pos = off + __popcll(m & ((1ull << ln) - 1ull));
  __syncthreads();
  total += all;
  return pos;
}

// Periodic box (nbqplist_box :10327-10553, nbqwlist_box :10554-10700): a
// solute charge group enters when any of its atoms (in the group's order)
// lies within Rcq of the Q switching atom (minimum image), and that atom is
// kept as the group's shift atom (qp_shatom); a water when its O does.
// with_waters = 0 rebuilds the Q-solute list only (a rejected volume move
// keeps the Q-water list of the scaled coordinates, as MC_volume does).
__global__ __launch_bounds__(1024) void k_q_lists(const DevSys* __restrict__ sys, int with_waters) {
  const DevSys& S = sys[blockIdx.y];
  if (!S.q_dynamic) return;
  __shared__ int wave_cnt[16];
  unsigned char* fl = const_cast<unsigned char*>(S.flags);
  const double p0 = S.xpcent[0], p1 = S.xpcent[1], p2 = S.xpcent[2];
  auto inside = [&](int a) {
    if (S.pbc) {
      const int q3 = 3 * S.qswitch;
      const double d0 = pbc_image(S, S.x[3 * a] - S.x[q3], 0), d1 = pbc_image(S, S.x[3 * a + 1] - S.x[q3 + 1], 1),
                   d2 = pbc_image(S, S.x[3 * a + 2] - S.x[q3 + 2], 2);
      return d0 * d0 + d1 * d1 + d2 * d2 <= S.rcq2;
    }
    const double d0 = S.x[3 * a] - p0, d1 = S.x[3 * a + 1] - p1, d2 = S.x[3 * a + 2] - p2;
    return !(fl[a] & F_EXCL) && d0 * d0 + d1 * d1 + d2 * d2 <= S.rcq2;
  };
  // groups first (their switch atoms' flags are read before any flag of
  // this instance is rewritten), then the atoms
  for (int g = threadIdx.x; g < S.ncgp_solute; g += blockDim.x) {
    if (S.pbc) {
      // the whole group's atoms (all slices of a split group see the same)
      int sa = -1;
      for (int a = S.cgp_cfirst[g]; a <= S.cgp_clast[g] && sa < 0; ++a)
        if (inside(S.cgpatom[a])) sa = S.cgpatom[a];
      S.qp_shatom[g] = sa;
      S.qp_gin[g] = sa >= 0 ? 1 : 0;
    } else {
      S.qp_gin[g] = inside(S.cgp_sw[g]) ? 1 : 0;
    }
  }
  __threadfence_block();
  __syncthreads();
  int nqp = 0;
  for (int base = 0; base < S.nat_solute; base += blockDim.x) {
    const int i = base + threadIdx.x;
    bool take = false;
    if (i < S.nat_solute) {
      take = S.qp_cand[i] && S.qp_gin[S.iwhich_cgp[i]];
      fl[i] = (unsigned char)(take ? (fl[i] | F_QP) : (fl[i] & ~F_QP));
    }
    const int pos = ordered_compact(take, wave_cnt, nqp);
    if (take) S.qp_atoms[pos] = i;
  }
  if (!with_waters) {
    if (threadIdx.x == 0) S.q_counts[0] = nqp;
    return;
  }
  int nqw = 0;
  for (int base = 0; base < S.nwat; base += blockDim.x) {
    const int iw = base + threadIdx.x;
    bool take = false;
    if (iw < S.nwat) {
      const int o = S.nat_solute + 3 * iw;
      take = inside(o);
      fl[o] = (unsigned char)(take ? (fl[o] | F_QW) : (fl[o] & ~F_QW));
    }
    const int pos = ordered_compact(take, wave_cnt, nqw);
    if (take) S.qw_waters[pos] = iw;
  }
  if (threadIdx.x == 0) { S.q_counts[0] = nqp; S.q_counts[1] = nqw; }
}

// nonbond_monitor (src/md.f90:11228-11350): Coulomb and LJ energies of every
// atom pair of a monitored group pair, per state (Q charges and, with Q-atom
// types, Q LJ parameters and soft-core of the state), FP64; then the
// lambda-weighted sums.  One workgroup per (pair, instance); excluded atoms
// frozen by excluded_freeze are at their topology positions, as after
// fix_shell.
__global__ __launch_bounds__(256) void k_monitor(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int par = blockIdx.x;
  if (par >= S.mon_n) return;
  const int ns = S.nstates > 0 ? S.nstates : 1;
  const int ni = S.mon_ni[par], nj = S.mon_nj[par];
  const int* ai = S.mon_at + S.mon_fi[par];
  const int* aj = S.mon_at + S.mon_fj[par];
  const signed char* ljc = S.mon_lj + S.mon_lo[par];
  auto pos = [&](int a, int c) {
    return (S.freeze && (S.flags[a] & F_EXCL)) ? S.xtop[3 * a + c] : S.x[3 * a + c];
  };
  double vel[kMaxQStates], vlj[kMaxQStates];
  for (int st = 0; st < kMaxQStates; ++st) { vel[st] = 0.0; vlj[st] = 0.0; }
  for (int p = threadIdx.x; p < ni * nj; p += blockDim.x) {
    const int i = p / nj, j = p % nj;
    const int a = ai[i], b = aj[j];
    const int qa = S.iqatom_d[a], qb = S.iqatom_d[b];
    double dx1 = pos(a, 0) - pos(b, 0), dx2 = pos(a, 1) - pos(b, 1), dx3 = pos(a, 2) - pos(b, 2);
    if (S.pbc) { dx1 = pbc_image(S, dx1, 0); dx2 = pbc_image(S, dx2, 1); dx3 = pbc_image(S, dx3, 2); }   // :11269-11273
    const double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    const double r = sqrt(1.0 / r2);
    const double r6_hc = r2 * r2 * r2;
    for (int st = 0; st < ns; ++st) {
      const int lj = ljc[((size_t)st * ni + i) * nj + j];
      const int ta = S.iac[a], tb = S.iac[b];
      double aLJi = S.avdw[ta * 3 + lj - 1], bLJi = S.bvdw[ta * 3 + lj - 1];
      double aLJj = S.avdw[tb * 3 + lj - 1], bLJj = S.bvdw[tb * 3 + lj - 1];
      const double qi = qa ? S.qcrg[(qa - 1) * S.nstates + st] : S.crg[a];
      const double qj = qb ? S.qcrg[(qb - 1) * S.nstates + st] : S.crg[b];
      double r6 = 1.0 / r6_hc;
      if (S.qvdw && (qa || qb)) {
        int sc1, sc2;
        if (qa) {
          const int t = S.qiac[(qa - 1) * S.nstates + st];
          aLJi = S.qavdw[3 * t + lj - 1]; bLJi = S.qbvdw[3 * t + lj - 1];
          sc1 = qa - 1;
          if (qb) {
            const int u = S.qiac[(qb - 1) * S.nstates + st];
            aLJj = S.qavdw[3 * u + lj - 1]; bLJj = S.qbvdw[3 * u + lj - 1];
            sc2 = S.natyps + qb - 1;
          } else {
            sc2 = tb;
          }
        } else {
          const int u = S.qiac[(qb - 1) * S.nstates + st];
          aLJj = S.qavdw[3 * u + lj - 1]; bLJj = S.qbvdw[3 * u + lj - 1];
          sc1 = qb - 1;
          sc2 = ta;
        }
        r6 = 1.0 / (r6_hc + sc_alpha(S, sc1, sc2, st));
      }
      vel[st] = vel[st] + qi * qj * r;
      vlj[st] = vlj[st] + (aLJi * aLJj * r6 * r6 - bLJi * bLJj * r6);
    }
  }
  __shared__ double red[2 * kMaxQStates][4];
  double* o = S.mon_out + (size_t)par * (kMon + 2 * ns);
  double wel = 0.0, wlj = 0.0;
  for (int st = 0; st < ns; ++st) {
    const double e1 = block_sum_256(vel[st], red[2 * st]);
    const double e2 = block_sum_256(vlj[st], red[2 * st + 1]);
    if (threadIdx.x == 0) {
      const double lam = S.nstates > 0 ? S.lambda[st] : 1.0;
      o[kMon + 2 * st] = e1;
      o[kMon + 2 * st + 1] = e2;
      wel = wel + e1 * lam;
      wlj = wlj + e2 * lam;
    }
  }
  if (threadIdx.x == 0) { o[0] = wlj + wel; o[1] = wel; o[2] = wlj; }
}

// Sum of the Q-atom chunks: gradient added to d, per-state energies stored.
// With shadow != 0 (FP64 re-evaluation of a mixed-precision batch on an
// energy-file step) only the energies are stored, into e_q64.
__global__ __launch_bounds__(64) void k_q_reduce(const DevSys* __restrict__ sys, int shadow) {
  const DevSys& S = sys[blockIdx.y];
  if (shadow && !shadow_record_step(S)) return;
  const int iq = blockIdx.x * blockDim.x + threadIdx.x;
  if (iq >= S.nqat) return;
  const int ns = S.nstates;
  const int n = 3 + 4 * ns;
  const int qi = S.iqseq[iq];
  const double* part = shadow ? S.q_part64 : S.q_part;
  double* eq = shadow ? S.e_q64 : S.e_q;
  for (int c = shadow ? 3 : 0; c < n; ++c) {
    double v = 0.0;
    for (int k = 0; k < kQChunks; ++k) v = v + part[(size_t)(iq * kQChunks + k) * kQPartStride + c];
    if (!(S.term_mask & TM_QX)) v = 0.0;
    if (c < 3) S.d[3 * qi + c] = S.d[3 * qi + c] + v;
    else eq[(size_t)iq * ns * 4 + (c - 3)] = v;
  }
}

// ---------------------------------------------------------------------------
// Bonded terms (bond :1332-1405, angle :1130-1246, torsion :17907-18052,
// improper :3130-3416) and Q-Q / Q-neighbour pairs (nonbond_qq
// :13491-13636), one thread per term: the term's energy is stored per term
// and the gradient contribution of each of its atoms into that atom's slot
// (slots of term n are consecutive: bonds i j, angles i j k, torsions and
// impropers i j k l, Q pairs i j); k_term_gather
// then adds every atom's slots to d in a fixed order.  FP64 in both
// precisions.
__device__ __forceinline__ void put_slot(const DevSys& S, int slot, double g0, double g1, double g2) {
  double* o = S.slot_g + 3 * (size_t)S.slot_pos[slot];
  o[0] = g0; o[1] = g1; o[2] = g2;
}

// Dihedral angle i-j-k-l and the 12 components of its derivative times
// -1/sin(phi), as torsion/improper form them (src/md.f90:3160-3244).
__device__ double dihedral_dp(const double* x, int i, int j, int k, int l, double dp[12]) {
  double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
  for (int c = 0; c < 3; ++c) {
    rji[c] = x[3 * i + c] - x[3 * j + c];
    rjk[c] = x[3 * k + c] - x[3 * j + c];
    rkl[c] = x[3 * l + c] - x[3 * k + c];
  }
  rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
  rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
  rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
  rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
  rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
  rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
  const double bj2inv = 1.0 / (rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
  const double bk2inv = 1.0 / (rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
  const double bjinv = sqrt(bj2inv), bkinv = sqrt(bk2inv);
  double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) * (bjinv * bkinv);
  if (scp > 1.0) scp = 1.0;
  else if (scp < -1.0) scp = -1.0;
  double phi = acos(scp);
  if (rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
          rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]) < 0)
    phi = -phi;
  double f1 = sin(phi);
  if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
  f1 = -1.0 / f1;
  double di[3], dl[3], rki[3], rlj[3];
  for (int c = 0; c < 3; ++c) {
    di[c] = f1 * (rnk[c] * bjinv * bkinv - scp * rnj[c] * bj2inv);
    dl[c] = f1 * (rnj[c] * bjinv * bkinv - scp * rnk[c] * bk2inv);
    rki[c] = rji[c] - rjk[c];
    rlj[c] = -rjk[c] - rkl[c];
  }
  dp[0] = rjk[1] * di[2] - rjk[2] * di[1];
  dp[1] = rjk[2] * di[0] - rjk[0] * di[2];
  dp[2] = rjk[0] * di[1] - rjk[1] * di[0];
  dp[3] = rki[1] * di[2] - rki[2] * di[1] + rkl[1] * dl[2] - rkl[2] * dl[1];
  dp[4] = rki[2] * di[0] - rki[0] * di[2] + rkl[2] * dl[0] - rkl[0] * dl[2];
  dp[5] = rki[0] * di[1] - rki[1] * di[0] + rkl[0] * dl[1] - rkl[1] * dl[0];
  dp[6] = rlj[1] * dl[2] - rlj[2] * dl[1] - rji[1] * di[2] + rji[2] * di[1];
  dp[7] = rlj[2] * dl[0] - rlj[0] * dl[2] - rji[2] * di[0] + rji[0] * di[2];
  dp[8] = rlj[0] * dl[1] - rlj[1] * dl[0] - rji[0] * di[1] + rji[1] * di[0];
  dp[9] = rjk[1] * dl[2] - rjk[2] * dl[1];
  dp[10] = rjk[2] * dl[0] - rjk[0] * dl[2];
  dp[11] = rjk[0] * dl[1] - rjk[1] * dl[0];
  return phi;
}

// Dihedral as qtorsion/qimproper form it (divisions by bj*bk and bj**2,
// src/md.f90:17295-17338): phi and its derivative components times -1/sin.
__device__ double qdihedral_dp(const double* x, int i, int j, int k, int l, double dp[12]) {
  double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
  for (int c = 0; c < 3; ++c) {
    rji[c] = x[3 * i + c] - x[3 * j + c];
    rjk[c] = x[3 * k + c] - x[3 * j + c];
    rkl[c] = x[3 * l + c] - x[3 * k + c];
  }
  rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
  rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
  rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
  rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
  rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
  rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
  const double bj = sqrt(rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
  const double bk = sqrt(rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
  double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) / (bj * bk);
  if (scp > 1.0) scp = 1.0;
  if (scp < -1.0) scp = -1.0;
  double phi = acos(scp);
  if (rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
          rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]) < 0)
    phi = -phi;
  double f1 = sin(phi);
  if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
  f1 = -1.0 / f1;
  double di[3], dl[3], rki[3], rlj[3];
  for (int c = 0; c < 3; ++c) {
    di[c] = f1 * (rnk[c] / (bj * bk) - scp * rnj[c] / (bj * bj));
    dl[c] = f1 * (rnj[c] / (bj * bk) - scp * rnk[c] / (bk * bk));
    rki[c] = rji[c] - rjk[c];
    rlj[c] = -rjk[c] - rkl[c];
  }
  dp[0] = rjk[1] * di[2] - rjk[2] * di[1];
  dp[1] = rjk[2] * di[0] - rjk[0] * di[2];
  dp[2] = rjk[0] * di[1] - rjk[1] * di[0];
  dp[3] = rki[1] * di[2] - rki[2] * di[1] + rkl[1] * dl[2] - rkl[2] * dl[1];
  dp[4] = rki[2] * di[0] - rki[0] * di[2] + rkl[2] * dl[0] - rkl[0] * dl[2];
  dp[5] = rki[0] * di[1] - rki[1] * di[0] + rkl[0] * dl[1] - rkl[1] * dl[0];
  dp[6] = rlj[1] * dl[2] - rlj[2] * dl[1] - rji[1] * di[2] + rji[2] * di[1];
  dp[7] = rlj[2] * dl[0] - rlj[0] * dl[2] - rji[2] * di[0] + rji[0] * di[2];
  dp[8] = rlj[0] * dl[1] - rlj[1] * dl[0] - rji[0] * di[1] + rji[1] * di[0];
  dp[9] = rjk[1] * dl[2] - rjk[2] * dl[1];
  dp[10] = rjk[2] * dl[0] - rjk[0] * dl[2];
  dp[11] = rjk[0] * dl[1] - rjk[1] * dl[0];
  return phi;
}

// Morse Q bond ib of parameters P = (Dmz, amz, r0, fk) at state weight lam
// (qbond, src/md.f90:17213-17242): energy, gradient factor dv (gradient on
// j = dv * rij), and the Morse factor record (EMorseD, dMorse_i, dMorse_j).
__device__ double qbond_eval(const DevSys& S, int ib, const double* P, double lam, double rij[3], double& dv,
                             double rec[kMorseRec]) {
  const double* x = S.xq;
  const int i3 = 3 * S.qb_atoms[2 * ib], j3 = 3 * S.qb_atoms[2 * ib + 1];
  for (int c = 0; c < 3; ++c) rij[c] = x[j3 + c] - x[i3 + c];
  const double b = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
  const double db = b - P[2];
  const double fexp = exp(-P[1] * db);
  const double be = P[0] * (fexp * fexp - 2.0 * fexp) + 0.5 * P[3] * (db * db);
  dv = (2.0 * P[0] * P[1] * (fexp - fexp * fexp) + P[3] * db) * lam / b;
  if (db > 0) {
    rec[0] = -(fexp * fexp - 2.0 * fexp);
    for (int c = 0; c < 3; ++c) {
      rec[1 + c] = +2.0 * P[1] * (fexp - fexp * fexp) * lam / b * rij[c];
      rec[4 + c] = -2.0 * P[1] * (fexp - fexp * fexp) * lam / b * rij[c];
    }
  } else {
    rec[0] = 1.0;
    for (int c = 0; c < 6; ++c) rec[1 + c] = 0.0;
  }
  return be;
}

// Morse factor record of the bond a coupled term sees (see DevSys::xt).
__device__ void coupled_morse(const DevSys& S, const int* R, const double* P, double rec[kMorseRec]) {
  const int ib = R[6], seff = R[8];
  if (seff >= 0) {
    double rij[3], dv;
    qbond_eval(S, ib, P + 4, S.lambda[seff], rij, dv, rec);
  } else {
    for (int c = 0; c < kMorseRec; ++c) rec[c] = S.qb_prev[kMorseRec * ib + c];
  }
}

// One extra term (XtKind record n): gradient into its slots, energy e_xt[n].
__device__ void xterm(const DevSys& S, int n) {
  const int* R = S.xt + (size_t)kXtInts * n;
  const double* P = S.xp + (size_t)kXtPars * n;
  const double* x = S.xq;
  const int kind = R[0], s0 = R[9];
  double e = 0.0;
  switch (kind) {
    case XT_QBOND: {
      // R: kind i j ib last-present-state state
      double rij[3], dv, rec[kMorseRec];
      e = qbond_eval(S, R[3], P, S.lambda[R[5]], rij, dv, rec);
      put_slot(S, s0, -dv * rij[0], -dv * rij[1], -dv * rij[2]);
      put_slot(S, s0 + 1, dv * rij[0], dv * rij[1], dv * rij[2]);
      if (R[4]) for (int c = 0; c < kMorseRec; ++c) S.qb_cur[kMorseRec * R[3] + c] = rec[c];
      break;
    }
    case XT_QANGLE: {
      // qangle (src/md.f90:17022-17114); R[6] coupled bond, R[7] flag, R[8] its state
      double gamma = 1.0, rec[kMorseRec];
      if (R[6] >= 0) {
        coupled_morse(S, R, P, rec);
        gamma = rec[0];
        if (R[7] == 1) gamma = 1.0 - gamma;
      }
      const int i = R[1], j = R[2], k = R[3];
      double rji[3], rjk[3];
      for (int c = 0; c < 3; ++c) { rji[c] = x[3 * i + c] - x[3 * j + c]; rjk[c] = x[3 * k + c] - x[3 * j + c]; }
      const double bji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
      const double bjk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
      double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]);
      scp = scp / (bji * bjk);
      if (scp > 1.0) scp = 1.0;
      if (scp < -1.0) scp = -1.0;
      const double ang = acos(scp);
      const double da = ang - P[1];
      const double ae = 0.5 * P[0] * (da * da);
      e = ae * gamma;
      const double dv = gamma * P[0] * da * S.lambda[R[5]];
      double f1 = sin(ang);
      if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
      f1 = -1.0 / f1;
      double di[3], dk[3];
      for (int c = 0; c < 3; ++c) {
        di[c] = f1 * (rjk[c] / (bji * bjk) - scp * rji[c] / (bji * bji));
        dk[c] = f1 * (rji[c] / (bji * bjk) - scp * rjk[c] / (bjk * bjk));
      }
      put_slot(S, s0, dv * di[0], dv * di[1], dv * di[2]);
      put_slot(S, s0 + 1, -dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2]));
      put_slot(S, s0 + 2, dv * dk[0], dv * dk[1], dv * dk[2]);
      if (R[6] >= 0) {
        put_slot(S, s0 + 3, rec[1] * ae, rec[2] * ae, rec[3] * ae);
        put_slot(S, s0 + 4, rec[4] * ae, rec[5] * ae, rec[6] * ae);
      }
      break;
    }
    case XT_QTORSION:
    case XT_QIMPROPER: {
      // qtorsion (src/md.f90:17395-17540), qimproper (:17249-17392)
      double gamma = 1.0, rec[kMorseRec];
      if (R[6] >= 0) {
        coupled_morse(S, R, P, rec);
        gamma = rec[0];
        if (R[7] == 1) gamma = 1 - gamma;
      }
      double dp[12];
      const double phi = qdihedral_dp(x, R[1], R[2], R[3], R[4], dp);
      double pe, dv;
      if (kind == XT_QTORSION) {
        const double arg = P[1] * phi - P[2];
        pe = P[0] * (1.0 + cos(arg));
        dv = -P[1] * P[0] * sin(arg) * gamma * S.lambda[R[5]];
      } else {
        double arg = phi - P[1];
        arg = arg - 2.0 * kPiD * round(arg / (2.0 * kPiD));
        dv = P[0] * arg;
        pe = 0.5 * dv * arg;
        dv = dv * gamma * S.lambda[R[5]];
      }
      e = pe * gamma;
      for (int m = 0; m < 4; ++m) put_slot(S, s0 + m, dv * dp[3 * m], dv * dp[3 * m + 1], dv * dp[3 * m + 2]);
      if (R[6] >= 0) {
        put_slot(S, s0 + 4, rec[1] * pe, rec[2] * pe, rec[3] * pe);
        put_slot(S, s0 + 5, rec[4] * pe, rec[5] * pe, rec[6] * pe);
      }
      break;
    }
    case XT_RST_SEQ: {
      // sequence restraint of one atom to its topology position (:15549-15575)
      const int i3 = 3 * R[1];
      double d0 = x[i3] - S.xtop[i3], d1 = x[i3 + 1] - S.xtop[i3 + 1], d2 = x[i3 + 2] - S.xtop[i3 + 2];
      // periodic box: the minimal periodic distance (:15561-15565)
      if (S.pbc) { d0 = pbc_image(S, d0, 0); d1 = pbc_image(S, d1, 1); d2 = pbc_image(S, d2, 2); }
      e = 0.5 * P[0] * (d0 * d0 + d1 * d1 + d2 * d2);
      put_slot(S, s0, P[0] * d0, P[0] * d1, P[0] * d2);
      break;
    }
    case XT_RST_SEQC: {
      // sequence restraint of a geometric (R[3] = 1) or mass (2) centre
      // (:15486-15547): R[1] first entry in rst_atoms, R[2] count
      const int* A = S.rst_atoms + R[1];
      const int na = R[2];
      double dr[3] = {0, 0, 0}, norm = 0;
      for (int m = 0; m < na; ++m) {
        const int i3 = 3 * A[m];
        const double w = R[3] == 1 ? 1.0 : S.mass[A[m]];
        for (int c = 0; c < 3; ++c)
          dr[c] = R[3] == 1 ? dr[c] + x[i3 + c] - S.xtop[i3 + c] : dr[c] + (x[i3 + c] - S.xtop[i3 + c]) * w;
        norm = norm + w;
      }
      for (int c = 0; c < 3; ++c) dr[c] = dr[c] / norm;
      e = 0.5 * P[0] * (dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      for (int m = 0; m < na; ++m) {
        const double ms = S.mass[A[m]];
        if (R[3] == 1) put_slot(S, s0 + m, P[0] * dr[0] * ms / 12.010, P[0] * dr[1] * ms / 12.010, P[0] * dr[2] * ms / 12.010);
        else put_slot(S, s0 + m, P[0] * dr[0], P[0] * dr[1], P[0] * dr[2]);
      }
      break;
    }
    case XT_RST_POS: {
      // position restraint (:15580-15619), P: x0[3] fk[3]
      const int i3 = 3 * R[1];
      const double wgt = R[5] != 0 ? S.lambda[R[5] - 1] : 1.0;
      double dr[3];
      for (int c = 0; c < 3; ++c) dr[c] = x[i3 + c] - P[c];
      e = 0.5 * P[3] * (dr[0] * dr[0]) + 0.5 * P[4] * (dr[1] * dr[1]) + 0.5 * P[5] * (dr[2] * dr[2]);
      put_slot(S, s0, P[3] * dr[0] * wgt, P[4] * dr[1] * wgt, P[5] * dr[2] * wgt);
      break;
    }
    case XT_RST_DIST: {
      // flat-bottom distance restraint (:15622-15670), P: d1 d2 fk
      const int i3 = 3 * R[1], j3 = 3 * R[2];
      const double wgt = R[5] != 0 ? S.lambda[R[5] - 1] : 1.0;
      double dr[3];
      for (int c = 0; c < 3; ++c) dr[c] = x[j3 + c] - x[i3 + c];
      if (S.pbc) for (int c = 0; c < 3; ++c) dr[c] = pbc_image(S, dr[c], c);   // :15630-15634
      const double b = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double db = 0.0;
      if (b < P[0]) db = b - P[0];
      else if (b > P[1]) db = b - P[1];
      const double dv = db != 0.0 ? wgt * P[2] * db / b : 0.0;
      e = db != 0.0 ? 0.5 * P[2] * (db * db) : 0.0;
      put_slot(S, s0, -dr[0] * dv, -dr[1] * dv, -dr[2] * dv);
      put_slot(S, s0 + 1, dr[0] * dv, dr[1] * dv, dr[2] * dv);
      break;
    }
    case XT_RST_ANG: {
      // angle restraint (:15673-15762), P: theta0 (radians) fk
      const int i3 = 3 * R[1], j3 = 3 * R[2], k3 = 3 * R[3];
      const double wgt = R[5] != 0 ? S.lambda[R[5] - 1] : 1.0;
      double dr[3], dr2[3];
      for (int c = 0; c < 3; ++c) { dr[c] = x[i3 + c] - x[j3 + c]; dr2[c] = x[k3 + c] - x[j3 + c]; }
      if (S.pbc)   // :15694-15702
        for (int c = 0; c < 3; ++c) { dr[c] = pbc_image(S, dr[c], c); dr2[c] = pbc_image(S, dr2[c], c); }
      const double r2ij = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      const double r2jk = dr2[0] * dr2[0] + dr2[1] * dr2[1] + dr2[2] * dr2[2];
      const double rij = sqrt(r2ij), rjk = sqrt(r2jk);
      double scp = (dr[0] * dr2[0] + dr[1] * dr2[1] + dr[2] * dr2[2]);
      scp = scp / (rij * rjk);
      if (scp > 1.0) scp = 1.0;
      if (scp < -1.0) scp = -1.0;
      const double theta = acos(scp);
      const double db = theta - P[0];
      e = 0.5 * P[1] * (db * db);
      const double dv = wgt * P[1] * db;
      double f1 = sin(theta);
      if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
      f1 = -1.0 / f1;
      double di[3], dk[3];
      for (int c = 0; c < 3; ++c) {
        di[c] = f1 * (dr2[c] / (rij * rjk) - scp * dr[c] / r2ij);
        dk[c] = f1 * (dr[c] / (rij * rjk) - scp * dr2[c] / r2jk);
      }
      put_slot(S, s0, dv * di[0], dv * di[1], dv * di[2]);
      put_slot(S, s0 + 1, -dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2]));
      put_slot(S, s0 + 2, dv * dk[0], dv * dk[1], dv * dk[2]);
      break;
    }
    case XT_RST_WALL: {
      // half-harmonic wall with Morse attraction (:15765-15797), P: d fk aMorse dMorse
      const int i3 = 3 * R[1];
      double dr[3];
      for (int c = 0; c < 3; ++c) dr[c] = x[i3 + c] - S.xwcent[c];
      const double b = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      const double db = b - P[0];
      double dv;
      if (db > 0) {
        e = 0.5 * P[1] * (db * db) - P[3];
        dv = P[1] * db / b;
      } else {
        const double fexp = exp(P[2] * db);
        e = P[3] * (fexp * fexp - 2.0 * fexp);
        dv = -2.0 * P[3] * P[2] * (fexp - fexp * fexp) / b;
      }
      put_slot(S, s0, dv * dr[0], dv * dr[1], dv * dr[2]);
      break;
    }
    default: break;
  }
  S.e_xt[n] = e;
}

__global__ __launch_bounds__(256) void k_terms(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  const double* x = S.x;
  if (n < S.nbnd) {
    const int i = S.bnd[3 * n], j = S.bnd[3 * n + 1], cod = S.bnd[3 * n + 2] - 1;
    const double fk = S.bondlib[2 * cod], b0 = S.bondlib[2 * cod + 1];
    const double r0 = x[3 * j] - x[3 * i], r1 = x[3 * j + 1] - x[3 * i + 1], r2 = x[3 * j + 2] - x[3 * i + 2];
    const double b = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    const double db = b - b0;
    const double dv = fk * db / b;
    put_slot(S, 2 * n, -r0 * dv, -r1 * dv, -r2 * dv);
    put_slot(S, 2 * n + 1, r0 * dv, r1 * dv, r2 * dv);
    S.e_bterm[n] = 0.5 * fk * (db * db);
    return;
  }
  n -= S.nbnd;
  if (n < S.nang) {
    const int i = S.ang[4 * n], j = S.ang[4 * n + 1], kk = S.ang[4 * n + 2], cod = S.ang[4 * n + 3] - 1;
    const double fk = S.anglib[2 * cod], a0 = S.anglib[2 * cod + 1];
    double rji[3], rjk[3];
    for (int c = 0; c < 3; ++c) { rji[c] = x[3 * i + c] - x[3 * j + c]; rjk[c] = x[3 * kk + c] - x[3 * j + c]; }
    const double bji2inv = 1.0 / (rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    const double bjk2inv = 1.0 / (rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    const double bjiinv = sqrt(bji2inv), bjkinv = sqrt(bjk2inv);
    double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]);
    scp = scp * bjiinv * bjkinv;
    if (scp > 1.0) scp = 1.0;
    else if (scp < -1.0) scp = -1.0;
    const double angv = acos(scp);
    const double da = angv - a0;
    const double dv = fk * da;
    double f1 = sin(angv);
    if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dk[3];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (rjk[c] * bjiinv * bjkinv - scp * rji[c] * bji2inv);
      dk[c] = f1 * (rji[c] * bjiinv * bjkinv - scp * rjk[c] * bjk2inv);
    }
    const int s0 = 2 * S.nbnd + 3 * n;
    put_slot(S, s0, dv * di[0], dv * di[1], dv * di[2]);
    put_slot(S, s0 + 1, -dv * (di[0] + dk[0]), -dv * (di[1] + dk[1]), -dv * (di[2] + dk[2]));
    put_slot(S, s0 + 2, dv * dk[0], dv * dk[1], dv * dk[2]);
    S.e_bterm[S.nbnd + n] = 0.5 * fk * (da * da);
    return;
  }
  n -= S.nang;
  if (n < S.ntor) {
    const int* T = S.tor + 5 * n;
    const int i = T[0], j = T[1], kk = T[2], l = T[3], cod = T[4] - 1;
    const double* L = S.torlib + 4 * cod;
    double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
    for (int c = 0; c < 3; ++c) {
      rji[c] = x[3 * i + c] - x[3 * j + c];
      rjk[c] = x[3 * kk + c] - x[3 * j + c];
      rkl[c] = x[3 * l + c] - x[3 * kk + c];
    }
    rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
    rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
    rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
    rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
    rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
    rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
    const double bj2inv = 1.0 / (rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
    const double bk2inv = 1.0 / (rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
    const double bjinv = sqrt(bj2inv), bkinv = sqrt(bk2inv);
    double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) * (bjinv * bkinv);
    double phi;
    if (scp > 1.0) { scp = 1.0; phi = acos(1.0); }
    else if (scp < -1.0) { scp = -1.0; phi = acos(-1.0); }
    else phi = acos(scp);
    if (rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
            rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]) < 0)
      phi = -phi;
    const double arg = L[1] * phi - L[2];
    const double dv = -L[1] * L[0] * sin(arg) * L[3];
    double f1 = sin(phi);
    if (fabs(f1) < qsys::kQrealEps) f1 = qsys::kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dl[3], rki[3], rlj[3];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (rnk[c] * (bjinv * bkinv) - scp * rnj[c] * bj2inv);
      dl[c] = f1 * (rnj[c] * (bjinv * bkinv) - scp * rnk[c] * bk2inv);
      rki[c] = rji[c] - rjk[c];
      rlj[c] = -rjk[c] - rkl[c];
    }
    const int s0 = 2 * S.nbnd + 3 * S.nang + 4 * n;
    put_slot(S, s0, dv * (rjk[1] * di[2] - rjk[2] * di[1]), dv * (rjk[2] * di[0] - rjk[0] * di[2]),
             dv * (rjk[0] * di[1] - rjk[1] * di[0]));
    put_slot(S, s0 + 1, dv * (rki[1] * di[2] - rki[2] * di[1] + rkl[1] * dl[2] - rkl[2] * dl[1]),
             dv * (rki[2] * di[0] - rki[0] * di[2] + rkl[2] * dl[0] - rkl[0] * dl[2]),
             dv * (rki[0] * di[1] - rki[1] * di[0] + rkl[0] * dl[1] - rkl[1] * dl[0]));
    put_slot(S, s0 + 2, dv * (rlj[1] * dl[2] - rlj[2] * dl[1] - rji[1] * di[2] + rji[2] * di[1]),
             dv * (rlj[2] * dl[0] - rlj[0] * dl[2] - rji[2] * di[0] + rji[0] * di[2]),
             dv * (rlj[0] * dl[1] - rlj[1] * dl[0] - rji[0] * di[1] + rji[1] * di[0]));
    put_slot(S, s0 + 3, dv * (rjk[1] * dl[2] - rjk[2] * dl[1]), dv * (rjk[2] * dl[0] - rjk[0] * dl[2]),
             dv * (rjk[0] * dl[1] - rjk[1] * dl[0]));
    S.e_bterm[S.nbnd + S.nang + n] = L[0] * (1.0 + cos(arg)) * L[3];
    return;
  }
  n -= S.ntor;
  if (n < S.nimp) {
    // improper (src/md.f90:3130-3272, harmonic in the wrapped angle) or
    // improper2 (:3276-3416, periodic fk (1 + cos(2 phi - imp0)))
    const int* T = S.imp + 5 * n;
    const double* L = S.implib + 2 * (T[4] - 1);
    double dp[12];
    const double phi = dihedral_dp(x, T[0], T[1], T[2], T[3], dp);
    double dv, e;
    if (S.ff_type == 2) {
      const double arg = 2 * phi - L[1];
      e = L[0] * (1 + cos(arg));
      dv = -2 * L[0] * sin(arg);
    } else {
      double arg = phi - L[1];
      arg = arg - 2.0 * kPiD * round(arg / (2.0 * kPiD));
      dv = L[0] * arg;
      e = 0.5 * dv * arg;
    }
    const int s0 = 2 * S.nbnd + 3 * S.nang + 4 * S.ntor + 4 * n;
    for (int m = 0; m < 4; ++m) put_slot(S, s0 + m, dv * dp[3 * m], dv * dp[3 * m + 1], dv * dp[3 * m + 2]);
    S.e_bterm[S.nbnd + S.nang + S.ntor + n] = e;
    return;
  }
  n -= S.nimp;
  // the terms below follow fix_shell in pot_energy: with excluded_freeze
  // they see the excluded atoms back at their topology positions
  x = S.xq;
  if (n < S.nqq) {
    const int p = n;
    const int st = S.qq_state[p], iq = S.qq_iq[p], j = S.qq_j[p], jq = S.qq_jq[p], lj = S.qq_lj[p];
    const int i = S.iqseq[iq];
    const int l = lj - 1;
    const double qi = S.qcrg[iq * S.nstates + st];
    double qi_used = qi;
    double aLJ, bLJ, qj;
    if (!S.qvdw) {
      aLJ = S.avdw[S.iac[i] * 3 + l] * S.avdw[S.iac[j] * 3 + l];
      bLJ = S.bvdw[S.iac[i] * 3 + l] * S.bvdw[S.iac[j] * 3 + l];
      qj = jq >= 0 ? S.qcrg[jq * S.nstates + st] : S.crg[j];
    } else {
      // Q-atom types (src/md.f90:13531-13550)
      const int qti = S.qiac[iq * S.nstates + st];
      aLJ = S.qavdw[3 * qti + l];
      bLJ = S.qbvdw[3 * qti + l];
      if (jq >= 0) {
        const int qtj = S.qiac[jq * S.nstates + st];
        aLJ = aLJ * S.qavdw[3 * qtj + l];
        bLJ = bLJ * S.qbvdw[3 * qtj + l];
        qj = S.qcrg[jq * S.nstates + st];
      } else {
        if (lj == 2) {
          aLJ = S.qavdw[3 * qti];
          bLJ = S.qbvdw[3 * qti];
        }
        aLJ = aLJ * S.avdw[S.iac[j] * 3 + l];
        bLJ = bLJ * S.bvdw[S.iac[j] * 3 + l];
        qj = S.crg[j];
      }
    }
    if (S.qq_lib) {
      // nonbond_qq_lib_charges (src/md.f90:13640-13700)
      qi_used = jq >= 0 ? S.crg[i] : qi;
      qj = S.crg[j];
    }
    const double dx1 = x[3 * j] - x[3 * i], dx2 = x[3 * j + 1] - x[3 * i + 1], dx3 = x[3 * j + 2] - x[3 * i + 2];
    double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    const double r6_hc = r2 * r2 * r2;
    // soft-core column natyps + jq (jq = -1 for a non-Q partner reads the
    // last topology type's column, as the reference's natyps + 0)
    double r6 = r2 * r2 * r2 + sc_alpha(S, iq, S.natyps + jq, st);
    r6 = 1.0 / r6;
    const double r12 = r6 * r6;
    r2 = 1.0 / r2;
    const double r = sqrt(r2);
    double Vel = qi_used * qj * r * S.qq_elscale[p];
    if (lj == 3) Vel = Vel * S.el14_scale;
    double Va, Vb, dv;
    if (S.qvdw && jq >= 0 && lj == 2) {
      Va = aLJ * exp(-bLJ / r);
      Vb = 0.0;
      dv = r2 * (-Vel - bLJ * Va / r) * S.lambda[st];
    } else {
      Va = aLJ * r12;
      Vb = bLJ * r6;
      dv = r2 * (-Vel - (12.0 * Va - 6.0 * Vb) * r6_hc * r6) * S.lambda[st];
    }
    const int s0 = 2 * S.nbnd + 3 * S.nang + 4 * S.ntor + 4 * S.nimp + 2 * p;
    put_slot(S, s0, -dv * dx1, -dv * dx2, -dv * dx3);
    put_slot(S, s0 + 1, dv * dx1, dv * dx2, dv * dx3);
    S.e_qq[2 * p] = Vel;
    S.e_qq[2 * p + 1] = Va - Vb;
    return;
  }
  n -= S.nqq;
  if (n < S.nxt) xterm(S, n);
}

// Adds each atom's term slots to its gradient, in slot order, and keeps the
// Q bonds' Morse factors of this evaluation for the next one.
__global__ __launch_bounds__(256) void k_term_gather(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < S.nqbond && !halted(S, PH_TERMS))
    for (int c = 0; c < kMorseRec; ++c) S.qb_prev[kMorseRec * t + c] = S.qb_cur[kMorseRec * t + c];
  if (t >= S.nslot_atoms) return;
  const int me = S.slot_atom[t];
  double g0 = 0, g1 = 0, g2 = 0;
  const gconst<double> o = gptr(S.slot_g) + 3 * (size_t)S.slot_first[t];
  const int n = S.slot_first[t + 1] - S.slot_first[t];
  // four slots' loads in flight at a time; the sums keep slot order
  int k = 0;
  for (; k + 4 <= n; k += 4) {
    double v[12];
    for (int u = 0; u < 12; ++u) v[u] = o[3 * k + u];
    for (int u = 0; u < 4; ++u) { g0 = g0 + v[3 * u]; g1 = g1 + v[3 * u + 1]; g2 = g2 + v[3 * u + 2]; }
  }
  for (; k < n; ++k) {
    g0 = g0 + o[3 * k]; g1 = g1 + o[3 * k + 1]; g2 = g2 + o[3 * k + 2];
  }
  if (!(S.term_mask & TM_OTHER)) return;
  S.d[3 * me] = S.d[3 * me] + g0;
  S.d[3 * me + 1] = S.d[3 * me + 1] + g1;
  S.d[3 * me + 2] = S.d[3 * me + 2] + g2;
}

}  // namespace qgpu
