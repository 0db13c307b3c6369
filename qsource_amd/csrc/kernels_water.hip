// CDNA4 kernels for the solvent path of qdyn's MD step: charge-group
// centres, water pair lists with the LRF multipole gather, the water
// nonbonded + LRF Taylor + radial restraint kernel, the water-polarisation
// restraint, the fused leapfrog/Berendsen + SHAKE + kinetic-energy update,
// the temperature/Berendsen-scale finalisation and the energy reduction.
//
// Every kernel reads its instance from sys[blockIdx.y], so a batch of
// uneven systems (lambda windows, sphere sizes) shares each launch.
#include "kernels_common.hpp"

namespace qgpu {

// ---------------------------------------------------------------------------
// cgp_centers (src/md.f90:2253-2288): centre of each charge group (the
// slices of a split group take the whole group's); LRF multipoles zeroed
// (they are overwritten by the list kernels for every group that is not
// excluded).
__global__ void k_cgp_centers(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  double c0 = 0, c1 = 0, c2 = 0;
  int f = S.cgp_cfirst[g], l = S.cgp_clast[g];
  for (int a = f; a <= l; ++a) {
    int i3 = 3 * S.cgpatom[a];
    c0 = c0 + S.x[i3];
    c1 = c1 + S.x[i3 + 1];
    c2 = c2 + S.x[i3 + 2];
  }
  double n = (double)(l - f + 1);
  double* L = S.lrf + (size_t)g * kLrfStride;
  L[0] = c0 / n; L[1] = c1 / n; L[2] = c2 / n;
  for (int k = 3; k < kLrfStride; ++k) L[k] = 0.0;
}

// FP32 copy of the coordinates read by the mixed-precision pair kernels
// (refreshed by k_integrate every step, and by this kernel after host writes).
__global__ void k_x_to_f32(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < 3 * S.natom) S.xf[c] = (float)S.x[c];
  // the solute slot records (mixed precision, standard waters)
  if (S.srec && c < S.nat_solute) {
    float* r = S.srec + (size_t)kSrec * S.srec_of_atom[c];
    for (int k = 0; k < 3; ++k) r[k] = (float)S.x[3 * c + k];
  }
  // the water records (mixed precision)
  if (S.xw && c < S.nwat) water_xf_store(S, c, S.x + 3 * (S.nat_solute + 3 * c));
}

// excluded_freeze (fix_shell src/md.f90:2831-2839): the terms evaluated
// after fix_shell in pot_energy see the excluded atoms at their topology
// positions; xq holds those coordinates.
__global__ void k_freeze_x(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!S.freeze || c >= 3 * S.natom) return;
  S.xq[c] = (S.flags[c / 3] & F_EXCL) ? S.xtop[c] : S.x[c];
}

// put_back_in_box (src/md.f90:18743-18930), periodic box, at list updates:
// the box centre follows the solute (all atoms without one) unless it is
// rigid; every molecule of the chosen kinds whose centre of mass lies
// beyond a face moves by one box length along that axis.  One workgroup per
// instance: the centre is summed in atom order by three threads, then each
// thread takes molecules (each summed in atom order), as the reference does.
// The FP32 copies are refreshed afterwards (k_x_to_f32); the LRF centres
// are formed again by the list build that follows.
// boxc[3] holds the step of the last put-back, so that a list update redone
// after a list regrowth does not put molecules back a second time.
__global__ __launch_bounds__(256) void k_put_back(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  if (!S.pbc || !(S.put_solute_back || S.put_solvent_back)) return;
  if (halted(S, PH_LISTS)) return;
  const double st = (double)*S.step_ctr;
  if (S.boxc[3] == st) return;
  __shared__ double boxc[3];
  const int t = threadIdx.x;
  if (t < 3) {
    double c;
    if (!S.rigid_boxc) {
      const int last = S.nat_solute > 0 ? S.nat_solute : S.natom;
      c = 0.0;
      for (int i = 0; i < last; ++i) c = c + S.x[3 * i + t];
      c = c / (double)last;
      S.boxc[t] = c;
    } else {
      c = S.boxc[t];
    }
    boxc[t] = c;
  }
  __syncthreads();
  double hi[3], lo[3];
  for (int k = 0; k < 3; ++k) { hi[k] = boxc[k] + S.boxl[k] / 2; lo[k] = boxc[k] - S.boxl[k] / 2; }
  const int m0 = S.put_solute_back ? 0 : S.nsolute_mol;
  const int m1 = S.put_solvent_back ? S.nmol : S.nsolute_mol;
  for (int m = m0 + t; m < m1; m += blockDim.x) {
    const int a0 = S.mol_start[m], a1 = S.mol_start[m + 1];
    double cm[3] = {0.0, 0.0, 0.0};
    for (int j = a0; j < a1; ++j)
      for (int k = 0; k < 3; ++k) cm[k] = cm[k] + S.x[3 * j + k] * S.mass[j];
    for (int k = 0; k < 3; ++k) {
      cm[k] = cm[k] * S.mol_minv[m];
      double mv = 0.0;
      if (cm[k] > hi[k]) mv = -S.boxl[k];
      else if (cm[k] < lo[k]) mv = S.boxl[k];
      if (mv != 0.0)
        for (int j = a0; j < a1; ++j) S.x[3 * j + k] = S.x[3 * j + k] + mv;
    }
  }
  __syncthreads();
  if (t == 0) S.boxc[3] = st;
}

// A constant-pressure volume move (MC_volume, src/md.f90:19081-19107,
// molecule-based scaling): every molecule of an instance with mc_move set
// moves so that its centre of mass scales about the box centre by the
// ratio of the new to the old box length per axis, ((cm - c) L / L_old +
// c) - cm.  One workgroup per instance, a thread per molecule (its centre
// summed in atom order).  The FP32 copies are refreshed afterwards.
__global__ __launch_bounds__(256) void k_mc_scale(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  if (!S.mc_move) return;
  for (int m = threadIdx.x; m < S.nmol; m += blockDim.x) {
    const int a0 = S.mol_start[m], a1 = S.mol_start[m + 1];
    double cm[3] = {0.0, 0.0, 0.0};
    for (int j = a0; j < a1; ++j)
      for (int k = 0; k < 3; ++k) cm[k] = cm[k] + S.x[3 * j + k] * S.mass[j];
    double mv[3];
    for (int k = 0; k < 3; ++k) {
      cm[k] = cm[k] * S.mol_minv[m];
      mv[k] = ((cm[k] - S.boxc[k]) * S.boxl[k] / S.boxl_old[k] + S.boxc[k]) - cm[k];
    }
    for (int j = a0; j < a1; ++j)
      for (int k = 0; k < 3; ++k) S.x[3 * j + k] = S.x[3 * j + k] + mv[k];
  }
}

// ---------------------------------------------------------------------------
// Water lists (nbwwlist_lrf src/md.f90:10897-11021 and the water side of
// nbpwlist_lrf :9300-9429).  kWLLanes lanes per water (8 in the FP64
// build, 16 in the FP32 one, whose per-lane state is smaller): each
// iteration tests the wavefront's waters against kWLLanes partner waters
// (or solute charge groups), so every partner's coordinates are read once
// for all the wavefront's waters.  Partners inside the cut-off are appended to the water's
// row in ascending order (ballot over the water's 8 lanes); partners between
// the cut-off and RcLRF add their non-Q atoms' multipoles about this water's
// charge-group centre (20 FP64 accumulators per lane, summed over the 8
// lanes at the end).  P = float evaluates each partner group's multipole
// terms in FP32 before adding them to the FP64 sums (mixed precision); list
// membership is decided in FP64 in both.  BOX: periodic box (nbwwlist_box,
// nbpwlist_box_lrf: minimum-image distances, shifted LRF partners).
template <typename P, bool BOX>
__device__ __forceinline__ void water_lists_body(const DevSys& S, const int base) {
  constexpr int kWLLanes = sizeof(P) == 4 ? kWLLanes32 : kWLLanes64;
  const int t = threadIdx.x;
  if (base + (t >> 6) * (64 / kWLLanes) >= S.nwat) return;   // whole wavefronts past the last water
  const int lane = t & 63, sub = t & (kWLLanes - 1), seg = lane & ~(kWLLanes - 1);
  const int iw = base + t / kWLLanes;
  const gconst<double> gx = gptr(S.x);
  const gconst<unsigned char> gfl = gptr(S.flags);
  const int io = S.nat_solute + 3 * (iw < S.nwat ? iw : 0);
  const bool active = iw < S.nwat && !(gfl[io] & F_EXCL);
  const double xi0 = gx[3 * io], xi1 = gx[3 * io + 1], xi2 = gx[3 * io + 2];
  const int ig = S.iwhich_cgp[io];
  double* Lg = S.lrf + (size_t)ig * kLrfStride;
  const double c0 = Lg[0], c1 = Lg[1], c2 = Lg[2];
  const float cf[3] = {(float)c0, (float)c1, (float)c2};
  const double cc[3] = {c0, c1, c2};
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long mine = ((1ull << kWLLanes) - 1ull) << seg;
  LrfAcc acc;
  acc.zero();
  // FP32 build: the raw moments of the far standard waters of kLrfPend
  // partner tiles are summed in FP32, and their components (like those of
  // other far water groups) into FP32 component sums, which join the FP64
  // sums once after the water partners.  The rounding of these sums stays well below that
  // of each partner's FP32 terms, whose atoms' contributions cancel.
  LrfRaw<float> pend;
  pend.zero();
  LrfTerms<float> sum32;
  sum32.zero();
  // waters
  int n_in = 0;
  auto row = gptr_w(S.ww_nb + (size_t)(iw < S.nwat ? iw : 0) * S.ww_cap);
  for (int base = 0, it = 1; base < S.nwat; base += kWLLanes, ++it) {
    const int jw = base + sub;
    bool inside = false, far = false;
    if (active && jw < S.nwat && jw != iw) {
      const int jo = S.nat_solute + 3 * jw;
      if (!(gfl[jo] & F_EXCL)) {
        double a = xi0 - gx[3 * jo], b = xi1 - gx[3 * jo + 1], c = xi2 - gx[3 * jo + 2];
        if constexpr (BOX) { a = pbc_image(S, a, 0); b = pbc_image(S, b, 1); c = pbc_image(S, c, 2); }
        const double r2 = a * a + b * b + c * c;
        inside = r2 <= S.rcww2;
        far = !inside && r2 <= S.rclrf2;
      }
    }
    const unsigned long long m = __ballot(inside) & mine;
    if (inside) {
      const int pos = n_in + __popcll(m & below);
      if (pos < S.ww_cap) row[pos] = jw;
    }
    n_in += __popcll(m);
    if (far && S.use_lrf) {
      const int jo = S.nat_solute + 3 * jw;
      if constexpr (sizeof(P) == 4) {
        if (S.std_water) {
          lrf_raw_water<BOX>(S, jw, cc, cf, pend);
        } else {
          LrfRaw<float> part;
          part.zero();
          lrf_raw_group<BOX>(S, S.iwhich_cgp[jo], c0, c1, c2, part);
          sum32.add(part.terms());
        }
      } else {
        if (S.std_water) lrf_gather_water<P, BOX>(S, jo, jw, c0, c1, c2, cf, acc);
        else lrf_gather_group<P, BOX>(S, S.iwhich_cgp[jo], c0, c1, c2, acc);
      }
    }
    if constexpr (sizeof(P) == 4) {
      // every kLrfPend partner tiles (the same iterations for every lane)
      if (it % kLrfPend == 0) {
        sum32.add(pend.terms());
        pend.zero();
      }
    }
  }
  if constexpr (sizeof(P) == 4) {
    sum32.add(pend.terms());
    acc.add(sum32);
  }
  // solute charge groups (water side of the solute-water list)
  int n_wp = 0;
  auto wrow = gptr_w(S.wp_nb + (size_t)(iw < S.nwat ? iw : 0) * S.wp_cap);
  for (int base = 0; base < S.ncgp_solute; base += kWLLanes) {
    const int g = base + sub;
    bool inside = false, far = false;
    if (active && g < S.ncgp_solute) {
      const int sw = S.cgp_sw[g];
      if (!(gfl[sw] & F_EXCL)) {
        // r2 as the reference forms it: solute switch atom minus water O
        double a = gx[3 * sw] - xi0, b = gx[3 * sw + 1] - xi1, c = gx[3 * sw + 2] - xi2;
        if constexpr (BOX) { a = pbc_image(S, a, 0); b = pbc_image(S, b, 1); c = pbc_image(S, c, 2); }
        const double r2 = a * a + b * b + c * c;
        inside = r2 <= S.rcpw2;
        far = !inside && r2 <= S.rclrf2;
      }
    }
    const unsigned long long m = __ballot(inside) & mine;
    if (inside) {
      const int pos = n_wp + __popcll(m & below);
      if (pos < S.wp_cap) wrow[pos] = g;
    }
    n_wp += __popcll(m);
    if (far && S.use_lrf) lrf_gather_group<P, BOX>(S, g, c0, c1, c2, acc);
  }
  acc.segn_reduce<kWLLanes>();
  if (iw < S.nwat && sub == 0) {
    if (!active) {
      S.ww_cnt[iw] = 0;
      S.wp_cnt[iw] = 0;
    } else {
      if (S.use_lrf) acc.store(Lg);
      S.ww_cnt[iw] = n_in < S.ww_cap ? n_in : S.ww_cap;
      S.wp_cnt[iw] = n_wp < S.wp_cap ? n_wp : S.wp_cap;
      if (n_in > S.ww_cap || n_wp > S.wp_cap) {
        atomicMax(S.status + 2, n_in);
        atomicMax(S.status + 3, n_wp);
        record_failure(S, ST_LIST_OVERFLOW, PH_LISTS);
      }
    }
  }
}
template <typename P, int MAPPED>
__global__ __launch_bounds__(256) void k_water_lists(const DevSys* __restrict__ sys, const int2* __restrict__ map) {
  constexpr int kWLLanes = sizeof(P) == 4 ? kWLLanes32 : kWLLanes64;
  const DevSys* Sp;
  int base;
  if (!block_work<MAPPED>(sys, map, blockDim.x / kWLLanes, Sp, base)) return;
  if (Sp->pbc) water_lists_body<P, true>(*Sp, base);
  else water_lists_body<P, false>(*Sp, base);
}
template __global__ void k_water_lists<double, 0>(const DevSys*, const int2*);
template __global__ void k_water_lists<float, 0>(const DevSys*, const int2*);
template __global__ void k_water_lists<double, 1>(const DevSys*, const int2*);
template __global__ void k_water_lists<float, 1>(const DevSys*, const int2*);

// ---------------------------------------------------------------------------
// Water forces (nonbond_ww_spc src/md.f90:15021-15119, water side of
// nonbond_pw :13282-13365 and nonbond_qw_spc :14411-14570, lrf_taylor
// :5824-5891, restrain_solvent :18055-18102).  Eight waters per wavefront,
// eight lanes per water: the lanes of a water take its partner waters (9 atom
// pairs each, LJ on O-O only), its listed solute groups and the Q atoms in
// turn, lanes 0-2 add the LRF Taylor term of atoms O, H1, H2 and lane 3 the
// radial restraint; the water's 9 gradient components are summed over its 8
// lanes and written once.  P is the pair-arithmetic type: double (plain,
// reference operation order inside each pair, every partner's contribution
// added to the FP64 gradient) or float (mixed: FP32 coordinates and FMA pair
// arithmetic; a lane sums its contributions in FP32 over a tile: all its
// partner waters, one solute group, or its Q atoms, and adds each tile's sum
// to the FP64 gradient, so the FP64 gradient is not held in registers while
// the partner waters are visited).
// W: wavefronts per SIMD the FP32 build is fitted to (6: registers spill to
// scratch; 4: no spills); which is faster is measured by the launch tuner.
// AA: all-atom solvent LJ (nonbond_3atomsolvent :15238-15333 and
// nonbond_qw_3atom :14749-14874): LJ on all nine atom pairs of two waters
// and between each Q atom and every water atom.
// en != 0: the launch also forms the energy partials (e_water); steps whose
// energies are not read launch with en = 0 (the gradient is the same).  The
// two cases are separate instantiations of the body (EN), so a force-only
// step carries no energy arithmetic.
// BOX: periodic box (nonbond_ww_spc_box :15123-15236, nonbond_pw_box
// :13369-13490, nonbond_qw_spc_box :14574-14748): each water pair's shift
// from its O-O vector, each solute group's from the group's switching atom
// and the water O, the Q atoms' from the water O and the Q switching atom;
// no solvent restraints.
template <typename P, int PF, int AA, bool EN, bool BOX>
__device__ __forceinline__ void water_forces_body(const DevSys& S, const int base) {
  constexpr bool F32 = sizeof(P) == 4;
  const int t = threadIdx.x;
  const int sub = t & 7;
  const int iw = base + (t >> 3);
  const bool active = iw < S.nwat;
  const int io = S.nat_solute + 3 * (active ? iw : 0);
  const gconst<P> X = gptr(pair_x<P>(S));
  double g[9];
  for (int k = 0; k < 9; ++k) g[k] = 0.0;
  // FP32 build: the current tile's gradient, joined to g by flush()
  float gf[9];
  for (int k = 0; k < 9; ++k) gf[k] = 0.0f;
  auto flush = [&]() {
    if constexpr (F32) {
      for (int k = 0; k < 9; ++k) {
        g[k] = g[k] + (double)gf[k];
        gf[k] = 0.0f;
      }
    }
  };
  // add c to gradient component k (a constant index after unrolling)
  auto gadd = [&](const int k, const P c) {
    if constexpr (F32) gf[k] = gf[k] + c;
    else g[k] = g[k] + c;
  };
  double e_el = 0.0, e_vdw = 0.0, e_lrf = 0.0, e_rad = 0.0;
  // coordinates of water w: FP64, or the FP32 water record
  auto load_water = [&](const int w, P* o) {
    if constexpr (F32) {
      water_xf(S, w, o);
    } else {
      const int wo = S.nat_solute + 3 * w;
      for (int u = 0; u < 9; ++u) o[u] = X[3 * wo + u];
    }
  };
  P xi[9];
  load_water(active ? iw : 0, xi);
  // charge products formed in P from the P charges: with q_O = -2 q_H exact
  // in either precision, all nine products share one rounding, so the large
  // O-O / O-H / H-H sums still cancel as they do in FP64
  const P qs3[3] = {(P)S.chg_solv[0], (P)S.chg_solv[1], (P)S.chg_solv[2]};
  P qq[9];
  for (int la = 0; la < 3; ++la)
    for (int ip = 0; ip < 3; ++ip) qq[3 * la + ip] = qs3[la] * qs3[ip];
  const P aOO = (P)S.aOO, bOO = (P)S.bOO;
  P aww[9], bww[9];
  if constexpr (AA != 0)
    for (int k = 0; k < 9; ++k) { aww[k] = (P)S.aww[k]; bww[k] = (P)S.bww[k]; }
  // water-water
  const int nn = active && (S.term_mask & TM_WW) ? S.ww_cnt[iw] : 0;
  const gconst<int> row = gptr(S.ww_nb) + (size_t)iw * S.ww_cap;
  auto ww_pair = [&](const int jw, const P* xj) {
    P el = 0, vdw = 0;
    P s0 = 0, s1 = 0, s2 = 0;
    if constexpr (BOX) {
      s0 = pbc_shift(S, xj[0] - xi[0], 0);
      s1 = pbc_shift(S, xj[1] - xi[1], 1);
      s2 = pbc_shift(S, xj[2] - xi[2], 2);
    }
    for (int la = 0; la < 3; ++la) {
      P gl0 = 0, gl1 = 0, gl2 = 0;
      for (int ip = 0; ip < 3; ++ip) {
        P dx1 = xj[3 * ip] - xi[3 * la];
        P dx2 = xj[3 * ip + 1] - xi[3 * la + 1];
        P dx3 = xj[3 * ip + 2] - xi[3 * la + 2];
        if constexpr (BOX) { dx1 = dx1 - s0; dx2 = dx2 - s1; dx3 = dx3 - s2; }
        const P r2sq = pnorm2(dx1, dx2, dx3);
        P r;
        const P r2 = inv_r2(r2sq, r);
        const P Vel = qq[3 * la + ip] * r;
        P dv = r2 * (-Vel);
        if (AA != 0 || (la == 0 && ip == 0)) {
          P r6;
          if constexpr (sizeof(P) == 8) r6 = (P)1 / (r2sq * r2sq * r2sq);
          else r6 = r2 * r2 * r2;
          const P Va = (AA != 0 ? aww[3 * la + ip] : aOO) * r6 * r6;
          const P Vb = (AA != 0 ? bww[3 * la + ip] : bOO) * r6;
          if constexpr (EN) vdw = vdw + Va - Vb;
          dv = dv + r2 * ((P)-12 * Va + (P)6 * Vb);
        }
        if constexpr (EN) el = el + Vel;
        gl0 = pmsub(dv, dx1, gl0);
        gl1 = pmsub(dv, dx2, gl1);
        gl2 = pmsub(dv, dx3, gl2);
      }
      gadd(3 * la, gl0);
      gadd(3 * la + 1, gl1);
      gadd(3 * la + 2, gl2);
    }
    if (EN && owns_pair(iw, jw)) {
      e_el = e_el + (double)el;
      e_vdw = e_vdw + (double)vdw;
    }
  };
  if constexpr (PF != 0) {
    // software pipeline over the lane's partners k, k+8, k+16, ... (the
    // same partners in the same order as without it): two coordinate
    // buffers alternate, so while one partner's pairs are computed the next
    // partner's coordinates and the index of the one after are in flight
    int k = sub;
    int j0 = 0, j1 = 0, jn = 0;
    P x0[9], x1[9];
    if (k < nn) {
      j0 = row[k];
      load_water(j0, x0);
    }
    if (k + 8 < nn) {
      j1 = row[k + 8];
      load_water(j1, x1);
    }
    if (k + 16 < nn) jn = row[k + 16];
    while (k < nn) {
      ww_pair(j0, x0);
      if (k + 16 < nn) {
        j0 = jn;
        load_water(j0, x0);
      }
      if (k + 24 < nn) jn = row[k + 24];
      if (k + 8 >= nn) break;
      ww_pair(j1, x1);
      if (k + 24 < nn) {
        j1 = jn;
        load_water(j1, x1);
      }
      if (k + 32 < nn) jn = row[k + 32];
      k += 16;
    }
  } else {
    for (int k = sub; k < nn; k += 8) {
      const int jw = row[k];
      P xj[9];
      load_water(jw, xj);
      ww_pair(jw, xj);
    }
  }
  flush();
  // water-solute: every non-Q atom of each listed solute group (the solute
  // atom is i of the reference pair)
  const int nwp = active && (S.term_mask & TM_PW) ? S.wp_cnt[iw] : 0;
  const gconst<int> wrow = gptr(S.wp_nb) + (size_t)iw * S.wp_cap;
  if (F32 && S.srec) {
    // FP32 records in slot order: the group's atoms are consecutive records,
    // so no load waits on another (the same pairs and arithmetic as below)
    for (int k = sub; k < nwp; k += 8) {
      const int gs = wrow[k];
      const int a0 = S.cgp_first[gs], a1 = S.cgp_last[gs];
      const gconst<f32v4> rec = gptr(reinterpret_cast<const f32v4*>(S.srec));
      P s0 = 0, s1 = 0, s2 = 0;
      if constexpr (BOX) {
        const int sw = S.cgp_sw[gs];
        s0 = pbc_shift(S, xi[0] - X[3 * sw], 0);
        s1 = pbc_shift(S, xi[1] - X[3 * sw + 1], 1);
        s2 = pbc_shift(S, xi[2] - X[3 * sw + 2], 2);
      }
      for (int a = a0; a <= a1; ++a) {
        const f32v4 r0 = rec[3 * (size_t)a], r1 = rec[3 * (size_t)a + 1], r2v = rec[3 * (size_t)a + 2];
        if (r2v.z != 0.0f) continue;   // a Q atom
        const P ra[3] = {(P)r1.x, (P)r1.y, (P)r1.z}, rb[3] = {(P)r1.w, (P)r2v.x, (P)r2v.y};
        for (int ip = 0; ip < 3; ++ip) {
          P dx1 = xi[3 * ip] - (P)r0.x, dx2 = xi[3 * ip + 1] - (P)r0.y, dx3 = xi[3 * ip + 2] - (P)r0.z;
          if constexpr (BOX) { dx1 = dx1 - s0; dx2 = dx2 - s1; dx3 = dx3 - s2; }
          P r;
          const P r2 = inv_r2(pnorm2(dx1, dx2, dx3), r);
          const P r6 = r2 * r2 * r2;
          const P r12 = r6 * r6;
          const P Vel = (P)r0.w * (P)S.chg_solv[ip] * r;
          const P Va = ra[ip] * r12;
          const P Vb = rb[ip] * r6;
          const P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb);
          gadd(3 * ip, dv * dx1);
          gadd(3 * ip + 1, dv * dx2);
          gadd(3 * ip + 2, dv * dx3);
        }
      }
      flush();
    }
  } else {
  for (int k = sub; k < nwp; k += 8) {
    const int gs = wrow[k];
    P s0 = 0, s1 = 0, s2 = 0;
    if constexpr (BOX) {
      const int sw = S.cgp_sw[gs];
      s0 = pbc_shift(S, xi[0] - X[3 * sw], 0);
      s1 = pbc_shift(S, xi[1] - X[3 * sw + 1], 1);
      s2 = pbc_shift(S, xi[2] - X[3 * sw + 2], 2);
    }
    for (int a = S.cgp_first[gs]; a <= S.cgp_last[gs]; ++a) {
      const int i = S.cgpatom[a];
      if (S.flags[i] & F_Q) continue;
      const P xs0 = X[3 * i], xs1 = X[3 * i + 1], xs2 = X[3 * i + 2];
      const int ti = S.iac[i];
      const P qs = (P)S.crg[i];
      const P* wlj = S.std_water ? solute_water_lj<P>(S) + 6 * (size_t)i : nullptr;
      for (int ip = 0; ip < 3; ++ip) {
        const int j = io + ip;
        P dx1 = xi[3 * ip] - xs0, dx2 = xi[3 * ip + 1] - xs1, dx3 = xi[3 * ip + 2] - xs2;
        if constexpr (BOX) { dx1 = dx1 - s0; dx2 = dx2 - s1; dx3 = dx3 - s2; }
        P aLJ, bLJ;
        if (wlj) {
          aLJ = wlj[ip];
          bLJ = wlj[3 + ip];
        } else {
          const int tj = S.iac[j];
          const int lj = S.ljcod[ti * S.natyps + tj] - 1;
          aLJ = (P)(S.avdw[ti * 3 + lj] * S.avdw[tj * 3 + lj]);
          bLJ = (P)(S.bvdw[ti * 3 + lj] * S.bvdw[tj * 3 + lj]);
        }
        P r;
        const P r2 = inv_r2(pnorm2(dx1, dx2, dx3), r);
        const P r6 = r2 * r2 * r2;
        const P r12 = r6 * r6;
        const P Vel = qs * (P)S.crg[j] * r;
        const P Va = aLJ * r12;
        const P Vb = bLJ * r6;
        const P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb);
        gadd(3 * ip, dv * dx1);
        gadd(3 * ip + 1, dv * dx2);
        gadd(3 * ip + 2, dv * dx3);
      }
    }
    flush();
  }
  }
  // Q-water pairs seen from the water, Q atoms taken in turn by the lanes
  // (the Q atom is i of each pair): O with LJ + Coulomb, H Coulomb only (LJ
  // too for an all-atom solvent).  Only the gradient is formed here, so with
  // fixed Q LJ parameters the states fold into one lambda-weighted charge
  // (sum_s lambda_s qcrg(iq,s)) and one lambda sum for the LJ part; with
  // Q-atom types (soft-core per state) each state is evaluated.
  if (active && (S.flags[io] & F_QW) && (S.term_mask & TM_QX)) {
    const int ns = S.nstates;
    const bool qv = S.qvdw != 0;
    P q0 = 0, q1 = 0, q2s = 0;
    if constexpr (BOX) {
      const int qs = S.qswitch;
      q0 = pbc_shift(S, xi[0] - X[3 * qs], 0);
      q1 = pbc_shift(S, xi[1] - X[3 * qs + 1], 1);
      q2s = pbc_shift(S, xi[2] - X[3 * qs + 2], 2);
    }
    for (int iq = sub; iq < S.nqat; iq += 8) {
      const int qi = S.iqseq[iq];
      const P xq0 = X[3 * qi], xq1 = X[3 * qi + 1], xq2 = X[3 * qi + 2];
      const gconst<P> R = gptr(qw_rec<P>(S)) + (size_t)iq * kQwRec;
      const P qe = R[0], lsum = R[1];
      for (int ip = 0; ip < 3; ++ip) {
        P ex1 = xi[3 * ip] - xq0, ex2 = xi[3 * ip + 1] - xq1, ex3 = xi[3 * ip + 2] - xq2;
        if constexpr (BOX) { ex1 = ex1 - q0; ex2 = ex2 - q1; ex3 = ex3 - q2s; }
        const P h2 = pnorm2(ex1, ex2, ex3);
        P rr;
        const P q2 = inv_r2(h2, rr);
        const bool lj = ip == 0 || AA != 0;
        P dv = 0;
        if (qv) {
          const int lh = S.qw_lj[3 * iq + ip];
          const P h6_hc = h2 * h2 * h2;
          for (int st = 0; st < ns; ++st) {
            P Vel, Va, Vb;
            if (lj) {
              dv = dv + qw_qvdw_state<P>(S, iq, st, ip, lh, q2, rr, h6_hc, Vel, Va, Vb);
            } else {
              Vel = (P)S.chg_solv[ip] * (P)S.qcrg[iq * ns + st] * rr;
              dv = dv - q2 * Vel * (P)S.lambda[st];
            }
          }
        } else {
          const P Vel = (P)S.chg_solv[ip] * qe * rr;
          if (lj) {
            P h6;
            if constexpr (sizeof(P) == 8) h6 = (P)1 / (h2 * h2 * h2);
            else h6 = q2 * q2 * q2;
            const P Va = R[2 + 2 * ip] * h6 * h6;
            const P Vb = R[3 + 2 * ip] * h6;
            dv = q2 * (-Vel - ((P)12 * Va - (P)6 * Vb) * lsum);
          } else {
            dv = -q2 * Vel;
          }
        }
        gadd(3 * ip, dv * ex1);
        gadd(3 * ip + 1, dv * ex2);
        gadd(3 * ip + 2, dv * ex3);
      }
    }
    flush();
  }
  // LRF Taylor expansion (FP64), atom `sub` of the water on lanes 0..2
  if (active && S.use_lrf && sub < 3 && (S.term_mask & TM_LRF)) {
    const int i = io + sub;
    if ((S.flags[i] & (F_EXCL | F_Q)) == 0) {
      const double* L = S.lrf + (size_t)S.iwhich_cgp[io] * kLrfStride;
      const double xa[3] = {S.x[3 * i], S.x[3 * i + 1], S.x[3 * i + 2]};
      double df[3];
      const double V = lrf_taylor_atom(L, xa, df);
      if constexpr (EN) e_lrf = 0.5 * S.crg[i] * V;
      for (int c = 0; c < 3; ++c) {
        const double f = S.crg[i] * df[c];
        if (sub == 0) g[c] = g[c] - f;
        else if (sub == 1) g[3 + c] = g[3 + c] - f;
        else g[6 + c] = g[6 + c] - f;
      }
    }
  }
  // radial solvent restraint on the water O (switching atom), lane 3 (FP64)
  if (!BOX && active && sub == 3 && !(S.flags[io] & F_EXCL) && (S.term_mask & TM_OTHER)) {
    const double shift = S.fk_wsphere != 0.0 ? sqrt(qsys::kBoltz * S.scal[2] / S.fk_wsphere) : 0.0;
    const double dr0 = S.x[3 * io] - S.xwcent[0], dr1 = S.x[3 * io + 1] - S.xwcent[1],
                 dr2 = S.x[3 * io + 2] - S.xwcent[2];
    const double b = sqrt(dr0 * dr0 + dr1 * dr1 + dr2 * dr2);
    const double db = b - (S.rwat - shift);
    double erst, dv;
    if (db > 0) {
      erst = 0.5 * S.fk_wsphere * db * db - S.dwmz;
      dv = S.fk_wsphere * db / b;
    } else if (b > 0.0) {
      const double fexp = exp(S.awmz * db);
      erst = S.dwmz * (fexp * fexp - 2.0 * fexp);
      dv = -2.0 * S.dwmz * S.awmz * (fexp - fexp * fexp) / b;
    } else {
      erst = 0.0; dv = 0.0;
    }
    e_rad = erst;
    g[0] = g[0] + dv * dr0;
    g[1] = g[1] + dv * dr1;
    g[2] = g[2] + dv * dr2;
  }
  // excluded water atoms: fix_shell's restraint to the topology position
  // (src/md.f90:2816-2862), atom sub-4 on lanes 4..6 (FP64)
  double e_fix = 0.0;
  if (active && sub >= 4 && sub < 7 && (S.flags[io + sub - 4] & F_EXCL) && (S.term_mask & TM_OTHER)) {
    const int i = io + sub - 4;
    const double fk = S.freeze ? 0.0 : S.fk_fix;
    const double dr0 = S.x[3 * i] - S.xtop[3 * i], dr1 = S.x[3 * i + 1] - S.xtop[3 * i + 1],
                 dr2 = S.x[3 * i + 2] - S.xtop[3 * i + 2];
    e_fix = 0.5 * fk * (dr0 * dr0 + dr1 * dr1 + dr2 * dr2);
    // constant indices only: a lane-dependent index would move g to scratch
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (sub - 4 == a) {
        g[3 * a] = g[3 * a] + fk * dr0;
        g[3 * a + 1] = g[3 * a + 1] + fk * dr1;
        g[3 * a + 2] = g[3 * a + 2] + fk * dr2;
      }
  }
  for (int k = 0; k < 9; ++k) g[k] = seg8_sum(g[k]);
  if constexpr (EN) {
    e_el = seg8_sum(e_el);
    e_vdw = seg8_sum(e_vdw);
    e_lrf = seg8_sum(e_lrf);
    e_rad = seg8_sum(e_rad);
    e_fix = seg8_sum(e_fix);
  }
  if (active && sub == 0) {
    for (int k = 0; k < 9; ++k) S.d[3 * io + k] = g[k];
    if constexpr (EN) {
      double* ew = S.e_water + (size_t)iw * kEWater;
      ew[0] = e_el; ew[1] = e_vdw; ew[2] = e_lrf; ew[3] = e_rad; ew[4] = e_fix;
    }
  }
}

template <typename P, int MAPPED, int PF, int W, int AA>
__global__ __launch_bounds__(256, sizeof(P) == 4 ? W : 1) void k_water_forces(const DevSys* __restrict__ sys,
                                                                              const int2* __restrict__ map,
                                                                              int en) {
  const DevSys* Sp;
  int base;
  if (!block_work<MAPPED>(sys, map, blockDim.x >> 3, Sp, base)) return;
  const DevSys& S = *Sp;
  if (base + ((threadIdx.x >> 6) << 3) >= S.nwat) return;   // whole wavefronts past the last water
  if (S.pbc) {
    if (en) water_forces_body<P, PF, AA, true, true>(S, base);
    else water_forces_body<P, PF, AA, false, true>(S, base);
  } else {
    if (en) water_forces_body<P, PF, AA, true, false>(S, base);
    else water_forces_body<P, PF, AA, false, false>(S, base);
  }
}
template __global__ void k_water_forces<double, 0, 0, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 0, 0, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 0, 1, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 0, 1, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 1, 0, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 1, 0, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 1, 1, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<double, 1, 1, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 0, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 0, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 0, 4, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 0, 4, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 1, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 1, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 1, 4, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 0, 1, 4, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 0, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 0, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 0, 4, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 0, 4, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 1, 6, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 1, 6, 1>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 1, 4, 0>(const DevSys*, const int2*, int);
template __global__ void k_water_forces<float, 1, 1, 4, 1>(const DevSys*, const int2*, int);

// ---------------------------------------------------------------------------
// Water polarisation restraint (watpol, src/md.f90:18372-18541), part 1:
// the angle between each water dipole and its radial vector, and its shell.
__global__ void k_watpol_theta(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= S.nwat || !S.wpol_restr) return;
  int i = S.nat_solute + 3 * iw;
  S.theta[iw] = 0.0;
  S.shell_of[iw] = -1;
  if (S.flags[i] & F_EXCL) return;
  double scp, rc, rm;
  double rmu[3], rcu[3];
  watpol_geometry(S, i, rmu, rcu, rm, rc, scp);
  S.theta[iw] = acos(scp);
  const DevShell& last = S.shell[S.nshell - 1];
  if (rc > last.rout - last.dr) {
    int is;
    for (is = S.nshell - 1; is >= 1; --is)
      if (rc <= S.shell[is].rout) break;
    if (is < 1) is = 0;
    S.shell_of[iw] = is;
    // the parts of the member's gradient that do not depend on its rank,
    // kept in its gradient record until the shell kernel completes it:
    // f1a + f2 on O (3), f1b on each H (3), -1 / sin(theta)
    double* g = S.wpol_g + 9 * (size_t)iw;
    for (int k = 0; k < 3; ++k) {
      double f1a = -2.0 * (rcu[k] - rmu[k] * scp) / rm;
      double f1b = (rcu[k] - rmu[k] * scp) / rm;
      double f2 = (rmu[k] - rcu[k] * scp) / rc;
      g[k] = f1a + f2;
      g[3 + k] = f1b;
    }
    double f0 = sin(acos(scp));
    if (fabs(f0) < qsys::kQrealEps) f0 = qsys::kQrealEps;
    g[6] = -1.0 / f0;
  }
}

// Part 2: kWpolSplit workgroups per shell.  Each compacts the shell's
// members in water order; a member's position in qdyn's selection sort
// (ascending theta, ties to the lower list index) is its rank, counted by
// four lanes over interleaved quarters of the members; the target angle
// follows from the rank.  Workgroup sp takes the sp-th chunk of the
// members.  The shell record is read only: its sums (energy, theta, member
// count) go to per-workgroup partials that k_watpol_shell folds into it in
// a fixed order after this kernel.
constexpr int kShellCap = 4096;
__global__ __launch_bounds__(kWpolBlock) void k_watpol_force(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int is = blockIdx.x / kWpolSplit, sp = blockIdx.x % kWpolSplit;
  if (!S.wpol_restr || is >= S.nshell || halted(S, PH_WATPOL)) return;
  const int istep = *S.step_ctr;
  const bool update = istep != 0 && istep % qsys::kItdisUpdate == 0;
  __shared__ int mem_iw[kShellCap];
  __shared__ double mem_th[kShellCap];
  __shared__ int wave_tot[kWpolBlock / 64];
  __shared__ double red[2][kWpolBlock / 64];
  // ordered compaction of this shell's members: each thread takes a
  // contiguous run of waters (all its loads issued together), counts its
  // members, and writes them after the members of the threads before it
  // (block-wide exclusive scan of the counts)
  const int per = (S.nwat + kWpolBlock - 1) / kWpolBlock;
  const int w0 = threadIdx.x * per, w1 = min(S.nwat, w0 + per);
  unsigned long long own = 0ull;   // membership of the run's first 64 waters
  int cnt = 0;
  for (int c0 = w0; c0 < w1; c0 += 8) {
    int sh[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) sh[u] = c0 + u < w1 ? S.shell_of[c0 + u] : -1;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (sh[u] == is) {
        ++cnt;
        if (c0 + u - w0 < 64) own |= 1ull << (c0 + u - w0);
      }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int incl = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) wave_tot[wv] = incl;
  __syncthreads();
  int off = incl - cnt, n = 0;
  for (int w = 0; w < kWpolBlock / 64; ++w) {
    if (w < wv) off += wave_tot[w];
    n += wave_tot[w];
  }
  for (int iw = w0, pos = off; iw < w1 && pos < off + cnt; ++iw) {
    const bool mem = iw - w0 < 64 ? ((own >> (iw - w0)) & 1ull) != 0 : S.shell_of[iw] == is;
    if (mem) {
      if (pos < kShellCap) mem_iw[pos] = iw;
      ++pos;
    }
  }
  __syncthreads();
  if (n > kShellCap) {
    if (threadIdx.x == 0 && sp == 0) record_failure(S, ST_SHELL_OVERFLOW, PH_WATPOL);
    n = kShellCap;
  }
  for (int m = threadIdx.x; m < n; m += kWpolBlock) mem_th[m] = S.theta[mem_iw[m]];
  __syncthreads();
  // theta_corr as this step uses it: updated every itdis_update steps
  // (k_watpol_shell stores the same update into the record)
  const DevShell& W = S.shell[is];
  const double cstb = W.cstb;
  double tcorr = W.theta_corr;
  if (update) tcorr = tcorr + W.avtheta / (double)qsys::kItdisUpdate - acos(cstb);
  const int chunk = (n + kWpolSplit - 1) / kWpolSplit;
  const int m0 = sp * chunk, m1 = min(n, m0 + chunk);
  const int q = threadIdx.x & 3;
  double e = 0.0, tsum = 0.0;
  for (int mb = m0; mb < m1; mb += kWpolBlock / 4) {
    const int m = mb + (int)(threadIdx.x >> 2);
    const bool have = m < m1;
    const double th = have ? mem_th[m] : 0.0;
    int rank = 0;
    if (have) {
#pragma unroll 8
      for (int k = q; k < n; k += 4) {
        const double tk = mem_th[k];
        rank += (tk < th || (tk == th && k < m)) ? 1 : 0;
      }
    }
    rank += __shfl_xor(rank, 1);
    rank += __shfl_xor(rank, 2);
    if (!have || q != 0) continue;
    const int iw = mem_iw[m];
    double arg = 1.0 + (1.0 - 2.0 * (double)(rank + 1)) / (double)n;
    double th0 = acos(arg);
    th0 = th0 - 3.0 * sin(th0) * cstb / 2.0;
    if (th0 < 0.0) th0 = 0.0;
    if (th0 > kPiD) th0 = kPiD;
    tsum = tsum + th;
    double dth = th - th0 + tcorr;
    e = e + 0.5 * S.fkwpol * dth * dth;
    double dv = S.fkwpol * (th - th0 + tcorr);
    // the water's gradient contribution (added to d by the integrator, or
    // by k_watpol_gather outside a step), completed from the rank-free
    // parts k_watpol_theta left in the record
    double* g = S.wpol_g + 9 * (size_t)iw;
    const double f0 = dv * g[6];
    const bool on = (S.term_mask & TM_OTHER) != 0;
    double ga[3], gb[3];
    for (int k = 0; k < 3; ++k) { ga[k] = g[k]; gb[k] = g[3 + k]; }
    for (int k = 0; k < 3; ++k) {
      g[k] = on ? f0 * ga[k] : 0.0;
      g[3 + k] = on ? f0 * gb[k] : 0.0;
      g[6 + k] = on ? f0 * gb[k] : 0.0;
    }
  }
  e = block_sum<kWpolBlock / 64>(e, red[0]);
  tsum = block_sum<kWpolBlock / 64>(tsum, red[1]);
  if (threadIdx.x == 0) {
    double* part = S.wpol_part + 3 * ((size_t)is * kWpolSplit + sp);
    part[0] = e;
    part[1] = tsum;
    part[2] = (double)n;
  }
}

// The shell records after k_watpol_force (one lane per shell): the
// itdis_update step's theta_corr update, then this step's mean theta and
// member count, and the restraint energy, from the workgroups' partials
// summed in workgroup order.
__global__ void k_watpol_shell(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int is = threadIdx.x;
  if (!S.wpol_restr || is >= S.nshell || halted(S, PH_WATPOL)) return;
  const int istep = *S.step_ctr;
  const bool update = istep != 0 && istep % qsys::kItdisUpdate == 0;
  DevShell& W = S.shell[is];
  const double* part = S.wpol_part + 3 * (size_t)is * kWpolSplit;
  double e = 0.0, tsum = 0.0;
  for (int sp = 0; sp < kWpolSplit; ++sp) {
    e = e + part[3 * sp];
    tsum = tsum + part[3 * sp + 1];
  }
  const int n = (int)part[2];
  if (update) {
    W.avtheta = W.avtheta / (double)qsys::kItdisUpdate;
    W.avn_insh = W.avn_insh / (double)qsys::kItdisUpdate;
    W.theta_corr = W.theta_corr + W.avtheta - acos(W.cstb);
    W.avtheta = 0.0;
    W.avn_insh = 0.0;
  }
  W.n_insh = n;
  if (n > 0) {
    W.avtheta = W.avtheta + tsum / (double)n;
    W.avn_insh = W.avn_insh + n;
  }
  S.e_wpol[is] = e;
}

// The polarisation gradient of shell members added to d (force evaluations
// outside a step; a step's integrator adds it itself).
__global__ void k_watpol_gather(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int iw = blockIdx.x * blockDim.x + threadIdx.x;
  if (iw >= S.nwat || !S.wpol_restr || S.shell_of[iw] < 0) return;
  double* d = S.d + 3 * (S.nat_solute + 3 * iw);
  const double* g = S.wpol_g + 9 * (size_t)iw;
  for (int c = 0; c < 9; ++c) d[c] = d[c] + g[c];
}

// ---------------------------------------------------------------------------
// Leapfrog/Berendsen update (newvel_ber src/md.f90:6194-6223), SHAKE
// (shake :17581-17679), v = (x - xx)/dt, and the kinetic-energy partial
// sums of temperature (:6098-6190), fused per SHAKE unit.  Units are SHAKE
// molecules (waters: 3 atoms, 3 constraints) or single atoms.
// The gradient of atom a as the step uses it: d plus the polarisation
// restraint and the sum of the atom's bonded-term slots, formed in
// registers (the integrator stores it back into d after the update, so no
// load waits behind a store).
__device__ __forceinline__ void step_gradient(const DevSys& S, int a, double* g) {
  g[0] = S.d[3 * a]; g[1] = S.d[3 * a + 1]; g[2] = S.d[3 * a + 2];
  // the polarisation restraint first (it was added right after the water
  // forces before it ran beside them)
  if (S.wpol_restr && a >= S.nat_solute) {
    const int rel = a - S.nat_solute, iw = rel / 3;
    if (S.shell_of[iw] >= 0) {
      const double* w = S.wpol_g + 9 * (size_t)iw + 3 * (rel % 3);
      g[0] = g[0] + w[0];
      g[1] = g[1] + w[1];
      g[2] = g[2] + w[2];
    }
  }
  // the atom's bonded/Q-Q/restraint term slots, summed in slot order
  const int t = S.slot_of_atom[a];
  if (t < 0 || !(S.term_mask & TM_OTHER)) return;
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
  g[0] = g[0] + g0;
  g[1] = g[1] + g1;
  g[2] = g[2] + g2;
}

// step_gradient for the three atoms of water unit atoms a..a+2, with every
// independent load of the water issued first (gradient, shell, slots) and
// the dependent ones (polarisation gradient, slot sums) after them; the
// same sums in the same order as step_gradient.
__device__ __forceinline__ void water_step_gradient(const DevSys& S, int a, double* gd) {
  for (int c = 0; c < 9; ++c) gd[c] = S.d[3 * a + c];
  const int iw = (a - S.nat_solute) / 3;
  const int sh = (S.wpol_restr && a >= S.nat_solute) ? S.shell_of[iw] : -1;
  int so[3];
  for (int k = 0; k < 3; ++k) so[k] = S.slot_of_atom[a + k];
  if (sh >= 0) {
    const double* w = S.wpol_g + 9 * (size_t)iw;
    double wv[9];
    for (int c = 0; c < 9; ++c) wv[c] = w[c];
    for (int c = 0; c < 9; ++c) gd[c] = gd[c] + wv[c];
  }
  if (!(S.term_mask & TM_OTHER)) return;
  for (int k = 0; k < 3; ++k) {
    const int t = so[k];
    if (t < 0) continue;
    double g0 = 0, g1 = 0, g2 = 0;
    const gconst<double> o = gptr(S.slot_g) + 3 * (size_t)S.slot_first[t];
    const int n = S.slot_first[t + 1] - S.slot_first[t];
    for (int j = 0; j < n; ++j) { g0 = g0 + o[3 * j]; g1 = g1 + o[3 * j + 1]; g2 = g2 + o[3 * j + 2]; }
    gd[3 * k] = gd[3 * k] + g0;
    gd[3 * k + 1] = gd[3 * k + 1] + g1;
    gd[3 * k + 2] = gd[3 * k + 2] + g2;
  }
}

// shake (src/md.f90:17581-17679) of one molecule, constraints c0..c1 of
// the instance's SHAKE tables (at most 64), on coordinates x with the
// pre-step coordinates xx: returns the molecule's iteration count, or -1
// when it reaches the iteration limit with a constraint still unmet.
__device__ __forceinline__ int shake_molecule(const DevSys& S, double* __restrict__ x,
                                              const double* __restrict__ xx, int c0, int c1) {
  unsigned long long ready = 0ull;
  const int nc = c1 - c0;
  const unsigned long long all = nc >= 64 ? ~0ull : ((1ull << nc) - 1ull);
  int nits = 0;
  for (;;) {
    for (int ic = 0; ic < nc; ++ic) {
      if (ready & (1ull << ic)) continue;
      const int i = S.sh_i[c0 + ic], j = S.sh_j[c0 + ic];
      const double d2 = S.sh_d2[c0 + ic];
      double* xi = x + 3 * i;
      double* xj = x + 3 * j;
      const double* xxi = xx + 3 * i;
      const double* xxj = xx + 3 * j;
      double a0_ = xi[0] - xj[0], a1_ = xi[1] - xj[1], a2_ = xi[2] - xj[2];
      double xij2 = a0_ * a0_ + a1_ * a1_ + a2_ * a2_;
      double diff = d2 - xij2;
      if (fabs(diff) < qsys::kShakeTol * d2) ready |= 1ull << ic;
      double b0 = xxi[0] - xxj[0], b1 = xxi[1] - xxj[1], b2 = xxi[2] - xxj[2];
      double scp = a0_ * b0 + a1_ * b1 + a2_ * b2;
      double wi = S.winv[i], wj = S.winv[j];
      double corr = diff / (2.0 * scp * (wi + wj));
      xi[0] = xi[0] + b0 * corr * wi;
      xi[1] = xi[1] + b1 * corr * wi;
      xi[2] = xi[2] + b2 * corr * wi;
      xj[0] = xj[0] - b0 * corr * wj;
      xj[1] = xj[1] - b1 * corr * wj;
      xj[2] = xj[2] - b2 * corr * wj;
    }
    ++nits;
    if ((ready & all) == all) return nits;
    if (nits >= S.shake_maxiter) return -1;
  }
}

// shake(xx, x) on arrays of one instance outside the step (the boundary
// call that mirrors the reference's function): one thread per SHAKE unit;
// out[0] accumulates the molecules' iteration counts, out[1] is set where a
// molecule fails.
__global__ __launch_bounds__(256) void k_shake(const DevSys* __restrict__ sys, const double* __restrict__ xx,
                                               double* __restrict__ x, int* __restrict__ out) {
  const DevSys& S = *sys;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= S.nunits) return;
  const int c0 = S.con_first[u], c1 = S.con_first[u + 1];
  if (c1 <= c0) return;
  const int n = shake_molecule(S, x, xx, c0, c1);
  if (n < 0) atomicOr(out + 1, 1);
  else atomicAdd(out, n);
}

// randm (src/md.f90:17543-17577) as the affine map state' = (31415821 state
// + 1) mod 1e8, and its real value: state/10*10 / 1e8, 0 outside (0, 1].
__device__ __forceinline__ double randm_next(long long& st) {
  st = (st * 31415821LL + 1) % 100000000LL;
  double r = (double)(st / 10) * 10 / 100000000.0;
  if (r <= 0.0 || r > 1.0) r = 0.0;
  return r;
}

// Langevin random forces of atom i (newvel_lan, src/md.f90:6232-6248): its
// draws start after those of atoms 0..i-1 of this step; uniform
// rands (u - 1/2) or Gaussian (gauss: twelve draws, (sum - 6) rands).
__device__ __forceinline__ void langevin_forces(const DevSys& S, int i, double* re) {
  long long st = ((long long)S.rng_jump[2 * i] * (long long)S.rng[0] + (long long)S.rng_jump[2 * i + 1]) %
                 100000000LL;
  const double sd = S.rands[i];
  for (int c = 0; c < 3; ++c) {
    if (!S.lgauss) {
      re[c] = sd * (randm_next(st) - 0.5);
    } else {
      double a = 0.0;
      for (int k = 0; k < 12; ++k) a = a + randm_next(st);
      re[c] = (a - 6.0) * sd + 0.0;
    }
  }
}

// The thermostat's velocity update of one component (newvel_ber,
// newvel_lan, newvel_nos: src/md.f90:6194-6301) with gradient g, inverse
// mass w, Berendsen scale ts and Langevin random force re.
__device__ __forceinline__ double vel_update(const DevSys& S, double v, double g, double w, double ts, double re) {
  if (S.thermostat == 1) return (v * S.dv_friction) - (g - re) * w * S.dv_mod;
  if (S.thermostat == 2) return (v - g * w * S.dv_mod);
  return (v - g * w * S.dv_mod) * ts;
}

// nh_prop (src/md.f90:6303-6344) for the Nose-Hoover instances, before
// k_integrate: the chain's half-step from the last temperature and the
// velocity scale exp(-vnh_1 dt/2); with SHAKE the second propagation md_run
// makes after SHAKE (from the same temperature) too.  Neither reads the
// velocities, so both scales are formed here: nh[3n] before the velocity
// update, nh[3n+1] after SHAKE.  One thread per instance.
__device__ double nh_half(const DevSys& S) {
  const int n = S.nhn;
  double* xnh = S.nh;
  double* vnh = S.nh + n;
  double* gnh = S.nh + 2 * n;
  const double* q = S.qnh;
  const double dt2 = 0.5 * S.dt, dt4 = 0.5 * dt2, dt8 = 0.5 * dt4, T0 = S.temp0, temp = S.scal[3];
  gnh[0] = S.ndegf * qsys::kBoltz * (temp - T0) / q[0];
  for (int i = 1; i < n; ++i) gnh[i] = (q[i - 1] * (vnh[i - 1] * vnh[i - 1]) - S.kbt) / q[i];
  vnh[n - 1] = vnh[n - 1] + gnh[n - 1] * dt4;
  xnh[n - 1] = xnh[n - 1] + vnh[n - 1] * dt2;
  double expf = 0.0;
  for (int i = n - 2; i >= 0; --i) {
    expf = exp(-(dt8 * vnh[i + 1]));
    vnh[i] = (vnh[i] * expf + gnh[i] * dt4) * expf;
    xnh[i] = xnh[i] + vnh[i] * dt2;
  }
  const double sc = exp(-(vnh[0] * dt2));
  gnh[0] = S.ndegf * qsys::kBoltz * (sc * sc * temp - T0) / q[0];
  vnh[0] = (vnh[0] * expf + gnh[0] * dt4) * expf;
  for (int i = 1; i < n - 1; ++i) {
    expf = exp(-(dt8 * vnh[i + 1]));
    gnh[i] = (q[i - 1] * (vnh[i - 1] * vnh[i - 1]) - S.kbt) / q[i];
    vnh[i] = (vnh[i] * expf + gnh[i] * dt4) * expf;
  }
  gnh[n - 1] = (q[n - 2] * (vnh[n - 2] * vnh[n - 2]) - S.kbt) / q[n - 1];
  vnh[n - 1] = vnh[n - 1] + gnh[n - 1] * dt4;
  return sc;
}
__global__ void k_nh_prop(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  if (threadIdx.x != 0 || S.thermostat != 2 || halted(S, PH_INTEGRATE)) return;
  const int n = S.nhn;
  S.nh[3 * n] = nh_half(S);
  S.nh[3 * n + 1] = S.has_shake ? nh_half(S) : 1.0;
}

__global__ __launch_bounds__(kIntegrateBlock) void k_integrate(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  if (halted(S, PH_INTEGRATE)) return;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  double ke[6] = {0, 0, 0, 0, 0, 0};  // solute tot, solute free, solvent tot, solvent free, spare
  if (u < S.nunits && S.unit_kind[u] == 1) {
    // water: the same arithmetic as the generic path below, with the 9
    // coordinates, the 9 pre-step coordinates and the velocities in registers
    const double dt = S.dt;
    const int a = S.unit_head[u];
    const int c0 = S.con_first[u];
    const double d2[3] = {S.sh_d2[c0], S.sh_d2[c0 + 1], S.sh_d2[c0 + 2]};
    const double ts = a < S.nat_solute ? S.scal[0] : S.scal[1];
    // Nose-Hoover: the velocities scaled by the chain first
    const double s1 = S.thermostat == 2 ? S.nh[3 * S.nhn] : 1.0;
    double x[9], xo[9], v[9], w[3], gd[9], re[9];
    for (int k = 0; k < 3; ++k) w[k] = S.winv[a + k];
    for (int c = 0; c < 9; ++c) { v[c] = S.v[3 * a + c]; xo[c] = S.x[3 * a + c]; }
    water_step_gradient(S, a, gd);
    for (int c = 0; c < 9; ++c) re[c] = 0.0;
    if (S.thermostat == 1)
      for (int k = 0; k < 3; ++k) langevin_forces(S, a + k, re + 3 * k);
    for (int c = 0; c < 9; ++c) {
      const int ic = 3 * a + c;
      // excluded_freeze: an excluded atom restarts from rest at its
      // topology position (fix_shell, src/md.f90:2831-2839)
      const bool frz = S.freeze && (S.flags[a + c / 3] & F_EXCL);
      double v0 = frz ? 0.0 : v[c];
      if (S.thermostat == 2) v0 = v0 * s1;
      double vv = vel_update(S, v0, gd[c], w[c / 3], ts, re[c]);
      if (frz) xo[c] = S.xtop[ic];
      x[c] = xo[c] + vv * dt;
      if (S.vverlet) vv = vel_update(S, vv, gd[c], w[c / 3], ts, re[c]);
      v[c] = vv;
    }
    const int ci[3] = {0, 0, 1}, cj[3] = {1, 2, 2};
    unsigned ready = 0u;
    int nits = 0;
    for (;;) {
      for (int ic = 0; ic < 3; ++ic) {
        if (ready & (1u << ic)) continue;
        const int i = ci[ic], j = cj[ic];
        double a0 = x[3 * i] - x[3 * j], a1 = x[3 * i + 1] - x[3 * j + 1], a2 = x[3 * i + 2] - x[3 * j + 2];
        double xij2 = a0 * a0 + a1 * a1 + a2 * a2;
        double diff = d2[ic] - xij2;
        if (fabs(diff) < qsys::kShakeTol * d2[ic]) ready |= 1u << ic;
        double b0 = xo[3 * i] - xo[3 * j], b1 = xo[3 * i + 1] - xo[3 * j + 1], b2 = xo[3 * i + 2] - xo[3 * j + 2];
        double scp = a0 * b0 + a1 * b1 + a2 * b2;
        double corr = diff / (2.0 * scp * (w[i] + w[j]));
        x[3 * i] = x[3 * i] + b0 * corr * w[i];
        x[3 * i + 1] = x[3 * i + 1] + b1 * corr * w[i];
        x[3 * i + 2] = x[3 * i + 2] + b2 * corr * w[i];
        x[3 * j] = x[3 * j] - b0 * corr * w[j];
        x[3 * j + 1] = x[3 * j + 1] - b1 * corr * w[j];
        x[3 * j + 2] = x[3 * j + 2] - b2 * corr * w[j];
      }
      ++nits;
      if (ready == 7u) break;
      if (nits >= S.shake_maxiter) { record_failure(S, ST_SHAKE, PH_INTEGRATE); break; }
    }
    const double s2 = S.thermostat == 2 ? S.nh[3 * S.nhn + 1] : 1.0;
    for (int c = 0; c < 9; ++c) {
      const int ic = 3 * a + c;
      if (S.has_shake) {
        v[c] = (x[c] - xo[c]) / dt;
        if (S.thermostat == 2) v[c] = v[c] * s2;
      }
      S.x[ic] = x[c];
      S.xf[ic] = (float)x[c];
      S.v[ic] = v[c];
      S.d[ic] = gd[c];
    }
    // solvent waters start at nat_solute, three atoms each
    if (S.xw && a >= S.nat_solute) water_xf_store(S, (a - S.nat_solute) / 3, x);
    for (int k = 0; k < 3; ++k) {
      const int i = a + k;
      const double ek = 0.5 * S.mass[i] * (v[3 * k] * v[3 * k] + v[3 * k + 1] * v[3 * k + 1] + v[3 * k + 2] * v[3 * k + 2]);
      const bool ex = (S.flags[i] & F_EXCL) != 0;
      if (i < S.nat_solute) { ke[0] += ek; if (!ex) ke[1] += ek; }
      else { ke[2] += ek; if (!ex) ke[3] += ek; }
    }
  } else if (u < S.nunits) {
    const double dt = S.dt;
    const double ts_solute = S.scal[0], ts_solvent = S.scal[1];
    const double s1 = S.thermostat == 2 ? S.nh[3 * S.nhn] : 1.0;
    const int a0 = S.unit_first[u], a1 = S.unit_first[u + 1];
    for (int k = a0; k < a1; ++k) {
      const int i = S.unit_atoms[k];
      double gd[3], re[3] = {0.0, 0.0, 0.0};
      step_gradient(S, i, gd);
      if (S.thermostat == 1) langevin_forces(S, i, re);
      const double ts = i < S.nat_solute ? ts_solute : ts_solvent;
      const double w = S.winv[i];
      const bool frz = S.freeze && (S.flags[i] & F_EXCL);
      for (int c = 0; c < 3; ++c) {
        const int ic = 3 * i + c;
        double v0 = frz ? 0.0 : S.v[ic];
        if (S.thermostat == 2) v0 = v0 * s1;
        double vv = vel_update(S, v0, gd[c], w, ts, re[c]);
        const double xo = frz ? S.xtop[ic] : S.x[ic];
        S.xx[ic] = xo;
        S.x[ic] = xo + vv * dt;
        if (S.vverlet) vv = vel_update(S, vv, gd[c], w, ts, re[c]);
        S.v[ic] = vv;
        S.d[ic] = gd[c];
      }
    }
    const int c0 = S.con_first[u], c1 = S.con_first[u + 1];
    if (c1 > c0 && shake_molecule(S, S.x, S.xx, c0, c1) < 0) record_failure(S, ST_SHAKE, PH_INTEGRATE);
    const double s2 = S.thermostat == 2 ? S.nh[3 * S.nhn + 1] : 1.0;
    for (int k = a0; k < a1; ++k) {
      const int i = S.unit_atoms[k];
      double vv[3];
      for (int c = 0; c < 3; ++c) {
        const int ic = 3 * i + c;
        if (S.has_shake) {
          S.v[ic] = (S.x[ic] - S.xx[ic]) / dt;
          if (S.thermostat == 2) S.v[ic] = S.v[ic] * s2;
        }
        vv[c] = S.v[ic];
        S.xf[ic] = (float)S.x[ic];
        if (S.srec && i < S.nat_solute) S.srec[(size_t)kSrec * S.srec_of_atom[i] + c] = (float)S.x[ic];
        if (S.xw && i >= S.nat_solute) {
          const int rel = i - S.nat_solute;
          S.xw[16 * (rel / 3) + 3 * (rel % 3) + c] = (float)S.x[ic];
        }
      }
      const double ek = 0.5 * S.mass[i] * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
      const bool ex = (S.flags[i] & F_EXCL) != 0;
      if (i < S.nat_solute) { ke[0] += ek; if (!ex) ke[1] += ek; }
      else { ke[2] += ek; if (!ex) ke[3] += ek; }
    }
  }
  __shared__ double red[6][kIntegrateBlock / 64];
  for (int c = 0; c < 4; ++c) {
    double s = block_sum<kIntegrateBlock / 64>(ke[c], red[c]);
    if (threadIdx.x == 0) S.ke_part[(size_t)blockIdx.x * 6 + c] = s;
  }
}

// temperature(): Ekin totals, T, Tfree and the Berendsen scale factors for
// the next step (src/md.f90:6098-6190).  One workgroup per instance.
__global__ __launch_bounds__(256) void k_temperature(const DevSys* __restrict__ sys, int nblocks) {
  const DevSys& S = sys[blockIdx.y];
  // a failure earlier in this step: the step is not completed, and every
  // later launch returns at once (key 0)
  if (halted(S, PH_TEMPERATURE)) {
    if (threadIdx.x == 0) atomicMin(S.fail, 0ull);
    return;
  }
  // the Q bonds' Morse factors of this step's evaluation, kept for the next
  for (int t = threadIdx.x; t < S.nqbond; t += blockDim.x)
    for (int c = 0; c < kMorseRec; ++c) S.qb_prev[kMorseRec * t + c] = S.qb_cur[kMorseRec * t + c];
  __shared__ double red[4][4];
  double p[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int c = 0; c < 4; ++c) p[c] += S.ke_part[(size_t)b * 6 + c];
  for (int c = 0; c < 4; ++c) p[c] = block_sum_256(p[c], red[c]);
  if (threadIdx.x != 0) return;
  double temp_solute = p[0], tfree_solute = p[1], temp_solvent = p[2], tfree_solvent = p[3];
  double tfree = tfree_solvent + tfree_solute;
  double temp = temp_solute + temp_solvent;
  S.energies[E_KINETIC] = temp;
  temp = 2.0 * temp / qsys::kBoltz / (double)S.ndegf;
  tfree = 2.0 * tfree / qsys::kBoltz / (double)S.ndegfree;
  if (S.detail_temps) {
    tfree_solute = 2.0 * tfree_solute / qsys::kBoltz / (double)S.ndegfree_solute;
    tfree_solvent = 2.0 * tfree_solvent / qsys::kBoltz / (double)S.ndegfree_solvent;
  }
  double tsu = S.scal[0], tsv = S.scal[1];
  const double r = S.dt / S.tau_t;
  if (S.separate_scaling) {
    if (tfree_solvent != 0.0) tsv = S.temp0 / tfree_solvent - 1.0;
    tsv = sqrt(1.0 + r * tsv);
    if (tfree_solute != 0.0) tsu = S.temp0 / tfree_solute - 1.0;
    tsu = sqrt(1.0 + r * tsu);
  } else {
    if (tfree != 0.0) tsv = S.temp0 / tfree - 1.0;
    tsv = sqrt(1.0 + r * tsv);
    tsu = tsv;
  }
  S.scal[0] = tsu; S.scal[1] = tsv; S.scal[2] = tfree; S.scal[3] = temp;
  // Langevin: randm's state after this step's draws
  if (S.thermostat == 1)
    S.rng[0] = (int)((S.rng_step[0] * (long long)S.rng[0] + S.rng_step[1]) % 100000000LL);
  // a NaN temperature stops the run after this step (md_run's test,
  // src/md.f90:6654-6656, made at every step)
  if (temp != temp) record_failure(S, ST_NAN, PH_TEMPERATURE);
  // end of the step: advance this instance's step counter
  *S.step_ctr = *S.step_ctr + 1;
}

// Sum the per-item energy partials into the instance's energy terms.
// With record != 0 it runs only on energy-file steps (step counter a
// positive multiple of iene) and also appends the frame's put_ene records.
__global__ __launch_bounds__(256) void k_energy_reduce(const DevSys* __restrict__ sys, int record) {
  const DevSys& S = sys[blockIdx.y];
  __shared__ double red[12][4];
  int frame = -1;
  if (record) {
    if (halted(S, PH_RECORD)) return;
    const int st = *S.step_ctr - S.eseries_base;
    if (S.iene <= 0 || st <= 0 || st % S.iene != 0) return;
    frame = st / S.iene - 1;
  }
  double w[kEWater] = {0, 0, 0, 0, 0};
  for (int iw = threadIdx.x; iw < S.nwat; iw += blockDim.x)
    for (int c = 0; c < kEWater; ++c) w[c] += S.e_water[(size_t)iw * kEWater + c];
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < S.nat_solute; i += blockDim.x)
    for (int c = 0; c < 8; ++c) p[c] += S.e_solute[(size_t)i * 8 + c];
  // bonded terms: p_bond p_angle p_torsion w_bond w_angle w_torsion
  double ef = 0.0;   // learned far field (far_mode)
  if (S.far_mode)
    for (int g = threadIdx.x; g < S.ncgp; g += blockDim.x) ef += S.e_far[g];
  double b[6] = {0, 0, 0, 0, 0, 0};
  for (int n = threadIdx.x; n < S.nbnd; n += blockDim.x) b[n < S.nbnd_solute ? 0 : 3] += S.e_bterm[n];
  for (int n = threadIdx.x; n < S.nang; n += blockDim.x) b[n < S.nang_solute ? 1 : 4] += S.e_bterm[S.nbnd + n];
  for (int n = threadIdx.x; n < S.ntor; n += blockDim.x)
    b[n < S.ntor_solute ? 2 : 5] += S.e_bterm[S.nbnd + S.nang + n];
  double bi[2] = {0, 0};   // impropers: solute, solvent
  for (int n = threadIdx.x; n < S.nimp; n += blockDim.x)
    bi[n < S.nimp_solute ? 0 : 1] += S.e_bterm[S.nbnd + S.nang + S.ntor + n];
  for (int c = 0; c < 4; ++c) w[c] = block_sum_256(w[c], red[c]);
  for (int c = 0; c < 8; ++c) p[c] = block_sum_256(p[c], red[4 + c]);
  w[4] = block_sum_256(w[4], red[0]);
  for (int c = 0; c < 6; ++c) b[c] = block_sum_256(b[c], red[c]);
  for (int c = 0; c < 2; ++c) bi[c] = block_sum_256(bi[c], red[6 + c]);
  ef = block_sum_256(ef, red[8]);
  // extra terms: restraints outside the FEP states (sequence, wall, and the
  // state-independent ones of a run without states, src/md.f90:15612-15618)
  double rp = 0.0;
  for (int n = threadIdx.x; n < S.nxt; n += blockDim.x) {
    const int kind = S.xt[kXtInts * n], tag = S.xt[kXtInts * n + 5];
    if (kind == XT_RST_SEQ || kind == XT_RST_SEQC || kind == XT_RST_WALL ||
        (kind >= XT_RST_POS && kind <= XT_RST_ANG && tag == 0 && S.nstates == 0))
      rp += S.e_xt[n];
  }
  rp = block_sum_256(rp, red[9]);
  // per-state Q energies: qq, qp (Q-neighbour pairs of nonbond_qq and the
  // Q-solute list), qw; pass 1 forms the FP64 re-evaluated frame of a
  // mixed-precision batch from e_q64 (Q-Q pairs are FP64 in both)
  for (int pass = 0; pass < 2; ++pass)
  for (int st = 0; st < S.nstates; ++st) {
    if (pass == 1 && !(frame >= 0 && S.eseries64)) break;
    const double* eq = pass == 0 ? S.e_q : S.e_q64;
    // qq qp qw (el, vdw), Q bond angle torsion improper, restraint
    double q[kQTerms] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = threadIdx.x; n < S.nxt; n += blockDim.x) {
      const int kind = S.xt[kXtInts * n], tag = S.xt[kXtInts * n + 5];
      if (kind <= XT_QIMPROPER) {
        if (tag == st) q[5 + kind] += S.e_xt[n];
      } else if (kind >= XT_RST_POS && kind <= XT_RST_ANG && (tag == 0 || tag == st + 1)) {
        q[10] += S.e_xt[n];
      }
    }
    for (int iq = threadIdx.x; iq < S.nqat; iq += blockDim.x) {
      const double* e = eq + ((size_t)iq * S.nstates + st) * 4;
      q[2] += e[0]; q[3] += e[1]; q[4] += e[2]; q[5] += e[3];
    }
    for (int pr = threadIdx.x; pr < S.nqq; pr += blockDim.x) {
      if (S.qq_state[pr] != st) continue;
      const int o = S.qq_jq[pr] >= 0 ? 0 : 2;
      q[o] += S.e_qq[2 * pr];
      q[o + 1] += S.e_qq[2 * pr + 1];
    }
    for (int c = 0; c < kQTerms; ++c) {
      q[c] = block_sum_256(q[c], red[c]);
      if (threadIdx.x == 0 && pass == 0) S.qenergies[st * kQTerms + c] = q[c];
    }
    if (threadIdx.x == 0 && frame >= 0 && frame < S.eseries_cap) {
      // put_ene record: total, q(bond angle torsion improper), qx, qq, qp,
      // qw (el, vdw each), restraint (src/nrgy.f90:128-130; EQ totals as
      // pot_energy forms them, src/md.f90:16186-16194)
      double* r = (pass == 0 ? S.eseries : S.eseries64) + ((size_t)frame * S.nstates + st) * kEneRecord;
      const double qx_el = q[0] + q[2] + q[4], qx_vdw = q[1] + q[3] + q[5];
      r[0] = q[6] + q[7] + q[8] + q[9] + qx_el + qx_vdw + q[10];
      r[1] = q[6]; r[2] = q[7]; r[3] = q[8]; r[4] = q[9];
      r[5] = qx_el; r[6] = qx_vdw;
      r[7] = q[0]; r[8] = q[1]; r[9] = q[2]; r[10] = q[3]; r[11] = q[4]; r[12] = q[5];
      r[13] = q[10];
    }
  }
  // EVB off-diagonal elements of the frame (offdiag, src/md.f90:15444-15470)
  if (threadIdx.x == 0 && frame >= 0 && frame < S.eseries_cap && S.eseries_offd) {
    for (int io = 0; io < S.noffd; ++io) {
      const int k3 = 3 * S.offd_kl[2 * io], l3 = 3 * S.offd_kl[2 * io + 1];
      const double a = S.x[l3] - S.x[k3], b2 = S.x[l3 + 1] - S.x[k3 + 1], c = S.x[l3 + 2] - S.x[k3 + 2];
      const double r = sqrt(a * a + b2 * b2 + c * c);
      S.eseries_offd[((size_t)frame * S.noffd + io) * 2] = S.offd_amu[2 * io] * exp(-S.offd_amu[2 * io + 1] * r);
      S.eseries_offd[((size_t)frame * S.noffd + io) * 2 + 1] = r;
    }
  }
  if (threadIdx.x == 0 && frame >= 0) *S.eseries_n = frame + 1;
  if (threadIdx.x == 0) {
    S.energies[E_P_BOND] = b[0]; S.energies[E_P_ANGLE] = b[1]; S.energies[E_P_TORSION] = b[2];
    S.energies[E_W_BOND] = b[3]; S.energies[E_W_ANGLE] = b[4]; S.energies[E_W_TORSION] = b[5];
    S.energies[E_P_IMPROPER] = bi[0];
    S.energies[E_W_IMPROPER] = bi[1];
    double* E = S.energies;
    E[E_WW_EL] = w[0]; E[E_WW_VDW] = w[1]; E[E_LRF] = w[2] + p[4] + ef; E[E_RST_RADIAL] = w[3];
    E[E_PP_EL] = p[0]; E[E_PP_VDW] = p[1]; E[E_PW_EL] = p[2]; E[E_PW_VDW] = p[3];
    E[E_RST_FIX] = p[5] + w[4]; E[E_RST_SHELL] = p[6]; E[E_RST_PROTEIN] = rp;
    double ew = 0.0;
    for (int s = 0; s < S.nshell; ++s) ew = ew + S.e_wpol[s];
    E[E_RST_WPOL] = S.wpol_restr ? ew : 0.0;
  }
}

}  // namespace qgpu
