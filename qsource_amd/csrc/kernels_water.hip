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
// cgp_centers (src/md.f90:2253-2288): centre of each charge group; LRF
// multipoles zeroed (they are overwritten by the list kernels for every
// group that is not excluded).
__global__ void k_cgp_centers(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ncgp) return;
  double c0 = 0, c1 = 0, c2 = 0;
  int f = S.cgp_first[g], l = S.cgp_last[g];
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

// ---------------------------------------------------------------------------
// Water lists (nbwwlist_lrf src/md.f90:10897-11021 and the water side of
// nbpwlist_lrf :9300-9429).  One wavefront per water: lanes scan the other
// waters and the solute charge groups; partners inside the cut-off are
// appended in ascending order by a ballot compaction, partners between the
// cut-off and RcLRF contribute their non-Q atoms' multipoles to this water's
// charge-group centre (gathered in registers, shuffle-reduced once).
__global__ __launch_bounds__(256) void k_water_lists(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int iw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (iw >= S.nwat) return;
  const int io = S.nat_solute + 3 * iw;
  if (S.flags[io] & F_EXCL) {
    if (lane == 0) { S.ww_cnt[iw] = 0; S.wp_cnt[iw] = 0; }
    return;
  }
  const double xi0 = S.x[3 * io], xi1 = S.x[3 * io + 1], xi2 = S.x[3 * io + 2];
  const int ig = S.iwhich_cgp[io];
  double* Lg = S.lrf + (size_t)ig * kLrfStride;
  const double c0 = Lg[0], c1 = Lg[1], c2 = Lg[2];
  LrfAcc acc;
  acc.zero();
  // waters
  int n_in = 0;
  int* row = S.ww_nb + (size_t)iw * S.ww_cap;
  for (int base = 0; base < S.nwat; base += 64) {
    int jw = base + lane;
    bool inside = false, far = false;
    if (jw < S.nwat && jw != iw) {
      int jo = S.nat_solute + 3 * jw;
      if (!(S.flags[jo] & F_EXCL)) {
        double a = xi0 - S.x[3 * jo], b = xi1 - S.x[3 * jo + 1], c = xi2 - S.x[3 * jo + 2];
        double r2 = a * a + b * b + c * c;
        inside = r2 <= S.rcww2;
        far = !inside && r2 <= S.rclrf2;
      }
    }
    unsigned long long m = __ballot(inside);
    if (inside) {
      int pos = n_in + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < S.ww_cap) row[pos] = jw;
    }
    n_in += __popcll(m);
    if (far && S.use_lrf) {
      int jg = S.iwhich_cgp[S.nat_solute + 3 * jw];
      lrf_gather_group(S, jg, c0, c1, c2, acc);
    }
  }
  // solute charge groups (water side of the solute-water list)
  int n_wp = 0;
  int* wrow = S.wp_nb + (size_t)iw * S.wp_cap;
  for (int base = 0; base < S.ncgp_solute; base += 64) {
    int g = base + lane;
    bool inside = false, far = false;
    if (g < S.ncgp_solute) {
      int sw = S.cgp_sw[g];
      if (!(S.flags[sw] & F_EXCL)) {
        // r2 as the reference forms it: solute switch atom minus water O
        double a = S.x[3 * sw] - xi0, b = S.x[3 * sw + 1] - xi1, c = S.x[3 * sw + 2] - xi2;
        double r2 = a * a + b * b + c * c;
        inside = r2 <= S.rcpw2;
        far = !inside && r2 <= S.rclrf2;
      }
    }
    unsigned long long m = __ballot(inside);
    if (inside) {
      int pos = n_wp + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < S.wp_cap) wrow[pos] = g;
    }
    n_wp += __popcll(m);
    if (far && S.use_lrf) lrf_gather_group(S, g, c0, c1, c2, acc);
  }
  acc.wave_reduce();
  if (lane == 0) {
    if (S.use_lrf) acc.store(Lg);
    S.ww_cnt[iw] = n_in < S.ww_cap ? n_in : S.ww_cap;
    S.wp_cnt[iw] = n_wp < S.wp_cap ? n_wp : S.wp_cap;
    if (n_in > S.ww_cap || n_wp > S.wp_cap) atomicOr(S.overflow, 1);
  }
}

// ---------------------------------------------------------------------------
// Water forces (nonbond_ww_spc src/md.f90:15021-15119, water side of
// nonbond_pw :13282-13365 and nonbond_qw_spc :14411-14570, lrf_taylor
// :5824-5891, restrain_solvent :18055-18102).  One wavefront per water:
// lanes take the partner waters (9 atom pairs each, LJ on O-O only), the
// listed solute groups, the Q atoms, and the per-atom LRF and restraint
// terms; the water's 9 gradient components are shuffle-reduced and written
// once.  P is the pair-arithmetic type: double (plain) or float (mixed, the
// 9 contributions of one partner are summed in FP32 and accumulated in FP64).
template <typename P>
__global__ __launch_bounds__(256) void k_water_forces(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int iw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (iw >= S.nwat) return;
  const int io = S.nat_solute + 3 * iw;
  double g[9];
  for (int k = 0; k < 9; ++k) g[k] = 0.0;
  double e_el = 0.0, e_vdw = 0.0, e_lrf = 0.0, e_rad = 0.0;
  double xi[9];
  for (int k = 0; k < 9; ++k) xi[k] = S.x[3 * io + k];
  const P q[3] = {(P)S.chg_solv[0], (P)S.chg_solv[1], (P)S.chg_solv[2]};
  const P aOO = (P)S.aOO, bOO = (P)S.bOO;
  // water-water
  const int nn = S.ww_cnt[iw];
  const int* row = S.ww_nb + (size_t)iw * S.ww_cap;
  for (int k = lane; k < nn; k += 64) {
    const int jw = row[k];
    const int jo = S.nat_solute + 3 * jw;
    const bool own = owns_pair(iw, jw);
    double xj[9];
    for (int t = 0; t < 9; ++t) xj[t] = S.x[3 * jo + t];
    P el = 0, vdw = 0;
    for (int la = 0; la < 3; ++la) {
      P gl0 = 0, gl1 = 0, gl2 = 0;
      for (int ip = 0; ip < 3; ++ip) {
        P dx1 = (P)(xj[3 * ip] - xi[3 * la]);
        P dx2 = (P)(xj[3 * ip + 1] - xi[3 * la + 1]);
        P dx3 = (P)(xj[3 * ip + 2] - xi[3 * la + 2]);
        P r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
        P r6 = r2 * r2 * r2;
        P r;
        r2 = inv_r2(r2, r);
        P Vel = q[la] * q[ip] * r;
        P dv = r2 * (-Vel);
        if (la == 0 && ip == 0) {
          r6 = (P)1 / r6;
          P Va = aOO * r6 * r6;
          P Vb = bOO * r6;
          vdw = vdw + Va - Vb;
          dv = dv + r2 * ((P)-12 * Va + (P)6 * Vb);
        }
        el = el + Vel;
        gl0 = gl0 - dv * dx1;
        gl1 = gl1 - dv * dx2;
        gl2 = gl2 - dv * dx3;
      }
      g[3 * la] = g[3 * la] + (double)gl0;
      g[3 * la + 1] = g[3 * la + 1] + (double)gl1;
      g[3 * la + 2] = g[3 * la + 2] + (double)gl2;
    }
    if (own) {
      e_el = e_el + (double)el;
      e_vdw = e_vdw + (double)vdw;
    }
  }
  // water-solute: every non-Q atom of each listed solute group
  const int nwp = S.wp_cnt[iw];
  const int* wrow = S.wp_nb + (size_t)iw * S.wp_cap;
  for (int k = lane; k < nwp; k += 64) {
    const int gs = wrow[k];
    for (int a = S.cgp_first[gs]; a <= S.cgp_last[gs]; ++a) {
      const int i = S.cgpatom[a];
      if (S.flags[i] & F_Q) continue;
      const double xs0 = S.x[3 * i], xs1 = S.x[3 * i + 1], xs2 = S.x[3 * i + 2];
      const int ti = S.iac[i];
      for (int ip = 0; ip < 3; ++ip) {
        const int j = io + ip;
        const int tj = S.iac[j];
        const int lj = S.ljcod[ti * S.natyps + tj] - 1;
        // reference orientation: i = solute atom, j = water atom
        P dx1 = (P)(xi[3 * ip] - xs0), dx2 = (P)(xi[3 * ip + 1] - xs1), dx3 = (P)(xi[3 * ip + 2] - xs2);
        const P aLJ = (P)(S.avdw[ti * 3 + lj] * S.avdw[tj * 3 + lj]);
        const P bLJ = (P)(S.bvdw[ti * 3 + lj] * S.bvdw[tj * 3 + lj]);
        P r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
        P r;
        r2 = inv_r2(r2, r);
        P r6 = r2 * r2 * r2;
        P r12 = r6 * r6;
        P Vel = (P)S.crg[i] * (P)S.crg[j] * r;
        P Va = aLJ * r12;
        P Vb = bLJ * r6;
        P dv = r2 * (-Vel - (P)12 * Va + (P)6 * Vb);
        g[3 * ip] = g[3 * ip] + (double)(dv * dx1);
        g[3 * ip + 1] = g[3 * ip + 1] + (double)(dv * dx2);
        g[3 * ip + 2] = g[3 * ip + 2] + (double)(dv * dx3);
      }
    }
  }
  // Q-water pairs seen from the water, one Q atom per lane (the Q atom is i
  // of each pair): O with LJ + Coulomb, H Coulomb only
  if (S.flags[io] & F_QW) {
    const int ns = S.nstates;
    for (int iq = lane; iq < S.nqat; iq += 64) {
      const int qi = S.iqseq[iq];
      const double* xq = S.x + 3 * qi;
      const int lo = S.qw_lj[iq];
      const int ti = S.iac[qi];
      P dx1 = (P)(xi[0] - xq[0]), dx2 = (P)(xi[1] - xq[1]), dx3 = (P)(xi[2] - xq[2]);
      P r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
      P r6_hc = r2 * r2 * r2;
      P r;
      r2 = inv_r2(r2, r);
      P r6 = (P)1 / r6_hc;
      P Va = (P)S.avdw[ti * 3 + lo - 1] * (P)S.alj_O[lo - 1] * r6 * r6;
      P Vb = (P)S.bvdw[ti * 3 + lo - 1] * (P)S.blj_O[lo - 1] * r6;
      P dv = 0;
      for (int st = 0; st < ns; ++st) {
        P Vel = q[0] * (P)S.qcrg[iq * ns + st] * r;
        dv = dv + r2 * (-Vel - ((P)12 * Va - (P)6 * Vb) * (r6_hc / r6_hc)) * (P)S.lambda[st];
      }
      g[0] = g[0] + (double)(dv * dx1);
      g[1] = g[1] + (double)(dv * dx2);
      g[2] = g[2] + (double)(dv * dx3);
      for (int ip = 1; ip < 3; ++ip) {
        P ex1 = (P)(xi[3 * ip] - xq[0]), ex2 = (P)(xi[3 * ip + 1] - xq[1]), ex3 = (P)(xi[3 * ip + 2] - xq[2]);
        P rr;
        P q2 = inv_r2(ex1 * ex1 + ex2 * ex2 + ex3 * ex3, rr);
        P dvh = 0;
        for (int st = 0; st < ns; ++st) {
          P Vel = q[ip] * (P)S.qcrg[iq * ns + st] * rr;
          dvh = dvh - q2 * Vel * (P)S.lambda[st];
        }
        g[3 * ip] = g[3 * ip] + (double)(dvh * ex1);
        g[3 * ip + 1] = g[3 * ip + 1] + (double)(dvh * ex2);
        g[3 * ip + 2] = g[3 * ip + 2] + (double)(dvh * ex3);
      }
    }
  }
  // LRF Taylor expansion, one water atom per lane 0..2
  if (S.use_lrf && lane < 3) {
    const int i = io + lane;
    if ((S.flags[i] & (F_EXCL | F_Q)) == 0) {
      const double* L = S.lrf + (size_t)S.iwhich_cgp[io] * kLrfStride;
      double df[3];
      const double xa[3] = {xi[3 * lane], xi[3 * lane + 1], xi[3 * lane + 2]};
      double V = lrf_taylor_atom(L, xa, df);
      e_lrf = 0.5 * S.crg[i] * V;
      for (int c = 0; c < 3; ++c) {
        const double f = S.crg[i] * df[c];
        if (lane == 0) g[c] = g[c] - f;
        else if (lane == 1) g[3 + c] = g[3 + c] - f;
        else g[6 + c] = g[6 + c] - f;
      }
    }
  }
  // radial solvent restraint on the water O (switching atom), lane 3
  if (lane == 3 && !(S.flags[io] & F_EXCL)) {
    double shift = S.fk_wsphere != 0.0 ? sqrt(qsys::kBoltz * S.scal[2] / S.fk_wsphere) : 0.0;
    double dr0 = xi[0] - S.xwcent[0], dr1 = xi[1] - S.xwcent[1], dr2 = xi[2] - S.xwcent[2];
    double b = sqrt(dr0 * dr0 + dr1 * dr1 + dr2 * dr2);
    double db = b - (S.rwat - shift);
    double erst, dv;
    if (db > 0) {
      erst = 0.5 * S.fk_wsphere * db * db - S.dwmz;
      dv = S.fk_wsphere * db / b;
    } else if (b > 0.0) {
      double fexp = exp(S.awmz * db);
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
  for (int k = 0; k < 9; ++k) g[k] = wave_sum(g[k]);
  e_el = wave_sum(e_el);
  e_vdw = wave_sum(e_vdw);
  e_lrf = wave_sum(e_lrf);
  e_rad = wave_sum(e_rad);
  if (lane == 0) {
    for (int k = 0; k < 9; ++k) S.d[3 * io + k] = g[k];
    double* ew = S.e_water + (size_t)iw * 4;
    ew[0] = e_el; ew[1] = e_vdw; ew[2] = e_lrf; ew[3] = e_rad;
  }
}
template __global__ void k_water_forces<double>(const DevSys*);
template __global__ void k_water_forces<float>(const DevSys*);

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
  }
}

// Part 2: one workgroup per shell.  Members are compacted in water order;
// each member's position in qdyn's selection sort (ascending theta, ties to
// the lower list index) is its rank; the target angle follows from the rank.
constexpr int kShellCap = 4096;
__global__ __launch_bounds__(1024) void k_watpol_force(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int is = blockIdx.x;
  if (!S.wpol_restr || is >= S.nshell) return;
  const int istep = *S.step_ctr;
  const bool update = istep != 0 && istep % qsys::kItdisUpdate == 0;
  __shared__ int mem_iw[kShellCap];
  __shared__ double mem_th[kShellCap];
  __shared__ int s_cnt;
  __shared__ double red[2][16];
  DevShell& W = S.shell[is];
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  // ordered compaction of this shell's members
  for (int base = 0; base < S.nwat; base += blockDim.x) {
    int iw = base + threadIdx.x;
    bool mine = iw < S.nwat && S.shell_of[iw] == is;
    // block-wide ordered prefix: wave ballots then a serial pass over waves
    unsigned long long m = __ballot(mine);
    __shared__ int wave_cnt[16];
    int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) wave_cnt[wv] = __popcll(m);
    __syncthreads();
    int off = s_cnt;
    for (int w = 0; w < wv; ++w) off += wave_cnt[w];
    if (mine) {
      int pos = off + __popcll(m & ((1ull << ln) - 1ull));
      if (pos < kShellCap) { mem_iw[pos] = iw; mem_th[pos] = S.theta[iw]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) tot += wave_cnt[w];
      s_cnt += tot;
    }
    __syncthreads();
  }
  int n = s_cnt;
  if (n > kShellCap) { if (threadIdx.x == 0) atomicOr(S.overflow, 2); n = kShellCap; }
  // theta_corr update every itdis_update steps (before the forces use it)
  if (update && threadIdx.x == 0) {
    W.avtheta = W.avtheta / (double)qsys::kItdisUpdate;
    W.avn_insh = W.avn_insh / (double)qsys::kItdisUpdate;
    W.theta_corr = W.theta_corr + W.avtheta - acos(W.cstb);
    W.avtheta = 0.0;
    W.avn_insh = 0.0;
  }
  __syncthreads();
  const double tcorr = W.theta_corr, cstb = W.cstb;
  double e = 0.0, tsum = 0.0;
  for (int m = threadIdx.x; m < n; m += blockDim.x) {
    const double th = mem_th[m];
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      double tk = mem_th[k];
      rank += (tk < th || (tk == th && k < m)) ? 1 : 0;
    }
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
    const int i = S.nat_solute + 3 * iw;
    double scp, rc, rm, rmu[3], rcu[3];
    watpol_geometry(S, i, rmu, rcu, rm, rc, scp);
    double f0 = sin(acos(scp));
    if (fabs(f0) < qsys::kQrealEps) f0 = qsys::kQrealEps;
    f0 = -1.0 / f0;
    f0 = dv * f0;
    double* d = S.d + 3 * i;
    for (int k = 0; k < 3; ++k) {
      double f1a = -2.0 * (rcu[k] - rmu[k] * scp) / rm;
      double f1b = (rcu[k] - rmu[k] * scp) / rm;
      double f2 = (rmu[k] - rcu[k] * scp) / rc;
      d[k] = d[k] + f0 * (f1a + f2);
      d[3 + k] = d[3 + k] + f0 * (f1b);
      d[6 + k] = d[6 + k] + f0 * (f1b);
    }
  }
  e = block_sum_1024(e, red[0]);
  tsum = block_sum_1024(tsum, red[1]);
  if (threadIdx.x == 0) {
    W.n_insh = n;
    if (n > 0) {
      W.avtheta = W.avtheta + tsum / (double)n;
      W.avn_insh = W.avn_insh + n;
    }
    S.e_wpol[is] = e;
  }
}

// ---------------------------------------------------------------------------
// Leapfrog/Berendsen update (newvel_ber src/md.f90:6194-6223), SHAKE
// (shake :17581-17679), v = (x - xx)/dt, and the kinetic-energy partial
// sums of temperature (:6098-6190), fused per SHAKE unit.  Units are SHAKE
// molecules (waters: 3 atoms, 3 constraints) or single atoms.
__global__ __launch_bounds__(256) void k_integrate(const DevSys* __restrict__ sys) {
  const DevSys& S = sys[blockIdx.y];
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  double ke[6] = {0, 0, 0, 0, 0, 0};  // solute tot, solute free, solvent tot, solvent free, spare
  if (u < S.nunits && S.unit_kind[u] == 1) {
    // water: the same arithmetic as the generic path below, with the 9
    // coordinates, the 9 pre-step coordinates and the velocities in registers
    const double dt = S.dt;
    const int a = S.unit_atoms[S.unit_first[u]];
    const double ts = a < S.nat_solute ? S.scal[0] : S.scal[1];
    double x[9], xo[9], v[9], w[3];
    for (int k = 0; k < 3; ++k) w[k] = S.winv[a + k];
    for (int c = 0; c < 9; ++c) {
      const int ic = 3 * a + c;
      double vv = (S.v[ic] - S.d[ic] * w[c / 3] * dt) * ts;
      v[c] = vv;
      xo[c] = S.x[ic];
      x[c] = xo[c] + vv * dt;
    }
    const int c0 = S.con_first[u];
    const double d2[3] = {S.sh_d2[c0], S.sh_d2[c0 + 1], S.sh_d2[c0 + 2]};
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
      if (nits >= qsys::kShakeMaxIter) { atomicOr(S.shake_fail, 1); break; }
    }
    for (int c = 0; c < 9; ++c) {
      const int ic = 3 * a + c;
      if (S.has_shake) v[c] = (x[c] - xo[c]) / dt;
      S.x[ic] = x[c];
      S.v[ic] = v[c];
    }
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
    const int a0 = S.unit_first[u], a1 = S.unit_first[u + 1];
    for (int k = a0; k < a1; ++k) {
      const int i = S.unit_atoms[k];
      const double ts = i < S.nat_solute ? ts_solute : ts_solvent;
      const double w = S.winv[i];
      for (int c = 0; c < 3; ++c) {
        const int ic = 3 * i + c;
        double vv = (S.v[ic] - S.d[ic] * w * dt) * ts;
        S.v[ic] = vv;
        S.xx[ic] = S.x[ic];
        S.x[ic] = S.x[ic] + vv * dt;
      }
    }
    const int c0 = S.con_first[u], c1 = S.con_first[u + 1];
    if (c1 > c0) {
      // constraints of one molecule: at most 64 handled in registers/bits
      unsigned long long ready = 0ull;
      const int nc = c1 - c0;
      const unsigned long long all = nc >= 64 ? ~0ull : ((1ull << nc) - 1ull);
      int nits = 0;
      for (;;) {
        for (int ic = 0; ic < nc; ++ic) {
          if (ready & (1ull << ic)) continue;
          const int i = S.sh_i[c0 + ic], j = S.sh_j[c0 + ic];
          const double d2 = S.sh_d2[c0 + ic];
          double* xi = S.x + 3 * i;
          double* xj = S.x + 3 * j;
          const double* xxi = S.xx + 3 * i;
          const double* xxj = S.xx + 3 * j;
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
        if ((ready & all) == all) break;
        if (nits >= qsys::kShakeMaxIter) { atomicOr(S.shake_fail, 1); break; }
      }
    }
    for (int k = a0; k < a1; ++k) {
      const int i = S.unit_atoms[k];
      double vv[3];
      for (int c = 0; c < 3; ++c) {
        const int ic = 3 * i + c;
        if (S.has_shake) S.v[ic] = (S.x[ic] - S.xx[ic]) / dt;
        vv[c] = S.v[ic];
      }
      const double ek = 0.5 * S.mass[i] * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
      const bool ex = (S.flags[i] & F_EXCL) != 0;
      if (i < S.nat_solute) { ke[0] += ek; if (!ex) ke[1] += ek; }
      else { ke[2] += ek; if (!ex) ke[3] += ek; }
    }
  }
  __shared__ double red[6][4];
  for (int c = 0; c < 4; ++c) {
    double s = block_sum_256(ke[c], red[c]);
    if (threadIdx.x == 0) S.ke_part[(size_t)blockIdx.x * 6 + c] = s;
  }
}

// temperature(): Ekin totals, T, Tfree and the Berendsen scale factors for
// the next step (src/md.f90:6098-6190).  One workgroup per instance.
__global__ __launch_bounds__(256) void k_temperature(const DevSys* __restrict__ sys, int nblocks) {
  const DevSys& S = sys[blockIdx.y];
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
  // end of the step: advance the batch's step counter once
  if (blockIdx.y == 0) *S.step_ctr = *S.step_ctr + 1;
}

// Sum the per-item energy partials into the instance's energy terms.
// With record != 0 it runs only on energy-file steps (step counter a
// positive multiple of iene) and also appends the frame's put_ene records.
__global__ __launch_bounds__(256) void k_energy_reduce(const DevSys* __restrict__ sys, int record) {
  const DevSys& S = sys[blockIdx.y];
  __shared__ double red[12][4];
  int frame = -1;
  if (record) {
    const int st = *S.step_ctr - S.eseries_base;
    if (S.iene <= 0 || st <= 0 || st % S.iene != 0) return;
    frame = st / S.iene - 1;
  }
  double w[4] = {0, 0, 0, 0};
  for (int iw = threadIdx.x; iw < S.nwat; iw += blockDim.x)
    for (int c = 0; c < 4; ++c) w[c] += S.e_water[(size_t)iw * 4 + c];
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < S.nat_solute; i += blockDim.x)
    for (int c = 0; c < 8; ++c) p[c] += S.e_solute[(size_t)i * 8 + c];
  double b[6] = {0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < S.nbonded_atoms; t += blockDim.x)
    for (int c = 0; c < 6; ++c) b[c] += S.e_bonded[(size_t)t * 6 + c];
  for (int c = 0; c < 4; ++c) w[c] = block_sum_256(w[c], red[c]);
  for (int c = 0; c < 8; ++c) p[c] = block_sum_256(p[c], red[4 + c]);
  for (int c = 0; c < 6; ++c) b[c] = block_sum_256(b[c], red[c]);
  // per-state Q energies: qq, qp (Q-neighbour pairs of nonbond_qq and the
  // Q-solute list), qw
  for (int st = 0; st < S.nstates; ++st) {
    double q[6] = {0, 0, 0, 0, 0, 0};
    for (int iq = threadIdx.x; iq < S.nqat; iq += blockDim.x) {
      const double* e = S.e_q + ((size_t)iq * S.nstates + st) * 4;
      q[2] += e[0]; q[3] += e[1]; q[4] += e[2]; q[5] += e[3];
    }
    for (int pr = threadIdx.x; pr < S.nqq; pr += blockDim.x) {
      if (S.qq_state[pr] != st) continue;
      const int o = S.qq_jq[pr] >= 0 ? 0 : 2;
      q[o] += S.e_qq[2 * pr];
      q[o + 1] += S.e_qq[2 * pr + 1];
    }
    for (int c = 0; c < 6; ++c) {
      q[c] = block_sum_256(q[c], red[c]);
      if (threadIdx.x == 0) S.qenergies[st * kQTerms + c] = q[c];
    }
    if (threadIdx.x == 0 && frame >= 0 && frame < S.eseries_cap) {
      // put_ene record: total, q(bond angle torsion improper), qx, qq, qp,
      // qw (el, vdw each), restraint (src/nrgy.f90:128-130; EQ totals as
      // pot_energy forms them, src/md.f90:16186-16194)
      double* r = S.eseries + ((size_t)frame * S.nstates + st) * kEneRecord;
      const double qx_el = q[0] + q[2] + q[4], qx_vdw = q[1] + q[3] + q[5];
      r[0] = 0.0 + 0.0 + 0.0 + 0.0 + qx_el + qx_vdw + 0.0;
      r[1] = 0.0; r[2] = 0.0; r[3] = 0.0; r[4] = 0.0;
      r[5] = qx_el; r[6] = qx_vdw;
      r[7] = q[0]; r[8] = q[1]; r[9] = q[2]; r[10] = q[3]; r[11] = q[4]; r[12] = q[5];
      r[13] = 0.0;
    }
  }
  if (threadIdx.x == 0 && frame >= 0) *S.eseries_n = frame + 1;
  if (threadIdx.x == 0) {
    S.energies[E_P_BOND] = b[0]; S.energies[E_P_ANGLE] = b[1]; S.energies[E_P_TORSION] = b[2];
    S.energies[E_W_BOND] = b[3]; S.energies[E_W_ANGLE] = b[4]; S.energies[E_W_TORSION] = b[5];
    S.energies[E_P_IMPROPER] = 0.0;
    double* E = S.energies;
    E[E_WW_EL] = w[0]; E[E_WW_VDW] = w[1]; E[E_LRF] = w[2] + p[4]; E[E_RST_RADIAL] = w[3];
    E[E_PP_EL] = p[0]; E[E_PP_VDW] = p[1]; E[E_PW_EL] = p[2]; E[E_PW_VDW] = p[3];
    E[E_RST_FIX] = p[5]; E[E_RST_SHELL] = p[6];
    double ew = 0.0;
    for (int s = 0; s < S.nshell; ++s) ew = ew + S.e_wpol[s];
    E[E_RST_WPOL] = S.wpol_restr ? ew : 0.0;
  }
}

}  // namespace qgpu
