// ORACLE — test infrastructure only.  A plain FP64 CPU restatement of serial
// reference qdyn's MD step for spherical boundaries (the path the GPU engine
// replaces), order-faithful so that it reproduces reference qdyn on the same
// inputs.  Only the tests, the smoke check and the benchmark's CPU-baseline
// leg load this library; the GPU product never links or calls it.
//
// Compile with -ffp-contract=off: every expression keeps the Fortran
// operation order and no multiply-add is fused.
//
// Functions and the reference code each one follows (src/md.f90):
//   md_run step loop            6471-6664
//   make_pair_lists             5985-6071   cgp_centers 2253-2288
//   nbpplist_lrf                7998-8188   nbpwlist_lrf 9300-9429
//   nbqplist                   10215-10322  nbqwlist    10464-10549
//   nbwwlist_lrf               10897-11021  lrf_update   5896-5981
//   pot_energy                 16015-16221  pot_energy_nonbonds 16261-16344
//   nonbond_pp 13069-13155, nonbond_pw 13282-13365, nonbond_qp 13788-13921,
//   nonbond_ww_spc 15021-15119, nonbond_qw_spc 14411-14570,
//   nonbond_3atomsolvent 15238-15333, nonbond_qw_3atom 14749-14874,
//   nonbond_qq 13491-13636, lrf_taylor 5824-5891
//   bond 1332-1405, angle 1130-1246, torsion 17907-18052, improper 3130-3272
//   fix_shell 2816-2862, restrain_solvent 18055-18102, watpol 18372-18541
//   p_restrain 15473-15802
//   qbond 17192-17245, qangle 17022-17114, qtorsion 17395-17540,
//   qimproper 17249-17392
//   newvel_ber 6194-6223, shake 17581-17679, temperature 6098-6190
#include <array>
#include "../qsource_amd/csrc/host/qsys.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using namespace qsys;

namespace {

const double kPi = 4.0 * std::atan(1.0);

// cgp: periodic box only, the charge-group pair of the entry (its shift)
struct NB { int i, j, ljcod; int cgp = -1; };
struct NBQ { int iq, j, ljcod; int cgp = -1; };  // Q-atom number, partner atom (or first atom of water)

struct LRF {
  double cent[3];
  double phi0, phi1[3], phi2[9], phi3[27];
};

struct Energies {
  // classical
  double pp_el = 0, pp_vdw = 0, pw_el = 0, pw_vdw = 0, ww_el = 0, ww_vdw = 0;
  double p_bond = 0, p_angle = 0, p_torsion = 0, p_improper = 0;
  double w_bond = 0, w_angle = 0, w_torsion = 0, w_improper = 0;
  double q_bond = 0, q_angle = 0, q_torsion = 0, q_improper = 0;
  double qx_el = 0, qx_vdw = 0, lrf = 0;
  double rst_fix = 0, rst_shell = 0, rst_protein = 0, rst_radial = 0, rst_wpol = 0, rst_total = 0;
  double potential = 0, kinetic = 0;
};
struct QEnergies {
  double qq_el = 0, qq_vdw = 0, qp_el = 0, qp_vdw = 0, qw_el = 0, qw_vdw = 0;
  double qx_el = 0, qx_vdw = 0, total = 0, restraint = 0;
  double bond = 0, angle = 0, torsion = 0, improper = 0;
};

struct Oracle {
  Setup s;
  std::vector<double> x, v, xx, d;
  std::vector<NB> nbpp, nbpw, nbww;
  std::vector<NBQ> nbqp, nbqw;
  bool qp_done = false, qw_done = false;
  // periodic box: the switching atoms of each listed charge-group pair
  // (nbpp_cgp, nbpw_cgp) and the partner atom of each Q-solute group
  // (nbqp_cgp); shifts formed from them at every force evaluation
  std::vector<std::array<int, 2>> pp_cgp, pw_cgp;
  std::vector<int> qp_cgp;
  std::vector<LRF> lrf;
  Energies E;
  std::vector<QEnergies> EQ;
  double Temp = 0, Tfree = 0, Temp_solute = 0, Tfree_solute = 0, Texcl_solute = 0;
  double Temp_solvent = 0, Tfree_solvent = 0, Texcl_solvent = 0;
  double Tscale_solute = 0, Tscale_solvent = 0, Ekinmax = 0;
  int istep = 0;
  // watpol work arrays
  std::vector<double> theta, theta0, tdum;
  std::vector<std::vector<int>> list_sh, nsort;
  std::vector<double> cgp_cent_dummy;
  int last_shake_iter = 0;
  std::string err;
  // constant pressure: volume moves tried and accepted, the accept flag of
  // each try, and randm's argument (the stream itself is s.rng)
  int volume_try = 0, volume_acc = 0;
  std::vector<int> volume_log;
  int pressure_seed = 0;
  // randm's argument for the Langevin forces; the Nose-Hoover chain
  int iseed = 0;
  std::vector<double> xnh, vnh, gnh;

  const Topology& t() const { return s.top; }
  int ncgp() const { return (int)s.top.cgp.size(); }
  int natyps() const { return s.natyps; }
  int ljcod(int a, int b) const { return s.ljcod[a * s.natyps + b]; }
  const AtomType& typ(int i) const { return s.top.types[s.iac[i]]; }
  // sc_lookup(iq, column, state), 0-based column (topology types, then Q atoms)
  double sc(int iq, int col, int st) const {
    return s.sc_lookup[((size_t)iq * (s.natyps + s.nqat) + col) * s.nstates + st];
  }

  explicit Oracle(const std::string& inp) : s(prepare(inp)) {
    x = s.x; v = s.v;
    xx.assign(x.size(), 0.0);
    d.assign(x.size(), 0.0);
    lrf.assign(ncgp(), LRF{});
    EQ.assign(s.nstates, QEnergies{});
    theta.assign(s.nwat, 0.0); theta0.assign(s.nwat, 0.0); tdum.assign(s.nwat, 0.0);
    list_sh.assign(s.wshell.size(), {});
    nsort.assign(s.wshell.size(), {});
    Ekinmax = 1000.0 * s.ndegf * kBoltz * s.inp.temp0 / 2.0 / (double)s.natom;
    pressure_seed = s.inp.pressure_seed;
    iseed = s.inp.iseed;
    xnh.assign(s.qnh.size(), 0.0); vnh.assign(s.qnh.size(), 0.0); gnh.assign(s.qnh.size(), 0.0);
    temperature();
  }

  // ---------------------------------------------------------------- lists
  bool excluded_pair(int i, int j) const {
    if (std::abs(j - i) <= kMaxNbrRange) {
      if (i < j) return (s.top.listex[i] >> (j - i - 1)) & 1u;
      return (s.top.listex[j] >> (i - j - 1)) & 1u;
    }
    for (auto& p : s.top.listexlong)
      if ((p.first == i && p.second == j) || (p.first == j && p.second == i)) return true;
    return false;
  }
  bool pair14(int i, int j) const {
    if (std::abs(j - i) <= kMaxNbrRange) {
      if (i < j) return (s.top.list14[i] >> (j - i - 1)) & 1u;
      return (s.top.list14[j] >> (i - j - 1)) & 1u;
    }
    bool r = false;
    for (auto& p : s.top.list14long)
      if ((p.first == i && p.second == j) || (p.first == j && p.second == i)) r = true;
    return r;
  }
  // the box shift of one component: boxlength * nint(d / boxlength)
  // (Fortran nint rounds half away from zero, as std::round)
  double bshift(double dd, int k) const { return s.boxl[k] * std::round(dd * s.inv_boxl[k]); }
  // minimum image of one component
  double mimage(double dd, int k) const { return dd - s.boxl[k] * std::round(dd * s.inv_boxl[k]); }

  static bool skip_parity(int ig, int jg) {
    // 1-based parity rule (src/md.f90:10977-10980) with 0-based arguments
    int a = ig + 1, b = jg + 1;
    return (a > b && (a + b) % 2 == 0) || (a < b && (a + b) % 2 == 1);
  }

  void cgp_centers() {
    const auto& T = s.top;
    for (int ig = 0; ig < ncgp(); ++ig) {
      LRF& L = lrf[ig];
      L.cent[0] = L.cent[1] = L.cent[2] = 0;
      L.phi0 = 0;
      for (double& p : L.phi1) p = 0;
      for (double& p : L.phi2) p = 0;
      for (double& p : L.phi3) p = 0;
      for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
        int i3 = 3 * T.cgpatom[a];
        L.cent[0] = L.cent[0] + x[i3];
        L.cent[1] = L.cent[1] + x[i3 + 1];
        L.cent[2] = L.cent[2] + x[i3 + 2];
      }
      double n = (double)(T.cgp[ig].last - T.cgp[ig].first + 1);
      for (int k = 0; k < 3; ++k) L.cent[k] = L.cent[k] / n;
    }
  }

  void lrf_update(int g1, int g2) {
    const auto& T = s.top;
    LRF& L = lrf[g2];
    // periodic box: the shift of group 1's switching atom against group 2's
    // centre (src/md.f90:5904-5917)
    double sh[3] = {0.0, 0.0, 0.0};
    if (s.pbc) {
      const int w3 = 3 * T.cgp[g1].iswitch;
      for (int k = 0; k < 3; ++k) sh[k] = bshift(x[w3 + k] - L.cent[k], k);
    }
    for (int ia = T.cgp[g1].first; ia <= T.cgp[g1].last; ++ia) {
      int i = T.cgpatom[ia];
      if (s.iqatom[i] != 0) continue;
      int i3 = 3 * i;
      double dr[3] = {x[i3] - L.cent[0] - sh[0], x[i3 + 1] - L.cent[1] - sh[1], x[i3 + 2] - L.cent[2] - sh[2]};
      double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      double f0 = s.crg[i] / (r2 * std::sqrt(r2));
      double f1 = 3.0 * f0 / r2;
      double f2 = -f1 / r2;
      L.phi0 = L.phi0 + f0 * r2;
      for (int k = 0; k < 3; ++k) L.phi1[k] = L.phi1[k] - f0 * dr[k];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double v2 = L.phi2[3 * a + b] + f1 * dr[a] * dr[b];
          L.phi2[3 * a + b] = (a == b) ? v2 - f0 : v2;
        }
      // phi3(9a+3b+c) += f2*(5 dr_a dr_b dr_c - r2*(delta terms))
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            double t5 = 5.0 * dr[a] * dr[b] * dr[c];
            double term;
            if (a == b && b == c) term = t5 - r2 * 3.0 * dr[a];
            else if (a == b) term = t5 - r2 * dr[c];
            else if (a == c) term = t5 - r2 * dr[b];
            else if (b == c) term = t5 - r2 * dr[a];
            else term = t5;
            L.phi3[9 * a + 3 * b + c] = L.phi3[9 * a + 3 * b + c] + f2 * term;
          }
    }
  }

  void nbwwlist_lrf() {
    nbww.clear();
    const double rcut2 = s.inp.rcww * s.inp.rcww;
    const double rclrf2 = s.inp.rclrf * s.inp.rclrf;
    const int ns = s.nat_solute;
    for (int iw = 0; iw < s.nwat; ++iw) {
      int is = ns + 3 * iw;
      if (!s.pbc && s.excl[is]) continue;
      int is3 = 3 * is;
      int ig = s.iwhich_cgp[is];
      for (int jw = 0; jw < s.nwat; ++jw) {
        int ja = ns + 3 * jw;
        if (!s.pbc && s.excl[ja]) continue;
        int jg = s.iwhich_cgp[ja];
        int j3 = 3 * ja;
        if (skip_parity(iw, jw) || iw == jw) continue;
        double r2;
        if (s.pbc) {
          // nbwwlist_box(_lrf) (src/md.f90:10793-10895, 11023-11148)
          const double dx = mimage(x[is3] - x[j3], 0), dy = mimage(x[is3 + 1] - x[j3 + 1], 1),
                       dz = mimage(x[is3 + 2] - x[j3 + 2], 2);
          r2 = dx * dx + dy * dy + dz * dz;
        } else {
          r2 = (x[is3] - x[j3]) * (x[is3] - x[j3]) + (x[is3 + 1] - x[j3 + 1]) * (x[is3 + 1] - x[j3 + 1])
             + (x[is3 + 2] - x[j3 + 2]) * (x[is3 + 2] - x[j3 + 2]);
        }
        if (r2 <= rcut2) {
          for (int la = 0; la < 3; ++la)
            for (int ka = 0; ka < 3; ++ka)
              nbww.push_back({ns + 3 * iw + la, ns + 3 * jw + ka, ljcod(s.iac[ns + la], s.iac[ns + ka])});
        } else if (s.inp.use_lrf && r2 <= rclrf2) {
          lrf_update(ig, jg);
          lrf_update(jg, ig);
        }
      }
    }
  }

  // make_pair_lists (src/md.f90:6030-6058): with LRF off, nbpplist,
  // nbpwlist and nbwwlist (:7631-7806, :9047-9164, :10698-10791) build the
  // same lists as their _lrf forms and collect no multipoles
  void make_pair_lists() {
    if (s.inp.use_lrf) cgp_centers();
    if (s.pbc) {
      // periodic box (src/md.f90:6003-6029)
      nbpplist_box();
      nbpwlist_box();
      nbqplist_box();
      nbwwlist_lrf();
      nbqwlist_box();
      return;
    }
    nbpplist_lrf();
    nbpwlist_lrf();
    nbqplist();
    nbwwlist_lrf();
    nbqwlist();
  }

  // put_back_in_box (src/md.f90:18743-18930): the box centre follows the
  // solute (or all atoms) unless it is rigid; molecules whose centre of
  // mass left the box move by a box length per axis; with LRF the moved
  // molecules' group centres are formed again
  void put_back_in_box() {
    const auto& T = s.top;
    double boxc[3];
    if (!s.inp.rigid_box_centre) {
      const int last = s.nat_solute > 0 ? s.nat_solute : s.natom;
      boxc[0] = boxc[1] = boxc[2] = 0.0;
      for (int i = 0; i < last; ++i)
        for (int k = 0; k < 3; ++k) boxc[k] = boxc[k] + x[3 * i + k];
      for (int k = 0; k < 3; ++k) boxc[k] = boxc[k] / (double)last;
      for (int k = 0; k < 3; ++k) s.boxc[k] = boxc[k];
    } else {
      for (int k = 0; k < 3; ++k) boxc[k] = s.boxc[k];
    }
    double hi[3], lo[3];
    for (int k = 0; k < 3; ++k) { hi[k] = boxc[k] + s.boxl[k] / 2; lo[k] = boxc[k] - s.boxl[k] / 2; }
    const int m0 = s.inp.put_solute_back ? 0 : s.nsolute_mol;
    const int m1 = s.inp.put_solvent_back ? s.nmol : s.nsolute_mol;
    std::vector<char> moved(s.nmol, 0);
    for (int m = m0; m < m1; ++m) {
      double cm[3] = {0.0, 0.0, 0.0};
      const int a0 = T.istart_mol[m], a1 = T.istart_mol[m + 1];
      for (int j = a0; j < a1; ++j)
        for (int k = 0; k < 3; ++k) cm[k] = cm[k] + x[3 * j + k] * s.mass[j];
      for (int k = 0; k < 3; ++k) cm[k] = cm[k] * s.mol_minv[m];
      for (int k = 0; k < 3; ++k) {
        if (cm[k] > hi[k]) {
          for (int j = a0; j < a1; ++j) x[3 * j + k] = x[3 * j + k] - s.boxl[k];
          moved[m] = 1;
        } else if (cm[k] < lo[k]) {
          for (int j = a0; j < a1; ++j) x[3 * j + k] = x[3 * j + k] + s.boxl[k];
          moved[m] = 1;
        }
      }
    }
    if (s.inp.use_lrf)
      for (int m = m0; m < m1; ++m) {
        if (!moved[m]) continue;
        const int g0 = s.iwhich_cgp[T.istart_mol[m]], g1 = s.iwhich_cgp[T.istart_mol[m + 1] - 1];
        for (int ig = g0; ig <= g1; ++ig) {
          LRF& L = lrf[ig];
          L.cent[0] = L.cent[1] = L.cent[2] = 0.0;
          for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a)
            for (int k = 0; k < 3; ++k) L.cent[k] = L.cent[k] + x[3 * T.cgpatom[a] + k];
          for (int k = 0; k < 3; ++k) L.cent[k] = L.cent[k] / (double)(T.cgp[ig].last - T.cgp[ig].first + 1);
        }
      }
  }

  // solute lists: filled in below
  void nbpplist_lrf();
  void nbpwlist_lrf();
  void nbqplist();
  void nbqwlist();
  void nbpplist_box();
  void nbpwlist_box();
  void nbqplist_box();
  void nbqwlist_box();

  // ---------------------------------------------------------------- forces
  // with a periodic box (nonbond_ww_spc_box, src/md.f90:15123-15236) the
  // shift of the water pair is formed from the O-O vector of its first row
  // and kept for the pair's other rows
  void nonbond_ww_spc() {
    int ichg = 1;
    double bs[3] = {0.0, 0.0, 0.0};
    const size_t n = nbww.size();
    for (size_t iw = 0; iw < n; iw += 3) {
      if (ichg > 3) ichg = 1;
      int i = nbww[iw].i, j = nbww[iw].j;
      int i3 = 3 * i;
      double c1 = s.chg_solv[ichg - 1];
      for (int ip = 0; ip < 3; ++ip) {
        int ja = j + ip, j3 = 3 * ja;
        double c2 = s.chg_solv[ip];
        double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
        if (s.pbc) {
          if (ichg == 1 && ip == 0) { bs[0] = bshift(dx1, 0); bs[1] = bshift(dx2, 1); bs[2] = bshift(dx3, 2); }
          dx1 = dx1 - bs[0]; dx2 = dx2 - bs[1]; dx3 = dx3 - bs[2];
        }
        double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
        double r6 = r2 * r2 * r2;
        r2 = 1.0 / r2;
        double r = std::sqrt(r2);
        double Vel = c1 * c2 * r;
        E.ww_el = E.ww_el + Vel;
        double dv = r2 * (-Vel);
        if (ichg == 1 && ip == 0) {
          int iLJ = ljcod(s.iac[i], s.iac[j]);
          double a1 = s.alj_solv[0][iLJ - 1], a2 = s.alj_solv[0][iLJ - 1];
          double b1 = s.blj_solv[0][iLJ - 1], b2 = s.blj_solv[0][iLJ - 1];
          r6 = 1.0 / r6;
          double Va = a1 * a2 * r6 * r6;
          double Vb = b1 * b2 * r6;
          E.ww_vdw = E.ww_vdw + Va - Vb;
          dv = dv + r2 * (-12.0 * Va + 6.0 * Vb);
        }
        d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
        d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
      }
      ++ichg;
    }
  }

  void lrf_taylor() {
    for (int i = 0; i < s.natom; ++i) {
      // excluded atoms take part in a periodic box (src/md.f90:5837)
      if ((!s.pbc && s.excl[i]) || s.iqatom[i] != 0) continue;
      int i3 = 3 * i;
      const LRF& L = lrf[s.iwhich_cgp[i]];
      double dr[3] = {L.cent[0] - x[i3], L.cent[1] - x[i3 + 1], L.cent[2] - x[i3 + 2]};
      const double* p2 = L.phi2;
      const double* p3 = L.phi3;
      double Vij = L.phi0 + L.phi1[0] * dr[0] + L.phi1[1] * dr[1] + L.phi1[2] * dr[2]
                 + 0.5 * (p2[0] * dr[0] + p2[1] * dr[1] + p2[2] * dr[2]) * dr[0]
                 + 0.5 * (p2[3] * dr[0] + p2[4] * dr[1] + p2[5] * dr[2]) * dr[1]
                 + 0.5 * (p2[6] * dr[0] + p2[7] * dr[1] + p2[8] * dr[2]) * dr[2];
      E.lrf = E.lrf + 0.5 * s.crg[i] * Vij;
      double df[3];
      for (int a = 0; a < 3; ++a) {
        const double* q = p3 + 9 * a;
        df[a] = L.phi1[a] + p2[3 * a] * dr[0] + p2[3 * a + 1] * dr[1] + p2[3 * a + 2] * dr[2]
              + 0.5 * (q[0] * dr[0] + q[1] * dr[1] + q[2] * dr[2]) * dr[0]
              + 0.5 * (q[3] * dr[0] + q[4] * dr[1] + q[5] * dr[2]) * dr[1]
              + 0.5 * (q[6] * dr[0] + q[7] * dr[1] + q[8] * dr[2]) * dr[2];
      }
      d[i3] = d[i3] - s.crg[i] * df[0];
      d[i3 + 1] = d[i3 + 1] - s.crg[i] * df[1];
      d[i3 + 2] = d[i3 + 2] - s.crg[i] * df[2];
    }
  }

  void nonbond_pp();
  void nonbond_pw();
  void nonbond_qp();
  void nonbond_qw_spc();
  void nonbond_qw_3atom();

  // nonbond_3atomsolvent (src/md.f90:15238-15333): LJ between every atom
  // pair of two waters, the LJ code from atom i and the other water's
  // first atom, parameters of the atoms' positions in the water
  void nonbond_3atomsolvent() {
    int ichg = 1;
    const size_t n = nbww.size();
    for (size_t iw = 0; iw < n; iw += 3) {
      if (ichg > 3) ichg = 1;
      int i = nbww[iw].i, j = nbww[iw].j;
      int i3 = 3 * i;
      double c1 = s.chg_solv[ichg - 1];
      for (int ip = 0; ip < 3; ++ip) {
        int ja = j + ip, j3 = 3 * ja;
        double c2 = s.chg_solv[ip];
        double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
        double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
        double r6 = r2 * r2 * r2;
        r2 = 1.0 / r2;
        double r = std::sqrt(r2);
        double Vel = c1 * c2 * r;
        double dv = r2 * (-Vel);
        int iLJ = ljcod(s.iac[i], s.iac[j]);
        double a1 = s.alj_solv[ichg - 1][iLJ - 1], a2 = s.alj_solv[ip][iLJ - 1];
        double b1 = s.blj_solv[ichg - 1][iLJ - 1], b2 = s.blj_solv[ip][iLJ - 1];
        r6 = 1.0 / r6;
        double Va = a1 * a2 * r6 * r6;
        double Vb = b1 * b2 * r6;
        E.ww_el = E.ww_el + Vel;
        E.ww_vdw = E.ww_vdw + Va - Vb;
        dv = dv + r2 * (-12.0 * Va + 6.0 * Vb);
        d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
        d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
      }
      ++ichg;
    }
  }
  void nonbond_qq();
  double bond(int a, int b);
  double angle(int a, int b);
  double torsion(int a, int b);
  double improper(int a, int b);
  double improper2(int a, int b);
  void nonbond_qp_qvdw();
  void offdiag();
  // EVB off-diagonal elements of the last offdiag() (per [off_diagonals]
  // line: Hij and r_kl)
  std::vector<double> offd_hij, offd_rkl;

  // Morse factors of the Q bonds, kept between calls as the reference's
  // module arrays EMorseD and dMorse_i/j (src/qatom.f90:201-203), zero at start
  std::vector<double> emorse, dmorse_i, dmorse_j;
  void qbond(int st);
  void qangle(int st);
  void qtorsion(int st);
  void qimproper(int st);
  void p_restrain();

  void fix_shell() {
    for (int i = 0; i < s.natom; ++i) {
      if (!(s.excl[i] || s.shell[i])) continue;
      int i3 = 3 * i;
      double fk;
      if (s.excl[i]) {
        fk = s.inp.fk_fix;
        if (s.inp.freeze) {
          v[i3] = v[i3 + 1] = v[i3 + 2] = 0;
          x[i3] = s.xtop[i3]; x[i3 + 1] = s.xtop[i3 + 1]; x[i3 + 2] = s.xtop[i3 + 2];
          fk = 0;
        }
      } else {
        fk = s.inp.fk_pshell;
      }
      double dr[3] = {x[i3] - s.xtop[i3], x[i3 + 1] - s.xtop[i3 + 1], x[i3 + 2] - s.xtop[i3 + 2]};
      double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      double erst = 0.5 * fk * r2;
      if (s.excl[i]) E.rst_fix = E.rst_fix + erst;
      if (s.shell[i]) E.rst_shell = E.rst_shell + erst;
      d[i3] = d[i3] + fk * dr[0];
      d[i3 + 1] = d[i3 + 1] + fk * dr[1];
      d[i3 + 2] = d[i3 + 2] + fk * dr[2];
    }
  }

  void restrain_solvent() {
    double shift = s.fk_wsphere != 0 ? std::sqrt(kBoltz * Tfree / s.fk_wsphere) : 0.0;
    const auto& T = s.top;
    for (int iw = T.ncgp_solute; iw < ncgp(); ++iw) {
      int i = T.cgp[iw].iswitch;
      if (s.excl[i]) continue;
      int i3 = 3 * i;
      double dr[3] = {x[i3] - T.xwcent[0], x[i3 + 1] - T.xwcent[1], x[i3 + 2] - T.xwcent[2]};
      double b = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double db = b - (s.rwat - shift);
      double erst, dv;
      if (db > 0) {
        erst = 0.5 * s.fk_wsphere * db * db - s.dwmz;
        dv = s.fk_wsphere * db / b;
      } else if (b > 0) {
        double fexp = std::exp(s.awmz * db);
        erst = s.dwmz * (fexp * fexp - 2.0 * fexp);
        dv = -2.0 * s.dwmz * s.awmz * (fexp - fexp * fexp) / b;
      } else {
        dv = 0; erst = 0;
      }
      E.rst_radial = E.rst_radial + erst;
      d[i3] = d[i3] + dv * dr[0];
      d[i3 + 1] = d[i3 + 1] + dv * dr[1];
      d[i3 + 2] = d[i3 + 2] + dv * dr[2];
    }
  }

  void watpol() {
    const int nsh = (int)s.wshell.size();
    const auto& xw = s.top.xwcent;
    for (auto& w : s.wshell) w.n_insh = 0;
    for (auto& l : list_sh) l.clear();
    for (int iw = 0; iw < s.nwat; ++iw) {
      theta[iw] = 0; theta0[iw] = 0;
      int i = s.nat_solute + 3 * iw;
      if (s.excl[i]) continue;
      int i3 = 3 * i;
      double rmu[3], rcu[3];
      for (int k = 0; k < 3; ++k) rmu[k] = x[i3 + 3 + k] + x[i3 + 6 + k] - 2.0 * x[i3 + k];
      double rm = std::sqrt(rmu[0] * rmu[0] + rmu[1] * rmu[1] + rmu[2] * rmu[2]);
      for (int k = 0; k < 3; ++k) rmu[k] = rmu[k] / rm;
      for (int k = 0; k < 3; ++k) rcu[k] = x[i3 + k] - xw[k];
      double rc = std::sqrt(rcu[0] * rcu[0] + rcu[1] * rcu[1] + rcu[2] * rcu[2]);
      for (int k = 0; k < 3; ++k) rcu[k] = rcu[k] / rc;
      double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
      if (scp > 1.0) scp = 1.0;
      if (scp < -1.0) scp = -1.0;
      theta[iw] = std::acos(scp);
      tdum[iw] = theta[iw];
      if (rc > s.wshell[nsh - 1].rout - s.wshell[nsh - 1].dr) {
        int is;
        for (is = nsh - 1; is >= 1; --is)
          if (rc <= s.wshell[is].rout) break;
        if (is < 1) is = 0;
        s.wshell[is].n_insh += 1;
        list_sh[is].push_back(iw);
      }
    }
    for (int is = 0; is < nsh; ++is) {
      int n = s.wshell[is].n_insh;
      nsort[is].assign(n, 0);
      int jmin = 0;
      for (int il = 0; il < n; ++il) {
        double tmin = 2.0 * kPi;
        for (int jl = 0; jl < n; ++jl) {
          int jw = list_sh[is][jl];
          if (tdum[jw] < tmin) { jmin = jw; tmin = theta[jw]; }
        }
        nsort[is][il] = jmin;
        tdum[jmin] = 99999.0;
      }
    }
    if (istep != 0 && istep % kItdisUpdate == 0) {
      for (auto& w : s.wshell) {
        w.avtheta = w.avtheta / (double)kItdisUpdate;
        w.avn_insh = w.avn_insh / (double)kItdisUpdate;
        w.theta_corr = w.theta_corr + w.avtheta - std::acos(w.cstb);
        w.avtheta = 0; w.avn_insh = 0;
      }
    }
    for (int is = 0; is < nsh; ++is) {
      WaterShell& w = s.wshell[is];
      if (w.n_insh == 0) continue;
      double avtdum = 0;
      for (int il = 0; il < w.n_insh; ++il) {
        int iw = nsort[is][il];
        double arg = 1.0 + (1.0 - 2.0 * (double)(il + 1)) / (double)w.n_insh;
        double th0 = std::acos(arg);
        th0 = th0 - 3.0 * std::sin(th0) * w.cstb / 2.0;
        if (th0 < 0) th0 = 0;
        if (th0 > kPi) th0 = kPi;
        theta0[il] = th0;
        avtdum = avtdum + theta[iw];
        double dth = theta[iw] - th0 + w.theta_corr;
        E.rst_wpol = E.rst_wpol + 0.5 * s.fkwpol * dth * dth;
        double dv = s.fkwpol * (theta[iw] - th0 + w.theta_corr);
        int i = s.nat_solute + 3 * iw;
        int i3 = 3 * i;
        double rmu[3], rcu[3];
        for (int k = 0; k < 3; ++k) rmu[k] = x[i3 + 3 + k] + x[i3 + 6 + k] - 2.0 * x[i3 + k];
        double rm = std::sqrt(rmu[0] * rmu[0] + rmu[1] * rmu[1] + rmu[2] * rmu[2]);
        for (int k = 0; k < 3; ++k) rmu[k] = rmu[k] / rm;
        for (int k = 0; k < 3; ++k) rcu[k] = x[i3 + k] - xw[k];
        double rc = std::sqrt(rcu[0] * rcu[0] + rcu[1] * rcu[1] + rcu[2] * rcu[2]);
        for (int k = 0; k < 3; ++k) rcu[k] = rcu[k] / rc;
        double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
        if (scp > 1.0) scp = 1.0;
        if (scp < -1.0) scp = -1.0;
        double f0 = std::sin(std::acos(scp));
        if (std::fabs(f0) < kQrealEps) f0 = kQrealEps;
        f0 = -1.0 / f0;
        f0 = dv * f0;
        double f1[9], f2[3];
        for (int k = 0; k < 3; ++k) {
          f1[k] = -2.0 * (rcu[k] - rmu[k] * scp) / rm;
          f1[3 + k] = (rcu[k] - rmu[k] * scp) / rm;
          f1[6 + k] = (rcu[k] - rmu[k] * scp) / rm;
          f2[k] = (rmu[k] - rcu[k] * scp) / rc;
        }
        for (int k = 0; k < 3; ++k) d[i3 + k] = d[i3 + k] + f0 * (f1[k] + f2[k]);
        for (int k = 3; k < 9; ++k) d[i3 + k] = d[i3 + k] + f0 * (f1[k]);
      }
      w.avtheta = w.avtheta + avtdum / (double)w.n_insh;
      w.avn_insh = w.avn_insh + w.n_insh;
    }
  }

  // pot_energy_nonbonds (src/md.f90:16261-16344; sphere, geometric rule):
  // adds into d, E and EQ
  void pot_energy_nonbonds() {
    nonbond_pp();
    nonbond_pw();
    if (s.qvdw_flag) nonbond_qp_qvdw();
    else nonbond_qp();
    if (s.natom > s.nat_solute) {
      if (s.top.solvent_type == 1) {
        if (s.pbc) throw std::runtime_error("3-atom solvent in a periodic box is not supported");
        nonbond_3atomsolvent();
        nonbond_qw_3atom();
      } else {
        nonbond_ww_spc();
        nonbond_qw_spc();
      }
    }
    if (s.inp.use_lrf) lrf_taylor();
  }

  void pot_energy() {
    const double kinetic = E.kinetic;  // pot_energy leaves E%kinetic alone (src/md.f90:16025)
    E = Energies{};
    E.kinetic = kinetic;
    for (auto& q : EQ) q = QEnergies{};
    std::fill(d.begin(), d.end(), 0.0);
    pot_energy_nonbonds();
    // bonded (FF_GROMOS/default impropers)
    E.p_bond = bond(0, s.nbonds_solute);
    E.w_bond = bond(s.nbonds_solute, (int)s.bnd.size());
    E.p_angle = angle(0, s.nangles_solute);
    E.w_angle = angle(s.nangles_solute, (int)s.ang.size());
    E.p_torsion = torsion(0, s.ntors_solute);
    E.w_torsion = torsion(s.ntors_solute, (int)s.tor.size());
    // pot_energy_bonds (src/md.f90:16224-16256): harmonic impropers for
    // GROMOS, periodic for AMBER
    if (s.ff_type == 2) {
      E.p_improper = improper2(0, s.nimps_solute);
      E.w_improper = improper2(s.nimps_solute, (int)s.imp.size());
    } else {
      E.p_improper = improper(0, s.nimps_solute);
      E.w_improper = improper(s.nimps_solute, (int)s.imp.size());
    }
    // no shell or solvent restraints in a periodic box (src/md.f90:16114-16126)
    if (!s.pbc) fix_shell();
    p_restrain();
    if (s.nwat > 0 && !s.pbc) {
      restrain_solvent();
      if (s.wpol_restr) watpol();
    }
    nonbond_qq();
    for (int st = 0; st < s.nstates; ++st) {
      qbond(st);
      qangle(st);
      qtorsion(st);
      qimproper(st);
    }
    for (int st = 0; st < s.nstates; ++st) {
      QEnergies& q = EQ[st];
      q.qx_el = q.qq_el + q.qp_el + q.qw_el;
      q.qx_vdw = q.qq_vdw + q.qp_vdw + q.qw_vdw;
      q.total = q.bond + q.angle + q.torsion + q.improper + q.qx_el + q.qx_vdw + q.restraint;
      double lam = s.lambda[st];
      E.q_bond = E.q_bond + q.bond * lam;
      E.q_angle = E.q_angle + q.angle * lam;
      E.q_torsion = E.q_torsion + q.torsion * lam;
      E.q_improper = E.q_improper + q.improper * lam;
      E.qx_el = E.qx_el + q.qx_el * lam;
      E.qx_vdw = E.qx_vdw + q.qx_vdw * lam;
      E.rst_protein = E.rst_protein + q.restraint * lam;
    }
    E.rst_total = E.rst_fix + E.rst_shell + E.rst_protein + E.rst_radial + E.rst_wpol;
    E.potential = E.p_bond + E.w_bond + E.p_angle + E.w_angle + E.p_torsion + E.p_improper + E.pp_el
                + E.pp_vdw + E.pw_el + E.pw_vdw + E.ww_el + E.ww_vdw + E.q_bond + E.q_angle
                + E.q_torsion + E.q_improper + E.qx_el + E.qx_vdw + E.rst_total + E.lrf;
  }

  // ---------------------------------------------------------------- update
  // newvel_ber (src/md.f90:6194-6223); velocity-Verlet repeats the velocity
  // update after the position update
  void newvel_ber(double step, double tscale, int a0, int a1) {
    for (int i = a0; i < a1; ++i) {
      int i3 = 3 * i;
      for (int k = 0; k < 3; ++k) {
        v[i3 + k] = (v[i3 + k] - d[i3 + k] * s.winv[i] * step) * tscale;
        xx[i3 + k] = x[i3 + k];
        x[i3 + k] = x[i3 + k] + v[i3 + k] * s.dt;
      }
      if (s.velocity_verlet)
        for (int k = 0; k < 3; ++k) v[i3 + k] = (v[i3 + k] - d[i3 + k] * s.winv[i] * step) * tscale;
    }
  }

  // newvel_lan (src/md.f90:6225-6271): random forces drawn atom by atom
  // from randm's stream, uniform (rands (u - 1/2)) or Gaussian (gauss with
  // standard deviation rands), then friction-scaled leapfrog
  void newvel_lan(double step, double lfriction) {
    std::vector<double> rande(3 * (size_t)s.natom);
    for (int i = 0; i < s.natom; ++i)
      for (int j = 0; j < 3; ++j) {
        if (!s.langevin_gauss) {
          const double u = s.rng.next(iseed);
          rande[3 * i + j] = s.rands[i] * (u - 0.5);
        } else {
          double a = 0.0;
          for (int k = 0; k < 12; ++k) a = a + s.rng.next(iseed);
          rande[3 * i + j] = (a - 6.0) * s.rands[i] + 0.0;
        }
      }
    for (int i = 0; i < s.natom; ++i) {
      int i3 = 3 * i;
      for (int k = 0; k < 3; ++k) {
        v[i3 + k] = (v[i3 + k] * lfriction) - (d[i3 + k] - rande[i3 + k]) * s.winv[i] * step;
        xx[i3 + k] = x[i3 + k];
        x[i3 + k] = x[i3 + k] + v[i3 + k] * s.dt;
      }
      if (s.velocity_verlet)
        for (int k = 0; k < 3; ++k)
          v[i3 + k] = (v[i3 + k] * lfriction) - (d[i3 + k] - rande[i3 + k]) * s.winv[i] * step;
    }
  }

  // newvel_nos (src/md.f90:6273-6301)
  void newvel_nos(double step) {
    for (int i = 0; i < s.natom; ++i) {
      int i3 = 3 * i;
      for (int k = 0; k < 3; ++k) {
        v[i3 + k] = (v[i3 + k] - d[i3 + k] * s.winv[i] * step);
        xx[i3 + k] = x[i3 + k];
        x[i3 + k] = x[i3 + k] + v[i3 + k] * s.dt;
      }
      if (s.velocity_verlet)
        for (int k = 0; k < 3; ++k) v[i3 + k] = (v[i3 + k] - d[i3 + k] * s.winv[i] * step);
    }
  }

  // nh_prop (src/md.f90:6303-6344): half-step propagation of the
  // Nose-Hoover chain from the last temperature, velocities scaled by
  // exp(-vnh_1 dt/2)
  void nh_prop() {
    const int n = (int)s.qnh.size();
    const double dt2 = 0.5 * s.dt, dt4 = 0.5 * dt2, dt8 = 0.5 * dt4, T0 = s.inp.temp0;
    const std::vector<double>& q = s.qnh;
    gnh[0] = s.ndegf * kBoltz * (Temp - T0) / q[0];
    for (int i = 1; i < n; ++i) gnh[i] = (q[i - 1] * (vnh[i - 1] * vnh[i - 1]) - s.kbt) / q[i];
    vnh[n - 1] = vnh[n - 1] + gnh[n - 1] * dt4;
    xnh[n - 1] = xnh[n - 1] + vnh[n - 1] * dt2;
    double expf = 0.0;
    for (int i = n - 2; i >= 0; --i) {
      expf = std::exp(-(dt8 * vnh[i + 1]));
      vnh[i] = (vnh[i] * expf + gnh[i] * dt4) * expf;
      xnh[i] = xnh[i] + vnh[i] * dt2;
    }
    const double sc = std::exp(-(vnh[0] * dt2));
    for (double& w : v) w = w * sc;
    gnh[0] = s.ndegf * kBoltz * (sc * sc * Temp - T0) / q[0];
    vnh[0] = (vnh[0] * expf + gnh[0] * dt4) * expf;
    for (int i = 1; i < n - 1; ++i) {
      expf = std::exp(-(dt8 * vnh[i + 1]));
      gnh[i] = (q[i - 1] * (vnh[i - 1] * vnh[i - 1]) - s.kbt) / q[i];
      vnh[i] = (vnh[i] * expf + gnh[i] * dt4) * expf;
    }
    gnh[n - 1] = (q[n - 2] * (vnh[n - 2] * vnh[n - 2]) - s.kbt) / q[n - 1];
    vnh[n - 1] = vnh[n - 1] + gnh[n - 1] * dt4;
  }

  int shake() {
    int total = 0;
    for (const ShakeMol& m : s.shake_mol) {
      int nc = (int)m.bond.size();
      std::vector<char> ready(nc, 0);
      int nits = 0;
      for (;;) {
        for (int ic = 0; ic < nc; ++ic) {
          if (ready[ic]) continue;
          int i = m.bond[ic].i, j = m.bond[ic].j, i3 = 3 * i, j3 = 3 * j;
          double a0 = x[i3] - x[j3], a1 = x[i3 + 1] - x[j3 + 1], a2 = x[i3 + 2] - x[j3 + 2];
          double xij2 = a0 * a0 + a1 * a1 + a2 * a2;
          double diff = m.bond[ic].dist2 - xij2;
          if (std::fabs(diff) < kShakeTol * m.bond[ic].dist2) ready[ic] = 1;
          double b0 = xx[i3] - xx[j3], b1 = xx[i3 + 1] - xx[j3 + 1], b2 = xx[i3 + 2] - xx[j3 + 2];
          double scp = a0 * b0 + a1 * b1 + a2 * b2;
          double corr = diff / (2.0 * scp * (s.winv[i] + s.winv[j]));
          x[i3] = x[i3] + b0 * corr * s.winv[i];
          x[i3 + 1] = x[i3 + 1] + b1 * corr * s.winv[i];
          x[i3 + 2] = x[i3 + 2] + b2 * corr * s.winv[i];
          x[j3] = x[j3] - b0 * corr * s.winv[j];
          x[j3 + 1] = x[j3 + 1] - b1 * corr * s.winv[j];
          x[j3 + 2] = x[j3 + 2] - b2 * corr * s.winv[j];
        }
        ++nits;
        bool all = true;
        for (int ic = 0; ic < nc; ++ic) all = all && ready[ic];
        if (all) break;
        if (nits >= kShakeMaxIter) return -1;
      }
      total += nits;
    }
    return total;
  }

  void temperature() {
    Temp = 0; Temp_solute = 0; Tfree_solute = 0; Texcl_solute = 0;
    for (int i = 0; i < s.nat_solute; ++i) {
      int i3 = 3 * i;
      double ek = 0.5 * typ(i).mass * (v[i3] * v[i3] + v[i3 + 1] * v[i3 + 1] + v[i3 + 2] * v[i3 + 2]);
      Temp_solute = Temp_solute + ek;
      if (!s.excl[i]) Tfree_solute = Tfree_solute + ek; else Texcl_solute = Texcl_solute + ek;
    }
    Tfree_solvent = 0; Temp_solvent = 0; Texcl_solvent = 0;
    for (int i = s.nat_solute; i < s.natom; ++i) {
      int i3 = 3 * i;
      double ek = 0.5 * typ(i).mass * (v[i3] * v[i3] + v[i3 + 1] * v[i3 + 1] + v[i3 + 2] * v[i3 + 2]);
      Temp_solvent = Temp_solvent + ek;
      if (!s.excl[i]) Tfree_solvent = Tfree_solvent + ek; else Texcl_solvent = Texcl_solvent + ek;
    }
    Tfree = Tfree_solvent + Tfree_solute;
    Temp = Temp_solute + Temp_solvent;
    E.kinetic = Temp;
    Temp = 2.0 * Temp / kBoltz / (double)s.ndegf;
    Tfree = 2.0 * Tfree / kBoltz / (double)s.ndegfree;
    if (s.detail_temps) {
      Temp_solute = 2.0 * Temp_solute / kBoltz / (double)s.ndegf_solute;
      Tfree_solute = 2.0 * Tfree_solute / kBoltz / (double)s.ndegfree_solute;
      Temp_solvent = 2.0 * Temp_solvent / kBoltz / (double)s.ndegf_solvent;
      Tfree_solvent = 2.0 * Tfree_solvent / kBoltz / (double)s.ndegfree_solvent;
    }
    const double T0 = s.inp.temp0, dt = s.dt, tau = s.tau_t;
    if (s.separate_scaling) {
      if (Tfree_solvent != 0) Tscale_solvent = T0 / Tfree_solvent - 1.0;
      Tscale_solvent = std::sqrt(1 + dt / tau * Tscale_solvent);
      if (Tfree_solute != 0) Tscale_solute = T0 / Tfree_solute - 1.0;
      Tscale_solute = std::sqrt(1 + dt / tau * Tscale_solute);
    } else {
      if (Tfree != 0) Tscale_solvent = T0 / Tfree - 1.0;
      Tscale_solvent = std::sqrt(1 + dt / tau * Tscale_solvent);
      Tscale_solute = Tscale_solvent;
    }
  }

  // new_potential (src/md.f90:19276-19434, geometric rule): the nonbonded,
  // LRF, restraint and Q-Q energies at the current coordinates with the
  // lists as they are; the potential takes the bonded and Q-bonded energies
  // of `old` (the last pot_energy), without its solvent torsions and
  // impropers, as the reference sums it
  void new_potential(const Energies& old) {
    E.potential = 0; E.pp_el = 0; E.pp_vdw = 0; E.pw_el = 0; E.pw_vdw = 0; E.ww_el = 0; E.ww_vdw = 0;
    E.qx_el = 0; E.qx_vdw = 0; E.rst_protein = 0; E.lrf = 0;
    E.p_bond = 0; E.w_bond = 0; E.p_angle = 0; E.w_angle = 0; E.p_torsion = 0; E.w_torsion = 0;
    E.p_improper = 0; E.w_improper = 0;
    for (auto& q : EQ) {
      q.qq_el = q.qq_vdw = q.qp_el = q.qp_vdw = q.qw_el = q.qw_vdw = 0;
      q.restraint = 0;
    }
    std::fill(d.begin(), d.end(), 0.0);
    pot_energy_nonbonds();
    E.p_bond = bond(0, s.nbonds_solute);
    E.w_bond = bond(s.nbonds_solute, (int)s.bnd.size());
    E.p_angle = angle(0, s.nangles_solute);
    E.w_angle = angle(s.nangles_solute, (int)s.ang.size());
    E.p_torsion = torsion(0, s.ntors_solute);
    E.w_torsion = torsion(s.ntors_solute, (int)s.tor.size());
    if (s.ff_type == 2) {
      E.p_improper = improper2(0, s.nimps_solute);
      E.w_improper = improper2(s.nimps_solute, (int)s.imp.size());
    } else {
      E.p_improper = improper(0, s.nimps_solute);
      E.w_improper = improper(s.nimps_solute, (int)s.imp.size());
    }
    p_restrain();
    nonbond_qq();
    for (int st = 0; st < s.nstates; ++st) {
      QEnergies& q = EQ[st];
      q.qx_el = q.qq_el + q.qp_el + q.qw_el;
      q.qx_vdw = q.qq_vdw + q.qp_vdw + q.qw_vdw;
      E.qx_el = E.qx_el + q.qx_el * s.lambda[st];
      E.qx_vdw = E.qx_vdw + q.qx_vdw * s.lambda[st];
      E.rst_protein = E.rst_protein + q.restraint * s.lambda[st];
    }
    E.potential = old.p_bond + old.w_bond + old.p_angle + old.w_angle + old.p_torsion + old.p_improper
                + E.pp_el + E.pp_vdw + E.pw_el + E.pw_vdw + E.ww_el + E.ww_vdw + old.q_bond + old.q_angle
                + old.q_torsion + old.q_improper + E.qx_el + E.qx_vdw + E.rst_protein + E.lrf;
  }

  // MC_volume (src/md.f90:18931-19272), molecule-based scaling: a random
  // volume change of at most max_volume_displ, every molecule's centre of
  // mass scaled about the box centre, the lists and LRF built again (with
  // LRF on), and the move accepted by the Metropolis test on
  // deltaE + P deltaV - nmol kT ln(V'/V).  A rejected move restores x, the
  // box, and (with LRF) every list and the multipoles except the Q-water
  // list, which the reference keeps from the scaled coordinates.
  bool mc_volume() {
    const std::vector<double> old_x = x;
    const Energies previous_E = E;
    double old_boxl[3], old_inv[3];
    for (int k = 0; k < 3; ++k) { old_boxl[k] = s.boxl[k]; old_inv[k] = s.inv_boxl[k]; }
    const bool lrf_on = s.inp.use_lrf;
    std::vector<NB> o_pp, o_pw, o_ww;
    std::vector<NBQ> o_qp;
    std::vector<std::array<int, 2>> o_ppc, o_pwc;
    std::vector<int> o_qpc;
    std::vector<LRF> o_lrf;
    if (lrf_on) {
      o_pp = nbpp; o_pw = nbpw; o_ww = nbww; o_qp = nbqp;
      o_ppc = pp_cgp; o_pwc = pw_cgp; o_qpc = qp_cgp; o_lrf = lrf;
    }
    new_potential(previous_E);
    const Energies old_E = E;
    const double old_V = old_boxl[0] * old_boxl[1] * old_boxl[2];
    double r = s.rng.next(pressure_seed);
    r = r * 2 - 1;
    const double deltaV = r * s.inp.max_vol_displ;
    const double new_V = deltaV + old_V;
    const double cubr = std::pow(new_V / old_V, 1.0 / 3.0);
    for (int k = 0; k < 3; ++k) s.boxl[k] = s.boxl[k] * cubr;
    for (int k = 0; k < 3; ++k) s.inv_boxl[k] = 1.0 / s.boxl[k];
    const double bmin = std::min(s.boxl[0], std::min(s.boxl[1], s.boxl[2]));
    const MDInput& in = s.inp;
    const char* small = nullptr;
    if (!(bmin > in.rcpp * 2)) small = "Solute-solute cut-off radii too large";
    else if (!(bmin > in.rcww * 2)) small = "Solvent-solvent cut-off radii too large";
    else if (!(bmin > in.rcpw * 2)) small = "Solute-solvent cut-off radii too large";
    else if (!(bmin > in.rcq * 2)) small = "Q-atom cut-off radii too large";
    else if (in.use_lrf && !(bmin > in.rclrf * 2)) small = "LRF cut-off radii too large";
    if (small) { err = small; return false; }
    const auto& T = s.top;
    for (int m = 0; m < s.nmol; ++m) {
      double cm[3] = {0.0, 0.0, 0.0};
      const int a0 = T.istart_mol[m], a1 = T.istart_mol[m + 1];
      for (int j = a0; j < a1; ++j)
        for (int k = 0; k < 3; ++k) cm[k] = cm[k] + x[3 * j + k] * s.mass[j];
      for (int k = 0; k < 3; ++k) cm[k] = cm[k] * s.mol_minv[m];
      double mv[3];
      for (int k = 0; k < 3; ++k)
        mv[k] = ((cm[k] - s.boxc[k]) * s.boxl[k] / old_boxl[k] + s.boxc[k]) - cm[k];
      for (int j = a0; j < a1; ++j)
        for (int k = 0; k < 3; ++k) x[3 * j + k] = x[3 * j + k] + mv[k];
    }
    if (lrf_on) make_pair_lists();
    new_potential(previous_E);
    const double deltaE = E.potential - old_E.potential;
    const double deltaW = deltaE + in.pressure * deltaV - s.nmol * kBoltz * in.temp0 * std::log(new_V / old_V);
    bool acc = true, drawn = false;
    if (!(deltaW <= 0.0)) {
      const double u = s.rng.next(pressure_seed);
      acc = !(u > std::exp(-deltaW / kBoltz / in.temp0));
      drawn = true;
    }
    volume_try += 1;
    volume_log.push_back(acc ? (drawn ? 2 : 1) : 0);
    if (acc) {
      volume_acc += 1;
    } else {
      x = old_x;
      E = previous_E;
      for (int k = 0; k < 3; ++k) { s.boxl[k] = old_boxl[k]; s.inv_boxl[k] = old_inv[k]; }
      if (lrf_on) {
        nbpp = o_pp; nbpw = o_pw; nbww = o_ww; nbqp = o_qp;
        pp_cgp = o_ppc; pw_cgp = o_pwc; qp_cgp = o_qpc; lrf = o_lrf;
      }
    }
    return true;
  }

  // one md_run iteration (src/md.f90:6471-6612) without output
  bool step() {
    if (istep % s.inp.nbcycle == 0) {
      if (s.pbc && (s.inp.put_solute_back || s.inp.put_solvent_back)) put_back_in_box();
      make_pair_lists();
      // a volume move at list updates (src/md.f90:6538-6544)
      if (s.pbc && s.inp.constant_pressure && istep % s.inp.ivolume_cycle == 0 && istep > 0)
        if (!mc_volume()) return false;
    }
    pot_energy();
    // thermostat (src/md.f90:6578-6588)
    if (s.thermostat == 0) {
      newvel_ber(s.dv_mod, Tscale_solute, 0, s.nat_solute);
      newvel_ber(s.dv_mod, Tscale_solvent, s.nat_solute, s.natom);
    } else if (s.thermostat == 1) {
      newvel_lan(s.dv_mod, s.dv_friction);
    } else {
      nh_prop();
      newvel_nos(s.dv_mod);
    }
    if (s.shake_constraints > 0) {
      int n = shake();
      if (n < 0) { err = "shake failure"; return false; }
      last_shake_iter = n / s.nmol;
      for (size_t k = 0; k < x.size(); ++k) v[k] = (x[k] - xx[k]) / s.dt;
      if (s.thermostat == 2) nh_prop();   // velocities scaled again after SHAKE
    }
    temperature();
    ++istep;
    return true;
  }
};

// ------------------------------------------------------------------ solute
void Oracle::nbpplist_lrf() {
  nbpp.clear();
  const auto& T = s.top;
  const double rcut2 = s.inp.rcpp * s.inp.rcpp, rclrf2 = s.inp.rclrf * s.inp.rclrf;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    int ia = T.cgp[ig].iswitch;
    if (s.excl[ia]) continue;
    int i3 = 3 * ia;
    for (int jg = 0; jg < T.ncgp_solute; ++jg) {
      if (skip_parity(ig, jg)) continue;
      int ja = T.cgp[jg].iswitch;
      if (s.excl[ja]) continue;
      int j3 = 3 * ja;
      double r2 = (x[i3] - x[j3]) * (x[i3] - x[j3]) + (x[i3 + 1] - x[j3 + 1]) * (x[i3 + 1] - x[j3 + 1])
                + (x[i3 + 2] - x[j3 + 2]) * (x[i3 + 2] - x[j3 + 2]);
      if (r2 <= rcut2) {
        for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
          int i = T.cgpatom[a];
          if (s.iqatom[i] != 0) continue;
          for (int b = T.cgp[jg].first; b <= T.cgp[jg].last; ++b) {
            int j = T.cgpatom[b];
            if (ig == jg && i >= j) continue;
            if (s.iqatom[j] != 0) continue;
            int lj = ljcod(s.iac[i], s.iac[j]);
            if (s.crg[i] * s.crg[j] == 0.0 && typ(i).avdw[lj - 1] * typ(j).avdw[lj - 1] == 0.0 &&
                typ(i).bvdw[lj - 1] * typ(j).bvdw[lj - 1] == 0.0)
              continue;
            if (excluded_pair(i, j)) continue;
            nbpp.push_back({i, j, pair14(i, j) ? 3 : lj});
          }
        }
      } else if (s.inp.use_lrf && r2 <= rclrf2) {
        lrf_update(ig, jg);
        lrf_update(jg, ig);
      }
    }
  }
}

void Oracle::nbpwlist_lrf() {
  nbpw.clear();
  const auto& T = s.top;
  const double rcut2 = s.inp.rcpw * s.inp.rcpw, rclrf2 = s.inp.rclrf * s.inp.rclrf;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    int ia = T.cgp[ig].iswitch;
    if (s.excl[ia]) continue;
    int i3 = 3 * ia;
    for (int jw = 0; jw < s.nwat; ++jw) {
      int ja = s.nat_solute + 3 * jw;
      if (s.excl[ja]) continue;
      int j3 = 3 * ja;
      int jg = s.iwhich_cgp[ja];
      double r2 = (x[i3] - x[j3]) * (x[i3] - x[j3]) + (x[i3 + 1] - x[j3 + 1]) * (x[i3 + 1] - x[j3 + 1])
                + (x[i3 + 2] - x[j3 + 2]) * (x[i3 + 2] - x[j3 + 2]);
      if (r2 <= rcut2) {
        for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
          int i = T.cgpatom[a];
          if (s.iqatom[i] != 0) continue;
          for (int k = 0; k < 3; ++k) {
            int j = ja + k;
            int lj = ljcod(s.iac[i], s.iac[j]);
            if (s.crg[i] * s.crg[j] == 0.0 && typ(i).avdw[lj - 1] * typ(j).avdw[lj - 1] == 0.0 &&
                typ(i).bvdw[lj - 1] * typ(j).bvdw[lj - 1] == 0.0)
              continue;
            nbpw.push_back({i, j, lj});
          }
        }
      } else if (s.inp.use_lrf && r2 <= rclrf2) {
        lrf_update(ig, jg);
        lrf_update(jg, ig);
      }
    }
  }
}

// connectivity of solute atom a to Q atom iq in state st (make_qconn)
static inline int qconn_of(const Setup& s, int st, int iq, int a) {
  return s.qconn[((size_t)st * s.nqat + iq) * s.nat_solute + a];
}

void Oracle::nbqplist() {
  // static once Rcq > rexcl_o (src/md.f90:10245)
  if (qp_done && s.inp.rcq > s.top.rexcl_o) return;
  nbqp.clear();
  if (s.nqat == 0) { qp_done = true; return; }
  const auto& T = s.top;
  const double rcut2 = s.inp.rcq * s.inp.rcq;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    int ia = T.cgp[ig].iswitch;
    if (s.excl[ia]) continue;
    int i3 = 3 * ia;
    double r2 = (x[i3] - T.xpcent[0]) * (x[i3] - T.xpcent[0]) + (x[i3 + 1] - T.xpcent[1]) * (x[i3 + 1] - T.xpcent[1])
              + (x[i3 + 2] - T.xpcent[2]) * (x[i3 + 2] - T.xpcent[2]);
    if (r2 > rcut2) continue;
    for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
      int i = T.cgpatom[a];
      bool conn = false;
      for (int st = 0; st < s.nstates && !conn; ++st)
        for (int iq = 0; iq < s.nqat && !conn; ++iq) conn = qconn_of(s, st, iq, i) <= 3;
      if (conn) continue;
      for (int iq = 0; iq < s.nqat; ++iq) {
        int lj = ljcod(s.iac[i], s.iac[s.iqseq[iq]]);
        if (qconn_of(s, 0, iq, i) == 4) lj = 3;
        nbqp.push_back({iq, i, lj});
      }
    }
  }
  qp_done = true;
}

void Oracle::nbqwlist() {
  if (qw_done && s.inp.rcq > s.top.rexcl_o) return;
  nbqw.clear();
  if (s.nqat == 0) { qw_done = true; return; }
  const auto& T = s.top;
  const double rcut2 = s.inp.rcq * s.inp.rcq;
  for (int iw = 0; iw < s.nwat; ++iw) {
    int ia = s.nat_solute + 3 * iw;
    if (s.excl[ia]) continue;
    int i3 = 3 * ia;
    double r2 = (x[i3] - T.xpcent[0]) * (x[i3] - T.xpcent[0]) + (x[i3 + 1] - T.xpcent[1]) * (x[i3 + 1] - T.xpcent[1])
              + (x[i3 + 2] - T.xpcent[2]) * (x[i3 + 2] - T.xpcent[2]);
    if (r2 <= rcut2)
      for (int iq = 0; iq < s.nqat; ++iq)
        for (int is = 0; is < 3; ++is)
          nbqw.push_back({iq, ia + is, ljcod(s.iac[s.nat_solute + is], s.iac[s.iqseq[iq]])});
  }
  qw_done = true;
}
// ------------------------------------------------------------- periodic box
// nbpplist_box(_lrf) (src/md.f90:7810-7997, 8190-8396): charge groups by
// the minimum-image distance of their switching atoms, no excluded atoms;
// each listed group pair keeps its switching atoms for the shift.
void Oracle::nbpplist_box() {
  nbpp.clear();
  pp_cgp.clear();
  const auto& T = s.top;
  const double rcut2 = s.inp.rcpp * s.inp.rcpp, rclrf2 = s.inp.rclrf * s.inp.rclrf;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    const int ig_sw = T.cgp[ig].iswitch, i3 = 3 * ig_sw;
    for (int jg = 0; jg < T.ncgp_solute; ++jg) {
      if (skip_parity(ig, jg)) continue;
      const int jg_sw = T.cgp[jg].iswitch, j3 = 3 * jg_sw;
      const double dx = mimage(x[i3] - x[j3], 0), dy = mimage(x[i3 + 1] - x[j3 + 1], 1),
                   dz = mimage(x[i3 + 2] - x[j3 + 2], 2);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 <= rcut2) {
        pp_cgp.push_back({ig_sw, jg_sw});
        const int cg = (int)pp_cgp.size() - 1;
        for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
          int i = T.cgpatom[a];
          if (s.iqatom[i] != 0) continue;
          for (int b = T.cgp[jg].first; b <= T.cgp[jg].last; ++b) {
            int j = T.cgpatom[b];
            if (ig == jg && i >= j) continue;
            if (s.iqatom[j] != 0) continue;
            int lj = ljcod(s.iac[i], s.iac[j]);
            if (s.crg[i] * s.crg[j] == 0.0 && typ(i).avdw[lj - 1] * typ(j).avdw[lj - 1] == 0.0 &&
                typ(i).bvdw[lj - 1] * typ(j).bvdw[lj - 1] == 0.0)
              continue;
            if (excluded_pair(i, j)) continue;
            nbpp.push_back({i, j, pair14(i, j) ? 3 : lj, cg});
          }
        }
      } else if (s.inp.use_lrf && r2 <= rclrf2) {
        lrf_update(ig, jg);
        lrf_update(jg, ig);
      }
    }
  }
}

// nbpwlist_box(_lrf) (src/md.f90:9169-9299, 9431-9580)
void Oracle::nbpwlist_box() {
  nbpw.clear();
  pw_cgp.clear();
  const auto& T = s.top;
  const double rcut2 = s.inp.rcpw * s.inp.rcpw, rclrf2 = s.inp.rclrf * s.inp.rclrf;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    const int ig_sw = T.cgp[ig].iswitch, i3 = 3 * ig_sw;
    for (int jw = 0; jw < s.nwat; ++jw) {
      const int ja = s.nat_solute + 3 * jw, j3 = 3 * ja, jg = s.iwhich_cgp[ja];
      const double dx = mimage(x[i3] - x[j3], 0), dy = mimage(x[i3 + 1] - x[j3 + 1], 1),
                   dz = mimage(x[i3 + 2] - x[j3 + 2], 2);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 <= rcut2) {
        pw_cgp.push_back({ig_sw, ja});
        const int cg = (int)pw_cgp.size() - 1;
        for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
          int i = T.cgpatom[a];
          if (s.iqatom[i] != 0) continue;
          for (int k = 0; k < 3; ++k) {
            int j = ja + k;
            int lj = ljcod(s.iac[i], s.iac[j]);
            if (s.crg[i] * s.crg[j] == 0.0 && typ(i).avdw[lj - 1] * typ(j).avdw[lj - 1] == 0.0 &&
                typ(i).bvdw[lj - 1] * typ(j).bvdw[lj - 1] == 0.0)
              continue;
            nbpw.push_back({i, j, lj, cg});
          }
        }
      } else if (s.inp.use_lrf && r2 <= rclrf2) {
        lrf_update(ig, jg);
        lrf_update(jg, ig);
      }
    }
  }
}

// nbqplist_box (src/md.f90:10327-10553): a solute charge group enters when
// any of its atoms lies within Rcq of the Q switching atom (minimum image);
// that atom gives the group's shift.  Rebuilt at every list update (built
// once only for Rcq < 0, which this restatement refuses).
void Oracle::nbqplist_box() {
  nbqp.clear();
  qp_cgp.clear();
  if (s.nqat == 0) return;
  const auto& T = s.top;
  const double rcut2 = s.inp.rcq * s.inp.rcq;
  const int q3 = 3 * s.qswitch;
  for (int ig = 0; ig < T.ncgp_solute; ++ig) {
    bool inside = false;
    for (int a = T.cgp[ig].first; a <= T.cgp[ig].last && !inside; ++a) {
      const int i = T.cgpatom[a], i3 = 3 * i;
      const double dx = mimage(x[i3] - x[q3], 0), dy = mimage(x[i3 + 1] - x[q3 + 1], 1),
                   dz = mimage(x[i3 + 2] - x[q3 + 2], 2);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 <= rcut2) {
        inside = true;
        qp_cgp.push_back(i);
      }
    }
    if (!inside) continue;
    const int cg = (int)qp_cgp.size() - 1;
    for (int a = T.cgp[ig].first; a <= T.cgp[ig].last; ++a) {
      int i = T.cgpatom[a];
      bool conn = false;
      for (int st = 0; st < s.nstates && !conn; ++st)
        for (int iq = 0; iq < s.nqat && !conn; ++iq) conn = qconn_of(s, st, iq, i) <= 3;
      if (conn) continue;
      for (int iq = 0; iq < s.nqat; ++iq) {
        int lj = ljcod(s.iac[i], s.iac[s.iqseq[iq]]);
        if (qconn_of(s, 0, iq, i) == 4) lj = 3;
        nbqp.push_back({iq, i, lj, cg});
      }
    }
  }
}

// nbqwlist_box (src/md.f90:10554-10700): waters whose O lies within Rcq of
// the Q switching atom (minimum image); rebuilt at every list update
void Oracle::nbqwlist_box() {
  nbqw.clear();
  if (s.nqat == 0) return;
  const double rcut2 = s.inp.rcq * s.inp.rcq;
  const int q3 = 3 * s.qswitch;
  for (int iw = 0; iw < s.nwat; ++iw) {
    const int ia = s.nat_solute + 3 * iw, i3 = 3 * ia;
    const double dx = mimage(x[i3] - x[q3], 0), dy = mimage(x[i3 + 1] - x[q3 + 1], 1),
                 dz = mimage(x[i3 + 2] - x[q3 + 2], 2);
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 <= rcut2)
      for (int iq = 0; iq < s.nqat; ++iq)
        for (int is = 0; is < 3; ++is)
          nbqw.push_back({iq, ia + is, ljcod(s.iac[s.nat_solute + is], s.iac[s.iqseq[iq]])});
  }
}

// With a periodic box (nonbond_pp_box, src/md.f90:13159-13280) every listed
// group pair's shift is formed from its switching atoms first.
void Oracle::nonbond_pp() {
  std::vector<std::array<double, 3>> sh(s.pbc ? pp_cgp.size() : 0);
  for (size_t g = 0; g < sh.size(); ++g) {
    const int a3 = 3 * pp_cgp[g][0], b3 = 3 * pp_cgp[g][1];
    for (int k = 0; k < 3; ++k) sh[g][k] = bshift(x[b3 + k] - x[a3 + k], k);
  }
  for (const NB& p : nbpp) {
    int i = p.i, j = p.j, i3 = 3 * i, j3 = 3 * j;
    double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
    if (s.pbc) { dx1 = dx1 - sh[p.cgp][0]; dx2 = dx2 - sh[p.cgp][1]; dx3 = dx3 - sh[p.cgp][2]; }
    double r2 = 1.0 / (dx1 * dx1 + dx2 * dx2 + dx3 * dx3);
    double r = std::sqrt(r2);
    double r6 = r2 * r2 * r2;
    double Vel = s.crg[i] * s.crg[j] * r;
    if (p.ljcod == 3) Vel = Vel * s.top.el14_scale;
    double Va = r6 * r6 * typ(i).avdw[p.ljcod - 1] * typ(j).avdw[p.ljcod - 1];
    double Vb = r6 * typ(i).bvdw[p.ljcod - 1] * typ(j).bvdw[p.ljcod - 1];
    double dv = r2 * (-Vel - 12.0 * Va + 6.0 * Vb);
    d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
    d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
    E.pp_el = E.pp_el + Vel;
    E.pp_vdw = E.pp_vdw + Va - Vb;
  }
}
// periodic box: nonbond_pw_box (src/md.f90:13369-13490)
void Oracle::nonbond_pw() {
  std::vector<std::array<double, 3>> sh(s.pbc ? pw_cgp.size() : 0);
  for (size_t g = 0; g < sh.size(); ++g) {
    const int a3 = 3 * pw_cgp[g][0], b3 = 3 * pw_cgp[g][1];
    for (int k = 0; k < 3; ++k) sh[g][k] = bshift(x[b3 + k] - x[a3 + k], k);
  }
  for (const NB& p : nbpw) {
    int i = p.i, j = p.j, i3 = 3 * i, j3 = 3 * j;
    double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
    if (s.pbc) { dx1 = dx1 - sh[p.cgp][0]; dx2 = dx2 - sh[p.cgp][1]; dx3 = dx3 - sh[p.cgp][2]; }
    double aLJ = typ(i).avdw[p.ljcod - 1] * typ(j).avdw[p.ljcod - 1];
    double bLJ = typ(i).bvdw[p.ljcod - 1] * typ(j).bvdw[p.ljcod - 1];
    double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    r2 = 1.0 / r2;
    double r = std::sqrt(r2);
    double r6 = r2 * r2 * r2;
    double r12 = r6 * r6;
    double Vel = s.crg[i] * s.crg[j] * r;
    double Va = aLJ * r12;
    double Vb = bLJ * r6;
    double dv = r2 * (-Vel - 12.0 * Va + 6.0 * Vb);
    d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
    d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
    E.pw_el = E.pw_el + Vel;
    E.pw_vdw = E.pw_vdw + Va - Vb;
  }
}
// periodic box: nonbond_qp_box (src/md.f90:13925-14096), shifts of the
// listed groups against the Q switching atom
void Oracle::nonbond_qp() {
  const int ns = s.nstates;
  std::vector<std::array<double, 3>> sh(s.pbc ? qp_cgp.size() : 0);
  for (size_t g = 0; g < sh.size(); ++g) {
    const int a3 = 3 * qp_cgp[g], q3 = 3 * s.qswitch;
    for (int k = 0; k < 3; ++k) sh[g][k] = bshift(x[a3 + k] - x[q3 + k], k);
  }
  for (const NBQ& p : nbqp) {
    int iq = p.iq, i = s.iqseq[iq], j = p.j, i3 = 3 * i, j3 = 3 * j;
    int iLJ = p.ljcod;
    double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
    if (s.pbc) { dx1 = dx1 - sh[p.cgp][0]; dx2 = dx2 - sh[p.cgp][1]; dx3 = dx3 - sh[p.cgp][2]; }
    double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    r2 = 1.0 / r2;
    double r6 = r2 * r2 * r2;
    double r = std::sqrt(r2);
    for (int st = 0; st < ns; ++st) {
      double aLJ = typ(i).avdw[iLJ - 1] * typ(j).avdw[iLJ - 1];
      double bLJ = typ(i).bvdw[iLJ - 1] * typ(j).bvdw[iLJ - 1];
      double Vel = s.qcrg[iq * ns + st] * s.crg[j] * r;
      if (iLJ == 3) Vel = Vel * s.top.el14_scale;
      double Va = aLJ * r6 * r6;
      double Vb = bLJ * r6;
      double dv = r2 * (-Vel - 12.0 * Va + 6.0 * Vb) * s.lambda[st];
      d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
      d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
      EQ[st].qp_el = EQ[st].qp_el + Vel;
      EQ[st].qp_vdw = EQ[st].qp_vdw + Va - Vb;
    }
  }
}

// periodic box: nonbond_qw_spc_box (src/md.f90:14574-14748), the shift of
// the water O against the Q switching atom for all three water atoms
void Oracle::nonbond_qw_spc() {
  const int ns = s.nstates;
  for (size_t iw = 0; iw < nbqw.size(); iw += 3) {
    int iq = nbqw[iw].iq, i = s.iqseq[iq], i3 = 3 * i;
    int j = nbqw[iw].j, j3 = 3 * j;
    double bs[3] = {0.0, 0.0, 0.0};
    if (s.pbc) {
      const int q3 = 3 * s.qswitch;
      for (int k = 0; k < 3; ++k) bs[k] = bshift(x[j3 + k] - x[q3 + k], k);
    }
    double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
    if (s.pbc) { dx1 = dx1 - bs[0]; dx2 = dx2 - bs[1]; dx3 = dx3 - bs[2]; }
    double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    double r6_hc = r2 * r2 * r2;
    r2 = 1.0 / r2;
    double r = std::sqrt(r2);
    double r6_sc = r6_hc;
    int iLJ = nbqw[iw].ljcod;
    double Va = 0, Vb = 0;
    if (!s.qvdw_flag) {
      double r6 = 1.0 / r6_hc;
      double aLJ = typ(i).avdw[iLJ - 1];
      double bLJ = typ(i).bvdw[iLJ - 1];
      Va = aLJ * s.alj_solv[0][iLJ - 1] * r6 * r6;
      Vb = bLJ * s.blj_solv[0][iLJ - 1] * r6;
    }
    double dv = 0;
    for (int st = 0; st < ns; ++st) {
      if (s.qvdw_flag) {
        // Q-atom type of this state, LJ code 1, soft-core column iLJ
        // (src/md.f90:14497-14505)
        const int qt = s.qiac[iq * ns + st];
        double aLJ = s.qavdw[3 * qt];
        double bLJ = s.qbvdw[3 * qt];
        r6_sc = r6_hc + sc(iq, iLJ - 1, st);
        Va = aLJ * s.alj_solv[0][iLJ - 1] / (r6_sc * r6_sc);
        Vb = bLJ * s.blj_solv[0][iLJ - 1] / (r6_sc);
      }
      double Vel = s.chg_solv[0] * s.qcrg[iq * ns + st] * r;
      dv = dv + r2 * (-Vel - ((12.0 * Va - 6.0 * Vb) * (r6_hc / r6_sc))) * s.lambda[st];
      EQ[st].qw_el = EQ[st].qw_el + Vel;
      EQ[st].qw_vdw = EQ[st].qw_vdw + Va - Vb;
    }
    d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
    d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
    for (int ip = 1; ip < 3; ++ip) {
      int ja = j + ip, ja3 = 3 * ja;
      double ex1 = x[ja3] - x[i3], ex2 = x[ja3 + 1] - x[i3 + 1], ex3 = x[ja3 + 2] - x[i3 + 2];
      if (s.pbc) { ex1 = ex1 - bs[0]; ex2 = ex2 - bs[1]; ex3 = ex3 - bs[2]; }
      double q2 = 1.0 / (ex1 * ex1 + ex2 * ex2 + ex3 * ex3);
      double rr = std::sqrt(q2);
      double dvh = 0;
      for (int st = 0; st < ns; ++st) {
        double Vel = s.chg_solv[ip] * s.qcrg[iq * ns + st] * rr;
        EQ[st].qw_el = EQ[st].qw_el + Vel;
        dvh = dvh - q2 * Vel * s.lambda[st];
      }
      d[i3] = d[i3] - dvh * ex1; d[i3 + 1] = d[i3 + 1] - dvh * ex2; d[i3 + 2] = d[i3 + 2] - dvh * ex3;
      d[ja3] = d[ja3] + dvh * ex1; d[ja3 + 1] = d[ja3 + 1] + dvh * ex2; d[ja3 + 2] = d[ja3 + 2] + dvh * ex3;
    }
  }
}

// nonbond_qw_3atom (src/md.f90:14749-14874): LJ and Coulomb on every water
// atom, the LJ code of each (water atom, Q atom) pair from the list
void Oracle::nonbond_qw_3atom() {
  const int ns = s.nstates;
  for (size_t iw = 0; iw < nbqw.size(); iw += 3) {
    int iq = nbqw[iw].iq, i = s.iqseq[iq], i3 = 3 * i;
    int j = nbqw[iw].j;
    for (int ip = 0; ip < 3; ++ip) {
      int ja = j + ip, j3 = 3 * ja;
      double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
      double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
      double r6_hc = r2 * r2 * r2;
      r2 = 1.0 / r2;
      double r = std::sqrt(r2);
      double r6_sc = r6_hc;
      int iLJ = nbqw[iw + ip].ljcod;
      double Va = 0, Vb = 0;
      if (!s.qvdw_flag) {
        double r6 = 1.0 / r6_hc;
        double aLJ = typ(i).avdw[iLJ - 1];
        double bLJ = typ(i).bvdw[iLJ - 1];
        Va = aLJ * s.alj_solv[ip][iLJ - 1] * r6 * r6;
        Vb = bLJ * s.blj_solv[ip][iLJ - 1] * r6;
      }
      double dv = 0;
      for (int st = 0; st < ns; ++st) {
        if (s.qvdw_flag) {
          const int qt = s.qiac[iq * ns + st];
          double aLJ = s.qavdw[3 * qt];
          double bLJ = s.qbvdw[3 * qt];
          r6_sc = r6_hc + sc(iq, iLJ - 1, st);
          Va = aLJ * s.alj_solv[ip][iLJ - 1] / (r6_sc * r6_sc);
          Vb = bLJ * s.blj_solv[ip][iLJ - 1] / (r6_sc);
        }
        double Vel = s.chg_solv[ip] * s.qcrg[iq * ns + st] * r;
        EQ[st].qw_el = EQ[st].qw_el + Vel;
        EQ[st].qw_vdw = EQ[st].qw_vdw + Va - Vb;
        dv = dv + r2 * (-Vel - ((12.0 * Va - 6.0 * Vb) * (r6_hc / r6_sc))) * s.lambda[st];
      }
      d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
      d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
    }
  }
}

void Oracle::nonbond_qq() {
  const int ns = s.nstates;
  for (int st = 0; st < ns; ++st) {
    for (const NbqPair& p : s.nbqq[st]) {
      int iq = p.iq, i = s.iqseq[iq], j = p.j, jq = p.jq;
      int i3 = 3 * i, j3 = 3 * j;
      int iLJ = p.ljcod;
      double qi = s.qcrg[iq * ns + st];
      double el_scale = p.el_scale;
      double aLJ, bLJ, qj;
      if (!s.qvdw_flag) {
        aLJ = typ(i).avdw[iLJ - 1] * typ(j).avdw[iLJ - 1];
        bLJ = typ(i).bvdw[iLJ - 1] * typ(j).bvdw[iLJ - 1];
        qj = jq >= 0 ? s.qcrg[jq * ns + st] : s.crg[j];
      } else {
        // Q-atom types (src/md.f90:13531-13550)
        const int qti = s.qiac[iq * ns + st];
        aLJ = s.qavdw[3 * qti + iLJ - 1];
        bLJ = s.qbvdw[3 * qti + iLJ - 1];
        if (jq >= 0) {
          const int qtj = s.qiac[jq * ns + st];
          aLJ = aLJ * s.qavdw[3 * qtj + iLJ - 1];
          bLJ = bLJ * s.qbvdw[3 * qtj + iLJ - 1];
          qj = s.qcrg[jq * ns + st];
        } else {
          if (iLJ == 2) {
            aLJ = s.qavdw[3 * qti];
            bLJ = s.qbvdw[3 * qti];
          }
          aLJ = aLJ * typ(j).avdw[iLJ - 1];
          bLJ = bLJ * typ(j).bvdw[iLJ - 1];
          qj = s.crg[j];
        }
      }
      if (s.qq_lib_charges) {
        // nonbond_qq_lib_charges (src/md.f90:13640-13700): library charges
        // for Q-Q pairs, the state's charge of i against non-Q atoms
        qi = jq >= 0 ? s.crg[i] : s.qcrg[iq * ns + st];
        qj = s.crg[j];
      }
      double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
      double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
      double r6_hc = r2 * r2 * r2;
      // soft-core column natyps + jq (1-based jq, 0 for a non-Q atom: the
      // reference then reads the column of the last topology type)
      double r6 = r2 * r2 * r2 + sc(iq, s.natyps + jq, st);
      r6 = 1.0 / r6;
      double r12 = r6 * r6;
      r2 = 1.0 / r2;
      double r = std::sqrt(r2);
      double Vel = qi * qj * r * el_scale;
      if (iLJ == 3) Vel = Vel * s.top.el14_scale;
      double Va, Vb, dv;
      if (s.qvdw_flag && jq >= 0 && iLJ == 2) {
        Va = aLJ * std::exp(-bLJ / r);
        Vb = 0.0;
        dv = r2 * (-Vel - bLJ * Va / r) * s.lambda[st];
      } else {
        Va = aLJ * r12;
        Vb = bLJ * r6;
        dv = r2 * (-Vel - (12.0 * Va - 6.0 * Vb) * r6_hc * r6) * s.lambda[st];
      }
      d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
      d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
      if (jq >= 0) {
        EQ[st].qq_el = EQ[st].qq_el + Vel;
        EQ[st].qq_vdw = EQ[st].qq_vdw + Va - Vb;
      } else {
        EQ[st].qp_el = EQ[st].qp_el + Vel;
        EQ[st].qp_vdw = EQ[st].qp_vdw + Va - Vb;
      }
    }
  }
}

double Oracle::bond(int a, int b) {
  double e = 0.0;
  for (int ib = a; ib < b; ++ib) {
    const Bond& B = s.bnd[ib];
    const BondLib& L = s.top.bondlib[B.cod - 1];
    int i3 = 3 * B.i, j3 = 3 * B.j;
    double r0 = x[j3] - x[i3], r1 = x[j3 + 1] - x[i3 + 1], r2 = x[j3 + 2] - x[i3 + 2];
    double bb = std::sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    double db = bb - L.b0;
    e = e + 0.5 * L.fk * (db * db);
    double dv = L.fk * db / bb;
    d[j3] = d[j3] + r0 * dv; d[j3 + 1] = d[j3 + 1] + r1 * dv; d[j3 + 2] = d[j3 + 2] + r2 * dv;
    d[i3] = d[i3] - r0 * dv; d[i3 + 1] = d[i3 + 1] - r1 * dv; d[i3 + 2] = d[i3 + 2] - r2 * dv;
  }
  return e;
}

double Oracle::angle(int a, int b) {
  double e = 0.0;
  for (int ia = a; ia < b; ++ia) {
    const Angle& A = s.ang[ia];
    const AngLib& L = s.top.anglib[A.cod - 1];
    int i3 = 3 * A.i, j3 = 3 * A.j, k3 = 3 * A.k;
    double rji[3], rjk[3];
    for (int c = 0; c < 3; ++c) { rji[c] = x[i3 + c] - x[j3 + c]; rjk[c] = x[k3 + c] - x[j3 + c]; }
    double bji2inv = 1.0 / (rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    double bjk2inv = 1.0 / (rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    double bjiinv = std::sqrt(bji2inv), bjkinv = std::sqrt(bjk2inv);
    double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]);
    scp = scp * bjiinv * bjkinv;
    if (scp > 1.0) scp = 1.0;
    else if (scp < -1.0) scp = -1.0;
    double angv = std::acos(scp);
    double da = angv - L.ang0;
    e = e + 0.5 * L.fk * (da * da);
    double dv = L.fk * da;
    double f1 = std::sin(angv);
    if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dk[3];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (rjk[c] * bjiinv * bjkinv - scp * rji[c] * bji2inv);
      dk[c] = f1 * (rji[c] * bjiinv * bjkinv - scp * rjk[c] * bjk2inv);
    }
    for (int c = 0; c < 3; ++c) d[i3 + c] = d[i3 + c] + dv * di[c];
    for (int c = 0; c < 3; ++c) d[k3 + c] = d[k3 + c] + dv * dk[c];
    for (int c = 0; c < 3; ++c) d[j3 + c] = d[j3 + c] - dv * (di[c] + dk[c]);
  }
  return e;
}

double Oracle::torsion(int a, int b) {
  double e = 0.0;
  for (int ip = b - 1; ip >= a; --ip) {
    const Torsion& t = s.tor[ip];
    const TorLib& L = s.top.torlib[t.cod - 1];
    int i3 = 3 * t.i, j3 = 3 * t.j, k3 = 3 * t.k, l3 = 3 * t.l;
    double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
    for (int c = 0; c < 3; ++c) {
      rji[c] = x[i3 + c] - x[j3 + c];
      rjk[c] = x[k3 + c] - x[j3 + c];
      rkl[c] = x[l3 + c] - x[k3 + c];
    }
    rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
    rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
    rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
    rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
    rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
    rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
    double bj2inv = 1.0 / (rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
    double bk2inv = 1.0 / (rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
    double bjinv = std::sqrt(bj2inv), bkinv = std::sqrt(bk2inv);
    double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) * (bjinv * bkinv);
    double phi;
    if (scp > 1.0) { scp = 1.0; phi = std::acos(1.0); }
    else if (scp < -1.0) { scp = -1.0; phi = std::acos(-1.0); }
    else phi = std::acos(scp);
    if (rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
            rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]) < 0)
      phi = -phi;
    double arg = L.rmult * phi - L.deltor;
    e = e + L.fk * (1.0 + std::cos(arg)) * L.paths;
    double dv = -L.rmult * L.fk * std::sin(arg) * L.paths;
    double f1 = std::sin(phi);
    if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dl[3], rki[3], rlj[3], dp[12];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (rnk[c] * (bjinv * bkinv) - scp * rnj[c] * bj2inv);
      dl[c] = f1 * (rnj[c] * (bjinv * bkinv) - scp * rnk[c] * bk2inv);
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
    for (int c = 0; c < 3; ++c) d[i3 + c] = d[i3 + c] + dv * dp[c];
    for (int c = 0; c < 3; ++c) d[j3 + c] = d[j3 + c] + dv * dp[3 + c];
    for (int c = 0; c < 3; ++c) d[k3 + c] = d[k3 + c] + dv * dp[6 + c];
    for (int c = 0; c < 3; ++c) d[l3 + c] = d[l3 + c] + dv * dp[9 + c];
  }
  return e;
}

// nonbond_qp_qvdw (src/md.f90:14097-14230): Q-solute pairs with the Q atom's
// type of each state and soft-core r^6 + alpha.
void Oracle::nonbond_qp_qvdw() {
  const int ns = s.nstates;
  std::vector<std::array<double, 3>> sh(s.pbc ? qp_cgp.size() : 0);
  for (size_t g = 0; g < sh.size(); ++g) {
    const int a3 = 3 * qp_cgp[g], q3 = 3 * s.qswitch;
    for (int k = 0; k < 3; ++k) sh[g][k] = bshift(x[a3 + k] - x[q3 + k], k);
  }
  for (const NBQ& p : nbqp) {
    int iq = p.iq, i = s.iqseq[iq], j = p.j, i3 = 3 * i, j3 = 3 * j;
    int iacj = s.iac[j];
    int iLJ = p.ljcod;
    int qLJ = iLJ == 2 ? 1 : iLJ;
    double dx1 = x[j3] - x[i3], dx2 = x[j3 + 1] - x[i3 + 1], dx3 = x[j3 + 2] - x[i3 + 2];
    if (s.pbc) { dx1 = dx1 - sh[p.cgp][0]; dx2 = dx2 - sh[p.cgp][1]; dx3 = dx3 - sh[p.cgp][2]; }
    double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
    double r6 = r2 * r2 * r2;
    r2 = 1.0 / r2;
    double r = std::sqrt(r2);
    for (int st = 0; st < ns; ++st) {
      const int iaci = s.qiac[iq * ns + st];
      double aLJ = s.qavdw[3 * iaci + qLJ - 1] * typ(j).avdw[iLJ - 1];
      double bLJ = s.qbvdw[3 * iaci + qLJ - 1] * typ(j).bvdw[iLJ - 1];
      double Vel = s.qcrg[iq * ns + st] * s.crg[j] * r;
      if (iLJ == 3) Vel = Vel * s.top.el14_scale;
      double r6_sc = r6 + sc(iq, iacj, st);
      double Va = aLJ / (r6_sc * r6_sc);
      double Vb = bLJ / (r6_sc);
      double dv = r2 * (-Vel - (12.0 * Va - 6.0 * Vb) * (r6 / r6_sc)) * s.lambda[st];
      d[i3] = d[i3] - dv * dx1; d[i3 + 1] = d[i3 + 1] - dv * dx2; d[i3 + 2] = d[i3 + 2] - dv * dx3;
      d[j3] = d[j3] + dv * dx1; d[j3 + 1] = d[j3 + 1] + dv * dx2; d[j3 + 2] = d[j3 + 2] + dv * dx3;
      EQ[st].qp_el = EQ[st].qp_el + Vel;
      EQ[st].qp_vdw = EQ[st].qp_vdw + Va - Vb;
    }
  }
}

// Dihedral geometry shared by the improper forms (as in torsion): phi and
// the 12 derivative components dp of phi scaled by -1/sin(phi).
static double dihedral(const std::vector<double>& x, int i, int j, int k, int l, double dp[12]) {
  const int i3 = 3 * i, j3 = 3 * j, k3 = 3 * k, l3 = 3 * l;
  double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
  for (int c = 0; c < 3; ++c) {
    rji[c] = x[i3 + c] - x[j3 + c];
    rjk[c] = x[k3 + c] - x[j3 + c];
    rkl[c] = x[l3 + c] - x[k3 + c];
  }
  rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
  rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
  rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
  rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
  rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
  rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
  double bj2inv = 1.0 / (rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
  double bk2inv = 1.0 / (rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
  double bjinv = std::sqrt(bj2inv), bkinv = std::sqrt(bk2inv);
  double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) * (bjinv * bkinv);
  if (scp > 1.0) scp = 1.0;
  else if (scp < -1.0) scp = -1.0;
  double phi = std::acos(scp);
  if (rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
          rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]) < 0)
    phi = -phi;
  double f1 = std::sin(phi);
  if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
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

// improper (src/md.f90:3130-3272): harmonic in the wrapped angle
// phi - imp0, terms visited from the last to the first.
double Oracle::improper(int a, int b) {
  double e = 0.0;
  for (int ip = b - 1; ip >= a; --ip) {
    const Torsion& t = s.imp[ip];
    const ImpLib& L = s.top.implib[t.cod - 1];
    double dp[12];
    double phi = dihedral(x, t.i, t.j, t.k, t.l, dp);
    double arg = phi - L.imp0;
    arg = arg - 2.0 * kPi * std::round(arg / (2.0 * kPi));
    double dv = L.fk * arg;
    e = e + 0.5 * dv * arg;
    const int at[4] = {t.i, t.j, t.k, t.l};
    for (int m = 0; m < 4; ++m)
      for (int c = 0; c < 3; ++c) d[3 * at[m] + c] = d[3 * at[m] + c] + dv * dp[3 * m + c];
  }
  return e;
}

// improper2 (src/md.f90:3276-3416): periodic fk (1 + cos(2 phi - imp0)).
double Oracle::improper2(int a, int b) {
  double e = 0.0;
  for (int ip = b - 1; ip >= a; --ip) {
    const Torsion& t = s.imp[ip];
    const ImpLib& L = s.top.implib[t.cod - 1];
    double dp[12];
    double phi = dihedral(x, t.i, t.j, t.k, t.l, dp);
    double arg = 2 * phi - L.imp0;
    e = e + L.fk * (1 + std::cos(arg));
    double dv = -2 * L.fk * std::sin(arg);
    const int at[4] = {t.i, t.j, t.k, t.l};
    for (int m = 0; m < 4; ++m)
      for (int c = 0; c < 3; ++c) d[3 * at[m] + c] = d[3 * at[m] + c] + dv * dp[3 * m + c];
  }
  return e;
}

// qbond (src/md.f90:17192-17245): Morse + harmonic Q bonds of one state;
// sets the bond's Morse factor and its gradient for coupled terms.
void Oracle::qbond(int st) {
  const size_t nb = s.qbnd.size();
  if (emorse.size() != nb) { emorse.assign(nb, 0.0); dmorse_i.assign(3 * nb, 0.0); dmorse_j.assign(3 * nb, 0.0); }
  const double lam = s.lambda[st];
  for (size_t ib = 0; ib < nb; ++ib) {
    const QTerm& b = s.qbnd[ib];
    const int ic = b.cod[st];
    if (ic <= 0) continue;
    const QBondLib& L = s.qbondlib[ic - 1];
    const int i3 = 3 * b.i, j3 = 3 * b.j;
    double rij[3] = {x[j3] - x[i3], x[j3 + 1] - x[i3 + 1], x[j3 + 2] - x[i3 + 2]};
    double bl = std::sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
    double db = bl - L.r0;
    double fexp = std::exp(-L.amz * db);
    double be = L.dmz * (fexp * fexp - 2.0 * fexp) + 0.5 * L.fk * (db * db);
    emorse[ib] = -(fexp * fexp - 2.0 * fexp);
    EQ[st].bond = EQ[st].bond + be;
    double dv = (2.0 * L.dmz * L.amz * (fexp - fexp * fexp) + L.fk * db) * lam / bl;
    for (int c = 0; c < 3; ++c) {
      d[i3 + c] = d[i3 + c] - dv * rij[c];
      d[j3 + c] = d[j3 + c] + dv * rij[c];
    }
    if (db > 0) {
      emorse[ib] = -(fexp * fexp - 2.0 * fexp);
      for (int c = 0; c < 3; ++c) {
        dmorse_i[3 * ib + c] = +2.0 * L.amz * (fexp - fexp * fexp) * lam / bl * rij[c];
        dmorse_j[3 * ib + c] = -2.0 * L.amz * (fexp - fexp * fexp) * lam / bl * rij[c];
      }
    } else {
      emorse[ib] = 1.0;
      for (int c = 0; c < 3; ++c) { dmorse_i[3 * ib + c] = 0.0; dmorse_j[3 * ib + c] = 0.0; }
    }
  }
}

// qangle (src/md.f90:17022-17114)
void Oracle::qangle(int st) {
  const double lam = s.lambda[st];
  for (size_t ia = 0; ia < s.qang.size(); ++ia) {
    const QTerm& a = s.qang[ia];
    const int ic = a.cod[st];
    if (ic <= 0) continue;
    const QAngLib& L = s.qanglib[ic - 1];
    const int ib = s.qang_cb[ia];
    double gamma = 1.0;
    if (ib >= 0) {
      gamma = emorse[ib];
      if (s.qang_cf[ia] == 1) gamma = 1.0 - gamma;
    }
    const int i3 = 3 * a.i, j3 = 3 * a.j, k3 = 3 * a.k;
    double rji[3], rjk[3];
    for (int c = 0; c < 3; ++c) { rji[c] = x[i3 + c] - x[j3 + c]; rjk[c] = x[k3 + c] - x[j3 + c]; }
    double bji = std::sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    double bjk = std::sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    double scp = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]);
    scp = scp / (bji * bjk);
    if (scp > 1.0) scp = 1.0;
    if (scp < -1.0) scp = -1.0;
    double ang = std::acos(scp);
    double da = ang - L.ang0;
    double ae = 0.5 * L.fk * (da * da);
    EQ[st].angle = EQ[st].angle + ae * gamma;
    double dv = gamma * L.fk * da * lam;
    double f1 = std::sin(ang);
    if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dk[3];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (rjk[c] / (bji * bjk) - scp * rji[c] / (bji * bji));
      dk[c] = f1 * (rji[c] / (bji * bjk) - scp * rjk[c] / (bjk * bjk));
    }
    for (int c = 0; c < 3; ++c) {
      d[i3 + c] = d[i3 + c] + dv * di[c];
      d[k3 + c] = d[k3 + c] + dv * dk[c];
      d[j3 + c] = d[j3 + c] - dv * (di[c] + dk[c]);
    }
    if (ib >= 0) {
      const int bi3 = 3 * s.qbnd[ib].i, bj3 = 3 * s.qbnd[ib].j;
      for (int c = 0; c < 3; ++c) d[bi3 + c] = d[bi3 + c] + dmorse_i[3 * ib + c] * ae;
      for (int c = 0; c < 3; ++c) d[bj3 + c] = d[bj3 + c] + dmorse_j[3 * ib + c] * ae;
    }
  }
}

// Dihedral geometry as qtorsion/qimproper form it: divisions by bj*bk and
// bj**2 (src/md.f90:17295-17338).
static double qdihedral(const std::vector<double>& x, int i, int j, int k, int l, double dp[12]) {
  const int i3 = 3 * i, j3 = 3 * j, k3 = 3 * k, l3 = 3 * l;
  double rji[3], rjk[3], rkl[3], rnj[3], rnk[3];
  for (int c = 0; c < 3; ++c) {
    rji[c] = x[i3 + c] - x[j3 + c];
    rjk[c] = x[k3 + c] - x[j3 + c];
    rkl[c] = x[l3 + c] - x[k3 + c];
  }
  rnj[0] = rji[1] * rjk[2] - rji[2] * rjk[1];
  rnj[1] = rji[2] * rjk[0] - rji[0] * rjk[2];
  rnj[2] = rji[0] * rjk[1] - rji[1] * rjk[0];
  rnk[0] = -rjk[1] * rkl[2] + rjk[2] * rkl[1];
  rnk[1] = -rjk[2] * rkl[0] + rjk[0] * rkl[2];
  rnk[2] = -rjk[0] * rkl[1] + rjk[1] * rkl[0];
  double bj = std::sqrt(rnj[0] * rnj[0] + rnj[1] * rnj[1] + rnj[2] * rnj[2]);
  double bk = std::sqrt(rnk[0] * rnk[0] + rnk[1] * rnk[1] + rnk[2] * rnk[2]);
  double scp = (rnj[0] * rnk[0] + rnj[1] * rnk[1] + rnj[2] * rnk[2]) / (bj * bk);
  if (scp > 1.0) scp = 1.0;
  if (scp < -1.0) scp = -1.0;
  double phi = std::acos(scp);
  double sgn = rjk[0] * (rnj[1] * rnk[2] - rnj[2] * rnk[1]) + rjk[1] * (rnj[2] * rnk[0] - rnj[0] * rnk[2]) +
               rjk[2] * (rnj[0] * rnk[1] - rnj[1] * rnk[0]);
  if (sgn < 0) phi = -phi;
  double f1 = std::sin(phi);
  if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
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

// qtorsion (src/md.f90:17395-17540)
void Oracle::qtorsion(int st) {
  const double lam = s.lambda[st];
  for (size_t ip = 0; ip < s.qtor.size(); ++ip) {
    const QTerm& t = s.qtor[ip];
    const int ic = t.cod[st];
    if (ic <= 0) continue;
    const QTorLib& L = s.qtorlib[ic - 1];
    const int ib = s.qtor_cb[ip];
    double gamma = 1.0;
    if (ib >= 0) {
      gamma = emorse[ib];
      if (s.qtor_cf[ip] == 1) gamma = 1 - gamma;
    }
    double dp[12];
    double phi = qdihedral(x, t.i, t.j, t.k, t.l, dp);
    double arg = L.rmult * phi - L.deltor;
    double pe = L.fk * (1.0 + std::cos(arg));
    EQ[st].torsion = EQ[st].torsion + pe * gamma;
    double dv = -L.rmult * L.fk * std::sin(arg) * gamma * lam;
    const int at[4] = {t.i, t.j, t.k, t.l};
    for (int m = 0; m < 4; ++m)
      for (int c = 0; c < 3; ++c) d[3 * at[m] + c] = d[3 * at[m] + c] + dv * dp[3 * m + c];
    if (ib >= 0) {
      const int bi3 = 3 * s.qbnd[ib].i, bj3 = 3 * s.qbnd[ib].j;
      for (int c = 0; c < 3; ++c) d[bi3 + c] = d[bi3 + c] + dmorse_i[3 * ib + c] * pe;
      for (int c = 0; c < 3; ++c) d[bj3 + c] = d[bj3 + c] + dmorse_j[3 * ib + c] * pe;
    }
  }
}

// qimproper (src/md.f90:17249-17392): harmonic in the wrapped angle
void Oracle::qimproper(int st) {
  const double lam = s.lambda[st];
  for (size_t ip = 0; ip < s.qimp.size(); ++ip) {
    const QTerm& t = s.qimp[ip];
    const int ic = t.cod[st];
    if (ic <= 0) continue;
    const QImpLib& L = s.qimplib[ic - 1];
    const int ib = s.qimp_cb[ip];
    double gamma = 1.0;
    if (ib >= 0) {
      gamma = emorse[ib];
      if (s.qimp_cf[ip] == 1) gamma = 1 - gamma;
    }
    double dp[12];
    double phi = qdihedral(x, t.i, t.j, t.k, t.l, dp);
    double arg = phi - L.imp0;
    arg = arg - 2.0 * kPi * std::round(arg / (2.0 * kPi));
    double dv = L.fk * arg;
    double pe = 0.5 * dv * arg;
    EQ[st].improper = EQ[st].improper + pe * gamma;
    dv = dv * gamma * lam;
    const int at[4] = {t.i, t.j, t.k, t.l};
    for (int m = 0; m < 4; ++m)
      for (int c = 0; c < 3; ++c) d[3 * at[m] + c] = d[3 * at[m] + c] + dv * dp[3 * m + c];
    if (ib >= 0) {
      const int bi3 = 3 * s.qbnd[ib].i, bj3 = 3 * s.qbnd[ib].j;
      for (int c = 0; c < 3; ++c) d[bi3 + c] = d[bi3 + c] + dmorse_i[3 * ib + c] * pe;
      for (int c = 0; c < 3; ++c) d[bj3 + c] = d[bj3 + c] + dmorse_j[3 * ib + c] * pe;
    }
  }
}

// p_restrain (src/md.f90:15473-15802), spherical boundary
void Oracle::p_restrain() {
  const double deg2rad = kPi / 180.0;
  auto add_state = [&](int istate, double edum) {
    if (istate == 0) {
      for (int k = 0; k < s.nstates; ++k) EQ[k].restraint = EQ[k].restraint + edum;
      if (s.nstates == 0) E.rst_protein = E.rst_protein + edum;
    } else {
      EQ[istate - 1].restraint = EQ[istate - 1].restraint + edum;
    }
  };
  auto wgt_of = [&](int istate) { return istate != 0 ? s.lambda[istate - 1] : 1.0; };
  for (const RstSeq& r : s.rstseq) {
    const double fk = r.fk;
    auto use = [&](int i) { return s.heavy[i] || (r.ih == 1 && !(s.excl[i] && s.freeze)); };
    if (r.to_centre == 1 || r.to_centre == 2) {
      double dr[3] = {0, 0, 0};
      double norm = 0;
      int n = 0;
      for (int i = r.i; i <= r.j; ++i) {
        if (!use(i)) continue;
        const double w = r.to_centre == 1 ? 1.0 : s.mass[i];
        if (r.to_centre == 1) {
          ++n;
          for (int c = 0; c < 3; ++c) dr[c] = dr[c] + x[3 * i + c] - s.xtop[3 * i + c];
        } else {
          norm = norm + w;
          for (int c = 0; c < 3; ++c) dr[c] = dr[c] + (x[3 * i + c] - s.xtop[3 * i + c]) * w;
        }
      }
      if (r.to_centre == 1 ? n > 0 : norm > 0) {
        for (int c = 0; c < 3; ++c) dr[c] = r.to_centre == 1 ? dr[c] / n : dr[c] / norm;
        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        E.rst_protein = E.rst_protein + 0.5 * fk * r2;
        for (int i = r.i; i <= r.j; ++i) {
          if (!use(i)) continue;
          for (int c = 0; c < 3; ++c)
            d[3 * i + c] = r.to_centre == 1 ? d[3 * i + c] + fk * dr[c] * s.mass[i] / 12.010
                                            : d[3 * i + c] + fk * dr[c];
        }
      }
    } else {
      for (int i = r.i; i <= r.j; ++i) {
        if (!use(i)) continue;
        const int i3 = 3 * i;
        double dr[3] = {x[i3] - s.xtop[i3], x[i3 + 1] - s.xtop[i3 + 1], x[i3 + 2] - s.xtop[i3 + 2]};
        // periodic box: the minimal periodic distance (src/md.f90:15561-15565)
        if (s.pbc) for (int c = 0; c < 3; ++c) dr[c] = mimage(dr[c], c);
        double r2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        E.rst_protein = E.rst_protein + 0.5 * fk * r2;
        for (int c = 0; c < 3; ++c) d[i3 + c] = d[i3 + c] + fk * dr[c];
      }
    }
  }
  for (const RstPos& r : s.rstpos) {
    const int i3 = 3 * r.i;
    double dr[3] = {x[i3] - r.x[0], x[i3 + 1] - r.x[1], x[i3 + 2] - r.x[2]};
    const double wgt = wgt_of(r.ipsi);
    double edum = 0.5 * r.fk[0] * (dr[0] * dr[0]) + 0.5 * r.fk[1] * (dr[1] * dr[1]) + 0.5 * r.fk[2] * (dr[2] * dr[2]);
    for (int c = 0; c < 3; ++c) d[i3 + c] = d[i3 + c] + r.fk[c] * dr[c] * wgt;
    add_state(r.ipsi, edum);
  }
  for (const RstDis& r : s.rstdis) {
    const int i3 = 3 * r.i, j3 = 3 * r.j;
    double dr[3] = {x[j3] - x[i3], x[j3 + 1] - x[i3 + 1], x[j3 + 2] - x[i3 + 2]};
    if (s.pbc) for (int c = 0; c < 3; ++c) dr[c] = mimage(dr[c], c);   // src/md.f90:15630-15634
    const double wgt = wgt_of(r.ipsi);
    double b = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    double db;
    if (b < r.d1) db = b - r.d1;
    else if (b > r.d2) db = b - r.d2;
    else continue;
    double edum = 0.5 * r.fk * (db * db);
    double dv = wgt * r.fk * db / b;
    for (int c = 0; c < 3; ++c) {
      d[j3 + c] = d[j3 + c] + dr[c] * dv;
      d[i3 + c] = d[i3 + c] - dr[c] * dv;
    }
    add_state(r.ipsi, edum);
  }
  for (const RstAng& r : s.rstang) {
    const int i3 = 3 * r.i, j3 = 3 * r.j, k3 = 3 * r.k;
    double dr[3], dr2[3];
    for (int c = 0; c < 3; ++c) { dr[c] = x[i3 + c] - x[j3 + c]; dr2[c] = x[k3 + c] - x[j3 + c]; }
    if (s.pbc)   // src/md.f90:15694-15702
      for (int c = 0; c < 3; ++c) { dr[c] = mimage(dr[c], c); dr2[c] = mimage(dr2[c], c); }
    const double wgt = wgt_of(r.ipsi);
    double r2ij = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    double r2jk = dr2[0] * dr2[0] + dr2[1] * dr2[1] + dr2[2] * dr2[2];
    double rij = std::sqrt(r2ij), rjk = std::sqrt(r2jk);
    double scp = (dr[0] * dr2[0] + dr[1] * dr2[1] + dr[2] * dr2[2]);
    scp = scp / (rij * rjk);
    if (scp > 1.0) scp = 1.0;
    if (scp < -1.0) scp = -1.0;
    double theta = std::acos(scp);
    double db = theta - (r.ang) * deg2rad;
    double edum = 0.5 * r.fk * (db * db);
    double dv = wgt * r.fk * db;
    double f1 = std::sin(theta);
    if (std::fabs(f1) < kQrealEps) f1 = kQrealEps;
    f1 = -1.0 / f1;
    double di[3], dk[3];
    for (int c = 0; c < 3; ++c) {
      di[c] = f1 * (dr2[c] / (rij * rjk) - scp * dr[c] / r2ij);
      dk[c] = f1 * (dr[c] / (rij * rjk) - scp * dr2[c] / r2jk);
    }
    for (int c = 0; c < 3; ++c) {
      d[i3 + c] = d[i3 + c] + dv * di[c];
      d[k3 + c] = d[k3 + c] + dv * dk[c];
      d[j3 + c] = d[j3 + c] - dv * (di[c] + dk[c]);
    }
    add_state(r.ipsi, edum);
  }
  // wall restraints only for spheres (src/md.f90:15768)
  for (const RstWal& r : s.rstwal) {
    if (s.pbc) break;
    const double fk = r.fk;
    for (int i = r.i; i <= r.j; ++i) {
      if (!(s.heavy[i] || r.ih == 1)) continue;
      const int i3 = 3 * i;
      double dr[3] = {x[i3] - s.top.xwcent[0], x[i3 + 1] - s.top.xwcent[1], x[i3 + 2] - s.top.xwcent[2]};
      double b = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double db = b - r.d;
      double erst, dv;
      if (db > 0) {
        erst = 0.5 * fk * (db * db) - r.dMorse;
        dv = fk * db / b;
      } else {
        double fexp = std::exp(r.aMorse * db);
        erst = r.dMorse * (fexp * fexp - 2.0 * fexp);
        dv = -2.0 * r.dMorse * r.aMorse * (fexp - fexp * fexp) / b;
      }
      E.rst_protein = E.rst_protein + erst;
      for (int c = 0; c < 3; ++c) d[i3 + c] = d[i3 + c] + dv * dr[c];
    }
  }
}

// offdiag (src/md.f90:15444-15470): H_ij = A exp(-mu r_kl) per function.
void Oracle::offdiag() {
  offd_hij.assign(s.offd.size(), 0.0);
  offd_rkl.assign(s.offd.size(), 0.0);
  for (size_t io = 0; io < s.offd.size(); ++io) {
    const OffDiag& o = s.offd[io];
    int k3 = 3 * s.iqseq[o.k], l3 = 3 * s.iqseq[o.l];
    double r = std::sqrt((x[l3] - x[k3]) * (x[l3] - x[k3]) + (x[l3 + 1] - x[k3 + 1]) * (x[l3 + 1] - x[k3 + 1]) +
                         (x[l3 + 2] - x[k3 + 2]) * (x[l3 + 2] - x[k3 + 2]));
    offd_hij[io] = o.A * std::exp(-o.mu * r);
    offd_rkl[io] = r;
  }
}

thread_local std::string g_err;

}  // namespace

// ------------------------------------------------------------------ C API
// nonbond_monitor (src/md.f90:11228-11350) at the current coordinates: per
// monitored pair Vwsum, Vwel, Vwlj, then (Vel, Vlj) of every state.  The
// reference's working variables (charges, type indices, r6) are carried
// from one state or partner to the next exactly as there.
static int oracle_monitor_eval(const Setup& s, const double* x, double* out, int cap) {
  const int ns = std::max(1, s.nstates);
  const int stride = 3 + 2 * ns;
  const int np = (int)s.mon.size();
  if (np * stride > cap) return -np * stride;
  const auto& T = s.top;
  for (int par = 0; par < np; ++par) {
    const auto& m = s.mon[par];
    const int ni = (int)m.ai.size(), nj = (int)m.aj.size();
    std::vector<double> Vel(ns, 0.0), Vlj(ns, 0.0);
    for (int i = 0; i < ni; ++i) {
      const int atomi = m.ai[i];
      const int qatomi = s.iqatom[atomi];
      double qi = s.crg[atomi];
      int iaci = s.iac[atomi];
      for (int j = 0; j < nj; ++j) {
        const int atomj = m.aj[j];
        const int qatomj = s.iqatom[atomj];
        int iacj = s.iac[atomj];
        double qj = s.crg[atomj];
        double dx1 = x[3 * atomi] - x[3 * atomj], dx2 = x[3 * atomi + 1] - x[3 * atomj + 1],
               dx3 = x[3 * atomi + 2] - x[3 * atomj + 2];
        if (s.pbc) {   // src/md.f90:11269-11273
          dx1 = dx1 - s.boxl[0] * std::round(dx1 * s.inv_boxl[0]);
          dx2 = dx2 - s.boxl[1] * std::round(dx2 * s.inv_boxl[1]);
          dx3 = dx3 - s.boxl[2] * std::round(dx3 * s.inv_boxl[2]);
        }
        const double r2 = dx1 * dx1 + dx2 * dx2 + dx3 * dx3;
        const double r = std::sqrt(1.0 / r2);
        double r6 = r2 * r2 * r2;
        const double r6_hc = r6;
        r6 = 1.0 / r6;
        int sc_1 = 0, sc_2 = 0;
        for (int st = 0; st < ns; ++st) {
          bool do_sc = false;
          const int lj = m.ljcod[((size_t)st * ni + i) * nj + j];
          // topology parameters of the (possibly carried-over) type indices;
          // with Q-atom types a Q atom's are replaced below
          auto tav = [&](int ty, int c) { return ty < (int)T.types.size() ? T.types[ty].avdw[c] : 0.0; };
          auto tbv = [&](int ty, int c) { return ty < (int)T.types.size() ? T.types[ty].bvdw[c] : 0.0; };
          double aLJi = tav(iaci, lj - 1), bLJi = tbv(iaci, lj - 1);
          double aLJj = tav(iacj, lj - 1), bLJj = tbv(iacj, lj - 1);
          if (qatomi) qi = s.qcrg[(qatomi - 1) * s.nstates + st];
          if (qatomj) qj = s.qcrg[(qatomj - 1) * s.nstates + st];
          if (s.qvdw_flag) {
            if (qatomi) {
              iaci = s.qiac[(qatomi - 1) * s.nstates + st];
              aLJi = s.qavdw[3 * iaci + lj - 1];
              bLJi = s.qbvdw[3 * iaci + lj - 1];
              do_sc = true;
              sc_1 = qatomi - 1;
            } else {
              sc_2 = iaci;
            }
            if (qatomj) {
              iacj = s.qiac[(qatomj - 1) * s.nstates + st];
              aLJj = s.qavdw[3 * iacj + lj - 1];
              bLJj = s.qbvdw[3 * iacj + lj - 1];
              if (do_sc) {
                sc_2 = qatomj - 1 + s.natyps;
              } else {
                do_sc = true;
                sc_1 = qatomj - 1;
              }
            } else {
              sc_2 = iacj;
            }
            if (do_sc) {
              r6 = r6_hc + s.sc_lookup[((size_t)sc_1 * (s.natyps + s.nqat) + sc_2) * s.nstates + st];
              r6 = 1.0 / r6;
            }
          }
          const double V = qi * qj * r;
          const double Vv = aLJi * aLJj * r6 * r6 - bLJi * bLJj * r6;
          Vel[st] = Vel[st] + V;
          Vlj[st] = Vlj[st] + Vv;
        }
      }
    }
    double wel = 0.0, wlj = 0.0;
    for (int st = 0; st < ns; ++st) {
      const double lam = st < (int)s.lambda.size() ? s.lambda[st] : 1.0;
      wel = wel + Vel[st] * lam;
      wlj = wlj + Vlj[st] * lam;
    }
    double* o = out + (size_t)par * stride;
    o[0] = wlj + wel; o[1] = wel; o[2] = wlj;
    for (int st = 0; st < ns; ++st) { o[3 + 2 * st] = Vel[st]; o[4 + 2 * st] = Vlj[st]; }
  }
  return np * stride;
}

extern "C" {

void* oracle_create(const char* inp) {
  try {
    return new Oracle(inp);
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
const char* oracle_last_error() { return g_err.c_str(); }
void oracle_destroy(void* h) { delete static_cast<Oracle*>(h); }
int oracle_natom(void* h) { return static_cast<Oracle*>(h)->s.natom; }
int oracle_nstates(void* h) { return static_cast<Oracle*>(h)->s.nstates; }
// pot_energy_nonbonds at coordinates x over the current lists: d += the
// gradient; E7 = pp el/vdW, pw el/vdW, ww el/vdW, LRF; EQ4 per state = qp
// el/vdW, qw el/vdW.  x becomes the oracle's coordinates.
int oracle_pot_energy_nonbonds(void* h, const double* x, double* d, double* E7, double* EQ4) {
  auto* o = static_cast<Oracle*>(h);
  try {
    std::copy(x, x + o->x.size(), o->x.begin());
    const double kinetic = o->E.kinetic;
    o->E = Energies{};
    o->E.kinetic = kinetic;
    for (auto& q : o->EQ) q = QEnergies{};
    std::vector<double> d0(o->d.size(), 0.0);
    std::swap(d0, o->d);
    o->pot_energy_nonbonds();
    for (size_t k = 0; k < o->d.size(); ++k) d[k] = d[k] + o->d[k];
    std::swap(d0, o->d);
    const Energies& E = o->E;
    const double e7[7] = {E.pp_el, E.pp_vdw, E.pw_el, E.pw_vdw, E.ww_el, E.ww_vdw, E.lrf};
    std::copy(e7, e7 + 7, E7);
    for (int st = 0; st < o->s.nstates; ++st) {
      const QEnergies& q = o->EQ[st];
      EQ4[4 * st] = q.qp_el; EQ4[4 * st + 1] = q.qp_vdw; EQ4[4 * st + 2] = q.qw_el; EQ4[4 * st + 3] = q.qw_vdw;
    }
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
  return 0;
}
// shake(xx, x) on caller arrays (src/md.f90:17581-17679): 0 with *niter =
// iterations summed over the molecules / number of molecules, 1 on a failure
int oracle_shake(void* h, const double* xx, double* x, int* niter) {
  auto* o = static_cast<Oracle*>(h);
  std::vector<double> sx(xx, xx + o->x.size()), cx(x, x + o->x.size());
  std::swap(sx, o->xx);
  std::swap(cx, o->x);
  const int n = o->shake();
  std::swap(sx, o->xx);
  std::swap(cx, o->x);
  std::copy(cx.begin(), cx.end(), x);
  if (n < 0) { g_err = "shake failure"; return 1; }
  *niter = o->s.nmol > 0 ? n / o->s.nmol : 0;
  return 0;
}
// the sizes write_out's layout depends on: waters, Q atoms, solute atoms
void oracle_sizes(void* h, int* out) {
  const auto& s = static_cast<Oracle*>(h)->s;
  out[0] = s.nwat; out[1] = s.nqat; out[2] = s.nat_solute;
}

// run n steps; returns 0 on success
int oracle_run(void* h, int n) {
  auto* o = static_cast<Oracle*>(h);
  try {
    for (int k = 0; k < n; ++k)
      if (!o->step()) { g_err = o->err; return 1; }
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
  return 0;
}
// pot_energy at the current coordinates over the current lists
int oracle_pot_energy(void* h) {
  auto* o = static_cast<Oracle*>(h);
  try {
    o->pot_energy();
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
  return 0;
}
// the final list build + energy of md_run (src/md.f90:6690-6700)
int oracle_final_energy(void* h) {
  auto* o = static_cast<Oracle*>(h);
  try {
    o->make_pair_lists();
    o->pot_energy();
  } catch (const std::exception& e) {
    g_err = e.what();
    return 2;
  }
  return 0;
}
void oracle_get_xv(void* h, double* x, double* v) {
  auto* o = static_cast<Oracle*>(h);
  std::memcpy(x, o->x.data(), 8 * o->x.size());
  std::memcpy(v, o->v.data(), 8 * o->v.size());
}
void oracle_get_d(void* h, double* d) {
  auto* o = static_cast<Oracle*>(h);
  std::memcpy(d, o->d.data(), 8 * o->d.size());
}
// energies in write_out order; returns the count written (<= cap)
int oracle_get_energies(void* h, double* e, int cap) {
  auto* o = static_cast<Oracle*>(h);
  const Energies& E = o->E;
  double v[] = {E.pp_el, E.pp_vdw, E.p_bond, E.p_angle, E.p_torsion, E.p_improper,
                E.ww_el, E.ww_vdw, E.w_bond, E.w_angle, E.w_torsion, E.w_improper,
                E.pw_el, E.pw_vdw, E.lrf,
                E.qx_el, E.qx_vdw, E.q_bond, E.q_angle, E.q_torsion, E.q_improper,
                E.rst_total, E.rst_fix, E.rst_radial, E.rst_wpol, E.rst_shell, E.rst_protein,
                E.potential + E.kinetic, E.potential, E.kinetic};
  int n = (int)(sizeof(v) / sizeof(v[0]));
  if (n > cap) n = cap;
  std::memcpy(e, v, 8 * n);
  return n;
}
// per-state Q energies: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw total bond
// angle torsion improper restraint (12 per state)
int oracle_get_qenergies(void* h, double* e, int cap) {
  auto* o = static_cast<Oracle*>(h);
  int k = 0;
  for (auto& q : o->EQ) {
    double v[] = {q.qq_el, q.qq_vdw, q.qp_el, q.qp_vdw, q.qw_el, q.qw_vdw, q.total,
                  q.bond, q.angle, q.torsion, q.improper, q.restraint};
    for (double z : v) if (k < cap) e[k++] = z;
  }
  return k;
}
// list sizes: nbpp nbpw nbww nbqp nbqw
void oracle_list_sizes(void* h, long* n) {
  auto* o = static_cast<Oracle*>(h);
  n[0] = (long)o->nbpp.size(); n[1] = (long)o->nbpw.size(); n[2] = (long)o->nbww.size();
  n[3] = (long)o->nbqp.size(); n[4] = (long)o->nbqw.size();
}
// copy list `which` (0 pp, 1 pw, 2 ww) as (i, j, ljcod) int triples
void oracle_get_list(void* h, int which, int* out) {
  auto* o = static_cast<Oracle*>(h);
  const std::vector<NB>& L = which == 0 ? o->nbpp : (which == 1 ? o->nbpw : o->nbww);
  for (size_t k = 0; k < L.size(); ++k) {
    out[3 * k] = L[k].i; out[3 * k + 1] = L[k].j; out[3 * k + 2] = L[k].ljcod;
  }
}
// LRF multipoles: 43 doubles per charge group (centre, phi0, phi1, phi2, phi3)
void oracle_get_lrf(void* h, double* out) {
  auto* o = static_cast<Oracle*>(h);
  for (size_t g = 0; g < o->lrf.size(); ++g) std::memcpy(out + 43 * g, &o->lrf[g], 43 * 8);
}
double oracle_temperature(void* h) { return static_cast<Oracle*>(h)->Temp; }
// EVB off-diagonal elements at the current coordinates: Hij and r_kl per
// [off_diagonals] line; returns the count.
int oracle_offdiag(void* h, double* hij, double* rkl, int cap) {
  Oracle* o = static_cast<Oracle*>(h);
  o->offdiag();
  int n = (int)o->offd_hij.size();
  for (int k = 0; k < n && k < cap; ++k) { hij[k] = o->offd_hij[k]; rkl[k] = o->offd_rkl[k]; }
  return n;
}

void oracle_theta_corr(void* h, double* out) {
  auto* o = static_cast<Oracle*>(h);
  for (size_t k = 0; k < o->s.wshell.size(); ++k) out[k] = o->s.wshell[k].theta_corr;
}
int oracle_nshells(void* h) { return (int)static_cast<Oracle*>(h)->s.wshell.size(); }
// periodic box: box lengths and centre (6 doubles); returns 1 for a box, 0 for a sphere
int oracle_box(void* h, double* out) {
  const Setup& s = static_cast<Oracle*>(h)->s;
  for (int k = 0; k < 3; ++k) { out[k] = s.boxl[k]; out[3 + k] = s.boxc[k]; }
  return s.pbc ? 1 : 0;
}
// volume moves: tries, accepted, then per try 0 (rejected), 1 (accepted,
// deltaW <= 0) or 2 (accepted by the random draw), as many as cap allows;
// returns the number of tries
int oracle_volume_moves(void* h, int* out, int cap) {
  const Oracle* o = static_cast<Oracle*>(h);
  if (cap > 0) out[0] = o->volume_try;
  if (cap > 1) out[1] = o->volume_acc;
  for (int k = 0; k < (int)o->volume_log.size() && k + 2 < cap; ++k) out[k + 2] = o->volume_log[k];
  return o->volume_try;
}
int oracle_monitor(void* h, double* out, int cap) {
  Oracle* o = static_cast<Oracle*>(h);
  return oracle_monitor_eval(o->s, o->x.data(), out, cap);
}

}  // extern "C"
