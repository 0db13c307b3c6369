// qdyn's initialisation sequence for a spherical-boundary run
// (src/qdyn.f90:108-132): topology conversions, restrained shell, restart or
// topology coordinates, FEP charges, solvent parameters and water shells,
// inverse masses, charge scaling, SHAKE molecules and degrees of freedom,
// Q-atom connectivity and Q-Q lists, topology shrinking, and Maxwell
// velocities with initial SHAKE.  The result is a Setup that both the CPU
// oracle and the GPU engine start from.
#include "qsys.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace qsys {

static const double kPi = 4.0 * std::atan(1.0);        // src/md.f90:527
static const double kDeg2Rad = kPi / 180.0;             // src/md.f90:528

static std::string dir_of(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}
static std::string join(const std::string& d, const std::string& f) {
  if (f.empty() || f[0] == '/') return f;
  return d + "/" + f;
}

// randm (src/md.f90:17543-17577)
double Randm::next(int& ig) {
  const int m = 100000000, m1 = 10000, mult = 31415821;
  if (!started) { started = true; irand = std::abs(ig) % m; }
  int irandh = irand / m1, irandl = irand % m1;
  int multh = mult / m1, multl = mult % m1;
  irand = ((irandh * multl + irandl * multh) % m1) * m1 + irandl * multl;
  irand = (irand + 1) % m;
  double r = (double)(irand / 10) * 10 / (double)m;
  if (r <= 0.0 || r > 1.0) r = 0.0;
  ig = irand;
  return r;
}

// shake (src/md.f90:17581-17679), used here for initial_shaking only
int shake_cpu(const Setup& s, const std::vector<double>& xx, std::vector<double>& x) {
  int total = 0;
  for (const ShakeMol& m : s.shake_mol) {
    int nc = (int)m.bond.size();
    std::vector<char> ready(nc, 0);
    int nits = 0;
    for (;;) {
      for (int ic = 0; ic < nc; ++ic) {
        if (ready[ic]) continue;
        int i = m.bond[ic].i, j = m.bond[ic].j;
        int i3 = 3 * i, j3 = 3 * j;
        double xij0 = x[i3] - x[j3], xij1 = x[i3+1] - x[j3+1], xij2_ = x[i3+2] - x[j3+2];
        double xij2 = xij0 * xij0 + xij1 * xij1 + xij2_ * xij2_;
        double diff = m.bond[ic].dist2 - xij2;
        if (std::fabs(diff) < kShakeTol * m.bond[ic].dist2) ready[ic] = 1;
        double xx0 = xx[i3] - xx[j3], xx1 = xx[i3+1] - xx[j3+1], xx2 = xx[i3+2] - xx[j3+2];
        double scp = xij0 * xx0 + xij1 * xx1 + xij2_ * xx2;
        double corr = diff / (2.0 * scp * (s.winv[i] + s.winv[j]));
        x[i3]   = x[i3]   + xx0 * corr * s.winv[i];
        x[i3+1] = x[i3+1] + xx1 * corr * s.winv[i];
        x[i3+2] = x[i3+2] + xx2 * corr * s.winv[i];
        x[j3]   = x[j3]   - xx0 * corr * s.winv[j];
        x[j3+1] = x[j3+1] - xx1 * corr * s.winv[j];
        x[j3+2] = x[j3+2] - xx2 * corr * s.winv[j];
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

// qatom_load_fep's completion of Q terms around breaking bonds
// (src/qatom.f90:957-1046, 1180-1265, 1323-1467, 1512-1676) and get_fep's
// removal of the redefined topology terms (src/md.f90:2937-3030).
static void q_bonded_setup(Setup& s, const FepData& f) {
  Topology& t = s.top;
  const int nst = s.nstates;
  s.qbondlib = f.qbondlib; s.qanglib = f.qanglib; s.qtorlib = f.qtorlib; s.qimplib = f.qimplib;
  s.qbnd = f.qbnd; s.qang = f.qang; s.qtor = f.qtor; s.qimp = f.qimp;
  for (const QTerm& b : s.qbnd)
    if (b.i >= s.nat_solute || b.j >= s.nat_solute) throw std::runtime_error("Q bonds must join solute atoms");
  // breaking bonds: the first adjacent pair of states where the bond is
  // present in one and absent in the other (the reference's second test,
  // (a .eqv. b .eqv. .false.), equals a .neqv. b)
  struct Break { int i, j, brk, set; };
  std::vector<Break> brk;
  for (const QTerm& b : s.qbnd)
    for (int st = 0; st + 1 < nst; ++st) {
      bool a0 = b.cod[st] > 0, a1 = b.cod[st + 1] > 0;
      if (a0 != a1) { brk.push_back({b.i, b.j, a0 ? st + 1 : st, a0 ? st : st + 1}); break; }
    }
  auto new_term = [&](int i, int j, int k, int l, int set, int code) {
    QTerm q{i, j, k, l, std::vector<int>(nst, 0)};
    q.cod[set] = code;
    return q;
  };
  // angles of the topology at a breaking bond that the FEP file does not redefine
  {
    std::vector<int> hits;
    std::vector<const Break*> hb;
    for (const Break& b : brk)
      for (size_t a = 0; a < t.ang.size(); ++a) {
        const Angle& g = t.ang[a];
        if (((g.i == b.i && g.j == b.j) || (g.j == b.i && g.k == b.j) || (g.k == b.i && g.j == b.j) ||
             (g.j == b.i && g.i == b.j)) && g.cod != 0) { hits.push_back((int)a); hb.push_back(&b); }
      }
    for (size_t h = 0; h < hits.size(); ++h) {
      const Angle& g = t.ang[hits[h]];
      bool set = false;
      for (const QTerm& q : s.qang)
        if ((g.i == q.i && g.j == q.j && g.k == q.k) || (g.i == q.k && g.j == q.j && g.k == q.i)) set = true;
      if (set) continue;
      const AngLib& L = t.anglib[g.cod - 1];
      if (L.ureyfk != 0.0) throw std::runtime_error("Urey-Bradley Q-angle terms (CHARMM) are not supported");
      s.qanglib.push_back({L.fk, L.ang0, L.ureyfk, L.ureyr0});
      s.qang.push_back(new_term(g.i, g.j, g.k, 0, hb[h]->set, (int)s.qanglib.size()));
    }
  }
  // torsions
  {
    std::vector<int> hits;
    std::vector<const Break*> hb;
    for (const Break& b : brk)
      for (size_t a = 0; a < t.tor.size(); ++a) {
        const Torsion& g = t.tor[a];
        if (((g.i == b.i && g.j == b.j) || (g.j == b.i && g.k == b.j) || (g.k == b.i && g.l == b.j) ||
             (g.l == b.i && g.k == b.j) || (g.k == b.i && g.j == b.j) || (g.j == b.i && g.i == b.j)) &&
            g.cod != 0) { hits.push_back((int)a); hb.push_back(&b); }
      }
    for (size_t h = 0; h < hits.size(); ++h) {
      const Torsion& g = t.tor[hits[h]];
      bool set = false;
      for (const QTerm& q : s.qtor)
        if ((g.i == q.i && g.j == q.j && g.k == q.k && g.l == q.l) ||
            (g.i == q.l && g.j == q.k && g.k == q.j && g.l == q.i)) set = true;
      if (set) continue;
      const TorLib& L = t.torlib[g.cod - 1];
      s.qtorlib.push_back({L.fk, L.rmult, L.deltor});
      s.qtor.push_back(new_term(g.i, g.j, g.k, g.l, hb[h]->set, (int)s.qtorlib.size()));
    }
  }
  // impropers; the reference compares against the first nqtor Q impropers
  // (src/qatom.f90:1591), which this set-up repeats where it stays in range
  {
    std::vector<int> hits;
    std::vector<const Break*> hb;
    for (const Break& b : brk)
      for (size_t a = 0; a < t.imp.size(); ++a) {
        const Torsion& g = t.imp[a];
        if (((g.i == b.i && g.j == b.j) || (g.i == b.i && g.j == b.i) || (g.j == b.i && g.k == b.j) ||
             (g.k == b.i && g.j == b.j) || (g.j == b.i && g.l == b.j) || (g.l == b.i && g.j == b.j)) &&
            g.cod != 0) { hits.push_back((int)a); hb.push_back(&b); }
      }
    const int nqtor = (int)s.qtor.size();
    for (size_t h = 0; h < hits.size(); ++h) {
      const Torsion& g = t.imp[hits[h]];
      if (nqtor > (int)s.qimp.size())
        throw std::runtime_error("impropers at a breaking bond with fewer Q impropers than Q torsions are not supported");
      bool set = false;
      for (int j = 0; j < nqtor; ++j) {
        const QTerm& q = s.qimp[j];
        if ((g.i == q.i && g.j == q.j && g.k == q.k && g.l == q.l) ||
            (g.i == q.k && g.j == q.j && g.k == q.i && g.l == q.l) ||
            (g.i == q.k && g.j == q.j && g.k == q.l && g.l == q.i) ||
            (g.i == q.l && g.j == q.j && g.k == q.i && g.l == q.k) ||
            (g.i == q.l && g.j == q.j && g.k == q.k && g.l == q.i)) set = true;
      }
      if (set) continue;
      const ImpLib& L = t.implib[g.cod - 1];
      s.qimplib.push_back({L.fk, L.imp0});
      s.qimp.push_back(new_term(g.i, g.j, g.k, g.l, hb[h]->set, (int)s.qimplib.size()));
    }
  }
  // couplings: an angle takes its first listed coupling, torsions and
  // impropers their last (src/md.f90:17042-17052, 17264-17273, 17420-17429)
  auto coupl = [&](const std::vector<QCoupling>& cl, int nterm, bool first, std::vector<int>& cb, std::vector<int>& cf) {
    cb.assign(nterm, -1); cf.assign(nterm, 0);
    for (const QCoupling& c : cl) {
      if (c.term >= nterm) throw std::runtime_error("Q-term coupling refers to a missing term");
      if (first && cb[c.term] >= 0) continue;
      cb[c.term] = c.bond; cf[c.term] = c.flag;
    }
  };
  coupl(f.ang_coupl, (int)s.qang.size(), true, s.qang_cb, s.qang_cf);
  coupl(f.tor_coupl, (int)s.qtor.size(), false, s.qtor_cb, s.qtor_cf);
  coupl(f.imp_coupl, (int)s.qimp.size(), false, s.qimp_cb, s.qimp_cf);
  // remove the redefined interactions from the topology
  for (Bond& b : t.bnd)
    for (const QTerm& q : s.qbnd)
      if ((b.i == q.i && b.j == q.j) || (b.i == q.j && b.j == q.i)) b.cod = 0;
  for (Angle& a : t.ang)
    for (const QTerm& q : s.qang)
      if ((a.i == q.i && a.j == q.j && a.k == q.k) || (a.i == q.k && a.j == q.j && a.k == q.i)) a.cod = 0;
  for (Torsion& a : t.tor)
    for (const QTerm& q : s.qtor)
      if (((a.i == q.i && a.j == q.j && a.k == q.k && a.l == q.l) ||
           (a.i == q.l && a.j == q.k && a.k == q.j && a.l == q.i)) && a.cod != 0) a.cod = 0;
  for (Torsion& a : t.imp)
    for (const QTerm& q : s.qimp)
      if ((a.i == q.i && a.j == q.j && a.k == q.k && a.l == q.l) ||
          (a.i == q.k && a.j == q.j && a.k == q.i && a.l == q.l) ||
          (a.i == q.k && a.j == q.j && a.k == q.l && a.l == q.i) ||
          (a.i == q.l && a.j == q.j && a.k == q.i && a.l == q.k) ||
          (a.i == q.l && a.j == q.j && a.k == q.k && a.l == q.i)) a.cod = 0;
}

// Extra restraints: atom numbers resolved and checked (src/md.f90:16431-16446).
static void restraints_setup(Setup& s) {
  const MDInput& in = s.inp;
  const Topology& t = s.top;
  s.rstseq = in.rstseq; s.rstpos = in.rstpos; s.rstdis = in.rstdis; s.rstang = in.rstang; s.rstwal = in.rstwal;
  auto resolve = [&](const std::string& txt) {
    size_t c = txt.find(':');
    if (c == std::string::npos) return (int)std::strtol(txt.c_str(), nullptr, 10) - 1;
    int res = (int)std::strtol(txt.substr(0, c).c_str(), nullptr, 10);
    int at = (int)std::strtol(txt.substr(c + 1).c_str(), nullptr, 10);
    const int nres = (int)t.res_start.size();
    if (res < 1 || res > nres) return -1;
    int first = t.res_start[res - 1], next = res < nres ? t.res_start[res] : s.natom;
    if (at < 1 || at > next - first) return -1;
    return first + at - 1;
  };
  for (size_t k = 0; k < s.rstdis.size(); ++k)
    if (!in.rstdis_text[k].first.empty()) {
      s.rstdis[k].i = resolve(in.rstdis_text[k].first);
      s.rstdis[k].j = resolve(in.rstdis_text[k].second);
    }
  auto ok = [&](int a) { return a >= 0 && a < s.natom; };
  auto okst = [&](int p) { return p >= 0 && p <= s.nstates; };
  for (auto& r : s.rstseq)
    if (!ok(r.i) || !ok(r.j)) throw std::runtime_error("invalid atom in [sequence_restraints]");
  for (auto& r : s.rstpos)
    if (!ok(r.i) || !okst(r.ipsi)) throw std::runtime_error("invalid atom or state in [atom_restraints]");
  for (auto& r : s.rstdis)
    if (!ok(r.i) || !ok(r.j) || !okst(r.ipsi)) throw std::runtime_error("invalid atom or state in [distance_restraints]");
  for (auto& r : s.rstang)
    if (!ok(r.i) || !ok(r.j) || !ok(r.k) || !okst(r.ipsi)) throw std::runtime_error("invalid atom or state in [angle_restraints]");
  for (auto& r : s.rstwal)
    if (!ok(r.i) || !ok(r.j)) throw std::runtime_error("invalid atom in [wall_restraints]");
}

// LJ codes of the monitored atom pairs per state (nbmonitorlist,
// src/md.f90:11151-11226): the topology code; code 1 instead of 2 between a
// Q and a non-Q atom with Q-atom types; code 3 for 1-4 pairs of two non-Q
// atoms, looked up as the reference does (list14 row of the larger number
// for a near pair, the long 1-4 list otherwise); for a pair with a Q atom the
// code of the first matching pair of the state's Q-Q list.
static void monitor_setup(Setup& s) {
  const auto& t = s.top;
  const int ns = std::max(1, s.nstates);
  auto is14 = [&](int i, int j) {   // 0-based atoms, i != j
    const int d = i > j ? i - j : j - i;
    if (d <= kMaxNbrRange) {
      const int row = i > j ? i : j;   // the reference's list14(|i-j|, max(i, j))
      if (row >= (int)t.list14.size()) return false;
      return ((t.list14[row] >> (d - 1)) & 1u) != 0;
    }
    for (auto& p : t.list14long)
      if ((p.first == i && p.second == j) || (p.first == j && p.second == i)) return true;
    return false;
  };
  for (auto& m : s.mon) {
    const int ni = (int)m.ai.size(), nj = (int)m.aj.size();
    m.ljcod.assign((size_t)ns * ni * nj, 1);
    for (int i = 0; i < ni; ++i)
      for (int j = 0; j < nj; ++j) {
        const int a = m.ai[i], b = m.aj[j];
        if (a == b) throw std::runtime_error("two paired monitor atom groups contain the same atom");
        for (int st = 0; st < ns; ++st) {
          int lj = s.ljcod[s.iac[a] * s.natyps + s.iac[b]];
          const bool qa = s.iqatom[a] != 0, qb = s.iqatom[b] != 0;
          if (s.qvdw_flag && lj == 2 && qa != qb) lj = 1;
          if (!qa && !qb) {
            if (is14(a, b)) lj = 3;
          } else if (st < (int)s.nbqq.size()) {
            for (auto& p : s.nbqq[st]) {
              const int qi = s.iqseq[p.iq];
              if ((qi == a && p.j == b) || (p.j == a && qi == b)) { lj = p.ljcod; break; }
            }
          }
          m.ljcod[((size_t)st * ni + i) * nj + j] = lj;
        }
      }
  }
}

Setup prepare(const std::string& input_path) {
  Setup s;
  const std::string dir = dir_of(input_path);
  s.inp = read_input(input_path);
  MDInput& in = s.inp;
  s.top = read_topology(join(dir, in.top_file));
  Topology& t = s.top;
  if (t.solvent_type != 0 && t.solvent_type != 1)
    throw std::runtime_error("only SPC (solvent_type 0) and 3-atom all-LJ (1) solvents are supported");
  if (t.ivdw_rule != 1) throw std::runtime_error("only the geometric vdW rule is supported");
  if (t.iuse_switch_atom != 1) throw std::runtime_error("only switching-atom charge-group cut-offs are supported");
  s.natom = t.nat_pro;
  s.nat_solute = t.nat_solute;
  s.solv_atom = t.solv_atom;
  if (s.natom > s.nat_solute && s.solv_atom != 3) throw std::runtime_error("solvent must be 3-atom water");
  s.nwat = (s.natom - s.nat_solute) / s.solv_atom;
  s.nmol = (int)t.istart_mol.size();
  t.istart_mol.push_back(s.natom);
  // periodic box (src/md.f90:17869-17905): the same boundary in topology and
  // input, box from the topology (or the restart below), Q-atom cut-off
  // default -1 (every atom)
  s.pbc = t.use_pbc;
  if (s.pbc != in.box) throw std::runtime_error("Must have same boundary (sphere or box) in topology and input file");
  if (s.pbc) {
    for (int k = 0; k < 3; ++k) { s.boxl[k] = t.boxlength[k]; s.boxc[k] = t.boxcentre[k]; }
    if (!in.have_rcq) in.rcq = -1.0;
    if (s.nwat > 0 && t.solvent_type != 0) throw std::runtime_error("periodic boxes need SPC solvent (solvent_type 0)");
  }

  // topology(): unit conversions and LJ code table (src/md.f90:17819-17905)
  for (auto& a : t.anglib) a.ang0 = kDeg2Rad * a.ang0;
  for (auto& q : t.torlib) { q.paths = 1.0 / q.paths; q.deltor = kDeg2Rad * q.deltor; }
  for (auto& q : t.implib) q.imp0 = kDeg2Rad * q.imp0;
  s.natyps = (int)t.types.size();
  s.ljcod.assign((size_t)s.natyps * s.natyps, 1);
  for (auto& p : t.lj2) {
    s.ljcod[p.first * s.natyps + p.second] = 2;
    s.ljcod[p.second * s.natyps + p.first] = 2;
  }
  s.iac = t.iac;
  s.heavy.resize(s.natom);
  for (int i = 0; i < s.natom; ++i) s.heavy[i] = !(t.types[s.iac[i]].mass < 4.0);
  s.excl = t.excl;
  s.dt = 0.020462 * in.stepsize;
  s.tau_t = 0.020462 * in.tau_t;
  if (in.temp0 <= 0) throw std::runtime_error("no dynamics at zero temperature");
  if (s.tau_t < s.dt) throw std::runtime_error("bath coupling time shorter than the step");

  // prep_coord: topology coordinates for restraining, optionally refreshed
  // from an external restart-format file (src/md.f90:16380-16425), then the
  // restrained shell on them (src/md.f90:16595-16622)
  s.xtop = t.xtop;
  if (!in.restraint_file.empty()) {
    std::vector<double> xr, vr;
    read_restart(join(dir, in.restraint_file), s.natom, xr, vr, nullptr);
    s.xtop = xr;
  }
  s.mass.resize(s.natom);
  for (int i = 0; i < s.natom; ++i) s.mass[i] = t.types[s.iac[i]].mass;
  s.freeze = in.freeze;
  s.rexcl_i = in.have_shell_radius ? in.shell_radius : 0.85;
  if (s.rexcl_i < 0 && !s.pbc) throw std::runtime_error("inner radius of restrained shell must be >= 0");
  s.shell.assign(s.natom, 0);
  if (s.rexcl_i > t.rexcl_o && !s.pbc)
    throw std::runtime_error("inner radius of restrained shell must be < exclusion radius");
  if (s.rexcl_i <= 1.0) s.rexcl_i = s.rexcl_i * t.rexcl_o;
  if (!s.pbc) {
    double rin2 = s.rexcl_i * s.rexcl_i;
    for (int ig = 0; ig < t.ncgp_solute; ++ig) {
      int sw = t.cgp[ig].iswitch;
      if (!s.excl[sw] && s.heavy[sw]) {
        int i3 = 3 * sw;
        double r2 = (s.xtop[i3] - t.xpcent[0]) * (s.xtop[i3] - t.xpcent[0])
                  + (s.xtop[i3+1] - t.xpcent[1]) * (s.xtop[i3+1] - t.xpcent[1])
                  + (s.xtop[i3+2] - t.xpcent[2]) * (s.xtop[i3+2] - t.xpcent[2]);
        if (r2 > rin2)
          for (int a = t.cgp[ig].first; a <= t.cgp[ig].last; ++a) s.shell[t.cgpatom[a]] = 1;
      }
    }
  }
  std::vector<double> theta_corr_rst;
  bool restart = !in.restart_file.empty();
  if (restart) {
    double box[6];
    read_restart(join(dir, in.restart_file), s.natom, s.x, s.v, &theta_corr_rst, s.pbc ? box : nullptr);
    if (s.pbc)
      for (int k = 0; k < 3; ++k) { s.boxl[k] = box[k]; s.boxc[k] = box[3 + k]; }
  } else {
    s.x = s.xtop;
    s.v.assign(3 * s.natom, 0.0);
  }

  s.ff_type = t.ff_type;
  if (s.ff_type == 3) throw std::runtime_error("CHARMM force fields (Urey-Bradley terms) are not supported");
  // get_fep: Q atoms and their per-state charges (src/md.f90:2884-2990)
  s.iqatom.assign(s.natom, 0);
  s.crg = t.crg;
  std::vector<std::pair<int, int>> s_soft_pairs;
  std::vector<FepData::ElScale> s_el_scale;
  std::vector<FepData::ExSpec> s_exspec;
  std::vector<FepData::QShake> s_qshake;
  if (!in.fep_file.empty()) {
    FepData f = read_fep(join(dir, in.fep_file), s.natom);
    if (s.pbc) {
      if (f.qswitch < 0) throw std::runtime_error("Section PBC (switching_atom) is required when using periodic boundary");
      s.qswitch = f.qswitch;
      // Rcq < 0 (every atom, the lists built once) leaves the reference's
      // Q-solute groups without a shift record (src/md.f90:10359-10400)
      if (in.rcq < 0 && !f.iqseq.empty())
        throw std::runtime_error("a periodic box with Q atoms needs a positive q_atom cut-off");
    }
    s.nstates = f.nstates;
    for (auto& pr : f.mon_pairs) {
      Setup::MonPair m;
      m.ai = f.mon_groups[pr.first];
      m.aj = f.mon_groups[pr.second];
      s.mon.push_back(m);
    }
    if (!in.lambdas.empty()) {
      if ((int)in.lambdas.size() != s.nstates)
        throw std::runtime_error("number of lambdas differs from FEP states");
      s.lambda = in.lambdas;
    } else {
      if (s.nstates != 1) throw std::runtime_error("[lambdas] required for more than one state");
      s.lambda = {1.0};
    }
    s.nqat = (int)f.iqseq.size();
    s.iqseq = f.iqseq;
    for (int iq = 0; iq < s.nqat; ++iq) {
      if (s.iqseq[iq] >= s.nat_solute) throw std::runtime_error("Q atoms must be solute atoms");
      s.iqatom[s.iqseq[iq]] = iq + 1;
    }
    s.qcrg.assign((size_t)s.nqat * s.nstates, 0.0);
    for (int iq = 0; iq < s.nqat; ++iq)
      for (int st = 0; st < s.nstates; ++st)
        s.qcrg[iq * s.nstates + st] = f.have_charges ? f.qcrg[iq][st] : t.crg[s.iqseq[iq]];
    // Q-atom types (qvdw_flag) and the soft-core table
    s.qvdw_flag = f.qvdw;
    s.offd = f.offd;
    if (f.qvdw) {
      if (t.ivdw_rule != 1) throw std::runtime_error("Q-atom types need the geometric combination rule");
      s.qiac.assign((size_t)s.nqat * s.nstates, 0);
      for (int iq = 0; iq < s.nqat; ++iq)
        for (int st = 0; st < s.nstates; ++st) s.qiac[iq * s.nstates + st] = f.qiac[iq][st];
      for (size_t k = 0; k < f.qtac.size(); ++k)
        for (int c = 0; c < 3; ++c) { s.qavdw.push_back(f.qavdw[k][c]); s.qbvdw.push_back(f.qbvdw[k][c]); }
    }
    const int nt = (int)t.types.size(), nq = s.nqat, nst = s.nstates, ncol = nt + nq;
    s.sc_lookup.assign((size_t)nq * ncol * nst, 0.0);
    if (f.softcore) {
      // src/qatom.f90:1898-1970: explicit alphas, or alphas from the
      // desired maximum vdW potential (geometric rule)
      auto sc = [&](int i, int col, int st) -> double& { return s.sc_lookup[((size_t)i * ncol + col) * nst + st]; };
      auto qa = [&](int i, int st) { return f.qavdw[f.qiac[i][st]][0]; };
      auto qb = [&](int i, int st) { return f.qbvdw[f.qiac[i][st]][0]; };
      for (int i = 0; i < nq; ++i)
        for (int st = 0; st < nst; ++st) {
          const double am = f.alpha_max[i][st];
          for (int j = 0; j < nt; ++j) {
            if (f.softcore_max_potential) {
              if (am != 0) {
                const double aq = qa(i, st), bq = qb(i, st), aj = t.types[j].avdw[0], bj = t.types[j].bvdw[0];
                sc(i, j, st) = (-bq * bj + std::sqrt(bq * bq * bj * bj + 4 * am * aq * aj)) / (2 * am);
              }
            } else {
              sc(i, j, st) = am;
            }
          }
          for (int j = 0; j < nq; ++j) {
            const double aj_max = f.alpha_max[j][st];
            if (am > 1e-6 || aj_max > 1e-6) {
              double at;
              if (f.softcore_max_potential) {
                at = std::min(am, aj_max);
                if (am == 0 || aj_max == 0) at = std::max(am, aj_max);
                const double aq = qa(i, st), bq = qb(i, st), aj = qa(j, st), bj = qb(j, st);
                sc(i, nt + j, st) = (-bq * bj + std::sqrt(bq * bq * bj * bj + 4 * at * aq * aj)) / (2 * at);
              } else {
                sc(i, nt + j, st) = std::max(am, aj_max);
              }
            }
          }
        }
    }
    q_bonded_setup(s, f);
    s.qq_lib_charges = f.qq_lib_charges;
    // special exclusions involving non-Q atoms: all states or none, then
    // into the topology's exclusion lists (src/md.f90:3039-3066)
    for (const auto& e : f.exspec) {
      if (e.i < 0 || e.i >= s.natom || e.j < 0 || e.j >= s.natom) throw std::runtime_error("invalid special exclusion data");
      if (s.iqatom[e.i] == 0 || s.iqatom[e.j] == 0) {
        bool any = false, all = true;
        for (char fl : e.flag) { any = any || fl; all = all && fl; }
        if (!any) continue;
        if (!all) throw std::runtime_error("Non-Q-atom special exclusion pairs must be on in all or no states");
        if (e.i >= s.nat_solute || e.j >= s.nat_solute) throw std::runtime_error("special exclusions must join solute atoms");
        const int lo = std::min(e.i, e.j), hi = std::max(e.i, e.j);
        if (hi - lo <= kMaxNbrRange) t.listex[lo] |= 1u << (hi - lo - 1);
        else t.listexlong.push_back({e.i, e.j});
      }
    }
    s_soft_pairs = f.soft_pairs; s_el_scale = f.el_scale; s_exspec = f.exspec; s_qshake = f.qshake;
  } else if (!in.lambdas.empty()) {
    throw std::runtime_error("[lambdas] given without an FEP file");
  }
  restraints_setup(s);

  // prep_sim: solvent parameters from the first water (src/md.f90:16704-16713)
  if (s.nwat > 0) {
    for (int iw = 0; iw < 3; ++iw) {
      int a = s.nat_solute + iw;
      s.chg_solv[iw] = s.crg[a];
      for (int k = 0; k < 2; ++k) {
        s.alj_solv[iw][k] = t.types[s.iac[a]].avdw[k];
        s.blj_solv[iw][k] = t.types[s.iac[a]].bvdw[k];
      }
    }
  }
  // wat_sphere (src/md.f90:18106-18185)
  s.rwat = t.rwat;
  if (in.rwat_in > 0) s.rwat = in.rwat_in;
  {
    double crgtot = 0, crgexcl = 0;
    for (int i = 0; i < s.nat_solute; ++i) {
      if (!s.excl[i]) { if (s.iqatom[i] == 0) crgtot = crgtot + s.crg[i]; }
      else crgexcl = crgexcl + s.crg[i];
    }
    double crgqtot = 0;
    for (int iq = 0; iq < s.nqat; ++iq)
      for (int st = 0; st < s.nstates; ++st) {
        crgtot = crgtot + s.qcrg[iq * s.nstates + st] * s.lambda[st];
        crgqtot = crgqtot + s.qcrg[iq * s.nstates + st] * s.lambda[st];
      }
    if (!in.wpol_born) crgqtot = 0;
    s.crgtot = crgtot;
    s.crgqtot = crgqtot;
  }
  s.fk_wsphere = in.fk_wsphere == -1 ? 60.0 : in.fk_wsphere;
  s.fkwpol = in.fkwpol == -1 ? 20.0 : in.fkwpol;
  s.dwmz = in.dwmz == -1 ? 0.26 * std::exp(-0.19 * (s.rwat - 15.0)) + 0.74 : in.dwmz;
  s.awmz = in.awmz == -1 ? 0.2 / (1.0 + std::exp(0.4 * (s.rwat - 25.0))) + 0.3 : in.awmz;
  s.wpol_restr = in.wpol_restr && !s.pbc;   // no solvent restraints in a periodic box
  // wat_shells (src/md.f90:18189-18368)
  if (s.nwat > 0 && s.wpol_restr) {
    int bndcodw = t.bnd.back().cod, angcodw = t.ang.back().cod;
    s.mu_w = -s.chg_solv[0] * t.bondlib[bndcodw - 1].b0 * std::cos(t.anglib[angcodw - 1].ang0 / 2);
    double drs = kWpolrLayer / kDrout;
    int nsh = (int)(-0.5 + std::sqrt(2 * drs + 0.25));
    s.wshell.resize(nsh);
    bool tc_ok = restart && (int)theta_corr_rst.size() == nsh;
    double rout = s.rwat;
    int nmax = 0;
    for (int is = 0; is < nsh; ++is) {
      WaterShell& w = s.wshell[is];
      w.avtheta = 0; w.avn_insh = 0; w.n_insh = 0;
      w.theta_corr = tc_ok ? theta_corr_rst[is] : 0.0;
      w.rout = rout;
      double dr = kDrout * (is + 1);
      double ri = rout - dr;
      w.dr = dr;
      double vshell = rout * rout * rout - ri * ri * ri;
      int n_insh = (int)(4 * kPi / 3 * vshell * kRhoWat);
      if (n_insh > nmax) nmax = n_insh;
      double rshell = std::pow(0.5 * (rout * rout * rout + ri * ri * ri), 1.0 / 3.0);
      w.cstb = s.crgqtot * 0.98750 / (kRhoWat * s.mu_w * 4.0 * kPi * rshell * rshell);
      rout = rout - dr;
    }
    s.n_max_insh = (int)(nmax * 1.5);
  }
  if (s.pbc) {
    // inverse box lengths and the cut-off checks (src/md.f90:17873-17905)
    for (int k = 0; k < 3; ++k) s.inv_boxl[k] = s.boxl[k] == 0.0 ? 0.0 : 1.0 / s.boxl[k];
    if (s.boxl[0] == 0.0 || s.boxl[1] == 0.0 || s.boxl[2] == 0.0)
      for (int k = 0; k < 3; ++k) s.inv_boxl[k] = 0.0;
    const double bmin = std::min(s.boxl[0], std::min(s.boxl[1], s.boxl[2]));
    if (!(bmin > in.rcpp * 2)) throw std::runtime_error("Solute-solute cut-off radii too large");
    if (!(bmin > in.rcww * 2)) throw std::runtime_error("Solvent-solvent cut-off radii too large");
    if (!(bmin > in.rcpw * 2)) throw std::runtime_error("Solute-solvent cut-off radii too large");
    if (!(bmin > in.rcq * 2)) throw std::runtime_error("Q-atom cut-off radii too large");
    if (in.use_lrf && !(bmin > in.rclrf * 2)) throw std::runtime_error("LRF cut-off radii too large");
    // per molecule 1/mass (prep_sim, src/md.f90:16769-16784) and the solute
    // molecules (src/md.f90:16758-16759)
    s.mol_minv.assign(s.nmol, 0.0);
    for (int m = 0; m < s.nmol; ++m) {
      double ms = 0.0;
      for (int j = t.istart_mol[m]; j < t.istart_mol[m + 1]; ++j) ms = ms + t.types[s.iac[j]].mass;
      s.mol_minv[m] = 1.0 / ms;
    }
    const int nres = (int)t.res_start.size();
    s.nsolute_mol = s.nmol - (nres - t.nres_solute);
  }
  s.iwhich_cgp.assign(s.natom, 0);
  for (int ig = 0; ig < (int)t.cgp.size(); ++ig)
    for (int a = t.cgp[ig].first; a <= t.cgp[ig].last; ++a) s.iwhich_cgp[t.cgpatom[a]] = ig;
  s.winv.resize(s.natom);
  for (int i = 0; i < s.natom; ++i) s.winv[i] = 1.0 / t.types[s.iac[i]].mass;
  const double sq = std::sqrt(t.coulomb_constant);
  for (auto& c : s.crg) c = c * sq;
  if (s.nwat > 0) for (auto& c : s.chg_solv) c = c * sq;
  for (auto& c : s.qcrg) c = c * sq;

  // init_shake (src/md.f90:4242-4426)
  std::vector<Bond> bnd = t.bnd;
  std::vector<Angle> ang = t.ang;
  {
    const auto& ist = t.istart_mol;
    std::vector<ShakeMol> mols(s.nmol);
    double exclshk = 0;
    int mol = -1;
    for (auto& b : bnd) {
      int ia = b.i, ja = b.j;
      while (ia >= ist[mol + 1]) ++mol;
      if (b.cod == 0) continue;
      bool hyd = in.shake_hydrogens && (!s.heavy[ia] || !s.heavy[ja]);
      if (hyd || (in.shake_solute && ia < s.nat_solute) || (in.shake_solvent && ia >= s.nat_solute)) {
        if (s.excl[ia]) exclshk += 0.5;
        if (s.excl[ja]) exclshk += 0.5;
        mols[mol].bond.push_back({ia, ja, t.bondlib[b.cod - 1].b0 * t.bondlib[b.cod - 1].b0});
        b.cod = -1;
      }
    }
    // extra constraints of the FEP file, at the lambda-weighted distance:
    // a pair already constrained is overwritten in place, otherwise appended
    // to the molecule of its first atom (src/md.f90:4328-4354)
    for (const auto& q : s_qshake) {
      int m = 0;
      while (q.i >= ist[m + 1]) ++m;
      double d = 0.0;
      for (int st = 0; st < s.nstates; ++st) d = d + s.lambda[st] * q.dist[st];
      ShakeBond nb{q.i, q.j, d * d};
      bool found = false;
      for (auto& c : mols[m].bond)
        if ((c.i == q.i && c.j == q.j) || (c.i == q.j && c.j == q.i)) { c = nb; found = true; break; }
      if (!found) mols[m].bond.push_back(nb);
    }
    int solute_shake = 0;
    for (int m = 0; m < s.nmol - s.nwat; ++m) solute_shake += (int)mols[m].bond.size();
    // stable compaction of molecules with constraints (src/md.f90:4363-4378)
    int trg = 0, src = 1;
    while (src < s.nmol) {
      if (mols[trg].bond.empty()) {
        mols[trg] = mols[src];
        mols[src].bond.clear();
        ++src;
      } else {
        ++trg;
        if (trg == src) ++src;
      }
    }
    int nmols = s.nmol > 0 ? trg + 1 : 0;
    mols.resize(nmols);
    s.shake_mol = mols;
    int nc = 0;
    for (auto& m : s.shake_mol) nc += (int)m.bond.size();
    s.shake_constraints = nc;
    s.exclshk = exclshk;
    int nexats = 0;
    for (int i = 0; i < s.natom; ++i) nexats += s.excl[i] ? 1 : 0;
    s.ndegf = 3 * s.natom - nc;
    s.ndegfree = (int)(s.ndegf - 3 * t.nexats + exclshk);
    s.ndegf_solvent = s.ndegf - 3 * s.nat_solute + solute_shake;
    s.ndegf_solute = s.ndegf - s.ndegf_solvent;
    s.ndegfree_solvent = 3 * (s.natom - s.nat_solute) - (nc - solute_shake);
    s.ndegfree_solute = s.ndegfree - s.ndegfree_solvent;
    s.separate_scaling = in.separate_scaling;
    if (s.ndegfree_solvent * s.ndegfree_solute == 0) {
      s.detail_temps = false;
      s.separate_scaling = false;
    } else {
      s.detail_temps = true;
    }
    // thermostat and integrator constants (src/md.f90:4646-4708,
    // 4416-4424, 6446-6462)
    s.velocity_verlet = in.integrator == "velocity-verlet";
    s.dv_mod = s.velocity_verlet ? 0.5 * s.dt : s.dt;
    if (in.thermostat == "langevin") {
      if (in.freeze) throw std::runtime_error("the Langevin thermostat with excluded_freeze is not supported");
      s.thermostat = 1;
      s.friction = in.have_friction ? in.friction : 1 / s.tau_t;
      s.langevin_gauss = in.langevin_gauss;
      s.dv_friction = (1.0 - s.friction * s.dv_mod);
      const double gkt = s.langevin_gauss ? 2.0 * s.friction * kBoltz * in.temp0 / s.dv_mod
                                          : 24.0 * s.friction * kBoltz * in.temp0 / s.dv_mod;
      s.rands.resize(s.natom);
      for (int i = 0; i < s.natom; ++i) s.rands[i] = std::sqrt(gkt / s.winv[i]);
    } else if (in.thermostat == "nose-hoover") {
      s.thermostat = 2;
      s.kbt = kBoltz * in.temp0;
      const double nhq = in.have_nhq ? in.nhq : s.kbt / (s.tau_t * s.tau_t);
      s.qnh.assign(in.nhchains, nhq);
      s.qnh[0] = s.ndegf * nhq;
    }
    // clear angles whose i-k pair is constrained
    for (auto& m : s.shake_mol)
      for (auto& c : m.bond)
        for (auto& a : ang)
          if ((a.i == c.i && a.k == c.j) || (a.i == c.j && a.k == c.i)) { a.cod = 0; break; }
  }

  // make_nbqqlist (src/md.f90:9582-9804): Q connectivity over solute bonds
  if (s.nqat > 0) {
    const int ns = s.nat_solute, nq = s.nqat, nst = s.nstates;
    s.qconn.assign((size_t)nst * nq * ns, 9);
    // find_bonded (src/md.f90:9626-9681): solute bonds that were not
    // redefined (shaken ones included) plus the Q bonds present in the state
    std::vector<std::vector<int>> nbr0(ns);
    for (int b = 0; b < t.nbonds_solute; ++b) {
      if (bnd[b].cod == 0) continue;
      nbr0[bnd[b].i].push_back(bnd[b].j);
      nbr0[bnd[b].j].push_back(bnd[b].i);
    }
    auto qc = [&](int st, int iq, int a) -> int& { return s.qconn[((size_t)st * nq + iq) * ns + a]; };
    for (int st = 0; st < nst; ++st) {
      std::vector<std::vector<int>> nbr = nbr0;
      for (const QTerm& b : s.qbnd)
        if (b.cod[st] > 0) { nbr[b.i].push_back(b.j); nbr[b.j].push_back(b.i); }
      for (int iq = 0; iq < nq; ++iq) {
        int o = s.iqseq[iq];
        qc(st, iq, o) = 1;
        std::vector<int> frontier{o};
        for (int lvl = 2; lvl <= 4; ++lvl) {
          std::vector<int> nxt;
          for (int u : frontier)
            for (int w : nbr[u])
              if (qc(st, iq, w) > lvl) { qc(st, iq, w) = lvl; nxt.push_back(w); }
          frontier = nxt;
        }
      }
    }
    // special exclusions of Q atoms: connectivity 0 in the flagged states
    // (make_qconn, src/md.f90:9598-9616)
    for (const auto& e : s_exspec) {
      for (int side = 0; side < 2; ++side) {
        const int a = side ? e.j : e.i, b = side ? e.i : e.j;
        const int iq = s.iqatom[a];
        if (iq == 0) continue;
        if (b >= ns) throw std::runtime_error("special exclusions of Q atoms must be with solute atoms");
        for (int st = 0; st < nst; ++st)
          if (e.flag[st]) qc(st, iq - 1, b) = 0;
      }
    }
    s.nbqq.assign(nst, {});
    for (int iq = 0; iq < nq - 1; ++iq)
      for (int jq = iq + 1; jq < nq; ++jq) {
        int j = s.iqseq[jq];
        for (int st = 0; st < nst; ++st) {
          int c = qc(st, iq, j);
          if (c > 3) {
            NbqPair p{iq, j, jq, 0, 1.0};
            // LJ code: 1-4 pairs 3; with Q-atom types 1, or 2 for a soft
            // pair; else the topology code (nbqqlist, src/md.f90:9734-9778)
            if (c == 4) p.ljcod = 3;
            else if (!s.qvdw_flag) p.ljcod = s.ljcod[s.iac[s.iqseq[iq]] * s.natyps + s.iac[j]];
            else {
              p.ljcod = 1;
              for (const auto& sp : s_soft_pairs)
                if ((iq == sp.first && jq == sp.second) || (jq == sp.first && iq == sp.second)) { p.ljcod = 2; break; }
            }
            for (const auto& es : s_el_scale)
              if ((iq == es.iq && jq == es.jq) || (iq == es.jq && jq == es.iq)) { p.el_scale = es.scale[st]; break; }
            s.nbqq[st].push_back(p);
          }
        }
      }
    for (int j = 0; j < ns; ++j) {
      if (s.iqatom[j] > 0) continue;
      bool any = false;
      for (int st = 0; st < nst && !any; ++st)
        for (int iq = 0; iq < nq && !any; ++iq) any = qc(st, iq, j) <= 3;
      if (!any) continue;
      for (int iq = 0; iq < nq; ++iq)
        for (int st = 0; st < nst; ++st) {
          int c = qc(st, iq, j);
          if (c >= 4) {
            NbqPair p{iq, j, -1, 0, 1.0};
            p.ljcod = (c == 4) ? 3 : s.qvdw_flag ? 1 : s.ljcod[s.iac[s.iqseq[iq]] * s.natyps + s.iac[j]];
            s.nbqq[st].push_back(p);
          }
        }
    }
  }

  // shrink_topology (src/md.f90:17683-17783): swap-with-last removal
  {
    int nb = (int)bnd.size(), nbs = t.nbonds_solute;
    int i = 0;
    while (i < nb) {
      if (bnd[i].cod <= 0) {
        if (i < nbs) { bnd[i] = bnd[nbs - 1]; bnd[nbs - 1] = bnd[nb - 1]; --nbs; }
        else bnd[i] = bnd[nb - 1];
        --nb;
        continue;
      }
      ++i;
    }
    bnd.resize(nb);
    s.bnd = bnd; s.nbonds_solute = nbs;
    int na = (int)ang.size(), nas = t.nangles_solute;
    i = 0;
    while (i < na) {
      if (ang[i].cod == 0) {
        if (i < nas) { ang[i] = ang[nas - 1]; ang[nas - 1] = ang[na - 1]; --nas; }
        else ang[i] = ang[na - 1];
        --na;
        continue;
      }
      ++i;
    }
    ang.resize(na);
    s.ang = ang; s.nangles_solute = nas;
    auto shrink4 = [&](std::vector<Torsion> v, int nsol, std::vector<Torsion>& out, int& nsol_out) {
      int n = (int)v.size();
      int k = 0;
      while (k < n) {
        bool all_ex = in.exclude_bonded && s.excl[v[k].i] && s.excl[v[k].j] && s.excl[v[k].k] && s.excl[v[k].l];
        if (all_ex || v[k].cod == 0) {
          if (k < nsol) { v[k] = v[nsol - 1]; v[nsol - 1] = v[n - 1]; --nsol; }
          else v[k] = v[n - 1];
          --n;
          continue;
        }
        ++k;
      }
      v.resize(n);
      out = v; nsol_out = nsol;
    };
    shrink4(t.tor, t.ntors_solute, s.tor, s.ntors_solute);
    shrink4(t.imp, t.nimps_solute, s.imp, s.nimps_solute);
  }

  // maxwell, initial_shaking, stop_cm_translation (src/qdyn.f90:128-132)
  if (in.iseed > 0) {
    Randm rnd;
    int ig = in.iseed;
    double kT = kBoltz * in.tmaxw;
    for (int i = 0; i < s.natom; ++i) {
      double sd = std::sqrt(kT / t.types[s.iac[i]].mass);
      for (int j = 0; j < 3; ++j) {
        double a = 0.0;
        for (int k = 0; k < 12; ++k) a = a + rnd.next(ig);
        s.v[3 * i + j] = (a - 6.0) * sd + 0.0;
      }
    }
    s.rng = rnd;
    {
      std::vector<double> xx = s.x;
      if (shake_cpu(s, xx, s.x) < 0) throw std::runtime_error("initial shake failure");
      for (int k = 0; k < 3 * s.natom; ++k) xx[k] = s.x[k] - s.dt * s.v[k];
      if (shake_cpu(s, s.x, xx) < 0) throw std::runtime_error("initial shake failure");
      for (int k = 0; k < 3 * s.natom; ++k) s.v[k] = (s.x[k] - xx[k]) / s.dt;
    }
    double totmass = 0, vcm[3] = {0, 0, 0};
    for (int i = 0; i < s.natom; ++i) {
      double rm = t.types[s.iac[i]].mass;
      totmass = totmass + rm;
      for (int j = 0; j < 3; ++j) vcm[j] = vcm[j] + rm * s.v[3 * i + j];
    }
    for (int j = 0; j < 3; ++j) vcm[j] = vcm[j] / totmass;
    for (int i = 0; i < s.natom; ++i)
      for (int j = 0; j < 3; ++j) s.v[3 * i + j] = s.v[3 * i + j] - vcm[j];
  }
  monitor_setup(s);
  return s;
}

}  // namespace qsys
