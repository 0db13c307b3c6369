// Readers for qdyn's topology, input, FEP and restart files.
//
// The topology reader follows the record structure of topo_read
// (src/topo.f90:515-1074): every Fortran READ starts at a new record (line)
// and a list-directed READ of n items continues over as many records as it
// needs, discarding the rest of the last one.  The input reader follows the
// sectioned key/value format of prmfile (src/prmfile.f90:834-963) and the
// keys and defaults of initialize (src/md.f90:4453-5358).
#include "qsys.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace qsys {
namespace {

struct Records {
  std::vector<std::string> lines;
  size_t pos = 0;
  explicit Records(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string l;
    while (std::getline(f, l)) {
      if (!l.empty() && l.back() == '\r') l.pop_back();
      lines.push_back(l);
    }
  }
  bool eof() const { return pos >= lines.size(); }
  const std::string& next() {
    if (pos >= lines.size()) throw std::runtime_error("unexpected end of topology file");
    return lines[pos++];
  }
  void skip() { next(); }
  static void tokens(const std::string& l, std::vector<std::string>& out) {
    size_t i = 0;
    while (i < l.size()) {
      while (i < l.size() && (std::isspace((unsigned char)l[i]) || l[i] == ',')) ++i;
      if (i >= l.size()) break;
      size_t j = i;
      while (j < l.size() && !std::isspace((unsigned char)l[j]) && l[j] != ',') ++j;
      out.push_back(l.substr(i, j - i));
      i = j;
    }
  }
  // list-directed READ of n values starting at a new record
  std::vector<std::string> items(size_t n) {
    std::vector<std::string> out;
    while (out.size() < n) {
      std::vector<std::string> t;
      tokens(next(), t);
      for (auto& s : t) {
        if (out.size() < n) out.push_back(s);
      }
    }
    return out;
  }
};

double to_d(const std::string& s) {
  std::string t = s;
  for (auto& c : t) if (c == 'd' || c == 'D') c = 'e';
  char* end = nullptr;
  double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str()) throw std::runtime_error("bad number '" + s + "'");
  return v;
}
long to_i(const std::string& s) {
  char* end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || (*end && *end != '.')) throw std::runtime_error("bad integer '" + s + "'");
  return v;
}
// leading integers of a record, stopping at the first non-integer token
std::vector<long> lead_ints(const std::string& l) {
  std::vector<std::string> t;
  Records::tokens(l, t);
  std::vector<long> out;
  for (auto& s : t) {
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end) break;
    out.push_back(v);
  }
  return out;
}
std::vector<double> lead_reals(const std::string& l) {
  std::vector<std::string> t;
  Records::tokens(l, t);
  std::vector<double> out;
  for (auto& s : t) {
    std::string u = s;
    for (auto& c : u) if (c == 'd' || c == 'D') c = 'e';
    char* end = nullptr;
    double v = std::strtod(u.c_str(), &end);
    if (end == u.c_str() || *end) break;
    out.push_back(v);
  }
  return out;
}

// (80i1)/(80l1) fixed-format bit tables: nat_solute rows of kMaxNbrRange
void read_bits(Records& r, int nrows, std::vector<uint32_t>& rows, bool logical) {
  rows.assign(nrows, 0u);
  size_t total = (size_t)nrows * kMaxNbrRange, k = 0;
  while (k < total) {
    const std::string& l = r.next();
    for (size_t c = 0; c < 80 && k < total; ++c, ++k) {
      char ch = c < l.size() ? l[c] : ' ';
      bool on = logical ? (ch == 'T' || ch == 't') : (ch == '1');
      if (on) rows[k / kMaxNbrRange] |= 1u << (k % kMaxNbrRange);
    }
  }
}

}  // namespace

Topology read_topology(const std::string& path) {
  Records r(path);
  Topology t;
  std::string title = r.next();
  if (title.rfind("Q topology file", 0) == 0) {
    for (;;) {
      std::string l = r.next();
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      std::string key = tk[0];
      for (auto& c : key) c = (char)std::toupper((unsigned char)c);
      if (key == "VERSION" && tk.size() > 1) t.version = to_d(tk[1]);
      else if (key == "FF_TYPE" && tk.size() > 1) t.ff_type = (int)to_i(tk[1]);
      else if (key == "END") break;
    }
  } else {
    t.version = 2;
  }
  if (t.version < 4.15) throw std::runtime_error("topology version < 4.15");
  {
    auto v = lead_ints(r.next());
    if (v.empty()) throw std::runtime_error("bad atom count record");
    t.nat_pro = (int)v[0];
    t.nat_solute = v.size() > 1 ? (int)v[1] : t.nat_pro;
    t.solv_atom = v.size() > 2 ? (int)v[2] : 3;
  }
  int N = t.nat_pro;
  {
    auto it = r.items(3 * (size_t)N);
    t.xtop.resize(3 * N);
    for (int i = 0; i < 3 * N; ++i) t.xtop[i] = to_d(it[i]);
  }
  r.skip();
  {
    auto it = r.items(N);
    t.iac.resize(N);
    for (int i = 0; i < N; ++i) t.iac[i] = (int)to_i(it[i]) - 1;
  }
  // bonds
  {
    auto v = lead_ints(r.next());
    int nb = (int)v.at(0);
    t.nbonds_solute = v.size() > 1 ? (int)v[1] : nb;
    auto it = r.items(3 * (size_t)nb);
    t.bnd.resize(nb);
    for (int b = 0; b < nb; ++b)
      t.bnd[b] = {(int)to_i(it[3*b]) - 1, (int)to_i(it[3*b+1]) - 1, (int)to_i(it[3*b+2])};
    int nc = (int)lead_ints(r.next()).at(0);
    t.bondlib.resize(nc);
    for (int c = 0; c < nc; ++c) {
      auto rv = lead_reals(r.next());
      t.bondlib[c] = {rv.at(1), rv.at(2)};
    }
  }
  // angles
  {
    auto v = lead_ints(r.next());
    int na = (int)v.at(0);
    t.nangles_solute = v.size() > 1 ? (int)v[1] : na;
    auto it = r.items(4 * (size_t)na);
    t.ang.resize(na);
    for (int a = 0; a < na; ++a)
      t.ang[a] = {(int)to_i(it[4*a]) - 1, (int)to_i(it[4*a+1]) - 1, (int)to_i(it[4*a+2]) - 1,
                  (int)to_i(it[4*a+3])};
    int nc = (int)lead_ints(r.next()).at(0);
    t.anglib.resize(nc);
    for (int c = 0; c < nc; ++c) {
      auto rv = lead_reals(r.next());
      AngLib al{rv.at(1), rv.at(2), 0.0, 0.0};
      if (rv.size() > 3) al.ureyfk = rv[3];
      if (rv.size() > 4) al.ureyr0 = rv[4];
      t.anglib[c] = al;
    }
  }
  // torsions
  {
    auto v = lead_ints(r.next());
    int nt = (int)v.at(0);
    t.ntors_solute = v.size() > 1 ? (int)v[1] : nt;
    auto it = r.items(5 * (size_t)nt);
    t.tor.resize(nt);
    for (int a = 0; a < nt; ++a)
      t.tor[a] = {(int)to_i(it[5*a]) - 1, (int)to_i(it[5*a+1]) - 1, (int)to_i(it[5*a+2]) - 1,
                  (int)to_i(it[5*a+3]) - 1, (int)to_i(it[5*a+4])};
    int nc = (int)to_i(r.items(1)[0]);
    t.torlib.resize(nc);
    for (int c = 0; c < nc; ++c) {
      auto it2 = r.items(5);
      t.torlib[c] = {to_d(it2[1]), to_d(it2[2]), to_d(it2[3]), to_d(it2[4])};
    }
  }
  // impropers
  {
    auto v = lead_ints(r.next());
    int ni = (int)v.at(0);
    t.nimps_solute = v.size() > 1 ? (int)v[1] : ni;
    auto it = r.items(5 * (size_t)ni);
    t.imp.resize(ni);
    for (int a = 0; a < ni; ++a)
      t.imp[a] = {(int)to_i(it[5*a]) - 1, (int)to_i(it[5*a+1]) - 1, (int)to_i(it[5*a+2]) - 1,
                  (int)to_i(it[5*a+3]) - 1, (int)to_i(it[5*a+4])};
    auto v2 = lead_ints(r.next());
    int nc = (int)v2.at(0);
    t.imp_type = v2.size() > 1 ? (int)v2[1] : 1;
    t.implib.resize(nc);
    for (int c = 0; c < nc; ++c) {
      auto it2 = r.items(3);
      t.implib[c] = {to_d(it2[1]), to_d(it2[2])};
    }
  }
  // charges
  {
    r.skip();
    auto it = r.items(N);
    t.crg.resize(N);
    for (int i = 0; i < N; ++i) t.crg[i] = to_d(it[i]);
  }
  // charge groups
  {
    auto v = lead_ints(r.next());
    int ncgp = (int)v.at(0);
    t.ncgp_solute = v.size() > 1 ? (int)v[1] : ncgp;
    t.iuse_switch_atom = v.size() > 2 ? (int)v[2] : 1;
    t.cgp.resize(ncgp);
    int k = 0;
    for (int g = 0; g < ncgp; ++g) {
      auto h = r.items(2);
      int n = (int)to_i(h[0]);
      t.cgp[g] = {(int)to_i(h[1]) - 1, k, k + n - 1};
      auto at = r.items(n);
      for (int a = 0; a < n; ++a) t.cgpatom.push_back((int)to_i(at[a]) - 1);
      k += n;
    }
  }
  // atom types
  {
    int nt = (int)to_i(r.items(1)[0]);
    t.types.resize(nt);
    t.ivdw_rule = (int)to_i(r.items(1)[0]);
    auto rv = lead_reals(r.next());
    t.el14_scale = rv.at(0);
    t.coulomb_constant = (rv.size() > 1 && rv[1] > 0) ? rv[1] : 332.0;
    r.skip();
    auto m = r.items(nt);
    for (int i = 0; i < nt; ++i) t.types[i].mass = to_d(m[i]);
    for (int j = 0; j < 3; ++j) {
      r.skip();
      auto a = r.items(nt);
      for (int i = 0; i < nt; ++i) t.types[i].avdw[j] = to_d(a[i]);
      r.skip();
      auto b = r.items(nt);
      for (int i = 0; i < nt; ++i) t.types[i].bvdw[j] = to_d(b[i]);
    }
    int nlj2 = (int)to_i(r.items(1)[0]);
    for (int i = 0; i < nlj2; ++i) {
      auto p = r.items(2);
      t.lj2.push_back({(int)to_i(p[0]) - 1, (int)to_i(p[1]) - 1});
    }
  }
  // neighbour / exclusion lists
  {
    r.items(1);
    if (t.nat_solute > 0) read_bits(r, t.nat_solute, t.list14, false);
    else t.list14.clear();
    int n14long = (int)to_i(r.items(1)[0]);
    for (int i = 0; i < n14long; ++i) {
      auto p = r.items(2);
      t.list14long.push_back({(int)to_i(p[0]) - 1, (int)to_i(p[1]) - 1});
    }
    r.items(1);
    if (t.nat_solute > 0) read_bits(r, t.nat_solute, t.listex, false);
    int nexlong = (int)to_i(r.items(1)[0]);
    for (int i = 0; i < nexlong; ++i) {
      auto p = r.items(2);
      t.listexlong.push_back({(int)to_i(p[0]) - 1, (int)to_i(p[1]) - 1});
    }
  }
  // residues and molecules
  {
    std::string l = r.next();
    if (l.find_first_not_of(" \t") == std::string::npos) l = r.next();
    auto v = lead_ints(l);
    int nres = (int)v.at(0);
    t.nres_solute = v.size() > 1 ? (int)v[1] : nres;
    if (nres > 0) {
      auto it = r.items(nres);
      for (int i = 0; i < nres; ++i) t.res_start.push_back((int)to_i(it[i]) - 1);
    }
    r.skip();
    for (int i = 0; i < (nres + 15) / 16; ++i) r.skip();
    int nmol = (int)to_i(r.items(1)[0]);
    if (nmol > 0) {
      auto it = r.items(nmol);
      for (int i = 0; i < nmol; ++i) t.istart_mol.push_back((int)to_i(it[i]) - 1);
    }
    int nt = (int)t.types.size();
    r.skip();
    for (int i = 0; i < (nt + 7) / 8; ++i) r.skip();
    r.skip();
    for (int i = 0; i < (nt + 12) / 13; ++i) r.skip();
  }
  if (t.version < 4) throw std::runtime_error("topology version < 4 has no boundary data");
  t.solvent_type = (int)to_i(r.items(1)[0]);
  {
    std::string l = r.next();
    std::vector<std::string> tk;
    Records::tokens(l, tk);
    if (!tk.empty() && tk[0] == "PBC") {
      // periodic box: box lengths and centre; no excluded atoms
      // (src/topo.f90:1022-1032)
      t.use_pbc = true;
      auto bl = r.items(3);
      for (int k = 0; k < 3; ++k) t.boxlength[k] = to_d(bl[k]);
      auto bc = r.items(3);
      for (int k = 0; k < 3; ++k) t.boxcentre[k] = to_d(bc[k]);
      t.excl.assign(N, 0);
      return t;
    }
    auto rv = lead_reals(l);
    if (t.version < 5.01) {
      t.rexcl_o = rv.at(0); t.rexcl_i = rv.at(1); t.rwat = rv.at(2);
    } else {
      t.rexcl_o = rv.at(0); t.rwat = rv.at(1);
    }
    auto a = r.items(3);
    for (int k = 0; k < 3; ++k) t.xpcent[k] = to_d(a[k]);
    auto b = r.items(3);
    for (int k = 0; k < 3; ++k) t.xwcent[k] = to_d(b[k]);
    auto c = r.items(2);
    t.nexats = (int)to_i(c[0]);
    t.nexwat = (int)to_i(c[1]);
    t.excl.assign(N, 0);
    int k = 0;
    while (k < N) {
      const std::string& fl = r.next();
      for (size_t q = 0; q < 80 && k < N; ++q, ++k) {
        char ch = q < fl.size() ? fl[q] : 'F';
        t.excl[k] = (ch == 'T' || ch == 't');
      }
    }
  }
  return t;
}

// --------------------------------------------------------------------------
// input file (prmfile semantics)
namespace {
struct Sections {
  std::map<std::string, std::vector<std::string>> sec;
  explicit Sections(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string l, cur;
    bool any = false;
    while (std::getline(f, l)) {
      if (!l.empty() && l.back() == '\r') l.pop_back();
      size_t s = l.find_first_not_of(" \t");
      if (s == std::string::npos) continue;
      char c = l[s];
      if (c == '!' || c == '#' || c == '*') continue;
      if (c == '[') {
        size_t e = l.find(']');
        cur = l.substr(s + 1, e - s - 1);
        for (auto& ch : cur) ch = (char)std::toupper((unsigned char)ch);
        sec[cur];
        any = true;
        continue;
      }
      if (!any) throw std::runtime_error("data outside a section in " + path);
      // strip a trailing comment: '!' anywhere, '#'/'*' after whitespace
      size_t cs = std::string::npos;
      for (size_t k = s + 1; k < l.size(); ++k) {
        if (l[k] == '!' || ((l[k] == '#' || l[k] == '*') && (l[k-1] == ' ' || l[k-1] == '\t'))) {
          cs = k; break;
        }
      }
      if (cs != std::string::npos) l = l.substr(0, cs);
      sec[cur].push_back(l);
    }
  }
  bool has(const std::string& s) const { return sec.count(s) > 0; }
  bool get(const std::string& s, const std::string& key, std::string& val) const {
    auto it = sec.find(s);
    if (it == sec.end()) return false;
    for (auto& l : it->second) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.size() >= 2 && tk[0] == key) { val = tk[1]; return true; }
    }
    return false;
  }
};
bool parse_onoff(const std::string& v, bool& out) {
  if (v == "on" || v == "ON" || v == "1") { out = true; return true; }
  if (v == "off" || v == "OFF" || v == "0") { out = false; return true; }
  return false;
}
}  // namespace

// Extra restraints (src/md.f90:5211-5354); atoms 0-based.
static void read_restraints(const Sections& S, MDInput& m) {
  auto lines = [&](const char* name) {
    std::vector<std::vector<std::string>> out;
    if (!S.has(name)) return out;
    for (auto& l : S.sec.at(name)) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (!tk.empty()) out.push_back(tk);
    }
    return out;
  };
  auto atom = [](const std::string& t) { return (int)to_i(t) - 1; };
  // "i j fk ih [to_centre]" (to_centre 0 unless given)
  for (auto& tk : lines("SEQUENCE_RESTRAINTS")) {
    if (tk.size() < 4) throw std::runtime_error("bad [sequence_restraints] line");
    m.rstseq.push_back({atom(tk[0]), atom(tk[1]), to_d(tk[2]), (int)to_i(tk[3]),
                        tk.size() >= 5 ? (int)to_i(tk[4]) : 0});
  }
  for (auto& tk : lines("ATOM_RESTRAINTS")) {
    if (tk.size() < 8) throw std::runtime_error("Invalid atom restraint data.");
    RstPos r{atom(tk[0]), {to_d(tk[1]), to_d(tk[2]), to_d(tk[3])}, {to_d(tk[4]), to_d(tk[5]), to_d(tk[6])},
             (int)to_i(tk[7])};
    m.rstpos.push_back(r);
  }
  // "i j d1 d2 fk ipsi"; i and j may be residue:atom (resolved at set-up)
  for (auto& tk : lines("DISTANCE_RESTRAINTS")) {
    if (tk.size() < 6) throw std::runtime_error("Invalid distance restraint data.");
    bool text = tk[0].find(':') != std::string::npos || tk[1].find(':') != std::string::npos;
    m.rstdis.push_back({text ? -1 : atom(tk[0]), text ? -1 : atom(tk[1]), to_d(tk[2]), to_d(tk[3]),
                        to_d(tk[4]), (int)to_i(tk[5])});
    m.rstdis_text.push_back(text ? std::make_pair(tk[0], tk[1]) : std::make_pair(std::string(), std::string()));
  }
  for (auto& tk : lines("ANGLE_RESTRAINTS")) {
    if (tk.size() < 6) throw std::runtime_error("Invalid angle restraint data.");
    m.rstang.push_back({atom(tk[0]), atom(tk[1]), atom(tk[2]), to_d(tk[3]), to_d(tk[4]), (int)to_i(tk[5])});
  }
  for (auto& tk : lines("WALL_RESTRAINTS")) {
    if (tk.size() < 7) throw std::runtime_error("Invalid wall restraint data.");
    m.rstwal.push_back({atom(tk[0]), atom(tk[1]), to_d(tk[2]), to_d(tk[3]), to_d(tk[4]), to_d(tk[5]),
                        (int)to_i(tk[6])});
  }
}

MDInput read_input(const std::string& path) {
  Sections S(path);
  MDInput m;
  std::string v;
  if (S.has("PBC")) {
    // [PBC] (src/md.f90:4516-4598); a changed box size (control_box) and
    // atom-based volume scaling are not implemented
    m.box = true;
    if (S.get("PBC", "rigid_box_centre", v)) parse_onoff(v, m.rigid_box_centre);
    if (S.get("PBC", "constant_pressure", v)) parse_onoff(v, m.constant_pressure);
    if (m.constant_pressure) {
      if (!S.get("PBC", "max_volume_displ", v))
        throw std::runtime_error("maximum volume displacement not specified (section PBC)");
      m.max_vol_displ = to_d(v);
      if (S.get("PBC", "pressure_seed", v)) m.pressure_seed = (int)to_i(v);
      if (S.get("PBC", "pressure", v)) m.pressure = to_d(v);
      // bar to the internal unit (src/md.f90:4553)
      m.pressure = m.pressure * 1.43836e-5;
      bool atom_based = false;
      if (S.get("PBC", "atom_based_scaling", v)) parse_onoff(v, atom_based);
      if (atom_based) throw std::runtime_error("atom-based volume scaling is not supported");
    } else if (S.get("PBC", "control_box", v)) {
      throw std::runtime_error("control_box is not supported");
    }
    if (S.get("PBC", "put_solvent_back_in_box", v)) parse_onoff(v, m.put_solvent_back);
    if (S.get("PBC", "put_solute_back_in_box", v)) parse_onoff(v, m.put_solute_back);
  }
  if (!S.has("MD")) throw std::runtime_error("input has no [MD] section");
  if (!S.get("MD", "steps", v)) throw std::runtime_error("[MD] steps missing");
  m.nsteps = (int)to_i(v);
  if (!S.get("MD", "stepsize", v)) throw std::runtime_error("[MD] stepsize missing");
  m.stepsize = to_d(v);
  if (!S.get("MD", "temperature", v)) throw std::runtime_error("[MD] temperature missing");
  m.temp0 = to_d(v);
  if (S.get("MD", "bath_coupling", v)) m.tau_t = to_d(v);
  // thermostat and integrator (src/md.f90:4646-4708); a name the reference
  // does not recognise falls back to Berendsen only if it is "nose_hoover",
  // as the reference's test of the name lets it
  if (S.get("MD", "thermostat", v)) m.thermostat = v;
  if (m.thermostat == "nose_hoover") m.thermostat = "berendsen";
  if (m.thermostat != "berendsen" && m.thermostat != "langevin" && m.thermostat != "nose-hoover")
    throw std::runtime_error("this thermostat does not exist in Q: " + m.thermostat);
  if (m.thermostat == "langevin") {
    if (S.get("MD", "langevin_friction", v)) { m.have_friction = true; m.friction = to_d(v); }
    if (S.get("MD", "langevin_random", v)) parse_onoff(v, m.langevin_gauss);
  }
  if (m.thermostat == "nose-hoover") {
    if (S.get("MD", "nhchains", v)) m.nhchains = (int)to_i(v);
    if (S.get("MD", "nose-hoover_mass", v)) { m.have_nhq = true; m.nhq = to_d(v); }
    if (m.nhchains < 2) throw std::runtime_error("nhchains must be at least 2");
  }
  if (S.get("MD", "integrator", v)) m.integrator = v;
  if (m.integrator != "leap-frog" && m.integrator != "velocity-verlet")
    throw std::runtime_error("no such integrator exists in Q: " + m.integrator);
  if (S.get("MD", "separate_scaling", v)) parse_onoff(v, m.separate_scaling);
  m.iseed = 1;
  if (S.get("MD", "random_seed", v)) m.iseed = (int)to_i(v);
  if (S.get("MD", "initial_temperature", v)) { m.have_tmaxw = true; m.tmaxw = to_d(v); }
  else m.iseed = 0;
  if (S.get("MD", "shake_solvent", v)) parse_onoff(v, m.shake_solvent);
  if (S.get("MD", "shake_solute", v)) parse_onoff(v, m.shake_solute);
  if (S.get("MD", "shake_hydrogens", v)) parse_onoff(v, m.shake_hydrogens);
  if (S.get("MD", "lrf", v)) parse_onoff(v, m.use_lrf);
  if (S.has("CUT-OFFS")) {
    if (S.get("CUT-OFFS", "solute_solute", v)) m.rcpp = to_d(v);
    if (S.get("CUT-OFFS", "solvent_solvent", v)) m.rcww = to_d(v);
    if (S.get("CUT-OFFS", "solute_solvent", v)) m.rcpw = to_d(v);
    if (S.get("CUT-OFFS", "q_atom", v)) { m.rcq = to_d(v); m.have_rcq = true; }
    if (m.use_lrf && S.get("CUT-OFFS", "lrf", v)) m.rclrf = to_d(v);
  }
  if (m.use_lrf && (m.rclrf < m.rcpp || m.rclrf < m.rcpw || m.rclrf < m.rcww))
    throw std::runtime_error("LRF cut-off must not be smaller than solute or solvent cut-offs");
  if (S.has("SPHERE")) {
    if (S.get("SPHERE", "shell_radius", v)) { m.have_shell_radius = true; m.shell_radius = to_d(v); }
    if (S.get("SPHERE", "shell_force", v)) m.fk_pshell = to_d(v);
    if (S.get("SPHERE", "excluded_force", v)) m.fk_fix = to_d(v);
    if (m.fk_fix <= 0) m.fk_fix = 200.0;
    if (S.get("SPHERE", "excluded_freeze", v)) parse_onoff(v, m.freeze);
    if (S.get("SPHERE", "exclude_bonded", v)) parse_onoff(v, m.exclude_bonded);
  }
  std::string ws = S.has("SOLVENT") ? "SOLVENT" : (S.has("WATER") ? "WATER" : "");
  if (!ws.empty()) {
    if (S.get(ws, "radius", v)) m.rwat_in = to_d(v);
    if (S.get(ws, "radial_force", v)) m.fk_wsphere = to_d(v);
    if (S.get(ws, "polarisation", v)) parse_onoff(v, m.wpol_restr);
    m.wpol_born = m.wpol_restr;
    if (S.get(ws, "charge_correction", v)) parse_onoff(v, m.wpol_born);
    if (S.get(ws, "polarisation_force", v)) m.fkwpol = to_d(v);
    if (S.get(ws, "morse_depth", v)) m.dwmz = to_d(v);
    if (S.get(ws, "morse_width", v)) m.awmz = to_d(v);
  }
  if (S.has("INTERVALS")) {
    if (S.get("INTERVALS", "non_bond", v)) m.nbcycle = (int)to_i(v);
    // volume moves every volume_change steps, read only under constant
    // pressure (src/md.f90:4955-4960)
    if (m.constant_pressure && S.get("INTERVALS", "volume_change", v)) m.ivolume_cycle = (int)to_i(v);
    if (S.get("INTERVALS", "output", v)) m.iout_cycle = (int)to_i(v);
    m.itemp_cycle = m.iout_cycle;
    if (S.get("INTERVALS", "temperature", v)) m.itemp_cycle = (int)to_i(v);
    if (S.get("INTERVALS", "energy", v)) m.iene_cycle = (int)to_i(v);
    if (S.get("INTERVALS", "trajectory", v)) m.itrj_cycle = (int)to_i(v);
  }
  if (!S.has("FILES")) throw std::runtime_error("input has no [files] section");
  if (!S.get("FILES", "topology", m.top_file)) throw std::runtime_error("[files] topology missing");
  S.get("FILES", "restart", m.restart_file);
  S.get("FILES", "final", m.final_file);
  S.get("FILES", "energy", m.energy_file);
  S.get("FILES", "fep", m.fep_file);
  if (m.restart_file.empty() && m.iseed == 0)
    throw std::runtime_error("need a restart file or random_seed + initial_temperature");
  if (S.has("LAMBDAS")) {
    for (auto& l : S.sec.at("LAMBDAS")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      for (auto& s : tk) m.lambdas.push_back(to_d(s));
    }
  }
  S.get("FILES", "restraint", m.restraint_file);
  S.get("FILES", "trajectory", m.trj_file);
  if (m.itrj_cycle > 0 && m.trj_file.empty())
    throw std::runtime_error("Trajectory file name required to write trajectory!");
  if (S.has("TRAJECTORY_ATOMS"))
    for (auto l : S.sec.at("TRAJECTORY_ATOMS")) {
      size_t a = l.find_first_not_of(" \t"), b = l.find_last_not_of(" \t");
      if (a != std::string::npos) m.trajectory_atoms.push_back(l.substr(a, b - a + 1));
    }
  read_restraints(S, m);
  return m;
}

// --------------------------------------------------------------------------
namespace {
const double kPiR = 4.0 * std::atan(1.0);
const double kDeg2RadR = kPiR / 180.0;     // src/qatom.f90:218-219

// Lines of a type section "[xxx_types]": type number (1-based) then values;
// types may come in any order (prm_max_enum, src/qatom.f90:923).
template <class Lib, class Fill>
void read_types(const Sections& S, const char* name, std::vector<Lib>& lib, std::vector<char>& have, Fill fill) {
  if (!S.has(name)) return;
  for (auto& l : S.sec.at(name)) {
    std::vector<std::string> tk;
    Records::tokens(l, tk);
    if (tk.empty()) continue;
    int j = (int)to_i(tk[0]);
    if (j < 1) throw std::runtime_error(std::string("bad type number in [") + name + "]");
    if ((int)lib.size() < j) { lib.resize(j); have.resize(j, 0); }
    fill(tk, lib[j - 1]);
    have[j - 1] = 1;
  }
}

// "[change_xxx]" lines: natoms topology atoms then one type code per state.
void read_changes(const Sections& S, const char* name, int natoms, int nstates, int natom,
                  const std::vector<char>& have, std::vector<QTerm>& out) {
  if (!S.has(name)) return;
  for (auto& l : S.sec.at(name)) {
    std::vector<std::string> tk;
    Records::tokens(l, tk);
    if (tk.empty()) continue;
    if ((int)tk.size() < natoms + nstates) throw std::runtime_error(std::string("bad [") + name + "] line");
    int a[4] = {0, 0, 0, 0};
    for (int m = 0; m < natoms; ++m) {
      a[m] = (int)to_i(tk[m]) - 1;
      if (a[m] < 0 || a[m] >= natom) throw std::runtime_error(std::string("invalid atom in [") + name + "]");
    }
    QTerm t{a[0], a[1], a[2], a[3], std::vector<int>(nstates, 0)};
    for (int st = 0; st < nstates; ++st) {
      int c = (int)to_i(tk[natoms + st]);
      if (c < 0 || (c > 0 && (c > (int)have.size() || !have[c - 1])))
        throw std::runtime_error(std::string("undefined type code in [") + name + "]");
      t.cod[st] = c;
    }
    out.push_back(t);
  }
}

void read_couplings(const Sections& S, const char* name, int nterms, const FepData& f,
                    std::vector<QCoupling>& out) {
  if (!S.has(name)) return;
  for (auto& l : S.sec.at(name)) {
    std::vector<std::string> tk;
    Records::tokens(l, tk);
    if (tk.empty()) continue;
    if (tk.size() < 2) throw std::runtime_error(std::string("bad [") + name + "] line");
    QCoupling c{(int)to_i(tk[0]) - 1, (int)to_i(tk[1]) - 1, tk.size() >= 3 ? (int)to_i(tk[2]) : 0};
    if (c.term < 0 || c.term >= nterms || c.bond < 0 || c.bond >= (int)f.qbnd.size())
      throw std::runtime_error(std::string("invalid term or bond in [") + name + "]");
    // coupled bonds must be Morse in every state where present
    // (bond_harmonic_in_any_state, src/qatom.f90:2039-2054)
    for (int st = 0; st < f.nstates; ++st) {
      int cod = f.qbnd[c.bond].cod[st];
      if (cod != 0 && f.qbondlib[cod - 1].fk != 0.0)
        throw std::runtime_error(std::string("coupled Q bond is harmonic in [") + name + "]");
    }
    out.push_back(c);
  }
}
}  // namespace

// Q-bonded sections of the FEP file (src/qatom.f90:920-1790).
static void read_fep_bonded(const Sections& S, FepData& f, int natom) {
  std::vector<char> hb, ha, ht, hi;
  // [bond_types]: Morse "j Dmz amz r0", else harmonic "j fk r0"
  read_types(S, "BOND_TYPES", f.qbondlib, hb, [](const std::vector<std::string>& tk, QBondLib& b) {
    if (tk.size() >= 4) b = {to_d(tk[1]), to_d(tk[2]), to_d(tk[3]), 0.0};
    else if (tk.size() == 3) b = {0.0, 0.0, to_d(tk[2]), to_d(tk[1])};
    else throw std::runtime_error("bad [bond_types] line");
  });
  // [angle_types]: "j fk theta0 [ureyfk urey_r0]"
  read_types(S, "ANGLE_TYPES", f.qanglib, ha, [](const std::vector<std::string>& tk, QAngLib& a) {
    if (tk.size() < 3) throw std::runtime_error("bad [angle_types] line");
    a = {to_d(tk[1]), kDeg2RadR * to_d(tk[2]), tk.size() >= 4 ? to_d(tk[3]) : 0.0, tk.size() >= 5 ? to_d(tk[4]) : 0.0};
  });
  read_types(S, "TORSION_TYPES", f.qtorlib, ht, [](const std::vector<std::string>& tk, QTorLib& t) {
    if (tk.size() < 4) throw std::runtime_error("bad [torsion_types] line");
    t = {to_d(tk[1]), to_d(tk[2]), kDeg2RadR * to_d(tk[3])};
  });
  read_types(S, "IMPROPER_TYPES", f.qimplib, hi, [](const std::vector<std::string>& tk, QImpLib& t) {
    if (tk.size() < 3) throw std::runtime_error("bad [improper_types] line");
    t = {to_d(tk[1]), kDeg2RadR * to_d(tk[2])};
  });
  for (auto& a : f.qanglib)
    if (a.ureyfk != 0.0) throw std::runtime_error("Urey-Bradley Q-angle terms (CHARMM) are not supported");
  read_changes(S, "CHANGE_BONDS", 2, f.nstates, natom, hb, f.qbnd);
  read_changes(S, "CHANGE_ANGLES", 3, f.nstates, natom, ha, f.qang);
  read_changes(S, "CHANGE_TORSIONS", 4, f.nstates, natom, ht, f.qtor);
  read_changes(S, "CHANGE_IMPROPERS", 4, f.nstates, natom, hi, f.qimp);
  read_couplings(S, "ANGLE_COUPLINGS", (int)f.qang.size(), f, f.ang_coupl);
  read_couplings(S, "TORSION_COUPLINGS", (int)f.qtor.size(), f, f.tor_coupl);
  read_couplings(S, "IMPROPER_COUPLINGS", (int)f.qimp.size(), f, f.imp_coupl);
}

FepData read_fep(const std::string& path, int natom) {
  Sections S(path);
  FepData f;
  std::string v;
  f.nstates = 1;
  if (S.get("FEP", "states", v)) f.nstates = (int)to_i(v);
  if (f.nstates < 1 || f.nstates > kMaxStates) throw std::runtime_error("bad number of FEP states");
  if (S.get("FEP", "offset", v) || S.get("FEP", "offset_name", v) || S.get("FEP", "offset_residue", v))
    throw std::runtime_error("FEP atom offsets are not supported");
  if (!S.has("ATOMS")) throw std::runtime_error("FEP file has no [atoms] section");
  for (auto& l : S.sec.at("ATOMS")) {
    std::vector<std::string> tk;
    Records::tokens(l, tk);
    if (tk.size() < 2) continue;
    int iq = (int)to_i(tk[0]);
    int at = (int)to_i(tk[1]);
    if (iq != (int)f.iqseq.size() + 1) throw std::runtime_error("Q atoms must be numbered 1..n in order");
    if (at < 1 || at > natom) throw std::runtime_error("invalid Q-atom topology number");
    f.iqseq.push_back(at - 1);
  }
  int nq = (int)f.iqseq.size();
  f.qcrg.assign(nq, std::vector<double>(f.nstates, 0.0));
  if (S.get("FEP", "qq_use_library_charges", v)) parse_onoff(v, f.qq_lib_charges);
  // the Q atoms' switching atom for periodic boxes (src/qatom.f90:672-690)
  if (S.get("PBC", "switching_atom", v)) {
    const int q = (int)to_i(v);
    if (q < 1 || q > natom) throw std::runtime_error("invalid PBC switching atom");
    f.qswitch = q - 1;
  }
  // monitored atom groups and group pairs (qatom_load_fep, src/qatom.f90:1987-2020)
  if (S.has("MONITOR_GROUPS"))
    for (auto& l : S.sec.at("MONITOR_GROUPS")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      std::vector<int> g;
      for (auto& t : tk) {
        const int a = (int)to_i(t);
        if (a < 1 || a > natom) throw std::runtime_error("bad atom in [monitor_groups]");
        g.push_back(a - 1);
      }
      f.mon_groups.push_back(g);
    }
  if (S.has("MONITOR_GROUP_PAIRS"))
    for (auto& l : S.sec.at("MONITOR_GROUP_PAIRS")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      if (tk.size() < 2) throw std::runtime_error("bad [monitor_group_pairs] line");
      const int gi = (int)to_i(tk[0]), gj = (int)to_i(tk[1]);
      if (gi < 1 || gj < 1 || gi > (int)f.mon_groups.size() || gj > (int)f.mon_groups.size())
        throw std::runtime_error("[monitor_group_pairs] names an undefined group");
      f.mon_pairs.push_back({gi - 1, gj - 1});
    }
  if (S.has("CHANGE_CHARGES")) {
    f.have_charges = true;
    for (auto& l : S.sec.at("CHANGE_CHARGES")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      int iq = (int)to_i(tk[0]);
      if (iq < 1 || iq > nq || (int)tk.size() < 1 + f.nstates) throw std::runtime_error("bad [change_charges] line");
      for (int s = 0; s < f.nstates; ++s) f.qcrg[iq - 1][s] = to_d(tk[1 + s]);
    }
  }
  // Q-atom type library: name, (A, B) for LJ codes 1..3, mass
  if (S.has("ATOM_TYPES")) {
    for (auto& l : S.sec.at("ATOM_TYPES")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      if (tk.size() < 8) throw std::runtime_error("bad [atom_types] line in " + path);
      f.qtac.push_back(tk[0]);
      f.qavdw.push_back({to_d(tk[1]), to_d(tk[3]), to_d(tk[5])});
      f.qbvdw.push_back({to_d(tk[2]), to_d(tk[4]), to_d(tk[6])});
      f.qmass.push_back(to_d(tk[7]));
    }
  }
  // per-state Q-atom types: a line for every Q atom (qvdw_flag)
  if (S.has("CHANGE_ATOMS") && !S.sec.at("CHANGE_ATOMS").empty()) {
    if (f.qtac.empty())
      throw std::runtime_error("[change_atoms] needs an [atom_types] library (topology types are not supported)");
    f.qvdw = true;
    f.qiac.assign(nq, std::vector<int>(f.nstates, -1));
    for (auto& l : S.sec.at("CHANGE_ATOMS")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      int iq = (int)to_i(tk[0]);
      if (iq < 1 || iq > nq || (int)tk.size() < 1 + f.nstates) throw std::runtime_error("bad [change_atoms] line");
      for (int s = 0; s < f.nstates; ++s) {
        auto it = std::find(f.qtac.begin(), f.qtac.end(), tk[1 + s]);
        if (it == f.qtac.end()) throw std::runtime_error("Q-atom type " + tk[1 + s] + " has not been defined");
        f.qiac[iq - 1][s] = (int)(it - f.qtac.begin());
      }
    }
    for (auto& r : f.qiac)
      for (int t : r)
        if (t < 0) throw std::runtime_error("Atom types of Q-atoms must be given for every Q-atom");
  }
  // soft-core alphas for every Q atom and state
  if (S.get("FEP", "softcore_use_max_potential", v)) parse_onoff(v, f.softcore_max_potential);
  if (S.has("SOFTCORE") && !S.sec.at("SOFTCORE").empty()) {
    if (!f.qvdw) throw std::runtime_error("Q-atom types must be redefined in [change_atoms] for [softcore]");
    f.softcore = true;
    f.alpha_max.assign(nq, std::vector<double>(f.nstates, 0.0));
    int n = 0;
    for (auto& l : S.sec.at("SOFTCORE")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      int iq = (int)to_i(tk[0]);
      if (iq < 1 || iq > nq || (int)tk.size() < 1 + f.nstates) throw std::runtime_error("bad [softcore] line");
      for (int s = 0; s < f.nstates; ++s) f.alpha_max[iq - 1][s] = to_d(tk[1 + s]);
      ++n;
    }
    if (n != nq) throw std::runtime_error("Alpha must be given for every Q-atom");
  }
  read_fep_bonded(S, f, natom);
  // [soft_pairs] "iq jq" (src/qatom.f90:832-858)
  auto rows = [&](const char* name) {
    std::vector<std::vector<std::string>> out;
    if (!S.has(name)) return out;
    for (auto& l : S.sec.at(name)) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (!tk.empty()) out.push_back(tk);
    }
    return out;
  };
  for (auto& tk : rows("SOFT_PAIRS")) {
    int a = (int)to_i(tk.at(0)), b = tk.size() > 1 ? (int)to_i(tk[1]) : 0;
    if (a < 1 || a > nq || b < 1 || b > nq) throw std::runtime_error("Invalid q-atom number in [soft_pairs]");
    f.soft_pairs.push_back({a - 1, b - 1});
  }
  // [el_scale] "iq jq scale(1..nstates)" (src/qatom.f90:862-888)
  for (auto& tk : rows("EL_SCALE")) {
    if ((int)tk.size() < 2 + f.nstates) throw std::runtime_error("bad [el_scale] line");
    int a = (int)to_i(tk[0]), b = (int)to_i(tk[1]);
    if (a < 1 || a > nq || b < 1 || b > nq) throw std::runtime_error("Invalid q-atom number in [el_scale]");
    FepData::ElScale e{a - 1, b - 1, {}};
    for (int st = 0; st < f.nstates; ++st) e.scale.push_back(to_d(tk[2 + st]));
    f.el_scale.push_back(e);
  }
  // [shake_constraints] "i j distance(1..nstates)" (src/qatom.f90:1792-1806)
  for (auto& tk : rows("SHAKE_CONSTRAINTS")) {
    if ((int)tk.size() < 2 + f.nstates) throw std::runtime_error("bad [shake_constraints] line");
    FepData::QShake q{(int)to_i(tk[0]) - 1, (int)to_i(tk[1]) - 1, {}};
    if (q.i < 0 || q.i >= natom || q.j < 0 || q.j >= natom) throw std::runtime_error("invalid atom in [shake_constraints]");
    for (int st = 0; st < f.nstates; ++st) q.dist.push_back(to_d(tk[2 + st]));
    f.qshake.push_back(q);
  }
  // [excluded_pairs] "i j flag(1..nstates)", flags 0/1 (src/qatom.f90:891-918)
  for (auto& tk : rows("EXCLUDED_PAIRS")) {
    if ((int)tk.size() < 2 + f.nstates) throw std::runtime_error("bad [excluded_pairs] line");
    FepData::ExSpec e{(int)to_i(tk[0]) - 1, (int)to_i(tk[1]) - 1, {}};
    for (int st = 0; st < f.nstates; ++st) {
      long fl = to_i(tk[2 + st]);
      if (fl != 0 && fl != 1) throw std::runtime_error("Special exclusion state flags are invalid.");
      e.flag.push_back((char)fl);
    }
    f.exspec.push_back(e);
  }
  // EVB off-diagonal functions
  if (S.has("OFF_DIAGONALS")) {
    for (auto& l : S.sec.at("OFF_DIAGONALS")) {
      std::vector<std::string> tk;
      Records::tokens(l, tk);
      if (tk.empty()) continue;
      if (tk.size() < 6) throw std::runtime_error("bad [off_diagonals] line");
      OffDiag o{(int)to_i(tk[0]) - 1, (int)to_i(tk[1]) - 1, (int)to_i(tk[2]) - 1, (int)to_i(tk[3]) - 1,
                to_d(tk[4]), to_d(tk[5])};
      if (o.i < 0 || o.i >= f.nstates || o.j < 0 || o.j >= f.nstates || o.i == o.j)
        throw std::runtime_error("Invalid combination of states in [off_diagonals]");
      if (o.k < 0 || o.k >= nq || o.l < 0 || o.l >= nq) throw std::runtime_error("Invalid q-atom number in [off_diagonals]");
      f.offd.push_back(o);
    }
    if ((int)f.offd.size() > kMaxStates) throw std::runtime_error("Too many off-diagonal functions");
  }
  return f;
}

// --------------------------------------------------------------------------
// restart (Fortran unformatted sequential records, 4-byte markers)
namespace {
bool read_rec(FILE* fp, std::vector<char>& buf) {
  int32_t n;
  if (std::fread(&n, 4, 1, fp) != 1) return false;
  buf.resize(n);
  if (n && std::fread(buf.data(), 1, n, fp) != (size_t)n) return false;
  int32_t n2;
  if (std::fread(&n2, 4, 1, fp) != 1 || n2 != n) return false;
  return true;
}
void write_rec(FILE* fp, const void* p, int32_t n) {
  std::fwrite(&n, 4, 1, fp);
  if (n) std::fwrite(p, 1, n, fp);
  std::fwrite(&n, 4, 1, fp);
}
}  // namespace

bool read_restart(const std::string& path, int natom, std::vector<double>& x,
                  std::vector<double>& v, std::vector<double>* theta_corr, double* box) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("cannot open restart " + path);
  std::vector<char> b;
  if (!read_rec(fp, b) || b.size() != 4) { std::fclose(fp); throw std::runtime_error("bad restart header"); }
  int32_t canary;
  std::memcpy(&canary, b.data(), 4);
  if (canary != -1337) { std::fclose(fp); throw std::runtime_error("restart is not double precision (-1337)"); }
  auto arr = [&](std::vector<double>& out) {
    if (!read_rec(fp, b)) throw std::runtime_error("short restart");
    int32_t nat3;
    std::memcpy(&nat3, b.data(), 4);
    if (nat3 != 3 * natom) throw std::runtime_error("wrong number of atoms in restart file");
    out.resize(nat3);
    std::memcpy(out.data(), b.data() + 4, 8 * (size_t)nat3);
  };
  arr(x);
  arr(v);
  if (box) {
    // box lengths, then box centre (3 doubles each)
    for (int k = 0; k < 2; ++k) {
      if (!read_rec(fp, b) || b.size() != 24) { std::fclose(fp); throw std::runtime_error("restart has no box record"); }
      std::memcpy(box + 3 * k, b.data(), 24);
    }
  } else if (theta_corr) {
    theta_corr->clear();
    if (read_rec(fp, b) && b.size() >= 4) {
      int32_t n;
      std::memcpy(&n, b.data(), 4);
      if (b.size() == 4 + 8 * (size_t)n) {
        theta_corr->resize(n);
        std::memcpy(theta_corr->data(), b.data() + 4, 8 * (size_t)n);
      }
    }
  }
  std::fclose(fp);
  return true;
}

void write_restart(const std::string& path, const std::vector<double>& x,
                   const std::vector<double>& v, const std::vector<double>* theta_corr, const double* box) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw std::runtime_error("cannot write " + path);
  int32_t canary = -1337;
  write_rec(fp, &canary, 4);
  for (const auto* a : {&x, &v}) {
    std::vector<char> b(4 + 8 * a->size());
    int32_t n = (int32_t)a->size();
    std::memcpy(b.data(), &n, 4);
    std::memcpy(b.data() + 4, a->data(), 8 * a->size());
    write_rec(fp, b.data(), (int32_t)b.size());
  }
  if (theta_corr && !theta_corr->empty()) {
    std::vector<char> b(4 + 8 * theta_corr->size());
    int32_t n = (int32_t)theta_corr->size();
    std::memcpy(b.data(), &n, 4);
    std::memcpy(b.data() + 4, theta_corr->data(), 8 * theta_corr->size());
    write_rec(fp, b.data(), (int32_t)b.size());
  }
  if (box) {
    write_rec(fp, box, 24);
    write_rec(fp, box + 3, 24);
  }
  std::fclose(fp);
}

}  // namespace qsys
