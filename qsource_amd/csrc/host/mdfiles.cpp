// qdyn's trajectory output: the atom mask of [trajectory_atoms]
// (src/mask.f90:129-361) and the DCD file of the trj module
// (src/trj.f90:72-233), for a prepared Setup.
#include "qsys.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace qsys {

namespace {
bool to_int(const std::string& w, int& v) {
  char* e = nullptr;
  long r = std::strtol(w.c_str(), &e, 10);
  if (e == w.c_str() || *e != '\0') return false;
  v = (int)r;
  return true;
}

// One mask row: property words (ALL, NOT/-, RESIDUE, SOLUTE, EXCLUDED,
// RESTRAINED, HEAVY) and an optional first [last] atom or residue range.
// A property asked for plainly selects atoms that have it, after NOT those
// that do not; a property not named selects both.
void mask_add(const Setup& s, std::vector<char>& m, const std::string& line) {
  struct Prop { bool a = true, b = false; };   // (a, b) as set%x1, set%x2
  Prop solute, excl, restr, heavy;
  bool notflag = false, empty = true, resnum = false;
  int first = 1, last = s.natom;
  const int nres = (int)s.top.res_start.size();
  std::istringstream in(line);
  std::string w;
  auto set = [&](Prop& p) { p.a = p.b = !notflag; empty = false; notflag = false; };
  while (in >> w) {
    std::string u = w;
    for (auto& c : u) c = (char)std::toupper((unsigned char)c);
    if (u == "ALL") { notflag = false; empty = false; }
    else if (u == "NOT" || u == "-") notflag = true;
    else if (u == "RESIDUE" || u == "RES" || u == "RESI") { resnum = true; notflag = false; }
    else if (u == "SOLUTE") set(solute);
    else if (u == "EXCLUDED" || u == "EXCL") set(excl);
    else if (u == "RESTRAINED") set(restr);
    else if (u == "HEAVY") set(heavy);
    else if (u == "SYBYL") throw std::runtime_error("SYBYL trajectory masks are not supported");
    else {
      if (!to_int(w, first)) throw std::runtime_error("unrecognised trajectory mask property " + w);
      if (resnum ? (first < 1 || first > nres) : (first < 1 || first > s.natom))
        throw std::runtime_error("invalid first atom or residue in trajectory mask: " + w);
      std::string w2;
      empty = false;
      if (!(in >> w2)) { last = first; break; }
      if (!to_int(w2, last) || (resnum && last > nres) || last < first || last > s.natom)
        throw std::runtime_error("invalid last atom or residue in trajectory mask: " + w2);
      break;
    }
  }
  if (empty) return;
  if (resnum) {
    first = s.top.res_start[first - 1] + 1;
    last = last == nres ? s.natom : s.top.res_start[last];
  }
  auto ok = [](const Prop& p, bool v) { return v == p.a || v == p.b; };
  for (int i = first - 1; i < last; ++i) {
    const bool sol = i < s.nat_solute, ex = s.excl[i] != 0, sh = s.shell[i] != 0 || ex, hv = s.heavy[i] != 0;
    if (ok(solute, sol) && ok(excl, ex) && ok(restr, sh) && ok(heavy, hv)) m[i] = 1;
  }
}

void rec(FILE* fp, const void* p, int32_t n) {
  std::fwrite(&n, 4, 1, fp);
  if (n) std::fwrite(p, 1, n, fp);
  std::fwrite(&n, 4, 1, fp);
}

std::string pad80(const std::string& t) {
  std::string r = t.substr(0, std::min<size_t>(80, t.size()));
  r.resize(80, ' ');
  return r;
}
}  // namespace

std::vector<char> trajectory_mask(const Setup& s, const std::vector<std::string>& rows) {
  std::vector<char> m(s.natom, 0);
  if (rows.empty()) mask_add(s, m, "all");
  for (const auto& r : rows) mask_add(s, m, r);
  return m;
}

DcdWriter::DcdWriter(const Setup& s, const std::string& path) {
  const MDInput& in = s.inp;
  if (in.itrj_cycle <= 0) throw std::runtime_error("no trajectory interval in the input");
  std::vector<std::string> rows = in.trajectory_atoms;
  if (rows.empty()) rows.push_back("all");
  if (rows.size() > 20) throw std::runtime_error("Too many trajectory atom mask rows (max 20).");
  std::vector<char> m = trajectory_mask(s, in.trajectory_atoms);
  for (int i = 0; i < s.natom; ++i)
    if (m[i]) atoms_.push_back(i);
  fp_ = std::fopen(path.c_str(), "wb");
  if (!fp_) throw std::runtime_error("failed to open trajectory " + path);
  // REC1: 'CORD', frames, steps before, interval, steps, velocity check,
  // 2 unused, degrees of freedom, fixed atoms, unused, constant p, 8 unused,
  // CHARMM version
  int32_t r1[20] = {0};
  r1[0] = in.nsteps / in.itrj_cycle;
  r1[1] = in.itrj_cycle;
  r1[2] = in.itrj_cycle;
  r1[3] = in.nsteps;
  r1[7] = s.ndegfree;
  std::string b1 = "CORD";
  b1.append(reinterpret_cast<const char*>(r1), sizeof(r1));
  rec(fp_, b1.data(), (int32_t)b1.size());
  int32_t nt = 2 + (int32_t)rows.size();
  std::string b2(reinterpret_cast<const char*>(&nt), 4);
  b2 += pad80("Q DCD trajectory version 5.6");
  b2 += pad80(in.top_file);
  for (const auto& r : rows) b2 += pad80(r);
  rec(fp_, b2.data(), (int32_t)b2.size());
  int32_t na = (int32_t)atoms_.size();
  rec(fp_, &na, 4);
}

DcdWriter::~DcdWriter() {
  if (fp_) std::fclose(fp_);
}

void DcdWriter::write(const double* x) {
  std::vector<float> c(atoms_.size());
  for (int d = 0; d < 3; ++d) {
    for (size_t k = 0; k < atoms_.size(); ++k) c[k] = (float)x[3 * atoms_[k] + d];
    rec(fp_, c.data(), (int32_t)(4 * c.size()));
  }
  std::fflush(fp_);
}

}  // namespace qsys
