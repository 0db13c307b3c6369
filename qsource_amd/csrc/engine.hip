// Host side of the engine: uploads prepared qdyn systems into the HBM layout
// of engine.hpp, sequences the kernels of one MD step for a whole batch, and
// exports the C ABI declared in include/qgpu.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qgpu.h"
#include "engine.hpp"

namespace qgpu {

// kernels (kernels_water.hip, kernels_solute.hip)
__global__ void k_cgp_centers(const DevSys*);
__global__ void k_water_lists(const DevSys*);
template <typename P> __global__ void k_water_forces(const DevSys*);
__global__ void k_watpol_theta(const DevSys*);
__global__ void k_watpol_force(const DevSys*, int);
__global__ void k_integrate(const DevSys*);
__global__ void k_temperature(const DevSys*, int);
__global__ void k_energy_reduce(const DevSys*);

#define HIPCHECK(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Batch::Inst {
  std::vector<void*> allocs;
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    HIPCHECK(hipMemset(p, 0, n * sizeof(T)));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  template <typename T>
  T* upload(const std::vector<T>& v) {
    T* p = alloc<T>(v.size());
    if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  ~Inst() {
    for (void* p : allocs) (void)hipFree(p);
  }
  int ke_blocks = 0;
  int natom = 0, nwat = 0;
};

// Initial temperature() of md_run (src/md.f90:6400-6404): Berendsen scale
// factors for the first step from the starting velocities.
static void initial_scales(const qsys::Setup& s, double* scal) {
  double temp_solute = 0, tfree_solute = 0, temp_solvent = 0, tfree_solvent = 0;
  for (int i = 0; i < s.natom; ++i) {
    const double* v = &s.v[3 * i];
    double ek = 0.5 * s.top.types[s.iac[i]].mass * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (i < s.nat_solute) { temp_solute += ek; if (!s.excl[i]) tfree_solute += ek; }
    else { temp_solvent += ek; if (!s.excl[i]) tfree_solvent += ek; }
  }
  double tfree = tfree_solvent + tfree_solute;
  double temp = temp_solute + temp_solvent;
  scal[4] = temp;
  temp = 2.0 * temp / qsys::kBoltz / (double)s.ndegf;
  tfree = 2.0 * tfree / qsys::kBoltz / (double)s.ndegfree;
  if (s.detail_temps) {
    tfree_solute = 2.0 * tfree_solute / qsys::kBoltz / (double)s.ndegfree_solute;
    tfree_solvent = 2.0 * tfree_solvent / qsys::kBoltz / (double)s.ndegfree_solvent;
  }
  double tsu = 0.0, tsv = 0.0;
  const double r = s.dt / s.tau_t, T0 = s.inp.temp0;
  if (s.separate_scaling) {
    if (tfree_solvent != 0) tsv = T0 / tfree_solvent - 1.0;
    tsv = std::sqrt(1.0 + r * tsv);
    if (tfree_solute != 0) tsu = T0 / tfree_solute - 1.0;
    tsu = std::sqrt(1.0 + r * tsu);
  } else {
    if (tfree != 0) tsv = T0 / tfree - 1.0;
    tsv = std::sqrt(1.0 + r * tsv);
    tsu = tsv;
  }
  scal[0] = tsu; scal[1] = tsv; scal[2] = tfree; scal[3] = temp;
}

Batch::Batch(const std::vector<std::string>& inputs, int precision, int device) {
  if (inputs.empty()) throw std::runtime_error("no instances");
  if (precision != 0 && precision != 1) throw std::runtime_error("precision must be QGPU_FP64 or QGPU_MIXED");
  precision_ = precision;
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&ev0));
  HIPCHECK(hipEventCreate(&ev1));
  for (const auto& f : inputs) setups_.push_back(new qsys::Setup(qsys::prepare(f)));
  for (auto* s : setups_) {
    max_nwat_ = std::max(max_nwat_, s->nwat);
    max_ncgp_ = std::max(max_ncgp_, (int)s->top.cgp.size());
    max_ncgp_solute_ = std::max(max_ncgp_solute_, s->top.ncgp_solute);
    max_natom_ = std::max(max_natom_, s->natom);
  }
  h_sys_.resize(setups_.size());
  for (size_t k = 0; k < setups_.size(); ++k) upload((int)k);
  HIPCHECK(hipMalloc(&d_sys_, h_sys_.size() * sizeof(DevSys)));
  HIPCHECK(hipMemcpy(d_sys_, h_sys_.data(), h_sys_.size() * sizeof(DevSys), hipMemcpyHostToDevice));
}

Batch::~Batch() {
  for (auto* i : inst_) delete i;
  for (auto* s : setups_) delete s;
  if (d_sys_) (void)hipFree(d_sys_);
  if (stream_) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    (void)hipStreamDestroy(stream_);
  }
}

void Batch::upload(int k) {
  const qsys::Setup& s = *setups_[k];
  const qsys::Topology& t = s.top;
  if (s.nat_solute > 0) throw std::runtime_error("solute atoms are not yet supported by the GPU engine");
  if (s.nqat > 0) throw std::runtime_error("Q atoms are not yet supported by the GPU engine");
  if (s.inp.freeze) throw std::runtime_error("excluded_freeze is not supported");
  if (!s.inp.use_lrf) throw std::runtime_error("runs without LRF are not supported");
  auto* I = new Inst();
  inst_.push_back(I);
  DevSys& D = h_sys_[k];
  std::memset(&D, 0, sizeof(DevSys));
  D.natom = s.natom; D.nat_solute = s.nat_solute; D.nwat = s.nwat;
  D.ncgp = (int)t.cgp.size(); D.ncgp_solute = t.ncgp_solute;
  D.nqat = s.nqat; D.nstates = s.nstates;
  D.precision = precision_;
  const int N = s.natom;
  D.x = I->upload(s.x);
  D.v = I->upload(s.v);
  D.xx = I->alloc<double>(3 * N);
  D.d = I->alloc<double>(3 * N);
  D.xtop = I->upload(s.xtop);
  D.crg = I->upload(s.crg);
  D.winv = I->upload(s.winv);
  std::vector<double> mass(N);
  std::vector<unsigned char> flags(N);
  for (int i = 0; i < N; ++i) {
    mass[i] = t.types[s.iac[i]].mass;
    flags[i] = (s.excl[i] ? F_EXCL : 0) | (s.shell[i] ? F_SHELL : 0) | (s.iqatom[i] ? F_Q : 0) |
               (s.heavy[i] ? F_HEAVY : 0);
  }
  D.mass = I->upload(mass);
  D.flags = I->upload(flags);
  D.iac = I->upload(s.iac);
  std::vector<int> cf, cl, cs;
  for (auto& g : t.cgp) { cf.push_back(g.first); cl.push_back(g.last); cs.push_back(g.iswitch); }
  D.cgp_first = I->upload(cf);
  D.cgp_last = I->upload(cl);
  D.cgp_sw = I->upload(cs);
  D.cgpatom = I->upload(t.cgpatom);
  D.iwhich_cgp = I->upload(s.iwhich_cgp);
  D.lrf = I->alloc<double>((size_t)D.ncgp * kLrfStride);
  D.natyps = s.natyps;
  std::vector<double> av, bv;
  for (auto& ty : t.types) for (int j = 0; j < 3; ++j) { av.push_back(ty.avdw[j]); bv.push_back(ty.bvdw[j]); }
  D.avdw = I->upload(av);
  D.bvdw = I->upload(bv);
  D.ljcod = I->upload(s.ljcod);
  D.el14_scale = t.el14_scale;
  // lists
  D.ww_cap = std::max(1, std::min(s.nwat, 448));
  D.wp_cap = std::max(1, std::min(t.ncgp_solute, 1024));
  D.pp_cap = std::max(1, t.ncgp_solute);
  D.pw_cap = std::max(1, std::min(s.nwat, 448));
  D.ww_cnt = I->alloc<int>(s.nwat);
  D.ww_nb = I->alloc<int>((size_t)s.nwat * D.ww_cap);
  D.wp_cnt = I->alloc<int>(s.nwat);
  D.wp_nb = I->alloc<int>((size_t)s.nwat * D.wp_cap);
  D.pp_cnt = I->alloc<int>(t.ncgp_solute);
  D.pp_nb = I->alloc<int>((size_t)t.ncgp_solute * D.pp_cap);
  D.pw_cnt = I->alloc<int>(t.ncgp_solute);
  D.pw_nb = I->alloc<int>((size_t)t.ncgp_solute * D.pw_cap);
  D.overflow = I->alloc<int>(1);
  D.rcpp2 = s.inp.rcpp * s.inp.rcpp;
  D.rcpw2 = s.inp.rcpw * s.inp.rcpw;
  D.rcww2 = s.inp.rcww * s.inp.rcww;
  D.rclrf2 = s.inp.rclrf * s.inp.rclrf;
  D.use_lrf = s.inp.use_lrf;
  // solvent
  if (s.nwat > 0) {
    for (int a = 0; a < 3; ++a) D.chg_solv[a] = s.chg_solv[a];
    int o = s.nat_solute;
    int lj = s.ljcod[s.iac[o] * s.natyps + s.iac[o]];
    D.aOO = s.alj_solv[0][lj - 1] * s.alj_solv[0][lj - 1];
    D.bOO = s.blj_solv[0][lj - 1] * s.blj_solv[0][lj - 1];
  }
  D.rwat = s.rwat; D.fk_wsphere = s.fk_wsphere; D.dwmz = s.dwmz; D.awmz = s.awmz; D.fkwpol = s.fkwpol;
  for (int c = 0; c < 3; ++c) D.xwcent[c] = t.xwcent[c];
  D.wpol_restr = (s.wpol_restr && s.nwat > 0) ? 1 : 0;
  D.nshell = (int)s.wshell.size();
  std::vector<DevShell> sh;
  for (auto& w : s.wshell) sh.push_back({w.rout, w.dr, w.cstb, w.avtheta, w.avn_insh, w.theta_corr, 0, 0});
  D.shell = I->upload(sh);
  D.theta = I->alloc<double>(s.nwat);
  D.shell_of = I->alloc<int>(s.nwat);
  D.fk_fix = s.inp.fk_fix; D.fk_pshell = s.inp.fk_pshell; D.freeze = s.inp.freeze;
  // integrator
  D.dt = s.dt; D.temp0 = s.inp.temp0; D.tau_t = s.tau_t;
  D.separate_scaling = s.separate_scaling; D.detail_temps = s.detail_temps;
  D.ndegf = s.ndegf; D.ndegfree = s.ndegfree; D.ndegf_solute = s.ndegf_solute;
  D.ndegf_solvent = s.ndegf_solvent; D.ndegfree_solute = s.ndegfree_solute;
  D.ndegfree_solvent = s.ndegfree_solvent;
  // SHAKE units
  std::vector<int> ufirst, uatoms, cfirst, shi, shj;
  std::vector<double> shd;
  std::vector<char> used(N, 0);
  for (auto& m : s.shake_mol) {
    if (m.bond.size() > 64) throw std::runtime_error("SHAKE molecules with more than 64 constraints are not supported");
    ufirst.push_back((int)uatoms.size());
    cfirst.push_back((int)shi.size());
    std::vector<int> atoms;
    for (auto& c : m.bond) {
      for (int a : {c.i, c.j})
        if (!used[a]) { used[a] = 1; atoms.push_back(a); }
      shi.push_back(c.i); shj.push_back(c.j); shd.push_back(c.dist2);
    }
    std::sort(atoms.begin(), atoms.end());
    uatoms.insert(uatoms.end(), atoms.begin(), atoms.end());
  }
  for (int i = 0; i < N; ++i)
    if (!used[i]) {
      ufirst.push_back((int)uatoms.size());
      cfirst.push_back((int)shi.size());
      uatoms.push_back(i);
    }
  ufirst.push_back((int)uatoms.size());
  cfirst.push_back((int)shi.size());
  D.nunits = (int)ufirst.size() - 1;
  D.has_shake = s.shake_constraints > 0;
  max_units_ = std::max(max_units_, D.nunits);
  D.unit_first = I->upload(ufirst);
  D.unit_atoms = I->upload(uatoms);
  D.con_first = I->upload(cfirst);
  D.sh_i = I->upload(shi);
  D.sh_j = I->upload(shj);
  D.sh_d2 = I->upload(shd);
  D.shake_fail = I->alloc<int>(1);
  // energies and reductions
  D.e_water = I->alloc<double>((size_t)std::max(1, s.nwat) * 4);
  D.e_solute = I->alloc<double>((size_t)std::max(1, t.ncgp_solute) * 8);
  D.e_wpol = I->alloc<double>(std::max(1, D.nshell));
  I->ke_blocks = 0;
  D.ke_part = nullptr;  // sized after all instances are known (see below)
  std::vector<double> scal(8, 0.0);
  initial_scales(s, scal.data());
  D.scal = I->upload(scal);
  std::vector<double> e0(E_NTERMS, 0.0);
  e0[E_KINETIC] = scal[4];
  D.energies = I->upload(e0);
  I->natom = N; I->nwat = s.nwat;
  // the kinetic partial buffer needs the batch-wide unit count: allocate
  // generously from this instance's own count, re-checked at launch
  int blocks = (std::max(D.nunits, 1) + 255) / 256;
  I->ke_blocks = blocks;
  D.ke_part = I->alloc<double>((size_t)blocks * 6 + 6 * 4096);
  D.ke_blocks = blocks + 4096;
}

void Batch::launch_lists() {
  const int ni = (int)h_sys_.size();
  dim3 g0((max_ncgp_ + 255) / 256, ni);
  hipLaunchKernelGGL(k_cgp_centers, g0, dim3(256), 0, stream_, d_sys_);
  if (max_nwat_ > 0) {
    dim3 g1((max_nwat_ * 64 + 255) / 256, ni);
    hipLaunchKernelGGL(k_water_lists, g1, dim3(256), 0, stream_, d_sys_);
  }
}

void Batch::launch_forces() {
  const int ni = (int)h_sys_.size();
  if (max_nwat_ > 0) {
    dim3 g1((max_nwat_ * 64 + 255) / 256, ni);
    if (precision_ == 0)
      hipLaunchKernelGGL(k_water_forces<double>, g1, dim3(256), 0, stream_, d_sys_);
    else
      hipLaunchKernelGGL(k_water_forces<float>, g1, dim3(256), 0, stream_, d_sys_);
    dim3 g2((max_nwat_ + 255) / 256, ni);
    hipLaunchKernelGGL(k_watpol_theta, g2, dim3(256), 0, stream_, d_sys_);
    int update = (istep_ != 0 && istep_ % qsys::kItdisUpdate == 0) ? 1 : 0;
    dim3 g3(kNumShellMax, ni);
    hipLaunchKernelGGL(k_watpol_force, g3, dim3(1024), 0, stream_, d_sys_, update);
  }
}

void Batch::launch_integrate() {
  const int ni = (int)h_sys_.size();
  int blocks = (std::max(max_units_, 1) + 255) / 256;
  for (size_t k = 0; k < h_sys_.size(); ++k)
    if (blocks > h_sys_[k].ke_blocks) throw std::runtime_error("kinetic partial buffer too small");
  hipLaunchKernelGGL(k_integrate, dim3(blocks, ni), dim3(256), 0, stream_, d_sys_);
  hipLaunchKernelGGL(k_temperature, dim3(1, ni), dim3(256), 0, stream_, d_sys_, blocks);
}

void Batch::make_lists() { launch_lists(); }
void Batch::forces() { launch_forces(); }

void Batch::step(int nsteps) {
  const int nb = setups_[0]->inp.nbcycle;
  HIPCHECK(hipEventRecord(ev0, stream_));
  for (int n = 0; n < nsteps; ++n) {
    if (istep_ % nb == 0) launch_lists();
    launch_forces();
    launch_integrate();
    ++istep_;
  }
  HIPCHECK(hipEventRecord(ev1, stream_));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  for (size_t k = 0; k < h_sys_.size(); ++k) {
    int flags[2] = {0, 0};
    HIPCHECK(hipMemcpy(&flags[0], h_sys_[k].overflow, 4, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&flags[1], h_sys_[k].shake_fail, 4, hipMemcpyDeviceToHost));
    if (flags[0]) throw std::runtime_error("pair-list capacity exceeded");
    if (flags[1]) throw std::runtime_error("shake failure");
  }
}

void Batch::get_state(int k, double* x, double* v) {
  HIPCHECK(hipStreamSynchronize(stream_));
  size_t n = 3 * (size_t)h_sys_[k].natom;
  if (x) HIPCHECK(hipMemcpy(x, h_sys_[k].x, n * 8, hipMemcpyDeviceToHost));
  if (v) HIPCHECK(hipMemcpy(v, h_sys_[k].v, n * 8, hipMemcpyDeviceToHost));
}

void Batch::set_state(int k, const double* x, const double* v) {
  HIPCHECK(hipStreamSynchronize(stream_));
  size_t n = 3 * (size_t)h_sys_[k].natom;
  if (x) HIPCHECK(hipMemcpy(h_sys_[k].x, x, n * 8, hipMemcpyHostToDevice));
  if (v) HIPCHECK(hipMemcpy(h_sys_[k].v, v, n * 8, hipMemcpyHostToDevice));
}

void Batch::get_gradient(int k, double* d) {
  HIPCHECK(hipStreamSynchronize(stream_));
  HIPCHECK(hipMemcpy(d, h_sys_[k].d, 3 * (size_t)h_sys_[k].natom * 8, hipMemcpyDeviceToHost));
}

void Batch::energies(int k, double* out, int cap) {
  const int ni = (int)h_sys_.size();
  hipLaunchKernelGGL(k_energy_reduce, dim3(1, ni), dim3(256), 0, stream_, d_sys_);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  double E[E_NTERMS];
  HIPCHECK(hipMemcpy(E, h_sys_[k].energies, sizeof(E), hipMemcpyDeviceToHost));
  // bonded and Q terms are zero until their kernels exist
  double p_bond = 0, w_bond = 0, p_angle = 0, w_angle = 0, p_tor = 0, w_tor = 0, p_imp = 0, w_imp = 0;
  double q_bond = 0, q_angle = 0, q_tor = 0, q_imp = 0, qx_el = 0, qx_vdw = 0, rst_protein = 0;
  double rst_total = E[E_RST_FIX] + E[E_RST_SHELL] + rst_protein + E[E_RST_RADIAL] + E[E_RST_WPOL];
  double pot = p_bond + w_bond + p_angle + w_angle + p_tor + p_imp + E[E_PP_EL] + E[E_PP_VDW] +
               E[E_PW_EL] + E[E_PW_VDW] + E[E_WW_EL] + E[E_WW_VDW] + q_bond + q_angle + q_tor + q_imp +
               qx_el + qx_vdw + rst_total + E[E_LRF];
  double v[QGPU_E_COUNT] = {E[E_PP_EL], E[E_PP_VDW], p_bond, p_angle, p_tor, p_imp,
                            E[E_WW_EL], E[E_WW_VDW], w_bond, w_angle, w_tor, w_imp,
                            E[E_PW_EL], E[E_PW_VDW], E[E_LRF],
                            qx_el, qx_vdw, q_bond, q_angle, q_tor, q_imp,
                            rst_total, E[E_RST_FIX], E[E_RST_RADIAL], E[E_RST_WPOL], E[E_RST_SHELL],
                            rst_protein, pot + E[E_KINETIC], pot, E[E_KINETIC]};
  int n = std::min(cap, (int)QGPU_E_COUNT);
  std::memcpy(out, v, 8 * n);
}

void Batch::export_list(int k, int which, std::vector<int>& out) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  const qsys::Setup& s = *setups_[k];
  out.clear();
  if (which == 2) {
    std::vector<int> cnt(D.nwat), nb((size_t)D.nwat * D.ww_cap);
    HIPCHECK(hipMemcpy(cnt.data(), D.ww_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nb.data(), D.ww_nb, 4 * nb.size(), hipMemcpyDeviceToHost));
    const int ns = s.nat_solute;
    for (int iw = 0; iw < D.nwat; ++iw)
      for (int k2 = 0; k2 < cnt[iw]; ++k2) {
        int jw = nb[(size_t)iw * D.ww_cap + k2];
        bool own = (iw < jw) == (((iw + jw) & 1) == 0);
        if (!own) continue;
        for (int la = 0; la < 3; ++la)
          for (int ka = 0; ka < 3; ++ka) {
            out.push_back(ns + 3 * iw + la);
            out.push_back(ns + 3 * jw + ka);
            out.push_back(s.ljcod[s.iac[ns + la] * s.natyps + s.iac[ns + ka]]);
          }
      }
  } else if (which == 0 || which == 1) {
    // no solute lists yet
  } else {
    throw std::runtime_error("unknown list");
  }
}

double Batch::temperature(int k) {
  HIPCHECK(hipStreamSynchronize(stream_));
  double sc[4];
  HIPCHECK(hipMemcpy(sc, h_sys_[k].scal, sizeof(sc), hipMemcpyDeviceToHost));
  return sc[3];
}

void Batch::theta_corr(int k, double* out) {
  HIPCHECK(hipStreamSynchronize(stream_));
  std::vector<DevShell> sh(h_sys_[k].nshell);
  if (!sh.empty())
    HIPCHECK(hipMemcpy(sh.data(), h_sys_[k].shell, sh.size() * sizeof(DevShell), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < sh.size(); ++i) out[i] = sh[i].theta_corr;
}

}  // namespace qgpu

// ---------------------------------------------------------------------------
// C ABI
struct qgpu_handle {
  std::unique_ptr<qgpu::Batch> b;
  double last_ms = 0;
};

static thread_local std::string g_err;

template <typename F>
static int guarded(F&& f) {
  try {
    g_err.clear();
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

#define CHECK_K(h, k)                                                        \
  do {                                                                       \
    if (!(h) || (k) < 0 || (k) >= (h)->b->size())                            \
      throw std::runtime_error("invalid handle or instance index");          \
  } while (0)

extern "C" {

int qgpu_create(const char* const* input_files, int n_instances, int precision, qgpu_handle** out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("null output handle");
    if (n_instances < 1) throw std::runtime_error("need at least one instance");
    std::vector<std::string> in;
    for (int k = 0; k < n_instances; ++k) in.push_back(input_files[k]);
    auto* h = new qgpu_handle();
    try {
      h->b.reset(new qgpu::Batch(in, precision, 0));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int qgpu_destroy(qgpu_handle* h) {
  return guarded([&] { delete h; });
}

int qgpu_natom(qgpu_handle* h, int k, int* natom) {
  return guarded([&] { CHECK_K(h, k); *natom = h->b->setup(k).natom; });
}

int qgpu_make_pair_lists(qgpu_handle* h) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    h->b->make_lists();
  });
}

int qgpu_pot_energy(qgpu_handle* h) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    h->b->forces();
  });
}

int qgpu_step(qgpu_handle* h, int nsteps) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    if (nsteps < 0) throw std::runtime_error("negative step count");
    h->b->step(nsteps);
    float ms = 0;
    if (nsteps > 0) {
      HIPCHECK(hipEventElapsedTime(&ms, h->b->ev0, h->b->ev1));
      h->last_ms = ms / nsteps;
    }
  });
}

int qgpu_get_state(qgpu_handle* h, int k, double* x, double* v) {
  return guarded([&] { CHECK_K(h, k); h->b->get_state(k, x, v); });
}
int qgpu_set_state(qgpu_handle* h, int k, const double* x, const double* v) {
  return guarded([&] { CHECK_K(h, k); h->b->set_state(k, x, v); });
}
int qgpu_get_gradient(qgpu_handle* h, int k, double* d) {
  return guarded([&] { CHECK_K(h, k); h->b->get_gradient(k, d); });
}
int qgpu_energies(qgpu_handle* h, int k, double* e, int cap) {
  return guarded([&] { CHECK_K(h, k); h->b->energies(k, e, cap); });
}
int qgpu_qenergies(qgpu_handle* h, int k, double* e, int cap) {
  return guarded([&] {
    CHECK_K(h, k);
    (void)e; (void)cap;
    if (h->b->setup(k).nstates > 0) throw std::runtime_error("Q energies are not yet available");
  });
}
int qgpu_get_pairlist(qgpu_handle* h, int k, int which, int* out, long* n) {
  return guarded([&] {
    CHECK_K(h, k);
    std::vector<int> v;
    h->b->export_list(k, which, v);
    long cnt = (long)(v.size() / 3);
    if (out) {
      if (*n < cnt) throw std::runtime_error("output buffer too small");
      std::memcpy(out, v.data(), 4 * v.size());
    }
    *n = cnt;
  });
}
int qgpu_temperature(qgpu_handle* h, int k, double* temp) {
  return guarded([&] { CHECK_K(h, k); *temp = h->b->temperature(k); });
}
int qgpu_theta_corr(qgpu_handle* h, int k, double* out, int* n) {
  return guarded([&] {
    CHECK_K(h, k);
    *n = (int)h->b->setup(k).wshell.size();
    if (out) h->b->theta_corr(k, out);
  });
}
int qgpu_last_step_ms(qgpu_handle* h, double* ms) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    *ms = h->last_ms;
  });
}
const char* qgpu_last_error(void) { return g_err.c_str(); }

}  // extern "C"
