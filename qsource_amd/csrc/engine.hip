// Host side of the engine: uploads prepared qdyn systems into the HBM layout
// of engine.hpp, sequences the kernels of one MD step for a whole batch, and
// exports the C ABI declared in include/qgpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qgpu.h"
#include "engine.hpp"
#include "mlp_host.hpp"

namespace qgpu {

// kernels (kernels_water.hip, kernels_solute.hip)
__global__ void k_cgp_centers(const DevSys*);
__global__ void k_x_to_f32(const DevSys*);
template <typename P, int MAPPED> __global__ void k_water_lists(const DevSys*, const int2*);
template <typename P, int MAPPED, int PF, int W, int AA> __global__ void k_water_forces(const DevSys*, const int2*, int);
__global__ void k_watpol_theta(const DevSys*);
__global__ void k_watpol_force(const DevSys*);
__global__ void k_watpol_shell(const DevSys*);
__global__ void k_watpol_gather(const DevSys*);
__global__ void k_integrate(const DevSys*);
__global__ void k_shake(const DevSys*, const double*, double*, int*);
__global__ void k_temperature(const DevSys*, int);
__global__ void k_energy_reduce(const DevSys*, int);
template <typename P, int MAPPED> __global__ void k_solute_lists(const DevSys*, const int2*);
template <typename P, int MAPPED, int PF> __global__ void k_solute_forces(const DevSys*, const int2*, int);
template <typename P, int NS> __global__ void k_q_forces(const DevSys*, int);
__global__ void k_q_reduce(const DevSys*, int);
__global__ void k_q_lists(const DevSys*, int);
__global__ void k_mc_scale(const DevSys*);
__global__ void k_nh_prop(const DevSys*);
__global__ void k_put_back(const DevSys*);
__global__ void k_monitor(const DevSys*);
__global__ void k_freeze_x(const DevSys*);
__global__ void k_terms(const DevSys*);
__global__ void k_term_gather(const DevSys*);
__global__ void k_far_features(const DevSys*);
__global__ void k_far_targets(const DevSys*);
__global__ void k_far_lrf_eval(const DevSys*);
__global__ void k_far_apply(const DevSys*, const float*, const float*);

#define HIPCHECK(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device charge groups.  The solute kernels hold the atom pairs of two
// groups as 8 x 8 bit masks, so a solute group of more than kMaxGroupAtoms
// atoms is split into consecutive slices of at most kMaxGroupAtoms atoms.
// Every slice keeps its group's switch atom, so that each list decision
// (cut-off or LRF, nbpplist_lrf src/md.f90:7998-8188) is the group's, and
// its group's atom range for the centre (cgp_centers, :2253-2288), so that
// the LRF expansion of every slice is taken about the group's centre; the
// slice's own atoms are the ones it brings to pairs and to the LRF sums of
// other groups.  Water groups follow unchanged.
struct DevGroups {
  std::vector<int> first, last, sw, cfirst, clast, ref;   // per device group
  std::vector<int> which;       // atom -> device group
  std::vector<int> dev_first;   // reference group -> its first device group (+ end)
  int nsolute = 0;
};
static DevGroups device_groups(const qsys::Setup& s) {
  const qsys::Topology& t = s.top;
  DevGroups G;
  G.which.assign(std::max(1, s.natom), -1);
  for (int g = 0; g < (int)t.cgp.size(); ++g) {
    const auto& c = t.cgp[g];
    G.dev_first.push_back((int)G.first.size());
    const int step = g < t.ncgp_solute ? kMaxGroupAtoms : std::max(1, c.last - c.first + 1);
    for (int f = c.first; f <= c.last; f += step) {
      const int l = std::min(c.last, f + step - 1);
      const int d = (int)G.first.size();
      G.first.push_back(f); G.last.push_back(l); G.sw.push_back(c.iswitch);
      G.cfirst.push_back(c.first); G.clast.push_back(c.last); G.ref.push_back(g);
      for (int a = f; a <= l; ++a) G.which[t.cgpatom[a]] = d;
    }
    if (g + 1 == t.ncgp_solute) G.nsolute = (int)G.first.size();
  }
  G.dev_first.push_back((int)G.first.size());
  for (int i = 0; i < s.natom; ++i)
    if (G.which[i] < 0) G.which[i] = G.dev_first[s.iwhich_cgp[i]];
  return G;
}

struct Batch::Inst {
  std::vector<void*> allocs;
  DevGroups groups;
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc(&p, n * sizeof(T)));
    HIPCHECK(hipMemset(p, 0, n * sizeof(T)));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  // frees one allocation made by alloc (buffers that are resized)
  void release(void* p) {
    if (!p) return;
    for (size_t k = 0; k < allocs.size(); ++k)
      if (allocs[k] == p) {
        (void)hipFree(p);
        allocs.erase(allocs.begin() + (long)k);
        return;
      }
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
  int eseries_alloc = -1;   // frames the energy-series buffers hold (-1: none yet)
};

// Initial temperature() of md_run (src/md.f90:6400-6404): Berendsen scale
// factors for the first step from the starting velocities.
static void initial_scales(const qsys::Setup& s, const double* vel, double* scal) {
  double temp_solute = 0, tfree_solute = 0, temp_solvent = 0, tfree_solvent = 0;
  for (int i = 0; i < s.natom; ++i) {
    const double* v = &vel[3 * i];
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

// Longest rows of the water-water, water-solute and solute-water lists at
// the starting coordinates (the switch-atom / water-O distance tests of the
// list kernels), counted over a uniform grid of cut-off sized cells.
struct ListRows { int ww = 0, wp = 0, pw = 0; };
static ListRows longest_rows(const qsys::Setup& s, const DevGroups& G) {
  ListRows out;
  const qsys::Topology& t = s.top;
  if (s.nwat == 0) return out;
  if (s.pbc) {
    // periodic box: every pair by its minimum image
    auto d2 = [&](int a, int b) {
      double r2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        double d = s.x[3 * a + c] - s.x[3 * b + c];
        d = d - s.boxl[c] * std::round(d * s.inv_boxl[c]);
        r2 += d * d;
      }
      return r2;
    };
    const double rww2 = s.inp.rcww * s.inp.rcww, rpw2 = s.inp.rcpw * s.inp.rcpw;
    std::vector<int> nww(s.nwat, 0), nwp(s.nwat, 0), npw(t.ncgp_solute, 0);
    for (int w = 0; w < s.nwat; ++w) {
      const int o = s.nat_solute + 3 * w;
      for (int v = w + 1; v < s.nwat; ++v)
        if (d2(o, s.nat_solute + 3 * v) <= rww2) { ++nww[w]; ++nww[v]; }
      for (int g = 0; g < t.ncgp_solute; ++g)
        if (d2(o, t.cgp[g].iswitch) <= rpw2) { nwp[w] += G.dev_first[g + 1] - G.dev_first[g]; ++npw[g]; }
    }
    for (int w = 0; w < s.nwat; ++w) { out.ww = std::max(out.ww, nww[w]); out.wp = std::max(out.wp, nwp[w]); }
    for (int g = 0; g < t.ncgp_solute; ++g) out.pw = std::max(out.pw, npw[g]);
    return out;
  }
  const double rww = s.inp.rcww, rpw = s.inp.rcpw;
  const double cell = std::max(1.0, std::max(rww, rpw));
  double lo[3] = {1e300, 1e300, 1e300};
  std::vector<int> ow;   // water O atoms that take part (not excluded)
  for (int w = 0; w < s.nwat; ++w) {
    const int o = s.nat_solute + 3 * w;
    if (s.excl[o]) continue;
    ow.push_back(w);
    for (int c = 0; c < 3; ++c) lo[c] = std::min(lo[c], s.x[3 * o + c]);
  }
  std::vector<int> sw;   // solute charge groups by switch atom
  for (int g = 0; g < t.ncgp_solute; ++g) {
    const int a = t.cgp[g].iswitch;
    if (s.excl[a]) continue;
    sw.push_back(g);
    for (int c = 0; c < 3; ++c) lo[c] = std::min(lo[c], s.x[3 * a + c]);
  }
  auto cell_of = [&](const double* x, int* ci) {
    for (int c = 0; c < 3; ++c) ci[c] = (int)std::floor((x[c] - lo[c]) / cell);
  };
  int dim[3] = {1, 1, 1};
  auto grow_dim = [&](const double* x) {
    int ci[3];
    cell_of(x, ci);
    for (int c = 0; c < 3; ++c) dim[c] = std::max(dim[c], ci[c] + 1);
  };
  for (int w : ow) grow_dim(&s.x[3 * (s.nat_solute + 3 * w)]);
  for (int g : sw) grow_dim(&s.x[3 * t.cgp[g].iswitch]);
  const size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
  std::vector<std::vector<int>> wcell(ncell), gcell(ncell);
  auto index = [&](const int* ci) { return ((size_t)ci[0] * dim[1] + ci[1]) * dim[2] + ci[2]; };
  for (int w : ow) { int ci[3]; cell_of(&s.x[3 * (s.nat_solute + 3 * w)], ci); wcell[index(ci)].push_back(w); }
  for (int g : sw) { int ci[3]; cell_of(&s.x[3 * t.cgp[g].iswitch], ci); gcell[index(ci)].push_back(g); }
  auto count = [&](const double* x, const std::vector<std::vector<int>>& cells, bool water, double rc, int self) {
    int ci[3];
    cell_of(x, ci);
    int n = 0;
    for (int a = ci[0] - 1; a <= ci[0] + 1; ++a)
      for (int b = ci[1] - 1; b <= ci[1] + 1; ++b)
        for (int c = ci[2] - 1; c <= ci[2] + 1; ++c) {
          if (a < 0 || b < 0 || c < 0 || a >= dim[0] || b >= dim[1] || c >= dim[2]) continue;
          const int cc[3] = {a, b, c};
          for (int m : cells[index(cc)]) {
            if (water && m == self) continue;
            const int at = water ? s.nat_solute + 3 * m : t.cgp[m].iswitch;
            const double d0 = s.x[3 * at] - x[0], d1 = s.x[3 * at + 1] - x[1], d2 = s.x[3 * at + 2] - x[2];
            // a split solute group stands in a water's row once per slice
            if (d0 * d0 + d1 * d1 + d2 * d2 <= rc * rc) n += water ? 1 : G.dev_first[m + 1] - G.dev_first[m];
          }
        }
    return n;
  };
  for (int w : ow) {
    const double* xo = &s.x[3 * (s.nat_solute + 3 * w)];
    out.ww = std::max(out.ww, count(xo, wcell, true, rww, w));
    out.wp = std::max(out.wp, count(xo, gcell, false, rpw, -1));
  }
  for (int g : sw) out.pw = std::max(out.pw, count(&s.x[3 * t.cgp[g].iswitch], wcell, true, rpw, -1));
  return out;
}

Batch::Batch(const std::vector<std::string>& inputs, int precision, int device) {
  if (inputs.empty()) throw std::runtime_error("no instances");
  if (precision != 0 && precision != 1) throw std::runtime_error("precision must be QGPU_FP64 or QGPU_MIXED");
  precision_ = precision;
  try {
    init(inputs, device);
  } catch (...) {
    // a constructor that throws runs no destructor: free what was made
    release();
    throw;
  }
}

void Batch::init(const std::vector<std::string>& inputs, int device) {
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&ev0));
  HIPCHECK(hipEventCreate(&ev1));
  HIPCHECK(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
  HIPCHECK(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
  ls_ = stream_;
  const char* br = std::getenv("QGPU_BRANCHES");
  branches_ = !(br && br[0] == '0');
  for (const auto& f : inputs) setups_.push_back(new qsys::Setup(qsys::prepare(f)));
  for (auto* s : setups_) {
    max_nwat_ = std::max(max_nwat_, s->nwat);
    const DevGroups G = device_groups(*s);
    max_ncgp_ = std::max(max_ncgp_, (int)G.first.size());
    max_ncgp_solute_ = std::max(max_ncgp_solute_, G.nsolute);
    max_natom_ = std::max(max_natom_, s->natom);
    max_nat_solute_ = std::max(max_nat_solute_, s->nat_solute);
    max_nstates_ = std::max(max_nstates_, s->nstates);
  }
  // one water-force kernel serves the whole batch: SPC and all-atom solvents
  // cannot share a batch
  int stype = -1;
  for (auto* s : setups_) {
    if (s->nwat == 0) continue;
    if (stype >= 0 && s->top.solvent_type != stype)
      throw std::runtime_error("instances of a batch must share the solvent type (SPC or all-atom)");
    stype = s->top.solvent_type;
  }
  allatom_ = stype == 1;
  for (auto* s : setups_) {
    q_dynamic_ = q_dynamic_ || (s->nqat > 0 && (s->pbc || !(s->inp.rcq > s->top.rexcl_o)));
    put_back_ = put_back_ || (s->pbc && (s->inp.put_solute_back || s->inp.put_solvent_back));
    freeze_ = freeze_ || s->inp.freeze;
    nose_hoover_ = nose_hoover_ || s->thermostat == 2;
  }
  nbcycle_ = setups_[0]->inp.nbcycle;
  for (auto* s : setups_)
    if (s->inp.nbcycle != nbcycle_ || nbcycle_ < 1)
      throw std::runtime_error("all instances of a batch need the same non_bond interval");
  // constant pressure: one volume-move interval and one list treatment
  // (LRF or not) for the batch's constant-pressure instances, whose moves
  // are tried together
  int npt_lrf = -1;
  for (auto* s : setups_) {
    if (!(s->pbc && s->inp.constant_pressure)) continue;
    if (s->inp.ivolume_cycle < 1) throw std::runtime_error("volume_change must be at least 1");
    if (npt_ && s->inp.ivolume_cycle != ivol_)
      throw std::runtime_error("constant-pressure instances of a batch need the same volume_change interval");
    if (npt_lrf >= 0 && npt_lrf != (s->inp.use_lrf ? 1 : 0))
      throw std::runtime_error("constant-pressure instances of a batch must all use LRF or all not");
    npt_ = true;
    ivol_ = s->inp.ivolume_cycle;
    npt_lrf = s->inp.use_lrf ? 1 : 0;
  }
  const char* ng = std::getenv("QGPU_NO_GRAPH");
  use_graph_ = !(ng && ng[0] == '1');
  const char* at = std::getenv("QGPU_AUTOTUNE");
  autotune_on_ = !(at && at[0] == '0');
  const size_t ni = setups_.size();
  HIPCHECK(hipMalloc(&d_step_, ni * sizeof(int)));
  HIPCHECK(hipMemset(d_step_, 0, ni * sizeof(int)));
  HIPCHECK(hipMalloc(&d_status_, ni * kStatusInts * sizeof(int)));
  HIPCHECK(hipMemset(d_status_, 0, ni * kStatusInts * sizeof(int)));
  HIPCHECK(hipMalloc(&d_fail_, sizeof(unsigned long long)));
  HIPCHECK(hipMemcpy(d_fail_, &kNoFailure, sizeof(unsigned long long), hipMemcpyHostToDevice));
  h_sys_.resize(ni);
  for (size_t k = 0; k < ni; ++k) {
    upload((int)k);
    h_sys_[k].step_ctr = d_step_ + k;
    h_sys_[k].status = d_status_ + k * kStatusInts;
    h_sys_[k].fail = d_fail_;
  }
  // one kinetic-energy partial per integrator workgroup of the widest instance
  const int kblocks = (std::max(max_units_, 1) + kIntegrateBlock - 1) / kIntegrateBlock;
  for (size_t k = 0; k < h_sys_.size(); ++k) {
    h_sys_[k].ke_part = inst_[k]->alloc<double>((size_t)kblocks * 6);
    h_sys_[k].ke_blocks = kblocks;
  }
  vol_.resize(ni);
  for (size_t k = 0; k < ni; ++k) {
    const qsys::Setup& s = *setups_[k];
    if (!(s.pbc && s.inp.constant_pressure)) continue;
    vol_[k].on = true;
    vol_[k].seed = s.inp.pressure_seed;
    vol_[k].x_saved = inst_[k]->alloc<double>(3 * (size_t)s.natom);
  }
  HIPCHECK(hipMalloc(&d_sys_, h_sys_.size() * sizeof(DevSys)));
  sync_sys();
  refresh_xf();
}

void Batch::sync_sys() {
  HIPCHECK(hipMemcpy(d_sys_, h_sys_.data(), h_sys_.size() * sizeof(DevSys), hipMemcpyHostToDevice));
}

void Batch::refresh_xf() {
  hipLaunchKernelGGL(k_x_to_f32, dim3((3 * max_natom_ + 255) / 256, (int)h_sys_.size()), dim3(256), 0,
                     stream_, d_sys_);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
}

void Batch::set_energy_interval(int iene, int max_frames) {
  if (iene < 0 || max_frames < 0) throw std::runtime_error("negative energy interval or frame count");
  HIPCHECK(hipStreamSynchronize(stream_));
  for (size_t k = 0; k < h_sys_.size(); ++k) {
    DevSys& D = h_sys_[k];
    Inst& I = *inst_[k];
    D.iene = iene;
    D.eseries_base = istep_;
    D.eseries_cap = max_frames;
    // the frame buffers are kept while they hold max_frames, else replaced
    if (max_frames > I.eseries_alloc) {
      I.release(D.eseries); I.release(D.eseries64); I.release(D.eseries_offd); I.release(D.eseries_n);
      const size_t n = (size_t)std::max(1, max_frames) * std::max(1, D.nstates) * kEneRecord;
      D.eseries = I.alloc<double>(n);
      D.eseries64 = precision_ == 1 && D.nqat > 0 ? I.alloc<double>(n) : nullptr;
      D.eseries_offd = D.noffd > 0 ? I.alloc<double>((size_t)std::max(1, max_frames) * D.noffd * 2) : nullptr;
      D.eseries_n = I.alloc<int>(1);
      I.eseries_alloc = max_frames;
    } else {
      HIPCHECK(hipMemset(D.eseries_n, 0, sizeof(int)));
    }
  }
  iene_ = iene;
  ebase_ = istep_;
  sync_sys();
  drop_graphs();
}

int Batch::energy_series(int k, double* out, int cap, bool fp64) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  if (!D.eseries_n) return 0;
  if (fp64 && !D.eseries64)
    throw std::runtime_error("FP64 re-evaluated frames exist only for mixed-precision batches with Q atoms");
  int n = 0;
  HIPCHECK(hipMemcpy(&n, D.eseries_n, sizeof(int), hipMemcpyDeviceToHost));
  n = std::min(n, D.eseries_cap);
  size_t per = (size_t)D.nstates * kEneRecord;
  size_t want = std::min((size_t)cap / std::max<size_t>(per, 1), (size_t)n);
  if (out && want) HIPCHECK(hipMemcpy(out, fp64 ? D.eseries64 : D.eseries, want * per * 8, hipMemcpyDeviceToHost));
  return n;
}

int Batch::offdiag_series(int k, double* out, int cap) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  if (!D.eseries_n) return 0;
  int n = 0;
  HIPCHECK(hipMemcpy(&n, D.eseries_n, sizeof(int), hipMemcpyDeviceToHost));
  n = std::min(n, D.eseries_cap);
  const size_t per = (size_t)D.noffd * 2;
  const size_t want = per ? std::min((size_t)cap / per, (size_t)n) : 0;
  if (out && want) HIPCHECK(hipMemcpy(out, D.eseries_offd, want * per * 8, hipMemcpyDeviceToHost));
  return n;
}

Batch::~Batch() { release(); }

void Batch::release() {
  drop_graphs();
  for (auto& m : maps_)
    if (m.second.first) (void)hipFree(m.second.first);
  if (far_x_) (void)hipFree(far_x_);
  if (far_y_) (void)hipFree(far_y_);
  if (far_frame_) (void)hipFree(far_frame_);
  for (auto* i : inst_) delete i;
  for (auto* s : setups_) delete s;
  if (d_sys_) (void)hipFree(d_sys_);
  if (d_step_) (void)hipFree(d_step_);
  if (d_status_) (void)hipFree(d_status_);
  if (d_fail_) (void)hipFree(d_fail_);
  d_status_ = nullptr; d_fail_ = nullptr;
  if (stream2_) (void)hipStreamDestroy(stream2_);
  if (ev_fork_) (void)hipEventDestroy(ev_fork_);
  if (ev_join_) (void)hipEventDestroy(ev_join_);
  stream2_ = nullptr; ev_fork_ = ev_join_ = nullptr;

  if (stream_) {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    (void)hipStreamDestroy(stream_);
  }
  maps_.clear(); inst_.clear(); setups_.clear();
  far_x_ = far_y_ = nullptr; far_frame_ = nullptr;
  d_sys_ = nullptr; d_step_ = nullptr; stream_ = nullptr; ev0 = ev1 = nullptr;
}

void Batch::upload(int k) {
  const qsys::Setup& s = *setups_[k];
  const qsys::Topology& t = s.top;
  if (s.nstates > kMaxQStates) throw std::runtime_error("too many FEP states");
  if (s.pbc && s.top.solvent_type != 0) throw std::runtime_error("periodic boxes need SPC solvent");
  auto* I = new Inst();
  inst_.push_back(I);
  I->groups = device_groups(s);
  const DevGroups& G = I->groups;
  DevSys& D = h_sys_[k];
  std::memset(&D, 0, sizeof(DevSys));
  D.natom = s.natom; D.nat_solute = s.nat_solute; D.nwat = s.nwat;
  D.ncgp = (int)G.first.size(); D.ncgp_solute = G.nsolute;
  D.nqat = s.nqat; D.nstates = s.nstates;
  D.precision = precision_;
  D.term_mask = TM_ALL;
  const int N = s.natom;
  D.x = I->upload(s.x);
  D.xf = I->alloc<float>(3 * N);
  D.xw = (precision_ == 1 && s.nwat > 0) ? I->alloc<float>((size_t)16 * s.nwat) : nullptr;
  D.v = I->upload(s.v);
  D.xx = I->alloc<double>(3 * N);
  D.d = I->alloc<double>(3 * N);
  D.xtop = I->upload(s.xtop);
  D.xq = s.inp.freeze ? I->alloc<double>(3 * (size_t)N) : D.x;
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
  D.cgp_first = I->upload(G.first);
  D.cgp_last = I->upload(G.last);
  D.cgp_sw = I->upload(G.sw);
  D.cgp_cfirst = I->upload(G.cfirst);
  D.cgp_clast = I->upload(G.clast);
  D.cgpatom = I->upload(t.cgpatom);
  D.iwhich_cgp = I->upload(G.which);
  D.lrf = I->alloc<double>((size_t)D.ncgp * kLrfStride);
  D.natyps = s.natyps;
  std::vector<double> av, bv;
  for (auto& ty : t.types) for (int j = 0; j < 3; ++j) { av.push_back(ty.avdw[j]); bv.push_back(ty.bvdw[j]); }
  D.avdw = I->upload(av);
  D.bvdw = I->upload(bv);
  D.avdw_f = I->upload(std::vector<float>(av.begin(), av.end()));
  D.bvdw_f = I->upload(std::vector<float>(bv.begin(), bv.end()));
  D.crg_f = I->upload(std::vector<float>(s.crg.begin(), s.crg.end()));
  D.ljcod = I->upload(s.ljcod);
  D.el14_scale = t.el14_scale;
  // exclusion and 1-4 tables (solute rows) and long lists
  D.listex = I->upload(std::vector<unsigned>(t.listex.begin(), t.listex.end()));
  D.list14 = I->upload(std::vector<unsigned>(t.list14.begin(), t.list14.end()));
  {
    // long exclusion / 1-4 pairs as a CSR over the lower atom
    std::vector<std::vector<int>> rows(std::max(1, s.nat_solute));
    auto add = [&](int i, int j, int bit) {
      int lo = std::min(i, j), hi = std::max(i, j);
      for (int& c : rows[lo])
        if ((c >> 2) == hi) { c |= bit; return; }
      rows[lo].push_back(hi << 2 | bit);
    };
    for (auto& p : t.listexlong) add(p.first, p.second, 1);
    for (auto& p : t.list14long) add(p.first, p.second, 2);
    std::vector<int> first(1, 0), code;
    for (auto& r : rows) {
      code.insert(code.end(), r.begin(), r.end());
      first.push_back((int)code.size());
    }
    D.exl_first = I->upload(first);
    D.exl_code = I->upload(code);
    std::vector<int> slot(std::max(1, s.nat_solute), 0);
    for (int g = 0; g < G.nsolute; ++g)
      for (int a = G.first[g]; a <= G.last[g]; ++a) slot[t.cgpatom[a]] = a - G.first[g];
    D.cgp_slot = I->upload(slot);
  }
  // lists: row capacities from the longest rows of the starting
  // coordinates (twice that, for density fluctuations); rows that still
  // outgrow them are regrown at the failing list update (grow_lists)
  {
    const ListRows r = longest_rows(s, G);
    auto cap = [](int longest, int most) {
      const int c = (2 * longest + 32 + 15) / 16 * 16;
      return std::max(1, std::min(most, c));
    };
    D.ww_cap = cap(r.ww, s.nwat);
    D.wp_cap = cap(r.wp, G.nsolute);
    D.pw_cap = cap(r.pw, s.nwat);
  }
  D.pp_cap = std::max(1, G.nsolute);
  D.ww_cnt = I->alloc<int>(s.nwat);
  D.ww_nb = I->alloc<int>((size_t)s.nwat * D.ww_cap);
  D.wp_cnt = I->alloc<int>(s.nwat);
  D.wp_nb = I->alloc<int>((size_t)s.nwat * D.wp_cap);
  D.pp_cnt = I->alloc<int>(G.nsolute);
  D.pp_nb = I->alloc<int>((size_t)G.nsolute * D.pp_cap);
  D.pp_mask = I->alloc<ulonglong2>((size_t)G.nsolute * D.pp_cap);
  D.pw_cnt = I->alloc<int>(G.nsolute);
  D.pw_nb = I->alloc<int>((size_t)G.nsolute * D.pw_cap);
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
    D.solv_allatom = t.solvent_type == 1 ? 1 : 0;
    if (D.solv_allatom) {
      // the kernels take one LJ code per position in the water: every water
      // must carry the first water's atom types
      for (int w = 1; w < s.nwat; ++w)
        for (int k = 0; k < 3; ++k)
          if (s.iac[o + 3 * w + k] != s.iac[o + k])
            throw std::runtime_error("all-atom solvent: waters with differing atom types are not supported");
      for (int la = 0; la < 3; ++la) {
        const int c = s.ljcod[s.iac[o + la] * s.natyps + s.iac[o]] - 1;
        for (int ka = 0; ka < 3; ++ka) {
          D.aww[3 * la + ka] = s.alj_solv[la][c] * s.alj_solv[ka][c];
          D.bww[3 * la + ka] = s.blj_solv[la][c] * s.blj_solv[ka][c];
        }
      }
    }
    // standard waters: one charge group of (O, H1, H2) in this order with the
    // solvent charges and no Q atom; the LRF gather then reads them directly
    bool std_w = true;
    for (int w = 0; w < s.nwat && std_w; ++w) {
      const int a = s.nat_solute + 3 * w;
      const auto& g = t.cgp[s.iwhich_cgp[a]];
      std_w = g.last - g.first == 2;
      for (int k = 0; k < 3 && std_w; ++k)
        std_w = t.cgpatom[g.first + k] == a + k && s.iwhich_cgp[a + k] == s.iwhich_cgp[a] &&
                !s.iqatom[a + k] && s.crg[a + k] == s.chg_solv[k];
    }
    for (int w = 0; w < s.nwat && std_w; ++w)
      for (int k = 0; k < 3 && std_w; ++k) std_w = s.iac[s.nat_solute + 3 * w + k] == s.iac[s.nat_solute + k];
    D.std_water = std_w ? 1 : 0;
    if (std_w && s.nat_solute > 0) {
      std::vector<double> wd((size_t)s.nat_solute * 6);
      std::vector<float> wf((size_t)s.nat_solute * 6);
      std::vector<unsigned char> wz(s.nat_solute, 0);
      for (int i = 0; i < s.nat_solute; ++i) {
        const int ti = s.iac[i];
        for (int k = 0; k < 3; ++k) {
          const int j = s.nat_solute + k, tj = s.iac[j];
          const int lj = s.ljcod[ti * s.natyps + tj] - 1;
          const double a = t.types[ti].avdw[lj] * t.types[tj].avdw[lj];
          const double b = t.types[ti].bvdw[lj] * t.types[tj].bvdw[lj];
          wd[6 * i + k] = a; wd[6 * i + 3 + k] = b;
          wf[6 * i + k] = (float)a; wf[6 * i + 3 + k] = (float)b;
          if (s.crg[i] * s.crg[j] == 0.0 && a == 0.0 && b == 0.0) wz[i] |= (unsigned char)(1u << k);
        }
      }
      D.swp_d = I->upload(wd);
      D.swp_f = I->upload(wf);
      D.swp_zero = I->upload(wz);
      if (precision_ == 1) {
        // solute slot records for the water side of the solute-water pairs:
        // the constant fields here, x y z written with xf
        int nslot = 0;
        for (int g = 0; g < G.nsolute; ++g) nslot = std::max(nslot, G.last[g] + 1);
        std::vector<float> rec((size_t)kSrec * std::max(1, nslot), 0.0f);
        std::vector<int> of(s.nat_solute, -1);
        for (int g = 0; g < G.nsolute; ++g)
          for (int a = G.first[g]; a <= G.last[g]; ++a) {
            const int i = t.cgpatom[a];
            if (i < 0 || i >= s.nat_solute) throw std::runtime_error("solute charge group holds a solvent atom");
            of[i] = a;
            float* r = rec.data() + (size_t)kSrec * a;
            for (int c = 0; c < 3; ++c) r[c] = (float)s.x[3 * i + c];
            r[3] = (float)s.crg[i];
            for (int c = 0; c < 6; ++c) r[4 + c] = wf[6 * i + c];
            r[10] = s.iqatom[i] ? 1.0f : 0.0f;
          }
        // every solute atom must have its slot (each is in one charge group);
        // otherwise the pairs take the per-atom path
        if (std::find(of.begin(), of.end(), -1) == of.end()) {
          D.srec = I->upload(rec);
          D.srec_of_atom = I->upload(of);
        }
      }
    }
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
  D.wpol_g = I->alloc<double>((size_t)9 * std::max(1, s.nwat));
  D.fk_fix = s.inp.fk_fix; D.fk_pshell = s.inp.fk_pshell; D.freeze = s.inp.freeze;
  for (int a = 0; a < 3; ++a)
    for (int k2 = 0; k2 < 2; ++k2) { D.alj_w[a][k2] = s.alj_solv[a][k2]; D.blj_w[a][k2] = s.blj_solv[a][k2]; }
  // Q atoms: static Q-solute and Q-water sets as nbqplist/nbqwlist build
  // them at the first list update (src/md.f90:10215-10322, 10464-10549)
  D.iqseq = I->upload(s.iqseq);
  D.qcrg = I->upload(s.qcrg);
  D.lambda = I->upload(s.lambda);
  D.qvdw = s.qvdw_flag ? 1 : 0;
  D.qiac = I->upload(s.qiac);
  D.qavdw = I->upload(s.qavdw);
  D.qbvdw = I->upload(s.qbvdw);
  D.sc_ncol = s.natyps + s.nqat;
  D.qq_lib = s.qq_lib_charges ? 1 : 0;
  D.sc_lookup = I->upload(s.sc_lookup);
  {
    std::vector<int> kl;
    std::vector<double> amu;
    for (const auto& o : s.offd) {
      kl.push_back(s.iqseq[o.k]); kl.push_back(s.iqseq[o.l]);
      amu.push_back(o.A); amu.push_back(o.mu);
    }
    D.noffd = (int)s.offd.size();
    D.offd_kl = I->upload(kl);
    D.offd_amu = I->upload(amu);
  }
  std::vector<int> qq_atom_i, qq_atom_j;   // atoms of each Q pair, for the term slots
  {
    const int nq = s.nqat, nsl = s.nat_solute;
    auto qconn = [&](int st, int iq, int a) { return s.qconn[((size_t)st * nq + iq) * nsl + a]; };
    std::vector<signed char> qp_lj((size_t)std::max(1, nsl * nq), 0), qw_lj(std::max(1, 3 * nq), 0);
    std::vector<int> qp_atoms, qw_waters;
    const double rcq2 = s.inp.rcq * s.inp.rcq;
    auto dist2c = [&](int a) {
      double d0 = s.x[3 * a] - t.xpcent[0], d1 = s.x[3 * a + 1] - t.xpcent[1], d2 = s.x[3 * a + 2] - t.xpcent[2];
      return d0 * d0 + d1 * d1 + d2 * d2;
    };
    // atoms that can enter the Q-solute list: not bonded or angled to a Q
    // atom in any state (nbqplist src/md.f90:10295); their LJ codes do not
    // depend on the coordinates
    std::vector<unsigned char> qp_cand(std::max(1, nsl), 0);
    if (nq > 0) {
      for (int i = 0; i < nsl; ++i) {
        bool conn = false;
        for (int st = 0; st < s.nstates && !conn; ++st)
          for (int iq = 0; iq < nq && !conn; ++iq) conn = qconn(st, iq, i) <= 3;
        if (conn) continue;
        qp_cand[i] = 1;
        for (int iq = 0; iq < nq; ++iq) {
          int lj = s.ljcod[s.iac[i] * s.natyps + s.iac[s.iqseq[iq]]];
          if (qconn(0, iq, i) == 4) lj = 3;
          qp_lj[(size_t)i * nq + iq] = (signed char)lj;
        }
      }
      for (int g = 0; g < t.ncgp_solute; ++g) {
        int sw = t.cgp[g].iswitch;
        if (s.excl[sw] || dist2c(sw) > rcq2) continue;
        for (int a = t.cgp[g].first; a <= t.cgp[g].last; ++a) {
          int i = t.cgpatom[a];
          if (!qp_cand[i]) continue;
          qp_atoms.push_back(i);
          flags[i] |= F_QP;
        }
      }
      std::sort(qp_atoms.begin(), qp_atoms.end());
      for (int iw = 0; iw < s.nwat; ++iw) {
        int o = s.nat_solute + 3 * iw;
        if (s.excl[o] || dist2c(o) > rcq2) continue;
        qw_waters.push_back(iw);
        flags[o] |= F_QW;
      }
      // nbqwlist's code per water atom: ljcod(iac(water atom ip), iac(Q atom))
      for (int iq = 0; iq < nq && s.nwat > 0; ++iq)
        for (int ip = 0; ip < 3; ++ip)
          qw_lj[3 * iq + ip] = (signed char)s.ljcod[s.iac[s.nat_solute + ip] * s.natyps + s.iac[s.iqseq[iq]]];
      HIPCHECK(hipMemcpy(const_cast<unsigned char*>(D.flags), flags.data(), N, hipMemcpyHostToDevice));
    }
    D.qp_lj = I->upload(qp_lj);
    D.qw_lj = I->upload(qw_lj);
    {
      // the water side needs only the lambda-weighted sum over states of
      // each Q-water pair's gradient; with fixed Q LJ parameters the states
      // differ only in the Q charge
      std::vector<double> rec((size_t)std::max(1, nq) * kQwRec, 0.0);
      for (int iq = 0; iq < nq && s.nwat > 0; ++iq) {
        double* r = rec.data() + (size_t)iq * kQwRec;
        for (int st = 0; st < s.nstates; ++st) {
          r[0] += s.lambda[st] * s.qcrg[iq * s.nstates + st];
          r[1] += s.lambda[st];
        }
        const int ti = s.iac[s.iqseq[iq]];
        for (int ip = 0; ip < 3; ++ip) {
          const int lj = qw_lj[3 * iq + ip] - 1;
          r[2 + 2 * ip] = t.types[ti].avdw[lj] * s.alj_solv[ip][lj];
          r[3 + 2 * ip] = t.types[ti].bvdw[lj] * s.blj_solv[ip][lj];
        }
      }
      std::vector<float> recf(rec.begin(), rec.end());
      D.qw_rec_d = I->upload(rec);
      D.qw_rec_f = I->upload(recf);
    }
    // Rcq > rexcl_o: the sets above hold for the whole run (nbqplist and
    // nbqwlist return after their first call); otherwise k_q_lists rebuilds
    // them at every list update, in capacity-sized arrays
    // a periodic box rebuilds them at every list update (nbqplist_box and
    // nbqwlist_box, with a positive Rcq)
    D.q_dynamic = (nq > 0 && (s.pbc || !(s.inp.rcq > t.rexcl_o))) ? 1 : 0;
    D.rcq2 = rcq2;
    {
      std::vector<int> cnt = {(int)qp_atoms.size(), (int)qw_waters.size()};
      D.q_counts = I->upload(cnt);
    }
    if (D.q_dynamic) {
      qp_atoms.resize(std::max<size_t>(qp_atoms.size(), (size_t)nsl), 0);
      qw_waters.resize(std::max<size_t>(qw_waters.size(), (size_t)s.nwat), 0);
    }
    D.qp_cand = I->upload(qp_cand);
    D.qp_gin = I->alloc<unsigned char>(std::max(1, G.nsolute));
    for (int c = 0; c < 3; ++c) D.xpcent[c] = t.xpcent[c];
    // periodic box
    D.pbc = s.pbc ? 1 : 0;
    for (int c = 0; c < 3; ++c) {
      D.boxl[c] = s.boxl[c]; D.inv_boxl[c] = s.inv_boxl[c];
      D.boxl_f[c] = (float)s.boxl[c]; D.inv_boxl_f[c] = (float)s.inv_boxl[c];
    }
    D.mc_move = 0;
    for (int c = 0; c < 3; ++c) D.boxl_old[c] = s.boxl[c];
    D.qswitch = s.pbc ? std::max(0, s.qswitch) : 0;
    D.qp_shatom = I->alloc<int>(std::max(1, G.nsolute));
    D.boxc = I->upload(std::vector<double>{s.boxc[0], s.boxc[1], s.boxc[2], -1.0});
    D.rigid_boxc = s.inp.rigid_box_centre ? 1 : 0;
    D.put_solute_back = s.inp.put_solute_back ? 1 : 0;
    D.put_solvent_back = s.inp.put_solvent_back ? 1 : 0;
    D.nmol = s.nmol;
    D.nsolute_mol = s.nsolute_mol;
    D.mol_start = I->upload(t.istart_mol);
    D.mol_minv = I->upload(s.pbc ? s.mol_minv : std::vector<double>(1, 0.0));
    D.qp_atoms = I->upload(qp_atoms);
    D.qw_waters = I->upload(qw_waters);
    D.e_q = I->alloc<double>((size_t)std::max(1, nq * s.nstates) * 4);
    D.q_part = I->alloc<double>((size_t)std::max(1, nq) * kQChunks * kQPartStride);
    if (precision_ == 1) {
      D.q_part64 = I->alloc<double>((size_t)std::max(1, nq) * kQChunks * kQPartStride);
      D.e_q64 = I->alloc<double>((size_t)std::max(1, nq * s.nstates) * 4);
    }
    // Q-Q / Q-neighbour pairs, flattened over states, with atom incidence
    std::vector<int> qst, qiq, qj, qjq, qlj;
    std::vector<double> qel;
    for (int st = 0; st < s.nstates; ++st)
      for (auto& pr : s.nbqq[st]) {
        qst.push_back(st); qiq.push_back(pr.iq); qj.push_back(pr.j); qjq.push_back(pr.jq);
        qlj.push_back(pr.ljcod); qel.push_back(pr.el_scale);
      }
    D.nqq = (int)qst.size();
    D.qq_state = I->upload(qst); D.qq_iq = I->upload(qiq); D.qq_j = I->upload(qj);
    D.qq_jq = I->upload(qjq); D.qq_lj = I->upload(qlj); D.qq_elscale = I->upload(qel);
    D.e_qq = I->alloc<double>((size_t)std::max(1, D.nqq) * 2);
    for (int p = 0; p < D.nqq; ++p) { qq_atom_i.push_back(s.iqseq[qiq[p]]); qq_atom_j.push_back(qj[p]); }
    max_nqat_ = std::max(max_nqat_, nq);
  }
  // bonded terms after shrink_topology, with atom incidence
  {
    std::vector<int> bv, av2, tv;
    std::vector<double> bl, al, tl;
    for (auto& b : s.bnd) { bv.push_back(b.i); bv.push_back(b.j); bv.push_back(b.cod); }
    for (auto& a : s.ang) { av2.push_back(a.i); av2.push_back(a.j); av2.push_back(a.k); av2.push_back(a.cod); }
    for (auto& q : s.tor) { tv.push_back(q.i); tv.push_back(q.j); tv.push_back(q.k); tv.push_back(q.l); tv.push_back(q.cod); }
    for (auto& b : t.bondlib) { bl.push_back(b.fk); bl.push_back(b.b0); }
    for (auto& a : t.anglib) { al.push_back(a.fk); al.push_back(a.ang0); }
    for (auto& q : t.torlib) { tl.push_back(q.fk); tl.push_back(q.rmult); tl.push_back(q.deltor); tl.push_back(q.paths); }
    std::vector<int> iv;
    std::vector<double> il;
    for (auto& q : s.imp) { iv.push_back(q.i); iv.push_back(q.j); iv.push_back(q.k); iv.push_back(q.l); iv.push_back(q.cod); }
    for (auto& q : t.implib) { il.push_back(q.fk); il.push_back(q.imp0); }
    D.imp = I->upload(iv); D.implib = I->upload(il);
    D.nimp = (int)s.imp.size(); D.nimp_solute = s.nimps_solute;
    D.ff_type = s.ff_type;
    D.bnd = I->upload(bv); D.ang = I->upload(av2); D.tor = I->upload(tv);
    D.bondlib = I->upload(bl); D.anglib = I->upload(al); D.torlib = I->upload(tl);
    D.nbnd = (int)s.bnd.size(); D.nbnd_solute = s.nbonds_solute;
    D.nang = (int)s.ang.size(); D.nang_solute = s.nangles_solute;
    D.ntor = (int)s.tor.size(); D.ntor_solute = s.ntors_solute;
    // extra terms: Q bonds, angles, torsions, impropers of every state where
    // present, then the restraints of p_restrain; each with the atoms of
    // its gradient slots
    std::vector<int> xt;
    std::vector<double> xpar;
    std::vector<std::vector<int>> xatoms;
    std::vector<int> rst_atoms;
    auto add_xt = [&](std::array<int, kXtInts> r, std::array<double, kXtPars> pr, std::vector<int> atoms) {
      xt.insert(xt.end(), r.begin(), r.end());
      xpar.insert(xpar.end(), pr.begin(), pr.end());
      xatoms.push_back(std::move(atoms));
    };
    const int nst = s.nstates;
    std::vector<int> qb_atoms;
    for (size_t ib = 0; ib < s.qbnd.size(); ++ib) {
      const qsys::QTerm& b = s.qbnd[ib];
      qb_atoms.push_back(b.i); qb_atoms.push_back(b.j);
      int last = -1;
      for (int st = 0; st < nst; ++st) if (b.cod[st] > 0) last = st;
      for (int st = 0; st < nst; ++st) {
        if (b.cod[st] <= 0) continue;
        const qsys::QBondLib& L = s.qbondlib[b.cod[st] - 1];
        add_xt({XT_QBOND, b.i, b.j, (int)ib, st == last ? 1 : 0, st, 0, 0, 0, 0},
               {L.dmz, L.amz, L.r0, L.fk, 0, 0, 0, 0}, {b.i, b.j});
      }
    }
    // the state whose Morse factor a coupled term of state st sees
    auto coupled = [&](int cb, int st, std::array<double, kXtPars>& pr) {
      if (cb < 0) return -1;
      for (int sp = st; sp >= 0; --sp)
        if (s.qbnd[cb].cod[sp] > 0) {
          const qsys::QBondLib& L = s.qbondlib[s.qbnd[cb].cod[sp] - 1];
          pr[4] = L.dmz; pr[5] = L.amz; pr[6] = L.r0; pr[7] = L.fk;
          return sp;
        }
      return -1;
    };
    auto with_bond = [&](std::vector<int> a, int cb) {
      if (cb >= 0) { a.push_back(s.qbnd[cb].i); a.push_back(s.qbnd[cb].j); }
      return a;
    };
    for (size_t ia = 0; ia < s.qang.size(); ++ia) {
      const qsys::QTerm& q = s.qang[ia];
      const int cb = s.qang_cb[ia];
      for (int st = 0; st < nst; ++st) {
        if (q.cod[st] <= 0) continue;
        const qsys::QAngLib& L = s.qanglib[q.cod[st] - 1];
        std::array<double, kXtPars> pr{L.fk, L.ang0, 0, 0, 0, 0, 0, 0};
        const int seff = coupled(cb, st, pr);
        add_xt({XT_QANGLE, q.i, q.j, q.k, 0, st, cb, s.qang_cf[ia], seff, 0}, pr, with_bond({q.i, q.j, q.k}, cb));
      }
    }
    for (size_t it = 0; it < s.qtor.size(); ++it) {
      const qsys::QTerm& q = s.qtor[it];
      const int cb = s.qtor_cb[it];
      for (int st = 0; st < nst; ++st) {
        if (q.cod[st] <= 0) continue;
        const qsys::QTorLib& L = s.qtorlib[q.cod[st] - 1];
        std::array<double, kXtPars> pr{L.fk, L.rmult, L.deltor, 0, 0, 0, 0, 0};
        const int seff = coupled(cb, st, pr);
        add_xt({XT_QTORSION, q.i, q.j, q.k, q.l, st, cb, s.qtor_cf[it], seff, 0}, pr,
               with_bond({q.i, q.j, q.k, q.l}, cb));
      }
    }
    for (size_t it = 0; it < s.qimp.size(); ++it) {
      const qsys::QTerm& q = s.qimp[it];
      const int cb = s.qimp_cb[it];
      for (int st = 0; st < nst; ++st) {
        if (q.cod[st] <= 0) continue;
        const qsys::QImpLib& L = s.qimplib[q.cod[st] - 1];
        std::array<double, kXtPars> pr{L.fk, L.imp0, 0, 0, 0, 0, 0, 0};
        const int seff = coupled(cb, st, pr);
        add_xt({XT_QIMPROPER, q.i, q.j, q.k, q.l, st, cb, s.qimp_cf[it], seff, 0}, pr,
               with_bond({q.i, q.j, q.k, q.l}, cb));
      }
    }
    const double deg2rad = 4.0 * std::atan(1.0) / 180.0;
    for (const qsys::RstSeq& r : s.rstseq) {
      std::vector<int> use;
      for (int i = r.i; i <= r.j; ++i)
        if (s.heavy[i] || (r.ih == 1 && !(s.excl[i] && s.freeze))) use.push_back(i);
      if (r.to_centre == 1 || r.to_centre == 2) {
        if (use.empty()) continue;
        if (r.to_centre == 2) {
          double tm = 0;
          for (int i : use) tm = tm + t.types[s.iac[i]].mass;
          if (!(tm > 0)) continue;
        }
        add_xt({XT_RST_SEQC, (int)rst_atoms.size(), (int)use.size(), r.to_centre, 0, -1, 0, 0, 0, 0},
               {r.fk, 0, 0, 0, 0, 0, 0, 0}, use);
        rst_atoms.insert(rst_atoms.end(), use.begin(), use.end());
      } else {
        for (int i : use) add_xt({XT_RST_SEQ, i, 0, 0, 0, -1, 0, 0, 0, 0}, {r.fk, 0, 0, 0, 0, 0, 0, 0}, {i});
      }
    }
    for (const qsys::RstPos& r : s.rstpos)
      add_xt({XT_RST_POS, r.i, 0, 0, 0, r.ipsi, 0, 0, 0, 0}, {r.x[0], r.x[1], r.x[2], r.fk[0], r.fk[1], r.fk[2], 0, 0},
             {r.i});
    for (const qsys::RstDis& r : s.rstdis)
      add_xt({XT_RST_DIST, r.i, r.j, 0, 0, r.ipsi, 0, 0, 0, 0}, {r.d1, r.d2, r.fk, 0, 0, 0, 0, 0}, {r.i, r.j});
    for (const qsys::RstAng& r : s.rstang)
      add_xt({XT_RST_ANG, r.i, r.j, r.k, 0, r.ipsi, 0, 0, 0, 0}, {r.ang * deg2rad, r.fk, 0, 0, 0, 0, 0, 0},
             {r.i, r.j, r.k});
    // wall restraints only for spheres (src/md.f90:15768)
    for (const qsys::RstWal& r : s.rstwal)
      for (int i = r.i; i <= r.j && !s.pbc; ++i)
        if (s.heavy[i] || r.ih == 1)
          add_xt({XT_RST_WALL, i, 0, 0, 0, -1, 0, 0, 0, 0}, {r.d, r.fk, r.aMorse, r.dMorse, 0, 0, 0, 0}, {i});
    D.nxt = (int)xatoms.size();
    D.nqbond = (int)s.qbnd.size();
    D.qb_atoms = I->upload(qb_atoms);
    D.qb_cur = I->alloc<double>((size_t)std::max(1, D.nqbond) * kMorseRec);
    D.qb_prev = I->alloc<double>((size_t)std::max(1, D.nqbond) * kMorseRec);
    D.rst_atoms = I->upload(rst_atoms);
    D.e_xt = I->alloc<double>((size_t)std::max(1, D.nxt));
    D.freeze = s.freeze;
    // term slots (bonds, angles, torsions, impropers, the Q pairs, then the
    // extra terms) and the CSR of every atom's slots in slot order
    D.e_bterm = I->alloc<double>((size_t)std::max(1, D.nbnd + D.nang + D.ntor + D.nimp));
    std::vector<std::vector<int>> inc(N);
    int slot = 0;
    for (int n = 0; n < D.nbnd; ++n) for (int a : {s.bnd[n].i, s.bnd[n].j}) inc[a].push_back(slot++);
    for (int n = 0; n < D.nang; ++n) for (int a : {s.ang[n].i, s.ang[n].j, s.ang[n].k}) inc[a].push_back(slot++);
    for (int n = 0; n < D.ntor; ++n)
      for (int a : {s.tor[n].i, s.tor[n].j, s.tor[n].k, s.tor[n].l}) inc[a].push_back(slot++);
    for (int n = 0; n < D.nimp; ++n)
      for (int a : {s.imp[n].i, s.imp[n].j, s.imp[n].k, s.imp[n].l}) inc[a].push_back(slot++);
    for (int p = 0; p < D.nqq; ++p)
      for (int a : {qq_atom_i[p], qq_atom_j[p]}) inc[a].push_back(slot++);
    for (int n = 0; n < D.nxt; ++n) {
      xt[(size_t)kXtInts * n + 9] = slot;
      for (int a : xatoms[n]) inc[a].push_back(slot++);
    }
    D.xt = I->upload(xt);
    D.xp = I->upload(xpar);
    std::vector<int> sa, sf, sl;
    for (int i = 0; i < N; ++i)
      if (!inc[i].empty()) {
        sa.push_back(i);
        sf.push_back((int)sl.size());
        sl.insert(sl.end(), inc[i].begin(), inc[i].end());
      }
    sf.push_back((int)sl.size());
    // each term slot is stored at its position in the atom's CSR run, so the
    // gather reads every atom's contributions contiguously
    std::vector<int> pos(std::max(1, slot), 0);
    for (size_t k = 0; k < sl.size(); ++k) pos[sl[k]] = (int)k;
    D.slot_g = I->alloc<double>((size_t)std::max(1, slot) * 3);
    D.nslot_atoms = (int)sa.size();
    D.slot_atom = I->upload(sa); D.slot_first = I->upload(sf); D.slot_pos = I->upload(pos);
    {
      // the integrator gathers the slots of the atoms it moves
      std::vector<int> of(N, -1);
      for (size_t k = 0; k < sa.size(); ++k) of[sa[k]] = (int)k;
      D.slot_of_atom = I->upload(of);
    }
    {
      // monitored group pairs
      std::vector<int> at, fi, fj, ni, nj, lo;
      std::vector<signed char> lj;
      for (const auto& m : s.mon) {
        fi.push_back((int)at.size()); at.insert(at.end(), m.ai.begin(), m.ai.end());
        fj.push_back((int)at.size()); at.insert(at.end(), m.aj.begin(), m.aj.end());
        ni.push_back((int)m.ai.size()); nj.push_back((int)m.aj.size());
        lo.push_back((int)lj.size());
        for (int c : m.ljcod) lj.push_back((signed char)c);
      }
      D.mon_n = (int)s.mon.size();
      D.iqatom_d = I->upload(s.iqatom);
      D.mon_at = I->upload(at); D.mon_fi = I->upload(fi); D.mon_fj = I->upload(fj);
      D.mon_ni = I->upload(ni); D.mon_nj = I->upload(nj); D.mon_lo = I->upload(lo);
      D.mon_lj = I->upload(lj);
      D.mon_out = I->alloc<double>((size_t)std::max(1, D.mon_n) * (kMon + 2 * std::max(1, s.nstates)));
      max_mon_ = std::max(max_mon_, D.mon_n);
    }
    max_terms_ = std::max(max_terms_, D.nbnd + D.nang + D.ntor + D.nimp + D.nqq + D.nxt);
    max_slot_atoms_ = std::max(max_slot_atoms_, std::max(D.nslot_atoms, D.nqbond));
  }
  // integrator
  D.dt = s.dt; D.temp0 = s.inp.temp0; D.tau_t = s.tau_t;
  // thermostat and integrator
  D.thermostat = s.thermostat;
  D.vverlet = s.velocity_verlet ? 1 : 0;
  D.dv_mod = s.dv_mod;
  D.dv_friction = s.dv_friction;
  D.lgauss = s.langevin_gauss ? 1 : 0;
  D.rands = nullptr; D.rng = nullptr; D.rng_jump = nullptr;
  D.rng_step[0] = 1; D.rng_step[1] = 0;
  if (s.thermostat == 1) {
    // randm as an affine map mod 1e8 (state' = 31415821 state + 1): the
    // jump to atom i's first draw and over a whole step's draws
    const long long m = 100000000LL, mult = 31415821LL;
    const int per_atom = s.langevin_gauss ? 36 : 3;
    std::vector<int> jump(2 * (size_t)s.natom);
    long long A = 1, C = 0;
    for (int i = 0; i < s.natom; ++i) {
      jump[2 * i] = (int)A;
      jump[2 * i + 1] = (int)C;
      for (int k = 0; k < per_atom; ++k) { A = A * mult % m; C = (C * mult + 1) % m; }
    }
    D.rng_step[0] = A; D.rng_step[1] = C;
    D.rng_jump = I->upload(jump);
    D.rands = I->upload(s.rands);
    // the stream continues from the Maxwell draws; unstarted, the first
    // draw seeds it from random_seed as randm does
    qsys::Randm r = s.rng;
    if (!r.started) { r.started = true; r.irand = std::abs(s.inp.iseed) % 100000000; }
    D.rng = I->upload(std::vector<int>{r.irand});
  }
  D.nhn = s.thermostat == 2 ? (int)s.qnh.size() : 0;
  D.qnh = D.nhn ? I->upload(s.qnh) : nullptr;
  D.nh = D.nhn ? I->upload(std::vector<double>(3 * (size_t)D.nhn + 2, 0.0)) : nullptr;
  D.kbt = s.kbt;
  D.separate_scaling = s.separate_scaling; D.detail_temps = s.detail_temps;
  D.ndegf = s.ndegf; D.ndegfree = s.ndegfree; D.ndegf_solute = s.ndegf_solute;
  D.ndegf_solvent = s.ndegf_solvent; D.ndegfree_solute = s.ndegfree_solute;
  D.ndegfree_solvent = s.ndegfree_solvent;
  // SHAKE units
  std::vector<int> ufirst, uatoms, cfirst, shi, shj;
  std::vector<double> shd;
  std::vector<signed char> ukind;
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
    bool water = atoms.size() == 3 && m.bond.size() == 3 && atoms[1] == atoms[0] + 1 && atoms[2] == atoms[0] + 2;
    if (water) {
      const int a = atoms[0];
      water = m.bond[0].i == a && m.bond[0].j == a + 1 && m.bond[1].i == a && m.bond[1].j == a + 2 &&
              m.bond[2].i == a + 1 && m.bond[2].j == a + 2;
    }
    ukind.push_back(water ? 1 : 0);
    uatoms.insert(uatoms.end(), atoms.begin(), atoms.end());
  }
  for (int i = 0; i < N; ++i)
    if (!used[i]) {
      ufirst.push_back((int)uatoms.size());
      cfirst.push_back((int)shi.size());
      uatoms.push_back(i);
      ukind.push_back(0);
    }
  D.unit_kind = I->upload(ukind);
  ufirst.push_back((int)uatoms.size());
  cfirst.push_back((int)shi.size());
  D.nunits = (int)ufirst.size() - 1;
  D.has_shake = s.shake_constraints > 0;
  max_units_ = std::max(max_units_, D.nunits);
  D.unit_first = I->upload(ufirst);
  D.unit_atoms = I->upload(uatoms);
  {
    std::vector<int> head(D.nunits);
    for (int u = 0; u < D.nunits; ++u) head[u] = uatoms[ufirst[u]];
    D.unit_head = I->upload(head);
  }
  D.con_first = I->upload(cfirst);
  D.sh_i = I->upload(shi);
  D.sh_j = I->upload(shj);
  D.sh_d2 = I->upload(shd);
  D.shake_maxiter = qsys::kShakeMaxIter;
  if (const char* m = std::getenv("QGPU_SHAKE_MAXITER")) D.shake_maxiter = std::max(1, std::atoi(m));
  // energies and reductions
  D.e_water = I->alloc<double>((size_t)std::max(1, s.nwat) * kEWater);
  D.e_solute = I->alloc<double>((size_t)std::max(1, s.nat_solute) * 8);
  D.e_wpol = I->alloc<double>(std::max(1, D.nshell));
  D.wpol_part = I->alloc<double>((size_t)std::max(1, D.nshell) * kWpolSplit * 3);
  max_nshell_ = std::max(max_nshell_, D.nshell);
  D.e_far = I->alloc<double>(std::max(1, D.ncgp));
  I->ke_blocks = 0;
  std::vector<double> scal(8, 0.0);
  initial_scales(s, s.v.data(), scal.data());
  D.scal = I->upload(scal);
  std::vector<double> e0(E_NTERMS, 0.0);
  e0[E_KINETIC] = scal[4];
  D.energies = I->upload(e0);
  D.qenergies = I->alloc<double>((size_t)std::max(1, s.nstates) * kQTerms);
  I->natom = N; I->nwat = s.nwat;
  // the kinetic partial buffer needs the batch-wide unit count: allocated
  // once every instance is known (Batch::init)
  D.ke_part = nullptr;
  D.ke_blocks = 0;
}

// Items and lanes per item of the tuned kernels, per instance, and the
// relative work of an instance (for balancing instances over the XCDs).
static void tuned_items(int kernel, const DevSys& D, bool f64, int& items, int& lanes, double& work) {
  switch (kernel) {
    case Batch::TK_WATER_LISTS:
      items = D.nwat; lanes = f64 ? kWLLanes64 : kWLLanes32;
      work = (double)D.nwat * (D.nwat + D.ncgp_solute); break;
    case Batch::TK_SOLUTE_LISTS:
      items = D.ncgp_solute; lanes = 64; work = (double)D.ncgp_solute * (D.ncgp_solute + D.nwat); break;
    case Batch::TK_WATER_FORCES:
      items = D.nwat; lanes = 8; work = (double)D.nwat; break;
    default:
      items = D.nat_solute; lanes = kSFLanes; work = (double)D.nat_solute; break;
  }
  if (items == 0 || (kernel == Batch::TK_SOLUTE_LISTS && D.ncgp_solute == 0)) work = 0;
}

// Work map of a tuned kernel: one (instance, first item) entry per
// workgroup.  Map 1 lists the workgroups instance by instance.  Map 2
// assigns whole instances to the 8 XCDs (largest work first, each to the
// least loaded XCD) and interleaves the XCD queues, since workgroup b of a
// 1-D dispatch runs on XCD b mod 8; short queues are padded with empty
// entries (instance -1).
const std::pair<int2*, int>& Batch::work_map(int kernel, int map, int block) {
  const std::array<int, 3> key{kernel, map, block};
  for (auto& m : maps_)
    if (m.first == key) return m.second;
  const int ni = (int)h_sys_.size();
  std::vector<std::vector<int2>> per(ni);
  std::vector<double> work(ni);
  for (int k = 0; k < ni; ++k) {
    int items, lanes;
    tuned_items(kernel, h_sys_[k], precision_ == 0, items, lanes, work[k]);
    const int ipb = block / lanes;
    for (int b = 0; b * ipb < items; ++b) per[k].push_back(make_int2(k, b * ipb));
  }
  std::vector<int2> table;
  if (map == 1) {
    for (auto& v : per) table.insert(table.end(), v.begin(), v.end());
  } else {
    constexpr int kXcd = 8;
    std::vector<int> order(ni);
    for (int k = 0; k < ni; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    std::vector<double> load(kXcd, 0.0);
    std::vector<std::vector<int2>> q(kXcd);
    for (int k : order) {
      if (per[k].empty()) continue;
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += work[k];
      q[x].insert(q[x].end(), per[k].begin(), per[k].end());
    }
    size_t n = 0;
    for (auto& v : q) n = std::max(n, v.size());
    table.assign(n * kXcd, make_int2(-1, 0));
    for (int x = 0; x < kXcd; ++x)
      for (size_t sI = 0; sI < q[x].size(); ++sI) table[sI * kXcd + x] = q[x][sI];
  }
  int2* d = nullptr;
  if (!table.empty()) {
    HIPCHECK(hipMalloc(&d, table.size() * sizeof(int2)));
    HIPCHECK(hipMemcpy(d, table.data(), table.size() * sizeof(int2), hipMemcpyHostToDevice));
  }
  maps_.push_back({key, {d, (int)table.size()}});
  return maps_.back().second;
}

void Batch::launch_tuned(int kernel, const LaunchShape& sh) {
  const int ni = (int)h_sys_.size();
  const int blk = sh.block;
  int max_items = 0, lanes = 8;
  for (int k = 0; k < ni; ++k) {
    int items, l;
    double w;
    tuned_items(kernel, h_sys_[k], precision_ == 0, items, l, w);
    if (w > 0) max_items = std::max(max_items, items);
    lanes = l;
  }
  if (max_items == 0) return;
  const int2* map = nullptr;
  dim3 grid((max_items * lanes + blk - 1) / blk, ni);
  if (sh.map != 0) {
    const auto& m = work_map(kernel, sh.map, blk);
    if (m.second == 0) return;
    map = m.first;
    grid = dim3(m.second, 1);
  }
  const bool f64 = precision_ == 0;
  const bool mp = sh.map != 0;
  const bool pf = sh.pf != 0;
  const int en = launch_energies_ ? 1 : 0;
#define QGPU_LAUNCH(K)                                                                         \
  do {                                                                                         \
    if (f64 && mp) hipLaunchKernelGGL((K<double, 1>), grid, dim3(blk), 0, ls_, d_sys_, map); \
    else if (f64) hipLaunchKernelGGL((K<double, 0>), grid, dim3(blk), 0, ls_, d_sys_, map);  \
    else if (mp) hipLaunchKernelGGL((K<float, 1>), grid, dim3(blk), 0, ls_, d_sys_, map);    \
    else hipLaunchKernelGGL((K<float, 0>), grid, dim3(blk), 0, ls_, d_sys_, map);            \
  } while (0)
#define QGPU_LAUNCH_PF(K)                                                                              \
  do {                                                                                                 \
    if (f64 && mp && pf) hipLaunchKernelGGL((K<double, 1, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en); \
    else if (f64 && mp) hipLaunchKernelGGL((K<double, 1, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);  \
    else if (f64 && pf) hipLaunchKernelGGL((K<double, 0, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en);  \
    else if (f64) hipLaunchKernelGGL((K<double, 0, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);        \
    else if (mp && pf) hipLaunchKernelGGL((K<float, 1, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en);    \
    else if (mp) hipLaunchKernelGGL((K<float, 1, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);          \
    else if (pf) hipLaunchKernelGGL((K<float, 0, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en);          \
    else hipLaunchKernelGGL((K<float, 0, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);                  \
  } while (0)
  // water forces: the FP32 build fitted to 6 or 4 wavefronts per SIMD
#define QGPU_WF(PP, M, F)                                                                                 \
  do {                                                                                                    \
    if (sh.occ && sizeof(PP) == 4 && allatom_)                                                            \
      hipLaunchKernelGGL((k_water_forces<PP, M, F, 4, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en);     \
    else if (sh.occ && sizeof(PP) == 4)                                                                   \
      hipLaunchKernelGGL((k_water_forces<PP, M, F, 4, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);     \
    else if (allatom_)                                                                                    \
      hipLaunchKernelGGL((k_water_forces<PP, M, F, 6, 1>), grid, dim3(blk), 0, ls_, d_sys_, map, en);     \
    else                                                                                                  \
      hipLaunchKernelGGL((k_water_forces<PP, M, F, 6, 0>), grid, dim3(blk), 0, ls_, d_sys_, map, en);     \
  } while (0)
  switch (kernel) {
    case TK_WATER_LISTS: QGPU_LAUNCH(k_water_lists); break;
    case TK_SOLUTE_LISTS: QGPU_LAUNCH(k_solute_lists); break;
    case TK_WATER_FORCES:
      if (f64 && mp && pf) QGPU_WF(double, 1, 1);
      else if (f64 && mp) QGPU_WF(double, 1, 0);
      else if (f64 && pf) QGPU_WF(double, 0, 1);
      else if (f64) QGPU_WF(double, 0, 0);
      else if (mp && pf) QGPU_WF(float, 1, 1);
      else if (mp) QGPU_WF(float, 1, 0);
      else if (pf) QGPU_WF(float, 0, 1);
      else QGPU_WF(float, 0, 0);
      break;
    default: QGPU_LAUNCH_PF(k_solute_forces); break;
  }
#undef QGPU_WF
#undef QGPU_LAUNCH
#undef QGPU_LAUNCH_PF
}

void Batch::autotune(int reps, double* ms) {
  static const int kMaps[3] = {0, 1, 2};
  static const int kBlocks[2] = {256, 128};
  // the tuning launches rebuild the pair lists from the current coordinates,
  // which the next step does too only at a list step
  if (istep_ % nbcycle_ != 0) throw std::runtime_error("launch shapes can be tuned only at a pair-list step");
  HIPCHECK(hipStreamSynchronize(stream_));
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  // lists need the group centres, forces need lists; the force kernels are
  // timed as most steps run them (without the energy partials)
  launch_lists();
  launch_energies_ = false;
  for (int kernel = 0; kernel < TK_COUNT; ++kernel) {
    double best = 1e30, dflt = 0.0;
    LaunchShape pick = shape_[kernel];
    const int npf = (kernel == TK_WATER_FORCES || kernel == TK_SOLUTE_FORCES) ? 2 : 1;
    const int nocc = (kernel == TK_WATER_FORCES && precision_ == 1) ? 2 : 1;
    for (int m : kMaps)
      for (int blk : kBlocks)
        for (int pf = 0; pf < npf; ++pf)
        for (int oc = 0; oc < nocc; ++oc) {
        const LaunchShape sh{m, blk, pf, oc};
        launch_tuned(kernel, sh);   // warm-up (and map upload)
        double per = 1e30;
        for (int pass = 0; pass < 2; ++pass) {   // the faster of two timed passes
          HIPCHECK(hipEventRecord(a, stream_));
          for (int r = 0; r < reps; ++r) launch_tuned(kernel, sh);
          HIPCHECK(hipEventRecord(b, stream_));
          HIPCHECK(hipEventSynchronize(b));
          float t = 0.f;
          HIPCHECK(hipEventElapsedTime(&t, a, b));
          per = std::min(per, (double)t / reps);
        }
        if (m == 0 && blk == 256 && pf == 0 && oc == 0) dflt = per;
        if (per < best) { best = per; pick = sh; }
        }
    shape_[kernel] = pick;
    if (ms) { ms[2 * kernel] = best; ms[2 * kernel + 1] = dflt; }
  }
  HIPCHECK(hipGetLastError());
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  tuned_ = true;
  // captured graphs hold the old launch shapes
  drop_graphs();
}

void Batch::set_shape(int k, int map, int block, int pf, int occ) {
  if (k < 0 || k >= TK_COUNT || map < 0 || map > 2 || (block != 128 && block != 256) || pf < 0 || pf > 1 ||
      occ < 0 || occ > 1)
    throw std::runtime_error("invalid launch shape");
  shape_[k] = LaunchShape{map, block, (k == TK_WATER_FORCES || k == TK_SOLUTE_FORCES) ? pf : 0,
                          k == TK_WATER_FORCES ? occ : 0};
  tuned_ = true;
  drop_graphs();
}

// Two branches of independent kernels: the water side stays on the engine's
// stream, the solute side runs on a second stream forked from it and joined
// back (in a captured graph: parallel graph branches).  The branches write
// disjoint buffers.  Profiling runs every kernel on one stream.
void Batch::fork() {
  if (prof_ || !branches_) return;
  HIPCHECK(hipEventRecord(ev_fork_, stream_));
  HIPCHECK(hipStreamWaitEvent(stream2_, ev_fork_, 0));
  ls_ = stream2_;
}

void Batch::join() {
  if (ls_ == stream_) return;
  HIPCHECK(hipEventRecord(ev_join_, stream2_));
  HIPCHECK(hipStreamWaitEvent(stream_, ev_join_, 0));
  ls_ = stream_;
}

// put_back_in_box of the periodic-box instances at a step's list update
// (md_run, src/md.f90:6483-6485), then the FP32 coordinate copies
void Batch::launch_put_back() {
  if (!put_back_) return;
  const int ni = (int)h_sys_.size();
  hipLaunchKernelGGL(k_put_back, dim3(1, ni), dim3(256), 0, stream_, d_sys_);
  hipLaunchKernelGGL(k_x_to_f32, dim3((3 * max_natom_ + 255) / 256, ni), dim3(256), 0, stream_, d_sys_);
}

void Batch::launch_lists(bool q_waters) {
  const int ni = (int)h_sys_.size();
  dim3 g0((max_ncgp_ + 255) / 256, ni);
  { pb(0); hipLaunchKernelGGL(k_cgp_centers, g0, dim3(256), 0, stream_, d_sys_); pe(0); }
  // water lists (ww, water side of pw, water multipoles) on the engine's
  // stream; solute lists (pp, solute side of pw, solute multipoles) beside
  fork();
  if (max_ncgp_solute_ > 0) { pb(2); launch_tuned(TK_SOLUTE_LISTS, shape_[TK_SOLUTE_LISTS]); pe(2); }
  if (q_dynamic_) hipLaunchKernelGGL(k_q_lists, dim3(1, ni), dim3(1024), 0, ls_, d_sys_, q_waters ? 1 : 0);
  const hipStream_t side = ls_;
  ls_ = stream_;
  if (max_nwat_ > 0) { pb(1); launch_tuned(TK_WATER_LISTS, shape_[TK_WATER_LISTS]); pe(1); }
  ls_ = side;
  join();
}

// Whether step s can be an energy-file step, for leaving the frame kernels
// out of the launches of other steps (they check the step counter on the
// device too).  A captured cycle is replayed every NBcycle steps, so its
// frame positions are fixed only when iene divides NBcycle.
bool Batch::record_possible(long s) const {
  if (iene_ <= 0) return false;
  if (use_graph_ && nbcycle_ % iene_ != 0) return true;
  return (s - ebase_) % iene_ == 0;
}

void Batch::drop_graphs() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_exec_final_) (void)hipGraphExecDestroy(graph_exec_final_);
  graph_exec_ = graph_exec_final_ = nullptr;
}

void Batch::launch_forces(bool step, long s, bool energies) {
  const int ni = (int)h_sys_.size();
  launch_energies_ = energies;
  // excluded_freeze: the coordinates the terms after fix_shell see
  if (freeze_)
    hipLaunchKernelGGL(k_freeze_x, dim3((3 * max_natom_ + 255) / 256, ni), dim3(256), 0, stream_, d_sys_);
  // side branch: the short bonded/Q-Q term and Q-atom kernels first, so
  // that they overlap the water forces; then the solute forces (assign the
  // solute atoms' gradient) and the Q-atom sums (added to the Q atoms'
  // gradient after the solute forces)
  fork();
  if (max_terms_ > 0)
    { pb(5); hipLaunchKernelGGL(k_terms, dim3((max_terms_ + 255) / 256, ni), dim3(256), 0, ls_, d_sys_); pe(5); }
  const dim3 gq((max_nqat_ * kQChunks * 64 + 255) / 256, ni), gr((max_nqat_ + 63) / 64, ni);
  auto q_fp64 = [&](int shadow) {
    if (max_nstates_ <= 2) hipLaunchKernelGGL((k_q_forces<double, 2>), gq, dim3(256), 0, ls_, d_sys_, shadow);
    else if (max_nstates_ <= 4) hipLaunchKernelGGL((k_q_forces<double, 4>), gq, dim3(256), 0, ls_, d_sys_, shadow);
    else hipLaunchKernelGGL((k_q_forces<double, kMaxQStates>), gq, dim3(256), 0, ls_, d_sys_, shadow);
  };
  // mixed precision with energy frames: the Q energies of each recorded
  // frame are also evaluated with FP64 pair arithmetic (the kernels return
  // at once on other steps)
  const bool shadow = max_nqat_ > 0 && precision_ == 1 && record_possible(s);
  if (max_nqat_ > 0) {
    pb(6);
    if (precision_ == 0) {
      q_fp64(0);
    } else {
      if (max_nstates_ <= 2) hipLaunchKernelGGL((k_q_forces<float, 2>), gq, dim3(256), 0, ls_, d_sys_, 0);
      else if (max_nstates_ <= 4) hipLaunchKernelGGL((k_q_forces<float, 4>), gq, dim3(256), 0, ls_, d_sys_, 0);
      else hipLaunchKernelGGL((k_q_forces<float, kMaxQStates>), gq, dim3(256), 0, ls_, d_sys_, 0);
    }
    pe(6);
    if (shadow) q_fp64(1);
  }
  if (max_ncgp_solute_ > 0) { pb(4); launch_tuned(TK_SOLUTE_FORCES, shape_[TK_SOLUTE_FORCES]); pe(4); }
  if (max_nqat_ > 0) {
    hipLaunchKernelGGL(k_q_reduce, gr, dim3(64), 0, ls_, d_sys_, 0);
    if (shadow) hipLaunchKernelGGL(k_q_reduce, gr, dim3(64), 0, ls_, d_sys_, 1);
  }
  // main branch: water forces (assign the waters' gradient), then the
  // polarisation restraint (its own per-water gradient buffer, which the
  // integrator adds)
  const hipStream_t side = ls_;
  ls_ = stream_;
  if (max_nwat_ > 0) {
    pb(3); launch_tuned(TK_WATER_FORCES, shape_[TK_WATER_FORCES]); pe(3);
    dim3 g2((max_nwat_ + 255) / 256, ni);
    { pb(8); hipLaunchKernelGGL(k_watpol_theta, g2, dim3(256), 0, stream_, d_sys_); pe(8); }
    if (max_nshell_ > 0) {
      dim3 g3(max_nshell_ * kWpolSplit, ni);
      pb(9);
      hipLaunchKernelGGL(k_watpol_force, g3, dim3(kWpolBlock), 0, stream_, d_sys_);
      hipLaunchKernelGGL(k_watpol_shell, dim3(1, ni), dim3(64), 0, stream_, d_sys_);
      pe(9);
    }
  }
  ls_ = side;
  join();
  // outside a step the polarisation gradient and then the term slots (which
  // may hold water atoms: unconstrained water bonds) are added here, in the
  // order the integrator adds them within a step
  if (max_nwat_ > 0 && !step)
    hipLaunchKernelGGL(k_watpol_gather, dim3((max_nwat_ + 255) / 256, ni), dim3(256), 0, stream_, d_sys_);
  if (max_slot_atoms_ > 0 && !step)
    { pb(7); hipLaunchKernelGGL(k_term_gather, dim3((max_slot_atoms_ + 255) / 256, ni), dim3(256), 0, stream_, d_sys_); pe(7); }
  if (far_model_) launch_far_model();
}

// Batch-wide far-field arrays (one row per charge group of every instance).
void Batch::ensure_far() {
  if (far_x_) return;
  int off = 0;
  for (auto& D : h_sys_) { D.far_off = off; off += D.ncgp; }
  far_total_ = off;
  HIPCHECK(hipMalloc(&far_x_, (size_t)std::max(1, off) * 4 * sizeof(float)));
  HIPCHECK(hipMalloc(&far_y_, (size_t)std::max(1, off) * 3 * sizeof(float)));
  HIPCHECK(hipMalloc(&far_frame_, (size_t)std::max(1, off) * 9 * sizeof(double)));
  for (auto& D : h_sys_) { D.far_x = far_x_; D.far_y = far_y_; D.far_frame = far_frame_; }
  sync_sys();
}

// Learned far field of every instance: group features, one batched MLP
// forward pass over all groups, per-atom energy and gradient.
void Batch::launch_far_model() {
  const int ni = (int)h_sys_.size();
  const dim3 g((max_ncgp_ + 255) / 256, ni);
  hipLaunchKernelGGL(k_far_features, g, dim3(256), 0, stream_, d_sys_);
  far_model_->predict_dev(far_x_, far_total_, far_y_, stream_);
  hipLaunchKernelGGL(k_far_apply, g, dim3(256), 0, stream_, d_sys_, far_model_->norm + 8, far_model_->ystd_d);
}

int Batch::far_samples(int k, float* X, float* Y, int cap, bool lrf_estimate) {
  ensure_far();
  const int ni = (int)h_sys_.size();
  const dim3 g((max_ncgp_ + 255) / 256, ni);
  if (!lrf_estimate) {
    if (far_model_) throw std::runtime_error("exact far-field samples need the LRF multipoles (no model attached)");
    launch_lists();
    hipLaunchKernelGGL(k_far_features, g, dim3(256), 0, stream_, d_sys_);
    hipLaunchKernelGGL(k_far_targets, g, dim3(256), 0, stream_, d_sys_);
  } else {
    hipLaunchKernelGGL(k_far_features, g, dim3(256), 0, stream_, d_sys_);
    hipLaunchKernelGGL(k_far_lrf_eval, g, dim3(256), 0, stream_, d_sys_);
  }
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  const int n = std::min(cap, D.ncgp);
  if (X) HIPCHECK(hipMemcpy(X, far_x_ + (size_t)D.far_off * 4, (size_t)n * 4 * sizeof(float), hipMemcpyDeviceToHost));
  if (Y) HIPCHECK(hipMemcpy(Y, far_y_ + (size_t)D.far_off * 3, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost));
  return D.ncgp;
}

void Batch::set_far_model(Mlp* m) {
  HIPCHECK(hipStreamSynchronize(stream_));
  ensure_far();
  far_model_ = m;
  for (size_t k = 0; k < h_sys_.size(); ++k) {
    h_sys_[k].far_mode = m ? 1 : 0;
    h_sys_[k].use_lrf = m ? false : setups_[k]->inp.use_lrf;
  }
  sync_sys();
  drop_graphs();
}

void Batch::launch_integrate(long s) {
  const int ni = (int)h_sys_.size();
  const int blocks = (std::max(max_units_, 1) + kIntegrateBlock - 1) / kIntegrateBlock;
  // energy-file frame of this step (the kernel checks the step counter)
  if (record_possible(s)) { pb(10); hipLaunchKernelGGL(k_energy_reduce, dim3(1, ni), dim3(256), 0, stream_, d_sys_, 1); pe(10); }
  // Nose-Hoover chains: this step's velocity scales
  if (nose_hoover_) hipLaunchKernelGGL(k_nh_prop, dim3(1, ni), dim3(64), 0, stream_, d_sys_);
  { pb(11); hipLaunchKernelGGL(k_integrate, dim3(blocks, ni), dim3(kIntegrateBlock), 0, stream_, d_sys_); pe(11); }
  { pb(12); hipLaunchKernelGGL(k_temperature, dim3(1, ni), dim3(256), 0, stream_, d_sys_, blocks); pe(12); }
}

// Energies of the monitored group pairs at the current coordinates
// (nonbond_monitor, called by md_run at every output step after pot_energy,
// src/md.f90:6562-6564): per pair Vwsum, Vwel, Vwlj, then Vel and Vlj of
// every state.  Returns the number of values (written only if cap allows).
int Batch::monitor(int k, double* out, int cap) {
  const DevSys& D = h_sys_[k];
  const int stride = kMon + 2 * std::max(1, D.nstates);
  const int need = D.mon_n * stride;
  if (need == 0 || need > cap) return need;
  HIPCHECK(hipStreamSynchronize(stream_));
  hipLaunchKernelGGL(k_monitor, dim3(max_mon_, (int)h_sys_.size()), dim3(256), 0, stream_, d_sys_);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(out, D.mon_out, need * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
  return need;
}

void Batch::status(int k, int* bits, int* at_step) {
  HIPCHECK(hipStreamSynchronize(stream_));
  int r[kStatusInts];
  HIPCHECK(hipMemcpy(r, h_sys_[k].status, sizeof(r), hipMemcpyDeviceToHost));
  if (bits) *bits = r[0];
  if (at_step) *at_step = r[1] - 1;
}

// Replaces instance k's list rows by rows of the given capacities (0 keeps
// one); the next list update fills them.
void Batch::set_list_caps(int k, int ww, int wp, int pw) {
  HIPCHECK(hipStreamSynchronize(stream_));
  DevSys& D = h_sys_[k];
  Inst& I = *inst_[k];
  if (ww < 0 || wp < 0 || pw < 0) throw std::runtime_error("negative list capacity");
  if (ww) { ww = std::min(ww, std::max(1, D.nwat)); I.release(D.ww_nb); D.ww_nb = I.alloc<int>((size_t)D.nwat * ww); D.ww_cap = ww; }
  if (wp) { wp = std::min(wp, std::max(1, D.ncgp_solute)); I.release(D.wp_nb); D.wp_nb = I.alloc<int>((size_t)D.nwat * wp); D.wp_cap = wp; }
  if (pw) { pw = std::min(pw, std::max(1, D.nwat)); I.release(D.pw_nb); D.pw_nb = I.alloc<int>((size_t)D.ncgp_solute * pw); D.pw_cap = pw; }
  sync_sys();
}

void Batch::make_lists() {
  if (!dead_.empty()) throw std::runtime_error(dead_);
  launch_lists();
  HIPCHECK(hipStreamSynchronize(stream_));
  // rows regrown: built again from the same coordinates
  if (check_failures()) {
    launch_lists();
    HIPCHECK(hipStreamSynchronize(stream_));
    check_failures();
  }
}
// A force evaluation that includes only the term groups of `mask` (TM_*),
// for per-term diagnostics; the batch then evaluates every term again.
void Batch::forces_terms(int mask) {
  if (mask <= 0 || mask > TM_ALL) throw std::runtime_error("invalid term mask");
  HIPCHECK(hipStreamSynchronize(stream_));
  for (auto& D : h_sys_) D.term_mask = mask;
  sync_sys();
  try {
    forces();
  } catch (...) {
    for (auto& D : h_sys_) D.term_mask = TM_ALL;
    sync_sys();
    throw;
  }
  for (auto& D : h_sys_) D.term_mask = TM_ALL;
  sync_sys();
}

void Batch::forces() {
  if (!dead_.empty()) throw std::runtime_error(dead_);
  launch_forces(false, istep_);
  HIPCHECK(hipStreamSynchronize(stream_));
  check_failures();
}

void Batch::one_step(bool energies) {
  if (istep_ % nbcycle_ == 0) {
    launch_put_back();
    launch_lists();
    // a volume move at this list update (md_run, src/md.f90:6538-6544)
    if (npt_ && istep_ > 0 && istep_ % ivol_ == 0) mc_volume();
  }
  launch_forces(true, istep_, energies || npt_energy_step(istep_));
  launch_integrate(istep_);
  ++istep_;
}

// Whether step s is the step before a volume move: its pot_energy's bonded
// and Q-bonded energies enter the move's potentials, so it forms energies.
bool Batch::npt_energy_step(long s) const {
  const long n = s + 1;
  return npt_ && n % nbcycle_ == 0 && n % ivol_ == 0;
}

// new_potential (src/md.f90:19276-19434) of instance k after a force
// evaluation with energies: its nonbonded, LRF, Q and restraint energies,
// with the bonded and Q-bonded energies of old_e (qgpu_energies order; the
// last pot_energy), summed as the reference sums them (no solvent torsions
// or impropers).
double Batch::new_potential(int k, const double* old_e) {
  double e[QGPU_E_COUNT];
  energies(k, e, QGPU_E_COUNT);
  return old_e[QGPU_E_P_BOND] + old_e[QGPU_E_W_BOND] + old_e[QGPU_E_P_ANGLE] + old_e[QGPU_E_W_ANGLE] +
         old_e[QGPU_E_P_TORSION] + old_e[QGPU_E_P_IMPROPER] + e[QGPU_E_PP_EL] + e[QGPU_E_PP_VDW] +
         e[QGPU_E_PW_EL] + e[QGPU_E_PW_VDW] + e[QGPU_E_WW_EL] + e[QGPU_E_WW_VDW] + old_e[QGPU_E_Q_BOND] +
         old_e[QGPU_E_Q_ANGLE] + old_e[QGPU_E_Q_TORSION] + old_e[QGPU_E_Q_IMPROPER] + e[QGPU_E_QX_EL] +
         e[QGPU_E_QX_VDW] + e[QGPU_E_RST_PROTEIN] + e[QGPU_E_LRF];
}

// MC_volume (src/md.f90:18931-19272) of the constant-pressure instances,
// at the list update of a volume-move step (after put_back_in_box and the
// lists).  The host reads the potentials back, draws the move from the
// instance's randm stream and judges it; the device scales the molecules
// (k_mc_scale) and builds the lists again.  Per instance: the potential
// with the lists of this update; a volume change of at most
// max_volume_displ, the box edges scaled by the cube root of the volume
// ratio; every molecule's centre of mass scaled about the box centre; with
// LRF, every list and the multipoles built again; the potential again; the
// move accepted when deltaW = deltaE + P deltaV - nmol kT ln(V'/V) <= 0,
// else when a second draw does not exceed exp(-deltaW / kT).  A rejected
// move restores x and the box and (with LRF) builds every list again from
// the restored coordinates except the Q-water list, which the reference
// keeps from the scaled ones.  Instances without constant pressure keep
// their state: their lists are built again from unchanged coordinates.
void Batch::mc_volume() {
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  if (check_failures()) make_lists();   // list rows regrown: this update's lists built again
  const int ni = size();
  std::vector<std::array<double, QGPU_E_COUNT>> prev(ni);
  std::vector<double> old_pot(ni, 0.0), old_v(ni, 0.0), dv(ni, 0.0), new_v(ni, 0.0);
  bool lrf_on = false;
  for (int k = 0; k < ni; ++k) {
    if (!vol_[k].on) continue;
    energies(k, prev[k].data(), QGPU_E_COUNT);
    const DevSys& D = h_sys_[k];
    HIPCHECK(hipMemcpyAsync(vol_[k].x_saved, D.x, 3 * (size_t)D.natom * sizeof(double), hipMemcpyDeviceToDevice,
                            stream_));
    lrf_on = setups_[k]->inp.use_lrf;
  }
  forces();
  for (int k = 0; k < ni; ++k)
    if (vol_[k].on) old_pot[k] = new_potential(k, prev[k].data());
  // the moves: new box lengths, checked against the cut-offs
  for (int k = 0; k < ni; ++k) {
    if (!vol_[k].on) continue;
    DevSys& D = h_sys_[k];
    qsys::Setup& s = *setups_[k];
    // Langevin instances: randm's stream is on the device
    if (D.rng) {
      HIPCHECK(hipMemcpy(&s.rng.irand, D.rng, sizeof(int), hipMemcpyDeviceToHost));
      s.rng.started = true;
    }
    old_v[k] = D.boxl[0] * D.boxl[1] * D.boxl[2];
    double r = s.rng.next(vol_[k].seed);
    r = r * 2 - 1;
    dv[k] = r * s.inp.max_vol_displ;
    new_v[k] = dv[k] + old_v[k];
    const double cubr = std::pow(new_v[k] / old_v[k], 1.0 / 3.0);
    for (int c = 0; c < 3; ++c) {
      D.boxl_old[c] = D.boxl[c];
      D.boxl[c] = D.boxl[c] * cubr;
      D.inv_boxl[c] = 1.0 / D.boxl[c];
      D.boxl_f[c] = (float)D.boxl[c];
      D.inv_boxl_f[c] = (float)D.inv_boxl[c];
    }
    const double bmin = std::min(D.boxl[0], std::min(D.boxl[1], D.boxl[2]));
    const qsys::MDInput& in = s.inp;
    const char* small = nullptr;
    if (!(bmin > in.rcpp * 2)) small = "Solute-solute cut-off radii too large";
    else if (!(bmin > in.rcww * 2)) small = "Solvent-solvent cut-off radii too large";
    else if (!(bmin > in.rcpw * 2)) small = "Solute-solvent cut-off radii too large";
    else if (!(bmin > in.rcq * 2)) small = "Q-atom cut-off radii too large";
    else if (in.use_lrf && !(bmin > in.rclrf * 2)) small = "LRF cut-off radii too large";
    if (small) {
      dead_ = "run stopped: instance " + std::to_string(k) + ": " + small + " at the volume move of step " +
              std::to_string(istep_);
      throw std::runtime_error(dead_);
    }
    D.mc_move = 1;
  }
  sync_sys();
  hipLaunchKernelGGL(k_mc_scale, dim3(1, ni), dim3(256), 0, stream_, d_sys_);
  HIPCHECK(hipGetLastError());
  refresh_xf();
  for (auto& D : h_sys_) D.mc_move = 0;
  sync_sys();
  if (lrf_on) make_lists();
  forces();
  // Metropolis test
  bool any_rejected = false;
  for (int k = 0; k < ni; ++k) {
    if (!vol_[k].on) continue;
    qsys::Setup& s = *setups_[k];
    const qsys::MDInput& in = s.inp;
    const double delta_e = new_potential(k, prev[k].data()) - old_pot[k];
    const double delta_w =
        delta_e + in.pressure * dv[k] - s.nmol * qsys::kBoltz * in.temp0 * std::log(new_v[k] / old_v[k]);
    bool acc = true, drawn = false;
    if (!(delta_w <= 0.0)) {
      const double u = s.rng.next(vol_[k].seed);
      acc = !(u > std::exp(-delta_w / qsys::kBoltz / in.temp0));
      drawn = true;
    }
    vol_[k].tries += 1;
    vol_[k].log.push_back(acc ? (drawn ? 2 : 1) : 0);
    if (h_sys_[k].rng) HIPCHECK(hipMemcpy(h_sys_[k].rng, &s.rng.irand, sizeof(int), hipMemcpyHostToDevice));
    if (acc) {
      vol_[k].accepted += 1;
      continue;
    }
    any_rejected = true;
    DevSys& D = h_sys_[k];
    HIPCHECK(hipMemcpyAsync(D.x, vol_[k].x_saved, 3 * (size_t)D.natom * sizeof(double), hipMemcpyDeviceToDevice,
                            stream_));
    for (int c = 0; c < 3; ++c) {
      D.boxl[c] = D.boxl_old[c];
      D.inv_boxl[c] = 1.0 / D.boxl[c];
      D.boxl_f[c] = (float)D.boxl[c];
      D.inv_boxl_f[c] = (float)D.inv_boxl[c];
    }
  }
  if (!any_rejected) return;
  HIPCHECK(hipStreamSynchronize(stream_));
  sync_sys();
  refresh_xf();
  if (lrf_on) {
    launch_lists(false);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream_));
    if (check_failures()) {
      launch_lists(false);
      HIPCHECK(hipStreamSynchronize(stream_));
      check_failures();
    }
  }
}

int Batch::volume_moves(int k, int* out, int cap) const {
  const VolumeMoves& m = vol_[k];
  if (cap > 0) out[0] = m.tries;
  if (cap > 1) out[1] = m.accepted;
  for (int i = 0; i < (int)m.log.size() && i + 2 < cap; ++i) out[i + 2] = m.log[i];
  return m.tries;
}

// One list-update cycle (NBcycle steps) captured as a hipGraph: the kernels
// read the step number from the device counter, so the same graph serves
// every cycle.  Rebuilt when the energy interval changes.
void Batch::launch_cycle_graph(bool final_energies) {
  // both variants are captured together, so a run never instantiates one
  // between timed calls
  for (int fin = 0; fin < 2; ++fin) {
    hipGraphExec_t& ge = fin ? graph_exec_final_ : graph_exec_;
    if (ge) continue;
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nbcycle_; ++k) {
      if (k == 0) {
        launch_put_back();
        launch_lists();
      }
      launch_forces(true, istep_ + k, fin && k == nbcycle_ - 1);
      launch_integrate(istep_ + k);
    }
    HIPCHECK(hipStreamEndCapture(stream_, &g));
    HIPCHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIPCHECK(hipGraphDestroy(g));
  }
  HIPCHECK(hipGraphLaunch(final_energies ? graph_exec_final_ : graph_exec_, stream_));
  istep_ += nbcycle_;
}

void Batch::profile(int nsteps, double* ms, int cap) {
  HIPCHECK(hipStreamSynchronize(stream_));
  for (auto& e : pev_)
    if (!e) HIPCHECK(hipEventCreate(&e));
  double tot[kNumKernels] = {0};
  int cnt[kNumKernels] = {0};
  prof_ = true;
  try {
    for (int n = 0; n < nsteps; ++n) {
      for (int& u : pused_) u = 0;
      one_step();
      HIPCHECK(hipStreamSynchronize(stream_));
      for (int k = 0; k < kNumKernels; ++k)
        if (pused_[k]) {
          float t = 0;
          HIPCHECK(hipEventElapsedTime(&t, pev_[2 * k], pev_[2 * k + 1]));
          tot[k] += t;
          cnt[k] += 1;
        }
    }
  } catch (...) {
    prof_ = false;
    throw;
  }
  prof_ = false;
  check_failures();
  // cost of an event pair with nothing between them on this stream, removed
  // from every kernel's time
  double ovh = 0.0;
  for (int n = 0; n < 20; ++n) {
    HIPCHECK(hipEventRecord(pev_[0], stream_));
    HIPCHECK(hipEventRecord(pev_[1], stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, pev_[0], pev_[1]));
    ovh += t / 20.0;
  }
  for (int k = 0; k < kNumKernels && k < cap; ++k)
    ms[k] = cnt[k] ? std::max(0.0, tot[k] / cnt[k] - ovh) : 0.0;
}

// Launches nsteps steps: single steps up to a list step, whole NBcycles as
// the captured graph, single steps for the rest.
void Batch::launch_steps(int nsteps, bool final_energies) {
  // energy partials only in the last step, when final_energies
  auto last = [&](int n) { return final_energies && n == nsteps - 1; };
  // volume moves need the host at their list updates: no captured cycles
  const bool graph = use_graph_ && !npt_;
  int n = 0;
  while (n < nsteps && (!graph || istep_ % nbcycle_ != 0 || nsteps - n < nbcycle_)) {
    one_step(last(n));
    ++n;
  }
  while (graph && nsteps - n >= nbcycle_) {
    launch_cycle_graph(final_energies && nsteps - n == nbcycle_);
    n += nbcycle_;
  }
  while (n < nsteps) {
    one_step(last(n));
    ++n;
  }
}

// After a synchronisation: whether a kernel recorded a failure.  A list row
// that outgrew its capacity is regrown and the batch resumes at the list
// update where every instance stopped (returns true: steps to redo).  Any
// other failure (SHAKE, NaN temperature, shell capacity) stops the batch for
// good at the failing step, as die() stops qdyn (src/md.f90:554-613), and
// throws.
bool Batch::check_failures() {
  unsigned long long f = kNoFailure;
  HIPCHECK(hipMemcpy(&f, d_fail_, sizeof(f), hipMemcpyDeviceToHost));
  if (f == kNoFailure) return false;
  const int ni = (int)h_sys_.size();
  std::vector<int> st((size_t)ni * kStatusInts), ctr(ni);
  HIPCHECK(hipMemcpy(st.data(), d_status_, st.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(ctr.data(), d_step_, ctr.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::string msg;
  bool fatal = false;
  for (int k = 0; k < ni; ++k) {
    const int* r = &st[(size_t)k * kStatusInts];
    if (!r[0]) continue;
    if (r[0] & ~ST_LIST_OVERFLOW) fatal = true;
    static const char* kWhat[4] = {"pair-list capacity exceeded", "water-shell capacity exceeded",
                                   "shake failure", "NaN temperature"};
    for (int b = 0; b < 4; ++b)
      if (r[0] & (1 << b))
        msg += "instance " + std::to_string(k) + ": " + kWhat[b] + " at step " + std::to_string(r[1] - 1) + "; ";
  }
  // every instance stopped at the same step
  istep_ = ctr[0];
  if (fatal) {
    dead_ = "run stopped: " + msg.substr(0, msg.size() - 2);
    throw std::runtime_error(dead_);
  }
  for (int k = 0; k < ni; ++k) {
    const int* r = &st[(size_t)k * kStatusInts];
    if (r[0] & ST_LIST_OVERFLOW) grow_lists(k, r[2], r[3], r[4]);
  }
  HIPCHECK(hipMemset(d_status_, 0, st.size() * sizeof(int)));
  HIPCHECK(hipMemcpy(d_fail_, &kNoFailure, sizeof(kNoFailure), hipMemcpyHostToDevice));
  return true;
}

// New list rows for instance k that hold rows of the given lengths (with
// headroom); the next list update fills them.
void Batch::grow_lists(int k, int need_ww, int need_wp, int need_pw) {
  DevSys& D = h_sys_[k];
  Inst& I = *inst_[k];
  auto bigger = [](int cap, int need, int most) {
    if (need <= cap) return cap;
    return std::max(1, std::min(most, (need + need / 4 + 32 + 15) / 16 * 16));
  };
  const int ww = bigger(D.ww_cap, need_ww, D.nwat), wp = bigger(D.wp_cap, need_wp, D.ncgp_solute);
  const int pw = bigger(D.pw_cap, need_pw, D.nwat);
  if (ww != D.ww_cap) { I.release(D.ww_nb); D.ww_nb = I.alloc<int>((size_t)D.nwat * ww); D.ww_cap = ww; }
  if (wp != D.wp_cap) { I.release(D.wp_nb); D.wp_nb = I.alloc<int>((size_t)D.nwat * wp); D.wp_cap = wp; }
  if (pw != D.pw_cap) { I.release(D.pw_nb); D.pw_nb = I.alloc<int>((size_t)D.ncgp_solute * pw); D.pw_cap = pw; }
  ++list_regrowths_;
  sync_sys();
}

void Batch::step(int nsteps) {
  if (!dead_.empty()) throw std::runtime_error(dead_);
  if (!tuned_ && autotune_on_ && istep_ % nbcycle_ == 0 && nsteps > 0) autotune(3);
  // work maps are uploaded here, outside any graph capture
  for (int k = 0; k < TK_COUNT; ++k)
    if (shape_[k].map != 0) work_map(k, shape_[k].map, shape_[k].block);
  HIPCHECK(hipEventRecord(ev0, stream_));
  // failures are checked every kCheckSteps steps and at the end; after a
  // failure every kernel that changes the state returns at once, so the
  // steps launched before the check do nothing
  constexpr int kCheckSteps = 1000;
  const int target = istep_ + nsteps;
  while (istep_ < target) {
    const int m = std::min(target - istep_, kCheckSteps);
    launch_steps(m, istep_ + m == target);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream_));
    check_failures();
  }
  HIPCHECK(hipEventRecord(ev1, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
}

// Periodic box of instance k: box lengths and the current centre (6
// values); false for a sphere.
bool Batch::box(int k, double* out) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  for (int c = 0; c < 3; ++c) out[c] = D.boxl[c];
  HIPCHECK(hipMemcpy(out + 3, D.boxc, 3 * sizeof(double), hipMemcpyDeviceToHost));
  return D.pbc != 0;
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
  if (v) {
    HIPCHECK(hipMemcpy(h_sys_[k].v, v, n * 8, hipMemcpyHostToDevice));
    // temperature of the new velocities, as md_run computes it before its
    // first step (src/md.f90:6400-6404): the Berendsen scales and Tfree
    // (which the radial water restraint reads) follow the state
    std::vector<double> scal(8, 0.0);
    HIPCHECK(hipMemcpy(scal.data(), h_sys_[k].scal, 8 * sizeof(double), hipMemcpyDeviceToHost));
    initial_scales(*setups_[k], v, scal.data());
    HIPCHECK(hipMemcpy(h_sys_[k].scal, scal.data(), 8 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h_sys_[k].energies + E_KINETIC, &scal[4], sizeof(double), hipMemcpyHostToDevice));
  }
  if (x) refresh_xf();
}

int Batch::shake(int k, const double* xx, double* x) {
  if (!dead_.empty()) throw std::runtime_error(dead_);
  HIPCHECK(hipStreamSynchronize(stream_));
  const DevSys& D = h_sys_[k];
  const size_t n = 3 * (size_t)D.natom;
  double* buf = nullptr;
  HIPCHECK(hipMalloc(&buf, 2 * n * sizeof(double) + 2 * sizeof(int)));
  int res[2] = {0, 0};
  try {
    int* cnt = reinterpret_cast<int*>(buf + 2 * n);
    HIPCHECK(hipMemcpy(buf, xx, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(buf + n, x, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(cnt, res, sizeof(res), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_shake, dim3((D.nunits + 255) / 256), dim3(256), 0, stream_, d_sys_ + k, buf, buf + n, cnt);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(x, buf + n, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(res, cnt, sizeof(res), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  HIPCHECK(hipFree(buf));
  return res[1] ? -1 : res[0];
}

void Batch::pot_energy_nonbonds(int k, const double* x, double* d, double* e7, double* eq4) {
  set_state(k, x, nullptr);
  forces_terms(TM_WW | TM_PW | TM_PP | TM_QX | TM_LRF);
  const size_t n = 3 * (size_t)h_sys_[k].natom;
  std::vector<double> g(n);
  get_gradient(k, g.data());
  for (size_t i = 0; i < n; ++i) d[i] = d[i] + g[i];
  double E[QGPU_E_COUNT];
  energies(k, E, QGPU_E_COUNT);
  const int pick[7] = {QGPU_E_PP_EL, QGPU_E_PP_VDW, QGPU_E_PW_EL, QGPU_E_PW_VDW,
                       QGPU_E_WW_EL, QGPU_E_WW_VDW, QGPU_E_LRF};
  for (int c = 0; c < 7; ++c) e7[c] = E[pick[c]];
  // per-state Q-solute and Q-water sums of this evaluation's per-Q-atom
  // energies (the Q-neighbour pairs of nonbond_qq, which pot_energy also
  // adds into EQ%qp, are not part of pot_energy_nonbonds)
  const int ns = setups_[k]->nstates, nq = h_sys_[k].nqat;
  if (ns > 0 && eq4) {
    std::vector<double> q((size_t)std::max(1, nq) * ns * 4, 0.0);
    if (nq > 0) HIPCHECK(hipMemcpy(q.data(), h_sys_[k].e_q, 8 * q.size(), hipMemcpyDeviceToHost));
    for (int st = 0; st < ns; ++st)
      for (int c = 0; c < 4; ++c) {
        double sum = 0.0;
        for (int iq = 0; iq < nq; ++iq) sum += q[((size_t)iq * ns + st) * 4 + c];
        eq4[4 * st + c] = sum;
      }
  }
}

void Batch::get_gradient(int k, double* d) {
  HIPCHECK(hipStreamSynchronize(stream_));
  HIPCHECK(hipMemcpy(d, h_sys_[k].d, 3 * (size_t)h_sys_[k].natom * 8, hipMemcpyDeviceToHost));
}

void Batch::energies(int k, double* out, int cap) {
  const int ni = (int)h_sys_.size();
  hipLaunchKernelGGL(k_energy_reduce, dim3(1, ni), dim3(256), 0, stream_, d_sys_, 0);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  double E[E_NTERMS];
  HIPCHECK(hipMemcpy(E, h_sys_[k].energies, sizeof(E), hipMemcpyDeviceToHost));
  const qsys::Setup& s = *setups_[k];
  std::vector<double> q((size_t)std::max(1, s.nstates) * kQTerms, 0.0);
  if (s.nstates > 0)
    HIPCHECK(hipMemcpy(q.data(), h_sys_[k].qenergies, 8 * q.size(), hipMemcpyDeviceToHost));
  double p_bond = E[E_P_BOND], w_bond = E[E_W_BOND], p_angle = E[E_P_ANGLE], w_angle = E[E_W_ANGLE];
  double p_tor = E[E_P_TORSION], w_tor = E[E_W_TORSION], p_imp = E[E_P_IMPROPER], w_imp = E[E_W_IMPROPER];
  // lambda-weighted Q sums of pot_energy (src/md.f90:16186-16211)
  double q_bond = 0, q_angle = 0, q_tor = 0, q_imp = 0, qx_el = 0, qx_vdw = 0, rst_protein = E[E_RST_PROTEIN];
  for (int st = 0; st < s.nstates; ++st) {
    const double* e = &q[(size_t)st * kQTerms];
    const double lam = s.lambda[st];
    double el = e[0] + e[2] + e[4], vdw = e[1] + e[3] + e[5];
    q_bond = q_bond + e[6] * lam;
    q_angle = q_angle + e[7] * lam;
    q_tor = q_tor + e[8] * lam;
    q_imp = q_imp + e[9] * lam;
    qx_el = qx_el + el * lam;
    qx_vdw = qx_vdw + vdw * lam;
    rst_protein = rst_protein + e[10] * lam;
  }
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

// per state: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw total (EQ(s)%total,
// src/md.f90:16190-16194), Q bond, angle, torsion, improper, restraint;
// from the last energy reduction
void Batch::qenergies(int k, double* out, int cap) {
  const int ni = (int)h_sys_.size();
  hipLaunchKernelGGL(k_energy_reduce, dim3(1, ni), dim3(256), 0, stream_, d_sys_, 0);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream_));
  const qsys::Setup& s = *setups_[k];
  std::vector<double> q((size_t)std::max(1, s.nstates) * kQTerms, 0.0);
  if (s.nstates > 0)
    HIPCHECK(hipMemcpy(q.data(), h_sys_[k].qenergies, 8 * q.size(), hipMemcpyDeviceToHost));
  int w = 0;
  for (int st = 0; st < s.nstates; ++st) {
    const double* e = &q[(size_t)st * kQTerms];
    double qx_el = e[0] + e[2] + e[4], qx_vdw = e[1] + e[3] + e[5];
    double tot = e[6] + e[7] + e[8] + e[9] + qx_el + qx_vdw + e[10];
    double v[12] = {e[0], e[1], e[2], e[3], e[4], e[5], tot, e[6], e[7], e[8], e[9], e[10]};
    for (int c = 0; c < 12 && w < cap; ++c) out[w++] = v[c];
  }
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
  } else if (which == 3) {
    // nbpp as the force kernel sees it: the device rows and pair masks, read
    // back in reference groups (the slices of a split group joined)
    const qsys::Topology& t = s.top;
    const DevGroups& G = inst_[k]->groups;
    const int nd = G.nsolute;
    std::vector<int> cnt(nd), nb((size_t)nd * D.pp_cap);
    std::vector<ulonglong2> mk((size_t)nd * D.pp_cap);
    if (nd > 0) {
      HIPCHECK(hipMemcpy(cnt.data(), D.pp_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(nb.data(), D.pp_nb, 4 * nb.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(mk.data(), D.pp_mask, sizeof(ulonglong2) * mk.size(), hipMemcpyDeviceToHost));
    }
    // the mask of device pair (di, dj) from di's row
    auto mask_of = [&](int di, int dj, ulonglong2& m) {
      for (int k2 = 0; k2 < cnt[di]; ++k2)
        if (nb[(size_t)di * D.pp_cap + k2] == dj) { m = mk[(size_t)di * D.pp_cap + k2]; return true; }
      return false;
    };
    for (int g = 0; g < t.ncgp_solute; ++g) {
      const int d0 = G.dev_first[g];
      std::vector<int> others;   // reference groups of the first slice's row, in order
      for (int k2 = 0; k2 < cnt[d0]; ++k2) {
        const int o = G.ref[nb[(size_t)d0 * D.pp_cap + k2]];
        if (others.empty() || others.back() != o) others.push_back(o);
      }
      const auto& cg = t.cgp[g];
      for (int other : others) {
        if (!(other == g || ((g < other) == (((g + other) & 1) == 0)))) continue;
        const auto& co = t.cgp[other];
        for (int a = cg.first; a <= cg.last; ++a)
          for (int b = co.first; b <= co.last; ++b) {
            const int i = t.cgpatom[a], j = t.cgpatom[b];
            if (other == g && i >= j) continue;
            const int ra = a - cg.first, rb = b - co.first;
            ulonglong2 m;
            if (!mask_of(d0 + ra / kMaxGroupAtoms, G.dev_first[other] + rb / kMaxGroupAtoms, m)) continue;
            const int bit = (ra % kMaxGroupAtoms) * kMaxGroupAtoms + rb % kMaxGroupAtoms;
            if ((m.x >> bit) & 1ull) continue;
            out.push_back(i); out.push_back(j);
            out.push_back(((m.y >> bit) & 1ull) ? 3 : s.ljcod[s.iac[i] * s.natyps + s.iac[j]]);
          }
      }
    }
  } else if (which == 0 || which == 1) {
    const qsys::Topology& t = s.top;
    const DevGroups& G = inst_[k]->groups;
    const int nd = G.nsolute, nc = t.ncgp_solute;
    const int cap = which == 0 ? D.pp_cap : D.pw_cap;
    std::vector<int> cnt(nd), nb((size_t)nd * cap);
    if (nd > 0) {
      HIPCHECK(hipMemcpy(cnt.data(), which == 0 ? D.pp_cnt : D.pw_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(nb.data(), which == 0 ? D.pp_nb : D.pw_nb, 4 * nb.size(), hipMemcpyDeviceToHost));
    }
    // the row of each reference group: its first slice's row, with the
    // slices of split solute groups joined (every slice of a group shares
    // the group's switch atom, so all of them stand in the same rows)
    std::vector<std::vector<int>> row(nc);
    for (int g = 0; g < nc; ++g) {
      const int d0 = G.dev_first[g];
      for (int k2 = 0; k2 < cnt[d0]; ++k2) {
        int o = nb[(size_t)d0 * cap + k2];
        if (which == 0) o = G.ref[o];
        if (which == 1 || row[g].empty() || row[g].back() != o) row[g].push_back(o);
      }
    }
    auto lj_of = [&](int i, int j) { return s.ljcod[s.iac[i] * s.natyps + s.iac[j]]; };
    auto zero = [&](int i, int j, int lj) {
      const auto& ti = t.types[s.iac[i]];
      const auto& tj = t.types[s.iac[j]];
      return s.crg[i] * s.crg[j] == 0.0 && ti.avdw[lj - 1] * tj.avdw[lj - 1] == 0.0 &&
             ti.bvdw[lj - 1] * tj.bvdw[lj - 1] == 0.0;
    };
    auto in_bits = [&](const std::vector<uint32_t>& rows, const std::vector<std::pair<int, int>>& lng,
                       int i, int j) {
      if (std::abs(j - i) <= qsys::kMaxNbrRange) {
        int lo = std::min(i, j), hi = std::max(i, j);
        return ((rows[lo] >> (hi - lo - 1)) & 1u) != 0;
      }
      for (auto& p : lng)
        if ((p.first == i && p.second == j) || (p.first == j && p.second == i)) return true;
      return false;
    };
    for (int g = 0; g < nc; ++g)
      for (int other : row[g]) {
        if (which == 0) {
          bool own = other == g || ((g < other) == (((g + other) & 1) == 0));
          if (!own) continue;
          for (int a = t.cgp[g].first; a <= t.cgp[g].last; ++a) {
            int i = t.cgpatom[a];
            if (s.iqatom[i]) continue;
            for (int b = t.cgp[other].first; b <= t.cgp[other].last; ++b) {
              int j = t.cgpatom[b];
              if (other == g && i >= j) continue;
              if (s.iqatom[j]) continue;
              int lj = lj_of(i, j);
              if (zero(i, j, lj)) continue;
              if (in_bits(t.listex, t.listexlong, i, j)) continue;
              if (in_bits(t.list14, t.list14long, i, j)) lj = 3;
              out.push_back(i); out.push_back(j); out.push_back(lj);
            }
          }
        } else {
          int jo = s.nat_solute + 3 * other;
          for (int a = t.cgp[g].first; a <= t.cgp[g].last; ++a) {
            int i = t.cgpatom[a];
            if (s.iqatom[i]) continue;
            for (int ip = 0; ip < 3; ++ip) {
              int lj = lj_of(i, jo + ip);
              if (zero(i, jo + ip, lj)) continue;
              out.push_back(i); out.push_back(jo + ip); out.push_back(lj);
            }
          }
        }
      }
  } else {
    throw std::runtime_error("unknown list");
  }
}

void Batch::list_counts(int k, long* counts) {
  std::vector<int> v;
  export_list(k, 0, v);
  counts[0] = (long)(v.size() / 3);
  export_list(k, 1, v);
  counts[1] = (long)(v.size() / 3);
  const DevSys& D = h_sys_[k];
  std::vector<int> cnt(D.nwat), nb((size_t)D.nwat * D.ww_cap);
  if (D.nwat > 0) {
    HIPCHECK(hipMemcpy(cnt.data(), D.ww_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nb.data(), D.ww_nb, 4 * nb.size(), hipMemcpyDeviceToHost));
  }
  long ww = 0;
  for (int iw = 0; iw < D.nwat; ++iw)
    for (int c = 0; c < cnt[iw]; ++c) {
      int jw = nb[(size_t)iw * D.ww_cap + c];
      if ((iw < jw) == (((iw + jw) & 1) == 0)) ++ww;
    }
  counts[2] = 9 * ww;
  int qc[2] = {0, 0};
  HIPCHECK(hipMemcpy(qc, D.q_counts, sizeof(qc), hipMemcpyDeviceToHost));
  counts[3] = (long)qc[0] * D.nqat;
  counts[4] = 3L * qc[1] * D.nqat;
}

// Work of one list update of instance k at its current coordinates,
// counted on the host with the list kernels' tests: [0] atom-to-centre
// multipole evaluations (lrf_update) of the water-side list kernel,
// [1] of the solute-side one, [2] water-water list entries (both
// directions), [3] water-solute group entries, [4] solute group pairs
// within the cut-off (both directions), [5] solute-water entries.
void Batch::work_counts(int k, double* out) {
  const DevSys& D = h_sys_[k];
  const qsys::Setup& s = *setups_[k];
  const qsys::Topology& t = s.top;
  std::vector<double> x(3 * (size_t)D.natom);
  get_state(k, x.data(), nullptr);
  for (int c = 0; c < 6; ++c) out[c] = 0.0;
  auto d2 = [&](int a, int b) {
    const double p = x[3 * a] - x[3 * b], q = x[3 * a + 1] - x[3 * b + 1], r = x[3 * a + 2] - x[3 * b + 2];
    return p * p + q * q + r * r;
  };
  std::vector<int> nonq(t.cgp.size(), 0);
  for (size_t g = 0; g < t.cgp.size(); ++g)
    for (int a = t.cgp[g].first; a <= t.cgp[g].last; ++a) nonq[g] += s.iqatom[t.cgpatom[a]] ? 0 : 1;
  const bool lrf = D.use_lrf;
  const int ns = s.nat_solute;
  for (int iw = 0; iw < s.nwat; ++iw) {
    const int io = ns + 3 * iw;
    if (s.excl[io]) continue;
    for (int jw = 0; jw < s.nwat; ++jw) {
      const int jo = ns + 3 * jw;
      if (jw == iw || s.excl[jo]) continue;
      const double r2 = d2(io, jo);
      if (r2 <= D.rcww2) out[2] += 1;
      else if (lrf && r2 <= D.rclrf2) out[0] += nonq[s.iwhich_cgp[jo]];
    }
    for (int g = 0; g < t.ncgp_solute; ++g) {
      const int sw = t.cgp[g].iswitch;
      if (s.excl[sw]) continue;
      const double r2 = d2(sw, io);
      if (r2 <= D.rcpw2) out[3] += 1;
      else if (lrf && r2 <= D.rclrf2) out[0] += nonq[g];
    }
  }
  for (int g = 0; g < t.ncgp_solute; ++g) {
    const int sw = t.cgp[g].iswitch;
    if (s.excl[sw]) continue;
    for (int jg = 0; jg < t.ncgp_solute; ++jg) {
      const int sj = t.cgp[jg].iswitch;
      if (jg == g || s.excl[sj]) continue;
      const double r2 = d2(sw, sj);
      if (r2 <= D.rcpp2) out[4] += 1;
      else if (lrf && r2 <= D.rclrf2) out[1] += nonq[jg];
    }
    for (int jw = 0; jw < s.nwat; ++jw) {
      const int jo = ns + 3 * jw;
      if (s.excl[jo]) continue;
      const double r2 = d2(sw, jo);
      if (r2 <= D.rcpw2) out[5] += 1;
      else if (lrf && r2 <= D.rclrf2) out[1] += nonq[s.iwhich_cgp[jo]];
    }
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

int qgpu_system_sizes(qgpu_handle* h, int k, int* sizes, int cap) {
  return guarded([&] {
    CHECK_K(h, k);
    const qsys::Setup& s = h->b->setup(k);
    const int v[7] = {s.natom, s.nat_solute, s.nwat, (int)s.top.cgp.size(), s.top.ncgp_solute, s.nqat, s.nstates};
    for (int i = 0; i < 7 && i < cap; ++i) sizes[i] = v[i];
  });
}

int qgpu_nstates(qgpu_handle* h, int k, int* nstates) {
  return guarded([&] { CHECK_K(h, k); *nstates = h->b->setup(k).nstates; });
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

int qgpu_pot_energy_terms(qgpu_handle* h, int mask) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    h->b->forces_terms(mask);
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
  return guarded([&] { CHECK_K(h, k); h->b->qenergies(k, e, cap); });
}
int qgpu_profile_kernels(qgpu_handle* h, int nsteps, double* ms, int cap) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    h->b->profile(nsteps, ms, cap);
  });
}
int qgpu_set_energy_interval(qgpu_handle* h, int iene, int max_frames) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    h->b->set_energy_interval(iene, max_frames);
  });
}
int qgpu_energy_series(qgpu_handle* h, int k, double* out, int cap, int* nframes) {
  return guarded([&] {
    CHECK_K(h, k);
    *nframes = h->b->energy_series(k, out, cap);
  });
}
int qgpu_energy_series_fp64(qgpu_handle* h, int k, double* out, int cap, int* nframes) {
  return guarded([&] {
    CHECK_K(h, k);
    *nframes = h->b->energy_series(k, out, cap, true);
  });
}
int qgpu_noffdiag(qgpu_handle* h, int k, int* n) {
  return guarded([&] { CHECK_K(h, k); *n = h->b->noffdiag(k); });
}
int qgpu_offdiag_series(qgpu_handle* h, int k, double* out, int cap, int* nframes) {
  return guarded([&] {
    CHECK_K(h, k);
    *nframes = h->b->offdiag_series(k, out, cap);
  });
}
int qgpu_far_samples(qgpu_handle* h, int k, float* X, float* Y, int cap, int* n) {
  return guarded([&] { CHECK_K(h, k); *n = h->b->far_samples(k, X, Y, cap, false); });
}
int qgpu_far_lrf_estimate(qgpu_handle* h, int k, float* X, float* Y, int cap, int* n) {
  return guarded([&] { CHECK_K(h, k); *n = h->b->far_samples(k, X, Y, cap, true); });
}
int qgpu_set_far_model(qgpu_handle* h, qgpu_mlp* m) {
  return guarded([&] { h->b->set_far_model(m ? m->m : nullptr); });
}
int qgpu_list_counts(qgpu_handle* h, int k, long* counts) {
  return guarded([&] {
    CHECK_K(h, k);
    h->b->list_counts(k, counts);
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
int qgpu_shake(qgpu_handle* h, int k, const double* xx, double* x, int* niter) {
  return guarded([&] {
    CHECK_K(h, k);
    if (!xx || !x) throw std::runtime_error("null coordinate array");
    const int nmol = h->b->setup(k).nmol;
    const int total = h->b->shake(k, xx, x);
    if (total < 0) throw std::runtime_error("shake failure");
    if (niter) *niter = nmol > 0 ? total / nmol : 0;
  });
}
int qgpu_pot_energy_nonbonds(qgpu_handle* h, int k, const double* x, double* d, double* E, double* EQ,
                             int eq_cap) {
  return guarded([&] {
    CHECK_K(h, k);
    if (!x || !d || !E) throw std::runtime_error("null array");
    const int ns = h->b->setup(k).nstates;
    if (ns > 0 && (!EQ || eq_cap < 4 * ns)) throw std::runtime_error("EQ needs 4 values per FEP state");
    h->b->pot_energy_nonbonds(k, x, d, E, EQ);
  });
}
int qgpu_last_step_ms(qgpu_handle* h, double* ms) {
  return guarded([&] {
    if (!h) throw std::runtime_error("null handle");
    *ms = h->last_ms;
  });
}
int qgpu_autotune(qgpu_handle* h, int reps, double* ms, int* shapes) {
  return guarded([&] {
    if (!h) throw std::runtime_error("invalid handle");
    h->b->autotune(reps > 0 ? reps : 3, ms);
    if (shapes)
      for (int k = 0; k < qgpu::Batch::TK_COUNT; ++k) {
        shapes[4 * k] = h->b->shape(k).map;
        shapes[4 * k + 1] = h->b->shape(k).block;
        shapes[4 * k + 2] = h->b->shape(k).pf;
        shapes[4 * k + 3] = h->b->shape(k).occ;
      }
  });
}

int qgpu_set_launch_shape(qgpu_handle* h, int kernel, int map, int block, int prefetch, int occupancy) {
  return guarded([&] {
    if (!h) throw std::runtime_error("invalid handle");
    h->b->set_shape(kernel, map, block, prefetch, occupancy);
  });
}

int qgpu_monitor(qgpu_handle* h, int k, double* out, int cap, int* n) {
  return guarded([&] {
    CHECK_K(h, k);
    *n = h->b->monitor(k, out, cap);
  });
}

int qgpu_work_counts(qgpu_handle* h, int k, double* counts, int cap) {
  return guarded([&] {
    CHECK_K(h, k);
    double c[6];
    h->b->work_counts(k, c);
    for (int i = 0; i < 6 && i < cap; ++i) counts[i] = c[i];
  });
}

int qgpu_instance_status(qgpu_handle* h, int k, int* bits, int* at_step) {
  return guarded([&] { CHECK_K(h, k); h->b->status(k, bits, at_step); });
}

int qgpu_list_capacity(qgpu_handle* h, int k, int* caps, int* regrowths) {
  return guarded([&] {
    CHECK_K(h, k);
    if (caps) h->b->list_caps(k, caps);
    if (regrowths) *regrowths = h->b->list_regrowths();
  });
}

int qgpu_set_list_capacity(qgpu_handle* h, int k, int ww, int wp, int pw) {
  return guarded([&] { CHECK_K(h, k); h->b->set_list_caps(k, ww, wp, pw); });
}

int qgpu_offdiag(qgpu_handle* h, int k, double* out, int cap, int* n) {
  return guarded([&] {
    CHECK_K(h, k);
    const qsys::Setup& s = h->b->setup(k);
    const auto& od = s.offd;
    *n = (int)od.size();
    if (!out || od.empty()) return;
    std::vector<double> x(3 * (size_t)s.natom);
    h->b->get_state(k, x.data(), nullptr);
    for (int m = 0; m < (int)od.size() && 6 * m + 5 < cap; ++m) {
      // offdiag (src/md.f90:15444-15470) at the current coordinates
      const int k3 = 3 * s.iqseq[od[m].k], l3 = 3 * s.iqseq[od[m].l];
      const double r = std::sqrt((x[l3] - x[k3]) * (x[l3] - x[k3]) + (x[l3 + 1] - x[k3 + 1]) * (x[l3 + 1] - x[k3 + 1]) +
                                 (x[l3 + 2] - x[k3 + 2]) * (x[l3 + 2] - x[k3 + 2]));
      out[6 * m] = od[m].i + 1; out[6 * m + 1] = od[m].j + 1;
      out[6 * m + 2] = od[m].k + 1; out[6 * m + 3] = od[m].l + 1;
      out[6 * m + 4] = od[m].A * std::exp(-od[m].mu * r);
      out[6 * m + 5] = r;
    }
  });
}

int qgpu_offdiag_states(qgpu_handle* h, int k, int* ij, int cap, int* n) {
  return guarded([&] {
    CHECK_K(h, k);
    const auto& od = h->b->setup(k).offd;
    *n = (int)od.size();
    for (int m = 0; m < (int)od.size() && 2 * m + 1 < cap; ++m) { ij[2 * m] = od[m].i + 1; ij[2 * m + 1] = od[m].j + 1; }
  });
}

}  // extern "C"

struct qgpu_trj {
  qsys::Setup setup;
  std::unique_ptr<qsys::DcdWriter> w;
};

extern "C" {

int qgpu_trj_create(const char* input_path, const char* trj_path, qgpu_trj** out) {
  return guarded([&] {
    if (!input_path || !out) throw std::runtime_error("null argument");
    auto t = std::make_unique<qgpu_trj>();
    t->setup = qsys::prepare(input_path);
    std::string path;
    if (trj_path) {
      path = trj_path;
    } else {
      const std::string ip(input_path), f = t->setup.inp.trj_file;
      if (f.empty()) throw std::runtime_error("the input names no trajectory file");
      const size_t sl = ip.find_last_of('/');
      path = (f[0] == '/' || sl == std::string::npos) ? f : ip.substr(0, sl) + "/" + f;
    }
    t->w = std::make_unique<qsys::DcdWriter>(t->setup, path);
    *out = t.release();
  });
}

int qgpu_trj_natoms(qgpu_trj* t) { return t && t->w ? t->w->natoms() : 0; }

int qgpu_trj_write(qgpu_trj* t, const double* x) {
  return guarded([&] {
    if (!t || !t->w || !x) throw std::runtime_error("invalid trajectory or null coordinates");
    t->w->write(x);
  });
}

int qgpu_trj_close(qgpu_trj* t) {
  return guarded([&] { delete t; });
}

int qgpu_write_restart(const char* path, int natom, const double* x, const double* v, int nshell,
                       const double* theta_corr) {
  return guarded([&] {
    if (!path || !x || !v || natom <= 0 || nshell < 0) throw std::runtime_error("invalid restart arguments");
    std::vector<double> xv(x, x + 3 * (size_t)natom), vv(v, v + 3 * (size_t)natom);
    std::vector<double> tc;
    if (nshell > 0) tc.assign(theta_corr, theta_corr + nshell);
    qsys::write_restart(path, xv, vv, nshell > 0 ? &tc : nullptr);
  });
}

int qgpu_write_restart_box(const char* path, int natom, const double* x, const double* v, const double* box) {
  return guarded([&] {
    if (!path || !x || !v || !box || natom <= 0) throw std::runtime_error("invalid restart arguments");
    std::vector<double> xv(x, x + 3 * (size_t)natom), vv(v, v + 3 * (size_t)natom);
    qsys::write_restart(path, xv, vv, nullptr, box);
  });
}

int qgpu_box(qgpu_handle* h, int k, double* out, int* is_box) {
  return guarded([&] {
    CHECK_K(h, k);
    if (!out || !is_box) throw std::runtime_error("null output");
    *is_box = h->b->box(k, out) ? 1 : 0;
  });
}

int qgpu_volume_moves(qgpu_handle* h, int k, int* out, int cap, int* tries) {
  return guarded([&] {
    CHECK_K(h, k);
    if (!tries || (cap > 0 && !out)) throw std::runtime_error("null output");
    *tries = h->b->volume_moves(k, out, cap);
  });
}

const char* qgpu_last_error(void) { return g_err.c_str(); }

}  // extern "C"
