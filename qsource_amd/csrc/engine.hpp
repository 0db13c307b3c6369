// Device-side layout of one simulation instance (one system at one set of
// lambdas) and the host engine that drives a batch of them on one GPU.
//
// HBM layout per instance (all FP64 unless noted):
//   x, v, xx, d      3N AoS, same element order as qdyn's x(:)/v(:)/d(:)
//   xtop             3N, topology coordinates (restraint reference)
//   crg, winv, mass  N (crg already scaled by sqrt(coulomb_constant))
//   flags            N bytes: bit0 excluded, bit1 restrained shell,
//                    bit2 Q atom, bit3 heavy
//   lrf              ncgp x 44: centre(3) phi0 phi1(3) phi2(9) phi3(27) pad
//   ww_nb            nwat x ww_cap int32 neighbour waters (full, ascending)
//   wp_nb            nwat x wp_cap int32 solute charge groups within Rcpw
//   pp_nb            ncgp_solute x pp_cap int32 solute charge groups within Rcpp
//   pw_nb            ncgp_solute x pw_cap int32 waters within Rcpw
// Lists are "full" (each owner sees all its partners) so that every atom's
// gradient is produced by exactly one wavefront without atomics; the
// reference's half lists (charge-group parity ownership, src/md.f90:
// 10977-10980) are recovered by the ownership predicate and are what the
// energies sum over and what the list export returns.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "host/qsys.hpp"

namespace qgpu {

// F_QP: solute atom in the static Q-solute list set (nbqplist);
// F_QW: water (flag on its O) in the static Q-water list set (nbqwlist)
enum : unsigned char { F_EXCL = 1, F_SHELL = 2, F_Q = 4, F_HEAVY = 8, F_QP = 16, F_QW = 32 };

constexpr int kMaxGroupAtoms = 8;   // atoms per device solute group (larger groups are split)
constexpr int kMaxQStates = 10;
// wavefronts per Q atom in the Q-atom kernel, and the per-wavefront partial
// record: gradient (3) and per-state qp_el qp_vdw qw_el qw_vdw
constexpr int kQChunks = 8;
constexpr int kQPartStride = 3 + 4 * kMaxQStates;

constexpr int kLrfStride = 44;
constexpr int kNumShellMax = 8;

// per-instance energy slots reduced on demand (index into DevSys::eacc)
enum EnergyTerm {
  E_PP_EL, E_PP_VDW, E_PW_EL, E_PW_VDW, E_WW_EL, E_WW_VDW, E_LRF,
  E_P_BOND, E_P_ANGLE, E_P_TORSION, E_P_IMPROPER, E_W_BOND, E_W_ANGLE, E_W_TORSION, E_W_IMPROPER,
  E_RST_FIX, E_RST_SHELL, E_RST_RADIAL, E_RST_WPOL, E_RST_PROTEIN,
  E_KINETIC,
  E_NTERMS
};
// per-state Q terms: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw, Q-bonded
// bond angle torsion improper, restraint
constexpr int kQTerms = 11;

// extra FP64 terms of the term kernel: Q-bonded terms per state
// (src/md.f90:17022-17540) and the restraints of p_restrain (:15473-15802).
// Record of kXtInts ints: kind, atoms a0..a3, state, aux0, aux1, aux2,
// first slot; kXtPars doubles of parameters.
enum XtKind {
  XT_QBOND = 1, XT_QANGLE, XT_QTORSION, XT_QIMPROPER,
  XT_RST_SEQ, XT_RST_SEQC, XT_RST_POS, XT_RST_DIST, XT_RST_ANG, XT_RST_WALL
};
constexpr int kXtInts = 10;
constexpr int kXtPars = 8;
constexpr int kMorseRec = 7;   // Morse factor and its gradient on the bond's two atoms
constexpr int kQwRec = 8;      // Q-water record per Q atom (qw_rec_d / qw_rec_f)
constexpr int kSrec = 12;   // floats per solute slot record (x y z q, a0 a1 a2 b0, b1 b2 skip pad)
constexpr int kEWater = 5;     // energy partials per water
constexpr int kMon = 3;        // Vwsum, Vwel, Vwlj before the per-state values
constexpr int kIntegrateBlock = 64;  // threads per k_integrate workgroup
constexpr int kWpolBlock = 256;      // threads per k_watpol_force workgroup
constexpr int kWpolSplit = 4;        // k_watpol_force workgroups per water shell
// lanes per water in k_water_lists: FP64 build, FP32 (mixed) build
constexpr int kWLLanes64 = 8, kWLLanes32 = 16;
// lanes per solute atom in k_solute_forces
constexpr int kSFLanes = 16;

struct DevShell {
  double rout, dr, cstb, avtheta, avn_insh, theta_corr;
  int n_insh;
  int pad;
};

struct DevSys {
  int natom, nat_solute, nwat, ncgp, ncgp_solute;   // ncgp*: device charge groups
  int nqat, nstates;
  int precision;                 // 0 = FP64, 1 = mixed FP32 pair / FP64 accumulation
  int term_mask;                 // TM_* terms a force evaluation includes (all but in term diagnostics)
  double *x, *v, *xx, *d, *xtop;
  float* xf;                     // FP32 copy of x for the mixed-precision pair kernels
  // mixed precision: FP32 coordinates of each water (O, H1, H2) as one
  // 64-byte record of 16 floats (9 used), so that a partner water is read
  // with three 16-byte loads of one cache line
  float* xw;
  // monitored group pairs (nonbond_monitor): atoms of group i and j of pair
  // p at mon_at[mon_fi[p] ..) and [mon_fj[p] ..), counts mon_ni/mon_nj, LJ
  // codes [(s*ni + i)*nj + j] from mon_lj[mon_lo[p]]; results kMon + 2 nstates
  // per pair into mon_out
  const int* iqatom_d;           // N: Q index + 1 of each atom (0 = not a Q atom)
  int mon_n;
  const int *mon_at, *mon_fi, *mon_fj, *mon_ni, *mon_nj, *mon_lo;
  const signed char* mon_lj;
  double* mon_out;
  const int* slot_of_atom;       // N: the atom's index in slot_atom, or -1
  double* xq;                    // coordinates of the terms after fix_shell: x, or with
                                 // excluded_freeze a copy holding the excluded atoms at xtop
  const double *crg, *winv, *mass;
  const unsigned char* flags;
  const int *iac;                // type index per atom
  // device charge groups (solute groups of more than kMaxGroupAtoms atoms
  // are split into slices): the group's atoms cgpatom[cgp_first..cgp_last],
  // switch atom, and the range whose mean is the centre (the whole group)
  const int *cgp_first, *cgp_last, *cgp_sw, *cgpatom, *iwhich_cgp;
  const int *cgp_cfirst, *cgp_clast;
  double* lrf;
  // type tables (natyps x 3 each), ljcod natyps x natyps
  int natyps;
  const double *avdw, *bvdw;
  const float *avdw_f, *bvdw_f, *crg_f;   // FP32 copies for the mixed-precision solute pairs
  const int* ljcod;
  // solute exclusion / 1-4 bit rows (pairs at most kMaxNbrRange apart) and,
  // for pairs further apart, a per-atom CSR over the lower atom of each pair:
  // exl_code[exl_first[lo] .. exl_first[lo+1]) = hi << 2 | 1 (excluded) | 2 (1-4)
  const unsigned* listex; const unsigned* list14;
  const int *exl_first, *exl_code;
  double el14_scale;
  const int* cgp_slot;           // solute atom -> its position within its charge group
  // lists
  int ww_cap, wp_cap, pp_cap, pw_cap;
  int *ww_cnt, *ww_nb, *wp_cnt, *wp_nb, *pp_cnt, *pp_nb, *pw_cnt, *pw_nb;
  // per pp entry (group g, neighbour group jg): bit a*8+b of .x set when the
  // atom pair (slot a of g, slot b of jg) is not evaluated (same atom, Q atom,
  // zero interaction or excluded), of .y when it is a 1-4 pair
  ulonglong2* pp_mask;
  // failure record of this instance (kStatusInts ints): [0] ST_* bits,
  // [1] step + 1 of the first failure, [2..4] longest ww, wp and pw row seen
  // at the failing list build (the capacity the rows need)
  int* status;
  // batch-wide failure key (step * 16 + launch phase of the first failure,
  // ULLONG_MAX while none): a state-changing kernel of phase p at step s
  // returns at once when *fail < s * 16 + p, so every instance stops at the
  // same point of the failing step; 0 halts everything
  unsigned long long* fail;
  // cut-offs squared
  double rcpp2, rcpw2, rcww2, rclrf2;
  bool use_lrf;
  // solvent
  double chg_solv[3];
  int std_water;                 // every water is one charge group (O, H1, H2) with chg_solv, no Q atom
  // with standard waters: per solute atom i the LJ products with water atoms
  // O, H1, H2 (aLJ[3] then bLJ[3]) in FP64 and FP32, and a 3-bit mask of
  // the water atoms it has a zero interaction with (nbpwlist's skip rule)
  const double* swp_d;
  const float* swp_f;
  const unsigned char* swp_zero;
  // mixed precision with standard waters: one kSrec-float record per solute
  // charge-group slot a (the atom cgpatom[a]), read by the water side of the
  // solute-water pairs in slot order: FP32 x, y, z (refreshed with xf), the
  // charge, the LJ products with water atoms O, H1, H2 (a then b) and 1.0 in
  // the skip field for a Q atom; srec_of_atom maps a solute atom to its slot
  float* srec;
  const int* srec_of_atom;
  double aOO, bOO;               // alj_solv(1,iLJ)^2, blj_solv(1,iLJ)^2 for iLJ=ljcod(O,O)
  // all-atom solvent LJ (topology solvent type 1): LJ between every atom
  // pair of two waters, A of pair (la, ka) = aLJ_solv(la,iLJ) aLJ_solv(ka,iLJ)
  // with iLJ = ljcod(atom la, water O) (src/md.f90:15291-15300)
  int solv_allatom;
  double aww[9], bww[9];
  double rwat, fk_wsphere, dwmz, awmz, fkwpol;
  double xwcent[3];
  int wpol_restr;
  int nshell;
  DevShell* shell;
  double* theta;                 // per water
  int* shell_of;                 // per water, -1 when outside the shells
  double* wpol_g;                // per water 9: this step's polarisation-restraint gradient
                                 // (valid for shell members, added to d by the integrator)
  // restraints
  double fk_fix, fk_pshell;
  // Q atoms
  const int* iqseq;
  const double* qcrg;            // nqat x nstates
  const double* lambda;
  const signed char* qp_lj;      // nat_solute x nqat LJ code of the (Q, solute atom) pair
  const signed char* qw_lj;      // nqat x 3: LJ code of the (water atom ip, Q atom) pair
  // per Q atom, for the water side of the Q-water forces with fixed Q LJ
  // parameters: sum_s lambda_s qcrg(iq,s), sum_s lambda_s, then A and B of
  // the Q atom's pair with water atoms O, H1, H2 (kQwRec values)
  const double* qw_rec_d;
  const float* qw_rec_f;
  double alj_w[3][2], blj_w[3][2]; // aLJ_solv(ip,:), bLJ_solv(ip,:) of the three water atoms
  int* q_counts;                 // [0] solute atoms in the Q-solute list, [1] waters in the Q-water list
  int q_dynamic;                 // Rcq <= rexcl_o: both lists rebuilt at every list update (k_q_lists)
  double rcq2;
  const unsigned char* qp_cand;  // nat_solute: atom may enter the Q-solute list
  unsigned char* qp_gin;         // ncgp_solute: group's switch atom within Rcq (k_q_lists)
  double xpcent[3];              // solute centre (topology)
  // periodic box (pbc != 0): edge lengths and their inverses (FP64 and FP32
  // copies), the Q atoms' switching atom, per solute device group the atom
  // whose image gives the group's Q-atom shift (-1: group not in the Q-solute
  // list; k_q_lists), the box centre (moves with put_back_in_box), and the
  // molecule table put_back_in_box works on
  int pbc;
  double boxl[3], inv_boxl[3];
  float boxl_f[3], inv_boxl_f[3];
  int qswitch;
  int* qp_shatom;
  double* boxc;
  int rigid_boxc, put_solute_back, put_solvent_back;
  int nmol, nsolute_mol;
  const int* mol_start;          // nmol + 1 (0-based first atom of each molecule, then natom)
  const double* mol_minv;        // 1 / mass of each molecule
  // constant pressure: a volume move scales this instance's molecules
  // (mc_move != 0) from the box lengths boxl_old to boxl (k_mc_scale)
  int mc_move;
  double boxl_old[3];
  int* qp_atoms;                 // solute atoms in the Q-solute list, ascending
  int* qw_waters;                // waters in the Q-water list, ascending
  // Q-atom LJ types (qvdw_flag): qiac [iq*nstates+s]; qavdw/qbvdw [type*3 +
  // LJ code-1]; soft-core sc_lookup [(iq*sc_ncol + column)*nstates + s] with
  // columns natyps topology types then nqat Q atoms (all 0 without softcore)
  int qvdw;
  int qq_lib;                    // nonbond_qq_lib_charges: library charges for Q-Q pairs
  const int* qiac;
  const double *qavdw, *qbvdw;
  const double* sc_lookup;
  int sc_ncol;
  double* e_q;                   // nqat x nstates x 4: qp_el qp_vdw qw_el qw_vdw
  double* q_part;                // nqat x kQChunks x kQPartStride partial sums
  double *q_part64, *e_q64;      // mixed precision: FP64 re-evaluation on energy-file steps
  // Q-Q and Q-neighbour pairs of all states (nbqq), and per-atom incidence
  int nqq;                       // total pairs over states
  const int *qq_state, *qq_iq, *qq_j, *qq_jq, *qq_lj;
  const double* qq_elscale;
  double* e_qq;                  // nqq x 2: el, vdw (of the pair)
  // bonded terms (after shrink_topology) and per-atom incidence
  const int* bnd;                // nbnd x 3 (i, j, cod)
  const int* ang;                // nang x 4
  const int* tor;                // ntor x 5
  const double* bondlib;         // fk, b0
  const double* anglib;          // fk, ang0
  const double* torlib;          // fk, rmult, deltor, paths
  const int* imp;                // nimp x 5 (i, j, k, l, cod)
  const double* implib;          // fk, imp0 (radians)
  int ff_type;                   // 2: periodic impropers (AMBER), else harmonic
  int nbnd, nbnd_solute, nang, nang_solute, ntor, ntor_solute, nimp, nimp_solute;
  double* e_bterm;               // energy per bonded term (bonds, angles, torsions, impropers)
  // gradient slots of the term kernel: 3 FP64 per (term, atom of the term),
  // numbered bonds 2, angles 3, torsions 4, impropers 4, Q pairs 2 per term; slot n is
  // stored at slot_pos[n], so that atom slot_atom[t]'s contributions occupy
  // slot_g[slot_first[t] .. slot_first[t+1]) in slot order
  double* slot_g;
  int nslot_atoms;
  const int *slot_atom, *slot_first, *slot_pos;
  // extra terms (XtKind records) with per-item energies.  Q bonds carry
  // their Morse factor (EMorseD, dMorse_i/j, src/qatom.f90:201-203) to the
  // coupled Q angles, torsions and impropers: a coupled term recomputes the
  // factor at the latest state <= its own where the bond is present; with
  // none it takes the value the bond's last present state left in the
  // previous evaluation (qb_prev, kept by the reference's module arrays)
  int nxt;
  const int* xt;
  const double* xp;
  double* e_xt;
  int nqbond;
  const int* qb_atoms;           // nqbond x 2
  double *qb_cur, *qb_prev;      // nqbond x kMorseRec
  const int* rst_atoms;          // atoms of the centre restraints
  int freeze;
  // integrator
  double dt;
  double temp0, tau_t;
  // thermostat (0 Berendsen, 1 Langevin, 2 Nose-Hoover) and velocity-Verlet
  // (the velocity update applied again after the position update); dv_mod
  // the velocity update's step
  int thermostat, vverlet;
  double dv_mod;
  // Langevin: friction factor, per-atom random force scale, Gaussian (1)
  // or uniform (0) forces; randm's state (rng[0], advanced once per step by
  // the step's draws: rng_step = {A, C}, state' = (A state + C) mod 1e8)
  // and per atom the jump to its first draw (rng_jump[2i], rng_jump[2i+1])
  double dv_friction;
  const double* rands;
  int lgauss;
  int* rng;
  const int* rng_jump;
  long long rng_step[2];
  // Nose-Hoover chain: nhn links with masses qnh; nh holds xnh, vnh, Gnh
  // (nhn each) and the step's two velocity scales (before the update, and
  // after SHAKE), formed by k_nh_prop before k_integrate
  int nhn;
  const double* qnh;
  double* nh;
  double kbt;
  int separate_scaling, detail_temps;
  int ndegf, ndegfree, ndegf_solute, ndegf_solvent, ndegfree_solute, ndegfree_solvent;
  // SHAKE units: every atom belongs to exactly one unit (a SHAKE molecule or
  // a single unconstrained atom); unit u spans atoms unit_atoms[unit_first[u]
  // .. unit_first[u+1]) and constraints sh_*[con_first[u] .. con_first[u+1])
  int nunits;
  int has_shake;                 // any SHAKE constraint: v = (x - xx)/dt for every atom
  const int *unit_first, *unit_atoms, *con_first, *sh_i, *sh_j;
  // 1 for a water unit (atoms a, a+1, a+2 with constraints (a,a+1),
  // (a,a+2), (a+1,a+2) in that order): integrated entirely in registers
  const signed char* unit_kind;
  const int* unit_head;          // first atom of each unit
  const double* sh_d2;
  int shake_maxiter;             // SHAKE_MAX_ITER (1000); QGPU_SHAKE_MAXITER lowers it to force a failure
  // energies: eacc[term] partial sums written per item, reduced on demand
  double* e_water;     // nwat x kEWater : ww_el, ww_vdw, lrf(water atoms), radial, excluded-atom restraint
  double* e_solute;    // nat_solute x 8 : pp_el pp_vdw pw_el pw_vdw lrf fix shell spare
  double* e_wpol;      // nshell
  double* wpol_part;   // nshell x kWpolSplit x 3: energy, theta sum, member count of each k_watpol_force workgroup
  double* ke_part;     // kinetic partial sums: nblocks x 6
  int ke_blocks;
  double* scal;        // [0]Tscale_solute [1]Tscale_solvent [2]Tfree [3]Temp [4]Ekin [5..] per-class temps
  double* energies;    // E_NTERMS reduced values
  double* qenergies;   // nstates x kQTerms reduced values
  // step counter shared by the batch (incremented once per step on the
  // device, so a captured multi-step graph needs no host-side step number)
  int* step_ctr;
  // energy-file frames recorded on the device every iene steps (istep > 0,
  // as md_run calls put_ene, src/md.f90:6628-6630): per frame and state the
  // 14 values of a put_ene record (src/nrgy.f90:128-130)
  int iene;
  int eseries_base;    // step counter when recording was switched on
  int eseries_cap;
  double* eseries;
  double* eseries64;   // mixed precision: the same frames with FP64 Q pair arithmetic
  // EVB off-diagonal functions (Q atoms k, l; A, mu) and, per recorded
  // frame, Hij and r_kl of each (offdiag, src/md.f90:15444-15470)
  int noffd;
  const int* offd_kl;
  const double* offd_amu;
  double* eseries_offd;
  // far field: far_mode 1 replaces the LRF multipoles and Taylor terms by the
  // learned surrogate; per group features far_x (4 floats), outputs far_y (3
  // floats), frame far_frame (e_r, t, centre: 9 doubles) at offset far_off
  // of the batch-wide arrays; e_far per group energy
  int far_mode;
  int far_off;
  float *far_x, *far_y;
  double* far_frame;
  double* e_far;
  int* eseries_n;
};

constexpr int kEneRecord = 14;

// Failure bits of DevSys::status and the launch phases of the failure key.
enum : int { ST_LIST_OVERFLOW = 1, ST_SHELL_OVERFLOW = 2, ST_SHAKE = 4, ST_NAN = 8 };
constexpr int kStatusInts = 8;
// Term groups of a force evaluation (DevSys::term_mask): water-water,
// solute-water, solute-solute, Q-atom nonbonded (Q-solute and Q-water),
// LRF Taylor, and the rest (bonded, Q-Q, Q-bonded, restraints, polarisation)
enum : int { TM_WW = 1, TM_PW = 2, TM_PP = 4, TM_QX = 8, TM_LRF = 16, TM_OTHER = 32, TM_ALL = 63 };
enum : int { PH_LISTS = 1, PH_TERMS = 2, PH_WATPOL = 3, PH_RECORD = 4, PH_INTEGRATE = 5, PH_TEMPERATURE = 7 };
constexpr unsigned long long kNoFailure = ~0ull;

// One GPU batch of instances sharing every launch.
class Batch {
 public:
  Batch(const std::vector<std::string>& inputs, int precision, int device);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  int size() const { return (int)inst_.size(); }
  void step(int nsteps);
  void make_lists();
  void forces();
  void forces_terms(int mask);
  bool box(int k, double* out);
  void get_state(int k, double* x, double* v);
  void set_state(int k, const double* x, const double* v);
  void get_gradient(int k, double* d);
  void energies(int k, double* e, int cap);
  void qenergies(int k, double* e, int cap);
  void export_list(int k, int which, std::vector<int>& out);
  void list_counts(int k, long* counts);
  void work_counts(int k, double* out);
  double temperature(int k);
  void theta_corr(int k, double* out);
  // shake(xx, x) of instance k on host arrays; returns the summed
  // iterations of its SHAKE molecules, or -1 on a failure
  int shake(int k, const double* xx, double* x);
  // pot_energy_nonbonds of instance k at host coordinates x: d += gradient,
  // e7 = pp, pw, ww (el, vdW) and LRF, eq4 = per state qp and qw (el, vdW)
  void pot_energy_nonbonds(int k, const double* x, double* d, double* e7, double* eq4);
  int istep() const { return istep_; }
  // constant pressure: volume moves of instance k so far (tries, accepted)
  // and per try 0 (rejected), 1 (accepted, deltaW <= 0) or 2 (accepted by
  // the random draw); returns the number of tries
  int volume_moves(int k, int* out, int cap) const;
  // failure record of instance k: ST_* bits and the step they were found at
  void status(int k, int* bits, int* at_step);
  int list_regrowths() const { return list_regrowths_; }
  void set_list_caps(int k, int ww, int wp, int pw);
  void list_caps(int k, int* caps) const {
    caps[0] = h_sys_[k].ww_cap; caps[1] = h_sys_[k].wp_cap; caps[2] = h_sys_[k].pw_cap;
  }
  const qsys::Setup& setup(int k) const { return *setups_[k]; }
  hipStream_t stream() const { return stream_; }
  // timing helpers
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // energy-file frames recorded on the device every iene steps
  void set_energy_interval(int iene, int max_frames);
  int energy_series(int k, double* out, int cap, bool fp64 = false);
  int offdiag_series(int k, double* out, int cap);
  // far-field samples of instance k at its current coordinates: features and
  // the exact far-field targets (lists rebuilt), or the estimate of the
  // current LRF multipoles; returns the number of charge groups
  int far_samples(int k, float* X, float* Y, int cap, bool lrf_estimate);
  void set_far_model(class Mlp* m);
  int noffdiag(int k) const { return h_sys_[k].noffd; }
  void set_graphs(bool on) { use_graph_ = on; }
  // per-kernel average device time (ms per launch) over nsteps steps run
  // with an event pair around every launch; order of kKernelNames
  void profile(int nsteps, double* ms, int cap);
  static constexpr int kNumKernels = 13;

  // Launch shapes of the four large kernels (water and solute lists and
  // forces): workgroup mapping (0: 2-D grid with instance = blockIdx.y;
  // 1: instance-major work map without empty workgroups; 2: work map that
  // keeps each instance's workgroups on one XCD, instances balanced over the
  // XCDs), workgroup size and, for the force kernels, whether each lane
  // loads its next partner during the current one.  Every shape computes each item with the same
  // arithmetic in the same order, so results do not depend on the choice.
  enum TunedKernel { TK_WATER_LISTS, TK_SOLUTE_LISTS, TK_WATER_FORCES, TK_SOLUTE_FORCES, TK_COUNT };
  // pf: partner prefetch (force kernels); occ: 1 = water forces fitted to 4
  // instead of 6 wavefronts per SIMD (FP32)
  struct LaunchShape { int map = 0; int block = 256; int pf = 0; int occ = 0; };
  // times every shape of every tuned kernel on this batch (a few launches
  // each; the kernels' outputs do not change) and keeps the fastest; ms
  // receives, per kernel, the best and the default shape's time
  void autotune(int reps, double* ms = nullptr);
  const LaunchShape& shape(int k) const { return shape_[k]; }
  // fixes one kernel's launch shape (and switches the tuning at the first step off)
  void set_shape(int k, int map, int block, int pf, int occ);
  bool tuned() const { return tuned_; }

 private:
  void pb(int id) {
    if (prof_) (void)hipEventRecord(pev_[2 * id], stream_);
  }
  void pe(int id) {
    if (prof_) { (void)hipEventRecord(pev_[2 * id + 1], stream_); pused_[id] = 1; }
  }
  bool prof_ = false;
  hipEvent_t pev_[2 * 13] = {};
  int pused_[13] = {};
  struct Inst;
  void init(const std::vector<std::string>& inputs, int device);
  void release();
  void fork();
  void join();
  void one_step(bool energies = true);
  void launch_steps(int nsteps, bool final_energies);
  void drop_graphs();
  bool check_failures();
  void grow_lists(int k, int need_ww, int need_wp, int need_pw);
  void launch_cycle_graph(bool final_energies);
  void sync_sys();
  void refresh_xf();
  int nbcycle_ = 10;
  bool allatom_ = false;   // all-atom solvent LJ (topology solvent type 1)
  bool q_dynamic_ = false; // some instance rebuilds its Q lists at list updates
  bool put_back_ = false;  // some periodic-box instance puts molecules back at list updates
  bool freeze_ = false;    // some instance has excluded_freeze on
  bool nose_hoover_ = false;   // some instance has a Nose-Hoover thermostat
  int iene_ = 0;
  long ebase_ = 0;         // step of the last set_energy_interval
  bool use_graph_ = true;
  // captured list cycle; the _final one forms the energy partials in its
  // last step (the energies a caller reads after qgpu_step)
  hipGraphExec_t graph_exec_ = nullptr, graph_exec_final_ = nullptr;
  // whether the force kernels of the launches being issued form the
  // nonbonded/LRF/restraint energy partials (dynamics needs only gradients)
  bool launch_energies_ = true;
  int* d_step_ = nullptr;              // per-instance step counters
  int* d_status_ = nullptr;            // per-instance failure records (kStatusInts each)
  unsigned long long* d_fail_ = nullptr;   // batch-wide failure key
  std::string dead_;                   // set when a failure stopped the batch for good
  void upload(int k);
  // every list and the LRF multipoles; q_waters = false leaves the
  // Q-water list as it is
  void launch_lists(bool q_waters = true);
  void launch_forces(bool step, long s, bool energies = true);
 public:
  int monitor(int k, double* out, int cap);   // monitored pair energies of instance k
  int max_mon_ = 0;
 private:
  bool record_possible(long s) const;
  void launch_tuned(int kernel, const LaunchShape& sh);
  void launch_put_back();
  const std::pair<int2*, int>& work_map(int kernel, int map, int block);
  LaunchShape shape_[TK_COUNT];
  bool tuned_ = false;
  bool autotune_on_ = true;       // tune at the first step (QGPU_AUTOTUNE=0 switches it off)
  std::vector<std::pair<std::array<int, 3>, std::pair<int2*, int>>> maps_;
  void launch_integrate(long s);
  // constant pressure (MC_volume): some periodic-box instance tries a
  // volume move every ivol_ steps at its list update; per instance the
  // move's saved coordinates, randm's argument and the move record
  bool npt_ = false;
  int ivol_ = 0;
  struct VolumeMoves {
    bool on = false;
    double* x_saved = nullptr;
    int seed = 0;
    std::vector<int> log;
    int tries = 0, accepted = 0;
  };
  std::vector<VolumeMoves> vol_;
  bool npt_energy_step(long s) const;
  void mc_volume();
  double new_potential(int k, const double* old_e);
  void ensure_far();
  void launch_far_model();
  class Mlp* far_model_ = nullptr;
  float *far_x_ = nullptr, *far_y_ = nullptr;
  double* far_frame_ = nullptr;
  int far_total_ = 0;
  void reduce_energies();
  std::vector<qsys::Setup*> setups_;
  std::vector<Inst*> inst_;
  DevSys* d_sys_ = nullptr;
  std::vector<DevSys> h_sys_;
  hipStream_t stream_ = nullptr;
  hipStream_t stream2_ = nullptr;   // side branch of the step (solute, Q, terms)
  hipStream_t ls_ = nullptr;        // stream the next launch goes to
  hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
  bool branches_ = true;            // QGPU_BRANCHES=0 runs every kernel on one stream
  int istep_ = 0;
  int list_regrowths_ = 0;
  int precision_ = 0;
  int max_nwat_ = 0, max_ncgp_ = 0, max_ncgp_solute_ = 0, max_units_ = 0, max_natom_ = 0;
  int max_nshell_ = 0;
  int max_nat_solute_ = 0, max_nstates_ = 0;
  int max_nqat_ = 0, max_terms_ = 0, max_slot_atoms_ = 0;
};

}  // namespace qgpu
