// Host-side description of one qdyn simulation: the topology, MD input and
// FEP data as read from qdyn's own files, and the prepared per-run state
// (scaled charges, SHAKE molecules, degrees of freedom, water shells, Q-atom
// lists) that qdyn's initialisation builds before md_run.
//
// Layout conventions: atom, charge-group and Q-atom indices are 0-based;
// coordinates are AoS FP64 [x0,y0,z0,x1,...] of length 3N exactly as qdyn's
// x(:) (src/md.f90:309-313); the gradient d = dV/dx (force = -d).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <array>
#include <vector>

namespace qsys {

constexpr int kMaxNbrRange = 25;   // src/topo.f90:36
constexpr int kMaxStates = 10;     // src/qatom.f90:39-41
constexpr double kBoltz = 0.001986;        // src/md.f90:54
constexpr double kRhoWat = 0.0335;         // src/md.f90:53
constexpr double kShakeTol = 0.0001;       // src/md.f90:460
constexpr int kShakeMaxIter = 1000;        // src/md.f90:461
constexpr double kWpolrLayer = 3.0001;     // src/md.f90:190
constexpr double kDrout = 0.5;             // src/md.f90:191
constexpr int kItdisUpdate = 100;          // src/md.f90:189
constexpr double kQrealEps = 1.0e-14;      // src/sizes.f90:48

struct AtomType {
  double mass;
  double avdw[3];
  double bvdw[3];
};

struct Bond { int i, j, cod; };
struct Angle { int i, j, k, cod; };
struct Torsion { int i, j, k, l, cod; };
struct BondLib { double fk, b0; };
struct AngLib { double fk, ang0, ureyfk, ureyr0; };
struct TorLib { double fk, rmult, deltor, paths; };
struct ImpLib { double fk, imp0; };
struct Cgp { int iswitch, first, last; };  // first..last index cgpatom

// Contents of a qdyn topology file (src/topo.f90:515-1074), 0-based.
struct Topology {
  double version = 0;
  int ff_type = 1;
  int nat_pro = 0, nat_solute = 0, solv_atom = 3;
  std::vector<double> xtop;
  std::vector<int> iac;  // 0-based type index
  std::vector<Bond> bnd; int nbonds_solute = 0;
  std::vector<BondLib> bondlib;
  std::vector<Angle> ang; int nangles_solute = 0;
  std::vector<AngLib> anglib;
  std::vector<Torsion> tor; int ntors_solute = 0;
  std::vector<TorLib> torlib;
  std::vector<Torsion> imp; int nimps_solute = 0;
  std::vector<ImpLib> implib; int imp_type = 1;
  std::vector<double> crg;
  std::vector<Cgp> cgp; int ncgp_solute = 0; int iuse_switch_atom = 1;
  std::vector<int> cgpatom;
  std::vector<AtomType> types;
  int ivdw_rule = 1;
  double el14_scale = 0.5, coulomb_constant = 332.0;
  std::vector<std::pair<int,int>> lj2;
  // list14/listex: bit k-1 of row i set <=> (i, i+k) listed; rows nat_solute
  std::vector<uint32_t> list14, listex;
  std::vector<std::pair<int,int>> list14long, listexlong;
  std::vector<int> res_start; int nres_solute = 0;
  std::vector<int> istart_mol;  // 0-based start atom of each molecule
  int solvent_type = 0;
  bool use_pbc = false;
  double boxlength[3] = {0, 0, 0}, boxcentre[3] = {0, 0, 0};   // periodic box (topology)
  double rexcl_o = 0, rexcl_i = 0, rwat = 0;
  double xpcent[3] = {0,0,0}, xwcent[3] = {0,0,0};
  int nexats = 0, nexwat = 0;
  std::vector<char> excl;
};

// The [MD]/[cut-offs]/[sphere]/[solvent]/[intervals]/[files]/[lambdas]
// settings of a qdyn input file (src/md.f90:4453-5358) that this engine uses.
// Extra restraints of the input file (src/md.f90:259-296), 0-based atoms;
// ipsi is the FEP state (1-based, 0 = every state) as in the file.
struct RstSeq { int i, j; double fk; int ih, to_centre; };
struct RstPos { int i; double x[3], fk[3]; int ipsi; };
struct RstDis { int i, j; double d1, d2, fk; int ipsi; };
struct RstAng { int i, j, k; double ang, fk; int ipsi; };   // ang in degrees
struct RstWal { int i, j; double d, fk, aMorse, dMorse; int ih; };

struct MDInput {
  int nsteps = 0;
  double stepsize = 1.0;       // fs
  double temp0 = 300.0;
  double tau_t = 100.0;        // fs, converted by 0.020462 at prep
  bool separate_scaling = true;
  int iseed = 1;
  bool have_tmaxw = false; double tmaxw = 0;
  bool shake_solvent = true, shake_solute = false, shake_hydrogens = false;
  bool use_lrf = true;
  double rcpp = 10, rcww = 10, rcpw = 10, rcq = 99, rclrf = 99;
  bool have_shell_radius = false; double shell_radius = 0.85;
  double fk_pshell = 10.0, fk_fix = 200.0;
  bool freeze = false, exclude_bonded = false;
  double rwat_in = -1, fk_wsphere = -1, dwmz = -1, awmz = -1, fkwpol = -1;
  bool wpol_restr = true, wpol_born = true;
  int nbcycle = 10, iout_cycle = 10, itemp_cycle = 10, iene_cycle = 0, itrj_cycle = 0;
  std::string top_file, restart_file, final_file, energy_file, fep_file;
  std::vector<double> lambdas;
  std::string thermostat = "berendsen", integrator = "leap-frog";
  // Langevin: friction (default 1/tau_T) and Gaussian instead of uniform
  // random forces; Nose-Hoover: chain length and thermostat mass (default
  // kT/tau_T^2)
  bool have_friction = false, langevin_gauss = false, have_nhq = false;
  double friction = 0.0, nhq = 0.0;
  int nhchains = 10;
  std::string restraint_file;            // [files] restraint: xtop from a restart file
  std::string trj_file;                  // [files] trajectory
  std::vector<std::string> trajectory_atoms;  // [trajectory_atoms] mask rows
  std::vector<RstSeq> rstseq;
  std::vector<RstPos> rstpos;
  std::vector<RstDis> rstdis;
  std::vector<RstAng> rstang;
  std::vector<RstWal> rstwal;
  // residue:atom texts of distance restraints, resolved against the topology
  std::vector<std::pair<std::string, std::string>> rstdis_text;
  // [PBC] (src/md.f90:4516-4598)
  bool box = false, rigid_box_centre = false, put_solute_back = false, put_solvent_back = false;
  bool have_rcq = false;
  // constant pressure (MC_volume): volume moves of at most max_vol_displ
  // A^3 every ivolume_cycle steps; pressure in kcal/mol/A^3
  bool constant_pressure = false;
  double max_vol_displ = 0.0, pressure = 1.0;
  int pressure_seed = 3781, ivolume_cycle = 10;
};

// Q-bonded terms of the FEP file (src/qatom.f90:920-1790): atoms 0-based,
// per-state type codes 1-based (0 = term absent in that state).
struct QBondLib { double dmz, amz, r0, fk; };
struct QAngLib { double fk, ang0, ureyfk, ureyr0; };   // ang0 in radians
struct QTorLib { double fk, rmult, deltor; };          // deltor in radians
struct QImpLib { double fk, imp0; };                   // imp0 in radians
struct QTerm { int i, j, k, l; std::vector<int> cod; };
// coupling of a Q angle/torsion/improper to a Morse Q bond: term, bond
// (0-based), flag 1 = scale by 1 - the bond's Morse factor
struct QCoupling { int term, bond, flag; };

// EVB off-diagonal function H_ij = A exp(-mu r_kl) (src/qatom.f90:1809-1850);
// states and Q atoms 0-based.
struct OffDiag { int i, j, k, l; double A, mu; };

// FEP file content used by this engine: [FEP] states and
// softcore_use_max_potential, [atoms], [change_charges], [atom_types],
// [change_atoms], [softcore] and [off_diagonals] (src/qatom.f90:301-1975).
struct FepData {
  int nstates = 0;
  std::vector<int> iqseq;                 // topology atom (0-based) per Q atom
  std::vector<std::vector<double>> qcrg;  // [nqat][nstates], unscaled
  bool have_charges = false;
  // Q-atom LJ types: qavdw/qbvdw [nqlib][3] (LJ codes 1..3), qmass; with
  // [change_atoms] every Q atom has a type per state (qvdw_flag)
  std::vector<std::string> qtac;
  std::vector<std::array<double, 3>> qavdw, qbvdw;
  std::vector<double> qmass;
  bool qvdw = false;
  std::vector<std::vector<int>> qiac;     // [nqat][nstates], 0-based rows of qtac
  bool softcore = false, softcore_max_potential = false;
  std::vector<std::vector<double>> alpha_max;  // [nqat][nstates]
  std::vector<OffDiag> offd;
  std::vector<QBondLib> qbondlib;
  std::vector<QAngLib> qanglib;
  std::vector<QTorLib> qtorlib;
  std::vector<QImpLib> qimplib;
  std::vector<QTerm> qbnd, qang, qtor, qimp;
  std::vector<QCoupling> ang_coupl, tor_coupl, imp_coupl;
  // [soft_pairs] Q-Q pairs with exponential repulsion (0-based Q numbers);
  // [el_scale] Q-Q electrostatic scale per state; [excluded_pairs] atom
  // pairs (0-based) excluded in the states flagged; qq_use_library_charges
  std::vector<std::pair<int, int>> soft_pairs;
  struct ElScale { int iq, jq; std::vector<double> scale; };
  std::vector<ElScale> el_scale;
  struct ExSpec { int i, j; std::vector<char> flag; };
  std::vector<ExSpec> exspec;
  bool qq_lib_charges = false;
  int qswitch = -1;                       // [PBC] switching_atom (0-based), periodic box only
  // [shake_constraints]: atom pair (0-based) and distance per state
  struct QShake { int i, j; std::vector<double> dist; };
  std::vector<QShake> qshake;
  // [monitor_groups]: atom groups (0-based atoms); [monitor_group_pairs]:
  // pairs of 0-based group numbers whose nonbonded energies are monitored
  std::vector<std::vector<int>> mon_groups;
  std::vector<std::pair<int, int>> mon_pairs;
};

struct ShakeBond { int i, j; double dist2; };
struct ShakeMol { std::vector<ShakeBond> bond; };

struct WaterShell {
  double rout, dr, cstb, avtheta, avn_insh, theta_corr;
  int n_insh;
};

// Q-Q / Q-neighbour pair for one state (NBQ_TYPE, src/md.f90:347-352).
struct NbqPair { int iq, j, jq, ljcod; double el_scale; };

// qdyn's integer LCG (randm, src/md.f90:17543-17577) with its saved state.
struct Randm {
  int irand = 0; bool started = false;
  double next(int& ig);
};

// Everything md_run needs, after initialize/topology/prep_coord/get_fep/
// prep_sim/init_shake/make_nbqqlist/shrink_topology (src/qdyn.f90:108-132).
struct Setup {
  Topology top;
  MDInput inp;
  int natom = 0, nat_solute = 0, nwat = 0, solv_atom = 3;
  std::vector<double> x, v, xtop, winv;
  std::vector<double> crg;            // scaled by sqrt(coulomb_constant)
  std::vector<int> iac;
  std::vector<char> heavy, excl, shell;
  std::vector<int> iwhich_cgp;
  std::vector<int> ljcod;             // natyps*natyps, values 1 or 2
  int natyps = 0;
  double dt = 0, tau_t = 0;
  // thermostat (0 Berendsen, 1 Langevin, 2 Nose-Hoover) and integrator
  // (velocity-Verlet: two half-step velocity updates); dv_mod the velocity
  // update's step (dt, or dt/2 with velocity-Verlet), the Langevin friction
  // factor 1 - friction dv_mod and per-atom random force scale
  // sqrt(gkT / winv), the Nose-Hoover chain masses and kT (src/md.f90:
  // 4646-4708, 4416-4424, 6446-6462)
  int thermostat = 0;
  bool velocity_verlet = false;
  double dv_mod = 0, friction = 0, dv_friction = 0, kbt = 0;
  bool langevin_gauss = false;
  std::vector<double> rands;
  std::vector<double> qnh;
  // solvent
  double chg_solv[3] = {0,0,0}, alj_solv[3][2] = {{0}}, blj_solv[3][2] = {{0}};
  double rwat = 0, fk_wsphere = 0, dwmz = 0, awmz = 0, fkwpol = 0;
  bool wpol_restr = true;
  double crgtot = 0, crgqtot = 0, mu_w = 0;
  std::vector<WaterShell> wshell; int n_max_insh = 0;
  double rexcl_i = 0;
  // bonded terms after shrink_topology
  std::vector<Bond> bnd; int nbonds_solute = 0;
  std::vector<Angle> ang; int nangles_solute = 0;
  std::vector<Torsion> tor; int ntors_solute = 0;
  std::vector<Torsion> imp; int nimps_solute = 0;
  // SHAKE
  std::vector<ShakeMol> shake_mol; int shake_constraints = 0; int nmol = 0;
  int ndegf = 0, ndegfree = 0, ndegf_solute = 0, ndegf_solvent = 0;
  int ndegfree_solute = 0, ndegfree_solvent = 0;
  double exclshk = 0;
  bool detail_temps = false, separate_scaling = true;
  // Q atoms
  int nqat = 0, nstates = 0;
  std::vector<int> iqseq, iqatom;     // iqatom: Q index+1 per atom (0 = not Q)
  std::vector<double> qcrg;           // [iq*nstates+s], scaled
  std::vector<double> lambda;
  std::vector<int> qconn;             // [s][iq][atom] connectivity (1..9)
  std::vector<std::vector<NbqPair>> nbqq;  // per state
  // monitored group pairs: the atoms of both groups and the LJ code of each
  // (state, i, j) atom pair, [(s*ni + i)*nj + j] (nbmonitorlist,
  // src/md.f90:11151-11226)
  struct MonPair { std::vector<int> ai, aj; std::vector<int> ljcod; };
  std::vector<MonPair> mon;
  // Q-atom types (qvdw_flag): qiac [iq*nstates+s] (0-based), qavdw/qbvdw
  // [type*3 + LJ code-1]; soft-core table sc_lookup(iq, column, s) with
  // columns 0..natyps-1 topology types and natyps+jq Q atoms, stored
  // [(iq*(natyps+nqat) + column)*nstates + s] (src/qatom.f90:1852-1975)
  bool qvdw_flag = false;
  std::vector<int> qiac;
  std::vector<double> qavdw, qbvdw;
  std::vector<double> sc_lookup;
  std::vector<OffDiag> offd;
  bool qq_lib_charges = false;         // nonbond_qq_lib_charges
  int ff_type = 1;
  // Q-bonded terms after the break-bond completion of qatom_load_fep, with
  // their libraries; coupled bond per angle/torsion/improper (-1 = none)
  // and its flag
  std::vector<QBondLib> qbondlib;
  std::vector<QAngLib> qanglib;
  std::vector<QTorLib> qtorlib;
  std::vector<QImpLib> qimplib;
  std::vector<QTerm> qbnd, qang, qtor, qimp;
  std::vector<int> qang_cb, qang_cf, qtor_cb, qtor_cf, qimp_cb, qimp_cf;
  // extra restraints (p_restrain), atoms resolved
  std::vector<RstSeq> rstseq;
  std::vector<RstPos> rstpos;
  std::vector<RstDis> rstdis;
  std::vector<RstAng> rstang;
  std::vector<RstWal> rstwal;
  std::vector<double> mass;           // topology type mass per atom (iaclib(iac)%mass)
  // periodic box: box lengths, centre and inverse lengths (src/md.f90:17873-17880),
  // the Q-atom switching atom, per molecule 1/mass, and the number of solute
  // molecules (nmol - (nres - nres_solute))
  bool pbc = false;
  double boxl[3] = {0, 0, 0}, boxc[3] = {0, 0, 0}, inv_boxl[3] = {0, 0, 0};
  int qswitch = -1;
  std::vector<double> mol_minv;
  int nsolute_mol = 0;
  bool freeze = false;
  // randm's saved state after the initialisation (the Maxwell velocities
  // draw from it); the volume moves continue the same stream
  Randm rng;
};

// Readers.  Each throws std::runtime_error with a message on bad input.
Topology read_topology(const std::string& path);
MDInput read_input(const std::string& path);
FepData read_fep(const std::string& path, int natom);
// Restart files (write_xfin src/md.f90:18705-18736, prep_coord :16477-16577):
// x, v, then theta_corr of the polarisation shells (sphere) or the box
// lengths and centre (periodic box, box != nullptr: 6 doubles).
bool read_restart(const std::string& path, int natom, std::vector<double>& x,
                  std::vector<double>& v, std::vector<double>* theta_corr, double* box = nullptr);
void write_restart(const std::string& path, const std::vector<double>& x,
                   const std::vector<double>& v, const std::vector<double>* theta_corr,
                   const double* box = nullptr);

// Trajectory atom mask of the [trajectory_atoms] rows (mask_add/update,
// src/mask.f90:129-361): per atom 1 if written; "all" when rows is empty.
std::vector<char> trajectory_mask(const Setup& s, const std::vector<std::string>& rows);

// DCD trajectory of qdyn's trj module (trj_initialize/trj_create/trj_write,
// src/trj.f90:72-233): header records, then per frame the masked x, y and z
// as REAL(4) records.
class DcdWriter {
 public:
  DcdWriter(const Setup& s, const std::string& path);
  ~DcdWriter();
  void write(const double* x);
  int natoms() const { return (int)atoms_.size(); }

 private:
  FILE* fp_ = nullptr;
  std::vector<int> atoms_;
};


// Build the Setup from an input file (paths in it are relative to its
// directory).  Runs the whole qdyn initialisation sequence, including
// Maxwell velocities, initial SHAKE and centre-of-mass stop when the input
// asks for them (those call the engine-independent CPU SHAKE below).
Setup prepare(const std::string& input_path);

// Plain CPU SHAKE used only at initialisation (initial_shaking,
// src/md.f90:4430-4450); returns total iterations or -1 on failure.
int shake_cpu(const Setup& s, const std::vector<double>& xx, std::vector<double>& x);

}  // namespace qsys
