/*
 * qgpu — C ABI of the MI355X engine for qdyn's nonbonded / SHAKE / step path.
 *
 * The reference has no FFI for this path: its step is a set of argument-less
 * subroutines on module globals in src/md.f90.  Each entry point below
 * replaces one reference call site (cited per function); a Fortran host binds
 * them through ISO_C_BINDING (interface module qsource_amd/fortran/qgpu.f90).
 *
 * Conventions
 *  - every function returns 0 on success, nonzero on failure; the message is
 *    available from qgpu_last_error() (the Fortran side calls die(msg) on a
 *    nonzero status, as src/md.f90:554-613 does);
 *  - coordinate, velocity and gradient arrays are FP64, AoS, length 3*natom,
 *    in qdyn's own order (x(:), v(:), d(:) of src/md.f90:309-313); d is the
 *    gradient dV/dx (force = -d);
 *  - one handle holds a batch of instances (one instance = one qdyn input:
 *    topology + FEP file + lambdas); instance indices are 0-based;
 *  - the library owns device memory; no torch types cross this boundary.
 */
#ifndef QGPU_H
#define QGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qgpu_handle qgpu_handle;

enum qgpu_precision {
  QGPU_FP64 = 0,   /* plain: all arithmetic FP64, reproduces the reference */
  QGPU_MIXED = 1   /* FP32 pair arithmetic, FP64 accumulation/SHAKE/integrator/LRF */
};

/* Energy terms returned by qgpu_energies, in this order (E%... of
 * src/nrgy.f90:48-63 as printed by write_out, src/md.f90:18545-18691). */
enum qgpu_energy_term {
  QGPU_E_PP_EL, QGPU_E_PP_VDW, QGPU_E_P_BOND, QGPU_E_P_ANGLE, QGPU_E_P_TORSION, QGPU_E_P_IMPROPER,
  QGPU_E_WW_EL, QGPU_E_WW_VDW, QGPU_E_W_BOND, QGPU_E_W_ANGLE, QGPU_E_W_TORSION, QGPU_E_W_IMPROPER,
  QGPU_E_PW_EL, QGPU_E_PW_VDW, QGPU_E_LRF,
  QGPU_E_QX_EL, QGPU_E_QX_VDW, QGPU_E_Q_BOND, QGPU_E_Q_ANGLE, QGPU_E_Q_TORSION, QGPU_E_Q_IMPROPER,
  QGPU_E_RST_TOTAL, QGPU_E_RST_FIX, QGPU_E_RST_RADIAL, QGPU_E_RST_WPOL, QGPU_E_RST_SHELL,
  QGPU_E_RST_PROTEIN,
  QGPU_E_TOTAL, QGPU_E_POTENTIAL, QGPU_E_KINETIC,
  QGPU_E_COUNT
};

/* Replaces qdyn's initialisation sequence (src/qdyn.f90:108-141:
 * initialize, topology, prep_coord, get_fep, prep_sim, init_shake,
 * make_nbqqlist, shrink_topology, maxwell/initial_shaking/
 * stop_cm_translation).  Reads each qdyn input file (and the topology, FEP
 * and restart files it names) and uploads the systems as one batch. */
int qgpu_create(const char* const* input_files, int n_instances, int precision,
                qgpu_handle** out);

/* Releases the handle and its device memory. */
int qgpu_destroy(qgpu_handle* h);

/* Number of atoms of instance k (nat_pro, src/topo.f90:104). */
int qgpu_natom(qgpu_handle* h, int k, int* natom);

/* Sizes of instance k (sizes[0..6]): atoms, solute atoms (nat_solute),
 * waters, charge groups, solute charge groups (ncgp_solute), Q atoms and
 * FEP states (src/topo.f90:104-130, src/qatom.f90:301-420). */
int qgpu_system_sizes(qgpu_handle* h, int k, int* sizes, int cap);

/* Number of FEP states of instance k (nstates, src/qatom.f90:301-420;
 * 0 without an FEP file). */
int qgpu_nstates(qgpu_handle* h, int k, int* nstates);

/* Replaces make_pair_lists (src/md.f90:5985-6071, called at :6505): charge-
 * group centres, solute/solvent pair lists and LRF multipoles, from the
 * instance coordinates currently on the device. */
int qgpu_make_pair_lists(qgpu_handle* h);

/* Replaces pot_energy_nonbonds + restraints + nonbond_qq (src/md.f90:
 * 16096-16140): computes the gradient for every instance on the device.
 * The gradient is fetched with qgpu_get_gradient. */
int qgpu_pot_energy(qgpu_handle* h);

/* Term groups of qgpu_pot_energy_terms. */
enum qgpu_term_group {
  QGPU_TERMS_WW = 1,     /* water-water (nonbond_ww_spc / nonbond_3atomsolvent) */
  QGPU_TERMS_PW = 2,     /* solute-water (nonbond_pw) */
  QGPU_TERMS_PP = 4,     /* solute-solute (nonbond_pp) */
  QGPU_TERMS_QX = 8,     /* Q-solute and Q-water (nonbond_qp, nonbond_qw_*) */
  QGPU_TERMS_LRF = 16,   /* LRF Taylor (lrf_taylor) */
  QGPU_TERMS_OTHER = 32, /* bonded, Q-Q, Q-bonded, restraints, polarisation */
  QGPU_TERMS_ALL = 63
};

/* As qgpu_pot_energy, with only the term groups in `mask` in the gradient
 * (per-term error reports of the mixed-precision path against the plain
 * one).  Later evaluations include every term again.  No counterpart in
 * the reference. */
int qgpu_pot_energy_terms(qgpu_handle* h, int mask);

/* Replaces pot_energy's call of pot_energy_nonbonds (src/md.f90:16096,
 * 16261-16344): the solute-solute, solute-water, Q-solute, water-water and
 * Q-water pair terms over the current pair lists and the LRF Taylor term,
 * for instance k at the coordinates x (3*natom FP64; they become the
 * instance's coordinates, as the reference's x is the module global).
 * d (3*natom) accumulates the gradient, as the reference adds into the d
 * that pot_energy zeroed.  E[7] = E%pp%el, E%pp%vdw, E%pw%el, E%pw%vdw,
 * E%ww%el, E%ww%vdw, E%LRF; EQ[4*nstates] = per state EQ%qp%el, %qp%vdw,
 * EQ%qw%el, %qw%vdw (eq_cap >= 4*nstates).  nonbond_qq, the bonded terms
 * and the restraints are not included (qgpu_pot_energy_terms selects any
 * group). */
int qgpu_pot_energy_nonbonds(qgpu_handle* h, int k, const double* x, double* d, double* E, double* EQ,
                             int eq_cap);

/* Replaces the reference's `integer function shake(xx, x)`
 * (src/md.f90:17581-17679), which md_run calls after the update
 * (src/md.f90:6596): every SHAKE molecule of instance k is iterated until
 * each constraint meets |d0^2 - r^2| < 1e-4 d0^2, moving x (3*natom FP64,
 * in place) along the bonds of the pre-step coordinates xx.  *niter (may be
 * NULL) = the iterations summed over the molecules, divided by the number
 * of molecules (integer division, as the function's result).  A molecule
 * that needs 1000 iterations fails: the call returns nonzero with
 * qgpu_last_error() = "shake failure" (die, src/md.f90:17673). */
int qgpu_shake(qgpu_handle* h, int k, const double* xx, double* x, int* niter);

/* Replaces md_run's loop body (src/md.f90:6471-6612): nsteps steps of
 * list update (every NBcycle), forces, leapfrog/Berendsen, SHAKE and
 * temperature, with x, v and d resident on the device.
 * Failures stop the batch where die() stops qdyn (src/md.f90:554-613): a
 * SHAKE failure (shake, :17645-17673), a NaN temperature (tested at every
 * step; md_run tests it at printed steps, :6654-6656) or a full water
 * shell leaves every instance at the failing step and returns nonzero,
 * and so does every later qgpu_step; qgpu_get_state still returns that
 * state (for the restart die writes).  A pair-list row that outgrows its
 * capacity is not a failure: the rows are regrown and the list update is
 * redone, as the reference reallocates its lists (:900-1128). */
int qgpu_step(qgpu_handle* h, int nsteps);

/* Failure record of instance k: *bits = 1 pair-list row over capacity,
 * 2 water-shell capacity, 4 SHAKE failure, 8 NaN temperature (0 = none);
 * *at_step = the step it was found at (-1 when none).  No counterpart in the
 * reference, whose die() ends the program. */
int qgpu_instance_status(qgpu_handle* h, int k, int* bits, int* at_step);

/* Row capacities of instance k's water-water, water-solute and solute-water
 * lists (caps[3]) and the number of times this batch regrew list rows.
 * The library sizes the rows from the starting coordinates; the reference
 * sizes its lists from counts at distribute_nonbonds (src/md.f90:2696-2714). */
int qgpu_list_capacity(qgpu_handle* h, int k, int* caps, int* regrowths);

/* Replaces the row capacities of instance k (0 keeps one); the next list
 * update fills the new rows and regrows any that are too short. */
int qgpu_set_list_capacity(qgpu_handle* h, int k, int ww, int wp, int pw);

/* State transfer of instance k (x, v: 3*natom FP64 each; either may be
 * NULL).  Setting v also sets the temperature, Tfree and the Berendsen
 * scales of the next step from it, as md_run's first temperature() call
 * does for the starting velocities (src/md.f90:6400-6404); Tfree enters the
 * radial water restraint (src/md.f90:18055-18102). */
int qgpu_get_state(qgpu_handle* h, int k, double* x, double* v);
int qgpu_set_state(qgpu_handle* h, int k, const double* x, const double* v);

/* Gradient d (3*natom) of instance k from the last force evaluation. */
int qgpu_get_gradient(qgpu_handle* h, int k, double* d);

/* Energy terms of instance k from the last force evaluation, ordered as
 * enum qgpu_energy_term; writes min(cap, QGPU_E_COUNT) values. */
int qgpu_energies(qgpu_handle* h, int k, double* e, int cap);

/* Per-state Q-atom energies of instance k (EQ(s), src/qatom.f90:197-200):
 * 12 values per state: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw total, then
 * the Q-bonded bond angle torsion improper (qbond/qangle/qtorsion/qimproper,
 * src/md.f90:17022-17540) and the state's restraint energy (p_restrain,
 * src/md.f90:15473-15802). */
int qgpu_qenergies(qgpu_handle* h, int k, double* e, int cap);

/* Replaces the energy-file writes of md_run (put_ene every iene_cycle
 * steps, src/md.f90:6628-6630) for batched runs: every `iene` steps
 * (istep > 0) each instance's per-state Q energies are recorded on the
 * device, up to max_frames frames, with no host synchronisation inside
 * qgpu_step.  Steps are counted from this call (the reference counts from
 * the start of the run).  iene = 0 switches recording off. */
int qgpu_set_energy_interval(qgpu_handle* h, int iene, int max_frames);

/* Recorded frames of instance k: nframes x nstates x 14 values in put_ene
 * order (total, q bond/angle/torsion/improper, qx el/vdw, qq el/vdw,
 * qp el/vdw, qw el/vdw, restraint; src/nrgy.f90:128-130).  Writes at most
 * `cap` doubles; *nframes gets the number recorded. */
int qgpu_energy_series(qgpu_handle* h, int k, double* out, int cap, int* nframes);

/* Mixed-precision batches only: the same frames as qgpu_energy_series, with
 * the per-state Q energies re-evaluated on the frame's coordinates with FP64
 * pair arithmetic (the plain engine's arithmetic).  Gives the FP64 dU of
 * exactly the configurations the mixed run sampled, for the precision error
 * of dU and dG.  No counterpart in the reference (it has one precision). */
int qgpu_energy_series_fp64(qgpu_handle* h, int k, double* out, int cap, int* nframes);

/* Number of EVB off-diagonal functions ([off_diagonals] lines of the FEP
 * file, src/qatom.f90:1809-1850) of instance k. */
int qgpu_noffdiag(qgpu_handle* h, int k, int* n);

/* Off-diagonal elements of the recorded energy frames of instance k:
 * nframes x noffd x (Hij, r_kl), computed on the frame's coordinates as
 * offdiag does (src/md.f90:15444-15470) and written after each frame's state
 * records by put_ene (src/nrgy.f90:136-138).  Writes at most `cap` doubles. */
int qgpu_offdiag_series(qgpu_handle* h, int k, double* out, int cap, int* nframes);

/* Exports pair list `which` of instance k in the reference's layout and
 * order (NB_TYPE i, j, LJcod; 0-based atoms): 0 = nbpp, 1 = nbpw, 2 = nbww
 * (src/md.f90:329-433); 3 = nbpp as the force kernel evaluates it (charge-
 * group list plus the per-entry pair masks formed on the device), which
 * equals list 0.  Call with out == NULL to get the count in *n. */
int qgpu_get_pairlist(qgpu_handle* h, int k, int which, int* out, long* n);

/* Sizes of the pair lists of instance k in the reference's layout (the
 * counts md_run prints after make_pair_lists, src/md.f90:6508-6512):
 * counts[0..4] = nbpp, nbpw, nbww, nbqp, and Q-water atom pairs. */
int qgpu_list_counts(qgpu_handle* h, int k, long* counts);

/* Work of one list update of instance k at its current coordinates, for
 * roofline accounting (no counterpart in the reference): counts[0..5] =
 * atom-to-centre multipole evaluations (lrf_update, src/md.f90:5896-5981)
 * of the water-side and of the solute-side list kernel, water-water list
 * entries (both directions), water-solute group entries, solute group pairs
 * within the cut-off (both directions), solute-water entries. */
int qgpu_work_counts(qgpu_handle* h, int k, double* counts, int cap);

/* Current temperature (Temp, src/md.f90:6159-6163) of instance k. */
int qgpu_temperature(qgpu_handle* h, int k, double* temp);

/* Water-polarisation shell corrections theta_corr (restart record,
 * src/md.f90:18729-18731) of instance k; *n gets the shell count. */
int qgpu_theta_corr(qgpu_handle* h, int k, double* out, int* n);

/* Runs nsteps steps (no graph) with a HIP event pair around every kernel
 * launch on the engine's stream and writes the average ms per launch of
 * each kernel, in this order: cgp_centers, water_lists, solute_lists,
 * water_forces, solute_forces, terms (bonded terms and Q-Q pairs), q_forces,
 * term_gather, watpol_theta,
 * watpol_force, energy_record, integrate, temperature (0 if not launched),
 * less the measured time of an empty event pair on the same stream.
 * A measurement aid with no counterpart in the reference (its -DPROFILING
 * build times whole sections, src/md.f90:6370-6381). */
int qgpu_profile_kernels(qgpu_handle* h, int nsteps, double* ms, int cap);

/* Average device time per step (ms) of the last qgpu_step call. */
int qgpu_last_step_ms(qgpu_handle* h, double* ms);

/* ---- Learned far-field surrogate ------------------------------------------
 * No counterpart in the reference: an alternative to the LRF Taylor far field
 * (lrf_update/lrf_taylor, src/md.f90:5824-5981).  A 4-128-128-3 SiLU MLP maps
 * per-charge-group features (net charge, dipole along and across the radial
 * direction, radial position / sphere radius) to the far-field potential and
 * the field along and across the radial direction at the group centre.  The
 * 128x128 layer runs on BF16 matrix cores with FP32 accumulation; training
 * (Adam, mean squared error of standardised outputs) runs on the GPU. */
typedef struct qgpu_mlp qgpu_mlp;

int qgpu_mlp_create(int seed, qgpu_mlp** out);
int qgpu_mlp_destroy(qgpu_mlp* m);
/* Number of parameters (W1[128][4] b1 W2[128][128] b2 W3[3][128] b3). */
int qgpu_mlp_nparams(int* n);
/* FP32 parameters in that order; norm (14 floats): input means and standard
 * deviations (4 + 4), output means and standard deviations (3 + 3); may be
 * NULL. */
int qgpu_mlp_get_params(qgpu_mlp* m, float* params, float* norm);
int qgpu_mlp_set_params(qgpu_mlp* m, const float* params, const float* norm);
/* Training samples (X: n x 4 features, Y: n x 3 targets); sets the
 * normalisation from their means and standard deviations. */
int qgpu_mlp_set_data(qgpu_mlp* m, const float* X, const float* Y, int n);
/* Adam epochs over shuffled mini-batches; *loss = last epoch's mean squared
 * error of the standardised outputs. */
int qgpu_mlp_train(qgpu_mlp* m, int epochs, int batch, float lr, int seed, double* loss);
/* Gradient of the mean squared error over (X, Y) w.r.t. all parameters. */
int qgpu_mlp_gradient(qgpu_mlp* m, const float* X, const float* Y, int n, float* grad, double* loss);
/* As qgpu_mlp_gradient, allocating no device memory.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving or training loops).  The training activations and
 * the staging copies of X and Y live in a static device array of about
 * 5.1 MiB per GPU (held until the process ends), for at most 2048 samples
 * per call; larger n is refused.  Calls that overlap on one GPU share that
 * array and race.  Same kernels and launch shapes as qgpu_mlp_gradient:
 * the gradient and loss are bitwise the same. */
int qgpu_mlp_gradient_noalloc(qgpu_mlp* m, const float* X, const float* Y, int n, float* grad, double* loss);
/* Predictions (n x 3, de-standardised) for features X (n x 4). */
int qgpu_mlp_predict(qgpu_mlp* m, const float* X, int n, float* Y);
const char* qgpu_mlp_last_error(void);

/* Far-field samples of instance k at its current coordinates (pair lists and
 * LRF multipoles rebuilt, so each target is the exact potential and field of
 * all far charges at the group centre): X (ncgp x 4), Y (ncgp x 3); *n =
 * ncgp.  qgpu_far_lrf_estimate gives, for the same groups, the estimate of
 * the LRF multipoles of the last list update at the current centres. */
int qgpu_far_samples(qgpu_handle* h, int k, float* X, float* Y, int cap, int* n);
int qgpu_far_lrf_estimate(qgpu_handle* h, int k, float* X, float* Y, int cap, int* n);
/* Far field of every instance from model m (kept alive by the caller)
 * instead of the LRF multipoles; m = NULL restores LRF. */
int qgpu_set_far_model(qgpu_handle* h, qgpu_mlp* m);

/* Launch-shape tuner (no counterpart in the reference).  The four large
 * kernels (water lists, solute lists, water forces, solute forces) have
 * launch shapes: workgroup mapping (0 2-D grid, 1 instance-major work map,
 * 2 instances kept on one XCD each) x workgroup size (256, 128); for the
 * two force kernels partner prefetch (0 off, 1 on); for the FP32 water
 * forces the register budget (0: 6 wavefronts per SIMD, 1: 4).  Every shape
 * gives bitwise the same results.  Each shape is timed over `reps` launches
 * on this batch and the fastest is kept; qgpu_step does this by itself at
 * its first call unless QGPU_AUTOTUNE=0.  Only at a pair-list step (the
 * tuning launches rebuild the lists).  ms (8 values): per kernel the chosen
 * and the default shape's ms per launch; shapes (16): per kernel mapping,
 * workgroup size, prefetch and register budget. */
int qgpu_autotune(qgpu_handle* h, int reps, double* ms, int* shapes);
/* Fixes the launch shape of tuned kernel `kernel` (0-3 as above) and
 * switches the tuning at the first step off. */
int qgpu_set_launch_shape(qgpu_handle* h, int kernel, int map, int block, int prefetch, int occupancy);

/* Replaces offdiag (src/md.f90:15444-15470, called by md_run at :6568):
 * every [off_diagonals] function of instance k at its current coordinates,
 * 6 values each: state i, state j, Q atoms k and l (1-based), Hij =
 * A exp(-mu r_kl) and r_kl.  *n = number of functions; out is written
 * while cap allows. */
int qgpu_offdiag(qgpu_handle* h, int k, double* out, int cap, int* n);

/* State indices (i, j, 1-based) of every [off_diagonals] function of
 * instance k, as the energy file's off-diagonal record holds them
 * (OFFDIAG_SAVE, src/nrgy.f90:73-78); *n = number of functions. */
int qgpu_offdiag_states(qgpu_handle* h, int k, int* ij, int cap, int* n);

/* Energies of the FEP file's monitored group pairs ([monitor_groups],
 * [monitor_group_pairs]) at the current coordinates of instance k; replaces
 * md_run's `call nonbond_monitor` at output steps (src/md.f90:6562-6564,
 * nonbond_monitor :11228-11350; printed by write_out :18653-18663).  Per pair:
 * Vwsum, Vwel, Vwvdw (lambda-weighted), then Vel and Vvdw of every state
 * (3 + 2 nstates values).  *n receives the number of values; out is written
 * when cap >= *n. */
int qgpu_monitor(qgpu_handle* h, int k, double* out, int cap, int* n);

/* ---- qdyn's output files (host only; no device needed) ------------------
 * Trajectory: replaces trj_initialize + trj_commit_mask + trj_create of
 * init_trj (src/md.f90:16668-16684, src/trj.f90:72-204) for the run of
 * `input_path`: the [trajectory_atoms] mask (all atoms when absent) and the
 * DCD header; trj_path NULL takes [files] trajectory of the input.
 * qgpu_trj_write replaces write_trj/trj_write (src/md.f90:18695-18701,
 * src/trj.f90:211-233): one frame of x (3*natom FP64) as REAL(4) records. */
typedef struct qgpu_trj qgpu_trj;
int qgpu_trj_create(const char* input_path, const char* trj_path, qgpu_trj** out);
int qgpu_trj_natoms(qgpu_trj* t);
int qgpu_trj_write(qgpu_trj* t, const double* x);
int qgpu_trj_close(qgpu_trj* t);

/* Restart file as write_xfin writes it (src/md.f90:18705-18736): canary
 * -1337, x, v (3*natom FP64 each) and, when nshell > 0, the water
 * polarisation shells' theta_corr. */
int qgpu_write_restart(const char* path, int natom, const double* x, const double* v, int nshell,
                       const double* theta_corr);

/* Restart file of a periodic box (write_xfin, src/md.f90:18705-18736):
 * canary -1337, x, v, then the box lengths and the box centre (box: 6 FP64). */
int qgpu_write_restart_box(const char* path, int natom, const double* x, const double* v, const double* box);

/* Periodic box of instance k (src/md.f90:18732-18735): out[0..2] the box
 * lengths, out[3..5] the current box centre (it follows the solute at list
 * updates unless rigid_box_centre is on); *is_box = 0 for a sphere. */
int qgpu_box(qgpu_handle* h, int k, double* out, int* is_box);

/* Constant-pressure volume moves of instance k so far (MC_volume,
 * src/md.f90:18931-19272, called by md_run at :6538-6544; the counters
 * volume_try and volume_acc that write_out prints at :18678-18686):
 * *tries = the number of moves; out[0] = tries, out[1] = accepted, then per
 * move 0 (rejected), 1 (accepted, deltaW <= 0) or 2 (accepted by the random
 * draw), as many as cap allows. */
int qgpu_volume_moves(qgpu_handle* h, int k, int* out, int cap, int* tries);

/* Message of the last failure in this thread ("" when none). */
const char* qgpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QGPU_H */
