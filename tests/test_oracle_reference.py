"""The FP64 oracle against reference qdyn (golden fixtures written by
tests/golden/make_golden.py from the reference's own serial qdyn5).

Pins: every step's printed energy summary (F12.2) and the final restart
(x, v, theta_corr) bit-for-bit; for the energy-file cases every recorded
frame's per-state Q energies (nonbonded, Q-bonded, restraint) in FP64.  A restart that is bitwise identical after
100-250 steps of chaotic MD leaves no room for a dropped term, a wrong sign
or index, or a changed operation order anywhere on the path.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import gensys, md

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# periodic-box runs: 30 steps of 0.5 fs at 50 K, lists every 5 steps, Q-atom
# and LRF cut-offs inside half the box edge
BOX_INP = dict(steps=30, stepsize=0.5, temperature=50.0, initial_temperature=50.0, bath_coupling=5.0,
               non_bond=5, rcq=12.0, rclrf=13.0, lambdas=[0.6, 0.4])
# 300 K runs of 1 fs steps from Maxwell velocities
HOT_INP = dict(steps=30, lambdas=[0.6, 0.4], stepsize=1.0, temperature=300.0, bath_coupling=10.0,
               initial_temperature=300.0)
# constant pressure in the box: volume moves of up to 60 A^3 at 1 bar
NPT = {"constant_pressure": True, "max_volume_displ": 60, "pressure_seed": 4242, "pressure": 1.0,
       "put_solvent_back_in_box": True, "put_solute_back_in_box": False}

CASES = {
    "c1_r15_100": (dict(radius=15.0, n_solute=0, seed=1001), dict(steps=100)),
    "c1_r10_250": (dict(radius=10.0, n_solute=0, seed=1001), dict(steps=250)),
    "c2_r14_q8_30": (dict(radius=14.0, n_solute=40, seed=2001, n_q=8, n_states=2),
                     dict(steps=30, lambdas=[0.7, 0.3], stepsize=0.5,
                          temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)),
    # plain cut-off: pair lists without LRF multipoles, no Taylor term
    "c2_r14_q8_nolrf_30": (dict(radius=14.0, n_solute=40, seed=2003, n_q=8, n_states=2),
                           dict(steps=30, lambdas=[0.5, 0.5], stepsize=0.5, lrf=False,
                                temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)),
    "c5_r12_q8s4_gromos_30": (dict(radius=12.0, n_solute=40, seed=2101, n_q=8, n_states=4, q_types=True,
                                   softcore=20.0, n_offdiag=3, impropers=1),
                              dict(steps=30, lambdas=[0.4, 0.3, 0.2, 0.1], stepsize=0.5, temperature=50.0,
                                   bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                   energy_interval=1)),
    "c5_r12_q8s4_amber_30": (dict(radius=12.0, n_solute=40, seed=2102, n_q=8, n_states=4, q_types=True,
                                  softcore=20.0, n_offdiag=3, impropers=2),
                             dict(steps=30, lambdas=[0.25, 0.25, 0.25, 0.25], stepsize=0.5, temperature=50.0,
                                  bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                  energy_interval=1)),
    # Q-bonded terms (Morse bonds, one breaking and one forming, Q angles,
    # torsions, impropers, couplings) and every kind of input restraint
    "c6_r12_q8s2_qbonded_30": (dict(radius=12.0, n_solute=40, seed=2201, n_q=8, n_states=2, q_types=True,
                                    softcore=60.0, impropers=1, q_bonded=True, restraints=True),
                               dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                    bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                    energy_interval=1, trajectory="case.dcd", trajectory_interval=10,
                                    trajectory_atoms=["1 20", "not heavy", "residue 2 5"])),
    # Q-Q pair options: soft repulsion pairs (Q-atom types), per-state
    # electrostatic scales, special exclusions (Q-Q in one state; Q/non-Q and
    # non-Q/non-Q in all); and the same with library charges for Q-Q pairs
    "c7_r12_q8s2_fepx_30": (dict(radius=12.0, n_solute=40, seed=2301, n_q=8, n_states=2, impropers=1,
                                 fep_pairs=True, q_types=True, softcore=40.0),
                            dict(steps=30, lambdas=[0.5, 0.5], stepsize=0.5, temperature=50.0,
                                 bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                 energy_interval=1)),
    "c7_r12_q8s2_qqlib_30": (dict(radius=12.0, n_solute=40, seed=2301, n_q=8, n_states=2, impropers=1,
                                  fep_pairs=True, qq_lib_charges=True),
                             dict(steps=30, lambdas=[0.5, 0.5], stepsize=0.5, temperature=50.0,
                                  bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                  energy_interval=1)),
    # FEP SHAKE constraints (lambda-weighted distances) with solute SHAKE on:
    # one overwrites a shaken bond, one adds a 1-3 constraint
    "c7_r12_q8s2_qshake_30": (dict(radius=12.0, n_solute=40, seed=2401, n_q=8, n_states=2, q_shake=True),
                              dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                   bath_coupling=2.0, initial_temperature=50.0, shake_solute=True,
                                   energy="case.en", energy_interval=1)),
    # all-atom solvent LJ (solvent type 1): every water atom pair, and the Q
    # atoms with every water atom, fixed Q LJ parameters or Q-atom types
    "c8_r12_q8s2_allatom_30": (dict(radius=12.0, n_solute=40, seed=2501, n_q=8, n_states=2,
                                    solvent_allatom=True),
                               dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                    bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                    energy_interval=1)),
    "c8_r12_q8s2_allatom_qvdw_30": (dict(radius=12.0, n_solute=40, seed=2502, n_q=8, n_states=2,
                                         q_types=True, softcore=30.0, solvent_allatom=True),
                                    dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                         bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                         energy_interval=1)),
    # Rcq below the exclusion radius: the Q-solute and Q-water lists are
    # rebuilt at every list update (every 5 steps) and change during the run
    "c9_r14_q8_dynq_60": (dict(radius=14.0, n_solute=40, seed=2601, n_q=8, n_states=2),
                          dict(steps=60, lambdas=[0.6, 0.4], stepsize=1.0, temperature=300.0,
                               bath_coupling=10.0, initial_temperature=300.0, energy="case.en",
                               energy_interval=1, rcq=7.0, non_bond=5)),
    # a solute reaching through the sphere boundary: charge groups beyond
    # the exclusion radius excluded (fix_shell's excluded-atom restraint),
    # heavy groups beyond 0.85 rexcl_o in the restrained shell, waters near
    # the boundary excluded; and the same with excluded_freeze
    "c10_r12_excl_30": (dict(radius=12.0, n_solute=120, seed=2702, n_q=8, n_states=2, solute_rmax=16.0,
                             exclude_outside=True, exclude_water_radius=11.3),
                        dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                             bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                             energy_interval=1)),
    "c10_r12_excl_freeze_30": (dict(radius=12.0, n_solute=120, seed=2702, n_q=8, n_states=2, solute_rmax=16.0,
                                    exclude_outside=True, exclude_water_radius=11.3),
                               dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                    bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                    energy_interval=1, excluded_freeze=True)),
    # solute hydrogens (non-heavy atoms): SHAKE of bonds to hydrogen with one
    # Berendsen scale for solute and solvent; hydrogens with sequence
    # restraints, Q-atom types and solute SHAKE; bonded terms of excluded
    # atoms removed (exclude_bonded)
    "c11_r12_hyd_shake_30": (dict(radius=12.0, n_solute=60, seed=2801, n_q=8, n_states=2, hydrogens=True),
                             dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                  bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                  energy_interval=1, shake_hydrogens=True, separate_scaling=False)),
    "c11_r12_hyd_rst_30": (dict(radius=12.0, n_solute=60, seed=2802, n_q=8, n_states=2, hydrogens=True,
                                restraints=True, q_types=True, softcore=30.0),
                           dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                energy_interval=1, shake_solute=True, shake_hydrogens=True)),
    "c11_r12_hyd_exclbonded_30": (dict(radius=12.0, n_solute=120, seed=2803, n_q=8, n_states=2, hydrogens=True,
                                       solute_rmax=16.0, exclude_outside=True),
                                  dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                       bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                       energy_interval=1, exclude_bonded=True)),
    # monitored group pairs (FEP [monitor_groups] / [monitor_group_pairs]),
    # with fixed Q LJ parameters and with Q-atom types and soft-core
    "c12_r12_q8s2_monitor_30": (dict(radius=12.0, n_solute=40, seed=2901, n_q=8, n_states=2, monitor=True),
                                dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                     bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                     energy_interval=1)),
    "c12_r12_q8s2_monitor_qvdw_30": (dict(radius=12.0, n_solute=40, seed=2902, n_q=8, n_states=2, monitor=True,
                                          q_types=True, softcore=30.0),
                                     dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                          bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                          energy_interval=1)),
    # solute charge groups of up to 12 atoms (the device splits those of
    # more than 8 into slices), one of them partly in the Q region
    "c14_r14_q8s2_group12_30": (dict(radius=14.0, n_solute=80, seed=3201, n_q=8, n_states=2, group_size=12),
                                dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0,
                                     bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                     energy_interval=1)),
    # a breaking bond with non-Q neighbours: topology torsions through it
    # become Q terms (qatom_load_fep's completion); periodic impropers
    "c6_r12_q6s2_qbreak_30": (dict(radius=12.0, n_solute=40, seed=2202, n_q=6, n_states=2, q_types=True,
                                   softcore=60.0, impropers=2, q_bonded=True, restraints=True),
                              dict(steps=30, lambdas=[0.5, 0.5], stepsize=0.5, temperature=50.0,
                                   bath_coupling=2.0, initial_temperature=50.0, energy="case.en",
                                   energy_interval=1)),
    # periodic box (28 A cube, 60-atom solute, 6 Q atoms, 2 states): LRF
    # with shifted multipoles and without; the box centre floating with an
    # off-centre solute; a rigid box centre with every atom shifted and the
    # solute put back too; Q-atom types with soft-core and an energy file;
    # input restraints (minimum-image distances, no wall restraint)
    "c15_box28_q6_30": (dict(radius=14.0, n_solute=60, seed=9101, n_q=6, n_states=2, box=28.0),
                        dict(BOX_INP)),
    "c15_box28_q6_nolrf_30": (dict(radius=14.0, n_solute=60, seed=9101, n_q=6, n_states=2, box=28.0),
                              dict(BOX_INP, lrf=False)),
    "c15_box28_q6_float_30": (dict(radius=14.0, n_solute=60, seed=9101, n_q=6, n_states=2, box=28.0,
                                   solute_offset=(4.0, -3.0, 2.5)), dict(BOX_INP)),
    "c15_box28_q6_rigid_30": (dict(radius=14.0, n_solute=60, seed=9101, n_q=6, n_states=2, box=28.0,
                                   shift=(3.0, -2.5, 1.5)),
                              dict(BOX_INP, pbc={"rigid_box_centre": True, "put_solvent_back_in_box": True,
                                                 "put_solute_back_in_box": True})),
    "c15_box28_q6_qvdw_30": (dict(radius=14.0, n_solute=60, seed=9102, n_q=6, n_states=2, box=28.0, q_types=True,
                                  softcore=20.0, solute_offset=(1.0, 2.0, -2.0)),
                             dict(BOX_INP, energy="case.en", energy_interval=1)),
    "c15_box28_q6_rst_30": (dict(radius=14.0, n_solute=60, seed=9103, n_q=6, n_states=2, box=28.0,
                                 restraints=True, solute_offset=(-3.0, 0.0, 2.0)), dict(BOX_INP)),
    "c15_box28_q6_monitor_30": (dict(radius=14.0, n_solute=60, seed=9104, n_q=6, n_states=2, box=28.0,
                                     monitor=True, solute_offset=(2.0, -4.0, 1.0)), dict(BOX_INP)),
    # constant pressure (MC_volume) every 5 steps: with LRF (a rejected
    # move), at 300 K with moves accepted by the random draw, and without
    # LRF every 10 steps at 500 bar (lists kept through the moves)
    "c15_box28_q6_npt_30": (dict(radius=14.0, n_solute=60, seed=9105, n_q=6, n_states=2, box=28.0,
                                 solute_offset=(1.5, -2.0, 0.5)),
                            dict(BOX_INP, pbc=dict(NPT, max_volume_displ=4), volume_change=5)),
    "c15_box28_q6_npt_hot_30": (dict(radius=14.0, n_solute=60, seed=9107, n_q=6, n_states=2, box=28.0,
                                     solute_offset=(1.5, -2.0, 0.5)),
                                dict(BOX_INP, temperature=300.0, initial_temperature=300.0,
                                     pbc=dict(NPT, max_volume_displ=3, pressure=500.0, pressure_seed=7),
                                     volume_change=5)),
    "c15_box28_q6_npt_nolrf_30": (dict(radius=14.0, n_solute=60, seed=9106, n_q=6, n_states=2, box=28.0),
                                  dict(BOX_INP, lrf=False, pbc=dict(NPT, pressure_seed=99, pressure=500.0),
                                       volume_change=10)),
    # thermostats and integrators at 300 K, 1 fs: Langevin (uniform random
    # forces) with SHAKE; Langevin with Gaussian forces, a set friction and
    # velocity-Verlet; Nose-Hoover with SHAKE (the chain propagated again
    # after SHAKE); Nose-Hoover with velocity-Verlet, 4 chains and a set
    # mass; Berendsen with velocity-Verlet and SHAKE; Langevin at constant
    # pressure (volume moves and random forces share randm's stream)
    "c16_r14_q8_lan_30": (dict(radius=14.0, n_solute=40, seed=3101, n_q=8, n_states=2),
                          dict(HOT_INP, thermostat="langevin")),
    "c16_r14_q8_lan_gauss_vv_30": (dict(radius=14.0, n_solute=40, seed=3102, n_q=8, n_states=2),
                                   dict(HOT_INP, thermostat="langevin", integrator="velocity-verlet",
                                        md_extra={"langevin_random": True, "langevin_friction": 0.5})),
    "c16_r14_q8_nh_shake_30": (dict(radius=14.0, n_solute=40, seed=3103, n_q=8, n_states=2),
                               dict(HOT_INP, thermostat="nose-hoover", shake_hydrogens=True)),
    "c16_r14_q8_nh_vv_30": (dict(radius=14.0, n_solute=40, seed=3104, n_q=8, n_states=2),
                            dict(HOT_INP, thermostat="nose-hoover", integrator="velocity-verlet",
                                 md_extra={"nhchains": 4, "nose-hoover_mass": 2.5})),
    "c16_r14_q8_ber_vv_30": (dict(radius=14.0, n_solute=40, seed=3105, n_q=8, n_states=2),
                             dict(HOT_INP, integrator="velocity-verlet", shake_hydrogens=True)),
    "c16_box28_q6_lan_npt_30": (dict(radius=14.0, n_solute=60, seed=3106, n_q=6, n_states=2, box=28.0),
                                dict(HOT_INP, stepsize=0.5, non_bond=5, rcq=12.0, rclrf=13.0, thermostat="langevin",
                                     pbc=dict(NPT, max_volume_displ=3, pressure=500.0, pressure_seed=7),
                                     volume_change=5)),
}

ORACLE_COLS = {"ww_el": "ww_el", "ww_vdw": "ww_vdw", "lrf": "lrf", "rst_total": "rst_total",
               "rst_fix": "rst_fix", "rst_radial": "rst_radial", "rst_wpol": "rst_wpol",
               "total": "total", "potential": "potential", "kinetic": "kinetic",
               "pp_el": "pp_el", "pp_vdw": "pp_vdw", "pw_el": "pw_el", "pw_vdw": "pw_vdw",
               "p_bond": "p_bond", "p_angle": "p_angle", "p_torsion": "p_torsion",
               "p_improper": "p_improper", "qx_el": "qx_el", "qx_vdw": "qx_vdw",
               "rst_shell": "rst_shell", "rst_protein": "rst_protein", "q_bond": "q_bond",
               "q_angle": "q_angle", "q_torsion": "q_torsion", "q_improper": "q_improper"}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_qdyn(name, tmp_path):
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    sysargs, inpargs = CASES[name]
    s = gensys.solvated_system(**sysargs)
    paths = gensys.write_system(s, str(tmp_path), name, output=1, temp_interval=1, **inpargs)
    sha = hashlib.sha256(open(paths["top"], "rb").read()).hexdigest()
    assert sha == str(g["top_sha"]), "generator output changed; regenerate the golden fixtures"
    cols = [str(c) for c in g["columns"]]
    ref = g["energies"]
    o = Oracle(paths["inp"])
    nsteps = inpargs["steps"]
    for k in range(nsteps + 1):
        if k < nsteps:
            o.run(1)
        else:
            o.final_energy()
        e = o.energies()
        for c, key in ORACLE_COLS.items():
            r = ref[k, cols.index(c)]
            # F12.2 rounding: the printed value is the oracle's, rounded
            assert abs(round(e[key], 2) - r) <= 0.0100001, (k, c, e[key], r)
    x, v = o.state()
    assert np.array_equal(x, g["x"].ravel()), "final coordinates differ from reference qdyn"
    assert np.array_equal(v, g["v"].ravel()), "final velocities differ from reference qdyn"
    if "box" in g.files:
        # periodic box: the restart's box lengths and (moving) centre
        bl, bc = o.box()
        assert np.array_equal(np.concatenate([bl, bc]), g["box"])
    if "volume_moves" in g.files:
        # constant pressure: the verdict of every volume move as printed
        tries, acc, log = o.volume_moves()
        assert [1 if f else 0 for f in log] == list(g["volume_moves"]) and tries == len(log)
        assert acc == sum(1 for f in log if f)
        assert 0 in log or name != "c15_box28_q6_npt_30"        # a rejected move with LRF
        assert 2 in log or name != "c15_box28_q6_npt_hot_30"    # moves accepted by the random draw
        bl, _ = o.box()
        assert md.volume_summary(bl[0] * bl[1] * bl[2], tries, acc) == str(g["volume_summary"])


ENERGY_FILE_CASES = sorted(k for k, (_, inp) in CASES.items() if inp.get("energy"))


@pytest.mark.parametrize("name", ENERGY_FILE_CASES)
def test_oracle_energy_file_frames(name, tmp_path):
    """Per-state Q energies and EVB off-diagonal elements of every recorded
    step, against the reference's own energy file (FP64 values): frame k is
    written at step k (k >= 1) from the coordinates that step starts with."""
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    sysargs, inpargs = CASES[name]
    s = gensys.solvated_system(**sysargs)
    paths = gensys.write_system(s, str(tmp_path), name, output=1, temp_interval=1, **inpargs)
    rows, offd = g["eq_frames"], g["offd_frames"]
    assert rows.shape[0] == inpargs["steps"] - 1 and offd.shape[1] == sysargs.get("n_offdiag", 0)
    o = Oracle(paths["inp"])
    o.run(1)
    for k in range(1, inpargs["steps"]):
        od = o.offdiag()
        assert np.allclose(od[:, 0], offd[k - 1, :, 2], rtol=1e-13, atol=0), (k, od, offd[k - 1])
        assert np.allclose(od[:, 1], offd[k - 1, :, 3], rtol=1e-13, atol=0)
        o.run(1)
        # (nstates, 12): qq el/vdw, qp el/vdw, qw el/vdw, total, Q bond angle
        # torsion improper, restraint; put_ene columns: 0 total, 1-4 Q
        # bonded, 7-12 qq qp qw (el, vdw), 13 restraint
        q = o.qenergies()
        ref = rows[k - 1]
        assert np.allclose(q[:, :6], ref[:, 7:13], rtol=1e-12, atol=1e-10), (k, q[:, :6], ref[:, 7:13])
        assert np.allclose(q[:, 6], ref[:, 0], rtol=1e-12, atol=1e-10)
        assert np.allclose(q[:, 7:11], ref[:, 1:5], rtol=1e-12, atol=1e-10), (k, q[:, 7:11], ref[:, 1:5])
        assert np.allclose(q[:, 11], ref[:, 13], rtol=1e-12, atol=1e-10), (k, q[:, 11], ref[:, 13])
    assert np.array_equal(o.theta_corr(), g["theta_corr"])


MONITOR_CASES = sorted(k for k, (sa, _) in CASES.items() if sa.get("monitor"))


@pytest.mark.parametrize("name", MONITOR_CASES)
def test_oracle_monitor_matches_reference(name, tmp_path):
    """nonbond_monitor: the oracle's monitored-pair energies at every output
    step equal the reference's printed table (2 decimals; overflowed fields
    skipped)."""
    g = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    ref = g["monitor"]
    sysargs, inpargs = CASES[name]
    s = gensys.solvated_system(**sysargs)
    paths = gensys.write_system(s, str(tmp_path), name, output=1, temp_interval=1, **inpargs)
    o = Oracle(paths["inp"])
    steps = inpargs["steps"]
    for k in range(steps):
        m = np.round(o.monitor(), 2)
        ok = np.isnan(ref[k]) | (np.abs(m - ref[k]) <= 0.0100001)
        assert ok.all(), (k, m, ref[k])
        o.run(1)
    o.close()
