"""Writes the golden fixtures of reference qdyn runs used by the CPU tests.

Runs the reference's serial qdyn5 (built from the reference sources by
oracle/Makefile.ref into oracle/_ref/) on systems written by
qsource_amd.gensys, and stores, per case:

* ``energies``: the write_out energy summary of every step, as printed by
  qdyn (F12.2): ww_el ww_vdw lrf rst_total rst_fix rst_radial rst_wpol
  total potential kinetic, plus pp/pw/Q terms where present;
* ``x``, ``v``: the final restart coordinates and velocities (FP64);
* ``theta_corr``: the water-polarisation shell corrections of the restart;
* ``top_sha``: SHA-256 of the generated topology, so a change of the
  generator is detected instead of silently comparing different systems;
* ``eq_frames``, ``offd_frames`` (cases that write an energy file): the
  reference's energy-file frames (per-state Q energies in FP64 and the
  off-diagonal elements) of every recorded step;
* ``re_bytes``, ``en_bytes``, ``dcd_bytes`` (cases that write a
  trajectory): the reference's restart, energy and trajectory files as
  written, for byte-exact checks of this package's writers.

and, for the reference's own basic test system (prep_SPH, basic_tests.py),
copies of its data files under prep_SPH/ and ``prep_sph_eq1.npz``: the
summary blocks qdyn5 printed for eq1 (leap-frog, Berendsen, sphere) and
its final restart.

Usage:  python tests/golden/make_golden.py
"""
import hashlib
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from qsource_amd import fep, gensys  # noqa: E402
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import basic_tests as bt  # noqa: E402

QDYN = os.path.join(ROOT, "oracle", "_ref", "qdyn5")
OUT = os.path.dirname(os.path.abspath(__file__))

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
    # name: (builder kwargs, input kwargs)
    "c1_r15_100": (dict(radius=15.0, n_solute=0, seed=1001), dict(steps=100)),
    "c1_r10_250": (dict(radius=10.0, n_solute=0, seed=1001), dict(steps=250)),
    "c2_r14_q8_30": (dict(radius=14.0, n_solute=40, seed=2001, n_q=8, n_states=2),
                     dict(steps=30, lambdas=[0.7, 0.3], stepsize=0.5,
                          temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)),
    # plain cut-off: pair lists without LRF multipoles, no Taylor term
    "c2_r14_q8_nolrf_30": (dict(radius=14.0, n_solute=40, seed=2003, n_q=8, n_states=2),
                           dict(steps=30, lambdas=[0.5, 0.5], stepsize=0.5, lrf=False,
                                temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)),
    # EVB-style Q region: 4 states with per-state Q-atom types (qvdw),
    # soft-core alphas and off-diagonal couplings; harmonic (GROMOS) or
    # periodic (AMBER) impropers; energy file written every step
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
    # all-atom solvent LJ (topology solvent type 1, TIP3P-like H types):
    # LJ between every atom pair of two waters and between the Q atoms and
    # every water atom, with fixed Q-atom LJ parameters or Q-atom types
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

def parse_volume_moves(log: str):
    """MC_volume's verdicts in the order printed (src/md.f90:19189-19196):
    1 for "Volume change accepted", 0 for "Volume change rejected"."""
    out = []
    for line in log.splitlines():
        if line.startswith("Volume change accepted"):
            out.append(1)
        elif line.startswith("Volume change rejected"):
            out.append(0)
    return np.array(out, dtype=np.int32)


def volume_summary_block(log: str) -> str:
    """The volume-change summary of the final write_out (src/md.f90:
    18678-18690), from its heading to its closing line of '='."""
    lines = log.splitlines()
    i = next(k for k, l in enumerate(lines) if "VOLUME CHANGE SUMMARY" in l)
    j = next(k for k in range(i + 1, len(lines)) if lines[k] == "=" * 79)
    return "\n".join(lines[i:j + 1])


def parse_monitor(log: str):
    """The monitored-pair table of every energy summary (write_out's formats
    i2, f10.2, f8.2, f9.2, then 2f8.2 per state; overflowed fields are NaN)."""
    lines = log.splitlines()
    out = []
    for k, l in enumerate(lines):
        if "Monitoring selected groups" not in l:
            continue
        rows, j = [], k + 2
        while not lines[j].startswith("====="):
            t = lines[j]
            f = [t[2:12], t[12:20], t[20:29]] + [t[29 + 8 * q:37 + 8 * q] for q in range((len(t) - 29) // 8)]
            rows.append([float(x) if "*" not in x else float("nan") for x in f])
            j += 1
        out.append(rows)
    return np.array(out)


ROW = re.compile(r"^(solute|solvent|solute-solvent|LRF|Q-atom|restraints|SUM)\s+(.*)$")


def parse_summaries(log: str):
    """Energy summaries of every step: dict per step."""
    out = []
    for blk in re.split(r"=+\s+Energy summary at step\s+\d+\s+=+|=+\s+FINAL\s+Energy summary\s+=+", log)[1:]:
        rows = {}
        for line in blk.splitlines():
            m = ROW.match(line.strip())
            if m and m.group(1) not in rows:
                rows[m.group(1)] = [float(t) for t in m.group(2).split()]
            if line.startswith("===") and rows:
                break
        out.append(rows)
    return out


def flatten(rows):
    sv = rows.get("solvent", [0] * 6)
    so = rows.get("solute", [0] * 6)
    pw = rows.get("solute-solvent", [0, 0])
    lrf = rows.get("LRF", [0])[0]
    q = rows.get("Q-atom", [0] * 6)
    r = rows["restraints"]
    s = rows["SUM"]
    return [sv[0], sv[1], lrf, r[0], r[1], r[2], r[3], s[0], s[1], s[2],
            so[0], so[1], so[2], so[3], so[4], so[5], pw[0], pw[1], q[0], q[1], r[4], r[5],
            q[2], q[3], q[4], q[5]]


COLUMNS = ["ww_el", "ww_vdw", "lrf", "rst_total", "rst_fix", "rst_radial", "rst_wpol",
           "total", "potential", "kinetic", "pp_el", "pp_vdw", "p_bond", "p_angle",
           "p_torsion", "p_improper", "pw_el", "pw_vdw", "qx_el", "qx_vdw",
           "rst_shell", "rst_protein", "q_bond", "q_angle", "q_torsion", "q_improper"]


def build_case(name, sysargs, inpargs, workdir):
    s = gensys.solvated_system(**sysargs)
    paths = gensys.write_system(s, workdir, name, output=1, temp_interval=1, **inpargs)
    return s, paths


# Cases whose printed energy summaries (write_out, src/md.f90:18545-18691)
# are kept as text: every block from its header to its closing line.
WRITE_OUT_CASES = {
    "c13_r12_q8s4_writeout_12": (dict(radius=12.0, n_solute=40, seed=3101, n_q=8, n_states=4, q_types=True,
                                      softcore=20.0, n_offdiag=3, impropers=1, monitor=True),
                                 dict(steps=12, lambdas=[0.4, 0.3, 0.2, 0.1], stepsize=0.5, temperature=50.0,
                                      bath_coupling=2.0, initial_temperature=50.0)),
}


def summary_blocks(log: str):
    """The text of every write_out block: from an 'Energy summary' header
    line to the 79-character closing line of '='."""
    lines = log.splitlines()
    out, cur = [], None
    for l in lines:
        if re.match(r"^=+ (FINAL )?\s*Energy summary", l):
            cur = [l]
            continue
        if cur is not None:
            cur.append(l)
            if l == "=" * 79:
                out.append("\n".join(cur))
                cur = None
    return out


def make_write_out_cases(only):
    for name, (sysargs, inpargs) in WRITE_OUT_CASES.items():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            s, paths = build_case(name, sysargs, inpargs, d)
            r = subprocess.run([QDYN, name + ".inp"], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            if "terminated normally" not in r.stdout:
                sys.exit("reference run failed for %s:\n%s" % (name, r.stdout[-3000:]))
            blocks = summary_blocks(r.stdout)
            sha = hashlib.sha256(open(paths["top"], "rb").read()).hexdigest()
            np.savez_compressed(os.path.join(OUT, name + ".npz"), blocks=np.array(blocks), top_sha=np.array(sha))
            print(name, len(blocks), "blocks", sha[:12])


def make_prep_sph(only):
    """Copies the basic test system's data files beside basic_tests.py and
    runs reference qdyn5 on eq1 as run_test.sh prepares it for the sphere
    with leap-frog and Berendsen: keeps every printed summary block and the
    final restart."""
    name = "prep_sph_eq1"
    if only and name not in only:
        return
    if not bt.reference_present():
        sys.exit("reference basic tests not found under " + bt.REFERENCE)
    os.makedirs(bt.FIXTURE, exist_ok=True)
    shutil.copyfile(os.path.join(bt.REFERENCE, "eq1.inp"), os.path.join(bt.FIXTURE, "eq1.inp"))
    for f in ("lig_w.top", "lig_w.fep"):
        shutil.copyfile(os.path.join(bt.REFERENCE, "prep_SPH", f), os.path.join(bt.FIXTURE, f))
    with tempfile.TemporaryDirectory() as d:
        inp = bt.sph_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE, d)
        r = subprocess.run([QDYN, inp], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if "terminated normally" not in r.stdout:
            sys.exit("reference run failed for %s:\n%s" % (name, r.stdout[-3000:]))
        blocks = summary_blocks(r.stdout)
        x, v, tc = gensys.read_restart(os.path.join(d, "eq1.re"))
        sha = hashlib.sha256(open(os.path.join(bt.FIXTURE, "lig_w.top"), "rb").read()).hexdigest()
        np.savez_compressed(os.path.join(OUT, name + ".npz"), blocks=np.array(blocks), x=x, v=v,
                            theta_corr=tc, top_sha=np.array(sha))
        print(name, len(blocks), "blocks", x.shape, sha[:12])


def make_prep_pbc(only):
    """The periodic-box basic test system (prep_PBC): copies its topology and
    FEP file beside basic_tests.py and runs reference qdyn5 on eq1 as
    run_test.sh prepares it for the box with leap-frog and Berendsen; keeps
    every printed summary block, SHA-256 digests of the final x and v bytes
    (FP64, bit-exact pins without storing 12 262 x 6 doubles) and the box."""
    name = "prep_pbc_eq1"
    if only and name not in only:
        return
    if not bt.reference_present():
        sys.exit("reference basic tests not found under " + bt.REFERENCE)
    os.makedirs(bt.FIXTURE_PBC, exist_ok=True)
    for f in ("lig_w.top", "lig_w.fep"):
        shutil.copyfile(os.path.join(bt.REFERENCE, "prep_PBC", f), os.path.join(bt.FIXTURE_PBC, f))
    with tempfile.TemporaryDirectory() as d:
        inp = bt.pbc_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE_PBC, d)
        r = subprocess.run([QDYN, inp], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if "terminated normally" not in r.stdout:
            sys.exit("reference run failed for %s:\n%s" % (name, r.stdout[-3000:]))
        blocks = summary_blocks(r.stdout)
        x, v, _ = gensys.read_restart(os.path.join(d, "eq1.re"))
        box = gensys.read_restart_box(os.path.join(d, "eq1.re"))
        sha = hashlib.sha256(open(os.path.join(bt.FIXTURE_PBC, "lig_w.top"), "rb").read()).hexdigest()
        digest = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
        np.savez_compressed(os.path.join(OUT, name + ".npz"), blocks=np.array(blocks), box=box,
                            x_sha=np.array(digest(x)), v_sha=np.array(digest(v)), top_sha=np.array(sha))
        print(name, len(blocks), "blocks", x.shape, sha[:12])


def make_test5_npt(only):
    """The reference's constant-pressure test (tests/test5: the prep_PBC
    system with its own FEP file and eq1.inp), shortened to 100 steps with
    lists and volume moves every 10 steps; copies eq1.inp and the FEP file
    beside basic_tests.py and keeps the printed summary blocks, the verdict
    of every volume move, SHA-256 digests of the final x and v bytes and the
    final box."""
    name = "test5_npt_eq1"
    if only and name not in only:
        return
    os.makedirs(bt.FIXTURE_NPT, exist_ok=True)
    shutil.copyfile(os.path.join(bt.REFERENCE_NPT, "eq1.inp"), os.path.join(bt.FIXTURE_NPT, "eq1.inp"))
    shutil.copyfile(os.path.join(bt.REFERENCE_NPT, "prep", "lig_w.fep"), os.path.join(bt.FIXTURE_NPT, "lig_w.fep"))
    with tempfile.TemporaryDirectory() as d:
        inp = bt.npt_input(d, **bt.NPT_SHORT)
        r = subprocess.run([QDYN, inp], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if "terminated normally" not in r.stdout:
            sys.exit("reference run failed for %s:\n%s" % (name, r.stdout[-3000:]))
        blocks = summary_blocks(r.stdout)
        x, v, _ = gensys.read_restart(os.path.join(d, "eq1.re"))
        box = gensys.read_restart_box(os.path.join(d, "eq1.re"))
        digest = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
        np.savez_compressed(os.path.join(OUT, name + ".npz"), blocks=np.array(blocks), box=box,
                            volume_moves=parse_volume_moves(r.stdout),
                            volume_summary=np.array(volume_summary_block(r.stdout)),
                            x_sha=np.array(digest(x)), v_sha=np.array(digest(v)))
        print(name, len(blocks), "blocks", parse_volume_moves(r.stdout))


def make_sph_thermostats(only):
    """The reference's basic sphere test (eq1 of tests/basic_tests) with each
    integrator and thermostat of its test matrix (run_test.sh:193-200)
    besides leap-frog/Berendsen (prep_sph_eq1): every printed summary block
    and SHA-256 digests of the final x and v bytes, per combination."""
    name = "sph_eq1_thermostats"
    if only and name not in only:
        return
    out = {}
    for integ in ("leap-frog", "velocity-verlet"):
        for th in ("berendsen", "langevin", "nose-hoover"):
            if (integ, th) == ("leap-frog", "berendsen"):
                continue
            key = "%s_%s" % (integ.replace("-", "_"), th.replace("-", "_"))
            with tempfile.TemporaryDirectory() as d:
                inp = bt.sph_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE, d, integrator=integ,
                                   thermostat=th)
                r = subprocess.run([QDYN, inp], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
                if "terminated normally" not in r.stdout:
                    sys.exit("reference run failed for %s %s:\n%s" % (name, key, r.stdout[-3000:]))
                x, v, _ = gensys.read_restart(os.path.join(d, "eq1.re"))
                digest = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
                out[key + "_blocks"] = np.array(summary_blocks(r.stdout))
                out[key + "_x_sha"] = np.array(digest(x))
                out[key + "_v_sha"] = np.array(digest(v))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    print(name, sorted(k[:-7] for k in out if k.endswith("_blocks")))


QFEP = os.path.join(ROOT, "oracle", "_ref", "qfep5")
QFEP_KT = 0.596
QFEP_LAMBDAS = [0.0, 0.25, 0.5, 0.75, 1.0]


def make_qfep(only):
    """Five lambda windows of a small two-state system run by reference
    qdyn5 with an energy frame every step (the engine writes these files
    byte for byte the same), then reference qfep5 on them: keeps the energy
    files and qfep's Part 1 table (lambda, dGf, sum dGf, dGr, sum dGr, <dG>
    per file)."""
    name = "qfep_5win"
    if only and name not in only:
        return
    if not os.path.exists(QFEP):
        sys.exit("reference qfep5 not built: make -f oracle/Makefile.ref OUT=oracle/_ref")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        s = gensys.solvated_system(radius=12.0, n_solute=40, seed=2401, n_q=8, n_states=2)
        files = []
        for k, lam in enumerate(QFEP_LAMBDAS):
            p = gensys.write_system(s, d, "w%d" % k, steps=60, stepsize=0.5, temperature=50.0, bath_coupling=2.0,
                                    initial_temperature=50.0, lambdas=[1 - lam, lam], energy="w%d.en" % k,
                                    energy_interval=1)
            r = subprocess.run([QDYN, p["inp"]], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            if "terminated normally" not in r.stdout:
                sys.exit("reference run failed for %s window %d" % (name, k))
            files.append("w%d.en" % k)
            out["en%d" % k] = np.frombuffer(open(os.path.join(d, files[-1]), "rb").read(), dtype=np.uint8)
        # qfep's prompts: files; states, off-diagonals; kT, points to skip;
        # gap bins; points per bin; alpha of state 2; off-diagonal count;
        # reaction coordinate; file names
        stdin = "%d\n2 0\n%.3f 0\n1\n1\n0.0\n0\n1 -1\n%s\n" % (len(files), QFEP_KT, "\n".join(files))
        r = subprocess.run([QFEP], cwd=d, input=stdin, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        lines = r.stdout.splitlines()
        i = next(k for k, l in enumerate(lines) if l.startswith("# lambda(1)      dGf"))
        table = [[float(t) for t in lines[i + 1 + k].split()] for k in range(len(files))]
    np.savez_compressed(os.path.join(OUT, name + ".npz"), part1=np.array(table), kT=np.array(QFEP_KT),
                        lambdas=np.array(QFEP_LAMBDAS), **out)
    print(name, np.array(table)[:, 5])


def main():
    if not os.path.exists(QDYN):
        sys.exit("reference qdyn5 not built: make -f oracle/Makefile.ref OUT=oracle/_ref")
    only = set(sys.argv[1:])   # case names to (re)make; all when none given
    make_write_out_cases(only)
    make_prep_sph(only)
    make_prep_pbc(only)
    make_test5_npt(only)
    make_sph_thermostats(only)
    make_qfep(only)
    for name, (sysargs, inpargs) in CASES.items():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            s, paths = build_case(name, sysargs, inpargs, d)
            r = subprocess.run([QDYN, name + ".inp"], cwd=d, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            if "terminated normally" not in r.stdout:
                sys.exit("reference run failed for %s:\n%s" % (name, r.stdout[-3000:]))
            summ = parse_summaries(r.stdout)
            en = np.array([flatten(x) for x in summ])
            x, v, tc = gensys.read_restart(paths["re"])
            sha = hashlib.sha256(open(paths["top"], "rb").read()).hexdigest()
            extra = {}
            if inpargs.get("energy"):
                # the reference's own energy file: per-state Q energies (FP64)
                # and off-diagonal elements of every recorded step
                _, rows, offd = fep.read_energy_frames(os.path.join(d, inpargs["energy"]),
                                                       len(inpargs["lambdas"]))
                extra = dict(eq_frames=rows, offd_frames=offd)
            if sysargs.get("monitor"):
                extra["monitor"] = parse_monitor(r.stdout)
            if inpargs.get("pbc", {}).get("constant_pressure"):
                extra["volume_moves"] = parse_volume_moves(r.stdout)
                extra["volume_summary"] = np.array(volume_summary_block(r.stdout))
            box = gensys.read_restart_box(paths["re"])
            if box is not None:
                extra["box"] = box
            if inpargs.get("trajectory"):
                raw = lambda f: np.frombuffer(open(os.path.join(d, f), "rb").read(), dtype=np.uint8)
                extra.update(re_bytes=raw(name + ".re"), en_bytes=raw(inpargs["energy"]),
                             dcd_bytes=raw(inpargs["trajectory"]))
            np.savez_compressed(os.path.join(OUT, name + ".npz"), energies=en, columns=np.array(COLUMNS),
                                x=x, v=v, theta_corr=tc if tc is not None else np.zeros(0),
                                top_sha=np.array(sha), **extra)
            print(name, en.shape, x.shape, sha[:12])


if __name__ == "__main__":
    main()
