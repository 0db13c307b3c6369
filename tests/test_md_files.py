"""This package's writers of qdyn's output files, byte for byte against the
files reference qdyn wrote for the same run (tests/golden, case with
Q-bonded terms, restraints, a trajectory mask and an energy file).  The
oracle, pinned bitwise to reference qdyn by test_oracle_reference.py,
supplies the coordinates and energies, so every difference is a writer's:

* restart (write_xfin): canary, x, v, theta_corr records;
* DCD trajectory (trj_create/trj_write): header records (CORD block, title,
  topology name and mask rows, atom count) and the REAL(4) x/y/z records
  of the [trajectory_atoms] mask (ranges, NOT HEAVY, residue numbering), at
  the loop steps md_run writes them;
* energy file (put_ene): every record except the 80-character version
  string of the header.
Host-side code only: no device is used.
"""
import os

import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import fep, gensys, md

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAME = "c6_r12_q8s2_qbonded_30"
SYSARGS = dict(radius=12.0, n_solute=40, seed=2201, n_q=8, n_states=2, q_types=True, softcore=60.0,
               impropers=1, q_bonded=True, restraints=True)
INPARGS = dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0, bath_coupling=2.0,
               initial_temperature=50.0, energy="case.en", energy_interval=1, trajectory="case.dcd",
               trajectory_interval=10, trajectory_atoms=["1 20", "not heavy", "residue 2 5"])


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("files")
    s = gensys.solvated_system(**SYSARGS)
    paths = gensys.write_system(s, str(d), NAME, output=1, temp_interval=1, **INPARGS)
    g = np.load(os.path.join(GOLD, NAME + ".npz"), allow_pickle=False)
    o = Oracle(paths["inp"])
    frames_x = {}
    eq = []
    targets = md.trajectory_steps(30, 10)
    for k in range(1, 31):
        o.run(1)
        if k > 1:
            # the energy frame of loop step k-1 (from the coordinates it starts with)
            eq.append(o.qenergies())
        if k in targets:
            frames_x[k] = o.state()[0].copy()
    x, v = o.state()
    return dict(paths=paths, g=g, x=x, v=v, tc=o.theta_corr(), frames_x=frames_x, eq=eq,
                targets=targets, s=s, dir=str(d))


def test_trajectory_steps_follow_md_run():
    # loop steps 10, 20 (coordinates after steps 11, 21) and the final image
    assert md.trajectory_steps(30, 10) == [11, 21, 30]
    assert md.trajectory_steps(31, 10) == [11, 21, 31]
    assert md.trajectory_steps(3, 1) == [2, 3, 3]
    assert md.trajectory_steps(5, 0) == []


def test_restart_bytes(run):
    path = os.path.join(run["dir"], "ours.re")
    qe.write_restart(path, run["x"], run["v"], run["tc"])
    ours = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    assert np.array_equal(ours, run["g"]["re_bytes"])


def test_dcd_bytes(run):
    path = os.path.join(run["dir"], "ours.dcd")
    t = qe.Trajectory(run["paths"]["inp"], path)
    s = run["s"]
    # the mask: atoms 1-20, every non-heavy atom, residues 2-5 (waters 1-4)
    heavy = np.array([s.masses[s.iac[i] - 1] >= 4.0 for i in range(s.natom)])
    m = np.zeros(s.natom, bool)
    m[:20] = True
    m[~heavy] = True
    m[s.nat_solute:s.nat_solute + 12] = True
    assert t.natoms == int(m.sum())
    for k in run["targets"]:
        t.write(run["frames_x"][k])
    t.close()
    ours = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    ref = run["g"]["dcd_bytes"]
    assert ours.size == ref.size
    assert np.array_equal(ours, ref)


def test_energy_file_bytes(run):
    path = os.path.join(run["dir"], "ours.en")
    rows = []
    for q in run["eq"]:
        # put_ene: total, q(bond angle torsion improper), qx, qq, qp, qw, restraint
        qx_el, qx_vdw = q[:, 0] + q[:, 2] + q[:, 4], q[:, 1] + q[:, 3] + q[:, 5]
        rows.append(np.stack([q[:, 6], q[:, 7], q[:, 8], q[:, 9], q[:, 10], qx_el, qx_vdw,
                              q[:, 0], q[:, 1], q[:, 2], q[:, 3], q[:, 4], q[:, 5], q[:, 11]], axis=1))
    fep.write_energy_file(path, INPARGS["lambdas"], np.array(rows))
    ours = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    ref = run["g"]["en_bytes"]
    assert ours.size == ref.size
    # header record: marker, 7 integers, then the version string (80 bytes)
    keep = np.ones(ref.size, bool)
    keep[4 + 28:4 + 28 + 80] = False
    assert np.array_equal(ours[keep], ref[keep])
