"""The GPU MD run loop (qsource_amd.md.run) against the files reference qdyn
wrote for the same run (tests/golden): the FP64 engine's final restart
within 1e-9 of the reference's x and v, the same theta_corr record layout;
the DCD header byte-identical and every frame within REAL(4) rounding of the
reference's; the energy file's records within 1e-9 (FP64 values) with the
same record structure.  A second run chained from the first's restart
continues it (restart consistency: its step-0 energy terms equal the first
run's final ones).
"""
import os

import numpy as np
import pytest

from qsource_amd import engine as qe
from qsource_amd import fep, gensys, md

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAME = "c6_r12_q8s2_qbonded_30"
SYSARGS = dict(radius=12.0, n_solute=40, seed=2201, n_q=8, n_states=2, q_types=True, softcore=60.0,
               impropers=1, q_bonded=True, restraints=True)
INPARGS = dict(steps=30, lambdas=[0.6, 0.4], stepsize=0.5, temperature=50.0, bath_coupling=2.0,
               initial_temperature=50.0, energy="case.en", energy_interval=1, trajectory="case.dcd",
               trajectory_interval=10, trajectory_atoms=["1 20", "not heavy", "residue 2 5"])


def _records(b):
    out, p = [], 0
    while p < len(b):
        n = int(np.frombuffer(b[p:p + 4].tobytes(), dtype="<i4")[0])
        out.append(b[p + 4:p + 4 + n])
        p += 8 + n
    return out


def test_run_writes_qdyn_files(tmp_path):
    print("case: md.run FP64, Q-bonded + restraints + trajectory mask, 30 steps", flush=True)
    g = np.load(os.path.join(GOLD, NAME + ".npz"), allow_pickle=False)
    s = gensys.solvated_system(**SYSARGS)
    paths = gensys.write_system(s, str(tmp_path), NAME, output=10, **INPARGS)
    fin = md.run([paths["inp"]], qe.QGPU_FP64)
    x, v, tc = gensys.read_restart(paths["re"])
    xr, vr = g["x"].ravel(), g["v"].ravel()
    assert np.abs(x.ravel() - xr).max() <= 1e-9 * np.abs(xr).max()
    assert np.abs(v.ravel() - vr).max() <= 1e-9 * np.abs(vr).max()
    assert np.allclose(tc, g["theta_corr"], rtol=1e-9, atol=1e-12)
    # trajectory: header identical, frames within REAL(4) rounding
    ours = _records(np.frombuffer(open(os.path.join(str(tmp_path), "case.dcd"), "rb").read(), dtype=np.uint8))
    ref = _records(g["dcd_bytes"])
    assert len(ours) == len(ref) == 3 + 3 * 3
    for a, b in zip(ours[:3], ref[:3]):
        assert np.array_equal(a, b)
    for a, b in zip(ours[3:], ref[3:]):
        fa, fb = np.frombuffer(a.tobytes(), "<f4"), np.frombuffer(b.tobytes(), "<f4")
        assert np.allclose(fa, fb, rtol=2e-7, atol=2e-6)
    # energy file: same records, values within 1e-9
    _, rows, _ = fep.read_energy_frames(os.path.join(str(tmp_path), "case.en"), 2)
    assert rows.shape == g["eq_frames"].shape
    assert np.allclose(rows, g["eq_frames"], rtol=1e-9, atol=1e-9)
    # chained run from the restart: its first energy is the first run's final one
    p2 = gensys.write_system(s, str(tmp_path / "next"), NAME, output=10, restart=paths["re"],
                             **dict(INPARGS, steps=10, initial_temperature=None))
    e2 = qe.Engine([p2["inp"]], qe.QGPU_FP64)
    e2.make_pair_lists()
    e2.pot_energy()
    en = e2.energies(0)
    # (Q angles/torsions coupled to a bond absent in the earlier states read
    # the Morse factor the previous evaluation left, which a new run starts
    # at zero, as qdyn's module arrays do: those terms are not compared)
    for t in ("q_bond", "rst_protein", "rst_radial", "qx_el", "qx_vdw", "lrf", "pp_el", "ww_el", "pw_vdw"):
        assert abs(en[t] - fin[0][t]) <= 1e-9 * max(1.0, abs(fin[0][t])), (t, en[t], fin[0][t])
    e2.close()


WO_NAME = "c13_r12_q8s4_writeout_12"
WO_SYSARGS = dict(radius=12.0, n_solute=40, seed=3101, n_q=8, n_states=4, q_types=True, softcore=20.0,
                  n_offdiag=3, impropers=1, monitor=True)
WO_INPARGS = dict(steps=12, lambdas=[0.4, 0.3, 0.2, 0.1], stepsize=0.5, temperature=50.0, bath_coupling=2.0,
                  initial_temperature=50.0)


def _numbers(block):
    import re
    return [float(t) for t in re.findall(r"-?\d+\.\d+", block)]


def test_run_prints_energy_summaries(tmp_path):
    """The run loop's printed summaries (FP64 engine) against the blocks
    reference qdyn printed for the same run: the same lines, and every
    printed number within 0.01 of the reference's (the engine's energies
    agree with the reference's to 1e-9, so only a value on a rounding
    boundary of the last digit can print differently)."""
    import io
    print("case: md.run FP64 printed summaries, 4 states + off-diagonals + monitored pairs, 12 steps", flush=True)
    g = np.load(os.path.join(GOLD, WO_NAME + ".npz"), allow_pickle=False)
    ref = [str(b) for b in g["blocks"]]
    s = gensys.solvated_system(**WO_SYSARGS)
    paths = gensys.write_system(s, str(tmp_path), WO_NAME, output=1, temp_interval=1, **WO_INPARGS)
    log = io.StringIO()
    md.run([paths["inp"]], qe.QGPU_FP64, logs=[log])
    text = log.getvalue()
    ours = []
    cur = None
    for line in text.splitlines():
        if "Energy summary" in line and line.startswith("="):
            cur = [line]
            continue
        if cur is not None:
            cur.append(line)
            if line == "=" * 79:
                ours.append("\n".join(cur))
                cur = None
    assert len(ours) == len(ref) == 13
    for k, (a, b) in enumerate(zip(ours, ref)):
        la, lb = a.splitlines(), b.splitlines()
        assert len(la) == len(lb), k
        for x, y in zip(la, lb):
            assert len(x) == len(y) and x.split()[:1] == y.split()[:1], (k, x, y)
        na, nb = _numbers(a), _numbers(b)
        assert len(na) == len(nb) and np.allclose(na, nb, rtol=0, atol=0.0100001), (k, a, b)
