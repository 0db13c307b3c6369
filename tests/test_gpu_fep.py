"""GPU parity of the EVB-style FEP features against the FP64 oracle (which
test_oracle_reference.py pins bitwise to reference qdyn on the same
systems): per-state Q-atom types (qvdw), soft-core, off-diagonal couplings,
harmonic (GROMOS) and periodic (AMBER) impropers.

Bars: step 0 energies, per-state Q energies and gradient within 1e-12
relative; 30 FP64 steps within 1e-9; energy-file frames (per-state Q
energies and off-diagonal Hij, r_kl) within 1e-9 of the oracle on the same
steps; mixed precision gradient RMS error below 1e-5 of the gradient RMS.
"""
import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

COOL = dict(stepsize=0.5, temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)
TERMS = ["pp_el", "pp_vdw", "p_bond", "p_angle", "p_torsion", "p_improper", "pw_el", "pw_vdw",
         "ww_el", "ww_vdw", "lrf", "qx_el", "qx_vdw", "rst_radial", "rst_wpol", "potential"]


def _say(msg):
    print(msg, flush=True)


def _system(tmp_path, impropers, seed):
    s = gensys.solvated_system(radius=12.0, n_solute=40, seed=seed, n_q=8, n_states=4, q_types=True,
                               softcore=20.0, n_offdiag=3, impropers=impropers)
    return gensys.write_system(s, str(tmp_path), "evb%d" % impropers, steps=30, output=10,
                               lambdas=[0.4, 0.3, 0.2, 0.1], **COOL)


@pytest.mark.parametrize("impropers,seed", [(1, 2101), (2, 2102)])
def test_evb_step0_and_30_steps_fp64(tmp_path, impropers, seed):
    _say("case: 4-state Q types + softcore + offdiag, impropers=%d, R=12, FP64" % impropers)
    paths = _system(tmp_path, impropers, seed)
    o = Oracle(paths["inp"])
    o.final_energy()
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    eo, eg = o.energies(), e.energies(0)
    assert abs(eo["p_improper"]) > 0.1
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    qg = np.array(e.qenergies(0)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-12, atol=1e-11), (qg, qo)
    do, dg = o.gradient(), e.gradient(0)
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    e.close()
    # 30 steps from the start
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    e.close()


def test_evb_energy_frames_and_offdiag(tmp_path):
    """Frames recorded on the device every step (from step 1) against the
    oracle's per-state Q energies and off-diagonal elements on the same
    steps."""
    _say("case: energy frames + off-diagonals, R=12 EVB system, FP64, 10 steps")
    paths = _system(tmp_path, 1, 2101)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.set_energy_interval(1, 16)
    e.step(10)
    fr = e.energy_series(0)
    od = e.offdiag_series(0)
    assert fr.shape[0] == 9 and od.shape == (9, 3, 2)
    o = Oracle(paths["inp"])
    o.run(1)
    for k in range(1, 10):
        ro = o.offdiag()
        assert np.allclose(od[k - 1], ro, rtol=1e-9, atol=0), (k, od[k - 1], ro)
        o.run(1)
        q = o.qenergies()
        assert np.allclose(fr[k - 1, :, 7:13], q[:, :6], rtol=1e-9, atol=1e-9), (k, fr[k - 1, :, 7:13], q[:, :6])
        assert np.allclose(fr[k - 1, :, 0], q[:, 6], rtol=1e-9, atol=1e-9)
    e.close()


def test_evb_mixed_precision(tmp_path):
    _say("case: EVB system, mixed precision step 0")
    paths = _system(tmp_path, 2, 2102)
    o = Oracle(paths["inp"])
    o.final_energy()
    e = qe.Engine([paths["inp"]], qe.QGPU_MIXED)
    e.make_pair_lists()
    e.pot_energy()
    do, dg = o.gradient(), e.gradient(0)
    rms = np.sqrt(np.mean(do ** 2))
    assert np.sqrt(np.mean((dg - do) ** 2)) <= 1e-5 * rms
    qo = o.qenergies()
    qg = np.array(e.qenergies(0)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-5, atol=1e-4), (qg, qo)
    e.close()
