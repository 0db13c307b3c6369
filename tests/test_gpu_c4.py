"""GPU parity and the learned far field at the far-field workload's full size.

The system (R = 30 A TIP3P sphere around a 2000-atom branched solute, 12 Q
atoms, 2 FEP states) is built, equilibrated on the GPU and run from its
restart at production settings (1 fs, 300 K):

* one lambda window as an FP64 instance of a batch of three: x, v, every
  energy term and the per-state Q energies within 1e-9 relative of the
  oracle after 30 steps;
* the three-window mixed-precision batch on the same coordinates: gradient
  RMS error at most 1e-5 of the gradient RMS, and per-state Q energies
  within 1e-5 relative;
* the far-field treatments against exact all-pairs evaluation (every
  cut-off beyond the sphere) on held-out frames, lists built five steps
  before the evaluated state: LRF's gradient error is below the plain
  cut-off's, and a surrogate trained on the GPU on exact far-field targets
  of the same system predicts the far potential at the charge-group centres
  with a smaller RMS error than the plain cut-off (which takes it as zero).
"""
import os

import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import equilibrate, farfield, gensys

pytestmark = pytest.mark.gpu

TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "p_bond", "p_angle",
         "p_torsion", "qx_el", "qx_vdw", "rst_radial", "rst_wpol", "potential", "kinetic"]
LAMS = [0.0, 0.5, 1.0]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c4(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c4"))
    _say("building the R=30 A / 2000-atom solute system")
    s = gensys.solvated_system(radius=30.0, n_solute=2000, seed=4000, n_q=12, n_states=2)
    _say("equilibrating on the GPU (%d atoms, %d waters)" % (s.natom, s.nwat))
    rst = equilibrate.equilibrate(s, d, "c4", lambdas=[0.5, 0.5], seed=31, progress=_say)
    inputs = []
    for k, lam in enumerate(LAMS):
        p = gensys.write_system(s, d, "c4_w%d" % k, steps=1000, output=1000, lambdas=[1.0 - lam, lam],
                                initial_temperature=300.0, random_seed=301 + k, restart=os.path.basename(rst))
        inputs.append(p["inp"])
    return {"inputs": inputs, "system": s, "restart": rst, "dir": d}


def test_c4_fp64_window_vs_oracle(c4):
    _say("case: R=30 / 2000-atom solute, FP64 batch of 3 windows, lambda 0.5 vs oracle, 30 steps")
    inputs = c4["inputs"]
    e = qe.Engine(inputs, qe.QGPU_FP64)
    e.step(30)
    k = 1
    o = Oracle(inputs[k])
    o.run(30)
    xo, vo = o.state()
    xg, vg = e.state(k)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    eo, eg = o.energies(), e.energies(k)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    qg = np.array(e.qenergies(k)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-9, atol=1e-9), (qg, qo)
    o.close()
    e.close()


def test_c4_mixed_batch_gradient(c4):
    _say("case: R=30, 3-window mixed batch vs FP64 batch on identical coordinates")
    inputs = c4["inputs"]
    ef = qe.Engine(inputs, qe.QGPU_FP64)
    ef.step(20)
    states = [ef.state(k) for k in range(len(inputs))]
    em = qe.Engine(inputs, qe.QGPU_MIXED)
    for k, (x, v) in enumerate(states):
        em.set_state(k, x, v)
    for e in (ef, em):
        e.make_pair_lists()
        e.pot_energy()
    for k in range(len(inputs)):
        d64, d32 = ef.gradient(k), em.gradient(k)
        rms = np.sqrt(np.mean(d64 ** 2))
        err = np.sqrt(np.mean((d32 - d64) ** 2))
        _say("  window %d: gradient RMS error %.3e of RMS %.3e" % (k, err, rms))
        assert err <= 1e-5 * rms, k
        q64 = np.array(ef.qenergies(k)).reshape(-1, 12)
        q32 = np.array(em.qenergies(k)).reshape(-1, 12)
        assert np.allclose(q32, q64, rtol=1e-5, atol=1e-4), (k, q32, q64)
    ef.close()
    em.close()


def test_c4_far_field_treatments_vs_exact(c4):
    _say("case: R=30, surrogate trained on 30 FP64 frames; LRF / surrogate / cut-off vs exact on 3 frames")
    Xs, Ys, _ = farfield.sample_frames(c4["inputs"][1], 30, 25)
    model, loss = farfield.train_surrogate(np.concatenate(Xs), np.concatenate(Ys), epochs=20, seed=5)
    assert np.isfinite(loss)
    tin = farfield.treatment_inputs(c4["system"], c4["dir"], "c4_ff", c4["restart"], [0.5, 0.5], seed=777)
    frames = farfield.held_out_frames(tin["lrf"], nframes=3, stride=60, gap=5)
    fe = farfield.force_errors(tin, frames, model)
    model.close()
    _say("  gradient RMS error vs exact: lrf %.4f, surrogate %.4f, cut-off %.4f (exact RMS %.3f)" % (
        fe["lrf"]["gradient_rms_err"], fe["surrogate"]["gradient_rms_err"], fe["cutoff"]["gradient_rms_err"],
        fe["gradient_rms_exact"]))
    cf = fe["centre_field"]
    _say("  centre potential RMS error: surrogate %.4f, lrf %.4f, cut-off %.4f" % (
        cf["potential_rms_err_surrogate"], cf["potential_rms_err_lrf"], cf["potential_rms_err_cutoff"]))
    assert fe["gradient_rms_exact"] > 0
    assert fe["lrf"]["gradient_rms_err"] < fe["cutoff"]["gradient_rms_err"]
    assert cf["potential_rms_err_surrogate"] < cf["potential_rms_err_cutoff"]
    assert cf["potential_rms_err_lrf"] < cf["potential_rms_err_cutoff"]
