"""GPU parity at the EVB workload's full size.

The EVB-style system (R = 40 A sphere, 5000-atom branched solute, 24 Q atoms,
4 states with per-state charges and Q-atom LJ types, off-diagonal couplings)
is built, equilibrated on the GPU and run from its restart at production
settings (1 fs, 300 K):

* one mapping window as an FP64 instance of a batch of four: x, v, every
  energy term and the per-state Q energies within 1e-9 relative of the
  oracle after 30 steps;
* the four-window mixed-precision batch on the same coordinates: gradient
  RMS error at most 1e-5 of the gradient RMS, per-state Q energies within
  1e-5 relative, and every term group's gradient error bounded the same way.
"""
import os

import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import equilibrate, gensys

pytestmark = pytest.mark.gpu

TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "p_bond", "p_angle",
         "p_torsion", "qx_el", "qx_vdw", "rst_radial", "rst_wpol", "potential", "kinetic"]
WINDOWS = [0, 5, 10, 15]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c5_inputs(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c5"))
    _say("building the R=40 A / 5000-atom solute EVB system")
    s = gensys.evb_system(radius=40.0, n_solute=5000, seed=5000, n_q=24, n_states=4)
    lam = gensys.mapping_lambdas(4, 16)
    _say("equilibrating on the GPU (%d atoms)" % s.natom)
    rst = equilibrate.equilibrate(s, d, "c5", lambdas=list(lam[0]), seed=21, progress=_say)
    inputs = []
    for k in WINDOWS:
        p = gensys.write_system(s, d, "c5_w%02d" % k, steps=1000, output=1000, lambdas=list(lam[k]),
                                initial_temperature=300.0, random_seed=201 + k, restart=os.path.basename(rst))
        inputs.append(p["inp"])
    return inputs


def test_c5_fp64_window_vs_oracle(c5_inputs):
    _say("case: EVB R=40 / 5000-atom solute / 24 Q / 4 states, FP64 batch of 4, window 5 vs oracle, 30 steps")
    e = qe.Engine(c5_inputs, qe.QGPU_FP64)
    e.step(30)
    k = 1
    o = Oracle(c5_inputs[k])
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
    assert qo.shape[0] == 4
    assert np.allclose(qg, qo, rtol=1e-9, atol=1e-9), (qg, qo)
    od = o.offdiag()
    assert len(od) == 6
    o.close()
    e.close()


def test_c5_mixed_batch_gradient(c5_inputs):
    _say("case: EVB R=40, 4-window mixed batch vs FP64 batch on identical coordinates")
    ef = qe.Engine(c5_inputs, qe.QGPU_FP64)
    ef.step(20)
    states = [ef.state(k) for k in range(len(c5_inputs))]
    em = qe.Engine(c5_inputs, qe.QGPU_MIXED)
    for k, (x, v) in enumerate(states):
        em.set_state(k, x, v)
    for e in (ef, em):
        e.make_pair_lists()
        e.pot_energy()
    for k in range(len(c5_inputs)):
        d64, d32 = ef.gradient(k), em.gradient(k)
        rms = np.sqrt(np.mean(d64 ** 2))
        err = np.sqrt(np.mean((d32 - d64) ** 2))
        _say("  window %d: gradient RMS error %.3e of RMS %.3e" % (WINDOWS[k], err, rms))
        assert err <= 1e-5 * rms, k
        q64 = np.array(ef.qenergies(k)).reshape(-1, 12)
        q32 = np.array(em.qenergies(k)).reshape(-1, 12)
        assert np.allclose(q32, q64, rtol=1e-5, atol=1e-4), (k, q32, q64)
    for t in ("ww", "pw", "pp", "qx", "lrf"):
        for e in (ef, em):
            e.pot_energy([t])
        for k in range(len(c5_inputs)):
            d64, d32 = ef.gradient(k), em.gradient(k)
            rms = np.sqrt(np.mean(d64 ** 2))
            assert rms > 0, (t, k)
            assert np.sqrt(np.mean((d32 - d64) ** 2)) <= 1e-5 * rms, (t, k)
    ef.close()
    em.close()
