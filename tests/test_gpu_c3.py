"""GPU parity at the batched-FEP workload's full system sizes.

The three systems of the batched FEP benchmark (TIP3P spheres of R = 18, 22
and 28 A around branched solutes of 150, 300 and 600 atoms, 12 Q atoms, 2
states) are equilibrated on the GPU and run from their restarts at
production settings (1 fs, 300 K), two lambda windows each, as one batch of
six uneven instances:

* the FP64 batch after 30 steps: x, v, every energy term and the per-state
  Q energies within 1e-9 relative of the oracle, run per window;
* the mixed-precision batch on the same coordinates: gradient RMS error at
  most 1e-5 of the gradient RMS, per-state Q energies within 1e-5 relative,
  and every term group's gradient error bounded the same way.
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
SYSTEMS = [(18.0, 150), (22.0, 300), (28.0, 600)]
LAMBDAS = [0.25, 0.8]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c3_inputs(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c3"))
    inputs = []
    for si, (R, ns) in enumerate(SYSTEMS):
        _say("building and equilibrating R=%.0f A / %d-atom solute on the GPU" % (R, ns))
        s = gensys.solvated_system(radius=R, n_solute=ns, seed=3000 + si, n_q=12, n_states=2)
        rst = equilibrate.equilibrate(s, d, "s%d" % si, lambdas=[0.5, 0.5], seed=11 + si)
        for k, lam in enumerate(LAMBDAS):
            p = gensys.write_system(s, d, "s%d_w%d" % (si, k), steps=1000, output=1000, lambdas=[1.0 - lam, lam],
                                    initial_temperature=300.0, random_seed=101 + k,
                                    restart=os.path.basename(rst))
            inputs.append(p["inp"])
    return inputs


def test_c3_fp64_batch_vs_oracle(c3_inputs):
    _say("case: C3 systems, 6-instance FP64 batch, 30 steps at 1 fs / 300 K vs oracle")
    e = qe.Engine(c3_inputs, qe.QGPU_FP64)
    e.step(30)
    for k, inp in enumerate(c3_inputs):
        _say("  instance %d: %d atoms" % (k, e.natom[k]))
        o = Oracle(inp)
        o.run(30)
        xo, vo = o.state()
        xg, vg = e.state(k)
        assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max(), k
        assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max(), k
        eo, eg = o.energies(), e.energies(k)
        for t in TERMS:
            assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (k, t, eg[t], eo[t])
        qo = o.qenergies()
        qg = np.array(e.qenergies(k)).reshape(qo.shape)
        assert np.allclose(qg, qo, rtol=1e-9, atol=1e-9), (k, qg, qo)
        o.close()
    e.close()


def test_c3_mixed_batch_gradient(c3_inputs):
    _say("case: C3 systems, 6-instance mixed batch vs FP64 batch on identical coordinates")
    ef = qe.Engine(c3_inputs, qe.QGPU_FP64)
    ef.step(30)
    states = [ef.state(k) for k in range(len(c3_inputs))]
    em = qe.Engine(c3_inputs, qe.QGPU_MIXED)
    for k, (x, v) in enumerate(states):
        em.set_state(k, x, v)
    for e in (ef, em):
        e.make_pair_lists()
        e.pot_energy()
    for k in range(len(c3_inputs)):
        d64, d32 = ef.gradient(k), em.gradient(k)
        rms = np.sqrt(np.mean(d64 ** 2))
        err = np.sqrt(np.mean((d32 - d64) ** 2))
        _say("  instance %d: gradient RMS error %.3e of RMS %.3e" % (k, err, rms))
        assert err <= 1e-5 * rms, k
        q64 = np.array(ef.qenergies(k)).reshape(-1, 12)
        q32 = np.array(em.qenergies(k)).reshape(-1, 12)
        assert np.allclose(q32, q64, rtol=1e-5, atol=1e-4), (k, q32, q64)
    for t in ("ww", "pw", "pp", "qx", "lrf"):
        for e in (ef, em):
            e.pot_energy([t])
        for k in range(len(c3_inputs)):
            d64, d32 = ef.gradient(k), em.gradient(k)
            rms = np.sqrt(np.mean(d64 ** 2))
            assert rms > 0, (t, k)
            assert np.sqrt(np.mean((d32 - d64) ** 2)) <= 1e-5 * rms, (t, k)
    ef.close()
    em.close()
