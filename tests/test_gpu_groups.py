"""Solute charge groups of more than 8 atoms: the device splits them into
slices of at most 8 atoms that keep the group's switch atom and centre.
System: R = 14 A sphere around an 80-atom solute in charge groups of up to
12 atoms, 8 Q atoms (part of one group), 2 states (the oracle equals the
reference's qdyn5 on it bit for bit: golden case c14_r14_q8s2_group12_30).

* Step 0: the solute-solute, solute-water and water-water lists read back in
  reference groups equal the oracle's, and so does the list the device pair
  masks give; energies and gradient within 1e-9 relative.
* 30 steps in FP64: x, v, every energy term and the per-state Q energies
  within 1e-9 relative of the oracle.
* Mixed precision on the same coordinates: gradient RMS error at most 1e-5
  of the gradient RMS.
"""
import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "p_bond", "p_angle",
         "p_torsion", "qx_el", "qx_vdw", "rst_radial", "rst_wpol", "potential", "kinetic"]


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def big_groups(tmp_path):
    s = gensys.solvated_system(radius=14.0, n_solute=80, seed=3201, n_q=8, n_states=2, group_size=12)
    assert max(len(g) for g in s.cgps[:s.ncgp_solute]) == 12
    return gensys.write_system(s, str(tmp_path), "bg", steps=30, output=10, lambdas=[0.6, 0.4], stepsize=0.5,
                               temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)


def test_split_groups_step0(big_groups):
    _say("case: 12-atom solute groups, step-0 lists, energies and gradient vs oracle (FP64)")
    o = Oracle(big_groups["inp"])
    o.final_energy()
    e = qe.Engine([big_groups["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    for which in (0, 1, 2):
        ref = o.pairlist(which)
        assert ref.shape[0] > 0, which
        assert np.array_equal(e.pairlist(0, which), ref), which
    assert np.array_equal(e.pairlist(0, 3), o.pairlist(0))
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS[:-1]:
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    d = o.gradient()
    assert np.abs(e.gradient(0) - d).max() <= 1e-9 * np.abs(d).max()
    e.close()
    o.close()


def test_split_groups_run(big_groups):
    _say("case: 12-atom solute groups, 30 FP64 steps vs oracle; mixed gradient")
    e = qe.Engine([big_groups["inp"]], qe.QGPU_FP64)
    e.step(30)
    o = Oracle(big_groups["inp"])
    o.run(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    assert np.allclose(np.array(e.qenergies(0)).reshape(qo.shape), qo, rtol=1e-9, atol=1e-9)
    em = qe.Engine([big_groups["inp"]], qe.QGPU_MIXED)
    em.set_state(0, xg, vg)
    for x in (e, em):
        x.make_pair_lists()
        x.pot_energy()
    d64, d32 = e.gradient(0), em.gradient(0)
    rms = np.sqrt(np.mean(d64 ** 2))
    err = np.sqrt(np.mean((d32 - d64) ** 2))
    _say("  mixed gradient RMS error %.3e of RMS %.3e" % (err, rms))
    assert err <= 1e-5 * rms
    e.close()
    em.close()
    o.close()
