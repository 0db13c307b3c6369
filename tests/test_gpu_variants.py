"""GPU parity on input variants of the step against the FP64 oracle (pinned
bitwise to reference qdyn on the same systems by test_oracle_reference.py):

* LRF off (plain cut-off lists, no multipoles or Taylor term,
  src/md.f90:6030-6058): step 0 energies and gradient within 1e-12, 30 FP64
  steps within 1e-9; mixed precision gradient RMS error below 1e-5 of the
  gradient RMS;
* the FEP file's Q-Q pair options: [soft_pairs] (exponential repulsion with
  Q-atom types), [el_scale], [excluded_pairs] (Q-Q in one state, Q/non-Q
  and non-Q/non-Q in all) and qq_use_library_charges
  (nonbond_qq_lib_charges, src/md.f90:13640-13786), and the FEP file's
  SHAKE constraints with solute SHAKE on (src/md.f90:4328-4354): step 0
  within 1e-12, 30 FP64 steps within 1e-9;
* an all-atom solvent (topology solvent type 1, TIP3P-like H with LJ:
  nonbond_3atomsolvent and nonbond_qw_3atom, src/md.f90:15238-15333,
  14749-14874) with fixed Q-atom LJ parameters and with Q-atom types: step 0
  within 1e-12, 30 FP64 steps within 1e-9, mixed precision gradient RMS
  error below 1e-5 of the gradient RMS; an SPC and an all-atom system in one
  batch are refused;
* Rcq below the exclusion radius (nbqplist / nbqwlist rebuilt at every list
  update, src/md.f90:10245, 10488): the Q list sizes equal the oracle's at
  every list update while waters leave the Q sphere, 60 FP64 steps within
  1e-9, per-state Q energies within 1e-9;
* excluded atoms (solute groups beyond the exclusion radius, waters near
  the boundary) and the restrained shell (fix_shell, src/md.f90:2816-2862),
  with and without excluded_freeze: step 0 within 1e-12, 30 FP64 steps
  within 1e-9, mixed precision gradient within 1e-5 (RMS);
* solute hydrogens: SHAKE of bonds to hydrogen with one Berendsen scale,
  sequence restraints with hydrogens and solute SHAKE, exclude_bonded:
  30 FP64 steps within 1e-9;
* monitored group pairs (nonbond_monitor, src/md.f90:11228-11350) with
  fixed Q LJ parameters and with Q-atom types: the engine's values equal the
  oracle's within 1e-10 relative after 0 and 15 steps.
"""
import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

COOL = dict(stepsize=0.5, temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)
TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "qx_el", "qx_vdw",
         "rst_radial", "rst_wpol", "potential"]


def _say(msg):
    print(msg, flush=True)


def _compare(paths, precision=qe.QGPU_FP64):
    o = Oracle(paths["inp"])
    o.final_energy()
    e = qe.Engine([paths["inp"]], precision)
    e.make_pair_lists()
    e.pot_energy()
    eo, eg = o.energies(), e.energies(0)
    do, dg = o.gradient(), e.gradient(0)
    e.close()
    return eo, eg, do, dg


def test_no_lrf(tmp_path):
    _say("case: LRF off, R=14 + 40-atom solute + 8 Q, FP64 step 0, 30 steps, mixed step 0")
    s = gensys.solvated_system(radius=14.0, n_solute=40, seed=2003, n_q=8, n_states=2)
    paths = gensys.write_system(s, str(tmp_path), "nolrf", steps=30, lambdas=[0.5, 0.5], lrf=False, **COOL)
    eo, eg, do, dg = _compare(paths)
    assert eo["lrf"] == 0.0 and eg["lrf"] == 0.0
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    e.close()
    eo, eg, do, dg = _compare(paths, qe.QGPU_MIXED)
    rms = np.sqrt(np.mean(do ** 2))
    assert np.sqrt(np.mean((dg - do) ** 2)) <= 1e-5 * rms


FEP_PAIRS = {
    "fepx": dict(radius=12.0, n_solute=40, seed=2301, n_q=8, n_states=2, impropers=1, fep_pairs=True,
                 q_types=True, softcore=40.0),
    "qqlib": dict(radius=12.0, n_solute=40, seed=2301, n_q=8, n_states=2, impropers=1, fep_pairs=True,
                  qq_lib_charges=True),
    "qshake": dict(radius=12.0, n_solute=40, seed=2401, n_q=8, n_states=2, q_shake=True),
}
EXTRA_INP = {"qshake": dict(shake_solute=True)}


@pytest.mark.parametrize("name", sorted(FEP_PAIRS))
def test_fep_pair_options(tmp_path, name):
    """[soft_pairs], [el_scale], [excluded_pairs] and qq_use_library_charges:
    step 0 energies, per-state Q energies and gradient within 1e-12; 30 FP64
    steps within 1e-9."""
    _say("case: FEP pair options (%s), R=12 + 40-atom solute + 8 Q, FP64" % name)
    s = gensys.solvated_system(**FEP_PAIRS[name])
    paths = gensys.write_system(s, str(tmp_path), name, steps=30, lambdas=[0.5, 0.5], **COOL,
                                **EXTRA_INP.get(name, {}))
    o = Oracle(paths["inp"])
    o.final_energy()
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    qg = np.array(e.qenergies(0)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-12, atol=1e-11), (qg, qo)
    do, dg = o.gradient(), e.gradient(0)
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    e.close()
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    e.close()


ALLATOM = {
    "fixed": dict(radius=12.0, n_solute=40, seed=2501, n_q=8, n_states=2, solvent_allatom=True),
    "qvdw": dict(radius=12.0, n_solute=40, seed=2502, n_q=8, n_states=2, q_types=True, softcore=30.0,
                 solvent_allatom=True),
}


@pytest.mark.parametrize("name", sorted(ALLATOM))
def test_allatom_solvent(tmp_path, name):
    """LJ on every water atom: step 0 energies, per-state Q energies and
    gradient within 1e-12 of the oracle, 30 FP64 steps within 1e-9, mixed
    precision gradient within 1e-5 (RMS)."""
    _say("case: all-atom solvent (%s), R=12 + 40-atom solute + 8 Q, FP64 and mixed" % name)
    s = gensys.solvated_system(**ALLATOM[name])
    assert s.solvent_type == 1
    paths = gensys.write_system(s, str(tmp_path), name, steps=30, lambdas=[0.6, 0.4], **COOL)
    o = Oracle(paths["inp"])
    o.final_energy()
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    eo, eg = o.energies(), e.energies(0)
    # the H LJ is present: the SPC formula would miss a visible part of ww_vdw
    assert eo["ww_vdw"] != 0.0
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    qg = np.array(e.qenergies(0)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-12, atol=1e-11), (qg, qo)
    do, dg = o.gradient(), e.gradient(0)
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    e.close()
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    e.close()
    eo, eg, do, dg = _compare(paths, qe.QGPU_MIXED)
    rms = np.sqrt(np.mean(do ** 2))
    assert np.sqrt(np.mean((dg - do) ** 2)) <= 1e-5 * rms
    for t in ("ww_vdw", "qx_vdw"):
        assert abs(eg[t] - eo[t]) <= 1e-5 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])


def test_mixed_solvent_batch_refused(tmp_path):
    _say("case: SPC and all-atom solvent in one batch (refused at set-up)")
    a = gensys.solvated_system(radius=10.0, n_solute=0, seed=11)
    b = gensys.solvated_system(radius=10.0, n_solute=0, seed=12, solvent_allatom=True)
    pa = gensys.write_system(a, str(tmp_path), "spc", steps=1)
    pb = gensys.write_system(b, str(tmp_path), "aa", steps=1)
    with pytest.raises(RuntimeError, match="solvent type"):
        qe.Engine([pa["inp"], pb["inp"]], qe.QGPU_FP64)


def test_dynamic_q_lists(tmp_path):
    """Q-solute and Q-water lists rebuilt on the device every 5 steps (Rcq 7
    inside a 14 A sphere at 300 K): list sizes match the oracle's at each
    update, and the run stays within 1e-9 of the FP64 oracle."""
    _say("case: dynamic Q lists (Rcq 7 < rexcl_o 14), R=14 + 40-atom solute + 8 Q, FP64 60 steps")
    s = gensys.solvated_system(radius=14.0, n_solute=40, seed=2601, n_q=8, n_states=2)
    paths = gensys.write_system(s, str(tmp_path), "dynq", steps=60, lambdas=[0.6, 0.4], stepsize=1.0,
                                temperature=300.0, bath_coupling=10.0, initial_temperature=300.0,
                                rcq=7.0, non_bond=5)
    o = Oracle(paths["inp"])
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    sizes_o, sizes_g = [], []
    for blk in range(12):
        if blk == 0:
            o.final_energy()
            e.make_pair_lists()
        sizes_o.append(list(o.list_sizes()[3:]))
        c = e.list_counts(0)
        sizes_g.append([c["qp"], c["qw"]])
        o.run(5)
        e.step(5)
    assert sizes_g == sizes_o, (sizes_g, sizes_o)
    assert len(set(tuple(v) for v in sizes_o)) > 1, "the Q lists did not change during the run"
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    qo = o.qenergies()
    qg = np.array(e.qenergies(0)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=1e-9, atol=1e-8), (qg, qo)
    e.close()


@pytest.mark.parametrize("freeze", [False, True])
def test_excluded_atoms(tmp_path, freeze):
    """Excluded solute groups and waters with the excluded-atom restraint (or
    frozen at their topology positions) and shell atoms with the shell
    restraint: the GPU step against the oracle."""
    _say("case: excluded atoms, R=12 + 120-atom solute reaching past the boundary, freeze=%s" % freeze)
    s = gensys.solvated_system(radius=12.0, n_solute=120, seed=2702, n_q=8, n_states=2, solute_rmax=16.0,
                               exclude_outside=True, exclude_water_radius=11.3)
    assert s.excl[:s.nat_solute].any() and s.excl[s.nat_solute:].any()
    paths = gensys.write_system(s, str(tmp_path), "excl", steps=30, lambdas=[0.6, 0.4], excluded_freeze=freeze,
                                **COOL)
    eo, eg, do, dg = _compare(paths)
    for t in TERMS + ["rst_fix", "rst_shell"]:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    eo2, eg2 = o.energies(), e.energies(0)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    for t in ("rst_fix", "rst_shell", "potential"):
        assert abs(eg2[t] - eo2[t]) <= 1e-9 * max(1.0, abs(eo2[t])), (t, eg2[t], eo2[t])
    if not freeze:
        assert eo2["rst_fix"] > 0.0
    e.close()
    eo, eg, do, dg = _compare(paths, qe.QGPU_MIXED)
    rms = np.sqrt(np.mean(do ** 2))
    assert np.sqrt(np.mean((dg - do) ** 2)) <= 1e-5 * rms


HYDROGENS = {
    "shake_h": (dict(radius=12.0, n_solute=60, seed=2801, n_q=8, n_states=2, hydrogens=True),
                dict(shake_hydrogens=True, separate_scaling=False)),
    "rst_h": (dict(radius=12.0, n_solute=60, seed=2802, n_q=8, n_states=2, hydrogens=True, restraints=True,
                   q_types=True, softcore=30.0), dict(shake_solute=True, shake_hydrogens=True)),
    "exclude_bonded": (dict(radius=12.0, n_solute=120, seed=2803, n_q=8, n_states=2, hydrogens=True,
                            solute_rmax=16.0, exclude_outside=True), dict(exclude_bonded=True)),
}


@pytest.mark.parametrize("name", sorted(HYDROGENS))
def test_hydrogen_variants(tmp_path, name):
    _say("case: solute hydrogens (%s), R=12, FP64 step 0 and 30 steps" % name)
    sysargs, inpargs = HYDROGENS[name]
    s = gensys.solvated_system(**sysargs)
    paths = gensys.write_system(s, str(tmp_path), name, steps=30, lambdas=[0.6, 0.4], **COOL, **inpargs)
    eo, eg, do, dg = _compare(paths)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()
    o = Oracle(paths["inp"])
    o.run(30)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(30)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    e.close()


@pytest.mark.parametrize("qvdw", [False, True])
def test_monitor_groups(tmp_path, qvdw):
    _say("case: monitored group pairs, R=12 + 40-atom solute + 8 Q, qvdw=%s, FP64" % qvdw)
    kw = dict(q_types=True, softcore=30.0) if qvdw else {}
    s = gensys.solvated_system(radius=12.0, n_solute=40, seed=2902 if qvdw else 2901, n_q=8, n_states=2,
                               monitor=True, **kw)
    paths = gensys.write_system(s, str(tmp_path), "mon", steps=15, lambdas=[0.6, 0.4], **COOL)
    o = Oracle(paths["inp"])
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    for k in range(2):
        mo, mg = o.monitor(), e.monitor(0)
        assert mg.shape == mo.shape == (3, 7)
        assert np.allclose(mg, mo, rtol=1e-10, atol=1e-9), (k, mg, mo)
        o.run(15)
        e.step(15)
    e.close()
