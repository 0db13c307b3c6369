"""GPU parity for the periodic box (the box kernels of the reference:
nb*list_box(_lrf), nonbond_*_box, lrf_update with shifts, put_back_in_box,
src/md.f90:7810-15444, 18743-18930).

The systems are the golden periodic-box cases (pinned bitwise to reference
qdyn by the CPU tests) and the reference's own basic PBC test system:

* every golden box case as an FP64 instance: x, v, every energy term, the
  per-state Q energies and the box centre within 1e-9 relative of the
  oracle after 30 steps (lists every 5 steps, molecules put back in the box);
* the same cases as one FP64 batch (box and sphere instances together)
  equal to the single-instance runs;
* the mixed-precision engine on the same coordinates: gradient RMS error at
  most 1e-5 of the gradient RMS;
* the reference basic PBC system (12 262 atoms, 49.6 A box, LRF, the
  solvent put back every 25 steps): 100 FP64 steps against the oracle;
* the monitored group pairs of a box case against the oracle;
* constant pressure (MC_volume): every golden NPT case as an FP64 instance
  with the oracle's verdict for every volume move and the same state, box
  and energies within 1e-9; two NPT cases as one batch equal to their
  single runs; the reference's NPT test (tests/test5, nine moves) over 100
  FP64 steps against the oracle; md.run printing the reference's
  volume-change summary; a mixed-precision NPT run.
"""
import os
import sys

import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import gensys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
sys.path.insert(0, HERE)
import basic_tests as bt  # noqa: E402
from test_oracle_reference import CASES  # noqa: E402

pytestmark = pytest.mark.gpu

BOX_CASES = sorted(k for k, (sa, _) in CASES.items() if sa.get("box"))
NPT_CASES = sorted(k for k, (_, ia) in CASES.items() if ia.get("pbc", {}).get("constant_pressure"))
TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "p_bond", "p_angle", "p_torsion",
         "qx_el", "qx_vdw", "rst_protein", "potential", "kinetic"]


def _say(msg):
    print(msg, flush=True)


def _write(name, d, **over):
    sysargs, inpargs = CASES[name]
    inpargs = dict(inpargs, **over)
    s = gensys.solvated_system(**sysargs)
    return gensys.write_system(s, d, name, **inpargs)["inp"], inpargs["steps"]


def _compare(e, k, o, tol=1e-9):
    xo, vo = o.state()
    xg, vg = e.state(k)
    assert np.abs(xg - xo).max() <= tol * np.abs(xo).max(), np.abs(xg - xo).max()
    assert np.abs(vg - vo).max() <= tol * np.abs(vo).max(), np.abs(vg - vo).max()
    eo, eg = o.energies(), e.energies(k)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= tol * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    qo = o.qenergies()
    qg = np.array(e.qenergies(k)).reshape(qo.shape)
    assert np.allclose(qg, qo, rtol=tol, atol=tol), (qg, qo)
    bo, bg = o.box(), e.box(k)
    assert np.array_equal(bg[0], bo[0])
    assert np.allclose(bg[1], bo[1], rtol=0, atol=1e-9), (bg[1], bo[1])


@pytest.mark.parametrize("name", BOX_CASES)
def test_box_case_fp64_vs_oracle(name, tmp_path):
    _say("case: %s, FP64 single instance vs oracle" % name)
    inp, n = _write(name, str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(n)
    o = Oracle(inp)
    o.run(n)
    _compare(e, 0, o)
    o.close()
    e.close()


def test_box_batch_equals_single_instances(tmp_path):
    names = ["c15_box28_q6_float_30", "c15_box28_q6_qvdw_30", "c2_r14_q8_30"]
    _say("case: FP64 batch %s vs single instances" % names)
    # one list interval for the batch: the sphere case's lists every 5 steps too
    inps = [_write(n, str(tmp_path), non_bond=5)[0] for n in names]
    eb = qe.Engine(inps, qe.QGPU_FP64)
    eb.step(30)
    for k, inp in enumerate(inps):
        es = qe.Engine([inp], qe.QGPU_FP64)
        es.step(30)
        xb, vb = eb.state(k)
        xs, vs = es.state(0)
        assert np.abs(xb - xs).max() <= 1e-12 * np.abs(xs).max(), names[k]
        assert np.abs(vb - vs).max() <= 1e-12 * np.abs(vs).max(), names[k]
        es.close()
    eb.close()


def test_box_mixed_gradient(tmp_path):
    names = ["c15_box28_q6_30", "c15_box28_q6_qvdw_30", "c15_box28_q6_rigid_30"]
    _say("case: mixed vs FP64 gradient on identical coordinates, %s" % names)
    inps = [_write(n, str(tmp_path))[0] for n in names]
    ef = qe.Engine(inps, qe.QGPU_FP64)
    ef.step(20)
    states = [ef.state(k) for k in range(len(inps))]
    em = qe.Engine(inps, qe.QGPU_MIXED)
    for k, (x, v) in enumerate(states):
        em.set_state(k, x, v)
    for e in (ef, em):
        e.make_pair_lists()
        e.pot_energy()
    for k in range(len(inps)):
        d64, d32 = ef.gradient(k), em.gradient(k)
        rms = np.sqrt(np.mean(d64 ** 2))
        err = np.sqrt(np.mean((d32 - d64) ** 2))
        _say("  %s: gradient RMS error %.3e of RMS %.3e" % (names[k], err, rms))
        assert err <= 1e-5 * rms, names[k]
        q64 = np.array(ef.qenergies(k)).reshape(-1, 12)
        q32 = np.array(em.qenergies(k)).reshape(-1, 12)
        assert np.allclose(q32, q64, rtol=1e-5, atol=1e-4), (k, q32, q64)
    em.step(30)
    for k in range(len(inps)):
        assert np.isfinite(em.energies(k)["potential"])
    ef.close()
    em.close()


def test_box_monitor_groups(tmp_path):
    _say("case: monitored group pairs in a periodic box vs oracle")
    inp, _ = _write("c15_box28_q6_monitor_30", str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(10)
    o = Oracle(inp)
    o.run(10)
    mg, mo = e.monitor(0), o.monitor()
    assert mg.shape == mo.shape
    assert np.allclose(mg, mo, rtol=1e-9, atol=1e-9), (mg, mo)
    o.close()
    e.close()


def test_reference_pbc_system_100_steps(tmp_path):
    _say("case: reference basic PBC system (12262 atoms, 49.6 A box), FP64 100 steps vs oracle")
    inp = bt.pbc_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE_PBC, str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(100)
    o = Oracle(inp)
    o.run(100)
    _compare(e, 0, o)
    o.close()
    e.close()
    _say("case: reference basic PBC system, mixed precision 100 steps")
    em = qe.Engine([inp], qe.QGPU_MIXED)
    em.step(100)
    assert np.isfinite(em.energies(0)["potential"])
    em.close()


def _moves(e, o):
    mg, mo = e.volume_moves(0), o.volume_moves()
    assert mg == mo, (mg, mo)
    return mg


@pytest.mark.parametrize("name", NPT_CASES)
def test_npt_case_fp64_vs_oracle(name, tmp_path):
    _say("case: %s, constant pressure, FP64 single instance vs oracle" % name)
    inp, n = _write(name, str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(n)
    o = Oracle(inp)
    o.run(n)
    tries, acc, log = _moves(e, o)
    assert tries > 0
    _compare(e, 0, o)
    o.close()
    e.close()


def test_npt_batch_equals_single_instances(tmp_path):
    names = ["c15_box28_q6_npt_30", "c15_box28_q6_npt_hot_30", "c15_box28_q6_30"]
    _say("case: FP64 batch %s vs single instances" % names)
    inps = [_write(n, str(tmp_path))[0] for n in names]
    eb = qe.Engine(inps, qe.QGPU_FP64)
    eb.step(30)
    for k, inp in enumerate(inps):
        es = qe.Engine([inp], qe.QGPU_FP64)
        es.step(30)
        xb, vb = eb.state(k)
        xs, vs = es.state(0)
        assert np.abs(xb - xs).max() <= 1e-12 * np.abs(xs).max(), names[k]
        assert np.abs(vb - vs).max() <= 1e-12 * np.abs(vs).max(), names[k]
        assert eb.volume_moves(k) == es.volume_moves(0), names[k]
        assert np.array_equal(eb.box(k)[0], es.box(0)[0]), names[k]
        es.close()
    eb.close()


def test_reference_npt_system_100_steps(tmp_path):
    _say("case: reference NPT test (tests/test5 eq1, 12262 atoms, 9 volume moves), FP64 100 steps vs oracle")
    inp = bt.npt_input(str(tmp_path), **bt.NPT_SHORT)
    g = np.load(os.path.join(HERE, "golden", "test5_npt_eq1.npz"), allow_pickle=False)
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(100)
    o = Oracle(inp)
    o.run(100)
    tries, acc, log = _moves(e, o)
    assert [1 if f else 0 for f in log] == list(g["volume_moves"])
    _compare(e, 0, o)
    o.close()
    e.close()


def test_npt_md_run_prints_volume_summary(tmp_path):
    import io

    from qsource_amd import md
    name = "c15_box28_q6_npt_hot_30"
    _say("case: md.run of %s prints the reference's volume-change summary" % name)
    inp, _ = _write(name, str(tmp_path))
    g = np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)
    log = io.StringIO()
    md.run([inp], precision=qe.QGPU_FP64, logs=[log])
    assert str(g["volume_summary"]) in log.getvalue()


def test_npt_mixed_precision_run(tmp_path):
    name = "c15_box28_q6_npt_30"
    _say("case: %s, mixed precision" % name)
    inp, n = _write(name, str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_MIXED)
    e.step(n)
    tries, acc, log = e.volume_moves(0)
    assert tries == 5   # list updates at steps 5, 10, 15, 20, 25
    assert np.isfinite(e.energies(0)["potential"])
    e.close()
