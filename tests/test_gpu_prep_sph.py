"""The GPU engine on the reference's own basic test system: prep_SPH (a
27-atom Q ligand in a 16 A water sphere, 3397 atoms) and eq1.inp as
run_test.sh prepares it for the sphere with leap-frog and Berendsen
(run_test.sh:176-189), from the copies in tests/golden/prep_SPH.

* md.run in FP64: the step-0 Q-surr. row of the log equals eval_test.sh's
  known answer (eval_test.sh:96-101); every number of the three printed
  summaries (steps 0, 50 and the final one) agrees with what reference
  qdyn5 printed to within one unit of the last printed digit; the final
  restart (x, v, theta_corr after the istep = 100 polarisation update)
  agrees with the reference's to 1e-9 relative.
* The ligand's 14-atom charge group, split into slices on the device: the
  pair lists read back equal the oracle's, and the step-0 energies and
  gradient agree to 1e-9 relative.
* Mixed precision on the start coordinates: gradient RMS error at most
  1e-5 of the gradient RMS against FP64, and md.run's step-0 Q-surr. row
  again equals the known answer.
"""
import io
import os
import re
import sys

import numpy as np
import pytest

from qsource_amd import engine as qe
from qsource_amd import gensys, md

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import basic_tests as bt  # noqa: E402

NUM = re.compile(r"-?\d+\.\d+")


def _say(msg):
    print(msg, flush=True)


def _blocks(text):
    out, cur = [], None
    for l in text.splitlines():
        if re.match(r"^=+ (FINAL )?\s*Energy summary", l):
            cur = [l]
            continue
        if cur is not None:
            cur.append(l)
            if l == "=" * 79:
                out.append("\n".join(cur))
                cur = None
    return out


def _qsurr_step0(log_text):
    blk = _blocks(log_text)[0]
    return [l for l in blk.splitlines() if l.startswith("Q-surr.")]


@pytest.fixture
def eq1(tmp_path):
    return bt.sph_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE, str(tmp_path))


def test_sph_eq1_fp64_run(eq1, tmp_path):
    _say("case: prep_SPH eq1 (3397 atoms, 100 steps of 0.1 fs), md.run FP64 vs reference qdyn5")
    g = np.load(os.path.join(GOLD, "prep_sph_eq1.npz"), allow_pickle=False)
    log = io.StringIO()
    md.run([eq1], qe.QGPU_FP64, logs=[log])
    text = log.getvalue()
    assert _qsurr_step0(text) == [bt.INIT_QSURR_SPH]
    mine, ref = _blocks(text), [str(b) for b in g["blocks"]]
    assert len(mine) == len(ref) == 3
    for k, (a, b) in enumerate(zip(mine, ref)):
        la, lb = a.splitlines(), b.splitlines()
        assert len(la) == len(lb), k
        for x, y in zip(la, lb):
            na, nb = NUM.findall(x), NUM.findall(y)
            assert NUM.sub("#", x) == NUM.sub("#", y), (k, x, y)
            assert len(na) == len(nb)
            for p, q in zip(na, nb):
                assert abs(float(p) - float(q)) <= 0.0101 * 10 ** (2 - len(p.split(".")[1])), (k, x, y)
    x, v, tc = gensys.read_restart(os.path.join(str(tmp_path), "eq1.re"))
    assert np.abs(x - g["x"]).max() <= 1e-9 * np.abs(g["x"]).max()
    assert np.abs(v - g["v"]).max() <= 1e-9 * np.abs(g["v"]).max()
    assert np.abs(tc - g["theta_corr"]).max() <= 1e-9


def test_sph_split_group_lists(eq1):
    # the ligand's 14-atom charge group is split into slices of 8 and 6
    # atoms on the device; the lists read back in reference groups equal the
    # oracle's (all 46 solute atoms are Q atoms: no solute-solute or
    # solute-water pairs), and so does the list the device pair masks give
    _say("case: prep_SPH pair lists with the 14-atom solute group split on the device")
    from oracle import Oracle
    o = Oracle(eq1)
    o.final_energy()
    e = qe.Engine([eq1], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    for which in (0, 1, 2):
        assert np.array_equal(e.pairlist(0, which), o.pairlist(which)), which
    assert np.array_equal(e.pairlist(0, 3), o.pairlist(0))
    eo, eg = o.energies(), e.energies(0)
    for t in ("pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "potential"):
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    d = o.gradient()
    assert np.abs(e.gradient(0) - d).max() <= 1e-9 * np.abs(d).max()
    e.close()
    o.close()


def test_sph_eq1_mixed(eq1):
    _say("case: prep_SPH eq1, mixed gradient vs FP64 on the start coordinates; mixed md.run step-0 Q-surr.")
    ef = qe.Engine([eq1], qe.QGPU_FP64)
    em = qe.Engine([eq1], qe.QGPU_MIXED)
    for e in (ef, em):
        e.make_pair_lists()
        e.pot_energy()
    d64, d32 = ef.gradient(0), em.gradient(0)
    rms = np.sqrt(np.mean(d64 ** 2))
    err = np.sqrt(np.mean((d32 - d64) ** 2))
    _say("  gradient RMS error %.3e of RMS %.3e" % (err, rms))
    assert err <= 1e-5 * rms
    ef.close()
    em.close()
    log = io.StringIO()
    md.run([eq1], qe.QGPU_MIXED, logs=[log])
    assert _qsurr_step0(log.getvalue()) == [bt.INIT_QSURR_SPH]
