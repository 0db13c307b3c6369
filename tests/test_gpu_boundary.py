"""The two boundary calls that mirror the reference's call sites one to one,
through the C ABI against the oracle:

* qgpu_pot_energy_nonbonds (pot_energy_nonbonds, src/md.f90:16261-16344):
  at coordinates moved away from the list build, the accumulated gradient
  and the seven nonbonded energies plus the per-state Q-solute and Q-water
  energies within 1e-9 relative of the oracle in FP64 (1e-5 of the
  gradient RMS in mixed precision); the coordinates become the instance's;
* qgpu_shake (shake(xx, x), src/md.f90:17581-17679): constrained
  coordinates within 1e-12 A of the oracle's and the same average
  iteration count; with the iteration limit lowered from the host
  (QGPU_SHAKE_MAXITER=1) the call fails with "shake failure", as the
  reference dies, and every kernel exits normally.
"""
import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

COOL = dict(stepsize=0.5, temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def fep_system(tmp_path):
    s = gensys.solvated_system(radius=14.0, n_solute=40, seed=2001, n_q=8, n_states=2, q_types=True,
                               softcore=20.0)
    return gensys.write_system(s, str(tmp_path), "bd", steps=10, output=10, lambdas=[0.7, 0.3], **COOL)


@pytest.mark.parametrize("precision", [qe.QGPU_FP64, qe.QGPU_MIXED])
def test_pot_energy_nonbonds(fep_system, precision):
    _say("case: pot_energy_nonbonds, R=14 + 40-atom solute, 8 Q atoms, 2 states, precision %d" % precision)
    inp = fep_system["inp"]
    e = qe.Engine([inp, inp], precision)
    o = Oracle(inp)
    e.make_pair_lists()
    o.final_energy()               # lists and energies at the start coordinates
    x0, _ = o.state()
    rng = np.random.default_rng(7)
    x1 = x0 + 0.02 * rng.standard_normal(x0.size)
    d0 = rng.standard_normal(x0.size)
    dg, do = d0.copy(), d0.copy()
    eg, eqg = e.pot_energy_nonbonds(1, x1, dg)
    eo, eqo = o.pot_energy_nonbonds(x1, do)
    assert eqg.shape == eqo.shape == (2, 4)
    gg, go = dg - d0, do - d0
    rms = np.sqrt(np.mean(go ** 2))
    if precision == qe.QGPU_FP64:
        assert np.abs(gg - go).max() <= 1e-9 * np.abs(go).max()
        assert np.allclose(eg, eo, rtol=1e-9, atol=1e-9), (eg, eo)
        assert np.allclose(eqg, eqo, rtol=1e-9, atol=1e-9), (eqg, eqo)
    else:
        err = np.sqrt(np.mean((gg - go) ** 2))
        _say("  gradient RMS error %.3e of RMS %.3e" % (err, rms))
        assert err <= 1e-5 * rms
        assert np.allclose(eg, eo, rtol=1e-5, atol=1e-4), (eg, eo)
        assert np.allclose(eqg, eqo, rtol=1e-5, atol=1e-4), (eqg, eqo)
    assert np.all(eo[[2, 3, 4, 5, 6]] != 0) and np.all(eqo != 0)
    assert np.array_equal(e.state(1)[0], x1)
    # instance 0 is untouched by a call for instance 1
    assert np.array_equal(e.state(0)[0], x0)
    e.close()
    o.close()


def test_shake_matches_oracle(fep_system):
    _say("case: shake(xx, x), R=14 + 40-atom solute, waters constrained, two displacement sizes")
    inp = fep_system["inp"]
    e = qe.Engine([inp], qe.QGPU_FP64)
    o = Oracle(inp)
    xx, _ = o.state()
    rng = np.random.default_rng(11)
    for scale in (0.005, 0.05):
        x = xx + scale * rng.standard_normal(xx.size)
        xg, xo = x.copy(), x.copy()
        ng = e.shake(0, xx, xg)
        no = o.shake(xx, xo)
        _say("  displacement %.3f A: %d iterations per molecule" % (scale, no))
        assert ng == no and no >= 1
        assert np.abs(xg - xo).max() <= 1e-12
        assert not np.array_equal(xg, x)
    e.close()
    o.close()


def test_shake_failure_from_host(fep_system, monkeypatch):
    _say("case: shake with the iteration limit lowered to 1 fails with 'shake failure'")
    monkeypatch.setenv("QGPU_SHAKE_MAXITER", "1")
    inp = fep_system["inp"]
    e = qe.Engine([inp], qe.QGPU_FP64)
    xx, _ = e.state(0)
    x = xx + 0.05 * np.random.default_rng(3).standard_normal(xx.size)
    with pytest.raises(qe.QgpuError, match="shake failure"):
        e.shake(0, xx, x)
    # the boundary call's failure does not stop the batch
    e.make_pair_lists()
    e.pot_energy()
    assert e.status(0) == (0, -1)
    e.close()
