"""GPU parity of the thermostats and integrators (newvel_ber, newvel_lan,
newvel_nos, nh_prop and velocity-Verlet, src/md.f90:6194-6344, dispatched
at :6578-6600):

* every generated thermostat case (pinned bitwise to reference qdyn by the
  CPU tests) as an FP64 instance: x, v, every energy term and the per-state
  Q energies within 1e-9 relative of the oracle after 30 steps.  Langevin
  draws its random forces on the device from randm's stream (each atom
  jumps to its own draws), so the forces are the oracle's;
* the reference's basic sphere test with every integrator/thermostat pair
  of its test matrix: 100 FP64 steps against the oracle;
* a batch that mixes thermostats and integrators equal to its single runs;
* mixed precision: Langevin and Nose-Hoover runs stay finite and draw the
  same random stream as FP64 (the state after the run is the oracle's).
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

THERMO_CASES = sorted(k for k in CASES if k.startswith("c16_"))
TERMS = ["pp_el", "pp_vdw", "pw_el", "pw_vdw", "ww_el", "ww_vdw", "lrf", "p_bond", "p_angle", "p_torsion",
         "qx_el", "qx_vdw", "potential", "kinetic"]


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
    if o.nstates > 0:
        qo = o.qenergies()
        qg = np.array(e.qenergies(k)).reshape(qo.shape)
        assert np.allclose(qg, qo, rtol=tol, atol=tol), (qg, qo)


@pytest.mark.parametrize("name", THERMO_CASES)
def test_thermostat_case_fp64_vs_oracle(name, tmp_path):
    _say("case: %s, FP64 single instance vs oracle" % name)
    inp, n = _write(name, str(tmp_path))
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(n)
    o = Oracle(inp)
    o.run(n)
    _compare(e, 0, o)
    if "npt" in name:
        assert e.volume_moves(0) == o.volume_moves()
    o.close()
    e.close()


SPH = [(i, t) for i in ("leap-frog", "velocity-verlet") for t in ("berendsen", "langevin", "nose-hoover")
       if (i, t) != ("leap-frog", "berendsen")]


@pytest.mark.parametrize("integrator,thermostat", SPH)
def test_reference_sphere_thermostats_100_steps(integrator, thermostat, tmp_path):
    _say("case: reference basic sphere test, %s / %s, FP64 100 steps vs oracle" % (integrator, thermostat))
    inp = bt.sph_input(os.path.join(bt.FIXTURE, "eq1.inp"), bt.FIXTURE, str(tmp_path), integrator=integrator,
                       thermostat=thermostat)
    e = qe.Engine([inp], qe.QGPU_FP64)
    e.step(100)
    o = Oracle(inp)
    o.run(100)
    _compare(e, 0, o)
    o.close()
    e.close()


def test_mixed_thermostat_batch_equals_single_instances(tmp_path):
    names = ["c16_r14_q8_lan_30", "c16_r14_q8_nh_shake_30", "c16_r14_q8_ber_vv_30", "c16_r14_q8_lan_gauss_vv_30"]
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
        es.close()
    eb.close()


def test_thermostats_mixed_precision(tmp_path):
    names = ["c16_r14_q8_lan_30", "c16_r14_q8_nh_shake_30"]
    _say("case: mixed precision %s" % names)
    inps = [_write(n, str(tmp_path))[0] for n in names]
    em = qe.Engine(inps, qe.QGPU_MIXED)
    em.step(30)
    for k, inp in enumerate(inps):
        o = Oracle(inp)
        o.run(30)
        xo, _ = o.state()
        xm, _ = em.state(k)
        assert np.isfinite(em.energies(k)["potential"])
        # FP32 pair arithmetic: the trajectory stays close to FP64's over 30 fs
        assert np.abs(xm - xo).max() < 1e-3, (names[k], np.abs(xm - xo).max())
        o.close()
    em.close()
