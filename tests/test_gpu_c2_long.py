"""The solvated-solute workload at full size over its 5000 steps: a 25 A
TIP3P sphere around a 300-atom branched solute, 12 Q atoms, 2 FEP states,
LRF on, equilibrated on the GPU, then 5000 steps of 1 fs at 300 K from the
same restart and Maxwell velocities in FP64 and in mixed precision
(FP32 pair arithmetic, FP64 accumulation).

The two trajectories are the same for the first steps and then separate
(MD is chaotic; the FP32 rounding seeds it), so the long run is compared
through what the free-energy workload uses: the per-step energy-gap dU
between the states, and the temperature and potential it samples.

* the first 20 steps: coordinates within 1e-4 A of the FP64 run;
* 5000 steps: no failure, finite energies; the mixed run's mean dU within
  5 block standard errors (blocks of 500 steps) plus 0.05 kcal/mol of the
  FP64 run's, the final temperature within 10 % of 300 K in both;
* FP64 re-evaluation of the mixed run's recorded frames: its dU within
  1e-3 kcal/mol of the mixed dU on every frame (the precision error alone).
"""
import os

import numpy as np
import pytest

from qsource_amd import engine as qe
from qsource_amd import equilibrate, gensys

pytestmark = pytest.mark.gpu

STEPS = 5000


def _say(msg):
    print(msg, flush=True)


def _block_se(a, nblk=10):
    b = np.array([np.mean(c) for c in np.array_split(np.asarray(a), nblk)])
    return float(np.std(b, ddof=1) / np.sqrt(nblk))


@pytest.fixture(scope="module")
def c2(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c2long"))
    _say("building the R=25 A / 300-atom solute system")
    s = gensys.solvated_system(radius=25.0, n_solute=300, seed=2500, n_q=12, n_states=2)
    _say("equilibrating on the GPU (%d atoms, %d waters)" % (s.natom, s.nwat))
    rst = equilibrate.equilibrate(s, d, "c2", lambdas=[0.5, 0.5], seed=25, progress=_say)
    p = gensys.write_system(s, d, "c2_run", steps=STEPS, output=STEPS, lambdas=[0.5, 0.5],
                            initial_temperature=300.0, random_seed=251, restart=os.path.basename(rst))
    return p["inp"]


def _run(inp, precision):
    e = qe.Engine([inp], precision)
    e.set_energy_interval(1, STEPS)
    e.step(20)
    x20 = e.state(0)[0]
    e.step(STEPS - 20)
    return e, x20


def test_c2_5000_steps_mixed_vs_fp64(c2):
    _say("case: R=25 / 300-atom solute, 5000 steps FP64")
    e64, x64 = _run(c2, qe.QGPU_FP64)
    _say("case: R=25 / 300-atom solute, 5000 steps mixed precision")
    e32, x32 = _run(c2, qe.QGPU_MIXED)
    assert np.abs(x32 - x64).max() < 1e-4, np.abs(x32 - x64).max()
    f64 = e64.energy_series(0)
    f32 = e32.energy_series(0)
    f32r = e32.energy_series(0, fp64=True)
    # a frame at every step after the first (put_ene at istep > 0)
    assert f64.shape == f32.shape == (STEPS - 1, 2, 14)
    du64 = f64[:, 1, 0] - f64[:, 0, 0]
    du32 = f32[:, 1, 0] - f32[:, 0, 0]
    du32r = f32r[:, 1, 0] - f32r[:, 0, 0]
    assert np.all(np.isfinite(du32))
    prec = float(np.abs(du32 - du32r).max())
    d_mean = float(np.mean(du32) - np.mean(du64))
    se = float(np.hypot(_block_se(du32), _block_se(du64)))
    E64, E32 = e64.energies(0), e32.energies(0)
    T64, T32 = e64.temperature(0), e32.temperature(0)
    _say("  dU mean: FP64 %.3f, mixed %.3f (difference %.3f, block stderr %.3f); precision error %.2e" % (
        np.mean(du64), np.mean(du32), d_mean, se, prec))
    _say("  final: potential FP64 %.2f mixed %.2f; T FP64 %.1f mixed %.1f" % (
        E64["potential"], E32["potential"], T64, T32))
    assert prec < 1e-3
    assert abs(d_mean) <= 5 * se + 0.05
    assert abs(T64 - 300.0) < 30.0 and abs(T32 - 300.0) < 30.0
    assert np.isfinite(E32["potential"])
    e64.close()
    e32.close()
