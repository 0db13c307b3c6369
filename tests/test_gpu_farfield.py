"""The learned far-field surrogate on the GPU.

* Training kernels (BF16 matrix cores, FP32 accumulation): forward outputs,
  loss and every parameter gradient against the CPU restatement
  (oracle/mlp.py) with the same BF16 rounding points, within 2e-3 of each
  tensor's largest entry (FP32 accumulation order and exp differ).
* Targets: the far-field potential and field at every charge-group centre
  (from a fresh list build) against a brute-force FP64 sum over the far
  atoms, within 1e-5 relative (FP32 storage).
* Gain: trained on frames of the engine's own trajectory.  Local features
  (group charge, dipole, radial position) fix the mean far field (dielectric
  reaction and the boundary's radial polarisation), not the instantaneous
  field of distant moving dipoles.  So the bar is: the radially binned mean
  potential and radial field on held-out frames are closer to the exact ones
  than with no far field (plain cut-off), and the network is never worse than
  the cut-off sample by sample in RMS.  The LRF estimate's instantaneous
  error is reported beside it.
* The engine runs MD with the surrogate in place of LRF.
"""
import numpy as np
import pytest

from oracle import mlp as omlp
from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

COOL = dict(stepsize=0.5, temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)


def _say(msg):
    print(msg, flush=True)


def test_mlp_kernels_match_oracle():
    _say("case: MLP forward/backward kernels vs oracle, n=300 (padded to 320)")
    rng = np.random.default_rng(4)
    X = rng.normal(size=(300, 4)).astype(np.float32) * np.array([2, 1, 0.5, 0.3], np.float32) + 0.1
    Y = rng.normal(size=(300, 3)).astype(np.float32)
    m = qe.FarFieldMLP(seed=9)
    m.set_data(X, Y)            # sets the normalisation from (X, Y)
    p, nm = m.params()
    g, loss = m.gradient(X, Y)
    xn, yn = omlp.normalise(X, Y, nm)
    lo, go = omlp.loss_and_grad(p, xn, yn, bf=True)
    assert abs(loss - lo) <= 1e-4 * abs(lo), (loss, lo)
    for a, b in [(omlp.OFF_W1, omlp.OFF_B1), (omlp.OFF_B1, omlp.OFF_W2), (omlp.OFF_W2, omlp.OFF_B2),
                 (omlp.OFF_B2, omlp.OFF_W3), (omlp.OFF_W3, omlp.OFF_B3), (omlp.OFF_B3, omlp.NPARAM)]:
        scale = np.abs(go[a:b]).max()
        assert np.abs(g[a:b] - go[a:b]).max() <= 2e-3 * scale, (a, np.abs(g[a:b] - go[a:b]).max(), scale)
    yo, _ = omlp.forward(p, xn, bf=True)
    yp = (m.predict(X) - nm[8:11]) / nm[11:14]
    assert np.abs(yp - yo).max() <= 1e-3 * max(1.0, np.abs(yo).max())
    m.close()


def _groups(s):
    groups = [[a - 1 for a in g] for g in s.cgps]
    switch = [g[0] for g in groups]
    qat = np.zeros(s.natom, dtype=bool)
    if s.qatoms is not None:
        qat[np.asarray(s.qatoms) - 1] = True
    return groups, switch, qat


def test_far_field_targets_exact(tmp_path):
    _say("case: exact far-field targets vs brute force, R=12 + 40-atom solute + 8 Q")
    s = gensys.solvated_system(radius=12.0, n_solute=40, seed=2001, n_q=8, n_states=2)
    paths = gensys.write_system(s, str(tmp_path), "ff", steps=10, lambdas=[0.5, 0.5], **COOL)
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    e.step(3)
    x, _ = e.state(0)
    X, Y = qe.far_samples(e, 0)
    groups, switch, qat = _groups(s)
    crg = np.round(s.crg, 4) * np.sqrt(332.0)   # the topology holds charges to 4 decimals
    cent, phi, E = omlp.far_field_exact(x, crg, groups, switch, np.zeros(s.natom, bool), qat,
                                        10.0, 10.0, 10.0, s.nat_solute)
    er = cent / np.linalg.norm(cent, axis=1)[:, None]
    assert np.allclose(Y[:, 0], phi, rtol=1e-5, atol=1e-5 * np.abs(phi).max())
    assert np.allclose(Y[:, 1], np.sum(E * er, 1), rtol=1e-5, atol=1e-5 * np.abs(E).max())
    # features: non-Q charge and dipole about the centre, radial position
    q = np.array([sum(crg[a] for a in g if not qat[a]) for g in groups])
    assert np.allclose(X[:, 0], q, rtol=1e-5, atol=1e-5)
    assert np.allclose(X[:, 3], np.linalg.norm(cent, axis=1) / s.rwat, rtol=1e-6)
    e.close()


def _frames(paths, nframes, stride):
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    Xs, Ys, Ls = [], [], []
    for _ in range(nframes):
        e.step(stride)             # stride is not a multiple of the list interval
        Xl, Yl = qe.far_samples(e, 0, lrf_estimate=True)
        Xe, Ye = qe.far_samples(e, 0)
        Xs.append(Xe); Ys.append(Ye); Ls.append(Yl)
    e.close()
    return Xs, Ys, Ls


def test_surrogate_gain_over_cutoff(tmp_path):
    _say("case: train surrogate on 24 frames, evaluate on 8 held-out frames, R=15 + 60-atom solute")
    s = gensys.solvated_system(radius=15.0, n_solute=60, seed=4001)
    paths = gensys.write_system(s, str(tmp_path), "tr", steps=1000, **COOL)
    Xs, Ys, Ls = _frames(paths, 32, 7)
    Xtr, Ytr = np.concatenate(Xs[:24]), np.concatenate(Ys[:24])
    Xte, Yte, Lte = np.concatenate(Xs[24:]), np.concatenate(Ys[24:]), np.concatenate(Ls[24:])
    m = qe.FarFieldMLP(seed=3)
    m.set_data(Xtr, Ytr)
    loss = m.train(epochs=60, batch=512, lr=2e-3, seed=5)
    P = m.predict(Xte)
    rms = lambda a: float(np.sqrt(np.mean(a ** 2)))
    err_none_E, err_mlp_E, err_lrf_E = rms(Yte[:, 1:]), rms(P[:, 1:] - Yte[:, 1:]), rms(Lte[:, 1:] - Yte[:, 1:])
    err_none_V, err_mlp_V = rms(Yte[:, 0]), rms(P[:, 0] - Yte[:, 0])
    _say("train loss %.4f; field RMS error: none %.4f  mlp %.4f  lrf(stale) %.4f; potential: none %.4f mlp %.4f"
         % (loss, err_none_E, err_mlp_E, err_lrf_E, err_none_V, err_mlp_V))
    # systematic part: the radially binned mean potential and radial field
    bins = np.clip((Xte[:, 3] * 10).astype(int), 0, 11)
    sys_err = {}
    for name, col in (("V", 0), ("Er", 1)):
        mt = np.array([Yte[bins == b, col].mean() for b in np.unique(bins)])
        mp = np.array([P[bins == b, col].mean() for b in np.unique(bins)])
        sys_err[name] = (rms(mt), rms(mp - mt))
    _say("binned mean error: V none %.4f mlp %.4f; Er none %.4f mlp %.4f"
         % (sys_err["V"] + sys_err["Er"]))
    assert err_mlp_E <= err_none_E and err_mlp_V <= err_none_V
    assert sys_err["V"][1] < sys_err["V"][0] and sys_err["Er"][1] < sys_err["Er"][0]
    # a second run with the same seeds is bitwise identical (fixed-order reductions)
    m2 = qe.FarFieldMLP(seed=3)
    m2.set_data(Xtr, Ytr)
    m2.train(epochs=2, batch=512, lr=2e-3, seed=5)
    m3 = qe.FarFieldMLP(seed=3)
    m3.set_data(Xtr, Ytr)
    m3.train(epochs=2, batch=512, lr=2e-3, seed=5)
    assert np.array_equal(m2.params()[0], m3.params()[0])
    # the engine runs MD with the surrogate in place of LRF
    e = qe.Engine([paths["inp"]], qe.QGPU_FP64)
    qe.set_far_model(e, m)
    e.step(50)
    x, v = e.state(0)
    en = e.energies(0)
    assert np.all(np.isfinite(x)) and np.isfinite(en["lrf"]) and en["lrf"] != 0.0
    assert 20.0 < e.temperature(0) < 120.0
    e.close()
    for mm in (m, m2, m3):
        mm.close()
