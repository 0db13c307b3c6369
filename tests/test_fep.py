"""Free-energy estimators and the energy-file format (CPU).

Pins: a constant energy gap gives dG = gap for Zwanzig and BAR exactly;
Gaussian gaps (linear response) give dG = mu - sigma^2/(2kT) in the limit
of many samples; BAR and Zwanzig agree on overlapping windows; the energy
file round-trips and holds qdyn's record layout (header + per-state records
of 4 + 8*15 bytes + an empty off-diagonal record per frame); on five lambda
windows written by reference qdyn5 (the engine writes such files byte for
byte the same), fep.read_energy_file and fep.zwanzig reproduce every
number of reference qfep5's Part 1 table (dGf, its running sum, dGr, its
running sum, <dG>) at qfep's printed precision.
"""
import os
import struct

import numpy as np

from qsource_amd import fep


def test_constant_gap():
    lam = np.array([[1.0, 0.0], [0.5, 0.5], [0.0, 1.0]])
    rng = np.random.default_rng(0)
    v = []
    for k in range(3):
        e0 = rng.normal(-50, 3, 500)
        v.append(np.stack([e0, e0 + 2.5], 1))  # state 2 is 2.5 kcal/mol above state 1
    z = fep.zwanzig(lam, v, 300.0)
    b = fep.bar(lam, v, 300.0)
    assert abs(z["total"] - 2.5) < 1e-12
    assert abs(b["total"] - 2.5) < 1e-8


def test_gaussian_linear_response():
    kT = fep.GAS_CONSTANT * 300.0
    mu, sig = 1.0, 0.8
    rng = np.random.default_rng(1)
    n = 400000
    # state 0 samples of the gap and state 1 samples (shifted by -sig^2/kT)
    g0 = rng.normal(mu, sig, n)
    g1 = rng.normal(mu - sig ** 2 / kT, sig, n)
    lam = np.array([[1.0, 0.0], [0.0, 1.0]])
    v = [np.stack([np.zeros(n), g0], 1), np.stack([np.zeros(n), g1], 1)]
    exact = mu - sig ** 2 / (2 * kT)
    z = fep.zwanzig(lam, v, 300.0)
    b = fep.bar(lam, v, 300.0)
    assert abs(z["total"] - exact) < 0.02
    assert abs(b["total"] - exact) < 0.01


def test_energy_file_roundtrip(tmp_path):
    lam = [0.7, 0.3]
    frames = np.random.default_rng(2).normal(size=(5, 2, 14))
    p = str(tmp_path / "e.en")
    fep.write_energy_file(p, lam, frames)
    raw = open(p, "rb").read()
    (n0,) = struct.unpack_from("<i", raw, 0)
    assert n0 == 7 * 4 + 80
    # first state record after the header
    off = 8 + n0
    (n1,) = struct.unpack_from("<i", raw, off)
    assert n1 == 4 + 8 * 15
    l2, tot = fep.read_energy_file(p, 2)
    assert np.allclose(l2, lam)
    assert np.array_equal(tot, frames[:, :, 0])


def test_energy_file_offdiag_roundtrip(tmp_path):
    """Off-diagonal records (OFFDIAG_SAVE: int32 i, j; FP64 Hij, rkl) written
    after each frame's state records read back unchanged, and a file the
    reference qdyn wrote (golden fixture frames) survives write -> read."""
    lam = [0.4, 0.3, 0.2, 0.1]
    rng = np.random.default_rng(3)
    frames = rng.normal(size=(4, 4, 14))
    offd = np.zeros((4, 3, 4))
    offd[:, :, 0], offd[:, :, 1] = 1, 4
    offd[:, :, 2:] = rng.uniform(0.1, 9.0, size=(4, 3, 2))
    p = str(tmp_path / "o.en")
    fep.write_energy_file(p, lam, frames, offd=offd)
    l2, rows, od = fep.read_energy_frames(p, 4)
    assert np.allclose(l2, lam) and np.array_equal(rows, frames) and np.array_equal(od, offd)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c5_r12_q8s4_gromos_30.npz"))
    p2 = str(tmp_path / "g.en")
    fep.write_energy_file(p2, lam, g["eq_frames"], offd=g["offd_frames"])
    _, rows2, od2 = fep.read_energy_frames(p2, 4)
    assert np.array_equal(rows2, g["eq_frames"]) and np.array_equal(od2, g["offd_frames"])


def test_zwanzig_matches_reference_qfep(tmp_path):
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qfep_5win.npz"),
                allow_pickle=False)
    lams = g["lambdas"]
    lam, tots = [], []
    for k in range(len(lams)):
        path = os.path.join(str(tmp_path), "w%d.en" % k)
        with open(path, "wb") as f:
            f.write(g["en%d" % k].tobytes())
        lk, tk = fep.read_energy_file(path, 2)
        assert np.allclose(lk, [1 - lams[k], lams[k]])
        lam.append(lk)
        tots.append(tk)
    kt = float(g["kT"])
    z = fep.zwanzig(np.array(lam), tots, temperature=kt / fep.GAS_CONSTANT)
    n = len(lams)
    dgf, dgr, dg = z["dGf"], z["dGr"], z["dG"]
    sum_f = np.concatenate([[0.0], np.cumsum(dgf[:-1])])
    sum_r = np.array([sum(dgr[j] for j in range(k + 1, n)) for k in range(n)])
    # qfep prints per file k: lambda(1), dGf(k-1), sum dGf up to k, dGr(k+1), sum dGr from k, <dG>(k)
    for k, row in enumerate(g["part1"]):
        ours = [lam[k][0], dgf[k - 1] if k > 0 else 0.0, sum_f[k], dgr[k + 1] if k + 1 < n else 0.0, sum_r[k],
                dg[k]]
        for a, b in zip(ours, row):
            assert abs(round(a, 3) - b) <= 0.0010001, (k, ours, list(row))
