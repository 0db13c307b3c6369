"""Training data and training of the learned far-field surrogate.

The targets are the engine's own exact far field: after a fresh pair-list
build, the LRF multipole sums phi0 and phi1 at each charge-group centre run
over every atom beyond the cut-off (lrf_update, src/md.f90:5896-5981), so
they are the all-pairs potential and field of the far charges.  Frames come
from an FP64 run of the engine; the network is trained on the GPU by the
library's BF16 matrix-core kernels (qgpu_mlp_train).
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from . import engine as qe


def sample_frames(inp: str, nframes: int, stride: int, precision: int = qe.QGPU_FP64,
                  with_lrf: bool = False) -> Tuple[List[np.ndarray], List[np.ndarray], List[np.ndarray]]:
    """Features X (ncgp, 4) and exact far-field targets Y (ncgp, 3) of
    `nframes` frames `stride` steps apart of the run of `inp`; with
    with_lrf, also the LRF estimate of the last list update at each frame
    (the far field the reference would use there)."""
    e = qe.Engine([inp], precision)
    Xs, Ys, Ls = [], [], []
    for _ in range(nframes):
        e.step(stride)
        if with_lrf:
            Ls.append(qe.far_samples(e, 0, lrf_estimate=True)[1])
        X, Y = qe.far_samples(e, 0)
        Xs.append(X)
        Ys.append(Y)
    e.close()
    return Xs, Ys, Ls


def train_surrogate(X: np.ndarray, Y: np.ndarray, epochs: int = 40, batch: int = 1024, lr: float = 2e-3,
                    seed: int = 1) -> Tuple["qe.FarFieldMLP", float]:
    """A surrogate trained on (X, Y); returns (model, last epoch's loss)."""
    m = qe.FarFieldMLP(seed=seed)
    m.set_data(X, Y)
    loss = m.train(epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return m, loss


def rms(a: np.ndarray) -> float:
    return float(np.sqrt(np.mean(np.asarray(a, dtype=np.float64) ** 2)))


def field_errors(P: np.ndarray, Y: np.ndarray, L: np.ndarray = None) -> dict:
    """RMS errors of a far-field prediction P against the exact Y (per
    sample: potential, and field over the radial and tangential components),
    next to the plain cut-off (no far field) and, given, the LRF estimate."""
    out = {"potential_rms_exact": rms(Y[:, 0]), "field_rms_exact": rms(Y[:, 1:]),
           "potential_rms_err_surrogate": rms(P[:, 0] - Y[:, 0]),
           "field_rms_err_surrogate": rms(P[:, 1:] - Y[:, 1:]),
           "potential_rms_err_cutoff": rms(Y[:, 0]), "field_rms_err_cutoff": rms(Y[:, 1:])}
    if L is not None:
        out["potential_rms_err_lrf"] = rms(L[:, 0] - Y[:, 0])
        out["field_rms_err_lrf"] = rms(L[:, 1:] - Y[:, 1:])
    return out


def treatment_inputs(system, directory: str, name: str, restart: str, lambdas, seed: int = 999,
                     exact_cutoff: float = 200.0) -> dict:
    """qdyn inputs of one window for comparing far-field treatments, all from
    the same restart (new Maxwell velocities from `seed`): the reference
    cut-offs with LRF ("lrf"), the same without LRF ("cutoff"), and every
    cut-off beyond the sphere with LRF off, i.e. exact all-pairs
    electrostatics and LJ ("exact")."""
    import os

    from . import gensys
    common = dict(steps=1000000, output=1000000, lambdas=list(lambdas), restart=os.path.basename(restart),
                  initial_temperature=300.0, random_seed=seed)
    out = {"lrf": gensys.write_system(system, directory, name + "_lrf", **common)["inp"],
           "cutoff": gensys.write_system(system, directory, name + "_cut", lrf=False, **common)["inp"],
           "exact": gensys.write_system(system, directory, name + "_exact", lrf=False, rcpp=exact_cutoff,
                                        rcww=exact_cutoff, rcpw=exact_cutoff, rcq=exact_cutoff,
                                        **common)["inp"]}
    return out


def held_out_frames(inp: str, nframes: int, stride: int, gap: int = 5):
    """Pairs of states of an FP64 run of inp, `stride` steps apart: (x, v)
    at a list step and (x, v) `gap` steps later."""
    e = qe.Engine([inp], qe.QGPU_FP64)
    frames = []
    for _ in range(nframes):
        e.step(stride)
        a = e.state(0)
        e.step(gap)
        frames.append((a, e.state(0)))
    e.close()
    return frames


def force_errors(inputs: dict, frames, model) -> dict:
    """Gradients of the far-field treatments against exact all-pairs
    evaluation on the same coordinates (FP64 pair arithmetic throughout, so
    only the treatment of the far field differs).  For every frame pair the
    lists (and the LRF multipoles) are built at the first state and the
    forces evaluated at the second, as between two list updates of a run.
    Per treatment ("lrf", "surrogate", "cutoff"): the RMS over all gradient
    components of (gradient - exact gradient) in kcal/mol/A, that RMS
    relative to the RMS of the exact gradient, and the mean and spread of
    the potential-energy difference.  Also the far field at the
    charge-group centres of the second state (field_errors): exact (a fresh
    list build there), the surrogate's prediction and the LRF estimate of
    the list build at the first state."""
    eng = {k: qe.Engine([inputs[k]], qe.QGPU_FP64) for k in ("exact", "lrf", "cutoff")}
    eng["surrogate"] = qe.Engine([inputs["lrf"]], qe.QGPU_FP64)
    qe.set_far_model(eng["surrogate"], model)
    sq = {k: 0.0 for k in eng}
    sq_ref = 0.0
    dpot = {k: [] for k in eng}
    n = 0
    P, Y, L = [], [], []
    for (xa, va), (xb, vb) in frames:
        res = {}
        for k, e in eng.items():
            e.set_state(0, xa, va)
            e.make_pair_lists()
            e.set_state(0, xb, vb)
            e.pot_energy()
            res[k] = (e.gradient(0), e.energies(0)["potential"])
        e = eng["lrf"]
        L.append(qe.far_samples(e, 0, lrf_estimate=True)[1])
        e.make_pair_lists()
        X, Yx = qe.far_samples(e, 0)
        P.append(model.predict(X))
        Y.append(Yx)
        ge, pe = res["exact"]
        for k in eng:
            sq[k] += float(np.sum((res[k][0] - ge) ** 2))
            dpot[k].append(res[k][1] - pe)
        sq_ref += float(np.sum(ge ** 2))
        n += ge.size
    for e in eng.values():
        e.close()
    ref = float(np.sqrt(sq_ref / n))
    out = {"frames": len(frames), "gradient_rms_exact": ref}
    for k in ("lrf", "surrogate", "cutoff"):
        r = float(np.sqrt(sq[k] / n))
        out[k] = {"gradient_rms_err": r, "gradient_rms_err_rel": r / ref,
                  "potential_mean_diff": float(np.mean(dpot[k])), "potential_sd_diff": float(np.std(dpot[k]))}
    out["centre_field"] = field_errors(np.concatenate(P), np.concatenate(Y), np.concatenate(L))
    return out
