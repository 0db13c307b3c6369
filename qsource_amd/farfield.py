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
