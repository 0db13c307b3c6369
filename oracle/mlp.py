"""ORACLE — test infrastructure only.

CPU restatement (numpy, FP64) of the far-field surrogate network of
qsource_amd/csrc/kernels_mlp.hip: x (4) -> SiLU(W1 x + b1) (128) ->
SiLU(W2 h1 + b2) (128) -> W3 h2 + b3 (3), inputs and outputs standardised,
loss = mean over samples and outputs of the squared error.  With bf16=True it
rounds at the points where the GPU kernels hold BF16 values (all weights,
the layer-1 activations and the layer-2 output gradient that feed the matrix
cores), so the GPU forward and backward passes can be compared with it at
FP32 accumulation accuracy.  The analytic gradient is pinned by finite
differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 4, 128, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NOUT, NH), p[OFF_B3:NPARAM])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def normalise(X, Y, norm):
    nm = np.asarray(norm, dtype=np.float64)
    xn = (np.asarray(X, dtype=np.float64) - nm[0:4]) / nm[4:8]
    yn = None if Y is None else (np.asarray(Y, dtype=np.float64) - nm[8:11]) / nm[11:14]
    return xn, yn


def forward(p, xn, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    z1 = xn @ r(W1).T + b1
    h1 = silu(z1)
    z2 = r(h1) @ r(W2).T + b2
    h2 = silu(z2)
    y = h2 @ r(W3).T + b3
    return y, (z1, h1, z2, h2)


def loss_and_grad(p, xn, yn, bf=False):
    """Mean squared error and its gradient w.r.t. the flat parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    y, (z1, h1, z2, h2) = forward(p, xn, bf)
    n = len(xn)
    e = y - yn
    loss = float(np.sum(e * e) / (3.0 * n))
    dy = 2.0 * e / (3.0 * n)
    dh2 = dy @ r(W3)
    dz2 = dh2 * dsilu(z2)
    dh1 = r(dz2) @ r(W2)
    dz1 = dh1 * dsilu(z1)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (dz1.T @ xn).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (r(dz2).T @ r(h1)).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = (dy.T @ h2).ravel()
    g[OFF_B3:NPARAM] = dy.sum(0)
    return loss, g


def far_field_exact(x, crg, groups, switch, excluded, qatom, cut_pp, cut_pw, cut_ww, nat_solute, rclrf=99.0):
    """Potential and field at every charge-group centre from the non-Q atoms of
    all far groups (switch-atom distance beyond the pair's cut-off, within
    RcLRF), as the LRF multipoles phi0/phi1 of a fresh list build hold them
    (lrf_update, src/md.f90:5896-5981).  Brute force over group pairs."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    ng = len(groups)
    cent = np.array([x[list(g)].mean(0) for g in groups])
    sw = np.array([x[s] for s in switch])
    solute = np.array([switch[k] < nat_solute for k in range(ng)])
    phi = np.zeros(ng)
    E = np.zeros((ng, 3))
    for a in range(ng):
        if excluded[switch[a]]:
            continue
        d2 = np.sum((sw - sw[a]) ** 2, axis=1)
        for b in range(ng):
            if b == a or excluded[switch[b]]:
                continue
            rc = cut_pp if (solute[a] and solute[b]) else (cut_ww if not (solute[a] or solute[b]) else cut_pw)
            if d2[b] <= rc * rc or d2[b] > rclrf * rclrf:
                continue
            for j in groups[b]:
                if qatom[j]:
                    continue
                dv = cent[a] - x[j]
                r = np.sqrt(dv @ dv)
                phi[a] += crg[j] / r
                E[a] += crg[j] * dv / r ** 3
    return cent, phi, E
