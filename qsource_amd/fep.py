"""Free-energy estimators over FEP windows, and qdyn's binary energy file.

* ``write_energy_file`` / ``read_energy_file``: the energy-file layout of
  qdyn (header record of prep_sim_version, src/md.f90:16989-17017; one
  record per state per frame as put_ene writes it, src/nrgy.f90:113-142,
  followed by the off-diagonal record), so qfep can read this engine's files.
* ``zwanzig``: qfep's forward/reverse exponential averaging over adjacent
  windows and its <dG> = sum of (dGf - dGr)/2 (src/qfep.f90:601-662).
* ``bar``: Bennett's acceptance ratio for each adjacent window pair (not in
  the reference; solved by bisection on the standard self-consistency
  equation with equal sample counts handled through the log-ratio term).
"""
from __future__ import annotations

import struct
from typing import Optional, List, Sequence, Tuple

import numpy as np

GAS_CONSTANT = 0.0019872  # kcal/(mol K), qfep's rt = gas constant * T


def _rec(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_energy_file(path: str, lambdas: Sequence[float], frames: np.ndarray,
                      version: str = "qgpu", offd: Optional[np.ndarray] = None) -> None:
    """frames: (nframes, nstates, 14) rows of
    total q(bond angle torsion improper) qx(el vdw) qq(el vdw) qp(el vdw)
    qw(el vdw) restraint, for one energy array (no group contributions);
    offd: optional (nframes, noffd, 4) off-diagonal entries (state i, state
    j, Hij, rkl), written as OFFDIAG_SAVE records (src/nrgy.f90:73-78)."""
    nstates = len(lambdas)
    # canary 1337, arrays=1, totresid=1, types(1)=NOGC(=3), numres(1)=1,
    # resid(1)=-1, gcnum(1)=-1, version char(80)
    head = struct.pack("<iiiiiii", 1337, 1, 1, 3, 1, -1, -1) + version.encode()[:80].ljust(80)
    out = [_rec(head)]
    for k, f in enumerate(frames):
        for s in range(nstates):
            row = f[s]
            out.append(_rec(struct.pack("<i", s + 1) + struct.pack("<d", lambdas[s]) +
                            np.asarray(row, dtype="<f8").tobytes()))
        if offd is None or len(offd) == 0:
            out.append(_rec(b""))  # no off-diagonal elements
        else:
            out.append(_rec(b"".join(struct.pack("<iidd", int(o[0]), int(o[1]), float(o[2]), float(o[3]))
                                     for o in offd[k])))
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def read_energy_frames(path: str, nstates: int):
    """All frames of a qdyn energy file (put_ene, src/nrgy.f90:113-142):
    returns (lambdas (nstates,), rows (nframes, nstates, 14), offd (nframes,
    noffd, 4) with state i, state j, Hij, rkl)."""
    data = open(path, "rb").read()
    recs: List[bytes] = []
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        recs.append(data[p + 4:p + 4 + n])
        p += 8 + n
    body = recs[1:]
    per = nstates + 1
    nfr = len(body) // per
    lam = np.zeros(nstates)
    rows = np.zeros((nfr, nstates, 14))
    noffd = len(body[nstates]) // 24 if nfr else 0
    offd = np.zeros((nfr, noffd, 4))
    for k in range(nfr):
        for s in range(nstates):
            r = body[k * per + s]
            lam[s] = struct.unpack_from("<d", r, 4)[0]
            rows[k, s] = np.frombuffer(r, dtype="<f8", count=14, offset=12)
        ro = body[k * per + nstates]
        for m in range(noffd):
            i, j, h, rkl = struct.unpack_from("<iidd", ro, 24 * m)
            offd[k, m] = (i, j, h, rkl)
    return lam, rows, offd


def read_energy_file(path: str, nstates: int) -> Tuple[np.ndarray, np.ndarray]:
    """Returns (lambdas (nstates,), totals (nframes, nstates))."""
    data = open(path, "rb").read()
    recs: List[bytes] = []
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        recs.append(data[p + 4:p + 4 + n])
        p += 8 + n
    body = recs[1:]
    per = nstates + 1
    nfr = len(body) // per
    lam = np.zeros(nstates)
    tot = np.zeros((nfr, nstates))
    for k in range(nfr):
        for s in range(nstates):
            r = body[k * per + s]
            lam[s] = struct.unpack_from("<d", r, 4)[0]
            tot[k, s] = struct.unpack_from("<d", r, 12)[0]
    return lam, tot


def zwanzig(lambdas: np.ndarray, v: List[np.ndarray], temperature: float,
            nskip: int = 0) -> dict:
    """qfep's exponential averaging.  lambdas: (nwin, nstates) lambda vectors
    of the windows; v[k]: (npts, nstates) EQ totals sampled in window k."""
    rt = GAS_CONSTANT * temperature
    n = len(v)
    dgf = np.zeros(n)
    dgr = np.zeros(n)
    for k in range(n - 1):
        e = v[k][nskip:]
        dv = e @ lambdas[k + 1] - e @ lambdas[k]
        dgf[k] = -rt * np.log(np.mean(np.exp(-dv / rt)))
    for k in range(n - 1, 0, -1):
        e = v[k][nskip:]
        dv = e @ lambdas[k - 1] - e @ lambdas[k]
        dgr[k] = -rt * np.log(np.mean(np.exp(-dv / rt)))
    dg = np.zeros(n)
    for k in range(1, n):
        dg[k] = dg[k - 1] + 0.5 * (dgf[k - 1] - dgr[k])
    return {"dGf": dgf, "dGr": dgr, "dG": dg, "total": float(dg[-1])}


def bar(lambdas: np.ndarray, v: List[np.ndarray], temperature: float, nskip: int = 0,
        tol: float = 1e-10) -> dict:
    """Bennett acceptance ratio between adjacent windows; returns per-pair
    dG and the sum."""
    beta = 1.0 / (GAS_CONSTANT * temperature)
    n = len(v)
    dgs = np.zeros(n - 1)
    for k in range(n - 1):
        ef = v[k][nskip:]
        er = v[k + 1][nskip:]
        wf = beta * (ef @ lambdas[k + 1] - ef @ lambdas[k])     # forward work
        wr = beta * (er @ lambdas[k] - er @ lambdas[k + 1])     # reverse work
        nf, nr = len(wf), len(wr)
        m = np.log(nf / nr)

        def f(df):  # BAR implicit equation in units of kT
            a = np.sum(1.0 / (1.0 + np.exp(m + wf - df)))
            b = np.sum(1.0 / (1.0 + np.exp(-m + wr + df)))
            return a - b

        lo, hi = -50.0, 50.0
        # bracket around the mean works
        c = 0.5 * (np.mean(wf) - np.mean(wr))
        lo, hi = c - 50.0, c + 50.0
        flo, fhi = f(lo), f(hi)
        if flo * fhi > 0:
            raise ValueError("BAR: root not bracketed")
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            fm = f(mid)
            if fm == 0 or hi - lo < tol:
                break
            if (fm > 0) == (fhi > 0):
                hi, fhi = mid, fm
            else:
                lo, flo = mid, fm
        dgs[k] = 0.5 * (lo + hi) / beta
    return {"dG": dgs, "total": float(dgs.sum())}
