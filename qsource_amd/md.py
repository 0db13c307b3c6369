"""Runs qdyn inputs on the GPU engine as md_run does (src/md.f90:6471-6704)
and writes qdyn's own output files, so runs chain like qdyn's:

* trajectory (DCD) every `trajectory` steps of [intervals]: the frame of
  loop step istep (istep > 0, istep % itrj == 0) holds the coordinates after
  that step's update, plus a last frame after the final step when nsteps is
  a multiple of itrj (src/md.f90:6625-6628, 6680-6683);
* energy file every `energy` steps (put_ene at istep % iene == 0, istep > 0,
  src/md.f90:6629-6632): the frames the engine records on the device;
* the final restart (write_xfin, src/md.f90:18705-18736): x, v and the
  water-polarisation theta_corr after the last step.

A batch runs many inputs in one set of launches per step; they must agree
on steps and output intervals.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import engine as qe
from . import fep


def read_settings(path: str) -> Dict:
    """[MD] steps, [intervals] trajectory/energy, [files] final, trajectory,
    energy (paths resolved against the input's directory) and [lambdas] of
    a qdyn input (prmfile layout: sections, '!' comments)."""
    sec: Dict[str, List[str]] = {}
    cur = None
    for raw in open(path):
        line = raw.split("!")[0].strip()
        if not line or line[0] in "#*":
            continue
        if line.startswith("["):
            cur = line[1:line.index("]")].strip().lower()
            sec.setdefault(cur, [])
            continue
        if cur is not None:
            sec[cur].append(line)

    def get(s, key, default=None):
        for l in sec.get(s, []):
            t = l.split()
            if len(t) >= 2 and t[0].lower() == key:
                return t[1]
        return default

    d = os.path.dirname(os.path.abspath(path))
    res = lambda f: None if f is None else (f if os.path.isabs(f) else os.path.join(d, f))
    lam = [float(v) for l in sec.get("lambdas", []) for v in l.split()]
    return dict(steps=int(get("md", "steps", 0)), itrj=int(get("intervals", "trajectory", 0)),
                iene=int(get("intervals", "energy", 0)), final=res(get("files", "final")),
                trajectory=res(get("files", "trajectory")), energy=res(get("files", "energy")),
                lambdas=lam)


def trajectory_steps(nsteps: int, itrj: int) -> List[int]:
    """Numbers of completed steps at which md_run writes a trajectory frame."""
    if itrj <= 0:
        return []
    out = [m * itrj + 1 for m in range(1, nsteps) if m * itrj <= nsteps - 1]
    if nsteps % itrj == 0:
        out.append(nsteps)      # with itrj = 1 the last state is written twice, as qdyn does
    return out


def run(inputs: Sequence[str], precision: int = qe.QGPU_MIXED,
        far_model: Optional["qe.FarFieldMLP"] = None, version: str = "qgpu") -> List[Dict[str, float]]:
    """Runs every input for its [MD] steps on one GPU and writes its files;
    returns the final energies of each instance (the closing list update and
    pot_energy of md_run)."""
    st = [read_settings(p) for p in inputs]
    for key in ("steps", "itrj", "iene"):
        if len({s[key] for s in st}) != 1:
            raise ValueError("batched inputs must share [MD] steps and output intervals (%s)" % key)
    nsteps, itrj, iene = st[0]["steps"], st[0]["itrj"], st[0]["iene"]
    e = qe.Engine(list(inputs), precision)
    try:
        if far_model is not None:
            qe.set_far_model(e, far_model)
        if iene > 0:
            e.set_energy_interval(iene, max(1, nsteps // iene))
        trj = [qe.Trajectory(p) if itrj > 0 else None for p in inputs]
        done = 0
        for target in trajectory_steps(nsteps, itrj):
            if target > done:
                e.step(target - done)
            done = target
            for k, t in enumerate(trj):
                t.write(e.state(k)[0])
        if nsteps > done:
            e.step(nsteps - done)
        for t in trj:
            if t is not None:
                t.close()
        for k, s in enumerate(st):
            if iene > 0 and s["energy"]:
                frames = e.energy_series(k)
                od = e.offdiag_series(k)
                ij = e.offdiag_states(k)
                offd = None
                if od.shape[1] > 0:
                    offd = np.concatenate([np.broadcast_to(ij, od.shape[:2] + (2,)), od], axis=2)
                fep.write_energy_file(s["energy"], s["lambdas"], frames, version=version, offd=offd)
            if s["final"]:
                x, v = e.state(k)
                qe.write_restart(s["final"], x, v, e.theta_corr(k))
        e.make_pair_lists()
        e.pot_energy()
        return [e.energies(k) for k in range(len(inputs))]
    finally:
        e.close()
