"""Staged equilibration of a generated system on the GPU engine.

A lattice-built droplet releases the water orientation energy within tens of
femtoseconds.  Like the reference's own equilibration protocol (short steps
and tight bath coupling first, tests/basic_tests/eq1.inp-eq5.inp), the
stages below heat the system gradually; each stage is an ordinary qdyn input
run through the engine, and each hands over a qdyn restart file (x, v and
the polarisation-shell theta_corr) to the next.
"""
from __future__ import annotations

import os
from typing import Callable, Optional, Sequence, Tuple

from . import gensys
from .engine import QGPU_FP64, Engine

# (steps, stepsize fs, temperature K, bath coupling fs)
DEFAULT_STAGES: Tuple[Tuple[int, float, float, float], ...] = (
    (1500, 0.2, 10.0, 1.0),
    (1000, 0.5, 100.0, 5.0),
    (1000, 1.0, 300.0, 10.0),
    (1000, 1.0, 300.0, 100.0),
)


def equilibrate(system: gensys.System, directory: str, name: str,
                stages: Sequence[Tuple[int, float, float, float]] = DEFAULT_STAGES,
                lambdas: Optional[Sequence[float]] = None, seed: int = 1,
                precision: int = QGPU_FP64, progress: Optional[Callable[[str], None]] = None) -> str:
    """Runs the stages and returns the path of the final restart file (in
    `directory`, named `name`_eq.re); `progress` receives a line per stage."""
    rst = None
    for k, (n, dt, temp, tau) in enumerate(stages):
        kw = dict(steps=n, stepsize=dt, temperature=temp, bath_coupling=tau, output=1000000)
        if lambdas is not None:
            kw["lambdas"] = list(lambdas)
        if rst is None:
            kw.update(initial_temperature=temp, random_seed=seed)
        else:
            kw.update(initial_temperature=None, restart=os.path.basename(rst))
        paths = gensys.write_system(system, directory, "%s_eq%d" % (name, k), **kw)
        e = Engine([paths["inp"]], precision)
        e.step(n)
        x, v = e.state(0)
        tc = e.theta_corr(0)
        rst = os.path.join(directory, "%s_eq%d.re" % (name, k))
        gensys.write_restart(rst, x, v, tc if len(tc) else None)
        e.close()
        if progress:
            progress("%s: equilibration stage %d (%d steps of %.1f fs at %.0f K) done" % (name, k, n, dt, temp))
    final = os.path.join(directory, name + "_eq.re")
    os.replace(rst, final)
    return final
