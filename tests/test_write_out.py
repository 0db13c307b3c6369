"""write_out's printed energy summary (src/md.f90:18545-18691) against the
reference's own printed text, byte for byte: the blocks reference qdyn5
printed at every step and at the end of a 12-step run (4 FEP states with
Q-atom types, off-diagonal couplings, monitored group pairs, LRF, solute
and water), kept by tests/golden/make_golden.py.  The values come from the
oracle, which reproduces the reference's energies; the layout comes from
qsource_amd.md.write_out, which the run loop uses with the engine's values.
"""
import hashlib
import os

import numpy as np

from oracle import Oracle
from qsource_amd import gensys, md

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAME = "c13_r12_q8s4_writeout_12"
SYSARGS = dict(radius=12.0, n_solute=40, seed=3101, n_q=8, n_states=4, q_types=True, softcore=20.0, n_offdiag=3,
               impropers=1, monitor=True)
INPARGS = dict(steps=12, lambdas=[0.4, 0.3, 0.2, 0.1], stepsize=0.5, temperature=50.0, bath_coupling=2.0,
               initial_temperature=50.0)


def test_write_out_text_matches_reference(tmp_path):
    g = np.load(os.path.join(GOLD, NAME + ".npz"), allow_pickle=False)
    blocks = [str(b) for b in g["blocks"]]
    s = gensys.solvated_system(**SYSARGS)
    paths = gensys.write_system(s, str(tmp_path), NAME, output=1, temp_interval=1, **INPARGS)
    assert hashlib.sha256(open(paths["top"], "rb").read()).hexdigest() == str(g["top_sha"])
    nsteps = INPARGS["steps"]
    assert len(blocks) == nsteps + 1
    ijkl = np.array([o[:4] for o in s.offdiag], dtype=float)
    kw = dict(nwat=s.nwat, nqat=8, nat_solute=s.nat_solute, use_lrf=True)
    o = Oracle(paths["inp"])
    offd = mon = None
    for k in range(nsteps):
        # md_run: pot_energy, nonbond_monitor and offdiag on the step's
        # starting coordinates, then the update; write_out after it
        mon = o.monitor()
        offd = np.hstack([ijkl, o.offdiag()])
        o.run(1)
        text = md.write_out(o.energies(), o.qenergies(), INPARGS["lambdas"], step=k, final=False,
                            offd=offd, monitor=mon, **kw)
        assert text == blocks[k], "step %d:\n%s\n---- reference ----\n%s" % (k, text, blocks[k])
    o.final_energy()
    text = md.write_out(o.energies(), o.qenergies(), INPARGS["lambdas"], step=nsteps, final=True,
                        offd=offd, monitor=mon, **kw)
    assert text == blocks[-1], "final:\n%s\n---- reference ----\n%s" % (text, blocks[-1])
    o.close()


def test_fortran_fields():
    # overflowing F fields print as asterisks; A fields right-justify
    assert md._f(123456.789, 8, 2) == "********"
    assert md._f(-0.001, 10, 2) == "     -0.00"
    assert md._a("el", 10) == "        el"
    assert md._centered("Monitoring selected groups of nonbonded interactions") == \
        "=" * 13 + " Monitoring selected groups of nonbonded interactions " + "=" * 12
