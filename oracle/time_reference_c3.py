"""Times the reference's serial qdyn5 (oracle/_ref/qdyn5, built from the
reference sources by oracle/Makefile.ref) on windows of the batched FEP
workload's middle system (TIP3P sphere of R = 22 A around a 300-atom
solute, 12 Q atoms, 2 states), on the host's cores.

The system is built exactly as bench.py builds it and equilibrated with the
same staged protocol (qsource_amd.equilibrate.DEFAULT_STAGES), each stage run
by qdyn5 itself.  Then one window alone, and `--windows` lambda windows
(0.5 +- k * 0.05) at once, one qdyn5 process per window, run `--steps`
production steps (1 fs, 300 K, restart from the equilibration); each
window's rate comes from the "Total time of main loop" qdyn5 prints.  Prints
one `result:` JSON line.

Usage:  python oracle/time_reference_c3.py [--steps 1000] [--windows 8]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from qsource_amd import equilibrate, gensys  # noqa: E402

QDYN = os.path.join(ROOT, "oracle", "_ref", "qdyn5")
LOOP = re.compile(r"Total time of main loop:\s+([0-9.]+)")


def run_qdyn(inp, cwd):
    r = subprocess.run([QDYN, os.path.basename(inp)], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if "terminated normally" not in r.stdout:
        sys.exit("qdyn5 failed on %s:\n%s" % (inp, r.stdout[-2000:]))
    return r.stdout


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--windows", type=int, default=8)
    args = ap.parse_args()
    if not os.path.exists(QDYN):
        sys.exit("reference qdyn5 not built: make -f oracle/Makefile.ref")
    with tempfile.TemporaryDirectory() as d:
        # bench.py's middle C3 system (seed_base 3000 + system index 1)
        s = gensys.solvated_system(radius=22.0, n_solute=300, seed=3001, n_q=12, n_states=2)
        rst = None
        for k, (n, dt, temp, tau) in enumerate(equilibrate.DEFAULT_STAGES):
            kw = dict(steps=n, stepsize=dt, temperature=temp, bath_coupling=tau, output=1000000,
                      lambdas=[0.5, 0.5])
            if rst is None:
                kw.update(initial_temperature=temp, random_seed=12)
            else:
                kw.update(initial_temperature=None, restart=rst)
            p = gensys.write_system(s, d, "eq%d" % k, **kw)
            run_qdyn(p["inp"], d)
            rst = "eq%d.re" % k
            print("equilibration stage %d done" % k, flush=True)
        def windows(count):
            procs = []
            for w in range(count):
                lam = 0.5 + 0.05 * (w - count // 2)
                p = gensys.write_system(s, d, "w%d" % w, steps=args.steps, output=1000000,
                                        lambdas=[1.0 - lam, lam], initial_temperature=300.0,
                                        random_seed=101 + w, restart=rst)
                procs.append(subprocess.Popen([QDYN, os.path.basename(p["inp"])], cwd=d,
                                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
            secs = []
            for pr in procs:
                out = pr.communicate()[0]
                m = LOOP.search(out)
                if "terminated normally" not in out or not m:
                    sys.exit("qdyn5 window failed:\n" + out[-2000:])
                secs.append(float(m.group(1)))
            return secs
        one = windows(1)
        print("one window done", flush=True)
        secs = windows(args.windows)
        per = [args.steps * 1e-6 * 86400.0 / t for t in secs]
        res = {"value": sum(per), "unit": "ns/day", "cores": args.windows, "kind": "reference",
               "per_window_ns_per_day": per, "main_loop_seconds": secs,
               "one_window_alone_ns_per_day": args.steps * 1e-6 * 86400.0 / one[0],
               "sample": "reference serial qdyn5, %d windows of the R=22 A / 300-atom system at once, one "
                         "process per core, %d steps each at 1 fs / 300 K from a qdyn5-equilibrated restart"
                         % (args.windows, args.steps),
               "natom": s.natom}
        print("result: " + json.dumps(res))


if __name__ == "__main__":
    main()
