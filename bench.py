"""Benchmark: aggregate ns/day of qdyn MD windows on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c1batch]

A step is one MD step (list update every NBcycle=10 steps, all forces,
leapfrog/Berendsen, SHAKE, temperature) of every instance in the batch, with
coordinates resident in HBM.  value = instances x steps x dt_fs x 1e-6 x
86400 / wall seconds.  The CPU baseline is the reference's serial qdyn5
(oracle/_ref) on the same system for a bounded number of steps, one core.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def make_workload(config, d, steps):
    from qsource_amd import gensys
    if config in ("c1", "c1batch"):
        s = gensys.water_droplet(15.0, 1001)
        n = 1 if config == "c1" else 64
        inputs = []
        for k in range(n):
            p = gensys.write_system(s, d, "c1_%02d" % k, steps=steps, output=1000000,
                                    random_seed=1 + k)
            inputs.append(p["inp"])
        return inputs, {"workload": "tip3p_droplet_r15", "atoms": s.natom, "waters": s.nwat,
                        "instances": n, "stepsize_fs": 1.0, "nbcycle": 10}
    raise SystemExit("unknown config " + config)


def cpu_baseline(inp_dir, name, steps):
    """Reference qdyn5 (serial, 1 thread) on the same system."""
    from oracle import REF_QDYN
    if not os.path.exists(REF_QDYN):
        return None
    t0 = time.time()
    r = subprocess.run([REF_QDYN, name + ".inp"], cwd=inp_dir, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    wall = time.time() - t0
    if "terminated normally" not in r.stdout:
        return None
    loop = wall
    for line in r.stdout.splitlines():
        if "Total time of main loop" in line:
            loop = float(line.split()[-1].replace("(s)", ""))
    return {"value": steps * 1.0 * 1e-6 * 86400.0 / loop, "unit": "ns/day", "cores": 1,
            "kind": "reference", "sample": "%d steps of one window (qdyn5 serial, loop time %.2f s)" % (steps, loop)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--config", default="c1batch")
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the lambda windows of this workload are independent and "
              "would be sharded across GPUs with no exchange; --gpus %d is not run." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from qsource_amd import engine as qe
    with tempfile.TemporaryDirectory() as d:
        inputs, cfg = make_workload(args.config, d, args.steps + args.warmup)
        prec = qe.QGPU_MIXED if args.precision == "mixed" else qe.QGPU_FP64
        e = qe.Engine(inputs, prec)
        e.step(args.warmup)
        t0 = time.perf_counter()
        e.step(args.steps)
        wall = time.perf_counter() - t0
        ms_dev = e.last_step_ms()
        n = len(inputs)
        value = n * args.steps * cfg["stepsize_fs"] * 1e-6 * 86400.0 / wall
        cpu = None
        if os.path.exists(os.path.join(ROOT, "oracle", "_ref", "qdyn5")):
            cd = os.path.join(d, "cpu")
            os.makedirs(cd, exist_ok=True)
            from qsource_amd import gensys
            s = gensys.water_droplet(15.0, 1001)
            gensys.write_system(s, cd, "cpu", steps=args.cpu_steps, output=1000000)
            cpu = cpu_baseline(cd, "cpu", args.cpu_steps)
    cfg["precision"] = args.precision
    out = {
        "metric": "aggregate ns/day over batched FEP windows at |ddG|<0.05 kcal/mol vs FP64",
        "value": value, "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "device_ms_per_step": ms_dev,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-pair/fp64-accumulate" if args.precision == "mixed" else "f64",
        "data": "synthetic", "config": cfg, "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
