"""Benchmark: aggregate ns/day of batched FEP lambda windows on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3|c1|c1batch]

Default workload (c3): three solvated systems (TIP3P spheres of radius 18,
22 and 28 A around branched solutes of 150, 300 and 600 atoms, 12-atom Q
region, 2 FEP states), each equilibrated on the GPU and then run as 21
lambda windows (lambda = 0, 0.05, ..., 1): 63 uneven instances advanced in
one batched, hipGraph-captured step.  A step is one MD step of every
instance (pair lists every 10 steps, all forces, leapfrog/Berendsen, SHAKE,
temperature), coordinates resident in HBM, per-state Q energies recorded on
the device every 10 steps.

value = instances x steps x dt_fs x 1e-6 x 86400 / wall seconds of the
timed region (mixed precision: FP32 pair arithmetic, FP64 accumulation).
The same windows run in FP64 give the reference dG per system (Zwanzig and
BAR over the recorded frames); |ddG| = |dG_mixed - dG_fp64| is reported.
The plain FP64 engine on one window is timed as the unbatched baseline, and
the reference's serial qdyn5 on one window of the R=22 system (host cores,
1 thread) as the CPU baseline.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"f64": 78.6, "mixed": 157.3}   # MI355X vector FP64 / FP32 (published)
FLOP_COUL, FLOP_LJ, FLOP_LRF_ATOM = 24, 34, 90


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_c3(d, seed_base=3000, equil=True):
    """Writes the 63 window inputs; returns (inputs, window lambdas per
    system, cfg)."""
    from qsource_amd import equilibrate, gensys
    systems = [(18.0, 150), (22.0, 300), (28.0, 600)]
    lams = [round(0.05 * k, 2) for k in range(21)]
    inputs, groups, atoms = [], [], []
    for si, (R, ns) in enumerate(systems):
        t0 = time.time()
        s = gensys.solvated_system(radius=R, n_solute=ns, seed=seed_base + si, n_q=12, n_states=2)
        name = "s%d" % si
        rst = None
        if equil:
            rst = equilibrate.equilibrate(s, d, name, lambdas=[0.5, 0.5], seed=11 + si)
        log("system %d: R=%.0f A, %d solute atoms, %d atoms, equilibrated in %.1f s" %
            (si, R, ns, s.natom, time.time() - t0))
        group = []
        for k, lam in enumerate(lams):
            kw = dict(steps=1000000, output=1000000, lambdas=[1.0 - lam, lam],
                      initial_temperature=300.0, random_seed=101 + k)
            if rst:
                kw["restart"] = os.path.basename(rst)
            p = gensys.write_system(s, d, "%s_w%02d" % (name, k), **kw)
            inputs.append(p["inp"])
            group.append(len(inputs) - 1)
        groups.append(group)
        atoms.append(s.natom)
    cfg = {"workload": "batched_fep_windows", "systems": "R18/150, R22/300, R28/600 solute atoms",
           "atoms_per_system": atoms, "windows_per_system": 21, "instances": len(inputs),
           "q_atoms": 12, "states": 2, "stepsize_fs": 1.0, "nbcycle": 10, "energy_interval": 10}
    return inputs, groups, np.array([[1.0 - l, l] for l in lams]), cfg


def run_batch(inputs, precision, steps, warmup, iene=10):
    from qsource_amd import engine as qe
    e = qe.Engine(inputs, precision)
    e.step(warmup)
    e.set_energy_interval(iene, steps // iene + 1)
    t0 = time.perf_counter()
    e.step(steps)
    wall = time.perf_counter() - t0
    return e, wall


def dg_per_system(e, groups, lam, temperature=300.0):
    from qsource_amd import fep
    out = []
    for g in groups:
        v = [e.energy_series(k)[:, :, 0] for k in g]
        z = fep.zwanzig(lam, v, temperature)
        try:
            b = fep.bar(lam, v, temperature)["total"]
        except ValueError:
            b = float("nan")
        out.append((z["total"], b))
    return out


def water_force_flops(e, k):
    """Algorithmic FLOP of one water-force launch for instance k: reference
    ww pairs (per water pair 8 Coulomb-only + 1 Coulomb+LJ atom pairs) and
    the LRF Taylor evaluation of the water atoms."""
    ww = e.list_counts(k)["ww"] // 9
    nat = e.natom[k]
    return ww * (8 * FLOP_COUL + FLOP_LJ), ww, nat


def cpu_baseline(d, steps):
    """Reference qdyn5 (serial, one thread) on one window of the R=22 A system."""
    from oracle import REF_QDYN
    if not os.path.exists(REF_QDYN):
        return None
    from qsource_amd import gensys
    s = gensys.solvated_system(radius=22.0, n_solute=300, seed=3001, n_q=12, n_states=2)
    cd = os.path.join(d, "cpu")
    os.makedirs(cd, exist_ok=True)
    # short, tightly coupled steps keep the lattice start stable for the sample
    gensys.write_system(s, cd, "cpu", steps=steps, output=1000000, lambdas=[0.5, 0.5],
                        stepsize=0.5, temperature=100.0, bath_coupling=2.0, initial_temperature=100.0)
    r = subprocess.run([REF_QDYN, "cpu.inp"], cwd=cd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if "terminated normally" not in r.stdout:
        log("reference qdyn5 failed:\n" + r.stdout[-2000:])
        return None
    loop = None
    for line in r.stdout.splitlines():
        if "Total time of main loop" in line:
            loop = float(line.split()[-1].replace("(s)", ""))
    if not loop:
        return None
    # ns/day at the production step of 1 fs (time per step is independent of dt)
    return {"value": steps * 1.0 * 1e-6 * 86400.0 / loop, "unit": "ns/day", "cores": 1,
            "kind": "reference",
            "sample": "qdyn5 serial, one window of the R=22 A / 300-atom system (%d atoms), %d steps, "
                      "loop %.2f s, scaled to 1 fs steps" % (s.natom, steps, loop)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--skip-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU.  The lambda windows are independent and would be "
              "sharded across GPUs with no exchange; --gpus %d is not run here." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from qsource_amd import engine as qe
    out = {"metric": "aggregate ns/day over batched FEP windows at |ddG|<0.05 kcal/mol vs FP64",
           "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32-pair/fp64-accumulate", "data": "synthetic"}
    with tempfile.TemporaryDirectory() as d:
        if args.config != "c3":
            raise SystemExit("unknown config " + args.config)
        inputs, groups, lam, cfg = build_c3(d)
        n = len(inputs)
        log("mixed batch: %d instances" % n)
        em, wall = run_batch(inputs, qe.QGPU_MIXED, args.steps, args.warmup)
        value = n * args.steps * cfg["stepsize_fs"] * 1e-6 * 86400.0 / wall
        log("mixed: %.3f ms/step, %.1f ns/day aggregate" % (wall / args.steps * 1e3, value))
        dg_m = dg_per_system(em, groups, lam)
        prof = em.profile_kernels(args.profile_steps)
        dom = max((k for k in prof if k not in ("cgp_centers", "water_lists", "solute_lists")),
                  key=lambda k: prof[k])
        flops = 0
        wpairs = 0
        for k in range(n):
            f, w, _ = water_force_flops(em, k)
            flops += f
            wpairs += w
        ach = flops / (prof["water_forces"] * 1e-3) / 1e12 if prof["water_forces"] > 0 else 0.0
        roof = {"bound": "valu", "kernel": "water_forces", "achieved": ach,
                "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["mixed"],
                "traffic": None, "flop_per_launch": flops, "water_pairs_per_launch": wpairs,
                "kernel_ms": prof["water_forces"], "dominant_kernel": dom}
        em.close()
        fp64 = {}
        if not args.skip_fp64:
            log("fp64 batch")
            ef, wall64 = run_batch(inputs, qe.QGPU_FP64, args.steps, args.warmup)
            dg_f = dg_per_system(ef, groups, lam)
            ef.close()
            ddg = [abs(a[0] - b[0]) for a, b in zip(dg_m, dg_f)]
            ddg_bar = [abs(a[1] - b[1]) for a, b in zip(dg_m, dg_f)]
            fp64 = {"fp64_batched_ns_per_day": n * args.steps * 1e-6 * 86400.0 / wall64,
                    "dG_mixed_zwanzig": [a[0] for a in dg_m], "dG_fp64_zwanzig": [b[0] for b in dg_f],
                    "dG_mixed_bar": [a[1] for a in dg_m], "dG_fp64_bar": [b[1] for b in dg_f],
                    "abs_ddG_zwanzig": ddg, "abs_ddG_bar": ddg_bar,
                    "gate_ddG_lt_0.05": bool(max(ddg) < 0.05)}
            log("ddG (Zwanzig) per system: %s" % ddg)
        # plain FP64 engine, one window at a time
        ep, wallp = run_batch([inputs[groups[1][10]]], qe.QGPU_FP64, max(200, args.steps // 4), 50)
        plain = max(200, args.steps // 4) * 1e-6 * 86400.0 / wallp
        ep.close()
        cpu = cpu_baseline(d, args.cpu_steps) if args.cpu_steps > 0 else None
    cfg["precision"] = "mixed"
    out.update({"value": value, "ms_per_step": wall / args.steps * 1e3, "config": cfg,
                "roofline": roof, "cpu_baseline": cpu, "kernel_ms": prof,
                "plain_fp64_single_window_ns_per_day": plain,
                "speedup_vs_plain_fp64": value / plain})
    out.update(fp64)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
