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
Accuracy gate: every recorded frame of the mixed run is also evaluated with
FP64 pair arithmetic on the same coordinates; |ddG| = |dG(mixed dU) -
dG(FP64 dU)| over identical frames (Zwanzig and BAR) must stay below 0.05
kcal/mol.  The same windows run as an FP64 batch give independent
trajectories; their dG difference is reported with its bootstrap standard
error (it is dominated by sampling noise at these run lengths).
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


def measured_traffic(name="k_water_forces_c3_63win_mixed_pmc_traffic.md"):
    """HBM bytes per launch of the FP32 water-force kernel, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (two counter runs,
    calibrated on a plain copy); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


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
    # launch shapes picked by the engine's tuner on this batch (the first
    # step would do the same); kept on the engine for the report
    e.tuned_shapes = e.autotune(reps=3)
    e.step(warmup)
    e.set_energy_interval(iene, steps // iene + 1)
    t0 = time.perf_counter()
    e.step(steps)
    wall = time.perf_counter() - t0
    return e, wall


def series(e, groups, fp64=False):
    """Per system, per window: (nframes, nstates) EQ totals of the frames."""
    return [[e.energy_series(k, fp64=fp64)[:, :, 0] for k in g] for g in groups]


def dg_of(ser, lam, temperature=300.0):
    """(Zwanzig, BAR) dG per system."""
    from qsource_amd import fep
    out = []
    for v in ser:
        z = fep.zwanzig(lam, v, temperature)["total"]
        try:
            b = fep.bar(lam, v, temperature)["total"]
        except ValueError:
            b = float("nan")
        out.append((z, b))
    return out


def zwanzig_stderr(ser, lam, temperature=300.0, block=10, nboot=200, seed=7):
    """Standard error of the Zwanzig dG per system: bootstrap over blocks of
    `block` consecutive frames within every window (frames are correlated)."""
    from qsource_amd import fep
    rng = np.random.default_rng(seed)
    out = []
    for v in ser:
        vals = []
        for _ in range(nboot):
            w = []
            for e in v:
                nb = max(1, len(e) // block)
                pick = rng.integers(0, nb, nb)
                w.append(np.concatenate([e[b * block:(b + 1) * block] for b in pick]))
            vals.append(fep.zwanzig(lam, w, temperature)["total"])
        out.append(float(np.std(vals)))
    return out


def force_error(inputs, states):
    """Mixed vs FP64 gradient and energies on identical coordinates (one
    window per system): relative RMS force error and per-term energy
    differences."""
    from qsource_amd import engine as qe
    res = []
    engines = [qe.Engine(inputs, p) for p in (qe.QGPU_FP64, qe.QGPU_MIXED)]
    for e in engines:
        for k, (x, v) in enumerate(states):
            e.set_state(k, x, v)
        e.make_pair_lists()
        e.pot_energy()
    for k in range(len(inputs)):
        d64, d32 = engines[0].gradient(k), engines[1].gradient(k)
        e64, e32 = engines[0].energies(k), engines[1].energies(k)
        rms = float(np.sqrt(np.mean((d32 - d64) ** 2)) / np.sqrt(np.mean(d64 ** 2)))
        terms = {t: float(e32[t] - e64[t]) for t in ("ww_el", "ww_vdw", "pw_el", "pw_vdw", "pp_el",
                                                      "pp_vdw", "qx_el", "qx_vdw", "lrf", "potential")}
        res.append({"force_rms_rel": rms, "energy_diff": terms})
    for e in engines:
        e.close()
    return res


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
    ap.add_argument("--cpu-steps", type=int, default=1000)
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
        tuned = {k: {"mapping": ["grid2d", "instance_map", "xcd_map"][m], "block": b, "prefetch": pf,
                     "waves_per_simd": (4 if oc else 6) if k == "water_forces" else None,
                     "ms_chosen": best, "ms_default": dflt}
                 for k, (m, b, pf, oc, best, dflt) in em.tuned_shapes.items()}
        log("tuned launch shapes: %s" % tuned)
        value = n * args.steps * cfg["stepsize_fs"] * 1e-6 * 86400.0 / wall
        log("mixed: %.3f ms/step, %.1f ns/day aggregate" % (wall / args.steps * 1e3, value))
        ser_m = series(em, groups)
        ser_m64 = series(em, groups, fp64=True)
        dg_m = dg_of(ser_m, lam)
        dg_m64 = dg_of(ser_m64, lam)
        ddg_frames = [abs(a[0] - b[0]) for a, b in zip(dg_m, dg_m64)]
        ddg_frames_bar = [abs(a[1] - b[1]) for a, b in zip(dg_m, dg_m64)]
        du_err = max(float(np.abs((a @ np.array([-1.0, 1.0])) - (b @ np.array([-1.0, 1.0]))).max())
                     for sm, sf in zip(ser_m, ser_m64) for a, b in zip(sm, sf))
        log("ddG on identical frames (mixed vs FP64 dU): %s" % ddg_frames)
        mid = [g[len(g) // 2] for g in groups]
        states = [em.state(k) for k in mid]
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
                "traffic": measured_traffic(), "traffic_source": "profiles/k_water_forces_c3_63win_mixed_pmc_traffic.md",
                "flop_per_launch": flops, "water_pairs_per_launch": wpairs,
                "kernel_ms": prof["water_forces"], "dominant_kernel": dom}
        em.close()
        ferr = force_error([inputs[k] for k in mid], states)
        log("force RMS error (mixed vs FP64, same coordinates): %s" % [f["force_rms_rel"] for f in ferr])
        acc = {"dG_mixed_zwanzig": [a[0] for a in dg_m], "dG_mixed_bar": [a[1] for a in dg_m],
               "dG_fp64_same_frames_zwanzig": [b[0] for b in dg_m64],
               "dG_fp64_same_frames_bar": [b[1] for b in dg_m64],
               "abs_ddG_same_frames_zwanzig": ddg_frames, "abs_ddG_same_frames_bar": ddg_frames_bar,
               "max_abs_dU_error_kcal": du_err,
               "gate_ddG_lt_0.05": bool(max(ddg_frames + ddg_frames_bar) < 0.05),
               "force_error_one_window_per_system": ferr}
        if not args.skip_fp64:
            log("fp64 batch (independent trajectories)")
            ef, wall64 = run_batch(inputs, qe.QGPU_FP64, args.steps, args.warmup)
            ser_f = series(ef, groups)
            ef.close()
            dg_f = dg_of(ser_f, lam)
            se = [float(np.hypot(a, b)) for a, b in zip(zwanzig_stderr(ser_m, lam), zwanzig_stderr(ser_f, lam))]
            acc.update({"fp64_batched_ns_per_day": n * args.steps * 1e-6 * 86400.0 / wall64,
                        "dG_fp64_traj_zwanzig": [b[0] for b in dg_f], "dG_fp64_traj_bar": [b[1] for b in dg_f],
                        "abs_ddG_traj_zwanzig": [abs(a[0] - b[0]) for a, b in zip(dg_m, dg_f)],
                        "stderr_ddG_traj_zwanzig": se})
            log("ddG of independent trajectories (Zwanzig) per system: %s +- %s" %
                (acc["abs_ddG_traj_zwanzig"], se))
        # plain FP64 engine, one window at a time
        ep, wallp = run_batch([inputs[groups[1][10]]], qe.QGPU_FP64, max(200, args.steps // 4), 50)
        plain = max(200, args.steps // 4) * 1e-6 * 86400.0 / wallp
        ep.close()
        cpu = cpu_baseline(d, args.cpu_steps) if args.cpu_steps > 0 else None
    cfg["precision"] = "mixed"
    out.update({"value": value, "ms_per_step": wall / args.steps * 1e3, "config": cfg,
                "roofline": roof, "cpu_baseline": cpu, "kernel_ms": prof, "autotune": tuned,
                "plain_fp64_single_window_ns_per_day": plain,
                "speedup_vs_plain_fp64": value / plain})
    out["accuracy"] = acc
    print(json.dumps(out))


if __name__ == "__main__":
    main()
