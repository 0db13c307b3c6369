"""Benchmark: aggregate ns/day of batched FEP lambda windows on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3|c4|c5]
                [--accuracy-steps N [--accuracy-control]]

Default workload (c3): three solvated systems (TIP3P spheres of radius 18,
22 and 28 A around branched solutes of 150, 300 and 600 atoms, 12-atom Q
region, 2 FEP states), each equilibrated on the GPU and then run as 21
lambda windows (lambda = 0, 0.05, ..., 1): 63 uneven instances advanced in
one batched, hipGraph-captured step.  A step is one MD step of every
instance (pair lists every 10 steps, all forces, leapfrog/Berendsen, SHAKE,
temperature), coordinates resident in HBM, per-state Q energies recorded on
the device every 10 steps.  Config c5: one EVB-style system (R = 40 A
sphere, 5000-atom solute, 24 Q atoms, 4 states with Q-atom types and
off-diagonal couplings) as 16 mapping windows.  Config c4: one system (R =
30 A sphere, 2000-atom solute) as 11 lambda windows with the learned far
field (trained on the GPU on 200 frames of exact far-field targets), its
gradient error against exact all-pairs evaluation next to LRF and the plain
cut-off, and dG over the windows with each far-field treatment.

value = instances x steps x dt_fs x 1e-6 x 86400 / wall seconds of the
timed region (mixed precision: FP32 pair arithmetic, FP64 accumulation).
Energy recording, launch tuning and at least one whole captured list cycle
happen before the timed region, so it replays the captured step graph.

Plain FP64 engine path: the same windows run one at a time in FP64; its
aggregate rate is timed on one window of every system (the simulated time
over the summed wall time), and speedup_vs_plain_fp64 = value / that rate.

Accuracy: every recorded frame of the mixed run is also re-evaluated with
FP64 pair arithmetic on the same coordinates, which gives the precision
error of dU and dG alone (gate_ddG_lt_0.05_precision_only); the per-term
force errors (mixed vs FP64 on identical coordinates) are reported per
system.  The dG of each system from the mixed batch is also compared with
the dG of the same windows run as an FP64 batch (independent trajectories,
Zwanzig and BAR), with a block-bootstrap standard error.  That gate
(gate_ddG_lt_0.05_trajectories) is established only where the standard
error is below 0.02 kcal/mol; a short run cannot reach that, and then it is
taken from the committed long-run summary of the same workload (profiles/,
written by --accuracy-steps, with --accuracy-control for the spread between
runs of one precision).
CPU baseline: the FP64 CPU oracle (an order-faithful restatement of serial
qdyn) on this host's cores, one window per core of the R = 22 A system, from
the same inputs and restart as the GPU batch.  The reference qdyn5 itself
(oracle/_ref, compiled from the reference sources) stays in the build
container; its timing there, for 1 and 8 windows at once, is read from the
committed summary (cpu_reference_build_host).
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
FLOP_COUL, FLOP_LJ, FLOP_LRF_UPDATE, FLOP_LRF_ATOM = 24, 34, 60, 90
IENE = 10
SE_GATE = 0.02        # kcal/mol: stderr below which the trajectory ddG gate is established
DDG_GATE = 0.05
PROFILES = os.path.join(ROOT, "profiles")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_value(name, key):
    """A `key: value` line of a committed profile summary (None if absent)."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return line.split(":", 1)[1].strip()
    return None


def profile_json(name):
    """The JSON object on the `result:` line of a committed summary."""
    v = profile_value(name, "result")
    return json.loads(v) if v else None


# ---------------------------------------------------------------------------
# workloads

def build_c3(d, seed_base=3000, equil=True):
    """The 63 window inputs; returns (inputs, window indices per system,
    lambda vectors per window, config)."""
    from qsource_amd import equilibrate, gensys
    systems = [(18.0, 150), (22.0, 300), (28.0, 600)]
    lams = [round(0.05 * k, 2) for k in range(21)]
    inputs, groups, atoms = [], [], []
    for si, (R, ns) in enumerate(systems):
        t0 = time.time()
        s = gensys.solvated_system(radius=R, n_solute=ns, seed=seed_base + si, n_q=12, n_states=2)
        name = "s%d" % si
        rst = equilibrate.equilibrate(s, d, name, lambdas=[0.5, 0.5], seed=11 + si, progress=log) if equil else None
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
           "q_atoms": 12, "states": 2, "stepsize_fs": 1.0, "nbcycle": 10, "energy_interval": IENE}
    lam = np.array([[1.0 - l, l] for l in lams])
    return inputs, groups, [lam] * len(systems), cfg


def build_c5(d, equil=True, windows=16):
    """One EVB-style system (R = 40 A, 5000-atom solute, 24 Q atoms, 4 states
    with Q-atom types, off-diagonal couplings) as `windows` mapping windows
    moving state 1 -> 2 -> 3 -> 4."""
    from qsource_amd import equilibrate, gensys
    t0 = time.time()
    s = gensys.evb_system(radius=40.0, n_solute=5000, seed=5000, n_q=24, n_states=4)
    log("c5 system: R=40 A, 5000 solute atoms, %d atoms, %d waters, built in %.1f s" %
        (s.natom, s.nwat, time.time() - t0))
    lam = gensys.mapping_lambdas(4, windows)
    rst = equilibrate.equilibrate(s, d, "c5", lambdas=list(lam[0]), seed=21, progress=log) if equil else None
    inputs = []
    for k in range(windows):
        kw = dict(steps=1000000, output=1000000, lambdas=list(lam[k]), initial_temperature=300.0,
                  random_seed=201 + k)
        if rst:
            kw["restart"] = os.path.basename(rst)
        inputs.append(gensys.write_system(s, d, "c5_w%02d" % k, **kw)["inp"])
    cfg = {"workload": "evb_mapping_windows", "systems": "R40/5000 solute atoms, EVB 4 states",
           "atoms_per_system": [s.natom], "waters": s.nwat, "windows_per_system": windows, "instances": windows,
           "q_atoms": 24, "states": 4, "offdiagonals": len(s.offdiag), "stepsize_fs": 1.0, "nbcycle": 10,
           "energy_interval": IENE, "far_field": "learned surrogate (BF16-MFMA MLP) in place of LRF"}
    return inputs, [list(range(windows))], [lam], cfg


def build_c4(d, equil=True, windows=11):
    """The learned far-field workload: one system (R = 30 A TIP3P sphere
    around a 2000-atom solute, 12 Q atoms, 2 FEP states, built like the
    batched-window systems) as `windows` lambda windows (lambda = 0, 0.1,
    ..., 1); the windows also as plain cut-off inputs (LRF off), and the
    system, its restart and the directory for the far-field comparison."""
    from qsource_amd import equilibrate, gensys
    t0 = time.time()
    s = gensys.solvated_system(radius=30.0, n_solute=2000, seed=4000, n_q=12, n_states=2)
    log("c4 system: R=30 A, 2000 solute atoms, %d atoms, %d waters, built in %.1f s" %
        (s.natom, s.nwat, time.time() - t0))
    lams = [round(k / (windows - 1), 4) for k in range(windows)]
    rst = equilibrate.equilibrate(s, d, "c4", lambdas=[0.5, 0.5], seed=31, progress=log) if equil else None
    inputs, cut = [], []
    for k, lam in enumerate(lams):
        kw = dict(steps=1000000, output=1000000, lambdas=[1.0 - lam, lam], initial_temperature=300.0,
                  random_seed=301 + k)
        if rst:
            kw["restart"] = os.path.basename(rst)
        inputs.append(gensys.write_system(s, d, "c4_w%02d" % k, **kw)["inp"])
        cut.append(gensys.write_system(s, d, "c4_cut_w%02d" % k, lrf=False, **kw)["inp"])
    cfg = {"workload": "learned_far_field_windows", "systems": "R30/2000 solute atoms",
           "atoms_per_system": [s.natom], "waters": s.nwat, "windows_per_system": windows, "instances": windows,
           "q_atoms": 12, "states": 2, "stepsize_fs": 1.0, "nbcycle": 10, "energy_interval": IENE,
           "far_field": "learned surrogate (BF16-MFMA MLP) in place of LRF"}
    lam = np.array([[1.0 - l, l] for l in lams])
    extra = {"system": s, "restart": rst, "dir": d, "cutoff_inputs": cut}
    return inputs, [list(range(windows))], [lam], cfg, extra


def c4_farfield(extra, model, inputs, groups, lams, ser_sur, steps, warmup):
    """Far-field treatments of the c4 system compared: gradient errors
    against exact all-pairs evaluation on held-out frames (farfield.
    force_errors), and dG over the windows with the surrogate, with LRF and
    with the plain cut-off (independent mixed-precision trajectories of the
    same length, block-bootstrap errors)."""
    from qsource_amd import engine as qe
    from qsource_amd import farfield
    t0 = time.time()
    tin = farfield.treatment_inputs(extra["system"], extra["dir"], "c4_ff", extra["restart"], [0.5, 0.5])
    frames = farfield.held_out_frames(tin["lrf"], nframes=10, stride=100, gap=5)
    fe = farfield.force_errors(tin, frames, model)
    log("far-field force errors vs exact (%.1f s): %s" % (time.time() - t0, json.dumps(fe)))
    res = {"force_error_vs_exact": fe}
    sers = {"surrogate": ser_sur}
    for name, inp in (("lrf", inputs), ("cutoff", extra["cutoff_inputs"])):
        e, wall, warm = run_batch(inp, qe.QGPU_MIXED, steps, warmup)
        sers[name] = series(e, groups, warm // IENE)
        e.close()
        log("c4 %s batch: %d steps" % (name, steps))
    dg = {}
    for name, ser in sers.items():
        z = dg_of(ser, lams)[0]
        b = bootstrap_dg(ser, lams)[0]
        dg[name] = {"dG_zwanzig": z[0], "dG_bar": z[1], "stderr": float(np.std(b)) if b is not None else None}
    for name in ("surrogate", "cutoff"):
        a, r = dg[name], dg["lrf"]
        dg[name]["ddG_vs_lrf"] = a["dG_zwanzig"] - r["dG_zwanzig"]
        dg[name]["stderr_ddG_vs_lrf"] = (float(np.hypot(a["stderr"], r["stderr"]))
                                         if a["stderr"] is not None and r["stderr"] is not None else None)
    res["dG_windows"] = dg
    res["frames_per_window"] = int(min(len(w) for w in ser_sur[0]))
    return res


def far_surrogate(inp, nframes=40, stride=25):
    """A far-field surrogate trained on the GPU on exact FP64 far-field
    targets of `nframes` frames, `stride` steps apart, of the FP64 run of
    window `inp`."""
    from qsource_amd import farfield
    t0 = time.time()
    Xs, Ys, _ = farfield.sample_frames(inp, nframes, stride)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    model, loss = farfield.train_surrogate(X, Y, epochs=30, batch=1024, lr=2e-3, seed=11)
    info = {"training_frames": nframes, "frame_stride_steps": stride, "samples": int(len(X)),
            "train_loss": loss, "seconds": time.time() - t0}
    log("surrogate: %s" % info)
    return model, info


# ---------------------------------------------------------------------------
# runs

def run_batch(inputs, precision, steps, warmup, far_model=None, chunk=0):
    """Engine over `inputs`: energy recording set up, launch shapes tuned and
    at least one whole list cycle run before the timed region.  Returns
    (engine, wall seconds of the timed steps, warm-up steps run).  With
    chunk > 0 the timed steps run in calls of `chunk` steps, each logged."""
    from qsource_amd import engine as qe
    e = qe.Engine(inputs, precision)
    nb = 10
    warm = max(nb, -(-warmup // nb) * nb)
    e.set_energy_interval(IENE, (warm + steps) // IENE + 2)
    if far_model is not None:
        qe.set_far_model(e, far_model)
    e.tuned_shapes = e.autotune(reps=3)
    e.step(warm)
    t0 = time.perf_counter()
    if chunk > 0:
        done = 0
        while done < steps:
            m = min(chunk, steps - done)
            e.step(m)
            done += m
            log("  %d / %d steps, %.1f s" % (done, steps, time.perf_counter() - t0))
    else:
        e.step(steps)
    wall = time.perf_counter() - t0
    return e, wall, warm


def series(e, groups, skip, fp64=False):
    """Per system, per window: (nframes, nstates) EQ totals of the frames
    recorded after the first `skip` frames."""
    return [[e.energy_series(k, fp64=fp64)[skip:, :, 0] for k in g] for g in groups]


def dg_of(ser, lams, temperature=300.0):
    """(Zwanzig, BAR) dG per system."""
    from qsource_amd import fep
    out = []
    for v, lam in zip(ser, lams):
        z = fep.zwanzig(lam, v, temperature)["total"]
        try:
            b = fep.bar(lam, v, temperature)["total"]
        except ValueError:
            b = float("nan")
        out.append((z, b))
    return out


def bootstrap_dg(ser, lams, temperature=300.0, nblocks=20, nboot=200, seed=7):
    """Bootstrap replicas of the Zwanzig dG per system: every window's frames
    are cut into `nblocks` consecutive blocks (frames are correlated) and the
    blocks are resampled.  None for a system whose windows hold fewer than
    2 * nblocks frames."""
    from qsource_amd import fep
    rng = np.random.default_rng(seed)
    out = []
    for v, lam in zip(ser, lams):
        n = min(len(w) for w in v)
        if n < 2 * nblocks:
            out.append(None)
            continue
        bl = n // nblocks
        reps = []
        for _ in range(nboot):
            w = []
            for e in v:
                pick = rng.integers(0, nblocks, nblocks)
                w.append(np.concatenate([e[b * bl:(b + 1) * bl] for b in pick]))
            reps.append(fep.zwanzig(lam, w, temperature)["total"])
        out.append(np.array(reps))
    return out


def trajectory_ddg(ser_m, ser_f, lams):
    """ddG of independent mixed and FP64 trajectories per system, its
    bootstrap standard error, and whether the gate is established."""
    dg_m, dg_f = dg_of(ser_m, lams), dg_of(ser_f, lams)
    bm, bf = bootstrap_dg(ser_m, lams), bootstrap_dg(ser_f, lams)
    ddg = [a[0] - b[0] for a, b in zip(dg_m, dg_f)]
    se = [float(np.hypot(np.std(a), np.std(b))) if a is not None and b is not None else None
          for a, b in zip(bm, bf)]
    established = all(s is not None and s < SE_GATE for s in se)
    return {"dG_mixed_zwanzig": [a[0] for a in dg_m], "dG_mixed_bar": [a[1] for a in dg_m],
            "dG_fp64_zwanzig": [b[0] for b in dg_f], "dG_fp64_bar": [b[1] for b in dg_f],
            "ddG_zwanzig": ddg, "ddG_bar": [a[1] - b[1] for a, b in zip(dg_m, dg_f)],
            "stderr_ddG_zwanzig": se, "frames_per_window": int(min(len(w) for v in ser_m for w in v)),
            "gate_established": established,
            "gate_pass": bool(established and all(abs(x) < DDG_GATE for x in ddg))}


def force_error(inputs, states, terms=("ww", "pw", "pp", "qx", "lrf")):
    """Mixed vs FP64 gradients on identical coordinates (one window per
    system): for the whole gradient and for each term group alone, the RMS
    of the difference (kcal/mol/A) and that RMS relative to the RMS of the
    FP64 term gradient; and per-term energy differences."""
    from qsource_amd import engine as qe
    engines = [qe.Engine(inputs, p) for p in (qe.QGPU_FP64, qe.QGPU_MIXED)]
    for e in engines:
        for k, (x, v) in enumerate(states):
            e.set_state(k, x, v)
        e.make_pair_lists()
    res = [{"per_term": {}} for _ in inputs]
    for term in (None,) + tuple(terms):
        for e in engines:
            e.pot_energy(None if term is None else [term])
        for k in range(len(inputs)):
            d64, d32 = engines[0].gradient(k), engines[1].gradient(k)
            rms_ref = float(np.sqrt(np.mean(d64 ** 2)))
            rms = float(np.sqrt(np.mean((d32 - d64) ** 2)))
            rec = {"rms_kcal_mol_A": rms, "rms_rel": rms / rms_ref if rms_ref > 0 else 0.0}
            if term is None:
                res[k]["force_rms_rel"] = rec["rms_rel"]
                e64, e32 = engines[0].energies(k), engines[1].energies(k)
                res[k]["energy_diff"] = {t: float(e32[t] - e64[t]) for t in (
                    "ww_el", "ww_vdw", "pw_el", "pw_vdw", "pp_el", "pp_vdw", "qx_el", "qx_vdw", "lrf",
                    "potential")}
            else:
                res[k]["per_term"][term] = rec
    for e in engines:
        e.close()
    return res


# ---------------------------------------------------------------------------
# roofline accounting

def kernel_rooflines(e, prof, nbcycle=10, lrf=True):
    """Roofline entries of the four large kernels, from the reference's
    list sizes (counted by the engine) and the per-unit FLOP counts:
    water forces: ww water pairs x 226, solute-water atom pairs x 34 (the
    water side), Q-water atom pairs x 24 and, with LRF, the Taylor term of
    each water atom (90); solute forces: solute-solute and solute-water atom
    pairs x 34 each and the Taylor term of each solute atom; water and solute
    lists: lrf_update atom-to-centre evaluations x 60."""
    acc = {"water_forces": 0.0, "solute_forces": 0.0, "water_lists": 0.0, "solute_lists": 0.0}
    units = {"water_pairs": 0, "pw_atom_pairs": 0, "pp_atom_pairs": 0, "qw_atom_pairs": 0,
             "lrf_evals_water_side": 0.0, "lrf_evals_solute_side": 0.0}
    for k in range(e.n):
        lc = e.list_counts(k)
        ww = lc["ww"] // 9
        nw = (e.natom[k] - e.nat_solute(k)) // 3
        units["water_pairs"] += ww
        units["pw_atom_pairs"] += lc["pw"]
        units["pp_atom_pairs"] += lc["pp"]
        units["qw_atom_pairs"] += lc["qw"]
        acc["water_forces"] += ww * (8 * FLOP_COUL + FLOP_LJ) + lc["pw"] * FLOP_LJ + lc["qw"] * FLOP_COUL + \
            (3 * nw * FLOP_LRF_ATOM if lrf else 0)
        acc["solute_forces"] += (lc["pp"] + lc["pw"]) * FLOP_LJ + (e.nat_solute(k) * FLOP_LRF_ATOM if lrf else 0)
        if lrf:
            wc = e.work_counts(k)
            units["lrf_evals_water_side"] += wc["lrf_evals_water_side"]
            units["lrf_evals_solute_side"] += wc["lrf_evals_solute_side"]
    acc["water_lists"] = units["lrf_evals_water_side"] * FLOP_LRF_UPDATE
    acc["solute_lists"] = units["lrf_evals_solute_side"] * FLOP_LRF_UPDATE
    out = {}
    for name, flop in acc.items():
        ms = prof.get(name, 0.0)
        ach = flop / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        per_step = ms / nbcycle if name.endswith("lists") else ms
        out[name] = {"bound": "valu", "kernel": name, "achieved": ach, "peak": PEAK_TFLOPS["mixed"],
                     "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["mixed"], "flop_per_launch": flop,
                     "kernel_ms": ms, "ms_per_step": per_step}
    out["units_per_launch"] = units
    tw = profile_value("k_water_forces_c3_63win_mixed_pmc_traffic.md", "traffic_bytes_per_launch")
    tl = profile_value("k_water_lists_c3_63win_mixed_pmc_traffic.md", "traffic_bytes_per_launch")
    out["water_forces"]["traffic"] = int(tw) if tw else None
    out["water_lists"]["traffic"] = int(tl) if tl else None
    return out


def per_step_ms(prof, nbcycle=10):
    """Kernel time per step: list kernels run once per list cycle, the
    energy-frame kernel once per energy interval."""
    once = ("cgp_centers", "water_lists", "solute_lists")
    return {k: (v / nbcycle if k in once else v / IENE if k == "energy_record" else v)
            for k, v in prof.items()}


# ---------------------------------------------------------------------------
# CPU baseline

_CPU_CHILD = r"""
import sys, time
sys.path.insert(0, sys.argv[1])
from oracle import Oracle
o = Oracle(sys.argv[2])
n = int(sys.argv[3])
t0 = time.perf_counter()
o.run(n)
print("SECONDS %.6f" % (time.perf_counter() - t0), flush=True)
"""


def cpu_baseline(inputs, seconds=15.0, procs=8):
    """The FP64 CPU oracle (a port of serial qdyn's step) on this host: one
    window per process, `procs` processes at once, each running about
    `seconds` of steps; aggregate ns/day at 1 fs."""
    from oracle import LIB_PATH
    if not os.path.exists(LIB_PATH) or not inputs:
        return None
    procs = max(1, min(procs, len(inputs), os.cpu_count() or 1, 16))

    def run(jobs):
        ps = [subprocess.Popen([sys.executable, "-c", _CPU_CHILD, ROOT, inp, str(n)],
                               stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for inp, n in jobs]
        t0 = time.perf_counter()
        outs = [p.communicate()[0] for p in ps]
        wall = time.perf_counter() - t0
        secs = []
        for o, p in zip(outs, ps):
            if p.returncode != 0 or "SECONDS" not in o:
                log("oracle CPU run failed:\n" + o[-2000:])
                return None, None
            secs.append(float(o.split("SECONDS")[1].split()[0]))
        return wall, secs
    _, cal = run([(inputs[0], 20)])
    if cal is None:
        return None
    n = max(20, int(round(seconds / (cal[0] / 20) / 10.0)) * 10)
    wall, secs = run([(inputs[k % len(inputs)], n) for k in range(procs)])
    if wall is None:
        return None
    return {"value": procs * n * 1e-6 * 86400.0 / max(secs), "unit": "ns/day", "cores": procs, "kind": "port",
            "sample": "FP64 CPU oracle (order-faithful port of serial qdyn), %d windows of the R=22 A / 300-atom "
                      "system at once, one per core, %d steps each at 1 fs / 300 K from the equilibrated restart "
                      "(longest %.1f s)" % (procs, n, max(secs))}


def trace_summary(path):
    """Per-step timeline of a rocprofv3 kernel trace (CSV) of a bench run:
    the steady-state steps are the run of mixed-precision water-force
    launches of the chosen variant; prints the wall time per step, the time
    per step at least one kernel runs, the idle fraction, and each kernel's
    busy time per step."""
    import collections
    import csv
    ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                 r["Kernel_Name"].split("(")[0].replace("void ", "").replace("qgpu::", ""))
                for r in csv.DictReader(open(path)))
    cnt = collections.Counter(k[2] for k in ks if k[2].startswith("k_water_forces<float"))
    marker = cnt.most_common(1)[0][0]
    wf = [k for k in ks if k[2] == marker]
    runs, cur = [], [wf[0]]
    for a, b in zip(wf, wf[1:]):
        if b[0] - a[0] < 4000000:
            cur.append(b)
        else:
            runs.append(cur)
            cur = [b]
    runs.append(cur)
    run = max(runs, key=len)
    t0, t1 = run[1][0], run[-1][0]
    nsteps = len(run) - 2
    inwin = [k for k in ks if k[1] > t0 and k[0] < t1]
    busy, cs, ce = 0, None, None
    for s, e in sorted((max(s, t0), min(e, t1)) for s, e, _ in inwin):
        if cs is None:
            cs, ce = s, e
        elif s <= ce:
            ce = max(ce, e)
        else:
            busy += ce - cs
            cs, ce = s, e
    if cs is not None:
        busy += ce - cs
    per = collections.defaultdict(lambda: [0, 0])
    for s, e, n in inwin:
        per[n][0] += min(e, t1) - max(s, t0)
        per[n][1] += 1
    out = {"steps": nsteps, "wall_us_per_step": (t1 - t0) / nsteps / 1e3,
           "busy_us_per_step": busy / nsteps / 1e3, "idle_fraction": 1.0 - busy / (t1 - t0),
           "kernel_us_per_step": {n: round(v[0] / nsteps / 1e3, 2) for n, v in
                                  sorted(per.items(), key=lambda x: -x[1][0])}}
    print(json.dumps(out))


# ---------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", default="c3", choices=["c3", "c4", "c5"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=8)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=0,
                    help="long accuracy run: mixed and FP64 batches of this many steps (no timing report)")
    ap.add_argument("--accuracy-control", action="store_true",
                    help="with --accuracy-steps: also a mixed and an FP64 batch from other initial velocities")
    ap.add_argument("--trace-summary", default=None,
                    help="summarise a rocprofv3 kernel-trace CSV of a bench run (no GPU work)")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
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
        c4x = None
        if args.config == "c3":
            inputs, groups, lams, cfg = build_c3(d)
        elif args.config == "c4":
            inputs, groups, lams, cfg, c4x = build_c4(d)
        else:
            inputs, groups, lams, cfg = build_c5(d)
        n = len(inputs)
        if args.accuracy_steps > 0:
            accuracy_run(inputs, groups, lams, cfg, args.accuracy_steps, control=args.accuracy_control)
            return
        cpu = None
        if args.cpu_seconds > 0 and args.config == "c3":
            log("CPU baseline: oracle, %d windows" % args.cpu_procs)
            cpu = cpu_baseline([inputs[k] for k in groups[1]], args.cpu_seconds, args.cpu_procs)
            log("CPU baseline: %s" % cpu)
        model = None
        if args.config == "c5":
            model, cfg["surrogate"] = far_surrogate(inputs[0])
        elif args.config == "c4":
            # 200 frames of the middle window's FP64 run
            model, cfg["surrogate"] = far_surrogate(inputs[len(inputs) // 2], nframes=200, stride=25)
        log("mixed batch: %d instances" % n)
        em, wall, warm = run_batch(inputs, qe.QGPU_MIXED, args.steps, args.warmup, far_model=model)
        skip = warm // IENE
        tuned = {k: {"mapping": ["grid2d", "instance_map", "xcd_map"][m], "block": b, "prefetch": pf,
                     "waves_per_simd": (4 if oc else 6) if k == "water_forces" else None,
                     "ms_chosen": best, "ms_default": dflt}
                 for k, (m, b, pf, oc, best, dflt) in em.tuned_shapes.items()}
        log("tuned launch shapes: %s" % tuned)
        value = n * args.steps * cfg["stepsize_fs"] * 1e-6 * 86400.0 / wall
        log("mixed: %.3f ms/step, %.1f ns/day aggregate (warm-up %d steps)" % (wall / args.steps * 1e3, value, warm))
        ser_m = series(em, groups, skip)
        ser_m64 = series(em, groups, skip, fp64=True)
        dg_m, dg_m64 = dg_of(ser_m, lams), dg_of(ser_m64, lams)
        du_err = max(float(np.abs(a - b).max()) for sm, sf in zip(ser_m, ser_m64) for a, b in zip(sm, sf))
        mid = [g[len(g) // 2] for g in groups]
        states = [em.state(k) for k in mid]
        prof = em.profile_kernels(args.profile_steps)
        step_ms = per_step_ms(prof)
        dom = max(step_ms, key=lambda k: step_ms[k])
        roofs = kernel_rooflines(em, prof, lrf=model is None)
        roof = roofs.get(dom, roofs["water_forces"])
        roof = dict(roof, dominant_kernel=dom)
        em.close()
        farfield_res = None
        if c4x is not None:
            farfield_res = c4_farfield(c4x, model, inputs, groups, lams, ser_m, args.steps, args.warmup)
        ferr = force_error([inputs[k] for k in mid], states)
        log("force RMS error per term (mixed vs FP64, same coordinates): %s" %
            [{t: round(v["rms_rel"], 9) for t, v in f["per_term"].items()} for f in ferr])
        acc = {"precision_only": {
            "what": "dG from the mixed run's dU vs dG from FP64 re-evaluation of the same frames",
            "abs_ddG_zwanzig": [abs(a[0] - b[0]) for a, b in zip(dg_m, dg_m64)],
            "abs_ddG_bar": [abs(a[1] - b[1]) for a, b in zip(dg_m, dg_m64)],
            "max_abs_dU_error_kcal": du_err},
            "force_error_one_window_per_system": ferr}
        if not args.skip_fp64:
            log("fp64 batch (independent trajectories)")
            ef, wall64, _ = run_batch(inputs, qe.QGPU_FP64, args.steps, args.warmup, far_model=model)
            ser_f = series(ef, groups, skip)
            ef.close()
            tr = trajectory_ddg(ser_m, ser_f, lams)
            tr["fp64_batched_ns_per_day"] = n * args.steps * 1e-6 * 86400.0 / wall64
            acc["trajectory_this_run"] = tr
            log("ddG of independent trajectories this run: %s +- %s (established: %s)" %
                (tr["ddG_zwanzig"], tr["stderr_ddG_zwanzig"], tr["gate_established"]))
        long_name = "ddG_%s_long_run.md" % args.config
        long_run = profile_json(long_name)
        this = acc.get("trajectory_this_run")
        # the gate on identical frames: the error the mixed arithmetic adds
        acc["gate_ddG_lt_0.05_precision_only"] = bool(
            max(acc["precision_only"]["abs_ddG_zwanzig"] + acc["precision_only"]["abs_ddG_bar"]) < DDG_GATE)
        # the gate on independent trajectories, where sampling noise enters
        if this and this["gate_established"]:
            gate, source = this["gate_pass"], "this run"
        elif long_run:
            gate, source = long_run.get("gate_pass"), "profiles/" + long_name
            acc["trajectory_long_run"] = {k: long_run.get(k) for k in (
                "steps_per_window", "ddG_zwanzig", "stderr_ddG_zwanzig", "rms_ddG_cross_precision",
                "rms_ddG_same_precision", "gate_established", "gate_note")}
        else:
            gate, source = None, "unestablished (no run long enough)"
        acc["gate_ddG_lt_0.05_trajectories"] = gate
        acc["gate_trajectories_source"] = source
        # plain FP64 engine path: the windows run one at a time, so the
        # job's aggregate rate is the simulated time over the summed wall
        # time of its windows; one window of every system is timed
        psteps = max(500, args.steps // 2)
        plain_walls = []
        for g in groups:
            ep, wallp, _ = run_batch([inputs[g[len(g) // 2]]], qe.QGPU_FP64, psteps, 50)
            plain_walls.append(wallp)
            ep.close()
        plain = len(groups) * psteps * 1e-6 * 86400.0 / sum(plain_walls)
        plain_per_system = [psteps * 1e-6 * 86400.0 / w for w in plain_walls]
        if model is not None:
            model.close()
    cfg["precision"] = "mixed"
    ref_cpu = profile_json("cpu_reference_qdyn5_%s.md" % args.config)
    out.update({"value": value, "ms_per_step": wall / args.steps * 1e3, "config": cfg,
                "roofline": roof, "roofline_kernels": roofs, "cpu_baseline": cpu,
                "cpu_reference_build_host": ref_cpu, "kernel_ms": prof, "kernel_ms_per_step": step_ms,
                "autotune": tuned, "plain_fp64_one_window_at_a_time_ns_per_day": plain,
                "plain_fp64_ns_per_day_per_system": plain_per_system,
                "speedup_vs_plain_fp64": value / plain, "accuracy": acc})
    if farfield_res is not None:
        out["farfield"] = farfield_res
    print(json.dumps(out))


def reseeded(inputs, offset):
    """Copies of the window inputs (same restarts) whose initial velocities
    come from other seeds: random_seed + offset."""
    out = []
    for p in inputs:
        lines = open(p).read().splitlines()
        for i, line in enumerate(lines):
            w = line.split()
            if len(w) == 2 and w[0] == "random_seed":
                lines[i] = "%-24s %d" % ("random_seed", int(w[1]) + offset)
        q = p[:-len(".inp")] + "_b.inp"
        with open(q, "w") as f:
            f.write("\n".join(lines) + "\n")
        out.append(q)
    return out


def accuracy_run(inputs, groups, lams, cfg, steps, control=False):
    """Long mixed and FP64 batches of the workload: the trajectory ddG with
    its bootstrap standard error, printed as one JSON line (committed as the
    long-run summary that the short benchmark cites).  With `control`, a
    second mixed and a second FP64 batch start from other initial
    velocities (B): the ddG between runs of the same precision (FP64 B - FP64
    A, mixed B - mixed A) measures the sampling noise that the cross-precision
    ddG (mixed - FP64, A and B) is to be compared with."""
    from qsource_amd import engine as qe
    res = {"workload": cfg["workload"], "steps_per_window": steps, "energy_interval": IENE}
    sers = {}
    runs = [("mixed", qe.QGPU_MIXED, inputs), ("fp64", qe.QGPU_FP64, inputs)]
    if control:
        other = reseeded(inputs, 500)
        runs += [("mixed_b", qe.QGPU_MIXED, other), ("fp64_b", qe.QGPU_FP64, other)]
    for name, prec, inp in runs:
        log("accuracy run: %s batch, %d steps" % (name, steps))
        e, wall, warm = run_batch(inp, prec, steps, 200, chunk=20000)
        res["%s_ns_per_day" % name] = len(inp) * steps * 1e-6 * 86400.0 / wall
        sers[name] = series(e, groups, warm // IENE)
        e.close()
        log("  %.1f s" % wall)
    tr = trajectory_ddg(sers["mixed"], sers["fp64"], lams)
    res.update(tr)
    if control:
        pairs = {"mixed_b-fp64_b": ("mixed_b", "fp64_b"), "fp64_b-fp64": ("fp64_b", "fp64"),
                 "mixed_b-mixed": ("mixed_b", "mixed")}
        res["control"] = {}
        for key, (a, b) in pairs.items():
            c = trajectory_ddg(sers[a], sers[b], lams)
            res["control"][key] = {"ddG_zwanzig": c["ddG_zwanzig"], "stderr_ddG_zwanzig": c["stderr_ddG_zwanzig"]}
        same = [x for key in ("fp64_b-fp64", "mixed_b-mixed") for x in res["control"][key]["ddG_zwanzig"]]
        cross = tr["ddG_zwanzig"] + res["control"]["mixed_b-fp64_b"]["ddG_zwanzig"]
        res["rms_ddG_same_precision"] = float(np.sqrt(np.mean(np.square(same))))
        res["rms_ddG_cross_precision"] = float(np.sqrt(np.mean(np.square(cross))))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
