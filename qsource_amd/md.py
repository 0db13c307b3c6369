"""Runs qdyn inputs on the GPU engine as md_run does (src/md.f90:6471-6704)
and writes qdyn's own output, so runs chain like qdyn's:

* the printed energy summary (write_out, src/md.f90:18545-18691) of every
  `output` step of [intervals] and the final one, in qdyn's layout;
* a backup restart after every step whose number is a multiple of 1000
  (src/md.f90:6639-6641), and a restart at a failure (die, :554-613);
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
import sys
from typing import Dict, List, Optional, Sequence, TextIO

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
    # without [lambdas] a one-state FEP file runs at lambda 1
    lam = [float(v) for l in sec.get("lambdas", []) for v in l.split()] or [1.0]
    onoff = lambda v: v is not None and v.lower() in ("on", "1", "true", ".true.")
    return dict(steps=int(get("md", "steps", 0)), itrj=int(get("intervals", "trajectory", 0)),
                iene=int(get("intervals", "energy", 0)),
                iout=int(get("intervals", "output", IOUT_DEFAULT)),
                use_lrf=get("md", "lrf") is None or onoff(get("md", "lrf")),
                npt=onoff(get("pbc", "constant_pressure")),
                force_rms=onoff(get("intervals", "force_rms")) or onoff(get("md", "force_rms")),
                final=res(get("files", "final")),
                trajectory=res(get("files", "trajectory")), energy=res(get("files", "energy")),
                lambdas=lam)


IOUT_DEFAULT = 10    # iout_cycle_default (src/md.f90:225)


# ---------------------------------------------------------------------------
# write_out's printed energy summary (src/md.f90:18545-18691), field by field
# as its Fortran formats lay it out.

def _f(v: float, w: int, d: int) -> str:
    """Fw.d (a value too wide for the field prints as w asterisks)."""
    t = "%*.*f" % (w, d, v)
    return t if len(t) <= w else "*" * w


def _i(v: int, w: int) -> str:
    t = "%*d" % (w, v)
    return t if len(t) <= w else "*" * w


def _a(t: str, w: int) -> str:
    """Aw: right-justified in w columns (truncated to its leftmost w)."""
    return t.rjust(w)[:w] if len(t) <= w else t[:w]


def _t17(name: str) -> str:
    """A then T17: the name, then the record continues at column 17."""
    return name.ljust(16) if len(name) < 16 else name[:16]


def _heading(title: str, final: bool, step: int) -> str:
    if final:
        return "=========================== FINAL " + _a(title, 15) + " ============================="
    return "======================= " + _a(title, 15) + " at step " + _i(step, 6) + " ========================"


def _centered(msg: str, fill: str = "=") -> str:
    """centered_heading (src/misc.f90:46-54)."""
    n = (78 - len(msg)) // 2
    return fill * n + " " + msg + " " + fill * max(0, n - 1)


def volume_summary(volume: float, tries: int, accepted: int) -> str:
    """The volume-change summary write_out prints after the final energies
    of a constant-pressure run (src/md.f90:18678-18690): the final volume,
    then the moves tried and accepted and their ratio."""
    ratio = accepted / tries if tries else float("nan")
    r = "       NaN" if tries == 0 else _f(ratio, 10, 3)
    L = ["=========================== VOLUME CHANGE SUMMARY ===========================",
         "Final volume: " + _f(volume, 10, 3), "",
         " " * 16 + _a("total", 10) + _a("accepted", 10) + _a("ratio", 10),
         _t17("Attempts") + _i(tries, 10) + _i(accepted, 10) + r, "=" * 79]
    return "\n".join(L)


def _reverting(items: List[str], per_record: int) -> List[str]:
    """A format with a repeated group of per_record items: further items
    continue on new records (format reversion)."""
    return ["".join(items[k:k + per_record]) for k in range(0, len(items), per_record)] or [""]


def write_out(E: Dict[str, float], EQ: Optional[np.ndarray], lambdas: Sequence[float], *, step: int,
              final: bool, nwat: int, nqat: int, nat_solute: int, use_lrf: bool = True,
              offd: Optional[np.ndarray] = None, monitor: Optional[np.ndarray] = None,
              grms: Optional[float] = None) -> str:
    """The energy summary qdyn prints (write_out): E holds the energy terms
    (qgpu_energies names); EQ (nstates, 12) the per-state Q energies
    (qgpu_qenergies rows); offd (noffd, 6) the off-diagonal functions (state
    i, j, Q atoms k, l, Hij, r_kl) of the last offdiag call; monitor
    (npairs, 3 + 2 nstates) the monitored group pairs of the last
    nonbond_monitor call; grms the RMS gradient when force_rms is on."""
    L: List[str] = []
    row = lambda name, vals: L.append(_t17(name) + "".join(_f(v, 12, 2) for v in vals))
    legend = lambda names: L.append(" " * 16 + "".join(_a(n, 10) for n in names))
    L.append(_heading("Energy summary", final, step))
    legend(["el", "vdW", "bond", "angle", "torsion", "improper"])
    row("solute", [E["pp_el"], E["pp_vdw"], E["p_bond"], E["p_angle"], E["p_torsion"], E["p_improper"]])
    if nwat > 0:
        row("solvent", [E["ww_el"], E["ww_vdw"], E["w_bond"], E["w_angle"], E["w_torsion"], E["w_improper"]])
    row("solute-solvent", [E["pw_el"], E["pw_vdw"]])
    if use_lrf:
        row("LRF", [E["lrf"]])
    if nqat > 0:
        row("Q-atom", [E["qx_el"], E["qx_vdw"], E["q_bond"], E["q_angle"], E["q_torsion"], E["q_improper"]])
    L.append("")
    legend(["total", "fix", "slvnt_rad", "slvnt_pol", "shell", "solute"])
    row("restraints", [E["rst_total"], E["rst_fix"], E["rst_radial"], E["rst_wpol"], E["rst_shell"],
                       E["rst_protein"]])
    L.append("")
    if grms is not None:
        legend(["total", "potential", "kinetic", "", "RMS force"])
        L.append(_t17("SUM") + _f(E["potential"] + E["kinetic"], 10, 2) + _f(E["potential"], 10, 2) +
                 _f(E["kinetic"], 10, 2) + " " * 10 + _f(grms, 10, 2))
    else:
        legend(["total", "potential", "kinetic"])
        row("SUM", [E["potential"] + E["kinetic"], E["potential"], E["kinetic"]])
    ns = 0 if EQ is None else len(EQ)
    if ns > 0:
        L.append(_heading("Q-atom energies", final, step))
        L.append("type   st lambda" + "".join(_a(n, 10) for n in ["el", "vdW", "bond", "angle", "torsion",
                                                                   "improper"]))
        q32 = lambda name, s, a, b: L.append(name.ljust(7)[:7] + _i(s + 1, 2) + _f(lambdas[s], 7, 4) +
                                             _f(a, 10, 2) + _f(b, 10, 2))
        for s in range(ns):
            q32("Q-Q", s, EQ[s, 0], EQ[s, 1])
        L.append("")
        if nat_solute > nqat:
            for s in range(ns):
                q32("Q-prot", s, EQ[s, 2], EQ[s, 3])
            L.append("")
        if nwat > 0:
            for s in range(ns):
                q32("Q-wat", s, EQ[s, 4], EQ[s, 5])
            L.append("")
        for s in range(ns):
            q32("Q-surr.", s, EQ[s, 2] + EQ[s, 4], EQ[s, 3] + EQ[s, 5])
        L.append("")
        for s in range(ns):
            qx_el, qx_vdw = EQ[s, 0] + EQ[s, 2] + EQ[s, 4], EQ[s, 1] + EQ[s, 3] + EQ[s, 5]
            L.append("Q-any".ljust(7) + _i(s + 1, 2) + _f(lambdas[s], 7, 4) +
                     "".join(_f(v, 10, 2) for v in (qx_el, qx_vdw, EQ[s, 7], EQ[s, 8], EQ[s, 9], EQ[s, 10])))
        L.append("")
        L.append("type   st lambda" + _a("total", 10) + _a("restraint", 10))
        for s in range(ns):
            q32("Q-SUM", s, EQ[s, 6], EQ[s, 11])
        if offd is not None and len(offd):
            # Hij(i,j) holds the value of the last function of that state pair
            hij = {}
            for o in offd:
                hij[(int(o[0]), int(o[1]))] = o[4]
            for o in offd:
                L.append("H(" + _i(int(o[0]), 2) + "," + _i(int(o[1]), 2) + ") =" +
                         _f(hij[(int(o[0]), int(o[1]))], 8, 2) + " dist. between Q-atoms" +
                         _i(int(o[2]), 4) + _i(int(o[3]), 4) + " =" + _f(o[5], 8, 2))
    if monitor is not None and len(monitor):
        L.append(_centered("Monitoring selected groups of nonbonded interactions"))
        nsm = (monitor.shape[1] - 3) // 2
        head = _reverting([_i(s + 1, 4) + ":Vel" + _i(s + 1, 3) + ":Vvdw" for s in range(nsm)], 3)
        L.append("pair   Vwsum    Vwel    Vwvdw" + head[0])
        L.extend(head[1:])
        for p, m in enumerate(monitor):
            vals = _reverting([_f(m[3 + 2 * s], 8, 2) + _f(m[4 + 2 * s], 8, 2) for s in range(nsm)], 3)
            L.append(_i(p + 1, 2) + _f(m[0], 10, 2) + _f(m[1], 8, 2) + _f(m[2], 9, 2) + vals[0])
            L.extend(vals[1:])
    L.append("=" * 79)
    return "\n".join(L)


def trajectory_steps(nsteps: int, itrj: int) -> List[int]:
    """Numbers of completed steps at which md_run writes a trajectory frame."""
    if itrj <= 0:
        return []
    out = [m * itrj + 1 for m in range(1, nsteps) if m * itrj <= nsteps - 1]
    if nsteps % itrj == 0:
        out.append(nsteps)      # with itrj = 1 the last state is written twice, as qdyn does
    return out


def _abnormal(log: TextIO, msg: str, at_step: int, restart: Optional[str]) -> None:
    """die's closing lines (src/md.f90:554-613)."""
    log.write("\n" + _centered("ABNORMAL TERMINATION", "!") + "\n")
    if restart:
        log.write(" restart file written at step %d\n" % at_step)
    log.write("!" * 79 + "\nABNORMAL TERMINATION of Qdyn5\nTerminating due to " + msg + "\n" + "!" * 79 + "\n")


def _box(e, k):
    """The restart's box record of a periodic-box instance (lengths, centre)."""
    b = e.box(k)
    return None if b is None else np.concatenate(b)


def run(inputs: Sequence[str], precision: int = qe.QGPU_MIXED,
        far_model: Optional["qe.FarFieldMLP"] = None, version: str = "qgpu",
        logs: Optional[Sequence[TextIO]] = None) -> List[Dict[str, float]]:
    """Runs every input for its [MD] steps on one GPU and writes its files;
    returns the final energies of each instance (the closing list update and
    pot_energy of md_run).  With `logs` (one text stream per input), each
    instance's energy summaries are written there as md_run prints them: at
    every loop step istep with istep % output == 0 (the energies of that
    step's pot_energy, the kinetic energy after its update, the
    off-diagonals of its offdiag call and the monitored pairs of its
    nonbond_monitor call), and the final summary after the closing list
    update and pot_energy (src/md.f90:6560-6568, 6634-6636, 6686-6692).
    A failure (die) writes each instance's restart at the failing step and
    raises."""
    st = [read_settings(p) for p in inputs]
    for key in ("steps", "itrj", "iene", "iout"):
        if len({s[key] for s in st}) != 1:
            raise ValueError("batched inputs must share [MD] steps and output intervals (%s)" % key)
    nsteps, itrj, iene, iout = st[0]["steps"], st[0]["itrj"], st[0]["iene"], st[0]["iout"]
    if logs is not None and len(logs) != len(inputs):
        raise ValueError("one log stream per input")
    e = qe.Engine(list(inputs), precision)
    try:
        if far_model is not None:
            qe.set_far_model(e, far_model)
        if iene > 0:
            e.set_energy_interval(iene, max(1, nsteps // iene))
        n = len(inputs)
        sizes = [e.sizes(k) for k in range(n)]
        noffd = [len(e.offdiag_states(k)) for k in range(n)]
        nmon = [len(e.monitor(k)) for k in range(n)]
        trj = [qe.Trajectory(p) if itrj > 0 else None for p in inputs]
        # the loop steps at which the host acts, as numbers of completed steps
        out_steps = [i for i in range(0, nsteps, iout)] if (logs is not None and iout > 0) else []
        pre = set(out_steps)
        if logs is not None and any(noffd) and nsteps > 0:
            pre.add(nsteps - 1)            # the final summary shows the last offdiag call
        post = {i + 1 for i in out_steps}
        frames = set(trajectory_steps(nsteps, itrj))
        backup = {i + 1 for i in range(0, nsteps, 1000)} if any(s["final"] for s in st) else set()
        stops = sorted(pre | post | frames | backup | {nsteps})
        last_offd: List[Optional[np.ndarray]] = [None] * n
        last_mon: List[Optional[np.ndarray]] = [None] * n
        done = 0

        def summary(k, step, final):
            grms = None
            if st[k]["force_rms"]:
                d = e.gradient(k)
                grms = float(np.sqrt(np.dot(d, d) / len(d)))
            z = sizes[k]
            return write_out(e.energies(k), e.qenergies(k) if z["nstates"] > 0 else None, st[k]["lambdas"],
                             step=step, final=final, nwat=z["nwat"], nqat=z["nqat"], nat_solute=z["nat_solute"],
                             use_lrf=st[k]["use_lrf"] and far_model is None, offd=last_offd[k],
                             monitor=last_mon[k], grms=grms) + "\n"

        for stop in stops:
            if stop > done:
                try:
                    e.step(stop - done)
                except qe.QgpuError as err:
                    for k, s in enumerate(st):
                        at = e.status(k)[1]
                        if s["final"]:
                            x, v = e.state(k)
                            qe.write_restart(s["final"], x, v, e.theta_corr(k), box=_box(e, k))
                        if logs is not None:
                            _abnormal(logs[k], str(err), at, s["final"])
                    raise
            done = stop
            if done in post and logs is not None:
                for k in range(n):
                    logs[k].write(summary(k, done - 1, False))
            if done in pre and done < nsteps:
                for k in range(n):
                    if noffd[k]:
                        last_offd[k] = e.offdiag(k)
                    if nmon[k] and done % max(iout, 1) == 0:
                        last_mon[k] = e.monitor(k)
            if done in frames:
                for k, t in enumerate(trj):
                    t.write(e.state(k)[0])
            if done in backup:
                for k, s in enumerate(st):
                    if s["final"]:
                        x, v = e.state(k)
                        qe.write_restart(s["final"], x, v, e.theta_corr(k), box=_box(e, k))
        for t in trj:
            if t is not None:
                t.close()
        for k, s in enumerate(st):
            if iene > 0 and s["energy"]:
                fr = e.energy_series(k)
                od = e.offdiag_series(k)
                ij = e.offdiag_states(k)
                offd = None
                if od.shape[1] > 0:
                    offd = np.concatenate([np.broadcast_to(ij, od.shape[:2] + (2,)), od], axis=2)
                fep.write_energy_file(s["energy"], s["lambdas"], fr, version=version, offd=offd)
        e.make_pair_lists()
        e.pot_energy()
        if logs is not None:
            for k in range(n):
                logs[k].write("\n" + summary(k, nsteps, True))
                if st[k]["npt"]:
                    bl, _ = e.box(k)
                    tries, acc, _ = e.volume_moves(k)
                    logs[k].write("\n" + volume_summary(bl[0] * bl[1] * bl[2], tries, acc) + "\n")
        # write_xfin comes after the closing pot_energy, whose water
        # polarisation update (istep = nsteps) the restart's theta_corr holds
        for k, s in enumerate(st):
            if s["final"]:
                x, v = e.state(k)
                qe.write_restart(s["final"], x, v, e.theta_corr(k), box=_box(e, k))
        return [e.energies(k) for k in range(n)]
    finally:
        e.close()
