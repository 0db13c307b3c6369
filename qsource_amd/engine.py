"""Python view of the qgpu C ABI (include/qgpu.h) over ctypes.

This is the host-side mirror of qdyn's step interface for tests, the smoke
check and the benchmark.  It loads the in-tree ``libqgpu.so`` and fails
loudly when the library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqgpu.so")

QGPU_FP64 = 0
QGPU_MIXED = 1

ENERGY_NAMES = [
    "pp_el", "pp_vdw", "p_bond", "p_angle", "p_torsion", "p_improper",
    "ww_el", "ww_vdw", "w_bond", "w_angle", "w_torsion", "w_improper",
    "pw_el", "pw_vdw", "lrf",
    "qx_el", "qx_vdw", "q_bond", "q_angle", "q_torsion", "q_improper",
    "rst_total", "rst_fix", "rst_radial", "rst_wpol", "rst_shell", "rst_protein",
    "total", "potential", "kinetic",
]

# exported symbols declared in include/qgpu.h
SYMBOLS = [
    "qgpu_create", "qgpu_destroy", "qgpu_natom", "qgpu_nstates", "qgpu_make_pair_lists", "qgpu_pot_energy",
    "qgpu_step", "qgpu_get_state", "qgpu_set_state", "qgpu_get_gradient", "qgpu_energies",
    "qgpu_qenergies", "qgpu_get_pairlist", "qgpu_temperature", "qgpu_theta_corr",
    "qgpu_last_step_ms", "qgpu_last_error", "qgpu_set_energy_interval", "qgpu_energy_series",
    "qgpu_energy_series_fp64", "qgpu_noffdiag", "qgpu_offdiag_series",
    "qgpu_mlp_create", "qgpu_mlp_destroy", "qgpu_mlp_nparams", "qgpu_mlp_get_params", "qgpu_mlp_set_params",
    "qgpu_mlp_set_data", "qgpu_mlp_train", "qgpu_mlp_gradient", "qgpu_mlp_predict", "qgpu_mlp_last_error",
    "qgpu_far_samples", "qgpu_far_lrf_estimate", "qgpu_set_far_model",
    "qgpu_profile_kernels", "qgpu_list_counts", "qgpu_offdiag_states", "qgpu_monitor",
    "qgpu_trj_create", "qgpu_trj_natoms", "qgpu_trj_write", "qgpu_trj_close", "qgpu_write_restart",
    "qgpu_write_restart_box", "qgpu_box", "qgpu_volume_moves",
    "qgpu_autotune", "qgpu_set_launch_shape",
    "qgpu_instance_status", "qgpu_list_capacity", "qgpu_set_list_capacity", "qgpu_pot_energy_terms",
    "qgpu_work_counts", "qgpu_mlp_gradient_noalloc", "qgpu_system_sizes", "qgpu_offdiag",
    "qgpu_pot_energy_nonbonds", "qgpu_shake",
]
TERM_GROUPS = {"ww": 1, "pw": 2, "pp": 4, "qx": 8, "lrf": 16, "other": 32}
TUNED_KERNELS = ["water_lists", "solute_lists", "water_forces", "solute_forces"]

KERNEL_NAMES = ["cgp_centers", "water_lists", "solute_lists", "water_forces", "solute_forces",
                "terms", "q_forces", "term_gather", "watpol_theta", "watpol_force",
                "energy_record", "integrate", "temperature"]

_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libqgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libqgpu.so not built: run python -m qsource_amd.build")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    lib.qgpu_create.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(P)]
    lib.qgpu_destroy.argtypes = [P]
    lib.qgpu_natom.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_nstates.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_make_pair_lists.argtypes = [P]
    lib.qgpu_pot_energy.argtypes = [P]
    lib.qgpu_step.argtypes = [P, ctypes.c_int]
    lib.qgpu_get_state.argtypes = [P, ctypes.c_int, dp, dp]
    lib.qgpu_set_state.argtypes = [P, ctypes.c_int, dp, dp]
    lib.qgpu_get_gradient.argtypes = [P, ctypes.c_int, dp]
    lib.qgpu_energies.argtypes = [P, ctypes.c_int, dp, ctypes.c_int]
    lib.qgpu_qenergies.argtypes = [P, ctypes.c_int, dp, ctypes.c_int]
    lib.qgpu_get_pairlist.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_long)]
    lib.qgpu_temperature.argtypes = [P, ctypes.c_int, dp]
    lib.qgpu_theta_corr.argtypes = [P, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_last_step_ms.argtypes = [P, dp]
    lib.qgpu_set_energy_interval.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.qgpu_profile_kernels.argtypes = [P, ctypes.c_int, dp, ctypes.c_int]
    lib.qgpu_list_counts.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
    lib.qgpu_energy_series.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_energy_series_fp64.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_noffdiag.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_offdiag_series.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.qgpu_mlp_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.qgpu_mlp_destroy.argtypes = [P]
    lib.qgpu_mlp_nparams.argtypes = [ip]
    lib.qgpu_mlp_get_params.argtypes = [P, fp, fp]
    lib.qgpu_mlp_set_params.argtypes = [P, fp, fp]
    lib.qgpu_mlp_set_data.argtypes = [P, fp, fp, ctypes.c_int]
    lib.qgpu_mlp_train.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_int, dp]
    lib.qgpu_mlp_gradient.argtypes = [P, fp, fp, ctypes.c_int, fp, dp]
    lib.qgpu_mlp_predict.argtypes = [P, fp, ctypes.c_int, fp]
    lib.qgpu_mlp_last_error.restype = ctypes.c_char_p
    lib.qgpu_far_samples.argtypes = [P, ctypes.c_int, fp, fp, ctypes.c_int, ip]
    lib.qgpu_far_lrf_estimate.argtypes = [P, ctypes.c_int, fp, fp, ctypes.c_int, ip]
    lib.qgpu_set_far_model.argtypes = [P, P]
    lib.qgpu_offdiag_states.argtypes = [P, ctypes.c_int, ip, ctypes.c_int, ip]
    lib.qgpu_monitor.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ip]
    lib.qgpu_autotune.argtypes = [P, ctypes.c_int, dp, ip]
    lib.qgpu_set_launch_shape.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.qgpu_pot_energy_terms.argtypes = [P, ctypes.c_int]
    lib.qgpu_pot_energy_nonbonds.argtypes = [P, ctypes.c_int, dp, dp, dp, dp, ctypes.c_int]
    lib.qgpu_shake.argtypes = [P, ctypes.c_int, dp, dp, ip]
    lib.qgpu_system_sizes.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    lib.qgpu_offdiag.argtypes = [P, ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.qgpu_work_counts.argtypes = [P, ctypes.c_int, dp, ctypes.c_int]
    lib.qgpu_instance_status.argtypes = [P, ctypes.c_int, ip, ip]
    lib.qgpu_list_capacity.argtypes = [P, ctypes.c_int, ip, ip]
    lib.qgpu_set_list_capacity.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.qgpu_trj_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(P)]
    lib.qgpu_trj_natoms.argtypes = [P]
    lib.qgpu_trj_write.argtypes = [P, dp]
    lib.qgpu_trj_close.argtypes = [P]
    lib.qgpu_write_restart.argtypes = [ctypes.c_char_p, ctypes.c_int, dp, dp, ctypes.c_int, dp]
    lib.qgpu_write_restart_box.argtypes = [ctypes.c_char_p, ctypes.c_int, dp, dp, dp]
    lib.qgpu_box.argtypes = [P, ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.qgpu_volume_moves.argtypes = [P, ctypes.c_int, ip, ctypes.c_int, ip]
    lib.qgpu_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class QgpuError(RuntimeError):
    pass


class Engine:
    """A batch of qdyn instances (one per input file) on the GPU."""

    def __init__(self, inputs: Sequence[str], precision: int = QGPU_FP64):
        self.lib = load_library()
        arr = (ctypes.c_char_p * len(inputs))(*[os.fsencode(p) for p in inputs])
        h = ctypes.c_void_p()
        self._check(self.lib.qgpu_create(arr, len(inputs), precision, ctypes.byref(h)))
        self.h = h
        self.n = len(inputs)
        self.natom = []
        self.nstates = []
        for k in range(self.n):
            n = ctypes.c_int()
            self._check(self.lib.qgpu_natom(self.h, k, ctypes.byref(n)))
            self.natom.append(n.value)
            self._check(self.lib.qgpu_nstates(self.h, k, ctypes.byref(n)))
            self.nstates.append(n.value)

    def _check(self, rc: int):
        if rc != 0:
            raise QgpuError(self.lib.qgpu_last_error().decode())

    def sizes(self, k: int = 0) -> dict:
        """Atoms, solute atoms, waters, charge groups, solute charge groups,
        Q atoms and FEP states of instance k."""
        v = (ctypes.c_int * 7)()
        self._check(self.lib.qgpu_system_sizes(self.h, k, v, 7))
        return dict(zip(["natom", "nat_solute", "nwat", "ncgp", "ncgp_solute", "nqat", "nstates"], list(v)))

    def nat_solute(self, k: int = 0) -> int:
        return self.sizes(k)["nat_solute"]

    def close(self):
        if getattr(self, "h", None):
            self.lib.qgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        self._check(self.lib.qgpu_step(self.h, n))

    def make_pair_lists(self):
        self._check(self.lib.qgpu_make_pair_lists(self.h))

    def pot_energy(self, terms: Optional[Sequence[str]] = None):
        """Gradient and energies at the current coordinates; with `terms`
        (names of TERM_GROUPS) only those term groups enter the gradient."""
        if terms is None:
            self._check(self.lib.qgpu_pot_energy(self.h))
        else:
            self._check(self.lib.qgpu_pot_energy_terms(self.h, sum(TERM_GROUPS[t] for t in set(terms))))

    def pot_energy_nonbonds(self, k: int, x: np.ndarray, d: np.ndarray):
        """pot_energy_nonbonds of instance k at coordinates x (which become
        its coordinates): adds the gradient into d (in place) and returns
        (E, EQ): E = pp_el pp_vdw pw_el pw_vdw ww_el ww_vdw lrf, EQ
        (nstates, 4) = qp_el qp_vdw qw_el qw_vdw."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        assert d.dtype == np.float64 and d.flags.c_contiguous and d.size == x.size == 3 * self.natom[k]
        ns = self.nstates[k]
        e = np.zeros(7)
        eq = np.zeros(max(4 * ns, 1))
        self._check(self.lib.qgpu_pot_energy_nonbonds(self.h, k, _dp(x), _dp(d), _dp(e), _dp(eq), 4 * ns))
        return e, eq[:4 * ns].reshape(ns, 4)

    def shake(self, k: int, xx: np.ndarray, x: np.ndarray) -> int:
        """shake(xx, x) of instance k: constrains x in place and returns the
        average iterations per molecule."""
        xx = np.ascontiguousarray(xx, dtype=np.float64).ravel()
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == xx.size == 3 * self.natom[k]
        n = ctypes.c_int(0)
        self._check(self.lib.qgpu_shake(self.h, k, _dp(xx), _dp(x), ctypes.byref(n)))
        return n.value

    def state(self, k: int = 0):
        x = np.zeros(3 * self.natom[k])
        v = np.zeros(3 * self.natom[k])
        self._check(self.lib.qgpu_get_state(self.h, k, _dp(x), _dp(v)))
        return x, v

    def set_state(self, k: int, x: np.ndarray, v: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        self._check(self.lib.qgpu_set_state(self.h, k, _dp(x), _dp(v)))

    def gradient(self, k: int = 0):
        d = np.zeros(3 * self.natom[k])
        self._check(self.lib.qgpu_get_gradient(self.h, k, _dp(d)))
        return d

    def energies(self, k: int = 0) -> dict:
        e = np.zeros(len(ENERGY_NAMES))
        self._check(self.lib.qgpu_energies(self.h, k, _dp(e), len(e)))
        return dict(zip(ENERGY_NAMES, e.tolist()))

    def qenergies(self, k: int = 0) -> np.ndarray:
        """Per-state Q energies: rows of qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw
        total bond angle torsion improper restraint."""
        e = np.zeros(12 * max(1, self.nstates[k]))
        self._check(self.lib.qgpu_qenergies(self.h, k, _dp(e), len(e)))
        return e[:12 * self.nstates[k]].reshape(-1, 12)

    def pairlist(self, k: int, which: int) -> np.ndarray:
        n = ctypes.c_long(0)
        self._check(self.lib.qgpu_get_pairlist(self.h, k, which, None, ctypes.byref(n)))
        out = np.zeros(3 * max(n.value, 1), dtype=np.int32)
        self._check(self.lib.qgpu_get_pairlist(self.h, k, which,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                               ctypes.byref(n)))
        return out[:3 * n.value].reshape(-1, 3)

    def list_counts(self, k: int = 0) -> dict:
        """Reference-layout list sizes: nbpp, nbpw, nbww, nbqp, qw atom pairs."""
        c = (ctypes.c_long * 5)()
        self._check(self.lib.qgpu_list_counts(self.h, k, c))
        return dict(zip(["pp", "pw", "ww", "qp", "qw"], list(c)))

    def work_counts(self, k: int = 0) -> dict:
        """Work of one list update of instance k (roofline accounting)."""
        c = np.zeros(6)
        self._check(self.lib.qgpu_work_counts(self.h, k, _dp(c), 6))
        return dict(zip(["lrf_evals_water_side", "lrf_evals_solute_side", "ww_entries", "wp_entries",
                         "pp_group_pairs", "pw_entries"], c.tolist()))

    def temperature(self, k: int = 0) -> float:
        t = ctypes.c_double()
        self._check(self.lib.qgpu_temperature(self.h, k, ctypes.byref(t)))
        return t.value

    def box(self, k: int = 0):
        """Periodic box of instance k: (lengths, current centre), or None for
        a sphere."""
        out = np.zeros(6)
        isb = ctypes.c_int()
        self._check(self.lib.qgpu_box(self.h, k, _dp(out), ctypes.byref(isb)))
        return (out[:3].copy(), out[3:].copy()) if isb.value else None

    def volume_moves(self, k: int = 0):
        """Constant-pressure volume moves of instance k: (tries, accepted,
        per move 0 = rejected, 1 = accepted with deltaW <= 0, 2 = accepted
        by the random draw)."""
        n = ctypes.c_int()
        self._check(self.lib.qgpu_volume_moves(self.h, k, None, 0, ctypes.byref(n)))
        out = (ctypes.c_int * (n.value + 2))()
        self._check(self.lib.qgpu_volume_moves(self.h, k, out, n.value + 2, ctypes.byref(n)))
        return out[0], out[1], [int(f) for f in out[2:n.value + 2]]

    def theta_corr(self, k: int = 0) -> np.ndarray:
        n = ctypes.c_int()
        self._check(self.lib.qgpu_theta_corr(self.h, k, None, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._check(self.lib.qgpu_theta_corr(self.h, k, _dp(out), ctypes.byref(n)))
        return out[:n.value]

    def set_energy_interval(self, iene: int, max_frames: int):
        self._check(self.lib.qgpu_set_energy_interval(self.h, iene, max_frames))
        self._max_frames = max_frames

    def offdiag_series(self, k: int = 0) -> np.ndarray:
        """(nframes, noffd, 2): EVB off-diagonal Hij and r_kl of each recorded
        frame."""
        m = ctypes.c_int()
        self._check(self.lib.qgpu_noffdiag(self.h, k, ctypes.byref(m)))
        cap = max(1, getattr(self, "_max_frames", 0)) * max(1, m.value) * 2
        out = np.zeros(cap)
        n = ctypes.c_int()
        self._check(self.lib.qgpu_offdiag_series(self.h, k, _dp(out), cap, ctypes.byref(n)))
        if m.value == 0:
            return np.zeros((n.value, 0, 2))
        return out[:n.value * m.value * 2].reshape(n.value, m.value, 2)

    def autotune(self, reps: int = 3) -> dict:
        """Times every launch shape of the four large kernels on this batch
        and keeps the fastest; returns per kernel (mapping, workgroup size,
        prefetch, register budget, chosen ms, default ms)."""
        ms = np.zeros(8)
        sh = np.zeros(16, dtype=np.int32)
        self._check(self.lib.qgpu_autotune(self.h, reps, _dp(ms), sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return {name: (int(sh[4 * k]), int(sh[4 * k + 1]), int(sh[4 * k + 2]), int(sh[4 * k + 3]),
                       float(ms[2 * k]), float(ms[2 * k + 1]))
                for k, name in enumerate(TUNED_KERNELS)}

    def set_launch_shape(self, kernel: str, mapping: int, block: int, prefetch: int = 0, occupancy: int = 0):
        self._check(self.lib.qgpu_set_launch_shape(self.h, TUNED_KERNELS.index(kernel), mapping, block, prefetch,
                                                   occupancy))

    def monitor(self, k: int = 0) -> np.ndarray:
        """Monitored group pairs at the current coordinates: per pair Vwsum,
        Vwel, Vwvdw, then Vel and Vvdw of every state (nonbond_monitor)."""
        n = ctypes.c_int()
        self._check(self.lib.qgpu_monitor(self.h, k, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._check(self.lib.qgpu_monitor(self.h, k, _dp(out), len(out), ctypes.byref(n)))
        ns = max(self.nstates[k], 1)
        return out[:n.value].reshape(-1, 3 + 2 * ns) if n.value else np.zeros((0, 3 + 2 * ns))

    def offdiag(self, k: int = 0) -> np.ndarray:
        """(noffd, 6): state i, state j, Q atoms k, l (1-based), Hij and r_kl
        of every off-diagonal function at the current coordinates."""
        n = ctypes.c_int()
        self._check(self.lib.qgpu_offdiag(self.h, k, None, 0, ctypes.byref(n)))
        out = np.zeros(6 * max(1, n.value))
        self._check(self.lib.qgpu_offdiag(self.h, k, _dp(out), len(out), ctypes.byref(n)))
        return out[:6 * n.value].reshape(-1, 6)

    def offdiag_states(self, k: int = 0) -> np.ndarray:
        """(noffd, 2) 1-based state pairs (i, j) of the off-diagonal functions."""
        ij = np.zeros(2 * 16, dtype=np.int32)
        n = ctypes.c_int()
        self._check(self.lib.qgpu_offdiag_states(self.h, k, ij.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 len(ij), ctypes.byref(n)))
        return ij[:2 * n.value].reshape(-1, 2)

    def energy_series(self, k: int = 0, fp64: bool = False) -> np.ndarray:
        """(nframes, nstates, 14) put_ene records recorded on the device.
        With fp64=True (mixed-precision batches): the same frames with the Q
        energies re-evaluated with FP64 pair arithmetic."""
        ns = max(1, self.nstates[k])
        cap = max(1, getattr(self, "_max_frames", 0)) * ns * 14
        out = np.zeros(cap)
        n = ctypes.c_int()
        fn = self.lib.qgpu_energy_series_fp64 if fp64 else self.lib.qgpu_energy_series
        self._check(fn(self.h, k, _dp(out), cap, ctypes.byref(n)))
        return out[:n.value * ns * 14].reshape(n.value, ns, 14)

    def profile_kernels(self, nsteps: int) -> dict:
        """Average ms per launch of each kernel over nsteps event-timed steps."""
        ms = np.zeros(len(KERNEL_NAMES))
        self._check(self.lib.qgpu_profile_kernels(self.h, nsteps, _dp(ms), len(ms)))
        return dict(zip(KERNEL_NAMES, ms.tolist()))

    def last_step_ms(self) -> float:
        t = ctypes.c_double()
        self._check(self.lib.qgpu_last_step_ms(self.h, ctypes.byref(t)))
        return t.value

    def status(self, k: int = 0):
        """(failure bits, step found at): 1 list row over capacity, 2 water
        shell capacity, 4 SHAKE failure, 8 NaN temperature; (0, -1) when none."""
        b, at = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.qgpu_instance_status(self.h, k, ctypes.byref(b), ctypes.byref(at)))
        return b.value, at.value

    def list_capacity(self, k: int = 0):
        """((ww, wp, pw) row capacities of instance k, list regrowths of the batch)."""
        caps = (ctypes.c_int * 3)()
        n = ctypes.c_int()
        self._check(self.lib.qgpu_list_capacity(self.h, k, caps, ctypes.byref(n)))
        return tuple(caps), n.value

    def set_list_capacity(self, k: int, ww: int = 0, wp: int = 0, pw: int = 0):
        self._check(self.lib.qgpu_set_list_capacity(self.h, k, ww, wp, pw))


class FarFieldMLP:
    """The learned far-field surrogate (4-128-128-3 SiLU MLP, BF16 matrix-core
    kernels) trained and evaluated on the GPU."""

    def __init__(self, seed: int = 1):
        self.lib = load_library()
        h = ctypes.c_void_p()
        self._check(self.lib.qgpu_mlp_create(seed, ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int()
        self.lib.qgpu_mlp_nparams(ctypes.byref(n))
        self.nparams = n.value

    def _check(self, rc: int):
        if rc != 0:
            raise QgpuError(self.lib.qgpu_mlp_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.qgpu_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        """(FP32 parameters, normalisation: xmean[4] xstd[4] ymean[3] ystd[3])."""
        p = np.zeros(self.nparams, dtype=np.float32)
        nm = np.zeros(14, dtype=np.float32)
        self._check(self.lib.qgpu_mlp_get_params(self.h, _fp(p), _fp(nm)))
        return p, nm

    def set_params(self, p: np.ndarray, norm: Optional[np.ndarray] = None):
        p = np.ascontiguousarray(p, dtype=np.float32)
        nm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self._check(self.lib.qgpu_mlp_set_params(self.h, _fp(p), _fp(nm) if nm is not None else None))

    def set_data(self, X: np.ndarray, Y: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        self._check(self.lib.qgpu_mlp_set_data(self.h, _fp(X), _fp(Y), len(X)))

    def train(self, epochs: int, batch: int = 1024, lr: float = 2e-3, seed: int = 1) -> float:
        loss = ctypes.c_double()
        self._check(self.lib.qgpu_mlp_train(self.h, epochs, batch, ctypes.c_float(lr), seed, ctypes.byref(loss)))
        return loss.value

    def gradient(self, X: np.ndarray, Y: np.ndarray):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self.lib.qgpu_mlp_gradient(self.h, _fp(X), _fp(Y), len(X), _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 3), dtype=np.float32)
        self._check(self.lib.qgpu_mlp_predict(self.h, _fp(X), len(X), _fp(Y)))
        return Y


def far_samples(engine: "Engine", k: int = 0, lrf_estimate: bool = False):
    """Features (ncgp, 4) and far-field (ncgp, 3) of instance k: exact at the
    current coordinates, or the current LRF multipoles' estimate."""
    cap = 1 << 20
    X = np.zeros((cap, 4), dtype=np.float32)
    Y = np.zeros((cap, 3), dtype=np.float32)
    n = ctypes.c_int()
    fn = engine.lib.qgpu_far_lrf_estimate if lrf_estimate else engine.lib.qgpu_far_samples
    engine._check(fn(engine.h, k, _fp(X), _fp(Y), cap, ctypes.byref(n)))
    return X[:n.value].copy(), Y[:n.value].copy()


def set_far_model(engine: "Engine", model: Optional[FarFieldMLP]):
    engine._check(engine.lib.qgpu_set_far_model(engine.h, model.h if model is not None else None))


class Trajectory:
    """qdyn DCD trajectory of the run of `input_path` (mask of its
    [trajectory_atoms] section; written to `path`, or to the input's
    [files] trajectory)."""

    def __init__(self, input_path: str, path: Optional[str] = None):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.qgpu_trj_create(os.fsencode(input_path), os.fsencode(path) if path else None,
                                      ctypes.byref(h))
        if rc != 0:
            raise QgpuError(self.lib.qgpu_last_error().decode())
        self.h = h
        self.natoms = self.lib.qgpu_trj_natoms(self.h)

    def write(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        if self.lib.qgpu_trj_write(self.h, _dp(x)) != 0:
            raise QgpuError(self.lib.qgpu_last_error().decode())

    def close(self):
        if self.h:
            self.lib.qgpu_trj_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_restart(path: str, x: np.ndarray, v: np.ndarray, theta_corr: Optional[np.ndarray] = None,
                  box: Optional[np.ndarray] = None):
    """Restart file as qdyn's write_xfin writes it; `box` (6 values: lengths
    and centre) for a periodic box."""
    lib = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    if box is not None:
        b = np.ascontiguousarray(box, dtype=np.float64).ravel()
        if lib.qgpu_write_restart_box(os.fsencode(path), x.size // 3, _dp(x), _dp(v), _dp(b)) != 0:
            raise QgpuError(lib.qgpu_last_error().decode())
        return
    tc = np.ascontiguousarray(theta_corr if theta_corr is not None else np.zeros(0), dtype=np.float64)
    if lib.qgpu_write_restart(os.fsencode(path), x.size // 3, _dp(x), _dp(v), tc.size,
                              _dp(tc) if tc.size else None) != 0:
        raise QgpuError(lib.qgpu_last_error().decode())
