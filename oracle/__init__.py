"""ORACLE — test infrastructure only.

ctypes view of ``oracle/liboracle.so``, the FP64 CPU restatement of serial
reference qdyn's MD step (oracle/qdyn_oracle.cpp).  Imported only by the
tests, the smoke check and the benchmark's CPU-baseline leg, always as the
checker or the timed CPU baseline, never as the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_QDYN = os.path.join(HERE, "_ref", "qdyn5")

ENERGY_NAMES = [
    "pp_el", "pp_vdw", "p_bond", "p_angle", "p_torsion", "p_improper",
    "ww_el", "ww_vdw", "w_bond", "w_angle", "w_torsion", "w_improper",
    "pw_el", "pw_vdw", "lrf",
    "qx_el", "qx_vdw", "q_bond", "q_angle", "q_torsion", "q_improper",
    "rst_total", "rst_fix", "rst_radial", "rst_wpol", "rst_shell", "rst_protein",
    "total", "potential", "kinetic",
]

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.c_char_p]
        lib.oracle_last_error.restype = ctypes.c_char_p
        for name in ["oracle_destroy", "oracle_natom", "oracle_nstates", "oracle_run",
                     "oracle_final_energy", "oracle_get_xv", "oracle_get_d", "oracle_get_energies",
                     "oracle_get_qenergies", "oracle_list_sizes", "oracle_get_list", "oracle_get_lrf",
                     "oracle_temperature", "oracle_theta_corr", "oracle_nshells", "oracle_offdiag",
                     "oracle_sizes", "oracle_pot_energy_nonbonds", "oracle_shake", "oracle_pot_energy",
                     "oracle_box"]:
            getattr(lib, name).argtypes = None
        lib.oracle_temperature.restype = ctypes.c_double
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    def __init__(self, inp: str):
        self.lib = _load()
        h = self.lib.oracle_create(os.fsencode(inp))
        if not h:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        self.h = ctypes.c_void_p(h)
        self.natom = self.lib.oracle_natom(self.h)
        self.nstates = self.lib.oracle_nstates(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n: int):
        rc = self.lib.oracle_run(self.h, ctypes.c_int(n))
        if rc:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def pot_energy(self):
        """pot_energy at the current coordinates over the current lists."""
        if self.lib.oracle_pot_energy(self.h):
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def final_energy(self):
        if self.lib.oracle_final_energy(self.h):
            raise RuntimeError(self.lib.oracle_last_error().decode())

    def state(self):
        x = np.zeros(3 * self.natom)
        v = np.zeros(3 * self.natom)
        self.lib.oracle_get_xv(self.h, _dp(x), _dp(v))
        return x, v

    def gradient(self):
        d = np.zeros(3 * self.natom)
        self.lib.oracle_get_d(self.h, _dp(d))
        return d

    def energies(self) -> dict:
        e = np.zeros(len(ENERGY_NAMES))
        self.lib.oracle_get_energies(self.h, _dp(e), ctypes.c_int(len(e)))
        return dict(zip(ENERGY_NAMES, e.tolist()))

    def qenergies(self) -> np.ndarray:
        """Per state: qq_el qq_vdw qp_el qp_vdw qw_el qw_vdw total bond angle
        torsion improper restraint."""
        e = np.zeros(12 * max(self.nstates, 1))
        n = self.lib.oracle_get_qenergies(self.h, _dp(e), ctypes.c_int(len(e)))
        return e[:n].reshape(-1, 12)

    def monitor(self) -> np.ndarray:
        """Monitored group pairs at the current coordinates: per pair Vwsum,
        Vwel, Vwlj, then Vel and Vlj of every state."""
        cap = 4096
        out = np.zeros(cap)
        n = self.lib.oracle_monitor(self.h, _dp(out), ctypes.c_int(cap))
        if n < 0:
            raise RuntimeError("too many monitored pairs")
        ns = max(self.nstates, 1)
        return out[:n].reshape(-1, 3 + 2 * ns)

    def list_sizes(self):
        n = (ctypes.c_long * 5)()
        self.lib.oracle_list_sizes(self.h, n)
        return list(n)

    def pairlist(self, which: int) -> np.ndarray:
        n = self.list_sizes()[which]
        out = np.zeros(3 * max(n, 1), dtype=np.int32)
        self.lib.oracle_get_list(self.h, ctypes.c_int(which), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out[:3 * n].reshape(-1, 3)

    def lrf(self, ncgp: int) -> np.ndarray:
        """(ncgp, 43) LRF records of the last list build: centre, phi0,
        phi1(3), phi2(9), phi3(27)."""
        out = np.zeros(43 * ncgp)
        self.lib.oracle_get_lrf(self.h, _dp(out))
        return out.reshape(ncgp, 43)

    def pot_energy_nonbonds(self, x: np.ndarray, d: np.ndarray):
        """pot_energy_nonbonds at x over the current lists; adds into d and
        returns (E7, EQ (nstates, 4)) as the engine's entry point does."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        e = np.zeros(7)
        eq = np.zeros(max(4 * self.nstates, 1))
        if self.lib.oracle_pot_energy_nonbonds(self.h, _dp(x), _dp(d), _dp(e), _dp(eq)):
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return e, eq[:4 * self.nstates].reshape(self.nstates, 4)

    def shake(self, xx: np.ndarray, x: np.ndarray) -> int:
        xx = np.ascontiguousarray(xx, dtype=np.float64).ravel()
        n = ctypes.c_int(0)
        if self.lib.oracle_shake(self.h, _dp(xx), _dp(x), ctypes.byref(n)):
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return n.value

    def sizes(self) -> dict:
        n = (ctypes.c_int * 3)()
        self.lib.oracle_sizes(self.h, n)
        return dict(nwat=n[0], nqat=n[1], nat_solute=n[2])

    def temperature(self) -> float:
        return self.lib.oracle_temperature(self.h)

    def theta_corr(self) -> np.ndarray:
        n = self.lib.oracle_nshells(self.h)
        out = np.zeros(max(n, 1))
        self.lib.oracle_theta_corr(self.h, _dp(out))
        return out[:n]

    def box(self):
        """Periodic box: (box lengths, box centre), or None for a sphere."""
        out = np.zeros(6)
        if not self.lib.oracle_box(self.h, _dp(out)):
            return None
        return out[:3].copy(), out[3:].copy()

    def volume_moves(self):
        """Constant pressure: (tries, accepted, per try 0 = rejected,
        1 = accepted with deltaW <= 0, 2 = accepted by the random draw)."""
        n = self.lib.oracle_volume_moves(self.h, (ctypes.c_int * 2)(), ctypes.c_int(2))
        out = (ctypes.c_int * (n + 2))()
        self.lib.oracle_volume_moves(self.h, out, ctypes.c_int(n + 2))
        return out[0], out[1], [int(f) for f in out[2:n + 2]]

    def offdiag(self) -> np.ndarray:
        """(noffd, 2): H_ij and r_kl of each [off_diagonals] line at the
        current coordinates."""
        h = np.zeros(16)
        r = np.zeros(16)
        n = self.lib.oracle_offdiag(self.h, _dp(h), _dp(r), ctypes.c_int(16))
        return np.stack([h[:n], r[:n]], axis=1)
