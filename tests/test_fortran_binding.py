"""The ISO_C_BINDING module (qsource_amd/fortran/qgpu.f90) compiles with the
image's flang, links against libqgpu.so and calls through the C ABI."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "qsource_amd", "fortran")
LIBDIR = os.path.join(ROOT, "qsource_amd")

flang = shutil.which("flang")
pytestmark = pytest.mark.skipif(flang is None or not os.path.exists(os.path.join(LIBDIR, "libqgpu.so")),
                                reason="flang or libqgpu.so missing")


def _build(tmp_path):
    subprocess.run([flang, "-c", os.path.join(FDIR, "qgpu.f90"), "-o", "qgpu.o"], cwd=tmp_path, check=True)
    subprocess.run([flang, os.path.join(FDIR, "qgpu_check.f90"), "qgpu.o", "-o", "qgpu_check",
                    "-L" + LIBDIR, "-lqgpu", "-Wl,-rpath," + LIBDIR], cwd=tmp_path, check=True)
    return os.path.join(str(tmp_path), "qgpu_check")


def test_fortran_error_path(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, str(tmp_path / "missing.inp"), "0"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode != 0
    assert "qgpu error:" in r.stdout


@pytest.mark.gpu
def test_fortran_host_runs_steps(tmp_path, c1_small):
    """A Fortran host creates the engine from a qdyn input, runs 20 steps and
    reads the potential energy; it equals the same run through ctypes."""
    print("case: Fortran host, R=10 droplet, 20 FP64 steps", flush=True)
    from qsource_amd import engine as qe
    exe = _build(tmp_path)
    r = subprocess.run([exe, c1_small["inp"], "20"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120, cwd=os.path.dirname(c1_small["inp"]))
    assert r.returncode == 0, r.stdout
    pot_f = float(r.stdout.split("potential")[1].split()[0])
    e = qe.Engine([c1_small["inp"]], qe.QGPU_FP64)
    e.step(20)
    pot = e.energies(0)["potential"]
    e.close()
    assert np.isclose(pot_f, pot, rtol=0, atol=2e-6), (pot_f, pot)
