"""The allocation-free MLP gradient entry point (qgpu_mlp_gradient_noalloc)
against qgpu_mlp_gradient, through the C ABI: two calls in a row, the second
with fewer samples, each bitwise equal to the allocating entry point's
gradient and loss (same kernels, same launch shapes, same sample order); a
sample count above the static scratch array's maximum is refused."""
import ctypes

import numpy as np
import pytest

from qsource_amd import engine as qe

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _noalloc(m, X, Y):
    lib = m.lib
    lib.qgpu_mlp_gradient_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    g = np.zeros(m.nparams, dtype=np.float32)
    loss = ctypes.c_double()
    rc = lib.qgpu_mlp_gradient_noalloc(m.h, _fp(X), _fp(Y), len(X), _fp(g), ctypes.byref(loss))
    if rc != 0:
        raise qe.QgpuError(lib.qgpu_mlp_last_error().decode())
    return g, loss.value


def test_gradient_noalloc_matches_allocating_entry():
    print("case: MLP gradient without allocation, 1500 then 300 samples", flush=True)
    rng = np.random.default_rng(5)
    m = qe.FarFieldMLP(seed=3)
    X = rng.normal(size=(1500, 4)).astype(np.float32)
    Y = rng.normal(size=(1500, 3)).astype(np.float32)
    m.set_data(X, Y)             # normalisation from these samples
    for n in (1500, 300):
        g_ref, l_ref = m.gradient(X[:n], Y[:n])
        g, loss = _noalloc(m, X[:n], Y[:n])
        assert np.array_equal(g, g_ref), n
        assert loss == l_ref, (n, loss, l_ref)
        assert np.abs(g).max() > 0
    with pytest.raises(qe.QgpuError, match="at most 2048 samples"):
        _noalloc(m, np.zeros((2049, 4)), np.zeros((2049, 3)))
    m.close()
