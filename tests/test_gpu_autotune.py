"""The launch-shape tuner: every launch shape of the four large kernels
(2-D grid, instance-major work map, XCD-grouped work map; workgroups of 256
or 128; partner prefetch in the force kernels; the FP32 water-force build
fitted to 6 or 4 wavefronts per SIMD) gives bitwise the same trajectory on a batch of uneven instances,
in FP64 and mixed precision, and the tuner's choice is never slower than
the default shape it was measured against."""
import numpy as np
import pytest

from qsource_amd import engine as qe
from qsource_amd import gensys

pytestmark = pytest.mark.gpu

COOL = dict(stepsize=0.5, temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)


@pytest.fixture(scope="module")
def batch(tmp_path_factory):
    d = tmp_path_factory.mktemp("tune")
    inputs = []
    for k, (R, ns) in enumerate([(10.0, 20), (13.0, 40), (15.0, 60)]):
        s = gensys.solvated_system(radius=R, n_solute=ns, seed=4100 + k, n_q=4, n_states=2)
        p = gensys.write_system(s, str(d), "b%d" % k, steps=20, lambdas=[0.5, 0.5], **COOL)
        inputs.append(p["inp"])
    return inputs


def _run(inputs, precision, shape=None):
    e = qe.Engine(inputs, precision)
    for kname in qe.TUNED_KERNELS:
        e.set_launch_shape(kname, *(shape if shape is not None else (0, 256, 0, 0)))
    e.step(12)                        # crosses a pair-list step
    out = [np.concatenate(e.state(k)) for k in range(len(inputs))]
    e.close()
    return out


@pytest.mark.parametrize("precision", [qe.QGPU_FP64, qe.QGPU_MIXED])
def test_every_launch_shape_is_bitwise_identical(batch, precision):
    print("case: launch shapes, 3-instance batch, precision %d" % precision, flush=True)
    ref = _run(batch, precision)
    for mapping in (0, 1, 2):
        for block in (256, 128):
            for pf in (0, 1):
                for occ in (0, 1):
                    got = _run(batch, precision, (mapping, block, pf, occ))
                    for a, b in zip(ref, got):
                        assert np.array_equal(a, b), (mapping, block, pf, occ)


def test_autotune_choice(batch):
    print("case: autotune on a 3-instance batch", flush=True)
    e = qe.Engine(batch, qe.QGPU_MIXED)
    res = e.autotune(reps=3)
    for name, (mapping, block, pf, occ, best, dflt) in res.items():
        assert mapping in (0, 1, 2) and block in (128, 256) and pf in (0, 1) and occ in (0, 1)
        assert 0 < best <= dflt, (name, best, dflt)
    e.step(12)
    e.close()
