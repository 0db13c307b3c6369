"""The oracle's boundary calls, pinned by what the mathematics fixes (the
oracle's step functions themselves are pinned bit for bit against the
reference's qdyn5 runs in test_oracle_reference.py and
test_reference_basic.py):

* pot_energy_nonbonds (src/md.f90:16261-16344) at coordinates away from the
  list build, without LRF: the accumulated gradient equals central finite
  differences of the energy it returns (pp, pw, ww el + vdW plus the
  lambda-weighted Q-solute and Q-water energies of each state, which is the
  energy whose gradient the reference's lambda-weighted Q forces are), to
  1e-7 relative; the classical and Q-water energies equal pot_energy's at
  the same coordinates; d is added into, not overwritten;
* shake(xx, x) (src/md.f90:17581-17679): the constrained start geometry is
  a fixed point; after a random displacement the molecules' mass-weighted
  centres do not move (the corrections of a constraint are opposite and
  scaled by 1/mass), every correction lies in the plane of the old water
  (it runs along the old bond vectors), and the constraint errors drop by
  three orders of magnitude, with over 99 % of them below 1e-4 d0^2.
"""
import numpy as np

from oracle import Oracle
from qsource_amd import gensys

LAM = [0.7, 0.3]


def _system(tmp_path, lrf):
    s = gensys.solvated_system(radius=14.0, n_solute=40, seed=2003, n_q=8, n_states=2)
    p = gensys.write_system(s, str(tmp_path), "nl", steps=10, output=10, lambdas=LAM, lrf=lrf, stepsize=0.5,
                            temperature=50.0, bath_coupling=2.0, initial_temperature=50.0)
    return s, p["inp"]


def test_nonbonds_gradient_is_energy_derivative(tmp_path):
    _, inp = _system(tmp_path, lrf=False)
    o = Oracle(inp)
    o.final_energy()
    x0, _ = o.state()
    rng = np.random.default_rng(1)
    x1 = x0 + 0.02 * rng.standard_normal(x0.size)
    lam = np.array(LAM)

    def V(x):
        d = np.zeros_like(x)
        e, eq = o.pot_energy_nonbonds(x, d)
        return e[:6].sum() + (lam[:, None] * eq).sum(), d, e

    _, d, e = V(x1)
    assert e[6] == 0.0                       # no LRF in this input
    h = 1e-5
    for i in rng.choice(x0.size, 24, replace=False):
        xp, xm = x1.copy(), x1.copy()
        xp[i] += h
        xm[i] -= h
        fd = (V(xp)[0] - V(xm)[0]) / (2 * h)
        assert abs(fd - d[i]) <= 1e-7 * max(1.0, abs(d[i])), (i, fd, d[i])
    # d accumulates
    d1 = np.full(x0.size, 2.5)
    V(x1)
    e2, _ = o.pot_energy_nonbonds(x1, d1)
    assert np.array_equal(d1 - 2.5, d) or np.abs(d1 - 2.5 - d).max() <= 1e-12 * np.abs(d).max()
    o.close()


def test_nonbonds_energies_equal_pot_energy(tmp_path):
    _, inp = _system(tmp_path, lrf=True)
    o = Oracle(inp)
    o.final_energy()
    E = o.energies()
    EQ = o.qenergies()
    x0, _ = o.state()
    e, eq = o.pot_energy_nonbonds(x0, np.zeros_like(x0))
    assert e.tolist() == [E["pp_el"], E["pp_vdw"], E["pw_el"], E["pw_vdw"], E["ww_el"], E["ww_vdw"], E["lrf"]]
    # Q-water equal; pot_energy's EQ%qp also holds the Q-neighbour pairs that
    # nonbond_qq adds (src/md.f90:13491-13636), which are not nonbonds'
    assert np.array_equal(eq[:, 2:4], EQ[:, 4:6])
    o.close()


def test_shake_properties(tmp_path):
    s, inp = _system(tmp_path, lrf=True)
    o = Oracle(inp)
    xx, _ = o.state()
    z = o.sizes()
    ns = z["nat_solute"]
    r2 = lambda X, a, b: np.sum((X[3 * a:3 * a + 3] - X[3 * b:3 * b + 3]) ** 2)
    mass = np.array([15.9994, 1.008, 1.008])
    # the start geometry is (to rounding) the constrained one: one
    # iteration per water, 341 waters over 342 molecules -> 0 by integer
    # division, and the coordinates barely move
    x = xx.copy()
    assert o.shake(xx, x) == 0
    assert np.abs(x - xx).max() <= 1e-6
    rng = np.random.default_rng(5)
    x = xx + 0.05 * rng.standard_normal(xx.size)
    x_in = x.copy()
    assert o.shake(xx, x) >= 2
    before, after = [], []
    for iw in range(z["nwat"]):
        a = ns + 3 * iw
        for i, j in ((a, a + 1), (a, a + 2), (a + 1, a + 2)):
            d0 = r2(xx, i, j)
            before.append(abs(r2(x_in, i, j) - d0) / d0)
            after.append(abs(r2(x, i, j) - d0) / d0)
        dx = (x - x_in)[3 * a:3 * a + 9].reshape(3, 3)
        # opposite corrections scaled by 1/mass: the centre of mass stays
        assert np.abs(mass @ dx).max() <= 1e-12, iw
        # corrections run along the old bond vectors: in the old plane
        P = xx[3 * a:3 * a + 9].reshape(3, 3)
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        assert np.abs(dx @ (nrm / np.linalg.norm(nrm))).max() <= 1e-12, iw
    before, after = np.array(before), np.array(after)
    # the constraints are met, but not all of them: each constraint's test
    # is taken before its correction, and a constraint found in range is
    # not looked at again while the others of its molecule move its atoms
    assert before.mean() > 0.05 and after.mean() < 1e-3 * before.mean()
    assert np.mean(after < 1e-4) > 0.99
    # solute atoms carry no constraints here and stay where they were
    assert np.array_equal(x[:3 * ns], x_in[:3 * ns])
    o.close()
