"""Synthetic qdyn systems: TIP3P droplets with optional branched solutes and
FEP Q regions, written in qdyn's own file formats.

Every system is generated from a seeded numpy ``default_rng`` and written as
the four files qdyn reads:

* the topology (``.top``), in the layout read by ``topo_read``
  (src/topo.f90:515-1074);
* the MD input (``.inp``), sectioned key/value as read by ``initialize``
  (src/md.f90:4453-5358);
* the FEP file (``.fep``): ``[FEP] states``, ``[atoms]`` and
  ``[change_charges]`` as read by ``qatom_load_atoms``/``qatom_load_fep``
  (src/qatom.f90:301-760);
* optionally a restart (``.re``): Fortran unformatted sequential records,
  canary -1337, ``nat3, x``, ``nat3, v``, ``nwpolr_shell, theta_corr``
  (src/md.f90:18705-18736).

The same files drive the reference qdyn (CPU) and this engine, so parity is
checked on identical inputs.
"""
from __future__ import annotations

import dataclasses
import math
import os
import struct
from typing import List, Optional, Sequence

import numpy as np

# TIP3P geometry and parameters
R_OH = 0.9572
THETA_HOH = 104.52
Q_O, Q_H = -0.834, 0.417
SIGMA_O, EPS_O = 3.1506, 0.1521
MASS_O, MASS_H = 15.9994, 1.008
SIGMA_H, EPS_H = 0.4, 0.046
DENSITY = 0.0334  # molecules / A^3
MAX_NBR_RANGE = 25  # src/topo.f90:36


def lj_sqrt_ab(sigma: float, eps: float):
    """Geometric-rule topology parameters: sqrt(A)=2 sqrt(eps) sigma^6,
    sqrt(B)=2 sqrt(eps) sigma^3, so A_ij=avdw_i*avdw_j (src/md.f90:13340)."""
    return 2.0 * math.sqrt(eps) * sigma ** 6, 2.0 * math.sqrt(eps) * sigma ** 3


@dataclasses.dataclass
class System:
    """A qdyn topology plus run settings (all atom indices are 1-based, as in
    the files)."""
    x: np.ndarray                  # (N,3) coordinates
    iac: np.ndarray                # (N,) atom type codes, 1-based
    crg: np.ndarray                # (N,) charges
    nat_solute: int
    masses: np.ndarray             # (natyps,)
    avdw: np.ndarray               # (natyps,3)
    bvdw: np.ndarray               # (natyps,3)
    tac: List[str]
    bonds: np.ndarray              # (nb,3) i j cod
    nbonds_solute: int
    bondlib: np.ndarray            # (nbc,2) fk b0
    angles: np.ndarray             # (na,4)
    nangles_solute: int
    anglib: np.ndarray             # (nac,2) fk theta0(deg)
    tors: np.ndarray               # (nt,5)
    ntors_solute: int
    torlib: np.ndarray             # (ntc,4) fk rmult deltor(deg) paths
    imps: np.ndarray               # (ni,5)
    nimps_solute: int
    implib: np.ndarray             # (nic,2) fk imp0(deg)
    cgps: List[Sequence[int]]      # atoms of each charge group; first = switch
    ncgp_solute: int
    list14: np.ndarray             # (nat_solute, 25) bool: list14[i-1, k-1] <=> (i, i+k) 1-4
    listex: np.ndarray             # (nat_solute, 25) bool
    list14long: np.ndarray         # (n,2)
    listexlong: np.ndarray         # (n,2)
    res_start: np.ndarray
    res_names: List[str]
    nres_solute: int
    mol_start: np.ndarray
    rexcl_o: float
    rwat: float
    xpcent: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    xwcent: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    el14_scale: float = 0.5
    coulomb_constant: float = 332.0
    # Q region (optional)
    qatoms: Optional[np.ndarray] = None    # (nqat,) topology atom numbers
    qcrg: Optional[np.ndarray] = None      # (nqat, nstates)
    # optional per-state Q-atom types ([atom_types] + [change_atoms]):
    # qtypes rows (name, A1, B1, A2, B2, A3, B3, mass), qtype_of (nqat, nstates)
    # 0-based rows of qtypes
    qtypes: Optional[list] = None
    qtype_of: Optional[np.ndarray] = None
    softcore: Optional[np.ndarray] = None  # (nqat, nstates) alphas ([softcore])
    offdiag: Optional[list] = None         # (state_i, state_j, q_k, q_l, A, mu), 1-based
    ff_type: int = 1                       # 1 GROMOS (harmonic impropers), 2 AMBER (periodic)
    # Q-bonded FEP sections (optional): type rows and terms with one type
    # code per state (0 = absent): qbond_types (Dmz, alpha, r0) with qbonds
    # (i, j, codes); qangle_types (fk, theta0) with qangles (i, j, k, codes);
    # qtorsion_types (fk, mult, delta) with qtorsions (i, j, k, l, codes);
    # qimproper_types (fk, imp0) with qimpropers (i, j, k, l, codes);
    # qcouplings {"angle"|"torsion"|"improper": [(term, bond, flag)]}, 1-based
    qbond_types: Optional[list] = None
    qbonds: Optional[list] = None
    qangle_types: Optional[list] = None
    qangles: Optional[list] = None
    qtorsion_types: Optional[list] = None
    qtorsions: Optional[list] = None
    qimproper_types: Optional[list] = None
    qimpropers: Optional[list] = None
    qcouplings: Optional[dict] = None
    # input-file restraints ([sequence|atom|distance|angle|wall]_restraints rows)
    restraints: Optional[dict] = None
    # FEP Q-Q options: [soft_pairs] (iq, jq), [el_scale] (iq, jq, scales),
    # [excluded_pairs] (i, j, flags) rows, 1-based; qq_use_library_charges
    soft_pairs: Optional[list] = None
    el_scale: Optional[list] = None
    excluded_pairs: Optional[list] = None
    qq_lib_charges: bool = False
    # [shake_constraints] rows (i, j, distance per state), 1-based
    shake_constraints: Optional[list] = None
    # [monitor_groups] (lists of atoms) and [monitor_group_pairs] (group
    # numbers), 1-based
    monitor_groups: Optional[list] = None
    monitor_pairs: Optional[list] = None
    # topology solvent type: 0 SPC-like (LJ on the water O only), 1 all-atom
    # LJ on every solvent atom (TIP3P-like H types)
    solvent_type: int = 0
    # excluded atoms (topology flags, one per atom), None = none
    excl: Optional[np.ndarray] = None
    # periodic box: edge lengths (x, y, z) and centre; None = sphere
    box: Optional[np.ndarray] = None
    box_centre: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))

    @property
    def natom(self) -> int:
        return int(self.x.shape[0])

    @property
    def nwat(self) -> int:
        return (self.natom - self.nat_solute) // 3


# ---------------------------------------------------------------------------
# geometry


def _random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
                  np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3),
                  np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    a, b, c, d = q[:, 3], q[:, 0], q[:, 1], q[:, 2]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = a * a + b * b - c * c - d * d
    R[:, 0, 1] = 2 * (b * c - a * d)
    R[:, 0, 2] = 2 * (b * d + a * c)
    R[:, 1, 0] = 2 * (b * c + a * d)
    R[:, 1, 1] = a * a - b * b + c * c - d * d
    R[:, 1, 2] = 2 * (c * d - a * b)
    R[:, 2, 0] = 2 * (b * d - a * c)
    R[:, 2, 1] = 2 * (c * d + a * b)
    R[:, 2, 2] = a * a - b * b - c * c + d * d
    return R


def water_lattice(radius: float, rng, avoid: Optional[np.ndarray] = None,
                  avoid_r: float = 3.0):
    """Waters on a cubic lattice of spacing DENSITY^(-1/3) with O inside the
    sphere, random orientations; lattice sites within avoid_r of any solute
    atom are dropped.  Returns (nw,3,3) coordinates O,H1,H2."""
    a = DENSITY ** (-1.0 / 3.0)
    n = int(math.ceil(radius / a)) + 1
    g = np.arange(-n, n + 1) * a
    O = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    O = O[np.linalg.norm(O, axis=1) <= radius]
    if avoid is not None and len(avoid):
        # in blocks of lattice sites (the full site x atom array of a large
        # solute would not fit in memory); the same distances as one pass
        keep = np.zeros(len(O), dtype=bool)
        for b0 in range(0, len(O), 1024):
            d2 = ((O[b0:b0 + 1024, None, :] - avoid[None, :, :]) ** 2).sum(-1)
            keep[b0:b0 + 1024] = d2.min(1) > avoid_r ** 2
        O = O[keep]
    nw = O.shape[0]
    half = math.radians(THETA_HOH) / 2
    local = np.array([[0, 0, 0],
                      [R_OH * math.sin(half), 0, R_OH * math.cos(half)],
                      [-R_OH * math.sin(half), 0, R_OH * math.cos(half)]])
    rot = _random_rotations(rng, nw)
    w = O[:, None, :] + np.einsum("nij,kj->nki", rot, local)
    return w


def water_box(length: float, rng, avoid: Optional[np.ndarray] = None, avoid_r: float = 3.0):
    """Waters filling a periodic cube of edge `length` centred at the
    origin: a cubic lattice of round(length * DENSITY^(1/3)) sites per edge
    (spacing length / n, so the lattice continues across the faces), random
    orientations; sites within avoid_r of a solute atom are dropped.
    Returns (nw,3,3) coordinates O,H1,H2."""
    n = max(1, int(round(length * DENSITY ** (1.0 / 3.0))))
    a = length / n
    g = -length / 2 + (np.arange(n) + 0.5) * a
    O = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if avoid is not None and len(avoid):
        keep = np.zeros(len(O), dtype=bool)
        for b0 in range(0, len(O), 1024):
            d = O[b0:b0 + 1024, None, :] - avoid[None, :, :]
            d -= length * np.round(d / length)
            keep[b0:b0 + 1024] = (d ** 2).sum(-1).min(1) > avoid_r ** 2
        O = O[keep]
    nw = O.shape[0]
    half = math.radians(THETA_HOH) / 2
    local = np.array([[0, 0, 0],
                      [R_OH * math.sin(half), 0, R_OH * math.cos(half)],
                      [-R_OH * math.sin(half), 0, R_OH * math.cos(half)]])
    rot = _random_rotations(rng, nw)
    return O[:, None, :] + np.einsum("nij,kj->nki", rot, local)


def _branched_solute(n_atoms: int, rmax: float, rng, bond=1.53, min_dist=3.2):
    """A branched chain grown as a self-avoiding walk confined to r < rmax:
    the chain grows from its tip, a one-atom branch is added on every other
    backbone atom, and when the tip is trapped the walk resumes from another
    backbone atom with a free valence (a branch point).  Atoms closer than
    three bonds to the new atom may sit closer than min_dist; all others may
    not.  Atoms are numbered in growth order.  Returns coords (n,3) and the
    bond list (0-based pairs)."""
    coords = np.zeros((n_atoms, 3))
    nbr: List[List[int]] = [[] for _ in range(n_atoms)]
    bonds = []
    backbone: List[int] = [0]
    n = 1
    # placed atoms by cell of edge min_dist: an atom closer than min_dist to
    # a candidate lies in one of the 27 cells around the candidate's
    cells: dict = {(0, 0, 0): [0]}

    def cell_of(p):
        return (int(math.floor(p[0] / min_dist)), int(math.floor(p[1] / min_dist)),
                int(math.floor(p[2] / min_dist)))

    def close_atoms(p, skip):
        c = cell_of(p)
        out = []
        for a in (c[0] - 1, c[0], c[0] + 1):
            for b in (c[1] - 1, c[1], c[1] + 1):
                for d in (c[2] - 1, c[2], c[2] + 1):
                    out.extend(k for k in cells.get((a, b, d), ()) if k not in skip)
        return out

    def near(parent):
        # atoms within two bonds of the parent: 1-2, 1-3 and 1-4 of the new atom
        s1 = {parent}
        s2 = set(nbr[parent])
        s3 = set(k for a in s2 for k in nbr[a])
        return s1 | s2 | s3

    def place_from(parent, tries=60):
        skip = near(parent)
        # 1-3 partners at >= 2.35 A (angle >~ 100 deg), 1-4 partners >= 2.6 A
        n13 = np.array(nbr[parent], dtype=int)
        n14 = np.array(sorted(skip - {parent} - set(nbr[parent])), dtype=int)
        c13 = coords[n13] if len(n13) else np.zeros((0, 3))
        c14 = coords[n14] if len(n14) else np.zeros((0, 3))
        prev = nbr[parent][0] if nbr[parent] else None
        for _ in range(tries):
            if prev is None:
                d = rng.normal(size=3)
            else:
                u = coords[parent] - coords[prev]
                u /= np.linalg.norm(u)
                t = rng.normal(size=3)
                t -= t.dot(u) * u
                t /= np.linalg.norm(t)
                ang = math.radians(180 - 111.0 + rng.normal(0, 4))
                d = math.cos(ang) * u + math.sin(ang) * t
            d /= np.linalg.norm(d)
            p = coords[parent] + bond * d
            if np.linalg.norm(p) > rmax:
                continue
            close = close_atoms(p, skip)
            if close and ((coords[close] - p) ** 2).sum(1).min() < min_dist ** 2:
                continue
            if len(c13) and ((c13 - p) ** 2).sum(1).min() < 2.40 ** 2:
                continue
            if len(c14) and ((c14 - p) ** 2).sum(1).min() < 2.9 ** 2:
                continue
            return p
        return None

    def add(parent, p):
        nonlocal n
        coords[n] = p
        cells.setdefault(cell_of(p), []).append(n)
        nbr[parent].append(n)
        nbr[n].append(parent)
        bonds.append((parent, n))
        n += 1
        return n - 1

    tip = 0
    stuck = 0
    while n < n_atoms:
        p = place_from(tip)
        if p is None:
            stuck += 1
            if stuck > 50000:
                raise RuntimeError("solute growth failed; increase rmax")
            open_atoms = [a for a in backbone if len(nbr[a]) < 3]
            tip = open_atoms[int(rng.integers(len(open_atoms)))]
            continue
        i = add(tip, p)
        backbone.append(i)
        tip = i
        if len(backbone) % 2 == 0 and n < n_atoms:
            q = place_from(i)
            if q is not None:
                add(i, q)
    c = coords.copy()
    c -= c.mean(0)
    return c, bonds


def _renumber_and_group(coords, bonds, size=4):
    """Renumbers the solute atoms in depth-first order of the bond tree (so
    bonded atoms get close indices) and partitions them, in that order, into
    connected charge groups of at most `size` atoms: a group is closed when
    it is full or when the next atom is not bonded to any of its members.
    Compact groups keep every atom near its group centre, which the LRF
    Taylor expansion about the group centre requires."""
    n = len(coords)
    nbr = [[] for _ in range(n)]
    for i, j in bonds:
        nbr[i].append(j)
        nbr[j].append(i)
    order, seen, stack = [], [False] * n, [0]
    while stack:
        a = stack.pop()
        if seen[a]:
            continue
        seen[a] = True
        order.append(a)
        for b in sorted(nbr[a], reverse=True):
            if not seen[b]:
                stack.append(b)
    new = {old: k for k, old in enumerate(order)}
    c2 = coords[order]
    b2 = sorted((min(new[i], new[j]), max(new[i], new[j])) for i, j in bonds)
    nb2 = [set() for _ in range(n)]
    for i, j in b2:
        nb2[i].add(j)
        nb2[j].add(i)
    groups, cur = [], []
    for a in range(n):
        if cur and (len(cur) >= size or not any(a in nb2[m] for m in cur)):
            groups.append(cur)
            cur = []
        cur.append(a)
    if cur:
        groups.append(cur)
    return c2, b2, groups


def _graph_terms(n, bonds):
    nbr = [[] for _ in range(n)]
    for i, j in bonds:
        nbr[i].append(j)
        nbr[j].append(i)
    angles = []
    for j in range(n):
        for a in range(len(nbr[j])):
            for b in range(a + 1, len(nbr[j])):
                angles.append((nbr[j][a], j, nbr[j][b]))
    tors = []
    for j, k in bonds:
        for i in nbr[j]:
            if i == k:
                continue
            for l in nbr[k]:
                if l == j or l == i:
                    continue
                tors.append((i, j, k, l))
    # topological distance up to 3
    dist = {}
    for s in range(n):
        frontier = {s}
        seen = {s: 0}
        for d in range(1, 4):
            nxt = set()
            for u in frontier:
                for w in nbr[u]:
                    if w not in seen:
                        seen[w] = d
                        nxt.add(w)
            frontier = nxt
        for w, d in seen.items():
            if w > s and d > 0:
                dist[(s, w)] = d
    return angles, tors, dist


# ---------------------------------------------------------------------------
# system builders


def water_droplet(radius: float = 15.0, seed: int = 1001) -> System:
    """TIP3P droplet (configuration 1: R=15 A, ~470 waters)."""
    return solvated_system(radius=radius, n_solute=0, seed=seed)


def solvated_system(radius: float, n_solute: int, seed: int, n_q: int = 0,
                    n_states: int = 0, n_types: int = 250,
                    q_sd: float = 0.3, q_types: bool = False,
                    softcore: Optional[float] = None, n_offdiag: int = 0,
                    impropers: int = 0, q_bonded: bool = False, restraints: bool = False,
                    fep_pairs: bool = False, qq_lib_charges: bool = False,
                    q_shake: bool = False, solvent_allatom: bool = False,
                    solute_rmax: Optional[float] = None, exclude_outside: bool = False,
                    exclude_water_radius: Optional[float] = None, hydrogens: bool = False,
                    monitor: bool = False, group_size: int = 4, box: Optional[float] = None,
                    solute_offset: Optional[Sequence[float]] = None,
                    shift: Optional[Sequence[float]] = None) -> System:
    """A TIP3P sphere of the given radius around an optional branched solute
    (charges N(0,q_sd) neutralised per 4-atom charge group, sigma~U[2.5,3.8],
    eps~U[0.05,0.2] quantised into at most n_types atom types), with an
    optional Q region of n_q solute atoms and n_states FEP states whose
    charges are drawn independently and neutralised per 4-atom block.

    q_types: every Q atom also gets an LJ type per state (sigma, eps drawn as
    for the solute), which switches qdyn to Q-atom types (qvdw);
    softcore: with q_types, soft-core alphas drawn from U[0, softcore] per Q
    atom and state; n_offdiag: EVB couplings A*exp(-mu r) with
    A~U[10,60], mu~U[0,0.5] between random state pairs and Q-atom pairs;
    impropers: 1 adds harmonic (GROMOS) and 2 periodic (AMBER) impropers on
    the solute's branch points; solvent_allatom: the water H gets the LJ
    type of CHARMM's TIP3P H (sigma 0.4 A, eps 0.046 kcal/mol) and the
    topology declares solvent type 1 (all-atom solvent LJ); solute_rmax: the
    solute grows within this radius instead of radius - 5 (so it may reach
    through the sphere boundary); exclude_outside: every solute charge group
    whose switch atom lies beyond the exclusion radius is excluded;
    exclude_water_radius: waters whose O lies beyond it are excluded too;
    hydrogens: about half of the solute's terminal atoms become hydrogens
    (mass 1.008, sigma 2.4 A, eps 0.022, own atom type and bond type);
    monitor: three monitored atom groups (the Q atoms, the solute atoms
    bonded to them, and the first eight solute atoms) and their three
    pairs; group_size: the most atoms of a solute charge group (connected
    atoms in bond-tree order); box: a periodic cube of this edge instead of
    the sphere (water_box; the solute grows within radius - 5 of its
    centre, moved by solute_offset before the waters are placed; every atom
    moved by shift at the end, so molecules may start outside the box)."""
    rng = np.random.default_rng(seed)
    # --- solute
    if n_solute > 0:
        rmax = solute_rmax if solute_rmax is not None else max(6.0, radius - 5.0)
        for attempt in range(20):
            try:
                sx, sb = _branched_solute(n_solute, rmax, rng)
                break
            except RuntimeError:
                if attempt == 19:
                    raise
                rmax *= 1.05
        sx, sb, groups0 = _renumber_and_group(sx, sb, size=group_size)
        angles0, tors0, dist = _graph_terms(n_solute, sb)
    else:
        sx = np.zeros((0, 3))
        sb, angles0, tors0, dist, groups0 = [], [], [], {}, []
    ns = n_solute
    if box is not None:
        if solute_offset is not None and ns:
            sx = sx + np.asarray(solute_offset, dtype=float)
        wat = water_box(box, rng, avoid=sx if ns else None)
    else:
        wat = water_lattice(radius, rng, avoid=sx if ns else None)
    nw = wat.shape[0]
    x = np.concatenate([sx, wat.reshape(-1, 3)], 0)
    N = ns + 3 * nw

    # --- atom types: water O, water H, then quantised solute types
    masses = [MASS_O, MASS_H]
    a_O, b_O = lj_sqrt_ab(SIGMA_O, EPS_O)
    a_H, b_H = lj_sqrt_ab(SIGMA_H, EPS_H) if solvent_allatom else (0.0, 0.0)
    avdw = [[a_O] * 3, [a_H] * 3]
    bvdw = [[b_O] * 3, [b_H] * 3]
    tac = ["OW", "HW"]
    iac = np.zeros(N, dtype=np.int64)
    crg = np.zeros(N)
    if ns:
        sig = rng.uniform(2.5, 3.8, ns)
        eps = rng.uniform(0.05, 0.2, ns)
        ntyp = min(n_types - 2, ns)
        # quantise (sigma, eps) on a grid: sqrt(ntyp) bins each
        nb = max(1, int(math.floor(math.sqrt(ntyp))))
        si = np.minimum((sig - 2.5) / 1.3 * nb, nb - 1).astype(int)
        ei = np.minimum((eps - 0.05) / 0.15 * nb, nb - 1).astype(int)
        key = si * nb + ei
        uniq = sorted(set(key.tolist()))
        tmap = {k: t for t, k in enumerate(uniq)}
        for k in uniq:
            s = 2.5 + (k // nb + 0.5) * 1.3 / nb
            e = 0.05 + (k % nb + 0.5) * 0.15 / nb
            a, b = lj_sqrt_ab(s, e)
            # 1-4 pairs (LJ type 3): sigma x 0.8, eps x 0.5
            a14, b14 = lj_sqrt_ab(0.8 * s, 0.5 * e)
            avdw.append([a, a, a14])
            bvdw.append([b, b, b14])
            masses.append(12.011)
            tac.append("S%d" % len(tac))
        iac[:ns] = np.array([3 + tmap[k] for k in key.tolist()])
        q = rng.normal(0.0, q_sd, ns)
        for grp in groups0:
            q[grp] -= q[grp].mean()
        crg[:ns] = np.round(q, 4)
        # keep groups exactly neutral after rounding
        for grp in groups0:
            crg[grp[-1]] -= round(float(crg[grp].sum()), 4)
    hset = set()
    if hydrogens and ns:
        deg = np.zeros(ns, dtype=int)
        for i, j in sb:
            deg[i] += 1
            deg[j] += 1
        hset = set(a for a in range(1, ns) if deg[a] == 1 and rng.random() < 0.5)
        if hset:
            a_h, b_h = lj_sqrt_ab(2.4, 0.022)
            a14, b14 = lj_sqrt_ab(0.8 * 2.4, 0.5 * 0.022)
            avdw.append([a_h, a_h, a14])
            bvdw.append([b_h, b_h, b14])
            masses.append(1.008)
            tac.append("HS")
            for a in hset:
                iac[a] = len(masses)
    iac[ns::3] = 1
    iac[ns + 1::3] = 2
    iac[ns + 2::3] = 2
    crg[ns::3] = Q_O
    crg[ns + 1::3] = Q_H
    crg[ns + 2::3] = Q_H

    # --- bonded terms: solute harmonic bonds/angles, 3-fold torsions
    bondlib = [[600.0, 1.53]]
    bonds = [(i + 1, j + 1, 1) for i, j in sb]
    anglib = [[80.0, 111.0]]
    angles = [(i + 1, j + 1, k + 1, 1) for i, j, k in angles0]
    torlib = [[0.3, 3.0, 0.0, 1.0]]
    tors = [(i + 1, j + 1, k + 1, l + 1, 1) for i, j, k, l in tors0]
    nbs, nas, nts = len(bonds), len(angles), len(tors)
    # water: O-H1, O-H2, H1-H2 (last bond of the topology is H1-H2), one angle
    hh = 2 * R_OH * math.sin(math.radians(THETA_HOH) / 2)
    bondlib += [[1106.0, R_OH], [0.0, round(hh, 4)]]
    if hset:
        # solute bonds to a hydrogen: their own type (code 4)
        bondlib.append([340.0, 1.53])
        bonds = [(i, j, 4 if (i - 1 in hset or j - 1 in hset) else c) for i, j, c in bonds]
    anglib += [[200.0, THETA_HOH]]
    for w in range(nw):
        o = ns + 3 * w + 1
        bonds += [(o, o + 1, 2), (o, o + 2, 2), (o + 1, o + 2, 3)]
        angles += [(o + 1, o, o + 2, 2)]

    # --- charge groups: connected solute groups of <= 4 atoms (first atom
    # switches), waters
    cgps = [[a + 1 for a in grp] for grp in groups0]
    ncgp_solute = len(cgps)
    cgps += [[ns + 3 * w + 1, ns + 3 * w + 2, ns + 3 * w + 3] for w in range(nw)]

    # --- neighbour lists
    list14 = np.zeros((ns, MAX_NBR_RANGE), dtype=bool)
    listex = np.zeros((ns, MAX_NBR_RANGE), dtype=bool)
    l14long, lexlong = [], []
    for (i, j), d in dist.items():
        k = j - i
        if d <= 2:
            if k <= MAX_NBR_RANGE:
                listex[i, k - 1] = True
            else:
                lexlong.append((i + 1, j + 1))
        elif d == 3:
            if k <= MAX_NBR_RANGE:
                list14[i, k - 1] = True
            else:
                l14long.append((i + 1, j + 1))

    # --- residues/molecules: solute is one residue and one molecule
    res_start = ([1] if ns else []) + [ns + 3 * w + 1 for w in range(nw)]
    res_names = (["SOL"] if ns else []) + ["HOH"] * nw
    mol_start = list(res_start)

    sysm = System(
        x=x, iac=iac, crg=crg, nat_solute=ns,
        masses=np.array(masses), avdw=np.array(avdw), bvdw=np.array(bvdw), tac=tac,
        bonds=np.array(bonds, dtype=np.int64).reshape(-1, 3), nbonds_solute=nbs,
        bondlib=np.array(bondlib),
        angles=np.array(angles, dtype=np.int64).reshape(-1, 4), nangles_solute=nas,
        anglib=np.array(anglib),
        tors=np.array(tors, dtype=np.int64).reshape(-1, 5), ntors_solute=nts,
        torlib=np.array(torlib),
        imps=np.zeros((0, 5), dtype=np.int64), nimps_solute=0,
        implib=np.zeros((0, 2)),
        cgps=cgps, ncgp_solute=ncgp_solute,
        list14=list14, listex=listex,
        list14long=np.array(l14long, dtype=np.int64).reshape(-1, 2),
        listexlong=np.array(lexlong, dtype=np.int64).reshape(-1, 2),
        res_start=np.array(res_start), res_names=res_names,
        nres_solute=1 if ns else 0, mol_start=np.array(mol_start),
        rexcl_o=radius, rwat=radius)
    if box is not None:
        sysm.box = np.array([box, box, box], dtype=float)
        if shift is not None:
            sysm.x = sysm.x + np.asarray(shift, dtype=float)

    if n_q > 0 and n_states > 0:
        # Q region: the n_q solute atoms nearest the centre, in whole
        # 4-atom charge groups
        gc = np.array([sx[np.array(g) - 1].mean(0) for g in cgps[:ncgp_solute]])
        order = np.argsort(np.linalg.norm(gc, axis=1))
        qa = []
        for gi in order:
            qa += cgps[gi]
            if len(qa) >= n_q:
                break
        qa = np.array(sorted(qa[:n_q]))
        qc = np.zeros((len(qa), n_states))
        for s in range(n_states):
            q = rng.normal(0.0, q_sd, len(qa))
            for b0 in range(0, len(qa), 4):
                q[b0:b0 + 4] -= q[b0:b0 + 4].mean()
            q = np.round(q, 4)
            for b0 in range(0, len(qa), 4):
                b1 = min(len(qa), b0 + 4)
                q[b1 - 1] -= round(float(q[b0:b1].sum()), 4)
            qc[:, s] = q
        sysm.qatoms = qa
        sysm.qcrg = qc
        if q_types:
            rows, tof = [], np.zeros((len(qa), n_states), dtype=np.int64)
            for k, a in enumerate(qa):
                mass = float(sysm.masses[sysm.iac[a - 1] - 1])
                for st in range(n_states):
                    sig, eps = rng.uniform(2.5, 3.8), rng.uniform(0.05, 0.2)
                    a1, b1 = lj_sqrt_ab(sig, eps)
                    a3, b3 = lj_sqrt_ab(0.8 * sig, 0.5 * eps)
                    tof[k, st] = len(rows)
                    rows.append(("Q%d_%d" % (k + 1, st + 1), a1, b1, a1, b1, a3, b3, mass))
            sysm.qtypes, sysm.qtype_of = rows, tof
            if softcore:
                sysm.softcore = np.round(rng.uniform(0.0, softcore, (len(qa), n_states)), 2)
        if n_offdiag > 0 and n_states > 1:
            off = []
            for _ in range(n_offdiag):
                si, sj = rng.choice(n_states, 2, replace=False)
                qk, ql = rng.choice(len(qa), 2, replace=False)
                off.append((int(si) + 1, int(sj) + 1, int(qk) + 1, int(ql) + 1,
                            round(float(rng.uniform(10, 60)), 2), round(float(rng.uniform(0, 0.5)), 3)))
            sysm.offdiag = off
    if impropers and n_solute > 0:
        # one improper per branch point j with neighbours (i, k, l): i-j-k-l
        nbr = [[] for _ in range(n_solute)]
        for i, j in sb:
            nbr[i].append(j)
            nbr[j].append(i)
        imps = [(nb[0] + 1, j + 1, nb[1] + 1, nb[2] + 1, 1) for j, nb in enumerate(nbr) if len(nb) >= 3]
        sysm.imps = np.array(imps, dtype=np.int64).reshape(-1, 5)
        sysm.nimps_solute = len(imps)
        sysm.ff_type = 1 if impropers == 1 else 2
        sysm.implib = np.array([[3.0, 0.0]]) if impropers == 1 else np.array([[1.5, 180.0]])
    if q_bonded and sysm.qatoms is not None and n_states >= 2:
        _q_bonded_terms(sysm, rng, n_states)
    if restraints and n_solute >= 40:
        sysm.restraints = _example_restraints(sysm, max(n_states, 0))
    if fep_pairs and sysm.qatoms is not None and n_states >= 2:
        _fep_pairs(sysm, rng, n_states)
    sysm.qq_lib_charges = bool(qq_lib_charges and sysm.qatoms is not None)
    if q_shake and sysm.qatoms is not None and n_states >= 2:
        # a bonded Q pair (1.53 A in the first state, 2 % longer in the
        # others) and a Q 1-3 pair at its current distance (+0.02 A later)
        qs = set(int(a) for a in sysm.qatoms)
        rows = []
        for i, j, _ in sysm.bonds[:sysm.nbonds_solute]:
            if int(i) in qs and int(j) in qs:
                rows.append((int(i), int(j), [1.53] + [round(1.53 * 1.02, 3)] * (n_states - 1)))
                break
        for i, j, k, _ in sysm.angles[:sysm.nangles_solute]:
            if int(i) in qs and int(k) in qs:
                d = float(np.linalg.norm(sysm.x[int(i) - 1] - sysm.x[int(k) - 1]))
                rows.append((int(i), int(k), [round(d, 3)] + [round(d + 0.02, 3)] * (n_states - 1)))
                break
        sysm.shake_constraints = rows
    sysm.solvent_type = 1 if solvent_allatom else 0
    if monitor and sysm.qatoms is not None and ns >= 16:
        qs = set(int(a) for a in sysm.qatoms)
        nbr = sorted(set(int(b) for i, j, _ in sysm.bonds[:sysm.nbonds_solute] for a, b in ((i, j), (j, i))
                         if int(a) in qs and int(b) not in qs))
        first = [a for a in range(1, ns + 1) if a not in qs and a not in nbr][:8]
        sysm.monitor_groups = [sorted(qs), nbr[:8], first]
        sysm.monitor_pairs = [(1, 2), (1, 3), (2, 3)]
    if (exclude_outside and ns) or exclude_water_radius is not None:
        excl = np.zeros(N, dtype=bool)
        if exclude_outside:
            for g in cgps[:ncgp_solute]:
                if np.linalg.norm(x[g[0] - 1] - sysm.xpcent) > sysm.rexcl_o:
                    excl[np.array(g) - 1] = True
        if exclude_water_radius is not None:
            for w in range(nw):
                o = ns + 3 * w
                if np.linalg.norm(x[o] - sysm.xwcent) > exclude_water_radius:
                    excl[o:o + 3] = True
        sysm.excl = excl
    return sysm


def _fep_pairs(sysm: System, rng, n_states: int) -> None:
    """[soft_pairs], [el_scale] and [excluded_pairs] over the Q region: the
    two Q-Q pairs furthest apart in the bond graph repel softly (with Q-atom
    types, whose LJ-code-2 parameters become C, alpha of C_i C_j
    exp(-alpha_i alpha_j r)); two pairs get per-state electrostatic scales;
    one Q-Q pair is excluded in the last state only, and a Q/non-Q and a
    non-Q/non-Q solute pair (the latter more than 25 atoms apart) in every
    state."""
    ns = sysm.nat_solute
    nbr = [set() for _ in range(ns + 1)]
    for i, j, _ in sysm.bonds[:sysm.nbonds_solute]:
        nbr[int(i)].add(int(j))
        nbr[int(j)].add(int(i))

    def dist(a):
        d = {a: 0}
        front = [a]
        while front:
            nxt = []
            for u in front:
                for w in nbr[u]:
                    if w not in d:
                        d[w] = d[u] + 1
                        nxt.append(w)
            front = nxt
        return d

    qa = [int(a) for a in sysm.qatoms]
    pairs = []
    for m, a in enumerate(qa):
        da = dist(a)
        for n in range(m + 1, len(qa)):
            pairs.append((da.get(qa[n], 99), m + 1, n + 1))
    pairs.sort(key=lambda t: (-t[0], t[1], t[2]))
    far = [(m, n) for d, m, n in pairs if d >= 4]
    if sysm.qtypes is not None:
        rows = []
        for name, a1, b1, a2, b2, a3, b3, mass in sysm.qtypes:
            rows.append((name, a1, b1, round(float(np.sqrt(rng.uniform(500, 2000))), 4),
                         round(float(np.sqrt(rng.uniform(2.0, 3.0))), 4), a3, b3, mass))
        sysm.qtypes = rows
        sysm.soft_pairs = far[:2]
    sysm.el_scale = [(m, n, [round(float(rng.uniform(0.5, 1.0)), 2) for _ in range(n_states)])
                     for m, n in (far[2:4] if len(far) >= 4 else far[:2])]
    ex = []
    if len(far) > 4:
        m, n = far[4]
        ex.append((qa[m - 1], qa[n - 1], [0] * (n_states - 1) + [1]))
    qs = set(qa)
    d0 = dist(qa[0])
    nq = [a for a in range(1, ns + 1) if a not in qs and d0.get(a, 99) >= 4 and abs(a - qa[0]) <= 25]
    if nq:
        ex.append((qa[0], nq[0], [1] * n_states))
    for a in range(1, ns + 1):
        if a in qs:
            continue
        da = dist(a)
        b = [c for c in range(a + 26, ns + 1) if c not in qs and da.get(c, 99) >= 4]
        if b:
            ex.append((a, b[0], [1] * n_states))
            break
    sysm.excluded_pairs = ex


def _example_restraints(sysm: System, n_states: int) -> dict:
    """One of every p_restrain kind on the solute, around its current
    geometry: sequence restraints to topology positions (heavy atoms), to the
    geometric and to the mass centre; position restraints; flat-bottom
    distance restraints active below and above their range and one in range
    (residue:atom notation); angle restraints; a wall with Morse attraction.
    FEP states are used where the system has them."""
    x = sysm.x
    st = lambda k: k if n_states >= k else 0
    dist = lambda i, j: float(np.linalg.norm(x[i - 1] - x[j - 1]))

    def ang(i, j, k):
        u, w = x[i - 1] - x[j - 1], x[k - 1] - x[j - 1]
        return float(np.degrees(np.arccos(np.dot(u, w) / np.linalg.norm(u) / np.linalg.norm(w))))

    f = lambda v: "%.4f" % v
    return {
        "sequence": [(1, 10, 2.0, 0), (11, 20, 1.5, 1, 1), (21, 30, 1.0, 0, 2)],
        "atom": [(5, f(x[4, 0] + 0.3), f(x[4, 1] - 0.2), f(x[4, 2] + 0.1), 1.0, 2.0, 3.0, 0),
                 (12, f(x[11, 0] - 0.2), f(x[11, 1] + 0.1), f(x[11, 2]), 2.0, 2.0, 2.0, st(2))],
        "distance": [(3, 15, f(dist(3, 15) - 0.5), f(dist(3, 15) - 0.2), 5.0, 0),
                     (7, 33, f(dist(7, 33) + 0.3), f(dist(7, 33) + 1.0), 4.0, st(1)),
                     ("1:12", "1:25", 0.0, f(dist(12, 25) + 0.5), 2.0, st(2))],
        "angle": [(2, 9, 4, f(ang(2, 9, 4) + 5.0), 10.0, st(1)), (6, 17, 8, f(ang(6, 17, 8) - 3.0), 20.0, 0)],
        "wall": [(1, 40, 9.0, 5.0, 1.5, 0.3, 1)],
    }


def _q_bonded_terms(sysm: System, rng, n_states: int) -> None:
    """Q-bonded FEP sections over the Q region (EVB-style): every solute
    bond, angle, torsion and improper whose atoms are all Q atoms gets a type
    per state; one topology bond breaks (Morse, present in the first state
    only) and one bond forms between the closest Q atoms not bonded in the
    topology (present in the last state only, r0 at their distance, with a
    new angle through it); the terms through the breaking bond are absent
    where it is, and angles/torsions/impropers are coupled to both bonds,
    including terms present in every state.  A bond that a topology improper
    touches is not chosen to break."""
    qs = set(int(a) for a in sysm.qatoms)
    sb = [(int(i), int(j)) for i, j, _ in sysm.bonds[:sysm.nbonds_solute]]
    qb = [b for b in sb if b[0] in qs and b[1] in qs]
    imp_atoms = [set(int(a) for a in r[:4]) for r in sysm.imps]
    free = [b for b in qb if not any(b[0] in ia and b[1] in ia for ia in imp_atoms)]
    nbr = {a: set() for a in range(1, sysm.nat_solute + 1)}
    for i, j in sb:
        nbr[i].add(j)
        nbr[j].add(i)
    # prefer a breaking bond with non-Q neighbours: the topology terms
    # through it that the FEP file does not redefine become Q terms
    outer = [b for b in free if (nbr[b[0]] | nbr[b[1]]) - qs]
    brk = outer[0] if outer else (free[0] if free else None)
    # forming bond: closest Q pair at topological distance >= 3
    x = sysm.x
    frm, frm_r = None, 1e9
    ql = sorted(qs)
    for m, i in enumerate(ql):
        for j in ql[m + 1:]:
            if j in nbr[i] or nbr[i] & nbr[j]:
                continue
            r = float(np.linalg.norm(x[i - 1] - x[j - 1]))
            if r < frm_r:
                frm, frm_r = (i, j), r
    bonds = list(qb) + ([frm] if frm else [])
    present = {b: [True] * n_states for b in bonds}
    if brk:
        present[brk] = [st == 0 for st in range(n_states)]
    if frm:
        present[frm] = [st == n_states - 1 for st in range(n_states)]

    def has(atoms, b):
        return any({atoms[m], atoms[m + 1]} == set(b) for m in range(len(atoms) - 1))

    def term_present(atoms):
        return [all(present[b][st] for b in bonds if has(atoms, b)) for st in range(n_states)]

    btypes, qbonds = [], []
    for b in bonds:
        codes = []
        for st in range(n_states):
            if present[b][st]:
                r0 = round(frm_r, 3) if b == frm else round(1.53 * float(rng.uniform(0.97, 1.03)), 3)
                btypes.append((round(float(rng.uniform(60, 110)), 2), round(float(rng.uniform(1.6, 2.4)), 3), r0))
                codes.append(len(btypes))
            else:
                codes.append(0)
        qbonds.append((b[0], b[1], codes))
    atypes, qangles = [], []
    for i, j, k, _ in sysm.angles[:sysm.nangles_solute]:
        a = (int(i), int(j), int(k))
        if not all(v in qs for v in a):
            continue
        pres = term_present(a)
        codes = []
        for st in range(n_states):
            if pres[st]:
                atypes.append((round(float(rng.uniform(50, 90)), 2), round(111.0 + float(rng.uniform(-4, 4)), 2)))
                codes.append(len(atypes))
            else:
                codes.append(0)
        qangles.append(a + (codes,))
    if frm:
        # a new angle n - i - j through the forming bond, at its current value
        i, j = frm
        n = min(nbr[i])
        u, w = x[n - 1] - x[i - 1], x[j - 1] - x[i - 1]
        th = float(np.degrees(np.arccos(np.dot(u, w) / np.linalg.norm(u) / np.linalg.norm(w))))
        atypes.append((40.0, round(th, 2)))
        qangles.append((n, i, j, [0] * (n_states - 1) + [len(atypes)]))
    ttypes, qtors = [], []
    for i, j, k, l, _ in sysm.tors[:sysm.ntors_solute]:
        t = (int(i), int(j), int(k), int(l))
        if not all(v in qs for v in t):
            continue
        pres = term_present(t)
        codes = []
        for st in range(n_states):
            if pres[st]:
                ttypes.append((round(float(rng.uniform(0.2, 1.5)), 3), float(rng.integers(1, 4)),
                               float(rng.choice([0.0, 180.0]))))
                codes.append(len(ttypes))
            else:
                codes.append(0)
        qtors.append(t + (codes,))
    itypes, qimps = [], []
    for r in sysm.imps[:sysm.nimps_solute]:
        t = tuple(int(v) for v in r[:4])
        if not all(v in qs for v in t):
            continue
        codes = []
        for st in range(n_states):
            itypes.append((round(float(rng.uniform(2, 6)), 3), round(float(rng.uniform(-10, 10)), 2)))
            codes.append(len(itypes))
        qimps.append(t + (codes,))
    coupl = {"angle": [], "torsion": [], "improper": []}
    ib = {b: n + 1 for n, b in enumerate(bonds)}
    for name, terms in (("angle", qangles), ("torsion", qtors), ("improper", qimps)):
        always = [n for n, t in enumerate(terms) if all(c > 0 for c in t[-1])]
        for n, t in enumerate(terms):
            atoms = t[:-1]
            if brk and has(atoms, brk):
                coupl[name].append((n + 1, ib[brk], 0))
            elif frm and has(atoms, frm):
                coupl[name].append((n + 1, ib[frm], 1))
        # terms present in every state coupled to the forming bond (the first
        # state sees the factor the bond left in the previous evaluation)
        # and to the breaking bond (later states see the first state's)
        for m, bnd, flag in zip(always[:2], (frm, brk), (0, 1)):
            if bnd and not any(c[0] == m + 1 for c in coupl[name]):
                coupl[name].append((m + 1, ib[bnd], flag))
    sysm.qbond_types, sysm.qbonds = btypes, qbonds
    sysm.qangle_types, sysm.qangles = atypes, qangles
    sysm.qtorsion_types, sysm.qtorsions = ttypes, qtors
    sysm.qimproper_types, sysm.qimpropers = itypes, qimps
    sysm.qcouplings = coupl


# ---------------------------------------------------------------------------
# writers


def _fmt_list(vals, per_line, fmt):
    out = []
    for k in range(0, len(vals), per_line):
        out.append("".join(fmt % v for v in vals[k:k + per_line]))
    return out


def write_top(s: System, path: str) -> None:
    """Write a qdyn topology (version 5.06 layout, sphere boundary)."""
    L = ["Q topology file",
         "TITLE      synthetic",
         "DATE       20260101",
         "VERSION     5.06",
         "FORCEFIELD synthetic",
         "FF_TYPE    %d" % s.ff_type,
         "END        of header"]
    N = s.natom
    L.append("%8d%8d%8d = Total no. of atoms, no. of solute atoms, solvent atoms" % (N, s.nat_solute, 3))
    L += _fmt_list([v for v in s.x.reshape(-1)], 6, "%12.6f")
    L.append("%8d = No. of integer atom codes. iac's:" % N)
    L += _fmt_list(list(s.iac), 16, "%5d")
    L.append("%8d%8d = No. of bonds, no. of solute bonds. i - j - icode: (5 per line)" % (len(s.bonds), s.nbonds_solute))
    L += _fmt_list([tuple(b) for b in s.bonds], 5, "%6d%6d%4d")
    L.append("%8d = No. of bond codes. Parameters:" % len(s.bondlib))
    for k, (fk, b0) in enumerate(s.bondlib):
        L.append("%5d%10.3f%11.4f" % (k + 1, fk, b0))
    L.append("%8d%8d = No. of angles, no. of solute angles. i - j - k - icode: (3 per line)" % (len(s.angles), s.nangles_solute))
    L += _fmt_list([tuple(a) for a in s.angles], 3, "%6d%6d%6d%4d")
    L.append("%8d = No. of angle codes. Parameters:" % len(s.anglib))
    for k, (fk, t0) in enumerate(s.anglib):
        L.append("%5d%10.3f%10.3f" % (k + 1, fk, t0))
    L.append("%8d%8d = No. of torsions, solute torsions. i - j - k - l - icode: (2 per line)" % (len(s.tors), s.ntors_solute))
    L += _fmt_list([tuple(t) for t in s.tors], 2, "%6d%6d%6d%6d%4d")
    L.append("%8d = No. of torsion codes. Parameters:" % len(s.torlib))
    for k, (fk, m, dt, p) in enumerate(s.torlib):
        L.append("%5d%10.3f%10.3f%10.3f%10.3f" % (k + 1, fk, m, dt, p))
    L.append("%8d%8d = No. of impropers, solute impr. i - j - k - l - icode: (2 per line)" % (len(s.imps), s.nimps_solute))
    L += _fmt_list([tuple(t) for t in s.imps], 2, "%6d%6d%6d%6d%4d")
    L.append("%8d%8d = No. of improper codes, type (1=harmonic,2=periodic). Parameters:" % (len(s.implib), s.ff_type))
    for k, (fk, i0) in enumerate(s.implib):
        L.append("%5d%10.3f%10.3f" % (k + 1, fk, i0))
    L.append("%8d = No. of atomic charges" % N)
    L += _fmt_list(list(s.crg), 10, "%9.4f")
    L.append("%8d%8d%8d = No. of charge groups, no. of solute cgps, switch atoms flag." % (len(s.cgps), s.ncgp_solute, 1))
    for g in s.cgps:
        L.append("%6d%6d" % (len(g), g[0]))
        L += _fmt_list(list(g), 16, "%6d")
    nt = len(s.masses)
    L.append("%8d = No. of atom types" % nt)
    L.append("%8d = vdW combination rule (1 = Geom. / 2 = Arit.)" % 1)
    L.append("%9.5f%10.4f = Electrostatic 1-4 scaling factor and  Coulomb constant" % (s.el14_scale, s.coulomb_constant))
    L.append("Masses:")
    L += _fmt_list(list(s.masses), 10, "%10.4f")
    for j in range(3):
        L.append("sqrt (Aii) normal:" if j == 0 else ("sqrt (Aii) polar:" if j == 1 else "sqrt (Aii) 1-4:"))
        L += _fmt_list(list(s.avdw[:, j]), 8, "%13.5f")
        L.append("sqrt (Bii) normal:" if j == 0 else ("sqrt (Bii) polar:" if j == 1 else "sqrt (Bii) 1-4:"))
        L += _fmt_list(list(s.bvdw[:, j]), 8, "%13.5f")
    L.append("%8d = No. of type-2 vdW interactions. pairs:" % 0)
    ns = s.nat_solute

    def bits(arr):
        flat = "".join("1" if b else "0" for b in arr.reshape(-1))
        return [flat[k:k + 80] for k in range(0, len(flat), 80)]
    L.append("%8d = No. of 1-4 neighbours. nborlist (range=max_nbr_range):" % int(s.list14.sum()))
    if ns:
        L += bits(s.list14)
    L.append("%8d = No. of long 1-4 nbrs (>max_nbr_range). pairlist:" % len(s.list14long))
    for i, j in s.list14long:
        L.append("%6d%6d" % (i, j))
    L.append("%8d = No. of exclusions. exclusion list (range=max_nbr_range):" % int(s.listex.sum()))
    if ns:
        L += bits(s.listex)
    L.append("%8d = No. of long exclusions (>max_nbr_range). pairlist:" % len(s.listexlong))
    for i, j in s.listexlong:
        L.append("%6d%6d" % (i, j))
    L.append("%8d%8d = No. of residues, No of solute residues. start atoms:" % (len(s.res_start), s.nres_solute))
    L += _fmt_list(list(s.res_start), 13, "%6d")
    L.append("Sequence:")
    L += _fmt_list(list(s.res_names), 16, "%-4s ")
    L.append("%8d = No. of separate molecules. start atoms:" % len(s.mol_start))
    L += _fmt_list(list(s.mol_start), 13, "%6d")
    L.append("%8d = No. of atom types:" % nt)
    L += _fmt_list(list(s.tac), 8, "%-8s ")
    L.append("%8d = No. of SYBYL atom types:" % nt)
    L += _fmt_list(["" for _ in range(nt)], 13, "%-5s ")
    L.append("%8d = solvent type (0=SPC,1=3-atom,2=general)" % s.solvent_type)
    if s.box is not None:
        # periodic box (src/topo.f90:1350-1360): kind, edge lengths, centre
        L.append("%8s = kind of boundary" % "PBC")
        L.append("%8.3f%8.3f%8.3f = Size of box, x y z" % tuple(s.box))
        L.append("%8.3f%8.3f%8.3f = Centre coordinate of box" % tuple(s.box_centre))
        with open(path, "w") as f:
            f.write("\n".join(L) + "\n")
        return
    L.append("%8.3f%8.3f = Exclusion, solvent radii" % (s.rexcl_o, s.rwat))
    L.append("%8.3f%8.3f%8.3f = Solute centre" % tuple(s.xpcent))
    L.append("%8.3f%8.3f%8.3f = Solvent centre" % tuple(s.xwcent))
    ex = s.excl if s.excl is not None else np.zeros(N, dtype=bool)
    nexwat = sum(1 for w in range(s.nwat) if ex[s.nat_solute + 3 * w])
    L.append("%8d%8d = No. of excluded atoms (incl. water), no. of excluded waters" % (int(ex.sum()), nexwat))
    flags = "".join("T" if e else "F" for e in ex)
    L += [flags[k:k + 80] for k in range(0, N, 80)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_fep(s: System, path: str) -> None:
    """FEP file with [FEP] states, [atoms] (Q number, topology atom),
    [change_charges] (Q number, charge per state) and, when present,
    [atom_types] / [change_atoms] (Q-atom LJ types per state), [softcore]
    (alpha per state) and [off_diagonals] (states i j, Q atoms k l, A, mu),
    in the layout qatom_load_fep reads (src/qatom.f90:700-1975)."""
    nq, ns = s.qcrg.shape
    L = ["[FEP]", "states %d" % ns]
    if s.softcore is not None:
        L.append("softcore_use_max_potential off")
    if s.qq_lib_charges:
        L.append("qq_use_library_charges on")
    L += ["", "[atoms]"]
    for k, a in enumerate(s.qatoms):
        L.append("%d %d" % (k + 1, a))
    L += ["", "[change_charges]"]
    for k in range(nq):
        L.append("%d " % (k + 1) + " ".join("%.4f" % c for c in s.qcrg[k]))
    if s.qtypes is not None:
        L += ["", "[atom_types]"]
        for name, a1, b1, a2, b2, a3, b3, m in s.qtypes:
            L.append("%-8s %12.4f %10.4f %12.4f %10.4f %12.4f %10.4f %8.3f" % (name, a1, b1, a2, b2, a3, b3, m))
        L += ["", "[change_atoms]"]
        for k in range(nq):
            L.append("%d " % (k + 1) + " ".join(s.qtypes[t][0] for t in s.qtype_of[k]))
    if s.softcore is not None:
        L += ["", "[softcore]"]
        for k in range(nq):
            L.append("%d " % (k + 1) + " ".join("%.2f" % a for a in s.softcore[k]))
    if s.offdiag:
        L += ["", "[off_diagonals]"]
        for si, sj, qk, ql, A, mu in s.offdiag:
            L.append("%d %d %d %d %.2f %.3f" % (si, sj, qk, ql, A, mu))
    if s.soft_pairs:
        L += ["", "[soft_pairs]"] + ["%d %d" % p for p in s.soft_pairs]
    if s.el_scale:
        L += ["", "[el_scale]"] + ["%d %d " % (m, n) + " ".join("%.2f" % v for v in sc) for m, n, sc in s.el_scale]
    if s.excluded_pairs:
        L += ["", "[excluded_pairs]"] + ["%d %d " % (i, j) + " ".join("%d" % f for f in fl)
                                         for i, j, fl in s.excluded_pairs]
    if s.monitor_groups:
        L += ["", "[monitor_groups]"] + [" ".join("%d" % a for a in g) for g in s.monitor_groups]
        L += ["", "[monitor_group_pairs]"] + ["%d %d" % p for p in s.monitor_pairs]
    if s.shake_constraints:
        L += ["", "[shake_constraints]"] + ["%d %d " % (i, j) + " ".join("%.3f" % v for v in d)
                                            for i, j, d in s.shake_constraints]
    if s.qbonds is not None:
        sec = [("bond_types", s.qbond_types, "change_bonds", s.qbonds),
               ("angle_types", s.qangle_types, "change_angles", s.qangles),
               ("torsion_types", s.qtorsion_types, "change_torsions", s.qtorsions),
               ("improper_types", s.qimproper_types, "change_impropers", s.qimpropers)]
        for tname, types, cname, terms in sec:
            if not terms:
                continue
            L += ["", "[%s]" % tname]
            for k, row in enumerate(types):
                L.append("%d " % (k + 1) + " ".join("%.6g" % v for v in row))
            L += ["", "[%s]" % cname]
            for t in terms:
                L.append(" ".join("%d" % a for a in t[:-1]) + "  " + " ".join("%d" % c for c in t[-1]))
        for name in ("angle", "torsion", "improper"):
            rows = (s.qcouplings or {}).get(name) or []
            if rows:
                L += ["", "[%s_couplings]" % name]
                for term, bond, flag in rows:
                    L.append("%d %d %d" % (term, bond, flag))
    if s.box is not None:
        # the Q atoms' switching atom for the periodic lists (src/qatom.f90:672-690)
        L += ["", "[PBC]", "switching_atom %d" % int(s.qatoms[0])]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_inp(path: str, top: str, final: str, *, steps: int, stepsize: float = 1.0,
              temperature: float = 300.0, bath_coupling: float = 100.0,
              random_seed: Optional[int] = 1, initial_temperature: Optional[float] = 300.0,
              restart: Optional[str] = None, fep: Optional[str] = None,
              lambdas: Optional[Sequence[float]] = None, energy: Optional[str] = None,
              energy_interval: int = 0, output: int = 10, non_bond: int = 10,
              rcpp: float = 10.0, rcww: float = 10.0, rcpw: float = 10.0,
              rcq: float = 99.0, rclrf: float = 99.0, lrf: bool = True,
              shake_solvent: bool = True, shake_solute: bool = False,
              shake_hydrogens: bool = False, separate_scaling: bool = True,
              radial_force: Optional[float] = 60.0, polarisation: bool = True,
              polarisation_force: Optional[float] = 20.0,
              shell_force: Optional[float] = None, shell_radius: Optional[float] = None,
              excluded_freeze: bool = False, excluded_force: Optional[float] = None,
              exclude_bonded: bool = False,
              temp_interval: Optional[int] = None, restraints: Optional[dict] = None,
              trajectory: Optional[str] = None, trajectory_interval: int = 0,
              trajectory_atoms: Optional[Sequence[str]] = None,
              pbc: Optional[dict] = None, volume_change: Optional[int] = None,
              thermostat: Optional[str] = None, integrator: Optional[str] = None,
              md_extra: Optional[dict] = None) -> None:
    """qdyn input file; `restraints` maps sequence/atom/distance/angle/wall
    to lists of rows written into the [<name>_restraints] sections;
    `trajectory_atoms` rows form the [trajectory_atoms] mask; `pbc` (a dict
    of [PBC] keys, e.g. put_solvent_back_in_box) writes a periodic-box input:
    a [PBC] section in place of [sphere] and [solvent]; `volume_change` is
    the interval of constant-pressure volume moves; `thermostat`,
    `integrator` and `md_extra` (more [MD] keys, e.g. nhchains) choose the
    dynamics."""
    on = lambda b: "on" if b else "off"
    L = ["[MD]", "steps %d" % steps, "stepsize %.6g" % stepsize,
         "temperature %.6g" % temperature, "bath_coupling %.6g" % bath_coupling,
         "separate_scaling %s" % on(separate_scaling)]
    if thermostat:
        L.append("thermostat %s" % thermostat)
    if integrator:
        L.append("integrator %s" % integrator)
    for k, v in (md_extra or {}).items():
        L.append("%s %s" % (k, on(v) if isinstance(v, bool) else v))
    if initial_temperature is not None:
        L += ["random_seed %d" % random_seed, "initial_temperature %.6g" % initial_temperature]
    L += ["shake_solvent %s" % on(shake_solvent), "shake_solute %s" % on(shake_solute),
          "shake_hydrogens %s" % on(shake_hydrogens), "lrf %s" % on(lrf), "",
          "[cut-offs]", "solute_solute %.6g" % rcpp, "solvent_solvent %.6g" % rcww,
          "solute_solvent %.6g" % rcpw, "q_atom %.6g" % rcq]
    if lrf:
        L.append("lrf %.6g" % rclrf)
    if pbc is not None:
        L += ["", "[PBC]"] + ["%s %s" % (k, on(v) if isinstance(v, bool) else v) for k, v in pbc.items()]
    else:
        L += ["", "[sphere]"]
        if shell_radius is not None:
            L.append("shell_radius %.6g" % shell_radius)
        if shell_force is not None:
            L.append("shell_force %.6g" % shell_force)
        if excluded_force is not None:
            L.append("excluded_force %.6g" % excluded_force)
        if excluded_freeze:
            L.append("excluded_freeze on")
        if exclude_bonded:
            L.append("exclude_bonded on")
        L += ["", "[solvent]"]
        if radial_force is not None:
            L.append("radial_force %.6g" % radial_force)
        L.append("polarisation %s" % on(polarisation))
        if polarisation_force is not None:
            L.append("polarisation_force %.6g" % polarisation_force)
    L += ["", "[intervals]", "non_bond %d" % non_bond, "output %d" % output]
    if temp_interval is not None:
        L.append("temperature %d" % temp_interval)
    if volume_change is not None:
        L.append("volume_change %d" % volume_change)
    if energy:
        L.append("energy %d" % energy_interval)
    if trajectory:
        L.append("trajectory %d" % trajectory_interval)
    L += ["", "[files]", "topology %s" % top, "final %s" % final]
    if trajectory:
        L.append("trajectory %s" % trajectory)
    if restart:
        L.append("restart %s" % restart)
    if energy:
        L.append("energy %s" % energy)
    if fep:
        L.append("fep %s" % fep)
    if lambdas is not None:
        L += ["", "[lambdas]", " ".join("%.6f" % l for l in lambdas)]
    if trajectory_atoms:
        L += ["", "[trajectory_atoms]"] + list(trajectory_atoms)
    for name in ("sequence", "atom", "distance", "angle", "wall"):
        rows = (restraints or {}).get(name)
        if rows:
            L += ["", "[%s_restraints]" % name]
            L += [" ".join(str(v) for v in row) for row in rows]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


# ---------------------------------------------------------------------------
# restart files (Fortran unformatted sequential: 4-byte record markers)


def _rec(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_restart(path: str, x: np.ndarray, v: np.ndarray,
                  theta_corr: Optional[np.ndarray] = None, box: Optional[np.ndarray] = None) -> None:
    """qdyn restart (write_xfin, src/md.f90:18705-18736): x, v, then the
    polarisation shells' theta_corr (sphere) or the box lengths and centre
    (periodic box, box = 6 values)."""
    x = np.ascontiguousarray(x, dtype="<f8").reshape(-1)
    v = np.ascontiguousarray(v, dtype="<f8").reshape(-1)
    nat3 = x.size
    out = _rec(struct.pack("<i", -1337))
    out += _rec(struct.pack("<i", nat3) + x.tobytes())
    out += _rec(struct.pack("<i", nat3) + v.tobytes())
    if theta_corr is not None:
        tc = np.ascontiguousarray(theta_corr, dtype="<f8")
        out += _rec(struct.pack("<i", tc.size) + tc.tobytes())
    if box is not None:
        b = np.ascontiguousarray(box, dtype="<f8").reshape(-1)
        out += _rec(b[:3].tobytes()) + _rec(b[3:6].tobytes())
    with open(path, "wb") as f:
        f.write(out)


def read_restart(path: str):
    """Returns (x (N,3), v (N,3), theta_corr or None); a periodic box's
    restart (two 24-byte box records after v) gives theta_corr None (see
    read_restart_box)."""
    with open(path, "rb") as f:
        data = f.read()
    recs = []
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        recs.append(data[p + 4:p + 4 + n])
        p += 8 + n
    if struct.unpack("<i", recs[0])[0] != -1337:
        raise ValueError("restart file without a double-precision canary")
    nat3 = struct.unpack_from("<i", recs[1])[0]
    x = np.frombuffer(recs[1], dtype="<f8", offset=4, count=nat3).copy()
    v = np.frombuffer(recs[2], dtype="<f8", offset=4, count=nat3).copy()
    tc = None
    if len(recs) > 3 and not (len(recs) == 5 and len(recs[3]) == 24 and len(recs[4]) == 24):
        nsh = struct.unpack_from("<i", recs[3])[0]
        tc = np.frombuffer(recs[3], dtype="<f8", offset=4, count=nsh).copy()
    return x.reshape(-1, 3), v.reshape(-1, 3), tc


def read_restart_box(path: str) -> Optional[np.ndarray]:
    """The box lengths and centre (6 values) of a periodic box's restart,
    or None."""
    with open(path, "rb") as f:
        data = f.read()
    recs = []
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        recs.append(data[p + 4:p + 4 + n])
        p += 8 + n
    if len(recs) == 5 and len(recs[3]) == 24 and len(recs[4]) == 24:
        return np.concatenate([np.frombuffer(recs[3], "<f8"), np.frombuffer(recs[4], "<f8")])
    return None


def mapping_lambdas(n_states: int, windows: int) -> np.ndarray:
    """EVB mapping windows that move the system from state 1 through 2, ...
    to state n_states: window k sits at t = k / (windows - 1) of the path,
    on the segment between states s and s + 1 with weight u, so its lambda
    vector is (1 - u) e_s + u e_{s+1}.  Returns (windows, n_states)."""
    lam = np.zeros((windows, n_states))
    seg = n_states - 1
    for k in range(windows):
        t = k / max(1, windows - 1) * seg
        sI = min(seg - 1, int(math.floor(t)))
        u = t - sI
        lam[k, sI] = 1.0 - u
        lam[k, sI + 1] = u
    return np.round(lam, 6)


def evb_system(radius: float, n_solute: int, seed: int, n_q: int = 24, n_states: int = 4) -> System:
    """The EVB-style workload: a solvated branched solute (as
    solvated_system) with an n_q-atom Q region of n_states states carrying
    per-state charges and Q-atom LJ types, and n_states (n_states - 1) / 2
    off-diagonal couplings A exp(-mu r) (A ~ U[10,60] kcal/mol,
    mu ~ U[0,0.5] 1/A) between random state pairs and Q-atom pairs."""
    return solvated_system(radius=radius, n_solute=n_solute, seed=seed, n_q=n_q, n_states=n_states,
                           q_types=True, n_offdiag=n_states * (n_states - 1) // 2)


def write_system(s: System, directory: str, name: str, **inp) -> dict:
    """Write name.top (+ name.fep) and name.inp into directory; returns paths."""
    os.makedirs(directory, exist_ok=True)
    top = os.path.join(directory, name + ".top")
    write_top(s, top)
    paths = {"top": top}
    if s.qatoms is not None:
        fep = os.path.join(directory, name + ".fep")
        write_fep(s, fep)
        paths["fep"] = fep
        inp.setdefault("fep", name + ".fep")
    inpp = os.path.join(directory, name + ".inp")
    inp.setdefault("steps", 100)
    if s.box is not None:
        inp.setdefault("pbc", {"put_solvent_back_in_box": True, "put_solute_back_in_box": False})
    if s.restraints:
        inp.setdefault("restraints", s.restraints)
    write_inp(inpp, name + ".top", name + ".re", **inp)
    paths["inp"] = inpp
    paths["re"] = os.path.join(directory, name + ".re")
    return paths
