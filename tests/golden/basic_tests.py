"""The reference's basic test system (tests/basic_tests): a 27-atom ligand,
every atom a Q atom, in a 16 A water sphere (prep_SPH, 3397 atoms), and the
first equilibration input eq1.inp (100 steps of 0.1 fs from 0.1 K, LRF on,
a sequence restraint over atoms 1-46).

eq1.inp carries placeholders that the reference's run_test.sh replaces at
run time (modify_inputs, tests/basic_tests/run_test.sh:176-189);
`sph_input` makes the same replacements for the spherical boundary.  The
data files (prep_SPH/lig_w.top, prep_SPH/lig_w.fep, eq1.inp) are kept
beside this module as copied by make_golden.py; the reference's own tree
is read in place where it is present.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "prep_SPH")          # eq1.inp, lig_w.top, lig_w.fep
FIXTURE_PBC = os.path.join(HERE, "prep_PBC")      # lig_w.top, lig_w.fep of the periodic box
REFERENCE = "/root/reference/tests/basic_tests"   # eq1.inp; prep_SPH/lig_w.{top,fep}
FILES = ("eq1.inp", "lig_w.top", "lig_w.fep")

# eval_test.sh:96-101: the Q-surr. row of the step-0 Q-atom energies of
# eq1 that every SPH (PBC) run must print
INIT_QSURR_SPH = "Q-surr. 1 1.0000      3.12    139.43"
INIT_QSURR_PBC = "Q-surr. 1 1.0000    -31.30    228.64"


def reference_present() -> bool:
    return all(os.path.exists(p) for p in (os.path.join(REFERENCE, "eq1.inp"),
                                           os.path.join(REFERENCE, "prep_SPH", "lig_w.top"),
                                           os.path.join(REFERENCE, "prep_SPH", "lig_w.fep")))


def sph_input(inp: str, data_dir: str, dest_dir: str, final: str = "eq1.re",
              integrator: str = "leap-frog", thermostat: str = "berendsen", name: str = None) -> str:
    """Writes `inp` into dest_dir as run_test.sh prepares it for the sphere:
    the boundary becomes [sphere] with shell_radius 16 and shell_force 20,
    the Q-atom and LRF cut-offs 99, and the integrator and thermostat the
    given ones; the topology and FEP file are taken from data_dir by
    absolute path and the final restart goes to dest_dir/final.  Returns
    the path of the written input."""
    t = open(inp).read()
    subst = [("invalid-boundary", "[sphere]\nshell_radius\t16\nshell_force\t20"),
             ("invalid-qatom", "q_atom\t99"), ("invalid-lrf", "lrf\t99"),
             ("invalid-integrator", integrator), ("invalid-thermostat", thermostat)]
    for a, b in subst:
        t = t.replace(a, b)
    out = []
    for line in t.splitlines():
        w = line.split()
        if len(w) == 2 and w[0] in ("topology", "fep"):
            line = "%-24s %s" % (w[0], os.path.join(os.path.abspath(data_dir), w[1]))
        elif len(w) == 2 and w[0] == "final":
            line = "%-24s %s" % (w[0], os.path.join(os.path.abspath(dest_dir), final))
        out.append(line)
    path = os.path.join(dest_dir, name or os.path.basename(inp))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path


def pbc_input(inp: str, data_dir: str, dest_dir: str, final: str = "eq1.re",
              integrator: str = "leap-frog", thermostat: str = "berendsen") -> str:
    """As sph_input for the periodic box (run_test.sh:179-182): the boundary
    becomes [PBC] with the solvent (not the solute) put back in the box at
    list updates, the Q-atom and LRF cut-offs 24."""
    t = open(inp).read()
    t = t.replace("invalid-boundary", "[PBC]\nput_solvent_back_in_box\ton\nput_solute_back_in_box\toff")
    t = t.replace("invalid-qatom", "q_atom\t24").replace("invalid-lrf", "lrf\t24")
    tmp = os.path.join(dest_dir, "_pbc_" + os.path.basename(inp))
    with open(tmp, "w") as f:
        f.write(t)
    path = sph_input(tmp, data_dir, dest_dir, final=final, integrator=integrator, thermostat=thermostat,
                     name=os.path.basename(inp))
    os.remove(tmp)
    return path


FIXTURE_NPT = os.path.join(HERE, "test5_NPT")      # eq1.inp, lig_w.fep of the constant-pressure test
REFERENCE_NPT = "/root/reference/tests/test5"
# the constant-pressure test shortened: 100 steps, lists and volume moves
# every 10 steps (nine moves)
NPT_SHORT = dict(steps=100, non_bond=10, volume_change=10)


def npt_input(dest_dir: str, final: str = "eq1.re", **over) -> str:
    """The reference's constant-pressure test input (tests/test5/eq1.inp:
    the prep_PBC ligand in its 50 A water box, 0.1 fs from 0.1 K, LRF on,
    volume moves of up to 5 A^3 at 1.01 bar from pressure seed 1337) written
    into dest_dir with the topology (identical to prep_PBC's) and its own
    FEP file by absolute path.  `over` replaces or adds keys, e.g.
    steps=100, non_bond=10, volume_change=10."""
    sect = {"steps": "MD", "non_bond": "intervals", "output": "intervals", "volume_change": "intervals",
            "max_volume_displ": "PBC", "temperature": "MD", "stepsize": "MD"}
    lines = open(os.path.join(FIXTURE_NPT, "eq1.inp")).read().splitlines()
    out, cur, done = [], None, set()
    files = {"topology": os.path.join(FIXTURE_PBC, "lig_w.top"), "fep": os.path.join(FIXTURE_NPT, "lig_w.fep"),
             "final": os.path.join(os.path.abspath(dest_dir), final)}

    def flush(section):
        blank = 0
        while out and not out[-1].strip():
            out.pop()
            blank += 1
        for k, v in over.items():
            if sect.get(k) == section and k not in done:
                out.append("%-24s %s" % (k, v))
                done.add(k)
        out.extend([""] * blank)

    for line in lines:
        w = line.split()
        if line.strip().startswith("["):
            flush(cur)
            cur = line.strip()[1:-1]
        elif len(w) >= 2 and w[0] in files:
            line = "%-24s %s" % (w[0], files[w[0]])
        elif len(w) >= 2 and w[0] in over:
            line = "%-24s %s" % (w[0], over[w[0]])
            done.add(w[0])
        out.append(line)
    flush(cur)
    path = os.path.join(dest_dir, "npt_eq1.inp")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path


def benchmark_row(path: str, row: int = 1):
    """Row `row` of a reference benchmark energy file (the lower and upper
    bound of QEL QVdW EL VdW QBND QANG QTOR QIMP BND ANG TOR IMP TOT KIN
    POT): returns {column: (lower, upper)}."""
    names = None
    for line in open(path):
        w = line.split()
        if line.startswith("#") and "QEL" in line:
            names = w[1:]
            continue
        if w and not line.startswith("#") and int(w[0]) == row:
            v = [float(t) for t in w[1:]]
            return {n: (v[2 * i], v[2 * i + 1]) for i, n in enumerate(names)}
    raise KeyError(row)
