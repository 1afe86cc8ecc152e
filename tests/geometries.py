"""Geometries used by the reference's tests (tests/test001-010.py) and
helpers to restate them at another Angstrom->bohr factor."""
import numpy as np

H2 = """
0 1
H 0.0 0.0 0.74
H 0.0 0.0 0.00
"""

WATER = """
0 1
O    0.000000      -0.075791844    0.000000
H    0.866811829    0.601435779    0.000000
H   -0.866811829    0.601435779    0.000000
"""

METHANE = """
0 1
C    0.000000    0.000000    0.000000
H    0.626425042   -0.626425042   -0.626425042
H    0.626425042    0.626425042    0.626425042
H   -0.626425042    0.626425042   -0.626425042
H   -0.626425042   -0.626425042    0.626425042
"""

FORMALDEHYDE = """
0 1
C          0.0000000000        0.0000000000       -0.5265526741
O          0.0000000000        0.0000000000        0.6555124750
H          0.0000000000       -0.9325664988       -1.1133424527
H          0.0000000000        0.9325664988       -1.1133424527
"""

HELIUM = """
0 1
He    0.000000    0.000000    0.000000
"""

# The reference's water/methane/formaldehyde energies were produced at the
# geometry obtained with a Bohr radius of 0.529177249 A (CODATA 1986), while
# mmd/molecule.py:224 divides by 0.52917721092.  Re-expressing the Angstrom
# input by this ratio feeds the reference parser the geometry those numbers
# belong to.
BOHR_CODE = 0.52917721092
BOHR_1986 = 0.529177249

# Water of the reference's README in bohr (the geometry its STO-3G energy
# -74.942079928192 and dipole 1.53400931 D belong to).
WATER_BOHR = [("O", 0.0, -0.143225816552, 0.0), ("H", 1.638036840407, 1.136548822547, 0.0),
              ("H", -1.638036840407, 1.136548822547, 0.0)]


def parse(geom):
    rows = [ln.split() for ln in geom.split("\n") if ln.strip()]
    return rows[0], [(r[0], float(r[1]), float(r[2]), float(r[3])) for r in rows[1:]]


def rescale(geom, bohr):
    """Same geometry string, Angstrom values scaled so that dividing by the
    reference's factor yields coordinates/bohr as if divided by `bohr`."""
    head, atoms = parse(geom)
    s = BOHR_CODE / bohr
    return "\n" + " ".join(head) + "\n" + "\n".join(
        "%s %.17g %.17g %.17g" % (a, x * s, y * s, z * s) for a, x, y, z in atoms) + "\n"


def from_bohr(atoms_bohr, charge=0, mult=1):
    return "\n%d %d\n" % (charge, mult) + "\n".join(
        "%s %.17g %.17g %.17g" % (a, x * BOHR_CODE, y * BOHR_CODE, z * BOHR_CODE)
        for a, x, y, z in atoms_bohr) + "\n"


def water_cluster(n, box, seed, min_oo=2.7, perturb=False):
    """(H2O)_n: monomers at the README geometry, each rotated uniformly
    (Shoemake quaternion) about its O, O placed uniformly in [0, box]^3 A
    with all O-O distances >= min_oo (rejection sampling).  With `perturb`,
    O-H lengths get N(0, 0.02 A) and the H-O-H angle N(0, 2 deg) first."""
    rng = np.random.default_rng(seed)
    O = np.array([0.0, -0.075791844, 0.0])
    H1 = np.array([0.866811829, 0.601435779, 0.0])
    H2v = np.array([-0.866811829, 0.601435779, 0.0])
    mono0 = np.array([O, H1, H2v]) - O
    placed = []
    out = []
    while len(placed) < n:
        mono = mono0.copy()
        if perturb:
            r1 = np.linalg.norm(mono[1]) + rng.normal(0.0, 0.02)
            r2 = np.linalg.norm(mono[2]) + rng.normal(0.0, 0.02)
            ang0 = np.arccos(np.dot(mono[1], mono[2]) / np.linalg.norm(mono[1]) / np.linalg.norm(mono[2]))
            ang = ang0 + np.deg2rad(rng.normal(0.0, 2.0))
            bis = (mono[1] / np.linalg.norm(mono[1]) + mono[2] / np.linalg.norm(mono[2]))
            bis /= np.linalg.norm(bis)
            perp = np.cross(bis, [0.0, 0.0, 1.0])
            perp /= np.linalg.norm(perp)
            mono[1] = r1 * (np.cos(ang / 2) * bis + np.sin(ang / 2) * perp)
            mono[2] = r2 * (np.cos(ang / 2) * bis - np.sin(ang / 2) * perp)
        u1, u2, u3 = rng.random(3)
        q = np.array([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                      np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)])
        x, y, z, w = q
        Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                       [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                       [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        pos = rng.random(3) * box
        if any(np.linalg.norm(pos - p) < min_oo for p in placed):
            continue
        placed.append(pos)
        for sym, r in zip("OHH", mono):
            out.append((sym,) + tuple(pos + Rm @ r))
    return "\n0 1\n" + "\n".join("%s %.12f %.12f %.12f" % a for a in out) + "\n"
