"""Geometries used by the reference's tests (tests/test001-010.py) and
helpers to restate them at another Angstrom->bohr factor."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mcmurchie-davidson_amd"))

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


from mmd.workloads import water_cluster  # noqa: E402,F401
