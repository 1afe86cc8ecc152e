"""GPU nuclear gradient (mmd_rhf_gradient, Molecule.forces) against the
reference's H2 forces (tests/test001.py:15-18) and the oracle's gradient
(tests/golden/gradient_water.npz, made by make_gradient_fixtures.py from the
restatement of mmd/forces.py and cython/grad.pyx, itself pinned by
tests/test_oracle.py against finite differences).  Tolerances: the
reference test's own for H2 (assert_allclose atol 1e-12); 1e-10 against
the oracle on identical P and F; invariances 1e-10; central differences of
the GPU energy (h = 1e-4 bohr) 1e-7, with the SCFs of the differences
unscreened (acc2e = 0: the screen's decisions change with the geometry and
their 1e-12-sized steps, divided by 2h, would dominate the quotient)."""
import os

import numpy as np
import pytest
from numpy.testing import assert_allclose

from geometries import H2, water_cluster

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


def test_h2_forces_reference(native):
    """tests/test001.py: energy and forces of H2 / STO-3G."""
    from mmd.molecule import Molecule
    mol = Molecule(geometry=H2, basis="sto-3g")
    mol.RHF(doPrint=False)
    assert_allclose(mol.energy.real, -1.11675930740, atol=1e-12)
    mol.forces()
    ref = np.array([[0.000000000, -0.000000000, -0.027679601], [-0.000000000, 0.000000000, 0.027679601]])
    print("H2 forces", mol._forces[:, 2], "device ms %.3f" % mol.gradient_ms)
    assert_allclose(mol._forces, ref, atol=1e-12)


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g", "6-31gs"])
def test_gradient_vs_oracle(native, basis):
    """The same P and F as the oracle's: electronic gradient + nuclear
    repulsion within 1e-10 of the oracle's dE/dX (6-31G* has Cartesian d)."""
    from mmd.forces import basis_atoms, nuclear_repulsion_gradient
    from mmd.molecule import Molecule
    fx = np.load(os.path.join(GOLDEN, "gradient_water.npz"))
    key = basis.replace("-", "")
    mol = Molecule(str(fx["geometry"]), basis)
    P, F = fx["P_" + key], fx["F_" + key]
    W = P @ F @ P
    g = mol.native.rhf_gradient(P, W, basis_atoms(mol)) + nuclear_repulsion_gradient(mol.atoms)
    err = np.max(np.abs(g - fx["grad_" + key]))
    print("%s: max |dg| %.2e, max |g| %.2e" % (basis, err, np.abs(g).max()))
    assert err < 1e-10


def test_gradient_cluster_invariance_and_differences(native):
    """(H2O)2 / 6-31G: forces sum to zero and exert no torque; two
    components against central differences of the converged GPU energy."""
    from mmd.molecule import Molecule
    geom = water_cluster(2, 5.0, seed=3, perturb=True)
    mol = Molecule(geom, "6-31g")
    mol.RHF(doPrint=False, direct=True, conver=1e-10, acc2e=0.0)
    f = mol.forces()
    xyz = np.array([a.origin for a in mol.atoms])
    print("(H2O)2 6-31G: sum %s torque %s, device ms %.2f" % (f.sum(0), np.cross(xyz, f).sum(0), mol.gradient_ms))
    assert np.max(np.abs(f.sum(0))) < 1e-10
    assert np.max(np.abs(np.cross(xyz, f).sum(0))) < 1e-10
    h = 1e-4
    rows = geom.strip().split("\n")
    for a, x in ((0, 0), (4, 2)):
        e = []
        for sgn in (1, -1):
            lines = list(rows)
            parts = lines[1 + a].split()
            vals = [float(v) for v in parts[1:]]
            vals[x] += sgn * h * 0.52917721092
            lines[1 + a] = parts[0] + " " + " ".join("%.12f" % v for v in vals)
            m = Molecule("\n".join(lines) + "\n", "6-31g")
            m.RHF(doPrint=False, direct=True, conver=1e-10, acc2e=0.0)
            e.append(m.energy.real)
        fd = (e[0] - e[1]) / (2 * h)
        print("atom %d axis %d: analytic %.10f, differences %.10f" % (a, x, -f[a, x], fd))
        assert abs(-f[a, x] - fd) < 1e-7
