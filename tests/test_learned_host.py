"""Host-side parts of the learned initial density (CPU only): the numpy
restatement of the perceptron against finite differences, and the
geometry handling of mmd.learned (local frames, rotation invariance of
features and of local-frame density blocks, block assembly)."""
import numpy as np
import pytest

import mlp_ref
from mmd.workloads import water_cluster


def test_numpy_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 4]
    nparam = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    flat = rng.standard_normal(nparam) * 0.5
    X = rng.standard_normal((9, 5))
    Y = rng.standard_normal((9, 4))
    _, g = mlp_ref.loss_grad(flat, dims, X, Y)
    h = 1e-6
    for k in rng.choice(nparam, 25, replace=False):
        e = np.zeros(nparam)
        e[k] = h
        fd = (mlp_ref.loss_grad(flat + e, dims, X, Y)[0] - mlp_ref.loss_grad(flat - e, dims, X, Y)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd))


def _rotated(geom, Q):
    lines = geom.strip().split("\n")
    out = [lines[0]]
    for ln in lines[1:]:
        s, x, y, z = ln.split()
        r = Q @ np.array([float(x), float(y), float(z)])
        out.append("%s %.12f %.12f %.12f" % (s, r[0], r[1], r[2]))
    return "\n".join(out) + "\n"


def _global_rotation(mol, Q):
    """Matrix that rotates every Cartesian p triple of the basis by Q."""
    N = mol.nbasis
    U = np.eye(N)
    k = 0
    while k < N:
        if sum(mol.bfs[k].shell) == 1:
            U[k:k + 3, k:k + 3] = Q
            k += 3
        else:
            k += 1
    return U


@pytest.fixture
def rotation():
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def test_features_and_local_blocks_rotation_invariant(rotation):
    from mmd import learned
    from mmd.molecule import Molecule
    geom = water_cluster(3, 6.0, 21, perturb=True)
    a = Molecule(geom, "6-31g")
    b = Molecule(_rotated(geom, rotation), "6-31g")
    assert np.allclose(learned.features(a), learned.features(b), atol=1e-9)
    rng = np.random.default_rng(2)
    P = rng.standard_normal((a.nbasis, a.nbasis))
    P = P + P.T
    U = _global_rotation(a, rotation)
    Pb = U @ P @ U.T   # the same density expressed for the rotated geometry
    assert np.allclose(learned.local_blocks(a, P), learned.local_blocks(b, Pb), atol=1e-9)


def test_assemble_inverts_local_blocks():
    from mmd import learned
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(2, 5.0, 4, perturb=True), "6-31g")
    rng = np.random.default_rng(1)
    P = rng.standard_normal((mol.nbasis, mol.nbasis))
    P = P + P.T
    R = learned.assemble(mol, np.zeros_like(P), learned.local_blocks(mol, P))
    for _, bf in learned.monomers(mol):
        assert np.allclose(R[np.ix_(bf, bf)], P[np.ix_(bf, bf)])
    assert learned.features(mol).shape == (2, 15)
    assert learned.local_blocks(mol, P).shape == (2, 91)


def test_learned_guess_rejects_non_water():
    from mmd import learned
    from mmd.molecule import Molecule
    from geometries import METHANE
    with pytest.raises(ValueError):
        learned.monomers(Molecule(METHANE, "6-31g"))
