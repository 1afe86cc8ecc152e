"""Host-side logic and the C-ABI boundary (CPU only, no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from geometries import METHANE, WATER, water_cluster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BASIS = os.path.join(ROOT, "mcmurchie-davidson_amd", "mmd", "basis")


def _gbs(name):
    return os.path.join(BASIS, name + ".gbs")


@pytest.mark.parametrize("geom,basis", [(WATER, "sto-3g"), (WATER, "cc-pvdz"), (METHANE, "6-31g"),
                                        (water_cluster(3, 7.0, 2), "6-31g"), (WATER, "6-31ppgss")])
def test_basis_order_and_normalisation_match_oracle(oracle, geom, basis):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    om = oracle.OracleMolecule(geom, _gbs(basis))
    assert mol.nbasis == om.nbasis
    assert mol.nelec == om.nelec and mol.nocc == om.nocc
    for i, b in enumerate(mol.bfs):
        assert np.array_equal(b.shell, om.shell[i])
        assert np.array_equal(b.origin, om.origin[i])
        o, k = om.poff[i], om.np[i]
        assert b.num_exps == k
        assert np.array_equal(b.exps, om.exps[o:o + k])
        assert np.allclose(b.coefs, om.coefs[o:o + k], rtol=1e-14, atol=0)
        assert np.allclose(b.norm, om.norm[o:o + k], rtol=1e-14, atol=0)
    assert np.allclose(mol.center_of_charge, om.center_of_charge, rtol=1e-15, atol=1e-15)


def test_cartesian_counts():
    """Cartesian d: water cc-pVDZ has 25 functions, (H2O)8 200."""
    from mmd.molecule import Molecule
    assert Molecule(WATER, "cc-pvdz").nbasis == 25
    assert Molecule(water_cluster(8, 7.0, 0), "cc-pvdz").nbasis == 200
    assert Molecule(water_cluster(2, 7.0, 0), "6-31g").nbasis == 26


def test_input_errors():
    from mmd.molecule import Molecule
    with pytest.raises(ValueError):
        Molecule("\n0 1\nXx 0 0 0\n", "sto-3g")
    with pytest.raises(FileNotFoundError):
        Molecule(WATER, "no-such-basis")
    with pytest.raises(AssertionError):
        Molecule("\n0 1 2\nH 0 0 0\n", "sto-3g")
    with pytest.raises(AssertionError):
        Molecule("\n0 1\nH 0 0\n", "sto-3g")


def test_dropin_type_errors_before_device():
    from mmd.integrals.fock import formPT
    from mmd.integrals.twoe import ERI, Basis, doERIs
    b = Basis([0, 0, 0], (0, 0, 0), 1, [1.0], [1.0])
    with pytest.raises(TypeError):
        ERI(b, b, b, "x")
    with pytest.raises(ValueError):
        doERIs(1, np.zeros((1, 1, 1, 1), dtype=np.float32), [b])
    with pytest.raises(TypeError):
        doERIs(1, np.zeros((1, 1, 1, 1)), (b,))
    with pytest.raises(ValueError):
        formPT(np.zeros((1, 1)), np.zeros((1, 1), complex), [b], 1, {0: 1.0}, 1e-12)


def test_basis_normalisation_self_overlap():
    """S_ii = 1 in closed form for every Cartesian component (s, p, d)."""
    from mmd.integrals.twoe import Basis
    e = np.array([3.0, 0.9, 0.21])
    c = np.array([0.3, 0.5, 0.4])
    for lmn in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 0, 0), (1, 1, 0), (0, 1, 1)]:
        b = Basis([0, 0, 0], lmn, 3, e, c)
        l, m, n = lmn
        L = l + m + n
        df = np.prod([max(1, np.prod(np.arange(2 * x - 1, 0, -2))) for x in lmn])
        w = b.norm * b.coefs
        s = np.sum(np.outer(w, w) * df * np.pi ** 1.5 / 2 ** L / (e[:, None] + e[None, :]) ** (L + 1.5))
        assert s == pytest.approx(1.0, rel=1e-14)


def test_631g_is_derived_from_631gs(tmp_path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_631g", os.path.join(BASIS, "make_631g.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = mod.derive(dst=str(tmp_path / "6-31g.gbs"))
    assert open(out).read() == open(_gbs("6-31g")).read()
    assert "\nD " not in open(out).read()


def test_c_abi_exports_every_declared_symbol():
    from mmd import _native
    header = open(os.path.join(ROOT, "include", "mmd.h")).read()
    declared = sorted(set(re.findall(r"\b(mmd_[a-z_0-9]+)\s*\(", header)))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_native.EXPORTED) == declared
    assert lib.mmd_version() == 1


def test_context_without_device_raises():
    from mmd import _native
    if _native.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError):
        _native.Context()
    assert "device" in _native.lib.mmd_last_error().decode()


def test_screen_dict_matches_reference_layout():
    from mmd.molecule import ScreenDict
    rng = np.random.default_rng(0)
    A = rng.random((5, 5))
    d = ScreenDict(A + A.T)
    assert isinstance(d, dict) and len(d) == 15
    for p in range(5):
        for q in range(p + 1):
            assert d[p * (p + 1) // 2 + q] == (A + A.T)[p, q]


def test_batch_molecule_arrays_match_molecule(oracle, basis_dir):
    """The batch driver's per-molecule arrays (element templates) are the
    basis and nuclear repulsion that Molecule and the oracle build."""
    import os
    import numpy as np
    from geometries import water_cluster
    from mmd.batch import _molecule_arrays
    from mmd.molecule import Molecule
    geom = water_cluster(3, 6.2, seed=4, perturb=True)
    Z, xyz, origin, lmn, nprim, exps, w, nocc, enuc = _molecule_arrays(geom, "6-31g")
    mol = Molecule(geom, "6-31g")
    assert np.array_equal(origin, np.array([b.origin for b in mol.bfs]))
    assert np.array_equal(lmn, np.array([b.shell for b in mol.bfs]))
    assert np.array_equal(nprim, np.array([b.num_exps for b in mol.bfs]))
    assert np.array_equal(exps, np.concatenate([b.exps for b in mol.bfs]))
    assert np.allclose(w, np.concatenate([b.norm * b.coefs for b in mol.bfs]), rtol=1e-15, atol=0)
    assert nocc == mol.nocc
    om = oracle.OracleMolecule(geom, os.path.join(basis_dir, "6-31g.gbs"))
    assert abs(enuc - om.nuc_energy) < 1e-12


def test_batched_learned_guess_matches_per_molecule(monkeypatch):
    """The batch driver's learned guess (features of all clusters, one forward
    pass, vectorised frames and assembly) equals learned_density()'s steps
    molecule by molecule: features() and assemble() on the same network
    output.  The network is replaced by a fixed linear map (no GPU here) and
    the atomic blocks by fixed symmetric matrices."""
    import numpy as np
    from geometries import water_cluster
    from mmd import _native, learned
    from mmd.batch import _molecule_arrays
    from mmd.molecule import Molecule
    rng = np.random.default_rng(3)
    model = learned.default_model()
    nout = model.dims[-1]
    Wmap = rng.standard_normal((15, nout))
    monkeypatch.setattr(_native, "mlp_forward", lambda ctx, m, X: X.astype(np.float64) @ Wmap)
    blocks = {}

    def element_density(Z, basis):
        if Z not in blocks:
            n = {8: 9, 1: 2}[Z]
            A = rng.standard_normal((n, n))
            blocks[Z] = A + A.T
        return blocks[Z]

    geoms = [water_cluster(k, 2.0 + 1.4 * k, seed=50 + k, perturb=True) for k in (1, 2, 4)]
    arrs = [_molecule_arrays(g, "6-31g") for g in geoms]
    got = learned.batch_learned_densities([(a[0], a[1], a[3], 2 * a[7]) for a in arrs], "6-31g", model,
                                          element_density=element_density)
    for g, a, P in zip(geoms, arrs, got):
        mol = Molecule(g, "6-31g")
        F = learned.features(mol)
        Fb, _ = learned.batch_features(a[0], a[1])
        assert np.allclose(Fb, F, rtol=1e-12, atol=1e-12)
        base = np.zeros((mol.nbasis, mol.nbasis))
        o = 0
        for atom in mol.atoms:
            b = element_density(atom.charge, "6-31g")
            base[o:o + len(b), o:o + len(b)] = b
            o += len(b)
        X = ((F - model.x_mean) / model.x_std).astype(np.float32)
        corr = (X.astype(np.float64) @ Wmap) * float(model.y_std)
        ref = learned.assemble(mol, base, corr)
        assert np.allclose(P, ref, rtol=0, atol=1e-10)


@pytest.mark.parametrize("activation", ["tanh", "silu"])
def test_mlp_reference_gradient_matches_finite_differences(activation):
    """The numpy perceptron that the GPU trainer is checked against: its
    analytic gradient against central differences (both activations)."""
    import mlp_ref
    dims = [5, 7, 6, 3]
    rng = np.random.default_rng(3)
    n = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    flat = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((11, 5))
    Y = rng.standard_normal((11, 3))
    _, g = mlp_ref.loss_grad(flat, dims, X, Y, activation)
    h = 1e-6
    fd = np.empty(n)
    for k in range(n):
        e = np.zeros(n)
        e[k] = h
        fd[k] = (mlp_ref.loss_grad(flat + e, dims, X, Y, activation)[0]
                 - mlp_ref.loss_grad(flat - e, dims, X, Y, activation)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


def _rotation(seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_pair_guess_invariants_round_trip_and_rotation():
    """Atom-pair guess geometry (mmd.pairguess): the bond-frame invariants of
    a density are unchanged when the cluster and the density are rotated
    together, and blocks built from invariants give back the invariants."""
    from mmd import pairguess as pg
    from mmd.learned import _monomer_rotation
    lay = {8: (9, ([0, 1, 5], [2, 6])), 1: (2, ([0, 1], []))}
    rng = np.random.default_rng(4)
    Z = np.array([8, 1, 1, 8, 1, 1])
    xyz = rng.standard_normal((6, 3)) * 2.5
    n = 26
    A = rng.standard_normal((n, n))
    P = A + A.T
    data = pg.training_pairs(Z, xyz, P, lay)
    assert set(data) == {(8, 8), (8, 1), (1, 8), (1, 1)}
    sizes = {(8, 8): 29, (8, 1): 10, (1, 8): 10, (1, 1): 4}
    for k, (F, Y) in data.items():
        assert F.shape[1] == pg.NRBF and Y.shape[1] == sizes[k]
        sh = pg.PairShape(lay[k[0]][1], lay[k[1]][1], lay[k[0]][0], lay[k[1]][0])
        assert np.allclose(sh.invariants(sh.blocks(Y)), Y, atol=1e-14)
    # rotate the cluster and the density (p triples of every atom by Q)
    Q = _rotation(5)
    lmn = np.array([[0, 0, 0]] * 2 + [[1, 0, 0], [0, 1, 0], [0, 0, 1]] + [[0, 0, 0]]
                   + [[1, 0, 0], [0, 1, 0], [0, 0, 1]] + [[0, 0, 0]] * 4)
    U = _monomer_rotation(lmn[:13], Q[None])[0]
    Ufull = np.kron(np.eye(2), U)
    P2 = Ufull @ P @ Ufull.T
    data2 = pg.training_pairs(Z, xyz @ Q.T, P2, lay)
    for k in data:
        assert np.allclose(data[k][0], data2[k][0], atol=1e-12)
        assert np.allclose(data[k][1], data2[k][1], atol=1e-10)
