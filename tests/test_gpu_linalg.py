"""Matrix-core products and what is built on them, through the C ABI:

- mmd_dgemm (v_mfma_f64_16x16x4_f64) against numpy for every transpose
  combination, ragged sizes and beta != 0 (tolerance 1e-13 relative to the
  size of the terms);
- the overlap powers X = S^-1/2 and U = S^1/2 (mmd/molecule.py:284-286,
  Newton-Schulz on the GPU) against the eigendecomposition formula the
  reference's fractional_matrix_power evaluates, and X S X = I, U U = S;
- the integral files of Molecule.save_integrals (mmd/molecule.py:295-315):
  the Crawford-group eri.dat layout, one line "i j k l (ij|kl)" per element
  with 1-based indices in row-major order, and the values of the oracle's
  tensor (1e-12).
"""
import os

import numpy as np
import pytest

from geometries import WATER, water_cluster

pytestmark = pytest.mark.gpu

BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mcmurchie-davidson_amd",
                     "mmd", "basis")


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


@pytest.fixture(scope="module")
def ctx(native):
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "sto-3g")
    return mol.native


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 45), (100, 1000, 100),
                                   (200, 200, 200), (3, 129, 257)])
def test_dgemm_vs_numpy(ctx, M, N, K):
    rng = np.random.default_rng(M * 1000 + N + K)
    for tA in (False, True):
        for tB in (False, True):
            A = rng.standard_normal((K, M) if tA else (M, K))
            B = rng.standard_normal((N, K) if tB else (K, N))
            C0 = rng.standard_normal((M, N))
            ref = 0.7 * (A.T if tA else A) @ (B.T if tB else B) - 1.3 * C0
            got = ctx.dgemm(A, B, transA=tA, transB=tB, alpha=0.7, beta=-1.3, C=C0)
            scale = np.sqrt(K) * 4.0
            err = np.max(np.abs(got - ref))
            assert err < 1e-13 * scale * 10, (tA, tB, err)


@pytest.mark.parametrize("geom,basis", [(WATER, "cc-pvdz"), (WATER, "6-31ppgss"),
                                        (water_cluster(8, 7.0, 0), "cc-pvdz")])
def test_overlap_powers_vs_eigh(native, geom, basis):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.one_electron_integrals()
    S = mol.S
    s, V = np.linalg.eigh(S)
    Xr = (V * s ** -0.5) @ V.T
    Ur = (V * s ** 0.5) @ V.T
    X, U, it = mol.native.overlap_powers()
    N = S.shape[0]
    print("%s N=%d cond(S)=%.1e: %d iterations, max|dX| %.2e, max|dU| %.2e"
          % (basis, N, s[-1] / s[0], it, np.abs(X - Xr).max(), np.abs(U - Ur).max()))
    # both sides carry rounding of order eps * cond(S) relative to |X|
    assert np.abs(X - Xr).max() < 1e-14 * (s[-1] / s[0]) * np.abs(Xr).max() + 1e-12
    assert np.abs(U - Ur).max() < 1e-12
    assert np.abs(X @ S @ X - np.eye(N)).max() < 1e-14 * (s[-1] / s[0]) + 1e-12
    assert np.abs(U @ U - S).max() < 1e-12
    assert np.array_equal(X, X.T) and np.array_equal(U, U.T)
    # the molecule's own X and U are these
    assert np.array_equal(mol.X, X) and np.array_equal(mol.U, U)
    # an explicit S gives the same result; an indefinite one is rejected
    X2, _, _ = mol.native.overlap_powers(S)
    assert np.abs(X2 - X).max() < 1e-14
    with pytest.raises(ValueError):
        mol.native.overlap_powers(-S)


def test_save_integrals_eri_dat(native, oracle, tmp_path):
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "sto-3g")
    mol.build()
    folder = str(tmp_path / "ints")
    mol.save_integrals(folder=folder)
    N = mol.nbasis
    T = oracle.OracleMolecule(WATER, os.path.join(BASIS, "sto-3g.gbs")).eri_tensor()
    with open(os.path.join(folder, "eri.dat")) as f:
        lines = f.read().splitlines()
    assert len(lines) == N ** 4
    n = 0
    for i in range(N):
        for j in range(N):
            for k in range(N):
                for l in range(N):
                    fields = lines[n].split()
                    n += 1
                    assert [int(x) for x in fields[:4]] == [i + 1, j + 1, k + 1, l + 1]
                    assert len(fields) == 5
                    assert abs(float(fields[4]) - T[i, j, k, l]) < 1e-12
                    # printed at full precision: the value reads back exactly
                    assert float(fields[4]) == mol.TwoE[i, j, k, l]
    assert int(np.loadtxt(os.path.join(folder, "nbf.dat"))) == N
    assert int(np.loadtxt(os.path.join(folder, "nelec.dat"))) == mol.nelec
    assert float(np.loadtxt(os.path.join(folder, "enuc.dat"))) == pytest.approx(mol.nuc_energy, abs=1e-15)
    for name, M in (("s", mol.S), ("t", mol.T), ("v", mol.V)):
        assert np.abs(np.loadtxt(os.path.join(folder, name + ".dat")) - M).max() < 1e-15
    with open(os.path.join(folder, "geometry.txt")) as f:
        assert f.read().strip() == mol.geometry_input.strip()
