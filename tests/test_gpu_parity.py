"""GPU parity: libmmd's HIP kernels against the CPU oracle and the
reference's own known answers, through the C ABI (ctypes).

Tolerances (stated per test): ERIs and Fock contributions 1e-10 absolute or
tighter, RHF energies 1e-8 Eh against the oracle, Boys values 5e-13 against
the reference's literature table (tests/test012.py of the reference).
"""
import os

import numpy as np
import pytest

from geometries import (FORMALDEHYDE, H2, HELIUM, METHANE, WATER, WATER_BOHR, BOHR_1986, from_bohr,
                        rescale, water_cluster)

pytestmark = pytest.mark.gpu

BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mcmurchie-davidson_amd",
                     "mmd", "basis")


def _gbs(name):
    return os.path.join(BASIS, name + ".gbs")


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


# --------------------------------------------------------------------------
# Boys function
# --------------------------------------------------------------------------
BOYS_LITERATURE = [  # (m, T, F_m(T)) -- tests/test012.py:12-112 of the reference
    (0.5, 6.8, 7.34475165333247E-02), (13, 14.1, 1.56775160456192E-07), (20.6, 32.4, 2.17602798734846E-14),
    (25, 6.4, 4.28028518677348E-05), (64, 50, 5.67024356263279E-24), (75.5, 40.8, 2.63173492081630E-20),
    (80.3, 78.2, 6.35062774057122E-36), (4, 7, 8.03538503977806E-04), (8.5, 3.6, 2.31681539108704E-03),
    (15.3, 20.7, 5.40914879973724E-10), (1.8, 25.3, 3.45745419193244E-04), (30, 26, 3.57321060811178E-13),
    (46.8, 37.6, 1.91851951160577E-18), (100, 125.1, 7.75391047694625E-55), (8, 16, 4.02308592502660E-07),
    (15, 27, 1.08359515555596E-11), (20, 30, 1.37585444267909E-13), (25, 13, 8.45734447905704E-08),
    (31, 34, 2.90561943091301E-16), (11, 38, 4.04561442253925E-12), (42, 32, 5.02183610419087E-16),
    (75, 30, 1.01429517438537E-15), (100, 33, 3.42689684943483E-17), (20, 1.4E-3, 2.43577075309547E-02),
    (45, 6.4E-5, 1.09883228385254E-02), (100, 2.6E-7, 4.97512309732144E-03), (8, 42, 1.11826597752251E-10),
    (16, 50, 2.40509456111904E-16), (21, 56, 1.43739730342730E-19), (12, 60, 4.05791663779760E-15),
    (15, 53, 3.14434039868936E-16), (18, 58, 1.78336953967902E-18), (8, 63, 3.56261924865627E-12),
    (14, 68, 3.09783511327517E-17), (20, 73, 1.71295886102040E-21), (33, 85, 1.74268831008018E-29),
    (36, 100, 3.08919970425521E-33), (100, 120, 4.97723065221079E-53), (5.7, 13.3, 9.02296149898981E-06),
    (0.5, 23.6, 2.11864406767677E-02), (23.8, 3.4, 7.92593349658604E-04), (25.8, 0.4, 1.29331240687006E-02),
    (28.3, 0.002, 1.73275865107165E-02), (36.6, 42.7, 7.12651246345736E-20), (43.2, 54.2, 1.53021328677383E-24),
    (64.3, 75.4, 5.52165865464571E-34), (104.6, 115.4, 1.26350192129925E-51), (115.6, 5.4, 2.03911971791491E-05),
]


def test_boys_general_literature(native):
    from mmd.integrals.onee import _boys
    for m, T, ref in BOYS_LITERATURE:
        assert abs(_boys(m, T) - ref) < 5e-13, (m, T)


def test_boys_table_vs_oracle(native, oracle):
    ctx = native.default_context()
    T = np.concatenate([np.linspace(0, 40, 4001), [1e-12, 1e-6, 35.99, 36.0, 36.01, 50, 100, 1e3, 1e5]])
    for m in range(9):
        got = ctx.boys(np.full(T.shape, float(m)), T, table=True)
        ref = np.array([oracle.boys(m, t) for t in T])
        assert np.max(np.abs(got - ref)) < 2e-14, m
        gen = ctx.boys(np.full(T.shape, float(m)), T, table=False)
        assert np.max(np.abs(gen - ref)) < 2e-15, m


# --------------------------------------------------------------------------
# One-electron integrals and the ERI tensor
# --------------------------------------------------------------------------
@pytest.mark.parametrize("geom,basis", [(WATER, "sto-3g"), (WATER, "cc-pvdz"), (FORMALDEHYDE, "6-31gs")])
def test_one_electron_vs_oracle(native, oracle, geom, basis):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.one_electron_integrals()
    om = oracle.OracleMolecule(geom, _gbs(basis))
    S, T, V, M, L = om.one_electron()
    for got, ref in ((mol.S, S), (mol.T, T), (mol.V, V), (mol.M, M), (mol.L, L)):
        assert np.max(np.abs(got - ref)) < 1e-11


@pytest.mark.parametrize("geom,basis,tol", [(WATER, "sto-3g", 1e-12), (WATER, "cc-pvdz", 1e-11),
                                            (HELIUM, "cc-pvtz", 1e-11)])
def test_eri_tensor_vs_oracle(native, oracle, geom, basis, tol):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.two_electron_integrals()
    ref = oracle.OracleMolecule(geom, _gbs(basis)).eri_tensor()
    assert np.max(np.abs(mol.TwoE - ref)) < tol


def test_doeris_and_eri_dropin(native):
    from mmd.integrals.twoe import ERI, doERIs
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "sto-3g")
    N = mol.nbasis
    T = np.zeros((N, N, N, N))
    mv = doERIs(N, T, mol.bfs)
    assert isinstance(mv, memoryview)
    T = np.asarray(mv)
    for (i, j, k, l) in [(0, 0, 0, 0), (1, 0, 2, 1), (4, 3, 2, 2), (6, 5, 4, 3), (2, 2, 6, 6)]:
        assert abs(ERI(mol.bfs[i], mol.bfs[j], mol.bfs[k], mol.bfs[l]) - T[i, j, k, l]) < 1e-13
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(T, T.transpose(perm))
    with pytest.raises(TypeError):
        ERI(mol.bfs[0], mol.bfs[1], mol.bfs[2], 3)
    with pytest.raises(ValueError):
        doERIs(N, np.zeros((N, N, N, N), dtype=np.float32), mol.bfs)
    with pytest.raises(TypeError):
        doERIs(N, T.copy(), tuple(mol.bfs))


def test_eri_list_vs_oracle_cluster(native, oracle):
    """Random basis-function quartets of (H2O)2/cc-pVDZ (all s,p,d classes)."""
    geom = water_cluster(2, 7.0, 0)
    from mmd.molecule import Molecule
    mol = Molecule(geom, "cc-pvdz")
    om = oracle.OracleMolecule(geom, _gbs("cc-pvdz"))
    rng = np.random.default_rng(1)
    idx = rng.integers(0, mol.nbasis, size=(3000, 4))
    got = mol.native.eri_list(idx)
    ref = om.eri_list(idx)
    assert np.max(np.abs(got - ref)) < 1e-11


def test_eri_tensor_cluster_vs_generic_and_oracle(native, oracle):
    """(H2O)2/cc-pVDZ stored tensor (shell-quartet class kernels, (dd|dd) on
    the wavefront kernel): every element against the per-function generic
    kernel, random and d-only elements against the oracle, 1e-11 absolute."""
    geom = water_cluster(2, 7.0, 0)
    from mmd.molecule import Molecule
    mol = Molecule(geom, "cc-pvdz")
    mol.two_electron_integrals()
    T = mol.TwoE
    N = mol.nbasis
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(T, T.transpose(perm))
    for i in range(N):   # every element, one leading index at a time
        idx = np.indices((1, N, N, N)).reshape(4, -1).T
        idx[:, 0] = i
        got = mol.native.eri_list(idx)
        assert np.max(np.abs(T[i].ravel() - got)) < 1e-11, i
    om = oracle.OracleMolecule(geom, _gbs("cc-pvdz"))
    rng = np.random.default_rng(5)
    dfun = np.array([k for k, b in enumerate(mol.bfs) if sum(b.shell) == 2])
    assert len(dfun) == 12
    for idx in (rng.integers(0, N, size=(3000, 4)), dfun[rng.integers(0, len(dfun), size=(3000, 4))]):
        ref = om.eri_list(idx)
        assert np.max(np.abs(T[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]] - ref)) < 1e-11


def test_schwarz_vs_oracle(native, oracle):
    from mmd.molecule import Molecule
    geom = water_cluster(2, 7.0, 3)
    mol = Molecule(geom, "cc-pvdz")
    d = mol.native.schwarz()
    ref = oracle.OracleMolecule(geom, _gbs("cc-pvdz")).schwarz()
    p, q = np.tril_indices(mol.nbasis)
    assert np.max(np.abs(d[p, q] - ref)) < 1e-11


# --------------------------------------------------------------------------
# Direct Fock build (formPT)
# --------------------------------------------------------------------------
def _sym_random(N, rng, scale=1.0):
    A = rng.standard_normal((N, N)) * scale
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("geom,basis", [(WATER, "sto-3g"), (WATER, "cc-pvdz"), (METHANE, "sto-3g")])
def test_formpt_raw_vs_oracle(native, oracle, geom, basis):
    """formPT's unsymmetrised output, same screening, same degeneracies."""
    from mmd.integrals.fock import formPT
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.build(direct=True)
    om = oracle.OracleMolecule(geom, _gbs(basis))
    rng = np.random.default_rng(7)
    N = mol.nbasis
    P = _sym_random(N, rng).astype(complex)
    P_old = _sym_random(N, rng, 1e-3).astype(complex)
    G = formPT(P, P_old, mol.bfs, N, mol.screen, 1e-12)
    Gref, _ = om.form_pt(P, P_old, om.schwarz(), 1e-12)
    assert G.dtype == np.complex128
    assert np.max(np.abs(G - Gref)) < 1e-10


def test_formpt_complex_vs_oracle(native, oracle):
    from mmd.integrals.fock import formPT
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "sto-3g")
    mol.build(direct=True)
    om = oracle.OracleMolecule(WATER, _gbs("sto-3g"))
    rng = np.random.default_rng(3)
    N = mol.nbasis
    A = rng.standard_normal((N, N))
    P = (0.5 * (A + A.T) + 1j * 0.5 * (A - A.T)).astype(complex)
    G = formPT(P, np.zeros_like(P), mol.bfs, N, mol.screen, 1e-12)
    Gref, _ = om.form_pt(P, np.zeros_like(P), om.schwarz(), 1e-12)
    assert np.max(np.abs(G - Gref)) < 1e-10


@pytest.mark.parametrize("geom,basis", [(WATER, "cc-pvdz"), (FORMALDEHYDE, "6-31gs")])
def test_fock_sym_vs_oracle(native, oracle, geom, basis):
    """Symmetrised G(dP) of the class kernels against the oracle's formPT,
    symmetrised as mmd/scf.py:93 does."""
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.build(direct=True)
    om = oracle.OracleMolecule(geom, _gbs(basis))
    rng = np.random.default_rng(11)
    N = mol.nbasis
    dP = _sym_random(N, rng)
    G, _, nsurv = mol.native.fock_delta(dP, None, tol=1e-12, raw=False)
    Gref, _ = om.form_pt(dP.astype(complex), np.zeros((N, N), complex), om.schwarz(), 1e-12)
    Gref = 0.5 * (Gref + Gref.T).real
    assert nsurv > 0
    assert np.max(np.abs(G - Gref)) < 1e-10
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 5)])
def test_fock_sym_vs_generic_cluster(native, n, seed):
    """Class kernels (symmetrised) against the independent per-function kernel
    (formPT's raw form, symmetrised) on water clusters with every s/p/d class."""
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(n, 7.0, seed), "cc-pvdz")
    mol.build(direct=True)
    rng = np.random.default_rng(seed)
    N = mol.nbasis
    for scale, tol in ((1.0, 1e-12), (1e-5, 1e-12), (1.0, 0.0)):
        dP = _sym_random(N, rng, scale)
        G, _, _ = mol.native.fock_delta(dP, None, tol=tol, raw=False)
        Gr, _, _ = mol.native.fock_delta(dP, None, tol=tol, raw=True)
        Gr = 0.5 * (Gr + Gr.T)
        assert np.max(np.abs(G - Gr)) < 1e-10 * max(1.0, scale)


def test_fock_linearity_and_zero(native):
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(2, 7.0, 1), "cc-pvdz")
    mol.build(direct=True)
    rng = np.random.default_rng(5)
    N = mol.nbasis
    A, B = _sym_random(N, rng), _sym_random(N, rng)
    GA, _, _ = mol.native.fock_delta(A, None, tol=0.0)
    GB, _, _ = mol.native.fock_delta(B, None, tol=0.0)
    GAB, _, _ = mol.native.fock_delta(A + B, None, tol=0.0)
    assert np.max(np.abs(GA + GB - GAB)) < 1e-10
    Z, _, ns = mol.native.fock_delta(np.zeros((N, N)), None, tol=1e-12)
    assert ns == 0 and not np.any(Z)


# --------------------------------------------------------------------------
# RHF energies
# --------------------------------------------------------------------------
def test_rhf_water_sto3g_reference_values(native):
    """tests/test002.py: -74.942079928192 (atol 1e-12) and the README dipole,
    at the geometry those values belong to."""
    from mmd.molecule import Molecule
    mol = Molecule(from_bohr(WATER_BOHR), "sto-3g")
    mol.RHF(doPrint=False, conver=1e-14)
    assert abs(mol.energy.real - (-74.942079928192)) < 1e-10
    assert abs(mol.mu[1].real - 1.53400931) < 1e-7
    mol.RHF(doPrint=False, direct=True, conver=1e-12)
    assert abs(mol.energy.real - (-74.942079928192)) < 1e-10


@pytest.mark.parametrize("geom,basis,direct", [
    (H2, "sto-3g", False), (WATER, "sto-3g", True), (METHANE, "sto-3g", True), (METHANE, "sto-3g", False),
    (WATER, "dz", False), (FORMALDEHYDE, "sto-3g", True), (HELIUM, "cc-pvtz", False)])
def test_rhf_vs_oracle(native, oracle, geom, basis, direct):
    from mmd.molecule import Molecule
    mol = Molecule(geom, basis)
    mol.RHF(doPrint=False, direct=direct, conver=1e-10)
    ref = oracle.rhf(oracle.OracleMolecule(geom, _gbs(basis)), direct=False, conver=1e-10)
    assert mol.is_converged
    assert abs(mol.energy.real - ref["energy"].real) < 1e-8
    assert np.max(np.abs(mol.mu - ref["mu"])) < 1e-6


def test_rhf_reference_known_answers(native):
    """The reference's own energies (tests/test001,003,010.py) at 1e-8 Eh
    (the input geometries those values were produced at differ from
    mmd/molecule.py's conversion by up to ~1e-8 Eh in energy).  Methane is
    converged tightly, so that the comparison does not depend on where an
    SCF stopped inside the 1e-8 convergence window (fully converged, the
    oracle gives -39.72685031531 Eh here)."""
    from mmd.molecule import Molecule
    mol = Molecule(H2, "sto-3g")
    mol.RHF(doPrint=False)
    assert abs(mol.energy.real - (-1.11675930740)) < 1e-10
    mol = Molecule(METHANE, "sto-3g")
    mol.RHF(doPrint=False, direct=True, conver=1e-12)
    assert abs(mol.energy.real - (-39.72685031531)) < 1e-10
    assert abs(mol.energy.real - (-39.726850324347)) < 2e-8
    mol = Molecule(rescale(WATER, BOHR_1986), "dz")
    mol.RHF(doPrint=False, conver=1e-12)
    assert abs(mol.energy.real - (-75.977878975377)) < 1e-9


def test_rhf_water_ccpvdz_direct_vs_conventional_oracle(native, oracle):
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "cc-pvdz")
    mol.RHF(doPrint=False, direct=True, conver=1e-10)
    ref = oracle.rhf(oracle.OracleMolecule(WATER, _gbs("cc-pvdz")), direct=False, conver=1e-10)
    assert abs(mol.energy.real - ref["energy"].real) < 1e-8


# --------------------------------------------------------------------------
# SAD initial guess (no reference counterpart): same converged energy as the
# reference's core guess, within 1e-8 Eh, in fewer iterations.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("geom,basis", [(WATER, "cc-pvdz"), (water_cluster(3, 6.0, 4), "6-31g")])
def test_sad_guess_matches_core_guess(native, geom, basis):
    from mmd.molecule import Molecule
    core = Molecule(geom, basis)
    core.RHF(doPrint=False, direct=True, conver=1e-10)
    sad = Molecule(geom, basis)
    sad.RHF(doPrint=False, direct=True, conver=1e-10, guess="sad")
    print("core %d iterations, SAD %d" % (core.iterations, sad.iterations))
    assert core.is_converged and sad.is_converged
    assert abs(sad.energy.real - core.energy.real) < 1e-8
    assert sad.iterations < core.iterations


# --------------------------------------------------------------------------
# Batched SCF (no reference counterpart): every molecule's energy equals its
# own single-molecule run (1e-10 Eh) and, for a (H2O)2 member, the oracle's
# conventional RHF (1e-8 Eh).
# --------------------------------------------------------------------------
def test_batch_matches_single_runs_and_oracle(native, oracle):
    from mmd.batch import run_batch
    from mmd.molecule import Molecule
    geoms = [water_cluster(n, 6.0, 100 + i, perturb=True) for i, n in enumerate([2, 3, 2, 4, 3, 2])]
    res = run_batch(geoms, "6-31g", workers=3, conver=1e-10)
    for g, r in zip(geoms, res):
        mol = Molecule(g, "6-31g")
        mol.RHF(doPrint=False, direct=True, conver=1e-10)
        assert r["converged"] and abs(r["energy"] - mol.energy.real) < 1e-10
        assert r["iterations"] == mol.iterations
    ref = oracle.rhf(oracle.OracleMolecule(geoms[0], _gbs("6-31g")), direct=False, conver=1e-10)
    assert abs(res[0]["energy"] - ref["energy"].real) < 1e-8


# --------------------------------------------------------------------------
# FP32 tier (no reference counterpart): quartets with screening bound below
# 1e-4 take FP32 integrals.  The Fock matrix stays within 1e-7 of FP64 and
# the converged energy within 1 uEh; fp32_below = 0 is bit-for-bit FP64.
# --------------------------------------------------------------------------
def test_fp32_tier_fock_close_to_fp64(native):
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(4, 6.0, 7), "cc-pvdz")
    mol.build(direct=True)
    ctx = mol.native
    rng = np.random.default_rng(3)
    dP = _sym_random(mol.nbasis, rng, 0.05)
    G64, _, n64 = ctx.fock_delta(dP, None, tol=1e-12)
    ctx.set_precision(1e-4)
    G32, _, n32 = ctx.fock_delta(dP, None, tol=1e-12)
    ntot, nfp32 = ctx.fock_tiers()
    ctx.set_precision(0.0)
    G64b, _, _ = ctx.fock_delta(dP, None, tol=1e-12)
    print("FP32 tier: %d of %d shell quartets, max |dG| %.2e" % (nfp32, ntot, np.max(np.abs(G32 - G64))))
    assert n32 == n64 == ntot and 0 < nfp32 < ntot
    assert np.max(np.abs(G32 - G64)) < 1e-7
    assert np.max(np.abs(G64b - G64)) < 1e-12
    with pytest.raises(ValueError):
        ctx.set_precision(-1.0)


def test_fp32_tier_energy_within_microhartree(native):
    from mmd.molecule import Molecule
    geom = water_cluster(4, 6.0, 11)
    ref = Molecule(geom, "cc-pvdz")
    ref.RHF(doPrint=False, direct=True, conver=1e-8)
    mix = Molecule(geom, "cc-pvdz")
    mix.native.set_precision(1e-4)
    mix.RHF(doPrint=False, direct=True, conver=1e-8)
    print("FP64 %.10f  tier %.10f  (%d / %d iterations)" % (ref.energy.real, mix.energy.real, ref.iterations,
                                                            mix.iterations))
    assert mix.is_converged
    assert abs(mix.energy.real - ref.energy.real) < 1e-6


# --------------------------------------------------------------------------
# Perceptron of the learned guess (no reference counterpart): HIP forward
# and backward against the numpy restatement (FP32: 1e-4 relative), and
# training by the HIP Adam loop.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("activation", ["tanh", "silu"])
def test_mlp_forward_and_gradient_vs_numpy(native, activation):
    import mlp_ref
    dims = [15, 32, 24, 91]
    net = native.MLP(dims, activation=activation)
    rng = np.random.default_rng(8)
    flat = (rng.standard_normal(net.nparam) * 0.3).astype(np.float32)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)
    X = rng.standard_normal((300, 15)).astype(np.float32)
    Y = rng.standard_normal((300, 91)).astype(np.float32)
    ref = mlp_ref.forward(flat.astype(np.float64), dims, X.astype(np.float64), activation)[-1]
    assert np.max(np.abs(net.forward(X) - ref)) < 1e-4 * np.max(np.abs(ref))
    loss, g = net.loss_grad(X, Y)
    loss_ref, g_ref = mlp_ref.loss_grad(flat.astype(np.float64), dims, X.astype(np.float64), Y.astype(np.float64),
                                        activation)
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert np.max(np.abs(g - g_ref)) < 1e-4 * np.max(np.abs(g_ref))
    with pytest.raises(ValueError):
        net.forward(np.zeros((3, 14), dtype=np.float32))


def test_mlp_training_fits_smooth_map(native):
    import mlp_ref
    rng = np.random.default_rng(9)
    dims = [4, 32, 32, 3]
    X = rng.uniform(-1, 1, (4096, 4)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0] + X[:, 1]), X[:, 2] * X[:, 3], np.cos(2 * X[:, 0]) * X[:, 3]], axis=1)
    net = native.MLP(dims)
    nparam = net.nparam
    net.set_params((rng.standard_normal(nparam) * 0.3).astype(np.float32))
    losses = net.train(X, Y.astype(np.float32), epochs=60, batch=256, lr=3e-3, seed=1)
    assert losses.shape == (60,) and losses[-1] < 0.1 * losses[0]
    final = mlp_ref.loss_grad(net.get_params().astype(np.float64), dims, X.astype(np.float64), Y)[0]
    assert final < 2 * losses[-1] + 1e-4


# --------------------------------------------------------------------------
# Learned guess (no reference counterpart): on held-out water clusters
# (generator seed 2, as the training script draws them) the learned start
# converges to the core-guess energy within 1e-8 Eh in fewer iterations
# than SAD in total.
# --------------------------------------------------------------------------
def test_learned_guess_gain_on_heldout(native):
    from mmd.molecule import Molecule
    rng = np.random.default_rng(2)
    ns = rng.integers(2, 7, size=6)
    geoms = [water_cluster(int(k), 2.0 + 1.4 * int(k), seed=2 * 1000003 + i, perturb=True) for i, k in enumerate(ns)]
    it = {"sad": 0, "learned": 0}
    for g in geoms:
        core = Molecule(g, "6-31g")
        core.RHF(doPrint=False, direct=True, conver=1e-10)
        for guess in ("sad", "learned"):
            m = Molecule(g, "6-31g")
            m.RHF(doPrint=False, direct=True, conver=1e-10, guess=guess)
            assert m.is_converged and abs(m.energy.real - core.energy.real) < 1e-8
            it[guess] += m.iterations
    print("iterations over %d held-out clusters: %s" % (len(geoms), it))
    assert it["learned"] < it["sad"]


def _random_pair_model(seed):
    from mmd import pairguess as pg
    from mmd.learned import Model
    layouts = pg.element_layouts("6-31g")
    nets = {}
    rng = np.random.default_rng(seed)
    for zi in pg.ELEMENTS:
        for zj in pg.ELEMENTS:
            nout = pg.PairShape(layouts[zi][1], layouts[zj][1], layouts[zi][0], layouts[zj][0]).size
            dims = [pg.NRBF] + list(pg.HIDDEN) + [nout]
            m = Model.init(dims, np.zeros(pg.NRBF), np.ones(pg.NRBF), 0.05, seed=int(rng.integers(1 << 30)))
            m.activation = "silu"
            nets[(zi, zj)] = m
    return pg.PairModel(nets, layouts)


def test_pair_guess_rotates_with_the_cluster(native):
    """Atom-pair guess on the GPU networks (random weights): the density of a
    rotated cluster is the rotated density (p functions of every atom turned
    by the rotation), and it is symmetric."""
    from mmd import pairguess as pg
    from mmd.batch import _molecule_arrays
    from mmd.learned import _monomer_rotation
    model = _random_pair_model(1)
    g = water_cluster(3, 7.0, 5)
    Z, xyz, _, lmn, _, _, _, nocc, _ = _molecule_arrays(g, "6-31g")
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    Q = Q @ np.array([[1.0, 0.0, 0.0], [0.0, np.cos(0.4), -np.sin(0.4)], [0.0, np.sin(0.4), np.cos(0.4)]])
    P1, P2 = pg.batch_pair_densities([(Z, xyz, lmn, 2 * nocc), (Z, xyz @ Q.T, lmn, 2 * nocc)], "6-31g", model)
    U = np.kron(np.eye(3), _monomer_rotation(lmn[:13], Q[None])[0])
    assert np.max(np.abs(P1 - P1.T)) < 1e-14
    assert np.max(np.abs(P1[:9, 9:11])) > 1e-3           # O-H blocks filled
    assert np.max(np.abs(U @ P1 @ U.T - P2)) < 2e-5 * np.max(np.abs(P1))


def test_pair_guess_converges_to_the_core_start_energy(native):
    """The trained atom-pair networks (mmd/data/pair_guess_631g.npz) as the
    start of the per-molecule and the batched SCF: same energies as the core
    start."""
    from mmd.batch import run_batch_device
    from mmd.molecule import Molecule
    geoms = [water_cluster(k, 2.0 + 1.4 * k, seed=17 + k, perturb=True) for k in (2, 4)]
    batch = run_batch_device(geoms, "6-31g", guess="pair", conver=1e-10)
    for g, b in zip(geoms, batch):
        core = Molecule(g, "6-31g")
        core.RHF(doPrint=False, direct=True, conver=1e-10)
        m = Molecule(g, "6-31g")
        m.RHF(doPrint=False, direct=True, conver=1e-10, guess="pair")
        assert m.is_converged and b["converged"]
        assert abs(m.energy.real - core.energy.real) < 1e-8
        assert abs(b["energy"] - core.energy.real) < 1e-8


# --------------------------------------------------------------------------
# The GPU-resident SCF loop (bases of >= 100 functions, mmd_scf_*) steps as
# the host loop: same iterations, energies within 1e-10 Eh, same density.
# --------------------------------------------------------------------------
def test_device_scf_loop_matches_host_loop(native):
    from mmd.molecule import Molecule
    geom = water_cluster(5, 6.5, 3)
    dev = Molecule(geom, "cc-pvdz")
    assert dev.nbasis >= dev.device_loop_min_nbasis
    dev.RHF(doPrint=False, direct=True, conver=1e-9)
    host = Molecule(geom, "cc-pvdz")
    host.device_loop_min_nbasis = None
    host.RHF(doPrint=False, direct=True, conver=1e-9)
    print("device %d / host %d iterations, dE %.2e" % (dev.iterations, host.iterations,
                                                       abs(dev.energy.real - host.energy.real)))
    assert dev.is_converged and host.is_converged and dev.iterations == host.iterations
    assert abs(dev.energy.real - host.energy.real) < 1e-10
    assert np.max(np.abs(dev.P - host.P)) < 1e-8 and dev.P.dtype == np.complex128
    assert np.max(np.abs(dev.F - host.F)) < 1e-8 and np.max(np.abs(dev.mu - host.mu)) < 1e-7
    sad = Molecule(geom, "cc-pvdz")
    sad.RHF(doPrint=False, direct=True, conver=1e-9, guess="sad")
    assert abs(sad.energy.real - host.energy.real) < 1e-8


# --------------------------------------------------------------------------
# Four-index transformation (PostSCF.ao2mo, mmd/postscf.py:21-43) on the GPU:
# against numpy's einsum transform (1e-12) and the reference's MP2 value of
# tests/test014.py (methane / 3-21G).
# --------------------------------------------------------------------------
def test_ao2mo_and_mp2_reference(native):
    from mmd.molecule import Molecule
    from mmd.postscf import PostSCF
    mol = Molecule(METHANE, "3-21g")
    mol.RHF(doPrint=False, direct=False)
    PostSCF(mol).MP2()
    assert abs(mol.energy.real - (-39.9768654272)) < 4e-6
    assert abs(mol.emp2.real - (-40.076963354817)) < 4e-6
    C = mol.C.real
    ref = np.einsum("pqrs,pP,qQ,rR,sS->PQRS", mol.TwoE, C, C, C, C, optimize=True)
    assert np.max(np.abs(mol.single_bar - ref)) < 1e-12
    e = mol.emp2.real
    PostSCF(mol).MP2(spin_orbital=True)
    assert abs(mol.emp2.real - e) < 1e-10


# --------------------------------------------------------------------------
# Real-time TDHF (mmd/realtime.py) on the GPU J/K of complex densities:
# tests/test011.py of the reference (energy conserved without a field), and
# the GPU propagation of water under a pulse against the oracle-free host
# contraction of the same stored tensor.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("method", ["Magnus2", "Magnus4"])
def test_realtime_reference_trivial(native, method):
    from mmd.molecule import Molecule
    from mmd.realtime import RealTime
    mol = Molecule("\n0 1\nH 0.0 0.0 0.74\nH 0.0 0.0 0.00\n", "sto-3g")
    rt = RealTime(mol, numsteps=100, stepsize=0.2, field=0.0001, pulse=None)
    e0 = rt.mol.energy.real
    getattr(rt, method)(direction='z')
    np.testing.assert_allclose(e0 * np.ones_like(rt.Energy), rt.Energy)


def test_realtime_complex_jk_vs_einsum(native):
    from mmd.molecule import Molecule
    mol = Molecule(WATER, "3-21g")
    mol.RHF(doPrint=False, direct=False)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((mol.nbasis, mol.nbasis)) + 1j * rng.standard_normal((mol.nbasis, mol.nbasis))
    P = 0.5 * (A + A.conj().T) * 0.1
    mol.P = P
    mol.buildFock()
    J = np.einsum("pqrs,sr->pq", mol.TwoE, P)
    K = np.einsum("psqr,sr->pq", mol.TwoE, P)
    assert np.max(np.abs(mol.G - (2 * J - K))) < 1e-11
