"""GPU parity at the benchmark sizes: (H2O)8 / cc-pVDZ (N = 200) and
(H2O)16 / cc-pVDZ (N = 400), through the C ABI.

The fixtures in tests/golden/ hold oracle values (tests/golden/make_fixtures.py):
contracted ERIs of basis-function quartets of all 81 ordered (l_i l_j | l_k l_l)
combinations (cython/twoe.pyx:34-50) and elements of the symmetrised direct
Fock contribution with the reference's screen (cython/fock.pyx:38-85,
mmd/scf.py:93).  Tolerances: ERIs 1e-10 absolute; Fock elements 1e-10
absolute for the full density and 1e-14 absolute for the density scaled by
1e-6 (elements of order 1e-6, where the screen drops most quartets).

The class kernels are also compared with the independent per-function
kernel (formPT's own loop, symmetrised) under reduced LDS budgets, which send
classes to the kernel variants that large bases use: without the ket-side
Coulomb buffer, and on the wavefront-per-quartet kernel with survivor lists.
"""
import os

import numpy as np
import pytest

from geometries import water_cluster

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def density_model(origin, N):
    """The fixtures' test density (tests/golden/make_fixtures.py)."""
    idx = np.arange(1, N + 1, dtype=np.float64)
    C = np.cos(0.731 * np.outer(idx, idx))
    d2 = ((origin[:, None, :] - origin[None, :, :]) ** 2).sum(-1)
    return C * np.exp(-0.2 * d2)


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


_MOLS = {}


def _molecule(name):
    from mmd.molecule import Molecule
    if name not in _MOLS:
        fx = _load(name)
        mol = Molecule(str(fx["geometry"]), str(fx["basis"]))
        mol.build(direct=True)
        assert mol.nbasis == int(fx["nbasis"])
        _MOLS[name] = (mol, fx)
    return _MOLS[name]


@pytest.mark.parametrize("name", ["h2o8_ccpvdz", "h2o16_ccpvdz"])
def test_eri_sample_vs_oracle(native, name):
    mol, fx = _molecule(name)
    q = fx["quartets"].astype(np.int64)
    got = mol.native.eri_list(q)
    err = np.abs(got - fx["eri"])
    print("%s: %d quartets, max |dERI| %.2e, max |ERI| %.2e" % (name, len(q), err.max(), np.abs(fx["eri"]).max()))
    assert err.max() < 1e-10


@pytest.mark.parametrize("name", ["h2o8_ccpvdz", "h2o16_ccpvdz"])
def test_fock_elements_vs_oracle(native, name):
    mol, fx = _molecule(name)
    N = mol.nbasis
    origin = np.array([b.origin for b in mol.bfs])
    P = density_model(origin, N)
    pq = fx["pq"].astype(np.int64)
    tol = float(fx["tol"])
    for scale, ref, atol in ((1.0, fx["g_full"], 1e-10), (float(fx["small"]), fx["g_small"], 1e-14)):
        G, _, nsurv = mol.native.fock_delta(P * scale, None, tol=tol, raw=False)
        got = G[pq[:, 0], pq[:, 1]]
        err = np.abs(got - ref)
        print("%s scale %g: %d elements, %d shell quartets kept, max |dG| %.2e, max |G| %.2e"
              % (name, scale, len(pq), nsurv, err.max(), np.abs(ref).max()))
        assert err.max() < atol
        assert np.array_equal(G, G.T)


def _sym_random(N, rng, scale=1.0):
    A = rng.standard_normal((N, N)) * scale
    return A + A.T


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 5)])
def test_fock_lds_fallbacks_vs_generic(native, n, seed):
    """Every LDS budget gives the same G as the per-function kernel: the full
    budget (row kernels with the ket-side Coulomb buffer), budgets that
    leave the buffer out for some or all classes, and one that sends every
    class to the wavefront-per-quartet kernel."""
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(n, 7.0, seed), "cc-pvdz")
    mol.build(direct=True)
    ctx = mol.native
    rng = np.random.default_rng(seed + 40)
    dP = _sym_random(mol.nbasis, rng)
    Gr, _, _ = ctx.fock_delta(dP, None, tol=1e-12, raw=True)
    Gr = 0.5 * (Gr + Gr.T)
    try:
        for limit in (0, 48 * 1024, 40 * 1024, 32 * 1024, 1024):
            ctx.set_lds_limit(limit)
            G, _, _ = ctx.fock_delta(dP, None, tol=1e-12, raw=False)
            err = np.max(np.abs(G - Gr))
            print("LDS limit %6d: max |G - G_generic| %.2e" % (limit, err))
            assert err < 1e-10
    finally:
        ctx.set_lds_limit(0)


def test_fock_class_impl_modes_vs_generic(native):
    """The six classes with both kernels give the generic kernel's G on the
    row kernels (mode 0), on the matrix cores (mode 1) and with the measured
    choice (mode 2, over the builds that time each variant and after)."""
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(3, 7.0, 2), "cc-pvdz")
    mol.build(direct=True)
    ctx = mol.native
    rng = np.random.default_rng(11)
    dP = _sym_random(mol.nbasis, rng)
    Gr, _, _ = ctx.fock_delta(dP, None, tol=1e-12, raw=True)
    Gr = 0.5 * (Gr + Gr.T)
    try:
        for mode, builds in ((0, 1), (1, 1), (2, 4)):
            ctx.set_class_impl(mode)
            for b in range(builds):
                G, _, _ = ctx.fock_delta(dP, None, tol=1e-12, raw=False)
                err = np.max(np.abs(G - Gr))
                print("mode %d build %d: max |G - G_generic| %.2e, choice %s" % (mode, b, err,
                                                                              ctx.set_class_impl().tolist()))
                assert err < 1e-10
        ch = ctx.set_class_impl()
        assert all(c in (0, 1) for c in ch)
    finally:
        ctx.set_class_impl(2)


def test_fock_h2o8_classes_vs_generic(native):
    """The symmetrised class kernels at N = 200 against the per-function kernel
    (formPT's loop, symmetrised) on a density of the size of an SCF update."""
    mol, _ = _molecule("h2o8_ccpvdz")
    N = mol.nbasis
    origin = np.array([b.origin for b in mol.bfs])
    dP = density_model(origin, N) * 1e-3
    G, _, _ = mol.native.fock_delta(dP, None, tol=1e-12, raw=False)
    Gr, _, _ = mol.native.fock_delta(dP, None, tol=1e-12, raw=True)
    Gr = 0.5 * (Gr + Gr.T)
    err = np.max(np.abs(G - Gr))
    print("(H2O)8: max |G - G_generic| %.2e, max |G| %.2e" % (err, np.abs(G).max()))
    assert err < 1e-12


# Converged direct-SCF energy of (H2O)8 / cc-pVDZ, seed 0, 7 A cube (N = 200),
# acc2e 1e-12, conver 1e-8: the value of this FP64 path, whose Fock builds the
# oracle fixtures above pin element by element.  A regression pin (1e-8 Eh).
E_H2O8_CCPVDZ = -607.9348800362


def test_h2o8_direct_energy(native):
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(8, 7.0, 0), "cc-pvdz")
    mol.RHF(doPrint=False, direct=True, conver=1e-8)
    print("(H2O)8 cc-pVDZ: E = %.10f, %d iterations" % (mol.energy.real, mol.iterations))
    assert mol.is_converged
    assert abs(mol.energy.real - E_H2O8_CCPVDZ) < 1e-8


def test_h2o16_fp32_tier_energy(native):
    """(H2O)16 / cc-pVDZ (N = 400): the FP32 tier (shell quartets with bound
    below 1e-4) converges to the FP64 energy within 1 microhartree."""
    from mmd.molecule import Molecule
    geom = water_cluster(16, 9.0, 0)
    ref = Molecule(geom, "cc-pvdz")
    ref.RHF(doPrint=False, direct=True, conver=1e-8)
    mix = Molecule(geom, "cc-pvdz")
    mix.native.set_precision(1e-4)
    mix.RHF(doPrint=False, direct=True, conver=1e-8)
    ntot, n32 = mix.native.fock_tiers()
    print("(H2O)16: FP64 %.10f  tier %.10f  |dE| %.2e  FP32 share of the last build %d / %d"
          % (ref.energy.real, mix.energy.real, abs(mix.energy.real - ref.energy.real), n32, ntot))
    assert ref.is_converged and mix.is_converged
    assert n32 > 0
    assert abs(mix.energy.real - ref.energy.real) < 1e-6


def test_fp32_matrix_core_classes_close_to_fp64(native):
    """With every quartet in the FP32 tier (threshold 1), the FP32 class
    kernels (FP32 MFMA for the high-L classes, FP32 VALU for the others)
    stay within FP32 rounding of the FP64 build: 1e-5 of max |G|."""
    from mmd.molecule import Molecule
    mol = Molecule(water_cluster(2, 7.0, 0), "cc-pvdz")
    mol.build(direct=True)
    ctx = mol.native
    rng = np.random.default_rng(9)
    dP = _sym_random(mol.nbasis, rng, 0.1)
    G64, _, n64 = ctx.fock_delta(dP, None, tol=1e-12)
    ctx.set_precision(1.0)
    try:
        G32, _, n32 = ctx.fock_delta(dP, None, tol=1e-12)
        ntot, nfp32 = ctx.fock_tiers()
    finally:
        ctx.set_precision(0.0)
    err = np.max(np.abs(G32 - G64))
    print("all-FP32: %d of %d shell quartets in FP32, max |dG| %.2e of max |G| %.2e"
          % (nfp32, ntot, err, np.abs(G64).max()))
    assert n32 == n64 and nfp32 > 0.9 * ntot
    assert err < 1e-5 * np.abs(G64).max()
