"""Initial density guesses for ``SCF.RHF``.

The reference starts every SCF from the core Hamiltonian (P = 0, F = H;
mmd/scf.py:23-24,44-45).  ``sad_density`` builds the superposition of
atomic densities (SAD): a spherically averaged, fractionally occupied
restricted SCF on each distinct element in the molecule's own basis, its
density placed in the atom's diagonal block.  Densities follow the
reference's convention P = C_occ C_occ^T (half the electron density), so a
doubly occupied orbital enters with weight 1 and an electron with 1/2.

All integrals of the atomic problems come from the molecule's compute
context (libmmd); the small dense algebra runs on the host, as the SCF
driver's does.
"""
import numpy as np

_ATOM_CACHE = {}


def _occupations(eps, nelec, tol=1e-5):
    """Electrons per spatial orbital (0..2) by aufbau; orbitals whose
    energies agree within tol share the electrons of a partly filled level
    equally (spherical averaging of open p and d shells)."""
    n = np.zeros(len(eps))
    left = float(nelec)
    i = 0
    while left > 1e-12 and i < len(eps):
        j = i + 1
        while j < len(eps) and abs(eps[j] - eps[i]) < tol * max(1.0, abs(eps[i])):
            j += 1
        cap = 2.0 * (j - i)
        take = min(cap, left)
        n[i:j] = take / (j - i)
        left -= take
        i = j
    return n


def atomic_scf(S, H, TwoE, nelec, jk, maxiter=200, conver=1e-8):
    """Fractional-occupation restricted SCF of one atom.

    S, H: overlap and core Hamiltonian; TwoE: (pq|rs) tensor; jk(TwoE, P)
    returns (J, K) as the conventional Fock build does (mmd/scf.py:97-98).
    Returns the converged half density."""
    s, V = np.linalg.eigh(S)
    X = np.dot(V * s ** -0.5, V.T)
    P = np.zeros_like(S)
    for it in range(maxiter):
        J, K = jk(TwoE, P)
        F = H + 2.0 * J - K
        eps, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        occ = _occupations(eps, nelec)
        P_new = (C * (0.5 * occ)) @ C.T
        dP = np.linalg.norm(P_new - P)
        P = P_new if it == 0 else 0.5 * (P + P_new)
        if dP < conver:
            return P_new
    return P


def _default_atom_factory(symbol, basis):
    from mmd.molecule import Molecule
    atom = Molecule("0 1\n%s 0.0 0.0 0.0\n" % symbol, basis)
    atom.one_electron_integrals()
    atom.two_electron_integrals()
    return atom


def _native_jk(ctx):
    def jk(T, P):
        J, _, K, _ = ctx.jk_tensor(T, np.ascontiguousarray(P))
        return J, K
    return jk


def atomic_density(mol, atom, atom_factory=None):
    """Half density of the free atom of ``atom``'s element in ``mol``'s basis."""
    from mmd.molecule import _SYMBOLS
    key = (atom.charge, mol.basis_name)
    if key not in _ATOM_CACHE:
        make = atom_factory or _default_atom_factory
        a = make(_SYMBOLS[atom.charge], mol.basis_name)
        P = atomic_scf(np.real(a.S), np.real(a.T + a.V), a.TwoE, atom.charge, _native_jk(a.native))
        _ATOM_CACHE[key] = P
    return _ATOM_CACHE[key]


def sad_density(mol, atom_factory=None):
    """Block-diagonal SAD half density for ``mol`` (basis functions of an atom
    are contiguous, mmd/molecule.py:57-88), scaled to the molecule's
    electron count."""
    N = mol.nbasis
    P = np.zeros((N, N))
    i = 0
    for atom in mol.atoms:
        Pa = atomic_density(mol, atom, atom_factory)
        n = Pa.shape[0]
        P[i:i + n, i:i + n] = Pa
        i += n
    assert i == N
    ztot = sum(a.charge for a in mol.atoms)
    if ztot > 0 and ztot != mol.nelec:
        P *= mol.nelec / ztot
    return P
