"""CPU oracle for the McMurchie-Davidson ERI / Fock-build path.

TEST INFRASTRUCTURE ONLY.  This package is the checker the HIP product path
(the ``mmd`` package over ``libmmd.so``) is compared against; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and never as the thing that is measured or shipped.

It restates the reference's algorithm:

* ``mmd_oracle.c`` -- Boys, E, R, primitive and contracted ERIs, ``doERIs``,
  ``formPT``, normalisation and the one-electron integrals
  (``cython/util.pxi``, ``cython/twoe.pyx``, ``cython/fock.pyx``,
  ``cython/basis.pxi``, ``cython/onee.pyx``);
* this module -- geometry / basis-file parsing and basis-function ordering
  (``mmd/molecule.py:57-233``) and the RHF driver (``mmd/scf.py:10-180``),
  in numpy.

Every function is pinned by checks in ``tests/test_oracle.py`` (the
reference's own known answers plus closed forms and invariants); none is
"parity unpinned".
"""
import ctypes
import itertools
import os

import numpy as np
import scipy.linalg
from scipy.linalg import fractional_matrix_power as _mat_pow

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libmmd_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_pi = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle library (gcc, no GPU)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_boys.restype = _d
        L.orc_boys.argtypes = [_d, _d]
        L.orc_fact2.restype = _d
        L.orc_fact2.argtypes = [ctypes.c_int]
        L.orc_E.restype = _d
        L.orc_E.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d, ctypes.c_int, _d]
        L.orc_R.restype = _d
        L.orc_R.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d, _d, _d]
        L.orc_normalize.restype = None
        L.orc_normalize.argtypes = [_i64, _i64, _i64, _i64, _pd, _pd, _pd]
        bfs = [_pd, _pi, _pi, _pi, _pd, _pd, _pd]
        L.orc_eri_list.restype = None
        L.orc_eri_list.argtypes = bfs + [_i64, _pi, _pd]
        L.orc_do_eris.restype = None
        L.orc_do_eris.argtypes = bfs + [_i64, _pd]
        L.orc_schwarz.restype = None
        L.orc_schwarz.argtypes = bfs + [_i64, _pd]
        L.orc_form_pt_survivors.restype = _i64
        L.orc_form_pt_survivors.argtypes = [_i64, _pd, _pd, _d]
        L.orc_form_pt.restype = _i64
        L.orc_form_pt.argtypes = bfs + [_i64, _pd, ctypes.c_void_p, _pd, _d, _pd,
                                        ctypes.c_void_p, _i64, _i64]
        L.orc_fock_elements.restype = None
        L.orc_fock_elements.argtypes = bfs + [_i64, _pd, _pd, _d, _i64, _pi, _pd]
        L.orc_rhf_gradient.restype = None
        L.orc_rhf_gradient.argtypes = bfs + [_i64, _i64, _pd, _pd, _pi, _pd, _pd, _pd]
        L.orc_one_electron.restype = None
        L.orc_one_electron.argtypes = bfs + [_i64, _i64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd]
        _lib = L
    return _lib


def boys(m, T):
    """F_m(T) (cython/util.pxi:53-55)."""
    return lib().orc_boys(float(m), float(T))


def fact2(n):
    return lib().orc_fact2(int(n))


def E(i, j, t, Qx, a, b, n=0, Ax=0.0):
    """Hermite coefficient (cython/util.pxi:13-28)."""
    return lib().orc_E(int(i), int(j), int(t), float(Qx), float(a), float(b), int(n), float(Ax))


def R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    """Hermite Coulomb integral (cython/util.pxi:33-50)."""
    return lib().orc_R(int(t), int(u), int(v), int(n), float(p), float(PCx), float(PCy),
                       float(PCz), float(RPC))


# --------------------------------------------------------------------------
# Geometry and basis files (mmd/molecule.py:104-233)
# --------------------------------------------------------------------------
SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
           "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"]
BOHR = 0.52917721092
SHELLS = {
    "S": [(0, 0, 0)],
    "P": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "D": [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
}


def read_geometry(geometry):
    """(charge, multiplicity, [(Z, xyz_bohr)]) as mmd/molecule.py:191-233."""
    lines = [ln for ln in geometry.split("\n") if ln]
    atoms = []
    charge = mult = None
    for n, line in enumerate(lines):
        f = line.split()
        if n == 0:
            assert len(f) == 2
            charge, mult = int(f[0]), int(f[1])
            continue
        if len(f) == 0:
            break
        assert len(f) == 4
        Z = SYMBOLS.index(f[0])
        atoms.append((Z, np.array([float(f[1]), float(f[2]), float(f[3])]) / BOHR))
    return charge, mult, atoms


def read_gbs(path):
    """{Z: [(mom, [(exp, coef), ...]), ...]} as mmd/molecule.py:135-189."""
    with open(path) as fh:
        blocks = fh.read().split("****")
    basis = {}
    for blk in blocks[1:]:
        rows = [x.split() for x in blk.split("\n")[1:-1]]
        new = True
        atom = None
        for idx, f in enumerate(rows):
            if not f:
                continue
            if idx == 0:
                assert len(f) == 2
                atom = SYMBOLS.index(f[0]) if f[0] in SYMBOLS else -1
                basis[atom] = []
                new = True
            elif new:
                mom, nprim, cnt, p1, p2, new = f[0], int(f[1]), 0, [], [], False
            else:
                val = [float(x.replace("D", "E")) for x in f]
                cnt += 1
                if mom == "SP":
                    p1.append((val[0], val[1]))
                    p2.append((val[0], val[2]))
                    if cnt == nprim:
                        basis[atom].append(("S", p1))
                        basis[atom].append(("P", p2))
                        new = True
                else:
                    p1.append((val[0], val[1]))
                    if cnt == nprim:
                        basis[atom].append((mom, p1))
                        new = True
    return basis


class OracleMolecule:
    """Basis functions in reference order plus the one-electron quantities
    (mmd/molecule.py:29-102, 235-286)."""

    def __init__(self, geometry, basis_path):
        self.charge, self.multiplicity, self.atoms = read_geometry(geometry)
        self.nelec = sum(Z for Z, _ in self.atoms) - self.charge
        self.nocc = self.nelec // 2
        data = read_gbs(basis_path)
        origin, shell, nps, exps, coefs, bf_atom = [], [], [], [], [], []
        for ia, (Z, xyz) in enumerate(self.atoms):
            for mom, prims in data[Z]:
                e = [p[0] for p in prims]
                c = [p[1] for p in prims]
                for lmn in SHELLS[mom]:
                    bf_atom.append(ia)
                    origin.append(xyz)
                    shell.append(lmn)
                    nps.append(len(e))
                    exps.extend(e)
                    coefs.extend(c)
        self.nbasis = len(shell)
        self.bf_atom = np.array(bf_atom, dtype=np.int64)
        self.origin = np.ascontiguousarray(np.array(origin, dtype=np.float64).reshape(-1, 3))
        self.shell = np.ascontiguousarray(np.array(shell, dtype=np.int64).reshape(-1, 3))
        self.np = np.array(nps, dtype=np.int64)
        self.poff = np.concatenate([[0], np.cumsum(self.np)[:-1]]).astype(np.int64)
        self.exps = np.array(exps, dtype=np.float64)
        self.coefs = np.array(coefs, dtype=np.float64)
        self.norm = np.zeros_like(self.exps)
        L = lib()
        for i in range(self.nbasis):
            o, k = self.poff[i], self.np[i]
            e = np.ascontiguousarray(self.exps[o:o + k])
            c = np.ascontiguousarray(self.coefs[o:o + k])
            nr = np.zeros(k)
            L.orc_normalize(int(self.shell[i, 0]), int(self.shell[i, 1]), int(self.shell[i, 2]),
                            int(k), e, c, nr)
            self.coefs[o:o + k] = c
            self.norm[o:o + k] = nr
        Zs = np.array([Z for Z, _ in self.atoms], dtype=np.float64)
        xyz = np.array([x for _, x in self.atoms], dtype=np.float64).reshape(-1, 3)
        self.Z = Zs
        self.xyz = np.ascontiguousarray(xyz)
        self.center_of_charge = (Zs[:, None] * xyz).sum(0) / Zs.sum()
        self.nuc_energy = 0.0
        for (Za, xa), (Zb, xb) in itertools.combinations(self.atoms, 2):
            self.nuc_energy += Za * Zb / np.linalg.norm(xa - xb)

    def _bfs(self):
        return (self.origin, self.shell, self.np, self.poff, self.exps, self.coefs, self.norm)

    def rhf_gradient(self, P, F):
        """dE/dX, shape (natom, 3), of the converged RHF energy for density P
        and Fock matrix F (mmd/forces.py:8-97, cython/grad.pyx); the
        reference's forces are its negative."""
        N = self.nbasis
        P = np.ascontiguousarray(np.real(P), dtype=np.float64)
        W = np.ascontiguousarray(P @ np.real(F) @ P)
        g = np.zeros((len(self.atoms), 3))
        lib().orc_rhf_gradient(*self._bfs(), N, len(self.atoms), self.Z, self.xyz, self.bf_atom, P, W, g)
        return g

    def one_electron(self):
        N = self.nbasis
        S, T, V = np.zeros((N, N)), np.zeros((N, N)), np.zeros((N, N))
        M, Lm = np.zeros((3, N, N)), np.zeros((3, N, N))
        lib().orc_one_electron(*self._bfs(), N, len(self.atoms), self.Z, self.xyz,
                               np.ascontiguousarray(self.center_of_charge), S, T, V, M, Lm)
        return S, T, V, M, Lm

    def eri_tensor(self):
        N = self.nbasis
        T = np.zeros((N, N, N, N))
        lib().orc_do_eris(*self._bfs(), N, T)
        return T

    def eri_list(self, ijkl):
        ijkl = np.ascontiguousarray(np.asarray(ijkl, dtype=np.int64).reshape(-1, 4))
        out = np.zeros(len(ijkl))
        lib().orc_eri_list(*self._bfs(), len(ijkl), ijkl, out)
        return out

    def schwarz(self):
        N = self.nbasis
        s = np.zeros(N * (N + 1) // 2)
        lib().orc_schwarz(*self._bfs(), N, s)
        return s

    def form_pt(self, P, P_old, screen, tol, limit=-1, start_ij=0):
        """formPT (cython/fock.pyx:13-87); returns (G complex unsymmetrised, survivors)."""
        N = self.nbasis
        dP = np.asarray(P) - np.asarray(P_old)
        dPr = np.ascontiguousarray(dP.real, dtype=np.float64)
        cplx = np.iscomplexobj(dP) and np.any(dP.imag != 0)
        dPi = np.ascontiguousarray(dP.imag, dtype=np.float64) if cplx else None
        Gr = np.zeros((N, N))
        Gi = np.zeros((N, N)) if cplx else None
        nsurv = lib().orc_form_pt(*self._bfs(), N, dPr,
                                  dPi.ctypes.data if cplx else None,
                                  np.ascontiguousarray(screen, dtype=np.float64), float(tol), Gr,
                                  Gi.ctypes.data if cplx else None, int(limit), int(start_ij))
        G = Gr.astype(complex)
        if cplx:
            G = G + 1j * Gi
        return G, nsurv


def form_pt_survivors(dP, screen, tol):
    """Quartets formPT keeps for the real difference dP = P - P_old
    (cython/fock.pyx:38-57), counted without computing integrals."""
    dP = np.ascontiguousarray(np.real(dP), dtype=np.float64)
    return int(lib().orc_form_pt_survivors(dP.shape[0], dP, np.ascontiguousarray(screen, dtype=np.float64),
                                           float(tol)))


def fock_elements(mol, P, S_diag, tol, pq):
    """Elements (p, q) of the symmetrised formPT(P, 0) for a real symmetric
    P (cython/fock.pyx:38-85, mmd/scf.py:93); S_diag is the full N*N
    Schwarz diagonal (pq|pq)."""
    N = mol.nbasis
    pq = np.ascontiguousarray(np.asarray(pq, dtype=np.int64).reshape(-1, 2))
    out = np.zeros(len(pq))
    lib().orc_fock_elements(*mol._bfs(), N, np.ascontiguousarray(P, dtype=np.float64),
                            np.ascontiguousarray(S_diag, dtype=np.float64), float(tol), len(pq), pq, out)
    return out


def schwarz_full(mol):
    """The Schwarz diagonal as a full symmetric N*N matrix."""
    N = mol.nbasis
    s = mol.schwarz()
    S = np.zeros((N, N))
    iu = np.tril_indices(N)
    S[iu] = s
    return S + np.tril(S, -1).T


def rhf(mol, DIIS=True, direct=False, conver=1e-8, acc2e=1e-12, maxiter=100, eri=None):
    """RHF restated from mmd/scf.py:10-80 (buildFock 82-100, DIIS 146-180).
    Returns a dict with energy, P, F, C, MO, mu, is_converged, iterations."""
    N = mol.nbasis
    S, T, V, M, _ = mol.one_electron()
    Core = T + V
    X = _mat_pow(S, -0.5)
    if direct:
        screen = mol.schwarz()
    else:
        TwoE = mol.eri_tensor() if eri is None else eri
    P_old = np.zeros((N, N), dtype=complex)
    P = P_old
    F = Core.astype(complex)
    fockSet, errorSet = [], []
    energy = None
    P_RMS = 1e20
    delta_energy = 1e20
    converged = False
    FO = None
    step = 0
    for step in range(maxiter):
        if step > 0:
            F_old = F
            energy_old = energy
            if direct:
                G, _ = mol.form_pt(P, P_old, screen, acc2e)
                G = 0.5 * (G + G.T)
                F = F_old + G
            else:
                J = np.einsum("pqrs,sr->pq", TwoE.astype(complex), P)
                K = np.einsum("psqr,sr->pq", TwoE.astype(complex), P)
                F = Core.astype(complex) + 2.0 * J - K
            P_old = P
            if DIIS:
                FPS = F @ P @ S
                err = X @ (FPS - FPS.conj().T) @ X
                fockSet.append(F)
                errorSet.append(err)
                if len(fockSet) > 8:
                    del fockSet[0]
                    del errorSet[0]
                n = len(fockSet)
                B = np.zeros((n + 1, n + 1))
                B[-1, :] = B[:, -1] = -1.0
                B[-1, -1] = 0.0
                for i in range(n):
                    for j in range(i + 1):
                        B[i, j] = B[j, i] = np.real(np.trace(errorSet[i].conj().T @ errorSet[j]))
                rhs = np.zeros(n + 1)
                rhs[-1] = -1.0
                w = np.linalg.solve(B, rhs)
                Fd = np.zeros((N, N), dtype=complex)
                for i, Fk in enumerate(fockSet):
                    Fd += w[i] * Fk
                FO = X.T @ Fd @ X
        if not DIIS or step == 0:
            FO = X.T @ F @ X
        e, CO = scipy.linalg.eigh(FO)
        C = X @ CO
        P = C[:, :mol.nocc] @ np.conjugate(C[:, :mol.nocc]).T
        energy = np.einsum("pq,qp", Core + F, P) + mol.nuc_energy
        if step > 0:
            delta_energy = energy - energy_old
            P_RMS = np.linalg.norm(P - P_old)
        if np.abs(P_RMS) < conver:
            converged = True
            break
    mu = np.zeros(3, dtype=complex)
    for i in range(3):
        mu[i] = -2 * np.trace(P @ M[i]) + sum(
            Z * (x[i] - mol.center_of_charge[i]) for Z, x in mol.atoms)
    mu *= 2.541765
    return dict(energy=energy, P=P, F=F, C=C, MO=e, mu=mu, is_converged=converged,
                iterations=step, delta_energy=delta_energy, P_RMS=P_RMS, S=S, Core=Core, X=X)
