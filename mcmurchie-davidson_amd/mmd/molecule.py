"""``Molecule(geometry, basis)``: input parsing, basis construction and the
integral build, with the API of the reference's mmd/molecule.py.

Host-side logic (parsing, basis-function order, normalisation) follows
mmd/molecule.py:13-233; every integral is computed by libmmd on the GPU.
"""
import itertools
import os
import sys

import numpy as np

from mmd import _native
from mmd.integrals.twoe import Basis
from mmd.forces import Forces
from mmd.scf import SCF

BOHR_ANGSTROM = 0.52917721092   # mmd/molecule.py:224-226

_SYMBOLS = [
    "X", "H", "He",
    "Li", "Be", "B", "C", "N", "O", "F", "Ne",
    "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar",
    "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
    "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
    "In", "Sn", "Sb", "Te", "I", "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu",
    "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu",
    "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg",
    "Tl", "Pb", "Bi", "Po", "At", "Rn"]

_MASSES = [0.0, 1.008, 4.003, 6.941, 9.012, 10.812, 12.011, 14.007, 15.999,
           18.998, 20.180, 22.990, 24.305, 26.982, 28.086, 30.974, 32.066,
           35.453, 39.948]

# Cartesian components per shell type, x exponent descending then y
# descending (mmd/molecule.py:108-114).
_SHELLS = {
    "S": [(0, 0, 0)],
    "P": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "D": [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
    "F": [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2),
          (0, 3, 0), (0, 2, 1), (0, 1, 2), (0, 0, 3)],
}


_BASIS_FILES = {}


class Atom(object):
    """An atom: nuclear charge, mass (a.u.) and position (bohr)."""

    def __init__(self, charge, mass, origin=np.zeros(3)):
        self.charge = charge
        self.origin = origin
        self.mass = mass
        self.forces = np.zeros(3)
        self.saved_forces = np.zeros(3)
        self.velocities = np.zeros(3)


class Molecule(SCF, Forces):
    """A closed-shell molecule: geometry (Angstrom), basis set and integrals.

    Attributes after ``build()``: S, T, V, M, L, Core, X, U, nuc_energy and,
    for the conventional path, TwoE; for the direct path, screen.
    """

    def __init__(self, geometry, basis='sto-3g'):
        charge, multiplicity, atomlist = self.read_molecule(geometry)
        self.charge = charge
        self.multiplicity = multiplicity
        self.atoms = atomlist
        self.nelec = sum([atom.charge for atom in atomlist]) - charge
        self.nocc = self.nelec // 2
        self.is_built = False
        self.geometry_input = geometry
        self._ctx = None
        self.basis_name = str(basis).lower()
        basis_file = os.path.join(os.path.dirname(__file__), 'basis', self.basis_name + '.gbs')
        self.basis_data = self.getBasis(basis_file)
        self.formBasis()

    # -- native context ----------------------------------------------------------
    @property
    def native(self):
        """The libmmd context holding this molecule's basis on the GPU."""
        if self._ctx is None:
            ctx = _native.Context()
            ctx.set_basis_from_bfs(self.bfs)
            ctx.set_atoms([a.charge for a in self.atoms], np.array([a.origin for a in self.atoms]))
            _native.register_bfs_context(self.bfs, ctx)
            self._ctx = ctx
        return self._ctx

    def attach_context(self, ctx):
        """Use an existing libmmd context (e.g. one reused across a batch of
        molecules): this molecule's basis and nuclei replace its contents."""
        ctx.set_basis_from_bfs(self.bfs)
        ctx.set_atoms([a.charge for a in self.atoms], np.array([a.origin for a in self.atoms]))
        ctx._screen_key = None
        _native.register_bfs_context(self.bfs, ctx)
        self._ctx = ctx
        return ctx

    @property
    def _forces(self):
        return np.concatenate([atom.forces for atom in self.atoms]).reshape(-1, 3)

    # -- basis -------------------------------------------------------------------
    def formBasis(self):
        """Basis functions in the reference's order: atoms, then shells as
        listed in the basis file, then Cartesian components
        (mmd/molecule.py:57-88)."""
        self.bfs = []
        for atom in self.atoms:
            for momentum, prims in self.basis_data[atom.charge]:
                exps = [e for e, c in prims]
                coefs = [c for e, c in prims]
                for shell in self.momentum2shell(momentum):
                    self.bfs.append(Basis(np.asarray(atom.origin), np.asarray(shell), len(exps),
                                          np.asarray(exps), np.asarray(coefs)))
        self.nbasis = len(self.bfs)
        idx = 0
        for atom in self.atoms:
            atom.mask = np.zeros(self.nbasis)
            for momentum, prims in self.basis_data[atom.charge]:
                for shell in self.momentum2shell(momentum):
                    atom.mask[idx] = 1.0
                    idx += 1
        Z = np.array([atom.charge for atom in self.atoms], dtype=np.float64)
        R = np.array([atom.origin for atom in self.atoms], dtype=np.float64).reshape(-1, 3)
        self.center_of_charge = (Z[:, None] * R).sum(axis=0) / Z.sum()
        self._ctx = None

    def momentum2shell(self, momentum):
        return _SHELLS[str(momentum)]

    def sym2num(self, sym):
        return _SYMBOLS.index(str(sym))

    def getBasis(self, filename):
        """Parse a Gaussian-94 basis file: {Z: [(momentum, [(exp, coef), ...]), ...]}.
        SP shells become an S and a P entry (mmd/molecule.py:135-189).  A file
        is parsed once per process; the result is shared read-only."""
        key = os.path.abspath(filename)
        if key not in _BASIS_FILES:
            _BASIS_FILES[key] = self._parse_basis(filename)
        return _BASIS_FILES[key]

    def _parse_basis(self, filename):
        basis = {}
        with open(filename, 'r') as fh:
            data = fh.read().split('****')
        for block in data[1:]:
            rows = [x.split() for x in block.split('\n')[1:-1]]
            new_shell = True
            atom = None
            for idx, f in enumerate(rows):
                if not f:
                    continue
                if idx == 0:
                    assert len(f) == 2
                    atom = self.sym2num(f[0])
                    basis[atom] = []
                    new_shell = True
                elif new_shell:
                    momentum, nprims, count, prims, prims_p, new_shell = f[0], int(f[1]), 0, [], [], False
                else:
                    vals = [float(x.replace('D', 'E')) for x in f]
                    count += 1
                    if momentum == 'SP':
                        prims.append((vals[0], vals[1]))
                        prims_p.append((vals[0], vals[2]))
                        if count == nprims:
                            basis[atom].append(('S', prims))
                            basis[atom].append(('P', prims_p))
                            new_shell = True
                    else:
                        prims.append((vals[0], vals[1]))
                        if count == nprims:
                            basis[atom].append((momentum, prims))
                            new_shell = True
        return basis

    def read_molecule(self, geometry):
        """Charge, multiplicity and atoms from the geometry string; coordinates
        in Angstrom are converted to bohr (mmd/molecule.py:191-233)."""
        lines = filter(None, geometry.split('\n'))
        atomlist = []
        charge = multiplicity = None
        for n, line in enumerate(lines):
            f = line.split()
            if n == 0:
                assert len(f) == 2
                charge = int(f[0])
                multiplicity = int(f[1])
            else:
                if len(f) == 0:
                    break
                assert len(f) == 4
                Z = self.sym2num(str(f[0]))
                mass = _MASSES[Z] * 1822.8885 if Z < len(_MASSES) else 0.0
                xyz = np.asarray([float(f[1]), float(f[2]), float(f[3])]) / BOHR_ANGSTROM
                atomlist.append(Atom(charge=Z, mass=mass, origin=xyz))
        return charge, multiplicity, atomlist

    # -- integrals ---------------------------------------------------------------
    def build(self, direct=False):
        """One-electron integrals, then the Schwarz screen (direct) or the full
        ERI tensor (conventional) (mmd/molecule.py:90-102)."""
        self.one_electron_integrals()
        if direct:
            diag = self.native.schwarz()   # also installs the screen in the context
            self.screen = ScreenDict(diag)
            self.native._screen_key = (id(self.screen), len(self.screen))
        else:
            self.two_electron_integrals()
        self.is_built = True

    def one_electron_integrals(self):
        """S, T, V, dipole M, angular momentum L, nuclear repulsion, Core = T+V,
        X = S^-1/2 and U = S^1/2 (mmd/molecule.py:235-286)."""
        N = self.nbasis
        self.S, self.T, self.V, self.M, self.L = self.native.one_electron(self.center_of_charge)
        self.mu = np.zeros(3, dtype='complex')
        self.nuc_energy = 0.0
        for a, b in itertools.combinations(self.atoms, 2):
            self.nuc_energy += a.charge * b.charge / np.linalg.norm(a.origin - b.origin)
        self.Core = self.T + self.V
        # S^(-1/2) and S^(1/2), the principal powers of the symmetric positive
        # definite overlap (fractional_matrix_power in mmd/molecule.py:285-286),
        # on the GPU from the overlap still resident there
        self.X, self.U, _ = self.native.overlap_powers()
        assert self.S.shape == (N, N)

    def two_electron_integrals(self):
        """The full (ij|kl) tensor (mmd/molecule.py:288-293)."""
        N = self.nbasis
        self.TwoE = self.native.eri_tensor(np.zeros((N, N, N, N)))

    def save_integrals(self, folder=None):
        """Write enuc/nbf/nelec/s/t/v/eri .dat files and the geometry
        (mmd/molecule.py:295-315)."""
        if folder is None:
            sys.exit('Please provide a folder to save the integrals.')
        if not self.is_built:
            self.build()
        os.makedirs(folder, exist_ok=True)
        np.savetxt(folder + '/enuc.dat', np.asarray(self.nuc_energy).reshape(1,))
        np.savetxt(folder + '/nbf.dat', np.asarray(self.nbasis, dtype=int).reshape(1,), fmt='%d')
        np.savetxt(folder + '/nelec.dat', np.asarray(self.nelec, dtype=int).reshape(1,), fmt='%d')
        np.savetxt(folder + '/s.dat', self.S)
        np.savetxt(folder + '/t.dat', self.T)
        np.savetxt(folder + '/v.dat', self.V)
        with open(folder + '/eri.dat', 'w') as f:
            for i, j, k, l in itertools.product(range(self.nbasis), repeat=4):
                print(i + 1, j + 1, k + 1, l + 1, self.TwoE[i, j, k, l], file=f)
        with open(folder + '/geometry.txt', 'w') as f:
            print(self.geometry_input, file=f)


class ScreenDict(dict):
    """The reference's screen dictionary {p(p+1)/2+q: (pq|pq)} (a real dict),
    carrying the full symmetric matrix it was made from in ``matrix`` so that
    the Fock build does not rebuild it."""

    def __init__(self, diag):
        N = diag.shape[0]
        p, q = np.tril_indices(N)
        super().__init__(zip((p * (p + 1) // 2 + q).tolist(), diag[p, q].tolist()))
        self.matrix = diag
