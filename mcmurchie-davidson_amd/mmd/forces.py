"""Nuclear gradient of the converged RHF energy (mmd/forces.py:8-97 of the
reference), with the derivative integrals on the GPU.

``Forces.forces()`` sets ``atom.forces`` (= -dE/dX, hartree/bohr, not
mass-weighted) for every atom, as the reference does, and ``_forces`` lists
them.  The electronic part

    sum_pq P_pq 2 dH_pq/dX + sum_pqrs Gamma_pqrs d(pq|rs)/dX - 2 sum_pq W_pq dS_pq/dX

(P = C_occ C_occ^T, W = P F P, Gamma the density products of
forces.py:62-84) comes from libmmd's mmd_rhf_gradient in one pass over all
atoms and directions; the nuclear repulsion term is added here
(forces.py:52-60).
"""
import numpy as np


def nuclear_repulsion_gradient(atoms):
    """d(sum_A<B Z_A Z_B / R_AB)/dX, shape (natom, 3) (forces.py:52-60)."""
    g = np.zeros((len(atoms), 3))
    for a, atom in enumerate(atoms):
        for other in atoms:
            d = np.asarray(atom.origin, dtype=float) - np.asarray(other.origin, dtype=float)
            r = np.linalg.norm(d)
            if not np.allclose(r, 0.0):
                g[a] += -d * atom.charge * other.charge / r ** 3
    return g


def basis_atoms(mol):
    """The atom of each basis function (the reference's atom.mask)."""
    bf_atom = np.zeros(mol.nbasis, dtype=np.int32)
    for ia, atom in enumerate(mol.atoms):
        bf_atom[np.asarray(atom.mask) > 0] = ia
    return bf_atom


class Forces(object):
    """Nuclear gradient methods of the molecule object."""

    def forces(self):
        """Compute the nuclear forces (mmd/forces.py:8-97)."""
        if not getattr(self, "is_converged", False):
            raise RuntimeError("Need to converge SCF before computing gradient")
        P = np.real(self.P)
        W = P @ np.real(self.F) @ P
        ctx = self.native
        g = ctx.rhf_gradient(P, W, basis_atoms(self)) + nuclear_repulsion_gradient(self.atoms)
        self.gradient_ms = ctx.last_gradient_ms()
        for a, atom in enumerate(self.atoms):
            atom.forces = -g[a]
        return self._forces
