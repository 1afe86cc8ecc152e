"""Post-SCF entry points on the stored ERI tensor: ``PostSCF(mol).ao2mo()``
and ``MP2()`` with the API of the reference's mmd/postscf.py:13-71.

The four-index transformation runs on the GPU (libmmd ``mmd_ao2mo``: four
FP64 GEMM quarter transforms); the spin-orbital blocks and the MP2 sums are
host numpy, as in the reference.  Configuration interaction and the other
post-SCF methods of the reference are not part of this package.
"""
import sys

import numpy as np


class PostSCF(object):
    """Post-SCF quantities of a converged conventional RHF molecule
    (``mol.RHF(direct=False)``, which stores ``mol.TwoE``)."""

    def __init__(self, mol):
        self.mol = mol
        if not self.mol.is_converged:
            sys.exit("SCF not converged, skipping Post-SCF")
        self.ao2mo()

    def ao2mo(self):
        """MO-basis integrals (mmd/postscf.py:21-43): ``single_bar`` (PQ|RS)
        in chemists' notation; spin-orbital ``double_bar`` <pq||rs>, core
        Hamiltonian ``Hp`` and orbital energies ``fs`` (spin orbital 2i + s
        has spatial orbital i and spin s)."""
        mol = self.mol
        if not hasattr(mol, "TwoE"):
            raise AttributeError("PostSCF needs the stored ERI tensor: run RHF(direct=False) first")
        C = np.asarray(mol.C)
        if np.iscomplexobj(C):
            if np.any(C.imag):
                raise NotImplementedError("complex MO coefficients are not supported by the GPU transform")
            C = C.real
        mol.single_bar = mol.native.ao2mo(np.asarray(mol.TwoE), C)
        mol.norb = mol.nbasis * 2
        spat = np.arange(mol.norb) // 2
        spin = np.arange(mol.norb) % 2
        same = (spin[:, None] == spin[None, :])
        # (pq|rs) over spin orbitals: spatial integral times spin delta(p,q) delta(r,s)
        sb = mol.single_bar[np.ix_(spat, spat, spat, spat)] * same[:, :, None, None] * same[None, None, :, :]
        # <pq||rs> = (pr|qs) - (ps|qr)
        mol.double_bar = sb.transpose(0, 2, 1, 3) - sb.transpose(0, 2, 3, 1)
        h = np.dot(C.T, np.dot(np.real(mol.Core), C))
        mol.Hp = h[np.ix_(spat, spat)] * same
        mol.fs = np.diag(np.asarray(mol.MO)[spat])

    def MP2(self, spin_orbital=False):
        """Second-order Moller-Plesset energy from the RHF reference
        (mmd/postscf.py:45-71); sets ``mol.emp2`` (total energy)."""
        mol = self.mol
        if spin_orbital:
            o = mol.nelec
            e = np.diag(mol.fs)
            v = np.arange(o, mol.norb)
            denom = e[:o, None, None, None] + e[None, :o, None, None] - e[None, None, v, None] - e[None, None, None, v]
            numer = mol.double_bar[:o, :o, o:, o:] ** 2
            mol.emp2 = 0.25 * np.sum(numer / denom) + mol.energy
        else:
            o = mol.nocc
            e = np.asarray(mol.MO)
            iajb = mol.single_bar[:o, o:, :o, o:]
            ibja = iajb.transpose(0, 3, 2, 1)
            denom = e[:o, None, None, None] - e[None, o:, None, None] + e[None, None, :o, None] - e[None, None, None, o:]
            mol.emp2 = np.sum(iajb * (2.0 * iajb - ibja) / denom) + mol.energy
        print('E(MP2) = ', mol.emp2.real)
