"""Restricted Hartree-Fock driver with the API of the reference's
mmd/scf.py: ``RHF``, ``buildFock``, DIIS and the helpers used by the
real-time code.

The two-electron part of every Fock build runs in libmmd: the direct path
computes the symmetrised incremental G(P - P_old) with the GPU Fock kernels
(formPT semantics, cython/fock.pyx:10-87), the conventional path contracts
the stored tensor on the GPU (mmd/scf.py:97-98).
"""
import contextlib

import numpy as np
import scipy.linalg
import scipy.linalg.lapack

from mmd.integrals.fock import formPT

try:  # BLAS thread counts are chosen per operation (see SCF.RHF)
    from threadpoolctl import ThreadpoolController as _Controller
except ImportError:  # pragma: no cover
    _Controller = None
_ctl = None


def _blas_threads(n):
    """Context that limits the BLAS/LAPACK thread pool to n threads (None:
    leave it as it is)."""
    global _ctl
    if _Controller is None or n is None:
        return contextlib.nullcontext()
    if _ctl is None:
        _ctl = _Controller()
    return _ctl.limit(limits=n, user_api="blas")


def _thread_plan(N):
    """(threads for the matrix products, threads for the eigensolver): the
    products of N >= 128 use 4 threads; the divide-and-conquer eigensolver
    is no faster on more threads at N <= 400 (and slower at N ~ 200), so it
    runs on one."""
    return (1 if N < 128 else 4), 1


_EVD_LWORK = {}


def _eigh(A):
    """Eigenvalues (ascending) and eigenvectors of a symmetric matrix by
    LAPACK's divide and conquer, as scipy.linalg.eigh(A, driver="evd")
    computes them (same routine, lower triangle, same results, and the same
    ValueError for non-finite input), without the workspace query on every
    call; complex input goes to scipy.linalg.eigh."""
    A = np.asarray(A)
    if A.dtype != np.float64 or A.ndim != 2:
        return scipy.linalg.eigh(A, driver="evd")
    if not np.isfinite(A).all():
        raise ValueError("array must not contain infs or NaNs")
    n = A.shape[0]
    lw = _EVD_LWORK.get(n)
    if lw is None:
        lw = _EVD_LWORK[n] = scipy.linalg.lapack.dsyevd_lwork(n, compute_v=1, lower=1)[:2]
    w, v, info = scipy.linalg.lapack.dsyevd(A, compute_v=1, lower=1, lwork=int(lw[0]), liwork=int(lw[1]))
    if info != 0:
        raise np.linalg.LinAlgError("dsyevd failed (info = %d)" % info)
    return w, v


def _re(A):
    """A itself, or its real part when the imaginary part is identically zero
    (the closed-shell SCF is real; arithmetic then runs in real FP64)."""
    A = np.asarray(A)
    if np.iscomplexobj(A) and not np.any(A.imag):
        return np.ascontiguousarray(A.real)
    return A


class SCF(object):
    """SCF methods and routines for a Molecule."""

    def RHF(self, doPrint=True, DIIS=True, direct=False, conver=1e-8, acc2e=1e-12, guess='core'):
        """Closed-shell SCF (mmd/scf.py:10-80).  Converged when
        ||P - P_old||_F < conver; at most 100 iterations.

        guess='core' starts from the core Hamiltonian as the reference does;
        guess='sad' starts from the superposition of atomic densities
        (mmd.guess.sad_density): F = H + G(P_sad), and the first incremental
        build is G(P_1 - P_sad); guess='learned' starts from SAD corrected
        per water monomer by the trained perceptron (mmd.learned);
        guess='pair' from SAD with the atom-pair blocks of the atom-pair
        networks (mmd.pairguess)."""
        if guess not in ('core', 'sad', 'learned', 'pair'):
            raise ValueError("guess must be 'core', 'sad', 'learned' or 'pair'")
        gemm_threads, self._eigh_threads = _thread_plan(self.nbasis)
        try:
            with _blas_threads(gemm_threads):
                return self._rhf(doPrint, DIIS, direct, conver, acc2e, guess)
        finally:
            self._eigh_threads = None

    # Direct SCFs with at least this many basis functions keep H, S, X, F, P
    # and the DIIS window in device memory between Fock builds
    # (mmd_scf_* of libmmd); None runs every step on the host.
    device_loop_min_nbasis = 100

    def _rhf(self, doPrint, DIIS, direct, conver, acc2e, guess='core'):
        self._rhf_setup(DIIS, direct, acc2e, guess)
        if (direct and self.device_loop_min_nbasis is not None and self.nbasis >= self.device_loop_min_nbasis
                and not np.any(np.imag(self.F)) and not np.any(np.imag(self.P))):
            return self._rhf_loop_device(doPrint, DIIS, conver)
        return self._rhf_loop(doPrint, DIIS, conver)

    def _rhf_setup(self, DIIS, direct, acc2e, guess):
        self.is_converged = False
        self.delta_energy = 1e20
        self.P_RMS = 1e20
        self.P_old = np.zeros((self.nbasis, self.nbasis), dtype='complex')
        self.maxiter = 100
        self.direct = direct
        if self.direct:
            self.incFockRst = False
        self.scrTol = acc2e
        self.fock_ms = []
        self.build(self.direct)

        self.P = self.P_old
        self.F = self.Core.astype('complex')
        if guess in ('sad', 'learned', 'pair'):
            if guess == 'sad':
                from mmd.guess import sad_density
                self.P = sad_density(self).astype('complex')
            elif guess == 'pair':
                from mmd.pairguess import pair_density
                self.P = pair_density(self, getattr(self, 'pair_model', None)).astype('complex')
            else:
                from mmd.learned import learned_density
                self.P = learned_density(self, getattr(self, 'guess_model', None)).astype('complex')
            self.F_old = self.F
            self.buildFock()          # F = H + G(P_sad)
            self.P_old = self.P

        if DIIS:
            self.fockSet = []
            self.errorSet = []

    def _rhf_loop(self, doPrint, DIIS, conver):
        for step in range(self.maxiter):
            if step > 0:
                self.F_old = self.F
                energy_old = self.energy
                self.buildFock()
                self.P_old = self.P
                if DIIS:
                    F_diis = _re(self.updateDIIS(self.F, self.P))
                    self.FO = np.dot(self.X.T, np.dot(F_diis, self.X))
            if not DIIS or step == 0:
                self.orthoFock()

            with _blas_threads(getattr(self, "_eigh_threads", None)):
                E, self.CO = _eigh(_re(self.FO))
            C = np.dot(self.X, self.CO)
            self.C = C
            self.MO = E
            Cocc = C[:, :self.nocc]
            self.P = np.dot(Cocc, np.conjugate(Cocc).T).astype('complex')
            self.computeEnergy()

            if step > 0:
                self.delta_energy = self.energy - energy_old
                self.P_RMS = np.linalg.norm(self.P - self.P_old)
            if np.abs(self.P_RMS) < conver or step == (self.maxiter - 1):
                if step == (self.maxiter - 1):
                    print("NOT CONVERGED")
                else:
                    self.is_converged = True
                    FPS = np.dot(self.F, np.dot(self.P, self.S))
                    error = FPS - self.adj(FPS)
                    self.computeDipole()
                    if doPrint:
                        self._print_converged(step, error)
                    break
        self.iterations = step

    def _print_converged(self, step, error):
        print("E(SCF)    = ", "{0:.12f}".format(self.energy.real) + " in " + str(step)
              + " iterations")
        print("  Convergence:")
        print("    FPS-SPF  = ", np.linalg.norm(error))
        print("    RMS(P)   = ", "{0:.2e}".format(self.P_RMS.real))
        print("    dE(SCF)  = ", "{0:.2e}".format(self.delta_energy.real))
        print("  Dipole X = ", "{0:.8f}".format(self.mu[0].real))
        print("  Dipole Y = ", "{0:.8f}".format(self.mu[1].real))
        print("  Dipole Z = ", "{0:.8f}".format(self.mu[2].real))

    def _rhf_loop_device(self, doPrint, DIIS, conver):
        """The loop of _rhf_loop with the N x N matrices in device memory
        (mmd_scf_*): Fock build, F update, DIIS and the X^T F X products on
        the GPU; the orthogonalised Fock matrix comes to the host for the same
        LAPACK eigensolver and its eigenvectors go back.  Same iteration
        semantics as the host loop (mmd/scf.py:10-80).  After the loop, F, P,
        P_old, G, C, CO, FO and MO are set as the host loop sets them; the DIIS
        window stays on the device (fockSet / errorSet are left empty)."""
        ctx = self.native
        N = self.nbasis
        from mmd.integrals.fock import screen_to_matrix
        if getattr(ctx, "_screen_key", None) != (id(self.screen), len(self.screen)):
            ctx.set_schwarz(screen_to_matrix(self.screen, N))
            ctx._screen_key = (id(self.screen), len(self.screen))
        ctx.scf_begin(self.nocc, self.Core, self.S, self.X, self.F, self.P_old)
        FO = np.empty((N, N))
        for step in range(self.maxiter):
            if step > 0:
                energy_old = self.energy
                ctx.scf_step(self.scrTol, True, DIIS, FO)
                self.fock_ms.append(ctx.last_fock_ms())
            else:
                ctx.scf_step(self.scrTol, False, False, FO)
            with _blas_threads(getattr(self, "_eigh_threads", None)):
                E, CO = _eigh(FO)
            e_el, rms = ctx.scf_density(CO)
            self.el_energy = np.complex128(e_el)
            self.energy = self.el_energy + self.nuc_energy
            if step > 0:
                self.delta_energy = self.energy - energy_old
                self.P_RMS = np.float64(rms)
            if np.abs(self.P_RMS) < conver or step == (self.maxiter - 1):
                break
        F, P, P_old, G = ctx.scf_get()
        self.F, self.P, self.P_old, self.G = (a.astype('complex') for a in (F, P, P_old, G))
        self.FO, self.CO, self.MO = FO.copy(), CO, E
        self.C = np.dot(self.X, CO)
        if step == (self.maxiter - 1):
            print("NOT CONVERGED")
        else:
            self.is_converged = True
            FPS = np.dot(self.F, np.dot(self.P, self.S))
            error = FPS - self.adj(FPS)
            self.computeDipole()
            if doPrint:
                self._print_converged(step, error)
        self.iterations = step

    def _direct_G(self, P, P_old):
        """Symmetrised G(P - P_old).  A real difference takes the GPU's
        symmetrised build; a complex one takes formPT's exact form."""
        dP = P - P_old
        if np.iscomplexobj(dP) and np.any(dP.imag != 0.0):
            G = formPT(np.asarray(P, dtype=complex), np.asarray(P_old, dtype=complex), self.bfs,
                       self.nbasis, self.screen, self.scrTol)
            return 0.5 * (G + G.T)
        ctx = self.native
        if getattr(ctx, "_screen_key", None) != (id(self.screen), len(self.screen)):
            from mmd.integrals.fock import screen_to_matrix
            ctx.set_schwarz(screen_to_matrix(self.screen, self.nbasis))
            ctx._screen_key = (id(self.screen), len(self.screen))
        G, _, _ = ctx.fock_delta(np.real(dP), None, tol=self.scrTol, raw=False)
        self.fock_ms.append(ctx.last_fock_ms())
        return G.astype('complex')

    def buildFock(self):
        """Fock matrix from the current density (mmd/scf.py:82-100)."""
        if self.direct:
            if self.incFockRst:
                self.G = self._direct_G(self.P, np.zeros_like(self.P))
                self.F = self.Core.astype('complex') + self.G
            else:
                self.G = self._direct_G(self.P, self.P_old)
                self.F = self.F_old + self.G
        else:
            P = np.asarray(self.P)
            Pim = P.imag if np.iscomplexobj(P) and np.any(P.imag != 0.0) else None
            Jr, Ji, Kr, Ki = self.native.jk_tensor(self.TwoE, np.real(P), Pim)
            self.J = Jr.astype('complex')
            self.K = Kr.astype('complex')
            if Pim is not None:
                self.J = self.J + 1j * Ji
                self.K = self.K + 1j * Ki
            self.G = 2. * self.J - self.K
            self.F = self.Core.astype('complex') + self.G

    def orthoFock(self):
        self.FO = np.dot(self.X.T, np.dot(_re(self.F), self.X))

    def unOrthoFock(self):
        self.F = np.dot(self.U.T, np.dot(self.FO, self.U))

    def orthoDen(self):
        self.PO = np.dot(self.U, np.dot(self.P, self.U.T))

    def unOrthoDen(self):
        self.P = np.dot(self.X, np.dot(self.PO, self.X.T))

    def computeEnergy(self):
        """E = sum_pq (H + F)_pq P_qp + E_nuc (mmd/scf.py:118-121)."""
        self.el_energy = np.einsum('pq,qp', self.Core + self.F, self.P)
        self.energy = self.el_energy + self.nuc_energy

    def computeDipole(self):
        """Dipole moment in Debye about the centre of nuclear charge
        (mmd/scf.py:123-129)."""
        self.el_energy = np.einsum('pq,qp', self.Core + self.F, self.P)
        for i in range(3):
            self.mu[i] = -2 * np.trace(np.dot(self.P, self.M[i])) + sum(
                [atom.charge * (atom.origin[i] - self.center_of_charge[i]) for atom in self.atoms])
        self.mu *= 2.541765

    def adj(self, x):
        assert x.shape[0] == x.shape[1]
        return np.conjugate(x).T

    def comm(self, A, B):
        return np.dot(A, B) - np.dot(B, A)

    def updateFock(self):
        self.unOrthoDen()
        self.buildFock()
        self.orthoFock()

    def updateDIIS(self, F, P):
        """Pulay DIIS on the orthogonalised FPS - SPF, subspace of 8
        (mmd/scf.py:146-180).  B_ij = tr(e_i^H e_j) is evaluated as the
        elementwise inner product (same value, O(N^2))."""
        FPS = np.dot(_re(F), np.dot(_re(P), self.S))
        SPF = self.adj(FPS)
        error = np.dot(self.X, np.dot(FPS - SPF, self.X))
        self.fockSet.append(self.F)
        self.errorSet.append(error)
        numFock = len(self.fockSet)
        if numFock > 8:
            del self.fockSet[0]
            del self.errorSet[0]
            numFock -= 1
        B = np.zeros((numFock + 1, numFock + 1))
        B[-1, :] = B[:, -1] = -1.0
        B[-1, -1] = 0.0
        for i in range(numFock):
            for j in range(i + 1):
                B[i, j] = B[j, i] = np.real(np.vdot(self.errorSet[i], self.errorSet[j]))
        residual = np.zeros(numFock + 1)
        residual[-1] = -1.0
        weights = np.linalg.solve(B, residual)
        assert np.isclose(sum(weights[:-1]), 1.0)
        F = np.zeros((self.nbasis, self.nbasis), dtype='complex')
        for i, Fock in enumerate(self.fockSet):
            F += weights[i] * Fock
        return F
