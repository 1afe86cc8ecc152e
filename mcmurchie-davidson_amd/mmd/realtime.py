"""Real-time TDHF propagation with the API of the reference's mmd/realtime.py
(``RealTime(mol, numsteps, stepsize, field, pulse)``, ``Magnus2``,
``Magnus4``, ``addField``).

Each step rebuilds the Fock matrix of a complex density through
``SCF.updateFock`` (the conventional J/K contraction of the stored tensor on
the GPU, real and imaginary parts).  The integrators are the explicit second-
and fourth-order Magnus methods of Blanes and Casas, *A Concise Introduction
to Geometric Numerical Integration* (CRC Press, 2016), Eqs. (4.61) and
(4.62); the orthonormal-basis density is propagated as U PO U^dagger.
"""
import numpy as np
from scipy.linalg import expm


class RealTime(object):
    """Real-time propagation of a converged RHF molecule (mmd/realtime.py:6-31)."""

    def __init__(self, mol, numsteps=1000, stepsize=0.1, field=0.0001, pulse=None):
        self.mol = mol
        self.field = field
        self.stepsize = stepsize
        self.numSteps = numsteps
        self.time = np.arange(0, self.numSteps) * self.stepsize
        self.pulse = pulse if pulse else (lambda t: 0.0)
        self.reset()

    def reset(self):
        """Ground-state SCF again and empty property records."""
        self.mol.RHF(doPrint=False)
        self.dipole = []
        self.angmom = []
        self.Energy = []
        self.shape = []

    # -- helpers -------------------------------------------------------------
    def _record(self, time, direction):
        axis = {'x': 0, 'y': 1, 'z': 2}.get(direction.lower())
        if axis is not None:
            self.mol.computeDipole()
            self.dipole.append(np.real(self.mol.mu[axis]))
        self.shape.append(self.pulse(time))

    def _evolve(self, rho0, L):
        """PO = exp(L) rho0 exp(L)^dagger, then the Fock matrix of that density."""
        U = expm(L)
        self.mol.PO = np.dot(U, np.dot(rho0, self.mol.adj(U)))
        self.mol.updateFock()

    def _finish_step(self):
        self.mol.unOrthoFock()
        self.mol.unOrthoDen()
        self.mol.computeEnergy()
        self.Energy.append(np.real(self.mol.energy))

    # -- integrators -----------------------------------------------------------
    def Magnus2(self, direction='x'):
        """Second-order explicit Magnus (Blanes & Casas Eq. 4.61): a predictor
        step with the Fock matrix at t, then the midpoint average of the
        generators at t and t + dt."""
        self.reset()
        self.mol.orthoDen()
        self.mol.orthoFock()
        h = -1j * self.stepsize
        for t in self.time:
            self._record(t, direction)
            rho0 = np.copy(self.mol.PO)
            self.addField(t, direction=direction)
            k1 = h * self.mol.FO
            self._evolve(rho0, k1)
            self.addField(t + self.stepsize, direction=direction)
            self._evolve(rho0, 0.5 * (k1 + h * self.mol.FO))
            self._finish_step()

    def Magnus4(self, direction='x'):
        """Fourth-order explicit Magnus (Blanes & Casas Eq. 4.62): six Fock
        evaluations per step at t, t + dt/2 and t + dt, with the commutator
        corrections of the scheme."""
        self.reset()
        self.mol.orthoDen()
        self.mol.orthoFock()
        h = -1j * self.stepsize
        dt = self.stepsize
        comm = self.mol.comm
        for t in self.time:
            self._record(t, direction)
            rho0 = np.copy(self.mol.PO)
            self.addField(t, direction=direction)
            k1 = h * self.mol.FO
            Q1 = k1
            self._evolve(rho0, 0.5 * Q1)
            self.addField(t + 0.5 * dt, direction=direction)
            k2 = h * self.mol.FO
            Q2 = k2 - k1
            self._evolve(rho0, 0.5 * Q1 + 0.25 * Q2)
            self.addField(t + 0.5 * dt, direction=direction)
            k3 = h * self.mol.FO
            Q3 = k3 - k2
            self._evolve(rho0, Q1 + Q2)
            self.addField(t + dt, direction=direction)
            k4 = h * self.mol.FO
            Q4 = k4 - 2 * k2 + k1
            self._evolve(rho0, 0.5 * Q1 + 0.25 * Q2 + Q3 / 3.0 - Q4 / 24.0 - comm(Q1, Q2) / 48.0)
            self.addField(t + 0.5 * dt, direction=direction)
            k5 = h * self.mol.FO
            Q5 = k5 - k2
            self._evolve(rho0, Q1 + Q2 + (2.0 / 3.0) * Q3 + Q4 / 6.0 - comm(Q1, Q2) / 6.0)
            self.addField(t + dt, direction=direction)
            k6 = h * self.mol.FO
            Q6 = k6 - 2 * k2 + k1
            self._evolve(rho0, Q1 + Q2 + (2.0 / 3.0) * Q5 + Q6 / 6.0 - comm(Q1, Q2 - Q3 + Q5 + 0.5 * Q6) / 6.0)
            self._finish_step()

    def addField(self, time, direction='x'):
        """F += field * pulse(time) * dipole integrals along ``direction``,
        then the orthonormal-basis Fock matrix (mmd/realtime.py:154-186)."""
        scale = self.field * self.pulse(time)
        axis = {'x': 0, 'y': 1, 'z': 2}.get(direction.lower())
        if axis is not None:
            self.mol.F = self.mol.F + scale * self.mol.M[axis]
        self.mol.orthoFock()
