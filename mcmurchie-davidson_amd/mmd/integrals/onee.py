"""One-electron integrals of ``Basis`` pairs: S, T, V, Mu, RxDel and the
Boys-function test hook ``_boys`` (cython/onee.pyx:14-69, 195-196),
evaluated by libmmd on the GPU."""
import numpy as np

from mmd import _native
from mmd.integrals.twoe import Basis


def _pair(a, b, C=None, charge=0.0):
    for x in (a, b):
        if not isinstance(x, Basis):
            raise TypeError("Argument has incorrect type (expected Basis, got %s)" % type(x).__name__)
    ctx = _native.default_context()
    # order [b, a]: element [1, 0] is the integral with `a` on the left
    ctx.set_basis_from_bfs([b, a])
    center = np.zeros(3) if C is None else np.asarray(C, dtype=np.float64).reshape(3)
    ctx.set_atoms(np.array([charge]), center.reshape(1, 3))
    return ctx.one_electron(center)


def S(a, b):
    """Overlap (cython/onee.pyx:14-22)."""
    return float(_pair(a, b)[0][1, 0])


def T(a, b):
    """Kinetic energy (cython/onee.pyx:48-56)."""
    return float(_pair(a, b)[1][1, 0])


def V(a, b, C):
    """Nuclear attraction of a unit positive charge at C, without the -Z
    factor that Molecule applies (cython/onee.pyx:61-69)."""
    return float(_pair(a, b, C, charge=-1.0)[2][1, 0])


_DIR = {"x": 0, "y": 1, "z": 2}


def Mu(a, b, C, direction):
    """Dipole integral about origin C along 'x', 'y' or 'z' (cython/onee.pyx:27-34)."""
    if not isinstance(direction, str):
        raise TypeError("Argument 'direction' has incorrect type (expected str)")
    d = _DIR.get(direction.lower())
    if d is None:
        return 0.0
    return float(_pair(a, b, C)[3][d, 1, 0])


def RxDel(a, b, C, direction):
    """Angular-momentum integral (cython/onee.pyx:36-43)."""
    d = _DIR.get(str(direction).lower())
    if d is None:
        return None
    return float(_pair(a, b, C)[4][d, 1, 0])


def _boys(n, T):
    """F_n(T) for real n >= 0 (cython/onee.pyx:195-196)."""
    return float(_native.default_context().boys(float(n), float(T), table=False)[0])
