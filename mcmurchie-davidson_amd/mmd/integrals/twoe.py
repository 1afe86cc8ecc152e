"""Two-electron integrals: ``Basis``, ``ERI`` and ``doERIs``.

Drop-in for the reference's ``mmd.integrals.twoe`` extension module
(cython/twoe.pyx, cython/basis.pxi); the integrals are evaluated by libmmd on
the GPU.
"""
import numpy as np

from mmd import _native


def _fact2(n):
    """Double factorial n!! with (-1)!! = 1."""
    r = 1.0
    k = int(n)
    while k > 1:
        r *= k
        k -= 2
    return r


# normalised (norm, coefs) per (shell, exponents, coefficients): the same
# contracted function recurs on every atom of an element
_NORMALIZED = {}


class Basis(object):
    """One Cartesian component of a contracted Gaussian (cython/basis.pxi:6-85).

    ``coefs`` are rescaled in place by the contraction normalisation and
    ``norm`` holds the primitive norms (cython/basis.pxi:87-120), so that
    ``norm[k]*coefs[k]`` is the weight of primitive k.
    """

    __slots__ = ("_origin", "_shell", "_num_exps", "_exps", "_coefs", "_norm", "__weakref__")

    def __init__(self, origin, shell, num_exps, exps, coefs):
        self._origin = np.array([float(origin[i]) for i in range(3)], dtype=np.float64)
        self._shell = np.array([int(shell[i]) for i in range(3)], dtype=np.int64)
        self._num_exps = int(num_exps)
        self._exps = np.array(exps[:self._num_exps], dtype=np.float64)
        self._coefs = np.array(coefs[:self._num_exps], dtype=np.float64)
        if self._exps.shape != (self._num_exps,) or self._coefs.shape != (self._num_exps,):
            raise ValueError("exps and coefs need num_exps entries")
        key = (tuple(self._shell.tolist()), self._exps.tobytes(), self._coefs.tobytes())
        cached = _NORMALIZED.get(key)
        if cached is None:
            self._norm = np.zeros(self._num_exps, dtype=np.float64)
            self.normalize()
            _NORMALIZED[key] = (self._norm.copy(), self._coefs.copy())
        else:
            self._norm = cached[0].copy()
            self._coefs = cached[1].copy()

    origin = property(lambda self: self._origin)
    shell = property(lambda self: self._shell)
    num_exps = property(lambda self: self._num_exps)
    exps = property(lambda self: self._exps)
    coefs = property(lambda self: self._coefs)
    norm = property(lambda self: self._norm)

    def normalize(self):
        """Primitive norms, then the contraction normalisation (Valeev,
        eq. 1.44), as cython/basis.pxi:87-120."""
        l, m, n = (int(x) for x in self._shell)
        L = l + m + n
        df = _fact2(2 * l - 1) * _fact2(2 * m - 1) * _fact2(2 * n - 1)
        a = self._exps
        self._norm[:] = np.sqrt(np.power(2.0, 2 * L + 1.5) * np.power(a, L + 1.5) / df / np.power(np.pi, 1.5))
        pref = np.power(np.pi, 1.5) * df / np.power(2.0, L)
        w = self._norm * self._coefs
        N = pref * np.sum(np.outer(w, w) / np.power(a[:, None] + a[None, :], L + 1.5))
        self._coefs *= np.power(N, -0.5)


def ERI(a, b, c, d):
    """Contracted (ab|cd) of four ``Basis`` objects (cython/twoe.pyx:34-50)."""
    for x in (a, b, c, d):
        if not isinstance(x, Basis):
            raise TypeError("Argument has incorrect type (expected Basis, got %s)" % type(x).__name__)
    pc = _native.per_call_basis()
    with pc.lock:
        idx = pc.indices([a, b, c, d])
        return float(pc.ctx.eri_list(np.array([idx], dtype=np.int64))[0])


def doERIs(N, TwoE, bfs):
    """Fill the caller's [N,N,N,N] float64 buffer with (ij|kl) using the
    8-fold symmetry (cython/twoe.pyx:9-31); returns a memoryview of it."""
    if not isinstance(bfs, list):
        raise TypeError("Argument 'bfs' has incorrect type (expected list, got %s)" % type(bfs).__name__)
    T = np.asarray(TwoE)
    if T.dtype != np.float64:
        raise ValueError("Buffer dtype mismatch, expected 'double' but got '%s'" % T.dtype)
    if T.ndim != 4:
        raise ValueError("Buffer has wrong number of dimensions (expected 4, got %d)" % T.ndim)
    if not T.flags.writeable:
        raise ValueError("buffer source array is read-only")
    N = int(N)
    if T.shape != (N, N, N, N) or len(bfs) != N:
        raise ValueError("TwoE must have shape (N,N,N,N) with N = len(bfs)")
    ctx = _native.context_for_bfs(bfs)
    if T.flags.c_contiguous:
        ctx.eri_tensor(out=T)
    else:
        T[...] = ctx.eri_tensor()
    return memoryview(T)
