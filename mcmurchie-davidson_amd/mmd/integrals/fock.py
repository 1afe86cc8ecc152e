"""Direct incremental Fock build: ``formPT`` (cython/fock.pyx:10-87).

The unsymmetrised G(dP) of the reference, dP = P - P_old, with the same
Schwarz x density screening per basis-function quartet
(cython/fock.pyx:46-57), degeneracy weights (cython/fock.pyx:60-70) and
Coulomb / exchange updates (cython/fock.pyx:79-85), evaluated on the GPU.
"""
import numpy as np

from mmd import _native


def screen_to_matrix(screen, N):
    """The reference's screen dictionary {p(p+1)/2+q: (pq|pq)} as an N x N
    symmetric matrix."""
    if isinstance(screen, np.ndarray) and screen.shape == (N, N):
        return np.ascontiguousarray(screen, dtype=np.float64)
    vals = getattr(screen, "matrix", None)
    if vals is not None:
        return np.ascontiguousarray(vals, dtype=np.float64)
    M = np.empty((N, N))
    p, q = np.tril_indices(N)
    pq = p * (p + 1) // 2 + q
    v = np.fromiter((screen[int(k)] for k in pq), dtype=np.float64, count=len(pq))
    M[p, q] = v
    M[q, p] = v
    return M


def formPT(P, P_old, bfs, nbasis, screen, tol):
    """G(P - P_old) before symmetrisation, complex128 [N, N]."""
    for name, x in (("P", P), ("P_old", P_old)):
        if not isinstance(x, np.ndarray):
            raise TypeError("Argument '%s' has incorrect type (expected numpy.ndarray)" % name)
        if x.dtype != np.complex128:
            raise ValueError("Buffer dtype mismatch, expected 'complex' but got '%s'" % x.dtype)
        if x.ndim != 2:
            raise ValueError("Buffer has wrong number of dimensions (expected 2, got %d)" % x.ndim)
    if not isinstance(bfs, list):
        raise TypeError("Argument 'bfs' has incorrect type (expected list, got %s)" % type(bfs).__name__)
    if not isinstance(screen, dict):
        raise TypeError("Argument 'screen' has incorrect type (expected dict, got %s)" % type(screen).__name__)
    N = int(nbasis)
    dP = P - P_old
    ctx = _native.context_for_bfs(bfs)
    key = id(screen)
    if getattr(ctx, "_screen_key", None) != (key, len(screen)):
        ctx.set_schwarz(screen_to_matrix(screen, N))
        ctx._screen_key = (key, len(screen))
    dPi = dP.imag if np.any(dP.imag != 0.0) else None
    Gr, Gi, _ = ctx.fock_delta(dP.real, dPi, tol=float(tol), raw=True)
    G = Gr.astype(np.complex128)
    if Gi is not None:
        G += 1j * Gi
    return G
