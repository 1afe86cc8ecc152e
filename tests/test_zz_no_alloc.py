"""mmd_fock_delta_noalloc: the symmetrised Fock build with every work buffer
in the library's static per-device scratch (no allocation in the call),
against mmd_fock_delta on the same inputs, through the C ABI.

Two calls in a row on one GPU, the second with a smaller basis (the scratch
keeps the first call's contents, so what must start at zero is cleared by
the call), and the error for a build whose survivor lists exceed the scratch.
Tolerance: 1e-12 absolute (the builds add into G with atomics, so the two
entry points agree to rounding, not bit for bit).
"""
import ctypes

import numpy as np
import pytest

from geometries import water_cluster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from mmd import _native
    if _native.device_count() < 1:
        pytest.fail("no HIP device: the GPU tests need an MI355X")
    return _native


def _noalloc(native):
    fn = native.lib.mmd_fock_delta_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                   ctypes.POINTER(ctypes.c_int64)]
    return fn


class _Dev:
    """Device buffers through the HIP runtime that libmmd itself loads (torch
    ships its own copy of the runtime, and torch could not initialise the
    device in a process that already used libmmd)."""

    def __init__(self):
        try:
            self.hip = ctypes.CDLL("libamdhip64.so.7")
        except OSError:
            self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _sym(N, seed):
    A = np.random.default_rng(seed).standard_normal((N, N))
    return A + A.T


def test_noalloc_matches_fock_delta_twice(native):
    from mmd.molecule import Molecule
    fn = _noalloc(native)
    dev = _Dev()
    for n, seed in ((3, 5), (1, 2)):   # the second call has the smaller basis
        mol = Molecule(water_cluster(n, 7.0, seed), "cc-pvdz")
        mol.build(direct=True)
        ctx = mol.native
        N = mol.nbasis
        dP = _sym(N, seed)
        G_ref, _, ns_ref = ctx.fock_delta(dP, None, tol=1e-12, raw=False)
        d = dev.upload(dP)
        g = dev.upload(np.full((N, N), np.nan))
        ns = ctypes.c_int64(-1)
        native.check(fn(ctx._h, d, 1e-12, g, ctypes.byref(ns)))
        G = dev.download(g, (N, N))
        err = np.max(np.abs(G - G_ref))
        print("N = %d: max |G_noalloc - G| %.2e, survivors %d / %d" % (N, err, ns.value, ns_ref))
        assert ns.value == ns_ref
        assert err < 1e-12
        assert np.array_equal(G, G.T)
    dev.free()


def test_noalloc_rejects_lists_beyond_scratch(native):
    """With every class on the survivor-list kernel (LDS budget 1 KiB),
    (H2O)8 / cc-pVDZ needs about 1e7 list entries, more than the 16 MiB
    scratch holds: a clear ValueError, and the context still works."""
    from mmd.molecule import Molecule
    fn = _noalloc(native)
    dev = _Dev()
    mol = Molecule(water_cluster(8, 7.0, 0), "cc-pvdz")
    mol.build(direct=True)
    ctx = mol.native
    N = mol.nbasis
    d = dev.upload(np.zeros((N, N)))
    g = dev.upload(np.full((N, N), np.nan))
    ctx.set_lds_limit(1024)
    try:
        with pytest.raises(ValueError, match="scratch"):
            native.check(fn(ctx._h, d, 1e-12, g, None))
    finally:
        ctx.set_lds_limit(0)
    ns = ctypes.c_int64(-1)
    native.check(fn(ctx._h, d, 1e-12, g, ctypes.byref(ns)))
    G = dev.download(g, (N, N))
    dev.free()
    assert ns.value == 0 and not np.any(G)
