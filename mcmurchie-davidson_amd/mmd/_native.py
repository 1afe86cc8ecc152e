"""ctypes binding of libmmd.so (the C ABI of include/mmd.h).

The library is the only compute path of this package: every integral and
Fock build runs in its HIP kernels.  Importing works without a GPU (so that
host-side code can be used and tested); creating a context on a machine
without a HIP device raises RuntimeError.  A missing library raises
ImportError -- there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MMD_LIB", os.path.join(os.path.dirname(_PKG_DIR), "lib", "libmmd.so"))

MMD_OK = 0
MMD_ERR_ARG = -1
MMD_ERR_NOMEM = -3
MMD_ERR_DEVICE = -4
MMD_ERR_STATE = -5
MMD_ERR_UNSUPPORTED = -6

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libmmd.so not found at %s; build it with `make -C mcmurchie-davidson_amd/csrc` "
        "(or __graft_entry__.build())" % LIB_PATH)

# The Fock build spreads its shell-quartet classes over min(8,
# GPU_MAX_HW_QUEUES) streams (four when the variable is unset, HIP's own
# default); a process that wants eight sets the variable before HIP starts.
# The library itself leaves the environment alone.
lib = ctypes.CDLL(LIB_PATH)

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_ctx_p = ctypes.c_void_p

lib.mmd_last_error.restype = ctypes.c_char_p
lib.mmd_last_error.argtypes = []
lib.mmd_version.restype = _c_int
lib.mmd_device_count.restype = _c_int
lib.mmd_device_count.argtypes = [ctypes.POINTER(_c_int)]
lib.mmd_ctx_create.restype = _c_int
lib.mmd_ctx_create.argtypes = [ctypes.POINTER(_ctx_p), _c_int]
lib.mmd_ctx_destroy.restype = None
lib.mmd_ctx_destroy.argtypes = [_ctx_p]
lib.mmd_set_basis.restype = _c_int
lib.mmd_set_basis.argtypes = [_ctx_p, _c_int, _vp, _vp, _vp, _vp, _vp]
lib.mmd_set_atoms.restype = _c_int
lib.mmd_set_atoms.argtypes = [_ctx_p, _c_int, _vp, _vp]
lib.mmd_one_electron.restype = _c_int
lib.mmd_one_electron.argtypes = [_ctx_p, _vp, _vp, _vp, _vp, _vp, _vp]
lib.mmd_schwarz.restype = _c_int
lib.mmd_schwarz.argtypes = [_ctx_p, _vp]
lib.mmd_set_molecules.restype = _c_int
lib.mmd_set_molecules.argtypes = [_ctx_p, _c_int, _vp, _vp]
lib.mmd_batch_rhf.restype = _c_int
lib.mmd_batch_rhf.argtypes = [_ctx_p, _vp, _vp, _vp, _c_double, _c_double, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                              _vp, _vp]
lib.mmd_overlap_powers.restype = _c_int
lib.mmd_overlap_powers.argtypes = [_ctx_p, _vp, _vp, _vp, ctypes.POINTER(_c_int)]
lib.mmd_dgemm.restype = _c_int
lib.mmd_dgemm.argtypes = [_ctx_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _vp, _c_int, _vp, _c_int,
                          _c_double, _vp, _c_int, _c_int]
lib.mmd_set_lds_limit.restype = _c_int
lib.mmd_set_lds_limit.argtypes = [_ctx_p, _c_i64]
lib.mmd_set_schwarz.restype = _c_int
lib.mmd_set_schwarz.argtypes = [_ctx_p, _vp]
lib.mmd_eri_tensor.restype = _c_int
lib.mmd_eri_tensor.argtypes = [_ctx_p, _vp, _c_int]
lib.mmd_eri_list.restype = _c_int
lib.mmd_eri_list.argtypes = [_ctx_p, _c_i64, _vp, _vp]
lib.mmd_fock_delta.restype = _c_int
lib.mmd_fock_delta.argtypes = [_ctx_p, _vp, _vp, _c_double, _c_int, _vp, _vp, _c_int, ctypes.POINTER(_c_i64)]
lib.mmd_jk_tensor.restype = _c_int
lib.mmd_jk_tensor.argtypes = [_ctx_p, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp]
lib.mmd_boys.restype = _c_int
lib.mmd_boys.argtypes = [_ctx_p, _c_int, _vp, _vp, _c_int, _vp]
lib.mmd_set_profiling.restype = _c_int
lib.mmd_set_profiling.argtypes = [_ctx_p, _c_int]
lib.mmd_fock_stats.restype = _c_int
lib.mmd_fock_stats.argtypes = [_ctx_p, _vp, _vp, _vp, _vp]
lib.mmd_set_class_impl.restype = _c_int
lib.mmd_set_class_impl.argtypes = [_ctx_p, _c_int, _vp]
lib.mmd_set_ket_slices.restype = _c_int
lib.mmd_set_ket_slices.argtypes = [_ctx_p, _vp, _vp]
lib.mmd_rhf_gradient.restype = _c_int
lib.mmd_rhf_gradient.argtypes = [_ctx_p, _vp, _vp, _vp, _vp]
lib.mmd_last_gradient_ms.restype = _c_int
lib.mmd_last_gradient_ms.argtypes = [_ctx_p, _vp]
lib.mmd_fock_profile_sums.restype = _c_int
lib.mmd_fock_profile_sums.argtypes = [_ctx_p, _vp, _vp, _vp, _vp, _vp]
lib.mmd_last_fock_ms.restype = _c_int
lib.mmd_last_fock_ms.argtypes = [_ctx_p, ctypes.POINTER(_c_double)]
lib.mmd_set_precision.restype = _c_int
lib.mmd_set_precision.argtypes = [_ctx_p, _c_double]
lib.mmd_set_streams.restype = _c_int
lib.mmd_set_streams.argtypes = [_ctx_p, _c_int]
lib.mmd_fock_tiers.restype = _c_int
lib.mmd_fock_tiers.argtypes = [_ctx_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]
lib.mmd_scf_begin.restype = _c_int
lib.mmd_scf_begin.argtypes = [_ctx_p, _c_int, _vp, _vp, _vp, _vp, _vp]
lib.mmd_scf_step.restype = _c_int
lib.mmd_scf_step.argtypes = [_ctx_p, _c_double, _c_int, _c_int, _vp]
lib.mmd_scf_density.restype = _c_int
lib.mmd_scf_density.argtypes = [_ctx_p, _vp, _vp]
lib.mmd_scf_get.restype = _c_int
lib.mmd_scf_get.argtypes = [_ctx_p, _vp, _vp, _vp, _vp]
lib.mmd_ao2mo.restype = _c_int
lib.mmd_ao2mo.argtypes = [_ctx_p, _vp, _c_int, _vp, _vp, _c_int, ctypes.POINTER(_c_double)]
lib.mmd_mlp_create.restype = _c_int
lib.mmd_mlp_create.argtypes = [ctypes.POINTER(_vp), _c_int, _vp, _c_int]
lib.mmd_mlp_destroy.restype = None
lib.mmd_mlp_destroy.argtypes = [_vp]
lib.mmd_mlp_nparam.restype = _c_i64
lib.mmd_mlp_nparam.argtypes = [_vp]
lib.mmd_mlp_set_activation.restype = _c_int
lib.mmd_mlp_set_activation.argtypes = [_vp, _c_int]
lib.mmd_mlp_set_params.restype = _c_int
lib.mmd_mlp_set_params.argtypes = [_vp, _vp]
lib.mmd_mlp_get_params.restype = _c_int
lib.mmd_mlp_get_params.argtypes = [_vp, _vp]
lib.mmd_mlp_forward.restype = _c_int
lib.mmd_mlp_forward.argtypes = [_vp, _c_int, _vp, _vp]
lib.mmd_mlp_loss_grad.restype = _c_int
lib.mmd_mlp_loss_grad.argtypes = [_vp, _c_int, _vp, _vp, _vp, _vp]
lib.mmd_mlp_train.restype = _c_int
lib.mmd_mlp_train.argtypes = [_vp, _c_int, _vp, _vp, _c_int, _c_int, ctypes.c_float, ctypes.c_uint64, _vp]

EXPORTED = [
    "mmd_last_error", "mmd_version", "mmd_device_count", "mmd_ctx_create", "mmd_ctx_destroy",
    "mmd_set_basis", "mmd_set_atoms", "mmd_one_electron", "mmd_schwarz", "mmd_set_schwarz",
    "mmd_eri_tensor", "mmd_eri_list", "mmd_fock_delta", "mmd_fock_delta_noalloc", "mmd_jk_tensor", "mmd_boys",
    "mmd_last_fock_ms", "mmd_set_profiling", "mmd_fock_stats", "mmd_fock_profile_sums", "mmd_rhf_gradient", "mmd_set_class_impl", "mmd_set_ket_slices",
    "mmd_last_gradient_ms", "mmd_set_precision", "mmd_fock_tiers",
    "mmd_set_streams", "mmd_set_lds_limit", "mmd_overlap_powers", "mmd_dgemm", "mmd_set_molecules",
    "mmd_batch_rhf", "mmd_scf_begin", "mmd_scf_step", "mmd_scf_density", "mmd_scf_get", "mmd_ao2mo",
    "mmd_mlp_create", "mmd_mlp_destroy", "mmd_mlp_nparam", "mmd_mlp_set_activation", "mmd_mlp_set_params", "mmd_mlp_get_params",
    "mmd_mlp_forward", "mmd_mlp_loss_grad", "mmd_mlp_train",
]


def check(rc):
    """Map a libmmd status code to the reference's exception types."""
    if rc == MMD_OK:
        return
    msg = lib.mmd_last_error().decode(errors="replace")
    if rc == MMD_ERR_ARG:
        raise ValueError(msg)
    if rc == MMD_ERR_NOMEM:
        raise MemoryError(msg)
    if rc == MMD_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def device_count():
    n = _c_int(0)
    check(lib.mmd_device_count(ctypes.byref(n)))
    return n.value


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One libmmd context (device, stream, basis, screening data)."""

    def __init__(self, device=0):
        h = _ctx_p()
        check(lib.mmd_ctx_create(ctypes.byref(h), int(device)))
        self._h = h
        self.device = int(device)
        self.nbf = 0

    def close(self):
        if getattr(self, "_h", None):
            lib.mmd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- basis ---------------------------------------------------------------
    def set_basis(self, origin, lmn, nprim, exps, weights):
        origin = _f64(origin).reshape(-1, 3)
        lmn = np.ascontiguousarray(lmn, dtype=np.int32).reshape(-1, 3)
        nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        exps = _f64(exps)
        weights = _f64(weights)
        check(lib.mmd_set_basis(self._h, len(nprim), _ptr(origin), _ptr(lmn), _ptr(nprim),
                                _ptr(exps), _ptr(weights)))
        self.nbf = len(nprim)

    def set_basis_from_bfs(self, bfs):
        origin = np.array([b.origin for b in bfs], dtype=np.float64)
        lmn = np.array([b.shell for b in bfs], dtype=np.int32)
        nprim = np.array([b.num_exps for b in bfs], dtype=np.int32)
        exps = np.concatenate([np.asarray(b.exps, dtype=np.float64) for b in bfs])
        w = np.concatenate([np.asarray(b.norm, dtype=np.float64) * np.asarray(b.coefs, dtype=np.float64)
                            for b in bfs])
        self.set_basis(origin, lmn, nprim, exps, w)

    def set_atoms(self, Z, xyz):
        Z = _f64(Z)
        xyz = _f64(xyz).reshape(-1, 3)
        check(lib.mmd_set_atoms(self._h, len(Z), _ptr(Z), _ptr(xyz)))
        self.natom = len(Z)

    # -- integrals -------------------------------------------------------------
    def one_electron(self, center):
        N = self.nbf
        c = _f64(center)
        S, T, V = np.empty((N, N)), np.empty((N, N)), np.empty((N, N))
        M, L = np.empty((3, N, N)), np.empty((3, N, N))
        check(lib.mmd_one_electron(self._h, _ptr(c), _ptr(S), _ptr(T), _ptr(V), _ptr(M), _ptr(L)))
        return S, T, V, M, L

    def set_molecules(self, nbf_per_mol, natom_per_mol):
        """The basis and nuclei are those of several molecules, block-diagonally."""
        nb = np.ascontiguousarray(nbf_per_mol, dtype=np.int32)
        na = np.ascontiguousarray(natom_per_mol, dtype=np.int32)
        check(lib.mmd_set_molecules(self._h, len(nb), _ptr(nb), _ptr(na)))
        self.nmol = len(nb)
        self.mol_nbf = nb

    def batch_rhf(self, nocc, enuc, P0=None, conver=1e-8, acc2e=1e-12, maxiter=100, diis=True, use_graph=True,
                  keep_density=False):
        """Direct RHF of every molecule of the basis at once (mmd_batch_rhf);
        returns (energy, iterations, converged[, packed densities])."""
        M = self.nmol
        no = np.ascontiguousarray(nocc, dtype=np.int32)
        en = _f64(enuc)
        E = np.empty(M)
        it = np.empty(M, dtype=np.int32)
        cv = np.empty(M, dtype=np.int32)
        ptot = int(np.sum(self.mol_nbf.astype(np.int64) ** 2))
        P = np.empty(ptot) if keep_density else None
        p0 = None if P0 is None else _f64(P0)
        check(lib.mmd_batch_rhf(self._h, _ptr(no), _ptr(en), _ptr(p0), float(conver), float(acc2e), int(maxiter),
                                int(bool(diis)), int(bool(use_graph)), _ptr(E), _ptr(it), _ptr(cv), _ptr(P), None))
        return (E, it, cv.astype(bool), P) if keep_density else (E, it, cv.astype(bool))

    def overlap_powers(self, S=None):
        """(X, U, iterations): S^-1/2 and S^1/2 on the GPU (Newton-Schulz on
        the matrix cores); S None uses the overlap of the last one_electron."""
        N = self.nbf
        X, U = np.empty((N, N)), np.empty((N, N))
        it = _c_int(0)
        src = None if S is None else _f64(S).reshape(N, N)
        check(lib.mmd_overlap_powers(self._h, None if src is None else _ptr(src), _ptr(X), _ptr(U),
                                     ctypes.byref(it)))
        return X, U, it.value

    def dgemm(self, A, B, transA=False, transB=False, alpha=1.0, beta=0.0, C=None):
        """alpha op(A) op(B) + beta C on the matrix cores (host arrays)."""
        A, B = _f64(A), _f64(B)
        M = A.shape[1] if transA else A.shape[0]
        K = A.shape[0] if transA else A.shape[1]
        N = B.shape[0] if transB else B.shape[1]
        Cm = np.zeros((M, N)) if C is None else np.array(C, dtype=np.float64, order="C")
        check(lib.mmd_dgemm(self._h, int(transA), int(transB), M, N, K, float(alpha), _ptr(A), A.shape[1],
                            _ptr(B), B.shape[1], float(beta), _ptr(Cm), max(1, N), 0))
        return Cm

    def schwarz(self):
        d = np.empty((self.nbf, self.nbf))
        check(lib.mmd_schwarz(self._h, _ptr(d)))
        return d

    def set_schwarz(self, diag):
        d = _f64(diag)
        check(lib.mmd_set_schwarz(self._h, _ptr(d)))

    def eri_tensor(self, out=None):
        N = self.nbf
        if out is None:
            out = np.empty((N, N, N, N))
        check(lib.mmd_eri_tensor(self._h, _ptr(out), 0))
        return out

    def eri_tensor_device(self, out_ptr):
        check(lib.mmd_eri_tensor(self._h, ctypes.c_void_p(out_ptr), 1))

    def eri_list(self, ijkl):
        ijkl = np.ascontiguousarray(ijkl, dtype=np.int64).reshape(-1, 4)
        out = np.empty(len(ijkl))
        check(lib.mmd_eri_list(self._h, len(ijkl), _ptr(ijkl), _ptr(out)))
        return out

    def fock_delta(self, dP_re, dP_im=None, tol=1e-12, raw=False):
        N = self.nbf
        dr = _f64(dP_re)
        di = None if dP_im is None else _f64(dP_im)
        Gr = np.empty((N, N))
        Gi = None if di is None else np.empty((N, N))
        ns = _c_i64(0)
        check(lib.mmd_fock_delta(self._h, _ptr(dr), _ptr(di), float(tol), 1 if raw else 0,
                                 _ptr(Gr), _ptr(Gi), 0, ctypes.byref(ns)))
        return Gr, Gi, ns.value

    def fock_delta_device(self, dP_ptr, G_ptr, tol=1e-12):
        """Symmetrised G(dP) with dP and G already in device memory."""
        ns = _c_i64(0)
        check(lib.mmd_fock_delta(self._h, ctypes.c_void_p(dP_ptr), None, float(tol), 0,
                                 ctypes.c_void_p(G_ptr), None, 1, ctypes.byref(ns)))
        return ns.value

    def jk_tensor(self, T, P_re, P_im=None, t_device_ptr=None):
        N = self.nbf
        pr = _f64(P_re)
        pi = None if P_im is None else _f64(P_im)
        Jr, Kr = np.empty((N, N)), np.empty((N, N))
        Ji = Ki = None
        if pi is not None:
            Ji, Ki = np.empty((N, N)), np.empty((N, N))
        if t_device_ptr is not None:
            check(lib.mmd_jk_tensor(self._h, ctypes.c_void_p(t_device_ptr), 1, _ptr(pr), _ptr(pi),
                                    _ptr(Jr), _ptr(Ji), _ptr(Kr), _ptr(Ki)))
        else:
            t = _f64(T)
            check(lib.mmd_jk_tensor(self._h, _ptr(t), 0, _ptr(pr), _ptr(pi), _ptr(Jr), _ptr(Ji),
                                    _ptr(Kr), _ptr(Ki)))
        return Jr, Ji, Kr, Ki

    def boys(self, m, T, table=False):
        m = _f64(np.atleast_1d(m))
        T = _f64(np.atleast_1d(T))
        m, T = np.broadcast_arrays(m, T)
        m, T = _f64(m), _f64(T)
        out = np.empty(m.shape)
        check(lib.mmd_boys(self._h, m.size, _ptr(m), _ptr(T), 1 if table else 0, _ptr(out)))
        return out

    def set_profiling(self, on=True, serial=False):
        """Per-class statistics and event times of every symmetrised build;
        serial=True runs the classes one after the other (isolated times)."""
        check(lib.mmd_set_profiling(self._h, (2 if serial else 1) if on else 0))

    def fock_stats(self):
        """Per-class (l_a, l_b, l_c, l_d), surviving shell quartets, surviving
        primitive quartets and kernel ms of the last symmetrised build."""
        lc = np.zeros(84, dtype=np.int32)
        nq = np.zeros(21, dtype=np.int64)
        npr = np.zeros(21, dtype=np.int64)
        ms = np.zeros(21)
        n = lib.mmd_fock_stats(self._h, _ptr(lc), _ptr(nq), _ptr(npr), _ptr(ms))
        if n < 0:
            check(n)
        return lc.reshape(21, 4)[:n], nq[:n], npr[:n], ms[:n]

    def set_class_impl(self, mode=None):
        """Kernel choice of the classes with both a row and a matrix-core
        kernel: 0 row, 1 matrix cores, 2 measured per class (default); None
        keeps the mode.  Returns the per-class choice (21 entries: 0 row,
        1 matrix cores, -1 undecided)."""
        ch = np.zeros(21, dtype=np.int32)
        check(lib.mmd_set_class_impl(self._h, -1 if mode is None else int(mode), _ptr(ch)))
        return ch

    def set_ket_slices(self, slices=None):
        """Ket primitive pairs per lane item of the row kernels' sliced
        classes: slices (21 entries, 0 keeps a class's value) or None;
        returns the values in force."""
        cur = np.zeros(21, dtype=np.int32)
        src = None if slices is None else np.ascontiguousarray(slices, dtype=np.int32)
        if src is not None and src.size != 21:
            raise ValueError("21 entries expected")
        check(lib.mmd_set_ket_slices(self._h, None if src is None else _ptr(src), _ptr(cur)))
        return cur

    def rhf_gradient(self, P, W, bf_atom):
        """Electronic part of the RHF nuclear gradient (natom, 3) for the
        half density P, W = P F P and the atom of each basis function."""
        N = self.nbf
        P = np.ascontiguousarray(np.real(P), dtype=np.float64)
        W = np.ascontiguousarray(np.real(W), dtype=np.float64)
        at = np.ascontiguousarray(bf_atom, dtype=np.int32)
        if P.shape != (N, N) or W.shape != (N, N) or at.shape != (N,):
            raise ValueError("P and W must be N x N and bf_atom of length N")
        g = np.zeros(3 * self.natom)
        check(lib.mmd_rhf_gradient(self._h, _ptr(P), _ptr(W), _ptr(at), _ptr(g)))
        return g.reshape(-1, 3)

    def last_gradient_ms(self):
        v = _c_double(0.0)
        check(lib.mmd_last_gradient_ms(self._h, ctypes.byref(v)))
        return v.value

    def fock_profile_sums(self):
        """Sums since profiling was turned on, per class: (l_a, l_b, l_c, l_d),
        shell quartets, primitive quartets and kernel ms, each as (FP64
        launches, FP32-tier launches), and the number of builds."""
        lc = np.zeros(84, dtype=np.int32)
        nq = np.zeros(42, dtype=np.int64)
        npr = np.zeros(42, dtype=np.int64)
        ms = np.zeros(42)
        nb = np.zeros(21, dtype=np.int64)
        n = lib.mmd_fock_profile_sums(self._h, _ptr(lc), _ptr(nq), _ptr(npr), _ptr(ms), _ptr(nb))
        if n < 0:
            check(n)
        return (lc.reshape(21, 4)[:n], nq.reshape(21, 2)[:n], npr.reshape(21, 2)[:n], ms.reshape(21, 2)[:n],
                nb[:n])

    def last_fock_ms(self):
        v = _c_double(0.0)
        check(lib.mmd_last_fock_ms(self._h, ctypes.byref(v)))
        return v.value

    def set_precision(self, fp32_below=0.0):
        """FP32 tier: shell quartets with screening bound below fp32_below
        get FP32 integrals (FP64 digestion); 0 keeps everything FP64."""
        check(lib.mmd_set_precision(self._h, float(fp32_below)))

    def set_streams(self, n):
        """Streams (1..8) the Fock-build classes are spread over."""
        check(lib.mmd_set_streams(self._h, int(n)))

    # -- GPU-resident SCF loop -------------------------------------------------
    def scf_begin(self, nocc, H, S, X, F, P):
        """Upload H, S, X, F and P (= P_old) for mmd_scf_step / mmd_scf_density."""
        arrs = [_f64(np.real(a)) for a in (H, S, X, F, P)]   # kept alive over the call
        check(lib.mmd_scf_begin(self._h, int(nocc), *[_ptr(a) for a in arrs]))

    def scf_step(self, tol, build, diis, FO):
        check(lib.mmd_scf_step(self._h, float(tol), 1 if build else 0, 1 if diis else 0, _ptr(FO)))
        return FO

    def scf_density(self, CO):
        stats = np.zeros(2)
        co = _f64(CO)
        check(lib.mmd_scf_density(self._h, _ptr(co), _ptr(stats)))
        return float(stats[0]), float(stats[1])

    def scf_get(self):
        N = self.nbf
        F, P, Po, G = (np.empty((N, N)) for _ in range(4))
        check(lib.mmd_scf_get(self._h, _ptr(F), _ptr(P), _ptr(Po), _ptr(G)))
        return F, P, Po, G

    def ao2mo(self, T, C, out=None):
        """(PQ|RS) = sum C_pP C_qQ C_rR C_sS (pq|rs) of an N^4 tensor (GPU GEMMs)."""
        N = self.nbf
        t = _f64(T)
        c = _f64(np.real(C))
        if t.shape != (N, N, N, N) or c.shape != (N, N):
            raise ValueError("T must be (N,N,N,N) and C (N,N) with N = %d" % N)
        if out is None:
            out = np.empty((N, N, N, N))
        ms = _c_double(0.0)
        check(lib.mmd_ao2mo(self._h, _ptr(t), 0, _ptr(c), _ptr(out), 0, ctypes.byref(ms)))
        self.last_ao2mo_ms = ms.value
        return out

    def set_lds_limit(self, nbytes=0):
        """LDS budget of a Fock class launch (0: the device's limit)."""
        check(lib.mmd_set_lds_limit(self._h, int(nbytes)))

    def fock_tiers(self):
        """(surviving shell quartets, of which FP32) of the last symmetrised build."""
        a, b = _c_i64(0), _c_i64(0)
        check(lib.mmd_fock_tiers(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class MLP:
    """The learned guesses' perceptron on the GPU (mmd_mlp_* of include/mmd.h):
    tanh or SiLU hidden layers (activation 'tanh' or 'silu'), linear output,
    FP32."""

    def __init__(self, dims, device=0, activation="tanh"):
        self.dims = [int(d) for d in dims]
        if activation not in ("tanh", "silu"):
            raise ValueError("activation must be 'tanh' or 'silu'")
        h = _vp()
        d = np.ascontiguousarray(self.dims, dtype=np.int32)
        check(lib.mmd_mlp_create(ctypes.byref(h), len(self.dims) - 1, _ptr(d), device))
        self._h = h
        self.activation = activation
        check(lib.mmd_mlp_set_activation(h, 1 if activation == "silu" else 0))
        self.nparam = int(lib.mmd_mlp_nparam(h))

    def close(self):
        if getattr(self, "_h", None):
            lib.mmd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.nparam:
            raise ValueError("expected %d parameters, got %d" % (self.nparam, flat.size))
        check(lib.mmd_mlp_set_params(self._h, _ptr(flat)))

    def get_params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        check(lib.mmd_mlp_get_params(self._h, _ptr(out)))
        return out

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != self.dims[0]:
            raise ValueError("input must be (n, %d)" % self.dims[0])
        Y = np.empty((X.shape[0], self.dims[-1]), dtype=np.float32)
        check(lib.mmd_mlp_forward(self._h, X.shape[0], _ptr(X), _ptr(Y)))
        return Y

    def loss_grad(self, X, Yt):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Yt = np.ascontiguousarray(Yt, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        check(lib.mmd_mlp_loss_grad(self._h, X.shape[0], _ptr(X), _ptr(Yt), _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def train(self, X, Yt, epochs, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Yt = np.ascontiguousarray(Yt, dtype=np.float32)
        losses = np.zeros(max(1, epochs), dtype=np.float32)
        check(lib.mmd_mlp_train(self._h, X.shape[0], _ptr(X), _ptr(Yt), int(epochs), int(batch), float(lr),
                                int(seed), _ptr(losses)))
        return losses[:epochs]


def mlp_forward(ctx, model, X):
    """Forward pass of a mmd.learned.Model on the GPU of ``ctx``."""
    dev = getattr(ctx, "device", 0)
    net = getattr(model, "_gpu", None)
    act = getattr(model, "activation", "tanh")
    if net is None or net.dims != model.dims or net.activation != act:
        net = MLP(model.dims, dev, activation=act)
        net.set_params(model.flat())
        model._gpu = net
    return net.forward(X)


_default = {}
_lock = threading.Lock()


def default_context(device=0):
    """Process-wide scratch context (for per-call entry points such as ERI())."""
    with _lock:
        ctx = _default.get(device)
        if ctx is None:
            ctx = Context(device)
            _default[device] = ctx
        return ctx


class PerCallBasis(object):
    """Device-resident basis for the per-call entry points (ERI(), S(), ...).

    The Basis objects seen so far stay on the device of one context, so that
    a loop over integrals of the same functions (the reference's
    tests/test013.py style) uploads each function once instead of on every
    call.  A function is recognised by identity and by its contents (origin,
    exponents, weights), so an object changed in place is uploaded again.
    """

    MAX_FUNCTIONS = 512

    def __init__(self, device=0):
        self.ctx = Context(device)
        self.bfs = []       # the resident functions, in device order (kept alive)
        self.index = {}     # id -> (position, fingerprint)
        self.lock = threading.Lock()

    @staticmethod
    def _fingerprint(b):
        return (np.asarray(b.origin, dtype=np.float64).tobytes(), np.asarray(b.shell).tobytes(),
                np.asarray(b.exps, dtype=np.float64).tobytes(),
                (np.asarray(b.norm, dtype=np.float64) * np.asarray(b.coefs, dtype=np.float64)).tobytes())

    def indices(self, fns):
        """Positions of `fns` in the resident basis, uploading it again when
        a function is new or changed.  Call with self.lock held."""
        pos = []
        stale = False
        fps = [self._fingerprint(b) for b in fns]
        for b, fp in zip(fns, fps):
            hit = self.index.get(id(b))
            if hit is None or hit[1] != fp or self.bfs[hit[0]] is not b:
                stale = True
                break
            pos.append(hit[0])
        if not stale:
            return pos
        if len(self.bfs) + len(fns) > self.MAX_FUNCTIONS:
            self.bfs, self.index = [], {}
        for b, fp in zip(fns, fps):
            hit = self.index.get(id(b))
            if hit is not None and self.bfs[hit[0]] is b:
                if hit[1] != fp:       # changed in place: new slot
                    self.bfs.append(b)
                    self.index[id(b)] = (len(self.bfs) - 1, fp)
            else:
                self.bfs.append(b)
                self.index[id(b)] = (len(self.bfs) - 1, fp)
        self.ctx.set_basis_from_bfs(self.bfs)
        return [self.index[id(b)][0] for b in fns]


_percall = {}


def per_call_basis(device=0):
    with _lock:
        pc = _percall.get(device)
        if pc is None:
            pc = _percall[device] = PerCallBasis(device)
        return pc


# Contexts keyed by the identity of a basis-function list, so that formPT,
# which receives only `bfs`, reuses the basis already resident on the device.
_bfs_cache = {}
_BFS_CACHE_MAX = 8


def context_for_bfs(bfs, device=0):
    key = tuple(id(b) for b in bfs)
    with _lock:
        hit = _bfs_cache.get(key)
        if hit is not None:
            return hit[0]
    ctx = Context(device)
    ctx.set_basis_from_bfs(bfs)
    register_bfs_context(bfs, ctx)
    return ctx


def register_bfs_context(bfs, ctx):
    key = tuple(id(b) for b in bfs)
    with _lock:
        if len(_bfs_cache) >= _BFS_CACHE_MAX:
            _bfs_cache.pop(next(iter(_bfs_cache)))
        # keep the Basis objects alive so that their ids stay unique
        _bfs_cache[key] = (ctx, list(bfs))
