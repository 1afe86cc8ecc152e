"""Pins of the CPU oracle (CPU only): the reference's known answers, closed
forms and invariants chosen so that a dropped term, a wrong sign or index,
or a transposed operand in the restatement fails at least one of them."""
import os

import numpy as np
import pytest

from geometries import (FORMALDEHYDE, H2, HELIUM, METHANE, WATER, WATER_BOHR, BOHR_1986, from_bohr,
                        rescale, water_cluster)
from test_gpu_parity import BOYS_LITERATURE

BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mcmurchie-davidson_amd",
                     "mmd", "basis")


def _gbs(name):
    return os.path.join(BASIS, name + ".gbs")


# --------------------------------------------------------------------------
# Boys function
# --------------------------------------------------------------------------
def test_boys_literature_table(oracle):
    """tests/test012.py of the reference: 50 values, places=12."""
    for m, T, ref in BOYS_LITERATURE:
        assert abs(oracle.boys(m, T) - ref) < 5e-13, (m, T)


def test_boys_closed_forms(oracle):
    from scipy.special import erf
    for m in range(12):
        assert oracle.boys(m, 0.0) == pytest.approx(1.0 / (2 * m + 1), rel=1e-15)
    for T in [1e-6, 0.3, 1.0, 7.7, 30.0, 120.0, 900.0]:
        assert oracle.boys(0, T) == pytest.approx(0.5 * np.sqrt(np.pi / T) * erf(np.sqrt(T)), rel=1e-14)
        for m in range(10):   # downward recursion F_m = (2T F_{m+1} + e^-T)/(2m+1)
            lhs = oracle.boys(m, T)
            rhs = (2 * T * oracle.boys(m + 1, T) + np.exp(-T)) / (2 * m + 1)
            assert lhs == pytest.approx(rhs, rel=1e-13)


# --------------------------------------------------------------------------
# Normalisation and one-electron integrals
# --------------------------------------------------------------------------
C_SE = [3047.5249000, 457.3695100, 103.9486900, 29.2101550, 9.2866630, 3.1639270]
C_SC = [0.0018347, 0.0140373, 0.0688426, 0.2321844, 0.4679413, 0.3623120]
C_PE = [7.8682724, 1.8812885, 0.5442493]
C_PC = [0.0689991, 0.3164240, 0.7443083]


class _Mini:
    """A hand-made set of basis functions (carbon 6-31G** s, p, d on two
    centres 2 bohr apart, tests/test013.py of the reference)."""

    def __init__(self, oracle, bfl):
        o = oracle.OracleMolecule.__new__(oracle.OracleMolecule)
        o.origin = np.ascontiguousarray(np.array([b[0] for b in bfl], float))
        o.shell = np.ascontiguousarray(np.array([b[1] for b in bfl], np.int64))
        o.np = np.array([len(b[2]) for b in bfl], np.int64)
        o.poff = np.concatenate([[0], np.cumsum(o.np)[:-1]]).astype(np.int64)
        o.exps = np.concatenate([b[2] for b in bfl]).astype(float)
        o.coefs = np.concatenate([b[3] for b in bfl]).astype(float)
        o.norm = np.zeros_like(o.exps)
        for i in range(len(bfl)):
            a, k = o.poff[i], o.np[i]
            e = np.ascontiguousarray(o.exps[a:a + k])
            c = np.ascontiguousarray(o.coefs[a:a + k])
            nr = np.zeros(k)
            oracle.lib().orc_normalize(*[int(x) for x in o.shell[i]], int(k), e, c, nr)
            o.coefs[a:a + k] = c
            o.norm[a:a + k] = nr
        o.nbasis = len(bfl)
        o.atoms = [(1, np.array([-0.34523, 0.4, 2.0]))]
        o.Z = np.array([1.0])
        o.xyz = np.array([[-0.34523, 0.4, 2.0]])
        o.center_of_charge = np.zeros(3)
        self.o = o


def _carbon(oracle):
    bfl = [([0, 0, 0], (0, 0, 0), C_SE, C_SC), ([0, 0, 0], (1, 0, 0), C_PE, C_PC), ([0, 0, 0], (2, 0, 0), [0.8], [1.0]),
           ([2, 0, 0], (0, 0, 0), C_SE, C_SC), ([2, 0, 0], (1, 0, 0), C_PE, C_PC), ([2, 0, 0], (2, 0, 0), [0.8], [1.0])]
    return _Mini(oracle, bfl).o


def test_overlap_kinetic_literals(oracle):
    """tests/test013.py:46-87 of the reference (assertAlmostEqual: 7 places)."""
    S, T, V, M, L = _carbon(oracle).one_electron()
    s1, p1x, d1xx, s2, p2x, d2xx = range(6)
    for i in (s2, p2x, d2xx):
        assert S[i, i] == pytest.approx(1.0, abs=1e-14)
    lit_S = {(s1, s2): 0.000257231, (s1, p2x): -0.115018581165, (s1, d2xx): 0.177080256092,
             (p1x, p2x): -0.3739122083249, (p1x, d2xx): 0.45586902867, (d1xx, d2xx): 0.46032406102}
    lit_T = {(s2, s2): 16.2075631760, (p2x, p2x): 2.178332480866, (d2xx, d2xx): 1.7333333333333327,
             (s1, s2): -0.004453163473, (s1, p2x): -0.02917787713063, (s1, d2xx): 0.122632029183,
             (p1x, p2x): -0.6425660209155, (p1x, d2xx): 0.803805220928, (d1xx, d2xx): 0.5532502983768}
    for (i, j), v in lit_S.items():
        assert round(S[i, j] - v, 7) == 0
    for (i, j), v in lit_T.items():
        assert round(T[i, j] - v, 7) == 0


def test_nuclear_attraction_derivative_relation(oracle):
    """(p_x|V_C|s) = (1/2a) d/dA_x (s|V_C|s) with the closed form
    (s|V_C|s) = 2pi/p K_AB F_0(p R_PC^2) (Richardson-extrapolated)."""
    a, b = 1.3, 0.7
    A = np.array([0.1, -0.3, 0.2])
    B = np.array([1.0, 0.4, -0.5])
    C = np.array([-0.4, 0.9, 0.3])

    def ssv(Ax):
        p = a + b
        P = (a * Ax + b * B) / p
        return 2 * np.pi / p * np.exp(-a * b / p * np.sum((Ax - B) ** 2)) * oracle.boys(0, p * np.sum((P - C) ** 2))

    def prim(l1):
        p = a + b
        P = (a * A + b * B) / p
        R = np.linalg.norm(P - C)
        val = 0.0
        for t in range(l1[0] + 1):
            for u in range(l1[1] + 1):
                for v in range(l1[2] + 1):
                    val += (oracle.E(l1[0], 0, t, A[0] - B[0], a, b) * oracle.E(l1[1], 0, u, A[1] - B[1], a, b)
                            * oracle.E(l1[2], 0, v, A[2] - B[2], a, b) * oracle.R(t, u, v, 0, p, *(P - C), R))
        return val * 2 * np.pi / p

    for x in range(3):
        def f(s):
            Ap = A.copy()
            Ap[x] = s
            return ssv(Ap)
        d1 = (f(A[x] + 2e-3) - f(A[x] - 2e-3)) / 4e-3
        d2 = (f(A[x] + 1e-3) - f(A[x] - 1e-3)) / 2e-3
        l = [0, 0, 0]
        l[x] = 1
        assert (4 * d2 - d1) / 3 / (2 * a) == pytest.approx(prim(l), abs=1e-12)


def test_angular_momentum_from_overlap_derivatives(oracle):
    """RxDel (cython/onee.pyx:36-43, 139-192) for single primitives of every
    (l, l') <= d against the overlap:
      L_x(a, b) = -<a| (y - C_y) d/dz - (z - C_z) d/dy |b>,
      <a| (y - C_y) d/dz |b> = -d/dB_z [ <a y+1| b> + (A_y - C_y) <a|b> ]
    for unnormalised primitives (d/dz b(r - B) = -d/dB_z b), with the centre
    derivative taken by Richardson-extrapolated central differences of the
    oracle's overlap, and L antisymmetric."""
    A0 = np.array([0.3, -0.2, 0.5])
    B0 = np.array([-0.4, 0.6, 0.1])
    C = np.array([0.25, -0.15, 0.35])
    a, b = 0.9, 1.4
    shells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (0, 1, 2)]

    def ints(la, lb, B):
        m = _Mini(oracle, [(A0, la, [a], [1.0]), (B, lb, [b], [1.0])]).o
        m.center_of_charge = C
        S, T, V, M, L = m.one_electron()
        w = m.norm * m.coefs
        return S[0, 1] / (w[0] * w[1]), L[:, 0, 1] / (w[0] * w[1])

    def s_prim(la, lb, B):
        return ints(la, lb, B)[0]

    def d_dB(la, lb, x):
        def f(h):
            B = B0.copy()
            B[x] += h
            return s_prim(la, lb, B)
        d1 = (f(2e-3) - f(-2e-3)) / 4e-3
        d2 = (f(1e-3) - f(-1e-3)) / 2e-3
        return (4 * d2 - d1) / 3

    def r_grad(la, lb, y, z):
        # <a| (r_y - C_y) d/dr_z |b>
        up = list(la)
        up[y] += 1
        return -(d_dB(tuple(up), lb, z) + (A0[y] - C[y]) * d_dB(la, lb, z))

    for la in shells:
        for lb in shells:
            _, L = ints(la, lb, B0)
            for x, (y, z) in enumerate(((1, 2), (2, 0), (0, 1))):
                expect = -(r_grad(la, lb, y, z) - r_grad(la, lb, z, y))
                assert L[x] == pytest.approx(expect, abs=1e-9)
    # antisymmetry on a contracted cluster basis
    om = oracle.OracleMolecule(WATER, _gbs("cc-pvdz"))
    L = om.one_electron()[4]
    assert np.max(np.abs(L + L.transpose(0, 2, 1))) < 1e-13


# --------------------------------------------------------------------------
# ERIs
# --------------------------------------------------------------------------
def _prim_eri(oracle):
    import ctypes
    L = oracle.lib()
    L.orc_electron_repulsion.restype = ctypes.c_double
    pd = np.ctypeslib.ndpointer(dtype=np.float64)
    pi = np.ctypeslib.ndpointer(dtype=np.int64)
    d = ctypes.c_double
    L.orc_electron_repulsion.argtypes = [d, pi, pd, d, pi, pd, d, pi, pd, d, pi, pd]

    def f(a, la, A, b, lb, B, c, lc, C, dd, ld, D):
        i = lambda l: np.array(l, dtype=np.int64)
        return L.orc_electron_repulsion(a, i(la), A, b, i(lb), B, c, i(lc), C, dd, i(ld), D)
    return f


EX = (1.3, 0.7, 0.9, 2.1)
CEN = (np.array([0.1, -0.3, 0.2]), np.array([1.0, 0.4, -0.5]), np.array([-0.4, 0.9, 0.3]), np.array([0.5, 0.5, 1.5]))


def _ssss(oracle, cen):
    a, b, c, d = EX
    A, B, C, D = cen
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    al = p * q / (p + q)
    return (2 * np.pi ** 2.5 / (p * q * np.sqrt(p + q)) * np.exp(-a * b / p * np.sum((A - B) ** 2))
            * np.exp(-c * d / q * np.sum((C - D) ** 2)) * oracle.boys(0, al * np.sum((P - Q) ** 2)))


def test_eri_ssss_closed_form(oracle):
    f = _prim_eri(oracle)
    z = [0, 0, 0]
    a, b, c, d = EX
    A, B, C, D = CEN
    assert f(a, z, A, b, z, B, c, z, C, d, z, D) == pytest.approx(_ssss(oracle, CEN), rel=1e-14)


def test_eri_p_derivative_relations(oracle):
    """[p s|s s] etc. from d/dX of the closed-form (ss|ss): each of the four
    centres and three directions (sign, index and operand placement)."""
    f = _prim_eri(oracle)
    for w in range(4):
        for x in range(3):
            def g(s):
                cc = [v.copy() for v in CEN]
                cc[w][x] = s
                return _ssss(oracle, cc)
            c0 = CEN[w][x]
            d1 = (g(c0 + 2e-3) - g(c0 - 2e-3)) / 4e-3
            d2 = (g(c0 + 1e-3) - g(c0 - 1e-3)) / 2e-3
            ls = [[0, 0, 0] for _ in range(4)]
            ls[w][x] = 1
            v = f(EX[0], ls[0], CEN[0], EX[1], ls[1], CEN[1], EX[2], ls[2], CEN[2], EX[3], ls[3], CEN[3])
            assert (4 * d2 - d1) / 3 / (2 * EX[w]) == pytest.approx(v, abs=1e-12)


def test_eri_d_second_derivative(oracle):
    """[d_xy s|s s] = (1/4a^2) d^2/dA_x dA_y (ss|ss) (mixed second derivative)."""
    f = _prim_eri(oracle)
    a = EX[0]
    h = 1e-3

    def g(dx, dy):
        cc = [v.copy() for v in CEN]
        cc[0][0] += dx
        cc[0][1] += dy
        return _ssss(oracle, cc)
    mixed = lambda h: (g(h, h) - g(h, -h) - g(-h, h) + g(-h, -h)) / (4 * h * h)
    est = (4 * mixed(h) - mixed(2 * h)) / 3 / (4 * a * a)
    z = [0, 0, 0]
    v = f(a, [1, 1, 0], CEN[0], EX[1], z, CEN[1], EX[2], z, CEN[2], EX[3], z, CEN[3])
    assert est == pytest.approx(v, abs=1e-10)


def test_eri_tensor_symmetry_schwarz_psd(oracle):
    om = oracle.OracleMolecule(WATER, _gbs("sto-3g"))
    T = om.eri_tensor()
    N = om.nbasis
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.array_equal(T, T.transpose(perm))
    M = T.reshape(N * N, N * N)
    d = np.diag(M)
    assert np.all(np.abs(M) <= np.sqrt(np.outer(d, d)) * (1 + 1e-12) + 1e-15)
    assert np.min(np.linalg.eigvalsh(M)) > -1e-12
    p, q = np.tril_indices(N)
    assert np.allclose(om.schwarz(), T[p, q, p, q], rtol=0, atol=1e-15)


def test_eri_rotation_invariance_p(oracle):
    """Rotating the molecule rotates p components covariantly:
    (p_i s|p_j s)' = R_ia R_jb (p_a s|p_b s)."""
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    geo = [("O", 0.1, 0.2, -0.3), ("H", 0.9, -0.4, 0.2), ("O", -1.1, 0.7, 0.4)]

    def mol(atoms):
        g = "\n0 1\n" + "\n".join("%s %.15f %.15f %.15f" % a for a in atoms) + "\n"
        return oracle.OracleMolecule(g, _gbs("sto-3g"))
    m0 = mol(geo)
    m1 = mol([(s,) + tuple(Rz @ np.array([x, y, z])) for s, x, y, z in geo])
    # O sto-3g: s, s, px, py, pz -> functions 2..4 (first O) and 8..10 (second O)
    P1, P2 = [2, 3, 4], [8, 9, 10]
    s_h = 5
    idx = [(P1[a], s_h, P2[b], 0) for a in range(3) for b in range(3)]
    v0 = m0.eri_list(idx).reshape(3, 3)
    v1 = m1.eri_list(idx).reshape(3, 3)
    assert np.allclose(v1, Rz @ v0 @ Rz.T, atol=1e-13)


# --------------------------------------------------------------------------
# formPT and RHF
# --------------------------------------------------------------------------
def test_formpt_equals_2j_minus_k(oracle):
    """For real symmetric P: sym(formPT(P, 0)) == 2J - K from the tensor
    (mmd/scf.py:97-99); with tol = 0 nothing is screened."""
    om = oracle.OracleMolecule(WATER, _gbs("sto-3g"))
    T = om.eri_tensor()
    rng = np.random.default_rng(0)
    A = rng.standard_normal((om.nbasis,) * 2)
    P = 0.5 * (A + A.T)
    J = np.einsum("pqrs,sr->pq", T, P)
    K = np.einsum("psqr,sr->pq", T, P)
    G, nsurv = om.form_pt(P.astype(complex), np.zeros_like(P, complex), om.schwarz(), 0.0)
    assert nsurv == (28 * 29) // 2
    assert np.max(np.abs(0.5 * (G + G.T) - (2 * J - K))) < 1e-13
    # linearity in dP
    B = rng.standard_normal(P.shape)
    Q = 0.5 * (B + B.T)
    G1, _ = om.form_pt(P.astype(complex), np.zeros_like(P, complex), om.schwarz(), 0.0)
    G2, _ = om.form_pt(Q.astype(complex), np.zeros_like(P, complex), om.schwarz(), 0.0)
    G3, _ = om.form_pt((P + Q).astype(complex), np.zeros_like(P, complex), om.schwarz(), 0.0)
    assert np.max(np.abs(G1 + G2 - G3)) < 1e-13


def test_fock_elements_equal_formpt(oracle):
    """The elementwise restatement of the symmetrised formPT (used for the
    large-basis fixtures) equals the symmetrised orc_form_pt with the screen
    active (tol 1e-12, and 1e-6 where it drops quartets)."""
    om = oracle.OracleMolecule(WATER, _gbs("cc-pvdz"))
    N = om.nbasis
    rng = np.random.default_rng(4)
    A = rng.standard_normal((N, N))
    P = (A + A.T) * 1e-3
    S = oracle.schwarz_full(om)
    pq = [(p, q) for p in range(0, N, 3) for q in range(0, p + 1, 2)]
    for tol in (1e-12, 1e-6):
        G, nsurv = om.form_pt(P.astype(complex), np.zeros((N, N), complex), om.schwarz(), tol)
        Gs = 0.5 * (G + G.T).real
        e = oracle.fock_elements(om, P, S, tol, pq)
        ref = np.array([Gs[p, q] for p, q in pq])
        assert np.max(np.abs(e - ref)) < 1e-15
    assert nsurv < N * (N + 1) // 2 * (N * (N + 1) // 2 + 1) // 2


def test_form_pt_survivor_count(oracle):
    """The integral-free survivor count (used by the CPU baseline) equals the
    number of quartets orc_form_pt keeps, at thresholds that keep all, some
    and none; with tol = 0 it is every canonical quartet."""
    om = oracle.OracleMolecule(WATER, _gbs("6-31g"))
    N = om.nbasis
    rng = np.random.default_rng(8)
    A = rng.standard_normal((N, N))
    dP = (A + A.T) * 1e-4
    scr = om.schwarz()
    npair = N * (N + 1) // 2
    assert oracle.form_pt_survivors(dP, scr, 0.0) == npair * (npair + 1) // 2
    counts = []
    for tol in (1e-6, 1e-5, 1e-4, 1e3):
        _, nsurv = om.form_pt(dP.astype(complex), np.zeros((N, N), complex), scr, tol)
        assert oracle.form_pt_survivors(dP, scr, tol) == nsurv
        counts.append(nsurv)
    assert npair * (npair + 1) // 2 > counts[0] > counts[1] > counts[2] > counts[3] == 0


def test_rhf_water_sto3g_reference(oracle):
    """tests/test002.py (E = -74.942079928192, atol 1e-12) and the README
    dipole 1.53400931 D, at the geometry those values belong to."""
    om = oracle.OracleMolecule(from_bohr(WATER_BOHR), _gbs("sto-3g"))
    r = oracle.rhf(om, conver=1e-14)
    assert abs(r["energy"].real - (-74.942079928192)) < 1e-12
    assert abs(r["mu"][1].real - 1.53400931) < 5e-9
    r = oracle.rhf(om, direct=True, conver=1e-12)
    assert abs(r["energy"].real - (-74.942079928192)) < 1e-11


@pytest.mark.parametrize("geom,basis,ref,tol", [
    (H2, "sto-3g", -1.11675930740, 1e-11),                              # tests/test001.py
    (HELIUM, "cc-pvtz", -2.86115357403, 1e-11),                         # tests/test007.py
    (rescale(WATER, BOHR_1986), "dz", -75.977878975377, 1e-10),         # tests/test004.py
    (rescale(WATER, BOHR_1986), "3-21g", -75.5613125965, 5e-9),         # tests/test005.py
    (METHANE, "sto-3g", -39.726850324347, 2e-8),                        # tests/test003.py
    (FORMALDEHYDE, "sto-3g", -112.351590112, 5e-8),                     # tests/test008.py
])
def test_rhf_reference_energies(oracle, geom, basis, ref, tol):
    r = oracle.rhf(oracle.OracleMolecule(geom, _gbs(basis)), conver=1e-12)
    assert r["is_converged"]
    assert abs(r["energy"].real - ref) < tol


def test_rhf_direct_methane(oracle):
    """tests/test010.py: the direct (formPT) path, default conver."""
    r = oracle.rhf(oracle.OracleMolecule(METHANE, _gbs("sto-3g")), direct=True)
    assert abs(r["energy"].real - (-39.726850324347)) < 2e-8
    r = oracle.rhf(oracle.OracleMolecule(METHANE, _gbs("sto-3g")), direct=True, conver=1e-11)
    rc = oracle.rhf(oracle.OracleMolecule(METHANE, _gbs("sto-3g")), direct=False, conver=1e-11)
    assert abs(r["energy"].real - rc["energy"].real) < 1e-10


@pytest.mark.slow
def test_rhf_water_631ppgss_d_functions(oracle):
    """tests/test006.py (6-31++G**, Cartesian d): multi-centre d integrals."""
    r = oracle.rhf(oracle.OracleMolecule(WATER, _gbs("6-31ppgss")), conver=1e-10)
    assert abs(r["energy"].real - (-75.9924381487)) < 3e-8
    assert abs(r["mu"][1].real - 2.4046) < 1e-4


def test_rhf_invariants(oracle):
    om = oracle.OracleMolecule(WATER, _gbs("sto-3g"))
    r = oracle.rhf(om, conver=1e-12)
    P, S, F = r["P"].real, r["S"], r["F"].real
    assert np.allclose(P @ S @ P, P, atol=1e-10)
    assert np.trace(P @ S) == pytest.approx(om.nocc, abs=1e-10)
    assert np.max(np.abs(F @ P @ S - S @ P @ F)) < 1e-8


def test_cluster_generator():
    g = water_cluster(8, 7.0, 0)
    rows = [ln.split() for ln in g.strip().split("\n")[1:]]
    assert len(rows) == 24
    O = np.array([[float(x) for x in r[1:]] for r in rows if r[0] == "O"])
    d = np.linalg.norm(O[:, None] - O[None], axis=-1) + np.eye(8) * 10
    assert d.min() >= 2.7 and O.min() >= 0 and O.max() <= 7.0 + 1.2
    assert water_cluster(8, 7.0, 0) == g


# --------------------------------------------------------------------------
# Nuclear gradient (mmd/forces.py:8-97, cython/grad.pyx)
# --------------------------------------------------------------------------
WATER_BENT = """0 1
O 0.05 -0.07 0.01
H 0.86 0.62 -0.03
H -0.88 0.58 0.04
"""


def test_gradient_h2_reference(oracle):
    """H2 / STO-3G forces of tests/test001.py:15-18, with that test's
    assert_allclose(atol=1e-12) (and numpy's default rtol)."""
    from numpy.testing import assert_allclose
    om = oracle.OracleMolecule(H2, _gbs("sto-3g"))
    r = oracle.rhf(om)
    assert abs(r["energy"].real - (-1.11675930740)) < 1e-10
    forces = -om.rhf_gradient(r["P"], r["F"])
    ref = np.array([[0.000000000, -0.000000000, -0.027679601], [-0.000000000, 0.000000000, 0.027679601]])
    assert_allclose(forces, ref, atol=1e-12)


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g"])
def test_gradient_finite_differences(oracle, basis):
    """The analytic gradient of a bent, off-axis water against central
    differences of the converged energy (h = 1e-4 bohr; the O(h^2) error is
    about 1e-9), and its translational and rotational invariance."""
    om = oracle.OracleMolecule(WATER_BENT, _gbs(basis))
    r = oracle.rhf(om, conver=1e-11)
    g = om.rhf_gradient(r["P"], r["F"])
    assert np.max(np.abs(g.sum(0))) < 1e-12
    assert np.max(np.abs(np.cross(om.xyz, g).sum(0))) < 1e-12
    h = 1e-4
    for a, x in ((0, 0), (1, 1), (2, 2), (0, 1)):
        e = []
        for sgn in (1, -1):
            lines = WATER_BENT.strip().split("\n")
            parts = lines[1 + a].split()
            vals = [float(v) for v in parts[1:]]
            vals[x] += sgn * h * 0.52917721092
            lines[1 + a] = parts[0] + " " + " ".join("%.12f" % v for v in vals)
            e.append(oracle.rhf(oracle.OracleMolecule("\n".join(lines) + "\n", _gbs(basis)), conver=1e-11)["energy"].real)
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(g[a, x] - fd) < 1e-8, (a, x, g[a, x], fd)
