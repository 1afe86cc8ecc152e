/*
 * mmd_oracle.c -- CPU restatement of the McMurchie-Davidson integral path of
 * jjgoings/McMurchie-Davidson.  TEST INFRASTRUCTURE ONLY: this file is the
 * checker that the HIP product path is compared against.  Nothing in the
 * product (the mmd package, libmmd.so) links, loads or calls it.
 *
 * Each function follows the reference routine cited beside it and keeps the
 * reference's algorithm (recursive Hermite coefficients, recursive Hermite
 * Coulomb integrals, per-basis-function quartet loops).  It is deliberately
 * written in plain C, scalar and single-threaded, like the reference.
 *
 * Pins (tests/test_oracle.py):
 *   orc_boys          50 literature values of the reference's Boys test, F_m(0),
 *                     the erf closed form of F_0, the downward recursion.
 *   orc_E / orc_R /
 *   orc_electron_repulsion / orc_eri
 *                     closed-form (ss|ss), 8-fold symmetry, Schwarz inequality,
 *                     rotation/translation invariance, and every reference
 *                     RHF energy (H2, H2O, CH4, H2CO STO-3G, H2O DZ/3-21G/
 *                     6-31++G**, He cc-pVTZ).
 *   orc_normalize     S_ii = 1 and the S/T literals of the reference's
 *                     integral test.
 *   orc_form_pt       sym(formPT(P,0)) == 2J-K from the full tensor, linearity,
 *                     and the reference's direct-SCF methane energy.
 *   orc_fock_elements equal to the symmetrised orc_form_pt, element by element,
 *                     with screening active.
 *   one-electron      S/T literals, energies.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.141592653589793238462643383279;

/* ------------------------------------------------------------------ */
/* Double factorial with (-1)!! = 1 (the semantics the reference's     */
/* normalization relies on; cython/basis.pxi:101-110).                 */
/* ------------------------------------------------------------------ */
double orc_fact2(int n)
{
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= (double)k;
    return r;
}

/* ------------------------------------------------------------------ */
/* Boys function  F_m(T) = 1F1(m+1/2; m+3/2; -T)/(2m+1)                 */
/* (cython/util.pxi:53-55, SciPy hyp1f1).                              */
/* Restated through the lower incomplete gamma function:               */
/*   F_m(T) = gamma(a, T) / (2 T^a),  a = m + 1/2.                      */
/* For T < a + 1: the power series                                      */
/*   gamma(a,T) = T^a e^-T sum_k T^k / (a (a+1) ... (a+k)).            */
/* Otherwise Gamma(a) - Gamma(a,T), with the upper part from Legendre's */
/* continued fraction evaluated by the modified Lentz method.           */
/* Extended precision (long double) throughout.                         */
/* ------------------------------------------------------------------ */
double orc_boys(double m, double T)
{
    long double a = (long double)m + 0.5L;
    long double x = (long double)T;
    if (x <= 0.0L) return (double)(1.0L / (2.0L * a));
    if (x < a + 1.0L) {
        /* F = e^-x sum_k x^k / (2 a (a+1) ... (a+k)) */
        long double term = 1.0L / (2.0L * a);
        long double sum = term;
        for (int k = 1; k < 100000; ++k) {
            term *= x / (a + (long double)k);
            sum += term;
            if (term < sum * 1e-21L) break;
        }
        return (double)(expl(-x) * sum);
    }
    /* Upper incomplete gamma, continued fraction:
       Gamma(a,x) = e^-x x^a / (x+1-a- 1(1-a)/(x+3-a- 2(2-a)/(x+5-a- ...))) */
    const long double tiny = 1e-4000L;
    long double b = x + 1.0L - a;
    long double c = 1.0L / tiny;
    long double d = 1.0L / b;
    long double h = d;
    for (int i = 1; i < 100000; ++i) {
        long double an = -(long double)i * ((long double)i - a);
        b += 2.0L;
        d = an * d + b;
        if (fabsl(d) < tiny) d = tiny;
        c = b + an / c;
        if (fabsl(c) < tiny) c = tiny;
        d = 1.0L / d;
        long double del = d * c;
        h *= del;
        if (fabsl(del - 1.0L) < 1e-21L) break;
    }
    /* Q = Gamma(a,x)/Gamma(a) */
    long double lg = lgammal(a);
    long double Q = expl(-x + a * logl(x) - lg) * h;
    long double F = expl(lg - a * logl(x)) * 0.5L * (1.0L - Q);
    return (double)F;
}

/* ------------------------------------------------------------------ */
/* Hermite expansion coefficients E^{ij}_t (cython/util.pxi:10-28),     */
/* recursive, including the multipole branch n > 0 used by RxDel.       */
/* ------------------------------------------------------------------ */
double orc_E(int i, int j, int t, double Qx, double a, double b, int n, double Ax)
{
    double p = a + b;
    double u = a * b / p;
    if (n == 0) {
        if ((t < 0) || (t > (i + j))) return 0.0;
        if (i == 0 && j == 0 && t == 0) return exp(-u * Qx * Qx);
        if (j == 0)
            return (1.0 / (2.0 * p)) * orc_E(i - 1, j, t - 1, Qx, a, b, 0, 0.0)
                 - (u * Qx / a) * orc_E(i - 1, j, t, Qx, a, b, 0, 0.0)
                 + (t + 1) * orc_E(i - 1, j, t + 1, Qx, a, b, 0, 0.0);
        return (1.0 / (2.0 * p)) * orc_E(i, j - 1, t - 1, Qx, a, b, 0, 0.0)
             + (u * Qx / b) * orc_E(i, j - 1, t, Qx, a, b, 0, 0.0)
             + (t + 1) * orc_E(i, j - 1, t + 1, Qx, a, b, 0, 0.0);
    }
    return orc_E(i + 1, j, t, Qx, a, b, n - 1, Ax) + Ax * orc_E(i, j, t, Qx, a, b, n - 1, Ax);
}

/* ------------------------------------------------------------------ */
/* Hermite Coulomb integrals R^n_{tuv} (cython/util.pxi:30-50).          */
/* ------------------------------------------------------------------ */
double orc_R(int t, int u, int v, int n, double p, double PCx, double PCy, double PCz, double RPC)
{
    double T = p * RPC * RPC;
    double val = 0.0;
    if (t == 0 && u == 0 && v == 0) {
        val += pow(-2.0 * p, n) * orc_boys((double)n, T);
    } else if (t == 0 && u == 0) {
        if (v > 1) val += (v - 1) * orc_R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC);
        val += PCz * orc_R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC);
    } else if (t == 0) {
        if (u > 1) val += (u - 1) * orc_R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC);
        val += PCy * orc_R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC);
    } else {
        if (t > 1) val += (t - 1) * orc_R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC);
        val += PCx * orc_R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC);
    }
    return val;
}

/* ------------------------------------------------------------------ */
/* Primitive [ab|cd] (cython/twoe.pyx:52-94).                           */
/* ------------------------------------------------------------------ */
double orc_electron_repulsion(double a, const int64_t *lmn1, const double *A,
                              double b, const int64_t *lmn2, const double *B,
                              double c, const int64_t *lmn3, const double *C,
                              double d, const int64_t *lmn4, const double *D)
{
    long l1 = lmn1[0], m1 = lmn1[1], n1 = lmn1[2];
    long l2 = lmn2[0], m2 = lmn2[1], n2 = lmn2[2];
    long l3 = lmn3[0], m3 = lmn3[1], n3 = lmn3[2];
    long l4 = lmn4[0], m4 = lmn4[1], n4 = lmn4[2];
    double p = a + b;
    double q = c + d;
    double alpha = p * q / (p + q);
    double Px = (a * A[0] + b * B[0]) / p;
    double Py = (a * A[1] + b * B[1]) / p;
    double Pz = (a * A[2] + b * B[2]) / p;
    double Qx = (c * C[0] + d * D[0]) / q;
    double Qy = (c * C[1] + d * D[1]) / q;
    double Qz = (c * C[2] + d * D[2]) / q;
    double RPQ = sqrt(pow(Px - Qx, 2) + pow(Py - Qy, 2) + pow(Pz - Qz, 2));
    double val = 0.0;
    for (long t = 0; t < l1 + l2 + 1; ++t)
        for (long u = 0; u < m1 + m2 + 1; ++u)
            for (long v = 0; v < n1 + n2 + 1; ++v)
                for (long tau = 0; tau < l3 + l4 + 1; ++tau)
                    for (long nu = 0; nu < m3 + m4 + 1; ++nu)
                        for (long phi = 0; phi < n3 + n4 + 1; ++phi)
                            val += orc_E(l1, l2, t, A[0] - B[0], a, b, 0, 0.0)
                                 * orc_E(m1, m2, u, A[1] - B[1], a, b, 0, 0.0)
                                 * orc_E(n1, n2, v, A[2] - B[2], a, b, 0, 0.0)
                                 * orc_E(l3, l4, tau, C[0] - D[0], c, d, 0, 0.0)
                                 * orc_E(m3, m4, nu, C[1] - D[1], c, d, 0, 0.0)
                                 * orc_E(n3, n4, phi, C[2] - D[2], c, d, 0, 0.0)
                                 * pow(-1.0, (double)(tau + nu + phi))
                                 * orc_R(t + tau, u + nu, v + phi, 0, alpha,
                                         Px - Qx, Py - Qy, Pz - Qz, RPQ);
    val *= 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q));
    return val;
}

/* ------------------------------------------------------------------ */
/* Basis-function set.  Arrays are laid out per basis function (BF):    */
/*   origin[3*i], shell[3*i] (l,m,n), poff[i] .. poff[i]+np[i] into     */
/*   exps/coefs/norm.  coefs are the contraction-normalised values,     */
/*   norm the primitive norms, exactly as cython/basis.pxi stores them. */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *origin;
    const int64_t *shell;
    const int64_t *np;
    const int64_t *poff;
    const double *exps;
    const double *coefs;
    const double *norm;
} orc_bfs;

/* Normalisation of one basis function (cython/basis.pxi:87-120). */
void orc_normalize(int64_t l, int64_t m, int64_t n, int64_t K,
                   const double *exps, double *coefs, double *norm)
{
    int64_t L = l + m + n;
    for (int64_t ia = 0; ia < K; ++ia)
        norm[ia] = sqrt(pow(2.0, 2.0 * (double)L + 1.5) * pow(exps[ia], (double)L + 1.5)
                        / orc_fact2((int)(2 * l - 1)) / orc_fact2((int)(2 * m - 1))
                        / orc_fact2((int)(2 * n - 1)) / pow(PI, 1.5));
    double prefactor = pow(PI, 1.5) * orc_fact2((int)(2 * l - 1)) * orc_fact2((int)(2 * m - 1))
                     * orc_fact2((int)(2 * n - 1)) / pow(2.0, (double)L);
    double N = 0.0;
    for (int64_t ia = 0; ia < K; ++ia)
        for (int64_t ib = 0; ib < K; ++ib)
            N += norm[ia] * norm[ib] * coefs[ia] * coefs[ib] / pow(exps[ia] + exps[ib], (double)L + 1.5);
    N *= prefactor;
    N = pow(N, -0.5);
    for (int64_t ia = 0; ia < K; ++ia) coefs[ia] *= N;
}

/* Contracted (ab|cd) (cython/twoe.pyx:34-50). */
static double eri_bf(const orc_bfs *B, int64_t i, int64_t j, int64_t k, int64_t l)
{
    double eri = 0.0;
    int64_t oi = B->poff[i], oj = B->poff[j], ok = B->poff[k], ol = B->poff[l];
    for (int64_t ja = 0; ja < B->np[i]; ++ja)
        for (int64_t jb = 0; jb < B->np[j]; ++jb)
            for (int64_t jc = 0; jc < B->np[k]; ++jc)
                for (int64_t jd = 0; jd < B->np[l]; ++jd)
                    eri += B->norm[oi + ja] * B->norm[oj + jb] * B->norm[ok + jc] * B->norm[ol + jd]
                         * B->coefs[oi + ja] * B->coefs[oj + jb] * B->coefs[ok + jc] * B->coefs[ol + jd]
                         * orc_electron_repulsion(B->exps[oi + ja], B->shell + 3 * i, B->origin + 3 * i,
                                                  B->exps[oj + jb], B->shell + 3 * j, B->origin + 3 * j,
                                                  B->exps[ok + jc], B->shell + 3 * k, B->origin + 3 * k,
                                                  B->exps[ol + jd], B->shell + 3 * l, B->origin + 3 * l);
    return eri;
}

#define ORC_BFS_ARGS const double *origin, const int64_t *shell, const int64_t *np, \
                     const int64_t *poff, const double *exps, const double *coefs, const double *norm
#define ORC_BFS_INIT orc_bfs B = {origin, shell, np, poff, exps, coefs, norm}

/* Contracted ERIs for a list of BF quartets (ERI(), twoe.pyx:36-50). */
void orc_eri_list(ORC_BFS_ARGS, int64_t n, const int64_t *ijkl, double *out)
{
    ORC_BFS_INIT;
    for (int64_t q = 0; q < n; ++q)
        out[q] = eri_bf(&B, ijkl[4 * q], ijkl[4 * q + 1], ijkl[4 * q + 2], ijkl[4 * q + 3]);
}

/* Full tensor with 8-fold scatter (doERIs, cython/twoe.pyx:9-31). */
void orc_do_eris(ORC_BFS_ARGS, int64_t N, double *T)
{
    ORC_BFS_INIT;
#define T4(a, b, c, d) T[(((a) * N + (b)) * N + (c)) * N + (d)]
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            int64_t ij = i * (i + 1) / 2 + j;
            for (int64_t k = 0; k < N; ++k)
                for (int64_t l = 0; l <= k; ++l) {
                    int64_t kl = k * (k + 1) / 2 + l;
                    if (ij >= kl) {
                        double val = eri_bf(&B, i, j, k, l);
                        T4(i, j, k, l) = val; T4(k, l, i, j) = val;
                        T4(j, i, l, k) = val; T4(l, k, j, i) = val;
                        T4(j, i, k, l) = val; T4(l, k, i, j) = val;
                        T4(i, j, l, k) = val; T4(k, l, j, i) = val;
                    }
                }
        }
#undef T4
}

/* Schwarz diagonal screen[pq] = (pq|pq), pq = p(p+1)/2+q (mmd/molecule.py:93-99). */
void orc_schwarz(ORC_BFS_ARGS, int64_t N, double *screen)
{
    ORC_BFS_INIT;
    for (int64_t p = 0; p < N; ++p)
        for (int64_t q = 0; q <= p; ++q)
            screen[p * (p + 1) / 2 + q] = eri_bf(&B, p, q, p, q);
}

/* Direct incremental G(dP) (formPT, cython/fock.pyx:10-87).
 * dP = P - P_old is given as real and imaginary parts (dPi may be NULL);
 * G (re/im) receives the unsymmetrised result, as formPT returns it.
 * `limit` (<0: all) stops after that many visited quartets, used for a
 * bounded timing sample; the return value is the number of quartets that
 * survived screening. */
int64_t orc_form_pt(ORC_BFS_ARGS, int64_t N, const double *dPr, const double *dPi,
                    const double *screen, double tol, double *Gr, double *Gi,
                    int64_t limit, int64_t start_ij)
{
    ORC_BFS_INIT;
    int64_t survivors = 0, visited = 0;
    memset(Gr, 0, sizeof(double) * N * N);
    if (Gi) memset(Gi, 0, sizeof(double) * N * N);
#define M(A, a, b) ((A)[(a) * N + (b)])
#define CABS(a, b) (dPi ? hypot(M(dPr, a, b), M(dPi, a, b)) : fabs(M(dPr, a, b)))
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            int64_t ij = i * (i + 1) / 2 + j;
            if (ij < start_ij) continue;
            for (int64_t k = 0; k < N; ++k)
                for (int64_t l = 0; l <= k; ++l) {
                    int64_t kl = k * (k + 1) / 2 + l;
                    if (ij >= kl) {
                        if (limit >= 0 && visited >= limit) return survivors;
                        ++visited;
                        double bound = sqrt(screen[ij]) * sqrt(screen[kl]);
                        double dm = 4.0 * CABS(i, j);
                        double x;
                        x = 4.0 * CABS(k, l); if (x > dm) dm = x;
                        x = CABS(i, k); if (x > dm) dm = x;
                        x = CABS(i, l); if (x > dm) dm = x;
                        x = CABS(j, k); if (x > dm) dm = x;
                        x = CABS(j, l); if (x > dm) dm = x;
                        bound *= dm;
                        if (bound < tol) continue;
                        ++survivors;
                        double s12 = (i == j) ? 1.0 : 2.0;
                        double s34 = (k == l) ? 1.0 : 2.0;
                        double s1234 = (i == k) ? ((j == l) ? 1.0 : 2.0) : 2.0;
                        double eri = s12 * s34 * s1234 * eri_bf(&B, i, j, k, l);
                        M(Gr, i, j) += M(dPr, k, l) * eri;
                        M(Gr, k, l) += M(dPr, i, j) * eri;
                        M(Gr, i, k) += -0.25 * M(dPr, j, l) * eri;
                        M(Gr, j, l) += -0.25 * M(dPr, i, k) * eri;
                        M(Gr, i, l) += -0.25 * M(dPr, j, k) * eri;
                        M(Gr, k, j) += -0.25 * M(dPr, i, l) * eri;
                        if (dPi && Gi) {
                            M(Gi, i, j) += M(dPi, k, l) * eri;
                            M(Gi, k, l) += M(dPi, i, j) * eri;
                            M(Gi, i, k) += -0.25 * M(dPi, j, l) * eri;
                            M(Gi, j, l) += -0.25 * M(dPi, i, k) * eri;
                            M(Gi, i, l) += -0.25 * M(dPi, j, k) * eri;
                            M(Gi, k, j) += -0.25 * M(dPi, i, l) * eri;
                        }
                    }
                }
        }
#undef CABS
#undef M
    return survivors;
}


/* The number of quartets formPT's screen keeps for a real dP, without
 * computing any integral: the loop and test of orc_form_pt
 * (cython/fock.pyx:38-57).  Used to state the survivors of each iteration
 * next to the timed formPT samples of the CPU baseline. */
int64_t orc_form_pt_survivors(int64_t N, const double *dPr, const double *screen, double tol)
{
    int64_t survivors = 0;
#define M(A, a, b) ((A)[(a) * N + (b)])
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            const int64_t ij = i * (i + 1) / 2 + j;
            const double sij = sqrt(screen[ij]);
            const double pij = 4.0 * fabs(M(dPr, i, j));
            for (int64_t k = 0; k <= i; ++k)
                for (int64_t l = 0; l <= k; ++l) {
                    const int64_t kl = k * (k + 1) / 2 + l;
                    if (kl > ij) break;
                    double dm = pij, x;
                    x = 4.0 * fabs(M(dPr, k, l)); if (x > dm) dm = x;
                    x = fabs(M(dPr, i, k)); if (x > dm) dm = x;
                    x = fabs(M(dPr, i, l)); if (x > dm) dm = x;
                    x = fabs(M(dPr, j, k)); if (x > dm) dm = x;
                    x = fabs(M(dPr, j, l)); if (x > dm) dm = x;
                    if (sij * sqrt(screen[kl]) * dm >= tol) ++survivors;
                }
        }
#undef M
    return survivors;
}

/* Selected elements of the symmetrised direct Fock contribution
 * 0.5 (G + G^T) with G = formPT(P, 0) (cython/fock.pyx:38-85 followed by
 * mmd/scf.py:93) for a real symmetric P, one element at a time:
 *   G_pq = sum_rs [ 2 m(pq,rs) (pq|rs) - m(pr,qs) (pr|qs) ] P_rs
 * where m(ij,kl) = 1 when formPT keeps the canonical quartet of (ij|kl):
 *   sqrt(S_ij) sqrt(S_kl) max(4|P_ij|, 4|P_kl|, |P_ik|, |P_il|, |P_jk|, |P_jl|) >= tol
 * (S = full N*N Schwarz diagonal).  The test is invariant under the eight
 * index permutations for symmetric P, so every member of an orbit carries its
 * canonical quartet's decision; formPT's degeneracy factors and -1/4
 * exchange weights then sum to the expression above.  Costs about 1.5 N^2
 * contracted ERIs per element, so a few elements of a large basis can be
 * checked without the N^4 loop.  pq: 2*n indices; out: n values. */
static int keep_quartet(const double *S, const double *P, int64_t N, double tol, int64_t i, int64_t j,
                        int64_t k, int64_t l)
{
#define M(A, a, b) ((A)[(a) * N + (b)])
    double bound = sqrt(M(S, i, j)) * sqrt(M(S, k, l));
    double dm = 4.0 * fabs(M(P, i, j)), x;
    x = 4.0 * fabs(M(P, k, l)); if (x > dm) dm = x;
    x = fabs(M(P, i, k)); if (x > dm) dm = x;
    x = fabs(M(P, i, l)); if (x > dm) dm = x;
    x = fabs(M(P, j, k)); if (x > dm) dm = x;
    x = fabs(M(P, j, l)); if (x > dm) dm = x;
#undef M
    return bound * dm >= tol;
}

void orc_fock_elements(ORC_BFS_ARGS, int64_t N, const double *P, const double *S, double tol, int64_t n,
                       const int64_t *pq, double *out)
{
    ORC_BFS_INIT;
    for (int64_t e = 0; e < n; ++e) {
        const int64_t p = pq[2 * e], q = pq[2 * e + 1];
        double g = 0.0;
        for (int64_t r = 0; r < N; ++r)
            for (int64_t s = 0; s <= r; ++s) {
                if (!keep_quartet(S, P, N, tol, p, q, r, s)) continue;
                const double w = (r == s ? 1.0 : 2.0) * P[r * N + s];
                if (w != 0.0) g += 2.0 * w * eri_bf(&B, p, q, r, s);
            }
        for (int64_t r = 0; r < N; ++r)
            for (int64_t s = 0; s < N; ++s) {
                if (P[r * N + s] == 0.0 || !keep_quartet(S, P, N, tol, p, r, q, s)) continue;
                g -= P[r * N + s] * eri_bf(&B, p, r, q, s);
            }
        out[e] = g;
    }
}

/* ------------------------------------------------------------------ */
/* One-electron primitives (cython/onee.pyx:71-192).                    */
/* ------------------------------------------------------------------ */
static double prim_overlap(double a, const int64_t *l1, const double *A,
                           double b, const int64_t *l2, const double *B)
{
    double S1 = orc_E((int)l1[0], (int)l2[0], 0, A[0] - B[0], a, b, 0, 0.0);
    double S2 = orc_E((int)l1[1], (int)l2[1], 0, A[1] - B[1], a, b, 0, 0.0);
    double S3 = orc_E((int)l1[2], (int)l2[2], 0, A[2] - B[2], a, b, 0, 0.0);
    return S1 * S2 * S3 * pow(PI / (a + b), 1.5);
}

static double prim_dipole(double a, const int64_t *l1, const double *A,
                          double b, const int64_t *l2, const double *B,
                          const double *C, int dir)
{
    double p = a + b;
    double P[3];
    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * B[x]) / p;
    double f[3];
    for (int x = 0; x < 3; ++x) {
        if (x == dir)
            f[x] = orc_E((int)l1[x], (int)l2[x], 1, A[x] - B[x], a, b, 0, 0.0)
                 + (P[x] - C[x]) * orc_E((int)l1[x], (int)l2[x], 0, A[x] - B[x], a, b, 0, 0.0);
        else
            f[x] = orc_E((int)l1[x], (int)l2[x], 0, A[x] - B[x], a, b, 0, 0.0);
    }
    return f[0] * f[1] * f[2] * pow(PI / p, 1.5);
}

static double prim_kinetic(double a, const int64_t *l1, const double *A,
                           double b, const int64_t *l2, const double *B)
{
    double Tx3[3];
    double E0[3];
    for (int x = 0; x < 3; ++x)
        E0[x] = orc_E((int)l1[x], (int)l2[x], 0, A[x] - B[x], a, b, 0, 0.0);
    for (int x = 0; x < 3; ++x) {
        double cA = (2.0 * (double)l2[x] + 1.0) * b;
        double cB = -2.0 * pow(b, 2);
        double cC = -0.5 * (double)l2[x] * ((double)l2[x] - 1.0);
        double t = cA * orc_E((int)l1[x], (int)l2[x], 0, A[x] - B[x], a, b, 0, 0.0)
                 + cB * orc_E((int)l1[x], (int)l2[x] + 2, 0, A[x] - B[x], a, b, 0, 0.0)
                 + cC * orc_E((int)l1[x], (int)l2[x] - 2, 0, A[x] - B[x], a, b, 0, 0.0);
        Tx3[x] = t * E0[(x + 1) % 3] * E0[(x + 2) % 3];
    }
    return (Tx3[0] + Tx3[1] + Tx3[2]) * pow(PI / (a + b), 1.5);
}

static double prim_angular(double a, const int64_t *l1, const double *A,
                           double b, const int64_t *l2, const double *B,
                           const double *C, int dir)
{
    double S0[3], S1[3], D1[3];
    for (int x = 0; x < 3; ++x) {
        int i = (int)l1[x], j = (int)l2[x];
        double Q = A[x] - B[x];
        S0[x] = orc_E(i, j, 0, Q, a, b, 0, 0.0);
        S1[x] = orc_E(i, j, 0, Q, a, b, 1, A[x] - C[x]);
        D1[x] = j * orc_E(i, j - 1, 0, Q, a, b, 0, 0.0) - 2.0 * b * orc_E(i, j + 1, 0, Q, a, b, 0, 0.0);
    }
    double pre = pow(PI / (a + b), 1.5);
    if (dir == 0) return -S0[0] * (S1[1] * D1[2] - S1[2] * D1[1]) * pre;
    if (dir == 1) return -S0[1] * (S1[2] * D1[0] - S1[0] * D1[2]) * pre;
    return -S0[2] * (S1[0] * D1[1] - S1[1] * D1[0]) * pre;
}

static double prim_nuclear(double a, const int64_t *l1, const double *A,
                           double b, const int64_t *l2, const double *B, const double *C)
{
    double p = a + b;
    double P[3];
    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * B[x]) / p;
    double RPC = sqrt((P[0] - C[0]) * (P[0] - C[0]) + (P[1] - C[1]) * (P[1] - C[1])
                      + (P[2] - C[2]) * (P[2] - C[2]));
    double val = 0.0;
    for (int t = 0; t < l1[0] + l2[0] + 1; ++t)
        for (int u = 0; u < l1[1] + l2[1] + 1; ++u)
            for (int v = 0; v < l1[2] + l2[2] + 1; ++v)
                val += orc_E((int)l1[0], (int)l2[0], t, A[0] - B[0], a, b, 0, 0.0)
                     * orc_E((int)l1[1], (int)l2[1], u, A[1] - B[1], a, b, 0, 0.0)
                     * orc_E((int)l1[2], (int)l2[2], v, A[2] - B[2], a, b, 0, 0.0)
                     * orc_R(t, u, v, 0, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], RPC);
    val *= 2.0 * PI / p;
    return val;
}

/* Contracted one-electron matrices as assembled by
 * Molecule.one_electron_integrals (mmd/molecule.py:235-286) from
 * S/T/V/Mu/RxDel (cython/onee.pyx:14-69).
 *   S,T,V: N*N;  M, L: 3*N*N;  atoms: natom charges Z and xyz (bohr);
 *   cc: centre of nuclear charge. */
void orc_one_electron(ORC_BFS_ARGS, int64_t N, int64_t natom, const double *Z,
                      const double *axyz, const double *cc,
                      double *S, double *T, double *V, double *Mx, double *Lx)
{
    ORC_BFS_INIT;
    (void)B;
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            double s = 0, t = 0, v = 0, mu[3] = {0, 0, 0}, lam[3] = {0, 0, 0};
            int64_t oi = poff[i], oj = poff[j];
            for (int64_t ia = 0; ia < np[i]; ++ia)
                for (int64_t ib = 0; ib < np[j]; ++ib) {
                    double a = exps[oi + ia], b = exps[oj + ib];
                    double w = norm[oi + ia] * norm[oj + ib] * coefs[oi + ia] * coefs[oj + ib];
                    const int64_t *l1 = shell + 3 * i, *l2 = shell + 3 * j;
                    const double *A = origin + 3 * i, *Bc = origin + 3 * j;
                    s += w * prim_overlap(a, l1, A, b, l2, Bc);
                    t += w * prim_kinetic(a, l1, A, b, l2, Bc);
                    for (int x = 0; x < 3; ++x) {
                        mu[x] += w * prim_dipole(a, l1, A, b, l2, Bc, cc, x);
                        lam[x] += w * prim_angular(a, l1, A, b, l2, Bc, cc, x);
                    }
                }
            /* V summed over atoms, each contracted separately as in
               mmd/molecule.py:265-266 */
            for (int64_t at = 0; at < natom; ++at) {
                double vat = 0.0;
                for (int64_t ia = 0; ia < np[i]; ++ia)
                    for (int64_t ib = 0; ib < np[j]; ++ib) {
                        double a = exps[oi + ia], b = exps[oj + ib];
                        double w = norm[oi + ia] * norm[oj + ib] * coefs[oi + ia] * coefs[oj + ib];
                        vat += w * prim_nuclear(a, shell + 3 * i, origin + 3 * i,
                                                b, shell + 3 * j, origin + 3 * j, axyz + 3 * at);
                    }
                v += -Z[at] * vat;
            }
            S[i * N + j] = S[j * N + i] = s;
            T[i * N + j] = T[j * N + i] = t;
            V[i * N + j] = V[j * N + i] = v;
            for (int x = 0; x < 3; ++x) {
                Mx[x * N * N + i * N + j] = Mx[x * N * N + j * N + i] = mu[x];
                Lx[x * N * N + i * N + j] = lam[x];
                Lx[x * N * N + j * N + i] = -lam[x];
            }
        }
}

/* ------------------------------------------------------------------ */
/* Nuclear gradient of the RHF energy (mmd/forces.py:8-97 with the      */
/* derivative integrals of cython/grad.pyx:11-350).                     */
/* ------------------------------------------------------------------ */
/* First derivative of a 1-D Hermite coefficient with respect to the    */
/* centre of the first (q = 1) or second (r = 1) Gaussian               */
/* (cython/grad.pyx:97-103).                                            */
static double orc_Ex(int i, int j, int t, double Q, double a, double b, int q, int r)
{
    if (q == 1) return 2.0 * a * orc_E(i + 1, j, t, Q, a, b, 0, 0.0) - i * orc_E(i - 1, j, t, Q, a, b, 0, 0.0);
    if (r == 1) return 2.0 * b * orc_E(i, j + 1, t, Q, a, b, 0, 0.0) - j * orc_E(i, j - 1, t, Q, a, b, 0, 0.0);
    return orc_E(i, j, t, Q, a, b, 0, 0.0);
}

/* 1-D coefficient along axis x of a primitive pair: plain, or the      */
/* derivative for the centre `ctr` (1 = first, 2 = second) when x is    */
/* the differentiated axis `dx`.                                        */
static double coef1(int x, int dx, int ctr, const int64_t *l1, const int64_t *l2, int dj, int t,
                    const double *A, const double *B, double a, double b)
{
    int i = (int)l1[x], j = (int)l2[x] + dj;
    if (j < 0) return 0.0;
    if (x == dx && ctr == 1) return orc_Ex(i, j, t, A[x] - B[x], a, b, 1, 0);
    if (x == dx && ctr == 2) return orc_Ex(i, j, t, A[x] - B[x], a, b, 0, 1);
    return orc_E(i, j, t, A[x] - B[x], a, b, 0, 0.0);
}

/* overlapX (cython/grad.pyx:352-398) */
static double prim_overlap_x(double a, const int64_t *l1, const double *A, double b, const int64_t *l2,
                             const double *B, int dx, int ctr)
{
    double s = 1.0;
    for (int x = 0; x < 3; ++x) s *= coef1(x, dx, ctr, l1, l2, 0, 0, A, B, a, b);
    return s * pow(PI / (a + b), 1.5);
}

/* kineticX (cython/grad.pyx:400-507) */
static double prim_kinetic_x(double a, const int64_t *l1, const double *A, double b, const int64_t *l2,
                             const double *B, int dx, int ctr)
{
    double Tk[3], S0[3];
    for (int x = 0; x < 3; ++x) {
        double j = (double)l2[x];
        Tk[x] = (2.0 * j + 1.0) * b * coef1(x, dx, ctr, l1, l2, 0, 0, A, B, a, b)
              - 2.0 * b * b * coef1(x, dx, ctr, l1, l2, 2, 0, A, B, a, b)
              - 0.5 * j * (j - 1.0) * coef1(x, dx, ctr, l1, l2, -2, 0, A, B, a, b);
        S0[x] = coef1(x, dx, ctr, l1, l2, 0, 0, A, B, a, b);
    }
    return (Tk[0] * S0[1] * S0[2] + Tk[1] * S0[0] * S0[2] + Tk[2] * S0[0] * S0[1]) * pow(PI / (a + b), 1.5);
}

/* nuclear_attractionXa (operator derivative, cython/grad.pyx:509-551) and
 * nuclear_attractionXb (basis-function derivative, 553-): both without
 * the nuclear charge. */
static double prim_nuclear_xa(double a, const int64_t *l1, const double *A, double b, const int64_t *l2,
                              const double *B, const double *C, int dx)
{
    double p = a + b, P[3];
    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * B[x]) / p;
    double RPC = sqrt((P[0] - C[0]) * (P[0] - C[0]) + (P[1] - C[1]) * (P[1] - C[1]) + (P[2] - C[2]) * (P[2] - C[2]));
    double val = 0.0;
    for (int t = 0; t <= l1[0] + l2[0]; ++t)
        for (int u = 0; u <= l1[1] + l2[1]; ++u)
            for (int v = 0; v <= l1[2] + l2[2]; ++v)
                val -= orc_E((int)l1[0], (int)l2[0], t, A[0] - B[0], a, b, 0, 0.0)
                     * orc_E((int)l1[1], (int)l2[1], u, A[1] - B[1], a, b, 0, 0.0)
                     * orc_E((int)l1[2], (int)l2[2], v, A[2] - B[2], a, b, 0, 0.0)
                     * orc_R(t + (dx == 0), u + (dx == 1), v + (dx == 2), 0, p, P[0] - C[0], P[1] - C[1],
                             P[2] - C[2], RPC);
    return val * 2.0 * PI / p;
}

static double prim_nuclear_xb(double a, const int64_t *l1, const double *A, double b, const int64_t *l2,
                              const double *B, const double *C, int dx, int ctr)
{
    double p = a + b, P[3];
    for (int x = 0; x < 3; ++x) P[x] = (a * A[x] + b * B[x]) / p;
    double RPC = sqrt((P[0] - C[0]) * (P[0] - C[0]) + (P[1] - C[1]) * (P[1] - C[1]) + (P[2] - C[2]) * (P[2] - C[2]));
    double val = 0.0;
    for (int t = 0; t <= l1[0] + l2[0] + (dx == 0); ++t)
        for (int u = 0; u <= l1[1] + l2[1] + (dx == 1); ++u)
            for (int v = 0; v <= l1[2] + l2[2] + (dx == 2); ++v)
                val += coef1(0, dx, ctr, l1, l2, 0, t, A, B, a, b) * coef1(1, dx, ctr, l1, l2, 0, u, A, B, a, b)
                     * coef1(2, dx, ctr, l1, l2, 0, v, A, B, a, b)
                     * orc_R(t, u, v, 0, p, P[0] - C[0], P[1] - C[1], P[2] - C[2], RPC);
    return val * 2.0 * PI / p;
}

/* electron_repulsionX (cython/grad.pyx:107-350): the derivative of the
 * primitive [ab|cd] with respect to centre `ctr` (0..3 = a..d) along dx. */
static double prim_eri_x(double ea, const int64_t *l1, const double *A, double eb, const int64_t *l2,
                         const double *B, double ec, const int64_t *l3, const double *C, double ed,
                         const int64_t *l4, const double *D, int dx, int ctr)
{
    double p = ea + eb, q = ec + ed, alpha = p * q / (p + q);
    double P[3], Q[3];
    for (int x = 0; x < 3; ++x) {
        P[x] = (ea * A[x] + eb * B[x]) / p;
        Q[x] = (ec * C[x] + ed * D[x]) / q;
    }
    double RPQ = sqrt((P[0] - Q[0]) * (P[0] - Q[0]) + (P[1] - Q[1]) * (P[1] - Q[1]) + (P[2] - Q[2]) * (P[2] - Q[2]));
    int cb = ctr == 0 ? 1 : (ctr == 1 ? 2 : 0);   /* bra centre derivative */
    int ck = ctr == 2 ? 1 : (ctr == 3 ? 2 : 0);   /* ket centre derivative */
    int nb[3], nk[3];
    for (int x = 0; x < 3; ++x) {
        nb[x] = (int)(l1[x] + l2[x]) + (cb && x == dx);
        nk[x] = (int)(l3[x] + l4[x]) + (ck && x == dx);
    }
    double val = 0.0;
    for (int t = 0; t <= nb[0]; ++t)
        for (int u = 0; u <= nb[1]; ++u)
            for (int v = 0; v <= nb[2]; ++v) {
                double eb3 = coef1(0, dx, cb, l1, l2, 0, t, A, B, ea, eb) * coef1(1, dx, cb, l1, l2, 0, u, A, B, ea, eb)
                           * coef1(2, dx, cb, l1, l2, 0, v, A, B, ea, eb);
                if (eb3 == 0.0) continue;
                for (int tau = 0; tau <= nk[0]; ++tau)
                    for (int nu = 0; nu <= nk[1]; ++nu)
                        for (int phi = 0; phi <= nk[2]; ++phi)
                            val += eb3 * coef1(0, dx, ck, l3, l4, 0, tau, C, D, ec, ed)
                                 * coef1(1, dx, ck, l3, l4, 0, nu, C, D, ec, ed)
                                 * coef1(2, dx, ck, l3, l4, 0, phi, C, D, ec, ed)
                                 * (((tau + nu + phi) & 1) ? -1.0 : 1.0)
                                 * orc_R(t + tau, u + nu, v + phi, 0, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2],
                                         RPQ);
            }
    return val * 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q));
}

/* dE/dX of the converged RHF energy, natom x 3 (the reference stores
 * -dE/dX as the force, mmd/forces.py:93-97):
 *   sum_pq P_qp (2 Hx + Gx)_pq - 2 sum_pq Sx_pq W_qp + dV_NN/dX,
 * Gx = 2 Jx - Kx from the derivative ERIs, W = P F P (forces.py:81-91).
 * The derivative tensor is never stored: each canonical quartet's
 * derivatives (the four centres' terms of forces.py:68-71) are weighted by
 * the density products of the positions it would fill in dTwoEx
 * (forces.py:73-80), which is the same sum as the einsums of 83-84.
 * bf_atom: the atom of each basis function (atom.mask, molecule.py:74-80). */
void orc_rhf_gradient(ORC_BFS_ARGS, int64_t N, int64_t natom, const double *Z, const double *axyz,
                      const int64_t *bf_atom, const double *P, const double *W, double *grad)
{
    ORC_BFS_INIT;
    memset(grad, 0, sizeof(double) * 3 * natom);
#define Pm(a, b) P[(a) * N + (b)]
    /* one-electron terms, per pair i >= j (weight 2 off the diagonal) */
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j) {
            double f = (i == j) ? 1.0 : 2.0;
            int64_t oi = poff[i], oj = poff[j];
            const int64_t *l1 = shell + 3 * i, *l2 = shell + 3 * j;
            const double *A = origin + 3 * i, *Bc = origin + 3 * j;
            for (int dx = 0; dx < 3; ++dx) {
                double sA = 0, sB = 0, hA = 0, hB = 0;
                double *hC = (double *)calloc((size_t)natom, sizeof(double));
                for (int64_t ia = 0; ia < np[i]; ++ia)
                    for (int64_t ib = 0; ib < np[j]; ++ib) {
                        double a = exps[oi + ia], b = exps[oj + ib];
                        double w = norm[oi + ia] * norm[oj + ib] * coefs[oi + ia] * coefs[oj + ib];
                        sA += w * prim_overlap_x(a, l1, A, b, l2, Bc, dx, 1);
                        sB += w * prim_overlap_x(a, l1, A, b, l2, Bc, dx, 2);
                        hA += w * prim_kinetic_x(a, l1, A, b, l2, Bc, dx, 1);
                        hB += w * prim_kinetic_x(a, l1, A, b, l2, Bc, dx, 2);
                        for (int64_t c = 0; c < natom; ++c) {
                            const double *C = axyz + 3 * c;
                            hA -= w * Z[c] * prim_nuclear_xb(a, l1, A, b, l2, Bc, C, dx, 1);
                            hB -= w * Z[c] * prim_nuclear_xb(a, l1, A, b, l2, Bc, C, dx, 2);
                            hC[c] -= w * Z[c] * prim_nuclear_xa(a, l1, A, b, l2, Bc, C, dx);
                        }
                    }
                grad[3 * bf_atom[i] + dx] += f * (2.0 * Pm(i, j) * hA - 2.0 * W[i * N + j] * sA);
                grad[3 * bf_atom[j] + dx] += f * (2.0 * Pm(i, j) * hB - 2.0 * W[i * N + j] * sB);
                for (int64_t c = 0; c < natom; ++c) grad[3 * c + dx] += f * 2.0 * Pm(i, j) * hC[c];
                free(hC);
            }
        }
    /* two-electron terms over canonical quartets */
    for (int64_t i = 0; i < N; ++i)
        for (int64_t j = 0; j <= i; ++j)
            for (int64_t k = 0; k < N; ++k)
                for (int64_t l = 0; l <= k; ++l) {
                    int64_t ij = i * (i + 1) / 2 + j, kl = k * (k + 1) / 2 + l;
                    if (ij < kl) continue;
                    /* the distinct positions of the quartet's orbit */
                    int64_t perm[8][4] = {{i, j, k, l}, {k, l, i, j}, {j, i, l, k}, {l, k, j, i},
                                          {j, i, k, l}, {l, k, i, j}, {i, j, l, k}, {k, l, j, i}};
                    double wgt = 0.0;
                    for (int s = 0; s < 8; ++s) {
                        int dup = 0;
                        for (int r = 0; r < s && !dup; ++r)
                            dup = perm[r][0] == perm[s][0] && perm[r][1] == perm[s][1] && perm[r][2] == perm[s][2]
                               && perm[r][3] == perm[s][3];
                        if (dup) continue;
                        int64_t a = perm[s][0], b = perm[s][1], c = perm[s][2], d = perm[s][3];
                        wgt += 2.0 * Pm(b, a) * Pm(d, c) - Pm(c, a) * Pm(b, d);
                    }
                    if (wgt == 0.0) continue;
                    const int64_t idx[4] = {i, j, k, l};
                    for (int ctr = 0; ctr < 4; ++ctr)
                        for (int dx = 0; dx < 3; ++dx) {
                            double d = 0.0;
                            int64_t o[4];
                            for (int s = 0; s < 4; ++s) o[s] = poff[idx[s]];
                            for (int64_t ja = 0; ja < np[i]; ++ja)
                                for (int64_t jb = 0; jb < np[j]; ++jb)
                                    for (int64_t jc = 0; jc < np[k]; ++jc)
                                        for (int64_t jd = 0; jd < np[l]; ++jd)
                                            d += norm[o[0] + ja] * norm[o[1] + jb] * norm[o[2] + jc] * norm[o[3] + jd]
                                               * coefs[o[0] + ja] * coefs[o[1] + jb] * coefs[o[2] + jc]
                                               * coefs[o[3] + jd]
                                               * prim_eri_x(exps[o[0] + ja], shell + 3 * i, origin + 3 * i,
                                                            exps[o[1] + jb], shell + 3 * j, origin + 3 * j,
                                                            exps[o[2] + jc], shell + 3 * k, origin + 3 * k,
                                                            exps[o[3] + jd], shell + 3 * l, origin + 3 * l, dx, ctr);
                            grad[3 * bf_atom[idx[ctr]] + dx] += wgt * d;
                        }
                }
#undef Pm
    (void)B;
    /* nuclear repulsion (forces.py:52-60) */
    for (int64_t a = 0; a < natom; ++a)
        for (int64_t c = 0; c < natom; ++c) {
            if (a == c) continue;
            double dx3[3], r2 = 0.0;
            for (int x = 0; x < 3; ++x) {
                dx3[x] = axyz[3 * a + x] - axyz[3 * c + x];
                r2 += dx3[x] * dx3[x];
            }
            double r = sqrt(r2);
            if (r < 1e-12) continue;
            for (int x = 0; x < 3; ++x) grad[3 * a + x] += -dx3[x] * Z[a] * Z[c] / (r2 * r);
        }
}
