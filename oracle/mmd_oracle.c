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
