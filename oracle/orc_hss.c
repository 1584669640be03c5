/*
 * CPU ORACLE -- test infrastructure only (see orc.h).
 *
 * Restatement of the eigenvector Cauchy->HSS construction and the HSS x dense
 * multiply of HSSPACK:
 *   tree        npart/n2tree/btree/child/hsslevel/GetTreeInfo   aux_hss.f90:238-827
 *   DHSSEVC     construction half                                CauchyHssEig.f90:2444-2583
 *   Cauchy2hssEig_omp (run sequentially, node order = level list) CauchyHssEig.f90:1945-2250
 *   ComprcauchyEig, rrluCauchyEigrow/col, CauchyPivtEig(m)(_CP),
 *   searchMax2Eig(m), CauchyMaxEig*, computdd/computfd,
 *   CauchylikeEig(Q)(Q2), invp                                   CauchyHssEig.f90:7-1942, 2253-2440
 *   fastHssmmL_omp                                               ConstructHssd.f90:1394-1630
 *
 * Arrays are addressed 1-based through the A1() macro to keep the statement
 * order of the Fortran.  One deliberate difference: CauchyPivtEig reads the
 * current pivot as F(PL(k)) with the node-local pole array F indexed by a global
 * index (CauchyHssEig.f90:665), which reads outside the node's poles; the global
 * pole array FF(PL(k)) is used here, as in the column variant (:821-825).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include "orc.h"

#define A1(a, i) (a)[(i) - 1]

/* ---------------- tree ---------------- */
static void btree(int n, int *tr)
{
    if (n > 3) {
        int m = (n - 1) / 2;
        btree(m, tr);
        for (int i = 1; i <= m; i++) A1(tr, m + i) = A1(tr, i) + m;
        A1(tr, m) = n;
        A1(tr, m + m) = n;
        A1(tr, n) = 0;
    } else {
        A1(tr, 1) = 3; A1(tr, 2) = 3; A1(tr, 3) = 0;
    }
}

static void n2tree(int n, int *tr)
{
    if (n == 1) { A1(tr, 1) = 0; return; }
    int n1 = (1 << (int)floorf(logf((float)n) / logf(2.0f))) - 1;
    btree(n1, tr);
    n2tree(n - n1 - 1, tr + n1);
    for (int i = n1 + 1; i <= n - 1; i++) A1(tr, i) += n1;
    A1(tr, n1) = n;
    A1(tr, n - 1) = n;
    A1(tr, n) = 0;
}

static void child(const int *tr, int *ch, int n)
{
    for (int i = 1; i <= n; i++) { ch[2 * (i - 1)] = 0; ch[2 * (i - 1) + 1] = 0; }
    for (int i = 1; i <= n - 1; i++) {
        int p = A1(tr, i);
        if (ch[2 * (p - 1)] == 0) ch[2 * (p - 1)] = i;
        else ch[2 * (p - 1) + 1] = i;
    }
}
#define CH1(ch, i) (ch)[2 * ((i) - 1)]
#define CH2(ch, i) (ch)[2 * ((i) - 1) + 1]

/* npart (aux_hss.f90:803) with ltr = 0 on entry */
static void npart(int n, int ni, int *tr, int *ltr, int *m, int *lm)
{
    int l = n / ni;
    for (int i = 1; i <= l; i++) A1(m, i) = ni;
    int mr = n % ni;
    if (mr >= ni / 2) { l++; A1(m, l) = mr; }
    else A1(m, l) += mr;
    *lm = l;
    *ltr = 2 * l - 1;
    n2tree(*ltr, tr);
}

/* ---------------- accurate differences ---------------- */
/* F(j)-D(k) = f_{PL(j)} - lambda_{PU(k)}  (computfd, CauchyHssEig.f90:519) */
static inline double computfd(const double *FF, int j, int k, const int *PL, const int *PU,
                              const double *DIFL, const double *DIFR, const double *S)
{
    int pd = A1(PL, j), pw = A1(PU, k);
    double r = (pd <= pw) ? (A1(FF, pd) - A1(FF, pw)) - A1(DIFL, pw) : (A1(FF, pd) - A1(FF, pw + 1)) - A1(DIFR, pw);
    return S ? r * (A1(FF, pd) + A1(S, pw)) : r;
}
/* D(j)-D(k) = lambda_{PU(j)} - lambda_{PU(k)}  (computdd, CauchyHssEig.f90:470) */
static inline double computdd(const double *FF, int j, int k, const int *PU,
                              const double *DIFL, const double *DIFR, const double *S)
{
    int pj = A1(PU, j), pk = A1(PU, k);
    double r = (pj < pk) ? ((A1(FF, pj + 1) - A1(FF, pk)) + A1(DIFR, pj)) - A1(DIFL, pk)
                         : ((A1(FF, pj) - A1(FF, pk + 1)) + A1(DIFL, pj)) - A1(DIFR, pk);
    return S ? r * (A1(S, pj) + A1(S, pk)) : r;
}
/* difference of two poles: f_a - f_b, or f_a^2 - f_b^2 as (f_a - f_b)(f_a + f_b) */
static inline double pp(double a, double b, const double *S) { return S ? (a - b) * (a + b) : a - b; }

static inline void dswp(double *a, int i, int j) { double t = A1(a, i); A1(a, i) = A1(a, j); A1(a, j) = t; }
static inline void iswp(int *a, int i, int j) { int t = A1(a, i); A1(a, i) = A1(a, j); A1(a, j) = t; }

/* max |.| and first location (MAXVAL/MAXLOC); n == 0 gives -huge, 0 */
#define MAXLOC_UPDATE(val, idx)                                   \
    do {                                                          \
        double _v = fabs(val);                                    \
        if (first || _v > best) { best = _v; loc = (idx); first = 0; } \
    } while (0)

/* CauchyMaxEigmc: column entries U(i)/(f_{PL(i)} - lambda_{PU(1)}) */
static void maxEigmc(const double *FF, const int *PL, const int *PU, const double *DIFL,
                     const double *DIFR, const double *S, const double *U, int n, double *junk, int *jj)
{
    double best = -DBL_MAX; int loc = 0, first = 1;
    int dk = A1(PU, 1);
    for (int i = 1; i <= n; i++) {
        int pli = A1(PL, i);
        double den = (pli <= dk) ? (A1(FF, pli) - A1(FF, dk)) - A1(DIFL, dk) : (A1(FF, pli) - A1(FF, dk + 1)) - A1(DIFR, dk);
        if (S) den = den * (A1(FF, pli) + A1(S, dk));
        double v = A1(U, i) / den;
        MAXLOC_UPDATE(v, i);
    }
    *junk = best; *jj = loc;
}
/* CauchyMaxEigm: row entries V(i)/(f_{Fk} - lambda_{PU(i)}) */
static void maxEigm(const double *FF, int fk, const int *PU, const double *DIFL,
                    const double *DIFR, const double *S, const double *V, int n, double *junk, int *jj)
{
    double best = -DBL_MAX; int loc = 0, first = 1;
    double f = A1(FF, fk);
    for (int i = 1; i <= n; i++) {
        int pui = A1(PU, i);
        double den = (fk <= pui) ? (f - A1(FF, pui)) - A1(DIFL, pui) : (f - A1(FF, pui + 1)) - A1(DIFR, pui);
        if (S) den = den * (f + A1(S, pui));
        double v = A1(V, i) / den;
        MAXLOC_UPDATE(v, i);
    }
    *junk = best; *jj = loc;
}
/* CauchyMaxEig: U(i)/(D(i)-F) */
static void maxEig(const double *D, double f, const double *S, const double *U, int n, double *junk, int *jj)
{
    double best = -DBL_MAX; int loc = 0, first = 1;
    for (int i = 1; i <= n; i++) {
        double v = A1(U, i) / pp(A1(D, i), f, S);
        MAXLOC_UPDATE(v, i);
    }
    *junk = best; *jj = loc;
}
/* CauchyMaxEigmm: U(i)/(D(i+N)-D(k)) via computdd */
static void maxEigmm(const double *FF, const int *PL, int k, const double *DIFL,
                     const double *DIFR, const double *S, const double *U, int m, int n, double *junk, int *jj)
{
    double best = -DBL_MAX; int loc = 0, first = 1;
    for (int i = 1; i <= m; i++) {
        double tt = computdd(FF, i + n, k, PL, DIFL, DIFR, S);
        double v = A1(U, i) / tt;
        MAXLOC_UPDATE(v, i);
    }
    *junk = best; *jj = loc;
}

/* CauchyPivtEig (CauchyHssEig.f90:571): rook-style pivot for the row RRLU */
static void pivtEig(double *F, double *D, double *U, double *V, int *PL, int *PU, int k,
                    double *A, int M, int N, int *PQ, const double *FF, const double *DIFL,
                    const double *DIFR, const double *S)
{
    double junkL, junkU, piv;
    int jjL, jjU;
    maxEigmc(FF, PL + k - 1, PU + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &junkL, &jjL);
    maxEigm(FF, A1(PL, k), PU + k - 1, DIFL, DIFR, S, V + k - 1, N - k + 1, &junkU, &jjU);
    junkL = junkL * fabs(A1(V, k));
    junkU = junkU * fabs(A1(U, k));
    {
        int a = A1(PL, k), b = A1(PU, k);
        double den = (a <= b) ? (A1(FF, a) - A1(FF, b)) - A1(DIFL, b) : (A1(FF, a) - A1(FF, b + 1)) - A1(DIFR, b);
        if (S) den = den * (A1(FF, a) + A1(S, b));
        piv = fabs(A1(U, k) * A1(V, k) / den);
    }
    if (junkL <= piv && junkU <= piv) return;
    int flg = (junkL > junkU) ? 1 : 0;
    for (;;) {
        if (flg == 1) {
            int jj = jjL;
            dswp(F, k, jj + k - 1); dswp(U, k, jj + k - 1); dswp(A, k, jj + k - 1);
            iswp(PL, k, jj + k - 1); iswp(PQ, k, jj + k - 1);
            maxEigm(FF, A1(PL, k), PU + k - 1, DIFL, DIFR, S, V + k - 1, N - k + 1, &junkU, &jjU);
            if (jjU == 1) return;
            flg = 0;
        }
        int jj = jjU;
        dswp(D, k, jj + k - 1); dswp(V, k, jj + k - 1); iswp(PU, k, jj + k - 1);
        maxEigmc(FF, PL + k - 1, PU + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &junkL, &jjL);
        if (jjL == 1) return;
        flg = 1;
    }
}

/* CauchyPivtEig_CP (CauchyHssEig.f90:1049): complete pivoting on the Schur complement */
static void pivtEigCP(double *F, double *D, double *U, double *V, int *PL, int *PU, int k,
                      double *A, int M, int N, int *PQ, const double *FF, const double *DIFL,
                      const double *DIFR, const double *S)
{
    double ju, junk;
    int jju, ii, jj = 1;
    maxEigmc(FF, PL + k - 1, PU + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &ju, &jju);
    junk = fabs(ju * A1(V, k));
    ii = jju;
    for (int j = 2; j <= N - k + 1; j++) {
        maxEigmc(FF, PL + k - 1, PU + k + j - 2, DIFL, DIFR, S, U + k - 1, M - k + 1, &ju, &jju);
        double jt = fabs(ju * A1(V, k + j - 1));
        if (jt > junk) { junk = jt; ii = jju; jj = j; }
    }
    dswp(F, k, ii + k - 1); dswp(U, k, ii + k - 1); dswp(A, k, ii + k - 1);
    iswp(PL, k, ii + k - 1); iswp(PQ, k, ii + k - 1);
    dswp(D, k, jj + k - 1); dswp(V, k, jj + k - 1); iswp(PU, k, jj + k - 1);
}

/* CauchyPivtEigm (CauchyHssEig.f90:737): rook-style pivot for the column RRLU */
static void pivtEigm(double *D, double *F, double *U, double *V, int *PL, int *PU, int k,
                     double *A, int M, int N, int *PQ, const double *FF, const double *DIFL,
                     const double *DIFR, const double *S)
{
    double junkL, junkU, piv;
    int jjL, jjU;
    maxEigm(FF, A1(PU, k), PL + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &junkL, &jjL);
    maxEigmc(FF, PU + k - 1, PL + k - 1, DIFL, DIFR, S, V + k - 1, N - k + 1, &junkU, &jjU);
    junkL = junkL * fabs(A1(V, k));
    junkU = junkU * fabs(A1(U, k));
    {
        int a = A1(PL, k), b = A1(PU, k);  /* root a, pole b */
        double den = (a < b) ? (A1(FF, a + 1) - A1(FF, b)) + A1(DIFR, a) : (A1(FF, a) - A1(FF, b)) + A1(DIFL, a);
        if (S) den = den * (A1(S, a) + A1(FF, b));
        piv = fabs(A1(U, k) * A1(V, k) / den);
    }
    if (junkL <= piv && junkU <= piv) return;
    int flg = (junkL > junkU) ? 1 : 0;
    for (;;) {
        if (flg == 1) {
            int jj = jjL;
            dswp(D, k, jj + k - 1); dswp(U, k, jj + k - 1); dswp(A, k, jj + k - 1);
            iswp(PL, k, jj + k - 1); iswp(PQ, k, jj + k - 1);
            maxEigmc(FF, PU + k - 1, PL + k - 1, DIFL, DIFR, S, V + k - 1, N - k + 1, &junkU, &jjU);
            if (jjU == 1) return;
            flg = 0;
        }
        int jj = jjU;
        dswp(F, k, jj + k - 1); dswp(V, k, jj + k - 1); iswp(PU, k, jj + k - 1);
        maxEigm(FF, A1(PU, k), PL + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &junkL, &jjL);
        if (jjL == 1) return;
        flg = 1;
    }
}

/* CauchyPivtEigm_CP (CauchyHssEig.f90:861) */
static void pivtEigmCP(double *D, double *F, double *U, double *V, int *PL, int *PU, int k,
                       double *A, int M, int N, int *PQ, const double *FF, const double *DIFL,
                       const double *DIFR, const double *S)
{
    double ju, junk;
    int jju, ii, jj = 1;
    maxEigm(FF, A1(PU, k), PL + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &ju, &jju);
    junk = fabs(ju * A1(V, k));
    ii = jju;
    for (int j = 2; j <= N - k + 1; j++) {
        maxEigm(FF, A1(PU, k + j - 1), PL + k - 1, DIFL, DIFR, S, U + k - 1, M - k + 1, &ju, &jju);
        double jt = fabs(ju * A1(V, k + j - 1));
        if (jt > junk) { junk = jt; ii = jju; jj = j; }
    }
    dswp(D, k, ii + k - 1); dswp(U, k, ii + k - 1); dswp(A, k, ii + k - 1);
    iswp(PL, k, ii + k - 1); iswp(PQ, k, ii + k - 1);
    dswp(F, k, jj + k - 1); dswp(V, k, jj + k - 1); iswp(PU, k, jj + k - 1);
}

/* rrluCauchyEigrow (CauchyHssEig.f90:7-212) */
static void rrluRow(double *F, double *D, double *A, double *V, const double *DIFL,
                    const double *DIFR, const double *S, double tol, double *U, double *W, const double *FF,
                    int *PL, int *PU, int M, int N, int *Rk, int *PQ, long long *nflops,
                    long long *nswap)
{
    int mn = M < N ? M : N, k, j, ii, jj;
    double amax = 0.0, rho = 1.5, ukk, temp, temp2, nmax;
    const int prd = 10;
    for (j = 1; j <= M; j++) { A1(U, j) = A1(A, j); A1(PQ, j) = j; }
    *Rk = mn;
    for (k = 1; k <= mn; k++) {
        pivtEig(F, D, U, V, PL, PU, k, A, M, N, PQ, FF, DIFL, DIFR, S);
        temp = computfd(FF, k, k, PL, PU, DIFL, DIFR, S);
        ukk = A1(U, k) * A1(V, k) / temp;
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            pivtEigCP(F, D, U, V, PL, PU, k, A, M, N, PQ, FF, DIFL, DIFR, S);
            temp = computfd(FF, k, k, PL, PU, DIFL, DIFR, S);
            ukk = A1(U, k) * A1(V, k) / temp;
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { *Rk = k - 1; break; }
        }
        for (j = k + 1; j <= M; j++) {
            temp = computfd(FF, j, k, PL, PU, DIFL, DIFR, S);
            temp = pp(A1(F, j), A1(F, k), S) / temp;
            A1(U, j) = A1(U, j) * temp;
        }
        for (j = k + 1; j <= N; j++) {
            temp = computdd(FF, j, k, PU, DIFL, DIFR, S);
            temp2 = -1.0 * computfd(FF, k, j, PL, PU, DIFL, DIFR, S);
            temp = temp / temp2;
            A1(V, j) = A1(V, j) * temp;
        }
        for (j = 1; j <= k - 1; j++) {
            temp = -1.0 * computfd(FF, j, k, PL, PU, DIFL, DIFR, S);
            A1(W, j) = A1(W, j) * (temp / pp(A1(F, k), A1(F, j), S));
        }
        temp = computfd(FF, k, k, PL, PU, DIFL, DIFR, S);
        A1(W, k) = temp / A1(A, k);
        for (j = 1; j <= k - 1; j++) {
            temp = -1.0 * computfd(FF, k, j, PL, PU, DIFL, DIFR, S);
            A1(W, k) = A1(W, k) * (temp / pp(A1(F, j), A1(F, k), S));
        }
        *nflops += 5 * (M - k) + 7 * (N - k) + 2 * 6 * (k - 1);
        int flgl = k % prd;
        while (flgl < 1) {
            flgl = 1;
            /* searchMax2Eig(U(k+1),W,F(k+1),F,M-k,k,...) */
            double junk; int jjL;
            maxEig(F + k, A1(F, 1), S, U + k, M - k, &junk, &jjL);
            junk = junk * fabs(A1(W, 1));
            nmax = junk; ii = 1; jj = jjL;
            for (j = 2; j <= k; j++) {
                maxEig(F + k, A1(F, j), S, U + k, M - k, &junk, &jjL);
                junk = junk * fabs(A1(W, j));
                if (junk > nmax) { nmax = junk; ii = j; jj = jjL; }
            }
            if (nmax > rho) {
                (*nswap)++;
                flgl = 0;
                jj = jj + k;
                for (j = k + 1; j <= N; j++) {
                    temp = -1.0 * computfd(FF, ii, j, PL, PU, DIFL, DIFR, S);
                    temp2 = -1.0 * computfd(FF, jj, j, PL, PU, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(V, j) = A1(V, j) * temp;
                }
                for (j = k + 1; j <= jj - 1; j++)
                    A1(U, j) = A1(U, j) * (pp(A1(F, j), A1(F, jj), S) / pp(A1(F, j), A1(F, ii), S));
                for (j = jj + 1; j <= M; j++)
                    A1(U, j) = A1(U, j) * (pp(A1(F, j), A1(F, jj), S) / pp(A1(F, j), A1(F, ii), S));
                for (j = 1; j <= ii - 1; j++)
                    A1(W, j) = A1(W, j) * (pp(A1(F, j), A1(F, ii), S) / pp(A1(F, j), A1(F, jj), S));
                for (j = ii + 1; j <= k; j++)
                    A1(W, j) = A1(W, j) * (pp(A1(F, j), A1(F, ii), S) / pp(A1(F, j), A1(F, jj), S));
                temp = computfd(FF, ii, ii, PL, PU, DIFL, DIFR, S);
                A1(U, jj) = A1(A, ii) * (pp(A1(F, ii), A1(F, jj), S) / temp);
                temp = computfd(FF, jj, ii, PL, PU, DIFL, DIFR, S);
                A1(W, ii) = temp / A1(A, jj);
                for (j = 1; j <= k; j++) {
                    if (j == ii) continue;
                    temp = computfd(FF, ii, j, PL, PU, DIFL, DIFR, S);
                    A1(U, jj) = A1(U, jj) * (pp(A1(F, ii), A1(F, j), S) / temp);
                    temp = -1.0 * computfd(FF, jj, j, PL, PU, DIFL, DIFR, S);
                    A1(W, ii) = A1(W, ii) * (temp / pp(A1(F, j), A1(F, jj), S));
                }
                iswp(PL, ii, jj); iswp(PQ, ii, jj);
                dswp(F, ii, jj); dswp(A, ii, jj);
                *nflops += 10 * (N - k) + 8 * M + 18 * k;
            }
        }
    }
}

/* Test hook: the row (rowcol 0) or column RRLU above on given generators, S = NULL;
 * the tests compare it step for step with the reference's own routines. */
void orc_rrlu(int rowcol, double *F, double *D, double *A, double *V, const double *DIFL, const double *DIFR,
              double tol, double *U, double *W, const double *FF, int *PL, int *PU, int M, int N, int *Rk,
              int *PQ, long long *nflops, long long *nswap);

/* rrluCauchyEigcol (CauchyHssEig.f90:215-467) */
static void rrluCol(double *D, double *F, double *A, double *V, const double *DIFL,
                    const double *DIFR, const double *S, double tol, double *U, double *W, const double *FF,
                    int *PL, int *PU, int M, int N, int *Rk, int *PQ, long long *nflops,
                    long long *nswap)
{
    int mn = M < N ? M : N, k, j, ii, jj;
    double amax = 0.0, rho = 1.5, ukk, temp, temp2, nmax;
    const int prd = 10;
    for (j = 1; j <= M; j++) { A1(U, j) = A1(A, j); A1(PQ, j) = j; }
    *Rk = mn;
    for (k = 1; k <= mn; k++) {
        pivtEigm(D, F, U, V, PL, PU, k, A, M, N, PQ, FF, DIFL, DIFR, S);
        temp = -1.0 * computfd(FF, k, k, PU, PL, DIFL, DIFR, S);
        ukk = A1(U, k) * A1(V, k) / temp;
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            pivtEigmCP(D, F, U, V, PL, PU, k, A, M, N, PQ, FF, DIFL, DIFR, S);
            temp = -1.0 * computfd(FF, k, k, PU, PL, DIFL, DIFR, S);
            ukk = A1(U, k) * A1(V, k) / temp;
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { *Rk = k - 1; break; }
        }
        for (j = k + 1; j <= M; j++) {
            temp = computdd(FF, j, k, PL, DIFL, DIFR, S);
            temp2 = -1.0 * computfd(FF, k, j, PU, PL, DIFL, DIFR, S);
            temp = temp / temp2;
            A1(U, j) = A1(U, j) * temp;
        }
        for (j = k + 1; j <= N; j++) {
            temp = computfd(FF, j, k, PU, PL, DIFL, DIFR, S);
            A1(V, j) = A1(V, j) * (pp(A1(F, j), A1(F, k), S) / temp);
        }
        for (j = 1; j <= k - 1; j++) {
            temp = computfd(FF, k, j, PU, PL, DIFL, DIFR, S);
            temp2 = computdd(FF, k, j, PL, DIFL, DIFR, S);
            temp = temp / temp2;
            A1(W, j) = A1(W, j) * temp;
        }
        temp = -1.0 * computfd(FF, k, k, PU, PL, DIFL, DIFR, S);
        A1(W, k) = temp / A1(A, k);
        for (j = 1; j <= k - 1; j++) {
            temp = computfd(FF, j, k, PU, PL, DIFL, DIFR, S);
            temp2 = computdd(FF, j, k, PL, DIFL, DIFR, S);
            temp = temp / temp2;
            A1(W, k) = A1(W, k) * temp;
        }
        *nflops += 6 * (M - k) + 5 * (N - k) + 2 * 7 * (k - 1);
        int flgl = k % prd;
        while (flgl < 1) {
            flgl = 1;
            double junk; int jjL;
            maxEigmm(FF, PL, 1, DIFL, DIFR, S, U + k, M - k, k, &junk, &jjL);
            junk = junk * fabs(A1(W, 1));
            nmax = junk; ii = 1; jj = jjL;
            for (j = 2; j <= k; j++) {
                maxEigmm(FF, PL, j, DIFL, DIFR, S, U + k, M - k, k, &junk, &jjL);
                junk = junk * fabs(A1(W, j));
                if (junk > nmax) { nmax = junk; ii = j; jj = jjL; }
            }
            if (nmax > rho) {
                (*nswap)++;
                flgl = 0;
                jj = jj + k;
                for (j = k + 1; j <= N; j++) {
                    temp = computfd(FF, j, ii, PU, PL, DIFL, DIFR, S);
                    temp2 = computfd(FF, j, jj, PU, PL, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(V, j) = A1(V, j) * temp;
                }
                for (j = k + 1; j <= M; j++) {
                    if (j == jj) continue;
                    temp = computdd(FF, j, jj, PL, DIFL, DIFR, S);
                    temp2 = computdd(FF, j, ii, PL, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(U, j) = A1(U, j) * temp;
                }
                for (j = 1; j <= k; j++) {
                    if (j == ii) continue;
                    temp = computdd(FF, j, ii, PL, DIFL, DIFR, S);
                    temp2 = computdd(FF, j, jj, PL, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(W, j) = A1(W, j) * temp;
                }
                temp = computdd(FF, ii, jj, PL, DIFL, DIFR, S);
                temp2 = -1.0 * computfd(FF, ii, ii, PU, PL, DIFL, DIFR, S);
                temp = temp / temp2;
                A1(U, jj) = A1(A, ii) * temp;
                temp = -1.0 * computfd(FF, ii, jj, PU, PL, DIFL, DIFR, S);
                A1(W, ii) = temp / A1(A, jj);
                for (j = 1; j <= k; j++) {
                    if (j == ii) continue;
                    temp = computdd(FF, ii, j, PL, DIFL, DIFR, S);
                    temp2 = -1.0 * computfd(FF, j, ii, PU, PL, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(U, jj) = A1(U, jj) * temp;
                    temp = computfd(FF, j, jj, PU, PL, DIFL, DIFR, S);
                    temp2 = computdd(FF, j, jj, PL, DIFL, DIFR, S);
                    temp = temp / temp2;
                    A1(W, ii) = A1(W, ii) * temp;
                }
                iswp(PL, ii, jj); iswp(PQ, ii, jj);
                dswp(D, ii, jj); dswp(A, ii, jj);
                *nflops += 10 * (N - k) + 12 * M + 22 * k;
            }
        }
    }
}

static void invp(int *P, int n)
{
    int *ip = (int *)calloc(n, sizeof(int));
    for (int i = 1; i <= n; i++) A1(ip, A1(P, i)) = i;
    memcpy(P, ip, sizeof(int) * n);
    free(ip);
}

/* CauchylikeEig (CauchyHssEig.f90:1688): A(i,j) = U(i)V(j)/(f_{PL(i)} - lambda_{PU(j)}) */
static void cauchylikeEig(double *a, int lda, const double *F, const double *U, const double *V,
                          const double *DIFL, const double *DIFR, const double *S, const int *PL, const int *PU,
                          int M, int N)
{
    for (int j = 1; j <= N; j++) {
        int pdjj = A1(PU, j);
        double fj = A1(F, pdjj), vj = A1(V, j);
        for (int i = 1; i <= M; i++) {
            int pfii = A1(PL, i);
            double den;
            if (pfii <= pdjj) den = (A1(F, pfii) - fj) - A1(DIFL, pdjj);
            else den = (A1(F, pfii) - A1(F, pdjj + 1)) - A1(DIFR, pdjj);
            if (S) den = den * (A1(F, pfii) + A1(S, pdjj));
            a[(i - 1) + (long)(j - 1) * lda] = A1(U, i) * vj / den;
        }
    }
}

/* ---------------- HSS storage ---------------- */
static long push(orc_hss *H, long *cap, long n)
{
    long off = H->nh;
    if (H->nh + n > *cap) {
        while (H->nh + n > *cap) *cap = *cap * 2 + 1024;
        H->h = (double *)realloc(H->h, sizeof(double) * (*cap));
    }
    H->nh += n;
    return off;
}

/* ComprcauchyEig (CauchyHssEig.f90:2253-2440) */
static void comprEig(orc_hss *H, long *cap, char rc, double *F, double *D, double *U, double *V,
                     const double *DIFL, const double *DIFR, const double *S, const double *FF, int *PL, int *PU,
                     double tol, int M, int N, int nodi, int *Rk)
{
    int mn = M < N ? M : N, i, j;
    if (rc == 'r') {
        double *Z = (double *)calloc(M + 1, sizeof(double)), *W = (double *)calloc(mn + 1, sizeof(double));
        int *PQ = (int *)malloc(sizeof(int) * (M + 1));
        rrluRow(F, D, U, V, DIFL, DIFR, S, tol, Z, W, FF, PL, PU, M, N, Rk, PQ, &H->nflops, &H->nswap);
        invp(PQ, M);
        int r = *Rk;
        if (r < mn) {
            double *Q = (double *)calloc((long)M * (r > 0 ? r : 1), sizeof(double));
            /* CauchylikeEigQ(Q, F(Rk+1), F, Z(Rk+1), W, M, Rk) */
            for (j = 1; j <= r; j++) {
                Q[(j - 1) + (long)(j - 1) * M] = 1.0;
                for (i = r + 1; i <= M; i++)
                    Q[(i - 1) + (long)(j - 1) * M] = A1(Z, i) * A1(W, j) / pp(A1(F, i), A1(F, j), S);
            }
            long off = push(H, cap, (long)M * r);
            for (j = 0; j < r; j++)
                for (i = 1; i <= M; i++) H->h[off + (i - 1) + (long)j * M] = Q[(A1(PQ, i) - 1) + (long)j * M];
            H->usz[2 * (nodi - 1)] = M; H->usz[2 * (nodi - 1) + 1] = r; H->pu[nodi - 1] = off;
            free(Q);
        } else {
            *Rk = M;
            long off = push(H, cap, (long)M * M);
            for (j = 0; j < M; j++)
                for (i = 1; i <= M; i++) H->h[off + (i - 1) + (long)j * M] = (A1(PQ, i) - 1 == j) ? 1.0 : 0.0;
            H->usz[2 * (nodi - 1)] = M; H->usz[2 * (nodi - 1) + 1] = M; H->pu[nodi - 1] = off;
        }
        free(Z); free(W); free(PQ);
    } else {
        /* column compression: the rows of the RRLU are this node's N columns */
        double *Z = (double *)calloc(N + 1, sizeof(double)), *W = (double *)calloc(mn + 1, sizeof(double));
        int *PQ = (int *)malloc(sizeof(int) * (N + 1));
        rrluCol(D, F, V, U, DIFL, DIFR, S, tol, Z, W, FF, PU, PL, N, M, Rk, PQ, &H->nflops, &H->nswap);
        invp(PQ, N);
        int r = *Rk;
        if (r < mn) {
            double *Q = (double *)calloc((long)(r > 0 ? r : 1) * N, sizeof(double));
            /* CauchylikeEigQ2(Q, D, D(Rk+1), W, Z(Rk+1), Rk, N, FF, DIFL, DIFR, PU) */
            for (j = 1; j <= r; j++) Q[(j - 1) + (long)(j - 1) * r] = 1.0;
            for (j = r + 1; j <= N; j++) {
                int pdjj = A1(PU, j);
                for (i = 1; i <= r; i++) {
                    int pdii = A1(PU, i);
                    double den;
                    if (pdii < pdjj) den = (A1(FF, pdii + 1) - A1(FF, pdjj)) + (A1(DIFR, pdii) - A1(DIFL, pdjj));
                    else den = (A1(FF, pdii) - A1(FF, pdjj + 1)) + (A1(DIFL, pdii) - A1(DIFR, pdjj));
                    if (S) den = den * (A1(S, pdii) + A1(S, pdjj));
                    Q[(i - 1) + (long)(j - 1) * r] = -1.0 * A1(W, i) * A1(Z, j) / den;
                }
            }
            long off = push(H, cap, (long)r * N);
            for (j = 1; j <= N; j++)
                for (i = 0; i < r; i++) H->h[off + i + (long)(j - 1) * r] = Q[i + (long)(A1(PQ, j) - 1) * r];
            H->vsz[2 * (nodi - 1)] = r; H->vsz[2 * (nodi - 1) + 1] = N; H->pv[nodi - 1] = off;
            free(Q);
        } else {
            *Rk = N;
            long off = push(H, cap, (long)N * N);
            for (j = 1; j <= N; j++)
                for (i = 0; i < N; i++) H->h[off + i + (long)(j - 1) * N] = (A1(PQ, j) - 1 == i) ? 1.0 : 0.0;
            H->vsz[2 * (nodi - 1)] = N; H->vsz[2 * (nodi - 1) + 1] = N; H->pv[nodi - 1] = off;
        }
        free(Z); free(W); free(PQ);
    }
}

void orc_hss_free(orc_hss *H)
{
    free(H->tr); free(H->m); free(H->ch); free(H->lentr); free(H->ttr);
    free(H->dsz); free(H->usz); free(H->vsz); free(H->bsz);
    free(H->pu); free(H->pv); free(H->pb); free(H->pd);
    free(H->h); free(H->dd);
    memset(H, 0, sizeof(*H));
}

#define TTR(H, LL, i) (H)->ttr[((LL) - 1) * (H)->ltr + (i) - 1]

/* The post-ordered HSS tree of K indices in leaves of about ni (npart, child,
 * hsslevel, GetTreeInfo: aux_hss.f90:238-827) and the generator directory and
 * storage of H.  Returns the number of levels (< 1: a single leaf). */
static int hss_tree_setup(orc_hss *H, int K, int ni)
{
    memset(H, 0, sizeof(*H));
    H->k = K;
    int lm0 = K / ni + 1;
    H->m = (int *)calloc(lm0 + 2, sizeof(int));
    H->tr = (int *)calloc(2 * lm0 + 8, sizeof(int));
    int ltr, lm;
    npart(K, ni, H->tr, &ltr, H->m, &lm);
    H->ltr = ltr; H->lm = lm;
    H->ch = (int *)calloc(2 * ltr, sizeof(int));
    child(H->tr, H->ch, ltr);
    /* node depths (hsslevel) and the number of levels */
    int *lv = (int *)calloc(ltr + 1, sizeof(int));
    int lvl = 0;
    A1(lv, ltr) = 0;
    for (int i = ltr - 1; i >= 1; i--) { A1(lv, i) = A1(lv, A1(H->tr, i)) + 1; if (A1(lv, i) > lvl) lvl = A1(lv, i); }
    H->lvl = lvl;
    if (lvl < 1) { free(lv); return -1; }
    H->lentr = (int *)calloc(lvl + 1, sizeof(int));
    H->ttr = (int *)calloc((long)lvl * ltr, sizeof(int));
    for (int LL = lvl; LL >= 1; LL--) {
        int it = 0;
        for (int j = 1; j <= ltr - 1; j++) if (A1(lv, j) == LL) { it++; TTR(H, LL, it) = j; }
        A1(H->lentr, LL) = it;
    }
    free(lv);
    H->dsz = (int *)calloc(ltr, sizeof(int));
    H->usz = (int *)calloc(2 * ltr, sizeof(int));
    H->vsz = (int *)calloc(2 * ltr, sizeof(int));
    H->bsz = (int *)calloc(2 * ltr, sizeof(int));
    H->pu = (long *)calloc(ltr, sizeof(long));
    H->pv = (long *)calloc(ltr, sizeof(long));
    H->pb = (long *)calloc(ltr, sizeof(long));
    H->pd = (long *)calloc(ltr, sizeof(long));
    long tnk = 0;
    for (int i = 1; i <= lm; i++) tnk += (long)A1(H->m, i) * A1(H->m, i);
    H->dd = (double *)malloc(sizeof(double) * tnk);
    H->ndd = tnk;
    long cap = 10L * ni * K;
    H->h = (double *)malloc(sizeof(double) * cap);
    H->nh = 0;
    return lvl;
}

/* DHSSEVC construction + Cauchy2hssEig_omp, sequential.  S == NULL: the
 * eigenvector matrix u_i v_j / (f_i - d_j); S == d0: the SVD's vector matrices
 * u_i v_j / (f_i^2 - d_j^2), every displacement formed as the accurate difference
 * times the sum (the same RRLU in the variables f^2 and d^2). */
static int hss_build(orc_hss *H, int K, int ni, const double *f, const double *d0, const double *u0,
                     const double *v0, const double *DIFL, const double *DIFR, const double *S,
                     double tol)
{
    if (hss_tree_setup(H, K, ni) < 1) return -1;
    const int ltr = H->ltr, lvl = H->lvl;
    long cap = 10L * ni * K;
    int *ch = H->ch;
    double *D = (double *)malloc(sizeof(double) * K), *U = (double *)malloc(sizeof(double) * K),
           *V = (double *)malloc(sizeof(double) * K);
    int *PL = (int *)malloc(sizeof(int) * K), *PU = (int *)malloc(sizeof(int) * K);
    int *lr = (int *)malloc(sizeof(int) * 2 * ltr), *lc = (int *)malloc(sizeof(int) * 2 * ltr);
    int *indx = (int *)malloc(sizeof(int) * K);
    double *Di = (double *)malloc(sizeof(double) * K), *Fi = (double *)malloc(sizeof(double) * K),
           *Ui = (double *)malloc(sizeof(double) * K), *Vi = (double *)malloc(sizeof(double) * K);
    int *PLi = (int *)malloc(sizeof(int) * K), *PUi = (int *)malloc(sizeof(int) * K);
    memcpy(D, d0, sizeof(double) * K);
    memcpy(U, u0, sizeof(double) * K);
    memcpy(V, v0, sizeof(double) * K);
    for (int j = 1; j <= K; j++) { A1(PL, j) = j; A1(PU, j) = j; }
#define LR(i, c) lr[2 * ((i) - 1) + (c) - 1]
#define LC(i, c) lc[2 * ((i) - 1) + (c) - 1]
    int lt = 1, itl = 1;
    long pdt = 0;
    for (int i = 1; i <= ltr; i++) {
        if (CH1(ch, i) == 0 && CH2(ch, i) == 0) {
            LR(i, 1) = lt; LR(i, 2) = lt + A1(H->m, itl) - 1;
            lt = LR(i, 2) + 1;
            H->pd[i - 1] = pdt;
            pdt += (long)A1(H->m, itl) * A1(H->m, itl);
            itl++;
        } else {
            LR(i, 1) = LR(CH1(ch, i), 1); LR(i, 2) = LR(CH2(ch, i), 2);
        }
    }
    memcpy(lc, lr, sizeof(int) * 2 * ltr);
    for (int i = 1; i <= ltr; i++) {
        if (CH1(ch, i) == 0 && CH2(ch, i) == 0) {
            int mi = LR(i, 2) - LR(i, 1) + 1;
            cauchylikeEig(H->dd + H->pd[i - 1], mi, f, U + LR(i, 1) - 1, V + LR(i, 1) - 1, DIFL, DIFR, S,
                          PL + LR(i, 1) - 1, PU + LR(i, 1) - 1, mi, mi);
            H->dsz[i - 1] = mi;
        }
    }
    int lda = K, lca = K;
    for (int LL = lvl; LL >= 1; LL--) {
        int n1 = A1(H->lentr, LL);
        /* row compression */
        for (int i = 1; i <= n1; i++) {
            int nodi = TTR(H, LL, i), Rk;
            if (CH1(ch, nodi) != 0) {
                LR(nodi, 1) = LR(CH1(ch, nodi), 1); LR(nodi, 2) = LR(CH2(ch, nodi), 2);
                LC(nodi, 1) = LC(CH1(ch, nodi), 1); LC(nodi, 2) = LC(CH2(ch, nodi), 2);
            }
            int mi = LR(nodi, 2) - LR(nodi, 1) + 1;
            int nic = LC(nodi, 2) - LC(nodi, 1) + 1;
            int nt = lca - nic, q = 0;
            for (int j = 1; j <= LC(nodi, 1) - 1; j++) indx[q++] = j;
            for (int j = LC(nodi, 2) + 1; j <= lca; j++) indx[q++] = j;
            for (int j = 0; j < nt; j++) {
                PUi[j] = A1(PU, indx[j]); Vi[j] = A1(V, indx[j]); Di[j] = A1(D, indx[j]);
            }
            for (int j = 0; j < mi; j++) Fi[j] = A1(f, A1(PL, LR(nodi, 1) + j));
            comprEig(H, &cap, 'r', Fi, Di, U + LR(nodi, 1) - 1, Vi, DIFL, DIFR, S, f,
                     PL + LR(nodi, 1) - 1, PUi, tol, mi, nt, nodi, &Rk);
        }
        /* compact surviving row generators */
        lt = 1;
        for (int i = 1; i <= n1; i++) {
            int nodi = TTR(H, LL, i);
            int Rk = H->usz[2 * (nodi - 1) + 1];
            for (int j = 0; j < Rk; j++) { Ui[lt - 1 + j] = A1(U, LR(nodi, 1) + j); PLi[lt - 1 + j] = A1(PL, LR(nodi, 1) + j); }
            LR(nodi, 1) = lt; LR(nodi, 2) = lt + Rk - 1; lt += Rk;
        }
        for (int LLt = LL - 1; LLt >= 1; LLt--) {
            for (int i = 1; i <= A1(H->lentr, LLt); i++) {
                int nodi = TTR(H, LLt, i);
                if (CH1(ch, nodi) == 0 && CH2(ch, nodi) == 0) {
                    int mi = LR(nodi, 2) - LR(nodi, 1) + 1;
                    for (int j = 0; j < mi; j++) { Ui[lt - 1 + j] = A1(U, LR(nodi, 1) + j); PLi[lt - 1 + j] = A1(PL, LR(nodi, 1) + j); }
                    LR(nodi, 1) = lt; LR(nodi, 2) = lt + mi - 1; lt += mi;
                }
            }
        }
        memcpy(U, Ui, sizeof(double) * (lt - 1));
        memcpy(PL, PLi, sizeof(int) * (lt - 1));
        lda = lt - 1;
        /* column compression */
        for (int i = 1; i <= n1; i++) {
            int nodi = TTR(H, LL, i), Rk;
            int mi = LC(nodi, 2) - LC(nodi, 1) + 1;
            int nir = LR(nodi, 2) - LR(nodi, 1) + 1;
            int nt = lda - nir, q = 0;
            for (int j = 1; j <= LR(nodi, 1) - 1; j++) indx[q++] = j;
            for (int j = LR(nodi, 2) + 1; j <= lda; j++) indx[q++] = j;
            for (int j = 0; j < nt; j++) {
                Ui[j] = A1(U, indx[j]); PLi[j] = A1(PL, indx[j]); Fi[j] = A1(f, A1(PL, indx[j]));
            }
            comprEig(H, &cap, 'c', Fi, D + LC(nodi, 1) - 1, Ui, V + LC(nodi, 1) - 1, DIFL, DIFR, S, f,
                     PLi, PU + LC(nodi, 1) - 1, tol, nt, mi, nodi, &Rk);
        }
        /* compact surviving column generators */
        lt = 1;
        for (int i = 1; i <= n1; i++) {
            int nodi = TTR(H, LL, i);
            int Rk = H->vsz[2 * (nodi - 1)];
            for (int j = 0; j < Rk; j++) {
                Vi[lt - 1 + j] = A1(V, LC(nodi, 1) + j); Di[lt - 1 + j] = A1(D, LC(nodi, 1) + j);
                PUi[lt - 1 + j] = A1(PU, LC(nodi, 1) + j);
            }
            LC(nodi, 1) = lt; LC(nodi, 2) = lt + Rk - 1; lt += Rk;
        }
        for (int LLt = LL - 1; LLt >= 1; LLt--) {
            for (int i = 1; i <= A1(H->lentr, LLt); i++) {
                int nodi = TTR(H, LLt, i);
                if (CH1(ch, nodi) == 0 && CH2(ch, nodi) == 0) {
                    int mi = LC(nodi, 2) - LC(nodi, 1) + 1;
                    for (int j = 0; j < mi; j++) {
                        Vi[lt - 1 + j] = A1(V, LC(nodi, 1) + j); Di[lt - 1 + j] = A1(D, LC(nodi, 1) + j);
                        PUi[lt - 1 + j] = A1(PU, LC(nodi, 1) + j);
                    }
                    LC(nodi, 1) = lt; LC(nodi, 2) = lt + mi - 1; lt += mi;
                }
            }
        }
        memcpy(V, Vi, sizeof(double) * (lt - 1));
        memcpy(D, Di, sizeof(double) * (lt - 1));
        memcpy(PU, PUi, sizeof(int) * (lt - 1));
        lca = lt - 1;
        /* B generators */
        for (int i = 1; i <= n1; i++) {
            int nodi = TTR(H, LL, i);
            int par = A1(H->tr, nodi);
            int nodj = (nodi == CH1(ch, par)) ? CH2(ch, par) : CH1(ch, par);
            int mi = LR(nodi, 2) - LR(nodi, 1) + 1;
            int nj = LC(nodj, 2) - LC(nodj, 1) + 1;
            long off = push(H, &cap, (long)mi * nj);
            cauchylikeEig(H->h + off, mi, f, U + LR(nodi, 1) - 1, V + LC(nodj, 1) - 1, DIFL, DIFR, S,
                          PL + LR(nodi, 1) - 1, PU + LC(nodj, 1) - 1, mi, nj);
            H->bsz[2 * (nodi - 1)] = mi; H->bsz[2 * (nodi - 1) + 1] = nj; H->pb[nodi - 1] = off;
            H->nflops += 6L * mi * nj;
        }
    }
#undef LR
#undef LC
    free(D); free(U); free(V); free(PL); free(PU); free(lr); free(lc); free(indx);
    free(Di); free(Fi); free(Ui); free(Vi); free(PLi); free(PUi);
    return 0;
}

int orc_hss_build(orc_hss *H, int K, int ni, const double *f, const double *d0, const double *u0,
                  const double *v0, const double *DIFL, const double *DIFR, double tol)
{
    return hss_build(H, K, ni, f, d0, u0, v0, DIFL, DIFR, NULL, tol);
}

int orc_hss_build_sq(orc_hss *H, int K, int ni, const double *f, const double *d0, const double *u0,
                     const double *v0, const double *DIFL, const double *DIFR, double tol)
{
    return hss_build(H, K, ni, f, d0, u0, v0, DIFL, DIFR, d0, tol);
}

static void gemm_acc(int m, int n, int k, const double *a, int lda, const double *b, int ldb,
                     double beta, double *c, int ldc)
{
    /* independent columns split over the threads; each summed in one thread's order */
#pragma omp parallel for schedule(static) if ((long)m * n * k > 4000000L)
    for (int j = 0; j < n; j++) {
        double *cj = c + (long)j * ldc;
        if (beta == 0.0) for (int i = 0; i < m; i++) cj[i] = 0.0;
        for (int l = 0; l < k; l++) {
            double blj = b[l + (long)j * ldb];
            const double *al = a + (long)l * lda;
            for (int i = 0; i < m; i++) cj[i] += al[i] * blj;
        }
    }
}

/* fastHssmmL_omp (ConstructHssd.f90:1394-1630): X := X * H.  No early return for
 * shallow trees: the sweeps below are valid for any depth >= 1. */
int orc_hss_mmL(const orc_hss *H, double *X, int ldx, int nrow)
{
    int ltr = H->ltr, lvl = H->lvl;
    const int *ch = H->ch, *tr = H->tr;
    int *lx = (int *)calloc(ltr + 1, sizeof(int)), *lg = (int *)calloc(ltr + 1, sizeof(int)),
        *lf = (int *)calloc(ltr + 1, sizeof(int));
    int lt = 1, i = 1;
    for (int nd = 1; nd <= ltr; nd++)
        if (CH1(ch, nd) == 0) { A1(lx, nd) = lt; lt += A1(H->m, i); i++; }
    int it = 1;
    for (int LL = lvl; LL >= 1; LL--)
        for (i = 1; i <= A1(H->lentr, LL); i++) { int nd = TTR(H, LL, i); A1(lg, nd) = it; it += H->usz[2 * (nd - 1) + 1]; }
    int lcg = it;
    it = 1;
    for (int LL = 1; LL <= lvl; LL++)
        for (i = 1; i <= A1(H->lentr, LL); i++) {
            int nd = TTR(H, LL, i), par = A1(tr, nd);
            int nj = (nd == CH1(ch, par)) ? CH2(ch, par) : CH1(ch, par);
            A1(lf, nd) = it; it += H->bsz[2 * (nj - 1) + 1];
        }
    int lcf = it;
    double *G = (double *)calloc((long)nrow * (lcg > 1 ? lcg : 1), sizeof(double));
    double *Fm = (double *)calloc((long)nrow * (lcf > 1 ? lcf : 1), sizeof(double));
    /* upsweep */
    for (int LL = lvl; LL >= 1; LL--)
        for (i = 1; i <= A1(H->lentr, LL); i++) {
            int nd = TTR(H, LL, i), c1 = CH1(ch, nd);
            int ur = H->usz[2 * (nd - 1)], uc = H->usz[2 * (nd - 1) + 1];
            const double *src = (c1 != 0) ? G + (long)(A1(lg, c1) - 1) * nrow : X + (long)(A1(lx, nd) - 1) * ldx;
            int lds = (c1 != 0) ? nrow : ldx;
            gemm_acc(nrow, uc, ur, src, lds, H->h + H->pu[nd - 1], ur, 0.0, G + (long)(A1(lg, nd) - 1) * nrow, nrow);
        }
    /* sibling products and downsweep */
    for (int LL = 1; LL <= lvl; LL++)
        for (i = 1; i <= A1(H->lentr, LL); i++) {
            int nd = TTR(H, LL, i), par = A1(tr, nd);
            int nj = (nd == CH1(ch, par)) ? CH2(ch, par) : CH1(ch, par);
            int br = H->bsz[2 * (nj - 1)], bc = H->bsz[2 * (nj - 1) + 1];
            gemm_acc(nrow, bc, br, G + (long)(A1(lg, nj) - 1) * nrow, nrow, H->h + H->pb[nj - 1], br,
                     LL == 1 ? 0.0 : 1.0, Fm + (long)(A1(lf, nd) - 1) * nrow, nrow);
            int c1 = CH1(ch, nd);
            if (c1 != 0) {
                int vr = H->vsz[2 * (nd - 1)], vc = H->vsz[2 * (nd - 1) + 1];
                gemm_acc(nrow, vc, vr, Fm + (long)(A1(lf, nd) - 1) * nrow, nrow, H->h + H->pv[nd - 1], vr,
                         0.0, Fm + (long)(A1(lf, c1) - 1) * nrow, nrow);
            }
        }
    /* leaves: X_i := F_i V_i + X_i D_i */
    double *T2 = (double *)malloc(sizeof(double) * (long)nrow * (H->k > 0 ? H->k : 1));
    for (int nd = 1; nd <= ltr - 1; nd++) {
        if (CH1(ch, nd) != 0) continue;
        int vr = H->vsz[2 * (nd - 1)], vc = H->vsz[2 * (nd - 1) + 1], md = H->dsz[nd - 1];
        gemm_acc(nrow, vc, vr, Fm + (long)(A1(lf, nd) - 1) * nrow, nrow, H->h + H->pv[nd - 1], vr, 0.0, T2, nrow);
        gemm_acc(nrow, md, md, X + (long)(A1(lx, nd) - 1) * ldx, ldx, H->dd + H->pd[nd - 1], md, 1.0, T2, nrow);
        for (int j = 0; j < md; j++)
            memcpy(X + (long)(A1(lx, nd) - 1 + j) * ldx, T2 + (long)j * nrow, sizeof(double) * nrow);
    }
    free(T2); free(G); free(Fm); free(lx); free(lg); free(lf);
    return 0;
}

void orc_hss_dense(const orc_hss *H, double *a, int lda)
{
    int K = H->k;
    for (int j = 0; j < K; j++)
        for (int i = 0; i < K; i++) a[i + (long)j * lda] = (i == j) ? 1.0 : 0.0;
    orc_hss_mmL(H, a, lda, K);
}

void orc_cauchy_dense(int k, const double *f, const double *u, const double *v,
                      const double *difl, const double *difr, double *a, int lda)
{
    int *p = (int *)malloc(sizeof(int) * k);
    for (int i = 0; i < k; i++) p[i] = i + 1;
    cauchylikeEig(a, lda, f, u, v, difl, difr, NULL, p, p, k, k);
    free(p);
}

void orc_cauchy_dense_sq(int k, const double *f, const double *d0, const double *u, const double *v,
                         const double *difl, const double *difr, double *a, int lda)
{
    int *p = (int *)malloc(sizeof(int) * k);
    for (int i = 0; i < k; i++) p[i] = i + 1;
    cauchylikeEig(a, lda, f, u, v, difl, difr, d0, p, p, k, k);
    free(p);
}

/* =====================================================================
 * General Cauchy-like matrices A(i,j) = u(i) v(j) / (d(i) - f(j)) with direct
 * differences (CauchyHss_VP2.f90): the structured RRLU rrluCauchy (:7-134, rho =
 * 1.02, a swap check every 10 steps) with its pivot searches CauchyPivtG (:191-243),
 * CauchyPivtG_CP (:246-296), searchMax2 (:137-163), CauchyMax (:166-188); the
 * interpolative generators of comprcauchy (:616-725) through Cauchylike2
 * (:743-778); and the stack-ordered post-order construction cauchy2hss
 * (:387-613).  Statement order is the Fortran's; the result goes into the same
 * orc_hss layout as the eigenvector construction (U_i: M x r, V_i: r x M, B_i of a
 * node: its row skeleton against its sibling's column skeleton, leaf D_i dense).
 * ===================================================================== */

/* CauchyMax: the largest |u(i) / (D(i) - F)|, i = 1..n, and its first index */
static void cmaxG(const double *D, double F, const double *u, int n, double *junk, int *jj)
{
    double best = 0.0;
    int bi = 1;
    for (int i = 1; i <= n; i++) {
        double x = fabs(A1(u, i) / (A1(D, i) - F));
        if (i == 1 || x > best) { best = x; bi = i; }
    }
    *junk = best;
    *jj = bi;
}

/* CauchyPivtG: rook pivoting on the k-th Schur complement */
static void pivtG(double *D, double *F, double *u, double *v, int *PL, int *PU, int k, double *A, int M, int N)
{
    double junkL, junkU;
    int jjL, jjU;
    cmaxG(D + k - 1, A1(F, k), u + k - 1, M - k + 1, &junkL, &jjL);
    cmaxG(F + k - 1, A1(D, k), v + k - 1, N - k + 1, &junkU, &jjU);
    junkL = junkL * fabs(A1(v, k));
    junkU = junkU * fabs(A1(u, k));
    double piv = fabs(A1(u, k) * A1(v, k) / (A1(D, k) - A1(F, k)));
    if (junkL <= piv && junkU <= piv) return;
    int flg = (junkL > junkU) ? 1 : 0;
    for (;;) {
        if (flg == 1) {
            int jj = jjL;
            dswp(D, k, jj + k - 1); dswp(u, k, jj + k - 1); dswp(A, k, jj + k - 1); iswp(PL, k, jj + k - 1);
            cmaxG(F + k - 1, A1(D, k), v + k - 1, N - k + 1, &junkU, &jjU);
            if (jjU == 1) return;
            flg = 0;
        }
        int jj = jjU;
        dswp(F, k, jj + k - 1); dswp(v, k, jj + k - 1); iswp(PU, k, jj + k - 1);
        cmaxG(D + k - 1, A1(F, k), u + k - 1, M - k + 1, &junkL, &jjL);
        if (jjL == 1) return;
        flg = 1;
    }
}

/* CauchyPivtG_CP: complete pivoting over the k-th Schur complement */
static void pivtGCP(double *D, double *F, double *u, double *v, int *PL, int *PU, int k, double *A, int M, int N)
{
    double best = 0.0;
    int ii = 1, jj = 1;
    for (int j = 1; j <= N - k + 1; j++) {
        double junk;
        int iiu;
        cmaxG(D + k - 1, A1(F, k + j - 1), u + k - 1, M - k + 1, &junk, &iiu);
        junk = fabs(junk * A1(v, k + j - 1));
        if (j == 1 || junk > best) { best = junk; ii = iiu; jj = j; }
    }
    dswp(D, k, ii + k - 1); dswp(u, k, ii + k - 1); dswp(A, k, ii + k - 1); iswp(PL, k, ii + k - 1);
    dswp(F, k, jj + k - 1); dswp(v, k, jj + k - 1); iswp(PU, k, jj + k - 1);
}

/* searchMax2: the largest entry of Mat_N(i, j) = U(i) W(j) / (D2(i) - D1(j)), column by column */
static void searchMax2G(const double *U, const double *W, const double *D2, const double *D1, int ldu, int ldw,
                        int *ii, int *jj, double *nmax)
{
    double junk;
    int jjl;
    cmaxG(D2, A1(D1, 1), U, ldu, &junk, &jjl);
    junk = junk * fabs(A1(W, 1));
    *nmax = junk; *ii = 1; *jj = jjl;
    for (int j = 2; j <= ldw; j++) {
        cmaxG(D2, A1(D1, j), U, ldu, &junk, &jjl);
        junk = junk * fabs(A1(W, j));
        if (junk > *nmax) { *nmax = junk; *ii = j; *jj = jjl; }
    }
}

/* rrluCauchy: structured RRLU of the M x N Cauchy-like matrix A(i)V(j)/(D(i)-F(j)) */
static void rrluG(double *D, double *F, double *A, double *V, double tol, double *U, double *W, int *PL, int *PU,
                  int M, int N, int *Rk, long long *nswap)
{
    int mn = M < N ? M : N, j, k, ii, jj;
    double amax = 0.0, ukk, nmax;
    const double rho = 1.02;
    const int prd = 10;
    for (j = 1; j <= M; j++) { A1(U, j) = A1(A, j); A1(PL, j) = j; }
    for (j = 1; j <= N; j++) A1(PU, j) = j;
    *Rk = mn;
    for (k = 1; k <= mn; k++) {
        pivtG(D, F, U, V, PL, PU, k, A, M, N);
        ukk = A1(U, k) * A1(V, k) / (A1(D, k) - A1(F, k));
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            pivtGCP(D, F, U, V, PL, PU, k, A, M, N);
            ukk = A1(U, k) * A1(V, k) / (A1(D, k) - A1(F, k));
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { *Rk = k - 1; break; }
        }
        for (j = k + 1; j <= M; j++) A1(U, j) = A1(U, j) * (A1(D, j) - A1(D, k)) / (A1(D, j) - A1(F, k));
        for (j = k + 1; j <= N; j++) A1(V, j) = A1(V, j) * (A1(F, j) - A1(F, k)) / (A1(F, j) - A1(D, k));
        for (j = 1; j <= k - 1; j++) A1(W, j) = A1(W, j) * ((A1(F, k) - A1(D, j)) / (A1(D, k) - A1(D, j)));
        A1(W, k) = (A1(D, k) - A1(F, k)) / A1(A, k);
        for (j = 1; j <= k - 1; j++) A1(W, k) = A1(W, k) * (A1(F, j) - A1(D, k)) / (A1(D, j) - A1(D, k));
        int flgl = k % prd;
        while (flgl < 1) {
            flgl = 1;
            searchMax2G(U + k, W, D + k, D, M - k, k, &ii, &jj, &nmax);
            if (nmax > rho) {
                if (nswap) (*nswap)++;
                flgl = 0;
                jj = jj + k;
                for (j = k + 1; j <= N; j++) A1(V, j) = A1(V, j) * ((A1(F, j) - A1(D, ii)) / (A1(F, j) - A1(D, jj)));
                for (j = k + 1; j <= jj - 1; j++)
                    A1(U, j) = A1(U, j) * ((A1(D, j) - A1(D, jj)) / (A1(D, j) - A1(D, ii)));
                for (j = jj + 1; j <= M; j++) A1(U, j) = A1(U, j) * ((A1(D, j) - A1(D, jj)) / (A1(D, j) - A1(D, ii)));
                for (j = 1; j <= ii - 1; j++) A1(W, j) = A1(W, j) * ((A1(D, j) - A1(D, ii)) / (A1(D, j) - A1(D, jj)));
                for (j = ii + 1; j <= k; j++) A1(W, j) = A1(W, j) * ((A1(D, j) - A1(D, ii)) / (A1(D, j) - A1(D, jj)));
                A1(U, jj) = A1(A, ii) * ((A1(D, ii) - A1(D, jj)) / (A1(D, ii) - A1(F, ii)));
                A1(W, ii) = (A1(D, jj) - A1(F, ii)) / A1(A, jj);
                for (j = 1; j <= ii - 1; j++) {
                    A1(U, jj) = A1(U, jj) * ((A1(D, ii) - A1(D, j)) / (A1(D, ii) - A1(F, j)));
                    A1(W, ii) = A1(W, ii) * ((A1(F, j) - A1(D, jj)) / (A1(D, j) - A1(D, jj)));
                }
                for (j = ii + 1; j <= k; j++) {
                    A1(U, jj) = A1(U, jj) * ((A1(D, ii) - A1(D, j)) / (A1(D, ii) - A1(F, j)));
                    A1(W, ii) = A1(W, ii) * ((A1(F, j) - A1(D, jj)) / (A1(D, j) - A1(D, jj)));
                }
                iswp(PL, ii, jj);
                dswp(D, ii, jj);
                dswp(A, ii, jj);
            }
        }
    }
}

/* comprcauchy: rows ('r', M node rows against N outside columns) or columns ('c', M
 * node columns against N outside rows) of one node; generators into H */
static void comprG(orc_hss *H, long *cap, char rc, double *D, double *F, double *U, double *V, double tol, int M,
                   int N, int nodi, int *Rk)
{
    int mn = M < N ? M : N;
    double *Z = (double *)calloc(M + 1, sizeof(double)), *W = (double *)calloc(mn + 1, sizeof(double));
    int *PL = (int *)malloc(sizeof(int) * (M + 1)), *PU = (int *)malloc(sizeof(int) * (N + 1));
    if (rc == 'r') rrluG(D, F, U, V, tol, Z, W, PL, PU, M, N, Rk, &H->nswap);
    else rrluG(F, D, V, U, tol, Z, W, PL, PU, M, N, Rk, &H->nswap);
    invp(PL, M);
    int r = *Rk;
    const double *X = (rc == 'r') ? D : F;  /* the compressed side's points */
    if (rc == 'r') {
        long off = push(H, cap, (long)M * (r < M ? r : M));
        if (r < M) {
            /* Cauchylike2(Q, D(r+1), D, Z(r+1), W, M, r), then rows Q(PL, :) */
            for (int j = 1; j <= r; j++)
                for (int i = 1; i <= M; i++) {
                    int s = A1(PL, i);
                    double q;
                    if (s <= r) q = (s == j) ? 1.0 : 0.0;
                    else q = A1(Z, s) * A1(W, j) / (A1(X, s) - A1(X, j));
                    H->h[off + (i - 1) + (long)(j - 1) * M] = q;
                }
            H->usz[2 * (nodi - 1)] = M; H->usz[2 * (nodi - 1) + 1] = r;
        } else {
            for (int j = 1; j <= M; j++)
                for (int i = 1; i <= M; i++) H->h[off + (i - 1) + (long)(j - 1) * M] = (A1(PL, i) == j) ? 1.0 : 0.0;
            H->usz[2 * (nodi - 1)] = M; H->usz[2 * (nodi - 1) + 1] = M;
        }
        H->pu[nodi - 1] = off;
    } else {
        long off = push(H, cap, (long)(r < M ? r : M) * M);
        if (r < M) {
            /* Cauchylike2(Q, F(r+1), F, Z(r+1), W, r, M), then columns Q(:, PL) */
            for (int j = 1; j <= M; j++) {
                int s = A1(PL, j);
                for (int i = 1; i <= r; i++) {
                    double q;
                    if (s <= r) q = (s == i) ? 1.0 : 0.0;
                    else q = A1(W, i) * A1(Z, s) / (A1(X, s) - A1(X, i));
                    H->h[off + (i - 1) + (long)(j - 1) * r] = q;
                }
            }
            H->vsz[2 * (nodi - 1)] = r; H->vsz[2 * (nodi - 1) + 1] = M;
        } else {
            for (int j = 1; j <= M; j++)
                for (int i = 1; i <= M; i++) H->h[off + (i - 1) + (long)(j - 1) * M] = (A1(PL, j) == i) ? 1.0 : 0.0;
            H->vsz[2 * (nodi - 1)] = M; H->vsz[2 * (nodi - 1) + 1] = M;
        }
        H->pv[nodi - 1] = off;
    }
    free(Z); free(W); free(PL); free(PU);
}

/* Cauchylike (Cauchylike.f90:1-13): A(i,j) = u(i) v(j) / (D(i) - F(j)) */
static void cauchyG(double *a, int lda, const double *D, const double *F, const double *u, const double *v, int M,
                    int N)
{
    for (int j = 1; j <= N; j++)
        for (int i = 1; i <= M; i++) a[(i - 1) + (long)(j - 1) * lda] = A1(u, i) * A1(v, j) / (A1(D, i) - A1(F, j));
}

int orc_cauchy2hss(orc_hss *H, int n, int ni, const double *d0, const double *f0, const double *u0, const double *v0,
                   double tol)
{
    if (hss_tree_setup(H, n, ni) < 1) return -1;
    const int ltr = H->ltr, lvl = H->lvl, lda = n;
    long cap = 10L * ni * n;
    int *ch = H->ch;
    double *D = (double *)malloc(sizeof(double) * n), *F = (double *)malloc(sizeof(double) * n),
           *U = (double *)malloc(sizeof(double) * n), *V = (double *)malloc(sizeof(double) * n);
    memcpy(D, d0, sizeof(double) * n); memcpy(F, f0, sizeof(double) * n);
    memcpy(U, u0, sizeof(double) * n); memcpy(V, v0, sizeof(double) * n);
    int *l = (int *)calloc(2 * ltr + 2, sizeof(int)), *lsr = (int *)calloc(lvl + 3, sizeof(int)),
        *lsc = (int *)calloc(lvl + 3, sizeof(int)), *indx1 = (int *)malloc(sizeof(int) * n),
        *indx2 = (int *)malloc(sizeof(int) * n);
    double *Di = (double *)malloc(sizeof(double) * n), *Fi = (double *)malloc(sizeof(double) * n),
           *Ui = (double *)malloc(sizeof(double) * n), *Vi = (double *)malloc(sizeof(double) * n);
#define L(i, c) l[2 * ((i) - 1) + (c) - 1]
    int lt = 1, it = 1;
    long pnd = 0;
    for (int i = 1; i <= ltr; i++) {
        if (CH1(ch, i) == 0 && CH2(ch, i) == 0) {
            L(i, 1) = lt; L(i, 2) = lt + A1(H->m, it) - 1;
            lt = L(i, 2) + 1;
            it++;
        } else {
            L(i, 1) = L(CH1(ch, i), 1); L(i, 2) = L(CH2(ch, i), 2);
        }
    }
    int ns = 1, Rk, mi, nii, nt;
    A1(lsr, ns) = 1;
    A1(lsc, ns) = 1;
    for (int i = 1; i <= ltr; i++) {
        if (CH1(ch, i) == 0) {  /* leaf */
            mi = L(i, 2) - L(i, 1) + 1;
            nii = lda - L(i, 2);
            for (int j = 0; j < mi; j++) indx1[j] = L(i, 1) + j;
            cauchyG(H->dd + pnd, mi, D + L(i, 1) - 1, F + L(i, 1) - 1, U + L(i, 1) - 1, V + L(i, 1) - 1, mi, mi);
            H->dsz[i - 1] = mi;
            H->pd[i - 1] = pnd;
            pnd += (long)mi * mi;
            /* off-diagonal row compression: the stacked column skeletons, then the columns to the right */
            if (ns == 1) {
                for (int j = 0; j < nii; j++) indx2[j] = L(i, 2) + 1 + j;
                nt = nii;
            } else {
                for (int j = 0; j < A1(lsc, ns) - 1; j++) indx2[j] = j + 1;
                for (int j = 0; j < nii; j++) indx2[A1(lsc, ns) - 1 + j] = L(i, 2) + 1 + j;
                nt = nii + A1(lsc, ns) - 1;
            }
            for (int j = 0; j < mi; j++) { Di[j] = A1(D, indx1[j]); Ui[j] = A1(U, indx1[j]); }
            for (int j = 0; j < nt; j++) { Fi[j] = A1(F, indx2[j]); Vi[j] = A1(V, indx2[j]); }
            comprG(H, &cap, 'r', Di, Fi, Ui, Vi, tol, mi, nt, i, &Rk);
            A1(lsr, ns + 1) = A1(lsr, ns) + Rk;
            for (int j = 0; j < Rk; j++) { A1(D, A1(lsr, ns) + j) = Di[j]; A1(U, A1(lsr, ns) + j) = Ui[j]; }
            /* off-diagonal column compression */
            if (ns == 1) {
                for (int j = 0; j < nii; j++) indx2[j] = L(i, 2) + 1 + j;
                nt = nii;
            } else {
                for (int j = 0; j < A1(lsr, ns) - 1; j++) indx2[j] = j + 1;
                for (int j = 0; j < nii; j++) indx2[A1(lsr, ns) - 1 + j] = L(i, 2) + 1 + j;
                nt = nii + A1(lsr, ns) - 1;
            }
            for (int j = 0; j < nt; j++) { Di[j] = A1(D, indx2[j]); Ui[j] = A1(U, indx2[j]); }
            for (int j = 0; j < mi; j++) { Fi[j] = A1(F, indx1[j]); Vi[j] = A1(V, indx1[j]); }
            comprG(H, &cap, 'c', Di, Fi, Ui, Vi, tol, mi, nt, i, &Rk);
            A1(lsc, ns + 1) = A1(lsc, ns) + Rk;
            for (int j = 0; j < Rk; j++) { A1(F, A1(lsc, ns) + j) = Fi[j]; A1(V, A1(lsc, ns) + j) = Vi[j]; }
            ns++;
        } else {  /* parent: the B generators of its children, then its own compressions */
            int c1 = CH1(ch, i), c2 = CH2(ch, i);
            mi = A1(lsr, ns - 1) - A1(lsr, ns - 2);
            nii = A1(lsc, ns) - A1(lsc, ns - 1);
            long off = push(H, &cap, (long)mi * nii);
            cauchyG(H->h + off, mi, D + A1(lsr, ns - 2) - 1, F + A1(lsc, ns - 1) - 1, U + A1(lsr, ns - 2) - 1,
                    V + A1(lsc, ns - 1) - 1, mi, nii);
            H->bsz[2 * (c1 - 1)] = mi; H->bsz[2 * (c1 - 1) + 1] = nii; H->pb[c1 - 1] = off;
            mi = A1(lsr, ns) - A1(lsr, ns - 1);
            nii = A1(lsc, ns - 1) - A1(lsc, ns - 2);
            off = push(H, &cap, (long)mi * nii);
            cauchyG(H->h + off, mi, D + A1(lsr, ns - 1) - 1, F + A1(lsc, ns - 2) - 1, U + A1(lsr, ns - 1) - 1,
                    V + A1(lsc, ns - 2) - 1, mi, nii);
            H->bsz[2 * (c2 - 1)] = mi; H->bsz[2 * (c2 - 1) + 1] = nii; H->pb[c2 - 1] = off;
            if (i == ltr) break;
            /* off-diagonal row compression of the merged skeleton rows */
            mi = A1(lsr, ns) - A1(lsr, ns - 2);
            nii = lda - L(i, 2);
            for (int j = 0; j < mi; j++) indx1[j] = A1(lsr, ns - 2) + j;
            if (ns == 3) {
                for (int j = 0; j < nii; j++) indx2[j] = L(i, 2) + 1 + j;
                nt = nii;
            } else {
                for (int j = 0; j < A1(lsc, ns - 2) - 1; j++) indx2[j] = j + 1;
                for (int j = 0; j < nii; j++) indx2[A1(lsc, ns - 2) - 1 + j] = L(i, 2) + 1 + j;
                nt = nii + A1(lsc, ns - 2) - 1;
            }
            for (int j = 0; j < mi; j++) { Di[j] = A1(D, indx1[j]); Ui[j] = A1(U, indx1[j]); }
            for (int j = 0; j < nt; j++) { Fi[j] = A1(F, indx2[j]); Vi[j] = A1(V, indx2[j]); }
            comprG(H, &cap, 'r', Di, Fi, Ui, Vi, tol, mi, nt, i, &Rk);
            A1(lsr, ns - 1) = A1(lsr, ns - 2) + Rk;
            for (int j = 0; j < Rk; j++) { A1(D, A1(lsr, ns - 2) + j) = Di[j]; A1(U, A1(lsr, ns - 2) + j) = Ui[j]; }
            /* off-diagonal column compression of the merged skeleton columns */
            mi = A1(lsc, ns) - A1(lsc, ns - 2);
            nii = lda - L(i, 2);
            for (int j = 0; j < mi; j++) indx1[j] = A1(lsc, ns - 2) + j;
            if (ns == 3) {
                for (int j = 0; j < nii; j++) indx2[j] = L(i, 2) + 1 + j;
                nt = nii;
            } else {
                for (int j = 0; j < A1(lsr, ns - 2) - 1; j++) indx2[j] = j + 1;
                for (int j = 0; j < nii; j++) indx2[A1(lsr, ns - 2) - 1 + j] = L(i, 2) + 1 + j;
                nt = nii + A1(lsr, ns - 2) - 1;
            }
            for (int j = 0; j < nt; j++) { Di[j] = A1(D, indx2[j]); Ui[j] = A1(U, indx2[j]); }
            for (int j = 0; j < mi; j++) { Fi[j] = A1(F, indx1[j]); Vi[j] = A1(V, indx1[j]); }
            comprG(H, &cap, 'c', Di, Fi, Ui, Vi, tol, mi, nt, i, &Rk);
            A1(lsc, ns - 1) = A1(lsc, ns - 2) + Rk;
            for (int j = 0; j < Rk; j++) { A1(F, A1(lsc, ns - 2) + j) = Fi[j]; A1(V, A1(lsc, ns - 2) + j) = Vi[j]; }
            ns--;
        }
    }
#undef L
    free(D); free(F); free(U); free(V); free(l); free(lsr); free(lsc); free(indx1); free(indx2);
    free(Di); free(Fi); free(Ui); free(Vi);
    return 0;
}

void orc_cauchy_dense_general(int n, const double *d, const double *f, const double *u, const double *v, double *a,
                              int lda)
{
    cauchyG(a, lda, d, f, u, v, n, n);
}

/* ------------------------------------------------------------------------------------
 * Dense matrix -> HSS by truncated SVDs (the construction half of DHSSMM / DHSSMML,
 * DHSSMM.f90:1-125, DHSSMML.f90:2-122, mat2hssvd ConstructHssd.f90:371-696), restated
 * level by level: the tree is npart's (the same as cauchy2hss's); every node of one
 * depth is compressed from the current compressed matrix, whose row and column blocks
 * are the active nodes' bases (a raw leaf: the identity).  A node's row basis U spans
 * the left singular vectors of its block row (its rows against every other active
 * node's columns) whose singular values exceed tol, its column basis W the right
 * singular vectors of its block column likewise (mat2hssvd's count(svs > tol)); the
 * generator V is W^T.  A sibling pair's B is the compressed block between them after
 * both are compressed, a leaf's D its diagonal block of A.  Each SVD is Householder QR
 * of the tall side, then one-sided Jacobi on R (orc_jacobi_svd).
 * ------------------------------------------------------------------------------------ */

/* One-sided Jacobi on the p columns (length l >= p) of c (ldc): pairs in round-robin
 * order, a pair rotates when |c_a . c_b| > l eps |c_a| |c_b|; afterwards the columns are
 * orthogonal, their norms the singular values.  sv (p) descending, u (l x p, ldu) the
 * normalised columns in that order (zero for a zero column).  Returns the sweeps, < 0
 * when 60 did not converge. */
static int jacobi_cols(int l, int p, double *c, long ldc, double *sv, double *u, long ldu)
{
    const int P = p + (p & 1);
    int *pos = (int *)malloc(sizeof(int) * (P > 0 ? P : 1)), conv = 0, sweeps;
    for (sweeps = 1; sweeps <= 60 && !conv; sweeps++) {
        int rot = 0;
        for (int i = 0; i < P; i++) pos[i] = i;
        for (int r = 0; r < P - 1; r++) {
            for (int i = 0; i < P / 2; i++) {
                int a = pos[i], b = pos[P - 1 - i];
                if (a > b) { int t = a; a = b; b = t; }
                if (b >= p) continue;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = 0; k < l; k++) {
                    const double x = c[k + (long)a * ldc], y = c[k + (long)b * ldc];
                    al += x * x; be += y * y; ga += x * y;
                }
                if (ga == 0.0 || !(fabs(ga) > l * 0x1p-53 * sqrt(al * be))) continue;
                const double ze = (be - al) / (2.0 * ga);
                const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                const double cc = 1.0 / sqrt(1.0 + t * t), ss = cc * t;
                for (int k = 0; k < l; k++) {
                    const double x = c[k + (long)a * ldc], y = c[k + (long)b * ldc];
                    c[k + (long)a * ldc] = cc * x - ss * y;
                    c[k + (long)b * ldc] = ss * x + cc * y;
                }
                rot++;
            }
            const int last = pos[P - 1];
            for (int i = P - 1; i > 1; i--) pos[i] = pos[i - 1];
            if (P > 1) pos[1] = last;
        }
        if (rot == 0) conv = 1;
    }
    int *ord = (int *)malloc(sizeof(int) * (p > 0 ? p : 1));
    double *nr = (double *)malloc(sizeof(double) * (p > 0 ? p : 1));
    for (int j = 0; j < p; j++) {
        double q = 0.0;
        for (int k = 0; k < l; k++) q += c[k + (long)j * ldc] * c[k + (long)j * ldc];
        nr[j] = sqrt(q);
        ord[j] = j;
    }
    for (int i = 0; i + 1 < p; i++) {  /* descending, ties by index */
        int k = i;
        for (int j = i + 1; j < p; j++) if (nr[ord[j]] > nr[ord[k]]) k = j;
        const int t = ord[i]; ord[i] = ord[k]; ord[k] = t;
    }
    for (int j = 0; j < p; j++) {
        const int s0 = ord[j];
        sv[j] = nr[s0];
        for (int k = 0; k < l; k++) u[k + (long)j * ldu] = nr[s0] > 0.0 ? c[k + (long)s0 * ldc] / nr[s0] : 0.0;
    }
    free(pos); free(ord); free(nr);
    return conv ? sweeps - 1 : -1;
}

/* right singular vectors of the w x m matrix T (ldt, overwritten): v (m x m, columns in
 * descending singular value order, the first min(w, m) meaningful), sv (m); returns the
 * count above tol.  w >= m: Householder QR of T, then one-sided Jacobi on R's columns
 * (their rotations are the right vectors); w < m: T^T has fewer columns than rows and
 * its own columns are orthogonalised (their normalised results are the vectors) -- R
 * would hold m columns in a w-dimensional space, which Jacobi cannot make orthogonal. */
static int right_basis(int w, int m, double *T, long ldt, double tol, double *v, double *sv)
{
    if (w < m) {
        double *X = (double *)malloc(sizeof(double) * (size_t)m * (w > 0 ? w : 1));
        for (int j = 0; j < w; j++)
            for (int i = 0; i < m; i++) X[i + (long)j * m] = T[j + (long)i * ldt];
        for (long e = 0; e < (long)m * m; e++) v[e] = 0.0;
        for (int j = 0; j < m; j++) sv[j] = 0.0;
        int cnt = 0;
        if (w > 0 && jacobi_cols(m, w, X, m, sv, v, m) < 0) cnt = -1;
        else
            for (int j = 0; j < w; j++) cnt += sv[j] > tol;
        free(X);
        return cnt;
    }
    const int p = m;
    for (int k = 0; k < p; k++) {
        double *x = T + k + (long)k * ldt;
        const int len = w - k;
        double xn = 0.0;
        for (int i = 1; i < len; i++) xn += x[i] * x[i];
        xn = sqrt(xn);
        const double alpha = x[0];
        if (xn == 0.0) continue;  /* H = I */
        const double beta = -copysign(hypot(alpha, xn), alpha);
        const double tau = (beta - alpha) / beta, scal = 1.0 / (alpha - beta);
        for (int i = 1; i < len; i++) x[i] *= scal;  /* v = (1, x(1:)) */
        for (int j = k + 1; j < m; j++) {
            double *c = T + k + (long)j * ldt;
            double dot = c[0];
            for (int i = 1; i < len; i++) dot += x[i] * c[i];
            dot *= tau;
            c[0] -= dot;
            for (int i = 1; i < len; i++) c[i] -= dot * x[i];
        }
        x[0] = beta;
    }
    double *R = (double *)calloc((size_t)m * m, sizeof(double)), *u = (double *)malloc(sizeof(double) * m * m);
    for (int j = 0; j < m; j++)
        for (int i = 0; i <= j && i < p; i++) R[i + (long)j * m] = T[i + (long)j * ldt];
    int cnt = 0;
    if (orc_jacobi_svd(m, R, m, sv, u, m, v, m) < 0) cnt = -1;
    else
        for (int j = 0; j < m; j++) cnt += sv[j] > tol;
    free(R); free(u);
    return cnt;
}

/* Entry i of mat2hssvdr's flat array of uniform [0, 1) numbers: splitmix64 of
 * seed + (i + 1) 0x9E3779B97F4A7C15, top 53 bits (the reference draws it unseeded with
 * random_number, ConstructHssd.f90:100). */
static double m2h_omega(unsigned long long seed, unsigned long long i)
{
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

/* The basis of one side: from T (w x m, context x own) itself, or -- mat2hssvdr's
 * sampling test, a context longer than 50 and than m + 10 (ConstructHssd.f90:101, 115,
 * 144, ...) -- from the sketch S = Omega T (m + 10 rows), Omega read from the flat
 * array with leading dimension m + 10 for a block column (side 1, :104) and as the
 * transpose of a w x (m + 10) array for a block row (side 0, :118). */
static int side_basis(int w, int m, double *T, double tol, double *v, double *sv, int rsvd, unsigned long long seed,
                      int side)
{
    if (!(rsvd && w > 50 && w > m + 10)) return right_basis(w, m, T, w > 0 ? w : 1, tol, v, sv);
    const int pp = m + 10;
    double *S = (double *)malloc(sizeof(double) * (size_t)pp * m);
    for (int t = 0; t < m; t++)
        for (int r = 0; r < pp; r++) {
            double acc = 0.0;
            for (int o = 0; o < w; o++) {
                const unsigned long long idx = side ? (unsigned long long)r + (unsigned long long)o * pp
                                                    : (unsigned long long)o + (unsigned long long)r * w;
                acc += m2h_omega(seed, idx) * T[o + (long)t * w];
            }
            S[r + (long)t * pp] = acc;
        }
    const int rk = right_basis(pp, m, S, pp, tol, v, sv);
    free(S);
    return rk;
}

static int mat2hss_impl(orc_hss *H, int n, int ni, const double *A, int lda, double tol, int rsvd,
                        unsigned long long seed);

int orc_mat2hss(orc_hss *H, int n, int ni, const double *A, int lda, double tol)
{
    return mat2hss_impl(H, n, ni, A, lda, tol, 0, 0);
}

int orc_mat2hss_rand(orc_hss *H, int n, int ni, const double *A, int lda, double tol, unsigned long long seed)
{
    return mat2hss_impl(H, n, ni, A, lda, tol, 1, seed);
}

static int mat2hss_impl(orc_hss *H, int n, int ni, const double *A, int lda, double tol, int rsvd,
                        unsigned long long seed)
{
    if (hss_tree_setup(H, n, ni) < 1) return -1;
    const int ltr = H->ltr, lvl = H->lvl;
    long cap = 10L * ni * n;
    const int *ch = H->ch;
    /* leaf ranges and depths */
    int *r0 = (int *)calloc(ltr + 1, sizeof(int)), *dep = (int *)calloc(ltr + 1, sizeof(int));
    {
        int lt = 0, it = 1;
        for (int nd = 1; nd <= ltr; nd++)
            if (CH1(ch, nd) == 0) { A1(r0, nd) = lt; lt += A1(H->m, it); it++; }
        for (int i = ltr - 1; i >= 1; i--) A1(dep, i) = A1(dep, A1(H->tr, i)) + 1;
    }
    /* active nodes in left-to-right order with their row / column blocks in the
     * compressed matrix C (nr x nc, column-major) */
    int na = 0;
    int *act = (int *)malloc(sizeof(int) * ltr), *ars = (int *)malloc(sizeof(int) * ltr),
        *arn = (int *)malloc(sizeof(int) * ltr), *acs = (int *)malloc(sizeof(int) * ltr),
        *acn = (int *)malloc(sizeof(int) * ltr);
    {
        int it = 1;
        for (int nd = 1; nd <= ltr; nd++)
            if (CH1(ch, nd) == 0) {
                act[na] = nd; ars[na] = acs[na] = A1(r0, nd); arn[na] = acn[na] = A1(H->m, it); na++; it++;
            }
    }
    int nr = n, nc = n;
    double *C = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; j++) memcpy(C + (size_t)j * n, A + (size_t)j * lda, sizeof(double) * n);
    long pnd = 0;
    /* leaf D blocks of A */
    {
        int it = 1;
        for (int nd = 1; nd <= ltr; nd++)
            if (CH1(ch, nd) == 0) {
                const int mi = A1(H->m, it), s0 = A1(r0, nd);
                for (int j = 0; j < mi; j++)
                    for (int i = 0; i < mi; i++) H->dd[pnd + i + (long)j * mi] = A[(s0 + i) + (long)(s0 + j) * lda];
                H->dsz[nd - 1] = mi; H->pd[nd - 1] = pnd; pnd += (long)mi * mi; it++;
            }
    }
    int info = 0;
    for (int LL = lvl; LL >= 1 && !info; LL--) {
        /* the nodes of this depth and, for each, its constituents in the active list */
        int nn = A1(H->lentr, LL);
        int *nds = (int *)malloc(sizeof(int) * nn), *a0 = (int *)malloc(sizeof(int) * nn),
            *a1 = (int *)malloc(sizeof(int) * nn);
        for (int q = 0; q < nn; q++) {
            const int nd = TTR(H, LL, q + 1);
            nds[q] = nd;
            const int f = CH1(ch, nd) ? CH1(ch, nd) : nd, l = CH1(ch, nd) ? CH2(ch, nd) : nd;
            a0[q] = a1[q] = -1;
            for (int t = 0; t < na; t++) { if (act[t] == f) a0[q] = t; if (act[t] == l) a1[q] = t; }
        }
        double **Ub = (double **)calloc(nn, sizeof(double *)), **Wb = (double **)calloc(nn, sizeof(double *));
        int *ru = (int *)calloc(nn, sizeof(int)), *rw = (int *)calloc(nn, sizeof(int)),
            *mr = (int *)calloc(nn, sizeof(int)), *mc = (int *)calloc(nn, sizeof(int));
        for (int q = 0; q < nn && !info; q++) {
            const int rs = ars[a0[q]], rn = ars[a1[q]] + arn[a1[q]] - rs;
            const int cs = acs[a0[q]], cn = acs[a1[q]] + acn[a1[q]] - cs;
            mr[q] = rn; mc[q] = cn;
            /* block row: T = (C(rows, other columns))^T, (nc - cn) x rn */
            const int w1 = nc - cn, w2 = nr - rn;
            double *T = (double *)malloc(sizeof(double) * (size_t)(w1 > 0 ? w1 : 1) * rn);
            for (int t = 0; t < rn; t++) {
                int o = 0;
                for (int c = 0; c < nc; c++) {
                    if (c >= cs && c < cs + cn) continue;
                    T[o + (long)t * w1] = C[(rs + t) + (size_t)c * nr];
                    o++;
                }
            }
            double *v = (double *)malloc(sizeof(double) * rn * rn), *sv = (double *)malloc(sizeof(double) * rn);
            ru[q] = side_basis(w1, rn, T, tol, v, sv, rsvd, seed, 0);
            if (ru[q] < 0) info = 1;
            Ub[q] = v;
            free(T); free(sv);
            /* block column: T = C(other rows, columns), (nr - rn) x cn */
            T = (double *)malloc(sizeof(double) * (size_t)(w2 > 0 ? w2 : 1) * cn);
            for (int t = 0; t < cn; t++) {
                int o = 0;
                for (int r = 0; r < nr; r++) {
                    if (r >= rs && r < rs + rn) continue;
                    T[o + (long)t * w2] = C[r + (size_t)(cs + t) * nr];
                    o++;
                }
            }
            v = (double *)malloc(sizeof(double) * cn * cn);
            sv = (double *)malloc(sizeof(double) * cn);
            rw[q] = side_basis(w2, cn, T, tol, v, sv, rsvd, seed, 1);
            if (rw[q] < 0) info = 1;
            Wb[q] = v;
            free(T); free(sv);
            /* generators: U (rn x ru), V = W^T (rw x cn) */
            const int nd = nds[q];
            long off = push(H, &cap, (long)rn * (ru[q] > 0 ? ru[q] : 1));
            for (int j = 0; j < ru[q]; j++)
                for (int i = 0; i < rn; i++) H->h[off + i + (long)j * rn] = Ub[q][i + (long)j * rn];
            H->usz[2 * (nd - 1)] = rn; H->usz[2 * (nd - 1) + 1] = ru[q]; H->pu[nd - 1] = off;
            off = push(H, &cap, (long)(rw[q] > 0 ? rw[q] : 1) * cn);
            for (int j = 0; j < cn; j++)
                for (int i = 0; i < rw[q]; i++) H->h[off + i + (long)j * rw[q]] = Wb[q][j + (long)i * cn];
            H->vsz[2 * (nd - 1)] = rw[q]; H->vsz[2 * (nd - 1) + 1] = cn; H->pv[nd - 1] = off;
        }
        if (!info) {
            /* the new active list and the compressed matrix C' = blockdiag(U)^T C blockdiag(W) */
            int nna = 0, nnr = 0, nnc = 0;
            int *nact = (int *)malloc(sizeof(int) * ltr), *nrs = (int *)malloc(sizeof(int) * ltr),
                *nrn = (int *)malloc(sizeof(int) * ltr), *ncs = (int *)malloc(sizeof(int) * ltr),
                *ncn = (int *)malloc(sizeof(int) * ltr), *src = (int *)malloc(sizeof(int) * ltr);
            for (int t = 0; t < na;) {
                int q = -1;
                for (int z = 0; z < nn; z++) if (a0[z] == t) q = z;
                if (q >= 0) {
                    nact[nna] = nds[q]; nrn[nna] = ru[q]; ncn[nna] = rw[q]; src[nna] = q;
                    t = a1[q] + 1;
                } else {
                    nact[nna] = act[t]; nrn[nna] = arn[t]; ncn[nna] = acn[t]; src[nna] = -1 - t;
                    t++;
                }
                nrs[nna] = nnr; ncs[nna] = nnc; nnr += nrn[nna]; nnc += ncn[nna]; nna++;
            }
            double *Tr = (double *)calloc((size_t)(nnr > 0 ? nnr : 1) * (nc > 0 ? nc : 1), sizeof(double));
            for (int a = 0; a < nna; a++) {  /* rows */
                const int s = src[a];
                for (int c = 0; c < nc; c++)
                    for (int i = 0; i < nrn[a]; i++) {
                        double acc;
                        if (s >= 0) {
                            const int rs = ars[a0[s]];
                            acc = 0.0;
                            for (int t = 0; t < mr[s]; t++) acc += Ub[s][t + (long)i * mr[s]] * C[(rs + t) + (size_t)c * nr];
                        } else {
                            acc = C[(ars[-1 - s] + i) + (size_t)c * nr];
                        }
                        Tr[(nrs[a] + i) + (size_t)c * nnr] = acc;
                    }
            }
            double *Cn = (double *)calloc((size_t)(nnr > 0 ? nnr : 1) * (nnc > 0 ? nnc : 1), sizeof(double));
            for (int b = 0; b < nna; b++) {  /* columns */
                const int s = src[b];
                for (int j = 0; j < ncn[b]; j++)
                    for (int r = 0; r < nnr; r++) {
                        double acc;
                        if (s >= 0) {
                            const int cs = acs[a0[s]];
                            acc = 0.0;
                            for (int t = 0; t < mc[s]; t++) acc += Tr[r + (size_t)(cs + t) * nnr] * Wb[s][t + (long)j * mc[s]];
                        } else {
                            acc = Tr[r + (size_t)(acs[-1 - s] + j) * nnr];
                        }
                        Cn[r + (size_t)(ncs[b] + j) * nnr] = acc;
                    }
            }
            free(Tr); free(C);
            C = Cn; nr = nnr; nc = nnc;
            na = nna;
            memcpy(act, nact, sizeof(int) * nna); memcpy(ars, nrs, sizeof(int) * nna);
            memcpy(arn, nrn, sizeof(int) * nna); memcpy(acs, ncs, sizeof(int) * nna);
            memcpy(acn, ncn, sizeof(int) * nna);
            free(nact); free(nrs); free(nrn); free(ncs); free(ncn); free(src);
            /* B of every sibling pair of this depth: the compressed block between them */
            for (int q = 0; q < nn; q++) {
                const int nd = nds[q], par = A1(H->tr, nd);
                const int sib = (CH1(ch, par) == nd) ? CH2(ch, par) : CH1(ch, par);
                int ia = -1, ib = -1;
                for (int t = 0; t < na; t++) { if (act[t] == nd) ia = t; if (act[t] == sib) ib = t; }
                const int br = arn[ia], bc = acn[ib];
                long off = push(H, &cap, (long)(br > 0 ? br : 1) * (bc > 0 ? bc : 1));
                for (int j = 0; j < bc; j++)
                    for (int i = 0; i < br; i++) H->h[off + i + (long)j * br] = C[(ars[ia] + i) + (size_t)(acs[ib] + j) * nr];
                H->bsz[2 * (nd - 1)] = br; H->bsz[2 * (nd - 1) + 1] = bc; H->pb[nd - 1] = off;
            }
        }
        for (int q = 0; q < nn; q++) { free(Ub[q]); free(Wb[q]); }
        free(Ub); free(Wb); free(ru); free(rw); free(mr); free(mc); free(nds); free(a0); free(a1);
    }
    free(C); free(act); free(ars); free(arn); free(acs); free(acn); free(r0); free(dep);
    return info;
}

void orc_rrlu(int rowcol, double *F, double *D, double *A, double *V, const double *DIFL, const double *DIFR,
              double tol, double *U, double *W, const double *FF, int *PL, int *PU, int M, int N, int *Rk,
              int *PQ, long long *nflops, long long *nswap)
{
    if (rowcol == 0)
        rrluRow(F, D, A, V, DIFL, DIFR, NULL, tol, U, W, FF, PL, PU, M, N, Rk, PQ, nflops, nswap);
    else
        rrluCol(D, F, A, V, DIFL, DIFR, NULL, tol, U, W, FF, PL, PU, M, N, Rk, PQ, nflops, nswap);
}
