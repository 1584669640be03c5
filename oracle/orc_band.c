/*
 * CPU ORACLE -- test infrastructure only (see orc.h).
 *
 * Restatement of HSSPACK's banded divide and conquer MDSBEVD (SRC/mdsbevd.f90:183-498)
 * with its merge MDSBED1 (SRC/mdsbed1.f90:109-312) and the per-update steps of
 * MDSBED2 / MDSBED22 (deflation, SRC/mdsbed2.f90, mdsbed22.f90) and MDSBED3 / MDSBED4
 * (secular equation, Loewner z, vectors, SRC/mdsbed3.f90, mdsbed4.f90):
 *
 *   tree      halve n until every block is <= smlsiz (mdsbevd.f90:303-318); the leaves
 *             must hold at least KD rows, which the caller's smlsiz >= 2 KD - 1 ensures
 *   split     each cut's KD x KD coupling block C = A(ic-KD+1:ic, ic+1:ic+KD) (DSB2FLC,
 *             SFtransf.f:140) is factored C = U S V^T, and DSymSplit (BasicMM.f90:1195-1307)
 *             takes U S U^T from the block before the cut and V S V^T from the one after,
 *             so A = diag(A_i) + sum_k s_k [u_k; v_k][u_k; v_k]^T per cut
 *   leaves    the dense symmetric eigenproblem of each block (mdsbevd.f90:356-377)
 *   merge     Z = [Q1(last KD rows)^T U; Q2(first KD rows)^T V], then KD rank-one
 *             updates with rho = s_k and z = Z(:, k): z / sqrt(2), rho = 2 rho; DLAED2's
 *             deflation with each Givens rotation also applied to the rows of the
 *             remaining columns of Z; the secular roots (DLAED4), Loewner z and the
 *             normalised vectors of the secular problem S; Q(:, 1:K) := Q(:, 1:K) S,
 *             Z(1:K, k+1:KD) := S^T Z(1:K, k+1:KD); INDXQ = DLAMRG(K, N-K, D, 1, -1)
 *
 * Two parts of the reference call LAPACK routines it does not contain (DGESVD for the
 * coupling block, DSYEVD for the leaves); they are replaced here by one-sided Jacobi
 * (Hestenes) and cyclic two-sided Jacobi, both pinned against LAPACK in the tests.
 * The reference's first update groups columns by type (mdsbed2.f90:378-420) to skip the
 * zero blocks of Q; here every update multiplies the dense columns, which changes only
 * the rounding.  When an update deflates completely (K = 0) the reference reorders Q and
 * D but not the remaining columns of Z (mdsbed22.f90:233-248); here Z's rows follow the
 * same permutation, which keeps Z expressed in the basis of Q.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define EPSB 0x1p-53

/* A(i, j) of the symmetric band in upper storage AB(kd+1+i-j, j), 0-based */
static double band_at(const double *ab, int ldab, int kd, int i, int j)
{
    if (i > j) { int t = i; i = j; j = t; }
    if (j - i > kd) return 0.0;
    return ab[(kd + i - j) + (long)j * ldab];
}

/* Cyclic two-sided Jacobi on the symmetric m x m matrix a (lda), round-robin pairs:
 * in each round every pair's rotation comes from the round's starting values, all
 * row rotations are applied, then all column rotations, then the pairs' off-diagonal
 * entries are set to zero.  A pair rotates when |a_pq| > m eps sqrt(|a_pp a_qq|).
 * Eigenvalues ascending in w, eigenvectors in v (ldv).  Returns the sweeps, < 0 when
 * 40 sweeps did not converge. */
int orc_sym_jacobi(int m, double *a, int lda, double *w, double *v, int ldv)
{
    int sweeps = 0, P = m + (m & 1);  /* players, a dummy when m is odd */
    int *pos = (int *)malloc(sizeof(int) * P);
    double *cs = (double *)malloc(sizeof(double) * P), *sn = (double *)malloc(sizeof(double) * P);
    int *pp = (int *)malloc(sizeof(int) * P), *pq = (int *)malloc(sizeof(int) * P);
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) v[i + (long)j * ldv] = (i == j) ? 1.0 : 0.0;
    int conv = 0;
    for (sweeps = 1; sweeps <= 40 && !conv; sweeps++) {
        int rot = 0;
        for (int i = 0; i < P; i++) pos[i] = i;
        for (int r = 0; r < P - 1; r++) {
            /* pairs of this round: (pos[i], pos[P-1-i]) */
            int np = 0;
            for (int i = 0; i < P / 2; i++) {
                int p = pos[i], q = pos[P - 1 - i];
                if (p > q) { int t = p; p = q; q = t; }
                if (q >= m) continue;
                double apq = a[p + (long)q * lda], app = a[p + (long)p * lda], aqq = a[q + (long)q * lda];
                if (apq == 0.0 || !(fabs(apq) > m * EPSB * sqrt(fabs(app * aqq)))) continue;
                double th = (aqq - app) / (2.0 * apq);
                double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                pp[np] = p; pq[np] = q; cs[np] = c; sn[np] = s; np++;
            }
            rot += np;
            for (int e = 0; e < np; e++) {  /* rows p, q */
                int p = pp[e], q = pq[e];
                double c = cs[e], s = sn[e];
                for (int j = 0; j < m; j++) {
                    double x = a[p + (long)j * lda], y = a[q + (long)j * lda];
                    a[p + (long)j * lda] = c * x - s * y;
                    a[q + (long)j * lda] = s * x + c * y;
                }
            }
            for (int e = 0; e < np; e++) {  /* columns p, q of a and v */
                int p = pp[e], q = pq[e];
                double c = cs[e], s = sn[e];
                for (int i = 0; i < m; i++) {
                    double x = a[i + (long)p * lda], y = a[i + (long)q * lda];
                    a[i + (long)p * lda] = c * x - s * y;
                    a[i + (long)q * lda] = s * x + c * y;
                    x = v[i + (long)p * ldv];
                    y = v[i + (long)q * ldv];
                    v[i + (long)p * ldv] = c * x - s * y;
                    v[i + (long)q * ldv] = s * x + c * y;
                }
            }
            for (int e = 0; e < np; e++) {
                a[pp[e] + (long)pq[e] * lda] = 0.0;
                a[pq[e] + (long)pp[e] * lda] = 0.0;
            }
            /* round-robin: player 0 stays, the others rotate one place */
            int last = pos[P - 1];
            for (int i = P - 1; i > 1; i--) pos[i] = pos[i - 1];
            pos[1] = last;
        }
        if (rot == 0) conv = 1;
    }
    /* ascending order, ties by index (selection sort keeps it deterministic) */
    for (int i = 0; i < m; i++) w[i] = a[i + (long)i * lda];
    for (int i = 0; i < m - 1; i++) {
        int k = i;
        for (int j = i + 1; j < m; j++) if (w[j] < w[k]) k = j;
        if (k != i) {
            double t = w[i]; w[i] = w[k]; w[k] = t;
            for (int r = 0; r < m; r++) {
                t = v[r + (long)i * ldv]; v[r + (long)i * ldv] = v[r + (long)k * ldv]; v[r + (long)k * ldv] = t;
            }
        }
    }
    free(pos); free(cs); free(sn); free(pp); free(pq);
    return conv ? sweeps - 1 : -1;
}

/* One-sided Jacobi SVD of the m x m matrix c (ldc): c = u diag(s) v^T, s descending.
 * Pairs in round-robin order; a pair rotates when |c_p . c_q| > m eps |c_p| |c_q|.
 * Zero singular values get u columns of zeros.  Returns the sweeps (< 0: no convergence). */
int orc_jacobi_svd(int m, double *c, int ldc, double *s, double *u, int ldu, double *v, int ldv)
{
    int P = m + (m & 1), conv = 0, sweeps;
    int *pos = (int *)malloc(sizeof(int) * P);
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) v[i + (long)j * ldv] = (i == j) ? 1.0 : 0.0;
    for (sweeps = 1; sweeps <= 60 && !conv; sweeps++) {
        int rot = 0;
        for (int i = 0; i < P; i++) pos[i] = i;
        for (int r = 0; r < P - 1; r++) {
            for (int i = 0; i < P / 2; i++) {
                int p = pos[i], q = pos[P - 1 - i];
                if (p > q) { int t = p; p = q; q = t; }
                if (q >= m) continue;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = 0; k < m; k++) {
                    double x = c[k + (long)p * ldc], y = c[k + (long)q * ldc];
                    al += x * x; be += y * y; ga += x * y;
                }
                if (ga == 0.0 || !(fabs(ga) > m * EPSB * sqrt(al * be))) continue;
                double ze = (be - al) / (2.0 * ga);
                double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                double cc = 1.0 / sqrt(1.0 + t * t), ss = cc * t;
                for (int k = 0; k < m; k++) {
                    double x = c[k + (long)p * ldc], y = c[k + (long)q * ldc];
                    c[k + (long)p * ldc] = cc * x - ss * y;
                    c[k + (long)q * ldc] = ss * x + cc * y;
                    x = v[k + (long)p * ldv];
                    y = v[k + (long)q * ldv];
                    v[k + (long)p * ldv] = cc * x - ss * y;
                    v[k + (long)q * ldv] = ss * x + cc * y;
                }
                rot++;
            }
            int last = pos[P - 1];
            for (int i = P - 1; i > 1; i--) pos[i] = pos[i - 1];
            pos[1] = last;
        }
        if (rot == 0) conv = 1;
    }
    for (int j = 0; j < m; j++) {
        double nr = 0.0;
        for (int k = 0; k < m; k++) nr += c[k + (long)j * ldc] * c[k + (long)j * ldc];
        s[j] = sqrt(nr);
        for (int k = 0; k < m; k++) u[k + (long)j * ldu] = (s[j] > 0.0) ? c[k + (long)j * ldc] / s[j] : 0.0;
    }
    /* descending, as DGESVD returns them */
    for (int i = 0; i < m - 1; i++) {
        int k = i;
        for (int j = i + 1; j < m; j++) if (s[j] > s[k]) k = j;
        if (k != i) {
            double t = s[i]; s[i] = s[k]; s[k] = t;
            for (int r = 0; r < m; r++) {
                t = u[r + (long)i * ldu]; u[r + (long)i * ldu] = u[r + (long)k * ldu]; u[r + (long)k * ldu] = t;
                t = v[r + (long)i * ldv]; v[r + (long)i * ldv] = v[r + (long)k * ldv]; v[r + (long)k * ldv] = t;
            }
        }
    }
    free(pos);
    return conv ? sweeps - 1 : -1;
}

/* one rank-one update of the merge block: diag(D) + rho z z^T in the basis Q (n x n,
 * ldq), the remaining columns Z(:, c0:kd) carried along; indx: the sorted order of D
 * (1-based positions into the physical columns).  On exit D, Q, Z hold the updated
 * eigensystem in physical order (the K new vectors first, then the deflated ones) and
 * indxq the DLAMRG order.  Returns K, or < 0 when a root did not converge. */
static int band_update(int n, double *D, double *Q, int ldq, double *Z, int ldz, int c0, int kd, double rho,
                       const int *indx, int *indxq, int maxit, long long *iters, int *nrot_out)
{
    const int kz = c0;  /* column of Z being applied */
    double *zs = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) zs[i] = Z[i + (long)kz * ldz] / sqrt(2.0);
    rho = fabs(2.0 * rho);
    int imax = 0, jmax = 0;
    for (int i = 1; i < n; i++) {
        if (fabs(zs[i]) > fabs(zs[imax])) imax = i;
        if (fabs(D[i]) > fabs(D[jmax])) jmax = i;
    }
    const double tol = 8.0 * EPSB * fmax(fabs(D[jmax]), fabs(zs[imax]));
    int *indxp = (int *)malloc(sizeof(int) * n);
    double *dlamda = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
    int k = 0, nrot = 0;
    if (rho * fabs(zs[imax]) <= tol) {
        for (int j = 0; j < n; j++) indxp[j] = indx[j];
    } else {
        int k2 = n, pj = -1, j = 0;
        for (; j < n; j++) {
            int nj = indx[j] - 1;
            if (rho * fabs(zs[nj]) <= tol) { indxp[--k2] = nj + 1; }
            else { pj = nj; break; }
        }
        for (j = j + 1; j < n; j++) {
            int nj = indx[j] - 1;
            if (rho * fabs(zs[nj]) <= tol) { indxp[--k2] = nj + 1; continue; }
            double s = zs[pj], c = zs[nj];
            double tau = hypot(c, s), t = D[nj] - D[pj];
            c = c / tau;
            s = -s / tau;
            if (fabs(t * c * s) <= tol) {
                zs[nj] = tau;
                zs[pj] = 0.0;
                for (int r = 0; r < n; r++) {  /* DROT on the columns pj, nj of Q */
                    double x = Q[r + (long)pj * ldq], y = Q[r + (long)nj * ldq];
                    Q[r + (long)pj * ldq] = c * x + s * y;
                    Q[r + (long)nj * ldq] = c * y - s * x;
                }
                for (int cz = kz + 1; cz < kd; cz++) {  /* and on the rows pj, nj of Z's later columns */
                    double x = Z[pj + (long)cz * ldz], y = Z[nj + (long)cz * ldz];
                    Z[pj + (long)cz * ldz] = c * x + s * y;
                    Z[nj + (long)cz * ldz] = c * y - s * x;
                }
                nrot++;
                t = D[pj] * c * c + D[nj] * s * s;
                D[nj] = D[pj] * s * s + D[nj] * c * c;
                D[pj] = t;
                k2--;
                int i = 1;
                for (;;) {
                    if (k2 + i < n) {
                        if (D[pj] < D[indxp[k2 + i] - 1]) {
                            indxp[k2 + i - 1] = indxp[k2 + i];
                            indxp[k2 + i] = pj + 1;
                            i++;
                            continue;
                        }
                        indxp[k2 + i - 1] = pj + 1;
                    } else {
                        indxp[k2 + i - 1] = pj + 1;
                    }
                    break;
                }
                pj = nj;
            } else {
                dlamda[k] = D[pj];
                w[k] = zs[pj];
                indxp[k] = pj + 1;
                k++;
                pj = nj;
            }
        }
        if (pj >= 0) {
            dlamda[k] = D[pj];
            w[k] = zs[pj];
            indxp[k] = pj + 1;
            k++;
        }
    }
    *nrot_out = nrot;
    /* physical reorder by INDXP: the K kept columns first, then the deflated ones */
    double *Q2 = (double *)malloc(sizeof(double) * (long)n * n), *Z2 = (double *)malloc(sizeof(double) * (long)n * kd);
    double *D2 = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) {
        int js = indxp[j] - 1;
        memcpy(Q2 + (long)j * n, Q + (long)js * ldq, sizeof(double) * n);
        for (int cz = 0; cz < kd; cz++) Z2[j + (long)cz * n] = Z[js + (long)cz * ldz];
        D2[j] = D[js];
    }
    for (int j = 0; j < n; j++) memcpy(Q + (long)j * ldq, Q2 + (long)j * n, sizeof(double) * n);
    for (int cz = 0; cz < kd; cz++)
        for (int j = 0; j < n; j++) Z[j + (long)cz * ldz] = Z2[j + (long)cz * n];
    memcpy(D, D2, sizeof(double) * n);
    int info = 0;
    if (k > 0) {
        /* secular roots and the vectors of the secular problem (DLAED3 with the Loewner z) */
        double *del = (double *)malloc(sizeof(double) * (long)k * k), *lam = (double *)malloc(sizeof(double) * k);
        double *S = (double *)malloc(sizeof(double) * (long)k * k);
        for (int j = 1; j <= k; j++) {
            int org, nit;
            double tau;
            if (orc_laed4(k, j, dlamda, w, del + (long)(j - 1) * k, rho, &lam[j - 1], &org, &tau, &nit, maxit)) info = 1;
            if (iters) *iters += nit;
        }
        if (k == 1) {
            S[0] = 1.0;
        } else if (k == 2) {
            /* DLAED4 returns the normalised eigenvector in DELTA for K = 2 (DLAED5) */
            for (int j = 0; j < 2; j++) {
                S[0 + (long)j * 2] = del[0 + (long)j * 2];
                S[1 + (long)j * 2] = del[1 + (long)j * 2];
            }
        } else {
            double *zh = (double *)malloc(sizeof(double) * k);
            for (int i = 0; i < k; i++) {
                double wi = del[i + (long)i * k];
                for (int j = 0; j < k; j++)
                    if (j != i) wi = wi * (del[i + (long)j * k] / (dlamda[i] - dlamda[j]));
                zh[i] = copysign(sqrt(-wi), w[i]);
            }
            for (int j = 0; j < k; j++) {
                double nr = 0.0;
                for (int i = 0; i < k; i++) {
                    double x = zh[i] / del[i + (long)j * k];
                    S[i + (long)j * k] = x;
                    nr += x * x;
                }
                nr = sqrt(nr);
                for (int i = 0; i < k; i++) S[i + (long)j * k] /= nr;
            }
            free(zh);
        }
        /* Q(:, 1:K) := Q(:, 1:K) S; Z(1:K, later) := S^T Z(1:K, later) */
        double *T = (double *)malloc(sizeof(double) * (long)n * k);
#pragma omp parallel for schedule(static)
        for (int j = 0; j < k; j++) {
            double *tj = T + (long)j * n;
            for (int r = 0; r < n; r++) tj[r] = 0.0;
            for (int i = 0; i < k; i++) {
                const double sij = S[i + (long)j * k], *qi = Q + (long)i * ldq;
                for (int r = 0; r < n; r++) tj[r] += qi[r] * sij;
            }
        }
        for (int j = 0; j < k; j++) memcpy(Q + (long)j * ldq, T + (long)j * n, sizeof(double) * n);
        for (int cz = kz + 1; cz < kd; cz++) {
            for (int j = 0; j < k; j++) {
                double acc = 0.0;
                for (int i = 0; i < k; i++) acc += S[i + (long)j * k] * Z[i + (long)cz * ldz];
                T[j] = acc;
            }
            for (int j = 0; j < k; j++) Z[j + (long)cz * ldz] = T[j];
        }
        for (int j = 0; j < k; j++) D[j] = lam[j];
        free(T); free(del); free(lam); free(S);
        orc_dlamrg(k, n - k, D, 1, -1, indxq);
    } else {
        for (int i = 0; i < n; i++) indxq[i] = i + 1;
    }
    free(zs); free(indxp); free(dlamda); free(w); free(Q2); free(Z2); free(D2);
    return info ? -1 : k;
}

int orc_sbevd(int n, int kd, const double *ab, int ldab, double *d, double *q, int ldq, int smlsiz,
              orc_band_trace *tr)
{
    if (n < 0) return -3;
    if (kd < 0) return -4;
    if (ldab < kd + 1) return -6;
    if (ldq < (n > 1 ? n : 1)) return -9;
    if (n == 0) return 0;
    if (kd > 0 && smlsiz < 2 * kd - 1) return -20;
    /* tree: halve until every block is <= smlsiz (mdsbevd.f90:303-318) */
    int nb = 1;
    int *ends = (int *)malloc(sizeof(int) * 2 * (n + 1));
    ends[0] = n;
    int tlvls = 0;
    while (ends[nb - 1] > smlsiz && kd > 0) {
        for (int j = nb; j >= 1; j--) {
            ends[2 * j - 1] = (ends[j - 1] + 1) / 2;
            ends[2 * j - 2] = ends[j - 1] / 2;
        }
        tlvls++;
        nb *= 2;
    }
    for (int j = 1; j < nb; j++) ends[j] += ends[j - 1];
    int *lend = (int *)malloc(sizeof(int) * nb);  /* the leaf ends, for the cuts' factors */
    memcpy(lend, ends, sizeof(int) * nb);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) q[i + (long)j * ldq] = 0.0;
    /* split: SVD of every cut's coupling block (DSB2FLC 'L'), DSymSplit's corrections */
    const int K2 = kd * kd;
    double *us = (double *)calloc((size_t)(nb > 1 ? nb - 1 : 1) * (2 * K2 + kd), sizeof(double));
    double *C = (double *)malloc(sizeof(double) * (K2 > 0 ? K2 : 1));
    for (int i = 0; i + 1 < nb; i++) {
        int ic = ends[i];
        double *U = us + (long)i * (2 * K2 + kd), *V = U + K2, *S = V + K2;
        for (int c = 0; c < kd; c++)
            for (int r = 0; r < kd; r++) C[r + c * kd] = band_at(ab, ldab, kd, ic - kd + r, ic + c);
        if (orc_jacobi_svd(kd, C, kd, S, U, kd, V, kd) < 0) { free(ends); free(lend); free(us); free(C); return 1; }
        if (tr && tr->sv) memcpy(tr->sv + (long)i * kd, S, sizeof(double) * kd);
    }
    /* leaves: dense blocks minus the corrections of their cuts, Jacobi */
    int *indxq = (int *)malloc(sizeof(int) * n);
    int st = 0;
    for (int i = 0; i < nb; i++) {
        int m = ends[i] - st;
        double *A = (double *)malloc(sizeof(double) * m * m), *Vv = (double *)malloc(sizeof(double) * m * m);
        for (int c = 0; c < m; c++)
            for (int r = 0; r < m; r++) A[r + c * m] = band_at(ab, ldab, kd, st + r, st + c);
        if (i + 1 < nb) {  /* bottom-right: U S U^T of the cut after the block */
            const double *U = us + (long)i * (2 * K2 + kd), *S = U + 2 * K2;
            for (int c = 0; c < kd; c++)
                for (int r = 0; r < kd; r++) {
                    double acc = 0.0;
                    for (int t = 0; t < kd; t++) acc += U[r + t * kd] * S[t] * U[c + t * kd];
                    A[(m - kd + r) + (m - kd + c) * m] -= acc;
                }
        }
        if (i > 0) {  /* top-left: V S V^T of the cut before it */
            const double *V = us + (long)(i - 1) * (2 * K2 + kd) + K2, *S = V + K2;
            for (int c = 0; c < kd; c++)
                for (int r = 0; r < kd; r++) {
                    double acc = 0.0;
                    for (int t = 0; t < kd; t++) acc += V[r + t * kd] * S[t] * V[c + t * kd];
                    A[r + c * m] -= acc;
                }
        }
        int sw = orc_sym_jacobi(m, A, m, d + st, Vv, m);
        if (sw < 0) { free(A); free(Vv); free(ends); free(lend); free(us); free(C); free(indxq); return 1; }
        for (int c = 0; c < m; c++) memcpy(q + st + (long)(st + c) * ldq, Vv + (long)c * m, sizeof(double) * m);
        for (int j = 0; j < m; j++) indxq[st + j] = j + 1;
        free(A); free(Vv);
        st = ends[i];
    }
    /* merges, level by level (mdsbevd.f90:383-438), KD rank-one updates each (MDSBED1) */
    int info = 0, curlvl = 1, nmerge = 0, nupd = 0;
    int subpbs = nb;
    double *Z = NULL;
    while (subpbs > 1) {
        for (int i = 0; i <= subpbs - 2; i += 2) {
            int submat, matsiz, msd2;
            if (i == 0) { submat = 0; matsiz = ends[1]; msd2 = ends[0]; }
            else { submat = ends[i - 1]; matsiz = ends[i + 1] - ends[i - 1]; msd2 = matsiz / 2; }
            const int N = matsiz, N1 = msd2, cut = submat + N1;
            double *Qb = q + submat + (long)submat * ldq;
            Z = (double *)realloc(Z, sizeof(double) * (long)N * (kd > 0 ? kd : 1));
            const double *U = NULL, *V = NULL, *S = NULL;
            {
                /* the merge's cut is the end of its left half's last leaf */
                int j = 0;
                while (j + 1 < nb && lend[j] != cut) j++;
                U = us + (long)j * (2 * K2 + kd); V = U + K2; S = V + K2;
            }
            for (int c = 0; c < kd; c++) {
                for (int r = 0; r < N1; r++) {
                    double acc = 0.0;
                    for (int t = 0; t < kd; t++) acc += Qb[(N1 - kd + t) + (long)r * ldq] * U[t + c * kd];
                    Z[r + (long)c * N] = acc;
                }
                for (int r = N1; r < N; r++) {
                    double acc = 0.0;
                    for (int t = 0; t < kd; t++) acc += Qb[(N1 + t) + (long)r * ldq] * V[t + c * kd];
                    Z[r + (long)c * N] = acc;
                }
            }
            int *iq = indxq + submat;
            int *indx = (int *)malloc(sizeof(int) * N), *indxc = (int *)malloc(sizeof(int) * N);
            double *dl = (double *)malloc(sizeof(double) * N);
            for (int upd = 0; upd < kd; upd++) {
                if (upd == 0) {  /* the two sorted halves merged (mdsbed2.f90:245-256) */
                    for (int r = N1; r < N; r++) iq[r] += N1;
                    for (int r = 0; r < N; r++) dl[r] = d[submat + iq[r] - 1];
                    orc_dlamrg(N1, N - N1, dl, 1, 1, indxc);
                    for (int r = 0; r < N; r++) indx[r] = iq[indxc[r] - 1];
                } else {
                    memcpy(indx, iq, sizeof(int) * N);
                }
                long long its = 0;
                int nrot = 0;
                int kk = band_update(N, d + submat, Qb, ldq, Z, N, upd, kd, S[upd], indx, iq, 100, &its, &nrot);
                if (kk < 0) { info = submat * (n + 1) + submat + matsiz - 1; break; }
                if (tr) {
                    if (tr->nupd < tr->cap) {
                        tr->k[tr->nupd] = kk;
                        tr->n[tr->nupd] = N;
                        tr->nrot[tr->nupd] = nrot;
                    }
                    tr->nupd++;
                    tr->iters += its;
                }
                nupd++;
            }
            free(indx); free(indxc); free(dl);
            if (info) break;
            nmerge++;
            ends[i / 2] = ends[i + 1];
        }
        if (info) break;
        subpbs /= 2;
        curlvl++;
        (void)curlvl;
    }
    free(Z);
    if (!info) {
        /* final reordering by INDXQ (mdlaed0.f90:355-362 in MDLAED0's form) */
        double *W = (double *)malloc(sizeof(double) * n), *Q2 = (double *)malloc(sizeof(double) * (long)n * n);
        for (int i = 0; i < n; i++) {
            int j = indxq[i] - 1;
            W[i] = d[j];
            memcpy(Q2 + (long)i * n, q + (long)j * ldq, sizeof(double) * n);
        }
        memcpy(d, W, sizeof(double) * n);
        for (int j = 0; j < n; j++) memcpy(q + (long)j * ldq, Q2 + (long)j * n, sizeof(double) * n);
        free(W); free(Q2);
    }
    if (tr) tr->nmerge = nmerge;
    free(ends); free(lend); free(us); free(C); free(indxq);
    return info;
}
