// Banded divide and conquer on gfx950: HSSPACK's MDSBEVD (mdsbevd.f90:183-498) with
// its merge MDSBED1 (mdsbed1.f90:109-312).
//
//   split    one workgroup per cut: the KD x KD coupling block in LDS, its one-sided
//            Jacobi SVD (DGESVD in the reference), singular values descending
//   leaves   one workgroup per leaf (KD <= m <= 128 rows): the block minus DSymSplit's
//            corrections (BasicMM.f90:1195-1307) in LDS, cyclic two-sided Jacobi
//            (DSYEVD in the reference), the vectors straight into Q's diagonal block
//   merges   level by level; every merge of a level takes its KD rank-one updates
//            together, one launch per step over all merges of the level:
//              deflation scan (MDSBED2 / MDSBED22) -- one workgroup per merge
//              the Givens rotations on Q's columns and on the rows of Z's later columns
//              the gather of the kept columns into X (the deflated ones into place)
//              secular roots, Loewner z, vectors S of the secular problem -- the
//              tridiagonal path's k_secular, k_loewner, k_sbuild on this update's arrays
//              Q(:, 1:K) := X S and Z(1:K, later) := S^T Z(1:K, later): batched FP64
//              MFMA GEMMs (k_bgemm) or, for K >= hss_min_k, the HSS product of the
//              Cauchy-like S (MDSBED5 -> DHSSEVC, mdsbed5.f90:277)
//              INDXQ := DLAMRG(K, N-K, D, 1, -1)
//            The tables of every launch after the scan are written by the device from
//            the K it found, so a dense update never waits for the host.
//
// Q, Z and D are ping-pong buffers (an update reads one and writes the other), so the
// deflation permutation is one pass of gathers, not a copy back.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#include "hsdc_internal.h"
#include "hsdc_hss.h"
#include "../../include/hsdc.h"

#define BWG 256
#define BKD_MAX 64    // largest KD: the coupling block and its right vectors stay in LDS
#define BLEAF_MAX 128 // largest leaf: its m x m block stays in LDS
#define DCH 2048      // sorted entries staged in LDS per chunk of the deflation scan
#define SBW 1024      // words of the deflation pre-pass's pair flags (32768 survivors)

__global__ void k_secular(const Merge *, const int *, int, const MergeOut *, Work, int, int *, const SecCtl *,
                          unsigned int *);
__global__ void k_loewner(const Merge *, const int *, int, const MergeOut *, Work);
__global__ void k_alpha(const Merge *, const int *, int, const MergeOut *, Work);
__global__ void k_sbuild(const Merge *, const int *, int, const MergeOut *, const long *, Work, double *);
__global__ void k_bgemm(const BG *);
template <int BM, int BN, int BK, int DB, int TA, int TB>
__global__ void k_bgemm_w(const BG *, int);
__global__ void k_fill_u64(unsigned long long *, long, unsigned long long);

namespace {

// A(i, j) of the symmetric band, 0-based: upper storage AB(kd+1+i-j, j) for i <= j,
// lower storage AB(1+i-j, j) for i >= j
__device__ __forceinline__ double band_at(const double *ab, int ldab, int kd, int lower, int i, int j)
{
    if (lower) {
        if (i < j) { int t = i; i = j; j = t; }
        if (i - j > kd) return 0.0;
        return ab[(i - j) + (long)j * ldab];
    }
    if (i > j) { int t = i; i = j; j = t; }
    if (j - i > kd) return 0.0;
    return ab[(kd + i - j) + (long)j * ldab];
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// round-robin pairing of P players (P even): in round r, slot i meets slot P-1-i; player
// 0 stays in slot 0 and the others move one slot on per round
__device__ __forceinline__ int rr_player(int slot, int r, int P)
{
    if (slot == 0) return 0;
    return 1 + (slot - 1 + (P - 1) - (r % (P - 1))) % (P - 1);
}

__device__ __forceinline__ int seg_of(const int *start, int nm, int b)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// DLAMRG(n1, n2, a, s1, s2, idx) by the whole workgroup: list 1 is a(0:n1) read forwards
// (s1 = 1) or backwards (s1 = -1), list 2 is a(n1:n1+n2) likewise; idx gets the 1-based
// positions in a of the merged ascending order, ties to list 1.  When both lists are
// ascending each element's place comes from a binary search in the other list; when
// either is not (rounding can leave a deflated tail out of order), thread 0 runs the
// sequential merge, which is what DLAMRG does in every case.
template <int NT>
__device__ void blk_lamrg(int n1, int n2, const double *a, int s1, int s2, int *idx, int *flag)
{
    const int tid = threadIdx.x;
    auto p1 = [&](int i) { return s1 > 0 ? i : n1 - 1 - i; };
    auto p2 = [&](int j) { return s2 > 0 ? n1 + j : n1 + n2 - 1 - j; };
    if (tid == 0) *flag = 0;
    __syncthreads();
    int bad = 0;
    for (int i = tid; i + 1 < n1; i += NT) bad |= !(a[p1(i)] <= a[p1(i + 1)]);
    for (int j = tid; j + 1 < n2; j += NT) bad |= !(a[p2(j)] <= a[p2(j + 1)]);
    if (bad) atomicOr(flag, 1);
    __syncthreads();
    if (*flag == 0) {
        for (int i = tid; i < n1; i += NT) {
            const double v = a[p1(i)];
            int lo = 0, hi = n2;  // list-2 entries < v
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a[p2(mid)] < v) lo = mid + 1; else hi = mid;
            }
            idx[i + lo] = p1(i) + 1;
        }
        for (int j = tid; j < n2; j += NT) {
            const double v = a[p2(j)];
            int lo = 0, hi = n1;  // list-1 entries <= v
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a[p1(mid)] <= v) lo = mid + 1; else hi = mid;
            }
            idx[j + lo] = p2(j) + 1;
        }
    } else if (tid == 0) {
        int i1 = 0, i2 = 0, q = 0;
        while (i1 < n1 && i2 < n2) {
            if (a[p1(i1)] <= a[p2(i2)]) { idx[q++] = p1(i1) + 1; i1++; }
            else { idx[q++] = p2(i2) + 1; i2++; }
        }
        while (i1 < n1) { idx[q++] = p1(i1) + 1; i1++; }
        while (i2 < n2) { idx[q++] = p2(i2) + 1; i2++; }
    }
    __syncthreads();
}

}  // namespace

// One merge of a level: block [s, s+n), cut point n1, cut index into the factors.
struct SbMerge {
    int s, n, n1, cut;
};

// Results of one merge's deflation scan.
struct SbUpd {
    int k, nrot;
};

// Device tables of one update, rewritten by k_sb_tab from the K of each merge.
struct SbTab {
    int *secblk;  // nm + 1: first secular block of each merge (SEC_WG / sec_team(K) roots a block)
    int *chunk;   // nm + 1: first 256-wide chunk of each merge (Loewner, S)
    BG *bq, *bz;  // nm each: Q(:, 1:K) := X S and Z(1:K, later) := S^T Xz(1:K, later)
};

// ---------------------------------------------------------------------------
// Split: one workgroup per cut.  C(r, c) = A(ic - kd + r, ic + c) (DSB2FLC 'L',
// SFtransf.f:140), one-sided Jacobi C V = U S with a wavefront per column pair (lanes
// over the kd rows); a pair rotates when |c_p . c_q| > kd eps |c_p| |c_q|; singular values
// sorted descending (ties by index).  U, V, S of cut i at cuts + i (2 kd^2 + kd).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BWG) void k_sb_split(const double *ab, int ldab, int kd, int lower, const int *cut_ic,
                                                  double *cuts, int *info)
{
    __shared__ double C[BKD_MAX * BKD_MAX], V[BKD_MAX * BKD_MAX], sv[BKD_MAX];
    __shared__ int perm[BKD_MAX], rot;
    const int ic = cut_ic[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int e = tid; e < kd * kd; e += BWG) {
        const int r = e % kd, c = e / kd;
        C[e] = band_at(ab, ldab, kd, lower, ic - kd + r, ic + c);
        V[e] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int P = kd + (kd & 1);
    bool conv = false;
    for (int sweep = 0; sweep < 60 && !conv; sweep++) {
        if (tid == 0) rot = 0;
        __syncthreads();
        for (int r = 0; r < P - 1; r++) {
            for (int pr = wv; pr < P / 2; pr += BWG / 64) {
                int p = rr_player(pr, r, P), q = rr_player(P - 1 - pr, r, P);
                if (p > q) { int t = p; p = q; q = t; }
                if (q >= kd) continue;  // uniform over the wavefront
                const double x = lane < kd ? C[lane + p * kd] : 0.0, y = lane < kd ? C[lane + q * kd] : 0.0;
                const double al = wave_sum(x * x), be = wave_sum(y * y), ga = wave_sum(x * y);
                if (ga == 0.0 || !(fabs(ga) > kd * HSDC_EPS * sqrt(al * be))) continue;
                const double ze = (be - al) / (2.0 * ga);
                const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                const double cc = 1.0 / sqrt(1.0 + t * t), ss = cc * t;
                if (lane < kd) {
                    C[lane + p * kd] = cc * x - ss * y;
                    C[lane + q * kd] = ss * x + cc * y;
                    const double vx = V[lane + p * kd], vy = V[lane + q * kd];
                    V[lane + p * kd] = cc * vx - ss * vy;
                    V[lane + q * kd] = ss * vx + cc * vy;
                }
                if (lane == 0) atomicAdd(&rot, 1);
            }
            __syncthreads();
        }
        conv = (rot == 0);
        __syncthreads();
    }
    if (!conv && tid == 0) atomicOr(info, 1);
    for (int j = wv; j < kd; j += BWG / 64) {
        const double x = lane < kd ? C[lane + j * kd] : 0.0;
        const double nr = sqrt(wave_sum(x * x));
        if (lane == 0) sv[j] = nr;
    }
    __syncthreads();
    if (tid == 0) {
        for (int j = 0; j < kd; j++) perm[j] = j;
        for (int i = 0; i + 1 < kd; i++) {
            int k = i;
            for (int j = i + 1; j < kd; j++) if (sv[perm[j]] > sv[perm[k]]) k = j;
            const int t = perm[i]; perm[i] = perm[k]; perm[k] = t;
        }
    }
    __syncthreads();
    double *U = cuts + (long)blockIdx.x * (2 * kd * kd + kd), *Vo = U + kd * kd, *S = Vo + kd * kd;
    for (int e = tid; e < kd * kd; e += BWG) {
        const int r = e % kd, src = perm[e / kd];
        const double s = sv[src];
        U[e] = s > 0.0 ? C[r + src * kd] / s : 0.0;
        Vo[e] = V[r + src * kd];
    }
    for (int j = tid; j < kd; j += BWG) S[j] = sv[perm[j]];
}

// ---------------------------------------------------------------------------
// Leaves: one workgroup per leaf [st, st + m): the band's block minus V S V^T of the cut
// before it (top-left kd x kd) and U S U^T of the cut after it (bottom-right), in LDS.
// Cyclic two-sided Jacobi: each round's rotations come from the round's starting values
// (|a_pq| > m eps sqrt|a_pp a_qq|), then all row rotations, all column rotations (and
// the vectors), and the pairs' off-diagonal entries set to zero.  Eigenvalues ascending
// into D (ties by index), vectors into Q's diagonal block, INDXQ = 1..m.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BWG) void k_sb_leaf(const double *ab, int ldab, int kd, int lower, const int *leaf_st,
                                                 const int *leaf_m, int nleaf, const double *cuts, double *Q, long ldq,
                                                 double *D, int *indxq, int *info)
{
    extern __shared__ double A[];  // m x m
    __shared__ double rc[BLEAF_MAX / 2], rs[BLEAF_MAX / 2], dg[BLEAF_MAX];
    __shared__ int rp[BLEAF_MAX / 2], rq[BLEAF_MAX / 2], perm[BLEAF_MAX], np, rot;
    const int i = blockIdx.x, st = leaf_st[i], m = leaf_m[i], tid = threadIdx.x;
    const long cs = 2L * kd * kd + kd;
    double *Qb = Q + st + (long)st * ldq;
    for (int e = tid; e < m * m; e += BWG) {
        const int r = e % m, c = e / m;
        double a = band_at(ab, ldab, kd, lower, st + r, st + c);
        if (i > 0 && r < kd && c < kd) {  // V S V^T of the cut before the leaf
            const double *V = cuts + (long)(i - 1) * cs + kd * kd, *S = V + kd * kd;
            double acc = 0.0;
            for (int t = 0; t < kd; t++) acc += V[r + t * kd] * S[t] * V[c + t * kd];
            a -= acc;
        }
        if (i + 1 < nleaf && r >= m - kd && c >= m - kd) {  // U S U^T of the cut after it
            const double *U = cuts + (long)i * cs, *S = U + 2 * kd * kd;
            const int rr = r - (m - kd), cc = c - (m - kd);
            double acc = 0.0;
            for (int t = 0; t < kd; t++) acc += U[rr + t * kd] * S[t] * U[cc + t * kd];
            a -= acc;
        }
        A[e] = a;
        Qb[r + (long)c * ldq] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int P = m + (m & 1);
    bool conv = false;
    for (int sweep = 0; sweep < 40 && !conv; sweep++) {
        if (tid == 0) rot = 0;
        __syncthreads();
        for (int r = 0; r < P - 1; r++) {
            if (tid == 0) np = 0;
            __syncthreads();
            if (tid < P / 2) {
                int p = rr_player(tid, r, P), q = rr_player(P - 1 - tid, r, P);
                if (p > q) { int t = p; p = q; q = t; }
                if (q < m) {
                    const double apq = A[p + q * m], app = A[p + p * m], aqq = A[q + q * m];
                    if (apq != 0.0 && fabs(apq) > m * HSDC_EPS * sqrt(fabs(app * aqq))) {
                        const double th = (aqq - app) / (2.0 * apq);
                        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
                        const double c = 1.0 / sqrt(1.0 + t * t);
                        const int slot = atomicAdd(&np, 1);
                        rp[slot] = p; rq[slot] = q; rc[slot] = c; rs[slot] = t * c;
                    }
                }
            }
            __syncthreads();
            const int npp = np;
            if (npp > 0) {
                for (int e = tid; e < npp * m; e += BWG) {  // rows p, q
                    const int k = e / m, j = e % m, p = rp[k], q = rq[k];
                    const double c = rc[k], s = rs[k], x = A[p + j * m], y = A[q + j * m];
                    A[p + j * m] = c * x - s * y;
                    A[q + j * m] = s * x + c * y;
                }
                __syncthreads();
                for (int e = tid; e < npp * m; e += BWG) {  // columns p, q of A and of the vectors
                    const int k = e / m, ii = e % m, p = rp[k], q = rq[k];
                    const double c = rc[k], s = rs[k];
                    double x = A[ii + p * m], y = A[ii + q * m];
                    A[ii + p * m] = c * x - s * y;
                    A[ii + q * m] = s * x + c * y;
                    double *qp = Qb + ii + (long)p * ldq, *qq = Qb + ii + (long)q * ldq;
                    x = *qp;
                    y = *qq;
                    *qp = c * x - s * y;
                    *qq = s * x + c * y;
                }
                __syncthreads();
                for (int k = tid; k < npp; k += BWG) {
                    A[rp[k] + rq[k] * m] = 0.0;
                    A[rq[k] + rp[k] * m] = 0.0;
                }
                if (tid == 0) rot += npp;
            }
            __syncthreads();
        }
        conv = (rot == 0);
        __syncthreads();
    }
    if (!conv && tid == 0) atomicOr(info, 1);
    for (int j = tid; j < m; j += BWG) dg[j] = A[j + j * m];
    __syncthreads();
    if (tid == 0) {
        for (int j = 0; j < m; j++) perm[j] = j;
        for (int a = 0; a + 1 < m; a++) {
            int k = a;
            for (int j = a + 1; j < m; j++) if (dg[perm[j]] < dg[perm[k]]) k = j;
            const int t = perm[a]; perm[a] = perm[k]; perm[k] = t;
        }
    }
    __syncthreads();
    for (int e = tid; e < m * m; e += BWG) A[e] = Qb[(e % m) + (long)(e / m) * ldq];
    __syncthreads();
    for (int e = tid; e < m * m; e += BWG) {
        const int r = e % m, c = e / m;
        Qb[r + (long)c * ldq] = A[r + perm[c] * m];
    }
    for (int j = tid; j < m; j += BWG) {
        D[st + j] = dg[perm[j]];
        indxq[st + j] = j + 1;
    }
}

// Start of a merge level: the off-diagonal quadrants of each merge's block in Q set to
// zero (one workgroup per column of a merge), so its columns are whole vectors of the
// merged block.
__global__ __launch_bounds__(BWG) void k_sb_zero_off(const SbMerge *mg, const int *colchunk, int nm, double *Q,
                                                     long ldq)
{
    const int m = seg_of(colchunk, nm, blockIdx.x);
    const SbMerge M = mg[m];
    const int c = blockIdx.x - colchunk[m];
    const int r0 = c < M.n1 ? M.n1 : 0, r1 = c < M.n1 ? M.n : M.n1;
    double *qc = Q + M.s + (long)(M.s + c) * ldq;
    for (int r = r0 + threadIdx.x; r < r1; r += BWG) qc[r] = 0.0;
}

// Z of each merge: Z(r, c) = sum_t Q(N1-kd+t, r) U(t, c) for r < N1 (the last kd rows of
// the left block's vectors) and sum_t Q(N1+t, r) V(t, c) otherwise (mdsbed1.f90:185-215).
// A workgroup covers 256 rows of a merge, each wavefront 64 of them in turn with one lane
// per column c.  INDXC = 1..N for the S kernel.
__global__ __launch_bounds__(BWG) void k_sb_zmat(const SbMerge *mg, const int *rowchunk, int nm, const double *Q,
                                                 long ldq, const double *cuts, int kd, double *Z, long ldz, int *indxc)
{
    const int m = seg_of(rowchunk, nm, blockIdx.x);
    const SbMerge M = mg[m];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = (blockIdx.x - rowchunk[m]) * BWG + wv * 64;
    const long cs = 2L * kd * kd + kd;
    const double *U = cuts + (long)M.cut * cs, *V = U + kd * kd;
    for (int rr = 0; rr < 64; rr++) {
        const int r = r0 + rr;
        if (r >= M.n) break;
        const bool top = r < M.n1;
        const double *qc = Q + (M.s + (top ? M.n1 - kd : M.n1)) + (long)(M.s + r) * ldq;
        const double *F = top ? U : V;
        if (lane < kd) {
            double acc = 0.0;
            for (int t = 0; t < kd; t++) acc += qc[t] * F[t + lane * kd];
            Z[(M.s + r) + (long)lane * ldz] = acc;
        }
        if (lane == 0) indxc[M.s + r] = r + 1;  // the secular vectors' rows are X's columns in order
    }
}

// ---------------------------------------------------------------------------
// Deflation of update `up` (MDSBED2 for up = 0, MDSBED22 after), one workgroup per merge.
//   z = Z(:, up) / sqrt(2), rho = |2 s_up|, tol = 8 eps max(max|D|, max|z|)
//   up = 0: INDXQ of the right half shifted by N1, the halves merged (DLAMRG 1, 1)
//   the sorted z and D staged in LDS chunk by chunk for thread 0's scan in the
//   reference's order: rho |z| <= tol deflates; a close pair (|t c s| <= tol) is rotated
//   (c, s = z_nj / tau, -z_pj / tau; tau = DLAPY2) and the rotated-out value inserted in
//   order among the deflated; otherwise the pending entry is kept (DLAMDA, W, INDXP).
// The rotation log (pj, nj, c, s) goes to k_sb_rotate; the deflated values to Dn[K:N)
// in INDXP order; K and the secular scalars of the merge to upd and mo.
// ---------------------------------------------------------------------------
struct SbDefl {
    const double *Z;
    long ldz;
    double *D, *Dn;
    int *indxq, *indx, *indxp;
    double *dlamda, *w, *srtd, *srtz;
    int *rpj, *rnj;
    double *rcs, *rsn;
    const double *cuts;
    int kd, up;
    SbUpd *upd;
    MergeOut *mo;
    Merge *mg2;  // the secular kernels' view of the merges (rho written here)
};

template <int NT>
__global__ __launch_bounds__(NT) void k_sb_deflate(const SbMerge *mg, SbDefl A)
{
    __shared__ double zc[DCH], dc[DCH];
    __shared__ int nc[DCH];
    __shared__ double red[2][NT / 64];
    __shared__ int flag, s_k, s_nrot, s_cnt, s_rot, pre[NT];
    // bit j: the pre-pass's rotation test of survivors j-1 and j passed (survivor
    // numbers below 32 x SBW)
    __shared__ unsigned int pfl[SBW];
    const SbMerge M = mg[blockIdx.x];
    const int s = M.s, n = M.n, n1 = M.n1, tid = threadIdx.x, up = A.up;
    const double rho0 = A.cuts[(long)M.cut * (2L * A.kd * A.kd + A.kd) + 2 * A.kd * A.kd + up];
    const double *zcol = A.Z + s + (long)up * A.ldz;
    double *d = A.D + s;
    int *iq = A.indxq + s, *ix = A.indx + s, *ip = A.indxp + s;
    double az = 0.0, ad = 0.0;
    for (int j = tid; j < n; j += NT) {
        az = fmax(az, fabs(zcol[j] / sqrt(2.0)));
        ad = fmax(ad, fabs(d[j]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        az = fmax(az, __shfl_xor(az, off));
        ad = fmax(ad, __shfl_xor(ad, off));
    }
    if ((tid & 63) == 0) { red[0][tid >> 6] = az; red[1][tid >> 6] = ad; }
    if (up == 0)
        for (int j = n1 + tid; j < n; j += NT) iq[j] += n1;
    __syncthreads();
    az = 0.0; ad = 0.0;
    for (int q = 0; q < NT / 64; q++) { az = fmax(az, red[0][q]); ad = fmax(ad, red[1][q]); }
    const double rho = fabs(2.0 * rho0);
    const double tol = 8.0 * HSDC_EPS * fmax(ad, az);
    // the sorted order
    if (up == 0) {
        double *dl = A.srtd + s;
        for (int r = tid; r < n; r += NT) dl[r] = d[iq[r] - 1];
        __syncthreads();
        blk_lamrg<NT>(n1, n - n1, dl, 1, 1, ip, &flag);  // INDXC, held in INDXP for now
        for (int r = tid; r < n; r += NT) ix[r] = iq[ip[r] - 1];
    } else {
        for (int r = tid; r < n; r += NT) ix[r] = iq[r];
    }
    __syncthreads();
    for (int r = tid; r < n; r += NT) {
        const int p = ix[r] - 1;
        A.srtz[s + r] = zcol[p] / sqrt(2.0);
        A.srtd[s + r] = d[p];
    }
    __syncthreads();
    if (rho * az <= tol) {  // all deflated (MDSBED2's early exit): INDXP = INDX, K = 0
        for (int r = tid; r < n; r += NT) ip[r] = ix[r];
        if (tid == 0) { s_k = 0; s_nrot = 0; }
        __syncthreads();
    } else {
        // Parallel pre-pass: the survivors (rho |z| > tol) in sorted order, numbered by a
        // block-wide prefix count, and each survivor's rotation test against the survivor
        // before it on the staged values.  The scan's first rotation would be one of these
        // tests, so when none passes there is no rotation at all and the scan's result is
        // the survivors kept in order, the others deflated from the back in order.
        int *pos = A.rpj + s, *L = A.rnj + s;  // scratch until a rotation is logged
        const int seg = (n + NT - 1) / NT, r0 = min(n, tid * seg), r1 = min(n, r0 + seg);
        int cnt = 0;
        for (int r = r0; r < r1; r++) cnt += !(rho * fabs(A.srtz[s + r]) <= tol);
        pre[tid] = cnt;
        if (tid == 0) s_rot = 0;
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int t = 0; t < NT; t++) { const int c = pre[t]; pre[t] = acc; acc += c; }
            s_cnt = acc;
        }
        __syncthreads();
        int q = pre[tid];
        for (int r = r0; r < r1; r++) {
            const bool big = !(rho * fabs(A.srtz[s + r]) <= tol);
            pos[r] = big ? q : -1 - (r - q);  // survivor number, or -1 - number among the deflated
            if (big) L[q++] = r;
        }
        __syncthreads();
        const int ns = s_cnt;
        for (int w = tid; w < SBW; w += NT) pfl[w] = 0u;
        __syncthreads();
        int any = 0;
        for (int j = 1 + tid; j < ns; j += NT) {
            const int rp = L[j - 1], rn = L[j];
            const double zp = A.srtz[s + rp], zn = A.srtz[s + rn];
            const double tau = hypot(zn, zp), t0 = A.srtd[s + rn] - A.srtd[s + rp];
            const double c = zn / tau, sn = -zp / tau;
            if (fabs(t0 * c * sn) <= tol) {
                any = 1;
                if (j < 32 * SBW) atomicOr(&pfl[j >> 5], 1u << (j & 31));
            }
        }
        if (any) atomicOr(&s_rot, 1);
        __syncthreads();
        if (!s_rot) {
            for (int r = tid; r < n; r += NT) {
                const int qq = pos[r];
                if (qq >= 0) {
                    A.dlamda[s + qq] = A.srtd[s + r];
                    A.w[s + qq] = A.srtz[s + r];
                    ip[qq] = ix[r];
                } else {
                    ip[n + qq] = ix[r];  // n - 1 - (-1 - qq)
                }
            }
            if (tid == 0) { s_k = ns; s_nrot = 0; }
            __syncthreads();
        }
    }
    if (!(rho * az <= tol) && s_rot) {  // a rotation happens: thread 0's scan in order
        // While the pending survivor is the previous one as staged (no rotation since it
        // was kept), the pair test is the pre-pass's: an unflagged pair keeps it without
        // being evaluated again
        int k = 0, k2 = n, pj = -1, nrot = 0, qs = 0;
        bool mod = false;
        double zp = 0.0, dp = 0.0;
        for (int c0 = 0; c0 < n; c0 += DCH) {
            const int cnt = min(DCH, n - c0);
            for (int t = tid; t < cnt; t += NT) {
                zc[t] = A.srtz[s + c0 + t];
                dc[t] = A.srtd[s + c0 + t];
                nc[t] = ix[c0 + t] - 1;
            }
            __syncthreads();
            if (tid == 0) {
                for (int t = 0; t < cnt; t++) {
                    const int nj = nc[t];
                    const double zn = zc[t], dn = dc[t];
                    if (rho * fabs(zn) <= tol) { ip[--k2] = nj + 1; continue; }
                    if (pj < 0) { pj = nj; zp = zn; dp = dn; qs = 1; continue; }
                    const bool skip = !mod && qs < 32 * SBW && !((pfl[qs >> 5] >> (qs & 31)) & 1u);
                    qs++;
                    if (skip) {
                        A.dlamda[s + k] = dp; A.w[s + k] = zp; ip[k] = pj + 1; k++;
                        pj = nj; zp = zn; dp = dn;
                        continue;
                    }
                    const double tau = hypot(zn, zp), t0 = dn - dp;
                    const double c = zn / tau, sn = -zp / tau;
                    if (fabs(t0 * c * sn) <= tol) {
                        A.rpj[s + nrot] = pj; A.rnj[s + nrot] = nj; A.rcs[s + nrot] = c; A.rsn[s + nrot] = sn;
                        nrot++;
                        const double dpn = dp * c * c + dn * sn * sn;
                        const double dnn = dp * sn * sn + dn * c * c;
                        d[pj] = dpn;
                        k2--;
                        int i = 1;
                        for (;;) {
                            if (k2 + i < n && dpn < d[ip[k2 + i] - 1]) {
                                ip[k2 + i - 1] = ip[k2 + i];
                                ip[k2 + i] = pj + 1;
                                i++;
                                continue;
                            }
                            ip[k2 + i - 1] = pj + 1;
                            break;
                        }
                        pj = nj; zp = tau; dp = dnn;
                        mod = true;
                    } else {
                        A.dlamda[s + k] = dp; A.w[s + k] = zp; ip[k] = pj + 1; k++;
                        pj = nj; zp = zn; dp = dn;
                        mod = false;
                    }
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            if (pj >= 0) { A.dlamda[s + k] = dp; A.w[s + k] = zp; ip[k] = pj + 1; k++; }
            s_k = k;
            s_nrot = nrot;
        }
    }
    __syncthreads();
    const int k = s_k;
    if (tid == 0) {
        A.upd[blockIdx.x] = SbUpd{k, s_nrot};
        MergeOut o;
        o.k = k; o.ctot[0] = o.ctot[1] = o.ctot[2] = o.ctot[3] = 0;
        o.hss = 0; o.nrot = s_nrot; o.flags = 0; o.rho = rho; o.tol = tol;
        A.mo[blockIdx.x] = o;
        A.mg2[blockIdx.x].rho = rho;
    }
    for (int j = k + tid; j < n; j += NT) A.Dn[s + j] = d[ip[j] - 1];
}

template __global__ void k_sb_deflate<1024>(const SbMerge *, SbDefl);

// Launch tables of the update from each merge's K (one thread): secular blocks, 256-wide
// chunks, and the two GEMM items of every merge.  Merges with hss[m] set take MDLAED3's
// HSS branch (mo[m].hss: HSS Loewner form, alpha, no dense S) and get empty GEMM items.
// K and the rotation count of each merge go to the trace (trc[2 (tbase + m)]).
__global__ void k_sb_tab(const SbMerge *mg, int nm, const SbUpd *upd, SbTab T, double *X, const long *xoff,
                         double *S, const long *soff, double *Qn, long ldq, double *Xz, double *Zn, long ldz, int up,
                         int kd, const int *hss, MergeOut *mo, int *trc, int tbase)
{
    if (threadIdx.x != 0) return;
    int bs = 0, cs = 0;
    for (int m = 0; m < nm; m++) {
        const SbMerge M = mg[m];
        const int K = upd[m].k;
        trc[2 * (tbase + m)] = K;
        trc[2 * (tbase + m) + 1] = upd[m].nrot;
        mo[m].hss = (hss && hss[m]) ? 1 : 0;
        T.secblk[m] = bs;
        T.chunk[m] = cs;
        if (K > 0) {
            bs += (K + SEC_WG / sec_team(K) - 1) / (SEC_WG / sec_team(K));
            cs += (K + BWG - 1) / BWG;
        }
        const bool dense = !(hss && hss[m]);
        BG q;
        q.A = X + xoff[m]; q.lda = M.n;
        q.B = S + soff[m]; q.ldb = K > 0 ? K : 1; q.btrans = 1;
        q.C = Qn + M.s + (long)M.s * ldq; q.ldc = ldq;
        q.nrow = M.n; q.n = dense ? K : 0; q.k = K; q.beta = 0;
        q.A2 = nullptr; q.B2 = nullptr; q.lda2 = 0; q.ldb2 = 0; q.k2 = 0;
        q.cmap = nullptr; q.cmoff = 0;
        T.bq[m] = q;
        BG z = q;
        z.A = S + soff[m]; z.lda = K > 0 ? K : 1;
        z.B = Xz + M.s + (long)(up + 1) * ldz; z.ldb = ldz; z.btrans = 0;
        z.C = Zn + M.s + (long)(up + 1) * ldz; z.ldc = ldz;
        z.nrow = dense ? K : 0; z.n = kd - up - 1; z.k = K;
        T.bz[m] = z;
    }
    T.secblk[nm] = bs;
    T.chunk[nm] = cs;
}

// Givens rotations of an update in the scan's order (DROT: (x, y) := (c x + s y, c y - s x)):
// on Q's columns pj, nj (one thread per row of the merge) and on the rows pj, nj of Z's
// later columns (the first kd - up - 1 threads of the merge's first block, one column each).
__global__ __launch_bounds__(BWG) void k_sb_rotate(const SbMerge *mg, const int *rowchunk, int nm, const SbUpd *upd,
                                                   const int *rpj, const int *rnj, const double *rcs,
                                                   const double *rsn, double *Q, long ldq, double *Z, long ldz, int up,
                                                   int kd)
{
    const int m = seg_of(rowchunk, nm, blockIdx.x);
    const SbMerge M = mg[m];
    const int nr = upd[m].nrot;
    if (nr == 0) return;
    const int r = (blockIdx.x - rowchunk[m]) * BWG + threadIdx.x;
    const int *pj = rpj + M.s, *nj = rnj + M.s;
    const double *cs = rcs + M.s, *sn = rsn + M.s;
    if (r < M.n) {
        double *qr = Q + (M.s + r) + (long)M.s * ldq;
        for (int e = 0; e < nr; e++) {
            const double c = cs[e], s = sn[e];
            double *xp = qr + (long)pj[e] * ldq, *yp = qr + (long)nj[e] * ldq;
            const double x = *xp, y = *yp;
            *xp = c * x + s * y;
            *yp = c * y - s * x;
        }
    }
    if (r < kd - up - 1) {
        double *zc = Z + M.s + (long)(up + 1 + r) * ldz;
        for (int e = 0; e < nr; e++) {
            const double c = cs[e], s = sn[e];
            const double x = zc[pj[e]], y = zc[nj[e]];
            zc[pj[e]] = c * x + s * y;
            zc[nj[e]] = c * y - s * x;
        }
    }
}

// The permuted columns (one workgroup per column j of a merge, source INDXP(j)): the K
// kept columns of Q into X (N x K per merge), the deflated ones straight into Qn; the
// rows of Z's later columns likewise into Xz (kept) and Zn (deflated).
__global__ __launch_bounds__(BWG) void k_sb_gather(const SbMerge *mg, const int *colchunk, int nm, const SbUpd *upd,
                                                   const int *indxp, const double *Q, double *Qn, long ldq,
                                                   const double *Z, double *Zn, long ldz, double *X, const long *xoff,
                                                   double *Xz, int up, int kd)
{
    const int m = seg_of(colchunk, nm, blockIdx.x);
    const SbMerge M = mg[m];
    const int j = blockIdx.x - colchunk[m], n = M.n, k = upd[m].k;
    const int src = indxp[M.s + j] - 1;
    const double *qs = Q + M.s + (long)(M.s + src) * ldq;
    double *dst = (j < k) ? X + xoff[m] + (long)j * n : Qn + M.s + (long)(M.s + j) * ldq;
    for (int r = threadIdx.x; r < n; r += BWG) dst[r] = qs[r];
    for (int c = up + 1 + threadIdx.x; c < kd; c += BWG) {
        const double v = Z[(M.s + src) + (long)c * ldz];
        if (j < k) Xz[(M.s + j) + (long)c * ldz] = v;
        else Zn[(M.s + j) + (long)c * ldz] = v;
    }
}

// HSS branch: the kept rows of Z's later columns transposed below each kept column of X
// ([X; Xz^T], leading dimension N + kd - up - 1).  One workgroup per column j < K of a
// merge that takes the HSS branch (hm lists them, hcol their first blocks).
__global__ __launch_bounds__(BWG) void k_sb_hss_stack(const SbMerge *mg, const int *hm, const int *hcol, int nh,
                                                      const double *X, const long *xoff, double *XS,
                                                      const long *yoff, const double *Xz, long ldz, int up, int kd)
{
    const int h = seg_of(hcol, nh, blockIdx.x);
    const int mi = hm[h];
    const SbMerge M = mg[mi];
    const int j = blockIdx.x - hcol[h], nx = M.n + kd - up - 1;
    const double *x = X + xoff[mi] + (long)j * M.n;
    double *xs = XS + yoff[h] + (long)j * nx;
    for (int r = threadIdx.x; r < M.n; r += BWG) xs[r] = x[r];
    for (int c = threadIdx.x; c < kd - up - 1; c += BWG) xs[M.n + c] = Xz[(M.s + j) + (long)(up + 1 + c) * ldz];
}

// HSS branch: Y = [X; Xz^T] S back into place -- the first N rows of each column are
// Q's new column, the rest the new row of Z's later columns.
__global__ __launch_bounds__(BWG) void k_sb_hss_scatter(const SbMerge *mg, const int *hm, const int *hcol, int nh,
                                                        const double *Y, const long *yoff, double *Qn, long ldq,
                                                        double *Zn, long ldz, int up, int kd)
{
    const int h = seg_of(hcol, nh, blockIdx.x);
    const SbMerge M = mg[hm[h]];
    const int j = blockIdx.x - hcol[h], nx = M.n + kd - up - 1;
    const double *y = Y + yoff[h] + (long)j * nx;
    double *q = Qn + M.s + (long)(M.s + j) * ldq;
    for (int r = threadIdx.x; r < M.n; r += BWG) q[r] = y[r];
    for (int c = threadIdx.x; c < kd - up - 1; c += BWG) Zn[(M.s + j) + (long)(up + 1 + c) * ldz] = y[M.n + c];
}

// After the products: Dn[0:K) = the roots, INDXQ = DLAMRG(K, N-K, Dn, 1, -1)
// (mdsbed1.f90:235-237); K = 0 leaves INDXQ = 1..N.  One workgroup per merge.
__global__ __launch_bounds__(BWG) void k_sb_post(const SbMerge *mg, const SbUpd *upd, const double *lam, double *Dn,
                                                 int *indxq)
{
    __shared__ int flag;
    const SbMerge M = mg[blockIdx.x];
    const int n = M.n, k = upd[blockIdx.x].k;
    double *a = Dn + M.s;
    int *out = indxq + M.s;
    for (int j = threadIdx.x; j < k; j += BWG) a[j] = lam[M.s + j];
    __syncthreads();
    if (k == 0) {
        for (int j = threadIdx.x; j < n; j += BWG) out[j] = j + 1;
        return;
    }
    blk_lamrg<BWG>(k, n - k, a, 1, -1, out, &flag);
}

// Eigenvalues in INDXQ order and the matching columns of Q into the output (one
// workgroup per column).
__global__ __launch_bounds__(BWG) void k_sb_final(int n, const int *indxq, const double *D, const double *Q, long ldq,
                                                  double *dout, double *Zout, long ldz)
{
    const int j = blockIdx.x;
    const int src = indxq[j] - 1;
    if (threadIdx.x == 0) dout[j] = D[src];
    if (!Zout) return;
    const double *qs = Q + (long)src * ldq;
    double *zd = Zout + (long)j * ldz;
    for (int r = threadIdx.x; r < n; r += BWG) zd[r] = qs[r];
}

// KD = 0: the diagonal itself in ascending order (ties by index) -- each entry's place
// is the count of entries before it -- and the matching unit vectors.  The diagonal is
// row 0 of AB in lower storage and row kd = 0 in upper storage alike.
__global__ __launch_bounds__(BWG) void k_sb_diag(int n, const double *ab, int ldab, double *dout, double *Zout,
                                                 long ldz)
{
    const int i = blockIdx.x * BWG + threadIdx.x;
    if (i >= n) return;
    const double v = ab[(long)i * ldab];
    int pos = 0;
    for (int j = 0; j < n; j++) {
        const double u = ab[(long)j * ldab];
        pos += (u < v) || (u == v && j < i);
    }
    dout[pos] = v;
    if (Zout)
        for (long r = 0; r < n; r++) Zout[r + (long)pos * ldz] = (r == i) ? 1.0 : 0.0;
}

// ===========================================================================
// Host driver
// ===========================================================================
int hsdc_hss_batch(hsdc_ctx *c, hipStream_t s, const std::vector<HssSpec> &sp, int ni, double tol,
                   std::vector<int> *rank, int *dense);

namespace {

struct SBuf {
    void *p = nullptr;
    size_t cap = 0;
};

int sb_ensure(SBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return HSDC_ERR_NOMEM;
    }
    b.cap = bytes;
    return 0;
}

void sb_free(SBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

inline long cdivl(long a, long b) { return (a + b - 1) / b; }
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// The DC tree (mdsbevd.f90:310-330): halve every block until the last one is <= smlsiz.
struct SbLevelPlan {
    std::vector<SbMerge> m;
    int nmax = 0;  // largest merge order
    long xcap = 0; // sum of N^2 (X and S blocks)
    size_t off = 0;  // byte offset of the level's tables in the static table buffer
};
struct SbPlan {
    int n = 0, kd = 0, nb = 0;
    std::vector<int> leaf_st, leaf_m, cut_ic;
    std::vector<SbLevelPlan> lv;
};

void sb_plan(SbPlan &P, int n, int kd, int smlsiz)
{
    P.n = n; P.kd = kd;
    std::vector<int> ends(1, n);
    while (ends.back() > smlsiz) {
        std::vector<int> e2(2 * ends.size());
        for (size_t j = 0; j < ends.size(); j++) {
            e2[2 * j] = ends[j] / 2;
            e2[2 * j + 1] = (ends[j] + 1) / 2;
        }
        ends.swap(e2);
    }
    const int nb = (int)ends.size();
    P.nb = nb;
    for (int j = 1; j < nb; j++) ends[j] += ends[j - 1];
    const std::vector<int> lend = ends;
    P.leaf_st.assign(nb, 0);
    P.leaf_m.assign(nb, 0);
    for (int i = 0; i < nb; i++) {
        P.leaf_st[i] = i ? lend[i - 1] : 0;
        P.leaf_m[i] = lend[i] - P.leaf_st[i];
    }
    P.cut_ic.assign(nb > 1 ? nb - 1 : 0, 0);
    for (int i = 0; i + 1 < nb; i++) P.cut_ic[i] = lend[i];
    P.lv.clear();
    int subpbs = nb;
    while (subpbs > 1) {
        SbLevelPlan L;
        for (int i = 0; i <= subpbs - 2; i += 2) {
            int submat, matsiz, msd2;
            if (i == 0) { submat = 0; matsiz = ends[1]; msd2 = ends[0]; }
            else { submat = ends[i - 1]; matsiz = ends[i + 1] - ends[i - 1]; msd2 = matsiz / 2; }
            const int cutpt = submat + msd2;
            int j = 0;
            while (j + 1 < nb && lend[j] != cutpt) j++;
            L.m.push_back(SbMerge{submat, matsiz, msd2, j});
            L.nmax = std::max(L.nmax, matsiz);
            L.xcap += (long)matsiz * matsiz;
            ends[i / 2] = ends[i + 1];
        }
        P.lv.push_back(L);
        subpbs /= 2;
    }
}

}  // namespace

struct hsdc_sb_ctx {
    hsdc_sb_params prm;
    hsdc_sb_stats st;
    hipStream_t stream = 0;
    SBuf q[2], z[2], xz, x, s, xs, ys, dv, iv, cuts, tabs, dyn, trc, ab, dd;
    void *pin = nullptr;  // pinned staging: K of a level's merges, HSS tables
    size_t pin_cap = 0;
    hsdc_ctx *hctx = nullptr;  // HSS construction and multiply (created on the first HSS update)
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<int> trace;    // per update and merge: s, n, k, nrot, hss rank
};

#define SCK(x) do { int _r = (x); if (_r) return _r; } while (0)
#define SHK(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); return HSDC_ERR_HIP; } } while (0)

namespace {

int sb_pin(hsdc_sb_ctx *c, size_t bytes)
{
    if (c->pin_cap >= bytes) return 0;
    if (c->pin) (void)hipHostFree(c->pin);
    c->pin = nullptr;
    c->pin_cap = 0;
    if (hipHostMalloc(&c->pin, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->pin = nullptr;
        return HSDC_ERR_NOMEM;
    }
    c->pin_cap = bytes;
    return 0;
}

// The whole solve on device arrays: ab (ldab x n band storage), eigenvalues ascending
// into dD, vectors into dQ (ldq; may be null).  Returns 0, 1 (a Jacobi sweep limit or a
// secular root failed: INFO = 1) or a negative error.
int sb_solve(hsdc_sb_ctx *c, int lower, int n, int kd, const double *dab, long ldab, double *dD, double *dQ, long ldq)
{
    hipStream_t st = c->stream;
    memset(&c->st, 0, sizeof c->st);
    c->trace.clear();
    if (!c->ev[0]) {
        SHK(hipEventCreate(&c->ev[0]));
        SHK(hipEventCreate(&c->ev[1]));
    }
    SHK(hipEventRecord(c->ev[0], st));
    if (kd == 0) {
        hipLaunchKernelGGL(k_sb_diag, dim3((unsigned)cdivl(n, BWG)), dim3(BWG), 0, st, n, dab, (int)ldab, dD, dQ, ldq);
        SHK(hipGetLastError());
        SHK(hipEventRecord(c->ev[1], st));
        SHK(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        SHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        c->st.ms_total = ms;
        c->st.leaves = n;
        return 0;
    }
    const int smlsiz = c->prm.smlsiz;
    SbPlan P;
    sb_plan(P, n, kd, smlsiz);
    const int nb = P.nb, nlv = (int)P.lv.size();
    c->st.leaves = nb;
    c->st.levels = nlv;
    // ---- device memory
    const long nn = (long)n * n;
    SCK(sb_ensure(c->q[0], sizeof(double) * nn));
    SCK(sb_ensure(c->q[1], nlv ? sizeof(double) * nn : 8));
    long xcap = 1, nmx = 1;
    int nmax_m = 1;
    for (const SbLevelPlan &L : P.lv) {
        xcap = std::max(xcap, L.xcap);
        nmx = std::max(nmx, (long)L.nmax);
        nmax_m = std::max(nmax_m, (int)L.m.size());
    }
    if (nlv) {
        SCK(sb_ensure(c->z[0], sizeof(double) * (size_t)n * kd));
        SCK(sb_ensure(c->z[1], sizeof(double) * (size_t)n * kd));
        SCK(sb_ensure(c->xz, sizeof(double) * (size_t)n * kd));
        SCK(sb_ensure(c->x, sizeof(double) * (size_t)xcap));
        SCK(sb_ensure(c->s, sizeof(double) * (size_t)xcap));
    }
    const int ND = 14, NI = 8;
    SCK(sb_ensure(c->dv, sizeof(double) * (size_t)n * ND));
    SCK(sb_ensure(c->iv, sizeof(int) * (size_t)n * NI));
    SCK(sb_ensure(c->cuts, sizeof(double) * (size_t)std::max(nb - 1, 1) * (2 * kd * kd + kd)));
    double *dv = (double *)c->dv.p;
    int *iv = (int *)c->iv.p;
    double *Dp[2] = {dv, dv + n};
    Work W;
    memset(&W, 0, sizeof W);
    W.dlamda = dv + 2L * n; W.w = dv + 3L * n; W.lam = dv + 4L * n; W.tau = dv + 5L * n; W.wnew = dv + 6L * n;
    W.difl = dv + 7L * n; W.difr = dv + 8L * n; W.alpha = dv + 9L * n;
    double *srtd = dv + 10L * n, *srtz = dv + 11L * n, *rcs = dv + 12L * n, *rsn = dv + 13L * n;
    int *indxq = iv, *indx = iv + n, *indxp = iv + 2L * n;
    W.indxc = iv + 3L * n; W.org = iv + 4L * n; W.nit = iv + 5L * n;
    int *rpj = iv + 6L * n, *rnj = iv + 7L * n;
    // ---- static tables: leaves, cuts, then per level merges (two views), row and
    // column chunks, X and S offsets
    std::vector<unsigned char> hb;
    auto put = [&](const void *src, size_t bytes) {
        size_t o = al256(hb.size());
        hb.resize(o + al256(bytes));
        if (bytes) memcpy(hb.data() + o, src, bytes);
        return o;
    };
    const size_t o_lst = put(P.leaf_st.data(), sizeof(int) * nb);
    const size_t o_lm = put(P.leaf_m.data(), sizeof(int) * nb);
    const size_t o_cut = put(P.cut_ic.data(), sizeof(int) * std::max(nb - 1, 1));
    struct LvOff { size_t mg, mg2, rc, cc, xo, so; int nrowb, ncol; };
    std::vector<LvOff> lo(nlv);
    for (int l = 0; l < nlv; l++) {
        const SbLevelPlan &L = P.lv[l];
        const int nm = (int)L.m.size();
        std::vector<Merge> m2(nm);
        std::vector<int> rc(nm + 1), cc(nm + 1);
        std::vector<long> xo(nm);
        long xacc = 0;
        rc[0] = cc[0] = 0;
        for (int m = 0; m < nm; m++) {
            const SbMerge &M = L.m[m];
            m2[m].s = M.s; m2[m].n = M.n; m2[m].n1 = M.n1; m2[m].prob = 0; m2[m].lvl = 0; m2[m].rho = 0.0;
            rc[m + 1] = rc[m] + (int)cdivl(M.n, BWG);
            cc[m + 1] = cc[m] + M.n;
            xo[m] = xacc;
            xacc += (long)M.n * M.n;
        }
        lo[l].mg = put(L.m.data(), sizeof(SbMerge) * nm);
        lo[l].mg2 = put(m2.data(), sizeof(Merge) * nm);
        lo[l].rc = put(rc.data(), sizeof(int) * (nm + 1));
        lo[l].cc = put(cc.data(), sizeof(int) * (nm + 1));
        lo[l].xo = put(xo.data(), sizeof(long) * nm);
        lo[l].so = lo[l].xo;  // S blocks at the same offsets as the X blocks
        lo[l].nrowb = rc[nm];
        lo[l].ncol = cc[nm];
    }
    SCK(sb_ensure(c->tabs, hb.size()));
    unsigned char *tb = (unsigned char *)c->tabs.p;
    SHK(hipMemcpyAsync(tb, hb.data(), hb.size(), hipMemcpyHostToDevice, st));
    // ---- dynamic tables of one update, secular control, INFO, trace
    long ntr = 0;
    for (const SbLevelPlan &L : P.lv) ntr += (long)L.m.size() * kd;
    size_t dsz = 0;
    auto dput = [&](size_t bytes) { size_t o = al256(dsz); dsz = o + al256(bytes); return o; };
    const size_t o_mo = dput(sizeof(MergeOut) * nmax_m), o_upd = dput(sizeof(SbUpd) * nmax_m);
    const size_t o_sb = dput(sizeof(int) * (nmax_m + 1)), o_ch = dput(sizeof(int) * (nmax_m + 1));
    const size_t o_bq = dput(sizeof(BG) * nmax_m), o_bz = dput(sizeof(BG) * nmax_m);
    const size_t o_hf = dput(sizeof(int) * nmax_m), o_hm = dput(sizeof(int) * nmax_m);
    const size_t o_hc = dput(sizeof(int) * (nmax_m + 1)), o_yo = dput(sizeof(long) * nmax_m);
    const size_t o_gate = dput(sizeof(int) * 2);  // gate, INFO
    const size_t o_sec = dput(sizeof(SecCtl)), o_cnt = dput(sizeof(unsigned long long) * 4 * SEC_SLOTS);
    const size_t o_lvm = dput(sizeof(unsigned int) * SEC_SLOTS);
    SCK(sb_ensure(c->dyn, dsz));
    SCK(sb_ensure(c->trc, sizeof(int) * 2 * (size_t)std::max(ntr, 1L)));
    unsigned char *db = (unsigned char *)c->dyn.p;
    MergeOut *dmo = (MergeOut *)(db + o_mo);
    SbUpd *dupd = (SbUpd *)(db + o_upd);
    SbTab T;
    T.secblk = (int *)(db + o_sb); T.chunk = (int *)(db + o_ch);
    T.bq = (BG *)(db + o_bq); T.bz = (BG *)(db + o_bz);
    int *dhf = (int *)(db + o_hf), *dhm = (int *)(db + o_hm), *dhc = (int *)(db + o_hc);
    long *dyo = (long *)(db + o_yo);
    int *dgate = (int *)(db + o_gate), *dinfo = dgate + 1;
    W.gate = dgate;
    unsigned long long *dcnt = (unsigned long long *)(db + o_cnt);
    {
        SecCtl C;
        memset(&C, 0, sizeof C);
        C.w = nullptr;
        C.iters = dcnt; C.used = dcnt + SEC_SLOTS; C.wave_iters = dcnt + 2 * SEC_SLOTS; C.evals = dcnt + 3 * SEC_SLOTS;
        C.hf = nullptr; C.ht = nullptr; C.feat = nullptr; C.count = nullptr; C.cap = 0;
        C.kmin = 1 << 30; C.gate = 0.0f;
        SHK(hipMemsetAsync(db + o_gate, 0, o_lvm + sizeof(unsigned int) * SEC_SLOTS - o_gate, st));
        SHK(hipMemcpyAsync(db + o_sec, &C, sizeof C, hipMemcpyHostToDevice, st));
        SHK(hipStreamSynchronize(st));  // C and hb are host locals
    }
    const SecCtl *dsec = (const SecCtl *)(db + o_sec);
    unsigned int *dlvm = (unsigned int *)(db + o_lvm);
    int *dtrc = (int *)c->trc.p;
    double *cuts = (double *)c->cuts.p;
    // ---- split and leaves
    if (nb > 1)
        hipLaunchKernelGGL(k_sb_split, dim3(nb - 1), dim3(BWG), 0, st, dab, (int)ldab, kd, lower,
                           (const int *)(tb + o_cut), cuts, dinfo);
    int mmax = 1;
    for (int m : P.leaf_m) mmax = std::max(mmax, m);
    double *Qc[2] = {(double *)c->q[0].p, (double *)c->q[1].p};
    double *Zc[2] = {(double *)c->z[0].p, (double *)c->z[1].p};
    hipLaunchKernelGGL(k_sb_leaf, dim3(nb), dim3(BWG), sizeof(double) * (size_t)mmax * mmax, st, dab, (int)ldab, kd,
                       lower, (const int *)(tb + o_lst), (const int *)(tb + o_lm), nb, (const double *)cuts, Qc[0],
                       (long)n, Dp[0], indxq, dinfo);
    SHK(hipGetLastError());
    // ---- merges
    hsdc_stats h0;  // the HSS context's totals before this solve (its figures are the change)
    memset(&h0, 0, sizeof h0);
    if (c->hctx) {
        hsdc_params hp;
        hsdc_default_params(&hp);
        hp.profile = c->prm.profile;
        SCK(hsdc_set_params(c->hctx, &hp));
        SCK(hsdc_get_stats(c->hctx, &h0));
    }
    int cur = 0, zc = 0;
    long tbase = 0;
    const int hmin = c->prm.hss_min_k;
    std::vector<int> rk;
    for (int l = 0; l < nlv; l++) {
        const SbLevelPlan &L = P.lv[l];
        const int nm = (int)L.m.size();
        const SbMerge *dm = (const SbMerge *)(tb + lo[l].mg);
        Merge *dm2 = (Merge *)(tb + lo[l].mg2);
        const int *rowc = (const int *)(tb + lo[l].rc), *colc = (const int *)(tb + lo[l].cc);
        const long *xo = (const long *)(tb + lo[l].xo), *so = (const long *)(tb + lo[l].so);
        const bool hss_lvl = hmin > 0 && L.nmax >= hmin;
        hipLaunchKernelGGL(k_sb_zero_off, dim3(lo[l].ncol), dim3(BWG), 0, st, dm, colc, nm, Qc[cur], (long)n);
        hipLaunchKernelGGL(k_sb_zmat, dim3(lo[l].nrowb), dim3(BWG), 0, st, dm, rowc, nm, (const double *)Qc[cur],
                           (long)n, (const double *)cuts, kd, Zc[zc], (long)n, W.indxc);
        SHK(hipGetLastError());
        const unsigned gsec = (unsigned)std::min<long>(4096, [&] {
            long b = 0;
            for (const SbMerge &M : L.m) b += cdivl(M.n, SEC_WG / sec_team(M.n));
            return std::max(b, 1L);
        }());
        const unsigned gch = (unsigned)std::min<long>(4096, std::max<long>(lo[l].nrowb, 1));
        for (int up = 0; up < kd; up++) {
            SbDefl A;
            A.Z = Zc[zc]; A.ldz = n; A.D = Dp[cur]; A.Dn = Dp[cur ^ 1];
            A.indxq = indxq; A.indx = indx; A.indxp = indxp;
            A.dlamda = W.dlamda; A.w = W.w; A.srtd = srtd; A.srtz = srtz;
            A.rpj = rpj; A.rnj = rnj; A.rcs = rcs; A.rsn = rsn;
            A.cuts = cuts; A.kd = kd; A.up = up; A.upd = dupd; A.mo = dmo; A.mg2 = dm2;
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sb_deflate<1024>), dim3(nm), dim3(1024), 0, st, dm, A);
            std::vector<int> hsel;  // merges of this update that take the HSS branch
            if (hss_lvl) {
                SCK(sb_pin(c, sizeof(SbUpd) * nm + sizeof(int) * (4 * nm + 2) + sizeof(long) * nm + 64));
                SbUpd *hu = (SbUpd *)c->pin;
                SHK(hipMemcpyAsync(hu, dupd, sizeof(SbUpd) * nm, hipMemcpyDeviceToHost, st));
                SHK(hipStreamSynchronize(st));
                int *hf = (int *)(hu + nm);
                for (int m = 0; m < nm; m++) {
                    hf[m] = hu[m].k >= std::max(hmin, 3) ? 1 : 0;  // the HSS Loewner form needs K >= 3
                    if (hf[m]) hsel.push_back(m);
                }
                SHK(hipMemcpyAsync(dhf, hf, sizeof(int) * nm, hipMemcpyHostToDevice, st));
            }
            hipLaunchKernelGGL(k_sb_tab, dim3(1), dim3(64), 0, st, dm, nm, (const SbUpd *)dupd, T, (double *)c->x.p,
                               xo, (double *)c->s.p, so, Qc[cur ^ 1], (long)n, (double *)c->xz.p, Zc[zc ^ 1], (long)n,
                               up, kd, (const int *)(hss_lvl ? dhf : nullptr), dmo, dtrc, (int)tbase);
            hipLaunchKernelGGL(k_sb_rotate, dim3(lo[l].nrowb), dim3(BWG), 0, st, dm, rowc, nm, (const SbUpd *)dupd,
                               (const int *)rpj, (const int *)rnj, (const double *)rcs, (const double *)rsn, Qc[cur],
                               (long)n, Zc[zc], (long)n, up, kd);
            hipLaunchKernelGGL(k_sb_gather, dim3(lo[l].ncol), dim3(BWG), 0, st, dm, colc, nm, (const SbUpd *)dupd,
                               (const int *)indxp, (const double *)Qc[cur], Qc[cur ^ 1], (long)n,
                               (const double *)Zc[zc], Zc[zc ^ 1], (long)n, (double *)c->x.p, xo, (double *)c->xz.p,
                               up, kd);
            hipLaunchKernelGGL(k_secular, dim3(gsec), dim3(SEC_WG), 0, st, (const Merge *)dm2, (const int *)T.secblk,
                               nm, (const MergeOut *)dmo, W, c->prm.maxit, dinfo, dsec, dlvm);
            hipLaunchKernelGGL(k_loewner, dim3(gch), dim3(BWG), 0, st, (const Merge *)dm2, (const int *)T.chunk, nm,
                               (const MergeOut *)dmo, W);
            if (!hsel.empty())
                hipLaunchKernelGGL(k_alpha, dim3(gch), dim3(BWG), 0, st, (const Merge *)dm2, (const int *)T.chunk,
                                   nm, (const MergeOut *)dmo, W);
            hipLaunchKernelGGL(k_sbuild, dim3(gch), dim3(BWG), 0, st, (const Merge *)dm2, (const int *)T.chunk, nm,
                               (const MergeOut *)dmo, so, W, (double *)c->s.p);
            if (L.nmax >= 4096)  // wide tiles where they fill the chip (same bits as k_bgemm)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<256, 128, 16, 1, 2, 4>),
                                   dim3((unsigned)cdivl(L.nmax, 256), nm, (unsigned)cdivl(L.nmax, 128)), dim3(1024), 0,
                                   st, (const BG *)T.bq, 0);
            else
                hipLaunchKernelGGL(k_bgemm, dim3((unsigned)cdivl(L.nmax, 128), nm, (unsigned)cdivl(L.nmax, 64)),
                                   dim3(BWG), 0, st, (const BG *)T.bq);
            if (up + 1 < kd)
                hipLaunchKernelGGL(k_bgemm, dim3((unsigned)cdivl(L.nmax, 128), nm, 1), dim3(BWG), 0, st,
                                   (const BG *)T.bz);
            SHK(hipGetLastError());
            if (!hsel.empty()) {
                // [X; Xz^T] S through the HSS form of S (MDSBED5 -> DHSSEVC)
                const SbUpd *hu = (const SbUpd *)c->pin;
                const int nh = (int)hsel.size();
                int *hm = (int *)c->pin + 2 * nm + nm;  // after hu (2 ints each) and hf
                int *hc = hm + nh;
                long *yo = (long *)(((uintptr_t)(hc + nh + 1) + 7) & ~(uintptr_t)7);
                long yacc = 0;
                hc[0] = 0;
                for (int h = 0; h < nh; h++) {
                    const int m = hsel[h], K = hu[m].k, nx = L.m[m].n + kd - up - 1;
                    hm[h] = m;
                    hc[h + 1] = hc[h] + K;
                    yo[h] = yacc;
                    yacc += (long)nx * K;
                }
                SCK(sb_ensure(c->xs, sizeof(double) * (size_t)std::max(yacc, 1L)));
                SCK(sb_ensure(c->ys, sizeof(double) * (size_t)std::max(yacc, 1L)));
                SHK(hipMemcpyAsync(dhm, hm, sizeof(int) * nh, hipMemcpyHostToDevice, st));
                SHK(hipMemcpyAsync(dhc, hc, sizeof(int) * (nh + 1), hipMemcpyHostToDevice, st));
                SHK(hipMemcpyAsync(dyo, yo, sizeof(long) * nh, hipMemcpyHostToDevice, st));
                double *XS = (double *)c->xs.p, *YS = (double *)c->ys.p;
                hipLaunchKernelGGL(k_sb_hss_stack, dim3(hc[nh]), dim3(BWG), 0, st, dm, (const int *)dhm,
                                   (const int *)dhc, nh, (const double *)c->x.p, xo, XS, (const long *)dyo,
                                   (const double *)c->xz.p, (long)n, up, kd);
                SHK(hipGetLastError());
                std::vector<HssSpec> sp(nh);
                for (int h = 0; h < nh; h++) {
                    const SbMerge &M = L.m[hsel[h]];
                    const int K = hu[hsel[h]].k, nx = M.n + kd - up - 1;
                    HssSpec &q = sp[h];
                    q.K = K;
                    q.FF = W.dlamda + M.s; q.D0 = W.lam + M.s; q.U0 = W.wnew + M.s; q.V0 = W.alpha + M.s;
                    q.DIFL = W.difl + M.s; q.DIFR = W.difr + M.s;
                    q.X = XS + yo[h]; q.ldx = nx; q.nrow = nx;
                    q.Y = YS + yo[h]; q.ldy = nx;
                    q.sq = 0;
                }
                if (!c->hctx) {
                    SCK(hsdc_ctx_create(&c->hctx));
                    hsdc_params hp;
                    hsdc_default_params(&hp);
                    hp.profile = c->prm.profile;
                    SCK(hsdc_set_params(c->hctx, &hp));
                    SCK(hsdc_get_stats(c->hctx, &h0));
                }
                int nd = 0;
                SCK(hsdc_hss_batch(c->hctx, st, sp, c->prm.hss_leaf, c->prm.hss_tol, &rk, &nd));
                hipLaunchKernelGGL(k_sb_hss_scatter, dim3(hc[nh]), dim3(BWG), 0, st, dm, (const int *)dhm,
                                   (const int *)dhc, nh, (const double *)YS, (const long *)dyo, Qc[cur ^ 1], (long)n,
                                   Zc[zc ^ 1], (long)n, up, kd);
                SHK(hipGetLastError());
                c->st.hss_updates += nh;
                for (int h = 0; h < nh; h++) c->st.hss_rank_max = std::max(c->st.hss_rank_max, rk[h]);
                for (int h = 0; h < nh; h++) {  // (trace index, rank) pairs until the end of the solve
                    c->trace.push_back((int)(tbase + hsel[h]));
                    c->trace.push_back(rk[h]);
                }
            }
            hipLaunchKernelGGL(k_sb_post, dim3(nm), dim3(BWG), 0, st, dm, (const SbUpd *)dupd, (const double *)W.lam,
                               Dp[cur ^ 1], indxq);
            SHK(hipGetLastError());
            cur ^= 1;
            zc ^= 1;
            tbase += nm;
            c->st.updates += nm;
        }
        c->st.merges += nm;
    }
    hipLaunchKernelGGL(k_sb_final, dim3(n), dim3(BWG), 0, st, n, (const int *)indxq, (const double *)Dp[cur],
                       (const double *)Qc[cur], (long)n, dD, dQ, ldq);
    SHK(hipGetLastError());
    SHK(hipEventRecord(c->ev[1], st));
    SHK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    SHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->st.ms_total = ms;
    // statistics and the per-update trace
    int info = 0;
    SHK(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> cnt(4 * SEC_SLOTS);
    SHK(hipMemcpy(cnt.data(), dcnt, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < SEC_SLOTS; i++) c->st.secular_iters += (long long)cnt[i];
    std::vector<int> tr(2 * (size_t)std::max(ntr, 1L));
    SHK(hipMemcpy(tr.data(), dtrc, sizeof(int) * 2 * ntr, hipMemcpyDeviceToHost));
    if (c->hctx) {
        hsdc_stats h1;
        SCK(hsdc_get_stats(c->hctx, &h1));
        c->st.hss_nflops = h1.hss_nflops - h0.hss_nflops;
        c->st.hss_mm_flops = h1.hss_mm_flops - h0.hss_mm_flops;
        c->st.ms_hss_build = h1.ms_hss_build - h0.ms_hss_build;
        c->st.ms_hss_mm = h1.ms_hss_mm - h0.ms_hss_mm;
    }
    std::vector<int> hr(ntr, 0);
    for (size_t i = 0; i + 1 < c->trace.size(); i += 2) hr[c->trace[i]] = c->trace[i + 1];
    c->trace.clear();
    long t = 0;
    for (int l = 0; l < nlv; l++) {
        const SbLevelPlan &L = P.lv[l];
        for (int up = 0; up < kd; up++)
            for (size_t m = 0; m < L.m.size(); m++, t++) {
                const int k = tr[2 * t], nr = tr[2 * t + 1];
                c->trace.insert(c->trace.end(), {L.m[m].s, L.m[m].n, k, nr, hr[t]});
                c->st.roots += k;
                c->st.deflated += L.m[m].n - k;
                if (hr[t] == 0)  // dense branch: Q(:, 1:K) := X S and Z(1:K, later) := S^T Z
                    c->st.dense_flops += 2LL * L.m[m].n * k * k + 2LL * k * k * (kd - up - 1);
                c->st.rotations += nr;
                if (l == nlv - 1 && up == kd - 1) c->st.k_top = k;
            }
    }
    return info ? 1 : 0;
}

}  // namespace

static thread_local hsdc_sb_ctx *t_sb_ctx = nullptr;

// releases the calling thread's banded context (hsdc_thread_ctx_release calls it)
extern "C" __attribute__((visibility("hidden"))) void hsdc_sb_thread_release(void)
{
    if (t_sb_ctx) (void)hsdc_sb_ctx_destroy(t_sb_ctx);
    t_sb_ctx = nullptr;
}

extern "C" {

void hsdc_sb_default_params(hsdc_sb_params *p)
{
    p->smlsiz = 128;
    p->maxit = 100;
    p->hss_min_k = 1500;            // BSMLZ, mdsbed1.f90:134
    p->hss_leaf = 40;               // Ni, mdsbed5.f90:276
    p->hss_tol = (double)1.0e-17f;  // TOL of DHSSEVC, a REAL literal (CauchyHssEig.f90:2523)
    p->profile = 0;
}

int hsdc_sb_ctx_create(hsdc_sb_ctx **ctx)
{
    if (!ctx) return HSDC_ERR_INTERNAL;
    hsdc_sb_ctx *c = new hsdc_sb_ctx();
    hsdc_sb_default_params(&c->prm);
    memset(&c->st, 0, sizeof c->st);
    *ctx = c;
    return 0;
}

int hsdc_sb_ctx_destroy(hsdc_sb_ctx *c)
{
    if (!c) return 0;
    SBuf *bufs[] = {&c->q[0], &c->q[1], &c->z[0], &c->z[1], &c->xz, &c->x, &c->s, &c->xs, &c->ys, &c->dv, &c->iv,
                    &c->cuts, &c->tabs, &c->dyn, &c->trc, &c->ab, &c->dd};
    for (SBuf *b : bufs) sb_free(*b);
    if (c->pin) (void)hipHostFree(c->pin);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->hctx) (void)hsdc_ctx_destroy(c->hctx);
    delete c;
    return 0;
}

int hsdc_sb_set_params(hsdc_sb_ctx *c, const hsdc_sb_params *p)
{
    if (!c || !p) return HSDC_ERR_INTERNAL;
    if (p->smlsiz < 1 || p->smlsiz > BLEAF_MAX) return -1;
    if (p->maxit < 1) return -2;
    if (p->hss_min_k < 0) return -3;
    if (p->hss_leaf < 8) return -4;
    if (!(p->hss_tol > 0.0)) return -5;
    if (p->profile != 0 && p->profile != 1) return -6;
    c->prm = *p;
    return 0;
}

int hsdc_sb_get_stats(hsdc_sb_ctx *c, hsdc_sb_stats *s)
{
    if (!c || !s) return HSDC_ERR_INTERNAL;
    *s = c->st;
    return 0;
}

int hsdc_sb_set_stream(hsdc_sb_ctx *c, void *stream)
{
    if (!c) return HSDC_ERR_INTERNAL;
    c->stream = (hipStream_t)stream;
    return 0;
}

hsdc_sb_ctx *hsdc_sb_thread_ctx(void)
{
    if (!t_sb_ctx && hsdc_sb_ctx_create(&t_sb_ctx) != 0) t_sb_ctx = nullptr;
    if (t_sb_ctx) {
        hsdc_sb_params p;
        hsdc_sb_default_params(&p);
        (void)hsdc_sb_set_params(t_sb_ctx, &p);
        (void)hsdc_sb_set_stream(t_sb_ctx, nullptr);
    }
    return t_sb_ctx;
}

static bool sb_have_device()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

// checks shared by both entry points; 0 when the call may go on
static int sb_args(hsdc_sb_ctx *c, char uplo, int n, int kd, int ldab, int ldq, bool vec)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (uplo != 'U' && uplo != 'u' && uplo != 'L' && uplo != 'l') return -2;
    if (n < 0) return -3;
    if (kd < 0 || kd > BKD_MAX) return -4;
    if (ldab < kd + 1) return -6;
    if (vec && ldq < std::max(1, n)) return -9;
    if (kd > 0 && n > c->prm.smlsiz && c->prm.smlsiz < 2 * kd - 1) return -20;
    if (n > 0 && n <= c->prm.smlsiz && n > BLEAF_MAX) return -20;
    return 0;
}

int hsdc_sbevd_dev(hsdc_sb_ctx *c, char uplo, int n, int kd, const double *d_ab, int ldab, double *d_d, double *d_q,
                   int ldq)
{
    const int rc = sb_args(c, uplo, n, kd, ldab, ldq, d_q != nullptr);
    if (rc) return rc;
    memset(&c->st, 0, sizeof c->st);
    c->trace.clear();
    if (n == 0) return 0;
    if (!d_ab) return -5;
    if (!d_d) return -7;
    if (!sb_have_device()) return HSDC_ERR_NODEVICE;
    const int lower = (uplo == 'L' || uplo == 'l');
    return sb_solve(c, lower, n, kd, d_ab, ldab, d_d, d_q, ldq);
}

int hsdc_sbevd(hsdc_sb_ctx *c, char compz, char uplo, int n, int kd, const double *ab, int ldab, double *d, double *q,
               int ldq, double *work, int lwork, int *iwork, int liwork)
{
    (void)work; (void)lwork; (void)iwork; (void)liwork;
    if (compz != 'N' && compz != 'n' && compz != 'I' && compz != 'i') return -1;
    const bool vec = (compz == 'I' || compz == 'i');
    const int rc = sb_args(c, uplo, n, kd, ldab, ldq, vec);
    if (rc) return rc;
    memset(&c->st, 0, sizeof c->st);
    c->trace.clear();
    if (n == 0) return 0;
    if (!ab) return -5;
    if (!d) return -7;
    if (vec && !q) return -8;
    if (!sb_have_device()) return HSDC_ERR_NODEVICE;
    hipStream_t st = c->stream;
    SCK(sb_ensure(c->ab, sizeof(double) * (size_t)ldab * n));
    SCK(sb_ensure(c->dd, sizeof(double) * (size_t)n * (vec ? n + 1 : 1)));
    double *dab = (double *)c->ab.p, *dd = (double *)c->dd.p, *dq = vec ? dd + n : nullptr;
    SHK(hipMemcpyAsync(dab, ab, sizeof(double) * (size_t)ldab * n, hipMemcpyHostToDevice, st));
    const int lower = (uplo == 'L' || uplo == 'l');
    const int info = sb_solve(c, lower, n, kd, dab, ldab, dd, dq, n);
    if (info < 0) return info;
    SHK(hipMemcpyAsync(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (vec)
        SHK(hipMemcpy2DAsync(q, sizeof(double) * ldq, dq, sizeof(double) * n, sizeof(double) * n, n,
                             hipMemcpyDeviceToHost, st));
    SHK(hipStreamSynchronize(st));
    return info;
}

int hsdc_sb_trace(hsdc_sb_ctx *c, int *s, int *n, int *k, int *nrot, int *hss_rank, int cap)
{
    if (!c) return HSDC_ERR_INTERNAL;
    const int cnt = (int)(c->trace.size() / 5);
    for (int i = 0; i < cnt && i < cap; i++) {
        const int *t = c->trace.data() + 5L * i;
        if (s) s[i] = t[0];
        if (n) n[i] = t[1];
        if (k) k[i] = t[2];
        if (nrot) nrot[i] = t[3];
        if (hss_rank) hss_rank[i] = t[4];
    }
    return cnt;
}

}  // extern "C"
