// HIP kernels of the divide-and-conquer merge path for gfx950.
//
// Every kernel works on all merges of one DC level at once (a table of Merge
// descriptors); launching it with a one-entry table gives the merge-at-a-time
// schedule of the reference (mdlaed0.f90:298-337).
//
// Arithmetic that decides integer outcomes (deflation, pivoting) or feeds the
// next level (roots, Loewner z) is written operation for operation like the CPU
// oracle and compiled with -ffp-contract=off, so a merge given identical inputs
// returns bit-identical roots, index sets and rotations.
#include <climits>
#include <cmath>
#include "hsdc_internal.h"
#include "hsdc_mlp.h"

#define WG 256

__device__ __forceinline__ double lapy2_d(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    double r = z / w;
    return w * sqrt(1.0 + r * r);
}

__device__ __forceinline__ double qload(const Prob &P, const int *lo, const int *hi, int row, int slot)
{
    if (row < lo[slot] || row >= hi[slot]) return 0.0;
    return P.q[(long)(row - P.off) + (long)(slot - P.off) * P.ld];
}

__device__ __forceinline__ int find_seg(const int *start, int n, int b)
{
    // Largest m < n with start[m] <= b (start nondecreasing, start[0] <= b).  All 64
    // lanes of a wavefront call it with the same arguments: each round they probe 64
    // evenly spaced entries and the ballot narrows the range 64-fold, so a table of
    // 4096 merges costs two dependent loads instead of twelve.
    const int lane = threadIdx.x & 63;
    int lo = 0, len = n;  // the answer lies in [lo, lo + len)
    while (len > 1) {
        const int step = (len + 63) >> 6;
        const int idx = lo + lane * step;
        const bool ok = idx < lo + len && start[idx] <= b;
        const int cnt = __popcll(__ballot(ok));  // a prefix of the lanes, lane 0 included
        lo += (cnt - 1) * step;
        len = min(step, len - (cnt - 1) * step);
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Device-resident input of MDSTEDC (mdstedc.f90:305-358) for a plan built without
// looking at the values (one problem, no split).  k_inp_scan: ORGNRM = max|d|,|e|
// (atomicMax on the bit patterns of non-negative doubles into maux tail word 6)
// and a screen for splits: |e_i| <= 2 eps sqrt|d_i| sqrt|d_i+1| or a NaN marks the
// input (tail word 7), twice MDSTEDC's own bound so that no rounding of the square
// roots lets a split through.  k_inp_scale: the scaling by 1/ORGNRM, the rank-one
// tears at the leaf boundaries (tflag bit 0: first row of a leaf, bit 1: last) in
// the order the host applies them, RHO of every merge and the scale-back factor;
// a marked input, ORGNRM = 0 or a non-finite ORGNRM sets the gate to INT_MIN
// instead, so every kernel of the solve returns and the host solves it itself.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_inp_scan(const double *d, const double *e, int n, Work W)
{
    const int i = blockIdx.x * WG + threadIdx.x;
    double m = 0.0;
    int sus = 0;
    if (i < n) {
        m = fabs(d[i]);
        if (i < n - 1) {
            const double ae = fabs(e[i]);
            m = fmax(m, ae);
            const double tiny = HSDC_EPS * sqrt(fabs(d[i])) * sqrt(fabs(d[i + 1]));
            sus = !(ae > 2.0 * tiny);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off));
        sus |= __shfl_xor(sus, off);
    }
    unsigned long long *tail = (unsigned long long *)W.gate;
    if ((threadIdx.x & 63) == 0) {
        if (m > 0.0) atomicMax(&tail[6], (unsigned long long)__double_as_longlong(m));
        if (sus) atomicOr(&tail[7], 1ULL);
    }
}

__global__ __launch_bounds__(WG) void k_inp_scale(const double *d, const double *e, int n, const int *tflag,
                                                  Merge *mg, int nmtot, double *dd, double *de, double *scale,
                                                  Work W)
{
    const unsigned long long *tail = (const unsigned long long *)W.gate;
    const double onrm = __longlong_as_double((long long)tail[6]);
    if (tail[7] || !(onrm > 0.0) || !(onrm < INFINITY)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *W.gate = INT_MIN;
        return;
    }
    const double mul = 1.0 / onrm;
    const int i = blockIdx.x * WG + threadIdx.x;
    if (i < n) {
        double v = d[i] * mul;
        const int f = tflag[i];
        if (f & 1) v -= fabs(e[i - 1] * mul);
        if (f & 2) v -= fabs(e[i] * mul);
        dd[i] = v;
        de[i] = (i < n - 1) ? e[i] * mul : 0.0;
    }
    if (i < nmtot) mg[i].rho = e[mg[i].s + mg[i].n1 - 1] * mul;
    if (i == 0) scale[0] = onrm;
}

// ---------------------------------------------------------------------------
// Leaves: implicit-shift QL (the tql2 iteration) on each leaf of the DC tree
// with eigenvectors, then selection sort (DSTEQR at mdlaed0.f90:263).  One wave
// per leaf; the scalar recurrence is evaluated redundantly by every lane, lane r
// owns row r of the leaf's eigenvector block in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_leaf(const Leaf *leaves, int nleaf, const Prob *probs,
                                             const double *din, const double *ein, Work W, int cur, int G,
                                             int lcap)
{
    if (*W.gate) return;  // the device-resident input was sent back to the host
    // 64 / G leaves per wavefront, G lanes (G >= leaf order) per leaf: lane r of a
    // group owns row r of its leaf's eigenvector block; every lane of the group runs
    // the same scalar QL iteration on the group's d and e (identical values written
    // by every lane), so after the initial fill no lane reads another lane's row and
    // the iteration needs no barriers.  Groups iterate independently (masked).
    extern __shared__ double lds_leaf[];
    __shared__ double sdl[64], sel[64];
    const int g = threadIdx.x / G, r = threadIdx.x - g * G, li = blockIdx.x * (64 / G) + g;
    Leaf L = {0, 0, 0};
    Prob P = {nullptr, 0, 0, 0, 0};
    int n = 0;
    if (li < nleaf) {
        L = leaves[li];
        P = probs[L.prob];
        n = L.m;
    }
    double *sz = lds_leaf + (long)g * lcap * lcap;  // n x n, column-major
    double *sd = sdl + g * G, *se = sel + g * G;
    for (int i = r; i < n * n; i += G) sz[i] = ((i % n) == (i / n)) ? 1.0 : 0.0;
    if (r < n) {
        sd[r] = din[L.s + r];
        se[r] = (r < n - 1) ? ein[L.s + r] : 0.0;
    }
    __syncthreads();
    int fail = 0;
    for (int l = 0; l < n && !fail; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; m++) {
                double dd = fabs(sd[m]) + fabs(sd[m + 1]);
                if (fabs(se[m]) <= HSDC_EPS * dd) break;
            }
            if (m == l) break;
            if (++iter > 30 * n) { fail = 1; break; }
            double g = (sd[l + 1] - sd[l]) / (2.0 * se[l]);
            double rr = lapy2_d(g, 1.0);
            g = sd[m] - sd[l] + se[l] / (g + (g >= 0.0 ? fabs(rr) : -fabs(rr)));
            double s = 1.0, c = 1.0, p = 0.0;
            int underflow = 0;
            for (int i = m - 1; i >= l; i--) {
                double f = s * se[i], b = c * se[i];
                rr = lapy2_d(f, g);
                se[i + 1] = rr;
                if (rr == 0.0) {
                    double t = sd[i + 1] - p;
                    sd[i + 1] = t;
                    se[m] = 0.0;
                    underflow = 1;
                    break;
                }
                s = f / rr;
                c = g / rr;
                g = sd[i + 1] - p;
                rr = (sd[i] - g) * s + 2.0 * c * b;
                p = s * rr;
                double dn = g + p;
                sd[i + 1] = dn;
                g = c * rr - b;
                if (r < n) {
                    double zi = sz[r + i * n], fz = sz[r + (i + 1) * n];
                    sz[r + (i + 1) * n] = s * zi + c * fz;
                    sz[r + i * n] = c * zi - s * fz;
                }
            }
            if (underflow) continue;
            double dl = sd[l] - p;
            sd[l] = dl;
            se[l] = g;
            se[m] = 0.0;
        }
    }
    // selection sort, ascending
    for (int ii = 1; ii < n; ii++) {
        int i = ii - 1, k = i;
        double p = sd[i];
        for (int j = ii; j < n; j++)
            if (sd[j] < p) { k = j; p = sd[j]; }
        if (k != i) {
            double di = sd[i];
            sd[k] = di;
            sd[i] = p;
            if (r < n) {
                double t = sz[r + i * n];
                sz[r + i * n] = sz[r + k * n];
                sz[r + k * n] = t;
            }
        }
    }
    __syncthreads();
    if (r < n) {
        long base = (long)(L.s - P.off);
        for (int c = 0; c < n; c++) P.q[base + r + (base + c) * P.ld] = sz[r + c * n];
        W.Dv[cur][L.s + r] = sd[r];
        W.vm[cur][L.s + r] = L.s + r;
        W.indxq[L.s + r] = r + 1;
        W.supp_lo[L.s + r] = L.s;
        W.supp_hi[L.s + r] = L.s + n;
    }
}

// ---------------------------------------------------------------------------
// Merge set-up (mdlaed1.f90:178-180, mdlaed2.f90:222-245): z from the last row of
// Q1 and the first row of Q2, sign flip for rho < 0, scaling by 1/sqrt(2),
// INDXQ of the second half shifted by n1, DLAMDA = D(INDXQ).
// Flattened over all merges: chunk_start[m] = first 256-element chunk of merge m.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_prep(const Merge *mg, const int *chunk_start, int nm,
                                             const Prob *probs, Work W, int cur)
{
    if (gate_closed(W, mg)) return;
    int m = find_seg(chunk_start, nm, blockIdx.x);
    Merge M = mg[m];
    int i = (blockIdx.x - chunk_start[m]) * WG + threadIdx.x;
    double az = 0.0, ad = 0.0;
    if (i < M.n) {
        int g = M.s + i;
        double zz;
        if (probs) {
            Prob P = probs[M.prob];
            int slot = W.vm[cur][g];
            int row = (i < M.n1) ? (M.s + M.n1 - 1) : (M.s + M.n1);
            zz = qload(P, W.supp_lo, W.supp_hi, row, slot);
        } else {
            zz = W.z[g];  // single-merge entry: z supplied by the caller
        }
        if (M.rho < 0.0 && i >= M.n1) zz = -1.0 * zz;
        const double t = 1.0 / sqrt(2.0);
        zz = t * zz;
        W.z[g] = zz;
        int q2 = W.indxq[g] + (i >= M.n1 ? M.n1 : 0);
        W.indxq2[g] = q2;
        const double v = W.Dv[cur][M.s + q2 - 1];
        W.dlamda[g] = v;
        // DLAMRG's runs must be ascending for the merge path; flag the merge otherwise
        if (i + 1 < M.n && i + 1 != M.n1) {
            int q2n = W.indxq[g + 1] + (i + 1 >= M.n1 ? M.n1 : 0);
            if (v > W.Dv[cur][M.s + q2n - 1]) W.maux[4 * m + 2] = 1ULL;
        }
        az = fabs(zz);
        ad = fabs(v);
    }
    // max |z| and max |D| of the merge (the deflation tolerance, mdlaed2.f90:264-266);
    // non-negative doubles order as their bit patterns
    for (int off = 32; off > 0; off >>= 1) {
        az = fmax(az, __shfl_xor(az, off));
        ad = fmax(ad, __shfl_xor(ad, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&W.maux[4 * m + 0], (unsigned long long)__double_as_longlong(az));
        atomicMax(&W.maux[4 * m + 1], (unsigned long long)__double_as_longlong(ad));
    }
}

__device__ __forceinline__ double maux_tol(const Work &W, int m)
{
    double zmax = __longlong_as_double((long long)W.maux[4 * m + 0]);
    double dmax = __longlong_as_double((long long)W.maux[4 * m + 1]);
    return 8.0 * HSDC_EPS * (dmax > zmax ? dmax : zmax);
}

// Stable merge of two sorted runs (LAPACK DLAMRG semantics, ties to the first
// run) by the merge path, one block per merge; falls back to the sequential
// merge when a run is out of order so the permutation stays DLAMRG's.
// a1 = base of run 1 read with stride s1 (s1 = -1 reads a1[n1-1 .. 0]).
__device__ void block_merge(const double *a, int n1, int n2, int s1, int s2, int *out, int *sflag)
{
    int n = n1 + n2, tid = threadIdx.x;
    auto A1 = [&](int p) { return s1 > 0 ? a[p] : a[n1 - 1 - p]; };
    auto A2 = [&](int q) { return s2 > 0 ? a[n1 + q] : a[n - 1 - q]; };
    auto I1 = [&](int p) { return s1 > 0 ? p + 1 : n1 - p; };
    auto I2 = [&](int q) { return s2 > 0 ? n1 + q + 1 : n - q; };
    if (tid == 0) *sflag = 0;
    __syncthreads();
    int bad = 0;
    for (int p = tid; p + 1 < n1; p += blockDim.x) bad |= (A1(p) > A1(p + 1));
    for (int q = tid; q + 1 < n2; q += blockDim.x) bad |= (A2(q) > A2(q + 1));
    if (bad) atomicOr(sflag, 1);
    __syncthreads();
    if (*sflag) {
        if (tid == 0) {
            int i1 = 0, i2 = 0, o = 0;
            while (i1 < n1 && i2 < n2) {
                if (A1(i1) <= A2(i2)) out[o++] = I1(i1++);
                else out[o++] = I2(i2++);
            }
            while (i1 < n1) out[o++] = I1(i1++);
            while (i2 < n2) out[o++] = I2(i2++);
        }
    } else {
        for (int p = tid; p < n1; p += blockDim.x) {
            double v = A1(p);
            int lo = 0, hi = n2;  // count of run-2 values strictly below v
            while (lo < hi) { int md = (lo + hi) >> 1; if (A2(md) < v) lo = md + 1; else hi = md; }
            out[p + lo] = I1(p);
        }
        for (int q = tid; q < n2; q += blockDim.x) {
            double v = A2(q);
            int lo = 0, hi = n1;  // count of run-1 values <= v
            while (lo < hi) { int md = (lo + hi) >> 1; if (A1(md) <= v) lo = md + 1; else hi = md; }
            out[q + lo] = I2(q);
        }
    }
    __syncthreads();
}

// A merge with a run out of order: DLAMRG(N1,N2,DLAMDA,1,1,INDXC) by one block,
// then the gathers of k_mergepath.
__device__ void sort_merge(const Merge &M, int m, const Work &W)
{
    __shared__ int sflag;
    block_merge(W.dlamda + M.s, M.n1, M.n - M.n1, 1, 1, W.indxc + M.s, &sflag);
    const double rho = fabs(2.0 * M.rho), tol = maux_tol(W, m);
    for (int i = threadIdx.x; i < M.n; i += WG) {
        int nj = W.indxq2[M.s + W.indxc[M.s + i] - 1];
        W.indx[M.s + i] = nj;
        double zz = W.z[M.s + nj - 1];
        W.zs[M.s + i] = zz;
        W.ds[M.s + i] = W.dlamda[M.s + W.indxc[M.s + i] - 1];
        W.coltyp[M.s + nj - 1] = (rho * fabs(zz) <= tol) ? 4 : ((nj <= M.n1) ? 1 : 3);
    }
}

// The merged order (DLAMRG(N1, N2, DLAMDA, 1, 1, INDXC), mdlaed2.f90:252-255) of every
// merge with ascending runs, one thread per element over all merges: each element's
// rank in the other run by binary search (ties to the first run).  Also the sorted
// gathers of the deflation scan: INDX = INDXQ(INDXC), z and D in that order, and the
// small-z column type (mdlaed2.f90:264-300).
__global__ __launch_bounds__(WG) void k_mergepath(const Merge *mg, const int *chunk_start, int nm, Work W)
{
    if (gate_closed(W, mg)) return;
    int m = find_seg(chunk_start, nm, blockIdx.x);
    if (W.maux[4 * m + 2]) {  // a run out of order (k_prep): the first block merges it alone
        if (blockIdx.x == chunk_start[m]) sort_merge(mg[m], m, W);
        return;
    }
    Merge M = mg[m];
    int i = (blockIdx.x - chunk_start[m]) * WG + threadIdx.x;
    if (i >= M.n) return;
    const double *a = W.dlamda + M.s;
    const int n1 = M.n1, n2 = M.n - M.n1;
    const double v = a[i];
    int out, ic;
    if (i < n1) {
        int lo = 0, hi = n2;  // run-2 values strictly below v
        while (lo < hi) { int md = (lo + hi) >> 1; if (a[n1 + md] < v) lo = md + 1; else hi = md; }
        out = i + lo;
        ic = i + 1;
    } else {
        int q = i - n1, lo = 0, hi = n1;  // run-1 values <= v
        while (lo < hi) { int md = (lo + hi) >> 1; if (a[md] <= v) lo = md + 1; else hi = md; }
        out = q + lo;
        ic = n1 + q + 1;
    }
    const int s = M.s;
    W.indxc[s + out] = ic;
    const int nj = W.indxq2[s + ic - 1];
    W.indx[s + out] = nj;
    const double zz = W.z[s + nj - 1];
    W.zs[s + out] = zz;
    W.ds[s + out] = v;  // D(INDX(out)) = DLAMDA(INDXC(out))
    const double rho = fabs(2.0 * M.rho), tol = maux_tol(W, m);
    W.coltyp[s + nj - 1] = (rho * fabs(zz) <= tol) ? 4 : ((nj <= n1) ? 1 : 3);
}

// ---------------------------------------------------------------------------
// Deflation (mdlaed2.f90:264-465), one block per merge, same decisions as the
// reference's sequential PJ/NJ scan:
//  1. parallel: tolerance, sorted gather, small-z screen, list of survivors;
//  2. parallel: the close-pair test of every consecutive survivor pair on the
//     unrotated values; before the first pair that deflates no value has been
//     modified, so those survivors are non-deflated as computed;
//  3. one lane continues the scan from that pair over survivors only;
//  4. the deflated region INDXP(K+1:N), which the reference keeps with an
//     insertion step (mdlaed2.f90:357-370), is replayed by one lane over runs of
//     small-z columns (binary search inside a run) and written out in parallel.
// A merge whose replay would exceed the LDS segment table runs the plain
// sequential scan instead.
// ---------------------------------------------------------------------------
#define DCH 1024
#define MAXSEG 1536
#define DU 8  // sub-tiles of 64 per step of the ballot passes
#define SPW 1024  // words of the speculative pair flags (merges with < 32768 survivors)
#define RUNS 256  // runs of kept survivors the scan may record
template <int NT>
__global__ __launch_bounds__(NT) void k_deflate(const Merge *mg, Work W, int cur, MergeOut *mo,
                                                int hss_min_k, int ds_cap)
{
    if (gate_closed(W, mg)) return;
    constexpr int NW = NT / 64;  // wavefronts
    extern __shared__ double dsl[];  // ds_cap doubles: sorted D for the replay searches
    __shared__ int sk, sk2, scnt[4], wcnt[4][NW], sL, sfirst, sfallback, snseg;
    __shared__ int sng;
    __shared__ int seg_a[MAXSEG], seg_b[MAXSEG], seg_nx[MAXSEG], seg_pos[MAXSEG], seg_ord[MAXSEG];
    __shared__ double seg_d[MAXSEG];
    // survivor positions with small-z columns before them, for the replay; seg_ord is
    // written only after the replay
    int *const gq = seg_ord;
    __shared__ double cz[DCH], cd[DCH];
    __shared__ int ci[DCH], cty[DCH], lsn[DCH];
    __shared__ int snr, snrun;
    __shared__ unsigned int spb[SPW];
    __shared__ int run_k[RUNS], run_q[RUNS], run_n[RUNS], run_o[RUNS + 1];
    Merge M = mg[blockIdx.x];
    const int n = M.n, n1 = M.n1, s = M.s, tid = threadIdx.x;
    const double *D = W.Dv[cur] + s;
    double *z = W.z + s;
    int *indx = W.indx + s, *indxp = W.indxp + s, *coltyp = W.coltyp + s;
    double *zs = W.zs + s, *ds = W.ds + s, *ddef = W.ddef + s;
    int *nsl = W.nsl + s;
    (void)z;
    (void)D;
    const double rho = fabs(2.0 * M.rho);
    const double zmax = __longlong_as_double((long long)W.maux[4 * blockIdx.x + 0]);
    const double tol = maux_tol(W, blockIdx.x);
    MergeOut *o = mo + blockIdx.x;
    __syncthreads();
    if (tid == 0) {
        // the per-merge record is consumed: clear it for the next launch
        W.maux[4 * blockIdx.x + 0] = W.maux[4 * blockIdx.x + 1] = W.maux[4 * blockIdx.x + 2] = 0ULL;
        W.maux[4 * blockIdx.x + 3] = 0ULL;  // k_newcols' order flag of this level
    }
    if (rho * zmax <= tol) {
        if (tid == 0) {
            o->k = 0; o->ctot[0] = o->ctot[1] = o->ctot[2] = o->ctot[3] = 0;
            o->hss = 0; o->nrot = 0; o->flags = 1; o->rho = rho; o->tol = tol;
        }
        return;
    }
    // Wavefront w owns sorted positions [w0, w1); the passes below walk them 64 at a
    // time and compact with ballots, which keeps every list in ascending order.
    const int lane = tid & 63, wv = tid >> 6;
    const unsigned long long lt = (1ULL << lane) - 1ULL;
    const int wper = (n + NW - 1) / NW;
    const int w0 = min(n, wv * wper), w1 = min(n, w0 + wper);
    // ---- 1. survivors of the small-z screen (zs, ds, INDX and the small-z column
    // types come from the merge kernels); DLAMDA must be ascending in scan order for
    // the run replay below
    if (tid == 0) sfallback = 0;
    __syncthreads();
    // (DU sub-tiles of 64 per step: their loads are independent and issue together)
    int cnt = 0, bad = 0;
    for (int j0 = w0; j0 < w1; j0 += 64 * DU) {
        bool sv[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            const int j = j0 + 64 * u + lane;
            sv[u] = false;
            if (j < w1) {
                const double a = ds[j], b = (j + 1 < n) ? ds[j + 1] : a;
                sv[u] = rho * fabs(zs[j]) > tol;
                if (a > b) bad = 1;
            }
        }
#pragma unroll
        for (int u = 0; u < DU; u++) cnt += __popcll(__ballot(sv[u]));
    }
    if (bad) sfallback = 1;
    if (lane == 0) wcnt[0][wv] = cnt;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; w++) base += wcnt[0][w];
    int L = 0;
    for (int w = 0; w < NW; w++) L += wcnt[0][w];
    // survivors from the front of NSL in order; small-z positions from the back
    // (rank r at NSL(n-1-r)), which is their order in the deflated region when no
    // rotation interleaves (mdlaed2.f90:304-325)
    int nbase = w0 - base;
    for (int j0 = w0; j0 < w1; j0 += 64 * DU) {
        bool sv[DU], in[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            const int j = j0 + 64 * u + lane;
            in[u] = j < w1;
            sv[u] = in[u] && rho * fabs(zs[j]) > tol;
        }
#pragma unroll
        for (int u = 0; u < DU; u++) {
            const int j = j0 + 64 * u + lane;
            unsigned long long msk = __ballot(sv[u]);
            unsigned long long nmk = __ballot(in[u] && !sv[u]);
            if (sv[u]) nsl[base + __popcll(msk & lt)] = j;
            else if (in[u]) nsl[n - 1 - (nbase + __popcll(nmk & lt))] = j;
            base += __popcll(msk);
            nbase += __popcll(nmk);
        }
    }
    if (tid == 0) { sL = L; sfirst = L - 1; }
    __syncthreads();
    // ---- 2. speculative close-pair test of every consecutive survivor pair on the
    // unrotated values: bit q of spb = pair (q, q+1) would deflate.  Where the scan
    // reaches a pair with its left survivor unmodified, this is the scan's own test.
    const bool have_bits = L < 32 * SPW;
    if (have_bits)
        for (int w = tid; w < SPW; w += NT) spb[w] = 0u;
    __syncthreads();
    int myfirst = L - 1;
    for (int q = tid; q + 1 < L; q += NT) {
        int p = nsl[q], r = nsl[q + 1];
        double ss = zs[p], cc = zs[r];
        double tau = lapy2_d(cc, ss);
        double t = ds[r] - ds[p];
        cc = cc / tau;
        ss = -ss / tau;
        if (fabs(t * cc * ss) <= tol) {
            myfirst = min(myfirst, q);
            if (have_bits) atomicOr(&spb[q >> 5], 1u << (q & 31));
            else break;
        }
    }
    atomicMin(&sfirst, myfirst);
    __syncthreads();
    const int first = sfirst;
    const double *dsr = ds;
    if (n <= ds_cap) {
        for (int i = tid; i < n; i += NT) dsl[i] = ds[i];
        __syncthreads();
        dsr = dsl;
    }
    for (int q = tid; q < first; q += NT) {
        int p = nsl[q];
        W.dlamda[s + q] = ds[p]; W.w[s + q] = zs[p]; indxp[q] = indx[p];
    }
    // ---- 3. sequential continuation over survivors, staged through LDS in chunks
    int K = first, nrot = 0, PJ = 0, ctP = 0, nrun = 0;
    bool mod = false;  // PJ's values changed by a rotation
    double zP = 0.0, DP = 0.0;
    if (tid == 0 && !sfallback) {
        int p0 = nsl[first];
        PJ = indx[p0];
        zP = zs[p0];
        DP = ds[p0];
        ctP = coltyp[PJ - 1];
    }
    if (!sfallback) {
        for (int c0 = first + 1; c0 < L; c0 += DCH) {
            const int cl = min(DCH, L - c0);
            __syncthreads();
            // a survivor's type is still its initial one when its chunk is staged: only
            // earlier survivors have been changed
            for (int u = tid; u < cl; u += NT) {
                int r = nsl[c0 + u];
                int nj = indx[r];
                ci[u] = nj;
                cz[u] = zs[r];
                cd[u] = ds[r];
                cty[u] = coltyp[nj - 1];
            }
            __syncthreads();
            if (tid == 0) {
                for (int u = 0; u < cl; u++) {
                    const int q = c0 + u;
                    if (!mod && have_bits && nrun < RUNS) {
                        // PJ = survivor q-1 unmodified: pairs up to the next speculative
                        // deflation keep their left survivor as it is; record them as a run
                        // b: first deflating pair (b, b+1) with b >= q-1, or lim + 1
                        int b = q - 1;
                        const int lim = c0 + cl - 1;
                        for (;;) {
                            unsigned int wd = spb[b >> 5] & (~0u << (b & 31));
                            if (wd) { b = (b & ~31) + __ffs(wd) - 1; break; }
                            b = (b | 31) + 1;
                            if (b > lim) break;
                        }
                        if (b > lim) b = lim + 1;
                        const int len = b - (q - 1);  // kept: survivors q-1 .. b-1
                        if (len > 0) {
                            run_k[nrun] = K; run_q[nrun] = q - 1; run_n[nrun] = len;
                            nrun++;
                            K += len;
                            // PJ becomes survivor b, unmodified (in this chunk or the next)
                            if (b <= lim) {
                                const int ub = b - c0;
                                PJ = ci[ub]; zP = cz[ub]; DP = cd[ub]; ctP = cty[ub];
                                u = ub;  // the loop's u++ makes the next pair (b, b+1)
                            } else {
                                const int ub = lim - c0;
                                PJ = ci[ub]; zP = cz[ub]; DP = cd[ub]; ctP = cty[ub];
                                u = cl;  // the rest of this chunk is in the run
                                // b = lim + 1 is the first survivor of the next chunk; the
                                // run ends at lim: survivor lim is PJ now, not kept yet
                                run_n[nrun - 1] = len - 1;
                                K -= 1;
                            }
                            continue;
                        }
                    }
                    int NJ = ci[u], ctN = cty[u];
                    double zN = cz[u], dN = cd[u];
                    double ss = zP, cc = zN;
                    double tau = lapy2_d(cc, ss);
                    double t = dN - DP;
                    cc = cc / tau;
                    ss = -ss / tau;
                    if (fabs(t * cc * ss) <= tol) {
                        if (ctN != ctP) { coltyp[NJ - 1] = 2; ctN = 2; }
                        coltyp[PJ - 1] = 4;
                        double tP = DP * (cc * cc) + dN * (ss * ss);
                        double dNn = DP * (ss * ss) + dN * (cc * cc);
                        W.rot_pj[s + nrot] = PJ; W.rot_nj[s + nrot] = NJ;
                        W.rot_c[s + nrot] = cc; W.rot_s[s + nrot] = ss;
                        W.evq[s + nrot] = q; W.evd[s + nrot] = tP;
                        nrot++;
                        PJ = NJ; zP = tau; DP = dNn; ctP = ctN;
                        mod = true;
                    } else {
                        W.dlamda[s + K] = DP; W.w[s + K] = zP; indxp[K] = PJ;
                        K++;
                        PJ = NJ; zP = zN; DP = dN; ctP = ctN;
                        mod = false;
                    }
                }
            }
        }
        if (tid == 0) { snr = nrot; snrun = nrun; }
        __syncthreads();
        // kept survivors of the runs, in parallel: run r keeps survivors run_q + o at
        // positions run_k + o, o < run_n (run_o: prefix of run_n)
        if (tid == 0) {
            int a = 0;
            for (int r = 0; r < snrun; r++) { run_o[r] = a; a += run_n[r]; }
            run_o[snrun] = a;
        }
        __syncthreads();
        {
            const int nr = snrun, tot = run_o[snrun];
#pragma unroll 4
            for (int x = tid; x < tot; x += NT) {
                int lo = 0, hi = nr - 1;
                while (lo < hi) { int md2 = (lo + hi + 1) >> 1; if (run_o[md2] <= x) lo = md2; else hi = md2 - 1; }
                const int o = x - run_o[lo], kk = run_k[lo] + o, j = nsl[run_q[lo] + o];
                W.dlamda[s + kk] = ds[j]; W.w[s + kk] = zs[j]; indxp[kk] = indx[j];
            }
        }
        // survivors and rotation events into LDS for the replay below
        if (L <= DCH)
            for (int u = tid; u < L; u += NT) lsn[u] = nsl[u];
        if (snr <= DCH)
            for (int u = tid; u < snr; u += NT) { cty[u] = W.evq[s + u]; cz[u] = W.evd[s + u]; ci[u] = W.rot_pj[s + u]; }
        __syncthreads();
        // with many survivors, the positions q (1 <= q < L) with small-z columns between
        // survivor q-1 and q, in ascending order: the replay then visits only these and
        // the rotation events (a walk over every survivor costs more only above a few
        // hundred)
        if (snr != 0 && L > 256) {
            const int Lq = L - 1, gper = (Lq + NW - 1) / NW;
            const int g0 = 1 + min(Lq, wv * gper), g1 = 1 + min(Lq, (wv + 1) * gper);
            int gc = 0;
            for (int q0 = g0; q0 < g1; q0 += 64) {
                const int q = q0 + lane;
                const bool gp = q < g1 && nsl[q] > nsl[q - 1] + 1;
                gc += __popcll(__ballot(gp));
            }
            if (lane == 0) wcnt[1][wv] = gc;
            __syncthreads();
            int gb = 0, gt = 0;
            for (int w = 0; w < NW; w++) {
                if (w < wv) gb += wcnt[1][w];
                gt += wcnt[1][w];
            }
            for (int q0 = g0; q0 < g1; q0 += 64) {
                const int q = q0 + lane;
                const bool gp = q < g1 && nsl[q] > nsl[q - 1] + 1;
                const unsigned long long m = __ballot(gp);
                const int at = gb + __popcll(m & lt);
                if (gp && at < MAXSEG) gq[at] = q;
                gb += __popcll(m);
            }
            if (tid == 0) sng = gt;
            __syncthreads();
        }
    }
    if (tid == 0 && !sfallback) {
        W.dlamda[s + K] = DP; W.w[s + K] = zP; indxp[K] = PJ;
        K++;
        sk = K; sk2 = nrot;
        if (nrot == 0) sfallback = 2;  // no rotation: the region is the small-z columns reversed
        if (nrot != 0) {
        // ---- 4. replay the deflated region as a list of segments, front first.
        // Segment: small-z run of sorted positions [a, b) (front holds b-1) or,
        // with a = -1, one rotated column b with value d.
        int nseg = 0, head = -1;
        auto push_run = [&](int a, int b) {
            if (a >= b) return;
            if (nseg >= MAXSEG) { nseg = MAXSEG + 1; return; }
            seg_a[nseg] = a; seg_b[nseg] = b; seg_nx[nseg] = head; head = nseg++;
        };
        int ev = 0;
        const int *nsr = (L <= DCH) ? lsn : nsl;
        const bool evl = nrot <= DCH;
        push_run(0, nsr[0]);
        // survivor positions in ascending order, each with the run before it, then its
        // rotation events; with the gap list (L > 256) only the positions that have either
        const bool sparse = L > 256;
        const int ng = sparse ? sng : 0;
        if (ng > MAXSEG) nseg = MAXSEG + 1;
        int gi = 0, qn = 1;
        while (nseg <= MAXSEG) {
            int q = qn++;
            if (sparse) {
                const int qg = gi < ng ? gq[gi] : L;
                const int qe = ev < nrot ? (evl ? cty[ev] : W.evq[s + ev]) : L;
                q = min(qg, qe);
            }
            if (q >= L) break;
            if (!sparse || (gi < ng && gq[gi] == q)) {
                push_run(nsr[q - 1] + 1, nsr[q]);
                gi++;
            }
            while (ev < nrot && (evl ? cty[ev] : W.evq[s + ev]) == q && nseg <= MAXSEG) {
                double dx = evl ? cz[ev] : W.evd[s + ev];
                int x = evl ? ci[ev] : W.rot_pj[s + ev];
                // walk from the front past every element with D > dx
                int prev = -1, cseg = head, split_t = -1;
                while (cseg >= 0) {
                    if (seg_a[cseg] < 0) {
                        if (dx < seg_d[cseg]) { prev = cseg; cseg = seg_nx[cseg]; continue; }
                        break;
                    }
                    int a = seg_a[cseg], b = seg_b[cseg];
                    // front order ds[b-1] >= ds[b-2] >= ...; t = first offset with ds <= dx
                    int lo = 0, hi = b - a;
                    while (lo < hi) { int md2 = (lo + hi) >> 1; if (dsr[b - 1 - md2] <= dx) hi = md2; else lo = md2 + 1; }
                    if (lo == b - a) { prev = cseg; cseg = seg_nx[cseg]; continue; }
                    split_t = lo;
                    break;
                }
                if (nseg + 2 > MAXSEG) { nseg = MAXSEG + 1; break; }
                int xs = nseg++;
                seg_a[xs] = -1; seg_b[xs] = x; seg_d[xs] = dx;
                if (split_t > 0) {
                    // cseg [a,b) -> front part [b-t, b), x, rest [a, b-t)
                    int rest = nseg++;
                    int a = seg_a[cseg], b = seg_b[cseg];
                    seg_a[rest] = a; seg_b[rest] = b - split_t; seg_nx[rest] = seg_nx[cseg];
                    seg_a[cseg] = b - split_t;
                    seg_nx[cseg] = xs;
                    seg_nx[xs] = rest;
                } else {
                    seg_nx[xs] = cseg;
                    if (prev < 0) head = xs; else seg_nx[prev] = xs;
                }
                ev++;
            }
        }
        if (nseg <= MAXSEG) push_run(nsr[L - 1] + 1, n);
        if (nseg > MAXSEG) {
            sfallback = 1;
        } else {
            // front-to-back order and region positions (INDXP(K+1..N) from the front)
            int pos = K, cnt2 = 0;
            for (int g = head; g >= 0; g = seg_nx[g]) {
                seg_ord[cnt2] = g;
                seg_pos[cnt2] = pos;
                pos += (seg_a[g] < 0) ? 1 : (seg_b[g] - seg_a[g]);
                cnt2++;
            }
            snseg = cnt2;
        }
        }
    }
    __syncthreads();
    if (sfallback == 2) {
        // no rotation: the r-th small-z position goes to INDXP(n-1-r)
#pragma unroll 8
        for (int q = sL + tid; q < n; q += NT) {
            const int j = nsl[q];
            indxp[q] = indx[j];
            ddef[q] = ds[j];
        }
    } else if (!sfallback) {
        // every position of the region at once: its segment by binary search over
        // the segment start positions (LDS), then its source
        const int K0 = sk, ns = snseg;
#pragma unroll 4
        for (int p = K0 + tid; p < n; p += NT) {
            int lo = 0, hi = ns - 1;
            while (lo < hi) { int md2 = (lo + hi + 1) >> 1; if (seg_pos[md2] <= p) lo = md2; else hi = md2 - 1; }
            const int g = seg_ord[lo];
            const int a = seg_a[g], b = seg_b[g];
            if (a < 0) {
                indxp[p] = b;
                ddef[p] = seg_d[g];
            } else {
                const int j = b - 1 - (p - seg_pos[lo]);
                indxp[p] = indx[j];
                ddef[p] = ds[j];
            }
        }
    } else {
        // ---- fallback: the reference's scan over every column in sorted order
        for (int i = tid; i < n; i += NT) coltyp[i] = (i < n1) ? 1 : 3;
        int K2 = n + 1, PJ = 0, have = 0, nrot = 0, KK = 0;
        double zP = 0.0, DP = 0.0;
        for (int c0 = 0; c0 < n; c0 += DCH) {
            int cl = min(DCH, n - c0);
            __syncthreads();
            for (int j = tid; j < cl; j += NT) { ci[j] = indx[c0 + j]; cz[j] = zs[c0 + j]; cd[j] = ds[c0 + j]; }
            __syncthreads();
            if (tid == 0) {
                for (int j = 0; j < cl; j++) {
                    int NJ = ci[j];
                    double zN = cz[j], dN = cd[j];
                    if (rho * fabs(zN) <= tol) {
                        K2--; coltyp[NJ - 1] = 4; indxp[K2 - 1] = NJ; ddef[K2 - 1] = dN;
                        continue;
                    }
                    if (!have) { PJ = NJ; zP = zN; DP = dN; have = 1; continue; }
                    double ss = zP, cc = zN;
                    double tau = lapy2_d(cc, ss);
                    double t = dN - DP;
                    cc = cc / tau;
                    ss = -ss / tau;
                    if (fabs(t * cc * ss) <= tol) {
                        if (coltyp[NJ - 1] != coltyp[PJ - 1]) coltyp[NJ - 1] = 2;
                        coltyp[PJ - 1] = 4;
                        W.rot_pj[s + nrot] = PJ; W.rot_nj[s + nrot] = NJ;
                        W.rot_c[s + nrot] = cc; W.rot_s[s + nrot] = ss;
                        nrot++;
                        double tP = DP * (cc * cc) + dN * (ss * ss);
                        double dNn = DP * (ss * ss) + dN * (cc * cc);
                        K2--;
                        int i = 1;
                        for (;;) {
                            if (K2 + i <= n) {
                                if (tP < ddef[K2 + i - 1]) {
                                    indxp[K2 + i - 2] = indxp[K2 + i - 1]; ddef[K2 + i - 2] = ddef[K2 + i - 1];
                                    indxp[K2 + i - 1] = PJ; ddef[K2 + i - 1] = tP;
                                    i++;
                                    continue;
                                }
                                indxp[K2 + i - 2] = PJ; ddef[K2 + i - 2] = tP;
                            } else {
                                indxp[K2 + i - 2] = PJ; ddef[K2 + i - 2] = tP;
                            }
                            break;
                        }
                        PJ = NJ; zP = tau; DP = dNn;
                    } else {
                        KK++;
                        W.dlamda[s + KK - 1] = DP; W.w[s + KK - 1] = zP; indxp[KK - 1] = PJ;
                        PJ = NJ; zP = zN; DP = dN;
                    }
                }
            }
        }
        if (tid == 0) {
            KK++;
            W.dlamda[s + KK - 1] = DP; W.w[s + KK - 1] = zP; indxp[KK - 1] = PJ;
            sk = KK; sk2 = nrot;
        }
    }
    __syncthreads();
    // CTOT and the type grouping (mdlaed2.f90:393-431) as a stable partition
    int c4[4] = {0, 0, 0, 0};
    for (int j0 = w0; j0 < w1; j0 += 64 * DU) {
        int js[DU], ct[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            const int j = j0 + 64 * u + lane;
            js[u] = (j < w1) ? indxp[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < DU; u++) ct[u] = js[u] ? coltyp[js[u] - 1] - 1 : -1;
#pragma unroll
        for (int u = 0; u < DU; u++)
            for (int t = 0; t < 4; t++) c4[t] += __popcll(__ballot(ct[u] == t));
    }
    if (lane == 0)
        for (int t = 0; t < 4; t++) wcnt[t][wv] = c4[t];
    __syncthreads();
    if (tid < 4) {
        int a = 0;
        for (int w = 0; w < NW; w++) a += wcnt[tid][w];
        scnt[tid] = a;
    }
    __syncthreads();
    const int KF = n - scnt[3];
    const int hss = (KF >= hss_min_k);
    if (!hss) {
        int pos[4];
        pos[0] = 0; pos[1] = scnt[0]; pos[2] = pos[1] + scnt[1]; pos[3] = pos[2] + scnt[2];
        for (int t = 0; t < 4; t++)
            for (int w = 0; w < wv; w++) pos[t] += wcnt[t][w];
        for (int j0 = w0; j0 < w1; j0 += 64 * DU) {
            int js[DU], ct[DU];
#pragma unroll
            for (int u = 0; u < DU; u++) {
                const int j = j0 + 64 * u + lane;
                js[u] = (j < w1) ? indxp[j] : 0;
            }
#pragma unroll
            for (int u = 0; u < DU; u++) ct[u] = js[u] ? coltyp[js[u] - 1] - 1 : -1;
#pragma unroll
            for (int u = 0; u < DU; u++) {
                unsigned long long mk[4];
                for (int t = 0; t < 4; t++) mk[t] = __ballot(ct[u] == t);
                if (ct[u] >= 0) {
                    int p = pos[ct[u]] + __popcll(mk[ct[u]] & lt);
                    indx[p] = js[u];
                    W.indxc[s + p] = j0 + 64 * u + lane + 1;
                }
                for (int t = 0; t < 4; t++) pos[t] += __popcll(mk[t]);
            }
        }
    } else {
#pragma unroll 8
        for (int j = tid; j < n; j += NT) indx[j] = indxp[j];
    }
    if (tid == 0) {
        o->k = KF;
        for (int t = 0; t < 4; t++) o->ctot[t] = scnt[t];
        o->hss = hss; o->nrot = sk2; o->flags = 0; o->rho = rho; o->tol = tol;
    }
}

template __global__ void k_deflate<256>(const Merge *, Work, int, MergeOut *, int, int);
template __global__ void k_deflate<1024>(const Merge *, Work, int, MergeOut *, int, int);

// Zero the rows of the merge block outside the support of every column that a
// Givens rotation touches, then widen its support to the block.
__global__ __launch_bounds__(WG) void k_rot_prep(const Merge *mg, const Prob *probs, Work W, int cur,
                                                 const MergeOut *mo)
{
    if (gate_closed(W, mg)) return;
    Merge M = mg[blockIdx.x];
    int nrot = mo[blockIdx.x].nrot;
    if (nrot == 0 || mo[blockIdx.x].flags) return;
    Prob P = probs[M.prob];
    for (int r = 0; r < 2 * nrot; r++) {
        int j = (r & 1) ? W.rot_nj[M.s + (r >> 1)] : W.rot_pj[M.s + (r >> 1)];
        int slot = W.vm[cur][M.s + j - 1];
        int lo = W.supp_lo[slot], hi = W.supp_hi[slot];
        __syncthreads();
        if (lo == M.s && hi == M.s + M.n) continue;
        double *col = P.q + (long)(slot - P.off) * P.ld - P.off;
        for (int row = M.s + threadIdx.x; row < M.s + M.n; row += WG)
            if (row < lo || row >= hi) col[row] = 0.0;
        __syncthreads();
        if (threadIdx.x == 0) { W.supp_lo[slot] = M.s; W.supp_hi[slot] = M.s + M.n; }
        __syncthreads();
    }
}

// DROT of every recorded pair in order (mdlaed2.f90:352); rows are independent.
__global__ __launch_bounds__(WG) void k_rotate(const Merge *mg, const int *chunk_start, int nm,
                                               const Prob *probs, Work W, int cur, const MergeOut *mo)
{
    if (gate_closed(W, mg)) return;
    int m = find_seg(chunk_start, nm, blockIdx.x);
    Merge M = mg[m];
    int nrot = mo[m].nrot;
    if (nrot == 0 || mo[m].flags) return;
    int i = (blockIdx.x - chunk_start[m]) * WG + threadIdx.x;
    if (i >= M.n) return;
    Prob P = probs[M.prob];
    long row = M.s + i - P.off;
    for (int r = 0; r < nrot; r++) {
        int ps = W.vm[cur][M.s + W.rot_pj[M.s + r] - 1], ns = W.vm[cur][M.s + W.rot_nj[M.s + r] - 1];
        double c = W.rot_c[M.s + r], sn = W.rot_s[M.s + r];
        double *xp = P.q + row + (long)(ps - P.off) * P.ld;
        double *yp = P.q + row + (long)(ns - P.off) * P.ld;
        double x = *xp, y = *yp;
        double tmp = c * x + sn * y;
        *yp = c * y - sn * x;
        *xp = tmp;
    }
}

// ---------------------------------------------------------------------------
// Secular equation (replaces DLAED4, mdlaed3.f90:208 / DLAED3): the rational
// "middle way" iteration with a bracketing safeguard, statement for statement as
// the oracle's orc_laed4.  A team of T = sec_team(K) lanes shares one root: lane l
// sums the poles i = l, l+T, ... (psi ascending, phi descending) and the team
// combines its partial sums by a butterfly, so every lane of the team holds the
// same iterate and takes the same branches.  A block holds SEC_WG / T roots of one
// merge; merges with K <= SEC_LDS read poles and z from LDS.
// ---------------------------------------------------------------------------
struct SecOut { double lam, tau, dl0, dl1, tc; int org, nit, info, used; };
struct SecInit { double t, lo, hi, wb; int ii; };

// the closed forms of K = 1 and K = 2 (LAPACK DLAED5 for K = 2)
__device__ SecOut sec_small(int n, int j, const double *d, const double *z, double rho)
{
    SecOut o;
    o.dl0 = o.dl1 = 0.0; o.info = 0; o.nit = 0; o.used = 0; o.tc = 0.0;
    if (n == 1) {
        o.lam = d[0] + rho * z[0] * z[0]; o.org = 1; o.tau = rho * z[0] * z[0];
        return o;
    }
    double del = d[1] - d[0], b, c, w, t, dl0, dl1;
    if (j == 1) {
        w = 1.0 + 2.0 * rho * (z[1] * z[1] - z[0] * z[0]) / del;
        if (w > 0.0) {
            b = del + rho * (z[0] * z[0] + z[1] * z[1]);
            c = rho * z[0] * z[0] * del;
            t = 2.0 * c / (b + sqrt(fabs(b * b - 4.0 * c)));
            o.lam = d[0] + t; o.org = 1; o.tau = t;
            dl0 = -z[0] / t; dl1 = z[1] / (del - t);
        } else {
            b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
            c = rho * z[1] * z[1] * del;
            if (b > 0.0) t = -2.0 * c / (b + sqrt(b * b + 4.0 * c));
            else t = (b - sqrt(b * b + 4.0 * c)) / 2.0;
            o.lam = d[1] + t; o.org = 2; o.tau = t;
            dl0 = -z[0] / (del + t); dl1 = -z[1] / t;
        }
    } else {
        b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
        c = rho * z[1] * z[1] * del;
        if (b > 0.0) t = (b + sqrt(b * b + 4.0 * c)) / 2.0;
        else t = 2.0 * c / (-b + sqrt(b * b + 4.0 * c));
        o.lam = d[1] + t; o.org = 2; o.tau = t;
        dl0 = -z[0] / (del + t); dl1 = -z[1] / t;
    }
    double sn = sqrt(dl0 * dl0 + dl1 * dl1);
    o.dl0 = dl0 / sn; o.dl1 = dl1 / sn;
    return o;
}

// classical starting iterate (n >= 3): the two nearest poles exactly, the others
// through their value at the bracket midpoint; origin pole and bracket
__device__ SecInit sec_init(int n, int j, const double *d, const double *z, double rho, int T, int l)
{
    SecInit q;
    const double rhoinv = 1.0 / rho;
    double t;
    if (j == n) {
        double zz = 0.0;
        for (int i = l; i < n; i += T) zz += z[i] * z[i];
        zz = team_sum(zz, T);
        double hh = rho * zz, mid = 0.5 * hh, dn = d[n - 1];
        double cs = 0.0;
        for (int i = l; i < n - 2; i += T) cs += z[i] * z[i] / ((d[i] - dn) - mid);
        double c = rhoinv + team_sum(cs, T);
        double dl = dn - d[n - 2];
        double z1 = z[n - 2] * z[n - 2], z2 = z[n - 1] * z[n - 1];
        if (c <= 0.0) {
            t = hh;
        } else {
            double b = c * dl - z1 - z2;
            double sq = sqrt(fabs(b * b + 4.0 * c * z2 * dl));
            if (b <= 0.0) t = (sq - b) / (2.0 * c);
            else t = 2.0 * z2 * dl / (b + sq);
        }
        if (!(t > 0.0 && t <= hh)) t = mid;
        q.ii = n; q.lo = 0.0; q.hi = hh; q.wb = hh;
    } else {
        double del = d[j] - d[j - 1], mid = 0.5 * del;
        q.wb = del;
        double cs = 0.0;
        for (int i = l; i < n; i += T) {
            if (i == j - 1 || i == j) continue;
            cs += z[i] * z[i] / ((d[i] - d[j - 1]) - mid);
        }
        double c = rhoinv + team_sum(cs, T);
        double zl = z[j - 1] * z[j - 1], zr = z[j] * z[j];
        double fmid = c + (zr - zl) / mid;
        if (fmid >= 0.0) {
            double b = c * del + zl + zr, cc = zl * del;
            double sq = sqrt(fabs(b * b - 4.0 * c * cc));
            if (b > 0.0) t = 2.0 * cc / (b + sq);
            else t = (b - sq) / (2.0 * c);
            if (!(t > 0.0 && t <= mid)) t = 0.5 * mid;
            q.ii = j; q.lo = 0.0; q.hi = mid;
        } else {
            double b = c * del - zl - zr, cc = zr * del;
            double sq = sqrt(fabs(b * b + 4.0 * c * cc));
            if (b <= 0.0) t = -2.0 * cc / (sq - b);
            else t = -(b + sq) / (2.0 * c);
            if (!(t >= -mid && t < 0.0)) t = -0.5 * mid;
            q.ii = j + 1; q.lo = -mid; q.hi = 0.0;
        }
    }
    q.t = t;
    return q;
}

// the iteration from a starting iterate inside the bracket of q
__device__ SecOut sec_iterate(int n, int j, const double *d, const double *z, double rho, int maxit, SecInit q,
                              int T, int l)
{
    SecOut o;
    o.dl0 = o.dl1 = 0.0; o.used = 0; o.tc = 0.0;
    const double rhoinv = 1.0 / rho;
    const int ii = q.ii;
    double t = q.t, lo = q.lo, hi = q.hi;
    // rounding of the team's butterfly: log2(T) more additions per sum
    int lgT = 0;
    while ((1 << lgT) < T) lgT++;
    const double dorg = d[ii - 1];
    const int split = (j == n) ? n - 1 : j;
    int it, info = 1;
    for (it = 0; it <= maxit; it++) {
        // psi over poles left of the root (ascending), phi over poles right of it
        // (descending), the origin pole excluded (it is the last psi or the last phi
        // pole); the two sums advance in the same loop for instruction-level
        // parallelism, each in its own fixed order
        double psi = 0.0, dpsi = 0.0, epsi = 0.0, phi = 0.0, dphi = 0.0, ephi = 0.0;
        const int na = split - ((ii - 1 < split) ? 1 : 0);          // psi: i = 0 .. na-1
        const int nb = (n - split) - ((ii - 1 >= split) ? 1 : 0);   // phi: i = n-1 .. n-nb
        const int nab = na > nb ? na : nb;
#pragma unroll 4
        for (int t2 = l; t2 < nab; t2 += T) {
            if (t2 < na) {
                double dl = (d[t2] - dorg) - t;
                double qq = z[t2] / dl;
                psi += z[t2] * qq;
                dpsi += qq * qq;
                epsi += psi;
            }
            if (t2 < nb) {
                int i = n - 1 - t2;
                double dl = (d[i] - dorg) - t;
                double qq = z[i] / dl;
                phi += z[i] * qq;
                dphi += qq * qq;
                ephi += phi;
            }
        }
        psi = team_sum(psi, T); dpsi = team_sum(dpsi, T); epsi = team_sum(epsi, T);
        phi = team_sum(phi, T); dphi = team_sum(dphi, T); ephi = team_sum(ephi, T);
        double qo = z[ii - 1] / (-t);
        double torg = z[ii - 1] * qo;
        double w = ((rhoinv + phi) + psi) + torg;
        double erretm = (8.0 + lgT) * (phi - psi) + (ephi - epsi) + 2.0 * rhoinv + 3.0 * fabs(torg) +
                        fabs(t) * ((dpsi + dphi) + qo * qo);
        if (ii - 1 < split) dpsi += qo * qo; else dphi += qo * qo;
        if (fabs(w) <= HSDC_EPS * erretm || it == maxit) {
            if (fabs(w) <= HSDC_EPS * erretm) info = 0;
            break;
        }
        if (w > 0.0) hi = t; else lo = t;
        double dl = (d[split - 1] - dorg) - t;
        double dr = (d[split] - dorg) - t;
        double a = (dl + dr) * w - dl * dr * (dpsi + dphi);
        double b = dl * dr * w;
        double c = w - dl * dpsi - dr * dphi;
        double eta;
        if (c == 0.0) {
            eta = (a != 0.0) ? b / a : -w / (dpsi + dphi);
        } else {
            double sq = sqrt(fabs(a * a - 4.0 * b * c));
            if (a <= 0.0) eta = (a - sq) / (2.0 * c);
            else eta = 2.0 * b / (a + sq);
        }
        if (w * eta >= 0.0) eta = -w / (dpsi + dphi);
        double tn = t + eta;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t) { info = 0; break; }
        t = tn;
    }
    o.nit = it + 1;
    o.lam = dorg + t; o.org = ii; o.tau = t; o.info = info;
    return o;
}

// block -> merge, root and team geometry of both secular kernels
struct SecGeom { int m, k, s, T, l, j; };
__device__ __forceinline__ SecGeom sec_geom(const Merge *mg, const int *block_start, int nm, const MergeOut *mo,
                                            int vb)
{
    SecGeom g;
    g.m = find_seg(block_start, nm, vb);
    g.k = mo[g.m].flags ? 0 : mo[g.m].k;
    g.s = mg[g.m].s;
    g.T = sec_team(g.k > 0 ? g.k : 1);
    g.l = threadIdx.x & (g.T - 1);
    g.j = (vb - block_start[g.m]) * (SEC_WG / g.T) + threadIdx.x / g.T + 1;
    return g;
}

// Learned starting iterates (merges with K >= max(3, sec->kmin), when weights are
// loaded or samples are harvested): the classical start, the 34 features, the
// network's proposal behind the gate; written per root for k_secular.
__device__ void secstart_block(const Merge *mg, const int *block_start, int nm, const MergeOut *mo, const Work &W,
                               const SecCtl *sec, int vb)
{
    const SecGeom g = sec_geom(mg, block_start, nm, mo, vb);
    const SecCtl C = *sec;
    if (g.k < 3 || g.k < C.kmin || g.j > g.k) return;
    const double *d = W.dlamda + g.s, *z = W.w + g.s;
    const double rho = mo[g.m].rho;
    SecInit q = sec_init(g.k, g.j, d, z, rho, g.T, g.l);
    const double tc = q.t;
    double xc = (q.ii == g.j) ? q.t / q.wb : 1.0 + q.t / q.wb;
    float fl[MLP_IN];
    sec_features(g.k, g.j - 1, d, z, rho, q.wb, xc, fl);
    int used = 0;
    if (C.w && g.T >= 8) {
        float y = mlp_forward_team(C.w, fl, g.T, g.l);
        y = y < -16.0f ? -16.0f : (y > 16.0f ? 16.0f : y);
        double tl = q.t * sec_exp((double)y);
        if (fabsf(y) > C.gate && tl > q.lo && tl < q.hi) { q.t = tl; used = 1; }
    }
    if (g.l == 0) {
        const int r = g.s + g.j - 1;
        W.st_t[r] = q.t; W.st_tc[r] = tc; W.st_lo[r] = q.lo; W.st_hi[r] = q.hi;
        W.st_ii[r] = q.ii | (used << 30);
        if (C.hf)
            for (int f = 0; f < MLP_IN; f++) C.feat[(long)r * MLP_IN + f] = fl[f];
    }
}

// virtual block vb of the level's secular launch (block_start[nm] of them)
__global__ __launch_bounds__(SEC_WG) void k_secstart(const Merge *mg, const int *block_start, int nm,
                                                     const MergeOut *mo, Work W, const SecCtl *sec)
{
    if (gate_closed(W, mg)) return;
    const int nvb = block_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x)
        secstart_block(mg, block_start, nm, mo, W, sec, vb);
}

// Root teams.  Merges that k_secstart covered start from its iterate; the others
// from the classical one.  sec->hf/ht: when non-null, (features, log(converged tau /
// classical starting tau)) of every root k_secstart covered are appended for
// training, up to sec->cap samples.  block_start[m]: first block of merge m
// (SEC_WG / sec_team(K) roots per block).
struct SecAcc {  // statistics of one wavefront (held by its lane 0)
    unsigned long long iters, used, wave_iters, evals;
    unsigned int itmax;
};

__device__ void secular_block(const Merge *mg, const int *block_start, int nm, const MergeOut *mo, const Work &W,
                              int maxit, int *info, const SecCtl *sec, SecAcc &acc, int vb, double *sd, double *sz)
{
    const SecGeom g = sec_geom(mg, block_start, nm, mo, vb);
    if (g.k == 0) return;  // uniform over the block
    const int k = g.k, s = g.s, T = g.T, l = g.l, j = g.j;
    const double *d = W.dlamda + s, *z = W.w + s;
    if (k <= SEC_LDS) {
        for (int i = threadIdx.x; i < k; i += SEC_WG) { sd[i] = d[i]; sz[i] = z[i]; }
        __syncthreads();
        d = sd;
        z = sz;
    }
    const SecCtl C = *sec;
    const double rho = mo[g.m].rho;
    int nit = 0, used = 0;
    if (j <= k) {
        SecOut r;
        const bool pre = k >= 3 && k >= C.kmin && (C.w || C.hf);
        if (k < 3) {
            r = sec_small(k, j, d, z, rho);
        } else if (pre) {
            const int rr = s + j - 1;
            SecInit q;
            q.t = W.st_t[rr]; q.lo = W.st_lo[rr]; q.hi = W.st_hi[rr]; q.wb = 0.0;
            q.ii = W.st_ii[rr] & ((1 << 30) - 1);
            r = sec_iterate(k, j, d, z, rho, maxit, q, T, l);
            r.used = W.st_ii[rr] >> 30;
            r.tc = W.st_tc[rr];
        } else {
            SecInit q = sec_init(k, j, d, z, rho, T, l);
            r = sec_iterate(k, j, d, z, rho, maxit, q, T, l);
            r.tc = q.t;
        }
        if (l == 0) {
            nit = r.nit;
            used = r.used;
            const double ratio = r.tau / r.tc;
            if (pre && C.hf && r.info == 0 && ratio > 0.0 && ratio < 1e300) {
                unsigned long long slot = atomicAdd(C.count, 1ULL);
                if (slot < C.cap) {
                    for (int q = 0; q < MLP_IN; q++)
                        C.hf[slot * MLP_IN + q] = C.feat[(long)(s + j - 1) * MLP_IN + q];
                    C.ht[slot] = (float)log(ratio);
                }
            }
            W.lam[s + j - 1] = r.lam;
            W.tau[s + j - 1] = r.tau;
            W.org[s + j - 1] = r.org;
            W.nit[s + j - 1] = r.nit;
            if (k == 2) { W.difl[s + j - 1] = r.dl0; W.difr[s + j - 1] = r.dl1; }
            if (r.info) atomicOr(info, 1);
        }
    }
    // statistics of the wavefront: total iterations, learned starts, and its largest
    // count (a wavefront runs as long as its slowest root, a launch as long as its
    // slowest wavefront)
    int mx = nit, sum = nit, nu = used;
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, __shfl_xor(mx, off));
        sum += __shfl_xor(sum, off);
        nu += __shfl_xor(nu, off);
    }
    acc.iters += (unsigned long long)sum;
    acc.used += (unsigned long long)nu;
    acc.wave_iters += (unsigned long long)mx;
    acc.evals += (unsigned long long)k * (unsigned long long)sum;  // pole evaluations: K per iteration
    acc.itmax = max(acc.itmax, (unsigned int)mx);
}

// The statistics leave a block as one atomic per counter, into one of SEC_SLOTS
// copies picked by the block (same-address atomics from every wavefront of a wide
// level serialise in L2); the host sums the copies (maximum for itmax).
__global__ __launch_bounds__(SEC_WG) void k_secular(const Merge *mg, const int *block_start, int nm,
                                                    const MergeOut *mo, Work W, int maxit, int *info,
                                                    const SecCtl *sec, unsigned int *itmax)
{
    if (gate_closed(W, mg)) return;
    __shared__ double sd[SEC_LDS], sz[SEC_LDS];
    __shared__ unsigned long long red[SEC_WG / 64][4];
    __shared__ unsigned int redm[SEC_WG / 64];
    SecAcc acc = {0, 0, 0, 0, 0};
    const int nvb = block_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        secular_block(mg, block_start, nm, mo, W, maxit, info, sec, acc, vb, sd, sz);
        __syncthreads();  // the next virtual block reloads the pole arrays
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = acc.iters; red[wv][1] = acc.used; red[wv][2] = acc.wave_iters; red[wv][3] = acc.evals;
        redm[wv] = acc.itmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        SecAcc t = {0, 0, 0, 0, 0};
        for (int w = 0; w < SEC_WG / 64; w++) {
            t.iters += red[w][0]; t.used += red[w][1]; t.wave_iters += red[w][2]; t.evals += red[w][3];
            t.itmax = max(t.itmax, redm[w]);
        }
        const SecCtl C = *sec;
        const int slot = blockIdx.x & (SEC_SLOTS - 1);
        if (t.iters) atomicAdd(C.iters + slot, t.iters);
        if (t.used) atomicAdd(C.used + slot, t.used);
        if (t.wave_iters) atomicAdd(C.wave_iters + slot, t.wave_iters);
        if (t.evals) atomicAdd(C.evals + slot, t.evals);
        if (t.itmax) atomicMax(itmax + slot, t.itmax);
    }
}

// delta_i(j) = d_i - lambda_j, formed as the secular solver forms it
__device__ __forceinline__ double delta_ij(const double *d, int i, const int *org, const double *tau, int j)
{
    return (d[i] - d[org[j] - 1]) - tau[j];
}

// Loewner z of the dense branch, LAPACK DLAED3 order: W(I) = DELTA_I(I) times
// DELTA_I(J)/(DLAMDA(I)-DLAMDA(J)) for J ascending, z = sign(sqrt(-W), w).
// Also the HSS branch (MDLAED3, mdlaed3.f90:207-230): product from 1 with the
// J = I factor taken in place, z = sign(sqrt|prod|, w).  One thread per I; the
// poles, tau and origins of a merge with K <= LW_LDS are staged in LDS first
// (every thread walks all K of them).
#define LW_LDS 2048
#define LW_U 8
__global__ __launch_bounds__(WG) void k_loewner(const Merge *mg, const int *chunk_start, int nm,
                                                const MergeOut *mo, Work W)
{
    if (gate_closed(W, mg)) return;
    // chunks of LW_LDS poles: d_j, d_org(j) and tau_j staged by the block, every
    // thread's product walks them in order
    __shared__ double sd[LW_LDS], so[LW_LDS], st[LW_LDS];
    const int nvb = chunk_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        const int m = find_seg(chunk_start, nm, vb);
        const MergeOut o = mo[m];
        if (o.flags || o.k < 3) continue;  // uniform over the block
        const int s = mg[m].s, K = o.k, i = (vb - chunk_start[m]) * WG + threadIdx.x;
        const double *d = W.dlamda + s, *tau = W.tau + s;
        const int *org = W.org + s;
        const bool own = i < K;
        const double di = own ? d[i] : 0.0;
        // delta_i(j) = (d_i - d_org(j)) - tau_j, as delta_ij forms it
        double wv = (own && !o.hss) ? delta_ij(d, i, org, tau, i) : 1.0;
        for (int c0 = 0; c0 < K; c0 += LW_LDS) {
            const int cn = min(LW_LDS, K - c0);
            for (int t = threadIdx.x; t < cn; t += WG) {
                sd[t] = d[c0 + t];
                so[t] = d[org[c0 + t] - 1];
                st[t] = tau[c0 + t];
            }
            __syncthreads();
            if (own) {
                // factor of pole J: the quotient, or at J = I nothing (dense: times 1.0,
                // which leaves the product's bits as they are) or the numerator (HSS).
                // LW_U quotients are formed independently, then multiplied into the
                // product in J order: a level launches one thread per root, too few
                // wavefronts to hide the divide and LDS latency of one pole at a time.
                // The divide's operands are selected, not its result (1.0 / 1.0 and
                // num / 1.0 are exact), so every pole runs the same straight-line code.
                const int me = i - c0;
                const bool hss = o.hss;
                int jj = 0;
                for (; jj + LW_U <= cn; jj += LW_U) {
                    double f[LW_U];
#pragma unroll
                    for (int u = 0; u < LW_U; u++) {
                        const bool self = jj + u == me;
                        const double num = (di - so[jj + u]) - st[jj + u], den = di - sd[jj + u];
                        f[u] = ((self && !hss) ? 1.0 : num) / (self ? 1.0 : den);
                    }
#pragma unroll
                    for (int u = 0; u < LW_U; u++) wv = wv * f[u];
                }
                for (; jj < cn; jj++) {
                    const bool self = jj == me;
                    const double num = (di - so[jj]) - st[jj], den = di - sd[jj];
                    wv = wv * (((self && !hss) ? 1.0 : num) / (self ? 1.0 : den));
                }
            }
            __syncthreads();  // the next chunk or virtual block restages
        }
        if (own) W.wnew[s + i] = o.hss ? copysign(sqrt(fabs(wv)), W.w[s + i]) : copysign(sqrt(-wv), W.w[s + i]);
    }
}

// Dense branch eigenvector matrix of the secular problem (DLAED3, after label
// 70): S(t,j) = (W(I)/DELTA_I(J)) / ||W./DELTA(:,J)||, I = INDXC(t); written
// row-major (S[t*K + j]) so one lane per column stores coalesced.  Poles, Loewner
// z and INDXC of a merge with K <= LW_LDS are staged in LDS.
__global__ __launch_bounds__(WG) void k_sbuild(const Merge *mg, const int *chunk_start, int nm,
                                               const MergeOut *mo, const long *soff, Work W, double *S)
{
    if (gate_closed(W, mg)) return;
    __shared__ double sd[LW_LDS], sw[LW_LDS];
    __shared__ int sc[LW_LDS];
    __shared__ double sdo[WG], stj[WG], stm[WG];  // per column of the block: d_org, tau, norm
    const int nvb = chunk_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        const int m = find_seg(chunk_start, nm, vb);
        const MergeOut o = mo[m];
        if (o.flags || o.hss || o.k == 0) continue;  // uniform over the block
        const int s = mg[m].s, K = o.k, j = (vb - chunk_start[m]) * WG + threadIdx.x;
        double *Sm = S + soff[m];
        const int *indxc = W.indxc + s;
        if (K <= 2) {
            if (j < K) {
                if (K == 1) {
                    Sm[0] = 1.0;
                } else {
                    double dv[2] = {W.difl[s + j], W.difr[s + j]};
                    for (int t = 0; t < 2; t++) Sm[t * K + j] = dv[indxc[t] - 1];
                }
            }
            continue;
        }
        const double *d = W.dlamda + s, *wz = W.wnew + s;
        if (K <= LW_LDS) {
            for (int t = threadIdx.x; t < K; t += WG) { sd[t] = d[t]; sw[t] = wz[t]; sc[t] = indxc[t]; }
            __syncthreads();
            d = sd; wz = sw; indxc = sc;
        }
        // column norms, one thread per column j (the sum runs over i in order);
        // delta_i(j) = (d_i - d_org(j)) - tau_j, as delta_ij forms it
        if (j < K) {
            const double dorg = d[W.org[s + j] - 1], tj = W.tau[s + j];
            double acc = 0.0;
            for (int i = 0; i < K; i++) {
                double v = wz[i] / ((d[i] - dorg) - tj);
                acc += v * v;
            }
            sdo[threadIdx.x] = dorg;
            stj[threadIdx.x] = tj;
            stm[threadIdx.x] = sqrt(acc);
        }
        __syncthreads();
        // the K x ncols entries, spread over all threads: WG / ncols rows per pass
        const int c0 = j - threadIdx.x, ncols = min(WG, K - c0), rpp = WG / ncols;
        const int jj = threadIdx.x % ncols, t0 = threadIdx.x / ncols;
        if (t0 < rpp) {
            const double dorg = sdo[jj], tj = stj[jj], temp = stm[jj];
            for (int t = t0; t < K; t += rpp) {
                int i = indxc[t] - 1;
                Sm[(long)t * K + c0 + jj] = (wz[i] / ((d[i] - dorg) - tj)) / temp;
            }
        }
        __syncthreads();  // the next virtual block restages
    }
}

// HSS branch scalars (mdlaed3.f90:215-249): DIFL, DIFR, alpha_j = 1/||col j||.
// One thread per column j; the poles and Loewner z are staged in chunks of LW_LDS.
__global__ __launch_bounds__(WG) void k_alpha(const Merge *mg, const int *chunk_start, int nm,
                                              const MergeOut *mo, Work W)
{
    if (gate_closed(W, mg)) return;
    __shared__ double sd[LW_LDS], sw[LW_LDS];
    const int nvb = chunk_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        const int m = find_seg(chunk_start, nm, vb);
        const MergeOut o = mo[m];
        if (o.flags || !o.hss) continue;  // uniform over the block
        const int s = mg[m].s, K = o.k, j = (vb - chunk_start[m]) * WG + threadIdx.x;
        const double *d = W.dlamda + s, *tau = W.tau + s, *wz = W.wnew + s;
        const int *org = W.org + s;
        const bool own = j < K;
        double difl = 0.0, difr = 0.0, dlamj = 0.0, difrj = 0.0, dlamjp = 0.0, wzj = 0.0;
        if (own) {
            difl = -delta_ij(d, j, org, tau, j);
            difr = (j + 1 < K) ? -delta_ij(d, j + 1, org, tau, j) : 0.0;
            W.difl[s + j] = difl;
            W.difr[s + j] = difr;
            dlamj = -d[j];
            if (j + 1 < K) { difrj = -difr; dlamjp = -d[j + 1]; }
            wzj = wz[j];
        }
        double acc = 0.0;
        for (int c0 = 0; c0 < K; c0 += LW_LDS) {
            const int cn = min(LW_LDS, K - c0);
            for (int t = threadIdx.x; t < cn; t += WG) { sd[t] = d[c0 + t]; sw[t] = wz[c0 + t]; }
            __syncthreads();
            // column entry i = j: wz_j / difl; i < j and i > j: the differences from
            // lambda_j's two neighbouring poles.  LW_U entries are formed
            // independently, then summed in i order (one thread per column, so few
            // wavefronts per level that one entry at a time leaves the latency bare).
            if (own) {
                int ii = 0;
                for (; ii + LW_U <= cn; ii += LW_U) {
                    double v[LW_U];
#pragma unroll
                    for (int u = 0; u < LW_U; u++) {
                        const int i = c0 + ii + u;
                        const double dlo = (sd[ii + u] + dlamj) - difl, dhi = (sd[ii + u] + dlamjp) + difrj;
                        v[u] = ((i == j) ? wzj : sw[ii + u]) / ((i == j) ? difl : (i < j) ? dlo : dhi);
                    }
#pragma unroll
                    for (int u = 0; u < LW_U; u++) acc += v[u] * v[u];
                }
                for (; ii < cn; ii++) {
                    const int i = c0 + ii;
                    const double den = (i == j) ? difl
                                     : (i < j) ? (sd[ii] + dlamj) - difl : (sd[ii] + dlamjp) + difrj;
                    const double v = ((i == j) ? wzj : sw[ii]) / den;
                    acc += v * v;
                }
            }
            __syncthreads();  // the next chunk or virtual block restages
        }
        if (own) W.alpha[s + j] = 1.0 / sqrt(acc);
    }
}

// ---------------------------------------------------------------------------
// Dense branch eigenvector update on FP64 MFMA (DLAED3's two DGEMMs):
//   rows of Q1: Ct(0:n1, :)  = Q2(0:n1, types 1,2) * S(types 1,2, :)
//   rows of Q2: Ct(n1:n, :)  = Q2(n1:n, types 2,3) * S(types 2,3, :)
// The A operand is gathered straight from the physical eigenvector columns
// (no packed Q2 copy), zero outside each column's support.
// Work item = one 64x64 output tile; 4 waves, each 32x32 as 2x2 MFMA 16x16x4.
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
struct GTile { int m, half, r0, c0; };

// Work item vt: tiles[vt] when the host built the list, else decoded from the
// level's tile table (per merge, the row tiles of the Q1 half then of the Q2 half,
// each across the ceil(K/64) column tiles).
__device__ __forceinline__ GTile gtile_at(const GTile *tiles, const int *tile_start, int nm, const Merge *mg,
                                          const MergeOut *mo, int vt)
{
    if (tiles) return tiles[vt];
    GTile T;
    T.m = find_seg(tile_start, nm, vt);
    const int loc = vt - tile_start[T.m];
    const int nc = (mo[T.m].k + 63) >> 6, rt1 = (mg[T.m].n1 + 63) >> 6;
    const int rt = loc / nc;
    T.c0 = (loc - rt * nc) * 64;
    if (rt < rt1) { T.half = 0; T.r0 = rt * 64; }
    else { T.half = 1; T.r0 = (rt - rt1) * 64; }
    return T;
}

__global__ __launch_bounds__(WG) void k_gemm(const Merge *mg, const GTile *tiles, int ntiles, const int *tile_start,
                                             int nm, const Prob *probs, const MergeOut *mo, const long *soff,
                                             const long *coff, Work W, int cur, const double *S, double *Ct)
{
    if (gate_closed(W, mg)) return;
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int total = tiles ? ntiles : tile_start[nm];
    for (int vt = blockIdx.x; vt < total; vt += gridDim.x) {
        GTile T = gtile_at(tiles, tile_start, nm, mg, mo, vt);
        Merge M = mg[T.m];
        const MergeOut o = mo[T.m];
        Prob P = probs[M.prob];
        const int K = o.k, n1 = M.n1, n2 = M.n - M.n1;
        const int nrow = T.half ? n2 : n1;
        const int rowg0 = M.s + (T.half ? n1 : 0);
        const int kdim = T.half ? (o.ctot[1] + o.ctot[2]) : (o.ctot[0] + o.ctot[1]);
        const int tbase = T.half ? o.ctot[0] : 0;  // first grouped column / S row used
        const double *Sm = S + soff[T.m];
        const int *indx = W.indx + M.s;
        const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
        const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
        d4 acc[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < kdim; k0 += 16) {
            for (int e = 0; e < 4; e++) {
                int idx = tid + WG * e;
                int r = idx & 63, kk = idx >> 6;
                double v = 0.0;
                int t = k0 + kk, row = T.r0 + r;
                if (t < kdim && row < nrow) {
                    int slot = W.vm[cur][M.s + indx[tbase + t] - 1];
                    v = qload(P, W.supp_lo, W.supp_hi, rowg0 + row, slot);
                }
                As[kk][r] = v;
                int c = idx & 63;
                double b = 0.0;
                int col = T.c0 + c;
                if (t < kdim && col < K) b = Sm[(long)(tbase + t) * K + col];
                Bs[kk][c] = b;
            }
            __syncthreads();
            for (int kq = 0; kq < 16; kq += 4) {
                int kl = kq + (lane >> 4);
                double a0 = As[kl][wr + (lane & 15)];
                double a1 = As[kl][wr + 16 + (lane & 15)];
                double b0 = Bs[kl][wc + (lane & 15)];
                double b1 = Bs[kl][wc + 16 + (lane & 15)];
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
            }
            __syncthreads();
        }
        double *C = Ct + coff[T.m];
        const int rowoff = T.half ? n1 : 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int i = 0; i < 4; i++) {
                    int row = T.r0 + wr + a * 16 + (lane >> 4) + 4 * i;
                    int col = T.c0 + wc + b * 16 + (lane & 15);
                    if (row < nrow && col < K) C[(long)(rowoff + row) + (long)col * M.n] = acc[a][b][i];
                }
    }
}

// ---------------------------------------------------------------------------
// New column bookkeeping of the merged block: virtual column v < K is root v
// (its vector from Ct), v >= K the deflated column INDXP(v) kept in place.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void newcols_at(const Merge *mg, const int *chunk_start, int nm, const MergeOut *mo,
                                           const Work &W, int cur, int b)
{
    int m = find_seg(chunk_start, nm, b);
    Merge M = mg[m];
    const MergeOut o = mo[m];
    int v = (b - chunk_start[m]) * WG + threadIdx.x;
    if (v >= M.n) return;
    int g = M.s + v, nxt = cur ^ 1;
    if (o.flags) {
        int src = W.indx[g];
        W.vm[nxt][g] = W.vm[cur][M.s + src - 1];
        W.Dv[nxt][g] = W.Dv[cur][M.s + src - 1];
        W.indxq[g] = v + 1;
        return;
    }
    int src = W.indxp[g];
    W.vm[nxt][g] = W.vm[cur][M.s + src - 1];
    const double dv = (v < o.k) ? W.lam[g] : W.ddef[g];
    W.Dv[nxt][g] = dv;
    // k_lamrg merges the roots (ascending) with the deflated values (read backwards);
    // a pair out of that order sends the merge to DLAMRG's sequential loop
    if (v + 1 < M.n && v + 1 != o.k) {
        const double nx = (v + 1 < o.k) ? W.lam[g + 1] : W.ddef[g + 1];
        if ((v < o.k) ? (dv > nx) : (dv < nx)) atomicOr(&W.maux[4 * m + 3], 1ULL);
    }
}

__global__ __launch_bounds__(WG) void k_newcols(const Merge *mg, const int *chunk_start, int nm,
                                                const MergeOut *mo, Work W, int cur)
{
    if (gate_closed(W, mg)) return;
    newcols_at(mg, chunk_start, nm, mo, W, cur, blockIdx.x);
}

// copy the K updated eigenvectors of a dense merge from Ct into their physical
// slots, and give every new vector (dense or HSS) the merge's row support;
// only = 1: the dense merges, 2: the HSS merges, 0: both
__device__ __forceinline__ void k_scatter_at(const Merge *mg, const int *col_start, int nm, const MergeOut *mo,
                                             const long *coff, const Prob *probs, const Work &W, int nxt,
                                             const double *Ct, int only, int vb)
{
    int m = find_seg(col_start, nm, vb);
    Merge M = mg[m];
    int v = vb - col_start[m];
    const bool hss = mo[m].hss;
    if (v >= mo[m].k || (only == 1 && hss) || (only == 2 && !hss)) return;
    Prob P = probs[M.prob];
    int slot = W.vm[nxt][M.s + v];
    if (!hss) {  // the HSS multiply writes its vectors into the slots itself
        const double *src = Ct + coff[m] + (long)v * M.n;
        double *dst = P.q + (long)(slot - P.off) * P.ld + (M.s - P.off);
        for (int r = threadIdx.x; r < M.n; r += WG) dst[r] = src[r];
    }
    if (threadIdx.x == 0) { W.supp_lo[slot] = M.s; W.supp_hi[slot] = M.s + M.n; }
}

__global__ __launch_bounds__(WG) void k_scatter(const Merge *mg, const int *col_start, int nm,
                                                const MergeOut *mo, const long *coff, const Prob *probs,
                                                Work W, int nxt, const double *Ct, int only)
{
    if (gate_closed(W, mg)) return;
    const int nvb = col_start[nm];
    for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x)
        k_scatter_at(mg, col_start, nm, mo, coff, probs, W, nxt, Ct, only, vb);
}

// The new column bookkeeping and, for the dense merges of a device-scheduled level,
// the copy of Ct into the slots, in one launch: blocks [0, nchunk) are k_newcols'
// chunks; the rest walk the elements of each merge's N x K block of Ct in chunks of
// SC_CH (sc_start: first chunk of every merge), so that the many short columns of
// the low levels do not each hold a block.  The copy finds the slot of new vector
// v as k_newcols does, vm[cur][INDXP(v)], without waiting for it.
__global__ __launch_bounds__(WG) void k_newcols_copy(const Merge *mg, const int *chunk_start, int nm, int nchunk,
                                                     const int *sc_start, const MergeOut *mo, const long *coff,
                                                     const Prob *probs, Work W, int cur, const double *Ct)
{
    if (gate_closed(W, mg)) return;
    if ((int)blockIdx.x < nchunk) {
        newcols_at(mg, chunk_start, nm, mo, W, cur, blockIdx.x);
        return;
    }
    const int nvb = sc_start[nm], g0 = blockIdx.x - nchunk, gs = gridDim.x - nchunk;
    for (int vb = g0; vb < nvb; vb += gs) {
        const int m = find_seg(sc_start, nm, vb);
        const Merge M = mg[m];
        const Prob P = probs[M.prob];
        const long ne = (long)M.n * mo[m].k, e0 = (long)(vb - sc_start[m]) * SC_CH;
        const double *src = Ct + coff[m];
        double *q = P.q + (M.s - P.off);
        for (long e = e0 + threadIdx.x; e < e0 + SC_CH && e < ne; e += WG) {
            // 32-bit division (N K < 2^31 for every dense merge a level can hold)
            const int v = (int)((unsigned)e / (unsigned)M.n), r = (int)e - v * M.n;
            const int slot = W.vm[cur][M.s + W.indxp[M.s + v] - 1];
            q[(long)(slot - P.off) * P.ld + r] = src[e];
            if (r == 0) { W.supp_lo[slot] = M.s; W.supp_hi[slot] = M.s + M.n; }
        }
    }
}

// INDXQ of the merged block: DLAMRG(K, N-K, D, 1, -1) (mdlaed1.f90:221), by the
// merge path over the whole grid (one thread per element, chunk_start as in
// k_newcols).  A merge whose runs k_newcols found out of order takes DLAMRG's
// sequential loop in the first thread of its first chunk, so the permutation is
// DLAMRG's in every case.
__global__ __launch_bounds__(WG) void k_lamrg(const Merge *mg, const int *chunk_start, int nm, const MergeOut *mo,
                                              Work W, int nxt, int *trc)
{
    if (gate_closed(W, mg)) return;
    const int m = find_seg(chunk_start, nm, blockIdx.x);
    const MergeOut o = mo[m];
    const int c = blockIdx.x - chunk_start[m];
    if (trc && c == 0 && threadIdx.x == 0) {  // the merge's trace entry: K, rotations, branch
        trc[3 * m] = o.flags ? 0 : o.k;
        trc[3 * m + 1] = o.flags ? 0 : o.nrot;
        trc[3 * m + 2] = (!o.flags && o.hss) ? 1 : 0;
    }
    if (o.flags) return;  // identity, set by k_newcols
    const Merge M = mg[m];
    const int n = M.n, n1 = o.k, n2 = M.n - o.k;
    const double *a = W.Dv[nxt] + M.s;
    int *out = W.indxq + M.s;
    if (W.maux[4 * m + 3]) {
        if (c == 0 && threadIdx.x == 0) {
            int i1 = 0, i2 = 0, q = 0;
            while (i1 < n1 && i2 < n2) {
                if (a[i1] <= a[n - 1 - i2]) { out[q++] = i1 + 1; i1++; }
                else { out[q++] = n - i2; i2++; }
            }
            while (i1 < n1) { out[q++] = i1 + 1; i1++; }
            while (i2 < n2) { out[q++] = n - i2; i2++; }
        }
        return;
    }
    const int i = c * WG + threadIdx.x;
    if (i >= n) return;
    if (i < n1) {
        const double v = a[i];
        int lo = 0, hi = n2;  // run-2 values strictly below v
        while (lo < hi) { int md = (lo + hi) >> 1; if (a[n - 1 - md] < v) lo = md + 1; else hi = md; }
        out[i + lo] = i + 1;
    } else {
        const int q = i - n1;
        const double v = a[n - 1 - q];
        int lo = 0, hi = n1;  // run-1 values <= v
        while (lo < hi) { int md = (lo + hi) >> 1; if (a[md] <= v) lo = md + 1; else hi = md; }
        out[q + lo] = n - q;
    }
}

// ---------------------------------------------------------------------------
// Device-scheduled levels.  k_lvltab turns the deflation results of a level into
// the tables its remaining kernels read (what the host builds after a round trip
// in the synchronised schedule), so a whole solve is queued without waiting.
// Per merge: chunks of 256 roots, secular blocks, columns, 64x64 GEMM tiles, S
// size, Ct size and copy chunks, for the dense merges (the HSS merges' vectors
// go straight into their slots once the host has seen the level).  Dense merges have
// K < hss_min_k, so sbcap and ctcap (doubles, sized by the host from the merge
// orders) bound the dense totals; a level beyond them does no work and sets bit
// 1 of *info.  One block; thread t owns a contiguous run of merges.
// ---------------------------------------------------------------------------
#define LT_NT 256
#define LT_Q 7
__device__ __forceinline__ void lt_counts(const Merge &M, const MergeOut &o, long *q)
{
    const int k = o.flags ? 0 : o.k;
    const bool hss = !o.flags && o.hss;
    const int rpb = SEC_WG / sec_team(k > 0 ? k : 1);
    q[0] = (k + WG - 1) / WG;
    q[1] = (k + rpb - 1) / rpb;
    q[2] = k;
    q[3] = (k > 0 && !hss) ? (long)(((M.n1 + 63) >> 6) + ((M.n - M.n1 + 63) >> 6)) * ((k + 63) >> 6) : 0;
    q[4] = (k > 0 && !hss) ? (long)k * k : 0;
    q[5] = hss ? 0 : (long)M.n * k;
    q[6] = hss ? 0 : ((long)M.n * k + SC_CH - 1) / SC_CH;
}

__global__ __launch_bounds__(LT_NT) void k_lvltab(const Merge *mg, int nm, const MergeOut *mo, Work W, LvlTab t,
                                                 long sbcap, long ctcap, int *info)
{
    if (gate_closed(W, mg)) return;
    __shared__ long wsum[LT_Q][LT_NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int per = (nm + LT_NT - 1) / LT_NT;
    const int a = min(nm, tid * per), b = min(nm, a + per);
    long own[LT_Q], q[LT_Q], base[LT_Q], tot[LT_Q];
    unsigned long long roots = 0, defl = 0, nh = 0;
#pragma unroll
    for (int i = 0; i < LT_Q; i++) own[i] = 0;
    for (int m = a; m < b; m++) {
        const Merge M = mg[m];
        const MergeOut o = mo[m];
        lt_counts(M, o, q);
#pragma unroll
        for (int i = 0; i < LT_Q; i++) own[i] += q[i];
        roots += q[2];
        defl += M.n - q[2];
        nh += (!o.flags && o.hss) ? 1 : 0;
        if (m == nm - 1) W.lstat[3] = (unsigned long long)q[2];
    }
    // exclusive prefix of every quantity: wavefront scan, then over wavefront totals
#pragma unroll
    for (int i = 0; i < LT_Q; i++) {
        long x = own[i];
        for (int off = 1; off < 64; off <<= 1) {
            long y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[i][wv] = x;
        base[i] = x - own[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < LT_Q; i++) {
        long pre = 0, all = 0;
#pragma unroll
        for (int w = 0; w < LT_NT / 64; w++) {
            pre += (w < wv) ? wsum[i][w] : 0;
            all += wsum[i][w];
        }
        base[i] += pre;
        tot[i] = all;
    }
    const bool over = tot[4] > sbcap || tot[5] > ctcap;
    for (int m = a; m < b; m++) {
        const MergeOut o = mo[m];
        lt_counts(mg[m], o, q);
        t.chunkK[m] = over ? 0 : (int)base[0];
        t.blockK[m] = over ? 0 : (int)base[1];
        t.colK[m] = over ? 0 : (int)base[2];
        t.tileK[m] = over ? 0 : (int)base[3];
        t.scK[m] = over ? 0 : (int)base[6];
        t.soff[m] = base[4];
        t.coff[m] = base[5];
#pragma unroll
        for (int i = 0; i < LT_Q; i++) base[i] += q[i];
    }
    if (tid == 0) {
        t.chunkK[nm] = over ? 0 : (int)tot[0];
        t.blockK[nm] = over ? 0 : (int)tot[1];
        t.colK[nm] = over ? 0 : (int)tot[2];
        t.tileK[nm] = over ? 0 : (int)tot[3];
        t.scK[nm] = over ? 0 : (int)tot[6];
        if (over) atomicOr(info, 2);
    }
    // statistics, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        roots += __shfl_xor(roots, off);
        defl += __shfl_xor(defl, off);
        nh += __shfl_xor(nh, off);
    }
    if (lane == 0 && roots) atomicAdd(&W.lstat[0], roots);
    if (lane == 0 && defl) atomicAdd(&W.lstat[1], defl);
    if (lane == 0 && nh) atomicAdd(&W.lstat[2], nh);
    // HSS merges: the levels above wait for the host (this level's own kernels do
    // not: they skip its HSS merges)
    const int anyh = __syncthreads_or(nh > 0);
    if (tid == 0 && anyh && !over) *W.gate = mg[0].lvl + 1;
}

// ---------------------------------------------------------------------------
// Zero fill of the rows x ncols output Z (leading dimension ldz), launched on a
// second stream when the solve starts: the DC levels are latency bound and leave
// HBM idle, so the 8 n^2 bytes of zeros are written under them, and the final
// reordering writes only the supported rows.  Consecutive blocks write
// consecutive 4 KB pieces (one contiguous front, as a plain device fill).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_zfill(double *Z, long rows, long ncols, long ldz)
{
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    const dbl2 z2 = {0.0, 0.0};
    if (ldz == rows && (((unsigned long)Z) & 15) == 0) {
        const long n2 = rows * ncols;
        dbl2 *d2 = (dbl2 *)Z;
        for (long q = blockIdx.x * (long)WG + threadIdx.x; q < (n2 >> 1); q += (long)gridDim.x * WG)
            __builtin_nontemporal_store(z2, d2 + q);
        if ((n2 & 1) && blockIdx.x == 0 && threadIdx.x == 0) Z[n2 - 1] = 0.0;
    } else {
        for (long col = blockIdx.x; col < ncols; col += gridDim.x)
            for (long r = threadIdx.x; r < rows; r += WG) __builtin_nontemporal_store(0.0, Z + col * ldz + r);
    }
}

// ---------------------------------------------------------------------------
// Final reordering (mdlaed0.f90:355-362 and the scale-back of mdstedc.f90:358):
// Z(:, i) = Q(:, slot of INDXQ(i)), zero outside the column's support.  One block
// per FIN_CH rows of an output column, consecutive blocks on consecutive rows of
// one column, so the blocks in flight write one contiguous front of Z (blocks
// owning whole columns write thousands of separate streams and reach about 5.1
// TB/s against 5.9 TB/s measured for this order).  zrows_global: one n x n output,
// cz chunks per column; else problem p owns blocks [fbs[p], fbs[p+1]) of its own
// output block.  zfilled: k_zfill has zeroed Z (on a second stream, during the DC
// levels), and one block per column writes only the rows of its slot's support.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_final(const Prob *probs, const int *col_prob, Work W, const int *outcol,
                                              double *Z, const long *zoff, const long *ldz, int zrows_global,
                                              double *dout, const double *scale, unsigned long long *rows_read,
                                              const int *fbs, int np, int cz, int zfilled)
{
    if (*W.gate) return;  // a level waits for the host
    int g, c, p;
    if (zfilled) {  // Z is zero already: one block per column writes the supported rows
        g = blockIdx.x;
        c = 0;
        p = col_prob[g];
    } else if (zrows_global) {
        g = blockIdx.x / cz;
        c = blockIdx.x - g * cz;
        p = col_prob[g];
    } else {
        p = find_seg(fbs, np, blockIdx.x);
        const int n = probs[p].n, cp = (n + FIN_CH - 1) / FIN_CH, loc = blockIdx.x - fbs[p];
        g = probs[p].off + loc / cp;
        c = loc - (loc / cp) * cp;
    }
    const Prob P = probs[p];
    const int v = W.indxq[g];  // local virtual column (1-based)
    const int vg = P.off + v - 1;
    const int slot = W.vm[P.par][vg];
    const int oc = outcol ? outcol[g] : g;  // output column (global numbering)
    const int lo = W.supp_lo[slot], hi = W.supp_hi[slot];
    if (c == 0 && threadIdx.x == 0) {
        dout[oc] = W.Dv[P.par][vg] * scale[p];
        atomicAdd(rows_read + (g & 63), (unsigned long long)(hi - lo));  // rows read, 64 counters
    }
    const double *src = P.q + (long)(slot - P.off) * P.ld - P.off;
    double *dst;
    int r0, r1;
    if (zrows_global) {  // one n x n output holding every problem (block diagonal)
        dst = Z + (long)oc * ldz[0];
        r0 = 0; r1 = zrows_global;
    } else {             // one output block per problem
        dst = Z + zoff[p] + (long)(oc - P.off) * ldz[p] - P.off;
        r0 = P.off; r1 = P.off + P.n;
    }
    if (zfilled) {
        const int a = max(lo, r0), b = min(hi, r1);
        for (int r = a + threadIdx.x; r < b; r += WG) __builtin_nontemporal_store(src[r], dst + r);
        return;
    }
    const int c0 = r0 + c * FIN_CH, c1 = min(r1, c0 + FIN_CH);  // this block's rows
    // streaming stores, 16 bytes per lane when the column start is 16-byte aligned
    // (FIN_CH is even, so every chunk start keeps that alignment)
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    if ((((unsigned long)(dst + r0)) & 15) == 0) {
        dbl2 *d2 = (dbl2 *)(dst + c0);
        const int nr = c1 - c0;
        for (int q = threadIdx.x; q < (nr >> 1); q += WG) {
            const int r = c0 + 2 * q;
            dbl2 val;
            val.x = (r >= lo && r < hi) ? src[r] : 0.0;
            val.y = (r + 1 >= lo && r + 1 < hi) ? src[r + 1] : 0.0;
            __builtin_nontemporal_store(val, d2 + q);
        }
        if ((nr & 1) && threadIdx.x == 0) {
            const int r = c1 - 1;
            __builtin_nontemporal_store((r >= lo && r < hi) ? src[r] : 0.0, dst + r);
        }
    } else {
        for (int r = c0 + threadIdx.x; r < c1; r += WG)
            __builtin_nontemporal_store((r >= lo && r < hi) ? src[r] : 0.0, dst + r);
    }
}

// The figures the host reads at the end of a solve, gathered by one wavefront into
// one record (mapped host memory): counters kept in SEC_SLOTS = 64 copies are summed
// (level maxima maximised over their copies, then summed over the levels).
__global__ __launch_bounds__(64) void k_readback(Work W, const int *info, const unsigned long long *rows,
                                                 const SecCtl *sec, const unsigned int *lvlmax, SolveRec *out)
{
    const int lane = threadIdx.x;
    const SecCtl C = *sec;
    unsigned long long r = rows[lane], it = C.iters[lane], us = C.used[lane], wi = C.wave_iters[lane],
                       ev = C.evals[lane];
    unsigned long long crit = 0;
    for (int l = 0; l < SEC_LEVELS; l++) {
        unsigned int m = lvlmax[l * SEC_SLOTS + lane];
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
        crit += m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off);
        it += __shfl_xor(it, off);
        us += __shfl_xor(us, off);
        wi += __shfl_xor(wi, off);
        ev += __shfl_xor(ev, off);
    }
    if (lane == 0) {
        out->gate = *W.gate;
        out->info = *info;
        out->rows = r;
        out->roots = W.lstat[0]; out->deflated = W.lstat[1]; out->hss = W.lstat[2]; out->ktop = W.lstat[3];
        out->iters = it; out->used = us; out->wave_iters = wi; out->crit = crit; out->evals = ev;
    }
}

// Plain C = A * B (column-major) on FP64 MFMA, for COMPZ = 'V' (Z := Zin * Q).
__global__ __launch_bounds__(WG) void k_dgemm_nn(int m, int n, int k, const double *A, long lda,
                                                 const double *B, long ldb, double *C, long ldc)
{
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
    d4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < k; k0 += 16) {
        for (int e = 0; e < 4; e++) {
            int idx = tid + WG * e;
            int r = idx & 63, kk = idx >> 6;
            int t = k0 + kk;
            As[kk][r] = (t < k && r0 + r < m) ? A[(long)(r0 + r) + (long)t * lda] : 0.0;
            Bs[kk][r] = (t < k && c0 + r < n) ? B[(long)t + (long)(c0 + r) * ldb] : 0.0;
        }
        __syncthreads();
        for (int kq = 0; kq < 16; kq += 4) {
            int kl = kq + (lane >> 4);
            double a0 = As[kl][wr + (lane & 15)], a1 = As[kl][wr + 16 + (lane & 15)];
            double b0 = Bs[kl][wc + (lane & 15)], b1 = Bs[kl][wc + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int i = 0; i < 4; i++) {
                int row = r0 + wr + a * 16 + (lane >> 4) + 4 * i;
                int col = c0 + wc + b * 16 + (lane & 15);
                if (row < m && col < n) C[(long)row + (long)col * ldc] = acc[a][b][i];
            }
}

__global__ __launch_bounds__(WG) void k_fill_u64(unsigned long long *p, long n, unsigned long long v)
{
    long i = (long)blockIdx.x * WG + threadIdx.x;
    if (i < n) p[i] = v;
}

// B (cols x rows, ldb) := A^T for A (rows x cols, lda), column-major: 64 x 64 tiles
// through LDS (coalesced reads and writes), one tile per block.
__global__ __launch_bounds__(WG) void k_transpose(int rows, int cols, const double *A, long lda, double *B, long ldb)
{
    __shared__ double t[64][65];
    const long r0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int j = ty; j < 64; j += WG / 64) {
        const long r = r0 + tx, cc = c0 + j;
        if (r < rows && cc < cols) t[j][tx] = A[r + cc * lda];
    }
    __syncthreads();
    for (int j = ty; j < 64; j += WG / 64) {
        const long cc = c0 + tx, r = r0 + j;
        if (r < rows && cc < cols) B[cc + r * ldb] = t[tx][j];
    }
}
