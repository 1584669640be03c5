// Bidiagonal divide-and-conquer SVD on gfx950: HSSPACK's MDBDSDC (SRC/mdbdsdc.f90)
// -> MDLASD0 (mdlasd0.f90) -> MDLASD1 (mdlasd1.f90) -> MDLASD2 (mdlasd2.f90) and the
// dense merge DLASD3/DLASD4 it calls below BSMLZ, as HIP kernels batched over every
// merge of one tree level.
//
// Layout.  U (n x n) and V = VT^T (n x n) are column-major, one buffer each.  A
// singular vector ("virtual column" j of its merge block) lives in a physical slot
// found through a map (vmU / vmV, ping-pong by level); every slot has a row support
// [lo, hi), the rows of the tree node that last wrote it (the whole merge block once
// a deflation rotation has touched it), and is read as zero outside it, so no buffer
// is ever cleared.  A deflated column keeps its slot: only the map changes.  The K
// new columns of a merge are formed in a staging buffer and scattered into the
// slots that merge frees (its non-deflated inputs and its centre column).  The
// slots of a block are the block's own columns, so merges of one level never share
// one.  Working vectors live at the global index of the merge block (s + i).
//
// Per level: k_bd_prep (scaling, z, sort: mdlasd1.f90:171-184, mdlasd2.f90:216-268),
// k_bd_deflate (mdlasd2.f90:270-470, in parallel), a read of K, k_bd_rotate (the
// deflation's Givens rotations), k_bd_znorm, k_bd_secular (DLASD4), k_bd_loewner,
// k_bd_vec (DLASD3's vector formulas), k_bd_gemm (DLASD3's products, gathering
// columns through the deflation's permutation into staging), k_bd_newcol (the
// columns formed without a product), k_bd_scatter (staging -> freed slots) and
// k_bd_post (unscaling, DLAMRG (mdlasd1.f90:282-286), the new maps and supports).
// U and VT are cleared on a second stream while the levels run; k_bd_fin (U) and
// k_bd_fin_vt (VT, transposed through LDS tiles) then write the supported rows of
// every vector in MDBDSDC's final order.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/hsdc.h"
#include "hsdc_internal.h"
#include "hsdc_hss.h"

#define BD_EPS 0x1p-53
#define BD_LDS_RUNS 19968  // doubles of dynamic LDS for a merge's sorted runs (156 KB)
#define BD_SAFMIN 0x1p-1022
#define JAC_TOL 0x1p-52
#define JAC_SWEEPS 60

// one merge of a level: block [s, s+n), n = nl+nr+1, m = n+sqre; U / V columns are read
// from buffer rd and written to buffer wr
struct BdMerge {
    int s, nl, nr, sqre, rd, wr;
};

// per-merge results of a level
struct BdOut {
    double orgnrm, alpha, beta, tol, z1, zr1, c, sn, rho;
    int k, nrot, ctot[4];
    int hss;  // 1: the merge takes the HSS branch (set by the host after deflation)
};

// one HSS merge of a level: merge index, offsets of its generators (5 K: d z, APHAU,
// APHAV, DIFL, DIFR) and of its gathered U2 (n x K) and V2 (m x K) in DSIGMA order
struct BdHss {
    int mi;
    long goff, xoff;
};

struct BdLeaf {
    int s, m, sqre, wr;
};

// launch tables of one level, built on the device from its deflation results
// (k_bd_lvltab): per merge, the first work item of every grid-stride kernel (nm + 1
// prefix counts each) and the offsets of its vector and staging blocks
struct BdTab {
    int *rot, *sec, *low, *tile, *sc;
    long *qoff, *soff;
};

// merge owning work item b of a prefix table (pre[0] = 0 <= b < pre[nm])
__device__ __forceinline__ int bd_find(const int *pre, int nm, int b)
{
    int lo = 0, hi = nm;
    while (hi - lo > 1) {
        const int md = (lo + hi) >> 1;
        if (pre[md] <= b) lo = md; else hi = md;
    }
    return lo;
}

struct BdWork {
    double *D, *E;                          // n
    int *idxq;                              // n (1-based values, local to each block)
    double *td, *dsg, *zt, *z, *zs, *zl;    // n
    double *sig, *tau;                      // n
    int *org, *ctyp, *idxc, *idx, *idxp;    // n
    int *surv, *srank, *clf, *rotf, *ti;   // n
    int *vmU[2], *vmV[2];                   // n: virtual column -> slot (ping-pong by level)
    int *sulo, *suhi, *svlo, *svhi;         // n: row support [lo, hi) of every U / V slot
    int *gU, *gV, *gUlo, *gUhi, *gVlo, *gVhi;  // n: slot and support of U2(:,J) / VT2(J,:)
    int *fU, *fV;                           // n: slots receiving a merge's new columns
    int *rp4;                               // 4 n: rotation slots (pU | fresh << 30, qU, pV, qV)
    double *rcs;                            // 2 n: rotation c, s
    unsigned long long *iters;              // secular iteration counter
};

// ------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------
// butterfly over the 32 lanes of a half wavefront; equals wsum64 of the same values
// when the other half is zero (its step adds exact zeros)
__device__ __forceinline__ double wsum32(double v)
{
    v = v + xlane<16>(v);
    v = v + xlane<8>(v);
    v = v + xlane<4>(v);
    v = v + xlane<2>(v);
    v = v + xlane<1>(v);
    return v;
}

__device__ __forceinline__ double wsum64(double v)
{
    v = v + xlane<32>(v);
    v = v + xlane<16>(v);
    v = v + xlane<8>(v);
    v = v + xlane<4>(v);
    v = v + xlane<2>(v);
    v = v + xlane<1>(v);
    return v;
}

__device__ __forceinline__ double tsum_team(double v, int T) { return team_bfly(v, T); }

__device__ __forceinline__ double bd_lapy2(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    double r = z / w;
    return w * sqrt(1.0 + r * r);
}

// the multipliers DLASCL applies in turn to scale by cto/cfrom (at most a few)
__host__ __device__ inline int lascl_muls(double cfrom, double cto, double *mul)
{
    const double smlnum = BD_SAFMIN, bignum = 1.0 / smlnum;
    double cfromc = cfrom, ctoc = cto;
    int nm = 0;
    for (;;) {
        double cfrom1 = cfromc * smlnum, m;
        bool done = false;
        if (cfrom1 == cfromc) {
            m = ctoc / cfromc;
            done = true;
        } else {
            double cto1 = ctoc / bignum;
            if (cto1 == ctoc) {
                m = ctoc;
                done = true;
                cfromc = 1.0;
            } else if (fabs(cfrom1) > fabs(ctoc) && ctoc != 0.0) {
                m = smlnum;
                cfromc = cfrom1;
            } else if (fabs(cto1) > fabs(cfromc)) {
                m = bignum;
                ctoc = cto1;
            } else {
                m = ctoc / cfromc;
                done = true;
                if (m == 1.0) return nm;
            }
        }
        if (nm < 8) mul[nm++] = m;
        if (done) return nm;
    }
}

__device__ __forceinline__ double apply_muls(double x, const double *mul, int nm)
{
    for (int i = 0; i < nm; i++) x *= mul[i];
    return x;
}

// exclusive block scan of one int per thread; returns the total
template <int NT>
__device__ int block_exscan(int v, int *sh, int &total)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int w = 0; w < NT / 64; w++) {
            int c = sh[w];
            sh[w] = acc;
            acc += c;
        }
        sh[NT / 64] = acc;
    }
    __syncthreads();
    int r = sh[wv] + x - v;
    total = sh[NT / 64];
    __syncthreads();
    return r;
}

// Ranks the flagged positions of [lo, hi) in position order: each wavefront walks one
// contiguous segment 64 positions at a time (ballot + popcount), a block scan over the
// wavefront counts gives each segment its base.  emit(p, rank) for every flagged p;
// returns the total.  The flags of a segment's first 64 steps are evaluated up front
// into one bit per step (independent loads, in flight together) and reused by the emit
// pass; flag must be free of side effects and must not read what emit writes.
template <typename F, typename G>
__device__ int wave_rank(int lo, int hi, F flag, G emit, int *sh)
{
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int len = hi > lo ? hi - lo : 0;
    const int seg = (((len + nw - 1) / nw) + 63) & ~63;
    const int a = lo + w * seg, b = min(a + seg, hi);
    const int nit = b > a ? (b - a + 63) / 64 : 0;
    unsigned long long fm = 0;
#pragma unroll 8
    for (int i = 0; i < min(nit, 64); i++) {
        const int p = a + 64 * i + lane;
        if (p < b && flag(p)) fm |= 1ull << i;
    }
    auto fl = [&](int i) -> bool {
        const int p = a + 64 * i + lane;
        return i < 64 ? ((fm >> i) & 1ull) != 0 : (p < b && flag(p));
    };
    int cnt = 0;
    for (int i = 0; i < nit; i++) cnt += __popcll(__ballot(fl(i)));
    if (lane == 0) sh[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < nw; i++) {
            int c = sh[i];
            sh[i] = acc;
            acc += c;
        }
        sh[nw] = acc;
    }
    __syncthreads();
    int base = sh[w];
    const int total = sh[nw];
    __syncthreads();
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int i = 0; i < nit; i++) {
        const int p = a + 64 * i + lane;
        const bool f = fl(i);
        const unsigned long long msk = __ballot(f);
        if (f) emit(p, base + __popcll(msk & below));
        base += __popcll(msk);
    }
    return total;
}

// for (i = i0 + threadIdx.x; i < n; i += NT): the loads of U consecutive iterations
// (load(i) -> value) are issued before their stores (store(i, value)), so their
// dependent gathers overlap instead of waiting on the previous iteration's stores.
// The iterations must be independent: no load reads what a store of the loop writes.
template <int U, int NT, typename L, typename S>
__device__ __forceinline__ void bd_batched(int i0, int n, L load, S store)
{
    for (int b = i0 + (int)threadIdx.x; b < n; b += U * NT) {
        decltype(load(0)) v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b + u * NT < n) v[u] = load(b + u * NT);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b + u * NT < n) store(b + u * NT, v[u]);
    }
}

// merge path: how many of the first diag outputs of the stable merge of ascending
// a (na) and b (nb) come from a (ties take a first); bget(j) reads b[j]
template <typename B>
__device__ int corank(int diag, const double *a, int na, B bget, int nb)
{
    int lo = diag - nb > 0 ? diag - nb : 0, hi = diag < na ? diag : na;
    while (lo < hi) {
        const int i = (lo + hi) >> 1, j = diag - i;
        if (j > 0 && a[i] <= bget(j - 1)) lo = i + 1;
        else hi = i;
    }
    return lo;
}

__device__ double block_max(double v, double *sh)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    if (t == 0) {
        double m = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, sh[w]);
        sh[0] = m;
    }
    __syncthreads();
    double r = sh[0];
    __syncthreads();
    return r;
}

// ------------------------------------------------------------------------------
// leaves: one-sided Jacobi in round-robin order (replaces DLASDQ, mdlasd0.f90:152-183),
// one block per leaf: lane r of every wavefront addresses row r of the LDS columns,
// and each wavefront rotates one pair of a round at a time
// ------------------------------------------------------------------------------
// pair i of round t of the circle method over P (even) columns: column 0 fixed, the
// others moved one place per round; p < q
__device__ __forceinline__ void bd_rr_pair(int t, int i, int P, int &p, int &q)
{
    int a, b;
    if (i == 0) {
        a = 0;
        b = (P - 2 + t) % (P - 1) + 1;
    } else {
        a = (i - 1 + t) % (P - 1) + 1;
        b = (P - 2 - i + t) % (P - 1) + 1;
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// HALF (leaves of at most 32 columns): each half of a wavefront rotates its own pair,
// 32 lanes per column; otherwise a wavefront per pair
template <bool HALF>
__global__ __launch_bounds__(1024) void k_bd_leaf(const BdLeaf *lv, int nleaf, BdWork W, double *U, double *V,
                                                  long ld, int pmax, int *fail)
{
    const int li = blockIdx.x;
    if (li >= nleaf) return;
    const BdLeaf L = lv[li];
    const int s = L.s, m = L.m, pc = m + L.sqre, r = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nw = (int)(blockDim.x >> 6);
    const int ppw = HALF ? 2 : 1;                  // pairs per wavefront at a time
    const int hf = HALF ? (r >> 5) : 0, rr = HALF ? (r & 31) : r;  // half, row
    extern __shared__ double lds[];
    double *Wc = lds, *Vc = lds + (size_t)pmax * 64, *nn = Vc + (size_t)pmax * 64, *sq = nn + 64;
    int *ord = (int *)(sq + 64);
    __shared__ int jn_s, rot_s;
    const double *d = W.D + s, *e = W.E + s;
    if (wv == 0)
        for (int j = 0; j < pc; j++) {
            double w = 0.0;
            if (r == j && j < m) w = d[j];
            if (j >= 1 && r == j - 1) w = e[j - 1];
            Wc[j * 64 + r] = w;
            Vc[j * 64 + r] = (r == j) ? 1.0 : 0.0;
        }
    double fro = 0.0;
    for (int j = 0; j < m; j++) fro = fro + d[j] * d[j];
    for (int j = 0; j < pc - 1; j++) fro = fro + e[j] * e[j];
    const double thr = (JAC_TOL * pc) * (JAC_TOL * pc) * fro;
    __syncthreads();
    // squared column norms, each the butterfly sum of the column as it now stands (the
    // same value a fresh sum would give), recomputed only when a rotation changes it;
    // every lane of the wavefront owning the column writes the same value
    if (wv == 0)
        for (int j = 0; j < pc; j++) {
            double w = Wc[j * 64 + r];
            const double a = wsum64(w * w);
            nn[j] = a;
            sq[j] = sqrt(a);   // the norm, kept beside its square
        }
    // round-robin sweeps: the P/2 pairs of a round touch disjoint columns, one pair per
    // wavefront (or half wavefront) at a time, a barrier between rounds
    const int P = pc + (pc & 1);
    int sweep, rot = 1;
    for (sweep = 0; sweep < JAC_SWEEPS && rot; sweep++) {
        if (threadIdx.x == 0) rot_s = 0;
        __syncthreads();
        for (int rd = 0; rd < P - 1; rd++) {
            for (int pb = wv * ppw; pb < P / 2; pb += nw * ppw) {
                const int pi = pb + hf;
                if (pi >= P / 2) continue;
                int p, q;
                bd_rr_pair(rd, pi, P, p, q);
                if (q >= pc) continue;
                double wp = Wc[p * 64 + rr], wq = Wc[q * 64 + rr];
                double al = nn[p];
                double be = nn[q];
                double ga = HALF ? wsum32(wp * wq) : wsum64(wp * wq);
                if (al <= thr || be <= thr || !(fabs(ga) > (JAC_TOL * pc) * (sq[p] * sq[q]))) continue;
                double zeta = (be - al) / (2.0 * ga), t;
                if (fabs(zeta) > 1e150) t = 0.5 / zeta;
                else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                const double np_ = c * wp - sn * wq, nq_ = sn * wp + c * wq;
                Wc[p * 64 + rr] = np_;
                Wc[q * 64 + rr] = nq_;
                double vp = Vc[p * 64 + rr], vq = Vc[q * 64 + rr];
                Vc[p * 64 + rr] = c * vp - sn * vq;
                Vc[q * 64 + rr] = sn * vp + c * vq;
                const double ap = HALF ? wsum32(np_ * np_) : wsum64(np_ * np_);
                const double aq = HALF ? wsum32(nq_ * nq_) : wsum64(nq_ * nq_);
                nn[p] = ap;
                nn[q] = aq;
                sq[p] = sqrt(ap);
                sq[q] = sqrt(aq);
                if (rr == 0) rot_s = 1;
            }
            __syncthreads();
        }
        rot = rot_s;
        __syncthreads();  // every thread has read rot_s before the next sweep clears it
    }
    if (rot && threadIdx.x == 0) atomicAdd(fail, 1);
    if (threadIdx.x == 0) {
        unsigned long long used = 0ULL;  // bit j: column j taken (pc <= 64); a mask, not an
                                         // array, so nothing goes to scratch memory
        int jn = -1;
        if (L.sqre) {
            jn = 0;
            for (int j = 1; j < pc; j++)
                if (sq[j] < sq[jn]) jn = j;
            used |= 1ULL << jn;
        }
        for (int k = 0; k < m; k++) {
            int b = -1;
            for (int j = 0; j < pc; j++)
                if (!((used >> j) & 1ULL) && (b < 0 || sq[j] < sq[b])) b = j;
            used |= 1ULL << b;
            ord[k] = b;
        }
        jn_s = jn;
    }
    __syncthreads();
    if (wv == 0) {
        for (int k = 0; k < m; k++) {
            int j = ord[k];
            double sg = sq[j];
            if (r == 0) W.D[s + k] = sg;
            if (r < m) U[(long)(s + k) * ld + s + r] = nn[j] > thr ? Wc[j * 64 + r] / sg : 0.0;
            if (r < pc) V[(long)(s + k) * ld + s + r] = Vc[j * 64 + r];
            if (r == 0) W.idxq[s + k] = k + 1;
        }
        if (L.sqre && r < pc) V[(long)(s + m) * ld + s + r] = Vc[jn_s * 64 + r];
        // slot = column, supported on the leaf's rows
        if (r < m) {
            W.vmU[L.wr][s + r] = s + r;
            W.sulo[s + r] = s;
            W.suhi[s + r] = s + m;
        }
        if (r < pc) {
            W.vmV[L.wr][s + r] = s + r;
            W.svlo[s + r] = s;
            W.svhi[s + r] = s + pc;
        }
    }
    __syncthreads();
    // left vectors of negligible columns: the first e_r keeping norm > 1/2 after two
    // projections off the other columns (wavefront 0; the loop bounds are uniform)
    for (int k = 0; k < m; k++) {
        if (nn[ord[k]] > thr) continue;
        if (wv == 0)
            for (int r0 = 0; r0 < m; r0++) {
                double x = (r == r0) ? 1.0 : 0.0;
                for (int pass = 0; pass < 2; pass++)
                    for (int k2 = 0; k2 < m; k2++) {
                        if (k2 == k) continue;
                        double uk = r < m ? U[(long)(s + k2) * ld + s + r] : 0.0;
                        double g = wsum64(r < m ? uk * x : 0.0);
                        if (r < m) x = x - g * uk;
                    }
                double nx = sqrt(wsum64(x * x));
                if (nx > 0.5) {
                    if (r < m) U[(long)(s + k) * ld + s + r] = x / nx;
                    break;
                }
            }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------
// merge step 1 (block per merge): ORGNRM scaling (mdlasd1.f90:171-184), z from the
// children's right vectors, the shift, DSIGMA by IDXQ, DLAMRG and the sort
// (mdlasd2.f90:216-262), TOL (mdlasd2.f90:266-268)
// ------------------------------------------------------------------------------
#define PREP_NT 1024
// lds_n: capacity in doubles of the dynamic LDS that holds the merged runs (0: they
// are read from global memory)
__global__ __launch_bounds__(PREP_NT) void k_bd_prep(const BdMerge *mg, BdWork W, BdOut *mo, const double *V,
                                                     long ld, int lds_n)
{
    __shared__ double shd[PREP_NT / 64 + 1];
    extern __shared__ double lruns[];
    const BdMerge M = mg[blockIdx.x];
    const int s = M.s, nl = M.nl, nr = M.nr, n = nl + nr + 1, m = n + M.sqre, t = threadIdx.x;
    const int *vm = W.vmV[M.rd];
    const double alpha0 = W.D[s + nl], beta0 = W.E[s + nl];
    double mx = 0.0;
    for (int i = t; i < n; i += PREP_NT)
        if (i != nl) mx = fmax(mx, fabs(W.D[s + i]));
    double orgnrm = fmax(fmax(fabs(alpha0), fabs(beta0)), block_max(mx, shd));
    double mul[8];
    const int nmul = lascl_muls(orgnrm, 1.0, mul);
    const double alpha = alpha0 / orgnrm, beta = beta0 / orgnrm;
    for (int i = t; i < n; i += PREP_NT) W.td[s + i] = apply_muls(i == nl ? 0.0 : W.D[s + i], mul, nmul);
    __syncthreads();
    // shift the left part one place, IDXQ adjusted, z, column types
    for (int i = t; i < n; i += PREP_NT) {
        double dv;
        int q;
        if (i >= 1 && i <= nl) {
            dv = W.td[s + i - 1];
            q = W.idxq[s + i - 1] + 1;
        } else {
            dv = W.td[s + i];
            q = (i >= nl + 1) ? W.idxq[s + i] + nl + 1 : W.idxq[s + i];
        }
        W.dsg[s + i] = dv;  // staging
        W.ti[s + i] = q;
        W.ctyp[s + i] = (i >= 1 && i <= nl) ? 1 : 2;
    }
    // row NLP1 (NLP2) of the children's right vectors; a slot reads as zero outside
    // its support (a column deflated below this level ends above those rows)
    bd_batched<8, PREP_NT>(0, m, [&](int i) {
        const int slot = vm[s + (i == 0 ? nl : (i <= nl ? i - 1 : i))];
        const int g = (i <= nl) ? s + nl : s + nl + 1;
        return (g >= W.svlo[slot] && g < W.svhi[slot]) ? V[(long)slot * ld + g] : 0.0;
    }, [&](int i, double v) { W.z[s + i] = (i <= nl ? alpha : beta) * v; });
    __syncthreads();
    struct DQ { double d; int q; };
    bd_batched<8, PREP_NT>(0, n, [&](int i) { return DQ{W.dsg[s + i], W.ti[s + i]}; },
                           [&](int i, const DQ &v) { W.D[s + i] = v.d; W.idxq[s + i] = v.q; });
    __syncthreads();
    // DSIGMA(I) = D(IDXQ(I)), U2(I,1) = Z(IDXQ(I)), IDXC(I) = COLTYP(IDXQ(I))
    struct DZC { double d, z; int c; };
    bd_batched<8, PREP_NT>(1, n, [&](int i) {
        const int q = W.idxq[s + i] - 1;
        return DZC{W.D[s + q], W.z[s + q], W.ctyp[s + q]};
    }, [&](int i, const DZC &v) { W.dsg[s + i] = v.d; W.zt[s + i] = v.z; W.idxc[s + i] = v.c; });
    __syncthreads();
    // DLAMRG(NL, NR, DSIGMA(2), 1, 1, IDX(2)): ties take the left run first; merge path,
    // each thread places a contiguous run of outputs.  The two runs are staged in LDS
    // when they fit: the co-rank search and the run read them with data-dependent
    // indices, one after another
    const bool inl = n - 1 <= lds_n;
    if (inl) {
        for (int i = t; i < n - 1; i += PREP_NT) lruns[i] = W.dsg[s + 1 + i];
        __syncthreads();
    }
    {
        const double *a = inl ? lruns : W.dsg + s + 1, *b = a + nl;
        const int tot = nl + nr, per = (tot + PREP_NT - 1) / PREP_NT;
        const int o0 = min(t * per, tot), o1 = min(o0 + per, tot);
        auto bget = [b](int j) { return b[j]; };
        int i = corank(o0, a, nl, bget, nr), j = o0 - i;
        for (int o = o0; o < o1; o++) {
            if (i < nl && (j >= nr || a[i] <= b[j])) W.idx[s + 1 + o] = ++i;
            else W.idx[s + 1 + o] = nl + (++j);
        }
    }
    __syncthreads();
    bd_batched<8, PREP_NT>(1, n, [&](int i) {
        const int ii = W.idx[s + i];  // 1-based into DSIGMA(2..) -> DSIGMA(1 + ii)
        return DZC{W.dsg[s + ii], W.zt[s + ii], W.idxc[s + ii]};
    }, [&](int i, const DZC &v) { W.D[s + i] = v.d; W.z[s + i] = v.z; W.ctyp[s + i] = v.c; });
    __syncthreads();
    if (t == 0) {
        double tol = fmax(fabs(alpha), fabs(beta));
        tol = 8.0 * BD_EPS * fmax(fabs(W.D[s + n - 1]), tol);
        BdOut &o = mo[blockIdx.x];
        o.orgnrm = orgnrm;
        o.hss = 0;
        o.alpha = alpha;
        o.beta = beta;
        o.tol = tol;
        o.z1 = W.z[s];
    }
}

// column of U (0-based, local) behind sorted position p (mdlasd2.f90:340-345)
__device__ __forceinline__ int bd_colidx(const BdWork &W, int s, int nl, int p)
{
    int v = W.idxq[s + W.idx[s + p]];
    if (v <= nl + 1) v--;
    return v - 1;
}

// ------------------------------------------------------------------------------
// merge step 2 (block per merge): deflation (mdlasd2.f90:270-470).  The sequential
// scan is restated in parallel: a position either has small z (deflated at itself)
// or survives; a survivor whose singular value lies within TOL of the previous
// survivor's rotates that one out (deflated at the survivor's position); the
// rotations form chains along runs of such survivors, each chain walked by one
// thread.  Deflation events fill IDXP from the back in position order.
// ------------------------------------------------------------------------------
#define DEF_NT 1024
__global__ __launch_bounds__(DEF_NT) void k_bd_deflate(const BdMerge *mg, BdWork W, BdOut *mo)
{
    __shared__ int sh[DEF_NT / 64 + 2];
    __shared__ int cnt[4];
    const BdMerge M = mg[blockIdx.x];
    const int s = M.s, nl = M.nl, nr = M.nr, n = nl + nr + 1, m = n + M.sqre, t = threadIdx.x;
    BdOut &o = mo[blockIdx.x];
    const double tol = o.tol;
    double *z = W.z + s;
    const double *D = W.D + s;
    int *surv = W.surv + s, *srank = W.srank + s, *clf = W.clf + s, *idxp = W.idxp + s, *ctyp = W.ctyp + s;
    for (int i = t; i < n; i += DEF_NT) {
        srank[i] = -1;
        W.rotf[s + i] = 0;
    }
    __syncthreads();
    // survivors of the small-z test, in position order
    const int ns = wave_rank(1, n, [&](int p) { return !(fabs(z[p]) <= tol); },
                             [&](int p, int r) { surv[r] = p; srank[p] = r; }, sh);
    __syncthreads();
    // a survivor within TOL of the previous survivor rotates that one out
    for (int q = t; q < ns; q += DEF_NT) clf[q] = (q >= 1) && (fabs(D[surv[q]] - D[surv[q - 1]]) <= tol);
    __syncthreads();
    // deflation events in position order fill IDXP from the back
    wave_rank(1, n, [&](int p) { int r = srank[p]; return r < 0 || clf[r]; },
              [&](int p, int e) {
                  int r = srank[p];
                  if (r < 0) {
                      idxp[n - 1 - e] = p + 1;
                      ctyp[p] = 4;
                  } else {
                      idxp[n - 1 - e] = surv[r - 1] + 1;
                  }
              }, sh);
    // rotation ranks and the non-deflated survivors, in survivor order
    const int nrot = wave_rank(0, ns, [&](int q) { return clf[q] != 0; }, [&](int q, int r) { W.ti[s + q] = r; }, sh);
    const int nk = wave_rank(0, ns, [&](int q) { return q == ns - 1 || !clf[q + 1]; },
                             [&](int q, int r) {
                                 idxp[1 + r] = surv[q] + 1;
                                 srank[surv[q]] = -2 - r;  // marks a non-deflated survivor
                             }, sh);
    __syncthreads();
    // chains of rotations
    for (int q = t; q < ns; q += DEF_NT) {
        if (clf[q] || q + 1 >= ns || !clf[q + 1]) continue;
        double zc = z[surv[q]];
        int ctc = ctyp[surv[q]];
        int q2 = q + 1;
        for (; q2 < ns && clf[q2]; q2++) {
            int pj = surv[q2 - 1], qj = surv[q2];
            double sv = zc, cv = z[qj];
            double tau = bd_lapy2(cv, sv);
            double cc = cv / tau, ss = -sv / tau;
            z[pj] = 0.0;
            zc = tau;
            int ctq = ctyp[qj];
            if (ctq != ctc) ctq = 3;
            ctyp[pj] = 4;
            ctc = ctq;
            int rk = W.ti[s + q2];
            int cp = bd_colidx(W, s, nl, pj), cq = bd_colidx(W, s, nl, qj);
            W.rp4[4 * (s + rk)] = W.vmU[M.rd][s + cp] | ((q2 == q + 1) << 30);
            W.rp4[4 * (s + rk) + 1] = W.vmU[M.rd][s + cq];
            W.rp4[4 * (s + rk) + 2] = W.vmV[M.rd][s + cp];
            W.rp4[4 * (s + rk) + 3] = W.vmV[M.rd][s + cq];
            W.rcs[2 * (s + rk)] = cc;
            W.rcs[2 * (s + rk) + 1] = ss;
            W.rotf[s + cp] = 1;
            W.rotf[s + cq] = 1;
        }
        z[surv[q2 - 1]] = zc;
        ctyp[surv[q2 - 1]] = ctc;
    }
    __syncthreads();
    // U2(K,1) = Z(JPREV) of the non-deflated survivors; column-type totals
    if (t < 4) cnt[t] = 0;
    __syncthreads();
    int lc[4] = {0, 0, 0, 0};
    for (int p = 1 + t; p < n; p += DEF_NT) {
        int r = srank[p];
        if (r <= -2) W.zt[s + 1 + (-2 - r)] = z[p];
        lc[ctyp[p] - 1]++;
    }
    for (int j = 0; j < 4; j++)
        if (lc[j]) atomicAdd(&cnt[j], lc[j]);
    __syncthreads();
    const int k = 1 + nk;
    const int ct0 = cnt[0], ct1 = cnt[1], ct2 = cnt[2], ct3 = cnt[3];
    // IDXC: positions J = 2..N grouped by the type of column IDXP(J), stable
    const int psm[4] = {1, 1 + ct0, 1 + ct0 + ct1, 1 + ct0 + ct1 + ct2};  // 0-based starts
    for (int ty = 0; ty < 4; ty++) {
        const int base = psm[ty];
        wave_rank(1, n, [&](int p) { return ctyp[idxp[p] - 1] == ty + 1; },
                  [&](int p, int r) { W.idxc[s + base + r] = p + 1; }, sh);
    }
    __syncthreads();
    // DSIGMA(J) = D(IDXP(J)); slots of U2(:,J) and VT2(J,:) and their supports (the
    // merge block for a column a rotation touched)
    struct GS { double d; int su, sv, ulo, uhi, vlo, vhi; };
    bd_batched<8, DEF_NT>(1, n, [&](int p) {
        GS g;
        g.d = W.D[s + W.idxp[s + p] - 1];
        const int src = bd_colidx(W, s, nl, W.idxp[s + W.idxc[s + p] - 1] - 1);
        g.su = W.vmU[M.rd][s + src];
        g.sv = W.vmV[M.rd][s + src];
        if (W.rotf[s + src]) {
            g.ulo = s; g.uhi = s + n; g.vlo = s; g.vhi = s + m;
        } else {
            g.ulo = W.sulo[g.su]; g.uhi = W.suhi[g.su]; g.vlo = W.svlo[g.sv]; g.vhi = W.svhi[g.sv];
        }
        return g;
    }, [&](int p, const GS &g) {
        W.dsg[s + p] = g.d;
        W.gU[s + p] = g.su;
        W.gV[s + p] = g.sv;
        W.gUlo[s + p] = g.ulo; W.gUhi[s + p] = g.uhi;
        W.gVlo[s + p] = g.vlo; W.gVhi[s + p] = g.vhi;
    });
    __syncthreads();
    if (t == 0) {
        W.dsg[s] = 0.0;
        const double hlftol = tol / 2.0;
        if (n > 1 && fabs(W.dsg[s + 1]) <= hlftol) W.dsg[s + 1] = hlftol;
        double z1 = o.z1, zr1, cq = 1.0, sq = 0.0;
        if (m > n) {
            double zm = W.z[s + m - 1];
            zr1 = bd_lapy2(z1, zm);
            if (zr1 <= tol) {
                zr1 = tol;
            } else {
                cq = z1 / zr1;
                sq = zm / zr1;
            }
        } else {
            zr1 = (fabs(z1) <= tol) ? tol : z1;
        }
        o.zr1 = zr1;
        o.c = cq;
        o.sn = sq;
        o.k = k;
        o.nrot = nrot;
        o.ctot[0] = ct0; o.ctot[1] = ct1; o.ctot[2] = ct2; o.ctot[3] = ct3;
        W.gU[s] = -1;
        W.gV[s] = -1;
        // slots freed for the new columns: the centre column of U, the left null
        // column of V and (M > N) the right one, then the non-deflated inputs
        W.fU[s] = W.vmU[M.rd][s + nl];
        W.fV[s] = W.vmV[M.rd][s + nl];
        if (m > n) W.fV[s + k] = W.vmV[M.rd][s + m - 1];
    }
    __syncthreads();
    // Z(1..K) for the secular equation (mdlasd2.f90:497-500)
    for (int i = t; i < k; i += DEF_NT) {
        W.z[s + i] = (i == 0) ? o.zr1 : W.zt[s + i];
        if (i > 0) {
            W.fU[s + i] = W.gU[s + i];
            W.fV[s + i] = W.gV[s + i];
        }
    }
}

// ------------------------------------------------------------------------------
// the deflation's Givens rotations (DROT at mdlasd2.f90:346-349) on U and V columns
// of the read buffer; thread per row, rotations of a merge in order.  A column met
// for the first time is read within its child's rows (zero elsewhere).
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_rotate(const BdMerge *mg, int nmg, BdTab tab, BdWork W, const BdOut *mo,
                                                   double *U, double *V, long ld)
{
  const int tot = tab.rot[nmg];
  for (int b = blockIdx.x; b < tot; b += gridDim.x) {
    const int mi = bd_find(tab.rot, nmg, b);
    const BdMerge M = mg[mi];
    const int s = M.s, nl = M.nl, n = nl + M.nr + 1, m = n + M.sqre;
    const int r = (b - tab.rot[mi]) * 256 + threadIdx.x;
    if (r >= m) continue;
    const int g = s + r;  // global row
    const int nrot = mo[mi].nrot;
    for (int i = 0; i < nrot; i++) {
        const int *rr = W.rp4 + 4 * (s + i);
        int pu = rr[0] & 0x3fffffff, fresh = rr[0] >> 30, qu = rr[1], pv = rr[2], qv = rr[3];
        double c = W.rcs[2 * (s + i)], sn = W.rcs[2 * (s + i) + 1];
        if (r < n) {
            bool inp = !fresh || (g >= W.sulo[pu] && g < W.suhi[pu]);
            bool inq = g >= W.sulo[qu] && g < W.suhi[qu];
            double x = inp ? U[(long)pu * ld + g] : 0.0;
            double y = inq ? U[(long)qu * ld + g] : 0.0;
            U[(long)pu * ld + g] = c * x + sn * y;
            U[(long)qu * ld + g] = c * y - sn * x;
        }
        bool inp = !fresh || (g >= W.svlo[pv] && g < W.svhi[pv]);
        bool inq = g >= W.svlo[qv] && g < W.svhi[qv];
        double x = inp ? V[(long)pv * ld + g] : 0.0;
        double y = inq ? V[(long)qv * ld + g] : 0.0;
        V[(long)pv * ld + g] = c * x + sn * y;
        V[(long)qv * ld + g] = c * y - sn * x;
    }
  }
}

// ------------------------------------------------------------------------------
// |z| and z / |z| (DLASD3: RHO = DNRM2, DLASCL), one block per merge with K >= 2; the
// norm in T-lane interleaved sums
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_bd_znorm(const BdMerge *mg, BdWork W, BdOut *mo)
{
    const int mi = blockIdx.x;
    const int s = mg[mi].s, k = mo[mi].k, T = sec_team(k), l = threadIdx.x;
    if (k < 2) return;
    double ps = 0.0;
    if (l < T)
        for (int i = l; i < k; i += T) ps += W.z[s + i] * W.z[s + i];
    double rho = __shfl(sqrt(tsum_team(l < T ? ps : 0.0, T)), 0);
    double mul[8];
    const int nm = lascl_muls(rho, 1.0, mul);
    for (int i = l; i < k; i += 64) {
        double zv = W.z[s + i];
        W.zs[s + i] = zv;
        W.z[s + i] = apply_muls(zv, mul, nm);
    }
    if (l == 0) mo[mi].rho = rho * rho;
}

// ------------------------------------------------------------------------------
// secular roots (replaces DLASD4): one team of T lanes per root; the poles of the
// equation in T interleaved partial sums combined by a butterfly
// ------------------------------------------------------------------------------
__device__ __forceinline__ double sv_den(double di, double dorg, double tau)
{
    return ((di - dorg) - tau) * ((di + dorg) + tau);
}

__global__ __launch_bounds__(256) void k_bd_secular(const BdMerge *mg, int nmg, BdTab tab, BdWork W, const BdOut *mo,
                                                    int maxit, int *fail)
{
  const int tot = tab.sec[nmg];
  for (int b = blockIdx.x; b < tot; b += gridDim.x) {
    const int mi = bd_find(tab.sec, nmg, b);
    const int s = mg[mi].s, k = mo[mi].k, T = sec_team(k);
    const int l = threadIdx.x & (T - 1);
    const int j = (b - tab.sec[mi]) * (256 / T) + threadIdx.x / T + 1;  // 1-based root
    if (j > k) continue;
    const double *d = W.dsg + s, *z = W.z + s;
    const double rho = mo[mi].rho, rhoinv = 1.0 / rho;
    // starting iterate
    int ii;
    double t, lo, hi;
    if (j == k) {
        const double dk = d[k - 1];
        hi = rho / (dk + sqrt(dk * dk + rho));
        const double mid = (0.5 * rho) / (dk + sqrt(dk * dk + 0.5 * rho));
        double ps = 0.0;
        for (int i = l; i < k - 2; i += T) ps += z[i] * z[i] / (((d[i] - dk) - mid) * ((d[i] + dk) + mid));
        const double c = rhoinv + tsum_team(ps, T);
        const double dl = (dk - d[k - 2]) * (dk + d[k - 2]);
        const double z1 = z[k - 2] * z[k - 2], z2 = z[k - 1] * z[k - 1];
        double t2;
        if (c <= 0.0) {
            t2 = rho;
        } else {
            double bb = c * dl - z1 - z2;
            double sq = sqrt(fabs(bb * bb + 4.0 * c * z2 * dl));
            if (bb <= 0.0) t2 = (sq - bb) / (2.0 * c);
            else t2 = 2.0 * z2 * dl / (bb + sq);
        }
        t = t2 / (dk + sqrt(dk * dk + t2));
        if (!(t > 0.0 && t <= hi)) t = mid;
        ii = k;
        lo = 0.0;
    } else {
        const double dj = d[j - 1], dj1 = d[j];
        const double delsq = (dj1 - dj) * (dj1 + dj);
        const double sq2 = sqrt((dj * dj + dj1 * dj1) / 2.0);
        const double tm = (delsq / 2.0) / (dj + sq2);
        double ps = 0.0;
        for (int i = l; i < k; i += T) {
            if (i == j - 1 || i == j) continue;
            ps += z[i] * z[i] / (((d[i] - dj) - tm) * ((d[i] + dj) + tm));
        }
        const double c = rhoinv + tsum_team(ps, T);
        const double zl = z[j - 1] * z[j - 1], zr = z[j] * z[j];
        const double fmid = c + (zr - zl) / (delsq / 2.0);
        if (fmid >= 0.0) {
            double bb = c * delsq + zl + zr, cc = zl * delsq;
            double sq = sqrt(fabs(bb * bb - 4.0 * c * cc));
            double t2 = (bb > 0.0) ? 2.0 * cc / (bb + sq) : (bb - sq) / (2.0 * c);
            t = t2 / (dj + sqrt(dj * dj + t2));
            if (!(t > 0.0 && t <= tm)) t = 0.5 * tm;
            ii = j;
            lo = 0.0;
            hi = tm;
        } else {
            double bb = c * delsq - zl - zr, cc = zr * delsq;
            double sq = sqrt(fabs(bb * bb + 4.0 * c * cc));
            double t2 = (bb <= 0.0) ? -2.0 * cc / (sq - bb) : -(bb + sq) / (2.0 * c);
            lo = -(delsq / 2.0) / (dj1 + sq2);
            t = t2 / (dj1 + sqrt(fabs(dj1 * dj1 + t2)));
            if (!(t >= lo && t < 0.0)) t = 0.5 * lo;
            ii = j + 1;
            hi = 0.0;
        }
    }
    const double dorg = d[ii - 1], dorg2 = dorg + dorg;
    const int split = (j == k) ? k - 1 : j;
    const int na = split - ((ii - 1 < split) ? 1 : 0), nb = (k - split) - ((ii - 1 >= split) ? 1 : 0);
    int lgT = 0;
    while ((1 << lgT) < T) lgT++;
    int it, info = 1;
    for (it = 0; it <= maxit; it++) {
        double psi = 0.0, dpsi = 0.0, epsi = 0.0, phi = 0.0, dphi = 0.0, ephi = 0.0;
        for (int t2 = l; t2 < na; t2 += T) {
            double dn = ((d[t2] - dorg) - t) * ((d[t2] + dorg) + t);
            double q = z[t2] / dn;
            psi += z[t2] * q;
            dpsi += q * q;
            epsi += psi;
        }
        for (int t2 = l; t2 < nb; t2 += T) {
            int i = k - 1 - t2;
            double dn = ((d[i] - dorg) - t) * ((d[i] + dorg) + t);
            double q = z[i] / dn;
            phi += z[i] * q;
            dphi += q * q;
            ephi += phi;
        }
        psi = tsum_team(psi, T);
        dpsi = tsum_team(dpsi, T);
        epsi = tsum_team(epsi, T);
        phi = tsum_team(phi, T);
        dphi = tsum_team(dphi, T);
        ephi = tsum_team(ephi, T);
        const double mu = t * (dorg2 + t);
        const double qo = z[ii - 1] / (-mu);
        const double torg = z[ii - 1] * qo;
        const double w = ((rhoinv + phi) + psi) + torg;
        const double erretm = (8.0 + lgT) * (phi - psi) + (ephi - epsi) + 2.0 * rhoinv + 3.0 * fabs(torg) +
                              fabs(mu) * ((dpsi + dphi) + qo * qo);
        if (ii - 1 < split) dpsi += qo * qo; else dphi += qo * qo;
        if (fabs(w) <= BD_EPS * erretm || it == maxit) {
            if (fabs(w) <= BD_EPS * erretm) info = 0;
            break;
        }
        if (w > 0.0) hi = t; else lo = t;
        const double dl = ((d[split - 1] - dorg) - t) * ((d[split - 1] + dorg) + t);
        const double dr = ((d[split] - dorg) - t) * ((d[split] + dorg) + t);
        const double a = (dl + dr) * w - dl * dr * (dpsi + dphi);
        const double bq = dl * dr * w;
        const double c = w - dl * dpsi - dr * dphi;
        double eta;
        if (c == 0.0) {
            eta = (a != 0.0) ? bq / a : -w / (dpsi + dphi);
        } else {
            double sq = sqrt(fabs(a * a - 4.0 * bq * c));
            if (a <= 0.0) eta = (a - sq) / (2.0 * c);
            else eta = 2.0 * bq / (a + sq);
        }
        if (w * eta >= 0.0) eta = -w / (dpsi + dphi);
        const double sg = dorg + t;
        double tn = t + eta / (sg + sqrt(fabs(sg * sg + eta)));
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t) { info = 0; break; }
        t = tn;
    }
    if (l == 0) {
        W.sig[s + j - 1] = dorg + t;
        W.org[s + j - 1] = ii;
        W.tau[s + j - 1] = t;
        atomicAdd(W.iters, (unsigned long long)(it + 1));
        if (info) atomicAdd(fail, 1);
    }
  }
}

// ------------------------------------------------------------------------------
// updated z (DLASD3, Loewner form): thread per index, products in DLASD3's order.
// DLASD3's two loops (J < I with pole D(J), J >= I with D(J+1)) are one product over
// J = 1 .. K-1 ascending; each lane picks the pole by its own I.  The block stages
// chunks of (D(ORG(J)), TAU(J), D(J)) in LDS and forms BL_U factors at a time before
// multiplying them in order: a level launches one thread per index, too few
// wavefronts to hide the latency of one factor at a time.
// ------------------------------------------------------------------------------
#define BL_LDS 2048
#define BL_U 8
__global__ __launch_bounds__(256) void k_bd_loewner(const BdMerge *mg, int nmg, BdTab tab, BdWork W, const BdOut *mo)
{
  __shared__ double so[BL_LDS], st[BL_LDS], sdd[BL_LDS + 1];
  const int tot = tab.low[nmg];
  for (int b = blockIdx.x; b < tot; b += gridDim.x) {
    const int mi = bd_find(tab.low, nmg, b);  // uniform over the block
    const int s = mg[mi].s, k = mo[mi].k;
    const int i = (b - tab.low[mi]) * 256 + threadIdx.x + 1;  // 1-based
    const bool own = i <= k;
    const double *ds = W.dsg + s;
    const double di = own ? ds[i - 1] : 0.0;
    double zi = own ? sv_den(di, ds[W.org[s + k - 1] - 1], W.tau[s + k - 1]) : 1.0;
    for (int c0 = 1; c0 <= k - 1; c0 += BL_LDS) {
      const int cn = min(BL_LDS, k - c0);  // J = c0 .. c0 + cn - 1
      for (int t = threadIdx.x; t < cn; t += 256) {
        const int j = c0 + t;
        so[t] = ds[W.org[s + j - 1] - 1];
        st[t] = W.tau[s + j - 1];
        sdd[t] = ds[j - 1];
      }
      if (threadIdx.x == 0) sdd[cn] = ds[c0 + cn - 1];
      __syncthreads();
      if (own) {
        // pole of J: D(J) = sdd[t] below I, D(J+1) = sdd[t + 1] from I on
        const int sw = i - c0;  // first t with J >= I
        int t = 0;
        for (; t + BL_U <= cn; t += BL_U) {
          double f[BL_U];
#pragma unroll
          for (int u = 0; u < BL_U; u++) {
            const double dd = (t + u < sw) ? sdd[t + u] : sdd[t + u + 1];
            f[u] = sv_den(di, so[t + u], st[t + u]) / (di - dd) / (di + dd);
          }
#pragma unroll
          for (int u = 0; u < BL_U; u++) zi = zi * f[u];
        }
        for (; t < cn; t++) {
          const double dd = (t < sw) ? sdd[t] : sdd[t + 1];
          zi = zi * (sv_den(di, so[t], st[t]) / (di - dd) / (di + dd));
        }
      }
      __syncthreads();  // the next chunk or block restages
    }
    if (own) {
      const double a = sqrt(fabs(zi));
      W.zl[s + i - 1] = W.zs[s + i - 1] >= 0.0 ? a : -a;
    }
  }
}

// ------------------------------------------------------------------------------
// singular vectors of the deflated problem (DLASD3): team of T lanes per root i;
// column i of Bu = left vector (rows in IDXC order), column i of Bv = row i of the
// right-vector matrix
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_vec(const BdMerge *mg, int nmg, BdTab tab, BdWork W, const BdOut *mo,
                                                double *Q)
{
  const long *qoff = tab.qoff;
  const int tot = tab.sec[nmg];
  for (int b = blockIdx.x; b < tot; b += gridDim.x) {
    const int mi = bd_find(tab.sec, nmg, b);
    if (mo[mi].hss) continue;  // its vectors come from the HSS forms
    const int s = mg[mi].s, k = mo[mi].k, T = sec_team(k);
    const int l = threadIdx.x & (T - 1);
    const int i = (b - tab.sec[mi]) * (256 / T) + threadIdx.x / T + 1;
    if (i > k) continue;
    const double *ds = W.dsg + s, *zl = W.zl + s;
    const int *idxc = W.idxc + s;
    const double dorg = ds[W.org[s + i - 1] - 1], ti = W.tau[s + i - 1];
    double *bu = Q + qoff[mi] + (long)(i - 1) * k;
    double *bv = Q + qoff[mi] + (long)k * k + (long)(i - 1) * k;
    double pu = 0.0, pv = 0.0;
    for (int j = l; j < k; j += T) {
        double dj = ds[j];
        double wv = zl[j] / ((dj - dorg) - ti) / ((dj + dorg) + ti);
        double x = j == 0 ? -1.0 : dj * wv;
        pu += x * x;
        pv += wv * wv;
    }
    const double tu = sqrt(tsum_team(pu, T)), tv = sqrt(tsum_team(pv, T));
    for (int j = l; j < k; j += T) {
        if (j == 0) {
            bu[0] = -1.0 / tu;
            double d0 = ds[0];
            double w0 = zl[0] / ((d0 - dorg) - ti) / ((d0 + dorg) + ti);
            bv[0] = w0 / tv;
        } else {
            int jc = idxc[j] - 1;
            double dj = ds[jc];
            double wv = zl[jc] / ((dj - dorg) - ti) / ((dj + dorg) + ti);
            bu[j] = dj * wv / tu;
            bv[j] = wv / tv;
        }
    }
  }
}

// ------------------------------------------------------------------------------
// U2 / VT2 column J (0-based) of a merge at local row r: the gathered child column
// inside its support, the special first columns (mdlasd2.f90:502-518)
// ------------------------------------------------------------------------------
__device__ __forceinline__ double bd_acol(const BdWork &W, const BdMerge &M, const BdOut &o, bool isV,
                                          const double *X, long ld, int J, int r)
{
    const int s = M.s, nl = M.nl, n = nl + M.nr + 1, m = n + M.sqre, g = s + r;
    if (J == 0) {
        if (!isV) return r == nl ? 1.0 : 0.0;
        const int *vm = W.vmV[M.rd];
        if (m > n) return r <= nl ? o.c * X[(long)vm[s + nl] * ld + g] : o.sn * X[(long)vm[s + m - 1] * ld + g];
        return r <= nl ? X[(long)vm[s + nl] * ld + g] : 0.0;
    }
    if (isV) return (g >= W.gVlo[s + J] && g < W.gVhi[s + J]) ? X[(long)W.gV[s + J] * ld + g] : 0.0;
    return (g >= W.gUlo[s + J] && g < W.gUhi[s + J]) ? X[(long)W.gU[s + J] * ld + g] : 0.0;
}

// ------------------------------------------------------------------------------
// DLASD3's products U(:,1:K) = U2 Bu, V(:,1:K) = VT2^T Bv: 64 x 64 output tiles, 16
// inner indices staged per step, each output summed over the inner index in
// ascending order
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_gemm(const BdMerge *mg, int nmg, BdTab tab, BdWork W, const BdOut *mo,
                                                 const double *Q, const double *U, const double *V, long ld,
                                                 double *stg)
{
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int ttot = tab.tile[nmg];
  for (int bi = blockIdx.x; bi < ttot; bi += gridDim.x) {
    // tiles of a merge: for every 64 output columns, the U row tiles then the V row tiles
    const int mi = bd_find(tab.tile, nmg, bi);
    const BdMerge M = mg[mi];
    const BdOut &o = mo[mi];
    const int n = M.nl + M.nr + 1, m = n + M.sqre, k = o.k;
    const int ntu = (n + 63) / 64, ntv = (m + 63) / 64, lt = bi - tab.tile[mi];
    const int rem = lt % (ntu + ntv);
    const bool isV = rem >= ntu;
    const int row0 = (isV ? rem - ntu : rem) * 64, col0 = (lt / (ntu + ntv)) * 64;
    const int rows = isV ? m : n;
    const double *X = isV ? V : U;
    double *Y = stg + tab.soff[2 * mi + (isV ? 1 : 0)];
    const double *B = Q + tab.qoff[mi] + (isV ? (long)k * k : 0);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
    for (int a = 0; a < 4; a++)
        for (int c = 0; c < 4; c++) acc[a][c] = 0.0;
    for (int j0 = 0; j0 < k; j0 += 16) {
        for (int e = threadIdx.x; e < 1024; e += 256) {
            int rr = e & 63, jj = e >> 6;
            int r = row0 + rr, J = j0 + jj;
            As[jj][rr] = (r < rows && J < k) ? bd_acol(W, M, o, isV, X, ld, J, r) : 0.0;
            int cc = e & 63, j2 = e >> 6;
            int col = col0 + cc, J2 = j0 + j2;
            Bs[j2][cc] = (col < k && J2 < k) ? B[(long)col * k + J2] : 0.0;
        }
        __syncthreads();
        const int jn = min(16, k - j0);
        for (int jj = 0; jj < jn; jj++) {
            double av[4], bv[4];
            for (int a = 0; a < 4; a++) av[a] = As[jj][tx + 16 * a];
            for (int c = 0; c < 4; c++) bv[c] = Bs[jj][ty + 16 * c];
            for (int a = 0; a < 4; a++)
                for (int c = 0; c < 4; c++) acc[a][c] = acc[a][c] + av[a] * bv[c];
        }
        __syncthreads();
    }
    for (int a = 0; a < 4; a++) {
        int r = row0 + tx + 16 * a;
        if (r >= rows) continue;
        for (int c = 0; c < 4; c++) {
            int col = col0 + ty + 16 * c;
            if (col < k) Y[(long)col * rows + r] = acc[a][c];
        }
    }
  }
}

// ------------------------------------------------------------------------------
// new columns formed without a product, into staging: U(:,1) and VT(1,:) when K = 1
// (DLASD3's quick return) and the last row of VT when M > N (mdlasd2.f90:504-512)
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_newcol(const BdMerge *mg, BdTab tab, BdWork W, const BdOut *mo,
                                                   const double *U, const double *V, long ld, double *stg)
{
    const int mi = blockIdx.x;
    const BdMerge M = mg[mi];
    const BdOut &o = mo[mi];
    const int s = M.s, nl = M.nl, n = nl + M.nr + 1, m = n + M.sqre, k = o.k;
    if (!(k == 1 || m > n)) return;
    double *SU = stg + tab.soff[2 * mi], *SV = stg + tab.soff[2 * mi + 1];
    const int *vm = W.vmV[M.rd];
    for (int r = threadIdx.x; r < m; r += 256) {
        if (k == 1) {
            if (r < n) {
                double u = bd_acol(W, M, o, false, U, ld, 0, r);
                SU[r] = o.zr1 > 0.0 ? u : -u;
            }
            SV[r] = bd_acol(W, M, o, true, V, ld, 0, r);
        }
        if (m > n) {
            const int g = s + r;
            SV[(long)k * m + r] = r <= nl ? -o.sn * V[(long)vm[s + nl] * ld + g] : o.c * V[(long)vm[s + m - 1] * ld + g];
        }
    }
}

// staging -> the slots freed by the merge (rows of the merge block)
__global__ __launch_bounds__(256) void k_bd_scatter(const BdMerge *mg, int nmg, BdTab tab, BdWork W,
                                                    const BdOut *mo, const double *stg, double *U, double *V, long ld)
{
  const int tot = tab.sc[nmg];
  for (int b = blockIdx.x; b < tot; b += gridDim.x) {
    // a merge's new columns: the K of U, then the K + sqre of V
    const int mi = bd_find(tab.sc, nmg, b);
    const BdMerge M = mg[mi];
    const int k = mo[mi].k, j = b - tab.sc[mi], isV = j >= k, i = isV ? j - k : j;
    const int s = M.s, n = M.nl + M.nr + 1, rows = isV ? n + M.sqre : n;
    const double *S = stg + tab.soff[2 * mi + isV] + (long)i * rows;
    double *Y = isV ? V + (long)W.fV[s + i] * ld : U + (long)W.fU[s + i] * ld;
    for (int r = threadIdx.x; r < rows; r += 256) Y[s + r] = S[r];
  }
}

// ------------------------------------------------------------------------------
// HSS branch (mdlasd1.f90:220-278, MDLASD8 mdlasd8.f90:201-278): per root i of an HSS
// merge the normalisers of its left and right vectors (the dense branch's sums),
// DIFL(i) = sigma_i - d_i, DIFR(i) = sigma_i - d_{i+1}, and d_i z_i
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_hssgen(const BdMerge *mg, const BdHss *hs, BdWork W, const BdOut *mo,
                                                   double *G)
{
    const BdHss h = hs[blockIdx.y];
    const int s = mg[h.mi].s, k = mo[h.mi].k, T = sec_team(k);
    const int l = threadIdx.x & (T - 1);
    const int i = blockIdx.x * (256 / T) + threadIdx.x / T + 1;
    if (i > k) return;
    const double *ds = W.dsg + s, *zl = W.zl + s;
    const double dorg = ds[W.org[s + i - 1] - 1], ti = W.tau[s + i - 1];
    double pu = 0.0, pv = 0.0;
    for (int j = l; j < k; j += T) {
        double dj = ds[j];
        double wv = zl[j] / ((dj - dorg) - ti) / ((dj + dorg) + ti);
        double x = j == 0 ? -1.0 : dj * wv;
        pu += x * x;
        pv += wv * wv;
    }
    const double tu = sqrt(tsum_team(pu, T)), tv = sqrt(tsum_team(pv, T));
    if (l == 0) {
        double *g = G + h.goff;
        g[i - 1] = ds[i - 1] * zl[i - 1];
        g[k + i - 1] = 1.0 / tu;
        g[2 * k + i - 1] = 1.0 / tv;
        g[3 * k + i - 1] = -((ds[i - 1] - dorg) - ti);
        g[4 * k + i - 1] = i < k ? -((ds[i] - dorg) - ti) : 0.0;
    }
}

// U2 and V2 = VT2^T of an HSS merge gathered with their columns in DSIGMA order
// (column idxc(J) takes U2(:,J)), zero outside the supports
__global__ __launch_bounds__(256) void k_bd_xgather(const BdMerge *mg, const BdHss *hs, BdWork W, const BdOut *mo,
                                                    const double *U, const double *V, long ld, double *X)
{
    const BdHss h = hs[blockIdx.y];
    const BdMerge M = mg[h.mi];
    const BdOut &o = mo[h.mi];
    const int s = M.s, n = M.nl + M.nr + 1, m = n + M.sqre, k = o.k;
    double *XU = X + h.xoff, *XV = XU + (long)n * k;
    // one output column per pass of the block: columns 0 .. k-1 of XU, then of XV
    for (int cb = blockIdx.x; cb < 2 * k; cb += gridDim.x) {
        const bool isV = cb >= k;
        const int J = isV ? cb - k : cb, rows = isV ? m : n;
        const int c = J == 0 ? 0 : W.idxc[s + J] - 1;
        double *dst = (isV ? XV : XU) + (long)c * rows;
        const double *src = isV ? V : U;
        for (int r = threadIdx.x; r < rows; r += 256) dst[r] = bd_acol(W, M, o, isV, src, ld, J, r);
    }
}

// U(:,1:K) = U2 H_U - U2(:,1) APHAU^T: U2(:,1) is the unit vector at the centre row
__global__ __launch_bounds__(256) void k_bd_hssfix(const BdMerge *mg, const BdHss *hs, BdTab tab, const BdOut *mo,
                                                   const double *G, double *stg)
{
    const BdHss h = hs[blockIdx.y];
    const BdMerge M = mg[h.mi];
    const int n = M.nl + M.nr + 1, k = mo[h.mi].k;
    double *SU = stg + tab.soff[2 * h.mi];
    const double *au = G + h.goff + k;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < k; i += gridDim.x * 256)
        SU[(long)i * n + M.nl] -= au[i];
}

// exclusive block scan of one 64-bit value per thread (1024 threads); returns the total
__device__ long block_exscan_l(long v, long *sh, long &total)
{
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, nw = blockDim.x >> 6;
    long x = v;
    for (int off = 1; off < 64; off <<= 1) {
        long y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (t == 0) {
        long acc = 0;
        for (int w = 0; w < nw; w++) {
            long c = sh[w];
            sh[w] = acc;
            acc += c;
        }
        sh[nw] = acc;
    }
    __syncthreads();
    long r = sh[wv] + x - v;
    total = sh[nw];
    __syncthreads();
    return r;
}

// An HSS tree over K columns in leaves of about ni has 2 to 512 leaves (the
// construction's limits; npart, aux_hss.f90:737-760).
__host__ __device__ inline bool bd_hss_tree_ok(int k, int ni)
{
    int l = k / ni;
    if (l < 1) return false;
    if (k % ni >= ni / 2) l++;
    return l >= 2 && l <= 512;
}

// the level's launch tables from its deflation results: per-merge work counts and
// their exclusive prefix sums, 1024 merges per pass; a merge with K >= hss_min_k (and
// a valid HSS tree) is flagged for the HSS branch here, on the device
__global__ __launch_bounds__(1024) void k_bd_lvltab(const BdMerge *mg, int nmg, BdOut *mo, BdTab tab, int hss_min_k,
                                                    int hss_leaf)
{
    __shared__ long sh[1024 / 64 + 1];
    long base[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i0 = 0; i0 < nmg; i0 += 1024) {
        const int i = i0 + threadIdx.x;
        long v[7] = {0, 0, 0, 0, 0, 0, 0};  // rot, sec, low, tile, sc, q, staging
        if (i < nmg) {
            const BdMerge M = mg[i];
            const int n = M.nl + M.nr + 1, m = n + M.sqre, k = mo[i].k;
            mo[i].hss = (k >= hss_min_k && bd_hss_tree_ok(k, hss_leaf)) ? 1 : 0;
            if (mo[i].nrot) v[0] = (m + 255) / 256;
            if (k >= 2) {
                const int T = sec_team(k);
                v[1] = (k + 256 / T - 1) / (256 / T);
                v[2] = (k + 255) / 256;
                if (!mo[i].hss) {  // the HSS branch forms no K x K vectors and no tiles
                    v[3] = (long)((k + 63) / 64) * ((n + 63) / 64 + (m + 63) / 64);
                    v[5] = 2 * (long)k * k;
                }
            }
            v[4] = 2 * k + M.sqre;
            v[6] = (long)n * k + (long)m * (k + M.sqre);
        }
        long ex[7], tot[7];
        for (int q = 0; q < 7; q++) ex[q] = base[q] + block_exscan_l(v[q], sh, tot[q]);
        if (i < nmg) {
            const BdMerge M = mg[i];
            const int n = M.nl + M.nr + 1, k = mo[i].k;
            tab.rot[i] = (int)ex[0]; tab.sec[i] = (int)ex[1]; tab.low[i] = (int)ex[2]; tab.tile[i] = (int)ex[3];
            tab.sc[i] = (int)ex[4];
            tab.qoff[i] = ex[5];
            tab.soff[2 * i] = ex[6];
            tab.soff[2 * i + 1] = ex[6] + (long)n * k;
        }
        for (int q = 0; q < 7; q++) base[q] += tot[q];
    }
    if (threadIdx.x == 0) {
        tab.rot[nmg] = (int)base[0]; tab.sec[nmg] = (int)base[1]; tab.low[nmg] = (int)base[2];
        tab.tile[nmg] = (int)base[3]; tab.sc[nmg] = (int)base[4];
    }
}

// ------------------------------------------------------------------------------
// merge step 4 (block per merge): D(1:K) = roots, D(K+1:N) = deflated values, both
// unscaled, IDXQ = DLAMRG(K, N-K, D, 1, -1) (mdlasd1.f90:282-286)
// ------------------------------------------------------------------------------
#define POST_NT 1024
__global__ __launch_bounds__(POST_NT) void k_bd_post(const BdMerge *mg, BdWork W, const BdOut *mo, int lds_n)
{
    extern __shared__ double lruns[];
    const BdMerge M = mg[blockIdx.x];
    const BdOut &o = mo[blockIdx.x];
    const int s = M.s, n = M.nl + M.nr + 1, k = o.k, t = threadIdx.x;
    double mul[8];
    const int nm = lascl_muls(1.0, o.orgnrm, mul);
    bd_batched<8, POST_NT>(0, n, [&](int i) {
        return (i < k) ? ((k == 1) ? fabs(o.zr1) : W.sig[s + i]) : W.dsg[s + i];
    }, [&](int i, double v) { W.D[s + i] = apply_muls(v, mul, nm); });
    __syncthreads();
    // DLAMRG(K, N-K, D, 1, -1, IDXQ): run 2 read backwards; merge path over D staged in
    // LDS when it fits
    const bool inl = n <= lds_n;
    if (inl) {
        for (int i = t; i < n; i += POST_NT) lruns[i] = W.D[s + i];
        __syncthreads();
    }
    {
        const double *a = inl ? lruns : W.D + s;
        const int n2 = n - k, per = (n + POST_NT - 1) / POST_NT;
        const int o0 = min(t * per, n), o1 = min(o0 + per, n);
        auto bget = [a, n](int j) { return a[n - 1 - j]; };
        int i = corank(o0, a, k, bget, n2), j = o0 - i;
        for (int o = o0; o < o1; o++) {
            if (i < k && (j >= n2 || a[i] <= a[n - 1 - j])) W.idxq[s + o] = ++i;
            else W.idxq[s + o] = n - (j++);
        }
    }
    // the block's new maps: new columns in freed slots (supported on the whole block),
    // deflated ones where they were (the block once a rotation touched them)
    const int m = n + M.sqre;
    int *vu = W.vmU[M.wr], *vv = W.vmV[M.wr];
    struct MS { int su, sv, ulo, uhi, vlo, vhi; };
    bd_batched<8, POST_NT>(0, n, [&](int i) {
        MS g;
        if (i < k) {
            g.su = W.fU[s + i]; g.sv = W.fV[s + i];
            g.ulo = s; g.uhi = s + n; g.vlo = s; g.vhi = s + m;
        } else {
            g.su = W.gU[s + i]; g.sv = W.gV[s + i];
            g.ulo = W.gUlo[s + i]; g.uhi = W.gUhi[s + i]; g.vlo = W.gVlo[s + i]; g.vhi = W.gVhi[s + i];
        }
        return g;
    }, [&](int i, const MS &g) {
        W.sulo[g.su] = g.ulo; W.suhi[g.su] = g.uhi;
        W.svlo[g.sv] = g.vlo; W.svhi[g.sv] = g.vhi;
        vu[s + i] = g.su;
        vv[s + i] = g.sv;
    });
    if (t == 0 && m > n) {
        int sv = W.fV[s + k];
        vv[s + m - 1] = sv;
        W.svlo[sv] = s;
        W.svhi[sv] = s + m;
    }
}

// ------------------------------------------------------------------------------
// output: zeros everywhere (second stream, beside the levels), then column i of U
// and row i of VT from the slot of virtual column perm(i), supported rows only
// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bd_zero(double *A, long lda, int n)
{
    const long tot = (long)n * n;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
        const long c = e / n, r = e - c * n;
        A[c * lda + r] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_bd_fin(const BdWork W, const double *Us, const double *Vs, long ld,
                                                const int *perm, int n, double *U, long ldu)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const int su = W.vmU[0][perm[i]], lo = W.sulo[su], hi = W.suhi[su];
    for (int r = lo + threadIdx.x; r < hi; r += 256) U[(long)i * ldu + r] = Us[(long)su * ld + r];
}

// VT(i, r) = V(r, slot of i) for r inside the support of i (the rest of VT is already
// zero): 64 x 64 tiles through LDS, read down V's columns and written along VT's
// columns, both coalesced (a thread per r would store with a stride of ldvt); a tile
// that meets no support exits after its 64 lookups.
__global__ __launch_bounds__(256) void k_bd_fin_vt(const BdWork W, const double *Vs, long ld, const int *perm, int n,
                                                   double *VT, long ldvt)
{
    __shared__ double t[64][65];
    __shared__ long sb[64];
    __shared__ int slo[64], shi[64], any;
    const int i0 = blockIdx.x * 64, r0 = blockIdx.y * 64, tid = threadIdx.x;
    if (tid == 0) any = 0;
    __syncthreads();
    if (tid < 64) {
        const int i = i0 + tid;
        int lo = 0, hi = 0;
        long b = 0;
        if (i < n) {
            const int sv = W.vmV[0][perm[i]];
            lo = max(W.svlo[sv], r0);
            hi = min(W.svhi[sv], min(r0 + 64, n));
            b = (long)sv * ld;
        }
        slo[tid] = lo; shi[tid] = hi; sb[tid] = b;
        if (lo < hi) any = 1;
    }
    __syncthreads();
    if (!any) return;
    const int q = tid & 63;
    for (int c = tid >> 6; c < 64; c += 4) {  // column i0 + c of the tile, row r0 + q
        const int r = r0 + q;
        if (r >= slo[c] && r < shi[c]) t[c][q] = Vs[sb[c] + r];
    }
    __syncthreads();
    for (int rr = tid >> 6; rr < 64; rr += 4) {  // row r0 + rr of V, VT entry (i0 + q, r0 + rr)
        const int r = r0 + rr;
        if (r >= slo[q] && r < shi[q]) VT[(long)r * ldvt + i0 + q] = t[q][rr];
    }
}

// MDBDSDC's final order for one subproblem (mdbdsdc.f90:383-404): with distinct values
// the selection sort puts the i-th largest singular value (and its vectors) at
// position i, i.e. the top merge's ascending IDXQ read backwards.  Writes the order,
// the unscaled values in it, and flags any two equal values (the selection sort's
// order of equal values is then reproduced on the host).
__global__ __launch_bounds__(1024) void k_bd_perm(const BdWork W, int n, double orgnrm, int scaled, int *perm,
                                                  double *Dout, int *tie)
{
    double mul[8];
    const int nm = scaled ? lascl_muls(1.0, orgnrm, mul) : 0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const int src = W.idxq[n - 1 - i] - 1;
        const double v = apply_muls(W.D[src], mul, nm);
        perm[i] = src;
        Dout[i] = v;
        if (i + 1 < n && v == apply_muls(W.D[W.idxq[n - 2 - i] - 1], mul, nm)) atomicOr(tie, 1);
    }
}

// identity maps of U's slots (the centre columns map to themselves)
static __global__ void k_bd_iota(int *a, int *b, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = b[i] = i;
}

// DLASR('L','V','B') on the columns of U: the rotations that made a lower bidiagonal
// upper (mdbdsdc.f90:413-414); thread per column
__global__ __launch_bounds__(256) void k_bd_lasr(double *U, long ldu, int n, const double *cs, const double *sn)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    double *u = U + (long)c * ldu;
    for (int j = n - 2; j >= 0; j--) {
        double ct = cs[j], st = sn[j];
        double t = u[j + 1];
        u[j + 1] = ct * t - st * u[j];
        u[j] = st * t + ct * u[j];
    }
}

// ------------------------------------------------------------------------------
// host driver
// ------------------------------------------------------------------------------
namespace {

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

int bensure(Buf &b, size_t bytes)
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

inline int bcdiv(long a, long b) { return (int)((a + b - 1) / b); }
inline size_t al8(size_t x) { return (x + 7) & ~(size_t)7; }

// LAPACK DLASDT: the tree of one subproblem (node arrays 0-based, values 1-based)
void lasdt(int n, int &lvl, int &nd, std::vector<int> &inode, std::vector<int> &ndiml, std::vector<int> &ndimr,
           int msub)
{
    int maxn = n > 1 ? n : 1;
    double temp = std::log((double)maxn / (double)(msub + 1)) / std::log(2.0);
    lvl = (int)temp + 1;
    inode.assign(n, 0);
    ndiml.assign(n, 0);
    ndimr.assign(n, 0);
    int i = n / 2;
    inode[0] = i + 1;
    ndiml[0] = i;
    ndimr[0] = n - i - 1;
    int il = -1, ir = 0, llst = 1;
    for (int l = 1; l <= lvl - 1; l++) {
        for (i = 0; i < llst; i++) {
            il += 2;
            ir += 2;
            int nc = llst + i - 1;
            ndiml[il] = ndiml[nc] / 2;
            ndimr[il] = ndiml[nc] - ndiml[il] - 1;
            inode[il] = inode[nc] - ndimr[il] - 1;
            ndiml[ir] = ndimr[nc] / 2;
            ndimr[ir] = ndimr[nc] - ndiml[ir] - 1;
            inode[ir] = inode[nc] + ndiml[ir] + 1;
        }
        llst *= 2;
    }
    nd = llst * 2 - 1;
}

// LAPACK DLARTG (3.10 form)
void lartg(double f, double g, double &c, double &s, double &r)
{
    const double rtmin = std::sqrt(BD_SAFMIN), rtmax = std::sqrt(1.0 / BD_SAFMIN / 2.0);
    double f1 = std::fabs(f), g1 = std::fabs(g);
    if (g == 0.0) {
        c = 1.0; s = 0.0; r = f;
    } else if (f == 0.0) {
        c = 0.0; s = g > 0.0 ? 1.0 : -1.0; r = g1;
    } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
        double d = std::sqrt(f * f + g * g);
        c = f1 / d;
        r = f > 0.0 ? d : -d;
        s = g / r;
    } else {
        double u = std::fmin(1.0 / BD_SAFMIN, std::fmax(BD_SAFMIN, std::fmax(f1, g1)));
        double fs = f / u, gs = g / u;
        double d = std::sqrt(fs * fs + gs * gs);
        c = std::fabs(fs) / d;
        r = fs > 0.0 ? d : -d;
        s = gs / r;
        r *= u;
    }
}

void host_lascl(double cfrom, double cto, int n, double *x)
{
    double mul[8];
    int nm = lascl_muls(cfrom, cto, mul);
    for (int s = 0; s < nm; s++)
        for (int i = 0; i < n; i++) x[i] *= mul[s];
}

}  // namespace

struct hsdc_bd_ctx {
    hsdc_bd_params prm;
    hsdc_bd_stats st;
    hipStream_t stream = 0;
    Buf ub[1], vb[1], wd, wi, mgb, mob, tab, qb, stg, misc, din, din2, ein, uo, vto, rot;
    hipStream_t zstream = nullptr;     // zero fill of the outputs beside the levels
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // solve start / end, fill fork / join
    std::vector<int> plan_key;   // shape of the tree tables on the device
    void *plan_buf = nullptr;
    double *hde = nullptr;       // pinned staging of d and e
    size_t hde_cap = 0;
    BdWork W;
    long wcap = 0;
    BdOut *hmo = nullptr;  // pinned
    size_t hmo_cap = 0;
    std::vector<int> tr;   // per merge of the last solve: block start, order, K, rotations, HSS rank
    std::vector<int> hrank;  // per merge: largest HSS generator rank (0: dense branch)
    hsdc_ctx *hctx = nullptr;  // HSS construction and multiply (created on the first HSS merge)
    hipEvent_t evh = nullptr;  // a level's deflation results are on the host
    Buf hgen, hx, htb;         // HSS generators, gathered U2 / V2, the level's BdHss table
};

#define BCK(x) do { int _r = (x); if (_r) return _r; } while (0)
#define BHK(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); return HSDC_ERR_HIP; } } while (0)

extern "C" {

void hsdc_bd_default_params(hsdc_bd_params *p)
{
    p->smlsiz = 25;
    p->maxit = 100;
    p->hss_min_k = 2000;              // BSMLZ, mdlasd2.f90:172
    p->hss_leaf = 40;                 // Ni, mdlasd8.f90:277
    p->hss_tol = (double)1.0e-17f;    // TOL, CauchyHssvd_VPS.f90:3694 (a REAL literal)
}

int hsdc_bd_ctx_create(hsdc_bd_ctx **ctx)
{
    if (!ctx) return HSDC_ERR_INTERNAL;
    hsdc_bd_ctx *c = new hsdc_bd_ctx();
    hsdc_bd_default_params(&c->prm);
    std::memset(&c->st, 0, sizeof c->st);
    *ctx = c;
    return 0;
}

int hsdc_bd_ctx_destroy(hsdc_bd_ctx *c)
{
    if (!c) return 0;
    Buf *all[] = {&c->ub[0], &c->vb[0], &c->wd, &c->wi, &c->mgb, &c->mob, &c->tab, &c->qb, &c->stg,
                  &c->misc, &c->din, &c->din2, &c->ein, &c->uo, &c->vto, &c->rot, &c->hgen, &c->hx, &c->htb};
    if (c->hctx) hsdc_ctx_destroy(c->hctx);
    if (c->evh) (void)hipEventDestroy(c->evh);
    for (Buf *b : all)
        if (b->p) (void)hipFree(b->p);
    for (int i = 0; i < 4; i++)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->zstream) (void)hipStreamDestroy(c->zstream);
    if (c->hmo) (void)hipHostFree(c->hmo);
    if (c->hde) (void)hipHostFree(c->hde);
    delete c;
    return 0;
}

int hsdc_bd_set_params(hsdc_bd_ctx *c, const hsdc_bd_params *p)
{
    if (!c || !p) return HSDC_ERR_INTERNAL;
    if (p->smlsiz < 3 || p->smlsiz > 63) return -2;
    if (p->maxit < 1) return -3;
    if (p->hss_min_k < 1) return -4;
    if (p->hss_leaf < 8) return -5;
    if (!(p->hss_tol >= 0.0)) return -6;
    c->prm = *p;
    return 0;
}

int hsdc_bd_get_stats(hsdc_bd_ctx *c, hsdc_bd_stats *s)
{
    if (!c || !s) return HSDC_ERR_INTERNAL;
    *s = c->st;
    return 0;
}

int hsdc_bd_set_stream(hsdc_bd_ctx *c, void *stream)
{
    if (!c) return HSDC_ERR_INTERNAL;
    c->stream = (hipStream_t)stream;
    return 0;
}

int hsdc_bd_trace(hsdc_bd_ctx *c, int *s, int *n, int *k, int *nrot, int *hss_rank, int cap)
{
    if (!c) return HSDC_ERR_INTERNAL;
    const int cnt = (int)c->tr.size() / 5;
    for (int i = 0; i < cnt && i < cap; i++) {
        if (s) s[i] = c->tr[5 * i];
        if (n) n[i] = c->tr[5 * i + 1];
        if (k) k[i] = c->tr[5 * i + 2];
        if (nrot) nrot[i] = c->tr[5 * i + 3];
        if (hss_rank) hss_rank[i] = c->tr[5 * i + 4];
    }
    return cnt;
}

}  // extern "C"

static int bd_alloc(hsdc_bd_ctx *c, int n)
{
    if (n <= c->wcap) return 0;
    const size_t nd = 13, ni = 36;
    BCK(bensure(c->wd, sizeof(double) * nd * (size_t)n));
    BCK(bensure(c->wi, sizeof(int) * ni * (size_t)n + 64));
    double *D = (double *)c->wd.p;
    int *I = (int *)c->wi.p;
    size_t N = n;
    BdWork &W = c->W;
    W.D = D; W.E = D + N; W.td = D + 2 * N; W.dsg = D + 3 * N; W.zt = D + 4 * N; W.z = D + 5 * N;
    W.zs = D + 6 * N; W.zl = D + 7 * N; W.sig = D + 8 * N; W.tau = D + 9 * N; W.rcs = D + 11 * N;
    W.idxq = I; W.org = I + N; W.ctyp = I + 2 * N; W.idxc = I + 3 * N; W.idx = I + 4 * N;
    W.idxp = I + 5 * N; W.surv = I + 6 * N; W.srank = I + 7 * N; W.clf = I + 8 * N;
    W.rotf = I + 9 * N; W.ti = I + 10 * N;
    W.vmU[0] = I + 11 * N; W.vmU[1] = I + 12 * N; W.vmV[0] = I + 13 * N; W.vmV[1] = I + 14 * N;
    W.sulo = I + 15 * N; W.suhi = I + 16 * N; W.svlo = I + 17 * N; W.svhi = I + 18 * N;
    W.gU = I + 19 * N; W.gV = I + 20 * N; W.gUlo = I + 21 * N; W.gUhi = I + 22 * N; W.gVlo = I + 23 * N;
    W.gVhi = I + 24 * N; W.fU = I + 25 * N; W.fV = I + 26 * N; W.rp4 = I + 27 * N;  // 4 N
    W.iters = (unsigned long long *)(I + 36 * N);
    c->wcap = n;
    return 0;
}

// the whole SVD of the upper bidiagonal already on the device (W.D, W.E, scaled when
// n > smlsiz); outputs U, VT (device, may be null for values only)
// The HSS branch of the merges `hs` of one level (mdlasd1.f90:220-278, DHSSVCS
// CauchyHssvd_VPS.f90:3587-3804), after their roots and updated z: generators, U2 and
// V2 gathered in DSIGMA order, the two HSS forms per merge built and applied as
// X := X * H into the merge's staging (the dense branch's layout), then the first
// row's rank-one term of the left vectors.  rank[mi] receives each merge's largest
// generator rank.
static int bd_hss_level(hsdc_bd_ctx *c, const BdMerge *mg, const BdOut *mo, const std::vector<BdMerge> &L,
                        const BdOut *ho, const std::vector<BdHss> &hs, BdTab tab, const double *U, const double *V,
                        long ld, double *S, int *rank)
{
    hipStream_t st = c->stream;
    const int nh = (int)hs.size();
    long gtot = 0, xtot = 0, gx = 1;
    int kmax = 1;
    for (const BdHss &h : hs) {
        const BdMerge &M = L[h.mi];
        const int k = ho[h.mi].k, n_ = M.nl + M.nr + 1, m_ = n_ + M.sqre;
        gtot += 5L * k;
        xtot += (long)(n_ + m_) * k;
        kmax = std::max(kmax, k);
        const int T = sec_team(k);
        gx = std::max(gx, (long)((k + 256 / T - 1) / (256 / T)));
    }
    BCK(bensure(c->hgen, sizeof(double) * (size_t)gtot));
    BCK(bensure(c->hx, sizeof(double) * (size_t)xtot));
    BCK(bensure(c->htb, sizeof(BdHss) * (size_t)nh));
    if (!c->hctx) BCK(hsdc_ctx_create(&c->hctx));
    BdHss *dh = (BdHss *)c->htb.p;
    double *G = (double *)c->hgen.p, *X = (double *)c->hx.p;
    BHK(hipMemcpyAsync(dh, hs.data(), sizeof(BdHss) * nh, hipMemcpyHostToDevice, st));
    BHK(hipStreamSynchronize(st));  // hs is pageable host memory owned by the caller
    hipLaunchKernelGGL(k_bd_hssgen, dim3((unsigned)gx, nh), dim3(256), 0, st, mg, (const BdHss *)dh, c->W, mo, G);
    hipLaunchKernelGGL(k_bd_xgather, dim3((unsigned)std::min(2L * kmax, 8192L), nh), dim3(256), 0, st, mg,
                       (const BdHss *)dh, c->W, mo, U, V, ld, X);
    BHK(hipGetLastError());
    // staging offsets of the level (k_bd_lvltab's prefix, from the K read back)
    std::vector<long> soff(L.size() + 1, 0);
    for (size_t i = 0; i < L.size(); i++) {
        const int n_ = L[i].nl + L[i].nr + 1, m_ = n_ + L[i].sqre, k = ho[i].k;
        soff[i + 1] = soff[i] + (long)n_ * k + (long)m_ * (k + L[i].sqre);
    }
    std::vector<HssSpec> sp;
    for (const BdHss &h : hs) {
        const BdMerge &M = L[h.mi];
        const int k = ho[h.mi].k, n_ = M.nl + M.nr + 1, m_ = n_ + M.sqre, s = M.s;
        const double *g = G + h.goff;
        HssSpec a;
        a.K = k;
        a.FF = c->W.dsg + s; a.D0 = c->W.sig + s; a.DIFL = g + 3L * k; a.DIFR = g + 4L * k;
        a.sq = 1;
        HssSpec b = a;
        a.U0 = g; a.V0 = g + k;                              // left: d z, APHAU
        a.X = X + h.xoff; a.ldx = n_; a.nrow = n_;
        a.Y = S + soff[h.mi]; a.ldy = n_;
        b.U0 = c->W.zl + s; b.V0 = g + 2L * k;               // right: z, APHAV
        b.X = X + h.xoff + (long)n_ * k; b.ldx = m_; b.nrow = m_;
        b.Y = S + soff[h.mi] + (long)n_ * k; b.ldy = m_;
        sp.push_back(a);
        sp.push_back(b);
    }
    std::vector<int> rk;
    int ndense = 0;
    BCK(hsdc_hss_batch(c->hctx, st, sp, c->prm.hss_leaf, c->prm.hss_tol, &rk, &ndense));
    hipLaunchKernelGGL(k_bd_hssfix, dim3((unsigned)bcdiv(kmax, 256), nh), dim3(256), 0, st, mg, (const BdHss *)dh,
                       tab, mo, (const double *)G, S);
    BHK(hipGetLastError());
    for (int i = 0; i < nh; i++) rank[hs[i].mi] = std::max(rk[2 * i], rk[2 * i + 1]);
    return 0;
}

static int bd_solve(hsdc_bd_ctx *c, int n, const std::vector<double> &he, double orgnrm, double *dU, long ldu,
                    double *dVT, long ldvt, double *dD, const double *rc, const double *rs)
{
    hipStream_t st = c->stream;
    const int smlsiz = c->prm.smlsiz;
    const long ld = n;
    std::memset(&c->st, 0, sizeof c->st);
    c->tr.clear();
    if (!c->ev[0])
        for (int i = 0; i < 4; i++) BHK(hipEventCreateWithFlags(&c->ev[i], i < 2 ? hipEventDefault : hipEventDisableTiming));
    if (!c->zstream) BHK(hipStreamCreateWithFlags(&c->zstream, hipStreamNonBlocking));
    BHK(hipEventRecord(c->ev[0], st));
    // U and VT cleared beside the levels
    if (dU || dVT) {
        BHK(hipEventRecord(c->ev[2], st));
        BHK(hipStreamWaitEvent(c->zstream, c->ev[2], 0));
        if (dU) hipLaunchKernelGGL(k_bd_zero, dim3(8192), dim3(256), 0, c->zstream, dU, ldu, n);
        if (dVT) hipLaunchKernelGGL(k_bd_zero, dim3(8192), dim3(256), 0, c->zstream, dVT, ldvt, n);
        BHK(hipGetLastError());
        BHK(hipEventRecord(c->ev[3], c->zstream));
    }
    BCK(bensure(c->ub[0], sizeof(double) * (size_t)n * n));
    BCK(bensure(c->vb[0], sizeof(double) * (size_t)n * n));
    double *U = (double *)c->ub[0].p, *V = (double *)c->vb[0].p;
    BdWork &W = c->W;
    // subproblems (mdbdsdc.f90:333-372) and their trees (mdlasd0.f90:166-221)
    std::vector<BdLeaf> leaves;
    std::vector<std::vector<BdMerge>> lv;
    int nprob = 0;
    auto add_problem = [&](int start, int size) {
        nprob++;
        if (size <= smlsiz) {
            leaves.push_back({start, size, 0, 0});
            return;
        }
        int nlvl, nd;
        std::vector<int> inode, ndiml, ndimr;
        lasdt(size, nlvl, nd, inode, ndiml, ndimr, smlsiz);
        const int wr0 = nlvl & 1;  // leaves use the map the last merge level ends in as map 0
        const int ndb1 = (nd + 1) / 2;
        for (int i = ndb1; i <= nd; i++) {
            int ic = inode[i - 1], nl = ndiml[i - 1], nr = ndimr[i - 1];
            int nlf = ic - nl, nrf = ic + 1;
            leaves.push_back({start + nlf - 1, nl, 1, wr0});
            leaves.push_back({start + nrf - 1, nr, (i == nd) ? 0 : 1, wr0});
        }
        if ((int)lv.size() < nlvl) lv.resize(nlvl);
        for (int lvl = nlvl; lvl >= 1; lvl--) {
            int lf = (lvl == 1) ? 1 : 1 << (lvl - 1), ll = (lvl == 1) ? 1 : 2 * lf - 1;
            int t = nlvl - lvl;  // merge step of this problem
            for (int i = lf; i <= ll; i++) {
                int ic = inode[i - 1], nl = ndiml[i - 1], nr = ndimr[i - 1];
                int nlf = ic - nl;
                int sq = (i == ll) ? 0 : 1;
                int rd = (wr0 + t) & 1, wr = (wr0 + t + 1) & 1;
                lv[t].push_back({start + nlf - 1, nl, nr, sq, rd, wr});
            }
        }
    };
    if (n <= smlsiz) {
        add_problem(0, n);
    } else {
        const double eps = 0.9 * BD_EPS;
        int start = 1;
        const int nm1 = n - 1;
        for (int i = 1; i <= nm1; i++) {
            if (std::fabs(he[i - 1]) < eps || i == nm1) {
                int nsize;
                if (i < nm1) nsize = i - start + 1;
                else if (std::fabs(he[i - 1]) >= eps) nsize = n - start + 1;
                else {
                    nsize = i - start + 1;
                    add_problem(n - 1, 1);
                }
                add_problem(start - 1, nsize);
                start = i + 1;
            }
        }
    }
    // device tables: merges of every step, leaves; U's centre columns map to themselves
    std::vector<int> moff(lv.size() + 1, 0);
    for (size_t t = 0; t < lv.size(); t++) moff[t + 1] = moff[t] + (int)lv[t].size();
    const int nmg = moff.back();
    BCK(bensure(c->mgb, sizeof(BdMerge) * (nmg + 1) + sizeof(BdLeaf) * (leaves.size() + 1) + 64));
    BdMerge *dmg = (BdMerge *)c->mgb.p;
    BdLeaf *dlv = (BdLeaf *)(dmg + nmg + 1);
    {
        // the tree tables stay on the device between solves of the same shape
        std::vector<int> key = {n, smlsiz, (int)leaves.size(), nmg};
        for (auto &L : leaves) key.insert(key.end(), {L.s, L.m, L.sqre, L.wr});
        for (auto &v : lv)
            for (auto &M : v) key.insert(key.end(), {M.s, M.nl, M.nr, M.sqre, M.rd, M.wr});
        if (key != c->plan_key || c->plan_buf != c->mgb.p) {
            std::vector<BdMerge> all;
            for (auto &v : lv) all.insert(all.end(), v.begin(), v.end());
            if (nmg) BHK(hipMemcpyAsync(dmg, all.data(), sizeof(BdMerge) * nmg, hipMemcpyHostToDevice, st));
            BHK(hipMemcpyAsync(dlv, leaves.data(), sizeof(BdLeaf) * leaves.size(), hipMemcpyHostToDevice, st));
            BHK(hipStreamSynchronize(st));  // the host vectors go out of scope
            c->plan_key.swap(key);
            c->plan_buf = c->mgb.p;
        }
        hipLaunchKernelGGL(k_bd_iota, dim3(bcdiv(n, 256)), dim3(256), 0, st, W.vmU[0], W.vmU[1], n);
    }
    BCK(bensure(c->mob, sizeof(BdOut) * (nmg + 1)));
    BdOut *dmo = (BdOut *)c->mob.p;
    if (c->hmo_cap < (size_t)(nmg + 1)) {
        if (c->hmo) (void)hipHostFree(c->hmo);
        c->hmo = nullptr;
        if (hipHostMalloc((void **)&c->hmo, sizeof(BdOut) * (nmg + 1)) != hipSuccess) {
            (void)hipGetLastError();
            c->hmo_cap = 0;
            return HSDC_ERR_NOMEM;
        }
        c->hmo_cap = nmg + 1;
    }
    int *dfail = (int *)(W.iters + 1);
    BHK(hipMemsetAsync(W.iters, 0, 16, st));
    c->hrank.assign(nmg + 1, 0);
    // leaves
    int pmax = 1;
    for (auto &L : leaves) pmax = std::max(pmax, L.m + L.sqre);
    size_t lds = sizeof(double) * (2 * (size_t)pmax * 64 + 128) + sizeof(int) * 64;
    // a round-robin round has P/2 pairs: one per half wavefront for leaves of <= 32
    // columns, else one per wavefront; at most 16 wavefronts
    const int npair = (pmax + (pmax & 1)) / 2;
    if (pmax <= 32)
        hipLaunchKernelGGL(k_bd_leaf<true>, dim3((unsigned)leaves.size()),
                           dim3(64 * std::max(std::min((npair + 1) / 2, 16), 1)), lds, st, dlv, (int)leaves.size(), W,
                           U, V, ld, pmax, dfail);
    else
        hipLaunchKernelGGL(k_bd_leaf<false>, dim3((unsigned)leaves.size()), dim3(64 * std::max(std::min(npair, 16), 1)),
                           lds, st, dlv, (int)leaves.size(), W, U, V, ld, pmax, dfail);
    BHK(hipGetLastError());
    // Every level is queued without a host round trip: k_bd_lvltab sizes the level's
    // work from its deflation results on the device, the kernels after it walk those
    // tables in grid-stride loops whose grids (and the vector and staging buffers) are
    // the host's bounds for K = N.
    const int nlv = (int)lv.size();
    std::vector<size_t> toff(nlv + 1, 0);
    long qcap = 1, scap = 1;
    std::vector<int> g_rot(nlv), g_sec(nlv), g_low(nlv), g_tile(nlv), g_sc(nlv);
    for (int t = 0; t < nlv; t++) {
        const int nm = (int)lv[t].size();
        toff[t + 1] = toff[t] + al8(sizeof(int) * 5 * (size_t)(nm + 1)) + sizeof(long) * 3 * (size_t)nm + 64;
        long q = 0, st = 0, br = 0, bs = 0, bl = 0, bt = 0, bc = 0;
        for (const BdMerge &M : lv[t]) {
            const long n_ = M.nl + M.nr + 1, m_ = n_ + M.sqre;
            const int T = sec_team((int)n_);
            q += 2 * n_ * n_;
            st += n_ * n_ + m_ * (n_ + M.sqre);
            br += (m_ + 255) / 256;
            bs += (n_ + 256 / T - 1) / (256 / T);
            bl += (n_ + 255) / 256;
            bt += ((n_ + 63) / 64) * ((n_ + 63) / 64 + (m_ + 63) / 64);
            bc += 2 * n_ + M.sqre;
        }
        qcap = std::max(qcap, q);
        scap = std::max(scap, st);
        g_rot[t] = (int)std::min(br, 4096L);
        g_sec[t] = (int)std::min(bs, 8192L);
        g_low[t] = (int)std::min(bl, 4096L);
        g_tile[t] = (int)std::min(bt, 4096L);
        g_sc[t] = (int)std::min(bc, 16384L);
    }
    BCK(bensure(c->tab, toff[nlv] + 64));
    BCK(bensure(c->qb, sizeof(double) * (size_t)qcap));
    BCK(bensure(c->stg, sizeof(double) * (size_t)scap));
    double *Q = (double *)c->qb.p, *S = (double *)c->stg.p;
    for (int t = 0; t < nlv; t++) {
        const int nm = (int)lv[t].size();
        if (!nm) continue;
        BdMerge *mg = dmg + moff[t];
        BdOut *mo = dmo + moff[t];
        char *tb = (char *)c->tab.p + toff[t];
        BdTab tab;
        int *ti = (int *)tb;
        tab.rot = ti; tab.sec = ti + (nm + 1); tab.low = ti + 2 * (nm + 1); tab.tile = ti + 3 * (nm + 1);
        tab.sc = ti + 4 * (nm + 1);
        tab.qoff = (long *)(tb + al8(sizeof(int) * 5 * (size_t)(nm + 1)));
        tab.soff = tab.qoff + nm;
        // the merged runs in LDS (gfx950: 160 KB per workgroup, a few KB kept for the
        // static arrays)
        int nmax = 0;
        for (const BdMerge &M : lv[t]) nmax = std::max(nmax, M.nl + M.nr + 1);
        const int lds_n = nmax <= BD_LDS_RUNS ? nmax : 0;
        hipLaunchKernelGGL(k_bd_prep, dim3(nm), dim3(PREP_NT), sizeof(double) * (size_t)lds_n, st, mg, W, mo,
                           (const double *)V, ld, lds_n);
        hipLaunchKernelGGL(k_bd_deflate, dim3(nm), dim3(DEF_NT), 0, st, mg, W, mo);
        // a level with a merge of order >= hss_min_k reads its deflation results back
        // (K and the HSS flags k_bd_lvltab sets): the HSS branch is planned on the host
        // (tree shape from K) while the level's secular and dense kernels run
        bool maybe_hss = false;
        for (const BdMerge &M : lv[t]) maybe_hss |= M.nl + M.nr + 1 >= c->prm.hss_min_k;
        hipLaunchKernelGGL(k_bd_lvltab, dim3(1), dim3(1024), 0, st, mg, nm, mo, tab, c->prm.hss_min_k,
                           c->prm.hss_leaf);
        BdOut *ho = c->hmo + moff[t];
        if (maybe_hss) {
            if (!c->evh) BHK(hipEventCreateWithFlags(&c->evh, hipEventDisableTiming));
            BHK(hipMemcpyAsync(ho, mo, sizeof(BdOut) * nm, hipMemcpyDeviceToHost, st));
            BHK(hipEventRecord(c->evh, st));
        }
        hipLaunchKernelGGL(k_bd_rotate, dim3(g_rot[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo, U, V, ld);
        hipLaunchKernelGGL(k_bd_znorm, dim3(nm), dim3(64), 0, st, mg, W, mo);
        hipLaunchKernelGGL(k_bd_secular, dim3(g_sec[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo,
                           c->prm.maxit, dfail);
        hipLaunchKernelGGL(k_bd_loewner, dim3(g_low[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo);
        hipLaunchKernelGGL(k_bd_vec, dim3(g_sec[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo, Q);
        hipLaunchKernelGGL(k_bd_gemm, dim3(g_tile[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo,
                           (const double *)Q, (const double *)U, (const double *)V, ld, S);
        if (maybe_hss) {
            BHK(hipEventSynchronize(c->evh));
            std::vector<BdHss> hsel;
            long goff = 0, xoff = 0;
            for (int i = 0; i < nm; i++) {
                if (!ho[i].hss) continue;
                const BdMerge &M = lv[t][i];
                const int k = ho[i].k, n_ = M.nl + M.nr + 1, m_ = n_ + M.sqre;
                hsel.push_back({i, goff, xoff});
                goff += 5L * k;
                xoff += (long)(n_ + m_) * k;
            }
            if (!hsel.empty())
                BCK(bd_hss_level(c, mg, mo, lv[t], ho, hsel, tab, U, V, ld, S, c->hrank.data() + moff[t]));
        }
        hipLaunchKernelGGL(k_bd_newcol, dim3(nm), dim3(256), 0, st, mg, tab, W, (const BdOut *)mo, (const double *)U,
                           (const double *)V, ld, S);
        hipLaunchKernelGGL(k_bd_scatter, dim3(g_sc[t]), dim3(256), 0, st, mg, nm, tab, W, (const BdOut *)mo,
                           (const double *)S, U, V, ld);
        hipLaunchKernelGGL(k_bd_post, dim3(nm), dim3(POST_NT), sizeof(double) * (size_t)lds_n, st, mg, W,
                           (const BdOut *)mo, lds_n);
        BHK(hipGetLastError());
        c->st.levels++;
        c->st.merges += nm;
    }
    // outputs in MDBDSDC's final order (mdbdsdc.f90:383-414)
    BCK(bensure(c->misc, sizeof(int) * (size_t)n + 64));
    int *dperm = (int *)c->misc.p, *dtie = dfail + 1;
    BCK(bensure(c->din2, sizeof(double) * (size_t)n));
    double *dDo = dD ? dD : (double *)c->din2.p;
    auto write_vectors = [&](bool clear) -> int {
        if (!(dU || dVT)) return 0;
        if (clear) {
            if (dU) hipLaunchKernelGGL(k_bd_zero, dim3(8192), dim3(256), 0, st, dU, ldu, n);
            if (dVT) hipLaunchKernelGGL(k_bd_zero, dim3(8192), dim3(256), 0, st, dVT, ldvt, n);
        } else {
            BHK(hipStreamWaitEvent(st, c->ev[3], 0));
        }
        if (dU)
            hipLaunchKernelGGL(k_bd_fin, dim3(n), dim3(256), 0, st, W, (const double *)U, (const double *)V, ld,
                               (const int *)dperm, n, dU, ldu);
        if (dVT)
            hipLaunchKernelGGL(k_bd_fin_vt, dim3(bcdiv(n, 64), bcdiv(n, 64)), dim3(256), 0, st, W, (const double *)V,
                               ld, (const int *)dperm, n, dVT, ldvt);
        BHK(hipGetLastError());
        if (rc && dU) {
            BCK(bensure(c->rot, sizeof(double) * 2 * (size_t)n));
            double *dc = (double *)c->rot.p, *ds = dc + n;
            BHK(hipMemcpyAsync(dc, rc, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st));
            BHK(hipMemcpyAsync(ds, rs, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_bd_lasr, dim3(bcdiv(n, 256)), dim3(256), 0, st, dU, ldu, n, (const double *)dc,
                               (const double *)ds);
            BHK(hipGetLastError());
        }
        return 0;
    };
    unsigned long long cnt[2];
    bool host_order = nprob != 1;
    if (!host_order) {
        hipLaunchKernelGGL(k_bd_perm, dim3(1), dim3(1024), 0, st, W, n, orgnrm, n > smlsiz ? 1 : 0, dperm, dDo, dtie);
        BHK(hipGetLastError());
        BCK(write_vectors(false));
    }
    BHK(hipMemcpyAsync(cnt, W.iters, 16, hipMemcpyDeviceToHost, st));
    BHK(hipStreamSynchronize(st));
    if ((int)(cnt[1] & 0xffffffffu)) {
        if (dU || dVT) BHK(hipStreamWaitEvent(st, c->ev[3], 0));
        return 1;
    }
    if ((int)(cnt[1] >> 32)) host_order = true;  // equal singular values
    // per-merge results of every level (trace and statistics)
    long long roots = 0, deflated = 0, rots = 0;
    int ktop = 0;
    if (nmg) {
        BHK(hipMemcpy(c->hmo, dmo, sizeof(BdOut) * nmg, hipMemcpyDeviceToHost));
        for (int t = 0; t < nlv; t++)
            for (size_t i = 0; i < lv[t].size(); i++) {
                const BdMerge &M = lv[t][i];
                const BdOut &o = c->hmo[moff[t] + i];
                const int n_ = M.nl + M.nr + 1;
                const int rk = c->hrank[moff[t] + i];
                c->tr.insert(c->tr.end(), {M.s, n_, o.k, o.nrot, rk});
                if (o.hss) c->st.hss_merges++;
                c->st.hss_rank_max = std::max(c->st.hss_rank_max, rk);
                roots += o.k;
                deflated += n_ - o.k;
                rots += o.nrot;
                ktop = o.k;
            }
    }
    if (host_order) {
        // selection sort on the host: position i takes the first largest of D(i..n); a
        // segment tree over (value, position) finds it, the swap puts D(i) where the
        // maximum was
        std::vector<double> D(n);
        BHK(hipMemcpyAsync(D.data(), W.D, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        BHK(hipStreamSynchronize(st));
        if (n > smlsiz) host_lascl(1.0, orgnrm, n, D.data());
        std::vector<int> perm(n);
        for (int i = 0; i < n; i++) perm[i] = i;
        int sz = 1;
        while (sz < n) sz <<= 1;
        std::vector<int> tr(2 * sz, -1);
        auto better = [&](int a, int b) {
            if (a < 0) return b;
            if (b < 0) return a;
            if (D[b] > D[a]) return b;
            if (D[a] > D[b]) return a;
            return a < b ? a : b;
        };
        for (int i = 0; i < n; i++) tr[sz + i] = i;
        for (int i = sz - 1; i >= 1; i--) tr[i] = better(tr[2 * i], tr[2 * i + 1]);
        auto update = [&](int q) {
            q += sz;
            for (q >>= 1; q >= 1; q >>= 1) tr[q] = better(tr[2 * q], tr[2 * q + 1]);
        };
        auto query = [&](int lo) {  // first maximum over [lo, n)
            int best = -1, l = lo + sz, r = n + sz;
            while (l < r) {
                if (l & 1) best = better(best, tr[l++]);
                if (r & 1) best = better(best, tr[--r]);
                l >>= 1;
                r >>= 1;
            }
            return best;
        };
        for (int i = 0; i < n - 1; i++) {
            int kk = query(i);
            if (kk != i && D[kk] > D[i]) {
                std::swap(D[kk], D[i]);
                std::swap(perm[kk], perm[i]);
                update(kk);
                update(i);
            }
        }
        BHK(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
        BHK(hipMemcpyAsync(dDo, D.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
        // vectors written already in IDXQ order are rewritten over cleared outputs
        BCK(write_vectors(nprob == 1));
        BHK(hipStreamSynchronize(st));
    }
    BHK(hipEventRecord(c->ev[1], st));
    BHK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->st.ms_total = ms;
    c->st.roots = roots;
    c->st.deflated = deflated;
    c->st.rotations = rots;
    c->st.k_top = ktop;
    c->st.secular_iters = (long long)cnt[0];
    c->st.leaves = (int)leaves.size();
    return 0;
}

static __global__ void k_bd_eye(double *A, long lda, int n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)n * n) return;
    const int r = (int)(i % n), c = (int)(i / n);
    A[(long)c * lda + r] = r == c ? 1.0 : 0.0;
}

extern "C" {

// MDBDSDC on device arrays (mdbdsdc.f90:1-422, COMPQ = 'I'); d_u / d_vt may be null
// for singular values only
int hsdc_bdsdc_dev(hsdc_bd_ctx *c, char uplo, int n, double *d_d, double *d_e, double *d_u, int ldu,
                   double *d_vt, int ldvt)
{
    if (!c) return HSDC_ERR_INTERNAL;
    int iuplo = (uplo == 'U' || uplo == 'u') ? 1 : (uplo == 'L' || uplo == 'l') ? 2 : 0;
    if (!iuplo) return -1;
    if (n < 0) return -3;
    if (d_u && ldu < (n > 1 ? n : 1)) return -7;
    if (d_vt && ldvt < (n > 1 ? n : 1)) return -9;
    if (n == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { (void)hipGetLastError(); return HSDC_ERR_NODEVICE; }
    hipStream_t st = c->stream;
    if (c->hde_cap < 2 * (size_t)n) {
        if (c->hde) (void)hipHostFree(c->hde);
        c->hde = nullptr;
        c->hde_cap = 0;
        if (hipHostMalloc((void **)&c->hde, sizeof(double) * 2 * (size_t)n) != hipSuccess) {
            (void)hipGetLastError();
            return HSDC_ERR_NOMEM;
        }
        c->hde_cap = 2 * (size_t)n;
    }
    BHK(hipMemcpyAsync(c->hde, d_d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (n > 1) BHK(hipMemcpyAsync(c->hde + n, d_e, sizeof(double) * (n - 1), hipMemcpyDeviceToHost, st));
    BHK(hipStreamSynchronize(st));
    std::vector<double> hd(c->hde, c->hde + n), he(n > 1 ? n - 1 : 1, 0.0);
    if (n > 1) std::copy(c->hde + n, c->hde + 2 * n - 1, he.begin());
    std::memset(&c->st, 0, sizeof c->st);
    c->tr.clear();
    if (n == 1) {
        double u = hd[0] >= 0.0 ? 1.0 : -1.0, v = 1.0, dv = std::fabs(hd[0]);
        if (d_u) BHK(hipMemcpyAsync(d_u, &u, sizeof(double), hipMemcpyHostToDevice, st));
        if (d_vt) BHK(hipMemcpyAsync(d_vt, &v, sizeof(double), hipMemcpyHostToDevice, st));
        BHK(hipMemcpyAsync(d_d, &dv, sizeof(double), hipMemcpyHostToDevice, st));
        BHK(hipStreamSynchronize(st));
        return 0;
    }
    // lower bidiagonal: rotate to upper from the left (mdbdsdc.f90:216-233)
    std::vector<double> rc, rs;
    if (iuplo == 2) {
        rc.resize(n - 1);
        rs.resize(n - 1);
        for (int i = 0; i < n - 1; i++) {
            double cs, sn, r;
            lartg(hd[i], he[i], cs, sn, r);
            hd[i] = r;
            he[i] = sn * hd[i + 1];
            hd[i + 1] = cs * hd[i + 1];
            rc[i] = cs;
            rs[i] = -sn;
        }
    }
    double orgnrm = 1.0;
    if (n > c->prm.smlsiz) {
        orgnrm = 0.0;
        for (int i = 0; i < n; i++) orgnrm = std::fmax(orgnrm, std::fabs(hd[i]));
        for (int i = 0; i < n - 1; i++) orgnrm = std::fmax(orgnrm, std::fabs(he[i]));
        if (orgnrm == 0.0) {
            // mdbdsdc.f90:292-294: U = VT = I, D unchanged
            if (d_u) hipLaunchKernelGGL(k_bd_eye, dim3(bcdiv((long)n * n, 256)), dim3(256), 0, st, d_u, (long)ldu, n);
            if (d_vt) hipLaunchKernelGGL(k_bd_eye, dim3(bcdiv((long)n * n, 256)), dim3(256), 0, st, d_vt, (long)ldvt, n);
            BHK(hipGetLastError());
            BHK(hipStreamSynchronize(st));
            return 0;
        }
        host_lascl(orgnrm, 1.0, n, hd.data());
        host_lascl(orgnrm, 1.0, n - 1, he.data());
        const double eps = 0.9 * BD_EPS;
        for (int i = 0; i < n; i++)
            if (std::fabs(hd[i]) < eps) hd[i] = hd[i] >= 0.0 ? eps : -eps;
    }
    BCK(bd_alloc(c, n));
    std::copy(hd.begin(), hd.end(), c->hde);
    std::copy(he.begin(), he.begin() + (n - 1), c->hde + n);
    BHK(hipMemcpyAsync(c->W.D, c->hde, sizeof(double) * n, hipMemcpyHostToDevice, st));
    BHK(hipMemcpyAsync(c->W.E, c->hde + n, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st));
    int info = bd_solve(c, n, he, orgnrm, d_u, ldu, d_vt, ldvt, d_d, iuplo == 2 ? rc.data() : nullptr,
                        iuplo == 2 ? rs.data() : nullptr);
    if (info) {
        // the fill of the outputs may still run on the second stream
        if (c->zstream) (void)hipStreamSynchronize(c->zstream);
        return info;
    }
    BHK(hipStreamSynchronize(st));
    return 0;
}

// MDBDSDC on host arrays (mdbdsdc.f90:1-2): COMPQ 'N' or 'I'.  q / iq (COMPQ = 'P') and
// work / iwork are not used: the context owns the device workspace.
int hsdc_bdsdc(hsdc_bd_ctx *c, char uplo, char compq, int n, double *d, double *e, double *u, int ldu,
               double *vt, int ldvt, double *q, int *iq, double *work, int *iwork)
{
    (void)q; (void)iq; (void)work; (void)iwork;
    if (!c) return HSDC_ERR_INTERNAL;
    int iuplo = (uplo == 'U' || uplo == 'u') ? 1 : (uplo == 'L' || uplo == 'l') ? 2 : 0;
    int icompq = (compq == 'N' || compq == 'n') ? 0 : (compq == 'P' || compq == 'p') ? 1
                 : (compq == 'I' || compq == 'i') ? 2 : -1;
    if (!iuplo) return -1;
    if (icompq < 0) return -2;
    if (icompq == 1) return -2;  // the compact form (DLASDA) is not provided
    if (n < 0) return -3;
    if (ldu < 1 || (icompq == 2 && ldu < n)) return -7;
    if (ldvt < 1 || (icompq == 2 && ldvt < n)) return -9;
    if (n == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { (void)hipGetLastError(); return HSDC_ERR_NODEVICE; }
    hipStream_t st = c->stream;
    BCK(bensure(c->din, sizeof(double) * 2 * (size_t)n));
    double *dd = (double *)c->din.p, *de = dd + n;
    double *du = nullptr, *dvt = nullptr;
    if (icompq == 2) {
        BCK(bensure(c->uo, sizeof(double) * (size_t)n * n));
        BCK(bensure(c->vto, sizeof(double) * (size_t)n * n));
        du = (double *)c->uo.p;
        dvt = (double *)c->vto.p;
    }
    BHK(hipMemcpyAsync(dd, d, sizeof(double) * n, hipMemcpyHostToDevice, st));
    if (n > 1) BHK(hipMemcpyAsync(de, e, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st));
    int info = hsdc_bdsdc_dev(c, uplo, n, dd, de, du, n, dvt, n);
    if (info) return info;
    BHK(hipMemcpyAsync(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (icompq == 2) {
        BHK(hipMemcpy2DAsync(u, sizeof(double) * ldu, du, sizeof(double) * n, sizeof(double) * n, n,
                             hipMemcpyDeviceToHost, st));
        BHK(hipMemcpy2DAsync(vt, sizeof(double) * ldvt, dvt, sizeof(double) * n, sizeof(double) * n, n,
                             hipMemcpyDeviceToHost, st));
    }
    BHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
