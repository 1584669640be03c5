// HSS form of the eigenvector matrix of a rank-one merge and its product with the
// old eigenvectors, for gfx950.
//
//   construction  Cauchy2hssEig_omp / ComprcauchyEig / rrluCauchyEigrow,col
//                 (CauchyHssEig.f90:7-467, 1945-2440): one workgroup per HSS node
//                 runs the structured rank-revealing LU on the Cauchy generators;
//                 every element of the Schur complement it touches is formed from
//                 the generators with the accurate differences of
//                 computfd / computdd (CauchyHssEig.f90:470-568), so nothing of the
//                 K x K matrix is ever stored.  Statement order follows the
//                 reference, which keeps pivots, ranks and generators identical to
//                 the CPU oracle's restatement.
//   multiply      fastHssmmL_omp (ConstructHssd.f90:1394-1630): X := X * H as
//                 batched FP64 MFMA GEMMs, one launch per tree level, over all
//                 nodes of all HSS merges of the DC level and all row tiles of X.
#include <float.h>
#include "hsdc_internal.h"
#include "hsdc_hss.h"

#define WG 256
#define A1(a, i) (a)[(i) - 1]

// Pointers typed by address space: global memory (1) and LDS (3).  Loads through an
// untyped pointer compile to FLAT instructions, which the compiler must order against
// every LDS access and every store; typed loads and stores let it issue the loads of
// many columns before their results are needed.
typedef const __attribute__((address_space(1))) double gdbl;
typedef __attribute__((address_space(1))) double gdblw;
typedef __attribute__((address_space(1))) int gint;
typedef const __attribute__((address_space(1))) int gcint;
typedef __attribute__((address_space(3))) double ldbl;

// ---------------------------------------------------------------------------
// accurate differences: f_{PL(j)} - lambda_{PU(k)} and lambda_{PU(j)} - lambda_{PU(k)}
// ---------------------------------------------------------------------------
__device__ __forceinline__ double fd(const HssMerge &H, int pd, int pw)
{
    gdbl *FF = (gdbl *)H.FF;
    if (pd <= pw) return (A1(FF, pd) - A1(FF, pw)) - A1((gdbl *)H.DIFL, pw);
    return (A1(FF, pd) - A1(FF, pw + 1)) - A1((gdbl *)H.DIFR, pw);
}
__device__ __forceinline__ double dd(const HssMerge &H, int pj, int pk)
{
    gdbl *FF = (gdbl *)H.FF;
    if (pj < pk) return ((A1(FF, pj + 1) - A1(FF, pk)) + A1((gdbl *)H.DIFR, pj)) - A1((gdbl *)H.DIFL, pk);
    return ((A1(FF, pj) - A1(FF, pk + 1)) + A1((gdbl *)H.DIFL, pj)) - A1((gdbl *)H.DIFR, pk);
}

// The same differences from values gathered once per root: a pass over many columns
// loads, for each column's root p, FF(p), FF(p + 1), DIFL(p), DIFR(p) (and D0(p) for the
// squared form) together, then forms every difference it needs from registers.  The
// operations and their order are those of fd / dd above, so the results are identical.
struct RootG { double f0, f1, dl, dr, s; };
template <bool SQ>
__device__ __forceinline__ RootG root_load(const HssMerge &H, int p)
{
    RootG r;
    r.f0 = A1((gdbl *)H.FF, p);
    r.f1 = A1((gdbl *)H.FF, p < H.K ? p + 1 : p);  // FF(p + 1) is read only for p < K
    r.dl = A1((gdbl *)H.DIFL, p);
    r.dr = A1((gdbl *)H.DIFR, p);
    r.s = SQ ? A1((gdbl *)H.D0, p) : 0.0;
    return r;
}
// fdq(pd, pw) with ffd = FF(pd) and w = root_load(pw)
template <bool SQ>
__device__ __forceinline__ double fd_r(double ffd, int pd, int pw, const RootG &w)
{
    const double r = (pd <= pw) ? (ffd - w.f0) - w.dl : (ffd - w.f1) - w.dr;
    return SQ ? r * (ffd + w.s) : r;
}
// Whether |z / den| can exceed the running maximum mv of a "first largest" search: the
// hardware reciprocal gives |z| / |den| to far better than 2^-10 relative, so an entry
// estimated below mv (1 - 2^-10) is below mv exactly and cannot replace it; every other
// entry (and any zero, infinite or NaN operand) takes the exact division.  The searches
// therefore choose the same entries as with a division for every element.
__device__ __forceinline__ bool may_exceed(double z, double den, double mv)
{
    const double az = fabs(z);
    const double ap = az * __builtin_amdgcn_rcp(fabs(den));
    return !(ap < mv * (1.0 - 0x1p-10)) || (ap == 0.0 && az != 0.0);
}
// ddq(pj, pk) with j = root_load(pj), k = root_load(pk)
template <bool SQ>
__device__ __forceinline__ double dd_r(int pj, const RootG &j, int pk, const RootG &k)
{
    const double r = (pj < pk) ? ((j.f1 - k.f0) + j.dr) - k.dl : ((j.f0 - k.f1) + j.dl) - k.dr;
    return SQ ? r * (j.s + k.s) : r;
}
// The SVD's vector matrices u_i v_j / (f_i^2 - sigma_j^2) are Cauchy-like in the
// squared variables: with SQ every displacement is the accurate difference times the
// sum (D0 holds the roots sigma): pole-root, root-root and pole-pole.
template <bool SQ>
__device__ __forceinline__ double fdq(const HssMerge &H, int pd, int pw)
{
    const double r = fd(H, pd, pw);
    return SQ ? r * (A1((gdbl *)H.FF, pd) + A1((gdbl *)H.D0, pw)) : r;
}
template <bool SQ>
__device__ __forceinline__ double ddq(const HssMerge &H, int pj, int pk)
{
    const double r = dd(H, pj, pk);
    return SQ ? r * (A1((gdbl *)H.D0, pj) + A1((gdbl *)H.D0, pk)) : r;
}
template <bool SQ>
__device__ __forceinline__ double ppq(double a, double b)
{
    return SQ ? (a - b) * (a + b) : a - b;
}

// block-wide (max |value|, first index); idx <= 0 means "no entry"
struct VI { double v; int i; };
__device__ __forceinline__ VI vi_better(VI a, VI b)
{
    if (b.i <= 0) return a;
    if (a.i <= 0) return b;
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}
template <int O>
__device__ __forceinline__ VI vi_step(VI x)
{
    VI y;
    y.v = xlane<O>(x.v);
    y.i = xlane32<O>(x.i);
    return vi_better(x, y);
}
// (value, index) maximum with ties to the smaller index is a total order, so the
// butterfly's pairing does not change the result
__device__ VI block_argmax(VI x, VI *sh)
{
    int tid = threadIdx.x;
    x = vi_step<32>(x);
    x = vi_step<16>(x);
    x = vi_step<8>(x);
    x = vi_step<4>(x);
    x = vi_step<2>(x);
    x = vi_step<1>(x);
    if ((tid & 63) == 0) sh[tid >> 6] = x;
    __syncthreads();
    VI r = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = vi_better(r, sh[w]);
    __syncthreads();
    return r;
}

// The same with one barrier: successive calls alternate between the two halves of sh
// (4 x blockDim / 64 entries; a half holds 2 x blockDim / 64 so that the two-value
// form below shares them), so a call's writes never meet the previous call's reads
// and the trailing barrier that protects a single buffer is not needed.  tog is
// uniform over the block and flips on every call.
// With wave set the block is one wavefront: every lane holds the maximum after the
// butterfly, and nothing goes through sh.
__device__ __forceinline__ VI block_argmax_t(VI x, VI *sh, int &tog, bool wave = false)
{
    const int tid = threadIdx.x, nw = (int)(blockDim.x >> 6);
    VI *b = sh + tog * 2 * nw;
    tog ^= 1;
    x = vi_step<32>(x);
    x = vi_step<16>(x);
    x = vi_step<8>(x);
    x = vi_step<4>(x);
    x = vi_step<2>(x);
    x = vi_step<1>(x);
    if (wave) return x;
    if ((tid & 63) == 0) b[tid >> 6] = x;
    __syncthreads();
    VI r = b[0];
    for (int w = 1; w < nw; w++) r = vi_better(r, b[w]);
    return r;
}
// Two independent (value, index) maxima under one barrier, in the same alternating
// halves of sh.
__device__ __forceinline__ void block_argmax2_t(VI &x, VI &y, VI *sh, int &tog, bool wave = false)
{
    const int tid = threadIdx.x, nw = (int)(blockDim.x >> 6);
    VI *b = sh + tog * 2 * nw;
    tog ^= 1;
    x = vi_step<32>(x); y = vi_step<32>(y);
    x = vi_step<16>(x); y = vi_step<16>(y);
    x = vi_step<8>(x); y = vi_step<8>(y);
    x = vi_step<4>(x); y = vi_step<4>(y);
    x = vi_step<2>(x); y = vi_step<2>(y);
    x = vi_step<1>(x); y = vi_step<1>(y);
    if (wave) return;
    if ((tid & 63) == 0) { b[tid >> 6] = x; b[nw + (tid >> 6)] = y; }
    __syncthreads();
    VI rx = b[0], ry = b[nw];
    for (int w = 1; w < nw; w++) { rx = vi_better(rx, b[w]); ry = vi_better(ry, b[nw + w]); }
    x = rx;
    y = ry;
}
// The block's barrier, or for a one-wavefront block the ordering of its LDS accesses
// (a wavefront executes in lockstep; the fence completes lane 0's writes before any
// lane reads them).
__device__ __forceinline__ void bsync(bool wave)
{
    if (wave) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}
// x * t[0] * t[1] * ... * t[n - 1], multiplied in that order (one thread), with the
// factors read from LDS eight at a time ahead of their multiplications
__device__ __forceinline__ double ordered_prod(double x, const double *t, int n)
{
    int j = 0;
    for (; j + 8 <= n; j += 8) {
        double f[8];
        #pragma unroll
        for (int u = 0; u < 8; u++) f[u] = t[j + u];
        #pragma unroll
        for (int u = 0; u < 8; u++) x = x * f[u];
    }
    for (; j < n; j++) x = x * t[j];
    return x;
}

// ---------------------------------------------------------------------------
// leaf D blocks (CauchylikeEig on the leaf rows/cols, CauchyHssEig.f90:2089-2100)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_hss_dgen(const HssMerge *hm, const HssNode *nodes, int nn)
{
    int b = blockIdx.x;
    if (b >= nn) return;
    HssNode N = nodes[b];
    const HssMerge H = hm[N.mrg];
    int m = N.leaf_m, r0 = N.leaf_r0;  // leaf rows/cols r0+1 .. r0+m (1-based)
    double *out = H.dd + H.pd[N.node - 1];
    for (int e = threadIdx.x; e < m * m; e += WG) {
        int i = e % m, j = e / m;
        int pf = r0 + 1 + i, pw = r0 + 1 + j;
        double den = (pf <= pw) ? (A1(H.FF, pf) - A1(H.FF, pw)) - A1(H.DIFL, pw)
                                : (A1(H.FF, pf) - A1(H.FF, pw + 1)) - A1(H.DIFR, pw);
        if (H.sq) den = den * (A1(H.FF, pf) + A1(H.D0, pw));
        out[e] = A1(H.U0, pf) * A1(H.V0, pw) / den;
    }
    if (threadIdx.x == 0) H.dsz[N.node - 1] = m;
}

// inlined: a call would give the merge record an address, and the compiler would then
// keep it in scratch memory and reload its fields there inside the pivot steps
__device__ __forceinline__ long arena_take(const HssMerge &H, long n)
{
    unsigned long long o = atomicAdd(H.arena_top, (unsigned long long)n);
    if ((long)(o + n) > H.arena_cap) { atomicOr(H.err, 2); return -1; }
    return (long)o;
}
__device__ __forceinline__ long scratch_take(const HssMerge &H, long n)
{
    unsigned long long o = atomicAdd(H.scratch_top, (unsigned long long)n);
    if ((long)(o + n) > H.scratch_cap) { atomicOr(H.err, 4); return -1; }
    return (long)o;
}

// ---------------------------------------------------------------------------
// Row compression of one node: rrluCauchyEigrow + CauchylikeEigQ (ComprcauchyEig 'r')
// rows: the node's M row generators (LDS); columns: the N column generators
// outside the node (private copies: V in LDS when it fits, else global; PU global,
// with positions 1..min(M, N) mirrored in LDS).
// The passes over the N columns take RB columns per thread at a time (7, or 6 in the
// squared form, for 512 threads; 4 for 256, whose launches put several workgroups on a
// CU): the loads of all RB (V, PU, then the roots' gathered values) are issued before
// any result is used, the next batch's V and PU before the current batch's results,
// and the stores of a batch after its loads.  A step the swap check does not follow runs the next step's
// row search inside its own row update and reduces it with the next step's first
// column search under one barrier.
// ---------------------------------------------------------------------------
template <int HSS_CWG, bool SQ>
__global__ __launch_bounds__(HSS_CWG) void k_hss_rowcomp(const HssMerge *hm, const HssNode *nodes, int nn,
                                                    double tol, int lds_cap)
{
    extern __shared__ double hss_dyn[];  // lds_cap doubles: V when the node's columns fit
    __shared__ double F[HSS_MAXM], A[HSS_MAXM], Z[HSS_MAXM], W[HSS_MAXM];
    __shared__ double T1[HSS_MAXM], T2[HSS_MAXM];  // factors of thread 0's ordered products
    __shared__ int PL[HSS_MAXM], PUs[HSS_MAXM];
    __shared__ short PQ[HSS_MAXM], IP[HSS_MAXM];   // positions within the node (<= HSS_MAXM)
    __shared__ VI red[4 * (HSS_CWG / 64)];
    int tog = 0;  // which half of red the next reduction uses
    __shared__ long s_off;
    __shared__ int s_M, s_N, s_Rk, s_flag, s_ii, s_jj;
    const int tid = threadIdx.x;
    constexpr int RB = HSS_CWG >= 512 ? (SQ ? 6 : 7) : 4;  // within 256 VGPRs
    if ((int)blockIdx.x >= nn) return;
    HssNode Nd = nodes[blockIdx.x];
    const HssMerge H = hm[Nd.mrg];
    int *lr = H.lr, *lc = H.lc;
    const int node = Nd.node;
    if (tid == 0) {
        if (Nd.c1 > 0) {
            lr[2 * (node - 1)] = lr[2 * (Nd.c1 - 1)]; lr[2 * (node - 1) + 1] = lr[2 * (Nd.c2 - 1) + 1];
            lc[2 * (node - 1)] = lc[2 * (Nd.c1 - 1)]; lc[2 * (node - 1) + 1] = lc[2 * (Nd.c2 - 1) + 1];
        }
        int M = lr[2 * (node - 1) + 1] - lr[2 * (node - 1)] + 1;
        int nic = lc[2 * (node - 1) + 1] - lc[2 * (node - 1)] + 1;
        s_M = M;
        s_N = *H.lca - nic;
        // a merge whose construction already failed is finished densely; skip it
        s_flag = (*H.err != 0 || M > HSS_MAXM || M < 1 || nic < 1 || s_N < 0) ? 1 : 0;
        s_off = s_flag ? -1 : scratch_take(H, 2L * (s_N > 0 ? s_N : 1));
        if (s_flag) atomicOr(H.err, 1);
    }
    __syncthreads();
    if (s_flag || s_off < 0) return;
    const int M = s_M, N = s_N, l0 = lr[2 * (node - 1)], c0 = lc[2 * (node - 1)], c1e = lc[2 * (node - 1) + 1];
    const int NS = N > 0 ? N : 1;
    const int mn = M < N ? M : N;
    // Private per column: the generator V (updated every step) and the root index PU;
    // the values of the root that its accurate differences read (FF, DIFL, DIFR and D0 at
    // PU and PU + 1) are shared by every node of the merge.  Kept in the dynamic LDS as
    // far as it holds them (lds_cap doubles): V; V and PU; or V, PU and the roots'
    // values, so that a small node's passes, swaps and pivot lookups stay in LDS.
    // Otherwise V and PU are global and the values are gathered by index.
    constexpr int RD = SQ ? 5 : 4;  // root values kept per column
    const bool vl = N <= lds_cap;   // all three uniform over the block
    const bool pl = (long)N + (N + 1) / 2 <= lds_cap;
    const bool sm = (long)N * (1 + RD) + (N + 1) / 2 <= lds_cap;
    // A node whose data all sit in LDS and whose outside part is small runs on one
    // wavefront: its steps need no block barrier (the other waves leave here)
    const bool wave = sm && N <= 384;
    if (wave && tid >= 64) return;
    const int NTH = wave ? 64 : HSS_CWG;
    ldbl *Vl = (ldbl *)hss_dyn;
    ldbl *Rl = Vl + NS;  // sm: f0, f1, dl, dr (and s) of every column, NS apart
    __attribute__((address_space(3))) int *PUl =
        (__attribute__((address_space(3))) int *)(sm ? Rl + RD * NS : Vl + NS);
    gdblw *Vg = (gdblw *)(H.scratch + s_off);
    gint *PUg = (gint *)(H.scratch + s_off + NS);
    auto vld = [&](int j) -> double { return vl ? A1(Vl, j) : A1(Vg, j); };
    auto vst = [&](int j, double x) {
        if (vl) A1(Vl, j) = x;
        else A1(Vg, j) = x;
    };
    auto pu = [&](int j) -> int { return pl ? A1(PUl, j) : A1(PUg, j); };
    auto pu_set = [&](int j, int x) {
        if (pl) A1(PUl, j) = x;
        else A1(PUg, j) = x;
    };
    auto rl_get = [&](int j) -> RootG {
        return RootG{Rl[j - 1], Rl[NS + j - 1], Rl[2 * NS + j - 1], Rl[3 * NS + j - 1], SQ ? Rl[4 * NS + j - 1] : 0.0};
    };
    // the root values of the column at position j <= min(M, N)
    auto croot = [&](int j) -> RootG { return sm ? rl_get(j) : root_load<SQ>(H, A1(PUs, j)); };
    for (int j = tid; j < N; j += NTH) {
        int src = (j < c0 - 1) ? j + 1 : j + 1 + (c1e - c0 + 1);
        vst(j + 1, A1((gdbl *)H.V, src));
        const int p = A1((gcint *)H.PU, src);
        pu_set(j + 1, p);
        if (j < mn) PUs[j] = p;
        if (sm) {
            const RootG r = root_load<SQ>(H, p);
            Rl[j] = r.f0; Rl[NS + j] = r.f1; Rl[2 * NS + j] = r.dl; Rl[3 * NS + j] = r.dr;
            if (SQ) Rl[4 * NS + j] = r.s;
        }
    }
    // column swap (one thread; a barrier follows)
    auto swapc = [&](int a, int c) {
        double t = vld(a); vst(a, vld(c)); vst(c, t);
        int u = pu(a), w = pu(c);
        pu_set(a, w); pu_set(c, u);
        if (a <= mn) A1(PUs, a) = w;
        if (c <= mn) A1(PUs, c) = u;
        if (sm) {
            #pragma unroll
            for (int q = 0; q < RD; q++) {
                t = Rl[q * NS + a - 1]; Rl[q * NS + a - 1] = Rl[q * NS + c - 1]; Rl[q * NS + c - 1] = t;
            }
        }
    };
    auto cload = [&](int j0, int *pw, double *v) {
        #pragma unroll
        for (int u = 0; u < RB; u++) {
            const int j = min(j0 + u * NTH, N);
            pw[u] = pu(j);
            v[u] = vld(j);
        }
    };
    // one batch with the roots' values (CP scan)
    auto bload = [&](int j0, int *pw, double *v, RootG *c) {
        cload(j0, pw, v);
        #pragma unroll
        for (int u = 0; u < RB; u++) c[u] = sm ? rl_get(min(j0 + u * NTH, N)) : root_load<SQ>(H, pw[u]);
    };
    // body(j, pw, v, c) for the columns j in [lo, N]; j may exceed N (loads clamped)
    auto cpass = [&](int lo, auto body) {
        int j0 = lo + tid;
        if (j0 > N) return;
        if (sm) {
            for (int j = j0; j <= N; j += NTH) body(j, A1(PUl, j), A1(Vl, j), rl_get(j));
            return;
        }
        int pw[RB];
        double v[RB];
        cload(j0, pw, v);
        for (;;) {
            RootG c[RB];
            #pragma unroll
            for (int u = 0; u < RB; u++) c[u] = root_load<SQ>(H, pw[u]);
            const int jn = j0 + RB * NTH;
            int pn[RB];
            double vn[RB];
            if (jn <= N) cload(jn, pn, vn);
            #pragma unroll
            for (int u = 0; u < RB; u++) body(j0 + u * NTH, pw[u], v[u], c[u]);
            if (jn > N) break;
            j0 = jn;
            #pragma unroll
            for (int u = 0; u < RB; u++) { pw[u] = pn[u]; v[u] = vn[u]; }
        }
    };
    // max over j in [k, N] of |V(j) / fd(pd, PU(j))| (index j - k + 1), ffd = FF(pd)
    auto colsearch = [&](int k, int pd, double ffd) -> VI {
        VI y = {0.0, 0};
        cpass(k, [&](int j, int pw, double v, const RootG &c) {
            if (j <= N) y = vi_better(y, VI{fabs(v / fd_r<SQ>(ffd, pd, pw, c)), j - k + 1});
        });
        return y;
    };
    // max over i in [k, M] of |Z(i) / fd(PL(i), pw)|, w = the values of root pw
    auto rowsearch = [&](int k, int pw, const RootG &w) -> VI {
        VI x = {0.0, 0};
        for (int i = k + tid; i <= M; i += NTH)
            x = vi_better(x, VI{fabs(A1(Z, i) / fd_r<SQ>(A1(F, i), A1(PL, i), pw, w)), i - k + 1});
        return x;
    };
    for (int i = tid; i < M; i += NTH) {
        PL[i] = A1((gcint *)H.PL, l0 + i);
        A[i] = A1((gdbl *)H.U, l0 + i);
        F[i] = A1((gdbl *)H.FF, PL[i]);
        Z[i] = A[i];
        PQ[i] = (short)(i + 1);
        W[i] = 0.0;
    }
    bsync(wave);
    double amax = 0.0;
    int Rk = mn;
    const double rho = 1.5;
    bool pre = false;     // pre_cl, pre_rw hold this step's first row and column searches
    VI pre_cl = {0.0, 0}, pre_rw = {0.0, 0};
    RootG cnext = {0.0, 0.0, 0.0, 0.0, 0.0};  // root data of PU(k + 1) when pre
    // flop count of the reference's model (nflops of rrluCauchyEigrow and its pivot
    // searches, CauchyHssEig.f90:105, 147, 205, 680-703); the loop is uniform over the
    // block, thread 0 adds it to the merge's total
    unsigned long long fl = 0;
    for (int k = 1; k <= mn; k++) {
        // ---- CauchyPivtEig: rook search on the k-th Schur complement
        fl += 4ull * (M + N - 2 * k + 3);
        RootG cw;  // root data of the column at position k (PU(k)), kept through the step
        {
            VI cl, rw;
            if (pre) {
                cl = pre_cl;
                rw = pre_rw;
                cw = cnext;
            } else {
                const int pwk = A1(PUs, k);
                cw = croot(k);
                cl = rowsearch(k, pwk, cw);
                rw = colsearch(k, A1(PL, k), A1(F, k));
                block_argmax2_t(cl, rw, red, tog, wave);
            }
            double junkL = cl.v * fabs(vld(k)), junkU = rw.v * fabs(A1(Z, k));
            double piv = fabs(A1(Z, k) * vld(k) / fd_r<SQ>(A1(F, k), A1(PL, k), A1(PUs, k), cw));
            if (!(junkL <= piv && junkU <= piv)) {
                int flg = (junkL > junkU) ? 1 : 0;
                int jjL = cl.i, jjU = rw.i;
                // Every thread has read position k for the tests above before any swap.
                // A row swap is made after the column search it feeds, which reads the
                // incoming row where it still is; every later swap follows a reduction,
                // whose barrier comes after all reads of the previous search.
                bsync(wave);
                for (;;) {
                    if (flg == 1) {
                        const int jr = jjL + k - 1;
                        jjU = block_argmax_t(colsearch(k, A1(PL, jr), A1(F, jr)), red, tog, wave).i;
                        fl += 4ull * (N - k + 1);
                        if (tid == 0) {
                            double t;
                            t = A1(F, k); A1(F, k) = A1(F, jr); A1(F, jr) = t;
                            t = A1(Z, k); A1(Z, k) = A1(Z, jr); A1(Z, jr) = t;
                            t = A1(A, k); A1(A, k) = A1(A, jr); A1(A, jr) = t;
                            int u = A1(PL, k); A1(PL, k) = A1(PL, jr); A1(PL, jr) = u;
                            short q = A1(PQ, k); A1(PQ, k) = A1(PQ, jr); A1(PQ, jr) = q;
                        }
                        if (jjU == 1) {
                            bsync(wave);
                            break;
                        }
                    }
                    if (tid == 0) swapc(k, jjU + k - 1);
                    bsync(wave);
                    const int pwk2 = A1(PUs, k);
                    cw = croot(k);
                    jjL = block_argmax_t(rowsearch(k, pwk2, cw), red, tog, wave).i;
                    fl += 4ull * (M - k + 1);
                    if (jjL == 1) break;
                    flg = 1;
                }
            }
        }
        // every swap of the search is followed by a barrier and the search only reads, so
        // the update pass below may start without another
        double ukk = A1(Z, k) * vld(k) / fd_r<SQ>(A1(F, k), A1(PL, k), A1(PUs, k), cw);
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            // ---- CauchyPivtEig_CP: complete pivoting over the Schur complement
            fl += 4ull * (M - k + 1) * (N - k + 1);
            VI best = {0.0, 0};
            VI bestrow = {0.0, 0};
            for (int j0 = k + tid; j0 <= N; j0 += RB * NTH) {
                int pw[RB];
                double v[RB], mv[RB];
                int mi[RB];
                RootG c[RB];
                bload(j0, pw, v, c);
                #pragma unroll
                for (int u = 0; u < RB; u++) { mv[u] = -DBL_MAX; mi[u] = 0; }
                for (int i = k; i <= M; i++) {
                    const double zi = A1(Z, i), fi = A1(F, i);
                    const int pli = A1(PL, i);
                    #pragma unroll
                    for (int u = 0; u < RB; u++) {
                        const double den = fd_r<SQ>(fi, pli, pw[u], c[u]);
                        if (mi[u] != 0 && !may_exceed(zi, den, mv[u])) continue;
                        double vv = fabs(zi / den);
                        if (mi[u] == 0 || vv > mv[u]) { mv[u] = vv; mi[u] = i - k + 1; }
                    }
                }
                #pragma unroll
                for (int u = 0; u < RB; u++) {
                    const int j = j0 + u * NTH;
                    if (j > N) continue;
                    double jt = fabs(mv[u] * v[u]);
                    VI cc = {jt, j - k + 1};
                    VI nb = vi_better(best, cc);
                    if (nb.i == cc.i) bestrow.i = mi[u];
                    best = nb;
                }
            }
            VI win = block_argmax_t(best, red, tog, wave);
            bsync(wave);
            if (win.i > 0 && best.i == win.i) { s_ii = bestrow.i; s_jj = win.i; }
            bsync(wave);
            if (tid == 0) {
                int ii = s_ii, jj = s_jj;
                int a = ii + k - 1, b = jj + k - 1;
                double t;
                t = A1(F, k); A1(F, k) = A1(F, a); A1(F, a) = t;
                t = A1(Z, k); A1(Z, k) = A1(Z, a); A1(Z, a) = t;
                t = A1(A, k); A1(A, k) = A1(A, a); A1(A, a) = t;
                int u = A1(PL, k); A1(PL, k) = A1(PL, a); A1(PL, a) = u;
                short q = A1(PQ, k); A1(PQ, k) = A1(PQ, a); A1(PQ, a) = q;
                swapc(k, b);
            }
            bsync(wave);
            cw = croot(k);
            ukk = A1(Z, k) * vld(k) / fd_r<SQ>(A1(F, k), A1(PL, k), A1(PUs, k), cw);
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { Rk = k - 1; break; }
        }
        // ---- generator updates of the (k+1)-th Schur complement; unless a swap check
        // follows, the pass over the rows also runs step k+1's row search and the pass
        // over the columns its first column search
        const double Fk = A1(F, k);
        const int plk = A1(PL, k), puk = A1(PUs, k);
        const RootG ck = cw;
        const bool fuse = k + 1 <= mn && !(k % 10 == 0 && M - k > 0);
        const int fk1 = fuse ? A1(PL, k + 1) : 0;
        const double ff1 = fuse ? A1(F, k + 1) : 0.0;
        const int pun = fuse ? A1(PUs, k + 1) : puk;
        const RootG cn = fuse ? croot(k + 1) : cw;
        VI x1 = {0.0, 0};
        for (int j = k + 1 + tid; j <= M; j += NTH) {
            double temp = fd_r<SQ>(A1(F, j), A1(PL, j), puk, ck);
            temp = ppq<SQ>(A1(F, j), Fk) / temp;
            const double zj = A1(Z, j) * temp;
            A1(Z, j) = zj;
            if (fuse) x1 = vi_better(x1, VI{fabs(zj / fd_r<SQ>(A1(F, j), A1(PL, j), pun, cn)), j - k});
        }
        VI y1 = {0.0, 0};
        cpass(k + 1, [&](int j, int pw, double v, const RootG &c) {
            double temp = dd_r<SQ>(pw, c, puk, ck);
            double temp2 = -1.0 * fd_r<SQ>(Fk, plk, pw, c);
            temp = temp / temp2;
            const double vj = v * temp;
            if (j <= N) {
                vst(j, vj);
                if (fuse) y1 = vi_better(y1, VI{fabs(vj / fd_r<SQ>(ff1, fk1, pw, c)), j - k});
            }
        });
        // j = k - 1 on thread 0, which wrote W(k - 1) at the end of the previous step: a
        // fused step has no barrier between that write and this loop.  For the same
        // reason the factors of W(k) go to T1 and T2 on alternate steps (thread 0 may
        // still read the previous step's; the swap check uses both only after a barrier)
        double *Tk = (k & 1) ? T2 : T1;
        for (int j = k - 1 - tid; j >= 1; j -= NTH) {
            double temp = -1.0 * fd_r<SQ>(A1(F, j), A1(PL, j), puk, ck);
            A1(W, j) = A1(W, j) * (temp / ppq<SQ>(Fk, A1(F, j)));
            // the factors of W(k) (their gathers in parallel; the product stays in order)
            const int puj = A1(PUs, j);
            const double t2 = -1.0 * fd_r<SQ>(Fk, plk, puj, croot(j));
            A1(Tk, j) = t2 / ppq<SQ>(A1(F, j), Fk);
        }
        fl += 5ull * (M - k) + 7ull * (N - k) + 12ull * (k - 1);
        pre = fuse;
        // The reduction's barrier also publishes the update pass (T1 for thread 0's
        // product), and the next step reads W(k) only after its own first barrier, so a
        // fused step needs no other; before a swap check, and at the last step, W(k) is
        // published here.
        if (fuse) {
            block_argmax2_t(x1, y1, red, tog, wave);
            pre_cl = x1;
            pre_rw = y1;
            cnext = cn;
        } else {
            bsync(wave);
        }
        if (tid == 0) {
            double wk = fd_r<SQ>(Fk, plk, puk, ck) / A1(A, k);
            wk = ordered_prod(wk, Tk, k - 1);
            A1(W, k) = wk;
        }
        if (!fuse) bsync(wave);
        // ---- every 10 steps: swap while max |Mat_N| > rho (searchMax2Eig)
        if (k % 10 == 0 && M - k > 0) {
            for (;;) {
                VI x = {0.0, 0};
                for (int jc = 1 + tid; jc <= k; jc += NTH) {
                    double fj = A1(F, jc), mv = 0.0;
                    int mi = 0;
                    for (int i = 1; i <= M - k; i++) {
                        double v = fabs(A1(Z, k + i) / ppq<SQ>(A1(F, k + i), fj));
                        if (mi == 0 || v > mv) { mv = v; mi = i; }
                    }
                    double junk = mv * fabs(A1(W, jc));
                    VI c = {junk, jc};
                    VI nb = vi_better(x, c);
                    if (nb.i == jc) IP[jc - 1] = (short)mi;
                    x = nb;
                }
                VI win = block_argmax_t(x, red, tog, wave);
                if (!(win.v > rho)) break;
                const int ii = win.i, jj = IP[ii - 1] + k;
                fl += 10ull * (N - k) + 8ull * M + 18ull * k;
                bsync(wave);
                const double Fii = A1(F, ii), Fjj = A1(F, jj);
                const int plii = A1(PL, ii), pljj = A1(PL, jj);
                cpass(k + 1, [&](int j, int pw, double v, const RootG &c) {
                    double temp = -1.0 * fd_r<SQ>(Fii, plii, pw, c);
                    double temp2 = -1.0 * fd_r<SQ>(Fjj, pljj, pw, c);
                    temp = temp / temp2;
                    if (j <= N) vst(j, v * temp);
                });
                for (int j = k + 1 + tid; j <= M; j += NTH) {
                    if (j == jj) continue;
                    A1(Z, j) = A1(Z, j) * (ppq<SQ>(A1(F, j), Fjj) / ppq<SQ>(A1(F, j), Fii));
                }
                for (int j = 1 + tid; j <= k; j += NTH) {
                    if (j == ii) continue;
                    A1(W, j) = A1(W, j) * (ppq<SQ>(A1(F, j), Fii) / ppq<SQ>(A1(F, j), Fjj));
                    const int puj = A1(PUs, j);
                    const RootG cj = croot(j);
                    double temp = fd_r<SQ>(Fii, plii, puj, cj);
                    A1(T1, j) = ppq<SQ>(Fii, A1(F, j)) / temp;
                    temp = -1.0 * fd_r<SQ>(Fjj, pljj, puj, cj);
                    A1(T2, j) = temp / ppq<SQ>(A1(F, j), Fjj);
                }
                bsync(wave);
                if (tid == 0) {
                    const int puii = A1(PUs, ii);
                    const RootG cii = croot(ii);
                    double uj = A1(A, ii) * (ppq<SQ>(Fii, Fjj) / fd_r<SQ>(Fii, plii, puii, cii));
                    double wi = fd_r<SQ>(Fjj, pljj, puii, cii) / A1(A, jj);
                    for (int j = 1; j <= k; j++) {
                        if (j == ii) continue;
                        uj = uj * A1(T1, j);
                        wi = wi * A1(T2, j);
                    }
                    A1(Z, jj) = uj;
                    A1(W, ii) = wi;
                    int u = A1(PL, ii); A1(PL, ii) = A1(PL, jj); A1(PL, jj) = u;
                    short q = A1(PQ, ii); A1(PQ, ii) = A1(PQ, jj); A1(PQ, jj) = q;
                    double t = A1(F, ii); A1(F, ii) = A1(F, jj); A1(F, jj) = t;
                    t = A1(A, ii); A1(A, ii) = A1(A, jj); A1(A, jj) = t;
                }
                bsync(wave);
            }
        }
    }
    // ---- U generator (CauchylikeEigQ, then rows back to the node order)
    if (tid == 0) {
        atomicAdd(H.flops, fl);
        for (int i = 1; i <= M; i++) IP[A1(PQ, i) - 1] = i;  // invp
        int r = Rk;
        if (r < mn) {
            s_Rk = r;
            s_off = arena_take(H, (long)M * (r > 0 ? r : 1));
        } else {
            s_Rk = M;
            s_off = arena_take(H, (long)M * M);
        }
        H.usz[2 * (node - 1)] = M;
        H.usz[2 * (node - 1) + 1] = s_Rk;
        H.pu[node - 1] = s_off;
    }
    bsync(wave);
    const long off = s_off;
    if (off >= 0) {
        double *Ug = H.arena + off;
        const int r = s_Rk;
        if (Rk < mn) {
            for (int e = tid; e < M * r; e += NTH) {
                int i = e % M + 1, j = e / M + 1;
                int src = IP[i - 1];  // row src of Q
                double q;
                if (src <= r) q = (src == j) ? 1.0 : 0.0;
                else q = A1(Z, src) * A1(W, j) / ppq<SQ>(A1(F, src), A1(F, j));
                Ug[e] = q;
            }
        } else {
            for (int e = tid; e < M * M; e += NTH) {
                int i = e % M + 1, j = e / M;
                Ug[e] = (IP[i - 1] - 1 == j) ? 1.0 : 0.0;
            }
        }
    }
    // permuted row generators and their indices back to the compacted arrays
    for (int i = tid; i < M; i += NTH) {
        A1(H.U, l0 + i) = A[i];
        A1(H.PL, l0 + i) = PL[i];
    }
}

// ---------------------------------------------------------------------------
// Column compression of one node: rrluCauchyEigcol + CauchylikeEigQ2
// (ComprcauchyEig 'c').  RRLU rows: the node's M columns (roots, alpha, LDS);
// RRLU columns: the N row generators outside the node (private copies, as in the row
// compression; their passes, the mirrored positions 1..min(M, N) and the fused next
// row search work the same way).
// ---------------------------------------------------------------------------
template <int HSS_CWG, bool SQ>
__global__ __launch_bounds__(HSS_CWG) void k_hss_colcomp(const HssMerge *hm, const HssNode *nodes, int nn,
                                                    double tol, int lds_cap)
{
    extern __shared__ double hss_dyn[];  // lds_cap doubles: V when the node's rows fit
    __shared__ double Dr[HSS_MAXM], A[HSS_MAXM], Z[HSS_MAXM], W[HSS_MAXM];
    // factors of thread 0's ordered products; during the swap check's scan, the scanned
    // rows' root data
    __shared__ double T12[2 * HSS_MAXM];
    double *const T1 = T12, *const T2 = T12 + HSS_MAXM;
    __shared__ int PL[HSS_MAXM], PUs[HSS_MAXM];
    __shared__ short PQ[HSS_MAXM], IP[HSS_MAXM];   // positions within the node (<= HSS_MAXM)
    __shared__ VI red[4 * (HSS_CWG / 64)];
    int tog = 0;  // which half of red the next reduction uses
    __shared__ long s_off;
    __shared__ int s_M, s_N, s_Rk, s_flag, s_ii, s_jj;
    const int tid = threadIdx.x;
    constexpr int RB = HSS_CWG >= 512 ? 8 : 4;
    if ((int)blockIdx.x >= nn) return;
    HssNode Nd = nodes[blockIdx.x];
    const HssMerge H = hm[Nd.mrg];
    int *lr = H.lr, *lc = H.lc;
    const int node = Nd.node;
    if (tid == 0) {
        int M = lc[2 * (node - 1) + 1] - lc[2 * (node - 1)] + 1;
        int nir = lr[2 * (node - 1) + 1] - lr[2 * (node - 1)] + 1;
        s_M = M;
        s_N = *H.lda - nir;
        s_flag = (*H.err != 0 || M > HSS_MAXM || M < 1 || nir < 1 || s_N < 0) ? 1 : 0;
        s_off = s_flag ? -1 : scratch_take(H, 2L * (s_N > 0 ? s_N : 1));
        if (s_flag) atomicOr(H.err, 1);
    }
    __syncthreads();
    if (s_flag || s_off < 0) return;
    const int M = s_M, N = s_N, c0 = lc[2 * (node - 1)], r0 = lr[2 * (node - 1)], r1e = lr[2 * (node - 1) + 1];
    const int mn = M < N ? M : N;
    // Private per outside row: its generator V (updated every step), pole index PU and
    // pole value FF(PU).  Kept in the dynamic LDS as far as it holds them: V; V and PU;
    // or V, PU, the pole values and the root values of the node's own M columns (FF,
    // DIFL, DIFR and D0 at PL and PL + 1), so that a small node's passes, swaps and row
    // searches stay in LDS.  Otherwise V and PU are global and the values gathered.
    constexpr int RD = SQ ? 5 : 4;
    gdbl *FF = (gdbl *)H.FF;
    const int NS = N > 0 ? N : 1;
    const bool vl = N <= lds_cap;  // all three uniform over the block
    const bool pl = (long)N + (N + 1) / 2 <= lds_cap;
    const bool sm = 2L * N + (N + 1) / 2 + (long)RD * M <= lds_cap;
    // A node whose data all sit in LDS and whose outside part is small runs on one
    // wavefront: its steps need no block barrier (the other waves leave here)
    const bool wave = sm && N <= 384;
    if (wave && tid >= 64) return;
    const int NTH = wave ? 64 : HSS_CWG;
    ldbl *Vl = (ldbl *)hss_dyn;
    ldbl *FFl = Vl + NS;     // sm: FF(PU(j))
    ldbl *RRl = FFl + NS;    // sm: f0, f1, dl, dr (and s) of the node's M columns, M apart
    __attribute__((address_space(3))) int *PUl =
        (__attribute__((address_space(3))) int *)(sm ? RRl + RD * M : Vl + NS);
    gdblw *Vg = (gdblw *)(H.scratch + s_off);
    gint *PUg = (gint *)(H.scratch + s_off + NS);
    auto vld = [&](int j) -> double { return vl ? A1(Vl, j) : A1(Vg, j); };
    auto vst = [&](int j, double x) {
        if (vl) A1(Vl, j) = x;
        else A1(Vg, j) = x;
    };
    auto pu = [&](int j) -> int { return pl ? A1(PUl, j) : A1(PUg, j); };
    auto pu_set = [&](int j, int x) {
        if (pl) A1(PUl, j) = x;
        else A1(PUg, j) = x;
    };
    // the root values of the node's column at position i (PL(i))
    auto rroot = [&](int i) -> RootG {
        if (sm)
            return RootG{RRl[i - 1], RRl[M + i - 1], RRl[2 * M + i - 1], RRl[3 * M + i - 1],
                         SQ ? RRl[4 * M + i - 1] : 0.0};
        return root_load<SQ>(H, A1(PL, i));
    };
    for (int j = tid; j < N; j += NTH) {
        int src = (j < r0 - 1) ? j + 1 : j + 1 + (r1e - r0 + 1);
        vst(j + 1, A1((gdbl *)H.U, src));
        const int p = A1((gcint *)H.PL, src);
        pu_set(j + 1, p);
        if (j < mn) PUs[j] = p;
        if (sm) FFl[j] = A1(FF, p);
    }
    // outside-row swap (one thread; a barrier follows)
    auto swapc = [&](int a, int c) {
        double t = vld(a); vst(a, vld(c)); vst(c, t);
        int u = pu(a), w = pu(c);
        pu_set(a, w); pu_set(c, u);
        if (a <= mn) A1(PUs, a) = w;
        if (c <= mn) A1(PUs, c) = u;
        if (sm) { t = A1(FFl, a); A1(FFl, a) = A1(FFl, c); A1(FFl, c) = t; }
    };
    // node-column swap of the root values (one thread, beside the swaps of Dr, Z, A, PL, PQ)
    auto swapr = [&](int a, int c) {
        if (sm) {
            #pragma unroll
            for (int q = 0; q < RD; q++) {
                double t = RRl[q * M + a - 1]; RRl[q * M + a - 1] = RRl[q * M + c - 1]; RRl[q * M + c - 1] = t;
            }
        }
    };
    auto cload = [&](int j0, int *pd, double *v) {
        #pragma unroll
        for (int u = 0; u < RB; u++) {
            const int j = min(j0 + u * NTH, N);
            pd[u] = pu(j);
            v[u] = vld(j);
        }
    };
    auto bload = [&](int j0, int *pd, double *v, double *ffd) {
        cload(j0, pd, v);
        #pragma unroll
        for (int u = 0; u < RB; u++) ffd[u] = sm ? A1(FFl, min(j0 + u * NTH, N)) : A1(FF, pd[u]);
    };
    // body(j, pd, v, FF(pd)) for the outside rows j in [lo, N]; j may exceed N
    auto cpass = [&](int lo, auto body) {
        int j0 = lo + tid;
        if (j0 > N) return;
        if (sm) {
            for (int j = j0; j <= N; j += NTH) body(j, A1(PUl, j), A1(Vl, j), A1(FFl, j));
            return;
        }
        int pd[RB];
        double v[RB];
        cload(j0, pd, v);
        for (;;) {
            double ffd[RB];
            #pragma unroll
            for (int u = 0; u < RB; u++) ffd[u] = A1(FF, pd[u]);
            const int jn = j0 + RB * NTH;
            int pn[RB];
            double vn[RB];
            if (jn <= N) cload(jn, pn, vn);
            #pragma unroll
            for (int u = 0; u < RB; u++) body(j0 + u * NTH, pd[u], v[u], ffd[u]);
            if (jn > N) break;
            j0 = jn;
            #pragma unroll
            for (int u = 0; u < RB; u++) { pd[u] = pn[u]; v[u] = vn[u]; }
        }
    };
    // max over j in [k, N] of |V(j) / fd(PU(j), pw)| (index j - k + 1), w = root values of pw
    auto colsearch = [&](int k, int pw, const RootG &w) -> VI {
        VI y = {0.0, 0};
        cpass(k, [&](int j, int pd, double v, double ffd) {
            if (j <= N) y = vi_better(y, VI{fabs(v / fd_r<SQ>(ffd, pd, pw, w)), j - k + 1});
        });
        return y;
    };
    // max over i in [k, M] of |Z(i) / fd(pd, PL(i))|, ffd = FF(pd)
    auto rowsearch = [&](int k, int pd, double ffd) -> VI {
        VI x = {0.0, 0};
        for (int i = k + tid; i <= M; i += NTH)
            x = vi_better(x, VI{fabs(A1(Z, i) / fd_r<SQ>(ffd, pd, A1(PL, i), rroot(i))), i - k + 1});
        return x;
    };
    for (int i = tid; i < M; i += NTH) {
        Dr[i] = A1((gdbl *)H.D, c0 + i);
        A[i] = A1((gdbl *)H.V, c0 + i);
        PL[i] = A1((gcint *)H.PU, c0 + i);
        Z[i] = A[i];
        PQ[i] = (short)(i + 1);
        W[i] = 0.0;
        if (sm) {
            const RootG r = root_load<SQ>(H, PL[i]);
            RRl[i] = r.f0; RRl[M + i] = r.f1; RRl[2 * M + i] = r.dl; RRl[3 * M + i] = r.dr;
            if (SQ) RRl[4 * M + i] = r.s;
        }
    }
    bsync(wave);
    double amax = 0.0;
    int Rk = mn;
    const double rho = 1.5;
    bool pre = false;     // pre_cl, pre_rw hold this step's first row and column searches
    VI pre_cl = {0.0, 0}, pre_rw = {0.0, 0};
    RootG cnext = {0.0, 0.0, 0.0, 0.0, 0.0};  // root data of PL(k + 1) when pre
    double ffnext = 0.0;                      // FF(PU(k + 1)) when pre
    // flop count of the reference's model (rrluCauchyEigcol, CauchyHssEig.f90:321, 371,
    // 460, and its pivot searches), added to the merge's total by thread 0
    unsigned long long fl = 0;
    for (int k = 1; k <= mn; k++) {
        // ---- CauchyPivtEigm
        fl += 4ull * (M + N - 2 * k + 3);
        RootG cr;    // root data of PL(k), kept through the step
        double ffp;  // FF(PU(k)), kept through the step
        {
            VI cl, rw;
            if (pre) {
                cl = pre_cl;
                rw = pre_rw;
                cr = cnext;
                ffp = ffnext;
            } else {
                const int puk = A1(PUs, k), plk = A1(PL, k);
                ffp = A1(FF, puk);
                cr = rroot(k);
                cl = rowsearch(k, puk, ffp);
                rw = colsearch(k, plk, cr);
                block_argmax2_t(cl, rw, red, tog, wave);
            }
            double junkL = cl.v * fabs(vld(k)), junkU = rw.v * fabs(A1(Z, k));
            double den;
            {
                const int plk = A1(PL, k), puk = A1(PUs, k);
                if (plk < puk) den = (cr.f1 - ffp) + cr.dr;
                else den = (cr.f0 - ffp) + cr.dl;
                if (SQ) den = den * (cr.s + ffp);
            }
            double piv = fabs(A1(Z, k) * vld(k) / den);
            if (!(junkL <= piv && junkU <= piv)) {
                int flg = (junkL > junkU) ? 1 : 0;
                int jjL = cl.i, jjU = rw.i;
                // barriers as in the row compression: one before the first swap, then a
                // row swap after the column search it feeds
                bsync(wave);
                for (;;) {
                    if (flg == 1) {
                        const int jr = jjL + k - 1;
                        const int plr = A1(PL, jr);
                        cr = rroot(jr);
                        jjU = block_argmax_t(colsearch(k, plr, cr), red, tog, wave).i;
                        fl += 4ull * (N - k + 1);
                        if (tid == 0) {
                            double t;
                            t = A1(Dr, k); A1(Dr, k) = A1(Dr, jr); A1(Dr, jr) = t;
                            t = A1(Z, k); A1(Z, k) = A1(Z, jr); A1(Z, jr) = t;
                            t = A1(A, k); A1(A, k) = A1(A, jr); A1(A, jr) = t;
                            int u = A1(PL, k); A1(PL, k) = A1(PL, jr); A1(PL, jr) = u;
                            short q = A1(PQ, k); A1(PQ, k) = A1(PQ, jr); A1(PQ, jr) = q;
                            swapr(k, jr);
                        }
                        if (jjU == 1) {
                            bsync(wave);
                            break;
                        }
                    }
                    if (tid == 0) swapc(k, jjU + k - 1);
                    bsync(wave);
                    const int puk2 = A1(PUs, k);
                    ffp = A1(FF, puk2);
                    jjL = block_argmax_t(rowsearch(k, puk2, ffp), red, tog, wave).i;
                    fl += 4ull * (M - k + 1);
                    if (jjL == 1) break;
                    flg = 1;
                }
            }
        }
        // every swap of the search is followed by a barrier and the search only reads, so
        // the update pass below may start without another
        double ukk = A1(Z, k) * vld(k) / (-1.0 * fd_r<SQ>(ffp, A1(PUs, k), A1(PL, k), cr));
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            // ---- CauchyPivtEigm_CP
            fl += 4ull * (M - k + 1) * (N - k + 1);
            VI best = {0.0, 0};
            int bestrow = 0;
            for (int j0 = k + tid; j0 <= N; j0 += RB * NTH) {
                int pd[RB];
                double v[RB], ffd[RB], mv[RB];
                int mi[RB];
                bload(j0, pd, v, ffd);
                #pragma unroll
                for (int u = 0; u < RB; u++) { mv[u] = 0.0; mi[u] = 0; }
                for (int i = k; i <= M; i++) {
                    const double zi = A1(Z, i);
                    const int pli = A1(PL, i);
                    const RootG ri = rroot(i);
                    #pragma unroll
                    for (int u = 0; u < RB; u++) {
                        const double den = fd_r<SQ>(ffd[u], pd[u], pli, ri);
                        if (mi[u] != 0 && !may_exceed(zi, den, mv[u])) continue;
                        double vv = fabs(zi / den);
                        if (mi[u] == 0 || vv > mv[u]) { mv[u] = vv; mi[u] = i - k + 1; }
                    }
                }
                #pragma unroll
                for (int u = 0; u < RB; u++) {
                    const int j = j0 + u * NTH;
                    if (j > N) continue;
                    VI c = {fabs(mv[u] * v[u]), j - k + 1};
                    VI nb = vi_better(best, c);
                    if (nb.i == c.i) bestrow = mi[u];
                    best = nb;
                }
            }
            VI win = block_argmax_t(best, red, tog, wave);
            if (win.i > 0 && best.i == win.i) { s_ii = bestrow; s_jj = win.i; }
            bsync(wave);
            if (tid == 0) {
                int a = s_ii + k - 1, b = s_jj + k - 1;
                double t;
                t = A1(Dr, k); A1(Dr, k) = A1(Dr, a); A1(Dr, a) = t;
                t = A1(Z, k); A1(Z, k) = A1(Z, a); A1(Z, a) = t;
                t = A1(A, k); A1(A, k) = A1(A, a); A1(A, a) = t;
                int u = A1(PL, k); A1(PL, k) = A1(PL, a); A1(PL, a) = u;
                short q = A1(PQ, k); A1(PQ, k) = A1(PQ, a); A1(PQ, a) = q;
                swapr(k, a);
                swapc(k, b);
            }
            bsync(wave);
            cr = rroot(k);
            ffp = A1(FF, A1(PUs, k));
            ukk = A1(Z, k) * vld(k) / (-1.0 * fd_r<SQ>(ffp, A1(PUs, k), A1(PL, k), cr));
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { Rk = k - 1; break; }
        }
        const int plk = A1(PL, k), puk = A1(PUs, k);
        const double Fk = ffp;
        const RootG ck = cr;
        const bool fuse = k + 1 <= mn && !(k % 10 == 0 && M - k > 0);
        const int plk1 = fuse ? A1(PL, k + 1) : plk;
        const int pun = fuse ? A1(PUs, k + 1) : puk;
        const RootG ck1 = fuse ? rroot(k + 1) : ck;
        const double ffn = A1(FF, pun);
        VI x1 = {0.0, 0};
        for (int j = k + 1 + tid; j <= M; j += NTH) {
            const int plj = A1(PL, j);
            const RootG rj = rroot(j);
            double temp = dd_r<SQ>(plj, rj, plk, ck);
            double temp2 = -1.0 * fd_r<SQ>(Fk, puk, plj, rj);
            temp = temp / temp2;
            const double zj = A1(Z, j) * temp;
            A1(Z, j) = zj;
            if (fuse) x1 = vi_better(x1, VI{fabs(zj / fd_r<SQ>(ffn, pun, plj, rj)), j - k});
        }
        VI y1 = {0.0, 0};
        cpass(k + 1, [&](int j, int pd, double v, double ffd) {
            double temp = fd_r<SQ>(ffd, pd, plk, ck);
            const double vj = v * (ppq<SQ>(ffd, Fk) / temp);
            if (j <= N) {
                vst(j, vj);
                if (fuse) y1 = vi_better(y1, VI{fabs(vj / fd_r<SQ>(ffd, pd, plk1, ck1)), j - k});
            }
        });
        // j = k - 1 on thread 0, and the factors of W(k) in T1 or T2 by the step's
        // parity, as in the row compression
        double *Tk = (k & 1) ? T2 : T1;
        for (int j = k - 1 - tid; j >= 1; j -= NTH) {
            const int plj = A1(PL, j), puj = A1(PUs, j);
            const RootG rj = rroot(j);
            double temp = fd_r<SQ>(Fk, puk, plj, rj);
            double temp2 = dd_r<SQ>(plk, ck, plj, rj);
            temp = temp / temp2;
            A1(W, j) = A1(W, j) * temp;
            // the factors of W(k) (their gathers in parallel; the product stays in order)
            temp = fd_r<SQ>(sm ? A1(FFl, j) : A1(FF, puj), puj, plk, ck);
            temp2 = dd_r<SQ>(plj, rj, plk, ck);
            A1(Tk, j) = temp / temp2;
        }
        fl += 6ull * (M - k) + 5ull * (N - k) + 14ull * (k - 1);
        pre = fuse;
        // as in the row compression: the reduction's barrier publishes the update pass
        if (fuse) {
            block_argmax2_t(x1, y1, red, tog, wave);
            pre_cl = x1;
            pre_rw = y1;
            cnext = ck1;
            ffnext = ffn;
        } else {
            bsync(wave);
        }
        if (tid == 0) {
            double wk = (-1.0 * fd_r<SQ>(Fk, puk, plk, ck)) / A1(A, k);
            wk = ordered_prod(wk, Tk, k - 1);
            A1(W, k) = wk;
        }
        if (!fuse) bsync(wave);
        if (k % 10 == 0 && M - k > 0) {
            for (;;) {
                // the scanned rows' root data, gathered once into T1 / T2 when they fit
                const int L = M - k;
                const bool stg = !sm && L * (SQ ? 5 : 4) <= 2 * HSS_MAXM;
                if (stg) {
                    for (int q = tid; q < L; q += NTH) {
                        const RootG r = root_load<SQ>(H, A1(PL, k + 1 + q));
                        T12[q] = r.f0;
                        T12[L + q] = r.f1;
                        T12[2 * L + q] = r.dl;
                        T12[3 * L + q] = r.dr;
                        if (SQ) T12[4 * L + q] = r.s;
                    }
                    bsync(wave);
                }
                VI x = {0.0, 0};
                for (int jc = 1 + tid; jc <= k; jc += NTH) {
                    int pljc = A1(PL, jc);
                    const RootG rc = rroot(jc);
                    double mv = 0.0;
                    int mi = 0;
                    for (int i = 1; i <= L; i++) {
                        double tt;
                        if (sm) {
                            tt = dd_r<SQ>(A1(PL, i + k), rroot(i + k), pljc, rc);
                        } else if (stg) {
                            const RootG ri = {T12[i - 1], T12[L + i - 1], T12[2 * L + i - 1], T12[3 * L + i - 1],
                                              SQ ? T12[4 * L + i - 1] : 0.0};
                            tt = dd_r<SQ>(A1(PL, i + k), ri, pljc, rc);
                        } else {
                            tt = ddq<SQ>(H, A1(PL, i + k), pljc);
                        }
                        double v = fabs(A1(Z, k + i) / tt);
                        if (mi == 0 || v > mv) { mv = v; mi = i; }
                    }
                    VI c = {mv * fabs(A1(W, jc)), jc};
                    VI nb = vi_better(x, c);
                    if (nb.i == jc) IP[jc - 1] = (short)mi;
                    x = nb;
                }
                VI win = block_argmax_t(x, red, tog, wave);
                if (!(win.v > rho)) break;
                const int ii = win.i, jj = IP[ii - 1] + k;
                fl += 10ull * (N - k) + 12ull * M + 22ull * k;
                bsync(wave);
                const int plii = A1(PL, ii), pljj = A1(PL, jj);
                const RootG cii = rroot(ii), cjj = rroot(jj);
                cpass(k + 1, [&](int j, int pd, double v, double ffd) {
                    double temp = fd_r<SQ>(ffd, pd, plii, cii);
                    double temp2 = fd_r<SQ>(ffd, pd, pljj, cjj);
                    temp = temp / temp2;
                    if (j <= N) vst(j, v * temp);
                });
                for (int j = k + 1 + tid; j <= M; j += NTH) {
                    if (j == jj) continue;
                    double temp = ddq<SQ>(H, A1(PL, j), pljj);
                    double temp2 = ddq<SQ>(H, A1(PL, j), plii);
                    temp = temp / temp2;
                    A1(Z, j) = A1(Z, j) * temp;
                }
                for (int j = 1 + tid; j <= k; j += NTH) {
                    if (j == ii) continue;
                    const int puj = A1(PUs, j);
                    double temp = ddq<SQ>(H, A1(PL, j), plii);
                    double temp2 = ddq<SQ>(H, A1(PL, j), pljj);
                    temp = temp / temp2;
                    A1(W, j) = A1(W, j) * temp;
                    temp = ddq<SQ>(H, plii, A1(PL, j));
                    temp2 = -1.0 * fdq<SQ>(H, puj, plii);
                    A1(T1, j) = temp / temp2;
                    temp = fdq<SQ>(H, puj, pljj);
                    temp2 = ddq<SQ>(H, A1(PL, j), pljj);
                    A1(T2, j) = temp / temp2;
                }
                bsync(wave);
                if (tid == 0) {
                    const int puii = A1(PUs, ii);
                    double temp = ddq<SQ>(H, plii, pljj);
                    double temp2 = -1.0 * fdq<SQ>(H, puii, plii);
                    double uj = A1(A, ii) * (temp / temp2);
                    double wi = (-1.0 * fdq<SQ>(H, puii, pljj)) / A1(A, jj);
                    for (int j = 1; j <= k; j++) {
                        if (j == ii) continue;
                        uj = uj * A1(T1, j);
                        wi = wi * A1(T2, j);
                    }
                    A1(Z, jj) = uj;
                    A1(W, ii) = wi;
                    int u = A1(PL, ii); A1(PL, ii) = A1(PL, jj); A1(PL, jj) = u;
                    short q = A1(PQ, ii); A1(PQ, ii) = A1(PQ, jj); A1(PQ, jj) = q;
                    swapr(ii, jj);
                    double t = A1(Dr, ii); A1(Dr, ii) = A1(Dr, jj); A1(Dr, jj) = t;
                    t = A1(A, ii); A1(A, ii) = A1(A, jj); A1(A, jj) = t;
                }
                bsync(wave);
            }
        }
    }
    // ---- V generator (CauchylikeEigQ2, columns back to the node order)
    if (tid == 0) {
        atomicAdd(H.flops, fl);
        for (int i = 1; i <= M; i++) IP[A1(PQ, i) - 1] = i;
        if (Rk < mn) {
            s_Rk = Rk;
            s_off = arena_take(H, (long)(Rk > 0 ? Rk : 1) * M);
        } else {
            s_Rk = M;
            s_off = arena_take(H, (long)M * M);
        }
        H.vsz[2 * (node - 1)] = s_Rk;
        H.vsz[2 * (node - 1) + 1] = M;
        H.pv[node - 1] = s_off;
    }
    bsync(wave);
    const long off = s_off;
    if (off >= 0) {
        double *Vg = H.arena + off;
        const int r = s_Rk;
        if (Rk < mn) {
            for (int e = tid; e < r * M; e += NTH) {
                int i = e % r + 1, j = e / r + 1;
                int src = IP[j - 1];  // column src of Q
                double q;
                if (src <= r) {
                    q = (src == i) ? 1.0 : 0.0;
                } else {
                    int pdjj = A1(PL, src), pdii = A1(PL, i);
                    double den;
                    if (pdii < pdjj) den = (A1(H.FF, pdii + 1) - A1(H.FF, pdjj)) + (A1(H.DIFR, pdii) - A1(H.DIFL, pdjj));
                    else den = (A1(H.FF, pdii) - A1(H.FF, pdjj + 1)) + (A1(H.DIFL, pdii) - A1(H.DIFR, pdjj));
                    if (SQ) den = den * (A1(H.D0, pdii) + A1(H.D0, pdjj));
                    q = -1.0 * A1(W, i) * A1(Z, src) / den;
                }
                Vg[e] = q;
            }
        } else {
            for (int e = tid; e < M * M; e += NTH) {
                int i = e % M, j = e / M + 1;
                Vg[e] = (IP[j - 1] - 1 == i) ? 1.0 : 0.0;
            }
        }
    }
    for (int i = tid; i < M; i += NTH) {
        A1(H.D, c0 + i) = Dr[i];
        A1(H.V, c0 + i) = A[i];
        A1(H.PU, c0 + i) = PL[i];
    }
}

// ---------------------------------------------------------------------------
// Compaction after a level (one block per merge): keep the first rank entries of
// every node of the level, then the full ranges of shallower leaves, in order.
// rows: U/PL with usz(2,:); cols: V/D/PU with vsz(1,:).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_hss_compact(const HssMerge *hm, const HssCompact *cp, const int *order,
                                                    int rows)
{
    // Every node's count and source start read in parallel, one thread per node; the
    // start of each node's slice by a prefix over LDS; then one pass over all elements,
    // each finding its node by binary search over the slice starts.
    __shared__ int s_start[1024 + 1], s_src[1024];
    HssCompact C = cp[blockIdx.x];
    const HssMerge H = hm[C.mrg];
    const int tid = threadIdx.x;
    int *rng = rows ? H.lr : H.lc;
    if (*H.err) return;
    // the host builds trees of at most 512 leaves, so a compaction lists at most 1024
    // nodes; a larger one would not fit the slice table and sends the merge dense
    if (C.count > 1024) {
        if (tid == 0) atomicOr(H.err, 1);
        return;
    }
    for (int q = tid; q < C.count; q += WG) {
        const int nd = order[C.start + q];
        const int r0 = rng[2 * (nd - 1)], have = rng[2 * (nd - 1) + 1] - r0 + 1;
        int cnt = have;
        if (q < C.nlevel) cnt = rows ? H.usz[2 * (nd - 1) + 1] : H.vsz[2 * (nd - 1)];
        if (cnt < 0 || cnt > have) cnt = 0;
        s_start[q] = cnt;
        s_src[q] = r0;  // 1-based
    }
    __syncthreads();
    if (tid == 0) {
        int lt = 0;
        for (int q = 0; q < C.count; q++) {
            const int c = s_start[q];
            s_start[q] = lt;
            lt += c;
        }
        s_start[C.count] = lt;
    }
    __syncthreads();
    double *Td = H.T, *Td2 = H.T + H.K;
    int *Ti = (int *)(H.T + 2 * H.K);
    const int tot = s_start[C.count];
    for (int d = tid; d < tot; d += WG) {
        int lo = 0, hi = C.count - 1;
        while (lo < hi) {
            const int md = (lo + hi + 1) >> 1;
            if (s_start[md] <= d) lo = md;
            else hi = md - 1;
        }
        const int sidx = s_src[lo] + (d - s_start[lo]);
        if (rows) { Td[d] = A1(H.U, sidx); Ti[d] = A1(H.PL, sidx); }
        else { Td[d] = A1(H.V, sidx); Td2[d] = A1(H.D, sidx); Ti[d] = A1(H.PU, sidx); }
    }
    __syncthreads();
    for (int d = tid; d < tot; d += WG) {
        if (rows) { H.U[d] = Td[d]; H.PL[d] = Ti[d]; }
        else { H.V[d] = Td[d]; H.D[d] = Td2[d]; H.PU[d] = Ti[d]; }
    }
    for (int q = tid; q < C.count; q += WG) {
        const int nd = order[C.start + q];
        rng[2 * (nd - 1)] = s_start[q] + 1;
        rng[2 * (nd - 1) + 1] = s_start[q + 1];
    }
    if (tid == 0) {
        if (rows) *H.lda = tot; else *H.lca = tot;
    }
}

// B generators of the nodes of one level: B_i = C(row skeleton of i, column
// skeleton of its sibling) (CauchyHssEig.f90:2229-2244).
__global__ __launch_bounds__(WG) void k_hss_bgen(const HssMerge *hm, const HssNode *nodes, int nn)
{
    __shared__ long s_off;
    if ((int)blockIdx.x >= nn) return;
    HssNode Nd = nodes[blockIdx.x];
    const HssMerge H = hm[Nd.mrg];
    const int node = Nd.node, sib = Nd.sib;
    const int r0 = H.lr[2 * (node - 1)], mi = H.lr[2 * (node - 1) + 1] - r0 + 1;
    const int c0 = H.lc[2 * (sib - 1)], nj = H.lc[2 * (sib - 1) + 1] - c0 + 1;
    if (*H.err || mi < 0 || nj < 0) return;
    if (threadIdx.x == 0) {
        s_off = arena_take(H, (long)mi * nj);
        H.bsz[2 * (node - 1)] = mi;
        H.bsz[2 * (node - 1) + 1] = nj;
        H.pb[node - 1] = s_off;
    }
    __syncthreads();
    if (s_off < 0) return;
    double *B = H.arena + s_off;
    for (int e = threadIdx.x; e < mi * nj; e += WG) {
        int i = e % mi, j = e / mi;
        int pf = A1(H.PL, r0 + i), pw = A1(H.PU, c0 + j);
        double fj = A1(H.FF, pw), vj = A1(H.V, c0 + j);
        double den = (pf <= pw) ? (A1(H.FF, pf) - fj) - A1(H.DIFL, pw)
                                : (A1(H.FF, pf) - A1(H.FF, pw + 1)) - A1(H.DIFR, pw);
        if (H.sq) den = den * (A1(H.FF, pf) + A1(H.D0, pw));
        B[e] = A1(H.U, r0 + i) * vj / den;
    }
}

// initial compacted arrays: D = roots, U = z, V = alpha, PL = PU = identity
__global__ __launch_bounds__(WG) void k_hss_init(const HssMerge *hm)
{
    const HssMerge H = hm[blockIdx.y];
    for (int i = blockIdx.x * WG + threadIdx.x; i < H.K; i += gridDim.x * WG) {
        H.D[i] = H.D0[i]; H.U[i] = H.U0[i]; H.V[i] = H.V0[i];
        H.PL[i] = i + 1; H.PU[i] = i + 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { *H.lda = H.K; *H.lca = H.K; }
}

// ---------------------------------------------------------------------------
// Batched GEMM of the HSS multiply: work item (blockIdx.y) = one BG, block
// (blockIdx.x, blockIdx.z) = one 128 x 64 tile of its output.  FP64 MFMA 16x16x4.
// The k range is walked in chunks of 32: the block stages a 128 x 32 slab of A and
// a 32 x 64 slab of B in LDS (coalesced loads by all 256 threads), and while the
// four wavefronts multiply out of LDS the next chunk is already loading into
// registers.  Each wavefront owns 32 x 64 outputs in 8 accumulators; the k-steps
// of 4 run in order, as the single-product reference order of the tests expects.
// ---------------------------------------------------------------------------
typedef double d4h __attribute__((ext_vector_type(4)));
// The operand pointers of a BG are read from the item table, so the compiler cannot see
// that they point to global memory and would use FLAT loads, which count in LGKM_CNT as
// well as VM_CNT: every wait for the LDS operand reads would then also wait for the next
// chunk's loads.  Global-address-space views keep the two counters apart.
// what a staged element outside the operands reads
static __device__ double bg_zero = 0.0;
#define BG_BK 32
#define BG_AP 132  // LDS pitch of the A slab (128 rows + 4: no bank conflicts between k rows)
#define BG_BP 68   // LDS pitch of the B slab (64 columns + 4)
__global__ __launch_bounds__(WG) void k_bgemm(const BG *items)
{
    __shared__ double As[BG_BK * BG_AP];
    __shared__ double Bs[BG_BK * BG_BP];
    const BG g = items[blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rb = blockIdx.x * 128, c0 = blockIdx.z * 64;
    if (rb >= g.nrow || c0 >= g.n) return;  // the whole block
    const int K1 = g.k, K = g.k + g.k2;
    const long lda = g.lda, lda2 = g.lda2, ldb = g.ldb, ldb2 = g.ldb2;
    // staging of one chunk: A element (row rb + (i & 127), k kc + (i >> 7)), i = tid + 256 q;
    // B element i = tid + 256 q: (k kc + (i & 31), column c0 + (i >> 5)) column-major,
    // (k kc + (i >> 6), column c0 + (i & 63)) for btrans
    double ra[16], rbv[8];
    gdbl *const zero = (gdbl *)&bg_zero;  // unconditional global loads (k_bgemm_w)
    auto stage = [&](int kc) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int i = tid + 256 * q, row = rb + (i & 127), t = kc + (i >> 7);
            const bool ok = row < g.nrow && t < K;
            gdbl *p = t < K1 ? (gdbl *)g.A + (row + (long)t * lda) : (gdbl *)g.A2 + (row + (long)(t - K1) * lda2);
            ra[q] = *(ok ? p : zero);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int i = tid + 256 * q;
            int t, col;
            if (g.btrans) { t = kc + (i >> 6); col = c0 + (i & 63); }
            else { t = kc + (i & 31); col = c0 + (i >> 5); }
            const bool ok = col < g.n && t < K;
            gdbl *p = t >= K1 ? (gdbl *)g.B2 + ((long)col * ldb2 + (t - K1))
                              : (gdbl *)g.B + (g.btrans ? (long)t * ldb + col : (long)col * ldb + t);
            rbv[q] = *(ok ? p : zero);
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int i = tid + 256 * q;
            As[(i >> 7) * BG_AP + (i & 127)] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int i = tid + 256 * q;
            if (g.btrans) Bs[(i >> 6) * BG_BP + (i & 63)] = rbv[q];
            else Bs[(i & 31) * BG_BP + (i >> 5)] = rbv[q];
        }
    };
    const int lr = lane & 15, lk = lane >> 4, wr = wv * 32;
    d4h acc[2][4];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 4; b++) acc[a][b] = (d4h){0.0, 0.0, 0.0, 0.0};
    stage(0);
    commit();
    __syncthreads();
    for (int kc = 0; kc < K; kc += BG_BK) {
        const bool more = kc + BG_BK < K;
        if (more) stage(kc + BG_BK);  // in flight while this chunk multiplies
        const int ksteps = min(BG_BK, ((K - kc) + 3) & ~3);
        for (int kq = 0; kq < ksteps; kq += 4) {
            const double *ap = As + (kq + lk) * BG_AP + wr + lr;
            const double *bp = Bs + (kq + lk) * BG_BP + lr;
            const double a0 = ap[0], a1 = ap[16];
            const double b0 = bp[0], b1 = bp[16], b2 = bp[32], b3 = bp[48];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[0][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b2, acc[0][2], 0, 0, 0);
            acc[0][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b3, acc[0][3], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
            acc[1][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b2, acc[1][2], 0, 0, 0);
            acc[1][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b3, acc[1][3], 0, 0, 0);
        }
        __syncthreads();  // every wavefront is done with this chunk
        if (more) {
            commit();
            __syncthreads();
        }
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 4; b++) {
            const int col = c0 + b * 16 + lr;
            if (col >= g.n) continue;
            const long pc = g.cmap ? (long)((const __attribute__((address_space(1))) int *)g.cmap)[col] - g.cmoff : (long)col;
            for (int i = 0; i < 4; i++) {
                int row = rb + wr + a * 16 + lk + 4 * i;
                if (row < g.nrow) {
                    gdblw *cp = (gdblw *)g.C + ((long)row + pc * g.ldc);
                    *cp = g.beta ? (*cp + acc[a][b][i]) : acc[a][b][i];
                }
            }
        }
}

// ---------------------------------------------------------------------------
// The same batched GEMM with wide tiles for large items: block = BM x BN outputs,
// wavefronts of (16 TA) x (16 TB) each (TA x TB accumulators of 16 x 16), k in chunks of
// BK staged through LDS (DB: two slabs, one barrier per chunk) with the next chunk in
// registers meanwhile.  nz > 0 numbers the blocks with the column block fastest
// (blockIdx.x = row block * nz + column block, blockIdx.z = 0), so the blocks of one row
// panel of A run together and the panel comes from HBM once; nz = 0 takes the row block
// from blockIdx.x and the column block from blockIdx.z.  Measured (FP64 MFMA peak 78.6
// TF/s):
//   * square products of order 8192: 256 x 128 with sixteen 32 x 64 wavefronts 52.7 TF/s,
//     eight 64 x 64 wavefronts 45.7, k_bgemm 34.8; below about 3000 rows and columns
//     there are too few blocks to fill the 256 CUs and k_bgemm is faster;
//   * the HSS multiply's steps at n = 32768: the tile follows the step's widest item
//     (128 x 64, 128 x 128 or 256 x 144), 24-38 % faster than k_bgemm.
// The k-steps of 4 run in the same order as in k_bgemm, so all give the same bits.
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int DB, int TA, int TB>
__global__ __launch_bounds__((BM / (16 * TA)) * (BN / (16 * TB)) * 64) void k_bgemm_w(const BG *items, int nz)
{
    constexpr int WM = 16 * TA, WN = 16 * TB;  // rows and columns of one wavefront's tile
    constexpr int NT = (BM / WM) * (BN / WN) * 64;
    constexpr int AP = BM + 4, BP = BN + 4;
    // staged elements per thread; a slab that does not split evenly over the threads
    // (BN = 144) leaves the last round partly idle
    constexpr int QA = (BM * BK + NT - 1) / NT, QB = (BN * BK + NT - 1) / NT;
    constexpr bool EA = QA * NT == BM * BK, EB = QB * NT == BN * BK;
    __shared__ double Asb[(DB ? 2 : 1) * BK * AP];
    __shared__ double Bsb[(DB ? 2 : 1) * BK * BP];
    double *As = Asb, *Bs = Bsb;
    const BG g = items[blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int bx = nz ? (int)blockIdx.x / nz : (int)blockIdx.x, bz = nz ? (int)blockIdx.x % nz : (int)blockIdx.z;
    const int rb = bx * BM, c0 = bz * BN;
    if (rb >= g.nrow || c0 >= g.n) return;  // the whole block
    const int K1 = g.k, K = g.k + g.k2;
    const long lda = g.lda, lda2 = g.lda2, ldb = g.ldb, ldb2 = g.ldb2;
    // staging: A element i = tid + NT q at (row rb + i % BM, k kc + i / BM); B element i at
    // (k kc + i / BN, column c0 + i % BN) for btrans, (k kc + i % BK, column c0 + i / BK)
    // column-major
    double ra[QA], rbv[QB];
    // Every staged element is one unconditional global load, an element outside the
    // operands reading a global zero, so no load sits in a branch and nothing uses the
    // loaded value before the commit (a branch around loads, or a select on their values,
    // makes the compiler wait for them at once, which serialises the prefetch with the
    // MFMAs).
    gdbl *const zero = (gdbl *)&bg_zero;
    auto stage = [&](int kc) {
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int i = tid + NT * q, row = rb + (i % BM), t = kc + i / BM;
            const bool ok = (EA || i < BM * BK) && row < g.nrow && t < K;
            gdbl *p = t < K1 ? (gdbl *)g.A + (row + (long)t * lda) : (gdbl *)g.A2 + (row + (long)(t - K1) * lda2);
            ra[q] = *(ok ? p : zero);
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int i = tid + NT * q;
            int t, col;
            if (g.btrans) { t = kc + i / BN; col = c0 + i % BN; }
            else { t = kc + i % BK; col = c0 + i / BK; }
            const bool ok = (EB || i < BN * BK) && col < g.n && t < K;
            gdbl *p = t >= K1 ? (gdbl *)g.B2 + ((long)col * ldb2 + (t - K1))
                              : (gdbl *)g.B + (g.btrans ? (long)t * ldb + col : (long)col * ldb + t);
            rbv[q] = *(ok ? p : zero);
        }
    };
    auto commit = [&](double *Ad, double *Bd) {
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int i = tid + NT * q;
            if (EA || i < BM * BK) Ad[(i / BM) * AP + (i % BM)] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int i = tid + NT * q;
            if (!EB && i >= BN * BK) continue;
            if (g.btrans) Bd[(i / BN) * BP + (i % BN)] = rbv[q];
            else Bd[(i % BK) * BP + (i / BK)] = rbv[q];
        }
    };
    const int lr = lane & 15, lk = lane >> 4;
    const int wr = (wv % (BM / WM)) * WM, wc = (wv / (BM / WM)) * WN;
    d4h acc[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; a++)
#pragma unroll
        for (int b = 0; b < TB; b++) acc[a][b] = (d4h){0.0, 0.0, 0.0, 0.0};
    stage(0);
    commit(As, Bs);
    __syncthreads();
    for (int kc = 0; kc < K; kc += BK) {
        const bool more = kc + BK < K;
        if (more) stage(kc + BK);  // in flight while this chunk multiplies
        const int ksteps = min(BK, ((K - kc) + 3) & ~3);
        for (int kq = 0; kq < ksteps; kq += 4) {
            const double *ap = As + (kq + lk) * AP + wr + lr;
            const double *bp = Bs + (kq + lk) * BP + wc + lr;
            double av[TA], bv[TB];
#pragma unroll
            for (int a = 0; a < TA; a++) av[a] = ap[16 * a];
#pragma unroll
            for (int b = 0; b < TB; b++) bv[b] = bp[16 * b];
#pragma unroll
            for (int a = 0; a < TA; a++)
#pragma unroll
                for (int b = 0; b < TB; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        if (DB) {
            // the other slab was last read before the previous barrier: fill it now, then
            // one barrier makes it readable and frees this one
            if (more) {
                double *An = (As == Asb) ? Asb + BK * AP : Asb, *Bn = (Bs == Bsb) ? Bsb + BK * BP : Bsb;
                commit(An, Bn);
                As = An;
                Bs = Bn;
            }
            __syncthreads();
        } else {
            __syncthreads();  // every wavefront is done with this chunk
            if (more) {
                commit(As, Bs);
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int a = 0; a < TA; a++)
#pragma unroll
        for (int b = 0; b < TB; b++) {
            const int col = c0 + wc + b * 16 + lr;
            if (col >= g.n) continue;
            const long pc = g.cmap ? (long)((const __attribute__((address_space(1))) int *)g.cmap)[col] - g.cmoff : (long)col;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = rb + wr + a * 16 + lk + 4 * i;
                if (row < g.nrow) {
                    gdblw *cp = (gdblw *)g.C + ((long)row + pc * g.ldc);
                    *cp = g.beta ? (*cp + acc[a][b][i]) : acc[a][b][i];
                }
            }
        }
}
template __global__ void k_bgemm_w<256, 128, 16, 1, 2, 4>(const BG *, int);
template __global__ void k_bgemm_w<128, 128, 16, 1, 2, 4>(const BG *, int);
template __global__ void k_bgemm_w<128, 64, 16, 1, 1, 4>(const BG *, int);
template __global__ void k_bgemm_w<128, 48, 16, 1, 1, 3>(const BG *, int);
template __global__ void k_bgemm_w<128, 32, 16, 1, 1, 2>(const BG *, int);
template __global__ void k_bgemm_w<128, 96, 16, 1, 2, 3>(const BG *, int);
template __global__ void k_bgemm_w<256, 144, 16, 1, 2, 9>(const BG *, int);

// X(:, t) = Q2(:, t): the K surviving old eigenvectors of each HSS merge packed
// densely (mdlaed2.f90:467-499), zero outside each column's support.
__global__ __launch_bounds__(WG) void k_gatherx(const Merge *mg, const int *hidx, const int *col_start, int nh,
                                                const Prob *probs, Work W, int cur, double *X, const long *xoff)
{
    int h = 0;
    {
        int lo = 0, hi = nh - 1, b = blockIdx.x;
        while (lo < hi) { int md = (lo + hi + 1) >> 1; if (col_start[md] <= b) lo = md; else hi = md - 1; }
        h = lo;
    }
    const Merge M = mg[hidx[h]];
    const int t = blockIdx.x - col_start[h];
    const Prob P = probs[M.prob];
    const int slot = W.vm[cur][M.s + W.indx[M.s + t] - 1];
    const int lo = W.supp_lo[slot], hi = W.supp_hi[slot];
    const double *src = P.q + (long)(slot - P.off) * P.ld - P.off;
    double *dst = X + xoff[h] + (long)t * M.n - M.s;
    for (int r = M.s + threadIdx.x; r < M.s + M.n; r += WG) dst[r] = (r >= lo && r < hi) ? src[r] : 0.0;
}

// dense C(i,j) = u_i v_j / (f_i - lambda_j) with accurate differences
// (CauchylikeEig with identity permutations), used when a merge's HSS form cannot
// be built within the LDS / arena limits
__global__ __launch_bounds__(WG) void k_cbuild(const HssMerge *hm, int h, double *C)
{
    const HssMerge H = hm[h];
    long KK = (long)H.K * H.K;
    for (long e = (long)blockIdx.x * WG + threadIdx.x; e < KK; e += (long)gridDim.x * WG) {
        int i = (int)(e % H.K) + 1, j = (int)(e / H.K) + 1;
        double den = (i <= j) ? (A1(H.FF, i) - A1(H.FF, j)) - A1(H.DIFL, j)
                              : (A1(H.FF, i) - A1(H.FF, j + 1)) - A1(H.DIFR, j);
        if (H.sq) den = den * (A1(H.FF, i) + A1(H.D0, j));
        C[e] = A1(H.U0, i) * A1(H.V0, j) / den;
    }
}

__global__ void k_hss_reset(const HssMerge *hm)
{
    if (threadIdx.x == 0) *hm[blockIdx.x].scratch_top = 0ULL;
}

// 256 threads per node when a launch has nodes for every CU twice over, 512 when
// it has fewer (the upper tree levels): the search over the node's columns is the
// sequential part of every pivot step
template __global__ void k_hss_rowcomp<256, false>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_rowcomp<512, false>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_colcomp<256, false>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_colcomp<512, false>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_rowcomp<256, true>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_rowcomp<512, true>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_colcomp<256, true>(const HssMerge *, const HssNode *, int, double, int);
template __global__ void k_hss_colcomp<512, true>(const HssMerge *, const HssNode *, int, double, int);

// ===========================================================================
// General Cauchy-like matrices A(i,j) = u(i) v(j) / (d(i) - f(j)), direct
// differences: cauchy2hss (CauchyHss_VP2.f90:387-613).  The reference walks the
// tree in post order with a stack of compressed skeleton rows and columns, so a
// node's compression sees the skeletons of the nodes before it; that order fixes
// the ranks, and it is kept: one workgroup walks the tree, and every RRLU step
// (rrluCauchy, :7-134, rho = 1.02, a swap check every 10 steps) runs across the
// workgroup's threads, its scalar decisions taken identically by every thread from
// the same global values.  Statement order is the Fortran's.
// ===========================================================================
#define C2H_WG 1024

// CauchyMax (:166-188): max |u(i) / (D(i) - F)|, i = 1..n, first index
__device__ VI g_cmax(const double *D, double F, const double *u, int n, VI *red)
{
    VI x = {0.0, 0};
    for (int i = 1 + threadIdx.x; i <= n; i += C2H_WG) x = vi_better(x, VI{fabs(A1(u, i) / (A1(D, i) - F)), i});
    return block_argmax(x, red);
}

__device__ __forceinline__ void g_dswp(double *a, int i, int j) { double t = A1(a, i); A1(a, i) = A1(a, j); A1(a, j) = t; }
__device__ __forceinline__ void g_iswp(int *a, int i, int j) { int t = A1(a, i); A1(a, i) = A1(a, j); A1(a, j) = t; }

// CauchyPivtG (:191-243): rook pivoting on the k-th Schur complement
__device__ void g_pivt(double *D, double *F, double *u, double *v, int *PL, int *PU, int k, double *A, int M, int N,
                       VI *red)
{
    const VI L = g_cmax(D + k - 1, A1(F, k), u + k - 1, M - k + 1, red);
    const VI R = g_cmax(F + k - 1, A1(D, k), v + k - 1, N - k + 1, red);
    const double junkL = L.v * fabs(A1(v, k)), junkU = R.v * fabs(A1(u, k));
    const double piv = fabs(A1(u, k) * A1(v, k) / (A1(D, k) - A1(F, k)));
    if (junkL <= piv && junkU <= piv) return;
    int flg = (junkL > junkU) ? 1 : 0, jjL = L.i, jjU = R.i;
    // every round moves a strictly larger entry into the pivot, so finite data ends
    // within M + N rounds; the bound only guarantees an exit for overflowed generators
    for (int round = 0; round < M + N + 2; round++) {
        if (flg == 1) {
            const int jj = jjL + k - 1;
            __syncthreads();  // every wave has read the pivot entries it decided on
            if (threadIdx.x == 0) { g_dswp(D, k, jj); g_dswp(u, k, jj); g_dswp(A, k, jj); g_iswp(PL, k, jj); }
            __syncthreads();
            jjU = g_cmax(F + k - 1, A1(D, k), v + k - 1, N - k + 1, red).i;
            if (jjU == 1) return;
            flg = 0;
        }
        const int jj = jjU + k - 1;
        __syncthreads();
        if (threadIdx.x == 0) { g_dswp(F, k, jj); g_dswp(v, k, jj); g_iswp(PU, k, jj); }
        __syncthreads();
        jjL = g_cmax(D + k - 1, A1(F, k), u + k - 1, M - k + 1, red).i;
        if (jjL == 1) return;
        flg = 1;
    }
}

// CauchyPivtG_CP (:246-296): complete pivoting, one thread per column
__device__ void g_pivt_cp(double *D, double *F, double *u, double *v, int *PL, int *PU, int k, double *A, int M,
                          int N, VI *red, int *s_ii)
{
    VI best = {0.0, 0};
    int bestrow = 1;
    for (int j = 1 + threadIdx.x; j <= N - k + 1; j += C2H_WG) {
        const double fj = A1(F, k + j - 1);
        double mv = 0.0;
        int mi = 1;
        for (int i = 1; i <= M - k + 1; i++) {
            const double x = fabs(A1(u, k + i - 1) / (A1(D, k + i - 1) - fj));
            if (i == 1 || x > mv) { mv = x; mi = i; }
        }
        const VI c = {fabs(mv * A1(v, k + j - 1)), j};
        const VI nb = vi_better(best, c);
        if (nb.i == c.i) bestrow = mi;
        best = nb;
    }
    const VI win = block_argmax(best, red);
    if (best.i == win.i && win.i > 0) *s_ii = bestrow;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int ii = *s_ii + k - 1, jj = win.i + k - 1;
        g_dswp(D, k, ii); g_dswp(u, k, ii); g_dswp(A, k, ii); g_iswp(PL, k, ii);
        g_dswp(F, k, jj); g_dswp(v, k, jj); g_iswp(PU, k, jj);
    }
    __syncthreads();
}

// rrluCauchy (:7-134) of the M x N matrix A(i) V(j) / (D(i) - F(j)); returns Rk
__device__ int g_rrlu(double *D, double *F, double *A, double *V, double tol, double *U, double *W, int *PL, int *PU,
                      int M, int N, VI *red, int *s_ii, unsigned long long &nsw)
{
    const int tid = threadIdx.x, mn = M < N ? M : N;
    const double rho = 1.02;
    for (int j = 1 + tid; j <= M; j += C2H_WG) { A1(U, j) = A1(A, j); A1(PL, j) = j; }
    for (int j = 1 + tid; j <= N; j += C2H_WG) A1(PU, j) = j;
    __syncthreads();
    int Rk = mn;
    double amax = 0.0;
    for (int k = 1; k <= mn; k++) {
        g_pivt(D, F, U, V, PL, PU, k, A, M, N, red);
        __syncthreads();
        double ukk = A1(U, k) * A1(V, k) / (A1(D, k) - A1(F, k));
        amax = fmax(amax, fabs(ukk));
        if (fabs(ukk) < tol * amax) {
            g_pivt_cp(D, F, U, V, PL, PU, k, A, M, N, red, s_ii);
            ukk = A1(U, k) * A1(V, k) / (A1(D, k) - A1(F, k));
            amax = fmax(amax, fabs(ukk));
            if (fabs(ukk) < tol * amax) { Rk = k - 1; break; }
        }
        const double Dk = A1(D, k), Fk = A1(F, k);
        for (int j = k + 1 + tid; j <= M; j += C2H_WG) A1(U, j) = A1(U, j) * (A1(D, j) - Dk) / (A1(D, j) - Fk);
        for (int j = k + 1 + tid; j <= N; j += C2H_WG) A1(V, j) = A1(V, j) * (A1(F, j) - Fk) / (A1(F, j) - Dk);
        for (int j = 1 + tid; j <= k - 1; j += C2H_WG) A1(W, j) = A1(W, j) * ((Fk - A1(D, j)) / (Dk - A1(D, j)));
        __syncthreads();
        if (tid == 0) {
            double wk = (Dk - Fk) / A1(A, k);
            for (int j = 1; j <= k - 1; j++) wk = wk * (A1(F, j) - Dk) / (A1(D, j) - Dk);
            A1(W, k) = wk;
        }
        __syncthreads();
        if (k % 10 != 0) continue;
        // each swap raises |det| of the pivot block by more than rho; bounded all the same
        for (int sw = 0; sw < 2 * M + 8; sw++) {
            // searchMax2 (:137-163): the largest entry of Mat_N(i, j) = U(k+i) W(j) / (D(k+i) - D(j))
            VI x = {0.0, 0};
            int xrow = 1;
            for (int jc = 1 + tid; jc <= k; jc += C2H_WG) {
                const double dj = A1(D, jc);
                double mv = 0.0;
                int mi = 1;
                for (int i = 1; i <= M - k; i++) {
                    const double t = fabs(A1(U, k + i) / (A1(D, k + i) - dj));
                    if (i == 1 || t > mv) { mv = t; mi = i; }
                }
                const VI c = {mv * fabs(A1(W, jc)), jc};
                const VI nb = vi_better(x, c);
                if (nb.i == c.i) xrow = mi;
                x = nb;
            }
            const VI win = block_argmax(x, red);
            if (x.i == win.i && win.i > 0) *s_ii = xrow;
            __syncthreads();
            if (!(win.v > rho)) break;
            nsw++;
            const int ii = win.i, jj = *s_ii + k;
            const double Dii = A1(D, ii), Djj = A1(D, jj);
            for (int j = k + 1 + tid; j <= N; j += C2H_WG) A1(V, j) = A1(V, j) * ((A1(F, j) - Dii) / (A1(F, j) - Djj));
            for (int j = k + 1 + tid; j <= M; j += C2H_WG)
                if (j != jj) A1(U, j) = A1(U, j) * ((A1(D, j) - Djj) / (A1(D, j) - Dii));
            for (int j = 1 + tid; j <= k; j += C2H_WG)
                if (j != ii) A1(W, j) = A1(W, j) * ((A1(D, j) - Dii) / (A1(D, j) - Djj));
            __syncthreads();
            if (tid == 0) {
                double uj = A1(A, ii) * ((Dii - Djj) / (Dii - A1(F, ii)));
                double wi = (Djj - A1(F, ii)) / A1(A, jj);
                for (int j = 1; j <= k; j++) {
                    if (j == ii) continue;
                    uj = uj * ((Dii - A1(D, j)) / (Dii - A1(F, j)));
                    wi = wi * ((A1(F, j) - Djj) / (A1(D, j) - Djj));
                }
                A1(U, jj) = uj;
                A1(W, ii) = wi;
                g_iswp(PL, ii, jj);
                g_dswp(D, ii, jj);
                g_dswp(A, ii, jj);
            }
            __syncthreads();
        }
    }
    __syncthreads();
    return Rk;
}

// comprcauchy (:616-725): rows ('r': M node rows against N outside columns) or
// columns ('c': M node columns against N outside rows) of node `nodi`; the
// interpolative generator (Cauchylike2, :743-778, rows or columns put back in the
// node's order) goes to the arena
__device__ int g_compr(const C2H &C, bool row, double *D, double *F, double *U, double *V, int M, int N, int nodi,
                       VI *red, int *s_ii, long &top, unsigned long long &nsw)
{
    const int tid = threadIdx.x, mn = M < N ? M : N;
    for (int j = 1 + tid; j <= M; j += C2H_WG) A1(C.Z, j) = 0.0;
    for (int j = 1 + tid; j <= mn; j += C2H_WG) A1(C.W, j) = 0.0;
    __syncthreads();
    const int Rk = row ? g_rrlu(D, F, U, V, C.tol, C.Z, C.W, C.PL, C.PU, M, N, red, s_ii, nsw)
                       : g_rrlu(F, D, V, U, C.tol, C.Z, C.W, C.PL, C.PU, M, N, red, s_ii, nsw);
    const double *X = row ? D : F;
    // invp: PU (free after the RRLU) takes the position of every original index
    for (int i = 1 + tid; i <= M; i += C2H_WG) A1(C.PU, A1(C.PL, i)) = i;
    const int r = Rk < M ? Rk : M;
    const long off = top;
    __syncthreads();
    if (off + (long)M * r > C.arena_cap) {
        if (tid == 0) atomicOr(C.err, 2);
        return -1;
    }
    double *Q = C.arena + off;
    if (row) {
        for (long e = tid; e < (long)M * r; e += C2H_WG) {
            const int i = (int)(e % M) + 1, j = (int)(e / M) + 1, s = A1(C.PU, i);
            double q;
            if (Rk >= M) q = (s == j) ? 1.0 : 0.0;
            else if (s <= r) q = (s == j) ? 1.0 : 0.0;
            else q = A1(C.Z, s) * A1(C.W, j) / (A1(X, s) - A1(X, j));
            Q[e] = q;
        }
        if (tid == 0) { C.usz[2 * (nodi - 1)] = M; C.usz[2 * (nodi - 1) + 1] = r; C.pu[nodi - 1] = off; }
    } else {
        for (long e = tid; e < (long)M * r; e += C2H_WG) {
            const int i = (int)(e % r) + 1, j = (int)(e / r) + 1, s = A1(C.PU, j);
            double q;
            if (Rk >= M) q = (s == i) ? 1.0 : 0.0;
            else if (s <= r) q = (s == i) ? 1.0 : 0.0;
            else q = A1(C.W, i) * A1(C.Z, s) / (A1(X, s) - A1(X, i));
            Q[e] = q;
        }
        if (tid == 0) { C.vsz[2 * (nodi - 1)] = r; C.vsz[2 * (nodi - 1) + 1] = M; C.pv[nodi - 1] = off; }
    }
    top = off + (long)M * r;
    __syncthreads();
    return Rk;
}

// Cauchylike (Cauchylike.f90:1-13) into a (lda), M x N
__device__ void g_cfill(double *a, int lda, const double *D, const double *F, const double *u, const double *v, int M,
                        int N)
{
    for (long e = threadIdx.x; e < (long)M * N; e += C2H_WG) {
        const int i = (int)(e % M) + 1, j = (int)(e / M) + 1;
        a[(i - 1) + (long)(j - 1) * lda] = A1(u, i) * A1(v, j) / (A1(D, i) - A1(F, j));
    }
}

__global__ __launch_bounds__(C2H_WG) void k_cauchy2hss(C2H C)
{
    __shared__ VI red[C2H_WG / 64];
    __shared__ int s_ii;
    const int tid = threadIdx.x, lda = C.n;
    long top = 0;               // arena bump, the same in every thread
    unsigned long long nsw = 0;
    int ns = 1;
    int *lsr = C.lsr, *lsc = C.lsc;
    if (tid == 0) { A1(lsr, 1) = 1; A1(lsc, 1) = 1; }
    __syncthreads();
    // gather the node's row side (D, U at idx(j)) and the outside column side
    // (the stacked column skeletons 1..cs-1, then the original columns after `last`)
    auto gather = [&](bool row, int m0, int mi, int cs, int last, int nt) {
        double *A = row ? C.D : C.F, *Au = row ? C.U : C.V, *B = row ? C.F : C.D, *Bv = row ? C.V : C.U;
        double *Ai = row ? C.Di : C.Fi, *Aui = row ? C.Ui : C.Vi, *Bi = row ? C.Fi : C.Di, *Bvi = row ? C.Vi : C.Ui;
        for (int j = tid; j < mi; j += C2H_WG) { Ai[j] = A1(A, m0 + j); Aui[j] = A1(Au, m0 + j); }
        for (int j = tid; j < nt; j += C2H_WG) {
            const int src = (j < cs - 1) ? j + 1 : last + 1 + (j - (cs - 1));
            Bi[j] = A1(B, src);
            Bvi[j] = A1(Bv, src);
        }
        __syncthreads();
    };
    for (int i = 1; i <= C.ltr; i++) {
        const int c1 = C.ch[2 * (i - 1)], c2 = C.ch[2 * (i - 1) + 1];
        const int l1 = C.rng[2 * (i - 1)], l2 = C.rng[2 * (i - 1) + 1];
        if (c1 == 0) {  // leaf
            const int mi = l2 - l1 + 1, nii = lda - l2;
            g_cfill(C.dd + C.pd[i - 1], mi, C.D + l1 - 1, C.F + l1 - 1, C.U + l1 - 1, C.V + l1 - 1, mi, mi);
            if (tid == 0) C.dsz[i - 1] = mi;
            int cs = (ns == 1) ? 1 : A1(lsc, ns);
            gather(true, l1, mi, cs, l2, nii + cs - 1);
            int Rk = g_compr(C, true, C.Di, C.Fi, C.Ui, C.Vi, mi, nii + cs - 1, i, red, &s_ii, top, nsw);
            if (Rk < 0) return;
            const int r0 = A1(lsr, ns);
            for (int j = tid; j < Rk; j += C2H_WG) { A1(C.D, r0 + j) = C.Di[j]; A1(C.U, r0 + j) = C.Ui[j]; }
            __syncthreads();
            if (tid == 0) A1(lsr, ns + 1) = r0 + Rk;
            __syncthreads();
            const int rs = (ns == 1) ? 1 : A1(lsr, ns);
            gather(false, l1, mi, rs, l2, nii + rs - 1);
            Rk = g_compr(C, false, C.Di, C.Fi, C.Ui, C.Vi, mi, nii + rs - 1, i, red, &s_ii, top, nsw);
            if (Rk < 0) return;
            const int q0 = A1(lsc, ns);
            for (int j = tid; j < Rk; j += C2H_WG) { A1(C.F, q0 + j) = C.Fi[j]; A1(C.V, q0 + j) = C.Vi[j]; }
            __syncthreads();
            if (tid == 0) A1(lsc, ns + 1) = q0 + Rk;
            __syncthreads();
            ns++;
        } else {  // parent: B of both children, then its own compressions
            int mi = A1(lsr, ns - 1) - A1(lsr, ns - 2), nii = A1(lsc, ns) - A1(lsc, ns - 1);
            if (top + (long)mi * nii > C.arena_cap) { if (tid == 0) atomicOr(C.err, 2); return; }
            g_cfill(C.arena + top, mi, C.D + A1(lsr, ns - 2) - 1, C.F + A1(lsc, ns - 1) - 1, C.U + A1(lsr, ns - 2) - 1,
                    C.V + A1(lsc, ns - 1) - 1, mi, nii);
            if (tid == 0) { C.bsz[2 * (c1 - 1)] = mi; C.bsz[2 * (c1 - 1) + 1] = nii; C.pb[c1 - 1] = top; }
            top += (long)mi * nii;
            mi = A1(lsr, ns) - A1(lsr, ns - 1);
            nii = A1(lsc, ns - 1) - A1(lsc, ns - 2);
            if (top + (long)mi * nii > C.arena_cap) { if (tid == 0) atomicOr(C.err, 2); return; }
            g_cfill(C.arena + top, mi, C.D + A1(lsr, ns - 1) - 1, C.F + A1(lsc, ns - 2) - 1, C.U + A1(lsr, ns - 1) - 1,
                    C.V + A1(lsc, ns - 2) - 1, mi, nii);
            if (tid == 0) { C.bsz[2 * (c2 - 1)] = mi; C.bsz[2 * (c2 - 1) + 1] = nii; C.pb[c2 - 1] = top; }
            top += (long)mi * nii;
            __syncthreads();
            if (i == C.ltr) break;
            // rows of the merged skeleton against the stacked columns before and the columns after
            mi = A1(lsr, ns) - A1(lsr, ns - 2);
            nii = lda - l2;
            int cs = (ns == 3) ? 1 : A1(lsc, ns - 2);
            gather(true, A1(lsr, ns - 2), mi, cs, l2, nii + cs - 1);
            int Rk = g_compr(C, true, C.Di, C.Fi, C.Ui, C.Vi, mi, nii + cs - 1, i, red, &s_ii, top, nsw);
            if (Rk < 0) return;
            const int r0 = A1(lsr, ns - 2);
            for (int j = tid; j < Rk; j += C2H_WG) { A1(C.D, r0 + j) = C.Di[j]; A1(C.U, r0 + j) = C.Ui[j]; }
            __syncthreads();
            if (tid == 0) A1(lsr, ns - 1) = r0 + Rk;
            __syncthreads();
            // columns of the merged skeleton
            mi = A1(lsc, ns) - A1(lsc, ns - 2);
            const int rs = (ns == 3) ? 1 : A1(lsr, ns - 2);
            gather(false, A1(lsc, ns - 2), mi, rs, l2, nii + rs - 1);
            Rk = g_compr(C, false, C.Di, C.Fi, C.Ui, C.Vi, mi, nii + rs - 1, i, red, &s_ii, top, nsw);
            if (Rk < 0) return;
            const int q0 = A1(lsc, ns - 2);
            for (int j = tid; j < Rk; j += C2H_WG) { A1(C.F, q0 + j) = C.Fi[j]; A1(C.V, q0 + j) = C.Vi[j]; }
            __syncthreads();
            if (tid == 0) A1(lsc, ns - 1) = q0 + Rk;
            __syncthreads();
            ns--;
        }
    }
    if (tid == 0) *C.nswap = nsw;
}

// the dense general Cauchy-like matrix (fallback for a tree of one leaf)
__global__ __launch_bounds__(WG) void k_cfill_general(int n, const double *d, const double *f, const double *u,
                                                      const double *v, double *a)
{
    for (long e = blockIdx.x * (long)WG + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * WG) {
        const int i = (int)(e % n), j = (int)(e / n);
        a[e] = u[i] * v[j] / (d[i] - f[j]);
    }
}
