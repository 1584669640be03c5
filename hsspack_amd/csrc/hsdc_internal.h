// Internal types shared by the hsdc HIP kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HSDC_EPS 0x1p-53 /* DLAMCH('Epsilon') */

// Lanes that share one secular root of a K-root merge: the largest power of two
// <= 64 that leaves every lane at least 4 poles.  Summation order inside the
// solver depends on it, so host tables and the kernel use this one rule.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline int sec_team(int k)
{
    int t = 1;
    while (t < 64 && 8 * t <= k) t *= 2;
    return t;
}
// the value lane l ^ OFF holds (OFF a power of two below 64): cross-lane moves
// without the LDS round trip where the pattern allows -- DPP quad permutations for 1
// and 2, a 16-lane row rotation by 8 for 8, ds_swizzle's xor mode (32-lane groups) for
// 4 and 16, ds_bpermute for 32.  Same values as __shfl_xor, so sums built from it
// keep their order and bits.
template <int OFF>
__device__ __forceinline__ int xlane32(int x)
{
    if constexpr (OFF == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    else if constexpr (OFF == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (OFF == 8) return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror:8
    else if constexpr (OFF == 4) return __builtin_amdgcn_ds_swizzle(x, 0x101F);   // and 0x1f, xor 4
    else if constexpr (OFF == 16) return __builtin_amdgcn_ds_swizzle(x, 0x401F);  // and 0x1f, xor 16
    else return __builtin_amdgcn_ds_bpermute((int)((threadIdx.x ^ OFF) << 2), x);
}

template <int OFF>
__device__ __forceinline__ double xlane(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = xlane32<OFF>((int)(unsigned)(u & 0xffffffffull));
    const int hi = xlane32<OFF>((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// butterfly over the T (power of two <= 64) lanes of a team, lanes aligned to T:
// lane l adds lane l ^ off for off = T/2 .. 1, every lane ends with the same value
__device__ __forceinline__ double team_bfly(double v, int T)
{
    if (T > 32) v = v + xlane<32>(v);
    if (T > 16) v = v + xlane<16>(v);
    if (T > 8) v = v + xlane<8>(v);
    if (T > 4) v = v + xlane<4>(v);
    if (T > 2) v = v + xlane<2>(v);
    if (T > 1) v = v + xlane<1>(v);
    return v;
}

#define SEC_WG 256   // threads per block of the secular kernel
#define SEC_LDS 2048 // merges with K <= SEC_LDS stage poles and z in LDS

// One eigenproblem of the batch; its eigenvector block Q is n x n, column-major.
struct Prob {
    double *q;  // device base of Q
    long ld;    // leading dimension (n)
    int off;    // first global index of the problem
    int n;
    int par;    // ping-pong buffer holding its final state (= DC levels mod 2)
};

// One rank-one merge of one DC level (mdlaed0.f90:298-337): block [s, s+n) with
// cut n1 and coupling rho = E(s+n1-1) (signed, before MDLAED2 rescales it).
struct Merge {
    int s, n, n1, prob;
    int lvl;  // DC level of the merge (0 = first merge level)
    double rho;
};

// A leaf of the DC tree (mdlaed0.f90:253-287).
struct Leaf {
    int s, m, prob;
};

// per-merge scalars written by the deflation kernel
struct MergeOut {
    int k;        // non-deflated count
    int ctot[4];  // column types 1..4 (mdlaed2.f90:393-400)
    int hss;      // 1 = HSS branch (k >= hss_min_k)
    int nrot;     // Givens rotations recorded
    int flags;    // bit0: rho*max|z| <= tol (K = 0 path)
    double rho;   // |2 rho|
    double tol;   // deflation tolerance
};

// workspace arrays indexed by global index (size ntot)
struct Work {
    double *Dv[2];   // eigenvalue of each virtual column (ping-pong)
    int *vm[2];      // physical slot of each virtual column (ping-pong)
    int *indxq;      // per block: sorted position -> local virtual column (1-based)
    int *supp_lo, *supp_hi;  // row support [lo, hi) of each physical slot
    double *z, *dlamda, *w, *ddef, *zs, *ds;
    int *indxq2, *indx, *indxc, *indxp, *coltyp;
    int *rot_pj, *rot_nj;
    int *nsl, *evq;     // deflation: survivor list, rotation event positions
    double *evd;        // deflation: value of each rotated-out column
    double *rot_c, *rot_s;
    double *lam, *tau, *wnew, *alpha, *difl, *difr;
    int *org, *nit;
    unsigned long long *maux;  // per merge of a launch: max|z|, max|D| (bit patterns), unsorted flag, pad
    double *st_t, *st_tc, *st_lo, *st_hi;  // per root: starting iterate, classical start, bracket
    int *st_ii;                // per root: origin pole | learned-start flag << 30
    // device-scheduled levels: nonzero gate = 1 + the level whose HSS merges wait for
    // the host; the kernels of the levels above it return at once while it is set
    int *gate;
    unsigned long long *lstat;  // roots, deflated, HSS merges, K of the last merge
};

// a kernel of the merges mg stands still: a level below waits for the host
__device__ __forceinline__ bool gate_closed(const Work &W, const Merge *mg)
{
    const int g = *W.gate;
    return g != 0 && mg[0].lvl >= g;
}

// Tables of one level derived from its deflation results (k_lvltab): first chunk
// (256 roots), secular block, column, GEMM tile and copy chunk of every merge,
// offsets of its S and Ct blocks (dense merges only).
#define SC_CH 2048    // elements per work item of the dense eigenvector copy
#define FIN_CH 1024   // rows per block of the final reordering copy
struct LvlTab {
    int *chunkK, *blockK, *colK, *tileK, *scK;  // nm + 1 entries each
    long *soff, *coff;                          // nm entries each
};

// End-of-solve figures, written by k_readback into mapped host memory: the gate,
// INFO, rows read by the final copy, level statistics and secular counters.
struct SolveRec {
    int gate, info;
    unsigned long long rows, roots, deflated, hss, ktop, iters, used, wave_iters, crit, evals;
};

// control block of the secular kernel: learned-start weights, statistics and
// optional harvesting of training samples
#define SEC_SLOTS 64    // copies of each statistics counter (summed at the end)
#define SEC_LEVELS 128  // DC levels with their own largest-iteration counters
struct SecCtl {
    const float *w;              // MLP weights or nullptr
    unsigned long long *iters;   // total iterations [SEC_SLOTS]
    unsigned long long *used;    // roots started from the learned iterate [SEC_SLOTS]
    unsigned long long *wave_iters;  // sum over wavefronts of their largest iteration count [SEC_SLOTS]
    unsigned long long *evals;   // pole evaluations: sum over roots of K x iterations [SEC_SLOTS]
    float *hf, *ht;              // harvest: features (cap x 34) and targets (cap)
    float *feat;                 // harvest: features of every root (34 per global index)
    unsigned long long *count;
    unsigned long long cap;
    int kmin;                    // merges with k < kmin use (and harvest) the classical start only
    float gate;                  // the proposal is taken when |log-correction| > gate
};
