// Device descriptors of the HSS construction and multiply.
#pragma once
#include <stdint.h>
#include <vector>
#include <hip/hip_runtime.h>

#define HSS_MAXM 512  // largest row/column block a node may compress (LDS tile)

// One HSS merge (the K x K Cauchy-like eigenvector matrix of one rank-one merge).
// Arrays are 1-based through A1() in the kernels, like the Fortran they restate.
struct HssMerge {
    const double *FF, *D0, *U0, *V0, *DIFL, *DIFR;  // poles, roots, z, alpha, differences (K)
    double *D, *U, *V;      // compacted column / row generators (K)
    int *PL, *PU;           // their original indices (K)
    double *T;              // compaction scratch (K doubles + K ints)
    int *lr, *lc;           // node row / column ranges in the compacted arrays (2*ltr)
    int *lda, *lca;         // current compacted row / column counts
    int *usz, *vsz, *bsz, *dsz;   // generator shapes (2 per node; dsz 1 per node)
    long *pu, *pv, *pb, *pd;      // generator offsets (arena / dd)
    double *arena;          // U, V, B generators
    unsigned long long *arena_top;
    long arena_cap;
    double *dd;             // leaf D blocks
    double *scratch;        // private copies of the compressed block rows / columns
    unsigned long long *scratch_top;
    long scratch_cap;
    unsigned long long *flops;  // the RRLU flop count of the reference's nflops model
    int *err;               // bit0 block too large, bit1 arena full, bit2 scratch full
    int K, ltr, lvl;
    int sq;                 // 1: u_i v_j / (f_i^2 - d_j^2), the SVD's vector matrices
};

// The general Cauchy-like matrix A(i,j) = u(i) v(j) / (d(i) - f(j)) of order n (direct
// differences) and its HSS construction by cauchy2hss's post-order traversal with a
// stack of compressed skeletons (CauchyHss_VP2.f90:387-613), one workgroup.  Arrays
// of n: the working generators (copies of the inputs), the node scratch; per node
// (1-based ids, 2 entries each where paired): children, original index range, and
// the generator directory in the layout of HssMerge.
struct C2H {
    int n, ltr, lvl;
    double tol;
    const int *ch;          // 2 per node: children (0, 0 for a leaf)
    const int *rng;         // 2 per node: first and last original index (1-based)
    double *D, *F, *U, *V;  // working generators (n)
    double *Di, *Fi, *Ui, *Vi, *Z, *W;  // node scratch (n)
    int *PL, *PU;           // permutations (n)
    int *lsr, *lsc;         // skeleton stacks (lvl + 3)
    double *arena;          // U, V, B generators
    long arena_cap;
    double *dd;             // leaf D blocks, at pd
    const long *pd;
    int *usz, *vsz, *bsz, *dsz;
    long *pu, *pv, *pb;
    int *err;                    // bit1: arena full
    unsigned long long *nswap;   // swaps of the RRLU
};

// One node of one merge's post-ordered HSS tree (aux_hss.f90:737-800).
struct HssNode {
    int mrg;        // index into the HssMerge table
    int node;       // 1-based post-order id
    int c1, c2;     // children (0 for a leaf)
    int sib;        // sibling
    int leaf_r0;    // leaf: first original index - 1
    int leaf_m;     // leaf: size
};

// compaction order of one merge at one level: the nodes of the level, then the
// leaves of shallower levels (CauchyHssEig.f90:2140-2168, 2195-2227)
struct HssCompact {
    int mrg;
    int start, count;  // slice of the order table
    int nlevel;        // the first nlevel entries are nodes of the level
};

// One GEMM of the batched HSS multiply: C(nrow x n) = A(nrow x k) B(k x n)
// [+ A2(nrow x k2) B2(k2 x n)] + beta C
struct BG {
    const double *A;
    const double *B;
    double *C;
    long lda, ldb, ldc;
    int nrow, n, k;
    int beta;
    int btrans;  // 1: B(t, j) = B[t * ldb + j]
    const double *A2, *B2;  // optional second product (k2 > 0), B2 column-major
    long lda2, ldb2;
    int k2;
    // optional column map of C: output column j goes to C + (cmap[j] - cmoff) * ldc
    const int *cmap = nullptr;
    long cmoff = 0;
};

// One matrix of an HSS batch: generators on the device (K each), the product
// Y = X * H with X nrow x K.  Column j of Y goes to Y + (ymap[j] - ymoff) * ldy when
// ymap is set.
struct HssSpec {
    int K;
    const double *FF, *D0, *U0, *V0, *DIFL, *DIFR;  // poles, roots, u, v, differences
    const double *X;
    double *Y;
    long ldx, ldy;
    int nrow;
    const int *ymap = nullptr;
    long ymoff = 0;
    int sq = 0;        // squared variables (HssMerge::sq)
};

// The HSS driver of the tridiagonal solver run for another solver's batch on stream s
// (defined in hsdc_api.hip): builds every matrix of sp and sets Y := X * H.  rank
// (may be null) gets per matrix its largest generator rank (0 when applied densely);
// *dense (may be null) the number of matrices applied densely.
struct hsdc_ctx;
int hsdc_hss_batch(hsdc_ctx *c, hipStream_t s, const std::vector<HssSpec> &sp, int ni, double tol,
                   std::vector<int> *rank, int *dense);
