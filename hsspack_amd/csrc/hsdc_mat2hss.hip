// Dense matrix -> HSS form by truncated SVDs on gfx950: the construction half of
// HSSPACK's DHSSMM / DHSSMML (DHSSMM.f90:1-125, DHSSMML.f90:2-122; mat2hssvd,
// ConstructHssd.f90:371-696), done depth by depth.  The compressed matrix C holds, for
// the active nodes in left-to-right order, C(a, b) = U_a^T A(a, b) W_b (a raw leaf's
// basis is the identity).  For every node of one depth, all in one launch:
//
//   k_m2h_gather   its block row (its rows against the other active columns) transposed
//                  and its block column, as tall matrices (context x own)
//   k_m2h_svd      one workgroup per (node, side): Householder QR of the tall matrix and
//                  one-sided Jacobi on R's columns with the rotations accumulated (the
//                  right singular vectors); for a matrix wider than tall the columns of
//                  its transpose are orthogonalised directly; the singular values
//                  descending, the rank = their count above tol (mat2hssvd's
//                  count(svs > tol)), the basis vectors in that order
//   k_m2h_rows,    C' = blockdiag(U)^T C blockdiag(W) over the new active list, one output
//   k_m2h_cols     element per thread
//   k_m2h_copy     generators: U (m x r), V = W^T (r x m), each sibling pair's B from C'
//
// With randomized range finding (mat2hssvdr), k_m2h_sketch forms Omega T for each side
// with a long context before k_m2h_svd, which then decomposes the sketch.
//
// The generators land in the layout the HSS multiply reads (U, V^T, B, D per node, the
// post-ordered tree of npart), so DHSSMM and DHSSMML run the same batched multiply as
// the eigenvector matrices.
#include <math.h>
#include "hsdc_internal.h"

#define MWG 256

namespace {

__device__ __forceinline__ double m2h_wsum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ int m2h_rr(int slot, int r, int P)
{
    if (slot == 0) return 0;
    return 1 + (slot - 1 + (P - 1) - (r % (P - 1))) % (P - 1);
}

}  // namespace

// One node's compression problem of one side: T (w x m, column-major, ld w) gathered
// from C; the result basis (m x m, ld m, columns in descending singular value order),
// its singular values and rank.  side 0: block row (basis U), 1: block column (W).
struct M2hJob {
    int rs, rn, cs, cn;  // the node's rows and columns in C
    int w, m;            // context length and own length of this side
    long toff;           // T, then R (m x m), V (m x m) at toff + w m, + m m
    long boff;           // basis (m x m) and singular values (m)
};

// ---------------------------------------------------------------------------
// Gather the tall matrices: side 0: T(o, t) = C(rs + t, col_o), col_o the o-th column
// outside [cs, cs + cn); side 1: T(o, t) = C(row_o, cs + t), row_o outside [rs, rs + rn).
// Grid: (chunks, jobs); one thread per element.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MWG) void k_m2h_gather(const M2hJob *jobs, const double *C, long ldc, double *S)
{
    const M2hJob J = jobs[blockIdx.y];
    const int side = blockIdx.y & 1;
    const long tot = (long)J.w * J.m;
    double *T = S + J.toff;
    for (long e = (long)blockIdx.x * MWG + threadIdx.x; e < tot; e += (long)gridDim.x * MWG) {
        const int o = (int)(e % J.w), t = (int)(e / J.w);
        if (side == 0) {
            const int col = o < J.cs ? o : o + J.cn;
            T[e] = C[(J.rs + t) + (long)col * ldc];
        } else {
            const int row = o < J.rs ? o : o + J.rn;
            T[e] = C[row + (long)(J.cs + t) * ldc];
        }
    }
}

// ---------------------------------------------------------------------------
// One workgroup per job.  w >= m: Householder QR of T in place (v = (1, x(1:)), tau =
// (beta - alpha) / beta, beta = -sign(alpha) |x|), R's upper triangle copied to the R
// area, then one-sided Jacobi on its m columns with the rotations accumulated in V
// (identity at the start); the basis = V's columns by descending column norm of the
// rotated R.  w < m: the m x w transpose of T is copied to the R area and its w columns
// orthogonalised; the basis = the normalised columns by descending norm (the rest zero).
// A pair rotates when |c_a . c_b| > L eps |c_a| |c_b|, L the column length; 60 sweeps at
// most (else *err is set).  rank[job] = the count of singular values above tol.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MWG) void k_m2h_svd(const M2hJob *jobs, double *S, double *Bs, double tol, int *rank,
                                                 int *err)
{
    __shared__ double red[MWG / 64];
    __shared__ int s_rot, perm_dummy;
    const M2hJob J = jobs[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, NW = MWG / 64;
    const int w = J.w, m = J.m;
    double *T = S + J.toff, *R = T + (long)w * m, *V = R + (long)m * m;
    double *B = Bs + J.boff, *sv = B + (long)m * m;
    const bool qr = w >= m;
    const int L = m, P = qr ? m : w;  // columns of length L to orthogonalise
    (void)perm_dummy;
    if (qr) {
        for (int k = 0; k < m; k++) {
            double *x = T + k + (long)k * w;
            const int len = w - k;
            double part = 0.0;
            for (int i = 1 + tid; i < len; i += MWG) part += x[i] * x[i];
            part = m2h_wsum(part);
            if (lane == 0) red[wv] = part;
            __syncthreads();
            double xn = 0.0;
            for (int q = 0; q < NW; q++) xn += red[q];
            xn = sqrt(xn);
            const double alpha = x[0];
            __syncthreads();
            if (xn == 0.0) continue;  // uniform: H = I
            const double beta = -copysign(hypot(alpha, xn), alpha);
            const double tau = (beta - alpha) / beta, scal = 1.0 / (alpha - beta);
            for (int i = 1 + tid; i < len; i += MWG) x[i] *= scal;
            __syncthreads();
            for (int j = k + 1 + wv; j < m; j += NW) {
                double *c = T + k + (long)j * w;
                double d = 0.0;
                for (int i = 1 + lane; i < len; i += 64) d += x[i] * c[i];
                d = m2h_wsum(d);
                d = (c[0] + d) * tau;
                if (lane == 0) c[0] -= d;
                for (int i = 1 + lane; i < len; i += 64) c[i] -= d * x[i];
            }
            __syncthreads();
            if (tid == 0) x[0] = beta;
            __syncthreads();
        }
        for (long e = tid; e < (long)m * m; e += MWG) {
            const int i = (int)(e % m), j = (int)(e / m);
            R[e] = (i <= j) ? T[i + (long)j * w] : 0.0;
            V[e] = (i == j) ? 1.0 : 0.0;
        }
    } else {
        for (long e = tid; e < (long)m * w; e += MWG) {
            const int i = (int)(e % m), j = (int)(e / m);  // R area as m x w: column j = row j of T
            R[i + (long)j * m] = T[j + (long)i * w];
        }
    }
    __syncthreads();
    const int PP = P + (P & 1);
    bool conv = P < 2;
    for (int sweep = 0; sweep < 60 && !conv; sweep++) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        for (int r = 0; r < PP - 1; r++) {
            for (int pr = wv; pr < PP / 2; pr += NW) {
                int a = m2h_rr(pr, r, PP), b = m2h_rr(PP - 1 - pr, r, PP);
                if (a > b) { const int t = a; a = b; b = t; }
                if (b >= P) continue;  // uniform over the wavefront
                double *ca = R + (long)a * m, *cb = R + (long)b * m;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = lane; k < L; k += 64) {
                    const double x = ca[k], y = cb[k];
                    al += x * x; be += y * y; ga += x * y;
                }
                al = m2h_wsum(al); be = m2h_wsum(be); ga = m2h_wsum(ga);
                if (ga == 0.0 || !(fabs(ga) > L * HSDC_EPS * sqrt(al * be))) continue;
                const double ze = (be - al) / (2.0 * ga);
                const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                const double cc = 1.0 / sqrt(1.0 + t * t), ss = cc * t;
                for (int k = lane; k < L; k += 64) {
                    const double x = ca[k], y = cb[k];
                    ca[k] = cc * x - ss * y;
                    cb[k] = ss * x + cc * y;
                }
                if (qr) {
                    double *va = V + (long)a * m, *vb = V + (long)b * m;
                    for (int k = lane; k < m; k += 64) {
                        const double x = va[k], y = vb[k];
                        va[k] = cc * x - ss * y;
                        vb[k] = ss * x + cc * y;
                    }
                }
                if (lane == 0) atomicAdd(&s_rot, 1);
            }
            __syncthreads();
        }
        conv = (s_rot == 0);
        __syncthreads();
    }
    if (!conv && tid == 0) atomicOr(err, 1);
    // singular values (column norms), descending with ties by index; basis columns
    for (int j = wv; j < P; j += NW) {
        double q = 0.0;
        for (int k = lane; k < L; k += 64) q += R[k + (long)j * m] * R[k + (long)j * m];
        q = sqrt(m2h_wsum(q));
        if (lane == 0) sv[j] = q;
    }
    __syncthreads();
    int *ord = (int *)(sv + m);  // P ints after the singular values
    if (tid == 0) {
        for (int j = 0; j < P; j++) ord[j] = j;
        for (int i = 0; i + 1 < P; i++) {
            int k = i;
            for (int j = i + 1; j < P; j++) if (sv[ord[j]] > sv[ord[k]]) k = j;
            const int t = ord[i]; ord[i] = ord[k]; ord[k] = t;
        }
        int cnt = 0;
        for (int j = 0; j < P; j++) cnt += sv[ord[j]] > tol;
        rank[blockIdx.x] = cnt;
    }
    __syncthreads();
    for (long e = tid; e < (long)m * m; e += MWG) {
        const int i = (int)(e % m), j = (int)(e / m);
        double v = 0.0;
        if (j < P) {
            const int s0 = ord[j];
            if (qr) v = V[i + (long)s0 * m];
            else {
                const double nrm = sv[s0];
                v = nrm > 0.0 ? R[i + (long)s0 * m] / nrm : 0.0;
            }
        }
        B[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Randomized range finding (mat2hssvdr, ConstructHssd.f90:8-368): a side whose context
// is longer than 50 and than m + 10 takes its basis from the sketch Omega T (pp x m,
// pp = m + 10) instead of T.  The reference draws one flat array of uniform [0, 1)
// numbers and reads it as a pp x w matrix (leading dimension pp) for a block column and
// as a w x pp matrix (leading dimension w) for a block row (:104, :129, ...); entry i of
// the flat array here is m2h_omega(seed, i), the splitmix64 stream of the seed, so the
// oracle draws the same numbers.  sk[job] is the sketch job (w = pp, toff at its
// scratch) or the job itself.  Grid: (chunks, jobs); one thread per element of S.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double m2h_omega(unsigned long long seed, unsigned long long i)
{
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

__global__ __launch_bounds__(MWG) void k_m2h_sketch(const M2hJob *jobs, const M2hJob *sk, const double *S0, double *S,
                                                    unsigned long long seed)
{
    const M2hJob J = jobs[blockIdx.y], K = sk[blockIdx.y];
    if (K.w == J.w) return;  // no sketch for this side
    const int side = blockIdx.y & 1, pp = K.w, w = J.w;
    const double *T = S0 + J.toff;
    double *Y = S + K.toff;
    const long tot = (long)pp * J.m;
    for (long e = (long)blockIdx.x * MWG + threadIdx.x; e < tot; e += (long)gridDim.x * MWG) {
        const int r = (int)(e % pp), t = (int)(e / pp);
        const double *tc = T + (long)t * w;
        double acc = 0.0;
        for (int o = 0; o < w; o++) {
            const unsigned long long idx = side ? (unsigned long long)r + (unsigned long long)o * pp
                                                : (unsigned long long)o + (unsigned long long)r * w;
            acc += m2h_omega(seed, idx) * tc[o];
        }
        Y[e] = acc;
    }
}

// New rows of C: for each new row block a, U_q^T C(rows of q, :) when it comes from a
// job pair q (src >= 0), the old rows otherwise.  out is nnr x nc (ld nnr).
struct M2hBlk {
    int start, count;  // new block in C' (rows or columns)
    int src;           // >= 0: job pair (2 src = row side, 2 src + 1 = column side); < 0: -1 - old block
    int ostart, ocount;  // the old rows / columns it is formed from
};

__global__ __launch_bounds__(MWG) void k_m2h_rows(const M2hBlk *blk, int nb, const M2hJob *jobs, const double *Bs,
                                                  const double *C, long ldc, int nc, double *out, long ldo)
{
    const long tot = ldo * (long)nc;
    for (long e = (long)blockIdx.x * MWG + threadIdx.x; e < tot; e += (long)gridDim.x * MWG) {
        const int r = (int)(e % ldo), c = (int)(e / ldo);
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (blk[mid].start <= r) lo = mid; else hi = mid - 1;
        }
        const M2hBlk b = blk[lo];
        const int i = r - b.start;
        double v;
        if (b.src >= 0) {
            const M2hJob J = jobs[2 * b.src];
            const double *U = Bs + J.boff;  // m x m, column i
            v = 0.0;
            for (int t = 0; t < J.m; t++) v += U[t + (long)i * J.m] * C[(b.ostart + t) + (long)c * ldc];
        } else {
            v = C[(b.ostart + i) + (long)c * ldc];
        }
        out[r + (long)c * ldo] = v;
    }
}

// New columns: C'(:, new block b) = Tr(:, old columns of q) W_q, or a copy.
__global__ __launch_bounds__(MWG) void k_m2h_cols(const M2hBlk *blk, int nb, const M2hJob *jobs, const double *Bs,
                                                  const double *Tr, long ldt, int nnc, double *out)
{
    const long tot = ldt * (long)nnc;
    for (long e = (long)blockIdx.x * MWG + threadIdx.x; e < tot; e += (long)gridDim.x * MWG) {
        const int r = (int)(e % ldt), c = (int)(e / ldt);
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (blk[mid].start <= c) lo = mid; else hi = mid - 1;
        }
        const M2hBlk b = blk[lo];
        const int j = c - b.start;
        double v;
        if (b.src >= 0) {
            const M2hJob J = jobs[2 * b.src + 1];
            const double *W = Bs + J.boff;  // m x m, column j
            v = 0.0;
            for (int t = 0; t < J.m; t++) v += Tr[r + (long)(b.ostart + t) * ldt] * W[t + (long)j * J.m];
        } else {
            v = Tr[r + (long)(b.ostart + j) * ldt];
        }
        out[r + (long)c * ldt] = v;
    }
}

// Copies of one level's generators into the arena: kind 0 U (m x r from the row basis),
// 1 V = W^T (r x m from the column basis), 2 B (a block of C'), 3 D (a block of A).
struct M2hCopy {
    int kind, rows, cols, job;  // job: the basis (kinds 0, 1)
    int r0, c0;                 // the block's origin in its source (kinds 2, 3)
    long dst;                   // arena / leaf-block offset
};

__global__ __launch_bounds__(MWG) void k_m2h_copy(const M2hCopy *cp, const M2hJob *jobs, const double *Bs,
                                                  const double *Cn, long ldcn, const double *A, long lda,
                                                  double *arena, double *dd)
{
    const M2hCopy P = cp[blockIdx.y];
    const long tot = (long)P.rows * P.cols;
    double *dst = (P.kind == 3 ? dd : arena) + P.dst;
    for (long e = (long)blockIdx.x * MWG + threadIdx.x; e < tot; e += (long)gridDim.x * MWG) {
        const int i = (int)(e % P.rows), j = (int)(e / P.rows);
        double v;
        if (P.kind == 0) {
            const M2hJob J = jobs[P.job];
            v = Bs[J.boff + i + (long)j * J.m];
        } else if (P.kind == 1) {
            const M2hJob J = jobs[P.job];
            v = Bs[J.boff + j + (long)i * J.m];
        } else if (P.kind == 2) {
            v = Cn[(P.r0 + i) + (long)(P.c0 + j) * ldcn];
        } else {
            v = A[(P.r0 + i) + (long)(P.c0 + j) * lda];
        }
        dst[e] = v;
    }
}
