/*
 * hsdc -- MI355X (gfx950) divide-and-conquer symmetric tridiagonal eigensolver:
 * the merge path of HSSPACK's MDSTEDC (SRC/mdstedc.f90 -> mdlaed0 -> mdlaed1 ->
 * mdlaed2 -> mdlaed3 -> DHSSEVC -> fastHssmmL) as HIP kernels behind a C ABI.
 *
 * Conventions
 *   - Plain pointers and sizes; int32 sizes; FP64 values; column-major matrices.
 *   - Index arrays carry 1-based values, as the Fortran routines they replace.
 *   - Return value: 0 on success, -i when argument i is illegal (the LAPACK
 *     convention of the routine replaced), > 0 for a numerical failure encoded as
 *     the replaced routine encodes it, and HSDC_ERR_* (below) for runtime errors.
 *   - Functions named *_dev take device pointers; all others take host pointers
 *     and stage through device memory owned by the context.
 *   - One context per host thread; calls on one context are serialised.
 */
#ifndef HSDC_H
#define HSDC_H

#ifdef __cplusplus
extern "C" {
#endif

/* runtime failures, disjoint from every INFO the replaced routines return */
#define HSDC_ERR_HIP       (-1000) /* a HIP call failed */
#define HSDC_ERR_NOMEM     (-1001) /* device allocation failed */
#define HSDC_ERR_NODEVICE  (-1002) /* no gfx950 device */
#define HSDC_ERR_INTERNAL  (-1003)

typedef struct hsdc_ctx hsdc_ctx;

/* Tunables.  Defaults are HSSPACK's compiled-in constants. */
typedef struct {
    int smlsiz;      /* DC leaf bound, ILAENV(9,'DSTEDC') = 25 (mdstedc.f90:211) */
    int hss_min_k;   /* K >= hss_min_k takes the HSS branch; BSMLZ = 1500 (mdlaed2.f90:175) */
    int hss_leaf;    /* HSS leaf block, Ni = 200 (mdlaed3.f90:253) */
    double hss_tol;  /* HSS compression tolerance, TOL = 1e-17 (CauchyHssEig.f90:2523) */
    int maxit;       /* secular iteration cap */
    int batched;     /* 1: one launch per DC level over all merges; 0: one merge at a time */
    int learned;     /* 1: MLP proposes the secular starting iterate */
    int profile;     /* 1: time the final reordering kernel with HIP events */
    int learned_min_k; /* merges with K below this keep the classical start (default 256) */
    double learned_gate; /* the proposal is taken when its |log-correction| exceeds this */
    int level_sync;  /* batched only.  0: the whole solve is queued at once, each level sizing its
                        launches on the device (one host round trip per solve, plus one per level
                        holding HSS merges); 1: the host reads every level's K before its launches */
    int graph;       /* 1 (default): a solve whose levels are all queued at once is captured as one
                        HIP graph per shape (plan, buffers, parameters) and replayed; 0, or profile = 1
                        (events around the final kernel): direct launches */
} hsdc_params;

/* Run statistics of the last call on a context. */
typedef struct {
    int levels, merges, hss_merges;
    long long roots, secular_iters;
    long long deflated;
    int k_top;          /* K of the last (top) merge */
    int rank_max;       /* largest HSS generator rank */
    double ms_total;    /* device time of the solve (events), ms */
    double ms_final;    /* final reordering kernel (events, profile = 1), ms */
    long long final_bytes;  /* its algorithmic HBM bytes: Z written (the supported rows only when
                               Z was zero-filled beside the levels) + Q rows read */
    long long learned_roots; /* roots started from the learned proposal */
    long long secular_wave_iters; /* sum over 64-root wavefronts of their largest count */
    long long secular_crit_iters; /* sum over DC levels of the largest count of the level */
    long long hss_mm_flops;  /* flops of the HSS multiplies (2 L (|H| + |DD|) per merge) */
    double ms_hss_mm;        /* their device time (events, profile = 1), ms */
    double ms_zfill;         /* zero fill of Z, run beside the DC levels (events, profile = 1), ms */
    long long zfill_bytes;   /* its bytes: 8 n^2 (device-resident single-output solves; else 0) */
    long long secular_evals; /* pole evaluations of the secular solves: sum over roots of K x iterations
                                (16 bytes each: one pole, one z) */
    double ms_secular;       /* device time of the secular launches (profile = 1), ms */
    double ms_loewner;       /* of the Loewner z and alpha launches (profile = 1), ms */
    long long loewner_pairs; /* (i, j) pairs those launches visit: K^2 per merge for the Loewner
                                products, K^2 more per HSS merge for alpha (16 bytes each) */
    long long hss_nflops;    /* the HSS construction's flops in HSSPACK's model: nflops of
                                rrluCauchyEigrow / rrluCauchyEigcol and their pivot searches
                                (CauchyHssEig.f90:105-205, 321-460, 680-703) */
    double ms_hss_build;     /* device time of the HSS construction (profile = 1), ms */
} hsdc_stats;

void hsdc_default_params(hsdc_params *p);
int  hsdc_ctx_create(hsdc_ctx **ctx);
int  hsdc_ctx_destroy(hsdc_ctx *ctx);
int  hsdc_set_params(hsdc_ctx *ctx, const hsdc_params *p);
int  hsdc_get_stats(hsdc_ctx *ctx, hsdc_stats *s);
/* HIP stream used by the context (hipStream_t); 0 = the null stream. */
int  hsdc_set_stream(hsdc_ctx *ctx, void *stream);

/* Merges of the last DC solve on the context (hsdc_stedc, hsdc_stedc_dev, hsdc_laed0,
 * hsdc_stedc_batched_dev), level by level and left to right within a level -- for one
 * unsplit problem the order of MDLAED0's merge loop (mdlaed0.f90:294-341), one MDLAED1
 * call (mdlaed1.f90:109) each: block start s (0-based global index), order n, K after
 * deflation (MDLAED2, mdlaed2.f90:149-409), the deflation rotations applied (DROT,
 * mdlaed2.f90:352), the branch (hss 1: MDLAED3 with DHSSEVC, mdlaed1.f90:210; 0: DLAED3,
 * :199) and the largest HSS generator rank (0 on the dense branch).  Any pointer may
 * be null; entries beyond cap are not written.  Returns the merge count (0 when a
 * single-merge call has reused the solve's tables since). */
int hsdc_trace(hsdc_ctx *ctx, int *s, int *n, int *k, int *nrot, int *hss, int *hss_rank, int cap);
/* INDXQ after the top merge of the last solve of one unsplit problem of order n
 * (1-based; the permutation MDLAED0 applies at mdlaed0.f90:355-362).  -3 when the
 * last solve had another order or the tables were reused. */
int hsdc_trace_indxq(hsdc_ctx *ctx, int *indxq, int n);

/* MDSTEDC (mdstedc.f90:139).  compz 'N', 'I' or 'V'; d(n), e(n-1), z(ldz,n) host
 * arrays; work/iwork follow the LAPACK workspace contract (lwork = -1 or
 * liwork = -1 is a size query that returns work[0], iwork[0]). */
int hsdc_stedc(hsdc_ctx *ctx, char compz, int n, double *d, double *e, double *z, int ldz,
               double *work, int lwork, int *iwork, int liwork);

/* MDLAED0 (mdlaed0.f90:121): DC of one unreduced block.  icompq 0: eigenvalues only;
 * 1: q(qsiz, n) := q * V; 2: q := V (the eigenvectors).  d(n), e(n-1), q(ldq, n) host
 * arrays; the routine's QSTORE/WORK/IWORK are not needed.  INFO as DLAED0 (-1, -2, -3, -7). */
int hsdc_laed0(hsdc_ctx *ctx, int icompq, int qsiz, int n, double *d, double *e, double *q, int ldq);

/* Device-resident MDSTEDC with COMPZ = 'I': d_d(n), d_e(n-1), d_z(ldz,n) device
 * arrays; eigenvalues overwrite d_d, e is destroyed. */
int hsdc_stedc_dev(hsdc_ctx *ctx, int n, double *d_d, double *d_e, double *d_z, int ldz);

/* Many independent tridiagonals in one solve (one launch per DC level over all
 * merges of all problems).  Problem p has order ns[p]; its diagonal is
 * d_d[offs[p] .. offs[p]+ns[p]), its off-diagonal d_e[offs[p] .. offs[p]+ns[p]-1),
 * its eigenvectors go to d_z + zoffs[p] with leading dimension ns[p].  Device
 * arrays except ns/offs/zoffs (host). */
int hsdc_stedc_batched_dev(hsdc_ctx *ctx, int nprob, const int *ns, const long *offs,
                           double *d_d, double *d_e, double *d_z, const long *zoffs);

/* Autotuner (no counterpart in HSSPACK, whose leaf sizes are compile-time constants:
 * ILAENV(9) = 25, Ni = 200 at mdlaed3.f90:253).  Times one batched solve of the given
 * problems (arguments as hsdc_stedc_batched_dev) for each HSS leaf size in {64, 128, 256}
 * and DC leaf bound smlsiz in {25, 40}, and keeps the fastest pair in the context's
 * parameters.  d_d/d_e are restored; d_z is overwritten.  times[6] (may be NULL)
 * receives each candidate's device time in ms, smlsiz-major. */
int hsdc_autotune(hsdc_ctx *ctx, int nprob, const int *ns, const long *offs, double *d_d, double *d_e,
                  double *d_z, const long *zoffs, int *best_leaf, int *best_smlsiz, double *times);

/* ---- single merge kernels (host pointers) ---- */

/* MDLAED2's deflation (mdlaed2.f90:149-409) on the merge data alone.  Inputs d(n),
 * z(n), indxq(n), rho, n1 as MDLAED2.  Outputs: *k, dlamda(k), w(k), indxp(n),
 * indx(n), indxc(n), coltyp(n) (first 4 = CTOT), d updated, *rho_out, and the
 * Givens rotations (*nrot, rot_pj, rot_nj (1-based), rot_c, rot_s) in the order
 * MDLAED2 applies them with DROT (mdlaed2.f90:352).  hss_min_k picks the column
 * grouping exactly as BSMLZ does. */
int hsdc_deflate(hsdc_ctx *ctx, int n, int n1, double *d, double *z, int *indxq, double rho,
                 int hss_min_k, int *k, double *rho_out, double *dlamda, double *w, int *indxp,
                 int *indx, int *indxc, int *coltyp, int *nrot, int *rot_pj, int *rot_nj,
                 double *rot_c, double *rot_s);

/* The secular equation of one merge: all k roots of 1 + rho*sum w_i^2/(dlamda_i - x)
 * (DLAED4 at mdlaed3.f90:208), as lam(k), origin pole org(k) (1-based), tau(k) with
 * lam = dlamda(org) + tau, and iterations per root niter(k). */
int hsdc_secular(hsdc_ctx *ctx, int k, const double *dlamda, const double *w, double rho,
                 double *lam, int *org, double *tau, int *niter);

/* LAPACK DLAED3 (dense branch, mdlaed1.f90:199) on device. */
int hsdc_laed3(hsdc_ctx *ctx, int k, int n, int n1, double *d, double *q, int ldq, double rho,
               double *dlamda, const double *q2, const int *indx, const int *ctot, double *w);

/* MDLAED3's secular / Loewner / alpha half (mdlaed3.f90:201-249): roots d(k),
 * updated z (in place), difl(k), difr(k), alpha(k). */
int hsdc_loewner(hsdc_ctx *ctx, int k, double rho, double *dlamda, double *z, double *d,
                 double *difl, double *difr, double *alpha);

/* No counterpart in HSSPACK.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * hsdc_loewner on device arrays with no context, allocating nothing.  Inputs
 * d_dlamda(k) (poles, ascending), d_z(k), rho; outputs d_d(k) roots, d_znew(k)
 * Loewner z, d_difl(k), d_difr(k), d_alpha(k), *d_info (0, or nonzero when a root
 * did not converge), all device arrays; the classical starting iterate, at most 100
 * iterations per root.  Work is queued on `stream` (hipStream_t, 0 = null stream)
 * and the call returns without waiting.  Its scratch (per root: tau, origin,
 * iteration count; a control block) is one static device array per GPU of
 * 16 KB + 16 bytes x HSDC_NOALLOC_KMAX (about 1 MB), held until the process ends:
 * k > HSDC_NOALLOC_KMAX returns -1, as does k < 3.  Calls on one GPU share that
 * array, so calls that overlap on one GPU (other streams, other threads) race. */
#define HSDC_NOALLOC_KMAX 65536
int hsdc_loewner_noalloc(int k, double rho, const double *d_dlamda, const double *d_z, double *d_d,
                         double *d_znew, double *d_difl, double *d_difr, double *d_alpha, int *d_info,
                         void *stream);

/* DHSSEVC (CauchyHssEig.f90:2444): build the HSS form of
 * C(i,j) = u(i) v(j) / (f(i) - d(j)) (accurate differences from difl/difr) with
 * leaf block ni and tolerance tol, then q(1:nrow, 1:k) := q * C (ldq).  On return
 * ranks[0] holds the largest generator rank. */
int hsdc_hss_evc(hsdc_ctx *ctx, int k, int ni, double tol, const double *f, const double *d,
                 const double *u, const double *v, const double *difl, const double *difr,
                 double *q, int ldq, int nrow, int *ranks);

/* DHSSVCS (CauchyHssvd_VPS.f90:3587-3804, the SVD merge's vector matrices): as
 * hsdc_hss_evc for C(i,j) = u(i) v(j) / (f(i)^2 - d(j)^2), every displacement
 * formed as the accurate difference times the sum (d(j) are the roots sigma_j). */
int hsdc_hss_svc(hsdc_ctx *ctx, int k, int ni, double tol, const double *f, const double *d,
                 const double *u, const double *v, const double *difl, const double *difr,
                 double *q, int ldq, int nrow, int *ranks);

/* Construct-and-multiply of the merge's vector matrices (the eigenvector form of
 * DHSSEVC, or with squared = 1 the SVD form of DHSSVCS), on either side: the HSS form
 * of C (squared = 0: as hsdc_hss_evc; 1: as hsdc_hss_svc), then side 'R': x (ncol x k,
 * ldx) := x * C (fastHssmmL, ConstructHssd.f90:1394-1630), or side 'L': x (k x ncol,
 * ldx) := C * x (fastHssmm, :926-1155).  These forms take their differences from
 * difl/difr and need interlaced f and d; a general Cauchy-like matrix goes through
 * hsdc_cauchyhssmm.  ranks as for hsdc_hss_evc. */
int hsdc_hss_cmm(hsdc_ctx *ctx, char side, int squared, int k, int ni, double tol, const double *f,
                 const double *d, const double *u, const double *v, const double *difl,
                 const double *difr, double *x, int ldx, int ncol, int *ranks);

/* CauchyHssMM (cauchyhssmm.f90:1-123): b (n x n, ldb) := A * b for the general
 * Cauchy-like A(i,j) = u(i) v(j) / (d(i) - f(j)) with direct differences.  A is put in
 * HSS form by cauchy2hss (CauchyHss_VP2.f90:387-613: the structured RRLU with rho =
 * 1.02, the post-ordered tree of leaf size ni walked with a stack of compressed
 * skeletons, compression tolerance tol), then multiplied, ncol columns of b at a time
 * or more.  d, f, u, v (n) are not modified (the reference destroys them).  time[0]
 * and time[1] (may be NULL) receive the construction and multiplication times in
 * seconds, ranks (may be NULL; 2 per tree node) the rank of each node's U and V, or
 * ranks[0] = -1 when A was applied densely (a tree of one leaf).  INFO: -i for an
 * illegal argument i; -1 also when some d(i) equals some f(j) or an input is not finite. */
int hsdc_cauchyhssmm(hsdc_ctx *ctx, const double *d, const double *f, const double *u, const double *v,
                     int n, int ni, double tol, int ncol, double *b, int ldb, double *time, int *ranks);

/* DHSSMM (DHSSMM.f90:1-125): b(1:n, 1:k) := A b for the dense a (n x n, lda; m must
 * equal n, as the reference assumes), A put in HSS form by truncated SVDs (mat2hssvd,
 * ConstructHssd.f90:371-696: npart's tree of leaf size ni, every basis the singular
 * vectors whose singular values exceed tol), then multiplied ncol columns of b at a
 * time or more.  a is not modified (the reference uses it as workspace); prank is
 * accepted and unused, as in mat2hssvd.  time[0..1] (may be NULL): construction and
 * multiplication seconds; ranks (may be NULL; 2 per tree node): the ranks of each node's
 * U and V, or ranks[0] = -1 when A was applied densely.  INFO: -i for an illegal
 * argument i (-1 a, -2 lda, -3 m != n, -4 n, -5 b, -6 ldb, -7 k, -8 tol, -10 ni, -11 ncol). */
int hsdc_dhssmm(hsdc_ctx *ctx, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol,
                int prank, int ni, int ncol, double *time, int *ranks);
/* DHSSMML (DHSSMML.f90:2-122): b(1:k, 1:n) := b A with the same HSS form of A; ldb >= k. */
int hsdc_dhssmml(hsdc_ctx *ctx, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol,
                 int prank, int ni, int ncol, double *time, int *ranks);
/* DHSSMM / DHSSMML with randomized range finding (mat2hssvdr, ConstructHssd.f90:8-368;
 * the reference calls it only from its SVD-updating code, CompSvdup.f90:97, and has no
 * driver of its own): a block row or column whose context is longer than 50 and than
 * its own length m + 10 takes its basis from the SVD of the sketch Omega T (m + 10 rows)
 * instead of T, the rank = the count of the sketch's singular values above tol.  Omega
 * is read from one flat stream of uniform [0, 1) numbers as the reference reads its
 * random_number array (leading dimension m + 10 for a block column, the context length
 * for a block row); entry i of the stream is splitmix64 of seed + (i + 1) 0x9E3779B97F4A7C15,
 * top 53 bits (the reference draws it unseeded).  side 0: b(1:n, 1:k) := A b (ldb >= n);
 * 1: b(1:k, 1:n) := b A (ldb >= k).  Other arguments and INFO as hsdc_dhssmm; -12 side. */
int hsdc_dhssmm_rand(hsdc_ctx *ctx, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol,
                     int ni, int ncol, int side, unsigned long long seed, double *time, int *ranks);

/* Contexts owned by the calling host thread, for callers that keep no handle (the
 * Fortran drop-ins): created on first use with the default parameters (reset to them
 * at every hsdc_thread_ctx / hsdc_bd_thread_ctx call) and kept, with their device
 * buffers, until hsdc_thread_ctx_release or the end of the process.  NULL when no
 * context could be created. */
hsdc_ctx *hsdc_thread_ctx(void);
void hsdc_thread_ctx_release(void);  /* also releases the thread's hsdc_bd_thread_ctx */

/* Dense expansion of the HSS form built by the last hsdc_hss_evc / hsdc_hss_svc call
 * (k x k). */
int hsdc_hss_dense(hsdc_ctx *ctx, double *a, int lda);

/* ---- learned starting iterate of the secular solver ----
 * These have no counterpart in HSSPACK, whose DLAED4 always starts from the
 * classical two-pole guess.  With params.learned = 1 and weights loaded, every root
 * of a merge with K >= params.learned_min_k starts from an FP32 network's proposal
 * (inputs: 16 nearest pole gaps, 16 normalised z^2, rho, the classical guess; two
 * hidden layers of 64, ReLU) when its log-correction y of the classical offset
 * exceeds params.learned_gate in magnitude and the proposal lies inside the
 * classical bracket, from the classical guess otherwise.  The FP64 iteration and its stopping test are
 * unchanged, so the result is a converged DLAED4 root either way. */
#define HSDC_MLP_IN 34
#define HSDC_MLP_NPARAM 6465  /* W1(64x34) b1(64) W2(64x64) b2(64) w3(64) b3 */

/* Weights (host float32, HSDC_MLP_NPARAM); w = NULL unloads them. */
int hsdc_mlp_set_weights(hsdc_ctx *ctx, const float *w, int nparam);
/* Returns 1 when no weights are loaded. */
int hsdc_mlp_get_weights(hsdc_ctx *ctx, float *w, int nparam);
/* Weights file: "HSDCMLP1", int32 34, int32 64, int32 nparam, uint32 FNV-1a of the
 * weight bytes, float32 weights (little-endian).  1 = bad file or checksum. */
int hsdc_mlp_load(hsdc_ctx *ctx, const char *path);
/* Training data from the solver's own merges: every later solve on this context
 * appends (features, log(converged tau / classical starting tau)) of each root of
 * each merge with K >= learned_min_k, up to cap samples; cap = 0 stops harvesting.
 * The network's output y proposes tau = tau_classical * exp(clamp(y, -16, 16)). */
int hsdc_mlp_harvest(hsdc_ctx *ctx, long long cap);
/* Copies up to maxn harvested samples: x(34, n) row-major per sample, t(n).  With
 * x or t NULL, *n receives the number of samples held and nothing is copied. */
int hsdc_mlp_samples(hsdc_ctx *ctx, float *x, float *t, long long maxn, long long *n);
/* Gradient of the loaded weights on a batch (device kernels):
 * L = (1/ns) sum_s (y_s - t_s)^2; grad(HSDC_MLP_NPARAM). */
int hsdc_mlp_grad(hsdc_ctx *ctx, const float *x, const float *t, int ns, float *grad, float *loss);
/* Adam training of the loaded weights (seeded initialisation when none are loaded)
 * on the device; loss_hist(epochs) receives the mean batch loss of each epoch. */
int hsdc_mlp_train(hsdc_ctx *ctx, const float *x, const float *t, long long ns, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_hist);

/* ---------------------------------------------------------------------------
 * Bidiagonal divide-and-conquer SVD: HSSPACK's MDBDSDC (SRC/mdbdsdc.f90) ->
 * MDLASD0 -> MDLASD1 -> MDLASD2, then per merge the dense branch (DLASD3 / DLASD4,
 * K < BSMLZ) or the HSS branch (MDLASD8 / DHSSVCS, K >= BSMLZ, mdlasd1.f90:220-278):
 * the left and right vector matrices in HSS form, U2 * H_U and V2 * H_V.
 * ------------------------------------------------------------------------- */
typedef struct hsdc_bd_ctx hsdc_bd_ctx;
/* The calling thread's own SVD context (see hsdc_thread_ctx). */
hsdc_bd_ctx *hsdc_bd_thread_ctx(void);

typedef struct {
    int smlsiz;     /* leaf bound, ILAENV(9,'DBDSDC') = 25 (mdbdsdc.f90:193); 3..63 */
    int maxit;      /* secular iteration cap */
    int hss_min_k;  /* a merge with K >= hss_min_k takes the HSS branch: BSMLZ = 2000
                       (mdlasd2.f90:172); >= 1 */
    int hss_leaf;   /* HSS leaf block, Ni = 40 (mdlasd8.f90:277); >= 8 */
    double hss_tol; /* HSS compression tolerance, 1.0E-17 (CauchyHssvd_VPS.f90:3694) */
} hsdc_bd_params;

typedef struct {
    int levels, merges, leaves, k_top;
    long long roots, deflated, rotations, secular_iters;
    double ms_total;  /* device time of the last solve (events), ms */
    int hss_merges, hss_rank_max;  /* merges of the HSS branch, largest generator rank */
} hsdc_bd_stats;

void hsdc_bd_default_params(hsdc_bd_params *p);
int  hsdc_bd_ctx_create(hsdc_bd_ctx **ctx);
int  hsdc_bd_ctx_destroy(hsdc_bd_ctx *ctx);
int  hsdc_bd_set_params(hsdc_bd_ctx *ctx, const hsdc_bd_params *p);
int  hsdc_bd_get_stats(hsdc_bd_ctx *ctx, hsdc_bd_stats *s);
int  hsdc_bd_set_stream(hsdc_bd_ctx *ctx, void *stream);

/* MDBDSDC (mdbdsdc.f90:1-2).  uplo 'U' or 'L'; compq 'N' (values) or 'I' (values and
 * vectors; 'P', the compact form, returns -2).  d(n): diagonal in, singular values
 * (descending) out; e(n-1) destroyed; u(ldu,n), vt(ldvt,n) host arrays.  q, iq,
 * work, iwork are accepted for the reference's argument list and not used.  INFO as
 * MDBDSDC: -i for an illegal argument i, 1 when the secular iteration or a leaf
 * did not converge. */
int hsdc_bdsdc(hsdc_bd_ctx *ctx, char uplo, char compq, int n, double *d, double *e, double *u, int ldu,
               double *vt, int ldvt, double *q, int *iq, double *work, int *iwork);
/* The same on device arrays (COMPQ = 'I'; d_u and d_vt may be NULL for values only). */
int hsdc_bdsdc_dev(hsdc_bd_ctx *ctx, char uplo, int n, double *d_d, double *d_e, double *d_u, int ldu,
                   double *d_vt, int ldvt);
/* Merges of the last solve, in launch order: block start (0-based), order N, K,
 * deflation rotations and the largest HSS generator rank (0: dense branch); any
 * pointer may be null.  Returns the merge count (entries beyond cap are not written). */
int  hsdc_bd_trace(hsdc_bd_ctx *ctx, int *s, int *n, int *k, int *nrot, int *hss_rank, int cap);

/* ---------------------------------------------------------------------------
 * Banded divide and conquer: HSSPACK's MDSBEVD (SRC/mdsbevd.f90:183-184) with its merge
 * MDSBED1 (SRC/mdsbed1.f90:2): each KD x KD coupling block split off by its SVD
 * (DSymSplit, BasicMM.f90:1195), the leaves solved densely, and every merge taken as KD
 * rank-one updates -- dense secular vectors below hss_min_k, the HSS form of the
 * Cauchy-like vector matrix at or above it (MDSBED5 -> DHSSEVC, mdsbed5.f90:277).
 * KD <= 64 and leaves of at most 128 rows (smlsiz); the leaves must hold KD rows, so
 * smlsiz >= 2 KD - 1 whenever the matrix is split.
 * ------------------------------------------------------------------------- */
typedef struct hsdc_sb_ctx hsdc_sb_ctx;
/* The calling thread's own banded context (see hsdc_thread_ctx). */
hsdc_sb_ctx *hsdc_sb_thread_ctx(void);

typedef struct {
    int smlsiz;     /* leaf bound, 2 KD - 1 <= smlsiz <= 128 (default 128; the reference takes
                       ILAENV(9) = 25 for every KD, mdsbevd.f90:256) */
    int maxit;      /* secular iteration cap */
    int hss_min_k;  /* an update with K >= hss_min_k takes the HSS branch: BSMLZ = 1500
                       (mdsbed1.f90:134); 0 = never */
    int hss_leaf;   /* HSS leaf block, Ni = 40 (mdsbed5.f90:276); >= 8 */
    double hss_tol; /* HSS compression tolerance, 1.0E-17 */
    int profile;    /* 1: time the HSS construction and multiply with events (ms_hss_*) */
} hsdc_sb_params;

typedef struct {
    int levels, merges, leaves, updates, k_top;
    long long roots, deflated, rotations, secular_iters;
    int hss_updates, hss_rank_max;  /* updates of the HSS branch, largest generator rank */
    double ms_total;                /* device time of the last solve (events), ms */
    long long hss_nflops;           /* HSS construction work in HSSPACK's RRLU flop model */
    long long hss_mm_flops;         /* HSS multiply: 2 rows n k over its GEMMs */
    long long dense_flops;          /* dense-branch products: 2 N K^2 + 2 K^2 (KD - j - 1) */
    double ms_hss_build, ms_hss_mm; /* profile = 1: device time of the two HSS phases */
} hsdc_sb_stats;

void hsdc_sb_default_params(hsdc_sb_params *p);
int  hsdc_sb_ctx_create(hsdc_sb_ctx **ctx);
int  hsdc_sb_ctx_destroy(hsdc_sb_ctx *ctx);
int  hsdc_sb_set_params(hsdc_sb_ctx *ctx, const hsdc_sb_params *p);
int  hsdc_sb_get_stats(hsdc_sb_ctx *ctx, hsdc_sb_stats *s);
int  hsdc_sb_set_stream(hsdc_sb_ctx *ctx, void *stream);

/* MDSBEVD (mdsbevd.f90:183).  compz 'I' (values and vectors) or 'N' (values only; the
 * reference prints a message and returns for 'N'); uplo 'U' or 'L' band storage
 * ab(ldab, n), ldab >= kd + 1 (the reference reads 'U' whatever uplo says).  d(n):
 * eigenvalues ascending; q(ldq, n): eigenvectors (compz 'I').  work, iwork are accepted
 * for the reference's argument list and not used.  INFO: -i for an illegal argument i
 * (-20: smlsiz below 2 KD - 1, or a single leaf above 128 rows), 1 when a Jacobi leaf or
 * split, or a secular root, did not converge. */
int hsdc_sbevd(hsdc_sb_ctx *ctx, char compz, char uplo, int n, int kd, const double *ab, int ldab, double *d,
               double *q, int ldq, double *work, int lwork, int *iwork, int liwork);
/* The same on device arrays (d_q may be NULL for values only). */
int hsdc_sbevd_dev(hsdc_sb_ctx *ctx, char uplo, int n, int kd, const double *d_ab, int ldab, double *d_d,
                   double *d_q, int ldq);
/* Rank-one updates of the last solve, level by level, update by update, merge by merge:
 * block start (0-based), order N, K, deflation rotations and the largest HSS generator
 * rank (0: dense); any pointer may be null.  Returns the update count (entries beyond
 * cap are not written). */
int  hsdc_sb_trace(hsdc_sb_ctx *ctx, int *s, int *n, int *k, int *nrot, int *hss_rank, int cap);

#ifdef __cplusplus
}
#endif
#endif
