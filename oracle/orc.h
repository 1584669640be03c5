/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of HSSPACK's tridiagonal divide-and-conquer merge path
 * (SRC/mdstedc.f90, mdlaed0.f90, mdlaed1.f90, mdlaed2.f90, mdlaed3.f90,
 * CauchyHssEig.f90, aux_hss.f90, ConstructHssd.f90) together with the LAPACK
 * routines that path calls (DLAED4, DLAED3, DLAMRG, DSTEQR).  It is the checker
 * for the HIP product library and the CPU baseline timed next to it.  Nothing in
 * the product links, loads or calls it; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg do.
 *
 * Index arrays (INDXQ, INDX, INDXC, INDXP, COLTYP) hold 1-based values exactly as
 * the Fortran routines do, so they compare directly with LAPACK's outputs.
 * Matrices are column-major.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_dlamrg   -- LAPACK dlamrg on random / tied inputs (bit-exact index sets)
 *   orc_laed4    -- LAPACK dlaed4 (roots and DELTA to a few ulps), K=1 and K=2 closed
 *                   forms, interlacing, mpmath 50-digit roots
 *   orc_steqr    -- LAPACK dsteqr (eigenvalues), orthogonality and residual
 *   orc_laed2    -- LAPACK dlaed2 (K, DLAMDA, W, INDXP, INDXC, COLTYP, CTOT bit-exact;
 *                   Q and Q2 to the rotation rounding)
 *   orc_laed3    -- LAPACK dlaed3 (eigenvalues and vectors within 1e-13)
 *   orc_mdlaed3  -- dense eig of diag(d)+rho*z*z' (the HSS product of the Cauchy
 *                   matrix against the explicitly formed one)
 *   orc_hss_*    -- dense expansion of the HSS form against the dense Cauchy-like
 *                   matrix, ranks monotone in tol, identity rows at skeletons
 *   orc_stedc    -- LAPACK dstedc (eigenvalues <= 4 n eps |T|; |Q'Q-I| and
 *                   |TQ-QL|/|T| <= 1e-12)
 */
#ifndef HSDC_ORACLE_H
#define HSDC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int smlsiz;   /* largest leaf of the DC tree, ILAENV(9,'DSTEDC') = 25 */
    int bsmlz;    /* a merge with K >= bsmlz takes the HSS branch (mdlaed2.f90:175) */
    int ni;       /* HSS leaf block size (mdlaed3.f90:253) */
    double tol;   /* HSS compression tolerance (CauchyHssEig.f90:2523) */
    int maxit;    /* secular iteration cap */
    int bd_bsmlz; /* an SVD merge with K >= bd_bsmlz takes the HSS branch (mdlasd2.f90:172) */
    int bd_ni;    /* its HSS leaf block size (mdlasd8.f90:277) */
} orc_params;

/* per-merge trace filled by orc_stedc when non-NULL; nrot (deflation rotations per
 * merge) and indxq (INDXQ of the top merge of an unsplit problem of order indxq_n,
 * the permutation of mdlaed0.f90:355-362) are optional (NULL) */
typedef struct {
    int cap, nmerge;
    int *level, *n, *k, *hss, *rank_max;
    long long roots, iters;
    int *nrot, *indxq;
    int indxq_n;
} orc_trace;

void orc_default_params(orc_params *p);

/* LAPACK DLAMRG: stable merge permutation of two sorted runs (strides +-1). */
void orc_dlamrg(int n1, int n2, const double *a, int s1, int s2, int *index);

/* Secular root J (1-based) of 1 + rho*sum z_i^2/(d_i - lambda).  delta(i) = d(i)-lambda
 * formed as (d(i)-d(org))-tau.  For n==2 delta holds the normalised eigenvector, as
 * LAPACK's DLAED5 does.  Returns 0, or 1 if the iteration did not converge. */
int orc_laed4(int n, int j, const double *d, const double *z, double *delta, double rho,
              double *dlam, int *org, double *tau, int *niter, int maxit);

/* Learned starting iterate (FP32 network 34 -> 64 -> 64 -> 1, ReLU): feature map of
 * root r (0-based) with bracket width W and classical start xc (bracket fraction),
 * forward pass y, and the secular solve started from tau_classical * exp(clamp(y, -16,
 * 16)) when |y| > gate and that lies inside the bracket (*used = 1), from the
 * classical start otherwise.  wmlp == NULL is orc_laed4.  orc_sec_exp is the exp both sides use. */
#define ORC_MLP_IN 34
#define ORC_MLP_H 64
#define ORC_MLP_W1 0
#define ORC_MLP_B1 (ORC_MLP_W1 + ORC_MLP_H * ORC_MLP_IN)
#define ORC_MLP_W2 (ORC_MLP_B1 + ORC_MLP_H)
#define ORC_MLP_B2 (ORC_MLP_W2 + ORC_MLP_H * ORC_MLP_H)
#define ORC_MLP_W3 (ORC_MLP_B2 + ORC_MLP_H)
#define ORC_MLP_B3 (ORC_MLP_W3 + ORC_MLP_H)
#define ORC_MLP_NPARAM (ORC_MLP_B3 + 1)
void orc_mlp_features(int n, int r, const double *d, const double *z, double rho, double W, double xc,
                      float *f);
float orc_mlp_forward(const float *w, const float *x);
float orc_mlp_forward_team(const float *w, const float *x, int T);
/* lanes per secular root on the device for an n-root equation (summation order) */
int orc_sec_team(int n);
double orc_sec_exp(double x);
int orc_laed4_mlp(int n, int j, const double *d, const double *z, double *delta, double rho,
                  double *dlam, int *org, double *tau, int *niter, int maxit, const float *wmlp,
                  float gate, int *used);

/* Implicit-shift QL on a symmetric tridiagonal of order n, eigenvalues ascending and
 * eigenvectors in z (ldz), z is overwritten with the eigenvector matrix. */
int orc_steqr(int n, double *d, double *e, double *z, int ldz);
/* Eigenvalues only, same iteration. */
int orc_sterf(int n, double *d, double *e);

/* MDLAED2 (mdlaed2.f90:149) == LAPACK DLAED2 when K < bsmlz.  *hflag=1 dense branch,
 * 0 HSS branch.  Optional rotation log (may be NULL). */
int orc_laed2(int *k, int n, int n1, double *d, double *q, int ldq, int *indxq, double *rho,
              double *z, double *dlamda, double *w, double *q2, int *indx, int *indxc,
              int *indxp, int *coltyp, int bsmlz, int *hflag,
              int *nrot, int *rot_pj, int *rot_nj, double *rot_c, double *rot_s);

/* LAPACK DLAED3 (dense branch called at mdlaed1.f90:199). iters (may be NULL) receives
 * the total secular iterations. */
int orc_laed3(int k, int n, int n1, double *d, double *q, int ldq, double rho,
              double *dlamda, const double *q2, const int *indx, const int *ctot,
              double *w, double *s, int maxit, long long *iters);

/* MDLAED3 (mdlaed3.f90:124), HSS branch: secular roots, Loewner z, alpha, DHSSEVC. */
int orc_mdlaed3(int k, int n, double *d, double *q, int ldq, double rho, double *dlamda,
                double *q2, double *z, double *difl, double *difr, double *alpha,
                const orc_params *p, int *rank_max, long long *iters);

/* The secular / Loewner / alpha half of MDLAED3 alone (mdlaed3.f90:201-249). */
int orc_mdlaed3_vectors(int k, double *d, double rho, double *dlamda, double *z, double *difl,
                        double *difr, double *alpha, int maxit, long long *iters);

/* MDLAED1 (mdlaed1.f90:109) */
int orc_laed1(int n, double *d, double *q, int ldq, int *indxq, double rho, int cutpnt,
              double *work, int *iwork, const orc_params *p, int *kout, int *hssout,
              int *rank_max, long long *iters);

/* MDLAED0 with ICOMPQ=2 (mdlaed0.f90:121) */
int orc_laed0(int n, double *d, double *e, double *q, int ldq, double *work, int *iwork,
              const orc_params *p, orc_trace *tr);

/* MDSTEDC (mdstedc.f90:139).  compz 'N','I' or 'V'.  lwork/liwork == -1 is a query. */
int orc_stedc(char compz, int n, double *d, double *e, double *z, int ldz, double *work,
              int lwork, int *iwork, int liwork, const orc_params *p, orc_trace *tr);

/* ---- HSS (CauchyHssEig.f90 / aux_hss.f90 / ConstructHssd.f90) ---- */
typedef struct {
    int k, ltr, lm, lvl;
    int *tr, *m;            /* post-ordered parent array (ltr), leaf sizes (lm) */
    int *ch;                /* children, 2*ltr (0 = none) */
    int *lentr, *ttr;       /* level lists: lentr[lvl], ttr[(LL-1)*ltr + i] */
    int *dsz, *usz, *vsz, *bsz;  /* D size, U/V/B dims (2 per node) */
    long *pu, *pv, *pb, *pd;     /* offsets into h and dd (0-based) */
    double *h, *dd;
    long nh, ndd;
    long long nflops, nswap;
} orc_hss;

/* DHSSEVC construction half: HSS approximation of A(i,j)=u(i)v(j)/(f(i)-d(j)). */
int orc_hss_build(orc_hss *H, int k, int ni, const double *f, const double *d0,
                  const double *u, const double *v, const double *difl, const double *difr,
                  double tol);
/* fastHssmmL: X(ldx x K, rows nrow) := X * H */
int orc_hss_mmL(const orc_hss *H, double *x, int ldx, int nrow);
/* Dense expansion of H into A (k x k, lda) for checks. */
void orc_hss_dense(const orc_hss *H, double *a, int lda);
void orc_hss_free(orc_hss *H);

/* Cauchy fill with accurate differences (CauchylikeEig, CauchyHssEig.f90:1688). */
void orc_cauchy_dense(int k, const double *f, const double *u, const double *v,
                      const double *difl, const double *difr, double *a, int lda);
/* The SVD's vector matrices A(i,j) = u(i)v(j)/(f(i)^2 - d0(j)^2) (DHSSVCS,
 * CauchyHssVD_VPS.f90:3587-3804): the same construction in the squared variables,
 * every displacement formed as accurate difference x sum. */
int orc_hss_build_sq(orc_hss *H, int k, int ni, const double *f, const double *d0,
                     const double *u, const double *v, const double *difl, const double *difr,
                     double tol);
void orc_cauchy_dense_sq(int k, const double *f, const double *d0, const double *u, const double *v,
                         const double *difl, const double *difr, double *a, int lda);
/* The general Cauchy-like matrix A(i,j) = u(i) v(j) / (d(i) - f(j)) of order n, direct
 * differences: its HSS form by cauchy2hss (CauchyHss_VP2.f90:387-613: structured RRLU
 * with rho = 1.02, post-order traversal with the stack of compressed skeletons), and
 * its dense fill (Cauchylike.f90:1-13).  Returns -1 for a tree of one leaf. */
int orc_cauchy2hss(orc_hss *H, int n, int ni, const double *d, const double *f, const double *u,
                   const double *v, double tol);
void orc_cauchy_dense_general(int n, const double *d, const double *f, const double *u, const double *v,
                              double *a, int lda);
/* A dense square matrix of order n (lda) in HSS form by truncated SVDs (DHSSMM /
 * DHSSMML's mat2hssvd, ConstructHssd.f90:371-696): npart's tree with leaf size ni, every
 * basis the singular vectors above tol (absolute).  Returns -1 for a tree of one leaf,
 * 1 when a Jacobi SVD did not converge. */
int orc_mat2hss(orc_hss *H, int n, int ni, const double *a, int lda, double tol);
/* The same with mat2hssvdr's randomized range finding (ConstructHssd.f90:8-368), the
 * uniform array drawn from the splitmix64 stream of seed. */
int orc_mat2hss_rand(orc_hss *H, int n, int ni, const double *a, int lda, double tol, unsigned long long seed);

/* ---- bidiagonal SVD (mdbdsdc.f90 / mdlasd0-2.f90 + LAPACK DLASD3/DLASD4/DLASDQ) ---- */
/* LAPACK DLASCL 'G' on a vector, DLARTG, DLASDT (node arrays 0-based, values 1-based). */
void orc_lascl(double cfrom, double cto, int n, double *x);
void orc_lartg(double f, double g, double *c, double *s, double *r);
void orc_lasdt(int n, int *lvl, int *nd, int *inode, int *ndiml, int *ndimr, int msub);
/* Leaf SVD of the m x (m+sqre) upper bidiagonal (replaces DLASDQ): singular values
 * ascending in d, U m x m, VT (m+sqre)^2 with the null vector last when sqre = 1.
 * One-sided Jacobi; returns the sweeps, < 0 on failure.  m + sqre <= 64. */
int orc_bd_leaf(int m, int sqre, double *d, const double *e, double *u, int ldu, double *vt,
                int ldvt);
/* Root j (1-based) of 1/rho + sum z_i^2/((d_i-s)(d_i+s)) (replaces DLASD4): s = d(org)+tau. */
int orc_lasd4(int k, int j, const double *d, const double *z, double rho, double *sigma, int *org,
              double *tau, int *niter, int maxit);
/* LAPACK DLASD2 == MDLASD2 (K < BSMLZ); rotation log optional (NULL). */
int orc_lasd2(int nl, int nr, int sqre, int *k, double *d, double *z, double alpha, double beta,
              double *u, int ldu, double *vt, int ldvt, double *dsigma, double *u2, int ldu2,
              double *vt2, int ldvt2, int *idxp, int *idx, int *idxc, int *idxq, int *coltyp,
              int *nrot, int *rot_p, int *rot_q, double *rot_c, double *rot_s);
/* LAPACK DLASD3 (dense branch); q: 2k^2 + 3k doubles. */
int orc_lasd3(int nl, int nr, int sqre, int k, double *d, double *q, const double *dsigma,
              double *u, int ldu, const double *u2, int ldu2, double *vt, int ldvt,
              const double *vt2, int ldvt2, const int *idxc, const int *ctot, double *z, int maxit,
              long long *iters);
/* MDLASD1 with the dense branch for every K. */
int orc_lasd1(int nl, int nr, int sqre, double *d, double *alpha, double *beta, double *u, int ldu,
              double *vt, int ldvt, int *idxq, const orc_params *p, int *kout, long long *iters,
              int *nrot, int *hss_rank);
/* MDLASD8 (mdlasd8.f90:201-278): roots d, updated z, DIFL, DIFR, APHAU, APHAV. */
int orc_lasd8(int k, double *d, const double *dsigma, double *z, double *difl, double *difr,
              double *aphau, double *aphav, int maxit, long long *iters);
/* 1 when an HSS tree of k columns in leaves of ni has 2 to 512 leaves. */
int orc_hss_tree_ok(int k, int ni);
/* MDLASD0; trace (may be NULL): level, n, k per merge, the deflation rotations in the
 * hss slot and the largest HSS generator rank in rank_max (0: dense branch). */
int orc_lasd0(int n, int sqre, double *d, double *e, double *u, int ldu, double *vt, int ldvt,
              const orc_params *p, orc_trace *tr);
/* MDBDSDC, COMPQ 'N' or 'I' ('N' still forms the vectors in u, vt). */
int orc_bdsdc(char uplo, char compq, int n, double *d, double *e, double *u, int ldu, double *vt,
              int ldvt, const orc_params *p, orc_trace *tr);

/* ---- banded divide and conquer (mdsbevd.f90, mdsbed1-5.f90) ---- */
/* per rank-one update (KD per merge, merges level by level): order n, K, rotations */
typedef struct {
    int cap, nupd, nmerge;
    int *k, *n, *nrot;
    long long iters;
    double *sv;   /* optional: the singular values of each cut's coupling block (kd per cut) */
} orc_band_trace;
/* Cyclic two-sided Jacobi eigensolver of a dense symmetric m x m matrix (replaces
 * DSYEVD at the leaves): eigenvalues ascending, vectors in v.  Sweeps, < 0 on failure. */
int orc_sym_jacobi(int m, double *a, int lda, double *w, double *v, int ldv);
/* One-sided Jacobi SVD of an m x m matrix (replaces DGESVD on the coupling blocks):
 * c = u diag(s) v^T with s descending. */
int orc_jacobi_svd(int m, double *c, int ldc, double *s, double *u, int ldu, double *v, int ldv);
/* MDSBEVD with COMPZ = 'I', UPLO = 'U': the band ab (ldab >= kd + 1, upper storage
 * AB(kd+1+i-j, j) = A(i,j)) of order n; eigenvalues ascending in d, vectors in q.
 * smlsiz >= 2 kd - 1 so that every leaf holds at least kd rows.  Returns INFO. */
int orc_sbevd(int n, int kd, const double *ab, int ldab, double *d, double *q, int ldq, int smlsiz,
              orc_band_trace *tr);

#ifdef __cplusplus
}
#endif
#endif
