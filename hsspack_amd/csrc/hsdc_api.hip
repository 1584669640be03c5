// Host driver and C ABI of the MI355X divide-and-conquer eigensolver.
// Boundary: include/hsdc.h.  The schedule follows MDSTEDC/MDLAED0
// (mdstedc.f90:139-436, mdlaed0.f90:121-380): split on negligible off-diagonals,
// scale, tear into leaves, solve the leaves, then merge level by level.  All
// merges of a level (of every problem in a batch) run in the same launches;
// the per-merge schedule (params.batched = 0) launches them one at a time.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <cmath>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <thread>
#include <vector>

#include "hsdc_internal.h"
#include "hsdc_hss.h"
#include "hsdc_hss_host.h"

#define HSS_LDS_V 16384  // doubles of dynamic LDS for an RRLU node's generator column
#include "hsdc_mlp.h"
#include "../../include/hsdc.h"

// kernels (hsdc_kern.hip)
struct GTile { int m, half, r0, c0; };
__global__ void k_leaf(const Leaf *, int, const Prob *, const double *, const double *, Work, int, int, int);
__global__ void k_prep(const Merge *, const int *, int, const Prob *, Work, int);
__global__ void k_mergepath(const Merge *, const int *, int, Work);
__global__ void k_fill_u64(unsigned long long *, long, unsigned long long);
template <int NT>
__global__ void k_deflate(const Merge *, Work, int, MergeOut *, int, int);
__global__ void k_rot_prep(const Merge *, const Prob *, Work, int, const MergeOut *);
__global__ void k_rotate(const Merge *, const int *, int, const Prob *, Work, int, const MergeOut *);
__global__ void k_secular(const Merge *, const int *, int, const MergeOut *, Work, int, int *, const SecCtl *,
                          unsigned int *);
__global__ void k_secstart(const Merge *, const int *, int, const MergeOut *, Work, const SecCtl *);
__global__ void k_loewner(const Merge *, const int *, int, const MergeOut *, Work);
__global__ void k_sbuild(const Merge *, const int *, int, const MergeOut *, const long *, Work, double *);
__global__ void k_alpha(const Merge *, const int *, int, const MergeOut *, Work);
__global__ void k_gemm(const Merge *, const GTile *, int, const int *, int, const Prob *, const MergeOut *,
                       const long *, const long *, Work, int, const double *, double *);
__global__ void k_newcols(const Merge *, const int *, int, const MergeOut *, Work, int);
__global__ void k_scatter(const Merge *, const int *, int, const MergeOut *, const long *, const Prob *,
                          Work, int, const double *, int);
__global__ void k_lamrg(const Merge *, const int *, int, const MergeOut *, Work, int, int *);
__global__ void k_readback(Work, const int *, const unsigned long long *, const SecCtl *, const unsigned int *,
                           SolveRec *);
__global__ void k_inp_scan(const double *, const double *, int, Work);
__global__ void k_inp_scale(const double *, const double *, int, const int *, Merge *, int, double *, double *,
                            double *, Work);
__global__ void k_newcols_copy(const Merge *, const int *, int, int, const int *, const MergeOut *, const long *,
                               const Prob *, Work, int, const double *);
__global__ void k_lvltab(const Merge *, int, const MergeOut *, Work, LvlTab, long, long, int *);
__global__ void k_final(const Prob *, const int *, Work, const int *, double *, const long *, const long *, int,
                        double *, const double *, unsigned long long *, const int *, int, int, int);
__global__ void k_zfill(double *, long, long, long);
__global__ void k_dgemm_nn(int, int, int, const double *, long, const double *, long, double *, long);
__global__ void k_transpose(int, int, const double *, long, double *, long);
__global__ void k_cauchy2hss(C2H);
struct M2hJob {
    int rs, rn, cs, cn;
    int w, m;
    long toff;
    long boff;
};
struct M2hBlk {
    int start, count;
    int src;
    int ostart, ocount;
};
struct M2hCopy {
    int kind, rows, cols, job;
    int r0, c0;
    long dst;
};
__global__ void k_m2h_gather(const M2hJob *, const double *, long, double *);
__global__ void k_m2h_svd(const M2hJob *, double *, double *, double, int *, int *);
__global__ void k_m2h_sketch(const M2hJob *, const M2hJob *, const double *, double *, unsigned long long);
__global__ void k_m2h_rows(const M2hBlk *, int, const M2hJob *, const double *, const double *, long, int, double *,
                           long);
__global__ void k_m2h_cols(const M2hBlk *, int, const M2hJob *, const double *, const double *, long, int, double *);
__global__ void k_m2h_copy(const M2hCopy *, const M2hJob *, const double *, const double *, long, const double *, long,
                           double *, double *);
__global__ void k_cfill_general(int, const double *, const double *, const double *, const double *, double *);
__global__ void k_hss_init(const HssMerge *);
__global__ void k_hss_reset(const HssMerge *);
__global__ void k_hss_dgen(const HssMerge *, const HssNode *, int);
template <int NT, bool SQ>
__global__ void k_hss_rowcomp(const HssMerge *, const HssNode *, int, double, int);
template <int NT, bool SQ>
__global__ void k_hss_colcomp(const HssMerge *, const HssNode *, int, double, int);
__global__ void k_hss_compact(const HssMerge *, const HssCompact *, const int *, int);
__global__ void k_hss_bgen(const HssMerge *, const HssNode *, int);
__global__ void k_bgemm(const BG *);
template <int BM, int BN, int BK, int DB, int TA, int TB>
__global__ void k_bgemm_w(const BG *, int);
__global__ void k_gatherx(const Merge *, const int *, const int *, int, const Prob *, Work, int, double *,
                          const long *);
__global__ void k_cbuild(const HssMerge *, int, double *);
// learned-start training (hsdc_mlp.hip)
__global__ void k_mlp_grad(const float *, const float *, const float *, const int *, int, float, float *, float *);
__global__ void k_mlp_sum(const float *, int, float *, const float *, float *);
__global__ void k_mlp_adam(float *, const float *, float *, float *, float, float, float);

#define WGS 256

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

int ensure(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return 0;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 8;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        if (hipMalloc(&b.p, bytes) != hipSuccess) { (void)hipGetLastError(); b.p = nullptr; return HSDC_ERR_NOMEM; }
        want = bytes;
    }
    b.cap = want;
    return 0;
}

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
};

int ensure_pin(PinBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return 0;
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    size_t want = bytes * 2 + 4096;
    if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) { b.p = nullptr; b.cap = 0; return HSDC_ERR_NOMEM; }
    b.cap = want;
    return 0;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
// every sub-table of a packed upload starts on a 64-byte boundary, so 8-byte
// fields and 64-bit atomics in it are naturally aligned
inline size_t al64(size_t x) { return (x + 63) & ~(size_t)63; }

}  // namespace

// host plan of one solve: its problems, the leaves and the merges of every level
struct HostProb {
    int off, n;        // global offset and order
    double scale;      // eigenvalues are multiplied back by this (mdstedc.f90:358)
    int levels;        // DC merge levels
};

struct Plan {
    int ntot = 0;
    std::vector<HostProb> probs;
    std::vector<Leaf> leaves;
    std::vector<std::vector<Merge>> lv;  // merges per level (level 0 = first merge level)
};

struct hsdc_ctx {
    hsdc_params prm;
    hsdc_stats st;
    hipStream_t stream = 0;
    // device storage
    DevBuf q, ct, sb, din, ein, wsd, wsi, tabs, lvl, mo, probs, misc, zout, dout;
    DevBuf hss, htab, xb, gf, cb, yb, tb;
    // generators of the last hsdc_hss_evc call (host copies), for hsdc_hss_dense
    std::vector<double> evc_gen;
    int evc_k = 0, evc_ni = 0, evc_sq = 0;
    double evc_tol = 0.0;
    PinBuf pin, pin2, pin3;
    PinBuf pinh;                        // the HSS generator directories of one batch
    PinBuf pinb;                        // the HSS multiply's GEMM item table
    hipEvent_t pinb_ev = nullptr;       // recorded after its upload
    PinBuf stage;                       // staging of large device-to-host copies
    hipEvent_t stev[3] = {nullptr, nullptr, nullptr};
    Work W;
    size_t wcap = 0;
    // learned secular start: weights (device + host copy), harvest buffer
    // (count | features | targets), secular control block of the running solve
    DevBuf mlpw, hv, sec, mlptr, maux;
    std::vector<float> hw;
    long long hv_cap = 0;
    const SecCtl *dsec = nullptr;
    unsigned int *dlvlmax = nullptr;  // largest iteration count of each level's secular launch
    int cur_level = 0;
    bool sec_pre = false;              // k_secstart runs before k_secular
    DevBuf featb;
    DevBuf ltab;  // tables of the device-scheduled level (k_lvltab)
    std::vector<long> tabkey;  // shape of the static tables kept on the device (device input)
    std::vector<Prob> hprobs;  // the problems of the running solve (host copy)
    Plan fplan;                // plan of a device-input solve, by order and leaf bound
    int fplan_n = -1, fplan_sml = -1;
    hipGraphExec_t gexec = nullptr;  // the last captured solve and its shape
    std::vector<long> gkey;
    hipStream_t gstream = nullptr;   // stream of graph solves called on the null stream
    hipEvent_t pf0 = nullptr, pf1 = nullptr;  // around the final kernel (profile = 1)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around a whole solve
    hipStream_t zstream = nullptr;             // zero fill of Z beside the DC levels
    hipEvent_t zev0 = nullptr, zev1 = nullptr; // its fork and join
    hipEvent_t pz0 = nullptr, pz1 = nullptr;   // around it (profile = 1)
    SolveRec *rec = nullptr, *drec = nullptr;  // end-of-solve record: mapped host memory, its device address
    // per-merge trace of the last solve: K, rotations and branch written by k_lamrg
    // into dtrc (3 ints per merge, level by level), the merges and their HSS ranks
    // kept by the host; valid until a single-merge call reuses the tables
    int *dtrc = nullptr;
    std::vector<Merge> trc_mg;
    std::vector<int> trc_rank;
    int trc_ntot = 0;
    bool trc_valid = false;
    // profile = 1: event pairs around the secular (kind 0) and Loewner / alpha (kind 1)
    // launches of a solve, read at its end
    std::vector<hipEvent_t> pev;
    std::vector<int> pkind;
    size_t pused = 0;
};


// DC tree of one problem (mdlaed0.f90:192-219): halve every block until the last
// one is <= smlsiz, tear with |E| at each cut.
static void plan_problem(Plan &pl, int pi, double *d, const double *e, int smlsiz)
{
    HostProb &P = pl.probs[pi];
    int n = P.n, off = P.off;
    std::vector<int> sz(1, n);
    int tl = 0;
    if (n > smlsiz) {
        while (sz.back() > smlsiz) {
            std::vector<int> nx(sz.size() * 2);
            for (size_t j = 0; j < sz.size(); j++) {
                nx[2 * j] = sz[j] / 2;
                nx[2 * j + 1] = (sz[j] + 1) / 2;
            }
            sz.swap(nx);
            tl++;
        }
    }
    P.levels = tl;
    std::vector<int> ends(sz.size());
    int acc = 0;
    for (size_t j = 0; j < sz.size(); j++) { acc += sz[j]; ends[j] = acc; }
    for (size_t i = 0; i + 1 < sz.size(); i++) {
        int sub = ends[i];  // 0-based index of first element of block i+1
        double ae = fabs(e[off + sub - 1]);
        d[off + sub - 1] -= ae;
        d[off + sub] -= ae;
    }
    int st = 0;
    for (size_t i = 0; i < sz.size(); i++) {
        pl.leaves.push_back(Leaf{off + st, sz[i], pi});
        st = ends[i];
    }
    // merges level by level (mdlaed0.f90:294-341)
    std::vector<int> bend = ends;
    for (int L = 0; L < tl; L++) {
        if ((int)pl.lv.size() <= L) pl.lv.resize(L + 1);
        std::vector<int> nb;
        for (size_t i = 0; i + 1 < bend.size(); i += 2) {
            int s = (i == 0) ? 0 : bend[i - 1];
            int matsiz = bend[i + 1] - s;
            int msd2 = (i == 0) ? bend[0] : matsiz / 2;
            Merge M;
            M.s = off + s; M.n = matsiz; M.n1 = msd2; M.prob = pi; M.lvl = L;
            M.rho = e[off + s + msd2 - 1];
            pl.lv[L].push_back(M);
            nb.push_back(bend[i + 1]);
        }
        bend.swap(nb);
    }
}

#define CK(x) do { int _r = (x); if (_r) return _r; } while (0)
#define HK(x) do { if ((x) != hipSuccess) { (void)hipGetLastError(); return HSDC_ERR_HIP; } } while (0)

static int have_device()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { (void)hipGetLastError(); return 0; }
    return 1;
}

static int alloc_work(hsdc_ctx *c, int ntot)
{
    if (!have_device()) return HSDC_ERR_NODEVICE;
    if ((size_t)ntot <= c->wcap && c->wcap) return 0;
    size_t nd = 22, ni = 19;
    CK(ensure(c->wsd, sizeof(double) * nd * (size_t)ntot));
    CK(ensure(c->wsi, sizeof(int) * ni * (size_t)ntot));
    double *D = (double *)c->wsd.p;
    int *I = (int *)c->wsi.p;
    Work &W = c->W;
    size_t n = ntot;
    W.Dv[0] = D + 0 * n; W.Dv[1] = D + 1 * n; W.z = D + 2 * n; W.dlamda = D + 3 * n; W.w = D + 4 * n;
    W.ddef = D + 5 * n; W.zs = D + 6 * n; W.ds = D + 7 * n; W.rot_c = D + 8 * n; W.rot_s = D + 9 * n;
    W.lam = D + 10 * n; W.tau = D + 11 * n; W.wnew = D + 12 * n; W.alpha = D + 13 * n;
    W.difl = D + 14 * n; W.difr = D + 15 * n; W.evd = D + 16 * n;
    W.vm[0] = I + 0 * n; W.vm[1] = I + 1 * n; W.indxq = I + 2 * n; W.supp_lo = I + 3 * n;
    W.supp_hi = I + 4 * n; W.indxq2 = I + 5 * n; W.indx = I + 6 * n; W.indxc = I + 7 * n;
    W.indxp = I + 8 * n; W.coltyp = I + 9 * n; W.rot_pj = I + 10 * n; W.rot_nj = I + 11 * n;
    W.org = I + 12 * n; W.nit = I + 13 * n; W.nsl = I + 14 * n; W.evq = I + 15 * n;
    W.st_t = D + 18 * n; W.st_tc = D + 19 * n; W.st_lo = D + 20 * n; W.st_hi = D + 21 * n;
    W.st_ii = I + 18 * n;
    // per-merge records of one launch (a level has at most ntot / 2 merges), kept
    // zero between launches by their consumer; then the level gate and statistics
    const size_t nmaux = 4 * ((size_t)ntot / 2 + 2);
    CK(ensure(c->maux, sizeof(unsigned long long) * (nmaux + 8)));
    HK(hipMemset(c->maux.p, 0, sizeof(unsigned long long) * (nmaux + 8)));
    W.maux = (unsigned long long *)c->maux.p;
    W.gate = (int *)(W.maux + nmaux);
    W.lstat = W.maux + nmaux + 1;
    c->wcap = ntot;
    return 0;
}

// Open the level gate and zero the level statistics (a kernel, not a memset, so
// that the stream may be under graph capture); gate_open leaves the statistics.
static void gate_reset(hsdc_ctx *c)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(WGS), 0, c->stream, (unsigned long long *)c->W.gate, 8L, 0ULL);
}

static void gate_open(hsdc_ctx *c)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(WGS), 0, c->stream, (unsigned long long *)c->W.gate, 1L, 0ULL);
}

// profile = 1: record the opening (kind >= 0) or closing (kind < 0) event of a timed
// launch group on the context's stream
static void prof_mark(hsdc_ctx *c, int kind)
{
    if (!c->prm.profile) return;
    if (c->pused == c->pev.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return; }
        c->pev.push_back(e);
        c->pkind.push_back(0);
    }
    c->pkind[c->pused] = kind;
    if (hipEventRecord(c->pev[c->pused], c->stream) != hipSuccess) { (void)hipGetLastError(); return; }
    c->pused++;
}

// the timed groups of the finished solve into the statistics
static void prof_collect(hsdc_ctx *c)
{
    for (size_t i = 0; i + 1 < c->pused; i++) {
        if (c->pkind[i] < 0 || c->pkind[i + 1] >= 0) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->pev[i], c->pev[i + 1]) != hipSuccess) { (void)hipGetLastError(); continue; }
        if (c->pkind[i] == 0) c->st.ms_secular += ms;
        else c->st.ms_loewner += ms;
    }
    c->pused = 0;
}

// Secular control block of the next solve: learned-start weights when enabled and
// loaded, the harvest buffer when harvesting, and zeroed iteration counters.
static int sec_setup(hsdc_ctx *c)
{
    // device layout: SecCtl | iters, used, wave_iters, evals [SEC_SLOTS each] | level maxima [SEC_LEVELS][SEC_SLOTS]
    const size_t hdr = al64(sizeof(SecCtl));
    const long ncnt = 4 * SEC_SLOTS + SEC_LEVELS * SEC_SLOTS / 2;  // in 8-byte words
    CK(ensure(c->sec, hdr + 8 * (size_t)ncnt));
    CK(ensure_pin(c->pin3, sizeof(SecCtl)));
    char *base = (char *)c->sec.p;
    SecCtl &h = *(SecCtl *)c->pin3.p;
    memset(&h, 0, sizeof(h));
    h.w = (c->prm.learned && !c->hw.empty()) ? (const float *)c->mlpw.p : nullptr;
    h.iters = (unsigned long long *)(base + hdr);
    h.used = h.iters + SEC_SLOTS;
    h.wave_iters = h.iters + 2 * SEC_SLOTS;
    h.evals = h.iters + 3 * SEC_SLOTS;
    if (c->hv_cap > 0) {
        char *hv = (char *)c->hv.p;
        h.count = (unsigned long long *)hv;
        h.hf = (float *)(hv + 64);
        h.ht = (float *)(hv + 64 + al64(sizeof(float) * MLP_IN * (size_t)c->hv_cap));
        h.cap = (unsigned long long)c->hv_cap;
        CK(ensure(c->featb, sizeof(float) * MLP_IN * std::max(c->wcap, (size_t)1)));
        h.feat = (float *)c->featb.p;
    }
    h.kmin = c->prm.learned_min_k;
    h.gate = (float)c->prm.learned_gate;
    // from pinned memory, and the counters cleared by a kernel: nothing waits here
    HK(hipMemcpyAsync(base, &h, sizeof(SecCtl), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(ncnt, WGS)), dim3(WGS), 0, c->stream, h.iters, ncnt, 0ULL);
    HK(hipGetLastError());
    c->dsec = (const SecCtl *)base;
    c->dlvlmax = (unsigned int *)(h.iters + 4 * SEC_SLOTS);
    c->cur_level = 0;
    c->sec_pre = (h.w != nullptr) || (h.hf != nullptr);
    return 0;
}

// iteration counters of the last solve into the statistics (stream synchronised)
static int sec_collect(hsdc_ctx *c)
{
    std::vector<unsigned long long> v(4 * SEC_SLOTS);
    std::vector<unsigned int> lm((size_t)SEC_LEVELS * SEC_SLOTS);
    HK(hipMemcpyAsync(v.data(), (char *)c->sec.p + al64(sizeof(SecCtl)), sizeof(unsigned long long) * v.size(),
                      hipMemcpyDeviceToHost, c->stream));
    HK(hipMemcpyAsync(lm.data(), c->dlvlmax, sizeof(unsigned int) * lm.size(), hipMemcpyDeviceToHost, c->stream));
    HK(hipStreamSynchronize(c->stream));
    long long crit = 0, it = 0, us = 0, wi = 0, ev = 0;
    for (int l = 0; l < SEC_LEVELS; l++) {
        unsigned int m = 0;
        for (int q = 0; q < SEC_SLOTS; q++) m = std::max(m, lm[(size_t)l * SEC_SLOTS + q]);
        crit += m;
    }
    for (int q = 0; q < SEC_SLOTS; q++) {
        it += (long long)v[q];
        us += (long long)v[SEC_SLOTS + q];
        wi += (long long)v[2 * SEC_SLOTS + q];
        ev += (long long)v[3 * SEC_SLOTS + q];
    }
    c->st.secular_crit_iters = crit;
    c->st.secular_iters = it;
    c->st.learned_roots = us;
    c->st.secular_wave_iters = wi;
    c->st.secular_evals = ev;
    return 0;
}

// ----------------------------------------------------------------------------
// HSS branch of one DC level (MDLAED3's DHSSEVC, CauchyHssEig.f90:2444-2598):
// build the HSS form of every HSS merge's eigenvector matrix, then
// Ct(:, 1:K) = Q2 * H.  `sel` lists the slice-local merges that take this branch.
// ----------------------------------------------------------------------------
struct HssPlanM {
    int li;              // slice-local merge index
    HssTree T;
    bool ok;
    HssMerge H;          // device pointers
    std::vector<int> usz, vsz, bsz, dsz;
    std::vector<long> pu, pv, pb, pd;
    int err = 0;
    long gcols = 0, fcols = 0;
    std::vector<int> lg, lf;
};

// One step of batched GEMM items (every item nrow <= rmax rows, n <= nmax columns): the
// tile is chosen by the widest item, since a column block that most items leave empty
// costs a pass over A for nothing.  Measured at n = 32768 (256 leaves of 128 or 254 of
// 129, 8.6 GB of X per step):
//   * n <= 64 (leaf upsweeps and siblings, ranks <= 64): 128 x 64 with eight 16 x 64
//     wavefronts, 4.3 / 2.9 ms against k_bgemm's 5.7 / 3.7;
//   * n <= 128: 128 x 128 with eight 32 x 64 wavefronts, 6.9 ms (upsweep, ranks 60-91)
//     and 9.1 ms (leaf products) against 9.3 / 13.9 for k_bgemm's two 64-column blocks;
//   * n <= 144 (leaves of 129-144 columns): 256 x 144 with eight 32 x 144 wavefronts,
//     11.2 ms against 18.0 for two 128-column blocks;
//   * wider: 128 x 128 tiles, the column block fastest when the last one is at least
//     half full (a row panel of A read once), else slowest -- with the column block
//     fastest, blocks that exit at once would take every other workgroup id, and ids go
//     round-robin to the 8 XCDs, so half of them would get no work.
// Every tile runs the k-steps in the same order, so all give the same bits.
static void bgemm_step(const BG *items, unsigned ni, int rmax, int nmax, hipStream_t s)
{
    if (nmax <= 32) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 32, 16, 1, 1, 2>), dim3(cdiv(rmax, 128), ni, 1), dim3(512), 0,
                           s, items, 1);
    } else if (nmax <= 48) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 48, 16, 1, 1, 3>), dim3(cdiv(rmax, 128), ni, 1), dim3(512), 0,
                           s, items, 1);
    } else if (nmax <= 64) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 64, 16, 1, 1, 4>), dim3(cdiv(rmax, 128), ni, 1), dim3(512), 0,
                           s, items, 1);
    } else if (nmax <= 96) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 96, 16, 1, 2, 3>), dim3(cdiv(rmax, 128), ni, 1), dim3(512), 0,
                           s, items, 1);
    } else if (nmax <= 128) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 128, 16, 1, 2, 4>), dim3(cdiv(rmax, 128), ni, 1), dim3(512),
                           0, s, items, 1);
    } else if (nmax <= 144) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<256, 144, 16, 1, 2, 9>), dim3(cdiv(rmax, 256), ni, 1), dim3(512),
                           0, s, items, 1);
    } else {
        const int nz = cdiv(nmax, 128);
        if (nmax - 128 * (nz - 1) >= 64)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 128, 16, 1, 2, 4>),
                               dim3((unsigned)(cdiv(rmax, 128) * nz), ni, 1), dim3(512), 0, s, items, nz);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgemm_w<128, 128, 16, 1, 2, 4>), dim3(cdiv(rmax, 128), ni, nz),
                               dim3(512), 0, s, items, 0);
    }
}

// The product of a batch of HSS matrices whose generators are on the device (P: tree,
// generator directory, arena and leaf blocks of each; a plan that is not ok takes the
// dense Cauchy product of its HssMerge): Y := X * H, or for left H X as (X^T H^T)^T on
// X^T.  fastHssmmL_omp (ConstructHssd.f90:1394-1630) and fastHssmm (:926-1155) as
// batched FP64 MFMA GEMMs, one launch per tree step over every matrix.
static int hss_mult(hsdc_ctx *c, const std::vector<HssSpec> &sp, std::vector<HssPlanM> &P, bool left)
{
    hipStream_t s = c->stream;
    // G / F column layout (fastHssmmL_omp lg / lf, ConstructHssd.f90:1472-1510)
    long gft = 0;
    std::vector<long> gbase(P.size()), fbase(P.size());
    long cbt = 0;
    std::vector<long> cbase(P.size(), -1);
    for (size_t h = 0; h < P.size(); h++) {
        HssPlanM &Q = P[h];
        if (!Q.ok) {
            cbase[h] = cbt;
            cbt += (long)sp[h].K * sp[h].K;
            continue;
        }
        const HssTree &T = Q.T;
        Q.lg.assign(T.ltr + 1, 0);
        Q.lf.assign(T.ltr + 1, 0);
        // X * H: G_i has the width of U_i's rank, F_i of the sibling's B columns; for
        // left (H * X as (X^T H^T)^T) G_i has V_i's rank and F_i its own B's rows
        long it = 0;
        for (int LL = T.lvl; LL >= 1; LL--)
            for (int nd : T.lev[LL]) {
                Q.lg[nd] = (int)it;
                it += left ? Q.vsz[2 * (nd - 1)] : Q.usz[2 * (nd - 1) + 1];
            }
        Q.gcols = it;
        it = 0;
        for (int LL = 1; LL <= T.lvl; LL++)
            for (int nd : T.lev[LL]) {
                int par = T.tr[nd - 1];
                int nj = (T.c1[par] == nd) ? T.c2[par] : T.c1[par];
                Q.lf[nd] = (int)it;
                it += left ? Q.bsz[2 * (nd - 1)] : Q.bsz[2 * (nj - 1) + 1];
            }
        Q.fcols = it;
        gbase[h] = gft;
        gft += (long)sp[h].nrow * Q.gcols;
        fbase[h] = gft;
        gft += (long)sp[h].nrow * Q.fcols;
    }
    CK(ensure(c->gf, sizeof(double) * std::max(gft, 1L)));
    CK(ensure(c->cb, sizeof(double) * std::max(cbt, 1L)));
    double *GF = (double *)c->gf.p, *CB = (double *)c->cb.p;
    // ---- GEMM steps
    std::vector<std::vector<BG>> steps;
    auto bg = [](const double *A, long lda, const double *B, long ldb, double *C, long ldc, int nrow, int n, int k,
                 int beta) {
        BG g; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc; g.nrow = nrow; g.n = n; g.k = k;
        g.beta = beta; g.btrans = 0; g.A2 = nullptr; g.B2 = nullptr; g.lda2 = 0; g.ldb2 = 0; g.k2 = 0; return g;
    };
    // dense fallback merges: C = Cauchy matrix, Y = X * C
    std::vector<BG> dense_items;
    for (size_t h = 0; h < P.size(); h++) {
        if (P[h].ok) continue;
        int K = sp[h].K;
        // the HssMerge of this merge may not be in the device table; upload it alone
        HssMerge H1 = P[h].H;
        CK(ensure(c->tabs, sizeof(HssMerge)));
        HK(hipMemcpyAsync(c->tabs.p, &H1, sizeof(HssMerge), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_cbuild, dim3(std::min(cdiv((long)K * K, WGS), 4096)), dim3(WGS), 0, s,
                           (const HssMerge *)c->tabs.p, 0, CB + cbase[h]);
        HK(hipStreamSynchronize(s));
        BG g = bg(sp[h].X, sp[h].ldx, CB + cbase[h], K, sp[h].Y, sp[h].ldy, sp[h].nrow, K, K, 0);
        g.btrans = left ? 1 : 0;  // X^T C^T
        g.cmap = sp[h].ymap;
        g.cmoff = sp[h].ymoff;
        dense_items.push_back(g);
    }
    if (!dense_items.empty()) steps.push_back(dense_items);
    int mlvl = 0;
    for (auto &Q : P) if (Q.ok) mlvl = std::max(mlvl, Q.T.lvl);
    if (left) {
        // H X = (X^T H^T)^T, with X^T in sp.X (nrow = the columns of X): the transposed
        // generators U' = V^T, V' = U^T, node i's B' = B_i^T (H's block (sibling, i) is
        // U_sib B_sib V_i, so H^T's block (i, sibling) is V_i^T B_sib^T U_sib^T), D' = D^T,
        // each read through btrans (fastHssmm, ConstructHssd.f90:926-1155)
        auto bgt = [&](const double *A, long lda, const double *B, long ldb, double *C, long ldc, int nrow, int n,
                       int k, int beta) {
            BG g = bg(A, lda, B, ldb, C, ldc, nrow, n, k, beta);
            g.btrans = 1;
            return g;
        };
        for (int st = 0; st < mlvl; st++) {
            std::vector<BG> v;
            for (size_t h = 0; h < P.size(); h++) {
                HssPlanM &Q = P[h];
                if (!Q.ok) continue;
                int LL = Q.T.lvl - st;
                if (LL < 1) continue;
                const long nr = sp[h].nrow;
                double *G = GF + gbase[h];
                const double *ar = Q.H.arena;
                for (int nd : Q.T.lev[LL]) {
                    int vr = Q.vsz[2 * (nd - 1)], vc = Q.vsz[2 * (nd - 1) + 1];
                    bool leaf = Q.T.c1[nd] == 0;
                    const double *A = leaf ? sp[h].X + (long)Q.T.leaf_r0[nd] * sp[h].ldx : G + (long)Q.lg[Q.T.c1[nd]] * nr;
                    long lda = leaf ? sp[h].ldx : nr;
                    if (vr > 0) v.push_back(bgt(A, lda, ar + Q.pv[nd - 1], vr, G + (long)Q.lg[nd] * nr, nr, (int)nr, vr, vc, 0));
                }
            }
            if (!v.empty()) steps.push_back(v);
        }
        for (int LL = 1; LL <= mlvl; LL++) {
            std::vector<BG> vs, vd;
            for (size_t h = 0; h < P.size(); h++) {
                HssPlanM &Q = P[h];
                if (!Q.ok || LL > Q.T.lvl) continue;
                const long nr = sp[h].nrow;
                double *G = GF + gbase[h], *F = GF + fbase[h];
                const double *ar = Q.H.arena;
                for (int nd : Q.T.lev[LL]) {
                    int par = Q.T.tr[nd - 1];
                    int nj = (Q.T.c1[par] == nd) ? Q.T.c2[par] : Q.T.c1[par];
                    int br = Q.bsz[2 * (nd - 1)], bc = Q.bsz[2 * (nd - 1) + 1];
                    if (br > 0)
                        vs.push_back(bgt(G + (long)Q.lg[nj] * nr, nr, ar + Q.pb[nd - 1], br, F + (long)Q.lf[nd] * nr, nr,
                                         (int)nr, br, bc, LL >= 2 ? 1 : 0));
                    if (Q.T.c1[nd] != 0) {
                        int ur = Q.usz[2 * (nd - 1)], uc = Q.usz[2 * (nd - 1) + 1];
                        if (ur > 0)
                            vd.push_back(bgt(F + (long)Q.lf[nd] * nr, nr, ar + Q.pu[nd - 1], ur,
                                             F + (long)Q.lf[Q.T.c1[nd]] * nr, nr, (int)nr, ur, uc, 0));
                    }
                }
            }
            if (!vs.empty()) steps.push_back(vs);
            if (!vd.empty()) steps.push_back(vd);
        }
        // leaves: Y_i = F_i U_i^T, then Y_i += X_i D_i^T
        std::vector<BG> v1, v2;
        for (size_t h = 0; h < P.size(); h++) {
            HssPlanM &Q = P[h];
            if (!Q.ok) continue;
            const long nr = sp[h].nrow;
            double *F = GF + fbase[h];
            const double *ar = Q.H.arena;
            for (int nd = 1; nd <= Q.T.ltr - 1; nd++) {
                if (Q.T.c1[nd] != 0) continue;
                int ur = Q.usz[2 * (nd - 1)], uc = Q.usz[2 * (nd - 1) + 1], md = Q.dsz[nd - 1];
                long lx = Q.T.leaf_r0[nd];
                v1.push_back(bgt(F + (long)Q.lf[nd] * nr, nr, ar + Q.pu[nd - 1], std::max(ur, 1),
                                 sp[h].Y + lx * sp[h].ldy, sp[h].ldy, (int)nr, md, uc, 0));
                v2.push_back(bgt(sp[h].X + lx * sp[h].ldx, sp[h].ldx, Q.H.dd + Q.pd[nd - 1], md,
                                 sp[h].Y + lx * sp[h].ldy, sp[h].ldy, (int)nr, md, md, 1));
            }
        }
        if (!v1.empty()) steps.push_back(v1);
        if (!v2.empty()) steps.push_back(v2);
    }
    // upsweep
    for (int st = 0; st < mlvl && !left; st++) {
        std::vector<BG> v;
        for (size_t h = 0; h < P.size(); h++) {
            HssPlanM &Q = P[h];
            if (!Q.ok) continue;
            int LL = Q.T.lvl - st;
            if (LL < 1) continue;
            const long nr = sp[h].nrow;
            double *G = GF + gbase[h];
            const double *ar = Q.H.arena;
            for (int nd : Q.T.lev[LL]) {
                int ur = Q.usz[2 * (nd - 1)], uc = Q.usz[2 * (nd - 1) + 1];
                bool leaf = Q.T.c1[nd] == 0;
                const double *A = leaf ? sp[h].X + (long)Q.T.leaf_r0[nd] * sp[h].ldx : G + (long)Q.lg[Q.T.c1[nd]] * nr;
                long lda = leaf ? sp[h].ldx : nr;
                if (uc > 0) v.push_back(bg(A, lda, ar + Q.pu[nd - 1], ur, G + (long)Q.lg[nd] * nr, nr, (int)nr, uc, ur, 0));
            }
        }
        if (!v.empty()) steps.push_back(v);
    }
    // siblings and downsweep, level by level from the root
    for (int LL = 1; LL <= mlvl && !left; LL++) {
        std::vector<BG> vs, vd;
        for (size_t h = 0; h < P.size(); h++) {
            HssPlanM &Q = P[h];
            if (!Q.ok || LL > Q.T.lvl) continue;
            const long nr = sp[h].nrow;
            double *G = GF + gbase[h], *F = GF + fbase[h];
            const double *ar = Q.H.arena;
            for (int nd : Q.T.lev[LL]) {
                int par = Q.T.tr[nd - 1];
                int nj = (Q.T.c1[par] == nd) ? Q.T.c2[par] : Q.T.c1[par];
                int br = Q.bsz[2 * (nj - 1)], bc = Q.bsz[2 * (nj - 1) + 1];
                if (bc > 0)
                    vs.push_back(bg(G + (long)Q.lg[nj] * nr, nr, ar + Q.pb[nj - 1], std::max(br, 1),
                                    F + (long)Q.lf[nd] * nr, nr, (int)nr, bc, br, LL >= 2 ? 1 : 0));
                if (Q.T.c1[nd] != 0) {
                    int vr = Q.vsz[2 * (nd - 1)], vc = Q.vsz[2 * (nd - 1) + 1];
                    if (vc > 0)
                        vd.push_back(bg(F + (long)Q.lf[nd] * nr, nr, ar + Q.pv[nd - 1], std::max(vr, 1),
                                        F + (long)Q.lf[Q.T.c1[nd]] * nr, nr, (int)nr, vc, vr, 0));
                }
            }
        }
        if (!vs.empty()) steps.push_back(vs);
        if (!vd.empty()) steps.push_back(vd);
    }
    // leaves: Y_i = F_i V_i + X_i D_i, both products in one GEMM item
    if (!left) {
        std::vector<BG> v1;
        for (size_t h = 0; h < P.size(); h++) {
            HssPlanM &Q = P[h];
            if (!Q.ok) continue;
            const long nr = sp[h].nrow;
            double *F = GF + fbase[h];
            double *Y = sp[h].Y;
            const double *ar = Q.H.arena;
            for (int nd = 1; nd <= Q.T.ltr - 1; nd++) {
                if (Q.T.c1[nd] != 0) continue;
                int vr = Q.vsz[2 * (nd - 1)], vc = Q.vsz[2 * (nd - 1) + 1], md = Q.dsz[nd - 1];
                long lx = Q.T.leaf_r0[nd];
                // V_i is vr x vc with vc = md (the leaf's columns)
                BG g = bg(F + (long)Q.lf[nd] * nr, nr, ar + Q.pv[nd - 1], std::max(vr, 1),
                          sp[h].ymap ? Y : Y + lx * sp[h].ldy, sp[h].ldy, (int)nr, md, (vc > 0) ? vr : 0, 0);
                if (sp[h].ymap) {
                    g.cmap = sp[h].ymap + lx;
                    g.cmoff = sp[h].ymoff;
                }
                g.A2 = sp[h].X + lx * sp[h].ldx; g.lda2 = sp[h].ldx;
                g.B2 = Q.H.dd + Q.pd[nd - 1]; g.ldb2 = md; g.k2 = md;
                v1.push_back(g);
            }
        }
        if (!v1.empty()) steps.push_back(v1);
    }
    size_t nitems = 0;
    for (auto &v : steps) nitems += v.size();
    CK(ensure(c->tabs, sizeof(BG) * std::max(nitems, (size_t)1)));
    {
        // staged in pinned memory; the host waits only until the previous multiply's
        // table has left it (callers that multiply column block after column block do
        // not synchronise in between)
        if (c->pinb_ev) HK(hipEventSynchronize(c->pinb_ev));
        CK(ensure_pin(c->pinb, sizeof(BG) * std::max(nitems, (size_t)1)));
        BG *all = (BG *)c->pinb.p;
        size_t o = 0;
        for (auto &v : steps) {
            memcpy(all + o, v.data(), sizeof(BG) * v.size());
            o += v.size();
        }
        HK(hipMemcpyAsync(c->tabs.p, all, sizeof(BG) * nitems, hipMemcpyHostToDevice, s));
        if (!c->pinb_ev) HK(hipEventCreateWithFlags(&c->pinb_ev, hipEventDisableTiming));
        HK(hipEventRecord(c->pinb_ev, s));
    }
    // flops of the multiply (2 nrow n k per GEMM: fastHssmmL's DGEMM calls exactly)
    for (auto &v : steps)
        for (auto &g : v) c->st.hss_mm_flops += 2LL * g.nrow * g.n * (g.k + g.k2);
    hipEvent_t m0 = nullptr, m1 = nullptr;
    if (c->prm.profile) {
        HK(hipEventCreate(&m0));
        HK(hipEventCreate(&m1));
        HK(hipEventRecord(m0, s));
    }
    size_t so = 0;
    for (auto &v : steps) {
        int rmax = 0, nmax = 0;
        for (auto &g : v) { rmax = std::max(rmax, g.nrow); nmax = std::max(nmax, g.n); }
        bgemm_step((const BG *)c->tabs.p + so, (unsigned)v.size(), rmax, nmax, s);
        so += v.size();
    }
    HK(hipGetLastError());
    if (c->prm.profile) HK(hipEventRecord(m1, s));
    // the tables above are host vectors copied synchronously; wait before they go
    HK(hipStreamSynchronize(s));
    if (c->prm.profile) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, m0, m1);
        c->st.ms_hss_mm += ms;
        (void)hipEventDestroy(m0);
        (void)hipEventDestroy(m1);
    }
    return 0;
}

static int hss_apply(hsdc_ctx *c, const std::vector<HssSpec> &sp, int ni, double tol,
                     std::vector<HssPlanM> *keep, bool left = false)
{
    hipStream_t s = c->stream;
    std::vector<HssPlanM> P(sp.size());
    // one batch is all eigenvector matrices or all SVD vector matrices
    const int sq = sp.empty() ? 0 : sp[0].sq;
    for (const HssSpec &q : sp)
        if (q.sq != sq) return HSDC_ERR_INTERNAL;
    // ---- device memory for every merge: the generator directories (node ranges,
    // generator shapes and offsets, counters) of all merges side by side at the front,
    // cleared by one fill and read back by one copy; then each merge's buffers
    size_t tot = 0, mtot = 0;
    std::vector<size_t> base(sp.size()), mbase(sp.size());
    for (size_t h = 0; h < sp.size(); h++) {
        P[h].li = (int)h;
        int K = sp[h].K;
        P[h].ok = hss_make_tree(P[h].T, K, ni);
        int ltr = std::max(P[h].T.ltr, 1);
        long dds = 0;
        for (int x : P[h].T.m) dds += (long)x * x;
        mbase[h] = mtot;
        mtot += al64(sizeof(int) * (11 * ltr + 8)) + al64(sizeof(long) * (4 * ltr)) +
                al64(sizeof(unsigned long long) * 3);
        size_t o = al64(6 * sizeof(double) * K) + al64(2 * sizeof(int) * K);
        o += al64(sizeof(double) * std::max(dds, 1L));
        o += al64(sizeof(double) * (size_t)12 * std::max(ni, 128) * K);  // arena
        o += al64(sizeof(double) * (size_t)8 * K * std::max(P[h].T.lm, 1));  // scratch
        base[h] = tot;
        tot += o;
    }
    mtot = al64(std::max(mtot, (size_t)64));
    CK(ensure(c->hss, mtot + tot));
    char *hb = (char *)c->hss.p;
    for (size_t h = 0; h < sp.size(); h++) {
        HssPlanM &Q = P[h];
        int K = sp[h].K, ltr = std::max(Q.T.ltr, 1);
        char *p = hb + mtot + base[h];
        HssMerge &H = Q.H;
        H.FF = sp[h].FF; H.D0 = sp[h].D0; H.U0 = sp[h].U0; H.V0 = sp[h].V0;
        H.DIFL = sp[h].DIFL; H.DIFR = sp[h].DIFR;
        double *dp = (double *)p;
        H.D = dp; H.U = dp + K; H.V = dp + 2 * K; H.T = dp + 3 * K;
        p += al64(6 * sizeof(double) * K);
        int *ip = (int *)p;
        H.PL = ip; H.PU = ip + K;
        p += al64(2 * sizeof(int) * K);
        char *mp = hb + mbase[h];
        ip = (int *)mp;
        H.lr = ip; H.lc = ip + 2 * ltr; H.usz = ip + 4 * ltr; H.vsz = ip + 6 * ltr; H.bsz = ip + 8 * ltr;
        H.dsz = ip + 10 * ltr; H.lda = ip + 11 * ltr; H.lca = ip + 11 * ltr + 1; H.err = ip + 11 * ltr + 2;
        mp += al64(sizeof(int) * (11 * ltr + 8));
        long *lp = (long *)mp;
        H.pu = lp; H.pv = lp + ltr; H.pb = lp + 2 * ltr; H.pd = lp + 3 * ltr;
        mp += al64(sizeof(long) * (4 * ltr));
        H.arena_top = (unsigned long long *)mp; H.scratch_top = H.arena_top + 1; H.flops = H.arena_top + 2;
        H.dd = (double *)p;
        long dds = 0;
        for (int x : Q.T.m) dds += (long)x * x;
        p += al64(sizeof(double) * std::max(dds, 1L));
        H.arena = (double *)p; H.arena_cap = (long)12 * std::max(ni, 128) * K;
        p += al64(sizeof(double) * (size_t)12 * std::max(ni, 128) * K);
        H.scratch = (double *)p; H.scratch_cap = (long)8 * K * std::max(Q.T.lm, 1);
        H.K = K; H.ltr = Q.T.ltr; H.lvl = Q.T.lvl; H.sq = sp[h].sq;
    }
    // ---- host tables: merges, leaves (with leaf lr/lc and D offsets), level steps
    std::vector<HssMerge> hh;
    std::vector<int> hmap;  // table index -> plan index
    for (size_t h = 0; h < P.size(); h++) if (P[h].ok) { hmap.push_back((int)h); hh.push_back(P[h].H); }
    const int nh = (int)hh.size();
    int maxlvl = 0;
    for (int t = 0; t < nh; t++) maxlvl = std::max(maxlvl, P[hmap[t]].T.lvl);
    std::vector<HssNode> nodes;   // leaves first, then per step
    std::vector<int> init_lrlc;   // per merge: 2*ltr lr then 2*ltr lc (leaves set, others 0)
    std::vector<long> init_pd;
    size_t leaf0 = 0, nleaf = 0;
    std::vector<size_t> step_off(maxlvl + 1), step_cnt(maxlvl + 1);
    std::vector<HssCompact> comps;
    std::vector<int> order;
    std::vector<size_t> comp_off(maxlvl + 1), comp_cnt(maxlvl + 1);
    for (int t = 0; t < nh; t++) {
        const HssTree &T = P[hmap[t]].T;
        long pdt = 0;
        int it = 0;
        for (int i = 1; i <= T.ltr; i++)
            if (T.c1[i] == 0) {
                HssNode nd;
                nd.mrg = t; nd.node = i; nd.c1 = 0; nd.c2 = 0; nd.sib = 0;
                nd.leaf_r0 = T.leaf_r0[i]; nd.leaf_m = T.m[it++];
                nodes.push_back(nd);
            }
        (void)pdt;
    }
    nleaf = nodes.size();
    for (int st = 0; st < maxlvl; st++) {
        step_off[st] = nodes.size();
        comp_off[st] = comps.size();
        for (int t = 0; t < nh; t++) {
            const HssTree &T = P[hmap[t]].T;
            int LL = T.lvl - st;
            if (LL < 1) continue;
            for (int nd_id : T.lev[LL]) {
                HssNode nd;
                nd.mrg = t; nd.node = nd_id; nd.c1 = T.c1[nd_id]; nd.c2 = T.c2[nd_id];
                int par = T.tr[nd_id - 1];
                nd.sib = (T.c1[par] == nd_id) ? T.c2[par] : T.c1[par];
                nd.leaf_r0 = T.leaf_r0[nd_id]; nd.leaf_m = 0;
                nodes.push_back(nd);
            }
            HssCompact cp;
            cp.mrg = t; cp.start = (int)order.size(); cp.nlevel = (int)T.lev[LL].size();
            for (int nd_id : T.lev[LL]) order.push_back(nd_id);
            for (int LLt = LL - 1; LLt >= 1; LLt--)
                for (int nd_id : T.lev[LLt]) if (T.c1[nd_id] == 0) order.push_back(nd_id);
            cp.count = (int)order.size() - cp.start;
            comps.push_back(cp);
        }
        step_cnt[st] = nodes.size() - step_off[st];
        comp_cnt[st] = comps.size() - comp_off[st];
    }
    // The directory area's first state (zeros, with every merge's leaf ranges and D
    // offsets) and the merge / node / compaction tables, staged in pinned memory: two
    // copies and no wait (the pinned area is next written by this call's readback, after
    // the stream has passed both copies)
    size_t t0 = 0, t1 = al64(t0 + sizeof(HssMerge) * std::max(nh, 1)),
           t2 = al64(t1 + sizeof(HssNode) * std::max(nodes.size(), (size_t)1)),
           t3 = al64(t2 + sizeof(HssCompact) * std::max(comps.size(), (size_t)1)),
           t4 = al64(t3 + sizeof(int) * std::max(order.size(), (size_t)1));
    CK(ensure(c->htab, t4));
    CK(ensure_pin(c->pinh, mtot + t4));
    {
        char *img = (char *)c->pinh.p;
        memset(img, 0, mtot);
        for (int t = 0; t < nh; t++) {
            const HssTree &T = P[hmap[t]].T;
            const int ltr1 = std::max(T.ltr, 1);
            int *lrlc = (int *)(img + mbase[hmap[t]]);
            long *pd = (long *)(img + mbase[hmap[t]] + al64(sizeof(int) * (11 * ltr1 + 8))) + 3 * ltr1;
            long pdt = 0;
            int it = 0;
            for (int i = 1; i <= T.ltr; i++)
                if (T.c1[i] == 0) {
                    lrlc[2 * (i - 1)] = T.leaf_r0[i] + 1;
                    lrlc[2 * (i - 1) + 1] = T.leaf_r0[i] + T.m[it];
                    pd[i - 1] = pdt;
                    pdt += (long)T.m[it] * T.m[it];
                    it++;
                }
            for (int i = 0; i < 2 * T.ltr; i++) lrlc[2 * T.ltr + i] = lrlc[i];
        }
        char *hbuf = img + mtot;
        memcpy(hbuf + t0, hh.data(), sizeof(HssMerge) * nh);
        memcpy(hbuf + t1, nodes.data(), sizeof(HssNode) * nodes.size());
        if (!comps.empty()) memcpy(hbuf + t2, comps.data(), sizeof(HssCompact) * comps.size());
        if (!order.empty()) memcpy(hbuf + t3, order.data(), sizeof(int) * order.size());
        HK(hipMemcpyAsync(hb, img, mtot, hipMemcpyHostToDevice, s));
        HK(hipMemcpyAsync(c->htab.p, hbuf, t4, hipMemcpyHostToDevice, s));
    }
    char *tb = (char *)c->htab.p;
    const HssMerge *dH = (const HssMerge *)(tb + t0);
    const HssNode *dN = (const HssNode *)(tb + t1);
    const HssCompact *dC = (const HssCompact *)(tb + t2);
    const int *dO = (const int *)(tb + t3);
    hipEvent_t b0 = nullptr, b1 = nullptr;
    if (c->prm.profile && nh > 0) {
        HK(hipEventCreate(&b0));
        HK(hipEventCreate(&b1));
        HK(hipEventRecord(b0, s));
    }
    if (nh > 0) {
        int kmax = 0;
        for (auto &H : hh) kmax = std::max(kmax, H.K);
        hipLaunchKernelGGL(k_hss_init, dim3(std::min(cdiv(kmax, WGS), 256), nh), dim3(WGS), 0, s, dH);
        hipLaunchKernelGGL(k_hss_dgen, dim3(nleaf), dim3(WGS), 0, s, dH, dN + leaf0, (int)nleaf);
        // the 512-thread RRLU launches (fewer nodes than two per CU) keep each node's
        // private column data in LDS as far as it fits: 160 KB per workgroup on gfx950,
        // 30.6 KB of it the kernels' static arrays.  A launch of at most one node per CU
        // takes the whole rest; more nodes than CUs take what the largest merge's
        // generator column needs, so that two workgroups can share a CU
        for (int st = 0; st < maxlvl; st++) {
            int nn = (int)step_cnt[st], nc = (int)comp_cnt[st];
            const int ldsn = nn <= 256 ? HSS_LDS_V : std::min(kmax, HSS_LDS_V);
            const size_t ldsb = sizeof(double) * (size_t)ldsn;
            const HssNode *sn = dN + step_off[st];
            const HssCompact *sc = dC + comp_off[st];
            hipLaunchKernelGGL(k_hss_reset, dim3(nh), dim3(64), 0, s, dH);
            if (sq && nn < 512)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_rowcomp<512, true>), dim3(nn), dim3(512), ldsb, s, dH, sn, nn,
                                   tol, ldsn);
            else if (sq)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_rowcomp<256, true>), dim3(nn), dim3(256), 0, s, dH, sn, nn, tol,
                                   0);
            else if (nn < 512)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_rowcomp<512, false>), dim3(nn), dim3(512), ldsb, s, dH, sn, nn,
                                   tol, ldsn);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_rowcomp<256, false>), dim3(nn), dim3(256), 0, s, dH, sn, nn, tol,
                                   0);
            hipLaunchKernelGGL(k_hss_compact, dim3(nc), dim3(WGS), 0, s, dH, sc, dO, 1);
            hipLaunchKernelGGL(k_hss_reset, dim3(nh), dim3(64), 0, s, dH);
            if (sq && nn < 512)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_colcomp<512, true>), dim3(nn), dim3(512), ldsb, s, dH, sn, nn,
                                   tol, ldsn);
            else if (sq)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_colcomp<256, true>), dim3(nn), dim3(256), 0, s, dH, sn, nn, tol,
                                   0);
            else if (nn < 512)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_colcomp<512, false>), dim3(nn), dim3(512), ldsb, s, dH, sn, nn,
                                   tol, ldsn);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_hss_colcomp<256, false>), dim3(nn), dim3(256), 0, s, dH, sn, nn, tol,
                                   0);
            hipLaunchKernelGGL(k_hss_compact, dim3(nc), dim3(WGS), 0, s, dH, sc, dO, 0);
            hipLaunchKernelGGL(k_hss_bgen, dim3(nn), dim3(WGS), 0, s, dH, sn, nn);
        }
        HK(hipGetLastError());
    }
    if (b1) HK(hipEventRecord(b1, s));
    // ---- read back the generator directories (one copy for every merge)
    CK(ensure_pin(c->pinh, mtot));
    HK(hipMemcpyAsync(c->pinh.p, hb, mtot, hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    for (int t = 0; t < nh; t++) {
        HssPlanM &Q = P[hmap[t]];
        int ltr = Q.T.ltr;
        const char *mh = (const char *)c->pinh.p + mbase[hmap[t]];
        std::vector<int> meta((const int *)mh, (const int *)mh + 11 * ltr + 8);
        mh += al64(sizeof(int) * (11 * std::max(ltr, 1) + 8));
        std::vector<long> offs((const long *)mh, (const long *)mh + 4 * ltr);
        mh += al64(sizeof(long) * (4 * std::max(ltr, 1)));
        const unsigned long long nfl = ((const unsigned long long *)mh)[2];
        c->st.hss_nflops += (long long)nfl;
        Q.usz.assign(meta.begin() + 4 * ltr, meta.begin() + 6 * ltr);
        Q.vsz.assign(meta.begin() + 6 * ltr, meta.begin() + 8 * ltr);
        Q.bsz.assign(meta.begin() + 8 * ltr, meta.begin() + 10 * ltr);
        Q.dsz.assign(meta.begin() + 10 * ltr, meta.begin() + 11 * ltr);
        Q.err = meta[11 * ltr + 2];
        Q.pu.assign(offs.begin(), offs.begin() + ltr);
        Q.pv.assign(offs.begin() + ltr, offs.begin() + 2 * ltr);
        Q.pb.assign(offs.begin() + 2 * ltr, offs.begin() + 3 * ltr);
        Q.pd.assign(offs.begin() + 3 * ltr, offs.begin() + 4 * ltr);
        if (Q.err) Q.ok = false;
        for (int i = 0; i < ltr - 1; i++) {
            c->st.rank_max = std::max(c->st.rank_max, Q.usz[2 * i + 1]);
            c->st.rank_max = std::max(c->st.rank_max, Q.vsz[2 * i]);
        }
    }
    if (b0) {
        float ms = 0;
        HK(hipStreamSynchronize(s));
        if (hipEventElapsedTime(&ms, b0, b1) == hipSuccess) c->st.ms_hss_build += ms;
        (void)hipGetLastError();
        (void)hipEventDestroy(b0);
        (void)hipEventDestroy(b1);
    }
    CK(hss_mult(c, sp, P, left));
    if (keep) *keep = std::move(P);
    return 0;
}

// HSS branch of the merges `sel` of one DC level: X = Q2 gathered from the
// eigenvector slots, then X * H written straight into the slots of the merge's new
// vectors (vm[cur ^ 1], which k_newcols has set).
static int run_hss(hsdc_ctx *c, const Merge *dmg, const std::vector<Merge> &hm, int m0,
                   const std::vector<MergeOut> &ho, const std::vector<int> &sel, Prob *dprobs, int cur, int gbase)
{
    hipStream_t s = c->stream;
    Work W = c->W;
    std::vector<int> hidx(sel.size()), xcol(sel.size() + 1, 0);
    std::vector<long> xoff(sel.size());
    long xt = 0;
    for (size_t h = 0; h < sel.size(); h++) {
        const Merge &M = hm[m0 + sel[h]];
        hidx[h] = sel[h];
        xoff[h] = xt;
        xt += (long)M.n * ho[sel[h]].k;
        xcol[h + 1] = xcol[h] + ho[sel[h]].k;
    }
    CK(ensure(c->xb, sizeof(double) * std::max(xt, 1L)));
    double *X = (double *)c->xb.p;
    {
        size_t u0 = 0, u1 = al64(sizeof(int) * sel.size()), u2 = al64(u1 + sizeof(int) * (sel.size() + 1)),
               u3 = al64(u2 + sizeof(long) * sel.size());
        std::vector<char> buf(u3);
        memcpy(buf.data() + u0, hidx.data(), sizeof(int) * sel.size());
        memcpy(buf.data() + u1, xcol.data(), sizeof(int) * (sel.size() + 1));
        memcpy(buf.data() + u2, xoff.data(), sizeof(long) * sel.size());
        CK(ensure(c->misc, u3));
        HK(hipMemcpyAsync(c->misc.p, buf.data(), u3, hipMemcpyHostToDevice, s));
        HK(hipStreamSynchronize(s));
        char *mp = (char *)c->misc.p;
        hipLaunchKernelGGL(k_gatherx, dim3(xcol[sel.size()]), dim3(WGS), 0, s, dmg, (const int *)(mp + u0),
                           (const int *)(mp + u1), (int)sel.size(), dprobs, W, cur, X, (const long *)(mp + u2));
        HK(hipGetLastError());
    }
    std::vector<HssSpec> sp(sel.size());
    for (size_t h = 0; h < sel.size(); h++) {
        const Merge &M = hm[m0 + sel[h]];
        HssSpec &q = sp[h];
        q.K = ho[sel[h]].k;
        q.FF = W.dlamda + M.s; q.D0 = W.lam + M.s; q.U0 = W.wnew + M.s; q.V0 = W.alpha + M.s;
        q.DIFL = W.difl + M.s; q.DIFR = W.difr + M.s;
        q.X = X + xoff[h]; q.ldx = M.n; q.nrow = M.n;
        // new vector v of the merge into its slot vm[cur ^ 1][s + v] (set by k_newcols)
        const Prob &PP = c->hprobs[M.prob];
        q.Y = PP.q + (M.s - PP.off); q.ldy = PP.ld;
        q.ymap = W.vm[cur ^ 1] + M.s; q.ymoff = PP.off;
    }
    std::vector<HssPlanM> kept;
    CK(hss_apply(c, sp, c->prm.hss_leaf, c->prm.hss_tol, &kept));
    // largest generator rank of each merge into the trace (0: dense Cauchy product)
    for (size_t h = 0; h < sel.size(); h++) {
        const size_t gm = (size_t)gbase + m0 + sel[h];
        if (gm >= c->trc_rank.size()) continue;
        int r = 0;
        const HssPlanM &Q = kept[h];
        if (Q.ok)
            for (int i = 0; i < Q.T.ltr; i++) r = std::max(r, std::max(Q.usz[2 * i + 1], Q.vsz[2 * i]));
        c->trc_rank[gm] = r;
    }
    return 0;
}

// one level (or one merge of it): merges [m0, m1) of table dm (device) / hm (host);
// dchunkN is the device chunk table of exactly this slice (first entry 0)
static int run_merges(hsdc_ctx *c, const Merge *dm, const std::vector<Merge> &hm, int m0, int m1,
                      const int *dchunkN, int totN, Prob *dprobs, int &cur, int *dinfo, int gbase)
{
    const int nm = m1 - m0;
    hipStream_t s = c->stream;
    const Merge *dmg = dm + m0;
    size_t need_mo = sizeof(MergeOut) * nm;
    CK(ensure(c->mo, need_mo));
    MergeOut *dmo = (MergeOut *)c->mo.p;
    Work W = c->W;
    hipLaunchKernelGGL(k_prep, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, dprobs, W, cur);
    hipLaunchKernelGGL(k_mergepath, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, W);
    // sorted D of merges up to 4096 in LDS for the deflated-region replay
    int nmax = 0;
    for (int i = m0; i < m1; i++) nmax = std::max(nmax, hm[i].n);
    const int ds_cap = nmax <= 4096 ? nmax : 0;
    // large merges get 16 wavefronts: their O(N) passes are latency bound
    if (nmax >= 2048)
        hipLaunchKernelGGL(k_deflate<1024>, dim3(nm), dim3(1024), sizeof(double) * ds_cap, s, dmg, W, cur, dmo,
                           c->prm.hss_min_k, ds_cap);
    else
        hipLaunchKernelGGL(k_deflate<256>, dim3(nm), dim3(WGS), sizeof(double) * ds_cap, s, dmg, W, cur, dmo,
                           c->prm.hss_min_k, ds_cap);
    // K of every merge back to the host
    CK(ensure_pin(c->pin2, need_mo));
    HK(hipMemcpyAsync(c->pin2.p, dmo, need_mo, hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    std::vector<MergeOut> ho(nm);
    memcpy(ho.data(), c->pin2.p, need_mo);
    // derived tables
    std::vector<int> chunkK(nm + 1), blockK(nm + 1), colK(nm + 1);
    std::vector<long> soff(nm), coff(nm);
    std::vector<GTile> tiles;
    long stot = 0, ctot = 0;
    int anyrot = 0, anyk = 0;
    chunkK[0] = blockK[0] = colK[0] = 0;
    for (int i = 0; i < nm; i++) {
        const MergeOut &o = ho[i];
        const Merge &M = hm[m0 + i];
        int k = o.flags ? 0 : o.k;
        chunkK[i + 1] = chunkK[i] + cdiv(k, WGS);
        blockK[i + 1] = blockK[i] + (k > 0 ? cdiv(k, SEC_WG / sec_team(k)) : 0);
        colK[i + 1] = colK[i] + k;
        anyrot |= (o.nrot > 0 && !o.flags);
        anyk |= (k > 0);
        soff[i] = stot;
        coff[i] = ctot;
        if (k > 0 && !o.hss) stot += (long)k * k;
        if (!o.hss) ctot += (long)M.n * k;  // HSS merges write their vectors into the slots
        c->st.roots += k;
        c->st.deflated += M.n - k;
        if (o.hss) c->st.hss_merges++;
        if (!o.hss && k > 0) {
            int n1 = M.n1, n2 = M.n - M.n1;
            for (int r0 = 0; r0 < n1; r0 += 64)
                for (int c0 = 0; c0 < k; c0 += 64) tiles.push_back(GTile{i, 0, r0, c0});
            for (int r0 = 0; r0 < n2; r0 += 64)
                for (int c0 = 0; c0 < k; c0 += 64) tiles.push_back(GTile{i, 1, r0, c0});
        }
        c->st.merges++;
        c->st.k_top = k;
    }
    // upload derived tables in one copy
    size_t o1 = 0, o2 = al64(o1 + sizeof(int) * (nm + 1)), o3 = al64(o2 + sizeof(int) * (nm + 1)),
           o4 = al64(o3 + sizeof(int) * (nm + 1)), o5 = al64(o4 + sizeof(long) * nm),
           o6 = al64(o5 + sizeof(long) * nm), o7 = al64(o6 + sizeof(GTile) * tiles.size());
    CK(ensure_pin(c->pin, o7));
    char *hp = (char *)c->pin.p;
    memcpy(hp + o1, chunkK.data(), sizeof(int) * (nm + 1));
    memcpy(hp + o2, blockK.data(), sizeof(int) * (nm + 1));
    memcpy(hp + o3, colK.data(), sizeof(int) * (nm + 1));
    memcpy(hp + o4, soff.data(), sizeof(long) * nm);
    memcpy(hp + o5, coff.data(), sizeof(long) * nm);
    if (!tiles.empty()) memcpy(hp + o6, tiles.data(), sizeof(GTile) * tiles.size());
    CK(ensure(c->lvl, o7));
    char *dp = (char *)c->lvl.p;
    HK(hipMemcpyAsync(dp, hp, o7, hipMemcpyHostToDevice, s));
    int *dchunkK = (int *)(dp + o1), *dblockK = (int *)(dp + o2), *dcolK = (int *)(dp + o3);
    long *dsoff = (long *)(dp + o4), *dcoff = (long *)(dp + o5);
    GTile *dtiles = (GTile *)(dp + o6);
    CK(ensure(c->sb, sizeof(double) * (size_t)std::max(stot, 1L)));
    CK(ensure(c->ct, sizeof(double) * (size_t)std::max(ctot, 1L)));
    if (anyrot) {
        hipLaunchKernelGGL(k_rot_prep, dim3(nm), dim3(WGS), 0, s, dmg, dprobs, W, cur, dmo);
        hipLaunchKernelGGL(k_rotate, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, dprobs, W, cur, dmo);
    }
    if (anyk) {
        prof_mark(c, 0);
        if (c->sec_pre)
            hipLaunchKernelGGL(k_secstart, dim3(blockK[nm]), dim3(SEC_WG), 0, s, dmg, dblockK, nm,
                               (const MergeOut *)dmo, W, c->dsec);
        hipLaunchKernelGGL(k_secular, dim3(blockK[nm]), dim3(SEC_WG), 0, s, dmg, dblockK, nm, dmo, W,
                           c->prm.maxit, dinfo, c->dsec,
                           c->dlvlmax + SEC_SLOTS * std::min(c->cur_level, SEC_LEVELS - 1));
        prof_mark(c, -1);
        prof_mark(c, 1);
        hipLaunchKernelGGL(k_loewner, dim3(chunkK[nm]), dim3(WGS), 0, s, dmg, dchunkK, nm, dmo, W);
        prof_mark(c, -1);
        hipLaunchKernelGGL(k_sbuild, dim3(chunkK[nm]), dim3(WGS), 0, s, dmg, dchunkK, nm, dmo, dsoff, W,
                           (double *)c->sb.p);
        if (!tiles.empty())
            hipLaunchKernelGGL(k_gemm, dim3(tiles.size()), dim3(WGS), 0, s, dmg, dtiles, (int)tiles.size(),
                               (const int *)nullptr, nm, dprobs, dmo, dsoff, dcoff, W, cur, (const double *)c->sb.p,
                               (double *)c->ct.p);
    }
    hipLaunchKernelGGL(k_newcols, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, dmo, W, cur);
    if (anyk) {
        std::vector<int> sel;
        for (int i = 0; i < nm; i++)
            if (!ho[i].flags && ho[i].hss && ho[i].k > 0) sel.push_back(i);
        if (!sel.empty()) {  // after k_newcols: the products go into the slots it assigned
            prof_mark(c, 1);
            hipLaunchKernelGGL(k_alpha, dim3(chunkK[nm]), dim3(WGS), 0, s, dmg, dchunkK, nm, dmo, W);
            prof_mark(c, -1);
            CK(run_hss(c, dmg, hm, m0, ho, sel, dprobs, cur, gbase));
        }
    }
    if (anyk)
        hipLaunchKernelGGL(k_scatter, dim3(colK[nm]), dim3(WGS), 0, s, dmg, dcolK, nm, dmo, dcoff, dprobs, W,
                           cur ^ 1, (const double *)c->ct.p, 0);
    hipLaunchKernelGGL(k_lamrg, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, (const MergeOut *)dmo, W, cur ^ 1,
                       c->dtrc ? c->dtrc + 3 * ((long)gbase + m0) : nullptr);
    HK(hipGetLastError());
    cur ^= 1;
    return 0;
}

// ----------------------------------------------------------------------------
// Device schedule of a level: launch sizes come from the merge orders alone, the
// level's kernels read the tables k_lvltab derives from the deflation results,
// and nothing returns to the host until the end of the solve.  The K-sized
// launches stride over their work items with at most DEV_GRID blocks.  A level
// whose dense S and Ct blocks could exceed DEV_BUDGET bytes keeps the
// synchronised schedule.
// ----------------------------------------------------------------------------
#define DEV_GRID 2048
#define DEV_BUDGET (16LL << 30)

struct LvlBound {
    long blocks, tiles, cols;  // secular blocks, GEMM tiles, columns with K = N
    long sbcap, ctcap;         // S and Ct doubles of the dense merges (K < hss_min_k)
    long scat;                 // copy chunks of the dense merges
};

static LvlBound level_bound(const std::vector<Merge> &hm, int hss_min_k)
{
    LvlBound b = {0, 0, 0, 0, 0, 0};
    const long kd = hss_min_k >= 1 ? hss_min_k - 1 : 0;  // largest dense K
    for (const Merge &M : hm) {
        const long kmax = std::min((long)M.n, kd);
        b.blocks += cdiv(M.n, SEC_WG / sec_team(M.n));  // nondecreasing in K
        b.tiles += (long)(cdiv(M.n1, 64) + cdiv(M.n - M.n1, 64)) * cdiv(kmax, 64);
        b.cols += M.n;
        b.sbcap += kmax * kmax;
        b.ctcap += (long)M.n * kmax;
        b.scat += ((long)M.n * kmax + SC_CH - 1) / SC_CH;
    }
    return b;
}

static inline int dev_grid(long n) { return (int)std::max(1L, std::min(n, (long)DEV_GRID)); }

static LvlTab level_tables(hsdc_ctx *c, int nm)
{
    LvlTab t;
    char *p = (char *)c->ltab.p;
    size_t ni = al64(sizeof(int) * (nm + 1)), nl = al64(sizeof(long) * nm);
    t.chunkK = (int *)p; t.blockK = (int *)(p + ni); t.colK = (int *)(p + 2 * ni); t.tileK = (int *)(p + 3 * ni);
    t.scK = (int *)(p + 4 * ni);
    t.soff = (long *)(p + 5 * ni); t.coff = (long *)(p + 5 * ni + nl);
    return t;
}

static size_t level_tables_bytes(int nm) { return 5 * al64(sizeof(int) * (nm + 1)) + 2 * al64(sizeof(long) * nm); }

// Queue one level (merges of table dmg) on the device schedule.
static int level_dev(hsdc_ctx *c, const Merge *dmg, const std::vector<Merge> &hm, int L, const int *dchunkN,
                     int totN, const LvlBound &B, Prob *dprobs, int cur, int *dinfo, int gbase)
{
    const int nm = (int)hm.size();
    hipStream_t s = c->stream;
    Work W = c->W;
    MergeOut *dmo = (MergeOut *)c->mo.p;
    const LvlTab t = level_tables(c, nm);
    int nmax = 0;
    for (const Merge &M : hm) nmax = std::max(nmax, M.n);
    const int ds_cap = nmax <= 4096 ? nmax : 0;
    hipLaunchKernelGGL(k_prep, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, dprobs, W, cur);
    hipLaunchKernelGGL(k_mergepath, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, W);
    if (nmax >= 2048)
        hipLaunchKernelGGL(k_deflate<1024>, dim3(nm), dim3(1024), sizeof(double) * ds_cap, s, dmg, W, cur, dmo,
                           c->prm.hss_min_k, ds_cap);
    else
        hipLaunchKernelGGL(k_deflate<256>, dim3(nm), dim3(WGS), sizeof(double) * ds_cap, s, dmg, W, cur, dmo,
                           c->prm.hss_min_k, ds_cap);
    hipLaunchKernelGGL(k_lvltab, dim3(1), dim3(256), 0, s, dmg, nm, (const MergeOut *)dmo, W, t, B.sbcap, B.ctcap,
                       dinfo);
    hipLaunchKernelGGL(k_rot_prep, dim3(nm), dim3(WGS), 0, s, dmg, dprobs, W, cur, (const MergeOut *)dmo);
    hipLaunchKernelGGL(k_rotate, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, dprobs, W, cur,
                       (const MergeOut *)dmo);
    prof_mark(c, 0);
    if (c->sec_pre && nmax >= std::max(3, c->prm.learned_min_k))  // else no merge can use it
        hipLaunchKernelGGL(k_secstart, dim3(dev_grid(B.blocks)), dim3(SEC_WG), 0, s, dmg, (const int *)t.blockK, nm,
                           (const MergeOut *)dmo, W, c->dsec);
    hipLaunchKernelGGL(k_secular, dim3(dev_grid(B.blocks)), dim3(SEC_WG), 0, s, dmg, (const int *)t.blockK, nm,
                       (const MergeOut *)dmo, W, c->prm.maxit, dinfo, c->dsec,
                       c->dlvlmax + SEC_SLOTS * std::min(L, SEC_LEVELS - 1));
    prof_mark(c, -1);
    prof_mark(c, 1);
    hipLaunchKernelGGL(k_loewner, dim3(totN), dim3(WGS), 0, s, dmg, (const int *)t.chunkK, nm,
                       (const MergeOut *)dmo, W);
    prof_mark(c, -1);
    hipLaunchKernelGGL(k_sbuild, dim3(totN), dim3(WGS), 0, s, dmg, (const int *)t.chunkK, nm, (const MergeOut *)dmo,
                       (const long *)t.soff, W, (double *)c->sb.p);
    if (B.tiles > 0)
        hipLaunchKernelGGL(k_gemm, dim3(dev_grid(B.tiles)), dim3(WGS), 0, s, dmg, (const GTile *)nullptr, 0,
                           (const int *)t.tileK, nm, dprobs, (const MergeOut *)dmo, (const long *)t.soff,
                           (const long *)t.coff, W, cur, (const double *)c->sb.p, (double *)c->ct.p);
    hipLaunchKernelGGL(k_newcols_copy, dim3(totN + (B.scat > 0 ? dev_grid(B.scat) : 0)), dim3(WGS), 0, s, dmg,
                       dchunkN, nm, totN, (const int *)t.scK, (const MergeOut *)dmo, (const long *)t.coff, dprobs, W,
                       cur, (const double *)c->ct.p);
    hipLaunchKernelGGL(k_lamrg, dim3(totN), dim3(WGS), 0, s, dmg, dchunkN, nm, (const MergeOut *)dmo, W, cur ^ 1,
                       c->dtrc ? c->dtrc + 3 * (long)gbase : nullptr);
    HK(hipGetLastError());
    return 0;
}

// The HSS merges of device-scheduled level L, which stopped at its gate: the host
// reads the level's results, builds and applies their HSS matrices and copies the
// vectors into place, then opens the gate.
static int level_hss(hsdc_ctx *c, const Merge *dmg, const std::vector<Merge> &hm, int totN, const LvlBound &B,
                     Prob *dprobs, int cur, int gbase)
{
    const int nm = (int)hm.size();
    hipStream_t s = c->stream;
    const LvlTab t = level_tables(c, nm);
    MergeOut *dmo = (MergeOut *)c->mo.p;
    const size_t bmo = sizeof(MergeOut) * nm;
    CK(ensure_pin(c->pin2, bmo));
    HK(hipMemcpyAsync(c->pin2.p, dmo, bmo, hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    std::vector<MergeOut> ho(nm);
    memcpy(ho.data(), c->pin2.p, bmo);
    std::vector<int> sel;
    for (int i = 0; i < nm; i++)
        if (!ho[i].flags && ho[i].hss && ho[i].k > 0) sel.push_back(i);
    gate_open(c);
    if (sel.empty()) return 0;
    Work W = c->W;
    prof_mark(c, 1);
    hipLaunchKernelGGL(k_alpha, dim3(totN), dim3(WGS), 0, s, dmg, (const int *)t.chunkK, nm, (const MergeOut *)dmo,
                       W);
    prof_mark(c, -1);
    CK(run_hss(c, dmg, hm, 0, ho, sel, dprobs, cur, gbase));  // writes the new vectors into their slots
    hipLaunchKernelGGL(k_scatter, dim3(dev_grid(B.cols)), dim3(WGS), 0, s, dmg, (const int *)t.colK, nm,
                       (const MergeOut *)dmo, (const long *)t.coff, dprobs, W, cur ^ 1, (const double *)c->ct.p, 2);
    HK(hipGetLastError());
    return 0;
}

// Solve every problem of the plan; hd/he are the torn, scaled host arrays.
// Output: either one global Z (zrows_global = ntot, ldz) or per-problem blocks.
// With din_d/din_e (device, one problem, plan built from its order alone) the
// device scans, scales and tears the input itself, the static tables stay on the
// device between calls of the same shape, and SOLVE_ON_HOST comes back when the
// input needs the host's split (or is zero or not finite).
#define SOLVE_ON_HOST 0x7fff0001
static int solve(hsdc_ctx *c, Plan &pl, const double *hd, const double *he, double *dZ,
                 const std::vector<long> &zoff, const std::vector<long> &ldz, int zglobal,
                 double *d_out_dev, const std::vector<int> *outcol, const double *din_d = nullptr,
                 const double *din_e = nullptr)
{
    const bool dev_in = din_d != nullptr;
    // graph capture needs a stream other than the null stream: a solve on the null
    // stream runs on the context's own stream once the null stream's work is done
    struct StreamRestore {
        hsdc_ctx *c;
        hipStream_t s0;
        ~StreamRestore() { c->stream = s0; }
    } restore{c, c->stream};
    // (an event recorded under stream capture does not time the replays: profile = 1
    // launches directly)
    const bool use_graph = c->prm.graph && c->prm.batched && !c->prm.level_sync && !c->prm.profile;
    if (use_graph && c->stream == nullptr) {
        HK(hipStreamSynchronize(nullptr));
        if (!c->gstream) HK(hipStreamCreate(&c->gstream));
        c->stream = c->gstream;
    }
    hipStream_t s = c->stream;
    const int ntot = pl.ntot;
    if (!c->ev0) {
        HK(hipEventCreate(&c->ev0));
        HK(hipEventCreate(&c->ev1));
    }
    if (!c->rec) {
        HK(hipHostMalloc((void **)&c->rec, sizeof(SolveRec), hipHostMallocMapped | hipHostMallocCoherent));
        HK(hipHostGetDevicePointer((void **)&c->drec, c->rec, 0));
    }
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    c->pused = 0;
    HK(hipEventRecord(e0, s));
    CK(alloc_work(c, ntot));
    // Q storage: one n x n block per problem
    size_t qtot = 0;
    std::vector<Prob> hp(pl.probs.size());
    for (size_t i = 0; i < pl.probs.size(); i++) qtot += (size_t)pl.probs[i].n * pl.probs[i].n;
    CK(ensure(c->q, sizeof(double) * qtot));
    size_t qo = 0;
    for (size_t i = 0; i < pl.probs.size(); i++) {
        hp[i].q = (double *)c->q.p + qo;
        hp[i].ld = pl.probs[i].n;
        hp[i].off = pl.probs[i].off;
        hp[i].n = pl.probs[i].n;
        hp[i].par = pl.probs[i].levels & 1;
        qo += (size_t)pl.probs[i].n * pl.probs[i].n;
    }
    c->hprobs = hp;
    // static tables: probs | leaves | merges of all levels | d | e | scale | colprob | zoff | ldz | outcol | info
    size_t nl = pl.leaves.size(), nmtot = 0;
    for (auto &L : pl.lv) nmtot += L.size();
    size_t np = pl.probs.size();
    size_t a0 = 0, a1 = al64(a0 + sizeof(Prob) * np), a2 = al64(a1 + sizeof(Leaf) * nl),
           a3 = al64(a2 + sizeof(Merge) * nmtot), a4 = al64(a3 + sizeof(double) * ntot),
           a5 = al64(a4 + sizeof(double) * ntot), a6 = al64(a5 + sizeof(double) * np),
           a7 = al64(a6 + sizeof(int) * ntot), a8 = al64(a7 + sizeof(long) * np),
           a9 = al64(a8 + sizeof(long) * np), a10 = al64(a9 + sizeof(int) * ntot), a11 = a10 + 64 + 512;
    // chunk tables: per level (nm+1 entries) followed by per merge {0, chunks}
    std::vector<int> chunks;
    std::vector<size_t> lvchunk, mchunk;
    for (auto &L : pl.lv) {
        lvchunk.push_back(chunks.size());
        int acc = 0;
        chunks.push_back(0);
        for (auto &M : L) { acc += cdiv(M.n, WGS); chunks.push_back(acc); }
        for (auto &M : L) { mchunk.push_back(chunks.size()); chunks.push_back(0); chunks.push_back(cdiv(M.n, WGS)); }
    }
    // final copy: row chunks of FIN_CH per output column; per problem, its first
    // block (one output block per problem) -- zglobal: one n x n output, ntot rows
    std::vector<int> fbs(np + 1, 0);
    for (size_t i = 0; i < np; i++) fbs[i + 1] = fbs[i] + pl.probs[i].n * cdiv(pl.probs[i].n, FIN_CH);
    const int fin_cz = cdiv(ntot, FIN_CH);
    const long fin_blocks = zglobal ? (long)ntot * fin_cz : (long)fbs[np];
    size_t a12 = al64(a11 + sizeof(int) * chunks.size()), a13 = al64(a12 + sizeof(int) * (np + 1));
    size_t a14 = dev_in ? al64(a13 + sizeof(int) * ntot) : a13;  // tear flags of the device input
    const size_t a15 = al64(a14 + sizeof(int) * 3 * std::max(nmtot, (size_t)1));  // per-merge trace
    CK(ensure(c->probs, a15));
    std::vector<long> tkey;
    if (dev_in)
        tkey = {(long)ntot, (long)nl, (long)nmtot, (long)np, (long)ldz[0], (long)zglobal, c->prm.smlsiz,
                (long)(uintptr_t)c->q.p, (long)(uintptr_t)c->probs.p};
    const bool cached = dev_in && tkey == c->tabkey;
    CK(ensure_pin(c->pin, a14));
    char *h = (char *)c->pin.p;
    size_t mo = 0;
    std::vector<size_t> lvoff;
    for (auto &L : pl.lv) {
        lvoff.push_back(mo);
        if (!cached) memcpy(h + a2 + sizeof(Merge) * mo, L.data(), sizeof(Merge) * L.size());
        mo += L.size();
    }
    if (!cached) {
        memcpy(h + a0, hp.data(), sizeof(Prob) * np);
        memcpy(h + a1, pl.leaves.data(), sizeof(Leaf) * nl);
        if (!dev_in) {
            memcpy(h + a3, hd, sizeof(double) * ntot);
            memcpy(h + a4, he, sizeof(double) * ntot);
        } else {
            std::vector<int> tf(ntot, 0);
            for (const Leaf &Lf : pl.leaves) {
                if (Lf.s > 0) tf[Lf.s] |= 1;
                if (Lf.s + Lf.m < ntot) tf[Lf.s + Lf.m - 1] |= 2;
            }
            memcpy(h + a13, tf.data(), sizeof(int) * ntot);
        }
        std::vector<double> scl(np);
        std::vector<int> colprob(ntot);
        for (size_t i = 0; i < np; i++) {
            scl[i] = pl.probs[i].scale;
            for (int j = 0; j < pl.probs[i].n; j++) colprob[pl.probs[i].off + j] = (int)i;
        }
        memcpy(h + a5, scl.data(), sizeof(double) * np);
        memcpy(h + a6, colprob.data(), sizeof(int) * ntot);
        std::vector<long> zo(np, 0), lz(np, 1);
        for (size_t i = 0; i < np && i < zoff.size(); i++) zo[i] = zoff[i];
        for (size_t i = 0; i < np && i < ldz.size(); i++) lz[i] = ldz[i];
        memcpy(h + a7, zo.data(), sizeof(long) * np);
        memcpy(h + a8, lz.data(), sizeof(long) * np);
        if (outcol) memcpy(h + a9, outcol->data(), sizeof(int) * ntot);
        memset(h + a10, 0, 64 + 512);
        if (!chunks.empty()) memcpy(h + a11, chunks.data(), sizeof(int) * chunks.size());
        memcpy(h + a12, fbs.data(), sizeof(int) * (np + 1));
    }
    char *dp = (char *)c->probs.p;
    if (!cached) HK(hipMemcpyAsync(dp, h, a14, hipMemcpyHostToDevice, s));
    c->tabkey = tkey;  // empty (nothing kept) for host input
    const int *dfbs = (const int *)(dp + a12);
    const int *dtflag = (const int *)(dp + a13);
    c->dtrc = (int *)(dp + a14);
    c->trc_mg.clear();
    for (auto &Lm : pl.lv) c->trc_mg.insert(c->trc_mg.end(), Lm.begin(), Lm.end());
    c->trc_rank.assign(nmtot, 0);
    c->trc_ntot = ntot;
    c->trc_valid = true;
    const int *dchunks = (const int *)(dp + a11);
    Prob *dprobs = (Prob *)(dp + a0);
    Leaf *dleaves = (Leaf *)(dp + a1);
    Merge *dmerges = (Merge *)(dp + a2);
    double *dd = (double *)(dp + a3), *de = (double *)(dp + a4), *dscale = (double *)(dp + a5);
    int *dcolprob = (int *)(dp + a6);
    long *dzoff = (long *)(dp + a7), *dldz = (long *)(dp + a8);
    int *doutcol = outcol ? (int *)(dp + a9) : nullptr;
    int *dinfo = (int *)(dp + a10);
    CK(sec_setup(c));
    // schedule of every level, and the buffers of the device-scheduled ones (sized
    // before the first launch: a buffer that grows moves)
    const int nlv = (int)pl.lv.size();
    std::vector<LvlBound> lb(nlv);
    std::vector<char> dev(nlv, 0);
    size_t nm_max = 1;
    long sbmax = 1, ctmax = 1;
    bool all_dev = true;
    for (int L = 0; L < nlv; L++) {
        nm_max = std::max(nm_max, pl.lv[L].size());
        lb[L] = level_bound(pl.lv[L], c->prm.hss_min_k);
        dev[L] = c->prm.batched && !c->prm.level_sync && 8 * (lb[L].sbcap + lb[L].ctcap) <= DEV_BUDGET;
        if (dev[L]) { sbmax = std::max(sbmax, lb[L].sbcap); ctmax = std::max(ctmax, lb[L].ctcap); }
        all_dev &= dev[L] != 0;
    }
    CK(ensure(c->mo, sizeof(MergeOut) * nm_max));
    CK(ensure(c->ltab, level_tables_bytes((int)nm_max)));
    CK(ensure(c->sb, sizeof(double) * (size_t)sbmax));
    CK(ensure(c->ct, sizeof(double) * (size_t)ctmax));
    int lmax = 1;
    for (auto &Lf : pl.leaves) lmax = std::max(lmax, Lf.m);
    const int zrows = zglobal ? ntot : 0;
    unsigned long long *drows = (unsigned long long *)(dp + a10 + 64);  // 64 counters
    // events around the final kernel (profile = 1), kept by the context
    if (c->prm.profile && !c->pf0) {
        HK(hipEventCreate(&c->pf0));
        HK(hipEventCreate(&c->pf1));
        HK(hipEventCreate(&c->pz0));
        HK(hipEventCreate(&c->pz1));
    }
    hipEvent_t f0 = c->pf0, f1 = c->pf1;
    // one output holding the whole solve: zero it on a second stream while the levels
    // run, the final reordering then writes the supported rows only
    const bool zfill = zglobal != 0;
    if (zfill && !c->zstream) {
        HK(hipStreamCreateWithFlags(&c->zstream, hipStreamNonBlocking));
        HK(hipEventCreateWithFlags(&c->zev0, hipEventDisableTiming));
        HK(hipEventCreateWithFlags(&c->zev1, hipEventDisableTiming));
    }
    // first launches: records and gate cleared, leaves
    auto launch_head = [&]() -> int {
        if (zfill) {  // fork: the fill starts after the stream's earlier work on Z
            HK(hipEventRecord(c->zev0, s));
            HK(hipStreamWaitEvent(c->zstream, c->zev0, 0));
            if (c->prm.profile) HK(hipEventRecord(c->pz0, c->zstream));
            hipLaunchKernelGGL(k_zfill, dim3(32768), dim3(WGS), 0, c->zstream, dZ, (long)ntot, (long)ntot, ldz[0]);
            HK(hipGetLastError());
            if (c->prm.profile) HK(hipEventRecord(c->pz1, c->zstream));
            HK(hipEventRecord(c->zev1, c->zstream));
        }
        long nmx = 4 * ((long)c->wcap / 2 + 2) + 8;  // per-merge records, gate, level statistics
        hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(nmx, WGS)), dim3(WGS), 0, s, c->W.maux, nmx, 0ULL);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(WGS), 0, s, (unsigned long long *)(dp + a10), 72L, 0ULL);
        if (dev_in) {
            hipLaunchKernelGGL(k_inp_scan, dim3(cdiv(ntot, WGS)), dim3(WGS), 0, s, din_d, din_e, ntot, c->W);
            hipLaunchKernelGGL(k_inp_scale, dim3(cdiv(ntot, WGS)), dim3(WGS), 0, s, din_d, din_e, ntot, dtflag,
                               dmerges, (int)nmtot, dd, de, dscale, c->W);
        }
        // 64 / G leaves per wavefront, G the smallest of 16, 32, 64 that holds the largest leaf
        const int G = lmax <= 16 ? 16 : (lmax <= 32 ? 32 : 64);
        hipLaunchKernelGGL(k_leaf, dim3(cdiv((long)nl, 64 / G)), dim3(64), sizeof(double) * (64 / G) * lmax * lmax,
                           s, dleaves, (int)nl, dprobs, dd, de, c->W, 0, G, lmax);
        HK(hipGetLastError());
        return 0;
    };
    auto launch_final = [&]() -> int {
        if (zfill) HK(hipStreamWaitEvent(s, c->zev1, 0));  // join
        if (c->prm.profile) HK(hipEventRecord(f0, s));
        hipLaunchKernelGGL(k_final, dim3(zfill ? (long)ntot : fin_blocks), dim3(WGS), 0, s, dprobs, dcolprob, c->W,
                           doutcol, dZ, dzoff, dldz, zrows, d_out_dev, dscale, drows, dfbs, (int)np, fin_cz,
                           zfill ? 1 : 0);
        HK(hipGetLastError());
        if (c->prm.profile) HK(hipEventRecord(f1, s));
        hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, s, c->W, (const int *)dinfo,
                           (const unsigned long long *)drows, c->dsec, (const unsigned int *)c->dlvlmax, c->drec);
        HK(hipGetLastError());
        return 0;
    };
    auto launch_dev = [&](int L) -> int {
        c->cur_level = L;
        const int nm = (int)pl.lv[L].size();
        return level_dev(c, dmerges + lvoff[L], pl.lv[L], L, dchunks + lvchunk[L], chunks[lvchunk[L] + nm], lb[L],
                         dprobs, L & 1, dinfo, (int)lvoff[L]);
    };
    // With every level on the device schedule the whole solve is one HIP graph,
    // captured once per shape (plan, buffers, parameters) and replayed.
    bool queued = false;  // levels and final kernel already in the stream
    if (all_dev && use_graph && s != nullptr) {
        std::vector<long> key = {(long)nlv, (long)np, (long)nl, (long)lmax, (long)ntot, (long)c->wcap, zrows,
                                 fin_blocks, (long)(outcol != nullptr), c->prm.maxit, c->prm.hss_min_k,
                                 (long)c->sec_pre, (long)c->prm.learned_min_k, (long)c->prm.profile, (long)zfill,
                                 (long)ldz[0], (long)(uintptr_t)c->zstream,
                                 (long)(uintptr_t)c->stream};
        for (const void *ptr : {(const void *)c->wsd.p, (const void *)c->wsi.p, (const void *)c->maux.p,
                                (const void *)c->probs.p, (const void *)c->mo.p, (const void *)c->ltab.p,
                                (const void *)c->sb.p, (const void *)c->ct.p, (const void *)c->dsec,
                                (const void *)c->dlvlmax, (const void *)dZ, (const void *)d_out_dev,
                                (const void *)din_d, (const void *)din_e})
            key.push_back((long)(uintptr_t)ptr);
        for (int L = 0; L < nlv; L++) {
            int nmax = 0;
            for (const Merge &M : pl.lv[L]) nmax = std::max(nmax, M.n);
            key.insert(key.end(), {(long)pl.lv[L].size(), (long)chunks[lvchunk[L] + pl.lv[L].size()], (long)nmax,
                                   lb[L].blocks, lb[L].tiles, lb[L].cols, lb[L].sbcap, lb[L].ctcap, lb[L].scat});
        }
        if (!c->gexec || key != c->gkey) {
            if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            int rc = HSDC_ERR_HIP;
            hipGraph_t g = nullptr;
            hipError_t ec = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (ec == hipSuccess) {
                rc = launch_head();
                for (int L = 0; L < nlv && !rc; L++) rc = launch_dev(L);
                if (!rc) rc = launch_final();
                ec = hipStreamEndCapture(s, &g);
            }
            if (!rc && ec == hipSuccess && hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
                c->gexec = nullptr;
                rc = HSDC_ERR_HIP;
            }
            if (g) (void)hipGraphDestroy(g);
            if (ec != hipSuccess) rc = HSDC_ERR_HIP;
            if (rc) {  // nothing ran: queue the same launches directly
                (void)hipGetLastError();
                if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
                c->gexec = nullptr;
                c->gkey.clear();
            } else {
                c->gkey = key;
            }
        }
        if (c->gexec) {
            HK(hipGraphLaunch(c->gexec, s));
            queued = true;
        }
    }
    if (!queued) {
        CK(launch_head());
        // the host tables above live in c->pin, which run_merges reuses: wait for the
        // upload (and, for device input, see whether the host has to take over)
        if (!all_dev) {
            int g = 0;
            HK(hipMemcpyAsync(&g, c->W.gate, sizeof(int), hipMemcpyDeviceToHost, s));
            HK(hipStreamSynchronize(s));
            if (g < 0) return SOLVE_ON_HOST;
        }
    }
    // Runs of device-scheduled levels are queued back to back; the host looks at the
    // gate once per run (normally once per solve, together with the final kernel).
    // Level L works on ping-pong parity L & 1.
    int L = 0, ndev = 0;
    for (;;) {
        int L1 = L;
        if (queued) {
            L1 = nlv;
            queued = false;
        } else {
            for (; L1 < nlv && dev[L1]; L1++) CK(launch_dev(L1));
            if (L1 == nlv) CK(launch_final());
        }
        const bool fin = L1 == nlv;
        if (L1 > L || fin) {
            int g = 0;
            if (fin) {  // the end-of-solve record carries the gate
                HK(hipEventRecord(e1, s));
                HK(hipStreamSynchronize(s));
                g = c->rec->gate;
            } else {
                HK(hipMemcpyAsync(&g, c->W.gate, sizeof(int), hipMemcpyDeviceToHost, s));
                HK(hipStreamSynchronize(s));
            }
            if (g < 0) return SOLVE_ON_HOST;  // the device input needs the host's split
            if (g) {  // level g - 1 holds HSS merges; the levels above it did nothing
                const int Lh = g - 1, nm = (int)pl.lv[Lh].size();
                for (int l = L; l <= Lh; l++) ndev += (int)pl.lv[l].size();
                CK(level_hss(c, dmerges + lvoff[Lh], pl.lv[Lh], chunks[lvchunk[Lh] + nm], lb[Lh], dprobs, Lh & 1,
                             (int)lvoff[Lh]));
                L = Lh + 1;
                continue;
            }
            for (int l = L; l < L1; l++) ndev += (int)pl.lv[l].size();
        }
        if (fin) break;
        // a level on the synchronised schedule
        const std::vector<Merge> &hm = pl.lv[L1];
        const Merge *dm = dmerges + lvoff[L1];
        const int nm = (int)hm.size();
        c->cur_level = L1;
        int cur = L1 & 1;
        if (c->prm.batched) {
            CK(run_merges(c, dm, hm, 0, nm, dchunks + lvchunk[L1], chunks[lvchunk[L1] + nm], dprobs, cur, dinfo,
                          (int)lvoff[L1]));
        } else {
            // one merge at a time, every merge keeping the same ping-pong parity
            for (int m = 0; m < nm; m++) {
                int cc = L1 & 1;
                const size_t mi = lvoff[L1] + m;
                CK(run_merges(c, dm, hm, m, m + 1, dchunks + mchunk[mi], chunks[mchunk[mi] + 1], dprobs, cc,
                              dinfo, (int)lvoff[L1]));
            }
        }
        L = L1 + 1;
    }
    c->st.levels += nlv;
    c->st.merges += ndev;
    const SolveRec R = *c->rec;
    const int hinfo = R.info;
    const unsigned long long rows = R.rows;
    c->st.roots += (long long)R.roots;
    c->st.deflated += (long long)R.deflated;
    c->st.hss_merges += (int)R.hss;
    if (nlv > 0 && dev[nlv - 1]) c->st.k_top = (int)R.ktop;  // K of the top merge
    c->st.secular_crit_iters = (long long)R.crit;
    c->st.secular_iters = (long long)R.iters;
    c->st.learned_roots = (long long)R.used;
    c->st.secular_wave_iters = (long long)R.wave_iters;
    c->st.secular_evals = (long long)R.evals;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c->st.ms_total = ms;
    long long zrows_written = 0;
    for (auto &P : pl.probs) zrows_written += (long long)(zglobal ? ntot : P.n) * P.n;
    if (zfill) {
        c->st.zfill_bytes = 8LL * zrows_written;
        c->st.final_bytes = 16LL * (long long)rows + 16LL * ntot;  // supported rows read and written
    } else {
        c->st.final_bytes = 8LL * (zrows_written + (long long)rows) + 16LL * ntot;
    }
    if (c->prm.profile) {
        prof_collect(c);
        // (i, j) pairs of the Loewner and alpha sweeps, from the merges' K and branch
        std::vector<int> tk(c->trc_mg.size()), th(c->trc_mg.size());
        const int ntr = hsdc_trace(c, nullptr, nullptr, tk.data(), nullptr, th.data(), nullptr, (int)tk.size());
        for (int i = 0; i < ntr; i++) {
            const long long k = tk[i];
            if (k >= 3) c->st.loewner_pairs += k * k * (th[i] ? 2 : 1);
        }
        float fm = 0, zm = 0;
        if (hipEventElapsedTime(&fm, f0, f1) != hipSuccess) { (void)hipGetLastError(); fm = 0; }
        c->st.ms_final = fm;
        if (zfill && hipEventElapsedTime(&zm, c->pz0, c->pz1) != hipSuccess) { (void)hipGetLastError(); zm = 0; }
        c->st.ms_zfill = zm;
    }
    if (hinfo & 2) return HSDC_ERR_INTERNAL;  // device-scheduled level over its bounds
    return hinfo ? 1 : 0;
}

// Device-to-host copy of a rows x cols column-major matrix (device leading dimension
// rows) into pageable host memory (leading dimension ldh).  A plain pageable copy runs
// through the runtime's single staging loop; here blocks of columns go by DMA into
// three pinned staging buffers while host threads copy the previous block out, so the
// PCIe transfer and the host copy overlap.
static int copy_d2h_matrix(hsdc_ctx *c, double *host, long ldh, const double *dev, long rows, long cols)
{
    if (rows <= 0 || cols <= 0) return 0;
    hipStream_t s = c->stream;
    const size_t colb = sizeof(double) * (size_t)rows;
    const long per = std::max(1L, (long)((64L << 20) / (long)colb));  // columns per block (64 MB)
    const int NB = 3;
    CK(ensure_pin(c->stage, (size_t)NB * per * colb));
    for (int i = 0; i < NB; i++)
        if (!c->stev[i]) HK(hipEventCreateWithFlags(&c->stev[i], hipEventDisableTiming));
    double *st = (double *)c->stage.p;
    const long nblk = (cols + per - 1) / per;
    auto issue = [&](long b) -> int {
        const long w = std::min(per, cols - b * per);
        HK(hipMemcpyAsync(st + (size_t)(b % NB) * per * rows, dev + (size_t)b * per * rows, colb * w,
                          hipMemcpyDeviceToHost, s));
        HK(hipEventRecord(c->stev[b % NB], s));
        return 0;
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int T = (int)std::min(8u, std::max(1u, hw / 2));
    for (long b = 0; b < std::min<long>(NB, nblk); b++) CK(issue(b));
    for (long b = 0; b < nblk; b++) {
        HK(hipEventSynchronize(c->stev[b % NB]));
        const long w = std::min(per, cols - b * per);
        const double *src = st + (size_t)(b % NB) * per * rows;
        double *dst = host + (size_t)b * per * ldh;
        // columns [w t / nt, w (t + 1) / nt) of the block, t < nt
        auto part = [&](int t, int nt) {
            const long c0 = w * t / nt, c1 = w * (t + 1) / nt;
            if (ldh == rows) memcpy(dst + (size_t)c0 * ldh, src + (size_t)c0 * rows, colb * (c1 - c0));
            else
                for (long j = c0; j < c1; j++) memcpy(dst + (size_t)j * ldh, src + (size_t)j * rows, colb);
        };
        if (T == 1 || (long)colb * w < (4L << 20)) {
            part(0, 1);
        } else {
            std::vector<std::thread> th;
            for (int t = 1; t < T; t++) th.emplace_back(part, t, T);
            part(0, T);
            for (auto &x : th) x.join();
        }
        if (b + NB < nblk) CK(issue(b + NB));
    }
    return 0;
}

// Splitting and scaling of MDSTEDC (mdstedc.f90:305-359) into a plan of problems.
static void make_plan_split(Plan &pl, int n, double *d, double *e, int smlsiz)
{
    int start = 0;
    while (start < n) {
        int finish = start;
        while (finish < n - 1) {
            double tiny = HSDC_EPS * sqrt(fabs(d[finish])) * sqrt(fabs(d[finish + 1]));
            if (fabs(e[finish]) > tiny) finish++;
            else break;
        }
        int m = finish - start + 1;
        HostProb P;
        P.off = start; P.n = m; P.scale = 1.0; P.levels = 0;
        if (m > smlsiz) {
            double onrm = 0.0;
            for (int i = 0; i < m; i++) onrm = fmax(onrm, fabs(d[start + i]));
            for (int i = 0; i < m - 1; i++) onrm = fmax(onrm, fabs(e[start + i]));
            double mul = 1.0 / onrm;
            for (int i = 0; i < m; i++) d[start + i] *= mul;
            for (int i = 0; i < m - 1; i++) e[start + i] *= mul;
            P.scale = onrm;
        }
        pl.probs.push_back(P);
        start = finish + 1;
    }
    pl.ntot = n;
    for (size_t i = 0; i < pl.probs.size(); i++) plan_problem(pl, (int)i, d, e, smlsiz);
}

// ============================================================================
// C ABI
// ============================================================================
// HSS batch of another solver on its stream (declared in hsdc_hss.h)
int hsdc_hss_batch(hsdc_ctx *c, hipStream_t s, const std::vector<HssSpec> &sp, int ni, double tol,
                   std::vector<int> *rank, int *dense)
{
    if (!c) return HSDC_ERR_INTERNAL;
    hipStream_t s0 = c->stream;
    c->stream = s;
    std::vector<HssPlanM> keep;
    const int rc = hss_apply(c, sp, ni, tol, &keep);
    c->stream = s0;
    if (rc) return rc;
    int nd = 0;
    if (rank) rank->assign(keep.size(), 0);
    for (size_t h = 0; h < keep.size(); h++) {
        const HssPlanM &Q = keep[h];
        if (!Q.ok) { nd++; continue; }
        int rm = 0;
        for (int i = 0; i < Q.T.ltr - 1; i++) rm = std::max(rm, std::max(Q.usz[2 * i + 1], Q.vsz[2 * i]));
        if (rank) (*rank)[h] = rm;
    }
    if (dense) *dense = nd;
    return 0;
}

extern "C" {

void hsdc_default_params(hsdc_params *p)
{
    p->smlsiz = 25;
    p->hss_min_k = 1500;
    p->hss_leaf = 200;
    p->hss_tol = (double)1.0e-17f;
    p->maxit = 100;
    p->batched = 1;
    p->learned = 0;
    p->profile = 0;
    p->learned_min_k = 256;
    p->learned_gate = 0.1;
    p->level_sync = 0;
    p->graph = 1;
}

// The context touches no device until the first solve, so argument checking and
// workspace queries work on a host without a GPU.
int hsdc_ctx_create(hsdc_ctx **ctx)
{
    if (!ctx) return -1;
    hsdc_ctx *c = new hsdc_ctx();
    hsdc_default_params(&c->prm);
    memset(&c->st, 0, sizeof(c->st));
    *ctx = c;
    return 0;
}

int hsdc_ctx_destroy(hsdc_ctx *c)
{
    if (!c) return 0;
    DevBuf *bufs[] = {&c->q,  &c->ct,  &c->sb,   &c->din, &c->ein,  &c->wsd,  &c->wsi, &c->tabs,
                      &c->lvl, &c->mo, &c->probs, &c->misc, &c->zout, &c->dout, &c->hss, &c->htab,
                      &c->xb, &c->gf,  &c->cb,   &c->yb, &c->tb,  &c->mlpw, &c->hv,   &c->sec, &c->mlptr,
                      &c->maux, &c->featb, &c->ltab};
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    if (c->pf0) (void)hipEventDestroy(c->pf0);
    if (c->pf1) (void)hipEventDestroy(c->pf1);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : {c->zev0, c->zev1, c->pz0, c->pz1}) if (e) (void)hipEventDestroy(e);
    if (c->zstream) (void)hipStreamDestroy(c->zstream);
    if (c->rec) (void)hipHostFree(c->rec);
    for (DevBuf *b : bufs) if (b->p) (void)hipFree(b->p);
    if (c->pin.p) (void)hipHostFree(c->pin.p);
    if (c->pin2.p) (void)hipHostFree(c->pin2.p);
    if (c->pin3.p) (void)hipHostFree(c->pin3.p);
    if (c->pinh.p) (void)hipHostFree(c->pinh.p);
    if (c->pinb_ev) (void)hipEventSynchronize(c->pinb_ev);
    if (c->pinb.p) (void)hipHostFree(c->pinb.p);
    if (c->pinb_ev) (void)hipEventDestroy(c->pinb_ev);
    if (c->stage.p) (void)hipHostFree(c->stage.p);
    for (hipEvent_t e : c->stev) if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->pev) if (e) (void)hipEventDestroy(e);
    delete c;
    return 0;
}

int hsdc_set_params(hsdc_ctx *c, const hsdc_params *p)
{
    if (!c || !p) return -1;
    if (p->smlsiz < 1 || p->smlsiz > 64) return -2;
    if (p->hss_leaf < 8 || p->hss_tol < 0 || p->maxit < 0) return -2;
    if (!(p->learned_gate >= 0.0)) return -2;
    c->prm = *p;
    if (c->prm.hss_min_k < 4) c->prm.hss_min_k = 4;
    return 0;
}

int hsdc_get_stats(hsdc_ctx *c, hsdc_stats *s)
{
    if (!c || !s) return -1;
    *s = c->st;
    return 0;
}

int hsdc_trace(hsdc_ctx *c, int *s, int *n, int *k, int *nrot, int *hss, int *hss_rank, int cap)
{
    if (!c) return -1;
    if (cap < 0) return -8;
    if (!c->trc_valid || !c->dtrc) return 0;
    const int nm = (int)c->trc_mg.size();
    std::vector<int> t(3 * (size_t)std::max(nm, 1));
    if (nm > 0) {
        HK(hipMemcpyAsync(t.data(), c->dtrc, sizeof(int) * 3 * nm, hipMemcpyDeviceToHost, c->stream));
        HK(hipStreamSynchronize(c->stream));
    }
    for (int i = 0; i < nm && i < cap; i++) {
        if (s) s[i] = c->trc_mg[i].s;
        if (n) n[i] = c->trc_mg[i].n;
        if (k) k[i] = t[3 * i];
        if (nrot) nrot[i] = t[3 * i + 1];
        if (hss) hss[i] = t[3 * i + 2];
        if (hss_rank) hss_rank[i] = c->trc_rank[i];
    }
    return nm;
}

int hsdc_trace_indxq(hsdc_ctx *c, int *indxq, int n)
{
    if (!c) return -1;
    if (!indxq) return -2;
    if (!c->trc_valid || n != c->trc_ntot || n < 1) return -3;
    HK(hipMemcpyAsync(indxq, c->W.indxq, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HK(hipStreamSynchronize(c->stream));
    return 0;
}

// Contexts owned by the calling host thread (created on first use, kept until
// hsdc_thread_ctx_release or the end of the process).  Device buffers persist between
// calls, so a repeated call of the same order allocates nothing.
static thread_local hsdc_ctx *t_ctx = nullptr;
__attribute__((visibility("hidden"))) void hsdc_sb_thread_release(void);  // the thread's banded context (hsdc_band.hip)
static thread_local hsdc_bd_ctx *t_bd_ctx = nullptr;

hsdc_ctx *hsdc_thread_ctx(void)
{
    if (!t_ctx && hsdc_ctx_create(&t_ctx) != 0) t_ctx = nullptr;
    if (t_ctx) {
        hsdc_params p;
        hsdc_default_params(&p);
        (void)hsdc_set_params(t_ctx, &p);
        (void)hsdc_set_stream(t_ctx, nullptr);
    }
    return t_ctx;
}

hsdc_bd_ctx *hsdc_bd_thread_ctx(void)
{
    if (!t_bd_ctx && hsdc_bd_ctx_create(&t_bd_ctx) != 0) t_bd_ctx = nullptr;
    if (t_bd_ctx) {
        hsdc_bd_params p;
        hsdc_bd_default_params(&p);
        (void)hsdc_bd_set_params(t_bd_ctx, &p);
        (void)hsdc_bd_set_stream(t_bd_ctx, nullptr);
    }
    return t_bd_ctx;
}

void hsdc_thread_ctx_release(void)
{
    if (t_ctx) (void)hsdc_ctx_destroy(t_ctx);
    if (t_bd_ctx) (void)hsdc_bd_ctx_destroy(t_bd_ctx);
    t_ctx = nullptr;
    t_bd_ctx = nullptr;
    hsdc_sb_thread_release();
}

int hsdc_set_stream(hsdc_ctx *c, void *stream)
{
    if (!c) return -1;
    c->stream = (hipStream_t)stream;
    return 0;
}

// Device-resident COMPZ='I'
int hsdc_stedc_dev(hsdc_ctx *c, int n, double *d_d, double *d_e, double *d_z, int ldz)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (n < 0) return -2;
    if (ldz < (n > 1 ? n : 1)) return -6;
    memset(&c->st, 0, sizeof(c->st));
    if (n == 0) return 0;
    if (n > c->prm.smlsiz && c->prm.batched && !c->prm.level_sync) {
        // The input stays on the device: the plan comes from the order alone (one
        // problem), the device scales and tears the input and screens it for splits;
        // an input with a split, or zero, or not finite, is solved below instead.
        if (c->fplan_n != n || c->fplan_sml != c->prm.smlsiz) {
            c->fplan = Plan();
            HostProb P;
            P.off = 0; P.n = n; P.scale = 1.0; P.levels = 0;
            c->fplan.probs.push_back(P);
            c->fplan.ntot = n;
            std::vector<double> zd(n, 0.0), ze(n, 0.0);
            plan_problem(c->fplan, 0, zd.data(), ze.data(), c->prm.smlsiz);  // values filled on the device
            c->fplan_n = n;
            c->fplan_sml = c->prm.smlsiz;
        }
        std::vector<long> zoff(1, 0), lz(1, ldz);
        int info = solve(c, c->fplan, nullptr, nullptr, d_z, zoff, lz, 1, d_d, nullptr, d_d, d_e);
        if (info != SOLVE_ON_HOST) return info;
        memset(&c->st, 0, sizeof(c->st));
    }
    std::vector<double> hd(n), he(n, 0.0);
    HK(hipMemcpyAsync(hd.data(), d_d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (n > 1) HK(hipMemcpyAsync(he.data(), d_e, sizeof(double) * (n - 1), hipMemcpyDeviceToHost, c->stream));
    HK(hipStreamSynchronize(c->stream));
    double orgnrm = 0.0;
    for (int i = 0; i < n; i++) orgnrm = fmax(orgnrm, fabs(hd[i]));
    for (int i = 0; i < n - 1; i++) orgnrm = fmax(orgnrm, fabs(he[i]));
    if (n == 1 || orgnrm == 0.0) {
        // Z = I, D unchanged
        HK(hipMemset2DAsync(d_z, sizeof(double) * ldz, 0, sizeof(double) * n, n, c->stream));
        std::vector<double> one(1, 1.0);
        for (int j = 0; j < n; j++)
            HK(hipMemcpyAsync(d_z + (long)j * ldz + j, one.data(), sizeof(double), hipMemcpyHostToDevice, c->stream));
        HK(hipStreamSynchronize(c->stream));
        return 0;
    }
    Plan pl;
    make_plan_split(pl, n, hd.data(), he.data(), c->prm.smlsiz);
    // splits: eigenvalues need a global selection sort (mdstedc.f90:397-425), done
    // after the solve by permuting columns; first solve into the plan order.
    std::vector<long> zoff(1, 0), lz(1, ldz);
    int info;
    if (pl.probs.size() == 1) {
        info = solve(c, pl, hd.data(), he.data(), d_z, zoff, lz, 1, d_d, nullptr);
    } else {
        // solve once to learn the eigenvalues, then once more writing sorted columns
        CK(ensure(c->dout, sizeof(double) * n));
        info = solve(c, pl, hd.data(), he.data(), d_z, zoff, lz, 1, (double *)c->dout.p, nullptr);
        if (info) return info;
        std::vector<double> lam(n);
        HK(hipMemcpy(lam.data(), c->dout.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        // selection sort as MDSTEDC: track where each column goes
        std::vector<int> pos(n);
        for (int i = 0; i < n; i++) pos[i] = i;
        for (int ii = 1; ii < n; ii++) {
            int i = ii - 1, k = i;
            double p = lam[i];
            for (int j = ii; j < n; j++) if (lam[j] < p) { k = j; p = lam[j]; }
            if (k != i) { lam[k] = lam[i]; lam[i] = p; std::swap(pos[i], pos[k]); }
        }
        std::vector<int> outcol(n);
        for (int i = 0; i < n; i++) outcol[pos[i]] = i;
        // re-run with the same plan (deterministic) writing columns to their place
        Plan pl3;
        std::vector<double> hd3(n), he3(n, 0.0);
        HK(hipMemcpy(hd3.data(), d_d, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (n > 1) HK(hipMemcpy(he3.data(), d_e, sizeof(double) * (n - 1), hipMemcpyDeviceToHost));
        make_plan_split(pl3, n, hd3.data(), he3.data(), c->prm.smlsiz);
        info = solve(c, pl3, hd3.data(), he3.data(), d_z, zoff, lz, 1, d_d, &outcol);
    }
    return info;
}

int hsdc_stedc_batched_dev(hsdc_ctx *c, int nprob, const int *ns, const long *offs, double *d_d,
                           double *d_e, double *d_z, const long *zoffs)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (nprob < 0) return -2;
    memset(&c->st, 0, sizeof(c->st));
    long tot = 0;
    for (int p = 0; p < nprob; p++) {
        if (ns[p] < 1) return -3;
        tot = std::max(tot, offs[p] + ns[p]);
    }
    std::vector<double> hd(tot), he(tot, 0.0);
    HK(hipMemcpy(hd.data(), d_d, sizeof(double) * tot, hipMemcpyDeviceToHost));
    HK(hipMemcpy(he.data(), d_e, sizeof(double) * tot, hipMemcpyDeviceToHost));
    // concatenate problems in a compact global index space
    Plan pl;
    int g = 0;
    std::vector<double> gd, ge;
    std::vector<long> zo, lz;
    for (int p = 0; p < nprob; p++) {
        int n = ns[p];
        HostProb P;
        P.off = g; P.n = n; P.scale = 1.0; P.levels = 0;
        double onrm = 0.0;
        for (int i = 0; i < n; i++) onrm = fmax(onrm, fabs(hd[offs[p] + i]));
        for (int i = 0; i < n - 1; i++) onrm = fmax(onrm, fabs(he[offs[p] + i]));
        double mul = (n > c->prm.smlsiz && onrm > 0) ? 1.0 / onrm : 1.0;
        for (int i = 0; i < n; i++) gd.push_back(hd[offs[p] + i] * mul);
        for (int i = 0; i < n - 1; i++) ge.push_back(he[offs[p] + i] * mul);
        ge.push_back(0.0);
        if (mul != 1.0) P.scale = onrm;
        pl.probs.push_back(P);
        zo.push_back(zoffs[p]);
        lz.push_back(n);
        g += n;
    }
    pl.ntot = g;
    for (int p = 0; p < nprob; p++) plan_problem(pl, p, gd.data(), ge.data(), c->prm.smlsiz);
    // eigenvalues back into d_d at each problem's offset: solve into a compact
    // buffer, then scatter
    CK(ensure(c->dout, sizeof(double) * g));
    int info = solve(c, pl, gd.data(), ge.data(), d_z, zo, lz, 0, (double *)c->dout.p, nullptr);
    for (int p = 0; p < nprob; p++)
        HK(hipMemcpyAsync(d_d + offs[p], (double *)c->dout.p + pl.probs[p].off, sizeof(double) * ns[p],
                          hipMemcpyDeviceToDevice, c->stream));
    HK(hipStreamSynchronize(c->stream));
    return info;
}

// Autotuner: one timed batched solve per candidate (HSS leaf size x DC leaf bound) on
// copies of the caller's problems; the fastest candidate goes into the context's
// parameters.  Inputs are restored; d_z holds the last trial's vectors.
int hsdc_autotune(hsdc_ctx *c, int nprob, const int *ns, const long *offs, double *d_d, double *d_e,
                  double *d_z, const long *zoffs, int *best_leaf, int *best_smlsiz, double *times)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (nprob < 1) return -2;
    long tot = 0;
    for (int p = 0; p < nprob; p++) {
        if (ns[p] < 1) return -3;
        tot = std::max(tot, offs[p] + ns[p]);
    }
    if (!have_device()) return HSDC_ERR_NODEVICE;
    DevBuf keep;
    if (ensure(keep, sizeof(double) * 2 * (size_t)tot)) return HSDC_ERR_NOMEM;
    double *kd = (double *)keep.p, *ke = kd + tot;
    int rc = 0;
    const int leaves[3] = {64, 128, 256}, smls[2] = {25, 40};
    hsdc_params p0 = c->prm, best = c->prm;
    double tbest = 1e300;
    int t = 0;
    if (hipMemcpy(kd, d_d, sizeof(double) * tot, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(ke, d_e, sizeof(double) * tot, hipMemcpyDeviceToDevice) != hipSuccess)
        rc = HSDC_ERR_HIP;
    for (int a = 0; a < 2 && rc == 0; a++)
        for (int b = 0; b < 3 && rc == 0; b++, t++) {
            hsdc_params q = p0;
            q.smlsiz = smls[a];
            q.hss_leaf = leaves[b];
            q.batched = 1;
            q.profile = 0;
            c->prm = q;
            // one untimed solve brings the buffers to size, the second is timed
            for (int rep = 0; rep < 2 && rc == 0; rep++) {
                if (hipMemcpy(d_d, kd, sizeof(double) * tot, hipMemcpyDeviceToDevice) != hipSuccess ||
                    hipMemcpy(d_e, ke, sizeof(double) * tot, hipMemcpyDeviceToDevice) != hipSuccess) {
                    rc = HSDC_ERR_HIP;
                    break;
                }
                rc = hsdc_stedc_batched_dev(c, nprob, ns, offs, d_d, d_e, d_z, zoffs);
            }
            if (rc) break;
            const double ms = c->st.ms_total;
            if (times) times[t] = ms;
            if (ms < tbest) { tbest = ms; best = q; }
        }
    (void)hipMemcpy(d_d, kd, sizeof(double) * tot, hipMemcpyDeviceToDevice);
    (void)hipMemcpy(d_e, ke, sizeof(double) * tot, hipMemcpyDeviceToDevice);
    (void)hipFree(keep.p);
    c->prm = p0;
    if (rc) return rc;
    c->prm.smlsiz = best.smlsiz;
    c->prm.hss_leaf = best.hss_leaf;
    if (best_leaf) *best_leaf = best.hss_leaf;
    if (best_smlsiz) *best_smlsiz = best.smlsiz;
    return 0;
}

// one merge of order n held at global index 0 of the work arrays
static int single_merge_setup(hsdc_ctx *c, int n, int n1, double rho, Merge **dm, int **dchunk)
{
    CK(alloc_work(c, std::max(n, 1)));
    gate_reset(c);
    CK(ensure(c->probs, 128));
    c->tabkey.clear();  // the kept tables of a device-input solve are overwritten
    c->trc_valid = false;  // and the trace of the last solve with them
    c->dtrc = nullptr;
    Merge M;
    M.s = 0; M.n = n; M.n1 = n1; M.prob = 0; M.lvl = 0; M.rho = rho;
    int ch[2] = {0, cdiv(n, WGS)};
    char *dp = (char *)c->probs.p;
    HK(hipMemcpy(dp, &M, sizeof(Merge), hipMemcpyHostToDevice));
    HK(hipMemcpy(dp + 64, ch, sizeof(ch), hipMemcpyHostToDevice));
    unsigned long long zero[4] = {0, 0, 0, 0};
    HK(hipMemcpy(c->W.maux, zero, sizeof(zero), hipMemcpyHostToDevice));
    *dm = (Merge *)dp;
    *dchunk = (int *)(dp + 64);
    return 0;
}

int hsdc_deflate(hsdc_ctx *c, int n, int n1, double *d, double *z, int *indxq, double rho, int hss_min_k,
                 int *k, double *rho_out, double *dlamda, double *w, int *indxp, int *indx, int *indxc,
                 int *coltyp, int *nrot, int *rot_pj, int *rot_nj, double *rot_c, double *rot_s)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (n < 0) return -2;
    if ((1 < n / 2 ? 1 : n / 2) > n1 || n / 2 < n1) return -3;
    if (n == 0) { *k = 0; return 0; }
    Merge *dm;
    int *dch;
    CK(single_merge_setup(c, n, n1, rho, &dm, &dch));
    Work W = c->W;
    CK(ensure(c->mo, sizeof(MergeOut)));
    MergeOut *dmo = (MergeOut *)c->mo.p;
    HK(hipMemcpy(W.Dv[0], d, sizeof(double) * n, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.z, z, sizeof(double) * n, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.indxq, indxq, sizeof(int) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_prep, dim3(cdiv(n, WGS)), dim3(WGS), 0, c->stream, (const Merge *)dm, (const int *)dch, 1,
                       (const Prob *)nullptr, W, 0);
    hipLaunchKernelGGL(k_mergepath, dim3(cdiv(n, WGS)), dim3(WGS), 0, c->stream, (const Merge *)dm,
                       (const int *)dch, 1, W);
    const int ds_cap = n <= 4096 ? n : 0;
    if (n >= 2048)
        hipLaunchKernelGGL(k_deflate<1024>, dim3(1), dim3(1024), sizeof(double) * ds_cap, c->stream,
                           (const Merge *)dm, W, 0, dmo, hss_min_k, ds_cap);
    else
        hipLaunchKernelGGL(k_deflate<256>, dim3(1), dim3(WGS), sizeof(double) * ds_cap, c->stream,
                           (const Merge *)dm, W, 0, dmo, hss_min_k, ds_cap);
    HK(hipGetLastError());
    MergeOut o;
    HK(hipMemcpy(&o, dmo, sizeof(MergeOut), hipMemcpyDeviceToHost));
    *k = o.k;
    *rho_out = o.rho;
    *nrot = o.nrot;
    HK(hipMemcpy(indx, W.indx, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int> iq2(n);
    HK(hipMemcpy(iq2.data(), W.indxq2, sizeof(int) * n, hipMemcpyDeviceToHost));
    memcpy(indxq, iq2.data(), sizeof(int) * n);
    if (o.flags) {
        // everything deflated: DLAMDA = D in sorted order (mdlaed2.f90:273-285)
        HK(hipMemcpy(dlamda, W.dlamda, sizeof(double) * n, hipMemcpyDeviceToHost));
        std::vector<double> dl(n);
        for (int j = 0; j < n; j++) dl[j] = d[indx[j] - 1];
        memcpy(dlamda, dl.data(), sizeof(double) * n);
        memcpy(d, dl.data(), sizeof(double) * n);
        return 0;
    }
    HK(hipMemcpy(dlamda, W.dlamda, sizeof(double) * o.k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(w, W.w, sizeof(double) * o.k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(indxp, W.indxp, sizeof(int) * n, hipMemcpyDeviceToHost));
    HK(hipMemcpy(indxc, W.indxc, sizeof(int) * n, hipMemcpyDeviceToHost));
    HK(hipMemcpy(coltyp, W.coltyp, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int t = 0; t < 4; t++) coltyp[t] = o.ctot[t];
    if (o.nrot) {
        HK(hipMemcpy(rot_pj, W.rot_pj, sizeof(int) * o.nrot, hipMemcpyDeviceToHost));
        HK(hipMemcpy(rot_nj, W.rot_nj, sizeof(int) * o.nrot, hipMemcpyDeviceToHost));
        HK(hipMemcpy(rot_c, W.rot_c, sizeof(double) * o.nrot, hipMemcpyDeviceToHost));
        HK(hipMemcpy(rot_s, W.rot_s, sizeof(double) * o.nrot, hipMemcpyDeviceToHost));
    }
    if (o.k < n) HK(hipMemcpy(d + o.k, W.ddef + o.k, sizeof(double) * (n - o.k), hipMemcpyDeviceToHost));
    return 0;
}

int hsdc_secular(hsdc_ctx *c, int k, const double *dlamda, const double *w, double rho, double *lam, int *org,
                 double *tau, int *niter)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (k < 0) return -2;
    if (k == 0) return 0;
    Merge *dm;
    int *dch;
    CK(single_merge_setup(c, k, k / 2, rho, &dm, &dch));
    Work W = c->W;
    CK(ensure(c->mo, al64(sizeof(MergeOut)) + 64));
    MergeOut o;
    memset(&o, 0, sizeof(o));
    o.k = k; o.rho = rho; o.hss = 0; o.flags = 0;
    MergeOut *dmo = (MergeOut *)c->mo.p;
    int *dinfo = (int *)((char *)c->mo.p + al64(sizeof(MergeOut)));
    int zero = 0;
    int bs[2] = {0, cdiv(k, SEC_WG / sec_team(k))};
    HK(hipMemcpy(dmo, &o, sizeof(o), hipMemcpyHostToDevice));
    HK(hipMemcpy(dinfo, &zero, sizeof(int), hipMemcpyHostToDevice));
    HK(hipMemcpy(dch, bs, sizeof(bs), hipMemcpyHostToDevice));
    HK(hipMemcpy(W.dlamda, dlamda, sizeof(double) * k, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.w, w, sizeof(double) * k, hipMemcpyHostToDevice));
    CK(sec_setup(c));
    if (c->sec_pre)
        hipLaunchKernelGGL(k_secstart, dim3(bs[1]), dim3(SEC_WG), 0, c->stream, (const Merge *)dm,
                           (const int *)dch, 1, (const MergeOut *)dmo, W, c->dsec);
    hipLaunchKernelGGL(k_secular, dim3(bs[1]), dim3(SEC_WG), 0, c->stream, (const Merge *)dm, (const int *)dch, 1,
                       (const MergeOut *)dmo, W, c->prm.maxit, dinfo, c->dsec, c->dlvlmax);
    HK(hipGetLastError());
    int info = 0;
    HK(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
    HK(hipMemcpy(lam, W.lam, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(tau, W.tau, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(org, W.org, sizeof(int) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(niter, W.nit, sizeof(int) * k, hipMemcpyDeviceToHost));
    CK(sec_collect(c));
    return info ? 1 : 0;
}

int hsdc_laed3(hsdc_ctx *c, int k, int n, int n1, double *d, double *q, int ldq, double rho, double *dlamda,
               const double *q2, const int *indx, const int *ctot, double *w)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (k < 0) return -1;
    if (n < k) return -2;
    if (ldq < (n > 1 ? n : 1)) return -6;
    if (k == 0) return 0;
    Merge *dm;
    int *dch;
    CK(single_merge_setup(c, n, n1, rho, &dm, &dch));
    Work W = c->W;
    hipStream_t s = c->stream;
    const int n2 = n - n1, n12 = ctot[0] + ctot[1], n23 = ctot[1] + ctot[2];
    const long q2len = (long)n1 * n12 + (long)n2 * n23;
    CK(ensure(c->mo, al64(sizeof(MergeOut)) + 128));
    MergeOut o;
    memset(&o, 0, sizeof(o));
    o.k = k; o.rho = rho; o.hss = 0;
    for (int t = 0; t < 4; t++) o.ctot[t] = ctot[t];
    MergeOut *dmo = (MergeOut *)c->mo.p;
    int *dinfo = (int *)((char *)c->mo.p + al64(sizeof(MergeOut)));
    int *dtab = dinfo + 16;  // block table {0, blocks} and chunk table {0, chunks}
    int zero = 0, tabs[4] = {0, cdiv(k, SEC_WG / sec_team(k)), 0, cdiv(k, WGS)};
    HK(hipMemcpy(dmo, &o, sizeof(o), hipMemcpyHostToDevice));
    HK(hipMemcpy(dinfo, &zero, sizeof(int), hipMemcpyHostToDevice));
    HK(hipMemcpy(dtab, tabs, sizeof(tabs), hipMemcpyHostToDevice));
    HK(hipMemcpy(W.dlamda, dlamda, sizeof(double) * k, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.w, w, sizeof(double) * k, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.indxc, indx, sizeof(int) * k, hipMemcpyHostToDevice));
    CK(ensure(c->sb, sizeof(double) * (size_t)k * k + 64));
    CK(ensure(c->xb, sizeof(double) * std::max(q2len, 1L)));
    CK(ensure(c->yb, sizeof(double) * (size_t)n * k));
    CK(ensure(c->lvl, 64));
    long zoff = 0;
    HK(hipMemcpy(c->lvl.p, &zoff, sizeof(long), hipMemcpyHostToDevice));
    if (q2len) HK(hipMemcpy(c->xb.p, q2, sizeof(double) * q2len, hipMemcpyHostToDevice));
    CK(sec_setup(c));
    if (c->sec_pre)
        hipLaunchKernelGGL(k_secstart, dim3(tabs[1]), dim3(SEC_WG), 0, s, (const Merge *)dm, (const int *)dtab,
                           1, (const MergeOut *)dmo, W, c->dsec);
    hipLaunchKernelGGL(k_secular, dim3(tabs[1]), dim3(SEC_WG), 0, s, (const Merge *)dm, (const int *)dtab, 1,
                       (const MergeOut *)dmo, W, c->prm.maxit, dinfo, c->dsec, c->dlvlmax);
    hipLaunchKernelGGL(k_loewner, dim3(tabs[3]), dim3(WGS), 0, s, (const Merge *)dm, (const int *)(dtab + 2), 1,
                       (const MergeOut *)dmo, W);
    hipLaunchKernelGGL(k_sbuild, dim3(tabs[3]), dim3(WGS), 0, s, (const Merge *)dm, (const int *)(dtab + 2), 1,
                       (const MergeOut *)dmo, (const long *)c->lvl.p, W, (double *)c->sb.p);
    // the two GEMMs of DLAED3 with the packed Q2 (types 1,2 rows on top, 2,3 below)
    const double *S = (const double *)c->sb.p, *Q2 = (const double *)c->xb.p;
    double *Y = (double *)c->yb.p;
    BG g[2];
    g[0].A = Q2; g[0].lda = std::max(n1, 1); g[0].B = S; g[0].ldb = k; g[0].C = Y; g[0].ldc = n;
    g[0].nrow = n1; g[0].n = k; g[0].k = n12; g[0].beta = 0; g[0].btrans = 1;
    g[1].A = Q2 + (long)n1 * n12; g[1].lda = std::max(n2, 1); g[1].B = S + (long)ctot[0] * k; g[1].ldb = k;
    g[1].C = Y + n1; g[1].ldc = n; g[1].nrow = n2; g[1].n = k; g[1].k = n23; g[1].beta = 0; g[1].btrans = 1;
    for (int t = 0; t < 2; t++) { g[t].A2 = nullptr; g[t].B2 = nullptr; g[t].lda2 = g[t].ldb2 = 0; g[t].k2 = 0; }
    CK(ensure(c->tabs, sizeof(g)));
    HK(hipMemcpy(c->tabs.p, g, sizeof(g), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_bgemm, dim3(cdiv(std::max(n1, n2), 128), 2, cdiv(k, 64)), dim3(WGS), 0, s,
                       (const BG *)c->tabs.p);
    HK(hipGetLastError());
    int info = 0;
    HK(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
    HK(hipMemcpy(d, W.lam, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy2D(q, sizeof(double) * ldq, Y, sizeof(double) * n, sizeof(double) * n, k, hipMemcpyDeviceToHost));
    if (k >= 3) HK(hipMemcpy(w, W.wnew, sizeof(double) * k, hipMemcpyDeviceToHost));
    return info ? 1 : 0;
}

int hsdc_loewner(hsdc_ctx *c, int k, double rho, double *dlamda, double *z, double *d, double *difl,
                 double *difr, double *alpha)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (k < 3) return -2;
    Merge *dm;
    int *dch;
    CK(single_merge_setup(c, k, k / 2, rho, &dm, &dch));
    Work W = c->W;
    hipStream_t s = c->stream;
    CK(ensure(c->mo, al64(sizeof(MergeOut)) + 128));
    MergeOut o;
    memset(&o, 0, sizeof(o));
    o.k = k; o.rho = rho; o.hss = 1;
    MergeOut *dmo = (MergeOut *)c->mo.p;
    int *dinfo = (int *)((char *)c->mo.p + al64(sizeof(MergeOut)));
    int *dtab = dinfo + 16;
    int zero = 0, tabs[4] = {0, cdiv(k, SEC_WG / sec_team(k)), 0, cdiv(k, WGS)};
    HK(hipMemcpy(dmo, &o, sizeof(o), hipMemcpyHostToDevice));
    HK(hipMemcpy(dinfo, &zero, sizeof(int), hipMemcpyHostToDevice));
    HK(hipMemcpy(dtab, tabs, sizeof(tabs), hipMemcpyHostToDevice));
    HK(hipMemcpy(W.dlamda, dlamda, sizeof(double) * k, hipMemcpyHostToDevice));
    HK(hipMemcpy(W.w, z, sizeof(double) * k, hipMemcpyHostToDevice));
    CK(sec_setup(c));
    if (c->sec_pre)
        hipLaunchKernelGGL(k_secstart, dim3(tabs[1]), dim3(SEC_WG), 0, s, (const Merge *)dm, (const int *)dtab,
                           1, (const MergeOut *)dmo, W, c->dsec);
    hipLaunchKernelGGL(k_secular, dim3(tabs[1]), dim3(SEC_WG), 0, s, (const Merge *)dm, (const int *)dtab, 1,
                       (const MergeOut *)dmo, W, c->prm.maxit, dinfo, c->dsec, c->dlvlmax);
    hipLaunchKernelGGL(k_loewner, dim3(tabs[3]), dim3(WGS), 0, s, (const Merge *)dm, (const int *)(dtab + 2), 1,
                       (const MergeOut *)dmo, W);
    hipLaunchKernelGGL(k_alpha, dim3(tabs[3]), dim3(WGS), 0, s, (const Merge *)dm, (const int *)(dtab + 2), 1,
                       (const MergeOut *)dmo, W);
    HK(hipGetLastError());
    int info = 0;
    HK(hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost));
    HK(hipMemcpy(d, W.lam, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(z, W.wnew, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(difl, W.difl, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(difr, W.difr, sizeof(double) * k, hipMemcpyDeviceToHost));
    HK(hipMemcpy(alpha, W.alpha, sizeof(double) * k, hipMemcpyDeviceToHost));
    return info ? 1 : 0;
}

// q (nrow x k) := q * H, or for left q (k x nrow) := H * q, computed as the transposed
// product on a host-transposed copy
static int hss_one(hsdc_ctx *c, int sq, int k, int ni, double tol, const double *f, const double *d, const double *u,
                   const double *v, const double *difl, const double *difr, double *q, int ldq, int nrow, int *ranks,
                   bool left = false)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (k < 1) return -2;
    if (ni < 8) return -3;
    if (nrow < 0) return -13;
    if (ldq < std::max(left ? k : nrow, 1)) return -12;
    CK(alloc_work(c, k));
    Work W = c->W;
    hipStream_t s = c->stream;
    const double *src[6] = {f, d, u, v, difl, difr};
    double *dst[6] = {W.dlamda, W.lam, W.wnew, W.alpha, W.difl, W.difr};
    for (int i = 0; i < 6; i++) HK(hipMemcpy(dst[i], src[i], sizeof(double) * k, hipMemcpyHostToDevice));
    c->evc_gen.assign(6 * (size_t)k, 0.0);
    for (int i = 0; i < 6; i++) memcpy(c->evc_gen.data() + (size_t)i * k, src[i], sizeof(double) * k);
    c->evc_k = k; c->evc_ni = ni; c->evc_tol = tol; c->evc_sq = sq;
    CK(ensure(c->xb, sizeof(double) * (size_t)std::max(nrow, 1) * k));
    CK(ensure(c->yb, sizeof(double) * (size_t)std::max(nrow, 1) * k));
    if (left) {  // q is k x nrow: staged as is, transposed on the device into X^T (nrow x k)
        CK(ensure(c->tb, sizeof(double) * (size_t)std::max(nrow, 1) * k));
        if (nrow > 0) {
            HK(hipMemcpy2D(c->tb.p, sizeof(double) * k, q, sizeof(double) * ldq, sizeof(double) * k, nrow,
                           hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_transpose, dim3(cdiv(k, 64), cdiv(nrow, 64)), dim3(WGS), 0, s, k, nrow,
                               (const double *)c->tb.p, (long)k, (double *)c->xb.p, (long)nrow);
            HK(hipGetLastError());
        }
    } else {
        HK(hipMemcpy2D(c->xb.p, sizeof(double) * nrow, q, sizeof(double) * ldq, sizeof(double) * nrow, k,
                       hipMemcpyHostToDevice));
    }
    HssSpec sp;
    sp.K = k;
    sp.FF = W.dlamda; sp.D0 = W.lam; sp.U0 = W.wnew; sp.V0 = W.alpha; sp.DIFL = W.difl; sp.DIFR = W.difr;
    sp.X = (const double *)c->xb.p; sp.ldx = nrow; sp.Y = (double *)c->yb.p; sp.ldy = nrow; sp.nrow = nrow;
    sp.sq = sq;
    std::vector<HssSpec> v1(1, sp);
    std::vector<HssPlanM> keep;
    CK(hss_apply(c, v1, ni, tol, &keep, left));
    if (left) {
        if (nrow > 0) {  // (X^T H^T)^T back to k x nrow on the device (the cb staging is free again)
            CK(ensure(c->tb, sizeof(double) * (size_t)nrow * k));
            hipLaunchKernelGGL(k_transpose, dim3(cdiv(nrow, 64), cdiv(k, 64)), dim3(WGS), 0, s, nrow, k,
                               (const double *)c->yb.p, (long)nrow, (double *)c->tb.p, (long)k);
            HK(hipGetLastError());
            HK(hipMemcpy2D(q, sizeof(double) * ldq, c->tb.p, sizeof(double) * k, sizeof(double) * k, nrow,
                           hipMemcpyDeviceToHost));
        }
    } else {
        HK(hipMemcpy2D(q, sizeof(double) * ldq, c->yb.p, sizeof(double) * nrow, sizeof(double) * nrow, k,
                       hipMemcpyDeviceToHost));
    }
    if (ranks) {
        const HssPlanM &Q = keep[0];
        if (!Q.ok) {
            ranks[0] = -1;
        } else {
            for (int i = 0; i < Q.T.ltr; i++) {
                ranks[2 * i] = Q.usz[2 * i + 1];
                ranks[2 * i + 1] = Q.vsz[2 * i];
            }
        }
    }
    return 0;
}

int hsdc_hss_evc(hsdc_ctx *c, int k, int ni, double tol, const double *f, const double *d, const double *u,
                 const double *v, const double *difl, const double *difr, double *q, int ldq, int nrow, int *ranks)
{
    return hss_one(c, 0, k, ni, tol, f, d, u, v, difl, difr, q, ldq, nrow, ranks);
}

int hsdc_hss_svc(hsdc_ctx *c, int k, int ni, double tol, const double *f, const double *d, const double *u,
                 const double *v, const double *difl, const double *difr, double *q, int ldq, int nrow, int *ranks)
{
    return hss_one(c, 1, k, ni, tol, f, d, u, v, difl, difr, q, ldq, nrow, ranks);
}

int hsdc_hss_cmm(hsdc_ctx *c, char side, int squared, int k, int ni, double tol, const double *f, const double *d,
                 const double *u, const double *v, const double *difl, const double *difr, double *x, int ldx,
                 int ncol, int *ranks)
{
    const bool left = side == 'L' || side == 'l';
    if (!c) return HSDC_ERR_INTERNAL;
    if (!left && side != 'R' && side != 'r') return -2;
    if (squared != 0 && squared != 1) return -3;
    if (k < 1) return -4;
    if (ni < 8) return -5;
    if (ncol < 0) return -15;
    if (ldx < std::max(left ? k : ncol, 1)) return -14;
    if (ncol == 0) {  // nothing to multiply: no HSS form is built, and none is left to expand
        c->evc_k = 0;
        return 0;
    }
    int rc = hss_one(c, squared, k, ni, tol, f, d, u, v, difl, difr, x, ldx, ncol, ranks, left);
    // hss_one numbers k, ni, ldx from 2 on; here two arguments come first
    return rc < 0 && rc > HSDC_ERR_HIP ? rc - 2 : rc;
}


// CauchyHssMM (cauchyhssmm.f90:1-123): B (n x n, ldb) := A * B for the general
// Cauchy-like A(i,j) = u(i) v(j) / (d(i) - f(j)).  The HSS form is built by
// cauchy2hss's traversal (k_cauchy2hss, one workgroup), then H * B runs as the
// transposed product (B^T H^T)^T through the batched multiply, ncol-wide column
// blocks at a time (at least 1024).  A tree of one leaf, more than 512 leaves or a
// full generator arena take the dense product instead.
int hsdc_cauchyhssmm(hsdc_ctx *c, const double *d, const double *f, const double *u, const double *v, int n, int ni,
                     double tol, int ncol, double *b, int ldb, double *time, int *ranks)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (!d) return -1;
    if (!f) return -2;
    if (!u) return -3;
    if (!v) return -4;
    if (n < 0) return -5;
    if (ni < 8) return -6;
    if (!(tol >= 0.0)) return -7;
    if (ncol < 1) return -8;
    if (ldb < std::max(n, 1)) return -10;
    if (time) { time[0] = 0.0; time[1] = 0.0; }
    c->evc_k = 0;  // hsdc_hss_dense expands only the eigenvector / SVD forms
    if (n == 0) return 0;
    // A must be finite and defined: no d(i) may equal an f(j)
    {
        std::vector<double> ds(d, d + n), fs(f, f + n);
        for (int i = 0; i < n; i++)
            if (!std::isfinite(d[i]) || !std::isfinite(f[i]) || !std::isfinite(u[i]) || !std::isfinite(v[i]))
                return -1;
        std::sort(ds.begin(), ds.end());
        std::sort(fs.begin(), fs.end());
        for (size_t i = 0, j = 0; i < ds.size() && j < fs.size();) {
            if (ds[i] == fs[j]) return -1;
            if (ds[i] < fs[j]) i++; else j++;
        }
    }
    if (!have_device()) return HSDC_ERR_NODEVICE;
    hipStream_t s = c->stream;
    if (!c->ev0) {
        HK(hipEventCreate(&c->ev0));
        HK(hipEventCreate(&c->ev1));
    }
    HssTree T;
    const bool tree = hss_make_tree(T, n, ni);
    const size_t nn = (size_t)n;
    // device storage: inputs, working generators and scratch (16 n), the matrix B
    CK(ensure(c->xb, sizeof(double) * 16 * nn));
    double *gd = (double *)c->xb.p;
    const double *hsrc[4] = {d, f, u, v};
    for (int i = 0; i < 4; i++) HK(hipMemcpyAsync(gd + i * nn, hsrc[i], sizeof(double) * nn, hipMemcpyHostToDevice, s));
    CK(ensure(c->q, sizeof(double) * nn * nn));
    double *dB = (double *)c->q.p;
    HK(hipMemcpy2DAsync(dB, sizeof(double) * nn, b, sizeof(double) * ldb, sizeof(double) * nn, nn,
                        hipMemcpyHostToDevice, s));
    std::vector<HssPlanM> P(1);
    HssPlanM &Q = P[0];
    Q.ok = false;
    float ms_build = 0.0f;
    if (tree) {
        const int ltr = T.ltr;
        std::vector<int> ch(2 * ltr), rng(2 * ltr);
        std::vector<long> pd(ltr, 0);
        long pdt = 0;
        int it = 0;
        for (int i = 1; i <= ltr; i++) {
            ch[2 * (i - 1)] = T.c1[i];
            ch[2 * (i - 1) + 1] = T.c2[i];
            if (T.c1[i] == 0) {
                rng[2 * (i - 1)] = T.leaf_r0[i] + 1;
                rng[2 * (i - 1) + 1] = T.leaf_r0[i] + T.m[it];
                pd[i - 1] = pdt;
                pdt += (long)T.m[it] * T.m[it];
                it++;
            } else {
                rng[2 * (i - 1)] = rng[2 * (T.c1[i] - 1)];
                rng[2 * (i - 1) + 1] = rng[2 * (T.c2[i] - 1) + 1];
            }
        }
        // tables: ch | rng | pd | usz vsz bsz (2 each) dsz | pu pv pb | lsr lsc | err | nswap
        const long acap = (long)12 * std::max(ni, 128) * n;
        size_t o0 = 0, o1 = al64(o0 + sizeof(int) * 2 * ltr), o2 = al64(o1 + sizeof(int) * 2 * ltr),
               o3 = al64(o2 + sizeof(long) * ltr), o4 = al64(o3 + sizeof(int) * 7 * ltr),
               o5 = al64(o4 + sizeof(long) * 3 * ltr), o6 = al64(o5 + sizeof(int) * 2 * (T.lvl + 8)),
               o7 = al64(o6 + 64), o8 = al64(o7 + sizeof(double) * std::max(pdt, 1L)),
               o9 = o8 + sizeof(double) * acap;
        CK(ensure(c->hss, o9));
        char *hb = (char *)c->hss.p;
        std::vector<char> tab(o3, 0);
        memcpy(tab.data() + o0, ch.data(), sizeof(int) * 2 * ltr);
        memcpy(tab.data() + o1, rng.data(), sizeof(int) * 2 * ltr);
        memcpy(tab.data() + o2, pd.data(), sizeof(long) * ltr);
        HK(hipMemcpyAsync(hb, tab.data(), o3, hipMemcpyHostToDevice, s));
        HK(hipMemsetAsync(hb + o3, 0, o7 - o3, s));
        C2H C;
        C.n = n; C.ltr = ltr; C.lvl = T.lvl; C.tol = tol;
        C.ch = (const int *)(hb + o0); C.rng = (const int *)(hb + o1); C.pd = (const long *)(hb + o2);
        int *sz = (int *)(hb + o3);
        C.usz = sz; C.vsz = sz + 2 * ltr; C.bsz = sz + 4 * ltr; C.dsz = sz + 6 * ltr;
        long *po = (long *)(hb + o4);
        C.pu = po; C.pv = po + ltr; C.pb = po + 2 * ltr;
        C.lsr = (int *)(hb + o5); C.lsc = C.lsr + T.lvl + 8;
        C.err = (int *)(hb + o6); C.nswap = (unsigned long long *)(hb + o6 + 8);
        C.dd = (double *)(hb + o7); C.arena = (double *)(hb + o8); C.arena_cap = acap;
        C.D = gd + 4 * nn; C.F = gd + 5 * nn; C.U = gd + 6 * nn; C.V = gd + 7 * nn;
        for (int i = 0; i < 4; i++)
            HK(hipMemcpyAsync(gd + (4 + i) * nn, gd + i * nn, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
        C.Di = gd + 8 * nn; C.Fi = gd + 9 * nn; C.Ui = gd + 10 * nn; C.Vi = gd + 11 * nn; C.Z = gd + 12 * nn;
        C.W = gd + 13 * nn;
        C.PL = (int *)(gd + 14 * nn); C.PU = (int *)(gd + 15 * nn);
        HK(hipEventRecord(c->ev0, s));
        hipLaunchKernelGGL(k_cauchy2hss, dim3(1), dim3(1024), 0, s, C);
        HK(hipGetLastError());
        HK(hipEventRecord(c->ev1, s));
        // generator directory back to the host
        std::vector<int> szh(7 * ltr);
        std::vector<long> poh(3 * ltr);
        int err = 0;
        HK(hipMemcpyAsync(szh.data(), sz, sizeof(int) * 7 * ltr, hipMemcpyDeviceToHost, s));
        HK(hipMemcpyAsync(poh.data(), po, sizeof(long) * 3 * ltr, hipMemcpyDeviceToHost, s));
        HK(hipMemcpyAsync(&err, C.err, sizeof(int), hipMemcpyDeviceToHost, s));
        HK(hipStreamSynchronize(s));
        (void)hipEventElapsedTime(&ms_build, c->ev0, c->ev1);
        if (!err) {
            Q.T = T;
            Q.ok = true;
            Q.usz.assign(szh.begin(), szh.begin() + 2 * ltr);
            Q.vsz.assign(szh.begin() + 2 * ltr, szh.begin() + 4 * ltr);
            Q.bsz.assign(szh.begin() + 4 * ltr, szh.begin() + 6 * ltr);
            Q.dsz.assign(szh.begin() + 6 * ltr, szh.begin() + 7 * ltr);
            Q.pu.assign(poh.begin(), poh.begin() + ltr);
            Q.pv.assign(poh.begin() + ltr, poh.begin() + 2 * ltr);
            Q.pb.assign(poh.begin() + 2 * ltr, poh.begin() + 3 * ltr);
            Q.pd = pd;
            memset(&Q.H, 0, sizeof(Q.H));
            Q.H.arena = C.arena;
            Q.H.dd = C.dd;
            Q.H.K = n;
            if (ranks)
                for (int i = 0; i < ltr; i++) {
                    ranks[2 * i] = Q.usz[2 * i + 1];
                    ranks[2 * i + 1] = Q.vsz[2 * i];
                }
        }
    }
    if (!Q.ok && ranks) ranks[0] = -1;
    // B := A B, in column blocks of nb: B_blk^T (nb x n) on the device, times H^T
    const int nb = std::min(n, std::max(ncol, 1024));
    CK(ensure(c->yb, sizeof(double) * nn * nb));
    CK(ensure(c->tb, sizeof(double) * nn * nb));
    double *Xt = (double *)c->tb.p, *Yt = (double *)c->yb.p;
    if (!Q.ok) {
        CK(ensure(c->cb, sizeof(double) * nn * nn));
        hipLaunchKernelGGL(k_cfill_general, dim3(std::min(cdiv((long)n * n, WGS), 8192)), dim3(WGS), 0, s, n,
                           (const double *)gd, (const double *)(gd + nn), (const double *)(gd + 2 * nn),
                           (const double *)(gd + 3 * nn), (double *)c->cb.p);
    }
    HK(hipEventRecord(c->ev0, s));
    for (int c0 = 0; c0 < n; c0 += nb) {
        const int w = std::min(nb, n - c0);
        double *Bb = dB + (size_t)c0 * nn;
        if (Q.ok) {
            hipLaunchKernelGGL(k_transpose, dim3(cdiv(n, 64), cdiv(w, 64)), dim3(WGS), 0, s, n, w, (const double *)Bb,
                               (long)n, Xt, (long)w);
            HssSpec sp;
            sp.K = n; sp.X = Xt; sp.ldx = w; sp.Y = Yt; sp.ldy = w; sp.nrow = w;
            std::vector<HssSpec> v1(1, sp);
            CK(hss_mult(c, v1, P, true));
            hipLaunchKernelGGL(k_transpose, dim3(cdiv(w, 64), cdiv(n, 64)), dim3(WGS), 0, s, w, n, (const double *)Yt,
                               (long)w, Bb, (long)n);
        } else {
            hipLaunchKernelGGL(k_dgemm_nn, dim3(cdiv(n, 64), cdiv(w, 64)), dim3(WGS), 0, s, n, w, n,
                               (const double *)c->cb.p, (long)n, (const double *)Bb, (long)n, Yt, (long)n);
            HK(hipMemcpyAsync(Bb, Yt, sizeof(double) * nn * w, hipMemcpyDeviceToDevice, s));
        }
        HK(hipGetLastError());
    }
    HK(hipEventRecord(c->ev1, s));
    HK(hipMemcpy2DAsync(b, sizeof(double) * ldb, dB, sizeof(double) * nn, sizeof(double) * nn, nn,
                        hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    float ms_mm = 0.0f;
    (void)hipEventElapsedTime(&ms_mm, c->ev0, c->ev1);
    if (time) { time[0] = ms_build * 1e-3; time[1] = ms_mm * 1e-3; }
    return 0;
}

// HSS form of the dense n x n matrix dA (device, ld n) by truncated SVDs (mat2hssvd of
// DHSSMM / DHSSMML, ConstructHssd.f90:371-696), depth by depth (the kernels of
// hsdc_mat2hss.hip): the generators into c->hss (arena and leaf blocks) and their
// directory into Q.  Q.ok stays false for a tree of one leaf or more than 512 leaves, a
// Jacobi sweep limit or a full arena; the caller then multiplies densely.  ranks (may be
// null; 2 per node): the ranks of each node's U and V.
// rsvd: randomized range finding (mat2hssvdr) with the Omega stream of seed
static int mat2hss_build(hsdc_ctx *c, const double *dA, int n, int ni, double tol, HssPlanM &Q, int *ranks,
                         float *ms_build, bool rsvd = false, unsigned long long seed = 0)
{
    hipStream_t s = c->stream;
    Q.ok = false;
    HssTree T;
    if (!hss_make_tree(T, n, ni)) return 0;
    const int ltr = T.ltr;
    std::vector<int> usz(2 * ltr, 0), vsz(2 * ltr, 0), bsz(2 * ltr, 0), dsz(ltr, 0);
    std::vector<long> pu(ltr, 0), pv(ltr, 0), pb(ltr, 0), pd(ltr, 0);
    // leaf blocks and the arena (12 max(ni, 128) n doubles, as for the Cauchy forms)
    long pdt = 0;
    {
        int it = 0;
        for (int i = 1; i <= ltr; i++)
            if (T.c1[i] == 0) { dsz[i - 1] = T.m[it]; pd[i - 1] = pdt; pdt += (long)T.m[it] * T.m[it]; it++; }
    }
    const long acap = (long)12 * std::max(ni, 128) * n;
    CK(ensure(c->hss, sizeof(double) * (size_t)(pdt + acap)));
    double *dd = (double *)c->hss.p, *arena = dd + pdt;
    long atop = 0;
    struct Act { int node, rs, rn, cs, cn; };
    std::vector<Act> act;
    {
        int it = 0;
        for (int i = 1; i <= ltr; i++)
            if (T.c1[i] == 0) { act.push_back({i, T.leaf_r0[i], T.m[it], T.leaf_r0[i], T.m[it]}); it++; }
    }
    int nr = n, nc = n;
    // C ping-pong: cb (current), misc (row-compressed and new); C starts as A
    CK(ensure(c->cb, sizeof(double) * (size_t)n * n));
    CK(ensure(c->misc, sizeof(double) * 2 * (size_t)n * n));
    double *Cc = (double *)c->cb.p, *Tr = (double *)c->misc.p, *Cn = Tr + (size_t)n * n;
    HK(hipMemcpyAsync(Cc, dA, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
    HK(hipEventRecord(c->ev0, s));
    // leaf D blocks straight from A
    {
        std::vector<M2hCopy> cp;
        for (int i = 1; i <= ltr; i++)
            if (T.c1[i] == 0) cp.push_back({3, dsz[i - 1], dsz[i - 1], 0, T.leaf_r0[i], T.leaf_r0[i], pd[i - 1]});
        CK(ensure(c->tabs, sizeof(M2hCopy) * cp.size()));
        HK(hipMemcpy(c->tabs.p, cp.data(), sizeof(M2hCopy) * cp.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_m2h_copy, dim3(16, (unsigned)cp.size()), dim3(WGS), 0, s, (const M2hCopy *)c->tabs.p,
                           (const M2hJob *)nullptr, (const double *)nullptr, (const double *)nullptr, 0L, dA, (long)n,
                           arena, dd);
        HK(hipGetLastError());
    }
    bool fail = false;
    for (int LL = T.lvl; LL >= 1 && !fail; LL--) {
        const std::vector<int> &nds = T.lev[LL];
        const int nn = (int)nds.size();
        std::vector<int> a0(nn), a1(nn);
        std::vector<M2hJob> jobs(2 * nn), sk(2 * nn);
        long so = 0, bo = 0;
        bool any_sketch = false;
        for (int q = 0; q < nn; q++) {
            const int nd = nds[q];
            const int f = T.c1[nd] ? T.c1[nd] : nd, l = T.c1[nd] ? T.c2[nd] : nd;
            a0[q] = a1[q] = -1;
            for (int t = 0; t < (int)act.size(); t++) {
                if (act[t].node == f) a0[q] = t;
                if (act[t].node == l) a1[q] = t;
            }
            if (a0[q] < 0 || a1[q] < 0) return HSDC_ERR_INTERNAL;
            const int rs = act[a0[q]].rs, rn = act[a1[q]].rs + act[a1[q]].rn - rs;
            const int cs = act[a0[q]].cs, cn = act[a1[q]].cs + act[a1[q]].cn - cs;
            for (int sd = 0; sd < 2; sd++) {
                M2hJob &J = jobs[2 * q + sd];
                J.rs = rs; J.rn = rn; J.cs = cs; J.cn = cn;
                J.m = sd == 0 ? rn : cn;
                J.w = sd == 0 ? nc - cn : nr - rn;
                J.toff = so;
                J.boff = bo;
                bo += (long)J.m * J.m + 2L * J.m;
                M2hJob &K = sk[2 * q + sd];
                K = J;
                if (rsvd && J.w > 50 && J.w > J.m + 10) {  // mat2hssvdr's sampling test
                    K.w = J.m + 10;
                    K.toff = so + (long)J.w * J.m;
                    so += (long)J.w * J.m + (long)K.w * J.m + 2L * J.m * J.m;
                    any_sketch = true;
                } else {
                    so += (long)J.w * J.m + 2L * J.m * J.m;
                }
            }
        }
        CK(ensure(c->xb, sizeof(double) * (size_t)std::max(so, 1L)));
        CK(ensure(c->yb, sizeof(double) * (size_t)std::max(bo, 1L)));
        double *S = (double *)c->xb.p, *Bs = (double *)c->yb.p;
        const size_t jb = al64(sizeof(M2hJob) * jobs.size());
        CK(ensure(c->tabs, 2 * jb + 64 + sizeof(int) * jobs.size()));
        char *tb = (char *)c->tabs.p;
        M2hJob *dj = (M2hJob *)tb, *dk = (M2hJob *)(tb + jb);
        int *derr = (int *)(tb + 2 * jb), *drank = derr + 16;
        HK(hipMemcpyAsync(dj, jobs.data(), sizeof(M2hJob) * jobs.size(), hipMemcpyHostToDevice, s));
        HK(hipMemcpyAsync(dk, sk.data(), sizeof(M2hJob) * sk.size(), hipMemcpyHostToDevice, s));
        HK(hipMemsetAsync(derr, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_m2h_gather, dim3(64, (unsigned)jobs.size()), dim3(WGS), 0, s, (const M2hJob *)dj,
                           (const double *)Cc, (long)nr, S);
        if (any_sketch)
            hipLaunchKernelGGL(k_m2h_sketch, dim3(32, (unsigned)jobs.size()), dim3(WGS), 0, s, (const M2hJob *)dj,
                               (const M2hJob *)dk, (const double *)S, S, seed);
        hipLaunchKernelGGL(k_m2h_svd, dim3((unsigned)jobs.size()), dim3(WGS), 0, s, (const M2hJob *)dk, S, Bs, tol,
                           drank, derr);
        HK(hipGetLastError());
        std::vector<int> rk(jobs.size());
        int err = 0;
        HK(hipMemcpyAsync(rk.data(), drank, sizeof(int) * rk.size(), hipMemcpyDeviceToHost, s));
        HK(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, s));
        HK(hipStreamSynchronize(s));
        if (err) { fail = true; break; }
        // generators of the level and the new active list
        std::vector<M2hCopy> cp;
        for (int q = 0; q < nn; q++) {
            const int nd = nds[q], ru = rk[2 * q], rw = rk[2 * q + 1];
            const M2hJob &Jr = jobs[2 * q], &Jc = jobs[2 * q + 1];
            usz[2 * (nd - 1)] = Jr.m; usz[2 * (nd - 1) + 1] = ru; pu[nd - 1] = atop;
            if ((long)Jr.m * ru > 0) cp.push_back({0, Jr.m, ru, 2 * q, 0, 0, atop});
            atop += (long)Jr.m * ru;
            vsz[2 * (nd - 1)] = rw; vsz[2 * (nd - 1) + 1] = Jc.m; pv[nd - 1] = atop;
            if ((long)rw * Jc.m > 0) cp.push_back({1, rw, Jc.m, 2 * q + 1, 0, 0, atop});
            atop += (long)rw * Jc.m;
        }
        std::vector<Act> nact;
        std::vector<M2hBlk> rb, cbk;
        int nnr = 0, nnc = 0;
        for (int t = 0; t < (int)act.size();) {
            int q = -1;
            for (int z = 0; z < nn; z++) if (a0[z] == t) q = z;
            Act a;
            M2hBlk br, bc;
            if (q >= 0) {
                a = {nds[q], nnr, rk[2 * q], nnc, rk[2 * q + 1]};
                br = {nnr, a.rn, q, jobs[2 * q].rs, jobs[2 * q].rn};
                bc = {nnc, a.cn, q, jobs[2 * q].cs, jobs[2 * q].cn};
                t = a1[q] + 1;
            } else {
                a = {act[t].node, nnr, act[t].rn, nnc, act[t].cn};
                br = {nnr, a.rn, -1 - t, act[t].rs, act[t].rn};
                bc = {nnc, a.cn, -1 - t, act[t].cs, act[t].cn};
                t++;
            }
            nact.push_back(a);
            if (a.rn > 0) rb.push_back(br);
            if (a.cn > 0) cbk.push_back(bc);
            nnr += a.rn;
            nnc += a.cn;
        }
        // B of every sibling pair of this depth: the compressed block between them
        for (int q = 0; q < nn; q++) {
            const int nd = nds[q], par = T.tr[nd - 1];
            const int sib = (T.c1[par] == nd) ? T.c2[par] : T.c1[par];
            int ia = -1, ib = -1;
            for (int t = 0; t < (int)nact.size(); t++) {
                if (nact[t].node == nd) ia = t;
                if (nact[t].node == sib) ib = t;
            }
            if (ia < 0 || ib < 0) return HSDC_ERR_INTERNAL;
            const int br = nact[ia].rn, bcn = nact[ib].cn;
            bsz[2 * (nd - 1)] = br; bsz[2 * (nd - 1) + 1] = bcn; pb[nd - 1] = atop;
            if ((long)br * bcn > 0) cp.push_back({2, br, bcn, 0, nact[ia].rs, nact[ib].cs, atop});
            atop += (long)br * bcn;
        }
        if (atop > acap) { fail = true; break; }
        // C' = blockdiag(U)^T C blockdiag(W)
        const size_t o1 = al64(sizeof(M2hJob) * jobs.size()), o2 = o1 + al64(sizeof(M2hBlk) * (rb.size() + 1)),
                     o3 = o2 + al64(sizeof(M2hBlk) * (cbk.size() + 1)), o4 = o3 + sizeof(M2hCopy) * (cp.size() + 1);
        CK(ensure(c->htab, o4));
        char *hb = (char *)c->htab.p;
        std::vector<char> up(o4, 0);
        memcpy(up.data(), jobs.data(), sizeof(M2hJob) * jobs.size());
        if (!rb.empty()) memcpy(up.data() + o1, rb.data(), sizeof(M2hBlk) * rb.size());
        if (!cbk.empty()) memcpy(up.data() + o2, cbk.data(), sizeof(M2hBlk) * cbk.size());
        if (!cp.empty()) memcpy(up.data() + o3, cp.data(), sizeof(M2hCopy) * cp.size());
        HK(hipMemcpyAsync(hb, up.data(), o4, hipMemcpyHostToDevice, s));
        const M2hJob *hj = (const M2hJob *)hb;
        if (nnr > 0 && nc > 0 && !rb.empty())
            hipLaunchKernelGGL(k_m2h_rows, dim3((unsigned)std::min<long>(cdiv((long)nnr * nc, WGS), 8192)), dim3(WGS), 0,
                               s, (const M2hBlk *)(hb + o1), (int)rb.size(), hj, (const double *)Bs,
                               (const double *)Cc, (long)nr, nc, Tr, (long)nnr);
        if (nnr > 0 && nnc > 0 && !cbk.empty())
            hipLaunchKernelGGL(k_m2h_cols, dim3((unsigned)std::min<long>(cdiv((long)nnr * nnc, WGS), 8192)), dim3(WGS),
                               0, s, (const M2hBlk *)(hb + o2), (int)cbk.size(), hj, (const double *)Bs,
                               (const double *)Tr, (long)nnr, nnc, Cn);
        if (!cp.empty())
            hipLaunchKernelGGL(k_m2h_copy, dim3(16, (unsigned)cp.size()), dim3(WGS), 0, s, (const M2hCopy *)(hb + o3),
                               hj, (const double *)Bs, (const double *)Cn, (long)std::max(nnr, 1), dA, (long)n, arena,
                               dd);
        HK(hipGetLastError());
        HK(hipStreamSynchronize(s));  // the host tables above are locals
        std::swap(Cc, Cn);
        nr = nnr;
        nc = nnc;
        act.swap(nact);
    }
    HK(hipEventRecord(c->ev1, s));
    HK(hipEventSynchronize(c->ev1));
    if (ms_build) (void)hipEventElapsedTime(ms_build, c->ev0, c->ev1);
    if (fail) return 0;
    Q.T = T;
    Q.ok = true;
    Q.usz = usz; Q.vsz = vsz; Q.bsz = bsz; Q.dsz = dsz;
    Q.pu = pu; Q.pv = pv; Q.pb = pb; Q.pd = pd;
    memset(&Q.H, 0, sizeof(Q.H));
    Q.H.arena = arena;
    Q.H.dd = dd;
    Q.H.K = n;
    if (ranks)
        for (int i = 0; i < ltr; i++) {
            ranks[2 * i] = usz[2 * i + 1];
            ranks[2 * i + 1] = vsz[2 * i];
        }
    return 0;
}

// Checks shared by DHSSMM and DHSSMML (DHSSMM.f90:1-125 takes LDA = M = N and a square B).
static int dhss_args(hsdc_ctx *c, const double *a, int lda, int m, int n, const double *b, int ldb, int k, double tol,
                     int ni, int ncol, bool left)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (!a) return -1;
    if (lda < std::max(1, n)) return -2;
    if (m != n) return -3;
    if (n < 0) return -4;
    if (!b) return -5;
    if (ldb < std::max(1, left ? n : k)) return -6;
    if (k < 0) return -7;
    if (!(tol >= 0.0)) return -8;
    if (ni < 1) return -10;
    if (ncol < 1) return -11;
    return 0;
}

// DHSSMM (DHSSMM.f90:1-125): b(1:n, 1:k) := A b with A (n x n) put in HSS form by
// truncated SVDs (leaf ni, tolerance tol), the product H b taken as (b^T H^T)^T in column
// blocks of at least 1024.  DHSSMML (DHSSMML.f90:2-122): b(1:k, 1:n) := b A, rows of b in
// blocks.  A is not modified (the reference uses it as workspace); prank is accepted and
// unused, as in mat2hssvd.
static int dhss_run(hsdc_ctx *c, const double *a, int lda, int n, double *b, int ldb, int k, double tol, int ni,
                    int ncol, double *time, int *ranks, bool left, bool rsvd = false, unsigned long long seed = 0)
{
    if (time) { time[0] = 0.0; time[1] = 0.0; }
    c->evc_k = 0;
    if (n == 0 || k == 0) {
        if (ranks) ranks[0] = -1;
        return 0;
    }
    if (!have_device()) return HSDC_ERR_NODEVICE;
    hipStream_t s = c->stream;
    if (!c->ev0) {
        HK(hipEventCreate(&c->ev0));
        HK(hipEventCreate(&c->ev1));
    }
    const size_t nn = (size_t)n, kk = (size_t)k;
    CK(ensure(c->zout, sizeof(double) * nn * nn));
    double *dA = (double *)c->zout.p;
    HK(hipMemcpy2DAsync(dA, sizeof(double) * nn, a, sizeof(double) * lda, sizeof(double) * nn, nn,
                        hipMemcpyHostToDevice, s));
    // b on the device: n x k (left) or k x n, leading dimension its row count
    const size_t br = left ? nn : kk, bc = left ? kk : nn;
    CK(ensure(c->q, sizeof(double) * br * bc));
    double *dB = (double *)c->q.p;
    HK(hipMemcpy2DAsync(dB, sizeof(double) * br, b, sizeof(double) * ldb, sizeof(double) * br, bc,
                        hipMemcpyHostToDevice, s));
    std::vector<HssPlanM> P(1);
    float ms_build = 0.0f;
    CK(mat2hss_build(c, dA, n, ni, tol, P[0], ranks, &ms_build, rsvd, seed));
    if (!P[0].ok && ranks) ranks[0] = -1;
    const int nb = (int)std::min<size_t>(kk, (size_t)std::max(ncol, 1024));
    CK(ensure(c->tb, sizeof(double) * nn * nb));
    CK(ensure(c->dout, sizeof(double) * nn * nb));
    double *Xt = (double *)c->tb.p, *Yt = (double *)c->dout.p;
    HK(hipEventRecord(c->ev0, s));
    for (int c0 = 0; c0 < k; c0 += nb) {
        const int w = std::min(nb, k - c0);
        if (left) {  // columns c0 .. c0 + w of b (n x w): b_blk := A b_blk
            double *Bb = dB + (size_t)c0 * nn;
            if (P[0].ok) {
                hipLaunchKernelGGL(k_transpose, dim3(cdiv(n, 64), cdiv(w, 64)), dim3(WGS), 0, s, n, w,
                                   (const double *)Bb, (long)n, Xt, (long)w);
                HssSpec sp;
                sp.K = n; sp.X = Xt; sp.ldx = w; sp.Y = Yt; sp.ldy = w; sp.nrow = w;
                std::vector<HssSpec> v1(1, sp);
                CK(hss_mult(c, v1, P, true));
                hipLaunchKernelGGL(k_transpose, dim3(cdiv(w, 64), cdiv(n, 64)), dim3(WGS), 0, s, w, n,
                                   (const double *)Yt, (long)w, Bb, (long)n);
            } else {
                hipLaunchKernelGGL(k_dgemm_nn, dim3(cdiv(n, 64), cdiv(w, 64)), dim3(WGS), 0, s, n, w, n,
                                   (const double *)dA, (long)n, (const double *)Bb, (long)n, Yt, (long)n);
                HK(hipMemcpyAsync(Bb, Yt, sizeof(double) * nn * w, hipMemcpyDeviceToDevice, s));
            }
        } else {  // rows c0 .. c0 + w of b (w x n, in b's k x n block): b_blk := b_blk A
            double *Bb = dB + c0;
            if (P[0].ok) {
                HssSpec sp;
                sp.K = n; sp.X = Bb; sp.ldx = k; sp.Y = Yt; sp.ldy = w; sp.nrow = w;
                std::vector<HssSpec> v1(1, sp);
                CK(hss_mult(c, v1, P, false));
            } else {
                hipLaunchKernelGGL(k_dgemm_nn, dim3(cdiv(w, 64), cdiv(n, 64)), dim3(WGS), 0, s, w, n, n,
                                   (const double *)Bb, (long)k, (const double *)dA, (long)n, Yt, (long)w);
            }
            HK(hipMemcpy2DAsync(Bb, sizeof(double) * kk, Yt, sizeof(double) * w, sizeof(double) * w, nn,
                                hipMemcpyDeviceToDevice, s));
        }
        HK(hipGetLastError());
    }
    HK(hipEventRecord(c->ev1, s));
    HK(hipMemcpy2DAsync(b, sizeof(double) * ldb, dB, sizeof(double) * br, sizeof(double) * br, bc,
                        hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    float ms_mm = 0.0f;
    (void)hipEventElapsedTime(&ms_mm, c->ev0, c->ev1);
    if (time) { time[0] = ms_build * 1e-3; time[1] = ms_mm * 1e-3; }
    return 0;
}

int hsdc_dhssmm(hsdc_ctx *c, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol, int prank,
                int ni, int ncol, double *time, int *ranks)
{
    (void)prank;
    const int rc = dhss_args(c, a, lda, m, n, b, ldb, k, tol, ni, ncol, true);
    if (rc) return rc;
    return dhss_run(c, a, lda, n, b, ldb, k, tol, ni, ncol, time, ranks, true);
}

int hsdc_dhssmml(hsdc_ctx *c, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol,
                 int prank, int ni, int ncol, double *time, int *ranks)
{
    (void)prank;
    const int rc = dhss_args(c, a, lda, m, n, b, ldb, k, tol, ni, ncol, false);
    if (rc) return rc;
    return dhss_run(c, a, lda, n, b, ldb, k, tol, ni, ncol, time, ranks, false);
}

int hsdc_dhssmm_rand(hsdc_ctx *c, const double *a, int lda, int m, int n, double *b, int ldb, int k, double tol,
                     int ni, int ncol, int side, unsigned long long seed, double *time, int *ranks)
{
    if (side != 0 && side != 1) return -12;
    const int rc = dhss_args(c, a, lda, m, n, b, ldb, k, tol, ni, ncol, side == 0);
    if (rc) return rc;
    return dhss_run(c, a, lda, n, b, ldb, k, tol, ni, ncol, time, ranks, side == 0, true, seed);
}

int hsdc_hss_dense(hsdc_ctx *c, double *a, int lda)
{
    if (!c) return HSDC_ERR_INTERNAL;
    int k = c->evc_k;
    if (k < 1) return -1;
    if (lda < k) return -3;
    std::vector<double> g = c->evc_gen;
    std::vector<double> eye((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) eye[(size_t)i * k + i] = 1.0;
    int rc = hss_one(c, c->evc_sq, k, c->evc_ni, c->evc_tol, g.data(), g.data() + k, g.data() + 2 * (size_t)k,
                     g.data() + 3 * (size_t)k, g.data() + 4 * (size_t)k, g.data() + 5 * (size_t)k, eye.data(), k,
                     k, nullptr);
    if (rc) return rc;
    for (int j = 0; j < k; j++) memcpy(a + (size_t)j * lda, eye.data() + (size_t)j * k, sizeof(double) * k);
    return 0;
}

// MDLAED0 (mdlaed0.f90:121-380) on host arrays: icompq 0 eigenvalues only, 1 Q(qsiz x n) :=
// Q * V, 2 Q := V.  LDQS and the WORK/IWORK arrays of the Fortran routine are not needed.
int hsdc_laed0(hsdc_ctx *c, int icompq, int qsiz, int n, double *d, double *e, double *q, int ldq)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (icompq < 0 || icompq > 2) return -1;
    if (icompq == 1 && qsiz < std::max(0, n)) return -2;
    if (n < 0) return -3;
    if (ldq < std::max(1, n)) return -7;
    if (icompq == 1 && ldq < qsiz) return -7;
    if (n == 0) return 0;
    if (n == 1) {
        if (icompq == 2) q[0] = 1.0;
        return 0;
    }
    if (!have_device()) return HSDC_ERR_NODEVICE;
    CK(ensure(c->din, sizeof(double) * 2 * n));
    double *dd = (double *)c->din.p, *de = dd + n;
    CK(ensure(c->zout, sizeof(double) * (size_t)n * n));
    double *dz = (double *)c->zout.p;
    HK(hipMemcpy(dd, d, sizeof(double) * n, hipMemcpyHostToDevice));
    HK(hipMemcpy(de, e, sizeof(double) * (n - 1), hipMemcpyHostToDevice));
    int info = hsdc_stedc_dev(c, n, dd, de, dz, n);
    if (info) return info;
    HK(hipMemcpy(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (icompq == 2) {
        HK(hipMemcpy2D(q, sizeof(double) * ldq, dz, sizeof(double) * n, sizeof(double) * n, n,
                       hipMemcpyDeviceToHost));
    } else if (icompq == 1) {
        CK(ensure(c->misc, sizeof(double) * 2 * (size_t)qsiz * n));
        double *qa = (double *)c->misc.p, *qc = qa + (size_t)qsiz * n;
        HK(hipMemcpy2D(qa, sizeof(double) * qsiz, q, sizeof(double) * ldq, sizeof(double) * qsiz, n,
                       hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_dgemm_nn, dim3(cdiv(qsiz, 64), cdiv(n, 64)), dim3(WGS), 0, c->stream, qsiz, n, n,
                           (const double *)qa, (long)qsiz, (const double *)dz, (long)n, qc, (long)qsiz);
        HK(hipGetLastError());
        HK(hipMemcpy2D(q, sizeof(double) * ldq, qc, sizeof(double) * qsiz, sizeof(double) * qsiz, n,
                       hipMemcpyDeviceToHost));
    }
    return 0;
}

int hsdc_stedc(hsdc_ctx *c, char compz, int n, double *d, double *e, double *z, int ldz, double *work,
               int lwork, int *iwork, int liwork)
{
    if (!c) return HSDC_ERR_INTERNAL;
    int icompz, info = 0, lwmin, liwmin, smlsiz = c->prm.smlsiz;
    int lquery = (lwork == -1 || liwork == -1);
    if (compz == 'N' || compz == 'n') icompz = 0;
    else if (compz == 'V' || compz == 'v') icompz = 1;
    else if (compz == 'I' || compz == 'i') icompz = 2;
    else icompz = -1;
    if (icompz < 0) info = -1;
    else if (n < 0) info = -2;
    else if (ldz < 1 || (icompz > 0 && ldz < (n > 1 ? n : 1))) info = -6;
    if (info == 0) {
        if (n <= 1 || icompz == 0) { liwmin = 1; lwmin = 1; }
        else if (n <= smlsiz) { liwmin = 1; lwmin = 2 * (n - 1); }
        else {
            int lgn = (int)(log((double)n) / log(2.0));
            if ((1 << lgn) < n) lgn++;
            if ((1 << lgn) < n) lgn++;
            if (icompz == 1) { lwmin = 1 + 3 * n + 2 * n * lgn + 4 * n * n; liwmin = 6 + 6 * n + 5 * n * lgn; }
            else { lwmin = 1 + 4 * n + n * n; liwmin = 3 + 5 * n; }
        }
        if (work) work[0] = lwmin;
        if (iwork) iwork[0] = liwmin;
        if (lwork < lwmin && !lquery) info = -8;
        else if (liwork < liwmin && !lquery) info = -10;
    }
    if (info != 0 || lquery) return info;
    if (n == 0) return 0;
    if (n == 1) { if (icompz != 0) z[0] = 1.0; return 0; }
    if (!have_device()) return HSDC_ERR_NODEVICE;
    // stage through device memory
    CK(ensure(c->din, sizeof(double) * 2 * n));
    double *dd = (double *)c->din.p, *de = dd + n;
    CK(ensure(c->zout, sizeof(double) * (size_t)n * n));
    double *dz = (double *)c->zout.p;
    HK(hipMemcpy(dd, d, sizeof(double) * n, hipMemcpyHostToDevice));
    HK(hipMemcpy(de, e, sizeof(double) * (n - 1), hipMemcpyHostToDevice));
    info = hsdc_stedc_dev(c, n, dd, de, dz, n);
    if (info) return info;
    HK(hipMemcpy(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (icompz == 2) {
        CK(copy_d2h_matrix(c, z, ldz, dz, n, n));
    } else if (icompz == 1) {
        // COMPZ = 'V': Z := Zin * Q (the back-transformation of DSTEDC) on the device
        CK(ensure(c->misc, sizeof(double) * 2 * (size_t)n * n));
        double *za = (double *)c->misc.p, *zc = za + (size_t)n * n;
        HK(hipMemcpy2D(za, sizeof(double) * n, z, sizeof(double) * ldz, sizeof(double) * n, n,
                       hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_dgemm_nn, dim3(cdiv(n, 64), cdiv(n, 64)), dim3(WGS), 0, c->stream, n, n, n,
                           (const double *)za, (long)n, (const double *)dz, (long)n, zc, (long)n);
        HK(hipGetLastError());
        CK(copy_d2h_matrix(c, z, ldz, zc, n, n));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Learned starting iterate of the secular solver
// ---------------------------------------------------------------------------
int hsdc_mlp_set_weights(hsdc_ctx *c, const float *w, int nparam)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (!w) { c->hw.clear(); return 0; }
    if (nparam != MLP_NPARAM) return -3;
    for (int i = 0; i < nparam; i++)
        if (!std::isfinite(w[i])) return -2;
    if (!have_device()) return HSDC_ERR_NODEVICE;
    CK(ensure(c->mlpw, sizeof(float) * MLP_NPARAM));
    HK(hipMemcpy(c->mlpw.p, w, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    c->hw.assign(w, w + MLP_NPARAM);
    return 0;
}

int hsdc_mlp_get_weights(hsdc_ctx *c, float *w, int nparam)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (nparam != MLP_NPARAM) return -3;
    if (c->hw.empty()) return 1;
    memcpy(w, c->hw.data(), sizeof(float) * MLP_NPARAM);
    return 0;
}

// FNV-1a over the weight bytes (the checksum of the weights file)
static unsigned int mlp_fnv(const float *w, int n)
{
    const unsigned char *b = (const unsigned char *)w;
    unsigned int h = 2166136261u;
    for (size_t i = 0; i < sizeof(float) * (size_t)n; i++) { h ^= b[i]; h *= 16777619u; }
    return h;
}

// Weights file: "HSDCMLP1", int32 inputs, int32 hidden, int32 nparam, uint32 FNV-1a of
// the weights, float32 weights; little-endian.
int hsdc_mlp_load(hsdc_ctx *c, const char *path)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (!path) return -2;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    char magic[8];
    int hdr[3];
    unsigned int sum = 0;
    std::vector<float> w(MLP_NPARAM);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "HSDCMLP1", 8) == 0 &&
              fread(hdr, sizeof(int), 3, f) == 3 && hdr[0] == MLP_IN && hdr[1] == MLP_H &&
              hdr[2] == MLP_NPARAM && fread(&sum, sizeof(sum), 1, f) == 1 &&
              fread(w.data(), sizeof(float), MLP_NPARAM, f) == (size_t)MLP_NPARAM;
    fclose(f);
    if (!ok || mlp_fnv(w.data(), MLP_NPARAM) != sum) return 1;
    return hsdc_mlp_set_weights(c, w.data(), MLP_NPARAM);
}

int hsdc_mlp_harvest(hsdc_ctx *c, long long cap)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (cap < 0) return -2;
    c->hv_cap = 0;
    if (cap == 0) return 0;
    if (!have_device()) return HSDC_ERR_NODEVICE;
    size_t bytes = 64 + al64(sizeof(float) * MLP_IN * (size_t)cap) + sizeof(float) * (size_t)cap;
    CK(ensure(c->hv, bytes));
    unsigned long long zero = 0;
    HK(hipMemcpy(c->hv.p, &zero, sizeof(zero), hipMemcpyHostToDevice));
    c->hv_cap = cap;
    return 0;
}

int hsdc_mlp_samples(hsdc_ctx *c, float *x, float *t, long long maxn, long long *n)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (!n) return -5;
    *n = 0;
    if (c->hv_cap == 0) return 1;
    unsigned long long cnt = 0;
    HK(hipMemcpy(&cnt, c->hv.p, sizeof(cnt), hipMemcpyDeviceToHost));
    long long m = std::min<long long>((long long)cnt, std::min(maxn, c->hv_cap));
    if (!x || !t) {
        *n = std::min<long long>((long long)cnt, c->hv_cap);
        return 0;
    }
    const char *hv = (const char *)c->hv.p;
    if (m > 0) {
        HK(hipMemcpy(x, hv + 64, sizeof(float) * MLP_IN * (size_t)m, hipMemcpyDeviceToHost));
        HK(hipMemcpy(t, hv + 64 + al64(sizeof(float) * MLP_IN * (size_t)c->hv_cap), sizeof(float) * (size_t)m,
                     hipMemcpyDeviceToHost));
    }
    *n = m;
    return 0;
}

// training workspace: X | T | idx | partials | block losses | grad | m | v | batch losses
struct MlpTrainBuf {
    float *X, *T, *part, *lossp, *g, *m, *v, *lb;
    int *idx;
    int nblk;
};

static int mlp_trainbuf(hsdc_ctx *c, long long ns, int nblk, int nbatch, MlpTrainBuf &b)
{
    size_t o1 = 0, o2 = al64(o1 + sizeof(float) * MLP_IN * (size_t)ns), o3 = al64(o2 + sizeof(float) * ns),
           o4 = al64(o3 + sizeof(int) * ns), o5 = al64(o4 + sizeof(float) * (size_t)nblk * MLP_NPARAM),
           o6 = al64(o5 + sizeof(float) * nblk), o7 = al64(o6 + sizeof(float) * MLP_NPARAM),
           o8 = al64(o7 + sizeof(float) * MLP_NPARAM), o9 = al64(o8 + sizeof(float) * MLP_NPARAM),
           o10 = al64(o9 + sizeof(float) * std::max(nbatch, 1));
    CK(ensure(c->mlptr, o10));
    char *p = (char *)c->mlptr.p;
    b.X = (float *)(p + o1); b.T = (float *)(p + o2); b.idx = (int *)(p + o3); b.part = (float *)(p + o4);
    b.lossp = (float *)(p + o5); b.g = (float *)(p + o6); b.m = (float *)(p + o7); b.v = (float *)(p + o8);
    b.lb = (float *)(p + o9);
    b.nblk = nblk;
    return 0;
}

static int mlp_grad_blocks(long long nb) { return (int)std::min<long long>(256, std::max<long long>(1, (nb + 31) / 32)); }

int hsdc_mlp_grad(hsdc_ctx *c, const float *x, const float *t, int ns, float *grad, float *loss)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (ns < 1) return -4;
    if (c->hw.empty()) return 1;
    if (!have_device()) return HSDC_ERR_NODEVICE;
    int nblk = mlp_grad_blocks(ns);
    MlpTrainBuf b;
    CK(mlp_trainbuf(c, ns, nblk, 1, b));
    hipStream_t s = c->stream;
    HK(hipMemcpyAsync(b.X, x, sizeof(float) * MLP_IN * (size_t)ns, hipMemcpyHostToDevice, s));
    HK(hipMemcpyAsync(b.T, t, sizeof(float) * ns, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, s, (const float *)c->mlpw.p, (const float *)b.X,
                       (const float *)b.T, (const int *)nullptr, ns, 1.0f / (float)ns, b.part, b.lossp);
    hipLaunchKernelGGL(k_mlp_sum, dim3(cdiv(MLP_NPARAM, 256)), dim3(256), 0, s, (const float *)b.part, nblk, b.g,
                       (const float *)b.lossp, b.lb);
    HK(hipGetLastError());
    HK(hipMemcpyAsync(grad, b.g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s));
    HK(hipMemcpyAsync(loss, b.lb, sizeof(float), hipMemcpyDeviceToHost, s));
    HK(hipStreamSynchronize(s));
    return 0;
}

static unsigned long long mlp_sm64(unsigned long long &st)
{
    unsigned long long z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Adam over shuffled mini-batches, learning rate lr * (1 + cos(pi e / epochs)) / 2 in
// epoch e.  Starts from the loaded weights, or from a seeded uniform initialisation
// (bound sqrt(6 / fan_in), last layer scaled by 0.1, zero biases) when none are loaded.
int hsdc_mlp_train(hsdc_ctx *c, const float *x, const float *t, long long ns, int epochs, int batch, float lr,
                   unsigned long long seed, float *loss_hist)
{
    if (!c) return HSDC_ERR_INTERNAL;
    if (ns < 1 || ns > 0x7fffffffLL) return -4;
    if (epochs < 1) return -5;
    if (batch < 1) return -6;
    if (!(lr > 0.0f)) return -7;
    if (!have_device()) return HSDC_ERR_NODEVICE;
    unsigned long long st = seed * 0x100000001ULL + 77;
    if (c->hw.empty()) {
        std::vector<float> w(MLP_NPARAM, 0.0f);
        auto uni = [&](float a) { return (float)(((mlp_sm64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); };
        for (int i = 0; i < MLP_H * MLP_IN; i++) w[MLP_W1 + i] = uni(sqrtf(6.0f / MLP_IN));
        for (int i = 0; i < MLP_H * MLP_H; i++) w[MLP_W2 + i] = uni(sqrtf(6.0f / MLP_H));
        for (int i = 0; i < MLP_H; i++) w[MLP_W3 + i] = uni(sqrtf(6.0f / MLP_H) * 0.1f);
        CK(hsdc_mlp_set_weights(c, w.data(), MLP_NPARAM));
    }
    const int bs = (int)std::min<long long>(batch, ns);
    const int nbatch = (int)((ns + bs - 1) / bs);
    const int nblk = mlp_grad_blocks(bs);
    MlpTrainBuf b;
    CK(mlp_trainbuf(c, ns, nblk, nbatch, b));
    hipStream_t s = c->stream;
    float *W = (float *)c->mlpw.p;
    HK(hipMemcpyAsync(b.X, x, sizeof(float) * MLP_IN * (size_t)ns, hipMemcpyHostToDevice, s));
    HK(hipMemcpyAsync(b.T, t, sizeof(float) * ns, hipMemcpyHostToDevice, s));
    std::vector<float> zeros(MLP_NPARAM, 0.0f);
    HK(hipMemcpyAsync(b.m, zeros.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s));
    HK(hipMemcpyAsync(b.v, zeros.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s));
    HK(hipStreamSynchronize(s));
    std::vector<int> perm(ns);
    std::vector<float> lb(nbatch);
    for (long long i = 0; i < ns; i++) perm[i] = (int)i;
    long long step = 0;
    for (int e = 0; e < epochs; e++) {
        for (long long i = ns - 1; i > 0; i--) std::swap(perm[i], perm[mlp_sm64(st) % (unsigned long long)(i + 1)]);
        HK(hipMemcpyAsync(b.idx, perm.data(), sizeof(int) * ns, hipMemcpyHostToDevice, s));
        const float lre = lr * 0.5f * (1.0f + cosf(3.14159265f * (float)e / (float)epochs));
        for (int bi = 0; bi < nbatch; bi++) {
            const int b0 = bi * bs, nb = (int)std::min<long long>(bs, ns - b0);
            step++;
            const float b1c = 1.0f - powf(0.9f, (float)step), b2c = 1.0f - powf(0.999f, (float)step);
            hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, s, (const float *)W, (const float *)b.X,
                               (const float *)b.T, (const int *)(b.idx + b0), nb, 1.0f / (float)nb, b.part, b.lossp);
            hipLaunchKernelGGL(k_mlp_sum, dim3(cdiv(MLP_NPARAM, 256)), dim3(256), 0, s, (const float *)b.part, nblk,
                               b.g, (const float *)b.lossp, b.lb + bi);
            hipLaunchKernelGGL(k_mlp_adam, dim3(cdiv(MLP_NPARAM, 256)), dim3(256), 0, s, W, (const float *)b.g, b.m,
                               b.v, lre, b1c, b2c);
        }
        HK(hipGetLastError());
        HK(hipMemcpyAsync(lb.data(), b.lb, sizeof(float) * nbatch, hipMemcpyDeviceToHost, s));
        HK(hipStreamSynchronize(s));
        if (loss_hist) {
            double a = 0.0;
            for (int bi = 0; bi < nbatch; bi++) a += lb[bi];
            loss_hist[e] = (float)(a / nbatch);
        }
    }
    HK(hipMemcpy(c->hw.data(), W, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    for (float v : c->hw)
        if (!std::isfinite(v)) return 2;
    return 0;
}

}  // extern "C"
