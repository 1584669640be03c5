// hsdc_loewner_noalloc: the secular / Loewner / alpha half of MDLAED3
// (mdlaed3.f90:201-249) on caller device arrays, with every piece of scratch in
// one file-scope device array, so that the call allocates nothing.  Its kernels
// are the ones hsdc_loewner runs (k_secular, k_loewner, k_alpha); only where their
// scratch lives differs.
#include <mutex>
#include <stddef.h>
#include <string.h>

#include "hsdc_internal.h"
#include "../../include/hsdc.h"

__global__ void k_secular(const Merge *, const int *, int, const MergeOut *, Work, int, int *, const SecCtl *,
                          unsigned int *);
__global__ void k_loewner(const Merge *, const int *, int, const MergeOut *, Work);
__global__ void k_alpha(const Merge *, const int *, int, const MergeOut *, Work);

namespace {

// Layout of the scratch array (byte offsets, every region 256-byte aligned):
//   control block (NA_HDR bytes): Merge | tables | MergeOut | gate | SecCtl | counters
//   tau (NA_KMAX doubles) | org (NA_KMAX ints) | iterations (NA_KMAX ints)
constexpr size_t NA_HDR = 16384;
constexpr size_t O_MERGE = 0, O_TAB = 256, O_MO = 512, O_GATE = 768, O_SEC = 1024, O_CNT = 2048;
constexpr size_t N_CNT = 4 * SEC_SLOTS;  // iters, used, wave_iters, evals (8 bytes each)
constexpr size_t O_LVL = O_CNT + 8 * N_CNT;  // largest iteration count per slot (4 bytes each)
static_assert(O_LVL + 4 * SEC_SLOTS <= NA_HDR, "control block");
constexpr size_t O_TAU = NA_HDR;
constexpr size_t O_ORG = O_TAU + 8 * (size_t)HSDC_NOALLOC_KMAX;
constexpr size_t O_NIT = O_ORG + 4 * (size_t)HSDC_NOALLOC_KMAX;
constexpr size_t NA_BYTES = O_NIT + 4 * (size_t)HSDC_NOALLOC_KMAX;

}  // namespace

// One copy per GPU, held until the process ends.  Host code never passes this name
// to a kernel or a copy: it uses the device address hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

constexpr int NA_MAXDEV = 64;
std::mutex na_mu;
void *na_ptr[NA_MAXDEV];  // device address of `scratch` on each device (null: not looked up yet)

// The scratch address on the current device, looked up once per device.
int na_scratch(unsigned char **p)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { (void)hipGetLastError(); return HSDC_ERR_NODEVICE; }
    if (dev < 0 || dev >= NA_MAXDEV) return HSDC_ERR_INTERNAL;
    std::lock_guard<std::mutex> g(na_mu);
    if (!na_ptr[dev]) {
        void *q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) {
            (void)hipGetLastError();
            return HSDC_ERR_HIP;
        }
        na_ptr[dev] = q;
    }
    *p = (unsigned char *)na_ptr[dev];
    return 0;
}

// Control block of one call, written by the device (no host copy, so the stream may
// be under capture): the merge of order k at index 0, its launch tables, its
// MergeOut (HSS branch: MDLAED3's formulas), the closed gate cleared, a SecCtl with
// the classical start and zeroed counters, INFO cleared.
__global__ void k_na_setup(unsigned char *base, int k, double rho, int *info)
{
    const int t = threadIdx.x;
    unsigned long long *cnt = (unsigned long long *)(base + O_CNT);
    for (int i = t; i < (int)N_CNT; i += blockDim.x) cnt[i] = 0;
    unsigned int *lvl = (unsigned int *)(base + O_LVL);
    for (int i = t; i < SEC_SLOTS; i += blockDim.x) lvl[i] = 0;
    if (t == 0) {
        Merge *M = (Merge *)(base + O_MERGE);
        M->s = 0; M->n = k; M->n1 = k / 2; M->prob = 0; M->lvl = 0; M->rho = rho;
        int *tab = (int *)(base + O_TAB);
        tab[0] = 0; tab[1] = (k + SEC_WG / sec_team(k) - 1) / (SEC_WG / sec_team(k));
        tab[2] = 0; tab[3] = (k + 255) / 256;
        MergeOut *o = (MergeOut *)(base + O_MO);
        o->k = k; o->ctot[0] = o->ctot[1] = o->ctot[2] = o->ctot[3] = 0;
        o->hss = 1; o->nrot = 0; o->flags = 0; o->rho = rho; o->tol = 0.0;
        *(int *)(base + O_GATE) = 0;
        SecCtl *C = (SecCtl *)(base + O_SEC);
        C->w = nullptr;
        C->iters = cnt; C->used = cnt + SEC_SLOTS; C->wave_iters = cnt + 2 * SEC_SLOTS;
        C->evals = cnt + 3 * SEC_SLOTS;
        C->hf = nullptr; C->ht = nullptr; C->feat = nullptr; C->count = nullptr; C->cap = 0;
        C->kmin = 1 << 30; C->gate = 0.0f;
        *info = 0;
    }
}

}  // namespace

extern "C" int hsdc_loewner_noalloc(int k, double rho, const double *d_dlamda, const double *d_z, double *d_d,
                                    double *d_znew, double *d_difl, double *d_difr, double *d_alpha, int *d_info,
                                    void *stream)
{
    if (k < 3) return -1;
    if (k > HSDC_NOALLOC_KMAX) return -1;
    if (!d_dlamda) return -3;
    if (!d_z) return -4;
    if (!d_d) return -5;
    if (!d_znew) return -6;
    if (!d_difl) return -7;
    if (!d_difr) return -8;
    if (!d_alpha) return -9;
    if (!d_info) return -10;
    unsigned char *b = nullptr;
    int rc = na_scratch(&b);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    Work W;
    memset(&W, 0, sizeof(W));
    W.dlamda = const_cast<double *>(d_dlamda);
    W.w = const_cast<double *>(d_z);
    W.lam = d_d;
    W.wnew = d_znew;
    W.difl = d_difl;
    W.difr = d_difr;
    W.alpha = d_alpha;
    W.tau = (double *)(b + O_TAU);
    W.org = (int *)(b + O_ORG);
    W.nit = (int *)(b + O_NIT);
    W.gate = (int *)(b + O_GATE);
    const Merge *dm = (const Merge *)(b + O_MERGE);
    const int *tab = (const int *)(b + O_TAB);
    const MergeOut *dmo = (const MergeOut *)(b + O_MO);
    const int team = sec_team(k);
    const int nsec = (k + SEC_WG / team - 1) / (SEC_WG / team), nch = (k + 255) / 256;
    hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(256), 0, s, b, k, rho, d_info);
    hipLaunchKernelGGL(k_secular, dim3(nsec), dim3(SEC_WG), 0, s, dm, tab, 1, dmo, W, 100, d_info,
                       (const SecCtl *)(b + O_SEC), (unsigned int *)(b + O_LVL));
    hipLaunchKernelGGL(k_loewner, dim3(nch), dim3(256), 0, s, dm, tab + 2, 1, dmo, W);
    hipLaunchKernelGGL(k_alpha, dim3(nch), dim3(256), 0, s, dm, tab + 2, 1, dmo, W);
    if (hipGetLastError() != hipSuccess) return HSDC_ERR_HIP;
    return 0;
}
