/*
 * CPU ORACLE -- test infrastructure only (see orc.h).
 *
 * Restatement of HSSPACK's bidiagonal divide-and-conquer SVD (SRC/mdbdsdc.f90,
 * mdlasd0.f90, mdlasd1.f90, mdlasd2.f90) and of the LAPACK routines that path
 * calls (DLASDT, DLASDQ, DLASD3, DLASD4, DLAMRG, DLASCL, DLARTG, DLASR).  For a
 * merge with K below the reference's HSS threshold (BSMLZ, mdlasd2.f90:176) MDLASD2
 * is LAPACK DLASD2 and MDLASD1 calls LAPACK DLASD3 (mdlasd1.f90:239-244), so the
 * whole path is LAPACK DBDSDC there; the checks against LAPACK pin it.  The HSS
 * branch of MDLASD1 (MDLASD8/DHSSVCS) is not restated: every merge takes the dense
 * branch.
 *
 * Own choices (the device follows them bit for bit):
 *   leaves  -- one-sided (Hestenes) Jacobi on the m x (m+sqre) bidiagonal instead of
 *              DLASDQ's implicit QR; dot products summed in the 64-lane butterfly
 *              order of one wavefront.
 *   roots   -- the rational middle-way iteration of orc_laed4 carried to the SVD
 *              secular function 1/rho + sum z_i^2/((d_i-s)(d_i+s)) (in the variable
 *              s^2 - d_org^2, steps mapped back to s as DLASD4 does); T-lane
 *              interleaved sums.
 *   vectors -- DLASD3's formulas; norms as T-lane interleaved sums; one dense
 *              product per singular-vector block (the structured zero blocks of
 *              DLASD3's products add exact zeros).
 *   final   -- MDBDSDC's selection sort (mdbdsdc.f90:383-404) literally.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define EPS 0x1p-53
#define SAFMIN 0x1p-1022

static double lapy2(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    double r = z / w;
    return w * sqrt(1.0 + r * r);
}

/* LAPACK DLASCL, TYPE 'G' on a vector: x := x * (cto / cfrom) in steps that neither
 * overflow nor underflow (the multiplier is 1/cfrom * cto for ordinary values). */
void orc_lascl(double cfrom, double cto, int n, double *x)
{
    const double smlnum = SAFMIN, bignum = 1.0 / smlnum;
    double cfromc = cfrom, ctoc = cto, mul;
    int done = 0;
    while (!done) {
        double cfrom1 = cfromc * smlnum;
        if (cfrom1 == cfromc) {
            mul = ctoc / cfromc;
            done = 1;
        } else {
            double cto1 = ctoc / bignum;
            if (cto1 == ctoc) {
                mul = ctoc;
                done = 1;
                cfromc = 1.0;
            } else if (fabs(cfrom1) > fabs(ctoc) && ctoc != 0.0) {
                mul = smlnum;
                cfromc = cfrom1;
            } else if (fabs(cto1) > fabs(cfromc)) {
                mul = bignum;
                ctoc = cto1;
            } else {
                mul = ctoc / cfromc;
                done = 1;
                if (mul == 1.0) return;
            }
        }
        for (int i = 0; i < n; i++) x[i] *= mul;
    }
}

/* LAPACK DLARTG (the 3.10 form): r = sign(f) sqrt(f^2+g^2), c = |f|/|r|, s = g/r. */
void orc_lartg(double f, double g, double *c, double *s, double *r)
{
    const double rtmin = sqrt(SAFMIN), rtmax = sqrt(1.0 / SAFMIN / 2.0);
    double f1 = fabs(f), g1 = fabs(g);
    if (g == 0.0) {
        *c = 1.0; *s = 0.0; *r = f;
    } else if (f == 0.0) {
        *c = 0.0; *s = g > 0.0 ? 1.0 : -1.0; *r = g1;
    } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
        double d = sqrt(f * f + g * g);
        *c = f1 / d;
        *r = f > 0.0 ? d : -d;
        *s = g / *r;
    } else {
        double u = fmin(1.0 / SAFMIN, fmax(SAFMIN, fmax(f1, g1)));
        double fs = f / u, gs = g / u;
        double d = sqrt(fs * fs + gs * gs);
        *c = fabs(fs) / d;
        *r = fs > 0.0 ? d : -d;
        *s = gs / *r;
        *r *= u;
    }
}

/* LAPACK DLASDT (called at mdlasd0.f90:166): the tree of subproblems.  Node arrays
 * are 0-based here, their values (centres, sizes) 1-based as in Fortran. */
void orc_lasdt(int n, int *lvl, int *nd, int *inode, int *ndiml, int *ndimr, int msub)
{
    int maxn = n > 1 ? n : 1;
    double temp = log((double)maxn / (double)(msub + 1)) / log(2.0);
    *lvl = (int)temp + 1;
    int i = n / 2;
    inode[0] = i + 1;
    ndiml[0] = i;
    ndimr[0] = n - i - 1;
    int il = -1, ir = 0, llst = 1;
    for (int l = 1; l <= *lvl - 1; l++) {
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
    *nd = llst * 2 - 1;
}

/* ------------------------------------------------------------------ */
/* Leaf SVD (replaces DLASDQ at mdlasd0.f90:152-183). */

/* butterfly over 64 lanes, lane l ^ off added to lane l for off = 32 .. 1 */
static double bfly64(double *v)
{
    double nv[64];
    for (int off = 32; off > 0; off >>= 1) {
        for (int l = 0; l < 64; l++) nv[l] = v[l] + v[l ^ off];
        for (int l = 0; l < 64; l++) v[l] = nv[l];
    }
    return v[0];
}

#define JAC_TOL 0x1p-52
#define JAC_SWEEPS 60

/* SVD of the m x (m+sqre) upper bidiagonal B (diagonal d, superdiagonal e, e has
 * m-1+sqre entries): B = U [diag(s) 0] VT, s ascending in d, U m x m (ldu), VT
 * (m+sqre) x (m+sqre) (ldvt) whose last row is the null vector when sqre = 1.
 * One-sided Jacobi on the columns of B in round-robin order (a sweep = P-1 rounds of
 * P/2 disjoint pairs, P = m+sqre rounded up to even, bd_rr_pair); a pair is rotated
 * while |w_p.w_q| > (m+sqre) 2^-52 |w_p||w_q| (the rounding level of the sums), until
 * a sweep rotates nothing; returns the number of sweeps, or -1 if JAC_SWEEPS did not
 * suffice. */
/* pair i of round t of the circle method over P (even) columns: column 0 fixed, the
 * others moved one place per round; p < q */
static void bd_rr_pair(int t, int i, int P, int *p, int *q)
{
    int a, b;
    if (i == 0) {
        a = 0;
        b = (P - 2 + t) % (P - 1) + 1;
    } else {
        a = (i - 1 + t) % (P - 1) + 1;
        b = (P - 2 - i + t) % (P - 1) + 1;
    }
    *p = a < b ? a : b;
    *q = a < b ? b : a;
}

int orc_bd_leaf(int m, int sqre, double *d, const double *e, double *u, int ldu, double *vt,
                int ldvt)
{
    const int pc = m + sqre;
    if (m < 1 || m > 64 || pc > 64) return -2;
    double W[64][64], V[64][64], a[64], sg[64];
    memset(W, 0, sizeof W);
    memset(V, 0, sizeof V);
    for (int j = 0; j < pc; j++) {
        V[j][j] = 1.0;
        if (j < m) W[j][j] = d[j];
        if (j >= 1) W[j][j - 1] = e[j - 1];
    }
    /* columns whose squared norm is below thr count as zero: rounding keeps such a
     * column (the null direction when sqre = 1) from ever looking orthogonal */
    double fro = 0.0;
    for (int j = 0; j < m; j++) fro = fro + d[j] * d[j];
    for (int j = 0; j < pc - 1; j++) fro = fro + e[j] * e[j];
    const double thr = (JAC_TOL * pc) * (JAC_TOL * pc) * fro;
    int sweep, rot = 1;
    const int P = pc + (pc & 1);
    for (sweep = 0; sweep < JAC_SWEEPS && rot; sweep++) {
        rot = 0;
        for (int rd = 0; rd < P - 1; rd++) {
            for (int pi = 0; pi < P / 2; pi++) {
                int p, q;
                bd_rr_pair(rd, pi, P, &p, &q);
                if (q >= pc) continue;
                for (int r = 0; r < 64; r++) a[r] = W[p][r] * W[p][r];
                double al = bfly64(a);
                for (int r = 0; r < 64; r++) a[r] = W[q][r] * W[q][r];
                double be = bfly64(a);
                for (int r = 0; r < 64; r++) a[r] = W[p][r] * W[q][r];
                double ga = bfly64(a);
                if (al <= thr || be <= thr || !(fabs(ga) > (JAC_TOL * pc) * (sqrt(al) * sqrt(be)))) continue;
                double zeta = (be - al) / (2.0 * ga), t;
                if (fabs(zeta) > 1e150) t = 0.5 / zeta;
                else t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < 64; r++) {
                    double wp = W[p][r], wq = W[q][r];
                    W[p][r] = c * wp - s * wq;
                    W[q][r] = s * wp + c * wq;
                    double vp = V[p][r], vq = V[q][r];
                    V[p][r] = c * vp - s * vq;
                    V[q][r] = s * vp + c * vq;
                }
                rot = 1;
            }
        }
    }
    double nn[64];
    for (int j = 0; j < pc; j++) {
        for (int r = 0; r < 64; r++) a[r] = W[j][r] * W[j][r];
        nn[j] = bfly64(a);
        sg[j] = sqrt(nn[j]);
    }
    /* the null column (sqre = 1): smallest norm, first of equals; the rest ascending */
    int used[64] = {0}, order[64], jn = -1;
    if (sqre) {
        jn = 0;
        for (int j = 1; j < pc; j++)
            if (sg[j] < sg[jn]) jn = j;
        used[jn] = 1;
    }
    for (int k = 0; k < m; k++) {
        int b = -1;
        for (int j = 0; j < pc; j++)
            if (!used[j] && (b < 0 || sg[j] < sg[b])) b = j;
        used[b] = 1;
        order[k] = b;
    }
    for (int k = 0; k < m; k++) {
        int j = order[k];
        d[k] = sg[j];
        for (int r = 0; r < m; r++) u[r + (long)k * ldu] = nn[j] > thr ? W[j][r] / sg[j] : 0.0;
        for (int c = 0; c < pc; c++) vt[k + (long)c * ldvt] = V[j][c];
    }
    /* a left vector of a negligible column: the first unit vector e_r that keeps norm
     * > 1/2 after two projections off the other columns */
    for (int k = 0; k < m; k++) {
        if (nn[order[k]] > thr) continue;
        for (int r0 = 0; r0 < m; r0++) {
            double x[64];
            for (int r = 0; r < 64; r++) x[r] = (r == r0) ? 1.0 : 0.0;
            for (int pass = 0; pass < 2; pass++)
                for (int k2 = 0; k2 < m; k2++) {
                    if (k2 == k) continue;
                    for (int r = 0; r < 64; r++) a[r] = r < m ? u[r + (long)k2 * ldu] * x[r] : 0.0;
                    double g = bfly64(a);
                    for (int r = 0; r < m; r++) x[r] = x[r] - g * u[r + (long)k2 * ldu];
                }
            for (int r = 0; r < 64; r++) a[r] = x[r] * x[r];
            double nx = sqrt(bfly64(a));
            if (nx > 0.5) {
                for (int r = 0; r < m; r++) u[r + (long)k * ldu] = x[r] / nx;
                break;
            }
        }
    }
    if (sqre)
        for (int c = 0; c < pc; c++) vt[m + (long)c * ldvt] = V[jn][c];
    return rot ? -1 : sweep;
}

/* ------------------------------------------------------------------ */
/* Secular equation of the SVD merge (replaces LAPACK DLASD4, called in DLASD3 at
 * mdlasd1.f90:241): root j (1-based) of f(s) = 1/rho + sum z_i^2/((d_i-s)(d_i+s)),
 * 0 = d_1 < d_2 < ... < d_k, |z| = 1.  The root is s = d_org + tau; the iteration
 * runs in mu = s^2 - d_org^2 whose poles sit at (d_i-d_org)(d_i+d_org), each
 * difference (d_i-s)(d_i+s) formed as ((d_i-d_org)-tau)*((d_i+d_org)+tau). */

static double tsumT(double *v, int T)
{
    double nv[64];
    for (int off = T >> 1; off > 0; off >>= 1) {
        for (int l = 0; l < T; l++) nv[l] = v[l] + v[l ^ off];
        for (int l = 0; l < T; l++) v[l] = nv[l];
    }
    return v[0];
}

/* starting iterate: two-pole model with the other poles frozen at the midpoint (in
 * s^2) of the root's interval -> origin (1-based), tau and its bracket */
static void sv_init(int k, int j, const double *d, const double *z, double rho, int T, int *orgp,
                    double *taup, double *lop, double *hip)
{
    double rhoinv = 1.0 / rho, ps[64];
    if (j == k) {
        /* s_k^2 in (d_k^2, d_k^2 + rho] */
        double dk = d[k - 1];
        double hi = rho / (dk + sqrt(dk * dk + rho));
        double mid = (0.5 * rho) / (dk + sqrt(dk * dk + 0.5 * rho));
        for (int l = 0; l < T; l++) {
            ps[l] = 0.0;
            for (int i = l; i < k - 2; i += T)
                ps[l] += z[i] * z[i] / (((d[i] - dk) - mid) * ((d[i] + dk) + mid));
        }
        double c = rhoinv + tsumT(ps, T);
        double dl = (dk - d[k - 2]) * (dk + d[k - 2]);
        double z1 = z[k - 2] * z[k - 2], z2 = z[k - 1] * z[k - 1];
        double t2;
        if (c <= 0.0) {
            t2 = rho;
        } else {
            /* c mu^2 + (c dl - z1 - z2) mu - z2 dl = 0, positive root */
            double b = c * dl - z1 - z2;
            double s = sqrt(fabs(b * b + 4.0 * c * z2 * dl));
            if (b <= 0.0) t2 = (s - b) / (2.0 * c);
            else t2 = 2.0 * z2 * dl / (b + s);
        }
        double t = t2 / (dk + sqrt(dk * dk + t2));
        if (!(t > 0.0 && t <= hi)) t = mid;
        *orgp = k; *taup = t; *lop = 0.0; *hip = hi;
        return;
    }
    double dj = d[j - 1], dj1 = d[j];
    double delsq = (dj1 - dj) * (dj1 + dj);
    double sq2 = sqrt((dj * dj + dj1 * dj1) / 2.0);
    double tm = (delsq / 2.0) / (dj + sq2);  /* s_mid - d_j */
    for (int l = 0; l < T; l++) {
        ps[l] = 0.0;
        for (int i = l; i < k; i += T) {
            if (i == j - 1 || i == j) continue;
            ps[l] += z[i] * z[i] / (((d[i] - dj) - tm) * ((d[i] + dj) + tm));
        }
    }
    double c = rhoinv + tsumT(ps, T);
    double zl = z[j - 1] * z[j - 1], zr = z[j] * z[j];
    double fmid = c + (zr - zl) / (delsq / 2.0);
    if (fmid >= 0.0) {
        /* origin d_j: c mu^2 - (c delsq + zl + zr) mu + zl delsq = 0, small root */
        double b = c * delsq + zl + zr, cc = zl * delsq;
        double s = sqrt(fabs(b * b - 4.0 * c * cc));
        double t2 = (b > 0.0) ? 2.0 * cc / (b + s) : (b - s) / (2.0 * c);
        double t = t2 / (dj + sqrt(dj * dj + t2));
        if (!(t > 0.0 && t <= tm)) t = 0.5 * tm;
        *orgp = j; *taup = t; *lop = 0.0; *hip = tm;
    } else {
        /* origin d_{j+1}: c mu^2 + (c delsq - zl - zr) mu - zr delsq = 0, negative root */
        double b = c * delsq - zl - zr, cc = zr * delsq;
        double s = sqrt(fabs(b * b + 4.0 * c * cc));
        double t2 = (b <= 0.0) ? -2.0 * cc / (s - b) : -(b + s) / (2.0 * c);
        double lo = -(delsq / 2.0) / (dj1 + sq2);
        double t = t2 / (dj1 + sqrt(fabs(dj1 * dj1 + t2)));
        if (!(t >= lo && t < 0.0)) t = 0.5 * lo;
        *orgp = j + 1; *taup = t; *lop = lo; *hip = 0.0;
    }
}

int orc_lasd4(int k, int j, const double *d, const double *z, double rho, double *sigma, int *org,
              double *tau, int *niter, int maxit)
{
    if (niter) *niter = 0;
    if (k == 1) {
        *sigma = sqrt(d[0] * d[0] + rho * z[0] * z[0]);
        *org = 1;
        *tau = *sigma - d[0];
        return 0;
    }
    const int T = orc_sec_team(k);
    int ii;
    double t, lo, hi;
    sv_init(k, j, d, z, rho, T, &ii, &t, &lo, &hi);
    const double rhoinv = 1.0 / rho, dorg = d[ii - 1], dorg2 = dorg + dorg;
    const int split = (j == k) ? k - 1 : j;
    const int na = split - ((ii - 1 < split) ? 1 : 0), nb = (k - split) - ((ii - 1 >= split) ? 1 : 0);
    int lgT = 0;
    while ((1 << lgT) < T) lgT++;
    int it, info = 1;
    for (it = 0; it <= maxit; it++) {
        double P[6][64];
        for (int l = 0; l < T; l++) {
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
            P[0][l] = psi; P[1][l] = dpsi; P[2][l] = epsi; P[3][l] = phi; P[4][l] = dphi; P[5][l] = ephi;
        }
        double psi = tsumT(P[0], T), dpsi = tsumT(P[1], T), epsi = tsumT(P[2], T);
        double phi = tsumT(P[3], T), dphi = tsumT(P[4], T), ephi = tsumT(P[5], T);
        double mu = t * (dorg2 + t);
        double qo = z[ii - 1] / (-mu);
        double torg = z[ii - 1] * qo;
        double w = ((rhoinv + phi) + psi) + torg;
        double erretm = (8.0 + lgT) * (phi - psi) + (ephi - epsi) + 2.0 * rhoinv + 3.0 * fabs(torg) +
                        fabs(mu) * ((dpsi + dphi) + qo * qo);
        if (ii - 1 < split) dpsi += qo * qo; else dphi += qo * qo;
        if (fabs(w) <= EPS * erretm || it == maxit) {
            if (fabs(w) <= EPS * erretm) info = 0;
            break;
        }
        if (w > 0.0) hi = t; else lo = t;
        /* poles either side of the root, as distances in mu from the iterate */
        double dl = ((d[split - 1] - dorg) - t) * ((d[split - 1] + dorg) + t);
        double dr = ((d[split] - dorg) - t) * ((d[split] + dorg) + t);
        double a = (dl + dr) * w - dl * dr * (dpsi + dphi);
        double b = dl * dr * w;
        double c = w - dl * dpsi - dr * dphi;
        double eta;
        if (c == 0.0) {
            eta = (a != 0.0) ? b / a : -w / (dpsi + dphi);
        } else {
            double s = sqrt(fabs(a * a - 4.0 * b * c));
            if (a <= 0.0) eta = (a - s) / (2.0 * c);
            else eta = 2.0 * b / (a + s);
        }
        if (w * eta >= 0.0) eta = -w / (dpsi + dphi);
        /* step in mu -> step in s: s' = sqrt(s^2 + eta) */
        double sg = dorg + t;
        double tn = t + eta / (sg + sqrt(fabs(sg * sg + eta)));
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t) { info = 0; break; }
        t = tn;
    }
    if (niter) *niter = it + 1;
    *sigma = dorg + t;
    *org = ii;
    *tau = t;
    return info;
}

/* (d_i - s)(d_i + s) for the root s = d_org + tau */
static inline double sv_den(double di, double dorg, double tau)
{
    return ((di - dorg) - tau) * ((di + dorg) + tau);
}

/* ------------------------------------------------------------------ */
/* LAPACK DLASD2 == MDLASD2 with HFLAG (mdlasd2.f90:1-548, K < BSMLZ branch): z, the
 * sort, deflation, the column grouping.  U (ldu) is the merge's n x n block, VT
 * (ldvt) its m x m block.  Index arrays 1-based.  nrot/rot_* (may be NULL) log the
 * Givens rotations of the deflation in order. */
int orc_lasd2(int nl, int nr, int sqre, int *kout, double *d, double *z, double alpha, double beta,
              double *u, int ldu, double *vt, int ldvt, double *dsigma, double *u2, int ldu2,
              double *vt2, int ldvt2, int *idxp, int *idx, int *idxc, int *idxq, int *coltyp,
              int *nrot, int *rot_p, int *rot_q, double *rot_c, double *rot_s)
{
#define D_(i) d[(i) - 1]
#define Z_(i) z[(i) - 1]
#define U_(i, j) u[((i) - 1) + (long)((j) - 1) * ldu]
#define VT_(i, j) vt[((i) - 1) + (long)((j) - 1) * ldvt]
#define U2_(i, j) u2[((i) - 1) + (long)((j) - 1) * ldu2]
#define VT2_(i, j) vt2[((i) - 1) + (long)((j) - 1) * ldvt2]
#define IX(a, i) a[(i) - 1]
    const int n = nl + nr + 1, m = n + sqre, nlp1 = nl + 1, nlp2 = nl + 2;
    int ctot[4], psm[4], i, j, k, k2, jprev = 0, nr_ = 0;
    double c, s, z1, tol, hlftol;
    z1 = alpha * VT_(nlp1, nlp1);
    Z_(1) = z1;
    for (i = nl; i >= 1; i--) {
        Z_(i + 1) = alpha * VT_(i, nlp1);
        D_(i + 1) = D_(i);
        IX(idxq, i + 1) = IX(idxq, i) + 1;
    }
    for (i = nlp2; i <= m; i++) Z_(i) = beta * VT_(i, nlp2);
    for (i = 2; i <= nlp1; i++) IX(coltyp, i) = 1;
    for (i = nlp2; i <= n; i++) IX(coltyp, i) = 2;
    for (i = nlp2; i <= n; i++) IX(idxq, i) += nlp1;
    for (i = 2; i <= n; i++) {
        dsigma[i - 1] = D_(IX(idxq, i));
        U2_(i, 1) = Z_(IX(idxq, i));
        IX(idxc, i) = IX(coltyp, IX(idxq, i));
    }
    orc_dlamrg(nl, nr, dsigma + 1, 1, 1, idx + 1);
    for (i = 2; i <= n; i++) {
        int idxi = 1 + IX(idx, i);
        D_(i) = dsigma[idxi - 1];
        Z_(i) = U2_(idxi, 1);
        IX(coltyp, i) = IX(idxc, idxi);
    }
    tol = fmax(fabs(alpha), fabs(beta));
    tol = 8.0 * EPS * fmax(fabs(D_(n)), tol);
    k = 1;
    k2 = n + 1;
    for (j = 2; j <= n; j++) {
        if (fabs(Z_(j)) <= tol) {
            k2--;
            IX(idxp, k2) = j;
            IX(coltyp, j) = 4;
            if (j == n) goto done;
        } else {
            jprev = j;
            break;
        }
    }
    j = jprev;
    for (;;) {
        j++;
        if (j > n) break;
        if (fabs(Z_(j)) <= tol) {
            k2--;
            IX(idxp, k2) = j;
            IX(coltyp, j) = 4;
        } else if (fabs(D_(j) - D_(jprev)) <= tol) {
            s = Z_(jprev);
            c = Z_(j);
            double tau = lapy2(c, s);
            c = c / tau;
            s = -s / tau;
            Z_(j) = tau;
            Z_(jprev) = 0.0;
            int idxjp = IX(idxq, IX(idx, jprev) + 1);
            int idxj = IX(idxq, IX(idx, j) + 1);
            if (idxjp <= nlp1) idxjp--;
            if (idxj <= nlp1) idxj--;
            for (i = 1; i <= n; i++) {
                double x = U_(i, idxjp), y = U_(i, idxj);
                U_(i, idxjp) = c * x + s * y;
                U_(i, idxj) = c * y - s * x;
            }
            for (i = 1; i <= m; i++) {
                double x = VT_(idxjp, i), y = VT_(idxj, i);
                VT_(idxjp, i) = c * x + s * y;
                VT_(idxj, i) = c * y - s * x;
            }
            if (rot_p) {
                rot_p[nr_] = idxjp; rot_q[nr_] = idxj; rot_c[nr_] = c; rot_s[nr_] = s;
            }
            nr_++;
            if (IX(coltyp, j) != IX(coltyp, jprev)) IX(coltyp, j) = 3;
            IX(coltyp, jprev) = 4;
            k2--;
            IX(idxp, k2) = jprev;
            jprev = j;
        } else {
            k++;
            U2_(k, 1) = Z_(jprev);
            dsigma[k - 1] = D_(jprev);
            IX(idxp, k) = jprev;
            jprev = j;
        }
    }
    k++;
    U2_(k, 1) = Z_(jprev);
    dsigma[k - 1] = D_(jprev);
    IX(idxp, k) = jprev;
done:
    for (j = 0; j < 4; j++) ctot[j] = 0;
    for (j = 2; j <= n; j++) ctot[IX(coltyp, j) - 1]++;
    psm[0] = 2;
    psm[1] = 2 + ctot[0];
    psm[2] = psm[1] + ctot[1];
    psm[3] = psm[2] + ctot[2];
    for (j = 2; j <= n; j++) {
        int jp = IX(idxp, j), ct = IX(coltyp, jp);
        IX(idxc, psm[ct - 1]) = j;
        psm[ct - 1]++;
    }
    for (j = 2; j <= n; j++) {
        int jp = IX(idxp, j);
        dsigma[j - 1] = D_(jp);
        int idxj = IX(idxq, IX(idx, IX(idxp, IX(idxc, j))) + 1);
        if (idxj <= nlp1) idxj--;
        for (i = 1; i <= n; i++) U2_(i, j) = U_(i, idxj);
        for (i = 1; i <= m; i++) VT2_(j, i) = VT_(idxj, i);
    }
    dsigma[0] = 0.0;
    hlftol = tol / 2.0;
    if (fabs(dsigma[1]) <= hlftol) dsigma[1] = hlftol;
    if (m > n) {
        Z_(1) = lapy2(z1, Z_(m));
        if (Z_(1) <= tol) {
            c = 1.0; s = 0.0; Z_(1) = tol;
        } else {
            c = z1 / Z_(1);
            s = Z_(m) / Z_(1);
        }
    } else {
        c = 1.0; s = 0.0;
        Z_(1) = (fabs(z1) <= tol) ? tol : z1;
    }
    for (i = 2; i <= k; i++) Z_(i) = U2_(i, 1);
    for (i = 1; i <= n; i++) U2_(i, 1) = 0.0;
    U2_(nlp1, 1) = 1.0;
    if (m > n) {
        for (i = 1; i <= nlp1; i++) {
            VT_(m, i) = -s * VT_(nlp1, i);
            VT2_(1, i) = c * VT_(nlp1, i);
        }
        for (i = nlp2; i <= m; i++) {
            VT2_(1, i) = s * VT_(m, i);
            VT_(m, i) = c * VT_(m, i);
        }
    } else {
        for (i = 1; i <= m; i++) VT2_(1, i) = VT_(nlp1, i);
    }
    if (m > n)
        for (i = 1; i <= m; i++) VT2_(m, i) = VT_(m, i);
    if (n > k) {
        for (i = k + 1; i <= n; i++) D_(i) = dsigma[i - 1];
        for (j = k + 1; j <= n; j++)
            for (i = 1; i <= n; i++) U_(i, j) = U2_(i, j);
        for (j = 1; j <= m; j++)
            for (i = k + 1; i <= n; i++) VT_(i, j) = VT2_(i, j);
    }
    for (j = 0; j < 4; j++) IX(coltyp, j + 1) = ctot[j];
    if (nrot) *nrot = nr_;
    *kout = k;
    return 0;
#undef IX
}

/* ------------------------------------------------------------------ */
/* The secular half of DLASD3 / MDLASD8: z scaled to unit norm (T-lane sums), the
 * roots d(1:k) with their origins and offsets, and z replaced by the updated z of the
 * Loewner products (DLASD3's order); zs keeps the signs of the input z.  Returns 0 or
 * the 1-based index of a root that did not converge. */
static int sv_solve(int k, double *d, const double *dsigma, double *z, double *zs, int *org,
                    double *tau, int maxit, long long *iters)
{
    int i, j;
    memcpy(zs, z, sizeof(double) * k);
    const int T = orc_sec_team(k);
    double ps[64];
    for (int l = 0; l < T; l++) {
        ps[l] = 0.0;
        for (i = l; i < k; i += T) ps[l] += z[i] * z[i];
    }
    double rho = sqrt(tsumT(ps, T));
    orc_lascl(rho, 1.0, k, z);
    rho = rho * rho;
    for (j = 1; j <= k; j++) {
        int nit = 0;
        double sg;
        int r = orc_lasd4(k, j, dsigma, z, rho, &sg, &org[j - 1], &tau[j - 1], &nit, maxit);
        d[j - 1] = sg;
        if (iters) *iters += nit;
        if (r) return j;
    }
    for (i = 1; i <= k; i++) {
        double di = dsigma[i - 1];
        double zi = sv_den(di, dsigma[org[k - 1] - 1], tau[k - 1]);
        for (j = 1; j <= i - 1; j++)
            zi = zi * (sv_den(di, dsigma[org[j - 1] - 1], tau[j - 1]) / (di - dsigma[j - 1]) /
                       (di + dsigma[j - 1]));
        for (j = i; j <= k - 1; j++)
            zi = zi * (sv_den(di, dsigma[org[j - 1] - 1], tau[j - 1]) / (di - dsigma[j]) /
                       (di + dsigma[j]));
        double a = sqrt(fabs(zi));
        z[i - 1] = zs[i - 1] >= 0.0 ? a : -a;
    }
    return 0;
}

/* 1 / 2-norms of the left vector (-1, d_2 w_2, ..., d_k w_k) and the right vector
 * (w_1, ..., w_k) of root i, w_j = z_j / ((d_j - sigma_i)(d_j + sigma_i)), in T-lane
 * interleaved sums; wv receives w. */
static void sv_norms(int k, const double *dsigma, const double *z, double dorg, double ti, double *wv,
                     double *tu, double *tv)
{
    const int T = orc_sec_team(k);
    double ps[64];
    int j;
    for (j = 1; j <= k; j++) {
        double dj = dsigma[j - 1];
        wv[j - 1] = z[j - 1] / ((dj - dorg) - ti) / ((dj + dorg) + ti);
    }
    for (int l = 0; l < T; l++) {
        ps[l] = 0.0;
        for (j = l; j < k; j += T) {
            double x = j == 0 ? -1.0 : dsigma[j] * wv[j];
            ps[l] += x * x;
        }
    }
    *tu = sqrt(tsumT(ps, T));
    for (int l = 0; l < T; l++) {
        ps[l] = 0.0;
        for (j = l; j < k; j += T) ps[l] += wv[j] * wv[j];
    }
    *tv = sqrt(tsumT(ps, T));
}

/* ------------------------------------------------------------------ */
/* DLASD3 (the dense branch, mdlasd1.f90:239-244): roots, updated z, singular vectors
 * of the deflated problem and their products with U2 / VT2.  q is workspace of
 * 2 k^2 + 3 k doubles. */
int orc_lasd3(int nl, int nr, int sqre, int k, double *d, double *q, const double *dsigma,
              double *u, int ldu, const double *u2, int ldu2, double *vt, int ldvt,
              const double *vt2, int ldvt2, const int *idxc, const int *ctot, double *z, int maxit,
              long long *iters)
{
    const int n = nl + nr + 1, m = n + sqre;
    int i, j;
    if (k == 1) {
        D_(1) = fabs(Z_(1));
        for (i = 1; i <= m; i++) VT_(1, i) = VT2_(1, i);
        for (i = 1; i <= n; i++) U_(i, 1) = Z_(1) > 0.0 ? U2_(i, 1) : -U2_(i, 1);
        return 0;
    }
    double *qu = q, *qv = q + (long)k * k, *zs = qv + (long)k * k, *tau = zs + k, *wv = tau + k;
    int *org = (int *)malloc(sizeof(int) * k);
    const int T = orc_sec_team(k);
    double ps[64];
    int info = sv_solve(k, d, dsigma, z, zs, org, tau, maxit, iters);
    if (info) { free(org); return info; }
    /* singular vectors of the deflated problem: column i of qu (left, rows in IDXC
     * order), row i of qv (right) */
    for (i = 1; i <= k; i++) {
        double tu, tv;
        sv_norms(k, dsigma, z, dsigma[org[i - 1] - 1], tau[i - 1], wv, &tu, &tv);
        qu[(long)(i - 1) * k] = -1.0 / tu;
        for (j = 2; j <= k; j++) {
            int jc = idxc[j - 1];
            qu[(j - 1) + (long)(i - 1) * k] = dsigma[jc - 1] * wv[jc - 1] / tu;
        }
        qv[(i - 1) + 0] = wv[0] / tv;
        for (j = 2; j <= k; j++) qv[(i - 1) + (long)(j - 1) * k] = wv[idxc[j - 1] - 1] / tv;
    }
    /* U(1:n,1:k) = U2(1:n,1:k) * qu ; VT(1:k,1:m) = qv * VT2(1:k,1:m), sums over the
     * inner index in ascending order */
    for (i = 1; i <= k; i++)
        for (int r = 1; r <= n; r++) {
            double acc = 0.0;
            for (j = 1; j <= k; j++) acc = acc + U2_(r, j) * qu[(j - 1) + (long)(i - 1) * k];
            U_(r, i) = acc;
        }
    for (int c = 1; c <= m; c++)
        for (i = 1; i <= k; i++) {
            double acc = 0.0;
            for (j = 1; j <= k; j++) acc = acc + VT2_(j, c) * qv[(i - 1) + (long)(j - 1) * k];
            VT_(i, c) = acc;
        }
    (void)ctot;
    (void)T;
    (void)ps;
    free(org);
    return 0;
}
#undef D_
#undef Z_
#undef U_
#undef VT_
#undef U2_
#undef VT2_

/* ------------------------------------------------------------------ */
/* MDLASD8 (mdlasd8.f90:201-278): roots d(1:k), the updated z, DIFL(j) = sigma_j - d_j
 * and DIFR(j) = sigma_j - d_{j+1} (from the root's origin and offset; DIFR(k) = 0) and
 * the normalisers APHAU / APHAV of the left / right vectors (1 / 2-norm, the dense
 * branch's sums).  Returns 0 or the index of a root that did not converge. */
int orc_lasd8(int k, double *d, const double *dsigma, double *z, double *difl, double *difr,
              double *aphau, double *aphav, int maxit, long long *iters)
{
    double *zs = malloc(sizeof(double) * 3 * (size_t)k), *tau = zs + k, *wv = tau + k;
    int *org = malloc(sizeof(int) * k);
    int info = sv_solve(k, d, dsigma, z, zs, org, tau, maxit, iters);
    if (!info) {
        for (int j = 1; j <= k; j++) {
            const double dorg = dsigma[org[j - 1] - 1], tj = tau[j - 1];
            double tu, tv;
            difl[j - 1] = -((dsigma[j - 1] - dorg) - tj);
            difr[j - 1] = j < k ? -((dsigma[j] - dorg) - tj) : 0.0;
            sv_norms(k, dsigma, z, dorg, tj, wv, &tu, &tv);
            aphau[j - 1] = 1.0 / tu;
            aphav[j - 1] = 1.0 / tv;
        }
    }
    free(zs);
    free(org);
    return info;
}

/* An HSS tree of K columns in leaves of about ni has at least two leaves and at most
 * 512 (npart, aux_hss.f90:737-760). */
int orc_hss_tree_ok(int k, int ni)
{
    if (ni < 1) return 0;
    int l = k / ni;
    if (l < 1) return 0;
    if (k % ni >= ni / 2) l++;
    return l >= 2 && l <= 512;
}

/* The HSS branch of MDLASD1 (mdlasd1.f90:220-278, K >= BSMLZ): MDLASD8, then
 * DHSSVCS's two HSS forms in the squared variables (CauchyHssvd_VPS.f90:3587-3804):
 *   right  H_V(j,i) = z_j APHAV_i / (d_j^2 - sigma_i^2),   V(:,1:k) = V2 H_V
 *   left   H_U(j,i) = d_j z_j APHAU_i / (d_j^2 - sigma_i^2), U(:,1:k) = U2 H_U - U2(:,1) APHAU^T
 * (d_1 = 0, so row 1 of H_U is zero and the left vectors' first entry -APHAU_i enters as
 * the rank-one term: DLASD3's U(1,i) = -1 before normalisation).  U2 / VT2 come from
 * MDLASD2 with their columns / rows in IDXC order; they are taken back to the order of
 * DSIGMA here.  *rank gets the largest generator rank of the two forms. */
static int lasd3_hss(int nl, int nr, int sqre, int k, double *d, const double *dsigma, double *u,
                     int ldu, const double *u2, int ldu2, double *vt, int ldvt, const double *vt2,
                     int ldvt2, const int *idxc, double *z, const orc_params *p, long long *iters,
                     int *rank)
{
    const int n = nl + nr + 1, m = n + sqre;
    double *g = malloc(sizeof(double) * 5 * (size_t)k);
    double *difl = g, *difr = g + k, *au = g + 2 * k, *av = g + 3 * k, *dz = g + 4 * k;
    int info = orc_lasd8(k, d, dsigma, z, difl, difr, au, av, p->maxit, iters);
    if (info) { free(g); return info; }
    for (int j = 0; j < k; j++) dz[j] = dsigma[j] * z[j];
    double *xu = malloc(sizeof(double) * (size_t)n * k), *xv = malloc(sizeof(double) * (size_t)m * k);
    for (int j = 1; j <= k; j++) {
        const int c = j == 1 ? 1 : idxc[j - 1];
        for (int r = 0; r < n; r++) xu[r + (long)(c - 1) * n] = u2[r + (long)(j - 1) * ldu2];
        for (int r = 0; r < m; r++) xv[r + (long)(c - 1) * m] = vt2[(j - 1) + (long)r * ldvt2];
    }
    int rk = 0;
    orc_hss H;
    info = orc_hss_build_sq(&H, k, p->bd_ni, dsigma, d, z, av, difl, difr, p->tol);
    if (!info) {
        for (int i = 0; i < H.ltr - 1; i++) {
            if (H.usz[2 * i + 1] > rk) rk = H.usz[2 * i + 1];
            if (H.vsz[2 * i] > rk) rk = H.vsz[2 * i];
        }
        orc_hss_mmL(&H, xv, m, m);
        orc_hss_free(&H);
        info = orc_hss_build_sq(&H, k, p->bd_ni, dsigma, d, dz, au, difl, difr, p->tol);
    }
    if (!info) {
        for (int i = 0; i < H.ltr - 1; i++) {
            if (H.usz[2 * i + 1] > rk) rk = H.usz[2 * i + 1];
            if (H.vsz[2 * i] > rk) rk = H.vsz[2 * i];
        }
        double *u21 = malloc(sizeof(double) * n);
        memcpy(u21, xu, sizeof(double) * n);
        orc_hss_mmL(&H, xu, n, n);
        orc_hss_free(&H);
        for (int i = 0; i < k; i++)
            for (int r = 0; r < n; r++) u[r + (long)i * ldu] = xu[r + (long)i * n] - u21[r] * au[i];
        for (int i = 0; i < k; i++)
            for (int c = 0; c < m; c++) vt[i + (long)c * ldvt] = xv[c + (long)i * m];
        free(u21);
    }
    free(xu); free(xv); free(g);
    if (rank) *rank = rk;
    return info ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* MDLASD1 (mdlasd1.f90:1-291), the dense branch (DLASD3) for K < bd_bsmlz, the HSS branch above. */
int orc_lasd1(int nl, int nr, int sqre, double *d, double *alpha, double *beta, double *u, int ldu,
              double *vt, int ldvt, int *idxq, const orc_params *p, int *kout, long long *iters,
              int *nrot, int *hss_rank)
{
    const int n = nl + nr + 1, m = n + sqre;
    double orgnrm = fmax(fabs(*alpha), fabs(*beta));
    d[nl] = 0.0;
    for (int i = 0; i < n; i++)
        if (fabs(d[i]) > orgnrm) orgnrm = fabs(d[i]);
    orc_lascl(orgnrm, 1.0, n, d);
    *alpha = *alpha / orgnrm;
    *beta = *beta / orgnrm;
    double *z = malloc(sizeof(double) * m), *dsigma = malloc(sizeof(double) * n);
    double *u2 = malloc(sizeof(double) * (long)n * n), *vt2 = malloc(sizeof(double) * (long)m * m);
    int *iw = malloc(sizeof(int) * 4 * n);
    int *idx = iw, *idxc = iw + n, *coltyp = iw + 2 * n, *idxp = iw + 3 * n;
    int k;
    orc_lasd2(nl, nr, sqre, &k, d, z, *alpha, *beta, u, ldu, vt, ldvt, dsigma, u2, n, vt2, m, idxp,
              idx, idxc, idxq, coltyp, nrot, NULL, NULL, NULL, NULL);
    int info, rk = 0;
    if (k >= p->bd_bsmlz && orc_hss_tree_ok(k, p->bd_ni)) {
        info = lasd3_hss(nl, nr, sqre, k, d, dsigma, u, ldu, u2, n, vt, ldvt, vt2, m, idxc, z, p, iters, &rk);
    } else {
        double *q = malloc(sizeof(double) * (2 * (long)k * k + 3 * (long)k));
        info = orc_lasd3(nl, nr, sqre, k, d, q, dsigma, u, ldu, u2, n, vt, ldvt, vt2, m, idxc, coltyp,
                         z, p->maxit, iters);
        free(q);
    }
    if (hss_rank) *hss_rank = rk;
    free(z); free(dsigma); free(u2); free(vt2); free(iw);
    if (info) return info;
    orc_lascl(1.0, orgnrm, n, d);
    orc_dlamrg(k, n - k, d, 1, -1, idxq);
    if (kout) *kout = k;
    return 0;
}

/* MDLASD0 (mdlasd0.f90:1-231): leaves, then the merges level by level. */
int orc_lasd0(int n, int sqre, double *d, double *e, double *u, int ldu, double *vt, int ldvt,
              const orc_params *p, orc_trace *tr)
{
    const int smlsiz = p->smlsiz;
    if (n <= smlsiz) {
        int r = orc_bd_leaf(n, sqre, d, e, u, ldu, vt, ldvt);
        return r < 0 ? 1 : 0;
    }
    int *inode = malloc(sizeof(int) * n), *ndiml = malloc(sizeof(int) * n), *ndimr = malloc(sizeof(int) * n);
    int *idxq = malloc(sizeof(int) * n);
    int nlvl, nd, info = 0;
    orc_lasdt(n, &nlvl, &nd, inode, ndiml, ndimr, smlsiz);
    int ndb1 = (nd + 1) / 2;
    for (int i = ndb1; i <= nd; i++) {
        int ic = inode[i - 1], nl = ndiml[i - 1], nr = ndimr[i - 1];
        int nlf = ic - nl, nrf = ic + 1;
        if (orc_bd_leaf(nl, 1, d + nlf - 1, e + nlf - 1, u + (nlf - 1) + (long)(nlf - 1) * ldu, ldu,
                        vt + (nlf - 1) + (long)(nlf - 1) * ldvt, ldvt) < 0) { info = 1; goto out; }
        for (int j = 1; j <= nl; j++) idxq[nlf - 2 + j] = j;
        int sqrei = (i == nd) ? sqre : 1;
        if (orc_bd_leaf(nr, sqrei, d + nrf - 1, e + nrf - 1, u + (nrf - 1) + (long)(nrf - 1) * ldu, ldu,
                        vt + (nrf - 1) + (long)(nrf - 1) * ldvt, ldvt) < 0) { info = 1; goto out; }
        for (int j = 1; j <= nr; j++) idxq[ic + j - 1] = j;
    }
    for (int lvl = nlvl; lvl >= 1; lvl--) {
        int lf = (lvl == 1) ? 1 : 1 << (lvl - 1), ll = (lvl == 1) ? 1 : 2 * lf - 1;
        for (int i = lf; i <= ll; i++) {
            int ic = inode[i - 1], nl = ndiml[i - 1], nr = ndimr[i - 1];
            int nlf = ic - nl;
            int sqrei = (sqre == 0 && i == ll) ? sqre : 1;
            double alpha = d[ic - 1], beta = e[ic - 1];
            int k = 0, nrot = 0, rk = 0;
            long long it = 0;
            info = orc_lasd1(nl, nr, sqrei, d + nlf - 1, &alpha, &beta,
                             u + (nlf - 1) + (long)(nlf - 1) * ldu, ldu,
                             vt + (nlf - 1) + (long)(nlf - 1) * ldvt, ldvt, idxq + nlf - 1, p, &k, &it,
                             &nrot, &rk);
            if (tr) {
                if (tr->nmerge < tr->cap) {
                    tr->level[tr->nmerge] = lvl;
                    tr->n[tr->nmerge] = nl + nr + 1;
                    tr->k[tr->nmerge] = k;
                    if (tr->hss) tr->hss[tr->nmerge] = nrot;
                    if (tr->rank_max) tr->rank_max[tr->nmerge] = rk;
                }
                tr->nmerge++;
                tr->roots += k;
                tr->iters += it;
            }
            if (info) goto out;
        }
    }
out:
    free(inode); free(ndiml); free(ndimr); free(idxq);
    return info;
}

/* MDBDSDC (mdbdsdc.f90:1-422) with COMPQ = 'N' or 'I'.  u, vt n x n (ldu, ldvt).
 * 'N' computes the vectors into u, vt as well (they are then workspace). */
int orc_bdsdc(char uplo, char compq, int n, double *d, double *e, double *u, int ldu, double *vt,
              int ldvt, const orc_params *p, orc_trace *tr)
{
    int iuplo = (uplo == 'U' || uplo == 'u') ? 1 : (uplo == 'L' || uplo == 'l') ? 2 : 0;
    int icompq = (compq == 'N' || compq == 'n') ? 0 : (compq == 'I' || compq == 'i') ? 2 : -1;
    if (iuplo == 0) return -1;
    if (icompq < 0) return -2;
    if (n < 0) return -3;
    if (ldu < 1 || ldu < n) return -7;
    if (ldvt < 1 || ldvt < n) return -9;
    if (n == 0) return 0;
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) {
            u[i + (long)j * ldu] = i == j;
            vt[i + (long)j * ldvt] = i == j;
        }
    }
    if (n == 1) {
        u[0] = d[0] >= 0.0 ? 1.0 : -1.0;
        vt[0] = 1.0;
        d[0] = fabs(d[0]);
        return 0;
    }
    const int nm1 = n - 1;
    double *rc = NULL, *rs = NULL;
    if (iuplo == 2) {
        rc = malloc(sizeof(double) * n);
        rs = malloc(sizeof(double) * n);
        for (int i = 0; i < n - 1; i++) {
            double cs, sn, r;
            orc_lartg(d[i], e[i], &cs, &sn, &r);
            d[i] = r;
            e[i] = sn * d[i + 1];
            d[i + 1] = cs * d[i + 1];
            rc[i] = cs;
            rs[i] = -sn;
        }
    }
    int info = 0;
    if (n <= p->smlsiz) {
        info = orc_bd_leaf(n, 0, d, e, u, ldu, vt, ldvt) < 0 ? 1 : 0;
    } else {
        double orgnrm = 0.0;
        for (int i = 0; i < n; i++) orgnrm = fmax(orgnrm, fabs(d[i]));
        for (int i = 0; i < nm1; i++) orgnrm = fmax(orgnrm, fabs(e[i]));
        if (orgnrm == 0.0) goto sort;
        orc_lascl(orgnrm, 1.0, n, d);
        orc_lascl(orgnrm, 1.0, nm1, e);
        const double eps = 0.9 * EPS;
        for (int i = 0; i < n; i++)
            if (fabs(d[i]) < eps) d[i] = d[i] >= 0.0 ? eps : -eps;
        int start = 1;
        for (int i = 1; i <= nm1; i++) {
            if (fabs(e[i - 1]) < eps || i == nm1) {
                int nsize;
                if (i < nm1) {
                    nsize = i - start + 1;
                } else if (fabs(e[i - 1]) >= eps) {
                    nsize = n - start + 1;
                } else {
                    nsize = i - start + 1;
                    u[(n - 1) + (long)(n - 1) * ldu] = d[n - 1] >= 0.0 ? 1.0 : -1.0;
                    vt[(n - 1) + (long)(n - 1) * ldvt] = 1.0;
                    d[n - 1] = fabs(d[n - 1]);
                }
                info = orc_lasd0(nsize, 0, d + start - 1, e + start - 1,
                                 u + (start - 1) + (long)(start - 1) * ldu, ldu,
                                 vt + (start - 1) + (long)(start - 1) * ldvt, ldvt, p, tr);
                if (info) goto fin;
                start = i + 1;
            }
        }
        orc_lascl(1.0, orgnrm, n, d);
    }
sort:
    for (int ii = 2; ii <= n; ii++) {
        int i = ii - 1, kk = i;
        double pv = d[i - 1];
        for (int j = ii; j <= n; j++)
            if (d[j - 1] > pv) { kk = j; pv = d[j - 1]; }
        if (kk != i) {
            d[kk - 1] = d[i - 1];
            d[i - 1] = pv;
            for (int r = 0; r < n; r++) {
                double t = u[r + (long)(i - 1) * ldu];
                u[r + (long)(i - 1) * ldu] = u[r + (long)(kk - 1) * ldu];
                u[r + (long)(kk - 1) * ldu] = t;
                t = vt[(i - 1) + (long)r * ldvt];
                vt[(i - 1) + (long)r * ldvt] = vt[(kk - 1) + (long)r * ldvt];
                vt[(kk - 1) + (long)r * ldvt] = t;
            }
        }
    }
    if (iuplo == 2) {
        /* DLASR('L','V','B'): rows j, j+1 of U rotated for j = n-1 .. 1 */
        for (int j = n - 2; j >= 0; j--) {
            double ct = rc[j], st = rs[j];
            for (int c = 0; c < n; c++) {
                double t = u[(j + 1) + (long)c * ldu];
                u[(j + 1) + (long)c * ldu] = ct * t - st * u[j + (long)c * ldu];
                u[j + (long)c * ldu] = st * t + ct * u[j + (long)c * ldu];
            }
        }
    }
fin:
    free(rc);
    free(rs);
    (void)icompq;
    return info;
}
