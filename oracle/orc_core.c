/*
 * CPU ORACLE -- test infrastructure only (see orc.h).  Restates the DC driver
 * chain of HSSPACK (mdstedc -> mdlaed0 -> mdlaed1 -> mdlaed2 -> {DLAED3 | mdlaed3})
 * and the LAPACK kernels it calls.  Compiled with -ffp-contract=off so that every
 * product and sum rounds separately, as the HIP kernels do.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define EPS 0x1p-53 /* DLAMCH('Epsilon') */

void orc_default_params(orc_params *p)
{
    p->smlsiz = 25;          /* ILAENV(9,'DSTEDC') */
    p->bsmlz = 1500;         /* mdlaed2.f90:175 */
    p->ni = 200;             /* mdlaed3.f90:253 */
    p->tol = (double)1.0e-17f; /* CauchyHssEig.f90:2523, a REAL literal */
    p->maxit = 100;
    p->bd_bsmlz = 2000;      /* mdlasd2.f90:172 */
    p->bd_ni = 40;           /* mdlasd8.f90:277 */
}

/* ------------------------------------------------------------------ */
/* DLAMRG: merge two sorted runs a(1..n1) (stride s1) and a(n1+1..n1+n2) (stride s2). */
void orc_dlamrg(int n1, int n2, const double *a, int s1, int s2, int *index)
{
    int n1sv = n1, n2sv = n2, ind1, ind2, i = 1;
    ind1 = (s1 > 0) ? 1 : n1;
    ind2 = (s2 > 0) ? 1 + n1 : n1 + n2;
    while (n1sv > 0 && n2sv > 0) {
        if (a[ind1 - 1] <= a[ind2 - 1]) {
            index[i - 1] = ind1; i++; ind1 += s1; n1sv--;
        } else {
            index[i - 1] = ind2; i++; ind2 += s2; n2sv--;
        }
    }
    if (n1sv == 0) {
        for (; n2sv > 0; n2sv--) { index[i - 1] = ind2; i++; ind2 += s2; }
    } else {
        for (; n1sv > 0; n1sv--) { index[i - 1] = ind1; i++; ind1 += s1; }
    }
}

/* sqrt(x^2+y^2) without destructive underflow/overflow (DLAPY2). */
static double lapy2(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    double r = z / w;
    return w * sqrt(1.0 + r * r);
}

/* ------------------------------------------------------------------ */
/* Secular equation solver.  Own implementation of the rational "middle way"
 * iteration (R.-C. Li, "Solving secular equations stably and efficiently",
 * LAWN 89) with a bracketing safeguard; the origin is the nearer pole, the
 * initial iterate comes from the two-pole model with the remaining poles frozen
 * at the interval midpoint.  Replaces LAPACK DLAED4 (called at mdlaed3.f90:208
 * and inside DLAED3, mdlaed1.f90:199). */
static void laed2x2(int j, const double *d, const double *z, double *delta, double rho,
                    double *dlam, int *org, double *tau)
{
    double del = d[1] - d[0], b, c, w, t, s;
    if (j == 1) {
        w = 1.0 + 2.0 * rho * (z[1] * z[1] - z[0] * z[0]) / del;
        if (w > 0.0) {
            b = del + rho * (z[0] * z[0] + z[1] * z[1]);
            c = rho * z[0] * z[0] * del;
            t = 2.0 * c / (b + sqrt(fabs(b * b - 4.0 * c)));
            *dlam = d[0] + t; *org = 1; *tau = t;
            delta[0] = -z[0] / t;
            delta[1] = z[1] / (del - t);
        } else {
            b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
            c = rho * z[1] * z[1] * del;
            if (b > 0.0) t = -2.0 * c / (b + sqrt(b * b + 4.0 * c));
            else t = (b - sqrt(b * b + 4.0 * c)) / 2.0;
            *dlam = d[1] + t; *org = 2; *tau = t;
            delta[0] = -z[0] / (del + t);
            delta[1] = -z[1] / t;
        }
    } else {
        b = -del + rho * (z[0] * z[0] + z[1] * z[1]);
        c = rho * z[1] * z[1] * del;
        if (b > 0.0) t = (b + sqrt(b * b + 4.0 * c)) / 2.0;
        else t = 2.0 * c / (-b + sqrt(b * b + 4.0 * c));
        *dlam = d[1] + t; *org = 2; *tau = t;
        delta[0] = -z[0] / (del + t);
        delta[1] = -z[1] / t;
    }
    s = sqrt(delta[0] * delta[0] + delta[1] * delta[1]);
    delta[0] /= s;
    delta[1] /= s;
}

/* Lanes sharing one root of an n-root secular equation on the device: the largest
 * power of two <= 64 that leaves each lane at least 4 poles.  The device sums the
 * poles of a root in T interleaved partial sums (lane l takes i = l, l+T, ...)
 * that a butterfly combines; the oracle sums in the same order, so both produce
 * the same bits.  T = 1 is the plain left-to-right order. */
int orc_sec_team(int n)
{
    int t = 1;
    while (t < 64 && 8 * t <= n) t *= 2;
    return t;
}

/* butterfly over T partial sums: step off adds lane l ^ off to lane l; every lane
 * ends with the same value (addition commutes) */
static double tsum(double *v, int T)
{
    double nv[64];
    for (int off = T >> 1; off > 0; off >>= 1) {
        for (int l = 0; l < T; l++) nv[l] = v[l] + v[l ^ off];
        for (int l = 0; l < T; l++) v[l] = nv[l];
    }
    return v[0];
}

static float tsumf(float *v, int T)
{
    float nv[64];
    for (int off = T >> 1; off > 0; off >>= 1) {
        for (int l = 0; l < T; l++) nv[l] = v[l] + v[l ^ off];
        for (int l = 0; l < T; l++) v[l] = nv[l];
    }
    return v[0];
}

/* initial iterate for root j (1-based, n>=3) -> origin, tau, bracket */
static void sec_init(int n, int j, const double *d, const double *z, double rho, int T,
                     int *orgp, double *taup, double *lop, double *hip)
{
    double rhoinv = 1.0 / rho, ps[64];
    if (j == n) {
        /* root in (d_n, d_n + rho*|z|^2] */
        for (int l = 0; l < T; l++) {
            ps[l] = 0.0;
            for (int i = l; i < n; i += T) ps[l] += z[i] * z[i];
        }
        double zz = tsum(ps, T);
        double hi = rho * zz, mid = 0.5 * hi, dn = d[n - 1];
        for (int l = 0; l < T; l++) {
            ps[l] = 0.0;
            for (int i = l; i < n - 2; i += T) ps[l] += z[i] * z[i] / ((d[i] - dn) - mid);
        }
        double c = rhoinv + tsum(ps, T);
        double dl = dn - d[n - 2];
        double z1 = z[n - 2] * z[n - 2], z2 = z[n - 1] * z[n - 1];
        double t;
        if (c <= 0.0) {
            t = hi;
        } else {
            double b = c * dl - z1 - z2;
            double s = sqrt(fabs(b * b + 4.0 * c * z2 * dl));
            if (b <= 0.0) t = (s - b) / (2.0 * c);
            else t = 2.0 * z2 * dl / (b + s);
        }
        if (!(t > 0.0 && t <= hi)) t = mid;
        *orgp = n; *taup = t; *lop = 0.0; *hip = hi;
        return;
    }
    double del = d[j] - d[j - 1], mid = 0.5 * del;
    for (int l = 0; l < T; l++) {
        ps[l] = 0.0;
        for (int i = l; i < n; i += T) {
            if (i == j - 1 || i == j) continue;
            ps[l] += z[i] * z[i] / ((d[i] - d[j - 1]) - mid);
        }
    }
    double c = rhoinv + tsum(ps, T);
    double zl = z[j - 1] * z[j - 1], zr = z[j] * z[j];
    double fmid = c + (zr - zl) / mid;
    double t;
    if (fmid >= 0.0) {
        /* c t^2 - b t + cc = 0, small positive root */
        double b = c * del + zl + zr, cc = zl * del;
        double s = sqrt(fabs(b * b - 4.0 * c * cc));
        if (b > 0.0) t = 2.0 * cc / (b + s);
        else t = (b - s) / (2.0 * c);
        if (!(t > 0.0 && t <= mid)) t = 0.5 * mid;
        *orgp = j; *taup = t; *lop = 0.0; *hip = mid;
    } else {
        /* c t^2 + b t - cc = 0, negative root of small magnitude */
        double b = c * del - zl - zr, cc = zr * del;
        double s = sqrt(fabs(b * b + 4.0 * c * cc));
        if (b <= 0.0) t = -2.0 * cc / (s - b);
        else t = -(b + s) / (2.0 * c);
        if (!(t >= -mid && t < 0.0)) t = -0.5 * mid;
        *orgp = j + 1; *taup = t; *lop = -mid; *hip = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Learned starting iterate of the secular iteration: feature map of one root and
 * the FP32 network 34 -> 64 -> 64 -> 1 (ReLU).  Weight layout: W1 (64x34, row u =
 * hidden unit), b1 (64), W2 (64x64), b2 (64), w3 (64), b3.  Features of root r
 * (0-based) with bracket width W:
 *   0..15  u/(1+|u|), u = W/(d[r+t]-d[r]) for t = -8..-1, 2..9 (0 when absent)
 *   16..31 z[r+t]^2 for t = -7..8 normalised to sum 1 (0 when absent)
 *   32     q/(1+q), q = W/(rho * that sum)
 *   33     the classical starting iterate as a fraction of the bracket. */
void orc_mlp_features(int n, int r, const double *d, const double *z, double rho, double W, double xc,
                      float *f)
{
    static const int T[16] = {-8, -7, -6, -5, -4, -3, -2, -1, 2, 3, 4, 5, 6, 7, 8, 9};
    for (int q = 0; q < 16; q++) {
        int i = r + T[q];
        float v = 0.0f;
        if (i >= 0 && i < n) {
            double u = W / (d[i] - d[r]);
            v = (float)(u / (1.0 + fabs(u)));
        }
        f[q] = v;
    }
    double zw[16], S = 0.0;
    for (int q = 0; q < 16; q++) {
        int i = r - 7 + q;
        zw[q] = (i >= 0 && i < n) ? z[i] * z[i] : 0.0;
        S = S + zw[q];
    }
    for (int q = 0; q < 16; q++) f[16 + q] = (float)(zw[q] / S);
    double qv = W / (rho * S);
    f[32] = (float)(qv / (1.0 + qv));
    f[33] = (float)xc;
}

/* e^x for |x| <= 16: 2^k e^r, k the nearest integer to x / ln 2, e^r by its Taylor
 * series to r^13 in Horner form (the device computes the same sequence) */
double orc_sec_exp(double x)
{
    /* ln 2 split in two (Cody and Waite): kf * LN2_HI is exact for |kf| < 2^11 */
    const double LN2 = 0.69314718055994530942, LN2_HI = 6.93147180369123816490e-01,
                 LN2_LO = 1.90821492927058770002e-10;
    double kf = floor(x / LN2 + 0.5);
    double r = (x - kf * LN2_HI) - kf * LN2_LO;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp(p, (int)kf);
}

/* forward pass in the device's order for a team of T lanes: units as one sum each;
 * y = b3 + butterfly over the lanes' partial sums of w3.h2 (lane l: units l, l+T, ...) */
float orc_mlp_forward_team(const float *w, const float *x, int T)
{
    float h1[ORC_MLP_H], yp[64];
    for (int u = 0; u < ORC_MLP_H; u++) {
        float a = w[ORC_MLP_B1 + u];
        for (int k = 0; k < ORC_MLP_IN; k++) a = a + w[ORC_MLP_W1 + u * ORC_MLP_IN + k] * x[k];
        h1[u] = a > 0.0f ? a : 0.0f;
    }
    for (int l = 0; l < T; l++) {
        yp[l] = 0.0f;
        for (int u = l; u < ORC_MLP_H; u += T) {
            float a = w[ORC_MLP_B2 + u];
            for (int v = 0; v < ORC_MLP_H; v++) a = a + w[ORC_MLP_W2 + u * ORC_MLP_H + v] * h1[v];
            a = a > 0.0f ? a : 0.0f;
            yp[l] = yp[l] + w[ORC_MLP_W3 + u] * a;
        }
    }
    return w[ORC_MLP_B3] + tsumf(yp, T);
}

float orc_mlp_forward(const float *w, const float *x)
{
    return orc_mlp_forward_team(w, x, 64);
}

int orc_laed4(int n, int j, const double *d, const double *z, double *delta, double rho,
              double *dlam, int *org, double *tau, int *niter, int maxit)
{
    return orc_laed4_mlp(n, j, d, z, delta, rho, dlam, org, tau, niter, maxit, NULL, 0.0f, NULL);
}

int orc_laed4_mlp(int n, int j, const double *d, const double *z, double *delta, double rho,
                  double *dlam, int *org, double *tau, int *niter, int maxit, const float *wmlp,
                  float gate, int *used)
{
    if (niter) *niter = 0;
    if (used) *used = 0;
    if (n == 1) {
        *dlam = d[0] + rho * z[0] * z[0];
        delta[0] = 1.0;
        if (org) *org = 1;
        if (tau) *tau = rho * z[0] * z[0];
        return 0;
    }
    if (n == 2) {
        int o; double t;
        laed2x2(j, d, z, delta, rho, dlam, &o, &t);
        if (org) *org = o;
        if (tau) *tau = t;
        return 0;
    }
    int ii;
    double t, lo, hi;
    const int T = orc_sec_team(n);
    sec_init(n, j, d, z, rho, T, &ii, &t, &lo, &hi);
    if (wmlp && T >= 8) {
        /* the proposal replaces the classical start only inside its bracket */
        double wb = (j == n) ? hi : d[j] - d[j - 1];
        double xc = (ii == j) ? t / wb : 1.0 + t / wb;
        float fl[ORC_MLP_IN];
        orc_mlp_features(n, j - 1, d, z, rho, wb, xc, fl);
        float y = orc_mlp_forward_team(wmlp, fl, T);
        y = y < -16.0f ? -16.0f : (y > 16.0f ? 16.0f : y);
        double tl = t * orc_sec_exp((double)y);
        if (fabsf(y) > gate && tl > lo && tl < hi) {
            t = tl;
            if (used) *used = 1;
        }
    }
    const double rhoinv = 1.0 / rho;
    const double dorg = d[ii - 1];
    const int split = (j == n) ? n - 1 : j; /* psi over 1..split, phi over split+1..n */
    /* the origin pole is the last psi pole (ii = j) or the last phi pole (ii = j+1) */
    const int na = split - ((ii - 1 < split) ? 1 : 0), nb = (n - split) - ((ii - 1 >= split) ? 1 : 0);
    int lgT = 0;
    while ((1 << lgT) < T) lgT++;
    int it, info = 1;
    for (it = 0; it <= maxit; it++) {
        /* sums over the poles left (psi, ascending) and right (phi, descending) of
         * the root in T interleaved partial sums, the origin pole's term kept apart
         * and added last */
        double P[6][64];
        for (int l = 0; l < T; l++) {
            double psi = 0.0, dpsi = 0.0, epsi = 0.0, phi = 0.0, dphi = 0.0, ephi = 0.0;
            for (int t2 = l; t2 < na; t2 += T) {
                double dl = (d[t2] - dorg) - t;
                double q = z[t2] / dl;
                psi += z[t2] * q;
                dpsi += q * q;
                epsi += psi;
            }
            for (int t2 = l; t2 < nb; t2 += T) {
                int i = n - 1 - t2;
                double dl = (d[i] - dorg) - t;
                double q = z[i] / dl;
                phi += z[i] * q;
                dphi += q * q;
                ephi += phi;
            }
            P[0][l] = psi; P[1][l] = dpsi; P[2][l] = epsi; P[3][l] = phi; P[4][l] = dphi; P[5][l] = ephi;
        }
        double psi = tsum(P[0], T), dpsi = tsum(P[1], T), epsi = tsum(P[2], T);
        double phi = tsum(P[3], T), dphi = tsum(P[4], T), ephi = tsum(P[5], T);
        double qo = z[ii - 1] / (-t);
        double torg = z[ii - 1] * qo;
        double w = ((rhoinv + phi) + psi) + torg;
        /* running-error bound of the evaluation of w, in the form of Li's analysis,
         * with log2(T) more additions per sum for the butterfly */
        double erretm = (8.0 + lgT) * (phi - psi) + (ephi - epsi) + 2.0 * rhoinv + 3.0 * fabs(torg) +
                        fabs(t) * ((dpsi + dphi) + qo * qo);
        if (ii - 1 < split) dpsi += qo * qo; else dphi += qo * qo;
        if (fabs(w) <= EPS * erretm || it == maxit) {
            if (fabs(w) <= EPS * erretm) info = 0;
            break;
        }
        if (w > 0.0) hi = t; else lo = t;
        double dl = (d[split - 1] - dorg) - t;
        double dr = (d[split] - dorg) - t;
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
        double tn = t + eta;
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (tn == t) { info = 0; break; }
        t = tn;
    }
    if (niter) *niter = it + 1;
    for (int i = 0; i < n; i++) delta[i] = (d[i] - dorg) - t;
    *dlam = dorg + t;
    if (org) *org = ii;
    if (tau) *tau = t;
    return info;
}

/* ------------------------------------------------------------------ */
/* Implicit-shift QL with Wilkinson shift (the classical tql2 iteration of
 * Bowdler, Martin, Reinsch and Wilkinson), eigenvalues sorted ascending with
 * the vectors by selection sort, like LAPACK DSTEQR's final step.  Used for the
 * leaves of the DC tree (mdlaed0.f90:263). */
static int ql_core(int n, double *d, double *e, double *z, int ldz, int want)
{
    for (int l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= EPS * dd) break;
            }
            if (m == l) break;
            if (++iter > 30 * n) return l + 1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = lapy2(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i, underflow = 0;
            for (i = m - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = lapy2(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (want) {
                    double *zi = z + (long)i * ldz, *zi1 = z + (long)(i + 1) * ldz;
                    for (int k = 0; k < n; k++) {
                        double fz = zi1[k];
                        zi1[k] = s * zi[k] + c * fz;
                        zi[k] = c * zi[k] - s * fz;
                    }
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    /* selection sort, ascending */
    for (int ii = 1; ii < n; ii++) {
        int i = ii - 1, k = i;
        double p = d[i];
        for (int j = ii; j < n; j++)
            if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            if (want) {
                double *zi = z + (long)i * ldz, *zk = z + (long)k * ldz;
                for (int r = 0; r < n; r++) { double t = zi[r]; zi[r] = zk[r]; zk[r] = t; }
            }
        }
    }
    return 0;
}

int orc_steqr(int n, double *d, double *e, double *z, int ldz)
{
    if (n <= 0) return 0;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) z[i + (long)j * ldz] = (i == j) ? 1.0 : 0.0;
    if (n == 1) return 0;
    double *ee = (double *)malloc(sizeof(double) * n);
    memcpy(ee, e, sizeof(double) * (n - 1));
    ee[n - 1] = 0.0;
    int info = ql_core(n, d, ee, z, ldz, 1);
    memcpy(e, ee, sizeof(double) * (n - 1));
    free(ee);
    return info;
}

int orc_sterf(int n, double *d, double *e)
{
    if (n <= 1) return 0;
    double *ee = (double *)malloc(sizeof(double) * n);
    memcpy(ee, e, sizeof(double) * (n - 1));
    ee[n - 1] = 0.0;
    int info = ql_core(n, d, ee, NULL, 1, 0);
    free(ee);
    return info;
}

/* ------------------------------------------------------------------ */
/* MDLAED2 (mdlaed2.f90:149-528).  Deflation of the rank-one merge.  The dense
 * branch (K < bsmlz) is LAPACK DLAED2; the HSS branch packs the K surviving
 * columns densely into Q2 (mdlaed2.f90:467-515). */
#define Q_(i, j) q[((i) - 1) + (long)((j) - 1) * ldq]
int orc_laed2(int *kout, int n, int n1, double *d, double *q, int ldq, int *indxq, double *rhop,
              double *z, double *dlamda, double *w, double *q2, int *indx, int *indxc,
              int *indxp, int *coltyp, int bsmlz, int *hflag,
              int *nrot, int *rot_pj, int *rot_nj, double *rot_c, double *rot_s)
{
    int ctot[4], psm[4];
    int k, k2, j, nj, pj = 0, i, ct, js, iq1, iq2, n2, n1p1;
    double c, s, t, tau, tol, rho = *rhop;
    *hflag = 1;
    if (nrot) *nrot = 0;
    if (n < 0) return -2;
    if (ldq < (n > 1 ? n : 1)) return -6;
    if ((1 < n / 2 ? 1 : n / 2) > n1 || n / 2 < n1) return -3;
    if (n == 0) return 0;
    n2 = n - n1;
    n1p1 = n1 + 1;
    if (rho < 0.0)
        for (i = n1p1; i <= n; i++) z[i - 1] = -1.0 * z[i - 1];
    t = 1.0 / sqrt(2.0);
    for (i = 0; i < n; i++) z[i] = t * z[i];
    rho = fabs(2.0 * rho);
    *rhop = rho;
    for (i = n1p1; i <= n; i++) indxq[i - 1] += n1;
    for (i = 1; i <= n; i++) dlamda[i - 1] = d[indxq[i - 1] - 1];
    orc_dlamrg(n1, n2, dlamda, 1, 1, indxc);
    for (i = 1; i <= n; i++) indx[i - 1] = indxq[indxc[i - 1] - 1];
    /* IDAMAX */
    int imax = 1, jmax = 1;
    for (i = 2; i <= n; i++) if (fabs(z[i - 1]) > fabs(z[imax - 1])) imax = i;
    for (i = 2; i <= n; i++) if (fabs(d[i - 1]) > fabs(d[jmax - 1])) jmax = i;
    double da = fabs(d[jmax - 1]), za = fabs(z[imax - 1]);
    tol = 8.0 * EPS * (da > za ? da : za);
    if (rho * fabs(z[imax - 1]) <= tol) {
        k = 0;
        iq2 = 0;
        for (j = 1; j <= n; j++) {
            i = indx[j - 1];
            memcpy(q2 + iq2, &Q_(1, i), sizeof(double) * n);
            dlamda[j - 1] = d[i - 1];
            iq2 += n;
        }
        for (j = 1; j <= n; j++) memcpy(&Q_(1, j), q2 + (long)(j - 1) * n, sizeof(double) * n);
        memcpy(d, dlamda, sizeof(double) * n);
        *kout = 0;
        return 0;
    }
    for (i = 1; i <= n1; i++) coltyp[i - 1] = 1;
    for (i = n1p1; i <= n; i++) coltyp[i - 1] = 3;
    k = 0;
    k2 = n + 1;
    for (j = 1; j <= n; j++) {
        nj = indx[j - 1];
        if (rho * fabs(z[nj - 1]) <= tol) {
            k2--;
            coltyp[nj - 1] = 4;
            indxp[k2 - 1] = nj;
            if (j == n) goto L100;
        } else {
            pj = nj;
            goto L80;
        }
    }
L80:
    j++;
    if (j > n) goto L100;
    nj = indx[j - 1];
    if (rho * fabs(z[nj - 1]) <= tol) {
        k2--;
        coltyp[nj - 1] = 4;
        indxp[k2 - 1] = nj;
    } else {
        s = z[pj - 1];
        c = z[nj - 1];
        tau = lapy2(c, s);
        t = d[nj - 1] - d[pj - 1];
        c = c / tau;
        s = -s / tau;
        if (fabs(t * c * s) <= tol) {
            z[nj - 1] = tau;
            z[pj - 1] = 0.0;
            if (coltyp[nj - 1] != coltyp[pj - 1]) coltyp[nj - 1] = 2;
            coltyp[pj - 1] = 4;
            /* DROT */
            for (i = 1; i <= n; i++) {
                double x = Q_(i, pj), y = Q_(i, nj);
                double tmp = c * x + s * y;
                Q_(i, nj) = c * y - s * x;
                Q_(i, pj) = tmp;
            }
            if (nrot) {  /* with no log arrays only the count is kept */
                if (rot_pj) {
                    rot_pj[*nrot] = pj; rot_nj[*nrot] = nj;
                    rot_c[*nrot] = c; rot_s[*nrot] = s;
                }
                (*nrot)++;
            }
            t = d[pj - 1] * (c * c) + d[nj - 1] * (s * s);
            d[nj - 1] = d[pj - 1] * (s * s) + d[nj - 1] * (c * c);
            d[pj - 1] = t;
            k2--;
            i = 1;
            for (;;) {
                if (k2 + i <= n) {
                    if (d[pj - 1] < d[indxp[k2 + i - 1] - 1]) {
                        indxp[k2 + i - 2] = indxp[k2 + i - 1];
                        indxp[k2 + i - 1] = pj;
                        i++;
                        continue;
                    } else {
                        indxp[k2 + i - 2] = pj;
                    }
                } else {
                    indxp[k2 + i - 2] = pj;
                }
                break;
            }
            pj = nj;
        } else {
            k++;
            dlamda[k - 1] = d[pj - 1];
            w[k - 1] = z[pj - 1];
            indxp[k - 1] = pj;
            pj = nj;
        }
    }
    goto L80;
L100:
    k++;
    dlamda[k - 1] = d[pj - 1];
    w[k - 1] = z[pj - 1];
    indxp[k - 1] = pj;

    for (j = 0; j < 4; j++) ctot[j] = 0;
    for (j = 1; j <= n; j++) ctot[coltyp[j - 1] - 1]++;
    psm[0] = 1;
    psm[1] = 1 + ctot[0];
    psm[2] = psm[1] + ctot[1];
    psm[3] = psm[2] + ctot[2];
    k = n - ctot[3];
    long liq1;
    if (k < bsmlz) {
        for (j = 1; j <= n; j++) {
            js = indxp[j - 1];
            ct = coltyp[js - 1];
            indx[psm[ct - 1] - 1] = js;
            indxc[psm[ct - 1] - 1] = j;
            psm[ct - 1]++;
        }
        i = 1;
        liq1 = 0;
        long liq2 = (long)(ctot[0] + ctot[1]) * n1;
        for (j = 1; j <= ctot[0]; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq1, &Q_(1, js), sizeof(double) * n1);
            z[i - 1] = d[js - 1];
            i++; liq1 += n1;
        }
        for (j = 1; j <= ctot[1]; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq1, &Q_(1, js), sizeof(double) * n1);
            memcpy(q2 + liq2, &Q_(n1 + 1, js), sizeof(double) * n2);
            z[i - 1] = d[js - 1];
            i++; liq1 += n1; liq2 += n2;
        }
        for (j = 1; j <= ctot[2]; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq2, &Q_(n1 + 1, js), sizeof(double) * n2);
            z[i - 1] = d[js - 1];
            i++; liq2 += n2;
        }
        liq1 = liq2;
        for (j = 1; j <= ctot[3]; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq2, &Q_(1, js), sizeof(double) * n);
            liq2 += n;
            z[i - 1] = d[js - 1];
            i++;
        }
        if (k < n) {
            for (j = 0; j < ctot[3]; j++)
                memcpy(&Q_(1, k + 1 + j), q2 + liq1 + (long)j * n, sizeof(double) * n);
            memcpy(d + k, z + k, sizeof(double) * (n - k));
        }
    } else {
        *hflag = 0;
        for (j = 1; j <= n; j++) indx[j - 1] = indxp[j - 1];
        i = 1;
        liq1 = 0;
        for (j = 1; j <= k; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq1, &Q_(1, js), sizeof(double) * n);
            z[i - 1] = d[js - 1];
            i++; liq1 += n;
        }
        long liq2 = (long)k * n;
        for (j = 1; j <= ctot[3]; j++) {
            js = indx[i - 1];
            memcpy(q2 + liq2, &Q_(1, js), sizeof(double) * n);
            liq2 += n;
            z[i - 1] = d[js - 1];
            i++;
        }
        for (j = 1; j <= n; j++) memcpy(&Q_(1, j), q2 + (long)(j - 1) * n, sizeof(double) * n);
        memcpy(d + k, z + k, sizeof(double) * (n - k));
    }
    for (j = 0; j < 4; j++) coltyp[j] = ctot[j];
    *kout = k;
    return 0;
}
#undef Q_

/* plain column-major GEMM C = A*B (beta = 0), k ascending */
static void gemm_nn(int m, int n, int k, const double *a, int lda, const double *b, int ldb,
                    double *c, int ldc)
{
    /* columns of C are independent: the threads split them, each column summed in
     * the same order as one thread would (the results do not depend on the count) */
#pragma omp parallel for schedule(static) if ((long)m * n * k > 4000000L)
    for (int j = 0; j < n; j++) {
        double *cj = c + (long)j * ldc;
        for (int i = 0; i < m; i++) cj[i] = 0.0;
        for (int l = 0; l < k; l++) {
            double blj = b[l + (long)j * ldb];
            const double *al = a + (long)l * lda;
            for (int i = 0; i < m; i++) cj[i] += al[i] * blj;
        }
    }
}

static double nrm2(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* LAPACK DLAED3 (called at mdlaed1.f90:199 for K < bsmlz). */
int orc_laed3(int k, int n, int n1, double *d, double *q, int ldq, double rho,
              double *dlamda, const double *q2, const int *indx, const int *ctot,
              double *w, double *s, int maxit, long long *iters)
{
#define Q_(i, j) q[((i) - 1) + (long)((j) - 1) * ldq]
    int i, j, info = 0, nit;
    if (k < 0) return -1;
    if (n < k) return -2;
    if (ldq < (n > 1 ? n : 1)) return -6;
    if (k == 0) return 0;
    for (j = 1; j <= k; j++) {
        info = orc_laed4(k, j, dlamda, w, &Q_(1, j), rho, &d[j - 1], NULL, NULL, &nit, maxit);
        if (iters) *iters += nit;
        if (info != 0) return info;
    }
    if (k == 1) goto L110;
    if (k == 2) {
        for (j = 1; j <= k; j++) {
            w[0] = Q_(1, j);
            w[1] = Q_(2, j);
            Q_(1, j) = w[indx[0] - 1];
            Q_(2, j) = w[indx[1] - 1];
        }
        goto L110;
    }
    memcpy(s, w, sizeof(double) * k);
    for (i = 1; i <= k; i++) w[i - 1] = Q_(i, i);
    for (j = 1; j <= k; j++) {
        for (i = 1; i <= j - 1; i++) w[i - 1] = w[i - 1] * (Q_(i, j) / (dlamda[i - 1] - dlamda[j - 1]));
        for (i = j + 1; i <= k; i++) w[i - 1] = w[i - 1] * (Q_(i, j) / (dlamda[i - 1] - dlamda[j - 1]));
    }
    for (i = 1; i <= k; i++) w[i - 1] = copysign(sqrt(-w[i - 1]), s[i - 1]);
    for (j = 1; j <= k; j++) {
        for (i = 1; i <= k; i++) s[i - 1] = w[i - 1] / Q_(i, j);
        double temp = nrm2(k, s);
        for (i = 1; i <= k; i++) Q_(i, j) = s[indx[i - 1] - 1] / temp;
    }
L110:;
    int n2 = n - n1, n12 = ctot[0] + ctot[1], n23 = ctot[1] + ctot[2];
    /* S := Q(ctot1+1 : ctot1+n23, 1:k) */
    for (j = 1; j <= k; j++)
        for (i = 1; i <= n23; i++) s[(i - 1) + (long)(j - 1) * n23] = Q_(ctot[0] + i, j);
    long iq2 = (long)n1 * n12;
    if (n23 != 0) gemm_nn(n2, k, n23, q2 + iq2, n2, s, n23, &Q_(n1 + 1, 1), ldq);
    else for (j = 1; j <= k; j++) for (i = 1; i <= n2; i++) Q_(n1 + i, j) = 0.0;
    for (j = 1; j <= k; j++)
        for (i = 1; i <= n12; i++) s[(i - 1) + (long)(j - 1) * n12] = Q_(i, j);
    if (n12 != 0) gemm_nn(n1, k, n12, q2, n1, s, n12, q, ldq);
    else for (j = 1; j <= k; j++) for (i = 1; i <= n1; i++) Q_(i, j) = 0.0;
    return 0;
#undef Q_
}

/* ------------------------------------------------------------------ */
/* MDLAED3 (mdlaed3.f90:124-261): roots, Loewner z (reference product order,
 * sqrt|prod|), alpha from DIFL/DIFR, then the HSS eigenvector update. */
int orc_mdlaed3(int k, int n, double *d, double *q, int ldq, double rho, double *dlamda,
                double *q2, double *z, double *difl, double *difr, double *alpha,
                const orc_params *p, int *rank_max, long long *iters)
{
    int i, j, info = 0, nit;
    if (k < 0) return -1;
    if (n < k) return -2;
    if (ldq < (n > 1 ? n : 1)) return -6;
    if (k == 0) return 0;
    double *work = (double *)malloc(sizeof(double) * 2 * (k + 1));
    double *prod = work + k + 1;
    for (i = 0; i < k; i++) prod[i] = 1.0;
    for (j = 1; j <= k; j++) {
        info = orc_laed4(k, j, dlamda, z, work, rho, &d[j - 1], NULL, NULL, &nit, p->maxit);
        if (iters) *iters += nit;
        if (info != 0) { free(work); return info; }
        difl[j - 1] = -work[j - 1];
        difr[j - 1] = (j < k) ? -work[j] : -prod[0];
        prod[j - 1] = prod[j - 1] * work[j - 1];
        for (i = 1; i <= j - 1; i++)
            prod[i - 1] = prod[i - 1] * (work[i - 1] / (dlamda[i - 1] - dlamda[j - 1]));
        for (i = j + 1; i <= k; i++)
            prod[i - 1] = prod[i - 1] * (work[i - 1] / (dlamda[i - 1] - dlamda[j - 1]));
    }
    for (i = 0; i < k; i++) z[i] = copysign(sqrt(fabs(prod[i])), z[i]);
    for (j = 1; j <= k; j++) {
        double difl_j = difl[j - 1], dlamj = -dlamda[j - 1], difr_j = 0.0, dlamjp = 0.0;
        if (j < k) { difr_j = -difr[j - 1]; dlamjp = -dlamda[j]; }
        work[j - 1] = z[j - 1] / difl_j;
        for (i = 1; i <= j - 1; i++) work[i - 1] = z[i - 1] / ((dlamda[i - 1] + dlamj) - difl_j);
        for (i = j + 1; i <= k; i++) work[i - 1] = z[i - 1] / ((dlamda[i - 1] + dlamjp) + difr_j);
        alpha[j - 1] = 1.0 / nrm2(k, work);
    }
    free(work);
    if (!q2) return 0;
    orc_hss H;
    info = orc_hss_build(&H, k, p->ni, dlamda, d, z, alpha, difl, difr, p->tol);
    if (info == 0) {
        if (rank_max) {
            int r = 0;
            for (i = 0; i < H.ltr; i++) {
                if (H.usz[2 * i + 1] > r) r = H.usz[2 * i + 1];
                if (H.vsz[2 * i] > r) r = H.vsz[2 * i];
            }
            *rank_max = r;
        }
        orc_hss_mmL(&H, q2, n, n);
        orc_hss_free(&H);
    } else {
        /* a tree of one leaf has no off-diagonal blocks: fastHssmmL_omp would
         * return X unchanged (ConstructHssd.f90:1455-1458); apply the dense
         * Cauchy-like matrix instead */
        double *cmat = (double *)malloc(sizeof(double) * (size_t)k * k);
        double *x = (double *)malloc(sizeof(double) * (size_t)n * k);
        orc_cauchy_dense(k, dlamda, z, alpha, difl, difr, cmat, k);
        memcpy(x, q2, sizeof(double) * (size_t)n * k);
        gemm_nn(n, k, k, x, n, cmat, k, q2, n);
        free(cmat);
        free(x);
        info = 0;
    }
    for (j = 0; j < k; j++) memcpy(q + (long)j * ldq, q2 + (long)j * n, sizeof(double) * n);
    return info;
}

int orc_mdlaed3_vectors(int k, double *d, double rho, double *dlamda, double *z, double *difl,
                        double *difr, double *alpha, int maxit, long long *iters)
{
    orc_params p;
    orc_default_params(&p);
    p.maxit = maxit;
    return orc_mdlaed3(k, k, d, NULL, k, rho, dlamda, NULL, z, difl, difr, alpha, &p, NULL, iters);
}

/* ------------------------------------------------------------------ */
/* MDLAED1 (mdlaed1.f90:109-233); *nrotout (may be NULL) receives the deflation
 * rotations MDLAED2 applied */
static int laed1_rot(int n, double *d, double *q, int ldq, int *indxq, double rho, int cutpnt,
                     double *work, int *iwork, const orc_params *p, int *kout, int *hssout,
                     int *rank_max, long long *iters, int *nrotout)
{
    int info, k = 0, hflag = 1, i;
    if (n < 0) return -1;
    if (ldq < (n > 1 ? n : 1)) return -4;
    if ((1 < n / 2 ? 1 : n / 2) > cutpnt || n / 2 < cutpnt) return -7;
    if (n == 0) return 0;
    double *zz = work, *dlamda = work + n, *w = work + 2 * n, *q2 = work + 3 * n;
    int *indx = iwork, *indxc = iwork + n, *coltyp = iwork + 2 * n, *indxp = iwork + 3 * n;
    for (i = 0; i < cutpnt; i++) zz[i] = q[(cutpnt - 1) + (long)i * ldq];
    for (i = cutpnt; i < n; i++) zz[i] = q[cutpnt + (long)i * ldq];
    int nrot = 0;
    info = orc_laed2(&k, n, cutpnt, d, q, ldq, indxq, &rho, zz, dlamda, w, q2, indx, indxc,
                     indxp, coltyp, p->bsmlz, &hflag, &nrot, NULL, NULL, NULL, NULL);
    if (kout) *kout = k;
    if (nrotout) *nrotout = nrot;
    if (hssout) *hssout = (k != 0 && !hflag);
    if (info != 0) return info;
    if (k != 0) {
        if (hflag) {
            long is = (long)(coltyp[0] + coltyp[1]) * cutpnt + (long)(coltyp[1] + coltyp[2]) * (n - cutpnt);
            info = orc_laed3(k, n, cutpnt, d, q, ldq, rho, dlamda, q2, indxc, coltyp, w,
                             q2 + is, p->maxit, iters);
        } else {
            /* mdlaed1.f90:205-207 places DIFL/DIFR/ALPHA at WORK(IQ2+K*N), which runs
             * past the documented LWORK = 1+4N+N**2 when K is close to N; they get
             * their own storage here. */
            double *difl = (double *)malloc(sizeof(double) * 3 * (size_t)k);
            double *difr = difl + k, *alpha = difr + k;
            info = orc_mdlaed3(k, n, d, q, ldq, rho, dlamda, q2, w, difl, difr, alpha, p,
                               rank_max, iters);
            free(difl);
        }
        if (info != 0) return info;
        orc_dlamrg(k, n - k, d, 1, -1, indxq);
    } else {
        for (i = 1; i <= n; i++) indxq[i - 1] = i;
    }
    return 0;
}

int orc_laed1(int n, double *d, double *q, int ldq, int *indxq, double rho, int cutpnt,
              double *work, int *iwork, const orc_params *p, int *kout, int *hssout,
              int *rank_max, long long *iters)
{
    return laed1_rot(n, d, q, ldq, indxq, rho, cutpnt, work, iwork, p, kout, hssout, rank_max, iters, NULL);
}

/* ------------------------------------------------------------------ */
/* MDLAED0 with ICOMPQ = 2 (mdlaed0.f90:121-380) */
int orc_laed0(int n, double *d, double *e, double *q, int ldq, double *work, int *iwork,
              const orc_params *p, orc_trace *tr)
{
    int subpbs, tlvls, i, j, spm1, submat, smm1, matsiz, msd2, info = 0;
    if (n < 0) return -3;
    if (ldq < (n > 1 ? n : 1)) return -7;
    if (n == 0) return 0;
    int smlsiz = p->smlsiz;
    iwork[0] = n;
    subpbs = 1;
    tlvls = 0;
    while (iwork[subpbs - 1] > smlsiz) {
        for (j = subpbs; j >= 1; j--) {
            iwork[2 * j - 1] = (iwork[j - 1] + 1) / 2;
            iwork[2 * j - 2] = iwork[j - 1] / 2;
        }
        tlvls++;
        subpbs *= 2;
    }
    for (j = 2; j <= subpbs; j++) iwork[j - 1] += iwork[j - 2];
    spm1 = subpbs - 1;
    for (i = 1; i <= spm1; i++) {
        submat = iwork[i - 1] + 1;
        smm1 = submat - 1;
        d[smm1 - 1] -= fabs(e[smm1 - 1]);
        d[submat - 1] -= fabs(e[smm1 - 1]);
    }
    int indxq = 4 * n + 3; /* 1-based position in IWORK */
    for (i = 0; i <= spm1; i++) {
        if (i == 0) { submat = 1; matsiz = iwork[0]; }
        else { submat = iwork[i - 1] + 1; matsiz = iwork[i] - iwork[i - 1]; }
        info = orc_steqr(matsiz, &d[submat - 1], &e[submat - 1],
                         &q[(submat - 1) + (long)(submat - 1) * ldq], ldq);
        if (info != 0) goto L130;
        int kk = 1;
        for (j = submat; j <= iwork[i]; j++) { iwork[indxq + j - 1] = kk; kk++; }
    }
    int curlvl = 1;
    while (subpbs > 1) {
        int spm2 = subpbs - 2;
        for (i = 0; i <= spm2; i += 2) {
            if (i == 0) { submat = 1; matsiz = iwork[1]; msd2 = iwork[0]; }
            else { submat = iwork[i - 1] + 1; matsiz = iwork[i + 1] - iwork[i - 1]; msd2 = matsiz / 2; }
            int kk = 0, hs = 0, rmax = 0, nr = 0;
            long long it0 = tr ? tr->iters : 0, itl = 0;
            info = laed1_rot(matsiz, &d[submat - 1], &q[(submat - 1) + (long)(submat - 1) * ldq], ldq,
                             &iwork[indxq + submat - 1], e[submat + msd2 - 2], msd2, work,
                             &iwork[subpbs], p, &kk, &hs, &rmax, &itl, &nr);
            if (tr) {
                tr->iters = it0 + itl;
                tr->roots += kk;
                if (tr->nmerge < tr->cap) {
                    tr->level[tr->nmerge] = curlvl; tr->n[tr->nmerge] = matsiz;
                    tr->k[tr->nmerge] = kk; tr->hss[tr->nmerge] = hs;
                    tr->rank_max[tr->nmerge] = rmax;
                    if (tr->nrot) tr->nrot[tr->nmerge] = nr;
                }
                tr->nmerge++;
            }
            if (info != 0) goto L130;
            iwork[i / 2] = iwork[i + 1];
        }
        subpbs /= 2;
        curlvl++;
    }
    if (tr && tr->indxq && tr->indxq_n == n) memcpy(tr->indxq, &iwork[indxq], sizeof(int) * n);
#pragma omp parallel for schedule(static)
    for (int ii = 1; ii <= n; ii++) {
        const int jj = iwork[indxq + ii - 1];
        work[ii - 1] = d[jj - 1];
        memcpy(work + (long)n * ii, &q[(long)(jj - 1) * ldq], sizeof(double) * n);
    }
    memcpy(d, work, sizeof(double) * n);
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < n; jj++) memcpy(&q[(long)jj * ldq], work + (long)n * (jj + 1), sizeof(double) * n);
    return 0;
L130:
    return submat * (n + 1) + submat + matsiz - 1;
}

/* ------------------------------------------------------------------ */
/* MDSTEDC (mdstedc.f90:139-436) */
int orc_stedc(char compz, int n, double *d, double *e, double *z, int ldz, double *work,
              int lwork, int *iwork, int liwork, const orc_params *p, orc_trace *tr)
{
    int icompz, info = 0, lwmin, liwmin, smlsiz = p->smlsiz;
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
    if (icompz == 0) return orc_sterf(n, d, e);
    if (icompz == 1) {
        /* Z_in * (eigenvectors of T): solve with 'I' into a copy, multiply back */
        double *zi = (double *)malloc(sizeof(double) * (long)n * n);
        double *zt = (double *)malloc(sizeof(double) * (long)n * n);
        for (int j = 0; j < n; j++) memcpy(zi + (long)j * n, z + (long)j * ldz, sizeof(double) * n);
        info = orc_stedc('I', n, d, e, zt, n, work, lwork, iwork, liwork, p, tr);
        if (info == 0) {
            double *c = (double *)malloc(sizeof(double) * (long)n * n);
            gemm_nn(n, n, n, zi, n, zt, n, c, n);
            for (int j = 0; j < n; j++) memcpy(z + (long)j * ldz, c + (long)j * n, sizeof(double) * n);
            free(c);
        }
        free(zi); free(zt);
        return info;
    }
    if (n <= smlsiz) return orc_steqr(n, d, e, z, ldz);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) z[i + (long)j * ldz] = (i == j) ? 1.0 : 0.0;
    double orgnrm = 0.0;
    for (int i = 0; i < n; i++) if (fabs(d[i]) > orgnrm || isnan(d[i])) orgnrm = fabs(d[i]);
    for (int i = 0; i < n - 1; i++) if (fabs(e[i]) > orgnrm || isnan(e[i])) orgnrm = fabs(e[i]);
    if (orgnrm == 0.0) return 0;
    int start = 1, finish, m = n;
    while (start <= n) {
        finish = start;
        while (finish < n) {
            double tiny = EPS * sqrt(fabs(d[finish - 1])) * sqrt(fabs(d[finish]));
            if (fabs(e[finish - 1]) > tiny) finish++;
            else break;
        }
        m = finish - start + 1;
        if (m == 1) { start = finish + 1; continue; }
        if (m > smlsiz) {
            double onrm = 0.0;
            for (int i = 0; i < m; i++) if (fabs(d[start - 1 + i]) > onrm) onrm = fabs(d[start - 1 + i]);
            for (int i = 0; i < m - 1; i++) if (fabs(e[start - 1 + i]) > onrm) onrm = fabs(e[start - 1 + i]);
            double mul = 1.0 / onrm;
            for (int i = 0; i < m; i++) d[start - 1 + i] *= mul;
            for (int i = 0; i < m - 1; i++) e[start - 1 + i] *= mul;
            info = orc_laed0(m, &d[start - 1], &e[start - 1],
                             &z[(start - 1) + (long)(start - 1) * ldz], ldz, work, iwork, p, tr);
            if (info != 0) {
                info = (info / (m + 1) + start - 1) * (n + 1) + info % (m + 1) + start - 1;
                return info;
            }
            for (int i = 0; i < m; i++) d[start - 1 + i] *= onrm;
        } else {
            info = orc_steqr(m, &d[start - 1], &e[start - 1],
                             &z[(start - 1) + (long)(start - 1) * ldz], ldz);
            if (info != 0) return start * (n + 1) + finish;
        }
        start = finish + 1;
    }
    if (m != n) {
        for (int ii = 2; ii <= n; ii++) {
            int i = ii - 1, k = i;
            double pp = d[i - 1];
            for (int j = ii; j <= n; j++) if (d[j - 1] < pp) { k = j; pp = d[j - 1]; }
            if (k != i) {
                d[k - 1] = d[i - 1];
                d[i - 1] = pp;
                for (int r = 0; r < n; r++) {
                    double t = z[r + (long)(i - 1) * ldz];
                    z[r + (long)(i - 1) * ldz] = z[r + (long)(k - 1) * ldz];
                    z[r + (long)(k - 1) * ldz] = t;
                }
            }
        }
    }
    return 0;
}
