// Learned starting iterate for the secular equation: features of one root and
// the forward pass of the FP32 network (inputs 34 -> 64 -> 64 -> 1, ReLU).
//
// Inputs of root r (0-based) with bracket width W (d[r+1]-d[r], or rho*|z|^2 for
// the last root):
//   0..15  the 16 nearest poles other than the bracket ends, as
//          u/(1+|u|) with u = W/(d[r+t]-d[r]), t = -8..-1, 2..9 (0 if absent)
//   16..31 z^2 of the 16 poles t = -7..8, normalised to sum 1
//   32     q/(1+q), q = W/(rho * sum of those z^2)
//   33     the classical two-pole iterate as a bracket fraction
// Output y: the log of the ratio of the root's offset from its origin pole to the
// classical starting offset, so the proposal is tau = tau_classical * exp(y) with y
// clamped to [-16, 16]; relative precision holds for roots at any distance from a pole.
// All arithmetic is plain IEEE double/float with no contraction and exp is the
// series below, so the CPU restatement gives bit-identical proposals.
#pragma once
#include "hsdc_internal.h"

#define MLP_IN 34
#define MLP_H 64
#define MLP_W1 0
#define MLP_B1 (MLP_W1 + MLP_H * MLP_IN)
#define MLP_W2 (MLP_B1 + MLP_H)
#define MLP_B2 (MLP_W2 + MLP_H * MLP_H)
#define MLP_W3 (MLP_B2 + MLP_H)
#define MLP_B3 (MLP_W3 + MLP_H)
#define MLP_NPARAM (MLP_B3 + 1)

__device__ inline void sec_features(int n, int r, const double *d, const double *z, double rho, double W,
                                    double xc, float *f)
{
    const int T[16] = {-8, -7, -6, -5, -4, -3, -2, -1, 2, 3, 4, 5, 6, 7, 8, 9};
    for (int q = 0; q < 16; q++) {
        int idx = r + T[q];
        float v = 0.0f;
        if (idx >= 0 && idx < n) {
            double u = W / (d[idx] - d[r]);
            v = (float)(u / (1.0 + fabs(u)));
        }
        f[q] = v;
    }
    double w[16], S = 0.0;
    for (int q = 0; q < 16; q++) {
        int idx = r - 7 + q;
        w[q] = (idx >= 0 && idx < n) ? z[idx] * z[idx] : 0.0;
        S = S + w[q];
    }
    for (int q = 0; q < 16; q++) f[16 + q] = (float)(w[q] / S);
    double qv = W / (rho * S);
    f[32] = (float)(qv / (1.0 + qv));
    f[33] = (float)xc;
}

// sum over the T lanes of a team (aligned group of T lanes): butterfly, every lane
// ends with the same value
__device__ inline double team_sum(double v, int T) { return team_bfly(v, T); }
__device__ inline float team_sumf(float v, int T)
{
    for (int off = T >> 1; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// Forward pass shared by the T lanes of a team: lane l computes the hidden units
// u = l, l+T, ...; the first layer's outputs reach every lane by shuffles; y = b3 +
// the team sum of the lanes' partial w3.h2.  Each unit sums its inputs in the same
// order as the one-lane form.
template <int T>
__device__ inline float mlp_forward_team_t(const float *w, const float *x, int l)
{
    constexpr int NU = MLP_H / T;  // units per lane
    float h1[NU];
#pragma unroll
    for (int m = 0; m < NU; m++) {
        int u = l + m * T;
        float a = w[MLP_B1 + u];
        for (int k = 0; k < MLP_IN; k++) a = a + w[MLP_W1 + u * MLP_IN + k] * x[k];
        h1[m] = a > 0.0f ? a : 0.0f;
    }
    const int base = (threadIdx.x & 63) & ~(T - 1);
    float yp = 0.0f;
#pragma unroll 1
    for (int m = 0; m < NU; m++) {
        int u = l + m * T;
        float a = w[MLP_B2 + u];
        // unit v = q * T + src lives on lane src of the team, slot q: v ascending
#pragma unroll
        for (int q = 0; q < NU; q++) {
#pragma unroll 2
            for (int src = 0; src < T; src++) {
                float hv = __shfl(h1[q], base + src, 64);
                a = a + w[MLP_W2 + u * MLP_H + q * T + src] * hv;
            }
        }
        a = a > 0.0f ? a : 0.0f;
        yp = yp + w[MLP_W3 + u] * a;
    }
    return w[MLP_B3] + team_sumf(yp, T);
}

__device__ inline float mlp_forward_team(const float *w, const float *x, int T, int l)
{
    switch (T) {
    case 8: return mlp_forward_team_t<8>(w, x, l);
    case 16: return mlp_forward_team_t<16>(w, x, l);
    case 32: return mlp_forward_team_t<32>(w, x, l);
    default: return mlp_forward_team_t<64>(w, x, l);
    }
}

// e^x for |x| <= 16: 2^k e^r with k the nearest integer to x/ln 2, |r| <= ln 2 / 2, and
// e^r from its Taylor series to r^13 (relative error below 1e-15)
__device__ inline double sec_exp(double x)
{
    // ln 2 split in two (Cody and Waite): kf * LN2_HI is exact for |kf| < 2^11
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
