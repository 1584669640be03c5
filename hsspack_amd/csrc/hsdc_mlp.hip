// Training kernels of the learned secular starting iterate (FP32, 34-64-64-1,
// ReLU).  Loss over a batch of B samples (features x_s, target t_s = log of the
// converged offset over the classical starting offset):
//   L = (1/B) sum_s (y_s - t_s)^2
// One block owns a contiguous slice of the batch and walks it in
// chunks of 32 samples held in LDS; every gradient entry is summed over samples
// in a fixed order, and block partials are added in block order, so a gradient is
// reproducible bit for bit.
#include <hip/hip_runtime.h>
#include "hsdc_mlp.h"

#define WG 256
#define CH 32
#define PPT ((MLP_NPARAM + WG - 1) / WG)

__global__ __launch_bounds__(WG) void k_mlp_grad(const float *wg, const float *X, const float *T, const int *idx,
                                                 int nb, float inv_b, float *partial, float *lossp)
{
    __shared__ float W[MLP_NPARAM];
    __shared__ float xs[CH][MLP_IN];
    __shared__ float h1[CH][MLP_H], h2[CH][MLP_H], d1[CH][MLP_H], d2[CH][MLP_H];
    __shared__ float gs[CH];
    __shared__ float lred[WG];
    const int tid = threadIdx.x;
    for (int i = tid; i < MLP_NPARAM; i += WG) W[i] = wg[i];
    const int per = (nb + gridDim.x - 1) / gridDim.x;
    const int s0 = blockIdx.x * per, s1 = min(nb, s0 + per);
    float acc[PPT];
    for (int q = 0; q < PPT; q++) acc[q] = 0.0f;
    float lacc = 0.0f;
    __syncthreads();
    for (int c0 = s0; c0 < s1; c0 += CH) {
        const int cn = min(CH, s1 - c0);
        for (int e = tid; e < cn * MLP_IN; e += WG) {
            int s = e / MLP_IN, k = e % MLP_IN;
            int src = idx ? idx[c0 + s] : c0 + s;
            xs[s][k] = X[(long)src * MLP_IN + k];
        }
        __syncthreads();
        for (int e = tid; e < cn * MLP_H; e += WG) {
            int s = e / MLP_H, u = e % MLP_H;
            float a = W[MLP_B1 + u];
            for (int k = 0; k < MLP_IN; k++) a = a + W[MLP_W1 + u * MLP_IN + k] * xs[s][k];
            h1[s][u] = a > 0.0f ? a : 0.0f;
        }
        __syncthreads();
        for (int e = tid; e < cn * MLP_H; e += WG) {
            int s = e / MLP_H, u = e % MLP_H;
            float a = W[MLP_B2 + u];
            for (int v = 0; v < MLP_H; v++) a = a + W[MLP_W2 + u * MLP_H + v] * h1[s][v];
            h2[s][u] = a > 0.0f ? a : 0.0f;
        }
        __syncthreads();
        if (tid < cn) {
            int s = tid;
            float y = W[MLP_B3];
            for (int u = 0; u < MLP_H; u++) y = y + W[MLP_W3 + u] * h2[s][u];
            int src = idx ? idx[c0 + s] : c0 + s;
            float r = y - T[src];
            gs[s] = 2.0f * r * inv_b;
            lacc = lacc + r * r * inv_b;
        }
        __syncthreads();
        for (int e = tid; e < cn * MLP_H; e += WG) {
            int s = e / MLP_H, u = e % MLP_H;
            d2[s][u] = (h2[s][u] > 0.0f) ? gs[s] * W[MLP_W3 + u] : 0.0f;
        }
        __syncthreads();
        for (int e = tid; e < cn * MLP_H; e += WG) {
            int s = e / MLP_H, v = e % MLP_H;
            float a = 0.0f;
            for (int u = 0; u < MLP_H; u++) a = a + W[MLP_W2 + u * MLP_H + v] * d2[s][u];
            d1[s][v] = (h1[s][v] > 0.0f) ? a : 0.0f;
        }
        __syncthreads();
        for (int q = 0; q < PPT; q++) {
            int p = tid + WG * q;
            if (p >= MLP_NPARAM) break;
            float a = 0.0f;
            if (p < MLP_B1) {
                int u = p / MLP_IN, k = p % MLP_IN;
                for (int s = 0; s < cn; s++) a = a + d1[s][u] * xs[s][k];
            } else if (p < MLP_W2) {
                int u = p - MLP_B1;
                for (int s = 0; s < cn; s++) a = a + d1[s][u];
            } else if (p < MLP_B2) {
                int u = (p - MLP_W2) / MLP_H, v = (p - MLP_W2) % MLP_H;
                for (int s = 0; s < cn; s++) a = a + d2[s][u] * h1[s][v];
            } else if (p < MLP_W3) {
                int u = p - MLP_B2;
                for (int s = 0; s < cn; s++) a = a + d2[s][u];
            } else if (p < MLP_B3) {
                int u = p - MLP_W3;
                for (int s = 0; s < cn; s++) a = a + gs[s] * h2[s][u];
            } else {
                for (int s = 0; s < cn; s++) a = a + gs[s];
            }
            acc[q] = acc[q] + a;
        }
        __syncthreads();
    }
    for (int q = 0; q < PPT; q++) {
        int p = tid + WG * q;
        if (p < MLP_NPARAM) partial[(long)blockIdx.x * MLP_NPARAM + p] = acc[q];
    }
    lred[tid] = lacc;
    __syncthreads();
    if (tid == 0) {
        float l = 0.0f;
        for (int i = 0; i < WG; i++) l = l + lred[i];
        lossp[blockIdx.x] = l;
    }
}

// grad = sum of the block partials in block order; loss likewise
__global__ __launch_bounds__(WG) void k_mlp_sum(const float *partial, int nblk, float *grad, const float *lossp,
                                                float *loss)
{
    int p = blockIdx.x * WG + threadIdx.x;
    if (p < MLP_NPARAM) {
        float a = 0.0f;
        for (int b = 0; b < nblk; b++) a = a + partial[(long)b * MLP_NPARAM + p];
        grad[p] = a;
    }
    if (p == 0) {
        float l = 0.0f;
        for (int b = 0; b < nblk; b++) l = l + lossp[b];
        *loss = l;
    }
}

// Adam (Kingma & Ba) with bias correction
__global__ __launch_bounds__(WG) void k_mlp_adam(float *w, const float *g, float *m, float *v, float lr, float b1c,
                                                 float b2c)
{
    int p = blockIdx.x * WG + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f;
    float gi = g[p];
    float mi = b1 * m[p] + (1.0f - b1) * gi;
    float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    float mh = mi / b1c, vh = vi / b2c;
    w[p] = w[p] - lr * mh / (sqrtf(vh) + 1e-8f);
}
