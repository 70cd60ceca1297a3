// fl_framelet_dev.h -- per-pixel bodies of the 2D framelet operators and the
// filter-tap table, shared by the single-problem and batched kernels.
#pragma once

#include "fl_common.h"

struct FlTaps {
    float D[5][5];   // decomposition filters, NT taps each
    float R[5][5];   // reconstruction filters
};

// Host: the tap table of a filter kind (FL_FRAME_HAAR / LINEAR / CUBIC).
FlTaps fl_make_taps(int kind);

__device__ __forceinline__ int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

__device__ __forceinline__ float shrinkf(float t, float tau)
{
    float m = fabsf(t) - tau;
    return m > 0.0f ? copysignf(m, t) : 0.0f;
}

// The NT x NT neighbourhood of (x, y) at dilation `step`, circular.
template <int NT, typename XT>
__device__ __forceinline__ void fl_gather2(const XT* __restrict__ u, int nx, int ny, int x, int y, int step,
                                           float (&nb)[NT][NT])
{
    constexpr int CC = (NT - 1) / 2;
#pragma unroll
    for (int ky = 0; ky < NT; ky++) {
        const int yy = wrapi(y + (ky - CC) * step, ny);
#pragma unroll
        for (int kx = 0; kx < NT; kx++) nb[ky][kx] = fl_load(u, (size_t)yy * nx + wrapi(x + (kx - CC) * step, nx));
    }
}

// All NF^2 analysis bands of one pixel: band(i,j) = sum D_j[ky] D_i[kx] nb[ky][kx]
// (FraDecMultiLevel2D.m:43-51: corr_y[D_j] after corr_x[D_i]).
template <int NF, int NT>
__device__ __forceinline__ void fl_bands2(const FlTaps& T, const float (&nb)[NT][NT], float (&out)[NF * NF])
{
    float tx[NF][NT];
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int ky = 0; ky < NT; ky++) {
            float s = 0.0f;
#pragma unroll
            for (int kx = 0; kx < NT; kx++) s = __builtin_fmaf(T.D[i][kx], nb[ky][kx], s);
            tx[i][ky] = s;
        }
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int j = 0; j < NF; j++) {
            float s = 0.0f;
#pragma unroll
            for (int ky = 0; ky < NT; ky++) s = __builtin_fmaf(T.D[j][ky], tx[i][ky], s);
            out[i + NF * j] = s;
        }
}

// Synthesis at one pixel: sum_i corr_x[R_i] sum_j corr_y[R_j] C_ij
// (FraRecMultiLevel2D.m:41-49); band b of C starts at C + b*band_stride,
// band 0 at lp instead where lp is given.
template <int NF, int NT, typename CT>
__device__ __forceinline__ float fl_synth2(const CT* __restrict__ C, size_t band_stride, int nx, int ny, int x, int y,
                                           int step, const FlTaps& T, const CT* lp = nullptr)
{
    constexpr int CC = (NT - 1) / 2;
    int xs[NT], ys[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        xs[k] = wrapi(x + (k - CC) * step, nx);
        ys[k] = wrapi(y + (k - CC) * step, ny);
    }
    float acc = 0.0f;
    for (int i = 0; i < NF; i++)
        for (int j = 0; j < NF; j++) {
            const CT* __restrict__ cb = (lp && i + NF * j == 0) ? lp : C + (size_t)(i + NF * j) * band_stride;
            float sx = 0.0f;
#pragma unroll
            for (int kx = 0; kx < NT; kx++) {
                float sy = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++) sy = __builtin_fmaf(T.R[j][ky], fl_load(cb, (size_t)ys[ky] * nx + xs[kx]), sy);
                sx = __builtin_fmaf(T.R[i][kx], sy, sx);
            }
            acc += sx;
        }
    return acc;
}

// Split-Bregman coefficient update at one pixel (solveSplitBregman.m:28-36):
// t = v + Wu, alpha = shrink(t) (low-pass copied), v <- v + (Wu - alpha),
// d = alpha - v.  Returns the pixel's share of |Wu - alpha|^2.  lp (may be
// null) receives the low-pass band of Wu, the input of the next level.
template <int NF, int NT, typename CT>
__device__ __forceinline__ double fl_bregman2_pixel(const float* __restrict__ u, CT* __restrict__ V,
                                                    CT* __restrict__ Dc, size_t band_stride, size_t idx, int nx,
                                                    int ny, int x, int y, int step, const FlTaps& T, float tau,
                                                    float* __restrict__ lp = nullptr)
{
    float nb[NT][NT];
    fl_gather2<NT>(u, nx, ny, x, y, step, nb);
    float w[NF * NF];
    fl_bands2<NF, NT>(T, nb, w);
    if (lp) lp[idx] = w[0];
    double r2 = 0.0;
#pragma unroll
    for (int b = 0; b < NF * NF; b++) {
        const size_t o = (size_t)b * band_stride + idx;
        const float v = fl_load(V, o);
        const float t = v + w[b];
        const float a = (b == 0) ? t : shrinkf(t, tau);
        const float e = w[b] - a;
        const float vn = v + e;
        r2 += (double)e * (double)e;
        fl_store(V, o, vn);
        fl_store(Dc, o, a - vn);
    }
    return r2;
}
