// framelet.hip -- undecimated B-spline tight-framelet transforms (2D/3D) and
// the fused split-Bregman / smoothed-l1 coefficient updates.
//
// Reference semantics (matlab/classes/+Transforms/@FrameletSystem):
//   filters  GenerateFrameletFilter.m:1-30 (haar 2 filters x 3 taps, linear
//            3 x 3, cubic 5 x 5; all boundaries circular);
//   1D pass  utils/ConvSymAsym3D.m:19-24: circular correlation with the taps
//            dilated by 2^(L-1), kernel centre floor((len+1)/2);
//   2D       FraDecMultiLevel2D.m:43-51: band(i,j) = corr_y[D_j] corr_x[D_i] u;
//            FraRecMultiLevel2D.m:41-49: u = sum_i corr_x[R_i] sum_j corr_y[R_j] C_ij;
//   3D       FraDecMultiLevel3D.m:35-47: band(i,j,k) = corr_z[D_k] corr_x[D_j] corr_y[D_i] u;
//            FraRecMultiLevel3D.m:35-48: u = sum_i corr_y[R_i] sum_j corr_x[R_j]
//            sum_k corr_z[R_k] C_ijk.
// Band b = i + NF*j (+ NF*NF*k) (MATLAB linear order of the cell {i,j,k});
// coefficients are stored band-major: C[b][z][y][x].
//
// One thread per voxel; analysis gathers the NT^d neighbourhood once and
// produces all NF^d bands from it (separably, in registers), so the input is
// read once and every band is written once (the transform is write-bound).
// 3D synthesis runs separably on LDS tiles (k_iwt3m, marching through z,
// for dilations 1 and 2, k_iwt3s beyond); 2D synthesis, and 3D levels whose
// halo outgrows the tiles, gather each band's neighbourhood per output.  All accumulate in
// FP32.
#include <algorithm>
#include <type_traits>

#include "fl_common.h"
#include "fl_framelet.h"
#include "fl_framelet_dev.h"

// The z-marching synthesis' target of resident waves and its use of
// full-row tiles (build-time tuning knobs).
#ifndef FL_IWT_MIN_WAVES
#define FL_IWT_MIN_WAVES 2048
#endif
#ifndef FL_IWT_FULL_ROWS
#define FL_IWT_FULL_ROWS 1
#endif

// ---------------------------------------------------------------- 3D ------

// Coordinate v in [-n, 2n) wrapped into [0, n) without a divide.
__device__ __forceinline__ int wrap_near(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Band b of a coefficient array with `nvox` elements per band; lp (may be
// null) stands in for band 0, the low-pass band (multi-level synthesis puts
// the reconstruction of the level above there, FraRecMultiLevel3D.m:14-17).
template <typename CT>
__device__ __forceinline__ const CT* band_ptr(const CT* C, int b, size_t nvox, const CT* lp)
{
    return (b == 0 && lp) ? lp : C + (size_t)b * nvox;
}

// NEAR: the taps reach at most one period beyond the volume (C*step <= n
// on every axis), so wrap_near serves.
template <int NF, int NT, bool NEAR = false, typename XT>
__device__ __forceinline__ void gather3(const XT* __restrict__ u, int nx, int ny, int nz, int x, int y,
                                        int z, int step, float (&nb)[NT][NT][NT])
{
    constexpr int C = (NT - 1) / 2;
    int xs[NT], ys[NT], zs[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        xs[k] = NEAR ? wrap_near(x + (k - C) * step, nx) : wrapi(x + (k - C) * step, nx);
        ys[k] = NEAR ? wrap_near(y + (k - C) * step, ny) : wrapi(y + (k - C) * step, ny);
        zs[k] = NEAR ? wrap_near(z + (k - C) * step, nz) : wrapi(z + (k - C) * step, nz);
    }
#pragma unroll
    for (int kz = 0; kz < NT; kz++)
#pragma unroll
        for (int ky = 0; ky < NT; ky++)
#pragma unroll
            for (int kx = 0; kx < NT; kx++)
                nb[kz][ky][kx] = fl_load(u, ((size_t)zs[kz] * ny + ys[ky]) * nx + xs[kx]);
}

// All NF^3 analysis bands of one voxel from its neighbourhood:
// band(i,j,k) = sum D_k[kz] D_j[kx] D_i[ky] nb[kz][ky][kx].
template <int NF, int NT>
__device__ __forceinline__ void bands3(const FlTaps& T, const float (&nb)[NT][NT][NT], float (&out)[NF * NF * NF])
{
    float ty[NF][NT][NT];   // corr_y with D_i: [i][kz][kx]
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int kz = 0; kz < NT; kz++)
#pragma unroll
            for (int kx = 0; kx < NT; kx++) {
                float s = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++) s = __builtin_fmaf(T.D[i][ky], nb[kz][ky][kx], s);
                ty[i][kz][kx] = s;
            }
    float tx[NF][NF][NT];   // then corr_x with D_j: [i][j][kz]
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int j = 0; j < NF; j++)
#pragma unroll
            for (int kz = 0; kz < NT; kz++) {
                float s = 0.0f;
#pragma unroll
                for (int kx = 0; kx < NT; kx++) s = __builtin_fmaf(T.D[j][kx], ty[i][kz][kx], s);
                tx[i][j][kz] = s;
            }
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int j = 0; j < NF; j++)
#pragma unroll
            for (int k = 0; k < NF; k++) {
                float s = 0.0f;
#pragma unroll
                for (int kz = 0; kz < NT; kz++) s = __builtin_fmaf(T.D[k][kz], tx[i][j][kz], s);
                out[i + NF * j + NF * NF * k] = s;
            }
}

// Workgroups of 64 x 4 (x, y) voxels of one slice: coordinates without
// divides.  Band stores are nontemporal (0.35 ms instead of 0.44 ms for 27
// FP32 bands of 256^3).
template <int NF, int NT, bool NEAR, typename XT, typename CT>
__global__ __launch_bounds__(256) void k_fwt3(const XT* __restrict__ u, CT* __restrict__ C, int nx, int ny,
                                             int nz, int z0, int step, FlTaps T)
{
    const size_t nvox = (size_t)nx * ny * nz;
    const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int y = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int z = z0 + (int)blockIdx.z;
    if (x >= nx || y >= ny) return;
    const size_t idx = ((size_t)z * ny + y) * nx + x;
    float nb[NT][NT][NT];
    gather3<NF, NT, NEAR>(u, nx, ny, nz, x, y, z, step, nb);
    float out[NF * NF * NF];
    bands3<NF, NT>(T, nb, out);
#pragma unroll
    for (int b = 0; b < NF * NF * NF; b++) fl_store_nt(C, (size_t)b * nvox + idx, out[b]);
}

// u = alpha * sum_b corr(R-stencil of b, C_b) + beta * base (base may be null)
// `base` may alias `u` (out = alpha*W^T C + beta*out): each thread reads and
// writes only its own element.
template <int NF, int NT, typename CT, typename XT>
__global__ __launch_bounds__(256) void k_iwt3(const CT* __restrict__ C, XT* u, int nx, int ny,
                                             int nz, int step, FlTaps T, float alpha, float beta,
                                             const float* base, const CT* lp)
{
    constexpr int CC = (NT - 1) / 2;
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nvox) return;
    const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
    size_t off[NT][NT][NT];
    {
        int xs[NT], ys[NT], zs[NT];
#pragma unroll
        for (int k = 0; k < NT; k++) {
            xs[k] = wrapi(x + (k - CC) * step, nx);
            ys[k] = wrapi(y + (k - CC) * step, ny);
            zs[k] = wrapi(z + (k - CC) * step, nz);
        }
#pragma unroll
        for (int kz = 0; kz < NT; kz++)
#pragma unroll
            for (int ky = 0; ky < NT; ky++)
#pragma unroll
                for (int kx = 0; kx < NT; kx++) off[kz][ky][kx] = ((size_t)zs[kz] * ny + ys[ky]) * nx + xs[kx];
    }
    float acc = 0.0f;
    for (int i = 0; i < NF; i++)
        for (int j = 0; j < NF; j++)
            for (int k = 0; k < NF; k++) {
                const CT* __restrict__ cb = band_ptr(C, i + NF * j + NF * NF * k, nvox, lp);
                // innermost corr_z[R_k], then corr_x[R_j], then corr_y[R_i]
                float sy = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++) {
                    float sx = 0.0f;
#pragma unroll
                    for (int kx = 0; kx < NT; kx++) {
                        float sz = 0.0f;
#pragma unroll
                        for (int kz = 0; kz < NT; kz++) sz = __builtin_fmaf(T.R[k][kz], fl_load(cb, off[kz][ky][kx]), sz);
                        sx = __builtin_fmaf(T.R[j][kx], sz, sx);
                    }
                    sy = __builtin_fmaf(T.R[i][ky], sx, sy);
                }
                acc += sy;
            }
    float v = alpha * acc;
    if (base) v = __builtin_fmaf(beta, base[idx], v);
    fl_store(u, idx, v);
}

// Separable 3D synthesis on LDS tiles.  The band sum factorises,
//   u = sum_i corr_y[R_i]( sum_j corr_x[R_j]( sum_k corr_z[R_k] C_ijk ) ),
// so a workgroup owning an IW_TX x IW_TY tile of (x, y) columns for `zc`
// consecutive slices forms, per slice and per y filter i: the z pass of the
// bands (i, j, *) for every j on the tile and its halo (in LDS), their x
// pass on the tile's and halo's rows (in LDS), and adds the y pass to its
// outputs.  Each coefficient is read NT times per halo position (its z
// taps, mostly from L2) instead of NT^3 times.  FP32 FMAs in another
// association than k_iwt3's: equal to it up to rounding.  Same in-place
// contract (`base` may alias u).
constexpr int IW_TX = 64, IW_TY = 16;
constexpr int IW_OUT = IW_TX * IW_TY / 256;   // outputs per thread

__host__ __device__ inline int iw_halo_positions(int h) { return (IW_TX + 2 * h) * (IW_TY + 2 * h); }
// LDS: z pass [NF][PH][PW], x pass [PH][IW_TX], halo offsets [PH][PW] (int)
template <int NF>
__host__ __device__ inline size_t iw_lds_bytes(int h)
{
    return ((size_t)(NF + 1) * iw_halo_positions(h) + (size_t)(IW_TY + 2 * h) * IW_TX) * sizeof(float);
}

template <int NF, int NT, typename CT, typename XT>
__global__ __launch_bounds__(256) void k_iwt3s(const CT* __restrict__ C, XT* u, int nx, int ny, int nz, int step,
                                              int zc, FlTaps T, float alpha, float beta, const float* base,
                                              const CT* lp)
{
    extern __shared__ float iw_lds[];
    constexpr int CC = (NT - 1) / 2;
    const int h = CC * step;
    const int PW = IW_TX + 2 * h, PH = IW_TY + 2 * h, P = PW * PH;
    float* tz = iw_lds;                                    // z pass of bands (i, j, *), per j
    float* sx = iw_lds + NF * P;                           // x pass
    int* poff = reinterpret_cast<int*>(sx + PH * IW_TX);   // halo position -> wrapped offset in a slice
    const size_t nxy = (size_t)nx * ny, nvox = nxy * nz;
    const int tilesx = (nx + IW_TX - 1) / IW_TX;
    const int x0 = (int)(blockIdx.x % tilesx) * IW_TX, y0 = (int)(blockIdx.x / tilesx) * IW_TY;
    const int tid = threadIdx.x;
    for (int p = tid; p < P; p += 256) poff[p] = wrapi(y0 - h + p / PW, ny) * nx + wrapi(x0 - h + p % PW, nx);
    __syncthreads();
    for (int z = blockIdx.y * zc; z < min(nz, (int)(blockIdx.y + 1) * zc); z++) {
        size_t zoff[NT];
#pragma unroll
        for (int kz = 0; kz < NT; kz++) zoff[kz] = (size_t)wrapi(z + (kz - CC) * step, nz) * nxy;
        float acc[IW_OUT];
#pragma unroll
        for (int r = 0; r < IW_OUT; r++) acc[r] = 0.0f;
#pragma unroll 1
        for (int i = 0; i < NF; i++) {
            for (int p = tid; p < P; p += 256) {
                const int off = poff[p];
#pragma unroll
                for (int j = 0; j < NF; j++) {
                    float t = 0.0f;
#pragma unroll
                    for (int k = 0; k < NF; k++) {
                        const CT* __restrict__ cb = band_ptr(C, i + NF * j + NF * NF * k, nvox, lp) + off;
#pragma unroll
                        for (int kz = 0; kz < NT; kz++) t = __builtin_fmaf(T.R[k][kz], fl_load(cb, zoff[kz]), t);
                    }
                    tz[j * P + p] = t;
                }
            }
            __syncthreads();
            for (int q = tid; q < PH * IW_TX; q += 256) {
                const float* row = tz + (q / IW_TX) * PW + q % IW_TX + h;
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < NF; j++)
#pragma unroll
                    for (int kx = 0; kx < NT; kx++) s = __builtin_fmaf(T.R[j][kx], row[j * P + (kx - CC) * step], s);
                sx[q] = s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < IW_OUT; r++) {
                const int o = tid + 256 * r, ox = o % IW_TX, oy = o / IW_TX;
                float s = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++)
                    s = __builtin_fmaf(T.R[i][ky], sx[(oy + h + (ky - CC) * step) * IW_TX + ox], s);
                acc[r] += s;
            }
        }
#pragma unroll
        for (int r = 0; r < IW_OUT; r++) {
            const int o = tid + 256 * r, x = x0 + o % IW_TX, y = y0 + o / IW_TX;
            if (x < nx && y < ny) {
                const size_t idx = (size_t)z * nxy + (size_t)y * nx + x;
                float v = alpha * acc[r];
                if (base) v = __builtin_fmaf(beta, base[idx], v);
                fl_store(u, idx, v);
            }
        }
    }
}

// The separable synthesis marching through z.  A workgroup owns a 64 x 16
// tile of (x, y) columns for `zs` consecutive output slices and walks its
// input slices z0 - H .. z0 + zs - 1 + H once, in order.  Per input slice
// and z filter k it forms the 2-D synthesis of the bands (*, *, k),
//   S_k = sum_i corr_y[R_i]( sum_j corr_x[R_j] C_ijk ),
// one y filter i at a time (the NF bands (i, *, k) staged in LDS, their x
// pass summed over j in LDS, the y pass into registers), and adds
// R_k[kz] * S_k to the 2H + 1 output slices that tap this input slice,
// held in registers; the oldest is then complete and is stored.  Each
// coefficient is read (zs + 2H) / zs times per halo position, and the
// loads of the next band triple are in flight while the current one is
// filtered.  Same in-place contract as k_iwt3 (`base` may alias u).
template <int NT, int STEP, int TXW>
struct ImGeom {
    static constexpr int CC = (NT - 1) / 2, H = CC * STEP;
    static constexpr int TX = TXW, ROWS = 4, TY = 4 * ROWS, NTH = 4 * TXW;   // four row groups of ROWS rows
    static constexpr int PW = TX + 2 * H, PH = TY + 2 * H, P = PW * PH;
    static constexpr int NPP = (P + NTH - 1) / NTH;   // halo positions per thread
    // LDS row pitch: a multiple of 4 floats and at least TX + 4, so the x
    // pass reads each row span as whole 16-byte words, conflict-free
    static constexpr int PWP = ((PW + 3) / 4) * 4 < TX + 4 ? TX + 4 : ((PW + 3) / 4) * 4;
    static constexpr int NV = (4 + 2 * H + 3) / 4;   // 16-byte words per x-pass span
    template <int NF>
    static constexpr size_t lds() { return 2 * ((size_t)NF * PH * PWP + (size_t)PH * TX) * sizeof(float); }
};

template <int NF, int NT, int STEP, int TXW, typename CT, typename XT>
__global__ __launch_bounds__(4 * TXW) void k_iwt3m(const CT* __restrict__ C, XT* u, int nx, int ny, int nz, int zs,
                                                  FlTaps T, float alpha, float beta, const float* base,
                                                  const CT* lp)
{
    using G = ImGeom<NT, STEP, TXW>;
    constexpr int CC = G::CC, H = G::H, PW = G::PW, PH = G::PH, P = G::P, NPP = G::NPP, TX = G::TX, TY = G::TY,
                  ROWS = G::ROWS, NTH = G::NTH, NR = 2 * H + 1, NS = NF * NF, PWP = G::PWP, NV = G::NV;
    static_assert(PWP % 4 == 0 && PWP >= 4 * NV + TX - 4, "x-pass spans must stay in their LDS row");
    // double-buffered: the bands (i, j, k), j = 0 .. NF-1, of one input
    // slice on the tile and its halo (rows of pitch PWP), and their x pass
    // summed over j
    __shared__ __align__(16) float sb[2][NF][PH * PWP];
    __shared__ __align__(16) float sx[2][PH][TX];
    const size_t nxy = (size_t)nx * ny, nvox = nxy * nz;
    const int tilesx = (nx + TX - 1) / TX;
    const uint32_t ntile = (uint32_t)tilesx * ((ny + TY - 1) / TY);
    // consecutive tiles (x fastest, then y, then z) on one XCD
    const uint32_t lin = fl_xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t tile = lin % ntile;
    const int x0 = (int)(tile % tilesx) * TX, y0 = (int)(tile / tilesx) * TY;
    const int z0 = (int)(lin / ntile) * zs, z1 = min(z0 + zs, nz);
    const int tid = threadIdx.x, ox = tid % TX, oy = (tid / TX) * ROWS;
    // this thread's halo positions -> wrapped offsets in a slice (-1: none)
    // and places in an LDS band
    int off[NPP], lix[NPP];
#pragma unroll
    for (int q = 0; q < NPP; q++) {
        const int p = tid + NTH * q;
        off[q] = p < P ? wrapi(y0 - H + p / PW, ny) * nx + wrapi(x0 - H + p % PW, nx) : -1;
        lix[q] = (p / PW) * PWP + p % PW;
    }
    // Steps s = 0 .. nsteps-1: input slice z0 - H + s / NS, z filter
    // k = (s / NF) % NF, y filter i = s % NF.  A step stages its bands
    // (phase A), forms their x pass (B) and adds their y pass to S_k (C);
    // iteration t runs A of step t, B of step t - 1 and C of step t - 2
    // between two barriers, on alternating buffers.  The raw coefficients
    // of step t + 2 load into one of two register sets during iteration t.
    const int nsteps = (z1 - z0 + 2 * H) * NS;
    CT preA[NF][NPP], preB[NF][NPP];
    auto fetch = [&](int s, CT (&pre)[NF][NPP]) {
        const int sl = s / NS, k = (s / NF) % NF, i = s % NF;
        const size_t zo = (size_t)wrapi(z0 - H + sl, nz) * nxy;
#pragma unroll
        for (int j = 0; j < NF; j++) {
            const CT* __restrict__ cz = band_ptr(C, i + NF * j + NF * NF * k, nvox, lp) + zo;
#pragma unroll
            for (int q = 0; q < NPP; q++) pre[j][q] = off[q] >= 0 ? cz[off[q]] : CT{};
        }
    };
    float ring[NR][ROWS];   // ring[s]: output slice zin - H + s
    float S[ROWS];
#pragma unroll
    for (int q = 0; q < NR; q++)
#pragma unroll
        for (int r = 0; r < ROWS; r++) ring[q][r] = 0.0f;
    // t: iteration; W: the buffer A of step t writes (t % 2)
    auto iter = [&](int t, auto wbuf, CT (&pre)[NF][NPP]) {
        constexpr int W = decltype(wbuf)::value, Rb = 1 - W;
        if (t < nsteps) {   // A(t)
#pragma unroll
            for (int j = 0; j < NF; j++)
#pragma unroll
                for (int q = 0; q < NPP; q++)
                    if (tid + NTH * q < P) sb[W][j][lix[q]] = fl_load(&pre[j][q], 0);
            if (t + 2 < nsteps) fetch(t + 2, pre);
        }
        if (t >= 1 && t <= nsteps) {   // B(t - 1): four adjacent outputs per work item
            for (int q = tid; q < PH * (TX / 4); q += NTH) {
                const int y = q / (TX / 4), x = (q % (TX / 4)) * 4;
                float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int j = 0; j < NF; j++) {
                    float v[4 * NV];
#pragma unroll
                    for (int m = 0; m < NV; m++) {
                        const float4 w = *reinterpret_cast<const float4*>(&sb[Rb][j][y * PWP + x + 4 * m]);
                        v[4 * m] = w.x; v[4 * m + 1] = w.y; v[4 * m + 2] = w.z; v[4 * m + 3] = w.w;
                    }
#pragma unroll
                    for (int o = 0; o < 4; o++)
#pragma unroll
                        for (int kx = 0; kx < NT; kx++) a[o] = __builtin_fmaf(T.R[j][kx], v[o + H + (kx - CC) * STEP], a[o]);
                }
                *reinterpret_cast<float4*>(&sx[Rb][y][x]) = make_float4(a[0], a[1], a[2], a[3]);
            }
        }
        if (t >= 2) {   // C(t - 2)
            const int s = t - 2, sl = s / NS, k = (s / NF) % NF, i = s % NF;
            if (i == 0) {
#pragma unroll
                for (int r = 0; r < ROWS; r++) S[r] = 0.0f;
            }
#pragma unroll
            for (int r = 0; r < ROWS; r++) {
                float a = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++) a = __builtin_fmaf(T.R[i][ky], sx[W][oy + r + H + (ky - CC) * STEP][ox], a);
                S[r] += a;
            }
            if (i == NF - 1) {
                // S_k feeds output slice zin - (kz - CC)*STEP, ring slot H - (kz - CC)*STEP
#pragma unroll
                for (int kz = 0; kz < NT; kz++)
#pragma unroll
                    for (int r = 0; r < ROWS; r++)
                        ring[H - (kz - CC) * STEP][r] =
                            __builtin_fmaf(T.R[k][kz], S[r], ring[H - (kz - CC) * STEP][r]);
                if (k == NF - 1) {
                    // output slice zin - H has all its taps
                    const int z = z0 - 2 * H + sl;
                    if (z >= z0) {
                        const int x = x0 + ox;
#pragma unroll
                        for (int r = 0; r < ROWS; r++) {
                            const int y = y0 + oy + r;
                            if (x < nx && y < ny) {
                                const size_t idx = (size_t)z * nxy + (size_t)y * nx + x;
                                float v = alpha * ring[0][r];
                                if (base) v = __builtin_fmaf(beta, base[idx], v);
                                fl_store(u, idx, v);
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q + 1 < NR; q++)
#pragma unroll
                        for (int r = 0; r < ROWS; r++) ring[q][r] = ring[q + 1][r];
#pragma unroll
                    for (int r = 0; r < ROWS; r++) ring[NR - 1][r] = 0.0f;
                }
            }
        }
        __syncthreads();
    };
    fetch(0, preA);
    if (nsteps > 1) fetch(1, preB);
#pragma unroll 1
    for (int t = 0; t < nsteps + 2; t += 2) {
        iter(t, std::integral_constant<int, 0>{}, preA);
        if (t + 1 < nsteps + 2) iter(t + 1, std::integral_constant<int, 1>{}, preB);
    }
}

// Launch of the marching synthesis: full-row tiles (TX = 256) where the
// rows are 256 wide or a multiple of it and their LDS fits 160 KB (no x
// halo: the taps wrap inside the tile), else 64-wide tiles.  Output slices
// per workgroup: the longest run (fewer halo re-reads) that still gives at
// least one workgroup per CU and `min_waves` waves, at least 4.
template <int NF, int NT, int STEP, typename CT, typename XT>
static void launch_iwt3m(const FlFrameDesc& f, const CT* C, XT* u, const FlTaps& T, float alpha, float beta,
                         const float* base, const CT* lp, hipStream_t st)
{
    auto go = [&](auto txw) {
        constexpr int TXW = decltype(txw)::value;
        constexpr int TY = ImGeom<NT, STEP, TXW>::TY;
        const unsigned tiles = (unsigned)(((f.nx + TXW - 1) / TXW) * ((f.ny + TY - 1) / TY));
        const size_t min_groups = std::max<size_t>(256, (size_t)FL_IWT_MIN_WAVES / (TXW / 16));
        int zs = 64;
        while (zs > 4 && (size_t)tiles * ((f.nz + zs - 1) / zs) < min_groups) zs /= 2;
        zs = std::min(zs, std::max(f.nz, 1));
        const dim3 grid(tiles * (unsigned)((f.nz + zs - 1) / zs));
        hipLaunchKernelGGL((k_iwt3m<NF, NT, STEP, TXW, CT, XT>), grid, dim3(4 * TXW), 0, st, C, u, f.nx, f.ny, f.nz,
                           zs, T, alpha, beta, base, lp);
    };
    if constexpr (FL_IWT_FULL_ROWS && ImGeom<NT, STEP, 256>::template lds<NF>() <= 160 * 1024) {
        if (f.nx % 256 == 0) {
            go(std::integral_constant<int, 256>{});
            return;
        }
    }
    go(std::integral_constant<int, 64>{});
}

// ---------------------------------------------------------------- 2D ------

template <int NF, int NT, typename XT, typename CT>
__global__ __launch_bounds__(256) void k_fwt2(const XT* __restrict__ u, CT* __restrict__ C, int nx, int ny,
                                             int step, FlTaps T)
{
    const size_t npix = (size_t)nx * ny;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= npix) return;
    const int x = (int)(idx % nx), y = (int)(idx / nx);
    float nb[NT][NT];
    fl_gather2<NT>(u, nx, ny, x, y, step, nb);
    float w[NF * NF];
    fl_bands2<NF, NT>(T, nb, w);
#pragma unroll
    for (int b = 0; b < NF * NF; b++) fl_store(C, (size_t)b * npix + idx, w[b]);
}

template <int NF, int NT, typename CT, typename XT>
__global__ __launch_bounds__(256) void k_iwt2(const CT* __restrict__ C, XT* u, int nx, int ny,
                                             int step, FlTaps T, float alpha, float beta,
                                             const float* base, const CT* lp)
{
    const size_t npix = (size_t)nx * ny;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= npix) return;
    const int x = (int)(idx % nx), y = (int)(idx / nx);
    float v = alpha * fl_synth2<NF, NT>(C, npix, nx, ny, x, y, step, T, lp);
    if (base) v = __builtin_fmaf(beta, base[idx], v);
    fl_store(u, idx, v);
}

// ----------------------------------------------- fused coefficient updates --
//
// Split Bregman (solveSplitBregman.m:28-36, shrink softThreshold.m:15-80):
//   Wu = W u;  alpha = shrink(v + Wu, tau) (low-pass band copied);
//   v <- v + (Wu - alpha);  d = alpha - v   (consumed by b = Atf + mu W^T d).
// Analysis, shrink and both updates happen in one pass; Wu and alpha are
// never stored.  resid accumulates per-block FP64 partials of |Wu - alpha|^2
// (the constraint residual used as the outer stopping test).  One launch per
// level (dilation `step`); `lp`, when given, receives the level's low-pass
// band of Wu, the input of the next level (FraDecMultiLevel3D.m:16-19).
//
// Smoothed l1 (gradient of lam * sum psi(Wu), psi(t) = sqrt(t^2 + eps)):
//   g_b = lam * psi'(Wu_b) = lam * Wu_b / sqrt(Wu_b^2 + eps), low-pass band 0
//   unless with_lowpass; loss partials of lam * sum psi.  A level below the
//   top (lp_pass) stores its low-pass band of Wu itself in band 0 instead:
//   the next level's input, later overwritten by that level's synthesis.

template <int NB>
__device__ __forceinline__ void block_reduce_store(double v, double* __restrict__ part)
{
    __shared__ double sred[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sred[w];
        part[blockIdx.x] = s;
    }
}

template <int NF, int NT, typename CT>
__global__ __launch_bounds__(256) void k_bregman3(const float* __restrict__ u, CT* __restrict__ V,
                                                 CT* __restrict__ Dc, int nx, int ny, int nz, int step,
                                                 FlTaps T, float tau, double* __restrict__ part,
                                                 float* __restrict__ lp)
{
    constexpr int NB = NF * NF * NF;
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    double r2 = 0.0;
    if (idx < nvox) {
        const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
        float nb[NT][NT][NT];
        gather3<NF, NT>(u, nx, ny, nz, x, y, z, step, nb);
        float wu[NB];
        bands3<NF, NT>(T, nb, wu);
        if (lp) lp[idx] = wu[0];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const size_t o = (size_t)b * nvox + idx;
            const float v = fl_load(V, o);
            const float t = v + wu[b];
            const float a = (b == 0) ? t : shrinkf(t, tau);
            const float e = wu[b] - a;
            const float vn = v + e;
            r2 += (double)e * (double)e;
            fl_store(V, o, vn);
            fl_store(Dc, o, a - vn);
        }
    }
    block_reduce_store<NB>(r2, part);
}

template <int NF, int NT, typename CT>
__global__ __launch_bounds__(256) void k_bregman2(const float* __restrict__ u, CT* __restrict__ V,
                                                 CT* __restrict__ Dc, int nx, int ny, int step, FlTaps T,
                                                 float tau, double* __restrict__ part, float* __restrict__ lp)
{
    const size_t npix = (size_t)nx * ny;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    double r2 = 0.0;
    if (idx < npix) {
        const int x = (int)(idx % nx), y = (int)(idx / nx);
        r2 = fl_bregman2_pixel<NF, NT>(u, V, Dc, npix, idx, nx, ny, x, y, step, T, tau, lp);
    }
    block_reduce_store<NF * NF>(r2, part);
}

template <int NF, int NT, typename CT>
__global__ __launch_bounds__(256) void k_l1grad3(const float* __restrict__ u, CT* __restrict__ G, int nx, int ny,
                                                int nz, int step, FlTaps T, float lam, float eps, int with_lp,
                                                int lp_pass, double* __restrict__ part)
{
    constexpr int NB = NF * NF * NF;
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    double loss = 0.0;
    if (idx < nvox) {
        const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
        float nb[NT][NT][NT];
        gather3<NF, NT>(u, nx, ny, nz, x, y, z, step, nb);
        float wu[NB];
        bands3<NF, NT>(T, nb, wu);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            float g = 0.0f;
            if (b == 0 && lp_pass) {
                g = wu[0];
            } else if (b != 0 || with_lp) {
                const float r = sqrtf(__builtin_fmaf(wu[b], wu[b], eps));
                g = lam * (wu[b] / r);
                loss += (double)lam * (double)r;
            }
            fl_store(G, (size_t)b * nvox + idx, g);
        }
    }
    block_reduce_store<NB>(loss, part);
}

template <int NF, int NT, typename CT>
__global__ __launch_bounds__(256) void k_l1grad2(const float* __restrict__ u, CT* __restrict__ G, int nx, int ny,
                                                int step, FlTaps T, float lam, float eps, int with_lp, int lp_pass,
                                                double* __restrict__ part)
{
    constexpr int CC = (NT - 1) / 2;
    const size_t npix = (size_t)nx * ny;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    double loss = 0.0;
    if (idx < npix) {
        const int x = (int)(idx % nx), y = (int)(idx / nx);
        float nb[NT][NT];
#pragma unroll
        for (int ky = 0; ky < NT; ky++) {
            const int yy = wrapi(y + (ky - CC) * step, ny);
#pragma unroll
            for (int kx = 0; kx < NT; kx++) nb[ky][kx] = u[(size_t)yy * nx + wrapi(x + (kx - CC) * step, nx)];
        }
#pragma unroll
        for (int i = 0; i < NF; i++) {
            float tx[NT];
#pragma unroll
            for (int ky = 0; ky < NT; ky++) {
                float s = 0.0f;
#pragma unroll
                for (int kx = 0; kx < NT; kx++) s = __builtin_fmaf(T.D[i][kx], nb[ky][kx], s);
                tx[ky] = s;
            }
#pragma unroll
            for (int j = 0; j < NF; j++) {
                float w = 0.0f;
#pragma unroll
                for (int ky = 0; ky < NT; ky++) w = __builtin_fmaf(T.D[j][ky], tx[ky], w);
                const int b = i + NF * j;
                float g = 0.0f;
                if (b == 0 && lp_pass) {
                    g = w;
                } else if (b != 0 || with_lp) {
                    const float r = sqrtf(__builtin_fmaf(w, w, eps));
                    g = lam * (w / r);
                    loss += (double)lam * (double)r;
                }
                fl_store(G, (size_t)b * npix + idx, g);
            }
        }
    }
    block_reduce_store<NF * NF>(loss, part);
}

// --------------------------------------------------------------- host side --


// The filter taps of GenerateFrameletFilter.m:1-30, D[i] (decomposition)
// and R[i] = flip(D[i]) (reconstruction), NT taps each, in double.
void fl_frame_taps_f64(int kind, double (&Dd)[5][5], double (&Rd)[5][5])
{
    const double s2 = 1.4142135623730951, s6 = 2.449489742783178;
    for (int i = 0; i < 5; i++)
        for (int k = 0; k < 5; k++) Dd[i][k] = Rd[i][k] = 0.0;
    if (kind == FL_FRAME_HAAR) {
        const double D[2][3] = {{0, 0.5, 0.5}, {0, -0.5, 0.5}};
        const double R[2][3] = {{0.5, 0.5, 0}, {0.5, -0.5, 0}};
        for (int i = 0; i < 2; i++)
            for (int k = 0; k < 3; k++) { Dd[i][k] = D[i][k]; Rd[i][k] = R[i][k]; }
    } else if (kind == FL_FRAME_LINEAR) {
        const double D[3][3] = {{0.25, 0.5, 0.25}, {-s2 / 4, 0, s2 / 4}, {-0.25, 0.5, -0.25}};
        const double R[3][3] = {{0.25, 0.5, 0.25}, {s2 / 4, 0, -s2 / 4}, {-0.25, 0.5, -0.25}};
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) { Dd[i][k] = D[i][k]; Rd[i][k] = R[i][k]; }
    } else {
        const double D[5][5] = {{1 / 16., 4 / 16., 6 / 16., 4 / 16., 1 / 16.},
                                {-1 / 8., -2 / 8., 0, 2 / 8., 1 / 8.},
                                {-s6 / 16, 0, 2 * s6 / 16, 0, -s6 / 16},
                                {-1 / 8., 2 / 8., 0, -2 / 8., 1 / 8.},
                                {-1 / 16., 4 / 16., -6 / 16., 4 / 16., -1 / 16.}};
        const double R[5][5] = {{1 / 16., 4 / 16., 6 / 16., 4 / 16., 1 / 16.},
                                {1 / 8., 2 / 8., 0, -2 / 8., -1 / 8.},
                                {-s6 / 16, 0, 2 * s6 / 16, 0, -s6 / 16},
                                {1 / 8., -2 / 8., 0, 2 / 8., -1 / 8.},
                                {-1 / 16., 4 / 16., -6 / 16., 4 / 16., -1 / 16.}};
        for (int i = 0; i < 5; i++)
            for (int k = 0; k < 5; k++) { Dd[i][k] = D[i][k]; Rd[i][k] = R[i][k]; }
    }
}

FlTaps fl_make_taps(int kind)
{
    FlTaps T = {};
    double D[5][5], R[5][5];
    fl_frame_taps_f64(kind, D, R);
    for (int i = 0; i < 5; i++)
        for (int k = 0; k < 5; k++) { T.D[i][k] = (float)D[i][k]; T.R[i][k] = (float)R[i][k]; }
    return T;
}

int fl_frame_nf(int kind) { return kind == FL_FRAME_HAAR ? 2 : kind == FL_FRAME_LINEAR ? 3 : 5; }

static unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

#define FL_FRAME_SWITCH(kind, CALL)                                                   \
    do {                                                                              \
        if (kind == FL_FRAME_HAAR) { constexpr int NF = 2, NT = 3; CALL; }            \
        else if (kind == FL_FRAME_LINEAR) { constexpr int NF = 3, NT = 3; CALL; }     \
        else { constexpr int NF = 5, NT = 5; CALL; }                                  \
    } while (0)

template <typename XT, typename CT>
static hipError_t fwt_t(const FlFrameDesc& f, const void* u, void* C, int level, hipStream_t st)
{
    const FlTaps T = fl_make_taps(f.kind);
    const int step = 1 << (level - 1);
    const size_t n = (size_t)f.nx * f.ny * (f.dim == 3 ? f.nz : 1);
    if (f.dim == 3)
        FL_FRAME_SWITCH(f.kind, {
            if ((f.ny + 3) / 4 > 65535) return hipErrorInvalidValue;
            const bool near = (NT - 1) / 2 * step <= std::min(f.nx, std::min(f.ny, f.nz));
            // slices in launches of at most 65535 (the grid's z limit)
            for (int z0 = 0; z0 < f.nz; z0 += 65535) {
                const dim3 grid((unsigned)((f.nx + 63) / 64), (unsigned)((f.ny + 3) / 4),
                                (unsigned)std::min(65535, f.nz - z0));
                if (near)
                    hipLaunchKernelGGL((k_fwt3<NF, NT, true, XT, CT>), grid, dim3(256), 0, st, (const XT*)u,
                                       (CT*)C, f.nx, f.ny, f.nz, z0, step, T);
                else
                    hipLaunchKernelGGL((k_fwt3<NF, NT, false, XT, CT>), grid, dim3(256), 0, st, (const XT*)u,
                                       (CT*)C, f.nx, f.ny, f.nz, z0, step, T);
            }
        });
    else
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_fwt2<NF, NT, XT, CT>), dim3(nblk(n)), dim3(256), 0, st,
                                                   (const XT*)u, (CT*)C, f.nx, f.ny, step, T));
    return hipGetLastError();
}

template <typename CT, typename XT>
static hipError_t iwt_t(const FlFrameDesc& f, const void* C, void* u, int level, float alpha, float beta,
                        const float* base, const void* lpv, hipStream_t st)
{
    const FlTaps T = fl_make_taps(f.kind);
    const int step = 1 << (level - 1);
    const size_t n = (size_t)f.nx * f.ny * (f.dim == 3 ? f.nz : 1);
    const CT* lp = (const CT*)lpv;
    if (f.dim == 3)
        FL_FRAME_SWITCH(f.kind, {
            // dilations 1 and 2: the z-marching tiles; then the separable
            // tiles while their halo tile fits 48 KB; the per-voxel gather
            // beyond
            if (step == 1) {
                (launch_iwt3m<NF, NT, 1, CT, XT>)(f, (const CT*)C, (XT*)u, T, alpha, beta, base, lp, st);
                return hipGetLastError();
            }
            if (step == 2) {
                (launch_iwt3m<NF, NT, 2, CT, XT>)(f, (const CT*)C, (XT*)u, T, alpha, beta, base, lp, st);
                return hipGetLastError();
            }
            const int h = ((NT - 1) / 2) * step;
            const size_t lds = iw_lds_bytes<NF>(h);
            if (lds <= 48 * 1024) {
                const unsigned tiles = (unsigned)(((f.nx + IW_TX - 1) / IW_TX) * ((f.ny + IW_TY - 1) / IW_TY));
                // a few slices per workgroup (their z taps overlap in L2),
                // while that leaves >= 2048 workgroups
                int zc = 1;
                while (zc < 8 && (size_t)tiles * ((f.nz + 2 * zc - 1) / (2 * zc)) >= 2048) zc *= 2;
                hipLaunchKernelGGL((k_iwt3s<NF, NT, CT, XT>), dim3(tiles, (f.nz + zc - 1) / zc), dim3(256), lds, st,
                                   (const CT*)C, (XT*)u, f.nx, f.ny, f.nz, step, zc, T, alpha, beta, base, lp);
            } else {
                hipLaunchKernelGGL((k_iwt3<NF, NT, CT, XT>), dim3(nblk(n)), dim3(256), 0, st, (const CT*)C, (XT*)u,
                                   f.nx, f.ny, f.nz, step, T, alpha, beta, base, lp);
            }
        });
    else
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_iwt2<NF, NT, CT, XT>), dim3(nblk(n)), dim3(256), 0, st,
                                                   (const CT*)C, (XT*)u, f.nx, f.ny, step, T, alpha, beta,
                                                   base, lp));
    return hipGetLastError();
}

hipError_t fl_launch_fwt(const FlFrameDesc& f, const void* u, int ut, void* C, int ct, int level, hipStream_t st)
{
    if (ut == FL_F32 && ct == FL_F32) return fwt_t<float, float>(f, u, C, level, st);
    if (ut == FL_F32 && ct == FL_BF16) return fwt_t<float, bf16_t>(f, u, C, level, st);
    if (ut == FL_F32 && ct == FL_F16) return fwt_t<float, f16_t>(f, u, C, level, st);
    if (ut == FL_BF16 && ct == FL_BF16) return fwt_t<bf16_t, bf16_t>(f, u, C, level, st);
    if (ut == FL_F16 && ct == FL_F16) return fwt_t<f16_t, f16_t>(f, u, C, level, st);
    return hipErrorInvalidValue;
}

hipError_t fl_launch_iwt(const FlFrameDesc& f, const void* C, int ct, void* u, int ut, int level, float alpha,
                         float beta, const float* base, hipStream_t st, const void* lp)
{
    if (ct == FL_F32 && ut == FL_F32) return iwt_t<float, float>(f, C, u, level, alpha, beta, base, lp, st);
    if (ct == FL_BF16 && ut == FL_F32) return iwt_t<bf16_t, float>(f, C, u, level, alpha, beta, base, lp, st);
    if (ct == FL_F16 && ut == FL_F32) return iwt_t<f16_t, float>(f, C, u, level, alpha, beta, base, lp, st);
    if (ct == FL_BF16 && ut == FL_BF16) return iwt_t<bf16_t, bf16_t>(f, C, u, level, alpha, beta, base, lp, st);
    if (ct == FL_F16 && ut == FL_F16) return iwt_t<f16_t, f16_t>(f, C, u, level, alpha, beta, base, lp, st);
    return hipErrorInvalidValue;
}

template <typename CT>
static hipError_t bregman_t(const FlFrameDesc& f, const float* u, void* V, void* D, float tau, int level,
                            float* lp, double* part, hipStream_t st)
{
    const FlTaps T = fl_make_taps(f.kind);
    const size_t n = (size_t)f.nx * f.ny * (f.dim == 3 ? f.nz : 1);
    const int step = 1 << (level - 1);
    if (f.dim == 3)
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_bregman3<NF, NT, CT>), dim3(nblk(n)), dim3(256), 0, st, u,
                                                   (CT*)V, (CT*)D, f.nx, f.ny, f.nz, step, T, tau, part, lp));
    else
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_bregman2<NF, NT, CT>), dim3(nblk(n)), dim3(256), 0, st, u,
                                                   (CT*)V, (CT*)D, f.nx, f.ny, step, T, tau, part, lp));
    return hipGetLastError();
}

hipError_t fl_launch_bregman(const FlFrameDesc& f, const float* u, void* V, void* D, int ct, float tau, int level,
                             float* lp, double* part, hipStream_t st)
{
    if (level < 1 || level > 16) return hipErrorInvalidValue;
    if (ct == FL_F32) return bregman_t<float>(f, u, V, D, tau, level, lp, part, st);
    if (ct == FL_BF16) return bregman_t<bf16_t>(f, u, V, D, tau, level, lp, part, st);
    if (ct == FL_F16) return bregman_t<f16_t>(f, u, V, D, tau, level, lp, part, st);
    return hipErrorInvalidValue;
}

hipError_t fl_launch_l1grad(const FlFrameDesc& f, const float* u, float* G, float lam, float eps, int with_lp,
                            int level, int lp_pass, double* part, hipStream_t st)
{
    if (level < 1 || level > 16) return hipErrorInvalidValue;
    const FlTaps T = fl_make_taps(f.kind);
    const size_t n = (size_t)f.nx * f.ny * (f.dim == 3 ? f.nz : 1);
    const int step = 1 << (level - 1);
    if (f.dim == 3)
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_l1grad3<NF, NT, float>), dim3(nblk(n)), dim3(256), 0, st, u,
                                                   G, f.nx, f.ny, f.nz, step, T, lam, eps, with_lp, lp_pass, part));
    else
        FL_FRAME_SWITCH(f.kind, hipLaunchKernelGGL((k_l1grad2<NF, NT, float>), dim3(nblk(n)), dim3(256), 0, st, u,
                                                   G, f.nx, f.ny, step, T, lam, eps, with_lp, lp_pass, part));
    return hipGetLastError();
}

// copy of one band (the low-pass of a level) into a volume buffer
__global__ void k_copy_f32(const float* __restrict__ a, float* __restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

hipError_t fl_launch_copy(const float* a, float* b, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_f32, dim3(nblk(n)), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

unsigned fl_frame_blocks(const FlFrameDesc& f)
{
    return nblk((size_t)f.nx * f.ny * (f.dim == 3 ? f.nz : 1));
}
