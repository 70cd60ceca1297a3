// adjoint.hip -- voxel-driven exact adjoints (backprojectors), atomic-free.
//
// Each thread owns one output pixel/voxel and gathers, view by view, the
// detector cells whose rays cross it, weighting each by the exact chord
// length.  No two threads write the same output, so there are no atomics
// and the result is deterministic.  The per-voxel summation order (views
// ascending, then detector rows, then columns) and the float arithmetic are
// those of the reference, so results are bit-identical to it:
//   fan:  matlab/mex/src/Atx_fan_mf_cpu_new.cpp:35-62
//   cone: matlab/mex/src/Atx_cone_mf_gpu_new.cu:41-84 (window from zr with
//         d=sqrt((1+dz^2)/2), ceil upper bounds, box sides (1,1,dz)).
#include <type_traits>

#include "fl_chord.h"
#include "fl_common.h"
#include "fl_geom.h"
#include "fl_fan.h"

// Workgroups per CU the register budget is sized for (a build-time tuning
// knob; the default is the measured optimum at 256^3).
#ifndef FL_EXACT_WG_PER_CU
#define FL_EXACT_WG_PER_CU 5
#endif

template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_fan(const YT* __restrict__ Y, XT* __restrict__ X, AdjGeom g)
{
    const int ix = blockIdx.x * 16 + threadIdx.x;
    const int iy = blockIdx.y * 16 + threadIdx.y;
    const int it = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    fl_store(X, (size_t)it * g.nvox + (size_t)iy * g.nx + ix, fl_fan_adj_pixel(Y, g, it, ix, iy));
}

// Cone, footprint-staged: per view the workgroup stages its detector
// footprint in LDS -- the rays' direction components, the six direction
// ratios find_l_3d divides out (each the IEEE quotient, computed once per
// workgroup instead of once per voxel), and the sinogram values.  Each voxel
// then evaluates, with the reference arithmetic and in the reference order,
// the chord of every cell in the intersection of the reference's window and
// the voxel's shadow on the detector (widened by 0.05 cells), or of the
// shadow alone where the plan has shown that the reference's window contains
// it (AdjGeom::win_ref clear).  Cells outside the shadow have exact-zero
// chords, and the reference adds l*y = +-0 for them, which leaves a sum that
// starts at +0 unchanged, so the result is bit-identical to the reference
// kernel.  (Per-cell miss tests before the
// chord were measured to cost more than the chords they saved.)
namespace {
constexpr int FP_A = 40;    // footprint columns held in LDS
constexpr int FP_B = 32;    // footprint rows

// wave-wide min / max (DPP row shifts and lane reads, no LDS round trips);
// called with every lane of the wave active
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }
}  // namespace

// Exact cone adjoint, z-column form.  A workgroup is an 8 x 8 x (4*KZ)
// voxel brick; each thread owns KZ z-adjacent voxels of one (x, y) column,
// which share the view rotation, the detector-column window and the
// workgroup's per-view footprint tables in LDS.  Per voxel, the cells and
// the order of accumulation are those of the reference kernel
// (Atx_cone_mf_gpu_new.cu:58-86: views in order, rows outer, columns
// inner), restricted to the cells whose chord can be non-zero.
// Five workgroups per CU (the LDS tables take 26.6 KB each; the register
// budget that allows five is the measured optimum at 256^3).
template <typename YT, typename XT, int KZ>
__global__ __launch_bounds__(256, FL_EXACT_WG_PER_CU) void k_adj_cone_zc(const YT* __restrict__ Y, XT* __restrict__ X, AdjGeom g)
{
    __shared__ float c_a[FP_A], c_b[FP_A], c_sba[FP_A], c_sab[FP_A];
    __shared__ float c_ra[FP_A], c_rb[FP_A];
    __shared__ float c_hx[FP_A], c_hy[FP_A];   // box half-sides signed by the ray's direction
    __shared__ float r_c[FP_B], r_rc[FP_B];
    __shared__ float4 q_r[FP_A * FP_B];
    __shared__ float q_y[FP_A * FP_B];
    __shared__ int red[2][4][4];
    __shared__ int s_special[2];   // per view (double-buffered): a footprint ray with no x or no y direction

    constexpr int BZ = 4 * KZ;
    const int bx = (g.nx + 7) / 8, by = (g.ny + 7) / 8, bz = (g.nz + BZ - 1) / BZ;
    const uint32_t nblk = (uint32_t)bx * by * bz;
    const uint32_t b = fl_xcd_remap(blockIdx.x, nblk);
    const int it = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ix = (int)(b % bx) * 8 + (tid & 7);
    const int iy = (int)((b / bx) % by) * 8 + ((tid >> 3) & 7);
    const int iz0 = (int)(b / ((uint32_t)bx * by)) * BZ + wv * KZ;
    const bool col = ix < g.nx && iy < g.ny;

    const int na = g.na, nb = g.nb, na2 = na / 2, nb2 = nb / 2;
    const float SD = g.SO + g.OD;
    const float d = sqrtf((1 + g.dz * g.dz) / 2);
    const float yc = (float)iy + 0.5f - (float)(g.ny / 2);
    const float xc = (float)ix + 0.5f - (float)(g.nx / 2);
    const float hz = g.dz / 2;
    const ShadowMap ma = fl_shadow_map(g.ya0, g.inv_pa, 0.05f), mb = fl_shadow_map(g.zb0, g.inv_pb, 0.05f);
    float zc[KZ], acc[KZ];
#pragma unroll
    for (int k = 0; k < KZ; k++) {
        zc[k] = ((float)(iz0 + k) + 0.5f - (float)(g.nz / 2)) * g.dz;
        acc[k] = 0.0f;
    }
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)g.id_S[it * g.tmp_size + j] * g.nd;
        const float c = g.cphi[iv], s = g.sphi[iv];
        const float sz = g.sdz[iv];
        int alo = na, ahi = -1;
        int blo[KZ], bhi[KZ];
        int ublo = nb, ubhi = -1;
#pragma unroll
        for (int k = 0; k < KZ; k++) { blo[k] = nb; bhi[k] = -1; }
        if (col) {
            const float xr = c * xc + s * yc;
            const float yr = -s * xc + c * yc;
            // shadow window (clamped to the panel), cut by the reference's
            // window where that can bind (the clamps of the latter are
            // implied by the former's)
            const ShadowCols sc = fl_shadow_cols(c, s, g.SO, SD, xc, yc, ma, na);
            alo = sc.alo;
            ahi = sc.ahi;
            if (g.win_ref & 1) {
                const float t = SD / ((xr + g.SO) * g.dy_det);
                ahi = min(ahi, (int)ceilf((yr + 1) * t - g.y_os + (float)na2));
                alo = max(alo, (int)floorf((yr - 1) * t - g.y_os + (float)na2));
            }
            const bool rows_ref = (g.win_ref & 2) != 0;
            float t = 0.0f;
            if (rows_ref) t = SD / ((xr + g.SO) * g.dz_det);
#pragma unroll
            for (int k = 0; k < KZ; k++) {
                if (iz0 + k < g.nz && alo <= ahi) {
                    const float zr = zc[k] - sz;
                    int bl, bh;
                    fl_shadow_rows(sc, zr, hz, mb, nb, bl, bh);
                    if (rows_ref) {
                        bh = min(bh, (int)ceilf((zr + d) * t + (float)nb2));
                        bl = max(bl, (int)floorf((zr - d) * t + (float)nb2));
                    }
                    if (bl <= bh) {
                        blo[k] = bl;
                        bhi[k] = bh;
                        ublo = min(ublo, bl);
                        ubhi = max(ubhi, bh);
                    }
                }
            }
            if (ubhi < 0) { alo = na; ahi = -1; }
        }
        // workgroup footprint = union of the voxels' windows
        {
            int m0 = wave_min(alo), m1 = wave_max(ahi), m2 = wave_min(ublo), m3 = wave_max(ubhi);
            if (lane == 0) {
                red[j & 1][wv][0] = m0; red[j & 1][wv][1] = m1;
                red[j & 1][wv][2] = m2; red[j & 1][wv][3] = m3;
            }
            // last read two views ago, before the barrier every thread has since passed
            if (tid == 0) s_special[j & 1] = 0;
        }
        __syncthreads();
        int A0 = red[j & 1][0][0], A1 = red[j & 1][0][1], B0 = red[j & 1][0][2], B1 = red[j & 1][0][3];
#pragma unroll
        for (int w = 1; w < 4; w++) {
            A0 = min(A0, red[j & 1][w][0]); A1 = max(A1, red[j & 1][w][1]);
            B0 = min(B0, red[j & 1][w][2]); B1 = max(B1, red[j & 1][w][3]);
        }
        const int nA = A1 - A0 + 1, nB = B1 - B0 + 1;
        if (nA <= 0 || nB <= 0) continue;   // uniform: no voxel of the brick sees this view
        const float x1 = c * (-g.SO), y1 = s * (-g.SO), z1 = sz;
        if (nA <= FP_A && nB <= FP_B) {
            for (int k = tid; k < nA; k += 256) {
                const float yd = g.ydet[A0 + k];
                const float x2 = c * g.OD - s * yd, y2 = s * g.OD + c * yd;
                const float a = x2 - x1, bb = y2 - y1;
                const float ra = 1.0f / a, rb = 1.0f / bb;
                c_a[k] = a; c_b[k] = bb;
                c_ra[k] = ra; c_rb[k] = rb;
                c_hx[k] = a > 0 ? 0.5f : -0.5f;
                c_hy[k] = bb > 0 ? 0.5f : -0.5f;
                if (a == 0 || bb == 0) s_special[j & 1] = 1;
                c_sba[k] = fl_div_r(bb, a, ra); c_sab[k] = fl_div_r(a, bb, rb);
            }
            for (int k = tid; k < nB; k += 256) {
                const float cc = (g.zdet[B0 + k] + sz) - z1;
                r_c[k] = cc;
                r_rc[k] = 1.0f / cc;
            }
            __syncthreads();
            // the ratios' IEEE quotients from the staged reciprocals
            const uint32_t mA = fl_small_div_m((uint32_t)nA);   // k * nA < 2^18 in these tables
            for (int k = tid; k < nA * nB; k += 256) {
                const int kb = fl_small_div(k, mA), ka = k - kb * nA;
                const float a = c_a[ka], bb = c_b[ka], cc = r_c[kb], rc = r_rc[kb];
                q_r[k] = make_float4(fl_div_r(cc, a, c_ra[ka]), fl_div_r(cc, bb, c_rb[ka]), fl_div_r(a, cc, rc),
                                     fl_div_r(bb, cc, rc));
                q_y[k] = fl_load(y, (size_t)(B0 + kb) * na + (A0 + ka));
            }
            __syncthreads();
            // the chord's 2-D branch only where the view has such a ray:
            // the common case runs without the per-cell test
            auto cells = [&](auto special) {
                const float h[3] = {0.5f, 0.5f, hz};
#pragma unroll
                for (int k = 0; k < KZ; k++) {
                    const float p[3] = {x1 - xc, y1 - yc, z1 - zc[k]};
                    for (int ib = blo[k]; ib <= bhi[k]; ib++) {
                        const int kb = ib - B0;
                        const float hzs = r_c[kb] > 0 ? hz : -hz;
                        for (int ia = alo; ia <= ahi; ia++) {
                            const int ka = ia - A0;
                            float l;
                            if (decltype(special)::value && (c_a[ka] == 0 || c_b[ka] == 0)) {
                                const float yd = g.ydet[ia];
                                l = fl_chord_box3(x1, y1, z1, c * g.OD - s * yd, s * g.OD + c * yd,
                                                  g.zdet[ib] + sz, 1.0f, 1.0f, g.dz, xc, yc, zc[k]);
                            } else {
                                const float4 r4 = q_r[kb * nA + ka];
                                FlRatiosS q;
                                q.r.sba = c_sba[ka]; q.r.sab = c_sab[ka];
                                q.r.sca = r4.x; q.r.scb = r4.y; q.r.sac = r4.z; q.r.sbc = r4.w;
                                q.hs[0] = c_hx[ka]; q.hs[1] = c_hy[ka]; q.hs[2] = hzs;
                                l = fl_chord_box3_fast(p, h, q, [&] { return fl_absf(c_a[ka]) > fl_absf(c_b[ka]); });
                            }
                            acc[k] = acc[k] + l * q_y[kb * nA + ka];
                        }
                    }
                }
            };
            if (alo <= ahi) {
                if (s_special[j & 1]) cells(std::true_type{});
                else cells(std::false_type{});
            }
        } else if (alo <= ahi) {
#pragma unroll
            for (int k = 0; k < KZ; k++) {
                for (int ib = blo[k]; ib <= bhi[k]; ib++) {
                    const float z2 = g.zdet[ib] + sz;
                    for (int ia = alo; ia <= ahi; ia++) {
                        const float yd = g.ydet[ia];
                        const float l = fl_chord_box3(x1, y1, z1, c * g.OD - s * yd, s * g.OD + c * yd, z2,
                                                      1.0f, 1.0f, g.dz, xc, yc, zc[k]);
                        acc[k] = acc[k] + l * fl_load(y, (size_t)ib * na + ia);
                    }
                }
            }
        }
    }
    if (col) {
#pragma unroll
        for (int k = 0; k < KZ; k++)
            if (iz0 + k < g.nz)
                fl_store(X, (size_t)it * g.nvox + ((size_t)(iz0 + k) * g.ny + iy) * g.nx + ix, acc[k] * g.scale);
    }
}

// Per (view, cell) of a fan plan, the two quotients of the cell's line
// chord that do not depend on the pixel (fl_chord_line2_q): the line's
// coefficients formed exactly as fl_fan_adj_pixel forms them.
__global__ __launch_bounds__(256) void k_fan_line_q(AdjGeom g, int nv, float2* __restrict__ out)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (size_t)nv * g.na) return;
    const int iv = (int)(k / g.na), id = (int)(k % g.na);
    const float c = g.cphi[iv], s = g.sphi[iv];
    const float sd = g.sdet[id], cd = g.cdet[id];
    const float Px = -(sd * c + cd * s);
    const float Py = cd * c - sd * s;
    const float a2 = fl_absf(Px), b2 = fl_absf(Py);
    const float hx = a2 / 2, hy = b2 / 2;
    out[k] = make_float2((hx < hy) ? 1.0f / b2 : 1.0f / a2, 1.0f / (a2 * b2));
}

// Per (view, cell) of a fan plan: (1/dx, 1/dy, |d|) of the ray, formed as
// fl_fan_matched_pixel forms them (zero direction components give infinite
// reciprocals, which the slab routine never uses).
__global__ __launch_bounds__(256) void k_fan_ray_q(FwdGeom f, int nv, float4* __restrict__ out)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= (size_t)nv * f.na) return;
    const int iv = (int)(k / f.na), id = (int)(k % f.na);
    const float c = f.cs[iv], s = f.sn[iv];
    const float x1 = c * (-f.SO), y1 = s * (-f.SO);
    const float yd = f.ydet[id];
    const float x2 = c * f.OD - s * yd;
    const float y2 = s * f.OD + c * yd;
    const float dx = x2 - x1, dy = y2 - y1;
    out[k] = make_float4(1.0f / dx, 1.0f / dy, sqrtf(dx * dx + dy * dy), 0.0f);
}

hipError_t fl_launch_fan_ray_q(const FwdGeom& f, int nv, float4* out, hipStream_t st)
{
    const size_t n = (size_t)nv * f.na;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fan_ray_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, f, nv, out);
    return hipGetLastError();
}

hipError_t fl_launch_fan_line_q(const AdjGeom& g, int nv, float2* out, hipStream_t st)
{
    const size_t n = (size_t)nv * g.na;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fan_line_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nv, out);
    return hipGetLastError();
}

template <typename YT, typename XT>
static hipError_t launch_adj_fan_t(const void* Y, void* X, const AdjGeom& g, hipStream_t st)
{
    dim3 grid((g.nx + 15) / 16, (g.ny + 15) / 16, g.nt);
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return hipSuccess;
    hipLaunchKernelGGL((k_adj_fan<YT, XT>), grid, dim3(16, 16), 0, st, (const YT*)Y, (XT*)X, g);
    return hipGetLastError();
}

template <typename YT, typename XT>
static hipError_t launch_adj_cone_t(const void* Y, void* X, const AdjGeom& g, hipStream_t st)
{
    constexpr int KZ = 4;
    uint32_t nblk = (uint32_t)((g.nx + 7) / 8) * ((g.ny + 7) / 8) * ((g.nz + 4 * KZ - 1) / (4 * KZ));
    if (nblk == 0 || g.nt == 0) return hipSuccess;
    hipLaunchKernelGGL((k_adj_cone_zc<YT, XT, KZ>), dim3(nblk, g.nt), dim3(256), 0, st, (const YT*)Y, (XT*)X, g);
    return hipGetLastError();
}

#define FL_DISPATCH_YX(FN, yt, xt, ...)                                                      \
    do {                                                                                     \
        if (yt == FL_F32 && xt == FL_F32) return FN<float, float>(__VA_ARGS__);               \
        if (yt == FL_BF16 && xt == FL_BF16) return FN<bf16_t, bf16_t>(__VA_ARGS__);           \
        if (yt == FL_F16 && xt == FL_F16) return FN<f16_t, f16_t>(__VA_ARGS__);               \
        if (yt == FL_BF16 && xt == FL_F32) return FN<bf16_t, float>(__VA_ARGS__);             \
        if (yt == FL_F16 && xt == FL_F32) return FN<f16_t, float>(__VA_ARGS__);               \
        if (yt == FL_F32 && xt == FL_BF16) return FN<float, bf16_t>(__VA_ARGS__);             \
        if (yt == FL_F32 && xt == FL_F16) return FN<float, f16_t>(__VA_ARGS__);               \
        return hipErrorInvalidValue;                                                         \
    } while (0)

hipError_t fl_launch_adj_fan(const void* Y, int yt, void* X, int xt, const AdjGeom& g, hipStream_t st)
{
    FL_DISPATCH_YX(launch_adj_fan_t, yt, xt, Y, X, g, st);
}

hipError_t fl_launch_adj_cone(const void* Y, int yt, void* X, int xt, const AdjGeom& g, hipStream_t st)
{
    FL_DISPATCH_YX(launch_adj_cone_t, yt, xt, Y, X, g, st);
}
