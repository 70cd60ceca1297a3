// matched.hip -- voxel-driven transpose of the Siddon forward projector.
//
// For voxel v and a ray crossing it, the Siddon walk's segment inside v runs
// from the largest entry-plane parameter to the smallest exit-plane
// parameter of v's three face pairs; no other plane parameter of the ray
// lies strictly between them (each axis' parameters are monotone in the
// plane index), so that interval is exactly one element pair of the forward
// walk.  Computing the plane parameters with the forward's float formula
// ((float)i*s - o)/d and the length as (a_out - a_in)*|d| therefore gives the
// forward's coefficient bit for bit, and the forward's midpoint rule decides
// whether the walk credits v with it.  The result is the transpose of the
// FP32 forward coefficient by coefficient (up to segments shorter than the
// float resolution whose rounded midpoint falls in a neighbouring voxel), so
// <Ax,y> = <x,A^T y> holds to summation rounding.  Gathers only: each thread
// owns one voxel, no atomics, deterministic.
//
// No reference counterpart: the reference pairs its Siddon forward with the
// exact-chord adjoint, whose float chords differ from the forward's lengths
// by ~1e-5 relative.  This is the backward used for gradients.
#include "fl_capi_internal.h"
#include "fl_common.h"
#include "fl_geom.h"
#include "fl_siddon.h"
#include "fl_chord.h"
#include "fl_fan.h"

// Workgroups per CU the register budget is sized for (a build-time tuning
// knob; the default is the measured optimum at 256^3).
#ifndef FL_MATCHED_WG_PER_CU
#define FL_MATCHED_WG_PER_CU 5
#endif



// Per (ray, voxel) pair of one view: the Siddon segment of the ray inside
// the voxel, credited when the forward's midpoint rule maps it there.
// in/out: entry/exit plane parameters per axis.  Returns the length
// coefficient (a_out - a_in)*|d|, or 0.  Branch-free: the three cell tests
// cost less than the divergent branches around them.
__device__ __forceinline__ float matched_coef(float inx, float outx, float iny, float outy, float inz, float outz,
                                              const SidAxis& ax, const SidAxis& ay, const SidAxis& az, int ix,
                                              int iy, int iz, bool dz_one, float L)
{
    const float ain = fmaxf(fmaxf(inx, iny), inz), aout = fminf(fminf(outx, outy), outz);
    const float m = (ain + aout) * 0.5f;
    // (every test evaluated: no short-circuit branches)
    const int hit = int(aout > ain) & int(sid_cell_xy(ax, m) == ix) & int(sid_cell_xy(ay, m) == iy) &
                    int(sid_cell_z(az, m, dz_one) == iz);
    return hit ? (aout - ain) * L : 0.0f;
}

namespace {
constexpr int MF_A = 48;   // footprint columns in LDS
constexpr int MF_B = 32;   // footprint rows
static_assert(MF_A <= 64 && MF_B <= 64, "footprint columns / rows are staged by one wave each");

// wave-wide min / max (DPP row shifts and lane reads, no LDS round trips);
// called with every lane of the wave active
__device__ __forceinline__ int mwave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int mwave_max(int v) { return __ockl_wfred_max_i32(v); }
}  // namespace

// 3D, footprint-staged, z-column form: a workgroup is an 8 x 8 x (4*KZ)
// voxel brick and each thread owns KZ z-adjacent voxels of one (x, y)
// column.  Per view the workgroup stages in LDS the plane parameters of its
// nine x planes and nine y planes for each footprint column, and for each
// footprint row the entry and exit parameters of each of its 4*KZ slices'
// z planes (each the forward's IEEE division, done once per workgroup),
// the ray lengths |d| and the sinogram values; each
// voxel then visits only the cells of its own shadow (the shadow of its
// column is shared by the four waves through LDS).  Views of each frame in
// order, columns outer; FP32 accumulation, one fused multiply-add per term.
// Five workgroups per CU (measured optimum of the register budget).
template <typename YT, typename XT, int KZ, bool DZ1>
__global__ __launch_bounds__(256, FL_MATCHED_WG_PER_CU) void k_adj_matched_cone(const YT* __restrict__ Y, XT* __restrict__ X,
                                                         FwdGeom f, AdjGeom g)
{
    constexpr int BZ = 4 * KZ;
    __shared__ float t_ax[MF_A][9], t_ay[MF_A][9];
    __shared__ float t_dx[MF_A], t_dy[MF_A], t_rx[MF_A], t_ry[MF_A], t_dxy2[MF_A];
    __shared__ float2 t_zio[MF_B][BZ];   // per row and brick slice: z entry / exit plane parameters
    __shared__ float t_dz[MF_B], t_rz[MF_B], t_dz2[MF_B];
    __shared__ float t_L[MF_A * MF_B], t_y[MF_A * MF_B];
    __shared__ int red[2][4][4];

    const int bx = (f.nx + 7) / 8, by = (f.ny + 7) / 8, bz = (f.nz + BZ - 1) / BZ;
    const uint32_t nblk = (uint32_t)bx * by * bz;
    const uint32_t b = fl_xcd_remap(blockIdx.x, nblk);
    const int it = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ix0 = (int)(b % bx) * 8, iy0 = (int)((b / bx) % by) * 8, iz0 = (int)(b / ((uint32_t)bx * by)) * BZ;
    const int lx = tid & 7, ly = (tid >> 3) & 7, lz0 = wv * KZ;
    const int ix = ix0 + lx, iy = iy0 + ly;
    const bool col = ix < f.nx && iy < f.ny;

    const int na = f.na, nb = f.nb;
    const float SD = f.SO + f.OD;
    const float xc = (float)ix + 0.5f - (float)(f.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(f.ny / 2);
    const float hz = 0.5f * f.dz0;
    const ShadowMap ma = fl_shadow_map(f.ya0, f.inv_pa, 0.05f), mb = fl_shadow_map(f.zb0, f.inv_pb, 0.05f);
    // DZ1: z planes one voxel apart, so the z cell index needs no division
    // (a compile-time choice: a runtime flag lets the divide be speculated)
    constexpr bool dz_one = DZ1;
    float acc[KZ];
#pragma unroll
    for (int k = 0; k < KZ; k++) acc[k] = 0.0f;
    const int nvf = g.Nv[it];
    // The shadow of a thread's (x, y) column is the same in all four waves
    // (lane -> column is wave-independent): one wave per view computes it a
    // view ahead, in turn, into double-buffered LDS, written before and read
    // after the view's first barrier.
    __shared__ __align__(16) ShadowCols s_cols[2][64];
    auto col_shadow = [&](int jj) {
        const int ivv = g.id_Y[it * g.tmp_size + jj];
        return fl_shadow_cols(f.cs[ivv], f.sn[ivv], f.SO, SD, xc, yc, ma, na);
    };
    if (wv == 0 && nvf > 0) s_cols[0][lane] = col_shadow(0);
    __syncthreads();
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)g.id_S[it * g.tmp_size + j] * g.nd;
        const float c = f.cs[iv], s = f.sn[iv];
        const float sz = f.sdz[iv];
        int alo = na, ahi = -1, ublo = nb, ubhi = -1;
        int blo[KZ], bhi[KZ];
#pragma unroll
        for (int k = 0; k < KZ; k++) { blo[k] = nb; bhi[k] = -1; }
        const ShadowCols sc = s_cols[j & 1][lane];
        if (wv == (j & 3) && j + 1 < nvf) s_cols[(j + 1) & 1][lane] = col_shadow(j + 1);
        if (col) {
            alo = sc.alo;
            ahi = sc.ahi;
#pragma unroll
            for (int k = 0; k < KZ; k++) {
                const int iz = iz0 + lz0 + k;
                if (iz < f.nz && alo <= ahi) {
                    const float zc = ((float)iz + 0.5f - (float)(f.nz / 2)) * f.dz0;
                    int bl, bh;
                    fl_shadow_rows(sc, zc - sz, hz, mb, nb, bl, bh);
                    if (bl <= bh) {
                        blo[k] = bl; bhi[k] = bh;
                        ublo = min(ublo, bl); ubhi = max(ubhi, bh);
                    }
                }
            }
            if (ubhi < 0) { alo = na; ahi = -1; }
        }
        {
            int m0 = mwave_min(alo), m1 = mwave_max(ahi), m2 = mwave_min(ublo), m3 = mwave_max(ubhi);
            if (lane == 0) {
                red[j & 1][wv][0] = m0; red[j & 1][wv][1] = m1;
                red[j & 1][wv][2] = m2; red[j & 1][wv][3] = m3;
            }
        }
        __syncthreads();
        int A0 = red[j & 1][0][0], A1 = red[j & 1][0][1], B0 = red[j & 1][0][2], B1 = red[j & 1][0][3];
#pragma unroll
        for (int w = 1; w < 4; w++) {
            A0 = min(A0, red[j & 1][w][0]); A1 = max(A1, red[j & 1][w][1]);
            B0 = min(B0, red[j & 1][w][2]); B1 = max(B1, red[j & 1][w][3]);
        }
        const int nA = A1 - A0 + 1, nB = B1 - B0 + 1;
        if (nA <= 0 || nB <= 0) continue;
        const float x1 = c * (-f.SO), y1 = s * (-f.SO), z1 = sz;
        if (nA <= MF_A && nB <= MF_B) {
            // columns (wave 0): direction components; rows (wave 1) likewise;
            // both fit one wave, and their load and divide latencies overlap
            if (tid < nA) {
                const int k = tid;
                const float yd = f.ydet[A0 + k];
                const float dx = (c * f.OD - s * yd) - x1, dy = (s * f.OD + c * yd) - y1;
                t_dx[k] = dx; t_dy[k] = dy;
                t_dxy2[k] = dx * dx + dy * dy;   // the first two terms of |d|^2, summed as the forward sums them
                t_rx[k] = 1.0f / dx; t_ry[k] = 1.0f / dy;
            }
            if (tid >= 64 && tid - 64 < nB) {
                const int k = tid - 64;
                const float dz = (f.zdet[B0 + k] + sz) - z1;
                t_dz[k] = dz;
                t_dz2[k] = dz * dz;
                t_rz[k] = 1.0f / dz;
            }
            __syncthreads();
            // plane coordinate i -> ((float)i - o)/d, the forward's IEEE
            // quotient, from the staged reciprocal of d
            for (int k = tid; k < nA * 9; k += 256) {
                const int ka = k / 9, q = k % 9;
                t_ax[ka][q] = fl_div_r((float)(ix0 + q - f.nx / 2) - x1, t_dx[ka], t_rx[ka]);
                t_ay[ka][q] = fl_div_r((float)(iy0 + q - f.ny / 2) - y1, t_dy[ka], t_ry[ka]);
            }
            // per row and brick slice the parameters of the slice's two z
            // planes (the forward's quotients), ordered as entry and exit
            // along the ray (+-inf for a ray parallel to them)
            for (int k = tid; k < nB * BZ; k += 256) {
                const int kb = k / BZ, q = k % BZ;
                const float dz = t_dz[kb], rz = t_rz[kb];
                const float lo = fl_div_r((float)(iz0 + q - f.nz / 2) * f.dz0 - z1, dz, rz);
                const float hi = fl_div_r((float)(iz0 + q + 1 - f.nz / 2) * f.dz0 - z1, dz, rz);
                t_zio[kb][q] = dz != 0.0f ? make_float2(dz > 0 ? lo : hi, dz > 0 ? hi : lo)
                                          : make_float2(-FL_INF, FL_INF);
            }
            const uint32_t mA = fl_small_div_m((uint32_t)nA);   // k * nA < 2^18 in these tables
            for (int k = tid; k < nA * nB; k += 256) {
                const int kb = fl_small_div(k, mA), ka = k - kb * nA;
                // |d| = sqrtf(dx*dx + dy*dy + dz*dz), the forward's ray length
                // bit for bit (left-to-right sum, correctly rounded root)
                t_L[k] = sqrtf(t_dxy2[ka] + t_dz2[kb]);
                t_y[k] = fl_load(y, (size_t)(B0 + kb) * na + (A0 + ka));
            }
            __syncthreads();
            // columns outermost (no reference order to keep): the column's x/y
            // slab interval is formed once and serves every row and voxel
            for (int ia = alo; ia <= ahi; ia++) {
                const int ka = ia - A0;
                const float dx = t_dx[ka], dy = t_dy[ka];
                SidAxis ax, ay;
                sid_axis_init(ax, x1, dx, 1.0f, f.nx);
                sid_axis_init(ay, y1, dy, 1.0f, f.ny);
                const float x_lo = t_ax[ka][lx], x_hi = t_ax[ka][lx + 1];
                const float y_lo = t_ay[ka][ly], y_hi = t_ay[ka][ly + 1];
                const float inx = ax.on ? (dx > 0 ? x_lo : x_hi) : -FL_INF;
                const float outx = ax.on ? (dx > 0 ? x_hi : x_lo) : FL_INF;
                const float iny = ay.on ? (dy > 0 ? y_lo : y_hi) : -FL_INF;
                const float outy = ay.on ? (dy > 0 ? y_hi : y_lo) : FL_INF;
                if (!(fminf(outx, outy) > fmaxf(inx, iny))) continue;   // misses the (x, y) column
#pragma unroll
                for (int k = 0; k < KZ; k++) {
                    const int lz = lz0 + k, iz = iz0 + lz;
                    for (int ib = blo[k]; ib <= bhi[k]; ib++) {
                        const int kb = ib - B0;
                        const float dz = t_dz[kb];
                        SidAxis az;
                        sid_axis_init(az, z1, dz, f.dz0, f.nz);
                        const float2 zio = t_zio[kb][lz];
                        const float inz = zio.x, outz = zio.y;
                        const int q = kb * nA + ka;
                        const float Si = matched_coef(inx, outx, iny, outy, inz, outz, ax, ay, az, ix, iy, iz,
                                                      dz_one, t_L[q]);
                        // one fused multiply-add (no reference rounding to keep;
                        // +-0 for a miss leaves the sum unchanged)
                        acc[k] = __builtin_fmaf(Si, t_y[q], acc[k]);
                    }
                }
            }
        } else if (alo <= ahi) {
#pragma unroll
            for (int k = 0; k < KZ; k++) {
                const int iz = iz0 + lz0 + k;
                for (int ib = blo[k]; ib <= bhi[k]; ib++) {
                    const float dz = (f.zdet[ib] + sz) - z1;
                    SidAxis az;
                    sid_axis_init(az, z1, dz, f.dz0, f.nz);
                    const FlAxisIv sv = fl_slab_r(az, iz, 1.0f / dz);
                    for (int ia = alo; ia <= ahi; ia++) {
                        const float yd = f.ydet[ia];
                        const float dx = (c * f.OD - s * yd) - x1, dy = (s * f.OD + c * yd) - y1;
                        SidAxis ax, ay;
                        sid_axis_init(ax, x1, dx, 1.0f, f.nx);
                        sid_axis_init(ay, y1, dy, 1.0f, f.ny);
                        const FlAxisIv sx = fl_slab_r(ax, ix, 1.0f / dx), sy = fl_slab_r(ay, iy, 1.0f / dy);
                        const float L = sqrtf(dx * dx + dy * dy + dz * dz);
                        const float Si = matched_coef(sx.in, sx.out, sy.in, sy.out, sv.in, sv.out, ax, ay, az, ix,
                                                      iy, iz, dz_one, L);
                        if (Si != 0.0f) acc[k] = __builtin_fmaf(Si, fl_load(y, (size_t)ib * na + ia), acc[k]);
                    }
                }
            }
        }
    }
    if (col) {
#pragma unroll
        for (int k = 0; k < KZ; k++) {
            const int iz = iz0 + lz0 + k;
            if (iz < f.nz) fl_store(X, (size_t)it * f.nvox + ((size_t)iz * f.ny + iy) * f.nx + ix, acc[k] * f.scale);
        }
    }
}

template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_matched_fan(const YT* __restrict__ Y, XT* __restrict__ X,
                                                        FwdGeom f, AdjGeom g)
{
    const int ix = blockIdx.x * 16 + threadIdx.x;
    const int iy = blockIdx.y * 16 + threadIdx.y;
    const int it = blockIdx.z;
    if (ix >= f.nx || iy >= f.ny) return;
    fl_store(X, (size_t)it * f.nvox + (size_t)iy * f.nx + ix, fl_fan_matched_pixel(Y, f, g, it, ix, iy));
}

template <typename YT, typename XT>
static hipError_t launch_matched_t(const void* Y, void* X, const fl_plan_s* p, hipStream_t st)
{
    const FwdGeom& f = p->fg;
    if (p->cone) {
        constexpr int KZ = 4;
        uint32_t nblk = (uint32_t)((f.nx + 7) / 8) * ((f.ny + 7) / 8) * ((f.nz + 4 * KZ - 1) / (4 * KZ));
        if (f.dz0 == 1.0f)
            hipLaunchKernelGGL((k_adj_matched_cone<YT, XT, KZ, true>), dim3(nblk, p->nt), dim3(256), 0, st,
                               (const YT*)Y, (XT*)X, f, p->ag);
        else
            hipLaunchKernelGGL((k_adj_matched_cone<YT, XT, KZ, false>), dim3(nblk, p->nt), dim3(256), 0, st,
                               (const YT*)Y, (XT*)X, f, p->ag);
    } else {
        dim3 grid((f.nx + 15) / 16, (f.ny + 15) / 16, p->nt);
        hipLaunchKernelGGL((k_adj_matched_fan<YT, XT>), grid, dim3(16, 16), 0, st, (const YT*)Y, (XT*)X, f,
                           p->ag);
    }
    return hipGetLastError();
}

hipError_t fl_launch_adj_matched(const void* Y, int yt, void* X, int xt, const fl_plan_s* p, hipStream_t st)
{
    if (yt == FL_F32 && xt == FL_F32) return launch_matched_t<float, float>(Y, X, p, st);
    if (yt == FL_BF16 && xt == FL_BF16) return launch_matched_t<bf16_t, bf16_t>(Y, X, p, st);
    if (yt == FL_F16 && xt == FL_F16) return launch_matched_t<f16_t, f16_t>(Y, X, p, st);
    if (yt == FL_BF16 && xt == FL_F32) return launch_matched_t<bf16_t, float>(Y, X, p, st);
    if (yt == FL_F16 && xt == FL_F32) return launch_matched_t<f16_t, float>(Y, X, p, st);
    if (yt == FL_F32 && xt == FL_BF16) return launch_matched_t<float, bf16_t>(Y, X, p, st);
    if (yt == FL_F32 && xt == FL_F16) return launch_matched_t<float, f16_t>(Y, X, p, st);
    return hipErrorInvalidValue;
}
