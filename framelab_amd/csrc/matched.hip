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

namespace {

struct AxisIv {
    float in, out;  // entry / exit parameters of the voxel slab along one axis
};

__device__ __forceinline__ AxisIv slab(const SidAxis& a, int cell)
{
    AxisIv r;
    if (!a.on) {
        r.in = -FL_INF;
        r.out = FL_INF;
        return r;
    }
    float lo = ((float)(cell - a.nh) * a.s - a.o) / a.d;
    float hi = ((float)(cell - a.nh + 1) * a.s - a.o) / a.d;
    r.in = a.st > 0 ? lo : hi;
    r.out = a.st > 0 ? hi : lo;
    return r;
}

}  // namespace

// 3D: one thread per voxel in 8x8x4 bricks; views of each frame in order.
template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_matched_cone(const YT* __restrict__ Y, XT* __restrict__ X,
                                                         FwdGeom f, AdjGeom g)
{
    const int bx = (f.nx + 7) / 8, by = (f.ny + 7) / 8, bz = (f.nz + 3) / 4;
    const uint32_t nblk = (uint32_t)bx * by * bz;
    const uint32_t b = fl_xcd_remap(blockIdx.x, nblk);
    const int it = blockIdx.y;
    const int ix = (int)(b % bx) * 8 + (threadIdx.x & 7);
    const int iy = (int)((b / bx) % by) * 8 + ((threadIdx.x >> 3) & 7);
    const int iz = (int)(b / ((uint32_t)bx * by)) * 4 + (threadIdx.x >> 6);
    if (ix >= f.nx || iy >= f.ny || iz >= f.nz) return;

    const int na = f.na, nb = f.nb;
    const float SD = f.SO + f.OD;
    const float xc = (float)ix + 0.5f - (float)(f.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(f.ny / 2);
    const float zc = ((float)iz + 0.5f - (float)(f.nz / 2)) * f.dz0;
    const float ya0 = f.ydet[0], zb0 = f.zdet[0];
    const float hz = 0.5f * f.dz0 + 0.5f;   // z half-extent plus a cell of margin (voxel units)
    const bool dz_one = f.dz0 == 1.0f;
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)iv * g.nd;
        const float c = f.cs[iv], s = f.sn[iv];
        const float sz = f.sdz[iv];
        // detector window of the voxel's shadow, with a margin of one cell
        const float xr = c * xc + s * yc, yr = -s * xc + c * yc;
        const float mag = SD / (xr + f.SO);
        int alo = (int)floorf(((yr - 1.0f) * mag - ya0) / g.dy_det) - 1;
        int ahi = (int)ceilf(((yr + 1.0f) * mag - ya0) / g.dy_det) + 1;
        const float zr = zc - sz;
        const float mz = SD / (xr + f.SO - 0.75f);
        const float zlo = fminf((zr - hz) * mz, (zr - hz) * mag), zhi = fmaxf((zr + hz) * mz, (zr + hz) * mag);
        int blo = (int)floorf((zlo - zb0) / g.dz_det) - 1;
        int bhi = (int)ceilf((zhi - zb0) / g.dz_det) + 1;
        if (alo < 0) alo = 0;
        if (ahi > na - 1) ahi = na - 1;
        if (blo < 0) blo = 0;
        if (bhi > nb - 1) bhi = nb - 1;

        const float x1 = c * (-f.SO), y1 = s * (-f.SO), z1 = sz;
        for (int ia = alo; ia <= ahi; ia++) {
            const float yd = f.ydet[ia];
            const float x2 = c * f.OD - s * yd;
            const float y2 = s * f.OD + c * yd;
            const float dx = x2 - x1, dy = y2 - y1;
            SidAxis ax, ay;
            sid_axis_init(ax, x1, dx, 1.0f, f.nx);
            sid_axis_init(ay, y1, dy, 1.0f, f.ny);
            const AxisIv sx = slab(ax, ix), sy = slab(ay, iy);
            const float in_xy = fmaxf(sx.in, sy.in), out_xy = fminf(sx.out, sy.out);
            if (!(out_xy > in_xy)) continue;
            const float dxy2 = dx * dx + dy * dy;
            for (int ib = blo; ib <= bhi; ib++) {
                const float z2 = f.zdet[ib] + sz;
                const float dz = z2 - z1;
                SidAxis az;
                sid_axis_init(az, z1, dz, f.dz0, f.nz);
                const AxisIv sv = slab(az, iz);
                const float ain = fmaxf(in_xy, sv.in), aout = fminf(out_xy, sv.out);
                if (!(aout > ain)) continue;
                const float m = (ain + aout) * 0.5f;
                if (sid_cell_xy(ax, m) != ix || sid_cell_xy(ay, m) != iy || sid_cell_z(az, m, dz_one) != iz)
                    continue;
                const float L = sqrtf(dxy2 + dz * dz);
                const float Si = (aout - ain) * L;
                acc = acc + Si * fl_load(y, (size_t)ib * na + ia);
            }
        }
    }
    fl_store(X, (size_t)it * f.nvox + ((size_t)iz * f.ny + iy) * f.nx + ix, acc * f.scale);
}

template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_matched_fan(const YT* __restrict__ Y, XT* __restrict__ X,
                                                        FwdGeom f, AdjGeom g)
{
    const int ix = blockIdx.x * 16 + threadIdx.x;
    const int iy = blockIdx.y * 16 + threadIdx.y;
    const int it = blockIdx.z;
    if (ix >= f.nx || iy >= f.ny) return;
    const int nd = f.na;
    const float SD = f.SO + f.OD;
    const float xc = (float)ix + 0.5f - (float)(f.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(f.ny / 2);
    const float ya0 = f.ydet[0];
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)iv * nd;
        const float c = f.cs[iv], s = f.sn[iv];
        const float xr = c * xc + s * yc, yr = -s * xc + c * yc;
        const float mag = SD / (xr + f.SO);
        int lo = (int)floorf(((yr - 1.0f) * mag - ya0) / g.dy_det) - 1;
        int hi = (int)ceilf(((yr + 1.0f) * mag - ya0) / g.dy_det) + 1;
        if (lo < 0) lo = 0;
        if (hi > nd - 1) hi = nd - 1;
        const float x1 = c * (-f.SO), y1 = s * (-f.SO);
        for (int id = lo; id <= hi; id++) {
            const float yd = f.ydet[id];
            const float x2 = c * f.OD - s * yd;
            const float y2 = s * f.OD + c * yd;
            const float dx = x2 - x1, dy = y2 - y1;
            SidAxis ax, ay;
            sid_axis_init(ax, x1, dx, 1.0f, f.nx);
            sid_axis_init(ay, y1, dy, 1.0f, f.ny);
            const AxisIv sx = slab(ax, ix), sy = slab(ay, iy);
            const float ain = fmaxf(sx.in, sy.in), aout = fminf(sx.out, sy.out);
            if (!(aout > ain)) continue;
            const float m = (ain + aout) * 0.5f;
            if (sid_cell_xy(ax, m) != ix || sid_cell_xy(ay, m) != iy) continue;
            const float L = sqrtf(dx * dx + dy * dy);
            acc = acc + ((aout - ain) * L) * fl_load(y, (size_t)id);
        }
    }
    fl_store(X, (size_t)it * f.nvox + (size_t)iy * f.nx + ix, acc * f.scale);
}

template <typename YT, typename XT>
static hipError_t launch_matched_t(const void* Y, void* X, const fl_plan_s* p, hipStream_t st)
{
    const FwdGeom& f = p->fg;
    if (p->cone) {
        uint32_t nblk = (uint32_t)((f.nx + 7) / 8) * ((f.ny + 7) / 8) * ((f.nz + 3) / 4);
        hipLaunchKernelGGL((k_adj_matched_cone<YT, XT>), dim3(nblk, p->nt), dim3(256), 0, st, (const YT*)Y,
                           (XT*)X, f, p->ag);
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
