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
#include "fl_chord.h"
#include "fl_common.h"
#include "fl_geom.h"

template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_fan(const YT* __restrict__ Y, XT* __restrict__ X, AdjGeom g)
{
    const int ix = blockIdx.x * 16 + threadIdx.x;
    const int iy = blockIdx.y * 16 + threadIdx.y;
    const int it = blockIdx.z;
    if (ix >= g.nx || iy >= g.ny) return;
    const int nd = g.na, nd2 = nd / 2;
    const float SD = g.SO + g.OD;
    const float xc = (float)ix + 0.5f - (float)(g.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(g.ny / 2);
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)iv * nd;
        const float c = g.cphi[iv], s = g.sphi[iv];
        const float xr = c * xc + s * yc;
        const float yr = -s * xc + c * yc;
        const float t = SD / ((xr + g.SO) * g.dy_det);
        int hi = (int)floorf((yr + 1) * t - g.y_os + (float)nd2);
        int lo = (int)floorf((yr - 1) * t - g.y_os + (float)nd2);
        if (lo < 0) lo = 0;
        if (hi > nd - 1) hi = nd - 1;
        for (int id = lo; id <= hi; id++) {
            const float sd = g.sdet[id], cd = g.cdet[id];
            const float Px = -(sd * c + cd * s);
            const float Py = cd * c - sd * s;
            const float Pd = g.SO * sd;
            const float l = fl_chord_line2(Px, Py, Pd, xc, yc);
            if (l > 0) acc = acc + l * fl_load(y, (size_t)id);
        }
    }
    fl_store(X, (size_t)it * g.nvox + (size_t)iy * g.nx + ix, acc * g.scale);
}

// Cone: a workgroup is an 8x8x4 voxel brick, so the 256 threads gather from
// one compact detector footprint per view (L1-resident).
template <typename YT, typename XT>
__global__ __launch_bounds__(256) void k_adj_cone(const YT* __restrict__ Y, XT* __restrict__ X, AdjGeom g)
{
    const int bx = (g.nx + 7) / 8, by = (g.ny + 7) / 8, bz = (g.nz + 3) / 4;
    const uint32_t nblk = (uint32_t)bx * by * bz;
    const uint32_t b = fl_xcd_remap(blockIdx.x, nblk);
    const int it = blockIdx.y;
    const int tbx = (int)(b % bx), tby = (int)((b / bx) % by), tbz = (int)(b / ((uint32_t)bx * by));
    const int ix = tbx * 8 + (threadIdx.x & 7);
    const int iy = tby * 8 + ((threadIdx.x >> 3) & 7);
    const int iz = tbz * 4 + (threadIdx.x >> 6);
    if (ix >= g.nx || iy >= g.ny || iz >= g.nz) return;

    const int na = g.na, nb = g.nb, na2 = na / 2, nb2 = nb / 2;
    const float SD = g.SO + g.OD;
    const float d = sqrtf((1 + g.dz * g.dz) / 2);
    const float zc = ((float)iz + 0.5f - (float)(g.nz / 2)) * g.dz;
    const float yc = (float)iy + 0.5f - (float)(g.ny / 2);
    const float xc = (float)ix + 0.5f - (float)(g.nx / 2);
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)iv * g.nd;
        const float c = g.cphi[iv], s = g.sphi[iv];
        const float sz = g.sdz[iv];
        const float xr = c * xc + s * yc;
        const float yr = -s * xc + c * yc;
        const float zr = zc - sz;
        float t = SD / ((xr + g.SO) * g.dy_det);
        int ahi = (int)ceilf((yr + 1) * t - g.y_os + (float)na2);
        int alo = (int)floorf((yr - 1) * t - g.y_os + (float)na2);
        t = SD / ((xr + g.SO) * g.dz_det);
        int bhi = (int)ceilf((zr + d) * t + (float)nb2);
        int blo = (int)floorf((zr - d) * t + (float)nb2);
        if (alo < 0) alo = 0;
        if (ahi > na - 1) ahi = na - 1;
        if (blo < 0) blo = 0;
        if (bhi > nb - 1) bhi = nb - 1;
        const float x1 = c * (-g.SO), y1 = s * (-g.SO), z1 = sz;
        for (int ib = blo; ib <= bhi; ib++) {
            const float z2 = g.zdet[ib] + sz;
            for (int ia = alo; ia <= ahi; ia++) {
                const float yd = g.ydet[ia];
                const float x2 = c * g.OD - s * yd;
                const float y2 = s * g.OD + c * yd;
                const float l = fl_chord_box3(x1, y1, z1, x2, y2, z2, 1.0f, 1.0f, g.dz, xc, yc, zc);
                acc = acc + l * fl_load(y, (size_t)ib * na + ia);
            }
        }
    }
    fl_store(X, (size_t)it * g.nvox + ((size_t)iz * g.ny + iy) * g.nx + ix, acc * g.scale);
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
    uint32_t nblk = (uint32_t)((g.nx + 7) / 8) * ((g.ny + 7) / 8) * ((g.nz + 3) / 4);
    if (nblk == 0 || g.nt == 0) return hipSuccess;
    hipLaunchKernelGGL((k_adj_cone<YT, XT>), dim3(nblk, g.nt), dim3(256), 0, st, (const YT*)Y, (XT*)X, g);
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
