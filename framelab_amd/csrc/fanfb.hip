// fanfb.hip -- finite-width-beam fan projector pair (reference version 2).
//
// Each detector cell is a strip between the source rays through its two
// edges; a pixel's weight is the area of the pixel inside the strip
// (find_area.cpp:18-69).  Forward: one thread per ray strip, walking the
// columns of the strip's major axis (Ax_fan_mf_cpu_new_fb.cpp:19-123).
// Adjoint: one thread per pixel gathering, view by view, the cells of its
// window (Atx_fan_mf_cpu_new_fb.cpp:21-66).  Float operations follow the
// reference CPU code in order, so both are bit-identical to it; the strip
// lines' trig values (sinf/cosf/atan2f) are tabulated on the host, where
// the reference evaluates them.
#include <hip/hip_runtime.h>

#include "fl_capi_internal.h"
#include "fl_common.h"
#include "fl_geom.h"

namespace {

__device__ __forceinline__ float fb_abs(float v) { return v > 0 ? v : -v; }

// Area of the unit pixel centred (x, y) on the far side of the line
// a*X + b*Y = c from the centre (at most 1/2), and the side of the centre.
__device__ __forceinline__ float fb_half(float a, float b, float c, float x, float y, int& sgn)
{
    const float aa = fb_abs(a), bb = fb_abs(b);
    const float hx = aa / 2, hy = bb / 2, dmax = hx + hy;
    const float t = c - a * x - b * y;
    const float d = fb_abs(t);
    sgn = t > 0 ? 1 : -1;
    if (!(d < dmax)) return 0.5f;
    const float d0 = fb_abs(hx - hy);
    const float inv = hx < hy ? 1.0f / bb : 1.0f / aa;
    if (d <= d0) return d * inv;
    return 0.5f - 0.5f * (dmax - d) * (dmax - d) / (aa * bb);
}

// Area of the pixel inside the strip between two lines.
__device__ __forceinline__ float fb_area(float a1, float b1, float c1, float a2, float b2, float c2, float x, float y)
{
    int s1, s2;
    const float A1 = fb_half(a1, b1, c1, x, y, s1);
    const float A2 = fb_half(a2, b2, c2, x, y, s2);
    return fb_abs(A1 - (float)(s1 * s2) * A2);
}

// Pixel range of the strip over one unit column [u1, u1+1] of the major
// axis: the four edge-ray crossings, kept within [0, n-1], starting from
// (n-1, 0).
__device__ __forceinline__ void fb_range(float o, float om, float s0, float s1, float u1, int n, int& lo, int& hi)
{
    const int n2 = n / 2;
    const float u2 = u1 + 1;
    const int v0 = (int)floorf(o + s0 * (u1 - om) + (float)n2), v1 = (int)floorf(o + s0 * (u2 - om) + (float)n2);
    const int v2 = (int)floorf(o + s1 * (u1 - om) + (float)n2), v3 = (int)floorf(o + s1 * (u2 - om) + (float)n2);
    lo = n - 1;
    hi = 0;
    if (v0 < lo && v0 >= 0) lo = v0;
    if (v0 > hi && v0 <= n - 1) hi = v0;
    if (v1 < lo && v1 >= 0) lo = v1;
    if (v1 > hi && v1 <= n - 1) hi = v1;
    if (v2 < lo && v2 >= 0) lo = v2;
    if (v2 > hi && v2 <= n - 1) hi = v2;
    if (v3 < lo && v3 >= 0) lo = v3;
    if (v3 > hi && v3 <= n - 1) hi = v3;
}

// Strip lines per (view, cell, edge): Px, Py at [(iv*nd + id)*2 + e], and
// Pd at [id*2 + e].
__global__ __launch_bounds__(64) void k_fwd_fan_fb(const float* __restrict__ X, float* __restrict__ Y, FwdGeom g,
                                                  const float* __restrict__ lpx, const float* __restrict__ lpy,
                                                  const float* __restrict__ lpd, float dy_det)
{
    const uint32_t tiles = (g.na + 63) / 64;
    const uint32_t b = fl_xcd_remap(blockIdx.x, gridDim.x);
    const int slot = (int)(b / tiles);
    const int id = (int)(b % tiles) * 64 + (int)threadIdx.x;
    if (slot >= g.nslots || id >= g.na) return;
    const int iv = g.slot_view[slot];
    const int nx = g.nx, ny = g.ny, nx2 = nx / 2, ny2 = ny / 2, nd = g.na;
    const float* __restrict__ x = X + (size_t)g.view_frame[iv] * g.nvox;
    const float c = g.cs[iv], s = g.sn[iv];
    const float x1 = c * (-g.SO), y1 = s * (-g.SO);
    float ex[2], ey[2], Px[2], Py[2], Pd[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
        const float u = e == 0 ? g.ydet[id] - dy_det / 2 : g.ydet[id] + dy_det / 2;
        ex[e] = c * g.OD - s * u;
        ey[e] = s * g.OD + c * u;
        Px[e] = lpx[((size_t)iv * nd + id) * 2 + e];
        Py[e] = lpy[((size_t)iv * nd + id) * 2 + e];
        Pd[e] = lpd[id * 2 + e];
    }
    float acc = 0.0f;
    int lo = 0, hi = -1;
    const float mx = (ex[0] + ex[1]) / 2, my = (ey[0] + ey[1]) / 2;
    if (fb_abs(x1 - mx) > fb_abs(y1 - my)) {
        const bool flat = ex[0] == x1 || ex[1] == x1;
        float s0 = 0.0f, s1 = 0.0f;
        if (!flat) { s0 = (ey[0] - y1) / (ex[0] - x1); s1 = (ey[1] - y1) / (ex[1] - x1); }
        for (int ix = 0; ix < nx; ix++) {
            const float xc = (float)ix + 0.5f - (float)nx2;
            if (flat) {
                if (fb_abs(x1 - xc) <= 0.5f) { lo = 0; hi = ny - 1; }
            } else {
                fb_range(y1, x1, s0, s1, (float)(ix - nx2), ny, lo, hi);
            }
            for (int iy = lo; iy <= hi; iy++) {
                const float yc = (float)iy + 0.5f - (float)ny2;
                const float l = fb_area(Px[0], Py[0], Pd[0], Px[1], Py[1], Pd[1], xc, yc);
                if (l > 0) acc = acc + l * x[(size_t)iy * nx + ix];
            }
        }
    } else {
        const bool flat = ey[0] == y1 || ey[1] == y1;
        float s0 = 0.0f, s1 = 0.0f;
        if (!flat) { s0 = (ex[0] - x1) / (ey[0] - y1); s1 = (ex[1] - x1) / (ey[1] - y1); }
        for (int iy = 0; iy < ny; iy++) {
            const float yc = (float)iy + 0.5f - (float)ny2;
            if (flat) {
                if (fb_abs(y1 - yc) <= 0.5f) { lo = 0; hi = nx - 1; }
            } else {
                fb_range(x1, y1, s0, s1, (float)(iy - ny2), nx, lo, hi);
            }
            for (int ix = lo; ix <= hi; ix++) {
                const float xc = (float)ix + 0.5f - (float)nx2;
                const float l = fb_area(Px[0], Py[0], Pd[0], Px[1], Py[1], Pd[1], xc, yc);
                if (l > 0) acc = acc + l * x[(size_t)iy * nx + ix];
            }
        }
    }
    Y[(size_t)slot * nd + id] = acc * g.scale;
}

__global__ __launch_bounds__(256) void k_adj_fan_fb(const float* __restrict__ Y, float* __restrict__ X, AdjGeom g,
                                                   const float* __restrict__ cd2, const float* __restrict__ sd2)
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
        const float* __restrict__ y = Y + (size_t)g.id_S[it * g.tmp_size + j] * nd;
        const float c = g.cphi[iv], s = g.sphi[iv];
        const float xr = c * xc + s * yc;
        const float yr = -s * xc + c * yc;
        const float t = SD / ((xr + g.SO) * g.dy_det);
        int hi = (int)floorf((yr + 1) * t + (float)nd2);
        int lo = (int)floorf((yr - 1) * t + (float)nd2);
        if (lo < 0) lo = 0;
        if (hi > nd - 1) hi = nd - 1;
        for (int id = lo; id <= hi; id++) {
            const float Px0 = -(sd2[id] * c + cd2[id] * s);
            const float Py0 = cd2[id] * c - sd2[id] * s;
            const float Pd0 = g.SO * sd2[id];
            const float Px1 = -(sd2[id + 1] * c + cd2[id + 1] * s);
            const float Py1 = cd2[id + 1] * c - sd2[id + 1] * s;
            const float Pd1 = g.SO * sd2[id + 1];
            const float l = fb_area(Px0, Py0, Pd0, Px1, Py1, Pd1, xc, yc);
            if (l > 0) acc = acc + l * y[id];
        }
    }
    X[(size_t)it * g.nvox + (size_t)iy * g.nx + ix] = acc * g.scale;
}

}  // namespace

hipError_t fl_launch_fwd_fan_fb(const float* X, float* Y, const fl_plan_s* p, hipStream_t st)
{
    const FwdGeom& g = p->fg;
    const uint32_t grid = ((g.na + 63) / 64) * (uint32_t)g.nslots;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fwd_fan_fb, dim3(grid), dim3(64), 0, st, X, Y, g, p->d_fb_px, p->d_fb_py, p->d_fb_pd,
                       p->ag.dy_det);
    return hipGetLastError();
}

hipError_t fl_launch_adj_fan_fb(const float* Y, float* X, const fl_plan_s* p, hipStream_t st)
{
    const AdjGeom& g = p->ag;
    dim3 grid((g.nx + 15) / 16, (g.ny + 15) / 16, g.nt);
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_fan_fb, grid, dim3(16, 16), 0, st, Y, X, g, p->d_cdet2, p->d_sdet2);
    return hipGetLastError();
}
