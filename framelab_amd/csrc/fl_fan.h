// fl_fan.h -- per-ray and per-pixel bodies of the 2D fan operators, shared by
// the single-problem kernels and the batched persistent kernels so both run
// exactly the same float arithmetic.
#pragma once

#include "fl_chord.h"
#include "fl_common.h"
#include "fl_geom.h"
#include "fl_siddon.h"

// Siddon ray sum of fan ray (view iv, cell id), scaled
// (Ax_fan_mf_cpu_siddon.cpp:32-91).
template <typename VT>
__device__ __forceinline__ float fl_fan_ray_sum(const VT* __restrict__ x, const FwdGeom& g, int iv, int id)
{
    const float c = g.cs[iv], s = g.sn[iv];
    const float x1 = c * (-g.SO), y1 = s * (-g.SO);
    const float yd = g.ydet[id];
    const float x2 = c * g.OD - s * yd;
    const float y2 = s * g.OD + c * yd;
    const float dx = x2 - x1, dy = y2 - y1;
    const float L = sqrtf(dx * dx + dy * dy);
    SidAxis ax, ay, az;
    sid_axis_init(ax, x1, dx, 1.0f, g.nx);
    sid_axis_init(ay, y1, dy, 1.0f, g.ny);
    sid_axis_init(az, 0.0f, 0.0f, 1.0f, 2);
    // y = y + S*x in walk order, each step's gather issued before the
    // step's plane advance and its term added after it (outside the image
    // from pixel 0, with the term +0, which leaves the sum bit-identical)
    struct Term {
        float S;
        VT x;
        bool in;
    };
    struct Visit {
        const VT* __restrict__ x;
        int nx;
        float L;
        float acc;
        __device__ Term operator()(int tx, int ty, int, float dal, bool in) const
        {
            return Term{dal * L, x[in ? ty * nx + tx : 0], in};
        }
        __device__ void commit(const Term& t) { acc = acc + (t.in ? t.S * fl_load(&t.x, 0) : 0.0f); }
    } visit{x, g.nx, L, 0.0f};
    sid_walk_all2(ax, ay, az, L, visit);
    return visit.acc * g.scale;
}

// Gao column walk of fan ray (iv, id), scaled (Ax_fan_mf_cpu_new.cpp:28-119).
// The y-major branch keeps the reference's `c2 == ny` test (cpu :111).
template <typename VT>
__device__ __forceinline__ float fl_fan_gao_sum(const VT* __restrict__ x, const FwdGeom& g, int iv, int id)
{
    const int nx = g.nx, ny = g.ny, nx2 = nx / 2, ny2 = ny / 2;
    const float c = g.cs[iv], s = g.sn[iv];
    const float x1 = c * (-g.SO), y1 = s * (-g.SO);
    const float yd = g.ydet[id];
    const float x2 = c * g.OD - s * yd;
    const float y2 = s * g.OD + c * yd;
    float acc = 0.0f;
    if (fl_absf(x1 - x2) > fl_absf(y1 - y2)) {
        const float slope = (y2 - y1) / (x2 - x1);
        for (int ix = 0; ix < nx; ix++) {
            const float u1 = (float)(ix - nx2), u2 = u1 + 1;
            float v1, v2;
            if (slope >= 0) { v1 = y1 + slope * (u1 - x1) + (float)ny2; v2 = y1 + slope * (u2 - x1) + (float)ny2; }
            else            { v1 = y1 + slope * (u2 - x1) + (float)ny2; v2 = y1 + slope * (u1 - x1) + (float)ny2; }
            const int c1 = (int)floorf(v1), c2 = (int)floorf(v2);
            if (c2 == c1) {
                if (c1 >= 0 && c1 <= ny - 1) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + l * fl_load(x, (size_t)c1 * nx + ix);
                }
            } else if (c2 > 0 && c2 < ny) {
                // c2 - c1 == 1 whenever |slope| < 1; a float slope that
                // rounds to 1 can give c1 == c2 - 2 = -1, which is skipped
                float d = v2 - v1, l = sqrtf(d * d + 1);
                if (c1 >= 0) acc = acc + (((float)c2 - v1) / d) * l * fl_load(x, (size_t)c1 * nx + ix);
                acc = acc + ((v2 - (float)c2) / d) * l * fl_load(x, (size_t)c2 * nx + ix);
            } else {
                if (c2 == 0) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + ((v2 - (float)c2) / d) * l * fl_load(x, (size_t)c2 * nx + ix);
                }
                if (c2 == ny) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + (((float)c2 - v1) / d) * l * fl_load(x, (size_t)c1 * nx + ix);
                }
            }
        }
    } else {
        const float slope = (x2 - x1) / (y2 - y1);
        for (int iy = 0; iy < ny; iy++) {
            const float u1 = (float)(iy - ny2), u2 = u1 + 1;
            float v1, v2;
            if (slope >= 0) { v1 = x1 + slope * (u1 - y1) + (float)nx2; v2 = x1 + slope * (u2 - y1) + (float)nx2; }
            else            { v1 = x1 + slope * (u2 - y1) + (float)nx2; v2 = x1 + slope * (u1 - y1) + (float)nx2; }
            const int c1 = (int)floorf(v1), c2 = (int)floorf(v2);
            if (c2 == c1) {
                if (c1 >= 0 && c1 <= nx - 1) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + l * fl_load(x, (size_t)iy * nx + c1);
                }
            } else if (c2 > 0 && c2 < nx) {
                float d = v2 - v1, l = sqrtf(d * d + 1);
                if (c1 >= 0) acc = acc + (((float)c2 - v1) / d) * l * fl_load(x, (size_t)iy * nx + c1);
                acc = acc + ((v2 - (float)c2) / d) * l * fl_load(x, (size_t)iy * nx + c2);
            } else {
                if (c2 == 0) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + ((v2 - (float)c2) / d) * l * fl_load(x, (size_t)iy * nx + c2);
                }
                // the reference tests ny here; with nx != ny only cells
                // c1 = ny-1 < nx are reachable, so the index stays in range
                if (c2 == ny && c1 >= 0 && c1 < nx) {
                    float d = v2 - v1, l = sqrtf(d * d + 1);
                    acc = acc + (((float)c2 - v1) / d) * l * fl_load(x, (size_t)iy * nx + c1);
                }
            }
        }
    }
    return acc * g.scale;
}

// Exact pixel-driven adjoint of pixel (ix, iy) of frame it, scaled
// (Atx_fan_mf_cpu_new.cpp:35-62).
// LQ: take the lines' pixel-independent quotients from g.line_q (one table
// read per cell instead of up to three divisions); the batched kernel,
// whose small persistent workgroups hide the extra read worse, divides.
template <typename YT, bool LQ = true>
__device__ __forceinline__ float fl_fan_adj_pixel(const YT* __restrict__ Y, const AdjGeom& g, int it, int ix, int iy)
{
    const int nd = g.na, nd2 = nd / 2;
    const float SD = g.SO + g.OD;
    const float xc = (float)ix + 0.5f - (float)(g.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(g.ny / 2);
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)g.id_S[it * g.tmp_size + j] * nd;
        const float c = g.cphi[iv], s = g.sphi[iv];
        const float xr = c * xc + s * yc;
        const float yr = -s * xc + c * yc;
        const float t = SD / ((xr + g.SO) * g.dy_det);
        int hi = (int)floorf((yr + 1) * t - g.y_os + (float)nd2);
        int lo = (int)floorf((yr - 1) * t - g.y_os + (float)nd2);
        if (lo < 0) lo = 0;
        if (hi > nd - 1) hi = nd - 1;
        const float2* __restrict__ lq = LQ ? g.line_q + (size_t)iv * nd : nullptr;
        for (int id = lo; id <= hi; id++) {
            const float sd = g.sdet[id], cd = g.cdet[id];
            const float Px = -(sd * c + cd * s);
            const float Py = cd * c - sd * s;
            const float Pd = g.SO * sd;
            const float l = LQ ? fl_chord_line2_q(Px, Py, Pd, xc, yc, lq[id]) : fl_chord_line2(Px, Py, Pd, xc, yc);
            // a select, not a branch around the load: lanes of a wave
            // disagree on l > 0 too often for a branch to pay
            const float yv = fl_load(y, (size_t)id);
            acc = l > 0 ? acc + l * yv : acc;
        }
    }
    return acc * g.scale;
}

struct FlAxisIv {
    float in, out;   // entry / exit parameters of a voxel slab along one axis
};

// Entry/exit plane parameters of cell `cell` along one walk axis, with the
// forward's plane formula ((float)i*s - o)/d.
__device__ __forceinline__ FlAxisIv fl_slab(const SidAxis& a, int cell)
{
    FlAxisIv r;
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

// As fl_slab, given rd = 1.0f / a.d: both quotients from the one reciprocal
// (fl_div_r, bit-identical to the divisions).
__device__ __forceinline__ FlAxisIv fl_slab_r(const SidAxis& a, int cell, float rd)
{
    FlAxisIv r;
    if (!a.on) {
        r.in = -FL_INF;
        r.out = FL_INF;
        return r;
    }
    const float lo = fl_div_r((float)(cell - a.nh) * a.s - a.o, a.d, rd);
    const float hi = fl_div_r((float)(cell - a.nh + 1) * a.s - a.o, a.d, rd);
    r.in = a.st > 0 ? lo : hi;
    r.out = a.st > 0 ? hi : lo;
    return r;
}

// Matched (Siddon-transpose) adjoint of pixel (ix, iy) of frame it, scaled.
// RQ: take each ray's reciprocals and length from g.ray_q instead of
// forming them per pixel (same values, the same IEEE operations made once).
template <typename YT, bool RQ = true>
__device__ __forceinline__ float fl_fan_matched_pixel(const YT* __restrict__ Y, const FwdGeom& f, const AdjGeom& g,
                                                      int it, int ix, int iy)
{
    const int nd = f.na;
    const float SD = f.SO + f.OD;
    const float xc = (float)ix + 0.5f - (float)(f.nx / 2);
    const float yc = (float)iy + 0.5f - (float)(f.ny / 2);
    const float ya0 = f.ydet[0];
    float acc = 0.0f;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int iv = g.id_Y[it * g.tmp_size + j];
        const YT* __restrict__ y = Y + (size_t)g.id_S[it * g.tmp_size + j] * nd;
        const float c = f.cs[iv], s = f.sn[iv];
        const float xr = c * xc + s * yc, yr = -s * xc + c * yc;
        const float mag = SD / (xr + f.SO);
        int lo = (int)floorf(((yr - 1.0f) * mag - ya0) * f.inv_pa) - 1;
        int hi = (int)ceilf(((yr + 1.0f) * mag - ya0) * f.inv_pa) + 1;
        if (lo > hi) { int t = lo; lo = hi; hi = t; }
        if (lo < 0) lo = 0;
        if (hi > nd - 1) hi = nd - 1;
        const float x1 = c * (-f.SO), y1 = s * (-f.SO);
        const float4* __restrict__ rq = RQ ? g.ray_q + (size_t)iv * nd : nullptr;
        for (int id = lo; id <= hi; id++) {
            const float yd = f.ydet[id];
            const float x2 = c * f.OD - s * yd;
            const float y2 = s * f.OD + c * yd;
            const float dx = x2 - x1, dy = y2 - y1;
            const float4 q = RQ ? rq[id] : make_float4(1.0f / dx, 1.0f / dy, 0.0f, 0.0f);
            SidAxis ax, ay;
            sid_axis_init(ax, x1, dx, 1.0f, f.nx);
            sid_axis_init(ay, y1, dy, 1.0f, f.ny);
            const FlAxisIv sx = fl_slab_r(ax, ix, q.x), sy = fl_slab_r(ay, iy, q.y);
            const float ain = fmaxf(sx.in, sy.in), aout = fminf(sx.out, sy.out);
            if (!(aout > ain)) continue;
            const float m = (ain + aout) * 0.5f;
            if (sid_cell_xy(ax, m) != ix || sid_cell_xy(ay, m) != iy) continue;
            const float L = RQ ? q.z : sqrtf(dx * dx + dy * dy);
            acc = acc + ((aout - ain) * L) * fl_load(y, (size_t)id);
        }
    }
    return acc * f.scale;
}
