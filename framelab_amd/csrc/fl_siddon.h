// fl_siddon.h -- incremental Siddon ray walk with the reference's exact
// traversal semantics (device code).
//
// Reference semantics (matlab/mex/src/Ax_cone_mf_cpu_siddon.cpp:19-134,
// Ax_fan_mf_cpu_siddon.cpp:19-93, sort_alpha.cpp:15-60):
//   * plane parameters alpha = (i*s - o)/d for planes i = -n/2..n/2 of each
//     axis with d != 0 (s = 1 for x and y, s = dz0 for z), in float;
//   * the sorted union of all axes' parameters (an exact cross-axis tie is
//     one element) defines the segments;
//   * a segment belongs to voxel floor(o + m*d + n/2) per axis, m the float
//     midpoint; out-of-volume segments are skipped;
//   * the ray sum accumulates fl(fl(len)*x) in traversal order, in float.
//
// Instead of materialising and merging the full parameter lists per ray,
// the walk streams the union with a 3-way minimum and starts at the volume
// entry.  Skipping the pre-entry and post-exit segments is exact: the
// computed voxel coordinate along each axis is a monotone function of the
// midpoint (a chain of correctly rounded monotone operations), so once an
// axis index is beyond the volume on the far side it stays there, and every
// segment before one whose index is on the near side is also outside.  A
// short backward scan from the entry plane recovers the rare segments that
// rounding places inside the volume before the mathematical entry.
#pragma once

#include "fl_common.h"

#ifndef FL_HD
#define FL_HD __host__ __device__ __forceinline__
#endif

// One axis of the walk.  Planes are visited in ascending-alpha order:
// plane j (0..n) sits at coordinate i = i0 + st*j.
struct SidAxis {
    float o;    // ray origin along the axis (source coordinate)
    float d;    // direction component (detector - source)
    float s;    // plane spacing multiplier (1 or dz0)
    int i0;     // first plane index in walk order (-n/2 or n/2)
    int st;     // +1 or -1
    int n;      // number of cells; planes j = 0..n
    int nh;     // n/2, the index offset of the cell coordinate
    bool on;    // d != 0
};

FL_HD void sid_axis_init(SidAxis& a, float o, float d, float s, int n)
{
    a.o = o; a.d = d; a.s = s; a.n = n; a.nh = n / 2;
    a.on = (d != 0.0f);
    a.st = d > 0.0f ? 1 : -1;
    a.i0 = d > 0.0f ? -a.nh : a.nh;
}

// Exact reference plane parameter: ((float)i * s - o) / d, IEEE division.
FL_HD float sid_alpha(const SidAxis& a, int j)
{
    float p = (float)(a.i0 + a.st * j) * a.s;
    return (p - a.o) / a.d;
}

// Number of planes with alpha <= t (strict=false) or alpha < t (strict=true).
FL_HD int sid_count(const SidAxis& a, float t, bool strict)
{
    float c = a.o + t * a.d;                         // coordinate at parameter t
    float est = (c / a.s - (float)a.i0) * (float)a.st;
    int j;
    if (!(est > -1.0f)) j = 0;
    else if (!(est < (float)a.n + 1.0f)) j = a.n + 1;
    else j = (int)floorf(est) + 1;
    if (j < 0) j = 0;
    if (j > a.n + 1) j = a.n + 1;
    // correct the estimate with exact parameters
    while (j > 0) {
        float v = sid_alpha(a, j - 1);
        if (strict ? (v >= t) : (v > t)) j--; else break;
    }
    while (j <= a.n) {
        float v = sid_alpha(a, j);
        if (strict ? (v < t) : (v <= t)) j++; else break;
    }
    return j;
}

// num / d, correctly rounded, from r = 1/d correctly rounded: q = num*r is
// within an ulp of the quotient, the fused residual num - q*d is exact, and
// q + residual*r rounds to the IEEE quotient (Markstein's theorem; holds
// without overflow or underflow, which plane parameters never reach).
FL_HD float sid_div(float num, float d, float r)
{
    const float q = num * r;
    const float e = fmaf(-q, d, num);
    return fmaf(e, r, q);
}

// Cell index of the midpoint along an x/y axis: floor(o + m*d + n/2).
FL_HD int sid_cell_xy(const SidAxis& a, float m)
{
    float t = a.o + m * a.d;
    return (int)floorf(t + (float)a.nh);
}
// z axis: floor((o + m*d)/dz0 + n/2); dz0 == 1 divides exactly, so it is skipped.
FL_HD int sid_cell_z(const SidAxis& a, float m, bool dz_one)
{
    float t = a.o + m * a.d;
    if (!dz_one) t = t / a.s;
    return (int)floorf(t + (float)a.nh);
}

FL_HD bool sid_in(const SidAxis& a, int t) { return (unsigned)t < (unsigned)a.n; }
// beyond the volume on the side the ray has not reached yet / has left
FL_HD bool sid_near(const SidAxis& a, int t)
{
    return a.on && (a.st > 0 ? t < 0 : t >= a.n);
}
FL_HD bool sid_far(const SidAxis& a, int t)
{
    return a.on && (a.st > 0 ? t >= a.n : t < 0);
}

// Walk state of one axis: parameter of the next plane (INF for an axis
// without planes or already past its last one at the start), that plane's
// coordinate index as a float (exact integer), and the constants of the
// branch-free advance.
struct SidNext {
    float a;     // alpha of the next plane
    float p;     // its index i (float)
    float st;    // +1 / -1
    float r;     // 1/d, correctly rounded
};

FL_HD void sid_next_init(SidNext& n, const SidAxis& ax, int j)
{
    n.st = (float)ax.st;
    n.p = (float)(ax.i0 + ax.st * j);
    n.r = ax.on ? 1.0f / ax.d : 0.0f;
    n.a = (ax.on && j <= ax.n) ? sid_alpha(ax, j) : FL_INF;
}

// If this axis's next plane is the one just reached (alpha == an), move to
// the following plane.  Branch-free; exact ties advance every tied axis.
// Past an axis's last plane the advance goes on to the plane one cell
// beyond it instead of ending that axis: the reference's list has no such
// plane, but every segment after the last plane lies outside the volume on
// that axis, and the walk stops at the first of them (its midpoint is at
// least half a cell beyond the boundary once it reaches the extra plane),
// so the visited segments are the same.
template <bool S1>
FL_HD void sid_next_step(SidNext& n, const SidAxis& ax, float an)
{
    const float pn = n.p + n.st;
    const float num = (S1 ? pn : pn * ax.s) - ax.o;
    const float v = sid_div(num, ax.d, n.r);
    const bool adv = n.a == an;
    n.a = adv ? v : n.a;
    n.p = adv ? pn : n.p;
}

// Generic walk over one ray.  NZ3 = true for the cone (3 axes); DZ1 = the
// z plane spacing is 1 (cell index needs no division).  The visitor
// receives (linear voxel index (z*ny+y)*nx+x, alpha difference of the
// segment) for every in-volume segment in traversal order; the length is
// that difference times |d|.
template <bool NZ3, bool DZ1, bool XYZ, bool ALL = false, class Visit>
FL_HD void sid_walk_t(const SidAxis& ax, const SidAxis& ay, const SidAxis& az, float L, Visit&& visit)
{
    constexpr bool dz_one = DZ1;
    // Inactive axes keep a constant cell index; if it is outside, nothing
    // of the ray is inside the volume.
    if (!ax.on && !sid_in(ax, sid_cell_xy(ax, 0.5f))) return;
    if (!ay.on && !sid_in(ay, sid_cell_xy(ay, 0.5f))) return;
    if (NZ3 && !az.on && !sid_in(az, sid_cell_z(az, 0.5f, dz_one))) return;

    // Entry parameter: the largest first-plane parameter of the active axes.
    float ae = -FL_INF;
    if (ax.on) ae = fmaxf(ae, sid_alpha(ax, 0));
    if (ay.on) ae = fmaxf(ae, sid_alpha(ay, 0));
    if (NZ3 && az.on) ae = fmaxf(ae, sid_alpha(az, 0));

    // Backward scan: move the start left while the preceding segment is not
    // provably before the volume.
    float cur = ae;
    for (;;) {
        float pred = -FL_INF;
        if (ax.on) { int c = sid_count(ax, cur, true); if (c > 0) pred = fmaxf(pred, sid_alpha(ax, c - 1)); }
        if (ay.on) { int c = sid_count(ay, cur, true); if (c > 0) pred = fmaxf(pred, sid_alpha(ay, c - 1)); }
        if (NZ3 && az.on) { int c = sid_count(az, cur, true); if (c > 0) pred = fmaxf(pred, sid_alpha(az, c - 1)); }
        if (pred == -FL_INF) break;
        float m = (pred + cur) * 0.5f;
        int tx = sid_cell_xy(ax, m), ty = sid_cell_xy(ay, m);
        bool near = sid_near(ax, tx) || sid_near(ay, ty);
        if (NZ3) near = near || sid_near(az, sid_cell_z(az, m, dz_one));
        if (near) break;
        cur = pred;
    }

    // Forward walk from `cur`.
    SidNext qx, qy, qz;
    sid_next_init(qx, ax, ax.on ? sid_count(ax, cur, false) : 0);
    sid_next_init(qy, ay, ay.on ? sid_count(ay, cur, false) : 0);
    if (NZ3) sid_next_init(qz, az, az.on ? sid_count(az, cur, false) : 0);
    else qz.a = FL_INF;
    const float rs = NZ3 && !DZ1 ? 1.0f / az.s : 1.0f;

    const int nx = ax.n, ny = ay.n;
    for (;;) {
        float an = fminf(qx.a, qy.a);
        if (NZ3) an = fminf(an, qz.a);
        if (an == FL_INF) break;
        float m = (cur + an) * 0.5f;
        int tx = sid_cell_xy(ax, m);
        int ty = sid_cell_xy(ay, m);
        int tz = 0;
        if (NZ3) {
            float t = az.o + m * az.d;
            if (!DZ1) t = sid_div(t, az.s, rs);
            tz = (int)floorf(t + (float)az.nh);
        }
        bool in = sid_in(ax, tx) && sid_in(ay, ty) && (!NZ3 || sid_in(az, tz));
        if constexpr (ALL) {
            // every step, with its in-volume flag: visit(...) issues the
            // step's work and returns a token that visit.commit consumes
            // after this step's advance, so what the visit waits for (a
            // gather) has the advance's arithmetic to complete behind
            const auto tok = visit(tx, ty, tz, an - cur, in);
            const bool stop = !in && (sid_far(ax, tx) || sid_far(ay, ty) || (NZ3 && sid_far(az, tz)));
            sid_next_step<true>(qx, ax, an);
            sid_next_step<true>(qy, ay, an);
            if (NZ3) sid_next_step<DZ1>(qz, az, an);
            cur = an;
            visit.commit(tok);
            if (stop) break;
            continue;
        } else if (in) {
            if constexpr (XYZ) visit(tx, ty, tz, an - cur);
            else visit((tz * ny + ty) * nx + tx, an - cur);
        } else {
            bool far = sid_far(ax, tx) || sid_far(ay, ty) || (NZ3 && sid_far(az, tz));
            if (far) break;
        }
        sid_next_step<true>(qx, ax, an);
        sid_next_step<true>(qy, ay, an);
        if (NZ3) sid_next_step<DZ1>(qz, az, an);
        cur = an;
    }
}

template <bool NZ3, class Visit>
FL_HD void sid_walk(const SidAxis& ax, const SidAxis& ay, const SidAxis& az, float L, bool dz_one, Visit&& visit)
{
    if (dz_one) sid_walk_t<NZ3, true, false>(ax, ay, az, L, visit);
    else sid_walk_t<NZ3, false, false>(ax, ay, az, L, visit);
}

// Every step of the walk: tok = visit(tx, ty, tz, dal, in) for the
// segments before and inside the volume (in = the cell is inside), up to
// the first one beyond it, and visit.commit(tok) once the step's plane
// advance is issued.
template <class Visit>
FL_HD void sid_walk_all3(const SidAxis& ax, const SidAxis& ay, const SidAxis& az, float L, bool dz_one,
                         Visit&& visit)
{
    if (dz_one) sid_walk_t<true, true, true, true>(ax, ay, az, L, visit);
    else sid_walk_t<true, false, true, true>(ax, ay, az, L, visit);
}

// The same for a 2-D walk (tz = 0).
template <class Visit>
FL_HD void sid_walk_all2(const SidAxis& ax, const SidAxis& ay, const SidAxis& az, float L, Visit&& visit)
{
    sid_walk_t<false, true, true, true>(ax, ay, az, L, visit);
}

// As sid_walk, but the visitor receives the cell coordinates (tx, ty, tz)
// instead of the linear index.
template <bool NZ3, class Visit>
FL_HD void sid_walk_xyz(const SidAxis& ax, const SidAxis& ay, const SidAxis& az, float L, bool dz_one,
                        Visit&& visit)
{
    if (dz_one) sid_walk_t<NZ3, true, true>(ax, ay, az, L, visit);
    else sid_walk_t<NZ3, false, true>(ax, ay, az, L, visit);
}
