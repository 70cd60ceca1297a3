// fl_chord.h -- exact ray/pixel and ray/voxel chord lengths (host+device).
//
// Same float arithmetic as the reference chord routines, so the voxel-driven
// adjoints are bit-identical to the reference's:
//   fl_chord_line2  <- matlab/mex/src/find_l.cpp:18-45 (line a*x+b*y=c through
//                      the unit pixel at (x,y));
//   fl_chord_seg2   <- matlab/mex/src/find_l_3d.cpp:64-137 (segment-line
//                      through a dx-by-dy cell);
//   fl_chord_box3   <- matlab/mex/src/find_l_3d.cpp:139-227 (line through a
//                      dx-by-dy-by-dz box; faces tried x,y,z or y,x,z).
#pragma once

#include "fl_common.h"

#ifndef FL_HD
#define FL_HD __host__ __device__ __forceinline__
#endif

FL_HD float fl_absf(float v) { return v > 0 ? v : -v; }

FL_HD float fl_chord_line2(float a, float b, float c, float x, float y)
{
    float a2 = fl_absf(a), b2 = fl_absf(b);
    float hx = a2 / 2, hy = b2 / 2;
    float dmax = hx + hy;
    float t = c - a * x - b * y;
    float d = fl_absf(t);
    float l = 0.0f;
    if (d < dmax) {
        float d0 = fl_absf(hx - hy);
        float flat = (hx < hy) ? 1.0f / b2 : 1.0f / a2;
        l = (d <= d0) ? flat : (dmax - d) / (a2 * b2);
    }
    return l;
}

FL_HD float fl_chord_seg2(float x1, float y1, float x2, float y2, float dx, float dy, float x, float y)
{
    float a = x2 - x1, b = y2 - y1;
    float hx = dx / 2, hy = dy / 2;
    if (a == 0) return (fl_absf(x1 - x) <= hx) ? dy : 0.0f;
    if (b == 0) return (fl_absf(y1 - y) <= hy) ? dx : 0.0f;
    float px = x1 - x, py = y1 - y;
    float xi0 = 0, yi0 = 0, xi1 = 0, yi1 = 0;
    int k = 0;
    auto put = [&](float u, float v) {
        if (k == 0) { xi0 = u; yi0 = v; } else { xi1 = u; yi1 = v; }
        k++;
    };
    const bool xfirst = fl_absf(a) > fl_absf(b);
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && k >= 2) break;
        const bool xf = (pass == 0) ? xfirst : !xfirst;
        if (xf) {
            float sl = b / a, m = sl * (-px) + py, w = sl * hx;
            if (fl_absf(m - w) <= hy) put(-hx, m - w);
            if ((pass == 0 || k < 2) && fl_absf(m + w) <= hy) put(hx, m + w);
        } else {
            float sl = a / b, m = sl * (-py) + px, w = sl * hy;
            if (fl_absf(m - w) <= hx) put(m - w, -hy);
            if ((pass == 0 || k < 2) && fl_absf(m + w) <= hx) put(m + w, hy);
        }
    }
    if (k == 2) {
        float u = xi1 - xi0, v = yi1 - yi0;
        return sqrtf(u * u + v * v);
    }
    return 0.0f;
}

// Intersection points collected by the face tests of fl_chord_box3.
struct FlPts {
    float p0[3], p1[3];
    int k;
};

// Faces normal to axis AX at -h and +h.  p: line point relative to the box
// centre, d: direction, h: half sides.  `later` marks the 2nd/3rd face pair,
// where the + face is tried only while fewer than two points were found.
template <int AX>
FL_HD void fl_face_pair(const float* p, const float* d, const float* h, bool later, FlPts& q)
{
    constexpr int U = (AX == 0) ? 1 : 0;
    constexpr int V = (AX == 2) ? 1 : 2;
    float su = d[U] / d[AX], sv = d[V] / d[AX];
    float mu = su * (-p[AX]) + p[U], wu = su * h[AX];
    float mv = sv * (-p[AX]) + p[V], wv = sv * h[AX];
    if (fl_absf(mu - wu) <= h[U] && fl_absf(mv - wv) <= h[V]) {
        float* o = q.k == 0 ? q.p0 : q.p1;
        o[AX] = -h[AX]; o[U] = mu - wu; o[V] = mv - wv;
        q.k++;
    }
    if (!later || q.k < 2) {
        if (fl_absf(mu + wu) <= h[U] && fl_absf(mv + wv) <= h[V]) {
            float* o = q.k == 0 ? q.p0 : q.p1;
            o[AX] = h[AX]; o[U] = mu + wu; o[V] = mv + wv;
            q.k++;
        }
    }
}

FL_HD float fl_chord_box3(float x1, float y1, float z1, float x2, float y2, float z2, float dx,
                          float dy, float dz, float x, float y, float z)
{
    float a = x2 - x1, b = y2 - y1, c = z2 - z1;
    if (a == 0) return fl_chord_seg2(y1, z1, y2, z2, dy, dz, y, z);
    if (b == 0) return fl_chord_seg2(x1, z1, x2, z2, dx, dz, x, z);
    const float p[3] = {x1 - x, y1 - y, z1 - z};
    const float d[3] = {a, b, c};
    const float h[3] = {dx / 2, dy / 2, dz / 2};
    FlPts q;
    q.k = 0;
    if (fl_absf(a) > fl_absf(b)) {
        fl_face_pair<0>(p, d, h, false, q);
        if (q.k < 2) fl_face_pair<1>(p, d, h, true, q);
    } else {
        fl_face_pair<1>(p, d, h, false, q);
        if (q.k < 2) fl_face_pair<0>(p, d, h, true, q);
    }
    if (q.k < 2) fl_face_pair<2>(p, d, h, true, q);
    if (q.k == 2) {
        float u = q.p1[0] - q.p0[0], v = q.p1[1] - q.p0[1], w = q.p1[2] - q.p0[2];
        return sqrtf(u * u + v * v + w * w);
    }
    return 0.0f;
}

// fl_chord_box3 with the six direction ratios supplied by the caller
// (sba=b/a, sca=c/a, sab=a/b, scb=c/b, sac=a/c, sbc=b/c, each computed by
// the same IEEE division the reference performs).  Identical result to
// fl_chord_box3 for a != 0 and b != 0; the caller handles a == 0 / b == 0.
struct FlRatios {
    float sba, sca, sab, scb, sac, sbc;
};

template <int AX>
FL_HD void fl_face_pair_r(const float* p, const float* h, float su, float sv, bool later, FlPts& q)
{
    constexpr int U = (AX == 0) ? 1 : 0;
    constexpr int V = (AX == 2) ? 1 : 2;
    float mu = su * (-p[AX]) + p[U], wu = su * h[AX];
    float mv = sv * (-p[AX]) + p[V], wv = sv * h[AX];
    if (fl_absf(mu - wu) <= h[U] && fl_absf(mv - wv) <= h[V]) {
        float* o = q.k == 0 ? q.p0 : q.p1;
        o[AX] = -h[AX]; o[U] = mu - wu; o[V] = mv - wv;
        q.k++;
    }
    if (!later || q.k < 2) {
        if (fl_absf(mu + wu) <= h[U] && fl_absf(mv + wv) <= h[V]) {
            float* o = q.k == 0 ? q.p0 : q.p1;
            o[AX] = h[AX]; o[U] = mu + wu; o[V] = mv + wv;
            q.k++;
        }
    }
}

FL_HD float fl_chord_box3_r(const float* p, const float* h, bool xmajor, const FlRatios& r)
{
    FlPts q;
    q.k = 0;
    if (xmajor) {
        fl_face_pair_r<0>(p, h, r.sba, r.sca, false, q);
        if (q.k < 2) fl_face_pair_r<1>(p, h, r.sab, r.scb, true, q);
    } else {
        fl_face_pair_r<1>(p, h, r.sab, r.scb, false, q);
        if (q.k < 2) fl_face_pair_r<0>(p, h, r.sba, r.sca, true, q);
    }
    if (q.k < 2) fl_face_pair_r<2>(p, h, r.sac, r.sbc, true, q);
    if (q.k == 2) {
        float u = q.p1[0] - q.p0[0], v = q.p1[1] - q.p0[1], w = q.p1[2] - q.p0[2];
        return sqrtf(u * u + v * v + w * w);
    }
    return 0.0f;
}

// Detector cells whose ray can meet a voxel box, from the projection of the
// box corners (shadow bounding box) widened by `margin` cells.  Rotated
// frame: xr along the source->centre axis, yr along the detector row; the
// ray of column a passes the detector at u = y_det[a] and meets a point
// (xr, yr) when u = yr*SD/(xr+SO) (likewise rows: v = z_det[b] and
// zr*SD/(xr+SO), zr relative to the source height).  Corner offsets of the
// rotated unit square are (c*dx+s*dy, -s*dx+c*dy), dx,dy = +-1/2.
// Hardware reciprocals are enough here: the margin (a twentieth of a cell)
// dwarfs their rounding.
struct ShadowWin {
    int alo, ahi, blo, bhi;
};

__device__ __forceinline__ ShadowWin fl_shadow_window(float c, float s, float SO, float SD, float xc, float yc,
                                                     float zr, float hz, float ya0, float inv_pa, float zb0,
                                                     float inv_pb, float margin, int na, int nb)
{
    const float xr = c * xc + s * yc, yr = -s * xc + c * yc;
    const float hc = 0.5f * c, hs = 0.5f * s;
    float umin = FL_INF, umax = -FL_INF;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float ox = (k & 1) ? hc : -hc, oy = (k & 2) ? hs : -hs;   // c*dx, s*dy
        const float xk = xr + ox + oy;
        const float yk = yr + ((k & 1) ? -hs : hs) + ((k & 2) ? hc : -hc);
        const float u = yk * SD * __builtin_amdgcn_rcpf(xk + SO);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
    }
    const float ex = 0.5f * (fabsf(c) + fabsf(s));
    const float rn = SD * __builtin_amdgcn_rcpf(xr - ex + SO), rf = SD * __builtin_amdgcn_rcpf(xr + ex + SO);
    const float zl = zr - hz, zh = zr + hz;
    const float vmax = zh > 0 ? zh * rn : zh * rf;
    const float vmin = zl < 0 ? zl * rn : zl * rf;
    ShadowWin w;
    // (a pitch may be negative: order the mapped ends)
    const float a1 = (umin - ya0) * inv_pa, a2 = (umax - ya0) * inv_pa;
    const float b1 = (vmin - zb0) * inv_pb, b2 = (vmax - zb0) * inv_pb;
    w.alo = (int)ceilf(fminf(a1, a2) - margin);
    w.ahi = (int)floorf(fmaxf(a1, a2) + margin);
    w.blo = (int)ceilf(fminf(b1, b2) - margin);
    w.bhi = (int)floorf(fmaxf(b1, b2) + margin);
    if (w.alo < 0) w.alo = 0;
    if (w.ahi > na - 1) w.ahi = na - 1;
    if (w.blo < 0) w.blo = 0;
    if (w.bhi > nb - 1) w.bhi = nb - 1;
    return w;
}
