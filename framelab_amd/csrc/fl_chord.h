// fl_chord.h -- exact ray/pixel and ray/voxel chord lengths (host+device).
//
// Same float arithmetic as the reference chord routines, so the voxel-driven
// adjoints are bit-identical to the reference's:
//   fl_chord_line2  <- matlab/mex/src/find_l.cpp:18-45 (line a*x+b*y=c through
//                      the unit pixel at (x,y));
//   fl_chord_seg2   <- matlab/mex/src/find_l_3d.cpp:19-92 (segment-line
//                      through a dx-by-dy cell);
//   fl_chord_box3   <- matlab/mex/src/find_l_3d.cpp:94-182 (line through a
//                      dx-by-dy-by-dz box; faces tried x,y,z or y,x,z).
// One deliberate difference: a line parallel to the x = const (y = const)
// planes is credited only to boxes whose x (y) slab holds it; the
// reference's fallback to the 2-D routine (find_l_3d.cpp:104-108) ignores
// that coordinate, crediting boxes the line misses.
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

// fl_chord_line2 with the line's own quotients supplied, q = (1/max(|a|,
// |b|), RN(1/(|a||b|))) (the same for every pixel a line meets): the flat
// value is q.x itself, and (dmax - d)/(|a||b|) is formed from q.y by the
// Markstein correction, bit-identical to the division.
__device__ __forceinline__ float fl_chord_line2_q(float a, float b, float c, float x, float y, float2 q)
{
    float a2 = fl_absf(a), b2 = fl_absf(b);
    float hx = a2 / 2, hy = b2 / 2;
    float dmax = hx + hy;
    float t = c - a * x - b * y;
    float d = fl_absf(t);
    float d0 = fl_absf(hx - hy);
    const float slope = fl_div_r(dmax - d, a2 * b2, q.y);
    const float l = (d <= d0) ? q.x : slope;
    return d < dmax ? l : 0.0f;
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
    // the line lies in the plane x = x1 (y = y1): outside the box's slab
    // along that axis it misses the box
    if (a == 0) return fl_absf(x1 - x) <= dx / 2 ? fl_chord_seg2(y1, z1, y2, z2, dy, dz, y, z) : 0.0f;
    if (b == 0) return fl_absf(y1 - y) <= dy / 2 ? fl_chord_seg2(x1, z1, x2, z2, dx, dz, x, z) : 0.0f;
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

// sqrtf(s) for finite s >= 0: the hardware square root and the
// round-to-nearest test of its two neighbours (the correctly rounded
// expansion the compiler emits), without that expansion's rescaling of
// inputs below 2^-96 -- those take sqrtf itself -- and without its
// special-value fix-up, which the one special input left, +0, does not need
// (the neighbour tests keep the hardware's +0).  Bit-identical to sqrtf.
__device__ __forceinline__ float fl_sqrt_nonneg(float s)
{
    float r = __builtin_amdgcn_sqrtf(s);
    const float rd = __uint_as_float(__float_as_uint(r) - 1u), ru = __uint_as_float(__float_as_uint(r) + 1u);
    const float ed = fmaf(-rd, r, s), eu = fmaf(-ru, r, s);
    r = ed <= 0.0f ? rd : r;
    r = eu > 0.0f ? ru : r;
    if (s < 0x1p-96f && s > 0.0f) {
        asm volatile("");   // keeps the rare path a branch (not speculated for every lane)
        r = sqrtf(s);
    }
    return r;
}

// Branch-free form of fl_chord_box3_r for the generic case.  For each face
// pair the candidate on the face the ray enters through (the -h face when
// the direction along that axis is positive) and the one it leaves through
// are formed with the same operations as in fl_face_pair_r: with the sign
// s of the direction, mu - s*wu equals the lo (s > 0) or hi (s < 0)
// candidate bit for bit, since x - (-y) and x + y round the same sum, and
// s*wu = r*(s*h) rounds to (r*h)*s exactly (a sign flip, and rounding is
// sign-symmetric).  When exactly one entry and one exit candidate pass,
// they are the only passing candidates, so they are the two points
// fl_chord_box3_r takes (in some order; the squared differences do not
// depend on it) and the length is bit-identical.  With fewer than two
// passing candidates the length is 0; every other pattern (ray through an
// edge or corner) takes the ordered routine, with xmajor() = |a| > |b|.
// hs = signed half sides (sign a * h[0], sign b * h[1], sign c * h[2]).
// The candidate counts are wave-wide lane masks (scalar-unit logic).
struct FlRatiosS {
    FlRatios r;
    float hs[3];
};

template <typename XMajor>
__device__ __forceinline__ float fl_chord_box3_fast(const float* p, const float* h, const FlRatiosS& q,
                                                    XMajor xmajor)
{
    const FlRatios& r = q.r;
    const float hx = q.hs[0], hy = q.hs[1], hz = q.hs[2];
    // x faces: U = y (slope sba), V = z (slope sca)
    const float mxu = r.sba * (-p[0]) + p[1], wxu = r.sba * hx;
    const float mxv = r.sca * (-p[0]) + p[2], wxv = r.sca * hx;
    // y faces: U = x (slope sab), V = z (slope scb)
    const float myu = r.sab * (-p[1]) + p[0], wyu = r.sab * hy;
    const float myv = r.scb * (-p[1]) + p[2], wyv = r.scb * hy;
    // z faces: U = x (slope sac), V = y (slope sbc)
    const float mzu = r.sac * (-p[2]) + p[0], wzu = r.sac * hz;
    const float mzv = r.sbc * (-p[2]) + p[1], wzv = r.sbc * hz;
    // entry (e) and exit (x) candidates
    const float ex_u = mxu - wxu, ex_v = mxv - wxv, xx_u = mxu + wxu, xx_v = mxv + wxv;
    const float ey_u = myu - wyu, ey_v = myv - wyv, xy_u = myu + wyu, xy_v = myv + wyv;
    const float ez_u = mzu - wzu, ez_v = mzv - wzv, xz_u = mzu + wzu, xz_v = mzv + wzv;
    // which candidates lie on their faces, as lane masks of the wave (one
    // compare each; the logic below runs on the scalar unit)
    auto in = [](float c, float hh) { return __builtin_amdgcn_ballot_w64(fl_absf(c) <= hh); };
    const uint64_t bex = in(ex_u, h[1]) & in(ex_v, h[2]), bxx = in(xx_u, h[1]) & in(xx_v, h[2]);
    const uint64_t bey = in(ey_u, h[0]) & in(ey_v, h[2]), bxy = in(xy_u, h[0]) & in(xy_v, h[2]);
    const uint64_t bez = in(ez_u, h[0]) & in(ez_v, h[1]), bxz = in(xz_u, h[0]) & in(xz_v, h[1]);
    const bool fex = __builtin_amdgcn_inverse_ballot_w64(bex), fxx = __builtin_amdgcn_inverse_ballot_w64(bxx);
    const bool fey = __builtin_amdgcn_inverse_ballot_w64(bey), fxy = __builtin_amdgcn_inverse_ballot_w64(bxy);
    // exactly one entry and one exit candidate (one), or at least two of
    // either kind (edge and corner rays)
    const uint64_t two_e = (bex & bey) | (bez & (bex | bey)), two_x = (bxx & bxy) | (bxz & (bxx | bxy));
    const uint64_t any_e = bex | bey | bez, any_x = bxx | bxy | bxz;
    const bool one = __builtin_amdgcn_inverse_ballot_w64(any_e & any_x & ~(two_e | two_x));
    const bool corner = __builtin_amdgcn_inverse_ballot_w64(two_e | two_x);
    // one entry and one exit candidate: the chord between them, formed
    // without a branch (lanes of other cases discard it)
    const float e0 = fex ? -hx : (fey ? ey_u : ez_u);
    const float e1 = fex ? ex_u : (fey ? -hy : ez_v);
    const float e2 = fex ? ex_v : (fey ? ey_v : -hz);
    const float x0 = fxx ? hx : (fxy ? xy_u : xz_u);
    const float x1 = fxx ? xx_u : (fxy ? hy : xz_v);
    const float x2 = fxx ? xx_v : (fxy ? xy_v : hz);
    const float u = x0 - e0, v = x1 - e1, w = x2 - e2;
    float l = one ? fl_sqrt_nonneg(u * u + v * v + w * w) : 0.0f;
    if (corner) l = fl_chord_box3_r(p, h, xmajor(), r);
    return l;
}

// Detector cells whose ray can meet a voxel box, from the projection of the
// box corners (shadow bounding box) widened by a margin in cells.  Rotated
// frame: xr along the source->centre axis, yr along the detector row; the
// ray of column a passes the detector at u = y_det[a] and meets a point
// (xr, yr) when u = yr*SD/(xr+SO) (likewise rows: v = z_det[b] and
// zr*SD/(xr+SO), zr relative to the source height).  Corner offsets of the
// rotated unit square are (c*dx+s*dy, -s*dx+c*dy), dx,dy = +-1/2.
// Hardware reciprocals and the one-FMA index maps are enough here: the
// margin (a twentieth of a cell) dwarfs their rounding.

// Coordinate -> widened cell index bounds of a uniform grid with first
// coordinate p0 and pitch 1/inv_p: index = u*inv_p + (klo or khi).
struct ShadowMap {
    float inv_p, klo, khi;
};

__host__ __device__ inline ShadowMap fl_shadow_map(float p0, float inv_p, float margin)
{
    ShadowMap m;
    m.inv_p = inv_p;
    m.klo = -p0 * inv_p - margin;
    m.khi = -p0 * inv_p + margin;
    return m;
}

// Cells [lo, hi] (clamped to [0, n-1]) whose coordinates the closed range
// [umin, umax], widened by the margin, reaches; a negative pitch swaps ends.
__device__ __forceinline__ void fl_shadow_range(const ShadowMap& m, float umin, float umax, int n, int& lo, int& hi)
{
    const bool up = m.inv_p > 0;
    lo = (int)ceilf(fmaf(up ? umin : umax, m.inv_p, m.klo));
    hi = (int)floorf(fmaf(up ? umax : umin, m.inv_p, m.khi));
    if (lo < 0) lo = 0;
    if (hi > n - 1) hi = n - 1;
}

// Column range of a column of voxels that share (xc, yc), and what their
// row ranges need: the rows depend also on each voxel's height zr
// (relative to the source) and half-height.
struct ShadowCols {
    int alo, ahi;
    float rn, rf;   // SD / depth of the near and far ends of the square
};

__device__ __forceinline__ ShadowCols fl_shadow_cols(float c, float s, float SO, float SD, float xc, float yc,
                                                     const ShadowMap& ma, int na)
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
    ShadowCols w;
    w.rn = SD * __builtin_amdgcn_rcpf(xr - ex + SO);
    w.rf = SD * __builtin_amdgcn_rcpf(xr + ex + SO);
    fl_shadow_range(ma, umin, umax, na, w.alo, w.ahi);
    return w;
}

// Row range of the voxel at height zr (half-height hz) in that column: the
// top edge zr+hz projects highest through the square's near end if it lies
// above the source height and through its far end otherwise; the bottom
// edge zr-hz lowest through the near end if below, the far end otherwise.
__device__ __forceinline__ void fl_shadow_rows(const ShadowCols& sc, float zr, float hz, const ShadowMap& mb, int nb,
                                               int& blo, int& bhi)
{
    const float zl = zr - hz, zh = zr + hz;
    const float vmax = zh * (zh > 0 ? sc.rn : sc.rf);
    const float vmin = zl * (zl < 0 ? sc.rn : sc.rf);
    fl_shadow_range(mb, vmin, vmax, nb, blo, bhi);
}
