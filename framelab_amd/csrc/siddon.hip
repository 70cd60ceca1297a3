// siddon.hip -- ray-driven Siddon forward projectors (fan and cone).
//
// One thread per ray; a workgroup covers a TA x TB patch of detector cells
// of one view so that its rays cross neighbouring voxels together and the
// sinogram row writes are coalesced.  Workgroups are remapped so that
// consecutive patches of a view run on the same XCD and share its L2.
//
// Arithmetic follows matlab/mex/src/Ax_cone_mf_cpu_siddon.cpp:35-131 and
// Ax_fan_mf_cpu_siddon.cpp:32-91 operation by operation (this file is built
// with -ffp-contract=off), so ray sums are bit-identical to the reference CPU
// path; see fl_siddon.h for the traversal.
#include <algorithm>

#include "fl_common.h"
#include "fl_geom.h"
#include "fl_siddon.h"
#include "fl_chord.h"
#include "fl_fan.h"

namespace {

struct RayGeo {
    float x1, y1, z1, dx, dy, dz, L;
};

// Source/detector geometry of one ray, in the reference's operation order
// (Ax_cone_mf_cpu_siddon.cpp:37-48, :112).
__device__ __forceinline__ RayGeo cone_ray(const FwdGeom& g, int iv, int ia, int ib)
{
    RayGeo r;
    float c = g.cs[iv], s = g.sn[iv];
    r.x1 = c * (-g.SO);
    r.y1 = s * (-g.SO);
    r.z1 = g.sdz[iv];
    float yd = g.ydet[ia];
    float x2 = c * g.OD - s * yd;
    float y2 = s * g.OD + c * yd;
    float z2 = g.zdet[ib] + g.sdz[iv];
    r.dx = x2 - r.x1;
    r.dy = y2 - r.y1;
    r.dz = z2 - r.z1;
    r.L = sqrtf(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
    return r;
}

__device__ __forceinline__ RayGeo fan_ray(const FwdGeom& g, int iv, int id)
{
    RayGeo r;
    float c = g.cs[iv], s = g.sn[iv];
    r.x1 = c * (-g.SO);
    r.y1 = s * (-g.SO);
    r.z1 = 0.0f;
    float yd = g.ydet[id];
    float x2 = c * g.OD - s * yd;
    float y2 = s * g.OD + c * yd;
    r.dx = x2 - r.x1;
    r.dy = y2 - r.y1;
    r.dz = 0.0f;
    r.L = sqrtf(r.dx * r.dx + r.dy * r.dy);
    return r;
}

}  // namespace

// Bricked volume layout for the ray gathers: the frame is cut into
// BX x BY x BZ bricks stored contiguously (x fastest inside a brick, then
// y, then z; bricks x-major, then y, then z), so a cache line holds a
// compact patch of voxels whatever the view angle and neighbouring rays
// share lines.  BX = BY = BZ = 1 is the plain layout.  Padding to whole
// bricks is never read.
template <int BX, int BY, int BZ = 1>
struct Bricks {
    int nbx, nby;
    unsigned row, plane;   // elements per row of bricks / per slab of BZ z planes
    size_t frame;          // elements per bricked frame
    __host__ __device__ Bricks(int nx, int ny, int nz)
        : nbx((nx + BX - 1) / BX), nby((ny + BY - 1) / BY), row((unsigned)nbx * BX * BY * BZ),
          plane(row * (unsigned)nby), frame((size_t)plane * ((nz + BZ - 1) / BZ)) {}
    // (tx, ty, tz) in the volume, so non-negative: unsigned shifts and masks,
    // and full-rate 24-bit multiplies -- bricked layouts are launched only
    // where every factor fits 24 bits (bricked_fits_u24), so no branch here
    // (a branch on the plane size made the compiler fall back to the
    // quarter-rate 32-bit multiply on every step of the walk)
    //
    // The offset (x/BX)*BV + x%BX + (y/BY)*row + (y%BY)*BX + (z/BZ)*plane +
    // (z%BZ)*BX*BY (BV = BX*BY*BZ) is formed as
    //   x + BX*y + BX*BY*z + (x/BX)*(BV-BX) + (y/BY)*(row-BX*BY) + (z/BZ)*(plane-BV)
    // (x%BX = x - (x/BX)*BX, likewise y and z): shifts, shift-adds and three
    // 24-bit products, about ten instructions.  Every partial sum is taken
    // modulo 2^32 and the offset itself is below 2^32, so it comes out exact.
    __device__ __forceinline__ unsigned at(int tx, int ty, int tz) const
    {
        const unsigned x = (unsigned)tx, y = (unsigned)ty, z = (unsigned)tz;
        if (BX == 1 && BY == 1 && BZ == 1) return (z * nby + y) * nbx + x;
        constexpr unsigned BV = BX * BY * BZ;
        const unsigned lin = x + BX * y + BX * BY * z;
        return __umul24(z / BZ, plane - BV) + __umul24(y / BY, row - BX * BY) + __umul24(x / BX, BV - BX) + lin;
    }
};

// Elements of the largest bricked frame of any launch shape (plan buffers).
size_t fl_bricked_frame_elems(int nx, int ny, int nz)
{
    return std::max(Bricks<8, 8>(nx, ny, nz).frame,
                    std::max(Bricks<4, 4, 4>(nx, ny, nz).frame, Bricks<4, 4, 2>(nx, ny, nz).frame));
}

// The bricked layouts' index arithmetic needs a brick slab of under 2^24
// elements, under 2^24 slabs and, as its in-frame offsets are 32-bit, a
// bricked frame of at most 2^32 elements (the padding to whole bricks can
// make a thin volume's frame several times its voxel count).
template <int BX, int BY, int BZ>
static bool bricked_fits_u24(int nx, int ny, int nz)
{
    const Bricks<BX, BY, BZ> br(nx, ny, nz);
    return br.plane < (1u << 24) && nz < (1 << 24) && br.frame <= ((size_t)1 << 32);
}

template <typename VT, int BX, int BY, int BZ>
__global__ __launch_bounds__(256) void k_to_bricks(const VT* __restrict__ x, VT* __restrict__ xb, int nx, int ny,
                                                  int nz, size_t nvox, size_t n)
{
    const Bricks<BX, BY, BZ> br(nx, ny, nz);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t f = i / nvox, v = i % nvox;
        const int tx = (int)(v % nx), ty = (int)((v / nx) % ny), tz = (int)(v / ((size_t)nx * ny));
        xb[f * br.frame + br.at(tx, ty, tz)] = x[i];
    }
}

// A workgroup covers a TA x TB patch of detector cells of one view; each
// wave covers a WA x (64/WA) sub-patch, so that its 64 rays run close
// together through the volume and their gathers share cache lines.  X is
// the input in the BX x BY x BZ brick layout.
template <typename VT, typename YT, int TA, int TB, int WA, int BX, int BY, int BZ>
__global__ __launch_bounds__(TA* TB) void k_fwd_cone(const VT* __restrict__ X, YT* __restrict__ Y,
                                                    FwdGeom g)
{
    constexpr int WB = 64 / WA, NWA = TA / WA;
    static_assert(TA % WA == 0 && TB % WB == 0 && (TA * TB) % 64 == 0, "tile shape");
    const uint32_t tiles_a = (g.na + TA - 1) / TA;
    const uint32_t tiles_b = (g.nb + TB - 1) / TB;
    const uint32_t tpv = tiles_a * tiles_b;
    const uint32_t b = fl_xcd_remap(blockIdx.x, gridDim.x);
    const int slot = (int)(b / tpv);
    const uint32_t t = b % tpv;
    const int w = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63;
    const int ia = (int)(t % tiles_a) * TA + (w % NWA) * WA + (l % WA);
    const int ib = (int)(t / tiles_a) * TB + (w / NWA) * WB + (l / WA);
    if (slot >= g.nslots || ia >= g.na || ib >= g.nb) return;
    const int iv = g.slot_view[slot];
    const Bricks<BX, BY, BZ> br(g.nx, g.ny, g.nz);
    const VT* __restrict__ x = X + (size_t)g.view_frame[iv] * br.frame;

    RayGeo r = cone_ray(g, iv, ia, ib);
    SidAxis ax, ay, az;
    sid_axis_init(ax, r.x1, r.dx, 1.0f, g.nx);
    sid_axis_init(ay, r.y1, r.dy, 1.0f, g.ny);
    sid_axis_init(az, r.z1, r.dz, g.dz0, g.nz);
    const float L = r.L;
    // y = y + S*x in walk order.  Every step issues its gather (outside the
    // volume from cell 0, whose term is then +0) before the step's plane
    // advance and adds its term after it, so the gather has the advance's
    // arithmetic to land behind.  Adding +0 leaves a sum that starts at +0
    // unchanged: the result is bit-identical to adding the in-volume terms.
    struct Term {
        float S;
        VT x;
        bool in;
    };
    struct Visit {
        const VT* __restrict__ x;
        const Bricks<BX, BY, BZ>& br;
        float L;
        float acc;
        __device__ Term operator()(int tx, int ty, int tz, float dal, bool in) const
        {
            // offset formed for every lane and masked to cell 0 outside
            // the volume (no branch around the offset arithmetic)
            return Term{dal * L, x[br.at(tx, ty, tz) & (0u - (unsigned)in)], in};
        }
        __device__ void commit(const Term& t) { acc = acc + (t.in ? t.S * fl_load(&t.x, 0) : 0.0f); }
    } visit{x, br, L, 0.0f};
    sid_walk_all3(ax, ay, az, L, g.dz0 == 1.0f, visit);
    const float acc = visit.acc;
    fl_store(Y, ((size_t)slot * g.nb + ib) * g.na + ia, acc * g.scale);
}

template <typename VT, typename YT>
__global__ __launch_bounds__(64) void k_fwd_fan(const VT* __restrict__ X, YT* __restrict__ Y, FwdGeom g)
{
    const uint32_t tiles = (g.na + 63) / 64;
    const uint32_t b = fl_xcd_remap(blockIdx.x, gridDim.x);
    const int slot = (int)(b / tiles);
    const int id = (int)(b % tiles) * 64 + (int)threadIdx.x;
    if (slot >= g.nslots || id >= g.na) return;
    const int iv = g.slot_view[slot];
    fl_store(Y, (size_t)slot * g.na + id, fl_fan_ray_sum(X + (size_t)g.view_frame[iv] * g.nvox, g, iv, id));
}

// Traversal log of selected rays (parity tests): ray k = (view, ia, ib)
// writes its in-volume segments (voxel index, length) at off[k] onwards,
// at most cap[k] of them; cnt[k] receives the true count.
__global__ void k_trace(FwdGeom g, int cone, int nrays, const int* __restrict__ rays,
                        const int* __restrict__ off, const int* __restrict__ cap,
                        int* __restrict__ J, float* __restrict__ S, int* __restrict__ cnt)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrays) return;
    int iv = rays[3 * k], ia = rays[3 * k + 1], ib = rays[3 * k + 2];
    RayGeo r = cone ? cone_ray(g, iv, ia, ib) : fan_ray(g, iv, ia);
    SidAxis ax, ay, az;
    sid_axis_init(ax, r.x1, r.dx, 1.0f, g.nx);
    sid_axis_init(ay, r.y1, r.dy, 1.0f, g.ny);
    const float L = r.L;
    int n = 0, base = off[k], lim = cap[k];
    auto log = [&](int j, float dal) {
        if (n < lim) { J[base + n] = j; S[base + n] = dal * L; }
        n++;
    };
    if (cone) {
        sid_axis_init(az, r.z1, r.dz, g.dz0, g.nz);
        sid_walk<true>(ax, ay, az, L, g.dz0 == 1.0f, log);
    } else {
        sid_axis_init(az, 0.0f, 0.0f, 1.0f, 2);
        sid_walk<false>(ax, ay, az, L, true, log);
    }
    cnt[k] = n;
}

// ---- host launchers ------------------------------------------------------

template <typename VT, typename YT, int TA, int TB, int WA, int BX, int BY, int BZ = 1>
static hipError_t launch_cone_shape(const void* X, void* Y, const FwdGeom& g, hipStream_t st)
{
    uint32_t tiles = ((g.na + TA - 1) / TA) * ((g.nb + TB - 1) / TB);
    uint32_t grid = tiles * (uint32_t)g.nslots;
    if (grid == 0) return hipSuccess;
    const VT* src = (const VT*)X;
    if (BX > 1 || BY > 1 || BZ > 1) {
        // nt frames into the plan's brick buffer (sized for 4-byte elements)
        const int nt = g.nframes;
        const size_t n = g.nvox * (size_t)nt;
        const unsigned gb = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL((k_to_bricks<VT, BX, BY, BZ>), dim3(gb), dim3(256), 0, st, (const VT*)X, (VT*)g.bricks,
                           g.nx, g.ny, g.nz, g.nvox, n);
        src = (const VT*)g.bricks;
    }
    hipLaunchKernelGGL((k_fwd_cone<VT, YT, TA, TB, WA, BX, BY, BZ>), dim3(grid), dim3(TA * TB), 0, st, src,
                       (YT*)Y, g);
    return hipGetLastError();
}

// Launch shapes, indexed by FwdGeom::tile (fl_plan_autotune times them):
// workgroup TA x TB rays, waves WA x 64/WA, volume bricks BX x BY x BZ.
template <typename VT, typename YT>
static hipError_t launch_cone_t(const void* X, void* Y, const FwdGeom& g, hipStream_t st)
{
    if (g.tile != 0 && g.tile != 3 && !g.bricks) return hipErrorInvalidValue;
    // slabs too large for the bricks' 24-bit index arithmetic take the
    // plain layout of the same workgroup shape family
    const bool fits = bricked_fits_u24<8, 8, 1>(g.nx, g.ny, g.nz) && bricked_fits_u24<4, 4, 4>(g.nx, g.ny, g.nz) &&
                      bricked_fits_u24<4, 4, 2>(g.nx, g.ny, g.nz);
    const int tile = (g.tile == 0 || g.tile == 3 || fits) ? g.tile : 3;
    switch (tile) {
    case 0: return launch_cone_shape<VT, YT, 64, 4, 64, 1, 1>(X, Y, g, st);
    case 1: return launch_cone_shape<VT, YT, 16, 16, 8, 8, 8>(X, Y, g, st);
    case 2: return launch_cone_shape<VT, YT, 32, 8, 8, 8, 8>(X, Y, g, st);
    case 3: return launch_cone_shape<VT, YT, 16, 16, 8, 1, 1>(X, Y, g, st);
    case 4: return launch_cone_shape<VT, YT, 16, 16, 8, 4, 4>(X, Y, g, st);
    case 5: return launch_cone_shape<VT, YT, 64, 4, 64, 8, 8>(X, Y, g, st);
    case 6: return launch_cone_shape<VT, YT, 16, 16, 8, 4, 4, 4>(X, Y, g, st);
    default: return launch_cone_shape<VT, YT, 16, 16, 8, 4, 4, 2>(X, Y, g, st);
    }
}

template <typename VT, typename YT>
static hipError_t launch_fan_t(const void* X, void* Y, const FwdGeom& g, hipStream_t st)
{
    uint32_t grid = ((g.na + 63) / 64) * (uint32_t)g.nslots;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_fwd_fan<VT, YT>), dim3(grid), dim3(64), 0, st, (const VT*)X, (YT*)Y, g);
    return hipGetLastError();
}

#define FL_DISPATCH2(FN, xt, yt, ...)                                                        \
    do {                                                                                     \
        if (xt == FL_F32 && yt == FL_F32) return FN<float, float>(__VA_ARGS__);               \
        if (xt == FL_BF16 && yt == FL_BF16) return FN<bf16_t, bf16_t>(__VA_ARGS__);           \
        if (xt == FL_F16 && yt == FL_F16) return FN<f16_t, f16_t>(__VA_ARGS__);               \
        if (xt == FL_BF16 && yt == FL_F32) return FN<bf16_t, float>(__VA_ARGS__);             \
        if (xt == FL_F16 && yt == FL_F32) return FN<f16_t, float>(__VA_ARGS__);               \
        if (xt == FL_F32 && yt == FL_BF16) return FN<float, bf16_t>(__VA_ARGS__);             \
        if (xt == FL_F32 && yt == FL_F16) return FN<float, f16_t>(__VA_ARGS__);               \
        return hipErrorInvalidValue;                                                         \
    } while (0)

hipError_t fl_launch_fwd_cone(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st)
{
    FL_DISPATCH2(launch_cone_t, xt, yt, X, Y, g, st);
}

hipError_t fl_launch_fwd_fan(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st)
{
    FL_DISPATCH2(launch_fan_t, xt, yt, X, Y, g, st);
}

hipError_t fl_launch_trace(const FwdGeom& g, int cone, int nrays, const int* rays, const int* off,
                           const int* cap, int* J, float* S, int* cnt, hipStream_t st)
{
    if (nrays <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace, dim3((nrays + 63) / 64), dim3(64), 0, st, g, cone, nrays, rays, off,
                       cap, J, S, cnt);
    return hipGetLastError();
}

// ---- Gao column-walk fan forward (class default, version 1) -------------
template <typename VT, typename YT>
__global__ __launch_bounds__(64) void k_fwd_fan_gao(const VT* __restrict__ X, YT* __restrict__ Y, FwdGeom g)
{
    const uint32_t tiles = (g.na + 63) / 64;
    const uint32_t b = fl_xcd_remap(blockIdx.x, gridDim.x);
    const int slot = (int)(b / tiles);
    const int id = (int)(b % tiles) * 64 + (int)threadIdx.x;
    if (slot >= g.nslots || id >= g.na) return;
    const int iv = g.slot_view[slot];
    fl_store(Y, (size_t)slot * g.na + id, fl_fan_gao_sum(X + (size_t)g.view_frame[iv] * g.nvox, g, iv, id));
}

template <typename VT, typename YT>
static hipError_t launch_fan_gao_t(const void* X, void* Y, const FwdGeom& g, hipStream_t st)
{
    uint32_t grid = ((g.na + 63) / 64) * (uint32_t)g.nslots;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((k_fwd_fan_gao<VT, YT>), dim3(grid), dim3(64), 0, st, (const VT*)X, (YT*)Y, g);
    return hipGetLastError();
}

hipError_t fl_launch_fwd_fan_gao(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st)
{
    FL_DISPATCH2(launch_fan_gao_t, xt, yt, X, Y, g, st);
}

// Per-ray source-detector distance |d| (the Siddon totallength), used for
// the normalised output yNorm = y / (|d| * scale).
__global__ void k_ynorm(const float* __restrict__ y, float* __restrict__ yn, FwdGeom g)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nd = (size_t)g.na * g.nb;
    if (k >= nd * (size_t)g.nslots) return;
    int slot = (int)(k / nd);
    int ib = (int)((k % nd) / g.na), ia = (int)(k % g.na);
    RayGeo r = cone_ray(g, g.slot_view[slot], ia, ib);
    yn[k] = y[k] / (r.L * g.scale);
}

hipError_t fl_launch_ynorm(const float* y, float* yn, const FwdGeom& g, hipStream_t st)
{
    size_t n = (size_t)g.na * g.nb * g.nslots;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ynorm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, yn, g);
    return hipGetLastError();
}
