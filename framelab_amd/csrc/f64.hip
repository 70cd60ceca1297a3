// f64.hip -- the FP64 projector pair and an FP64 CGLS: the accuracy
// reference the FP32 / BF16 / FP16 paths are measured against at sizes a
// CPU oracle cannot reach (256^3 and beyond).
//
//  * Forward: the Siddon ray sum in double with the reference's traversal
//    rule (matlab/mex/src/Ax_cone_mf_cpu_siddon.cpp:35-131, merge
//    sort_alpha.cpp:15-60): plane parameters ((double)i*s - o)/d, the sorted
//    union of all axes' parameters (an exact cross-axis tie is one element),
//    segment cell floor(o + m*d + n/2) at the midpoint m, length
//    (a_{k+1} - a_k)*|d|.  One thread per ray; the union is streamed with a
//    3-way minimum from the volume entry (segments before the entry lie
//    outside the volume up to a midpoint rounding far below FP64 accuracy).
//  * Adjoint: its transpose, voxel-driven and atomic-free.  Per (voxel, ray)
//    the ray's Siddon segment inside the voxel is [largest entry-plane
//    parameter, smallest exit-plane parameter] of the voxel's three face
//    pairs, each parameter formed with the forward's formula, so the
//    coefficient is the forward's (to FP64 rounding).  Candidate cells come
//    from the voxel's shadow window on the detector (a 0.05-cell margin).
//  * CGLS (matlab/classes/+Optimizers/@CGLS/solveCGLS.m:7-30) with every
//    vector in double and fixed-order reductions, and split Bregman
//    (solveSplitBregman.m:16-41) over it with a one-level framelet
//    analysis / synthesis in double.
//
// No reference counterpart: the reference has FP32 projectors only.
#include <hip/hip_runtime.h>

#include <math.h>

#include <mutex>
#include <vector>

#include "../../include/framelab_hip.h"
#include "fl_capi_internal.h"
#include "fl_chord.h"
#include "fl_common.h"
#include "fl_framelet.h"

namespace {

constexpr double kInf = __builtin_huge_val();

struct Geo64 {
    const double *c, *s, *sdz, *ydet, *zdet;   // double geometry
    const float *cs, *sn;                      // float trig of the same angles (shadow windows only)
    const int *slot_view, *view_frame, *id_Y, *id_S, *Nv;
    double SO, OD, scale, dz0;
    int nx, ny, nz, na, nb, tmp_size;
    size_t nvox, nd;
    float ya0, inv_pa, zb0, inv_pb, SOf, SDf;
};

// One axis of the FP64 walk: plane j (0..n, walk order) at coordinate
// index i0 + st*j, parameter ((double)i * s - o) / d.
struct Ax64 {
    double o, d, s;
    int n, i0, st;
    bool on;
};

__device__ __forceinline__ void ax64_init(Ax64& a, double o, double d, double s, int n)
{
    a.o = o; a.d = d; a.s = s; a.n = n;
    a.on = d != 0.0;
    a.st = d > 0.0 ? 1 : -1;
    a.i0 = d > 0.0 ? -(n / 2) : n / 2;
}

__device__ __forceinline__ double ax64_alpha(const Ax64& a, int j)
{
    return ((double)(a.i0 + a.st * j) * a.s - a.o) / a.d;
}

// number of planes (walk order) whose parameter is <= t: an estimate from
// the coordinate at t, corrected with the exact parameters
__device__ __forceinline__ int ax64_count(const Ax64& a, double t)
{
    const double est = ((a.o + t * a.d) / a.s - (double)a.i0) * (double)a.st;
    int j = !(est > -1.0) ? 0 : !(est < (double)a.n + 1.0) ? a.n + 1 : (int)floor(est) + 1;
    while (j > 0 && ax64_alpha(a, j - 1) > t) j--;
    while (j <= a.n && ax64_alpha(a, j) <= t) j++;
    return j;
}

__device__ __forceinline__ double ax64_next(const Ax64& a, int j) { return (a.on && j <= a.n) ? ax64_alpha(a, j) : kInf; }

// A workgroup covers 64 detector columns x 4 rows of one output slot.
template <bool CONE>
__global__ __launch_bounds__(256) void k_fwd_f64(const double* __restrict__ X, double* __restrict__ Y, Geo64 g)
{
    const int ia = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int ib = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    const int slot = (int)blockIdx.z;
    if (ia >= g.na || ib >= g.nb) return;
    const int iv = g.slot_view[slot];
    const double* __restrict__ x = X + (size_t)g.view_frame[iv] * g.nvox;
    const double c = g.c[iv], s = g.s[iv];
    const double x1 = -g.SO * c, y1 = -g.SO * s, z1 = CONE ? g.sdz[iv] : 0.0;
    const double yd = g.ydet[ia];
    const double x2 = c * g.OD - s * yd, y2 = s * g.OD + c * yd, z2 = CONE ? g.zdet[ib] + z1 : 0.0;
    const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    const int nx2 = g.nx / 2, ny2 = g.ny / 2, nz2 = g.nz / 2;
    Ax64 ax, ay, az;
    ax64_init(ax, x1, dx, 1.0, g.nx);
    ax64_init(ay, y1, dy, 1.0, g.ny);
    ax64_init(az, z1, dz, g.dz0, CONE ? g.nz : 1);
    if (!CONE) az.on = false;
    double acc = 0.0;
    bool miss = false;
    // an axis the ray does not advance along keeps one cell for the whole ray
    if (!ax.on) miss |= (unsigned)(int)floor(x1 + (double)nx2) >= (unsigned)g.nx;
    if (!ay.on) miss |= (unsigned)(int)floor(y1 + (double)ny2) >= (unsigned)g.ny;
    if (CONE && !az.on) miss |= (unsigned)(int)floor(z1 / g.dz0 + (double)nz2) >= (unsigned)g.nz;
    double ae = -kInf, aq = kInf;
    if (ax.on) { ae = fmax(ae, ax64_alpha(ax, 0)); aq = fmin(aq, ax64_alpha(ax, ax.n)); }
    if (ay.on) { ae = fmax(ae, ax64_alpha(ay, 0)); aq = fmin(aq, ax64_alpha(ay, ay.n)); }
    if (az.on) { ae = fmax(ae, ax64_alpha(az, 0)); aq = fmin(aq, ax64_alpha(az, az.n)); }
    if (!miss && ae < aq) {
        int jx = ax.on ? ax64_count(ax, ae) : 0, jy = ay.on ? ax64_count(ay, ae) : 0;
        int jz = az.on ? ax64_count(az, ae) : 0;
        double qx = ax64_next(ax, jx), qy = ax64_next(ay, jy), qz = ax64_next(az, jz);
        double cur = ae;
        while (cur < aq) {
            const double an = fmin(fmin(qx, qy), qz);
            if (an == kInf) break;
            const double m = 0.5 * (cur + an);
            const int tx = (int)floor(x1 + m * dx + (double)nx2);
            const int ty = (int)floor(y1 + m * dy + (double)ny2);
            const int tz = CONE ? (int)floor((z1 + m * dz) / g.dz0 + (double)nz2) : 0;
            if ((unsigned)tx < (unsigned)g.nx && (unsigned)ty < (unsigned)g.ny &&
                (!CONE || (unsigned)tz < (unsigned)g.nz))
                acc += ((an - cur) * L) * x[((size_t)tz * g.ny + ty) * g.nx + tx];
            if (qx == an) qx = ax64_next(ax, ++jx);
            if (qy == an) qy = ax64_next(ay, ++jy);
            if (qz == an) qz = ax64_next(az, ++jz);
            cur = an;
        }
    }
    Y[((size_t)slot * g.nb + ib) * g.na + ia] = g.scale * acc;
}

// Parameters of the two planes bounding cell `t` of an axis (coordinate
// index t - n/2 and t + 1 - n/2), as entry / exit along the ray; an axis the
// ray does not advance along bounds nothing if the ray stays in that cell.
__device__ __forceinline__ bool slab64(double o, double d, double s, int n, int t, double& tin, double& tout)
{
    const int lo = t - n / 2;
    if (d == 0.0) return (int)floor(o / s + (double)(n / 2)) == t;
    const double a = ((double)lo * s - o) / d, b = ((double)(lo + 1) * s - o) / d;
    tin = fmax(tin, d > 0.0 ? a : b);
    tout = fmin(tout, d > 0.0 ? b : a);
    return true;
}

// One thread per voxel (16 x 16 columns per workgroup, one slice), views of
// its frame in list order, then detector rows, then columns.
template <bool CONE>
__global__ __launch_bounds__(256) void k_adj_f64(const double* __restrict__ Y, double* __restrict__ X, Geo64 g)
{
    const int ix = (int)blockIdx.x * 16 + (int)(threadIdx.x & 15);
    const int iy = (int)blockIdx.y * 16 + (int)(threadIdx.x >> 4);
    const int nzs = CONE ? g.nz : 1;
    const int iz = (int)blockIdx.z % nzs, it = (int)blockIdx.z / nzs;
    if (ix >= g.nx || iy >= g.ny) return;
    const float xcf = (float)ix + 0.5f - (float)(g.nx / 2), ycf = (float)iy + 0.5f - (float)(g.ny / 2);
    const float zcf = CONE ? ((float)iz + 0.5f - (float)(g.nz / 2)) * (float)g.dz0 : 0.0f;
    const ShadowMap ma = fl_shadow_map(g.ya0, g.inv_pa, 0.05f), mb = fl_shadow_map(g.zb0, g.inv_pb, 0.05f);
    double acc = 0.0;
    const int nvf = g.Nv[it];
    for (int j = 0; j < nvf; j++) {
        const int e = it * g.tmp_size + j;
        const int iv = g.id_Y[e];
        const double* __restrict__ y = Y + (size_t)g.id_S[e] * g.nd;
        const ShadowCols sc = fl_shadow_cols(g.cs[iv], g.sn[iv], g.SOf, g.SDf, xcf, ycf, ma, g.na);
        int blo = 0, bhi = 0;
        if (CONE) fl_shadow_rows(sc, zcf - (float)g.sdz[iv], 0.5f * (float)g.dz0, mb, g.nb, blo, bhi);
        if (sc.alo > sc.ahi || blo > bhi) continue;
        const double c = g.c[iv], s = g.s[iv];
        const double x1 = -g.SO * c, y1 = -g.SO * s, z1 = CONE ? g.sdz[iv] : 0.0;
        for (int ib = blo; ib <= bhi; ib++) {
            const double z2 = CONE ? g.zdet[ib] + z1 : 0.0;
            const double dz = z2 - z1;
            for (int ia = sc.alo; ia <= sc.ahi; ia++) {
                const double yd = g.ydet[ia];
                const double dx = (c * g.OD - s * yd) - x1, dy = (s * g.OD + c * yd) - y1;
                double tin = -kInf, tout = kInf;
                bool in = slab64(x1, dx, 1.0, g.nx, ix, tin, tout) && slab64(y1, dy, 1.0, g.ny, iy, tin, tout);
                if (CONE) in = in && slab64(z1, dz, g.dz0, g.nz, iz, tin, tout);
                if (in && tout > tin) {
                    const double L = sqrt(dx * dx + dy * dy + dz * dz);
                    acc += ((tout - tin) * L) * y[(size_t)ib * g.na + ia];
                }
            }
        }
    }
    X[(size_t)it * g.nvox + ((size_t)iz * g.ny + iy) * g.nx + ix] = g.scale * acc;
}

// ------------------------------------------------------------------ CGLS --

constexpr int NP64 = 1024;   // workgroups of the reduction kernels

struct Cg64 {
    double gamma, alpha, beta, tol;
    int stop, iter;
};

__device__ __forceinline__ void part_store(double v, double* __restrict__ part)
{
    __shared__ double sw[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

__device__ __forceinline__ double part_sum(const double* __restrict__ part, int np)
{
    __shared__ double sp[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) v += part[i];
    sp[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sp[threadIdx.x] += sp[threadIdx.x + w];
        __syncthreads();
    }
    return sp[0];
}

#define GRID_STRIDE(i, n) for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < (n); i += (size_t)gridDim.x * 256)

__global__ __launch_bounds__(256) void k64_init(const double* __restrict__ q, const double* __restrict__ u,
                                               const double* __restrict__ b, double lam, double* __restrict__ r,
                                               double* __restrict__ p, size_t n, double* __restrict__ part)
{
    double acc = 0.0;
    GRID_STRIDE(i, n)
    {
        const double ri = (q[i] + lam * u[i]) - b[i];
        r[i] = ri;
        p[i] = -ri;
        acc += ri * ri;
    }
    part_store(acc, part);
}

__global__ __launch_bounds__(256) void k64_fin_gamma(const double* __restrict__ part, int np, Cg64* st, double tol,
                                                    double* __restrict__ log)
{
    const double g = part_sum(part, np);
    if (threadIdx.x == 0) {
        st->gamma = g; st->alpha = 0.0; st->beta = 0.0; st->tol = tol;
        st->iter = 0;
        st->stop = g < tol ? 1 : 0;
        if (log) log[0] = g;
    }
}

__global__ __launch_bounds__(256) void k64_pq(double* __restrict__ q, const double* __restrict__ p, double lam,
                                             size_t n, double* __restrict__ part, const Cg64* st)
{
    double acc = 0.0;
    if (!st->stop) {
        GRID_STRIDE(i, n)
        {
            const double qi = q[i] + lam * p[i];
            q[i] = qi;
            acc += p[i] * qi;
        }
    }
    part_store(acc, part);
}

__global__ __launch_bounds__(256) void k64_fin_alpha(const double* __restrict__ part, int np, Cg64* st)
{
    const double pq = part_sum(part, np);
    if (threadIdx.x == 0 && !st->stop) {
        if (pq == 0.0) { st->stop = 1; st->alpha = 0.0; }
        else st->alpha = st->gamma / pq;
    }
}

__global__ __launch_bounds__(256) void k64_update(double* __restrict__ u, double* __restrict__ r,
                                                 const double* __restrict__ p, const double* __restrict__ q,
                                                 size_t n, double* __restrict__ part, const Cg64* st)
{
    double acc = 0.0;
    if (!st->stop) {
        const double a = st->alpha;
        GRID_STRIDE(i, n)
        {
            u[i] += a * p[i];
            const double ri = r[i] + a * q[i];
            r[i] = ri;
            acc += ri * ri;
        }
    }
    part_store(acc, part);
}

__global__ __launch_bounds__(256) void k64_fin_beta(const double* __restrict__ part, int np, Cg64* st,
                                                   double* __restrict__ log)
{
    const double g = part_sum(part, np);
    if (threadIdx.x == 0 && !st->stop) {
        st->beta = g / st->gamma;
        st->gamma = g;
        st->iter += 1;
        if (log) log[st->iter] = g;
        if (g < st->tol) st->stop = 1;
    }
}

__global__ __launch_bounds__(256) void k64_p(double* __restrict__ p, const double* __restrict__ r, size_t n,
                                            const Cg64* st)
{
    if (st->stop) return;
    const double bt = st->beta;
    GRID_STRIDE(i, n) p[i] = -r[i] + bt * p[i];
}

unsigned grid64(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < (size_t)NP64 ? (b ? b : 1) : NP64);
}

// ------------------------------------------------------------------ host --

int ensure_tables(fl_plan_s* p)
{
    std::lock_guard<std::mutex> lk(p->f64_mu);
    if (p->d_c64) return FL_OK;
    const int nv = p->nv;
    std::vector<double> c(nv), s(nv), sdz(nv), ydet(p->h_ydet.size()), zdet(p->h_zdet.size());
    for (int v = 0; v < nv; v++) {
        c[v] = cos((double)p->h_sd_phi[v]);
        s[v] = sin((double)p->h_sd_phi[v]);
        sdz[v] = (double)p->h_sdz[v];
    }
    for (size_t k = 0; k < ydet.size(); k++) ydet[k] = (double)p->h_ydet[k];
    for (size_t k = 0; k < zdet.size(); k++) zdet[k] = (double)p->h_zdet[k];
    double* dst[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const std::vector<double>* src[5] = {&c, &s, &sdz, &ydet, &zdet};
    for (int k = 0; k < 5; k++) {
        const size_t bytes = src[k]->size() * sizeof(double);
        hipError_t e = hipMalloc((void**)&dst[k], bytes ? bytes : 8);
        if (e == hipSuccess && bytes) e = hipMemcpy(dst[k], src[k]->data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            for (int q = 0; q <= k; q++)
                if (dst[q]) (void)hipFree(dst[q]);
            return fl_hip_fail(e, "FP64 geometry tables");
        }
    }
    p->d_c64 = dst[0]; p->d_s64 = dst[1]; p->d_sdz64 = dst[2]; p->d_ydet64 = dst[3]; p->d_zdet64 = dst[4];
    return FL_OK;
}

Geo64 geo64(const fl_plan_s* p)
{
    Geo64 g{};
    g.c = p->d_c64; g.s = p->d_s64; g.sdz = p->d_sdz64; g.ydet = p->d_ydet64; g.zdet = p->d_zdet64;
    g.cs = p->fg.cs; g.sn = p->fg.sn;
    g.slot_view = p->fg.slot_view; g.view_frame = p->fg.view_frame;
    g.id_Y = p->ag.id_Y; g.id_S = p->ag.id_S; g.Nv = p->ag.Nv;
    g.SO = p->fg.SO; g.OD = p->fg.OD; g.scale = p->fg.scale; g.dz0 = p->fg.dz0;
    g.nx = p->fg.nx; g.ny = p->fg.ny; g.nz = p->fg.nz; g.na = p->fg.na; g.nb = p->fg.nb;
    g.tmp_size = p->ag.tmp_size;
    g.nvox = (size_t)p->nvox; g.nd = (size_t)p->ndet;
    g.ya0 = p->fg.ya0; g.inv_pa = p->fg.inv_pa; g.zb0 = p->fg.zb0; g.inv_pb = p->fg.inv_pb;
    g.SOf = p->fg.SO; g.SDf = p->fg.SO + p->fg.OD;
    return g;
}

int launch_fwd(const fl_plan_s* p, const double* x, double* y, hipStream_t st)
{
    const Geo64 g = geo64(p);
    const dim3 grid((unsigned)((g.na + 63) / 64), (unsigned)((g.nb + 3) / 4), (unsigned)p->nv);
    if (p->cone) hipLaunchKernelGGL(k_fwd_f64<true>, grid, dim3(256), 0, st, x, y, g);
    else hipLaunchKernelGGL(k_fwd_f64<false>, grid, dim3(256), 0, st, x, y, g);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FL_OK : fl_hip_fail(e, "FP64 forward");
}

int launch_adj(const fl_plan_s* p, const double* y, double* x, hipStream_t st)
{
    if (!p->slots_ok)
        return fl_fail(FL_ERR_UNSUPPORTED, "a frame lists a view more often than the forward's output slots hold it");
    if (p->det_dev > FL_DET_DEV_MAX)
        return fl_fail(FL_ERR_UNSUPPORTED, "the voxel-driven adjoints need uniformly spaced detector coordinates");
    const Geo64 g = geo64(p);
    const dim3 grid((unsigned)((g.nx + 15) / 16), (unsigned)((g.ny + 15) / 16),
                    (unsigned)((p->cone ? g.nz : 1) * p->nt));
    if (p->cone) hipLaunchKernelGGL(k_adj_f64<true>, grid, dim3(256), 0, st, y, x, g);
    else hipLaunchKernelGGL(k_adj_f64<false>, grid, dim3(256), 0, st, y, x, g);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FL_OK : fl_hip_fail(e, "FP64 adjoint");
}

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem()
    {
        for (void* q : ptrs) (void)hipFree(q);
    }
    template <typename T>
    int get(T** out, size_t count, const char* what)
    {
        void* q = nullptr;
        if (hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess)
            return fl_fail(FL_ERR_ALLOC, "hipMalloc(%s, %zu bytes)", what, count * sizeof(T));
        ptrs.push_back(q);
        *out = (T*)q;
        return FL_OK;
    }
};

#define CHK(expr)                                             \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return fl_hip_fail(e_, #expr);  \
    } while (0)

}  // namespace

void fl_plan_release_f64(fl_plan_s* p)
{
    for (double* q : {p->d_c64, p->d_s64, p->d_sdz64, p->d_ydet64, p->d_zdet64})
        if (q) (void)hipFree(q);
    p->d_c64 = p->d_s64 = p->d_sdz64 = p->d_ydet64 = p->d_zdet64 = nullptr;
}

extern "C" int fl_forward_f64(fl_plan p, const double* d_x, double* d_y, void* stream)
{
    if (!p || !d_x || !d_y) return fl_fail(FL_ERR_INVALID, "null plan or buffer");
    int rc = ensure_tables(p);
    if (rc) return rc;
    if ((rc = launch_fwd(p, d_x, d_y, (hipStream_t)stream))) return rc;
    return fl_ok();
}

extern "C" int fl_adjoint_f64(fl_plan p, const double* d_y, double* d_x, void* stream)
{
    if (!p || !d_x || !d_y) return fl_fail(FL_ERR_INVALID, "null plan or buffer");
    int rc = ensure_tables(p);
    if (rc) return rc;
    if ((rc = launch_adj(p, d_y, d_x, (hipStream_t)stream))) return rc;
    return fl_ok();
}

namespace {

// FP64 CG workspace (allocated per call: a reference path)
struct Cg64Ws {
    double *r = nullptr, *p = nullptr, *q = nullptr, *sino = nullptr, *part = nullptr;
    Cg64* state = nullptr;
};

int cg64_alloc(DevMem& m, Cg64Ws& w, size_t n, size_t ns)
{
    int rc;
    if ((rc = m.get(&w.r, n, "r")) || (rc = m.get(&w.p, n, "p")) || (rc = m.get(&w.q, n, "q")) ||
        (rc = m.get(&w.sino, ns, "sinogram")) || (rc = m.get(&w.part, NP64 + 8, "partials")) ||
        (rc = m.get(&w.state, 1, "state")))
        return rc;
    return FL_OK;
}

// `iters` CG steps on (A^T A + lam I) u = b from the u given (solveCGLS.m:7-30),
// FP64 pair; log (device, iters+1 doubles) optional
int cg64_run(const fl_plan_s* p, Cg64Ws& w, const double* b, double* u, double lam, int iters, double tol,
             double* log, hipStream_t st)
{
    const size_t n = (size_t)p->nvox * p->nt;
    const unsigned g = grid64(n);
    int rc;
    if ((rc = launch_fwd(p, u, w.sino, st)) || (rc = launch_adj(p, w.sino, w.q, st))) return rc;
    hipLaunchKernelGGL(k64_init, dim3(g), dim3(256), 0, st, w.q, u, b, lam, w.r, w.p, n, w.part);
    hipLaunchKernelGGL(k64_fin_gamma, dim3(1), dim3(256), 0, st, w.part, (int)g, w.state, tol, log);
    CHK(hipGetLastError());
    for (int k = 0; k < iters; k++) {
        if ((rc = launch_fwd(p, w.p, w.sino, st)) || (rc = launch_adj(p, w.sino, w.q, st))) return rc;
        hipLaunchKernelGGL(k64_pq, dim3(g), dim3(256), 0, st, w.q, w.p, lam, n, w.part, w.state);
        hipLaunchKernelGGL(k64_fin_alpha, dim3(1), dim3(256), 0, st, w.part, (int)g, w.state);
        hipLaunchKernelGGL(k64_update, dim3(g), dim3(256), 0, st, u, w.r, w.p, w.q, n, w.part, w.state);
        hipLaunchKernelGGL(k64_fin_beta, dim3(1), dim3(256), 0, st, w.part, (int)g, w.state, log);
        hipLaunchKernelGGL(k64_p, dim3(g), dim3(256), 0, st, w.p, w.r, n, w.state);
        CHK(hipGetLastError());
    }
    return FL_OK;
}

// ------------------------------------------------ FP64 framelets (level 1)
//
// The undecimated frame of GenerateFrameletFilter.m with circular
// correlation (ConvSymAsym3D.m:19-24), every value in double:
//   3-D band (i,j,k) = corr_z[D_k] corr_x[D_j] corr_y[D_i] u   (FraDecMultiLevel3D.m:35-47)
//   2-D band (i,j)   = corr_y[D_j] corr_x[D_i] u               (FraDecMultiLevel2D.m:43-51)
//   band b = i + NF j (+ NF^2 k); synthesis the transpose with R = flip(D)
//   (FraRecMultiLevel{3D,2D}.m), by per-voxel gathers.

struct Taps64 {
    double D[5][5], R[5][5];
};

__device__ __forceinline__ int wrap64(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// all NB bands of voxel (x, y, z) of u (dim 3) or pixel (x, y) (dim 2)
template <int DIM, int NF, int NT>
__device__ __forceinline__ void bands64(const double* __restrict__ u, int nx, int ny, int nz, int x, int y, int z,
                                        const Taps64& T, double* out)
{
    constexpr int C = (NT - 1) / 2;
    if (DIM == 3) {
        double nb[NT][NT][NT];
        for (int kz = 0; kz < NT; kz++)
            for (int ky = 0; ky < NT; ky++)
                for (int kx = 0; kx < NT; kx++)
                    nb[kz][ky][kx] = u[((size_t)wrap64(z + kz - C, nz) * ny + wrap64(y + ky - C, ny)) * nx +
                                       wrap64(x + kx - C, nx)];
        for (int i = 0; i < NF; i++) {
            double ty[NT][NT];   // [kz][kx]: y pass with D_i
            for (int kz = 0; kz < NT; kz++)
                for (int kx = 0; kx < NT; kx++) {
                    double s = 0.0;
                    for (int ky = 0; ky < NT; ky++) s += T.D[i][ky] * nb[kz][ky][kx];
                    ty[kz][kx] = s;
                }
            for (int j = 0; j < NF; j++) {
                double tx[NT];   // [kz]: x pass with D_j
                for (int kz = 0; kz < NT; kz++) {
                    double s = 0.0;
                    for (int kx = 0; kx < NT; kx++) s += T.D[j][kx] * ty[kz][kx];
                    tx[kz] = s;
                }
                for (int k = 0; k < NF; k++) {
                    double s = 0.0;
                    for (int kz = 0; kz < NT; kz++) s += T.D[k][kz] * tx[kz];
                    out[i + NF * j + NF * NF * k] = s;
                }
            }
        }
    } else {
        double nb[NT][NT];
        for (int ky = 0; ky < NT; ky++)
            for (int kx = 0; kx < NT; kx++)
                nb[ky][kx] = u[(size_t)wrap64(y + ky - C, ny) * nx + wrap64(x + kx - C, nx)];
        for (int i = 0; i < NF; i++) {
            double tx[NT];   // [ky]: x pass with D_i
            for (int ky = 0; ky < NT; ky++) {
                double s = 0.0;
                for (int kx = 0; kx < NT; kx++) s += T.D[i][kx] * nb[ky][kx];
                tx[ky] = s;
            }
            for (int j = 0; j < NF; j++) {
                double s = 0.0;
                for (int ky = 0; ky < NT; ky++) s += T.D[j][ky] * tx[ky];
                out[i + NF * j] = s;
            }
        }
    }
}

template <int DIM, int NF, int NT>
__global__ __launch_bounds__(256) void k64_fwt(const double* __restrict__ u, double* __restrict__ C, int nx, int ny,
                                              int nz, Taps64 T)
{
    constexpr int NB = DIM == 3 ? NF * NF * NF : NF * NF;
    const size_t n = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
    double w[NB];
    bands64<DIM, NF, NT>(u, nx, ny, nz, x, y, z, T, w);
    for (int b = 0; b < NB; b++) C[(size_t)b * n + idx] = w[b];
}

// out = alpha * W^T C + beta * base (base may alias out)
template <int DIM, int NF, int NT>
__global__ __launch_bounds__(256) void k64_iwt(const double* __restrict__ C, double* out, int nx, int ny, int nz,
                                              Taps64 T, double alpha, double beta, const double* base)
{
    constexpr int CC = (NT - 1) / 2;
    const size_t n = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
    double acc = 0.0;
    if (DIM == 3) {
        // u = sum_i corr_y[R_i] sum_j corr_x[R_j] sum_k corr_z[R_k] C_ijk
        for (int i = 0; i < NF; i++)
            for (int j = 0; j < NF; j++)
                for (int k = 0; k < NF; k++) {
                    const double* cb = C + (size_t)(i + NF * j + NF * NF * k) * n;
                    for (int ky = 0; ky < NT; ky++) {
                        if (T.R[i][ky] == 0.0) continue;
                        const int yy = wrap64(y + ky - CC, ny);
                        for (int kx = 0; kx < NT; kx++) {
                            if (T.R[j][kx] == 0.0) continue;
                            const int xx = wrap64(x + kx - CC, nx);
                            double s = 0.0;
                            for (int kz = 0; kz < NT; kz++)
                                s += T.R[k][kz] * cb[((size_t)wrap64(z + kz - CC, nz) * ny + yy) * nx + xx];
                            acc += T.R[i][ky] * T.R[j][kx] * s;
                        }
                    }
                }
    } else {
        // u = sum_i corr_x[R_i] sum_j corr_y[R_j] C_ij
        for (int i = 0; i < NF; i++)
            for (int j = 0; j < NF; j++) {
                const double* cb = C + (size_t)(i + NF * j) * n;
                for (int kx = 0; kx < NT; kx++) {
                    if (T.R[i][kx] == 0.0) continue;
                    const int xx = wrap64(x + kx - CC, nx);
                    double s = 0.0;
                    for (int ky = 0; ky < NT; ky++) s += T.R[j][ky] * cb[(size_t)wrap64(y + ky - CC, ny) * nx + xx];
                    acc += T.R[i][kx] * s;
                }
            }
    }
    double v = alpha * acc;
    if (base) v += beta * base[idx];
    out[idx] = v;
}

// Split-Bregman coefficient step (solveSplitBregman.m:28-36) in double:
// t = v + Wu, alpha = shrink(t, tau) (low-pass copied), v <- v + (Wu - alpha),
// d = alpha - v; partials of |Wu - alpha|^2
template <int DIM, int NF, int NT>
__global__ __launch_bounds__(256) void k64_bregman(const double* __restrict__ u, double* __restrict__ V,
                                                  double* __restrict__ Dc, int nx, int ny, int nz, Taps64 T, double tau,
                                                  double* __restrict__ part)
{
    constexpr int NB = DIM == 3 ? NF * NF * NF : NF * NF;
    const size_t n = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    double r2 = 0.0;
    if (idx < n) {
        const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
        double w[NB];
        bands64<DIM, NF, NT>(u, nx, ny, nz, x, y, z, T, w);
        for (int b = 0; b < NB; b++) {
            const size_t o = (size_t)b * n + idx;
            const double v = V[o], t = v + w[b];
            double a = t;
            if (b != 0) {
                const double m = fabs(t) - tau;
                a = m > 0.0 ? copysign(m, t) : 0.0;
            }
            const double e = w[b] - a, vn = v + e;
            r2 += e * e;
            V[o] = vn;
            Dc[o] = a - vn;
        }
    }
    part_store(r2, part);
}

__global__ __launch_bounds__(256) void k64_sum_to(const double* __restrict__ part, int np, double* out)
{
    const double v = part_sum(part, np);
    if (threadIdx.x == 0) *out = v;
}

#define FRAME64_SWITCH(dim, kind, CALL)                                                        \
    do {                                                                                       \
        if (dim == 3) {                                                                        \
            constexpr int DIM = 3;                                                             \
            if (kind == FL_FRAME_HAAR) { constexpr int NF = 2, NT = 3; CALL; }                \
            else if (kind == FL_FRAME_LINEAR) { constexpr int NF = 3, NT = 3; CALL; }         \
            else { constexpr int NF = 5, NT = 5; CALL; }                                       \
        } else {                                                                               \
            constexpr int DIM = 2;                                                             \
            if (kind == FL_FRAME_HAAR) { constexpr int NF = 2, NT = 3; CALL; }                \
            else if (kind == FL_FRAME_LINEAR) { constexpr int NF = 3, NT = 3; CALL; }         \
            else { constexpr int NF = 5, NT = 5; CALL; }                                       \
        }                                                                                      \
    } while (0)

}  // namespace

extern "C" int fl_cgls_f64(fl_plan p, const double* d_b, double* d_u, double lam, int iters, double tol,
                           double* gamma_log, int* iters_done, void* stream)
{
    if (!p || !d_b || !d_u || iters < 0) return fl_fail(FL_ERR_INVALID, "bad FP64 CGLS arguments");
    int rc = ensure_tables(p);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)p->nvox * p->nt, ns = (size_t)p->ndet * p->nv;
    DevMem m;
    Cg64Ws w;
    double* log;
    if ((rc = cg64_alloc(m, w, n, ns)) || (rc = m.get(&log, (size_t)iters + 1, "log"))) return rc;
    if ((rc = cg64_run(p, w, d_b, d_u, lam, iters, tol, log, st))) return rc;
    Cg64 s{};
    CHK(hipMemcpyAsync(&s, w.state, sizeof s, hipMemcpyDeviceToHost, st));
    if (gamma_log) CHK(hipMemcpyAsync(gamma_log, log, ((size_t)iters + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
    // the workspace is freed on return: wait for the solve
    CHK(hipStreamSynchronize(st));
    if (iters_done) *iters_done = s.iter;
    if (gamma_log)
        for (int k = s.iter + 1; k <= iters; k++) gamma_log[k] = s.gamma;
    return fl_ok();
}

extern "C" int fl_split_bregman_f64(fl_plan p, int kind, const double* d_f, double* d_u, double lam, double mu,
                                    int outer, int inner, double* resid_log, double* d_u_hist, void* stream)
{
    if (!p || !d_f || !d_u || outer < 0 || inner < 0 || !(mu > 0))
        return fl_fail(FL_ERR_INVALID, "bad FP64 split-Bregman arguments");
    if (kind != FL_FRAME_HAAR && kind != FL_FRAME_LINEAR && kind != FL_FRAME_CUBIC)
        return fl_fail(FL_ERR_INVALID, "bad framelet kind");
    if (p->nt != 1) return fl_fail(FL_ERR_UNSUPPORTED, "split Bregman on one frame");
    int rc = ensure_tables(p);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int dim = p->cone ? 3 : 2;
    const int nx = p->fg.nx, ny = p->fg.ny, nz = p->cone ? p->fg.nz : 1;
    const int nf = kind == FL_FRAME_HAAR ? 2 : kind == FL_FRAME_LINEAR ? 3 : 5;
    const int nb = dim == 3 ? nf * nf * nf : nf * nf;
    const size_t n = (size_t)p->nvox, ns = (size_t)p->ndet * p->nv;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    Taps64 T;
    fl_frame_taps_f64(kind, T.D, T.R);
    DevMem m;
    Cg64Ws w;
    double *V, *Dc, *Atf, *b, *fpart, *res;
    if ((rc = cg64_alloc(m, w, n, ns)) || (rc = m.get(&V, (size_t)nb * n, "v")) ||
        (rc = m.get(&Dc, (size_t)nb * n, "d")) || (rc = m.get(&Atf, n, "Atf")) || (rc = m.get(&b, n, "b")) ||
        (rc = m.get(&fpart, blocks, "framelet partials")) || (rc = m.get(&res, (size_t)(outer > 0 ? outer : 1), "residuals")))
        return rc;
    // v = alpha = W u0 (so alpha - v = 0 and b = A^T f)
    FRAME64_SWITCH(dim, kind, hipLaunchKernelGGL((k64_fwt<DIM, NF, NT>), dim3(blocks), dim3(256), 0, st, d_u, V, nx,
                                                 ny, nz, T));
    CHK(hipGetLastError());
    if ((rc = launch_adj(p, d_f, Atf, st))) return rc;
    CHK(hipMemcpyAsync(b, Atf, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < outer; k++) {
        if ((rc = cg64_run(p, w, b, d_u, mu, inner, 0.0, nullptr, st))) return rc;
        if (d_u_hist) CHK(hipMemcpyAsync(d_u_hist + (size_t)k * n, d_u, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        FRAME64_SWITCH(dim, kind, hipLaunchKernelGGL((k64_bregman<DIM, NF, NT>), dim3(blocks), dim3(256), 0, st, d_u,
                                                     V, Dc, nx, ny, nz, T, lam / mu, fpart));
        hipLaunchKernelGGL(k64_sum_to, dim3(1), dim3(256), 0, st, fpart, (int)blocks, res + k);
        // b = A^T f + mu W^T (alpha - v)
        FRAME64_SWITCH(dim, kind, hipLaunchKernelGGL((k64_iwt<DIM, NF, NT>), dim3(blocks), dim3(256), 0, st, Dc, b,
                                                     nx, ny, nz, T, mu, 1.0, Atf));
        CHK(hipGetLastError());
    }
    if (resid_log && outer > 0)
        CHK(hipMemcpyAsync(resid_log, res, (size_t)outer * sizeof(double), hipMemcpyDeviceToHost, st));
    // the workspace is freed on return: wait for the solve
    CHK(hipStreamSynchronize(st));
    return fl_ok();
}
