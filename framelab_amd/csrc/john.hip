// john.hip -- discretised John's equation on multi-helix cone-beam data.
//
// The normalised sinogram g (detector column a, row b, view angle theta,
// helix zeta; stored a fastest: g[((z*nv + t)*nb + b)*na + a]) of a
// multi-helix scan satisfies John's ultrahyperbolic equation.  Its
// discretisation (classes/+DataTypes/@CTData/applyJohn.m:38-106) at interior
// point (A, B, T, Z) is
//   R g_tb - R SO g_az - R h g_bz + 2 a g_b + a b g_bb + (a^2 + R^2) g_ab,
// with central differences (steps da, db, dtheta; dzeta across helices, a
// forward difference when there are only two) -- a 21-point stencil whose
// coefficients depend on a and b.  Output point (p, q, r, s) is interior
// point (p+1, q+1, r+1, s+1), or (p+1, q+1, r+1, 0) with two helices, and
// is stored p fastest on the (na-2, nb-2, nv-2, nz-2 or 1) grid.
//
// D gathers its stencil per output point; D^T gathers, per input point,
// the stencil entries of every output point that reads it (no atomics).
// FP64 arithmetic.
#include <hip/hip_runtime.h>

#include <new>

#include "../../include/framelab_hip.h"
#include "fl_capi_internal.h"
#include "fl_common.h"

struct fl_john_s {
    fl_john_geometry g;
    double *d_a = nullptr, *d_b = nullptr;
    int nzo = 0;   // output helix planes
};

namespace {

struct JohnK {
    int na, nb, nv, nz, nzo;
    double c1, c2, c3;          // R/(4 db dt), -R SO/(2 da dz'), -R h/(2 db dz')
    double inv_db, inv_db2, c6; // 1/db, 1/db^2, 1/(4 da db)
    double R2;
    int zp, zm;                 // helix offsets of the z difference (+1/-1, or +1/0 with two helices)
    const double *a, *b;
};

// Stencil entry: offsets (dA, dB, dT, dZ selector), term, weight.
// dZ: 0 = same plane, 1 = the z+ plane, 2 = the z- plane.
struct Ent {
    signed char dA, dB, dT, dZ, term, w;
};

__constant__ Ent kEnt[21] = {
    // T1: R g_tb
    {0, 1, 1, 0, 1, 1}, {0, -1, 1, 0, 1, -1}, {0, 1, -1, 0, 1, -1}, {0, -1, -1, 0, 1, 1},
    // T2: -R SO g_az
    {1, 0, 0, 1, 2, 1}, {-1, 0, 0, 1, 2, -1}, {1, 0, 0, 2, 2, -1}, {-1, 0, 0, 2, 2, 1},
    // T3: -R h g_bz
    {0, 1, 0, 1, 3, 1}, {0, -1, 0, 1, 3, -1}, {0, 1, 0, 2, 3, -1}, {0, -1, 0, 2, 3, 1},
    // T4: 2 a g_b
    {0, 1, 0, 0, 4, 1}, {0, -1, 0, 0, 4, -1},
    // T5: a b g_bb
    {0, 1, 0, 0, 5, 1}, {0, 0, 0, 0, 5, -2}, {0, -1, 0, 0, 5, 1},
    // T6: (a^2 + R^2) g_ab
    {1, 1, 0, 0, 6, 1}, {-1, 1, 0, 0, 6, -1}, {1, -1, 0, 0, 6, -1}, {-1, -1, 0, 0, 6, 1},
};

__device__ __forceinline__ double term_coef(const JohnK& k, int term, int A, int B)
{
    switch (term) {
    case 1: return k.c1;
    case 2: return k.c2;
    case 3: return k.c3;
    case 4: return k.a[A] * k.inv_db;
    case 5: return k.a[A] * k.b[B] * k.inv_db2;
    default: return (k.a[A] * k.a[A] + k.R2) * k.c6;
    }
}

__device__ __forceinline__ int zoff(const JohnK& k, int sel) { return sel == 0 ? 0 : (sel == 1 ? k.zp : k.zm); }

template <typename GT>
__global__ __launch_bounds__(256) void k_john_apply(const GT* __restrict__ g, double* __restrict__ out, JohnK k)
{
    const size_t nA = k.na - 2, nB = k.nb - 2, nT = k.nv - 2;
    const size_t n = nA * nB * nT * (size_t)k.nzo;
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < n; o += (size_t)gridDim.x * 256) {
        size_t v = o;
        const int A = (int)(v % nA) + 1; v /= nA;
        const int B = (int)(v % nB) + 1; v /= nB;
        const int T = (int)(v % nT) + 1; v /= nT;
        const int Z = k.nz > 2 ? (int)v + 1 : 0;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < 21; e++) {
            const Ent en = kEnt[e];
            const size_t gi = (((size_t)(Z + zoff(k, en.dZ)) * k.nv + (T + en.dT)) * k.nb + (B + en.dB)) * k.na + (A + en.dA);
            acc += term_coef(k, en.term, A, B) * (double)en.w * (double)g[gi];
        }
        out[o] = acc;
    }
}

__global__ __launch_bounds__(256) void k_john_adjoint(const double* __restrict__ u, double* __restrict__ out, JohnK k)
{
    const size_t nA = k.na - 2, nB = k.nb - 2, nT = k.nv - 2;
    const size_t n = (size_t)k.na * k.nb * k.nv * k.nz;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        size_t v = i;
        const int Xa = (int)(v % k.na); v /= k.na;
        const int Xb = (int)(v % k.nb); v /= k.nb;
        const int Xt = (int)(v % k.nv); v /= k.nv;
        const int Xz = (int)v;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < 21; e++) {
            const Ent en = kEnt[e];
            // the output point whose entry e reads this point
            const int A = Xa - en.dA, B = Xb - en.dB, T = Xt - en.dT, Z = Xz - zoff(k, en.dZ);
            if (A < 1 || A > k.na - 2 || B < 1 || B > k.nb - 2 || T < 1 || T > k.nv - 2) continue;
            int s;
            if (k.nz > 2) {
                if (Z < 1 || Z > k.nz - 2) continue;
                s = Z - 1;
            } else {
                if (Z != 0) continue;
                s = 0;
            }
            const size_t o = (((size_t)s * nT + (T - 1)) * nB + (B - 1)) * nA + (A - 1);
            acc += term_coef(k, en.term, A, B) * (double)en.w * u[o];
        }
        out[i] = acc;
    }
}

JohnK make_k(const fl_john_s* J)
{
    const fl_john_geometry& g = J->g;
    JohnK k;
    k.na = g.na; k.nb = g.nb; k.nv = g.nv; k.nz = g.nz; k.nzo = J->nzo;
    const double dz = g.nz > 2 ? 2.0 * g.dzeta : g.dzeta;
    k.c1 = g.R / (4.0 * g.db * g.dtheta);
    k.c2 = -g.R * g.SO / (2.0 * g.da * dz);
    k.c3 = -g.R * g.h / (2.0 * g.db * dz);
    // 2 a g_b with g_b = (g+ - g-)/(2 db): coefficient a/db
    k.inv_db = 1.0 / g.db;
    k.inv_db2 = 1.0 / (g.db * g.db);
    k.c6 = 1.0 / (4.0 * g.da * g.db);
    k.R2 = g.R * g.R;
    k.zp = 1;
    k.zm = g.nz > 2 ? -1 : 0;
    k.a = J->d_a;
    k.b = J->d_b;
    return k;
}

unsigned grid_of(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

}  // namespace

extern "C" int fl_john_create(fl_john* out, const fl_john_geometry* g)
{
    if (!out || !g || !g->a || !g->b) return fl_fail(FL_ERR_INVALID, "null John geometry");
    *out = nullptr;
    if (g->na < 3 || g->nb < 3 || g->nv < 3 || g->nz < 2)
        return fl_fail(FL_ERR_INVALID, "John's equation needs na, nb, nv >= 3 and at least two helices");
    if (!(g->da > 0) || !(g->db > 0) || !(g->dtheta > 0) || g->dzeta == 0)
        return fl_fail(FL_ERR_INVALID, "John steps must be non-zero (da, db, dtheta > 0)");
    fl_john_s* J = new (std::nothrow) fl_john_s();
    if (!J) return fl_fail(FL_ERR_ALLOC, "host allocation");
    J->g = *g;
    J->g.a = nullptr;
    J->g.b = nullptr;
    J->nzo = g->nz > 2 ? g->nz - 2 : 1;
    hipError_t e = hipMalloc((void**)&J->d_a, g->na * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&J->d_b, g->nb * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(J->d_a, g->a, g->na * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(J->d_b, g->b, g->nb * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (J->d_a) (void)hipFree(J->d_a);
        if (J->d_b) (void)hipFree(J->d_b);
        delete J;
        return fl_hip_fail(e, "John coordinate tables");
    }
    *out = J;
    return fl_ok();
}

extern "C" int fl_john_destroy(fl_john J)
{
    if (!J) return fl_ok();
    (void)hipFree(J->d_a);
    (void)hipFree(J->d_b);
    delete J;
    return fl_ok();
}

extern "C" int fl_john_sizes(fl_john J, long long* n_in, long long* n_out)
{
    if (!J) return fl_fail(FL_ERR_INVALID, "null John operator");
    const fl_john_geometry& g = J->g;
    if (n_in) *n_in = (long long)g.na * g.nb * g.nv * g.nz;
    if (n_out) *n_out = (long long)(g.na - 2) * (g.nb - 2) * (g.nv - 2) * J->nzo;
    return fl_ok();
}

extern "C" int fl_john_apply(fl_john J, const void* d_g, int g_dtype, double* d_out, void* stream)
{
    if (!J || !d_g || !d_out) return fl_fail(FL_ERR_INVALID, "null John operator or buffer");
    const JohnK k = make_k(J);
    long long no = 0;
    fl_john_sizes(J, nullptr, &no);
    hipStream_t st = (hipStream_t)stream;
    if (g_dtype == FL_F32)
        hipLaunchKernelGGL(k_john_apply<float>, dim3(grid_of(no)), dim3(256), 0, st, (const float*)d_g, d_out, k);
    else if (g_dtype == FL_F64)
        hipLaunchKernelGGL(k_john_apply<double>, dim3(grid_of(no)), dim3(256), 0, st, (const double*)d_g, d_out, k);
    else
        return fl_fail(FL_ERR_INVALID, "John input is FP32 or FP64");
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fl_ok() : fl_hip_fail(e, "John apply");
}

extern "C" int fl_john_adjoint(fl_john J, const double* d_u, double* d_out, void* stream)
{
    if (!J || !d_u || !d_out) return fl_fail(FL_ERR_INVALID, "null John operator or buffer");
    const JohnK k = make_k(J);
    long long ni = 0;
    fl_john_sizes(J, &ni, nullptr);
    hipLaunchKernelGGL(k_john_adjoint, dim3(grid_of(ni)), dim3(256), 0, (hipStream_t)stream, d_u, d_out, k);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? fl_ok() : fl_hip_fail(e, "John adjoint");
}

extern "C" int fl_john_normal(fl_john J, const void* d_g, int g_dtype, double* d_work, double* d_out, void* stream)
{
    if (!d_work) return fl_fail(FL_ERR_INVALID, "John normal operator needs a work buffer");
    int rc = fl_john_apply(J, d_g, g_dtype, d_work, stream);
    return rc ? rc : fl_john_adjoint(J, d_work, d_out, stream);
}
