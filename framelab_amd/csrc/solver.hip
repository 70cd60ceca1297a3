// solver.hip -- device-resident CGLS and split Bregman, gradients of the
// reconstruction objective, and the framelet C ABI.
//
// CGLS follows matlab/classes/+Optimizers/@CGLS/solveCGLS.m:7-30 (CG on
// (A^T A + lam I) u = b, r = A_loc(u0) - b, p = -r, stop when r.r < tol);
// split Bregman follows matlab/classes/+Optimizers/@SplitBregman/
// solveSplitBregman.m:16-41 (class form v <- v + (Wu - alpha)).
//
// Vectors are FP32 (optionally BF16/FP16 copies of the projector input and
// the sinogram); every dot product, alpha, beta, residual norm and stopping
// test is FP64 and stays on the GPU: each reduction writes one FP64 partial
// per workgroup and a single-workgroup finalize kernel sums them in a fixed
// order, so results are deterministic and no host round trip is needed
// inside the loop.  One CG iteration can be captured in a hipGraph and
// replayed; after the stopping test fires, the remaining replays leave the
// iterate unchanged (every update kernel reads the device stop flag).
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/framelab_hip.h"
#include "fl_capi_internal.h"
#include "fl_common.h"
#include "fl_framelet.h"

namespace {

constexpr int NPART = 2048;   // workgroups of the reduction kernels

struct CgState {
    double gamma, alpha, beta, pq, tol;
    int stop, iter;
};

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void store_partial(double v, double* __restrict__ part)
{
    __shared__ double s[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

// fixed-order sum of np partials by one 256-thread workgroup
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int np)
{
    __shared__ double s[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) v += part[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

template <typename LT>
__global__ __launch_bounds__(256) void k_cg_init(const float* __restrict__ q, const float* __restrict__ u,
                                                const float* __restrict__ b, float lam, float* __restrict__ r,
                                                float* __restrict__ p, LT* __restrict__ plo, size_t n,
                                                double* __restrict__ part)
{
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float ri = __builtin_fmaf(lam, u[i], q[i]) - b[i];
        r[i] = ri;
        p[i] = -ri;
        if (plo) fl_store(plo, i, -ri);
        acc += (double)ri * (double)ri;
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_fin_gamma(const double* __restrict__ part, int np, CgState* st,
                                                  double* __restrict__ log)
{
    double g = sum_partials(part, np);
    if (threadIdx.x == 0) {
        st->gamma = g;
        st->iter = 0;
        st->stop = (g < st->tol) ? 1 : 0;
        if (log) log[0] = g;
    }
}

__global__ __launch_bounds__(256) void k_cg_pq(float* __restrict__ q, const float* __restrict__ p, float lam,
                                              size_t n, double* __restrict__ part, const CgState* st)
{
    double acc = 0.0;
    if (!st->stop) {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
            float qi = __builtin_fmaf(lam, p[i], q[i]);
            q[i] = qi;
            acc += (double)p[i] * (double)qi;
        }
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_fin_alpha(const double* __restrict__ part, int np, CgState* st)
{
    double pq = sum_partials(part, np);
    if (threadIdx.x == 0 && !st->stop) {
        st->pq = pq;
        if (pq == 0.0) {
            st->stop = 1;      // p = 0: converged exactly, alpha would be 0/0
            st->alpha = 0.0;
        } else {
            st->alpha = st->gamma / pq;
        }
    }
}

__global__ __launch_bounds__(256) void k_cg_update(float* __restrict__ u, float* __restrict__ r,
                                                  const float* __restrict__ p, const float* __restrict__ q,
                                                  size_t n, double* __restrict__ part, const CgState* st)
{
    double acc = 0.0;
    if (!st->stop) {
        const float a = (float)st->alpha;
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
            u[i] = __builtin_fmaf(a, p[i], u[i]);
            float ri = __builtin_fmaf(a, q[i], r[i]);
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_fin_beta(const double* __restrict__ part, int np, CgState* st,
                                                 double* __restrict__ log, int maxlog)
{
    double g = sum_partials(part, np);
    if (threadIdx.x == 0 && !st->stop) {
        st->beta = g / st->gamma;
        st->gamma = g;
        st->iter += 1;
        if (log && st->iter < maxlog) log[st->iter] = g;
        if (g < st->tol) st->stop = 1;
    }
}

template <typename LT>
__global__ __launch_bounds__(256) void k_cg_p(float* __restrict__ p, const float* __restrict__ r,
                                             LT* __restrict__ plo, size_t n, const CgState* st)
{
    if (st->stop) return;   // converged: the iterate no longer changes
    const float bt = (float)st->beta;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float pi = __builtin_fmaf(bt, p[i], -r[i]);
        p[i] = pi;
        if (plo) fl_store(plo, i, pi);
    }
}

template <typename LT>
__global__ __launch_bounds__(256) void k_cast(const float* __restrict__ x, LT* __restrict__ y, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) fl_store(y, i, x[i]);
}

__global__ void k_state_init(CgState* st, double tol)
{
    st->gamma = 0.0; st->alpha = 0.0; st->beta = 0.0; st->pq = 0.0;
    st->tol = tol; st->stop = 0; st->iter = 0;
}

// residual for the data gradient: s <- s - b; partials of 0.5*|s|^2
__global__ __launch_bounds__(256) void k_resid(float* __restrict__ s, const float* __restrict__ b, size_t n,
                                              double* __restrict__ part)
{
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float e = s[i] - b[i];
        s[i] = e;
        acc += 0.5 * (double)e * (double)e;
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_sum_to(const double* __restrict__ part, int np, double* out)
{
    double v = sum_partials(part, np);
    if (threadIdx.x == 0) *out = v;
}

// CGLS vector steps with host-supplied scalars (CGLS over caller-supplied
// operators): the same element arithmetic as the plan solver's kernels.
__global__ __launch_bounds__(256) void k_gv_pq(float* __restrict__ q, const float* __restrict__ p, float lam, size_t n,
                                              double* __restrict__ part)
{
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float qi = __builtin_fmaf(lam, p[i], q[i]);
        q[i] = qi;
        acc += (double)p[i] * (double)qi;
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_gv_update(float* __restrict__ u, float* __restrict__ r,
                                                  const float* __restrict__ p, const float* __restrict__ q, float a,
                                                  size_t n, double* __restrict__ part)
{
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        u[i] = __builtin_fmaf(a, p[i], u[i]);
        const float ri = __builtin_fmaf(a, q[i], r[i]);
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    store_partial(acc, part);
}

__global__ __launch_bounds__(256) void k_gv_dir(float* __restrict__ p, const float* __restrict__ r, float bt, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = __builtin_fmaf(bt, p[i], -r[i]);
}

// soft threshold of every band but band 0 of each level, in place
// (softThreshold.m:15-80: sign(a) max(|a| - tau, 0), low-pass kept)
template <typename CT>
__global__ __launch_bounds__(256) void k_shrink_bands(CT* __restrict__ c, size_t band, int nbands, size_t total, float tau)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        if ((i / band) % (size_t)nbands == 0) continue;
        const float a = fl_load(c, i);
        const float m = fabsf(a) - tau;
        fl_store(c, i, m > 0.0f ? copysignf(m, a) : 0.0f);
    }
}

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b < (size_t)NPART ? (b ? b : 1) : NPART);
}

// ---- workspace -------------------------------------------------------------

struct Workspace {
    float *r = nullptr, *p = nullptr, *q = nullptr, *sino = nullptr;
    float* u = nullptr;      // the iterate while a solve runs (graphs capture this address)
    void* plo = nullptr;
    void* slo = nullptr;
    double* part = nullptr;
    CgState* st = nullptr;
    double* log = nullptr;
    int maxlog = 0;
    // split Bregman
    void* V = nullptr;
    void* D = nullptr;
    double* fpart = nullptr;
    size_t fpart_n = 0;
    double* res = nullptr;     // per outer step |Wu - alpha|^2
    size_t res_n = 0;
    float *Atf = nullptr, *b = nullptr;
    float* lp = nullptr;       // two low-pass bands of W u (levels > 1)
    size_t nb_bands = 0;       // coefficient elements of V and D
    int coef_type = FL_F32;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // cached graph of one CG iteration
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    float gkey_lam = 0.0f;
    int gkey_flags = -1;
};

struct PlanWs {
    Workspace w;
    std::mutex mu;
};

int ws_alloc(void** p, size_t bytes, const char* what)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    return e == hipSuccess ? FL_OK : fl_fail(FL_ERR_ALLOC, "hipMalloc(%s, %zu bytes)", what, bytes);
}

int lowp_type(int flags)
{
    if (flags & FL_SOLVE_LOWP_BF16) return FL_BF16;
    if (flags & FL_SOLVE_LOWP_F16) return FL_F16;
    return FL_F32;
}

void drop_graph(Workspace& w)
{
    if (w.gexec) (void)hipGraphExecDestroy(w.gexec);
    if (w.graph) (void)hipGraphDestroy(w.graph);
    w.gexec = nullptr;
    w.graph = nullptr;
    w.gkey_flags = -1;
}

int ensure_cg(fl_plan_s* pl, Workspace& w, int maxlog, int flags)
{
    const size_t nx = (size_t)pl->nvox * pl->nt, ns = (size_t)pl->ndet * pl->nv;
    int rc;
    if (!w.r) {
        if ((rc = ws_alloc((void**)&w.r, nx * 4, "r")) || (rc = ws_alloc((void**)&w.p, nx * 4, "p")) ||
            (rc = ws_alloc((void**)&w.q, nx * 4, "q")) || (rc = ws_alloc((void**)&w.sino, ns * 4, "sinogram")) ||
            (rc = ws_alloc((void**)&w.u, nx * 4, "iterate")) ||
            (rc = ws_alloc((void**)&w.part, (NPART + 8) * sizeof(double), "partials")) ||
            (rc = ws_alloc((void**)&w.st, sizeof(CgState), "state")))
            return rc;
        hipError_t e = hipStreamCreate(&w.stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev_in, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev_out, hipEventDisableTiming);
        if (e != hipSuccess) return fl_hip_fail(e, "solver stream");
    }
    if (lowp_type(flags) != FL_F32 && !w.plo) {
        if ((rc = ws_alloc(&w.plo, nx * 2, "p (low precision)")) || (rc = ws_alloc(&w.slo, ns * 2, "sinogram (low precision)")))
            return rc;
    }
    if (maxlog > w.maxlog) {
        drop_graph(w);   // the captured iteration holds the old log pointer and length
        if (w.log) (void)hipFree(w.log);
        w.log = nullptr;
        if ((rc = ws_alloc((void**)&w.log, (size_t)maxlog * sizeof(double), "log"))) return rc;
        w.maxlog = maxlog;
    }
    return FL_OK;
}

void free_ws(Workspace& w)
{
    void* ptrs[] = {w.r, w.p, w.q, w.sino, w.u, w.plo, w.slo, w.part, w.st, w.log, w.V, w.D, w.Atf, w.b, w.fpart,
                    w.res, w.lp};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    drop_graph(w);
    if (w.ev_in) (void)hipEventDestroy(w.ev_in);
    if (w.ev_out) (void)hipEventDestroy(w.ev_out);
    if (w.stream) (void)hipStreamDestroy(w.stream);
    w = Workspace();
}

PlanWs* get_ws(fl_plan_s* p)
{
    if (!p->ws) p->ws = new PlanWs();
    return (PlanWs*)p->ws;
}

#define HIPCHK(expr)                                          \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return fl_hip_fail(e_, #expr);  \
    } while (0)

// The low-precision copy of p a solve in storage type LT maintains: none in
// FP32 (the buffer a low-precision solve on the same plan left holds two
// bytes per element, too small for floats).
template <typename LT>
LT* lowp_copy(Workspace& w)
{
    return std::is_same<LT, float>::value ? nullptr : (LT*)w.plo;
}

// q = A^T (A x) for x the FP32 vector (or its low-precision copy)
int normal_op(fl_plan_s* pl, Workspace& w, const float* x, const void* xlo, int flags, float* out, hipStream_t st)
{
    const int lt = lowp_type(flags);
    const int kind = (flags & FL_SOLVE_MATCHED) ? FL_ADJ_MATCHED : FL_ADJ_EXACT;
    int rc;
    if (lt == FL_F32) {
        if ((rc = fl_forward(pl, x, FL_F32, w.sino, FL_F32, st))) return rc;
        return fl_adjoint(pl, w.sino, FL_F32, out, FL_F32, kind, st);
    }
    if ((rc = fl_forward(pl, xlo, lt, w.slo, lt, st))) return rc;
    return fl_adjoint(pl, w.slo, lt, out, FL_F32, kind, st);
}

template <typename LT>
int cg_body_t(fl_plan_s* pl, Workspace& w, float* u, float lam, int flags, hipStream_t st)
{
    const size_t n = (size_t)pl->nvox * pl->nt;
    const unsigned g = grid_for(n);
    int rc;
    if ((rc = normal_op(pl, w, w.p, w.plo, flags, w.q, st))) return rc;
    hipLaunchKernelGGL(k_cg_pq, dim3(g), dim3(256), 0, st, w.q, w.p, lam, n, w.part, w.st);
    hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(256), 0, st, w.part, (int)g, w.st);
    hipLaunchKernelGGL(k_cg_update, dim3(g), dim3(256), 0, st, u, w.r, w.p, w.q, n, w.part, w.st);
    hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(256), 0, st, w.part, (int)g, w.st, w.log, w.maxlog);
    hipLaunchKernelGGL((k_cg_p<LT>), dim3(g), dim3(256), 0, st, w.p, w.r, lowp_copy<LT>(w), n, w.st);
    HIPCHK(hipGetLastError());
    return FL_OK;
}

int cg_body(fl_plan_s* pl, Workspace& w, float* u, float lam, int flags, hipStream_t st)
{
    switch (lowp_type(flags)) {
    case FL_BF16: return cg_body_t<bf16_t>(pl, w, u, lam, flags, st);
    case FL_F16: return cg_body_t<f16_t>(pl, w, u, lam, flags, st);
    default: return cg_body_t<float>(pl, w, u, lam, flags, st);
    }
}

template <typename LT>
int cg_init_t(fl_plan_s* pl, Workspace& w, const float* b, float* u, float lam, int flags, double tol,
              hipStream_t st)
{
    const size_t n = (size_t)pl->nvox * pl->nt;
    const unsigned g = grid_for(n);
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(1), 0, st, w.st, tol);
    int rc;
    if (lowp_type(flags) == FL_F32) {
        rc = normal_op(pl, w, u, nullptr, flags, w.q, st);
    } else {
        hipLaunchKernelGGL((k_cast<LT>), dim3(g), dim3(256), 0, st, u, lowp_copy<LT>(w), n);
        rc = normal_op(pl, w, nullptr, w.plo, flags, w.q, st);
    }
    if (rc) return rc;
    hipLaunchKernelGGL((k_cg_init<LT>), dim3(g), dim3(256), 0, st, w.q, u, b, lam, w.r, w.p, lowp_copy<LT>(w), n,
                       w.part);
    hipLaunchKernelGGL(k_fin_gamma, dim3(1), dim3(256), 0, st, w.part, (int)g, w.st, w.log);
    HIPCHK(hipGetLastError());
    return FL_OK;
}

int cg_init(fl_plan_s* pl, Workspace& w, const float* b, float* u, float lam, int flags, double tol, hipStream_t st)
{
    switch (lowp_type(flags)) {
    case FL_BF16: return cg_init_t<bf16_t>(pl, w, b, u, lam, flags, tol, st);
    case FL_F16: return cg_init_t<f16_t>(pl, w, b, u, lam, flags, tol, st);
    default: return cg_init_t<float>(pl, w, b, u, lam, flags, tol, st);
    }
}

// Runs `iters` CG iterations on w.stream (already ordered after the caller).
// The iterate lives in w.u while the solve runs, so one captured iteration
// serves every call with the same lam and flags.
int cg_run(fl_plan_s* pl, Workspace& w, const float* b, float* u_caller, float lam, int iters, double tol,
           int flags)
{
    hipStream_t st = w.stream;
    const size_t n = (size_t)pl->nvox * pl->nt;
    float* u = w.u;
    HIPCHK(fl_launch_copy(u_caller, u, n, st));
    int rc;
    if ((rc = cg_init(pl, w, b, u, lam, flags, tol, st))) return rc;
    if (iters <= 0) return FL_OK;
    if (flags & FL_SOLVE_GRAPH) {
        const bool hit = w.gexec && w.gkey_lam == lam && w.gkey_flags == flags;
        if (!hit) {
            drop_graph(w);
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = cg_body(pl, w, u, lam, flags, st);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(st, &gr);
            if (rc) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            if (e != hipSuccess) return fl_hip_fail(e, "hipStreamEndCapture");
            w.graph = gr;
            HIPCHK(hipGraphInstantiate(&w.gexec, w.graph, nullptr, nullptr, 0));
            w.gkey_lam = lam;
            w.gkey_flags = flags;
        }
        // The captured forwards write the plan's brick buffer, and a stream
        // under capture takes no part in the plan's forward ordering: order
        // the replays after any forward in flight on another stream, and
        // later forwards after the replays.
        std::lock_guard<std::mutex> lk(pl->fwd_mu);
        if (pl->fwd_done && pl->fwd_pending) HIPCHK(hipStreamWaitEvent(st, pl->fwd_done, 0));
        for (int k = 0; k < iters; k++) HIPCHK(hipGraphLaunch(w.gexec, st));
        if (pl->fwd_done) {
            HIPCHK(hipEventRecord(pl->fwd_done, st));
            pl->fwd_pending = true;
        }
    } else {
        for (int k = 0; k < iters; k++)
            if ((rc = cg_body(pl, w, u, lam, flags, st))) return rc;
    }
    HIPCHK(fl_launch_copy(u, u_caller, n, st));
    return FL_OK;
}

int begin_on(Workspace& w, hipStream_t caller)
{
    HIPCHK(hipEventRecord(w.ev_in, caller));
    HIPCHK(hipStreamWaitEvent(w.stream, w.ev_in, 0));
    return FL_OK;
}

int end_on(Workspace& w, hipStream_t caller)
{
    HIPCHK(hipEventRecord(w.ev_out, w.stream));
    HIPCHK(hipStreamWaitEvent(caller, w.ev_out, 0));
    return FL_OK;
}

// per-device scratch for the plan-free framelet calls
std::mutex g_scratch_mu;
std::vector<std::pair<void*, size_t>> g_scratch;

// Grow-only per-device buffer of doubles for the plan-free reductions.
// Calls that use it on one device are serialised by the caller's stream
// order; the buffer lives until the process ends.
double* device_scratch_doubles(size_t count)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1, {nullptr, 0});
    const size_t bytes = count * sizeof(double);
    if (g_scratch[dev].second < bytes) {
        if (g_scratch[dev].first) {
            (void)hipDeviceSynchronize();
            (void)hipFree(g_scratch[dev].first);
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            g_scratch[dev] = {nullptr, 0};
            return nullptr;
        }
        g_scratch[dev] = {p, bytes};
    }
    return (double*)g_scratch[dev].first;
}

bool frame_ok(int dim, int kind, int nx, int ny, int nz)
{
    if (dim != 2 && dim != 3) return false;
    if (kind != FL_FRAME_HAAR && kind != FL_FRAME_LINEAR && kind != FL_FRAME_CUBIC) return false;
    return nx > 0 && ny > 0 && (dim == 2 || nz > 0);
}

// g += W^T (lam psi'(W x)) over `level` levels, loss partials in part
// (nbk per level) and their fixed-order sum in part[nbk * level].  work:
// level blocks of NB bands.  Level L analyses the low-pass band of level
// L-1, which the analysis of level L-1 leaves in its band 0 (unpenalised);
// synthesis then replaces each such band 0 by the reconstruction of the level
// above (FraRecMultiLevel3D.m:14-17), which is the chain rule's backward.
int grad_framelet_run(const FlFrameDesc& f, int level, const float* x, float lam, float eps, int with_lowpass,
                      float* g, float* work, double* part, unsigned nbk, hipStream_t st)
{
    const size_t n = (size_t)f.nx * f.ny * f.nz;
    const int nb = fl_framelet_num_bands(f.dim, f.kind);
    hipError_t e = hipSuccess;
    for (int L = 1; L <= level && e == hipSuccess; L++) {
        const float* src = L == 1 ? x : work + (size_t)(L - 2) * nb * n;
        e = fl_launch_l1grad(f, src, work + (size_t)(L - 1) * nb * n, lam, eps, with_lowpass, L, L < level ? 1 : 0,
                             part + (size_t)(L - 1) * nbk, st);
    }
    for (int L = level; L >= 2 && e == hipSuccess; L--)
        e = fl_launch_iwt(f, work + (size_t)(L - 1) * nb * n, FL_F32, work + (size_t)(L - 2) * nb * n, FL_F32, L, 1.0f,
                          0.0f, nullptr, st);
    if (e == hipSuccess) e = fl_launch_iwt(f, work, FL_F32, g, FL_F32, 1, 1.0f, 1.0f, g, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, st, part, (int)(nbk * level), part + (size_t)nbk * level);
        e = hipGetLastError();
    }
    return e == hipSuccess ? FL_OK : fl_hip_fail(e, "framelet gradient");
}

}  // namespace

void fl_plan_release_workspace(fl_plan_s* p)
{
    if (p && p->ws) {
        PlanWs* pw = (PlanWs*)p->ws;
        free_ws(pw->w);
        delete pw;
        p->ws = nullptr;
    }
}

// =============================================================== C ABI ====

extern "C" int fl_framelet_num_bands(int dim, int kind)
{
    if ((dim != 2 && dim != 3) || (kind != FL_FRAME_HAAR && kind != FL_FRAME_LINEAR && kind != FL_FRAME_CUBIC))
        return fl_fail(FL_ERR_INVALID, "bad framelet dim/kind");
    int nf = fl_frame_nf(kind);
    return dim == 3 ? nf * nf * nf : nf * nf;
}

extern "C" int fl_framelet_forward(int dim, int kind, int level, int nx, int ny, int nz, const void* d_u,
                                   int ut, void* d_c, int ct, void* stream)
{
    if (!frame_ok(dim, kind, nx, ny, nz) || level < 1 || !d_u || !d_c)
        return fl_fail(FL_ERR_INVALID, "bad framelet arguments");
    FlFrameDesc f = {dim, kind, nx, ny, dim == 3 ? nz : 1};
    const size_t n = (size_t)nx * ny * f.nz;
    const int nb = fl_framelet_num_bands(dim, kind);
    const size_t esz = ct == FL_F32 ? 4 : 2;
    hipStream_t st = (hipStream_t)stream;
    // level L decomposes the low-pass band (band 0) of level L-1
    const void* src = d_u;
    int st_t = ut;
    for (int L = 1; L <= level; L++) {
        char* dst = (char*)d_c + (size_t)(L - 1) * nb * n * esz;
        hipError_t e = fl_launch_fwt(f, src, st_t, dst, ct, L, st);
        if (e == hipErrorInvalidValue) return fl_fail(FL_ERR_UNSUPPORTED, "dtype combination %d/%d", ut, ct);
        if (e != hipSuccess) return fl_hip_fail(e, "framelet forward");
        src = dst;
        st_t = ct;
    }
    return fl_ok();
}

extern "C" int fl_framelet_adjoint(int dim, int kind, int level, int nx, int ny, int nz, const void* d_c, int ct,
                                   void* d_u, int ut, void* d_work, void* stream)
{
    if (!frame_ok(dim, kind, nx, ny, nz) || level < 1 || !d_c || !d_u)
        return fl_fail(FL_ERR_INVALID, "bad framelet arguments");
    if (level > 1 && (!d_work || ct != FL_F32))
        return fl_fail(FL_ERR_INVALID, "multi-level synthesis needs an FP32 coefficient array and a work buffer");
    FlFrameDesc f = {dim, kind, nx, ny, dim == 3 ? nz : 1};
    const size_t n = (size_t)nx * ny * f.nz;
    const int nb = fl_framelet_num_bands(dim, kind);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    if (level == 1) {
        e = fl_launch_iwt(f, d_c, ct, d_u, ut, 1, 1.0f, 0.0f, nullptr, st);
        if (e == hipErrorInvalidValue) return fl_fail(FL_ERR_UNSUPPORTED, "dtype combination %d/%d", ct, ut);
        if (e != hipSuccess) return fl_hip_fail(e, "framelet adjoint");
        return fl_ok();
    }
    // FraRecMultiLevel2D.m:16-18 / 3D.m:14-16: C{L-1}{1} <- Rec(C{L}),
    // ..., u = Rec(C{1}).  The caller's coefficients are left untouched:
    // each level's reconstruction goes to one of two volumes of the work
    // buffer, and the synthesis of the level below reads it in place of
    // its low-pass band.
    float* work = (float*)d_work;
    const float* C = (const float*)d_c;
    const float* lp = nullptr;
    for (int L = level; L >= 2; L--) {
        float* rec = work + (size_t)((level - L) & 1) * n;
        e = fl_launch_iwt(f, C + (size_t)(L - 1) * nb * n, FL_F32, rec, FL_F32, L, 1.0f, 0.0f, nullptr, st, lp);
        if (e != hipSuccess) return fl_hip_fail(e, "framelet adjoint");
        lp = rec;
    }
    e = fl_launch_iwt(f, C, FL_F32, d_u, ut, 1, 1.0f, 0.0f, nullptr, st, lp);
    if (e != hipSuccess) return fl_hip_fail(e, "framelet adjoint");
    return fl_ok();
}

extern "C" int fl_cgls(fl_plan pl, const float* d_b, float* d_u, float lam, int iters, double tol, int flags,
                       double* gamma_log, int* iters_done, void* stream)
{
    if (!pl || !d_b || !d_u || iters < 0) return fl_fail(FL_ERR_INVALID, "bad CGLS arguments");
    if ((flags & FL_SOLVE_LOWP_BF16) && (flags & FL_SOLVE_LOWP_F16))
        return fl_fail(FL_ERR_INVALID, "choose one low-precision storage type");
    PlanWs* pw = get_ws(pl);
    std::lock_guard<std::mutex> lk(pw->mu);
    Workspace& w = pw->w;
    int rc;
    if ((rc = ensure_cg(pl, w, iters + 1, flags))) return rc;
    hipStream_t caller = (hipStream_t)stream;
    if ((rc = begin_on(w, caller))) return rc;
    if ((rc = cg_run(pl, w, d_b, d_u, lam, iters, tol, flags))) return rc;
    if (gamma_log || iters_done) {
        CgState s;
        HIPCHK(hipMemcpyAsync(&s, w.st, sizeof s, hipMemcpyDeviceToHost, w.stream));
        if (gamma_log) HIPCHK(hipMemcpyAsync(gamma_log, w.log, (size_t)(iters + 1) * sizeof(double),
                                             hipMemcpyDeviceToHost, w.stream));
        HIPCHK(hipStreamSynchronize(w.stream));
        if (iters_done) *iters_done = s.iter;
        if (gamma_log)
            for (int k = s.iter + 1; k <= iters; k++) gamma_log[k] = s.gamma;
    }
    if ((rc = end_on(w, caller))) return rc;
    return fl_ok();
}

// b <- Atf + mu * W^T d for the multi-level coefficients d (level blocks of
// nbands bands): levels above the first are synthesised in place into the
// low-pass band of the level below (FraRecMultiLevel3D.m:14-17: C{L-1}{1}
// <- Rec(C{L})); d is workspace and rebuilt by the next Bregman step.
static hipError_t synth_rhs(const FlFrameDesc& f, int level, int nbands, void* D, int ct, float mu, float* b,
                            const float* Atf, hipStream_t st)
{
    const size_t n = (size_t)f.nx * f.ny * f.nz, esz = ct == FL_F32 ? 4 : 2;
    char* base = (char*)D;
    for (int L = level; L >= 2; L--) {
        const hipError_t e = fl_launch_iwt(f, base + (size_t)(L - 1) * nbands * n * esz, ct,
                                           base + (size_t)(L - 2) * nbands * n * esz, ct, L, 1.0f, 0.0f, nullptr, st);
        if (e != hipSuccess) return e;
    }
    return fl_launch_iwt(f, D, ct, b, FL_F32, 1, mu, 1.0f, Atf, st);
}

extern "C" int fl_split_bregman(fl_plan pl, int kind, int level, const float* d_f, float* d_u, float lam, float mu,
                                int outer, int inner, double cg_tol, int flags, double* resid_log, float* d_u_hist,
                                void* stream)
{
    if (!pl || !d_f || !d_u || outer < 0 || inner < 0 || !(mu > 0) || level < 1 || level > 8)
        return fl_fail(FL_ERR_INVALID, "bad split-Bregman arguments");
    if (pl->nt != 1) return fl_fail(FL_ERR_UNSUPPORTED, "split Bregman on one frame");
    const int dim = pl->cone ? 3 : 2;
    FlFrameDesc f = {dim, kind, pl->fg.nx, pl->fg.ny, pl->cone ? pl->fg.nz : 1};
    if (!frame_ok(dim, kind, f.nx, f.ny, f.nz)) return fl_fail(FL_ERR_INVALID, "bad framelet kind");
    const int nbands = fl_framelet_num_bands(dim, kind);
    const size_t n = (size_t)pl->nvox;
    const int ct = lowp_type(flags);
    const size_t esz = ct == FL_F32 ? 4 : 2;
    const size_t ncoef = (size_t)level * nbands * n;
    PlanWs* pw = get_ws(pl);
    std::lock_guard<std::mutex> lk(pw->mu);
    Workspace& w = pw->w;
    int rc;
    if ((rc = ensure_cg(pl, w, inner + 1, flags))) return rc;
    if (!w.V || w.nb_bands != ncoef || w.coef_type != ct || (level > 1 && !w.lp)) {
        for (void* q : {w.V, w.D, (void*)w.Atf, (void*)w.b, (void*)w.lp})
            if (q) (void)hipFree(q);
        w.V = w.D = nullptr;
        w.Atf = w.b = w.lp = nullptr;
        w.nb_bands = 0;
        if ((rc = ws_alloc(&w.V, ncoef * esz, "v")) || (rc = ws_alloc(&w.D, ncoef * esz, "d")) ||
            (rc = ws_alloc((void**)&w.Atf, n * 4, "Atf")) || (rc = ws_alloc((void**)&w.b, n * 4, "b")) ||
            (level > 1 && (rc = ws_alloc((void**)&w.lp, 2 * n * 4, "low-pass bands"))))
            return rc;
        w.nb_bands = ncoef;
        w.coef_type = ct;
    }
    const size_t fb = fl_frame_blocks(f);
    if (w.fpart_n < fb * level) {
        if (w.fpart) (void)hipFree(w.fpart);
        w.fpart = nullptr;
        w.fpart_n = 0;
        if ((rc = ws_alloc((void**)&w.fpart, fb * level * sizeof(double), "framelet partials"))) return rc;
        w.fpart_n = fb * level;
    }
    if (w.res_n < (size_t)(outer > 0 ? outer : 1)) {
        if (w.res) (void)hipFree(w.res);
        w.res = nullptr;
        w.res_n = 0;
        if ((rc = ws_alloc((void**)&w.res, (size_t)(outer > 0 ? outer : 1) * sizeof(double), "residuals"))) return rc;
        w.res_n = (size_t)(outer > 0 ? outer : 1);
    }
    hipStream_t caller = (hipStream_t)stream, st = w.stream;
    if ((rc = begin_on(w, caller))) return rc;
    const int kind_adj = (flags & FL_SOLVE_MATCHED) ? FL_ADJ_MATCHED : FL_ADJ_EXACT;
    // v = alpha = W u0 (so alpha - v = 0 and b = Atf), Atf = A^T f
    for (int L = 1; L <= level; L++) {
        const void* src = L == 1 ? (const void*)d_u : (const void*)((const char*)w.V + (size_t)(L - 2) * nbands * n * esz);
        hipError_t e = fl_launch_fwt(f, src, L == 1 ? FL_F32 : ct, (char*)w.V + (size_t)(L - 1) * nbands * n * esz, ct,
                                     L, st);
        if (e != hipSuccess) return fl_hip_fail(e, "W u0");
    }
    if ((rc = fl_adjoint(pl, d_f, FL_F32, w.Atf, FL_F32, kind_adj, st))) return rc;
    HIPCHK(fl_launch_copy(w.Atf, w.b, n, st));
    for (int k = 0; k < outer; k++) {
        if ((rc = cg_run(pl, w, w.b, d_u, mu, inner, cg_tol, flags))) break;
        if (d_u_hist) {
            const hipError_t eh = fl_launch_copy(d_u, d_u_hist + (size_t)k * n, n, st);
            if (eh != hipSuccess) {
                rc = fl_hip_fail(eh, "iterate history");
                break;
            }
        }
        // per level: analysis of the previous level's low-pass band of W u
        // (u itself at level 1), shrink, Bregman update
        hipError_t e = hipSuccess;
        for (int L = 1; L <= level && e == hipSuccess; L++) {
            const float* src = L == 1 ? d_u : w.lp + (size_t)((L - 2) & 1) * n;
            float* lp_out = L < level ? w.lp + (size_t)((L - 1) & 1) * n : nullptr;
            const size_t off = (size_t)(L - 1) * nbands * n * esz;
            e = fl_launch_bregman(f, src, (char*)w.V + off, (char*)w.D + off, ct, lam / mu, L, lp_out,
                                  w.fpart + (size_t)(L - 1) * fb, st);
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, st, w.fpart, (int)(fb * level), w.res + k);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = synth_rhs(f, level, nbands, w.D, ct, mu, w.b, w.Atf, st);
        if (e != hipSuccess) {
            rc = fl_hip_fail(e, "Bregman update");
            break;
        }
    }
    if (!rc && resid_log && outer > 0) {
        hipError_t e = hipMemcpyAsync(resid_log, w.res, (size_t)outer * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fl_hip_fail(e, "residual log");
    }
    if (rc) return rc;
    if ((rc = end_on(w, caller))) return rc;
    return fl_ok();
}

extern "C" int fl_grad_data(fl_plan pl, const float* d_x, const float* d_b, float* d_g, double* loss, int flags,
                            void* stream)
{
    if (!pl || !d_x || !d_b || !d_g) return fl_fail(FL_ERR_INVALID, "bad gradient arguments");
    PlanWs* pw = get_ws(pl);
    std::lock_guard<std::mutex> lk(pw->mu);
    Workspace& w = pw->w;
    int rc;
    if ((rc = ensure_cg(pl, w, 1, 0))) return rc;
    // the residual sinogram and the partials are plan workspace, shared with
    // the solvers: the gradient runs on the workspace stream, ordered after
    // the caller's stream and before the caller's later work
    hipStream_t caller = (hipStream_t)stream, st = w.stream;
    if ((rc = begin_on(w, caller))) return rc;
    const size_t ns = (size_t)pl->ndet * pl->nv;
    if ((rc = fl_forward(pl, d_x, FL_F32, w.sino, FL_F32, st))) return rc;
    const unsigned g = grid_for(ns);
    hipLaunchKernelGGL(k_resid, dim3(g), dim3(256), 0, st, w.sino, d_b, ns, w.part);
    double* out = w.part + NPART;
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, st, w.part, (int)g, out);
    HIPCHK(hipGetLastError());
    const int kind = (flags & FL_GRAD_EXACT_ADJOINT) ? FL_ADJ_EXACT : FL_ADJ_MATCHED;
    if ((rc = fl_adjoint(pl, w.sino, FL_F32, d_g, FL_F32, kind, st))) return rc;
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, out, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    if ((rc = end_on(w, caller))) return rc;
    return fl_ok();
}

extern "C" int fl_grad_framelet(int dim, int kind, int level, int nx, int ny, int nz, const float* d_x, float lam,
                                float eps, int with_lowpass, float* d_g, float* d_work, double* loss, void* stream)
{
    if (!frame_ok(dim, kind, nx, ny, nz) || !d_x || !d_g || !d_work || !(eps > 0) || level < 1 || level > 8)
        return fl_fail(FL_ERR_INVALID, "bad framelet gradient arguments");
    FlFrameDesc f = {dim, kind, nx, ny, dim == 3 ? nz : 1};
    const unsigned nbk = fl_frame_blocks(f);
    double* part = device_scratch_doubles((size_t)nbk * level + 1);
    if (!part) return fl_fail(FL_ERR_ALLOC, "framelet partials");
    const int rc = grad_framelet_run(f, level, d_x, lam, eps, with_lowpass, d_g, d_work, part, nbk, (hipStream_t)stream);
    if (rc) return rc;
    if (loss) {
        HIPCHK(hipMemcpyAsync(loss, part + (size_t)nbk * level, sizeof(double), hipMemcpyDeviceToHost,
                              (hipStream_t)stream));
        HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    }
    return fl_ok();
}

// ------------------------------------- CGLS over caller-supplied operators --

namespace {
// FP64 partials of the generic vector steps: summed in a fixed order on the
// device, the sum copied to the host (the step synchronises the stream)
int sum_to_host(const double* part, unsigned np, double* out, hipStream_t st)
{
    double* d_out = const_cast<double*>(part) + NPART;
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, st, part, (int)np, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return FL_OK;
}
}  // namespace

extern "C" int fl_vec_cgls_init(long long n, const float* d_q, const float* d_u, const float* d_b, float lam,
                                float* d_r, float* d_p, double* gamma, void* stream)
{
    if (n <= 0 || !d_q || !d_u || !d_b || !d_r || !d_p || !gamma)
        return fl_fail(FL_ERR_INVALID, "bad CGLS vector arguments");
    double* part = device_scratch_doubles(NPART + 1);
    if (!part) return fl_fail(FL_ERR_ALLOC, "CGLS partials");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)n);
    hipLaunchKernelGGL((k_cg_init<float>), dim3(g), dim3(256), 0, st, d_q, d_u, d_b, lam, d_r, d_p, (float*)nullptr,
                       (size_t)n, part);
    const int rc = sum_to_host(part, g, gamma, st);
    return rc ? rc : fl_ok();
}

extern "C" int fl_vec_cgls_pq(long long n, float* d_q, const float* d_p, float lam, double* pq, void* stream)
{
    if (n <= 0 || !d_q || !d_p || !pq) return fl_fail(FL_ERR_INVALID, "bad CGLS vector arguments");
    double* part = device_scratch_doubles(NPART + 1);
    if (!part) return fl_fail(FL_ERR_ALLOC, "CGLS partials");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)n);
    hipLaunchKernelGGL(k_gv_pq, dim3(g), dim3(256), 0, st, d_q, d_p, lam, (size_t)n, part);
    const int rc = sum_to_host(part, g, pq, st);
    return rc ? rc : fl_ok();
}

extern "C" int fl_vec_cgls_update(long long n, float* d_u, float* d_r, const float* d_p, const float* d_q,
                                  double alpha, double* gamma, void* stream)
{
    if (n <= 0 || !d_u || !d_r || !d_p || !d_q || !gamma) return fl_fail(FL_ERR_INVALID, "bad CGLS vector arguments");
    double* part = device_scratch_doubles(NPART + 1);
    if (!part) return fl_fail(FL_ERR_ALLOC, "CGLS partials");
    hipStream_t st = (hipStream_t)stream;
    const unsigned g = grid_for((size_t)n);
    hipLaunchKernelGGL(k_gv_update, dim3(g), dim3(256), 0, st, d_u, d_r, d_p, d_q, (float)alpha, (size_t)n, part);
    const int rc = sum_to_host(part, g, gamma, st);
    return rc ? rc : fl_ok();
}

extern "C" int fl_vec_cgls_dir(long long n, float* d_p, const float* d_r, double beta, void* stream)
{
    if (n <= 0 || !d_p || !d_r) return fl_fail(FL_ERR_INVALID, "bad CGLS vector arguments");
    hipLaunchKernelGGL(k_gv_dir, dim3(grid_for((size_t)n)), dim3(256), 0, (hipStream_t)stream, d_p, d_r, (float)beta,
                       (size_t)n);
    HIPCHK(hipGetLastError());
    return fl_ok();
}

extern "C" int fl_framelet_shrink(int nbands, int level, long long band_elems, void* d_c, int c_dtype, float tau,
                                  void* stream)
{
    if (nbands <= 0 || level < 1 || band_elems <= 0 || !d_c || !(tau >= 0))
        return fl_fail(FL_ERR_INVALID, "bad soft-threshold arguments");
    const size_t band = (size_t)band_elems, total = band * nbands * level;
    const unsigned g = grid_for(total);
    hipStream_t st = (hipStream_t)stream;
    if (c_dtype == FL_F32)
        hipLaunchKernelGGL(k_shrink_bands<float>, dim3(g), dim3(256), 0, st, (float*)d_c, band, nbands, total, tau);
    else if (c_dtype == FL_BF16)
        hipLaunchKernelGGL(k_shrink_bands<bf16_t>, dim3(g), dim3(256), 0, st, (bf16_t*)d_c, band, nbands, total, tau);
    else if (c_dtype == FL_F16)
        hipLaunchKernelGGL(k_shrink_bands<f16_t>, dim3(g), dim3(256), 0, st, (f16_t*)d_c, band, nbands, total, tau);
    else
        return fl_fail(FL_ERR_INVALID, "bad coefficient dtype");
    HIPCHK(hipGetLastError());
    return fl_ok();
}

// ------------------------------------- framelet gradient, no allocation --
//
// fl_grad_framelet_noalloc: the FP64 loss partials live in `scratch`, a
// file-scope device array of FL_NOALLOC_PARTIALS doubles (16 MiB) that each
// GPU holds until the process ends, instead of the grow-only buffer
// fl_grad_framelet allocates; the loss goes to a device pointer, so the call
// neither allocates nor synchronises.  Every partial is written by its
// workgroup before the fixed-order sum reads it, so nothing needs clearing.
// Calls that overlap on one GPU share the array and race.

static __device__ __align__(256) double scratch[FL_NOALLOC_PARTIALS];

namespace {
std::mutex g_noalloc_mu;
std::vector<double*> g_noalloc_ptr;   // per device: the device address of `scratch`

__global__ void k_copy_double(const double* __restrict__ a, double* __restrict__ b) { *b = *a; }

// device address of `scratch` on the current device (resolved once per device)
int noalloc_scratch(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fl_hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc_ptr.size() <= dev) g_noalloc_ptr.resize(dev + 1, nullptr);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return fl_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[dev] = (double*)ptr;
    }
    *out = g_noalloc_ptr[dev];
    return FL_OK;
}
}  // namespace

extern "C" int fl_grad_framelet_noalloc(int dim, int kind, int level, int nx, int ny, int nz, const float* d_x,
                                        float lam, float eps, int with_lowpass, float* d_g, float* d_work,
                                        double* d_loss, void* stream)
{
    if (!frame_ok(dim, kind, nx, ny, nz) || !d_x || !d_g || !d_work || !(eps > 0) || level < 1 || level > 8)
        return fl_fail(FL_ERR_INVALID, "bad framelet gradient arguments");
    FlFrameDesc f = {dim, kind, nx, ny, dim == 3 ? nz : 1};
    const unsigned nbk = fl_frame_blocks(f);
    const size_t need = (size_t)nbk * level + 1;
    if (need > (size_t)FL_NOALLOC_PARTIALS)
        return fl_fail(FL_ERR_INVALID, "volume too large for the no-allocation gradient: %zu partials of the %d in its "
                                       "static scratch (about %d voxels x levels)", need, FL_NOALLOC_PARTIALS,
                       (FL_NOALLOC_PARTIALS - 1) * 256);
    double* part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = grad_framelet_run(f, level, d_x, lam, eps, with_lowpass, d_g, d_work, part, nbk, st))) return rc;
    if (d_loss) {
        hipLaunchKernelGGL(k_copy_double, dim3(1), dim3(1), 0, st, part + (size_t)nbk * level, d_loss);
        HIPCHK(hipGetLastError());
    }
    return fl_ok();
}
