// diag.hip -- measurement and self-check kernels (libfl_diag.so; test and
// profiling infrastructure, not part of libframelab_hip.so): VALU issue-peak
// calibration for counter-based rooflines, and an exhaustive check of the
// exact adjoint's square root against sqrtf.
//
// The projectors are bound by vector-ALU issue, not HBM.  To state what
// fraction of the VALU peak they reach, the profiler's per-dispatch VALU
// instruction count is compared with these kernels' under the same
// counters: eight independent chains per thread of FP32 FMAs or of 32-bit
// integer shift/add/xor, no memory traffic but one store per thread, so
// each issues VALU instructions of its kind at the rate the hardware allows.
#include <hip/hip_runtime.h>

#include "../../framelab_amd/csrc/fl_chord.h"

namespace {

// KIND 0: FP32 FMA chains; KIND 1: 32-bit integer shift/add/xor chains.
template <int KIND>
__global__ __launch_bounds__(256) void k_valu_peak(float* __restrict__ out, int iters, float b, float c)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    float a[8];
    unsigned u[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        a[j] = (float)threadIdx.x + (float)j;
        u[j] = threadIdx.x + (unsigned)j;
    }
    for (int k = 0; k < iters; k++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (KIND == 0) a[j] = __builtin_fmaf(a[j], b, c);
            else u[j] = (u[j] >> 1) ^ (u[j] + (unsigned)iters);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++) s += KIND == 0 ? a[j] : (float)u[j];
    out[i] = s;
}

// Every non-negative finite float (bit patterns 0 .. 0x7f7fffff, grid-
// stride): does fl_sqrt_nonneg return the bits of the correctly rounded
// sqrtf?  Mismatches are counted per thread and added to *bad.
constexpr uint32_t kLastFinite = 0x7f7fffffu;

__global__ __launch_bounds__(256) void k_check_sqrt(unsigned long long* __restrict__ bad)
{
    const uint32_t stride = gridDim.x * 256u;   // 2^24: p + stride stays below 2^32
    unsigned long long n = 0;
    for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p <= kLastFinite; p += stride) {
        const float s = __uint_as_float(p);
        n += __float_as_uint(fl_sqrt_nonneg(s)) != __float_as_uint(sqrtf(s));
    }
    if (n) atomicAdd(bad, n);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ v) { *v = 0; }

}  // namespace

// Status: 0 ok, -1 invalid argument, -2 HIP launch error.
extern "C" int fl_check_chord_sqrt(unsigned long long* d_bad, void* stream)
{
    if (!d_bad) return -1;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, (hipStream_t)stream, d_bad);
    hipLaunchKernelGGL(k_check_sqrt, dim3(65536), dim3(256), 0, (hipStream_t)stream, d_bad);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

extern "C" int fl_calibrate_valu(float* d_out, int n_threads, int iters, int kind, void* stream)
{
    if (!d_out || n_threads <= 0 || (n_threads % 256) != 0 || iters <= 0 || (kind != 0 && kind != 1))
        return -1;   // n_threads: a positive multiple of 256; iters positive; kind 0 or 1
    const dim3 grid(n_threads / 256);
    if (kind == 0)
        hipLaunchKernelGGL(k_valu_peak<0>, grid, dim3(256), 0, (hipStream_t)stream, d_out, iters, 0.999f, 1e-3f);
    else
        hipLaunchKernelGGL(k_valu_peak<1>, grid, dim3(256), 0, (hipStream_t)stream, d_out, iters, 0.999f, 1e-3f);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
