// fl_common.h -- shared device helpers for the FrameLab MI355X kernels.
//
// Everything here is compiled with -ffp-contract=off: the projector
// arithmetic follows the reference's float operation order exactly (each
// multiply and add rounds), which is what makes voxel indices, lengths and
// ray sums bit-identical to the reference CPU mex path.  Places that want a
// fused multiply-add call __builtin_fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FL_INF __builtin_huge_valf()

// Storage element types of volumes and sinograms.  Arithmetic is always
// FP32; BF16/FP16 are storage formats only (round-to-nearest-even stores).
#ifndef FL_F32
#define FL_F32 0
#define FL_BF16 1
#define FL_F16 2
#endif

struct bf16_t { uint16_t v; };
struct f16_t { _Float16 v; };

// num / d given r = 1.0f / d (correctly rounded), bit-identical to the IEEE
// quotient: q = num*r lies within an ulp of it, the fused residual
// num - q*d is exact, and q + residual*r rounds to the quotient
// (Markstein's theorem, for operands whose quotient and residual neither
// overflow nor underflow, as geometry differences do not).  d = +-0 makes r
// infinite; the product num*r is then IEEE's signed infinity (or NaN for
// 0/0), which the fused step would turn into NaN, so it is kept.  Lets one
// division per denominator serve a whole table of quotients.
__device__ __forceinline__ float fl_div_r(float num, float d, float r)
{
    const float q = num * r;
    const float e = __builtin_fmaf(-q, d, num);
    const float v = __builtin_fmaf(e, r, q);
    return __builtin_isinf(r) ? q : v;
}

// k / n for 0 <= k, 1 <= n with k * n < 2^18, by one multiply and shift:
// with m = ceil(2^18 / n) = (2^18 + e) / n, e < n, k*m / 2^18 exceeds k / n
// by k*e / (n 2^18) < 1/n, which never reaches the next integer.
__device__ __forceinline__ uint32_t fl_small_div_m(uint32_t n) { return ((1u << 18) + n - 1) / n; }
__device__ __forceinline__ int fl_small_div(int k, uint32_t m) { return (int)(((uint32_t)k * m) >> 18); }

__device__ __forceinline__ float fl_load(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float fl_load(const bf16_t* p, size_t i)
{
    return __uint_as_float(((uint32_t)p[i].v) << 16);
}
__device__ __forceinline__ float fl_load(const f16_t* p, size_t i) { return (float)p[i].v; }

__device__ __forceinline__ uint16_t fl_f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void fl_store(float* p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ void fl_store(bf16_t* p, size_t i, float v) { p[i].v = fl_f32_to_bf16(v); }
__device__ __forceinline__ void fl_store(f16_t* p, size_t i, float v) { p[i].v = (_Float16)v; }
// Streaming (nontemporal) stores for outputs written once and not re-read
// by the same kernel: 27 band streams per wave write-combine better when
// they do not displace L2 lines.
__device__ __forceinline__ void fl_store_nt(float* p, size_t i, float v) { __builtin_nontemporal_store(v, p + i); }
__device__ __forceinline__ void fl_store_nt(bf16_t* p, size_t i, float v)
{
    __builtin_nontemporal_store(fl_f32_to_bf16(v), &p[i].v);
}
__device__ __forceinline__ void fl_store_nt(f16_t* p, size_t i, float v)
{
    __builtin_nontemporal_store((_Float16)v, &p[i].v);
}

// Map a linear workgroup id so that consecutive logical tiles land on the
// same XCD (the dispatcher deals workgroups round-robin over 8 XCDs, each
// with its own L2).  nblk must be the grid size along the remapped axis.
__device__ __forceinline__ uint32_t fl_xcd_remap(uint32_t b, uint32_t nblk)
{
    const uint32_t nx = 8;
    uint32_t per = nblk / nx, rem = nblk % nx;
    uint32_t xcd = b % nx, k = b / nx;
    // XCDs [0, rem) hold per+1 tiles, the rest hold per tiles.
    uint32_t base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + k;
}
