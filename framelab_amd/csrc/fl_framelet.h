// fl_framelet.h -- host-side declarations of the framelet launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>

#define FL_FRAME_HAAR 0
#define FL_FRAME_LINEAR 1
#define FL_FRAME_CUBIC 3

struct FlFrameDesc {
    int dim;    // 2 or 3
    int kind;   // FL_FRAME_*
    int nx, ny, nz;
};

int fl_frame_nf(int kind);
// the filter taps in double (decomposition D, reconstruction R; NT taps of NF filters)
void fl_frame_taps_f64(int kind, double (&D)[5][5], double (&R)[5][5]);
unsigned fl_frame_blocks(const FlFrameDesc& f);
hipError_t fl_launch_fwt(const FlFrameDesc& f, const void* u, int ut, void* C, int ct, int level, hipStream_t st);
hipError_t fl_launch_iwt(const FlFrameDesc& f, const void* C, int ct, void* u, int ut, int level, float alpha,
                         float beta, const float* base, hipStream_t st, const void* lp = nullptr);
// one level (dilation 2^(level-1)); lp (may be null): that level's low-pass band of W u
hipError_t fl_launch_bregman(const FlFrameDesc& f, const float* u, void* V, void* D, int ct, float tau, int level,
                             float* lp, double* part, hipStream_t st);
// one level; lp_pass: band 0 receives the low-pass band of W u (no penalty)
hipError_t fl_launch_l1grad(const FlFrameDesc& f, const float* u, float* G, float lam, float eps, int with_lp,
                            int level, int lp_pass, double* part, hipStream_t st);
hipError_t fl_launch_copy(const float* a, float* b, size_t n, hipStream_t st);
