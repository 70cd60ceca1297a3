// fl_capi_internal.h -- plan object and launcher declarations shared by the
// host-side translation units of libframelab_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "fl_geom.h"

struct fl_plan_s {
    int device = 0;
    bool cone = false;
    long long nvox = 0, ndet = 0;
    int nt = 1, nv = 0;
    FwdGeom fg{};
    AdjGeom ag{};
    std::vector<float> h_sd_phi;
    float *d_cs = nullptr, *d_sn = nullptr, *d_sdz = nullptr, *d_ydet = nullptr, *d_zdet = nullptr;
    float *d_cphi = nullptr, *d_sphi = nullptr, *d_cdet = nullptr, *d_sdet = nullptr;
    int *d_slot_view = nullptr, *d_view_frame = nullptr, *d_idY = nullptr, *d_Nv = nullptr;
    int* d_idS = nullptr;         // adjoint: sinogram slot per frame-list entry (AdjGeom::id_S)
    bool slots_ok = true;         // every frame-list view has a forward output slot to read
    void* d_bricks = nullptr;   // cone: bricked copy of the forward input (nt frames, bricks of the launch shape, 4-byte)
    // forwards share d_bricks: each one waits for the previous one (on
    // whatever stream it ran) through this event, recorded after it
    hipEvent_t fwd_done = nullptr;
    bool fwd_pending = false;
    std::mutex fwd_mu;
    // fan, finite-width beam: strip lines per (view, cell, edge) and cell
    // edge tables (built on first use / uploaded when given)
    float *d_fb_px = nullptr, *d_fb_py = nullptr, *d_fb_pd = nullptr;
    float *d_cdet2 = nullptr, *d_sdet2 = nullptr;
    float2* d_line_q = nullptr;   // fan: per (view, cell) line-chord quotients
    float4* d_ray_q = nullptr;    // fan: per (view, cell) ray reciprocals and length
    std::vector<float> h_ydet, h_zdet, h_sdz;
    // largest distance, in cells, of a detector coordinate from the uniform
    // grid through the first and last cell: the adjoints' window maps assume
    // a uniform pitch
    double det_dev = 0.0;
    // FP64 projector pair: geometry in double (cos/sin of the double angle,
    // source heights, detector coordinates), built on first use
    double *d_c64 = nullptr, *d_s64 = nullptr, *d_sdz64 = nullptr, *d_ydet64 = nullptr, *d_zdet64 = nullptr;
    std::mutex f64_mu;
    // lazily allocated workspace (solvers, gradients)
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

constexpr int FL_FWD_NTILES = 8;   // cone forward ray-tile shapes (siddon.hip)
// elements of the largest bricked frame of any of them (siddon.hip)
size_t fl_bricked_frame_elems(int nx, int ny, int nz);
// detector coordinates may stray from a uniform grid by at most this many
// cells for the voxel-driven adjoints (their windows widen by 0.05 cells)
constexpr double FL_DET_DEV_MAX = 0.01;

int fl_fail(int code, const char* fmt, ...);
int fl_ok();
int fl_hip_fail(hipError_t e, const char* what);
void fl_plan_release_workspace(fl_plan_s* p);

hipError_t fl_launch_fwd_cone(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st);
hipError_t fl_launch_fwd_fan(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st);
hipError_t fl_launch_fwd_fan_gao(const void* X, int xt, void* Y, int yt, const FwdGeom& g, hipStream_t st);
hipError_t fl_launch_trace(const FwdGeom& g, int cone, int nrays, const int* rays, const int* off,
                           const int* cap, int* J, float* S, int* cnt, hipStream_t st);
hipError_t fl_launch_ynorm(const float* y, float* yn, const FwdGeom& g, hipStream_t st);
hipError_t fl_launch_adj_fan(const void* Y, int yt, void* X, int xt, const AdjGeom& g, hipStream_t st);
hipError_t fl_launch_fan_line_q(const AdjGeom& g, int nv, float2* out, hipStream_t st);
hipError_t fl_launch_fan_ray_q(const FwdGeom& f, int nv, float4* out, hipStream_t st);
hipError_t fl_launch_adj_cone(const void* Y, int yt, void* X, int xt, const AdjGeom& g, hipStream_t st);
hipError_t fl_launch_adj_matched(const void* Y, int yt, void* X, int xt, const fl_plan_s* p, hipStream_t st);
hipError_t fl_launch_fwd_fan_fb(const float* X, float* Y, const fl_plan_s* p, hipStream_t st);
hipError_t fl_launch_adj_fan_fb(const float* Y, float* X, const fl_plan_s* p, hipStream_t st);
void fl_plan_release_f64(fl_plan_s* p);
