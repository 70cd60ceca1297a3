/*
 * framelab_hip.h -- C ABI of the MI355X-native FrameLab CT-operator library
 * (libframelab_hip.so).  Plain C: pointers, sizes and status codes only.
 *
 * Two layers:
 *   1. Drop-in kernel entry points with the reference's names and argument
 *      lists (host pointers, synchronous, outputs pre-allocated by the
 *      caller), so the reference's MEX gateways relink unchanged.
 *   2. A plan API on device pointers and a caller-supplied HIP stream
 *      (passed as void*), returning int status codes, with storage-type
 *      variants (FP32 / BF16 / FP16), adjoint variants, framelet transforms,
 *      device-resident solvers and the gradient of the reconstruction
 *      objective.
 *
 * Reference citations are relative to the reference repository root.
 */
#ifndef FRAMELAB_HIP_H
#define FRAMELAB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define FL_OK 0
#define FL_ERR_INVALID (-1)     /* invalid argument (sizes, pointers, enums) */
#define FL_ERR_HIP (-2)         /* a HIP runtime call failed */
#define FL_ERR_ALLOC (-3)       /* device allocation failed */
#define FL_ERR_UNSUPPORTED (-4) /* valid but not implemented combination */

/* Status of the last call made on this thread (drop-in entry points return
 * void, like the reference's, and report through this). */
int fl_get_last_error(void);
const char *fl_last_error_message(void);
const char *fl_error_string(int code);
int fl_version(void);

/* storage element types */
#define FL_F32 0
#define FL_BF16 1
#define FL_F16 2

/* ------------------------------------------------ 1. drop-in entry points */

/* 2D fan Siddon forward.  Replaces Ax_fan_mf_gpu_siddon
 * (matlab/mex/src/Ax_fan_mf.h:3; gateway call matlab/mex/src/Ax_fan_mf.cpp:49).
 * Output y[iv*nd+id]; image frame of view iv is id_X[iv]. */
void Ax_fan_mf_gpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                          int nv, float *sd_phi, int nd, float *y_det, int *id_X, int nt);

/* 2D fan forward, class default version 1 (Gao column walk).  Replaces
 * Ax_fan_mf_gpu_new (matlab/mex/src/Ax_fan_mf.h:6; Ax_fan_mf.cpp:55). */
void Ax_fan_mf_gpu_new(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                       int nv, float *sd_phi, int nd, float *y_det, int *id_X, int nt);

/* 2D pixel-driven exact adjoint.  Replaces Atx_fan_mf_gpu_new
 * (matlab/mex/src/Atx_fan_mf.h:3; gateway matlab/mex/src/Atx_fan_mf.cpp:60). */
void Atx_fan_mf_gpu_new(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                        float *sd_phi, int nd, float *y_det, float dy_det, float y_os, int nt,
                        int *id_Y, int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                        float *cos_det, float *sin_det);

/* 3D cone Siddon forward.  Replaces Ax_cone_mf_gpu_siddon
 * (matlab/mex/src/Ax_cone_mf.h:8; gateway matlab/mex/src/Ax_cone_mf.cpp:65).
 * Output slots follow the frame -> view lists id_Y/Nv, as the reference's
 * GPU wrapper orders them (matlab/mex/src/Ax_cone_mf_gpu_siddon.cu:167-186). */
void Ax_cone_mf_gpu_siddon(float *X, float *y, float *sd_phi, float *sd_z, float *y_det,
                           float *z_det, int *id_Y, int *Nv, float SO, float OD, float scale,
                           float dz, int nx, int ny, int nz, int nt, int na, int nb, int nv,
                           int tmp_size, int nv_block);

/* 3D cone forward, class default version 1, with the normalised second
 * output yNorm = y / (|source - detector cell| * scale).  Replaces
 * Ax_cone_mf_gpu_new (matlab/mex/src/Ax_cone_mf.h:5; Ax_cone_mf.cpp:71).
 * y is the exact Siddon ray sum (the reference's Gao cone column walk
 * misplaces lengths when the y and z slopes have opposite signs). */
void Ax_cone_mf_gpu_new(float *X, float *y, float *yNorm, float *sd_phi, float *sd_z,
                        float *y_det, float *z_det, int *id_Y, int *Nv, float SO, float OD,
                        float scale, float dz, int nx, int ny, int nz, int nt, int na, int nb,
                        int nv, int tmp_size, int nv_block);

/* 3D voxel-driven exact adjoint.  Replaces Atx_cone_mf_gpu_new
 * (matlab/mex/src/Atx_cone_mf.h:3; gateway matlab/mex/src/Atx_cone_mf.cpp:64). */
void Atx_cone_mf_gpu_new(float *X, float *y, float *cos_phi, float *sin_phi, float *sd_z,
                         float *y_det, float *z_det, int *id_Y, int *Nv, float SO, float OD,
                         float scale, float dy_det, float y_os, float dz_det, float dz, int nx,
                         int ny, int nz, int nt, int na, int nb, int nv, int tmp_size,
                         int nv_block);

/* ------------------------------------------------------------ 2. plan API */

/* Geometry: the fields of the reference `para` struct
 * (matlab/classes/+Operators/@ConeBeamScanner/setPara.m:1-83,
 *  @FanBeamScanner/setPara.m:1-56).  Fan: nz = 0, nb = 1, na = nd.
 * Host pointers; the plan copies what it needs.  slot_view may be NULL
 * (output slot k holds view k); otherwise forward output slot k holds view
 * slot_view[k].  cos_det/sin_det are needed by the fan adjoint only. */
typedef struct fl_geometry {
    int nx, ny, nz, nt;
    int nv, na, nb;
    float SO, OD, scale, dz, dy_det, dz_det, y_os;
    const float *sd_phi, *sd_z, *y_det, *z_det;
    const float *cos_phi, *sin_phi, *cos_det, *sin_det;
    const int *id_X, *id_Y, *Nv;
    int tmp_size;
    const int *slot_view;
} fl_geometry;

typedef struct fl_plan_s *fl_plan;

int fl_plan_create(fl_plan *plan, const fl_geometry *g);
int fl_plan_destroy(fl_plan plan);
/* sizes: voxels per frame, frames, detector cells per view, views/slots */
int fl_plan_sizes(fl_plan plan, long long *nvox, int *nt, long long *ndet, int *nv);

/* adjoint kinds */
#define FL_ADJ_EXACT 0   /* reference exact-chord voxel-driven adjoint (find_l / find_l_3d) */
#define FL_ADJ_MATCHED 1 /* voxel-driven transpose of the Siddon forward (same lengths) */

/* y = A x.  d_x: nt frames of nvox voxels, d_y: nv slots of ndet cells. */
int fl_forward(fl_plan plan, const void *d_x, int x_dtype, void *d_y, int y_dtype, void *stream);
/* Fan forward with the Gao column walk (class default version 1,
 * matlab/mex/src/Ax_fan_mf_cpu_new.cpp:18-122); fan plans only. */
int fl_forward_gao_fan(fl_plan plan, const void *d_x, int x_dtype, void *d_y, int y_dtype,
                       void *stream);
/* x = A^T y. */
int fl_adjoint(fl_plan plan, const void *d_y, int y_dtype, void *d_x, int x_dtype, int kind,
               void *stream);

/* Traversal log of selected rays (testing): rays = (view, ia, ib) triples
 * (host), cap entries per ray; J/S/cnt host outputs of nrays*cap / nrays. */
int fl_siddon_trace(fl_plan plan, int nrays, const int *rays, int cap, int *J, float *S, int *cnt);

#ifdef __cplusplus
}
#endif

#endif /* FRAMELAB_HIP_H */
