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
#define FL_F64 3 /* John's-equation input only */

/* ------------------------------------------------ 1. drop-in entry points */

/* 2D fan Siddon forward.  Replaces Ax_fan_mf_gpu_siddon
 * (matlab/mex/src/Ax_fan_mf.h:3; gateway call matlab/mex/src/Ax_fan_mf.cpp:51).
 * Output y[iv*nd+id]; image frame of view iv is id_X[iv]. */
void Ax_fan_mf_gpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                          int nv, float *sd_phi, int nd, float *y_det, int *id_X, int nt);

/* 2D fan forward, class default version 1 (Gao column walk).  Replaces
 * Ax_fan_mf_gpu_new (matlab/mex/src/Ax_fan_mf.h:7; Ax_fan_mf.cpp:57). */
void Ax_fan_mf_gpu_new(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                       int nv, float *sd_phi, int nd, float *y_det, int *id_X, int nt);

/* 2D pixel-driven exact adjoint.  Replaces Atx_fan_mf_gpu_new
 * (matlab/mex/src/Atx_fan_mf.h:4; gateway matlab/mex/src/Atx_fan_mf.cpp:59). */
void Atx_fan_mf_gpu_new(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                        float *sd_phi, int nd, float *y_det, float dy_det, float y_os, int nt,
                        int *id_Y, int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                        float *cos_det, float *sin_det);

/* 2D fan forward, finite-width beam (version 2).  Replaces
 * Ax_fan_mf_gpu_new_fb (matlab/mex/src/Ax_fan_mf.h:13; Ax_fan_mf.cpp:63). */
void Ax_fan_mf_gpu_new_fb(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                          int nv, float *sd_phi, int nd, float *y_det, int *id_X, int nt,
                          float dy_det);

/* 2D adjoint, finite-width beam (version 2).  Replaces Atx_fan_mf_gpu_new_fb
 * (matlab/mex/src/Atx_fan_mf.h:13; Atx_fan_mf.cpp:67): cos_det / sin_det are
 * the nd+1 edge tables the gateway reads from para.cos_det2 / sin_det2. */
void Atx_fan_mf_gpu_new_fb(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                           float *sd_phi, int nd, float *y_det, float dy_det, int nt, int *id_Y,
                           int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                           float *cos_det, float *sin_det);

/* 3D cone Siddon forward.  Replaces Ax_cone_mf_gpu_siddon
 * (matlab/mex/src/Ax_cone_mf.h:10; gateway matlab/mex/src/Ax_cone_mf.cpp:65).
 * Output slots follow the frame -> view lists id_Y/Nv, as the reference's
 * GPU wrapper orders them (matlab/mex/src/Ax_cone_mf_gpu_siddon.cu:167-186). */
void Ax_cone_mf_gpu_siddon(float *X, float *y, float *sd_phi, float *sd_z, float *y_det,
                           float *z_det, int *id_Y, int *Nv, float SO, float OD, float scale,
                           float dz, int nx, int ny, int nz, int nt, int na, int nb, int nv,
                           int tmp_size, int nv_block);

/* 3D cone forward, class default version 1, with the normalised second
 * output yNorm = y / (|source - detector cell| * scale).  Replaces
 * Ax_cone_mf_gpu_new (matlab/mex/src/Ax_cone_mf.h:3; Ax_cone_mf.cpp:71).
 * y is the exact Siddon ray sum (the reference's Gao cone column walk
 * misplaces lengths when the y and z slopes have opposite signs). */
void Ax_cone_mf_gpu_new(float *X, float *y, float *yNorm, float *sd_phi, float *sd_z,
                        float *y_det, float *z_det, int *id_Y, int *Nv, float SO, float OD,
                        float scale, float dz, int nx, int ny, int nz, int nt, int na, int nb,
                        int nv, int tmp_size, int nv_block);

/* 3D voxel-driven exact adjoint.  Replaces Atx_cone_mf_gpu_new
 * (matlab/mex/src/Atx_cone_mf.h:4; gateway matlab/mex/src/Atx_cone_mf.cpp:65). */
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
    /* fan, finite-width beam (version 2) adjoint: cos/sin of the nd+1
     * detector-cell edge angles (para.cos_det2 / sin_det2); optional */
    const float *cos_det2, *sin_det2;
} fl_geometry;

typedef struct fl_plan_s *fl_plan;

int fl_plan_create(fl_plan *plan, const fl_geometry *g);
int fl_plan_destroy(fl_plan plan);
/* sizes: voxels per frame, frames, detector cells per view, views/slots */
int fl_plan_sizes(fl_plan plan, long long *nvox, int *nt, long long *ndet, int *nv);

/* Times the cone-forward ray-tile shapes on this plan's geometry and keeps
 * the fastest (outputs optional; fan plans: no-op).  fl_plan_set_tile picks
 * a shape by index 0..7; a cone plan starts at shape 6 (16 x 16-ray
 * workgroups gathering from 4 x 4 x 4 voxel bricks).  No reference
 * counterpart (launch tuning). */
int fl_plan_autotune(fl_plan plan, int *fwd_tile, float *fwd_ms);
int fl_plan_set_tile(fl_plan plan, int tile);

/* adjoint kinds */
#define FL_ADJ_EXACT 0   /* reference exact-chord voxel-driven adjoint (find_l / find_l_3d) */
#define FL_ADJ_MATCHED 1 /* voxel-driven transpose of the Siddon forward (same lengths) */
#define FL_ADJ_FAN_FB 2  /* fan, finite-width beam strip areas (Atx_fan_mf_cpu_new_fb.cpp:21-66);
                            needs cos_det2/sin_det2; FP32 only */

/* y = A x.  d_x: nt frames of nvox voxels, d_y: nv slots of ndet cells. */
int fl_forward(fl_plan plan, const void *d_x, int x_dtype, void *d_y, int y_dtype, void *stream);
/* Fan forward, finite-width beam: strip-area weights per detector cell
 * (version 2, matlab/mex/src/Ax_fan_mf_cpu_new_fb.cpp:19-123); FP32. */
int fl_forward_fan_fb(fl_plan plan, const float *d_x, float *d_y, void *stream);
/* Fan forward with the Gao column walk (class default version 1,
 * matlab/mex/src/Ax_fan_mf_cpu_new.cpp:18-122); fan plans only. */
int fl_forward_gao_fan(fl_plan plan, const void *d_x, int x_dtype, void *d_y, int y_dtype,
                       void *stream);
/* x = A^T y. */
int fl_adjoint(fl_plan plan, const void *d_y, int y_dtype, void *d_x, int x_dtype, int kind,
               void *stream);

/* ------------------------------------------------ framelet transforms */

/* Frame codes of GenerateFrameletFilter
 * (matlab/classes/+Transforms/@FrameletSystem/GenerateFrameletFilter.m:1-30). */
#define FL_FRAME_HAAR 0
#define FL_FRAME_LINEAR 1
#define FL_FRAME_CUBIC 3

/* Bands per level: NF^dim (NF = 2, 3, 5 filters), or a negative status. */
int fl_framelet_num_bands(int dim, int kind);

/* Analysis W (FrameletSystem.forwardTransform, FrameletSystem.m:31-49 ->
 * FraDecMultiLevel{2D,3D}.m).  u: nz*ny*nx (nz = 1 for dim 2); c: `level`
 * blocks of NB band volumes, band b = i + NF*j (+ NF^2*k), level L using
 * dilation 2^(L-1) on the low-pass band of level L-1. */
int fl_framelet_forward(int dim, int kind, int level, int nx, int ny, int nz, const void *d_u,
                        int u_dtype, void *d_c, int c_dtype, void *stream);

/* Synthesis W^T (FrameletSystem.adjointTransform, FrameletSystem.m:51-64 ->
 * FraRecMultiLevel{2D,3D}.m).  For level > 1 the coefficients must be FP32
 * and d_work must hold 2*nz*ny*nx floats (two volumes); d_c is not modified. */
int fl_framelet_adjoint(int dim, int kind, int level, int nx, int ny, int nz, const void *d_c,
                        int c_dtype, void *d_u, int u_dtype, void *d_work, void *stream);

/* ------------------------------------------------------------- solvers */

#define FL_SOLVE_GRAPH 1     /* capture one iteration in a hipGraph and replay it */
#define FL_SOLVE_MATCHED 2   /* use the matched (Siddon-transpose) adjoint */
#define FL_SOLVE_LOWP_BF16 4 /* projector input and sinogram stored in BF16 */
#define FL_SOLVE_LOWP_F16 8  /* ... or in FP16 */

/* CGLS: CG on (A^T A + lam I) u = b.  Replaces CGLS.solveCGLS
 * (matlab/classes/+Optimizers/@CGLS/solveCGLS.m:1-31).  d_u holds u0 on entry
 * and the iterate on return.  FP64 alpha/beta/gamma on the device; the loop
 * stops once gamma = r.r < tol.  gamma_log (iters+1 doubles) and iters_done
 * are optional host outputs (requesting them synchronises). */
int fl_cgls(fl_plan plan, const float *d_b, float *d_u, float lam, int iters, double tol,
            int flags, double *gamma_log, int *iters_done, void *stream);

/* Split Bregman with the framelet analysis model.  Replaces
 * SplitBregman.solveSplitBregman (matlab/classes/+Optimizers/@SplitBregman/
 * solveSplitBregman.m:1-44): outer iterations of an `inner`-step CGLS with
 * lam = mu, shrink(v + Wu, lam/mu) (every level's low-pass band kept),
 * v <- v + (Wu - alpha), b = A^T f + mu W^T (alpha - v), with the `level`-
 * level frame of FrameletSystem(dim, kind, level) (FrameletSystem.m:18-27).
 * d_f: sinogram; d_u: u0 in, u out.  resid_log (outer doubles, optional):
 * |Wu - alpha|^2 over all levels after each outer step (requesting it
 * synchronises).  d_u_hist (device, outer volumes, optional) receives the
 * iterate after each outer step's CGLS. */
int fl_split_bregman(fl_plan plan, int frame_kind, int level, const float *d_f, float *d_u,
                     float lam, float mu, int outer, int inner, double cg_tol, int flags,
                     double *resid_log, float *d_u_hist, void *stream);

/* CGLS over caller-supplied operators (the class mirror's CGLS.solveCGLS
 * with arbitrary A / At): the vector steps of solveCGLS.m:7-30 on FP32
 * device vectors with FP64 dot products returned to the host (each call
 * with a host output synchronises the stream).  The caller forms
 * q = At(A(.)) between the calls.
 *   init:   r = q + lam u - b, p = -r, *gamma = r.r
 *   pq:     q = q + lam p, *pq = p.q
 *   update: u = u + alpha p, r = r + alpha q, *gamma = r.r
 *   dir:    p = -r + beta p */
int fl_vec_cgls_init(long long n, const float *d_q, const float *d_u, const float *d_b, float lam,
                     float *d_r, float *d_p, double *gamma, void *stream);
int fl_vec_cgls_pq(long long n, float *d_q, const float *d_p, float lam, double *pq, void *stream);
int fl_vec_cgls_update(long long n, float *d_u, float *d_r, const float *d_p, const float *d_q,
                       double alpha, double *gamma, void *stream);
int fl_vec_cgls_dir(long long n, float *d_p, const float *d_r, double beta, void *stream);

/* Soft threshold in place of framelet coefficients (`level` blocks of
 * `nbands` bands of band_elems elements): sign(a) max(|a| - tau, 0) on every
 * band but each level's low-pass band 0.  Replaces
 * FrameletExpansion.softThreshold (matlab/classes/+DataTypes/
 * @FrameletExpansion/softThreshold.m:1-80, shrinkFramelet2D/3D). */
int fl_framelet_shrink(int nbands, int level, long long band_elems, void *d_c, int c_dtype, float tau,
                       void *stream);

/* ----------------------------------------------------------- gradients */

#define FL_GRAD_EXACT_ADJOINT 1 /* use the exact-chord adjoint instead of the matched one */

/* g = A^T (A x - b) and loss = 0.5 |A x - b|^2 (loss optional; FP64).  No
 * reference counterpart: the backward of the forward projector. */
int fl_grad_data(fl_plan plan, const float *d_x, const float *d_b, float *d_g, double *loss,
                 int flags, void *stream);

/* g += W^T (lam * psi'(W x)) with psi(t) = sqrt(t^2 + eps) (smoothed l1)
 * over the bands of a `level`-level frame, loss = lam * sum psi (optional,
 * host; requesting it synchronises).  Every level's low-pass band is left
 * unpenalised (it is not a coefficient of the multi-level frame: the
 * synthesis replaces it) except the top level's when with_lowpass.  d_work:
 * level*NB*nz*ny*nx floats.  d_g is accumulated into (pass zeros for the
 * framelet term alone).  No reference counterpart: the backward of the
 * analysis transform. */
int fl_grad_framelet(int dim, int kind, int level, int nx, int ny, int nz, const float *d_x,
                     float lam, float eps, int with_lowpass, float *d_g, float *d_work,
                     double *loss, void *stream);

/* fl_grad_framelet for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results as
 * fl_grad_framelet except the loss, which goes to the device pointer d_loss
 * (one double, optional), so the call neither allocates nor synchronises.
 * Its FP64 partials live in a static device array of FL_NOALLOC_PARTIALS
 * doubles (16 MiB, held by each GPU the library runs on until the process
 * ends): ceil(nx*ny*nz/256)*level + 1 must not exceed it (FL_ERR_INVALID
 * otherwise), i.e. about 5.4e8 voxels x levels.  Calls that overlap on one
 * GPU share that array and race: order them on one stream. */
#define FL_NOALLOC_PARTIALS (1 << 21)
int fl_grad_framelet_noalloc(int dim, int kind, int level, int nx, int ny, int nz, const float *d_x,
                             float lam, float eps, int with_lowpass, float *d_g, float *d_work,
                             double *d_loss, void *stream);

/* ------------------------------------------------------------- batches
 * Many independent 2D fan-beam problems of uneven sizes (one frame each),
 * each operator one persistent launch over all of them.  Volumes and
 * sinograms are packed back to back in problem order (fl_batch_layout gives
 * the P+1 offsets, in floats).  Per-problem results are identical to the
 * single-plan calls (same per-ray / per-pixel arithmetic); CGLS and split
 * Bregman keep per-problem FP64 scalars and stopping tests.  No reference
 * counterpart: the reference runs one problem per call. */
typedef struct fl_batch_s *fl_batch;

int fl_batch_create(fl_batch *batch, int nprob, const fl_geometry *geoms);
int fl_batch_destroy(fl_batch batch);
int fl_batch_layout(fl_batch batch, long long *vol_offsets, long long *sino_offsets);
/* method 0: Siddon, 1: Gao column walk */
int fl_batch_forward(fl_batch batch, const float *d_x, float *d_y, int method, void *stream);
int fl_batch_adjoint(fl_batch batch, const float *d_y, float *d_x, int kind, void *stream);
/* gamma_log: nprob*(iters+1) doubles, iters_done: nprob ints (optional) */
int fl_batch_cgls(fl_batch batch, const float *d_b, float *d_u, float lam, int iters, double tol,
                  int flags, double *gamma_log, int *iters_done, void *stream);
/* resid_log: nprob*outer doubles (optional) */
int fl_batch_split_bregman(fl_batch batch, int frame_kind, const float *d_f, float *d_u, float lam,
                           float mu, int outer, int inner, double cg_tol, int flags,
                           double *resid_log, void *stream);
/* Times the candidate ray-tile widths and persistent grid sizes on this
 * batch and keeps the fastest (outputs optional).  Must not overlap other
 * calls on the batch. */
int fl_batch_autotune(fl_batch batch, int *fwd_tile, int *fwd_grid, int *adj_grid);

/* ------------------------------------------------- John's equation
 * Discretised John's equation on the normalised sinogram of a multi-helix
 * scan (classes/+DataTypes/@CTData/applyJohn.m:38-106 and
 * applyJohnNormal.m): g is (na, nb, nv views per helix, nz helices), a
 * fastest; D g is (na-2, nb-2, nv-2, nz-2, or 1 with two helices), p
 * fastest; FP64 results.  a, b: detector coordinates (host, na / nb
 * doubles); R = source-detector distance; h = vtab/(2 pi rps); dzeta =
 * phase shift between helices times h; dtheta = view step; da, db =
 * detector pitches (physical units). */
typedef struct {
    int na, nb, nv, nz;
    double R, SO, h, dzeta, dtheta, da, db;
    const double *a, *b;
} fl_john_geometry;
typedef struct fl_john_s *fl_john;

int fl_john_create(fl_john *john, const fl_john_geometry *g);
int fl_john_destroy(fl_john john);
int fl_john_sizes(fl_john john, long long *n_in, long long *n_out);
/* out = D g (g: FL_F32 or FL_F64) */
int fl_john_apply(fl_john john, const void *d_g, int g_dtype, double *d_out, void *stream);
/* out = D^T u */
int fl_john_adjoint(fl_john john, const double *d_u, double *d_out, void *stream);
/* out = D^T D g; d_work holds n_out doubles */
int fl_john_normal(fl_john john, const void *d_g, int g_dtype, double *d_work, double *d_out,
                   void *stream);

/* Normalised cone sinogram yNorm = y / (|d| * scale) per ray, the second
 * output of the reference's cone gateway (matlab/mex/src/Ax_cone_mf.cpp:55-57);
 * cone plans, FP32. */
int fl_forward_norm(fl_plan plan, const float *d_y, float *d_ynorm, void *stream);

/* ------------------------------------------------- FP64 reference pair
 * The Siddon forward in double (the reference's traversal rule:
 * matlab/mex/src/Ax_cone_mf_cpu_siddon.cpp:35-131 with every float
 * operation in double; view trig cos/sin of the angle in double) and its
 * voxel-driven transpose, on FP64 device buffers laid out as fl_forward /
 * fl_adjoint's; and CGLS (solveCGLS.m:7-30) with FP64 vectors over that
 * pair.  The accuracy reference for the FP32 / BF16 / FP16 paths at sizes
 * no CPU oracle reaches.  No reference counterpart (the reference is FP32).
 * fl_cgls_f64 allocates its workspace per call and returns after the solve
 * (it synchronises the stream). */
int fl_forward_f64(fl_plan plan, const double *d_x, double *d_y, void *stream);
int fl_adjoint_f64(fl_plan plan, const double *d_y, double *d_x, void *stream);
int fl_cgls_f64(fl_plan plan, const double *d_b, double *d_u, double lam, int iters, double tol,
                double *gamma_log, int *iters_done, void *stream);
/* Split Bregman (solveSplitBregman.m:16-41, as fl_split_bregman) over the
 * FP64 pair with a one-level framelet analysis / synthesis in double:
 * d_f and d_u (u0 in, u out) are FP64; resid_log and d_u_hist as
 * fl_split_bregman's.  Allocates its workspace per call and synchronises
 * the stream. */
int fl_split_bregman_f64(fl_plan plan, int frame_kind, const double *d_f, double *d_u, double lam,
                         double mu, int outer, int inner, double *resid_log, double *d_u_hist,
                         void *stream);

/* Traversal log of selected rays (testing): rays = (view, ia, ib) triples
 * (host), cap entries per ray; J/S/cnt host outputs of nrays*cap / nrays. */
int fl_siddon_trace(fl_plan plan, int nrays, const int *rays, int cap, int *J, float *S, int *cnt);

#ifdef __cplusplus
}
#endif

#endif /* FRAMELAB_HIP_H */
