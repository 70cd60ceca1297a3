// ref_shim.cpp -- C-linkage entry points onto the reference's own CPU kernels.
//
// TEST INFRASTRUCTURE ONLY.  Linked into oracle/_ref/libframelab_ref.so
// together with the reference sources, compiled where they lie under the
// reference checkout (see oracle/Makefile).  The prototypes below restate the
// signatures declared in the reference headers:
//   matlab/mex/src/sort_alpha.h, Ax_fan_mf.h:2, Ax_cone_mf.h:9,
//   find_l_3d.h, Atx_cone_mf.h:2, Ax_fan_mf.h:6 (Gao 2D forward), find_area.h,
//   Ax_fan_mf.h and Atx_fan_mf.h (finite-width beam, version 2).
int sort_alpha(int Nx, int Ny, float *alpha_x, float *alpha_y, float *alpha);
void Ax_fan_mf_cpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                          int nv, float *sd_phi, int nd, float *y_det, int *id_X);
void Ax_cone_mf_cpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                           int nz, int nv, float *sd_phi, float *sd_z, int na, int nb,
                           float *y_det, float *z_det, int *id_X, float dz0);
float find_l(float x1_0, float y1_0, float x2_0, float y2_0, float dx, float dy, float x, float y);
float find_l_3d(float x1_0, float y1_0, float z1_0, float x2_0, float y2_0, float z2_0,
                float dx, float dy, float dz, float x, float y, float z);
void Atx_cone_mf_cpu_new(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                         int nz, float *sd_z, float *y_det, float *z_det, float dy_det,
                         float y_os, float dz_det, float dz, int nt, int *id_Y, int *Nv,
                         int tmp_size, int nv, float *cos_phi, float *sin_phi, int na, int nb);
void Ax_fan_mf_cpu_new(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                       int nv, float *sd_phi, int nd, float *y_det, int *id_X);
float find_area(float a1, float b1, float c1, float a2, float b2, float c2, float x, float y);
void Ax_fan_mf_cpu_new_fb(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                          int nv, float *sd_phi, int nd, float *y_det, int *id_X, float dy_det);
void Atx_fan_mf_cpu_new_fb(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                           float *sd_phi, int nd, float *y_det, float dy_det, int nt, int *id_Y,
                           int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                           float *cos_det, float *sin_det);

extern "C" {
int ref_sort_alpha(int Nx, int Ny, float *ax, float *ay, float *out)
{
    return sort_alpha(Nx, Ny, ax, ay, out);
}
void ref_Ax_fan_mf_cpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx,
                              int ny, int nv, float *sd_phi, int nd, float *y_det, int *id_X)
{
    Ax_fan_mf_cpu_siddon(X, y, SO, OD, scale, nx, ny, nv, sd_phi, nd, y_det, id_X);
}
void ref_Ax_cone_mf_cpu_siddon(float *X, float *y, float SO, float OD, float scale, int nx,
                               int ny, int nz, int nv, float *sd_phi, float *sd_z, int na,
                               int nb, float *y_det, float *z_det, int *id_X, float dz0)
{
    Ax_cone_mf_cpu_siddon(X, y, SO, OD, scale, nx, ny, nz, nv, sd_phi, sd_z, na, nb, y_det,
                          z_det, id_X, dz0);
}
float ref_find_l2(float x1, float y1, float x2, float y2, float dx, float dy, float x, float y)
{
    return find_l(x1, y1, x2, y2, dx, dy, x, y);
}
float ref_find_l_3d(float x1, float y1, float z1, float x2, float y2, float z2, float dx,
                    float dy, float dz, float x, float y, float z)
{
    return find_l_3d(x1, y1, z1, x2, y2, z2, dx, dy, dz, x, y, z);
}
void ref_Atx_cone_mf_cpu_new(float *x, float *Y, float SO, float OD, float scale, int nx,
                             int ny, int nz, float *sd_z, float *y_det, float *z_det,
                             float dy_det, float y_os, float dz_det, float dz, int nt, int *id_Y,
                             int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                             int na, int nb)
{
    Atx_cone_mf_cpu_new(x, Y, SO, OD, scale, nx, ny, nz, sd_z, y_det, z_det, dy_det, y_os,
                        dz_det, dz, nt, id_Y, Nv, tmp_size, nv, cos_phi, sin_phi, na, nb);
}
void ref_Ax_fan_mf_cpu_new(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                           int nv, float *sd_phi, int nd, float *y_det, int *id_X)
{
    Ax_fan_mf_cpu_new(X, y, SO, OD, scale, nx, ny, nv, sd_phi, nd, y_det, id_X);
}
float ref_find_area(float a1, float b1, float c1, float a2, float b2, float c2, float x, float y)
{
    return find_area(a1, b1, c1, a2, b2, c2, x, y);
}
void ref_Ax_fan_mf_cpu_new_fb(float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                              int nv, float *sd_phi, int nd, float *y_det, int *id_X, float dy_det)
{
    Ax_fan_mf_cpu_new_fb(X, y, SO, OD, scale, nx, ny, nv, sd_phi, nd, y_det, id_X, dy_det);
}
void ref_Atx_fan_mf_cpu_new_fb(float *x, float *Y, float SO, float OD, float scale, int nx, int ny,
                               float *sd_phi, int nd, float *y_det, float dy_det, int nt, int *id_Y,
                               int *Nv, int tmp_size, int nv, float *cos_phi, float *sin_phi,
                               float *cos_det, float *sin_det)
{
    Atx_fan_mf_cpu_new_fb(x, Y, SO, OD, scale, nx, ny, sd_phi, nd, y_det, dy_det, nt, id_Y, Nv,
                          tmp_size, nv, cos_phi, sin_phi, cos_det, sin_det);
}
}
