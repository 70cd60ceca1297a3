// fl_geom.h -- device-side geometry bundles passed by value to the kernels.
#pragma once

#include <stddef.h>

// Forward projector geometry (fan: nz == 0, nb == 1).
struct FwdGeom {
    const float* cs;        // cosf(sd_phi[v]) per view (reference forward trig)
    const float* sn;        // sinf(sd_phi[v])
    const float* sdz;       // source z per view (cone)
    const float* ydet;      // detector column coordinates
    const float* zdet;      // detector row coordinates (cone)
    const int* slot_view;   // output slot -> view index
    const int* view_frame;  // view -> image frame (id_X)
    float SO, OD, scale, dz0;
    int nx, ny, nz, na, nb, nslots, nframes;
    size_t nvox;            // voxels per frame
    // detector index maps for voxel shadows: cell a is at y_det[0] + a*pitch
    float ya0, inv_pa, zb0, inv_pb;
    int tile;               // cone forward launch shape, 0..FL_FWD_NTILES-1
    void* bricks;           // cone: bricked copy of the input frames (plan-owned)
};

// Voxel-driven adjoint geometry (the reference para tables).
struct AdjGeom {
    const float* cphi;      // para.cos_phi
    const float* sphi;      // para.sin_phi
    const float* cdet;      // para.cos_det (fan)
    const float* sdet;      // para.sin_det (fan)
    const float* sdz;       // para.sd_z (cone)
    const float* ydet;
    const float* zdet;
    const int* id_Y;        // frame -> views, tmp_size per frame
    const int* id_S;        // sinogram slot holding the j-th view of frame it (id_Y's layout):
                            // the slot the forward writes that view to
    const int* Nv;          // views per frame
    float SO, OD, scale, dy_det, y_os, dz_det, dz;
    int nx, ny, nz, na, nb, nt, tmp_size;
    size_t nvox;
    size_t nd;              // detector cells per view
    float ya0, inv_pa, zb0, inv_pb;
    // cone exact adjoint: bit 0 (columns) / bit 1 (rows) set where the
    // reference's detector window may cut a voxel's shadow window and has
    // to be applied; clear where the plan's geometry rules that out
    int win_ref;
    // fan: per (view, cell) the view-only quotients of the line chord,
    // (1/max(|a|,|b|), 1/(|a||b|)), at [iv*nd + id]
    const float2* line_q;
    // fan: per (view, cell) (1/dx, 1/dy, |d|) of the ray, for the matched adjoint
    const float4* ray_q;
};
