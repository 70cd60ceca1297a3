// A C++ host program on the library's C ABI alone (include/framelab_hip.h and
// libframelab_hip.so, no Python): what a MEX gateway or another native caller
// does.  On a 32^3 cone geometry (24 views, 48 x 40 cells) it checks that
//   * the plan API's forward on device buffers equals the drop-in
//     Ax_cone_mf_gpu_siddon on host buffers bit for bit, and the plan's
//     exact adjoint equals the drop-in Atx_cone_mf_gpu_new;
//   * <A x, y> = <x, A^T y> for the matched adjoint (FP64 dots, <= 1e-5);
//   * five graph-replayed CGLS iterations run and decrease gamma;
//   * a bad geometry comes back as a negative status with a message.
// Prints "host example: OK" and exits 0, or names the failed check and exits 1.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "framelab_hip.h"

namespace {

int g_fail = 0;

void check(bool ok, const char* what)
{
    if (!ok) {
        std::printf("host example: FAILED %s\n", what);
        g_fail = 1;
    }
}

bool status_ok(int rc, const char* what)
{
    if (rc != FL_OK) {
        std::printf("host example: %s returned %d (%s)\n", what, rc, fl_last_error_message());
        g_fail = 1;
        return false;
    }
    return true;
}

bool same_bits(const std::vector<float>& a, const std::vector<float>& b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}

}  // namespace

int main()
{
    const int n = 32, na = 48, nb = 40, nv = 24;
    const float scale = 2.0f / n, SO = 4.0f / scale, OD = 4.0f / scale, pitch = 1.5f;
    std::vector<float> phi(nv), sdz(nv, 0.0f), cphi(nv), sphi(nv), ydet(na), zdet(nb);
    std::vector<int> idY(nv), Nv(1, nv);
    for (int v = 0; v < nv; v++) {
        phi[v] = (float)(2.0 * M_PI * v / nv);
        cphi[v] = (float)std::cos((double)phi[v]);
        sphi[v] = (float)std::sin((double)phi[v]);
        idY[v] = v;
    }
    for (int a = 0; a < na; a++) ydet[a] = ((float)a - na / 2 + 0.5f) * pitch;
    for (int b = 0; b < nb; b++) zdet[b] = ((float)b - nb / 2 + 0.5f) * pitch;

    const size_t nvox = (size_t)n * n * n, nsino = (size_t)nv * na * nb;
    std::vector<float> x(nvox), y(nsino);
    unsigned s = 12345u;   // small LCG: values in [0, 1) and [-1, 1)
    auto rnd = [&s] { s = s * 1664525u + 1013904223u; return (float)(s >> 8) / 16777216.0f; };
    for (auto& v : x) v = rnd();
    for (auto& v : y) v = 2.0f * rnd() - 1.0f;

    fl_geometry g;
    std::memset(&g, 0, sizeof g);
    g.nx = g.ny = g.nz = n; g.nt = 1; g.nv = nv; g.na = na; g.nb = nb;
    g.SO = SO; g.OD = OD; g.scale = scale; g.dz = 1.0f; g.dy_det = pitch; g.dz_det = pitch; g.y_os = 0.0f;
    g.sd_phi = phi.data(); g.sd_z = sdz.data(); g.y_det = ydet.data(); g.z_det = zdet.data();
    g.cos_phi = cphi.data(); g.sin_phi = sphi.data();
    g.id_Y = idY.data(); g.Nv = Nv.data(); g.tmp_size = nv;

    fl_plan plan = nullptr;
    if (!status_ok(fl_plan_create(&plan, &g), "fl_plan_create")) return 1;
    float *dx = nullptr, *dy = nullptr, *dxb = nullptr, *du = nullptr;
    if (hipMalloc(&dx, nvox * 4) != hipSuccess || hipMalloc(&dy, nsino * 4) != hipSuccess ||
        hipMalloc(&dxb, nvox * 4) != hipSuccess || hipMalloc(&du, nvox * 4) != hipSuccess) {
        std::printf("host example: hipMalloc failed\n");
        return 1;
    }
    hipStream_t st;
    (void)hipStreamCreate(&st);
    (void)hipMemcpy(dx, x.data(), nvox * 4, hipMemcpyHostToDevice);

    // plan forward vs drop-in forward
    std::vector<float> y_plan(nsino), y_drop(nsino);
    status_ok(fl_forward(plan, dx, FL_F32, dy, FL_F32, st), "fl_forward");
    (void)hipStreamSynchronize(st);
    (void)hipMemcpy(y_plan.data(), dy, nsino * 4, hipMemcpyDeviceToHost);
    Ax_cone_mf_gpu_siddon(x.data(), y_drop.data(), phi.data(), sdz.data(), ydet.data(), zdet.data(), idY.data(),
                          Nv.data(), SO, OD, scale, 1.0f, n, n, n, 1, na, nb, nv, nv, 4);
    check(fl_get_last_error() == FL_OK, "drop-in forward status");
    check(same_bits(y_plan, y_drop), "plan forward == drop-in forward");

    // plan exact adjoint vs drop-in exact adjoint
    std::vector<float> x_plan(nvox), x_drop(nvox);
    (void)hipMemcpy(dy, y.data(), nsino * 4, hipMemcpyHostToDevice);
    status_ok(fl_adjoint(plan, dy, FL_F32, dxb, FL_F32, FL_ADJ_EXACT, st), "fl_adjoint");
    (void)hipStreamSynchronize(st);
    (void)hipMemcpy(x_plan.data(), dxb, nvox * 4, hipMemcpyDeviceToHost);
    Atx_cone_mf_gpu_new(x_drop.data(), y.data(), cphi.data(), sphi.data(), sdz.data(), ydet.data(), zdet.data(),
                        idY.data(), Nv.data(), SO, OD, scale, pitch, 0.0f, pitch, 1.0f, n, n, n, 1, na, nb, nv, nv,
                        4);
    check(fl_get_last_error() == FL_OK, "drop-in adjoint status");
    check(same_bits(x_plan, x_drop), "plan exact adjoint == drop-in exact adjoint");

    // dot test of the forward with the matched adjoint
    std::vector<float> xm(nvox);
    status_ok(fl_adjoint(plan, dy, FL_F32, dxb, FL_F32, FL_ADJ_MATCHED, st), "fl_adjoint matched");
    (void)hipStreamSynchronize(st);
    (void)hipMemcpy(xm.data(), dxb, nvox * 4, hipMemcpyDeviceToHost);
    double lhs = 0.0, rhs = 0.0;
    for (size_t i = 0; i < nsino; i++) lhs += (double)y_plan[i] * (double)y[i];
    for (size_t i = 0; i < nvox; i++) rhs += (double)x[i] * (double)xm[i];
    const double rel = std::fabs(lhs - rhs) / std::fabs(lhs);
    std::printf("host example: dot test (matched) rel %.2e\n", rel);
    check(rel <= 1e-5, "dot test");

    // five graph-replayed CGLS iterations on b = A^T A x
    status_ok(fl_forward(plan, dx, FL_F32, dy, FL_F32, st), "fl_forward (rhs)");
    status_ok(fl_adjoint(plan, dy, FL_F32, dxb, FL_F32, FL_ADJ_MATCHED, st), "fl_adjoint (rhs)");
    (void)hipMemsetAsync(du, 0, nvox * 4, st);
    double gam[6] = {0};
    int done = 0;
    status_ok(fl_cgls(plan, dxb, du, 0.0f, 5, 0.0, FL_SOLVE_GRAPH | FL_SOLVE_MATCHED, gam, &done, st), "fl_cgls");
    std::printf("host example: CGLS gamma %.3e -> %.3e in %d iterations\n", gam[0], gam[done], done);
    check(done == 5 && gam[5] < gam[0], "CGLS decreases gamma");

    // a bad geometry is a status, not an abort
    fl_geometry bad = g;
    bad.nx = 31;   // odd sides are rejected
    fl_plan p2 = nullptr;
    const int rc = fl_plan_create(&p2, &bad);
    check(rc < 0 && std::strlen(fl_last_error_message()) > 0, "odd volume rejected with a message");

    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dxb); (void)hipFree(du);
    (void)hipStreamDestroy(st);
    fl_plan_destroy(plan);
    if (!g_fail) std::printf("host example: OK\n");
    return g_fail;
}
