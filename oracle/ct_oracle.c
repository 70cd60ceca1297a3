/*
 * ct_oracle.c -- CPU restatement of FrameLab's CT projector arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP product path is
 * compared with: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load the library built from it.  Nothing in
 * framelab_amd/ links it.
 *
 * Every routine restates the reference algorithm in FP32 with the reference's
 * operation order (no FMA contraction: built with -ffp-contract=off), so that
 * voxel indices, intersection lengths and ray sums are bit-identical to the
 * reference CPU mex path.  FP64 twins give the accuracy oracle.
 *
 * Reference files followed (paths relative to the reference repository):
 *   matlab/mex/src/sort_alpha.cpp:15-60            Siddon alpha merge
 *   matlab/mex/src/Ax_fan_mf_cpu_siddon.cpp:19-93  2D Siddon forward
 *   matlab/mex/src/Ax_cone_mf_cpu_siddon.cpp:19-134 3D Siddon forward
 *   matlab/mex/src/find_l.cpp:18-45                2D line/pixel chord
 *   matlab/mex/src/find_l_3d.cpp:19-182            3D ray/voxel chord
 *   matlab/mex/src/Atx_fan_mf_cpu_new.cpp:21-63    2D pixel-driven adjoint
 *   matlab/mex/src/Atx_cone_mf_gpu_new.cu:41-84    3D voxel-driven adjoint
 *
 * Pinning (see tests/test_oracle_*.py): bit-exact against the reference's
 * own CPU kernels compiled from source into oracle/_ref, against committed
 * golden vectors under tests/golden, and against closed forms (box, disc and
 * ball chords, brute-force segment clipping in FP64).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF (1.0f / 0.0f)

/* ------------------------------------------------------------------------ */
/* Siddon alpha lists and their merge                                        */
/* ------------------------------------------------------------------------ */

/* Sorted union of two ascending lists; a value present in both lists is
 * emitted once.  This is the observable behaviour of sort_alpha
 * (matlab/mex/src/sort_alpha.cpp:27-58: runs of one list are emitted while
 * <= the head of the other; an equal head of the other list is skipped). */
int orc_merge_union(int na, const float *a, int nb, const float *b, float *out)
{
    int i = 0, j = 0, n = 0;
    while (i < na || j < nb) {
        if (j >= nb) {
            out[n++] = a[i++];
        } else if (i >= na) {
            out[n++] = b[j++];
        } else if (a[i] < b[j]) {
            out[n++] = a[i++];
        } else if (b[j] < a[i]) {
            out[n++] = b[j++];
        } else {
            out[n++] = a[i];
            i++;
            j++;
        }
    }
    return n;
}

/* Plane-crossing parameters of one axis in ascending order.
 * Planes sit at coordinates i*s for i=-n/2..n/2 (s=1 for x,y; s=dz0 for z).
 * alpha = (i*s - o)/d, walked from -n/2 up when d>0 and from n/2 down when
 * d<0 (Ax_cone_mf_cpu_siddon.cpp:49-74).  Returns the count, 0 if d==0. */
static int axis_alphas(int n, float s, float o, float d, float *out)
{
    int h = n / 2, j = 0, i;
    if (d > 0) {
        for (i = -h; i <= h; i++) out[j++] = ((float)i * s - o) / d;
    } else if (d < 0) {
        for (i = h; i >= -h; i--) out[j++] = ((float)i * s - o) / d;
    }
    return j;
}

/* Build the merged alpha list of one ray (union of the active axes). */
static int ray_alphas(int nx, int ny, int nz, float dz0,
                      float x1, float y1, float z1, float dx, float dy, float dz,
                      float *ax, float *ay, float *az, float *axy, float *out)
{
    int cx = axis_alphas(nx, 1.0f, x1, dx, ax);
    int cy = axis_alphas(ny, 1.0f, y1, dy, ay);
    int cz = (nz > 0) ? axis_alphas(nz, dz0, z1, dz, az) : 0;
    int cxy = orc_merge_union(cx, ax, cy, ay, axy);
    return orc_merge_union(cxy, axy, cz, az, out);
}

/* Walk the merged list: each segment's voxel is the floor of its midpoint,
 * the length is (a[i+1]-a[i])*|d| (Ax_cone_mf_cpu_siddon.cpp:114-128).
 * nz==0 selects the 2D walk (Ax_fan_mf_cpu_siddon.cpp:77-88).
 * When J/S are non-null the in-volume segments are logged in order. */
static float walk(int n_alpha, const float *al, int nx, int ny, int nz, float dz0,
                  float x1, float y1, float z1, float dx, float dy, float dz,
                  float totallength, const float *x, int *J, float *S, int *count)
{
    int nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2, i, k = 0;
    float acc = 0.0f;
    for (i = 0; i < n_alpha - 1; i++) {
        float ac = (al[i] + al[i + 1]) * 0.5f;
        float fx = x1 + ac * dx;
        int tx = (int)floorf(fx + (float)nx2);
        if (tx < 0 || tx >= nx) continue;
        float fy = y1 + ac * dy;
        int ty = (int)floorf(fy + (float)ny2);
        if (ty < 0 || ty >= ny) continue;
        int Ji;
        if (nz > 0) {
            float fz = (z1 + ac * dz) / dz0;
            int tz = (int)floorf(fz + (float)nz2);
            if (tz < 0 || tz >= nz) continue;
            Ji = (tz * ny + ty) * nx + tx;
        } else {
            Ji = ty * nx + tx;
        }
        float Si = (al[i + 1] - al[i]) * totallength;
        if (x) acc = acc + Si * x[Ji];
        if (J) { J[k] = Ji; S[k] = Si; }
        k++;
    }
    if (count) *count = k;
    return acc;
}

/* Per-view trig the way the reference's C++ computes it: (float)cos(float)
 * resolves to the float overload (Ax_fan_mf_cpu_siddon.cpp:34). */
void orc_trig_f32(int nv, const float *phi, float *c, float *s)
{
    for (int i = 0; i < nv; i++) { c[i] = cosf(phi[i]); s[i] = sinf(phi[i]); }
}

/* 2D fan Siddon, Ax_fan_mf_cpu_siddon.cpp:19-93.  cos_phi/sin_phi are the
 * per-view values from orc_trig_f32. */
void orc_siddon_fan_f32(const float *X, float *y, float SO, float OD, float scale,
                        int nx, int ny, int nv, const float *cos_phi, const float *sin_phi,
                        int nd, const float *y_det, const int *id_X)
{
    int n = nx * ny;
#pragma omp parallel
    {
        float *ax = malloc((nx + 1) * sizeof(float));
        float *ay = malloc((ny + 1) * sizeof(float));
        float *al = malloc((nx + ny + 2) * sizeof(float));
#pragma omp for schedule(static)
        for (int iv = 0; iv < nv; iv++) {
            const float *x = X + (size_t)id_X[iv] * n;
            float c = cos_phi[iv], s = sin_phi[iv];
            float x1 = c * (-SO), y1 = s * (-SO);
            for (int id = 0; id < nd; id++) {
                float x2 = c * OD - s * y_det[id];
                float y2 = s * OD + c * y_det[id];
                float dx = x2 - x1, dy = y2 - y1;
                int cx = axis_alphas(nx, 1.0f, x1, dx, ax);
                int cy = axis_alphas(ny, 1.0f, y1, dy, ay);
                int na = orc_merge_union(cx, ax, cy, ay, al);
                float L = sqrtf(dx * dx + dy * dy);
                float acc = walk(na, al, nx, ny, 0, 1.0f, x1, y1, 0.0f, dx, dy, 0.0f, L, x, 0, 0, 0);
                y[(size_t)iv * nd + id] = acc * scale;
            }
        }
        free(ax); free(ay); free(al);
    }
}

/* 3D cone Siddon, Ax_cone_mf_cpu_siddon.cpp:19-134. */
void orc_siddon_cone_f32(const float *X, float *y, float SO, float OD, float scale,
                         int nx, int ny, int nz, int nv, const float *cos_phi,
                         const float *sin_phi, const float *sd_z, int na, int nb,
                         const float *y_det, const float *z_det, const int *id_X, float dz0)
{
    size_t n = (size_t)nx * ny * nz, nd = (size_t)na * nb;
    int rays = nv * nb;
#pragma omp parallel
    {
        float *ax = malloc((nx + 1) * sizeof(float));
        float *ay = malloc((ny + 1) * sizeof(float));
        float *az = malloc((nz + 1) * sizeof(float));
        float *axy = malloc((nx + ny + 2) * sizeof(float));
        float *al = malloc((nx + ny + nz + 3) * sizeof(float));
#pragma omp for schedule(dynamic, 4)
        for (int r = 0; r < rays; r++) {
            int iv = r / nb, ib = r % nb;
            const float *x = X + (size_t)id_X[iv] * n;
            float c = cos_phi[iv], s = sin_phi[iv];
            float x1 = c * (-SO), y1 = s * (-SO), z1 = sd_z[iv];
            for (int ia = 0; ia < na; ia++) {
                float x2 = c * OD - s * y_det[ia];
                float y2 = s * OD + c * y_det[ia];
                float z2 = z_det[ib] + sd_z[iv];
                float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
                int nal = ray_alphas(nx, ny, nz, dz0, x1, y1, z1, dx, dy, dz, ax, ay, az, axy, al);
                float L = sqrtf(dx * dx + dy * dy + dz * dz);
                float acc = walk(nal, al, nx, ny, nz, dz0, x1, y1, z1, dx, dy, dz, L, x, 0, 0, 0);
                y[(size_t)iv * nd + (size_t)ib * na + ia] = acc * scale;
            }
        }
        free(ax); free(ay); free(az); free(axy); free(al);
    }
}

/* Traversal log of a list of cone rays (nz>0) or fan rays (nz==0; ib ignored).
 * rays[3*k] = view, rays[3*k+1] = ia, rays[3*k+2] = ib.  Segment k of ray r
 * is written at J[off[r]+k], S[off[r]+k]; off has nrays+1 entries and
 * off[nrays] is the total count.  cap bounds the output arrays; returns -1
 * when exceeded. */
int orc_siddon_trace_f32(int nrays, const int *rays, float SO, float OD,
                         int nx, int ny, int nz, const float *cos_phi, const float *sin_phi,
                         const float *sd_z, const float *y_det, const float *z_det, float dz0,
                         int cap, int *J, float *S, int *off)
{
    float *ax = malloc((nx + 1) * sizeof(float));
    float *ay = malloc((ny + 1) * sizeof(float));
    float *az = malloc((nz + 2) * sizeof(float));
    float *axy = malloc((nx + ny + 2) * sizeof(float));
    float *al = malloc((nx + ny + nz + 3) * sizeof(float));
    int total = 0, rc = 0;
    for (int r = 0; r < nrays; r++) {
        int iv = rays[3 * r], ia = rays[3 * r + 1], ib = rays[3 * r + 2];
        float c = cos_phi[iv], s = sin_phi[iv];
        float x1 = c * (-SO), y1 = s * (-SO);
        float x2 = c * OD - s * y_det[ia];
        float y2 = s * OD + c * y_det[ia];
        float dx = x2 - x1, dy = y2 - y1;
        float z1 = 0.0f, dz = 0.0f, L;
        int nal;
        if (nz > 0) {
            z1 = sd_z[iv];
            float z2 = z_det[ib] + sd_z[iv];
            dz = z2 - z1;
            nal = ray_alphas(nx, ny, nz, dz0, x1, y1, z1, dx, dy, dz, ax, ay, az, axy, al);
            L = sqrtf(dx * dx + dy * dy + dz * dz);
        } else {
            int cx = axis_alphas(nx, 1.0f, x1, dx, ax);
            int cy = axis_alphas(ny, 1.0f, y1, dy, ay);
            nal = orc_merge_union(cx, ax, cy, ay, al);
            L = sqrtf(dx * dx + dy * dy);
        }
        off[r] = total;
        if (total + nal > cap) { rc = -1; break; }
        int cnt = 0;
        walk(nal, al, nx, ny, nz, dz0, x1, y1, z1, dx, dy, dz, L, 0, J + total, S + total, &cnt);
        total += cnt;
    }
    off[nrays] = total;
    free(ax); free(ay); free(az); free(axy); free(al);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* FP64 Siddon (accuracy oracle; same traversal rule in double)             */
/* ------------------------------------------------------------------------ */

static int axis_alphas_d(int n, double s, double o, double d, double *out)
{
    int h = n / 2, j = 0, i;
    if (d > 0) {
        for (i = -h; i <= h; i++) out[j++] = ((double)i * s - o) / d;
    } else if (d < 0) {
        for (i = h; i >= -h; i--) out[j++] = ((double)i * s - o) / d;
    }
    return j;
}

static int merge_union_d(int na, const double *a, int nb, const double *b, double *out)
{
    int i = 0, j = 0, n = 0;
    while (i < na || j < nb) {
        if (j >= nb) out[n++] = a[i++];
        else if (i >= na) out[n++] = b[j++];
        else if (a[i] < b[j]) out[n++] = a[i++];
        else if (b[j] < a[i]) out[n++] = b[j++];
        else { out[n++] = a[i]; i++; j++; }
    }
    return n;
}

/* One FP64 ray: returns sum over segments of len*x[J] (x may be NULL), and
 * when g is non-null scatters w*len into g[J] (the exact transpose). */
static double ray_f64(int nx, int ny, int nz, double dz0,
                      double x1, double y1, double z1, double x2, double y2, double z2,
                      const double *x, double *g, double w,
                      double *ax, double *ay, double *az, double *axy, double *al)
{
    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    int cx = axis_alphas_d(nx, 1.0, x1, dx, ax);
    int cy = axis_alphas_d(ny, 1.0, y1, dy, ay);
    int cz = nz > 0 ? axis_alphas_d(nz, dz0, z1, dz, az) : 0;
    int cxy = merge_union_d(cx, ax, cy, ay, axy);
    int n = merge_union_d(cxy, axy, cz, az, al);
    double L = sqrt(dx * dx + dy * dy + dz * dz), acc = 0.0;
    int nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
    for (int i = 0; i < n - 1; i++) {
        double ac = 0.5 * (al[i] + al[i + 1]);
        int tx = (int)floor(x1 + ac * dx + nx2);
        if (tx < 0 || tx >= nx) continue;
        int ty = (int)floor(y1 + ac * dy + ny2);
        if (ty < 0 || ty >= ny) continue;
        int Ji;
        if (nz > 0) {
            int tz = (int)floor((z1 + ac * dz) / dz0 + nz2);
            if (tz < 0 || tz >= nz) continue;
            Ji = (tz * ny + ty) * nx + tx;
        } else {
            Ji = ty * nx + tx;
        }
        double Si = (al[i + 1] - al[i]) * L;
        if (x) acc += Si * x[Ji];
        if (g) g[Ji] += w * Si;
    }
    return acc;
}

/* FP64 cone (nz>0) or fan (nz==0, nb==1, z ignored) forward; per-view
 * cos/sin of the view angle are inputs. */
void orc_siddon_f64(const double *X, double *y, double SO, double OD, double scale,
                    int nx, int ny, int nz, int nv, const double *cphi, const double *sphi, const double *sd_z,
                    int na, int nb, const double *y_det, const double *z_det, double dz0)
{
    size_t nd = (size_t)na * nb;
    int rays = nv * nb;
#pragma omp parallel
    {
        double *ax = malloc((nx + 1) * sizeof(double));
        double *ay = malloc((ny + 1) * sizeof(double));
        double *az = malloc((nz + 2) * sizeof(double));
        double *axy = malloc((nx + ny + 2) * sizeof(double));
        double *al = malloc((nx + ny + nz + 3) * sizeof(double));
#pragma omp for schedule(dynamic, 4)
        for (int r = 0; r < rays; r++) {
            int iv = r / nb, ib = r % nb;
            double c = cphi[iv], s = sphi[iv];
            double zs = nz > 0 ? sd_z[iv] : 0.0;
            for (int ia = 0; ia < na; ia++) {
                double x1 = -SO * c, y1 = -SO * s;
                double x2 = c * OD - s * y_det[ia], y2 = s * OD + c * y_det[ia];
                double z2 = nz > 0 ? z_det[ib] + zs : 0.0;
                y[(size_t)iv * nd + (size_t)ib * na + ia] =
                    scale * ray_f64(nx, ny, nz, dz0, x1, y1, zs, x2, y2, z2, X, 0, 0.0, ax, ay, az, axy, al);
            }
        }
        free(ax); free(ay); free(az); free(axy); free(al);
    }
}

/* FP64 exact transpose of orc_siddon_f64 (ray-driven scatter).  Runs
 * serially so the summation order is fixed. */
void orc_siddon_adj_f64(double *X, const double *y, double SO, double OD, double scale,
                        int nx, int ny, int nz, int nv, const double *cphi, const double *sphi, const double *sd_z,
                        int na, int nb, const double *y_det, const double *z_det, double dz0)
{
    size_t nvox = (size_t)nx * ny * (nz > 0 ? nz : 1), nd = (size_t)na * nb;
    double *ax = malloc((nx + 1) * sizeof(double));
    double *ay = malloc((ny + 1) * sizeof(double));
    double *az = malloc((nz + 2) * sizeof(double));
    double *axy = malloc((nx + ny + 2) * sizeof(double));
    double *al = malloc((nx + ny + nz + 3) * sizeof(double));
    memset(X, 0, nvox * sizeof(double));
    for (int iv = 0; iv < nv; iv++) {
        double c = cphi[iv], s = sphi[iv];
        double zs = nz > 0 ? sd_z[iv] : 0.0;
        for (int ib = 0; ib < nb; ib++)
            for (int ia = 0; ia < na; ia++) {
                double w = y[(size_t)iv * nd + (size_t)ib * na + ia] * scale;
                if (w == 0.0) continue;
                double x1 = -SO * c, y1 = -SO * s;
                double x2 = c * OD - s * y_det[ia], y2 = s * OD + c * y_det[ia];
                double z2 = nz > 0 ? z_det[ib] + zs : 0.0;
                ray_f64(nx, ny, nz, dz0, x1, y1, zs, x2, y2, z2, 0, X, w, ax, ay, az, axy, al);
            }
    }
    free(ax); free(ay); free(az); free(axy); free(al);
}

/* ------------------------------------------------------------------------ */
/* Exact chord functions                                                     */
/* ------------------------------------------------------------------------ */

#define FABS(v) ((v) > 0 ? (v) : -(v))

/* Chord of the line a*x+b*y=c (unit normal) through the unit pixel centred
 * at (x,y).  find_l.cpp:18-45. */
float orc_find_l(float a, float b, float c, float x, float y)
{
    float a2 = FABS(a), b2 = FABS(b);
    float hx = a2 / 2, hy = b2 / 2;
    float dmax = hx + hy;
    float t = c - a * x - b * y;
    float d = FABS(t);
    float l = 0;
    if (d < dmax) {
        float diff = hx - hy;
        float d0 = FABS(diff);
        float flat = (hx < hy) ? (float)1.0 / b2 : (float)1.0 / a2;
        l = (d <= d0) ? flat : (dmax - d) / (a2 * b2);
    }
    return l;
}

/* Chord of the segment-line (x1,y1)->(x2,y2) through the dx-by-dy cell
 * centred at (x,y).  find_l_3d.cpp:19-92 (2D helper). */
float orc_find_l2(float x1_0, float y1_0, float x2_0, float y2_0, float dx, float dy,
                  float x, float y)
{
    float a = x2_0 - x1_0, b = y2_0 - y1_0;
    float hx = dx / 2, hy = dy / 2;
    float xi[2], yi[2];
    int k = 0;
    if (a == 0) {
        float t = FABS(x1_0 - x);
        return (t <= hx) ? dy : 0.0f;
    }
    if (b == 0) {
        float t = FABS(y1_0 - y);
        return (t <= hy) ? dx : 0.0f;
    }
    float px = x1_0 - x, py = y1_0 - y;
    /* x-faces: y at x=-hx and x=+hx; y-faces: x at y=-hy and y=+hy.  The
     * face order depends on which of |a|,|b| dominates (find_l_3d.cpp:42-83). */
    int xfirst = FABS(a) > FABS(b);
    for (int pass = 0; pass < 2 && k < 2; pass++) {
        int xf = (pass == 0) ? xfirst : !xfirst;
        if (xf) {
            float sl = b / a, m = sl * (-px) + py, w = sl * hx;
            if (FABS(m - w) <= hy) { xi[k] = -hx; yi[k] = m - w; k++; }
            if ((pass == 0 || k < 2) && FABS(m + w) <= hy) { xi[k] = hx; yi[k] = m + w; k++; }
        } else {
            float sl = a / b, m = sl * (-py) + px, w = sl * hy;
            if (FABS(m - w) <= hx) { yi[k] = -hy; xi[k] = m - w; k++; }
            if ((pass == 0 || k < 2) && FABS(m + w) <= hx) { yi[k] = hy; xi[k] = m + w; k++; }
        }
    }
    if (k == 2) {
        float u = xi[1] - xi[0], v = yi[1] - yi[0];
        return sqrtf(u * u + v * v);
    }
    return 0.0f;
}

/* One face pair of find_l_3d: the line meets the two faces normal to axis
 * `ax` (0=x,1=y,2=z) at offsets -h[ax], +h[ax] from the box centre.  p = line
 * point relative to the centre, d = direction, h = half sides.  `second`
 * marks the 2nd/3rd pass, where the + face is only tried while fewer than
 * two points were found (find_l_3d.cpp:114-173). */
static void face_pair(int ax, const float *p, const float *d, const float *h, int second,
                      float pts[2][3], int *k)
{
    int u = (ax == 0) ? 1 : 0;          /* the two other axes, in x<y<z order */
    int v = (ax == 2) ? 1 : 2;
    float su = d[u] / d[ax], sv = d[v] / d[ax];
    float mu = su * (-p[ax]) + p[u], wu = su * h[ax];
    float mv = sv * (-p[ax]) + p[v], wv = sv * h[ax];
    if (FABS(mu - wu) <= h[u] && FABS(mv - wv) <= h[v]) {
        pts[*k][ax] = -h[ax]; pts[*k][u] = mu - wu; pts[*k][v] = mv - wv; (*k)++;
    }
    if (!second || *k < 2) {
        if (FABS(mu + wu) <= h[u] && FABS(mv + wv) <= h[v]) {
            pts[*k][ax] = h[ax]; pts[*k][u] = mu + wu; pts[*k][v] = mv + wv; (*k)++;
        }
    }
}

/* Chord of the line (x1,y1,z1)->(x2,y2,z2) through the box centred at
 * (x,y,z) with sides (dx,dy,dz).  find_l_3d.cpp:94-182: a==0 or b==0 fall
 * back to the 2D routine (:104-108); otherwise faces are tried x,y,z
 * (|a|>|b|) or y,x,z, stopping once two points are found (:114-173).
 * Deviation: with a==0 (b==0) the line lies in the plane x = x1 (y = y1)
 * and is credited only to boxes whose x (y) slab holds it; the reference's
 * 2D fallback ignores that coordinate and credits boxes the line misses. */
float orc_find_l_3d(float x1_0, float y1_0, float z1_0, float x2_0, float y2_0, float z2_0,
                    float dx, float dy, float dz, float x, float y, float z)
{
    float a = x2_0 - x1_0, b = y2_0 - y1_0, c = z2_0 - z1_0;
    if (a == 0) return FABS(x1_0 - x) <= dx / 2 ? orc_find_l2(y1_0, z1_0, y2_0, z2_0, dy, dz, y, z) : 0.0f;
    if (b == 0) return FABS(y1_0 - y) <= dy / 2 ? orc_find_l2(x1_0, z1_0, x2_0, z2_0, dx, dz, x, z) : 0.0f;
    float p[3] = {x1_0 - x, y1_0 - y, z1_0 - z};
    float d[3] = {a, b, c};
    float h[3] = {dx / 2, dy / 2, dz / 2};
    float pts[2][3];
    int k = 0;
    int first = FABS(a) > FABS(b) ? 0 : 1;
    face_pair(first, p, d, h, 0, pts, &k);
    if (k < 2) face_pair(1 - first, p, d, h, 1, pts, &k);
    if (k < 2) face_pair(2, p, d, h, 1, pts, &k);
    if (k == 2) {
        float u = pts[1][0] - pts[0][0], v = pts[1][1] - pts[0][1], w = pts[1][2] - pts[0][2];
        return sqrtf(u * u + v * v + w * w);
    }
    return 0.0f;
}

/* ------------------------------------------------------------------------ */
/* Voxel-driven exact adjoints                                               */
/* ------------------------------------------------------------------------ */

/* 2D pixel-driven adjoint, Atx_fan_mf_cpu_new.cpp:21-63. */
void orc_atx_fan_f32(float *x, const float *Y, float SO, float OD, float scale, int nx, int ny,
                     int nd, float dy_det, float y_os, int nt, const int *id_Y, const int *Nv,
                     int tmp_size, const float *cos_phi, const float *sin_phi,
                     const float *cos_det, const float *sin_det)
{
    int nd2 = nd / 2, nx2 = nx / 2, ny2 = ny / 2, n = nx * ny;
    float SD = SO + OD;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; iy++) {
        float yc = (float)(iy + 0.5 - ny2);
        for (int ix = 0; ix < nx; ix++) {
            float xc = (float)(ix + 0.5 - nx2);
            for (int it = 0; it < nt; it++) {
                float acc = 0;
                for (int j = 0; j < Nv[it]; j++) {
                    int iv = id_Y[it * tmp_size + j];
                    const float *y = Y + (size_t)iv * nd;
                    float c = cos_phi[iv], s = sin_phi[iv];
                    float xr = c * xc + s * yc;
                    float yr = -s * xc + c * yc;
                    float t = SD / ((xr + SO) * dy_det);
                    int hi = (int)floorf((yr + 1) * t - y_os + (float)nd2);
                    int lo = (int)floorf((yr - 1) * t - y_os + (float)nd2);
                    if (lo < 0) lo = 0;
                    if (hi > nd - 1) hi = nd - 1;
                    for (int id = lo; id <= hi; id++) {
                        float Px = -(sin_det[id] * c + cos_det[id] * s);
                        float Py = cos_det[id] * c - sin_det[id] * s;
                        float Pd = SO * sin_det[id];
                        float l = orc_find_l(Px, Py, Pd, xc, yc);
                        if (l > 0) acc = acc + l * y[id];
                    }
                }
                x[(size_t)it * n + iy * nx + ix] = acc * scale;
            }
        }
    }
}

/* 3D voxel-driven adjoint with the semantics of the reference GPU kernel,
 * Atx_cone_mf_gpu_new.cu:41-84 (window from zr=zc-sd_z, d=sqrt((1+dz^2)/2),
 * ceil for the upper bounds, box sides (1,1,dz)).  Views of frame `it` are
 * id_Y[it*tmp_size+j], j<Nv[it]; Y is indexed by view. */
void orc_atx_cone_f32(float *x, const float *Y, const float *cos_phi, const float *sin_phi,
                      const float *sd_z, const float *y_det, const float *z_det,
                      const int *id_Y, const int *Nv, float SO, float OD, float scale,
                      float dy_det, float y_os, float dz_det, float dz, int nx, int ny, int nz,
                      int nt, int na, int nb, int tmp_size)
{
    float SD = SO + OD;
    int na2 = na / 2, nb2 = nb / 2, nx2 = nx / 2, ny2 = ny / 2, nz2 = nz / 2;
    float d = sqrtf((1 + dz * dz) / 2);
    size_t n = (size_t)nx * ny * nz, nd = (size_t)na * nb;
    int planes = ny * nz;
#pragma omp parallel for schedule(dynamic, 1)
    for (int pyz = 0; pyz < planes; pyz++) {
        int iz = pyz / ny, iy = pyz % ny;
        float zc = (float)(iz + 0.5 - nz2) * dz;
        float yc = (float)(iy + 0.5 - ny2);
        for (int ix = 0; ix < nx; ix++) {
            float xc = (float)(ix + 0.5 - nx2);
            size_t idx = ((size_t)iz * ny + iy) * nx + ix;
            for (int it = 0; it < nt; it++) {
                float acc = 0;
                for (int j = 0; j < Nv[it]; j++) {
                    int iv = id_Y[it * tmp_size + j];
                    const float *y = Y + (size_t)iv * nd;
                    float c = cos_phi[iv], s = sin_phi[iv];
                    float xr = c * xc + s * yc;
                    float yr = -s * xc + c * yc;
                    float zr = zc - sd_z[iv];
                    float t = SD / ((xr + SO) * dy_det);
                    int ahi = (int)ceilf((yr + 1) * t - y_os + (float)na2);
                    int alo = (int)floorf((yr - 1) * t - y_os + (float)na2);
                    t = SD / ((xr + SO) * dz_det);
                    int bhi = (int)ceilf((zr + d) * t + (float)nb2);
                    int blo = (int)floorf((zr - d) * t + (float)nb2);
                    if (alo < 0) alo = 0;
                    if (ahi > na - 1) ahi = na - 1;
                    if (blo < 0) blo = 0;
                    if (bhi > nb - 1) bhi = nb - 1;
                    float x1 = c * (-SO), y1 = s * (-SO), z1 = sd_z[iv];
                    for (int ib = blo; ib <= bhi; ib++) {
                        float z2 = z_det[ib] + sd_z[iv];
                        for (int ia = alo; ia <= ahi; ia++) {
                            float x2 = c * OD - s * y_det[ia];
                            float y2 = s * OD + c * y_det[ia];
                            float l = orc_find_l_3d(x1, y1, z1, x2, y2, z2, 1.0f, 1.0f, dz, xc, yc, zc);
                            acc = acc + l * y[(size_t)ib * na + ia];
                        }
                    }
                }
                x[(size_t)it * n + idx] = acc * scale;
            }
        }
    }
}

/* FP64 voxel-driven exact chord (slab clipping), used by the FP64 adjoint
 * oracle and by closed-form tests.  Returns the length of the part of the
 * line through p0,p1 inside the box [c-h, c+h]. */
double orc_chord_box_f64(const double *p0, const double *p1, const double *c, const double *h, int dim)
{
    double t0 = -INFINITY, t1 = INFINITY, len2 = 0.0;
    for (int k = 0; k < dim; k++) {
        double d = p1[k] - p0[k];
        len2 += d * d;
        double lo = c[k] - h[k] - p0[k], hi = c[k] + h[k] - p0[k];
        if (d == 0.0) {
            if (lo > 0.0 || hi < 0.0) return 0.0;
            continue;
        }
        double a = lo / d, b = hi / d;
        if (a > b) { double tmp = a; a = b; b = tmp; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
    }
    if (t1 <= t0) return 0.0;
    return (t1 - t0) * sqrt(len2);
}

/* ---------------------------------------------------------------------
 * Finite-width beam (fan, version 2): strip-area weights.
 *
 * A line a*X + b*Y = c with unit normal (a, b) cuts the unit pixel centred
 * (x, y).  With t = c - a*x - b*y, the area of the pixel on the far side of
 * the line from the centre is, for |t| < (|a|+|b|)/2 (the line meets the
 * pixel): |t|/max(|a|,|b|) while |t| <= ||a|-|b||/2 (the line crosses two
 * opposite edges; a trapezoid), else 1/2 - (dmax-|t|)^2/(2|a||b|) (a corner
 * triangle); 1/2 when the line misses.  The strip between two such lines
 * covers |A1 - s1*s2*A2| of the pixel, s_k the side of the centre
 * (find_area.cpp:18-69, float operations in the same order).
 */
static float half_cut(float a, float b, float c, float x, float y, int *sgn)
{
    float aa = a > 0 ? a : -a, bb = b > 0 ? b : -b;
    float hx = aa / 2, hy = bb / 2, dmax = hx + hy;
    float t = c - a * x - b * y;
    float d = t > 0 ? t : -t;
    *sgn = t > 0 ? 1 : -1;
    if (!(d < dmax)) return 0.5f;
    float g = hx - hy;
    float d0 = g > 0 ? g : -g;
    float inv = hx < hy ? 1.0f / bb : 1.0f / aa;
    if (d <= d0) return d * inv;
    return 0.5f - 0.5f * (dmax - d) * (dmax - d) / (aa * bb);
}

float orc_find_area(float a1, float b1, float c1, float a2, float b2, float c2, float x, float y)
{
    int s1, s2;
    float A1 = half_cut(a1, b1, c1, x, y, &s1);
    float A2 = half_cut(a2, b2, c2, x, y, &s2);
    float t = A1 - (float)(s1 * s2) * A2;
    return t > 0 ? t : -t;
}

/* Column range of a beam bounded by the source rays through two edge
 * points, over the unit column [u1, u1+1] of the major axis: the four edge
 * crossings floor(o + slope_k*(u - o_major) + n/2), kept in [0, n-1];
 * (lo, hi) start at (n-1, 0) (Ax_fan_mf_cpu_new_fb.cpp:59-75). */
static void beam_range(float o, float om, const float *slope, float u1, int n, int *lo, int *hi)
{
    int n2 = n / 2;
    float u2 = u1 + 1;
    int v[4] = {(int)floorf(o + slope[0] * (u1 - om) + (float)n2), (int)floorf(o + slope[0] * (u2 - om) + (float)n2),
                (int)floorf(o + slope[1] * (u1 - om) + (float)n2), (int)floorf(o + slope[1] * (u2 - om) + (float)n2)};
    *lo = n - 1;
    *hi = 0;
    for (int i = 0; i < 4; i++) {
        if (v[i] < *lo && v[i] >= 0) *lo = v[i];
        if (v[i] > *hi && v[i] <= n - 1) *hi = v[i];
    }
}

/* Forward, finite-width beam (Ax_fan_mf_cpu_new_fb.cpp:19-123): cell id of
 * view iv is the strip between the source rays to its edges
 * y_det[id] -+ dy_det/2; the walk runs along the major axis of the strip's
 * centre line and sums area*pixel for pixels of the per-column range with a
 * positive area.  Trig in float (cosf, sinf, atan2f of float arguments).
 * When an edge point shares the source's major coordinate, the reference
 * takes the whole column if the source lies within half a pixel of it and
 * otherwise keeps the previous column's range; the range starts empty for
 * each ray here. */
void orc_fan_fb_fwd_f32(const float *X, float *y, float SO, float OD, float scale, int nx, int ny,
                        int nv, const float *sd_phi, int nd, const float *y_det, const int *id_X,
                        float dy_det)
{
    int nx2 = nx / 2, ny2 = ny / 2, n = nx * ny;
    float SD = SO + OD;
#pragma omp parallel for schedule(static)
    for (int iv = 0; iv < nv; iv++) {
        const float *x = X + (size_t)id_X[iv] * n;
        float c = cosf(sd_phi[iv]), s = sinf(sd_phi[iv]);
        float x1 = c * (-SO), y1 = s * (-SO);
        for (int id = 0; id < nd; id++) {
            float ex[2], ey[2], Px[2], Py[2], Pd[2], slope[2];
            for (int e = 0; e < 2; e++) {
                float u = e == 0 ? y_det[id] - dy_det / 2 : y_det[id] + dy_det / 2;
                ex[e] = c * OD - s * u;
                ey[e] = s * OD + c * u;
                float ang = atan2f(u, SD);
                Px[e] = -sinf(ang + sd_phi[iv]);
                Py[e] = cosf(ang + sd_phi[iv]);
                Pd[e] = SO * sinf(ang);
            }
            float acc = 0;
            int lo = 0, hi = -1;
            float mx = (ex[0] + ex[1]) / 2, my = (ey[0] + ey[1]) / 2;
            float gx = x1 - mx, gy = y1 - my;
            if ((gx > 0 ? gx : -gx) > (gy > 0 ? gy : -gy)) {
                int flat = ex[0] == x1 || ex[1] == x1;
                if (!flat) { slope[0] = (ey[0] - y1) / (ex[0] - x1); slope[1] = (ey[1] - y1) / (ex[1] - x1); }
                for (int ix = 0; ix < nx; ix++) {
                    float xc = (float)(ix + 0.5 - nx2);
                    if (flat) {
                        float g = x1 - xc;
                        if ((g > 0 ? g : -g) <= 0.5f) { lo = 0; hi = ny - 1; }
                    } else {
                        beam_range(y1, x1, slope, (float)(ix - nx2), ny, &lo, &hi);
                    }
                    for (int iy = lo; iy <= hi; iy++) {
                        float yc = (float)(iy + 0.5 - ny2);
                        float l = orc_find_area(Px[0], Py[0], Pd[0], Px[1], Py[1], Pd[1], xc, yc);
                        if (l > 0) acc = acc + l * x[iy * nx + ix];
                    }
                }
            } else {
                int flat = ey[0] == y1 || ey[1] == y1;
                if (!flat) { slope[0] = (ex[0] - x1) / (ey[0] - y1); slope[1] = (ex[1] - x1) / (ey[1] - y1); }
                for (int iy = 0; iy < ny; iy++) {
                    float yc = (float)(iy + 0.5 - ny2);
                    if (flat) {
                        float g = y1 - yc;
                        if ((g > 0 ? g : -g) <= 0.5f) { lo = 0; hi = nx - 1; }
                    } else {
                        beam_range(x1, y1, slope, (float)(iy - ny2), nx, &lo, &hi);
                    }
                    for (int ix = lo; ix <= hi; ix++) {
                        float xc = (float)(ix + 0.5 - nx2);
                        float l = orc_find_area(Px[0], Py[0], Pd[0], Px[1], Py[1], Pd[1], xc, yc);
                        if (l > 0) acc = acc + l * x[iy * nx + ix];
                    }
                }
            }
            y[(size_t)iv * nd + id] = acc * scale;
        }
    }
}

/* Adjoint, finite-width beam (Atx_fan_mf_cpu_new_fb.cpp:21-66): per pixel
 * and view, cells floor((yr-+1)*t + nd/2) (no y_os term), strip edges from
 * the nd+1 edge-angle tables cos_det2/sin_det2 (edge id and id+1 bound cell
 * id), area-weighted sum of positive areas, times scale. */
void orc_fan_fb_adj_f32(float *x, const float *Y, float SO, float OD, float scale, int nx, int ny,
                        int nd, float dy_det, int nt, const int *id_Y, const int *Nv, int tmp_size,
                        const float *cos_phi, const float *sin_phi, const float *cos_det2,
                        const float *sin_det2)
{
    int nd2 = nd / 2, nx2 = nx / 2, ny2 = ny / 2, n = nx * ny;
    float SD = SO + OD;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < ny; iy++) {
        float yc = (float)(iy + 0.5 - ny2);
        for (int ix = 0; ix < nx; ix++) {
            float xc = (float)(ix + 0.5 - nx2);
            for (int it = 0; it < nt; it++) {
                float acc = 0;
                for (int j = 0; j < Nv[it]; j++) {
                    int iv = id_Y[it * tmp_size + j];
                    const float *yv = Y + (size_t)iv * nd;
                    float c = cos_phi[iv], s = sin_phi[iv];
                    float xr = c * xc + s * yc;
                    float yr = -s * xc + c * yc;
                    float t = SD / ((xr + SO) * dy_det);
                    int hi = (int)floorf((yr + 1) * t + (float)nd2);
                    int lo = (int)floorf((yr - 1) * t + (float)nd2);
                    if (lo < 0) lo = 0;
                    if (hi > nd - 1) hi = nd - 1;
                    for (int id = lo; id <= hi; id++) {
                        float Px0 = -(sin_det2[id] * c + cos_det2[id] * s);
                        float Py0 = cos_det2[id] * c - sin_det2[id] * s;
                        float Pd0 = SO * sin_det2[id];
                        float Px1 = -(sin_det2[id + 1] * c + cos_det2[id + 1] * s);
                        float Py1 = cos_det2[id + 1] * c - sin_det2[id + 1] * s;
                        float Pd1 = SO * sin_det2[id + 1];
                        float l = orc_find_area(Px0, Py0, Pd0, Px1, Py1, Pd1, xc, yc);
                        if (l > 0) acc = acc + l * yv[id];
                    }
                }
                x[(size_t)it * n + iy * nx + ix] = acc * scale;
            }
        }
    }
}

/* Thread count of the OpenMP loops above (benchmarks time 1 and all cores). */
#include <omp.h>
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
