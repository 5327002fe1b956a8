/*
 * ch_ref.c -- FP64 CPU oracle for the collapsed-cone dose path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it as the checker and as the timed CPU baseline.  The
 * product (conehead_amd/) never links or calls it.
 *
 * Built with -O2 -ffp-contract=off and no -march flag, so every
 * multiply-add below rounds twice, exactly as the reference's Cython code
 * (compiled for baseline x86-64, which has no FMA) does.
 *
 * Pinning (see tests/test_oracle_*.py):
 *  - ch_ref_dda_3d: the reference notebook goldens (rays A and C of
 *    dda_benchmarking.ipynb), brute-force boundary-crossing merges on small
 *    grids, and axis-aligned closed forms.
 *  - ch_ref_convolve*: impulse response == tap table, telescoping sum for a
 *    single cone, linearity, x n_theta scaling, 1-D axis-aligned closed form,
 *    and scatter == gather identity.
 *  - ch_ref_terma: closed forms (rho = 1 radiological depth equals the
 *    geometric exit distance, inverse-square fluence on the central axis,
 *    blocked fraction of a square field in multiples of 1/samples^3).
 *  - ch_ref_convolve_rad (physics mode, no reference counterpart): an
 *    impulse against a forward walk of ch_ref_dda_3d with the cuda_dose
 *    index, rho = 1 against the geometric mode, rho = 2 against doubled
 *    voxels, and a 1-D slab closed form.
 */
#include "ch_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CH_REF_BIG 1000000000.0
#define CH_REF_MAX_INDEX 5995

/* conehead/dda_3d_c.pyx:33-99 */
int32_t ch_ref_dda_3d(const double direction[3], const int32_t grid_shape[3],
                      const int32_t first_voxel[3], const double voxel_size[3],
                      int32_t* vox_out, double* t_out, int32_t cap,
                      double* dir_abs_out)
{
    int32_t step[3], cur[3];
    double d[3], t[3], dt[3];
    int32_t n = 0;
    for (int a = 0; a < 3; ++a) {
        /* :33-36 step sign; -0.0 and +0.0 both give +1 */
        step[a] = direction[a] < 0 ? -1 : 1;
        /* :38-43 abs in place */
        d[a] = direction[a] < 0 ? -1 * direction[a] : direction[a];
        cur[a] = first_voxel[a];
    }
    if (dir_abs_out) {
        dir_abs_out[0] = d[0];
        dir_abs_out[1] = d[1];
        dir_abs_out[2] = d[2];
    }
    /* :52-70 */
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            t[a] = CH_REF_BIG;
            dt[a] = CH_REF_BIG;
        } else {
            t[a] = (voxel_size[a] / 2) / d[a];
            dt[a] = voxel_size[a] / d[a];
        }
    }
    /* :72-99 */
    while (cur[0] >= 0 && cur[0] < grid_shape[0] &&
           cur[1] >= 0 && cur[1] < grid_shape[1] &&
           cur[2] >= 0 && cur[2] < grid_shape[2]) {
        int a;
        if (t[0] < t[1]) {
            a = (t[0] < t[2]) ? 0 : 2;
        } else {
            a = (t[1] < t[2]) ? 1 : 2;
        }
        if (n < cap) {
            vox_out[3 * n + 0] = cur[0];
            vox_out[3 * n + 1] = cur[1];
            vox_out[3 * n + 2] = cur[2];
            t_out[n] = t[a];
        }
        t[a] += dt[a];
        cur[a] = cur[a] + step[a];
        ++n;
    }
    return n;
}

/* conehead/convolve_c.pyx:95-117 */
void ch_ref_cone_direction(double theta_deg, double phi_deg, double out[3])
{
    const double pi = 3.14159265359;                 /* :48 */
    double theta_rad = theta_deg * pi / 180.0;
    double phi_rad = phi_deg * pi / 180.0;
    double c_t = cos(theta_rad);
    double s_t = sin(theta_rad);
    double c_p = cos(phi_rad);
    double s_p = sin(phi_rad);
    double d0 = c_t * s_p, d1 = s_t * s_p, d2 = c_p; /* :101 */
    double N = d0 * d0 + d1 * d1 + d2 * d2;          /* :102-106 */
    double N_inv = 1.0 / N;                          /* :107 */
    d0 = d0 * N_inv;                                 /* :108-112 */
    d1 = d1 * N_inv;
    d2 = d2 * N_inv;
    out[0] = ceil(d0 * 100000) * 0.00001;            /* :113-117 */
    out[1] = ceil(d1 * 100000) * 0.00001;
    out[2] = ceil(d2 * 100000) * 0.00001;
}

/* conehead/convolve_c.pyx:138-145.  The cast of a double beyond the int32
 * range is undefined in C; such t-values (> 2e7 cm) are treated as skipped. */
int32_t ch_ref_kernel_index(double t)
{
    double f = floor(t * 100.0 - 50.0);
    if (!(f < 2147483647.0 && f > -2147483648.0))
        return -1;
    int32_t idx = (int32_t)f;
    idx = abs(idx);
    if (idx > CH_REF_MAX_INDEX)
        return -1;
    return idx;
}

static int64_t convolve_range(const double* terma, double* dose,
                              const int32_t shape[3], const double dims[3],
                              int32_t n_cones, const double* dirs,
                              const int32_t* cone_row, int32_t row_len,
                              const double* cum, int32_t x0, int32_t x1,
                              int32_t* vox, double* tv, int32_t* idxs,
                              int32_t cap)
{
    const int64_t sy = shape[2], sx = (int64_t)shape[1] * shape[2];
    int64_t updates = 0;
    for (int32_t x = x0; x < x1; ++x)
        for (int32_t y = 0; y < shape[1]; ++y)
            for (int32_t z = 0; z < shape[2]; ++z) {
                double T = terma[x * sx + y * sy + z];  /* :84 */
                if (T == 0.0)                           /* :85 `if T:` */
                    continue;
                for (int32_t c = 0; c < n_cones; ++c) {
                    int32_t first[3] = {x, y, z};
                    int32_t n = ch_ref_dda_3d(dirs + 3 * c, shape, first,
                                              dims, vox, tv, cap, NULL);
                    if (n > cap)
                        return -2;
                    /* :133-150: indices, skipping those > 5995 */
                    int32_t ni = 0;
                    for (int32_t m = 0; m < n; ++m) {
                        int32_t idx = ch_ref_kernel_index(tv[m]);
                        if (idx < 0)
                            continue;
                        idxs[ni++] = idx;
                    }
                    const double* row = cum + (int64_t)cone_row[c] * row_len;
                    /* :157-192 */
                    for (int32_t m = 0; m < ni; ++m) {
                        const int32_t* v = vox + 3 * m;
                        double k3;
                        if (m == 0)
                            k3 = row[idxs[0]];
                        else
                            k3 = row[idxs[m]] - row[idxs[m - 1]];
                        dose[v[0] * sx + v[1] * sy + v[2]] += T * k3;
                    }
                    updates += ni;
                }
            }
    return updates;
}

int64_t ch_ref_convolve_dirs(const double* terma, double* dose,
                             const int32_t shape[3], const double dims[3],
                             int32_t n_cones, const double* dirs,
                             const int32_t* cone_row, int32_t row_len,
                             const double* cum, int32_t n_threads)
{
    if (row_len < CH_REF_MAX_INDEX + 1 || shape[0] <= 0 || shape[1] <= 0 ||
        shape[2] <= 0 || n_cones < 0)
        return -1;
    const int32_t cap = shape[0] + shape[1] + shape[2] + 1;
    const int64_t nv = (int64_t)shape[0] * shape[1] * shape[2];
    if (n_threads <= 1) {
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * 3 * cap);
        double* tv = (double*)malloc(sizeof(double) * cap);
        int32_t* idxs = (int32_t*)malloc(sizeof(int32_t) * cap);
        int64_t r = convolve_range(terma, dose, shape, dims, n_cones, dirs,
                                   cone_row, row_len, cum, 0, shape[0], vox,
                                   tv, idxs, cap);
        free(vox);
        free(tv);
        free(idxs);
        return r;
    }
    /* Multi-threaded CPU baseline: private dose buffers, fixed-order sum. */
    double** priv = (double**)calloc((size_t)n_threads, sizeof(double*));
    int64_t* counts = (int64_t*)calloc((size_t)n_threads, sizeof(int64_t));
    int err = 0;
#pragma omp parallel num_threads(n_threads)
    {
        int tid = 0, nt = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        double* buf = (double*)calloc((size_t)nv, sizeof(double));
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * 3 * cap);
        double* tv = (double*)malloc(sizeof(double) * cap);
        int32_t* idxs = (int32_t*)malloc(sizeof(int32_t) * cap);
        priv[tid] = buf;
        int64_t cnt = 0;
        for (int32_t x = tid; x < shape[0]; x += nt) {
            int64_t r = convolve_range(terma, buf, shape, dims, n_cones, dirs,
                                       cone_row, row_len, cum, x, x + 1, vox,
                                       tv, idxs, cap);
            if (r < 0) {
#pragma omp atomic write
                err = 1;
                break;
            }
            cnt += r;
        }
        counts[tid] = cnt;
        free(vox);
        free(tv);
        free(idxs);
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nv; ++i) {
            double s = dose[i];
            for (int th = 0; th < nt; ++th)
                if (priv[th])
                    s += priv[th][i];
            dose[i] = s;
        }
    }
    int64_t total = 0;
    for (int th = 0; th < n_threads; ++th) {
        total += counts[th];
        free(priv[th]);
    }
    free(priv);
    free(counts);
    return err ? -2 : total;
}

int64_t ch_ref_convolve(const double* terma, double* dose,
                        const int32_t shape[3], const double dims[3],
                        const double* thetas, int32_t n_theta,
                        const double* phis, int32_t n_phi, int32_t row_len,
                        const double* cum, int32_t n_threads)
{
    int32_t n = n_theta * n_phi;
    double* dirs = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    for (int32_t i = 0; i < n_theta; ++i)        /* :88-89 theta, then phi */
        for (int32_t j = 0; j < n_phi; ++j) {
            ch_ref_cone_direction(thetas[i], phis[j],
                                  dirs + 3 * (i * n_phi + j));
            rows[i * n_phi + j] = j;
        }
    int64_t r = ch_ref_convolve_dirs(terma, dose, shape, dims, n, dirs, rows,
                                     row_len, cum, n_threads);
    free(dirs);
    free(rows);
    return r;
}

/* Tap table of one cone: the DDA from the origin voxel of an unbounded grid,
 * kept while every offset component can still land inside `shape` and the
 * kernel index is <= 5995.  Same arithmetic as ch_ref_dda_3d. */
int32_t ch_ref_cone_taps(const double dir[3], const int32_t shape[3],
                         const double dims[3], int32_t row_len,
                         const double* cum_row, int32_t* off_out,
                         double* k_out, int32_t cap)
{
    (void)row_len;
    int32_t step[3], cur[3] = {0, 0, 0};
    double d[3], t[3], dt[3];
    for (int a = 0; a < 3; ++a) {
        step[a] = dir[a] < 0 ? -1 : 1;
        d[a] = dir[a] < 0 ? -1 * dir[a] : dir[a];
        if (d[a] == 0.0) {
            t[a] = CH_REF_BIG;
            dt[a] = CH_REF_BIG;
        } else {
            t[a] = (dims[a] / 2) / d[a];
            dt[a] = dims[a] / d[a];
        }
    }
    int32_t n = 0, prev = -1;
    while (abs(cur[0]) < shape[0] && abs(cur[1]) < shape[1] &&
           abs(cur[2]) < shape[2]) {
        int a;
        if (t[0] < t[1])
            a = (t[0] < t[2]) ? 0 : 2;
        else
            a = (t[1] < t[2]) ? 1 : 2;
        int32_t idx = ch_ref_kernel_index(t[a]);
        if (idx < 0)
            break;
        if (n >= cap)
            return -1;
        off_out[3 * n + 0] = cur[0];
        off_out[3 * n + 1] = cur[1];
        off_out[3 * n + 2] = cur[2];
        k_out[n] = (prev < 0) ? cum_row[idx] : cum_row[idx] - cum_row[prev];
        prev = idx;
        ++n;
        t[a] += dt[a];
        cur[a] = cur[a] + step[a];
    }
    return n;
}

int64_t ch_ref_convolve_gather(const double* terma, double* dose,
                               const int32_t shape[3], const double dims[3],
                               int32_t n_cones, const double* dirs,
                               const int32_t* cone_row, int32_t row_len,
                               const double* cum)
{
    if (row_len < CH_REF_MAX_INDEX + 1)
        return -1;
    const int32_t cap = shape[0] + shape[1] + shape[2] + 1;
    const int64_t sy = shape[2], sx = (int64_t)shape[1] * shape[2];
    int32_t* off = (int32_t*)malloc(sizeof(int32_t) * 3 * cap);
    double* k = (double*)malloc(sizeof(double) * cap);
    int64_t updates = 0;
    for (int32_t c = 0; c < n_cones; ++c) {
        int32_t m_taps = ch_ref_cone_taps(dirs + 3 * c, shape, dims, row_len,
                                          cum + (int64_t)cone_row[c] * row_len,
                                          off, k, cap);
        if (m_taps < 0) {
            free(off);
            free(k);
            return -2;
        }
        for (int32_t x = 0; x < shape[0]; ++x)
            for (int32_t y = 0; y < shape[1]; ++y)
                for (int32_t z = 0; z < shape[2]; ++z) {
                    double acc = 0.0;
                    for (int32_t m = 0; m < m_taps; ++m) {
                        int32_t sxv = x - off[3 * m], syv = y - off[3 * m + 1],
                                szv = z - off[3 * m + 2];
                        if (sxv < 0 || sxv >= shape[0] || syv < 0 ||
                            syv >= shape[1] || szv < 0 || szv >= shape[2])
                            continue;
                        double T = terma[sxv * sx + syv * sy + szv];
                        if (T != 0.0) {
                            acc += T * k[m];
                            ++updates;
                        }
                    }
                    dose[x * sx + y * sy + z] += acc;
                }
    }
    free(off);
    free(k);
    return updates;
}

/* ---------------------------------------------------------------- TERMA */

static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* conehead/conehead.py:25-47 (plane point = isocentre at the origin). */
static int plane_hit(const double* start, const double* u, const double* n,
                     double* out)
{
    double ndotu = dot3(n, u);
    if (fabs(ndotu) < 1e-6)
        return -1;
    double si = -dot3(n, start) / ndotu;
    out[0] = start[0] + si * u[0] + 0.0;
    out[1] = start[1] + si * u[1] + 0.0;
    out[2] = start[2] + si * u[2] + 0.0;
    return 0;
}

/* conehead/conehead.py:50-68 on the 4000x4000 raster (negative index -1
 * wraps to the last row/column, as Python/Numba indexing does). */
static double block_transmission(double px, double pz, const float* raster)
{
    px = floor(px * 100.0) + 2000.0;
    pz = floor(pz * 100.0) + 2000.0;
    if (px < 0 || px > 3999 || pz < 0 || pz > 3999)
        return 0.0;
    int32_t i = (int32_t)px - 1, j = (int32_t)pz - 1;
    if (i < 0)
        i += 4000;
    if (j < 0)
        j += 4000;
    return (double)raster[(int64_t)i * 4000 + j];
}

int32_t ch_ref_terma(const ch_ref_beam* b, const int32_t shape[3],
                     const double origin[3], const double spacing[3],
                     const float* density, const float* raster,
                     double* terma, double* blocked_out, double* d_eff_out,
                     double* oad_out, double* fluence_out, int32_t n_threads)
{
    const int64_t sy = shape[2], sx = (int64_t)shape[1] * shape[2];
    const int32_t cap = shape[0] + shape[1] + shape[2] + 1;
    const int ns = b->samples;
    int err = 0;
    if (n_threads < 1)
        n_threads = 1;
#pragma omp parallel num_threads(n_threads)
    {
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * 3 * cap);
        double* tv = (double*)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 1)
        for (int32_t x = 0; x < shape[0]; ++x)
            for (int32_t y = 0; y < shape[1]; ++y)
                for (int32_t z = 0; z < shape[2]; ++z) {
                    const int64_t li = x * sx + y * sy + z;
                    int32_t idx3[3] = {x, y, z};
                    double pos[3];
                    for (int a = 0; a < 3; ++a)
                        pos[a] = origin[a] + spacing[a] * idx3[a];

                    /* hit test, :78-120 */
                    double off[3];
                    for (int a = 0; a < 3; ++a)
                        off[a] = spacing[a] / ns;
                    double blocked = 0.0;
                    const double n3 = (double)(ns * ns * ns);
                    for (int ix = 0; ix < ns; ++ix)
                        for (int iy = 0; iy < ns; ++iy)
                            for (int iz = 0; iz < ns; ++iz) {
                                int ii[3] = {ix, iy, iz};
                                double ps[3], u[3], pp[3];
                                for (int a = 0; a < 3; ++a) {
                                    ps[a] = pos[a] - spacing[a] / 2 +
                                            off[a] / 2 + off[a] * ii[a];
                                    u[a] = b->source_position[a] - ps[a];
                                }
                                if (plane_hit(b->source_position, u, b->v_y,
                                              pp)) {
#pragma omp atomic write
                                    err = 1;
                                    continue;
                                }
                                double q0 = dot3(b->transform + 0, pp);
                                double q2 = dot3(b->transform + 6, pp);
                                blocked = blocked +
                                          block_transmission(q0, q2, raster) /
                                              n3;
                            }

                    /* effective depth, :209-254 */
                    double r[3];
                    for (int a = 0; a < 3; ++a)
                        r[a] = b->source_position[a] - pos[a];
                    double mag = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                    double dir[3] = {r[0] / mag, r[1] / mag, r[2] / mag};
                    int32_t n = ch_ref_dda_3d(dir, shape, idx3, spacing, vox,
                                              tv, cap, NULL);
                    double deff = 0.0;
                    for (int32_t m = 0; m < n && m < cap; ++m) {
                        double rho = (double)density[vox[3 * m] * sx +
                                                     vox[3 * m + 1] * sy +
                                                     vox[3 * m + 2]];
                        deff += rho * (m == 0 ? tv[0] : tv[m] - tv[m - 1]);
                    }

                    /* off-axis distance, :286-314 */
                    double pp[3], oad = 0.0;
                    if (plane_hit(b->source_position, r, b->v_y, pp)) {
#pragma omp atomic write
                        err = 1;
                    } else {
                        double q0 = dot3(b->transform + 0, pp);
                        double q2 = dot3(b->transform + 6, pp);
                        oad = sqrt(q0 * q0 + q2 * q2);
                    }

                    /* fluence, :320-360 */
                    double sad = b->sad;
                    double fp = b->sPri * ((sad / mag) * (sad / mag));
                    double r_ann = oad * b->zAnn / sad;
                    double fa = 0.0;
                    if (r_ann >= b->rInner && r_ann <= b->rOuter)
                        fa = b->sAnn * ((sad - b->zAnn) * (sad - b->zAnn)) /
                             ((mag - b->zAnn) * (mag - b->zAnn));
                    double oad_c = oad < 2.0 ? 2.0 : oad;
                    double r_exp = oad_c * b->zExp / sad;
                    double fe = b->sExp / r_exp * exp(-b->kExp * r_exp) *
                                ((sad - b->zExp) * (sad - b->zExp)) /
                                ((mag - b->zExp) * (mag - b->zExp));
                    double flu = (fp + fa + fe) * blocked;

                    /* softening and horn, :641-655 */
                    double fs = oad < b->softLimit
                                    ? 1.0 / (1.0 - b->softRatio * oad)
                                    : 1.0;
                    double fh = 1.0 + oad * b->hornRatio;

                    /* TERMA, :364-377 */
                    double T = 0.0;
                    for (int i = 0; i < b->n_energy; ++i)
                        T += b->weight[i] * flu *
                             exp(-b->mu[i] * fs * deff) * b->energy[i] *
                             b->mu[i] * fh;
                    T = T * blocked;

                    terma[li] = T;
                    if (blocked_out)
                        blocked_out[li] = blocked;
                    if (d_eff_out)
                        d_eff_out[li] = deff;
                    if (oad_out)
                        oad_out[li] = oad;
                    if (fluence_out)
                        fluence_out[li] = flu;
                }
        free(vox);
        free(tv);
    }
    return err ? -1 : 0;
}

/* ------------------------------------------------------- physics mode */

/* Kernel index of cuda_dose (conehead/conehead.py:453-463): the radius r
 * looks up |floor(100 r - 5)|, clamped at 5995 (no truncation). */
int32_t ch_ref_kernel_index_cuda(double r)
{
    double f = floor(r * 100.0 - 5.0);
    if (!(f < 2147483647.0 && f > -2147483648.0))
        return CH_REF_MAX_INDEX;
    int32_t idx = abs((int32_t)f);
    return idx > CH_REF_MAX_INDEX ? CH_REF_MAX_INDEX : idx;
}

/* The DDA path of one cone from the origin voxel of an unbounded grid, kept
 * while every offset component can still land inside `shape`: offsets and
 * the t-values at which each voxel is left (same arithmetic as
 * ch_ref_dda_3d).  Returns the number of voxels, or -1 on overflow. */
int32_t ch_ref_cone_path(const double dir[3], const int32_t shape[3],
                         const double dims[3], int32_t* off_out,
                         double* t_out, int32_t cap)
{
    int32_t step[3], cur[3] = {0, 0, 0};
    double d[3], t[3], dt[3];
    for (int a = 0; a < 3; ++a) {
        step[a] = dir[a] < 0 ? -1 : 1;
        d[a] = dir[a] < 0 ? -1 * dir[a] : dir[a];
        if (d[a] == 0.0) {
            t[a] = CH_REF_BIG;
            dt[a] = CH_REF_BIG;
        } else {
            t[a] = (dims[a] / 2) / d[a];
            dt[a] = dims[a] / d[a];
        }
    }
    int32_t n = 0;
    while (abs(cur[0]) < shape[0] && abs(cur[1]) < shape[1] &&
           abs(cur[2]) < shape[2]) {
        int a;
        if (t[0] < t[1])
            a = (t[0] < t[2]) ? 0 : 2;
        else
            a = (t[1] < t[2]) ? 1 : 2;
        if (n >= cap)
            return -1;
        off_out[3 * n + 0] = cur[0];
        off_out[3 * n + 1] = cur[1];
        off_out[3 * n + 2] = cur[2];
        t_out[n] = t[a];
        ++n;
        t[a] += dt[a];
        cur[a] = cur[a] + step[a];
    }
    return n;
}

static void rad_range(const double* terma, const float* density, double* dose,
                      const int32_t shape[3], int32_t n_cones,
                      const int32_t* path_n, const int32_t* path_off,
                      const double* path_t, int32_t cap,
                      const int32_t* cone_row, int32_t row_len,
                      const double* cum, int32_t radiological, int32_t x0,
                      int32_t x1, int64_t* updates)
{
    const int64_t sy = shape[2], sx = (int64_t)shape[1] * shape[2];
    for (int32_t x = x0; x < x1; ++x)
        for (int32_t y = 0; y < shape[1]; ++y)
            for (int32_t z = 0; z < shape[2]; ++z) {
                double acc = 0.0;
                for (int32_t c = 0; c < n_cones; ++c) {
                    const int32_t* off = path_off + (int64_t)3 * cap * c;
                    const double* tv = path_t + (int64_t)cap * c;
                    const double* row = cum + (int64_t)cone_row[c] * row_len;
                    double r = 0.0, t_prev = 0.0, c_prev = 0.0;
                    for (int32_t m = 0; m < path_n[c]; ++m) {
                        /* the voxel m steps back from v along the cone */
                        const int32_t px = x - off[3 * m],
                                      py = y - off[3 * m + 1],
                                      pz = z - off[3 * m + 2];
                        if (px < 0 || px >= shape[0] || py < 0 ||
                            py >= shape[1] || pz < 0 || pz >= shape[2])
                            break;  /* the path only leaves the grid */
                        const int64_t p = px * sx + py * sy + pz;
                        if (radiological) {
                            const double seg = tv[m] - t_prev;
                            r = r + (double)density[p] * seg;
                            t_prev = tv[m];
                        } else {
                            r = tv[m];
                        }
                        const double cv = row[ch_ref_kernel_index_cuda(r)];
                        const double k = cv - c_prev;
                        c_prev = cv;
                        const double T = terma[p];
                        if (T != 0.0) {
                            acc += T * k;
                            ++*updates;
                        }
                    }
                }
                dose[x * sx + y * sy + z] += acc;
            }
}

int64_t ch_ref_convolve_rad(const double* terma, const float* density,
                            double* dose, const int32_t shape[3],
                            const double dims[3], int32_t n_cones,
                            const double* dirs, const int32_t* cone_row,
                            int32_t row_len, const double* cum,
                            int32_t radiological, int32_t n_threads)
{
    if (row_len < CH_REF_MAX_INDEX + 1 || shape[0] <= 0 || shape[1] <= 0 ||
        shape[2] <= 0 || n_cones < 0 || (radiological && !density))
        return -1;
    const int32_t cap = shape[0] + shape[1] + shape[2] + 1;
    int32_t* path_n = (int32_t*)malloc(sizeof(int32_t) * (n_cones + 1));
    int32_t* path_off =
        (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)cap * (n_cones + 1));
    double* path_t =
        (double*)malloc(sizeof(double) * (size_t)cap * (n_cones + 1));
    for (int32_t c = 0; c < n_cones; ++c) {
        path_n[c] = ch_ref_cone_path(dirs + 3 * c, shape, dims,
                                     path_off + (int64_t)3 * cap * c,
                                     path_t + (int64_t)cap * c, cap);
        if (path_n[c] < 0) {
            free(path_n);
            free(path_off);
            free(path_t);
            return -2;
        }
    }
    int64_t total = 0;
    if (n_threads < 1)
        n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic) \
    reduction(+ : total)
    for (int32_t x = 0; x < shape[0]; ++x) {
        int64_t u = 0;
        rad_range(terma, density, dose, shape, n_cones, path_n, path_off,
                  path_t, cap, cone_row, row_len, cum, radiological, x, x + 1,
                  &u);
        total += u;
    }
    free(path_n);
    free(path_off);
    free(path_t);
    return total;
}
