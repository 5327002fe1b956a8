// plan.hip -- per-cone tap tables of the collapsed-cone gather, built on the
// device in FP64.
//
// convolve_c (conehead/convolve_c.pyx:88-192) runs, for every source voxel s
// and cone c, the DDA from the centre of s and adds T(s) * k_m to the m-th
// visited voxel.  The DDA offsets o_m = v_m - s, the t-values and hence the
// kernel indices and increments depend only on (direction, voxel size), and
// the path is monotone per axis, so voxel s + o_m is visited exactly when it
// lies in the grid.  The convolution is therefore the translation-invariant
// stencil dose(v) = sum_c sum_m k_{c,m} T(v - o_{c,m}); this file builds the
// (o_{c,m}, k_{c,m}) lists once per (grid, voxel size, cone set, kernel).
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dda_core.h"

namespace ch {

// One lane per cone: walk the DDA from the origin of an unbounded grid while
// every offset component could still land inside the grid, stop at the first
// skipped kernel index (t >= 60.46 cm, convolve_c) or after the first
// clamped one (cuda_dose), record taps and first-step tables.  Radiological
// plans keep the whole path and store each voxel's path length instead of
// an increment.
__global__ __launch_bounds__(64) void build_taps_kernel(
    int32_t n_cones, const double* __restrict__ dirs,
    const int32_t* __restrict__ rows, const double* __restrict__ cum,
    int32_t row_len, int32_t X, int32_t Y, int32_t Z, double v0, double v1,
    double v2, int32_t semantics, int32_t tap_stride, int32_t f_stride,
    int32_t* __restrict__ lin, float* __restrict__ k32,
    double* __restrict__ k64, int32_t* __restrict__ off,
    int32_t* __restrict__ ftab, ConeMeta* __restrict__ meta)
{
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cones)
        return;
    Dda d;
    d.init(dirs[3 * c + 0], dirs[3 * c + 1], dirs[3 * c + 2], v0, v1, v2);
    const double* row = cum + (int64_t)rows[c] * row_len;
    const int32_t tb = c * tap_stride;
    const int32_t fb0 = c * f_stride;
    const int32_t fb1 = fb0 + X + 2;
    const int32_t fb2 = fb1 + Y + 2;
    int32_t c0 = 0, c1 = 0, c2 = 0;  // current offset
    int32_t n0 = 0, n1 = 0, n2 = 0;  // steps taken per axis (= |offset|)
    ftab[fb0] = 0;
    ftab[fb1] = 0;
    ftab[fb2] = 0;
    int32_t n = 0, prev = -1;
    double t_prev = 0.0;
    // strides of the padded TERMA copy the gather reads
    const int32_t sy = Z + 2 * kZPad, sx = (Y + 2 * kYPad) * sy;
    while (n0 < X && n1 < Y && n2 < Z && n < tap_stride - kTapPad) {
        const int a = d.axis();
        double t = a == 0 ? d.t0 : (a == 1 ? d.t1 : d.t2);
        bool clamped = false;
        double k;
        if (semantics == CH_SEM_RADIOLOGICAL) {
            // the path length inside the m-th voxel; the increments depend
            // on the densities and are formed by ch_convolve_rad
            k = t - t_prev;
            t_prev = t;
        } else {
            const int32_t idx = kernel_index(t, semantics, &clamped);
            if (idx < 0)
                break;
            k = prev < 0 ? row[idx] : row[idx] - row[prev];
            prev = idx;
        }
        lin[tb + n] = 4 * (c0 * sx + c1 * sy + c2);  // byte offset
        k64[tb + n] = k;
        k32[tb + n] = (float)k;
        off[3 * (tb + n) + 0] = c0;
        off[3 * (tb + n) + 1] = c1;
        off[3 * (tb + n) + 2] = c2;
        ++n;
        if (clamped)
            break;  // later taps would add cum[5995] - cum[5995] = 0
        d.advance(a, c0, c1, c2);
        // The next tap (index n) is the first one with this step count.
        if (a == 0) ftab[fb0 + (++n0)] = n;
        else if (a == 1) ftab[fb1 + (++n1)] = n;
        else ftab[fb2 + (++n2)] = n;
    }
    // Zero padding lets the gather read a whole unrolled chunk past the end.
    for (int32_t q = n; q < n + kTapPad; ++q) {
        lin[tb + q] = 0;
        k64[tb + q] = 0.0;
        k32[tb + q] = 0.0f;
        off[3 * (tb + q) + 0] = 0;
        off[3 * (tb + q) + 1] = 0;
        off[3 * (tb + q) + 2] = 0;
    }
    ftab[fb0 + n0 + 1] = n;
    ftab[fb1 + n1 + 1] = n;
    ftab[fb2 + n2 + 1] = n;
    ConeMeta m;
    m.tap_begin = tb;
    m.n_taps = n;
    m.f_begin[0] = fb0;
    m.f_begin[1] = fb1;
    m.f_begin[2] = fb2;
    m.f_len[0] = n0 + 2;
    m.f_len[1] = n1 + 2;
    m.f_len[2] = n2 + 2;
    m.step[0] = d.s0;
    m.step[1] = d.s1;
    m.step[2] = d.s2;
    m.s_ch = 0;  // set on the host by stage_cones
    m.s_ext[0] = m.s_ext[1] = m.s_ext[2] = 0;
    m.s_nstage = 0;
    meta[c] = m;
}

void plan_release(ch_plan* p)
{
    if (!p)
        return;
    (void)hipFree(p->d_lin);
    (void)hipFree(p->d_k);
    (void)hipFree(p->d_k64);
    (void)hipFree(p->d_off);
    (void)hipFree(p->d_lin2);
    (void)hipFree(p->d_ftab);
    (void)hipFree(p->d_ftT);
    (void)hipFree(p->d_meta);
    (void)hipFree(p->d_cum);
    (void)hipFree(p->d_perm);
    (void)hipFree(p->d_row_begin);
    spectrum_free(p);
    delete p;
}

// Chunking of every cone for the staged gather.  A chunk of L consecutive
// taps moves the source box by at most w_a = max |o_{m+L-1,a} - o_{m,a}|
// along axis a, so its TERMA box fits extents (kSBX, kSBY, kSBZ) + w.  Each
// cone takes the longest L <= kSChunk whose box fits the LDS buffer, the
// per-thread staging registers and the y/z padding; L = 1 always fits.
// lin2[m] is tap m's byte offset in that box layout (rows of kSPitch
// floats).
static void stage_cones(std::vector<ConeMeta>& meta,
                        const std::vector<int32_t>& off, int32_t tap_stride,
                        std::vector<int32_t>& lin2)
{
    lin2.assign(meta.size() * (size_t)tap_stride, 0);
    for (size_t c = 0; c < meta.size(); ++c) {
        ConeMeta& m = meta[c];
        const int32_t* o = off.data() + 3 * (size_t)m.tap_begin;
        const int32_t n = m.n_taps;
        int32_t w[3] = {0, 0, 0};
        int32_t ch = kSChunk;
        for (; ch >= 1; --ch) {
            const int32_t L = n < ch ? n : ch;
            w[0] = w[1] = w[2] = 0;
            for (int32_t i = 0; L > 0 && i + L - 1 < n; ++i)
                for (int a = 0; a < 3; ++a) {
                    const int32_t d = std::abs(o[3 * (i + L - 1) + a] -
                                               o[3 * i + a]);
                    w[a] = d > w[a] ? d : w[a];
                }
            const int32_t ex = kSBX + w[0], ey = kSBY + w[1],
                          ez = kSBZ + w[2];
            if (w[1] <= kSMaxWY && w[2] <= kSMaxWZ &&
                ex * ey * kSPitch <= kSLds && ex * ey * ez <= kSStage * 256)
                break;
        }
        if (ch < 1) {  // unreachable: L = 1 gives w = 0
            ch = 1;
            w[0] = w[1] = w[2] = 0;
        }
        m.s_ch = ch;
        m.s_ext[0] = kSBX + w[0];
        m.s_ext[1] = kSBY + w[1];
        m.s_ext[2] = kSBZ + w[2];
        m.s_nstage = (m.s_ext[0] * m.s_ext[1] * m.s_ext[2] + 255) / 256;
        const int32_t px = m.s_ext[1] * kSPitch;
        int32_t* l2 = lin2.data() + m.tap_begin;
        for (int32_t t = 0; t < n; ++t)
            l2[t] = 4 * (o[3 * t] * px + o[3 * t + 1] * kSPitch + o[3 * t + 2]);
    }
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_plan_create(const int32_t h_shape[3], const double h_vsize[3],
                         int32_t n_cones, const double* h_dirs,
                         const int32_t* h_cone_row, int32_t n_rows,
                         int32_t row_len, const double* h_cum,
                         int32_t semantics, ch_plan** plan_out)
{
    CH_REQUIRE(plan_out && h_shape && h_vsize, "ch_plan_create: null pointer");
    *plan_out = nullptr;
    CH_REQUIRE(h_shape[0] > 0 && h_shape[1] > 0 && h_shape[2] > 0,
               "ch_plan_create: empty grid %d x %d x %d", h_shape[0],
               h_shape[1], h_shape[2]);
    CH_REQUIRE((int64_t)h_shape[0] * (h_shape[1] + 2 * kYPad) *
                       (h_shape[2] + 2 * kZPad) < (1LL << 29),
               "ch_plan_create: grid too large for 32-bit byte offsets");
    CH_REQUIRE(n_cones > 0 && h_dirs && h_cone_row,
               "ch_plan_create: no cones");
    CH_REQUIRE(n_rows > 0 && h_cum, "ch_plan_create: no kernel rows");
    CH_REQUIRE(row_len >= kRowLenMin,
               "ch_plan_create: kernel rows have %d values, need >= %d",
               row_len, kRowLenMin);
    CH_REQUIRE(semantics == CH_SEM_CONVOLVE_C ||
                   semantics == CH_SEM_CUDA_DOSE ||
                   semantics == CH_SEM_RADIOLOGICAL,
               "ch_plan_create: unknown semantics %d", semantics);
    for (int a = 0; a < 3; ++a)
        CH_REQUIRE(h_vsize[a] > 0, "ch_plan_create: voxel size <= 0");
    for (int32_t c = 0; c < n_cones; ++c)
        CH_REQUIRE(h_cone_row[c] >= 0 && h_cone_row[c] < n_rows,
                   "ch_plan_create: cone %d uses kernel row %d of %d", c,
                   h_cone_row[c], n_rows);

    ch_plan* p = new ch_plan();
    p->release = [](GraphHeld* h) { plan_release(static_cast<ch_plan*>(h)); };
    memset(p->shape, 0, sizeof p->shape);
    for (int a = 0; a < 3; ++a) {
        p->shape[a] = h_shape[a];
        p->vsize[a] = h_vsize[a];
    }
    p->n_cones = n_cones;
    p->semantics = semantics;
    p->tap_stride = h_shape[0] + h_shape[1] + h_shape[2] + kTapPad;
    p->f_stride = h_shape[0] + h_shape[1] + h_shape[2] + 6;
    p->tile[0] = 1;
    p->tile[1] = 2;
    p->tile[2] = 32;
    p->accum_mode = 2;
    p->d_lin = nullptr;
    p->d_k = nullptr;
    p->d_k64 = nullptr;
    p->d_off = nullptr;
    p->d_lin2 = nullptr;
    p->d_ftab = nullptr;
    p->d_ftT = nullptr;
    p->d_meta = nullptr;
    p->d_cum = nullptr;
    p->d_perm = nullptr;
    p->d_row_begin = nullptr;
    p->n_rows = n_rows;
    p->row_len = row_len;
    p->fft_n[0] = p->fft_n[1] = p->fft_n[2] = 0;
    p->fft_kp = 0;
    p->d_khat = nullptr;
    p->d_tw[0] = p->d_tw[1] = p->d_tw[2] = nullptr;

    const size_t nt = (size_t)n_cones * p->tap_stride;
    const size_t nf = (size_t)n_cones * p->f_stride;
    double* d_dirs = nullptr;
    int32_t* d_rows = nullptr;
    double* d_cum = nullptr;
    hipError_t e = hipSuccess;
#define CH_TRY(x)                                                           \
    do {                                                                    \
        if (e == hipSuccess)                                                \
            e = (x);                                                        \
    } while (0)
    CH_TRY(hipMalloc((void**)&p->d_lin, nt * sizeof(int32_t)));
    CH_TRY(hipMalloc((void**)&p->d_k, nt * sizeof(float)));
    CH_TRY(hipMalloc((void**)&p->d_k64, nt * sizeof(double)));
    CH_TRY(hipMalloc((void**)&p->d_off, nt * 3 * sizeof(int32_t)));
    CH_TRY(hipMalloc((void**)&p->d_lin2, nt * sizeof(int32_t)));
    CH_TRY(hipMalloc((void**)&p->d_ftab, nf * sizeof(int32_t)));
    CH_TRY(hipMalloc((void**)&p->d_meta, n_cones * sizeof(ConeMeta)));
    CH_TRY(hipMalloc((void**)&d_dirs, (size_t)n_cones * 3 * sizeof(double)));
    CH_TRY(hipMalloc((void**)&d_rows, (size_t)n_cones * sizeof(int32_t)));
    CH_TRY(hipMalloc((void**)&d_cum,
                     (size_t)n_rows * row_len * sizeof(double)));
    CH_TRY(hipMemcpy(d_dirs, h_dirs, (size_t)n_cones * 3 * sizeof(double),
                     hipMemcpyHostToDevice));
    CH_TRY(hipMemcpy(d_rows, h_cone_row, (size_t)n_cones * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    CH_TRY(hipMemcpy(d_cum, h_cum, (size_t)n_rows * row_len * sizeof(double),
                     hipMemcpyHostToDevice));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(build_taps_kernel,
                           dim3((unsigned)((n_cones + 63) / 64)), dim3(64), 0,
                           nullptr, n_cones, d_dirs, d_rows, d_cum, row_len,
                           h_shape[0], h_shape[1], h_shape[2], h_vsize[0],
                           h_vsize[1], h_vsize[2], semantics, p->tap_stride,
                           p->f_stride, p->d_lin, p->d_k, p->d_k64, p->d_off,
                           p->d_ftab, p->d_meta);
        e = hipGetLastError();
    }
    p->h_meta.resize(n_cones);
    CH_TRY(hipMemcpy(p->h_meta.data(), p->d_meta,
                     (size_t)n_cones * sizeof(ConeMeta),
                     hipMemcpyDeviceToHost));
    if (e == hipSuccess) {
        std::vector<int32_t> off(nt * 3), lin2;
        CH_TRY(hipMemcpy(off.data(), p->d_off, nt * 3 * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
        if (e == hipSuccess)
            stage_cones(p->h_meta, off, p->tap_stride, lin2);
        CH_TRY(hipMemcpy(p->d_lin2, lin2.data(), nt * sizeof(int32_t),
                         hipMemcpyHostToDevice));
        CH_TRY(hipMemcpy(p->d_meta, p->h_meta.data(),
                         (size_t)n_cones * sizeof(ConeMeta),
                         hipMemcpyHostToDevice));
    }
    if (e == hipSuccess) {
        // cone-minor copy of the first-step tables
        std::vector<int32_t> ft(nf);
        CH_TRY(hipMemcpy(ft.data(), p->d_ftab, nf * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
        const int32_t X = h_shape[0], Y = h_shape[1], Z = h_shape[2];
        const size_t nt_t = (size_t)(X + 2 + Y + 2 + Z + 2) * n_cones;
        std::vector<int32_t> ftT(nt_t);
        const int32_t dims[3] = {X, Y, Z};
        for (int a = 0; a < 3; ++a) {
            const int64_t base = ftab_t_base(a, X, Y, n_cones);
            for (int32_t c = 0; c < n_cones; ++c) {
                const ConeMeta& m = p->h_meta[c];
                const int32_t* f = ft.data() + m.f_begin[a];
                for (int32_t q = 0; q < dims[a] + 2; ++q)
                    ftT[base + (int64_t)q * n_cones + c] =
                        f[q < m.f_len[a] ? q : m.f_len[a] - 1];
            }
        }
        CH_TRY(hipMalloc((void**)&p->d_ftT, nt_t * sizeof(int32_t)));
        CH_TRY(hipMemcpy(p->d_ftT, ftT.data(), nt_t * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    }
    if (semantics == CH_SEM_RADIOLOGICAL) {
        // cones ordered by kernel row, so a workgroup stages each row once
        std::vector<int32_t> begin(n_rows + 1, 0), perm(n_cones);
        for (int32_t c = 0; c < n_cones; ++c)
            ++begin[h_cone_row[c] + 1];
        for (int32_t r = 0; r < n_rows; ++r)
            begin[r + 1] += begin[r];
        std::vector<int32_t> fill(begin.begin(), begin.end() - 1);
        for (int32_t c = 0; c < n_cones; ++c)
            perm[fill[h_cone_row[c]]++] = c;
        CH_TRY(hipMalloc((void**)&p->d_perm, n_cones * sizeof(int32_t)));
        CH_TRY(hipMalloc((void**)&p->d_row_begin,
                         (n_rows + 1) * sizeof(int32_t)));
        CH_TRY(hipMemcpy(p->d_perm, perm.data(), n_cones * sizeof(int32_t),
                         hipMemcpyHostToDevice));
        CH_TRY(hipMemcpy(p->d_row_begin, begin.data(),
                         (n_rows + 1) * sizeof(int32_t),
                         hipMemcpyHostToDevice));
        if (e == hipSuccess) {
            p->d_cum = d_cum;  // kept: the increments are formed per launch
            d_cum = nullptr;
        }
    }
#undef CH_TRY
    (void)hipFree(d_dirs);
    (void)hipFree(d_rows);
    (void)hipFree(d_cum);
    if (e != hipSuccess) {
        plan_release(p);
        return hip_fail(e, "ch_plan_create", __FILE__, __LINE__);
    }
    p->n_taps_total = 0;
    p->max_taps = 0;
    for (int32_t c = 0; c < n_cones; ++c) {
        p->n_taps_total += p->h_meta[c].n_taps;
        if (p->h_meta[c].n_taps > p->max_taps)
            p->max_taps = p->h_meta[c].n_taps;
    }
    *plan_out = p;
    return CH_OK;
}

ch_status ch_plan_destroy(ch_plan* plan)
{
    // a plan read by a live graph is freed with the last such graph
    if (plan && plan_defer_free(plan))
        return CH_OK;
    plan_release(plan);
    return CH_OK;
}

ch_status ch_plan_graphs(const ch_plan* plan, int32_t* n_graphs)
{
    CH_REQUIRE(plan && n_graphs, "ch_plan_graphs: null pointer");
    *n_graphs = plan_pins(plan);
    return CH_OK;
}

ch_status ch_plan_info(const ch_plan* plan, int32_t* n_cones, int64_t* n_taps,
                       int32_t* max_taps)
{
    CH_REQUIRE(plan, "ch_plan_info: null plan");
    if (n_cones)
        *n_cones = plan->n_cones;
    if (n_taps)
        *n_taps = plan->n_taps_total;
    if (max_taps)
        *max_taps = plan->max_taps;
    return CH_OK;
}

ch_status ch_plan_cone_taps(const ch_plan* plan, int32_t cone,
                            int32_t* h_off_out, double* h_k_out, int32_t cap,
                            int32_t* n_out)
{
    CH_REQUIRE(plan && n_out, "ch_plan_cone_taps: null pointer");
    CH_REQUIRE(cone >= 0 && cone < plan->n_cones, "cone %d out of range",
               cone);
    const ConeMeta& m = plan->h_meta[cone];
    *n_out = m.n_taps;
    if (m.n_taps > cap) {
        set_error("ch_plan_cone_taps: %d taps exceed capacity %d", m.n_taps,
                  cap);
        return CH_ERR_CAPACITY;
    }
    if (m.n_taps == 0)
        return CH_OK;
    CH_HIP(hipMemcpy(h_off_out, plan->d_off + 3 * (size_t)m.tap_begin,
                     (size_t)m.n_taps * 3 * sizeof(int32_t),
                     hipMemcpyDeviceToHost));
    CH_HIP(hipMemcpy(h_k_out, plan->d_k64 + (size_t)m.tap_begin,
                     (size_t)m.n_taps * sizeof(double),
                     hipMemcpyDeviceToHost));
    return CH_OK;
}

ch_status ch_plan_set_tile(ch_plan* plan, int32_t tx, int32_t ty, int32_t tz)
{
    CH_REQUIRE(plan, "ch_plan_set_tile: null plan");
    const bool staged = tx == kSBX && ty == kSBY && tz == kSBZ;
    CH_REQUIRE(staged || (tx == 1 && ty * tz == 64 && ty >= 1 && ty <= 8 &&
                          (ty & (ty - 1)) == 0),
               "tile %dx%dx%d: either a wavefront owns 1 x TY x TZ = 64 "
               "voxels, TY in {1, 2, 4, 8}, or a workgroup owns the staged "
               "block %d x %d x %d",
               tx, ty, tz, kSBX, kSBY, kSBZ);
    plan->tile[0] = tx;
    plan->tile[1] = ty;
    plan->tile[2] = tz;
    return CH_OK;
}

ch_status ch_plan_get_tile(const ch_plan* plan, int32_t* tx, int32_t* ty,
                           int32_t* tz)
{
    CH_REQUIRE(plan, "ch_plan_get_tile: null plan");
    *tx = plan->tile[0];
    *ty = plan->tile[1];
    *tz = plan->tile[2];
    return CH_OK;
}

ch_status ch_plan_set_accum(ch_plan* plan, int32_t mode)
{
    CH_REQUIRE(plan, "ch_plan_set_accum: null plan");
    CH_REQUIRE(mode >= 0 && mode <= 2, "accumulation mode %d", mode);
    plan->accum_mode = mode;
    return CH_OK;
}

}  // extern "C"
