// physics.hip -- physics-mode convolution: cones aligned with the beam and
// kernel increments looked up at the radiological distance.
//
// The reference convolves with world-fixed cones and geometric distances
// and lists kernel tilting as a milestone (README.md:22).  Here the caller
// rotates the cone set into the beam frame (the directions handed to
// ch_plan_create), and for a receiving voxel v and cone c with path offsets
// o_m and exit t-values t_m the increments are
//   R_m = sum_{j<=m} rho(v - o_j) (t_j - t_{j-1}),  t_{-1} = 0,
//   k_m = cum[i(R_m)] - cum[i(R_{m-1})],  i(R) = min(|floor(100 R - 5)|, 5995)
// (the index of cuda_dose, conehead/conehead.py:453-463, taken at R instead
// of t), and dose[v] = sum_c sum_m k_m T(v - o_m).  The path lengths
// t_m - t_{m-1} come from the plan (CH_SEM_RADIOLOGICAL).
//
// The gather keeps the direct gather's layout: a wavefront owns a 1 x 4 x 16
// output tile and reads T and the densities from padded copies with
// range-checked buffer loads.  Every lane walks its cone from its own voxel
// (R needs the whole path), loading densities only up to the first tap of
// the tile's union range and both grids from there on.  A workgroup of 8
// wavefronts stages each kernel row (5996 FP64 values) in LDS once and runs
// the cones of that row.  R, the index and the increments are FP64 and
// round like the FP64 oracle (this file is built without contraction); the
// sums are FP32 with compensation.
#include <climits>

#include "ch_internal.h"

namespace ch {

typedef __attribute__((address_space(4))) const int32_t c_i32;
typedef __attribute__((address_space(4))) const double c_f64;

struct RadJob {
    const float* terma;    // padded copies, interior voxel (0, 0, 0)
    const float* density;
    float* dose;
    const int32_t* box;
    const int32_t* lin;    // per tap: byte offset into the padded grids
    const double* seg;     // per tap: path length inside the voxel
    const int32_t* ftT;    // first-step tables, cone-minor
    const ConeMeta* meta;
    const int32_t* perm;   // cones ordered by kernel row
    const int32_t* row_begin;
    const double* cum;
    int32_t row_len, n_rows;
    int32_t X, Y, Z, n_cones, accumulate, nty, ntz;
};

constexpr int kRadTY = 4, kRadTZ = 16, kRadWaves = 8;
constexpr int kRadRow = kMaxIndex + 1;

// trunc(x) as int32, saturating (v_cvt_i32_f64 clamps to the int32 range
// and reads NaN as 0; a C++ cast of an out-of-range value is undefined).
__device__ __forceinline__ int trunc_sat(double x)
{
    int r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// |floor(x - 5)| clamped to kMaxIndex, x = 100 R as the FP64 product.
__device__ __forceinline__ int rad_index_of(double x)
{
    const double f = floor(x - 5.0);
    if (!(f < 2147483647.0 && f > -2147483648.0))
        return kMaxIndex;
    int i = (int)f;
    i = i < 0 ? -i : i;
    return i > kMaxIndex ? kMaxIndex : i;
}

__device__ __forceinline__ int rad_index(double r)
{
    const double f = floor(r * 100.0 - 5.0);
    if (!(f < 2147483647.0 && f > -2147483648.0))
        return kMaxIndex;
    int i = (int)f;
    i = i < 0 ? -i : i;
    return i > kMaxIndex ? kMaxIndex : i;
}

__device__ __forceinline__ void rad_kahan(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Densities (and TERMA) of kUnroll taps from tap t of the tap tables ml,
// relative to the lane's voxel at byte offset base4 of the padded grids.
__device__ __forceinline__ void load_taps(__amdgpu_buffer_rsrc_t rr,
                                          uint32_t base4, const c_i32* ml,
                                          int t, float (&rho)[kUnroll])
{
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
        rho[u] = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(
            rr, (int)(base4 - (uint32_t)ml[t + u]), 0, 0));
}

__device__ __forceinline__ void load_taps2(__amdgpu_buffer_rsrc_t rr,
                                           __amdgpu_buffer_rsrc_t rt,
                                           uint32_t base4, const c_i32* ml,
                                           int t, float (&rho)[kUnroll],
                                           float (&tv)[kUnroll])
{
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        const int o = (int)(base4 - (uint32_t)ml[t + u]);
        rho[u] = __int_as_float(
            __builtin_amdgcn_raw_buffer_load_b32(rr, o, 0, 0));
        tv[u] = __int_as_float(
            __builtin_amdgcn_raw_buffer_load_b32(rt, o, 0, 0));
    }
}

__global__ __launch_bounds__(64 * kRadWaves) void gather_rad(RadJob j,
                                                            int n_tiles)
{
    __shared__ double s_cum[kRadRow];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kRadWaves + wave;
    const bool live = tile < n_tiles;
    int x = 0, y0 = 0, z0 = 0;
    if (live) {
        const int tz = tile % j.ntz;
        const int r = tile / j.ntz;
        y0 = (r % j.nty) * kRadTY;
        x = r / j.nty;
        z0 = tz * kRadTZ;
    }
    const int yl = min(y0 + kRadTY - 1, j.Y - 1);
    const int zl = min(z0 + kRadTZ - 1, j.Z - 1);
    const int y = y0 + lane / kRadTZ;
    const int z = z0 + lane % kRadTZ;
    const int Zp = j.Z + 2 * kZPad;
    const int Yp = j.Y + 2 * kYPad;

    int lo0 = 0, lo1 = 0, lo2 = 0;
    int hi0 = j.X - 1, hi1 = j.Y - 1, hi2 = j.Z - 1;
    if (j.box) {
        const c_i32* b = (const c_i32*)j.box;
        lo0 = max(lo0, b[0]);
        lo1 = max(lo1, b[1]);
        lo2 = max(lo2, b[2]);
        hi0 = min(hi0, b[3]);
        hi1 = min(hi1, b[4]);
        hi2 = min(hi2, b[5]);
    }
    const bool work = live && lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2;
    const uint32_t base4 =
        (uint32_t)(((x * Yp + y + kYPad) * Zp + z + kZPad) * 4);
    const size_t interior = (size_t)kYPad * Zp + kZPad;
    const int bytes = j.X * Yp * Zp * 4;
    const __amdgpu_buffer_rsrc_t rt = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(j.terma - interior), (short)0, bytes, kBufferFlags);
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(j.density - interior), (short)0, bytes, kBufferFlags);
    const c_i32* lin = (const c_i32*)j.lin;
    const c_f64* seg = (const c_f64*)j.seg;

    float sum = 0.0f, comp = 0.0f;
    for (int row = 0; row < j.n_rows; ++row) {
        __syncthreads();
        const double* src = j.cum + (size_t)row * j.row_len;
        for (int i = threadIdx.x; i < kRadRow; i += 64 * kRadWaves)
            s_cum[i] = src[i];
        __syncthreads();
        if (!work)
            continue;
        const int cb = j.row_begin[row], ce = j.row_begin[row + 1];
        for (int q = cb; q < ce; q += 64) {
            // lane l: the tile's union range of cone perm[q + l]
            int r_lo = 0, r_hi = 0, r_tb = 0;
            if (q + lane < ce) {
                const int c = j.perm[q + lane];
                const ConeMeta* m = j.meta + c;
                const int n = j.n_cones;
                const int32_t* f0 = j.ftT;
                const int32_t* f1 = j.ftT + ftab_t_base(1, j.X, j.Y, n);
                const int32_t* f2 = j.ftT + ftab_t_base(2, j.X, j.Y, n);
                int ax, bx, ay1, by1, ay2, by2, az1, bz1, az2, bz2;
                vrange_t(x, lo0, hi0, m->step[0], f0, j.X + 2, n, c, ax, bx);
                vrange_t(y0, lo1, hi1, m->step[1], f1, j.Y + 2, n, c, ay1,
                         by1);
                vrange_t(yl, lo1, hi1, m->step[1], f1, j.Y + 2, n, c, ay2,
                         by2);
                vrange_t(z0, lo2, hi2, m->step[2], f2, j.Z + 2, n, c, az1,
                         bz1);
                vrange_t(zl, lo2, hi2, m->step[2], f2, j.Z + 2, n, c, az2,
                         bz2);
                r_lo = max(max(ax, min(ay1, ay2)), min(az1, az2));
                r_hi = min(min(bx, max(by1, by2)), max(bz1, bz2));
                r_tb = m->tap_begin;
            }
            const int nc = min(64, ce - q);
            for (int i = 0; i < nc; ++i) {
                const int ulo = __builtin_amdgcn_readlane(r_lo, i);
                const int uhi = __builtin_amdgcn_readlane(r_hi, i);
                if (ulo >= uhi)
                    continue;
                const int tb = __builtin_amdgcn_readlane(r_tb, i);
                const c_i32* ml = lin + tb;
                const c_f64* ms = seg + tb;
                // the path from the receiving voxel up to the union range:
                // densities only (every source there is outside the box)
                double R = 0.0;
                int t = 0;
                for (; t + kUnroll <= ulo; t += kUnroll) {
                    float rho[kUnroll];
                    load_taps(rr, base4, ml, t, rho);
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u)
                        R = R + (double)rho[u] * ms[t + u];
                }
                for (; t < ulo; ++t) {
                    const float rho = __int_as_float(
                        __builtin_amdgcn_raw_buffer_load_b32(
                            rr, (int)(base4 - (uint32_t)ml[t]), 0, 0));
                    R = R + (double)rho * ms[t];
                }
                double cprev = ulo > 0 ? s_cum[rad_index(R)] : 0.0;
                // whole chunks: a tap past the union end has T = 0 for
                // every lane, and the zero path lengths of the tap padding
                // leave R unchanged
                float acc = 0.0f, ac = 0.0f;
                for (t = ulo; t < uhi; t += kUnroll) {
                    float rho[kUnroll], tv[kUnroll];
                    load_taps2(rr, rt, base4, ml, t, rho, tv);
                    // The index |floor(100 R - 5)| clamped to kMaxIndex.
                    // For x = 100 R >= 5 the subtraction of 5 is exact
                    // (the difference is a multiple of ulp(x) below x), so
                    // the index is min(trunc(x), 6000) - 5; a chunk with a
                    // tap below 5 (or NaN, which truncates to 0) takes
                    // rad_index_of for all its taps.
                    double x[kUnroll];
                    int idx[kUnroll];
                    bool slow = false;
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        R = R + (double)rho[u] * ms[t + u];
                        x[u] = R * 100.0;
                        const int ti = trunc_sat(x[u]);
                        slow = slow || ti < 5;
                        idx[u] = min(ti, kMaxIndex + 5) - 5;
                    }
                    if (__builtin_amdgcn_ballot_w64(slow) != 0) {
                        // a real branch: x passes through an empty asm
                        // here, so the compiler cannot hoist this work
                        // onto every chunk
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            double xu = x[u];
                            asm volatile("" : "+v"(xu));
                            idx[u] = rad_index_of(xu);
                        }
                    }
                    float part = 0.0f;
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        const double cv = s_cum[idx[u]];
                        const float k = (float)(cv - cprev);
                        cprev = cv;
                        part = fmaf(k, tv[u], part);
                    }
                    rad_kahan(acc, ac, part);
                }
                rad_kahan(sum, comp, acc);
            }
        }
    }
    if (live && y <= yl && z <= zl) {
        float* d = j.dose + ((size_t)x * j.Y + y) * j.Z + z;
        *d = (j.accumulate ? *d : 0.0f) + sum;
    }
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_convolve_rad(const ch_plan* plan, const float* d_terma,
                          const float* d_density, float* d_dose,
                          const int32_t* d_box, int32_t accumulate,
                          void* stream)
{
    CH_REQUIRE(plan && d_terma && d_density && d_dose,
               "ch_convolve_rad: null pointer");
    CH_REQUIRE(plan->semantics == CH_SEM_RADIOLOGICAL,
               "ch_convolve_rad: the plan was built with semantics %d, not "
               "CH_SEM_RADIOLOGICAL",
               plan->semantics);
    CH_REQUIRE(plan->d_cum && plan->d_perm && plan->d_row_begin &&
                   plan->row_len >= kRadRow,
               "ch_convolve_rad: plan has no kernel rows");
    hipStream_t s = (hipStream_t)stream;
    const float* pt = nullptr;
    const float* pr = nullptr;
    ch_status st = pad_pair(plan, d_terma, d_density, s, &pt, &pr);
    if (st != CH_OK)
        return st;
    capture_pin(s, plan);
    RadJob j;
    j.terma = pt;
    j.density = pr;
    j.dose = d_dose;
    j.box = d_box;
    j.lin = plan->d_lin;
    j.seg = plan->d_k64;
    j.ftT = plan->d_ftT;
    j.meta = plan->d_meta;
    j.perm = plan->d_perm;
    j.row_begin = plan->d_row_begin;
    j.cum = plan->d_cum;
    j.row_len = plan->row_len;
    j.n_rows = plan->n_rows;
    j.X = plan->shape[0];
    j.Y = plan->shape[1];
    j.Z = plan->shape[2];
    j.n_cones = plan->n_cones;
    j.accumulate = accumulate;
    j.nty = (j.Y + kRadTY - 1) / kRadTY;
    j.ntz = (j.Z + kRadTZ - 1) / kRadTZ;
    const int64_t n_tiles = (int64_t)j.X * j.nty * j.ntz;
    CH_REQUIRE(n_tiles < INT_MAX / 2, "ch_convolve_rad: too many tiles");
    const unsigned blocks =
        (unsigned)((n_tiles + kRadWaves - 1) / kRadWaves);
    LaunchTimer timer(s);
    timer.launch(gather_rad, dim3(blocks), dim3(64 * kRadWaves), 0, j,
                 (int)n_tiles);
    CH_HIP(hipGetLastError());
    timer.stop();
    return CH_OK;
}

}  // extern "C"
