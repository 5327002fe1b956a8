// resample.hip -- linear resampling of phantom densities onto a dose grid.
//
// The reference builds scipy's RegularGridInterpolator (method 'linear',
// bounds_error False, fill_value 0) on the phantom axes
// linspace(o_a, o_a + n_a s_a, n_a) (conehead/conehead.py:525-533) and
// evaluates it at the dose grid's axes built the same way, laid out by
// np.meshgrid with indexing 'xy' (:562-567), so output[i][j][k] is the
// value at (x_j, y_i, z_k).  One thread per output voxel.  The arithmetic
// follows numpy and scipy in FP64 without contraction: linspace nodes are
// k * step + start with the last node equal to stop, the cell of a
// coordinate is searchsorted(nodes, x) - 1 clamped to [0, n - 2], points
// outside [node_0, node_{n-1}] give 0, and the eight corner weights are
// products of (1 - t) and t accumulated corner by corner in
// itertools.product order.
#include "ch_internal.h"

#include <cmath>
#include <cstring>

namespace ch {

struct Axis {
    double start, stop;
    int n;
};

// numpy.linspace(start, stop, n)[k]
__device__ __forceinline__ double node(const Axis& a, int k)
{
    if (a.n == 1)
        return a.start;
    if (k == a.n - 1)
        return a.stop;
    const double step = (a.stop - a.start) / (double)(a.n - 1);
    return (double)k * step + a.start;
}

// Cell index and fraction of x on the axis; false when x is outside.
__device__ __forceinline__ bool locate(const Axis& a, double x, int& i,
                                       double& t)
{
    const double first = node(a, 0), last = node(a, a.n - 1);
    if (!(x >= first && x <= last))
        return false;
    if (a.n == 1) {
        i = 0;
        t = 0.0;
        return true;
    }
    // searchsorted(side='left'): first k with node(k) >= x
    int lo = 0, hi = a.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (node(a, mid) < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    i = lo - 1;
    i = i < 0 ? 0 : (i > a.n - 2 ? a.n - 2 : i);
    const double g0 = node(a, i), g1 = node(a, i + 1);
    t = (x - g0) / (g1 - g0);
    return true;
}

__global__ __launch_bounds__(256) void resample_kernel(
    const float* __restrict__ src, int sx, int sy, int sz, Axis ax, Axis ay,
    Axis az, float* __restrict__ dst, int dx, int dy, int dz, Axis bx,
    Axis by, Axis bz, int swap)
{
    const int64_t li = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (li >= (int64_t)dx * dy * dz)
        return;
    const int c = (int)(li % dz);
    const int b = (int)((li / dz) % dy);
    const int a = (int)(li / ((int64_t)dz * dy));
    // output index (a, b, c) -> sample point
    const double x = swap ? node(bx, b) : node(bx, a);
    const double y = swap ? node(by, a) : node(by, b);
    const double z = node(bz, c);
    int i0, i1, i2;
    double t0, t1, t2;
    float out = 0.0f;
    if (locate(ax, x, i0, t0) && locate(ay, y, i1, t1) &&
        locate(az, z, i2, t2)) {
        double v = 0.0;
        for (int e0 = 0; e0 < 2; ++e0)
            for (int e1 = 0; e1 < 2; ++e1)
                for (int e2 = 0; e2 < 2; ++e2) {
                    double w = 1.0;
                    w = w * (e0 ? t0 : 1.0 - t0);
                    w = w * (e1 ? t1 : 1.0 - t1);
                    w = w * (e2 ? t2 : 1.0 - t2);
                    const int j0 = sx == 1 ? 0 : i0 + e0;
                    const int j1 = sy == 1 ? 0 : i1 + e1;
                    const int j2 = sz == 1 ? 0 : i2 + e2;
                    const double val =
                        (double)src[((int64_t)j0 * sy + j1) * sz + j2];
                    v = v + val * w;
                }
        out = (float)v;
    }
    dst[li] = out;
}

// CT numbers to densities: a piecewise-linear calibration curve evaluated
// as numpy.interp does (clamped to the end values outside the curve, the
// node's own density at a node, slope * (x - x_j) + y_j between nodes, FP64
// without contraction), then a per-label override for the voxels of a
// label map (structures forced to a density).  Both tables travel in the
// kernel arguments.
struct CtTable {
    double hu[CH_CT_MAX_POINTS];
    double rho[CH_CT_MAX_POINTS];
    double slope[CH_CT_MAX_POINTS];
    float override_rho[256];  // NaN: keep the curve's density
    int n;
};

__device__ __forceinline__ float ct_density(const CtTable& c, double x)
{
    if (x < c.hu[0])
        return (float)c.rho[0];
    if (x >= c.hu[c.n - 1])
        return (float)c.rho[c.n - 1];
    int j = 0;  // the last node at or below x
    for (int k = 1; k < c.n - 1; ++k)
        j = x >= c.hu[k] ? k : j;
    if (x == c.hu[j])
        return (float)c.rho[j];
    return (float)(c.slope[j] * (x - c.hu[j]) + c.rho[j]);
}

template <typename T>
__global__ __launch_bounds__(256) void ct_kernel(
    const T* __restrict__ hu, const uint8_t* __restrict__ labels, int64_t n,
    CtTable c, float* __restrict__ rho)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float v = ct_density(c, (double)hu[i]);
    if (labels) {
        const float o = c.override_rho[labels[i]];
        if (o == o)
            v = o;
    }
    rho[i] = v;
}

static Axis axis(double origin, double spacing, int n)
{
    Axis a;
    a.start = origin;
    a.stop = origin + (double)n * spacing;
    a.n = n;
    return a;
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_resample_linear(const float* d_src, const int32_t src_shape[3],
                             const double src_origin[3],
                             const double src_spacing[3], float* d_dst,
                             const int32_t dst_shape[3],
                             const double dst_origin[3],
                             const double dst_spacing[3], int32_t xy_swap,
                             void* stream)
{
    CH_REQUIRE(d_src && d_dst && src_shape && src_origin && src_spacing &&
                   dst_shape && dst_origin && dst_spacing,
               "ch_resample_linear: null pointer");
    for (int k = 0; k < 3; ++k) {
        CH_REQUIRE(src_shape[k] > 0 && dst_shape[k] > 0,
                   "ch_resample_linear: empty grid");
        CH_REQUIRE(src_spacing[k] > 0 && dst_spacing[k] > 0,
                   "ch_resample_linear: spacing must be > 0");
    }
    CH_REQUIRE(!xy_swap || dst_shape[0] == dst_shape[1],
               "ch_resample_linear: the 'xy' layout needs size[0] == size[1] "
               "(got %d x %d)",
               dst_shape[0], dst_shape[1]);
    const int64_t n = (int64_t)dst_shape[0] * dst_shape[1] * dst_shape[2];
    CH_REQUIRE(n < (1LL << 31), "ch_resample_linear: grid too large");
    hipLaunchKernelGGL(
        resample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
        (hipStream_t)stream, d_src, src_shape[0], src_shape[1], src_shape[2],
        axis(src_origin[0], src_spacing[0], src_shape[0]),
        axis(src_origin[1], src_spacing[1], src_shape[1]),
        axis(src_origin[2], src_spacing[2], src_shape[2]), d_dst,
        dst_shape[0], dst_shape[1], dst_shape[2],
        axis(dst_origin[0], dst_spacing[0], dst_shape[0]),
        axis(dst_origin[1], dst_spacing[1], dst_shape[1]),
        axis(dst_origin[2], dst_spacing[2], dst_shape[2]), xy_swap ? 1 : 0);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_ct_to_density(const void* d_ct, int32_t ct_type, int64_t n,
                           const double* curve_hu, const double* curve_rho,
                           int32_t n_points, const uint8_t* d_labels,
                           const float* override_rho, float* d_rho,
                           void* stream)
{
    CH_REQUIRE(d_ct && d_rho && curve_hu && curve_rho,
               "ch_ct_to_density: null pointer");
    CH_REQUIRE(ct_type == CH_CT_I16 || ct_type == CH_CT_F32,
               "ch_ct_to_density: unknown CT number type %d", ct_type);
    CH_REQUIRE(n >= 0 && n < (1LL << 40), "ch_ct_to_density: bad count");
    CH_REQUIRE(n_points >= 1 && n_points <= CH_CT_MAX_POINTS,
               "ch_ct_to_density: the curve needs 1 to %d points (got %d)",
               CH_CT_MAX_POINTS, n_points);
    CH_REQUIRE(!d_labels || override_rho,
               "ch_ct_to_density: a label map needs its override table");
    CtTable c;
    memset(&c, 0, sizeof(c));
    c.n = n_points;
    for (int k = 0; k < n_points; ++k) {
        CH_REQUIRE(std::isfinite(curve_hu[k]) && std::isfinite(curve_rho[k]),
                   "ch_ct_to_density: curve point %d is not finite", k);
        CH_REQUIRE(k == 0 || curve_hu[k] > curve_hu[k - 1],
                   "ch_ct_to_density: curve CT numbers must increase "
                   "(point %d)", k);
        c.hu[k] = curve_hu[k];
        c.rho[k] = curve_rho[k];
    }
    for (int k = 0; k + 1 < n_points; ++k)
        c.slope[k] = (c.rho[k + 1] - c.rho[k]) / (c.hu[k + 1] - c.hu[k]);
    for (int k = 0; k < 256; ++k)
        c.override_rho[k] = override_rho ? override_rho[k] : NAN;
    if (n == 0)
        return CH_OK;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (ct_type == CH_CT_I16)
        hipLaunchKernelGGL(ct_kernel<int16_t>, grid, dim3(256), 0,
                           (hipStream_t)stream, (const int16_t*)d_ct,
                           d_labels, n, c, d_rho);
    else
        hipLaunchKernelGGL(ct_kernel<float>, grid, dim3(256), 0,
                           (hipStream_t)stream, (const float*)d_ct, d_labels,
                           n, c, d_rho);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

}  // extern "C"
