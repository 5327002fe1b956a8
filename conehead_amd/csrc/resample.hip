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

}  // extern "C"
