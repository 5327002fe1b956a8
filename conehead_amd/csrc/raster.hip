// raster.hip -- aperture transmission rasters (4000 x 4000 FP32, 0.1 mm
// pitch, row index ~ source x, column index ~ source z) built on the device,
// so a beam's aperture never crosses PCIe as a 64 MB array.
#include "ch_internal.h"

namespace ch {

// Block.set_square (conehead/block.py:64-85): block_values[x1:x2, y1:y2] = 1
// (numpy slice bounds clamped to the raster), other entries unchanged.
__global__ __launch_bounds__(256) void square_kernel(float* r, int x1, int x2,
                                                     int y1, int y2)
{
    const int w = y2 - y1;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)(x2 - x1) * w)
        return;
    const int row = x1 + (int)(i / w), col = y1 + (int)(i % w);
    r[row * 4000 + col] = 1.0f;
}

// Leaf transmission table of conehead/block.py:155-164 evaluated at one
// element: 1 - area, area = ones with edge rows 0.2/0.5/0.75 (the rows
// counted from the far edge are assigned last and win) and the last 15
// columns scaled by numpy.linspace(1, 0, 15).
__device__ __forceinline__ float leaf_value(int rr, int cc, int w, int h)
{
    float e = 1.0f;
    if (rr == w - 1) e = 0.20f;
    else if (rr == w - 2) e = 0.50f;
    else if (rr == w - 3) e = 0.75f;
    else if (rr == 0) e = 0.20f;
    else if (rr == 1) e = 0.50f;
    else if (rr == 2) e = 0.75f;
    if (cc >= h - 15) {
        const int q = cc - (h - 15);
        const double step = (0.0 - 1.0) / 14;
        const double lin = q == 14 ? 0.0 : (double)q * step + 1.0;
        e = (float)((double)e * lin);
    }
    return 1.0f - e;
}

// Leaf bank and jaws of conehead/block.py:130-196, in integer tenths of a
// millimetre.  The bank (width W) is centred on raster row 2000; rows
// outside it are blocked.
__global__ __launch_bounds__(256) void mlc_kernel(
    float* r, int n_pairs, const int32_t* __restrict__ bound,
    const int32_t* __restrict__ ea, const int32_t* __restrict__ eb, int jx0,
    int jx1, int jy0, int jy1)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 4000LL * 4000LL)
        return;
    const int row = (int)(i / 4000), col = (int)(i % 4000);
    const int b0 = bound[0];
    const int W = abs(b0 - bound[n_pairs]);
    const int rr = row - (2000 - W / 2);
    float v = 0.0f;
    if (rr >= 0 && rr < W) {
        v = 1.0f;
        const int shift = abs(b0);
        for (int n = 0; n < n_pairs; ++n) {          // bank A (:187-188)
            const int rmin = bound[n] + shift;
            const int w = abs(bound[n + 1] - bound[n]);
            const int cmax = 2000 + ea[n];
            if (rr >= rmin && rr < rmin + w && col < cmax)
                v = leaf_value(rr - rmin, col, w, cmax);
        }
        for (int n = 0; n < n_pairs; ++n) {          // bank B (:189-190)
            const int rmin = bound[n] + shift;
            const int w = abs(bound[n + 1] - bound[n]);
            const int cmin = 2000 + eb[n];
            const int h = 3999 - cmin;
            if (rr >= rmin && rr < rmin + w && col >= cmin && col < 3999)
                v = leaf_value(w - 1 - (rr - rmin), h - 1 - (col - cmin), w,
                               h);
        }
        if (col < 2000 + jx0 || col >= 2000 + jx1)   // jaws (:193-196)
            v = 0.0f;
        if (rr < jy0 || rr >= jy1)
            v = 0.0f;
    }
    r[i] = v;
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_raster_square(float* d_raster, double length, void* stream)
{
    CH_REQUIRE(d_raster, "ch_raster_square: null raster");
    // int((xnum / 2) - (length / 2) * xres) with xnum = 4000, xres = 100
    int x1 = (int)(2000.0 - (length / 2) * 100.0);
    int x2 = (int)(2000.0 + (length / 2) * 100.0);
    x1 = x1 < 0 ? 0 : (x1 > 4000 ? 4000 : x1);
    x2 = x2 < 0 ? 0 : (x2 > 4000 ? 4000 : x2);
    if (x2 <= x1)
        return CH_OK;
    const int64_t n = (int64_t)(x2 - x1) * (x2 - x1);
    hipLaunchKernelGGL(square_kernel, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, d_raster, x1, x2,
                       x1, x2);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_raster_mlc(float* d_raster, int32_t n_pairs,
                        const int32_t* h_bound, const int32_t* h_a,
                        const int32_t* h_b, const int32_t h_jaw_x[2],
                        const int32_t h_jaw_y[2], void* stream)
{
    CH_REQUIRE(d_raster && h_bound && h_a && h_b && h_jaw_x && h_jaw_y,
               "ch_raster_mlc: null pointer");
    CH_REQUIRE(n_pairs > 0 && n_pairs <= 512, "ch_raster_mlc: %d leaf pairs",
               n_pairs);
    const int W = abs(h_bound[0] - h_bound[n_pairs]);
    CH_REQUIRE(W > 0 && W <= 4000, "ch_raster_mlc: bank width %d", W);
    for (int n = 0; n < n_pairs; ++n) {
        CH_REQUIRE(h_bound[n + 1] > h_bound[n],
                   "ch_raster_mlc: boundaries must increase");
        CH_REQUIRE(2000 + h_a[n] >= 15 && 2000 + h_a[n] <= 4000 &&
                       3999 - (2000 + h_b[n]) >= 15 && 2000 + h_b[n] >= 0,
                   "ch_raster_mlc: leaf %d tips %d/%d outside the raster",
                   n, h_a[n], h_b[n]);
    }
    // Row limits of the Y jaws: int(mlc_width / 2 + jaw_y)
    const int jy0 = (int)(W / 2.0 + h_jaw_y[0]);
    const int jy1 = (int)(W / 2.0 + h_jaw_y[1]);
    int32_t* d_tab = nullptr;
    const size_t nb = (size_t)(3 * n_pairs + 1);
    CH_HIP(hipMalloc((void**)&d_tab, nb * sizeof(int32_t)));
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(d_tab, h_bound,
                                  (n_pairs + 1) * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tab + n_pairs + 1, h_a, n_pairs * sizeof(int32_t),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tab + 2 * n_pairs + 1, h_b,
                           n_pairs * sizeof(int32_t), hipMemcpyHostToDevice,
                           s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mlc_kernel, dim3(62500), dim3(256), 0, s, d_raster,
                           n_pairs, d_tab, d_tab + n_pairs + 1,
                           d_tab + 2 * n_pairs + 1, h_jaw_x[0], h_jaw_x[1],
                           jy0, jy1);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    hipFree(d_tab);
    if (e != hipSuccess)
        return hip_fail(e, "ch_raster_mlc", __FILE__, __LINE__);
    return CH_OK;
}

}  // extern "C"
