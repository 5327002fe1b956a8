// raster.hip -- aperture transmission rasters (4000 x 4000 FP32, 0.1 mm
// pitch, row index ~ source x, column index ~ source z) and leaf tables,
// built on the device so a beam's aperture never crosses PCIe as a 64 MB
// array.  Cell values come from aperture.h, the same functions the TERMA
// hit test evaluates analytically.
#include <cstdlib>
#include <vector>

#include "aperture.h"

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

__global__ __launch_bounds__(256) void mlc_kernel(float* r,
                                                  const int32_t* table)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 4000LL * 4000LL)
        return;
    r[i] = mlc_value(table, (int)(i / 4000), (int)(i % 4000));
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

int32_t ch_mlc_table_len(int32_t n_pairs)
{
    return mlc_row_table_start(n_pairs) + 4 * kMlcRows;
}

ch_status ch_mlc_table(int32_t* d_table, int32_t n_pairs,
                       const int32_t* h_bound, const int32_t* h_a,
                       const int32_t* h_b, const int32_t h_jaw_x[2],
                       const int32_t h_jaw_y[2], int32_t* h_bbox_out)
{
    CH_REQUIRE(d_table && h_bound && h_a && h_b && h_jaw_x && h_jaw_y,
               "ch_mlc_table: null pointer");
    CH_REQUIRE(n_pairs > 0 && n_pairs <= 512, "ch_mlc_table: %d leaf pairs",
               n_pairs);
    const int W = abs(h_bound[0] - h_bound[n_pairs]);
    CH_REQUIRE(W > 0 && W <= 4000, "ch_mlc_table: bank width %d", W);
    for (int n = 0; n < n_pairs; ++n) {
        CH_REQUIRE(h_bound[n + 1] > h_bound[n],
                   "ch_mlc_table: boundaries must increase");
        CH_REQUIRE(2000 + h_a[n] >= 15 && 2000 + h_a[n] <= 4000 &&
                       3999 - (2000 + h_b[n]) >= 15 && 2000 + h_b[n] >= 0,
                   "ch_mlc_table: leaf %d tips %d/%d outside the raster",
                   n, h_a[n], h_b[n]);
    }
    const int rt = mlc_row_table_start(n_pairs);
    std::vector<int32_t> t(rt + 4 * kMlcRows, 0);
    t[0] = n_pairs;
    t[1] = W;
    t[2] = 2000 - W / 2;                         // bank centred on row 2000
    t[3] = h_jaw_x[0];
    t[4] = h_jaw_x[1];
    t[5] = (int)(W / 2.0 + h_jaw_y[0]);          // int(mlc_width/2 + jaw_y)
    t[6] = (int)(W / 2.0 + h_jaw_y[1]);
    t[7] = rt;
    for (int n = 0; n <= n_pairs; ++n)
        t[kMlcHeader + n] = h_bound[n];
    for (int n = 0; n < n_pairs; ++n) {
        t[kMlcHeader + n_pairs + 1 + n] = h_a[n];
        t[kMlcHeader + 2 * n_pairs + 1 + n] = h_b[n];
    }
    // Row table: the bank rows of conehead/block.py:130-196 in raster rows
    // -- outside the bank or the Y jaws closed, otherwise the leaf pair
    // whose rows [bound[n] + shift, + width) hold the row (the last pair
    // starting at or before it), or open when none does.
    const int shift = abs(h_bound[0]);
    for (int row = 0; row < kMlcRows; ++row) {
        int32_t* e = &t[rt + 4 * row];
        const int rr = row - t[2];
        if (rr < 0 || rr >= W || rr < t[5] || rr >= t[6]) {
            e[0] = kRowClosed;
            continue;
        }
        e[0] = kRowOpen;
        if (h_bound[0] + shift > rr)
            continue;
        int lo = 0;
        while (lo + 1 < n_pairs && h_bound[lo + 1] + shift <= rr)
            ++lo;
        const int rmin = h_bound[lo] + shift;
        const int w = abs(h_bound[lo + 1] - h_bound[lo]);
        if (rr < rmin + w) {
            e[0] = kRowLeaf;
            e[1] = 2000 + h_a[lo];
            e[2] = 2000 + h_b[lo];
            e[3] = ((rr - rmin) << 16) | w;
        }
    }
    if (h_bbox_out) {
        // Every nonzero cell lies inside the bank rows, the Y-jaw rows and
        // the X-jaw columns (leaf edge rows transmit under the whole leaf).
        int r0 = t[2] + (t[5] > 0 ? t[5] : 0);
        int r1 = t[2] + (t[6] < W ? t[6] : W);
        int c0 = 2000 + h_jaw_x[0], c1 = 2000 + h_jaw_x[1];
        h_bbox_out[0] = r0 < 0 ? 0 : r0;
        h_bbox_out[1] = r1 > 4000 ? 4000 : r1;
        h_bbox_out[2] = c0 < 0 ? 0 : c0;
        h_bbox_out[3] = c1 > 4000 ? 4000 : c1;
    }
    CH_HIP(hipMemcpy(d_table, t.data(), t.size() * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    return CH_OK;
}

ch_status ch_raster_mlc(float* d_raster, const int32_t* d_table,
                        void* stream)
{
    CH_REQUIRE(d_raster && d_table, "ch_raster_mlc: null pointer");
    hipLaunchKernelGGL(mlc_kernel, dim3(62500), dim3(256), 0,
                       (hipStream_t)stream, d_raster, d_table);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

}  // extern "C"
