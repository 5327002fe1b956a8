// aperture.h -- beam aperture transmission at a raster cell (row i, col j)
// of the 4000 x 4000, 0.1 mm grid of conehead/block.py, evaluated from a
// compact description.  The raster builder (raster.hip) and the TERMA hit
// test (terma.hip) call the same functions, so the analytic lookups and the
// materialised raster agree bit for bit.
#pragma once

#include "ch_internal.h"

namespace ch {

// Leaf transmission table of conehead/block.py:155-164 at one element:
// 1 - area, area = ones with edge rows 0.2/0.5/0.75 (the rows counted from
// the far edge are assigned last and win) and the last 15 columns scaled
// by numpy.linspace(1, 0, 15) (arange * step + start, last point exact).
// Compiled without contraction.
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

// Leaf-table layout (int32): [0] n_pairs, [1] W (bank width), [2] roff
// (raster row of bank row 0), [3] jx0, [4] jx1 (column jaws, tenths of mm),
// [5] jy0, [6] jy1 (bank-row jaws), [7] start of the row table, then
// bound[n_pairs + 1], ends of bank A [n_pairs], ends of bank B [n_pairs],
// and at [7] (16-byte aligned) the row table: per raster row 4 int32
// {kind, column of bank A's tip, column of bank B's tip, (row within the
// leaf << 16) | leaf width}, kind 0 = closed (outside the bank or the
// Y jaws), 1 = open (no leaf), 2 = covered by a leaf pair.
constexpr int kMlcHeader = 8;
constexpr int kMlcRows = 4000;
constexpr int kRowClosed = 0, kRowOpen = 1, kRowLeaf = 2;

__host__ __device__ __forceinline__ int mlc_row_table_start(int n_pairs)
{
    return (kMlcHeader + 3 * n_pairs + 1 + 3) & ~3;
}

// Leaf bank and jaws of conehead/block.py:130-196 (bank centred on raster
// row 2000; rows outside the bank are blocked) at one cell: the row's
// entry of the row table (one 16-byte load) gives its leaf pair's tips and
// the row's place in the leaf; the leaf that contains a row is the one
// whose boundaries bracket it (leaves are contiguous and ordered).
__device__ __forceinline__ float mlc_value(const int32_t* __restrict__ t,
                                           int row, int col)
{
    const int4 e = reinterpret_cast<const int4*>(t + t[7])[row];
    if (e.x == kRowClosed)
        return 0.0f;
    float v = 1.0f;
    if (e.x == kRowLeaf) {
        const int w = e.w & 0xffff, rl = e.w >> 16;
        const int cmax = e.y;                     // bank A, :187-188
        if (col < cmax)
            v = leaf_value(rl, col, w, cmax);
        const int cmin = e.z;                     // bank B, :189-190
        const int h = 3999 - cmin;
        if (col >= cmin && col < 3999)
            v = leaf_value(w - 1 - rl, h - 1 - (col - cmin), w, h);
    }
    if (col < 2000 + t[3] || col >= 2000 + t[4])  // jaws, :193-196
        v = 0.0f;
    return v;
}

// Transmission of every cell of the block [i0, i1] x [j0, j1] (raster
// rows and columns, inside the raster) when it is one constant, else -1.
// A leaf table is constant on rectangles that avoid the bank edge, leaf
// edge rows, tips, tapers and jaw edges: the open gap of one leaf pair, or
// anything outside the bank or the jaws (closed).
__device__ __forceinline__ float mlc_block_value(const int32_t* __restrict__ t,
                                                 int i0, int i1, int j0,
                                                 int j1)
{
    const int n_pairs = t[0], W = t[1];
    const int r0 = i0 - t[2], r1 = i1 - t[2];
    if (r1 < 0 || r0 >= W)
        return 0.0f;  // outside the bank rows
    if (j1 < 2000 + t[3] || j0 >= 2000 + t[4] || r1 < t[5] || r0 >= t[6])
        return 0.0f;  // behind a jaw
    if (r0 < 0 || r1 >= W || r0 < t[5] || r1 >= t[6] || j0 < 2000 + t[3] ||
        j1 >= 2000 + t[4])
        return -1.0f;  // straddles a bank or jaw edge
    const int32_t* bound = t + kMlcHeader;
    const int32_t* ea = bound + n_pairs + 1;
    const int32_t* eb = ea + n_pairs;
    const int shift = abs(bound[0]);
    if (bound[0] + shift > r0)
        return -1.0f;
    int lo = 0, hi = n_pairs - 1;  // the leaf pair holding row r0
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bound[mid] + shift <= r0)
            lo = mid;
        else
            hi = mid - 1;
    }
    // Walk the leaf pairs the rows r0..r1 cross (pairs are contiguous):
    // the block is open (value 1, mlc_value leaves v = 1 there) when its
    // columns lie in the gap between the tips of every one of them.
    for (int p = lo; p < n_pairs; ++p) {
        const int rmin = bound[p] + shift;
        const int w = abs(bound[p + 1] - bound[p]);
        if (rmin > r1)
            return 1.0f;          // every row of the block was covered
        if (!(j0 >= 2000 + ea[p] && j1 < 2000 + eb[p]))
            return -1.0f;         // a tip or a leaf inside the block
        if (r1 < rmin + w)
            return 1.0f;
        if (p + 1 < n_pairs && bound[p + 1] + shift != rmin + w)
            return -1.0f;         // not contiguous: rows between pairs
    }
    return -1.0f;                 // rows beyond the last pair
}

__device__ __forceinline__ float rects_block_value(const int32_t* __restrict__ t,
                                                   int n, int i0, int i1,
                                                   int j0, int j1)
{
    bool any = false;
    for (int k = 0; k < n; ++k) {
        const int a = t[4 * k], b = t[4 * k + 1], c = t[4 * k + 2],
                  d = t[4 * k + 3];
        if (i0 >= a && i1 < b && j0 >= c && j1 < d)
            return 1.0f;  // inside one rectangle
        if (i1 >= a && i0 < b && j1 >= c && j0 < d)
            any = true;   // touches this rectangle
    }
    return any ? -1.0f : 0.0f;
}

// Union of half-open rectangles [r0, r1) x [c0, c1) of ones
// (Block.set_square, conehead/block.py:73-77), layout n x 4.
__device__ __forceinline__ float rects_value(const int32_t* __restrict__ t,
                                             int n, int row, int col)
{
    for (int k = 0; k < n; ++k)
        if (row >= t[4 * k] && row < t[4 * k + 1] && col >= t[4 * k + 2] &&
            col < t[4 * k + 3])
            return 1.0f;
    return 0.0f;
}

}  // namespace ch
