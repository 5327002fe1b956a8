// ccc.hip -- collapsed-cone convolution as a deterministic stencil gather.
//
// dose(v) = sum_c sum_m k_{c,m} * T(v - o_{c,m}) over sources v - o inside
// the grid (T is zero outside its support box, so sources are clipped to
// the box).  The gather reads a padded copy of T: every z-row carries kZPad
// zeros on both ends and every x-plane kYPad zero rows on both sides.
//
// One wavefront owns a 1 x TY x TZ tile of outputs (64 lanes, z fastest).
// The sources are clipped to x-slab support boxes of T (plane_box_kernel
// and slab_kernel below, one record per group of x planes).  For a box and
// a cone, x is uniform over the tile, so the box clips the tap list to one
// range from the per-axis first-step tables; along y and z the per-voxel
// ranges shift monotonically, and the union over the tile is taken from its
// end points.  Every tap of the union is applied to every lane without a
// predicate: a lane whose source falls outside the box along y or z reads a
// zero (T is zero there within the slab's planes, or it is padding, which
// is at least one tile wide), and a plane outside the grid reads zero
// through the range-checked buffer load.  Tap offsets and increments are
// wave-uniform scalar loads; T loads are TY runs of TZ contiguous floats.
// No atomics: every output is written once, in a fixed order, so results
// are bitwise reproducible alone or inside any batch.
#include <algorithm>
#include <climits>
#include <mutex>
#include <vector>

#include "ch_internal.h"

namespace ch {

// Constant-address-space views: wave-uniform loads through them become
// scalar (SMEM) loads; the data is not written while a kernel runs.
typedef __attribute__((address_space(4))) const int32_t c_i32;
typedef __attribute__((address_space(4))) const float c_f32;
typedef __attribute__((address_space(4))) const ConeMeta c_meta;
typedef __attribute__((address_space(4))) const ConvJob c_job;

template <typename T>
__device__ __forceinline__ auto as_const(const T* p)
{
    return (__attribute__((address_space(4))) const T*)(p);
}

// Field-wise copy out of the constant address space (a struct copy
// constructor cannot bind to an address_space(4) object).
__device__ __forceinline__ ConvJob load_job(const c_job* p)
{
    ConvJob j;
    j.terma = p->terma;
    j.dose = p->dose;
    j.box = p->box;
    j.lin = p->lin;
    j.k = p->k;
    j.ftab = p->ftab;
    j.ftT = p->ftT;
    j.slabs = p->slabs;
    j.meta = p->meta;
    j.off = p->off;
    j.lin2 = p->lin2;
    j.X = p->X;
    j.Y = p->Y;
    j.Z = p->Z;
    j.n_cones = p->n_cones;
    j.accumulate = p->accumulate;
    j.ntx = p->ntx;
    j.nty = p->nty;
    j.ntz = p->ntz;
    j.tile_begin = p->tile_begin;
    return j;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Tap range [a, b) of one axis for a uniform coordinate v and box [lo, hi]:
// taps whose source coordinate v - o lies in the box.  Out-of-range
// arguments clamp to the ends of the table, which yields empty ranges and
// keeps a and b monotone in v.
__device__ __forceinline__ void urange(int v, int lo, int hi, int step,
                                       const c_i32* f, int f_len, int& a,
                                       int& b)
{
    int L, U;
    if (step > 0) {
        L = v - hi;
        U = v - lo;
    } else {
        L = lo - v;
        U = hi - v;
    }
    const int last = f_len - 1;  // sentinel entry, = n_taps
    L = L < 0 ? 0 : (L > last ? last : L);
    U = U + 1 < 0 ? 0 : (U + 1 > last ? last : U + 1);
    a = f[L];
    b = f[U];
}


// Sum of one chunk of N taps (kUnroll unless stated) into a cone line's
// accumulator:
//   ACC 0  Kahan-compensated per tap,
//   ACC 1  plain FMA along the line (the line total is compensated),
//   ACC 2  plain sum of the chunk, Kahan-compensated across chunks.
template <int ACC, int N = kUnroll>
__device__ __forceinline__ void accumulate_chunk(float& acc, float& ac,
                                                 const float* kc,
                                                 const float* v)
{
    if (ACC == 0) {
#pragma unroll
        for (int u = 0; u < N; ++u)
            kahan_add(acc, ac, kc[u] * v[u]);
    } else if (ACC == 1) {
#pragma unroll
        for (int u = 0; u < N; ++u)
            acc = fmaf(kc[u], v[u], acc);
    } else {
        float part = kc[0] * v[0];
#pragma unroll
        for (int u = 1; u < N; ++u)
            part = fmaf(kc[u], v[u], part);
        kahan_add(acc, ac, part);
    }
}

// One wavefront owns a 1 x TY x TZ output tile (TY * TZ = 64 lanes, z
// fastest).  The tile's union tap range of 64 cones at a time is computed
// in parallel, one lane per cone, and read back with readlane, so the
// scalar unit only walks taps.  Taps are applied in whole chunks of
// kUnroll: a tap past the union end reads a zero (its source lies outside
// the support box along some axis for every lane, within the y/z padding,
// or outside the grid along x, where the range-checked buffer load returns
// 0), so no chunk needs a mask.
template <int TY, int TZ, int ACC, typename OutT>
__device__ __forceinline__ void gather_tile(const ConvJob& j, int tile,
                                            int lane)
{
    static_assert(TY * TZ == 64 && TY <= kYPad && TZ + kUnroll <= kZPad,
                  "tile");
    const int tz = tile % j.ntz;
    const int r = tile / j.ntz;
    const int ty = r % j.nty;
    const int x = r / j.nty;
    const int y0 = ty * TY, z0 = tz * TZ;
    const int yl = min(y0 + TY - 1, j.Y - 1);  // last valid y of the tile
    const int zl = min(z0 + TZ - 1, j.Z - 1);  // last valid z of the tile
    const int y = y0 + lane / TZ;
    const int z = z0 + lane % TZ;
    const int Zp = j.Z + 2 * kZPad;
    const int Yp = j.Y + 2 * kYPad;

    // byte offset of this lane's own voxel in the padded T allocation
    const uint32_t base4 =
        (uint32_t)(((x * Yp + y + kYPad) * Zp + z + kZPad) * 4);
    const float* alloc = j.terma - ((size_t)kYPad * Zp + kZPad);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)alloc, (short)0, j.X * Yp * Zp * 4, kBufferFlags);
    const c_i32* lin = as_const(j.lin);
    const c_f32* kk = as_const(j.k);
    const int n = j.n_cones;
    const int32_t* f0 = j.ftT;
    const int32_t* f1 = j.ftT + ftab_t_base(1, j.X, j.Y, n);
    const int32_t* f2 = j.ftT + ftab_t_base(2, j.X, j.Y, n);
    // support boxes: the x-slab boxes of T, or the one box (or the grid)
    const c_i32* sl = j.slabs ? as_const(j.slabs) : nullptr;
    const int ns = sl ? sl[0] : 1;

    float sum = 0.0f, comp = 0.0f;
    for (int cb = 0; cb < n; cb += 64) {
        for (int s = 0; s < ns; ++s) {
            int lo0 = 0, lo1 = 0, lo2 = 0;
            int hi0 = j.X - 1, hi1 = j.Y - 1, hi2 = j.Z - 1;
            const c_i32* b = sl ? sl + kSlabStride * (1 + s)
                                : (j.box ? as_const(j.box) : nullptr);
            if (b) {
                lo0 = max(lo0, b[0]);
                lo1 = max(lo1, b[1]);
                lo2 = max(lo2, b[2]);
                hi0 = min(hi0, b[3]);
                hi1 = min(hi1, b[4]);
                hi2 = min(hi2, b[5]);
            }
            if (lo0 > hi0 || lo1 > hi1 || lo2 > hi2)
                continue;
            // lane l: the tile's union range of cone cb + l within the box
            int r_lo = 0, r_hi = 0, r_tb = 0;
            const int c = cb + lane;
            if (c < n) {
                const ConeMeta* m = j.meta + c;
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
                // union of the tile's per-voxel ranges (each axis is
                // monotone, so its extremes sit at the tile's end points)
                r_lo = max(max(ax, min(ay1, ay2)), min(az1, az2));
                r_hi = min(min(bx, max(by1, by2)), max(bz1, bz2));
                r_tb = m->tap_begin;
            }
            // cones with taps in this box
            uint64_t live = __ballot(r_lo < r_hi);
            while (live) {
                const int i = __builtin_ctzll(live);
                live &= live - 1;
                const int ulo = __builtin_amdgcn_readlane(r_lo, i);
                const int uhi = __builtin_amdgcn_readlane(r_hi, i);
                const int tb = __builtin_amdgcn_readlane(r_tb, i);
                const c_i32* ml = lin + tb;
                const c_f32* mk = kk + tb;
                float acc = 0.0f, ac = 0.0f;
                int t = ulo;
                for (; t + kUnroll <= uhi; t += kUnroll) {
                    float v[kUnroll], kc[kUnroll];
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        kc[u] = mk[t + u];
                        v[u] = __int_as_float(
                            __builtin_amdgcn_raw_buffer_load_b32(
                                rsrc, (int)(base4 - (uint32_t)ml[t + u]), 0,
                                0));
                    }
                    accumulate_chunk<ACC>(acc, ac, kc, v);
                }
                // the rest in chunks of 4, the last one partial: taps past
                // the range end get a zero weight (their sources may lie in
                // the next slab's box); the loads stay in bounds (tap
                // padding, y/z padding, x range check)
                for (; t < uhi; t += 4) {
                    float v[4], kc[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        kc[u] = t + u < uhi ? mk[t + u] : 0.0f;
                        v[u] = __int_as_float(
                            __builtin_amdgcn_raw_buffer_load_b32(
                                rsrc, (int)(base4 - (uint32_t)ml[t + u]), 0,
                                0));
                    }
                    accumulate_chunk<ACC, 4>(acc, ac, kc, v);
                }
                kahan_add(sum, comp, acc);
            }
        }
    }
    if (y <= yl && z <= zl) {
        OutT* d = reinterpret_cast<OutT*>(j.dose) +
                  ((size_t)x * j.Y + y) * j.Z + z;
        const OutT prev = j.accumulate ? *d : (OutT)0;
        *d = prev + (OutT)sum;
    }
}

// Wavefronts (tiles) per workgroup of the direct gather: consecutive tiles
// along z, whose rows meet at shared L1 lines.  Measured on MI355X: 8 for a
// single grid (1.13 -> 1.09 ms at 128^3), 4 for batches (a 64-beam batch
// took 96 ms with 8 against 79 ms with 4).
constexpr int kGatherWaves = 8;
constexpr int kBatchWaves = 4;

template <int TY, int TZ, int ACC, typename OutT>
__global__ __launch_bounds__(64 * kGatherWaves) void gather_single(
    ConvJob j, int n_tiles)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * kGatherWaves + wave;
    if (tile >= n_tiles)
        return;
    gather_tile<TY, TZ, ACC, OutT>(j, tile, threadIdx.x & 63);
}

template <int TY, int TZ, int ACC, typename OutT>
__global__ __launch_bounds__(64 * kBatchWaves) void gather_batch(
    const ConvJob* jobs, int n_jobs, int n_tiles)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * kBatchWaves + wave;
    if (tile >= n_tiles)
        return;
    const c_job* J = as_const(jobs);
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J[mid].tile_begin <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ConvJob jb = load_job(J + lo);
    gather_tile<TY, TZ, ACC, OutT>(jb, tile - jb.tile_begin,
                                   threadIdx.x & 63);
}

// ------------------------------------------------------ staged gather ----
//
// One 256-thread workgroup owns a kSBX x kSBY x kSBZ output block: wave w
// owns x = bx + w and lane (j, k) the outputs (y = by + j + 4 r, z = bz + k),
// r = 0..3.  Per cone, the block's union tap range is cut into chunks of
// s_ch taps.  The TERMA a chunk reads is the block shifted by the chunk's
// offsets, a box of at most s_ext voxels.  It is staged into LDS one row
// (s_ext[2] <= kSPitch floats along z) per range-checked buffer load that
// writes LDS directly (x outside the grid reads 0; y and z stay inside the
// padding), then every lane applies the chunk's taps from LDS.  The box
// layout is fixed per cone (rows of kSPitch floats, planes of s_ext[1]
// rows), so tap m reads LDS at (lane base + chunk base - lin2[m]).  With
// rows of 48 floats the four y rows a wavefront reads per tap fall in
// disjoint bank ranges.
// Sums: plain FMA over a chunk's taps, Kahan-compensated across chunks.

template <typename OutT>
__device__ __forceinline__ void staged_block(const ConvJob& j, int blk,
                                             float* __restrict__ lds)
{
    const int bz = (blk % j.ntz) * kSBZ;
    const int rr = blk / j.ntz;
    const int by = (rr % j.nty) * kSBY;
    const int bx = (rr / j.nty) * kSBX;
    const int xl = min(bx + kSBX - 1, j.X - 1);
    const int yl = min(by + kSBY - 1, j.Y - 1);
    const int zl = min(bz + kSBZ - 1, j.Z - 1);
    const int tid = threadIdx.x;
    const int li = tid >> 6, lj = (tid >> 4) & 3, lk = tid & 15;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;

    int lo0 = 0, lo1 = 0, lo2 = 0;
    int hi0 = j.X - 1, hi1 = j.Y - 1, hi2 = j.Z - 1;
    if (j.box) {
        const c_i32* b = as_const(j.box);
        lo0 = max(lo0, b[0]);
        lo1 = max(lo1, b[1]);
        lo2 = max(lo2, b[2]);
        hi0 = min(hi0, b[3]);
        hi1 = min(hi1, b[4]);
        hi2 = min(hi2, b[5]);
    }
    const int Zp = j.Z + 2 * kZPad;
    const int Yp = j.Y + 2 * kYPad;
    const float* alloc = j.terma - ((size_t)kYPad * Zp + kZPad);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)alloc, (short)0, j.X * Yp * Zp * 4, kBufferFlags);

    float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float comp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2) {
        const c_meta* meta = as_const(j.meta);
        const c_i32* ftab = as_const(j.ftab);
        for (int c = 0; c < j.n_cones; ++c) {
            const c_meta* m = meta + c;
            int ax1, bx1, ax2, bx2, ay1, by1, ay2, by2, az1, bz1, az2, bz2;
            urange(bx, lo0, hi0, m->step[0], ftab + m->f_begin[0],
                   m->f_len[0], ax1, bx1);
            urange(xl, lo0, hi0, m->step[0], ftab + m->f_begin[0],
                   m->f_len[0], ax2, bx2);
            urange(by, lo1, hi1, m->step[1], ftab + m->f_begin[1],
                   m->f_len[1], ay1, by1);
            urange(yl, lo1, hi1, m->step[1], ftab + m->f_begin[1],
                   m->f_len[1], ay2, by2);
            urange(bz, lo2, hi2, m->step[2], ftab + m->f_begin[2],
                   m->f_len[2], az1, bz1);
            urange(zl, lo2, hi2, m->step[2], ftab + m->f_begin[2],
                   m->f_len[2], az2, bz2);
            // union of the block's per-voxel ranges (monotone per axis)
            const int ulo =
                max(max(min(ax1, ax2), min(ay1, ay2)), min(az1, az2));
            const int uhi =
                min(min(max(bx1, bx2), max(by1, by2)), max(bz1, bz2));
            if (ulo >= uhi)
                continue;
            const int ch = m->s_ch;
            const int ex = m->s_ext[0], ey = m->s_ext[1], ez = m->s_ext[2];
            const int px = ey * kSPitch;
            const int lane_base = li * px + lj * kSPitch + lk;
            const c_i32* off = as_const(j.off) + 3 * m->tap_begin;
            const c_i32* l2 = as_const(j.lin2) + m->tap_begin;
            const c_f32* kk = as_const(j.k) + m->tap_begin;
            for (int m0 = ulo; m0 < uhi; m0 += ch) {
                const int m1 = min(m0 + ch, uhi);
                // offsets are monotone per axis: the chunk's largest offset
                // sits at one of its ends
                const int P0 = max(off[3 * m0], off[3 * (m1 - 1)]);
                const int P1 = max(off[3 * m0 + 1], off[3 * (m1 - 1) + 1]);
                const int P2 = max(off[3 * m0 + 2], off[3 * (m1 - 1) + 2]);
                // element index of the box's first voxel in the padded T
                const int origin =
                    ((bx - P0) * Yp + (by - P1) + kYPad) * Zp + (bz - P2) +
                    kZPad;
                __syncthreads();  // every wave is done with the last chunk
                // wave w stages the box planes x = w, w + 4, ...: one row of
                // ez floats per instruction, straight into LDS (lanes >= ez
                // idle; planes outside the grid read zeros)
                if (lane < ez) {
                    for (int xi = w; xi < ex; xi += 4) {
                        const int g = origin + xi * Yp * Zp;
                        float* row = lds + xi * px;
                        for (int yi = 0; yi < ey; ++yi)
                            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                                rsrc,
                                (__attribute__((address_space(3))) void*)(
                                    row + yi * kSPitch),
                                4, (g + yi * Zp + lane) * 4, 0, 0, 0);
                    }
                }
                __builtin_amdgcn_s_waitcnt(0);  // the rows have landed
                __syncthreads();
                // byte offsets: lin2 holds 4 x the element offset
                const int base =
                    4 * (lane_base + P0 * px + P1 * kSPitch + P2);
                const char* L = reinterpret_cast<const char*>(lds);
                float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
                int t = m0;
                for (; t + 8 <= m1; t += 8) {
                    float v[8][4], kv[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float* q = reinterpret_cast<const float*>(
                            L + (base - l2[t + u]));
                        v[u][0] = q[0];
                        v[u][1] = q[4 * kSPitch];
                        v[u][2] = q[8 * kSPitch];
                        v[u][3] = q[12 * kSPitch];
                        kv[u] = kk[t + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        a0 = fmaf(kv[u], v[u][0], a0);
                        a1 = fmaf(kv[u], v[u][1], a1);
                        a2 = fmaf(kv[u], v[u][2], a2);
                        a3 = fmaf(kv[u], v[u][3], a3);
                    }
                }
                for (; t < m1; ++t) {
                    const float* q =
                        reinterpret_cast<const float*>(L + (base - l2[t]));
                    const float kv = kk[t];
                    a0 = fmaf(kv, q[0], a0);
                    a1 = fmaf(kv, q[4 * kSPitch], a1);
                    a2 = fmaf(kv, q[8 * kSPitch], a2);
                    a3 = fmaf(kv, q[12 * kSPitch], a3);
                }
                kahan_add(sum[0], comp[0], a0);
                kahan_add(sum[1], comp[1], a1);
                kahan_add(sum[2], comp[2], a2);
                kahan_add(sum[3], comp[3], a3);
            }
        }
    }
    const int x = bx + li, z = bz + lk;
    if (x < j.X && z < j.Z) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int y = by + lj + 4 * r;
            if (y < j.Y) {
                OutT* d = reinterpret_cast<OutT*>(j.dose) +
                          ((size_t)x * j.Y + y) * j.Z + z;
                const OutT prev = j.accumulate ? *d : (OutT)0;
                *d = prev + (OutT)sum[r];
            }
        }
    }
}

template <typename OutT>
__global__ __launch_bounds__(256) void staged_single(ConvJob j, int n_blocks)
{
    __shared__ float lds[kSLds];
    if ((int)blockIdx.x >= n_blocks)
        return;
    staged_block<OutT>(j, blockIdx.x, lds);
}

template <typename OutT>
__global__ __launch_bounds__(256) void staged_batch(const ConvJob* jobs,
                                                    int n_jobs, int n_blocks)
{
    __shared__ float lds[kSLds];
    const int b = blockIdx.x;
    if (b >= n_blocks)
        return;
    const c_job* J = as_const(jobs);
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J[mid].tile_begin <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ConvJob jb = load_job(J + lo);
    staged_block<OutT>(jb, b - jb.tile_begin, lds);
}

// --------------------------------------------------- beam-group gather ----
//
// Up to kGroup beams on one plan share every tap: their padded TERMA grids
// are interleaved ([X][Yp][Zp][kGroup] floats), so one aligned 16-byte load
// per lane and tap feeds all of them, and the tap walk (ranges, offsets,
// increments) is paid once per group.  Tap ranges come from the union of
// the beams' support boxes; each beam's sum follows the direct gather's
// accumulation scheme.

struct GroupJob {
    const float* terma;        // interleaved padded TERMA, interior (0,0,0)
    float* dose[kGroup];
    const int32_t* box[kGroup];
    const int32_t* lin;
    const float* k;
    const int32_t* ftab;
    const ConeMeta* meta;
    int32_t X, Y, Z;
    int32_t n_cones;
    int32_t n_beams;
    int32_t accumulate;
    int32_t nty, ntz;
    int32_t tile_begin;
};
typedef __attribute__((address_space(4))) const GroupJob c_gjob;

__device__ __forceinline__ GroupJob load_gjob(const c_gjob* p)
{
    GroupJob j;
    j.terma = p->terma;
#pragma unroll
    for (int b = 0; b < kGroup; ++b) {
        j.dose[b] = p->dose[b];
        j.box[b] = p->box[b];
    }
    j.lin = p->lin;
    j.k = p->k;
    j.ftab = p->ftab;
    j.meta = p->meta;
    j.X = p->X;
    j.Y = p->Y;
    j.Z = p->Z;
    j.n_cones = p->n_cones;
    j.n_beams = p->n_beams;
    j.accumulate = p->accumulate;
    j.nty = p->nty;
    j.ntz = p->ntz;
    j.tile_begin = p->tile_begin;
    return j;
}

// N taps from t of one cone for every beam of a group; with `masked`,
// taps at or past uhi read offset 0 with increment 0.
template <int ACC, int N>
__device__ __forceinline__ void group_chunk(int t, int uhi, bool masked,
                                            const c_i32* ml, const c_f32* mk,
                                            const char* Tb, uint32_t base16,
                                            float* acc, float* ac)
{
    float4 v[N];
    float kc[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
        const bool in = !masked || t + u < uhi;
        const uint32_t o = (uint32_t)ml[t + u] << 2;
        v[u] = *reinterpret_cast<const float4*>(Tb + (base16 - (in ? o : 0u)));
        kc[u] = in ? mk[t + u] : 0.0f;
    }
#pragma unroll
    for (int b = 0; b < kGroup; ++b) {
        float vb[N];
#pragma unroll
        for (int u = 0; u < N; ++u)
            vb[u] = b == 0 ? v[u].x
                           : (b == 1 ? v[u].y : (b == 2 ? v[u].z : v[u].w));
        accumulate_chunk<ACC, N>(acc[b], ac[b], kc, vb);
    }
}

template <int TY, int TZ, int ACC>
__device__ __forceinline__ void gather_tile_group(const GroupJob& j, int tile,
                                                  int lane)
{
    static_assert(TY * TZ == 64 && TY <= kYPad && TZ <= kZPad, "tile");
    const int tz = tile % j.ntz;
    const int r = tile / j.ntz;
    const int ty = r % j.nty;
    const int x = r / j.nty;
    const int y0 = ty * TY, z0 = tz * TZ;
    const int yl = min(y0 + TY - 1, j.Y - 1);
    const int zl = min(z0 + TZ - 1, j.Z - 1);
    const int y = y0 + lane / TZ;
    const int z = z0 + lane % TZ;
    const int Zp = j.Z + 2 * kZPad;
    const int Yp = j.Y + 2 * kYPad;

    // union of the beams' support boxes
    int lo0 = INT_MAX, lo1 = INT_MAX, lo2 = INT_MAX;
    int hi0 = INT_MIN, hi1 = INT_MIN, hi2 = INT_MIN;
#pragma unroll
    for (int b = 0; b < kGroup; ++b) {
        if (b >= j.n_beams)
            continue;
        int l0 = 0, l1 = 0, l2 = 0, h0 = j.X - 1, h1 = j.Y - 1, h2 = j.Z - 1;
        if (j.box[b]) {
            const c_i32* bb = as_const(j.box[b]);
            l0 = max(l0, bb[0]);
            l1 = max(l1, bb[1]);
            l2 = max(l2, bb[2]);
            h0 = min(h0, bb[3]);
            h1 = min(h1, bb[4]);
            h2 = min(h2, bb[5]);
        }
        if (l0 <= h0 && l1 <= h1 && l2 <= h2) {
            lo0 = min(lo0, l0);
            lo1 = min(lo1, l1);
            lo2 = min(lo2, l2);
            hi0 = max(hi0, h0);
            hi1 = max(hi1, h1);
            hi2 = max(hi2, h2);
        }
    }
    // byte offset of this lane's voxel (16 bytes per voxel) in the padded
    // interleaved allocation
    const uint32_t base16 =
        (uint32_t)(((x * Yp + y + kYPad) * Zp + z + kZPad)) * 16u;
    const char* __restrict__ Tb = reinterpret_cast<const char*>(j.terma) -
                                  (size_t)(kYPad * Zp + kZPad) * 16;
    const c_meta* meta = as_const(j.meta);
    const c_i32* ftab = as_const(j.ftab);
    const c_i32* lin = as_const(j.lin);
    const c_f32* kk = as_const(j.k);

    float sum[kGroup] = {}, comp[kGroup] = {};
    if (lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2) {
        for (int c = 0; c < j.n_cones; ++c) {
            const c_meta* m = meta + c;
            int ax, bx, ay1, by1, ay2, by2, az1, bz1, az2, bz2;
            urange(x, lo0, hi0, m->step[0], ftab + m->f_begin[0],
                   m->f_len[0], ax, bx);
            urange(y0, lo1, hi1, m->step[1], ftab + m->f_begin[1],
                   m->f_len[1], ay1, by1);
            urange(z0, lo2, hi2, m->step[2], ftab + m->f_begin[2],
                   m->f_len[2], az1, bz1);
            if (TY > 1)
                urange(yl, lo1, hi1, m->step[1], ftab + m->f_begin[1],
                       m->f_len[1], ay2, by2);
            else {
                ay2 = ay1;
                by2 = by1;
            }
            urange(zl, lo2, hi2, m->step[2], ftab + m->f_begin[2],
                   m->f_len[2], az2, bz2);
            const int ulo = max(max(ax, min(ay1, ay2)), min(az1, az2));
            const int uhi = min(min(bx, max(by1, by2)), max(bz1, bz2));
            if (ulo >= uhi)
                continue;
            const c_i32* ml = lin + m->tap_begin;
            const c_f32* mk = kk + m->tap_begin;
            float acc[kGroup] = {}, ac[kGroup] = {};
            // whole chunks of kUnroll, then chunks of 4 (the direct
            // gather's chunking, so a one-beam group sums like it); taps
            // past uhi read offset 0 with increment 0
            int t = ulo;
            for (; t + kUnroll <= uhi; t += kUnroll)
                group_chunk<ACC, kUnroll>(t, uhi, false, ml, mk, Tb, base16,
                                          acc, ac);
            for (; t < uhi; t += 4)
                group_chunk<ACC, 4>(t, uhi, true, ml, mk, Tb, base16, acc,
                                    ac);
#pragma unroll
            for (int b = 0; b < kGroup; ++b)
                kahan_add(sum[b], comp[b], acc[b]);
        }
    }
    if (y <= yl && z <= zl) {
        const size_t o = ((size_t)x * j.Y + y) * j.Z + z;
#pragma unroll
        for (int b = 0; b < kGroup; ++b)
            if (b < j.n_beams) {
                float* d = j.dose[b] + o;
                const float prev = j.accumulate ? *d : 0.0f;
                *d = prev + sum[b];
            }
    }
}

template <int TY, int TZ, int ACC>
__global__ __launch_bounds__(256) void gather_group(const GroupJob* jobs,
                                                    int n_jobs, int n_tiles)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= n_tiles)
        return;
    const c_gjob* J = (const c_gjob*)as_const(jobs);
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (J[mid].tile_begin <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    const GroupJob jb = load_gjob(J + lo);
    gather_tile_group<TY, TZ, ACC>(jb, tile - jb.tile_begin,
                                   threadIdx.x & 63);
}

// Up to kGroup TERMA grids -> one padded interleaved grid
// [X][Y + 2 kYPad][Z + 2 kZPad][kGroup] with zero borders and zero slots
// for absent beams.
__global__ __launch_bounds__(256) void pad_group_kernel(
    const float* __restrict__ T0, const float* __restrict__ T1,
    const float* __restrict__ T2, const float* __restrict__ T3,
    float4* __restrict__ P, int X, int Y, int Z)
{
    const int Zp = Z + 2 * kZPad, Yp = Y + 2 * kYPad;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)X * Yp * Zp)
        return;
    const int zz = (int)(i % Zp) - kZPad;
    const int64_t r = i / Zp;
    const int yy = (int)(r % Yp) - kYPad;
    const int64_t x = r / Yp;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (zz >= 0 && zz < Z && yy >= 0 && yy < Y) {
        const int64_t s = (x * Y + yy) * Z + zz;
        v.x = T0[s];
        if (T1)
            v.y = T1[s];
        if (T2)
            v.z = T2[s];
        if (T3)
            v.w = T3[s];
    }
    P[i] = v;
}

// T [X][Y][Z] -> padded [X][Y + 2 kYPad][Z + 2 kZPad] with zero borders.
__global__ __launch_bounds__(256) void pad_kernel(const float* __restrict__ T,
                                                  float* __restrict__ P,
                                                  int X, int Y, int Z)
{
    const int Zp = Z + 2 * kZPad, Yp = Y + 2 * kYPad;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (int64_t)X * Yp * Zp)
        return;
    const int zz = (int)(i % Zp) - kZPad;
    const int64_t r = i / Zp;
    const int yy = (int)(r % Yp) - kYPad;
    const int64_t x = r / Yp;
    P[i] = (zz >= 0 && zz < Z && yy >= 0 && yy < Y)
               ? T[(x * Y + yy) * Z + zz]
               : 0.0f;
}

// ------------------------------------------------------ slab boxes ----

// Extent of T != 0 in x plane blockIdx.x (within the caller's box when
// given): planes[4 p + 0..3] = min y, max y, min z, max z (min > max when
// the plane is empty).
__global__ __launch_bounds__(256) void plane_box_kernel(
    const float* __restrict__ T, int Y, int Z, const int32_t* box,
    int32_t* planes)
{
    const int p = blockIdx.x;
    const float* P = T + (size_t)p * Y * Z;
    int32_t v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    bool in_box = true;
    int ylo = 0, yhi = Y - 1, zlo = 0, zhi = Z - 1;
    if (box) {
        in_box = p >= box[0] && p <= box[3];
        ylo = max(ylo, box[1]);
        yhi = min(yhi, box[4]);
        zlo = max(zlo, box[2]);
        zhi = min(zhi, box[5]);
    }
    if (in_box)
        for (int y = ylo + (int)(threadIdx.x >> 6); y <= yhi; y += 4)
            for (int z = zlo + (int)(threadIdx.x & 63); z <= zhi; z += 64)
                if (P[(size_t)y * Z + z] != 0.0f) {
                    v[1] = min(v[1], y);
                    v[4] = max(v[4], y);
                    v[2] = min(v[2], z);
                    v[5] = max(v[5], z);
                }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v[1] = min(v[1], __shfl_xor(v[1], o, 64));
        v[2] = min(v[2], __shfl_xor(v[2], o, 64));
        v[4] = max(v[4], __shfl_xor(v[4], o, 64));
        v[5] = max(v[5], __shfl_xor(v[5], o, 64));
    }
    __shared__ int32_t part[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[w][0] = v[1];
        part[w][1] = v[4];
        part[w][2] = v[2];
        part[w][3] = v[5];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int a = threadIdx.x;  // 0, 2: minima; 1, 3: maxima
        int32_t r = part[0][a];
        for (int k = 1; k < 4; ++k)
            r = (a & 1) ? max(r, part[k][a]) : min(r, part[k][a]);
        planes[4 * p + a] = r;
    }
}

// Slab records from the plane extents (one block of 64 threads): lane s
// builds slab s, then lane 0 merges neighbours whose union box is at most
// 2 % larger than the two boxes together and writes the records.
__global__ __launch_bounds__(64) void slab_kernel(const int32_t* planes,
                                                  int X, int32_t* slabs)
{
    __shared__ int32_t raw[64][6];
    const int n_raw = (X + kSlabW - 1) / kSlabW;
    for (int s = threadIdx.x; s < n_raw; s += 64) {
        int32_t b[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN,
                        INT_MIN};
        for (int p = s * kSlabW; p < min(X, (s + 1) * kSlabW); ++p) {
            const int32_t* q = planes + 4 * p;
            if (q[0] > q[1])
                continue;
            b[0] = min(b[0], p);
            b[3] = max(b[3], p);
            b[1] = min(b[1], q[0]);
            b[4] = max(b[4], q[1]);
            b[2] = min(b[2], q[2]);
            b[5] = max(b[5], q[3]);
        }
        // slabs beyond the shared table (X > 64 kSlabW) are written in
        // place and merged below from global memory
        int32_t* dst = s < 64 ? raw[s] : slabs + kSlabStride * (1 + s);
        for (int a = 0; a < 6; ++a)
            dst[a] = b[a];
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    auto vol = [](const int32_t* b) {
        return (int64_t)(b[3] - b[0] + 1) * (b[4] - b[1] + 1) *
               (b[5] - b[2] + 1);
    };
    int n = 0;
    int32_t cur[6];
    for (int s = 0; s < n_raw; ++s) {
        const int32_t* b = s < 64 ? raw[s] : slabs + kSlabStride * (1 + s);
        int32_t nb[6];
        for (int a = 0; a < 6; ++a)
            nb[a] = b[a];
        if (nb[0] > nb[3])
            continue;  // no T in this slab
        if (n > 0) {
            int32_t u[6];
            for (int a = 0; a < 3; ++a) {
                u[a] = min(cur[a], nb[a]);
                u[3 + a] = max(cur[3 + a], nb[3 + a]);
            }
            if (50 * vol(u) <= 51 * (vol(cur) + vol(nb))) {
                for (int a = 0; a < 6; ++a)
                    cur[a] = u[a];
                continue;
            }
            for (int a = 0; a < 6; ++a)
                slabs[kSlabStride * n + a] = cur[a];  // record n of 1..
        }
        ++n;
        for (int a = 0; a < 6; ++a)
            cur[a] = nb[a];
    }
    if (n > 0)
        for (int a = 0; a < 6; ++a)
            slabs[kSlabStride * n + a] = cur[a];
    slabs[0] = n;
}

// ----------------------------------------------------------- dispatch ----

template <typename OutT>
struct Kernels {
    void (*single)(ConvJob, int);
    void (*batch)(const ConvJob*, int, int);
    // per launch of `single` and of `batch`: tiles per workgroup and its size
    int tiles_per_block, threads;
    int batch_tiles_per_block, batch_threads;
};

template <typename OutT>
static ch_status kernels_for(const ch_plan* p, Kernels<OutT>& k)
{
    if (p->semantics == CH_SEM_RADIOLOGICAL) {
        set_error("a CH_SEM_RADIOLOGICAL plan is used through "
                  "ch_convolve_rad");
        return CH_ERR_ARG;
    }
    const int tx = p->tile[0], ty = p->tile[1], tz = p->tile[2];
    if (tx == kSBX && ty == kSBY && tz == kSBZ) {
        k.single = staged_single<OutT>;
        k.batch = staged_batch<OutT>;
        k.tiles_per_block = k.batch_tiles_per_block = 1;
        k.threads = k.batch_threads = 256;
        return CH_OK;
    }
#define CH_PICK(TY, TZ, A)                                                  \
    if (tx == 1 && ty == TY && tz == TZ && p->accum_mode == A) {            \
        k.single = gather_single<TY, TZ, A, OutT>;                          \
        k.batch = gather_batch<TY, TZ, A, OutT>;                            \
        k.tiles_per_block = kGatherWaves;                                   \
        k.threads = 64 * kGatherWaves;                                      \
        k.batch_tiles_per_block = kBatchWaves;                              \
        k.batch_threads = 64 * kBatchWaves;                                 \
        return CH_OK;                                                       \
    }
#define CH_PICK3(TY, TZ) CH_PICK(TY, TZ, 0) CH_PICK(TY, TZ, 1) CH_PICK(TY, TZ, 2)
    CH_PICK3(1, 64)
    CH_PICK3(2, 32)
    CH_PICK3(4, 16)
    CH_PICK3(8, 8)
#undef CH_PICK3
#undef CH_PICK
    set_error("no gather kernel for tile %dx%dx%d, accumulation %d", tx, ty,
              tz, p->accum_mode);
    return CH_ERR_ARG;
}

// `padded` is the start of the padded allocation; the job points at its
// interior voxel (0, 0, 0).
static ConvJob make_job(const ch_plan* p, const float* padded, void* dose,
                        const int32_t* box, int32_t accumulate)
{
    const int Zp = p->shape[2] + 2 * kZPad;
    ConvJob j;
    j.terma = padded + (size_t)kYPad * Zp + kZPad;
    j.dose = dose;
    j.box = box;
    j.lin = p->d_lin;
    j.k = p->d_k;
    j.ftab = p->d_ftab;
    j.ftT = p->d_ftT;
    j.slabs = nullptr;
    j.meta = p->d_meta;
    j.off = p->d_off;
    j.lin2 = p->d_lin2;
    j.X = p->shape[0];
    j.Y = p->shape[1];
    j.Z = p->shape[2];
    j.n_cones = p->n_cones;
    j.accumulate = accumulate;
    j.ntx = (j.X + p->tile[0] - 1) / p->tile[0];
    j.nty = (j.Y + p->tile[1] - 1) / p->tile[1];
    j.ntz = (j.Z + p->tile[2] - 1) / p->tile[2];
    j.tile_begin = 0;
    return j;
}

static int64_t job_tiles(const ConvJob& j)
{
    return (int64_t)j.ntx * j.nty * j.ntz;
}

static size_t padded_elems(const ch_plan* p)
{
    return (size_t)p->shape[0] * (p->shape[1] + 2 * kYPad) *
           (p->shape[2] + 2 * kZPad);
}

// Per-stream workspaces: the padded TERMA copies and the batch job table
// (a byte buffer holding ConvJob or GroupJob records).  Kernels on one
// stream run in order, so a stream's workspace is reused by its next call
// without further synchronisation; the pinned job-table staging buffer is
// reused only after the event of its last copy.
struct Workspace {
    hipStream_t stream;
    float* pad;
    size_t pad_elems;
    void* d_tab;
    void* h_tab;
    size_t tab_bytes;
    hipEvent_t jobs_done;
};
static std::vector<Workspace*> g_ws;
static std::mutex g_ws_mu;

static ch_status workspace(hipStream_t s, size_t pad_elems, size_t tab_bytes,
                           Workspace** out)
{
    Workspace* w = nullptr;
    for (auto* x : g_ws)
        if (x->stream == s)
            w = x;
    if (!w) {
        w = new Workspace();
        w->stream = s;
        w->pad = nullptr;
        w->pad_elems = 0;
        w->d_tab = nullptr;
        w->h_tab = nullptr;
        w->tab_bytes = 0;
        CH_HIP(hipEventCreateWithFlags(&w->jobs_done, hipEventDisableTiming));
        g_ws.push_back(w);
    }
    if (capturing(s) && (w->pad_elems < pad_elems || tab_bytes > 0)) {
        set_error(tab_bytes > 0
                      ? "batched gathers stage a host job table and cannot "
                        "be captured; capture ch_convolve calls"
                      : "the gather workspace must grow while the stream is "
                        "being captured: run the pipeline once before "
                        "capturing it");
        return CH_ERR_ARG;
    }
    if (w->pad_elems < pad_elems && stream_graph_holds(s) > 0) {
        set_error("the gather workspace must grow, but a live graph "
                  "captured on this stream reads it: destroy the graph or "
                  "use another stream");
        return CH_ERR_ARG;
    }
    if (w->pad_elems < pad_elems) {
        // the old buffer may still be read by queued kernels of this stream
        CH_HIP(hipStreamSynchronize(s));
        CH_HIP(hipFree(w->pad));
        w->pad = nullptr;
        w->pad_elems = 0;
        CH_HIP(hipMalloc((void**)&w->pad, pad_elems * sizeof(float)));
        w->pad_elems = pad_elems;
    }
    if (tab_bytes > 0) {
        CH_HIP(hipEventSynchronize(w->jobs_done));
        if (w->tab_bytes < tab_bytes) {
            CH_HIP(hipStreamSynchronize(s));
            CH_HIP(hipFree(w->d_tab));
            CH_HIP(hipHostFree(w->h_tab));
            w->d_tab = nullptr;
            w->h_tab = nullptr;
            w->tab_bytes = 0;
            const size_t cap = std::max(tab_bytes, (size_t)16384);
            CH_HIP(hipMalloc(&w->d_tab, cap));
            CH_HIP(hipHostMalloc(&w->h_tab, cap, hipHostMallocDefault));
            w->tab_bytes = cap;
        }
    }
    *out = w;
    return CH_OK;
}

void release_gather_workspace(hipStream_t s)
{
    std::lock_guard<std::mutex> lock(g_ws_mu);
    for (size_t i = 0; i < g_ws.size(); ++i) {
        Workspace* w = g_ws[i];
        if (w->stream != s)
            continue;
        (void)hipStreamSynchronize(s);
        (void)hipFree(w->pad);
        (void)hipFree(w->d_tab);
        (void)hipHostFree(w->h_tab);
        (void)hipEventDestroy(w->jobs_done);
        delete w;
        g_ws.erase(g_ws.begin() + i);
        return;
    }
}

static ch_status pad_into(const float* T, float* P, const ch_plan* p,
                          hipStream_t s)
{
    const int64_t n = (int64_t)padded_elems(p);
    hipLaunchKernelGGL(pad_kernel, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, T, P, p->shape[0], p->shape[1],
                       p->shape[2]);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status pad_pair(const ch_plan* p, const float* a, const float* b,
                   hipStream_t s, const float** pa, const float** pb)
{
    std::lock_guard<std::mutex> lock(g_ws_mu);
    Workspace* w = nullptr;
    const size_t n = padded_elems(p);
    ch_status st = workspace(s, 2 * n, 0, &w);
    if (st != CH_OK)
        return st;
    st = pad_into(a, w->pad, p, s);
    if (st == CH_OK)
        st = pad_into(b, w->pad + n, p, s);
    if (st != CH_OK)
        return st;
    const size_t interior = (size_t)kYPad * (p->shape[2] + 2 * kZPad) + kZPad;
    *pa = w->pad + interior;
    *pb = w->pad + n + interior;
    return CH_OK;
}

// The x-slab boxes of T (used by the direct gather) into `ints`
// (slab_ints(X) of them); returns the slab records.
static int32_t* launch_slabs(const ch_plan* p, const float* T,
                             const int32_t* box, int32_t* ints,
                             hipStream_t s)
{
    int32_t* planes = ints;
    int32_t* slabs = ints + 4 * (size_t)p->shape[0];
    hipLaunchKernelGGL(plane_box_kernel, dim3((unsigned)p->shape[0]),
                       dim3(256), 0, s, T, p->shape[1], p->shape[2], box,
                       planes);
    hipLaunchKernelGGL(slab_kernel, dim3(1), dim3(64), 0, s,
                       (const int32_t*)planes, p->shape[0], slabs);
    return slabs;
}

static bool uses_slabs(const ch_plan* p) { return p->tile[0] == 1; }

template <typename OutT>
static ch_status convolve_one(const ch_plan* p, const float* d_terma,
                              OutT* d_dose, const int32_t* d_box,
                              int32_t accumulate, void* stream)
{
    CH_REQUIRE(p && d_terma && d_dose, "convolve: null pointer");
    Kernels<OutT> k;
    ch_status st = kernels_for<OutT>(p, k);
    if (st != CH_OK)
        return st;
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_ws_mu);
    Workspace* w = nullptr;
    const size_t np = padded_elems(p);
    st = workspace(s, np + slab_ints(p->shape[0]), 0, &w);
    if (st != CH_OK)
        return st;
    capture_pin(s, p);
    st = pad_into(d_terma, w->pad, p, s);
    if (st != CH_OK)
        return st;
    ConvJob j = make_job(p, w->pad, d_dose, d_box, accumulate);
    if (uses_slabs(p)) {
        j.slabs = launch_slabs(p, d_terma, d_box,
                               reinterpret_cast<int32_t*>(w->pad + np), s);
        CH_HIP(hipGetLastError());
    }
    const int64_t n_tiles = job_tiles(j);
    CH_REQUIRE(n_tiles < INT_MAX / 2, "too many tiles");
    const unsigned blocks =
        (unsigned)((n_tiles + k.tiles_per_block - 1) / k.tiles_per_block);
    LaunchTimer timer(s);
    timer.launch(k.single, dim3(blocks), dim3(k.threads), 0, j,
                 (int)n_tiles);
    CH_HIP(hipGetLastError());
    timer.stop();
    return CH_OK;
}

// ------------------------------------------------------- support box ----

__global__ __launch_bounds__(256) void box_kernel(const float* __restrict__ T,
                                                  int X, int Y, int Z,
                                                  int32_t* box)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)X * Y * Z;
    int32_t v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    if (i < n && T[i] != 0.0f) {
        const int z = (int)(i % Z);
        const int64_t r = i / Z;
        v[0] = v[3] = (int)(r / Y);
        v[1] = v[4] = (int)(r % Y);
        v[2] = v[5] = z;
    }
    // one partial box per block; box_finalize reduces them (no atomics)
    block_box_partial(v, box);
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_support_box(const float* d_terma, const int32_t h_shape[3],
                         int32_t* d_box, void* stream)
{
    CH_REQUIRE(d_terma && h_shape && d_box, "ch_support_box: null pointer");
    const int64_t n = (int64_t)h_shape[0] * h_shape[1] * h_shape[2];
    CH_REQUIRE(n > 0 && n < (1LL << 31), "ch_support_box: grid size");
    hipStream_t s = (hipStream_t)stream;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    void* parts = nullptr;
    ch_status st = stream_scratch(s, kScratchBoxPartials,
                                  (size_t)blocks * kBoxStride *
                                      sizeof(int32_t),
                                  &parts);
    if (st != CH_OK)
        return st;
    hipLaunchKernelGGL(box_kernel, dim3(blocks), dim3(256), 0, s, d_terma,
                       h_shape[0], h_shape[1], h_shape[2], (int32_t*)parts);
    CH_HIP(hipGetLastError());
    launch_box_finalize((const int32_t*)parts, (int)blocks, d_box, 0,
                        s);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_convolve(const ch_plan* plan, const float* d_terma,
                      float* d_dose, const int32_t* d_box, int32_t accumulate,
                      void* stream)
{
    return convolve_one<float>(plan, d_terma, d_dose, d_box, accumulate,
                               stream);
}

ch_status ch_convolve_f64(const ch_plan* plan, const float* d_terma,
                          double* d_dose, const int32_t* d_box, void* stream)
{
    return convolve_one<double>(plan, d_terma, d_dose, d_box, 1, stream);
}

ch_status ch_convolve_batch(int32_t n_jobs, const ch_conv_job* h_jobs,
                            void* stream)
{
    CH_REQUIRE(n_jobs >= 0, "n_jobs < 0");
    if (n_jobs == 0)
        return CH_OK;
    CH_REQUIRE(h_jobs, "ch_convolve_batch: null jobs");
    hipStream_t s = (hipStream_t)stream;
    for (int32_t i = 0; i < n_jobs; ++i) {
        CH_REQUIRE(h_jobs[i].plan, "job %d: null plan", i);
        CH_REQUIRE(h_jobs[i].d_terma && h_jobs[i].d_dose,
                   "job %d: null buffer", i);
    }
    std::lock_guard<std::mutex> lock(g_ws_mu);
    // Jobs that share the tile shape and accumulation mode run as one
    // launch (in order of first appearance); each job keeps its own output,
    // so grouping changes no result.
    std::vector<char> taken(n_jobs, 0);
    std::vector<int32_t> group;
    for (int32_t i0 = 0; i0 < n_jobs; ++i0) {
        if (taken[i0])
            continue;
        const ch_plan* p0 = h_jobs[i0].plan;
        group.clear();
        for (int32_t i = i0; i < n_jobs; ++i)
            if (!taken[i] && h_jobs[i].plan->tile[0] == p0->tile[0] &&
                h_jobs[i].plan->tile[1] == p0->tile[1] &&
                h_jobs[i].plan->tile[2] == p0->tile[2] &&
                h_jobs[i].plan->accum_mode == p0->accum_mode) {
                taken[i] = 1;
                group.push_back(i);
            }
        const int32_t ng = (int32_t)group.size();
        Kernels<float> k;
        ch_status st = kernels_for<float>(p0, k);
        if (st != CH_OK)
            return st;
        size_t pad_total = 0, ints_total = 0;
        for (int32_t i : group) {
            pad_total += padded_elems(h_jobs[i].plan);
            if (uses_slabs(h_jobs[i].plan))
                ints_total += slab_ints(h_jobs[i].plan->shape[0]);
        }
        Workspace* w = nullptr;
        st = workspace(s, pad_total + ints_total,
                       (size_t)ng * sizeof(ConvJob), &w);
        if (st != CH_OK)
            return st;
        ConvJob* tab = (ConvJob*)w->h_tab;
        int64_t tiles = 0;
        size_t pad_off = 0;
        int32_t* ints = reinterpret_cast<int32_t*>(w->pad + pad_total);
        for (int32_t g = 0; g < ng; ++g) {
            const ch_conv_job& hj = h_jobs[group[g]];
            float* P = w->pad + pad_off;
            st = pad_into(hj.d_terma, P, hj.plan, s);
            if (st != CH_OK)
                return st;
            pad_off += padded_elems(hj.plan);
            ConvJob j = make_job(hj.plan, P, hj.d_dose, hj.d_box,
                                 hj.accumulate);
            if (uses_slabs(hj.plan)) {
                j.slabs = launch_slabs(hj.plan, hj.d_terma, hj.d_box, ints,
                                       s);
                CH_HIP(hipGetLastError());
                ints += slab_ints(hj.plan->shape[0]);
            }
            j.tile_begin = (int32_t)tiles;
            tiles += job_tiles(j);
            CH_REQUIRE(tiles < INT_MAX / 2, "batch has too many tiles");
            tab[g] = j;
        }
        CH_HIP(hipMemcpyAsync(w->d_tab, w->h_tab,
                              (size_t)ng * sizeof(ConvJob),
                              hipMemcpyHostToDevice, s));
        CH_HIP(hipEventRecord(w->jobs_done, s));
        if (tiles == 0)
            continue;
        const unsigned blocks =
            (unsigned)((tiles + k.batch_tiles_per_block - 1) /
                       k.batch_tiles_per_block);
        LaunchTimer timer(s);
        timer.launch(k.batch, dim3(blocks), dim3(k.batch_threads), 0,
                     (const ConvJob*)w->d_tab, ng, (int)tiles);
        CH_HIP(hipGetLastError());
        timer.stop();
    }
    return CH_OK;
}

ch_status ch_convolve_groups(int32_t n_groups, const ch_conv_group* h_groups,
                             void* stream)
{
    CH_REQUIRE(n_groups >= 0, "n_groups < 0");
    if (n_groups == 0)
        return CH_OK;
    CH_REQUIRE(h_groups, "ch_convolve_groups: null groups");
    for (int32_t i = 0; i < n_groups; ++i) {
        const ch_conv_group& g = h_groups[i];
        CH_REQUIRE(g.plan, "group %d: null plan", i);
        CH_REQUIRE(g.plan->semantics != CH_SEM_RADIOLOGICAL,
                   "group %d: a CH_SEM_RADIOLOGICAL plan is used through "
                   "ch_convolve_rad", i);
        CH_REQUIRE(g.n_beams >= 1 && g.n_beams <= CH_MAX_GROUP,
                   "group %d: %d beams (1..%d)", i, g.n_beams, CH_MAX_GROUP);
        for (int b = 0; b < g.n_beams; ++b)
            CH_REQUIRE(g.d_terma[b] && g.d_dose[b],
                       "group %d beam %d: null buffer", i, b);
        CH_REQUIRE(padded_elems(g.plan) < (1ull << 28),
                   "group %d: grid too large for a beam group", i);
    }
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(g_ws_mu);
    // one launch per (direct tile, accumulation mode); plans tuned for the
    // staged gather use the 1 x 4 x 16 tile here
    auto tile_of = [](const ch_plan* p, int a) {
        if (p->tile[0] != 1)
            return a == 1 ? 4 : 16;
        return a == 1 ? p->tile[1] : p->tile[2];
    };
    std::vector<char> taken(n_groups, 0);
    std::vector<int32_t> sel;
    for (int32_t i0 = 0; i0 < n_groups; ++i0) {
        if (taken[i0])
            continue;
        const ch_plan* p0 = h_groups[i0].plan;
        const int ty = tile_of(p0, 1), tz = tile_of(p0, 2);
        sel.clear();
        for (int32_t i = i0; i < n_groups; ++i)
            if (!taken[i] && tile_of(h_groups[i].plan, 1) == ty &&
                tile_of(h_groups[i].plan, 2) == tz &&
                h_groups[i].plan->accum_mode == p0->accum_mode) {
                taken[i] = 1;
                sel.push_back(i);
            }
        void (*kern)(const GroupJob*, int, int) = nullptr;
#define CH_GPICK(TY, TZ, A)                                                 \
    if (ty == TY && tz == TZ && p0->accum_mode == A)                        \
        kern = gather_group<TY, TZ, A>;
#define CH_GPICK3(TY, TZ) CH_GPICK(TY, TZ, 0) CH_GPICK(TY, TZ, 1) CH_GPICK(TY, TZ, 2)
        CH_GPICK3(1, 64)
        CH_GPICK3(2, 32)
        CH_GPICK3(4, 16)
        CH_GPICK3(8, 8)
#undef CH_GPICK3
#undef CH_GPICK
        CH_REQUIRE(kern, "no group gather for tile 1x%dx%d", ty, tz);
        const int32_t ng = (int32_t)sel.size();
        size_t pad_total = 0;
        for (int32_t i : sel)
            pad_total += (size_t)kGroup * padded_elems(h_groups[i].plan);
        Workspace* w = nullptr;
        ch_status st =
            workspace(s, pad_total, (size_t)ng * sizeof(GroupJob), &w);
        if (st != CH_OK)
            return st;
        GroupJob* tab = (GroupJob*)w->h_tab;
        int64_t tiles = 0;
        size_t pad_off = 0;
        for (int32_t g = 0; g < ng; ++g) {
            const ch_conv_group& hg = h_groups[sel[g]];
            const ch_plan* p = hg.plan;
            float* P = w->pad + pad_off;
            const int64_t n = (int64_t)padded_elems(p);
            hipLaunchKernelGGL(
                pad_group_kernel, dim3((unsigned)((n + 255) / 256)),
                dim3(256), 0, s, hg.d_terma[0],
                hg.n_beams > 1 ? hg.d_terma[1] : nullptr,
                hg.n_beams > 2 ? hg.d_terma[2] : nullptr,
                hg.n_beams > 3 ? hg.d_terma[3] : nullptr, (float4*)P,
                p->shape[0], p->shape[1], p->shape[2]);
            CH_HIP(hipGetLastError());
            pad_off += (size_t)kGroup * padded_elems(p);
            GroupJob j;
            const int Zp = p->shape[2] + 2 * kZPad;
            j.terma = P + ((size_t)kYPad * Zp + kZPad) * kGroup;
            for (int b = 0; b < kGroup; ++b) {
                j.dose[b] = b < hg.n_beams ? hg.d_dose[b] : nullptr;
                j.box[b] = b < hg.n_beams ? hg.d_box[b] : nullptr;
            }
            j.lin = p->d_lin;
            j.k = p->d_k;
            j.ftab = p->d_ftab;
            j.meta = p->d_meta;
            j.X = p->shape[0];
            j.Y = p->shape[1];
            j.Z = p->shape[2];
            j.n_cones = p->n_cones;
            j.n_beams = hg.n_beams;
            j.accumulate = hg.accumulate;
            j.nty = (j.Y + ty - 1) / ty;
            j.ntz = (j.Z + tz - 1) / tz;
            j.tile_begin = (int32_t)tiles;
            tiles += (int64_t)j.X * j.nty * j.ntz;
            CH_REQUIRE(tiles < INT_MAX / 2, "groups have too many tiles");
            tab[g] = j;
        }
        CH_HIP(hipMemcpyAsync(w->d_tab, w->h_tab,
                              (size_t)ng * sizeof(GroupJob),
                              hipMemcpyHostToDevice, s));
        CH_HIP(hipEventRecord(w->jobs_done, s));
        LaunchTimer timer(s);
        timer.launch(kern, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0,
                     (const GroupJob*)w->d_tab, ng, (int)tiles);
        CH_HIP(hipGetLastError());
        timer.stop();
    }
    return CH_OK;
}

}  // extern "C"
