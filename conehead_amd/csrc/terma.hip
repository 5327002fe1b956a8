// terma.hip -- fused TERMA: hit test, FP64-DDA radiological depth, off-axis
// distance, three-source fluence, softening, horn and the polyenergetic sum,
// one lane per dose-grid voxel, one launch per beam.
//
// The reference runs these as six kernels with a host round trip after each
// (conehead/conehead.py:536-682).  Geometry (sub-sample positions, plane
// hits, ray directions, DDA alphas) is FP64 with contraction off so that the
// traversal and the aperture lookups match the FP64 oracle bit for bit; the
// depth integral uses FP32 with Kahan compensation and the fluence and TERMA
// arithmetic is FP32.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "aperture.h"
#include "dda_core.h"

namespace ch {

struct TermaParams {
    double src[3];
    double vy[3];
    double tr0[3];     // row 0 of the world -> source rotation
    double tr2[3];     // row 2
    double origin[3];
    double spacing[3];
    double sad;
    // Thresholds stay FP64 so the piecewise terms switch exactly where the
    // FP64 oracle switches them.
    double rInner, rOuter, softLimit, zAnn64;
    float sPri, sAnn, zAnn, zExp, sExp, kExp;
    float softRatio, hornRatio;
    int32_t samples, n_energy;
    int32_t X, Y, Z;
    float ew[CH_MAX_ENERGIES];   // weight * energy * mu, folded on the host
    float emu[CH_MAX_ENERGIES];  // mu
    int32_t ap_kind, ap_n;       // aperture (ch_aperture)
    int32_t ap_bbox[4];
    const int32_t* ap_table;
    const float* raster;
    const float* density;
    float* terma;
    int32_t* box;              // support-box partials of this beam or null
    int32_t* counts;           // this beam's counter state (kCounterInts)
    int32_t* list;             // and cost-class brick lists
    float* debug;              // 4 planes of X*Y*Z or null
    int32_t* status;           // geometry error flag
    int32_t accumulate;        // terma += this beam
    // per-launch constants, computed on the host with the kernel's own
    // operations: sub-sample steps spacing / n, n . source, n^3, and the
    // blocked fraction of a fully open voxel (1 / n^3 added n^3 times)
    double off[3];
    double ndotn_src;
    double n3;
    double blocked_one;
    float inv_n3, inv_ns;
    float src_vox[3];          // (src - origin) / spacing, for cost classes
    // 100 tr0.src, 100 tr2.src and -100 n.src, for fast_cells, which is
    // used only when |ci_src|, |cj_src| < 1e5 (fast_ok)
    double ci_src, cj_src, k100;
    int32_t fast_ok;
};

__device__ __forceinline__ double dot3(const double* a, double b0, double b1,
                                       double b2)
{
    return a[0] * b0 + a[1] * b1 + a[2] * b2;
}

// Voxel centre and unit direction toward the source (conehead.py:226-238),
// FP64 without contraction.
__device__ __forceinline__ void ray_to_source(const TermaParams& P, int x,
                                              int y, int z, double* pos,
                                              double* r, double& mag,
                                              double* dir)
{
    pos[0] = P.origin[0] + P.spacing[0] * x;
    pos[1] = P.origin[1] + P.spacing[1] * y;
    pos[2] = P.origin[2] + P.spacing[2] * z;
    r[0] = P.src[0] - pos[0];
    r[1] = P.src[1] - pos[1];
    r[2] = P.src[2] - pos[2];
    mag = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    dir[0] = r[0] / mag;
    dir[1] = r[1] / mag;
    dir[2] = r[2] / mag;
}

// The traversal of each TERMA ray, for parity checks of the depth integral.
__global__ __launch_bounds__(256) void terma_rays_kernel(
    TermaParams P, int32_t n, const int32_t* __restrict__ vox_in,
    int32_t cap, double* __restrict__ dirs, int32_t* __restrict__ vox,
    double* __restrict__ tout, int32_t* __restrict__ count)
{
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    int32_t c0 = vox_in[3 * r], c1 = vox_in[3 * r + 1], c2 = vox_in[3 * r + 2];
    double pos[3], rv[3], dir[3], mag;
    ray_to_source(P, c0, c1, c2, pos, rv, mag, dir);
    dirs[3 * r] = dir[0];
    dirs[3 * r + 1] = dir[1];
    dirs[3 * r + 2] = dir[2];
    Dda d;
    d.init(dir[0], dir[1], dir[2], P.spacing[0], P.spacing[1], P.spacing[2]);
    int32_t m = 0;
    while (c0 >= 0 && c0 < P.X && c1 >= 0 && c1 < P.Y && c2 >= 0 &&
           c2 < P.Z) {
        if (m < cap) {
            vox[((int64_t)r * cap + m) * 3 + 0] = c0;
            vox[((int64_t)r * cap + m) * 3 + 1] = c1;
            vox[((int64_t)r * cap + m) * 3 + 2] = c2;
        }
        const double t = d.advance(d.axis(), c0, c1, c2);
        if (m < cap)
            tout[(int64_t)r * cap + m] = t;
        ++m;
    }
    count[r] = m;
}

// Cell coordinates (row, column) times 100 -- before flooring -- of the
// isocentre-plane image of point p seen from the source, and n.u with
// u = source - p: the plane hit of conehead/conehead.py:25-47 and the
// source-frame rotation of :109-111, FP64 without contraction.  When
// |n.u| < 1e-6 the image is not used.
struct PlaneImage {
    double ci, cj, ndotu;
};

__device__ __forceinline__ PlaneImage plane_image(const TermaParams& P,
                                                  double ps0, double ps1,
                                                  double ps2,
                                                  double ndotn_src)
{
    PlaneImage c;
    const double u0 = P.src[0] - ps0, u1 = P.src[1] - ps1,
                 u2 = P.src[2] - ps2;
    c.ndotu = dot3(P.vy, u0, u1, u2);
    const double si = -ndotn_src / c.ndotu;
    const double pp0 = P.src[0] + si * u0 + 0.0;
    const double pp1 = P.src[1] + si * u1 + 0.0;
    const double pp2 = P.src[2] + si * u2 + 0.0;
    c.ci = dot3(P.tr0, pp0, pp1, pp2) * 100.0;
    c.cj = dot3(P.tr2, pp0, pp1, pp2) * 100.0;
    return c;
}

// Sub-sample position along one axis (conehead.py:82-96): voxel centre
// minus half a voxel plus half a sub-step plus k sub-steps.
__device__ __forceinline__ double sample_pos(const TermaParams& P, int a,
                                            int v, double off, int k)
{
    const double pos = P.origin[a] + P.spacing[a] * v;
    return pos - P.spacing[a] / 2 + off / 2 + off * k;
}

// Isocentre-plane image of a corner of a sub-sample box: as plane_image,
// with the plane-hit quotient from a reciprocal refined by one Newton step
// (relative error ~1e-15, so the cell coordinates are off by ~1e-11 cell).
// Only the enclosure of the corner images is used, with a margin of 1e-6
// cell, so the samples' exact images stay inside it.
__device__ __forceinline__ PlaneImage corner_image(const TermaParams& P,
                                                   double ps0, double ps1,
                                                   double ps2)
{
    PlaneImage c;
    const double u0 = P.src[0] - ps0, u1 = P.src[1] - ps1,
                 u2 = P.src[2] - ps2;
    c.ndotu = dot3(P.vy, u0, u1, u2);
    double r = __builtin_amdgcn_rcp(c.ndotu);
    r = r * (2.0 - c.ndotu * r);
    const double si = -P.ndotn_src * r;
    const double pp0 = P.src[0] + si * u0;
    const double pp1 = P.src[1] + si * u1;
    const double pp2 = P.src[2] + si * u2;
    c.ci = dot3(P.tr0, pp0, pp1, pp2) * 100.0;
    c.cj = dot3(P.tr2, pp0, pp1, pp2) * 100.0;
    return c;
}

// Raster cells floor(c -/+ 1e-6) + 1999 (a cell coordinate widened by the
// rounding margin), clamped to [-2, 3999]: every test below treats cells
// below -1 or above 3998 alike, so the clamp changes no decision.
__device__ __forceinline__ int cell_of(double c)
{
    const double f = fmin(fmax(floor(c) + 1999.0, -2.0), 3999.0);
    return (int)f;
}

// True when no raster cell of [a, b] can read a value inside the
// aperture's bounding cells [lo, hi).  Cells outside [-1, 3998] read 0
// (conehead.py:53-67) and cell -1 reads the last row or column (Python's
// negative indexing).
__device__ __forceinline__ bool axis_misses(int a, int b, int lo, int hi)
{
    if (b < -1 || a > 3998)
        return true;
    const int A = max(a, 0), B = min(b, 3998);
    const bool in = A <= B && B >= lo && A <= hi - 1;
    const bool wrap = a <= -1 && 3999 >= lo && 3999 <= hi - 1;
    return !(in || wrap);
}

// Cell box of a set of plane images and whether every n.u has one sign and
// |n.u| >= 1e-6 (then the central projection maps the convex hull of the
// points into the convex hull of their images).  Cells are monotone in the
// coordinate, so the box of the cells is the cells of the box.
struct CellBox {
    int a_i, b_i, a_j, b_j;
    int flags;  // 1: all |n.u| >= 1e-6, 2: all n.u > 0, 4: all n.u <= 0
    __device__ __forceinline__ void init(const PlaneImage& c)
    {
        a_i = cell_of(c.ci - 1e-6);
        b_i = cell_of(c.ci + 1e-6);
        a_j = cell_of(c.cj - 1e-6);
        b_j = cell_of(c.cj + 1e-6);
        flags = (fabs(c.ndotu) >= 1e-6 ? 1 : 0) | (c.ndotu > 0 ? 2 : 4);
    }
    __device__ __forceinline__ void add(const CellBox& o)
    {
        a_i = min(a_i, o.a_i);
        b_i = max(b_i, o.b_i);
        a_j = min(a_j, o.a_j);
        b_j = max(b_j, o.b_j);
        flags &= o.flags;
    }
    __device__ __forceinline__ bool ok() const
    {
        return (flags & 1) && (flags & 6);
    }
    // The transmission shared by every sub-sample whose image lies in the
    // box, or -1 when it is not one known constant: 0 outside the
    // aperture's bounding cells or the raster, or the leaf table's or the
    // rectangles' value over the block.
    __device__ __forceinline__ float value(const TermaParams& P) const
    {
        if (!ok())
            return -1.0f;
        if (axis_misses(a_i, b_i, P.ap_bbox[0], P.ap_bbox[1]) ||
            axis_misses(a_j, b_j, P.ap_bbox[2], P.ap_bbox[3]))
            return 0.0f;
        if (P.ap_kind == CH_APERTURE_RASTER ||
            !(a_i >= 0 && b_i <= 3998 && a_j >= 0 && b_j <= 3998))
            return -1.0f;
        return P.ap_kind == CH_APERTURE_MLC
                   ? mlc_block_value(P.ap_table, a_i, b_i, a_j, b_j)
                   : rects_block_value(P.ap_table, P.ap_n, a_i, b_i, a_j,
                                       b_j);
    }
};

// Reduction of a CellBox over each group of 8 lanes with DPP moves: lanes
// exchange within quads (xor 1, xor 2), then mirror within half rows of 8.
template <int CTRL>
__device__ __forceinline__ int dpp(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ void add_dpp(CellBox& b)
{
    CellBox o;
    o.a_i = dpp<CTRL>(b.a_i);
    o.b_i = dpp<CTRL>(b.b_i);
    o.a_j = dpp<CTRL>(b.a_j);
    o.b_j = dpp<CTRL>(b.b_j);
    o.flags = dpp<CTRL>(b.flags);
    b.add(o);
}
constexpr int kDppQuadXor1 = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadXor2 = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;   // row_half_mirror

// conehead.py:50-68 at cell coordinates (times 100, not floored) of an
// image: raster cell floor(c) + 2000 - 1, a -1 index wrapping to the last
// row or column, looked up in the beam's aperture.
__device__ __forceinline__ float transmission_at(const TermaParams& P,
                                                 double fi, double fj)
{
    const double px = fi + 2000.0;
    const double pz = fj + 2000.0;
    if (px < 0 || px > 3999 || pz < 0 || pz > 3999)
        return 0.0f;
    int i = (int)px - 1, j = (int)pz - 1;
    if (i < 0)
        i += 4000;
    if (j < 0)
        j += 4000;
    if (P.ap_kind == CH_APERTURE_MLC)
        return mlc_value(P.ap_table, i, j);
    if (P.ap_kind == CH_APERTURE_RECTS)
        return rects_value(P.ap_table, P.ap_n, i, j);
    return P.raster[i * 4000 + j];
}
// (fi, fj: the floors of the cell coordinates, all the lookup uses)
__device__ __forceinline__ float transmission(const TermaParams& P, double ci,
                                              double cj)
{
    return transmission_at(P, floor(ci), floor(cj));
}

// The floors of a sub-sample's cell coordinates (plane_image's ci, cj)
// from the same geometry rearranged: with u = source - p,
//   ci = 100 tr0.src + (-100 n.src) (tr0.u) / (n.u)
// (cj with tr2), the products fused and the quotient from a reciprocal
// refined by one Newton step.  For |n.u| >= 1 cm, |ci|, |cj| < 1e6 and
// |100 tr0.src|, |100 tr2.src| < 1e5 this and plane_image's rounded
// evaluation are each within 1e-8 cell of the exact value (a few dozen
// roundings of quantities below 2.2e6 at 2^-53), so when both coordinates
// lie more than 1e-6 cell from an integer their floors are plane_image's,
// and |n.u| >= 1e-6 (no geometry flag).  Returns false otherwise:
// plane_image then decides.
__device__ __forceinline__ bool fast_cells(const TermaParams& P, double p0,
                                           double p1, double p2, double& fi,
                                           double& fj)
{
    if (!P.fast_ok)
        return false;  // uniform over the launch
    const double u0 = P.src[0] - p0, u1 = P.src[1] - p1,
                 u2 = P.src[2] - p2;
    const double d = fma(P.vy[2], u2, fma(P.vy[1], u1, P.vy[0] * u0));
    const double ni = fma(P.tr0[2], u2, fma(P.tr0[1], u1, P.tr0[0] * u0));
    const double nj = fma(P.tr2[2], u2, fma(P.tr2[1], u1, P.tr2[0] * u0));
    double r = __builtin_amdgcn_rcp(d);
    r = r * fma(-d, r, 2.0);
    const double s = P.k100 * r;
    const double ci = fma(s, ni, P.ci_src), cj = fma(s, nj, P.cj_src);
    fi = floor(ci);
    fj = floor(cj);
    const double ei = ci - fi, ej = cj - fj;
    constexpr double kMargin = 1e-6;
    return fabs(d) >= 1.0 && fabs(ci) < 1e6 && fabs(cj) < 1e6 &&
           ei > kMargin && ei < 1.0 - kMargin && ej > kMargin &&
           ej < 1.0 - kMargin;
}

// LDS of one wavefront exchanged between its own lanes: the writes are
// complete and visible once every lane has passed this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Radiological depth from voxel (x, y, z) toward the source: the FP64 DDA
// of conehead/conehead.py:138-166 (the traversal of dda_3d_c.pyx:76-99,
// the same t-values, voxels and tie rules as ch_dda_3d), each voxel's
// density times its path length summed in FP32 with Kahan compensation, in
// path order (conehead.py:249-254).  The step needs no branch.  The tie
// rules (x if t0 < t1 and t0 < t2, else y if t1 <= t0 and t1 < t2, else z)
// take the smallest crossing t = min(t0, t1, t2) and, among the axes that
// reach it, z before y before x: z when t2 == t, else y when t1 == t, else
// x.  The axis taken advances by t_a = fma(dt_a, 1, t_a), the others by
// fma(dt_a, 0, t_a) = t_a, both rounded as t_a + dt_a and t_a are.
// kWalkBatch steps are traversed first and their density loads issued
// together before the first is used; lanes past the grid exit keep
// stepping with their sums frozen, their range-checked loads reading 0
// wherever they land.  Byte offsets are 32-bit (ch_terma requires
// 4 X Y Z < 2^32).  PACK (every axis of at most kPackMax voxels): the
// steps left before each axis leaves the grid are one word of three 10-bit
// fields, field a = 511 - left_a counting up, whose top bit (the guard) is
// set on the step that takes axis a past the grid, when left_a would go
// below zero; one select, one add and one test per step instead of three
// counters.  A field stops at most kWalkBatch - 1 steps past 512, far from
// carrying into the next.
constexpr int kWalkBatch = 8;
constexpr int kPackMax = 512;

template <bool PACK>
__device__ __forceinline__ float depth_walk_t(const float* __restrict__ rho,
                                              int X, int Y, int Z, int x,
                                              int y, int z, const Dda& d)
{
    double t0 = d.t0, t1 = d.t1, t2 = d.t2;
    const double dt0 = d.dt0, dt1 = d.dt1, dt2 = d.dt2;
    // steps left before each axis leaves the grid
    int32_t left0 = d.s0 > 0 ? X - 1 - x : x;
    int32_t left1 = d.s1 > 0 ? Y - 1 - y : y;
    int32_t left2 = d.s2 > 0 ? Z - 1 - z : z;
    constexpr uint32_t kGuards = (1u << 9) | (1u << 19) | (1u << 29);
    uint32_t pk = PACK ? (511u - (uint32_t)left0) |
                             ((511u - (uint32_t)left1) << 10) |
                             ((511u - (uint32_t)left2) << 20)
                       : 0u;
    // byte strides and offset
    const int32_t st0 = 4 * d.s0 * Y * Z, st1 = 4 * d.s1 * Z, st2 = 4 * d.s2;
    uint32_t off = 4u * (uint32_t)((x * Y + y) * Z + z);
    double tprev = 0.0;
    float de = 0.0f, dc = 0.0f;
    bool live = true;
    // range-checked loads: past the grid a read returns 0 (and is not used)
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)rho, (short)0, 4 * X * Y * Z, kBufferFlags);
    while (live) {
        uint32_t offs[kWalkBatch];
        float seg[kWalkBatch];
        bool use[kWalkBatch];
#pragma unroll
        for (int k = 0; k < kWalkBatch; ++k) {
            use[k] = live;
            const double t = fmin(fmin(t0, t1), t2);
            const bool m2 = t2 == t;
            const bool m1 = !m2 & (t1 == t);
            const bool m0 = !(m1 | m2);
            offs[k] = off;
            seg[k] = (float)(t - tprev);
            tprev = t;
            t0 = fma(dt0, m0 ? 1.0 : 0.0, t0);
            t1 = fma(dt1, m1 ? 1.0 : 0.0, t1);
            t2 = fma(dt2, m2 ? 1.0 : 0.0, t2);
            off += (uint32_t)(m0 ? st0 : (m1 ? st1 : st2));
            if (PACK) {
                pk += m0 ? 1u : (m1 ? (1u << 10) : (1u << 20));
                live = live & ((pk & kGuards) == 0u);
            } else {
                left0 -= (int32_t)m0;
                left1 -= (int32_t)m1;
                left2 -= (int32_t)m2;
                live = live & ((left0 | left1 | left2) >= 0);
            }
        }
        float r[kWalkBatch];
#pragma unroll
        for (int k = 0; k < kWalkBatch; ++k)
            r[k] = __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                rsrc, (int)offs[k], 0, 0));
#pragma unroll
        for (int k = 0; k < kWalkBatch; ++k) {
            // a step past the exit adds dc: yk = 0, de keeps its bits (dc
            // is not used after the walk)
            const float term = use[k] ? r[k] * seg[k] : dc;
            const float yk = term - dc;
            const float tk = de + yk;
            dc = (tk - de) - yk;
            de = tk;
        }
    }
    return de;
}
// Sub-sample values staged per wavefront by the compacted hit test.
constexpr int kSampleBuf = 512;

// TERMA in two phases.  The hit test (conehead.py:78-120) runs in three
// stages that each give the exact blocked fraction of the loop over n^3
// sub-samples: (1) the brick of 64 voxels at once (brick_value), (2) each
// voxel's own 8 corners by the same argument, (3) the sub-samples of the
// voxels still open, spread over all 64 lanes (a few voxels at a time),
// each lane evaluating one sub-sample into LDS; each voxel then sums its
// own values in the loop's order, so the result is bitwise that of the
// loop (when most of the lanes need samples the loop runs on each lane).
// Phase 1 sorts the bricks by stage 1: those outside the aperture are done
// there; the rest -- the voxels the beam reaches, a seventh of a 128^3 grid
// for a 10 x 10 cm field -- go to phase 2 in order of their estimated cost,
// the longest depth walks first, so they start together and the cheaper
// ones fill in behind them.
// Bricks of 4 x 4 x 4 voxels (one per wavefront in phase 2), indexed z
// fastest: the corner voxel of brick b, and lane l's voxel (x0 + l / 16,
// y0 + l / 4 % 4, z0 + l % 4).
struct Brick {
    int x0, y0, z0;
};
__device__ __forceinline__ Brick brick_of(const TermaParams& P, int b)
{
    const int nbz = (P.Z + 3) / 4, nby = (P.Y + 3) / 4;
    Brick k;
    k.z0 = (b % nbz) * 4;
    k.y0 = ((b / nbz) % nby) * 4;
    k.x0 = (b / nbz / nby) * 4;
    return k;
}

// Hit-test stage 1 for the whole wavefront (every lane must call it): the
// box spanned by every sub-sample of the brick's 64 voxels, corner (l & 7)
// per lane.  The sub-samples are convex combinations of those corners, and
// the projection through the source onto the isocentre plane keeps convex
// combinations while n.u keeps one sign (n.u is affine in the point), so
// when the corner images have one sign of n.u and their cell box (widened by
// 1e-6 cell for rounding) reads one known transmission v, every sub-sample
// of the brick reads v.  Returns v, or -1.
__device__ __forceinline__ float brick_value(const TermaParams& P,
                                             const Brick& k, int l)
{
    const int ns = P.samples;
    const int c = l & 7;
    const int xa = (c & 1) ? min(k.x0 + 3, P.X - 1) : k.x0;
    const int ya = (c & 2) ? min(k.y0 + 3, P.Y - 1) : k.y0;
    const int za = (c & 4) ? min(k.z0 + 3, P.Z - 1) : k.z0;
    const int s0 = (c & 1) ? ns - 1 : 0;
    const int s1 = (c & 2) ? ns - 1 : 0;
    const int s2 = (c & 4) ? ns - 1 : 0;
    CellBox b;
    b.init(corner_image(P, sample_pos(P, 0, xa, P.off[0], s0),
                        sample_pos(P, 1, ya, P.off[1], s1),
                        sample_pos(P, 2, za, P.off[2], s2)));
    add_dpp<kDppQuadXor1>(b);
    add_dpp<kDppQuadXor2>(b);
    add_dpp<kDppHalfMirror>(b);
    return b.value(P);
}

// Cost classes of the bricks that need work: kCostClasses buckets of the
// estimated depth-walk length of the brick's centre (voxels to the grid exit
// toward the source), plus 2 n^3 steps for a brick on the aperture's edge
// (its sub-samples cost about that much).
constexpr int kCostClasses = 16;
// Each class is split over kShards counters and sub-lists (the bricks of
// phase-1 workgroup w in shard w % kShards), so the appending atomics of
// phase 1 spread over kCostClasses x kShards addresses instead of queueing
// on a few.
constexpr int kShards = 16;
constexpr int kCounters = kCostClasses * kShards;   // 4 per lane
// A beam's counter state (int32, kept between calls, zero when allocated):
// the kCounters append counters, two base rows of kCounters, and the call
// word.  The counters only grow (unsigned, wrapping): a call's count of a
// list is its counter minus the base row of the call's parity, the totals
// the previous call's phase 2 left there, and its entries sit at the
// counter's values modulo the list's power-of-two capacity.  Phase 1 adds
// one to the call word (its block 0), phase 2 reads it and writes the
// totals into the other base row (its wavefront 0), so no launch has to
// clear the counters between calls.
constexpr int kBaseRows = kCounters;                // base rows at +256, +512
constexpr int kCallWord = 3 * kCounters;
constexpr int kCounterInts = 3 * kCounters + 64;    // 256-byte multiple

__device__ __forceinline__ int brick_cost_class(const TermaParams& P,
                                                const Brick& k, bool edge)
{
    const float c[3] = {k.x0 + 1.5f, k.y0 + 1.5f, k.z0 + 1.5f};
    const int n[3] = {P.X, P.Y, P.Z};
    float d[3], t = 3.0e38f;
    for (int a = 0; a < 3; ++a) {
        d[a] = P.src_vox[a] - c[a];
        if (d[a] > 0.0f)
            t = fminf(t, ((float)n[a] - 0.5f - c[a]) / d[a]);
        else if (d[a] < 0.0f)
            t = fminf(t, (c[a] + 0.5f) / -d[a]);
    }
    float steps = 0.0f;
    for (int a = 0; a < 3; ++a)
        steps += t * fabsf(d[a]);
    const int n3 = P.samples * P.samples * P.samples;
    const float full = (float)(P.X + P.Y + P.Z + 2 * n3);
    const float cost = steps + (edge ? 2.0f * n3 : 0.0f);
    const int q = (int)(cost * kCostClasses / full);
    return q < 0 ? 0 : (q >= kCostClasses ? kCostClasses - 1 : q);
}

// Phase 1: stage 1 of the hit test for eight bricks per wavefront, the
// 2 x 2 x 2 bricks of one super-brick, one group of 8 lanes each (the cell
// box reduction stays within each group), four super-bricks per workgroup.
// A brick that reads 0 everywhere gets T = 0 here; every other brick (all
// of them when the debug planes are asked for) is appended to its cost
// class for phase 2, with its stage-1 result: the workgroup ranks its
// bricks per class in LDS, then reserves their slots in its shard's lists
// with one device atomic per class.
struct SuperGrid {
    int nbx, nby, nbz;   // bricks per axis
    int sby, sbz;        // super-bricks per axis (y, z)
};
__device__ __forceinline__ SuperGrid super_grid(const TermaParams& P)
{
    SuperGrid g;
    g.nbx = (P.X + 3) / 4;
    g.nby = (P.Y + 3) / 4;
    g.nbz = (P.Z + 3) / 4;
    g.sby = (g.nby + 1) / 2;
    g.sbz = (g.nbz + 1) / 2;
    return g;
}
__host__ __device__ __forceinline__ int32_t super_bricks(int X, int Y, int Z)
{
    return (((X + 3) / 4 + 1) / 2) * (((Y + 3) / 4 + 1) / 2) *
           (((Z + 3) / 4 + 1) / 2);
}
constexpr int kSortSupers = 4;                   // super-bricks per workgroup
constexpr int kSortBricks = 8 * kSortSupers;     // bricks per workgroup
__host__ __device__ __forceinline__ int32_t sort_blocks(int X, int Y, int Z)
{
    return (super_bricks(X, Y, Z) + kSortSupers - 1) / kSortSupers;
}
// Capacity of one shard's list of one class: kSortBricks per workgroup of
// the shard, rounded up to a power of two (entries go to counter values
// modulo the capacity).
__host__ __device__ __forceinline__ int32_t shard_cap(int X, int Y, int Z)
{
    const int32_t need =
        (sort_blocks(X, Y, Z) + kShards - 1) / kShards * kSortBricks;
    int32_t cap = kSortBricks;
    while (cap < need)
        cap *= 2;
    return cap;
}

__device__ __forceinline__ void sort_block(const TermaParams& P, int blk,
                                           int32_t* s_cnt, int32_t* s_base)
{
    const int t = threadIdx.x, l = t & 63;
    if (t < kCostClasses)
        s_cnt[t] = 0;
    __syncthreads();
    const SuperGrid g = super_grid(P);
    const int sb = blk * kSortSupers + (t >> 6);
    int q = -1, rank = 0, entry = 0;
    if (sb < super_bricks(P.X, P.Y, P.Z)) {  // wavefront-uniform
        const int sxy = sb / g.sbz, sz = sb - sxy * g.sbz;
        const int sx = sxy / g.sby, sy = sxy - sx * g.sby;
        // lane l tests child l >> 3; every lane takes part in the group
        // reductions (a child past the grid tests its clamped neighbour and
        // its result is dropped)
        const int c = l >> 3;
        const int bx = 2 * sx + (c >> 2), by = 2 * sy + ((c >> 1) & 1),
                  bz = 2 * sz + (c & 1);
        const bool in = bx < g.nbx && by < g.nby && bz < g.nbz;
        Brick k;
        k.x0 = 4 * min(bx, g.nbx - 1);
        k.y0 = 4 * min(by, g.nby - 1);
        k.z0 = 4 * min(bz, g.nbz - 1);
        const float v = brick_value(P, k, l);
        const bool zero = in && v == 0.0f && P.debug == nullptr;
        // bit 8 c set when child c reads 0: lane l writes voxel l of each
        const uint64_t zmask = __ballot(zero) & 0x0101010101010101ull;
        if (zmask != 0 && !P.accumulate) {
            const int dx = l >> 4, dy = (l >> 2) & 3, dz = l & 3;
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) {
                if (!((zmask >> (8 * cc)) & 1))
                    continue;  // wavefront-uniform
                const int x = 4 * (2 * sx + (cc >> 2)) + dx,
                          y = 4 * (2 * sy + ((cc >> 1) & 1)) + dy,
                          z = 4 * (2 * sz + (cc & 1)) + dz;
                if (x < P.X && y < P.Y && z < P.Z)
                    P.terma[((int64_t)x * P.Y + y) * P.Z + z] = 0.0f;
            }
        }
        if (in && !zero && (l & 7) == 0) {
            q = brick_cost_class(P, k, v < 0.0f);
            // entry: brick index and stage-1 value (0 -> 0, 1 -> 1, -1 -> 2)
            const int b = (bx * g.nby + by) * g.nbz + bz;
            entry = b * 4 + (v > 0.0f ? 1 : (v < 0.0f ? 2 : 0));
            rank = atomicAdd(&s_cnt[q], 1);
        }
    }
    __syncthreads();
    const int shard = blk % kShards;
    if (t < kCostClasses) {
        const int n = s_cnt[t];
        s_base[t] = n > 0 ? (int32_t)atomicAdd(
                                (unsigned*)&P.counts[t * kShards + shard],
                                (unsigned)n)
                          : 0;
    }
    if (blk == 0 && t == 0)
        P.counts[kCallWord] += 1;  // read by this call's phase 2 only
    __syncthreads();
    if (q >= 0) {
        const int32_t cap = shard_cap(P.X, P.Y, P.Z);
        const uint32_t at = ((uint32_t)s_base[q] + (uint32_t)rank) &
                            (uint32_t)(cap - 1);
        P.list[(int64_t)(q * kShards + shard) * cap + at] = entry;
    }
}

__global__ __launch_bounds__(256) void terma_sort_kernel(TermaParams P)
{
    __shared__ int32_t s_cnt[kCostClasses], s_base[kCostClasses];
    sort_block(P, blockIdx.x, s_cnt, s_base);
}

// Phase 1 of several beams on one grid shape in one launch: the blocks of
// beam i are i * blocks_per_beam ... , its parameters tab[i] (a device
// table written before the launch and not written by it).
__global__ __launch_bounds__(256) void terma_sort_set_kernel(
    const TermaParams* __restrict__ tab, int32_t blocks_per_beam)
{
    __shared__ int32_t s_cnt[kCostClasses], s_base[kCostClasses];
    const int beam = blockIdx.x / blocks_per_beam;
    sort_block(tab[beam], blockIdx.x - beam * blocks_per_beam, s_cnt,
               s_base);
}

// TERMA of one brick (one wavefront, every lane calls it) whose stage-1 hit
// test gave v: hit-test stages 2 and 3, then per voxel the radiological
// depth, off-axis distance, fluence, softening, horn and the polyenergetic
// sum (conehead.py:71-377).  Returns the lane's T (0 outside the grid).
template <bool PACK>
__device__ __forceinline__ float terma_brick(const TermaParams& P,
                                             const Brick& k, float v, int l,
                                             double* __restrict__ s_samp,
                                             int32_t* __restrict__ s_lane,
                                             bool& bad)
{
    const int64_t nv = (int64_t)P.X * P.Y * P.Z;
    const int x0 = k.x0, y0 = k.y0, z0 = k.z0;
    const int x = x0 + (l >> 4);
    const int y = y0 + ((l >> 2) & 3);
    const int z = z0 + (l & 3);
    const bool inside = x < P.X && y < P.Y && z < P.Z;
    const int ns = P.samples;
    const int n3i = ns * ns * ns;
    const double n3 = P.n3;
    const double* off = P.off;
    const double ndotn_src = P.ndotn_src;
    float T = 0.0f;
        // 2. the voxel's own corners
        if (v < 0.0f && inside) {
            CellBox b;
            for (int k = 0; k < 8; ++k) {
                const int s0 = (k & 1) ? ns - 1 : 0;
                const int s1 = (k & 2) ? ns - 1 : 0;
                const int s2 = (k & 4) ? ns - 1 : 0;
                CellBox c;
                c.init(corner_image(P, sample_pos(P, 0, x, off[0], s0),
                                    sample_pos(P, 1, y, off[1], s1),
                                    sample_pos(P, 2, z, off[2], s2)));
                if (k == 0)
                    b = c;
                else
                    b.add(c);
                if (!b.ok())
                    break;
            }
            v = b.value(P);
        }
        // box values are 0, 1 or unknown (-1): a known box adds v / n^3
        // n^3 times, which is 0 or the open fraction
        double blocked = v > 0.0f ? P.blocked_one : 0.0;

        // 3. the sub-samples of the voxels still open: spread over the
        // lanes when at most two rounds of voxels need them, else each lane
        // loops over its own (more voxels than that fill the lanes anyway)
        const bool need = inside && v < 0.0f;
        const uint64_t need_mask = __ballot(need);
        const int n_need = __popcll(need_mask);
        const int G = n3i <= kSampleBuf ? kSampleBuf / n3i : 0;
        if (n_need > 0 && n_need <= 2 * G) {
            double* buf = s_samp;
            int32_t* lanes = s_lane;
            const int rank = __popcll(need_mask & ((1ull << l) - 1ull));
            if (need)
                lanes[rank] = l;
            wave_lds_sync();
            // k / n3 and s / ns by float reciprocals: for k, s < 512 and
            // divisors <= 512 the product (k + 0.5) / d is off by
            // < 512 * 2^-23 < 0.5 / 512, less than its distance to an
            // integer, so the truncation is the integer quotient
            const float inv_n3 = P.inv_n3;
            const float inv_ns = P.inv_ns;
            for (int g0 = 0; g0 < n_need; g0 += G) {
                const int gn = min(G, n_need - g0);
                for (int k = l; k < gn * n3i; k += 64) {
                    const int j = (int)(((float)k + 0.5f) * inv_n3);
                    const int smp = k - j * n3i;
                    const int q = (int)(((float)smp + 0.5f) * inv_ns);
                    const int iz = smp - q * ns;
                    const int ix = (int)(((float)q + 0.5f) * inv_ns);
                    const int iy = q - ix * ns;
                    const int src = lanes[g0 + j];
                    const double p0 =
                        sample_pos(P, 0, x0 + (src >> 4), off[0], ix);
                    const double p1 =
                        sample_pos(P, 1, y0 + ((src >> 2) & 3), off[1], iy);
                    const double p2 =
                        sample_pos(P, 2, z0 + (src & 3), off[2], iz);
                    double tv = 0.0;  // a sample in the source plane adds 0
                    double fi, fj;
                    if (fast_cells(P, p0, p1, p2, fi, fj)) {
                        tv = (double)transmission_at(P, fi, fj) / n3;
                    } else {
                        const PlaneImage c =
                            plane_image(P, p0, p1, p2, ndotn_src);
                        if (fabs(c.ndotu) < 1e-6)
                            bad = true;
                        else
                            tv = (double)transmission(P, c.ci, c.cj) / n3;
                    }
                    buf[k] = tv;
                }
                wave_lds_sync();
                if (need && rank >= g0 && rank < g0 + gn) {
                    const double* mine = buf + (rank - g0) * n3i;
                    for (int smp = 0; smp < n3i; ++smp)
                        blocked = blocked + mine[smp];
                }
                wave_lds_sync();
            }
        } else if (need) {
            for (int ix = 0; ix < ns; ++ix)
                for (int iy = 0; iy < ns; ++iy)
                    for (int iz = 0; iz < ns; ++iz) {
                        const double p0 = sample_pos(P, 0, x, off[0], ix);
                        const double p1 = sample_pos(P, 1, y, off[1], iy);
                        const double p2 = sample_pos(P, 2, z, off[2], iz);
                        double fi, fj;
                        if (fast_cells(P, p0, p1, p2, fi, fj)) {
                            blocked = blocked +
                                      (double)transmission_at(P, fi, fj) / n3;
                            continue;
                        }
                        const PlaneImage c =
                            plane_image(P, p0, p1, p2, ndotn_src);
                        if (fabs(c.ndotu) < 1e-6) {
                            bad = true;
                            continue;
                        }
                        blocked = blocked +
                                  (double)transmission(P, c.ci, c.cj) / n3;
                    }
        }

        const bool dbg = P.debug != nullptr;
        if (inside && (blocked != 0.0 || dbg)) {
            // radiological depth toward the source (conehead.py:209-254)
            double rv[3], dir[3], mag, pos2[3];
            ray_to_source(P, x, y, z, pos2, rv, mag, dir);
            const double r0 = rv[0], r1 = rv[1], r2 = rv[2];
            Dda d;
            d.init(dir[0], dir[1], dir[2], P.spacing[0], P.spacing[1],
                   P.spacing[2]);
            const float de =
                depth_walk_t<PACK>(P.density, P.X, P.Y, P.Z, x, y, z, d);

            // off-axis distance (conehead.py:286-314)
            double oad = 0.0;
            const double ndotr = dot3(P.vy, r0, r1, r2);
            if (fabs(ndotr) < 1e-6) {
                bad = true;
            } else {
                const double si = -ndotn_src / ndotr;
                const double pp0 = P.src[0] + si * r0 + 0.0;
                const double pp1 = P.src[1] + si * r1 + 0.0;
                const double pp2 = P.src[2] + si * r2 + 0.0;
                const double q0 = dot3(P.tr0, pp0, pp1, pp2);
                const double q2 = dot3(P.tr2, pp0, pp1, pp2);
                oad = sqrt(q0 * q0 + q2 * q2);
            }

            // fluence (conehead.py:320-360)
            const float magf = (float)mag, oadf = (float)oad;
            const float sad = (float)P.sad;
            const float g = sad / magf;
            const float fp = P.sPri * (g * g);
            const double r_ann = oad * P.zAnn64 / P.sad;
            float fa = 0.0f;
            if (r_ann >= P.rInner && r_ann <= P.rOuter) {
                const float a = sad - P.zAnn, b = magf - P.zAnn;
                fa = P.sAnn * (a * a) / (b * b);
            }
            const float oc = oadf < 2.0f ? 2.0f : oadf;
            const float r_exp = oc * P.zExp / sad;
            const float ea = sad - P.zExp, eb = magf - P.zExp;
            const float fe = P.sExp / r_exp * expf(-P.kExp * r_exp) *
                             (ea * ea) / (eb * eb);
            const float bl = (float)blocked;
            const float flu = (fp + fa + fe) * bl;

            // softening, horn (conehead.py:641-655), TERMA (:364-377)
            const float fs = oad < P.softLimit
                                 ? 1.0f / (1.0f - P.softRatio * oadf)
                                 : 1.0f;
            const float fh = 1.0f + oadf * P.hornRatio;
            const float sd = fs * de;
            float acc = 0.0f;
            for (int i = 0; i < P.n_energy; ++i)
                acc += P.ew[i] * flu * expf(-P.emu[i] * sd);
            T = acc * fh * bl;
            if (dbg) {
                const int64_t li = ((int64_t)x * P.Y + y) * P.Z + z;
                P.debug[li] = bl;
                P.debug[nv + li] = de;
                P.debug[2 * nv + li] = oadf;
                P.debug[3 * nv + li] = flu;
            }
        }
        if (inside) {
            const int64_t li = ((int64_t)x * P.Y + y) * P.Z + z;
            P.terma[li] = P.accumulate ? P.terma[li] + T : T;
        }
    return T;
}

// Phase 2: wavefront g takes the g-th brick of phase 1's lists in order of
// descending cost class, so the longest depth walks are dispatched first
// and the rest fill in behind them (wavefronts past the last brick exit at
// once).  Each wavefront writes its support-box partial.
template <bool PACK>
__device__ __forceinline__ void terma_wave(const TermaParams& P, int g,
                                           int l, double* __restrict__ s_samp,
                                           int32_t* __restrict__ s_lane,
                                           int32_t n_bricks)
{
    const int32_t* __restrict__ list = P.list;
    // Lane l holds positions j = 4 l .. 4 l + 3 of the counters in order of
    // descending cost class, shards in order (j = (kCostClasses - 1 - q) *
    // kShards + shard): counters q * kShards + 4 (l & 3) + i, i = 0 .. 3,
    // four consecutive ints read as one 16-byte load, with the same four of
    // both base rows (chosen by the call's parity once the loads are in, so
    // no load waits for another).  The count of a list is its counter minus
    // the base (the previous call's total).
    const int c0 = (kCostClasses - 1 - (l >> 2)) * kShards + 4 * (l & 3);
    const uint4 cnt = *(const uint4*)(P.counts + c0);
    const uint4 b0 = *(const uint4*)(P.counts + kBaseRows + c0);
    const uint4 b1 = *(const uint4*)(P.counts + 2 * kBaseRows + c0);
    const uint32_t call = (uint32_t)P.counts[kCallWord];
    const uint4 b = (call & 1) ? b1 : b0;
    if (g == 0) {
        // this call's totals become the next call's base (the other row)
        *(uint4*)(P.counts + kBaseRows * (1 + (~call & 1)) + c0) = cnt;
    }
    const uint32_t cv[4] = {cnt.x, cnt.y, cnt.z, cnt.w};
    const uint32_t bv[4] = {b.x, b.y, b.z, b.w};
    int c4[4], sum = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        c4[i] = (int)(cv[i] - bv[i]);
        sum += c4[i];
    }
    // incl = inclusive prefix of the lanes' sums
    int incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (l >= o)
            incl += u;
    }
    const int total = __shfl(incl, 63, 64);
    float T = 0.0f;
    int x = INT_MAX, y = 0, z = 0;
    if (g < total) {  // wavefront-uniform
        const uint64_t past = __ballot(incl > g);
        const int owner = __ffsll((unsigned long long)past) - 1;
        int before = incl - sum, pick = 0, rank = 0;
        uint32_t pbase = 0;
        bool found = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (!found && g < before + c4[i]) {
                pick = c0 + i;
                rank = g - before;
                pbase = bv[i];
                found = true;
            }
            before += c4[i];
        }
        const int c = __shfl(pick, owner, 64);
        rank = __shfl(rank, owner, 64);
        pbase = (uint32_t)__shfl((int)pbase, owner, 64);
        const int32_t cap = shard_cap(P.X, P.Y, P.Z);
        const uint32_t at = (pbase + (uint32_t)rank) & (uint32_t)(cap - 1);
        const int e = list[(int64_t)c * cap + at];
        // an entry of this call is a brick index of this grid; anything else
        // (a list slot no phase 1 of this state wrote) is skipped
        if ((unsigned)(e >> 2) < (unsigned)n_bricks) {  // wavefront-uniform
            const Brick k = brick_of(P, e >> 2);
            const float v =
                (e & 3) == 1 ? 1.0f : ((e & 3) == 2 ? -1.0f : 0.0f);
            bool bad = false;
            T = terma_brick<PACK>(P, k, v, l, s_samp, s_lane, bad);
            if (__ballot(bad) != 0 && l == 0)
                atomicOr(P.status, 1);
            x = k.x0 + (l >> 4);
            y = k.y0 + ((l >> 2) & 3);
            z = k.z0 + (l & 3);
        }
    }
    if (P.box) {
        // one partial box per wavefront; box_finalize reduces them
        int32_t bx[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN,
                         INT_MIN};
        if (T != 0.0f) {
            bx[0] = bx[3] = x;
            bx[1] = bx[4] = y;
            bx[2] = bx[5] = z;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bx[a] = min(bx[a], __shfl_xor(bx[a], o, 64));
                bx[3 + a] = max(bx[3 + a], __shfl_xor(bx[3 + a], o, 64));
            }
        int32_t mine = bx[0];
#pragma unroll
        for (int a = 1; a < 6; ++a)
            mine = l == a ? bx[a] : mine;
        if (l < 6)
            P.box[(int64_t)g * kBoxStride + l] = mine;
    }
}

// PACK: every axis of the grid has at most kPackMax voxels (depth_walk_t).
template <bool PACK>
__global__ __launch_bounds__(256) void terma_kernel(TermaParams P,
                                                    int32_t n_bricks)
{
    __shared__ double s_samp[4][kSampleBuf];
    __shared__ int32_t s_lane[4][64];
    const int w = threadIdx.x >> 6;
    terma_wave<PACK>(P, blockIdx.x * 4 + w, threadIdx.x & 63, s_samp[w],
                     s_lane[w], n_bricks);
}

// Phase 2 of several beams in one launch, beam by beam (the blocks of beam
// i are i * blocks_per_beam ...): the wavefronts in flight then serve one
// or two beams, whose rays toward one source cross the same density lines
// and read one aperture table, so they share the caches; the next beam's
// longest walks start while the last of the previous beam finish.
template <bool PACK>
__global__ __launch_bounds__(256) void terma_set_kernel(
    const TermaParams* __restrict__ tab, int32_t n_bricks,
    int32_t blocks_per_beam)
{
    __shared__ double s_samp[4][kSampleBuf];
    __shared__ int32_t s_lane[4][64];
    const int w = threadIdx.x >> 6;
    const int beam = blockIdx.x / blocks_per_beam;
    const int g = (blockIdx.x - beam * blocks_per_beam) * 4 + w;
    terma_wave<PACK>(tab[beam], g, threadIdx.x & 63, s_samp[w], s_lane[w],
               n_bricks);
}

__global__ __launch_bounds__(64) void terma_prep_kernel(int32_t* status)
{
    if (threadIdx.x == 0)
        status[0] = 0;
}

// Zeroes the beams' counter states before phase 1 where the scratch is
// shared by every call on the device (a kernel, so the call can be captured
// in a graph).
__global__ __launch_bounds__(256) void terma_clear_kernel(int32_t* counts,
                                                          int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        counts[i] = 0;
}

static TermaParams make_params(const ch_beam& B, const int32_t* h_shape,
                               const double* h_origin,
                               const double* h_spacing)
{
    TermaParams P;
    memset(&P, 0, sizeof P);
    for (int a = 0; a < 3; ++a) {
        P.src[a] = B.source_position[a];
        P.vy[a] = B.v_y[a];
        P.tr0[a] = B.transform[a];
        P.tr2[a] = B.transform[6 + a];
        P.origin[a] = h_origin[a];
        P.spacing[a] = h_spacing[a];
    }
    P.sad = B.sad;
    P.sPri = (float)B.sPri;
    P.sAnn = (float)B.sAnn;
    P.zAnn = (float)B.zAnn;
    P.zAnn64 = B.zAnn;
    P.rInner = B.rInner;
    P.rOuter = B.rOuter;
    P.zExp = (float)B.zExp;
    P.sExp = (float)B.sExp;
    P.kExp = (float)B.kExp;
    P.softRatio = (float)B.softRatio;
    P.softLimit = B.softLimit;
    P.hornRatio = (float)B.hornRatio;
    P.samples = B.samples;
    P.n_energy = B.n_energy;
    P.X = h_shape[0];
    P.Y = h_shape[1];
    P.Z = h_shape[2];
    for (int i = 0; i < B.n_energy; ++i) {
        P.ew[i] = (float)(B.weight[i] * B.energy[i] * B.mu[i]);
        P.emu[i] = (float)B.mu[i];
    }
    const int ns = B.samples > 0 ? B.samples : 1;
    for (int a = 0; a < 3; ++a)
        P.off[a] = P.spacing[a] / ns;
    P.ndotn_src = P.vy[0] * P.src[0] + P.vy[1] * P.src[1] + P.vy[2] * P.src[2];
    P.n3 = (double)(ns * ns * ns);
    // the hit test's loop adds (double)v / n^3 n^3 times, here for v = 1
    const double inc = 1.0 / P.n3;
    double one = 0.0;
    for (int k = 0; k < ns * ns * ns; ++k)
        one = one + inc;
    P.blocked_one = one;
    P.inv_n3 = 1.0f / (float)(ns * ns * ns);
    P.inv_ns = 1.0f / (float)ns;
    for (int a = 0; a < 3; ++a)
        P.src_vox[a] = (float)((P.src[a] - P.origin[a]) / P.spacing[a]);
    P.ci_src = 100.0 * (P.tr0[0] * P.src[0] + P.tr0[1] * P.src[1] +
                        P.tr0[2] * P.src[2]);
    P.cj_src = 100.0 * (P.tr2[0] * P.src[0] + P.tr2[1] * P.src[1] +
                        P.tr2[2] * P.src[2]);
    P.k100 = -100.0 * P.ndotn_src;
    P.fast_ok = fabs(P.ci_src) < 1e5 && fabs(P.cj_src) < 1e5;
    return P;
}

// Device memory TERMA launches need beyond their inputs and outputs, per
// beam of one launch: a counter state (kCounterInts, kept between calls),
// the brick lists of phase 1 and, with a support box, one partial box per
// phase-2 wavefront.
inline int32_t terma_bricks(const int32_t* shape)
{
    return ((shape[0] + 3) / 4) * ((shape[1] + 3) / 4) *
           ((shape[2] + 3) / 4);
}
inline size_t terma_list_ints(const int32_t* shape)
{
    return (size_t)kCounters * shard_cap(shape[0], shape[1], shape[2]);
}
inline size_t terma_list_bytes(int32_t n_beams, const int32_t* shape)
{
    return (size_t)n_beams * terma_list_ints(shape) * sizeof(int32_t);
}
// The beams' counter states (kCounterInts each), zero when allocated.
inline size_t terma_counter_bytes(int32_t n_beams)
{
    return (size_t)n_beams * kCounterInts * sizeof(int32_t);
}
inline size_t terma_box_ints(int32_t n_bricks)
{
    return (size_t)((n_bricks + 3) / 4) * 4 * kBoxStride;
}
inline size_t terma_box_bytes(int32_t n_beams, int32_t n_bricks)
{
    return (size_t)n_beams * terma_box_ints(n_bricks) * sizeof(int32_t);
}

// Points beam i of a launch at its part of the scratch (counters: the
// counter states; lists: the cost-class lists; boxes: partial boxes, null
// without any support box).
static void terma_assign(TermaParams& P, const ch_beam& B, int i,
                         const int32_t* shape, int32_t* counters,
                         int32_t* lists, int32_t* boxes)
{
    const int32_t n_bricks = terma_bricks(shape);
    P.counts = counters + (size_t)i * kCounterInts;
    P.list = lists + (size_t)i * terma_list_ints(shape);
    P.box = B.d_box && boxes ? boxes + (size_t)i * terma_box_ints(n_bricks)
                             : nullptr;
}

// Beams that share one launch: runs of consecutive beams with distinct
// TERMA grids (a beam that adds into the grid of an earlier beam of the
// run starts the next launch, so the two stay ordered), at most kSetBeams
// beams and kSetScratchBytes of scratch per launch.  Returns the start of
// every launch followed by n.
constexpr int kSetBeams = 64;
constexpr size_t kSetScratchBytes = (size_t)512 << 20;
static std::vector<int32_t> terma_groups(int32_t n, const ch_beam* beams,
                                         const int32_t* shape)
{
    const size_t per = terma_list_bytes(1, shape) +
                       terma_box_bytes(1, terma_bricks(shape)) +
                       terma_counter_bytes(1);
    const int32_t cap = (int32_t)std::max<size_t>(
        1, std::min<size_t>(kSetBeams, kSetScratchBytes / per));
    std::vector<int32_t> starts{0};
    for (int32_t b = 1; b < n; ++b) {
        bool split = b - starts.back() >= cap;
        for (int32_t a = starts.back(); a < b && !split; ++a)
            split = beams[a].d_terma == beams[b].d_terma;
        if (split)
            starts.push_back(b);
    }
    starts.push_back(n);
    return starts;
}

static int32_t largest_group(const std::vector<int32_t>& starts)
{
    int32_t m = 0;
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        m = std::max(m, starts[g + 1] - starts[g]);
    return m;
}

// Checks beam b and fills its kernel parameters.
static ch_status terma_params(const ch_beam& B, int b, const int32_t* shape,
                              const double* origin, const double* spacing,
                              float* d_debug, int32_t* stream_flag,
                              TermaParams* out)
{
    const ch_aperture& A = B.aperture;
    CH_REQUIRE(B.d_density && B.d_terma, "beam %d: null device buffer", b);
    CH_REQUIRE((A.kind == CH_APERTURE_RASTER && A.d_raster) ||
                   (A.kind == CH_APERTURE_RECTS && A.n >= 0 &&
                    (A.n == 0 || A.d_table)) ||
                   (A.kind == CH_APERTURE_MLC && A.d_table),
               "beam %d: bad aperture (kind %d)", b, A.kind);
    CH_REQUIRE(B.samples >= 1 && B.samples <= 16,
               "beam %d: fluenceResampling %d", b, B.samples);
    CH_REQUIRE(B.n_energy >= 0 && B.n_energy <= CH_MAX_ENERGIES,
               "beam %d: %d energies", b, B.n_energy);
    const int64_t nv = (int64_t)shape[0] * shape[1] * shape[2];
    TermaParams P = make_params(B, shape, origin, spacing);
    P.ap_kind = A.kind;
    P.ap_n = A.n;
    P.ap_table = A.d_table;
    P.raster = A.d_raster;
    for (int q = 0; q < 4; ++q)
        P.ap_bbox[q] = A.bbox[q];
    P.density = B.d_density;
    P.terma = B.d_terma;
    P.accumulate = B.accumulate;
    P.debug = d_debug ? d_debug + (size_t)b * 4 * nv : nullptr;
    P.status = B.d_status ? B.d_status : stream_flag;
    *out = P;
    return CH_OK;
}

// The launches of n beams on one grid: phase 1, phase 2 and the support
// boxes (first the clearing of their counter states when `clear`).  h holds
// the beams' parameters with their scratch assigned; with d_tab (a device
// copy of h) the beams share each launch, without it (n == 1) the
// parameters go by value.  ev0 / ev1 (nullable) are attached to the start
// of phase 1 and the end of phase 2.
static ch_status terma_launch(const TermaParams* h, const TermaParams* d_tab,
                              int32_t n, const ch_beam* beams,
                              int32_t n_bricks, hipEvent_t ev0,
                              hipEvent_t ev1, hipStream_t s,
                              bool clear = false)
{
    const unsigned bpb = (unsigned)((n_bricks + 3) / 4);
    const unsigned spb = (unsigned)sort_blocks(h[0].X, h[0].Y, h[0].Z);
    const bool pack =
        h[0].X <= kPackMax && h[0].Y <= kPackMax && h[0].Z <= kPackMax;
    if (clear) {
        const int nc = n * kCounterInts;
        hipLaunchKernelGGL(terma_clear_kernel,
                           dim3((unsigned)(nc + 255) / 256), dim3(256), 0, s,
                           h[0].counts, nc);
    }
    if (!d_tab) {
        hipExtLaunchKernelGGL(terma_sort_kernel, dim3(spb), dim3(256), 0, s,
                              ev0, nullptr, 0, h[0]);
        hipExtLaunchKernelGGL(pack ? terma_kernel<true> : terma_kernel<false>,
                              dim3(bpb), dim3(256), 0, s, nullptr, ev1, 0,
                              h[0], n_bricks);
    } else {
        hipExtLaunchKernelGGL(terma_sort_set_kernel, dim3(spb * n),
                              dim3(256), 0, s, ev0, nullptr, 0, d_tab,
                              (int32_t)spb);
        hipExtLaunchKernelGGL(pack ? terma_set_kernel<true>
                                   : terma_set_kernel<false>,
                              dim3(bpb * n), dim3(256), 0, s, nullptr, ev1, 0,
                              d_tab, n_bricks, (int32_t)bpb);
    }
    CH_HIP(hipGetLastError());
    for (int32_t i = 0; i < n; ++i)
        if (beams[i].d_box) {
            launch_box_finalize(h[i].box, (int)bpb * 4, beams[i].d_box,
                                beams[i].accumulate ? 1 : 0, s);
            CH_HIP(hipGetLastError());
        }
    return CH_OK;
}

// Checks the shared grid arguments of a TERMA call.
static ch_status terma_grid_ok(const char* fn, const int32_t* shape,
                               const double* origin, const double* spacing)
{
    CH_REQUIRE(shape && origin && spacing, "%s: null pointer", fn);
    CH_REQUIRE(shape[0] > 0 && shape[1] > 0 && shape[2] > 0,
               "%s: empty grid", fn);
    const int64_t nv = (int64_t)shape[0] * shape[1] * shape[2];
    CH_REQUIRE(nv < (1LL << 30),
               "%s: grid too large (4 X Y Z must stay below 2^32)", fn);
    return CH_OK;
}

// The allocation-free variant's scratch: one array per device, resolved to
// a device address per device once (hipGetSymbolAddress) and used only
// through that address.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_ptr_mu;

static ch_status noalloc_scratch(void** out)
{
    int dev = 0;
    CH_HIP(hipGetDevice(&dev));
    CH_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_ptr_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        CH_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        CH_REQUIRE(ptr, "hipGetSymbolAddress gave no address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return CH_OK;
}

// Called by ch_init with the device current.
ch_status terma_init_device()
{
    void* ptr = nullptr;
    return noalloc_scratch(&ptr);
}

}  // namespace ch

// A device-resident set of beams on one grid: their kernel parameters in a
// device table, written when the set is created or updated, and the scratch
// of their launches, so ch_terma_beamset queues kernels only (it can be
// captured in a graph and replayed).
struct ch_beamset : ch::GraphHeld {
    int32_t n = 0, n_bricks = 0;
    int32_t shape[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0}, spacing[3] = {0, 0, 0};
    std::vector<ch::TermaParams> h;   // host copy of the table
    std::vector<ch_beam> beams;       // d_terma, d_box, accumulate per beam
    std::vector<int32_t> starts;      // launch groups (terma_groups)
    ch::TermaParams* d_tab = nullptr;
    int32_t* counters = nullptr;      // counter states, largest group
    int32_t* lists = nullptr;         // cost-class lists, largest group
    int32_t* boxes = nullptr;         // partial boxes (null without boxes)
    int32_t* flags = nullptr;         // geometry flag per beam (own d_status
                                      // overrides)
    ch::TermaParams* h_stage = nullptr;  // pinned staging of updates
    hipEvent_t staged = nullptr;         // the last update's copy
};

namespace ch {

static void beamset_release(ch_beamset* set)
{
    (void)hipDeviceSynchronize();  // a queued launch may still read it
    (void)hipFree(set->d_tab);
    (void)hipFree(set->counters);
    (void)hipFree(set->lists);
    (void)hipFree(set->boxes);
    (void)hipFree(set->flags);
    (void)hipHostFree(set->h_stage);
    if (set->staged)
        (void)hipEventDestroy(set->staged);
    delete set;
}

// Parameters of every beam of a set with their scratch assigned; the
// launch groups must be `want` when it is given (an update keeps the
// launches a captured graph has recorded).
static ch_status beamset_params(ch_beamset* set, int32_t n,
                                const ch_beam* beams,
                                const std::vector<int32_t>* want,
                                std::vector<TermaParams>& hp,
                                std::vector<int32_t>& starts)
{
    hp.resize(n);
    for (int32_t b = 0; b < n; ++b) {
        ch_status st = terma_params(beams[b], b, set->shape, set->origin,
                                    set->spacing, nullptr, set->flags + b,
                                    &hp[b]);
        if (st != CH_OK)
            return st;
    }
    starts = terma_groups(n, beams, set->shape);
    if (want) {
        CH_REQUIRE(starts == *want,
                   "ch_beamset_update: the beams share their TERMA grids "
                   "differently from the set's (create a new set)");
        for (int32_t b = 0; b < n; ++b)
            CH_REQUIRE((beams[b].d_box != nullptr) ==
                           (set->beams[b].d_box != nullptr),
                       "ch_beamset_update: beam %d adds or drops its "
                       "support box (create a new set)", b);
    }
    for (size_t g = 0; g + 1 < starts.size(); ++g)
        for (int32_t b = starts[g]; b < starts[g + 1]; ++b)
            terma_assign(hp[b], beams[b], b - starts[g], set->shape,
                         set->counters, set->lists, set->boxes);
    return CH_OK;
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_terma(int32_t n_beams, const ch_beam* h_beams,
                   const int32_t h_shape[3], const double h_origin[3],
                   const double h_spacing[3], float* d_debug, void* stream)
{
    CH_REQUIRE(n_beams >= 0, "n_beams < 0");
    if (n_beams == 0)
        return CH_OK;
    CH_REQUIRE(h_beams, "ch_terma: null pointer");
    ch_status st = terma_grid_ok("ch_terma", h_shape, h_origin, h_spacing);
    if (st != CH_OK)
        return st;
    hipStream_t s = (hipStream_t)stream;
    // The stream's geometry flag stays set until ch_terma_check reads it, so
    // one check covers every call queued since the last one; it is zeroed
    // when it is allocated.  A beam with its own d_status uses that instead.
    void* status = nullptr;
    st = stream_scratch(s, kScratchStatus, 256, &status, true);
    if (st != CH_OK)
        return st;
    std::vector<TermaParams> hp(n_beams);
    bool any_box = false;
    for (int32_t b = 0; b < n_beams; ++b) {
        st = terma_params(h_beams[b], b, h_shape, h_origin, h_spacing,
                          d_debug, (int32_t*)status, &hp[b]);
        if (st != CH_OK)
            return st;
        any_box = any_box || h_beams[b].d_box;
    }
    // Runs of beams share each launch (their parameters staged into a
    // device table); under capture, where a staged table would be
    // recorded, every beam takes its own launches.
    const bool cap = capturing(s);
    const int32_t n_bricks = terma_bricks(h_shape);
    const std::vector<int32_t> starts =
        terma_groups(n_beams, h_beams, h_shape);
    const int32_t width = cap ? 1 : largest_group(starts);
    void* lists = nullptr;
    st = stream_scratch(s, kScratchTermaList,
                        terma_list_bytes(width, h_shape), &lists);
    if (st != CH_OK)
        return st;
    void* counters = nullptr;
    st = stream_scratch(s, kScratchTermaCounters, terma_counter_bytes(width),
                        &counters, true);
    if (st != CH_OK)
        return st;
    void* boxes = nullptr;
    if (any_box) {
        st = stream_scratch(s, kScratchBoxPartials,
                            terma_box_bytes(width, n_bricks), &boxes);
        if (st != CH_OK)
            return st;
    }
    // the caller's armed event pair (ch_set_launch_events) brackets the
    // first phase 1 and the last phase 2
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!cap)
        (void)take_launch_events(&ev0, &ev1);
    for (size_t g = 0; g + 1 < starts.size(); ++g) {
        const int32_t g0 = starts[g], ng = starts[g + 1] - g0;
        const bool last_group = g + 2 == starts.size();
        if (ng == 1 || cap) {
            for (int32_t i = 0; i < ng; ++i) {
                const int32_t b = g0 + i;
                terma_assign(hp[b], h_beams[b], 0, h_shape,
                             (int32_t*)counters, (int32_t*)lists,
                             (int32_t*)boxes);
                st = terma_launch(&hp[b], nullptr, 1, &h_beams[b], n_bricks,
                                  b == 0 ? ev0 : nullptr,
                                  b == n_beams - 1 ? ev1 : nullptr, s);
                if (st != CH_OK)
                    return st;
            }
            continue;
        }
        for (int32_t i = 0; i < ng; ++i)
            terma_assign(hp[g0 + i], h_beams[g0 + i], i, h_shape,
                         (int32_t*)counters, (int32_t*)lists,
                         (int32_t*)boxes);
        void* d_tab = nullptr;
        st = stage_table(s, &hp[g0], (size_t)ng * sizeof(TermaParams),
                         &d_tab);
        if (st != CH_OK)
            return st;
        st = terma_launch(&hp[g0], (const TermaParams*)d_tab, ng,
                          &h_beams[g0], n_bricks, g == 0 ? ev0 : nullptr,
                          last_group ? ev1 : nullptr, s);
        if (st != CH_OK)
            return st;
    }
    return CH_OK;
}

ch_status ch_terma_noalloc(int32_t n_beams, const ch_beam* h_beams,
                           const int32_t h_shape[3],
                           const double h_origin[3],
                           const double h_spacing[3], float* d_debug,
                           void* stream)
{
    CH_REQUIRE(n_beams >= 0, "n_beams < 0");
    if (n_beams == 0)
        return CH_OK;
    CH_REQUIRE(h_beams, "ch_terma_noalloc: null pointer");
    ch_status st =
        terma_grid_ok("ch_terma_noalloc", h_shape, h_origin, h_spacing);
    if (st != CH_OK)
        return st;
    const int32_t n_bricks = terma_bricks(h_shape);
    const size_t need = terma_counter_bytes(1) + terma_list_bytes(1, h_shape) +
                        terma_box_bytes(1, n_bricks);
    CH_REQUIRE(need <= kNoAllocScratchBytes,
               "ch_terma_noalloc: a %d x %d x %d grid needs %zu bytes of "
               "scratch, more than its %zu (use ch_terma)",
               h_shape[0], h_shape[1], h_shape[2], need,
               kNoAllocScratchBytes);
    for (int32_t b = 0; b < n_beams; ++b)
        CH_REQUIRE(h_beams[b].d_status,
                   "ch_terma_noalloc: beam %d has no d_status (the geometry "
                   "flag is the caller's output here)", b);
    void* base = nullptr;
    st = noalloc_scratch(&base);
    if (st != CH_OK)
        return st;
    // counter state, lists, boxes; the state is cleared by every call (the
    // array is shared by every call on the device)
    int32_t* counters = (int32_t*)base;
    int32_t* lists = (int32_t*)((char*)base + terma_counter_bytes(1));
    int32_t* boxes = (int32_t*)((char*)base + terma_counter_bytes(1) +
                                terma_list_bytes(1, h_shape));
    hipStream_t s = (hipStream_t)stream;
    // one beam at a time through the one scratch array (no device table)
    for (int32_t b = 0; b < n_beams; ++b) {
        TermaParams P;
        st = terma_params(h_beams[b], b, h_shape, h_origin, h_spacing,
                          d_debug, nullptr, &P);
        if (st != CH_OK)
            return st;
        terma_assign(P, h_beams[b], 0, h_shape, counters, lists, boxes);
        st = terma_launch(&P, nullptr, 1, &h_beams[b], n_bricks, nullptr,
                          nullptr, s, true);
        if (st != CH_OK)
            return st;
    }
    return CH_OK;
}

ch_status ch_beamset_create(int32_t n_beams, const ch_beam* h_beams,
                            const int32_t h_shape[3],
                            const double h_origin[3],
                            const double h_spacing[3], ch_beamset** out)
{
    CH_REQUIRE(out, "ch_beamset_create: null output");
    *out = nullptr;
    CH_REQUIRE(n_beams > 0 && h_beams, "ch_beamset_create: no beams");
    ch_status st =
        terma_grid_ok("ch_beamset_create", h_shape, h_origin, h_spacing);
    if (st != CH_OK)
        return st;
    ch_beamset* set = new ch_beamset();
    set->release = [](GraphHeld* h) {
        beamset_release(static_cast<ch_beamset*>(h));
    };
    set->n = n_beams;
    set->n_bricks = terma_bricks(h_shape);
    for (int a = 0; a < 3; ++a) {
        set->shape[a] = h_shape[a];
        set->origin[a] = h_origin[a];
        set->spacing[a] = h_spacing[a];
    }
    set->beams.assign(h_beams, h_beams + n_beams);
    bool any_box = false;
    for (int32_t b = 0; b < n_beams; ++b)
        any_box = any_box || h_beams[b].d_box;
    const std::vector<int32_t> starts =
        terma_groups(n_beams, h_beams, set->shape);
    const int32_t width = largest_group(starts);
    hipError_t e = hipMalloc((void**)&set->d_tab,
                             (size_t)n_beams * sizeof(TermaParams));
    if (e == hipSuccess)
        e = hipMalloc((void**)&set->lists,
                      terma_list_bytes(width, set->shape));
    if (e == hipSuccess)
        e = hipMalloc((void**)&set->counters, terma_counter_bytes(width));
    if (e == hipSuccess)
        e = hipMemset(set->counters, 0, terma_counter_bytes(width));
    if (e == hipSuccess && any_box)
        e = hipMalloc((void**)&set->boxes,
                      terma_box_bytes(width, set->n_bricks));
    if (e == hipSuccess)
        e = hipMalloc((void**)&set->flags, (size_t)n_beams * sizeof(int32_t));
    if (e == hipSuccess)
        e = hipMemset(set->flags, 0, (size_t)n_beams * sizeof(int32_t));
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&set->h_stage,
                          (size_t)n_beams * sizeof(TermaParams),
                          hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&set->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        beamset_release(set);
        return hip_fail(e, "ch_beamset_create", __FILE__, __LINE__);
    }
    std::vector<int32_t> got;
    st = beamset_params(set, n_beams, h_beams, nullptr, set->h, got);
    if (st != CH_OK) {
        beamset_release(set);
        return st;
    }
    set->starts = got;
    e = hipMemcpy(set->d_tab, set->h.data(),
                  (size_t)n_beams * sizeof(TermaParams),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        beamset_release(set);
        return hip_fail(e, "ch_beamset_create", __FILE__, __LINE__);
    }
    *out = set;
    return CH_OK;
}

ch_status ch_beamset_update(ch_beamset* set, int32_t n_beams,
                            const ch_beam* h_beams, void* stream)
{
    CH_REQUIRE(set && h_beams, "ch_beamset_update: null pointer");
    CH_REQUIRE(n_beams == set->n,
               "ch_beamset_update: %d beams for a set of %d", n_beams,
               set->n);
    hipStream_t s = (hipStream_t)stream;
    CH_REQUIRE(!capturing(s), "ch_beamset_update copies from the host: "
                              "call it outside the capture");
    std::vector<TermaParams> hp;
    std::vector<int32_t> starts;
    ch_status st = beamset_params(set, n_beams, h_beams, &set->starts, hp,
                                  starts);
    if (st != CH_OK)
        return st;
    // the staging buffer is rewritten once the previous copy has read it;
    // launches queued on `stream` before this call read the old table
    CH_HIP(hipEventSynchronize(set->staged));
    memcpy(set->h_stage, hp.data(), (size_t)n_beams * sizeof(TermaParams));
    CH_HIP(hipMemcpyAsync(set->d_tab, set->h_stage,
                          (size_t)n_beams * sizeof(TermaParams),
                          hipMemcpyHostToDevice, s));
    CH_HIP(hipEventRecord(set->staged, s));
    set->h.swap(hp);
    set->beams.assign(h_beams, h_beams + n_beams);
    return CH_OK;
}

ch_status ch_terma_beamset(const ch_beamset* cset, void* stream)
{
    CH_REQUIRE(cset, "ch_terma_beamset: null set");
    ch_beamset* set = const_cast<ch_beamset*>(cset);
    hipStream_t s = (hipStream_t)stream;
    capture_hold(s, set);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (!capturing(s))
        (void)take_launch_events(&ev0, &ev1);
    const std::vector<int32_t>& starts = set->starts;
    for (size_t g = 0; g + 1 < starts.size(); ++g) {
        const int32_t g0 = starts[g], ng = starts[g + 1] - g0;
        ch_status st = terma_launch(&set->h[g0], set->d_tab + g0, ng,
                                    &set->beams[g0], set->n_bricks,
                                    g == 0 ? ev0 : nullptr,
                                    g + 2 == starts.size() ? ev1 : nullptr,
                                    s);
        if (st != CH_OK)
            return st;
    }
    return CH_OK;
}

ch_status ch_beamset_flags(ch_beamset* set, int32_t* h_flags, void* stream)
{
    CH_REQUIRE(set && h_flags, "ch_beamset_flags: null pointer");
    hipStream_t s = (hipStream_t)stream;
    CH_REQUIRE(!capturing(s), "ch_beamset_flags waits for the stream");
    CH_HIP(hipMemcpyAsync(h_flags, set->flags, (size_t)set->n * 4,
                          hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(terma_clear_kernel,
                       dim3((unsigned)(set->n + 255) / 256), dim3(256), 0, s,
                       set->flags, set->n);
    CH_HIP(hipGetLastError());
    CH_HIP(hipStreamSynchronize(s));
    return CH_OK;
}

ch_status ch_beamset_destroy(ch_beamset* set)
{
    if (!set)
        return CH_OK;
    if (defer_free(set))  // read by a live graph: freed with the last one
        return CH_OK;
    beamset_release(set);
    return CH_OK;
}

ch_status ch_terma_rays(const ch_beam* h_beam, const int32_t h_shape[3],
                        const double h_origin[3], const double h_spacing[3],
                        int32_t n_rays, const int32_t* d_vox_in, int32_t cap,
                        double* d_dirs, int32_t* d_vox, double* d_t,
                        int32_t* d_count, void* stream)
{
    CH_REQUIRE(h_beam && h_shape && h_origin && h_spacing,
               "ch_terma_rays: null pointer");
    CH_REQUIRE(n_rays >= 0 && cap >= 0, "ch_terma_rays: bad sizes");
    if (n_rays == 0)
        return CH_OK;
    CH_REQUIRE(d_vox_in && d_dirs && d_vox && d_t && d_count,
               "ch_terma_rays: null device buffer");
    TermaParams P = make_params(*h_beam, h_shape, h_origin, h_spacing);
    hipLaunchKernelGGL(terma_rays_kernel,
                       dim3((unsigned)((n_rays + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, P, n_rays, d_vox_in, cap, d_dirs,
                       d_vox, d_t, d_count);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_terma_check(void* stream)
{
    void* status = stream_scratch_peek((hipStream_t)stream, kScratchStatus);
    if (!status)
        return CH_OK;
    int32_t flag = 0;
    CH_HIP(hipMemcpyAsync(&flag, status, sizeof flag,
                          hipMemcpyDeviceToHost, (hipStream_t)stream));
    CH_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (flag) {
        hipLaunchKernelGGL(terma_prep_kernel, dim3(1), dim3(64), 0,
                           (hipStream_t)stream, (int32_t*)status);
        CH_HIP(hipStreamSynchronize((hipStream_t)stream));
        set_error("no intersection or line is within plane");
        return CH_ERR_GEOMETRY;
    }
    return CH_OK;
}

}  // extern "C"
