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
#include <climits>
#include <cmath>
#include <cstring>

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
    int32_t* box;
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

// True when no raster cell of [a, b] (cells floor(c) + 1999, as doubles)
// can read a value inside the aperture's bounding cells [lo, hi).  Cells
// outside [-1, 3998] read 0 (conehead.py:53-67) and cell -1 reads the last
// row or column (Python's negative indexing).
__device__ __forceinline__ bool axis_misses(double a, double b, int lo,
                                            int hi)
{
    if (b < -1.0 || a > 3998.0)
        return true;
    const double A = fmax(a, 0.0), B = fmin(b, 3998.0);
    const bool in = A <= B && B >= (double)lo && A <= (double)(hi - 1);
    const bool wrap = a <= -1.0 && 3999 >= lo && 3999 <= hi - 1;
    return !(in || wrap);
}

// The transmission shared by every sub-sample whose cell coordinates lie
// in [lo_i, hi_i] x [lo_j, hi_j] (widened by 1e-6 cell for rounding), or -1
// when it is not one known constant.
__device__ __forceinline__ float box_value(const TermaParams& P, double lo_i,
                                           double hi_i, double lo_j,
                                           double hi_j)
{
    const double a_i = floor(lo_i - 1e-6) + 1999.0;
    const double b_i = floor(hi_i + 1e-6) + 1999.0;
    const double a_j = floor(lo_j - 1e-6) + 1999.0;
    const double b_j = floor(hi_j + 1e-6) + 1999.0;
    if (axis_misses(a_i, b_i, P.ap_bbox[0], P.ap_bbox[1]) ||
        axis_misses(a_j, b_j, P.ap_bbox[2], P.ap_bbox[3]))
        return 0.0f;
    if (P.ap_kind == CH_APERTURE_RASTER ||
        !(a_i >= 0.0 && b_i <= 3998.0 && a_j >= 0.0 && b_j <= 3998.0))
        return -1.0f;
    const int i0 = (int)a_i, i1 = (int)b_i, j0 = (int)a_j, j1 = (int)b_j;
    return P.ap_kind == CH_APERTURE_MLC
               ? mlc_block_value(P.ap_table, i0, i1, j0, j1)
               : rects_block_value(P.ap_table, P.ap_n, i0, i1, j0, j1);
}

// Extent of a set of plane images: cell box, and whether every n.u has
// one sign and |n.u| >= 1e-6 (then the central projection maps the convex
// hull of the points into the convex hull of their images).
struct ImageBox {
    double lo_i, hi_i, lo_j, hi_j;
    bool pos, neg, ok;
    __device__ __forceinline__ void init(const PlaneImage& c)
    {
        lo_i = hi_i = c.ci;
        lo_j = hi_j = c.cj;
        ok = fabs(c.ndotu) >= 1e-6;
        pos = c.ndotu > 0;
        neg = !pos;
    }
    __device__ __forceinline__ void add(const ImageBox& o)
    {
        lo_i = fmin(lo_i, o.lo_i);
        hi_i = fmax(hi_i, o.hi_i);
        lo_j = fmin(lo_j, o.lo_j);
        hi_j = fmax(hi_j, o.hi_j);
        ok = ok && o.ok;
        pos = pos && o.pos;
        neg = neg && o.neg;
    }
    __device__ __forceinline__ float value(const TermaParams& P) const
    {
        return ok && (pos || neg) ? box_value(P, lo_i, hi_i, lo_j, hi_j)
                                  : -1.0f;
    }
};

__device__ __forceinline__ ImageBox shfl_xor_box(const ImageBox& b, int m)
{
    ImageBox o;
    o.lo_i = __shfl_xor(b.lo_i, m, 64);
    o.hi_i = __shfl_xor(b.hi_i, m, 64);
    o.lo_j = __shfl_xor(b.lo_j, m, 64);
    o.hi_j = __shfl_xor(b.hi_j, m, 64);
    o.ok = __shfl_xor((int)b.ok, m, 64) != 0;
    o.pos = __shfl_xor((int)b.pos, m, 64) != 0;
    o.neg = __shfl_xor((int)b.neg, m, 64) != 0;
    return o;
}

// conehead.py:50-68 at cell coordinates (times 100, not floored) of an
// image: raster cell floor(c) + 2000 - 1, a -1 index wrapping to the last
// row or column, looked up in the beam's aperture.
__device__ __forceinline__ float transmission(const TermaParams& P, double ci,
                                              double cj)
{
    const double px = floor(ci) + 2000.0;
    const double pz = floor(cj) + 2000.0;
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
// path order (conehead.py:249-254).  The step needs no branch: the tie
// rules (x if t0 < t1 and t0 < t2, else y if t1 <= t0 and t1 < t2, else z)
// become three masks that select the crossing taken and the axis that
// advances.  kWalkBatch steps are traversed first and their density loads
// issued together before the first is used; lanes past the grid exit keep
// stepping harmlessly on voxel 0 with their sums frozen.  Byte offsets are
// 32-bit (ch_terma requires 4 X Y Z < 2^32).
constexpr int kWalkBatch = 8;

__device__ __forceinline__ float depth_walk(const float* __restrict__ rho,
                                            int X, int Y, int Z, int x, int y,
                                            int z, const Dda& d)
{
    double t0 = d.t0, t1 = d.t1, t2 = d.t2;
    const double dt0 = d.dt0, dt1 = d.dt1, dt2 = d.dt2;
    // steps left before each axis leaves the grid
    int32_t left0 = d.s0 > 0 ? X - 1 - x : x;
    int32_t left1 = d.s1 > 0 ? Y - 1 - y : y;
    int32_t left2 = d.s2 > 0 ? Z - 1 - z : z;
    const int32_t st0 = d.s0 * Y * Z, st1 = d.s1 * Z, st2 = d.s2;
    int32_t lin = (x * Y + y) * Z + z;
    double tprev = 0.0;
    float de = 0.0f, dc = 0.0f;
    bool live = true;
    const char* base = (const char*)rho;
    while (live) {
        uint32_t off[kWalkBatch];
        float seg[kWalkBatch];
        bool use[kWalkBatch];
#pragma unroll
        for (int k = 0; k < kWalkBatch; ++k) {
            use[k] = live;
            const bool c01 = t0 < t1, c02 = t0 < t2, c12 = t1 < t2;
            const bool m0 = c01 & c02;
            const bool m1 = !c01 & c12;
            const bool m2 = !(m0 | m1);
            const double t = m0 ? t0 : (m1 ? t1 : t2);
            off[k] = live ? 4u * (uint32_t)lin : 0u;
            seg[k] = (float)(t - tprev);
            tprev = t;
            const double n0 = t0 + dt0, n1 = t1 + dt1, n2 = t2 + dt2;
            t0 = m0 ? n0 : t0;
            t1 = m1 ? n1 : t1;
            t2 = m2 ? n2 : t2;
            lin += m0 ? st0 : (m1 ? st1 : st2);
            left0 -= (int32_t)m0;
            left1 -= (int32_t)m1;
            left2 -= (int32_t)m2;
            live = live & ((left0 | left1 | left2) >= 0);
        }
        float r[kWalkBatch];
#pragma unroll
        for (int k = 0; k < kWalkBatch; ++k)
            r[k] = *(const float*)(base + off[k]);
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

// Thread layout: a workgroup covers a 4 x 4 x 16 brick of voxels, each
// wavefront a 4 x 4 x 4 cube, so a wavefront's voxels are either all well
// inside, all outside or on the edge of the field and their rays toward
// the source have similar lengths.
//
// Hit test (conehead.py:78-120), in three stages that each give the exact
// blocked fraction of the loop over n^3 sub-samples:
//  1. the brick: 8 lanes project the 8 corners of the box spanned by every
//     sub-sample of the wavefront's 64 voxels.  The sub-samples are convex
//     combinations of those corners, and the projection through the source
//     onto the isocentre plane keeps convex combinations while n.u keeps
//     one sign (n.u is affine in the point), so when the corner images
//     have one sign of n.u and their cell box (widened by 1e-6 cell for
//     rounding) reads one known transmission v -- 0 outside the aperture's
//     bounding cells or the raster, or a constant of the leaf table or the
//     rectangles -- every sub-sample of the brick reads v and each voxel's
//     fraction is v / n^3 added n^3 times, as the loop adds it;
//  2. otherwise each voxel applies the same test to its own 8 corners;
//  3. the voxels still open have their sub-samples spread over all 64
//     lanes (a few voxels at a time), each lane evaluating one sub-sample
//     into LDS; each voxel then sums its own values in the loop's order,
//     so the result is bitwise that of the loop.  When most of the lanes
//     need samples the loop runs on each lane instead.
__global__ __launch_bounds__(256) void terma_kernel(TermaParams P)
{
    __shared__ double s_samp[4][kSampleBuf];
    __shared__ int32_t s_lane[4][64];
    const int64_t nv = (int64_t)P.X * P.Y * P.Z;
    const int nbz = (P.Z + 15) / 16, nby = (P.Y + 3) / 4;
    const int bz = blockIdx.x % nbz;
    const int br = blockIdx.x / nbz;
    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int x0 = (br / nby) * 4, y0 = (br % nby) * 4, z0 = bz * 16 + w * 4;
    const int x = x0 + (l >> 4);
    const int y = y0 + ((l >> 2) & 3);
    const int z = z0 + (l & 3);
    const bool inside = x < P.X && y < P.Y && z < P.Z;
    float T = 0.0f;
    if (z0 < P.Z) {  // wavefront-uniform: the brick holds grid voxels
        const int ns = P.samples;
        const int n3i = ns * ns * ns;
        const double n3 = P.n3;
        const double* off = P.off;
        const double ndotn_src = P.ndotn_src;
        bool bad = false;

        // 1. the brick's sub-sample box, corner (l & 7) per lane
        float v;
        {
            const int k = l & 7;
            const int xa = (k & 1) ? min(x0 + 3, P.X - 1) : x0;
            const int ya = (k & 2) ? min(y0 + 3, P.Y - 1) : y0;
            const int za = (k & 4) ? min(z0 + 3, P.Z - 1) : z0;
            const int s0 = (k & 1) ? ns - 1 : 0;
            const int s1 = (k & 2) ? ns - 1 : 0;
            const int s2 = (k & 4) ? ns - 1 : 0;
            ImageBox b;
            b.init(plane_image(P, sample_pos(P, 0, xa, off[0], s0),
                               sample_pos(P, 1, ya, off[1], s1),
                               sample_pos(P, 2, za, off[2], s2), ndotn_src));
            b.add(shfl_xor_box(b, 1));
            b.add(shfl_xor_box(b, 2));
            b.add(shfl_xor_box(b, 4));
            v = b.value(P);
        }

        // 2. the voxel's own corners
        if (v < 0.0f && inside) {
            ImageBox b;
            for (int k = 0; k < 8; ++k) {
                const int s0 = (k & 1) ? ns - 1 : 0;
                const int s1 = (k & 2) ? ns - 1 : 0;
                const int s2 = (k & 4) ? ns - 1 : 0;
                ImageBox c;
                c.init(plane_image(P, sample_pos(P, 0, x, off[0], s0),
                                   sample_pos(P, 1, y, off[1], s1),
                                   sample_pos(P, 2, z, off[2], s2),
                                   ndotn_src));
                if (k == 0)
                    b = c;
                else
                    b.add(c);
                if (!b.ok || !(b.pos || b.neg))
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
            double* buf = s_samp[w];
            int32_t* lanes = s_lane[w];
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
                    const PlaneImage c = plane_image(
                        P, sample_pos(P, 0, x0 + (src >> 4), off[0], ix),
                        sample_pos(P, 1, y0 + ((src >> 2) & 3), off[1], iy),
                        sample_pos(P, 2, z0 + (src & 3), off[2], iz),
                        ndotn_src);
                    double tv = 0.0;  // a sample in the source plane adds 0
                    if (fabs(c.ndotu) < 1e-6)
                        bad = true;
                    else
                        tv = (double)transmission(P, c.ci, c.cj) / n3;
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
                        const PlaneImage c = plane_image(
                            P, sample_pos(P, 0, x, off[0], ix),
                            sample_pos(P, 1, y, off[1], iy),
                            sample_pos(P, 2, z, off[2], iz), ndotn_src);
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
                depth_walk(P.density, P.X, P.Y, P.Z, x, y, z, d);

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
        if (__ballot(bad) != 0 && l == 0)
            atomicOr(P.status, 1);
    }
    if (P.box) {
        // one partial box per block; box_finalize reduces them (no atomics)
        int32_t bv[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
        if (inside && T != 0.0f) {
            bv[0] = bv[3] = x;
            bv[1] = bv[4] = y;
            bv[2] = bv[5] = z;
        }
        block_box_partial(bv, P.box);
    }
}

__global__ __launch_bounds__(64) void terma_prep_kernel(int32_t* status)
{
    if (threadIdx.x == 0)
        status[0] = 0;
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
    return P;
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
    CH_REQUIRE(h_beams && h_shape && h_origin && h_spacing,
               "ch_terma: null pointer");
    CH_REQUIRE(h_shape[0] > 0 && h_shape[1] > 0 && h_shape[2] > 0,
               "ch_terma: empty grid");
    const int64_t nv = (int64_t)h_shape[0] * h_shape[1] * h_shape[2];
    CH_REQUIRE(nv < (1LL << 30), "ch_terma: grid too large (4 X Y Z must "
                                 "stay below 2^32)");
    hipStream_t s = (hipStream_t)stream;
    // The stream's geometry flag stays set until ch_terma_check reads it, so
    // one check covers every call queued since the last one; it is zeroed
    // when it is allocated.  A beam with its own d_status uses that instead.
    void* status = nullptr;
    ch_status sst = stream_scratch(s, kScratchStatus, 256, &status, true);
    if (sst != CH_OK)
        return sst;
    for (int32_t b = 0; b < n_beams; ++b) {
        const ch_beam& B = h_beams[b];
        const ch_aperture& A = B.aperture;
        CH_REQUIRE(B.d_density && B.d_terma, "beam %d: null device buffer",
                   b);
        CH_REQUIRE((A.kind == CH_APERTURE_RASTER && A.d_raster) ||
                       (A.kind == CH_APERTURE_RECTS && A.n >= 0 &&
                        (A.n == 0 || A.d_table)) ||
                       (A.kind == CH_APERTURE_MLC && A.d_table),
                   "beam %d: bad aperture (kind %d)", b, A.kind);
        CH_REQUIRE(B.samples >= 1 && B.samples <= 16,
                   "beam %d: fluenceResampling %d", b, B.samples);
        CH_REQUIRE(B.n_energy >= 0 && B.n_energy <= CH_MAX_ENERGIES,
                   "beam %d: %d energies", b, B.n_energy);
        TermaParams P = make_params(B, h_shape, h_origin, h_spacing);
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
        P.status = B.d_status ? B.d_status : (int32_t*)status;
        const unsigned blocks = (unsigned)(((h_shape[0] + 3) / 4) *
                                           ((h_shape[1] + 3) / 4) *
                                           ((h_shape[2] + 15) / 16));
        P.box = nullptr;
        if (B.d_box) {
            void* parts = nullptr;
            ch_status st = stream_scratch(s, kScratchBoxPartials,
                                          (size_t)blocks * kBoxStride *
                                              sizeof(int32_t),
                                          &parts);
            if (st != CH_OK)
                return st;
            P.box = (int32_t*)parts;
        }
        hipLaunchKernelGGL(terma_kernel, dim3(blocks), dim3(256), 0, s, P);
        CH_HIP(hipGetLastError());
        if (B.d_box) {
            launch_box_finalize(P.box, (int)blocks, B.d_box,
                                B.accumulate ? 1 : 0, s);
            CH_HIP(hipGetLastError());
        }
    }
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
