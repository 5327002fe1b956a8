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
};

__device__ __forceinline__ double dot3(const double* a, double b0, double b1,
                                       double b2)
{
    return a[0] * b0 + a[1] * b1 + a[2] * b2;
}

// conehead/conehead.py:50-68: raster cell floor(100 p) + 2000 - 1 of the
// isocentre-plane point (a -1 index wraps to the last row/column as
// Python/Numba negative indexing does), looked up in the beam's aperture.
__device__ __forceinline__ float transmission(const TermaParams& P, double px,
                                              double pz)
{
    px = floor(px * 100.0) + 2000.0;
    pz = floor(pz * 100.0) + 2000.0;
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

// Thread layout: a workgroup covers a 4 x 4 x 16 brick of voxels, each
// wavefront a 4 x 4 x 4 cube, so a wavefront's voxels are either all well
// inside, all outside or on the edge of the field (the corner shortcut of
// the hit test pays for whole wavefronts) and their rays toward the source
// have similar lengths.
__global__ __launch_bounds__(256) void terma_kernel(TermaParams P)
{
    const int64_t nv = (int64_t)P.X * P.Y * P.Z;
    const int nbz = (P.Z + 15) / 16, nby = (P.Y + 3) / 4;
    const int bz = blockIdx.x % nbz;
    const int br = blockIdx.x / nbz;
    const int l = threadIdx.x & 63;
    const int x = (br / nby) * 4 + (l >> 4);
    const int y = (br % nby) * 4 + ((l >> 2) & 3);
    const int z = bz * 16 + (threadIdx.x >> 6) * 4 + (l & 3);
    const bool inside = x < P.X && y < P.Y && z < P.Z;
    const int64_t li = ((int64_t)x * P.Y + y) * P.Z + z;
    float T = 0.0f;
    if (inside) {
        double pos[3];
        pos[0] = P.origin[0] + P.spacing[0] * x;
        pos[1] = P.origin[1] + P.spacing[1] * y;
        pos[2] = P.origin[2] + P.spacing[2] * z;

        // hit test (conehead.py:78-120)
        const int ns = P.samples;
        const double n3 = (double)(ns * ns * ns);
        double off[3];
        for (int a = 0; a < 3; ++a)
            off[a] = P.spacing[a] / ns;
        const double ndotn_src = dot3(P.vy, P.src[0], P.src[1], P.src[2]);
        double blocked = 0.0;
        bool bad = false;
        // Early out: if the voxel centre projects so far outside the
        // aperture's bounding cells that no sub-sample can reach them, the
        // blocked fraction is exactly 0.  With u = s - p, a = n.u and
        // A = n.s, the plane image is s - A u / a; moving p by |d| <= h (half
        // the voxel diagonal) moves it by at most
        //   |A| h / |a'| (1 + |u'| / |a'|),  |a'| >= |a| - h, |u'| <= |u| + h.
        // For |a| >= 4 h and |u| <= sqrt(5) |a| that is <= 6 |A| h / |a|;
        // raster cells are plane coordinates times 100 (a rotation keeps
        // lengths), plus 4 cells for the floor and index offsets.
        bool skip = false;
        {
            const double h = 0.5 * sqrt(P.spacing[0] * P.spacing[0] +
                                        P.spacing[1] * P.spacing[1] +
                                        P.spacing[2] * P.spacing[2]);
            const double u0 = P.src[0] - pos[0], u1 = P.src[1] - pos[1],
                         u2 = P.src[2] - pos[2];
            const double ndotu = dot3(P.vy, u0, u1, u2);
            const double uu = u0 * u0 + u1 * u1 + u2 * u2;
            if (fabs(ndotu) >= 4.0 * h + 1.0 && uu <= 5.0 * ndotu * ndotu) {
                const double si = -ndotn_src / ndotu;
                const double pp0 = P.src[0] + si * u0;
                const double pp1 = P.src[1] + si * u1;
                const double pp2 = P.src[2] + si * u2;
                const double ci = dot3(P.tr0, pp0, pp1, pp2) * 100.0 + 1999.0;
                const double cj = dot3(P.tr2, pp0, pp1, pp2) * 100.0 + 1999.0;
                const double m =
                    600.0 * h * fabs(ndotn_src / ndotu) + 4.0;
                skip = ci + m < P.ap_bbox[0] || ci - m > P.ap_bbox[1] ||
                       cj + m < P.ap_bbox[2] || cj - m > P.ap_bbox[3];
            }
        }
        // Corner shortcut: the sub-samples form a lattice whose points are
        // convex combinations of its 8 corners, and the projection through
        // the source onto the isocentre plane keeps convex combinations
        // while n.u keeps one sign (n.u is affine in the point).  So when
        // the 8 corner images (computed exactly as below) have one sign of
        // n.u and their cell box, widened by 1e-6 cell for rounding, has one
        // constant transmission, every sample has it and the blocked
        // fraction is that value added n^3 times, as the loop would.
        bool done = skip;
        if (!skip && P.ap_kind != CH_APERTURE_RASTER) {
            double lo_i = 1e300, hi_i = -1e300, lo_j = 1e300, hi_j = -1e300;
            int sgn = 0;
            bool ok = true;
            for (int k = 0; k < 8 && ok; ++k) {
                const int ix = (k & 1) ? ns - 1 : 0;
                const int iy = (k & 2) ? ns - 1 : 0;
                const int iz = (k & 4) ? ns - 1 : 0;
                const double ps0 = pos[0] - P.spacing[0] / 2 + off[0] / 2 + off[0] * ix;
                const double ps1 = pos[1] - P.spacing[1] / 2 + off[1] / 2 + off[1] * iy;
                const double ps2 = pos[2] - P.spacing[2] / 2 + off[2] / 2 + off[2] * iz;
                const double u0 = P.src[0] - ps0, u1 = P.src[1] - ps1,
                             u2 = P.src[2] - ps2;
                const double ndotu = dot3(P.vy, u0, u1, u2);
                const int sg = ndotu > 0 ? 1 : -1;
                if (fabs(ndotu) < 1e-6 || (sgn != 0 && sg != sgn)) {
                    ok = false;
                    break;
                }
                sgn = sg;
                const double si = -ndotn_src / ndotu;
                const double pp0 = P.src[0] + si * u0 + 0.0;
                const double pp1 = P.src[1] + si * u1 + 0.0;
                const double pp2 = P.src[2] + si * u2 + 0.0;
                const double ci = dot3(P.tr0, pp0, pp1, pp2) * 100.0;
                const double cj = dot3(P.tr2, pp0, pp1, pp2) * 100.0;
                lo_i = fmin(lo_i, ci);
                hi_i = fmax(hi_i, ci);
                lo_j = fmin(lo_j, cj);
                hi_j = fmax(hi_j, cj);
            }
            // cells floor(c) + 1999 of the widened box, inside the raster
            // (no wrap-around, no out-of-raster zeros)
            const double a_i = floor(lo_i - 1e-6) + 1999.0;
            const double b_i = floor(hi_i + 1e-6) + 1999.0;
            const double a_j = floor(lo_j - 1e-6) + 1999.0;
            const double b_j = floor(hi_j + 1e-6) + 1999.0;
            if (ok && a_i >= 0.0 && b_i <= 3998.0 && a_j >= 0.0 &&
                b_j <= 3998.0) {
                const int i0 = (int)a_i, i1 = (int)b_i, j0 = (int)a_j,
                          j1 = (int)b_j;
                const float v =
                    P.ap_kind == CH_APERTURE_MLC
                        ? mlc_block_value(P.ap_table, i0, i1, j0, j1)
                        : rects_block_value(P.ap_table, P.ap_n, i0, i1, j0,
                                            j1);
                if (v >= 0.0f) {
                    const double inc = (double)v / n3;
                    for (int k = 0; k < ns * ns * ns; ++k)
                        blocked = blocked + inc;
                    done = true;
                }
            }
        }
        for (int ix = 0; ix < ns && !done; ++ix)
            for (int iy = 0; iy < ns; ++iy)
                for (int iz = 0; iz < ns; ++iz) {
                    const double ps0 = pos[0] - P.spacing[0] / 2 + off[0] / 2 + off[0] * ix;
                    const double ps1 = pos[1] - P.spacing[1] / 2 + off[1] / 2 + off[1] * iy;
                    const double ps2 = pos[2] - P.spacing[2] / 2 + off[2] / 2 + off[2] * iz;
                    const double u0 = P.src[0] - ps0, u1 = P.src[1] - ps1,
                                 u2 = P.src[2] - ps2;
                    const double ndotu = dot3(P.vy, u0, u1, u2);
                    if (fabs(ndotu) < 1e-6) {
                        bad = true;
                        continue;
                    }
                    const double si = -ndotn_src / ndotu;
                    const double pp0 = P.src[0] + si * u0 + 0.0;
                    const double pp1 = P.src[1] + si * u1 + 0.0;
                    const double pp2 = P.src[2] + si * u2 + 0.0;
                    const double q0 = dot3(P.tr0, pp0, pp1, pp2);
                    const double q2 = dot3(P.tr2, pp0, pp1, pp2);
                    blocked = blocked + transmission(P, q0, q2) / n3;
                }

        const bool dbg = P.debug != nullptr;
        if (blocked != 0.0 || dbg) {
            // radiological depth toward the source (conehead.py:209-254)
            double rv[3], dir[3], mag, pos2[3];
            ray_to_source(P, x, y, z, pos2, rv, mag, dir);
            const double r0 = rv[0], r1 = rv[1], r2 = rv[2];
            Dda d;
            d.init(dir[0], dir[1], dir[2], P.spacing[0], P.spacing[1],
                   P.spacing[2]);
            // Only the axis that steps can leave the grid: keep the linear
            // index and each axis's steps left to the boundary instead of
            // recomputing the index and testing all six bounds per voxel
            // (the same t-values, voxels and sums).
            int32_t c0 = x, c1 = y, c2 = z;
            int64_t lin = ((int64_t)x * P.Y + y) * P.Z + z;
            const int64_t st0 = (int64_t)d.s0 * P.Y * P.Z;
            const int64_t st1 = (int64_t)d.s1 * P.Z;
            const int64_t st2 = d.s2;
            int32_t left0 = d.s0 > 0 ? P.X - 1 - x : x;
            int32_t left1 = d.s1 > 0 ? P.Y - 1 - y : y;
            int32_t left2 = d.s2 > 0 ? P.Z - 1 - z : z;
            double tprev = 0.0;
            float de = 0.0f, dc = 0.0f;
            while (true) {
                const float rho = P.density[lin];
                const int a = d.axis();
                const double t = d.advance(a, c0, c1, c2);
                const float term = rho * (float)(t - tprev);
                tprev = t;
                const float yk = term - dc;
                const float tk = de + yk;
                dc = (tk - de) - yk;
                de = tk;
                int32_t left;
                if (a == 0) {
                    left = --left0;
                    lin += st0;
                } else if (a == 1) {
                    left = --left1;
                    lin += st1;
                } else {
                    left = --left2;
                    lin += st2;
                }
                if (left < 0)
                    break;
            }

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
                P.debug[li] = bl;
                P.debug[nv + li] = de;
                P.debug[2 * nv + li] = oadf;
                P.debug[3 * nv + li] = flu;
            }
        }
        P.terma[li] = P.accumulate ? P.terma[li] + T : T;
        if (bad)
            atomicOr(P.status, 1);
    }
    if (P.box) {
        // one partial box per block; box_finalize reduces them (no atomics)
        int32_t v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
        if (inside && T != 0.0f) {
            v[0] = v[3] = x;
            v[1] = v[4] = y;
            v[2] = v[5] = z;
        }
        block_box_partial(v, P.box);
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
    CH_REQUIRE(nv < (1LL << 31), "ch_terma: grid too large");
    hipStream_t s = (hipStream_t)stream;
    // The geometry flag is per stream and covers every beam of this call.
    // The geometry flag stays set until ch_terma_check reads it, so one
    // check covers every call queued since the last one; it is cleared
    // here only when the stream's flag is first allocated.
    const bool fresh = stream_scratch_peek(s, kScratchStatus) == nullptr;
    void* status = nullptr;
    ch_status sst = stream_scratch(s, kScratchStatus, 256, &status);
    if (sst != CH_OK)
        return sst;
    if (fresh)
        hipLaunchKernelGGL(terma_prep_kernel, dim3(1), dim3(64), 0, s,
                           (int32_t*)status);
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
        P.status = (int32_t*)status;
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
