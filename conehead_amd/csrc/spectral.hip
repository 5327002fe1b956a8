// Spectral form of the convolution: dose = T (*) K on zero-padded frames.
//
// convolve_c (conehead/convolve_c.pyx:77-192) adds, for every source voxel
// s with T != 0 and every cone c, k_cm * T(s) to the voxel s + o_cm of the
// cone's m-th DDA step.  The offsets o_cm and increments k_cm depend only
// on the cone (the plan's tap tables), so the dose is ONE linear 3D
// convolution of T with the kernel K(o) = sum over (c, m) with o_cm = o of
// k_cm, taken with T zero outside the grid.  Every offset component lies in
// (-n_a, n_a), so on a frame of F_a >= 2 n_a - 1 points per axis the
// circular convolution equals the linear one on the grid:
//   dose = IDFT( DFT(T zero-padded) * DFT(K wrapped) )[grid].
// The plan's taps are merged into K in FP64 on the host; K's transform
// (K^, FP32, scaled by 1 / (Fx Fy Fz)) is built once per plan on the GPU.
//
// Each beam then costs five HBM-bound passes over a complex grid of
// X x Fy x Kp points (Kp >= Fz / 2 + 1), nothing of O(taps):
//   z_fwd   real lines of T, two per complex FFT, split into half spectra;
//   col     y transforms of the x < X slabs (rows y >= Y are zero);
//   col     x transforms, times K^, inverse x transforms, rows x < X kept;
//   col     inverse y transforms, rows y < Y kept;
//   z_inv   two half spectra per complex inverse FFT into the dose lines.
// The transforms are mixed-radix (2, 3, 4, 5, 8, 16) Stockham FFTs in LDS,
// W lines per workgroup; frame lengths are products of 2, 3 and 5.  Where
// the frame is at least twice the grid, half of every forward transform's
// input is padding and half of every inverse transform's output is cropped:
// the first stage then skips those rows (half_in) and the last stage those
// outputs (half_out), when its radix is even.
#include "ch_internal.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace ch {
namespace {

constexpr int kFftThreads = 256;
constexpr int kFftMax = 1024;      // longest transform
// F * W per LDS buffer: the convolution pass takes up to 16 lines of 256
// (its radix-16 stages then keep all 256 threads busy and K^ is read in
// 128-byte rows); the other passes run faster with half of that.  Two
// buffers (ping-pong) unless single_buffer() allows one.
constexpr int kFftLdsElemsConv = 4096;
constexpr int kFftLdsElems = 2048;
// F * W a pass takes when more lines are asked for (the autotuner): with
// one buffer, 16 lines of 256 points keep all 256 threads busy in every
// radix-16 stage, which pays when many beams share a launch.
constexpr int kFftLdsElemsMax = 4096;
// Elements of one line group per thread in the unfused path (transforms
// of one stage, F <= 16, so F * W <= 256 and one per thread covers it).
constexpr int kPerThread = 8;
constexpr int kPerThreadConv = kFftLdsElemsConv / kFftThreads;
// Work grids one batched launch may use (beams beyond it take further
// launches).
constexpr size_t kSpectralWorkBytes = (size_t)4 << 30;

struct FftLen {
    int32_t n;
    int32_t nfac;
    int32_t fac[12];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b)
{
    return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b)
{
    return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s)
{
    return make_float2(a.x * s, a.y * s);
}
// -i * a for the forward transform, +i * a for the inverse
template <bool INV>
__device__ __forceinline__ float2 rot(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b)
{
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c,
                                     float2& d)
{
    const float2 t0 = cadd(a, c), t1 = csub(a, c);
    const float2 t2 = cadd(b, d), t3 = rot<INV>(csub(b, d));
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

// dft4 of (a, b, 0, 0), c and d not read: a + c and b + d are a and b, so
// every nonzero output has the bits dft4 gives it (only the sign of an
// exact zero may differ).
template <bool INV>
__device__ __forceinline__ void dft4h(float2& a, float2& b, float2& c,
                                      float2& d)
{
    const float2 t3 = rot<INV>(b);
    c = csub(a, b);
    d = csub(a, t3);
    const float2 a0 = a;
    a = cadd(a0, b);
    b = cadd(a0, t3);
}

// dft<INV, H>: the DFT of v in place.  H: the upper half v[R/2 ..] of an
// even-length v is zero and not read (the first layer of each
// decomposition drops its zero terms, as in dft4h).
template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[2])
{
    if (H)
        v[1] = v[0];
    else
        dft2<INV>(v[0], v[1]);
}

template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[4])
{
    if (H)
        dft4h<INV>(v[0], v[1], v[2], v[3]);
    else
        dft4<INV>(v[0], v[1], v[2], v[3]);
}

template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[8])
{
    // even and odd halves, then X[k] = E[k] + w^k O[k], X[k+4] = E - w^k O
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    if (H) {
        dft4h<INV>(e0, e1, e2, e3);
        dft4h<INV>(o0, o1, o2, o3);
    } else {
        dft4<INV>(e0, e1, e2, e3);
        dft4<INV>(o0, o1, o2, o3);
    }
    const float h = 0.70710678118654752f;
    // w^1 = (h, -h) forward, (h, h) inverse; w^3 = (-h, -h) / (-h, h)
    const float2 w1o = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
                           : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    const float2 w2o = rot<INV>(o2);
    const float2 w3o = INV
                           ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
                           : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1o);
    v[5] = csub(e1, w1o);
    v[2] = cadd(e2, w2o);
    v[6] = csub(e2, w2o);
    v[3] = cadd(e3, w3o);
    v[7] = csub(e3, w3o);
}

template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[3])
{
    static_assert(!H, "odd length");
    const float s3 = 0.86602540378443865f;  // sin(pi / 3)
    const float2 t1 = cadd(v[1], v[2]);
    const float2 t2 = csub(v[0], cscale(t1, 0.5f));
    const float2 t3 = rot<INV>(cscale(csub(v[1], v[2]), s3));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[5])
{
    static_assert(!H, "odd length");
    const float c1 = 0.30901699437494742f;   // cos(2 pi / 5)
    const float c2 = -0.80901699437494742f;  // cos(4 pi / 5)
    const float s1 = 0.95105651629515357f;   // sin(2 pi / 5)
    const float s2 = 0.58778525229247313f;   // sin(4 pi / 5)
    const float2 a = v[0];
    const float2 sbe = cadd(v[1], v[4]), dbe = csub(v[1], v[4]);
    const float2 scd = cadd(v[2], v[3]), dcd = csub(v[2], v[3]);
    const float2 a1 = cadd(a, cadd(cscale(sbe, c1), cscale(scd, c2)));
    const float2 a2 = cadd(a, cadd(cscale(sbe, c2), cscale(scd, c1)));
    const float2 b1 = rot<INV>(cadd(cscale(dbe, s1), cscale(dcd, s2)));
    const float2 b2 = rot<INV>(csub(cscale(dbe, s2), cscale(dcd, s1)));
    v[0] = cadd(a, cadd(sbe, scd));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <bool INV, bool H = false>
__device__ __forceinline__ void dft(float2 (&v)[16])
{
    // 4 x 4: DFT4 over n1 of x[4 n1 + n2], twiddle W16^(n2 k1), DFT4 over
    // n2, output k1 + 4 k2
    const float c8 = 0.92387953251128674f, s8 = 0.38268343236508977f;
    const float h = 0.70710678118654752f;
    const float sg = INV ? 1.0f : -1.0f;
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        if (H)
            dft4h<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
        else
            dft4<INV>(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    }
    // after this v[n2 + 4 k1] holds y[n2][k1]; W16^m = (cos, sg sin)
    const float2 w1 = make_float2(c8, sg * s8), w2 = make_float2(h, sg * h),
                 w3 = make_float2(s8, sg * c8), w6 = make_float2(-h, sg * h),
                 w9 = make_float2(-c8, -sg * s8);
    v[1 + 4] = cmul(v[1 + 4], w1);
    v[1 + 8] = cmul(v[1 + 8], w2);
    v[1 + 12] = cmul(v[1 + 12], w3);
    v[2 + 4] = cmul(v[2 + 4], w2);
    v[2 + 8] = rot<INV>(v[2 + 8]);
    v[2 + 12] = cmul(v[2 + 12], w6);
    v[3 + 4] = cmul(v[3 + 4], w3);
    v[3 + 8] = cmul(v[3 + 8], w6);
    v[3 + 12] = cmul(v[3 + 12], w9);
    float2 o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 a = v[4 * k1], b = v[4 * k1 + 1], c = v[4 * k1 + 2],
               d = v[4 * k1 + 3];
        dft4<INV>(a, b, c, d);
        o[k1] = a;
        o[k1 + 4] = b;
        o[k1 + 8] = c;
        o[k1 + 12] = d;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
        v[i] = o[i];
}

// One Stockham stage of radix R over W lines of length F in LDS, element
// (n, w) at n * W + w.  Sub-transforms of length Ns are done; this stage
// makes them Ns * R long, in natural order.  Each thread keeps one line
// (w = tid % W) and steps j by kFftThreads / W: no divisions.
template <int R, bool INV, int W>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ src,
                                          float2* __restrict__ dst, int F,
                                          int Ns, const float2* __restrict__ tw)
{
    const int nb = F / R;
    const int twstep = F / (Ns * R);
    const int w = threadIdx.x % W;
    const bool pow2 = (Ns & (Ns - 1)) == 0;
    for (int j = threadIdx.x / W; j < nb; j += kFftThreads / W) {
        const int jm = pow2 ? (j & (Ns - 1)) : (j % Ns);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = src[(j + r * nb) * W + w];
        if (Ns > 1) {
            const int step = jm * twstep;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 t = tw[r * step];
                if (INV)
                    t.y = -t.y;
                v[r] = cmul(v[r], t);
            }
        }
        dft<INV>(v);
        const int d0 = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r)
            dst[(d0 + r * Ns) * W + w] = v[r];
    }
}

// Full transform of the W lines held in `a`; returns the buffer (a or b)
// holding the result.  Ends with a barrier.
template <bool INV, int W>
__device__ float2* lds_fft(float2* a, float2* b, const float2* tw,
                           const FftLen& L)
{
    int Ns = 1;
    for (int s = 0; s < L.nfac; ++s) {
        const int R = L.fac[s];
        switch (R) {
        case 16:
            fft_stage<16, INV, W>(a, b, L.n, Ns, tw);
            break;
        case 8:
            fft_stage<8, INV, W>(a, b, L.n, Ns, tw);
            break;
        case 4:
            fft_stage<4, INV, W>(a, b, L.n, Ns, tw);
            break;
        case 2:
            fft_stage<2, INV, W>(a, b, L.n, Ns, tw);
            break;
        case 3:
            fft_stage<3, INV, W>(a, b, L.n, Ns, tw);
            break;
        default:
            fft_stage<5, INV, W>(a, b, L.n, Ns, tw);
            break;
        }
        __syncthreads();
        Ns *= R;
        float2* t = a;
        a = b;
        b = t;
    }
    return a;
}

// ---- stages fed from, or draining to, registers -------------------------
// A transform of factors f0 .. f(n-1) (n >= 2) runs its first stage from
// global memory (no LDS store of the input) and its last stage to global
// memory (no LDS read of the output); the convolution pass also fuses its
// forward last stage, the product with K^ and its inverse first stage,
// whose radices match because the inverse takes the factors in reverse.

// First stage (Ns = 1): rows n = j + r * F / R come from load(n).  H
// (R even): rows n >= F / 2, i.e. r >= R / 2, are zero padding and are
// neither loaded nor added.
template <int R, bool INV, int W, bool H, class Load>
__device__ __forceinline__ void first_stage(const Load& load,
                                            float2* __restrict__ dst, int F)
{
    const int nb = F / R;
    const int w = threadIdx.x % W;
    for (int j = threadIdx.x / W; j < nb; j += kFftThreads / W) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = (H && r >= R / 2) ? make_float2(0.0f, 0.0f)
                                     : load(j + r * nb);
        dft<INV, H>(v);
#pragma unroll
        for (int r = 0; r < R; ++r)
            dst[(j * R + r) * W + w] = v[r];
    }
}

// Last stage (Ns = F / R): outputs k = j + r * F / R go to store(k, v).
// H (R even): only outputs k < F / 2 (r < R / 2) are wanted; the others are
// neither stored nor computed.
template <int R, bool INV, int W, bool H, class Store>
__device__ __forceinline__ void last_stage(const float2* __restrict__ src,
                                           int F,
                                           const float2* __restrict__ tw,
                                           const Store& store)
{
    const int nb = F / R;
    const int w = threadIdx.x % W;
    for (int j = threadIdx.x / W; j < nb; j += kFftThreads / W) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = src[(j + r * nb) * W + w];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            float2 t = tw[r * j];
            if (INV)
                t.y = -t.y;
            v[r] = cmul(v[r], t);
        }
        dft<INV>(v);
#pragma unroll
        for (int r = 0; r < (H ? R / 2 : R); ++r)
            store(j + r * nb, v[r]);
    }
}

// Forward last stage, times K^ (kload(k) for output k), inverse first
// stage: LDS src -> LDS dst.
template <int R, int W, class KLoad>
__device__ __forceinline__ void conv_stage(const float2* __restrict__ src,
                                           float2* __restrict__ dst, int F,
                                           const float2* __restrict__ tw,
                                           const KLoad& kload)
{
    const int nb = F / R;
    const int w = threadIdx.x % W;
    for (int j = threadIdx.x / W; j < nb; j += kFftThreads / W) {
        float2 kv[R], v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            kv[r] = kload(j + r * nb);
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = src[(j + r * nb) * W + w];
#pragma unroll
        for (int r = 1; r < R; ++r)
            v[r] = cmul(v[r], tw[r * j]);
        dft<false>(v);
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = cmul(v[r], kv[r]);
        dft<true>(v);
#pragma unroll
        for (int r = 0; r < R; ++r)
            dst[(j * R + r) * W + w] = v[r];
    }
}

// In-place forms for a single LDS buffer, when every stage gives each
// thread at most one butterfly group ((F / R) * W <= kFftThreads): read
// into registers, barrier, write back.  Every thread reaches the barrier.
template <int R, bool INV, int W>
__device__ __forceinline__ void fft_stage_ip(float2* __restrict__ buf, int F,
                                             int Ns,
                                             const float2* __restrict__ tw)
{
    const int nb = F / R;
    const int twstep = F / (Ns * R);
    const int w = threadIdx.x % W;
    const int j = threadIdx.x / W;
    const bool act = j < nb;
    const int jm = (Ns & (Ns - 1)) == 0 ? (j & (Ns - 1)) : (j % Ns);
    float2 v[R];
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = buf[(j + r * nb) * W + w];
        if (Ns > 1) {
            const int step = jm * twstep;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 t = tw[r * step];
                if (INV)
                    t.y = -t.y;
                v[r] = cmul(v[r], t);
            }
        }
        dft<INV>(v);
    }
    __syncthreads();
    if (act) {
        const int d0 = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r)
            buf[(d0 + r * Ns) * W + w] = v[r];
    }
}

template <int R, int W, class KLoad>
__device__ __forceinline__ void conv_stage_ip(float2* __restrict__ buf, int F,
                                              const float2* __restrict__ tw,
                                              const KLoad& kload)
{
    const int nb = F / R;
    const int w = threadIdx.x % W;
    const int j = threadIdx.x / W;
    const bool act = j < nb;
    float2 v[R];
    if (act) {
        float2 kv[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            kv[r] = kload(j + r * nb);
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = buf[(j + r * nb) * W + w];
#pragma unroll
        for (int r = 1; r < R; ++r)
            v[r] = cmul(v[r], tw[r * j]);
        dft<false>(v);
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = cmul(v[r], kv[r]);
        dft<true>(v);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            buf[(j * R + r) * W + w] = v[r];
    }
}

template <int W, class KLoad>
__device__ __forceinline__ void conv_any_ip(int R, float2* buf, int F,
                                            const float2* tw,
                                            const KLoad& kload)
{
    switch (R) {
    case 16: conv_stage_ip<16, W>(buf, F, tw, kload); break;
    case 8: conv_stage_ip<8, W>(buf, F, tw, kload); break;
    case 4: conv_stage_ip<4, W>(buf, F, tw, kload); break;
    case 2: conv_stage_ip<2, W>(buf, F, tw, kload); break;
    case 3: conv_stage_ip<3, W>(buf, F, tw, kload); break;
    default: conv_stage_ip<5, W>(buf, F, tw, kload); break;
    }
}

// In-place LDS stages fac[s], s = s0, s0 + step, ... (stopping before s1);
// each ends with a barrier.
template <bool INV, int W>
__device__ void stages_range_ip(float2* buf, const float2* tw, int F,
                                const int32_t* fac, int s0, int s1, int step,
                                int& Ns)
{
    for (int s = s0; s != s1; s += step) {
        const int R = fac[s];
        switch (R) {
        case 16: fft_stage_ip<16, INV, W>(buf, F, Ns, tw); break;
        case 8: fft_stage_ip<8, INV, W>(buf, F, Ns, tw); break;
        case 4: fft_stage_ip<4, INV, W>(buf, F, Ns, tw); break;
        case 2: fft_stage_ip<2, INV, W>(buf, F, Ns, tw); break;
        case 3: fft_stage_ip<3, INV, W>(buf, F, Ns, tw); break;
        default: fft_stage_ip<5, INV, W>(buf, F, Ns, tw); break;
        }
        __syncthreads();
        Ns *= R;
    }
}

// H: rows n >= F / 2 read as zero and R is even (the half forms; R = 2
// stands for any radix the others do not name).
template <bool INV, int W, bool H, class Load>
__device__ __forceinline__ void first_any(int R, const Load& load,
                                          float2* dst, int F)
{
    if constexpr (H) {
        switch (R) {
        case 16: first_stage<16, INV, W, true>(load, dst, F); break;
        case 8: first_stage<8, INV, W, true>(load, dst, F); break;
        case 4: first_stage<4, INV, W, true>(load, dst, F); break;
        default: first_stage<2, INV, W, true>(load, dst, F); break;
        }
    } else {
        switch (R) {
        case 16: first_stage<16, INV, W, false>(load, dst, F); break;
        case 8: first_stage<8, INV, W, false>(load, dst, F); break;
        case 4: first_stage<4, INV, W, false>(load, dst, F); break;
        case 2: first_stage<2, INV, W, false>(load, dst, F); break;
        case 3: first_stage<3, INV, W, false>(load, dst, F); break;
        default: first_stage<5, INV, W, false>(load, dst, F); break;
        }
    }
}

// H: only outputs k < F / 2 are wanted and R is even.
template <bool INV, int W, bool H, class Store>
__device__ __forceinline__ void last_any(int R, const float2* src, int F,
                                         const float2* tw, const Store& st)
{
    if constexpr (H) {
        switch (R) {
        case 16: last_stage<16, INV, W, true>(src, F, tw, st); break;
        case 8: last_stage<8, INV, W, true>(src, F, tw, st); break;
        case 4: last_stage<4, INV, W, true>(src, F, tw, st); break;
        default: last_stage<2, INV, W, true>(src, F, tw, st); break;
        }
    } else {
        switch (R) {
        case 16: last_stage<16, INV, W, false>(src, F, tw, st); break;
        case 8: last_stage<8, INV, W, false>(src, F, tw, st); break;
        case 4: last_stage<4, INV, W, false>(src, F, tw, st); break;
        case 2: last_stage<2, INV, W, false>(src, F, tw, st); break;
        case 3: last_stage<3, INV, W, false>(src, F, tw, st); break;
        default: last_stage<5, INV, W, false>(src, F, tw, st); break;
        }
    }
}

// Whether a transform of length L reading rows n < n_in (writing outputs
// k < n_out) takes the half form of its first (last) stage: the frame is at
// least twice the grid and the stage's radix is even.  The kernels take the
// answer as a template flag, so each holds one form of the stage.
bool half_in(const FftLen& L, int n_in)
{
    return L.nfac >= 2 && 2 * n_in <= L.n && L.fac[0] % 2 == 0;
}
bool half_out(const FftLen& L, int n_out)
{
    return L.nfac >= 2 && 2 * n_out <= L.n && L.fac[L.nfac - 1] % 2 == 0;
}

template <int W, class KLoad>
__device__ __forceinline__ void conv_any(int R, const float2* src,
                                         float2* dst, int F, const float2* tw,
                                         const KLoad& kload)
{
    switch (R) {
    case 16: conv_stage<16, W>(src, dst, F, tw, kload); break;
    case 8: conv_stage<8, W>(src, dst, F, tw, kload); break;
    case 4: conv_stage<4, W>(src, dst, F, tw, kload); break;
    case 2: conv_stage<2, W>(src, dst, F, tw, kload); break;
    case 3: conv_stage<3, W>(src, dst, F, tw, kload); break;
    default: conv_stage<5, W>(src, dst, F, tw, kload); break;
    }
}

// LDS stages fac[s] for s = s0, s0 + step, ... (stopping before s1), Ns
// updated; each ends with a barrier.  Returns the buffer holding the data.
template <bool INV, int W>
__device__ float2* stages_range(float2* a, float2* b, const float2* tw,
                                int F, const int32_t* fac, int s0, int s1,
                                int step, int& Ns)
{
    for (int s = s0; s != s1; s += step) {
        const int R = fac[s];
        switch (R) {
        case 16: fft_stage<16, INV, W>(a, b, F, Ns, tw); break;
        case 8: fft_stage<8, INV, W>(a, b, F, Ns, tw); break;
        case 4: fft_stage<4, INV, W>(a, b, F, Ns, tw); break;
        case 2: fft_stage<2, INV, W>(a, b, F, Ns, tw); break;
        case 3: fft_stage<3, INV, W>(a, b, F, Ns, tw); break;
        default: fft_stage<5, INV, W>(a, b, F, Ns, tw); break;
        }
        __syncthreads();
        Ns *= R;
        float2* t = a;
        a = b;
        b = t;
    }
    return a;
}

__device__ __forceinline__ void load_twiddles(float2* s_tw, const float2* g,
                                              int F)
{
    for (int i = threadIdx.x; i < F; i += blockDim.x)
        s_tw[i] = g[i];
}

// The beams of one launch: TERMA and dose grids from device tables (a
// batch) or one pointer each (a single beam); work grid b at a_beam * b.
// Workgroups take blockIdx.x = group * n_beams + beam, so consecutive
// workgroups are the same columns of different beams (K^ chunks are read
// once per launch from HBM and then from L2).
struct BeamIO {
    const float* t;
    float* d;
    const float* const* t_tab;
    float* const* d_tab;
    const int32_t* acc_tab;
    int32_t accumulate;
    int32_t n_beams;
    int64_t a_beam;
};

__device__ __forceinline__ int beam_of_block(const BeamIO& io)
{
    return (int)(blockIdx.x % (unsigned)io.n_beams);
}
__device__ __forceinline__ int group_of_block(const BeamIO& io)
{
    return (int)(blockIdx.x / (unsigned)io.n_beams);
}

// Workgroups go to the 8 XCDs round-robin by linear id, and each XCD has
// its own L2.  The y passes take their work in XCD-major order instead:
// XCD k runs a contiguous range of logical blocks, so the two halves of a
// 128-byte line (neighbouring 8-column groups) meet in one L2.  The work
// per block is uniform, so the ranges stay balanced.
constexpr unsigned kXcds = 8;
struct BlockXY {
    unsigned x, y;
};
__device__ __forceinline__ BlockXY xcd_major_block(bool on)
{
    const unsigned nx = gridDim.x, total = gridDim.x * gridDim.y;
    unsigned l = blockIdx.x + nx * blockIdx.y;
    if (on && total % kXcds == 0)
        l = (l % kXcds) * (total / kXcds) + l / kXcds;
    BlockXY b;
    b.x = l % nx;
    b.y = l / nx;
    return b;
}

// The forward y pass (an even number of column groups per beam) puts the
// two halves of a 128-byte line (column groups 2 gp and 2 gp + 1 of one
// beam and slab) on one XCD, as consecutive blocks there, and deals these
// pairs to the XCDs round robin: the slabs whose transforms the support
// flags skip are then spread over all XCDs, where contiguous ranges would
// leave some XCDs idle.  Block x = group * n_beams + beam.  Other grids
// take xcd_major_block.
__device__ __forceinline__ BlockXY paired_block(unsigned n_beams)
{
    const unsigned nx = gridDim.x, total = gridDim.x * gridDim.y;
    if (nx % (2 * n_beams) != 0 || total % (2 * kXcds) != 0)
        return xcd_major_block(true);
    const unsigned h = blockIdx.x + nx * blockIdx.y;
    const unsigned k = h % kXcds, j = h / kXcds;
    const unsigned q = (j >> 1) * kXcds + k;   // pair (slab, gp, beam)
    const unsigned pairs = nx / (2 * n_beams);
    const unsigned beam = q % n_beams, rest = q / n_beams;
    BlockXY b;
    b.x = (2 * (rest % pairs) + (j & 1)) * n_beams + beam;
    b.y = rest / pairs;
    return b;
}

// Line pair w of a z-pass workgroup: lines (x, y) and (x, y + 1) of the
// grid, y even; valid = false past the last pair.
struct LinePair {
    int64_t x;
    int32_t y;
    bool valid;
};
__device__ __forceinline__ LinePair line_pair(int64_t p, int32_t X,
                                              int32_t Y)
{
    const int Yp = (Y + 1) / 2;
    LinePair lp;
    lp.valid = p < (int64_t)X * Yp;
    lp.x = lp.valid ? p / Yp : 0;
    lp.y = lp.valid ? 2 * (int)(p - lp.x * Yp) : 0;
    return lp;
}

// Zero support of a beam's work grid, so the forward passes skip what is
// zero: z_fwd records per line pair whether it holds a nonzero value (and
// writes the pair's spectra only then); the forward y pass reads those
// flags, skips a slab whose pairs are all zero (recording it in the slab
// flags) and reads only the nonzero rows of the others; the convolution
// pass reads only the rows x of nonzero slabs.  The flags are int32 at
// `off` (complex points after the beam's work grid): X * yp pair flags
// (pair p = x * yp + y / 2), then X slab flags at xyp = X * yp.  off < 0:
// no flags (every row is read).
struct Support {
    int64_t off;
    int32_t yp, xyp;
};
constexpr int kMaxFlagRows = 512;   // pairs per slab and slabs (X) with flags

__host__ __device__ __forceinline__ Support no_support()
{
    Support sp;
    sp.off = -1;
    sp.yp = sp.xyp = 0;
    return sp;
}

// Whether any thread of line pair w (threads tid % W == w) loaded a
// nonzero value: each wavefront's vote goes to s_vote[wave][w] before the
// barrier that follows the loads; pair_voted reads it after that barrier.
template <int W>
__device__ __forceinline__ void pair_vote(bool nz, int32_t (*s_vote)[16])
{
    const uint64_t m = __ballot(nz);
    const int lane = threadIdx.x & 63;
    constexpr uint64_t rep = ~0ull / ((1ull << W) - 1);  // lanes w mod W
    if (lane < W)
        s_vote[threadIdx.x >> 6][lane] = ((m >> lane) & rep) != 0;
}
__device__ __forceinline__ bool pair_voted(const int32_t (*s_vote)[16],
                                           int w)
{
    static_assert(kFftThreads == 256, "four wavefronts vote");
    return (s_vote[0][w] | s_vote[1][w] | s_vote[2][w] | s_vote[3][w]) != 0;
}

// Forward z transforms.  Lines (x, y) of T (X x Y x Z, z fastest) are taken
// in pairs (y even, y + 1) as one complex line T_y + i T_{y+1}, zero-padded
// to F; after the FFT the two half spectra (k = 0 .. F/2) are separated,
// X_y[k] = (C[k] + conj C[F-k]) / 2 and X_{y+1}[k] = (C[k] - conj C[F-k]) /
// 2i, and written to A at row (x * a_rows + y) with pitch kp.  Workgroup:
// W pairs, thread tid on pair tid % W.  H: the first stage's half form
// (half_in: rows z >= F / 2 are padding).
template <int W, bool H>
__global__ __launch_bounds__(kFftThreads) void z_fwd_kernel(
    BeamIO io, float2* __restrict__ A, int32_t X, int32_t Y, int32_t Z,
    int32_t a_rows, int32_t kp, FftLen L, const float2* __restrict__ g_tw,
    int32_t inplace, Support sp)
{
    extern __shared__ float2 sm[];
    __shared__ int32_t s_vote[kFftThreads / 64][16];
    const int F = L.n;
    float2* s_tw = sm;
    float2* s_a = sm + F;
    float2* s_b = s_a + W * F;
    load_twiddles(s_tw, g_tw, F);
    const int beam = beam_of_block(io);
    const float* __restrict__ T = io.t_tab ? io.t_tab[beam] : io.t;
    A += beam * io.a_beam;
    const int w = threadIdx.x % W;
    const int n0 = threadIdx.x / W;
    constexpr int kStep = kFftThreads / W;
    const int64_t pair = (int64_t)group_of_block(io) * W + w;
    const LinePair lp = line_pair(pair, X, Y);
    const bool two = lp.y + 1 < Y;
    const float* row = T + (lp.x * Y + lp.y) * (int64_t)Z;
    const int K = F / 2 + 1;
    float2* out = A + (lp.x * a_rows + lp.y) * (int64_t)kp;
    int32_t* flags = sp.off >= 0 ? (int32_t*)(A + sp.off) : nullptr;
    bool nz = false;
    if (L.nfac >= 2) {
        auto load = [&](int n) -> float2 {
            float2 v = make_float2(0.0f, 0.0f);
            if (lp.valid && n < Z) {
                v.x = row[n];
                if (two)
                    v.y = row[Z + n];
            }
            nz = nz || v.x != 0.0f || v.y != 0.0f;
            return v;
        };
        first_any<false, W, H>(L.fac[0], load, s_a, F);
        pair_vote<W>(nz, s_vote);
        __syncthreads();
        // a pair of zero lines has zero spectra: with flags, only the flag
        // is written, and a workgroup of zero pairs stops here
        const bool keep = !flags || pair_voted(s_vote, w);
        if (flags && lp.valid && n0 == 0)
            flags[pair] = keep ? 1 : 0;
        if (flags) {
            bool any = false;
            for (int q = 0; q < W; ++q)
                any = any || pair_voted(s_vote, q);
            if (!any)
                return;  // workgroup-uniform
        }
        int Ns = L.fac[0];
        const float2* r = s_a;
        if (inplace)
            stages_range_ip<false, W>(s_a, s_tw, F, L.fac, 1, L.nfac, 1, Ns);
        else
            r = stages_range<false, W>(s_a, s_b, s_tw, F, L.fac, 1, L.nfac, 1,
                                       Ns);
        if (!lp.valid || !keep)
            return;
        for (int k = n0; k < K; k += kStep) {
            const float2 c = r[k * W + w];
            const float2 cm = r[(k == 0 ? 0 : F - k) * W + w];
            out[k] = make_float2(0.5f * (c.x + cm.x), 0.5f * (c.y - cm.y));
            if (two)
                out[kp + k] =
                    make_float2(0.5f * (c.y + cm.y), -0.5f * (c.x - cm.x));
        }
        return;
    }
    // every load of this thread first, then the LDS stores
    float2 v[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
        const int n = n0 + i * kStep;
        v[i] = make_float2(0.0f, 0.0f);
        if (lp.valid && n < Z) {
            v[i].x = row[n];
            if (two)
                v[i].y = row[Z + n];
        }
        nz = nz || v[i].x != 0.0f || v[i].y != 0.0f;
    }
    pair_vote<W>(nz, s_vote);
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
        const int n = n0 + i * kStep;
        if (n < F)
            s_a[n * W + w] = v[i];
    }
    __syncthreads();
    const bool keep = !flags || pair_voted(s_vote, w);
    if (flags && lp.valid && n0 == 0)
        flags[pair] = keep ? 1 : 0;
    const float2* r = lds_fft<false, W>(s_a, s_b, s_tw, L);
    if (!lp.valid || !keep)
        return;
    for (int k = n0; k < K; k += kStep) {
        const float2 c = r[k * W + w];
        const float2 cm = r[(k == 0 ? 0 : F - k) * W + w];
        out[k] = make_float2(0.5f * (c.x + cm.x), 0.5f * (c.y - cm.y));
        if (two)
            out[kp + k] =
                make_float2(0.5f * (c.y + cm.y), -0.5f * (c.x - cm.x));
    }
}

// Inverse z transforms: the half spectra of lines y and y + 1 (Hermitian
// extension conj X[F-k] above F/2, imaginary parts of the k = 0 bin and,
// for even F, the k = F/2 bin dropped) as C = X_y + i X_{y+1}; the inverse
// FFT gives line y in the real and line y + 1 in the imaginary part; z < Z
// is stored or added.  H: the last stage's half form (half_out: z < Z <=
// F / 2 kept).
template <int W, bool H>
__global__ __launch_bounds__(kFftThreads) void z_inv_kernel(
    BeamIO io, const float2* __restrict__ A, int32_t X, int32_t Y,
    int32_t Z, int32_t a_rows, int32_t kp, FftLen L,
    const float2* __restrict__ g_tw, int32_t inplace)
{
    extern __shared__ float2 sm[];
    const int F = L.n;
    float2* s_tw = sm;
    float2* s_a = sm + F;
    float2* s_b = s_a + W * F;
    load_twiddles(s_tw, g_tw, F);
    const int beam = beam_of_block(io);
    float* __restrict__ D = io.d_tab ? io.d_tab[beam] : io.d;
    const int accumulate = io.acc_tab ? io.acc_tab[beam] : io.accumulate;
    A += beam * io.a_beam;
    const int w = threadIdx.x % W;
    const int n0 = threadIdx.x / W;
    constexpr int kStep = kFftThreads / W;
    const LinePair lp =
        line_pair((int64_t)group_of_block(io) * W + w, X, Y);
    const bool two = lp.y + 1 < Y;
    const float2* row = A + (lp.x * a_rows + lp.y) * (int64_t)kp;
    const int half = F / 2;
    float* out = D + (lp.x * Y + lp.y) * (int64_t)Z;
    if (L.nfac >= 2) {
        auto load = [&](int k) -> float2 {
            float2 a = make_float2(0.0f, 0.0f), b = a;
            if (lp.valid) {
                const int kk = k <= half ? k : F - k;
                a = row[kk];
                if (two)
                    b = row[kp + kk];
            }
            if (k > half) {
                a.y = -a.y;
                b.y = -b.y;
            }
            if (k == 0 || 2 * k == F) {
                a.y = 0.0f;
                b.y = 0.0f;
            }
            return make_float2(a.x - b.y, a.y + b.x);
        };
        first_any<true, W, false>(L.fac[0], load, s_a, F);
        __syncthreads();
        int Ns = L.fac[0];
        const float2* cur = s_a;
        if (inplace)
            stages_range_ip<true, W>(s_a, s_tw, F, L.fac, 1, L.nfac - 1, 1,
                                     Ns);
        else
            cur = stages_range<true, W>(s_a, s_b, s_tw, F, L.fac, 1,
                                        L.nfac - 1, 1, Ns);
        auto store = [&](int n, float2 c) {
            if (!lp.valid || n >= Z)
                return;
            if (accumulate) {
                out[n] += c.x;
                if (two)
                    out[Z + n] += c.y;
            } else {
                out[n] = c.x;
                if (two)
                    out[Z + n] = c.y;
            }
        };
        last_any<true, W, H>(L.fac[L.nfac - 1], cur, F, s_tw, store);
        return;
    }
    float2 va[kPerThread], vb[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
        const int k = n0 + i * kStep;
        va[i] = make_float2(0.0f, 0.0f);
        vb[i] = make_float2(0.0f, 0.0f);
        if (lp.valid && k < F) {
            const int kk = k <= half ? k : F - k;
            va[i] = row[kk];
            if (two)
                vb[i] = row[kp + kk];
        }
    }
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
        const int k = n0 + i * kStep;
        if (k >= F)
            continue;
        float2 a = va[i], b = vb[i];
        if (k > half) {
            a.y = -a.y;
            b.y = -b.y;
        }
        if (k == 0 || 2 * k == F) {
            a.y = 0.0f;
            b.y = 0.0f;
        }
        s_a[k * W + w] = make_float2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    const float2* r = lds_fft<true, W>(s_a, s_b, s_tw, L);
    if (!lp.valid)
        return;
    for (int n = n0; n < Z; n += kStep) {
        const float2 c = r[n * W + w];
        if (accumulate) {
            out[n] += c.x;
            if (two)
                out[Z + n] += c.y;
        } else {
            out[n] = c.x;
            if (two)
                out[Z + n] = c.y;
        }
    }
}

enum ColMode { kColFwd = 0, kColInv = 1, kColConv = 2 };

// Transforms along a strided axis: W adjacent columns (kz .. kz + W - 1) of
// the rows n = 0 .. F-1 at base + n * row_stride, base = A + outer *
// outer_stride + W * group, with (group, beam, outer) from the block id
// (XCD-major for the y passes).  Rows n >= n_in read as zero; rows n <
// n_out are written back in place.  kColConv multiplies by K^ (same
// layout, its own strides) between the forward and the inverse transform.
// `inplace`: the host allotted one LDS buffer (single_buffer()).  H: the
// half forms (col_half), bit 0 the first stage's (half_in), bit 1 the last
// stage's (half_out).  With support flags (sp), the forward y pass (outer =
// slab x, rows y) and the convolution pass (rows x) read only the rows the
// flags mark nonzero.
template <int W, int MODE, int H>
__global__ __launch_bounds__(kFftThreads) void col_kernel(
    BeamIO io, float2* __restrict__ A, int64_t row_stride,
    int64_t outer_stride, int32_t n_cols, int32_t n_in, int32_t n_out,
    FftLen L, const float2* __restrict__ g_tw,
    const float2* __restrict__ khat, int64_t k_row_stride,
    int64_t k_outer_stride, int32_t inplace, Support sp)
{
    constexpr int kPer = MODE == kColConv ? kPerThreadConv : kPerThread;
    // the forward transform reads rows n < F / 2 only (HI), the inverse one
    // keeps outputs k < F / 2 only (HO)
    constexpr bool HI = (H & 1) != 0 && MODE != kColInv;
    constexpr bool HO = (H & 2) != 0 && MODE != kColFwd;
    extern __shared__ float2 sm[];
    __shared__ int32_t s_rows[MODE == kColInv ? 1 : kMaxFlagRows];
    const int F = L.n;
    float2* s_tw = sm;
    float2* s_a = sm + F;
    float2* s_b = s_a + F * W;
    load_twiddles(s_tw, g_tw, F);
    // (the convolution pass reads whole 128-byte lines and measured no
    // faster in XCD-major order)
    const BlockXY blk = MODE == kColFwd ? paired_block(io.n_beams)
                                        : xcd_major_block(MODE == kColInv);
    const int beam = (int)(blk.x % (unsigned)io.n_beams);
    const int k0 = (int)(blk.x / (unsigned)io.n_beams) * W;
    if (k0 >= n_cols)
        return;  // a group added to make the count even
    const int w = threadIdx.x % W;
    const int n0 = threadIdx.x / W;
    constexpr int kStep = kFftThreads / W;
    const bool col_ok = w < n_cols - k0;
    float2* base = A + beam * io.a_beam + blk.y * outer_stride + k0 + w;
    // rows marked zero by the support flags (workgroup-uniform branches)
    const bool flagged = MODE != kColInv && sp.off >= 0;
    if (flagged) {
        int32_t* flags = (int32_t*)(A + beam * io.a_beam + sp.off);
        if (MODE == kColFwd) {
            // the line pairs of slab blk.y; a slab of zero pairs has zero
            // y transforms: only its flag is written
            int any = 0;
            for (int i = threadIdx.x; i < sp.yp; i += kFftThreads) {
                const int32_t f = flags[(int64_t)blk.y * sp.yp + i];
                s_rows[i] = f;
                any |= f;
            }
            any = __syncthreads_or(any);
            if (threadIdx.x == 0)
                flags[sp.xyp + blk.y] = any ? 1 : 0;
            if (!any)
                return;
        } else {
            for (int i = threadIdx.x; i < n_in; i += kFftThreads)
                s_rows[i] = flags[sp.xyp + i];
            __syncthreads();
        }
    }
    auto row_nz = [&](int n) -> bool {
        return !flagged || s_rows[MODE == kColFwd ? n >> 1 : n] != 0;
    };
    if (L.nfac >= 2) {
        const int nf = L.nfac;
        auto load = [&](int n) -> float2 {
            return (col_ok && n < n_in && row_nz(n))
                       ? base[n * row_stride]
                       : make_float2(0.0f, 0.0f);
        };
        auto store = [&](int n, float2 v) {
            if (col_ok && n < n_out)
                base[n * row_stride] = v;
        };
        first_any<MODE == kColInv, W, HI>(L.fac[0], load, s_a, F);
        __syncthreads();
        int Ns = L.fac[0];
        float2* cur;
        if (inplace) {
            // one LDS buffer (the host allotted F * W elements)
            if (MODE == kColConv) {
                stages_range_ip<false, W>(s_a, s_tw, F, L.fac, 1, nf - 1, 1,
                                          Ns);
                const float2* kb =
                    khat + blk.y * k_outer_stride + k0 + w;
                auto kload = [&](int k) -> float2 {
                    return col_ok ? kb[k * k_row_stride]
                                  : make_float2(0.0f, 0.0f);
                };
                conv_any_ip<W>(L.fac[nf - 1], s_a, F, s_tw, kload);
                __syncthreads();
                Ns = L.fac[nf - 1];
                stages_range_ip<true, W>(s_a, s_tw, F, L.fac, nf - 2, 0, -1,
                                         Ns);
                last_any<true, W, HO>(L.fac[0], s_a, F, s_tw, store);
            } else {
                stages_range_ip<MODE == kColInv, W>(s_a, s_tw, F, L.fac, 1,
                                                    nf - 1, 1, Ns);
                last_any<MODE == kColInv, W, HO>(L.fac[nf - 1], s_a, F,
                                                 s_tw, store);
            }
            return;
        }
        if (MODE == kColConv) {
            cur = stages_range<false, W>(s_a, s_b, s_tw, F, L.fac, 1, nf - 1,
                                         1, Ns);
            float2* oth = (cur == s_a) ? s_b : s_a;
            const float2* kb = khat + blk.y * k_outer_stride + k0 + w;
            auto kload = [&](int k) -> float2 {
                return col_ok ? kb[k * k_row_stride]
                              : make_float2(0.0f, 0.0f);
            };
            conv_any<W>(L.fac[nf - 1], cur, oth, F, s_tw, kload);
            __syncthreads();
            // the inverse takes the factors in reverse: its first stage
            // (radix fac[nf-1]) was fused above
            Ns = L.fac[nf - 1];
            cur = stages_range<true, W>(oth, cur, s_tw, F, L.fac, nf - 2, 0,
                                        -1, Ns);
            last_any<true, W, HO>(L.fac[0], cur, F, s_tw, store);
        } else {
            cur = stages_range<MODE == kColInv, W>(s_a, s_b, s_tw, F, L.fac,
                                                   1, nf - 1, 1, Ns);
            last_any<MODE == kColInv, W, HO>(L.fac[nf - 1], cur, F, s_tw,
                                             store);
        }
        return;
    }
    float2 v[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int n = n0 + i * kStep;
        v[i] = make_float2(0.0f, 0.0f);
        if (col_ok && n < n_in && row_nz(n))
            v[i] = base[n * row_stride];
    }
    // K^ for this thread's elements, loaded while the transform runs
    float2 kh[MODE == kColConv ? kPer : 1];
    if (MODE == kColConv) {
        const float2* kb = khat + blk.y * k_outer_stride + k0 + w;
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int n = n0 + i * kStep;
            kh[i] = make_float2(0.0f, 0.0f);
            if (col_ok && n < F)
                kh[i] = kb[n * k_row_stride];
        }
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
        const int n = n0 + i * kStep;
        if (n < F)
            s_a[n * W + w] = v[i];
    }
    __syncthreads();
    float2* r;
    if (MODE == kColInv)
        r = lds_fft<true, W>(s_a, s_b, s_tw, L);
    else
        r = lds_fft<false, W>(s_a, s_b, s_tw, L);
    if (MODE == kColConv) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int n = n0 + i * kStep;
            if (n < F)
                r[n * W + w] = cmul(r[n * W + w], kh[MODE == kColConv ? i : 0]);
        }
        __syncthreads();
        float2* o = (r == s_a) ? s_b : s_a;
        r = lds_fft<true, W>(r, o, s_tw, L);
    }
    if (!col_ok)
        return;
    for (int n = n0; n < n_out; n += kStep)
        base[n * row_stride] = r[n * W + w];
}

__global__ void scatter_frame_kernel(float* __restrict__ frame,
                                     const int64_t* __restrict__ idx,
                                     const float* __restrict__ val, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        frame[idx[i]] = val[i];
}

// ---------------------------------------------------------------- host --

// Lines per workgroup: by default the most (up to 16) with F * W <=
// limit; `want` (2, 4, 8 or 16) is taken when F * want <= kFftLdsElemsMax.
int lines_per_group(int F, int limit = kFftLdsElems, int want = 0)
{
    auto most = [F](int lim) {
        int w = 16;
        while (w > 2 && F * w > lim)
            w /= 2;
        return w;
    };
    if ((want == 2 || want == 4 || want == 8 || want == 16) &&
        want <= most(kFftLdsElemsMax))
        return want;
    return most(limit);
}

size_t z_lds_bytes(int F, int W)
{
    return (size_t)(F + 2 * W * F) * sizeof(float2);
}

size_t col_lds_bytes(int F, int W)
{
    return (size_t)(F + 2 * F * W) * sizeof(float2);
}

// One LDS buffer suffices when the fused path runs (two stages or more)
// and every stage gives each thread at most one butterfly group.
bool single_buffer(const FftLen& L, int W)
{
    bool ip = L.nfac >= 2;
    for (int i = 0; i < L.nfac; ++i)
        ip = ip && (L.n / L.fac[i]) * W <= kFftThreads;
    return ip;
}

template <typename K>
ch_status allow_lds(K kernel, size_t bytes)
{
    CH_HIP(hipFuncSetAttribute((const void*)kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
    return CH_OK;
}

// Radices 16, 8, 4, 2, then 3 and 5; when an odd radix would come last and
// two or more radices are even, the last even one moves to the end, so that
// the first and the last stage both have an even radix (half_in, half_out).
bool fft_factor(int n, FftLen& L)
{
    L.n = n;
    L.nfac = 0;
    int m = n;
    const int order[6] = {16, 8, 4, 2, 3, 5};
    for (int r : order)
        while (m % r == 0 && L.nfac < 12) {
            L.fac[L.nfac++] = r;
            m /= r;
        }
    int even = 0;
    while (even < L.nfac && L.fac[even] % 2 == 0)
        ++even;
    if (even >= 2 && even < L.nfac) {
        const int r = L.fac[even - 1];
        for (int i = even - 1; i + 1 < L.nfac; ++i)
            L.fac[i] = L.fac[i + 1];
        L.fac[L.nfac - 1] = r;
    }
    return m == 1;
}

FftLen plan_len(const ch_plan* p, int a)
{
    FftLen L;
    fft_factor(p->fft_n[a], L);
    return L;
}

template <int W>
ch_status launch_z_fwd(const BeamIO& io, float2* A, int X, int Y, int Z,
                       int a_rows, int kp, const FftLen& L, const float2* tw,
                       hipStream_t s, const Support& sp)
{
    const bool ip = single_buffer(L, W);
    const size_t lds = ip ? (size_t)(L.n + L.n * W) * sizeof(float2)
                          : z_lds_bytes(L.n, W);
    auto kernel = half_in(L, Z) ? z_fwd_kernel<W, true>
                                : z_fwd_kernel<W, false>;
    ch_status st = allow_lds(kernel, z_lds_bytes(L.n, W));
    if (st != CH_OK)
        return st;
    const int64_t groups = ((int64_t)X * ((Y + 1) / 2) + W - 1) / W;
    hipLaunchKernelGGL(kernel,
                       dim3((unsigned)(groups * io.n_beams)),
                       dim3(kFftThreads), lds, s, io, A, X, Y, Z, a_rows, kp,
                       L, tw, (int32_t)ip, sp);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

template <int W>
ch_status launch_z_inv(const BeamIO& io, const float2* A, int X, int Y,
                       int Z, int a_rows, int kp, const FftLen& L,
                       const float2* tw, hipStream_t s)
{
    const bool ip = single_buffer(L, W);
    const size_t lds = ip ? (size_t)(L.n + L.n * W) * sizeof(float2)
                          : z_lds_bytes(L.n, W);
    auto kernel = half_out(L, Z) ? z_inv_kernel<W, true>
                                 : z_inv_kernel<W, false>;
    ch_status st = allow_lds(kernel, z_lds_bytes(L.n, W));
    if (st != CH_OK)
        return st;
    const int64_t groups = ((int64_t)X * ((Y + 1) / 2) + W - 1) / W;
    hipLaunchKernelGGL(kernel,
                       dim3((unsigned)(groups * io.n_beams)),
                       dim3(kFftThreads), lds, s, io, A, X, Y, Z, a_rows, kp,
                       L, tw, (int32_t)ip);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

// The half forms of a column pass (col_kernel's H): bit 0 for its forward
// transform's first stage, bit 1 for its inverse transform's last stage.
template <int MODE>
int col_half(const FftLen& L, int n_in, int n_out)
{
    return (MODE != kColInv && half_in(L, n_in) ? 1 : 0) |
           (MODE != kColFwd && half_out(L, n_out) ? 2 : 0);
}

template <int W, int MODE>
auto col_variant(int h)
{
    if constexpr (MODE == kColFwd)
        return h ? col_kernel<W, MODE, 1> : col_kernel<W, MODE, 0>;
    else if constexpr (MODE == kColInv)
        return h ? col_kernel<W, MODE, 2> : col_kernel<W, MODE, 0>;
    else
        return h == 3   ? col_kernel<W, MODE, 3>
               : h == 2 ? col_kernel<W, MODE, 2>
               : h == 1 ? col_kernel<W, MODE, 1>
                        : col_kernel<W, MODE, 0>;
}

template <int W, int MODE>
ch_status launch_col(const BeamIO& io, float2* A, int64_t row_stride,
                     int64_t outer_stride, int n_outer, int n_cols, int n_in,
                     int n_out, const FftLen& L, const float2* tw,
                     const float2* khat, int64_t k_row_stride,
                     int64_t k_outer_stride, hipStream_t s,
                     LaunchTimer* timer, const Support& sp)
{
    // one LDS buffer when the fused path runs and every stage gives each
    // thread at most one butterfly group
    const bool ip = single_buffer(L, W);
    const size_t lds = ip ? (size_t)(L.n + L.n * W) * sizeof(float2)
                          : col_lds_bytes(L.n, W);
    auto kernel = col_variant<W, MODE>(col_half<MODE>(L, n_in, n_out));
    ch_status st = allow_lds(kernel, col_lds_bytes(L.n, W));
    if (st != CH_OK)
        return st;
    // the forward y pass takes an even number of column groups
    // (paired_block)
    int64_t groups = (n_cols + W - 1) / W;
    if (MODE == kColFwd)
        groups += groups & 1;
    const dim3 grid((unsigned)(groups * io.n_beams), (unsigned)n_outer);
    if (timer)
        timer->launch(kernel, grid, dim3(kFftThreads),
                      (std::uint32_t)lds, io, A, row_stride, outer_stride,
                      n_cols, n_in, n_out, L, tw, khat, k_row_stride,
                      k_outer_stride, (int32_t)ip, sp);
    else
        hipLaunchKernelGGL(kernel, grid, dim3(kFftThreads),
                           lds, s, io, A, row_stride, outer_stride, n_cols,
                           n_in, n_out, L, tw, khat, k_row_stride,
                           k_outer_stride, (int32_t)ip, sp);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

// W is chosen from the transform length at run time.
ch_status z_fwd(const BeamIO& io, float2* A, int X, int Y, int Z, int a_rows,
                int kp, const FftLen& L, const float2* tw, hipStream_t s,
                int want, const Support& sp)
{
    switch (lines_per_group(L.n, kFftLdsElems, want)) {
    case 16:
        return launch_z_fwd<16>(io, A, X, Y, Z, a_rows, kp, L, tw, s, sp);
    case 8:
        return launch_z_fwd<8>(io, A, X, Y, Z, a_rows, kp, L, tw, s, sp);
    case 4:
        return launch_z_fwd<4>(io, A, X, Y, Z, a_rows, kp, L, tw, s, sp);
    default:
        return launch_z_fwd<2>(io, A, X, Y, Z, a_rows, kp, L, tw, s, sp);
    }
}

ch_status z_inv(const BeamIO& io, const float2* A, int X, int Y, int Z,
                int a_rows, int kp, const FftLen& L, const float2* tw,
                hipStream_t s, int want)
{
    switch (lines_per_group(L.n, kFftLdsElems, want)) {
    case 16:
        return launch_z_inv<16>(io, A, X, Y, Z, a_rows, kp, L, tw, s);
    case 8:
        return launch_z_inv<8>(io, A, X, Y, Z, a_rows, kp, L, tw, s);
    case 4:
        return launch_z_inv<4>(io, A, X, Y, Z, a_rows, kp, L, tw, s);
    default:
        return launch_z_inv<2>(io, A, X, Y, Z, a_rows, kp, L, tw, s);
    }
}

template <int MODE>
ch_status col(const BeamIO& io, float2* A, int64_t row_stride,
              int64_t outer_stride, int n_outer, int n_cols, int n_in,
              int n_out, const FftLen& L, const float2* tw,
              const float2* khat, int64_t k_row_stride,
              int64_t k_outer_stride, hipStream_t s, int want,
              LaunchTimer* timer = nullptr,
              const Support& sp = no_support())
{
    switch (lines_per_group(
        L.n, MODE == kColConv ? kFftLdsElemsConv : kFftLdsElems, want)) {
    case 16:
        return launch_col<16, MODE>(io, A, row_stride, outer_stride, n_outer,
                                    n_cols, n_in, n_out, L, tw, khat,
                                    k_row_stride, k_outer_stride, s, timer,
                                    sp);
    case 8:
        return launch_col<8, MODE>(io, A, row_stride, outer_stride, n_outer,
                                   n_cols, n_in, n_out, L, tw, khat,
                                   k_row_stride, k_outer_stride, s, timer,
                                   sp);
    case 4:
        return launch_col<4, MODE>(io, A, row_stride, outer_stride, n_outer,
                                   n_cols, n_in, n_out, L, tw, khat,
                                   k_row_stride, k_outer_stride, s, timer,
                                   sp);
    default:
        return launch_col<2, MODE>(io, A, row_stride, outer_stride, n_outer,
                                   n_cols, n_in, n_out, L, tw, khat,
                                   k_row_stride, k_outer_stride, s, timer,
                                   sp);
    }
}

BeamIO single_beam(const float* t, float* d, int32_t accumulate)
{
    BeamIO io;
    io.t = t;
    io.d = d;
    io.t_tab = nullptr;
    io.d_tab = nullptr;
    io.acc_tab = nullptr;
    io.accumulate = accumulate;
    io.n_beams = 1;
    io.a_beam = 0;
    return io;
}

// Smallest n' >= n whose prime factors are 2, 3 and 5.
int32_t smooth_at_least(int32_t n)
{
    for (int32_t m = std::max(n, 1);; ++m) {
        int32_t k = m;
        for (int32_t q : {2, 3, 5})
            while (k % q == 0)
                k /= q;
        if (k == 1)
            return m;
    }
}

// The support flags of a beam's work grid (after its complex points), or
// none when a slab or its line pairs exceed the flag rows of the kernels.
Support plan_support(const ch_plan* p)
{
    const int X = p->shape[0], yp = (p->shape[1] + 1) / 2;
    if (X > kMaxFlagRows || yp > kMaxFlagRows)
        return no_support();
    Support sp;
    sp.off = (int64_t)X * p->fft_n[1] * p->fft_kp;
    sp.yp = yp;
    sp.xyp = X * yp;
    return sp;
}

// Complex points of one beam's work grid, its support flags included
// (a multiple of 32, so every beam's grid starts 256-byte aligned).
size_t beam_points(const ch_plan* p)
{
    const size_t grid = (size_t)p->shape[0] * p->fft_n[1] * p->fft_kp;
    const Support sp = plan_support(p);
    const size_t flags = sp.off < 0 ? 0 : ((size_t)sp.xyp + p->shape[0] + 1) / 2;
    return (grid + flags + 31) / 32 * 32;
}

// The five passes over the beams of `io` (all on plan p), work grids in A.
// time_conv: the caller's launch events (and ch_set_timing) bracket the
// convolution pass, the dominant kernel, instead of the whole call.
ch_status spectral_passes(const ch_plan* p, BeamIO io, float2* A,
                          hipStream_t s, bool time_conv)
{
    const int X = p->shape[0], Y = p->shape[1], Z = p->shape[2];
    const int Fy = p->fft_n[1];
    const int kz = p->fft_n[2] / 2 + 1, kp = p->fft_kp;
    const FftLen Lx = plan_len(p, 0), Ly = plan_len(p, 1), Lz = plan_len(p, 2);
    io.a_beam = (int64_t)beam_points(p);
    const int64_t slab = (int64_t)Fy * kp;
    const int32_t* W = p->fft_lines;
    const Support sp = plan_support(p);
    ch_status st =
        z_fwd(io, A, X, Y, Z, Fy, kp, Lz, p->d_tw[2], s, W[0], sp);
    if (st == CH_OK)
        st = col<kColFwd>(io, A, kp, slab, X, kz, Y, Fy, Ly, p->d_tw[1],
                          nullptr, 0, 0, s, W[1], nullptr, sp);
    if (st == CH_OK) {
        if (time_conv) {
            LaunchTimer timer(s);
            st = col<kColConv>(io, A, slab, kp, Fy, kz, X, X, Lx, p->d_tw[0],
                               p->d_khat, slab, kp, s, W[2], &timer, sp);
            timer.stop();
        } else {
            st = col<kColConv>(io, A, slab, kp, Fy, kz, X, X, Lx, p->d_tw[0],
                               p->d_khat, slab, kp, s, W[2], nullptr, sp);
        }
    }
    if (st == CH_OK)
        st = col<kColInv>(io, A, kp, slab, X, kz, Fy, Y, Ly, p->d_tw[1],
                          nullptr, 0, 0, s, W[3]);
    if (st == CH_OK)
        st = z_inv(io, A, X, Y, Z, Fy, kp, Lz, p->d_tw[2], s, W[4]);
    return st;
}

// The launches of a batch of FFT-form jobs: jobs on one plan share each
// launch, in order of first appearance, up to kSpectralWorkBytes of work
// grids per launch.  Each launch's device table (TERMA pointers, dose
// pointers, accumulate flags) is laid out in `tables` at L.table.
struct FftLaunch {
    const ch_plan* plan;
    int32_t nb;
    size_t work;          // work-grid bytes
    size_t table;         // offset of the table in the set's tables
    size_t table_bytes;
};

static ch_status fft_launches(int32_t n_jobs, const ch_conv_job* h_jobs,
                              const char* fn, std::vector<FftLaunch>& out,
                              std::vector<char>& tables)
{
    CH_REQUIRE(n_jobs >= 0 && (n_jobs == 0 || h_jobs), "%s: bad job list",
               fn);
    for (int32_t i = 0; i < n_jobs; ++i) {
        const ch_conv_job& j = h_jobs[i];
        CH_REQUIRE(j.plan && j.d_terma && j.d_dose,
                   "%s: job %d has a null pointer", fn, i);
        CH_REQUIRE(j.plan->d_khat,
                   "%s: job %d's plan has no spectrum (call "
                   "ch_plan_spectrum first)", fn, i);
    }
    std::vector<const ch_plan*> plans;
    std::vector<std::vector<int32_t>> members;
    for (int32_t i = 0; i < n_jobs; ++i) {
        size_t g = 0;
        while (g < plans.size() && plans[g] != h_jobs[i].plan)
            ++g;
        if (g == plans.size()) {
            plans.push_back(h_jobs[i].plan);
            members.emplace_back();
        }
        members[g].push_back(i);
    }
    out.clear();
    tables.clear();
    for (size_t g = 0; g < plans.size(); ++g) {
        const ch_plan* p = plans[g];
        const size_t per = beam_points(p) * sizeof(float2);
        const size_t cap = std::max<size_t>(1, kSpectralWorkBytes / per);
        const std::vector<int32_t>& m = members[g];
        for (size_t b0 = 0; b0 < m.size(); b0 += cap) {
            FftLaunch L;
            L.plan = p;
            L.nb = (int32_t)std::min(cap, m.size() - b0);
            L.work = per * L.nb;
            L.table = tables.size();
            const size_t ptrs = (size_t)L.nb * sizeof(void*);
            L.table_bytes = 2 * ptrs + (size_t)L.nb * sizeof(int32_t);
            tables.resize(L.table + ((L.table_bytes + 255) & ~(size_t)255));
            char* h = tables.data() + L.table;
            const float** ht = (const float**)h;
            float** hd = (float**)(h + ptrs);
            int32_t* ha = (int32_t*)(h + 2 * ptrs);
            for (int32_t b = 0; b < L.nb; ++b) {
                const ch_conv_job& j = h_jobs[m[b0 + b]];
                ht[b] = j.d_terma;
                hd[b] = j.d_dose;
                ha[b] = j.accumulate ? 1 : 0;
            }
            out.push_back(L);
        }
    }
    return CH_OK;
}

// The beams of a launch read through its device table at d.
static BeamIO table_io(void* d, int32_t nb)
{
    const size_t ptrs = (size_t)nb * sizeof(void*);
    BeamIO io = single_beam(nullptr, nullptr, 0);
    io.t_tab = (const float* const*)d;
    io.d_tab = (float* const*)((char*)d + ptrs);
    io.acc_tab = (const int32_t*)((char*)d + 2 * ptrs);
    io.n_beams = nb;
    return io;
}

}  // namespace
}  // namespace ch

using namespace ch;

// A device-resident set of FFT-form jobs: the launches of
// ch_convolve_fft_batch with their tables in device memory, written once,
// so ch_convolve_fft_set queues kernels only (it can be captured).
struct ch_convset : ch::GraphHeld {
    std::vector<ch::FftLaunch> launches;
    void* d_tables = nullptr;
};

static void convset_release(ch_convset* set)
{
    (void)hipDeviceSynchronize();  // a queued launch may still read it
    (void)hipFree(set->d_tables);
    delete set;
}

void spectrum_free(ch_plan* p)
{
    (void)hipFree(p->d_khat);
    for (int a = 0; a < 3; ++a)
        (void)hipFree(p->d_tw[a]);
    p->d_khat = nullptr;
    p->d_tw[0] = p->d_tw[1] = p->d_tw[2] = nullptr;
}

extern "C" {

ch_status ch_plan_spectrum(ch_plan* plan, void* stream)
{
    CH_REQUIRE(plan, "ch_plan_spectrum: null plan");
    if (plan->d_khat)
        return CH_OK;
    CH_REQUIRE(plan->semantics != CH_SEM_RADIOLOGICAL,
               "ch_plan_spectrum: radiological plans are not a convolution "
               "(the kernel distance depends on the densities)");
    hipStream_t s = (hipStream_t)stream;
    CH_REQUIRE(!capturing(s), "ch_plan_spectrum allocates: build the "
                              "spectrum before capturing");
    int32_t F[3];
    for (int a = 0; a < 3; ++a) {
        F[a] = smooth_at_least(2 * plan->shape[a] - 1);
        CH_REQUIRE(F[a] <= kFftMax,
                   "ch_plan_spectrum: axis %d of %d voxels needs a %d-point "
                   "transform (at most %d)", a, plan->shape[a], F[a], kFftMax);
    }
    const int32_t kz = F[2] / 2 + 1;
    const int32_t kp = (kz + 15) / 16 * 16;

    // K on the frame: every cone's taps at (o mod F), merged in FP64 and
    // scaled by 1 / (Fx Fy Fz), the inverse transform's normalisation
    const int nc = plan->n_cones;
    const size_t nt = (size_t)nc * plan->tap_stride;
    std::vector<int32_t> off(nt * 3);
    std::vector<double> k64(nt);
    CH_HIP(hipMemcpy(off.data(), plan->d_off, nt * 3 * sizeof(int32_t),
                     hipMemcpyDeviceToHost));
    CH_HIP(hipMemcpy(k64.data(), plan->d_k64, nt * sizeof(double),
                     hipMemcpyDeviceToHost));
    std::unordered_map<int64_t, double> acc;
    for (int c = 0; c < nc; ++c) {
        const ConeMeta& m = plan->h_meta[c];
        for (int32_t t = 0; t < m.n_taps; ++t) {
            const size_t i = (size_t)m.tap_begin + t;
            int64_t q = 0;
            for (int a = 0; a < 3; ++a) {
                const int32_t o = off[3 * i + a];
                q = q * F[a] + ((o % F[a]) + F[a]) % F[a];
            }
            acc[q] += k64[i];
        }
    }
    std::vector<std::pair<int64_t, double>> items(acc.begin(), acc.end());
    std::sort(items.begin(), items.end());
    const double scale = 1.0 / ((double)F[0] * F[1] * F[2]);
    std::vector<int64_t> h_idx(items.size());
    std::vector<float> h_val(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        h_idx[i] = items[i].first;
        h_val[i] = (float)(items[i].second * scale);
    }

    ch_status st = CH_OK;
    float* frame = nullptr;
    int64_t* d_idx = nullptr;
    float* d_val = nullptr;
    float2* khat = nullptr;
    float2* tw[3] = {nullptr, nullptr, nullptr};
    const size_t n_frame = (size_t)F[0] * F[1] * F[2];
    const size_t n_khat = (size_t)F[0] * F[1] * kp;
    hipError_t e = hipSuccess;
#define CH_TRY(x)                                                           \
    do {                                                                    \
        if (e == hipSuccess)                                                \
            e = (x);                                                        \
    } while (0)
    CH_TRY(hipMalloc((void**)&frame, n_frame * sizeof(float)));
    CH_TRY(hipMalloc((void**)&khat, n_khat * sizeof(float2)));
    CH_TRY(hipMalloc((void**)&d_idx, std::max<size_t>(1, h_idx.size()) *
                                         sizeof(int64_t)));
    CH_TRY(hipMalloc((void**)&d_val,
                     std::max<size_t>(1, h_val.size()) * sizeof(float)));
    for (int a = 0; a < 3; ++a) {
        std::vector<float2> h(F[a]);
        for (int32_t q = 0; q < F[a]; ++q) {
            const double ang = 2.0 * M_PI * (double)q / (double)F[a];
            h[q] = make_float2((float)std::cos(ang), (float)-std::sin(ang));
        }
        CH_TRY(hipMalloc((void**)&tw[a], F[a] * sizeof(float2)));
        CH_TRY(hipMemcpy(tw[a], h.data(), F[a] * sizeof(float2),
                         hipMemcpyHostToDevice));
    }
    CH_TRY(hipMemcpy(d_idx, h_idx.data(), h_idx.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    CH_TRY(hipMemcpy(d_val, h_val.data(), h_val.size() * sizeof(float),
                     hipMemcpyHostToDevice));
#undef CH_TRY
    if (e != hipSuccess) {
        set_error("ch_plan_spectrum: %s", hipGetErrorString(e));
        st = CH_ERR_HIP;
    }
    if (st == CH_OK)
        st = ch_zero(frame, n_frame * sizeof(float), stream);
    if (st == CH_OK && !h_idx.empty()) {
        const int64_t n = (int64_t)h_idx.size();
        hipLaunchKernelGGL(scatter_frame_kernel,
                           dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           frame, (const int64_t*)d_idx, (const float*)d_val,
                           n);
        if (hipGetLastError() != hipSuccess) {
            set_error("ch_plan_spectrum: scatter launch failed");
            st = CH_ERR_HIP;
        }
    }
    if (st == CH_OK) {
        plan->fft_n[0] = F[0];
        plan->fft_n[1] = F[1];
        plan->fft_n[2] = F[2];
        plan->fft_kp = kp;
        plan->d_tw[0] = tw[0];
        plan->d_tw[1] = tw[1];
        plan->d_tw[2] = tw[2];
        const FftLen Lx = plan_len(plan, 0), Ly = plan_len(plan, 1),
                     Lz = plan_len(plan, 2);
        const int64_t slab = (int64_t)F[1] * kp;
        // the whole frame is the grid here: no zero rows, all rows kept
        const BeamIO io = single_beam(frame, nullptr, 0);
        st = z_fwd(io, khat, F[0], F[1], F[2], F[1], kp, Lz, tw[2], s, 0,
                   no_support());
        if (st == CH_OK)
            st = col<kColFwd>(io, khat, kp, slab, F[0], kz, F[1], F[1], Ly,
                              tw[1], nullptr, 0, 0, s, 0);
        if (st == CH_OK)
            st = col<kColFwd>(io, khat, slab, kp, F[1], kz, F[0], F[0], Lx,
                              tw[0], nullptr, 0, 0, s, 0);
        if (st == CH_OK && hipStreamSynchronize(s) != hipSuccess) {
            set_error("ch_plan_spectrum: %s",
                      hipGetErrorString(hipGetLastError()));
            st = CH_ERR_HIP;
        }
    }
    (void)hipFree(frame);
    (void)hipFree(d_idx);
    (void)hipFree(d_val);
    if (st == CH_OK) {
        plan->d_khat = khat;
    } else {
        (void)hipFree(khat);
        for (int a = 0; a < 3; ++a)
            (void)hipFree(tw[a]);
        plan->d_tw[0] = plan->d_tw[1] = plan->d_tw[2] = nullptr;
        plan->fft_n[0] = plan->fft_n[1] = plan->fft_n[2] = 0;
    }
    return st;
}

ch_status ch_plan_spectrum_release(ch_plan* plan)
{
    CH_REQUIRE(plan, "ch_plan_spectrum_release: null plan");
    if (!plan->d_khat)
        return CH_OK;
    CH_REQUIRE(plan_pins(plan) == 0,
               "ch_plan_spectrum_release: %d live graph(s) read this plan's "
               "spectrum; destroy them first", plan_pins(plan));
    // kernels queued on any stream may still read K^
    CH_HIP(hipDeviceSynchronize());
    spectrum_free(plan);
    plan->fft_n[0] = plan->fft_n[1] = plan->fft_n[2] = 0;
    plan->fft_kp = 0;
    return CH_OK;
}

ch_status ch_plan_spectrum_info(const ch_plan* plan, int32_t* frame3,
                                int32_t* pitch, int64_t* work_bytes)
{
    CH_REQUIRE(plan, "ch_plan_spectrum_info: null plan");
    CH_REQUIRE(plan->d_khat, "ch_plan_spectrum_info: no spectrum (call "
                             "ch_plan_spectrum first)");
    if (frame3)
        for (int a = 0; a < 3; ++a)
            frame3[a] = plan->fft_n[a];
    if (pitch)
        *pitch = plan->fft_kp;
    if (work_bytes)
        *work_bytes = (int64_t)(beam_points(plan) * sizeof(float2));
    return CH_OK;
}

ch_status ch_convolve_fft(const ch_plan* plan, const float* d_terma,
                          float* d_dose, int32_t accumulate, void* stream)
{
    CH_REQUIRE(plan && d_terma && d_dose, "ch_convolve_fft: null pointer");
    CH_REQUIRE(plan->d_khat, "ch_convolve_fft: the plan has no spectrum "
                             "(call ch_plan_spectrum first)");
    hipStream_t s = (hipStream_t)stream;
    void* work = nullptr;
    ch_status st = stream_scratch(s, kScratchSpectrum,
                                  beam_points(plan) * sizeof(float2), &work);
    if (st != CH_OK)
        return st;
    capture_pin(s, plan);
    return spectral_passes(plan, single_beam(d_terma, d_dose, accumulate),
                           (float2*)work, s, true);
}

ch_status ch_convolve_fft_batch(int32_t n_jobs, const ch_conv_job* h_jobs,
                                void* stream)
{
    hipStream_t s = (hipStream_t)stream;
    CH_REQUIRE(!capturing(s), "ch_convolve_fft_batch stages a host table and "
                              "cannot be captured; capture ch_convolve_fft "
                              "or a ch_convset");
    std::vector<FftLaunch> launches;
    std::vector<char> tables;
    ch_status st = fft_launches(n_jobs, h_jobs, "ch_convolve_fft_batch",
                                launches, tables);
    if (st != CH_OK)
        return st;
    LaunchTimer timer(s, true);
    for (const FftLaunch& L : launches) {
        void* work = nullptr;
        st = stream_scratch(s, kScratchSpectrum, L.work, &work);
        if (st != CH_OK)
            return st;
        void* d = nullptr;
        st = stage_table(s, tables.data() + L.table, L.table_bytes, &d);
        if (st != CH_OK)
            return st;
        st = spectral_passes(L.plan, table_io(d, L.nb), (float2*)work, s,
                             false);
        if (st != CH_OK)
            return st;
    }
    timer.stop();
    return CH_OK;
}

ch_status ch_convset_create(int32_t n_jobs, const ch_conv_job* h_jobs,
                            ch_convset** out)
{
    CH_REQUIRE(out, "ch_convset_create: null output");
    *out = nullptr;
    CH_REQUIRE(n_jobs > 0, "ch_convset_create: no jobs");
    ch_convset* set = new ch_convset();
    set->release = [](GraphHeld* h) {
        convset_release(static_cast<ch_convset*>(h));
    };
    std::vector<char> tables;
    ch_status st = fft_launches(n_jobs, h_jobs, "ch_convset_create",
                                set->launches, tables);
    if (st != CH_OK) {
        convset_release(set);
        return st;
    }
    hipError_t e = hipMalloc(&set->d_tables, tables.size());
    if (e == hipSuccess)
        e = hipMemcpy(set->d_tables, tables.data(), tables.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        convset_release(set);
        return hip_fail(e, "ch_convset_create", __FILE__, __LINE__);
    }
    *out = set;
    return CH_OK;
}

ch_status ch_convolve_fft_set(const ch_convset* cset, void* stream)
{
    CH_REQUIRE(cset, "ch_convolve_fft_set: null set");
    ch_convset* set = const_cast<ch_convset*>(cset);
    hipStream_t s = (hipStream_t)stream;
    capture_hold(s, set);
    LaunchTimer timer(s, true);
    for (const FftLaunch& L : set->launches) {
        void* work = nullptr;
        ch_status st = stream_scratch(s, kScratchSpectrum, L.work, &work);
        if (st != CH_OK)
            return st;
        capture_pin(s, L.plan);
        st = spectral_passes(L.plan,
                             table_io((char*)set->d_tables + L.table, L.nb),
                             (float2*)work, s, false);
        if (st != CH_OK)
            return st;
    }
    timer.stop();
    return CH_OK;
}

ch_status ch_convset_destroy(ch_convset* set)
{
    if (!set)
        return CH_OK;
    if (defer_free(set))  // read by a live graph: freed with the last one
        return CH_OK;
    convset_release(set);
    return CH_OK;
}

ch_status ch_plan_set_fft_lines(ch_plan* plan, const int32_t lines[5])
{
    CH_REQUIRE(plan && lines, "ch_plan_set_fft_lines: null pointer");
    for (int k = 0; k < 5; ++k)
        CH_REQUIRE(lines[k] == 0 || lines[k] == 2 || lines[k] == 4 ||
                       lines[k] == 8 || lines[k] == 16,
                   "ch_plan_set_fft_lines: pass %d: %d lines (0, 2, 4, 8 "
                   "or 16)", k, lines[k]);
    for (int k = 0; k < 5; ++k)
        plan->fft_lines[k] = lines[k];
    return CH_OK;
}

ch_status ch_plan_get_fft_lines(const ch_plan* plan, int32_t lines[5])
{
    CH_REQUIRE(plan && lines, "ch_plan_get_fft_lines: null pointer");
    int32_t F[3];
    for (int a = 0; a < 3; ++a)
        F[a] = plan->fft_n[a] ? plan->fft_n[a]
                              : smooth_at_least(2 * plan->shape[a] - 1);
    const int* W = plan->fft_lines;
    lines[0] = lines_per_group(F[2], kFftLdsElems, W[0]);
    lines[1] = lines_per_group(F[1], kFftLdsElems, W[1]);
    lines[2] = lines_per_group(F[0], kFftLdsElemsConv, W[2]);
    lines[3] = lines_per_group(F[1], kFftLdsElems, W[3]);
    lines[4] = lines_per_group(F[2], kFftLdsElems, W[4]);
    return CH_OK;
}

}  // extern "C"
