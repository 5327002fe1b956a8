// ccc.hip -- collapsed-cone convolution as a deterministic stencil gather.
//
// dose(v) = sum_c sum_m k_{c,m} * T(v - o_{c,m}), over sources v - o inside
// the grid and inside the TERMA support box (T is zero outside it).  One
// wavefront owns a tile of 64 output voxels (TX x TY x TZ, z fastest); for
// each cone it clips every lane's tap range with the per-axis first-step
// tables (the path is monotone per axis, so the taps whose source lies in
// the box form one contiguous range), then walks the union of the ranges
// with wave-uniform tap loads and per-lane predicated, coalesced T loads.
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

// Field-wise copies out of the constant address space (a struct copy
// constructor cannot bind to an address_space(4) object).
__device__ __forceinline__ ConeMeta load_meta(const c_meta* p)
{
    ConeMeta m;
    m.tap_begin = p->tap_begin;
    m.n_taps = p->n_taps;
    for (int a = 0; a < 3; ++a) {
        m.f_begin[a] = p->f_begin[a];
        m.f_len[a] = p->f_len[a];
        m.step[a] = p->step[a];
    }
    m.pad = 0;
    return m;
}

__device__ __forceinline__ ConvJob load_job(const c_job* p)
{
    ConvJob j;
    j.terma = p->terma;
    j.dose = p->dose;
    j.box = p->box;
    j.lin = p->lin;
    j.k = p->k;
    j.ftab = p->ftab;
    j.meta = p->meta;
    j.X = p->X;
    j.Y = p->Y;
    j.Z = p->Z;
    j.n_cones = p->n_cones;
    j.accumulate = p->accumulate;
    j.ntx = p->ntx;
    j.nty = p->nty;
    j.ntz = p->ntz;
    j.tile_begin = p->tile_begin;
    j.pad = 0;
    return j;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Tap range [mlo, mhi) of one axis for coordinate v, box [lo, hi].
__device__ __forceinline__ void axis_range(int v, int lo, int hi, int step,
                                           const c_i32* f, int f_len,
                                           int& mlo, int& mhi)
{
    int L, U;
    if (step > 0) {
        L = v - hi;
        U = v - lo;
    } else {
        L = lo - v;
        U = hi - v;
    }
    const int last = f_len - 1;  // index of the sentinel (= n_taps)
    const int A = L > 0 ? L : 0;
    if (U < 0 || A >= last) {
        mhi = 0;
        return;
    }
    const int Ub = U + 1 < last ? U + 1 : last;
    mlo = max(mlo, f[A]);
    mhi = min(mhi, f[Ub]);
}

// Taps per unrolled step; every cone's tap list is padded with kUnroll
// zero taps (offset 0, increment 0) so a chunk may read past its end.
constexpr int kUnroll = 8;

// Accumulates taps [t0, t1) of one cone into (acc, ac).  PRED selects the
// per-lane predicate: 0 none (every lane active; t1 - t0 is a multiple of
// kUnroll), 1 t >= mlo && t < lim, 2 t < mhi.  Inactive lanes read their own voxel and add 0, so the loads
// of a chunk are unconditional and all in flight together.
template <int ACC, int PRED>
__device__ __forceinline__ void tap_run(const char* __restrict__ Tb,
                                        uint32_t base4, const c_i32* ml,
                                        const c_f32* mk, int t0, int t1,
                                        int mlo, int mhi, int lim,
                                        float& acc, float& ac)
{
    if (t0 >= t1)
        return;
    int o[kUnroll];
    float kv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
        o[u] = ml[t0 + u];
        kv[u] = mk[t0 + u];
    }
    for (int t = t0; t < t1; t += kUnroll) {
        float v[kUnroll], kc[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            bool act = true;
            if (PRED == 1)
                act = (t + u >= mlo) && (t + u < lim);
            else if (PRED == 2)
                act = t + u < mhi;
            const uint32_t off = act ? base4 - (uint32_t)o[u] : base4;
            v[u] = *reinterpret_cast<const float*>(Tb + off);
            kc[u] = act ? kv[u] : 0.0f;
        }
        // Prefetch the next chunk's taps while the T loads are in flight;
        // the padding keeps t + kUnroll .. t + 2 kUnroll - 1 readable.
        if (t + kUnroll < t1) {
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                o[u] = ml[t + kUnroll + u];
                kv[u] = mk[t + kUnroll + u];
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (ACC == 0)
                kahan_add(acc, ac, kc[u] * v[u]);
            else
                acc = fmaf(kc[u], v[u], acc);
        }
    }
}

template <int TX, int TY, int TZ, int ACC, typename OutT>
__device__ __forceinline__ void gather_tile(const ConvJob& j, int tile,
                                            int lane)
{
    static_assert(TX * TY * TZ == 64, "one wavefront per tile");
    const int tz = tile % j.ntz;
    const int r = tile / j.ntz;
    const int ty = r % j.nty;
    const int tx = r / j.nty;
    const int lz = lane % TZ;
    const int ly = (lane / TZ) % TY;
    const int lx = lane / (TZ * TY);
    const int x = tx * TX + lx, y = ty * TY + ly, z = tz * TZ + lz;
    const bool valid = x < j.X && y < j.Y && z < j.Z;

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
    const bool any = lo0 <= hi0 && lo1 <= hi1 && lo2 <= hi2;
    const int base = valid ? (x * j.Y + y) * j.Z + z : 0;
    const uint32_t base4 = (uint32_t)base * 4u;
    const char* __restrict__ Tb = reinterpret_cast<const char*>(j.terma);
    const c_meta* meta = as_const(j.meta);
    const c_i32* ftab = as_const(j.ftab);
    const c_i32* lin = as_const(j.lin);
    const c_f32* kk = as_const(j.k);

    float sum = 0.0f, comp = 0.0f;
    if (any) {
        for (int c = 0; c < j.n_cones; ++c) {
            const ConeMeta m = load_meta(meta + c);
            int mlo = 0, mhi = m.n_taps;
            axis_range(x, lo0, hi0, m.step[0], ftab + m.f_begin[0],
                       m.f_len[0], mlo, mhi);
            axis_range(y, lo1, hi1, m.step[1], ftab + m.f_begin[1],
                       m.f_len[1], mlo, mhi);
            axis_range(z, lo2, hi2, m.step[2], ftab + m.f_begin[2],
                       m.f_len[2], mlo, mhi);
            const bool part = valid && mhi > mlo;
            // union [ulo, uhi) and common part [blo, bhi) over the lanes
            // that have taps
            const int ulo = __builtin_amdgcn_readfirstlane(
                wave_min(part ? mlo : INT_MAX));
            const int uhi = __builtin_amdgcn_readfirstlane(
                wave_max(part ? mhi : INT_MIN));
            const int blo = __builtin_amdgcn_readfirstlane(
                wave_max(part ? mlo : INT_MIN));
            const int bhi = __builtin_amdgcn_readfirstlane(
                wave_min(part ? mhi : INT_MAX));
            if (ulo >= uhi)
                continue;  // no lane of this tile has a source in range
            const c_i32* ml = lin + m.tap_begin;
            const c_f32* mk = kk + m.tap_begin;
            float acc = 0.0f, ac = 0.0f;
            if (part) {
                if (blo < bhi) {
                    // head: lanes start at their own mlo
                    tap_run<ACC, 1>(Tb, base4, ml, mk, ulo, blo, mlo, mhi,
                                    blo, acc, ac);
                    // body: every participating lane is active
                    const int nb = (bhi - blo) / kUnroll * kUnroll;
                    tap_run<ACC, 0>(Tb, base4, ml, mk, blo, blo + nb, mlo,
                                    mhi, blo + nb, acc, ac);
                    // tail: rest of the body and lanes ending at their mhi
                    tap_run<ACC, 2>(Tb, base4, ml, mk, blo + nb, uhi, mlo,
                                    mhi, uhi, acc, ac);
                } else {
                    tap_run<ACC, 1>(Tb, base4, ml, mk, ulo, uhi, mlo, mhi,
                                    mhi, acc, ac);
                }
            }
            kahan_add(sum, comp, acc);
        }
    }
    if (valid) {
        OutT* d = reinterpret_cast<OutT*>(j.dose);
        const OutT prev = j.accumulate ? d[base] : (OutT)0;
        d[base] = prev + (OutT)sum;
    }
}

template <int TX, int TY, int TZ, int ACC, typename OutT>
__global__ __launch_bounds__(256) void gather_single(ConvJob j, int n_tiles)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= n_tiles)
        return;
    gather_tile<TX, TY, TZ, ACC, OutT>(j, tile, threadIdx.x & 63);
}

template <int TX, int TY, int TZ, int ACC, typename OutT>
__global__ __launch_bounds__(256) void gather_batch(const ConvJob* jobs,
                                                    int n_jobs, int n_tiles)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * 4 + wave;
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
    gather_tile<TX, TY, TZ, ACC, OutT>(jb, tile - jb.tile_begin,
                                       threadIdx.x & 63);
}

// ----------------------------------------------------------- dispatch ----

template <typename OutT>
struct Kernels {
    void (*single)(ConvJob, int);
    void (*batch)(const ConvJob*, int, int);
};

#define CH_TILE_LIST(X)                                                     \
    X(1, 1, 64) X(1, 2, 32) X(1, 4, 16) X(2, 2, 16) X(1, 8, 8) X(2, 4, 8)   \
    X(4, 4, 4) X(2, 8, 4) X(4, 8, 2) X(8, 8, 1)

template <int ACC, typename OutT>
static bool pick(int tx, int ty, int tz, Kernels<OutT>& k)
{
#define CH_CASE(a, b, c)                                                    \
    if (tx == a && ty == b && tz == c) {                                    \
        k.single = gather_single<a, b, c, ACC, OutT>;                       \
        k.batch = gather_batch<a, b, c, ACC, OutT>;                         \
        return true;                                                        \
    }
    CH_TILE_LIST(CH_CASE)
#undef CH_CASE
    return false;
}

template <typename OutT>
static ch_status kernels_for(const ch_plan* p, Kernels<OutT>& k)
{
    bool ok = p->accum_mode == 0
                  ? pick<0, OutT>(p->tile[0], p->tile[1], p->tile[2], k)
                  : pick<1, OutT>(p->tile[0], p->tile[1], p->tile[2], k);
    if (!ok) {
        set_error("no gather kernel for tile %dx%dx%d", p->tile[0],
                  p->tile[1], p->tile[2]);
        return CH_ERR_ARG;
    }
    return CH_OK;
}

static ConvJob make_job(const ch_plan* p, const float* terma, void* dose,
                        const int32_t* box, int32_t accumulate)
{
    ConvJob j;
    j.terma = terma;
    j.dose = dose;
    j.box = box;
    j.lin = p->d_lin;
    j.k = p->d_k;
    j.ftab = p->d_ftab;
    j.meta = p->d_meta;
    j.X = p->shape[0];
    j.Y = p->shape[1];
    j.Z = p->shape[2];
    j.n_cones = p->n_cones;
    j.accumulate = accumulate;
    j.ntx = (j.X + p->tile[0] - 1) / p->tile[0];
    j.nty = (j.Y + p->tile[1] - 1) / p->tile[1];
    j.ntz = (j.Z + p->tile[2] - 1) / p->tile[2];
    j.tile_begin = 0;
    j.pad = 0;
    return j;
}

// Optional HIP-event timing of the gather launches (bench and profiling).
struct LaunchTimer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s;
    bool on;
    explicit LaunchTimer(hipStream_t st) : s(st), on(timing_enabled())
    {
        if (on) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, s);
        }
    }
    void stop()
    {
        if (!on)
            return;
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = -1.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        record_kernel_ms(ms);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

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
    ConvJob j = make_job(p, d_terma, d_dose, d_box, accumulate);
    const int64_t n_tiles = (int64_t)j.ntx * j.nty * j.ntz;
    CH_REQUIRE(n_tiles < INT_MAX / 2, "too many tiles");
    const unsigned blocks = (unsigned)((n_tiles + 3) / 4);
    LaunchTimer timer((hipStream_t)stream);
    hipLaunchKernelGGL(k.single, dim3(blocks), dim3(256), 0,
                       (hipStream_t)stream, j, (int)n_tiles);
    CH_HIP(hipGetLastError());
    timer.stop();
    return CH_OK;
}

// Per-stream device job tables for batched launches.
struct JobTable {
    hipStream_t stream;
    ConvJob* d_jobs;
    ConvJob* h_jobs;  // pinned
    int cap;
    hipEvent_t done;
};
static std::vector<JobTable> g_tables;
static std::mutex g_tables_mu;

static ch_status job_table(hipStream_t s, int n, JobTable** out)
{
    for (auto& t : g_tables)
        if (t.stream == s) {
            if (t.cap >= n) {
                // The previous batch on this stream must have consumed the
                // pinned staging copy before it is overwritten.
                CH_HIP(hipEventSynchronize(t.done));
                *out = &t;
                return CH_OK;
            }
            CH_HIP(hipEventSynchronize(t.done));
            CH_HIP(hipFree(t.d_jobs));
            CH_HIP(hipHostFree(t.h_jobs));
            t.cap = std::max(n, 2 * t.cap);
            CH_HIP(hipMalloc((void**)&t.d_jobs, t.cap * sizeof(ConvJob)));
            CH_HIP(hipHostMalloc((void**)&t.h_jobs, t.cap * sizeof(ConvJob),
                                 hipHostMallocDefault));
            *out = &t;
            return CH_OK;
        }
    JobTable t;
    t.stream = s;
    t.cap = std::max(n, 64);
    CH_HIP(hipMalloc((void**)&t.d_jobs, t.cap * sizeof(ConvJob)));
    CH_HIP(hipHostMalloc((void**)&t.h_jobs, t.cap * sizeof(ConvJob),
                         hipHostMallocDefault));
    CH_HIP(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
    CH_HIP(hipEventRecord(t.done, s));
    g_tables.push_back(t);
    *out = &g_tables.back();
    return CH_OK;
}

// ------------------------------------------------------- support box ----

__global__ __launch_bounds__(64) void box_init_kernel(int32_t* box)
{
    if (threadIdx.x < 3)
        box[threadIdx.x] = INT_MAX;
    else if (threadIdx.x < 6)
        box[threadIdx.x] = INT_MIN;
}

__global__ __launch_bounds__(256) void box_kernel(const float* __restrict__ T,
                                                  int X, int Y, int Z,
                                                  int32_t* box)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)X * Y * Z;
    int mn0 = INT_MAX, mn1 = INT_MAX, mn2 = INT_MAX;
    int mx0 = INT_MIN, mx1 = INT_MIN, mx2 = INT_MIN;
    if (i < n && T[i] != 0.0f) {
        const int z = (int)(i % Z);
        const int64_t r = i / Z;
        const int y = (int)(r % Y);
        const int x = (int)(r / Y);
        mn0 = mx0 = x;
        mn1 = mx1 = y;
        mn2 = mx2 = z;
    }
    mn0 = wave_min(mn0);
    mn1 = wave_min(mn1);
    mn2 = wave_min(mn2);
    mx0 = wave_max(mx0);
    mx1 = wave_max(mx1);
    mx2 = wave_max(mx2);
    if ((threadIdx.x & 63) == 0 && mx0 != INT_MIN) {
        atomicMin(box + 0, mn0);
        atomicMin(box + 1, mn1);
        atomicMin(box + 2, mn2);
        atomicMax(box + 3, mx0);
        atomicMax(box + 4, mx1);
        atomicMax(box + 5, mx2);
    }
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_support_box(const float* d_terma, const int32_t h_shape[3],
                         int32_t* d_box, void* stream)
{
    CH_REQUIRE(d_terma && h_shape && d_box, "ch_support_box: null pointer");
    const int64_t n = (int64_t)h_shape[0] * h_shape[1] * h_shape[2];
    CH_REQUIRE(n > 0, "ch_support_box: empty grid");
    hipLaunchKernelGGL(box_init_kernel, dim3(1), dim3(64), 0,
                       (hipStream_t)stream, d_box);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(box_kernel, dim3(blocks), dim3(256), 0,
                       (hipStream_t)stream, d_terma, h_shape[0], h_shape[1],
                       h_shape[2], d_box);
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
    std::lock_guard<std::mutex> lock(g_tables_mu);
    // Group consecutive jobs that share tile shape and accumulation mode;
    // each group is one launch over all of its jobs' tiles.
    int32_t i0 = 0;
    while (i0 < n_jobs) {
        const ch_plan* p0 = h_jobs[i0].plan;
        CH_REQUIRE(p0, "job %d: null plan", i0);
        int32_t i1 = i0 + 1;
        while (i1 < n_jobs && h_jobs[i1].plan &&
               h_jobs[i1].plan->tile[0] == p0->tile[0] &&
               h_jobs[i1].plan->tile[1] == p0->tile[1] &&
               h_jobs[i1].plan->tile[2] == p0->tile[2] &&
               h_jobs[i1].plan->accum_mode == p0->accum_mode)
            ++i1;
        Kernels<float> k;
        ch_status st = kernels_for<float>(p0, k);
        if (st != CH_OK)
            return st;
        JobTable* tab = nullptr;
        st = job_table(s, i1 - i0, &tab);
        if (st != CH_OK)
            return st;
        int64_t tiles = 0;
        for (int32_t i = i0; i < i1; ++i) {
            const ch_conv_job& hj = h_jobs[i];
            CH_REQUIRE(hj.d_terma && hj.d_dose, "job %d: null buffer", i);
            ConvJob j = make_job(hj.plan, hj.d_terma, hj.d_dose, hj.d_box,
                                 hj.accumulate);
            j.tile_begin = (int32_t)tiles;
            tiles += (int64_t)j.ntx * j.nty * j.ntz;
            CH_REQUIRE(tiles < INT_MAX / 2, "batch has too many tiles");
            tab->h_jobs[i - i0] = j;
        }
        CH_HIP(hipMemcpyAsync(tab->d_jobs, tab->h_jobs,
                              (size_t)(i1 - i0) * sizeof(ConvJob),
                              hipMemcpyHostToDevice, s));
        const unsigned blocks = (unsigned)((tiles + 3) / 4);
        LaunchTimer timer(s);
        hipLaunchKernelGGL(k.batch, dim3(blocks), dim3(256), 0, s,
                           (const ConvJob*)tab->d_jobs, (int)(i1 - i0),
                           (int)tiles);
        CH_HIP(hipGetLastError());
        CH_HIP(hipEventRecord(tab->done, s));
        timer.stop();
        i0 = i1;
    }
    return CH_OK;
}

}  // extern "C"
