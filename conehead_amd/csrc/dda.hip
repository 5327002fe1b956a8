// dda.hip -- batched FP64 3D-DDA traces (the dda_3d_c counterpart).
#include "dda_core.h"

namespace ch {

// One ray per lane; waves of neighbouring rays have similar lengths.
__global__ __launch_bounds__(256) void dda_trace_kernel(
    int32_t n_rays, const double* __restrict__ dirs,
    const int32_t* __restrict__ first, int32_t X, int32_t Y, int32_t Z,
    double v0, double v1, double v2, int32_t cap, int32_t* __restrict__ vox,
    double* __restrict__ tout, int32_t* __restrict__ count)
{
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rays)
        return;
    Dda d;
    d.init(dirs[3 * r + 0], dirs[3 * r + 1], dirs[3 * r + 2], v0, v1, v2);
    int32_t c0 = first[3 * r + 0], c1 = first[3 * r + 1],
            c2 = first[3 * r + 2];
    int32_t n = 0;
    int32_t* vr = vox + (int64_t)r * cap * 3;
    double* tr = tout + (int64_t)r * cap;
    while (c0 >= 0 && c0 < X && c1 >= 0 && c1 < Y && c2 >= 0 && c2 < Z) {
        const int a = d.axis();
        if (n < cap) {
            vr[3 * n + 0] = c0;
            vr[3 * n + 1] = c1;
            vr[3 * n + 2] = c2;
        }
        const double t = d.advance(a, c0, c1, c2);
        if (n < cap)
            tr[n] = t;
        ++n;
    }
    count[r] = n;
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_dda_trace_batch(int32_t n_rays, const double* d_dirs,
                             const int32_t* d_first, const int32_t h_shape[3],
                             const double h_vsize[3], int32_t cap,
                             int32_t* d_vox, double* d_t, int32_t* d_count,
                             void* stream)
{
    CH_REQUIRE(n_rays >= 0, "n_rays < 0");
    if (n_rays == 0)
        return CH_OK;
    CH_REQUIRE(d_dirs && d_first && d_vox && d_t && d_count && h_shape &&
                   h_vsize,
               "ch_dda_trace_batch: null pointer");
    CH_REQUIRE(cap >= 0, "cap < 0");
    const unsigned blocks = (unsigned)((n_rays + 255) / 256);
    hipLaunchKernelGGL(dda_trace_kernel, dim3(blocks), dim3(256), 0,
                       (hipStream_t)stream, n_rays, d_dirs, d_first,
                       h_shape[0], h_shape[1], h_shape[2], h_vsize[0],
                       h_vsize[1], h_vsize[2], cap, d_vox, d_t, d_count);
    CH_HIP(hipGetLastError());
    return CH_OK;
}

ch_status ch_dda_3d(const double h_dir[3], const int32_t h_shape[3],
                    const int32_t h_first[3], const double h_vsize[3],
                    int32_t* h_vox_out, double* h_t_out, int32_t cap,
                    int32_t* n_out, double* h_dir_abs_out)
{
    CH_REQUIRE(h_dir && h_shape && h_first && h_vsize && n_out,
               "ch_dda_3d: null pointer");
    CH_REQUIRE(cap >= 0 && (cap == 0 || (h_vox_out && h_t_out)),
               "ch_dda_3d: bad output buffers");
    if (h_dir_abs_out)
        for (int a = 0; a < 3; ++a)
            h_dir_abs_out[a] = h_dir[a] < 0 ? -h_dir[a] : h_dir[a];
    // A path visits at most sum(shape) voxels; trace with that capacity so
    // the exact count is known even when the caller's capacity is smaller.
    int64_t full = (int64_t)(h_shape[0] > 0 ? h_shape[0] : 0) +
                   (h_shape[1] > 0 ? h_shape[1] : 0) +
                   (h_shape[2] > 0 ? h_shape[2] : 0) + 1;
    CH_REQUIRE(full < (1 << 28), "ch_dda_3d: grid too large");
    const int32_t dcap = (int32_t)full;
    void* buf = nullptr;
    const size_t bytes = 64 + 32 + sizeof(int32_t) * 4 +
                         (size_t)dcap * (3 * sizeof(int32_t) + sizeof(double));
    ch_status st = ch_malloc(&buf, bytes + 256);
    if (st != CH_OK)
        return st;
    char* base = (char*)buf;
    double* d_dir = (double*)base;
    int32_t* d_first = (int32_t*)(base + 32);
    int32_t* d_count = (int32_t*)(base + 48);
    double* d_t = (double*)(base + 64);
    int32_t* d_vox = (int32_t*)(base + 64 + (size_t)dcap * sizeof(double));
    hipStream_t s = nullptr;
    int32_t n = 0;
    hipError_t e = hipMemcpyAsync(d_dir, h_dir, 3 * sizeof(double),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_first, h_first, 3 * sizeof(int32_t),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(dda_trace_kernel, dim3(1), dim3(64), 0, s, 1,
                           d_dir, d_first, h_shape[0], h_shape[1], h_shape[2],
                           h_vsize[0], h_vsize[1], h_vsize[2], dcap, d_vox,
                           d_t, d_count);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(&n, d_count, sizeof(int32_t), hipMemcpyDeviceToHost);
    const int32_t ncopy = n < cap ? n : cap;
    if (e == hipSuccess && ncopy > 0)
        e = hipMemcpy(h_t_out, d_t, (size_t)ncopy * sizeof(double),
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && ncopy > 0)
        e = hipMemcpy(h_vox_out, d_vox, (size_t)ncopy * 3 * sizeof(int32_t),
                      hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(e, "ch_dda_3d", __FILE__, __LINE__);
    *n_out = n;
    if (n > cap) {
        set_error("ch_dda_3d: %d steps exceed capacity %d", n, cap);
        return CH_ERR_CAPACITY;
    }
    return CH_OK;
}

}  // extern "C"
