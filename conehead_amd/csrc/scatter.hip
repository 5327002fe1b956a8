// scatter.hip -- the plain scatter convolution (unbatched baseline path).
//
// The loop of conehead/convolve_c.pyx:77-192 run as one GPU thread per
// (voxel, cone): voxels with T == 0 return at once (:84-85); otherwise the
// thread walks the cone's ray from the voxel centre with the FP64 DDA of
// dda_3d_c (:119-131), turns each exit t into a kernel index (:133-150) and
// adds T * k into the dose with an FP32 atomic (:157-192), the way the
// reference's GPU kernel cuda_dose does (conehead/conehead.py:385-488).
// It needs no plan, so it is the path a straight port would take; its
// summation order is nondeterministic.
#include "dda_core.h"

namespace ch {

__global__ __launch_bounds__(256) void scatter_kernel(
    const float* __restrict__ T, float* __restrict__ dose, int X, int Y,
    int Z, double v0, double v1, double v2, int n_cones,
    const double* __restrict__ dirs, const int32_t* __restrict__ rows,
    const double* __restrict__ cum, int row_len, int semantics)
{
    const int64_t nv = (int64_t)X * Y * Z;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nv * n_cones)
        return;
    const int c = (int)(i / nv);
    const int64_t li = i - (int64_t)c * nv;
    const float t_src = T[li];
    if (t_src == 0.0f)
        return;
    int32_t c2 = (int32_t)(li % Z);
    int32_t c1 = (int32_t)((li / Z) % Y);
    int32_t c0 = (int32_t)(li / ((int64_t)Z * Y));
    Dda d;
    d.init(dirs[3 * c], dirs[3 * c + 1], dirs[3 * c + 2], v0, v1, v2);
    const double* row = cum + (int64_t)rows[c] * row_len;
    const double tv = (double)t_src;
    double prev = 0.0;
    while (c0 >= 0 && c0 < X && c1 >= 0 && c1 < Y && c2 >= 0 && c2 < Z) {
        float* dst = dose + ((int64_t)c0 * Y + c1) * Z + c2;
        const double t = d.advance(d.axis(), c0, c1, c2);
        bool clamped;
        const int32_t idx = kernel_index(t, semantics, &clamped);
        if (idx < 0)
            break;  // t only grows: every later index is skipped too
        const double cv = row[idx];
        atomicAdd(dst, (float)(tv * (cv - prev)));
        prev = cv;
    }
}

}  // namespace ch

using namespace ch;

extern "C" {

ch_status ch_convolve_scatter(const float* d_terma, float* d_dose,
                              const int32_t h_shape[3],
                              const double h_vsize[3], int32_t n_cones,
                              const double* d_dirs, const int32_t* d_rows,
                              const double* d_cum, int32_t row_len,
                              int32_t semantics, void* stream)
{
    CH_REQUIRE(d_terma && d_dose && h_shape && h_vsize,
               "ch_convolve_scatter: null pointer");
    CH_REQUIRE(h_shape[0] > 0 && h_shape[1] > 0 && h_shape[2] > 0,
               "ch_convolve_scatter: empty grid");
    CH_REQUIRE(n_cones >= 0, "ch_convolve_scatter: n_cones < 0");
    CH_REQUIRE(semantics == CH_SEM_CONVOLVE_C || semantics == CH_SEM_CUDA_DOSE,
               "ch_convolve_scatter: unknown semantics %d", semantics);
    CH_REQUIRE(row_len >= kRowLenMin,
               "ch_convolve_scatter: kernel rows need >= %d entries",
               kRowLenMin);
    CH_REQUIRE(h_vsize[0] > 0 && h_vsize[1] > 0 && h_vsize[2] > 0,
               "ch_convolve_scatter: voxel size must be > 0");
    if (n_cones == 0)
        return CH_OK;
    CH_REQUIRE(d_dirs && d_rows && d_cum, "ch_convolve_scatter: null table");
    const int64_t nv = (int64_t)h_shape[0] * h_shape[1] * h_shape[2];
    const int64_t n = nv * n_cones;
    CH_REQUIRE(n < (int64_t)INT32_MAX * 256, "ch_convolve_scatter: too big");
    hipStream_t s = (hipStream_t)stream;
    LaunchTimer timer(s);
    timer.launch(scatter_kernel, dim3((unsigned)((n + 255) / 256)),
                 dim3(256), 0, d_terma, d_dose, h_shape[0], h_shape[1],
                 h_shape[2], h_vsize[0], h_vsize[1], h_vsize[2], n_cones,
                 d_dirs, d_rows, d_cum, row_len, semantics);
    CH_HIP(hipGetLastError());
    timer.stop();
    return CH_OK;
}

}  // extern "C"
