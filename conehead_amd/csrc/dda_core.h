// dda_core.h -- FP64 3D-DDA step shared by the trace, plan and TERMA kernels.
//
// Same operations as conehead/dda_3d_c.pyx:33-99: step sign from d < 0,
// |d| in place, t = (s/2)/|d| and dt = s/|d| (1e9 for a zero component),
// then repeated t += dt (never a closed form) and the tie rules
//   t_x < t_y && t_x < t_z -> x;  t_x < t_y -> z;  t_y < t_z -> y;  else z.
// Files including this header are compiled with -ffp-contract=off.
#pragma once

#include "ch_internal.h"

namespace ch {

struct Dda {
    double t0, t1, t2;
    double dt0, dt1, dt2;
    int32_t s0, s1, s2;

    __device__ __forceinline__ void init(double d0, double d1, double d2,
                                         double v0, double v1, double v2)
    {
        s0 = d0 < 0 ? -1 : 1;
        s1 = d1 < 0 ? -1 : 1;
        s2 = d2 < 0 ? -1 : 1;
        d0 = d0 < 0 ? -d0 : d0;
        d1 = d1 < 0 ? -d1 : d1;
        d2 = d2 < 0 ? -d2 : d2;
        if (d0 == 0.0) { t0 = kBig; dt0 = kBig; }
        else { t0 = (v0 / 2) / d0; dt0 = v0 / d0; }
        if (d1 == 0.0) { t1 = kBig; dt1 = kBig; }
        else { t1 = (v1 / 2) / d1; dt1 = v1 / d1; }
        if (d2 == 0.0) { t2 = kBig; dt2 = kBig; }
        else { t2 = (v2 / 2) / d2; dt2 = v2 / d2; }
    }

    // Axis chosen at the current voxel (dda_3d_c.pyx:80-99).
    __device__ __forceinline__ int axis() const
    {
        if (t0 < t1)
            return (t0 < t2) ? 0 : 2;
        return (t1 < t2) ? 1 : 2;
    }

    // Exit t of the current voxel along `a`, then advance (t += dt).
    __device__ __forceinline__ double advance(int a, int32_t& c0, int32_t& c1,
                                              int32_t& c2)
    {
        double t;
        if (a == 0) { t = t0; t0 += dt0; c0 = c0 + s0; }
        else if (a == 1) { t = t1; t1 += dt1; c1 = c1 + s1; }
        else { t = t2; t2 += dt2; c2 = c2 + s2; }
        return t;
    }
};

// Kernel table index of a boundary-crossing t-value.
//  CH_SEM_CONVOLVE_C: |(int32)floor(100 t - 50)|, entries > 5995 are skipped
//    (conehead/convolve_c.pyx:138-145) -> returns -1.
//  CH_SEM_CUDA_DOSE: |floor(100 t - 5)| clamped at 5995
//    (conehead/conehead.py:453-463); *clamped is set when it clamps.
__device__ __forceinline__ int32_t kernel_index(double t, int32_t semantics,
                                                bool* clamped)
{
    *clamped = false;
    const double off = semantics == CH_SEM_CUDA_DOSE ? 5.0 : 50.0;
    double f = floor(t * 100.0 - off);
    if (!(f < 2147483647.0 && f > -2147483648.0)) {
        if (semantics == CH_SEM_CUDA_DOSE) {
            *clamped = true;
            return kMaxIndex;
        }
        return -1;
    }
    int32_t idx = (int32_t)f;
    idx = idx < 0 ? -idx : idx;
    if (idx > kMaxIndex) {
        if (semantics == CH_SEM_CUDA_DOSE) {
            *clamped = true;
            return kMaxIndex;
        }
        return -1;
    }
    return idx;
}

}  // namespace ch
