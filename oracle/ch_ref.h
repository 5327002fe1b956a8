/*
 * ch_ref.h -- FP64 CPU oracle for the collapsed-cone dose path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (conehead_amd/) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline, never as the thing that is measured or shipped.
 *
 * Every function restates one piece of samuelpeet/conehead in IEEE double
 * arithmetic with floating-point contraction disabled (-ffp-contract=off),
 * citing the reference file:line it follows.
 */
#ifndef CH_REF_H
#define CH_REF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 3D-DDA traversal; follows conehead/dda_3d_c.pyx:33-99.
 * Writes at most `cap` voxels (x,y,z triples) and t-values; returns the full
 * number of steps N (which may exceed cap: the caller checks).
 * `dir_abs_out` (nullable) receives |direction| the way dda_3d_c mutates it. */
int32_t ch_ref_dda_3d(const double direction[3], const int32_t grid_shape[3],
                      const int32_t first_voxel[3], const double voxel_size[3],
                      int32_t* vox_out, double* t_out, int32_t cap,
                      double* dir_abs_out);

/* Discretised cone direction; follows conehead/convolve_c.pyx:95-117. */
void ch_ref_cone_direction(double theta_deg, double phi_deg, double out[3]);

/* Kernel index of a boundary-crossing t-value; follows
 * conehead/convolve_c.pyx:138-145.  Returns -1 when the entry is skipped. */
int32_t ch_ref_kernel_index(double t);

/* Scatter-form collapsed-cone convolution; follows
 * conehead/convolve_c.pyx:77-192 (loop order x,y,z,theta,phi,m, FP64 sums).
 * terma, dose: C-order [X][Y][Z] doubles; dose is accumulated in place.
 * dirs: n_cones x 3 cone directions (already discretised), cone_row[c] is the
 * row of `cum` used by cone c; cum is n_rows x row_len (row_len >= 5996).
 * n_threads > 1 splits source voxels over OpenMP threads with private dose
 * buffers reduced in thread order.  Returns the number of dose += T*k
 * executions (the DDA-step update count), or a negative error code. */
int64_t ch_ref_convolve_dirs(const double* terma, double* dose,
                             const int32_t shape[3], const double dims[3],
                             int32_t n_cones, const double* dirs,
                             const int32_t* cone_row, int32_t row_len,
                             const double* cum, int32_t n_threads);

/* Same, taking thetas/phis in degrees exactly like convolve_c's signature
 * (cones ordered theta-major, phi-minor; row j of cum belongs to phis[j]). */
int64_t ch_ref_convolve(const double* terma, double* dose,
                        const int32_t shape[3], const double dims[3],
                        const double* thetas, int32_t n_theta,
                        const double* phis, int32_t n_phi,
                        int32_t row_len, const double* cum, int32_t n_threads);

/* Stencil (gather) form of the same sum: for every cone, the DDA offsets
 * and kernel increments are built once from the origin, then
 * dose[v] += sum_c sum_m k_{c,m} * terma[v - o_{c,m}].  Used to check the
 * translation-invariance identity against the scatter form. */
int64_t ch_ref_convolve_gather(const double* terma, double* dose,
                               const int32_t shape[3], const double dims[3],
                               int32_t n_cones, const double* dirs,
                               const int32_t* cone_row, int32_t row_len,
                               const double* cum);

/* Per-cone tap table (offset, increment) as used by the gather form;
 * returns the number of taps written (<= cap) or -1 on overflow. */
int32_t ch_ref_cone_taps(const double dir[3], const int32_t shape[3],
                         const double dims[3], int32_t row_len,
                         const double* cum_row, int32_t* off_out,
                         double* k_out, int32_t cap);

/* Physics mode: cones aligned with the beam (the caller rotates the
 * directions) and the kernel looked up at the radiological distance.
 * The reference lists both as milestones ("Kernel tilting", README.md:22)
 * and has no implementation; this is the build's own definition, stated in
 * gather form for output voxel v and cone c with path offsets o_m and exit
 * t-values t_m (ch_ref_cone_path):
 *   R_m = sum_{j<=m} rho(v - o_j) (t_j - t_{j-1}),  t_{-1} = 0,
 *   dose[v] += sum_c sum_m T(v - o_m) (cum[i(R_m)] - cum[i(R_{m-1})]),
 * with i() the cuda_dose index (ch_ref_kernel_index_cuda) and the walk ending
 * where v - o_m leaves the grid.  radiological = 0 uses R_m = t_m, which is
 * the cuda_dose convolution (conehead/conehead.py:385-488) in gather form.
 * density is float32 (the engine's densities); returns the number of
 * T != 0 updates, or a negative code. */
int32_t ch_ref_kernel_index_cuda(double r);
int32_t ch_ref_cone_path(const double dir[3], const int32_t shape[3],
                         const double dims[3], int32_t* off_out,
                         double* t_out, int32_t cap);
int64_t ch_ref_convolve_rad(const double* terma, const float* density,
                            double* dose, const int32_t shape[3],
                            const double dims[3], int32_t n_cones,
                            const double* dirs, const int32_t* cone_row,
                            int32_t row_len, const double* cum,
                            int32_t radiological, int32_t n_threads);

/* Beam description for the TERMA chain. */
typedef struct ch_ref_beam {
    double source_position[3];
    double v_y[3];
    double transform[9];   /* row-major world -> source rotation */
    double sad;
    double sPri, sAnn, zAnn, rInner, rOuter, zExp, sExp, kExp;
    double softRatio, softLimit, hornRatio;
    int32_t samples;       /* settings['fluenceResampling'] */
    int32_t n_energy;
    double energy[32];
    double weight[32];
    double mu[32];
} ch_ref_beam;

/* TERMA chain; follows conehead/conehead.py:71-120 (hit test), :209-254
 * (effective depth), :286-314 (off-axis distance), :320-360 (fluence),
 * :641-655 (softening, horn) and :364-377 (TERMA).
 * density: [X][Y][Z] float32 densities on the dose grid; raster: the
 * 4000x4000 float32 block raster (block.block_values).  Any of the optional
 * outputs may be NULL.  Returns 0, or -1 when a ray is parallel to the
 * isocentre plane (the reference raises RuntimeError there). */
int32_t ch_ref_terma(const ch_ref_beam* beam, const int32_t shape[3],
                     const double origin[3], const double spacing[3],
                     const float* density, const float* raster,
                     double* terma, double* blocked, double* d_eff,
                     double* oad, double* fluence, int32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif /* CH_REF_H */
