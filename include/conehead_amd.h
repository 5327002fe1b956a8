/*
 * conehead_amd.h -- C-ABI of the MI355X collapsed-cone dose engine
 * (libconehead_amd.so, hand-written HIP kernels for gfx950).
 *
 * Plain pointers and sizes only.  Arguments named d_* are device pointers
 * (from ch_malloc), h_* are host pointers.  `stream` is a hipStream_t from
 * ch_stream_create, or NULL for the null stream.  Every function returns a
 * ch_status (0 on success); ch_last_error() gives the message of the last
 * failure on the calling thread.  Nothing here calls exit().
 *
 * Array layout: C order [X][Y][Z] (z fastest), like every reference array.
 */
#ifndef CONEHEAD_AMD_H
#define CONEHEAD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ch_status;
#define CH_OK 0
#define CH_ERR_ARG (-1)      /* bad argument (shape, size, null pointer)   */
#define CH_ERR_CAPACITY (-2) /* output capacity too small                  */
#define CH_ERR_HIP (-3)      /* HIP runtime error                          */
#define CH_ERR_NOMEM (-4)    /* device allocation failed                   */
#define CH_ERR_GEOMETRY (-5) /* ray parallel to the isocentre plane        */

/* ------------------------------------------------------------ runtime -- */

const char* ch_last_error(void);
int32_t ch_version(void);
/* Selects the device for the calling thread and reports its properties. */
ch_status ch_init(int32_t device, int32_t* n_cu, int64_t* hbm_bytes,
                  char* name_out, int32_t name_cap);
ch_status ch_malloc(void** d_ptr, size_t bytes);
ch_status ch_free(void* d_ptr);
ch_status ch_mem_info(size_t* free_bytes, size_t* total_bytes);
ch_status ch_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes,
                        void* stream);
ch_status ch_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes,
                        void* stream);
/* Queues a device-to-host copy and returns at once.  With a pinned h_dst
 * (ch_host_alloc) the copy runs on the DMA engine in stream order; the
 * data is there once later work on the stream (an event) has completed. */
ch_status ch_memcpy_d2h_async(void* h_dst, const void* d_src, size_t bytes,
                              void* stream);
/* Device-to-device copy by a kernel (16-byte aligned pointers). */
ch_status ch_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes,
                        void* stream);
/* Zero-fills with a kernel (never hipMemset), so it is safe in captures. */
ch_status ch_zero(void* d_ptr, size_t bytes, void* stream);
/* Elementwise conversion of n values between FP32 and FP64 grids by a
 * kernel: d_dst = d_src, or d_dst += d_src with accumulate != 0 (the sum
 * rounded once in the destination type).  Used by the convolve_c drop-in
 * to bring its FP64 TERMA to FP32 and add the FP32 result into its FP64
 * dose on the device (conehead/convolve_c.pyx:17-20 takes and accumulates
 * float64 arrays). */
#define CH_F32 0
#define CH_F64 1
ch_status ch_convert(void* d_dst, int32_t dst_type, const void* d_src,
                     int32_t src_type, size_t n, int32_t accumulate,
                     void* stream);
ch_status ch_host_alloc(void** h_ptr, size_t bytes);   /* pinned */
ch_status ch_host_free(void* h_ptr);
ch_status ch_stream_create(void** stream);
ch_status ch_stream_destroy(void* stream);
ch_status ch_stream_sync(void* stream);
ch_status ch_device_sync(void);
ch_status ch_event_create(void** event);
ch_status ch_event_destroy(void* event);
ch_status ch_event_record(void* event, void* stream);
/* Later work on `stream` waits for `event` (recorded on another stream). */
ch_status ch_stream_wait(void* stream, void* event);
ch_status ch_event_sync(void* event);
/* 1 if the event has completed, 0 if not. */
ch_status ch_event_query(void* event, int32_t* done);
ch_status ch_event_elapsed(float* ms, void* start, void* stop);

/* -------------------------------------------------------------- DDA ---- */

/* One 3D-DDA traversal, run on the GPU in FP64.
 * Replaces: cdef void dda_3d_c(float64* direction, int32* grid_shape,
 *   int32* first_voxel, float64* voxel_size, result* r) nogil
 *   (conehead/dda_3d_c.pxd:8, conehead/dda_3d_c.pyx:13-99).
 * Differences: caller-owned flat outputs instead of the malloc'd `vector`
 * elements of conehead/vector.pyx:2-54; `h_dir` is not mutated (|dir| is
 * written to h_dir_abs_out when non-NULL, mirroring dda_3d_c.pyx:38-43);
 * capacity overflow returns CH_ERR_CAPACITY with *n_out = required size. */
ch_status ch_dda_3d(const double h_dir[3], const int32_t h_shape[3],
                    const int32_t h_first[3], const double h_vsize[3],
                    int32_t* h_vox_out, double* h_t_out, int32_t cap,
                    int32_t* n_out, double* h_dir_abs_out);

/* Many traversals in one launch (one ray per lane), same arithmetic.
 * d_dirs: n x 3, d_first: n x 3; ray r writes d_vox[r*cap*3 ...] and
 * d_t[r*cap ...] and its full step count to d_count[r] (entries beyond cap
 * are not written).  Replaces the per-ray dda_3d_c calls of
 * conehead/convolve_c.pyx:121-129 and the traversal of
 * conehead/conehead.py:123-206 (cuda_dda_3d). */
ch_status ch_dda_trace_batch(int32_t n_rays, const double* d_dirs,
                             const int32_t* d_first, const int32_t h_shape[3],
                             const double h_vsize[3], int32_t cap,
                             int32_t* d_vox, double* d_t, int32_t* d_count,
                             void* stream);

/* ------------------------------------------------------ cone plan ------ */

typedef struct ch_plan ch_plan;

#define CH_SEM_CONVOLVE_C 0 /* |floor(100 t - 50)|, skip > 5995
                               (conehead/convolve_c.pyx:138-145)       */
#define CH_SEM_CUDA_DOSE 1  /* |floor(100 t - 5)|, clamp at 5995
                               (conehead/conehead.py:453-463)          */
#define CH_SEM_RADIOLOGICAL 2 /* physics mode: the cuda_dose index taken
                               at the radiological distance R instead of
                               t; plans of this kind are used through
                               ch_convolve_rad only                    */

/* Builds the per-cone tap tables on the device: for every cone, the FP64
 * DDA from the origin (offsets o_m, kernel increments
 * k_m = cum[idx_m] - cum[idx_{m-1}], k_0 = cum[idx_0]) and the per-axis
 * first-step tables used to clip each output voxel's tap range.
 * h_dirs: n_cones x 3 discretised directions; h_cone_row[c]: row of h_cum
 * used by cone c; h_cum: n_rows x row_len (row_len >= 5996).
 * Replaces the per-call kernel staging of conehead/convolve_c.pyx:66-68,
 * :236-274 (process_kernel / get_single_cone_data) and the per-voxel index
 * computation of :133-150. */
ch_status ch_plan_create(const int32_t h_shape[3], const double h_vsize[3],
                         int32_t n_cones, const double* h_dirs,
                         const int32_t* h_cone_row, int32_t n_rows,
                         int32_t row_len, const double* h_cum,
                         int32_t semantics, ch_plan** plan_out);
ch_status ch_plan_destroy(ch_plan* plan);
/* Live graphs (ch_graph_end) that read the plan.  While there are any,
 * ch_plan_spectrum_release refuses and ch_plan_destroy is deferred to the
 * destruction of the last of them.  No reference counterpart. */
ch_status ch_plan_graphs(const ch_plan* plan, int32_t* n_graphs);
/* Totals of the plan: number of cones, total taps, longest tap list. */
ch_status ch_plan_info(const ch_plan* plan, int32_t* n_cones,
                       int64_t* n_taps, int32_t* max_taps);
/* Copies cone c's taps to the host (offsets n x 3, FP64 increments). */
ch_status ch_plan_cone_taps(const ch_plan* plan, int32_t cone,
                            int32_t* h_off_out, double* h_k_out, int32_t cap,
                            int32_t* n_out);

/* ------------------------------------------------------ convolution ---- */

/* Support box of a TERMA grid: d_box[0..2] = min (x,y,z) with T != 0,
 * d_box[3..5] = max; an all-zero grid gives lo > hi. */
ch_status ch_support_box(const float* d_terma, const int32_t h_shape[3],
                         int32_t* d_box, void* stream);

/* dose[v] (+)= sum_c sum_m k_{c,m} * terma[v - o_{c,m}] over in-grid
 * sources, i.e. convolve_c's result as a deterministic gather in FP32 with
 * compensated sums.  d_box (nullable: whole grid) restricts the sources to
 * the TERMA support box.  accumulate != 0 adds into d_dose like convolve_c
 * does (conehead/convolve_c.pyx:192); 0 overwrites.
 * Replaces: def convolve_c(float64[:,:,:] dose_grid_terma,
 *   float64[:,:,:] dose_grid_dose, float64[:] dose_grid_dim,
 *   float64[:] thetas, float64[:] phis, kernel)
 *   (conehead/convolve_c.pyx:17-20) and cuda_dose
 *   (conehead/conehead.py:385-488). */
ch_status ch_convolve(const ch_plan* plan, const float* d_terma,
                      float* d_dose, const int32_t* d_box, int32_t accumulate,
                      void* stream);

/* Same convolution with an FP64 dose accumulated in place, for the
 * convolve_c drop-in (its dose grid is float64). */
ch_status ch_convolve_f64(const ch_plan* plan, const float* d_terma,
                          double* d_dose, const int32_t* d_box, void* stream);

/* One gather launch over many (terma, dose) jobs that may use different
 * plans (grid shapes, voxel sizes, cone sets): the grid covers the tiles of
 * every job and each wavefront finds its job by binary search over the
 * jobs' first tiles.  Jobs whose plans use different tile shapes or
 * accumulation modes go into one launch per (tile, mode). */
typedef struct ch_conv_job {
    const ch_plan* plan;
    const float* d_terma;
    float* d_dose;
    const int32_t* d_box;  /* nullable */
    int32_t accumulate;
    int32_t reserved;
} ch_conv_job;
ch_status ch_convolve_batch(int32_t n_jobs, const ch_conv_job* h_jobs,
                            void* stream);

/* ------------------------------------------------- spectral form ------ */

/* Builds the plan's kernel spectrum: the taps of every cone summed into one
 * kernel K on a frame of F_a >= 2 n_a - 1 points per axis (prime factors 2,
 * 3, 5; F_a <= 1024), transformed once.  Idempotent.  Radiological plans
 * are refused: their kernel distance depends on the densities.  No
 * reference counterpart: it precomputes what ch_convolve_fft needs. */
ch_status ch_plan_spectrum(ch_plan* plan, void* stream);
/* Frees the plan's kernel spectrum (waits for the device first: queued
 * kernels may read it); ch_plan_spectrum builds it again.  Refused
 * (CH_ERR_ARG) while a live graph reads the plan (ch_plan_graphs).  No
 * reference counterpart: it bounds the HBM held by cached plans. */
ch_status ch_plan_spectrum_release(ch_plan* plan);
/* Frame lengths, half-spectrum row pitch and the work-grid bytes one beam
 * uses (held per stream) of a plan whose spectrum is built. */
ch_status ch_plan_spectrum_info(const ch_plan* plan, int32_t* frame3,
                                int32_t* pitch, int64_t* work_bytes);
/* The same dose as ch_convolve (convolve_c, conehead/convolve_c.pyx:17-197,
 * as dose[v] (+)= sum_c sum_m k_cm T[v - o_cm]) computed as one linear
 * convolution of T with K through zero-padded FP32 FFTs: five HBM passes
 * per beam whatever the number of cones or taps.  Needs ch_plan_spectrum;
 * float dose; accumulate adds into d_dose.  ch_set_launch_events and
 * ch_set_timing time its convolution pass (the x transforms with the
 * multiplication by K^), the dominant one of the five kernels. */
ch_status ch_convolve_fft(const ch_plan* plan, const float* d_terma,
                          float* d_dose, int32_t accumulate, void* stream);
/* ch_convolve_fft for many jobs (the ch_conv_job list of
 * ch_convolve_batch; d_box is not used): jobs on one plan share each pass's
 * launch, up to 4 GiB of work grids per launch.  Dose buffers must be
 * distinct within a call.  Stages a host table, so it cannot be captured;
 * the tables of a stream rotate through 64 pinned slots, so the call waits
 * on the host only when it runs 64 calls ahead of the stream. */
ch_status ch_convolve_fft_batch(int32_t n_jobs, const ch_conv_job* h_jobs,
                                void* stream);

/* A device-resident convolution set: the jobs of ch_convolve_fft_batch
 * (same grouping: one launch per FFT pass for all jobs of one plan) with
 * their pointer tables written to device memory once, at creation.
 * ch_convolve_fft_set then queues kernels only, so it can be captured in a
 * graph (ch_graph_begin) and replayed; the graph pins the set and its
 * plans (ch_convset_destroy is deferred to the last such graph).  The
 * stream's FFT work grids must already be large enough when it is
 * captured: run the set once on that stream first.  The jobs' buffers and
 * plans must outlive the set.  No counterpart in the reference, whose
 * calculate handles one beam per call (conehead/conehead.py:536-738). */
typedef struct ch_convset ch_convset;
ch_status ch_convset_create(int32_t n_jobs, const ch_conv_job* h_jobs,
                            ch_convset** out);
ch_status ch_convolve_fft_set(const ch_convset* set, void* stream);
ch_status ch_convset_destroy(ch_convset* set);

/* Up to CH_MAX_GROUP beams on the same plan convolved together: their
 * TERMA grids are interleaved so each tap is one 16-byte load for all of
 * them, and the tap walk is shared.  Tap ranges use the union of the
 * beams' support boxes (d_box[b] nullable: whole grid).  Each beam's dose
 * is written to d_dose[b] (accumulate != 0 adds).  The result equals
 * ch_convolve's to FP32 rounding (the union box can shift the blocks of
 * the compensated sum); it is deterministic for a given group. */
#define CH_MAX_GROUP 4
typedef struct ch_conv_group {
    const ch_plan* plan;
    int32_t n_beams;
    int32_t accumulate;
    const float* d_terma[CH_MAX_GROUP];
    float* d_dose[CH_MAX_GROUP];
    const int32_t* d_box[CH_MAX_GROUP];
} ch_conv_group;
ch_status ch_convolve_groups(int32_t n_groups, const ch_conv_group* h_groups,
                             void* stream);

/* Tile shape used by the gather kernels for this plan; set by the
 * autotuner.  1 x TY x TZ (TY in {1, 2, 4, 8}, TY * TZ = 64) selects the
 * direct gather, one wavefront per tile; 4 x 16 x 16 selects the staged
 * gather, one workgroup per block with the TERMA box of each chunk of taps
 * staged through LDS. */
ch_status ch_plan_set_tile(ch_plan* plan, int32_t tx, int32_t ty, int32_t tz);
ch_status ch_plan_get_tile(const ch_plan* plan, int32_t* tx, int32_t* ty,
                           int32_t* tz);
/* Lines per workgroup of the FFT form's five passes (z forward, y
 * forward, x convolution, y inverse, z inverse); set by the autotuner.
 * Each value is 0 (the default: the most lines, up to 16, whose frames fit
 * 2048 points, 4096 for the convolution pass) or 2, 4, 8 or 16, taken when
 * its frames fit 4096 points; a value whose frames do not fit falls back to
 * the default.  Every choice computes the same butterflies
 * on every line, so the dose does not depend on it.  get returns the
 * values the passes use. */
ch_status ch_plan_set_fft_lines(ch_plan* plan, const int32_t lines[5]);
ch_status ch_plan_get_fft_lines(const ch_plan* plan, int32_t lines[5]);
/* Accumulation mode: 0 = Kahan per tap, 1 = plain per cone line with Kahan
 * across lines, 2 = Kahan over blocks of 8 taps (default). */
ch_status ch_plan_set_accum(ch_plan* plan, int32_t mode);

/* Plain scatter convolution, the unbatched baseline: the loop of
 * conehead/convolve_c.pyx:77-192 as one thread per (voxel, cone) -- voxels
 * with T == 0 return, the others walk the cone's FP64 DDA ray from their
 * centre and add T * k into d_dose with FP32 atomics (the structure of the
 * reference GPU kernel cuda_dose, conehead/conehead.py:385-488).  No plan:
 * d_dirs [n_cones][3] (discretised, as ch_plan_create takes them), d_rows
 * [n_cones] and d_cum [rows][row_len] FP64 are device arrays.  d_dose is
 * accumulated (+=); the summation order is nondeterministic. */
ch_status ch_convolve_scatter(const float* d_terma, float* d_dose,
                              const int32_t shape[3], const double vsize[3],
                              int32_t n_cones, const double* d_dirs,
                              const int32_t* d_rows, const double* d_cum,
                              int32_t row_len, int32_t semantics,
                              void* stream);

/* Physics mode (no counterpart in the reference, which lists kernel
 * tilting as a milestone, README.md:22): the cone set is aligned with the
 * beam by the caller (the directions given to ch_plan_create are rotated
 * into the beam frame) and each increment is looked up at the
 * radiological distance from the receiving voxel v:
 *   R_m = sum_{j<=m} rho(v - o_j) (t_j - t_{j-1}),  t_{-1} = 0,
 *   dose[v] (+)= sum_c sum_m T(v - o_m) (cum[i(R_m)] - cum[i(R_{m-1})]),
 * i(R) = min(|floor(100 R - 5)|, 5995) as cuda_dose indexes t
 * (conehead/conehead.py:453-463).  R and the increments are FP64; the sum
 * is FP32 with compensation.  `plan` must be built with
 * CH_SEM_RADIOLOGICAL; d_density is the FP32 density grid of the same
 * shape; d_box (nullable) is the TERMA support box. */
ch_status ch_convolve_rad(const ch_plan* plan, const float* d_terma,
                          const float* d_density, float* d_dose,
                          const int32_t* d_box, int32_t accumulate,
                          void* stream);

/* --------------------------------------------------------- densities --- */

/* Linear resampling of phantom densities onto a dose grid, as the
 * reference's scipy RegularGridInterpolator (linear, fill 0 outside) on the
 * axes linspace(origin, origin + n * spacing, n) of both grids
 * (conehead/conehead.py:525-533, :561-568).  xy_swap != 0 lays the result
 * out as np.meshgrid(indexing='xy') does: d_dst[i][j][k] is the value at
 * (x_j, y_i, z_k), which needs dst_shape[0] == dst_shape[1] (the
 * reference's own grids are then square in x and y); 0 gives the
 * [x][y][z] layout. */
ch_status ch_resample_linear(const float* d_src, const int32_t src_shape[3],
                             const double src_origin[3],
                             const double src_spacing[3], float* d_dst,
                             const int32_t dst_shape[3],
                             const double dst_origin[3],
                             const double dst_spacing[3], int32_t xy_swap,
                             void* stream);

/* CT numbers to densities on the phantom grid, before resampling.  The
 * reference lists CT import and density overrides as milestones it has not
 * built (README.md:26-27); its calculate interpolates whatever density grid
 * the phantom holds (conehead/conehead.py:525-533), so this produces that
 * grid.  d_ct holds n CT numbers (CH_CT_I16: int16, CH_CT_F32: float32);
 * each is mapped through the calibration curve (curve_hu strictly
 * increasing, n_points <= CH_CT_MAX_POINTS, host arrays) exactly as
 * numpy.interp(ct, curve_hu, curve_rho) in FP64, rounded to FP32.  With a
 * label map d_labels (uint8 per voxel, nullable) the voxels whose label L
 * has a non-NaN override_rho[L] (256 host floats) take that density
 * instead: a density override per structure. */
#define CH_CT_I16 0
#define CH_CT_F32 1
#define CH_CT_MAX_POINTS 32
ch_status ch_ct_to_density(const void* d_ct, int32_t ct_type, int64_t n,
                           const double* curve_hu, const double* curve_rho,
                           int32_t n_points, const uint8_t* d_labels,
                           const float* override_rho, float* d_rho,
                           void* stream);

/* ------------------------------------------------------------ graphs --- */

/* Capture of a pipeline of calls on `stream` into a HIP graph, replayed by
 * one launch.  Between ch_graph_begin and ch_graph_end the calls made on
 * `stream` (ch_terma, ch_support_box, ch_convolve, ch_convolve_f64,
 * ch_zero, ch_convolve_scatter) are recorded, not run: every one of them
 * is a kernel launch (no memset, no host copy), with the device pointers
 * and sizes it was given.  Run the same pipeline once before capturing it,
 * so its stream scratch and gather workspace exist; a call that would have
 * to allocate or synchronise fails with CH_ERR_ARG instead, and the batch
 * entry points (host job tables) always do.  ch_graph_end ends the capture
 * in every case.  A graph holds what it reads until ch_graph_destroy: the
 * plans used in the capture are pinned (ch_plan_graphs), and the capture
 * stream's scratch and gather workspace are never reallocated (a call on
 * that stream that would need them larger fails with CH_ERR_ARG) nor freed
 * (ch_stream_destroy refuses).  Device buffers passed by the caller stay
 * the caller's to keep alive.  ch_graph_destroy waits for the device.  No
 * counterpart in the reference (its flow re-launches every kernel from
 * Python, conehead/conehead.py:536-682). */
typedef struct ch_graph ch_graph;
ch_status ch_graph_begin(void* stream);
ch_status ch_graph_end(void* stream, ch_graph** graph_out);
ch_status ch_graph_launch(const ch_graph* graph, void* stream);
ch_status ch_graph_destroy(ch_graph* graph);

/* ------------------------------------------------------------ TERMA ---- */

#define CH_MAX_ENERGIES 32

/* Beam aperture: the transmission of the 4000 x 4000 raster at 0.1 mm pitch
 * (block.block_values, conehead/block.py:14-26), given either as the
 * raster itself or analytically, so the hit test never reads a 64 MB array:
 *   CH_APERTURE_RASTER  d_raster, 4000 x 4000 float32;
 *   CH_APERTURE_RECTS   d_table = n x 4 int32 half-open cell rectangles
 *                       [row0, row1) x [col0, col1) of ones (squares);
 *   CH_APERTURE_MLC     d_table = a leaf table from ch_mlc_table.
 * bbox = [row0, row1, col0, col1): every nonzero cell lies inside it (the
 * hit test skips voxels whose projection provably misses it). */
#define CH_APERTURE_RASTER 0
#define CH_APERTURE_RECTS 1
#define CH_APERTURE_MLC 2
typedef struct ch_aperture {
    int32_t kind;
    int32_t n;
    const float* d_raster;
    const int32_t* d_table;
    int32_t bbox[4];
} ch_aperture;

/* Beam description.  Geometry in cm, world frame of the dose grid. */
typedef struct ch_beam {
    double source_position[3];
    double v_y[3];
    double transform[9];        /* row-major world -> source rotation */
    double sad;
    double sPri, sAnn, zAnn, rInner, rOuter, zExp, sExp, kExp;
    double softRatio, softLimit, hornRatio;
    int32_t samples;            /* settings['fluenceResampling'] */
    int32_t n_energy;
    double energy[CH_MAX_ENERGIES];
    double weight[CH_MAX_ENERGIES];
    double mu[CH_MAX_ENERGIES];
    ch_aperture aperture;
    const float* d_density;     /* [X][Y][Z] densities on the dose grid */
    float* d_terma;             /* output [X][Y][Z] */
    int32_t* d_box;             /* output support box (6 ints), nullable */
    int32_t accumulate;         /* 1: d_terma += this beam, d_box widened */
    int32_t reserved;
    int32_t* d_status;          /* nullable: this beam's own geometry flag
                                   (one int32 the caller zeroes; set to 1
                                   when a ray lies in the source plane)
                                   instead of the stream's flag that
                                   ch_terma_check reads */
} ch_beam;

/* TERMA for n beams on one dose grid, queued in order on `stream`: hit
 * test, FP64-DDA radiological depth toward the source, off-axis distance,
 * three-source fluence, softening, horn and the polyenergetic TERMA sum.
 * Runs of beams with distinct d_terma share one grid of workgroups (their
 * parameters staged into a per-stream device table); a beam that adds into
 * the grid of an earlier beam of the run starts a new launch, so
 * accumulating beams never race.  Under graph capture every beam takes its
 * own launches (see ch_beamset for a captured multi-beam launch).
 * Replaces the kernel chain cuda_hit_test, cuda_d_eff, cuda_oad,
 * cuda_fluence, cuda_terma and the host glue between them
 * (conehead/conehead.py:71-377, :536-682).
 * d_debug (nullable): n_beams x 4 x X*Y*Z floats receiving blocked, d_eff,
 * oad and fluence for each beam. */
ch_status ch_terma(int32_t n_beams, const ch_beam* h_beams,
                   const int32_t h_shape[3], const double h_origin[3],
                   const double h_spacing[3], float* d_debug, void* stream);

/* ch_terma without any allocation, for callers that cannot allocate device
 * memory (code under HIP-graph capture on a fresh stream, tight serving
 * loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  Same arguments, kernels and results as ch_terma, except:
 * every beam must carry d_status (its geometry flag, an output the caller
 * zeroes; ch_terma_check does not see these calls), and the launch scratch
 * (cost-class lists and support-box partials) is a fixed 16 MiB array per
 * GPU, held until the process ends, so grids needing more (about 11 M
 * voxels and beyond) are refused with CH_ERR_ARG.  Calls that overlap on
 * one GPU (on different streams or threads) share that array and race;
 * calls queued on one stream do not. */
ch_status ch_terma_noalloc(int32_t n_beams, const ch_beam* h_beams,
                           const int32_t h_shape[3],
                           const double h_origin[3],
                           const double h_spacing[3], float* d_debug,
                           void* stream);

/* A device-resident beam set: n beams on one grid (the ch_terma arguments
 * without debug planes) whose kernel parameters live in a device table and
 * whose launch scratch belongs to the set.  ch_terma_beamset queues the
 * same launches as ch_terma (beams with distinct TERMA grids share one
 * grid of workgroups) and nothing else, so it can be captured in a graph
 * and replayed; the graph pins the set (ch_beamset_destroy is deferred to
 * the last such graph).  ch_beamset_update rewrites the table for new beams
 * (new geometry, apertures or outputs with the same sharing of TERMA grids
 * and support boxes) by a copy queued on `stream`, so launches and replays
 * queued later on that stream see the new beams; it cannot be captured.
 * Beams without their own d_status flag into the set's per-beam flags,
 * which ch_beamset_flags copies to h_flags (n int32) and clears, waiting
 * for `stream`.  The beams' buffers must outlive the set.  The beams of
 * one call of the reference's calculate are its single beam
 * (conehead/conehead.py:536-682); the set is several such calls. */
typedef struct ch_beamset ch_beamset;
ch_status ch_beamset_create(int32_t n_beams, const ch_beam* h_beams,
                            const int32_t h_shape[3],
                            const double h_origin[3],
                            const double h_spacing[3], ch_beamset** out);
ch_status ch_beamset_update(ch_beamset* set, int32_t n_beams,
                            const ch_beam* h_beams, void* stream);
ch_status ch_terma_beamset(const ch_beamset* set, void* stream);
ch_status ch_beamset_flags(ch_beamset* set, int32_t* h_flags, void* stream);
ch_status ch_beamset_destroy(ch_beamset* set);

/* The FP64 traversals of the TERMA rays toward the source from n voxels
 * (d_vox_in: n x 3 indices), with the same direction arithmetic as
 * ch_terma: d_dirs (n x 3) receives each unit direction, d_vox / d_t /
 * d_count the traversal as in ch_dda_trace_batch.  Replaces the ray set-up
 * and cuda_dda_3d call of cuda_d_eff (conehead/conehead.py:226-246). */
ch_status ch_terma_rays(const ch_beam* h_beam, const int32_t h_shape[3],
                        const double h_origin[3], const double h_spacing[3],
                        int32_t n_rays, const int32_t* d_vox_in, int32_t cap,
                        double* d_dirs, int32_t* d_vox, double* d_t,
                        int32_t* d_count, void* stream);

/* Synchronises `stream` and returns CH_ERR_GEOMETRY when a ray of any
 * ch_terma call queued on it since the last check was parallel to the
 * isocentre plane (the reference raises RuntimeError there,
 * conehead/conehead.py:34-35); the flag is then cleared.  One check after a
 * batch of ch_terma calls covers all of them. */
ch_status ch_terma_check(void* stream);

/* Opens a centred square of side `length` cm (sets it to 1, leaves the
 * rest of the raster unchanged), exactly as Block.set_square
 * (conehead/block.py:64-85). */
ch_status ch_raster_square(float* d_raster, double length, void* stream);

/* Device leaf table of a multi-leaf collimator, all positions in integer
 * tenths of a millimetre (floor(mm * 10), as conehead/block.py:117-120
 * converts them): n_pairs leaf pairs with boundaries h_bound[n_pairs+1]
 * (across the leaves), bank A / B tip positions h_a, h_b (along leaf
 * travel) and jaws h_jaw_x[2], h_jaw_y[2].  Leaf edge transmission
 * (0.2/0.5/0.75) and the 15-column tip taper follow conehead/block.py:
 * 130-196; the bank is centred on raster row 2000 and rows outside it are
 * blocked.  d_table holds ch_mlc_table_len(n_pairs) int32; h_bbox_out
 * (nullable) receives the aperture's bounding cell box.  Replaces
 * Block._set_from_plan's raster. */
int32_t ch_mlc_table_len(int32_t n_pairs);
ch_status ch_mlc_table(int32_t* d_table, int32_t n_pairs,
                       const int32_t* h_bound, const int32_t* h_a,
                       const int32_t* h_b, const int32_t h_jaw_x[2],
                       const int32_t h_jaw_y[2], int32_t* h_bbox_out);

/* Materialises the raster of a leaf table (the same cell function the
 * analytic aperture evaluates). */
ch_status ch_raster_mlc(float* d_raster, const int32_t* d_table,
                        void* stream);

/* Kernel time of the last timed launch (the gather, the FFT form's
 * convolution pass, the physics gather or the scatter; a whole
 * ch_convolve_fft_batch call), measured with HIP events when timing is
 * enabled (the call then waits for the launch). */
ch_status ch_set_timing(int32_t enabled);
ch_status ch_last_kernel_ms(float* ms);

/* Arms a pair of events (ch_event_create) for the calling thread: the next
 * timed launch of this thread (as for ch_set_timing) takes them as the
 * kernel's own start and end timestamps (hipExtLaunchKernelGGL), so their
 * interval is the kernel's duration without the dispatch gap; a batched
 * FFT call records them around the whole call.  It does not wait, and it
 * disarms the pair (read it with ch_event_elapsed once `stop` completed).
 * Pass two nulls to disarm.  Launches that prepare the gather (padding,
 * support and slab boxes) and the FFT form's other passes fall outside the
 * pair.  ch_terma takes an armed pair too: `start` at the start of the
 * first beam's first TERMA kernel, `stop` at the end of the last beam's
 * last one.  No counterpart in the reference: it is the measurement hook of
 * the benchmark. */
ch_status ch_set_launch_events(void* start, void* stop);

#ifdef __cplusplus
}
#endif
#endif /* CONEHEAD_AMD_H */
