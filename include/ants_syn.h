/*
 * ants_syn.h -- C-ABI of the MI355X (gfx950) SyN + CC registration library.
 *
 * Drop-in for the SyN stage that antsRegistration drives through
 * RegistrationHelper (Examples/itkantsRegistrationHelper.hxx:1983-2220):
 * `--transform SyN[step,updateVariance,totalVariance]`
 * `--metric CC[fixed,moving,1,radius]` with `-c/-f/-s` multi-resolution
 * schedules.  Images are dense x-fastest float buffers with ITK's geometry
 * (size, spacing, origin, orthonormal direction).  Displacement fields are
 * ITK `Image<Vector<float,D>,D>` buffers: D interleaved float components per
 * voxel (x fastest), physical vectors (mm).
 *
 * Conventions: every function returns ANTS_OK (0) or an error code; no C++
 * exception crosses the ABI.  ants_last_error() returns the calling thread's
 * last message.  Pointers named d_* are device pointers (HBM); host_* are
 * host pointers.  `stream` is a hipStream_t (NULL = the default stream).
 * A handle is single-threaded; distinct handles may be used concurrently.
 */
#ifndef ANTS_SYN_H
#define ANTS_SYN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANTS_OK 0
#define ANTS_ERR_INVALID 1      /* bad argument or parameter combination */
#define ANTS_ERR_HIP 2          /* HIP runtime failure */
#define ANTS_ERR_OOM 3          /* device allocation or admission budget */
#define ANTS_ERR_UNSUPPORTED 4  /* valid in ANTs, not supported here */
#define ANTS_ERR_CANCELLED 5    /* iteration callback asked to stop */

#define ANTS_MAX_LEVELS 8

/* itk::Image geometry (ImageBase: size, spacing, origin, direction). */
typedef struct ants_image_desc {
  int dim;              /* 2 or 3 */
  int64_t size[3];      /* size[2] ignored (treated as 1) when dim == 2 */
  double spacing[3];
  double origin[3];
  double direction[9];  /* row-major dim x dim; all-zero means identity */
} ants_image_desc;

#define ANTS_PRECISION_FP32 0       /* FP32 images, fields and sums */
#define ANTS_PRECISION_BF16_PYRAMID 1 /* BF16 smoothed images, FP32 sums/fields */

/* Field models of the stage (-t SyN[...] / -t BSplineSyN[...]). */
#define ANTS_MODEL_SYN 0          /* Gaussian update/total field smoothing */
#define ANTS_MODEL_BSPLINE_SYN 1  /* B-spline fitted update/total fields */

/* The SyN stage parameters antsRegistration parses
 * (Examples/antsRegistrationTemplateHeader.h:542-815, 985-1086) and
 * RegistrationHelper hands to SyNImageRegistrationMethod
 * (Examples/itkantsRegistrationHelper.hxx:2141-2164). */
typedef struct ants_syn_params {
  int dim;                              /* 2 or 3 */
  int n_levels;                         /* -f/-s/-c entries */
  int shrink[ANTS_MAX_LEVELS];          /* -f 4x2x1 (isotropic factor; per-axis rule of the helper) */
  double sigma[ANTS_MAX_LEVELS];        /* -s 2x1x0 */
  int sigma_in_mm;                      /* -s ...mm (1) or ...vox (0) */
  int iterations[ANTS_MAX_LEVELS];      /* -c [100x70x20,...] */
  double convergence_threshold;         /* -c [...,1e-6,...] */
  int convergence_window;               /* -c [...,...,10] */
  double learning_rate;                 /* SyN[step,...] */
  double update_variance;               /* SyN[.,3,.] (voxel^2) */
  double total_variance;                /* SyN[.,.,0] (voxel^2) */
  int cc_radius;                        /* CC[f,m,1,radius] */
  int precision;                        /* ANTS_PRECISION_* */
  int preprocess;                       /* window both images to [0,1] as PreprocessImage does (1 = ANTs default) */
  int force_iterations;                 /* 1: never stop early (parity harness), 0: convergence monitor */
  /* BSplineSyN[step, updateMesh, totalMesh=0, order=3]
   * (Examples/antsRegistrationTemplateHeader.h:701-743, itkantsRegistrationHelper.hxx:2222-2340):
   * mesh sizes (spans per axis) at the first level, doubled at every later
   * level as the helper's transform adaptors do; a knot spacing is converted
   * to a mesh by the caller (RegistrationHelper::CalculateMeshSizeForSpecified-
   * KnotSpacing, hxx:3737-3767).  total_mesh all zero: the total field is not
   * fitted.  update_variance / total_variance are unused in this model. */
  int model;                            /* ANTS_MODEL_* (0 = SyN) */
  int update_mesh[3];
  int total_mesh[3];
  int spline_order;                     /* 1..3 (BSplineSyN default 3) */
} ants_syn_params;

/* Per-iteration observer (antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate
 * IterationEvent, Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:117-195).
 * inv_iters[4]: inner steps of the four field inversions (F2M inverse,
 * F2M forward, M2M inverse, M2M forward).  Return non-zero to cancel. */
typedef int (*ants_syn_iteration_cb)(void* user, int level, int iteration, double energy, double convergence,
                                     const int* inv_iters);

typedef struct ants_syn_handle ants_syn_handle;

/* ---- library / device helpers (no counterpart in ANTs: device plumbing) ---- */
const char* ants_last_error(void);
int ants_version(void);
int ants_device_malloc(void** d_ptr, size_t bytes);
int ants_device_free(void* d_ptr);
int ants_memcpy_h2d(void* d_dst, const void* host_src, size_t bytes, void* stream);
int ants_memcpy_d2h(void* host_dst, const void* d_src, size_t bytes, void* stream);
int ants_memset_device(void* d_dst, int value, size_t bytes, void* stream);
/* A plain kernel copy of `bytes` (a multiple of 16) between device buffers:
 * the known-size transfer that HBM byte counters are calibrated against. */
int ants_copy_calibration(void* d_dst, const void* d_src, size_t bytes, void* stream);
int ants_stream_create(void** stream);
int ants_stream_destroy(void* stream);
int ants_stream_synchronize(void* stream);
int ants_device_synchronize(void);
/* hipMemGetInfo of the current device (free and total bytes; either may be NULL). */
int ants_device_memory_info(size_t* free_bytes, size_t* total_bytes);
int ants_event_create(void** event);
int ants_event_destroy(void* event);
int ants_event_record(void* event, void* stream);
int ants_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---- the SyN stage ---- */

/* Level grid of level `level` (ShrinkImageFilter output information with
 * RegistrationHelper::CalculateShrinkFactorsPerDimension,
 * Examples/itkantsRegistrationHelper.hxx:16-39, 593-641). */
int ants_syn_level_geometry(const ants_syn_params* p, const ants_image_desc* fixed, int level, ants_image_desc* out);

/* Replaces SyNImageRegistrationMethod::New() + its setters
 * (Examples/itkantsRegistrationHelper.hxx:2001-2164).  The handle owns its
 * stream and device workspace. */
int ants_syn_create(const ants_syn_params* p, ants_syn_handle** out);
void ants_syn_destroy(ants_syn_handle* h);
/* Use a caller stream instead of the handle's own. */
int ants_syn_set_stream(ants_syn_handle* h, void* stream);

/* Replaces SyNImageRegistrationMethod::Update() + GetModifiableTransform()
 * (Examples/itkantsRegistrationHelper.hxx:2198-2218), host buffers in and
 * out.  out_fwd / out_inv: D x (fixed voxels) floats each (the output
 * DisplacementFieldTransform's displacement and inverse displacement
 * fields on the last level's grid).  Either may be NULL. */
int ants_syn_register(ants_syn_handle* h, const ants_image_desc* fixed, const float* host_fixed,
                      const ants_image_desc* moving, const float* host_moving, float* host_out_fwd,
                      float* host_out_inv, ants_syn_iteration_cb cb, void* user);

/* Same, with the images already resident in HBM and device outputs. */
int ants_syn_register_device(ants_syn_handle* h, const ants_image_desc* fixed, const float* d_fixed,
                             const ants_image_desc* moving, const float* d_moving, float* d_out_fwd,
                             float* d_out_inv, ants_syn_iteration_cb cb, void* user);

/* Registration state after a run: FixedToMiddle (forward, inverse) and
 * MovingToMiddle (forward, inverse) fields, as --save-state keeps them
 * (Examples/itkantsRegistrationHelper.hxx:2206-2215).  Device buffers of
 * D x voxels floats each; any may be NULL. */
int ants_syn_get_state_device(ants_syn_handle* h, float* d_f2m_fwd, float* d_f2m_inv, float* d_m2m_fwd,
                              float* d_m2m_inv);

/* Replaces SyNImageRegistrationMethod::SetFixedToMiddleTransform +
 * SetMovingToMiddleTransform when antsRegistration restores a saved state
 * (--restore-state with --initialize-transforms-per-stage,
 * Examples/itkantsRegistrationHelper.hxx:2091-2117): later runs start from
 * these four fields on `grid`, resampled onto their level-0 grid (as fields
 * are resampled between levels), instead of zero fields.  Device buffers of
 * D x (grid voxels) floats each, copied before the call returns; all four
 * NULL clears the state.  A run split into two stages this way equals the
 * single stage. */
int ants_syn_set_initial_state_device(ants_syn_handle* h, const ants_image_desc* grid, const float* d_f2m_fwd,
                                      const float* d_f2m_inv, const float* d_m2m_fwd, const float* d_m2m_inv);

/* WindowConvergenceMonitoringFunction as the SyN stage uses it to stop a
 * level (described at Examples/antsRegistration.cxx:397-404; the stage's
 * ConvergenceWindowSize and ConvergenceThreshold at
 * Examples/itkantsRegistrationHelper.hxx:2157-2158): the convergence value
 * after energies[0..n) were added in order, DBL_MAX while fewer than
 * `window` exist.  The engine stops a level when it is below the threshold. */
int ants_syn_convergence_value(const double* energies, int n, int window, double* out);

/* Work done by the last run: iterations per level, inner inversion steps,
 * voxel-iterations and the algorithmic HBM byte count (per voxel-iteration
 * B_vi = 2*g*e_i + 14*D*e_f + sum_j (3*K_j + 1)*D*e_f, summed). */
typedef struct ants_syn_stats {
  int levels_run;
  int iterations[ANTS_MAX_LEVELS];
  int64_t level_voxels[ANTS_MAX_LEVELS];
  int64_t inv_steps[ANTS_MAX_LEVELS];   /* sum over iterations of the 4 inner-step counts */
  double voxel_iterations;              /* sum_l N_l * iterations_l */
  double algorithmic_bytes;             /* sum over iterations of N_l * B_vi */
  int64_t graph_launches;               /* iterations replayed from a captured hipGraph */
  int64_t eager_iterations;             /* iterations launched kernel by kernel */
  int64_t inv_tails;                    /* iterations whose inversions ran past the steps launched with them */
} ants_syn_stats;
int ants_syn_get_stats(const ants_syn_handle* h, ants_syn_stats* out);

/* The last run's iteration records: the observer's arguments
 * (Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:117-195),
 * kept whether or not an observer was passed, so that a caller can read the
 * trace after the run instead of being called back between iterations.
 * Writes min(capacity, total) records to out (out may be NULL when capacity
 * is 0) and the total to *count. */
typedef struct ants_syn_iteration_record {
  int level;
  int iteration;
  double energy;
  double convergence;
  int inv_iters[4];
  double full_scale_cc;  /* the observer's full-scale CC r=4 (0 when not computed at this iteration) */
} ants_syn_iteration_record;
int ants_syn_get_trace(const ants_syn_handle* h, ants_syn_iteration_record* out, int64_t capacity, int64_t* count);

/* antsRegistration --print-similarity-measure-interval N
 * (Examples/antsRegistrationTemplateHeader.h:175-195, set on the SyN observer
 * at itkantsRegistrationHelper.hxx:2176-2181): at iteration 1, every N-th
 * iteration and the last iteration of each level, the observer's CC metric of
 * radius 4 between the ORIGINAL fixed and moving images, each resampled onto
 * the fixed image's grid through its composite [initial transforms, middle
 * transform inverse] (linear interpolation, 0 outside)
 * (Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:
 * 151-159, 252-397).  The value lands in the iteration record's
 * full_scale_cc and, inside the observer callback, in
 * ants_syn_last_full_scale_cc (0 at iterations where it is not computed).
 * interval 0 (default) turns it off.  Costs two full-grid resamples and a
 * full-grid CC per computed iteration. */
int ants_syn_set_full_scale_cc_interval(ants_syn_handle* h, int interval);
int ants_syn_last_full_scale_cc(const ants_syn_handle* h, double* value);

/* Kernel classes timed with HIP events on the handle's stream when enabled
 * by ants_syn_set_profiling (bit k enables class k).  A profiled run launches
 * its iterations kernel by kernel (no graph replay).  The inversion class is
 * timed as one span per iteration over all its launches. */
#define ANTS_KC_WARP 0     /* image warps through the inverse middle fields */
#define ANTS_KC_CC 1       /* CC value + both derivatives */
#define ANTS_KC_SMOOTH 2   /* update/total field Gaussian passes */
#define ANTS_KC_COMPOSE 3  /* scaled composition onto the forward fields */
#define ANTS_KC_INV 4      /* fused fixed-point inversion steps (active launches only) */
#define ANTS_KC_COUNT 5
typedef struct ants_syn_kernel_stats {
  double ms[ANTS_KC_COUNT];        /* summed event time of timed launches */
  int64_t launches[ANTS_KC_COUNT]; /* timed launches */
  double bytes[ANTS_KC_COUNT];     /* algorithmic HBM bytes of those launches */
} ants_syn_kernel_stats;
int ants_syn_set_profiling(ants_syn_handle* h, unsigned int class_mask);
/* Time only iterations 1, 1 + every, 1 + 2·every, ... of each level (default
 * 1: all).  Each timing span records two events, and an event between two
 * launches keeps the second from starting until the first has finished:
 * sampling keeps most iterations free of that stall. */
int ants_syn_set_profiling_stride(ants_syn_handle* h, int every);
int ants_syn_get_kernel_stats(const ants_syn_handle* h, ants_syn_kernel_stats* out);
/* The same restricted to the spans of pyramid level `level` of the last run. */
int ants_syn_get_kernel_stats_level(const ants_syn_handle* h, int level, ants_syn_kernel_stats* out);

/* ---- launch-shape tuning (no counterpart in ANTs: device launch plumbing) ----
 * A table of launch shapes per kernel and grid-size bucket
 * (log2_voxels = floor(log2(level voxels))), written by the project's
 * autotuner (ants_amd/autotune.py) and loaded when the library is loaded
 * from tuning.txt beside it (ANTS_SYN_TUNING=<file> or "none" overrides).
 * Knobs: 0 inversion blocks per problem, 1 CC z-chunk planes, 2 smoothing
 * z-chunk planes, 3 CC tile (1: 16x16, 2: 16x8), 4 smoothing tile
 * (1: 16x32, 2: 16x16); value 0 = no entry (the built-in launch model
 * decides).  Not thread-safe against running registrations: change the
 * table between runs. */
#define ANTS_TUNE_INV_BLOCKS 0
#define ANTS_TUNE_CC_ZCHUNK 1
#define ANTS_TUNE_SMOOTH_ZCHUNK 2
#define ANTS_TUNE_CC_TILE 3
#define ANTS_TUNE_SMOOTH_TILE 4
int ants_tuning_set(int knob, int log2_voxels, int value);
int ants_tuning_get(int knob, int log2_voxels, int* value);
int ants_tuning_clear(void);
int ants_tuning_load(const char* path);

/* ---- building blocks (device buffers) ----
 * Each replaces the ITK filter that SyN calls at the cited ANTs site. */

/* DiscreteGaussianImageFilter (variance sigma^2, MaximumError 0.01, width 32),
 * ImageRegistrationMethodv4 per-level smoothing (itkantsRegistrationHelper.hxx:2149-2150). */
int ants_op_smooth_image(const ants_image_desc* g, const float* d_in, float* d_out, double sigma, int sigma_in_mm,
                         void* stream);

/* SyN GaussianSmoothDisplacementField(field, variance): per-axis
 * GaussianOperator(MaxError 0.001), boundary zeroed, blend below 0.5
 * (variance set at itkantsRegistrationHelper.hxx:2161-2162). */
int ants_op_smooth_field(const ants_image_desc* g, const float* d_in, float* d_out, double variance, void* stream);

/* The centre of gravity of an image (ImageMomentsCalculator::GetCenterOf-
 * Gravity: the intensity-weighted mean physical point), as
 * CenteredTransformInitializer::MomentsOn uses it for antsRegistration
 * -r [fixed,moving,1] (Examples/itkantsRegistrationHelper.h:1296-1348: the
 * initial transform centred at the fixed image's centre of gravity,
 * translated by moving minus fixed).  center[dim] receives the point, mass
 * (optional) the intensity total; ANTS_ERR_INVALID when the total is 0. */
int ants_image_center_of_mass(const ants_image_desc* g, const float* d_img, double* center, double* mass,
                              void* stream);

/* BSplineSyN's BSplineSmoothDisplacementField(field, mesh + order control
 * points): DisplacementFieldToBSplineImageFilter with one fitting level,
 * stationary boundary (boundary voxels fitted as zero vectors of weight
 * 1e10), no confidence image, evaluated back on the field's grid -- as the
 * helper configures BSplineSyN (itkantsRegistrationHelper.hxx:2222-2340) and
 * as SmoothDisplacementField's B-spline mode calls the filter with one level
 * (Examples/SmoothDisplacementField.cxx:134-165).  mesh: spans per axis
 * (>= 1; every axis needs >= 2 voxels), order 1..3.  host_max (optional)
 * receives the output's max |u / spacing|. */
int ants_op_bspline_field(const ants_image_desc* g, const float* d_in, float* d_out, const int* mesh, int order,
                          double* host_max, void* stream);

/* ResampleImageFilter + LinearInterpolateImageFunction through x -> x + u(x)
 * (u on the output grid, NULL = identity), default value 0
 * (itkantsRegistrationHelper.hxx:719-763). */
int ants_op_warp_image(const ants_image_desc* g_img, const float* d_img, const ants_image_desc* g_out,
                       const float* d_field, float* d_out, void* stream);

/* The two resamplings of one SyN iteration (SyN ComputeMetricGradientField
 * with DownsampleImagesForMetricDerivatives, mirrored in
 * Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:342-381):
 * on level grid g_level, fixed_out(x) = fixed(x + u_fixed(x)) and
 * moving_out(x) = moving(x + u_moving(x)), trilinear, 0 outside each image's
 * buffer; u_* are D x (level voxels) fields.  bf16_images != 0 rounds both
 * images to BF16 first, as the BF16 pyramid form stores them. */
int ants_op_warp_pair(const ants_image_desc* g_level, const ants_image_desc* g_fixed, const float* d_fixed,
                      const float* d_u_fixed, float* d_fixed_out, const ants_image_desc* g_moving,
                      const float* d_moving, const float* d_u_moving, float* d_moving_out, int bf16_images,
                      void* stream);

/* ANTSNeighborhoodCorrelationImageToImageMetricv4::GetValueAndDerivative on
 * two images of grid g, both SyN roles at once (metric set up at
 * itkantsRegistrationHelper.hxx:931-948).  d_g_f2m / d_g_m2m: D x voxels.
 * host_value receives -mean(localCC). */
int ants_op_cc(const ants_image_desc* g, const float* d_fixed, const float* d_moving, int radius, float* d_g_f2m,
               float* d_g_m2m, double* host_value, void* stream);

/* ComposeDisplacementFieldsImageFilter: out = warp + disp(x + warp), 0 outside
 * (itkantsRegistrationHelper.hxx:3698-3711). */
int ants_op_compose(const ants_image_desc* g, const float* d_disp, const float* d_warp, float* d_out, void* stream);

/* Inverse-consistency residual of a forward / inverse field pair (the
 * DisplacementFieldTransform's displacement and inverse displacement
 * fields, itkantsRegistrationHelper.hxx:2206-2215): r(x) = fwd(x) +
 * inv(x + fwd(x)) with inv linearly interpolated and 0 outside its buffer,
 * |r| in voxel units over the voxels at least `margin` voxels from every
 * face; the sum is accumulated in FP64 on the device.  host_max, host_mean
 * (sum / count) and host_count may be NULL. */
int ants_op_inverse_consistency(const ants_image_desc* g, const float* d_fwd, const float* d_inv, int margin,
                                double* host_max, double* host_mean, int64_t* host_count, void* stream);

/* SyN ScaleUpdateField: u *= rate / max|u/spacing|; host_max receives the max. */
int ants_op_scale_update(const ants_image_desc* g, float* d_u, double rate, double* host_max, void* stream);

/* InvertDisplacementFieldImageFilter (20 steps, mean 0.001, max 0.1,
 * boundary enforced) as SyN configures it; d_init may be NULL (zero start). */
int ants_op_invert(const ants_image_desc* g, const float* d_field, const float* d_init, float* d_inv,
                   int* host_iters, double* host_max, double* host_mean, void* stream);

/* ants_op_invert without any device allocation.  No counterpart in ANTs:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The step
 * slots, per-block partials and candidate masks live in a fixed 8 MiB device
 * array that the library holds per GPU until the process ends, so grids of
 * up to about 33 million voxels are accepted (ANTS_ERR_INVALID beyond).
 * Same arguments, results and bits as ants_op_invert.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several streams or host
 * threads) race and must be serialised by the caller. */
int ants_op_invert_noalloc(const ants_image_desc* g, const float* d_field, const float* d_init, float* d_inv,
                           int* host_iters, double* host_max, double* host_mean, void* stream);

/* DisplacementFieldTransformParametersAdaptor: linear resample of a field onto
 * a new grid (itkantsRegistrationHelper.hxx:2033-2055). */
int ants_op_resample_field(const ants_image_desc* g_in, const float* d_in, const ants_image_desc* g_out,
                           float* d_out, void* stream);

/* PreprocessImage (itkantsRegistrationHelper.hxx:74-136) with the CLI
 * defaults: window [min, max + range/2560] -> [0,1]. */
int ants_op_preprocess(int64_t n, const float* d_in, float* d_out, void* stream);

/* antsApplyTransforms -n NearestNeighbor of an integer label map through
 * x -> x + u(x) (Examples/antsApplyTransforms.cxx:593-630): output grid g_out,
 * labels on g_lab, 0 outside. */
int ants_apply_transform_labels_nn(const ants_image_desc* g_out, const float* d_field, const ants_image_desc* g_lab,
                                   const int32_t* d_labels, int32_t* d_out, void* stream);

/* ---- antsApplyTransforms subset (Examples/antsApplyTransforms.cxx:546-640) ----
 * Transforms of a composite (itkantsRegistrationHelper.h:1262-1455):
 *   IDENTITY
 *   AFFINE        MatrixOffsetTransformBase: y = A (x - c) + t + c
 *   DISPLACEMENT  DisplacementFieldTransform: y = x + u(x), u linearly
 *                 interpolated in the field's grid, 0 outside its buffer */
#define ANTS_TRANSFORM_IDENTITY 0
#define ANTS_TRANSFORM_AFFINE 1
#define ANTS_TRANSFORM_DISPLACEMENT 2
#define ANTS_MAX_TRANSFORMS 8
typedef struct ants_transform {
  int kind;
  double matrix[9];       /* A, row-major dim x dim */
  double translation[3];  /* t */
  double center[3];       /* c (ITK fixed parameters) */
  ants_image_desc field;  /* DISPLACEMENT: geometry of the field */
  const float* d_field;   /* DISPLACEMENT: device buffer, dim interleaved physical components */
} ants_transform;

/* Interpolators (-n, make_interpolator_snip.tmpl): Linear, NearestNeighbor
 * (half-integers round up), GenericLabel[Linear] (the label whose indicator
 * image interpolates highest; ties go to the smaller label), BSpline[order]
 * (Examples/make_interpolator_snip.tmpl:30-40: BSplineInterpolateImageFunction
 * of order 0..5 over the input's BSplineDecompositionImageFilter coefficients,
 * mirror boundaries; ANTS_INTERP_BSPLINE is the default order 3). */
#define ANTS_INTERP_LINEAR 0
#define ANTS_INTERP_NEAREST 1
#define ANTS_INTERP_GENERIC_LABEL 2
#define ANTS_INTERP_BSPLINE_ORDER(q) (16 + (q))
#define ANTS_INTERP_BSPLINE ANTS_INTERP_BSPLINE_ORDER(3)
/* WindowedSincInterpolateImageFunction of radius 3 with a constant (zero)
 * boundary, make_interpolator_snip.tmpl:77-110: CosineWindowedSinc,
 * HammingWindowedSinc, LanczosWindowedSinc, BlackmanWindowedSinc,
 * WelchWindowedSinc. */
#define ANTS_INTERP_COSINE_WINDOWED_SINC 32
#define ANTS_INTERP_HAMMING_WINDOWED_SINC 33
#define ANTS_INTERP_LANCZOS_WINDOWED_SINC 34
#define ANTS_INTERP_BLACKMAN_WINDOWED_SINC 35
#define ANTS_INTERP_WELCH_WINDOWED_SINC 36

/* Resample d_in (grid `in`) onto grid `ref` (ResampleImageFilter): each
 * reference point p is mapped through transforms[0], then transforms[1], ...
 * -- the order of the -t options on the command line (the parser stores them
 * reversed and the CompositeTransform applies the last added first) -- and
 * the input is interpolated there, or default_value where the point falls
 * outside the input buffer.  Coordinates are formed in double.  Directions
 * may be any orthonormal matrices.  Asynchronous on `stream`; the transform
 * array is read before the call returns.  B-spline interpolation
 * holds the input's coefficients (8 bytes a voxel) in a stream-ordered
 * allocation for the call's duration. */
int ants_apply_transforms(const ants_image_desc* ref, const ants_image_desc* in, const float* d_in,
                          const ants_transform* transforms, int n_transforms, int interpolation,
                          double default_value, float* d_out, void* stream);

/* CreateJacobianDeterminantImage dim field out [doLogJacobian]
 * (Examples/CreateJacobianDeterminantImage.cxx:60-160 with
 * Utilities/itkDeformationFieldGradientTensorImageFilter.hxx and
 * Utilities/itkDeterminantTensorImageFilter.hxx): per voxel of the
 * displacement field d_field (grid `grid`, dim interleaved physical
 * components) the gradient by the five-point central difference with the
 * neighbours clamped to the buffer, divided by the spacing, each row rotated
 * by the grid's direction, plus the identity; d_out = max(det, 0), or
 * log(max(det, 0.001)) when do_log.  d_gradient (nullable; dim*dim floats
 * a voxel, row-major) receives that matrix: the tool's deformationGradient
 * output.  geometric != 0 is useGeometric
 * (Utilities/itkGeometricJacobianDeterminantImageFilter.hxx:88-335): the
 * |volume| of each voxel's triangle / tetrahedron moved by the linearly
 * interpolated field over its unmoved volume (no gradient output then).
 * Arithmetic in double, outputs float.  Asynchronous on `stream`. */
int ants_jacobian_determinant(const ants_image_desc* grid, const float* d_field, int do_log, int geometric,
                              float* d_out, float* d_gradient, void* stream);

/* antsApplyTransformsToPoints (Examples/antsApplyTransformsToPoints.cxx:145-190):
 * each of n_points physical points (dim doubles at d_in + i * dim) through
 * the transform list in -t order, CompositeTransform::TransformPoint, into
 * d_out (may equal d_in).  The list is the one ants_apply_transforms uses
 * for a reference image, so it maps reference-space points to input-space
 * points.  Asynchronous on `stream`. */
int ants_apply_transforms_to_points(int dim, const double* d_in, int64_t n_points,
                                    const ants_transform* transforms, int n_transforms, double* d_out,
                                    void* stream);

/* antsApplyTransforms -o [field,1] / -o DisplacementField[field]
 * (Examples/antsApplyTransforms.cxx:706-760, TransformToDisplacementFieldFilter):
 * d_out[i] = T(p_i) - p_i for every point p_i of grid `ref`, T the transform
 * list applied in order as in ants_apply_transforms; dim interleaved
 * physical components per voxel (the displacement-field layout).
 * Asynchronous on `stream`. */
int ants_compose_transforms_to_field(const ants_image_desc* ref, const ants_transform* transforms,
                                     int n_transforms, float* d_out, void* stream);

/* Replaces SyNImageRegistrationMethod::SetMovingInitialTransform and
 * SetFixedInitialTransform (Examples/itkantsRegistrationHelper.hxx:2119-2137,
 * 3583-3633): the composite of -r / -q initial transforms and the preceding
 * linear stages.  Each array lists the transforms in the order a point passes
 * through them (the -t order of ants_apply_transforms).  Every iteration then
 * samples the moving image at moving(T_m(x + u_m2m_inv(x))) and the fixed image
 * at fixed(T_f(x + u_f2m_inv(x))), T_m and T_f the two lists applied in order;
 * the output fields stay the SyN displacement fields
 * alone, as the composite keeps the initial transforms separately.  Identity,
 * affine and displacement-field transforms (an earlier deformable stage or a
 * restored composite, hxx:2119-2137); a list with a field holds at most 7
 * transforms, and its field buffers must stay valid while the handle uses
 * them.  Affine-only lists are folded into one map per side; a list with a
 * field is applied transform by transform.  Held by the handle for later
 * ants_syn_register* calls; n_moving = n_fixed = 0 clears them. */
int ants_syn_set_initial_transforms(ants_syn_handle* h, const ants_transform* moving, int n_moving,
                                    const ants_transform* fixed, int n_fixed);

/* ---- batched and served registration ----
 * No counterpart inside ANTs: the reference runs independent registrations
 * as separate processes or cluster jobs
 * (Scripts/antsMultivariateTemplateConstruction2.sh:1163-1280).  Here a batch
 * keeps `slots` registrations in flight on one GPU, each on its own stream
 * with a reusable engine, admits queued requests in order as slots and the
 * device-memory budget free up, and keeps fixed templates resident. */
typedef struct ants_batch_config {
  int slots;              /* concurrent registrations (streams); 1..64 */
  int64_t memory_budget;  /* device bytes admitted requests may hold; 0 = 90% of free device memory */
} ants_batch_config;

typedef struct ants_syn_request {
  ants_syn_params params;
  ants_image_desc fixed;               /* ignored when use_template */
  const float* host_fixed;
  ants_image_desc moving;
  const float* host_moving;
  const int32_t* host_moving_labels;   /* optional: label map on the moving grid, warped NN into the fixed grid */
  int use_template;                    /* 1: the fixed image is resident template `template_index` */
  int template_index;
  float* host_out_fwd;                 /* optional outputs on the last level grid */
  float* host_out_inv;
  int32_t* host_out_labels;            /* optional, last level grid */
  void* user;
  /* optional initial transforms, as ants_syn_set_initial_transforms takes
   * them; the arrays are read by ants_batch_submit, displacement-field
   * buffers must stay valid until the request's result is returned */
  const ants_transform* moving_initial;
  int n_moving_initial;
  const ants_transform* fixed_initial;
  int n_fixed_initial;
} ants_syn_request;

typedef struct ants_request_result {
  int64_t id;
  int status;                          /* ANTS_OK or an error code */
  int slot;
  double t_submit, t_start, t_end;     /* seconds on ants_batch_clock() */
  double voxel_iterations;
  double algorithmic_bytes;
  int64_t footprint_bytes;
  void* user;
  char error[256];
} ants_request_result;

typedef struct ants_batch ants_batch;
int ants_batch_create(const ants_batch_config* cfg, ants_batch** out);
void ants_batch_destroy(ants_batch* b);
/* Upload a fixed template once (index 0..7); requests may then use it.
 * Setting an index again publishes a new copy: requests submitted before
 * keep using the copy that was current at their submission (it is freed when
 * the last of them finishes), later ones use the new one.  Templates count
 * against the memory budget (ANTS_ERR_OOM when one does not fit). */
int ants_batch_set_template(ants_batch* b, int index, const ants_image_desc* g, const float* host_image);
/* Queue a request (host buffers must stay valid until its result is
 * returned).  ANTS_ERR_OOM when its footprint plus the resident templates
 * exceeds the memory budget. */
int ants_batch_submit(ants_batch* b, const ants_syn_request* r, int64_t* id);
/* Completed results: poll returns at once; wait blocks until at least one is
 * available or nothing is pending. */
int ants_batch_poll(ants_batch* b, ants_request_result* out, int max_results, int* n_out);
int ants_batch_wait(ants_batch* b, ants_request_result* out, int max_results, int* n_out);
int64_t ants_batch_pending(ants_batch* b);
/* Cancel request `id`: a queued request is dropped and returned with status
 * ANTS_ERR_CANCELLED; a running one stops at the end of its current
 * iteration and is returned with that status (the other requests are not
 * affected).  ANTS_ERR_INVALID when the request is neither queued nor running
 * (unknown, finished or already returned).  The clean failure return of a
 * registration: Examples/itkantsRegistrationHelper.hxx:2200-2204. */
int ants_batch_cancel(ants_batch* b, int64_t id);
/* Device bytes the batch holds or has reserved: resident templates, every
 * slot's retained workspace and staging buffers, and the footprint of each
 * running request.  Admission keeps this within the memory budget; an idle
 * slot frees its retained buffers when they keep the queue head out. */
int64_t ants_batch_device_bytes(ants_batch* b);
double ants_batch_clock(void);
/* Device bytes one registration of this geometry and schedule holds. */
int64_t ants_syn_footprint(const ants_syn_params* p, const ants_image_desc* fixed, const ants_image_desc* moving);

/* ---- linear stages (-t Translation[step] / Rigid[step] / Affine[step]) ----
 * The stages antsRegistrationSyN.sh runs before SyN
 * (Scripts/antsRegistrationSyN.sh:487-497), as RegistrationHelper builds them
 * (Examples/itkantsRegistrationHelper.hxx:1290-1372 optimizer, scales and
 * sampling; 1464-1732 the transforms; itkantsRegistrationHelper.h:905-1097
 * PrepareRegistrationMethod / AddLinearTransformToCompositeTransform):
 * ImageRegistrationMethodv4 with a ConjugateGradientLineSearchOptimizerv4
 * (line search over [0, 2] x the learning rate, epsilon 0.2), the learning
 * rate re-estimated at each level's first iteration so that the step moves
 * no corner of the virtual domain by more than `gradient_step`, parameter
 * scales from physical shifts, and the window convergence monitor.
 * Transforms: Translation (TranslationTransform), Rigid (Euler2DTransform /
 * Euler3DTransform, itkantsRegistrationHelper.h:1142-1180), Affine
 * (AffineTransform); a stage transform maps fixed points before the moving
 * initial transforms (y = M(S(x))). */
#define ANTS_LINEAR_TRANSLATION 0
#define ANTS_LINEAR_RIGID 1
#define ANTS_LINEAR_AFFINE 2
/* Metrics (hxx:926-1010): MeanSquares; GC = CorrelationImageToImageMetricv4
 * (minus the squared Pearson correlation); Mattes = Mattes mutual information
 * (zero-order fixed / cubic B-spline moving Parzen windows, 2 padding bins). */
#define ANTS_METRIC_MEAN_SQUARES 0
#define ANTS_METRIC_GC 1
#define ANTS_METRIC_MATTES 2
/* Metric sampling ([...,Regular|Random|None,percentage]): REGULAR takes every
 * ceil(1/percentage)-th virtual voxel in raster order, RANDOM
 * floor(N x percentage) voxels drawn with replacement; both move each point by
 * a normal variate of spacing/3 per axis (MT19937 seeded with `seed` at each
 * level, Box-Muller normals). */
#define ANTS_SAMPLING_NONE 0
#define ANTS_SAMPLING_REGULAR 1
#define ANTS_SAMPLING_RANDOM 2
typedef struct ants_linear_params {
  int dim;
  int transform;                        /* ANTS_LINEAR_* */
  double gradient_step;                 /* Rigid[step] */
  int metric;                           /* ANTS_METRIC_* */
  int bins;                             /* Mattes[f,m,w,bins,...] */
  int sampling;                         /* ANTS_SAMPLING_* */
  double sampling_percentage;           /* (0, 1] */
  uint32_t seed;                        /* --random-seed */
  int n_levels;
  int shrink[ANTS_MAX_LEVELS];          /* -f (isotropic factor, per-axis rule of the helper) */
  double sigma[ANTS_MAX_LEVELS];        /* -s */
  int sigma_in_mm;
  int iterations[ANTS_MAX_LEVELS];      /* -c [..,threshold,window] */
  double convergence_threshold;
  int convergence_window;
  int preprocess;                       /* window both images to [0,1] (PreprocessImage) */
} ants_linear_params;

/* One optimizer iteration (the 2DIAGNOSTIC line of
 * Examples/antsRegistrationOptimizerCommandIterationUpdate.h:158-171): the
 * metric value at the iteration's start, the convergence value, and the
 * learning rate the line search chose. */
typedef struct ants_linear_record {
  int level;
  int iteration;
  double value;
  double convergence;
  double learning_rate;
} ants_linear_record;

typedef struct ants_linear_result {
  int n_params;
  double params[12];                    /* the stage transform's ITK parameters */
  ants_transform transform;             /* the same as an AFFINE (matrix, translation, centre) */
  int iterations[ANTS_MAX_LEVELS];
  int64_t metric_evaluations;
} ants_linear_result;

/* Replaces registrationMethod->Update() of a linear stage
 * (itkantsRegistrationHelper.h:1085) and the transform it adds to the
 * composite (:1094).  d_fixed / d_moving: full-resolution images in HBM.
 * moving_initial / fixed_initial: identity or affine transforms in the
 * order a point passes them (the composite of -r / -q and earlier stages).
 * center: the stage transform's centre (NULL: the origin, as a new ITK
 * transform has); initial_params: NULL for identity.  Synchronous. */
int ants_linear_register(const ants_linear_params* p, const ants_image_desc* fixed, const float* d_fixed,
                         const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                         int n_moving, const ants_transform* fixed_initial, int n_fixed, const double* center,
                         const double* initial_params, ants_linear_result* out, ants_linear_record* trace,
                         int64_t trace_capacity, int64_t* trace_count, void* stream);
/* One metric evaluation on a level (the ImageToImageMetricv4
 * GetValueAndDerivative the optimizer calls): value and, with derivative,
 * minus its derivative by the stage parameters (ITK's v4 convention). */
int ants_linear_metric(const ants_linear_params* p, int level, const ants_image_desc* fixed, const float* d_fixed,
                       const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                       int n_moving, const ants_transform* fixed_initial, int n_fixed, const double* center,
                       const double* params, int derivative, double* value, double* der, int64_t* n_valid,
                       void* stream);
/* Host only: a level's sample points (physical, x fastest), and the
 * parameter scales of RegistrationParameterScalesFromPhysicalShift. */
int ants_linear_sample_points(const ants_linear_params* p, int level, const ants_image_desc* fixed, double* out,
                              int64_t capacity, int64_t* count);
int ants_linear_estimate_scales(const ants_linear_params* p, int level, const ants_image_desc* fixed,
                                const ants_transform* moving_initial, int n_moving, const double* center,
                                const double* params, double* scales);
/* Sample points are kept on the device between stages: they depend only on
 * the seed, the sampling strategy and percentage and the level grid, so the
 * Affine stage that follows a Rigid stage with the same metric sampling
 * reuses the Rigid stage's points (ITK draws the same points for both from
 * the same seed).  At most ANTS_LINEAR_POINT_CACHE sets (default 8; 0 turns
 * the cache off) are held per process.  This call frees every set that no
 * running stage holds.  No counterpart in the reference: ITK regenerates
 * the points in each stage. */
int ants_linear_release_points(void);

/* ---- dense field-transform stages with the CC metric ----
 * -t GaussianDisplacementField[step, updateVariance, totalVariance]
 *    (Examples/itkantsRegistrationHelper.hxx:1733-1847: ImageRegistrationMethodv4
 *    with GaussianSmoothingOnUpdateDisplacementFieldTransform and the
 *    conjugate-gradient line-search optimizer of hxx:1290-1372)
 * -t Exponential[step, updateVariance, velocityVariance, integrationSteps]
 *    (hxx:3112-3277: GaussianExponentialDiffeomorphicTransform with the
 *    gradient-descent optimizer of hxx:1405-1418; integration steps 0 =
 *    automatic), parsed at Examples/antsRegistrationTemplateHeader.h:692-699, 871-886.
 * Variances are in voxel^2 (variance space, as the CLI gives them). */
#define ANTS_FIELD_GAUSSIAN_DISPLACEMENT 0
#define ANTS_FIELD_EXPONENTIAL 1
/* The B-spline forms (hxx:1848-1982, 3278-3461): the update and total /
 * velocity fields fitted by B-splines instead of Gaussian smoothing
 * (BSplineSmoothingOnUpdateDisplacementFieldTransform, BSplineExponential-
 * DiffeomorphicTransform), meshes doubled per level as mesh << (level + 1);
 * a zero mesh means no fit of that field.
 * -t BSplineDisplacementField[step, updateMesh, totalMesh=0, order=3]
 * -t BSplineExponential[step, updateMesh, velocityMesh=0, integrationSteps, order=3] */
#define ANTS_FIELD_BSPLINE_DISPLACEMENT 2
#define ANTS_FIELD_BSPLINE_EXPONENTIAL 3
typedef struct ants_field_params {
  int dim;
  int model;                            /* ANTS_FIELD_* */
  int n_levels;
  int shrink[ANTS_MAX_LEVELS];          /* -f */
  double sigma[ANTS_MAX_LEVELS];        /* -s */
  int sigma_in_mm;
  int iterations[ANTS_MAX_LEVELS];      /* -c [..,threshold,window] */
  double convergence_threshold;
  int convergence_window;
  double learning_rate;                 /* step */
  double update_variance;
  double total_variance;                /* GaussianDisplacementField: total field; Exponential: velocity field */
  int integration_steps;                /* Exponential: squarings of exp(v); 0 = automatic */
  int cc_radius;                        /* CC[f,m,1,radius] */
  int preprocess;                       /* window both images to [0,1] (PreprocessImage) */
  int force_iterations;                 /* 1: never stop early (parity harness) */
  int update_mesh[3];                   /* B-spline forms: spans per axis at the base level */
  int total_mesh[3];                    /* B-spline forms: total / velocity field (0: no fit) */
  int spline_order;                     /* B-spline forms: 1..3 */
} ants_field_params;

/* Replaces registrationMethod->Update() of such a stage and the transform it
 * adds to the composite (hxx:1838-1846, 3268-3276).  d_fixed / d_moving:
 * full-resolution images in HBM; moving_initial / fixed_initial: identity
 * or affine transforms in the order a point passes them.  Outputs on the
 * last level's grid (ants_syn_level_geometry of the last level), D floats
 * per voxel, any may be NULL: d_out_field the displacement field (for
 * Exponential exp(v)), d_out_inverse exp(-v) (Exponential only),
 * d_out_velocity v (Exponential only).  trace: the 2DIAGNOSTIC records
 * (value, convergence, learning rate per iteration); iterations_out: per
 * level.  Synchronous. */
int ants_field_register(const ants_field_params* p, const ants_image_desc* fixed, const float* d_fixed,
                        const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                        int n_moving, const ants_transform* fixed_initial, int n_fixed, float* d_out_field,
                        float* d_out_inverse, float* d_out_velocity, ants_linear_record* trace,
                        int64_t trace_capacity, int64_t* trace_count, int* iterations_out, void* stream);
/* One metric evaluation of such a stage on a level for a given displacement
 * (level grid, D floats per voxel; NULL = zero): the CC value and, with
 * d_derivative, the derivative field (c1 * A^T grad M at each mapped
 * point).  No counterpart call in ANTs: the metric's GetValueAndDerivative
 * as the optimizer sees it, for tests. */
int ants_field_metric(const ants_field_params* p, int level, const ants_image_desc* fixed, const float* d_fixed,
                      const ants_image_desc* moving, const float* d_moving, const ants_transform* moving_initial,
                      int n_moving, const ants_transform* fixed_initial, int n_fixed, const float* d_field,
                      double* value, float* d_derivative, void* stream);
/* exp(v) (sign +1) or exp(-v) (sign -1) of a velocity field on grid g by
 * ExponentialDisplacementFieldImageFilter's scaling and squaring (steps 0 =
 * automatic); *steps_used receives the squaring count. */
int ants_op_exponential(const ants_image_desc* g, const float* d_v, float* d_out, int steps, int sign,
                        int* steps_used, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ANTS_SYN_H */
