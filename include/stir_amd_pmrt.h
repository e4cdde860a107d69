/*
 * stir_amd_pmrt -- C ABI of the MI355X ray-tracing-matrix projector pair
 * and OSEM update.
 *
 * This is the boundary a STIR ForwardProjectorByBin / BackProjectorByBin
 * plugin (and a GeneralisedObjectiveFunction plugin for the fused OSEM step)
 * binds. All entry points take plain pointers and sizes. Arrays use STIR's
 * contiguous layouts:
 *   image           [z][y][x]  float32 (VoxelsOnCartesianGrid, x fastest)
 *   projection data ProjDataInMemory order: segment sequence 0,+1,-1,+2,-2,...,
 *                   then axial position, view, tangential position (fastest)
 *                   (ProjDataInMemory.cxx:190-235, ProjData.cxx:641-661).
 * Every call returns 0 on success and a negative MI_ERR_* code otherwise;
 * mi_last_error() returns the message of the last failure on the calling
 * thread. No C++ exception crosses this boundary.
 *
 * Threading: a context is used by one host thread at a time. The entry
 * points with host buffers use the context's own stream and return when
 * their results are on the host; the *_dev entry points run on the stream
 * the caller passes (NULL = the HIP null stream) and share the context's
 * device scratch buffers, so their calls must be ordered (one stream, or
 * synchronised).
 */
#ifndef STIR_AMD_PMRT_H
#define STIR_AMD_PMRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_MAX_SEGMENTS 256
#define MI_MAX_TOF_BINS 64

enum {
  MI_OK = 0,
  MI_ERR_ARGUMENT = -1,  /* bad argument or geometry the projector does not support */
  MI_ERR_STATE = -2,     /* call order violated (e.g. projection before mi_set_up) */
  MI_ERR_DEVICE = -3,    /* HIP runtime error */
  MI_ERR_MEMORY = -4,    /* device or host allocation failed */
  MI_ERR_INTERNAL = -5
};

/* Geometry of one problem: ProjDataInfoCylindrical{NoArcCorr,ArcCorr} +
 * VoxelsOnCartesianGrid + "Ray Tracing Matrix Parameters"
 * (ProjMatrixByBinUsingRayTracing.cxx:76-104). Segment arrays are indexed by
 * segment_num - min_segment_num. Image index ranges: z in [min_z, min_z+nz),
 * y in [min_y, min_y+ny), x in [min_x, min_x+nx). */
typedef struct {
  int32_t num_rings;
  int32_t num_detectors_per_ring;
  float effective_ring_radius; /* mm, Scanner::get_effective_ring_radius */
  float ring_spacing;          /* mm */
  float intrinsic_tilt;        /* rad */
  int32_t arc_corrected;
  float bin_size; /* mm, tangential sampling of arc-corrected data */
  int32_t num_views;
  int32_t num_tangential_poss;
  int32_t min_tangential_pos_num;
  int32_t min_segment_num;
  int32_t max_segment_num;
  int32_t min_ring_diff[MI_MAX_SEGMENTS];
  int32_t max_ring_diff[MI_MAX_SEGMENTS];
  int32_t num_axial_poss[MI_MAX_SEGMENTS];
  int32_t nz, ny, nx;
  int32_t min_z, min_y, min_x;
  float vox_z, vox_y, vox_x;
  float origin_z, origin_y, origin_x;
  int32_t restrict_to_cylindrical_FOV;
  int32_t num_tangential_LORs;
  int32_t do_symmetry_90degrees_min_phi;
  int32_t do_symmetry_180degrees_min_phi;
  int32_t do_symmetry_swap_segment;
  int32_t do_symmetry_swap_s;
  int32_t do_symmetry_shift_z;
  /* Time of flight (ProjDataInfo::set_tof_mash_factor, ProjDataInfo.cxx:174-258;
   * ProjMatrixByBin::enable_tof, ProjMatrixByBin.cxx:66-75). num_tof_bins <= 1
   * means non-TOF data. TOF bin k (0-based: timing_pos_num - min_tof_pos_num)
   * covers [tof_low[k], tof_high[k]] mm along the LOR; tof_r_sqrt2_gauss_sigma
   * = 1 / (sigma * sqrt(2)), sigma in mm. TOF projection data are
   * num_tof_bins non-TOF blocks, TOF bin outermost. */
  int32_t num_tof_bins;
  float tof_r_sqrt2_gauss_sigma;
  float tof_low[MI_MAX_TOF_BINS];
  float tof_high[MI_MAX_TOF_BINS];
  /* "use actual detector boundaries" (ProjMatrixByBinUsingRayTracing.cxx:82,
   * 300-387,608-634,760-762): phi and s of each bin's LOR from its two
   * detectors, and twice the tangential spacing of multiple LORs. As in the
   * reference it applies only to non-arc-corrected data without view mashing
   * whose segments each hold one ring difference (else it is ignored). */
  int32_t use_actual_detector_boundaries;
} mi_geometry;

/* Sizes of a set-up context. */
typedef struct {
  int64_t num_bins;
  int64_t num_voxels;
  int64_t num_basic_rows;
  int64_t num_basic_elements;   /* Siddon elements stored for the basic rows */
  int64_t num_families;         /* (segment, view, tangential) groups sharing a z-shifted row */
  int64_t num_matrix_elements;  /* sum of row lengths over all bins, before z clipping */
  int64_t num_table_entries;    /* entries of the voxel-driven back-projection table */
  int64_t num_nonzeros;         /* matrix elements whose voxel lies inside the image (nnz after z clipping) */
  int64_t device_bytes;         /* device memory held by the context */
  int32_t num_subsets;
  int32_t symmetry_flags[5];    /* effective 90,180,swap_segment,swap_s,shift_z */
  int32_t nz_padded;
  int32_t planes_per_axial_position; /* shared by all segments (1 or 2), 0 when it varies */
  /* Bytes one projection of all subsets must move through HBM under this
   * design when every re-read is served by a cache: the matrix data it
   * streams (forward: the rows per task, task and family records;
   * back: the back-projection table and its offsets) plus one read of the
   * input and one write of the output. */
  int64_t fp_design_bytes;
  int64_t bp_design_bytes;
  /* back projection: 1 through LDS slabs of image tiles, 0 by image column
   * over the column table; the slab tile and its segment groups per view */
  int32_t slab_back_projection;
  int32_t slab_tile_x, slab_tile_y, slab_segment_groups;
  int32_t slab_lds_bytes; /* LDS of one slab workgroup */
  /* TOF: window widths of the stored forward and back records (TOF bins per
   * element; 0 when the weights are computed on the fly) and the widest
   * range of TOF bins with a non-zero weight of one element */
  int32_t tof_fp_window;
  int32_t tof_bp_window;
  int32_t tof_live_span;
  /* 1 when back projections run over 2 x 2 column blocks (MI_TUNE_BP_QUAD),
   * and the number of merged entries of that table */
  int32_t bp_quad;
  int64_t num_quad_entries;
} mi_info;

typedef struct mi_ctx mi_ctx;

/* ---- context ------------------------------------------------------------ */
/* Creates a context on HIP device `device`. */
int mi_ctx_create(int device, mi_ctx** out);
void mi_ctx_destroy(mi_ctx* ctx);
/* Message of the last failed call on this thread ("" if none). */
const char* mi_last_error(void);

/* ---- set-up ------------------------------------------------------------- */
/* Replaces ProjMatrixByBinUsingRayTracing::set_up
 * (ProjMatrixByBinUsingRayTracing.cxx:239-407) together with
 * ForwardProjectorByBin::set_up / BackProjectorByBin::set_up
 * (ForwardProjectorByBin.h:70-73, BackProjectorByBin.h:68-71): builds the
 * symmetry tables, traces the basic rows on the device (the reference's
 * basic-bin cache, ProjMatrixByBin.inl:48-113) and the voxel-driven
 * back-projection table for `num_subsets` OSEM subsets
 * (find_basic_vs_nums_in_subset.cxx:31-69). May be called again. */
int mi_set_up(mi_ctx* ctx, const mi_geometry* geometry, int num_subsets);
int mi_get_info(mi_ctx* ctx, mi_info* out);
/* Effective symmetries and subset balance, as
 * ForwardProjectorByBin::get_symmetries_used (ForwardProjectorByBin.h:80) and
 * subsets_are_approximately_balanced (...AndProjData.cxx:490-529). */
int mi_subsets_balanced(mi_ctx* ctx, int* balanced);
/* (view, segment) pairs of a subset in processing order. *count is set to the
 * number of pairs; at most cap are written. */
int mi_subset_view_segments(mi_ctx* ctx, int subset, int32_t* views, int32_t* segments, int cap, int* count);

/* Basic bins and symmetry operations of n bins, evaluated on the host plan
 * without a device: DataSymmetriesForBins_PET_CartesianGrid::
 * find_symmetry_operation_from_basic_bin (.inl:664-674). bins, basic: n x
 * (segment, view, axial, tangential); ops: n x (in-plane code, z sign, z add)
 * with z' = zsign*z + zadd and the in-plane codes 0 (x,y), 1 (-x,y), 2 (x,-y),
 * 3 (-x,-y), 4 (y,x), 5 (-y,x), 6 (y,-x), 7 (-y,-x). */
int mi_symmetry_operations(const mi_geometry* geometry, int n, const int32_t* bins, int32_t* basic, int32_t* ops);

/* Matrix row of one bin as the device uses it (basic row + symmetry
 * operation), ProjMatrixByBin::get_proj_matrix_elems_for_one_bin
 * (ProjMatrixByBin.inl:48-113). *n receives the length; at most cap written. */
int mi_get_row(mi_ctx* ctx, int segment, int view, int axial_pos, int tang_pos,
               int16_t* z, int16_t* y, int16_t* x, float* w, int cap, int* n);

/* ---- projections, host buffers (STIR layouts) ---------------------------- */
/* ForwardProjectorByBin::forward_project(ProjData&, image, subset_num,
 * num_subsets, zero=true) (ForwardProjectorByBin.cxx:97-103,168-230).
 * subset < 0: all bins; otherwise the bins of that subset are projected and
 * all other bins are set to zero. */
int mi_forward_project(mi_ctx* ctx, const float* image, float* proj_data, int subset);
/* BackProjectorByBin::back_project(image, ProjData)
 * (BackProjectorByBin.cxx:111-117,183-229,316-346); overwrites image. */
int mi_back_project(mi_ctx* ctx, const float* proj_data, float* image, int subset);

/* ---- projections, device buffers (STIR layouts, device pointers) --------- */
/* Same operations on device-resident arrays, issued on `stream` (a
 * hipStream_t; NULL is the HIP null stream, as for any HIP library call).
 * They are stream-ordered with the caller's other work on that stream and
 * return without waiting. Used by the PyTorch layer. */
int mi_forward_project_dev(mi_ctx* ctx, const float* d_image, float* d_proj_data, int subset, void* stream);
int mi_back_project_dev(mi_ctx* ctx, const float* d_proj_data, float* d_image, int subset, void* stream);

/* ---- projections of basic view-segments (the drop-in projector plugins) ---- */
/* detail::find_basic_vs_nums_in_subset (find_basic_vs_nums_in_subset.cxx:31-69)
 * for any subset / num_subsets with the context's symmetries: the basic
 * (view, segment) pairs, segment-major, views subset, subset + num_subsets,
 * ... *count is set to their number; at most cap are written. */
int mi_basic_view_segments_in_subset(mi_ctx* ctx, int subset, int num_subsets, int32_t* views, int32_t* segments,
                                     int cap, int* count);
/* ForwardProjectorByBin::set_input (ForwardProjectorByBin.cxx:306-318): the
 * image (host, STIR layout) that mi_forward_project_view_segments projects;
 * kept on the device until the next call. */
int mi_set_image(mi_ctx* ctx, const float* image);
/* ForwardProjectorByBin::forward_project of the RelatedViewgrams of n
 * view-segments (each replaced by its basic view-segment; repeats count
 * once), as forward_project(ProjData&, subset_num, num_subsets) does for the
 * basic view-segments of a subset (ForwardProjectorByBin.cxx:168-230): the
 * image of mi_set_image is projected into exactly those viewgrams (all TOF
 * bins) of proj_data (host, ProjDataInMemory order); no other bin of
 * proj_data is written. The work is that of those view-segments only. */
int mi_forward_project_view_segments(mi_ctx* ctx, int n, const int32_t* views, const int32_t* segments,
                                     float* proj_data);
/* BackProjectorByBin::back_project of the RelatedViewgrams of n view-segments
 * followed by get_output (BackProjectorByBin.cxx:183-229,316-346): only
 * those viewgrams of proj_data (host, ProjDataInMemory order) are read and
 * back-projected; image (host, STIR layout) is overwritten. When the
 * view-segments are exactly one subset of the context (mi_set_up's
 * num_subsets) only that subset's part of the back-projection table runs;
 * any other set runs the whole table over data zeroed outside the set. */
int mi_back_project_view_segments(mi_ctx* ctx, int n, const int32_t* views, const int32_t* segments,
                                  const float* proj_data, float* image);

/* ---- objective function / OSEM, data resident on the device -------------- */
/* Measured data and additive term of
 * PoissonLogLikelihoodWithLinearModelForMeanAndProjData (set_up_before_sensitivity,
 * ...AndProjData.cxx:580-686). additive may be NULL (then additive_constant is
 * used for every bin). zero_end_planes: "zero end planes of segment 0"
 * (distributable.cxx:145-225). */
int mi_obj_set_data(mi_ctx* ctx, const float* measured, const float* additive, float additive_constant,
                    int zero_end_planes);
/* Subset sensitivities BP_subset(1) (compute_sensitivities,
 * PoissonLogLikelihoodWithLinearModelForMean.cxx:374-455, use subset sensitivities). */
int mi_obj_compute_sensitivities(mi_ctx* ctx);
/* "use time-of-flight sensitivities" (...AndProjData.cxx:143, default 0): with
 * TOF data and 0 the sensitivities are non-TOF back projections
 * (:533-536,648-670), with 1 TOF back projections of TOF ones. Takes effect at
 * the next mi_obj_compute_sensitivities. */
int mi_obj_set_use_tof_sensitivities(mi_ctx* ctx, int use);
/* Multiplicative factors m of the model ybar = m (A x + b): the
 * normalisation's undo() applied to ones ("Bin Normalisation type",
 * ...AndProjData.cxx:158,551-565; get_viewgrams, distributable.cxx:197-223).
 * Host array, ProjDataInMemory order, all TOF bins; NULL removes them. They
 * enter the sensitivities (BP of m), the gradient (BP(y/(A x + b) - m),
 * RPC at ...AndProjData.cxx:1450-1465) and the log-likelihood
 * ((A x + b) m, :1484-1515); "zero end planes of segment 0" zeroes their end
 * sinograms. With TOF data and non-TOF sensitivities every TOF bin must carry
 * the same factors (:636-646). Sensitivities must be recomputed. */
int mi_obj_set_multiplicative(mi_ctx* ctx, const float* mult);
/* List-mode data of PoissonLogLikelihoodWithLinearModelForMeanAndListModeData-
 * WithProjMatrixByBin: prompt events as num_events x 5 int32 (segment, view,
 * axial, tangential position, timing position; timing 0 without TOF); events
 * outside the projection data are dropped (...ListModeDataWithProjMatrixByBin.cxx:
 * 495-511). additive: host projection data (ProjDataInMemory order) or NULL
 * for additive_constant; with TOF data every event gets the additive term of
 * the last timing position, as the reference's event cache (:570-596).
 * Afterwards mi_obj_compute_sensitivities (BP of the multiplicative factors,
 * no end-plane zeroing), mi_osem_*, mi_obj_subset_gradient (sum over the
 * subset's events of row^T / fwd, minus the subset sensitivity unless
 * add_sensitivity) and mi_obj_log_likelihood (the value the reference's
 * actual_compute_objective_function_without_penalty returns: -sum over events
 * of log fwd, log(1e-4) for events with 1 > 1e4 fwd, :756-787,876-911) work
 * on the events, fwd = row . x + additive. Every event contributes row^T / fwd
 * to the gradient: the reference's gradient instantiation
 * LM_gradient_and_value<true, false> applies the 1e4 cap only to the value
 * (:822-843). mi_obj_set_data switches back to projection data. */
int mi_lm_set_events(mi_ctx* ctx, const int32_t* events, int64_t num_events, const float* additive,
                     float additive_constant);
/* "use subset sensitivities" (default 1, ...ForMean.cxx:48). 0: every subset
 * uses the total sensitivity (accumulated subset by subset) divided by the
 * number of subsets (:374-455); subsets must be balanced. Takes effect at the
 * next mi_obj_compute_sensitivities. */
int mi_obj_set_use_subset_sensitivities(mi_ctx* ctx, int use);
/* Subset sensitivity (host image, STIR layout); subset -1: the total
 * sensitivity, accumulated subset by subset (get_sensitivity,
 * ...ForMean.cxx:147-153,419-455). */
int mi_obj_get_sensitivity(mi_ctx* ctx, int subset, float* image);
/* Current estimate (host, STIR layout). mi_obj_set_estimate applies
 * threshold_min_to_small_positive_value(x, 1e-6) when enforce_positivity
 * (OSMAPOSLReconstruction.cxx:291-292). */
int mi_obj_set_estimate(mi_ctx* ctx, const float* image, int enforce_positivity);
int mi_obj_get_estimate(mi_ctx* ctx, float* image);
/* One OSMAPOSL subiteration, prior-free (OSMAPOSLReconstruction::update_estimate,
 * OSMAPOSLReconstruction.cxx:373-516): x *= clamp(BP_s(y/(FP_s(x)+b))/sens_s).
 * subiteration_num is 1-based (thresholding is skipped for subiteration 1). */
int mi_osem_subset_step(mi_ctx* ctx, int subset, int subiteration_num);
/* OSMAPOSL keys (OSMAPOSLReconstruction.cxx:95-121): "MAP_model" (0 additive,
 * the default; 1 multiplicative) and "minimum/maximum relative change"
 * (defaults 0 and FLT_MAX) for mi_osem_subset_step / mi_osem_run. */
int mi_osem_set_options(mi_ctx* ctx, int map_model_multiplicative, float minimum_relative_change,
                        float maximum_relative_change);

/* Prior of OSMAPOSL's one-step-late update ("prior type" of
 * GeneralisedObjectiveFunction; QuadraticPrior.cxx, RelativeDifferencePrior.cxx). */
#define MI_PRIOR_NONE 0
#define MI_PRIOR_QUADRATIC 1
#define MI_PRIOR_RELATIVE_DIFFERENCE 2
typedef struct {
  int32_t type;               /* MI_PRIOR_* */
  float penalisation_factor;  /* beta */
  float gamma;                /* relative difference prior "gamma value" (reference default 2) */
  float epsilon;              /* relative difference prior "epsilon value" (default 0) */
  int32_t only_2D;            /* no neighbours in other planes */
  const float* weights;       /* 27 floats [dz+1][dy+1][dx+1], or NULL for vox_x / distance */
  const float* kappa;         /* host image (STIR layout) or NULL */
} mi_prior;
/* With a prior (type != NONE and beta != 0) mi_osem_subset_step / mi_osem_run
 * run the one-step-late update of OSMAPOSLReconstruction::update_estimate
 * (:403-516): u = BP_s(y/ybar); denominator from the prior's gradient g at
 * the current estimate -- additive: clamp(g/num_subsets + sens_s, sens_s/10,
 * 10 sens_s); multiplicative: sens_s clamp(1 + g, 0.1, 10) --; divide(u, d,
 * 1e-6); threshold; x *= u. NULL or type NONE removes the prior. */
int mi_obj_set_prior(mi_ctx* ctx, const mi_prior* prior);

/* num_subiterations subiterations starting at start_subiteration_num,
 * subset = (k + start_subset - 1) % num_subsets
 * (IterativeReconstruction.cxx:384-428,630-639). */
int mi_osem_run(mi_ctx* ctx, int start_subiteration_num, int num_subiterations, int start_subset);
/* compute_sub_gradient_without_penalty[_plus_sensitivity]
 * (...AndProjData.cxx:694-735, RPC at :1423-1481) at the current estimate;
 * subset < 0: all subsets. Output image in STIR layout (host). */
int mi_obj_subset_gradient(mi_ctx* ctx, int subset, int add_sensitivity, float* gradient);
/* actual_compute_objective_function_without_penalty
 * (...AndProjData.cxx:739-781, recon_array_functions.cxx:336-377) at the
 * current estimate, accumulated in FP64. subset < 0: all subsets. */
int mi_obj_log_likelihood(mi_ctx* ctx, int subset, double* value);

/* ---- differentiable pieces, device buffers (STIR layouts) ----------------- */
/* Calls on one context share its device scratch buffers: issue them on one
 * stream (or serialise them); calls on different streams that overlap race. */

/* Log-likelihood and its sub-gradient at a device image in one forward
 * projection (FP64 partial sums) plus one back projection:
 *   *d_value (device double, if not NULL) =
 *     actual_compute_objective_function_without_penalty (...AndProjData.cxx:739-781,
 *     recon_array_functions.cxx:336-377),
 *   d_gradient (device image, if not NULL) =
 *     compute_sub_gradient_without_penalty[_plus_sensitivity] (...AndProjData.cxx:694-735,
 *     RPC at :1423-1481).
 * Uses the data of mi_obj_set_data; subset < 0: all subsets. Nothing is
 * copied to the host. */
int mi_obj_value_and_gradient_dev(mi_ctx* ctx, const float* d_image, int subset, int add_sensitivity,
                                  float* d_gradient, double* d_value, void* stream);

/* QuadraticPrior::compute_value and compute_gradient (QuadraticPrior.cxx:229-367)
 * of a device image with the context's image geometry. weights: 27 host
 * floats [dz+1][dy+1][dx+1], or NULL for the default vox_x / distance
 * (compute_weights, :202-226); d_kappa: device image or NULL; only_2D: no
 * neighbours in other planes. *d_value (device double) and d_gradient
 * (device image) are each optional. */
int mi_quadratic_prior_dev(mi_ctx* ctx, const float* d_image, const float* d_kappa, const float* weights, float beta,
                           int only_2D, float* d_gradient, double* d_value, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mi_quadratic_prior_dev of a device image of nz x ny x nx
 * voxels (STIR layout, voxel sizes in mm for the default weights) with no
 * context and no allocation. The prior's FP64 partial sums, one per
 * workgroup of 64 x 4 voxels of one plane, live in a fixed array of
 * 65536 doubles (512 KB) that every GPU the library runs on holds until the
 * process ends: images of more than 65536 such workgroups (e.g. above
 * 256 x 256 x 1024) are rejected. Calls that overlap on one GPU (other
 * streams, other threads) race on that array; calls on one stream do not.
 * device becomes the current device. Results equal mi_quadratic_prior_dev's
 * on the same image. */
int mi_quadratic_prior_noalloc_dev(int device, int nz, int ny, int nx, float vox_z, float vox_y, float vox_x,
                                   const float* d_image, const float* d_kappa, const float* weights, float beta,
                                   int only_2D, float* d_gradient, double* d_value, void* stream);

/* RelativeDifferencePrior::compute_value and compute_gradient
 * (RelativeDifferencePrior.cxx:250-433) of a device image, arguments as
 * mi_quadratic_prior_dev plus gamma and epsilon. */
int mi_relative_difference_prior_dev(mi_ctx* ctx, const float* d_image, const float* d_kappa, const float* weights,
                                     float beta, float gamma, float epsilon, int only_2D, float* d_gradient,
                                     double* d_value, void* stream);

/* The data term's Hessian products on device images (STIR layout),
 * accumulated into d_output (d_output -= A^T q) for one subset, or all
 * subsets in one pass with subset = -1:
 *  approximate = 0: accumulate_sub_Hessian_times_input_without_penalty
 *    (PoissonLogLikelihoodWithLinearModelForMeanAndProjData.cxx:1066-1218),
 *    q = divide_and_truncate(y A d_input, (A d_current + b)^2);
 *  approximate = 1: add_multiplication_with_approximate_sub_Hessian_without_penalty
 *    (:957-1064), q = divide_and_truncate(A d_input, y / m / m) with m the
 *    multiplicative factors of mi_obj_set_multiplicative (at least 1e-20;
 *    none: y), d_current unused.
 * y and b are the data of mi_obj_set_data as given (end planes not zeroed);
 * divide_and_truncate works per viewgram (recon_array_functions.cxx:172-266).
 * A negative value in A d_input is an error (MI_ERR_ARGUMENT) that leaves
 * d_output unchanged, as the reference aborts; synchronizes the stream.
 * List-mode data (mi_lm_set_events): approximate = 0 only, the list-mode
 * objective's accumulate_sub_Hessian_times_input_without_penalty
 * (...ListModeDataWithProjMatrixByBin.cxx:790-822,963-1003): every event of
 * the subset with fwd = A_e d_current + b adds
 * A_e^T (-1 / fwd^2 * A_e d_input) to d_output, except events with
 * 1 > 1e4 fwd; no negativity check, no synchronization. */
int mi_obj_hessian_times_input_dev(mi_ctx* ctx, int approximate, const float* d_current, const float* d_input,
                                   int subset, float* d_output, void* stream);

/* Prior Hessian products on device images (STIR layout), accumulated into
 * d_output: accumulate_Hessian_times_input (QuadraticPrior.cxx:604-690,
 * RelativeDifferencePrior.cxx:535-614) of d_input at the estimate d_image,
 * or with approximate = 1 the quadratic prior's
 * add_multiplication_with_approximate_Hessian (:537-602; d_image unused).
 * type MI_PRIOR_QUADRATIC or MI_PRIOR_RELATIVE_DIFFERENCE; other arguments
 * as mi_relative_difference_prior_dev. */
int mi_prior_hessian_times_input_dev(mi_ctx* ctx, int type, int approximate, const float* d_image,
                                     const float* d_input, const float* d_kappa, const float* weights, float beta,
                                     float gamma, float epsilon, int only_2D, float* d_output, void* stream);

/* One OSMAPOSL subiteration on device images, out of place:
 * d_x_out = d_x_in * u, u = clamp(BP_s(y/(FP_s(x)+b)) / sens_s)
 * (OSMAPOSLReconstruction::update_estimate, OSMAPOSLReconstruction.cxx:373-516,
 * as mi_osem_subset_step); d_update (if not NULL) receives u. Needs
 * mi_obj_set_data and mi_obj_compute_sensitivities. d_x_out may equal d_x_in.
 * Prior-free (an error when a prior is set), with the relative-change bounds
 * of mi_osem_set_options. */
int mi_osem_subset_step_dev(mi_ctx* ctx, const float* d_x_in, float* d_x_out, float* d_update, int subset,
                            int subiteration_num, void* stream);
/* No counterpart in the reference (it has no derivatives of the OSEM map):
 * vector-Jacobian product of mi_osem_subset_step_dev with respect to x_in,
 *   d_grad_in = g * u + BP_s( r'(FP_s(x_in)+b) * FP_s(c * g * x_in / sens_s) ),
 * r' the derivative of the truncated quotient y/ybar, g = d_grad_out, u the
 * update written by the forward step, c the derivative of its clamp to the
 * context's relative-change bounds (mi_osem_set_options): 0 where
 * subiteration_num != 1 and u <= minimum or u >= maximum relative change,
 * else 1. Two forward and one back projection. */
int mi_osem_subset_step_vjp_dev(mi_ctx* ctx, const float* d_x_in, const float* d_update, const float* d_grad_out,
                                float* d_grad_in, int subset, int subiteration_num, void* stream);

/* Waits for all work of the context. */
int mi_synchronize(mi_ctx* ctx);

/* ---- batches of independent reconstructions ------------------------------- */
/* No counterpart in the reference (it reconstructs one problem at a time).
 * A batch holds n independent non-TOF problems, each a full context with its
 * own geometry (mi_batch_context gives it, e.g. for mi_osem_run on one
 * problem), and runs an OSMAPOSL subiteration of all of them with one launch
 * per lane-group size: forward projection by families, back projection by
 * image columns, packed several to a 64-lane wave when they are short
 * (at most 64 axial positions per segment and 64 planes). Results equal the
 * per-problem path's. All problems use num_subsets subsets. */
typedef struct mi_batch mi_batch;
int mi_batch_create(int device, int n, const mi_geometry* geometries, int num_subsets, mi_batch** out);
void mi_batch_destroy(mi_batch* batch);
int mi_batch_context(mi_batch* batch, int problem, mi_ctx** ctx);
/* mi_obj_set_data + mi_obj_compute_sensitivities of one problem. */
int mi_batch_set_data(mi_batch* batch, int problem, const float* measured, const float* additive,
                      float additive_constant, int zero_end_planes);
int mi_batch_set_estimate(mi_batch* batch, int problem, const float* image, int enforce_positivity);
int mi_batch_get_estimate(mi_batch* batch, int problem, float* image);
/* Launch shape: forward tasks of at most fp_group_max bins run in lane
 * groups (the others one per wave; row groups, below, take precedence),
 * fp_widen 1 doubles those groups; columns of images of at most 32 planes run
 * one per full wave (bp_pack 0), packed in lane groups (1), or one per wave
 * split into 2 or 4 parts that sum different table entries (2); tile_x x
 * tile_y orders the columns. Defaults 32, 0, 2, 32, 8. */
int mi_batch_set_launch(mi_batch* batch, int fp_group_max, int fp_widen, int bp_pack, int tile_x, int tile_y);
/* Forward projection in row groups: the families of one basic row with at
 * most max_axial axial positions (0: none) run up to 64 axial positions per
 * wave with the row's elements loaded once (default 32). */
int mi_batch_set_row_groups(mi_batch* batch, int max_axial);
/* Times the candidate launch shapes on subset 0 (no side effects on the
 * estimates) and keeps the fastest; *choice receives its index. */
int mi_batch_tune(mi_batch* batch, int* choice);
/* num_subiterations subiterations of every problem, as mi_osem_run;
 * use_graph: the launches are captured once into a HIP graph (per
 * argument triple) and replayed. Returns without waiting. */
int mi_batch_osem_run(mi_batch* batch, int start_subiteration_num, int num_subiterations, int start_subset,
                      int use_graph);
int mi_batch_synchronize(mi_batch* batch);

/* ---- launch tuning ---------------------------------------------------------- */
/* No counterpart in the reference. Launch shape of the back projection: image
 * columns are visited in tiles of MI_TUNE_BP_TILE_X x MI_TUNE_BP_TILE_Y
 * (default 8 x 64; 1 x 1 is raster order); MI_TUNE_BP_XCD_REMAP maps
 * workgroups to XCDs: 0 in hardware order, 1 in one contiguous range per XCD,
 * k > 1 in chunks of k workgroups (4 columns each) dealt to the XCDs in turn
 * (default 64: half an 8 x 64 tile per chunk); MI_TUNE_BP_SLOT_SKIP skips z slots
 * an entry cannot reach (default off). Results do not depend on these
 * settings. MI_TUNE_FP_XCD_CHUNK maps forward-projection workgroups (4 tasks
 * each) the same way: 0 one contiguous range per XCD, k > 0 chunks of k
 * (default 32).
 * MI_TUNE_BP_SLAB 1 switches the back projection of tables with one plane
 * per axial position (and images of at most 128 planes) to projection-data
 * slabs staged in LDS, one workgroup per image tile (default 0: by image
 * column over the column table, the faster of the two at mMR size);
 * MI_TUNE_SLAB_TILE_X / _Y set its tile (default 8 x 8 columns),
 * MI_TUNE_SLAB_COLUMNS_PER_WAVE (2, 4 or 8, default 4) the columns each wave
 * of the tile's workgroup sums, and MI_TUNE_SLAB_LDS_KB (16-160, default 80)
 * the LDS a slab may take (the segments are split into as few groups per
 * view as keep it there).
 * MI_TUNE_TOF_ROWS selects how TOF projections get each element's TOF
 * weights: 1 (default) from records built at set-up that hold the weights of
 * a window of consecutive TOF bins per row element and per table entry (the
 * narrowest window that holds every element's live bins), 2 the same with
 * wide windows (14 forward, 13 back), 0 computed on the fly in every
 * projection (also used when no window fits). Forward projections are the
 * same bit for bit in all three.
 * These take effect at the next mi_set_up (they decide the table's layout);
 * the sums of one setting are deterministic, other settings agree to float
 * rounding.
 * MI_TUNE_BP_QUAD 1 builds, at mi_set_up, a second table for non-TOF
 * one-plane tables that merges the column lists of 2 x 2 blocks of image
 * columns, and back-projects one block per wave (each projection-data run
 * gathered once for the block's four columns: 0.58 of the gathers at mMR
 * size, measured no faster); 0 (default) keeps the column kernel. Between a
 * set-up with 1 and the next, 0 and 1 switch the kernel at the next launch.
 * The two give the same sums bit for bit;
 * MI_TUNE_BP_QUAD_UNROLL (4, 6 or 8, default 8) sets the merged entries
 * whose gathers each wave keeps in flight.
 * MI_TUNE_BP_SPLIT_SUBSETS 1 back-projects several subsets (a plain back
 * projection of all subsets) in one launch per subset, each continuing the
 * sums the previous one left in the image (the same sums bit for bit as one
 * launch; a working set of one subset's projection data; 2 % faster at mMR
 * size); 0 (default) in one launch. */
enum { MI_TUNE_BP_TILE_X = 1, MI_TUNE_BP_TILE_Y = 2, MI_TUNE_BP_XCD_REMAP = 3, MI_TUNE_BP_SLOT_SKIP = 4,
       MI_TUNE_FP_XCD_CHUNK = 5, MI_TUNE_BP_SLAB = 6, MI_TUNE_SLAB_TILE_X = 7, MI_TUNE_SLAB_TILE_Y = 8,
       MI_TUNE_SLAB_COLUMNS_PER_WAVE = 9, MI_TUNE_SLAB_LDS_KB = 10, MI_TUNE_TOF_ROWS = 11, MI_TUNE_BP_QUAD = 12,
       MI_TUNE_BP_QUAD_UNROLL = 13, MI_TUNE_BP_SPLIT_SUBSETS = 14 };
int mi_set_tuning(mi_ctx* ctx, int param, int value);

/* ---- diagnostics ----------------------------------------------------------- */
/* No counterpart in the reference (its TimedObject timers are CPU/wall clocks,
 * TimedObject.h:45-57). When enabled, every forward- and back-projection kernel
 * launch of the context is bracketed by HIP events recorded on the stream it
 * runs on; mi_get_kernel_timing returns the summed kernel time (ms) and the
 * launch counts since the last call, and resets them. */
int mi_set_kernel_timing(mi_ctx* ctx, int enable);
int mi_get_kernel_timing(mi_ctx* ctx, double* fp_ms, int64_t* fp_launches, double* bp_ms, int64_t* bp_launches);

#ifdef __cplusplus
}
#endif
#endif
