/*
 * CPU oracle for the ray-tracing-matrix projector pair and OSEM.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of STIR's
 * ProjMatrixByBinUsingRayTracing (PMRT) projector pair, its PET
 * Cartesian-grid symmetries, the Poisson log-likelihood sub-gradient and
 * the OSMAPOSL multiplicative update.  The HIP path is compared against
 * it in the tests; it is never called by the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the STIR source tree, STIR 6.5.0).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - RPTsens_seg3_PM (reference-held golden sensitivity image): pins the
 *     Siddon tracer, FOV clipping, multi-tangential merging, the axial
 *     TOR handling for span-1 data, the symmetries and the back projector.
 *   - closed forms: axis-aligned rays, total row length, z-TOR overlaps.
 *   - brute force: an FP64 exact-interval tracer on tiny grids.
 *   - symmetric rows equal directly traced rows (reference invariant
 *     test_DataSymmetriesForBins_PET_CartesianGrid.cxx:112-122).
 *   - adjoint identity, finite-difference gradient, one-voxel EM closed
 *     form, OSEM likelihood monotonicity.
 *   - use_actual_detector_boundaries: rows against nominal plans (even
 *     tangential positions: the same scanner; odd ones: the scanner turned
 *     by -pi/N; wrapped detector pairs: the z-mirrored LOR), the reset
 *     rules, a planted detector-pair mistake (tests/test_actual_boundaries_oracle.py).
 *   - Hessian products (projection data and list mode): central differences
 *     of the gradients, symmetry and sign, closed forms, planted mistakes
 *     (tests/test_hessian_oracle.py).
 * Parity unpinned: the choice between float and double libm overloads in
 * get_s/get_tantheta/get_costheta (depends on header inclusion order in
 * the reference build, which cannot be compiled here), the OpenMP
 * reduction order of the reference back projector, and with
 * use_actual_detector_boundaries the doubled spacing of several tangential
 * LORs (checked only against the GPU's restatement of the same rule).
 */
#ifndef PMRT_ORACLE_H
#define PMRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_SEGMENTS 256
#define OR_MAX_TOF 64

/* Geometry of one reconstruction problem. Segment arrays are indexed by
 * (segment_num - min_segment_num). Image index ranges follow
 * VoxelsOnCartesianGrid (z in [min_z, min_z+nz), y in [min_y, ...), x ...). */
typedef struct {
  /* scanner / ProjDataInfoCylindrical */
  int32_t num_rings;
  int32_t num_detectors_per_ring;
  float effective_ring_radius; /* mm */
  float ring_spacing;          /* mm */
  float intrinsic_tilt;        /* rad */
  int32_t arc_corrected;       /* 1: ProjDataInfoCylindricalArcCorr */
  float bin_size;              /* mm, arc-corrected tangential sampling */
  int32_t num_views;
  int32_t num_tangential_poss;
  int32_t min_tangential_pos_num;
  int32_t min_segment_num;
  int32_t max_segment_num;
  int32_t min_ring_diff[OR_MAX_SEGMENTS];
  int32_t max_ring_diff[OR_MAX_SEGMENTS];
  int32_t num_axial_poss[OR_MAX_SEGMENTS];
  /* image / VoxelsOnCartesianGrid */
  int32_t nz, ny, nx;
  int32_t min_z, min_y, min_x;
  float vox_z, vox_y, vox_x;
  float origin_z, origin_y, origin_x;
  /* Ray Tracing Matrix Parameters (ProjMatrixByBinUsingRayTracing.cxx:76-89) */
  int32_t restrict_to_cylindrical_FOV;
  int32_t num_tangential_LORs;
  int32_t do_symmetry_90degrees_min_phi;
  int32_t do_symmetry_180degrees_min_phi;
  int32_t do_symmetry_swap_segment;
  int32_t do_symmetry_swap_s;
  int32_t do_symmetry_shift_z;
  /* time of flight (ProjDataInfo::set_tof_mash_factor, ProjDataInfo.cxx:174-258;
   * ProjMatrixByBin::enable_tof, ProjMatrixByBin.cxx:66-75). num_tof_bins <= 1:
   * non-TOF. Bin k (0-based, = timing_pos_num - min_tof_pos_num) covers
   * [tof_low[k], tof_high[k]] mm along the LOR; r_sqrt2_gauss_sigma =
   * 1/(sigma*sqrt(2)) with sigma the Gaussian kernel's standard deviation in mm.
   * TOF projection data hold num_tof_bins non-TOF blocks, TOF bin outermost. */
  int32_t num_tof_bins;
  float tof_r_sqrt2_gauss_sigma;
  float tof_low[OR_MAX_TOF];
  float tof_high[OR_MAX_TOF];
  int32_t use_actual_detector_boundaries; /* ProjMatrixByBinUsingRayTracing.cxx:608-634 */
} or_geometry;

typedef struct or_plan or_plan;

/* Build the plan: geometry, symmetries and all basic rows (the reference's
 * basic-bin cache, ProjMatrixByBin.inl:48-113). Returns NULL on error. */
or_plan* or_plan_create(const or_geometry* g, char* err, int errlen);
void or_plan_destroy(or_plan* p);

/* Effective symmetry flags after the reference's auto-disable rules
 * (DataSymmetriesForBins_PET_CartesianGrid.cxx:241-378). out[5]. */
void or_symmetry_flags(const or_plan* p, int32_t* out);

/* Total number of bins, and the STIR ProjDataInMemory index of a bin
 * (ProjDataInMemory.cxx:190-235, ProjData.cxx:641-661). */
int64_t or_num_bins(const or_plan* p);
int64_t or_bin_index(const or_plan* p, int seg, int view, int ax, int tang);

/* Row of one bin through basic bin + symmetry operation
 * (ProjMatrixByBin.inl:48-113). Returns the number of elements, or
 * -(needed) when cap is too small. */
int or_get_row(const or_plan* p, int seg, int view, int ax, int tang,
               int16_t* z, int16_t* y, int16_t* x, float* w, int cap);
/* Row computed directly for the given bin (no symmetry), as
 * calculate_proj_matrix_elems_for_one_bin (PMRT.cxx:582-829). */
/* TOF row of bin (seg, view, ax, tang, TOF bin tof in 0..num_tof_bins-1):
 * the basic row's elements with a positive TOF kernel value, weights value *
 * kernel (ProjMatrixByBin.inl:115-181), transformed to the bin. */
int or_get_row_tof(const or_plan* p, int seg, int view, int ax, int tang, int tof,
                   int16_t* z, int16_t* y, int16_t* x, float* w, int cap);
int or_trace_row(const or_plan* p, int seg, int view, int ax, int tang,
                 int16_t* z, int16_t* y, int16_t* x, float* w, int cap);
/* RayTraceVoxelsOnCartesianGrid on one segment (voxel units, coordinates
 * z,y,x), for closed-form and brute-force checks. Same return convention. */
int or_ray_trace(const float* start_zyx, const float* stop_zyx, const float* vox_zyx, float nc,
                 int16_t* z, int16_t* y, int16_t* x, float* w, int cap);
/* Basic bin (seg, view, ax, tang) and symmetry operation of a bin:
 * op = (in-plane code, z sign, z add) with z' = zsign*z + zadd and the in-plane
 * codes 0 id, 1 (-x,y), 2 (x,-y), 3 (-x,-y), 4 (y,x), 5 (-y,x), 6 (y,-x), 7 (-y,-x). */
void or_symmetry_op(const or_plan* p, int seg, int view, int ax, int tang, int32_t* basic4, int32_t* op3);
/* Per-bin geometry (s, phi, tantheta, costheta, t) as float[5]. */
void or_bin_geometry(const or_plan* p, int seg, int view, int ax, int tang, float* out5);

/* Subsets: number of (view,segment) pairs in subset and the list
 * (find_basic_vs_nums_in_subset.cxx:31-69 + related viewgrams
 * DataSymmetriesForBins_PET_CartesianGrid.inl:801-849). */
int or_subset_view_segments(const or_plan* p, int subset, int num_subsets,
                            int32_t* views, int32_t* segs, int cap);
/* Subset balance test (PoissonLogLikelihood...AndProjData.cxx:490-529). */
int or_subsets_balanced(const or_plan* p, int num_subsets);

/* Projections in STIR layouts: image [z][y][x], projection data
 * ProjDataInMemory order. subset < 0 means all bins. FP writes only the
 * bins of the subset. */
void or_forward_project(const or_plan* p, const float* img, float* sino, int subset, int num_subsets);
void or_forward_project_f64(const or_plan* p, const double* img, double* sino, int subset, int num_subsets);
void or_back_project(const or_plan* p, const float* sino, float* img, int subset, int num_subsets);
void or_back_project_f64(const or_plan* p, const double* sino, double* img, int subset, int num_subsets);
/* OpenMP variants: per-thread images reduced serially, as the reference
 * (BackProjectorByBin.cxx:316-346, distributable.cxx:414-434). */
void or_forward_project_omp(const or_plan* p, const float* img, float* sino, int subset, int num_subsets);
void or_back_project_omp(const or_plan* p, const float* sino, float* img, int subset, int num_subsets);

/* Number of matrix elements with z inside the image (nnz after clipping). */
int64_t or_nnz(const or_plan* p, int subset, int num_subsets);

/* Objective function pieces (PoissonLogLikelihoodWithLinearModelForMeanAndProjData).
 * bkg may be NULL (no additive term) and is a full sinogram otherwise.
 * zero_end_planes: "zero end planes of segment 0". */
void or_subset_sensitivity(const or_plan* p, int subset, int num_subsets, int zero_end_planes, float* img);
/* "use time-of-flight sensitivities" (...AndProjData.cxx:143): off by default,
 * then TOF data get the non-TOF sensitivity (:533-536,648-670). */
void or_set_use_tof_sensitivities(or_plan* p, int use);
/* The ratio projection data of the gradient RPC: divide_and_truncate(y,
 * FP(x) + b) per viewgram (minus mult or 1 unless add_sensitivity), the
 * subset's bins only (others 0); q holds all bins. */
void or_subset_ratio(const or_plan* p, const float* y, const float* bkg, const float* x,
                     int subset, int num_subsets, int zero_end_planes, int add_sensitivity, float* q);
/* compute_sub_gradient_without_penalty(_plus_sensitivity):
 * grad = BP_subset(y/ybar) (+sens) or BP_subset(y/ybar - 1). */
/* The data-term Hessian products of one subset (accumulate_sub_Hessian_times_
 * input_without_penalty when approximate == 0, add_multiplication_with_
 * approximate_sub_Hessian_without_penalty when 1): out -= BP(q). Returns 1
 * (out untouched) when the forward projection of input has a negative value. */
int or_hessian_times_input(const or_plan* p, const float* y, const float* bkg, const float* current,
                           const float* input, int subset, int nsub, int approximate, float* out);
void or_subset_gradient(const or_plan* p, const float* y, const float* bkg, const float* x,
                        int subset, int num_subsets, int zero_end_planes, int add_sensitivity,
                        float* grad);
double or_log_likelihood(const or_plan* p, const float* y, const float* bkg, const float* x,
                         int subset, int num_subsets, int zero_end_planes);
/* Multiplicative factors m of the model ybar = m (A x + b) (normalisation
 * undo() of ones, distributable.cxx:197-215), all TOF bins in measured order,
 * caller-owned; NULL removes them. Used by the sensitivities (BP of m), the
 * gradient (BP(y/(Ax+b) - m)) and the log-likelihood (ybar = (Ax+b) m). */
void or_set_multiplicative(or_plan* p, const float* mult);
/* "use subset sensitivities" (default 1); 0: every subset uses total / num_subsets. */
void or_set_use_subset_sensitivities(or_plan* p, int use);
/* OSMAPOSL "MAP_model" (0 additive, 1 multiplicative) and "minimum/maximum
 * relative change" (defaults 0, FLT_MAX). */
void or_set_osem_options(or_plan* p, int map_multiplicative, float min_rel, float max_rel);
/* Prior of the one-step-late update: type 0 none, 1 quadratic, 2 relative
 * difference (gamma, eps); w27 NULL = default weights; kappa caller-owned or NULL. */
void or_set_prior(or_plan* p, int type, float beta, float gamma, float eps, const float* kappa, const float* w27,
                  int only_2D);
/* OSMAPOSL (OSMAPOSLReconstruction.cxx:266-516, IterativeReconstruction.cxx:384-428,630-639)
 * with the settings above, initial positivity enforced on x. */
void or_osem(const or_plan* p, const float* y, const float* bkg, float* x,
             int num_subsets, int num_subiterations, int start_subset, int zero_end_planes,
             int use_omp);

/* QuadraticPrior (QuadraticPrior.cxx:202-367): weights w27[dz+1][dy+1][dx+1]
 * = vox_x/|offset|; value = beta * sum w (x_v - x_n)^2 / 4 (returned),
 * grad_v = beta * sum_n w (x_v - x_n) (written if grad != NULL), kappa
 * optional (NULL). Images [nz][ny][nx]. */
void or_quadratic_prior_weights(float vox_z, float vox_y, float vox_x, float* w27);
double or_quadratic_prior(const float* img, const float* kappa, int nz, int ny, int nx, const float* w27,
                          int only_2D, float beta, float* grad);
/* RelativeDifferencePrior (RelativeDifferencePrior.cxx:250-433): value =
 * beta * sum w 0.5 (x_v - x_n)^2 / (x_v + x_n + gamma |x_v - x_n| + eps)
 * (returned), grad_v = beta * sum_n w derivative_10(x_v, x_n). */
double or_relative_difference_prior(const float* img, const float* kappa, int nz, int ny, int nx, const float* w27,
                                    int only_2D, float beta, float gamma, float eps, float* grad);

/* Prior Hessian products accumulated into out: kind 1 quadratic and 2
 * relative difference accumulate_Hessian_times_input, kind 3 the quadratic
 * prior's add_multiplication_with_approximate_Hessian. */
void or_prior_hessian_times_input(int kind, const float* x, const float* in, const float* kappa, int nz, int ny, int nx,
                                  const float* w27, int only_2D, float beta, float gamma, float eps, float* out);
/* List mode: events n x (segment, view, axial, tangential, timing position).
 * or_lm_gradient accumulates sum over the subset's events of row^T (1/fwd)
 * into grad (if not NULL) and returns the reference's accumulated value
 * -sum log(fwd) (log(1/1e4) where 1 > 1e4 fwd); or_lm_osem runs prior-free
 * OSMAPOSL on the events (...ListModeDataWithProjMatrixByBin.cxx). */
double or_lm_gradient(const or_plan* p, const int32_t* ev, int64_t n, const float* add, const float* x, int subset,
                      int num_subsets, float* grad);
void or_lm_osem(const or_plan* p, const int32_t* ev, int64_t n, const float* add, float* x, int num_subsets,
                int num_subiterations, int start_subset);
/* list-mode Hessian product of the subset's events (LM_Hessian), accumulated into out */
void or_lm_hessian(const or_plan* p, const int32_t* ev, int64_t n, const float* add, const float* x,
                   const float* rhs, int subset, int nsub, float* out);

#ifdef __cplusplus
}
#endif
#endif
