// Device data structures and kernel launchers of the projector pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pmrt_plan.hpp"

namespace mi {

// One wavefront of forward-projection work: one family, or the two families
// of a (+segment, -segment) pair that share a basic row and in-plane
// operation (second = -1 when single). Lane l of slot q handles combined
// axial index g = l + 64*q: family `first` for g < nax(first), then `second`.
struct FpTask {
  int32_t first;
  int32_t second;
};

// One entry of the voxel-driven back-projection table: the family's
// projection data run, the image z of its axial position 0, and the weight.
// Voxel z receives w * proj[sino_base + a] for a = (z - k) / ppa when that is
// an integer in [0, nax).
struct BpEntry {
  int32_t sino_base;
  int32_t k;
  int32_t nax_ppa;  // nax | ppa << 16
  float w;
};
static_assert(sizeof(BpEntry) == 16, "BpEntry layout");

// One entry of the 2 x 2 column-block back-projection table (pmrt_quad.hip):
// a family's run as in BpEntry and the weights of the block's columns
// (x0, y0), (x0 + 1, y0), (x0, y0 + 1), (x0 + 1, y0 + 1), 0 for a column
// without an element of the family at that plane.
struct QuadEntry {
  int32_t sino_base;
  int32_t k;
  int32_t nax_ppa;
  int32_t pad;
  float w[4];
};
static_assert(sizeof(QuadEntry) == 32, "QuadEntry layout");

struct ImageDims {
  int32_t nz, ny, nx;
  int32_t min_z, min_y, min_x;
  int32_t nzp;  // padded z stride of the device image layout
  int32_t ncols;
};

// Time of flight (ProjMatrixByBin::apply_tof_kernel, ProjMatrixByBin.inl:115-197).
constexpr int kErfSamples = 200000;  // FastErf samples (ProjMatrixByBin.cxx:174-176)
constexpr double kErfMax = 5.0;      // FastErf maximum sample value (FastErf.h:57)
struct TofConsts {
  int32_t ntof;       // TOF bins (> 1)
  float r;            // 1 / (sigma * sqrt(2))
  float z_off;        // scanner z of image plane 0
  float vz, vy, vx;   // voxel size
  float oz, oy, ox;   // image origin
  int32_t nvg;        // viewgrams per TOF block
  int32_t shared_bounds;  // high[k] == low[k+1] for all k (one erf per boundary)
  int64_t block;      // bins per TOF block (the non-TOF number of bins)
  const double* erf;  // kErfSamples + 2 samples of erf on [-kErfMax, kErfMax]
  float low[kMaxTofBins];
  float high[kMaxTofBins];
};

// Forward-projection epilogues.
enum FpMode : int {
  FP_PLAIN = 0,            // out = A x
  FP_RATIO = 1,            // out = divide_and_truncate(y, A x + b)
  FP_GRADIENT = 2,         // out = divide_and_truncate(y, A x + b) - m
  FP_LOGLIK = 3,           // partial sums of the Poisson log-likelihood (FP64)
  FP_GRADIENT_LOGLIK = 4,  // FP_GRADIENT and FP_LOGLIK in one pass
  FP_RATIO_LOGLIK = 5,     // FP_RATIO and FP_LOGLIK in one pass
  FP_DRATIO = 6,           // out = d divide_and_truncate(y, ybar) / d ybar at ybar = A x + b
  FP_SCALE = 7,            // out = (A x) * scale
  // list mode with the events histogrammed into y (LM_gradient_and_value,
  // ...ListModeDataWithProjMatrixByBin.cxx:748-787): out = y / (A x + b),
  // 0 where y = 0 or 1 > 1e4 (A x + b)
  FP_LM_RATIO = 8,
  FP_LM_VALUE = 9          // partial sums of -y log(A x + b) (-y log(1e-4) where capped)
};

struct FpData {
  const float* measured;      // y, device layout (end planes already zeroed when requested)
  const float* additive;      // per-bin b, or nullptr
  float additive_constant;    // b when additive == nullptr
  const float* vg_small;      // per-viewgram small_value = max(max(y_viewgram)*1e-6, 0)
  const float* mult;          // m for FP_GRADIENT (ones, zero at end planes), or nullptr (= 1)
  double* partial;            // FP_LOGLIK: one double per workgroup
  const float* scale;         // FP_SCALE: per-bin factor, device layout
};

// Back-projection epilogues.
enum BpMode : int {
  BP_PLAIN = 0,       // out = A^T p
  BP_OSEM_UPDATE = 1  // x *= clamp(divide(A^T p, sens))
};

struct BpData {
  const float* sens;  // subset sensitivity, device image layout
  float* estimate;    // x, device image layout, updated in place
  float* update;      // if not null, receives the multiplicative update u
  int threshold;      // apply threshold_upper_lower(thr_lo, thr_hi)
  float thr_lo = 0.F;            // "minimum relative change"
  float thr_hi = 3.40282347e38F;  // "maximum relative change" (FLT_MAX)
};

// ---- row builder (pmrt_trace.hip)
void launch_trace_basic_rows(const BasicRay* d_rays, int first, int count, const TraceConsts& tc, uint64_t* scratch_k,
                             float* scratch_w, int cap, int* d_len, int* d_buf, int* d_overflow, hipStream_t stream);
void launch_pack_rows(int first, int count, const uint64_t* scratch_k, const float* scratch_w, int cap, const int* d_len,
                      const int* d_buf, const int64_t* d_row_start, int2* d_pool, int* d_range_error, int4 lim,
                      hipStream_t stream);

// ---- projections (pmrt_project.hip)
void launch_forward_project(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams, const int2* d_pool,
                            const float* d_img, float* d_out, const ImageDims& dims, const FpData& data,
                            hipStream_t stream, int xcd_chunk = 0, const int2* d_ops = nullptr,
                            int64_t nelem = 0);
// copies of the basic rows per in-plane operation for launch_forward_project
// (d_ops: 8 * n elements)
void launch_expand_rows(const int2* d_pool, int64_t n, const ImageDims& dims, int2* d_ops, hipStream_t stream);
// Launch shape of the back projection: image columns are visited in
// tile_x x tile_y tiles (1 x 1 = raster order), optionally remapped so each XCD
// gets a contiguous range of tiles, optionally skipping z slots an entry cannot reach.
// Defaults: 8 x 64 tiles (8 columns along x, 64 along y), chunks of 64
// workgroups dealt to the XCDs in turn: the fastest shape measured at mMR size
// (10.07 ms against 10.56 ms for 16 x 16 tiles and 17.6 ms for 64 x 8 tiles in
// one contiguous range per XCD).
struct BpTuning {
  int tile_x = 8;
  int tile_y = 64;
  int xcd_remap = 64;  // 0: hardware order, 1: contiguous ranges per XCD, k > 1: chunks of k blocks dealt to the XCDs
  int slot_skip = 0;
  int quad_unroll = 8;  // 2 x 2 block kernel: merged entries in flight per wave (4, 6 or 8)
  int split_subsets = 0;  // plain projection of several subsets: 1 one launch per subset
};

// ppa: planes per axial position shared by all families (1 or 2 select a
// specialised kernel), 0 when it varies between segments.
// proj_values: the number of floats at d_proj (0 if unknown), used to pick
// 32-bit run-buffer arithmetic when they lie inside one 4 GB-aligned window
void launch_back_project(BpMode mode, int ppa, const int64_t* d_list_off, const BpEntry* d_entries, int subset_begin,
                         int subset_end, const float* d_proj, float* d_img, const ImageDims& dims, const BpData& data,
                         const BpTuning& tune, hipStream_t stream, int64_t proj_values = 0);
// 2-D transpose used for the STIR <-> device layout conversions:
// dst[c * dst_ld + r] = src[r * src_ld + c] for r < rows, c < cols.
void launch_transpose(const float* src, float* dst, int rows, int cols, int src_ld, int dst_ld, hipStream_t stream);

// ---- time-of-flight projections (pmrt_tof.hip); tof.ntof > 1
void launch_forward_project_tof(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams,
                                const FamilyLor* d_lors, const int2* d_pool, const float* d_img, float* d_out,
                                const ImageDims& dims, const TofConsts& tof, const FpData& data, hipStream_t stream);
void launch_back_project_tof(BpMode mode, int ppa, const int64_t* d_list_off, const BpEntry* d_entries,
                             const float* d_entry_d, int subset_begin, int subset_end, const float* d_proj,
                             float* d_img, const ImageDims& dims, const TofConsts& tof, const BpData& data,
                             const BpTuning& tune, hipStream_t stream);
// signed TOF distance d of every back-projection table entry, from its
// family's LOR and its voxel (the column of the entry and z = k)
void launch_entry_tof_distance(const Family* d_fams, int fam_begin, int fam_end, const int64_t* d_fam_entry_off,
                               const FamilyLor* d_lors, const int2* d_pool, const ImageDims& dims,
                               const TofConsts& tof, float* d_entry_d, hipStream_t stream);
void launch_gather_floats(const float* d_src, const uint32_t* d_idx, float* d_dst, int64_t n, hipStream_t stream);

// Stored TOF rows: each element's TOF weights computed once (as the
// on-the-fly kernels compute them) and kept as a window of consecutive TOF
// bins.
//   forward record of a row element, 2 + W words (tof_fp_words):
//     [0] col | c << 20  (c: first bin of the window)
//     [1] z << 16 | r  (r: elements from this one to the end of its run of
//         equal c in the row, at most 65535)
//     [2..] the weights of TOF bins c .. c + W - 1
//   back record of a table entry, 3 + W words (tof_bp_words):
//     [0] klo * block + sino_base (the run of the window's first bin)
//     [1] k  [2] nax | ppa << 16 | klo << 20 | nk << 25
//     [3..] the weights of TOF bins klo .. klo + nk - 1, zeros up to W
// A forward window starts at c = min(klo, nt - W) so it stays inside the nt
// accumulators of the kernel; bins outside the live range carry weight 0.
constexpr int kTofFpNarrow = 6, kTofFpWide = 14;  // forward window widths (8- and 16-word records)
constexpr int kTofBpNarrow = 5, kTofBpWide = 13;  // back window widths (8- and 16-word records)
constexpr int tof_fp_words(int w) { return 2 + w; }
constexpr int tof_bp_words(int w) { return 3 + w; }
// widest live range (khi - klo + 1 over bins with a non-zero weight) of the
// table entries [0, n) into *d_span (atomic max; *d_span zeroed by the caller)
void launch_tof_live_span(const BpEntry* d_entries, const float* d_entry_d, int64_t n, const TofConsts& tof,
                          int* d_span, hipStream_t stream);
// forward records of the families' rows (rows must not overlap); nt, w: the
// forward kernel's accumulator count and window width; the run lengths are
// filled in by the second launch
void launch_tof_fp_records(const Family* d_fams, int nfam, const FamilyLor* d_lors, const int2* d_pool,
                           const ImageDims& dims, const TofConsts& tof, int nt, int w, int32_t* d_rec,
                           hipStream_t stream);
void launch_tof_bp_records(const BpEntry* d_entries, const float* d_entry_d, int64_t n, const TofConsts& tof, int w,
                           int32_t* d_rec, hipStream_t stream);
void launch_forward_project_tof_rows(FpMode mode, const FpTask* d_tasks, int ntasks, const Family* d_fams,
                                     const int32_t* d_rec, int nt, int w, const float* d_img, float* d_out,
                                     const ImageDims& dims, const TofConsts& tof, const FpData& data,
                                     hipStream_t stream);
// d_rec holds num_entries records followed by one all-zero record
void launch_back_project_tof_rows(BpMode mode, int ppa, const int64_t* d_list_off, const int32_t* d_rec, int w,
                                  int64_t num_entries, int subset_begin, int subset_end, const float* d_proj,
                                  float* d_img, const ImageDims& dims, const TofConsts& tof, const BpData& data,
                                  const BpTuning& tune, hipStream_t stream);

// ---- grouped kernels for many problems in one launch (pmrt_batch.hip)
// One problem of a batch as the grouped kernels see it (device pointers into
// that problem's own buffers).
struct ProblemDesc {
  const Family* fams;
  const int2* pool;
  const float* fp_in;        // forward-projection input image (device layout)
  float* fp_out;             // forward-projection output
  FpData fd;                 // epilogue data of the forward projection
  const int64_t* list_off;   // back-projection table offsets per (subset, column)
  const BpEntry* entries;
  const float* bp_in;        // back-projection input
  float* bp_out;             // BP_PLAIN output image
  const float* const* sens;  // per-subset sensitivity images (BP_OSEM_UPDATE)
  float* estimate;           // BP_OSEM_UPDATE: x, updated in place
  ImageDims d;
  int32_t ppa;               // planes per axial position shared by all families, 0 if they differ
  const int2* pool_ops;      // the rows per in-plane operation (or null), nelem elements each
  int64_t nelem;
  float thr_lo, thr_hi;      // OSMAPOSL "minimum / maximum relative change" of the problem
};
// work item of a grouped launch: (problem, family) or (problem, image column);
// index -1 is padding
struct WorkItem {
  int32_t problem;
  int32_t index;
};
// a row group of a batched forward projection: item_fams[first .. first+count)
// are families of one problem sharing one basic row, at most 64 axial
// positions in total
struct RowGroupItem {
  int32_t problem;
  int32_t first;
  int32_t count;
  int32_t pad;
};
void launch_fp_rowgroup(FpMode mode, const ProblemDesc* d_descs, const RowGroupItem* d_items, const int32_t* d_item_fams,
                        int nitems, hipStream_t stream);
// G lanes per work item (G in 8, 16, 32, 64): a forward-projection group
// handles one family with at most G axial positions, a back-projection group
// one image column with at most G planes.
void launch_fp_group(FpMode mode, int G, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork,
                     hipStream_t stream);
void launch_bp_group(BpMode mode, int G, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork, int subset,
                     int threshold, hipStream_t stream);
// One work item per wave with the per-problem kernels' bodies (pmrt_project.hip):
// forward items index each problem's task list (± pairs) and every task of a
// launch has at most 64 * nslot bins; back-projection items are columns of
// images with at most 64 * nslot planes and `ppa` planes per axial position
// (1, 2, or 0 for any); nslot = -2 or -4: images of at most 32 or 16 planes,
// a wave split into that many parts working on different table entries.
void launch_fp_batch(FpMode mode, int nslot, const ProblemDesc* d_descs, const FpTask* const* d_task_lists,
                     const WorkItem* d_work, int nwork, hipStream_t stream);
void launch_bp_batch(BpMode mode, int ppa, int nslot, const ProblemDesc* d_descs, const WorkItem* d_work, int nwork,
                     int subset, int threshold, hipStream_t stream);

// ---- back projection through LDS slabs (pmrt_slab.hip): tables of one
// plane per axial position whose families cover whole runs (z-shift
// symmetry), images of at most 128 planes
constexpr int kSlabPad = 64;               // zeros opening each segment of a slab and following every run
constexpr int kSlabLead = 128;             // zeros before the first segment and after the last one
constexpr int kSlabMaxBytes = 160 * 1024;  // LDS of a workgroup on gfx950
// table entry: byte offset in the slab of axial index -k of the entry's run
// (k = plane of its axial position 0), and its weight for each 64-plane slot
// (0 for a slot it does not reach)
struct SlabEntry {
  int32_t off4;
  float w0, w1;
};
static_assert(sizeof(SlabEntry) == 12, "SlabEntry layout");
// first slab index of a segment (after `before` segments of `cum` axial
// positions in total in its group) in a slab of T tangential positions
__host__ __device__ inline int slab_segment_start(int before, int cum, int T) {
  return kSlabLead + before * kSlabPad + T * (cum + before * kSlabPad);
}
struct SlabArgs {
  const uint2* hdr;        // per (subset, tile, phase, wave): first table entry, steps
  const SlabEntry* ent;
  const int2* rng;         // tangential index range per (subset, tile, view slot); y < x: no entries
  const int32_t* view_of;  // view of (subset, view slot)
  const int32_t* nviews;   // view slots per subset
  const int4* segs;        // per segment in storage order: device offset, nax, 0, group
  const int2* seg_pos;     // per segment: nax of the segments before it in its group, their number
  const int2* groups;      // per segment group: first segment slot, end slot
  const int32_t* tile_order;  // tiles with entries (most entries first), then the empty ones
  int32_t tx, ty, ntx, ntiles, nv_max, ngroups, nseg, ntang, cpw, nwaves;
  int32_t nbusy;        // tiles with entries
  int32_t slab_floats;  // LDS floats of a workgroup (the largest slab)
};
void launch_slab_range(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end, const int2* d_pool,
                       int subset, const ImageDims& dims, const SlabArgs& A, int2* d_rng, hipStream_t stream);
void launch_slab_count(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end, const int2* d_pool,
                       int subset, const ImageDims& dims, const SlabArgs& A, uint32_t* d_counts, hipStream_t stream);
void launch_slab_make(const Family* d_fams, const int4* d_fam_slab, int fam_begin, int fam_end,
                      const int64_t* d_fam_entry_off, const int2* d_pool, int subset, const ImageDims& dims,
                      const SlabArgs& A, uint32_t* d_keys, SlabEntry* d_entries, hipStream_t stream);
void launch_slab_steps(const uint32_t* d_counts, int64_t nblocks, int cpw, int tc, uint32_t* d_steps,
                       hipStream_t stream);
void launch_slab_place(const uint32_t* d_keys_sorted, const uint32_t* d_idx_sorted, const SlabEntry* d_raw,
                       const uint32_t* d_list_start, const uint2* d_hdr, int64_t n, int cpw, int tc,
                       SlabEntry* d_table, hipStream_t stream);
void launch_slab_fill(SlabEntry* d_table, int64_t n, hipStream_t stream);
void launch_back_project_slab(BpMode mode, const SlabArgs& A, int subset_begin, int subset_end, const float* d_proj,
                              float* d_img, const ImageDims& dims, const BpData& data, hipStream_t stream);

// ---- BP table build (pmrt_table.hip)
void launch_count_entries(const Family* d_fams, int fam_begin, int fam_end, const int2* d_pool, int subset,
                          const ImageDims& dims, int32_t* d_counts, hipStream_t stream);
void launch_make_entries(const Family* d_fams, int fam_begin, int fam_end, const int64_t* d_fam_entry_off,
                         const int2* d_pool, int subset, const ImageDims& dims, uint32_t* d_keys, BpEntry* d_entries,
                         hipStream_t stream);
void launch_iota(uint32_t* d_vals, int64_t n, hipStream_t stream);
void launch_count_nonzeros(const Family* d_fams, int nfam, const int2* d_pool, const ImageDims& dims,
                           unsigned long long* d_total, hipStream_t stream);
void launch_gather_entries(const BpEntry* d_src, const uint32_t* d_idx, BpEntry* d_dst, int64_t n, hipStream_t stream);
// ---- 2 x 2 column-block back projection (pmrt_quad.hip)
// merge keys of the sorted column table (d_perm: generation index of each entry)
void launch_quad_keys(const uint32_t* d_perm, int64_t n, const int64_t* d_famoff, int nfam, const Family* d_fams,
                      const int2* d_pool, uint64_t* d_key, hipStream_t stream);
// per (subset, block): count (d_qcount) or write (from d_qoff) the merged entries
void launch_quad_merge(bool count, const int64_t* d_list_off, const BpEntry* d_ent, const uint64_t* d_key,
                       const ImageDims& dims, int nqx, int nq, int nsub, const int64_t* d_qoff, int32_t* d_qcount,
                       QuadEntry* d_qent, int* d_bad, hipStream_t stream);
// one-plane tables only; tiles of tune.tile_x/2 x tune.tile_y/2 blocks
void launch_back_project_quad(BpMode mode, const int64_t* d_qoff, const QuadEntry* d_qent, int subset_begin,
                              int subset_end, const float* d_proj, float* d_img, const ImageDims& dims,
                              const BpData& data, const BpTuning& tune, hipStream_t stream, int64_t proj_values);

// ---- small utilities (pmrt_project.hip)
void launch_fill(float* d, int64_t n, float v, hipStream_t stream);
void launch_viewgram_max(const float* d_proj, const int64_t* d_vg_off, const int32_t* d_vg_len, int nvg, float* d_small,
                         hipStream_t stream);
void launch_zero_end_planes(float* d_proj, const int64_t* d_off, int count, int run, int stride, hipStream_t stream);
// the objective's data-term Hessian products (one workgroup per viewgram):
// q = divide_and_truncate(y * fin, (fcur + b)^2) (exact) or
// divide_and_truncate(fin, y / e / e) (approximate, e = max(1e-20, d_mult) or
// 1) over the nvg viewgrams d_vgs lists (the first nvg without a list); d_q
// may be d_fin; *d_negative set when fin holds a negative value
void launch_hessian_ratio(const int32_t* d_vgs, int nvg, const int64_t* d_vg_off, const int32_t* d_vg_len,
                          const float* d_y, const float* d_fin, const float* d_fcur, const float* d_add,
                          float add_const, const float* d_mult, int approximate, float* d_q, int* d_negative,
                          hipStream_t stream);
// the list-mode objective's Hessian product over histogrammed events y:
// q = y * (-1 / fwd^2 * fin) with fwd = fcur + b, 0 where y = 0,
// 1 > 1e4 fwd or fin = 0 (d_q may be d_fin)
void launch_lm_hessian(const int32_t* d_vgs, int nvg, const int64_t* d_vg_off, const int32_t* d_vg_len,
                       const float* d_y, const float* d_fin, const float* d_fcur, const float* d_add, float add_const,
                       float* d_q, hipStream_t stream);

// ---- objective-function helpers (pmrt_objective.hip)
// QuadraticPrior weights w[dz+1][dy+1][dx+1].
struct PriorWeights {
  float w[27];
};
// Image of a prior kernel: [z][y][x] (zfast = 0) or the device layout
// [y][x][zp] (zfast = 1, kappa and gradient in the same layout).
struct PriorImage {
  int nz, ny, nx, nzp;
  int zfast;
};
// Value partials without beta (one double per workgroup, count returned by
// prior_num_partials) and beta * gradient (if d_grad) of a neighbourhood
// prior: kind 1 quadratic, 2 relative difference (gamma, eps); kappa may be null.
int prior_num_partials(const PriorImage& im);
void launch_prior(int kind, const float* d_img, const float* d_kappa, const PriorImage& im, const PriorWeights& w,
                  int only_2D, float beta, float gamma, float eps, float* d_grad, double* d_partial,
                  hipStream_t stream);
// d_out += Hessian products of a [z][y][x] prior (kind 1 quadratic, 2 relative
// difference: accumulate_Hessian_times_input; 3: the quadratic prior's
// add_multiplication_with_approximate_Hessian)
void launch_prior_hessian(int kind, const float* d_x, const float* d_in, const float* d_kappa, const PriorImage& im,
                          const PriorWeights& w, int only_2D, float beta, float gamma, float eps, float* d_out,
                          hipStream_t stream);
// Allocation-free prior: FP64 partials in a fixed per-device array of
// kNoAllocPartials doubles (512 KB), one per prior workgroup; its device
// address on `device`, which must be current (null if it cannot be had).
constexpr int kNoAllocPartials = 65536;
constexpr int kMaxDevices = 64;
double* noalloc_partials(int device);
// *d_small = max(max(v) * small_num, 0) over n floats (d_part: max_num_partials() floats)
int max_num_partials();
void launch_small_value(const float* d_v, int64_t n, float small_num, float* d_part, float* d_small, hipStream_t stream);
// One-step-late OSMAPOSL update on the device layout (n = ncols * nzp elements)
void launch_osl_update(float* d_x, const float* d_upd, const float* d_pgrad, const float* d_sens, const float* d_small,
                       int nsub, int multiplicative, int threshold, float lo, float hi, int nz, int nzp, int64_t n,
                       hipStream_t stream);
// dst += src; dst /= d (element-wise, n elements)
void launch_accumulate(float* d_dst, const float* d_src, int64_t n, hipStream_t stream);
void launch_divide_scalar(float* d_dst, float d, int64_t n, hipStream_t stream);
void launch_subtract(float* d_dst, const float* d_src, int64_t n, hipStream_t stream);  // dst -= src
// *d_out = sum of n doubles in a fixed order (one workgroup); d_out may be null-checked by the caller.
void launch_sum_partials(const double* d_partial, int n, double scale, double* d_out, hipStream_t stream);
// v = g * x / sens (0 where sens == 0), device image layout, n elements; v may alias g or x
void launch_osem_vjp_scale(const float* d_g, const float* d_x, const float* d_sens, float* d_v, int64_t n,
                           hipStream_t stream);
// out += g * u, n elements
void launch_add_product(float* d_out, const float* d_g, const float* d_u, int64_t n, hipStream_t stream);
// out = g where the update u was not clamped by threshold_upper_lower(lo, hi)
// (lo < u < hi), 0 where it was (u <= lo or u >= hi): the clamp's derivative
void launch_osem_vjp_mask(const float* d_g, const float* d_u, float* d_out, int64_t n, float lo, float hi,
                          hipStream_t stream);

}  // namespace mi
