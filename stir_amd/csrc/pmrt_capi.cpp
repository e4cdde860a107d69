// C ABI of the MI355X ray-tracing-matrix projector pair (include/stir_amd_pmrt.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "pmrt_device.hpp"
#include "pmrt_plan.hpp"
#include "stir_amd_pmrt.h"

namespace mi {
hipError_t radix_sort_pairs(void* d_temp, size_t& temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, int n, int end_bit, hipStream_t stream);
hipError_t exclusive_scan_counts(void* d_temp, size_t& temp_bytes, const int32_t* in, int64_t* out, int n,
                                 hipStream_t stream);
hipError_t exclusive_scan_u32(void* d_temp, size_t& temp_bytes, const uint32_t* in, uint32_t* out, int n,
                              hipStream_t stream);
void launch_fill_int2(int2* d, int64_t n, int2 v, hipStream_t stream);
}  // namespace mi

namespace {

constexpr int kRowSlack = 16;  // elements after the row copies: the forward kernel reads a group ahead

thread_local std::string g_last_error;

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory) ? MI_ERR_MEMORY : MI_ERR_DEVICE;
    throw ApiError(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}
#define HIPCHECK(x) hip_check((x), #x)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return MI_OK;
  } catch (const ApiError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return MI_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return MI_ERR_ARGUMENT;
  }
}

// Owning device buffer.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    HIPCHECK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    n = count;
  }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace

struct mi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<mi::Plan> plan;
  mi::ImageDims dims{};
  // matrix
  DBuf<int2> pool;
  DBuf<int2> pool_ops;  // the rows per in-plane operation: ((column << 12) | (z + 2048), weight)
  DBuf<mi::Family> fams;
  DBuf<mi::FpTask> tasks;
  std::vector<int> task_off;  // per subset, size nsub+1
  // basic view-segment groups: the related viewgrams of one basic
  // (view, segment), the unit find_basic_vs_nums_in_subset hands to the
  // projectors (find_basic_vs_nums_in_subset.cxx:31-69)
  std::unordered_map<int64_t, int> vs_group;            // basic (segment, view) -> group
  std::vector<std::vector<std::pair<int, int>>> group_vs;  // group -> related (view, segment)
  std::vector<int> task_group;                          // forward task -> group
  std::vector<std::vector<int>> subset_groups;          // subset of the table -> its groups, ascending
  std::vector<mi::FpTask> h_tasks;
  DBuf<mi::FpTask> sel_tasks;  // forward tasks of a view-segment request
  bool have_image = false;     // ctx->img holds the image of mi_set_image
  DBuf<int64_t> list_off;
  DBuf<mi::BpEntry> entries;
  // back projection through LDS slabs (build_slab_table), or by image
  // column over the column table (list_off) otherwise
  bool slab_allowed = true;  // batches read the column table
  int slab_enable = 0;       // MI_TUNE_BP_SLAB
  int slab_tx = 8, slab_ty = 8, slab_cpw = 4, slab_lds_kb = 80;
  bool use_slab = false;
  DBuf<uint2> slab_hdr;
  DBuf<mi::SlabEntry> slab_ent;
  DBuf<int2> slab_rng, slab_seg_pos, slab_groups;
  DBuf<int32_t> slab_view_of, slab_nviews, slab_tiles;
  DBuf<int4> slab_segs;
  mi::SlabArgs slab{};
  void release_slab() {
    use_slab = false;
    slab_hdr.release(); slab_ent.release(); slab_rng.release(); slab_seg_pos.release(); slab_groups.release();
    slab_view_of.release(); slab_nviews.release(); slab_tiles.release(); slab_segs.release();
    slab = mi::SlabArgs{};
  }
  // back projection over 2 x 2 column blocks (pmrt_quad.hip) for one-plane
  // tables: MI_TUNE_BP_QUAD (1: build the block table at set-up and use it;
  // default 0: it halves the gathers but measured no faster at mMR size)
  int quad_enable = 0;
  DBuf<int64_t> quad_off;  // per (subset, block), size nsub * blocks + 1
  DBuf<mi::QuadEntry> quad_ent;
  int64_t quad_entries = 0;
  bool use_quad() const { return quad_enable && quad_ent.p != nullptr; }
  void build_quad_table(const uint32_t* d_perm, int64_t n, const int64_t* d_famoff);
  std::vector<mi::Family> h_fams;
  std::unordered_map<int64_t, int> fam_lookup;  // (seg,view,tang,a0) -> family
  int64_t num_basic_elements = 0;
  int64_t num_matrix_elements = 0;
  int64_t num_nonzeros = 0;
  int uniform_ppa = 0;  // planes per axial position if equal for all families, else 0
  // time of flight: ntof blocks of block_bins bins each
  int ntof = 1;
  int64_t block_bins = 0;
  DBuf<mi::FamilyLor> lors;
  DBuf<float> entry_d;   // TOF distance of every back-projection table entry
  // stored TOF rows (records with each element's TOF weights, pmrt_tof.hip):
  // mode 0 computes the weights on the fly, 1 stores the narrowest windows
  // that hold every element's live TOF bins, 2 the wide ones
  int tof_rows_mode = 1;
  DBuf<int32_t> tof_fp_rec, tof_bp_rec;
  int tof_fp_nt = 0, tof_fp_w = 0, tof_bp_w = 0;  // window widths, 0 when not stored
  int tof_live_span = 0;  // widest range of TOF bins with a non-zero weight of one element
  DBuf<double> erf_tab;  // FastErf samples
  mi::TofConsts tofc{};
  bool tof_sensitivities = false;  // "use time-of-flight sensitivities" (...AndProjData.cxx:143)
  int64_t total_bins() const { return block_bins * ntof; }
  mi::BpTuning bp_tuning;
  // the time-of-flight back projection: 32 x 8 tiles in hardware order (measured best for it)
  mi::BpTuning tof_bp_tuning{32, 8, 0, 0};
  int fp_xcd_chunk = 32;  // forward workgroups: 0 contiguous per XCD, k > 0 chunks of k dealt to the XCDs
  // kernel timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> fp_events, bp_events;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  void release_events() {
    for (auto& p : fp_events) { event_pool.push_back(p.first); event_pool.push_back(p.second); }
    for (auto& p : bp_events) { event_pool.push_back(p.first); event_pool.push_back(p.second); }
    fp_events.clear();
    bp_events.clear();
  }
  ~mi_ctx() {
    release_events();
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
  }
  // images / projection data (device layout) and staging (STIR layout)
  DBuf<float> img, img2, sino, ratio, stage_img, stage_sino;
  // objective function
  DBuf<float> y, add, mult, ones, vg_small, x;
  // the measured data and additive term as given, kept when "zero end planes
  // of segment 0" zeroes y and add (the Hessian products read them unzeroed)
  DBuf<float> y_raw, add_raw;
  // the Hessian products: each subset's viewgrams (TOF bins included) and
  // the flag of a negative forward projection
  DBuf<int32_t> subset_vgs;
  std::vector<int> subset_vg_off;
  DBuf<int32_t> hess_flag;
  std::vector<std::unique_ptr<DBuf<float>>> sens;
  DBuf<int64_t> vg_off;
  DBuf<int32_t> vg_len;
  DBuf<double> partial, prior_partial;
  bool have_data = false, have_sens = false, have_x = false;
  float add_const = 0.F;
  int zero_end_planes = 0;
  // multiplicative factors given by the caller (device layout), and whether
  // every TOF bin carries the same factors
  DBuf<float> mult_user;
  DBuf<float> sens_total;  // total sensitivity when subsets use total / num_subsets
  bool lm = false;         // the data are histogrammed list-mode events (mi_lm_set_events)
  mi::FpMode ratio_mode() const { return lm ? mi::FP_LM_RATIO : mi::FP_RATIO; }
  bool mult_tof_uniform = true;
  bool use_subset_sens = true;  // "use subset sensitivities"
  // OSMAPOSL: "MAP_model", "minimum/maximum relative change", the prior of the
  // one-step-late update (kind 0 none, 1 quadratic, 2 relative difference)
  int map_multiplicative = 0;
  float min_rel = 0.F, max_rel = FLT_MAX;
  int prior_kind = 0;
  float prior_beta = 0.F, prior_gamma = 2.F, prior_eps = 0.F;
  int prior_only_2D = 0;
  mi::PriorWeights prior_w{};
  DBuf<float> prior_kappa, prior_grad, small_part, small_val;
  bool prior_active() const { return prior_kind != 0 && prior_beta != 0.F; }
  mi::PriorImage device_prior_image() const {
    const mi_geometry& g = plan->geometry();
    return mi::PriorImage{g.nz, g.ny, g.nx, dims.nzp, 1};
  }
  mi::PriorImage stir_prior_image() const {
    const mi_geometry& g = plan->geometry();
    return mi::PriorImage{g.nz, g.ny, g.nx, g.nz, 0};
  }
  void rebuild_mult(hipStream_t st);

  void require_setup() const {
    if (!plan) throw ApiError(MI_ERR_STATE, "mi_set_up has not been called");
  }
  int subset_begin(int subset) const { return subset < 0 ? 0 : subset; }
  int subset_end(int subset) const { return subset < 0 ? plan->num_subsets() : subset + 1; }
  void check_subset(int subset) const {
    if (subset >= plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
  }
  // *_dev entry points: the caller's stream; NULL is the HIP null stream.
  static hipStream_t pick(void* s) { return static_cast<hipStream_t>(s); }

  // ---- layout conversions
  void image_to_device(const float* src_stir, float* dst, hipStream_t st) const {
    mi::launch_transpose(src_stir, dst, dims.nz, dims.ncols, dims.ncols, dims.nzp, st);
  }
  void image_from_device(const float* src, float* dst_stir, hipStream_t st) const {
    mi::launch_transpose(src, dst_stir, dims.ncols, dims.nz, dims.nzp, dims.ncols, st);
  }
  // projection data: each TOF block (one block without TOF) is converted on its own
  void proj_to_device(const float* src_stir, float* dst, hipStream_t st) const {
    const mi_geometry& g = plan->geometry();
    const int cols = g.num_views * g.num_tangential_poss;
    for (int k = 0; k < ntof; ++k)
      for (int seg : plan->segment_sequence()) {
        const int64_t off = plan->segment_offset(seg) + k * block_bins;
        const int nax = plan->num_axial(seg);
        mi::launch_transpose(src_stir + off, dst + off, nax, cols, cols, nax, st);
      }
  }
  void proj_from_device(const float* src, float* dst_stir, hipStream_t st) const {
    const mi_geometry& g = plan->geometry();
    const int cols = g.num_views * g.num_tangential_poss;
    for (int k = 0; k < ntof; ++k)
      for (int seg : plan->segment_sequence()) {
        const int64_t off = plan->segment_offset(seg) + k * block_bins;
        const int nax = plan->num_axial(seg);
        mi::launch_transpose(src + off, dst_stir + off, cols, nax, nax, cols, st);
      }
  }

  // ---- projections on device-layout buffers
  void fp(mi::FpMode mode, const float* d_img, float* d_out, int subset, const mi::FpData& data, hipStream_t st) {
    const int tb = task_off[subset_begin(subset)], te = task_off[subset_end(subset)];
    fp_tasks(mode, tasks.p + tb, te - tb, d_img, d_out, data, st);
  }
  void fp_tasks(mi::FpMode mode, const mi::FpTask* d_tasks, int ntasks, const float* d_img, float* d_out,
                const mi::FpData& data, hipStream_t st) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timing) { e0 = take_event(); e1 = take_event(); HIPCHECK(hipEventRecord(e0, st)); }
    if (ntof > 1 && tof_fp_w > 0)
      mi::launch_forward_project_tof_rows(mode, d_tasks, ntasks, fams.p, tof_fp_rec.p, tof_fp_nt, tof_fp_w, d_img, d_out,
                                          dims, tofc, data, st);
    else if (ntof > 1)
      mi::launch_forward_project_tof(mode, d_tasks, ntasks, fams.p, lors.p, pool.p, d_img, d_out, dims, tofc, data, st);
    else
      mi::launch_forward_project(mode, d_tasks, ntasks, fams.p, pool.p, d_img, d_out, dims, data, st, fp_xcd_chunk,
                                 pool_ops.p, num_basic_elements);
    if (timing) { HIPCHECK(hipEventRecord(e1, st)); fp_events.emplace_back(e0, e1); }
  }
  // groups of n (view, segment) pairs (each mapped to its basic
  // view-segment), ascending, without repeats
  std::vector<int> request_groups(int n, const int32_t* views, const int32_t* segments) const {
    const mi_geometry& g = plan->geometry();
    std::vector<int> out;
    for (int i = 0; i < n; ++i) {
      int v = views[i], sg = segments[i];
      if (sg < g.min_segment_num || sg > g.max_segment_num || v < 0 || v >= g.num_views)
        throw ApiError(MI_ERR_ARGUMENT, "view or segment out of range");
      plan->basic_view_segment(v, sg);
      auto it = vs_group.find(static_cast<int64_t>(sg + 512) * 65536 + v);
      if (it == vs_group.end()) throw ApiError(MI_ERR_INTERNAL, "basic view-segment without forward tasks");
      out.push_back(it->second);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }
  // one viewgram (segment, view, TOF bin k) between the device layout and
  // STIR's layout on the host: nax rows of ntang values, row pitch
  // views * ntang; staged contiguously at stage_sino + slot
  void viewgram_to_host(const float* d_proj, int seg, int view, int k, float* host, int64_t slot, hipStream_t st) {
    const mi_geometry& g = plan->geometry();
    const int nt = g.num_tangential_poss, nax = plan->num_axial(seg);
    const int64_t blk = k * block_bins + plan->segment_offset(seg);
    mi::launch_transpose(d_proj + blk + static_cast<int64_t>(view) * nt * nax, stage_sino.p + slot, nt, nax, nax, nt, st);
    HIPCHECK(hipMemcpy2DAsync(host + blk + static_cast<int64_t>(view) * nt, sizeof(float) * g.num_views * nt,
                              stage_sino.p + slot, sizeof(float) * nt, sizeof(float) * nt, nax, hipMemcpyDeviceToHost,
                              st));
  }
  void viewgram_from_host(const float* host, int seg, int view, int k, float* d_proj, int64_t slot, hipStream_t st) {
    const mi_geometry& g = plan->geometry();
    const int nt = g.num_tangential_poss, nax = plan->num_axial(seg);
    const int64_t blk = k * block_bins + plan->segment_offset(seg);
    HIPCHECK(hipMemcpy2DAsync(stage_sino.p + slot, sizeof(float) * nt, host + blk + static_cast<int64_t>(view) * nt,
                              sizeof(float) * g.num_views * nt, sizeof(float) * nt, nax, hipMemcpyHostToDevice, st));
    mi::launch_transpose(stage_sino.p + slot, d_proj + blk + static_cast<int64_t>(view) * nt * nax, nax, nt, nt, nax, st);
  }
  DBuf<float> input_img;  // the image of mi_set_image (device layout)
  // nontof: back-project one non-TOF block with the non-TOF values of the
  // table (the sensitivity of TOF data, unless TOF sensitivities are asked for)
  void bp(mi::BpMode mode, const float* d_proj, float* d_img, int subset, const mi::BpData& data, hipStream_t st,
          bool nontof = false) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timing) { e0 = take_event(); e1 = take_event(); HIPCHECK(hipEventRecord(e0, st)); }
    if (ntof > 1 && !nontof && tof_bp_w > 0)
      mi::launch_back_project_tof_rows(mode, uniform_ppa, list_off.p, tof_bp_rec.p, tof_bp_w,
                                       static_cast<int64_t>(tof_bp_rec.n / mi::tof_bp_words(tof_bp_w)) - 1,
                                       subset_begin(subset), subset_end(subset), d_proj, d_img, dims, tofc, data,
                                       tof_bp_tuning, st);
    else if (ntof > 1 && !nontof)
      mi::launch_back_project_tof(mode, uniform_ppa, list_off.p, entries.p, entry_d.p, subset_begin(subset),
                                  subset_end(subset), d_proj, d_img, dims, tofc, data, tof_bp_tuning, st);
    else if (use_slab)
      mi::launch_back_project_slab(mode, slab, subset_begin(subset), subset_end(subset), d_proj, d_img, dims, data,
                                   st);
    else if (use_quad())
      mi::launch_back_project_quad(mode, quad_off.p, quad_ent.p, subset_begin(subset), subset_end(subset), d_proj,
                                   d_img, dims, data, bp_tuning, st, block_bins);
    else
      mi::launch_back_project(mode, uniform_ppa, list_off.p, entries.p, subset_begin(subset), subset_end(subset), d_proj,
                              d_img, dims, data, bp_tuning, st, block_bins);
    if (timing) { HIPCHECK(hipEventRecord(e1, st)); bp_events.emplace_back(e0, e1); }
  }

  // HBM bytes of one projection over all subsets when every re-read hits a
  // cache: the matrix data streamed once, the input read once, the output
  // written once (mi_info::fp_design_bytes / bp_design_bytes)
  void design_bytes(int64_t& fp_bytes, int64_t& bp_bytes) const {
    const int64_t img_bytes = int64_t{4} * dims.ncols * dims.nzp;
    const int64_t proj_bytes = int64_t{4} * total_bins();
    fp_bytes = img_bytes + proj_bytes;
    const bool fp_rec = ntof > 1 && tof_fp_w > 0;
    for (const mi::FpTask& t : h_tasks) {
      const mi::Family& f = h_fams[t.first];
      fp_bytes += (fp_rec ? int64_t{4} * mi::tof_fp_words(tof_fp_w) : int64_t{8}) * f.row_len +
                  int64_t(sizeof(mi::FpTask)) + int64_t(sizeof(mi::Family)) * (t.second >= 0 ? 2 : 1) +
                  (ntof > 1 && !fp_rec ? int64_t(sizeof(mi::FamilyLor)) : 0);
    }
    if (ntof > 1 && tof_bp_w > 0)
      bp_bytes = img_bytes + proj_bytes + static_cast<int64_t>(tof_bp_rec.bytes() + list_off.bytes());
    else if (use_quad())
      bp_bytes = img_bytes + proj_bytes + static_cast<int64_t>(quad_ent.bytes() + quad_off.bytes());
    else
      bp_bytes = img_bytes + proj_bytes +
                 static_cast<int64_t>(entries.bytes() + list_off.bytes() + entry_d.bytes() + slab_hdr.bytes() +
                                      slab_ent.bytes());
  }
  void set_up(const mi_geometry& g, int nsub);
  void build_rows();
  void build_tasks();
  void build_table();
  bool build_slab_table();
  void build_tof_rows(int64_t n);
  mi::FpData fp_data() const {
    mi::FpData d{};
    d.measured = y.p;
    d.additive = add.p;
    d.additive_constant = add_const;
    d.vg_small = vg_small.p;
    d.mult = mult.p;
    d.partial = partial.p;
    return d;
  }
};

namespace {
// compute_weights (QuadraticPrior.cxx:202-226): vox_x / |offset|, the squared
// length summed in float and the quotient formed in double (the reference's
// unqualified sqrt resolves to the C library's double overload there).
void default_prior_weights(float vox_z, float vox_y, float vox_x, float* w27) {
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        float& w = w27[((z + 1) * 3 + (y + 1)) * 3 + (x + 1)];
        if (z == 0 && y == 0 && x == 0) {
          w = 0;
        } else {
          const float sx = static_cast<float>(x) * vox_x, sy = static_cast<float>(y) * vox_y;
          const float sz = static_cast<float>(z) * vox_z;
          w = static_cast<float>(static_cast<double>(vox_x) / std::sqrt(static_cast<double>(sx * sx + sy * sy + sz * sz)));
        }
      }
}

void default_prior_weights(const mi_geometry& g, float* w27) { default_prior_weights(g.vox_z, g.vox_y, g.vox_x, w27); }

int64_t fam_key(int seg, int view, int tang, int a0) {
  return (((static_cast<int64_t>(seg + 512) * 4096 + view) * 8192) + (tang + 4096)) * 8192 + a0;
}
}  // namespace

void mi_ctx::build_rows() {
  const auto& rays = plan->basic_rays();
  const int nb = static_cast<int>(rays.size());
  DBuf<mi::BasicRay> d_rays;
  d_rays.alloc(nb);
  HIPCHECK(hipMemcpyAsync(d_rays.p, rays.data(), nb * sizeof(mi::BasicRay), hipMemcpyHostToDevice, stream));
  DBuf<int> d_len, d_buf, d_flag;
  d_len.alloc(nb);
  d_buf.alloc(nb);
  d_flag.alloc(2);
  std::vector<int> h_len(nb), h_buf(nb);
  int cap = plan->row_length_bound();
  // allowed in-plane coordinates of basic-row elements: with any in-plane
  // symmetry the element's images (x,y negated or swapped) must stay inside
  // the image, so the range is the largest centred square that fits
  int4 lim;
  {
    const mi_geometry& g = plan->geometry();
    int32_t fl[5];
    plan->symmetry_flags(fl);
    const int max_x = g.min_x + g.nx - 1, max_y = g.min_y + g.ny - 1;
    if (fl[0] || fl[1] || fl[3]) {
      const int m = std::min(std::min(max_x, -g.min_x), std::min(max_y, -g.min_y));
      lim = make_int4(-m, m, -m, m);
    } else {
      lim = make_int4(g.min_x, max_x, g.min_y, max_y);
    }
    if (lim.x < -512 || lim.y > 511 || lim.z < -512 || lim.w > 511)
      throw ApiError(MI_ERR_ARGUMENT, "image x/y index range exceeds the packed element range");
  }
  const size_t budget = size_t(2) << 30;  // scratch bytes per batch
  struct Chunk {
    int64_t global_start;
    int64_t size;
    std::unique_ptr<DBuf<int2>> data;
  };
  std::vector<Chunk> chunks;
  std::vector<int64_t> row_start(nb);
  int64_t total = 0;
  DBuf<uint64_t> sk;
  DBuf<float> sw;
  DBuf<int64_t> d_start;
  d_start.alloc(nb);
  int first = 0;
  while (first < nb) {
    int batch = static_cast<int>(std::max<size_t>(64, budget / (size_t(3) * cap * (sizeof(uint64_t) + sizeof(float)))));
    batch = std::min(batch, nb - first);
    if (sk.n < size_t(3) * cap * batch) {
      sk.alloc(size_t(3) * cap * batch);
      sw.alloc(size_t(3) * cap * batch);
    }
    HIPCHECK(hipMemsetAsync(d_flag.p, 0, 2 * sizeof(int), stream));
    mi::launch_trace_basic_rows(d_rays.p, first, batch, plan->trace_consts(), sk.p, sw.p, cap, d_len.p, d_buf.p, d_flag.p,
                                stream);
    HIPCHECK(hipGetLastError());
    int flag[2];
    HIPCHECK(hipMemcpyAsync(flag, d_flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(h_len.data() + first, d_len.p + first, batch * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (flag[0]) {  // a row exceeded the scratch bound: retry the batch with more room
      if (cap > (1 << 22)) throw ApiError(MI_ERR_INTERNAL, "row length bound exceeded");
      cap *= 4;
      continue;
    }
    int64_t batch_total = 0;
    for (int i = first; i < first + batch; ++i) {
      row_start[i] = batch_total;
      batch_total += h_len[i];
    }
    Chunk ch{total, batch_total, std::make_unique<DBuf<int2>>()};
    ch.data->alloc(std::max<int64_t>(batch_total, 1));
    HIPCHECK(hipMemcpyAsync(d_start.p + first, row_start.data() + first, batch * sizeof(int64_t), hipMemcpyHostToDevice,
                            stream));
    mi::launch_pack_rows(first, batch, sk.p, sw.p, cap, d_len.p, d_buf.p, d_start.p, ch.data->p, d_flag.p + 1, lim, stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(flag, d_flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (flag[1]) throw ApiError(MI_ERR_ARGUMENT, "matrix element coordinates exceed the packed range");
    for (int i = first; i < first + batch; ++i) row_start[i] += total;
    total += batch_total;
    chunks.push_back(std::move(ch));
    first += batch;
  }
  sk.release();
  sw.release();
  if (total >= (int64_t(1) << 31)) throw ApiError(MI_ERR_ARGUMENT, "more than 2^31 basic-row elements");
  pool.alloc(std::max<int64_t>(total, 1));
  for (auto& ch : chunks)
    if (ch.size > 0)
      HIPCHECK(hipMemcpyAsync(pool.p + ch.global_start, ch.data->p, ch.size * sizeof(int2), hipMemcpyDeviceToDevice, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  num_basic_elements = total;
  // the rows once per in-plane operation for the (non-TOF) forward projection
  pool_ops.release();
  if (ntof == 1 && total > 0) {
    // 8 copies, then slack for the forward kernel's read-ahead of one element group
    pool_ops.alloc(8 * total + kRowSlack);
    HIPCHECK(hipMemsetAsync(pool_ops.p + 8 * total, 0, kRowSlack * sizeof(int2), stream));
    mi::launch_expand_rows(pool.p, total, dims, pool_ops.p, stream);
    HIPCHECK(hipGetLastError());
  }
  // families
  h_fams = plan->families();
  const auto& keys = plan->family_keys();
  num_matrix_elements = 0;
  fam_lookup.clear();
  fam_lookup.reserve(h_fams.size() * 2);
  for (size_t i = 0; i < h_fams.size(); ++i) {
    const int bi = keys[i].basic_index;
    h_fams[i].row_start = static_cast<int32_t>(row_start[bi]);
    h_fams[i].row_len = h_len[bi];
    num_matrix_elements += static_cast<int64_t>(h_len[bi]) * h_fams[i].nax;
    fam_lookup[fam_key(keys[i].seg, keys[i].view, keys[i].tang, keys[i].a0)] = static_cast<int>(i);
  }
  uniform_ppa = h_fams.empty() ? 0 : h_fams[0].ppa;
  for (const auto& f : h_fams)
    if (f.ppa != uniform_ppa) uniform_ppa = 0;
  fams.alloc(h_fams.size());
  HIPCHECK(hipMemcpyAsync(fams.p, h_fams.data(), h_fams.size() * sizeof(mi::Family), hipMemcpyHostToDevice, stream));
  {
    DBuf<unsigned long long> d_nnz;
    d_nnz.alloc(1);
    HIPCHECK(hipMemsetAsync(d_nnz.p, 0, sizeof(unsigned long long), stream));
    mi::launch_count_nonzeros(fams.p, static_cast<int>(h_fams.size()), pool.p, dims, d_nnz.p, stream);
    HIPCHECK(hipGetLastError());
    unsigned long long h = 0;
    HIPCHECK(hipMemcpyAsync(&h, d_nnz.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    num_nonzeros = static_cast<int64_t>(h);
  }
  HIPCHECK(hipStreamSynchronize(stream));
}

void mi_ctx::build_tasks() {
  // One task per family, or per (+segment, -segment) pair of families that
  // share the basic row and in-plane operation (the swap-segment symmetry);
  // the pair's bins together fit the 4 x 64 lanes of a wave.
  const int nsub = plan->num_subsets();
  const auto& keys = plan->family_keys();
  std::vector<mi::FpTask> h;
  std::vector<char> taken(h_fams.size(), 0);
  task_off.assign(nsub + 1, 0);
  for (int s = 0; s < nsub; ++s) {
    task_off[s] = static_cast<int>(h.size());
    for (int f = plan->subset_family_begin(s); f < plan->subset_family_begin(s + 1); ++f) {
      if (taken[f]) continue;
      taken[f] = 1;
      if (h_fams[f].nax > 256) throw ApiError(MI_ERR_ARGUMENT, "more than 256 axial positions in a segment");
      int partner = -1;
      if (ntof == 1 && keys[f].seg > 0) {
        auto it = fam_lookup.find(fam_key(-keys[f].seg, keys[f].view, keys[f].tang, keys[f].a0));
        if (it != fam_lookup.end()) {
          const int g = it->second;
          const mi::Family& A = h_fams[f];
          const mi::Family& B = h_fams[g];
          if (!taken[g] && g >= plan->subset_family_begin(s) && g < plan->subset_family_begin(s + 1) &&
              A.row_start == B.row_start && A.row_len == B.row_len && A.inplane == B.inplane && A.nax + B.nax <= 256)
            partner = g;
        }
      }
      if (partner >= 0) taken[partner] = 1;
      h.push_back({f, partner});
    }
    // launch order within the subset: by (view, tangential position), then
    // segment. The tasks of one (view, tang) follow the same path through the
    // image columns (only their z differs), so running them together keeps
    // those columns in L2; each task writes its own bins, so the order does
    // not change any projected value.
    std::stable_sort(h.begin() + task_off[s], h.end(), [&](const mi::FpTask& a, const mi::FpTask& b) {
      const auto& ka = keys[a.first];
      const auto& kb = keys[b.first];
      if (ka.view != kb.view) return ka.view < kb.view;
      if (ka.tang != kb.tang) return ka.tang < kb.tang;
      if (ka.seg != kb.seg) return ka.seg < kb.seg;
      return ka.a0 < kb.a0;
    });
  }
  task_off[nsub] = static_cast<int>(h.size());
  // the basic view-segment group of every task (a task's families are
  // related view-segments of one basic view-segment)
  vs_group.clear();
  group_vs.clear();
  task_group.assign(h.size(), -1);
  subset_groups.assign(nsub, {});
  auto group_of = [&](int fam) {
    int v = keys[fam].view, sg = keys[fam].seg;
    plan->basic_view_segment(v, sg);
    const int64_t key = static_cast<int64_t>(sg + 512) * 65536 + v;
    auto it = vs_group.emplace(key, static_cast<int>(group_vs.size()));
    if (it.second) group_vs.push_back(plan->related_view_segments(v, sg));
    return it.first->second;
  };
  for (int s = 0; s < nsub; ++s)
    for (int t = task_off[s]; t < task_off[s + 1]; ++t) {
      const int g = group_of(h[t].first);
      if (h[t].second >= 0 && group_of(h[t].second) != g)
        throw ApiError(MI_ERR_INTERNAL, "forward task across basic view-segments");
      task_group[t] = g;
      subset_groups[s].push_back(g);
    }
  for (auto& v : subset_groups) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  h_tasks = h;
  tasks.alloc(h.size());
  HIPCHECK(hipMemcpyAsync(tasks.p, h.data(), h.size() * sizeof(mi::FpTask), hipMemcpyHostToDevice, stream));
  partial.alloc(std::max<size_t>(h.size(), 1));
  HIPCHECK(hipStreamSynchronize(stream));
}

void mi_ctx::build_table() {
  release_slab();
  if (build_slab_table()) {
    use_slab = true;
    return;
  }
  release_slab();
  const int nsub = plan->num_subsets();
  const int ncols = dims.ncols;
  // entry offsets per family (family order == subset order)
  std::vector<int64_t> fam_off(h_fams.size() + 1, 0);
  for (size_t i = 0; i < h_fams.size(); ++i) fam_off[i + 1] = fam_off[i] + h_fams[i].row_len;
  const int64_t n = fam_off.back();
  if (n >= (int64_t(1) << 31)) throw ApiError(MI_ERR_ARGUMENT, "back-projection table above 2^31 entries");
  const int64_t nkeys = static_cast<int64_t>(nsub) * ncols;
  DBuf<int32_t> counts;
  counts.alloc(nkeys);
  HIPCHECK(hipMemsetAsync(counts.p, 0, nkeys * sizeof(int32_t), stream));
  DBuf<int64_t> d_famoff;
  d_famoff.alloc(fam_off.size());
  HIPCHECK(hipMemcpyAsync(d_famoff.p, fam_off.data(), fam_off.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
  DBuf<uint32_t> keys_in, keys_out, vals_in, vals_out;
  DBuf<mi::BpEntry> raw;
  keys_in.alloc(std::max<int64_t>(n, 1));
  raw.alloc(std::max<int64_t>(n, 1));
  DBuf<float> raw_d;
  if (ntof > 1) raw_d.alloc(std::max<int64_t>(n, 1));
  for (int s = 0; s < nsub; ++s) {
    const int fb = plan->subset_family_begin(s), fe = plan->subset_family_begin(s + 1);
    mi::launch_count_entries(fams.p, fb, fe, pool.p, s, dims, counts.p, stream);
    mi::launch_make_entries(fams.p, fb, fe, d_famoff.p + fb, pool.p, s, dims, keys_in.p, raw.p, stream);
    if (ntof > 1)
      mi::launch_entry_tof_distance(fams.p, fb, fe, d_famoff.p + fb, lors.p, pool.p, dims, tofc, raw_d.p, stream);
  }
  HIPCHECK(hipGetLastError());
  // list offsets = exclusive scan of counts, plus the total at the end
  list_off.alloc(nkeys + 1);
  {
    size_t tb = 0;
    HIPCHECK(mi::exclusive_scan_counts(nullptr, tb, counts.p, list_off.p, static_cast<int>(nkeys), stream));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tb, 1));
    HIPCHECK(mi::exclusive_scan_counts(tmp.p, tb, counts.p, list_off.p, static_cast<int>(nkeys), stream));
    HIPCHECK(hipMemcpyAsync(list_off.p + nkeys, &n, sizeof(int64_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  counts.release();
  entries.alloc(std::max<int64_t>(n, 1));
  if (n > 0) {
    keys_out.alloc(n);
    vals_in.alloc(n);
    vals_out.alloc(n);
    mi::launch_iota(vals_in.p, n, stream);
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < nkeys) ++end_bit;
    size_t tb = 0;
    HIPCHECK(mi::radix_sort_pairs(nullptr, tb, keys_in.p, keys_out.p, vals_in.p, vals_out.p, static_cast<int>(n), end_bit,
                                  stream));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tb, 1));
    HIPCHECK(mi::radix_sort_pairs(tmp.p, tb, keys_in.p, keys_out.p, vals_in.p, vals_out.p, static_cast<int>(n), end_bit,
                                  stream));
    mi::launch_gather_entries(raw.p, vals_out.p, entries.p, n, stream);
    if (ntof > 1) {
      entry_d.alloc(n);
      mi::launch_gather_floats(raw_d.p, vals_out.p, entry_d.p, n, stream);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));
    tmp.release();
    raw.release();
    raw_d.release();
    keys_in.release();
    keys_out.release();
    vals_in.release();
    build_quad_table(vals_out.p, n, d_famoff.p);
  }
  build_tof_rows(n);
}

// The 2 x 2 column-block table (pmrt_quad.hip) of a one-plane table without
// TOF, from the sorted column table (d_perm: each entry's generation index),
// when it fits in three quarters of the free device memory; otherwise, or if
// the merge finds a column whose element order it cannot keep, the column
// kernel stays in use.
void mi_ctx::build_quad_table(const uint32_t* d_perm, int64_t n, const int64_t* d_famoff) {
  quad_off.release();
  quad_ent.release();
  if (!quad_enable || ntof != 1 || uniform_ppa != 1 || n <= 0 || h_fams.empty()) return;
  const int nsub = plan->num_subsets();
  const int nqx = (dims.nx + 1) / 2, nqy = (dims.ny + 1) / 2, nq = nqx * nqy;
  const int64_t nlists = static_cast<int64_t>(nsub) * nq;
  if (nlists >= (int64_t{1} << 31) || h_fams.size() >= (size_t{1} << 40)) return;
  size_t free_b = 0, total_b = 0;
  HIPCHECK(hipMemGetInfo(&free_b, &total_b));
  // keys, merged entries (at most one per column entry), counts and offsets
  const int64_t need = 8 * n + int64_t(sizeof(mi::QuadEntry)) * n + 12 * nlists;
  if (need > static_cast<int64_t>(free_b / 4 * 3)) return;
  DBuf<uint64_t> key;
  key.alloc(n);
  mi::launch_quad_keys(d_perm, n, d_famoff, static_cast<int>(h_fams.size()), fams.p, pool.p, key.p, stream);
  DBuf<int32_t> counts;
  counts.alloc(nlists);
  DBuf<int> bad;
  bad.alloc(1);
  HIPCHECK(hipMemsetAsync(bad.p, 0, sizeof(int), stream));
  mi::launch_quad_merge(true, list_off.p, entries.p, key.p, dims, nqx, nq, nsub, nullptr, counts.p, nullptr, bad.p,
                        stream);
  quad_off.alloc(nlists + 1);
  {
    size_t tb = 0;
    HIPCHECK(mi::exclusive_scan_counts(nullptr, tb, counts.p, quad_off.p, static_cast<int>(nlists), stream));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tb, 1));
    HIPCHECK(mi::exclusive_scan_counts(tmp.p, tb, counts.p, quad_off.p, static_cast<int>(nlists), stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  int64_t last_off = 0;
  int32_t last_cnt = 0;
  HIPCHECK(hipMemcpy(&last_off, quad_off.p + nlists - 1, sizeof(int64_t), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(&last_cnt, counts.p + nlists - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
  const int64_t total = last_off + last_cnt;
  HIPCHECK(hipMemcpy(quad_off.p + nlists, &total, sizeof(int64_t), hipMemcpyHostToDevice));
  counts.release();
  quad_ent.alloc(std::max<int64_t>(total, 1));
  mi::launch_quad_merge(false, list_off.p, entries.p, key.p, dims, nqx, nq, nsub, quad_off.p, nullptr, quad_ent.p,
                        bad.p, stream);
  HIPCHECK(hipGetLastError());
  int h_bad = 0;
  HIPCHECK(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  if (h_bad) {
    quad_off.release();
    quad_ent.release();
  }
  quad_entries = total;
}

// Stored TOF rows when every family has a row of its own, the planes per
// axial position fit the record's 4 bits, the TOF data hold fewer than 2^31
// bins, the widest live range fits a
// window and the records fit in three quarters of the free device memory;
// otherwise the on-the-fly kernels stay in use.
void mi_ctx::build_tof_rows(int64_t n) {
  tof_fp_rec.release();
  tof_bp_rec.release();
  tof_fp_nt = tof_fp_w = tof_bp_w = tof_live_span = 0;
  if (ntof <= 1 || tof_rows_mode == 0 || h_fams.empty() || n <= 0) return;
  std::vector<std::pair<int64_t, int64_t>> rows;
  rows.reserve(h_fams.size());
  int64_t pool_end = 0;
  if (total_bins() >= (int64_t{1} << 31)) return;  // back records hold a bin index in 31 bits
  for (const mi::Family& f : h_fams) {
    if (f.ppa < 1 || f.ppa > 15) return;
    rows.emplace_back(f.row_start, f.row_len);
    pool_end = std::max<int64_t>(pool_end, int64_t{f.row_start} + f.row_len);
  }
  std::sort(rows.begin(), rows.end());
  for (size_t i = 1; i < rows.size(); ++i)
    if (rows[i - 1].first + rows[i - 1].second > rows[i].first) return;
  DBuf<int> span;
  span.alloc(1);
  HIPCHECK(hipMemsetAsync(span.p, 0, sizeof(int), stream));
  mi::launch_tof_live_span(entries.p, entry_d.p, n, tofc, span.p, stream);
  int s = 0;
  HIPCHECK(hipMemcpyAsync(&s, span.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  tof_live_span = s;
  int fnt, fw, bw;
  if (tof_rows_mode == 1 && ntof <= 12 && s <= mi::kTofFpNarrow) {
    fnt = 12;
    fw = mi::kTofFpNarrow;
  } else if (s <= mi::kTofFpWide) {
    fnt = 16;
    fw = mi::kTofFpWide;
  } else {
    return;
  }
  if (tof_rows_mode == 1 && s <= mi::kTofBpNarrow) bw = mi::kTofBpNarrow;
  else if (s <= mi::kTofBpWide) bw = mi::kTofBpWide;
  else return;
  // the records only where they fit beside everything else (a quarter of the
  // free memory is left for the data, images and sensitivities set up later)
  const size_t need = sizeof(int32_t) * (static_cast<size_t>(pool_end) * mi::tof_fp_words(fw) +
                                         static_cast<size_t>(n) * mi::tof_bp_words(bw));
  size_t free_b = 0, total_b = 0;
  HIPCHECK(hipMemGetInfo(&free_b, &total_b));
  if (need > free_b / 4 * 3) return;
  tof_fp_rec.alloc(pool_end * mi::tof_fp_words(fw));
  mi::launch_tof_fp_records(fams.p, static_cast<int>(h_fams.size()), lors.p, pool.p, dims, tofc, fnt, fw,
                            tof_fp_rec.p, stream);
  tof_bp_rec.alloc((n + 1) * mi::tof_bp_words(bw));  // + the all-zero record read past a list's end
  HIPCHECK(hipMemsetAsync(tof_bp_rec.p + n * mi::tof_bp_words(bw), 0, sizeof(int32_t) * mi::tof_bp_words(bw), stream));
  mi::launch_tof_bp_records(entries.p, entry_d.p, n, tofc, bw, tof_bp_rec.p, stream);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(stream));
  tof_fp_nt = fnt;
  tof_fp_w = fw;
  tof_bp_w = bw;
  entry_d.release();  // read only by the on-the-fly back projection
}

// The slab table (pmrt_slab.hip): for one plane per axial position, families
// that cover whole runs, images of at most 128 planes and slabs that fit the
// LDS of a workgroup; false when the geometry does not qualify (the column
// table is built instead).
bool mi_ctx::build_slab_table() {
  const mi_geometry& g = plan->geometry();
  if (!slab_allowed || !slab_enable || ntof > 1 || uniform_ppa != 1 || dims.nzp > 2 * 64 || h_fams.empty())
    return false;
  const auto& keys = plan->family_keys();
  for (size_t i = 0; i < h_fams.size(); ++i)
    if (keys[i].a0 != 0 || h_fams[i].nax != plan->num_axial(keys[i].seg) || h_fams[i].nax > 2 * 64) return false;
  const int nsub = plan->num_subsets();
  const int tx = slab_tx, ty = slab_ty, TC = tx * ty, cpw = slab_cpw;
  if (tx < 1 || ty < 1 || TC % cpw != 0 || TC / cpw > 16) return false;
  const int nwaves = TC / cpw;
  const int ntx = (g.nx + tx - 1) / tx, nty = (g.ny + ty - 1) / ty, ntiles = ntx * nty;
  // view slots: the distinct views of each subset's families, ascending
  std::vector<std::vector<int>> views(nsub);
  for (int s = 0; s < nsub; ++s) {
    for (int f = plan->subset_family_begin(s); f < plan->subset_family_begin(s + 1); ++f) views[s].push_back(keys[f].view);
    std::sort(views[s].begin(), views[s].end());
    views[s].erase(std::unique(views[s].begin(), views[s].end()), views[s].end());
  }
  int nv_max = 1;
  for (auto& v : views) nv_max = std::max<int>(nv_max, static_cast<int>(v.size()));
  // segments in storage order
  const std::vector<int>& seq = plan->segment_sequence();
  const int nseg = static_cast<int>(seq.size());
  std::vector<int4> hsegs(nseg);
  std::unordered_map<int, int> seg_slot;
  for (int j = 0; j < nseg; ++j) {
    hsegs[j] = make_int4(static_cast<int>(plan->segment_offset(seq[j])), plan->num_axial(seq[j]), 0, 0);
    seg_slot[seq[j]] = j;
  }
  // per family: view slot, tangential index, segment slot
  std::vector<int4> hfs(h_fams.size());
  for (int s = 0; s < nsub; ++s)
    for (int f = plan->subset_family_begin(s); f < plan->subset_family_begin(s + 1); ++f) {
      const int vs = static_cast<int>(std::lower_bound(views[s].begin(), views[s].end(), keys[f].view) - views[s].begin());
      hfs[f] = make_int4(vs, keys[f].tang - g.min_tangential_pos_num, seg_slot[keys[f].seg], 0);
    }
  std::vector<int32_t> hvo(static_cast<size_t>(nsub) * nv_max, 0), hnv(nsub);
  for (int s = 0; s < nsub; ++s) {
    hnv[s] = static_cast<int>(views[s].size());
    for (size_t v = 0; v < views[s].size(); ++v) hvo[static_cast<size_t>(s) * nv_max + v] = views[s][v];
  }
  DBuf<int4> d_fs;
  d_fs.alloc(hfs.size());
  HIPCHECK(hipMemcpyAsync(d_fs.p, hfs.data(), hfs.size() * sizeof(int4), hipMemcpyHostToDevice, stream));
  slab_view_of.alloc(hvo.size());
  HIPCHECK(hipMemcpyAsync(slab_view_of.p, hvo.data(), hvo.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  slab_nviews.alloc(nsub);
  HIPCHECK(hipMemcpyAsync(slab_nviews.p, hnv.data(), nsub * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  mi::SlabArgs A{};
  A.tx = tx; A.ty = ty; A.ntx = ntx; A.ntiles = ntiles; A.nv_max = nv_max; A.nseg = nseg;
  A.ntang = g.num_tangential_poss;
  A.cpw = cpw; A.nwaves = nwaves;
  A.view_of = slab_view_of.p;
  A.nviews = slab_nviews.p;
  // pass 1: tangential ranges per (subset, tile, view slot)
  const int64_t nrng = static_cast<int64_t>(nsub) * ntiles * nv_max;
  slab_rng.alloc(nrng);
  mi::launch_fill_int2(slab_rng.p, nrng, make_int2(0x7fffffff, -1), stream);
  A.rng = slab_rng.p;
  for (int s = 0; s < nsub; ++s)
    mi::launch_slab_range(fams.p, d_fs.p, plan->subset_family_begin(s), plan->subset_family_begin(s + 1), pool.p, s,
                          dims, A, slab_rng.p, stream);
  HIPCHECK(hipGetLastError());
  std::vector<int2> hr(nrng);
  HIPCHECK(hipMemcpyAsync(hr.data(), slab_rng.p, nrng * sizeof(int2), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  int tmax = 0;
  for (const int2& r : hr) tmax = std::max(tmax, r.y - r.x + 1);
  if (tmax <= 0) return false;
  // segment groups: consecutive segments of the storage order, as few as
  // keep every slab within the LDS budget at the widest tangential range
  auto group_floats = [&](int nax_sum, int n) {
    return int64_t{2} * mi::kSlabLead + int64_t{n} * mi::kSlabPad + int64_t{tmax} * (nax_sum + int64_t{n} * mi::kSlabPad);
  };
  int64_t budget = std::min<int64_t>(int64_t{slab_lds_kb} * 1024, mi::kSlabMaxBytes) / 4;
  for (int j = 0; j < nseg; ++j)
    if (group_floats(hsegs[j].y, 1) > budget) budget = mi::kSlabMaxBytes / 4;
  std::vector<int2> hgroups, hpos(nseg);
  int64_t slab_floats = 0;
  for (int j = 0; j < nseg;) {
    int e = j, nax_sum = 0;
    while (e < nseg && group_floats(nax_sum + hsegs[e].y, e - j + 1) <= budget) nax_sum += hsegs[e++].y;
    if (e == j) return false;  // one segment's slab does not fit the LDS
    int cum = 0;
    for (int k = j; k < e; ++k) {
      hsegs[k].w = static_cast<int>(hgroups.size());
      hpos[k] = make_int2(cum, k - j);
      cum += hsegs[k].y;
    }
    slab_floats = std::max(slab_floats, group_floats(nax_sum, e - j));
    hgroups.push_back(make_int2(j, e));
    j = e;
  }
  const int ngroups = static_cast<int>(hgroups.size());
  slab_segs.alloc(nseg);
  HIPCHECK(hipMemcpyAsync(slab_segs.p, hsegs.data(), nseg * sizeof(int4), hipMemcpyHostToDevice, stream));
  slab_seg_pos.alloc(nseg);
  HIPCHECK(hipMemcpyAsync(slab_seg_pos.p, hpos.data(), nseg * sizeof(int2), hipMemcpyHostToDevice, stream));
  slab_groups.alloc(ngroups);
  HIPCHECK(hipMemcpyAsync(slab_groups.p, hgroups.data(), ngroups * sizeof(int2), hipMemcpyHostToDevice, stream));
  A.segs = slab_segs.p;
  A.seg_pos = slab_seg_pos.p;
  A.groups = slab_groups.p;
  A.ngroups = ngroups;
  // pass 2: terms per (subset, tile, phase, column)
  const int64_t nph_max = int64_t{nv_max} * ngroups;
  const int64_t nkeys = int64_t{nsub} * ntiles * nph_max * TC;
  if (nkeys >= 0x7fffffffLL) return false;
  DBuf<uint32_t> counts;  // one more zero at the end: the scan then ends with the total
  counts.alloc(nkeys + 1);
  HIPCHECK(hipMemsetAsync(counts.p, 0, (nkeys + 1) * sizeof(uint32_t), stream));
  for (int s = 0; s < nsub; ++s)
    mi::launch_slab_count(fams.p, d_fs.p, plan->subset_family_begin(s), plan->subset_family_begin(s + 1), pool.p, s,
                          dims, A, counts.p, stream);
  HIPCHECK(hipGetLastError());
  // block of every (subset, tile, phase, wave): steps = the most terms of its columns
  const int64_t nblocks = nkeys / TC * nwaves;
  std::vector<uint32_t> hsteps(nblocks);
  {
    DBuf<uint32_t> d_steps;
    d_steps.alloc(nblocks);
    mi::launch_slab_steps(counts.p, nblocks, cpw, TC, d_steps.p, stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hsteps.data(), d_steps.p, nblocks * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  std::vector<uint2> hh(nblocks);
  std::vector<int64_t> load(ntiles, 0);
  int64_t total = 0;
  for (int64_t b = 0; b < nblocks; ++b) {
    hh[b] = make_uint2(static_cast<uint32_t>(total), hsteps[b]);
    total += int64_t{hsteps[b]} * cpw;
    if (total >= (int64_t(1) << 31)) throw ApiError(MI_ERR_ARGUMENT, "back-projection table above 2^31 entries");
    load[(b / nwaves / nph_max) % ntiles] += int64_t{hsteps[b]} * cpw;
  }
  slab_hdr.alloc(nblocks);
  HIPCHECK(hipMemcpyAsync(slab_hdr.p, hh.data(), nblocks * sizeof(uint2), hipMemcpyHostToDevice, stream));
  // rank of every term within its column's list: the list's first position
  // in the sorted order (exclusive scan of the counts)
  DBuf<uint32_t> list_start;
  list_start.alloc(nkeys + 1);
  {
    size_t tb = 0;
    HIPCHECK(mi::exclusive_scan_u32(nullptr, tb, counts.p, list_start.p, static_cast<int>(nkeys + 1), stream));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tb, 1));
    HIPCHECK(mi::exclusive_scan_u32(tmp.p, tb, counts.p, list_start.p, static_cast<int>(nkeys + 1), stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  counts.release();
  // pass 3: every element's key and entry in family order, element order;
  // stable sort by key; each sorted element to its block slot
  std::vector<int64_t> fam_off(h_fams.size() + 1, 0);
  for (size_t i = 0; i < h_fams.size(); ++i) fam_off[i + 1] = fam_off[i] + h_fams[i].row_len;
  const int64_t n = fam_off.back();
  if (n >= (int64_t(1) << 31)) throw ApiError(MI_ERR_ARGUMENT, "more than 2^31 matrix elements");
  slab_ent.alloc(std::max<int64_t>(total, 1));
  mi::launch_slab_fill(slab_ent.p, static_cast<int64_t>(slab_ent.n), stream);
  if (n > 0) {
    DBuf<int64_t> d_famoff;
    d_famoff.alloc(fam_off.size());
    HIPCHECK(hipMemcpyAsync(d_famoff.p, fam_off.data(), fam_off.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    DBuf<uint32_t> keys_in, keys_out, vals_in, vals_out;
    DBuf<mi::SlabEntry> raw;
    keys_in.alloc(n);
    raw.alloc(n);
    A.hdr = slab_hdr.p;
    for (int s = 0; s < nsub; ++s) {
      const int fb = plan->subset_family_begin(s), fe = plan->subset_family_begin(s + 1);
      mi::launch_slab_make(fams.p, d_fs.p, fb, fe, d_famoff.p + fb, pool.p, s, dims, A, keys_in.p, raw.p, stream);
    }
    HIPCHECK(hipGetLastError());
    keys_out.alloc(n);
    vals_in.alloc(n);
    vals_out.alloc(n);
    mi::launch_iota(vals_in.p, n, stream);
    size_t tb = 0;
    HIPCHECK(mi::radix_sort_pairs(nullptr, tb, keys_in.p, keys_out.p, vals_in.p, vals_out.p, static_cast<int>(n), 32,
                                  stream));
    DBuf<char> tmp;
    tmp.alloc(std::max<size_t>(tb, 1));
    HIPCHECK(mi::radix_sort_pairs(tmp.p, tb, keys_in.p, keys_out.p, vals_in.p, vals_out.p, static_cast<int>(n), 32,
                                  stream));
    // elements without an image voxel (key 0xffffffff) sort last: place the others
    uint32_t nvalid = 0;
    HIPCHECK(hipMemcpyAsync(&nvalid, list_start.p + nkeys, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (static_cast<int64_t>(nvalid) > n) throw ApiError(MI_ERR_INTERNAL, "slab table count");
    mi::launch_slab_place(keys_out.p, vals_out.p, raw.p, list_start.p, slab_hdr.p, nvalid, cpw, TC, slab_ent.p, stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));
  }
  // tiles with entries first, most entries first (the longest workgroups start first)
  std::vector<int32_t> order(ntiles);
  for (int t = 0; t < ntiles; ++t) order[t] = t;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return load[a] > load[b]; });
  int nbusy = 0;
  while (nbusy < ntiles && load[order[nbusy]] > 0) ++nbusy;
  slab_tiles.alloc(ntiles);
  HIPCHECK(hipMemcpyAsync(slab_tiles.p, order.data(), ntiles * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  A.hdr = slab_hdr.p;
  A.ent = slab_ent.p;
  A.tile_order = slab_tiles.p;
  A.nbusy = nbusy;
  A.slab_floats = static_cast<int32_t>(slab_floats);
  slab = A;
  return true;
}

void mi_ctx::set_up(const mi_geometry& g, int nsub) {
  HIPCHECK(hipSetDevice(device));
  // release everything of a previous set-up
  pool.release(); pool_ops.release(); fams.release(); tasks.release(); list_off.release(); entries.release();
  release_slab();
  img.release(); img2.release(); sino.release(); ratio.release(); stage_img.release(); stage_sino.release();
  y.release(); add.release(); mult.release(); ones.release(); vg_small.release(); x.release(); sens.clear();
  vg_off.release(); vg_len.release(); partial.release(); prior_partial.release();
  subset_vgs.release(); subset_vg_off.clear(); hess_flag.release(); y_raw.release(); add_raw.release();
  quad_off.release(); quad_ent.release();
  lors.release(); entry_d.release(); erf_tab.release(); sel_tasks.release();
  tof_fp_rec.release(); tof_bp_rec.release();
  tof_fp_nt = tof_fp_w = tof_bp_w = tof_live_span = 0;
  have_data = have_sens = have_x = have_image = false;
  plan.reset();
  plan = std::make_unique<mi::Plan>(g, nsub);
  // the back projection forms its gather offsets in 32 bits (bytes of one
  // sinogram's worth of bins)
  if (plan->num_bins() >= (int64_t{1} << 30)) {
    plan.reset();
    throw ApiError(MI_ERR_ARGUMENT, "more than 2^30 bins per TOF bin are not supported");
  }
  dims.nz = g.nz; dims.ny = g.ny; dims.nx = g.nx;
  dims.min_z = g.min_z; dims.min_y = g.min_y; dims.min_x = g.min_x;
  dims.nzp = plan->nz_padded();
  // the projections form image byte offsets in 32 bits (signed in the row
  // groups); the forward rows hold column indices in 20 bits
  if (static_cast<int64_t>(g.ny) * g.nx * dims.nzp >= (int64_t{1} << 29) ||
      static_cast<int64_t>(g.ny) * g.nx >= (int64_t{1} << 20)) {
    plan.reset();
    throw ApiError(MI_ERR_ARGUMENT, "images of 2^29 or more (z-padded) voxels or 2^20 or more columns are not supported");
  }
  dims.ncols = plan->num_columns();
  ntof = plan->num_tof_bins();
  block_bins = plan->num_bins();
  if (ntof > 1) {
    for (const mi::Family& f : plan->families()) {
      if (f.inplane != mi::IP_ID || f.zsign != 1) throw ApiError(MI_ERR_INTERNAL, "TOF family with an in-plane symmetry");
      if (f.nax > 128) throw ApiError(MI_ERR_ARGUMENT, "TOF data with more than 128 axial positions in a segment");
    }
    const auto& hl = plan->family_lors();
    lors.alloc(hl.size());
    HIPCHECK(hipMemcpyAsync(lors.p, hl.data(), hl.size() * sizeof(mi::FamilyLor), hipMemcpyHostToDevice, stream));
    // FastErf samples (FastErf.inl:30-40)
    std::vector<double> tab(mi::kErfSamples + 2);
    const double per = (2 * mi::kErfMax) / mi::kErfSamples;
    for (int i = 0; i < mi::kErfSamples + 2; ++i) tab[i] = std::erf(i * per - mi::kErfMax);
    erf_tab.alloc(tab.size());
    HIPCHECK(hipMemcpyAsync(erf_tab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    tofc = mi::TofConsts{};
    tofc.ntof = ntof;
    tofc.r = g.tof_r_sqrt2_gauss_sigma;
    tofc.z_off = plan->tof_z_offset();
    tofc.vz = g.vox_z; tofc.vy = g.vox_y; tofc.vx = g.vox_x;
    tofc.oz = g.origin_z; tofc.oy = g.origin_y; tofc.ox = g.origin_x;
    tofc.nvg = plan->num_viewgrams();
    tofc.block = block_bins;
    tofc.erf = erf_tab.p;
    for (int k = 0; k < ntof; ++k) { tofc.low[k] = g.tof_low[k]; tofc.high[k] = g.tof_high[k]; }
    tofc.shared_bounds = 1;
    for (int k = 0; k + 1 < ntof; ++k)
      if (g.tof_high[k] != g.tof_low[k + 1]) tofc.shared_bounds = 0;
    HIPCHECK(hipStreamSynchronize(stream));
  }
  build_rows();
  build_tasks();
  build_table();
  const size_t nvox_dev = static_cast<size_t>(dims.ncols) * dims.nzp;
  img.alloc(nvox_dev);
  img2.alloc(nvox_dev);
  HIPCHECK(hipMemsetAsync(img.p, 0, img.bytes(), stream));
  HIPCHECK(hipMemsetAsync(img2.p, 0, img2.bytes(), stream));
  sino.alloc(total_bins());
  ratio.alloc(total_bins());
  stage_img.alloc(plan->num_voxels());
  stage_sino.alloc(total_bins());
  prior_partial.alloc(std::max(mi::prior_num_partials(stir_prior_image()), mi::prior_num_partials(device_prior_image())));
  // viewgram extents in the device layout
  {
    // TOF viewgram (segment, view, TOF bin k) has index k * num_viewgrams + viewgram_index
    const int nvg1 = plan->num_viewgrams(), nvg = nvg1 * ntof;
    std::vector<int64_t> off(nvg, 0);
    std::vector<int32_t> len(nvg, 0);
    for (int k = 0; k < ntof; ++k)
      for (int seg = g.min_segment_num; seg <= g.max_segment_num; ++seg)
        for (int v = 0; v < g.num_views; ++v) {
          const int i = k * nvg1 + plan->viewgram_index(seg, v);
          off[i] = plan->device_bin_offset(seg, v, 0, g.min_tangential_pos_num) + k * block_bins;
          len[i] = g.num_tangential_poss * plan->num_axial(seg);
        }
    vg_off.alloc(nvg);
    vg_len.alloc(nvg);
    HIPCHECK(hipMemcpyAsync(vg_off.p, off.data(), nvg * sizeof(int64_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(vg_len.p, len.data(), nvg * sizeof(int32_t), hipMemcpyHostToDevice, stream));
  }
  HIPCHECK(hipStreamSynchronize(stream));
}

// ===========================================================================
extern "C" {

const char* mi_last_error(void) { return g_last_error.c_str(); }

int mi_ctx_create(int device, mi_ctx** out) {
  return guarded([&] {
    if (!out) throw ApiError(MI_ERR_ARGUMENT, "out is NULL");
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw ApiError(MI_ERR_ARGUMENT, "no such HIP device");
    auto ctx = std::make_unique<mi_ctx>();
    ctx->device = device;
    HIPCHECK(hipSetDevice(device));
    // a blocking stream: ordered with work on the HIP null stream, so host-buffer
    // calls and *_dev calls on the null stream see each other's results
    HIPCHECK(hipStreamCreate(&ctx->stream));
    *out = ctx.release();
  });
}

void mi_ctx_destroy(mi_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  hipStream_t s = ctx->stream;
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
}

int mi_set_up(mi_ctx* ctx, const mi_geometry* geometry, int num_subsets) {
  return guarded([&] {
    if (!ctx || !geometry) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->set_up(*geometry, num_subsets);
  });
}

int mi_get_info(mi_ctx* ctx, mi_info* out) {
  return guarded([&] {
    if (!ctx || !out) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    std::memset(out, 0, sizeof(*out));
    out->num_bins = ctx->total_bins();
    out->num_voxels = ctx->plan->num_voxels();
    out->num_basic_rows = static_cast<int64_t>(ctx->plan->basic_rays().size());
    out->num_basic_elements = ctx->num_basic_elements;
    out->num_families = static_cast<int64_t>(ctx->h_fams.size());
    out->num_matrix_elements = ctx->num_matrix_elements;
    out->num_table_entries = static_cast<int64_t>(ctx->use_slab ? ctx->slab_ent.n : ctx->entries.n);
    out->num_nonzeros = ctx->num_nonzeros;
    out->device_bytes = static_cast<int64_t>(ctx->pool.bytes() + ctx->pool_ops.bytes() + ctx->fams.bytes() + ctx->tasks.bytes() +
                                             ctx->list_off.bytes() + ctx->entries.bytes() + ctx->img.bytes() +
                                             ctx->img2.bytes() + ctx->sino.bytes() + ctx->ratio.bytes() +
                                             ctx->stage_img.bytes() + ctx->stage_sino.bytes() + ctx->y.bytes() +
                                             ctx->add.bytes() + ctx->mult.bytes() + ctx->ones.bytes() + ctx->x.bytes() +
                                             ctx->slab_hdr.bytes() + ctx->slab_ent.bytes() + ctx->slab_rng.bytes() +
                                             ctx->slab_tiles.bytes() + ctx->lors.bytes() + ctx->entry_d.bytes() +
                                             ctx->erf_tab.bytes() + ctx->tof_fp_rec.bytes() +
                                             ctx->tof_bp_rec.bytes());
    for (auto& s : ctx->sens) out->device_bytes += static_cast<int64_t>(s->bytes());
    out->num_subsets = ctx->plan->num_subsets();
    ctx->plan->symmetry_flags(out->symmetry_flags);
    out->nz_padded = ctx->dims.nzp;
    out->planes_per_axial_position = ctx->uniform_ppa;
    ctx->design_bytes(out->fp_design_bytes, out->bp_design_bytes);
    out->slab_back_projection = ctx->use_slab ? 1 : 0;
    if (ctx->use_slab) {
      out->slab_tile_x = ctx->slab.tx;
      out->slab_tile_y = ctx->slab.ty;
      out->slab_segment_groups = ctx->slab.ngroups;
      out->slab_lds_bytes = ctx->slab.slab_floats * 4;
    }
    out->tof_fp_window = ctx->tof_fp_w;
    out->tof_bp_window = ctx->tof_bp_w;
    out->tof_live_span = ctx->tof_live_span;
    out->bp_quad = ctx->use_quad() && !ctx->use_slab && ctx->ntof == 1 ? 1 : 0;
    out->num_quad_entries = ctx->quad_ent.p ? ctx->quad_entries : 0;
  });
}

int mi_subsets_balanced(mi_ctx* ctx, int* balanced) {
  return guarded([&] {
    if (!ctx || !balanced) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    *balanced = ctx->plan->subsets_balanced() ? 1 : 0;
  });
}

int mi_subset_view_segments(mi_ctx* ctx, int subset, int32_t* views, int32_t* segments, int cap, int* count) {
  return guarded([&] {
    if (!ctx || !count) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (subset < 0 || subset >= ctx->plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    const auto vs = ctx->plan->subset_view_segments(subset);
    *count = static_cast<int>(vs.size());
    for (int i = 0; i < cap && i < static_cast<int>(vs.size()); ++i) {
      if (views) views[i] = vs[i].first;
      if (segments) segments[i] = vs[i].second;
    }
  });
}

int mi_symmetry_operations(const mi_geometry* geometry, int n, const int32_t* bins, int32_t* basic, int32_t* ops) {
  return guarded([&] {
    if (!geometry || (n > 0 && (!bins || !basic || !ops))) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    mi::Plan plan(*geometry, 1);
    const mi_geometry& g = plan.geometry();
    for (int i = 0; i < n; ++i) {
      int seg = bins[4 * i], view = bins[4 * i + 1], ax = bins[4 * i + 2], tang = bins[4 * i + 3];
      if (seg < g.min_segment_num || seg > g.max_segment_num || view < 0 || view >= g.num_views)
        throw ApiError(MI_ERR_ARGUMENT, "bin out of range");
      const mi::SymOp op = plan.symmetry_op(seg, view, ax, tang);
      plan.basic_bin(seg, view, ax, tang);
      basic[4 * i] = seg; basic[4 * i + 1] = view; basic[4 * i + 2] = ax; basic[4 * i + 3] = tang;
      ops[3 * i] = op.inplane; ops[3 * i + 1] = op.zsign; ops[3 * i + 2] = op.zadd;
    }
  });
}

int mi_get_row(mi_ctx* ctx, int segment, int view, int axial_pos, int tang_pos, int16_t* z, int16_t* y, int16_t* x,
               float* w, int cap, int* n) {
  return guarded([&] {
    if (!ctx || !n) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    const mi_geometry& g = ctx->plan->geometry();
    if (segment < g.min_segment_num || segment > g.max_segment_num || view < 0 || view >= g.num_views ||
        tang_pos < g.min_tangential_pos_num || tang_pos > ctx->plan->max_tang() || axial_pos < 0 ||
        axial_pos >= ctx->plan->num_axial(segment))
      throw ApiError(MI_ERR_ARGUMENT, "bin out of range");
    int32_t flags[5];
    ctx->plan->symmetry_flags(flags);
    const int a0 = flags[4] ? 0 : axial_pos;
    auto it = ctx->fam_lookup.find(fam_key(segment, view, tang_pos, a0));
    if (it == ctx->fam_lookup.end()) throw ApiError(MI_ERR_INTERNAL, "family not found");
    const mi::Family& f = ctx->h_fams[it->second];
    *n = f.row_len;
    if (f.row_len == 0) return;
    std::vector<int2> row(f.row_len);
    HIPCHECK(hipMemcpyAsync(row.data(), ctx->pool.p + f.row_start, f.row_len * sizeof(int2), hipMemcpyDeviceToHost,
                            ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const int a = axial_pos - a0;
    for (int i = 0; i < f.row_len && i < cap; ++i) {
      const int c = row[i].x;
      const int bx = (c & 1023) - 512, by = ((c >> 10) & 1023) - 512;
      const int bz = static_cast<int>(static_cast<unsigned>(c) >> 20) - 2048;
      int xo = bx, yo = by;
      switch (f.inplane) {
        case mi::IP_ID: break;
        case mi::IP_XMX: xo = -bx; break;
        case mi::IP_YMY: yo = -by; break;
        case mi::IP_XMX_YMY: xo = -bx; yo = -by; break;
        case mi::IP_XY_YX: xo = by; yo = bx; break;
        case mi::IP_XMY_YX: xo = -by; yo = bx; break;
        case mi::IP_XY_YMX: xo = by; yo = -bx; break;
        default: xo = -by; yo = -bx; break;
      }
      if (z) z[i] = static_cast<int16_t>(f.zsign * bz + f.zadd0 + f.ppa * a);
      if (y) y[i] = static_cast<int16_t>(yo);
      if (x) x[i] = static_cast<int16_t>(xo);
      float wv;
      std::memcpy(&wv, &row[i].y, sizeof(float));
      if (w) w[i] = wv;
    }
  });
}

int mi_forward_project_dev(mi_ctx* ctx, const float* d_image, float* d_proj_data, int subset, void* stream) {
  return guarded([&] {
    if (!ctx || !d_image || !d_proj_data) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    ctx->image_to_device(d_image, ctx->img.p, st);
    if (subset >= 0) HIPCHECK(hipMemsetAsync(ctx->sino.p, 0, ctx->sino.bytes(), st));
    ctx->fp(mi::FP_PLAIN, ctx->img.p, ctx->sino.p, subset, mi::FpData{}, st);
    ctx->proj_from_device(ctx->sino.p, d_proj_data, st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_back_project_dev(mi_ctx* ctx, const float* d_proj_data, float* d_image, int subset, void* stream) {
  return guarded([&] {
    if (!ctx || !d_image || !d_proj_data) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    ctx->proj_to_device(d_proj_data, ctx->sino.p, st);
    ctx->bp(mi::BP_PLAIN, ctx->sino.p, ctx->img.p, subset, mi::BpData{}, st);
    ctx->image_from_device(ctx->img.p, d_image, st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_forward_project(mi_ctx* ctx, const float* image, float* proj_data, int subset) {
  return guarded([&] {
    if (!ctx || !image || !proj_data) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipMemcpyAsync(ctx->stage_img.p, image, ctx->stage_img.bytes(), hipMemcpyHostToDevice, ctx->stream));
    if (mi_forward_project_dev(ctx, ctx->stage_img.p, ctx->stage_sino.p, subset, ctx->stream) != MI_OK)
      throw ApiError(MI_ERR_DEVICE, g_last_error);
    HIPCHECK(hipMemcpyAsync(proj_data, ctx->stage_sino.p, ctx->stage_sino.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mi_back_project(mi_ctx* ctx, const float* proj_data, float* image, int subset) {
  return guarded([&] {
    if (!ctx || !image || !proj_data) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, proj_data, ctx->stage_sino.bytes(), hipMemcpyHostToDevice, ctx->stream));
    if (mi_back_project_dev(ctx, ctx->stage_sino.p, ctx->stage_img.p, subset, ctx->stream) != MI_OK)
      throw ApiError(MI_ERR_DEVICE, g_last_error);
    HIPCHECK(hipMemcpyAsync(image, ctx->stage_img.p, ctx->stage_img.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mi_basic_view_segments_in_subset(mi_ctx* ctx, int subset, int num_subsets, int32_t* views, int32_t* segments,
                                     int cap, int* count) {
  return guarded([&] {
    if (!ctx || !count) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (num_subsets < 1 || subset < 0 || subset >= num_subsets) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    const mi_geometry& g = ctx->plan->geometry();
    // find_basic_vs_nums_in_subset.cxx:31-69 (its TOF loop runs once)
    int n = 0;
    for (int seg = g.min_segment_num; seg <= g.max_segment_num; ++seg)
      for (int view = subset; view < g.num_views; view += num_subsets) {
        int v = view, sg = seg;
        if (ctx->plan->basic_view_segment(v, sg)) continue;
        if (n < cap) {
          if (views) views[n] = view;
          if (segments) segments[n] = seg;
        }
        ++n;
      }
    *count = n;
  });
}

int mi_set_image(mi_ctx* ctx, const float* image) {
  return guarded([&] {
    if (!ctx || !image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    if (ctx->input_img.n != ctx->img.n) ctx->input_img.alloc(ctx->img.n);
    HIPCHECK(hipMemcpyAsync(ctx->stage_img.p, image, ctx->stage_img.bytes(), hipMemcpyHostToDevice, ctx->stream));
    ctx->image_to_device(ctx->stage_img.p, ctx->input_img.p, ctx->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->have_image = true;
  });
}

int mi_forward_project_view_segments(mi_ctx* ctx, int n, const int32_t* views, const int32_t* segments,
                                     float* proj_data) {
  return guarded([&] {
    if (!ctx || !proj_data || (n > 0 && (!views || !segments))) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_image) throw ApiError(MI_ERR_STATE, "mi_set_image has not been called");
    HIPCHECK(hipSetDevice(ctx->device));
    const std::vector<int> groups = ctx->request_groups(n, views, segments);
    if (groups.empty()) return;
    hipStream_t st = ctx->stream;
    // the request's forward tasks, in the order of the table's subsets
    std::vector<char> want(ctx->group_vs.size(), 0);
    for (int gi : groups) want[gi] = 1;
    std::vector<mi::FpTask> sel;
    for (size_t t = 0; t < ctx->h_tasks.size(); ++t)
      if (want[ctx->task_group[t]]) sel.push_back(ctx->h_tasks[t]);
    if (ctx->sel_tasks.n < sel.size()) ctx->sel_tasks.alloc(sel.size());
    HIPCHECK(hipMemcpyAsync(ctx->sel_tasks.p, sel.data(), sel.size() * sizeof(mi::FpTask), hipMemcpyHostToDevice, st));
    ctx->fp_tasks(mi::FP_PLAIN, ctx->sel_tasks.p, static_cast<int>(sel.size()), ctx->input_img.p, ctx->sino.p,
                  mi::FpData{}, st);
    for (int gi : groups)
      for (const auto& vs : ctx->group_vs[gi])
        for (int k = 0; k < ctx->ntof; ++k)
          ctx->viewgram_to_host(ctx->sino.p, vs.second, vs.first, k, proj_data,
                                k * ctx->block_bins + ctx->plan->device_bin_offset(vs.second, vs.first, 0,
                                                                                   ctx->plan->geometry().min_tangential_pos_num),
                                st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}

int mi_back_project_view_segments(mi_ctx* ctx, int n, const int32_t* views, const int32_t* segments,
                                  const float* proj_data, float* image) {
  return guarded([&] {
    if (!ctx || !proj_data || !image || (n > 0 && (!views || !segments)))
      throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    const std::vector<int> groups = ctx->request_groups(n, views, segments);
    hipStream_t st = ctx->stream;
    // one subset of the table: its entries read only its own bins; otherwise
    // every other bin is zeroed and the whole table runs
    int subset = -1;
    for (int s = 0; s < ctx->plan->num_subsets(); ++s)
      if (ctx->subset_groups[s] == groups) subset = s;
    if (subset < 0) HIPCHECK(hipMemsetAsync(ctx->sino.p, 0, ctx->sino.bytes(), st));
    for (int gi : groups)
      for (const auto& vs : ctx->group_vs[gi])
        for (int k = 0; k < ctx->ntof; ++k)
          ctx->viewgram_from_host(proj_data, vs.second, vs.first, k, ctx->sino.p,
                                  k * ctx->block_bins + ctx->plan->device_bin_offset(vs.second, vs.first, 0,
                                                                                     ctx->plan->geometry().min_tangential_pos_num),
                                  st);
    if (groups.empty()) mi::launch_fill(ctx->img.p, static_cast<int64_t>(ctx->img.n), 0.F, st);
    else ctx->bp(mi::BP_PLAIN, ctx->sino.p, ctx->img.p, subset, mi::BpData{}, st);
    ctx->image_from_device(ctx->img.p, ctx->stage_img.p, st);
    HIPCHECK(hipMemcpyAsync(image, ctx->stage_img.p, ctx->stage_img.bytes(), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"

// The multiplicative viewgrams of the model (distributable.cxx:197-223): the
// caller's factors, or ones when only "zero end planes of segment 0" asks for
// them, with the end sinograms of segment 0 zeroed on request; none otherwise.
void mi_ctx::rebuild_mult(hipStream_t st) {
  const int64_t nb = total_bins();
  if (!mult_user.p && !zero_end_planes) {
    mult.release();
    return;
  }
  mult.alloc(nb);
  if (mult_user.p) HIPCHECK(hipMemcpyAsync(mult.p, mult_user.p, nb * sizeof(float), hipMemcpyDeviceToDevice, st));
  else mi::launch_fill(mult.p, nb, 1.F, st);
  if (zero_end_planes) {
    const mi_geometry& g = plan->geometry();
    DBuf<int64_t> off;
    off.alloc(ntof);
    std::vector<int64_t> o(ntof);
    for (int k = 0; k < ntof; ++k) o[k] = plan->segment_offset(0) + k * block_bins;
    HIPCHECK(hipMemcpyAsync(off.p, o.data(), o.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    const int runs = g.num_views * g.num_tangential_poss, run = plan->num_axial(0);
    for (int k = 0; k < ntof; ++k) mi::launch_zero_end_planes(mult.p, off.p + k, runs, run, 0, st);
  }
  HIPCHECK(hipStreamSynchronize(st));
}

extern "C" {

int mi_obj_set_data(mi_ctx* ctx, const float* measured, const float* additive, float additive_constant,
                    int zero_end_planes) {
  return guarded([&] {
    if (!ctx || !measured) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    const mi_geometry& g = ctx->plan->geometry();
    hipStream_t st = ctx->stream;
    const int64_t nb = ctx->total_bins();
    ctx->y.alloc(nb);
    HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, measured, nb * sizeof(float), hipMemcpyHostToDevice, st));
    ctx->proj_to_device(ctx->stage_sino.p, ctx->y.p, st);
    if (additive) {
      ctx->add.alloc(nb);
      HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, additive, nb * sizeof(float), hipMemcpyHostToDevice, st));
      ctx->proj_to_device(ctx->stage_sino.p, ctx->add.p, st);
    } else {
      ctx->add.release();
    }
    ctx->add_const = additive_constant;
    ctx->zero_end_planes = zero_end_planes;
    ctx->lm = false;
    if (zero_end_planes) {
      ctx->y_raw.alloc(nb);
      HIPCHECK(hipMemcpyAsync(ctx->y_raw.p, ctx->y.p, nb * sizeof(float), hipMemcpyDeviceToDevice, st));
      if (ctx->add.p) {
        ctx->add_raw.alloc(nb);
        HIPCHECK(hipMemcpyAsync(ctx->add_raw.p, ctx->add.p, nb * sizeof(float), hipMemcpyDeviceToDevice, st));
      } else {
        ctx->add_raw.release();
      }
    } else {
      ctx->y_raw.release();
      ctx->add_raw.release();
    }
    ctx->ones.alloc(nb);
    mi::launch_fill(ctx->ones.p, nb, 1.F, st);
    if (zero_end_planes) {
      // "zero end planes of segment 0" (distributable.cxx:218-223): y and b get
      // their first and last sinogram zeroed (the multiplicative viewgrams too,
      // in rebuild_mult)
      DBuf<int64_t> off;
      off.alloc(ctx->ntof);
      std::vector<int64_t> o(ctx->ntof);
      for (int k = 0; k < ctx->ntof; ++k) o[k] = ctx->plan->segment_offset(0) + k * ctx->block_bins;
      HIPCHECK(hipMemcpyAsync(off.p, o.data(), o.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
      const int runs = g.num_views * g.num_tangential_poss, run = ctx->plan->num_axial(0);
      for (int k = 0; k < ctx->ntof; ++k) {
        mi::launch_zero_end_planes(ctx->y.p, off.p + k, runs, run, 0, st);
        if (ctx->add.p) mi::launch_zero_end_planes(ctx->add.p, off.p + k, runs, run, 0, st);
      }
      HIPCHECK(hipStreamSynchronize(st));
    }
    ctx->rebuild_mult(st);
    const int nvg = ctx->plan->num_viewgrams() * ctx->ntof;
    ctx->vg_small.alloc(nvg);
    mi::launch_viewgram_max(ctx->y.p, ctx->vg_off.p, ctx->vg_len.p, nvg, ctx->vg_small.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    ctx->have_data = true;
    ctx->have_sens = false;
  });
}

int mi_obj_compute_sensitivities(mi_ctx* ctx) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_data) throw ApiError(MI_ERR_STATE, "mi_obj_set_data has not been called");
    HIPCHECK(hipSetDevice(ctx->device));
    const int nsub = ctx->plan->num_subsets();
    ctx->sens.clear();
    const float* in = ctx->mult.p ? ctx->mult.p : ctx->ones.p;
    if (!ctx->use_subset_sens && !ctx->plan->subsets_balanced())
      // compute_sensitivities (...ForMean.cxx:377-384)
      throw ApiError(MI_ERR_STATE, "subsets are not balanced: use subset sensitivities");
    if (ctx->ntof > 1 && !ctx->tof_sensitivities && ctx->mult_user.p && !ctx->mult_tof_uniform)
      // the non-TOF sensitivity needs a non-TOF normalisation (...AndProjData.cxx:636-646)
      throw ApiError(MI_ERR_STATE, "multiplicative factors differ between TOF bins: use TOF sensitivities");
    if (ctx->ntof > 1 && !ctx->tof_sensitivities) {
      // TOF data: the non-TOF sensitivity (...AndProjData.cxx:533-536,648-670),
      // computed by the non-TOF projector, whose symmetries decide the
      // view-segments of each subset (add_subset_sensitivity, :884-903). Each
      // subset's viewgrams of the first (non-TOF) block of ones / mult are
      // back-projected through the whole table.
      mi_geometry g1 = ctx->plan->geometry();
      g1.num_tof_bins = 1;
      const mi::Plan nontof_plan(g1, nsub);
      DBuf<float> mask;
      mask.alloc(ctx->block_bins);
      std::vector<int64_t> h_off(ctx->plan->num_viewgrams());
      std::vector<int32_t> h_len(ctx->plan->num_viewgrams());
      HIPCHECK(hipMemcpy(h_off.data(), ctx->vg_off.p, h_off.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
      HIPCHECK(hipMemcpy(h_len.data(), ctx->vg_len.p, h_len.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int s = 0; s < nsub; ++s) {
        HIPCHECK(hipMemsetAsync(mask.p, 0, mask.bytes(), ctx->stream));
        for (const auto& vs : nontof_plan.subset_view_segments(s)) {
          const int vg = ctx->plan->viewgram_index(vs.second, vs.first);
          HIPCHECK(hipMemcpyAsync(mask.p + h_off[vg], in + h_off[vg], h_len[vg] * sizeof(float),
                                  hipMemcpyDeviceToDevice, ctx->stream));
        }
        auto b = std::make_unique<DBuf<float>>();
        b->alloc(ctx->img.n);
        ctx->bp(mi::BP_PLAIN, mask.p, b->p, -1, mi::BpData{}, ctx->stream, true);
        ctx->sens.push_back(std::move(b));
      }
      HIPCHECK(hipStreamSynchronize(ctx->stream));
    } else {
      for (int s = 0; s < nsub; ++s) {
        auto b = std::make_unique<DBuf<float>>();
        b->alloc(ctx->img.n);
        ctx->bp(mi::BP_PLAIN, in, b->p, s, mi::BpData{}, ctx->stream);
        ctx->sens.push_back(std::move(b));
      }
    }
    if (!ctx->use_subset_sens) {
      // "use subset sensitivities" := 0 (...ForMean.cxx:386-455): the total,
      // accumulated subset by subset, divided by num_subsets for every subset
      for (int s = 1; s < nsub; ++s) mi::launch_accumulate(ctx->sens[0]->p, ctx->sens[s]->p, ctx->img.n, ctx->stream);
      ctx->sens_total.alloc(ctx->img.n);
      HIPCHECK(hipMemcpyAsync(ctx->sens_total.p, ctx->sens[0]->p, ctx->sens_total.bytes(), hipMemcpyDeviceToDevice,
                              ctx->stream));
      mi::launch_divide_scalar(ctx->sens[0]->p, static_cast<float>(nsub), ctx->img.n, ctx->stream);
      for (int s = 1; s < nsub; ++s)
        HIPCHECK(hipMemcpyAsync(ctx->sens[s]->p, ctx->sens[0]->p, ctx->sens[0]->bytes(), hipMemcpyDeviceToDevice,
                                ctx->stream));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->have_sens = true;
  });
}

int mi_obj_get_sensitivity(mi_ctx* ctx, int subset, float* image) {
  return guarded([&] {
    if (!ctx || !image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_sens) throw ApiError(MI_ERR_STATE, "sensitivities have not been computed");
    if (subset < -1 || subset >= ctx->plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    HIPCHECK(hipSetDevice(ctx->device));
    const float* src = nullptr;
    if (subset >= 0) {
      src = ctx->sens[subset]->p;
    } else if (!ctx->use_subset_sens) {
      src = ctx->sens_total.p;
    } else {
      // total = sens_0 + sens_1 + ... in subset order (...ForMean.cxx:419-438)
      HIPCHECK(hipMemcpyAsync(ctx->img2.p, ctx->sens[0]->p, ctx->img2.bytes(), hipMemcpyDeviceToDevice, ctx->stream));
      for (size_t s = 1; s < ctx->sens.size(); ++s)
        mi::launch_accumulate(ctx->img2.p, ctx->sens[s]->p, ctx->img2.n, ctx->stream);
      src = ctx->img2.p;
    }
    ctx->image_from_device(src, ctx->stage_img.p, ctx->stream);
    HIPCHECK(hipMemcpyAsync(image, ctx->stage_img.p, ctx->stage_img.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mi_obj_set_estimate(mi_ctx* ctx, const float* image, int enforce_positivity) {
  return guarded([&] {
    if (!ctx || !image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    const int64_t nv = ctx->plan->num_voxels();
    std::vector<float> h(image, image + nv);
    if (enforce_positivity) {
      // threshold_min_to_small_positive_value(x, 1e-6) (thresholding.h:101-109)
      const float small_number = 0.000001F;
      bool found = false;
      float minpos = 0;
      for (float v : h)
        if (v > 0 && (!found || v < minpos)) { minpos = v; found = true; }
      if (found) {
        const float thr = minpos * small_number;
        for (float& v : h)
          if (thr > v) v = thr;
      } else {
        std::fill(h.begin(), h.end(), small_number);
      }
    }
    if (ctx->x.n != ctx->img.n) ctx->x.alloc(ctx->img.n);  // keep the address: batches hold it
    HIPCHECK(hipMemsetAsync(ctx->x.p, 0, ctx->x.bytes(), ctx->stream));
    HIPCHECK(hipMemcpyAsync(ctx->stage_img.p, h.data(), nv * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    ctx->image_to_device(ctx->stage_img.p, ctx->x.p, ctx->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->have_x = true;
  });
}

int mi_obj_get_estimate(mi_ctx* ctx, float* image) {
  return guarded([&] {
    if (!ctx || !image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_x) throw ApiError(MI_ERR_STATE, "no estimate");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->image_from_device(ctx->x.p, ctx->stage_img.p, ctx->stream);
    HIPCHECK(hipMemcpyAsync(image, ctx->stage_img.p, ctx->stage_img.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

static void osem_step(mi_ctx* ctx, int subset, int subiteration_num) {
  ctx->fp(ctx->ratio_mode(), ctx->x.p, ctx->ratio.p, subset, ctx->fp_data(), ctx->stream);
  if (!ctx->prior_active()) {
    // prior-free: divide, threshold and multiply fused into the back projection
    mi::BpData bd{};
    bd.sens = ctx->sens[subset]->p;
    bd.estimate = ctx->x.p;
    bd.threshold = subiteration_num != 1;
    bd.thr_lo = ctx->min_rel;
    bd.thr_hi = ctx->max_rel;
    ctx->bp(mi::BP_OSEM_UPDATE, ctx->ratio.p, nullptr, subset, bd, ctx->stream);
    return;
  }
  // one-step-late MAP update (OSMAPOSLReconstruction.cxx:403-516): u = BP(y/ybar),
  // the prior's gradient at the current estimate, divide(u, denominator, 1e-6)
  ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, ctx->stream);
  mi::launch_prior(ctx->prior_kind, ctx->x.p, ctx->prior_kappa.p, ctx->device_prior_image(), ctx->prior_w,
                   ctx->prior_only_2D, ctx->prior_beta, ctx->prior_gamma, ctx->prior_eps, ctx->prior_grad.p,
                   ctx->prior_partial.p, ctx->stream);
  mi::launch_small_value(ctx->img2.p, ctx->img2.n, 0.000001F, ctx->small_part.p, ctx->small_val.p, ctx->stream);
  mi::launch_osl_update(ctx->x.p, ctx->img2.p, ctx->prior_grad.p, ctx->sens[subset]->p, ctx->small_val.p,
                        ctx->plan->num_subsets(), ctx->map_multiplicative, subiteration_num != 1, ctx->min_rel,
                        ctx->max_rel, ctx->dims.nz, ctx->dims.nzp, ctx->img.n, ctx->stream);
}

int mi_osem_subset_step(mi_ctx* ctx, int subset, int subiteration_num) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_data || !ctx->have_sens || !ctx->have_x)
      throw ApiError(MI_ERR_STATE, "data, sensitivities and estimate must be set first");
    if (subset < 0 || subset >= ctx->plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    HIPCHECK(hipSetDevice(ctx->device));
    osem_step(ctx, subset, subiteration_num);
    HIPCHECK(hipGetLastError());
  });
}

int mi_osem_run(mi_ctx* ctx, int start_subiteration_num, int num_subiterations, int start_subset) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    if (!ctx->have_data || !ctx->have_sens || !ctx->have_x)
      throw ApiError(MI_ERR_STATE, "data, sensitivities and estimate must be set first");
    HIPCHECK(hipSetDevice(ctx->device));
    const int nsub = ctx->plan->num_subsets();
    for (int k = start_subiteration_num; k < start_subiteration_num + num_subiterations; ++k) {
      const int subset = (k + start_subset - 1) % nsub;  // IterativeReconstruction.cxx:630-639
      osem_step(ctx, subset, k);
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_obj_subset_gradient(mi_ctx* ctx, int subset, int add_sensitivity, float* gradient) {
  return guarded([&] {
    if (!ctx || !gradient) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    if (!ctx->have_data || !ctx->have_x) throw ApiError(MI_ERR_STATE, "data and estimate must be set first");
    HIPCHECK(hipSetDevice(ctx->device));
    if (ctx->lm) {
      // list mode: sum of row^T/fwd over the subset's events, minus the subset
      // sensitivity (...ListModeDataWithProjMatrixByBin.cxx:913-961)
      if (!add_sensitivity && !ctx->have_sens) throw ApiError(MI_ERR_STATE, "sensitivities have not been computed");
      ctx->fp(mi::FP_LM_RATIO, ctx->x.p, ctx->ratio.p, subset, ctx->fp_data(), ctx->stream);
      ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, ctx->stream);
      if (!add_sensitivity)
        for (int s = ctx->subset_begin(subset); s < ctx->subset_end(subset); ++s)
          mi::launch_subtract(ctx->img2.p, ctx->sens[s]->p, ctx->img2.n, ctx->stream);
    } else {
      ctx->fp(add_sensitivity ? mi::FP_RATIO : mi::FP_GRADIENT, ctx->x.p, ctx->ratio.p, subset, ctx->fp_data(),
              ctx->stream);
      ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, ctx->stream);
    }
    ctx->image_from_device(ctx->img2.p, ctx->stage_img.p, ctx->stream);
    HIPCHECK(hipMemcpyAsync(gradient, ctx->stage_img.p, ctx->stage_img.bytes(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

int mi_obj_log_likelihood(mi_ctx* ctx, int subset, double* value) {
  return guarded([&] {
    if (!ctx || !value) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    if (!ctx->have_data || !ctx->have_x) throw ApiError(MI_ERR_STATE, "data and estimate must be set first");
    HIPCHECK(hipSetDevice(ctx->device));
    const int tb = ctx->task_off[ctx->subset_begin(subset)], te = ctx->task_off[ctx->subset_end(subset)];
    ctx->fp(ctx->lm ? mi::FP_LM_VALUE : mi::FP_LOGLIK, ctx->x.p, ctx->ratio.p, subset, ctx->fp_data(), ctx->stream);
    std::vector<double> h(te - tb);
    HIPCHECK(hipMemcpyAsync(h.data(), ctx->partial.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    double acc = 0;
    for (double v : h) acc += v;
    *value = acc;
  });
}

int mi_obj_value_and_gradient_dev(mi_ctx* ctx, const float* d_image, int subset, int add_sensitivity,
                                  float* d_gradient, double* d_value, void* stream) {
  return guarded([&] {
    if (!ctx || !d_image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (!d_gradient && !d_value) throw ApiError(MI_ERR_ARGUMENT, "neither a gradient nor a value was asked for");
    ctx->require_setup();
    ctx->check_subset(subset);
    if (ctx->lm) throw ApiError(MI_ERR_STATE, "the fused value and gradient work on projection data, not list mode");
    if (!ctx->have_data) throw ApiError(MI_ERR_STATE, "mi_obj_set_data has not been called");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    ctx->image_to_device(d_image, ctx->img.p, st);
    mi::FpMode mode;
    if (!d_gradient) mode = mi::FP_LOGLIK;
    else if (!d_value) mode = add_sensitivity ? mi::FP_RATIO : mi::FP_GRADIENT;
    else mode = add_sensitivity ? mi::FP_RATIO_LOGLIK : mi::FP_GRADIENT_LOGLIK;
    ctx->fp(mode, ctx->img.p, ctx->ratio.p, subset, ctx->fp_data(), st);
    if (d_value) {
      const int tb = ctx->task_off[ctx->subset_begin(subset)], te = ctx->task_off[ctx->subset_end(subset)];
      mi::launch_sum_partials(ctx->partial.p, te - tb, 1.0, d_value, st);
    }
    if (d_gradient) {
      ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, st);
      ctx->image_from_device(ctx->img2.p, d_gradient, st);
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_quadratic_prior_dev(mi_ctx* ctx, const float* d_image, const float* d_kappa, const float* weights, float beta,
                           int only_2D, float* d_gradient, double* d_value, void* stream) {
  return guarded([&] {
    if (!ctx || !d_image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (!d_gradient && !d_value) throw ApiError(MI_ERR_ARGUMENT, "neither a gradient nor a value was asked for");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    const mi_geometry& g = ctx->plan->geometry();
    if (beta == 0.F) {  // QuadraticPrior.cxx:233-237,305-309
      if (d_gradient) mi::launch_fill(d_gradient, ctx->plan->num_voxels(), 0.F, st);
      if (d_value) mi::launch_sum_partials(ctx->prior_partial.p, 0, 0.0, d_value, st);
    } else {
      mi::PriorWeights W;
      if (weights) std::memcpy(W.w, weights, sizeof(W.w));
      else default_prior_weights(g, W.w);
      mi::launch_prior(1, d_image, d_kappa, ctx->stir_prior_image(), W, only_2D, beta, 0.F, 0.F, d_gradient,
                       ctx->prior_partial.p, st);
      if (d_value)
        mi::launch_sum_partials(ctx->prior_partial.p, mi::prior_num_partials(ctx->stir_prior_image()),
                                static_cast<double>(beta), d_value, st);
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_quadratic_prior_noalloc_dev(int device, int nz, int ny, int nx, float vox_z, float vox_y, float vox_x,
                                   const float* d_image, const float* d_kappa, const float* weights, float beta,
                                   int only_2D, float* d_gradient, double* d_value, void* stream) {
  return guarded([&] {
    if (!d_image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (!d_gradient && !d_value) throw ApiError(MI_ERR_ARGUMENT, "neither a gradient nor a value was asked for");
    if (nz < 1 || ny < 1 || nx < 1) throw ApiError(MI_ERR_ARGUMENT, "empty image");
    const mi::PriorImage im{nz, ny, nx, nz, 0};
    const int nparts = mi::prior_num_partials(im);
    if (nparts > mi::kNoAllocPartials)
      throw ApiError(MI_ERR_ARGUMENT, "image too large for the fixed partial sums (" + std::to_string(nparts) +
                                          " workgroups, at most " + std::to_string(mi::kNoAllocPartials) + ")");
    if (device < 0 || device >= mi::kMaxDevices) throw ApiError(MI_ERR_ARGUMENT, "device out of range");
    HIPCHECK(hipSetDevice(device));
    double* partials = mi::noalloc_partials(device);
    if (!partials) throw ApiError(MI_ERR_DEVICE, "hipGetSymbolAddress of the partial sums failed");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (beta == 0.F) {  // QuadraticPrior.cxx:233-237,305-309
      if (d_gradient) mi::launch_fill(d_gradient, static_cast<int64_t>(nz) * ny * nx, 0.F, st);
      if (d_value) mi::launch_sum_partials(partials, 0, 0.0, d_value, st);
    } else {
      mi::PriorWeights W;
      if (weights) std::memcpy(W.w, weights, sizeof(W.w));
      else default_prior_weights(vox_z, vox_y, vox_x, W.w);
      mi::launch_prior(1, d_image, d_kappa, im, W, only_2D, beta, 0.F, 0.F, d_gradient, partials, st);
      if (d_value) mi::launch_sum_partials(partials, nparts, static_cast<double>(beta), d_value, st);
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_osem_subset_step_dev(mi_ctx* ctx, const float* d_x_in, float* d_x_out, float* d_update, int subset,
                            int subiteration_num, void* stream) {
  return guarded([&] {
    if (!ctx || !d_x_in || !d_x_out) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (ctx->lm) throw ApiError(MI_ERR_STATE, "the out-of-place OSEM step works on projection data, not list mode");
    if (ctx->prior_active()) throw ApiError(MI_ERR_STATE, "the out-of-place OSEM step is prior-free");
    ctx->require_setup();
    if (!ctx->have_data || !ctx->have_sens) throw ApiError(MI_ERR_STATE, "data and sensitivities must be set first");
    if (subset < 0 || subset >= ctx->plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    ctx->image_to_device(d_x_in, ctx->img.p, st);
    ctx->fp(mi::FP_RATIO, ctx->img.p, ctx->ratio.p, subset, ctx->fp_data(), st);
    mi::BpData bd{};
    bd.sens = ctx->sens[subset]->p;
    bd.estimate = ctx->img.p;
    bd.update = d_update ? ctx->img2.p : nullptr;
    bd.threshold = subiteration_num != 1;
    bd.thr_lo = ctx->min_rel;
    bd.thr_hi = ctx->max_rel;
    ctx->bp(mi::BP_OSEM_UPDATE, ctx->ratio.p, nullptr, subset, bd, st);
    ctx->image_from_device(ctx->img.p, d_x_out, st);
    if (d_update) ctx->image_from_device(ctx->img2.p, d_update, st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_osem_subset_step_vjp_dev(mi_ctx* ctx, const float* d_x_in, const float* d_update, const float* d_grad_out,
                                float* d_grad_in, int subset, int subiteration_num, void* stream) {
  return guarded([&] {
    if (!ctx || !d_x_in || !d_update || !d_grad_out || !d_grad_in) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (d_grad_in == d_grad_out || d_grad_in == d_update)
      throw ApiError(MI_ERR_ARGUMENT, "d_grad_in must not alias d_grad_out or d_update");
    ctx->require_setup();
    if (!ctx->have_data || !ctx->have_sens) throw ApiError(MI_ERR_STATE, "data and sensitivities must be set first");
    if (subset < 0 || subset >= ctx->plan->num_subsets()) throw ApiError(MI_ERR_ARGUMENT, "subset out of range");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    const int64_t ndev = static_cast<int64_t>(ctx->dims.ncols) * ctx->dims.nzp;
    // r'(FP_s(x) + b) on the subset's bins
    ctx->image_to_device(d_x_in, ctx->img.p, st);
    ctx->fp(mi::FP_DRATIO, ctx->img.p, ctx->ratio.p, subset, ctx->fp_data(), st);
    // v = g * x / sens_s where the update was not clamped (the clamp of
    // threshold_upper_lower has derivative 0 where it acts; it acts after
    // subiteration 1 only), then FP_s(v) * r'
    if (subiteration_num != 1) {
      mi::launch_osem_vjp_mask(d_grad_out, d_update, ctx->stage_img.p, ctx->plan->num_voxels(), ctx->min_rel,
                               ctx->max_rel, st);
      ctx->image_to_device(ctx->stage_img.p, ctx->img2.p, st);
    } else {
      ctx->image_to_device(d_grad_out, ctx->img2.p, st);
    }
    mi::launch_osem_vjp_scale(ctx->img2.p, ctx->img.p, ctx->sens[subset]->p, ctx->img.p, ndev, st);
    mi::FpData fd = ctx->fp_data();
    fd.scale = ctx->ratio.p;
    ctx->fp(mi::FP_SCALE, ctx->img.p, ctx->sino.p, subset, fd, st);
    ctx->bp(mi::BP_PLAIN, ctx->sino.p, ctx->img2.p, subset, mi::BpData{}, st);
    ctx->image_from_device(ctx->img2.p, d_grad_in, st);
    mi::launch_add_product(d_grad_in, d_grad_out, d_update, ctx->plan->num_voxels(), st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_obj_set_use_tof_sensitivities(mi_ctx* ctx, int use) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->tof_sensitivities = use != 0;
    ctx->have_sens = false;
  });
}

int mi_obj_set_multiplicative(mi_ctx* ctx, const float* mult) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    if (!mult) {
      ctx->mult_user.release();
      ctx->mult_tof_uniform = true;
    } else {
      const int64_t nb = ctx->total_bins();
      bool uniform = true;
      for (int k = 1; k < ctx->ntof && uniform; ++k)
        uniform = std::memcmp(mult, mult + k * ctx->block_bins, ctx->block_bins * sizeof(float)) == 0;
      ctx->mult_tof_uniform = uniform;
      ctx->mult_user.alloc(nb);
      HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, mult, nb * sizeof(float), hipMemcpyHostToDevice, st));
      ctx->proj_to_device(ctx->stage_sino.p, ctx->mult_user.p, st);
    }
    if (ctx->have_data) ctx->rebuild_mult(st);
    HIPCHECK(hipStreamSynchronize(st));
    ctx->have_sens = false;
  });
}

int mi_obj_set_use_subset_sensitivities(mi_ctx* ctx, int use) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->use_subset_sens = use != 0;
    ctx->have_sens = false;
  });
}

int mi_lm_set_events(mi_ctx* ctx, const int32_t* events, int64_t num_events, const float* additive,
                     float additive_constant) {
  return guarded([&] {
    if (!ctx || (!events && num_events > 0)) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (num_events < 0) throw ApiError(MI_ERR_ARGUMENT, "negative number of events");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    const mi_geometry& g = ctx->plan->geometry();
    const int64_t nb = ctx->total_bins();
    // histogram of the prompts, events outside the projection data dropped
    // (...ListModeDataWithProjMatrixByBin.cxx:495-511): y events of one bin
    // contribute as y events of value 1 do
    std::vector<float> counts(nb, 0.F);
    const int min_tof = -(ctx->ntof / 2);
    const int ntang = g.num_tangential_poss, min_tang = -(ntang / 2);
    int64_t kept = 0;
    for (int64_t i = 0; i < num_events; ++i) {
      const int32_t* e = events + 5 * i;
      const int seg = e[0], view = e[1], ax = e[2], tang = e[3], tof = e[4];
      if (seg < g.min_segment_num || seg > g.max_segment_num || view < 0 || view >= g.num_views ||
          tang < min_tang || tang >= min_tang + ntang || ax < 0 || ax >= ctx->plan->num_axial(seg))
        continue;
      if (ctx->ntof > 1 ? (tof < min_tof || tof >= min_tof + ctx->ntof) : tof != 0) continue;
      const int64_t k = ctx->ntof > 1 ? tof - min_tof : 0;
      const int64_t idx = k * ctx->block_bins + ctx->plan->segment_offset(seg) +
                          (static_cast<int64_t>(ax) * g.num_views + view) * ntang + (tang - min_tang);
      counts[idx] += 1.F;
      ++kept;
    }
    (void)kept;
    hipStream_t st = ctx->stream;
    ctx->y.alloc(nb);
    HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, counts.data(), nb * sizeof(float), hipMemcpyHostToDevice, st));
    ctx->proj_to_device(ctx->stage_sino.p, ctx->y.p, st);
    if (additive) {
      // the cache keeps the additive term of the last timing position for
      // every event (...ListModeDataWithProjMatrixByBin.cxx:570-596)
      std::vector<float> a(nb);
      for (int k = 0; k < ctx->ntof; ++k)
        std::memcpy(a.data() + k * ctx->block_bins, additive + (ctx->ntof - 1) * ctx->block_bins,
                    ctx->block_bins * sizeof(float));
      ctx->add.alloc(nb);
      HIPCHECK(hipMemcpyAsync(ctx->stage_sino.p, a.data(), nb * sizeof(float), hipMemcpyHostToDevice, st));
      ctx->proj_to_device(ctx->stage_sino.p, ctx->add.p, st);
      HIPCHECK(hipStreamSynchronize(st));
    } else {
      ctx->add.release();
    }
    ctx->add_const = additive_constant;
    ctx->zero_end_planes = 0;
    ctx->ones.alloc(nb);
    mi::launch_fill(ctx->ones.p, nb, 1.F, st);
    ctx->rebuild_mult(st);
    const int nvg = ctx->plan->num_viewgrams() * ctx->ntof;
    ctx->vg_small.alloc(nvg);
    mi::launch_viewgram_max(ctx->y.p, ctx->vg_off.p, ctx->vg_len.p, nvg, ctx->vg_small.p, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    ctx->y_raw.release();
    ctx->add_raw.release();
    ctx->lm = true;
    ctx->have_data = true;
    ctx->have_sens = false;
  });
}

int mi_osem_set_options(mi_ctx* ctx, int map_model_multiplicative, float minimum_relative_change,
                        float maximum_relative_change) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->map_multiplicative = map_model_multiplicative != 0;
    ctx->min_rel = minimum_relative_change;
    ctx->max_rel = maximum_relative_change;
  });
}

int mi_obj_set_prior(mi_ctx* ctx, const mi_prior* prior) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    if (!prior || prior->type == MI_PRIOR_NONE) {
      ctx->prior_kind = 0;
      ctx->prior_kappa.release();
      ctx->prior_grad.release();
      return;
    }
    if (prior->type != MI_PRIOR_QUADRATIC && prior->type != MI_PRIOR_RELATIVE_DIFFERENCE)
      throw ApiError(MI_ERR_ARGUMENT, "unknown prior type");
    const mi_geometry& g = ctx->plan->geometry();
    ctx->prior_kind = prior->type;
    ctx->prior_beta = prior->penalisation_factor;
    ctx->prior_gamma = prior->gamma;
    ctx->prior_eps = prior->epsilon;
    ctx->prior_only_2D = prior->only_2D != 0;
    if (prior->weights) std::memcpy(ctx->prior_w.w, prior->weights, sizeof(ctx->prior_w.w));
    else default_prior_weights(g, ctx->prior_w.w);
    hipStream_t st = ctx->stream;
    if (prior->kappa) {
      ctx->prior_kappa.alloc(ctx->img.n);
      HIPCHECK(hipMemsetAsync(ctx->prior_kappa.p, 0, ctx->prior_kappa.bytes(), st));
      HIPCHECK(hipMemcpyAsync(ctx->stage_img.p, prior->kappa, ctx->plan->num_voxels() * sizeof(float),
                              hipMemcpyHostToDevice, st));
      ctx->image_to_device(ctx->stage_img.p, ctx->prior_kappa.p, st);
    } else {
      ctx->prior_kappa.release();
    }
    ctx->prior_grad.alloc(ctx->img.n);
    HIPCHECK(hipMemsetAsync(ctx->prior_grad.p, 0, ctx->prior_grad.bytes(), st));
    ctx->small_part.alloc(mi::max_num_partials());
    ctx->small_val.alloc(1);
    HIPCHECK(hipStreamSynchronize(st));
  });
}

int mi_relative_difference_prior_dev(mi_ctx* ctx, const float* d_image, const float* d_kappa, const float* weights,
                                     float beta, float gamma, float epsilon, int only_2D, float* d_gradient,
                                     double* d_value, void* stream) {
  return guarded([&] {
    if (!ctx || !d_image) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (!d_gradient && !d_value) throw ApiError(MI_ERR_ARGUMENT, "neither a gradient nor a value was asked for");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    const mi_geometry& g = ctx->plan->geometry();
    if (beta == 0.F) {  // RelativeDifferencePrior.cxx:307-310,380-384
      if (d_gradient) mi::launch_fill(d_gradient, ctx->plan->num_voxels(), 0.F, st);
      if (d_value) mi::launch_sum_partials(ctx->prior_partial.p, 0, 0.0, d_value, st);
    } else {
      mi::PriorWeights W;
      if (weights) std::memcpy(W.w, weights, sizeof(W.w));
      else default_prior_weights(g, W.w);
      mi::launch_prior(2, d_image, d_kappa, ctx->stir_prior_image(), W, only_2D, beta, gamma, epsilon, d_gradient,
                       ctx->prior_partial.p, st);
      if (d_value)
        mi::launch_sum_partials(ctx->prior_partial.p, mi::prior_num_partials(ctx->stir_prior_image()),
                                static_cast<double>(beta), d_value, st);
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_obj_hessian_times_input_dev(mi_ctx* ctx, int approximate, const float* d_current, const float* d_input,
                                   int subset, float* d_output, void* stream) {
  return guarded([&] {
    if (!ctx || !d_input || !d_output || (!d_current && !approximate)) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    ctx->require_setup();
    ctx->check_subset(subset);
    if (ctx->lm && approximate)
      // the list-mode objective does not overload the approximate product
      // (GeneralisedObjectiveFunction.cxx:341-348)
      throw ApiError(MI_ERR_STATE, "the list-mode objective has no approximate Hessian product");
    if (!ctx->have_data) throw ApiError(MI_ERR_STATE, "mi_obj_set_data has not been called");
    HIPCHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->pick(stream);
    const int nsub = ctx->plan->num_subsets();
    const int nvg1 = ctx->plan->num_viewgrams();
    if (ctx->subset_vg_off.empty()) {
      // the viewgrams of each subset, every TOF bin (find_basic_viewgram_indices_in_subset
      // with their related viewgrams)
      std::vector<int32_t> h;
      ctx->subset_vg_off.push_back(0);
      for (int s = 0; s < nsub; ++s) {
        for (int k = 0; k < ctx->ntof; ++k)
          for (const auto& vs : ctx->plan->subset_view_segments(s))
            h.push_back(k * nvg1 + ctx->plan->viewgram_index(vs.second, vs.first));
        ctx->subset_vg_off.push_back(static_cast<int>(h.size()));
      }
      ctx->subset_vgs.alloc(h.size());
      HIPCHECK(hipMemcpy(ctx->subset_vgs.p, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      ctx->hess_flag.alloc(1);
    }
    const int v0 = ctx->subset_vg_off[ctx->subset_begin(subset)], v1 = ctx->subset_vg_off[ctx->subset_end(subset)];
    // A input and, for the exact product, A current
    ctx->image_to_device(d_input, ctx->img.p, st);
    ctx->fp(mi::FP_PLAIN, ctx->img.p, ctx->ratio.p, subset, mi::FpData{}, st);
    if (!approximate) {
      ctx->image_to_device(d_current, ctx->img.p, st);
      ctx->fp(mi::FP_PLAIN, ctx->img.p, ctx->sino.p, subset, mi::FpData{}, st);
    }
    if (ctx->lm) {
      // LM_Hessian over the histogrammed events (...ListModeDataWithProjMatrixByBin.cxx:790-822,963-1003):
      // output += A^T q, q <= 0, no negativity check
      mi::launch_lm_hessian(ctx->subset_vgs.p + v0, v1 - v0, ctx->vg_off.p, ctx->vg_len.p, ctx->y.p, ctx->ratio.p,
                            ctx->sino.p, ctx->add.p, ctx->add_const, ctx->ratio.p, st);
      ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, st);
      ctx->image_from_device(ctx->img2.p, ctx->stage_img.p, st);
      mi::launch_accumulate(d_output, ctx->stage_img.p, ctx->stage_img.n, st);
      HIPCHECK(hipGetLastError());
      return;
    }
    HIPCHECK(hipMemsetAsync(ctx->hess_flag.p, 0, sizeof(int32_t), st));
    const float* y = ctx->y_raw.p ? ctx->y_raw.p : ctx->y.p;
    const float* add = ctx->add_raw.p ? ctx->add_raw.p : ctx->add.p;
    mi::launch_hessian_ratio(ctx->subset_vgs.p + v0, v1 - v0, ctx->vg_off.p, ctx->vg_len.p, y, ctx->ratio.p,
                             ctx->sino.p, add, ctx->add_const, ctx->mult_user.p, approximate, ctx->ratio.p,
                             ctx->hess_flag.p, st);
    int32_t negative = 0;
    HIPCHECK(hipMemcpyAsync(&negative, ctx->hess_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (negative)
      // ...AndProjData.cxx:1043-1046,1143-1146: the output is left as it was
      throw ApiError(MI_ERR_ARGUMENT, "forward projection of input contains negatives: the result would be incorrect");
    ctx->bp(mi::BP_PLAIN, ctx->ratio.p, ctx->img2.p, subset, mi::BpData{}, st);
    ctx->image_from_device(ctx->img2.p, ctx->stage_img.p, st);
    mi::launch_subtract(d_output, ctx->stage_img.p, ctx->stage_img.n, st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_prior_hessian_times_input_dev(mi_ctx* ctx, int type, int approximate, const float* d_image,
                                     const float* d_input, const float* d_kappa, const float* weights, float beta,
                                     float gamma, float epsilon, int only_2D, float* d_output, void* stream) {
  return guarded([&] {
    if (!ctx || !d_input || !d_output || (!d_image && !approximate)) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (type != MI_PRIOR_QUADRATIC && type != MI_PRIOR_RELATIVE_DIFFERENCE)
      throw ApiError(MI_ERR_ARGUMENT, "unknown prior type");
    if (approximate && type != MI_PRIOR_QUADRATIC)
      // RelativeDifferencePrior.cxx:527-532
      throw ApiError(MI_ERR_ARGUMENT, "the approximate Hessian is implemented for the quadratic prior only");
    ctx->require_setup();
    HIPCHECK(hipSetDevice(ctx->device));
    if (beta == 0.F) return;  // accumulate_Hessian_times_input returns at once
    hipStream_t st = ctx->pick(stream);
    mi::PriorWeights W;
    if (weights) std::memcpy(W.w, weights, sizeof(W.w));
    else default_prior_weights(ctx->plan->geometry(), W.w);
    const int kind = approximate ? 3 : (type == MI_PRIOR_QUADRATIC ? 1 : 2);
    mi::launch_prior_hessian(kind, approximate ? d_input : d_image, d_input, d_kappa, ctx->stir_prior_image(), W,
                             only_2D, beta, gamma, epsilon, d_output, st);
    HIPCHECK(hipGetLastError());
  });
}

int mi_set_tuning(mi_ctx* ctx, int param, int value) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    switch (param) {
      case MI_TUNE_BP_TILE_X:
        if (value < 1 || value > 1024) throw ApiError(MI_ERR_ARGUMENT, "tile x out of range");
        ctx->bp_tuning.tile_x = value;
        break;
      case MI_TUNE_BP_TILE_Y:
        if (value < 1 || value > 1024) throw ApiError(MI_ERR_ARGUMENT, "tile y out of range");
        ctx->bp_tuning.tile_y = value;
        break;
      case MI_TUNE_BP_XCD_REMAP:
        if (value < 0 || value > 4096) throw ApiError(MI_ERR_ARGUMENT, "XCD remap out of range");
        ctx->bp_tuning.xcd_remap = value;
        break;
      case MI_TUNE_BP_SLOT_SKIP: ctx->bp_tuning.slot_skip = value != 0; break;
      case MI_TUNE_FP_XCD_CHUNK:
        if (value < 0 || value > 4096) throw ApiError(MI_ERR_ARGUMENT, "forward XCD chunk out of range");
        ctx->fp_xcd_chunk = value;
        break;
      case MI_TUNE_BP_SLAB: ctx->slab_enable = value != 0; break;
      case MI_TUNE_SLAB_TILE_X:
      case MI_TUNE_SLAB_TILE_Y:
        if (value < 1 || value > 64) throw ApiError(MI_ERR_ARGUMENT, "slab tile side out of range");
        (param == MI_TUNE_SLAB_TILE_X ? ctx->slab_tx : ctx->slab_ty) = value;
        break;
      case MI_TUNE_SLAB_COLUMNS_PER_WAVE:
        if (value != 2 && value != 4 && value != 8) throw ApiError(MI_ERR_ARGUMENT, "columns per wave must be 2, 4 or 8");
        ctx->slab_cpw = value;
        break;
      case MI_TUNE_SLAB_LDS_KB:
        if (value < 16 || value > 160) throw ApiError(MI_ERR_ARGUMENT, "slab LDS budget out of range");
        ctx->slab_lds_kb = value;
        break;
      case MI_TUNE_TOF_ROWS:
        if (value < 0 || value > 2) throw ApiError(MI_ERR_ARGUMENT, "TOF row mode must be 0, 1 or 2");
        ctx->tof_rows_mode = value;
        break;
      case MI_TUNE_BP_QUAD:
        if (value < 0 || value > 1) throw ApiError(MI_ERR_ARGUMENT, "the block back projection switch must be 0 or 1");
        ctx->quad_enable = value;
        break;
      case MI_TUNE_BP_QUAD_UNROLL:
        if (value != 4 && value != 6 && value != 8)
          throw ApiError(MI_ERR_ARGUMENT, "the block back projection keeps 4, 6 or 8 entries in flight");
        ctx->bp_tuning.quad_unroll = value;
        break;
      case MI_TUNE_BP_SPLIT_SUBSETS:
        if (value < 0 || value > 1) throw ApiError(MI_ERR_ARGUMENT, "the subset split switch must be 0 or 1");
        ctx->bp_tuning.split_subsets = value;
        break;
      default: throw ApiError(MI_ERR_ARGUMENT, "unknown tuning parameter");
    }
  });
}

int mi_set_kernel_timing(mi_ctx* ctx, int enable) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->device));
    ctx->timing = enable != 0;
  });
}

int mi_get_kernel_timing(mi_ctx* ctx, double* fp_ms, int64_t* fp_launches, double* bp_ms, int64_t* bp_launches) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->device));
    double f = 0, b = 0;
    for (auto& p : ctx->fp_events) {
      HIPCHECK(hipEventSynchronize(p.second));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, p.first, p.second));
      f += ms;
    }
    for (auto& p : ctx->bp_events) {
      HIPCHECK(hipEventSynchronize(p.second));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, p.first, p.second));
      b += ms;
    }
    if (fp_ms) *fp_ms = f;
    if (bp_ms) *bp_ms = b;
    if (fp_launches) *fp_launches = static_cast<int64_t>(ctx->fp_events.size());
    if (bp_launches) *bp_launches = static_cast<int64_t>(ctx->bp_events.size());
    ctx->release_events();
  });
}

int mi_synchronize(mi_ctx* ctx) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    HIPCHECK(hipSetDevice(ctx->device));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"

// ===========================================================================
// Batches of independent problems (grouped kernels, one launch per group size)
struct mi_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int nsub = 0;
  std::vector<mi_ctx*> probs;  // owned
  // launch choices (mi_batch_tune)
  int fp_widen = 0;             // forward groups: 0 = smallest group size holding the family, 1 = twice that
  int fp_group_max = 32;        // forward tasks with at most this many bins go to lane groups, the rest to full waves
  int bp_pack = 2;              // back projection of images of at most 32 planes: full waves (0), lane groups (1),
                                // waves split into parts over table entries (2)
  int tile_x = 32, tile_y = 8;  // back-projection column order
  int fp_rowgroup_max = 32;     // families of at most this many axial positions run in row groups (0: off)
  // device descriptors and work lists
  DBuf<mi::ProblemDesc> descs;
  std::vector<std::unique_ptr<DBuf<const float*>>> sens_lists;
  std::vector<std::array<int, 4>> fp_n;                       // [subset][group-size index]
  std::vector<std::array<std::unique_ptr<DBuf<mi::WorkItem>>, 4>> fp_work;
  std::array<int, 4> bp_n{};
  std::array<std::unique_ptr<DBuf<mi::WorkItem>>, 4> bp_work;
  // full-wave items: forward (problem, task) per subset and slot count,
  // back projection (problem, column) per planes-per-axial-position class
  // (ppa 0 = any, 1, 2) and slot count: [class * 4 + nslot - 1]
  DBuf<const mi::FpTask*> task_lists;
  std::vector<std::array<int, 4>> fpw_n;
  std::vector<std::array<std::unique_ptr<DBuf<mi::WorkItem>>, 4>> fpw_work;
  // [ppa class * 6 + variant], variants: 4 parts, 2 parts, 1..4 slots
  std::array<int, 18> bpw_n{};
  std::array<std::unique_ptr<DBuf<mi::WorkItem>>, 18> bpw_work;
  // forward row groups per subset: items and their family lists
  std::vector<int> rg_n;
  std::vector<std::unique_ptr<DBuf<mi::RowGroupItem>>> rg_items;
  std::vector<std::unique_ptr<DBuf<int32_t>>> rg_fams;
  bool lists_ready = false;
  hipGraphExec_t graph_exec = nullptr;
  int graph_key[3] = {0, 0, 0};  // start subiteration, count, start subset of the captured graph
  ~mi_batch() {
    if (stream) (void)hipStreamSynchronize(stream);  // queued runs still use the problems' buffers
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (mi_ctx* c : probs) delete c;
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
constexpr int kGroupSizes[4] = {8, 16, 32, 64};
int group_index_at_least(int n) {
  for (int i = 0; i < 4; ++i)
    if (n <= kGroupSizes[i]) return i;
  return -1;
}

// mi_batch_osem_run returns without waiting: anything that writes a
// problem's buffers (on the problem's own stream), rebuilds the work lists or
// destroys the graph first waits for the batch stream's queued work.
void batch_wait(mi_batch* b) { HIPCHECK(hipStreamSynchronize(b->stream)); }

void batch_invalidate(mi_batch* b) {
  batch_wait(b);
  b->lists_ready = false;
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
}

// descriptors (device pointers of every problem) and work lists
void batch_prepare(mi_batch* b) {
  if (b->lists_ready) return;
  const int np = static_cast<int>(b->probs.size());
  std::vector<mi::ProblemDesc> h(np);
  b->sens_lists.clear();
  for (int p = 0; p < np; ++p) {
    mi_ctx* c = b->probs[p];
    if (!c->have_data || !c->have_sens || !c->have_x)
      throw ApiError(MI_ERR_STATE, "data, sensitivities and estimate must be set for every problem");
    auto sl = std::make_unique<DBuf<const float*>>();
    sl->alloc(b->nsub);
    std::vector<const float*> hs(b->nsub);
    for (int s = 0; s < b->nsub; ++s) hs[s] = c->sens[s]->p;
    HIPCHECK(hipMemcpy(sl->p, hs.data(), hs.size() * sizeof(const float*), hipMemcpyHostToDevice));
    mi::ProblemDesc& d = h[p];
    d.fams = c->fams.p;
    d.pool = c->pool.p;
    d.pool_ops = c->pool_ops.p;
    d.nelem = c->num_basic_elements;
    d.fp_in = c->x.p;
    d.fp_out = c->ratio.p;
    d.fd = c->fp_data();
    d.list_off = c->list_off.p;
    d.entries = c->entries.p;
    d.bp_in = c->ratio.p;
    d.bp_out = c->img2.p;
    d.sens = sl->p;
    d.estimate = c->x.p;
    d.d = c->dims;
    d.ppa = c->uniform_ppa;
    d.thr_lo = c->min_rel;
    d.thr_hi = c->max_rel;
    b->sens_lists.push_back(std::move(sl));
  }
  b->descs.alloc(np);
  HIPCHECK(hipMemcpy(b->descs.p, h.data(), np * sizeof(mi::ProblemDesc), hipMemcpyHostToDevice));
  // task lists of every problem (full-wave forward items index them)
  {
    std::vector<const mi::FpTask*> tl(np);
    for (int p = 0; p < np; ++p) tl[p] = b->probs[p]->tasks.p;
    b->task_lists.alloc(np);
    HIPCHECK(hipMemcpy(b->task_lists.p, tl.data(), np * sizeof(const mi::FpTask*), hipMemcpyHostToDevice));
  }
  auto upload = [](const std::vector<mi::WorkItem>& v) {
    auto d = std::make_unique<DBuf<mi::WorkItem>>();
    if (!v.empty()) {
      d->alloc(v.size());
      HIPCHECK(hipMemcpy(d->p, v.data(), v.size() * sizeof(mi::WorkItem), hipMemcpyHostToDevice));
    }
    return d;
  };
  // forward work, per subset: tasks (single families or +/- pairs) with more
  // than fp_group_max bins run a full wave each; the families of the others
  // go to lane groups of the smallest size that holds them
  b->fp_n.assign(b->nsub, {0, 0, 0, 0});
  b->fp_work.clear();
  b->fp_work.resize(b->nsub);
  b->fpw_n.assign(b->nsub, {0, 0, 0, 0});
  b->fpw_work.clear();
  b->fpw_work.resize(b->nsub);
  b->rg_n.assign(b->nsub, 0);
  b->rg_items.clear();
  b->rg_fams.clear();
  for (int s = 0; s < b->nsub; ++s) {
    std::array<std::vector<mi::WorkItem>, 4> lists;
    std::array<std::vector<mi::WorkItem>, 4> wide;
    std::vector<mi::RowGroupItem> rg;
    std::vector<int32_t> rg_fam;
    for (int p = 0; p < np; ++p) {
      const mi_ctx* c = b->probs[p];
      std::vector<mi::FpTask> ht(c->task_off[s + 1] - c->task_off[s]);
      if (!ht.empty())
        HIPCHECK(hipMemcpy(ht.data(), c->tasks.p + c->task_off[s], ht.size() * sizeof(mi::FpTask), hipMemcpyDeviceToHost));
      // families sharing a basic row, in order of first appearance; keyed by
      // start and length (an empty row starts where the next row does)
      std::unordered_map<int64_t, int> group_of_row;
      std::vector<std::vector<int>> groups;
      for (size_t i = 0; i < ht.size(); ++i) {
        const mi::FpTask& t = ht[i];
        if (b->fp_rowgroup_max > 0 && c->h_fams[t.first].nax <= std::min(b->fp_rowgroup_max, 64)) {
          for (int f : {t.first, t.second}) {
            if (f < 0) continue;
            const int64_t key = (static_cast<int64_t>(c->h_fams[f].row_start) << 32) | c->h_fams[f].row_len;
            auto it = group_of_row.emplace(key, static_cast<int>(groups.size()));
            if (it.second) groups.emplace_back();
            groups[it.first->second].push_back(f);
          }
          continue;
        }
        const int total = c->h_fams[t.first].nax + (t.second >= 0 ? c->h_fams[t.second].nax : 0);
        if (total > b->fp_group_max) {
          const int nslot = std::min((total + 63) / 64, 4);
          wide[nslot - 1].push_back({p, c->task_off[s] + static_cast<int>(i)});
          continue;
        }
        for (int f : {t.first, t.second}) {
          if (f < 0) continue;
          const int gi = std::min(group_index_at_least(c->h_fams[f].nax) + b->fp_widen, 3);
          lists[gi].push_back({p, f});
        }
      }
      // pack each row group into items of at most 64 axial positions
      for (const auto& g : groups) {
        mi::RowGroupItem cur{p, static_cast<int32_t>(rg_fam.size()), 0, 0};
        int lanes = 0;
        for (int f : g) {
          const int n = c->h_fams[f].nax;
          if (lanes + n > 64 && cur.count > 0) {
            rg.push_back(cur);
            cur = mi::RowGroupItem{p, static_cast<int32_t>(rg_fam.size()), 0, 0};
            lanes = 0;
          }
          rg_fam.push_back(f);
          ++cur.count;
          lanes += n;
        }
        if (cur.count > 0) rg.push_back(cur);
      }
    }
    for (int gi = 0; gi < 4; ++gi) {
      b->fp_n[s][gi] = static_cast<int>(lists[gi].size());
      b->fp_work[s][gi] = upload(lists[gi]);
    }
    for (int k = 0; k < 4; ++k) {
      b->fpw_n[s][k] = static_cast<int>(wide[k].size());
      b->fpw_work[s][k] = upload(wide[k]);
    }
    b->rg_n[s] = static_cast<int>(rg.size());
    auto di = std::make_unique<DBuf<mi::RowGroupItem>>();
    auto df = std::make_unique<DBuf<int32_t>>();
    if (!rg.empty()) {
      di->alloc(rg.size());
      HIPCHECK(hipMemcpy(di->p, rg.data(), rg.size() * sizeof(mi::RowGroupItem), hipMemcpyHostToDevice));
      df->alloc(rg_fam.size());
      HIPCHECK(hipMemcpy(df->p, rg_fam.data(), rg_fam.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    b->rg_items.push_back(std::move(di));
    b->rg_fams.push_back(std::move(df));
  }
  // back-projection work: the columns of every problem in tile order; images
  // of at most 32 planes in lane groups (when packing), the others one column per wave
  std::array<std::vector<mi::WorkItem>, 4> lists;
  std::array<std::vector<mi::WorkItem>, 18> wide;
  for (int p = 0; p < np; ++p) {
    const mi::ImageDims& d = b->probs[p]->dims;
    const int gi = group_index_at_least(d.nzp);
    const bool grouped = b->bp_pack == 1 && gi >= 0 && gi <= 2;
    const int ppa = b->probs[p]->uniform_ppa;
    int variant = 1 + std::min((d.nzp + 63) / 64, 4);  // slots
    if (b->bp_pack == 2 && d.nzp <= 16) variant = 0;
    else if (b->bp_pack == 2 && d.nzp <= 32) variant = 1;
    const int wclass = (ppa == 1 || ppa == 2 ? ppa : 0) * 6 + variant;
    const int tx = std::max(b->tile_x, 1), ty = std::max(b->tile_y, 1);
    for (int y0 = 0; y0 < d.ny; y0 += ty)
      for (int x0 = 0; x0 < d.nx; x0 += tx)
        for (int y = y0; y < y0 + ty; ++y)
          for (int x = x0; x < x0 + tx; ++x)
            if (y < d.ny && x < d.nx) {
              const mi::WorkItem w{p, y * d.nx + x};
              if (grouped) lists[std::max(gi, 1)].push_back(w);
              else wide[wclass].push_back(w);
            }
  }
  for (int gi = 0; gi < 4; ++gi) {
    b->bp_n[gi] = static_cast<int>(lists[gi].size());
    b->bp_work[gi] = upload(lists[gi]);
  }
  for (int k = 0; k < 18; ++k) {
    b->bpw_n[k] = static_cast<int>(wide[k].size());
    b->bpw_work[k] = upload(wide[k]);
  }
  b->lists_ready = true;
}

void batch_fp(mi_batch* b, mi::FpMode mode, int s) {
  mi::launch_fp_rowgroup(mode, b->descs.p, b->rg_items[s]->p, b->rg_fams[s]->p, b->rg_n[s], b->stream);
  for (int k = 0; k < 4; ++k)
    mi::launch_fp_batch(mode, k + 1, b->descs.p, b->task_lists.p, b->fpw_work[s][k]->p, b->fpw_n[s][k], b->stream);
  for (int gi = 0; gi < 4; ++gi)
    mi::launch_fp_group(mode, kGroupSizes[gi], b->descs.p, b->fp_work[s][gi]->p, b->fp_n[s][gi], b->stream);
}
void batch_bp(mi_batch* b, mi::BpMode mode, int s, int threshold) {
  static const int kVariantSlots[6] = {-4, -2, 1, 2, 3, 4};
  for (int k = 0; k < 18; ++k)
    mi::launch_bp_batch(mode, k / 6, kVariantSlots[k % 6], b->descs.p, b->bpw_work[k]->p, b->bpw_n[k], s, threshold,
                        b->stream);
  for (int gi = 1; gi < 4; ++gi)
    mi::launch_bp_group(mode, kGroupSizes[gi], b->descs.p, b->bp_work[gi]->p, b->bp_n[gi], s, threshold, b->stream);
}
void batch_subiterations(mi_batch* b, int start_subiteration_num, int num_subiterations, int start_subset) {
  for (int k = start_subiteration_num; k < start_subiteration_num + num_subiterations; ++k) {
    const int s = (k + start_subset - 1) % b->nsub;  // IterativeReconstruction.cxx:630-639
    batch_fp(b, mi::FP_RATIO, s);
    batch_bp(b, mi::BP_OSEM_UPDATE, s, k != 1);
  }
}
mi_ctx* batch_problem(mi_batch* b, int i) {
  if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
  if (i < 0 || i >= static_cast<int>(b->probs.size())) throw ApiError(MI_ERR_ARGUMENT, "problem index out of range");
  return b->probs[i];
}
}  // namespace

extern "C" {

int mi_batch_create(int device, int n, const mi_geometry* geometries, int num_subsets, mi_batch** out) {
  return guarded([&] {
    if (!out || !geometries || n < 1) throw ApiError(MI_ERR_ARGUMENT, "bad argument");
    auto b = std::make_unique<mi_batch>();
    b->device = device;
    b->nsub = num_subsets;
    HIPCHECK(hipSetDevice(device));
    HIPCHECK(hipStreamCreate(&b->stream));
    for (int i = 0; i < n; ++i) {
      if (geometries[i].num_tof_bins > 1) throw ApiError(MI_ERR_ARGUMENT, "batched problems are non-TOF");
      mi_ctx* c = nullptr;
      if (mi_ctx_create(device, &c) != MI_OK) throw ApiError(MI_ERR_DEVICE, g_last_error);
      b->probs.push_back(c);
      c->slab_allowed = false;  // the batch kernels read the column table
      if (mi_set_up(c, &geometries[i], num_subsets) != MI_OK)
        throw ApiError(MI_ERR_ARGUMENT, "problem " + std::to_string(i) + ": " + g_last_error);
    }
    *out = b.release();
  });
}

void mi_batch_destroy(mi_batch* b) { delete b; }

int mi_batch_context(mi_batch* b, int problem, mi_ctx** ctx) {
  return guarded([&] {
    if (!ctx) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    *ctx = batch_problem(b, problem);
  });
}

int mi_batch_set_data(mi_batch* b, int problem, const float* measured, const float* additive, float additive_constant,
                      int zero_end_planes) {
  return guarded([&] {
    mi_ctx* c = batch_problem(b, problem);
    HIPCHECK(hipSetDevice(b->device));
    batch_invalidate(b);
    if (mi_obj_set_data(c, measured, additive, additive_constant, zero_end_planes) != MI_OK ||
        mi_obj_compute_sensitivities(c) != MI_OK)
      throw ApiError(MI_ERR_DEVICE, g_last_error);
  });
}

int mi_batch_set_estimate(mi_batch* b, int problem, const float* image, int enforce_positivity) {
  return guarded([&] {
    mi_ctx* c = batch_problem(b, problem);
    HIPCHECK(hipSetDevice(b->device));
    batch_wait(b);
    const bool had = c->have_x;
    if (mi_obj_set_estimate(c, image, enforce_positivity) != MI_OK) throw ApiError(MI_ERR_DEVICE, g_last_error);
    if (!had) batch_invalidate(b);  // the estimate buffer is allocated on first use
  });
}

int mi_batch_get_estimate(mi_batch* b, int problem, float* image) {
  return guarded([&] {
    mi_ctx* c = batch_problem(b, problem);
    HIPCHECK(hipStreamSynchronize(b->stream));
    if (mi_obj_get_estimate(c, image) != MI_OK) throw ApiError(MI_ERR_DEVICE, g_last_error);
  });
}

int mi_batch_set_launch(mi_batch* b, int fp_group_max, int fp_widen, int bp_pack, int tile_x, int tile_y) {
  return guarded([&] {
    if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (fp_group_max < 0 || fp_group_max > 64) throw ApiError(MI_ERR_ARGUMENT, "fp_group_max must be in [0, 64]");
    b->fp_group_max = fp_group_max;
    b->fp_widen = fp_widen ? 1 : 0;
    if (bp_pack < 0 || bp_pack > 2) throw ApiError(MI_ERR_ARGUMENT, "bp_pack must be 0, 1 or 2");
    b->bp_pack = bp_pack;
    b->tile_x = std::max(tile_x, 1);
    b->tile_y = std::max(tile_y, 1);
    batch_invalidate(b);
  });
}

int mi_batch_set_row_groups(mi_batch* b, int max_axial) {
  return guarded([&] {
    if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (max_axial < 0 || max_axial > 64) throw ApiError(MI_ERR_ARGUMENT, "max_axial must be in [0, 64]");
    b->fp_rowgroup_max = max_axial;
    batch_invalidate(b);
  });
}

int mi_batch_tune(mi_batch* b, int* choice) {
  return guarded([&] {
    if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    HIPCHECK(hipSetDevice(b->device));
    batch_wait(b);
    struct Cand { int rg, gmax, widen, pack, tx, ty; };
    const Cand cands[] = {{32, 32, 0, 2, 32, 8},  {32, 32, 0, 0, 32, 8}, {16, 32, 0, 2, 32, 8}, {0, 32, 0, 2, 32, 8},
                          {0, 0, 0, 0, 32, 8},    {32, 32, 0, 2, 8, 8},  {32, 32, 0, 2, 16, 16}, {32, 32, 0, 1, 32, 8},
                          {16, 32, 0, 2, 8, 64},  {16, 32, 0, 2, 8, 32}};
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    float best = 0;
    int best_i = 0;
    const int ncand = static_cast<int>(sizeof(cands) / sizeof(cands[0]));
    for (int i = 0; i < ncand; ++i) {
      b->fp_rowgroup_max = cands[i].rg;
      b->fp_group_max = cands[i].gmax;
      b->fp_widen = cands[i].widen;
      b->bp_pack = cands[i].pack;
      b->tile_x = cands[i].tx;
      b->tile_y = cands[i].ty;
      batch_invalidate(b);
      batch_prepare(b);
      // side-effect free timing: forward projection into the ratio buffers,
      // plain back projection into scratch images, subset 0
      batch_fp(b, mi::FP_PLAIN, 0);
      batch_bp(b, mi::BP_PLAIN, 0, 0);
      HIPCHECK(hipEventRecord(e0, b->stream));
      for (int r = 0; r < 3; ++r) {
        batch_fp(b, mi::FP_PLAIN, 0);
        batch_bp(b, mi::BP_PLAIN, 0, 0);
      }
      HIPCHECK(hipEventRecord(e1, b->stream));
      HIPCHECK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
      if (i == 0 || ms < best) { best = ms; best_i = i; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->fp_rowgroup_max = cands[best_i].rg;
    b->fp_group_max = cands[best_i].gmax;
    b->fp_widen = cands[best_i].widen;
    b->bp_pack = cands[best_i].pack;
    b->tile_x = cands[best_i].tx;
    b->tile_y = cands[best_i].ty;
    batch_invalidate(b);
    if (choice) *choice = best_i;
  });
}

int mi_batch_osem_run(mi_batch* b, int start_subiteration_num, int num_subiterations, int start_subset, int use_graph) {
  return guarded([&] {
    if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    if (num_subiterations < 0 || start_subiteration_num < 1) throw ApiError(MI_ERR_ARGUMENT, "bad subiteration range");
    HIPCHECK(hipSetDevice(b->device));
    // problems' own streams may still be writing data or estimates
    for (mi_ctx* c : b->probs) HIPCHECK(hipStreamSynchronize(c->stream));
    batch_prepare(b);
    if (!use_graph) {
      batch_subiterations(b, start_subiteration_num, num_subiterations, start_subset);
    } else {
      const int key[3] = {start_subiteration_num, num_subiterations, start_subset};
      if (!b->graph_exec || std::memcmp(key, b->graph_key, sizeof(key)) != 0) {
        if (b->graph_exec) {
          batch_wait(b);  // the previous graph may still be running
          HIPCHECK(hipGraphExecDestroy(b->graph_exec));
          b->graph_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        batch_subiterations(b, start_subiteration_num, num_subiterations, start_subset);
        HIPCHECK(hipStreamEndCapture(b->stream, &g));
        const hipError_t e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
        std::memcpy(b->graph_key, key, sizeof(key));
      }
      HIPCHECK(hipGraphLaunch(b->graph_exec, b->stream));
    }
    HIPCHECK(hipGetLastError());
  });
}

int mi_batch_synchronize(mi_batch* b) {
  return guarded([&] {
    if (!b) throw ApiError(MI_ERR_ARGUMENT, "NULL argument");
    HIPCHECK(hipStreamSynchronize(b->stream));
  });
}

}  // extern "C"
