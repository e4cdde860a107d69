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

struct ImageDims {
  int32_t nz, ny, nx;
  int32_t min_z, min_y, min_x;
  int32_t nzp;  // padded z stride of the device image layout
  int32_t ncols;
};

// Forward-projection epilogues.
enum FpMode : int {
  FP_PLAIN = 0,      // out = A x
  FP_RATIO = 1,      // out = divide_and_truncate(y, A x + b)
  FP_GRADIENT = 2,   // out = divide_and_truncate(y, A x + b) - m
  FP_LOGLIK = 3      // partial sums of the Poisson log-likelihood (FP64)
};

struct FpData {
  const float* measured;      // y, device layout (end planes already zeroed when requested)
  const float* additive;      // per-bin b, or nullptr
  float additive_constant;    // b when additive == nullptr
  const float* vg_small;      // per-viewgram small_value = max(max(y_viewgram)*1e-6, 0)
  const float* mult;          // m for FP_GRADIENT (ones, zero at end planes), or nullptr (= 1)
  double* partial;            // FP_LOGLIK: one double per workgroup
};

// Back-projection epilogues.
enum BpMode : int {
  BP_PLAIN = 0,       // out = A^T p
  BP_OSEM_UPDATE = 1  // x *= clamp(divide(A^T p, sens))
};

struct BpData {
  const float* sens;  // subset sensitivity, device image layout
  float* estimate;    // x, device image layout, updated in place
  int threshold;      // apply threshold_upper_lower(0, FLT_MAX)
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
                            hipStream_t stream);
// ppa: planes per axial position shared by all families (1 or 2 select a
// specialised kernel), 0 when it varies between segments.
void launch_back_project(BpMode mode, int ppa, const int64_t* d_list_off, const BpEntry* d_entries, int subset_begin,
                         int subset_end, const float* d_proj, float* d_img, const ImageDims& dims, const BpData& data,
                         hipStream_t stream);
// 2-D transpose used for the STIR <-> device layout conversions:
// dst[c * dst_ld + r] = src[r * src_ld + c] for r < rows, c < cols.
void launch_transpose(const float* src, float* dst, int rows, int cols, int src_ld, int dst_ld, hipStream_t stream);

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

// ---- small utilities (pmrt_project.hip)
void launch_fill(float* d, int64_t n, float v, hipStream_t stream);
void launch_viewgram_max(const float* d_proj, const int64_t* d_vg_off, const int32_t* d_vg_len, int nvg, float* d_small,
                         hipStream_t stream);
void launch_zero_end_planes(float* d_proj, const int64_t* d_off, int count, int run, int stride, hipStream_t stream);

}  // namespace mi
