// Host-side plan of the ray-tracing-matrix projector pair.
//
// Everything here is set-up work done once per geometry on the host: the
// projection geometry of ProjDataInfoCylindrical, the PET Cartesian-grid
// symmetries, the OSEM subsets, and the grouping of bins into "families".
// A family is a (segment, view, tangential position) whose bins over all
// axial positions share one basic row up to an integer z shift
// (do_symmetry_shift_z), so the device processes a family as one unit with
// lanes over axial positions.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "stir_amd_pmrt.h"

namespace mi {

// TOF bins the kernels support
constexpr int kMaxTofBins = 16;

// In-plane part of a symmetry operation: where (x,y) of a basic-row element
// goes. Named after SymmetryOperations_PET_CartesianGrid.h:48-464.
enum InPlane : int8_t {
  IP_ID = 0,       // (x, y)
  IP_XMX = 1,      // (-x, y)
  IP_YMY = 2,      // (x, -y)
  IP_XMX_YMY = 3,  // (-x, -y)
  IP_XY_YX = 4,    // (y, x)
  IP_XMY_YX = 5,   // (-y, x)
  IP_XY_YMX = 6,   // (y, -x)
  IP_XMY_YMX = 7   // (-y, -x)
};

// z' = zsign * z + zadd (z_shift included in zadd)
struct SymOp {
  int inplane = IP_ID;
  int zsign = 1;
  int zadd = 0;
};

// Per-basic-bin ray parameters handed to the device tracer. All values are
// produced with the reference's float evaluation order
// (ProjMatrixByBinUsingRayTracing.cxx:582-732).
struct BasicRay {
  float s;            // s_in_mm
  float t;            // t_in_mm
  float cphi, sphi;   // cos/sin(phi), float libm
  float costheta;     // 1/sqrt(1+tan^2) (float)
  float tantheta;
  float offset_in_z;  // incl. the -0.1 voxel nudge for direct rays on a plane boundary
  float samp_z;       // sampling_distance_of_adjacent_LORs_z
  float s_inc;        // tangential LOR spacing (multiple tangential LORs)
  int32_t num_lors;   // num_lors_per_axial_pos
  int32_t pad0, pad1;
};

// Global tracing constants.
struct TraceConsts {
  float vz, vy, vx;
  float fovrad;
  float origin_z_over_vz;  // origin.z()/voxel_size.z()
  float zmid;              // (max_index.z()+min_index.z())/2.F
  int32_t restrict_cyl;
  int32_t num_tlors;
};

// One family as the device sees it.
struct Family {
  int32_t sino_base;  // device projection-data offset of the first axial position
  int32_t row_start;  // first element of the basic row in the element pool
  int32_t row_len;
  int32_t zadd0;      // z' = zsign*z + zadd0 + ppa*a  (a = local axial index)
  int16_t nax;        // number of axial positions in the family
  int8_t inplane;
  int8_t zsign;
  int16_t ppa;        // planes per axial position
  int16_t pad0;
  int32_t viewgram;   // index of the (segment, view) viewgram, for per-viewgram thresholds
  int32_t pad1;
};
static_assert(sizeof(Family) == 32, "Family layout");

// Mid-point and unit direction (z, y, x, scanner coordinates, mm) of the LOR
// of a family's basic bin, as ProjMatrixByBin::apply_tof_kernel forms them
// (ProjMatrixByBin.inl:117-137). TOF only.
struct FamilyLor {
  float mz, my, mx;
  float uz, uy, ux;
};

// Device-layout position of a family in ProjDataInMemory order, for layout
// conversion and for mi_get_row.
struct FamilyKey {
  int32_t seg, view, tang, a0;
  int32_t basic_index;  // index into basic rays
};

class Plan {
 public:
  // Throws std::runtime_error on unsupported geometry.
  Plan(const mi_geometry& g, int num_subsets);

  const mi_geometry& geometry() const { return g_; }
  int num_subsets() const { return nsub_; }
  int nseg() const { return nseg_; }
  int max_tang() const { return max_tang_; }
  int nz_padded() const { return nzp_; }
  int64_t num_bins() const { return num_bins_; }
  int64_t num_voxels() const { return (int64_t)g_.nz * g_.ny * g_.nx; }
  int num_columns() const { return g_.ny * g_.nx; }
  int num_axial(int seg) const { return g_.num_axial_poss[seg - g_.min_segment_num]; }
  const std::vector<int>& segment_sequence() const { return seg_seq_; }
  int64_t segment_offset(int seg) const { return seg_off_[seg - g_.min_segment_num]; }
  void symmetry_flags(int32_t* out) const;

  // geometry getters (ProjDataInfoCylindrical*)
  float get_s(int tang) const;
  float get_phi(int view) const;
  float get_tantheta(int seg, int tang) const;
  float get_costheta(int seg, int tang) const;
  float get_axial_sampling(int seg) const;
  float get_m(int seg, int ax) const;
  float get_t(int seg, int ax, int tang) const;
  float get_sampling_in_s(int tang) const;

  // symmetries (DataSymmetriesForBins_PET_CartesianGrid)
  bool basic_view_segment(int& view, int& seg) const;  // returns true if changed
  void basic_bin(int& seg, int& view, int& ax, int& tang) const;
  SymOp symmetry_op(int seg, int view, int ax, int tang) const;
  std::vector<std::pair<int, int>> related_view_segments(int view, int seg) const;  // (view, seg)
  std::vector<std::pair<int, int>> subset_view_segments(int subset) const;          // (view, seg)
  bool subsets_balanced() const;

  // device projection-data offset of a bin
  int64_t device_bin_offset(int seg, int view, int ax, int tang) const {
    return segment_offset(seg) + ((int64_t)view * g_.num_tangential_poss + (tang - g_.min_tangential_pos_num)) * num_axial(seg) + ax;
  }
  int64_t stir_bin_offset(int seg, int view, int ax, int tang) const {
    return segment_offset(seg) + ((int64_t)ax * g_.num_views + view) * g_.num_tangential_poss + (tang - g_.min_tangential_pos_num);
  }

  // basic rays and families
  const std::vector<BasicRay>& basic_rays() const { return rays_; }
  const std::vector<std::array<int, 4>>& basic_bins() const { return basic_bins_; }
  const TraceConsts& trace_consts() const { return tc_; }
  int basic_index(int seg, int view, int ax, int tang) const;  // -1 if not basic
  std::vector<Family>& families() { return fams_; }
  const std::vector<Family>& families() const { return fams_; }
  const std::vector<FamilyKey>& family_keys() const { return fam_keys_; }
  // families of subset s are [subset_family_begin(s), subset_family_begin(s+1))
  int subset_family_begin(int s) const { return sub_fam_off_[s]; }
  // viewgram index of (segment, view): segment-sequence position * num_views + view
  int viewgram_index(int seg, int view) const { return seg_pos_[seg - g_.min_segment_num] * g_.num_views + view; }
  int num_viewgrams() const { return nseg_ * g_.num_views; }
  // upper bound on the length of any basic row (tracer scratch sizing)
  int row_length_bound() const { return row_bound_; }

  // time of flight: number of TOF bins (1 = non-TOF), the LOR of every
  // family (TOF only), and the scanner z of image plane 0
  int num_tof_bins() const { return ntof_; }
  const std::vector<FamilyLor>& family_lors() const { return fam_lor_; }
  float tof_z_offset() const { return z_off_; }
  void tof_lor(int seg, int view, int ax, int tang, FamilyLor& out) const;

 private:
  int find_transform_z(int seg, int ax) const;
  SymOp op_bin0(int seg, int view, int ax) const;
  SymOp op_general(int s, int seg, int view, int ax) const;
  void build_basic_rays();
  void build_families();

  mi_geometry g_;
  int nsub_;
  int nseg_;
  int max_tang_;
  int max_z_, max_y_, max_x_;
  int nzp_;
  float dphi_, phi_offset_, dalpha_, radius_, ring_spacing_;
  std::vector<float> ax_samp_, m_offset_, delta_, ap2z_;
  std::vector<int> ppa_;
  int ppr_;
  bool sym90_, sym180_, symseg_, syms_, symz_;
  bool actual_bounds_ = false;  // use_actual_detector_boundaries after the set-up's checks
  int ntof_ = 1;
  float z_off_ = 0.F;
  std::vector<FamilyLor> fam_lor_;
  std::vector<int> seg_seq_;
  std::vector<int64_t> seg_off_;
  std::vector<int> seg_pos_;  // position of a segment in the storage sequence
  int64_t num_bins_;
  TraceConsts tc_;
  std::vector<BasicRay> rays_;
  std::vector<std::array<int, 4>> basic_bins_;  // (seg, view, ax, tang)
  std::vector<int32_t> basic_lookup_;          // (seg, view, ax?, tang) -> basic index
  int lookup_na_ = 1;
  int tmax_ = 0;
  std::vector<Family> fams_;
  std::vector<FamilyKey> fam_keys_;
  std::vector<int> sub_fam_off_;
  int row_bound_;
};

}  // namespace mi
