// Host-side plan: projection geometry, symmetries, subsets and families.
// Built with -ffp-contract=off and without fast-math: the float expressions
// follow the reference's evaluation order so that the rays handed to the
// device tracer are the reference's rays.
#include "pmrt_plan.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mi {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// stir::round(float) -- round.inl:59-66
int stir_round(float x) { return x >= 0 ? static_cast<int>(x + 0.5F) : -static_cast<int>(-x + 0.5F); }

// stir::modulo(float,float) -- modulo.h:52-79
float stir_modulo(float a, float b) {
  const double r = std::fmod(static_cast<double>(a), static_cast<double>(b));
  float res = static_cast<float>(r < 0 ? r + std::fabs(static_cast<double>(b)) : r);
  const float abs_b = b >= 0 ? b : -b;
  if (res >= abs_b) res -= abs_b;
  return res;
}

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

// Symmetry operation choice tables (DataSymmetriesForBins_PET_CartesianGrid.inl:120-396).
// Region: 0 = (90,135], 1 = (45,90], 2 = (90,180) with 180-degree symmetry, 3 = basic range.
struct OpCode {
  int8_t inplane;
  int8_t zq;
};
// general bins: [region][segment class: 0 = positive (or no segment symmetry), 1 = negative, 2 = zero][s sign: 0 = s>0 (or no s symmetry), 1 = s<0]
constexpr OpCode kGeneral[4][3][2] = {
    {{{IP_XMY_YX, 0}, {IP_XY_YMX, 1}}, {{IP_XMY_YX, 1}, {IP_XY_YMX, 0}}, {{IP_XMY_YX, 0}, {IP_XY_YMX, 0}}},
    {{{IP_XY_YX, 1}, {IP_XMY_YMX, 0}}, {{IP_XY_YX, 0}, {IP_XMY_YMX, 1}}, {{IP_XY_YX, 0}, {IP_XMY_YMX, 0}}},
    {{{IP_XMX, 1}, {IP_YMY, 0}}, {{IP_XMX, 0}, {IP_YMY, 1}}, {{IP_XMX, 0}, {IP_YMY, 1}}},
    {{{IP_ID, 0}, {IP_XMX_YMY, 1}}, {{IP_ID, 1}, {IP_XMX_YMY, 0}}, {{IP_ID, 0}, {IP_XMX_YMY, 0}}},
};
// tangential position 0: [region][0 = segment >= 0 (or no segment symmetry), 1 = segment < 0]
constexpr OpCode kBin0[4][2] = {
    {{IP_XMY_YX, 0}, {IP_XMY_YX, 1}},
    {{IP_XY_YX, 1}, {IP_XY_YX, 0}},
    {{IP_XMX, 1}, {IP_XMX, 0}},
    {{IP_ID, 0}, {IP_ID, 1}},
};

}  // namespace

Plan::Plan(const mi_geometry& g, int num_subsets) : g_(g), nsub_(num_subsets) {
  if (num_subsets < 1) fail("number of subsets must be at least 1");
  nseg_ = g.max_segment_num - g.min_segment_num + 1;
  if (nseg_ < 1 || nseg_ > MI_MAX_SEGMENTS) fail("bad segment range");
  if (g.min_segment_num > 0 || g.max_segment_num < 0) fail("projection data must contain segment 0");
  if (g.do_symmetry_swap_segment && g.min_segment_num != -g.max_segment_num)
    fail("segment symmetry needs a symmetric segment range");
  if (g.num_views < 1 || g.num_tangential_poss < 1 || g.nx < 1 || g.ny < 1 || g.nz < 1) fail("empty geometry");
  if (g.num_tangential_LORs < 1) fail("num_tangential_LORs should be at least 1");
  if (g.nx > 1022 || g.ny > 1022) fail("image x/y size above 1022 is not supported by the element packing");
  if (g.nz > 256) fail("images with more than 256 planes are not supported");
  if (g.num_tof_bins > kMaxTofBins) fail("more than 16 TOF bins are not supported");
  ntof_ = g.num_tof_bins > 1 ? g.num_tof_bins : 1;
  // scanner z of image plane 0, subtracted from voxel positions in apply_tof_kernel
  // (ProjMatrixByBin.inl:128-131)
  if (ntof_ > 1) z_off_ = (g.num_rings - 1) / 2.0F * g.ring_spacing;
  max_tang_ = g.min_tangential_pos_num + g.num_tangential_poss - 1;
  max_z_ = g.min_z + g.nz - 1;
  max_y_ = g.min_y + g.ny - 1;
  max_x_ = g.min_x + g.nx - 1;
  nzp_ = (g.nz + 15) / 16 * 16;

  // ProjDataInfoCylindrical ctor (ProjDataInfoCylindrical.cxx:56-117)
  dphi_ = static_cast<float>(kPi / g.num_views);
  phi_offset_ = g.intrinsic_tilt;
  {
    const int ndet = g.num_detectors_per_ring;
    if (ndet > 2 && g.num_views * 2 != ndet && (ndet % (g.num_views * 2)) == 0) {
      const int mash = ndet / 2 / g.num_views;
      phi_offset_ += static_cast<float>(kPi / (ndet / 2) * (mash - 1) / 2.F);
    }
  }
  radius_ = g.effective_ring_radius;
  ring_spacing_ = g.ring_spacing;
  dalpha_ = static_cast<float>(kPi / g.num_detectors_per_ring);  // ProjDataInfoCylindricalNoArcCorr.cxx:77-93
  ax_samp_.resize(nseg_);
  m_offset_.resize(nseg_);
  delta_.resize(nseg_);
  for (int i = 0; i < nseg_; ++i) {
    if (g.num_axial_poss[i] < 1) fail("segment without axial positions");
    ax_samp_[i] = ring_spacing_ / (g.max_ring_diff[i] != g.min_ring_diff[i] ? 2 : 1);
    m_offset_[i] = ((g.num_axial_poss[i] - 1 + 0) * ax_samp_[i]) / 2;
    delta_[i] = (g.min_ring_diff[i] + g.max_ring_diff[i]) / 2.F;
  }
  // storage order 0,+1,-1,... (ProjData.cxx:641-661)
  seg_seq_.push_back(0);
  for (int s = 1; static_cast<int>(seg_seq_.size()) < nseg_; ++s) {
    if (s <= g.max_segment_num) seg_seq_.push_back(s);
    if (-s >= g.min_segment_num) seg_seq_.push_back(-s);
  }
  seg_off_.assign(nseg_, 0);
  seg_pos_.assign(nseg_, 0);
  int64_t off = 0;
  for (size_t k = 0; k < seg_seq_.size(); ++k) {
    const int s = seg_seq_[k];
    seg_pos_[s - g.min_segment_num] = static_cast<int>(k);
    seg_off_[s - g.min_segment_num] = off;
    off += static_cast<int64_t>(num_axial(s)) * g.num_views * g.num_tangential_poss;
  }
  num_bins_ = off;
  if (num_bins_ >= (int64_t(1) << 31)) fail("more than 2^31 bins");

  // symmetries (DataSymmetriesForBins_PET_CartesianGrid.cxx:241-378)
  sym90_ = g.do_symmetry_90degrees_min_phi != 0;
  sym180_ = g.do_symmetry_90degrees_min_phi || g.do_symmetry_180degrees_min_phi;
  symseg_ = g.do_symmetry_swap_segment != 0;
  syms_ = g.do_symmetry_swap_s != 0;
  symz_ = g.do_symmetry_shift_z != 0;
  // use_actual_detector_boundaries is reset when the data are arc-corrected,
  // view-mashed or axially compressed (ProjMatrixByBinUsingRayTracing.cxx:300-327)
  actual_bounds_ = g.use_actual_detector_boundaries && !g.arc_corrected && g.num_views * 2 == g.num_detectors_per_ring;
  for (int si = 0; si < nseg_; ++si)
    if (g.min_ring_diff[si] != g.max_ring_diff[si]) actual_bounds_ = false;
  {
    const float zo = g.origin_z / g.vox_z;
    if (std::fabs(stir_round(zo) - zo) > 1.E-3F) fail("the z origin must be a multiple of the plane separation");
    if (std::fabs(g.vox_y - g.vox_x) > 2.E-3F) sym90_ = false;
    if (g.num_views % 4 != 0) sym90_ = false;
    if (g.num_views % 2 != 0) sym180_ = false;
    if (std::fabs(get_tantheta(0, 0)) > 1.E-4F) fail("segment 0 must contain direct planes");
    for (int s = 1; s <= std::min(g.max_segment_num, -g.min_segment_num); ++s)
      if (std::fabs(get_tantheta(s, 0) + get_tantheta(-s, 0)) > 1.E-4F) fail("segments -s and s must be mirror images");
    if (g.min_tangential_pos_num < 0 && max_tang_ > 0 && std::fabs(get_s(1) + get_s(-1)) > 1.E-4F)
      fail("get_s must be odd in the tangential position");
    if (std::fabs(get_phi(0)) > 1.E-4F && (sym90_ || sym180_)) sym90_ = sym180_ = false;
    if (std::fabs(g.origin_x) > .01F || std::fabs(g.origin_y) > .01F) sym90_ = sym180_ = symseg_ = syms_ = false;
    // TOF keeps only the z shift (DataSymmetriesForBins_PET_CartesianGrid.cxx:345-366)
    if (ntof_ > 1) sym90_ = sym180_ = symseg_ = syms_ = false;
  }
  // ProjMatrixByBinUsingRayTracing.cxx:276-277
  if (std::fabs(g.origin_x) > .05F || std::fabs(g.origin_y) > .05F) fail("shifted x/y origin is not supported");
  // find_relation_between_coordinate_systems (DataSymmetriesForBins_PET_CartesianGrid.cxx:51-125)
  {
    const float ips = g.vox_z;
    const float pprf = ring_spacing_ / ips;
    ppr_ = stir_round(pprf);
    if (std::fabs(pprf - ppr_) > 1.E-2) fail("z-grid spacing must be the ring spacing divided by an integer");
    ppa_.resize(nseg_);
    ap2z_.resize(nseg_);
    for (int i = 0; i < nseg_; ++i) {
      const float f = ax_samp_[i] / ips;
      ppa_[i] = stir_round(f);
      if (std::fabs(f - ppa_[i]) > 1.E-2) fail("z-grid spacing must be the axial sampling divided by an integer");
      ap2z_[i] = (max_z_ + g.min_z) / 2.F - g.origin_z / ips - (ppa_[i] * (g.num_axial_poss[i] - 1 + 0) + ppr_ * delta_[i]) / 2;
    }
  }
  build_basic_rays();
  build_families();
}

void Plan::symmetry_flags(int32_t* out) const {
  out[0] = sym90_;
  out[1] = sym180_;
  out[2] = symseg_;
  out[3] = syms_;
  out[4] = symz_;
}

// ---------------------------------------------------------------------------
// geometry (ProjDataInfoCylindrical.inl:65-138, ProjDataInfo.inl:200-203,
// ProjDataInfoCylindricalNoArcCorr.inl:79-82, ProjDataInfoCylindricalArcCorr.inl:31-34).
// get_s/get_tantheta/get_costheta use the double libm overloads (the headers
// defining them are included before <math.h> in the reference's translation
// units); the ray tracer's own cos/sin/sqrt use the float overloads.
float Plan::get_s(int tang) const {
  if (g_.arc_corrected) return tang * g_.bin_size;
  return static_cast<float>(static_cast<double>(radius_) * std::sin(static_cast<double>(tang * dalpha_)));
}
float Plan::get_phi(int view) const { return view * dphi_ + phi_offset_; }
float Plan::get_axial_sampling(int seg) const { return ax_samp_[seg - g_.min_segment_num]; }
float Plan::get_tantheta(int seg, int tang) const {
  const float delta = delta_[seg - g_.min_segment_num];
  if (std::fabs(delta) < 0.0001F) return 0;
  const float s = get_s(tang);
  const float num = delta * ring_spacing_;
  return static_cast<float>(num / (2 * std::sqrt(static_cast<double>(radius_ * radius_ - s * s))));
}
float Plan::get_costheta(int seg, int tang) const {
  const float tt = get_tantheta(seg, tang);
  return static_cast<float>(1 / std::sqrt(static_cast<double>(1 + tt * tt)));
}
float Plan::get_m(int seg, int ax) const { return ax * get_axial_sampling(seg) - m_offset_[seg - g_.min_segment_num]; }
float Plan::get_t(int seg, int ax, int tang) const { return get_m(seg, ax) * get_costheta(seg, tang); }
float Plan::get_sampling_in_s(int tang) const {
  if (g_.arc_corrected) return g_.bin_size;
  return std::fabs(get_s(tang + 1) - get_s(tang - 1)) / 2;  // ProjDataInfo.cxx:110-115
}

// ---------------------------------------------------------------------------
// symmetries
int Plan::find_transform_z(int seg, int ax) const {
  // DataSymmetriesForBins_PET_CartesianGrid.inl:67-118
  const int i = seg - g_.min_segment_num;
  const float tz = (2 * ppa_[i] * (ax) + ppr_ * delta_[i] + 2 * ap2z_[i]);
  return static_cast<int>(std::floor(tz + 0.5));
}

static int view_region(int view, int nv, bool sym90, bool sym180) {
  const int v135 = nv / 4 * 3, v90 = nv / 2, v45 = nv / 4;
  if (sym90 && view > v90 && view <= v135) return 0;
  if (sym90 && view > v45 && view <= v90) return 1;
  if (sym180 && view > v90) return 2;
  return 3;
}

SymOp Plan::op_bin0(int seg, int view, int ax) const {
  const int q = find_transform_z(std::abs(seg), symz_ ? 0 : ax);
  const int z_shift = symz_ ? ppa_[seg - g_.min_segment_num] * ax : 0;
  const int region = view_region(view, g_.num_views, sym90_, sym180_);
  const OpCode c = kBin0[region][(symseg_ && seg < 0) ? 1 : 0];
  SymOp o;
  o.inplane = c.inplane;
  o.zsign = c.zq ? -1 : 1;
  o.zadd = c.zq ? q + z_shift : z_shift;
  return o;
}

SymOp Plan::op_general(int s, int seg, int view, int ax) const {
  const int q = find_transform_z(std::abs(seg), symz_ ? 0 : ax);
  const int z_shift = symz_ ? ppa_[seg - g_.min_segment_num] * ax : 0;
  const int region = view_region(view, g_.num_views, sym90_, sym180_);
  int segclass;
  if (!symseg_ || seg > 0) segclass = 0;
  else if (seg < 0) segclass = 1;
  else segclass = 2;
  // in the basic region the s-sign condition is "swap_s && s < 0"; elsewhere "!( !swap_s || s > 0 )"
  const int sclass = (region == 3) ? ((syms_ && s < 0) ? 1 : 0) : ((!syms_ || s > 0) ? 0 : 1);
  const OpCode c = kGeneral[region][segclass][sclass];
  SymOp o;
  o.inplane = c.inplane;
  o.zsign = c.zq ? -1 : 1;
  o.zadd = c.zq ? q + z_shift : z_shift;
  return o;
}

SymOp Plan::symmetry_op(int seg, int view, int ax, int tang) const {
  // find_symmetry_operation_from_basic_bin (.inl:664-674)
  return tang == 0 ? op_bin0(seg, view, ax) : op_general(tang, seg, view, ax);
}

bool Plan::basic_view_segment(int& view, int& seg) const {
  // find_basic_view_segment_numbers (.inl:398-447)
  bool change = false;
  const int nv = g_.num_views, v90 = nv >> 1, v45 = v90 >> 1, v135 = v90 + v45;
  if (symseg_ && seg < 0) {
    seg = -seg;
    change = true;
  }
  if (sym90_) {
    if (view >= v135) { view = nv - view; return true; }
    if (view >= v90) { view = view - v90; return true; }
    if (view > v45) { view = v90 - view; return true; }
  } else if (sym180_) {
    if (view > v90) { view = nv - view; return true; }
  }
  return change;
}

void Plan::basic_bin(int& seg, int& view, int& ax, int& tang) const {
  basic_view_segment(view, seg);  // (.inl:449-478)
  if (syms_ && tang < 0) tang = -tang;
  if (symz_) ax = 0;
}

std::vector<std::pair<int, int>> Plan::related_view_segments(int view, int seg) const {
  // get_related_view_segment_numbers (.inl:800-849)
  const int nv = g_.num_views;
  const bool symz = symseg_ && seg != 0;
  std::vector<std::pair<int, int>> out;
  auto add = [&](int v) {
    out.emplace_back(v, seg);
    if (symz) out.emplace_back(v, -seg);
  };
  add(view);
  if (sym180_ && sym90_ && (view % (nv / 2)) != nv / 4) add(view < nv / 2 ? view + nv / 2 : view - nv / 2);
  if (sym180_ && (view % (nv / 2)) != 0) add(nv - view);
  if (sym90_ && (view % (nv / 4)) != 0) add((nv / 2 - view + nv) % nv);
  // the segment-pair order of the reference: (v,s),(v,-s),(v2,s),(v2,-s),...
  return out;
}

std::vector<std::pair<int, int>> Plan::subset_view_segments(int subset) const {
  // find_basic_vs_nums_in_subset.cxx:31-69, then the related view-segments of each
  std::vector<std::pair<int, int>> out;
  for (int seg = g_.min_segment_num; seg <= g_.max_segment_num; ++seg)
    for (int view = subset; view < g_.num_views; view += nsub_) {
      int v = view, s = seg;
      if (basic_view_segment(v, s)) continue;
      for (auto& p : related_view_segments(view, seg)) out.push_back(p);
    }
  return out;
}

bool Plan::subsets_balanced() const {
  // ...AndProjData.cxx:490-529: equal numbers of segment-0 viewgrams per subset
  int first = -1;
  for (int sub = 0; sub < nsub_; ++sub) {
    int cnt = 0;
    for (int view = sub; view < g_.num_views; view += nsub_) {
      int v = view, s = 0;
      if (basic_view_segment(v, s)) continue;
      cnt += static_cast<int>(related_view_segments(view, 0).size());
    }
    if (first < 0) first = cnt;
    else if (cnt != first) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
int Plan::basic_index(int seg, int view, int ax, int tang) const {
  const int na = lookup_na_;
  const int ai = symz_ ? 0 : ax;
  if (tang < -tmax_ || tang > tmax_) return -1;
  const size_t idx = ((static_cast<size_t>(seg - g_.min_segment_num) * g_.num_views + view) * na + ai) * (2 * tmax_ + 1) +
                     (tang + tmax_);
  return basic_lookup_[idx];
}

void Plan::build_basic_rays() {
  const mi_geometry& g = g_;
  int max_ax = 0;
  for (int i = 0; i < nseg_; ++i) max_ax = std::max(max_ax, g.num_axial_poss[i]);
  const int na = symz_ ? 1 : max_ax;
  lookup_na_ = na;
  // the basic bin of a tangential position can lie outside the data's range
  // (-n/2 maps to +n/2 under swap_s); the reference traces it on demand
  tmax_ = std::max(max_tang_, -g.min_tangential_pos_num);
  basic_lookup_.assign(static_cast<size_t>(nseg_) * g.num_views * na * (2 * tmax_ + 1), -1);

  tc_.vz = g.vox_z;
  tc_.vy = g.vox_y;
  tc_.vx = g.vox_x;
  // fovrad (ProjMatrixByBinUsingRayTracing.cxx:731-732)
  tc_.fovrad = std::min((std::min(max_x_, -g.min_x)) * g.vox_x, (std::min(max_y_, -g.min_y)) * g.vox_y);
  tc_.origin_z_over_vz = g.origin_z / g.vox_z;
  tc_.zmid = (max_z_ + g.min_z) / 2.F;
  tc_.restrict_cyl = g.restrict_to_cylindrical_FOV;
  tc_.num_tlors = g.num_tangential_LORs;

  float max_tan = 0.F;
  int max_lors = 1;
  for (int seg = g.min_segment_num; seg <= g.max_segment_num; ++seg)
    for (int view = 0; view < g.num_views; ++view)
      for (int ax = 0; ax < (symz_ ? 1 : num_axial(seg)); ++ax)
        for (int tang = -tmax_; tang <= tmax_; ++tang) {
          int bs = seg, bv = view, ba = ax, bt = tang;
          basic_bin(bs, bv, ba, bt);
          if (bs != seg || bv != view || ba != ax || bt != tang) continue;
          // calculate_proj_matrix_elems_for_one_bin (ProjMatrixByBinUsingRayTracing.cxx:582-732)
          BasicRay r{};
          r.s = get_s(tang);
          float phi = get_phi(view);
          if (actual_bounds_) {
            // the LOR between the bin's two detectors (:613-634), with the
            // detector pair of ProjDataInfoCylindricalNoArcCorr.cxx:164-200
            // (>> of a negative int rounds towards minus infinity), in double
            // and rounded to float
            const int ndet = g.num_detectors_per_ring;
            const int det1 = (view + (tang >> 1) + ndet) % ndet;
            const int det2 = (view - ((tang + 1) >> 1) + ndet / 2) % ndet;
            phi = static_cast<float>((det1 + det2) * kPi / ndet - kPi / 2 + phi_offset_);
            r.s = static_cast<float>(g.effective_ring_radius * std::sin((det1 - det2) * kPi / ndet + kPi / 2));
          }
          r.cphi = std::cos(phi);  // float overloads
          r.sphi = std::sin(phi);
          r.tantheta = get_tantheta(seg, tang);
          r.costheta = 1 / std::sqrt(1 + r.tantheta * r.tantheta);
          r.t = get_t(seg, ax, tang);
          r.samp_z = (get_axial_sampling(seg) * get_costheta(seg, tang)) / r.costheta;
          r.num_lors = static_cast<int>(std::ceil(r.samp_z / g.vox_z - 1.E-3));
          if (std::fabs(r.samp_z / g.vox_z - r.num_lors) > 1E-3)
            fail("the axial sampling must be an integer multiple of the z voxel size");
          const float zfirst = (-r.samp_z / (2 * r.num_lors) * (r.num_lors - 1)) - g.origin_z;
          r.offset_in_z = zfirst + (max_z_ + g.min_z) / 2.F * g.vox_z;
          if (r.tantheta == 0) {
            if (std::fabs(stir_modulo((r.t + r.offset_in_z) / g.vox_z, 1.F) - .5) < .001) r.offset_in_z -= .1F * g.vox_z;
          }
          r.s_inc = (actual_bounds_ ? 2 : 1) * get_sampling_in_s(tang) / g.num_tangential_LORs;  // :760-762
          basic_lookup_[((static_cast<size_t>(seg - g.min_segment_num) * g.num_views + view) * na + (symz_ ? 0 : ax)) *
                            (2 * tmax_ + 1) +
                        (tang + tmax_)] = static_cast<int>(rays_.size());
          rays_.push_back(r);
          basic_bins_.push_back({seg, view, ax, tang});
          max_tan = std::max(max_tan, std::fabs(r.tantheta));
          max_lors = std::max(max_lors, r.num_lors);
        }
  // Upper bound of a row: per Siddon trace ceil|dz|+ceil|dy|+ceil|dx|+3
  // (RayTraceVoxelsOnCartesianGrid.cxx:73-74), times the tangential LORs and
  // the axial copies; the tracer reports overflow if this is ever exceeded.
  const double d = 2.0 * tc_.fovrad;
  const double per_trace = std::ceil(d / g.vox_x) + std::ceil(d / g.vox_y) + std::ceil(d * max_tan / g.vox_z) + 8;
  const int axial = max_lors + 3;
  row_bound_ = static_cast<int>(per_trace * g.num_tangential_LORs * 2) * axial + 64;
}

// to_0_2pi for floats: modulo(float, float) (modulo.h:62-76,136-141)
static float to_0_2pi_f(float phi) {
  const float b = static_cast<float>(2 * kPi);
  double r = std::fmod(static_cast<double>(phi), static_cast<double>(b));
  r = r < 0 ? r + std::fabs(static_cast<double>(b)) : r;
  float res = static_cast<float>(r);
  if (res >= b) res -= b;
  return res;
}

// LOR of a bin as apply_tof_kernel forms it (ProjMatrixByBin.inl:117-137):
// get_LOR (ProjDataInfoCylindrical.cxx:515-547), the sinogram -> cylinder ->
// two-point conversions (LORCoordinates.inl:106-139,287-298), mid-point and
// unit direction. Float arithmetic with the float libm overloads.
void Plan::tof_lor(int seg, int view, int ax, int tang, FamilyLor& out) const {
  const float s = get_s(tang), m = get_m(seg, ax), tt = get_tantheta(seg, tang), phi = get_phi(view);
  const float R = radius_;
  const float max_a = sqrtf(R * R - s * s), min_a = -max_a;
  float z1 = m - max_a * tt, z2 = m - min_a * tt;
  float ph = to_0_2pi_f(phi), beta = asinf(s / R);
  bool swapped = false;
  if (static_cast<double>(ph) >= kPi) {
    ph -= static_cast<float>(kPi);
    beta = -beta;
    std::swap(z1, z2);
    swapped = true;
  }
  const float psi1 = to_0_2pi_f(ph + beta), psi2 = to_0_2pi_f(ph - beta + static_cast<float>(kPi));
  float p1[3] = {z1, -R * cosf(psi1), R * sinf(psi1)};
  float p2[3] = {z2, -R * cosf(psi2), R * sinf(psi2)};
  if (swapped)
    for (int i = 0; i < 3; ++i) std::swap(p1[i], p2[i]);
  float mid[3], diff[3];
  for (int i = 0; i < 3; ++i) {
    mid[i] = (p1[i] + p2[i]) * 0.5F;
    diff[i] = p2[i] - mid[i];
  }
  float acc = 0.F;
  acc += diff[0] * diff[0];
  acc += diff[1] * diff[1];
  acc += diff[2] * diff[2];
  const float n = static_cast<float>(std::sqrt(static_cast<double>(acc)));
  out.mz = mid[0]; out.my = mid[1]; out.mx = mid[2];
  out.uz = diff[0] / n; out.uy = diff[1] / n; out.ux = diff[2] / n;
}

void Plan::build_families() {
  const mi_geometry& g = g_;
  sub_fam_off_.assign(nsub_ + 1, 0);
  for (int sub = 0; sub < nsub_; ++sub) {
    sub_fam_off_[sub] = static_cast<int>(fams_.size());
    for (auto& vs : subset_view_segments(sub)) {
      const int view = vs.first, seg = vs.second;
      const int nax = num_axial(seg);
      for (int tang = g.min_tangential_pos_num; tang <= max_tang_; ++tang) {
        for (int a0 = 0; a0 < (symz_ ? 1 : nax); ++a0) {
          int bs = seg, bv = view, ba = a0, bt = tang;
          basic_bin(bs, bv, ba, bt);
          const int bi = basic_index(bs, bv, ba, bt);
          if (bi < 0) fail("internal: basic bin not found for seg " + std::to_string(seg) + " view " + std::to_string(view) + " tang " + std::to_string(tang) + " -> " + std::to_string(bs) + "," + std::to_string(bv) + "," + std::to_string(ba) + "," + std::to_string(bt));
          const SymOp op = symmetry_op(seg, view, a0, tang);
          Family f{};
          f.sino_base = static_cast<int32_t>(device_bin_offset(seg, view, a0, tang));
          f.row_start = 0;
          f.row_len = 0;
          f.zadd0 = op.zadd;
          f.nax = static_cast<int16_t>(symz_ ? nax : 1);
          f.inplane = static_cast<int8_t>(op.inplane);
          f.zsign = static_cast<int8_t>(op.zsign);
          f.ppa = static_cast<int16_t>(ppa_[seg - g.min_segment_num]);
          f.viewgram = viewgram_index(seg, view);
          fams_.push_back(f);
          fam_keys_.push_back({seg, view, tang, a0, bi});
          if (ntof_ > 1) {
            FamilyLor lor;
            tof_lor(bs, bv, ba, bt, lor);
            fam_lor_.push_back(lor);
          }
        }
      }
    }
  }
  sub_fam_off_[nsub_] = static_cast<int>(fams_.size());
}

}  // namespace mi
