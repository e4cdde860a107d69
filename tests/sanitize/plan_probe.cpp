// Test driver for the host plan under AddressSanitizer / UBSan: builds the
// plan of a geometry and walks every table it holds (basic rays, families,
// family keys, subsets and their related view-segments, symmetry operations
// of every family's bin), returning counts the test compares with what the
// geometry implies.
#include <cstdint>
#include <cstring>
#include <exception>

#include "pmrt_plan.hpp"

extern "C" int plan_probe(const mi_geometry* g, int num_subsets, int64_t* out, char* err, int errlen) {
  try {
    mi::Plan plan(*g, num_subsets);
    int64_t nseg_axial = 0;
    for (int s = g->min_segment_num; s <= g->max_segment_num; ++s) nseg_axial += plan.num_axial(s);
    int64_t checksum = 0, related = 0;
    const auto& fams = plan.families();
    const auto& keys = plan.family_keys();
    for (size_t i = 0; i < fams.size(); ++i) {
      const mi::FamilyKey& k = keys[i];
      int seg = k.seg, view = k.view, ax = k.a0, tang = k.tang;
      plan.basic_bin(seg, view, ax, tang);
      const mi::SymOp op = plan.symmetry_op(k.seg, k.view, k.a0, k.tang);
      checksum += fams[i].sino_base + 3 * op.inplane + 7 * op.zsign + op.zadd + k.basic_index;
      if (k.basic_index < 0 || k.basic_index >= static_cast<int>(plan.basic_rays().size())) return -2;
    }
    for (int s = 0; s < num_subsets; ++s)
      for (const auto& vs : plan.subset_view_segments(s)) {
        int v = vs.first, sg = vs.second;
        plan.basic_view_segment(v, sg);
        related += static_cast<int64_t>(plan.related_view_segments(v, sg).size());
      }
    out[0] = plan.num_bins();
    out[1] = static_cast<int64_t>(plan.basic_rays().size());
    out[2] = static_cast<int64_t>(fams.size());
    out[3] = nseg_axial;
    out[4] = related;
    out[5] = checksum;
    out[6] = plan.subset_family_begin(num_subsets);
    return 0;
  } catch (const std::exception& e) {
    std::strncpy(err, e.what(), static_cast<size_t>(errlen) - 1);
    err[errlen - 1] = 0;
    return -1;
  }
}
