// Box rules shared by the host and the GPU BVH builders (host and device code).
#pragma once
#include "lrt_math.h"

namespace lrt {

// Box4f::surfaceArea (LiteMath.h:3440-3444)
LRT_HD float box_surface_area(float3 mn, float3 mx) {
  const float3 d = mx - mn;
  return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

// RelativeBoxOverlap (cbvh_embree2.cpp:39-48) over BoxBoxOverlap (LiteMath.h:3535-3562): the
// overlap box's area relative to the smaller box, at most 1; a disjoint pair overlaps in a
// zero-size box. The leaf-merge rule of both builders compares it with the over-split threshold.
LRT_HD float relative_box_overlap(float3 mn1, float3 mx1, float3 mn2, float3 mx2) {
  const float a1[3] = {mn1.x, mn1.y, mn1.z}, b1[3] = {mx1.x, mx1.y, mx1.z};
  const float a2[3] = {mn2.x, mn2.y, mn2.z}, b2[3] = {mx2.x, mx2.y, mx2.z};
  float omn[3], omx[3];
  bool disjoint = false;
  for (int a = 0; a < 3; ++a) {
    float lo1 = a1[a], hi1 = b1[a], lo2 = a2[a], hi2 = b2[a];
    if (lo2 < lo1) {
      float t = lo1; lo1 = lo2; lo2 = t;
      t = hi1; hi1 = hi2; hi2 = t;
    }
    if (hi1 < lo2) { disjoint = true; break; }
    omn[a] = lo2;
    omx[a] = fmin_ref(hi1, hi2);
  }
  const float sa1 = box_surface_area(mn1, mx1), sa2 = box_surface_area(mn2, mx2);
  const float sa3 = disjoint ? box_surface_area(mx1, mx1)
                             : box_surface_area(make3(omn[0], omn[1], omn[2]), make3(omx[0], omx[1], omx[2]));
  return fmin_ref(sa3 / fmin_ref(sa1, sa2), 1.0f);
}

// BuildBVHEmbree (cbvh_embree2.cpp:344-347): over-split threshold of the leaf merge.
LRT_HD float over_split_threshold(int prims_in_leaf) { return prims_in_leaf >= 4 ? 0.25f : 0.5f; }

}  // namespace lrt
