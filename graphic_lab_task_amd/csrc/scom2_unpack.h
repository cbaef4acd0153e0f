// Field access of the SCom2 packing (sdf/scom2/scom_shared.h:148-292), shared by the host checks
// and the kernels: node heads, link ends, child and data edges.
#pragma once
#include "lrt_types.h"

namespace lrt {

struct SCom2NodeHead {
  uint32_t base_type, children_types, base_links_end, children_active;
};
struct SCom2DataEdge {
  uint32_t data_offset, rotation_id;
  float add;
};
struct SCom2ChildEdge {
  uint32_t child_offset, rotation_id;
};

// unpack_node_head (scom_shared.h:183-200)
LRT_HD SCom2NodeHead scom2_node_head(const SCom2Header& h, uint32_t node0, uint32_t node1) {
  SCom2NodeHead r;
  r.base_type = node0 & 0xFFu;
  r.children_types = ((h.has_multi_nodes != 0 ? node0 : node1) >> h.children_types_shift) & h.children_types_mask;
  if (h.has_multi_nodes != 0)
    r.base_links_end = (node0 >> h.base_reference_shift) & h.reference_mask;
  else
    r.base_links_end = r.base_type == SCOM2_CHILD_EMPTY ? 0u : (h.has_surfaces == 0 ? h.has_channels : 1u + h.has_channels);
  r.children_active = (node0 >> h.children_active_bits_shift) & h.children_active_bits_mask;
  return r;
}

// unpack_data_edge (scom_shared.h:202-217): a unique brick has no rotation and no shift
LRT_HD SCom2DataEdge scom2_data_edge(const SCom2Header& h, float max_val, uint32_t edge0, uint32_t edge1) {
  SCom2DataEdge e;
  const bool is_unique = (edge0 & h.unique_brick_prefix) == h.unique_brick_prefix;
  const bool no_add = is_unique || h.child_add_mask == 0;
  const uint32_t code = (edge0 >> h.child_add_shift) & h.child_add_mask;
  const float frac = float(2u * code) / float(h.child_add_mask) - 1.0f;
  e.add = no_add ? 0.0f : (h.max_val > 0 ? h.max_val * frac : max_val * frac);
  e.rotation_id = is_unique ? 0u : (edge0 >> h.child_rot_shift) & h.child_rot_mask;
  e.data_offset = is_unique ? (edge1 & h.unique_brick_offset_mask) : (edge1 & h.child_offset_mask);
  return e;
}

// unpack_child_edge (scom_shared.h:219-225)
LRT_HD SCom2ChildEdge scom2_child_edge(const SCom2Header& h, uint32_t edge0, uint32_t edge1) {
  SCom2ChildEdge e;
  e.rotation_id = (edge0 >> h.child_rot_shift) & h.child_rot_mask;
  e.child_offset = edge1 & h.node_offset_mask;
  return e;
}

// get_links_end (scom_shared.h:237-258): end of child n's links after the node's own links
LRT_HD uint32_t scom2_links_end(const SCom2Header& h, uint32_t child_n, uint32_t node1, uint32_t node2) {
  if (h.has_multi_nodes != 0) {
    const uint32_t sh = child_n * h.reference_bits;
    return sh >= 32 ? ((node2 >> (sh % 32)) & h.reference_mask) : ((node1 >> sh) & h.reference_mask);
  }
  const uint32_t base = (node1 & 0xFFu) == SCOM2_CHILD_EMPTY ? 0u : 1u;
  const uint32_t active = (node1 >> h.children_active_bits_shift) & h.children_active_bits_mask;
  const uint32_t n = uint32_t(__builtin_popcount(active & ((1u << (1 + child_n)) - 1u)));
  return h.has_surfaces > 0 ? (1 + h.has_channels) * base + n : h.has_channels * (base + n);
}

// get_max_sdf_val (scom_shared.h:227-230)
LRT_HD float scom2_max_sdf_val(float level_sz) { return 1.7320508f * 2.0f / level_sz; }

}  // namespace lrt
