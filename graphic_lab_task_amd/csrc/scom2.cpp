// SCom2 trees on the host: the structural checks that keep every word the walk can read inside
// the tree and the geometry record (BVHRT::AddGeom_SCom2, MultiRenderer/BVH2Common_host.cpp:1133-1191).
#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "host_math.h"
#include "scene.h"
#include "scom2_unpack.h"

namespace lrt {

void coctree_rot_modifiers(int v, int4 out[COCTREE_ROT_COUNT]);

namespace {

// Follows SCom2Intersect (BVH2Common.cpp:2560-2818) node by node: for every node entry a ray can
// reach (through child edges, or through a leaf link read as a node when it has no surface, as
// the walk does) the head and reference words, each child's links, every surface link and its
// brick, the channel link, and the node's own links taken as an LOD leaf.
struct SCom2Checker {
  const SCom2Header& h;
  const uint32_t* D;      // nodes then bricks, as the scene holds them (local offsets)
  uint64_t total;
  uint64_t bricks_at;     // local first brick word
  uint32_t v3;
  uint32_t max_depth = 15;
  std::unordered_map<uint32_t, uint32_t> height;
  std::unordered_set<uint64_t> leaves_ok;

  [[noreturn]] void bad(const std::string& what) const { throw std::invalid_argument("SCom2: " + what); }
  uint32_t word(uint64_t i) const {
    if (i >= total) bad("word index out of range");
    return D[i];
  }

  void check_leaf(uint64_t link_off, uint32_t surf_count) {
    const uint64_t key = (link_off << 16) | surf_count;   // surf_count < 2^16
    if (leaves_ok.count(key)) return;
    for (uint32_t s = 0; s < surf_count; ++s) {
      const uint64_t l0 = link_off + uint64_t(s) * h.uints_per_link;
      const SCom2DataEdge e = scom2_data_edge(h, 1.0f, word(l0), word(l0 + h.child_offset_off));
      if (e.rotation_id >= COCTREE_ROT_COUNT) bad("brick rotation beyond the 48 rotations");
      const uint64_t rel = uint64_t(h.bricks_step) * e.data_offset;   // the walk adds it in 32 bits
      if (rel >= total || bricks_at + rel + (v3 - 1) / h.values_per_uint >= total) bad("brick values out of range");
    }
    if (h.has_channels) {
      const uint64_t ch = link_off + uint64_t(surf_count) * h.uints_per_link;
      const SCom2ChildEdge ce = scom2_child_edge(h, word(ch), word(ch + h.child_offset_off));
      if (ce.rotation_id >= COCTREE_ROT_COUNT) bad("channel rotation beyond the 48 rotations");
    }
    leaves_ok.insert(key);
  }

  uint32_t check_node(uint64_t n, uint32_t depth) {
    if (depth > max_depth) bad("deeper than 15 levels (or a cycle)");
    if (n >= total) bad("node index out of range");
    auto it = height.find(uint32_t(n));
    if (it != height.end()) {
      if (depth + it->second > max_depth) bad("deeper than 15 levels");
      return it->second;
    }
    const uint32_t node0 = word(n), nodeL = word(n + h.links_offset - 1), ref = word(n + h.references_offset);
    const SCom2NodeHead head = scom2_node_head(h, node0, nodeL);
    if (head.base_type != SCOM2_CHILD_EMPTY) {   // the node's own links, read when it is taken as an LOD leaf
      const uint32_t cnt = head.base_links_end - h.has_channels;
      if (cnt >= (1u << 16)) bad("node link count out of range");
      check_leaf(n + h.links_offset, cnt);
    }
    uint32_t hgt = 0;
    for (uint32_t c = 0; c < 8; ++c) {
      const uint32_t type = (head.children_types >> (SCOM2_CHILD_TYPE_BITS * c)) & SCOM2_CHILD_TYPE_MASK;
      if (type == SCOM2_CHILD_EMPTY) continue;
      const uint32_t start = c == 0 ? head.base_links_end : scom2_links_end(h, c - 1, ref, nodeL);
      const uint32_t end = scom2_links_end(h, c, ref, nodeL);
      const uint64_t link_off = n + h.links_offset + uint64_t(start) * h.uints_per_link;
      if (type == SCOM2_CHILD_NODE) {
        const SCom2ChildEdge ce = scom2_child_edge(h, word(link_off), word(link_off + h.child_offset_off));
        if (ce.rotation_id >= COCTREE_ROT_COUNT) bad("child rotation beyond the 48 rotations");
        hgt = std::max(hgt, 1 + check_node(uint64_t(ce.child_offset), depth + 1));
      } else {
        const uint32_t surf = end - start - h.has_channels;
        const bool surface = type == SCOM2_CHILD_LEAF_SURFACE;
        if (surf == 0 && !surface) {
          // the walk tells leaves from nodes by a non-zero count or surface flag: this link is
          // read as a node head
          hgt = std::max(hgt, 1 + check_node(link_off, depth + 1));
        } else {
          if (surf >= (1u << 16)) bad("leaf surface count out of range");
          if (depth + 1 > max_depth) bad("deeper than 15 levels");
          check_leaf(link_off, surf);
          hgt = std::max(hgt, 1u);
        }
      }
    }
    height[uint32_t(n)] = hgt;
    return hgt;
  }
};

// Follows SCom2IntersectFast (BVH2Common.cpp:2988-3231) for a tree it may be used on (no
// multi-nodes, no channels, 3^3 values, one-word links): a node's head word, one link per child
// that has data (its rank among them), the child node's head or the leaf's brick.
struct SCom2FastChecker {
  const SCom2Header& h;
  const uint32_t* D;
  uint64_t total, bricks_at;
  uint32_t max_depth = 15;
  std::unordered_map<uint32_t, uint32_t> height;

  [[noreturn]] void bad(const std::string& what) const { throw std::invalid_argument("SCom2: " + what); }
  uint32_t word(uint64_t i) const {
    if (i >= total) bad("word index out of range");
    return D[i];
  }
  uint32_t check_node(uint64_t n, uint32_t depth) {
    if (depth > max_depth) bad("deeper than 15 levels (or a cycle)");
    auto it = height.find(uint32_t(n));
    if (it != height.end()) {
      if (depth + it->second > max_depth) bad("deeper than 15 levels");
      return it->second;
    }
    const uint32_t node0 = word(n);
    const SCom2NodeHead head = scom2_node_head(h, node0, node0);
    const uint64_t links = n + h.links_offset + (head.base_type == SCOM2_CHILD_EMPTY ? 0 : 1);
    uint32_t rank = 0, hgt = 0;
    for (uint32_t c = 0; c < 8; ++c) {
      const uint32_t type = (head.children_types >> (2 * c)) & 3u;
      if (type == SCOM2_CHILD_EMPTY) continue;
      const uint32_t link = word(links + rank++);
      if (type & 1u) {
        const SCom2DataEdge e = scom2_data_edge(h, 1.0f, link, link);
        if (e.rotation_id >= COCTREE_ROT_COUNT) bad("brick rotation beyond the 48 rotations");
        const uint64_t rel = uint64_t(h.bricks_step) * e.data_offset;
        if (rel >= total || bricks_at + rel + (27 - 1) / h.values_per_uint >= total) bad("brick values out of range");
        hgt = std::max(hgt, 1u);
      } else {
        const SCom2ChildEdge ce = scom2_child_edge(h, link, link);
        if (ce.rotation_id >= COCTREE_ROT_COUNT) bad("child rotation beyond the 48 rotations");
        hgt = std::max(hgt, 1 + check_node(ce.child_offset, depth + 1));
      }
    }
    height[uint32_t(n)] = hgt;
    return hgt;
  }
};

}  // namespace

// The tree shape for which the reference may take SCom2IntersectFast (its preset-dependent switch,
// BVH2Common_host.cpp:1436-1451, also needs the tree to be the scene's only one and
// LAMBERT_NO_TEX rendering).
bool scom2_plain_tree(const SCom2Header& h) {
  return h.has_multi_nodes == 0 && h.has_channels == 0 && h.v_size == 3 && h.links_offset == 1 &&
         h.child_offset_off == 0;
}

// BVHRT::AddGeom_SCom2 (BVH2Common_host.cpp:1133-1191): nodes and bricks appended to the scene's
// word array, the header's nodes_arr_offset / bricks_arr_offset set to where they landed (the
// offsets inside the tree stay relative to them); the BLAS is the unit box plus a dummy box. The
// reference keeps the rotation modifiers of the first geometry added; here each SCom2 keeps its own.
uint32_t Scene::add_scom2(const SCom2Header& header, const uint32_t* nodes, size_t n_nodes, const uint32_t* bricks,
                          size_t n_bricks) {
  if (!n_nodes || !n_bricks) throw std::invalid_argument("SCom2: nodes and bricks must be non-empty");
  if (uint64_t(n_nodes) + n_bricks >= (1u << 30)) throw std::invalid_argument("SCom2: tree too large");
  const SCom2Header& h = header;
  if (h.dimension != 3) throw std::invalid_argument("SCom2: only 3-dimensional trees are rendered");
  if (h.brick_size < 1 || h.brick_size > 16) throw std::invalid_argument("SCom2: brick_size must be 1..16");
  if (h.bits_per_value < 1 || h.values_per_uint < 1 || uint64_t(h.bits_per_value) * h.values_per_uint > 32)
    throw std::invalid_argument("SCom2: values_per_uint x bits_per_value must fit one word");
  if (h.uints_per_link < 1 || h.links_offset < 1) throw std::invalid_argument("SCom2: links need at least one word");
  // the unpacking shifts words by these fields: 32 or more is no shift the walk could mean
  if (h.child_rot_shift >= 32 || h.child_add_shift >= 32 || h.children_types_shift >= 32 ||
      h.base_reference_shift >= 32 || h.children_active_bits_shift >= 32 || h.reference_bits >= 32)
    throw std::invalid_argument("SCom2: header shift fields must be below 32");
  std::vector<uint32_t> local(nodes, nodes + n_nodes);
  local.insert(local.end(), bricks, bricks + n_bricks);
  SCom2Header lh = h;
  lh.nodes_arr_offset = 0;
  lh.bricks_arr_offset = uint32_t(n_nodes);
  const uint32_t v = h.brick_size + 1;
  SCom2Checker ck{lh, local.data(), local.size(), n_nodes, v * v * v};
  ck.check_node(uint64_t(h.root_node_off), 0);
  if (scom2_plain_tree(h)) {
    if (h.brick_size != 2) throw std::invalid_argument("SCom2: a plain tree with 3^3 values needs brick_size 2");
    SCom2FastChecker fk{lh, local.data(), local.size(), n_nodes};
    fk.check_node(uint64_t(h.root_node_off), 0);
  }

  const uint32_t base = uint32_t(scom2Data.size());
  if (uint64_t(base) + local.size() >= (1ull << 32)) throw std::invalid_argument("SCom2: scene words exceed 2^32");
  SCom2Header gh = h;
  gh.nodes_arr_offset = base;
  gh.bricks_arr_offset = base + uint32_t(n_nodes);
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{base, uint32_t(scom2Headers.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SCOM2;
  geomData.push_back(g);
  scom2Headers.push_back(gh);
  scom2Data.insert(scom2Data.end(), local.begin(), local.end());
  int4 mods[3 * COCTREE_ROT_COUNT];
  coctree_rot_modifiers(int(h.brick_size + 1), mods);
  coctree_rot_modifiers(int(h.brick_size), mods + COCTREE_ROT_COUNT);
  coctree_rot_modifiers(2, mods + 2 * COCTREE_ROT_COUNT);
  scom2Mods.insert(scom2Mods.end(), mods, mods + 3 * COCTREE_ROT_COUNT);
  BVHNode boxes[2]{};
  boxes[0].boxMin = make3(-1, -1, -1);
  boxes[0].boxMax = make3(1, 1, 1);
  boxes[1].boxMin = make3(-1.0001f, -1.0001f, -1.0001f);
  boxes[1].boxMax = make3(-1, -1, -1);
  return add_aabb(TAG_SCOM2, boxes, 2);
}

}  // namespace lrt
