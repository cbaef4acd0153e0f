// Compact octrees (COctreeV3) on the host: the index modifiers of the 48 brick rotations, the
// structural checks that keep every word the walk reads inside the geometry, and the BLAS boxes
// (BVHRT::AddGeom_COctreeV3, MultiRenderer/BVH2Common_host.cpp:915-966).
#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "host_math.h"
#include "scene.h"

namespace lrt {

// scom::initialize_rot_modifiers (sdf/scom/similarity_compression.cpp:93-150) searches, for each
// rotation of a v^3 lattice about its centre (initialize_rot_transforms, :22-45), the linear map
// dot(m, (x, y, z, 1)) that gives the rotated point's row-major index (x v^2 + y v + z). Rotation
// i sends input axis a to output axis index_a with sign sign_a, so the map is known in closed
// form: output axis k weighs W[k] = (v^2, v, 1)[k]; an unmirrored input coordinate c adds
// W c, a mirrored one W (v - 1 - c). For v = 1 every map with m.w = 0 fits and the search keeps
// its last candidate, (-1, -1, -1, 0).
void coctree_rot_modifiers(int v, int4 out[COCTREE_ROT_COUNT]) {
  const int W[3] = {v * v, v, 1};
  for (int i = 0; i < int(COCTREE_ROT_COUNT); ++i) {
    if (v == 1) {
      out[i] = int4{-1, -1, -1, 0};
      continue;
    }
    int code = i;
    const int index_1 = code % 3; code /= 3;
    const int sign_1 = code % 2; code /= 2;
    const int index_2 = (index_1 + code % 2 + 1) % 3;
    const int index_3 = (index_1 + (code + 1) % 2 + 1) % 3; code /= 2;
    const int sign_2 = code % 2; code /= 2;
    const int sign_3 = code % 2;
    const int idx[3] = {index_1, index_2, index_3};
    const int sgn[3] = {sign_1, sign_2, sign_3};
    int coef[3], w = 0;
    for (int a = 0; a < 3; ++a) {
      const int wk = W[idx[a]];
      coef[a] = sgn[a] ? -wk : wk;
      if (sgn[a]) w += wk * (v - 1);
    }
    out[i] = int4{coef[0], coef[1], coef[2], w};
  }
}

namespace {

inline uint32_t popc(uint32_t x) { return uint32_t(__builtin_popcount(x)); }

// Rank of flag bit `local` among the two flag words, as the brick decoders form it
// (BVH2Common.cpp:1762-1771, 1837-1845); defined for local <= 64.
inline uint32_t flag_rank(uint32_t w0, uint32_t w1, uint32_t local) {
  const uint32_t b0 = local > 31 ? w0 : (w0 & ((1u << local) - 1u));
  const uint32_t b1 = local > 32 ? (w1 & ((1u << (local - 32)) - 1u)) : 0u;
  return popc(b0) + popc(b1);
}

struct CoctChecker {
  const COctreeV3Header& h;
  const uint32_t* D;
  uint64_t size;
  uint32_t v, p, vpi;
  std::unordered_map<uint32_t, uint32_t> node_height;
  std::unordered_set<uint64_t> leaves_ok;
  uint32_t max_depth = 15;   // 16-bit positions and level sizes of the walk

  [[noreturn]] void bad(const std::string& what) const {
    throw std::invalid_argument("COctreeV3: " + what);
  }
  uint32_t word(uint64_t i) const {
    if (i >= size) bad("word index out of range");
    return D[i];
  }

  // Every word the three brick decoders, the smoothed normal and the texture coordinates read
  // (BVH2Common.cpp:1653-1851, 1940-1999) must lie inside the geometry.
  void check_leaf(uint32_t offset, uint32_t type, uint32_t code) {
    const uint64_t key = (uint64_t(offset) << 2) | type;
    if ((code & h.rot_mask) >= COCTREE_ROT_COUNT) bad("leaf rotation index beyond the 48 rotations");
    if (leaves_ok.count(key)) return;
    const uint64_t o = offset;
    const uint64_t v3 = uint64_t(v) * v * v, v2 = uint64_t(v) * v, p3 = uint64_t(p) * p * p;
    const uint64_t pfs = (p3 + 31) / 32;
    uint64_t off5 = 0, last = 0;
    if (type == COCTREE_LEAF_TYPE_GRID) {
      off5 = 2 + 8ull * h.uv_size;
      last = o + off5 + (v3 - 1) / vpi;
    } else if (type == COCTREE_LEAF_TYPE_BIT_PACK) {
      if (v3 > 64) bad("bit-pack leaves need at most 64 values per brick");
      const uint64_t off1 = pfs, off3 = off1 + (v3 + 31) / 32;
      off5 = off3 + 2 + 8ull * h.uv_size;
      if (o + off5 > size) bad("leaf header out of range");
      const uint32_t w0 = word(o + off1), w1 = v3 > 32 ? word(o + off1 + 1) : 0u;
      last = o + off5 + flag_rank(w0, w1, uint32_t(v3 - 1)) / vpi;
    } else if (type == COCTREE_LEAF_TYPE_SLICES) {
      if (v2 > 64) bad("slice leaves need at most 64 values per slice");
      const uint64_t sfu = (v2 + 31) / 32, off1 = pfs, off2 = off1 + uint64_t(v) * sfu, off3 = off2 + (v + 1) / 2;
      off5 = off3 + 2 + 8ull * h.uv_size;
      if (o + off5 > size) bad("leaf header out of range");
      uint64_t mx = 0;
      for (uint32_t s = 0; s < v; ++s) {
        const uint32_t w0 = word(o + off1 + sfu * s), w1 = v2 > 32 ? word(o + off1 + sfu * s + 1) : 0u;
        const uint32_t so = (word(o + off2 + s / 2) >> (16 * (s % 2))) & 0xFFFFu;
        mx = std::max<uint64_t>(mx, uint64_t(so) + flag_rank(w0, w1, uint32_t(v2 - 1)));
      }
      last = o + off5 + mx / vpi;
    } else {
      bad("unknown leaf type");
    }
    if (last >= size) bad("leaf values out of range");
    if (h.uv_size > 0) {
      // texture coordinates sit where a slice leaf would keep them, whatever the leaf type
      const uint64_t sfu = (v2 + 31) / 32;
      const uint64_t off4 = pfs + uint64_t(v) * sfu + (v + 1) / 2 + 2;
      if (o + off4 + 8 > size) bad("leaf texture coordinates out of range");
    }
    leaves_ok.insert(key);
  }

  // Returns the levels below node n; the node sits `depth` levels under the root.
  uint32_t check_node(uint32_t n, uint32_t depth) {
    if (depth > max_depth) bad("deeper than 15 levels (or a cycle)");
    auto it = node_height.find(n);
    if (it != node_height.end()) {
      if (depth + it->second > max_depth) bad("deeper than 15 levels");
      return it->second;
    }
    const uint32_t info = word(n);
    if (h.lods) {
      const uint32_t type = (info >> COCTREE_LOD_LEAF_TYPE_SHIFT) & COCTREE_LEAF_TYPE_MASK;
      const uint32_t id = word(uint64_t(n) + 1);
      const uint32_t code = word(uint64_t(n) + 1 + h.trans_off) & 0x3FFFFFFFu;
      if (type != COCTREE_LEAF_TYPE_NOT_A_LEAF) check_leaf((id & h.idx_mask) >> h.idx_sh, type, code);
    }
    uint32_t height = 0;
    for (uint32_t c = 0; c < 8; ++c) {
      if (!(info & (1u << c))) continue;
      const uint32_t local = h.lods + popc(info & ((1u << c) - 1u));
      const uint64_t pos = uint64_t(n) + 1 + uint64_t(h.uints_per_child_info) * local;
      const uint32_t child = word(pos);
      const uint32_t code = word(pos + h.trans_off) & 0x3FFFFFFFu;
      const uint32_t type = (info >> (8 + COCTREE_LEAF_TYPE_BITS * c)) & COCTREE_LEAF_TYPE_MASK;
      if (type == COCTREE_LEAF_TYPE_NOT_A_LEAF) {
        height = std::max(height, 1 + check_node(child, depth + 1));
      } else {
        if (depth + 1 > max_depth) bad("deeper than 15 levels");
        check_leaf((child & h.idx_mask) >> h.idx_sh, type, code);
        height = std::max(height, 1u);
      }
    }
    node_height[n] = height;
    return height;
  }
};

// add_border_nodes_rec (BVH2Common_host.cpp:76-117): boxes of the nodes that have a leaf child
// or sit at the requested BLAS level, each pointing at its node.
void coct_border_boxes(const uint32_t* D, uint64_t size, uint32_t upc, uint32_t max_level, std::vector<BVHNode>& out,
                       uint32_t nodeId, float3 p, float d, uint32_t level) {
  if (nodeId >= size) return;   // the reference skips such a node with a message
  const uint32_t info = D[nodeId];
  const uint32_t children_leaves = info & 0xFFFF00u;
  if (children_leaves || max_level == level) {
    BVHNode b{};
    b.leftOffset = nodeId;
    b.boxMin = make3(-1, -1, -1) + (2.0f * p) * d;
    b.boxMax = make3(-1, -1, -1) + (2.0f * (p + make3(1, 1, 1))) * d;
    out.push_back(b);
    return;
  }
  uint32_t child_pos = 0;
  for (uint32_t i = 0; i < 8; i++) {
    if (!(info & (1u << i))) continue;
    const float ch_d = d / 2;
    const float3 ch_p = 2.0f * p + make3(float((i & 4) >> 2), float((i & 2) >> 1), float(i & 1));
    const uint64_t at = uint64_t(nodeId) + 1 + uint64_t(upc) * child_pos;
    if (at >= size) throw std::invalid_argument("COctreeV3: child entry out of range");
    coct_border_boxes(D, size, upc, max_level, out, D[at], ch_p, ch_d, level + 1);
    child_pos++;
  }
}

// fill_coctree_v3_header (sdf/scom/scom.h:27-61): the words node_pack_mode determines
bool header_matches_pack_mode(const COctreeV3Header& h) {
  struct P { uint32_t upc, idx_sh, trans_off, idx_mask, rot_mask, add_mask; };
  static const P modes[3] = {{1, 0, 0, 0xFFFFFFFFu, 0x00000000u, 0x00000000u},
                             {2, 0, 1, 0xFFFFFFFFu, 0x0000003Fu, 0x3FFFFF00u},
                             {1, 14, 0, 0xFFFFC000u, 0x0000003Fu, 0x00003FC0u}};
  if (h.node_pack_mode > 2) return false;
  const P& m = modes[h.node_pack_mode];
  return h.uints_per_child_info == m.upc && h.idx_sh == m.idx_sh && h.trans_off == m.trans_off &&
         h.idx_mask == m.idx_mask && h.rot_mask == m.rot_mask && h.add_mask == m.add_mask;
}

}  // namespace

// BVHRT::AddGeom_COctreeV3 (BVH2Common_host.cpp:915-966). The reference keeps one octree per
// scene (the words, header and leaf boxes of the last one added) and the rotation modifiers of
// the first; here every octree keeps its own (CoctGeom), identical for single-octree scenes.
uint32_t Scene::add_coctree_v3(const COctreeV3Header& h, const uint32_t* data, size_t size, uint32_t bvh_level) {
  if (size == 0 || size >= (1u << 28)) throw std::invalid_argument("COctreeV3: size out of range");
  if (h.brick_size < 1 || h.brick_size > 16) throw std::invalid_argument("COctreeV3: brick_size must be 1..16");
  if (h.brick_pad > 2) throw std::invalid_argument("COctreeV3: brick_pad must be 0..2");
  if (h.bits_per_value < 1 || h.bits_per_value > 32) throw std::invalid_argument("COctreeV3: bits_per_value must be 1..32");
  if (h.uv_size > 1) throw std::invalid_argument("COctreeV3: uv_size must be 0 or 1");
  if (h.sim_compression > 1 || h.lods > 1) throw std::invalid_argument("COctreeV3: sim_compression and lods are 0 or 1");
  if (!header_matches_pack_mode(h))
    throw std::invalid_argument("COctreeV3: header words do not match node_pack_mode");
  if (h.sim_compression && h.add_mask == 0)
    throw std::invalid_argument("COctreeV3: sim_compression needs a pack mode with an additive shift field");
  if (h.lods && bvh_level > 0)
    throw std::invalid_argument("COctreeV3: octrees with LOD leaves take BLAS level 0 (the box builder skips no LOD word)");
  CoctChecker ck{h, data, size, h.brick_size + 2 * h.brick_pad + 1, h.brick_size + 2 * h.brick_pad,
                 32u / h.bits_per_value, {}, {}};
  ck.check_node(0, 0);

  std::vector<BVHNode> boxes;
  coct_border_boxes(data, size, h.uints_per_child_info, bvh_level, boxes, 0, make3(0, 0, 0), 1.0f, 0);
  if (boxes.size() < 2) {
    boxes.resize(2);
    boxes[0].boxMin = make3(-1, -1, -1);
    boxes[0].boxMax = make3(1, 1, 1);
    boxes[0].leftOffset = 0;
    boxes[1].boxMin = make3(-1.0001f, -1.0001f, -1.0001f);
    boxes[1].boxMax = make3(-1, -1, -1);
    boxes[1].leftOffset = 0;
  }
  for (BVHNode& b : boxes) b.escapeIndex = 0;

  CoctGeom cg{};
  cg.base = uint32_t(coctData.size());
  cg.mods = uint32_t(coctMods.size());
  cg.size = uint32_t(size);
  cg.brick_size = h.brick_size;
  cg.brick_pad = h.brick_pad;
  cg.bits_per_value = h.bits_per_value;
  cg.uv_size = h.uv_size;
  cg.sim_compression = h.sim_compression;
  cg.lods = h.lods;
  cg.uints_per_child_info = h.uints_per_child_info;
  cg.idx_mask = h.idx_mask;
  cg.idx_sh = h.idx_sh;
  cg.trans_off = h.trans_off;
  cg.rot_mask = h.rot_mask;
  cg.add_mask = h.add_mask;
  if (uint64_t(cg.base) + size > 0xFFFFFFFFull) throw std::invalid_argument("COctreeV3: scene words exceed 2^32");

  int4 mods[3 * COCTREE_ROT_COUNT];
  coctree_rot_modifiers(int(h.brick_size + 2 * h.brick_pad + 1), mods);
  coctree_rot_modifiers(int(h.brick_size + 2 * h.brick_pad), mods + COCTREE_ROT_COUNT);
  coctree_rot_modifiers(2, mods + 2 * COCTREE_ROT_COUNT);

  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(coctGeoms.size()), uint32_t(origNodes.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_COCTREE_V3;
  geomData.push_back(g);
  coctGeoms.push_back(cg);
  coctData.insert(coctData.end(), data, data + size);
  coctMods.insert(coctMods.end(), mods, mods + 3 * COCTREE_ROT_COUNT);
  origNodes.insert(origNodes.end(), boxes.begin(), boxes.end());
  return add_aabb(TAG_COCTREE_BRICKED, boxes.data(), boxes.size());
}

}  // namespace lrt
