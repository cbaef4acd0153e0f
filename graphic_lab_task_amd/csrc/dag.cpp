// Octree DAGs (SdfDAG) on the host: the rotation composition table, the structural checks that
// keep every edge, node and distance a ray can reach inside the arrays, and the geometry record
// (BVHRT::AddGeom_SdfDAG, MultiRenderer/BVH2Common_host.cpp:1020-1131).
#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_math.h"
#include "scene.h"

namespace lrt {

void coctree_rot_modifiers(int v, int4 out[COCTREE_ROT_COUNT]);

// m_RotAddTable (BVH2Common_host.cpp:1057-1060): entry i * 48 + j is the rotation that applies
// rotation i and then rotation j, i.e. the Cayley table of the default transform subgroup
// (sdf/scom2/clustering/symmetric_groups.cpp:186-256, 294-310) read transposed. That subgroup's
// elements are the 48 legacy lattice rotations in their legacy order (octahedral_groups.h:27-35),
// so the table follows from composing their index maps on the 2^3 lattice, where they differ.
const std::vector<uint8_t>& rot_add_table() {
  static const std::vector<uint8_t> table = [] {
    int4 m[COCTREE_ROT_COUNT];
    coctree_rot_modifiers(2, m);
    auto map_of = [&](int r, int x, int y, int z) { return m[r].x * x + m[r].y * y + m[r].z * z + m[r].w; };
    std::vector<std::vector<int>> maps(COCTREE_ROT_COUNT, std::vector<int>(8));
    for (int r = 0; r < int(COCTREE_ROT_COUNT); ++r)
      for (int p = 0; p < 8; ++p) maps[r][p] = map_of(r, p >> 2, (p >> 1) & 1, p & 1);
    std::vector<uint8_t> t(COCTREE_ROT_COUNT * COCTREE_ROT_COUNT, 0);
    for (int i = 0; i < int(COCTREE_ROT_COUNT); ++i)
      for (int j = 0; j < int(COCTREE_ROT_COUNT); ++j) {
        std::vector<int> c(8);
        for (int p = 0; p < 8; ++p) c[p] = maps[j][maps[i][p]];   // i first, then j
        int k = 0;
        while (k < int(COCTREE_ROT_COUNT) && maps[k] != c) ++k;
        if (k == int(COCTREE_ROT_COUNT)) throw std::logic_error("rotations do not form a group");
        t[i * COCTREE_ROT_COUNT + j] = uint8_t(k);
      }
    return t;
  }();
  return table;
}

namespace {

struct DagChecker {
  const SdfDAGHeader& h;
  const SdfDAGNode* nodes;
  size_t n_nodes;
  const SdfDAGChildEdge* ce;
  size_t n_ce;
  const SdfDAGDataEdge* de;
  size_t n_de;
  size_t n_dist;
  uint32_t g3, v3, max_depth;
  std::unordered_map<uint32_t, uint32_t> height;

  [[noreturn]] void bad(const std::string& what) const { throw std::invalid_argument("SdfDAG: " + what); }

  // DAGIntersect (BVH2Common.cpp:2365-2557) reads a node's child edges when it has any, its data
  // edges when its voxel count is non-zero and the offset is set (leaves, and inner nodes taken as
  // LOD leaves), and its channel edge's rotation in the channel render modes.
  uint32_t check(uint32_t n, uint32_t depth) {
    if (depth > max_depth) bad("deeper than the walk's 16-bit positions allow (or a cycle)");
    auto it = height.find(n);
    if (it != height.end()) {
      if (depth + it->second > max_depth) bad("deeper than the walk's 16-bit positions allow");
      return it->second;
    }
    if (n >= n_nodes) bad("node index out of range");
    const SdfDAGNode& nd = nodes[n];
    if (nd.channels_edge.rotation_id >= COCTREE_ROT_COUNT) bad("channel edge rotation beyond the 48 rotations");
    const uint32_t count = nd.voxel_count_flags & 0xFFFFu;
    if (count != 0 && nd.data_edges_offset > 0) {
      if (uint64_t(nd.data_edges_offset) + count > n_de) bad("data edges out of range");
      for (uint32_t s = 0; s < count; ++s) {
        const SdfDAGDataEdge& e = de[nd.data_edges_offset + s];
        if (e.rotation_id >= COCTREE_ROT_COUNT) bad("data edge rotation beyond the 48 rotations");
        if (uint64_t(e.data_offset) + v3 > n_dist) bad("brick distances out of range");
      }
    }
    uint32_t hgt = 0;
    if (nd.children_edges_offset != 0) {
      if (uint64_t(nd.children_edges_offset) + g3 > n_ce) bad("child edges out of range");
      for (uint32_t c = 0; c < g3; ++c) {
        const SdfDAGChildEdge& e = ce[nd.children_edges_offset + c];
        if (e.child_offset == 0) continue;
        if (e.rotation_id >= COCTREE_ROT_COUNT) bad("child edge rotation beyond the 48 rotations");
        hgt = std::max(hgt, 1 + check(e.child_offset, depth + 1));
      }
    }
    height[n] = hgt;
    return hgt;
  }
};

}  // namespace

// BVHRT::AddGeom_SdfDAG (BVH2Common_host.cpp:1020-1131): arrays appended with their offsets
// moved by what the scene already holds (as the reference does, except that a brick starting at
// distance 0 is moved too; the reference leaves data offset 0 in place, which only matters from
// the second DAG of a scene on); the BLAS is the unit box plus a tiny dummy box.
uint32_t Scene::add_sdf_dag(const SdfDAGHeader& h, const SdfDAGNode* nodes, size_t n_nodes,
                            const SdfDAGChildEdge* child_edges, size_t n_child_edges,
                            const SdfDAGDataEdge* data_edges, size_t n_data_edges, const float* distances,
                            size_t n_distances) {
  if (!n_nodes || !n_child_edges || !n_data_edges || !n_distances)
    throw std::invalid_argument("SdfDAG: nodes, child edges, data edges and distances must be non-empty");
  if (n_nodes >= (1u << 28) || n_child_edges >= (1u << 28) || n_data_edges >= (1u << 28) || n_distances >= (1u << 30))
    throw std::invalid_argument("SdfDAG: arrays too large");
  if (h.dim != 3) throw std::invalid_argument("SdfDAG: only 3-dimensional DAGs are rendered");
  if (h.transform_subgroup != 0)
    throw std::invalid_argument("SdfDAG: only the default transform subgroup (48 rotations) is built");
  if (h.brick_size < 1 || h.brick_size > 16 || h.brick_pad > 2)
    throw std::invalid_argument("SdfDAG: brick_size must be 1..16 and brick_pad 0..2");
  if (h.node_grid_size < 2 || h.node_grid_size > 4) throw std::invalid_argument("SdfDAG: node_grid_size must be 2..4");
  const uint32_t v = h.brick_size + 2 * h.brick_pad + 1, g = h.node_grid_size;
  uint32_t max_depth = 0;
  for (uint64_t sz = g; sz <= 0xFFFFu; sz *= g) ++max_depth;   // level sizes and positions fit 16 bits
  DagChecker ck{h, nodes, n_nodes, child_edges, n_child_edges, data_edges, n_data_edges, n_distances,
                g * g * g, v * v * v, max_depth, {}};
  ck.check(0, 0);

  const uint32_t n_off = uint32_t(dagNodes.size()), ce_off = uint32_t(dagChildEdges.size());
  const uint32_t de_off = uint32_t(dagDataEdges.size()), d_off = uint32_t(dagDistances.size());
  if (uint64_t(n_off) + n_nodes >= (1u << 31) || uint64_t(d_off) + n_distances >= (1ull << 32))
    throw std::invalid_argument("SdfDAG: scene arrays too large");
  GeomData gd{};
  gd.boxMin = make4(-1, -1, -1, 1);
  gd.boxMax = make4(1, 1, 1, 1);
  gd.offset = uint2{n_off, uint32_t(dagGeoms.size())};
  gd.bvhOffset = uint32_t(nodePairs.size());
  gd.type = TYPE_SDF_DAG;
  geomData.push_back(gd);
  DagGeom dg{};
  dg.root = n_off;
  dg.brick_size = h.brick_size;
  dg.brick_pad = h.brick_pad;
  dg.node_grid_size = g;
  dg.mods = uint32_t(dagMods.size());
  dagGeoms.push_back(dg);
  int4 mods[2 * COCTREE_ROT_COUNT];
  coctree_rot_modifiers(int(v), mods);
  coctree_rot_modifiers(int(g), mods + COCTREE_ROT_COUNT);
  dagMods.insert(dagMods.end(), mods, mods + 2 * COCTREE_ROT_COUNT);
  for (size_t i = 0; i < n_nodes; ++i) {
    SdfDAGNode nd = nodes[i];
    if (nd.channels_edge.child_offset > 0) nd.channels_edge.child_offset += n_off;
    if (nd.children_edges_offset > 0) nd.children_edges_offset += ce_off;
    if (nd.data_edges_offset > 0) nd.data_edges_offset += de_off;
    dagNodes.push_back(nd);
  }
  for (size_t i = 0; i < n_child_edges; ++i) {
    SdfDAGChildEdge e = child_edges[i];
    if (e.child_offset > 0) e.child_offset += n_off;
    dagChildEdges.push_back(e);
  }
  for (size_t i = 0; i < n_data_edges; ++i) {
    SdfDAGDataEdge e = data_edges[i];
    e.data_offset += d_off;
    dagDataEdges.push_back(e);
  }
  dagDistances.insert(dagDistances.end(), distances, distances + n_distances);
  BVHNode boxes[2]{};
  boxes[0].boxMin = make3(-1, -1, -1);
  boxes[0].boxMax = make3(1, 1, 1);
  boxes[1].boxMin = make3(-1.0001f, -1.0001f, -1.0001f);
  boxes[1].boxMax = make3(-1, -1, -1);
  return add_aabb(TAG_SDF_DAG, boxes, 2);
}

}  // namespace lrt
