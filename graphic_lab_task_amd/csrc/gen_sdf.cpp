// SDF structure builders: the reference's sparse-octree builder and converters restated
// over a sphere-union distance function (the functor of the benchmark scenes).
#include <omp.h>

#include <cmath>
#include <limits>
#include <stdexcept>
#include <unordered_map>

#include "gen.h"
#include "lrt_math.h"

namespace lrt {

float SphereUnion::operator()(float3 p) const {
  float d = std::numeric_limits<float>::max();
  for (size_t i = 0; i < n; ++i) {
    const float* s = spheres + 4 * i;
    const float dx = p.x - s[0], dy = p.y - s[1], dz = p.z - s[2];
    const float di = std::sqrt(dx * dx + dy * dy + dz * dz) - s[3];
    d = fmin_ref(d, di);
  }
  return d;
}

namespace {

// is_border (sdf/build/sparse_octree_common.h:84-87), evaluated in double as there.
bool is_border(float distance, int level) {
  return level < 2 ? true : double(std::fabs(distance)) < std::sqrt(3.0) * std::pow(2.0, -level);
}

// add_global_node_rec (sdf/build/sparse_octree_builder.cpp:37-104), fill_internal_volume = false.
void add_global_node_rec(std::vector<GONode>& nodes, std::vector<float>& values, const SphereUnion& sdf,
                         uint32_t node_idx, uint32_t depth, uint32_t max_depth, float3 p, float d, int bs, int pad) {
  const float value_center = sdf(2.0f * ((p + make3(0.5f, 0.5f, 0.5f)) * d) - make3(1, 1, 1));
  float min_val = 1000, max_val = -1000;
  nodes[node_idx].val_off = uint32_t(values.size());
  nodes[node_idx].offset = 0;
  for (int i = -pad; i <= bs + pad; ++i)
    for (int j = -pad; j <= bs + pad; ++j)
      for (int k = -pad; k <= bs + pad; ++k) {
        const float3 q = p + make3(float(i) / float(bs), float(j) / float(bs), float(k) / float(bs));
        const float val = sdf(2.0f * (q * d) - make3(1, 1, 1));
        values.push_back(val);
        min_val = fmin_ref(min_val, val);
        max_val = fmax_ref(max_val, val);
      }
  bool is_parent = false;
  if (depth < max_depth && (std::fabs(value_center) < std::sqrt(3.0f) * d || min_val * max_val <= 0)) {
    is_parent = true;
    const uint32_t off = uint32_t(nodes.size());
    nodes[node_idx].offset = off;
    nodes.resize(nodes.size() + 8);
    for (uint32_t cid = 0; cid < 8; cid++)
      add_global_node_rec(nodes, values, sdf, off + cid, depth + 1, max_depth,
                          2.0f * p + make3(float((cid & 4) >> 2), float((cid & 2) >> 1), float(cid & 1)), d / 2, bs,
                          pad);
  }
  const bool active_as_leaf = min_val <= 0 && max_val >= 0;
  if (active_as_leaf) nodes[node_idx].type = is_parent ? 3 : 1;
  else nodes[node_idx].type = is_parent ? 2 : 0;
}

struct LargeNode {
  float3 p;
  float d;
  uint32_t level;
  uint32_t children_idx;
  float value;
};

}  // namespace

void sdf_to_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                          int threads, GlobalOctree& out) {
  if (depth < 1 || depth > 16) throw std::invalid_argument("octree depth must be 1..16");
  if (brick_size < 1 || brick_size > 16) throw std::invalid_argument("brick size must be 1..16");
  out.brick_size = brick_size;
  out.brick_pad = brick_pad;
  const uint32_t min_remove_level = std::min(depth, 4u);
  std::vector<LargeNode> large;
  large.push_back({make3(0, 0, 0), 1.0f, 0u, 0u, 1000.0f});
  for (size_t i = 0; i < large.size(); ++i) {
    if (large[i].level < min_remove_level) {
      large[i].children_idx = uint32_t(large.size());
      for (int j = 0; j < 8; j++) {
        const float ch_d = large[i].d / 2;
        const float3 ch_p = 2.0f * large[i].p + make3(float((j & 4) >> 2), float((j & 2) >> 1), float(j & 1));
        large.push_back({ch_p, ch_d, large[i].level + 1, 0u, 1000.0f});
      }
    }
  }
  std::vector<uint32_t> border;
  for (size_t i = 0; i < large.size(); i++) {
    const float3 pos = 2.0f * ((large[i].p + make3(0.5f, 0.5f, 0.5f)) * large[i].d) - make3(1, 1, 1);
    large[i].value = sdf(pos);
    if (large[i].children_idx == 0 && is_border(large[i].value, int(large[i].level))) border.push_back(uint32_t(i));
  }
  // Each border subtree is built independently (as the reference's per-thread groups are),
  // then merged in large-node order, so the numbering does not depend on the thread count.
  std::vector<std::vector<GONode>> sub_nodes(border.size());
  std::vector<std::vector<float>> sub_vals(border.size());
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (long b = 0; b < long(border.size()); ++b) {
    const LargeNode& ln = large[border[b]];
    sub_nodes[b].emplace_back();
    add_global_node_rec(sub_nodes[b], sub_vals[b], sdf, 0, ln.level, depth, ln.p, ln.d, int(brick_size),
                        int(brick_pad));
  }
  std::vector<int64_t> group_of(large.size(), -1);
  for (size_t b = 0; b < border.size(); ++b) group_of[border[b]] = int64_t(b);

  const int v_size = int(brick_size + 2 * brick_pad + 1);
  const size_t v3 = size_t(v_size) * v_size * v_size;
  out.nodes.assign(large.size(), GONode{3, 0, 0});
  out.values.clear();
  for (size_t i = 0; i < large.size(); i++) {
    out.nodes[i].val_off = uint32_t(out.values.size());
    out.nodes[i].type = 3;  // the reference marks every top node NODE (sparse_octree_builder.cpp:202)
    for (int i0 = -int(brick_pad); i0 <= int(brick_size + brick_pad); ++i0)
      for (int j0 = -int(brick_pad); j0 <= int(brick_size + brick_pad); ++j0)
        for (int k0 = -int(brick_pad); k0 <= int(brick_size + brick_pad); ++k0) {
          const float3 q = large[i].p + make3(float(i0) / float(brick_size), float(j0) / float(brick_size),
                                              float(k0) / float(brick_size));
          out.values.push_back(sdf(2.0f * (q * large[i].d) - make3(1, 1, 1)));
        }
    out.nodes[i].offset = large[i].children_idx;
    if (group_of[i] >= 0) {
      const std::vector<GONode>& sn = sub_nodes[size_t(group_of[i])];
      const std::vector<float>& sv = sub_vals[size_t(group_of[i])];
      if (sn.size() <= 1) continue;  // local root not split: stays a leaf
      const uint32_t prev = uint32_t(out.nodes.size());
      const int64_t shift = int64_t(prev) - 1;  // local index 1 -> global prev
      out.nodes[i].offset = uint32_t(int64_t(sn[0].offset) + shift);
      const uint32_t val_start = sn[1].val_off;
      const int64_t val_shift = int64_t(out.values.size()) - int64_t(val_start);
      for (size_t j = 1; j < sn.size(); ++j) {
        GONode n = sn[j];
        if (n.offset != 0) n.offset = uint32_t(int64_t(n.offset) + shift);
        n.val_off = uint32_t(int64_t(n.val_off) + val_shift);
        out.nodes.push_back(n);
      }
      out.values.insert(out.values.end(), sv.begin() + val_start, sv.begin() + val_start + (sn.size() - 1) * v3);
    }
  }
}

void build_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                         int threads, GlobalOctree& out) {
  if (threads < 0) gpu_sdf_to_global_octree(sdf, depth, brick_size, brick_pad, out);
  else sdf_to_global_octree(sdf, depth, brick_size, brick_pad, threads, out);
}

// global_octree_to_multi_octree (sdf/multi/sdf_multi.cpp:124-180)
void global_octree_to_multi_octree(const GlobalOctree& g, std::vector<SdfMultiOctreeNode>& nodes,
                                   std::vector<float>& values) {
  if (g.brick_size != 1 || g.brick_pad != 0) throw std::invalid_argument("multi octree needs brick size 1, pad 0");
  nodes.resize(g.nodes.size());
  for (size_t i = 0; i < g.nodes.size(); ++i) {
    nodes[i].children_offset = g.nodes[i].offset;
    nodes[i].data_offset = g.nodes[i].val_off;
    // add_global_node_rec: bricks_count = 1, reset to 0 when the node is not active
    nodes[i].voxel_count = (g.nodes[i].type == 1 || g.nodes[i].type == 3) ? 1u : 0u;
    nodes[i].flags = 0;   // is_surfaced is never set by the builder
  }
  values = g.values;
}

void global_octree_to_frame_octree(const GlobalOctree& g, std::vector<SdfFrameOctreeNode>& out) {
  if (g.brick_size != 1 || g.brick_pad != 0) throw std::invalid_argument("frame octree needs brick size 1, pad 0");
  out.resize(g.nodes.size());
  for (size_t i = 0; i < g.nodes.size(); ++i) {
    out[i].offset = g.nodes[i].offset;
    for (int j = 0; j < 8; j++) out[i].values[j] = g.values[g.nodes[i].val_off + j];
  }
}

namespace {

// Seeded smooth field in [0,1] at a brick corner (value noise), channel c.
double corner_field(uint32_t lod, uint32_t px, uint32_t py, uint32_t pz, uint32_t corner, uint64_t seed, uint64_t c) {
  const double s = 2.0 / lod;
  const double x = -1.0 + s * (px + ((corner & 4) ? 1 : 0));
  const double y = -1.0 + s * (py + ((corner & 2) ? 1 : 0));
  const double z = -1.0 + s * (pz + ((corner & 1) ? 1 : 0));
  return std::min(1.0, std::max(0.0, 0.5 + 0.45 * value_noise3(2.0 * x, 2.0 * y, 2.0 * z, seed * 3 + c)));
}

void sbs_dx_rec(const GlobalOctree& g, const SdfSBSHeader& h, uint64_t colour_seed, std::vector<SdfSBSNode>& nodes,
                std::vector<uint32_t>& values, uint32_t idx, uint32_t lod, uint32_t px, uint32_t py, uint32_t pz) {
  const GONode& n = g.nodes[idx];
  if (n.offset == 0) {
    if (n.type == 0) return;
    const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
    const uint32_t v3 = v_size * v_size * v_size;
    const uint32_t off = uint32_t(values.size());
    const float d_max = float(2 * std::sqrt(3.0) / lod);
    const uint32_t bits = 8 * h.bytes_per_value;
    const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
    const uint32_t vpi = 4 / h.bytes_per_value;
    const uint32_t layout = h.aux_data & SDF_SBS_NODE_LAYOUT_MASK;
    // DX_UV16 reserves 32 words (8 used), as the reference converter does; DX_RGB8 needs 8
    const uint32_t texs = layout == SDF_SBS_NODE_LAYOUT_DX_UV16 ? 32 : (layout == SDF_SBS_NODE_LAYOUT_DX_RGB8 ? 8 : 0);
    const uint32_t dist_size = (v3 + vpi - 1) / vpi;
    SdfSBSNode sn{};
    sn.data_offset = off;
    sn.pos_xy = (px << 16) | py;
    sn.pos_z_lod_size = (pz << 16) | lod;
    nodes.push_back(sn);
    values.resize(values.size() + dist_size + texs, 0u);
    for (uint32_t i = 0; i < v3; i++) {
      const float q = fmax_ref(0.0f, float(max_val) * ((g.values[n.val_off + i] + d_max) / (2 * d_max)));
      uint32_t c = q >= 4294967296.0f ? 0xFFFFFFFFu : uint32_t(q);
      c = std::min(c, max_val);
      values[off + i / vpi] |= c << (bits * (i % vpi) % 32);
    }
    // per-corner texture coordinates (u16|v16) or colour (r8|g8<<8|b8<<16); zero without a seed
    if (colour_seed != 0 && layout == SDF_SBS_NODE_LAYOUT_DX_UV16) {
      for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t u = uint32_t(65535.0 * corner_field(lod, px, py, pz, k, colour_seed, 0));
        const uint32_t v = uint32_t(65535.0 * corner_field(lod, px, py, pz, k, colour_seed, 1));
        values[off + dist_size + k] = (u << 16) | (v & 0xFFFFu);
      }
    } else if (colour_seed != 0 && layout == SDF_SBS_NODE_LAYOUT_DX_RGB8) {
      for (uint32_t k = 0; k < 8; ++k) {
        uint32_t w = 0;
        for (uint32_t c = 0; c < 3; ++c) w |= uint32_t(255.0 * corner_field(lod, px, py, pz, k, colour_seed, c) + 0.5) << (8 * c);
        values[off + dist_size + k] = w;
      }
    }
  } else {
    for (uint32_t i = 0; i < 8; i++)
      sbs_dx_rec(g, h, colour_seed, nodes, values, n.offset + i, 2 * lod, 2 * px + ((i & 4) >> 2), 2 * py + ((i & 2) >> 1),
                 2 * pz + (i & 1));
  }
}

struct BrickRef {
  uint32_t node, lod, px, py, pz;
};

void collect_bricks(const GlobalOctree& g, std::vector<BrickRef>& out, uint32_t idx, uint32_t lod, uint32_t px,
                    uint32_t py, uint32_t pz) {
  const GONode& n = g.nodes[idx];
  if (n.offset == 0) {
    if (n.type != 0) out.push_back({idx, lod, px, py, pz});
    return;
  }
  for (uint32_t i = 0; i < 8; i++)
    collect_bricks(g, out, n.offset + i, 2 * lod, 2 * px + ((i & 4) >> 2), 2 * py + ((i & 2) >> 1), 2 * pz + (i & 1));
}

}  // namespace

// global_octree_to_SBS_rec (sdf/simple/sdf_converter.cpp:156-215)
void global_octree_to_sbs_dx(const GlobalOctree& g, const SdfSBSHeader& header, uint64_t colour_seed,
                             std::vector<SdfSBSNode>& nodes, std::vector<uint32_t>& values) {
  if (header.brick_size != g.brick_size || header.brick_pad != g.brick_pad)
    throw std::invalid_argument("header brick size/pad differ from the octree's");
  nodes.clear();
  values.clear();
  sbs_dx_rec(g, header, colour_seed, nodes, values, 0, 1, 0, 0, 0);
}

// Indexed fp32 conversion (our replacement for SBS_col_to_SBS_ind, sdf_converter_exotic.cpp:432-518,
// which dequantises an 8-bit brick set and deduplicates by float position): values are kept in
// fp32, and lattice points shared by neighbouring bricks are deduplicated by their integer
// coordinate on the finest grid, so one distance value (and one albedo) exists per point.
// values_f = [distances (N)] [albedo RGB (3N)]; the 8 colour indices of a brick address the
// albedo of its corner points (layout ID32F_IRGB32F, sdf_simple_shared.h:17).
void global_octree_to_sbs_indexed(const GlobalOctree& g, uint64_t albedo_seed, std::vector<SdfSBSNode>& nodes,
                                  std::vector<uint32_t>& values, std::vector<float>& values_f,
                                  uint32_t* n_distance_values) {
  std::vector<BrickRef> bricks;
  collect_bricks(g, bricks, 0, 1, 0, 0, 0);
  const uint32_t bs = g.brick_size, pad = g.brick_pad;
  const uint32_t v_size = bs + 2 * pad + 1;
  const uint32_t v3 = v_size * v_size * v_size;
  uint32_t max_lod = 1;
  for (auto& b : bricks) max_lod = std::max(max_lod, b.lod);
  const uint64_t G = uint64_t(max_lod) * bs + 2 * pad * uint64_t(max_lod) + 1;  // points per axis incl. padding
  const bool dense = G * G * G <= (uint64_t(1) << 27);
  std::vector<uint32_t> dense_map;
  std::unordered_map<uint64_t, uint32_t> sparse_map;
  if (dense) dense_map.assign(G * G * G, 0xFFFFFFFFu);
  std::vector<float> dist, albedo;
  std::vector<uint32_t> corner_idx;
  const double inv_extent = 2.0 / (double(max_lod) * bs);
  // Albedo of a new lattice point: a smooth seeded colour field (value noise, 3 channels).
  auto add_point = [&](uint64_t fx, uint64_t fy, uint64_t fz, float v) -> uint32_t {
    const uint32_t id = uint32_t(dist.size());
    dist.push_back(v);
    const double x = -1.0 + (double(fx) - double(pad) * max_lod) * inv_extent;
    const double y = -1.0 + (double(fy) - double(pad) * max_lod) * inv_extent;
    const double z = -1.0 + (double(fz) - double(pad) * max_lod) * inv_extent;
    for (uint64_t c = 0; c < 3; ++c) {
      const double a = 0.5 + 0.45 * value_noise3(2.0 * x, 2.0 * y, 2.0 * z, albedo_seed * 3 + c);
      albedo.push_back(float(a));
    }
    return id;
  };
  nodes.clear();
  values.clear();
  nodes.reserve(bricks.size());
  values.reserve(bricks.size() * (v3 + 8));
  for (auto& b : bricks) {
    SdfSBSNode sn{};
    sn.data_offset = uint32_t(values.size());
    sn.pos_xy = (b.px << 16) | b.py;
    sn.pos_z_lod_size = (b.pz << 16) | b.lod;
    nodes.push_back(sn);
    const uint64_t scale = max_lod / b.lod;
    const GONode& n = g.nodes[b.node];
    for (uint32_t i = 0; i < v3; ++i) {
      const int64_t vx = int64_t(i / (v_size * v_size)) - int64_t(pad);
      const int64_t vy = int64_t((i / v_size) % v_size) - int64_t(pad);
      const int64_t vz = int64_t(i % v_size) - int64_t(pad);
      const uint64_t fx = uint64_t((int64_t(b.px) * bs + vx) * int64_t(scale) + int64_t(pad * max_lod));
      const uint64_t fy = uint64_t((int64_t(b.py) * bs + vy) * int64_t(scale) + int64_t(pad * max_lod));
      const uint64_t fz = uint64_t((int64_t(b.pz) * bs + vz) * int64_t(scale) + int64_t(pad * max_lod));
      const uint64_t key = (fx * G + fy) * G + fz;
      uint32_t id;
      if (dense) {
        uint32_t& slot = dense_map[key];
        if (slot == 0xFFFFFFFFu) slot = add_point(fx, fy, fz, g.values[n.val_off + i]);
        id = slot;
      } else {
        auto it = sparse_map.find(key);
        if (it == sparse_map.end()) {
          id = add_point(fx, fy, fz, g.values[n.val_off + i]);
          sparse_map.emplace(key, id);
        } else {
          id = it->second;
        }
      }
      values.push_back(id);
    }
    for (uint32_t c = 0; c < 8; ++c) {
      const uint32_t cx = pad + ((c & 4) ? bs : 0), cy = pad + ((c & 2) ? bs : 0), cz = pad + ((c & 1) ? bs : 0);
      corner_idx.push_back(values[sn.data_offset + (cx * v_size + cy) * v_size + cz]);
      values.push_back(0);  // patched below once the distance count is known
    }
  }
  const uint32_t N = uint32_t(dist.size());
  values_f.assign(size_t(N) * 4, 0.0f);
  std::copy(dist.begin(), dist.end(), values_f.begin());
  std::copy(albedo.begin(), albedo.end(), values_f.begin() + N);
  size_t ci = 0;
  for (auto& sn : nodes) {
    for (uint32_t c = 0; c < 8; ++c) values[sn.data_offset + v3 + c] = N + 3 * corner_idx[ci++];
  }
  if (n_distance_values) *n_distance_values = N;
}

}  // namespace lrt
