// Binned-SAH BVH builder (ours) that emits the reference's node layout contract.
//
// The reference builds its binary tree with Embree 3.13.5 rtcBuildBVH (binary, one
// primitive per leaf, SAH costs 1/1; utils/raytracing/cbvh_embree2.cpp:287-332). That
// library is not available, so the split decisions here are our own 32-bin SAH. Everything
// after the split decisions follows the reference:
//   * 2-primitive subtrees whose children's boxes overlap by more than the over-split
//     threshold become one leaf (ConvertTreeRec, cbvh_embree2.cpp:186-215, 39-48);
//   * nodes are laid out depth-first with the root dropped and siblings adjacent
//     (AlignNodes2, cbvh.cpp:104-169), escape index of a left child = its sibling,
//     of a right child = the parent's escape;
//   * fat pairs are numbered in depth-first pre-order (CreateFatTreeArray, cbvh_fat.cpp:100-142);
//   * triangle leaves hold PackOffsetAndSize(leafId,1) and a per-leaf start/count word
//     (cbvh.cpp:502-511); box leaves hold PackOffsetAndSize(boxId,1) (cbvh.cpp:645-655).
#include "bvh_build.h"
#include "bvh_rules.h"
#include "lrt_math.h"

#include <algorithm>
#include <cstring>
#include <limits>

namespace lrt {
namespace {

struct Box {
  float3 mn, mx;
};

Box empty_box() {
  const float inf = std::numeric_limits<float>::infinity();
  return {make3(inf, inf, inf), make3(-inf, -inf, -inf)};
}
inline void grow(Box& b, const Box& o) {
  b.mn = min3(b.mn, o.mn);
  b.mx = max3(b.mx, o.mx);
}
inline void grow(Box& b, float3 p) {
  b.mn = min3(b.mn, p);
  b.mx = max3(b.mx, p);
}
inline float comp(float3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline float surface_area(const Box& b) { return box_surface_area(b.mn, b.mx); }
float relative_box_overlap(const Box& b1, const Box& b2) { return lrt::relative_box_overlap(b1.mn, b1.mx, b2.mn, b2.mx); }

struct SahNode {
  Box box[2];
  int32_t child[2];  // >= 0: inner node index; < 0: leaf holding primitive -(c+1)
  uint32_t prims;    // primitives below this node (EvaluatePrimsInNode, cbvh_embree2.cpp:121-135)
};

class SahBuilder {
 public:
  SahBuilder(const Box* boxes, size_t n) : boxes_(boxes), idx_(n), cen_(n) {
    for (size_t i = 0; i < n; ++i) {
      idx_[i] = uint32_t(i);
      cen_[i] = (boxes[i].mn + boxes[i].mx) * 0.5f;
    }
    nodes.reserve(n);
  }
  // Requires n >= 2. Returns the root index (always 0).
  int32_t build() { return rec(0, uint32_t(idx_.size()), 1); }

  std::vector<SahNode> nodes;
  uint32_t max_depth = 0;

 private:
  static constexpr int kBins = 32;
  const Box* boxes_;
  std::vector<uint32_t> idx_;
  std::vector<float3> cen_;

  Box bounds(uint32_t b, uint32_t e) const {
    Box r = empty_box();
    for (uint32_t i = b; i < e; ++i) grow(r, boxes_[idx_[i]]);
    return r;
  }

  uint32_t split(uint32_t b, uint32_t e) {
    Box cb = empty_box();
    for (uint32_t i = b; i < e; ++i) grow(cb, cen_[idx_[i]]);
    float best = std::numeric_limits<float>::infinity();
    int best_axis = -1, best_bin = -1;
    for (int a = 0; a < 3; ++a) {
      const float lo = comp(cb.mn, a), ext = comp(cb.mx, a) - lo;
      if (!(ext > 0.0f)) continue;
      const float k = float(kBins) / ext;
      Box bb[kBins];
      uint32_t cnt[kBins];
      for (int i = 0; i < kBins; ++i) { bb[i] = empty_box(); cnt[i] = 0; }
      for (uint32_t i = b; i < e; ++i) {
        int bin = int((comp(cen_[idx_[i]], a) - lo) * k);
        bin = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
        grow(bb[bin], boxes_[idx_[i]]);
        cnt[bin]++;
      }
      float right_area[kBins];
      uint32_t right_cnt[kBins];
      Box acc = empty_box();
      uint32_t c = 0;
      for (int i = kBins - 1; i > 0; --i) {
        grow(acc, bb[i]); c += cnt[i];
        right_area[i] = c ? surface_area(acc) : 0.0f;
        right_cnt[i] = c;
      }
      acc = empty_box(); c = 0;
      for (int i = 0; i < kBins - 1; ++i) {
        grow(acc, bb[i]); c += cnt[i];
        if (c == 0 || right_cnt[i + 1] == 0) continue;
        const float cost = surface_area(acc) * float(c) + right_area[i + 1] * float(right_cnt[i + 1]);
        if (cost < best) { best = cost; best_axis = a; best_bin = i; }
      }
    }
    if (best_axis < 0) {
      // identical centroids: split the index range in half
      return b + (e - b) / 2;
    }
    const float lo = comp(cb.mn, best_axis);
    const float k = float(kBins) / (comp(cb.mx, best_axis) - lo);
    uint32_t* first = idx_.data() + b;
    uint32_t* mid = std::partition(first, idx_.data() + e, [&](uint32_t p) {
      int bin = int((comp(cen_[p], best_axis) - lo) * k);
      bin = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
      return bin <= best_bin;
    });
    uint32_t m = uint32_t(mid - idx_.data());
    if (m == b || m == e) m = b + (e - b) / 2;
    return m;
  }

  int32_t rec(uint32_t b, uint32_t e, uint32_t depth) {
    max_depth = std::max(max_depth, depth);
    const uint32_t m = split(b, e);
    const int32_t me = int32_t(nodes.size());
    nodes.push_back(SahNode{});
    nodes[me].prims = e - b;
    nodes[me].box[0] = bounds(b, m);
    nodes[me].box[1] = bounds(m, e);
    const int32_t l = (m - b == 1) ? -int32_t(idx_[b]) - 1 : rec(b, m, depth + 1);
    const int32_t r = (e - m == 1) ? -int32_t(idx_[m]) - 1 : rec(m, e, depth + 1);
    nodes[me].child[0] = l;
    nodes[me].child[1] = r;
    return me;
  }
};

void collect_prims(const std::vector<SahNode>& nodes, int32_t c, std::vector<uint32_t>& out) {
  if (c < 0) { out.push_back(uint32_t(-(c + 1))); return; }
  collect_prims(nodes, nodes[c].child[0], out);
  collect_prims(nodes, nodes[c].child[1], out);
}

BVHNode make_node(const Box& b) {
  BVHNode n;
  n.boxMin = b.mn; n.boxMax = b.mx;
  n.leftOffset = LEAF_NORMAL; n.escapeIndex = LEAF_NORMAL;
  return n;
}

// Emits the aligned node array (pairs at even offsets, root dropped) with leaves marked
// LEAF_NORMAL and their [start,count) ranges into `reordered` recorded per node.
struct Emitter {
  const std::vector<SahNode>& sah;
  uint32_t prims_in_leaf;
  float over_split;
  std::vector<BVHNode> nodes;
  std::vector<uint32_t> iv_start, iv_count;
  std::vector<uint32_t> reordered;
  uint32_t merged = 0;

  void emit_leaf(uint32_t slot, int32_t c) {
    std::vector<uint32_t> tmp;
    collect_prims(sah, c, tmp);
    iv_start[slot] = uint32_t(reordered.size());
    iv_count[slot] = uint32_t(tmp.size());
    reordered.insert(reordered.end(), tmp.begin(), tmp.end());
    nodes[slot].leftOffset = LEAF_NORMAL;
  }

  uint32_t emit(int32_t n, uint32_t parent_escape) {
    const uint32_t off = uint32_t(nodes.size());
    nodes.push_back(make_node(sah[n].box[0]));
    nodes.push_back(make_node(sah[n].box[1]));
    iv_start.resize(nodes.size(), 0);
    iv_count.resize(nodes.size(), 0);
    nodes[off + 0].escapeIndex = off + 1;
    nodes[off + 1].escapeIndex = parent_escape;
    for (int k = 0; k < 2; ++k) {
      const int32_t c = sah[n].child[k];
      const uint32_t slot = off + uint32_t(k);
      if (c < 0) { emit_leaf(slot, c); continue; }
      if (sah[c].prims <= prims_in_leaf &&
          relative_box_overlap(sah[c].box[0], sah[c].box[1]) > over_split) {
        merged++;
        emit_leaf(slot, c);
        continue;
      }
      const uint32_t child_off = emit(c, k == 0 ? off + 1 : parent_escape);
      nodes[slot].leftOffset = child_off;
    }
    return off;
  }
};

Emitter build_aligned(const std::vector<Box>& boxes, uint32_t prims_in_leaf, float over_split,
                      BuildStats* stats) {
  SahBuilder b(boxes.data(), boxes.size());
  b.build();
  Emitter em{b.nodes, prims_in_leaf, over_split, {}, {}, {}, {}, 0};
  em.nodes.reserve(boxes.size() * 2);
  em.emit(0, ESCAPE_ROOT);
  if (stats) {
    stats->max_depth = b.max_depth;
    stats->merged_leaves = em.merged;
  }
  return em;
}

// CreateFatTreeArray: with pairs at even offsets in pre-order, pair index = offset / 2.
std::vector<BVHNodePair> to_fat(const std::vector<BVHNode>& aligned) {
  std::vector<BVHNodePair> fat(aligned.size() / 2);
  for (size_t p = 0; p < fat.size(); ++p) {
    fat[p].left = aligned[2 * p];
    fat[p].right = aligned[2 * p + 1];
    if (!(fat[p].left.leftOffset & LEAF_BIT)) fat[p].left.leftOffset /= 2;
    if (!(fat[p].right.leftOffset & LEAF_BIT)) fat[p].right.leftOffset /= 2;
  }
  return fat;
}

}  // namespace

FatBlas build_blas_triangles(const float4* verts, size_t n_verts, const uint32_t* tri_idx, size_t n_idx,
                             int prims_in_leaf, BuildStats* stats) {
  FatBlas out;
  const size_t n_tris = n_idx / 3;
  if (n_tris == 0) return out;
  if (n_tris <= size_t(prims_in_leaf)) {
    // BuildBVHSmall (cbvh.cpp:325-360): one leaf with every triangle, one empty sibling.
    Box all = empty_box();
    for (size_t i = 0; i < n_verts; ++i) grow(all, xyz(verts[i]));
    BVHNodePair p;
    p.left = make_node(all);
    p.left.leftOffset = pack_offset_and_size(0, 1);
    p.left.escapeIndex = 1;
    std::memset(&p.right, 0, sizeof(p.right));
    p.right.leftOffset = LEAF_NORMAL;
    p.right.escapeIndex = ESCAPE_ROOT;
    out.nodes.push_back(p);
    for (size_t i = 0; i < n_tris; ++i) out.indices.push_back(uint32_t(i));
    out.startCount.push_back(pack_offset_and_size(0, uint32_t(n_tris)));
    if (stats) { stats->max_depth = 1; stats->leaves = 1; stats->merged_leaves = 0; }
    return out;
  }
  std::vector<Box> boxes(n_tris);
  for (size_t i = 0; i < n_tris; ++i) {
    const float3 a = xyz(verts[tri_idx[3 * i + 0]]);
    const float3 b = xyz(verts[tri_idx[3 * i + 1]]);
    const float3 c = xyz(verts[tri_idx[3 * i + 2]]);
    boxes[i].mn = min3(a, min3(b, c));
    boxes[i].mx = max3(a, max3(b, c));
  }
  // BuildBVHEmbree (cbvh_embree2.cpp:344-347): over-split threshold 0.5 below 4 prims per leaf.
  const float over_split = over_split_threshold(prims_in_leaf);
  Emitter em = build_aligned(boxes, uint32_t(prims_in_leaf), over_split, stats);
  uint32_t leaf_id = 0;
  for (size_t i = 0; i < em.nodes.size(); ++i) {
    if (em.nodes[i].leftOffset == LEAF_NORMAL) {
      em.nodes[i].leftOffset = pack_offset_and_size(leaf_id++, 1);
      out.startCount.push_back(pack_offset_and_size(em.iv_start[i], em.iv_count[i]));
    }
  }
  if (stats) stats->leaves = leaf_id;
  out.nodes = to_fat(em.nodes);
  out.indices = std::move(em.reordered);
  return out;
}

FatBlas build_blas_boxes(const BVHNode* in, size_t n, BuildStats* stats) {
  FatBlas out;
  if (n == 0) return out;
  if (n == 1) {
    BVHNodePair p;
    p.left = in[0];
    p.left.leftOffset = pack_offset_and_size(0, 1);
    p.left.escapeIndex = 1;
    std::memset(&p.right, 0, sizeof(p.right));
    p.right.leftOffset = LEAF_NORMAL;
    p.right.escapeIndex = ESCAPE_ROOT;
    out.nodes.push_back(p);
    out.indices.push_back(0);
    if (stats) { stats->max_depth = 1; stats->leaves = 1; }
    return out;
  }
  std::vector<Box> boxes(n);
  for (size_t i = 0; i < n; ++i) { boxes[i].mn = in[i].boxMin; boxes[i].mx = in[i].boxMax; }
  Emitter em = build_aligned(boxes, 1, 0.25f, stats);
  uint32_t leaves = 0;
  for (size_t i = 0; i < em.nodes.size(); ++i) {
    if (em.nodes[i].leftOffset == LEAF_NORMAL) {
      em.nodes[i].leftOffset = pack_offset_and_size(em.reordered[em.iv_start[i]], em.iv_count[i]);
      leaves++;
    }
  }
  if (stats) stats->leaves = leaves;
  out.nodes = to_fat(em.nodes);
  out.indices = std::move(em.reordered);
  return out;
}

std::vector<BVHNode> build_tlas(const float4* box_min, const float4* box_max, size_t n) {
  std::vector<BVHNode> out;
  if (n == 0) return out;
  if (n == 1) {
    BVHNode node;
    node.boxMin = xyz(box_min[0]);
    node.boxMax = xyz(box_max[0]);
    node.leftOffset = LEAF_BIT;
    node.escapeIndex = LEAF_NORMAL;
    out.push_back(node);
    return out;
  }
  std::vector<Box> boxes(n);
  for (size_t i = 0; i < n; ++i) { boxes[i].mn = xyz(box_min[i]); boxes[i].mx = xyz(box_max[i]); }
  Emitter em = build_aligned(boxes, 1, 0.25f, nullptr);
  for (size_t i = 0; i < em.nodes.size(); ++i)
    if (em.nodes[i].leftOffset == LEAF_NORMAL)
      em.nodes[i].leftOffset = pack_offset_and_size(em.reordered[em.iv_start[i]], em.iv_count[i]);
  return em.nodes;
}

}  // namespace lrt
