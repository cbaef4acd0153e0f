// GPU construction of the global sparse octree of a sphere-union SDF: the same tree, node
// numbering and value order as the host restatement of sdf_to_global_octree
// (sdf/build/sparse_octree_builder.cpp:37-264).
//
// The host recursion is depth-first; here the border subtrees grow level by level (one block per
// node samples its (bs + 2 pad + 1)^3 distances, reduces min/max and decides the split with the
// same float arithmetic), then the depth-first numbering is recovered from subtree sizes: a
// node's children block sits at 1 + 8 x (number of split nodes before it in pre-order), its
// values at v^3 x (its pre-order rank), exactly where the recursion would have put them.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "gen.h"
#include "lrt_math.h"

namespace lrt {

namespace {

// SphereUnion::operator() with the host's operation order.
__device__ __forceinline__ float sphere_union(const float* __restrict__ sp, uint32_t n, float3 p) {
  float d = 3.40282347e+38f;   // std::numeric_limits<float>::max()
  for (uint32_t i = 0; i < n; ++i) {
    const float dx = p.x - sp[4 * i + 0], dy = p.y - sp[4 * i + 1], dz = p.z - sp[4 * i + 2];
    const float di = std::sqrt(dx * dx + dy * dy + dz * dz) - sp[4 * i + 3];
    d = fmin_ref(d, di);
  }
  return d;
}

struct LevelNode {
  float px, py, pz;    // position in the level's grid (exact integers)
  uint32_t subtree;    // border subtree
  uint32_t parent;     // index in the previous level (subtree roots: the large node)
  uint32_t slot;       // child slot 0..7
};

struct NodeOut {       // per level node: decision and numbering
  uint32_t split, type, size_all, size_split;
  uint32_t pre, blockpre, pad0, pad1;
};

// Samples of one node: q = p + (i, j, k) / bs for i, j, k in [-pad, bs + pad], value
// sdf(2 (q d) - 1), in the host's loop order (x outer, z inner). Returns min and max through the
// block; the center value comes from thread 0.
__device__ void sample_node(const float* __restrict__ sp, uint32_t n_sp, float3 p, float d, int bs, int pad,
                            float* __restrict__ vals, float& vmin, float& vmax, float& center) {
  __shared__ float smin[256], smax[256];
  const int v = bs + 2 * pad + 1, v3 = v * v * v;
  float lo = 1000.0f, hi = -1000.0f;
  for (int idx = int(threadIdx.x); idx < v3; idx += int(blockDim.x)) {
    const int i = idx / (v * v) - pad, j = (idx / v) % v - pad, k = idx % v - pad;
    const float3 q = p + make3(float(i) / float(bs), float(j) / float(bs), float(k) / float(bs));
    const float val = sphere_union(sp, n_sp, 2.0f * (q * d) - make3(1, 1, 1));
    if (vals) vals[idx] = val;
    lo = fmin_ref(lo, val);
    hi = fmax_ref(hi, val);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      smin[threadIdx.x] = fmin_ref(smin[threadIdx.x], smin[threadIdx.x + st]);
      smax[threadIdx.x] = fmax_ref(smax[threadIdx.x], smax[threadIdx.x + st]);
    }
    __syncthreads();
  }
  vmin = smin[0];
  vmax = smax[0];
  center = sphere_union(sp, n_sp, 2.0f * ((p + make3(0.5f, 0.5f, 0.5f)) * d) - make3(1, 1, 1));
}

// Large (top) nodes: their values and centre values.
__global__ __launch_bounds__(256) void k_large_nodes(const float* __restrict__ sp, uint32_t n_sp,
                                                      const float4* __restrict__ pd, int bs, int pad,
                                                      float* __restrict__ values, float* __restrict__ centers) {
  const uint32_t node = blockIdx.x;
  const float4 q = pd[node];
  const int v = bs + 2 * pad + 1;
  float lo, hi, c;
  sample_node(sp, n_sp, make3(q.x, q.y, q.z), q.w, bs, pad, values + size_t(node) * v * v * v, lo, hi, c);
  if (threadIdx.x == 0) centers[node] = c;
}

// One level of the border subtrees: values, split decision and node type
// (add_global_node_rec, sparse_octree_builder.cpp:37-104, fill_internal_volume = false).
__global__ __launch_bounds__(256) void k_level(const float* __restrict__ sp, uint32_t n_sp,
                                               const LevelNode* __restrict__ nodes, float d, int bs, int pad,
                                               bool may_split, float* __restrict__ values, NodeOut* __restrict__ out) {
  const uint32_t node = blockIdx.x;
  const LevelNode nd = nodes[node];
  const int v = bs + 2 * pad + 1;
  float lo, hi, c;
  sample_node(sp, n_sp, make3(nd.px, nd.py, nd.pz), d, bs, pad, values + size_t(node) * v * v * v, lo, hi, c);
  if (threadIdx.x == 0) {
    const bool split = may_split && (std::fabs(c) < std::sqrt(3.0f) * d || lo * hi <= 0);
    const bool active = lo <= 0 && hi >= 0;
    NodeOut o{};
    o.split = split ? 1u : 0u;
    o.type = active ? (split ? 3u : 1u) : (split ? 2u : 0u);
    out[node] = o;
  }
}

// Children of the split nodes: 8 consecutive entries per split node, in slot order.
__global__ void k_spawn(const LevelNode* __restrict__ nodes, const NodeOut* __restrict__ out,
                        const uint32_t* __restrict__ child_base, uint32_t count, LevelNode* __restrict__ next) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count || !out[i].split) return;
  const LevelNode p = nodes[i];
  const uint32_t b = 8u * child_base[i];
  for (uint32_t c = 0; c < 8; ++c) {
    LevelNode ch;
    ch.px = 2.0f * p.px + float((c & 4) >> 2);
    ch.py = 2.0f * p.py + float((c & 2) >> 1);
    ch.pz = 2.0f * p.pz + float(c & 1);
    ch.subtree = p.subtree;
    ch.parent = i;
    ch.slot = c;
    next[b + c] = ch;
  }
}

// Subtree sizes, deepest level first: all nodes and split nodes below and including each node.
__global__ void k_sizes(NodeOut* __restrict__ out, const uint32_t* __restrict__ child_base,
                        const NodeOut* __restrict__ next_out, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  NodeOut o = out[i];
  o.size_all = 1u;
  o.size_split = o.split;
  if (o.split) {
    const uint32_t b = 8u * child_base[i];
    for (uint32_t c = 0; c < 8; ++c) {
      o.size_all += next_out[b + c].size_all;
      o.size_split += next_out[b + c].size_split;
    }
  }
  out[i].size_all = o.size_all;
  out[i].size_split = o.size_split;
}

// Pre-order ranks, top level first: a child's rank follows its parent's and its elder siblings'
// subtrees; blockpre counts the split nodes before a node in pre-order.
__global__ void k_preorder(const NodeOut* __restrict__ out, const uint32_t* __restrict__ child_base, uint32_t count,
                           NodeOut* __restrict__ next_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count || !out[i].split) return;
  const NodeOut o = out[i];
  uint32_t pre = o.pre + 1u, blockpre = o.blockpre + 1u;
  const uint32_t b = 8u * child_base[i];
  for (uint32_t c = 0; c < 8; ++c) {
    next_out[b + c].pre = pre;
    next_out[b + c].blockpre = blockpre;
    pre += next_out[b + c].size_all;
    blockpre += next_out[b + c].size_split;
  }
}

// Output records of the subtree nodes below their roots.
struct GOut {
  uint32_t type, val_off, offset;
};

__global__ void k_emit_level(const LevelNode* __restrict__ nodes, const NodeOut* __restrict__ out,
                             const NodeOut* __restrict__ parent_out, uint32_t count, const uint32_t* __restrict__ base,
                             const uint32_t* __restrict__ val_base, uint32_t v3, GOut* __restrict__ g) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const LevelNode nd = nodes[i];
  const NodeOut o = out[i];
  const uint32_t local = 1u + 8u * parent_out[nd.parent].blockpre + nd.slot;
  const uint32_t b = nd.subtree;
  GOut r;
  r.type = o.type;
  r.offset = o.split ? base[b] + (1u + 8u * o.blockpre) - 1u : 0u;
  r.val_off = val_base[b] + v3 + (o.pre - 1u) * v3;
  g[base[b] + local - 1u] = r;
}

__global__ void k_copy_values(const float* __restrict__ src, const NodeOut* __restrict__ out,
                              const LevelNode* __restrict__ nodes, const uint32_t* __restrict__ val_base, uint32_t v3,
                              float* __restrict__ dst) {
  const uint32_t node = blockIdx.x;
  const NodeOut o = out[node];
  const uint32_t dst0 = val_base[nodes[node].subtree] + v3 + (o.pre - 1u) * v3;
  for (uint32_t k = threadIdx.x; k < v3; k += blockDim.x) dst[size_t(dst0) + k] = src[size_t(node) * v3 + k];
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("GPU octree build: ") + what + ": " + hipGetErrorString(e));
}

struct Dev {
  void* p = nullptr;
  Dev() = default;
  explicit Dev(size_t bytes) { check(hipMalloc(&p, std::max<size_t>(bytes, 16)), "allocation"); }
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  Dev(Dev&& o) noexcept : p(o.p) { o.p = nullptr; }
  Dev& operator=(Dev&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~Dev() { if (p) (void)hipFree(p); }
  template <class T> T* get() const { return static_cast<T*>(p); }
};

inline dim3 grid_for(size_t n) { return dim3(uint32_t((n + 255) / 256)); }

uint32_t exclusive_scan(const uint32_t* flags_dev, uint32_t* out_dev, uint32_t n) {
  // out[i] = number of split flags before i; returns the total
  size_t temp = 0;
  check(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, flags_dev, out_dev, int(n)), "scan size");
  Dev t(temp);
  check(hipcub::DeviceScan::ExclusiveSum(t.p, temp, flags_dev, out_dev, int(n)), "scan");
  uint32_t last_base = 0, last_flag = 0;
  check(hipMemcpy(&last_base, out_dev + n - 1, 4, hipMemcpyDeviceToHost), "scan total");
  check(hipMemcpy(&last_flag, flags_dev + n - 1, 4, hipMemcpyDeviceToHost), "scan total");
  return last_base + last_flag;
}

__global__ void k_copy_large(const float* __restrict__ src, const uint32_t* __restrict__ dst_off, uint32_t v3,
                             float* __restrict__ dst) {
  const uint32_t node = blockIdx.x;
  for (uint32_t k = threadIdx.x; k < v3; k += blockDim.x) dst[size_t(dst_off[node]) + k] = src[size_t(node) * v3 + k];
}

__global__ void k_split_flags(const NodeOut* __restrict__ out, uint32_t count, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) flags[i] = out[i].split;
}

__global__ void k_root_rank(NodeOut* __restrict__ out, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    out[i].pre = 0u;
    out[i].blockpre = 0u;
  }
}

}  // namespace

// sdf_to_global_octree on the current HIP device; the result equals the host build's.
void gpu_sdf_to_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                              GlobalOctree& out) {
  if (depth < 1 || depth > 16) throw std::invalid_argument("octree depth must be 1..16");
  if (brick_size < 1 || brick_size > 16) throw std::invalid_argument("brick size must be 1..16");
  const int bs = int(brick_size), pad = int(brick_pad);
  const uint32_t v = brick_size + 2 * brick_pad + 1, v3 = v * v * v;
  // threads per node-block: the samples of one node, rounded to whole waves (a power of two for
  // the block reduction)
  const uint32_t tpb = v3 <= 64 ? 64u : (v3 <= 128 ? 128u : 256u);
  out.brick_size = brick_size;
  out.brick_pad = brick_pad;
  // large nodes: the complete tree down to level min(depth, 4), breadth first (host restatement)
  const uint32_t min_remove_level = std::min(depth, 4u);
  struct Large { float px, py, pz, d; uint32_t level, children; };
  std::vector<Large> large;
  large.push_back({0, 0, 0, 1.0f, 0u, 0u});
  for (size_t i = 0; i < large.size(); ++i) {
    if (large[i].level < min_remove_level) {
      large[i].children = uint32_t(large.size());
      for (int j = 0; j < 8; j++)
        large.push_back({2.0f * large[i].px + float((j & 4) >> 2), 2.0f * large[i].py + float((j & 2) >> 1),
                         2.0f * large[i].pz + float(j & 1), large[i].d / 2, large[i].level + 1, 0u});
    }
  }
  const uint32_t nl = uint32_t(large.size());
  Dev d_sp(sdf.n * 16);
  check(hipMemcpy(d_sp.p, sdf.spheres, sdf.n * 16, hipMemcpyHostToDevice), "spheres");
  std::vector<float4> pd(nl);
  for (uint32_t i = 0; i < nl; ++i) pd[i] = make4(large[i].px, large[i].py, large[i].pz, large[i].d);
  Dev d_pd(nl * sizeof(float4)), d_lv(size_t(nl) * v3 * 4), d_lc(nl * 4);
  check(hipMemcpy(d_pd.p, pd.data(), nl * sizeof(float4), hipMemcpyHostToDevice), "large nodes");
  hipLaunchKernelGGL(k_large_nodes, dim3(nl), dim3(tpb), 0, nullptr, d_sp.get<float>(), uint32_t(sdf.n),
                     d_pd.get<float4>(), bs, pad, d_lv.get<float>(), d_lc.get<float>());
  check(hipGetLastError(), "large nodes");
  std::vector<float> centers(nl);
  check(hipMemcpy(centers.data(), d_lc.p, nl * 4, hipMemcpyDeviceToHost), "centres");
  // border roots (is_border, sparse_octree_common.h:84-87, in double as the host)
  std::vector<uint32_t> border;
  std::vector<LevelNode> roots;
  for (uint32_t i = 0; i < nl; ++i) {
    const int level = int(large[i].level);
    const bool is_border = level < 2 ? true : double(std::fabs(centers[i])) < std::sqrt(3.0) * std::pow(2.0, -level);
    if (large[i].children == 0 && is_border) {
      roots.push_back({large[i].px, large[i].py, large[i].pz, uint32_t(border.size()), i, 0u});
      border.push_back(i);
    }
  }
  // level-by-level growth of every border subtree
  struct Level {
    uint32_t count = 0;
    float d = 0;
    Dev nodes, out, values, child_base;
  };
  std::vector<Level> levels;
  if (!roots.empty()) {
    Level l0;
    l0.count = uint32_t(roots.size());
    l0.d = 1.0f / float(1u << min_remove_level);
    l0.nodes = Dev(roots.size() * sizeof(LevelNode));
    check(hipMemcpy(l0.nodes.p, roots.data(), roots.size() * sizeof(LevelNode), hipMemcpyHostToDevice), "roots");
    levels.push_back(std::move(l0));
  }
  for (uint32_t lev = min_remove_level; !levels.empty(); ++lev) {
    Level& L = levels.back();
    L.out = Dev(size_t(L.count) * sizeof(NodeOut));
    L.values = Dev(size_t(L.count) * v3 * 4);
    hipLaunchKernelGGL(k_level, dim3(L.count), dim3(tpb), 0, nullptr, d_sp.get<float>(), uint32_t(sdf.n),
                       L.nodes.get<LevelNode>(), L.d, bs, pad, lev < depth, L.values.get<float>(), L.out.get<NodeOut>());
    check(hipGetLastError(), "level");
    Dev flags(size_t(L.count) * 4);
    L.child_base = Dev(size_t(L.count) * 4);
    hipLaunchKernelGGL(k_split_flags, grid_for(L.count), dim3(256), 0, nullptr, L.out.get<NodeOut>(), L.count,
                       flags.get<uint32_t>());
    const uint32_t n_split = exclusive_scan(flags.get<uint32_t>(), L.child_base.get<uint32_t>(), L.count);
    if (n_split == 0) break;
    if (uint64_t(n_split) * 8 >= (1ull << 31)) throw std::runtime_error("GPU octree build: too many nodes");
    Level next;
    next.count = 8 * n_split;
    next.d = L.d / 2;
    next.nodes = Dev(size_t(next.count) * sizeof(LevelNode));
    hipLaunchKernelGGL(k_spawn, grid_for(L.count), dim3(256), 0, nullptr, L.nodes.get<LevelNode>(),
                       L.out.get<NodeOut>(), L.child_base.get<uint32_t>(), L.count, next.nodes.get<LevelNode>());
    check(hipGetLastError(), "spawn");
    levels.push_back(std::move(next));
  }
  // sizes bottom-up, ranks top-down
  for (size_t l = levels.size(); l-- > 0;) {
    Level& L = levels[l];
    const NodeOut* next_out = l + 1 < levels.size() ? levels[l + 1].out.get<NodeOut>() : nullptr;
    hipLaunchKernelGGL(k_sizes, grid_for(L.count), dim3(256), 0, nullptr, L.out.get<NodeOut>(),
                       L.child_base.get<uint32_t>(), next_out, L.count);
  }
  if (!levels.empty())
    hipLaunchKernelGGL(k_root_rank, grid_for(levels[0].count), dim3(256), 0, nullptr, levels[0].out.get<NodeOut>(),
                       levels[0].count);
  for (size_t l = 0; l + 1 < levels.size(); ++l)
    hipLaunchKernelGGL(k_preorder, grid_for(levels[l].count), dim3(256), 0, nullptr, levels[l].out.get<NodeOut>(),
                       levels[l].child_base.get<uint32_t>(), levels[l].count, levels[l + 1].out.get<NodeOut>());
  check(hipGetLastError(), "numbering");
  // placement: large nodes first (breadth-first), then each border subtree below its root in
  // large-node order; values interleaved per large node as the host merge does
  std::vector<NodeOut> root_out(roots.size());
  if (!roots.empty())
    check(hipMemcpy(root_out.data(), levels[0].out.p, roots.size() * sizeof(NodeOut), hipMemcpyDeviceToHost), "roots");
  std::vector<uint32_t> base(roots.size()), val_base(roots.size());
  std::vector<uint32_t> large_val(nl);
  uint64_t n_nodes = nl, n_vals = 0;
  {
    size_t b = 0;
    for (uint32_t i = 0; i < nl; ++i) {
      large_val[i] = uint32_t(n_vals);
      n_vals += v3;
      if (b < border.size() && border[b] == i) {
        const uint32_t extra = root_out[b].size_all - 1u;
        base[b] = uint32_t(n_nodes);
        val_base[b] = large_val[i];
        n_nodes += extra;
        n_vals += uint64_t(extra) * v3;
        ++b;
      }
    }
  }
  if (n_vals >= (1ull << 32) || n_nodes >= (1ull << 31)) throw std::runtime_error("GPU octree build: too large");
  Dev d_g(size_t(n_nodes) * sizeof(GOut)), d_vals(size_t(n_vals) * 4), d_base(base.size() * 4 + 4),
      d_vbase(val_base.size() * 4 + 4);
  if (!base.empty()) {
    check(hipMemcpy(d_base.p, base.data(), base.size() * 4, hipMemcpyHostToDevice), "bases");
    check(hipMemcpy(d_vbase.p, val_base.data(), val_base.size() * 4, hipMemcpyHostToDevice), "bases");
  }
  for (size_t l = 1; l < levels.size(); ++l) {
    Level& L = levels[l];
    hipLaunchKernelGGL(k_emit_level, grid_for(L.count), dim3(256), 0, nullptr, L.nodes.get<LevelNode>(),
                       L.out.get<NodeOut>(), levels[l - 1].out.get<NodeOut>(), L.count, d_base.get<uint32_t>(),
                       d_vbase.get<uint32_t>(), v3, d_g.get<GOut>());
    hipLaunchKernelGGL(k_copy_values, dim3(L.count), dim3(tpb), 0, nullptr, L.values.get<float>(),
                       L.out.get<NodeOut>(), L.nodes.get<LevelNode>(), d_vbase.get<uint32_t>(), v3,
                       d_vals.get<float>());
  }
  Dev d_lvo(size_t(nl) * 4);
  check(hipMemcpy(d_lvo.p, large_val.data(), size_t(nl) * 4, hipMemcpyHostToDevice), "large offsets");
  hipLaunchKernelGGL(k_copy_large, dim3(nl), dim3(tpb), 0, nullptr, d_lv.get<float>(), d_lvo.get<uint32_t>(), v3,
                     d_vals.get<float>());
  check(hipGetLastError(), "emit");
  std::vector<GOut> g(n_nodes);
  out.values.resize(n_vals);
  check(hipMemcpy(g.data(), d_g.p, g.size() * sizeof(GOut), hipMemcpyDeviceToHost), "nodes");
  check(hipMemcpy(out.values.data(), d_vals.p, out.values.size() * 4, hipMemcpyDeviceToHost), "values");
  out.nodes.assign(n_nodes, GONode{3, 0, 0});
  size_t b = 0;
  for (uint32_t i = 0; i < nl; ++i) {
    out.nodes[i].type = 3;   // every top node is NODE (sparse_octree_builder.cpp:202)
    out.nodes[i].val_off = large_val[i];
    out.nodes[i].offset = large[i].children;
    if (b < border.size() && border[b] == i) {
      if (root_out[b].split) out.nodes[i].offset = base[b];
      ++b;
    }
  }
  for (uint64_t k = nl; k < n_nodes; ++k) {
    out.nodes[k].type = uint8_t(g[k].type);
    out.nodes[k].val_off = g[k].val_off;
    out.nodes[k].offset = g[k].offset;
  }
}

}  // namespace lrt
