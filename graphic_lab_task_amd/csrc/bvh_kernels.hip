// GPU construction of the fat BLAS layout and GPU refit of the TLAS.
//
// BLAS: a linear BVH (Karras 2012, "Maximizing Parallelism in the Construction of BVHs, Octrees
// and k-d Trees") over 63-bit Morton codes of the primitive centroids, then the same layout
// contract as the host builder: the 2-primitive merge rule of ConvertTreeRec
// (cbvh_embree2.cpp:186-215, 39-48), fat pairs in depth-first pre-order with the root dropped
// and siblings adjacent (AlignNodes2, cbvh.cpp:143-169; CreateFatTreeArray, cbvh_fat.cpp:128-142),
// escape index of a left child = its sibling, of a right child = the parent's escape, triangle
// leaves holding PackOffsetAndSize(leafId, 1) and a per-leaf start/count word (cbvh.cpp:502-511),
// box leaves PackOffsetAndSize(boxId, 1) (cbvh.cpp:645-655). Every kernel works on one node or
// one primitive per thread; the hierarchy's boxes are reduced bottom-up, each parent by the
// second of its two children to arrive (agent-scope acquire/release counters).
//
// TLAS refit: the node boxes of an existing TLAS recomputed from the instance boxes, bottom-up
// in the same way, for commits after BVHRT::UpdateInstance (BVH2Common_host.cpp:1303-1335)
// that keep the instance set.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "bvh_rules.h"
#include "kernels.h"

namespace lrt {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t LEAF_REF = 0x80000000u;   // child reference to leaf j of the sorted order

struct PBox {
  float3 mn;
  float pad0;
  float3 mx;
  float pad1;
};

__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmin_ref(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax_ref(v, __shfl_xor(v, o));
  return v;
}

// Centroid bounds: every wave reduces its lanes, lane 0 folds them into six ordered words
// (min x, y, z as ordered-min, max x, y, z as ordered-max).
__device__ void fold_centroid(bool valid, float3 c, uint32_t* cb) {
  const float inf = __uint_as_float(0x7F800000u);
  const float x0 = wave_min(valid ? c.x : inf), y0 = wave_min(valid ? c.y : inf), z0 = wave_min(valid ? c.z : inf);
  const float x1 = wave_max(valid ? c.x : -inf), y1 = wave_max(valid ? c.y : -inf), z1 = wave_max(valid ? c.z : -inf);
  if ((threadIdx.x & 63u) == 0u) {
    atomicMin(cb + 0, f2ord(x0)); atomicMin(cb + 1, f2ord(y0)); atomicMin(cb + 2, f2ord(z0));
    atomicMax(cb + 3, f2ord(x1)); atomicMax(cb + 4, f2ord(y1)); atomicMax(cb + 5, f2ord(z1));
  }
}

__global__ void k_init_bounds(uint32_t* cb) {
  if (threadIdx.x < 3) cb[threadIdx.x] = 0xFFFFFFFFu;
  else if (threadIdx.x < 6) cb[threadIdx.x] = 0u;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_tri_boxes(const float4* __restrict__ verts, const uint32_t* __restrict__ idx, uint32_t n,
                            PBox* __restrict__ boxes, uint32_t* cb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  float3 c = make3(0, 0, 0);
  if (valid) {
    const float3 a = xyz(verts[idx[3 * i + 0]]), b = xyz(verts[idx[3 * i + 1]]), d = xyz(verts[idx[3 * i + 2]]);
    PBox r;
    r.mn = min3(a, min3(b, d));
    r.mx = max3(a, max3(b, d));
    r.pad0 = r.pad1 = 0.0f;
    boxes[i] = r;
    c = (r.mn + r.mx) * 0.5f;
  }
  fold_centroid(valid, c, cb);
}

__global__ void k_aabb_boxes(const BVHNode* __restrict__ in, uint32_t n, PBox* __restrict__ boxes, uint32_t* cb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  float3 c = make3(0, 0, 0);
  if (valid) {
    PBox r;
    r.mn = in[i].boxMin;
    r.mx = in[i].boxMax;
    r.pad0 = r.pad1 = 0.0f;
    boxes[i] = r;
    c = (r.mn + r.mx) * 0.5f;
  }
  fold_centroid(valid, c, cb);
}

__device__ __forceinline__ uint64_t expand21(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t quantise21(float c, float lo, float hi) {
  const float ext = hi - lo;
  float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
  u = fmin_ref(fmax_ref(u * 2097152.0f, 0.0f), 2097151.0f);
  return uint32_t(u);
}

__global__ void k_morton(const PBox* __restrict__ boxes, uint32_t n, const uint32_t* __restrict__ cb,
                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PBox b = boxes[i];
  const float3 c = (b.mn + b.mx) * 0.5f;
  const uint32_t qx = quantise21(c.x, ord2f(cb[0]), ord2f(cb[3]));
  const uint32_t qy = quantise21(c.y, ord2f(cb[1]), ord2f(cb[4]));
  const uint32_t qz = quantise21(c.z, ord2f(cb[2]), ord2f(cb[5]));
  keys[i] = (expand21(qx) << 2) | (expand21(qy) << 1) | expand21(qz);
  vals[i] = i;
}

// Length of the common prefix of sorted keys i and j; equal keys continue with the indices.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz(uint32_t(i ^ j));
  return __clzll(a ^ b);
}

// Karras's node-parallel hierarchy: internal node i covers the sorted range [first, last] and
// splits it where the common prefix drops. Children: internal index, or LEAF_REF | j.
__global__ void k_hierarchy(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ child,
                            uint32_t* __restrict__ iparent, uint32_t* __restrict__ lparent,
                            uint32_t* __restrict__ first, uint32_t* __restrict__ last) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t c0 = (lo == gamma) ? (LEAF_REF | uint32_t(gamma)) : uint32_t(gamma);
  const uint32_t c1 = (hi == gamma + 1) ? (LEAF_REF | uint32_t(gamma + 1)) : uint32_t(gamma + 1);
  child[2 * i + 0] = c0;
  child[2 * i + 1] = c1;
  if (c0 & LEAF_REF) lparent[gamma] = uint32_t(i); else iparent[gamma] = uint32_t(i);
  if (c1 & LEAF_REF) lparent[gamma + 1] = uint32_t(i); else iparent[gamma + 1] = uint32_t(i);
  first[i] = uint32_t(lo);
  last[i] = uint32_t(hi);
  if (i == 0) iparent[0] = NONE;
}

struct NodeInfo {
  uint32_t pairs;      // fat pairs of the subtree (0 when the node is a leaf of its parent)
  uint32_t leaves;     // leaf slots of the subtree
  uint32_t collapsed;  // 1: merged into one leaf by the overlap rule
  uint32_t pad;
};

__device__ __forceinline__ void child_box(uint32_t c, const uint32_t* __restrict__ vals, const PBox* boxes,
                                          const PBox* nbox, float3& mn, float3& mx) {
  if (c & LEAF_REF) {
    const PBox b = boxes[vals[c & ~LEAF_REF]];
    mn = b.mn; mx = b.mx;
  } else {
    mn = nbox[c].mn; mx = nbox[c].mx;
  }
}

// Bottom-up reduction, one thread per leaf: the second child to reach a parent computes the
// parent's box, its merge decision and its subtree counts, then climbs on.
__global__ void k_bottom_up(uint32_t n, const uint32_t* __restrict__ vals, const PBox* __restrict__ boxes,
                            const uint32_t* __restrict__ child, const uint32_t* __restrict__ iparent,
                            const uint32_t* __restrict__ lparent, const uint32_t* __restrict__ first,
                            const uint32_t* __restrict__ last, uint32_t prims_in_leaf, float over_split,
                            PBox* nbox, NodeInfo* info, uint32_t* cnt) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t p = lparent[j];
  for (;;) {
    const uint32_t old = __hip_atomic_fetch_add(cnt + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0u) return;
    float3 mn0, mx0, mn1, mx1;
    const uint32_t c0 = child[2 * p], c1 = child[2 * p + 1];
    child_box(c0, vals, boxes, nbox, mn0, mx0);
    child_box(c1, vals, boxes, nbox, mn1, mx1);
    PBox b;
    b.mn = min3(mn0, mn1);
    b.mx = max3(mx0, mx1);
    b.pad0 = b.pad1 = 0.0f;
    nbox[p] = b;
    NodeInfo r;
    const uint32_t prims = last[p] - first[p] + 1u;
    r.collapsed = (p != 0u && prims <= prims_in_leaf && relative_box_overlap(mn0, mx0, mn1, mx1) > over_split) ? 1u : 0u;
    r.pairs = 1u;
    r.leaves = 0u;
    for (int k = 0; k < 2; ++k) {
      const uint32_t c = k ? c1 : c0;
      if ((c & LEAF_REF) || info[c].collapsed) {
        r.leaves += 1u;
      } else {
        r.pairs += info[c].pairs;
        r.leaves += info[c].leaves;
      }
    }
    r.pad = 0u;
    info[p] = r;
    if (p == 0u) return;
    p = iparent[p];
  }
}

__device__ __forceinline__ bool is_pair(uint32_t c, const NodeInfo* info) {
  return !(c & LEAF_REF) && !info[c].collapsed;
}

// Pre-order pair index and first leaf id of every emitted internal node (one that is neither
// merged nor below a merged node), from the subtree counts along its path to the root.
__global__ void k_offsets(uint32_t n, const uint32_t* __restrict__ child, const uint32_t* __restrict__ iparent,
                          const NodeInfo* __restrict__ info, uint32_t* __restrict__ pair_of,
                          uint32_t* __restrict__ leaf_base, uint32_t* max_depth) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n - 1) return;
  if (info[x].collapsed) { pair_of[x] = NONE; return; }
  uint32_t P = 0, L = 0, depth = 1, y = x;
  bool emitted = true;
  while (y != 0u) {
    const uint32_t p = iparent[y];
    if (info[p].collapsed) { emitted = false; break; }
    const uint32_t c0 = child[2 * p];
    if (c0 == y) {
      P += 1u;
    } else {
      const bool pr = is_pair(c0, info);
      P += 1u + (pr ? info[c0].pairs : 0u);
      L += pr ? info[c0].leaves : 1u;
    }
    y = p;
    depth++;
  }
  pair_of[x] = emitted ? P : NONE;
  leaf_base[x] = L;
  if (emitted) atomicMax(max_depth, depth);
}

__global__ void k_emit(uint32_t n, const uint32_t* __restrict__ child, const uint32_t* __restrict__ iparent,
                       const uint32_t* __restrict__ first, const uint32_t* __restrict__ last,
                       const uint32_t* __restrict__ vals, const PBox* __restrict__ boxes,
                       const PBox* __restrict__ nbox, const NodeInfo* __restrict__ info,
                       const uint32_t* __restrict__ pair_of, const uint32_t* __restrict__ leaf_base, uint32_t triangles,
                       BVHNodePair* __restrict__ out, uint32_t* __restrict__ start_count) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n - 1) return;
  const uint32_t P = pair_of[x];
  if (P == NONE) return;
  // escape of the right slot: 2 P(p) + 1 of the nearest ancestor p entered through its left child
  uint32_t y = x;
  while (y != 0u && child[2 * iparent[y] + 1] == y) y = iparent[y];
  const uint32_t escape = (y == 0u) ? ESCAPE_ROOT : 2u * pair_of[iparent[y]] + 1u;
  uint32_t leaf_id = leaf_base[x];
  BVHNodePair pr;
  for (int k = 0; k < 2; ++k) {
    const uint32_t c = child[2 * x + k];
    BVHNode nd;
    child_box(c, vals, boxes, nbox, nd.boxMin, nd.boxMax);
    nd.escapeIndex = k == 0 ? 2u * P + 1u : escape;
    if (is_pair(c, info)) {
      nd.leftOffset = pair_of[c];
      leaf_id += info[c].leaves;
    } else {
      const uint32_t f = (c & LEAF_REF) ? (c & ~LEAF_REF) : first[c];
      const uint32_t count = (c & LEAF_REF) ? 1u : last[c] - first[c] + 1u;
      if (triangles) {
        nd.leftOffset = pack_offset_and_size(leaf_id, 1u);
        start_count[leaf_id] = pack_offset_and_size(f, count);
      } else {
        nd.leftOffset = pack_offset_and_size(vals[f], count);
      }
      leaf_id += 1u;
    }
    if (k == 0) pr.left = nd; else pr.right = nd;
  }
  out[P] = pr;
}

// TLAS refit: one thread per node; leaves take their instance's box, parents the union of their
// two children, computed by the second child to arrive.
__global__ void k_tlas_refit(BVHNode* tlas, uint32_t n_nodes, const uint32_t* __restrict__ parent,
                             const InstanceData* __restrict__ inst, uint32_t* cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const uint32_t lo = tlas[i].leftOffset;
  if (!(lo & LEAF_BIT)) return;
  const InstanceData& d = inst[lo & START_MASK];
  tlas[i].boxMin = xyz(d.boxMin);
  tlas[i].boxMax = xyz(d.boxMax);
  uint32_t p = parent[i];
  while (p != NONE) {
    const uint32_t old = __hip_atomic_fetch_add(cnt + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0u) return;
    const uint32_t c = tlas[p].leftOffset;
    tlas[p].boxMin = min3(tlas[c].boxMin, tlas[c + 1].boxMin);
    tlas[p].boxMax = max3(tlas[c].boxMax, tlas[c + 1].boxMax);
    p = parent[p];
  }
}

// Device allocation owned by one build (build-time memory, freed on return).
struct Tmp {
  void* p = nullptr;
  Tmp() = default;
  Tmp(const Tmp&) = delete;
  Tmp& operator=(const Tmp&) = delete;
  ~Tmp() { if (p) (void)hipFree(p); }
  template <class T> T* get() const { return static_cast<T*>(p); }
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("GPU BVH build: ") + what + ": " + hipGetErrorString(e));
}

void alloc(Tmp& t, size_t bytes, const char* what) { check(hipMalloc(&t.p, std::max<size_t>(bytes, 16)), what); }

inline dim3 grid_for(size_t n, int block) { return dim3(uint32_t((n + size_t(block) - 1) / size_t(block))); }

// The common part of both BLAS builds: boxes are already in d_boxes, centroid bounds in d_cb.
FatBlas build_from_boxes(uint32_t n, Tmp& d_boxes, Tmp& d_cb, bool triangles, int prims_in_leaf, BuildStats* stats) {
  constexpr int B = 256;
  Tmp keys, keys2, vals, vals2, child, iparent, lparent, first, last, nbox, info, cnt, pair_of, leaf_base, depth;
  alloc(keys, size_t(n) * 8, "keys"); alloc(keys2, size_t(n) * 8, "keys");
  alloc(vals, size_t(n) * 4, "values"); alloc(vals2, size_t(n) * 4, "values");
  hipLaunchKernelGGL(k_morton, grid_for(n, B), dim3(B), 0, nullptr, d_boxes.get<PBox>(), n, d_cb.get<uint32_t>(),
                     keys.get<uint64_t>(), vals.get<uint32_t>());
  check(hipGetLastError(), "morton codes");
  size_t temp_bytes = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.get<uint64_t>(), keys2.get<uint64_t>(),
                                           vals.get<uint32_t>(), vals2.get<uint32_t>(), int(n), 0, 63, nullptr),
        "sort size");
  {
    Tmp temp;
    alloc(temp, temp_bytes, "sort scratch");
    check(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, keys.get<uint64_t>(), keys2.get<uint64_t>(),
                                             vals.get<uint32_t>(), vals2.get<uint32_t>(), int(n), 0, 63, nullptr),
          "sort");
  }
  const uint32_t ni = n - 1;
  alloc(child, size_t(ni) * 8, "children"); alloc(iparent, size_t(ni) * 4, "parents");
  alloc(lparent, size_t(n) * 4, "parents"); alloc(first, size_t(ni) * 4, "ranges"); alloc(last, size_t(ni) * 4, "ranges");
  alloc(nbox, size_t(ni) * sizeof(PBox), "node boxes"); alloc(info, size_t(ni) * sizeof(NodeInfo), "node info");
  alloc(cnt, size_t(ni) * 4, "counters"); alloc(pair_of, size_t(ni) * 4, "pair index");
  alloc(leaf_base, size_t(ni) * 4, "leaf index"); alloc(depth, 4, "depth");
  hipLaunchKernelGGL(k_hierarchy, grid_for(ni, B), dim3(B), 0, nullptr, keys2.get<uint64_t>(), int(n),
                     child.get<uint32_t>(), iparent.get<uint32_t>(), lparent.get<uint32_t>(), first.get<uint32_t>(),
                     last.get<uint32_t>());
  hipLaunchKernelGGL(k_fill_u32, grid_for(ni, B), dim3(B), 0, nullptr, cnt.get<uint32_t>(), ni, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, nullptr, depth.get<uint32_t>(), 1u, 0u);
  hipLaunchKernelGGL(k_bottom_up, grid_for(n, B), dim3(B), 0, nullptr, n, vals2.get<uint32_t>(), d_boxes.get<PBox>(),
                     child.get<uint32_t>(), iparent.get<uint32_t>(), lparent.get<uint32_t>(), first.get<uint32_t>(),
                     last.get<uint32_t>(), uint32_t(prims_in_leaf), over_split_threshold(prims_in_leaf),
                     nbox.get<PBox>(), info.get<NodeInfo>(), cnt.get<uint32_t>());
  hipLaunchKernelGGL(k_offsets, grid_for(ni, B), dim3(B), 0, nullptr, n, child.get<uint32_t>(),
                     iparent.get<uint32_t>(), info.get<NodeInfo>(), pair_of.get<uint32_t>(),
                     leaf_base.get<uint32_t>(), depth.get<uint32_t>());
  check(hipGetLastError(), "hierarchy");
  NodeInfo root;
  check(hipMemcpy(&root, info.get<NodeInfo>(), sizeof(root), hipMemcpyDeviceToHost), "root counts");
  FatBlas out;
  Tmp d_out, d_sc;
  alloc(d_out, size_t(root.pairs) * sizeof(BVHNodePair), "pairs");
  alloc(d_sc, size_t(root.leaves) * 4, "leaf words");
  hipLaunchKernelGGL(k_emit, grid_for(ni, B), dim3(B), 0, nullptr, n, child.get<uint32_t>(), iparent.get<uint32_t>(),
                     first.get<uint32_t>(), last.get<uint32_t>(), vals2.get<uint32_t>(), d_boxes.get<PBox>(),
                     nbox.get<PBox>(), info.get<NodeInfo>(), pair_of.get<uint32_t>(), leaf_base.get<uint32_t>(),
                     triangles ? 1u : 0u, d_out.get<BVHNodePair>(), d_sc.get<uint32_t>());
  check(hipGetLastError(), "emit");
  out.nodes.resize(root.pairs);
  out.indices.resize(n);
  check(hipMemcpy(out.nodes.data(), d_out.p, out.nodes.size() * sizeof(BVHNodePair), hipMemcpyDeviceToHost), "pairs");
  check(hipMemcpy(out.indices.data(), vals2.p, size_t(n) * 4, hipMemcpyDeviceToHost), "indices");
  if (triangles) {
    out.startCount.resize(root.leaves);
    check(hipMemcpy(out.startCount.data(), d_sc.p, size_t(root.leaves) * 4, hipMemcpyDeviceToHost), "leaf words");
  }
  if (stats) {
    uint32_t dmax = 0;
    check(hipMemcpy(&dmax, depth.p, 4, hipMemcpyDeviceToHost), "depth");
    stats->max_depth = dmax;
    stats->leaves = root.leaves;
    stats->merged_leaves = 0;
    for (uint32_t w : out.startCount) stats->merged_leaves += extract_count(w) > 1u ? 1u : 0u;
  }
  return out;
}

}  // namespace

FatBlas gpu_build_blas_triangles(const float4* verts, size_t n_verts, const uint32_t* tri_idx, size_t n_idx,
                                 int prims_in_leaf, BuildStats* stats) {
  const size_t n_tris = n_idx / 3;
  if (n_tris <= size_t(prims_in_leaf)) return build_blas_triangles(verts, n_verts, tri_idx, n_idx, prims_in_leaf, stats);
  if (n_tris >= (size_t(1) << 28)) throw std::invalid_argument("GPU BVH build: too many triangles");
  const uint32_t n = uint32_t(n_tris);
  Tmp d_verts, d_idx, d_boxes, d_cb;
  alloc(d_verts, n_verts * sizeof(float4), "vertices");
  alloc(d_idx, n_idx * 4, "indices");
  alloc(d_boxes, size_t(n) * sizeof(PBox), "boxes");
  alloc(d_cb, 6 * 4, "bounds");
  check(hipMemcpy(d_verts.p, verts, n_verts * sizeof(float4), hipMemcpyHostToDevice), "upload vertices");
  check(hipMemcpy(d_idx.p, tri_idx, n_idx * 4, hipMemcpyHostToDevice), "upload indices");
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, nullptr, d_cb.get<uint32_t>());
  hipLaunchKernelGGL(k_tri_boxes, grid_for(n, 256), dim3(256), 0, nullptr, d_verts.get<float4>(),
                     d_idx.get<uint32_t>(), n, d_boxes.get<PBox>(), d_cb.get<uint32_t>());
  check(hipGetLastError(), "triangle boxes");
  return build_from_boxes(n, d_boxes, d_cb, true, prims_in_leaf, stats);
}

FatBlas gpu_build_blas_boxes(const BVHNode* boxes, size_t n_boxes, BuildStats* stats) {
  if (n_boxes <= 1) return build_blas_boxes(boxes, n_boxes, stats);
  if (n_boxes >= (size_t(1) << 28)) throw std::invalid_argument("GPU BVH build: too many boxes");
  const uint32_t n = uint32_t(n_boxes);
  Tmp d_in, d_boxes, d_cb;
  alloc(d_in, size_t(n) * sizeof(BVHNode), "input boxes");
  alloc(d_boxes, size_t(n) * sizeof(PBox), "boxes");
  alloc(d_cb, 6 * 4, "bounds");
  check(hipMemcpy(d_in.p, boxes, size_t(n) * sizeof(BVHNode), hipMemcpyHostToDevice), "upload boxes");
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, nullptr, d_cb.get<uint32_t>());
  hipLaunchKernelGGL(k_aabb_boxes, grid_for(n, 256), dim3(256), 0, nullptr, d_in.get<BVHNode>(), n,
                     d_boxes.get<PBox>(), d_cb.get<uint32_t>());
  check(hipGetLastError(), "boxes");
  return build_from_boxes(n, d_boxes, d_cb, false, 1, stats);
}

hipError_t launch_tlas_refit(BVHNode* tlas, uint32_t n_nodes, const uint32_t* parent, const InstanceData* inst,
                             uint32_t* counters, hipStream_t stream) {
  if (n_nodes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, grid_for(n_nodes, 256), dim3(256), 0, stream, counters, n_nodes, 0u);
  hipLaunchKernelGGL(k_tlas_refit, grid_for(n_nodes, 256), dim3(256), 0, stream, tlas, n_nodes, parent, inst,
                     counters);
  return hipGetLastError();
}

}  // namespace lrt
