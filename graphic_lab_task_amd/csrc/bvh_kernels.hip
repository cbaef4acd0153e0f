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
// one primitive per thread; the hierarchy's boxes are reduced bottom-up: inside each 1024-leaf
// chunk by the second of a node's two children to arrive (work-group-scope acquire/release
// counters), above the chunks by one work-group climbing the same way with LDS counters.
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

// Centroid bounds: every wave reduces its lanes, the block's waves meet in LDS, and one lane per
// block folds the result into six ordered words (min x, y, z as ordered-min, max x, y, z as
// ordered-max). The box kernels are grid-strided so that few blocks contend on the six words.
__device__ void fold_centroid(float3 lo, float3 hi, uint32_t* cb) {
  __shared__ float part[6][16];
  const float v[6] = {wave_min(lo.x), wave_min(lo.y), wave_min(lo.z), wave_max(hi.x), wave_max(hi.y), wave_max(hi.z)};
  const uint32_t w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
  if ((threadIdx.x & 63u) == 0u)
    for (int k = 0; k < 6; ++k) part[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    float r[6] = {part[0][0], part[1][0], part[2][0], part[3][0], part[4][0], part[5][0]};
    for (uint32_t i = 1; i < nw; ++i) {
      for (int k = 0; k < 3; ++k) r[k] = fmin_ref(r[k], part[k][i]);
      for (int k = 3; k < 6; ++k) r[k] = fmax_ref(r[k], part[k][i]);
    }
    for (int k = 0; k < 3; ++k) atomicMin(cb + k, f2ord(r[k]));
    for (int k = 3; k < 6; ++k) atomicMax(cb + k, f2ord(r[k]));
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

__global__ __launch_bounds__(256) void k_tri_boxes(const float4* __restrict__ verts, const uint32_t* __restrict__ idx,
                                                   uint32_t n, PBox* __restrict__ boxes, uint32_t* cb) {
  const float inf = __uint_as_float(0x7F800000u);
  float3 lo = make3(inf, inf, inf), hi = make3(-inf, -inf, -inf);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float3 a = xyz(verts[idx[3 * i + 0]]), b = xyz(verts[idx[3 * i + 1]]), d = xyz(verts[idx[3 * i + 2]]);
    PBox r;
    r.mn = min3(a, min3(b, d));
    r.mx = max3(a, max3(b, d));
    r.pad0 = r.pad1 = 0.0f;
    boxes[i] = r;
    const float3 c = (r.mn + r.mx) * 0.5f;
    lo = min3(lo, c);
    hi = max3(hi, c);
  }
  fold_centroid(lo, hi, cb);
}

__global__ __launch_bounds__(256) void k_aabb_boxes(const BVHNode* __restrict__ in, uint32_t n,
                                                    PBox* __restrict__ boxes, uint32_t* cb) {
  const float inf = __uint_as_float(0x7F800000u);
  float3 lo = make3(inf, inf, inf), hi = make3(-inf, -inf, -inf);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    PBox r;
    r.mn = in[i].boxMin;
    r.mx = in[i].boxMax;
    r.pad0 = r.pad1 = 0.0f;
    boxes[i] = r;
    const float3 c = (r.mn + r.mx) * 0.5f;
    lo = min3(lo, c);
    hi = max3(hi, c);
  }
  fold_centroid(lo, hi, cb);
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

// Nodes whose leaf range spans two CHUNK-leaf chunks ("upper" nodes) are appended to `upper`.
constexpr uint32_t CHUNK = 1024;

__device__ __forceinline__ bool is_upper(uint32_t f, uint32_t l) { return f / CHUNK != l / CHUNK; }

// The sorted leaf range [lo, hi] of internal node i (the direction and length search of the
// hierarchy construction below).
__device__ __forceinline__ void node_range(const uint64_t* __restrict__ keys, int n, int i, int& lo, int& hi, int& d,
                                           int& l) {
  d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  lo = i < j ? i : j;
  hi = i < j ? j : i;
}


// Karras's node-parallel hierarchy: internal node i covers the sorted range [first, last] and
// splits it where the common prefix drops. Children: internal index, or LEAF_REF | j.
__global__ void k_hierarchy(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ child,
                            uint32_t* __restrict__ iparent, uint32_t* __restrict__ lparent,
                            uint32_t* __restrict__ first, uint32_t* __restrict__ last, uint32_t* __restrict__ upper,
                            uint32_t* upper_count, uint32_t* __restrict__ slot_of) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  const bool valid = i < n - 1;
  int lo = 0, hi = 0, d = 1, l = 0;
  if (valid) node_range(keys, n, i, lo, hi, d, l);
  const bool up = valid && is_upper(uint32_t(lo), uint32_t(hi));
  const uint64_t m = __ballot(up);   // every lane of the wave reaches this
  if (m != 0ull) {
    const uint32_t lane = threadIdx.x & 63u, leader = uint32_t(__ffsll((unsigned long long)m) - 1);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(upper_count, uint32_t(__popcll(m)));
    base = __shfl(base, int(leader));
    const uint32_t slot = base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    if (up) {
      upper[slot] = uint32_t(i);
      slot_of[i] = slot;
    }
  }
  if (!valid) return;
  const int j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
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

// The box, merge decision and subtree counts of internal node p from its two finished children.
__device__ __forceinline__ void finish_node(uint32_t p, const uint32_t* __restrict__ vals, const PBox* boxes,
                                            const uint32_t* __restrict__ child, const uint32_t* __restrict__ first,
                                            const uint32_t* __restrict__ last, uint32_t prims_in_leaf,
                                            float over_split, PBox* nbox, NodeInfo* info) {
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
}

// Bottom-up reduction inside each CHUNK-leaf chunk, one block per chunk and one thread per leaf:
// the second child to reach a parent finishes it and climbs on (work-group-scope counters: a
// node inside the chunk is reached only by this block's threads); a thread stops below the first
// upper node. Every node's data is written by the thread that then releases its parent's counter.
__global__ __launch_bounds__(CHUNK) void k_bottom_up(uint32_t n, const uint32_t* __restrict__ vals,
                                                     const PBox* __restrict__ boxes,
                                                     const uint32_t* __restrict__ child,
                                                     const uint32_t* __restrict__ iparent,
                                                     const uint32_t* __restrict__ lparent,
                                                     const uint32_t* __restrict__ first,
                                                     const uint32_t* __restrict__ last, uint32_t prims_in_leaf,
                                                     float over_split, PBox* nbox, NodeInfo* info, uint32_t* cnt) {
  const uint32_t j = blockIdx.x * CHUNK + threadIdx.x;
  if (j >= n) return;
  uint32_t p = lparent[j];
  for (;;) {
    if (is_upper(first[p], last[p])) return;
    const uint32_t old = __hip_atomic_fetch_add(cnt + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0u) return;
    finish_node(p, vals, boxes, child, first, last, prims_in_leaf, over_split, nbox, info);
    if (p == 0u) return;
    p = iparent[p];
  }
}

// The upper nodes in one work-group, climbing as inside the chunks but with the arrival counters
// in LDS (slot of node p: slot_of[p]); a counter starts at the number of the node's children that
// are leaves or chunk nodes (finished by the chunk pass), so the nodes starting at 2 are the
// first to finish. The host runs this when the upper nodes fit UPPER_LDS counters.
constexpr uint32_t UPPER_LDS = 16384;
constexpr uint32_t UPPER_FRONT = UPPER_LDS / 1024;   // slots per thread

__device__ __forceinline__ bool finished_by_chunks(uint32_t c, const uint32_t* __restrict__ first,
                                                   const uint32_t* __restrict__ last) {
  return (c & LEAF_REF) || !is_upper(first[c], last[c]);
}

__global__ __launch_bounds__(1024) void k_upper_climb(const uint32_t* __restrict__ upper,
                                                      const uint32_t* __restrict__ upper_count,
                                                      const uint32_t* __restrict__ slot_of,
                                                      const uint32_t* __restrict__ vals, const PBox* boxes,
                                                      const uint32_t* __restrict__ child,
                                                      const uint32_t* __restrict__ iparent,
                                                      const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ last, uint32_t prims_in_leaf,
                                                      float over_split, PBox* nbox, NodeInfo* info) {
  __shared__ uint32_t arrivals[UPPER_LDS];
  const uint32_t count = min(*upper_count, UPPER_LDS);
  for (uint32_t k = threadIdx.x; k < count; k += 1024u) {
    const uint32_t p = upper[k];
    arrivals[k] = (finished_by_chunks(child[2 * p], first, last) ? 1u : 0u) +
                  (finished_by_chunks(child[2 * p + 1], first, last) ? 1u : 0u);
  }
  __syncthreads();
  uint32_t front[UPPER_FRONT];
  uint32_t nf = 0;
  for (uint32_t k = threadIdx.x; k < count; k += 1024u)
    if (arrivals[k] == 2u) front[nf++] = upper[k];
  __syncthreads();   // the snapshot above precedes every increment below
  for (uint32_t f = 0; f < nf; ++f) {
    uint32_t p = front[f];
    for (;;) {
      finish_node(p, vals, boxes, child, first, last, prims_in_leaf, over_split, nbox, info);
      if (p == 0u) break;
      p = iparent[p];
      const uint32_t old = __hip_atomic_fetch_add(arrivals + slot_of[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;
    }
  }
}

// The upper nodes (a few per chunk) in one work-group: every round finishes the listed nodes
// whose children are done (a leaf, a chunk node, or an upper node of an earlier round), then the
// round's nodes are marked done (counter 2) behind a barrier. Rounds stop when no node is left or
// after `count` rounds; the tree's height above the chunks bounds the rounds needed.
constexpr int UPPER_BLOCK = 1024;
constexpr int UPPER_PER_THREAD = 16;   // nodes one thread finishes in one round (more wait)

__global__ __launch_bounds__(UPPER_BLOCK) void k_upper(const uint32_t* __restrict__ upper,
                                                       const uint32_t* __restrict__ upper_count,
                                                       const uint32_t* __restrict__ vals, const PBox* boxes,
                                                       const uint32_t* __restrict__ child,
                                                       const uint32_t* __restrict__ first,
                                                       const uint32_t* __restrict__ last, uint32_t prims_in_leaf,
                                                       float over_split, PBox* nbox, NodeInfo* info, uint32_t* cnt) {
  __shared__ uint32_t left;
  const uint32_t count = *upper_count;
  for (uint32_t round = 0; round <= count; ++round) {
    if (threadIdx.x == 0) left = 0u;
    __syncthreads();
    uint32_t done[UPPER_PER_THREAD];
    int nd = 0;
    for (uint32_t k = threadIdx.x; k < count; k += UPPER_BLOCK) {
      const uint32_t p = upper[k];
      if (cnt[p] == 2u) continue;
      const uint32_t c0 = child[2 * p], c1 = child[2 * p + 1];
      const bool ready = ((c0 & LEAF_REF) || cnt[c0] == 2u) && ((c1 & LEAF_REF) || cnt[c1] == 2u);
      if (ready && nd < UPPER_PER_THREAD) {
        finish_node(p, vals, boxes, child, first, last, prims_in_leaf, over_split, nbox, info);
        done[nd++] = p;
      } else {
        left = 1u;
      }
    }
    __syncthreads();
    for (int k = 0; k < nd; ++k) cnt[done[k]] = 2u;
    __syncthreads();
    if (left == 0u) break;
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

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("GPU BVH build: ") + what + ": " + hipGetErrorString(e));
}

// All device memory of one build in one allocation, carved in 256-B aligned pieces (freed on
// return).
struct Arena {
  void* base = nullptr;
  size_t used = 0, cap = 0;
  Arena() = default;
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  ~Arena() { if (base) (void)hipFree(base); }
  static size_t up(size_t b) { return (b + 255) & ~size_t(255); }
  void reserve(size_t bytes) { check(hipMalloc(&base, std::max<size_t>(bytes, 256)), "build memory"); cap = bytes; }
  template <class T> T* take(size_t count) {
    const size_t b = up(count * sizeof(T));
    if (used + b > cap) throw std::runtime_error("GPU BVH build: arena sized too small");
    T* p = reinterpret_cast<T*>(static_cast<char*>(base) + used);
    used += b;
    return p;
  }
};

inline dim3 grid_for(size_t n, int block) { return dim3(uint32_t((n + size_t(block) - 1) / size_t(block))); }
inline dim3 box_grid(size_t n) { return dim3(uint32_t(std::min<size_t>((n + 255) / 256, 256))); }

size_t sort_scratch_bytes(uint32_t n) {
  size_t temp_bytes = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, static_cast<uint64_t*>(nullptr),
                                           static_cast<uint64_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), int(n), 0, 63, nullptr),
        "sort size");
  return temp_bytes;
}

// Bytes of the arena build_from_boxes carves (besides the caller's own inputs).
size_t core_bytes(uint32_t n) {
  const size_t ni = n - 1;
  return 2 * Arena::up(size_t(n) * 8) + 2 * Arena::up(size_t(n) * 4) + Arena::up(sort_scratch_bytes(n)) +
         Arena::up(ni * 8) + 4 * Arena::up(ni * 4) + Arena::up(size_t(n) * 4) + Arena::up(ni * sizeof(PBox)) +
         Arena::up(ni * sizeof(NodeInfo)) + Arena::up(ni * 4) + Arena::up(4) + Arena::up(ni * sizeof(BVHNodePair)) +
         2 * Arena::up(ni * 4) + Arena::up(4) +
         Arena::up(size_t(n) * 4) + Arena::up(size_t(n) * sizeof(PBox)) + Arena::up(6 * 4);
}

// The common part of both BLAS builds: boxes are already in `boxes`, centroid bounds in `cb`.
FatBlas build_from_boxes(uint32_t n, Arena& A, PBox* boxes, uint32_t* cb, bool triangles, int prims_in_leaf,
                         BuildStats* stats) {
  constexpr int B = 256;
  const uint32_t ni = n - 1;
  uint64_t* keys = A.take<uint64_t>(n);
  uint64_t* keys2 = A.take<uint64_t>(n);
  uint32_t* vals = A.take<uint32_t>(n);
  uint32_t* vals2 = A.take<uint32_t>(n);
  size_t temp_bytes = sort_scratch_bytes(n);
  void* temp = A.take<char>(temp_bytes);
  uint32_t* child = A.take<uint32_t>(size_t(ni) * 2);
  uint32_t* iparent = A.take<uint32_t>(ni);
  uint32_t* first = A.take<uint32_t>(ni);
  uint32_t* last = A.take<uint32_t>(ni);
  uint32_t* pair_of = A.take<uint32_t>(ni);
  uint32_t* lparent = A.take<uint32_t>(n);
  PBox* nbox = A.take<PBox>(ni);
  NodeInfo* info = A.take<NodeInfo>(ni);
  uint32_t* cnt = A.take<uint32_t>(ni);
  uint32_t* depth = A.take<uint32_t>(1);
  BVHNodePair* d_out = A.take<BVHNodePair>(ni);   // at most n - 1 pairs
  uint32_t* d_sc = A.take<uint32_t>(n);            // at most n leaves
  uint32_t* upper = A.take<uint32_t>(ni);
  uint32_t* upper_count = A.take<uint32_t>(1);
  uint32_t* slot_of = A.take<uint32_t>(ni);
  uint32_t* leaf_base = cnt;                       // the counters are done with after the climb
  hipLaunchKernelGGL(k_morton, grid_for(n, B), dim3(B), 0, nullptr, boxes, n, cb, keys, vals);
  check(hipGetLastError(), "morton codes");
  check(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys2, vals, vals2, int(n), 0, 63, nullptr), "sort");
  hipLaunchKernelGGL(k_fill_u32, grid_for(ni, B), dim3(B), 0, nullptr, cnt, ni, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, nullptr, depth, 1u, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, nullptr, upper_count, 1u, 0u);
  hipLaunchKernelGGL(k_hierarchy, grid_for(ni, B), dim3(B), 0, nullptr, keys2, int(n), child, iparent, lparent, first,
                     last, upper, upper_count, slot_of);
  const float os = over_split_threshold(prims_in_leaf);
  hipLaunchKernelGGL(k_bottom_up, grid_for(n, CHUNK), dim3(CHUNK), 0, nullptr, n, vals2, boxes, child, iparent,
                     lparent, first, last, uint32_t(prims_in_leaf), os, nbox, info, cnt);
  uint32_t n_upper = 0;
  check(hipMemcpy(&n_upper, upper_count, 4, hipMemcpyDeviceToHost), "upper node count");
  if (n_upper <= UPPER_LDS)
    hipLaunchKernelGGL(k_upper_climb, dim3(1), dim3(1024), 0, nullptr, upper, upper_count, slot_of, vals2, boxes, child,
                       iparent, first, last, uint32_t(prims_in_leaf), os, nbox, info);
  else
    hipLaunchKernelGGL(k_upper, dim3(1), dim3(UPPER_BLOCK), 0, nullptr, upper, upper_count, vals2, boxes, child, first,
                       last, uint32_t(prims_in_leaf), os, nbox, info, cnt);
  hipLaunchKernelGGL(k_offsets, grid_for(ni, B), dim3(B), 0, nullptr, n, child, iparent, info, pair_of, leaf_base,
                     depth);
  hipLaunchKernelGGL(k_emit, grid_for(ni, B), dim3(B), 0, nullptr, n, child, iparent, first, last, vals2, boxes, nbox,
                     info, pair_of, leaf_base, triangles ? 1u : 0u, d_out, d_sc);
  check(hipGetLastError(), "hierarchy");
  NodeInfo root;
  check(hipMemcpy(&root, info, sizeof(root), hipMemcpyDeviceToHost), "root counts");
  if (root.pairs > ni || root.leaves > n) throw std::runtime_error("GPU BVH build: inconsistent subtree counts");
  FatBlas out;
  out.nodes.resize(root.pairs);
  out.indices.resize(n);
  check(hipMemcpy(out.nodes.data(), d_out, out.nodes.size() * sizeof(BVHNodePair), hipMemcpyDeviceToHost), "pairs");
  check(hipMemcpy(out.indices.data(), vals2, size_t(n) * 4, hipMemcpyDeviceToHost), "indices");
  if (triangles) {
    out.startCount.resize(root.leaves);
    check(hipMemcpy(out.startCount.data(), d_sc, size_t(root.leaves) * 4, hipMemcpyDeviceToHost), "leaf words");
  }
  if (stats) {
    uint32_t dmax = 0;
    check(hipMemcpy(&dmax, depth, 4, hipMemcpyDeviceToHost), "depth");
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
  Arena A;
  A.reserve(core_bytes(n) + Arena::up(n_verts * sizeof(float4)) + Arena::up(n_idx * 4));
  float4* d_verts = A.take<float4>(n_verts);
  uint32_t* d_idx = A.take<uint32_t>(n_idx);
  PBox* boxes = A.take<PBox>(n);
  uint32_t* cb = A.take<uint32_t>(6);
  check(hipMemcpy(d_verts, verts, n_verts * sizeof(float4), hipMemcpyHostToDevice), "upload vertices");
  check(hipMemcpy(d_idx, tri_idx, n_idx * 4, hipMemcpyHostToDevice), "upload indices");
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, nullptr, cb);
  hipLaunchKernelGGL(k_tri_boxes, box_grid(n), dim3(256), 0, nullptr, d_verts, d_idx, n, boxes, cb);
  check(hipGetLastError(), "triangle boxes");
  return build_from_boxes(n, A, boxes, cb, true, prims_in_leaf, stats);
}

FatBlas gpu_build_blas_boxes(const BVHNode* in, size_t n_boxes, BuildStats* stats) {
  if (n_boxes <= 1) return build_blas_boxes(in, n_boxes, stats);
  if (n_boxes >= (size_t(1) << 28)) throw std::invalid_argument("GPU BVH build: too many boxes");
  const uint32_t n = uint32_t(n_boxes);
  Arena A;
  A.reserve(core_bytes(n) + Arena::up(size_t(n) * sizeof(BVHNode)));
  BVHNode* d_in = A.take<BVHNode>(n);
  PBox* boxes = A.take<PBox>(n);
  uint32_t* cb = A.take<uint32_t>(6);
  check(hipMemcpy(d_in, in, size_t(n) * sizeof(BVHNode), hipMemcpyHostToDevice), "upload boxes");
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, nullptr, cb);
  hipLaunchKernelGGL(k_aabb_boxes, box_grid(n), dim3(256), 0, nullptr, d_in, n, boxes, cb);
  check(hipGetLastError(), "boxes");
  return build_from_boxes(n, A, boxes, cb, false, 1, stats);
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
