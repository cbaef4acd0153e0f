// Wave-cooperative BLAS traversal: one ray per wave64, the wave expands up to 64 frontier
// nodes per iteration. It returns exactly the hit of the sequential walk
// (BVHRT::BVH2TraverseF32, MultiRenderer/BVH2Common.cpp:3239-3320) by ranking candidates
// with the sequential visiting order:
//  * the sequential walk visits leaves in depth-first order, nearer child first
//    (tm0.x <= tm1.x); each node gets a 64-bit key = its path of near(0)/far(1) choices,
//    left-aligned, so comparing keys compares visiting order;
//  * triangle leaves: the sequential result is the smallest t, ties going to the earlier
//    visit (strict t < hit.t, BVH2Common.cpp:1600) - here min over (t, key, index in leaf);
//    with stop-on-first-hit it is the first leaf in order holding an accepted triangle;
//  * SDF leaves end the walk at the first leaf that hits (BVH2Common.cpp:3234-3237, 3310) -
//    here min over key of the leaves that hit.
// Nodes are culled only when they cannot hold a better candidate, so the set of candidates
// that can win is the same as in the sequential walk.
#pragma once
#include "trace_device.h"

namespace lrt {

constexpr int COOP_CAP = 1024;   // frontier entries per wave (LDS)

struct CoopEntry {
  uint32_t pair;
  uint32_t depth;
  uint64_t key;
};
// Entries live in LDS as 4 words: pair, depth, key low, key high.
typedef lds_u32 lds_entry;

__device__ __forceinline__ void put_entry(lds_entry* st, int i, uint32_t pair, uint32_t depth, uint64_t key) {
  st[4 * i + 0] = pair;
  st[4 * i + 1] = depth;
  st[4 * i + 2] = uint32_t(key);
  st[4 * i + 3] = uint32_t(key >> 32);
}

__device__ __forceinline__ CoopEntry get_entry(const lds_entry* st, int i) {
  CoopEntry e;
  e.pair = st[4 * i + 0];
  e.depth = st[4 * i + 1];
  e.key = uint64_t(st[4 * i + 2]) | (uint64_t(st[4 * i + 3]) << 32);
  return e;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}

__device__ __forceinline__ CRT_Hit shfl_hit(const CRT_Hit& h, int src) {
  CRT_Hit r;
  r.t = __shfl(h.t, src, 64);
  r.primId = __shfl(h.primId, src, 64);
  r.instId = __shfl(h.instId, src, 64);
  r.geomId = __shfl(h.geomId, src, 64);
  r.coords[0] = __shfl(h.coords[0], src, 64);
  r.coords[1] = __shfl(h.coords[1], src, 64);
  r.coords[2] = __shfl(h.coords[2], src, 64);
  r.coords[3] = __shfl(h.coords[3], src, 64);
  return r;
}

// One lane's candidate from the leaves it processed this iteration.
struct CoopCand {
  bool valid;
  float t;
  uint64_t key;
  CRT_Hit hit;
};

// Processes one leaf for this lane against the incoming hit h_in; keeps the lane's best
// candidate by (t, key) for triangles or by key for SDF/stop-on-first-hit.
template <uint32_t KINDS>
__device__ __forceinline__ void coop_leaf(const SceneView& s, float3 ray_pos, float3 ray_dir, float tNear,
                                          uint32_t instId, uint32_t geomId, const GeomRef& g, uint32_t tag,
                                          uint32_t leafWord,
                                          uint64_t key, bool by_key, const CRT_Hit& h_in, CoopCand& c,
                                          uint32_t& leaves, uint32_t& prims) {
  const float tNearSdf = fmax_ref(tNear, 0.01f);
  const uint32_t a_start = extract_start(leafWord);
  const uint32_t a_count = extract_count(leafWord);
  CRT_Hit h = h_in;
  leaves++;
  if ((KINDS & LK_TRI) && tag == TAG_TRIANGLE) {
    prims += a_count;
    intersect_tris(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, a_count, h);
  } else if ((KINDS & LK_FRAME) && tag == TAG_SDF_NODE) {
    frame_node_intersect<(KINDS & LK_FRAME_TEX) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, h);
  } else if ((KINDS & LK_SBS) && tag == TAG_SDF_BRICK) {
    sbs_brick_intersect<(KINDS & LK_SBS_SMOOTH) != 0, (KINDS & LK_TRICUBIC) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, h);
  } else if ((KINDS & LK_MULTI) && tag == TAG_SDF_OCTREE) {
    multi_octree_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, h);
  } else if ((KINDS & LK_GRID) && tag == TAG_SDF_GRID) {
    sdf_grid_intersect<(KINDS & LK_TRICUBIC) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, h);
  } else if ((KINDS & LK_GPRIM) && tag == TAG_GRAPHICS_PRIM) {
    gprim_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, h);
  } else if ((KINDS & LK_COCT) && tag == TAG_COCTREE_BRICKED) {
    coct_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, h);
  } else if ((KINDS & LK_DAG) && tag == TAG_SDF_DAG) {
    dag_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, h);
  } else if ((KINDS & LK_SCOM2) && tag == TAG_SCOM2) {
    if (s.scom2_fast) scom2_intersect_fast(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, h);
    else scom2_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, h);
  }
  if (s.preset.render_mode == MULTI_RENDER_MODE_INTERSECTION_COUNT)
    h.instId = (h.instId == 0xFFFFFFFFu) ? 1 : h.instId + 1;
  if (!(h.t < h_in.t)) return;
  const bool better = !c.valid || (by_key ? key < c.key : (h.t < c.t || (h.t == c.t && key < c.key)));
  if (better) {
    c.valid = true;
    c.t = h.t;
    c.key = key;
    c.hit = h;
  }
}

// Cooperative BVH2TraverseF32 for one BLAS. Every lane holds the same ray; on return every
// lane holds the same hit. Returns false if the frontier outgrew
// COOP_CAP or the tree is deeper than 63 levels (the caller then walks sequentially).
template <uint32_t KINDS>
__device__ bool coop_blas(const SceneView& s, lds_entry* st, float3 ray_pos, float3 ray_dir, float tNear,
                          uint32_t instId, uint32_t geomId, bool stopOnFirstHit, CRT_Hit& hit, RayCounts& cnt) {
  const uint32_t lane = __lane_id();
  const uint32_t bvhOffset = s.geoms[geomId].bvhOffset;
  const GeomRef g = geom_ref(s, geomId);
  const uint32_t tag = s.geom_tags[geomId];
  const bool first_hit = stopOnFirstHit || tag != TAG_TRIANGLE;   // first_hit_is_closest / stop-on-first
  const float3 inv = safe_inverse(ray_dir);
  const CRT_Hit h_in = hit;
  bool have = false;
  float best_t = h_in.t;
  uint64_t best_key = ~0ull;
  CRT_Hit best_hit = h_in;
  bool ok = true;
  uint32_t leaves = 0, prims = 0, nodes = 0;

  int top = 0;
  if (lane == 0) put_entry(st, 0, 0u, 0u, 0ull);
  top = 1;
  __syncthreads();
  while (top > 0) {
    const int take = top < 64 ? top : 64;
    const int base = top - take;
    CoopEntry e{0u, 0u, 0ull};
    const bool active = int(lane) < take;
    if (active) e = get_entry(st, base + int(lane));
    __syncthreads();
    top = base;
    // cull entries that cannot hold a winner any more
    bool live = active;
    if (live && have && first_hit && e.key > best_key) live = false;
    bool push0 = false, push1 = false, leaf0 = false, leaf1 = false;
    uint32_t c0 = 0, c1 = 0;
    uint64_t k0 = 0, k1 = 0;
    if (live) {
      if (e.depth >= 63) {
        ok = false;
      } else {
        nodes++;
        const BVHNodePair fat = load_pair(s.nodes + bvhOffset + e.pair);
        const float2 tm0 = ray_box2(ray_pos, inv, fat.left.boxMin, fat.left.boxMax);
        const float2 tm1 = ray_box2(ray_pos, inv, fat.right.boxMin, fat.right.boxMax);
        const float cull = first_hit ? h_in.t : best_t;
        const bool hit0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= cull);
        const bool hit1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= cull);
        const bool near0 = tm0.x <= tm1.x;
        const uint64_t bit = 1ull << (63 - e.depth);
        c0 = fat.left.leftOffset;
        c1 = fat.right.leftOffset;
        k0 = e.key | (near0 ? 0ull : bit);
        k1 = e.key | (near0 ? bit : 0ull);
        if (hit0 && c0 != 0xFFFFFFFFu) { if (c0 & LEAF_BIT) leaf0 = true; else push0 = true; }
        if (hit1 && c1 != 0xFFFFFFFFu) { if (c1 & LEAF_BIT) leaf1 = true; else push1 = true; }
      }
    }
    // push inner children (ballot compaction)
    const uint64_t m0 = __ballot(push0), m1 = __ballot(push1);
    const int n0 = __popcll(m0), n1 = __popcll(m1);
    const uint64_t lt = lanemask_lt();
    if (top + n0 + n1 > COOP_CAP) ok = false;
    if (ok) {
      if (push0) put_entry(st, top + __popcll(m0 & lt), c0, e.depth + 1, k0);
      if (push1) put_entry(st, top + n0 + __popcll(m1 & lt), c1, e.depth + 1, k1);
      top += n0 + n1;
    }
    // leaves: each lane processes its own, then the wave votes for the winner
    CoopCand c;
    c.valid = false;
    c.t = 0.0f;
    c.key = ~0ull;
    if (leaf0) coop_leaf<KINDS>(s, ray_pos, ray_dir, tNear, instId, geomId, g, tag, c0, k0, first_hit, h_in, c, leaves, prims);
    if (leaf1) coop_leaf<KINDS>(s, ray_pos, ray_dir, tNear, instId, geomId, g, tag, c1, k1, first_hit, h_in, c, leaves, prims);
    if (__ballot(c.valid)) {
      int winner;
      if (first_hit) {
        const uint64_t kmin = wave_min_u64(c.valid ? c.key : ~0ull);
        winner = __ffsll((long long)__ballot(c.valid && c.key == kmin)) - 1;
      } else {
        const float tmin = wave_min_f(c.valid ? c.t : __int_as_float(0x7f800000));
        const bool tie = c.valid && c.t == tmin;
        const uint64_t kmin = wave_min_u64(tie ? c.key : ~0ull);
        winner = __ffsll((long long)__ballot(tie && c.key == kmin)) - 1;
      }
      const uint64_t wkey = __shfl(c.key, winner, 64);
      const float wt = __shfl(c.t, winner, 64);
      const bool improves = !have || (first_hit ? wkey < best_key : (wt < best_t || (wt == best_t && wkey < best_key)));
      if (improves) {
        best_hit = shfl_hit(c.hit, winner);
        best_t = wt;
        best_key = wkey;
        have = true;
      }
    }
    if (!__ballot(ok) || __ballot(!ok)) {   // any lane failed: give up cooperatively
      ok = false;
      break;
    }
    __syncthreads();
  }
  cnt.blas_nodes += nodes;
  cnt.leaves += leaves;
  cnt.prims += prims;
  if (!ok) return false;
  if (have) hit = best_hit;
  return true;
}

}  // namespace lrt
