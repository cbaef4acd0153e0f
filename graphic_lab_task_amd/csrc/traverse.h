// Ray traversal device code shared by the trace and differentiable kernels: the reference's
// sequential BLAS walk, its TLAS walk and the hit finishing, plus small launch helpers.
#pragma once
#include "trace_coop.h"
#include "trace_device.h"

namespace lrt {


// BVHRT::BVH2TraverseF32 (MultiRenderer/BVH2Common.cpp:3239-3320): stack-based walk over
// fat nodes, nearer child first, far child pushed; leaves dispatched on the geometry tag.
// `steps` counts fat-node fetches and leaf visits; past `budget` the walk stops and sets
// `aborted` (budget 0 = unlimited).
template <uint32_t KINDS, int S, int BLOCK, bool COUNT, bool FAST = false>
__device__ __forceinline__ void blas_traverse(const SceneView& s, float3 ray_pos, float3 ray_dir, float tNear,
                                              uint32_t instId, uint32_t geomId, bool stopOnFirstHit, CRT_Hit& hit,
                                              LaneStack<S, BLOCK>& st, uint32_t& overflow,
                                              RayCounts& cnt, const lds_pair* top_pairs, uint32_t budget,
                                              uint32_t& steps, bool& aborted, lds_u32* ot_col = nullptr,
                                              SbsVoxelRec* rec = nullptr) {
  const uint32_t bvhOffset = s.geoms[geomId].bvhOffset;
  const GeomRef g = geom_ref(s, geomId);
  const BVHNodePair* nodes = s.nodes + bvhOffset;
  const uint32_t cached = (top_pairs != nullptr && geomId == s.cache_geom) ? s.cache_pairs : 0u;
  int top = 0;
  uint32_t leftNodeOffset = 0;
  bool hitFound = false;
  const float3 rayDirInv = safe_inverse(ray_dir);
  const uint32_t tag = s.geom_tags[geomId];
  // BVHRT::IntersectAllPrimitivesInLeaf (BVH2Common.cpp:47-124); device-layout leaves carry
  // PackOffsetAndSize(start, count) themselves. Returns false when the step budget ran out.
  auto leaf = [&](uint32_t word) -> bool {
    if (budget && ++steps > budget) return false;
    const uint32_t a_start = extract_start(word);
    const uint32_t a_count = extract_count(word);
    const float tNearSdf = fmax_ref(tNear, 0.01f);
    const float tPrev = hit.t;
    if (COUNT) {
      cnt.leaves++;
      if (tag == TAG_TRIANGLE) cnt.prims += a_count;
    }
    if ((KINDS & LK_TRI) && tag == TAG_TRIANGLE) {
      intersect_tris(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, a_count, hit);
    } else if ((KINDS & LK_FRAME) && tag == TAG_SDF_NODE) {
      frame_node_intersect<(KINDS & LK_FRAME_TEX) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, hit);
    } else if ((KINDS & LK_SBS) && tag == TAG_SDF_BRICK) {
      cnt.sdf_voxels += sbs_brick_intersect<(KINDS & LK_SBS_SMOOTH) != 0, (KINDS & LK_TRICUBIC) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId,
                                                                          geomId, a_start, hit, rec);
    } else if ((KINDS & LK_MULTI) && tag == TAG_SDF_OCTREE) {
      cnt.sdf_voxels += multi_octree_intersect<BLOCK>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, hit, ot_col);
    } else if ((KINDS & LK_GRID) && tag == TAG_SDF_GRID) {
      sdf_grid_intersect<(KINDS & LK_TRICUBIC) != 0>(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, hit);
    } else if ((KINDS & LK_GPRIM) && tag == TAG_GRAPHICS_PRIM) {
      gprim_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, hit);
    } else if ((KINDS & LK_COCT) && tag == TAG_COCTREE_BRICKED) {
      coct_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, hit);
    } else if ((KINDS & LK_DAG) && tag == TAG_SDF_DAG) {
      dag_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, hit);
    } else if ((KINDS & LK_SCOM2) && tag == TAG_SCOM2) {
      if (s.scom2_fast) scom2_intersect_fast(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, hit);
      else scom2_intersect(s, g, ray_pos, ray_dir, tNearSdf, instId, geomId, hit);
    }
    if (s.preset.render_mode == MULTI_RENDER_MODE_INTERSECTION_COUNT)
      hit.instId = (hit.instId == 0xFFFFFFFFu) ? 1 : hit.instId + 1;
    const uint32_t hitTag = (hit.t >= tPrev) ? TAG_NONE : tag;
    hitFound = (hitTag != TAG_NONE) && (hitTag != TAG_TRIANGLE || stopOnFirstHit);
    return true;
  };
  // SDF walks end at the first leaf that hits, and hit.t does not change before that. So a lane
  // that reaches a leaf notes it and keeps walking until every lane of the wave holds a noted
  // leaf; the noted leaves are then intersected together, each lane's in its walk order: the
  // same leaves in the same order as the sequential walk, with fewer wave-wide waits
  // (postponed leaves, Aila & Laine 2009). Triangle walks keep the plain order unless the
  // kernel is built with LK_TRI_POSTPONE: a closest-hit triangle walk then culls the boxes met
  // while a leaf waits against the older, larger hit.t, so it may visit extra leaves but never
  // skips one; leaves keep their walk order, and an extra leaf's triangles lie at or beyond its
  // box entry, past the final hit, so the result only differs where a triangle's t rounds below
  // its own box entry by a few ulp and ties the final hit that closely.
  const bool postpone = !COUNT && (((KINDS & LK_TRI_POSTPONE) != 0) ||
                                   ((KINDS & (LK_FRAME | LK_SBS | LK_MULTI | LK_GRID | LK_GPRIM | LK_COCT | LK_DAG | LK_SCOM2)) != 0 &&
                                    tag != TAG_TRIANGLE));
  uint32_t postponed = 0;   // the noted leaf word (leaf words always have LEAF_BIT set)
  constexpr bool LEAF_HALF = (KINDS & (LK_SBS | LK_MULTI | LK_GRID | LK_COCT | LK_DAG | LK_SCOM2)) == 0;
  auto note_leaf = [&]() {
    if (leftNodeOffset != 0xFFFFFFFFu) postponed = leftNodeOffset;
    top--;
    leftNodeOffset = st.get(top > 0 ? top : 0);
  };
  while (!hitFound && (top >= 0 || postponed != 0)) {
    if (postpone && postponed == 0 && top >= 0 && (leftNodeOffset & LEAF_BIT)) note_leaf();
    while (top >= 0 && ((leftNodeOffset & LEAF_BIT) == 0)) {
      if (budget && ++steps > budget) { aborted = true; return; }
      if (COUNT) cnt.blas_nodes++;
      const BVHNodePair fat = leftNodeOffset < cached ? load_pair_lds(top_pairs + leftNodeOffset)
                                                      : load_pair(nodes + leftNodeOffset);
      const uint32_t n0 = fat.left.leftOffset;
      const uint32_t n1 = fat.right.leftOffset;
      const float2 tm0 = FAST ? ray_box2_fast(ray_pos, rayDirInv, fat.left.boxMin, fat.left.boxMax)
                              : ray_box2(ray_pos, rayDirInv, fat.left.boxMin, fat.left.boxMax);
      const float2 tm1 = FAST ? ray_box2_fast(ray_pos, rayDirInv, fat.right.boxMin, fat.right.boxMax)
                              : ray_box2(ray_pos, rayDirInv, fat.right.boxMin, fat.right.boxMax);
      const bool hit0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= hit.t);
      const bool hit1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= hit.t);
      leftNodeOffset = hit0 ? n0 : n1;
      if (hit0 && hit1) {
        const bool first0 = tm0.x <= tm1.x;
        leftNodeOffset = first0 ? n0 : n1;
        if (top < STACK_SIZE) {
          st.put(top, first0 ? n1 : n0);
          top++;
        } else {
          overflow = 1;
        }
      }
      if (!hit0 && !hit1) {
        top--;
        leftNodeOffset = st.get(top > 0 ? top : 0);
      }
      if (postpone) {
        if (postponed == 0 && top >= 0 && (leftNodeOffset & LEAF_BIT)) note_leaf();
        // Frame-octree and triangle leaves are cheap: the noted leaves are taken once half of
        // the wave's walking lanes hold one (C2 -3%). Brick-set and multi-octree leaves are whole
        // marches and wait for every lane (taking them at half costs C3 +6%).
        if (LEAF_HALF ? __popcll(__ballot(postponed == 0)) * 2 <= __popcll(__ballot(true)) : !__any(postponed == 0))
          break;
      }
    }
    if (postpone) {
      if (postponed != 0) {
        const uint32_t w = postponed;
        postponed = 0;
        if (!leaf(w)) { aborted = true; return; }
      } else if (top >= 0 && (leftNodeOffset & LEAF_BIT)) {
        if (leftNodeOffset != 0xFFFFFFFFu && !leaf(leftNodeOffset)) { aborted = true; return; }
        top--;
        leftNodeOffset = st.get(top > 0 ? top : 0);
      }
    } else {
      if (top >= 0 && leftNodeOffset != 0xFFFFFFFFu && !leaf(leftNodeOffset)) { aborted = true; return; }
      top--;
      leftNodeOffset = st.get(top > 0 ? top : 0);
    }
  }
}

__device__ __forceinline__ CRT_Hit miss_hit(float far) {
  CRT_Hit hit;
  hit.t = far;
  hit.primId = 0xFFFFFFFFu;
  hit.instId = 0xFFFFFFFFu;
  hit.geomId = 0xFFFFFFFFu;
  hit.coords[0] = 1.0f;
  hit.coords[1] = 0.0f;
  hit.coords[2] = 0.0f;
  hit.coords[3] = 0.0f;
  return hit;
}

// Normal of a final triangle hit and the m_primIndices remap (BVH2Common.cpp:3382-3388).
template <uint32_t KINDS>
__device__ __forceinline__ void finish_hit(const SceneView& s, CRT_Hit& hit) {
  if ((KINDS & LK_TRI) && hit.geomId != 0xFFFFFFFFu && (hit.geomId >> SH_TYPE) == TYPE_MESH_TRIANGLE)
    resolve_tri_normal(s, hit);
  if (hit.geomId < 0xFFFFFFFFu && ((hit.geomId >> SH_TYPE) == TYPE_MESH_TRIANGLE)) {
    const uint2 go = s.geoms[hit.geomId & GEOM_ID_MASK].offset;
    hit.primId = s.prim_indices[go.x / 3 + hit.primId];
  }
}

// BVHRT::RayQuery_NearestHit (MultiRenderer/BVH2Common.cpp:3322-3392): stackless TLAS walk
// over escape indices, ray moved to instance space (direction not renormalised).
// COOP = true: the wave holds one ray and walks each BLAS cooperatively. FAST = true: box tests
// with hardware min/max (the caller checked box_test_fast for the world ray; each instance-space
// ray is checked again before its BLAS walk). `keep` (LDS, 6 words per lane at keep[k * BLOCK],
// may be null): the world ray is parked there and read back after each BLAS walk, so that its
// registers are free while the BLAS is walked (the inverse direction is recomputed, bit for bit);
// kernels that walk multi-octrees extend it by 2 * OT_LDS words per lane for the walk's stack.
// `rec` (brick-set kernels, may be null): the brick march records the voxel of every surface hit
// it makes; an SDF walk ends at its first hitting leaf and a later hit is always closer, so after
// the walk it holds the voxel of the final hit when that hit is a brick-set hit.
template <uint32_t KINDS, int S, int BLOCK, bool COUNT, bool COOP, bool FAST = false>
__device__ __forceinline__ CRT_Hit ray_query_nearest(const SceneView& s, float4 posAndNear, float4 dirAndFar,
                                                     LaneStack<S, BLOCK>& st, uint32_t& overflow, RayCounts& cnt,
                                                     const lds_pair* top_pairs, uint32_t budget, bool& aborted,
                                                     lds_entry* coop_stack, lds_u32* keep = nullptr,
                                                     SbsVoxelRec* rec = nullptr) {
  bool stopOnFirstHit = (dirAndFar.w <= 0.0f);
  if (stopOnFirstHit) dirAndFar.w *= -1.0f;
  if (keep) {
    keep[0 * BLOCK] = __float_as_uint(posAndNear.x);
    keep[1 * BLOCK] = __float_as_uint(posAndNear.y);
    keep[2 * BLOCK] = __float_as_uint(posAndNear.z);
    keep[3 * BLOCK] = __float_as_uint(dirAndFar.x);
    keep[4 * BLOCK] = __float_as_uint(dirAndFar.y);
    keep[5 * BLOCK] = __float_as_uint(dirAndFar.z);
  }
  CRT_Hit hit = miss_hit(dirAndFar.w);
  uint32_t steps = 0;
  uint32_t nodeIdx = 0;
  do {
    // compiler-only fence (emits nothing): the world ray is re-read below instead of being held
    // in registers across the BLAS walk
    if (keep) asm volatile("" ::: "memory");
    const float3 pos = keep ? make3(__uint_as_float(keep[0 * BLOCK]), __uint_as_float(keep[1 * BLOCK]),
                                    __uint_as_float(keep[2 * BLOCK]))
                            : xyz(posAndNear);
    const float3 dir = keep ? make3(__uint_as_float(keep[3 * BLOCK]), __uint_as_float(keep[4 * BLOCK]),
                                    __uint_as_float(keep[5 * BLOCK]))
                            : xyz(dirAndFar);
    const float3 rayDirInv = safe_inverse(dir);
    uint32_t travFlags = 0;
    uint32_t leftOffset = 0;
    do {
      if (COUNT) cnt.tlas_nodes++;
      const BVHNode cur = load_node(s.tlas + nodeIdx);
      const float2 bh = FAST ? ray_box2_fast(pos, rayDirInv, cur.boxMin, cur.boxMax)
                             : ray_box2(pos, rayDirInv, cur.boxMin, cur.boxMax);
      const bool intersects = (bh.x <= bh.y) && (bh.y > posAndNear.w) && (bh.x < hit.t);
      travFlags = (cur.leftOffset & LEAF_BIT) | uint32_t(intersects);
      leftOffset = cur.leftOffset;
      const bool leaf_or_miss = (travFlags & LEAF_BIT) != 0 || (travFlags & 0x1) == 0;
      nodeIdx = leaf_or_miss ? cur.escapeIndex : leftOffset;
    } while (travFlags != (LEAF_BIT | 0x1) && nodeIdx != 0 && nodeIdx < 0xFFFFFFFEu);
    if (travFlags == (LEAF_BIT | 0x1)) {
      if (COUNT) cnt.instances++;
      const uint32_t instId = extract_start(leftOffset);
      const InstanceData* inst = s.instances + instId;
      const uint32_t geomId = inst->geomId;
      float m[16];
      load_mat(inst->transformInv, m);
      const float3 ray_pos = matmul4x3(m, pos);
      const float3 ray_dir = matmul3x3(m, dir);
      if (COOP) {
        if (!coop_blas<KINDS>(s, coop_stack, ray_pos, ray_dir, posAndNear.w, instId, geomId, stopOnFirstHit, hit,
                              cnt)) {
          // frontier too wide or tree deeper than the key: every lane walks the same ray
          // sequentially on one shared LDS stack
          LaneStack<STACK_SIZE, 1> seq;
          seq.col = (lds_u32*)coop_stack;
          seq.spill = nullptr;
          uint32_t seq_steps = 0;
          bool seq_abort = false;
          blas_traverse<KINDS, STACK_SIZE, 1, false>(s, ray_pos, ray_dir, posAndNear.w, instId, geomId,
                                                     stopOnFirstHit, hit, seq, overflow, cnt, nullptr, 0u,
                                                     seq_steps, seq_abort);
          __syncthreads();
        }
      } else {
        lds_u32* ot_col = ((KINDS & LK_MULTI) && keep) ? keep + 6 * BLOCK : nullptr;
        if (FAST && box_test_fast(ray_pos, ray_dir))
          blas_traverse<KINDS, S, BLOCK, COUNT, true>(s, ray_pos, ray_dir, posAndNear.w, instId, geomId, stopOnFirstHit,
                                                      hit, st, overflow, cnt, top_pairs, budget, steps, aborted,
                                                      ot_col, rec);
        else
          blas_traverse<KINDS, S, BLOCK, COUNT, false>(s, ray_pos, ray_dir, posAndNear.w, instId, geomId,
                                                       stopOnFirstHit, hit, st, overflow, cnt, top_pairs, budget,
                                                       steps, aborted, ot_col, rec);
        if (aborted) return hit;
      }
    }
  } while (nodeIdx < 0xFFFFFFFEu && !(stopOnFirstHit && hit.primId != 0xFFFFFFFFu));
  finish_hit<KINDS>(s, hit);
  return hit;
}

// Ray of thread i when the lanes of each wave are dealt out in groups of 2^group_log2: group g
// of wave w takes the same lanes of 64-ray chunk (w + g * n64 / groups) mod n64, so a wave holds
// rays from 64 / 2^group_log2 distant parts of the batch. Neighbouring primary rays graze the
// same silhouettes and a wave steps at the pace of its slowest lane; spreading them shortens the
// slowest waves. A bijection on [0, 64 * n64); the last n mod 64 rays keep their thread.
__device__ __forceinline__ uint32_t interleave_ray(uint32_t i, uint32_t n, uint32_t group_log2) {
  const uint32_t n64 = n >> 6;
  const uint32_t groups = 64u >> group_log2;
  if (n64 < groups || i >= n64 * 64u) return i;
  const uint32_t w = i >> 6, l = i & 63u, g = l >> group_log2;
  uint32_t c = w + g * (n64 / groups);   // < 2 * n64
  if (c >= n64) c -= n64;
  return c * 64u + l;
}

// Logical block index so that each XCD (blocks are dealt round-robin to the 8 XCDs) works
// on one contiguous range of rays and keeps its own L2 warm with the nodes those rays touch.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ void store_hit(CRT_Hit* out, uint32_t i, const CRT_Hit& h) {
  float4* o = reinterpret_cast<float4*>(out + i);
  o[0] = make4(h.t, __uint_as_float(h.primId), __uint_as_float(h.instId), __uint_as_float(h.geomId));
  o[1] = make4(h.coords[0], h.coords[1], h.coords[2], h.coords[3]);
}

}  // namespace lrt
