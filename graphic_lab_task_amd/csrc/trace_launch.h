// Closest-hit / any-hit ray query kernels for gfx950 and their launch templates. Each leaf-kind
// instantiation is compiled in a translation unit of its own (trace_<kind>.hip).
//
// A trace call runs in two passes:
//  1. one ray per lane, the reference's sequential walk, with a step budget: the few rays
//     that need more steps than the budget (grazing rays crossing many boxes) leave their
//     index in a queue instead of holding their whole wave hostage;
//  2. a persistent kernel where each wave takes one queued ray and walks its BLAS
//     cooperatively (trace_coop.h), which returns the same hit in far fewer dependent steps.
// The instrumented (counted) variant runs pass 1 without a budget.
#pragma once
#include "kernels.h"
#include "traverse.h"

namespace lrt {

// Pass 1. `budget` 0: unlimited (no queue).
template <uint32_t KINDS, int S, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace_closest(SceneView s, const float4* __restrict__ pos,
                                                         const float4* __restrict__ dir, CRT_Hit* __restrict__ out,
                                                         uint32_t* __restrict__ any_out, uint32_t n,
                                                         uint32_t* overflow_count, uint32_t xcd_remap,
                                                         uint32_t* __restrict__ counts, uint32_t* __restrict__ spill,
                                                         uint32_t budget, TailQueue* queue, uint32_t group_log2) {
  __shared__ uint32_t lds[S * BLOCK];
  // world ray of each lane while its BLAS is walked (+ the top of the multi-octree walk's stack)
  __shared__ uint32_t keep[(6 + ((KINDS & LK_MULTI) ? 2 * OT_LDS : 0)) * BLOCK];
  // top BFS levels of the cached geometry's BLAS: dynamic, s.cache_pairs entries (none for a
  // scene without a cached geometry, which then keeps its LDS for more resident waves)
  extern __shared__ BVHNodePair top_pairs[];
  // stage the top BFS levels of the cached geometry's BLAS (every thread reaches the barrier)
  if (s.cache_pairs) {
    const uint4* src = reinterpret_cast<const uint4*>(s.nodes + s.geoms[s.cache_geom].bvhOffset);
    uint4* dst = reinterpret_cast<uint4*>(top_pairs);
    for (uint32_t k = threadIdx.x; k < s.cache_pairs * 4; k += BLOCK) dst[k] = src[k];
  }
  __syncthreads();
  const uint32_t b = xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t i = b * BLOCK + threadIdx.x;
  if (i >= n) return;
  if (!COUNT && group_log2 < 6) i = interleave_ray(i, n, group_log2);
  LaneStack<S, BLOCK> st;
  st.col = (lds_u32*)(lds) + threadIdx.x;
  st.spill = spill + size_t(i) * (STACK_SIZE - S);
  uint32_t overflow = 0;
  float4 p = pos[i];
  float4 d = dir[i];
  if (any_out) d.w *= -1.0f;   // RayQuery_AnyHit flips the sign of w (BVH2Common.cpp:3394-3399)
  RayCounts cnt{};
  bool aborted = false;
  const uint64_t t_start = COUNT ? wall_clock64() : 0;
  CRT_Hit h;
  if (!COUNT && box_test_fast(xyz(p), xyz(d)))
    h = ray_query_nearest<KINDS, S, BLOCK, COUNT, false, true>(s, p, d, st, overflow, cnt, (const lds_pair*)top_pairs,
                                                               budget, aborted, nullptr,
                                                               (lds_u32*)(keep) + threadIdx.x);
  else
    h = ray_query_nearest<KINDS, S, BLOCK, COUNT, false, false>(s, p, d, st, overflow, cnt,
                                                                (const lds_pair*)top_pairs, COUNT ? 0u : budget,
                                                                aborted, nullptr, (lds_u32*)(keep) + threadIdx.x);
  if (aborted) {
    queue->list[atomicAdd(&queue->count, 1u)] = i;
    return;
  }
  if (out) store_hit(out, i, h);
  if (COUNT) {
    const uint64_t t_end = wall_clock64();
    uint32_t* c = counts + size_t(i) * RAY_COUNTERS;
    c[0] = cnt.blas_nodes; c[1] = cnt.leaves; c[2] = cnt.prims;
    c[3] = cnt.tlas_nodes; c[4] = cnt.instances; c[5] = cnt.sdf_voxels;
    c[6] = uint32_t(t_start);   // device wall clock (100 MHz) at ray start, low 32 bits
    c[7] = uint32_t(t_end - t_start);
  }
  if (any_out) any_out[i] = (h.geomId != 0xFFFFFFFFu) ? 1u : 0u;
  if (overflow && overflow_count) atomicAdd(overflow_count, 1u);
}

// Pass 2: persistent waves; each takes the next queued ray and traces it cooperatively.
template <uint32_t KINDS>
__global__ __launch_bounds__(64) void k_trace_tail(SceneView s, const float4* __restrict__ pos,
                                                   const float4* __restrict__ dir, CRT_Hit* __restrict__ out,
                                                   uint32_t* __restrict__ any_out, uint32_t* overflow_count,
                                                   TailQueue* queue) {
  __shared__ uint32_t coop[COOP_CAP * 4];
  const uint32_t total = __atomic_load_n(&queue->count, __ATOMIC_RELAXED);
  for (;;) {
    uint32_t idx = 0;
    if (__lane_id() == 0) idx = atomicAdd(&queue->next, 1u);
    idx = __shfl(idx, 0, 64);
    if (idx >= total) {
      // Every wave ends on exactly one fetch past the queue, after its last ray; the fetch that
      // returns total + waves - 1 is the last of all, so that wave clears the queue for the next
      // call on this stream (no reset launch precedes pass 1).
      if (__lane_id() == 0 && idx == total + gridDim.x - 1) {
        queue->count = 0;
        queue->next = 0;
      }
      break;
    }
    const uint32_t i = queue->list[idx];
    float4 p = pos[i];
    float4 d = dir[i];
    if (any_out) d.w *= -1.0f;
    LaneStack<1, 64> unused;
    unused.col = nullptr;
    unused.spill = nullptr;
    RayCounts cnt{};
    uint32_t overflow = 0;
    bool aborted = false;
    const CRT_Hit h = ray_query_nearest<KINDS, 1, 64, false, true>(s, p, d, unused, overflow, cnt, nullptr, 0u,
                                                                   aborted, (lds_entry*)coop);
    if (__lane_id() == 0) {
      if (out) store_hit(out, i, h);
      if (any_out) any_out[i] = (h.geomId != 0xFFFFFFFFu) ? 1u : 0u;
      if (overflow && overflow_count) atomicAdd(overflow_count, 1u);
    }
    __syncthreads();
  }
}

// Persistent single pass: a grid of resident blocks whose lanes each take the next untraced ray
// as soon as their current one is finished (a wave fetches rays for all its idle lanes with one
// atomic), instead of a wave waiting for its slowest lane before new waves start. A lane's work
// is cut at TLAS-leg granularity: one leg walks the TLAS to the next instance leaf (or to the end)
// and then that instance's BLAS; a lane whose ray ended takes a new ray before the next leg.
// The stack spill is indexed by the lane's slot in the grid. counter[0] hands out rays,
// counter[1] counts finished waves; the last wave clears both for the next call on this stream.
template <uint32_t KINDS, int S, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_persistent(SceneView s, const float4* __restrict__ pos,
                                                            const float4* __restrict__ dir, CRT_Hit* __restrict__ out,
                                                            uint32_t* __restrict__ any_out, uint32_t n,
                                                            uint32_t* overflow_count, uint32_t* __restrict__ spill,
                                                            uint32_t* counter) {
  __shared__ uint32_t lds[S * BLOCK];
  __shared__ uint32_t keep[(6 + ((KINDS & LK_MULTI) ? 2 * OT_LDS : 0)) * BLOCK];
  extern __shared__ BVHNodePair top_pairs[];
  if (s.cache_pairs) {
    const uint4* src = reinterpret_cast<const uint4*>(s.nodes + s.geoms[s.cache_geom].bvhOffset);
    uint4* dst = reinterpret_cast<uint4*>(top_pairs);
    for (uint32_t k = threadIdx.x; k < s.cache_pairs * 4; k += BLOCK) dst[k] = src[k];
  }
  __syncthreads();
  LaneStack<S, BLOCK> st;
  st.col = (lds_u32*)(lds) + threadIdx.x;
  st.spill = spill + size_t(blockIdx.x * BLOCK + threadIdx.x) * (STACK_SIZE - S);
  lds_u32* kp = (lds_u32*)(keep) + threadIdx.x;
  lds_u32* ot_col = (KINDS & LK_MULTI) ? kp + 6 * BLOCK : nullptr;
  const uint32_t lane = __lane_id();
  bool exhausted = false;   // wave-uniform: the counter has passed n
  bool has_ray = false;
  uint32_t i = 0, nodeIdx = 0, overflow = 0;
  float tnear = 0.0f;
  bool first_hit = false, fast = false;
  CRT_Hit hit = miss_hit(0.0f);
  RayCounts cnt{};
  for (;;) {
    if (!exhausted) {
      const uint64_t need = __ballot(!has_ray);
      if (need) {
        const uint32_t leader = uint32_t(__ffsll((unsigned long long)need) - 1);
        const uint32_t count = uint32_t(__popcll(need));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counter, count);
        base = __shfl(base, leader, 64);
        if (!has_ray) {
          const uint32_t idx = base + uint32_t(__popcll(need & ((1ull << lane) - 1ull)));
          if (idx < n) {
            i = idx;
            has_ray = true;
            const float4 p = pos[i];
            float4 d = dir[i];
            if (any_out) d.w *= -1.0f;   // RayQuery_AnyHit flips the sign of w (BVH2Common.cpp:3394-3399)
            first_hit = d.w <= 0.0f;
            if (first_hit) d.w *= -1.0f;
            tnear = p.w;
            fast = box_test_fast(xyz(p), xyz(d));
            hit = miss_hit(d.w);
            nodeIdx = 0;
            overflow = 0;
            kp[0 * BLOCK] = __float_as_uint(p.x);
            kp[1 * BLOCK] = __float_as_uint(p.y);
            kp[2 * BLOCK] = __float_as_uint(p.z);
            kp[3 * BLOCK] = __float_as_uint(d.x);
            kp[4 * BLOCK] = __float_as_uint(d.y);
            kp[5 * BLOCK] = __float_as_uint(d.z);
          }
        }
        if (base + count >= n) exhausted = true;
      }
    }
    if (!__any(has_ray)) break;
    bool at_leaf = false;
    uint32_t leftOffset = 0;
    if (has_ray) {
      // one TLAS leg of BVHRT::RayQuery_NearestHit (BVH2Common.cpp:3339-3360)
      asm volatile("" ::: "memory");
      const float3 wpos = make3(__uint_as_float(kp[0 * BLOCK]), __uint_as_float(kp[1 * BLOCK]),
                                __uint_as_float(kp[2 * BLOCK]));
      const float3 wdir = make3(__uint_as_float(kp[3 * BLOCK]), __uint_as_float(kp[4 * BLOCK]),
                                __uint_as_float(kp[5 * BLOCK]));
      const float3 rayDirInv = safe_inverse(wdir);
      // The leaf decision is recorded in the iteration that finds the leaf. (A do-while whose
      // leaf test came after the loop, reading the flags of the lane's last iteration, returned
      // the decision of a later iteration for lanes that had left the loop early: rays lost the
      // instance they had reached.)
      for (;;) {
        const BVHNode cur = load_node(s.tlas + nodeIdx);
        const float2 bh = fast ? ray_box2_fast(wpos, rayDirInv, cur.boxMin, cur.boxMax)
                               : ray_box2(wpos, rayDirInv, cur.boxMin, cur.boxMax);
        const bool intersects = (bh.x <= bh.y) && (bh.y > tnear) && (bh.x < hit.t);
        const bool is_leaf = (cur.leftOffset & LEAF_BIT) != 0;
        if (is_leaf && intersects) {
          at_leaf = true;
          leftOffset = cur.leftOffset;
          nodeIdx = cur.escapeIndex;
          break;
        }
        nodeIdx = (is_leaf || !intersects) ? cur.escapeIndex : cur.leftOffset;
        if (nodeIdx == 0 || nodeIdx >= 0xFFFFFFFEu) break;
      }
    }
    if (at_leaf) {
      const uint32_t instId = extract_start(leftOffset);
      const InstanceData* inst = s.instances + instId;
      const uint32_t geomId = inst->geomId;
      float m[16];
      load_mat(inst->transformInv, m);
      const float3 wpos = make3(__uint_as_float(kp[0 * BLOCK]), __uint_as_float(kp[1 * BLOCK]),
                                __uint_as_float(kp[2 * BLOCK]));
      const float3 wdir = make3(__uint_as_float(kp[3 * BLOCK]), __uint_as_float(kp[4 * BLOCK]),
                                __uint_as_float(kp[5 * BLOCK]));
      const float3 ray_pos = matmul4x3(m, wpos);
      const float3 ray_dir = matmul3x3(m, wdir);
      uint32_t steps = 0;
      bool aborted = false;
      if (fast && box_test_fast(ray_pos, ray_dir))
        blas_traverse<KINDS, S, BLOCK, false, true>(s, ray_pos, ray_dir, tnear, instId, geomId, first_hit, hit, st,
                                                    overflow, cnt, (const lds_pair*)top_pairs, 0u, steps, aborted,
                                                    ot_col, nullptr);
      else
        blas_traverse<KINDS, S, BLOCK, false, false>(s, ray_pos, ray_dir, tnear, instId, geomId, first_hit, hit, st,
                                                     overflow, cnt, (const lds_pair*)top_pairs, 0u, steps, aborted,
                                                     ot_col, nullptr);
    }
    if (has_ray && !(nodeIdx < 0xFFFFFFFEu && !(first_hit && hit.primId != 0xFFFFFFFFu))) {
      finish_hit<KINDS>(s, hit);
      if (out) store_hit(out, i, hit);
      if (any_out) any_out[i] = (hit.geomId != 0xFFFFFFFFu) ? 1u : 0u;
      if (overflow && overflow_count) atomicAdd(overflow_count, 1u);
      has_ray = false;
    }
  }
  if (lane == 0) {
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    if (atomicAdd(counter + 1, 1u) == waves - 1) {
      atomicExch(counter, 0u);
      atomicExch(counter + 1, 0u);
    }
  }
}

// ---------------------------------------------------------------------------------------
// log2 of the lanes per group of interleave_ray; 6 (whole waves) means no interleaving
static inline uint32_t lanes_log2(uint32_t lanes) {
  uint32_t l = 0;
  while (l < 6 && (1u << l) < lanes) ++l;
  return lanes == 0 ? 6u : l;
}

template <uint32_t KINDS, int S, int BLOCK>
static hipError_t launch_trace_t(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  const uint32_t grid = (a.n + BLOCK - 1) / BLOCK;
  if (grid == 0) return hipSuccess;
  const size_t lds_cache = size_t(s.cache_pairs) * sizeof(BVHNodePair);
  if (a.counts) {
    if constexpr ((KINDS & LK_TRI_POSTPONE) != 0) {
      return hipErrorInvalidValue;   // counted runs take the plain walk (launch_trace routes them)
    } else {
      hipLaunchKernelGGL((k_trace_closest<KINDS, S, BLOCK, true>), dim3(grid), dim3(BLOCK), lds_cache, stream, s,
                         a.pos, a.dir, a.out, a.any_out, a.n, a.overflow_count, a.xcd_remap, a.counts, a.spill, 0u,
                         a.queue, 6u);
      return hipGetLastError();
    }
  }
  if (a.persistent && a.pcount != nullptr) {
    // resident blocks only: every block's lanes loop over the batch
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_trace_persistent<KINDS, S, BLOCK>,
                                                                BLOCK, lds_cache);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const uint32_t resident = uint32_t(per_cu > 0 ? per_cu : 1) * uint32_t(cus > 0 ? cus : 1);
    const uint32_t pgrid = grid < resident ? grid : resident;
    hipLaunchKernelGGL((k_trace_persistent<KINDS, S, BLOCK>), dim3(pgrid), dim3(BLOCK), lds_cache, stream, s, a.pos,
                       a.dir, a.out, a.any_out, a.n, a.overflow_count, a.spill, a.pcount);
    return hipGetLastError();
  }
  // the queue starts cleared (allocation) and every second pass leaves it cleared
  const bool tail = a.budget != 0 && a.queue != nullptr;
  hipLaunchKernelGGL((k_trace_closest<KINDS, S, BLOCK, false>), dim3(grid), dim3(BLOCK), lds_cache, stream, s,
                     a.pos, a.dir, a.out, a.any_out, a.n, a.overflow_count, a.xcd_remap, a.counts, a.spill,
                     tail ? a.budget : 0u, a.queue, lanes_log2(a.ray_groups));
  if (tail) {
    const uint32_t waves = a.tail_waves ? a.tail_waves : 1024u;
    // pass 2 walks cooperatively; the postponement option only shapes pass 1
    hipLaunchKernelGGL((k_trace_tail<KINDS & ~LK_TRI_POSTPONE>), dim3(waves), dim3(64), 0, stream, s, a.pos, a.dir,
                       a.out, a.any_out, a.overflow_count, a.queue);
  }
  return hipGetLastError();
}

template <uint32_t KINDS>
hipError_t launch_kinds(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  if (a.counts) return launch_trace_t<KINDS, 16, 256>(s, a, stream);   // one shape for the counted variant
  const int cfg = a.block * 1000 + a.lds_stack;
  switch (cfg) {
    case 64 * 1000 + 16:  return launch_trace_t<KINDS, 16, 64>(s, a, stream);
    case 128 * 1000 + 16: return launch_trace_t<KINDS, 16, 128>(s, a, stream);
    case 256 * 1000 + 16: return launch_trace_t<KINDS, 16, 256>(s, a, stream);
    case 64 * 1000 + 32:  return launch_trace_t<KINDS, 32, 64>(s, a, stream);
    case 128 * 1000 + 32: return launch_trace_t<KINDS, 32, 128>(s, a, stream);
    case 256 * 1000 + 32: return launch_trace_t<KINDS, 32, 256>(s, a, stream);
    case 64 * 1000 + 8:   return launch_trace_t<KINDS, 8, 64>(s, a, stream);
    case 128 * 1000 + 8:  return launch_trace_t<KINDS, 8, 128>(s, a, stream);
    case 256 * 1000 + 8:  return launch_trace_t<KINDS, 8, 256>(s, a, stream);
    default: return hipErrorInvalidValue;
  }
}

// One instantiation per translation unit (trace_tri.hip, trace_frame.hip, ...).
hipError_t launch_trace_tri(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_tri_postponed(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_frame(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_sbs(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_multi(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_mixed(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_coct(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_dag(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_scom2(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_sbs_smooth(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_mixed_smooth(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_sbs_tricubic(const SceneView& s, const TraceArgs& a, hipStream_t stream);
hipError_t launch_trace_mixed_tricubic(const SceneView& s, const TraceArgs& a, hipStream_t stream);

}  // namespace lrt
