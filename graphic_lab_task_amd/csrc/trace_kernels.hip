// Closest-hit / any-hit ray query kernels for gfx950 and the primary-ray generator.
//
// A trace call runs in two passes:
//  1. one ray per lane, the reference's sequential walk, with a step budget: the few rays
//     that need more steps than the budget (grazing rays crossing many boxes) leave their
//     index in a queue instead of holding their whole wave hostage;
//  2. a persistent kernel where each wave takes one queued ray and walks its BLAS
//     cooperatively (trace_coop.h), which returns the same hit in far fewer dependent steps.
// The instrumented (counted) variant runs pass 1 without a budget.
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
  __shared__ uint32_t keep[(6 + ((KINDS & LK_MULTI) ? 4 * OT_LDS : 0)) * BLOCK];
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

// SuperBlockIndex2DOpt (utils/raytracing/render_common.h:225-242): 4x4 blocks in 2x2 super blocks.
__device__ __forceinline__ uint32_t super_block_index(uint32_t x, uint32_t y, uint32_t w) {
  const uint32_t local = (y & 3u) * 4 + (x & 3u);
  const uint32_t bx = x >> 2, by = y >> 2;
  const uint32_t localH = (by & 1u) * 2 + (bx & 1u);
  const uint32_t wbH = (w >> 2) >> 1;
  return ((bx >> 1) + (by >> 1) * wbH) * 64 + localH * 16 + local;
}

// MultiRenderer::initEyeRay (MultiRenderer/MultiRenderer.cpp:142-183) with the cutting plane
// disabled: EyeRayDirNormalized + transform_ray3f (utils/raytracing/render_common.h:25-55).
__global__ void k_eye_rays(EyeRayParams prm, const float2* __restrict__ jitter, float4* __restrict__ pos,
                           float4* __restrict__ dir) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= prm.width || y >= prm.height) return;
  const uint32_t pix = y * prm.width + x;
  float2 dd = prm.d;
  if (jitter) dd = jitter[pix];
  const float fx = (float(x) + dd.x) / float(prm.packed_w);
  const float fy = (float(y) + dd.y) / float(prm.packed_h);
  float4 p4 = make4(2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f);
  p4 = mat_mul4(prm.proj_inv, p4);
  const float w = p4.w;
  p4 = make4(p4.x / w, p4.y / w, p4.z / w, p4.w / w);
  float3 rd = normalize3(xyz(p4));
  const float3 zero = make3(0, 0, 0);
  const float3 o = matmul4x3(prm.world_view_inv, zero);
  const float3 o2 = matmul4x3(prm.world_view_inv, zero + 100.0f * rd);
  rd = normalize3(o2 - o);
  const uint32_t idx = prm.order == 1 ? super_block_index(x, y, prm.packed_w) : pix;
  pos[idx] = make4(o.x, o.y, o.z, 1e-9f);
  dir[idx] = make4(rd.x, rd.y, rd.z, 1e9f);
}

// ---------------------------------------------------------------------------------------
// log2 of the lanes per group of interleave_ray; 6 (whole waves) means no interleaving
static uint32_t lanes_log2(uint32_t lanes) {
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
  const hipError_t e = hipGetLastError();
  // a call whose second pass did not launch must not leave queued rays to the next call
  if (e != hipSuccess && tail) (void)hipMemsetAsync(a.queue, 0, 2 * sizeof(uint32_t), stream);
  return e;
}

template <uint32_t KINDS>
static hipError_t launch_kinds(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
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

hipError_t launch_trace(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  const uint32_t k = s.leaf_kinds;
  if (k == LK_TRI)
    return a.postpone_tri && !a.counts ? launch_kinds<LK_TRI | LK_TRI_POSTPONE>(s, a, stream)
                                       : launch_kinds<LK_TRI>(s, a, stream);
  if (k == LK_FRAME) return launch_kinds<LK_FRAME>(s, a, stream);
  if (k == LK_SBS) return launch_kinds<LK_SBS>(s, a, stream);
  if (k == LK_MULTI) return launch_kinds<LK_MULTI>(s, a, stream);
  return launch_kinds<LK_TRI | LK_FRAME | LK_SBS | LK_MULTI>(s, a, stream);
}

hipError_t launch_eye_rays(const EyeRayParams& prm, const float2* jitter, float4* pos, float4* dir,
                           hipStream_t stream) {
  dim3 block(16, 16);
  dim3 grid((prm.width + 15) / 16, (prm.height + 15) / 16);
  hipLaunchKernelGGL(k_eye_rays, grid, block, 0, stream, prm, jitter, pos, dir);
  return hipGetLastError();
}

}  // namespace lrt
