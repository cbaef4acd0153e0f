// Closest-hit / any-hit ray query kernels for gfx950 and the primary-ray generator.
#include "trace_device.h"
#include "kernels.h"

namespace lrt {

// BVHRT::BVH2TraverseF32 (MultiRenderer/BVH2Common.cpp:3239-3320): stack-based walk over
// fat nodes, nearer child first, far child pushed; leaves dispatched on the geometry tag.
template <uint32_t KINDS, int S, bool COUNT>
__device__ __forceinline__ void blas_traverse(const SceneView& s, float3 ray_pos, float3 ray_dir, float tNear,
                                              uint32_t instId, uint32_t geomId, bool stopOnFirstHit, CRT_Hit& hit,
                                              PendingTri& pend, LaneStack<S>& st, uint32_t& overflow,
                                              RayCounts& cnt) {
  const uint32_t bvhOffset = s.geoms[geomId].bvhOffset;
  const BVHNodePair* nodes = s.nodes + bvhOffset;
  int top = 0;
  uint32_t leftNodeOffset = 0;
  bool hitFound = false;
  const float3 rayDirInv = safe_inverse(ray_dir);
  while (top >= 0 && !hitFound) {
    while (top >= 0 && ((leftNodeOffset & LEAF_BIT) == 0)) {
      if (COUNT) cnt.blas_nodes++;
      const BVHNodePair fat = load_pair(nodes + leftNodeOffset);
      const uint32_t n0 = fat.left.leftOffset;
      const uint32_t n1 = fat.right.leftOffset;
      const float2 tm0 = ray_box2(ray_pos, rayDirInv, fat.left.boxMin, fat.left.boxMax);
      const float2 tm1 = ray_box2(ray_pos, rayDirInv, fat.right.boxMin, fat.right.boxMax);
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
    }
    if (top >= 0 && leftNodeOffset != 0xFFFFFFFFu) {
      // BVHRT::IntersectAllPrimitivesInLeaf (BVH2Common.cpp:47-124)
      const uint32_t aabbId = extract_start(leftNodeOffset);
      const float tNearSdf = fmax_ref(tNear, 0.01f);
      const float tPrev = hit.t;
      const uint32_t packed = s.prim_id_count[s.start_end[geomId].x + aabbId];
      const uint32_t a_start = extract_start(packed);
      const uint32_t a_count = extract_count(packed);
      const uint32_t tag = s.geom_tags[geomId];
      if (COUNT) {
        cnt.leaves++;
        if (tag == TAG_TRIANGLE) cnt.prims += a_count;
      }
      if ((KINDS & LK_TRI) && tag == TAG_TRIANGLE) {
        intersect_tris(s, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, a_count, hit, pend);
      } else if ((KINDS & LK_FRAME) && tag == TAG_SDF_NODE) {
        frame_node_intersect(s, s.geoms[geomId].type, ray_pos, ray_dir, tNearSdf, instId, geomId, a_start, hit);
        if (hit.t < tPrev) pend.valid = 0;
      } else if ((KINDS & LK_SBS) && tag == TAG_SDF_BRICK) {
        cnt.sdf_voxels += sbs_brick_intersect(s, s.geoms[geomId].type, ray_pos, ray_dir, tNearSdf, instId, geomId,
                                              a_start, hit);
        if (hit.t < tPrev) pend.valid = 0;
      }
      if (s.preset.render_mode == MULTI_RENDER_MODE_INTERSECTION_COUNT)
        hit.instId = (hit.instId == 0xFFFFFFFFu) ? 1 : hit.instId + 1;
      const uint32_t hitTag = (hit.t >= tPrev) ? TAG_NONE : tag;
      hitFound = (hitTag != TAG_NONE) && (hitTag != TAG_TRIANGLE || stopOnFirstHit);
    }
    top--;
    leftNodeOffset = st.get(top > 0 ? top : 0);
  }
}

// BVHRT::RayQuery_NearestHit (MultiRenderer/BVH2Common.cpp:3322-3392): stackless TLAS walk
// over escape indices, ray moved to instance space (direction not renormalised).
template <uint32_t KINDS, int S, bool COUNT>
__device__ __forceinline__ CRT_Hit ray_query_nearest(const SceneView& s, float4 posAndNear, float4 dirAndFar,
                                                     LaneStack<S>& st, uint32_t& overflow, RayCounts& cnt) {
  bool stopOnFirstHit = (dirAndFar.w <= 0.0f);
  if (stopOnFirstHit) dirAndFar.w *= -1.0f;
  const float3 pos = xyz(posAndNear);
  const float3 dir = xyz(dirAndFar);
  const float3 rayDirInv = safe_inverse(dir);
  CRT_Hit hit;
  hit.t = dirAndFar.w;
  hit.primId = 0xFFFFFFFFu;
  hit.instId = 0xFFFFFFFFu;
  hit.geomId = 0xFFFFFFFFu;
  hit.coords[0] = 1.0f;
  hit.coords[1] = 0.0f;
  hit.coords[2] = 0.0f;
  hit.coords[3] = 0.0f;
  PendingTri pend;
  pend.valid = 0;
  uint32_t nodeIdx = 0;
  do {
    uint32_t travFlags = 0;
    uint32_t leftOffset = 0;
    do {
      if (COUNT) cnt.tlas_nodes++;
      const BVHNode cur = load_node(s.tlas + nodeIdx);
      const float2 bh = ray_box2(pos, rayDirInv, cur.boxMin, cur.boxMax);
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
      blas_traverse<KINDS, S, COUNT>(s, ray_pos, ray_dir, posAndNear.w, instId, geomId, stopOnFirstHit, hit, pend,
                                     st, overflow, cnt);
    }
  } while (nodeIdx < 0xFFFFFFFEu && !(stopOnFirstHit && hit.primId != 0xFFFFFFFFu));

  if ((KINDS & LK_TRI) && pend.valid && hit.geomId != 0xFFFFFFFFu && (hit.geomId >> SH_TYPE) == TYPE_MESH_TRIANGLE)
    resolve_tri_normal(s, pend, hit);
  if (hit.geomId < 0xFFFFFFFFu && ((hit.geomId >> SH_TYPE) == TYPE_MESH_TRIANGLE)) {
    const uint2 go = s.geoms[hit.geomId & GEOM_ID_MASK].offset;
    hit.primId = s.prim_indices[go.x / 3 + hit.primId];
  }
  return hit;
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

template <uint32_t KINDS, int S, int BLOCK, bool COUNT>
__global__ __launch_bounds__(BLOCK) void k_trace_closest(SceneView s, const float4* __restrict__ pos,
                                                         const float4* __restrict__ dir, CRT_Hit* __restrict__ out,
                                                         uint32_t* __restrict__ any_out, uint32_t n,
                                                         uint32_t* overflow_count, uint32_t xcd_remap,
                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t lds[S > 0 ? S * BLOCK : 1];
  const uint32_t b = xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint32_t i = b * BLOCK + threadIdx.x;
  if (i >= n) return;
  LaneStack<S> st;
  st.col = lds + threadIdx.x;
  st.stride = BLOCK;
  uint32_t overflow = 0;
  float4 p = pos[i];
  float4 d = dir[i];
  if (any_out) d.w *= -1.0f;   // RayQuery_AnyHit flips the sign of w (BVH2Common.cpp:3394-3399)
  RayCounts cnt{};
  const CRT_Hit h = ray_query_nearest<KINDS, S, COUNT>(s, p, d, st, overflow, cnt);
  if (out) store_hit(out, i, h);
  if (COUNT) {
    uint32_t* c = counts + size_t(i) * 6;
    c[0] = cnt.blas_nodes; c[1] = cnt.leaves; c[2] = cnt.prims;
    c[3] = cnt.tlas_nodes; c[4] = cnt.instances; c[5] = cnt.sdf_voxels;
  }
  if (any_out) any_out[i] = (h.geomId != 0xFFFFFFFFu) ? 1u : 0u;
  if (overflow && overflow_count) atomicAdd(overflow_count, 1u);
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
template <uint32_t KINDS, int S, int BLOCK>
static hipError_t launch_trace_t(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  const uint32_t grid = (a.n + BLOCK - 1) / BLOCK;
  if (grid == 0) return hipSuccess;
  if (a.counts)
    hipLaunchKernelGGL((k_trace_closest<KINDS, S, BLOCK, true>), dim3(grid), dim3(BLOCK), 0, stream, s, a.pos, a.dir,
                       a.out, a.any_out, a.n, a.overflow_count, a.xcd_remap, a.counts);
  else
    hipLaunchKernelGGL((k_trace_closest<KINDS, S, BLOCK, false>), dim3(grid), dim3(BLOCK), 0, stream, s, a.pos,
                       a.dir, a.out, a.any_out, a.n, a.overflow_count, a.xcd_remap, a.counts);
  return hipGetLastError();
}

template <uint32_t KINDS>
static hipError_t launch_kinds(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
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
  if (a.counts) {  // the counting variant has one launch shape
    if (k == LK_TRI) return launch_trace_t<LK_TRI, 16, 256>(s, a, stream);
    if (k == LK_FRAME) return launch_trace_t<LK_FRAME, 16, 256>(s, a, stream);
    if (k == LK_SBS) return launch_trace_t<LK_SBS, 16, 256>(s, a, stream);
    return launch_trace_t<LK_TRI | LK_FRAME | LK_SBS, 16, 256>(s, a, stream);
  }
  if (k == LK_TRI) return launch_kinds<LK_TRI>(s, a, stream);
  if (k == LK_FRAME) return launch_kinds<LK_FRAME>(s, a, stream);
  if (k == LK_SBS) return launch_kinds<LK_SBS>(s, a, stream);
  return launch_kinds<LK_TRI | LK_FRAME | LK_SBS>(s, a, stream);
}

hipError_t launch_eye_rays(const EyeRayParams& prm, const float2* jitter, float4* pos, float4* dir,
                           hipStream_t stream) {
  dim3 block(16, 16);
  dim3 grid((prm.width + 15) / 16, (prm.height + 15) / 16);
  hipLaunchKernelGGL(k_eye_rays, grid, block, 0, stream, prm, jitter, pos, dir);
  return hipGetLastError();
}

}  // namespace lrt
