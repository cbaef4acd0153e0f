// Differentiable sparse-brick-set rendering: forward colour and the adjoint of an image loss
// with respect to the brick set's distance values and per-lattice-point albedo.
//
// The forward image is C(ray) = trilinear albedo of the hit voxel at the hit point (miss: 0).
// The backward recomputes the hit (same traversal, then a second march over the hit brick
// that records the voxel; nothing is stored between the passes) and accumulates
//   dL/dA[k][c] += G_c * w_k(q)
//   dL/ds[j]    += -(sum_c G_c * dot(grad C_c(q), dir)) * w_j(q) / dot(grad f(q), dir)
// where q is the hit point in voxel coordinates, w the trilinear weights, f the trilinear
// distance and G = dL/dC (given, or 2(C - T) for the fused L2 loss): the hit point moves along
// the ray when a corner distance changes (implicit function theorem on f(q(t)) = 0).
#include "kernels.h"
#include "traverse.h"

namespace lrt {

namespace {

__device__ __forceinline__ void trilinear_weights(float3 q, float w[8], float3 gw[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float wx = (i & 4) ? q.x : 1.0f - q.x, dx = (i & 4) ? 1.0f : -1.0f;
    const float wy = (i & 2) ? q.y : 1.0f - q.y, dy = (i & 2) ? 1.0f : -1.0f;
    const float wz = (i & 1) ? q.z : 1.0f - q.z, dz = (i & 1) ? 1.0f : -1.0f;
    w[i] = wx * wy * wz;
    gw[i] = make3(dx * wy * wz, wx * dy * wz, wx * wy * dz);
  }
}

// Recomputes the voxel that produced hit `h` of the ray (p, d): one march over the hit brick
// with the same arithmetic as the traversal; returns false unless it lands on the same t.
__device__ bool recompute_hit_voxel(const SceneView& s, uint32_t diff_geom, float4 p, float4 d, const CRT_Hit& h,
                                    SbsVoxelRec& rec, float3& ray_dir_local) {
  if (h.geomId == 0xFFFFFFFFu) return false;
  const uint32_t geomId = h.geomId & GEOM_ID_MASK;
  if (geomId != diff_geom) return false;
  float m[16];
  load_mat(s.instances[h.instId].transformInv, m);
  const float3 ray_pos = matmul4x3(m, xyz(p));
  const float3 ray_dir = matmul3x3(m, xyz(d));
  CRT_Hit hh = miss_hit(fabs_ref(d.w));
  sbs_brick_intersect(s, geom_ref(s, geomId), ray_pos, ray_dir, fmax_ref(p.w, 0.01f), h.instId, geomId, h.primId, hh,
                      &rec);
  ray_dir_local = ray_dir;
  return __float_as_uint(hh.t) == __float_as_uint(h.t);
}

// MODE 0 forward, 1 backward with dL/dC, 2 backward of the L2 loss against a target, 3 both the
// forward colour and the L2 backward. Returns the ray's loss term |C - T|^2 (modes 2, 3).
// `rec`: the voxel record the traversal's own brick march left (t_local NaN if it recorded none);
// without one the hit brick is marched once more.
template <int MODE>
__device__ __forceinline__ float diff_ray(const SceneView& s, const SbsDiffArgs& a, uint32_t i, const CRT_Hit& h,
                                          SbsVoxelRec& rec) {
  const float4 p = a.pos[i];
  const float4 d = a.dir[i];
  float3 dir_l = make3(0, 0, 0);
  bool hit = false;
  if (h.geomId != 0xFFFFFFFFu && (h.geomId & GEOM_ID_MASK) == a.geom_id) {
    if (rec.t_local == rec.t_local) {   // recorded by the traversal
      float m[16];
      load_mat(s.instances[h.instId].transformInv, m);
      dir_l = matmul3x3(m, xyz(d));
      hit = true;
    } else {
      hit = recompute_hit_voxel(s, a.geom_id, p, d, h, rec, dir_l);
      if (!hit && a.counters) atomicAdd(a.counters, 1u);
    }
  }
  float3 C = make3(0, 0, 0);
  float w[8];
  float3 gw[8];
  float3 q = make3(0, 0, 0);
  if (hit) {
    q = rec.start_q + rec.t_local * dir_l;
    trilinear_weights(q, w, gw);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t ab = a.f_offset + a.n_dist + 3u * (rec.idx[k] - a.f_offset);
      C = C + w[k] * make3(a.params[ab], a.params[ab + 1], a.params[ab + 2]);
    }
  }
  if (MODE == 0 || MODE == 3) {
    if (a.colour) a.colour[i] = make4(C.x, C.y, C.z, hit ? 1.0f : 0.0f);
    if (MODE == 0) return 0.0f;
  }
  float3 G;
  float loss = 0.0f;
  if (MODE == 1) {
    G = xyz(a.dL_dcolour[i]);
  } else {
    const float3 T = xyz(a.target[i]);
    const float3 e = C - T;
    G = 2.0f * e;
    loss = dot3(e, e);
  }
  if (!hit) return loss;
  float3 gf = make3(0, 0, 0);
  float gCd = 0.0f;   // sum_c G_c * dot(grad C_c, dir)
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const uint32_t ab = a.f_offset + a.n_dist + 3u * (rec.idx[k] - a.f_offset);
    const float3 A = make3(a.params[ab], a.params[ab + 1], a.params[ab + 2]);
    gf = gf + a.params[rec.idx[k]] * gw[k];
    gCd += dot3(G, A) * dot3(gw[k], dir_l);
    unsafeAtomicAdd(a.grad + ab + 0, G.x * w[k]);
    unsafeAtomicAdd(a.grad + ab + 1, G.y * w[k]);
    unsafeAtomicAdd(a.grad + ab + 2, G.z * w[k]);
  }
  const float den = dot3(gf, dir_l);
  if (fabs_ref(den) > 1e-20f) {
    const float k0 = -gCd / den;
#pragma unroll
    for (int k = 0; k < 8; ++k) unsafeAtomicAdd(a.grad + rec.idx[k], k0 * w[k]);
  }
  return loss;
}

// Sum over the 64 lanes of a wave (every lane active).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int MODE, int S, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sbs_diff(SceneView s, SbsDiffArgs a) {
  __shared__ uint32_t lds[S * BLOCK];
  __shared__ BVHNodePair top_pairs[LDS_TOP_PAIRS];
  if (s.cache_pairs) {
    const uint4* src = reinterpret_cast<const uint4*>(s.nodes + s.geoms[s.cache_geom].bvhOffset);
    uint4* dst = reinterpret_cast<uint4*>(top_pairs);
    for (uint32_t k = threadIdx.x; k < s.cache_pairs * 4; k += BLOCK) dst[k] = src[k];
  }
  __syncthreads();
  const uint32_t b = a.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  uint32_t i = b * BLOCK + threadIdx.x;
  float loss = 0.0f;
  if (i < a.n) {
    i = interleave_ray(i, a.n, 3);   // groups of 8 lanes from distant parts of the batch
    LaneStack<S, BLOCK> st;
    st.col = (lds_u32*)(lds) + threadIdx.x;
    st.spill = a.spill + size_t(i) * (STACK_SIZE - S);
    uint32_t overflow = 0;
    RayCounts cnt{};
    bool aborted = false;
    SbsVoxelRec rec;
    rec.t_local = __int_as_float(0x7fc00000);
    const CRT_Hit h = ray_query_nearest<LK_SBS, S, BLOCK, false, false>(s, a.pos[i], a.dir[i], st, overflow, cnt,
                                                                        (const lds_pair*)top_pairs, 0u, aborted,
                                                                        nullptr, nullptr, &rec);
    loss = diff_ray<MODE>(s, a, i, h, rec);
  }
  if (MODE >= 2 && a.loss) {   // one atomic per wave instead of one per ray
    loss = wave_sum(loss);
    if (__lane_id() == 0) atomicAdd(a.loss, loss);
  }
}

// ---- batched launch ----
template <typename T>
__device__ __forceinline__ uint32_t find_item(const T* begin, uint32_t n_items, T x) {
  uint32_t lo = 0, hi = n_items;   // largest k with begin[k] <= x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (begin[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// The per-item set-up kernels run on a 2D grid, blockIdx.y = item, each item's blocks striding
// over its own range: no thread searches for its item.
__global__ void k_batch_zero(SbsBatchLaunch L) {
  const SbsBatchItem& item = L.items[blockIdx.y];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (blockIdx.y == 0) *L.counter = 0;
    if (item.args.loss) *item.args.loss = 0.0f;
  }
  float* g = item.args.grad;
  const uint64_t n = item.n_params;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t head = 0;   // 16-byte stores from the first 16-byte aligned element on
  if ((reinterpret_cast<uintptr_t>(g) & 15u) == 0u) {
    uint4* g4 = reinterpret_cast<uint4*>(g);
    for (uint64_t k = t; k < n / 4; k += stride) g4[k] = make_uint4(0u, 0u, 0u, 0u);
    head = n / 4 * 4;
  }
  for (uint64_t k = head + t; k < n; k += stride) g[k] = 0.0f;
}

// One thread per mask bit, 32 consecutive lanes per word (as k_sbs_mask_words); the loop bound
// is the same for a whole wave, so every lane reaches the ballot.
__global__ void k_batch_mask_words(SbsBatchLaunch L) {
  const SbsBatchItem& item = L.items[blockIdx.y];
  const SceneView& s = item.view;
  const SdfSBSHeader h = s.sbs_headers[0];
  const uint32_t bs = h.brick_size, v3 = bs * bs * bs, wpb = (v3 + 31u) / 32u;
  const uint64_t bits = uint64_t(item.n_bricks) * wpb * 32u;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint64_t wave = uint64_t(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u); wave < bits; wave += stride) {
    const uint64_t tid = wave + lane;
    const uint64_t word_id = tid >> 5;
    const bool in = tid < bits;
    const uint32_t brick = in ? uint32_t(word_id / wpb) : 0u, w = in ? uint32_t(word_id % wpb) : 0u;
    const uint32_t v = w * 32u + uint32_t(tid & 31u);
    const bool bit = in && v < v3 && sbs_voxel_occupied(s, h, s.sbs_roots[0], brick, v);
    const uint64_t m = __ballot(bit);
    if (in && (lane & 31u) == 0u) item.mask[size_t(brick) * (1u + wpb) + 1u + w] = uint32_t(m >> (lane & 32u));
  }
}

// One wave per brick of each item (blockIdx.y), as k_sbs_mask_bricks.
__global__ __launch_bounds__(64) void k_batch_mask_bricks(SbsBatchLaunch L) {
  extern __shared__ uint32_t flags[];
  const SbsBatchItem& item = L.items[blockIdx.y];
  const SceneView& s = item.view;
  const SdfSBSHeader h = s.sbs_headers[0];
  const uint32_t bs = h.brick_size, wpb = (bs * bs * bs + 31u) / 32u;
  const uint32_t root = s.sbs_roots[0];
  for (uint32_t b = blockIdx.x; b < item.n_bricks; b += gridDim.x)
    sbs_mask_brick(s, h, root, b, item.mask + size_t(b) * (1u + wpb), flags);
}

__global__ void k_batch_mask_summary(SbsBatchLaunch L) {
  const SbsBatchItem& item = L.items[blockIdx.y];
  const uint32_t bs = item.brick_size, wpb = (bs * bs * bs + 31u) / 32u;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < item.n_bricks; b += gridDim.x * blockDim.x) {
    uint32_t* m = item.mask + size_t(b) * (1u + wpb);
    uint32_t any = 0;
    for (uint32_t w = 0; w < wpb; ++w) any |= m[1 + w];
    m[0] = any != 0u ? 1u : 0u;
  }
}

// Persistent one-wave blocks: each takes the next 64-ray chunk of the whole batch from the
// global counter (one wave per block, so the block barrier that publishes the chunk never makes
// a wave wait for another's slow rays; the barrier also keeps the wave converged around the
// broadcast, which a lane-0 atomic followed by a lane read does not guarantee). Rays stay in
// batch order: here throughput counts, and neighbouring rays of similar cost share a chunk.
// S: stack entries in LDS. REC: the walk's own brick march records the hit voxel (12 more
// registers held through the walk) instead of a second march over the hit brick.
template <int S, int BLOCK, bool REC>
__global__ __launch_bounds__(BLOCK) void k_sbs_batch(SbsBatchLaunch L) {
  static_assert(BLOCK == 64 && SBS_BATCH_CHUNK == 64, "one wave per block, one chunk per wave");
  __shared__ uint32_t lds[S * BLOCK];
  __shared__ uint32_t chunk_sh;
  for (;;) {
    if (threadIdx.x == 0) chunk_sh = atomicAdd(L.counter, 1u);
    __syncthreads();
    const uint32_t chunk = chunk_sh;
    __syncthreads();
    if (chunk >= L.total_chunks) break;   // the same for the whole block: every block ends
    const uint32_t it = __builtin_amdgcn_readfirstlane(find_item(L.chunk_begin, L.n_items, chunk));
    const SbsBatchItem& item = L.items[it];
    const uint32_t i = (chunk - L.chunk_begin[it]) * SBS_BATCH_CHUNK + threadIdx.x;
    float loss = 0.0f;
    if (i < item.args.n) {
      LaneStack<S, BLOCK> st;
      st.col = (lds_u32*)(lds) + threadIdx.x;
      st.spill = L.spill + size_t(blockIdx.x * BLOCK + threadIdx.x) * (STACK_SIZE - S);
      uint32_t overflow = 0;
      RayCounts cnt{};
      bool aborted = false;
      SbsVoxelRec rec;
      rec.t_local = __int_as_float(0x7fc00000);
      const CRT_Hit h = ray_query_nearest<LK_SBS, S, BLOCK, false, false>(item.view, item.args.pos[i], item.args.dir[i],
                                                                          st, overflow, cnt, nullptr, 0u, aborted,
                                                                          nullptr, nullptr, REC ? &rec : nullptr);
      loss = diff_ray<3>(item.view, item.args, i, h, rec);
    }
    loss = wave_sum(loss);
    if (__lane_id() == 0 && item.args.loss) atomicAdd(item.args.loss, loss);
  }
}

template <int S, bool REC>
const void* batch_kernel() {
  return reinterpret_cast<const void*>(k_sbs_batch<S, SBS_BATCH_BLOCK, REC>);
}

template <int MODE>
hipError_t launch_mode(const SceneView& s, const SbsDiffArgs& a, hipStream_t stream) {
  constexpr int S = 16, BLOCK = 128;
  const uint32_t grid = (a.n + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL((k_sbs_diff<MODE, S, BLOCK>), dim3(grid), dim3(BLOCK), 0, stream, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t sbs_batch_blocks_per_cu(int lds_stack, int record, int* blocks) {
  const void* k = nullptr;
  switch (lds_stack * 2 + (record ? 1 : 0)) {
    case 16: k = batch_kernel<8, false>(); break;
    case 17: k = batch_kernel<8, true>(); break;
    case 32: k = batch_kernel<16, false>(); break;
    case 33: k = batch_kernel<16, true>(); break;
    case 64: k = batch_kernel<32, false>(); break;
    case 65: k = batch_kernel<32, true>(); break;
    default: return hipErrorInvalidValue;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k, SBS_BATCH_BLOCK, 0);
}

hipError_t launch_sbs_batch(const SbsBatchLaunch& L, hipStream_t stream) {
  if (L.n_items == 0) return hipSuccess;
  // x extents sized for the largest item (smaller items leave blocks idle)
  hipLaunchKernelGGL(k_batch_zero, dim3(L.zero_blocks, L.n_items), dim3(256), 0, stream, L);
  if (L.total_bricks && L.mask_flag_bytes) {
    hipLaunchKernelGGL(k_batch_mask_bricks, dim3(L.brick_blocks, L.n_items), dim3(64), L.mask_flag_bytes, stream, L);
  } else if (L.total_bricks) {   // very large bricks: one thread per voxel bit, then the summary words
    hipLaunchKernelGGL(k_batch_mask_words, dim3(L.mask_blocks, L.n_items), dim3(256), 0, stream, L);
    hipLaunchKernelGGL(k_batch_mask_summary, dim3(L.summary_blocks, L.n_items), dim3(256), 0, stream, L);
  }
  const dim3 g(L.grid), b(SBS_BATCH_BLOCK);
  switch (L.lds_stack * 2 + (L.record ? 1 : 0)) {
    case 16: hipLaunchKernelGGL((k_sbs_batch<8, SBS_BATCH_BLOCK, false>), g, b, 0, stream, L); break;
    case 17: hipLaunchKernelGGL((k_sbs_batch<8, SBS_BATCH_BLOCK, true>), g, b, 0, stream, L); break;
    case 32: hipLaunchKernelGGL((k_sbs_batch<16, SBS_BATCH_BLOCK, false>), g, b, 0, stream, L); break;
    case 33: hipLaunchKernelGGL((k_sbs_batch<16, SBS_BATCH_BLOCK, true>), g, b, 0, stream, L); break;
    case 64: hipLaunchKernelGGL((k_sbs_batch<32, SBS_BATCH_BLOCK, false>), g, b, 0, stream, L); break;
    case 65: hipLaunchKernelGGL((k_sbs_batch<32, SBS_BATCH_BLOCK, true>), g, b, 0, stream, L); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sbs_diff(const SceneView& s, const SbsDiffArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  SceneView v = s;
  v.sbs_data_f = a.params;   // distances come from the caller's parameter array
  if (a.colour && a.target) return launch_mode<3>(v, a, stream);
  if (a.colour) return launch_mode<0>(v, a, stream);
  if (a.dL_dcolour) return launch_mode<1>(v, a, stream);
  return launch_mode<2>(v, a, stream);
}

}  // namespace lrt
