// Voxel occupancy masks of sparse brick sets: for every brick one summary word and one bit per
// voxel, set when the minimum of the voxel's 8 corner values is <= 0 — exactly the voxels where
// the brick march (BVH2Common.cpp:884-1112) calls LocalSurfaceIntersection. The march then skips
// the value loads of empty voxels and whole empty bricks; its results do not change.
#include <algorithm>

#include "kernels.h"
#include "trace_device.h"

namespace lrt {

namespace {

// One thread per mask bit, 32 consecutive lanes per word: every voxel's 8 corner loads are in
// flight at once, and one wave ballot yields two words.
__global__ void k_sbs_mask_words(SceneView s, uint32_t sdf_id, uint32_t n_bricks, uint32_t* __restrict__ mask) {
  const SdfSBSHeader h = s.sbs_headers[sdf_id];
  const uint32_t bs = h.brick_size, v3 = bs * bs * bs, wpb = (v3 + 31u) / 32u;
  const uint64_t tid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t word_id = tid >> 5;
  const bool in = word_id < uint64_t(n_bricks) * wpb;
  const uint32_t brick = in ? uint32_t(word_id / wpb) : 0u, w = in ? uint32_t(word_id % wpb) : 0u;
  const uint32_t v = w * 32u + uint32_t(tid & 31u);
  const bool bit = in && v < v3 && sbs_voxel_occupied(s, h, s.sbs_roots[sdf_id], brick, v);
  const uint64_t m = __ballot(bit);   // every lane of the wave reaches this
  if (in && (threadIdx.x & 31u) == 0u) mask[size_t(brick) * (1u + wpb) + 1u + w] = uint32_t(m >> (threadIdx.x & 32u));
}

// One wave per brick (grid-strided): sbs_mask_brick writes the brick's bits and summary word.
__global__ __launch_bounds__(64) void k_sbs_mask_bricks(SceneView s, uint32_t sdf_id, uint32_t n_bricks,
                                                        uint32_t* __restrict__ mask) {
  extern __shared__ uint32_t flags[];
  const SdfSBSHeader h = s.sbs_headers[sdf_id];
  const uint32_t bs = h.brick_size, wpb = (bs * bs * bs + 31u) / 32u;
  const uint32_t root = s.sbs_roots[sdf_id];
  for (uint32_t b = blockIdx.x; b < n_bricks; b += gridDim.x)
    sbs_mask_brick(s, h, root, b, mask + size_t(b) * (1u + wpb), flags);
}

__global__ void k_sbs_mask_summary(uint32_t n_bricks, uint32_t wpb, uint32_t* __restrict__ mask) {
  const uint32_t brick = blockIdx.x * blockDim.x + threadIdx.x;
  if (brick >= n_bricks) return;
  uint32_t any = 0;
  const uint32_t* m = mask + size_t(brick) * (1u + wpb);
  for (uint32_t w = 0; w < wpb; ++w) any |= m[1 + w];
  mask[size_t(brick) * (1u + wpb)] = any != 0u ? 1u : 0u;
}

}  // namespace

hipError_t launch_sbs_masks(const SceneView& s, const SbsMaskGeom* geoms, uint32_t n_geoms, uint32_t* mask,
                            hipStream_t stream) {
  for (uint32_t g = 0; g < n_geoms; ++g) {
    const SbsMaskGeom& m = geoms[g];
    if (m.n_bricks == 0) continue;
    const uint32_t wpb = (m.brick_size * m.brick_size * m.brick_size + 31u) / 32u;
    const size_t flag_bytes = size_t(sbs_mask_flag_words(m.brick_size)) * 4u;
    if (flag_bytes <= SBS_MASK_LDS_MAX) {
      hipLaunchKernelGGL(k_sbs_mask_bricks, dim3(std::min<uint32_t>(m.n_bricks, 16384u)), dim3(64), flag_bytes, stream,
                         s, m.sdf_id, m.n_bricks, mask + m.base);
      continue;
    }
    // very large bricks: one thread per voxel bit, then the summary words
    const uint64_t bits = uint64_t(m.n_bricks) * wpb * 32u;
    hipLaunchKernelGGL(k_sbs_mask_words, dim3(uint32_t((bits + 255) / 256)), dim3(256), 0, stream, s, m.sdf_id,
                       m.n_bricks, mask + m.base);
    hipLaunchKernelGGL(k_sbs_mask_summary, dim3((m.n_bricks + 255) / 256), dim3(256), 0, stream, m.n_bricks, wpb,
                       mask + m.base);
  }
  return hipGetLastError();
}

}  // namespace lrt
