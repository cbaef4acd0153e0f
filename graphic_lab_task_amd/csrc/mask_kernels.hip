// Voxel occupancy masks of sparse brick sets: for every brick one summary word and one bit per
// voxel, set when the minimum of the voxel's 8 corner values is <= 0 — exactly the voxels where
// the brick march (BVH2Common.cpp:884-1112) calls LocalSurfaceIntersection. The march then skips
// the value loads of empty voxels and whole empty bricks; its results do not change.
#include "kernels.h"
#include "trace_device.h"

namespace lrt {

namespace {

__global__ void k_sbs_mask_words(SceneView s, uint32_t sdf_id, uint32_t n_bricks, uint32_t* __restrict__ mask) {
  const SdfSBSHeader h = s.sbs_headers[sdf_id];
  const uint32_t bs = h.brick_size, v3 = bs * bs * bs, wpb = (v3 + 31u) / 32u;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n_bricks * wpb) return;
  const uint32_t brick = tid / wpb, w = tid % wpb;
  const uint32_t node_id = brick + s.sbs_roots[sdf_id];
  const uint32_t v_size = bs + 2 * h.brick_pad + 1;
  const float sz = float(s.sbs_nodes[node_id].pos_z_lod_size & 0x0000FFFF);
  const float sz_inv = 2.0f / sz;
  uint32_t word = 0;
  for (uint32_t b = 0; b < 32u; ++b) {
    const uint32_t v = w * 32u + b;
    if (v >= v3) break;
    const float3 vp = make3(float(v / (bs * bs)), float((v / bs) % bs), float(v % bs));
    float values[8];
    const float vmin = sbs_load_values(s, h, node_id, vp, v_size, sz_inv, values);
    if (vmin <= 0.0f) word |= 1u << b;
  }
  mask[size_t(brick) * (1u + wpb) + 1u + w] = word;
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
    const uint32_t words = m.n_bricks * wpb;
    hipLaunchKernelGGL(k_sbs_mask_words, dim3((words + 255) / 256), dim3(256), 0, stream, s, m.sdf_id, m.n_bricks,
                       mask + m.base);
    hipLaunchKernelGGL(k_sbs_mask_summary, dim3((m.n_bricks + 255) / 256), dim3(256), 0, stream, m.n_bricks, wpb,
                       mask + m.base);
  }
  return hipGetLastError();
}

}  // namespace lrt
