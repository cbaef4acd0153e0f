// Workload generators and SDF structure builders (host side).
#pragma once
#include <cstdint>
#include <vector>

#include "lrt_types.h"

namespace lrt {

double value_noise3(double x, double y, double z, uint64_t seed);
void make_icosphere(uint32_t subdiv, std::vector<double>& v, std::vector<uint32_t>& f);
void make_geodesic(uint32_t n, std::vector<double>& v, std::vector<uint32_t>& f);
void displace_radial(std::vector<double>& v, double amplitude, double freq, uint64_t seed);

// Union of spheres: f(p) = min_i(|p - c_i| - r_i); spheres are (cx, cy, cz, r).
struct SphereUnion {
  const float* spheres;
  size_t n;
  float operator()(float3 p) const;
};

// GlobalOctree (sdf/common/global_octree.h:113-199), the subset the converters need.
struct GONode {
  uint8_t  type;        // 0 EMPTY, 1 LEAF, 2 EMPTY_NODE, 3 NODE
  uint32_t val_off;     // offset into values (v_size^3 values per node)
  uint32_t offset;      // children block, 0 = leaf
};
struct GlobalOctree {
  uint32_t brick_size = 1, brick_pad = 0;
  std::vector<GONode> nodes;
  std::vector<float> values;
};

// sdf_converter::sdf_to_global_octree (sdf/build/sparse_octree_builder.cpp:106-264)
void sdf_to_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                          int threads, GlobalOctree& out);
// The same build on the current HIP device (sdf_build_kernels.hip): level-synchronous growth of
// the border subtrees, depth-first numbering recovered from subtree sizes; equal to the host
// build byte for byte. Throws std::runtime_error on a HIP failure.
void gpu_sdf_to_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                              GlobalOctree& out);
// host (threads >= 0; 0 = all cores) or GPU (threads < 0) build
void build_global_octree(const SphereUnion& sdf, uint32_t depth, uint32_t brick_size, uint32_t brick_pad,
                         int threads, GlobalOctree& out);
// global_octree_to_frame_octree (sdf/simple/sdf_converter.cpp:53-65)
void global_octree_to_frame_octree(const GlobalOctree& g, std::vector<SdfFrameOctreeNode>& out);
// global_octree_to_multi_octree (sdf/multi/sdf_multi.cpp:124-180): one node per octree node,
// voxel_count 1 for active nodes (the builder's bricks_count), 0 for empty ones; values as is.
void global_octree_to_multi_octree(const GlobalOctree& g, std::vector<SdfMultiOctreeNode>& nodes,
                                   std::vector<float>& values);
// global_octree_to_SBS (sdf/simple/sdf_converter.cpp:156-225), quantised DX layout
// colour_seed != 0 fills the per-corner words of the DX_UV16 / DX_RGB8 layouts from a seeded field.
void global_octree_to_sbs_dx(const GlobalOctree& g, const SdfSBSHeader& header, uint64_t colour_seed,
                             std::vector<SdfSBSNode>& nodes, std::vector<uint32_t>& values);
// Indexed fp32 layout (ID32F_IRGB32F): shared lattice points stored once in values_f, addressed by
// v_size^3 indices per brick, followed by 8 colour indices per brick (brick corners, 3 floats each).
// values_f = [N distances][N albedo RGB triples]; the albedo is a seeded smooth colour field.
void global_octree_to_sbs_indexed(const GlobalOctree& g, uint64_t albedo_seed, std::vector<SdfSBSNode>& nodes,
                                  std::vector<uint32_t>& values, std::vector<float>& values_f,
                                  uint32_t* n_distance_values);

}  // namespace lrt
