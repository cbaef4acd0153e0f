// Host-side BVH construction that emits the reference's BLAS/TLAS layout contract.
#pragma once
#include "lrt_types.h"
#include <vector>

namespace lrt {

// Fat BLAS in the reference layout: pair 0 holds the root's two children
// (utils/raytracing/cbvh.cpp:143-169 AlignNodes2, cbvh_fat.cpp:128-142 CreateFatTreeArray).
struct FatBlas {
  std::vector<BVHNodePair> nodes;
  std::vector<uint32_t>    indices;     // primitive ids in leaf order (indicesReordered)
  std::vector<uint32_t>    startCount;  // per leaf: PackOffsetAndSize(start, count) (triangles only)
};

struct BuildStats {
  uint32_t max_depth = 0;
  uint32_t leaves = 0;
  uint32_t merged_leaves = 0;
};

// BLAS over triangles (BuildBVHFat, utils/raytracing/cbvh_fat.cpp:150-166 and
// BuildBVH, cbvh.cpp:430-539). Vertices are float4 (w ignored).
FatBlas build_blas_triangles(const float4* verts, size_t n_verts, const uint32_t* tri_idx, size_t n_idx,
                             int prims_in_leaf, BuildStats* stats = nullptr);

// BLAS over custom boxes, one box per leaf (BuildBVHFatCustom, cbvh_fat.cpp:168-183;
// BuildBVH(const BVHNode*), cbvh.cpp:597-664). Leaf leftOffset = PackOffsetAndSize(boxId, 1).
FatBlas build_blas_boxes(const BVHNode* boxes, size_t n, BuildStats* stats = nullptr);

// TLAS over instance boxes (BVHRT::CommitScene, MultiRenderer/BVH2Common_host.cpp:1230-1251):
// stackless BVH2_LEFT_OFFSET nodes with escape indices; a single instance gives one leaf node.
std::vector<BVHNode> build_tlas(const float4* box_min, const float4* box_max, size_t n);

// The same two builds on the current HIP device (bvh_kernels.hip): a linear BVH over 63-bit
// Morton codes of the primitive centroids with the leaf-merge rule and layout contract above.
// Small inputs (one leaf) take the host path. Throws std::runtime_error on a HIP failure.
FatBlas gpu_build_blas_triangles(const float4* verts, size_t n_verts, const uint32_t* tri_idx, size_t n_idx,
                                 int prims_in_leaf, BuildStats* stats = nullptr);
FatBlas gpu_build_blas_boxes(const BVHNode* boxes, size_t n, BuildStats* stats = nullptr);

}  // namespace lrt
