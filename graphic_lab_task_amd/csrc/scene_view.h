// Plain device-pointer view of a committed scene, passed to kernels by value.
#pragma once
#include "lrt_types.h"

namespace lrt {

// Leaf kinds present in a scene; selects the kernel instantiation so that a
// triangle-only scene does not pay registers for the SDF intersectors.
enum LeafKind : uint32_t {
  LK_TRI   = 1u << 0,
  LK_FRAME = 1u << 1,   // TAG_SDF_NODE over a frame octree
  LK_SBS   = 1u << 2,   // TAG_SDF_BRICK
  LK_MULTI = 1u << 3,   // TAG_SDF_OCTREE (multi-octree walk)
};

struct SceneView {
  // BVH
  const BVHNodePair*  nodes;        // m_allNodePairs
  const BVHNode*      tlas;         // m_nodesTLAS
  const InstanceData* instances;    // m_instanceData
  const GeomData*     geoms;        // m_geomData
  const uint32_t*     geom_tags;    // m_geomTags
  const uint2*        start_end;    // startEnd
  const uint32_t*     prim_id_count;// m_primIdCount
  // meshes
  const float4*       vert_pos;     // m_vertPos
  const float4*       vert_norm;    // m_vertNorm
  const uint32_t*     indices;      // m_indices (leaf order)
  const uint32_t*     prim_indices; // m_primIndices
  // frame octree
  const BVHNode*            orig_nodes;   // m_origNodes
  const SdfFrameOctreeNode* frame_nodes;  // m_SdfFrameOctreeNodes
  const uint32_t*           frame_roots;  // m_SdfFrameOctreeRoots
  // sparse brick sets
  const SdfSBSNode*   sbs_nodes;    // m_SdfSBSNodes
  const uint32_t*     sbs_data;     // m_SdfSBSData
  const float*        sbs_data_f;   // m_SdfSBSDataF
  const uint32_t*     sbs_roots;    // m_SdfSBSRoots
  const SdfSBSHeader* sbs_headers;  // m_SdfSBSHeaders
  // multi octree
  const SdfMultiOctreeNode* multi_nodes;  // m_SdfMultiOctreeNodes
  const float*              multi_values; // m_SdfMultiOctreeValues

  MultiRenderPreset preset;
  float    lod_thr;                 // m_precomputedLoDThr
  uint32_t leaf_kinds;              // LeafKind mask
  uint32_t n_instances;
};

}  // namespace lrt
