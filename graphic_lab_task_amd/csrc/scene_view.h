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
  // Walk option, not a leaf kind (never in SceneView::leaf_kinds): triangle leaves are
  // postponed like SDF leaves (traverse.h).
  LK_TRI_POSTPONE = 1u << 4,
  // Walk option set in SceneView::leaf_kinds when the preset asks for smoothed normals and a
  // brick set has neighbour indices (ID32F_IRGB32F_IN): selects the kernels that compile the
  // smoothing in (the others stay free of its registers).
  LK_SBS_SMOOTH = 1u << 5,
  LK_GRID  = 1u << 6,   // TAG_SDF_GRID (dense SDF grid, sphere tracing)
  LK_GPRIM = 1u << 7,   // TAG_GRAPHICS_PRIM (points, lines, segments, arrows, wire boxes)
  LK_COCT  = 1u << 8,   // TAG_COCTREE_BRICKED (compact octree COctreeV3)
  LK_DAG   = 1u << 9,   // TAG_SDF_DAG (octree DAG with rotated shared nodes and bricks)
  LK_SCOM2 = 1u << 10,  // TAG_SCOM2 (packed SCom2 tree)
  // Walk option set in SceneView::leaf_kinds when the preset interpolates tricubically and the
  // scene holds brick sets or grids: selects the kernels with the 4^3-sample interpolant.
  LK_TRICUBIC = 1u << 11,
  // TAG_SDF_NODE over a textured frame octree (always with LK_FRAME): its texture-coordinate step
  // is compiled only into the kernels that take this kind (the mixed ones), so the frame-octree
  // kernel keeps its registers.
  LK_FRAME_TEX = 1u << 12,
};

// Frame-octree leaf as the kernels read it: the leaf box, its 8 corner distances and the
// octree node index, gathered at commit into one 64-byte record per BLAS box. Leaves of textured
// frame octrees also name their 16 texture coordinates in frame_tex (4 float4 from 4 * tex).
struct alignas(16) FrameLeaf {
  float3   boxMin;
  float3   boxMax;
  float    values[8];
  uint32_t primId;
  uint32_t tex;
};
static_assert(sizeof(FrameLeaf) == 64, "FrameLeaf is 64 B");

// Number of BLAS pairs (top BFS levels of the cached geometry) the kernels keep in LDS.
constexpr int LDS_TOP_PAIRS = 63;

struct SceneView {
  // BVH, device layout: the reference's fat pairs with leaves holding PackOffsetAndSize(start,
  // count) directly and each geometry's pairs renumbered so its top BFS levels come first.
  const BVHNodePair*  nodes;
  const BVHNode*      tlas;         // m_nodesTLAS
  const InstanceData* instances;    // m_instanceData
  const GeomData*     geoms;        // m_geomData
  const uint32_t*     geom_tags;    // m_geomTags
  const uint2*        start_end;    // startEnd
  const uint32_t*     prim_id_count;// m_primIdCount
  // meshes
  const float4*       tri_soup;     // 3 float4 per leaf-ordered triangle: (A.xyz, idxA), (B.xyz, idxB), (C.xyz, idxC)
  const float4*       vert_pos;     // m_vertPos
  const float4*       vert_norm;    // m_vertNorm
  const uint32_t*     indices;      // m_indices (leaf order)
  const uint32_t*     prim_indices; // m_primIndices
  // frame octree
  const FrameLeaf*          frame_leaves; // per BLAS box, same index as m_origNodes
  const BVHNode*            orig_nodes;   // m_origNodes
  const SdfFrameOctreeNode* frame_nodes;  // m_SdfFrameOctreeNodes
  const uint32_t*           frame_roots;  // m_SdfFrameOctreeRoots
  const float4*             frame_tex;    // texture coordinates of textured frame-octree leaves, 4 float4 each
  // sparse brick sets
  const SdfSBSNode*   sbs_nodes;    // m_SdfSBSNodes
  const uint32_t*     sbs_data;     // m_SdfSBSData
  const float*        sbs_data_f;   // m_SdfSBSDataF
  const uint32_t*     sbs_roots;    // m_SdfSBSRoots
  const SdfSBSHeader* sbs_headers;  // m_SdfSBSHeaders
  // voxel occupancy (ours): per brick one summary word (non-zero if any voxel is occupied) and
  // ceil(bs^3/32) words with bit (x*bs+y)*bs+z set when min of the voxel's 8 corner values
  // <= 0, i.e. exactly the voxels the brick march intersects; null = march without it
  const uint32_t*     sbs_mask;
  const uint32_t*     sbs_mask_base;  // first mask word of each brick set (indexed like sbs_roots)
  // multi octree
  const SdfMultiOctreeNode* multi_nodes;  // m_SdfMultiOctreeNodes
  const float*              multi_values; // m_SdfMultiOctreeValues
  // dense SDF grids
  const float*              grid_data;    // m_SdfGridData
  const uint32_t*           grid_offsets; // m_SdfGridOffsets
  const uint32_t*           grid_sizes;   // m_SdfGridSizes, 3 words per grid
  // graphics primitives
  const GraphicsPrimHeader* gprim_headers; // m_GraphicsPrimHeaders
  const float4*             gprim_points;  // m_GraphicsPrimPoints
  // compact octrees (COctreeV3)
  const uint32_t*           coct_data;     // m_SdfCompactOctreeV3Data, every geometry's words in turn
  const CoctGeom*           coct_geoms;    // per COctreeV3 geometry (GeomData.offset.x)
  const int4*               coct_mods;     // m_SdfCompactOctreeRotModifiers, 3 x 48 per geometry
  // octree DAGs
  const SdfDAGNode*         dag_nodes;     // m_SdfDAGNodes
  const SdfDAGChildEdge*    dag_child_edges;
  const SdfDAGDataEdge*     dag_data_edges;
  const float*              dag_distances;
  const DagGeom*            dag_geoms;     // per DAG geometry (GeomData.offset.y)
  const int4*               dag_mods;      // 2 x 48 index modifiers per DAG geometry
  const uint8_t*            rot_add;       // m_RotAddTable: 48 x 48 rotation compositions
  // SCom2 trees
  const uint32_t*           scom2_data;    // m_SCom2Data: each tree's nodes then bricks
  const SCom2Header*        scom2_headers; // m_SCom2Headers (GeomData.offset.y), offsets made global
  const int4*               scom2_mods;    // 3 x 48 index modifiers per tree
  uint32_t                  scom2_fast;    // 1: SCom2 leaves take the fast traversal (support_fast_scom_traversal)

  MultiRenderPreset preset;
  float    lod_thr;                 // m_precomputedLoDThr
  uint32_t leaf_kinds;              // LeafKind mask
  uint32_t n_instances;
  uint32_t cache_geom;              // geometry whose top pairs are staged in LDS (~0u: none)
  uint32_t cache_pairs;             // number of staged pairs (<= LDS_TOP_PAIRS)
};

}  // namespace lrt
