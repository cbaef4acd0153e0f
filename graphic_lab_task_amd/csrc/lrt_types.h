// Byte-compatible restatements of the reference ray-query data types, shared by
// host assembly code and the gfx950 kernels. Every struct keeps the reference's
// field order and size so that arrays built here can be handed to reference code
// (and vice versa) without conversion.
#pragma once
#include <cstdint>
#include <cstddef>

#if defined(__HIPCC__)
#define LRT_HD __host__ __device__ __forceinline__
#else
#define LRT_HD inline
#endif

namespace lrt {

struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct alignas(16) float4 { float x, y, z, w; };
struct uint2 { uint32_t x, y; };
struct alignas(16) int4 { int32_t x, y, z, w; };

// Column-major 4x4 matrix, m[c*4 + r] (LiteMath float4x4, LiteMath.h:1931).
struct alignas(16) float4x4 { float m[16]; };

// CrossRT/CrossRT.h:26-33
struct CRT_Hit {
  float    t;
  uint32_t primId;
  uint32_t instId;
  uint32_t geomId;   // (type << 24) | geomId
  float    coords[4];
};

// utils/raytracing/cbvh.h:21-27
struct BVHNode {
  float3   boxMin;
  uint32_t leftOffset;   // LEAF_BIT set -> leaf, START/SIZE packed
  float3   boxMax;
  uint32_t escapeIndex;
};

// utils/raytracing/cbvh.h:29-33 (fat node: both children's boxes)
struct alignas(16) BVHNodePair {
  BVHNode left;
  BVHNode right;
};

// utils/raytracing/intersection_common.h:33-44
struct GeomData {
  float4   boxMin;
  float4   boxMax;
  uint32_t bvhOffset;
  uint32_t type;
  uint2    offset;
  uint32_t hasChannels;
  uint32_t _pad[3];
};

// utils/raytracing/intersection_common.h:47-56
struct InstanceData {
  float4   boxMin;
  float4   boxMax;
  uint32_t geomId;
  uint32_t _pad[7];
  float4x4 transform;
  float4x4 transformInv;
  float4x4 transformInvTransposed;
};

// sdf/simple/sdf_simple_shared.h:23-27
struct SdfFrameOctreeNode {
  float    values[8];
  uint32_t offset;     // children block offset, 0 = leaf
};
// SdfFrameOctreeTexNode (sdf/simple/sdf_simple_shared.h:57-63): a frame-octree node with a
// texture coordinate pair (u, v) per corner, corner j = (x << 2 | y << 1 | z) at [2j, 2j+1], 104 B
struct SdfFrameOctreeTexNode {
  float    tex_coords[16];
  float    values[8];
  uint32_t offset;       // children block offset, 0 = leaf
  uint32_t material_id;  // read by the material resolve only
};
static_assert(sizeof(SdfFrameOctreeTexNode) == 104, "SdfFrameOctreeTexNode is 104 B");
// SdfSVSNode (sdf/simple/sdf_simple_shared.h:30-35): one voxel of a sparse voxel set, 16 B
struct SdfSVSNode {
  uint32_t pos_xy;           // x << 16 | y, position of the voxel in its level
  uint32_t pos_z_lod_size;   // z << 16 | level size (2^level)
  uint32_t values[2];        // 8 distance values, 1 byte each
};
static_assert(sizeof(SdfSVSNode) == 16, "SdfSVSNode layout");

// GraphicsPrimHeader (utils/misc/graphics_primitive_common.h:19-23) and GraphicsPrimType (:7-16):
// points carry (centre, radius); lines, segments, arrows and boxes two float4 (the first with
// the radius); the *_COLOR kinds one more float4 with an RGB colour per primitive.
struct GraphicsPrimHeader {
  float    color[3];
  uint32_t prim_type;
};
constexpr uint32_t GRAPH_PRIM_POINT = 0, GRAPH_PRIM_LINE = 1, GRAPH_PRIM_LINE_SEGMENT = 2,
                   GRAPH_PRIM_LINE_SEGMENT_DIR = 3, GRAPH_PRIM_BOX = 4, GRAPH_PRIM_POINT_COLOR = 5,
                   GRAPH_PRIM_LINE_COLOR = 6, GRAPH_PRIM_LINE_SEGMENT_COLOR = 7,
                   GRAPH_PRIM_LINE_SEGMENT_DIR_COLOR = 8, GRAPH_PRIM_BOX_COLOR = 9;
// float4 records per primitive (IntersectGraphicPrims, BVH2Common.cpp:1236-1240)
LRT_HD uint32_t graphics_prim_len(uint32_t prim_type) {
  uint32_t len = 2u;
  if (prim_type == GRAPH_PRIM_POINT) len = 1u;
  if (prim_type >= GRAPH_PRIM_LINE_COLOR) len = 3u;
  return len;
}

// COctreeV3Header (sdf/scom/scom_shared.h:24-44): 15 words, the layout of the file header and
// of the ABI. The last six words follow from node_pack_mode (scom.h:27-61).
struct COctreeV3Header {
  uint32_t brick_size;       // voxels per brick edge
  uint32_t brick_pad;        // extra voxels stored on each side of a brick
  uint32_t bits_per_value;   // quantised distance width
  uint32_t uv_size;          // 0: untextured, 1: 8 packed 16-bit (u, v) pairs per brick
  uint32_t sim_compression;  // 1: leaves carry an additive distance shift
  uint32_t lods;             // 1: every node carries a coarse leaf of its own
  uint32_t default_leaf_type, fallback_leaf_type;
  uint32_t node_pack_mode;
  uint32_t uints_per_child_info, idx_mask, idx_sh, trans_off, rot_mask, add_mask;
};
static_assert(sizeof(COctreeV3Header) == 60, "COctreeV3Header layout");
constexpr uint32_t COCTREE_V3_VERSION = 5;   // scom.h:88
constexpr uint32_t COCTREE_NODE_PACK_MODE_DEFAULT = 0, COCTREE_NODE_PACK_MODE_SIM_COMP_FULL = 1,
                   COCTREE_NODE_PACK_MODE_SIM_COMP_SMALL = 2;
constexpr uint32_t COCTREE_LEAF_TYPE_NOT_A_LEAF = 0, COCTREE_LEAF_TYPE_GRID = 1, COCTREE_LEAF_TYPE_BIT_PACK = 2,
                   COCTREE_LEAF_TYPE_SLICES = 3;
constexpr uint32_t COCTREE_LEAF_TYPE_BITS = 2, COCTREE_LEAF_TYPE_MASK = 0x3, COCTREE_LOD_LEAF_TYPE_SHIFT = 30;
constexpr uint32_t COCTREE_ROT_COUNT = 48;   // axis permutations with signs (BVH2Common.h:286)

// One COctreeV3 geometry as the kernels read it (ours): the header fields the walk uses, where
// its words start in the scene's COctreeV3 word array and where its 3 x 48 index modifiers start
// (rotations of the value lattice, of the voxel lattice and of the 2^3 child lattice).
struct CoctGeom {
  uint32_t base, mods, size;
  uint32_t brick_size, brick_pad, bits_per_value, uv_size, sim_compression, lods;
  uint32_t uints_per_child_info, idx_mask, idx_sh, trans_off, rot_mask, add_mask;
  uint32_t _pad;
};
static_assert(sizeof(CoctGeom) == 64, "CoctGeom is 64 B");

// SdfDAG (sdf/scom2/scom.h:56-67, scom_shared.h:77-119): octree nodes whose children and leaf
// bricks are shared through edges carrying one of the 48 lattice rotations (and an additive
// shift for bricks).
struct SdfDAGHeader {
  uint32_t brick_size, brick_pad, node_grid_size, dim, transform_subgroup;
  float    user_params[7];
  int32_t  tex_id_off, mat_id_off, all_float_tex_id_off, all_int_mat_id_off;
};
struct SdfDAGChildEdge {
  uint32_t child_offset;   // node index, 0: no child
  uint32_t rotation_id;
};
struct SdfDAGDataEdge {
  uint32_t data_offset;    // first distance of the brick
  uint32_t rotation_id;
  uint32_t type_id;
  float    add;
};
struct SdfDAGNode {
  uint32_t children_edges_offset;   // first of node_grid_size^3 child edges, 0: a leaf
  uint32_t data_edges_offset;       // first data edge, 0: no surface
  SdfDAGChildEdge channels_edge;
  uint32_t voxel_count_flags;       // bits 0-15 data edges, bit 16 surface node, bit 17 inside
  uint32_t _pad;
};
static_assert(sizeof(SdfDAGHeader) == 64 && sizeof(SdfDAGNode) == 24 && sizeof(SdfDAGChildEdge) == 8 &&
              sizeof(SdfDAGDataEdge) == 16, "SdfDAG layouts");
// SCom2Header (sdf/scom2/scom_shared.h:21-75): 54 words describing how an SCom2 tree packs its
// node heads, links and bricks into one u32 array.
struct SCom2Header {
  uint32_t brick_size, v_size, bits_per_value, values_per_uint, value_mask, bitmask_len, dimension;
  uint32_t child_rot_shift, child_rot_mask, child_add_shift, child_add_mask, child_offset_mask, child_offset_off;
  uint32_t node_offset_mask, uints_per_link, unique_brick_prefix, unique_brick_offset_mask;
  uint32_t children_types_shift, children_types_mask, base_reference_shift, children_active_bits_shift;
  uint32_t children_active_bits_mask, references_offset, reference_bits, reference_mask, references_per_uint;
  uint32_t links_offset, max_surface_count, max_surface_count_per_leaf;
  uint32_t bricks_step, bricks_arr_offset, nodes_arr_offset, root_node_off;
  uint32_t has_channels, has_surfaces, has_multi_nodes;
  int32_t  tex_id_off, mat_id_off, all_float_tex_id_off, all_int_mat_id_off;
  float    max_val;
  uint32_t max_depth;
  float    user_params[7];
  uint32_t _pad0, _pad1, _pad2, _pad3, _pad4;
};
static_assert(sizeof(SCom2Header) == 216, "SCom2Header layout");
constexpr uint32_t SCOM2_CHILD_EMPTY = 0, SCOM2_CHILD_LEAF_VOLUME = 1, SCOM2_CHILD_NODE = 2,
                   SCOM2_CHILD_LEAF_SURFACE = 3, SCOM2_CHILD_TYPE_BITS = 2, SCOM2_CHILD_TYPE_MASK = 3;
constexpr uint32_t SCOM2_MAGIC_NUMBER = 0xffffdefau, SCOM2_VERSION = 4;

// One DAG geometry as the kernels read it (ours): its root node, brick shape, node grid size and
// the first of its 2 x 48 index modifiers (brick value lattice, node grid).
struct DagGeom {
  uint32_t root, brick_size, brick_pad, node_grid_size, mods, _pad[3];
};

// sdf/simple/sdf_simple_shared.h:39-45
struct SdfSBSNode {
  uint32_t pos_xy;          // (x << 16) | y, brick position in its LOD grid
  uint32_t pos_z_lod_size;  // (z << 16) | lod_size
  uint32_t data_offset;     // offset (in u32) of the brick's data block
  uint32_t _pad;
};

// sdf/simple/sdf_simple_shared.h:49-55
struct SdfSBSHeader {
  uint32_t brick_size;
  uint32_t brick_pad;
  uint32_t bytes_per_value;
  uint32_t aux_data;        // node layout (top 8 bits)
};

// sdf/multi/sdf_multi_shared.h:8-14
struct SdfMultiOctreeNode {
  uint32_t data_offset;
  uint32_t children_offset;
  uint32_t voxel_count;
  uint32_t flags;
};

// core/render_settings.h:116-136
struct MultiRenderPreset {
  uint32_t render_mode;
  uint32_t sdf_node_intersect;
  uint32_t normal_mode;
  uint32_t ray_gen_mode;
  uint32_t interpolation_mode;
  uint32_t representation_mode;
  uint32_t spp;
  uint32_t fixed_lod;
  float    level_of_detail;
  float    compact_sdf_eps;
  float    LOD_distance;
  uint32_t channel_render_mode;
  uint32_t channel_id;
  uint32_t component_id;
  uint32_t CTF_id;
};

static_assert(sizeof(CRT_Hit) == 32, "CRT_Hit is 32 B");
static_assert(sizeof(BVHNode) == 32, "BVHNode is 32 B");
static_assert(sizeof(BVHNodePair) == 64, "BVHNodePair is 64 B");
static_assert(sizeof(GeomData) == 64, "GeomData is 64 B");
static_assert(sizeof(InstanceData) == 256, "InstanceData is 256 B");
static_assert(sizeof(SdfFrameOctreeNode) == 36, "SdfFrameOctreeNode is 36 B");
static_assert(sizeof(SdfSBSNode) == 16, "SdfSBSNode is 16 B");
static_assert(sizeof(SdfSBSHeader) == 16, "SdfSBSHeader is 16 B");
static_assert(sizeof(SdfMultiOctreeNode) == 16, "SdfMultiOctreeNode is 16 B");
static_assert(sizeof(MultiRenderPreset) == 60, "MultiRenderPreset is 60 B");
static_assert(offsetof(InstanceData, transform) == 64, "InstanceData layout");
static_assert(offsetof(InstanceData, transformInv) == 128, "InstanceData layout");
static_assert(offsetof(InstanceData, transformInvTransposed) == 192, "InstanceData layout");

// ---- constants (utils/raytracing/raytrace_common.h:35-57, cbvh.h:17-19) ----
constexpr int      STACK_SIZE  = 80;
constexpr uint32_t START_MASK  = 0x0FFFFFFFu;
constexpr uint32_t SIZE_MASK   = 0x70000000u;
constexpr uint32_t LEAF_BIT    = 0x80000000u;
constexpr uint32_t LEAF_NORMAL = 0xFFFFFFFFu;
constexpr uint32_t LEAF_EMPTY  = 0xFFFFFFFDu;
constexpr uint32_t ESCAPE_ROOT = 0xFFFFFFFEu;

// ---- geometry types and tags (core/render_settings.h:3-42, BVH2Common.h:30-43) ----
constexpr uint32_t TYPE_MESH_TRIANGLE    = 0;
constexpr uint32_t TYPE_SDF_GRID         = 1;
constexpr uint32_t TYPE_GRAPHICS_PRIM    = 17;
constexpr uint32_t TYPE_SDF_FRAME_OCTREE = 3;
constexpr uint32_t TYPE_SDF_FRAME_OCTREE_TEX = 8;
constexpr uint32_t TYPE_SDF_SVS          = 5;
constexpr uint32_t TYPE_SDF_SBS          = 6;
constexpr uint32_t TYPE_SDF_MULTI_OCTREE = 9;
constexpr uint32_t TYPE_COCTREE_V3       = 20;
constexpr uint32_t TYPE_SDF_DAG          = 25;
constexpr uint32_t TYPE_SCOM2            = 26;
constexpr uint32_t MULTI_OCTREE_FLAG_SURFACE = 1;          // sdf_multi_shared.h:5
constexpr uint32_t MULTI_RENDER_MODE_MO_VOXEL_COUNT = 24;  // render_settings.h:85-86
constexpr uint32_t MULTI_RENDER_MODE_MO_FLAGS = 25;
constexpr uint32_t TYPE_SDF_SBS_TEX      = 10;
constexpr uint32_t TYPE_SDF_SBS_COL      = 11;
constexpr uint32_t SH_TYPE      = 24;
constexpr uint32_t GEOM_ID_MASK = 0x00FFFFFFu;

constexpr uint32_t TAG_NONE       = 0;
constexpr uint32_t TAG_TRIANGLE   = 1;
constexpr uint32_t TAG_COCTREE_BRICKED = 2;
constexpr uint32_t TAG_SDF_BRICK  = 4;
constexpr uint32_t TAG_SDF_NODE   = 8;
constexpr uint32_t TAG_SDF_GRID   = 11;
constexpr uint32_t TAG_GRAPHICS_PRIM = 9;
constexpr uint32_t TAG_SDF_OCTREE = 15;
constexpr uint32_t TAG_SDF_DAG    = 18;
constexpr uint32_t TAG_SCOM2      = 19;

// SBS node layouts (sdf/simple/sdf_simple_shared.h:13-21)
constexpr uint32_t SDF_SBS_NODE_LAYOUT_UNDEFINED        = 0u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_DX               = 1u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_DX_UV16          = 2u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_DX_RGB8          = 3u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F    = 4u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN = 5u << 24;
constexpr uint32_t SDF_SBS_NODE_LAYOUT_MASK             = 0xFF000000u;

constexpr uint32_t INVALID_IDX = 1u << 31;   // sdf/common/sdf_common_shared.h:30

// preset enums (core/render_settings.h:53-106)
constexpr uint32_t SDF_OCTREE_NODE_INTERSECT_ST       = 0;
constexpr uint32_t SDF_OCTREE_NODE_INTERSECT_ANALYTIC = 1;
constexpr uint32_t SDF_OCTREE_NODE_INTERSECT_NEWTON   = 2;
constexpr uint32_t SDF_OCTREE_NODE_INTERSECT_BBOX     = 3;
constexpr uint32_t SDF_OCTREE_NODE_INTERSECT_IT       = 4;
constexpr uint32_t MULTI_RENDER_MODE_MASK             = 0;
constexpr uint32_t MULTI_RENDER_MODE_LAMBERT_NO_TEX   = 1;
constexpr uint32_t MULTI_RENDER_MODE_NORMAL           = 8;
constexpr uint32_t MULTI_RENDER_MODE_ST_ITERATIONS    = 10;
constexpr uint32_t MULTI_RENDER_MODE_PBR_NO_TEX       = 11;
constexpr uint32_t MULTI_RENDER_MODE_PHONG_NO_TEX     = 12;
constexpr uint32_t MULTI_RENDER_MODE_LAMBERT          = 17;
constexpr uint32_t MULTI_RENDER_MODE_PHONG            = 18;
constexpr uint32_t MULTI_RENDER_MODE_LOD              = 20;
constexpr uint32_t MULTI_RENDER_MODE_GEOM             = 7;
constexpr uint32_t MULTI_RENDER_MODE_PBR              = 21;
constexpr uint32_t MULTI_RENDER_MODE_CHANNEL          = 23;
constexpr uint32_t MULTI_RENDER_MODE_INTERSECTION_COUNT = 26;
constexpr uint32_t MULTI_RENDER_MODE_CHANNEL_LAMBERT  = 27;
constexpr uint32_t MULTI_RENDER_MODE_DEPTH            = 2;
constexpr uint32_t MULTI_RENDER_MODE_LINEAR_DEPTH     = 3;
constexpr uint32_t MULTI_RENDER_MODE_INVERSE_LINEAR_DEPTH = 4;
constexpr uint32_t MULTI_RENDER_MODE_PRIMITIVE        = 5;
constexpr uint32_t MULTI_RENDER_MODE_TYPE             = 6;
constexpr uint32_t MULTI_RENDER_MODE_BARYCENTRIC      = 9;
constexpr uint32_t MULTI_RENDER_MODE_TEX_COORDS       = 15;
constexpr uint32_t MULTI_RENDER_MODE_HSV_DEPTH        = 19;
// Light types (core/IRenderer.h:16-18)
constexpr uint32_t LIGHT_TYPE_DIRECT  = 0;
constexpr uint32_t LIGHT_TYPE_POINT   = 1;
constexpr uint32_t LIGHT_TYPE_AMBIENT = 2;
constexpr uint32_t NORMAL_MODE_GEOMETRY     = 0;
constexpr uint32_t NORMAL_MODE_VERTEX       = 1;
constexpr uint32_t NORMAL_MODE_SDF_SMOOTHED = 2;
constexpr uint32_t INTERPOLATION_MODE_TRILINEAR = 0;
constexpr uint32_t INTERPOLATION_MODE_TRICUBIC  = 1;
constexpr uint32_t REPRESENTATION_MODE_SURFACE  = 0;
constexpr uint32_t REPRESENTATION_MODE_VOLUME   = 1;

// core/render_settings.h:138-160
inline MultiRenderPreset default_preset() {
  MultiRenderPreset p{};
  p.render_mode = MULTI_RENDER_MODE_LAMBERT_NO_TEX;
  p.sdf_node_intersect = SDF_OCTREE_NODE_INTERSECT_ANALYTIC;
  p.normal_mode = NORMAL_MODE_VERTEX;
  p.ray_gen_mode = 0;
  p.interpolation_mode = INTERPOLATION_MODE_TRILINEAR;
  p.representation_mode = REPRESENTATION_MODE_VOLUME;
  p.spp = 1;
  p.fixed_lod = 1;
  p.level_of_detail = 16.0f;
  p.compact_sdf_eps = 0.05f;
  p.LOD_distance = 10.0f;
  return p;
}

// MultiRenderer/BVH2Common.cpp:19-32
LRT_HD bool preset_needs_normal(uint32_t m) {
  return m == MULTI_RENDER_MODE_LAMBERT_NO_TEX || m == MULTI_RENDER_MODE_NORMAL ||
         m == MULTI_RENDER_MODE_PHONG_NO_TEX || m == MULTI_RENDER_MODE_LAMBERT ||
         m == MULTI_RENDER_MODE_LOD || m == MULTI_RENDER_MODE_GEOM ||
         m == MULTI_RENDER_MODE_PHONG || m == MULTI_RENDER_MODE_CHANNEL ||
         m == MULTI_RENDER_MODE_PBR || m == MULTI_RENDER_MODE_PBR_NO_TEX ||
         m == MULTI_RENDER_MODE_CHANNEL_LAMBERT;
}

LRT_HD uint32_t pack_offset_and_size(uint32_t start, uint32_t size) {
  return LEAF_BIT | ((size << 28) & SIZE_MASK) | (start & START_MASK);
}
LRT_HD uint32_t extract_start(uint32_t v) { return v & START_MASK; }
LRT_HD uint32_t extract_count(uint32_t v) { return (v & SIZE_MASK) >> 28; }

}  // namespace lrt
