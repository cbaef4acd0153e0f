// ORACLE - test infrastructure only. A CPU restatement of LiteRT's ray-query path used to
// check the gfx950 library; it is never linked into or called by the product. Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
//
// Independent of the product sources: its own structs and arithmetic, written from the
// reference files cited per function. Parity status of every function is listed in the
// header comment of oracle/__init__.py.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---- LiteMath subset (dependencies/1st-party/LiteMath/LiteMath.h) ----
struct f2 { float x, y; };
struct f3 {
  float x, y, z;
  f3() : x(0), y(0), z(0) {}
  f3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct f4 { float x, y, z, w; };
inline f3 operator+(f3 a, f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, f3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f3 operator/(f3 a, f3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline f3 operator*(f3 a, float k) { return f3(a.x * k, a.y * k, a.z * k); }
inline f3 operator*(float k, f3 a) { return f3(k * a.x, k * a.y, k * a.z); }
inline f3 operator/(f3 a, float k) { return f3(a.x / k, a.y / k, a.z / k); }
inline f3 operator-(f3 a, float k) { return f3(a.x - k, a.y - k, a.z - k); }
inline f3 operator+(f3 a, float k) { return f3(a.x + k, a.y + k, a.z + k); }
// LiteMath.h:133-138
inline float lm_sign(float x) { if (x == 0.0f) return 0.0f; else if (x < 0.0f) return -1.0f; else return 1.0f; }
// LiteMath.h:963-966
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
inline f3 normalize(f3 a) { float lenInv = float(1) / length(a); return a * lenInv; }
// LiteMath.h:987-992 shuffle_yzx(a*b_yzx - a_yzx*b)
inline f3 cross(f3 a, f3 b) {
  const f3 a_yzx(a.y, a.z, a.x), b_yzx(b.y, b.z, b.x);
  const f3 r = a * b_yzx - a_yzx * b;
  return f3(r.y, r.z, r.x);
}
inline f3 to3(f4 v) { return f3(v.x, v.y, v.z); }
// LiteMath.h:1813-1819, column-major B
inline f4 mul(const float* B, f4 V) {
  f4 r;
  r.x = V.x * B[0] + V.y * B[4] + V.z * B[8] + V.w * B[12];
  r.y = V.x * B[1] + V.y * B[5] + V.z * B[9] + V.w * B[13];
  r.z = V.x * B[2] + V.y * B[6] + V.z * B[10] + V.w * B[14];
  r.w = V.x * B[3] + V.y * B[7] + V.z * B[11] + V.w * B[15];
  return r;
}
// raytrace_common.h:84-92
inline f3 matmul3x3(const float* m, f3 v) { return to3(mul(m, f4{v.x, v.y, v.z, 0.0f})); }
inline f3 matmul4x3(const float* m, f3 v) { return to3(mul(m, f4{v.x, v.y, v.z, 1.0f})); }

// raytrace_common.h:59-67
inline f3 SafeInverse(f3 d) {
  const float ooeps = 1.0e-36f;
  f3 res;
  res.x = 1.0f / (std::abs(d.x) > ooeps ? d.x : std::abs(ooeps) * lm_sign(d.x));
  res.y = 1.0f / (std::abs(d.y) > ooeps ? d.y : std::abs(ooeps) * lm_sign(d.y));
  res.z = 1.0f / (std::abs(d.z) > ooeps ? d.z : std::abs(ooeps) * lm_sign(d.z));
  return res;
}
// raytrace_common.h:69-82
inline f2 RayBoxIntersection2(f3 o, f3 inv, f3 bmin, f3 bmax) {
  const float lo = inv.x * (bmin.x - o.x);
  const float hi = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  const float tmin = std::max(std::min(lo, hi), std::min(lo1, hi1));
  const float tmax = std::min(std::max(lo, hi), std::max(lo1, hi1));
  return f2{std::max(tmin, std::min(lo2, hi2)), std::min(tmax, std::max(lo2, hi2))};
}
// intersection_common.h:96-101
inline f2 encode_normal(f3 v) {
  const float k = 1.0f / (std::abs(v.x) + std::abs(v.y) + std::abs(v.z));
  const f2 p{v.x * k, v.y * k};
  const f2 s{(p.x >= 0.0f) ? +1.0f : -1.0f, (p.y >= 0.0f) ? +1.0f : -1.0f};
  return (v.z <= 0.0f) ? f2{(1.0f - std::abs(p.y)) * s.x, (1.0f - std::abs(p.x)) * s.y} : p;
}

// ---- reference layouts (CrossRT.h:26-33, cbvh.h:21-33, intersection_common.h:33-56,
//      sdf_simple_shared.h:23-55, render_settings.h:116-136) ----
struct Hit { float t; uint32_t primId, instId, geomId; float coords[4]; };
struct Node { f3 boxMin; uint32_t leftOffset; f3 boxMax; uint32_t escapeIndex; };
struct NodePair { Node left, right; };
struct Geom { f4 boxMin, boxMax; uint32_t bvhOffset, type; uint32_t offset[2]; uint32_t hasChannels, pad[3]; };
struct Inst { f4 boxMin, boxMax; uint32_t geomId, pad[7]; float transform[16], transformInv[16], transformInvTransposed[16]; };
struct FrameNode { float values[8]; uint32_t offset; };
struct FrameTexNode { float tex_coords[16]; float values[8]; uint32_t offset, material_id; };   // sdf_simple_shared.h:57-63
struct SBSNode { uint32_t pos_xy, pos_z_lod_size, data_offset, pad; };
struct SBSHeader { uint32_t brick_size, brick_pad, bytes_per_value, aux_data; };
struct MultiNode { uint32_t data_offset, children_offset, voxel_count, flags; };   // sdf_multi_shared.h:8-14
struct SVSNode { uint32_t pos_xy, pos_z_lod_size, values[2]; };                    // sdf_simple_shared.h:30-35
struct GPrimHeader { float color[3]; uint32_t prim_type; };                         // graphics_primitive_common.h:19-23
// One COctreeV3 geometry: where its words and 3 x 48 index modifiers start, and the
// COctreeV3Header fields the walk reads (sdf/scom/scom_shared.h:24-44).
struct CoctG {
  uint32_t base, mods, size, brick_size, brick_pad, bits_per_value, uv_size, sim_compression, lods;
  uint32_t uints_per_child_info, idx_mask, idx_sh, trans_off, rot_mask, add_mask, pad;
};
struct I4 { int32_t x, y, z, w; };
// SdfDAG records (sdf/scom2/scom_shared.h:77-103) and one DAG geometry: root node, brick shape,
// node grid size, first of its 2 x 48 index modifiers.
struct DagChildEdge { uint32_t child_offset, rotation_id; };
struct DagDataEdge { uint32_t data_offset, rotation_id, type_id; float add; };
struct DagNode { uint32_t children_edges_offset, data_edges_offset; DagChildEdge channels_edge; uint32_t voxel_count_flags, pad; };
struct DagG { uint32_t root, brick_size, brick_pad, node_grid_size, mods, pad[3]; };
// SCom2Header (sdf/scom2/scom_shared.h:21-75)
struct SC2Header {
  uint32_t brick_size, v_size, bits_per_value, values_per_uint, value_mask, bitmask_len, dimension;
  uint32_t child_rot_shift, child_rot_mask, child_add_shift, child_add_mask, child_offset_mask, child_offset_off;
  uint32_t node_offset_mask, uints_per_link, unique_brick_prefix, unique_brick_offset_mask;
  uint32_t children_types_shift, children_types_mask, base_reference_shift, children_active_bits_shift;
  uint32_t children_active_bits_mask, references_offset, reference_bits, reference_mask, references_per_uint;
  uint32_t links_offset, max_surface_count, max_surface_count_per_leaf;
  uint32_t bricks_step, bricks_arr_offset, nodes_arr_offset, root_node_off;
  uint32_t has_channels, has_surfaces, has_multi_nodes;
  int32_t tex_id_off, mat_id_off, all_float_tex_id_off, all_int_mat_id_off;
  float max_val;
  uint32_t max_depth;
  float user_params[7];
  uint32_t pad[5];
};
static_assert(sizeof(SC2Header) == 216, "layout");
struct Preset {
  uint32_t render_mode, sdf_node_intersect, normal_mode, ray_gen_mode, interpolation_mode, representation_mode, spp,
      fixed_lod;
  float level_of_detail, compact_sdf_eps, LOD_distance;
  uint32_t channel_render_mode, channel_id, component_id, CTF_id;
};
static_assert(sizeof(Hit) == 32 && sizeof(Node) == 32 && sizeof(NodePair) == 64, "layout");
static_assert(sizeof(Geom) == 64 && sizeof(Inst) == 256 && sizeof(FrameNode) == 36 && sizeof(Preset) == 60, "layout");

constexpr uint32_t LEAF_BIT = 0x80000000u, START_MASK = 0x0FFFFFFFu, SIZE_MASK = 0x70000000u;
constexpr int STACK_SIZE = 80;
constexpr uint32_t TAG_NONE = 0, TAG_TRIANGLE = 1, TAG_SDF_BRICK = 4, TAG_SDF_NODE = 8, TAG_SDF_OCTREE = 15;
constexpr uint32_t TAG_SDF_GRID = 11, TYPE_SDF_GRID = 1, TAG_GRAPHICS_PRIM = 9, TYPE_GRAPHICS_PRIM = 17;
constexpr uint32_t TAG_COCTREE_BRICKED = 2, TAG_SDF_DAG = 18, TYPE_SDF_DAG = 25, TYPE_SCOM2 = 26, TAG_SCOM2 = 19;
// GraphicsPrimType (utils/misc/graphics_primitive_common.h:7-16)
constexpr uint32_t GP_POINT = 0, GP_LINE = 1, GP_SEGMENT = 2, GP_SEGMENT_DIR = 3, GP_BOX = 4, GP_POINT_COLOR = 5,
                   GP_LINE_COLOR = 6, GP_SEGMENT_COLOR = 7, GP_SEGMENT_DIR_COLOR = 8, GP_BOX_COLOR = 9;
constexpr uint32_t TYPE_SDF_SVS = 5;
constexpr uint32_t TYPE_SDF_FRAME_OCTREE_TEX = 8;
constexpr uint32_t TYPE_MESH_TRIANGLE = 0, TYPE_SDF_FRAME_OCTREE = 3, SH_TYPE = 24, GEOM_ID_MASK = 0x00FFFFFFu;
constexpr uint32_t LAYOUT_DX_UV16 = 2u << 24, LAYOUT_DX_RGB8 = 3u << 24, LAYOUT_ID32F = 4u << 24, LAYOUT_ID32F_IN = 5u << 24;

}  // namespace

extern "C" {

// Host arrays of a committed scene (the BVHRT members, BVH2Common.h:262-365).
struct OracleScene {
  const NodePair* nodes;
  const Node* tlas;
  const Inst* instances;
  const Geom* geoms;
  const uint32_t* geom_tags;
  const uint32_t* start_end;     // uint2 per geometry
  const uint32_t* prim_id_count;
  const f4* vert_pos;
  const f4* vert_norm;
  const uint32_t* indices;
  const uint32_t* prim_indices;
  const Node* orig_nodes;
  const FrameNode* frame_nodes;
  const uint32_t* frame_roots;
  const FrameTexNode* frame_tex_nodes;   // m_SdfFrameOctreeTexNodes
  const uint32_t* frame_tex_roots;
  const SBSNode* sbs_nodes;
  const uint32_t* sbs_data;
  const float* sbs_data_f;
  const uint32_t* sbs_roots;
  const SBSHeader* sbs_headers;
  const MultiNode* multi_nodes;
  const float* multi_values;
  const SVSNode* svs_nodes;
  const uint32_t* svs_roots;
  const float* grid_data;
  const uint32_t* grid_offsets;
  const uint32_t* grid_sizes;     // 3 per grid
  const GPrimHeader* gprim_headers;
  const f4* gprim_points;
  const uint32_t* coct_data;      // m_SdfCompactOctreeV3Data of every COctreeV3 geometry in turn
  const CoctG* coct_geoms;
  const I4* coct_mods;            // m_SdfCompactOctreeRotModifiers, 3 x 48 per geometry
  const DagNode* dag_nodes;       // m_SdfDAGNodes and their edges and distances
  const DagChildEdge* dag_child_edges;
  const DagDataEdge* dag_data_edges;
  const float* dag_distances;
  const DagG* dag_geoms;
  const I4* dag_mods;             // 2 x 48 index modifiers per DAG (brick lattice, node grid)
  const uint8_t* rot_add;         // m_RotAddTable, 48 x 48
  const uint32_t* scom2_data;     // m_SCom2Data
  const SC2Header* scom2_headers; // m_SCom2Headers (offsets into scom2_data)
  const I4* scom2_mods;           // 3 x 48 index modifiers per tree
  Preset preset;
  uint32_t n_scom2;               // SCom2 trees in the scene
};

// Per-ray traversal counters (the RayInfo idea, BVH2Common.cpp:3403-3571) for the
// algorithmic-bytes model.
struct OracleCounts {
  uint32_t blas_nodes, leaves, prims, tlas_nodes, instances, sdf_voxels;
};

}  // extern "C"

namespace {

// The voxel of the latest brick-set hit (indexed layouts): value indices of its corners, entry
// point in voxel units, direction, local hit parameter and the hit t it produced.
struct VoxelRec {
  uint32_t hit;
  uint32_t idx[8];
  float start_q[3];
  float dir[3];
  float t_local;
  float t;
  float pad;
};
static_assert(sizeof(VoxelRec) == 72, "layout");

// scom::initialize_rot_transforms (sdf/scom2/clustering/octahedral_groups.cpp:37-101) in three
// dimensions: the axis permutations in lexicographic order (the recursive enumeration's order),
// each with the 8 sign patterns (bit d negates axis d), as a matrix R with R[perm[d]][d] = sign_d,
// turned about the lattice centre c = v/2 - 0.5 (T1 R T2, translations by +c and -c), then taken
// in the legacy order of legacy_to_universal_id_remap (octahedral_groups.h:27-35).
struct RotTransforms {
  float M[48][4][4];
  explicit RotTransforms(int v) {
    static const int legacy_to_universal[48] = {0, 24, 32, 1, 25, 33, 8, 16, 40, 9, 17, 41,
                                                2, 26, 34, 3, 27, 35, 10, 18, 42, 11, 19, 43,
                                                4, 28, 36, 5, 29, 37, 12, 20, 44, 13, 21, 45,
                                                6, 30, 38, 7, 31, 39, 14, 22, 46, 15, 23, 47};
    float U[48][4][4];
    int perm[3] = {0, 1, 2};
    int t_id = 0;
    const float c = v / 2.0f - 0.5f;
    do {
      for (int sign = 0; sign < 8; ++sign) {
        float R[4][4] = {}, T1[4][4] = {}, T2[4][4] = {}, A[4][4] = {};
        for (int d = 0; d < 3; ++d) R[perm[d]][d] = ((sign >> d) % 2 == 0) ? 1.0f : -1.0f;
        R[3][3] = 1.0f;
        for (int d = 0; d < 4; ++d) T1[d][d] = T2[d][d] = 1.0f;
        for (int d = 0; d < 3; ++d) { T1[d][3] = c; T2[d][3] = -c; }
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b)
            for (int k = 0; k < 4; ++k) A[a][b] += R[a][k] * T2[k][b];
        float* out = &U[t_id * 8 + sign][0][0];
        for (int a = 0; a < 4; ++a)
          for (int b = 0; b < 4; ++b) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += T1[a][k] * A[k][b];
            out[a * 4 + b] = acc;
          }
      }
      ++t_id;
    } while (std::next_permutation(perm, perm + 3));
    for (int i = 0; i < 48; ++i) std::memcpy(M[i], U[legacy_to_universal[i]], sizeof(M[i]));
  }
  // to_float3(M * float4(p, 1))
  f3 apply(int r, f3 p) const {
    const float (*m)[4] = M[r];
    return f3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3], m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
              m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]);
  }
};

// one table per lattice size, built on first use
inline const RotTransforms& rot_transforms(int v) {
  static RotTransforms* tables[33] = {};
  #pragma omp critical(oracle_rot_transforms)
  {
    if (!tables[v]) tables[v] = new RotTransforms(v);
  }
  return *tables[v];
}

struct Tracer {
  const OracleScene& s;
  OracleCounts* cnt;
  VoxelRec* rec = nullptr;     // records the voxel of each brick-set hit (test support)
  float last_t_local = 0.0f;   // local hit parameter of the latest LocalSurfaceIntersection hit
  const FrameTexNode* tex_node = nullptr;   // the textured frame-octree node being intersected

  // support_fast_scom_traversal (BVH2Common_host.cpp:1436-1451): one plain SCom2 tree, LAMBERT_NO_TEX
  bool scom2_fast() const {
    if (s.n_scom2 != 1 || s.preset.render_mode != 1) return false;
    const SC2Header& h = s.scom2_headers[0];
    return h.has_multi_nodes == 0 && h.has_channels == 0 && h.v_size == 3 && h.links_offset == 1 && h.child_offset_off == 0;
  }

  // BVH2Common.cpp:19-32
  bool need_normal() const {
    const uint32_t m = s.preset.render_mode;
    return m == 1 || m == 8 || m == 12 || m == 17 || m == 20 || m == 7 || m == 18 || m == 23 || m == 21 || m == 11 ||
           m == 27;
  }

  // BVH2Common.cpp:1571-1635
  void IntersectAllTrianglesInLeaf(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                                   uint32_t a_start, uint32_t a_count, Hit* pHit) {
    const uint32_t* go = s.geoms[geomId].offset;
    for (uint32_t triId = a_start; triId < a_start + a_count; triId++) {
      if (cnt) cnt->prims++;
      const uint32_t A = s.indices[go[0] + triId * 3 + 0];
      const uint32_t B = s.indices[go[0] + triId * 3 + 1];
      const uint32_t C = s.indices[go[0] + triId * 3 + 2];
      const f3 A_pos = to3(s.vert_pos[go[1] + A]);
      const f3 B_pos = to3(s.vert_pos[go[1] + B]);
      const f3 C_pos = to3(s.vert_pos[go[1] + C]);
      const f3 edge1 = B_pos - A_pos;
      const f3 edge2 = C_pos - A_pos;
      const f3 pvec = cross(ray_dir, edge2);
      const f3 tvec = ray_pos - A_pos;
      const f3 qvec = cross(tvec, edge1);
      const float invDet = 1.0f / dot(edge1, pvec);
      const float v = dot(tvec, pvec) * invDet;
      const float u = dot(qvec, ray_dir) * invDet;
      const float t = dot(edge2, qvec) * invDet;
      if (v >= -1e-6f && u >= -1e-6f && (u + v <= 1.0f + 1e-6f) && t > tNear && t < pHit->t) {
        pHit->t = t;
        pHit->primId = triId;
        pHit->instId = instId;
        pHit->geomId = geomId | (TYPE_MESH_TRIANGLE << SH_TYPE);
#ifdef ORACLE_PLANT_SWAP_UV
        // planted mistake, compiled only by the test that proves the checks catch it
        pHit->coords[0] = v;
        pHit->coords[1] = u;
#else
        pHit->coords[0] = u;
        pHit->coords[1] = v;
#endif
        if (need_normal()) {
          f3 n(1, 0, 0);
          if (s.preset.normal_mode == 0) {
            n = cross(edge1, edge2);
          } else if (s.preset.normal_mode == 1) {
            const f4 a = s.vert_norm[go[1] + A], b = s.vert_norm[go[1] + B], c = s.vert_norm[go[1] + C];
            const float w = 1.0f - u - v;
            // to_float3(nA*(1-u-v) + nB*v + u*nC)
            n = f3(a.x * w + b.x * v + u * c.x, a.y * w + b.y * v + u * c.y, a.z * w + b.z * v + u * c.z);
          }
          n = normalize(matmul4x3(s.instances[instId].transformInvTransposed, n));
          const f2 e = encode_normal(n);
          pHit->coords[2] = e.x;
          pHit->coords[3] = e.y;
        } else {
          pHit->coords[2] = 0;
          pHit->coords[3] = 0;
        }
      }
    }
  }

  // BVH2Common.cpp:126-136
  static float eval_dist_trilinear(const float v[8], f3 dp) {
    return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * v[0] + (1 - dp.x) * (1 - dp.y) * (dp.z) * v[1] +
           (1 - dp.x) * (dp.y) * (1 - dp.z) * v[2] + (1 - dp.x) * (dp.y) * (dp.z) * v[3] +
           (dp.x) * (1 - dp.y) * (1 - dp.z) * v[4] + (dp.x) * (1 - dp.y) * (dp.z) * v[5] +
           (dp.x) * (dp.y) * (1 - dp.z) * v[6] + (dp.x) * (dp.y) * (dp.z) * v[7];
  }

  // BVH2Common.cpp:345-725 (no cutting plane); textured frame octrees take their texture
  // coordinates (:689-710); the compact types (DAG, SCom2, :340-343) widen EPS on surface nodes
  void LocalSurfaceIntersection(uint32_t type, f3 ray_dir, uint32_t instId, uint32_t geomId, const float values[8],
                                uint32_t primId, float d, float qNear, float qFar, f2 fNearFar, f3 start_q, Hit* pHit,
                                bool is_surface_node = false) {
    const float BASE_EPS = 1e-6f;
    const bool compact = type == TYPE_SDF_DAG || type == TYPE_SCOM2;
    const float EPS = (is_surface_node && compact) ? std::max(BASE_EPS, s.preset.compact_sdf_eps * d) : BASE_EPS;
    const float d_inv = 1.0f / d;
    float t = qNear;
    bool early_hit = false, hit = false;
    unsigned iter = 0;
    float start_sign = 1;
    float start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
    if (is_surface_node || s.preset.representation_mode == 0) start_sign = lm_sign(start_dist);
    start_dist *= start_sign;
    const uint32_t mode = s.preset.sdf_node_intersect;
    if (start_dist <= EPS || mode == 3) {
      early_hit = true;
      hit = true;
    } else if (mode == 0) {
      float dist = start_dist;
      while (t < qFar && dist > EPS && iter < 256u) {
        t += dist * d_inv;
        dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
        iter++;
      }
      hit = (dist <= EPS);
    } else {
      float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
      float s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
      float tmp = s101 - s001;
      float k0 = s000, k1 = s100 - s000, k2 = s010 - s000, k3 = s110 - s010 - k1, k4 = k0 - s001, k5 = k1 - tmp;
      float k6 = k2 - (s011 - s001), k7 = k3 - (s111 - s011 - tmp);
#ifdef ORACLE_PLANT_DROP_K7
      k7 = 0.0f;   // planted mistake (the xyz term of the trilinear), compiled only by the test that catches it
#endif
      f3 o = start_q, d3 = ray_dir;
      float m0 = o.x * o.y, m1 = d3.x * d3.y, m2 = o.x * d3.y + o.y * d3.x;
      float m3 = k5 * o.z - k1, m4 = k6 * o.z - k2, m5 = k7 * o.z - k3;
      float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
      float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
      float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
      float c3 = k7 * m1 * d3.z;
      if (mode == 1) {
        float x1 = 1000, x2 = 1000, x3 = 1000;
        const float PI = 3.14159265358979323846f;
        if (std::abs(c3) > 1e-5f) {
          float a = c2 / c3, b = c1 / c3, c = c0 / c3;
          float Q = (a * a - 3 * b) / 9;
          float R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
          float Q3 = Q * Q * Q;
          if (R * R < Q3) {
            float theta = std::acos(R / std::sqrt(Q3));
            x1 = -2 * std::sqrt(Q) * std::cos(theta / 3) - a / 3;
            x2 = -2 * std::sqrt(Q) * std::cos((theta + 2 * PI) / 3) - a / 3;
            x3 = -2 * std::sqrt(Q) * std::cos((theta - 2 * PI) / 3) - a / 3;
          } else {
            float A = -lm_sign(R) * std::pow(std::abs(R) + std::sqrt(R * R - Q3), 1.0f / 3.0f);
            float B = std::abs(A) > EPS ? Q / A : 0;
            x1 = A + B - a / 3;
          }
        } else if (std::abs(c2) > 1e-7f) {
          float a = c2, b = c1, c = c0;
          float D = b * b - 4 * a * c;
          if (D > 0) {
            float q = -0.5f * (b + lm_sign(b) * std::sqrt(D));
            x1 = q / a;
            if (std::abs(q) > EPS) x2 = c / q;
          }
        } else if (std::abs(c1) > EPS) {
          x1 = -c0 / c1;
        }
        x1 = x1 < 0 ? 1000 : x1;
        x2 = x2 < 0 ? 1000 : x2;
        x3 = x3 < 0 ? 1000 : x3;
        t = std::min(x1, std::min(x2, x3));
        hit = (t >= 0 && t <= qFar);
      } else if (mode == 2) {
        float a = 3 * c3, b = 2 * c2, c = c1;
        float t0 = 0, t1 = qFar, t2 = qFar, t3 = qFar;
        float D = b * b - 4 * a * c;
        if (D >= 0) {
          float q = -0.5f * (b + lm_sign(b) * std::sqrt(D));
          t1 = std::abs(a) > EPS ? q / a : t0;
          t2 = std::abs(q) > EPS ? c / q : qFar;
          float min_t = std::min(t1, t2);
          t2 = std::max(t1, t2);
          t1 = min_t;
          t1 = std::min(std::max(t0, t1), t3);
          t2 = std::min(std::max(t0, t2), t3);
        }
        bool s0 = c0 > 0, s1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3))) > 0;
        bool s2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3))) > 0, s3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3))) > 0;
        float nmin = t0, nmax = t0;
        if (s0 != s1) { nmin = t0; nmax = t1; }
        else if (s1 != s2) { nmin = t1; nmax = t2; }
        else if (s2 != s3) { nmin = t2; nmax = t3; }
        if (nmin < nmax) {
          unsigned it = 0;
          float rtn = 0.5f * (nmin + nmax), f = 1000;
          while (it < 10u && std::abs(f) >= EPS) {
            f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
            float df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
            float dx = f / (df + lm_sign(df) * 1e-9f);
            rtn -= dx;
            it++;
          }
          t = rtn;
          hit = (t >= 0 && t <= qFar && std::abs(f) < EPS);
        } else {
          hit = false;
        }
      } else {
        const float k = 2;
        float e = 0.1f * qFar;
        float t_max = std::abs(c3) < EPS ? 1e6f : -c2 / (3 * c3);
        float df_max = 3 * c3 * t_max * t_max + 2 * c2 * t_max + c1;
        float dist = start_dist;
        while (t < qFar && dist > EPS && iter < 256u) {
          float df_1 = 3 * c3 * t * t + 2 * c2 * t + c1;
          float df_2 = 3 * c3 * (t + e) * (t + e) + 2 * c2 * (t + e) + c1;
          float L = (t_max > t && t_max < t + e) ? std::max(df_max, std::max(df_1, df_2)) : std::max(df_1, df_2);
          L = std::max(L, EPS);
          float st = std::min((dist * d_inv) / L, e);
          t += st;
          e = k * st;
          dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
          iter++;
        }
        hit = (dist <= EPS);
      }
    }
    const float tReal = fNearFar.x + d * t;
    if (t <= qFar && hit && tReal < pHit->t) {
      f3 norm(0, 0, 1);
      if (need_normal()) {
        if (early_hit && mode == 3) {
          norm = f3(0, 0, 0);
          f3 v(std::max(start_q.x, 1 - start_q.x), std::max(start_q.y, 1 - start_q.y), std::max(start_q.z, 1 - start_q.z));
          int dim = 0;
          if (v.y > v.x && v.y > v.z) dim = 1;
          else if (v.z > v.x && v.z > v.y) dim = 2;
          const float q = dim == 0 ? start_q.x : (dim == 1 ? start_q.y : start_q.z);
          const float nv = q < 0.5f ? -1.0f : 1.0f;
          if (dim == 0) norm.x = nv; else if (dim == 1) norm.y = nv; else norm.z = nv;
        } else {
          f3 p0 = start_q + t * ray_dir;
          const float h = 0.001f;
          float ddx = (eval_dist_trilinear(values, p0 + f3(h, 0, 0)) - eval_dist_trilinear(values, p0 + f3(-h, 0, 0))) / (2 * h);
          float ddy = (eval_dist_trilinear(values, p0 + f3(0, h, 0)) - eval_dist_trilinear(values, p0 + f3(0, -h, 0))) / (2 * h);
          float ddz = (eval_dist_trilinear(values, p0 + f3(0, 0, h)) - eval_dist_trilinear(values, p0 + f3(0, 0, -h))) / (2 * h);
          norm = start_sign * f3(ddx, ddy, ddz);
        }
        norm = normalize(matmul4x3(s.instances[instId].transformInvTransposed, norm));
      }
      f2 en = encode_normal(norm);
      last_t_local = t;
      pHit->t = tReal;
      pHit->primId = primId;
      pHit->instId = instId;
      pHit->geomId = geomId | (type << SH_TYPE);
      pHit->coords[0] = 0;
      pHit->coords[1] = 0;
      pHit->coords[2] = en.x;
      pHit->coords[3] = en.y;
      if (s.preset.render_mode == 10) pHit->primId = iter;
      // textured frame octrees: corner texture coordinates at the hit point (:689-710)
      if (type == TYPE_SDF_FRAME_OCTREE_TEX && tex_node) {
        const f3 dp = start_q + t * ray_dir;
        const float* tc = tex_node->tex_coords;
        pHit->coords[0] = (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * tc[0] + (1 - dp.x) * (1 - dp.y) * (dp.z) * tc[2] +
                          (1 - dp.x) * (dp.y) * (1 - dp.z) * tc[4] + (1 - dp.x) * (dp.y) * (dp.z) * tc[6] +
                          (dp.x) * (1 - dp.y) * (1 - dp.z) * tc[8] + (dp.x) * (1 - dp.y) * (dp.z) * tc[10] +
                          (dp.x) * (dp.y) * (1 - dp.z) * tc[12] + (dp.x) * (dp.y) * (dp.z) * tc[14];
        pHit->coords[1] = (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * tc[1] + (1 - dp.x) * (1 - dp.y) * (dp.z) * tc[3] +
                          (1 - dp.x) * (dp.y) * (1 - dp.z) * tc[5] + (1 - dp.x) * (dp.y) * (dp.z) * tc[7] +
                          (dp.x) * (1 - dp.y) * (1 - dp.z) * tc[9] + (dp.x) * (1 - dp.y) * (dp.z) * tc[11] +
                          (dp.x) * (dp.y) * (1 - dp.z) * tc[13] + (dp.x) * (dp.y) * (dp.z) * tc[15];
      }
      // channel render modes: interpolation quotients instead of texture coordinates (:714-723)
      if ((type == 9 || type == TYPE_SDF_DAG || type == TYPE_SCOM2) &&
          (s.preset.render_mode == 23 || s.preset.render_mode == 27)) {
        f3 dp = start_q + t * ray_dir;
        dp = f3(std::min(std::max(0.0f, float(0xFFFFu) * dp.x), float(0xFFFFu)),
                std::min(std::max(0.0f, float(0xFFFFu) * dp.y), float(0xFFFFu)),
                std::min(std::max(0.0f, float(0xFFFFu) * dp.z), float(0xFFFFu)));
        const int ux = int(dp.x), uy = int(dp.y), uz = int(dp.z);
        const uint32_t c0 = uint32_t((uy << 16) | ux), c1 = uint32_t(uz);
        std::memcpy(&pHit->coords[0], &c0, 4);
        std::memcpy(&pHit->coords[1], &c1, 4);
      }
    }
  }

  // BVH2Common.cpp:727-749 + 141-167. Leaf boxes of geometry g start at geoms[g].offset[1].
  // Textured frame octrees: :168-188, their nodes in frame_tex_nodes. SVS: BVH2Common.cpp:190-215
  // (box from the voxel's position, byte values decoded)
  void OctreeNodeIntersect(uint32_t type, f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                           uint32_t a_start, Hit* pHit) {
    if (type != TYPE_SDF_FRAME_OCTREE && type != TYPE_SDF_SVS && type != TYPE_SDF_FRAME_OCTREE_TEX) return;
    const uint32_t sdfId = s.geoms[geomId].offset[0];
    uint32_t primId = 0;
    f3 min_pos, max_pos;
    float values[8];
    tex_node = nullptr;
    if (type == TYPE_SDF_FRAME_OCTREE || type == TYPE_SDF_FRAME_OCTREE_TEX) {
      const Node& on = s.orig_nodes[s.geoms[geomId].offset[1] + a_start];
      primId = on.leftOffset;
      min_pos = on.boxMin;
      max_pos = on.boxMax;
      if (type == TYPE_SDF_FRAME_OCTREE) {
        const uint32_t nodeId = primId + s.frame_roots[sdfId];
        for (int i = 0; i < 8; i++) values[i] = s.frame_nodes[nodeId].values[i];
      } else {
        const uint32_t nodeId = primId + s.frame_tex_roots[sdfId];
        tex_node = &s.frame_tex_nodes[nodeId];
        for (int i = 0; i < 8; i++) values[i] = tex_node->values[i];
      }
    } else {
      primId = a_start;
      const SVSNode& nd = s.svs_nodes[primId + s.svs_roots[sdfId]];
      const float px = nd.pos_xy >> 16, py = nd.pos_xy & 0x0000FFFF;
      const float pz = nd.pos_z_lod_size >> 16, sz = nd.pos_z_lod_size & 0x0000FFFF;
      const float d_max = 2 * 1.73205081f / sz;
      min_pos = f3(-1, -1, -1) + 2.0f * f3(px, py, pz) / sz;
      max_pos = min_pos + 2.0f * f3(1, 1, 1) / sz;
      for (int i = 0; i < 8; i++)
        values[i] = -d_max + 2.0f * d_max * (1.0f / 255.0f) * ((nd.values[i / 4] >> (8 * (i % 4))) & 0xFF);
    }
    const f3 size = max_pos - min_pos;
    const f2 fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
    const f3 start_pos = ray_pos + fNearFar.x * ray_dir;
    const float d = std::max(size.x, std::max(size.y, size.z));
    const f3 start_q = (start_pos - min_pos) / d;
    const float qFar = (fNearFar.y - fNearFar.x) / d;
    const float qNear = tNear > fNearFar.x ? (tNear - fNearFar.x) / d : 0.0f;
    if (qNear > 0.0f) return;
    LocalSurfaceIntersection(type, ray_dir, instId, geomId, values, primId, d, qNear, qFar, fNearFar, start_q, pHit);
  }

  // BVH2Common.cpp:846-882
  float load_distance_values(uint32_t nodeId, f3 voxelPos, uint32_t v_size, float sz_inv, const SBSHeader& h,
                             float values[8]) {
    float vmin = 1e6f;
    const uint32_t v_off = s.sbs_nodes[nodeId].data_offset;
    for (int i = 0; i < 8; i++) {
      const uint32_t x = uint32_t(voxelPos.x) + ((i & 4) >> 2), y = uint32_t(voxelPos.y) + ((i & 2) >> 1),
                     z = uint32_t(voxelPos.z) + (i & 1);
      // SBS_v_to_i (sdf_common_shared.h:49-52)
      const uint32_t vId = (x + h.brick_pad) * v_size * v_size + (y + h.brick_pad) * v_size + (z + h.brick_pad);
      if (h.aux_data == LAYOUT_ID32F || h.aux_data == LAYOUT_ID32F_IN) {
        values[i] = s.sbs_data_f[s.sbs_data[v_off + vId]];
      } else {
        const uint32_t vpi = 4 / h.bytes_per_value, bits = 8 * h.bytes_per_value;
        const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
        const float d_max = 1.73205081f * sz_inv;
        const float mult = 2 * d_max / max_val;
        const uint32_t sh = bits * (vId % vpi);
        values[i] = -d_max + mult * ((sh < 32 ? s.sbs_data[v_off + vId / vpi] >> sh : 0u) & max_val);
      }
      vmin = std::min(vmin, values[i]);
    }
    return vmin;
  }

  // BVH2Common.cpp:790-844: 4^3 samples from one voxel before the voxel's corner, stored at
  // 16 z + 4 y + x. The reference converts the float position (which can be -1) to an unsigned
  // index before adding the padding; with brick_pad >= 1 (trace calls refuse less) that is the
  // integer arithmetic below.
  float load_tricubic_distance_values(uint32_t nodeId, f3 voxelPos, uint32_t v_size, float sz_inv, const SBSHeader& h,
                                      float values[64]) {
    float vmin = 1e6f;
    const uint32_t v_off = s.sbs_nodes[nodeId].data_offset;
    const int p0x = int(voxelPos.x) - 1, p0y = int(voxelPos.y) - 1, p0z = int(voxelPos.z) - 1;
    for (int x = 0; x < 4; x++)
      for (int y = 0; y < 4; y++)
        for (int z = 0; z < 4; z++) {
          const uint32_t vId = uint32_t(p0x + x + int(h.brick_pad)) * v_size * v_size +
                               uint32_t(p0y + y + int(h.brick_pad)) * v_size + uint32_t(p0z + z + int(h.brick_pad));
          float v;
          if (h.aux_data == LAYOUT_ID32F || h.aux_data == LAYOUT_ID32F_IN) {
            v = s.sbs_data_f[s.sbs_data[v_off + vId]];
          } else {
            const uint32_t vpi = 4 / h.bytes_per_value, bits = 8 * h.bytes_per_value;
            const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
            const float d_max = 1.73205081f * sz_inv;
            const float mult = 2 * d_max / max_val;
            const uint32_t sh = bits * (vId % vpi);
            v = -d_max + mult * ((sh < 32 ? s.sbs_data[v_off + vId / vpi] >> sh : 0u) & max_val);
          }
          values[16 * z + 4 * y + x] = v;
          vmin = std::min(vmin, v);
        }
    return vmin;
  }

  // tricubic_spline (BVH2Common.cpp:1526-1529): Catmull-Rom through p0..p3 at x in [0, 1]
  static float tricubic_spline(float p0, float p1, float p2, float p3, float x) {
    return p1 + 0.5f * x * (p2 - p0 + x * (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3 + x * (3.0f * (p1 - p2) + p3 - p0)));
  }

  // BVHRT::tricubicInterpolation (BVH2Common.cpp:1531-1569): along x for each (y, z) row, then
  // along y, then along z
  static float tricubicInterpolation(const float grid[64], const float dp[3]) {
    float values_yz[16], values_z[4];
    for (uint32_t j = 0; j < 4; ++j)
      for (uint32_t k = 0; k < 4; ++k)
        values_yz[4 * j + k] = tricubic_spline(grid[k * 16 + j * 4 + 0], grid[k * 16 + j * 4 + 1], grid[k * 16 + j * 4 + 2],
                                               grid[k * 16 + j * 4 + 3], dp[0]);
    for (uint32_t k = 0; k < 4; ++k)
      values_z[k] = tricubic_spline(values_yz[k], values_yz[4 + k], values_yz[8 + k], values_yz[12 + k], dp[1]);
    return tricubic_spline(values_z[0], values_z[1], values_z[2], values_z[3], dp[2]);
  }
  static float tricubic_at(const float grid[64], f3 p) {
    const float dp[3] = {p.x, p.y, p.z};
    return tricubicInterpolation(grid, dp);
  }

  // BVHRT::TricubicLocalSurfaceIntersection (BVH2Common.cpp:220-335): only ST mode marches; other
  // modes hit where the ray starts inside (BBOX at once)
  void TricubicLocalSurfaceIntersection(uint32_t type, f3 ray_dir, uint32_t instId, uint32_t geomId,
                                        const float values[64], uint32_t primId, float d, float qNear, float qFar,
                                        f2 fNearFar, f3 start_q, Hit* pHit) {
    const float EPS = 1e-6f;
    const float d_inv = 1.0f / d;
    float t = qNear;
    bool hit = false;
    unsigned iter = 0;
    float start_sign = 1;
    float start_dist = tricubic_at(values, start_q + t * ray_dir);
    if (s.preset.representation_mode == 0) start_sign = lm_sign(start_dist);   // REPRESENTATION_MODE_SURFACE
    start_dist *= start_sign;
    if (start_dist <= EPS || s.preset.sdf_node_intersect == 3) {
      hit = true;
    } else if (s.preset.sdf_node_intersect == 0) {
      const unsigned ST_max_iters = 256;
      float dist = start_dist;
      while (t < qFar && dist > EPS && iter < ST_max_iters) {
        t += dist * d_inv;
        dist = tricubic_at(values, start_q + t * ray_dir);
        iter++;
      }
      hit = (dist <= EPS);
    }
    const float tReal = fNearFar.x + d * t;
    if (t <= qFar && hit && tReal < pHit->t) {
      f3 norm(0, 0, 1);
      if (need_normal()) {
        const f3 p0 = start_q + t * ray_dir;
        const float h = 0.001f;
        const float ddx = (tricubic_at(values, f3((p0 + f3(h, 0, 0)).x, p0.y, p0.z)) -
                           tricubic_at(values, f3((p0 + f3(-h, 0, 0)).x, p0.y, p0.z))) / (2 * h);
        const float ddy = (tricubic_at(values, f3(p0.x, (p0 + f3(0, h, 0)).y, p0.z)) -
                           tricubic_at(values, f3(p0.x, (p0 + f3(0, -h, 0)).y, p0.z))) / (2 * h);
        const float ddz = (tricubic_at(values, f3(p0.x, p0.y, (p0 + f3(0, 0, h)).z)) -
                           tricubic_at(values, f3(p0.x, p0.y, (p0 + f3(0, 0, -h)).z))) / (2 * h);
        norm = start_sign * normalize(matmul4x3(s.instances[instId].transformInvTransposed, f3(ddx, ddy, ddz)));
      }
      const f2 en = encode_normal(norm);
      pHit->t = tReal;
      pHit->primId = primId;
      pHit->instId = instId;
      pHit->geomId = geomId | (type << SH_TYPE);
      pHit->coords[0] = 0;
      pHit->coords[1] = 0;
      pHit->coords[2] = en.x;
      pHit->coords[3] = en.y;
      if (s.preset.render_mode == 10) pHit->primId = iter;   // MULTI_RENDER_MODE_ST_ITERATIONS
    }
  }

  // rayCapsuleIntersect (BVH2Common.cpp:1184-1223)
  static f4 rayCapsuleIntersect(f3 ray_pos, f3 ray_dir, f3 pa, f3 pb, float ra) {
    f3 ba = pb - pa;
    f3 oa = ray_pos - pa;
    f3 ob = ray_pos - pb;
    float baba = dot(ba, ba);
    float bard = dot(ba, ray_dir);
    float baoa = dot(ba, oa);
    float rdoa = dot(ray_dir, oa);
    float oaoa = dot(oa, oa);
    float a = baba - bard * bard;
    float b = baba * rdoa - baoa * bard;
    float c = baba * oaoa - baoa * baoa - ra * ra * baba;
    float h = b * b - a * c;
    if (h >= 0.f) {
      float dist = (-b - std::sqrt(h)) / a;
      float y = baoa + dist * bard;
      if (y > 0.f && y < baba) {
        f3 n = normalize(oa + dist * ray_dir - ba * y / baba);
        return f4{n.x, n.y, n.z, dist};
      }
      f3 oc = (y <= 0.0f) ? oa : ob;
      b = dot(ray_dir, oc);
      c = dot(oc, oc) - ra * ra;
      h = b * b - c;
      if (h > 0.f) {
        dist = -b - std::sqrt(h);
        f3 n = normalize(ray_pos + dist * ray_dir + (oc - ray_pos));
        return f4{n.x, n.y, n.z, dist};
      }
    }
    return f4{2e15f, 2e15f, 2e15f, 2e15f};
  }

  // BVHRT::IntersectGraphicPrims (BVH2Common.cpp:1225-1440)
  void IntersectGraphicPrims(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, uint32_t a_start,
                             Hit* pHit) {
    const float EPS = 1e-5f, T_MAX = 1e15f;
    const GPrimHeader header = s.gprim_headers[s.geoms[geomId].offset[0]];
    f3 color(header.color[0], header.color[1], header.color[2]);
    const uint32_t pt = header.prim_type;
    uint32_t prim_len = 2u;
    if (pt == GP_POINT) prim_len = 1u;
    if (pt >= GP_LINE_COLOR) prim_len = 3u;
    const uint32_t start_index = s.geoms[geomId].offset[1] + prim_len * a_start;
    const uint32_t end_index = start_index + prim_len;
    const bool has_custom_color = pt >= GP_POINT_COLOR;
    auto store = [&](float t, f3 norm) {
      f2 e = encode_normal(norm);
      pHit->t = t;
      pHit->primId = a_start;
      pHit->instId = instId;
      pHit->geomId = geomId | (TYPE_GRAPHICS_PRIM << SH_TYPE);
      pHit->coords[0] = color.x + color.y / 256.0f;
      pHit->coords[1] = color.z;
      pHit->coords[2] = e.x;
      pHit->coords[3] = e.y;
    };
    if (pt == GP_POINT || pt == GP_POINT_COLOR) {
      for (uint32_t i = start_index; i < end_index; i += prim_len) {
        f4 point = s.gprim_points[i];
        if (has_custom_color) color = to3(s.gprim_points[i + 1]);
        f3 point3(point.x, point.y, point.z);
        float pt_radius = point.w;
        f3 pmp = ray_pos - point3;
        float b = 2.f * dot(ray_dir, pmp);
        float c = dot(pmp, pmp) - pt_radius * pt_radius;
        float D = b * b - 4 * c;
        if (D >= EPS) {
          D = std::sqrt(D);
          float t = -(b + D) * 0.5f;
          if (t < EPS) t = (-b + D) * 0.5f;
          if (t > tNear && t < pHit->t) store(t, normalize((ray_pos + t * ray_dir - point3) * 100.f));
        }
      }
      return;
    }
    for (uint32_t i = start_index; i < end_index; i += prim_len) {
      f4 point1 = s.gprim_points[i];
      f4 point2 = s.gprim_points[i + 1];
      if (has_custom_color) color = to3(s.gprim_points[i + 2]);
      f3 point1_3(point1.x, point1.y, point1.z);
      f3 point2_3(point2.x, point2.y, point2.z);
      const bool arrow = pt == GP_SEGMENT_DIR || pt == GP_SEGMENT_DIR_COLOR;
      if (arrow) point2_3 = point2_3 * 0.71f + point1_3 * 0.29f;
      float ra = point1.w;
      float t = T_MAX;
      f3 norm(0.f, 0.f, 0.f);
      if (pt == GP_LINE || pt == GP_SEGMENT || pt == GP_SEGMENT_DIR || pt == GP_LINE_COLOR || pt == GP_SEGMENT_COLOR ||
          pt == GP_SEGMENT_DIR_COLOR) {
        f3 oc = ray_pos - point1_3;
        f3 ba = point2_3 - point1_3;
        float baba = dot(ba, ba);
        float bard = dot(ba, ray_dir);
        float baoc = dot(ba, oc);
        float k2 = baba - (bard * bard);
        float k1 = baba * dot(oc, ray_dir) - baoc * bard;
        float k0 = baba * dot(oc, oc) - baoc * baoc - ra * ra * baba;
        float D = k1 * k1 - k2 * k0;
        if (D > 0.f) {
          float dist = -(k1 + std::sqrt(D)) / k2;
          float y = baoc + dist * bard;
          if (pt == GP_LINE || (y > 0.f && y < baba)) {
            t = dist;
            norm = normalize(oc + t * ray_dir - ba * y / baba);
          } else {
            dist = ((y < 0.f ? 0.f : baba) - baoc) / bard;
            if (std::abs(k1 + k2 * dist) < std::sqrt(D)) {
              t = dist;
              norm = normalize(ba * lm_sign(y));
            }
          }
        }
      }
      if (arrow) {
        ra = length(point2_3 - point1_3) * 0.15f;
        point2_3 = f3(point2.x, point2.y, point2.z);
        f3 pa = point2_3 * 0.7f + point1_3 * 0.3f;
        f3 ba = point2_3 - pa;
        f3 oa = ray_pos - pa;
        float m0 = dot(ba, ba), m1 = dot(oa, ba), m2 = dot(ray_dir, ba), m3 = dot(ray_dir, oa), m5 = dot(oa, oa);
        if (m1 < 0.f) {
          f3 tmp1 = oa * m2 - ray_dir * m1;
          if (dot(tmp1, tmp1) < (ra * ra * m2 * m2)) {
            t = -m1 / m2;
            norm = normalize(-1.0f * ba);
          }
        }
        float rr = ra;
        float hy = m0 + rr * rr;
        float k2 = m0 * m0 - m2 * m2 * hy;
        float k1 = m0 * m0 * m3 - m1 * m2 * hy + m0 * ra * (rr * m2);
        float k0 = m0 * m0 * m5 - m1 * m1 * hy + m0 * ra * (rr * m1 * 2.0f - m0 * ra);
        float D = k1 * k1 - k2 * k0;
        if (D > 0.f) {
          float dist = -(k1 + std::sqrt(D)) / k2;
          float y = m1 + dist * m2;
          if (y >= 0.f && y <= m0) {
            t = dist;
            norm = normalize(m0 * (m0 * (oa + t * ray_dir) + rr * ba * ra) - ba * hy * y);
          }
        }
      }
      if (pt == GP_BOX || pt == GP_BOX_COLOR) {
        f3 lo(std::min(point2_3.x, point1_3.x), std::min(point2_3.y, point1_3.y), std::min(point2_3.z, point1_3.z));
        f3 hi(std::max(point2_3.x, point1_3.x), std::max(point2_3.y, point1_3.y), std::max(point2_3.z, point1_3.z));
        float minmax[3][2] = {{lo.x, hi.x}, {lo.y, hi.y}, {lo.z, hi.z}};
        const int vert_indices[4] = {0, 3, 5, 6};
        for (int vi = 0; vi < 4; ++vi) {
          int ax[3] = {int((vert_indices[vi] & 4u) != 0), int((vert_indices[vi] & 2u) != 0), int((vert_indices[vi] & 1u) != 0)};
          f3 pa(minmax[0][ax[0]], minmax[1][ax[1]], minmax[2][ax[2]]);
          for (int axis = 0; axis < 3; ++axis) {
            f3 pb = pa;
            float& slot = axis == 0 ? pb.x : (axis == 1 ? pb.y : pb.z);
            slot = minmax[axis][1 - ax[axis]];
            f4 nd = rayCapsuleIntersect(ray_pos, ray_dir, pa, pb, ra);
            if (nd.w < t) {
              t = nd.w;
              norm = f3(nd.x, nd.y, nd.z);
            }
          }
        }
      }
      if (t > tNear && t < pHit->t) store(t, norm);
    }
  }

  // BVHRT::eval_distance_sdf_grid (BVH2Common.cpp:1442-1520): trilinear, or tricubic where the
  // 4^3 samples fit inside the grid (trilinear next to the faces)
  float eval_distance_sdf_grid(uint32_t grid_id, f3 pos) const {
    const uint32_t off = s.grid_offsets[grid_id];
    const uint32_t sx = s.grid_sizes[3 * grid_id], sy = s.grid_sizes[3 * grid_id + 1], sz = s.grid_sizes[3 * grid_id + 2];
    const f3 grid_size_f(float(sx - 1), float(sy - 1), float(sz - 1));
    f3 vox_f = grid_size_f * ((pos - f3(-1, -1, -1)) / f3(2, 2, 2));
    vox_f = f3(std::min(std::max(vox_f.x, 0.0f), grid_size_f.x - 1e-5f), std::min(std::max(vox_f.y, 0.0f), grid_size_f.y - 1e-5f),
               std::min(std::max(vox_f.z, 0.0f), grid_size_f.z - 1e-5f));
    const uint32_t vx = uint32_t(vox_f.x), vy = uint32_t(vox_f.y), vz = uint32_t(vox_f.z);
    const f3 dp = vox_f - f3(float(vx), float(vy), float(vz));
    float res = 0.0f;
    if (s.preset.interpolation_mode == 1 && vx < sx - 2 && vy < sy - 2 && vz < sz - 2 && vx > 0 && vy > 0 && vz > 0) {
      const uint32_t x0 = vx - 1u, y0 = vy - 1u, z0 = vz - 1u;
      const float f_dp[3] = {dp.x, dp.y, dp.z};
      float grid_part[64];
      for (uint32_t x = 0; x < 4; ++x)
        for (uint32_t y = 0; y < 4; ++y)
          for (uint32_t z = 0; z < 4; ++z)
            grid_part[z * 16 + y * 4 + x] = s.grid_data[off + (z0 + z) * sx * sy + (y0 + y) * sx + (x0 + x)];
      res = tricubicInterpolation(grid_part, f_dp);
    } else if (vx < sx - 1 && vy < sy - 1 && vz < sz - 1) {
      for (uint32_t i = 0; i < 2; i++)
        for (uint32_t j = 0; j < 2; j++)
          for (uint32_t k = 0; k < 2; k++) {
            float qx = (1 - dp.x + i * (2 * dp.x - 1));
            float qy = (1 - dp.y + j * (2 * dp.y - 1));
            float qz = (1 - dp.z + k * (2 * dp.z - 1));
            res += qx * qy * qz * s.grid_data[off + (vz + k) * sx * sy + (vy + j) * sx + (vx + i)];
          }
    } else {
      res += s.grid_data[off + vz * sx * sy + vy * sx + vx];
    }
    return res;
  }

  // BVHRT::box_intersects (BVH2Common.cpp:34-45)
  static f2 box_intersects(f3 min_pos, f3 max_pos, f3 origin, f3 dir) {
    f3 safe_dir(lm_sign(dir.x) * std::max(1e-9f, std::abs(dir.x)), lm_sign(dir.y) * std::max(1e-9f, std::abs(dir.y)),
                lm_sign(dir.z) * std::max(1e-9f, std::abs(dir.z)));
    f3 tMin = (min_pos - origin) / safe_dir;
    f3 tMax = (max_pos - origin) / safe_dir;
    f3 t1(std::min(tMin.x, tMax.x), std::min(tMin.y, tMax.y), std::min(tMin.z, tMax.z));
    f3 t2(std::max(tMin.x, tMax.x), std::max(tMin.y, tMax.y), std::max(tMin.z, tMax.z));
    return f2{std::max(t1.x, std::max(t1.y, t1.z)), std::min(t2.x, std::min(t2.y, t2.z))};
  }

  // BVHRT::IntersectSDFGrid (BVH2Common.cpp:1114-1182)
  void IntersectSDFGrid(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, Hit* pHit) {
    const uint32_t sdfId = s.geoms[geomId].offset[0];
    const float dir_len = length(ray_dir);
    const f3 dir = ray_dir / dir_len;
    const float EPS = 1e-5f;
    f2 tt = box_intersects(f3(-1, -1, -1), f3(1, 1, 1), ray_pos, dir);
    float t = std::max(tNear, tt.x);
    float tFar = tt.y;
    if (t > tFar) return;
    uint32_t iter = 0;
    float d = eval_distance_sdf_grid(sdfId, ray_pos + t * dir);
    while (iter < 1000 && d > EPS && t < tFar) {
      t += d + EPS;
      d = eval_distance_sdf_grid(sdfId, ray_pos + t * dir);
      iter++;
    }
    if (d > EPS) return;
    f3 p0 = ray_pos + t * dir;
    if (t > tNear && t < pHit->t) {
      f3 norm(0, 0, 1);
      if (need_normal()) {
        const float h = 0.001f;
        float ddx = (eval_distance_sdf_grid(sdfId, p0 + f3(h, 0, 0)) - eval_distance_sdf_grid(sdfId, p0 + f3(-h, 0, 0))) / (2 * h);
        float ddy = (eval_distance_sdf_grid(sdfId, p0 + f3(0, h, 0)) - eval_distance_sdf_grid(sdfId, p0 + f3(0, -h, 0))) / (2 * h);
        float ddz = (eval_distance_sdf_grid(sdfId, p0 + f3(0, 0, h)) - eval_distance_sdf_grid(sdfId, p0 + f3(0, 0, -h))) / (2 * h);
        norm = normalize(f3(ddx, ddy, ddz));
      }
      f3 n = normalize(matmul4x3(s.instances[instId].transformInvTransposed, norm));
      f2 e = encode_normal(n);
      pHit->t = t;
      pHit->primId = 0;
      pHit->instId = instId;
      pHit->geomId = geomId | (TYPE_SDF_GRID << SH_TYPE);
      pHit->coords[0] = 0;
      pHit->coords[1] = 0;
      pHit->coords[2] = e.x;
      pHit->coords[3] = e.y;
      if (s.preset.render_mode == 10) pHit->primId = iter;   // MULTI_RENDER_MODE_ST_ITERATIONS
    }
  }

  // eval_dist_trilinear_diff (BVH2Common.cpp:751-781)
  static f3 eval_dist_trilinear_diff(const float v[8], f3 dp) {
    float ddist_dx = -(1 - dp.y) * (1 - dp.z) * v[0] + -(1 - dp.y) * (dp.z) * v[1] + -(dp.y) * (1 - dp.z) * v[2] +
                     -(dp.y) * (dp.z) * v[3] + (1 - dp.y) * (1 - dp.z) * v[4] + (1 - dp.y) * (dp.z) * v[5] +
                     (dp.y) * (1 - dp.z) * v[6] + (dp.y) * (dp.z) * v[7];
    float ddist_dy = -(1 - dp.x) * (1 - dp.z) * v[0] + -(1 - dp.x) * (dp.z) * v[1] + (1 - dp.x) * (1 - dp.z) * v[2] +
                     (1 - dp.x) * (dp.z) * v[3] + -(dp.x) * (1 - dp.z) * v[4] + -(dp.x) * (dp.z) * v[5] +
                     (dp.x) * (1 - dp.z) * v[6] + (dp.x) * (dp.z) * v[7];
    float ddist_dz = -(1 - dp.x) * (1 - dp.y) * v[0] + (1 - dp.x) * (1 - dp.y) * v[1] + -(1 - dp.x) * (dp.y) * v[2] +
                     (1 - dp.x) * (dp.y) * v[3] + -(dp.x) * (1 - dp.y) * v[4] + (dp.x) * (1 - dp.y) * v[5] +
                     -(dp.x) * (dp.y) * v[6] + (dp.x) * (dp.y) * v[7];
    return f3(ddist_dx, ddist_dy, ddist_dz);
  }
  // step (BVH2Common.cpp:784-788)
  static float step(float edge0, float edge1, float x) {
    return std::min(std::max(0.f, (x - edge0) / (edge1 - edge0)), 1.f);
  }

  // BVH2Common.cpp:958-1028: smoothed normal of the ID32F_IRGB32F_IN layout (not transformed by
  // the instance, as in the reference)
  void SmoothedNormal(const SBSHeader& header, uint32_t nodeId, uint32_t data_offset, f3 voxelPos, uint32_t v_size,
                      float sz_inv, float d, f3 start_q, float f_near, f3 ray_dir, Hit* pHit) {
    const float beta = 0.5f;
    f3 dp = start_q + ((pHit->t - f_near) / d) * ray_dir;
    f3 normals[8];
    float values_n[8];
    for (int i = 0; i < 8; i++) normals[i] = f3(0, 0, 0);
    f3 nmq(0.5f, 0.5f, 0.5f);
    nmq.x = dp.x < 0.5f ? step(-beta, beta, dp.x) : step(-beta, beta, dp.x - 1.0f);
    nmq.y = dp.y < 0.5f ? step(-beta, beta, dp.y) : step(-beta, beta, dp.y - 1.0f);
    nmq.z = dp.z < 0.5f ? step(-beta, beta, dp.z) : step(-beta, beta, dp.z - 1.0f);
    const int bs = int(header.brick_size);
    const int vp0[3] = {dp.x < 0.5f ? int(voxelPos.x) - 1 : int(voxelPos.x), dp.y < 0.5f ? int(voxelPos.y) - 1 : int(voxelPos.y),
                        dp.z < 0.5f ? int(voxelPos.z) - 1 : int(voxelPos.z)};
    for (int i = 0; i < 8; i++) {
      const int VI[3] = {vp0[0] + ((i & 4) >> 2), vp0[1] + ((i & 2) >> 1), vp0[2] + (i & 1)};
      int BI[3];
      for (int a = 0; a < 3; a++) BI[a] = VI[a] >= 0 ? (VI[a] < bs ? 0 : 1) : -1;
      uint32_t neighborId = 0;
      neighborId += 3 * 3 * (BI[0] + 1);
      neighborId += 3 * (BI[1] + 1);
      neighborId += (BI[2] + 1);
      const f3 vpi = f3(float(VI[0]), float(VI[1]), float(VI[2]));
      f3 dVoxelPos = vpi - voxelPos;
      uint32_t neighbor_nodeId;
      f3 neighbor_voxelPos;
      if (neighborId == 9 + 3 + 1) {
        neighbor_nodeId = nodeId;
        neighbor_voxelPos = vpi;
      } else {
        neighbor_nodeId = s.sbs_data[data_offset + v_size * v_size * v_size + 8 + neighborId];
        neighbor_voxelPos = vpi - float(header.brick_size) * f3(float(BI[0]), float(BI[1]), float(BI[2]));
      }
      if (neighbor_nodeId != 0x80000000u) {   // INVALID_IDX, sdf_common_shared.h:30
        load_distance_values(neighbor_nodeId, neighbor_voxelPos, v_size, sz_inv, header, values_n);
        normals[i] = normalize(eval_dist_trilinear_diff(values_n, dp - dVoxelPos));
      }
    }
    f3 sm = ((1 - nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[0];
    sm = sm + ((1 - nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[1];
    sm = sm + ((1 - nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[2];
    sm = sm + ((1 - nmq.x) * (nmq.y) * (nmq.z)) * normals[3];
    sm = sm + ((nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[4];
    sm = sm + ((nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[5];
    sm = sm + ((nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[6];
    sm = sm + ((nmq.x) * (nmq.y) * (nmq.z)) * normals[7];
    f2 e = encode_normal(normalize(sm));
    pHit->coords[2] = e.x;
    pHit->coords[3] = e.y;
  }

  // BVH2Common.cpp:884-1112 (trilinear, or tricubic when the preset asks for it)
  void OctreeBrickIntersect(uint32_t type, f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                            uint32_t bvhNodeId, Hit* pHit) {
    const uint32_t sdfId = s.geoms[geomId].offset[0];
    const uint32_t primId = bvhNodeId;
    const uint32_t nodeId = primId + s.sbs_roots[sdfId];
    const SBSHeader header = s.sbs_headers[sdfId];
    const uint32_t v_size = header.brick_size + 2 * header.brick_pad + 1;
    const SBSNode& nd = s.sbs_nodes[nodeId];
    float px = nd.pos_xy >> 16, py = nd.pos_xy & 0x0000FFFF, pz = nd.pos_z_lod_size >> 16,
          sz = nd.pos_z_lod_size & 0x0000FFFF;
    float sz_inv = 2.0f / sz;
    float d = 2.0f / (sz * header.brick_size);
    f3 brick_min_pos = f3(-1, -1, -1) + sz_inv * f3(px, py, pz);
    f3 brick_max_pos = brick_min_pos + sz_inv * f3(1, 1, 1);
    f2 brick_fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), brick_min_pos, brick_max_pos);
    const float old_t = pHit->t;
    // step cap: a ray crosses at most 3*brick_size-2 voxels; only rounding-stalled marches hit it
    uint32_t steps = 0;
    const uint32_t max_steps = 4 * header.brick_size + 8;
    while (brick_fNearFar.x < brick_fNearFar.y && pHit->t == old_t && steps < max_steps) {
      steps++;
      if (cnt) cnt->sdf_voxels++;
      f3 hit_pos = ray_pos + brick_fNearFar.x * ray_dir;
      f3 local_pos = (hit_pos - brick_min_pos) * (0.5f * sz * header.brick_size);
      const float hi = header.brick_size - 1e-6f;
      f3 voxelPos(std::floor(std::min(std::max(1e-6f, local_pos.x), hi)), std::floor(std::min(std::max(1e-6f, local_pos.y), hi)),
                  std::floor(std::min(std::max(1e-6f, local_pos.z), hi)));
      f3 min_pos = brick_min_pos + d * voxelPos;
      f3 max_pos = min_pos + d * f3(1, 1, 1);
      float values[8];
      if (s.preset.interpolation_mode != 0) {
        float tricub_values[64];
        const float vmin3 = load_tricubic_distance_values(nodeId, voxelPos, v_size, sz_inv, header, tricub_values);
        const f2 fNF = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
        if (tNear < fNF.x && vmin3 <= 0.0f) {
          const f3 start_pos = ray_pos + fNF.x * ray_dir;
          const f3 start_q = (start_pos - min_pos) * (0.5f * sz * header.brick_size);
          const float qFar = (fNF.y - fNF.x) * (0.5f * sz * header.brick_size);
          TricubicLocalSurfaceIntersection(type, ray_dir, instId, geomId, tricub_values, primId, d, 0.0f, qFar, fNF,
                                           start_q, pHit);
        }
        brick_fNearFar.x += std::max(0.0f, fNF.y - brick_fNearFar.x) + 1e-6f;
        continue;
      }
      float vmin = load_distance_values(nodeId, voxelPos, v_size, sz_inv, header, values);
      f2 fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
      if (tNear < fNearFar.x && vmin <= 0.0f) {
        f3 start_pos = ray_pos + fNearFar.x * ray_dir;
        f3 start_q = (start_pos - min_pos) * (0.5f * sz * header.brick_size);
        float qFar = (fNearFar.y - fNearFar.x) * (0.5f * sz * header.brick_size);
        const float t_before = pHit->t;
        LocalSurfaceIntersection(type, ray_dir, instId, geomId, values, primId, d, 0.0f, qFar, fNearFar, start_q, pHit);
        if (s.preset.interpolation_mode == 0 && header.aux_data == LAYOUT_ID32F_IN && s.preset.normal_mode == 2 &&
            need_normal() && pHit->t != t_before)
          SmoothedNormal(header, nodeId, nd.data_offset, voxelPos, v_size, sz_inv, d, start_q, fNearFar.x, ray_dir, pHit);
        if (rec && pHit->t < t_before && (header.aux_data == LAYOUT_ID32F || header.aux_data == LAYOUT_ID32F_IN)) {
          rec->hit = 1;
          for (int i = 0; i < 8; i++) {
            const uint32_t x = uint32_t(voxelPos.x) + ((i & 4) >> 2), y = uint32_t(voxelPos.y) + ((i & 2) >> 1),
                           z = uint32_t(voxelPos.z) + (i & 1);
            const uint32_t vId = (x + header.brick_pad) * v_size * v_size + (y + header.brick_pad) * v_size +
                                 (z + header.brick_pad);
            rec->idx[i] = s.sbs_data[nd.data_offset + vId];
          }
          rec->start_q[0] = start_q.x; rec->start_q[1] = start_q.y; rec->start_q[2] = start_q.z;
          rec->dir[0] = ray_dir.x; rec->dir[1] = ray_dir.y; rec->dir[2] = ray_dir.z;
          rec->t_local = last_t_local;
          rec->t = pHit->t;
        }
      }
      brick_fNearFar.x += std::max(0.0f, fNearFar.y - brick_fNearFar.x) + 1e-6f;
    }
    if (pHit->t < old_t) {
      f3 pos = ray_pos + pHit->t * ray_dir;
      f3 dp = (pos - brick_min_pos) * (0.5f * sz);
      const float w[8] = {(1 - dp.x) * (1 - dp.y) * (1 - dp.z), (1 - dp.x) * (1 - dp.y) * (dp.z),
                          (1 - dp.x) * (dp.y) * (1 - dp.z),     (1 - dp.x) * (dp.y) * (dp.z),
                          (dp.x) * (1 - dp.y) * (1 - dp.z),     (dp.x) * (1 - dp.y) * (dp.z),
                          (dp.x) * (dp.y) * (1 - dp.z),         (dp.x) * (dp.y) * (dp.z)};
      if (header.aux_data == LAYOUT_DX_UV16) {
        const uint32_t vpi = 4 / header.bytes_per_value;
        const uint32_t t_off = nd.data_offset + (v_size * v_size * v_size + vpi - 1) / vpi;
        float u = w[0] * (1.5259022e-5f * float(s.sbs_data[t_off] >> 16));
        float v = w[0] * (1.5259022e-5f * float(s.sbs_data[t_off] & 0xFFFF));
        for (int i = 1; i < 8; ++i) {
          u = u + w[i] * (1.5259022e-5f * float(s.sbs_data[t_off + i] >> 16));
          v = v + w[i] * (1.5259022e-5f * float(s.sbs_data[t_off + i] & 0xFFFF));
        }
        pHit->coords[0] = u;
        pHit->coords[1] = v;
      } else if (header.aux_data == LAYOUT_DX_RGB8 || header.aux_data == LAYOUT_ID32F || header.aux_data == LAYOUT_ID32F_IN) {
        const bool rgb8 = header.aux_data == LAYOUT_DX_RGB8;
        const uint32_t vpi = 4 / header.bytes_per_value;
        const uint32_t t_off = rgb8 ? nd.data_offset + (v_size * v_size * v_size + vpi - 1) / vpi
                                    : nd.data_offset + v_size * v_size * v_size;
        f3 color;
        for (int i = 0; i < 8; ++i) {
          const uint32_t q = s.sbs_data[t_off + i];
          const f3 c = rgb8 ? f3(float((q >> 0) & 0xFF), float((q >> 8) & 0xFF), float((q >> 16) & 0xFF))
                            : f3(s.sbs_data_f[q + 0], s.sbs_data_f[q + 1], s.sbs_data_f[q + 2]);
          color = (i == 0) ? w[0] * c : color + w[i] * c;
        }
        const float k = rgb8 ? 1.0f : 255.0f;
        color = f3(std::min(std::max(0.0f, std::floor(k * color.x + 0.5f)), 255.0f),
                   std::min(std::max(0.0f, std::floor(k * color.y + 0.5f)), 255.0f),
                   std::min(std::max(0.0f, std::floor(k * color.z + 0.5f)), 255.0f));
        pHit->coords[0] = color.x + color.y / 256.0f;
        pHit->coords[1] = color.z;
      }
    }
  }

  // BVH2Common.cpp:2176-2363 with first_node/new_node of intersection_common.h:58-93: the
  // octree as [-1,1]^3, coordinates mirrored so the ray direction is positive (octant mask a),
  // children visited in parametric front-to-back order, the walk ending at the first leaf hit.
  static int FirstNode(f3 t0, f3 tm) {
    int answer = 0;
    if (t0.x > t0.y) {
      if (t0.x > t0.z) {
        if (t0.x > tm.y) answer |= 2;
        if (t0.x > tm.z) answer |= 1;
        return answer;
      }
    } else if (t0.y > t0.z) {
      if (tm.x < t0.y) answer |= 4;
      if (t0.y > tm.z) answer |= 1;
      return answer;
    }
    if (tm.x < t0.z) answer |= 4;
    if (tm.y < t0.z) answer |= 2;
    return answer;
  }
  static int NewNode(float txm, int x, float tym, int y, float tzm, int z) {
    return (txm < tym) ? (txm < tzm ? x : z) : (tym < tzm ? y : z);
  }

  void OctreeIntersect(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, Hit* pHit) {
    struct Elem { uint32_t nodeId, info, px, py, pz, size; };
    f3 rp = ray_pos, rd = ray_dir;
    uint32_t a = 0;
    if (ray_dir.x < 0) { rp.x *= -1; rd.x *= -1; a |= 4; }
    if (ray_dir.y < 0) { rp.y *= -1; rd.y *= -1; a |= 2; }
    if (ray_dir.z < 0) { rp.z *= -1; rd.z *= -1; a |= 1; }
    rp = rp * -1.0f;
    const f3 inv = SafeInverse(rd);
    const f3 T0 = rp * inv - inv, T1 = rp * inv + inv, L = T1 - T0;
    static const int nn[8][3] = {{4, 2, 1}, {5, 3, 8}, {6, 8, 3}, {7, 8, 8}, {8, 6, 5}, {8, 7, 8}, {8, 8, 7}, {8, 8, 8}};
    Elem stack[32];
    int top = 0;
    const float old_t = pHit->t;
    const uint32_t root = s.geoms[geomId].offset[0];
    stack[0] = Elem{root, 0, 0, 0, 0, 1};
    while (top >= 0) {
      Elem& e = stack[top];
      float d = 1.0f / float(e.size);
      const f3 pf(float(e.px), float(e.py), float(e.pz));
      const f3 t0 = T0 + d * pf * L;
      const f3 t1 = T0 + d * (pf + 1.0f) * L;
      if (std::min(std::min(t1.x, t1.y), t1.z) < tNear) { top--; continue; }
      bool leaf = e.info > 0;
      if (!leaf) {
        const float t = s.preset.fixed_lod > 0 ? 1.0f : std::max(t0.x, std::max(t0.y, t0.z));
        const float target = (std::pow(2.0f, s.preset.level_of_detail) - 0.01f) / t;
        if (float(e.size) > target) {
          e.info = s.multi_nodes[e.nodeId].data_offset;
          leaf = true;
          if (s.multi_nodes[e.nodeId].voxel_count == 0) { top--; continue; }
        }
      }
      if (leaf) {
        const uint32_t mask = e.size - 1;
        const uint32_t rx = (a & 4) ? (~e.px & mask) : e.px, ry = (a & 2) ? (~e.py & mask) : e.py,
                       rz = (a & 1) ? (~e.pz & mask) : e.pz;
        const float sz = 0.5f * float(e.size);
        d = 1.0f / sz;
        const f3 mn = f3(-1, -1, -1) + d * f3(float(rx), float(ry), float(rz));
        const f2 fnf{std::max(t0.x, std::max(t0.y, t0.z)), std::min(t1.x, std::min(t1.y, t1.z))};
        const f3 start_q = sz * ((ray_pos + fnf.x * ray_dir) - mn);
        const float qNear = std::max(0.0f, sz * (tNear - fnf.x));
        const float qFar = sz * (fnf.y - fnf.x);
        const MultiNode& n = s.multi_nodes[e.nodeId];
        for (uint32_t v = 0; v < n.voxel_count; v++) {
          float values[8];
          for (int i = 0; i < 8; i++) values[i] = s.multi_values[e.info + 8 * v + i];
          LocalSurfaceIntersection(9, ray_dir, instId, geomId, values, e.nodeId - root, d, qNear, qFar, fnf, start_q,
                                   pHit, (n.flags & 1u) != 0);
        }
        if (s.preset.render_mode == 24 && pHit->t < old_t) pHit->primId = n.voxel_count;
        else if (s.preset.render_mode == 25 && pHit->t < old_t) pHit->primId = n.flags;
        top = pHit->t < old_t ? -1 : top - 1;
      } else {
        Elem buf[4];
        int nb = 0;
        const f3 tm = 0.5f * (t0 + t1);
        int cur = FirstNode(t0, tm);
        const uint32_t children = s.multi_nodes[e.nodeId].children_offset;
        do {
          const uint32_t child = uint32_t(cur) ^ a;
          const MultiNode& c = s.multi_nodes[children + child];
          const bool is_node = c.children_offset > 0;
          if ((c.voxel_count > 0 || is_node) && nb < 4)
            buf[nb++] = Elem{children + child, is_node ? 0u : c.data_offset, 2 * e.px + ((cur & 4) ? 1u : 0u),
                             2 * e.py + ((cur & 2) ? 1u : 0u), 2 * e.pz + ((cur & 1) ? 1u : 0u), 2 * e.size};
          cur = NewNode((cur & 4) ? t1.x : tm.x, nn[cur][0], (cur & 2) ? t1.y : tm.y, nn[cur][1], (cur & 1) ? t1.z : tm.z,
                        nn[cur][2]);
        } while (cur < 8);
        for (int i = 0; i < nb; i++)
          if (top + i < 32) stack[top + i] = buf[nb - i - 1];
        top = std::min(top + nb - 1, 31);
      }
    }
  }

  // ---- COctreeV3 (BVH2Common.cpp:1637-2173) ----
  // COctreeV3_LoadDistanceValues{LeafGrid,LeafBitPack,LeafSlices} (BVH2Common.cpp:1637-1851). A
  // voxel outside the padded brick (asked for by the smoothed normal without padding, where the
  // reference reads outside the brick) counts as absent.
  float COctreeV3_LoadDistanceValues(const CoctG& h, uint32_t leafType, uint32_t brickOffset, int vx, int vy, int vz,
                                     uint32_t v_size, float sz_inv, uint32_t transform_code, float values[8]) const {
    const uint32_t* data = s.coct_data + h.base;
    const I4* mods = s.coct_mods + h.mods;
    const int pad = int(h.brick_pad), psz = int(h.brick_size + 2 * h.brick_pad);
    if (vx + pad < 0 || vy + pad < 0 || vz + pad < 0 || vx + pad >= psz || vy + pad >= psz || vz + pad >= psz)
      return 1e6f;
    auto dot4 = [](const I4& m, int x, int y, int z) { return m.x * x + m.y * y + m.z * z + m.w * 1; };
    const uint32_t rotIdx = transform_code & h.rot_mask;
    const uint32_t vals_per_int = 32 / h.bits_per_value;
    const uint32_t bits = h.bits_per_value;
    const uint32_t max_val = h.bits_per_value == 32 ? 0xFFFFFFFF : ((1u << bits) - 1);
    const float add_transform =
        h.sim_compression > 0
            ? 2.0f * 1.73205081f * sz_inv * (2.0f * (float(transform_code & h.add_mask) / float(h.add_mask)) - 1.0f)
            : 0.0f;
    if (leafType == 1) {   // COCTREE_LEAF_TYPE_GRID
      const unsigned off_3 = 0, off_4 = off_3 + 2, off_5 = off_4 + 8 * h.uv_size;
      const float min_val = -float(data[brickOffset + off_3 + 0]) / float(0xFFFFFFFFu) + add_transform;
      const float range = (float(data[brickOffset + off_3 + 1]) / float(0xFFFFFFFFu)) / max_val;
      const I4 m = mods[rotIdx];
      const uint32_t vId0 = uint32_t(dot4(m, vx + pad, vy + pad, vz + pad));
      const uint32_t ids[8] = {vId0,
                               vId0 + m.z,
                               vId0 + m.y,
                               vId0 + m.y + m.z,
                               vId0 + m.x,
                               vId0 + m.x + m.z,
                               vId0 + m.x + m.y,
                               vId0 + m.x + m.y + m.z};
      float vmin = 1e6f;
      for (int i = 0; i < 8; i++) {
        const uint32_t vId = ids[i];
        const uint32_t dist = (data[brickOffset + off_5 + vId / vals_per_int] >> (bits * (vId % vals_per_int))) & max_val;
        values[i] = min_val + range * dist;
        vmin = std::min(values[i], vmin);
      }
      return vmin;
    }
    if (leafType != 2 && leafType != 3) return 1e6f;
    const uint32_t p_size = h.brick_size + 2 * h.brick_pad;
    const uint32_t PFlagPos = uint32_t(dot4(mods[48 + rotIdx], vx + pad, vy + pad, vz + pad));
    if ((data[brickOffset + PFlagPos / 32] & (1u << (PFlagPos % 32))) == 0) return 1e6f;
    const uint32_t presence_flags_size_uints = (p_size * p_size * p_size + 32 - 1) / 32;
    const bool slices = leafType == 3;
    const uint32_t slice_distance_flags_uints = (v_size * v_size + 32 - 1) / 32;
    const uint32_t distance_flags_size_uints =
        slices ? v_size * slice_distance_flags_uints : (v_size * v_size * v_size + 32 - 1) / 32;
    const uint32_t distance_offsets_size_uints = slices ? (v_size + 2 - 1) / 2 : 0;
    const unsigned off_0 = 0, off_1 = off_0 + presence_flags_size_uints, off_2 = off_1 + distance_flags_size_uints;
    const unsigned off_3 = off_2 + distance_offsets_size_uints, off_4 = off_3 + 2, off_5 = off_4 + 8 * h.uv_size;
    const float min_val = -float(data[brickOffset + off_3 + 0]) / float(0xFFFFFFFFu) + add_transform;
    const float range = (float(data[brickOffset + off_3 + 1]) / float(0xFFFFFFFFu)) / max_val;
    for (int i = 0; i < 8; i++) {
      const int cx = vx + pad + ((i & 4) >> 2), cy = vy + pad + ((i & 2) >> 1), cz = vz + pad + (i & 1);
      const uint32_t voxelId = uint32_t(dot4(mods[rotIdx], cx, cy, cz));
      uint32_t sliceId = 0, localId = voxelId, flags = off_1, sliceOffset = 0;
      if (slices) {
        sliceId = voxelId / (v_size * v_size);
        localId = voxelId % (v_size * v_size);
        flags = off_1 + slice_distance_flags_uints * sliceId;
        sliceOffset = (data[brickOffset + off_2 + sliceId / 2] >> (16 * (sliceId % 2))) & ((1u << 16) - 1);
      }
      uint32_t b0 = data[brickOffset + flags];
      b0 = localId > 31 ? b0 : b0 & ((1u << localId) - 1);
      b0 = uint32_t(__builtin_popcount(b0));
      uint32_t b1 = localId > 32 ? data[brickOffset + flags + 1] & ((1u << (localId - 32)) - 1) : 0;
      b1 = uint32_t(__builtin_popcount(b1));
      const uint32_t vId0 = sliceOffset + b0 + b1;
      const uint32_t dist0 = (data[brickOffset + off_5 + vId0 / vals_per_int] >> (bits * (vId0 % vals_per_int))) & max_val;
      values[i] = min_val + range * dist0;
    }
    return -1.0f;
  }

  // COctreeV3_BrickIntersect (BVH2Common.cpp:1853-2003); the march stops after 4 bs + 8 voxels
  // (only marches whose 1e-5 step is lost to rounding get there; the reference would not end).
  void COctreeV3_BrickIntersect(const CoctG& h, f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                                uint32_t brickOffset, f3 p, float sz, uint32_t transform_code_leaf_type, Hit* pHit) {
    float values[8];
    const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
    const uint32_t leaf_type = transform_code_leaf_type & 3;
    const uint32_t transform_code = transform_code_leaf_type >> 2;
    const float sz_inv = 2.0f / sz;
    const float d = 2.0f / (sz * h.brick_size);
    const f3 brick_min_pos = f3(-1, -1, -1) + sz_inv * p;
    const f3 brick_max_pos = brick_min_pos + sz_inv * f3(1, 1, 1);
    f2 brick_fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), brick_min_pos, brick_max_pos);
    brick_fNearFar.x = std::max(tNear, brick_fNearFar.x);
    const float old_t = pHit->t;
    const float hi = h.brick_size - 1e-6f;
    uint32_t steps = 0;
    while (brick_fNearFar.x < brick_fNearFar.y && pHit->t == old_t && steps < 4 * h.brick_size + 8) {
      steps++;
      const f3 hit_pos = ray_pos + brick_fNearFar.x * ray_dir;
      const f3 local_pos = (hit_pos - brick_min_pos) * (0.5f * sz * h.brick_size);
      // LiteMath clamp(u, a, b) = min(max(a, u), b) (LiteMath.h:108)
      const f3 voxelPos(std::floor(std::min(std::max(1e-6f, local_pos.x), hi)),
                        std::floor(std::min(std::max(1e-6f, local_pos.y), hi)),
                        std::floor(std::min(std::max(1e-6f, local_pos.z), hi)));
      const f3 min_pos = brick_min_pos + d * voxelPos;
      const f3 max_pos = min_pos + d * f3(1, 1, 1);
      const int vx = int(voxelPos.x), vy = int(voxelPos.y), vz = int(voxelPos.z);
      const float vmin = COctreeV3_LoadDistanceValues(h, leaf_type, brickOffset, vx, vy, vz, v_size, sz_inv,
                                                      transform_code, values);
      const f2 fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
      if (tNear < fNearFar.x && vmin <= 0.0f) {
        const f3 start_pos = ray_pos + fNearFar.x * ray_dir;
        const f3 start_q = (start_pos - min_pos) * (0.5f * sz * h.brick_size);
        const float qFar = (fNearFar.y - fNearFar.x) * (0.5f * sz * h.brick_size);
        LocalSurfaceIntersection(TYPE_SDF_FRAME_OCTREE, ray_dir, instId, geomId, values, brickOffset, d, 0.0f, qFar,
                                 fNearFar, start_q, pHit);
        if (s.preset.normal_mode == 2 && need_normal() && pHit->t < old_t) {   // NORMAL_MODE_SDF_SMOOTHED
          const float beta = 0.5f;
          const f3 dp = start_q + ((pHit->t - fNearFar.x) / d) * ray_dir;
          f3 normals[8];
          float values_n[8];
          f3 nmq(0.5f, 0.5f, 0.5f);
          nmq.x = dp.x < 0.5f ? step(-beta, beta, dp.x) : step(-beta, beta, dp.x - 1.0f);
          nmq.y = dp.y < 0.5f ? step(-beta, beta, dp.y) : step(-beta, beta, dp.y - 1.0f);
          nmq.z = dp.z < 0.5f ? step(-beta, beta, dp.z) : step(-beta, beta, dp.z - 1.0f);
          const int p0x = int(dp.x < 0.5f ? voxelPos.x - 1 : voxelPos.x);
          const int p0y = int(dp.y < 0.5f ? voxelPos.y - 1 : voxelPos.y);
          const int p0z = int(dp.z < 0.5f ? voxelPos.z - 1 : voxelPos.z);
          for (int i = 0; i < 8; i++) {
            const int ix = p0x + ((i & 4) >> 2), iy = p0y + ((i & 2) >> 1), iz = p0z + (i & 1);
            const f3 dVoxelPos = f3(float(ix), float(iy), float(iz)) - voxelPos;
            const float vmin_n = COctreeV3_LoadDistanceValues(h, leaf_type, brickOffset, ix, iy, iz, v_size, sz_inv,
                                                              transform_code, values_n);
            if (vmin_n <= 0.0f) normals[i] = normalize(eval_dist_trilinear_diff(values_n, dp - dVoxelPos));
          }
          const f3 smoothed_norm = (1 - nmq.x) * (1 - nmq.y) * (1 - nmq.z) * normals[0] +
                                   (1 - nmq.x) * (1 - nmq.y) * (nmq.z) * normals[1] +
                                   (1 - nmq.x) * (nmq.y) * (1 - nmq.z) * normals[2] +
                                   (1 - nmq.x) * (nmq.y) * (nmq.z) * normals[3] +
                                   (nmq.x) * (1 - nmq.y) * (1 - nmq.z) * normals[4] +
                                   (nmq.x) * (1 - nmq.y) * (nmq.z) * normals[5] +
                                   (nmq.x) * (nmq.y) * (1 - nmq.z) * normals[6] +
                                   (nmq.x) * (nmq.y) * (nmq.z) * normals[7];
          const f2 en = encode_normal(normalize(matmul4x3(s.instances[instId].transformInvTransposed, smoothed_norm)));
          pHit->coords[2] = en.x;
          pHit->coords[3] = en.y;
        }
      }
      brick_fNearFar.x += std::max(0.0f, fNearFar.y - brick_fNearFar.x) + 1e-5f;
    }
    if (pHit->t < old_t && h.uv_size > 0) {
      const uint32_t* data = s.coct_data + h.base;
      const uint32_t p_size = h.brick_size + 2 * h.brick_pad;
      const uint32_t slice_distance_flags_uints = (v_size * v_size + 32 - 1) / 32;
      const uint32_t off_1 = (p_size * p_size * p_size + 32 - 1) / 32;
      const uint32_t off_2 = off_1 + v_size * slice_distance_flags_uints;
      const uint32_t off_3 = off_2 + (v_size + 2 - 1) / 2;
      const uint32_t off_4 = off_3 + 2;
      const f3 pos = ray_pos + pHit->t * ray_dir;
      const f3 dp = (pos - brick_min_pos) * (0.5f * sz);
      const uint32_t* tc = data + brickOffset + off_4;
      pHit->coords[0] = (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[0] >> 16)) +
                        (1 - dp.x) * (1 - dp.y) * (dp.z) * (1.5259022e-5f * float(tc[1] >> 16)) +
                        (1 - dp.x) * (dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[2] >> 16)) +
                        (1 - dp.x) * (dp.y) * (dp.z) * (1.5259022e-5f * float(tc[3] >> 16)) +
                        (dp.x) * (1 - dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[4] >> 16)) +
                        (dp.x) * (1 - dp.y) * (dp.z) * (1.5259022e-5f * float(tc[5] >> 16)) +
                        (dp.x) * (dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[6] >> 16)) +
                        (dp.x) * (dp.y) * (dp.z) * (1.5259022e-5f * float(tc[7] >> 16));
      pHit->coords[1] = (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[0] & 0xFFFF)) +
                        (1 - dp.x) * (1 - dp.y) * (dp.z) * (1.5259022e-5f * float(tc[1] & 0xFFFF)) +
                        (1 - dp.x) * (dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[2] & 0xFFFF)) +
                        (1 - dp.x) * (dp.y) * (dp.z) * (1.5259022e-5f * float(tc[3] & 0xFFFF)) +
                        (dp.x) * (1 - dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[4] & 0xFFFF)) +
                        (dp.x) * (1 - dp.y) * (dp.z) * (1.5259022e-5f * float(tc[5] & 0xFFFF)) +
                        (dp.x) * (dp.y) * (1 - dp.z) * (1.5259022e-5f * float(tc[6] & 0xFFFF)) +
                        (dp.x) * (dp.y) * (dp.z) * (1.5259022e-5f * float(tc[7] & 0xFFFF));
    }
  }

  // COctreeIntersectV3 (BVH2Common.cpp:2005-2173), starting at the BLAS box's node; pushes past
  // the 32-entry stack are dropped (the reference writes past its array).
  void COctreeIntersectV3(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, uint32_t bvhNodeId,
                          Hit* pHit) {
    struct OTStackElement { uint32_t nodeId, info, p_size_x, p_size_y; };
    const Geom& g = s.geoms[geomId];
    const CoctG& h = s.coct_geoms[g.offset[0]];
    const uint32_t* data = s.coct_data + h.base;
    f3 pos_ray_pos = ray_pos, pos_ray_dir = ray_dir;
    uint32_t a = 0;
    if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
    if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
    if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
    const Node& box = s.orig_nodes[g.offset[1] + bvhNodeId];
    const f3 min_pos = box.boxMin, max_pos = box.boxMax;
    const uint32_t start_sz = uint32_t(2.0f / (max_pos.x - min_pos.x));
    const f3 spf = float(start_sz) * 0.5f * (min_pos - f3(-1, -1, -1));
    const uint32_t spo[3] = {uint32_t(spf.x), uint32_t(spf.y), uint32_t(spf.z)};
    uint32_t p_mask = start_sz - 1;
    const uint32_t start_p[3] = {(a & 4) ? (~spo[0] & p_mask) : spo[0], (a & 2) ? (~spo[1] & p_mask) : spo[1],
                                 (a & 1) ? (~spo[2] & p_mask) : spo[2]};
    pos_ray_pos = pos_ray_pos * -1.0f;
    const f3 inv = SafeInverse(pos_ray_dir);
    const f3 _t0 = pos_ray_pos * inv - inv, _t1 = pos_ray_pos * inv + inv, _l = _t1 - _t0;
    static const int nn[8][3] = {{4, 2, 1}, {5, 3, 8}, {6, 8, 3}, {7, 8, 8}, {8, 6, 5}, {8, 7, 8}, {8, 8, 7}, {8, 8, 8}};
    OTStackElement stack[32];
    int top = 0;
    const float old_t = pHit->t;
    stack[0] = OTStackElement{box.leftOffset, 0, (start_p[0] << 16) | start_p[1], (start_p[2] << 16) | start_sz};
    while (top >= 0) {
      const uint32_t level_sz = stack[top].p_size_y & 0xFFFF;
      const uint32_t px = stack[top].p_size_x >> 16, py = stack[top].p_size_x & 0xFFFF, pz = stack[top].p_size_y >> 16;
      const float d = 1.0f / float(level_sz);
      const f3 p_f = f3(float(px), float(py), float(pz));
      const f3 t0 = _t0 + d * p_f * _l;
      const f3 t1 = _t0 + d * (p_f + 1.0f) * _l;
      if (std::min(std::min(t1.x, t1.y), t1.z) < tNear) { top--; continue; }
      bool is_leaf = (stack[top].info & 3) != 0;
      if (h.lods > 0 && !is_leaf) {
        const float t = s.preset.fixed_lod > 0 ? 1.0f : std::max(t0.x, std::max(t0.y, t0.z));
        const float target_lod_size = (std::pow(2.0f, s.preset.level_of_detail) - 0.01f) / t;
        if (float(level_sz * h.brick_size) > target_lod_size) {
          const uint32_t lodLeafId = data[stack[top].nodeId + 1];
          const uint32_t type = (data[stack[top].nodeId + 0] >> 30) & 3;
          const uint32_t infoOffset = stack[top].nodeId + 1 + h.trans_off;
          stack[top].nodeId = lodLeafId;
          stack[top].info = (data[infoOffset] << 2) | type;
          is_leaf = true;
          if (type == 0) { top--; continue; }
        }
      }
      if (is_leaf) {
        p_mask = level_sz - 1;
        const uint32_t rx = (a & 4) ? (~px & p_mask) : px, ry = (a & 2) ? (~py & p_mask) : py,
                       rz = (a & 1) ? (~pz & p_mask) : pz;
        const uint32_t leafOffset = (stack[top].nodeId & h.idx_mask) >> h.idx_sh;
        COctreeV3_BrickIntersect(h, ray_pos, ray_dir, tNear, instId, geomId, leafOffset,
                                 f3(float(rx), float(ry), float(rz)), float(level_sz), stack[top].info, pHit);
        if (pHit->t < old_t) {
          top = -1;
          if (s.preset.render_mode == 20) pHit->primId = uint32_t(std::log2(float(level_sz * h.brick_size)));
        } else {
          top--;
        }
      } else {
        OTStackElement tmp_buf[4];
        int buf_top = 0;
        const f3 tm = 0.5f * (t0 + t1);
        int currNode = FirstNode(t0, tm);
        do {
          const uint32_t childrenInfo = data[stack[top].nodeId + 0];
          const uint32_t childNode = uint32_t(currNode) ^ a;
          if ((childrenInfo & (1u << childNode)) > 0 && buf_top < 4) {
            const uint32_t localChildPos = h.lods + uint32_t(__builtin_popcount(childrenInfo & ((1u << childNode) - 1)));
            const uint32_t childPos = 1 + h.uints_per_child_info * localChildPos;
            const uint32_t childOffset = data[stack[top].nodeId + childPos];
            const uint32_t childType = (childrenInfo >> (8 + 2 * childNode)) & 3;
            const uint32_t childInfoOffset = stack[top].nodeId + childPos + h.trans_off;
            const uint32_t childInfo = (data[childInfoOffset] << 2) | childType;
            tmp_buf[buf_top] = OTStackElement{
                childOffset, childInfo,
                (stack[top].p_size_x << 1) | (((uint32_t(currNode) & 4) << (16 - 2)) | ((uint32_t(currNode) & 2) >> 1)),
                (stack[top].p_size_y << 1) | ((uint32_t(currNode) & 1) << 16)};
            buf_top++;
          }
          currNode = NewNode((currNode & 4) ? t1.x : tm.x, nn[currNode][0], (currNode & 2) ? t1.y : tm.y,
                             nn[currNode][1], (currNode & 1) ? t1.z : tm.z, nn[currNode][2]);
        } while (currNode < 8);
        for (int i = 0; i < buf_top; i++)
          if (top + i < 32) stack[top + i] = tmp_buf[buf_top - i - 1];
        top = std::min(top + buf_top - 1, 31);
      }
    }
  }

  // DAGIntersect (BVH2Common.cpp:2365-2557). The transpositions of the brick and node-grid
  // lattices (m_SdfDAGTranspositions) are the index maps of the rotations, read through the
  // modifiers: T_r[x v^2 + y v + z] = dot(M_r, (x, y, z, 1)). Pushes past 32 entries are dropped;
  // a brick march stops after 4 bs + 8 voxels, a child march after 256 cells per grid step.
  void DAGIntersect(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, Hit* pHit) {
    struct OTStackElement { uint32_t nodeId, info, p_size_x, p_size_y; };
    const Geom& gm = s.geoms[geomId];
    const DagG& header = s.dag_geoms[gm.offset[1]];
    const I4* bricks_T = s.dag_mods + header.mods;
    const I4* branch_T = s.dag_mods + header.mods + 48;
    const uint32_t type = TYPE_SDF_DAG;
    const f3 ray_dir_inv = SafeInverse(ray_dir);
    OTStackElement stack[32];
    int top = 0;
    f2 fNearFar{0.0f, 0.0f};
    float old_t = pHit->t;
    stack[top] = OTStackElement{gm.offset[0], 0, 0, 1};
    while (top >= 0) {
      const OTStackElement curElem = stack[top];
      top--;
      const DagNode node = s.dag_nodes[curElem.nodeId];
      const uint32_t level_sz = curElem.p_size_y & 0xFFFF;
      const float level_sz_f = float(level_sz);
      const f3 p(float(curElem.p_size_x >> 16), float(curElem.p_size_x & 0xFFFF), float(curElem.p_size_y >> 16));
      const float sz_inv = 2.0f / level_sz_f;
      const f3 brick_min_pos = f3(-1, -1, -1) + sz_inv * p;
      const f3 brick_max_pos = brick_min_pos + sz_inv * f3(1, 1, 1);
      f2 brick_fNearFar = RayBoxIntersection2(ray_pos, ray_dir_inv, brick_min_pos, brick_max_pos);
      brick_fNearFar.x = std::max(tNear, brick_fNearFar.x);
      if (brick_fNearFar.x >= brick_fNearFar.y) continue;
      bool is_leaf = node.children_edges_offset == 0;
      if (!is_leaf) {
        const float t = s.preset.fixed_lod > 0 ? 1.0f : brick_fNearFar.x;
        const float target_lod_size = (std::pow(2.0f, s.preset.level_of_detail) - 0.01f) / t;
        is_leaf = float(level_sz) > target_lod_size;
      }
      if (is_leaf) {
        if ((node.voxel_count_flags & 0xFFFFu) != 0 && node.data_edges_offset > 0) {
          const float sz = level_sz_f;
          const float d = 2.0f / (sz * header.brick_size);
          const bool is_surface = ((node.voxel_count_flags >> 16) & 0x1) > 0;
          const uint32_t surface_count = node.voxel_count_flags & 0xFFFFu;
          const uint32_t primId = node.channels_edge.child_offset;
          uint32_t steps = 0;
          while (brick_fNearFar.x < brick_fNearFar.y && pHit->t == old_t && steps < 4 * header.brick_size + 8) {
            steps++;
            const f3 hit_pos = ray_pos + brick_fNearFar.x * ray_dir;
            const f3 local_pos = (hit_pos - brick_min_pos) * (0.5f * sz * header.brick_size);
            const float hi = header.brick_size - 1e-6f;
            const f3 voxelPos(std::floor(std::min(std::max(1e-6f, local_pos.x), hi)),
                              std::floor(std::min(std::max(1e-6f, local_pos.y), hi)),
                              std::floor(std::min(std::max(1e-6f, local_pos.z), hi)));
            const f3 min_pos = brick_min_pos + d * voxelPos;
            const f3 max_pos = min_pos + d * f3(1, 1, 1);
            uint32_t s_id = 0;
            while (s_id < surface_count && pHit->t == old_t) {
              const DagDataEdge& e = s.dag_data_edges[node.data_edges_offset + s_id];
              const uint32_t rotIdx = s.rot_add[curElem.info * 48 + e.rotation_id];
              const I4& T = bricks_T[rotIdx];
              float values[8];
              float vmin = 1.0f;
              float vmax = is_surface ? -1.0f : 1.0f;
              for (int i = 0; i < 8; i++) {
                const int px = int(voxelPos.x) + ((i & 4) >> 2), py = int(voxelPos.y) + ((i & 2) >> 1),
                          pz = int(voxelPos.z) + (i & 1);
                const uint32_t vId0 = uint32_t(T.x * px + T.y * py + T.z * pz + T.w);
                const float val = s.dag_distances[e.data_offset + vId0] + e.add;
                values[i] = val - s.preset.compact_sdf_eps * d;
                vmin = std::min(vmin, val);
                vmax = std::max(vmax, val);
              }
              fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
              f3 start_q(0, 0, 0);
              if (fNearFar.x < fNearFar.y && vmin <= 0.0f && vmax >= 0.0f) {
                const f3 start_pos = ray_pos + fNearFar.x * ray_dir;
                const float vox_sz = (0.5f * sz * header.brick_size);
                start_q = vox_sz * (start_pos - min_pos);
                const float qNear = std::max(0.0f, vox_sz * (tNear - fNearFar.x));
                const float qFar = vox_sz * (fNearFar.y - fNearFar.x);
                LocalSurfaceIntersection(type, ray_dir, instId, geomId, values, primId, d, qNear, qFar, fNearFar,
                                         start_q, pHit, is_surface);
              }
              // channel render modes: quotients turned by the channel rotation (:2492-2505), with the
              // transforms of the brick_size + 1 lattice (BVH2Common_host.cpp:1027)
              if (pHit->t != old_t && (s.preset.render_mode == 23 || s.preset.render_mode == 27)) {
                const RotTransforms& rt = rot_transforms(header.brick_size + 1);
                const uint32_t channelRotId = s.rot_add[curElem.info * 48 + node.channels_edge.rotation_id];
                const float t = (pHit->t - fNearFar.x) / d;
                const f3 dp = start_q + t * ray_dir;
                f3 dp_rot = rt.apply(int(channelRotId), dp);
                dp_rot = f3(std::min(std::max(65535.0f * dp_rot.x, 0.0f), 65535.0f),
                            std::min(std::max(65535.0f * dp_rot.y, 0.0f), 65535.0f),
                            std::min(std::max(65535.0f * dp_rot.z, 0.0f), 65535.0f));
                const int qx = int(dp_rot.x), qy = int(dp_rot.y), qz = int(dp_rot.z);
                const uint32_t w0 = uint32_t((qy << 16) | qx), w1 = uint32_t(qz);
                std::memcpy(&pHit->coords[0], &w0, 4);
                std::memcpy(&pHit->coords[1], &w1, 4);
              }
              s_id++;
            }
            brick_fNearFar.x += std::max(0.0f, fNearFar.y - brick_fNearFar.x) + 1e-5f;
            if (pHit->t < old_t) top = -1;
          }
        }
      } else {
        const int v_sz = int(header.node_grid_size);
        const float t_neg = (brick_fNearFar.y + tNear + 1);
        const f3 neg_pos = ray_pos + t_neg * ray_dir;
        const f3 neg_dir = ray_dir * -1.0f;
        const f3 neg_dir_inv = ray_dir_inv * -1.0f;
        f2 neg_brick_fNearFar{t_neg - brick_fNearFar.y, t_neg - brick_fNearFar.x};
        uint32_t steps = 0;
        while (neg_brick_fNearFar.x < neg_brick_fNearFar.y && steps < uint32_t(256 * v_sz)) {
          steps++;
          const float d = 2.0f / (level_sz_f * v_sz);
          const f3 hit_pos = neg_pos + neg_brick_fNearFar.x * neg_dir;
          const f3 local_pos = (hit_pos - brick_min_pos) * (0.5f * level_sz_f * v_sz);
          const float hi = v_sz - 1e-6f;
          const f3 voxelPos(std::floor(std::min(std::max(1e-6f, local_pos.x), hi)),
                            std::floor(std::min(std::max(1e-6f, local_pos.y), hi)),
                            std::floor(std::min(std::max(1e-6f, local_pos.z), hi)));
          const I4& T = branch_T[curElem.info];
          const uint32_t voxId = uint32_t(T.x * int(voxelPos.x) + T.y * int(voxelPos.y) + T.z * int(voxelPos.z) + T.w);
          const f3 min_pos = brick_min_pos + d * voxelPos;
          const f3 max_pos = min_pos + d * f3(1, 1, 1);
          fNearFar = RayBoxIntersection2(neg_pos, neg_dir_inv, min_pos, max_pos);
          const DagChildEdge edge = s.dag_child_edges[node.children_edges_offset + voxId];
          if (tNear < fNearFar.x && edge.child_offset > 0 && top + 1 < 32) {
            top++;
            stack[top].nodeId = edge.child_offset;
            stack[top].info = s.rot_add[curElem.info * 48 + edge.rotation_id];
            stack[top].p_size_x = curElem.p_size_x * v_sz + ((uint32_t(voxelPos.x) << 16) | uint32_t(voxelPos.y));
            stack[top].p_size_y = curElem.p_size_y * v_sz + ((uint32_t(voxelPos.z) << 16) | 0);
          }
          neg_brick_fNearFar.x += std::max(0.0f, fNearFar.y - neg_brick_fNearFar.x) + 0.01f * d;
        }
      }
    }
  }

  // ---- SCom2 (BVH2Common.cpp:2560-2818, sdf/scom2/scom_shared.h:183-258) ----
  struct NodeHead { uint32_t base_type, children_types, base_links_end, children_active; };
  static NodeHead unpack_node_head(const SC2Header& h, uint32_t node0, uint32_t node1) {
    NodeHead r;
    r.base_type = node0 & 0xFFu;
    if (h.has_multi_nodes != 0) r.children_types = (node0 >> h.children_types_shift) & h.children_types_mask;
    else r.children_types = (node1 >> h.children_types_shift) & h.children_types_mask;
    if (h.has_multi_nodes != 0) r.base_links_end = (node0 >> h.base_reference_shift) & h.reference_mask;
    else r.base_links_end = (r.base_type == 0 ? 0 : (h.has_surfaces == 0 ? h.has_channels : 1 + h.has_channels));
    r.children_active = (node0 >> h.children_active_bits_shift) & h.children_active_bits_mask;
    return r;
  }
  struct DataEdge { uint32_t data_offset, rotation_id; float add; };
  static DataEdge unpack_data_edge(const SC2Header& h, float max_val, uint32_t edge0, uint32_t edge1) {
    DataEdge e;
    const bool is_unique = (edge0 & h.unique_brick_prefix) == h.unique_brick_prefix;
    const bool no_add = is_unique || (h.child_add_mask == 0);
    if (h.max_val > 0)
      e.add = no_add ? 0.0f : h.max_val * (2 * ((edge0 >> h.child_add_shift) & h.child_add_mask) / float(h.child_add_mask) - 1);
    else
      e.add = no_add ? 0.0f : max_val * (2 * ((edge0 >> h.child_add_shift) & h.child_add_mask) / float(h.child_add_mask) - 1);
    e.rotation_id = is_unique ? 0 : (edge0 >> h.child_rot_shift) & h.child_rot_mask;
    e.data_offset = is_unique ? edge1 & h.unique_brick_offset_mask : edge1 & h.child_offset_mask;
    return e;
  }
  struct ChildEdge { uint32_t child_offset, rotation_id; };
  static ChildEdge unpack_child_edge(const SC2Header& h, uint32_t edge0, uint32_t edge1) {
    return ChildEdge{edge1 & h.node_offset_mask, (edge0 >> h.child_rot_shift) & h.child_rot_mask};
  }
  static uint32_t get_links_end(const SC2Header& h, uint32_t child_n, uint32_t node1, uint32_t node2) {
    if (h.has_multi_nodes != 0) {
      const uint32_t sh = child_n * h.reference_bits;
      return sh >= 32 ? ((node2 >> sh % 32) & h.reference_mask) : ((node1 >> sh) & h.reference_mask);
    } else if (h.has_surfaces > 0) {
      const uint32_t base_links_end = ((node1 & 0xFFu) == 0 ? 0 : 1);
      const uint32_t ch_active = (node1 >> h.children_active_bits_shift) & h.children_active_bits_mask;
      const uint32_t ch_n = uint32_t(__builtin_popcount(ch_active & ((1u << (1 + child_n)) - 1)));
      return (1 + h.has_channels) * (base_links_end) + ch_n;
    } else {
      const uint32_t base_links_end = ((node1 & 0xFFu) == 0 ? 0 : 1);
      const uint32_t ch_active = (node1 >> h.children_active_bits_shift) & h.children_active_bits_mask;
      const uint32_t ch_n = uint32_t(__builtin_popcount(ch_active & ((1u << (1 + child_n)) - 1)));
      return (h.has_channels) * (base_links_end + ch_n);
    }
  }

  // SCom2Intersect. Pushes past the 32-entry stack are dropped; the brick march stops after
  // 4 bs + 8 voxels (only marches whose 1e-5 step is lost to rounding get there).
  void SCom2Intersect(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, Hit* pHit) {
    struct OTStackElement { uint32_t nodeId, info, p_size_x, p_size_y; };
    const uint32_t type = TYPE_SCOM2;
    const uint32_t hidx = s.geoms[geomId].offset[1];
    const SC2Header header = s.scom2_headers[hidx];
    const uint32_t* data = s.scom2_data;
    const I4* mods = s.scom2_mods + hidx * 144;
    f3 pos_ray_pos = ray_pos, pos_ray_dir = ray_dir;
    uint32_t a = 0;
    if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
    if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
    if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
    pos_ray_pos = pos_ray_pos * -1.0f;
    const f3 pos_ray_dir_inv = SafeInverse(pos_ray_dir);
    const f3 _t0 = pos_ray_pos * pos_ray_dir_inv - pos_ray_dir_inv;
    const f3 _t1 = pos_ray_pos * pos_ray_dir_inv + pos_ray_dir_inv;
    const f3 _l = _t1 - _t0;
    static const int nn[8][3] = {{4, 2, 1}, {5, 3, 8}, {6, 8, 3}, {7, 8, 8}, {8, 6, 5}, {8, 7, 8}, {8, 8, 7}, {8, 8, 8}};
    OTStackElement stack[32];
    int top = 0;
    stack[0] = OTStackElement{header.nodes_arr_offset + header.root_node_off, 0, 0, 1};
    while (top >= 0) {
      OTStackElement curElem = stack[top];
      top--;
      const uint32_t level_sz = curElem.p_size_y & 0xFFFF;
      const float level_sz_f = float(level_sz);
      const uint32_t px = curElem.p_size_x >> 16, py = curElem.p_size_x & 0xFFFF, pz = curElem.p_size_y >> 16;
      float d = 1.0f / float(level_sz);
      const f3 p_f = f3(float(px), float(py), float(pz));
      const f3 t0 = _t0 + d * p_f * _l;
      const f3 t1 = _t0 + d * (p_f + 1.0f) * _l;
      if (std::min(t1.x, std::min(t1.y, t1.z)) < tNear) continue;
      bool is_leaf = curElem.info > 0xFF;
      bool is_surface = curElem.info > 0xFFFFFFu;
      uint32_t surface_count = (curElem.info >> 8) & 0xFFFF;
      if (!is_leaf) {
        const float t = s.preset.fixed_lod > 0 ? 1.0f : std::max(t0.x, std::max(t0.y, t0.z));
        const float target_lod_size = (std::pow(2.0f, s.preset.level_of_detail) - 0.01f) / t;
        is_leaf = float(level_sz) > target_lod_size;
        if (is_leaf) {
          const NodeHead node = unpack_node_head(header, data[curElem.nodeId], data[curElem.nodeId + header.links_offset - 1]);
          if (node.base_type == 0) continue;
          is_surface = (node.base_type & 3) == 3;
          curElem.nodeId = curElem.nodeId + header.links_offset;
          surface_count = node.base_links_end - header.has_channels;
        }
      }
      if (is_leaf) {
        const float sz_inv = 2.0f / level_sz_f;
        const uint32_t p_mask = level_sz - 1;
        const uint32_t rp[3] = {(a & 4) ? (~px & p_mask) : px, (a & 2) ? (~py & p_mask) : py, (a & 1) ? (~pz & p_mask) : pz};
        const f3 brick_min_pos = f3(-1, -1, -1) + sz_inv * f3(float(rp[0]), float(rp[1]), float(rp[2]));
        float tmin = std::max(t0.x, std::max(t0.y, t0.z));
        const float tmax = std::min(t1.x, std::min(t1.y, t1.z));
        const uint32_t rotId = curElem.info & 0xFF;
        const float max_val = 1.7320508f * 2.0f / level_sz_f;   // get_max_sdf_val
        d = 2.0f / (level_sz_f * header.brick_size);
        const float old_t = pHit->t;
        uint32_t steps = 0;
        while (tmin < tmax && pHit->t == old_t && steps < 4 * header.brick_size + 8) {
          steps++;
          const f3 hit_pos = ray_pos + tmin * ray_dir;
          const f3 local_pos = (hit_pos - brick_min_pos) * (0.5f * level_sz_f * header.brick_size);
          const float hi = header.brick_size - 1e-6f;
          const f3 voxelPos(std::floor(std::min(std::max(1e-6f, local_pos.x), hi)),
                            std::floor(std::min(std::max(1e-6f, local_pos.y), hi)),
                            std::floor(std::min(std::max(1e-6f, local_pos.z), hi)));
          const f3 min_pos = brick_min_pos + d * voxelPos;
          const f3 max_pos = min_pos + d * f3(1, 1, 1);
          uint32_t s_id = 0;
          const f2 fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
          while (fNearFar.x < fNearFar.y && s_id < surface_count && pHit->t == old_t) {
            const uint32_t link_0 = curElem.nodeId + s_id * header.uints_per_link;
            const uint32_t link_1 = link_0 + header.child_offset_off;
            const DataEdge de = unpack_data_edge(header, max_val, data[link_0], data[link_1]);
            const uint32_t offset = header.bricks_arr_offset + header.bricks_step * de.data_offset;
            const uint32_t rotIdx = s.rot_add[rotId * 48 + de.rotation_id];
            float vmin = 1.0f;
            float vmax = is_surface ? -1.0f : 1.0f;
            float values[8];
            for (int i = 0; i < 8; i++) {
              const int pI[3] = {int(voxelPos.x) + ((i & 4) >> 2), int(voxelPos.y) + ((i & 2) >> 1), int(voxelPos.z) + (i & 1)};
              const I4& m = mods[rotIdx];
              const uint32_t vId0 = uint32_t(m.x * pI[0] + m.y * pI[1] + m.z * pI[2] + m.w * 1);
              const uint32_t p_val = data[offset + vId0 / header.values_per_uint];
              const uint32_t p_off = (vId0 % header.values_per_uint) * header.bits_per_value;
              const float val = max_val * (2.0f * ((p_val >> p_off) & header.value_mask) / float(header.value_mask) - 1) + de.add;
              values[i] = val - s.preset.compact_sdf_eps * d;
              vmin = std::min(vmin, val);
              vmax = std::max(vmax, val);
            }
            f3 start_q(0, 0, 0);
            if (vmin <= 0.0f && vmax >= 0.0f) {
              const f3 start_pos = ray_pos + fNearFar.x * ray_dir;
              const float vox_sz = (0.5f * level_sz_f * header.brick_size);
              start_q = vox_sz * (start_pos - min_pos);
              const float qNear = std::max(0.0f, vox_sz * (tNear - fNearFar.x));
              const float qFar = vox_sz * (fNearFar.y - fNearFar.x);
              LocalSurfaceIntersection(type, ray_dir, instId, geomId, values, curElem.nodeId, d, qNear, qFar, fNearFar,
                                       start_q, pHit, is_surface);
            }
            if (pHit->t != old_t && header.has_channels > 0 && s.preset.render_mode != 5) {
              const uint32_t ch_off = curElem.nodeId + surface_count * header.uints_per_link;
              const ChildEdge ce = unpack_child_edge(header, data[ch_off], data[ch_off + header.child_offset_off]);
              pHit->primId = ce.child_offset;
              // channel render modes: quotients turned by the channel rotation (:2731-2747), with the
              // transforms of the brick_size + 1 lattice (BVH2Common_host.cpp:1141)
              if (s.preset.render_mode == 23 || s.preset.render_mode == 27) {
                const RotTransforms& rt = rot_transforms(header.brick_size + 1);
                const uint32_t channelRotId = s.rot_add[rotId * 48 + ce.rotation_id];
                const float t = (pHit->t - fNearFar.x) / d;
                const f3 dp = start_q + t * ray_dir;
                f3 dp_rot = rt.apply(int(channelRotId), dp);
                dp_rot = f3(std::min(std::max(65535.0f * dp_rot.x, 0.0f), 65535.0f),
                            std::min(std::max(65535.0f * dp_rot.y, 0.0f), 65535.0f),
                            std::min(std::max(65535.0f * dp_rot.z, 0.0f), 65535.0f));
                const int qx = int(dp_rot.x), qy = int(dp_rot.y), qz = int(dp_rot.z);
                const uint32_t w0 = uint32_t((qy << 16) | qx), w1 = uint32_t(qz);
                std::memcpy(&pHit->coords[0], &w0, 4);
                std::memcpy(&pHit->coords[1], &w1, 4);
              }
            }
            s_id++;
          }
          tmin += std::max(0.0f, fNearFar.y - tmin) + 1e-5f;
          if (pHit->t < old_t) {
            top = -1;
            if (s.preset.render_mode == 20) pHit->primId = uint32_t(std::log2(float(level_sz * header.brick_size)));
          }
        }
      } else {
        const NodeHead node = unpack_node_head(header, data[curElem.nodeId], data[curElem.nodeId + header.links_offset - 1]);
        const uint32_t node1 = data[curElem.nodeId + header.references_offset];
        const uint32_t node2 = data[curElem.nodeId + header.links_offset - 1];
        OTStackElement tmp_buf[4];
        int buf_top = 0;
        const f3 tm = 0.5f * (t0 + t1);
        int currNode = FirstNode(t0, tm);
        do {
          const uint32_t rawChildNode = uint32_t(currNode) ^ a;
          const int vp[3] = {int((rawChildNode & 4) >> 2), int((rawChildNode & 2) >> 1), int(rawChildNode & 1)};
          const I4& m = mods[2 * 48 + curElem.info];
          const uint32_t voxId = uint32_t(m.x * vp[0] + m.y * vp[1] + m.z * vp[2] + m.w * 1);
          const uint32_t ch_type = (node.children_types >> 2 * voxId) & 3;
          const uint32_t ch_start_link = voxId == 0 ? node.base_links_end : get_links_end(header, voxId - 1, node1, node2);
          const uint32_t ch_end_link = get_links_end(header, voxId, node1, node2);
          if (ch_type != 0 && buf_top < 4) {
            tmp_buf[buf_top].p_size_x = (curElem.p_size_x << 1) | (((uint32_t(currNode) & 4) << (16 - 2)) | ((uint32_t(currNode) & 2) >> 1));
            tmp_buf[buf_top].p_size_y = (curElem.p_size_y << 1) | ((uint32_t(currNode) & 1) << 16);
            const uint32_t link_off = curElem.nodeId + header.links_offset + ch_start_link * header.uints_per_link;
            if (ch_type == 2) {
              const ChildEdge ce = unpack_child_edge(header, data[link_off], data[link_off + header.child_offset_off]);
              tmp_buf[buf_top].nodeId = header.nodes_arr_offset + ce.child_offset;
              tmp_buf[buf_top].info = s.rot_add[curElem.info * 48 + ce.rotation_id];
            } else {
              const uint32_t surf_count = ch_end_link - ch_start_link - header.has_channels;
              tmp_buf[buf_top].nodeId = link_off;
              tmp_buf[buf_top].info = (curElem.info & 0xFF) | (surf_count << 8) | (ch_type == 3 ? (1u << 24) : 0);
            }
            buf_top++;
          }
          currNode = NewNode((currNode & 4) ? t1.x : tm.x, nn[currNode][0], (currNode & 2) ? t1.y : tm.y, nn[currNode][1],
                             (currNode & 1) ? t1.z : tm.z, nn[currNode][2]);
        } while (currNode < 8);
        for (int i = 1; i <= buf_top; i++)
          if (top + i < 32) stack[top + i] = tmp_buf[buf_top - i];
        top = std::min(top + buf_top, 31);
      }
    }
  }

  // LocalSurfaceIntersectionFast (BVH2Common.cpp:2820-2986): the analytic root (or the box in
  // BBOX mode), the normal from central differences without the instance transform
  void LocalSurfaceIntersectionFast(uint32_t type, f3 ray_dir, uint32_t instId, uint32_t geomId, const float values[8],
                                    uint32_t primId, float d, float qNear, float qFar, f2 fNearFar, f3 start_q,
                                    Hit* pHit) {
    const float EPS = 1e-6f;
    const float d_inv = 1.0f / d;
    float t = qNear;
    bool hit = false;
    const float start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
    if (start_dist <= EPS || s.preset.sdf_node_intersect == 3) {
      hit = true;
    } else {
      const float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
      const float s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
      const float tmp = s101 - s001;
      const float k0 = s000, k1 = s100 - s000, k2 = s010 - s000, k3 = s110 - s010 - k1, k4 = k0 - s001;
      const float k5 = k1 - tmp, k6 = k2 - (s011 - s001), k7 = k3 - (s111 - s011 - tmp);
      const f3 o = start_q, d3 = ray_dir;
      const float m0 = o.x * o.y, m1 = d3.x * d3.y, m2 = o.x * d3.y + o.y * d3.x;
      const float m3 = k5 * o.z - k1, m4 = k6 * o.z - k2, m5 = k7 * o.z - k3;
      const float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
      const float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
      const float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
      const float c3 = k7 * m1 * d3.z;
      float x1 = 1000, x2 = 1000, x3 = 1000;
      const float PI = 3.14159265358979323846f;
      if (std::abs(c3) > 1e-5f) {
        float a = c2 / c3, b = c1 / c3, c = c0 / c3;
        float Q = (a * a - 3 * b) / 9;
        float R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
        float Q3 = Q * Q * Q;
        if (R * R < Q3) {
          float theta = std::acos(R / std::sqrt(Q3));
          x1 = -2 * std::sqrt(Q) * std::cos(theta / 3) - a / 3;
          x2 = -2 * std::sqrt(Q) * std::cos((theta + 2 * PI) / 3) - a / 3;
          x3 = -2 * std::sqrt(Q) * std::cos((theta - 2 * PI) / 3) - a / 3;
        } else {
          float A = -lm_sign(R) * std::pow(std::abs(R) + std::sqrt(R * R - Q3), 1.0f / 3.0f);
          float B = std::abs(A) > EPS ? Q / A : 0;
          x1 = A + B - a / 3;
        }
      } else if (std::abs(c2) > 1e-7f) {
        float a = c2, b = c1, c = c0;
        float D = b * b - 4 * a * c;
        if (D > 0) {
          float q = -0.5f * (b + lm_sign(b) * std::sqrt(D));
          x1 = q / a;
          if (std::abs(q) > EPS) x2 = c / q;
        }
      } else if (std::abs(c1) > EPS) {
        x1 = -c0 / c1;
      }
      x1 = x1 < 0 ? 1000 : x1;
      x2 = x2 < 0 ? 1000 : x2;
      x3 = x3 < 0 ? 1000 : x3;
      t = std::min(x1, std::min(x2, x3));
      hit = (t >= 0 && t <= qFar);
    }
    const float tReal = fNearFar.x + d * t;
    if (t <= qFar && hit && tReal < pHit->t) {
      const f3 p0 = start_q + t * ray_dir;
      const float h = 0.001f;
      const float ddx = (eval_dist_trilinear(values, p0 + f3(h, 0, 0)) - eval_dist_trilinear(values, p0 + f3(-h, 0, 0))) / (2 * h);
      const float ddy = (eval_dist_trilinear(values, p0 + f3(0, h, 0)) - eval_dist_trilinear(values, p0 + f3(0, -h, 0))) / (2 * h);
      const float ddz = (eval_dist_trilinear(values, p0 + f3(0, 0, h)) - eval_dist_trilinear(values, p0 + f3(0, 0, -h))) / (2 * h);
      const f2 en = encode_normal(normalize(f3(ddx, ddy, ddz)));
      pHit->t = tReal;
      pHit->primId = primId;
      pHit->instId = instId;
      pHit->geomId = geomId | (type << SH_TYPE);
      pHit->coords[0] = 0;
      pHit->coords[1] = 0;
      pHit->coords[2] = en.x;
      pHit->coords[3] = en.y;
    }
  }

  // new_node(i, txm, tym, tzm) (utils/raytracing/intersection_common.h:87-93)
  static uint32_t NewNodeI(uint32_t i, float txm, float tym, float tzm) {
    const uint32_t x = 4 + i, y = 2 + (i & 0x5) + ((i & 0x2) << 3), z = 1 + (i & 0x6) + ((i & 0x1) << 3);
    return (txm < tym) ? (txm < tzm ? x : z) : (tym < tzm ? y : z);
  }

  // SCom2IntersectFast (BVH2Common.cpp:2988-3231), taken for a scene whose only SCom2 tree is plain
  // (no multi-nodes, no channels, 3^3 values, one-word links) when the preset renders
  // LAMBERT_NO_TEX: one child per step, a 16-entry stack of the nodes still to finish, leaves
  // walked voxel by voxel in front-to-back order.
  void SCom2IntersectFast(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, Hit* pHit) {
    struct Ext { uint32_t linksOff, info, transform, p_size_x, p_size_y; };
    const uint32_t type = TYPE_SCOM2;
    const uint32_t hidx = s.geoms[geomId].offset[1];
    const SC2Header header = s.scom2_headers[hidx];
    const uint32_t* data = s.scom2_data;
    const I4* mods = s.scom2_mods + hidx * 144;
    f3 pos_ray_pos = ray_pos, pos_ray_dir = ray_dir;
    uint32_t a = 0;
    if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
    if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
    if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
    pos_ray_pos = pos_ray_pos * -1.0f;
    const f3 pos_ray_dir_inv = SafeInverse(pos_ray_dir);
    const f3 _t0 = pos_ray_pos * pos_ray_dir_inv - pos_ray_dir_inv;
    const f3 _t1 = pos_ray_pos * pos_ray_dir_inv + pos_ray_dir_inv;
    const f3 _l = _t1 - _t0;
    Ext stack[16];
    int top = 0;
    float d = 1.0f;
    f3 t0 = _t0;
    f3 tm = t0 + 0.5f * _l;
    Ext cur;
    {
      const uint32_t nodeData0 = data[header.root_node_off];
      const NodeHead node = unpack_node_head(header, nodeData0, nodeData0);
      const uint32_t currNode = uint32_t(FirstNode(t0, tm));
      cur.linksOff = header.root_node_off + header.links_offset + (node.base_type == 0 ? 0 : 1);
      cur.info = (0u << 24) | (node.children_types << 8) | currNode;
      cur.transform = 0;
      cur.p_size_x = 0;
      cur.p_size_y = 1;
    }
    stack[0] = Ext{0xFFFFFFFFu, 0, 0, 0, 0};
    while (top >= 0) {
      const uint32_t currNode = cur.info & 0x7;
      const uint32_t rawChildNode = currNode ^ a;
      const int vp[3] = {int((rawChildNode & 4) >> 2), int((rawChildNode & 2) >> 1), int(rawChildNode & 1)};
      const I4& mc = mods[2 * 48 + cur.transform];
      const uint32_t childN = uint32_t(mc.x * vp[0] + mc.y * vp[1] + mc.z * vp[2] + mc.w * 1);
      const uint32_t activeChildrenMask = (((cur.info & 0x00AAAA00u) >> 1) | (cur.info & 0x00555500u)) >> 8;
      const uint32_t childLink = cur.linksOff + uint32_t(__builtin_popcount(activeChildrenMask & ((1u << (2 * childN)) - 1)));
      const uint32_t childHasData = activeChildrenMask & (1u << (2 * childN));
      const uint32_t childIsLeaf = cur.info & (1u << (8 + 2 * childN));
      const uint32_t rotId = cur.info >> 24;
      d = 1.0f / float(cur.p_size_y & 0xFFFF);
      const f3 pf2 = f3(float(cur.p_size_x >> 16) + (((currNode & 4) > 0) ? 1.0f : 0.5f),
                        float(cur.p_size_x & 0xFFFF) + (((currNode & 2) > 0) ? 1.0f : 0.5f),
                        float(cur.p_size_y >> 16) + (((currNode & 1) > 0) ? 1.0f : 0.5f));
      const f3 tl = _t0 + d * pf2 * _l;
      if (std::min(std::min(tl.x, tl.y), tl.z) < tNear || childHasData == 0) {
        const uint32_t next = NewNodeI(currNode, tl.x, tl.y, tl.z);
        if (next >= 8) cur = stack[top--];
        else cur.info = next | (cur.info & 0xFFFFFF00u);
      } else if (childIsLeaf > 0) {
        const uint32_t level_sz = 2 * (cur.p_size_y & 0xFFFF);
        const uint32_t p[3] = {2 * (cur.p_size_x >> 16) + ((currNode & 4) >> 2), 2 * (cur.p_size_x & 0xFFFF) + ((currNode & 2) >> 1),
                               2 * (cur.p_size_y >> 16) + (currNode & 1)};
        d = 2.0f / float(level_sz);
        const f3 p_f = f3(float(p[0]), float(p[1]), float(p[2]));
        t0 = _t0 + 0.5f * d * p_f * _l;
        const f3 t1 = _t0 + 0.5f * d * (p_f + 1.0f) * _l;
        const float max_val = 1.7320508f * 2.0f / float(level_sz);
        const bool childIsSurface = (cur.info & (1u << (8 + 2 * childN + 1))) > 0;
        const uint32_t linkData0 = data[childLink];
        const DataEdge de = unpack_data_edge(header, max_val, linkData0, linkData0);
        const uint32_t offset = header.bricks_arr_offset + header.bricks_step * de.data_offset;
        const uint32_t rotIdx = s.rot_add[rotId * 48 + de.rotation_id];
        const float add = de.add;
        const f3 vox_tm = 0.5f * (t0 + t1);
        uint32_t vox_currNode = uint32_t(FirstNode(t0, vox_tm));
        while (vox_currNode < 8) {
          const f3 vox_tl((vox_currNode & 4) > 0 ? t1.x : vox_tm.x, (vox_currNode & 2) > 0 ? t1.y : vox_tm.y,
                          (vox_currNode & 1) > 0 ? t1.z : vox_tm.z);
          const uint32_t voxId = vox_currNode ^ a;
          float vmin = 1.0f;
          float vmax = childIsSurface ? -1.0f : 1.0f;
          float values[8];
          for (int i = 0; i < 8; i++) {
            const int pI[3] = {int((voxId & 4) >> 2) + ((i & 4) >> 2), int((voxId & 2) >> 1) + ((i & 2) >> 1),
                               int(voxId & 1) + (i & 1)};
            const I4& m = mods[rotIdx];
            const uint32_t vId0 = uint32_t(m.x * pI[0] + m.y * pI[1] + m.z * pI[2] + m.w * 1);
            const uint32_t p_val = data[offset + vId0 / header.values_per_uint];
            const uint32_t p_off = (vId0 % header.values_per_uint) * header.bits_per_value;
            const float val = max_val * (2.0f * ((p_val >> p_off) & header.value_mask) / float(header.value_mask) - 1) + add;
            values[i] = val - 0.5f * s.preset.compact_sdf_eps * d;
            vmin = std::min(vmin, val);
            vmax = std::max(vmax, val);
          }
          if (vmin <= 0.0f && vmax >= 0.0f) {
            const uint32_t p_mask = level_sz - 1;
            const uint32_t rp[3] = {(a & 4) ? (~p[0] & p_mask) : p[0], (a & 2) ? (~p[1] & p_mask) : p[1],
                                    (a & 1) ? (~p[2] & p_mask) : p[2]};
            const f3 vox_t0((vox_currNode & 4) > 0 ? vox_tm.x : t0.x, (vox_currNode & 2) > 0 ? vox_tm.y : t0.y,
                            (vox_currNode & 1) > 0 ? vox_tm.z : t0.z);
            const f3 vox_t1((vox_currNode & 4) > 0 ? t1.x : vox_tm.x, (vox_currNode & 2) > 0 ? t1.y : vox_tm.y,
                            (vox_currNode & 1) > 0 ? t1.z : vox_tm.z);
            const float tmin = std::max(vox_t0.x, std::max(vox_t0.y, vox_t0.z));
            const float tmax = std::min(vox_t1.x, std::min(vox_t1.y, vox_t1.z));
            const f3 start_pos = ray_pos + tmin * ray_dir;
            const f3 min_pos = f3(-1, -1, -1) + d * f3(float(rp[0]), float(rp[1]), float(rp[2])) +
                               0.5f * d * f3(float((voxId & 4) >> 2), float((voxId & 2) >> 1), float(voxId & 1));
            const float vox_sz = float(level_sz);
            const f3 start_q = vox_sz * (start_pos - min_pos);
            const float qNear = std::max(0.0f, vox_sz * (tNear - tmin));
            const float qFar = vox_sz * (tmax - tmin);
            const float last_t = pHit->t;
            LocalSurfaceIntersectionFast(type, ray_dir, instId, geomId, values, childLink, 0.5f * d, qNear, qFar,
                                         f2{tmin, tmax}, start_q, pHit);
            if (pHit->t < last_t) return;
          }
          vox_currNode = NewNodeI(vox_currNode, vox_tl.x, vox_tl.y, vox_tl.z);
        }
        const uint32_t next = NewNodeI(currNode, tl.x, tl.y, tl.z);
        if (next >= 8) cur = stack[top--];
        else cur.info = next | (cur.info & 0xFFFFFF00u);
      } else {
        const uint32_t next = NewNodeI(currNode, tl.x, tl.y, tl.z);
        if (next < 8) {
          cur.info = next | (cur.info & 0xFFFFFF00u);
          if (top + 1 < 16) stack[++top] = cur;
        }
        d = 0.5f / float(cur.p_size_y & 0xFFFF);
        const f3 p_f = f3(float(2 * (cur.p_size_x >> 16) + ((currNode & 4) >> 2)), float(2 * (cur.p_size_x & 0xFFFF) + ((currNode & 2) >> 1)),
                          float(2 * (cur.p_size_y >> 16) + (currNode & 1)));
        t0 = _t0 + d * p_f * _l;
        tm = t0 + 0.5f * d * _l;
        const uint32_t linkData0 = data[childLink];
        const ChildEdge ce = unpack_child_edge(header, linkData0, linkData0);
        const uint32_t nodeData0 = data[ce.child_offset];
        const NodeHead node = unpack_node_head(header, nodeData0, nodeData0);
        const uint32_t nextNode = uint32_t(FirstNode(t0, tm));
        const uint32_t rotIdx = s.rot_add[rotId * 48 + ce.rotation_id];
        cur.linksOff = ce.child_offset + header.links_offset + (node.base_type == 0 ? 0 : 1);
        cur.transform = s.rot_add[cur.transform * 48 + ce.rotation_id];
        cur.info = (rotIdx << 24) | (node.children_types << 8) | nextNode;
        cur.p_size_x = (cur.p_size_x << 1) | (((currNode & 4) << (16 - 2)) | ((currNode & 2) >> 1));
        cur.p_size_y = (cur.p_size_y << 1) | ((currNode & 1) << 16);
      }
    }
  }

  // BVH2Common.cpp:47-124
  uint32_t IntersectAllPrimitivesInLeaf(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                                        uint32_t aabbId, Hit* pHit) {
    const float tPrev = pHit->t;
    const uint32_t globalAABBId = s.start_end[2 * geomId] + aabbId;
    const uint32_t packed = s.prim_id_count[globalAABBId];
    const uint32_t a_start = packed & START_MASK, a_count = (packed & SIZE_MASK) >> 28;
    const uint32_t geomTag = s.geom_tags[geomId];
    if (cnt) cnt->leaves++;
    switch (geomTag) {
      case TAG_TRIANGLE: IntersectAllTrianglesInLeaf(ray_pos, ray_dir, tNear, instId, geomId, a_start, a_count, pHit); break;
      case TAG_SDF_NODE: OctreeNodeIntersect(s.geoms[geomId].type, ray_pos, ray_dir, tNear, instId, geomId, a_start, pHit); break;
      case TAG_SDF_BRICK: OctreeBrickIntersect(s.geoms[geomId].type, ray_pos, ray_dir, tNear, instId, geomId, a_start, pHit); break;
      case TAG_SDF_OCTREE: OctreeIntersect(ray_pos, ray_dir, tNear, instId, geomId, pHit); break;
      case TAG_SDF_GRID: IntersectSDFGrid(ray_pos, ray_dir, tNear, instId, geomId, pHit); break;
      case TAG_GRAPHICS_PRIM: IntersectGraphicPrims(ray_pos, ray_dir, tNear, instId, geomId, a_start, pHit); break;
      case TAG_COCTREE_BRICKED: COctreeIntersectV3(ray_pos, ray_dir, tNear, instId, geomId, a_start, pHit); break;
      case TAG_SDF_DAG: DAGIntersect(ray_pos, ray_dir, tNear, instId, geomId, pHit); break;
      case TAG_SCOM2:
        if (scom2_fast()) SCom2IntersectFast(ray_pos, ray_dir, tNear, instId, geomId, pHit);
        else SCom2Intersect(ray_pos, ray_dir, tNear, instId, geomId, pHit);
        break;
      default: break;
    }
    if (s.preset.render_mode == 26) pHit->instId = (pHit->instId == 0xFFFFFFFF) ? 1 : pHit->instId + 1;
    return (pHit->t >= tPrev) ? TAG_NONE : geomTag;
  }

  // BVH2Common.cpp:3239-3320. Pushes beyond STACK_SIZE are dropped (the reference writes past
  // its array); *overflow reports it.
  void BVH2TraverseF32(f3 ray_pos, f3 ray_dir, float tNear, uint32_t instId, uint32_t geomId, bool stopOnFirstHit,
                       Hit* pHit, bool* overflow) {
    const uint32_t bvhOffset = s.geoms[geomId].bvhOffset;
    uint32_t stack[STACK_SIZE];
    int top = 0;
    uint32_t leftNodeOffset = 0;
    bool hitFound = false;
    const f3 rayDirInv = SafeInverse(ray_dir);
    while (top >= 0 && !hitFound) {
      while (top >= 0 && ((leftNodeOffset & LEAF_BIT) == 0)) {
        if (cnt) cnt->blas_nodes++;
        const NodePair fatNode = s.nodes[bvhOffset + leftNodeOffset];
        const uint32_t node0_leftOffset = fatNode.left.leftOffset;
        const uint32_t node1_leftOffset = fatNode.right.leftOffset;
        const f2 tm0 = RayBoxIntersection2(ray_pos, rayDirInv, fatNode.left.boxMin, fatNode.left.boxMax);
        const f2 tm1 = RayBoxIntersection2(ray_pos, rayDirInv, fatNode.right.boxMin, fatNode.right.boxMax);
        const bool hitChild0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= pHit->t);
        const bool hitChild1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= pHit->t);
        leftNodeOffset = hitChild0 ? node0_leftOffset : node1_leftOffset;
        if (hitChild0 && hitChild1) {
          leftNodeOffset = (tm0.x <= tm1.x) ? node0_leftOffset : node1_leftOffset;
          if (top < STACK_SIZE) {
            stack[top] = (tm0.x <= tm1.x) ? node1_leftOffset : node0_leftOffset;
            top++;
          } else if (overflow) {
            *overflow = true;
          }
        }
        if (!hitChild0 && !hitChild1) {
          top--;
          leftNodeOffset = stack[std::max(top, 0)];
        }
      }
      if (top >= 0 && leftNodeOffset != 0xFFFFFFFF) {
        const float SDF_BIAS = 0.01f;
        const float tNearSdf = std::max(tNear, SDF_BIAS);
        const uint32_t hitTag =
            IntersectAllPrimitivesInLeaf(ray_pos, ray_dir, tNearSdf, instId, geomId, leftNodeOffset & START_MASK, pHit);
        hitFound = (hitTag != 0) && (hitTag != 1 || stopOnFirstHit);
      }
      top--;
      leftNodeOffset = stack[std::max(top, 0)];
    }
  }

  // BVH2Common.cpp:3322-3392
  Hit RayQuery_NearestHit(f4 posAndNear, f4 dirAndFar, bool* overflow) {
    bool stopOnFirstHit = (dirAndFar.w <= 0.0f);
    if (stopOnFirstHit) dirAndFar.w *= -1.0f;
    const f3 rayDirInv = SafeInverse(to3(dirAndFar));
    Hit hit;
    hit.t = dirAndFar.w;
    hit.primId = hit.instId = hit.geomId = uint32_t(-1);
    hit.coords[0] = 1.0f;
    hit.coords[1] = hit.coords[2] = hit.coords[3] = 0.0f;
    uint32_t nodeIdx = 0;
    do {
      uint32_t travFlags = 0, leftOffset = 0;
      do {
        if (cnt) cnt->tlas_nodes++;
        const Node currNode = s.tlas[nodeIdx];
        const f2 boxHit = RayBoxIntersection2(to3(posAndNear), rayDirInv, currNode.boxMin, currNode.boxMax);
        const bool intersects = (boxHit.x <= boxHit.y) && (boxHit.y > posAndNear.w) && (boxHit.x < hit.t);
        travFlags = (currNode.leftOffset & LEAF_BIT) | uint32_t(intersects);
        leftOffset = currNode.leftOffset;
        const bool leafOrMiss = (travFlags & LEAF_BIT) != 0 || (travFlags & 0x1) == 0;
        nodeIdx = leafOrMiss ? currNode.escapeIndex : leftOffset;
      } while (travFlags != (LEAF_BIT | 0x1) && nodeIdx != 0 && nodeIdx < 0xFFFFFFFE);
      if (travFlags == (LEAF_BIT | 0x1)) {
        if (cnt) cnt->instances++;
        const uint32_t instId = leftOffset & START_MASK;
        const uint32_t geomId = s.instances[instId].geomId;
        const f3 ray_pos = matmul4x3(s.instances[instId].transformInv, to3(posAndNear));
        const f3 ray_dir = matmul3x3(s.instances[instId].transformInv, to3(dirAndFar));
        BVH2TraverseF32(ray_pos, ray_dir, posAndNear.w, instId, geomId, stopOnFirstHit, &hit, overflow);
      }
    } while (nodeIdx < 0xFFFFFFFE && !(stopOnFirstHit && hit.primId != uint32_t(-1)));
    if (hit.geomId < uint32_t(-1) && ((hit.geomId >> SH_TYPE) == TYPE_MESH_TRIANGLE)) {
      const uint32_t* go = s.geoms[hit.geomId & GEOM_ID_MASK].offset;
      hit.primId = s.prim_indices[go[0] / 3 + hit.primId];
    }
    return hit;
  }
};

// LiteMath.h:2070-2154
void inverse4x4(const float* a, float* out) {
  auto m1 = [&](int r, int c) { return a[c * 4 + r]; };
  float m[16];
  auto M = [&](int r, int c) -> float& { return m[c * 4 + r]; };
  float tmp[12];
  tmp[0] = m1(2, 2) * m1(3, 3); tmp[1] = m1(2, 3) * m1(3, 2); tmp[2] = m1(2, 1) * m1(3, 3);
  tmp[3] = m1(2, 3) * m1(3, 1); tmp[4] = m1(2, 1) * m1(3, 2); tmp[5] = m1(2, 2) * m1(3, 1);
  tmp[6] = m1(2, 0) * m1(3, 3); tmp[7] = m1(2, 3) * m1(3, 0); tmp[8] = m1(2, 0) * m1(3, 2);
  tmp[9] = m1(2, 2) * m1(3, 0); tmp[10] = m1(2, 0) * m1(3, 1); tmp[11] = m1(2, 1) * m1(3, 0);
  M(0, 0) = tmp[0] * m1(1, 1) + tmp[3] * m1(1, 2) + tmp[4] * m1(1, 3);
  M(0, 0) -= tmp[1] * m1(1, 1) + tmp[2] * m1(1, 2) + tmp[5] * m1(1, 3);
  M(1, 0) = tmp[1] * m1(1, 0) + tmp[6] * m1(1, 2) + tmp[9] * m1(1, 3);
  M(1, 0) -= tmp[0] * m1(1, 0) + tmp[7] * m1(1, 2) + tmp[8] * m1(1, 3);
  M(2, 0) = tmp[2] * m1(1, 0) + tmp[7] * m1(1, 1) + tmp[10] * m1(1, 3);
  M(2, 0) -= tmp[3] * m1(1, 0) + tmp[6] * m1(1, 1) + tmp[11] * m1(1, 3);
  M(3, 0) = tmp[5] * m1(1, 0) + tmp[8] * m1(1, 1) + tmp[11] * m1(1, 2);
  M(3, 0) -= tmp[4] * m1(1, 0) + tmp[9] * m1(1, 1) + tmp[10] * m1(1, 2);
  M(0, 1) = tmp[1] * m1(0, 1) + tmp[2] * m1(0, 2) + tmp[5] * m1(0, 3);
  M(0, 1) -= tmp[0] * m1(0, 1) + tmp[3] * m1(0, 2) + tmp[4] * m1(0, 3);
  M(1, 1) = tmp[0] * m1(0, 0) + tmp[7] * m1(0, 2) + tmp[8] * m1(0, 3);
  M(1, 1) -= tmp[1] * m1(0, 0) + tmp[6] * m1(0, 2) + tmp[9] * m1(0, 3);
  M(2, 1) = tmp[3] * m1(0, 0) + tmp[6] * m1(0, 1) + tmp[11] * m1(0, 3);
  M(2, 1) -= tmp[2] * m1(0, 0) + tmp[7] * m1(0, 1) + tmp[10] * m1(0, 3);
  M(3, 1) = tmp[4] * m1(0, 0) + tmp[9] * m1(0, 1) + tmp[10] * m1(0, 2);
  M(3, 1) -= tmp[5] * m1(0, 0) + tmp[8] * m1(0, 1) + tmp[11] * m1(0, 2);
  tmp[0] = m1(0, 2) * m1(1, 3); tmp[1] = m1(0, 3) * m1(1, 2); tmp[2] = m1(0, 1) * m1(1, 3);
  tmp[3] = m1(0, 3) * m1(1, 1); tmp[4] = m1(0, 1) * m1(1, 2); tmp[5] = m1(0, 2) * m1(1, 1);
  tmp[6] = m1(0, 0) * m1(1, 3); tmp[7] = m1(0, 3) * m1(1, 0); tmp[8] = m1(0, 0) * m1(1, 2);
  tmp[9] = m1(0, 2) * m1(1, 0); tmp[10] = m1(0, 0) * m1(1, 1); tmp[11] = m1(0, 1) * m1(1, 0);
  M(0, 2) = tmp[0] * m1(3, 1) + tmp[3] * m1(3, 2) + tmp[4] * m1(3, 3);
  M(0, 2) -= tmp[1] * m1(3, 1) + tmp[2] * m1(3, 2) + tmp[5] * m1(3, 3);
  M(1, 2) = tmp[1] * m1(3, 0) + tmp[6] * m1(3, 2) + tmp[9] * m1(3, 3);
  M(1, 2) -= tmp[0] * m1(3, 0) + tmp[7] * m1(3, 2) + tmp[8] * m1(3, 3);
  M(2, 2) = tmp[2] * m1(3, 0) + tmp[7] * m1(3, 1) + tmp[10] * m1(3, 3);
  M(2, 2) -= tmp[3] * m1(3, 0) + tmp[6] * m1(3, 1) + tmp[11] * m1(3, 3);
  M(3, 2) = tmp[5] * m1(3, 0) + tmp[8] * m1(3, 1) + tmp[11] * m1(3, 2);
  M(3, 2) -= tmp[4] * m1(3, 0) + tmp[9] * m1(3, 1) + tmp[10] * m1(3, 2);
  M(0, 3) = tmp[2] * m1(2, 2) + tmp[5] * m1(2, 3) + tmp[1] * m1(2, 1);
  M(0, 3) -= tmp[4] * m1(2, 3) + tmp[0] * m1(2, 1) + tmp[3] * m1(2, 2);
  M(1, 3) = tmp[8] * m1(2, 3) + tmp[0] * m1(2, 0) + tmp[7] * m1(2, 2);
  M(1, 3) -= tmp[6] * m1(2, 2) + tmp[9] * m1(2, 3) + tmp[1] * m1(2, 0);
  M(2, 3) = tmp[6] * m1(2, 1) + tmp[11] * m1(2, 3) + tmp[3] * m1(2, 0);
  M(2, 3) -= tmp[10] * m1(2, 3) + tmp[2] * m1(2, 0) + tmp[7] * m1(2, 1);
  M(3, 3) = tmp[10] * m1(2, 2) + tmp[4] * m1(2, 0) + tmp[9] * m1(2, 1);
  M(3, 3) -= tmp[8] * m1(2, 1) + tmp[11] * m1(2, 2) + tmp[5] * m1(2, 0);
  const float k = 1.0f / (m1(0, 0) * M(0, 0) + m1(0, 1) * M(1, 0) + m1(0, 2) * M(2, 0) + m1(0, 3) * M(3, 0));
  for (int i = 0; i < 16; ++i) out[i] = m[i] * k;
}

}  // namespace

extern "C" {

int oracle_flavor(void) {
#ifdef ORACLE_FMA
  return 1;
#else
  return 0;
#endif
}

// Batched BVHRT::RayQuery_NearestHit over host arrays; OpenMP over rays with
// schedule(dynamic,256). Returns the number of rays whose stack
// reached the 80-entry limit. counts may be NULL.
uint64_t oracle_trace_closest(const OracleScene* sc, const float* pos, const float* dir, void* hits, uint64_t n,
                              int threads, OracleCounts* counts) {
  uint64_t overflow = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) reduction(+ : overflow)
  for (long i = 0; i < long(n); ++i) {
    OracleCounts c{};
    Tracer tr{*sc, counts ? &c : nullptr};
    bool ov = false;
    const f4 p{pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], pos[4 * i + 3]};
    const f4 d{dir[4 * i], dir[4 * i + 1], dir[4 * i + 2], dir[4 * i + 3]};
    static_cast<Hit*>(hits)[i] = tr.RayQuery_NearestHit(p, d, &ov);
    if (counts) counts[i] = c;
    overflow += ov ? 1 : 0;
  }
  return overflow;
}

// BVHRT::RayQuery_AnyHit (BVH2Common.cpp:3394-3399)
void oracle_trace_any(const OracleScene* sc, const float* pos, const float* dir, uint32_t* flags, uint64_t n,
                      int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
  for (long i = 0; i < long(n); ++i) {
    Tracer tr{*sc, nullptr};
    const f4 p{pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], pos[4 * i + 3]};
    f4 d{dir[4 * i], dir[4 * i + 1], dir[4 * i + 2], dir[4 * i + 3]};
    d.w *= -1.0f;
    const Hit h = tr.RayQuery_NearestHit(p, d, nullptr);
    flags[i] = h.geomId != uint32_t(-1) ? 1u : 0u;
  }
}

// Closest hit over every triangle of one mesh with the reference's triangle test
// (BVH2Common.cpp:1589-1600), no acceleration structure: the topology-independent answer.
// verts: float4 per vertex, tris: original (not reordered) index triples. tnear as in the
// BLAS leaf call (max(ray near, 0.01), BVH2Common.cpp:3306-3307). Writes t and the original
// triangle id (0xFFFFFFFF on miss) and the second-closest t (for the tie band).
void oracle_brute_force_tris(const float* verts, const uint32_t* tris, uint64_t n_tris, const float* pos,
                             const float* dir, uint64_t n, float* t_out, uint32_t* id_out, float* t2_out, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (long r = 0; r < long(n); ++r) {
    const f3 o(pos[4 * r], pos[4 * r + 1], pos[4 * r + 2]);
    const f3 d(dir[4 * r], dir[4 * r + 1], dir[4 * r + 2]);
    const float tNear = std::max(pos[4 * r + 3], 0.01f);
    float best = std::abs(dir[4 * r + 3]), second = best;
    uint32_t id = 0xFFFFFFFFu;
    for (uint64_t k = 0; k < n_tris; ++k) {
      const float* A = verts + 4 * tris[3 * k];
      const float* B = verts + 4 * tris[3 * k + 1];
      const float* C = verts + 4 * tris[3 * k + 2];
      const f3 a(A[0], A[1], A[2]);
      const f3 e1 = f3(B[0], B[1], B[2]) - a, e2 = f3(C[0], C[1], C[2]) - a;
      const f3 pvec = cross(d, e2), tvec = o - a, qvec = cross(tvec, e1);
      const float invDet = 1.0f / dot(e1, pvec);
      const float v = dot(tvec, pvec) * invDet, u = dot(qvec, d) * invDet, t = dot(e2, qvec) * invDet;
      if (v >= -1e-6f && u >= -1e-6f && (u + v <= 1.0f + 1e-6f) && t > tNear) {
        if (t < best) { second = best; best = t; id = uint32_t(k); }
        else if (t < second) second = t;
      }
    }
    t_out[r] = best;
    id_out[r] = id;
    t2_out[r] = second;
  }
}

// MultiRenderer::initEyeRay (MultiRenderer.cpp:142-183), cutting plane disabled, spp 1:
// EyeRayDirNormalized + transform_ray3f (render_common.h:25-55). order 1 places ray (x,y)
// at SuperBlockIndex2DOpt(x, y, packed width) (render_common.h:225-242).
void oracle_eye_rays(const float* proj_inv, const float* wv_inv, uint32_t width, uint32_t height, const float* jitter,
                     uint32_t order, float* pos, float* dir) {
  const uint32_t pw = 8 * ((width + 7) / 8), ph = 8 * ((height + 7) / 8);
  for (uint32_t y = 0; y < height; ++y)
    for (uint32_t x = 0; x < width; ++x) {
      const uint32_t pix = y * width + x;
      const f2 dd = jitter ? f2{jitter[2 * pix], jitter[2 * pix + 1]} : f2{0.5f, 0.5f};
      const float fx = (float(x) + dd.x) / float(pw), fy = (float(y) + dd.y) / float(ph);
      f4 p{2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f};
      p = mul(proj_inv, p);
      const float w = p.w;
      p = f4{p.x / w, p.y / w, p.z / w, p.w / w};
      f3 rd = normalize(to3(p));
      const f3 zero(0, 0, 0);
      const f3 o = matmul4x3(wv_inv, zero);
      const f3 o2 = matmul4x3(wv_inv, zero + 100.0f * rd);
      rd = normalize(o2 - o);
      uint32_t idx = pix;
      if (order == 1) {
        const uint32_t local = (y & 3u) * 4 + (x & 3u), bx = x >> 2, by = y >> 2;
        const uint32_t localH = (by & 1u) * 2 + (bx & 1u), wbH = (pw >> 2) >> 1;
        idx = ((bx >> 1) + (by >> 1) * wbH) * 64 + localH * 16 + local;
      }
      pos[4 * idx] = o.x; pos[4 * idx + 1] = o.y; pos[4 * idx + 2] = o.z; pos[4 * idx + 3] = 1e-9f;
      dir[4 * idx] = rd.x; dir[4 * idx + 1] = rd.y; dir[4 * idx + 2] = rd.z; dir[4 * idx + 3] = 1e9f;
    }
}

// MultiRenderer::UpdateCamera inverse matrices for lookAt/perspectiveMatrix
// (LiteMath.h:3315-3360, MultiRenderer_host.cpp:104-110).
void oracle_camera_inverse(const float* eye, const float* target, const float* up, float fovy, float aspect,
                           float znear, float zfar, float* proj_inv, float* wv_inv) {
  f3 z = normalize(f3(eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]));
  f3 y(up[0], up[1], up[2]);
  f3 x = cross(y, z);
  y = cross(z, x);
  x = normalize(x);
  y = normalize(y);
  const float view[16] = {x.x, y.x, z.x, 0.0f, x.y, y.y, z.y, 0.0f, x.z, y.z, z.z, 0.0f,
                          -x.x * eye[0] - x.y * eye[1] - x.z * eye[2], -y.x * eye[0] - y.y * eye[1] - y.z * eye[2],
                          -z.x * eye[0] - z.y * eye[1] - z.z * eye[2], 1.0f};
  const float ymax = znear * tanf(fovy * 3.14159265358979323846f / 360.0f);
  const float xmax = ymax * aspect;
  const float left = -xmax, right = +xmax, bottom = -ymax, top = +ymax;
  const float temp = 2.0f * znear, temp2 = right - left, temp3 = top - bottom, temp4 = zfar - znear;
  const float proj[16] = {temp / temp2, 0.0f, 0.0f, 0.0f, 0.0f, temp / temp3, 0.0f, 0.0f,
                          (right + left) / temp2, (top + bottom) / temp3, (-zfar - znear) / temp4, -1.0f,
                          0.0f, 0.0f, (-temp * zfar) / temp4, 0.0f};
  inverse4x4(proj, proj_inv);
  inverse4x4(view, wv_inv);
}

void oracle_inverse4x4(const float* m, float* out) { inverse4x4(m, out); }

}  // extern "C"

extern "C" {

// One LocalSurfaceIntersection call (BVH2Common.cpp:345-725) on a unit voxel, with an identity
// instance, for the float64 root checks: `values` are the 8 corner distances, start_q the entry
// point in voxel coordinates, d the voxel size. Writes the hit record (t = fNear + d * t_local
// when hit, else far = 1e9) and returns 1 on a hit.
int oracle_local_surface(const void* preset, const float* values, float d, float qNear, float qFar, float fNear,
                         const float* start_q, const float* ray_dir, void* hit_out) {
  OracleScene sc{};
  std::memcpy(&sc.preset, preset, sizeof(Preset));
  Inst inst{};
  for (int i = 0; i < 4; ++i) {
    inst.transform[5 * i] = 1.0f;
    inst.transformInv[5 * i] = 1.0f;
    inst.transformInvTransposed[5 * i] = 1.0f;
  }
  sc.instances = &inst;
  Tracer tr{sc, nullptr};
  Hit h{1e9f, uint32_t(-1), uint32_t(-1), uint32_t(-1), {1, 0, 0, 0}};
  tr.LocalSurfaceIntersection(TYPE_SDF_FRAME_OCTREE, f3(ray_dir[0], ray_dir[1], ray_dir[2]), 0, 0, values, 0, d, qNear,
                              qFar, f2{fNear, 0.0f}, f3(start_q[0], start_q[1], start_q[2]), &h);
  std::memcpy(hit_out, &h, sizeof(h));
  return h.primId != uint32_t(-1) ? 1 : 0;
}

// The 8 corner values of voxel (vx, vy, vz) of one COctreeV3 leaf (COctreeV3_LoadDistanceValues,
// BVH2Common.cpp:1637-1851): leaf type, brick word offset and transform code as a child entry
// gives them, sz the leaf's level size. Returns the loader's result (min value, -1 or 1e6).
float oracle_coct_voxel(const OracleScene* sc, uint32_t coct, uint32_t leaf_type, uint32_t offset, int vx, int vy,
                        int vz, float sz, uint32_t code, float* values) {
  Tracer tr{*sc, nullptr};
  const CoctG& h = sc->coct_geoms[coct];
  const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
  return tr.COctreeV3_LoadDistanceValues(h, leaf_type, offset, vx, vy, vz, v_size, 2.0f / sz, code, values);
}

}  // extern "C"

extern "C" {

// Closest hit plus the voxel record of brick-set hits (VoxelRec per ray; hit = 0 when the final
// hit is not from a recorded voxel), for the float64 adjoint checks.
void oracle_sbs_hit_voxels(const OracleScene* sc, const float* pos, const float* dir, uint64_t n, int threads,
                           void* hits, void* recs) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
  for (long i = 0; i < long(n); ++i) {
    VoxelRec r{};
    Tracer tr{*sc, nullptr};
    tr.rec = &r;
    const f4 p{pos[4 * i], pos[4 * i + 1], pos[4 * i + 2], pos[4 * i + 3]};
    const f4 d{dir[4 * i], dir[4 * i + 1], dir[4 * i + 2], dir[4 * i + 3]};
    const Hit h = tr.RayQuery_NearestHit(p, d, nullptr);
    if (r.hit && (h.primId == uint32_t(-1) || r.t != h.t)) r.hit = 0;
    static_cast<Hit*>(hits)[i] = h;
    static_cast<VoxelRec*>(recs)[i] = r;
  }
}

}  // extern "C"

// ---- g-buffer resolve and tonemap: MultiRenderer::kernel1D_ResolveCommon / ResolveDebug /
//      Tonemap (MultiRenderer/MultiRenderer_resolve.cpp:10-75, 529-1005), MultiRenderer.cpp:142-183 ----
namespace {

struct Col { float x, y, z, w; };
inline Col operator*(float a, Col b) { return Col{a * b.x, a * b.y, a * b.z, a * b.w}; }
inline Col operator+(Col a, Col b) { return Col{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline Col operator/(Col a, float b) { return Col{a.x / b, a.y / b, a.z / b, a.w / b}; }

// MultiRenderer.h:270-284
const uint32_t kPalette[20] = {0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0,
                               0xfff032e6, 0xffd2f53c, 0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8,
                               0xff800000, 0xffaaffc3, 0xff808000, 0xffffd8b1, 0xff000080, 0xff808080};
const uint32_t kLodPalette[20] = {0xff000000, 0xff000040, 0xff000080, 0xff0000c0, 0xff0040ff, 0xff0080ff, 0xff00c0ff,
                                  0xff00ffff, 0xff00ffc0, 0xff00ff80, 0xff00ff40, 0xff00ff00, 0xff00c000, 0xff008000,
                                  0xff004000, 0xff000000, 0xffff0000, 0xffff0000, 0xffff0000, 0xffff0000};
const f2 kHamilton[16] = {{0.500000f, 0.333333f}, {0.250000f, 0.666667f}, {0.750000f, 0.111111f},
                          {0.125000f, 0.444444f}, {0.625000f, 0.777778f}, {0.375000f, 0.222222f},
                          {0.875000f, 0.555556f}, {0.062500f, 0.888889f}, {0.562500f, 0.037037f},
                          {0.312500f, 0.370370f}, {0.812500f, 0.703704f}, {0.187500f, 0.148148f},
                          {0.687500f, 0.481481f}, {0.437500f, 0.814815f}, {0.937500f, 0.259259f},
                          {0.031250f, 0.592593f}};

// render_common.h:250-254
inline Col decode_RGBA8(uint32_t c) {
  const uint32_t r = c & 0xFF, g = (c >> 8) & 0xFF, b = (c >> 16) & 0xFF, a = (c >> 24) & 0xFF;
  return Col{r * (1.0f / 255.0f), g * (1.0f / 255.0f), b * (1.0f / 255.0f), a * (1.0f / 255.0f)};
}
// render_common.h:256-267
inline f3 decode_normal(f2 e) {
  f3 v(e.x, e.y, 1.0f - std::abs(e.x) - std::abs(e.y));
  if (v.z < 0) {
    const float vx = v.x;
    v.x = (1.0f - std::abs(v.y)) * ((v.x >= 0.0f) ? +1.0f : -1.0f);
    v.y = (1.0f - std::abs(vx)) * ((v.y >= 0.0f) ? +1.0f : -1.0f);
  }
  return normalize(v);
}
inline float lm_clamp(float u, float a, float b) { return std::min(std::max(a, u), b); }

struct ResolveLight { float space[3]; uint32_t type; float color[3]; uint32_t pad; };

struct Resolver {
  const OracleScene& s;
  const uint32_t* orig_indices;
  const float* proj_inv;
  const float* wv_inv;
  uint32_t width, height, pw, ph, aa_frame;
  const ResolveLight* lights;
  uint32_t n_lights;
  const float* bg;
  const uint32_t* hl;   // mode, prim, geom, inst
  Preset p;

  // MultiRenderer::initEyeRay for sample 0, cutting plane disabled
  void init_eye_ray(uint32_t x, uint32_t y, f3* pos, f3* dir) const {
    const uint32_t spp_sqrt = uint32_t(std::sqrt(float(p.spp)));
    const float i_spp_sqrt = 1.0f / spp_sqrt;
    const f2 d = (aa_frame >= 2) ? kHamilton[aa_frame % 16] : f2{i_spp_sqrt * float(0 / spp_sqrt + 0.5),
                                                                  i_spp_sqrt * float(0 % spp_sqrt + 0.5)};
    const float fx = (float(x) + d.x) / float(pw), fy = (float(y) + d.y) / float(ph);
    f4 q{2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f};
    q = mul(proj_inv, q);
    const float w = q.w;
    q = f4{q.x / w, q.y / w, q.z / w, q.w / w};
    f3 rd = normalize(to3(q));
    const f3 zero(0, 0, 0);
    const f3 o = matmul4x3(wv_inv, zero);
    const f3 o2 = matmul4x3(wv_inv, zero + 100.0f * rd);
    *pos = o;
    *dir = normalize(o2 - o);
  }

  // MultiRenderer::loadNormalTcFromGbuffer (no texture coordinates on meshes here)
  void load_normal_tc(const Hit& hit, f3 ray_dir, f3* norm, f2* tc) const {
    const uint32_t type = hit.geomId >> SH_TYPE, geomId = hit.geomId & GEOM_ID_MASK;
    if (type == TYPE_MESH_TRIANGLE) {
      const Geom& g = s.geoms[geomId];
      const uint32_t A = orig_indices[g.offset[0] + hit.primId * 3 + 0];
      const uint32_t B = orig_indices[g.offset[0] + hit.primId * 3 + 1];
      const uint32_t C = orig_indices[g.offset[0] + hit.primId * 3 + 2];
      f3 raw(1, 0, 0);
      if (p.normal_mode == 0) {
        raw = cross(to3(s.vert_pos[g.offset[1] + B]) - to3(s.vert_pos[g.offset[1] + A]),
                    to3(s.vert_pos[g.offset[1] + C]) - to3(s.vert_pos[g.offset[1] + A]));
      } else if (p.normal_mode == 1) {
        const f4 nA = s.vert_norm[g.offset[1] + A], nB = s.vert_norm[g.offset[1] + B], nC = s.vert_norm[g.offset[1] + C];
        const float u = hit.coords[0], v = hit.coords[1], k = 1.0f - u - v;
        raw = f3(k * nA.x + v * nB.x + u * nC.x, k * nA.y + v * nB.y + u * nC.y, k * nA.z + v * nB.z + u * nC.z);
      }
      *tc = f2{0.0f, 0.0f};
      // transpose(inverse4x4(M)) of the renderer's own copy: the transpose of transformInv
      const float* inv = s.instances[hit.instId].transformInv;
      float it[16];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) it[4 * r + c] = inv[4 * c + r];
      *norm = normalize(matmul4x3(it, raw));
      *norm = *norm * lm_sign(dot(-1.0f * ray_dir, *norm));
    } else {
      *tc = f2{hit.coords[0], hit.coords[1]};
      *norm = decode_normal(f2{hit.coords[2], hit.coords[3]});
    }
  }

  Col resolve(uint32_t tid, const Hit& hit) const {
    if (hit.primId == 0xFFFFFFFFu) return Col{bg[0], bg[1], bg[2], aa_frame == 0 ? bg[3] : -1.0f};
    const uint32_t type = hit.geomId >> SH_TYPE, geomId = hit.geomId & GEOM_ID_MASK;
    f3 rpos, rdir;
    init_eye_ray(tid % pw, tid / pw, &rpos, &rdir);
    f2 tc{0, 0};
    f3 norm(1, 0, 0);
    load_normal_tc(hit, rdir, &norm, &tc);
    const uint32_t m = p.render_mode;
    Col res{1, 0, 1, 1};
    if (m == 1 || m == 7 || m == 20 || m == 17) {
      f3 color(0, 0, 1);
      if (m == 1) color = f3(1, 1, 1);
      else if (m == 20) { const Col c = decode_RGBA8(kLodPalette[hit.primId % 20]); color = f3(c.x, c.y, c.z); }
      else if (m == 7) { const Col c = decode_RGBA8(kPalette[geomId % 20]); color = f3(c.x, c.y, c.z); }
      else {
        color.x = std::round(hit.coords[0]) / 255.0f;
        color.y = hit.coords[0] - std::floor(hit.coords[0]);
        color.z = std::round(hit.coords[1]) / 255.0f;
      }
      f3 fin(0, 0, 0);
      for (uint32_t i = 0; i < n_lights; i++) {
        const ResolveLight& L = lights[i];
        const f3 lc(L.color[0], L.color[1], L.color[2]), ls(L.space[0], L.space[1], L.space[2]);
        if (L.type == 0) {
          const float q = std::max(0.0f, dot(norm, ls));
          fin = fin + lc * color * q;
        } else if (L.type == 1) {
          const f3 surf = rpos + hit.t * rdir;
          f3 dir = ls - surf;
          const float l = length(dir);
          dir = dir / l;
          const float q = std::max(0.0f, dot(norm, dir));
          fin = fin + lc * color * q / (l * l);
        } else {
          fin = fin + lc * color;
        }
      }
      res = Col{fin.x, fin.y, fin.z, 1};
    } else {
      const float z = hit.t, z_near = 0.1f, z_far = 10.0f;
      switch (m) {
        case 0: res = Col{1.0f, 1.0f, 1.0f, 1.0f}; break;
        case 2: { const float d = (1.0f / z - 1.0f / z_near) / (1.0f / z_far - 1.0f / z_near); res = Col{d, d, d, 1}; } break;
        case 3: { const float d = ((z - z_near) / (z_far - z_near)); res = Col{d, d, d, 1.0f}; } break;
        case 19: {
          const f3 ip = rpos + z * rdir;
          float zt = (ip.z + 1.f) * 2.f;
          zt = zt - int(zt);
          if (zt < 0.f) zt = 1.f + zt;
          const f3 hsv(zt, 1.f, 1.f);
          float rc[4] = {hsv.z, hsv.z, hsv.z, 1.f};
          const float Vmin = (1.f - hsv.y) * hsv.z;
          float H_i = 0.f;
          const float a = (hsv.z - Vmin) * std::modf(hsv.x * 6.f, &H_i);
          uint32_t hi = uint32_t(H_i);
          hi = hi <= 5u ? hi : 5u;
          rc[(2u + (hi >> 1)) % 3u] = Vmin;
          if ((hi & 1u) != 0u) rc[hi >> 1] = hsv.z - a;
          else rc[(1u + (hi >> 1)) % 3u] = Vmin + a;
          res = Col{rc[0], rc[1], rc[2], 1.f};
        } break;
        case 4: { const float d = 1 - ((z - z_near) / (z_far - z_near)); res = Col{d, d, d, 1}; } break;
        case 5: res = decode_RGBA8(kPalette[hit.primId % 20]); break;
        case 6: res = decode_RGBA8(kPalette[type % 20]); break;
        case 8: {
          const f3 v(std::abs(norm.x), std::abs(norm.y), std::abs(norm.z));
          res = Col{0.25f, 0.25f, 0.25f, 0.5f} +
                0.5f * Col{std::pow(v.x, 1.0f / 1.5f), std::pow(v.y, 1.0f / 1.5f), std::pow(v.z, 1.0f / 1.5f), 1.0f};
        } break;
        case 9: res = Col{hit.coords[0], hit.coords[1], 1 - hit.coords[0] - hit.coords[1], 1}; break;
        case 10:
          if (hit.primId == 0) res = Col{0.5f, 0.5f, 0.5f, 1.0f};
          else if (hit.primId <= 64) {
            const float q = lm_clamp(float(hit.primId) / 64, 0.0f, 1.0f);
            res = q * Col{0, 0, 1, 1} + (1.0f - q) * Col{0, 1, 0, 1};
          } else {
            const float q = lm_clamp(float(hit.primId - 64) / (256 - 64), 0.0f, 1.0f);
            res = q * Col{1, 0, 0, 1} + (1.0f - q) * Col{0, 0, 1, 1};
          }
          break;
        case 15: res = Col{tc.x, tc.y, 0, 1}; break;
        case 24:
          if (hit.primId == 0) res = Col{0.5, 0.5, 0.5, 1};
          else if (hit.primId == 1) res = Col{0.5, 1, 0.5, 1};
          else if (hit.primId == 2) res = Col{0.5, 0.5, 1, 1};
          else if (hit.primId <= 64) {
            const float q = lm_clamp(float(hit.primId) / 64, 0.0f, 1.0f);
            res = q * Col{1, 0, 0, 1} + (1 - q) * Col{0.5, 0.5, 1, 1};
          } else res = Col{1, 0, 0, 1};
          break;
        case 25: res = decode_RGBA8(kPalette[hit.primId % 20]); break;
        case 26:
          if (hit.instId == 0) res = Col{0.3f, 0.3f, 0.3f, 1};
          else if (hit.instId == 1) res = Col{0, 100, 0, 255} / 255.0f;
          else if (hit.instId == 2) res = Col{0, 200, 0, 255} / 255.0f;
          else if (hit.instId == 3) res = Col{50, 250, 50, 255} / 255.0f;
          else if (hit.instId == 4) res = Col{150, 250, 150, 255} / 255.0f;
          else if (hit.instId == 5) res = Col{250, 250, 250, 255} / 255.0f;
          else if (hit.instId <= 20)
            res = Col{250, (1000 - 50 * hit.instId) * 0.333f, (1000 - 50 * hit.instId) * 0.333f, 255} / 255.0f;
          else res = Col{1, 0, std::min(1.0f, 0.025f * (hit.instId - 20)), 1};
          break;
        default: break;
      }
    }
    if (hl[0] == 1 && hl[1] == hit.primId && hl[2] == hit.geomId && hl[3] == hit.instId) res = res + Col{0.5, 0, 0, 0};
    res.w = aa_frame == 0 ? 1.0f : hit.t;
    return res;
  }
};

}  // namespace

extern "C" {

// colour: float4[width*height] (row-major); sample_id > 0 accumulates. Render mode, normal mode
// and spp from sc->preset; lights: 32-B Light records.
void oracle_resolve(const OracleScene* sc, const uint32_t* orig_indices, const void* gbuf, uint32_t order,
                    const float* proj_inv, const float* wv_inv, uint32_t width, uint32_t height, uint32_t sample_id,
                    uint32_t aa_frame, const void* lights, uint32_t n_lights, const float* bg, const uint32_t* hl,
                    float* colour) {
  Resolver r{*sc, orig_indices, proj_inv, wv_inv, width, height, 8 * ((width + 7) / 8), 8 * ((height + 7) / 8),
             aa_frame, static_cast<const ResolveLight*>(lights), n_lights, bg, hl, Preset{}};
  r.p = sc->preset;
  if (r.p.spp == 0) r.p.spp = 1;
  const Hit* g = static_cast<const Hit*>(gbuf);
  for (uint32_t tid = 0; tid < width * height; ++tid) {
    uint32_t slot = tid;
    if (order == 1) {
      const uint32_t x = tid % width, y = tid / width;
      const uint32_t local = (y & 3u) * 4 + (x & 3u), bx = x >> 2, by = y >> 2;
      const uint32_t localH = (by & 1u) * 2 + (bx & 1u), wbH = (r.pw >> 2) >> 1;
      slot = ((bx >> 1) + (by >> 1) * wbH) * 64 + localH * 16 + local;
    }
    const Col c = r.resolve(tid, g[slot]);
    float* o = colour + 4 * size_t(tid);
    if (sample_id == 0) { o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = c.w; }
    else { o[0] += c.x; o[1] += c.y; o[2] += c.z; o[3] += c.w; }
  }
}

// MultiRenderer::Tonemap (MultiRenderer_resolve.cpp:538-544)
void oracle_tonemap(const float* colour, uint64_t n, uint32_t samples, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const float* c = colour + 4 * i;
    const float s = float(samples);
    const float f[4] = {c[0] / s, c[1] / s, c[2] / s, c[3] / s};
    uint32_t col[4];
    for (int k = 0; k < 4; ++k) col[k] = uint32_t(255 * lm_clamp(f[k], 0.0f, 1.0f));
    out[i] = (col[3] << 24) | (col[2] << 16) | (col[1] << 8) | col[0];
  }
}

}  // extern "C"
