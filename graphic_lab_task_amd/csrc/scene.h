// Host-side scene assembly (the BVHRT host half) and its device mirror.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "kernels.h"
#include "scene_view.h"

namespace lrt {

// m_RotAddTable: 48 x 48 compositions of the lattice rotations (entry i * 48 + j: i, then j)
const std::vector<uint8_t>& rot_add_table();
// An SCom2 tree of the shape the reference's fast traversal takes (no multi-nodes or channels,
// 3^3 values, one-word links)
bool scom2_plain_tree(const SCom2Header& h);

class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  hipError_t upload(const void* src, size_t bytes);
  hipError_t reserve(size_t bytes);
  void release();
  void* ptr() const { return p_; }
  size_t bytes() const { return used_; }

 private:
  void* p_ = nullptr;
  size_t cap_ = 0;
  size_t used_ = 0;
};

// Mirrors the member arrays of BVHRT (MultiRenderer/BVH2Common.h:262-365).
struct Scene {
  std::string build_name, layout_name;
  int prims_in_leaf = 2;
  MultiRenderPreset preset = default_preset();
  float lod_thr = 0.0f;

  std::vector<GeomData> geomData;
  std::vector<uint32_t> geomTags;
  std::vector<uint2> startEnd;
  std::vector<float4> vertPos, vertNorm;
  std::vector<uint32_t> indices, primIndices, primIdCount;
  std::vector<uint32_t> origIndices;   // triangle indices in the caller's order (g-buffer resolve)
  std::vector<BVHNodePair> nodePairs;
  std::vector<BVHNode> origNodes;
  std::vector<SdfFrameOctreeNode> frameNodes;
  std::vector<uint32_t> frameRoots;
  std::vector<SdfFrameOctreeTexNode> frameTexNodes;   // m_SdfFrameOctreeTexNodes
  std::vector<uint32_t> frameTexRoots;                // m_SdfFrameOctreeTexRoots
  std::vector<SdfSVSNode> svsNodes;
  std::vector<uint32_t> svsRoots;
  std::vector<SdfSBSNode> sbsNodes;
  std::vector<uint32_t> sbsData;
  std::vector<float> sbsDataF;
  std::vector<uint2> sbsFSlice;     // per brick-set geometry: (first values_f entry, count)
  std::vector<uint8_t> sbsDiffChecked;
  std::vector<SbsMaskGeom> sbsMaskGeoms;  // occupancy-mask layout of every brick set (commit)
  std::vector<uint32_t> sbsMaskBase;      // per brick set: first mask word
  size_t sbsMaskWords = 0;  // per brick-set geometry: 0 unchecked, 1 usable by the adjoint, 2 not
  std::vector<uint32_t> sbsRoots;
  std::vector<SdfSBSHeader> sbsHeaders;
  std::vector<SdfMultiOctreeNode> multiNodes;
  std::vector<float> multiValues;
  std::vector<float> gridData;
  std::vector<uint32_t> gridOffsets, gridSizes;   // sizes: 3 words per grid
  std::vector<GraphicsPrimHeader> gprimHeaders;
  std::vector<float4> gprimPoints;
  std::vector<uint32_t> coctData;    // words of every COctreeV3 geometry in turn
  std::vector<CoctGeom> coctGeoms;
  std::vector<int4> coctMods;        // 3 x 48 index modifiers per COctreeV3 geometry
  std::vector<SdfDAGNode> dagNodes;
  std::vector<SdfDAGChildEdge> dagChildEdges;
  std::vector<SdfDAGDataEdge> dagDataEdges;
  std::vector<float> dagDistances;
  std::vector<DagGeom> dagGeoms;
  std::vector<int4> dagMods;         // 2 x 48 index modifiers per DAG geometry
  std::vector<uint32_t> scom2Data;   // each SCom2 tree's nodes then bricks
  std::vector<SCom2Header> scom2Headers;
  std::vector<int4> scom2Mods;       // 3 x 48 index modifiers per SCom2 tree
  std::vector<InstanceData> instances;
  std::vector<BVHNode> tlas;

  // Device layout derived at commit (same traversal, fewer dependent loads per step).
  std::vector<BVHNodePair> devNodes;
  std::vector<float4> triSoup;
  std::vector<FrameLeaf> frameLeaves;
  std::vector<float4> frameTex;      // 4 float4 per textured frame-octree leaf
  uint32_t cacheGeom = 0xFFFFFFFFu, cachePairs = 0;

  bool committed = false;
  int device = 0;
  std::map<std::string, DevBuf> dev;
  SceneView view{};
  uint32_t* d_overflow = nullptr;
  int block = 256, lds_stack = 16, xcd_remap = 1;
  uint32_t ray_groups = 32;  // pass-1 lanes per group dealt to distant parts of the batch (0: off)
  uint32_t budget = 64, tail_waves = 1024;
  // Triangle leaf postponement in pass 1: -1 automatic (on for a triangle-only scene with one
  // instance), 0 off, 1 on.
  int leaf_postpone = -1;
  int persistent = 0;   // 1: persistent single-pass trace kernel with per-lane ray refill

  // BLAS builder: false the host binned SAH, true the GPU linear BVH (build names "gpu...").
  bool gpu_builder = false;
  BuildStats lastBuild;            // stats of the most recent BLAS build
  // What the next commit must redo: geometry (every array) or the instance set (TLAS rebuild).
  // Instance updates alone leave both false: the commit then uploads the instances and refits
  // the TLAS on the GPU (tlas_update 1) or rebuilds it on the host (tlas_update 0).
  bool geom_dirty = true, inst_set_dirty = true;
  int tlas_update = 1;
  std::vector<uint32_t> tlasParent;   // parent node of every TLAS node (0xFFFFFFFF: root's children)
  double tlasAreaBuilt = 0.0;         // summed node surface area of the last full TLAS build
  double tlasAreaNow = 0.0;           // the same after the last refit
  uint32_t lastCommitKind = 0;        // LRT_COMMIT_* of the last commit
  bool postpone_triangles() const;

  // Work memory of one stream: the tail queue (256-B header + list), the deep-stack spill and
  // the occupancy masks of calls whose values differ from the committed ones. A call holds
  // `mu` from sizing the buffers until its last kernel is enqueued, so that two threads tracing
  // on one stream never interleave their passes over the same queue; a buffer only grows, and
  // only after the stream has drained the work that may still read it.
  struct StreamWork {
    std::mutex mu;
    DevBuf buf;
    size_t cap_rays = 0;
    DevBuf mask;
  };
  std::mutex work_mu;   // guards the map only
  std::map<void*, std::unique_ptr<StreamWork>> work;
  StreamWork& stream_work(void* stream);
  // With w.mu held: queue, spill and persistent-kernel counters for n rays on `stream`.
  hipError_t work_for(StreamWork& w, void* stream, size_t n, uint32_t** spill, TailQueue** queue, uint32_t** pcount);
  // With w.mu held: the stream's occupancy-mask buffer.
  hipError_t mask_for(StreamWork& w, void* stream, uint32_t** mask);

  // Staging for the host-buffer trace calls (lrt_trace_closest_host, lrt_query_one): grows
  // only, so a picking loop allocates on its first call and never again.
  std::mutex host_mu;
  DevBuf host_io;

  // Commit generation: bumped by every change that invalidates device pointers handed out
  // earlier (commit, clear, add, instance update). Batch plans record it and refuse to run once
  // it moved; the weak_ptr of a plan also sees the scene's destruction.
  std::shared_ptr<uint64_t> generation = std::make_shared<uint64_t>(0);
  void invalidate() {
    committed = false;
    ++*generation;
  }

  ~Scene();

  void clear_geom();
  void clear_scene();
  uint32_t add_triangles(const float* vpos, size_t n_verts, size_t v_stride_bytes, const float* vnorm,
                         const uint32_t* idx, size_t n_idx);
  uint32_t add_aabb(uint32_t tag, const BVHNode* boxes, size_t n);
  uint32_t add_sdf_frame_octree(const SdfFrameOctreeNode* nodes, size_t n);
  uint32_t add_sdf_frame_octree_tex(const SdfFrameOctreeTexNode* nodes, size_t n);
  uint32_t add_sdf_svs(const SdfSVSNode* nodes, size_t n);
  uint32_t add_sdf_grid(const uint32_t size[3], const float* data);
  uint32_t add_graphics_prims(const GraphicsPrimHeader& header, const float4* points, size_t n_points);
  uint32_t add_sdf_sbs(const SdfSBSHeader& header, const SdfSBSNode* nodes, size_t n_nodes, const uint32_t* values,
                       size_t n_values, const float* values_f, size_t n_values_f);
  uint32_t add_sdf_dag(const SdfDAGHeader& h, const SdfDAGNode* nodes, size_t n_nodes,
                       const SdfDAGChildEdge* child_edges, size_t n_child_edges, const SdfDAGDataEdge* data_edges,
                       size_t n_data_edges, const float* distances, size_t n_distances);
  uint32_t add_scom2(const SCom2Header& header, const uint32_t* nodes, size_t n_nodes, const uint32_t* bricks,
                     size_t n_bricks);
  uint32_t add_coctree_v3(const COctreeV3Header& h, const uint32_t* data, size_t size, uint32_t bvh_level);
  uint32_t add_sdf_multi_octree(const SdfMultiOctreeNode* nodes, size_t n_nodes, const float* values,
                                size_t n_values);
  uint32_t add_instance(uint32_t geom_id, const float4x4& m);
  void update_instance(uint32_t inst_id, const float4x4& m);
  void build_host();   // TLAS over the instance boxes (host only) and its parent array
  hipError_t commit_instances();   // commit after instance updates only (geometry unchanged)
  void update_cache_geom();
  void build_device_layout();
  hipError_t commit(); // device upload of every array
  void set_preset(const MultiRenderPreset& p);
  uint32_t leaf_kinds() const;
  SceneView make_view() const;
};

}  // namespace lrt
