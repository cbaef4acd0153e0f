// Host-side scene assembly: restates the BVHRT host methods
// (MultiRenderer/BVH2Common_host.cpp) over our own builder, then mirrors every array
// into device memory at commit.
#include "scene.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "host_math.h"

namespace lrt {

hipError_t DevBuf::reserve(size_t bytes) {
  if (bytes <= cap_) { used_ = bytes; return hipSuccess; }
  release();
  hipError_t e = hipMalloc(&p_, bytes);
  if (e != hipSuccess) { p_ = nullptr; return e; }
  cap_ = bytes;
  used_ = bytes;
  return hipSuccess;
}

hipError_t DevBuf::upload(const void* src, size_t bytes) {
  if (bytes == 0) { used_ = 0; return hipSuccess; }
  hipError_t e = reserve(bytes);
  if (e != hipSuccess) return e;
  return hipMemcpy(p_, src, bytes, hipMemcpyHostToDevice);
}

void DevBuf::release() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  cap_ = used_ = 0;
}

Scene::StreamWork& Scene::stream_work(void* stream) {
  std::lock_guard<std::mutex> lock(work_mu);
  std::unique_ptr<StreamWork>& w = work[stream];
  if (!w) w = std::make_unique<StreamWork>();
  return *w;
}

hipError_t Scene::work_for(StreamWork& w, void* stream, size_t n, uint32_t** spill, TailQueue** queue,
                           uint32_t** pcount) {
  // 256-B header: the tail queue's counters and list pointer at 0, the persistent kernel's two
  // counters at byte 64; then the queue list (4 B per ray) and the spill, sized for the smallest
  // LDS part so that a later launch-config change never overruns it, and for a persistent grid's
  // lane slots (up to SPILL_SLACK_RAYS past n)
  const size_t head = 256;
  const size_t list_bytes = (n * 4 + 255) / 256 * 256;
  const size_t need = head + list_bytes + (n + SPILL_SLACK_RAYS) * size_t(STACK_SIZE - 8) * 4;
  if (w.cap_rays < n || w.buf.bytes() < need) {
    // earlier calls on this stream may still be reading the old buffer
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return e;
    e = w.buf.reserve(need);
    if (e != hipSuccess) {
      w.cap_rays = 0;
      return e;
    }
    uint8_t header[256] = {};
    TailQueue q{0u, 0u, reinterpret_cast<uint32_t*>(static_cast<char*>(w.buf.ptr()) + head)};
    std::memcpy(header, &q, sizeof(q));
    e = hipMemcpy(w.buf.ptr(), header, sizeof(header), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      w.cap_rays = 0;
      return e;
    }
    w.cap_rays = n;
  }
  *queue = static_cast<TailQueue*>(w.buf.ptr());
  *pcount = reinterpret_cast<uint32_t*>(static_cast<char*>(w.buf.ptr()) + 64);
  *spill = reinterpret_cast<uint32_t*>(static_cast<char*>(w.buf.ptr()) + head + (w.cap_rays * 4 + 255) / 256 * 256);
  return hipSuccess;
}

Scene::~Scene() {
  work.clear();
  dev.clear();
  if (d_overflow) (void)hipFree(d_overflow);
}

namespace {

struct Box4 {
  float4 mn, mx;
};
Box4 empty_box4() {
  const float inf = std::numeric_limits<float>::infinity();
  return {make4(inf, inf, inf, inf), make4(-inf, -inf, -inf, -inf)};
}
// Box4f::include (LiteMath.h:3422-3426), all four components
void include4(Box4& b, float4 p) {
  b.mn = make4(fmin_ref(b.mn.x, p.x), fmin_ref(b.mn.y, p.y), fmin_ref(b.mn.z, p.z), fmin_ref(b.mn.w, p.w));
  b.mx = make4(fmax_ref(b.mx.x, p.x), fmax_ref(b.mx.y, p.y), fmax_ref(b.mx.z, p.z), fmax_ref(b.mx.w, p.w));
}

// add_border_nodes_rec for frame octrees (BVH2Common_host.cpp:1348-1383): one box per leaf
// whose corner values change sign; leftOffset keeps the octree node index. The textured octree's
// version (:1392-1427) tests a positive value against -sqrt(3)*d instead of 0, which every
// positive value passes, so both node types share this walk. Children must follow their parent
// (no cycle can then form) and the octree be at most FRAME_MAX_DEPTH levels deep, which bounds
// this recursion; the reference walks any offsets it is given.
constexpr uint32_t FRAME_MAX_DEPTH = 30;
template <class NodeT>
void add_border_nodes_rec(const NodeT* nodes, size_t n_nodes, std::vector<BVHNode>& out, uint32_t idx,
                          float3 p, float d, uint32_t depth = 0) {
  const uint32_t ofs = nodes[idx].offset;
  if (ofs != 0 && size_t(ofs) + 8 > n_nodes) throw std::invalid_argument("frame octree child offset out of range");
  if (ofs != 0 && ofs <= idx) throw std::invalid_argument("frame octree children must follow their parent");
  if (ofs != 0 && depth >= FRAME_MAX_DEPTH) throw std::invalid_argument("frame octree deeper than 30 levels");
  if (ofs == 0) {
    bool less = false, more = false;
    for (int i = 0; i < 8; i++) {
      if (nodes[idx].values[i] <= 0) less = true;
      else if (nodes[idx].values[i] >= 0) more = true;
    }
    if (less && more) {
      const float3 min_pos = 2.0f * (d * p) - 1.0f;
      const float3 max_pos = min_pos + 2.0f * d * make3(1, 1, 1);
      BVHNode n;
      n.boxMin = min_pos;
      n.boxMax = max_pos;
      n.leftOffset = idx;
      n.escapeIndex = 0;
      out.push_back(n);
    }
  } else {
    for (int i = 0; i < 8; i++) {
      const float ch_d = d / 2;
      const float3 ch_p = 2.0f * p + make3(float((i & 4) >> 2), float((i & 2) >> 1), float(i & 1));
      add_border_nodes_rec(nodes, n_nodes, out, ofs + i, ch_p, ch_d, depth + 1);
    }
  }
}

}  // namespace

void Scene::clear_scene() {
  instances.clear();
  tlas.clear();
  tlasParent.clear();
  inst_set_dirty = true;
  invalidate();
}

void Scene::clear_geom() {
  geomData.clear(); geomTags.clear(); startEnd.clear();
  vertPos.clear(); vertNorm.clear(); indices.clear(); primIndices.clear(); primIdCount.clear(); origIndices.clear();
  nodePairs.clear(); origNodes.clear();
  frameNodes.clear(); frameRoots.clear(); frameTexNodes.clear(); frameTexRoots.clear(); svsNodes.clear(); svsRoots.clear();
  sbsNodes.clear(); sbsData.clear(); sbsDataF.clear(); sbsRoots.clear(); sbsHeaders.clear(); sbsFSlice.clear(); sbsDiffChecked.clear();
  multiNodes.clear(); multiValues.clear();
  gridData.clear(); gridOffsets.clear(); gridSizes.clear();
  gprimHeaders.clear(); gprimPoints.clear();
  coctData.clear(); coctGeoms.clear(); coctMods.clear();
  dagNodes.clear(); dagChildEdges.clear(); dagDataEdges.clear(); dagDistances.clear(); dagGeoms.clear(); dagMods.clear();
  scom2Data.clear(); scom2Headers.clear(); scom2Mods.clear();
  geom_dirty = true;
  clear_scene();
}

// BVHRT::AddGeom_Triangles3f (BVH2Common_host.cpp:307-359) + AppendTreeData (:287-301)
uint32_t Scene::add_triangles(const float* vpos, size_t n_verts, size_t v_stride_bytes, const float* vnorm,
                              const uint32_t* idx, size_t n_idx) {
  if (v_stride_bytes % 4 != 0 || v_stride_bytes < 12) throw std::invalid_argument("vertex stride must be a multiple of 4 and >= 12");
  if (n_idx % 3 != 0) throw std::invalid_argument("index count must be a multiple of 3");
  for (size_t i = 0; i < n_idx; ++i)
    if (idx[i] >= n_verts) throw std::invalid_argument("triangle index out of range");
  const size_t vStride = v_stride_bytes / 4;
  const size_t oldSizeVert = vertPos.size();
  const size_t oldSizeInd = indices.size();
  vertPos.resize(oldSizeVert + n_verts);
  vertNorm.resize(oldSizeVert + n_verts);
  Box4 bbox = empty_box4();
  for (size_t i = 0; i < n_verts; i++) {
    const float4 v = make4(vpos[i * vStride + 0], vpos[i * vStride + 1], vpos[i * vStride + 2], 1.0f);
    vertPos[oldSizeVert + i] = v;
    vertNorm[oldSizeVert + i] = vnorm ? make4(vnorm[i * vStride + 0], vnorm[i * vStride + 1], vnorm[i * vStride + 2], 1.0f)
                                      : make4(1.0f, 0.0f, 0.0f, 1.0f);
    include4(bbox, v);
  }
  GeomData g{};
  g.boxMin = bbox.mn;
  g.boxMax = bbox.mx;
  g.offset = uint2{uint32_t(oldSizeInd), uint32_t(oldSizeVert)};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_MESH_TRIANGLE;
  geomData.push_back(g);
  geomTags.push_back(TAG_TRIANGLE);

  FatBlas bvh = gpu_builder ? gpu_build_blas_triangles(vertPos.data() + oldSizeVert, n_verts, idx, n_idx, prims_in_leaf, &lastBuild)
                            : build_blas_triangles(vertPos.data() + oldSizeVert, n_verts, idx, n_idx, prims_in_leaf, &lastBuild);
  nodePairs.insert(nodePairs.end(), bvh.nodes.begin(), bvh.nodes.end());
  primIndices.insert(primIndices.end(), bvh.indices.begin(), bvh.indices.end());
  indices.resize(oldSizeInd + bvh.indices.size() * 3);
  for (size_t i = 0; i < bvh.indices.size(); i++) {
    const uint32_t triId = bvh.indices[i];
    indices[oldSizeInd + 3 * i + 0] = idx[triId * 3 + 0];
    indices[oldSizeInd + 3 * i + 1] = idx[triId * 3 + 1];
    indices[oldSizeInd + 3 * i + 2] = idx[triId * 3 + 2];
  }
  origIndices.insert(origIndices.end(), idx, idx + n_idx);
  const size_t oldSize = primIdCount.size();
  primIdCount.insert(primIdCount.end(), bvh.startCount.begin(), bvh.startCount.end());
  startEnd.push_back(uint2{uint32_t(oldSize), uint32_t(primIdCount.size())});
  geom_dirty = true;
  invalidate();
  return uint32_t(startEnd.size() - 1);
}

// BVHRT::AddGeom_AABB (BVH2Common_host.cpp:441-459)
uint32_t Scene::add_aabb(uint32_t tag, const BVHNode* boxes, size_t n) {
  FatBlas bvh = gpu_builder ? gpu_build_blas_boxes(boxes, n, &lastBuild) : build_blas_boxes(boxes, n, &lastBuild);
  nodePairs.insert(nodePairs.end(), bvh.nodes.begin(), bvh.nodes.end());
  const size_t oldSize = primIdCount.size();
  primIdCount.resize(oldSize + n);
  for (size_t i = 0; i < n; i++) primIdCount[oldSize + i] = uint32_t(i);
  startEnd.push_back(uint2{uint32_t(oldSize), uint32_t(primIdCount.size())});
  geomTags.push_back(tag);
  geom_dirty = true;
  invalidate();
  return uint32_t(startEnd.size() - 1);
}

// BVHRT::AddGeom_SdfFrameOctree (BVH2Common_host.cpp:501-534). The reference keeps the leaf
// boxes of the most recently added octree only (m_origNodes = orig_nodes); here they are
// appended and GeomData.offset.y holds each octree's first box, which is 0 for the first
// octree and therefore identical for single-octree scenes.
uint32_t Scene::add_sdf_frame_octree(const SdfFrameOctreeNode* nodes, size_t n) {
  if (n == 0 || n >= (1u << 28)) throw std::invalid_argument("frame octree size out of range");
  std::vector<BVHNode> boxes;
  add_border_nodes_rec(nodes, n, boxes, 0, make3(0, 0, 0), 1.0f);
  if (boxes.empty()) throw std::invalid_argument("frame octree has no surface leaves");
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(frameRoots.size()), uint32_t(origNodes.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SDF_FRAME_OCTREE;
  geomData.push_back(g);
  const uint32_t n_offset = uint32_t(frameNodes.size());
  frameRoots.push_back(n_offset);
  frameNodes.insert(frameNodes.end(), nodes, nodes + n);
  for (size_t i = n_offset; i < frameNodes.size(); i++) frameNodes[i].offset += n_offset;
  origNodes.insert(origNodes.end(), boxes.begin(), boxes.end());
  return add_aabb(TAG_SDF_NODE, boxes.data(), boxes.size());
}

// BVHRT::AddGeom_SdfFrameOctreeTex (BVH2Common_host.cpp:697-731): a frame octree whose nodes also
// carry texture coordinates; traced as the frame octree, hits report TYPE_SDF_FRAME_OCTREE_TEX and
// the coordinates interpolated at the hit point. Leaf boxes are kept per octree as for frame
// octrees.
uint32_t Scene::add_sdf_frame_octree_tex(const SdfFrameOctreeTexNode* nodes, size_t n) {
  if (n == 0 || n >= (1u << 28)) throw std::invalid_argument("textured frame octree size out of range");
  std::vector<BVHNode> boxes;
  add_border_nodes_rec(nodes, n, boxes, 0, make3(0, 0, 0), 1.0f);
  if (boxes.empty()) throw std::invalid_argument("textured frame octree has no surface leaves");
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(frameTexRoots.size()), uint32_t(origNodes.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SDF_FRAME_OCTREE_TEX;
  geomData.push_back(g);
  const uint32_t n_offset = uint32_t(frameTexNodes.size());
  frameTexRoots.push_back(n_offset);
  frameTexNodes.insert(frameTexNodes.end(), nodes, nodes + n);
  for (size_t i = n_offset; i < frameTexNodes.size(); i++) frameTexNodes[i].offset += n_offset;
  origNodes.insert(origNodes.end(), boxes.begin(), boxes.end());
  return add_aabb(TAG_SDF_NODE, boxes.data(), boxes.size());
}

// BVHRT::AddGeom_SdfSVS (BVH2Common_host.cpp:536-597): every voxel is a BLAS leaf whose box
// follows from its position in its level. The boxes are also appended to the leaf-box array (the
// reference keeps them only for the BLAS build), where the device leaf records take them from;
// GeomData.offset.y holds the first of them, as for frame octrees.
uint32_t Scene::add_sdf_svs(const SdfSVSNode* nodes, size_t n) {
  if (n == 0 || n >= (1u << 28)) throw std::invalid_argument("sparse voxel set size out of range");
  std::vector<BVHNode> boxes(n);
  for (size_t i = 0; i < n; i++) {
    const float px = float(nodes[i].pos_xy >> 16), py = float(nodes[i].pos_xy & 0x0000FFFF);
    const float pz = float(nodes[i].pos_z_lod_size >> 16), sz = float(nodes[i].pos_z_lod_size & 0x0000FFFF);
    if (sz == 0.0f) throw std::invalid_argument("sparse voxel set node with level size 0");
    boxes[i].boxMin = make3(-1, -1, -1) + (2.0f * make3(px, py, pz)) / sz;
    boxes[i].boxMax = boxes[i].boxMin + (2.0f * make3(1, 1, 1)) / sz;
  }
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(svsRoots.size()), uint32_t(origNodes.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SDF_SVS;
  geomData.push_back(g);
  svsRoots.push_back(uint32_t(svsNodes.size()));
  svsNodes.insert(svsNodes.end(), nodes, nodes + n);
  origNodes.insert(origNodes.end(), boxes.begin(), boxes.end());
  return add_aabb(TAG_SDF_NODE, boxes.data(), boxes.size());
}

// BVHRT::AddGeom_SdfGrid (BVH2Common_host.cpp:466-499): a dense grid of distances over the unit
// cube, sphere-traced as a whole; the BLAS holds the cube's two halves (GetBoxes_SdfGrid,
// :1337-1346).
uint32_t Scene::add_sdf_grid(const uint32_t size[3], const float* data) {
  const uint64_t n = uint64_t(size[0]) * size[1] * size[2];
  if (n == 0 || n >= (1ull << 28)) throw std::invalid_argument("SDF grid size out of range");
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(gridOffsets.size()), 0};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SDF_GRID;
  geomData.push_back(g);
  gridOffsets.push_back(uint32_t(gridData.size()));
  gridSizes.insert(gridSizes.end(), size, size + 3);
  gridData.insert(gridData.end(), data, data + n);
  BVHNode boxes[2]{};
  boxes[0].boxMin = make3(-1, -1, -1);
  boxes[0].boxMax = make3(1, 1, 0);
  boxes[1].boxMin = make3(-1, -1, 0);
  boxes[1].boxMax = make3(1, 1, 1);
  return add_aabb(TAG_SDF_GRID, boxes, 2);
}

namespace {
// getDiskAABB (BVH2Common_host.cpp:754-765): box of a disk of radius `rad` around `pt` facing `norm`
BVHNode disk_box(float3 pt, float3 norm, float rad) {
  BVHNode r{};
  norm = normalize3(norm) * rad;
  const float xy = std::sqrt(norm.x * norm.x + norm.y * norm.y);
  const float xz = std::sqrt(norm.x * norm.x + norm.z * norm.z);
  const float yz = std::sqrt(norm.y * norm.y + norm.z * norm.z);
  r.boxMin = pt - make3(yz, xz, xy);
  r.boxMax = pt + make3(yz, xz, xy);
  return r;
}
inline float comp3(const float3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
}  // namespace

// BVHRT::AddGeom_GraphicsPrim (BVH2Common_host.cpp:767-903): one BLAS box per primitive, from the
// primitive's extent (a line: its part inside the unit cube, as the reference clips it).
uint32_t Scene::add_graphics_prims(const GraphicsPrimHeader& header, const float4* points, size_t n_points) {
  if (header.prim_type > GRAPH_PRIM_BOX_COLOR) throw std::invalid_argument("unknown graphics primitive type");
  const uint32_t step = graphics_prim_len(header.prim_type);
  if (n_points == 0 || n_points % step != 0) throw std::invalid_argument("graphics primitive points: not whole primitives");
  if (n_points / step >= (1u << 28)) throw std::invalid_argument("too many graphics primitives");
  float3 mn = make3(1, 1, 1), mx = make3(-1, -1, -1);
  std::vector<BVHNode> boxes;
  const uint32_t pt = header.prim_type;
  for (size_t i = 0; i < n_points; i += step) {
    BVHNode nd{};
    if (pt == GRAPH_PRIM_POINT || pt == GRAPH_PRIM_POINT_COLOR) {
      const float4 p = points[i];
      nd.boxMin = xyz(p) - p.w;
      nd.boxMax = xyz(p) + p.w;
    } else if (pt == GRAPH_PRIM_LINE || pt == GRAPH_PRIM_LINE_COLOR) {
      const float4 p1 = points[i], p2 = points[i + 1];
      const float3 pos = xyz(p2);
      const float3 dir = normalize3(xyz(p1) - pos);
      float tx = 0.f, ty = 0.f;
      bool on_pos = false, on_neg = false;
      for (int a = 0; a < 3; ++a) {
        const float da = comp3(dir, a), pa = comp3(pos, a);
        if (da > 1e-4f) {
          float dist = 1.f - pa;
          if (std::abs(dist) < 1e-4f) {
            on_pos = true;
          } else {
            const float t1 = dist / da;
            if (tx == 0.f || std::abs(t1) < std::abs(tx)) tx = t1;
          }
          dist = -1.f - pa;
          if (std::abs(dist) < 1e-4f) {
            on_neg = true;
          } else {
            const float t2 = dist / da;
            if (ty == 0.f || std::abs(t2) < std::abs(ty)) ty = t2;
          }
        }
      }
      const float3 i1 = on_pos ? pos : pos + tx * dir;
      const float3 i2 = on_neg ? pos : pos + ty * dir;
      nd = disk_box(i1, dir, p1.w);
      const BVHNode o = disk_box(i2, dir, p1.w);
      nd.boxMin = min3(nd.boxMin, o.boxMin);
      nd.boxMax = max3(nd.boxMax, o.boxMax);
    } else if (pt == GRAPH_PRIM_LINE_SEGMENT || pt == GRAPH_PRIM_LINE_SEGMENT_COLOR) {
      const float4 p1 = points[i], p2 = points[i + 1];
      const float3 dir = normalize3(make3(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z));
      nd = disk_box(xyz(p2), dir, p1.w);
      const BVHNode o = disk_box(xyz(p1), dir, p1.w);
      nd.boxMin = min3(nd.boxMin, o.boxMin);
      nd.boxMax = max3(nd.boxMax, o.boxMax);
    } else if (pt == GRAPH_PRIM_LINE_SEGMENT_DIR || pt == GRAPH_PRIM_LINE_SEGMENT_DIR_COLOR) {
      const float4 p1r = points[i];
      const float3 p2 = xyz(points[i + 1]), p1 = xyz(p1r);
      const float3 dir = p2 - p1;
      const float cone_ra = length3(p2 - p1) * 0.15f;
      const float3 pa = p2 * 0.7f + p1 * 0.3f;
      nd = disk_box(p1, dir, p1r.w);
      nd.boxMin = min3(nd.boxMin, p2);
      nd.boxMax = max3(nd.boxMax, p2);
      const BVHNode o = disk_box(pa, dir, cone_ra);
      nd.boxMin = min3(nd.boxMin, o.boxMin);
      nd.boxMax = max3(nd.boxMax, o.boxMax);
    } else {   // boxes
      const float4 p1r = points[i];
      const float3 p2 = xyz(points[i + 1]);
      nd.boxMin = min3(xyz(p1r), p2) - p1r.w;
      nd.boxMax = max3(xyz(p1r), p2) + p1r.w;
    }
    mn = min3(mn, nd.boxMin);
    mx = max3(mx, nd.boxMax);
    boxes.push_back(nd);
  }
  if (boxes.size() == 1) {
    boxes.resize(2);
    boxes[1].boxMin = boxes[0].boxMin - 0.001f;
    boxes[1].boxMax = boxes[0].boxMin;
  }
  GeomData g{};
  g.boxMin = make4(mn.x, mn.y, mn.z, 1);
  g.boxMax = make4(mx.x, mx.y, mx.z, 1);
  g.offset = uint2{uint32_t(gprimHeaders.size()), uint32_t(gprimPoints.size())};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_GRAPHICS_PRIM;
  geomData.push_back(g);
  gprimHeaders.push_back(header);
  gprimPoints.insert(gprimPoints.end(), points, points + n_points);
  return add_aabb(TAG_GRAPHICS_PRIM, boxes.data(), boxes.size());
}

// BVHRT::AddGeom_SdfMultiOctree (BVH2Common_host.cpp:970-1018): the octree is walked by its
// own leaf routine; the BLAS is the [-1,1]^3 box plus a tiny dummy box. Nodes must be stored
// after their parent (children_offset > index) and at most 15 levels deep (16-bit positions in
// the walk), and every node's voxel values must lie inside `values`.
uint32_t Scene::add_sdf_multi_octree(const SdfMultiOctreeNode* nodes, size_t n_nodes, const float* values,
                                     size_t n_values) {
  if (n_nodes == 0 || n_values == 0) throw std::invalid_argument("empty multi octree");
  if (n_nodes >= (1u << 28)) throw std::invalid_argument("multi octree too large");
  std::vector<uint8_t> depth(n_nodes, 0);
  for (size_t i = 0; i < n_nodes; ++i) {
    const SdfMultiOctreeNode& nd = nodes[i];
    if (uint64_t(nd.data_offset) + 8ull * nd.voxel_count > n_values)
      throw std::invalid_argument("multi octree node values out of range");
    if (nd.children_offset > 0) {
      if (nd.children_offset <= i || uint64_t(nd.children_offset) + 8 > n_nodes)
        throw std::invalid_argument("multi octree children must follow their parent and lie inside the node array");
      for (uint32_t c = 0; c < 8; ++c) {
        if (depth[i] + 1 > 15) throw std::invalid_argument("multi octree deeper than 15 levels");
        depth[nd.children_offset + c] = uint8_t(depth[i] + 1);
      }
    }
  }
  const uint32_t n_offset = uint32_t(multiNodes.size());
  const uint32_t v_offset = uint32_t(multiValues.size());
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{n_offset, 0};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = TYPE_SDF_MULTI_OCTREE;
  geomData.push_back(g);
  multiNodes.insert(multiNodes.end(), nodes, nodes + n_nodes);
  multiValues.insert(multiValues.end(), values, values + n_values);
  for (size_t i = n_offset; i < multiNodes.size(); i++) {
    if (multiNodes[i].children_offset > 0) multiNodes[i].children_offset += n_offset;
    multiNodes[i].data_offset += v_offset;
  }
  BVHNode boxes[2]{};
  boxes[0].boxMin = make3(-1, -1, -1);
  boxes[0].boxMax = make3(1, 1, 1);
  boxes[1].boxMin = make3(-1.0001f, -1.0001f, -1.0001f);
  boxes[1].boxMax = make3(-1, -1, -1);
  return add_aabb(TAG_SDF_OCTREE, boxes, 2);
}

// BVHRT::AddGeom_SdfSBS (BVH2Common_host.cpp:601-695)
uint32_t Scene::add_sdf_sbs(const SdfSBSHeader& header, const SdfSBSNode* nodes, size_t n_nodes,
                            const uint32_t* values, size_t n_values, const float* values_f, size_t n_values_f) {
  if (n_nodes == 0 || n_values == 0) throw std::invalid_argument("empty sparse brick set");
  if (header.brick_size < 1 || header.brick_size > 16) throw std::invalid_argument("brick_size must be 1..16");
  if (header.bytes_per_value != 1 && header.bytes_per_value != 2 && header.bytes_per_value != 4)
    throw std::invalid_argument("bytes_per_value must be 1, 2 or 4");
  const uint32_t layout = header.aux_data & SDF_SBS_NODE_LAYOUT_MASK;
  uint32_t type = TYPE_SDF_SBS;
  if (layout == SDF_SBS_NODE_LAYOUT_UNDEFINED || layout == SDF_SBS_NODE_LAYOUT_DX) type = TYPE_SDF_SBS;
  else if (layout == SDF_SBS_NODE_LAYOUT_DX_UV16) type = TYPE_SDF_SBS_TEX;
  else if (layout == SDF_SBS_NODE_LAYOUT_DX_RGB8 || layout == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F ||
           layout == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN)
    type = TYPE_SDF_SBS_COL;
  else
    throw std::invalid_argument("unsupported SBS node layout");
  const bool indexed = layout == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F || layout == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN;
  if (indexed && (values_f == nullptr || n_values_f == 0)) throw std::invalid_argument("indexed SBS layout needs values_f");
  // Bounds of every data access the intersector can make (checked on the host once).
  const uint32_t v_size = header.brick_size + 2 * header.brick_pad + 1;
  const uint32_t vpi = 4 / header.bytes_per_value;
  const bool with_neighbors = layout == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN;
  const size_t per_brick = indexed ? size_t(v_size) * v_size * v_size + 8 + (with_neighbors ? 27 : 0)
                                   : (size_t(v_size) * v_size * v_size + vpi - 1) / vpi +
                                         ((layout == SDF_SBS_NODE_LAYOUT_DX_RGB8 || layout == SDF_SBS_NODE_LAYOUT_DX_UV16) ? 8 : 0);
  for (size_t i = 0; i < n_nodes; ++i) {
    if (size_t(nodes[i].data_offset) + per_brick > n_values) throw std::invalid_argument("brick data out of range");
    const uint32_t lod = nodes[i].pos_z_lod_size & 0xFFFF;
    if (lod == 0) throw std::invalid_argument("brick lod_size is 0");
  }
  if (indexed) {
    const size_t v3 = size_t(v_size) * v_size * v_size;
    for (size_t i = 0; i < n_nodes; ++i) {
      const uint32_t* w = values + nodes[i].data_offset;
      for (size_t k = 0; k < v3; ++k)
        if (w[k] >= n_values_f) throw std::invalid_argument("distance index out of range");
      for (size_t k = v3; k < v3 + 8; ++k)
        if (size_t(w[k]) + 3 > n_values_f) throw std::invalid_argument("colour index out of range");
      // neighbour brick ids (ID32F_IRGB32F_IN, read by the smoothed normals as scene node ids):
      // every word of an indexed brick set is shifted by the set's first values_f entry below,
      // as in the reference, so they stay valid for the scene's first indexed set only
      const uint64_t f_shift = sbsDataF.size(), total_nodes = sbsNodes.size() + n_nodes;
      for (size_t k = v3 + 8; with_neighbors && k < v3 + 8 + 27; ++k) {
        const uint64_t nid = uint64_t(w[k]) + f_shift;
        if (nid != INVALID_IDX && nid >= total_nodes)
          throw std::invalid_argument("neighbour brick index out of range (or a later indexed brick set with neighbours)");
      }
    }
  }
  GeomData g{};
  g.boxMin = make4(-1, -1, -1, 1);
  g.boxMax = make4(1, 1, 1, 1);
  g.offset = uint2{uint32_t(sbsRoots.size()), 0};
  g.bvhOffset = uint32_t(nodePairs.size());
  g.type = type;
  geomData.push_back(g);
  const uint32_t n_offset = uint32_t(sbsNodes.size());
  const uint32_t v_offset = uint32_t(sbsData.size());
  sbsRoots.push_back(n_offset);
  sbsHeaders.push_back(header);
  sbsNodes.insert(sbsNodes.end(), nodes, nodes + n_nodes);
  sbsData.insert(sbsData.end(), values, values + n_values);
  for (size_t i = n_offset; i < sbsNodes.size(); i++) sbsNodes[i].data_offset += v_offset;
  if (indexed) {
    const uint32_t f_offset = uint32_t(sbsDataF.size());
    sbsDataF.insert(sbsDataF.end(), values_f, values_f + n_values_f);
    for (size_t i = v_offset; i < sbsData.size(); i++) sbsData[i] += f_offset;
    sbsFSlice.push_back({f_offset, uint32_t(n_values_f)});
  } else {
    sbsFSlice.push_back({0u, 0u});
  }
  std::vector<BVHNode> boxes(n_nodes);
  for (size_t i = 0; i < n_nodes; i++) {
    const float px = float(nodes[i].pos_xy >> 16);
    const float py = float(nodes[i].pos_xy & 0x0000FFFF);
    const float pz = float(nodes[i].pos_z_lod_size >> 16);
    const float sz = float(nodes[i].pos_z_lod_size & 0x0000FFFF);
    boxes[i].boxMin = make3(-1, -1, -1) + (2.0f * make3(px, py, pz)) / sz;
    boxes[i].boxMax = boxes[i].boxMin + (2.0f * make3(1, 1, 1)) / sz;
    boxes[i].leftOffset = 0;
    boxes[i].escapeIndex = 0;
  }
  if (boxes.size() == 1) {  // one-brick edge case (BVH2Common_host.cpp:684-692)
    boxes.resize(2);
    boxes[1].boxMin = boxes[0].boxMin - 0.001f * make3(1, 1, 1);
    boxes[1].boxMax = boxes[0].boxMin;
  }
  return add_aabb(TAG_SDF_BRICK, boxes.data(), boxes.size());
}

// BVHRT::AddInstance (BVH2Common_host.cpp:1264-1301)
uint32_t Scene::add_instance(uint32_t geom_id, const float4x4& m) {
  if (geom_id & 0x80000000u) geom_id &= 0x7fffffffu;
  if (geom_id >= geomData.size()) throw std::invalid_argument("invalid geometry id");
  const float4 bmin = geomData[geom_id].boxMin, bmax = geomData[geom_id].boxMax;
  const float4 corners[8] = {
      make4(bmin.x, bmin.y, bmin.z, 1.0f), make4(bmax.x, bmin.y, bmin.z, 1.0f), make4(bmin.x, bmax.y, bmin.z, 1.0f),
      make4(bmin.x, bmin.y, bmax.z, 1.0f), make4(bmax.x, bmax.y, bmin.z, 1.0f), make4(bmax.x, bmin.y, bmax.z, 1.0f),
      make4(bmin.x, bmax.y, bmax.z, 1.0f), make4(bmax.x, bmax.y, bmax.z, 1.0f)};
  Box4 box = empty_box4();
  for (int i = 0; i < 8; i++) include4(box, mat_mul4(m.m, corners[i]));
  InstanceData inst{};
  inst.boxMin = box.mn;
  inst.boxMax = box.mx;
  inst.geomId = geom_id;
  inst.transform = m;
  inst.transformInv = inverse4x4(m);
  inst.transformInvTransposed = transpose4x4(inverse4x4(m));
  instances.push_back(inst);
  inst_set_dirty = true;
  invalidate();
  return uint32_t(instances.size() - 1);
}

// BVHRT::UpdateInstance (BVH2Common_host.cpp:1303-1335). As in the reference, the
// normal matrix (transformInvTransposed) is left as it was.
void Scene::update_instance(uint32_t inst_id, const float4x4& m) {
  if (inst_id >= instances.size()) throw std::invalid_argument("bad instance id");
  const uint32_t geom_id = instances[inst_id].geomId;
  const float4 bmin = geomData[geom_id].boxMin, bmax = geomData[geom_id].boxMax;
  const float4 corners[8] = {
      make4(bmin.x, bmin.y, bmin.z, 1.0f), make4(bmax.x, bmin.y, bmin.z, 1.0f), make4(bmin.x, bmax.y, bmin.z, 1.0f),
      make4(bmin.x, bmin.y, bmax.z, 1.0f), make4(bmax.x, bmax.y, bmin.z, 1.0f), make4(bmax.x, bmin.y, bmax.z, 1.0f),
      make4(bmin.x, bmax.y, bmax.z, 1.0f), make4(bmax.x, bmax.y, bmax.z, 1.0f)};
  Box4 box = empty_box4();
  for (int i = 0; i < 8; i++) include4(box, mat_mul4(m.m, corners[i]));
  instances[inst_id].boxMin = box.mn;
  instances[inst_id].boxMax = box.mx;
  instances[inst_id].transform = m;
  instances[inst_id].transformInv = inverse4x4(m);
  invalidate();
}

// BVHRT::SetPreset (BVH2Common_host.cpp:1437-1440)
void Scene::set_preset(const MultiRenderPreset& p) {
  preset = p;
  lod_thr = std::pow(2.0f, preset.level_of_detail) - 0.01f;
}

uint32_t Scene::leaf_kinds() const {
  uint32_t k = 0;
  for (size_t g = 0; g < geomTags.size(); ++g) {
    const uint32_t t = geomTags[g];
    if (t == TAG_TRIANGLE) k |= LK_TRI;
    else if (t == TAG_SDF_NODE) k |= geomData[g].type == TYPE_SDF_FRAME_OCTREE_TEX ? (LK_FRAME | LK_FRAME_TEX) : LK_FRAME;
    else if (t == TAG_SDF_BRICK) k |= LK_SBS;
    else if (t == TAG_SDF_OCTREE) k |= LK_MULTI;
    else if (t == TAG_SDF_GRID) k |= LK_GRID;
    else if (t == TAG_GRAPHICS_PRIM) k |= LK_GPRIM;
    else if (t == TAG_COCTREE_BRICKED) k |= LK_COCT;
    else if (t == TAG_SDF_DAG) k |= LK_DAG;
    else if (t == TAG_SCOM2) k |= LK_SCOM2;
  }
  return k;
}

// Postponed triangle leaves help coherent rays into one mesh (fewer wave-wide waits) and hurt
// incoherent rays crossing many instances (more boxes pass the stale hit.t), hence the default.
bool Scene::postpone_triangles() const {
  if (leaf_postpone >= 0) return leaf_postpone != 0;
  return instances.size() == 1 && leaf_kinds() == LK_TRI;
}

// BVHRT::CommitScene (BVH2Common_host.cpp:1230-1262), then device upload.
namespace {
double tlas_area(const std::vector<BVHNode>& t) {
  double a = 0.0;
  for (const BVHNode& n : t) {
    const float3 d = n.boxMax - n.boxMin;
    a += 2.0 * (double(d.x) * d.y + double(d.x) * d.z + double(d.y) * d.z);
  }
  return a;
}
}  // namespace

void Scene::build_host() {
  if (instances.empty()) throw std::invalid_argument("commit with no instances");
  std::vector<float4> bmin(instances.size()), bmax(instances.size());
  for (size_t i = 0; i < instances.size(); ++i) { bmin[i] = instances[i].boxMin; bmax[i] = instances[i].boxMax; }
  tlas = build_tlas(bmin.data(), bmax.data(), instances.size());
  tlasParent.assign(tlas.size(), 0xFFFFFFFFu);
  for (uint32_t i = 0; i < uint32_t(tlas.size()); ++i) {
    const uint32_t lo = tlas[i].leftOffset;
    if (lo & LEAF_BIT) continue;
    tlasParent[lo] = i;
    tlasParent[lo + 1] = i;
  }
  tlasAreaBuilt = tlasAreaNow = tlas_area(tlas);
}

void Scene::update_cache_geom() {
  cacheGeom = 0xFFFFFFFFu;
  cachePairs = 0;
  if (instances.size() == 1) {
    cacheGeom = instances[0].geomId;
    const uint32_t b = geomData[cacheGeom].bvhOffset;
    const uint32_t e = cacheGeom + 1 < geomData.size() ? geomData[cacheGeom + 1].bvhOffset : uint32_t(nodePairs.size());
    cachePairs = std::min<uint32_t>(LDS_TOP_PAIRS, e - b);
  }
}

// Commit after UpdateInstance calls only: the instance array (and the resolve's inverse
// transposes) are uploaded, the BLAS arrays stay; the TLAS keeps its topology and its boxes are
// refitted on the GPU, then copied back so the host arrays describe what the kernels walk. A
// refit whose summed node area grows past twice that of the last full build, or tlas_update 0,
// rebuilds the TLAS on the host instead (CommitScene, BVH2Common_host.cpp:1230-1251).
hipError_t Scene::commit_instances() {
  invalidate();
  hipError_t e = dev["instances"].upload(instances.data(), instances.size() * sizeof(InstanceData));
  if (e != hipSuccess) return e;
  std::vector<float4x4> inst_inv_t(instances.size());
  for (size_t i = 0; i < instances.size(); ++i) inst_inv_t[i] = transpose4x4(instances[i].transformInv);
  e = dev["inst_inv_t"].upload(inst_inv_t.data(), inst_inv_t.size() * sizeof(float4x4));
  if (e != hipSuccess) return e;
  bool rebuild = tlas_update == 0;
  if (!rebuild) {
    DevBuf& cnt = dev["tlas_refit_cnt"];
    e = cnt.reserve(tlas.size() * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    e = launch_tlas_refit(static_cast<BVHNode*>(dev["tlas"].ptr()), uint32_t(tlas.size()),
                          static_cast<const uint32_t*>(dev["tlas_parent"].ptr()),
                          static_cast<const InstanceData*>(dev["instances"].ptr()), static_cast<uint32_t*>(cnt.ptr()),
                          nullptr);
    if (e != hipSuccess) return e;
    e = hipMemcpy(tlas.data(), dev["tlas"].ptr(), tlas.size() * sizeof(BVHNode), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    tlasAreaNow = tlas_area(tlas);
    rebuild = tlasAreaNow > 2.0 * tlasAreaBuilt;
    lastCommitKind = 1;
  }
  if (rebuild) {
    build_host();
    e = dev["tlas"].upload(tlas.data(), tlas.size() * sizeof(BVHNode));
    if (e != hipSuccess) return e;
    e = dev["tlas_parent"].upload(tlasParent.data(), tlasParent.size() * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    lastCommitKind = 2;
  }
  update_cache_geom();
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  committed = true;
  return hipSuccess;
}

// Derives the arrays the kernels read from the reference-layout host arrays. The traversal
// visits the same nodes in the same order and tests the same floats, so hits are unchanged:
//  * each geometry's pairs are renumbered: its top BFS levels first (they are staged in LDS
//    for single-instance scenes), then the remaining pairs in their depth-first order;
//  * triangle leaves hold PackOffsetAndSize(start, count) of their reordered triangles, taken
//    from m_primIdCount, so a leaf costs no indirection;
//  * triangles are stored as a vertex soup in leaf order (one 48-byte record per triangle,
//    vertex indices kept in .w for the normal of the final hit);
//  * frame-octree leaf boxes carry their 8 corner distances in one 64-byte record.
void Scene::build_device_layout() {
  devNodes = nodePairs;
  // pair ranges per geometry (geometries append their pairs in order)
  std::vector<std::pair<uint32_t, uint32_t>> ranges(geomData.size());
  for (size_t g = 0; g < geomData.size(); ++g) {
    const uint32_t b = geomData[g].bvhOffset;
    const uint32_t e = g + 1 < geomData.size() ? geomData[g + 1].bvhOffset : uint32_t(nodePairs.size());
    ranges[g] = {b, e};
  }
  for (size_t g = 0; g < geomData.size(); ++g) {
    const uint32_t b = ranges[g].first, c = ranges[g].second - ranges[g].first;
    if (c == 0) continue;
    // new order: BFS over the top levels (up to LDS_TOP_PAIRS pairs), then depth-first order
    std::vector<uint32_t> order;
    std::vector<uint8_t> taken(c, 0);
    std::vector<uint32_t> frontier{0};
    taken[0] = 1;
    while (!frontier.empty() && order.size() < size_t(LDS_TOP_PAIRS)) {
      std::vector<uint32_t> next;
      for (uint32_t p : frontier) {
        if (order.size() >= size_t(LDS_TOP_PAIRS)) break;
        order.push_back(p);
        for (const BVHNode* ch : {&nodePairs[b + p].left, &nodePairs[b + p].right})
          if (!(ch->leftOffset & LEAF_BIT) && ch->leftOffset < c) next.push_back(ch->leftOffset);
      }
      frontier.swap(next);
    }
    for (uint32_t p : order) taken[p] = 1;
    for (uint32_t p = 0; p < c; ++p)
      if (!taken[p]) order.push_back(p);
    std::vector<uint32_t> newIdx(c);
    for (uint32_t k = 0; k < c; ++k) newIdx[order[k]] = k;
    const bool tri = geomTags[g] == TAG_TRIANGLE;
    for (uint32_t k = 0; k < c; ++k) {
      BVHNodePair pr = nodePairs[b + order[k]];
      for (BVHNode* ch : {&pr.left, &pr.right}) {
        if (!(ch->leftOffset & LEAF_BIT)) {
          ch->leftOffset = newIdx[ch->leftOffset];
        } else if (tri && ch->leftOffset != LEAF_NORMAL) {
          const uint32_t w = primIdCount[startEnd[g].x + extract_start(ch->leftOffset)];
          ch->leftOffset = pack_offset_and_size(extract_start(w), extract_count(w));
        }
      }
      devNodes[b + k] = pr;
    }
  }
  // triangle soup in leaf order
  triSoup.assign(indices.size(), float4{0, 0, 0, 0});
  for (size_t g = 0; g < geomData.size(); ++g) {
    if (geomTags[g] != TAG_TRIANGLE) continue;
    const uint32_t ib = geomData[g].offset.x, vb = geomData[g].offset.y;
    size_t next_ib = indices.size();
    for (size_t h = 0; h < geomData.size(); ++h)
      if (geomTags[h] == TAG_TRIANGLE && geomData[h].offset.x > ib) next_ib = std::min<size_t>(next_ib, geomData[h].offset.x);
    for (size_t k = ib; k < next_ib; ++k) {
      const uint32_t vi = indices[k];
      const float4 p = vertPos[vb + vi];
      float4 r;
      r.x = p.x; r.y = p.y; r.z = p.z;
      std::memcpy(&r.w, &vi, 4);
      triSoup[k] = r;
    }
  }
  // frame-octree leaf records (textured octrees: values from their nodes, and the texture
  // coordinates of each leaf node gathered into frameTex)
  frameLeaves.assign(origNodes.size(), FrameLeaf{});
  frameTex.clear();
  for (size_t g = 0; g < geomData.size(); ++g) {
    const bool tex = geomData[g].type == TYPE_SDF_FRAME_OCTREE_TEX;
    if (geomTags[g] != TAG_SDF_NODE || (geomData[g].type != TYPE_SDF_FRAME_OCTREE && !tex)) continue;
    const uint32_t root = (tex ? frameTexRoots : frameRoots)[geomData[g].offset.x], ob = geomData[g].offset.y;
    const uint32_t nb = startEnd[g].y - startEnd[g].x;
    for (uint32_t k = 0; k < nb; ++k) {
      const BVHNode& on = origNodes[ob + k];
      FrameLeaf& fl = frameLeaves[ob + k];
      fl.boxMin = on.boxMin;
      fl.boxMax = on.boxMax;
      const float* vals = tex ? frameTexNodes[on.leftOffset + root].values : frameNodes[on.leftOffset + root].values;
      for (int j = 0; j < 8; ++j) fl.values[j] = vals[j];
      fl.primId = on.leftOffset;
      if (tex) {
        fl.tex = uint32_t(frameTex.size() / 4);
        const float* tc = frameTexNodes[on.leftOffset + root].tex_coords;
        for (int j = 0; j < 4; ++j) frameTex.push_back(make4(tc[4 * j], tc[4 * j + 1], tc[4 * j + 2], tc[4 * j + 3]));
      }
    }
  }
  // sparse-voxel-set leaf records: the voxel's box and its 8 byte-quantised values decoded as
  // BVHRT::RayNodeIntersection does (BVH2Common.cpp:190-215); primId = the voxel's index
  for (size_t g = 0; g < geomData.size(); ++g) {
    if (geomTags[g] != TAG_SDF_NODE || geomData[g].type != TYPE_SDF_SVS) continue;
    const uint32_t root = svsRoots[geomData[g].offset.x], ob = geomData[g].offset.y;
    const uint32_t nb = startEnd[g].y - startEnd[g].x;
    for (uint32_t k = 0; k < nb; ++k) {
      const SdfSVSNode& nd = svsNodes[root + k];
      FrameLeaf& fl = frameLeaves[ob + k];
      fl.boxMin = origNodes[ob + k].boxMin;
      fl.boxMax = origNodes[ob + k].boxMax;
      const float sz = float(nd.pos_z_lod_size & 0x0000FFFF);
      const float d_max = 2 * 1.73205081f / sz;
      for (int j = 0; j < 8; ++j)
        fl.values[j] = -d_max + 2.0f * d_max * (1.0f / 255.0f) * float((nd.values[j / 4] >> (8 * (j % 4))) & 0xFF);
      fl.primId = k;
    }
  }
  update_cache_geom();
}

hipError_t Scene::mask_for(StreamWork& w, void* stream, uint32_t** mask) {
  const size_t need = std::max<size_t>(sbsMaskWords, 1) * sizeof(uint32_t);
  if (w.mask.bytes() < need) {
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return e;
    e = w.mask.reserve(need);
    if (e != hipSuccess) return e;
  }
  *mask = static_cast<uint32_t*>(w.mask.ptr());
  return hipSuccess;
}

hipError_t Scene::commit() {
  if (!geom_dirty && !inst_set_dirty && !tlas.empty() && dev["tlas"].bytes() == tlas.size() * sizeof(BVHNode))
    return commit_instances();
  invalidate();   // buffers below may move
  build_host();
  build_device_layout();
  // occupancy-mask layout: per brick 1 summary word + ceil(bs^3 / 32) bit words
  sbsMaskGeoms.clear();
  sbsMaskBase.assign(sbsRoots.size(), 0u);
  sbsMaskWords = 0;
  for (size_t g = 0; g < sbsRoots.size(); ++g) {
    const uint32_t first = sbsRoots[g];
    const uint32_t last = g + 1 < sbsRoots.size() ? sbsRoots[g + 1] : uint32_t(sbsNodes.size());
    const uint32_t bs = sbsHeaders[g].brick_size;
    sbsMaskBase[g] = uint32_t(sbsMaskWords);
    sbsMaskGeoms.push_back({uint32_t(g), last - first, bs, uint32_t(sbsMaskWords)});
    sbsMaskWords += size_t(last - first) * (1 + (size_t(bs) * bs * bs + 31) / 32);
  }
  if (sbsMaskWords > 0xFFFFFFFFull) return hipErrorInvalidValue;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return e;
#define UP(name, vec)                                                        \
  e = dev[name].upload(vec.data(), vec.size() * sizeof(vec[0]));             \
  if (e != hipSuccess) return e;
  UP("nodes", devNodes) UP("tri_soup", triSoup) UP("frame_leaves", frameLeaves) UP("frame_tex", frameTex) UP("tlas", tlas) UP("instances", instances) UP("geoms", geomData) UP("geom_tags", geomTags)
  UP("start_end", startEnd) UP("prim_id_count", primIdCount) UP("vert_pos", vertPos) UP("vert_norm", vertNorm)
  UP("indices", indices) UP("prim_indices", primIndices) UP("orig_nodes", origNodes) UP("frame_nodes", frameNodes)
  UP("frame_roots", frameRoots) UP("sbs_nodes", sbsNodes) UP("sbs_data", sbsData) UP("sbs_data_f", sbsDataF)
  UP("sbs_roots", sbsRoots) UP("sbs_headers", sbsHeaders) UP("multi_nodes", multiNodes) UP("multi_values", multiValues)
  UP("sbs_mask_base", sbsMaskBase) UP("orig_indices", origIndices) UP("tlas_parent", tlasParent)
  UP("grid_data", gridData) UP("grid_offsets", gridOffsets) UP("grid_sizes", gridSizes)
  UP("gprim_headers", gprimHeaders) UP("gprim_points", gprimPoints)
  UP("coct_data", coctData) UP("coct_geoms", coctGeoms) UP("coct_mods", coctMods)
  UP("dag_nodes", dagNodes) UP("dag_child_edges", dagChildEdges) UP("dag_data_edges", dagDataEdges)
  UP("dag_distances", dagDistances) UP("dag_geoms", dagGeoms) UP("dag_mods", dagMods)
  UP("scom2_data", scom2Data) UP("scom2_headers", scom2Headers) UP("scom2_mods", scom2Mods)
  if (!dagGeoms.empty() || !scom2Headers.empty()) { UP("rot_add", rot_add_table()) }
  // MultiRenderer's own m_instanceTransformInvTransposed (MultiRenderer_host.cpp:870-878):
  // transpose(inverse4x4(matrix)), refreshed by UpdateInstance (unlike the BVHRT copy)
  std::vector<float4x4> inst_inv_t(instances.size());
  for (size_t i = 0; i < instances.size(); ++i) inst_inv_t[i] = transpose4x4(instances[i].transformInv);
  UP("inst_inv_t", inst_inv_t)
#undef UP
  e = dev["sbs_mask"].reserve(sbsMaskWords * sizeof(uint32_t));
  if (e != hipSuccess) return e;
  if (sbsMaskWords) {
    SceneView v = make_view();
    v.sbs_mask = nullptr;   // built from the committed values, without a mask
    e = launch_sbs_masks(v, sbsMaskGeoms.data(), uint32_t(sbsMaskGeoms.size()), static_cast<uint32_t*>(dev["sbs_mask"].ptr()),
                         nullptr);
    if (e != hipSuccess) return e;
  }
  if (!d_overflow) {
    e = hipMalloc(&d_overflow, 2 * sizeof(uint32_t));   // [0] stack overflows, [1] adjoint recompute misses
    if (e != hipSuccess) return e;
    e = hipMemset(d_overflow, 0, 2 * sizeof(uint32_t));
    if (e != hipSuccess) return e;
  }
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  geom_dirty = inst_set_dirty = false;
  lastCommitKind = 0;
  committed = true;
  return hipSuccess;
}

SceneView Scene::make_view() const {
  SceneView v{};
  auto P = [&](const char* n) -> const void* {
    auto it = dev.find(n);
    return (it == dev.end() || it->second.bytes() == 0) ? nullptr : it->second.ptr();
  };
  v.nodes = (const BVHNodePair*)P("nodes");
  v.tlas = (const BVHNode*)P("tlas");
  v.tri_soup = (const float4*)P("tri_soup");
  v.frame_leaves = (const FrameLeaf*)P("frame_leaves");
  v.frame_tex = (const float4*)P("frame_tex");
  v.instances = (const InstanceData*)P("instances");
  v.geoms = (const GeomData*)P("geoms");
  v.geom_tags = (const uint32_t*)P("geom_tags");
  v.start_end = (const uint2*)P("start_end");
  v.prim_id_count = (const uint32_t*)P("prim_id_count");
  v.vert_pos = (const float4*)P("vert_pos");
  v.vert_norm = (const float4*)P("vert_norm");
  v.indices = (const uint32_t*)P("indices");
  v.prim_indices = (const uint32_t*)P("prim_indices");
  v.orig_nodes = (const BVHNode*)P("orig_nodes");
  v.frame_nodes = (const SdfFrameOctreeNode*)P("frame_nodes");
  v.frame_roots = (const uint32_t*)P("frame_roots");
  v.sbs_nodes = (const SdfSBSNode*)P("sbs_nodes");
  v.sbs_data = (const uint32_t*)P("sbs_data");
  v.sbs_data_f = (const float*)P("sbs_data_f");
  v.sbs_roots = (const uint32_t*)P("sbs_roots");
  v.sbs_headers = (const SdfSBSHeader*)P("sbs_headers");
  v.sbs_mask = sbsMaskWords ? (const uint32_t*)P("sbs_mask") : nullptr;
  v.sbs_mask_base = (const uint32_t*)P("sbs_mask_base");
  v.multi_nodes = (const SdfMultiOctreeNode*)P("multi_nodes");
  v.multi_values = (const float*)P("multi_values");
  v.grid_data = (const float*)P("grid_data");
  v.grid_offsets = (const uint32_t*)P("grid_offsets");
  v.grid_sizes = (const uint32_t*)P("grid_sizes");
  v.gprim_headers = (const GraphicsPrimHeader*)P("gprim_headers");
  v.gprim_points = (const float4*)P("gprim_points");
  v.coct_data = (const uint32_t*)P("coct_data");
  v.coct_geoms = (const CoctGeom*)P("coct_geoms");
  v.coct_mods = (const int4*)P("coct_mods");
  v.dag_nodes = (const SdfDAGNode*)P("dag_nodes");
  v.dag_child_edges = (const SdfDAGChildEdge*)P("dag_child_edges");
  v.dag_data_edges = (const SdfDAGDataEdge*)P("dag_data_edges");
  v.dag_distances = (const float*)P("dag_distances");
  v.dag_geoms = (const DagGeom*)P("dag_geoms");
  v.dag_mods = (const int4*)P("dag_mods");
  v.rot_add = (const uint8_t*)P("rot_add");
  v.scom2_data = (const uint32_t*)P("scom2_data");
  v.scom2_headers = (const SCom2Header*)P("scom2_headers");
  v.scom2_mods = (const int4*)P("scom2_mods");
  // BVHRT::SetPreset (BVH2Common_host.cpp:1436-1451): one plain SCom2 tree rendered LAMBERT_NO_TEX
  v.scom2_fast = (scom2Headers.size() == 1 && preset.render_mode == MULTI_RENDER_MODE_LAMBERT_NO_TEX &&
                  scom2_plain_tree(scom2Headers[0])) ? 1u : 0u;
  v.preset = preset;
  v.lod_thr = lod_thr;
  v.leaf_kinds = leaf_kinds();
  // smoothed brick-set normals (BVH2Common.cpp:958-961): neighbour-index layout, trilinear,
  // normal mode SDF_SMOOTHED and a render mode that needs normals
  if ((v.leaf_kinds & LK_SBS) && preset.normal_mode == NORMAL_MODE_SDF_SMOOTHED &&
      preset.interpolation_mode == INTERPOLATION_MODE_TRILINEAR && preset_needs_normal(preset.render_mode)) {
    for (const SdfSBSHeader& h : sbsHeaders)
      if (h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN) v.leaf_kinds |= LK_SBS_SMOOTH;
  }
  // tricubic interpolation (BVH2Common.cpp:933-955, 1478-1499) is read by brick sets and grids only
  if ((v.leaf_kinds & (LK_SBS | LK_GRID)) && preset.interpolation_mode == INTERPOLATION_MODE_TRICUBIC)
    v.leaf_kinds |= LK_TRICUBIC;
  v.n_instances = uint32_t(instances.size());
  v.cache_geom = cacheGeom;
  v.cache_pairs = cachePairs;
  return v;
}

}  // namespace lrt
