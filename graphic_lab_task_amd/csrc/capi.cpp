// extern "C" entry points declared in include/lrt.h.
#include "../../include/lrt.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <exception>
#include <stdexcept>
#include <string>

#include "host_math.h"
#include "kernels.h"
#include "scene.h"
#include "tune.h"

struct lrt_scene {
  lrt::Scene s;
};

struct lrt_sbs_batch {
  int device = -1;
  std::vector<lrt::SbsBatchItem> items;
  // per item: the scene's commit generation when the plan was made (a re-commit, clear or
  // destruction of the scene invalidates the device pointers the plan holds)
  std::vector<std::weak_ptr<uint64_t>> scene_gen;
  std::vector<uint64_t> gen_at_create;
  lrt::DevBuf d_items, d_begins, d_masks, d_spill, d_counter;
  lrt::SbsBatchLaunch L{};
  uint32_t resident[6] = {};   // resident blocks per variant: (LDS stack 8 << v/2, record v%2)
};

namespace {

thread_local std::string g_err;

lrt_status fail(lrt_status code, const std::string& msg) {
  g_err = msg;
  return code;
}
lrt_status hip_fail(hipError_t e, const char* where) {
  return fail(LRT_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

template <class F>
lrt_status guarded(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::invalid_argument& e) {
    return fail(LRT_ERR_INVALID_ARG, e.what());
  } catch (const std::exception& e) {
    return fail(LRT_ERR_HIP, e.what());
  } catch (...) {
    return fail(LRT_ERR_HIP, "unknown error");
  }
}

lrt::float4x4 to_mat(const float* m16) {
  lrt::float4x4 m;
  std::memcpy(m.m, m16, sizeof(m.m));
  return m;
}

lrt_status check_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(LRT_ERR_NO_DEVICE, "no HIP device available");
  return LRT_OK;
}

// Launch arguments of a trace call from the scene's tuned configuration (queue and spill unset).
lrt::TraceArgs trace_args(lrt_scene* s, const float* pos, const float* dir, lrt_hit* hits, uint32_t* any, uint64_t n,
                          uint32_t* counts) {
  lrt::TraceArgs a{};
  a.pos = reinterpret_cast<const lrt::float4*>(pos);
  a.dir = reinterpret_cast<const lrt::float4*>(dir);
  a.out = reinterpret_cast<lrt::CRT_Hit*>(hits);
  a.any_out = any;
  a.n = uint32_t(n);
  a.overflow_count = s->s.d_overflow;
  a.block = s->s.block;
  a.lds_stack = s->s.lds_stack;
  a.xcd_remap = uint32_t(s->s.xcd_remap);
  a.counts = counts;
  // The cooperative pass reproduces every hit field except the per-leaf call count that the
  // INTERSECTION_COUNT debug mode accumulates in instId; that mode runs single-pass.
  a.budget = s->s.preset.render_mode == lrt::MULTI_RENDER_MODE_INTERSECTION_COUNT ? 0u : s->s.budget;
  a.tail_waves = s->s.tail_waves;
  a.ray_groups = s->s.ray_groups;
  // extra leaf visits would change the counts of that mode too
  a.postpone_tri = s->s.preset.render_mode != lrt::MULTI_RENDER_MODE_INTERSECTION_COUNT && s->s.postpone_triangles();
  a.persistent = s->s.persistent ? 1u : 0u;
  return a;
}

// Queries this library does not answer: tricubic interpolation of a brick set without padding
// (the 4^3 samples start one voxel before the brick, BVH2Common.cpp:790-844: the reference reads
// outside the brick there). Returns the reason, or null.
const char* unsupported_query(const lrt::Scene& sc) {
  if (sc.preset.interpolation_mode == lrt::INTERPOLATION_MODE_TRICUBIC)
    for (const lrt::SdfSBSHeader& h : sc.sbsHeaders)
      if (h.brick_pad < 1) return "tricubic interpolation needs brick sets with brick_pad >= 1";
  return nullptr;
}

lrt_status trace_common(lrt_scene* s, const float* pos, const float* dir, lrt_hit* hits, uint32_t* any, uint64_t n,
                        void* stream, uint32_t* counts = nullptr) {
  if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
  if (!s->s.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
  if (const char* why = unsupported_query(s->s)) return fail(LRT_ERR_UNSUPPORTED, why);
  if (n == 0) return LRT_OK;
  if (!pos || !dir || (!hits && !any)) return fail(LRT_ERR_INVALID_ARG, "null ray or hit buffer");
  if (n > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many rays in one call");
  lrt::TraceArgs a = trace_args(s, pos, dir, hits, any, n, counts);
  // held until both passes are enqueued: another thread on this stream must not slip its passes
  // between ours (they share the stream's tail queue) or grow the buffer under us
  lrt::Scene::StreamWork& w = s->s.stream_work(stream);
  std::lock_guard<std::mutex> lock(w.mu);
  hipError_t e = s->s.work_for(w, stream, size_t(n), &a.spill, &a.queue, &a.pcount);
  if (e != hipSuccess) return hip_fail(e, "work buffer");
  e = lrt::launch_trace(s->s.make_view(), a, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "trace launch");
  return LRT_OK;
}

// Resolves the differentiable brick-set geometry: indexed layout whose values_f slice holds
// n_dist distances then n_dist RGB albedos (4 n_dist floats).
lrt_status diff_geometry(lrt_scene* s, uint32_t geom_id, uint32_t* f_offset, uint32_t* n_dist) {
  if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
  lrt::Scene& sc = s->s;
  if (geom_id >= sc.geomData.size() || sc.geomTags[geom_id] != lrt::TAG_SDF_BRICK)
    return fail(LRT_ERR_INVALID_ARG, "geometry is not a sparse brick set");
  // the differentiable kernels walk brick-set leaves only: another leaf kind in the scene would be
  // skipped silently (an occluding mesh rendered invisible, its shadowed gradients wrong)
  if (sc.leaf_kinds() != lrt::LK_SBS)
    return fail(LRT_ERR_UNSUPPORTED, "differentiable calls need a scene whose geometries are all sparse brick sets");
  if (sc.preset.interpolation_mode != lrt::INTERPOLATION_MODE_TRILINEAR)
    return fail(LRT_ERR_UNSUPPORTED, "differentiable calls interpolate trilinearly (interpolation_mode 0)");
  const uint32_t sdf = sc.geomData[geom_id].offset.x;
  const lrt::uint2 sl = sc.sbsFSlice[sdf];
  if (sl.y == 0 || sl.y % 4 != 0)
    return fail(LRT_ERR_INVALID_ARG, "brick set needs the indexed fp32 layout with 4*n_dist values_f entries "
                                     "(distances, then RGB albedo per lattice point)");
  // every distance index of the geometry's bricks must address the distance part of its slice
  // (checked once per geometry; the kernels derive albedo addresses from these indices)
  if (sc.sbsDiffChecked.size() <= sdf) sc.sbsDiffChecked.resize(sdf + 1, 0);
  if (sc.sbsDiffChecked[sdf] == 0) {
    const lrt::SdfSBSHeader& h = sc.sbsHeaders[sdf];
    const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1, v3 = v_size * v_size * v_size;
    const size_t first = sc.sbsRoots[sdf];
    const size_t last = sdf + 1 < sc.sbsRoots.size() ? sc.sbsRoots[sdf + 1] : sc.sbsNodes.size();
    bool ok = true;
    for (size_t b = first; b < last && ok; ++b)
      for (uint32_t k = 0; k < v3 && ok; ++k) {
        const uint32_t idx = sc.sbsData[sc.sbsNodes[b].data_offset + k];
        ok = idx >= sl.x && idx < sl.x + sl.y / 4;
      }
    sc.sbsDiffChecked[sdf] = ok ? 1 : 2;
  }
  if (sc.sbsDiffChecked[sdf] != 1)
    return fail(LRT_ERR_INVALID_ARG, "brick distance indices must address the first n_dist values_f entries");
  if (f_offset) *f_offset = sl.x;
  if (n_dist) *n_dist = sl.y / 4;
  return LRT_OK;
}

lrt_status sbs_diff_common(lrt_scene* s, uint32_t geom_id, const float* params, const float* pos, const float* dir,
                           uint64_t n, float* colour, const float* dL_dcolour, const float* target, float* grad,
                           float* loss, void* stream) {
  uint32_t f_offset = 0, n_dist = 0;
  lrt_status st = diff_geometry(s, geom_id, &f_offset, &n_dist);
  if (st != LRT_OK) return st;
  if (!s->s.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
  if (n == 0) return LRT_OK;
  if (!params || !pos || !dir) return fail(LRT_ERR_INVALID_ARG, "null parameter or ray buffer");
  if (!colour && !grad) return fail(LRT_ERR_INVALID_ARG, "null output buffer");
  if (n > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many rays in one call");
  lrt::SbsDiffArgs a{};
  a.pos = reinterpret_cast<const lrt::float4*>(pos);
  a.dir = reinterpret_cast<const lrt::float4*>(dir);
  a.n = uint32_t(n);
  a.params = params;
  a.geom_id = geom_id;
  a.f_offset = f_offset;
  a.n_dist = n_dist;
  a.colour = reinterpret_cast<lrt::float4*>(colour);
  a.dL_dcolour = reinterpret_cast<const lrt::float4*>(dL_dcolour);
  a.target = reinterpret_cast<const lrt::float4*>(target);
  a.grad = grad;
  a.loss = loss;
  a.counters = s->s.d_overflow ? s->s.d_overflow + 1 : nullptr;
  a.xcd_remap = uint32_t(s->s.xcd_remap);
  lrt::TailQueue* q = nullptr;
  lrt::Scene::StreamWork& w = s->s.stream_work(stream);
  std::lock_guard<std::mutex> lock(w.mu);
  uint32_t* pcount = nullptr;
  hipError_t e = s->s.work_for(w, stream, size_t(n), &a.spill, &q, &pcount);
  if (e != hipSuccess) return hip_fail(e, "work buffer");
  // occupancy masks from the caller's parameters (they differ from the committed values)
  lrt::SceneView v = s->s.make_view();
  v.sbs_data_f = params;
  v.sbs_mask = nullptr;
  uint32_t* mask = nullptr;
  e = s->s.mask_for(w, stream, &mask);
  if (e != hipSuccess) return hip_fail(e, "mask buffer");
  e = lrt::launch_sbs_masks(v, s->s.sbsMaskGeoms.data(), uint32_t(s->s.sbsMaskGeoms.size()), mask,
                            reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "occupancy mask launch");
  v.sbs_mask = mask;
  e = lrt::launch_sbs_diff(v, a, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "differentiable brick-set launch");
  return LRT_OK;
}

}  // namespace

namespace lrt {
// Error message of the calling thread for the entry points outside this file (lrt_build.h).
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace lrt

extern "C" {

const char* lrt_last_error(void) { return g_err.c_str(); }

const char* lrt_version(void) { return "lrt-gfx950 0.1"; }

lrt_status lrt_create(const char* impl_name, const char* build_name, const char* layout_name, lrt_scene** out) {
  return guarded([&]() -> lrt_status {
    if (!out) return fail(LRT_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    const std::string impl = impl_name ? impl_name : "";
    if (impl != "BVH2Common") return fail(LRT_ERR_UNSUPPORTED, "unknown scene implementation '" + impl + "'");
    auto* sc = new lrt_scene();
    sc->s.build_name = build_name ? build_name : "";
    sc->s.layout_name = layout_name ? layout_name : "";
    // BuilderPresetsFromString (utils/raytracing/cbvh.cpp:57-69): last digit = prims per leaf
    const std::string& b = sc->s.build_name;
    sc->s.prims_in_leaf = 2;
    if (!b.empty() && b.back() >= '0' && b.back() <= '9') sc->s.prims_in_leaf = b.back() - '0';
    if (sc->s.prims_in_leaf < 1) sc->s.prims_in_leaf = 1;
    if (sc->s.prims_in_leaf > 7) sc->s.prims_in_leaf = 7;   // 3-bit leaf count field
    sc->s.gpu_builder = b.compare(0, 3, "gpu") == 0;
    sc->s.set_preset(lrt::default_preset());
    *out = sc;
    return LRT_OK;
  });
}

lrt_status lrt_destroy(lrt_scene* s) {
  return guarded([&]() -> lrt_status {
    delete s;
    return LRT_OK;
  });
}

lrt_status lrt_clear_geom(lrt_scene* s) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    s->s.clear_geom();
    return LRT_OK;
  });
}

lrt_status lrt_add_triangles(lrt_scene* s, const float* vpos, size_t n_verts, size_t v_byte_stride, const float* vnorm,
                             const uint32_t* indices, size_t n_indices, uint32_t build_flags, uint32_t* out_geom_id) {
  (void)build_flags;
  return guarded([&]() -> lrt_status {
    if (!s || !vpos || !indices) return fail(LRT_ERR_INVALID_ARG, "null argument");
    if (v_byte_stride == 0) v_byte_stride = 12;
    const uint32_t id = s->s.add_triangles(vpos, n_verts, v_byte_stride, vnorm, indices, n_indices);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_frame_octree(lrt_scene* s, const void* nodes, size_t n_nodes, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !nodes) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_frame_octree(static_cast<const lrt::SdfFrameOctreeNode*>(nodes), n_nodes);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_frame_octree_tex(lrt_scene* s, const void* nodes, size_t n_nodes, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !nodes) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_frame_octree_tex(static_cast<const lrt::SdfFrameOctreeTexNode*>(nodes), n_nodes);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_graphics_prims(lrt_scene* s, const void* header, const float* points, size_t n_points,
                                  uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !header || !points) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_graphics_prims(*static_cast<const lrt::GraphicsPrimHeader*>(header),
                                                reinterpret_cast<const lrt::float4*>(points), n_points);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_grid(lrt_scene* s, const uint32_t size[3], const float* data, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !size || !data) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_grid(size, data);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_coctree_v3(lrt_scene* s, const uint32_t header[15], const uint32_t* data, size_t size,
                              uint32_t bvh_level, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !header || !data) return fail(LRT_ERR_INVALID_ARG, "null argument");
    lrt::COctreeV3Header h;
    std::memcpy(&h, header, sizeof(h));
    const uint32_t id = s->s.add_coctree_v3(h, data, size, bvh_level);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_dag(lrt_scene* s, const void* header, const void* nodes, size_t n_nodes,
                           const void* child_edges, size_t n_child_edges, const void* data_edges, size_t n_data_edges,
                           const float* distances, size_t n_distances, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !header || !nodes || !child_edges || !data_edges || !distances)
      return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_dag(*static_cast<const lrt::SdfDAGHeader*>(header),
                                         static_cast<const lrt::SdfDAGNode*>(nodes), n_nodes,
                                         static_cast<const lrt::SdfDAGChildEdge*>(child_edges), n_child_edges,
                                         static_cast<const lrt::SdfDAGDataEdge*>(data_edges), n_data_edges, distances,
                                         n_distances);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_scom2(lrt_scene* s, const void* header, const uint32_t* nodes, size_t n_nodes,
                         const uint32_t* bricks, size_t n_bricks, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !header || !nodes || !bricks) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id =
        s->s.add_scom2(*static_cast<const lrt::SCom2Header*>(header), nodes, n_nodes, bricks, n_bricks);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_svs(lrt_scene* s, const void* nodes, size_t n_nodes, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !nodes) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_svs(static_cast<const lrt::SdfSVSNode*>(nodes), n_nodes);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_sbs(lrt_scene* s, const void* header, const void* nodes, size_t n_nodes, const uint32_t* values,
                           size_t n_values, const float* values_f, size_t n_values_f, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !header || !nodes || !values) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_sbs(*static_cast<const lrt::SdfSBSHeader*>(header),
                                         static_cast<const lrt::SdfSBSNode*>(nodes), n_nodes, values, n_values,
                                         values_f, n_values_f);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_add_sdf_multi_octree(lrt_scene* s, const void* nodes, size_t n_nodes, const float* values,
                                    size_t n_values, uint32_t* out_geom_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !nodes || !values) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_sdf_multi_octree(static_cast<const lrt::SdfMultiOctreeNode*>(nodes), n_nodes, values,
                                                  n_values);
    if (out_geom_id) *out_geom_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_clear_scene(lrt_scene* s) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    s->s.clear_scene();
    return LRT_OK;
  });
}

lrt_status lrt_add_instance(lrt_scene* s, uint32_t geom_id, const float* matrix16, uint32_t* out_inst_id) {
  return guarded([&]() -> lrt_status {
    if (!s || !matrix16) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const uint32_t id = s->s.add_instance(geom_id, to_mat(matrix16));
    if (out_inst_id) *out_inst_id = id;
    return LRT_OK;
  });
}

lrt_status lrt_update_instance(lrt_scene* s, uint32_t inst_id, const float* matrix16) {
  return guarded([&]() -> lrt_status {
    if (!s || !matrix16) return fail(LRT_ERR_INVALID_ARG, "null argument");
    s->s.update_instance(inst_id, to_mat(matrix16));
    return LRT_OK;
  });
}

lrt_status lrt_commit(lrt_scene* s, uint32_t build_flags) {
  (void)build_flags;
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    lrt_status st = check_device();
    if (st != LRT_OK) {
      s->s.build_host();
      return fail(st, "host arrays built; no HIP device to upload them to");
    }
    hipError_t e = s->s.commit();
    if (e != hipSuccess) return hip_fail(e, "commit");
    return LRT_OK;
  });
}

lrt_status lrt_set_tlas_update(lrt_scene* s, int mode) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (mode != 0 && mode != 1) return fail(LRT_ERR_INVALID_ARG, "mode must be 0 (rebuild) or 1 (refit)");
    s->s.tlas_update = mode;
    return LRT_OK;
  });
}

lrt_status lrt_get_commit_info(lrt_scene* s, lrt_commit_info* out) {
  return guarded([&]() -> lrt_status {
    if (!s || !out) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const lrt::Scene& sc = s->s;
    out->kind = sc.lastCommitKind;
    out->gpu_builder = sc.gpu_builder ? 1u : 0u;
    out->blas_max_depth = sc.lastBuild.max_depth;
    out->blas_leaves = sc.lastBuild.leaves;
    out->blas_merged_leaves = sc.lastBuild.merged_leaves;
    out->tlas_nodes = uint32_t(sc.tlas.size());
    out->tlas_area_ratio = sc.tlasAreaBuilt > 0.0 ? float(sc.tlasAreaNow / sc.tlasAreaBuilt) : 1.0f;
    return LRT_OK;
  });
}

lrt_status lrt_set_preset(lrt_scene* s, const void* preset) {
  return guarded([&]() -> lrt_status {
    if (!s || !preset) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const lrt::MultiRenderPreset& p = *static_cast<const lrt::MultiRenderPreset*>(preset);
    if (p.interpolation_mode != lrt::INTERPOLATION_MODE_TRILINEAR && p.interpolation_mode != lrt::INTERPOLATION_MODE_TRICUBIC)
      return fail(LRT_ERR_UNSUPPORTED, "interpolation_mode must be 0 (trilinear) or 1 (tricubic)");
    s->s.set_preset(p);
    return LRT_OK;
  });
}

lrt_status lrt_trace_closest_dev(lrt_scene* s, const float* pos_near, const float* dir_far, lrt_hit* hits, uint64_t n,
                                 void* stream) {
  return guarded([&]() -> lrt_status { return trace_common(s, pos_near, dir_far, hits, nullptr, n, stream); });
}

lrt_status lrt_trace_closest_noalloc_dev(lrt_scene* s, const float* pos_near, const float* dir_far, lrt_hit* hits,
                                         uint64_t n, void* stream) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (!s->s.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
    if (n == 0) return LRT_OK;
    if (!pos_near || !dir_far || !hits) return fail(LRT_ERR_INVALID_ARG, "null ray or hit buffer");
    if (n > LRT_NOALLOC_MAX_RAYS)
      return fail(LRT_ERR_INVALID_ARG, "lrt_trace_closest_noalloc_dev takes at most " +
                                           std::to_string(LRT_NOALLOC_MAX_RAYS) + " rays per call (got " +
                                           std::to_string(n) + ")");
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return hip_fail(e, "current device");
    if (device != s->s.device) return fail(LRT_ERR_INVALID_ARG, "the current device is not the scene's device");
    if (const char* why = unsupported_query(s->s)) return fail(LRT_ERR_UNSUPPORTED, why);
    const lrt::TraceArgs a = trace_args(s, pos_near, dir_far, hits, nullptr, n, nullptr);
    e = lrt::launch_trace_noalloc(s->s.make_view(), a, device, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "allocation-free trace launch");
    return LRT_OK;
  });
}

lrt_status lrt_trace_closest_counted_dev(lrt_scene* s, const float* pos_near, const float* dir_far, lrt_hit* hits,
                                         uint32_t* counters, uint64_t n, void* stream) {
  return guarded([&]() -> lrt_status {
    if (!counters) return fail(LRT_ERR_INVALID_ARG, "null counter buffer");
    return trace_common(s, pos_near, dir_far, hits, nullptr, n, stream, counters);
  });
}

lrt_status lrt_trace_any_dev(lrt_scene* s, const float* pos_near, const float* dir_far, uint32_t* flags, uint64_t n,
                             void* stream) {
  return guarded([&]() -> lrt_status {
    if (!flags) return fail(LRT_ERR_INVALID_ARG, "null flag buffer");
    return trace_common(s, pos_near, dir_far, nullptr, flags, n, stream);
  });
}

lrt_status lrt_trace_closest_host(lrt_scene* s, const float* pos_near, const float* dir_far, lrt_hit* hits,
                                  uint64_t n) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (!s->s.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
    if (n == 0) return LRT_OK;
    if (!pos_near || !dir_far || !hits) return fail(LRT_ERR_INVALID_ARG, "null buffer");
    if (n > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many rays in one call");
    // one grow-only staging buffer per scene: rays (2 x 16 B) then hits (32 B) per ray
    std::lock_guard<std::mutex> lock(s->s.host_mu);
    lrt::DevBuf& io = s->s.host_io;
    hipError_t e;
    if (io.bytes() < n * 64) {
      if ((e = io.reserve(n * 64)) != hipSuccess) return hip_fail(e, "host staging buffer");
    }
    char* base = static_cast<char*>(io.ptr());
    float* dp = reinterpret_cast<float*>(base);
    float* dd = reinterpret_cast<float*>(base + n * 16);
    lrt_hit* dh = reinterpret_cast<lrt_hit*>(base + n * 32);
    if ((e = hipMemcpy(dp, pos_near, n * 16, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "copy rays");
    if ((e = hipMemcpy(dd, dir_far, n * 16, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "copy rays");
    lrt_status st = trace_common(s, dp, dd, dh, nullptr, n, nullptr);
    if (st != LRT_OK) return st;
    if ((e = hipMemcpy(hits, dh, n * sizeof(lrt_hit), hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_fail(e, "copy hits");
    return LRT_OK;
  });
}

lrt_status lrt_query_one(lrt_scene* s, const float* pos_near4, const float* dir_far4, lrt_hit* out) {
  return lrt_trace_closest_host(s, pos_near4, dir_far4, out, 1);
}

lrt_status lrt_resolve_dev(lrt_scene* s, const lrt_hit* gbuffer, const lrt_resolve_desc* desc, float* colour,
                           void* stream) {
  return guarded([&]() -> lrt_status {
    if (!s || !gbuffer || !desc || !colour || !desc->proj_inv16 || !desc->world_view_inv16)
      return fail(LRT_ERR_INVALID_ARG, "null argument");
    lrt::Scene& sc = s->s;
    if (!sc.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
    const uint32_t w = desc->width, h = desc->height;
    if (w == 0 || h == 0 || w > 65535 || h > 65535) return fail(LRT_ERR_INVALID_ARG, "bad viewport");
    if (desc->order > 1) return fail(LRT_ERR_INVALID_ARG, "order must be 0 or 1");
    if (desc->order == 1 && (w % 8 != 0 || h % 8 != 0))
      return fail(LRT_ERR_INVALID_ARG, "packed order needs width and height multiples of 8");
    if (desc->n_lights > lrt::RESOLVE_MAX_LIGHTS || (desc->n_lights && !desc->lights))
      return fail(LRT_ERR_INVALID_ARG, "at most 8 lights");
    const uint32_t m = sc.preset.render_mode;
    using namespace lrt;
    const bool debug = m == MULTI_RENDER_MODE_MASK || m == MULTI_RENDER_MODE_DEPTH ||
                       m == MULTI_RENDER_MODE_LINEAR_DEPTH || m == MULTI_RENDER_MODE_INVERSE_LINEAR_DEPTH ||
                       m == MULTI_RENDER_MODE_PRIMITIVE || m == MULTI_RENDER_MODE_TYPE ||
                       m == MULTI_RENDER_MODE_NORMAL || m == MULTI_RENDER_MODE_BARYCENTRIC ||
                       m == MULTI_RENDER_MODE_ST_ITERATIONS || m == MULTI_RENDER_MODE_TEX_COORDS ||
                       m == MULTI_RENDER_MODE_HSV_DEPTH || m == MULTI_RENDER_MODE_MO_VOXEL_COUNT ||
                       m == MULTI_RENDER_MODE_MO_FLAGS || m == MULTI_RENDER_MODE_INTERSECTION_COUNT;
    bool common = m == MULTI_RENDER_MODE_LAMBERT_NO_TEX || m == MULTI_RENDER_MODE_GEOM || m == MULTI_RENDER_MODE_LOD;
    if (m == MULTI_RENDER_MODE_LAMBERT) {
      common = !sc.geomData.empty();
      for (const GeomData& gd : sc.geomData) common = common && gd.type == TYPE_SDF_SBS_COL;
      if (!common)
        return fail(LRT_ERR_UNSUPPORTED, "LAMBERT resolve needs material tables unless every geometry carries "
                                         "per-voxel colour");
    }
    if (!debug && !common)
      return fail(LRT_ERR_UNSUPPORTED, "render mode " + std::to_string(m) + " needs the renderer's material, "
                                       "texture or channel tables");
    ResolveParams P{};
    std::memcpy(P.proj_inv, desc->proj_inv16, sizeof(P.proj_inv));
    std::memcpy(P.world_view_inv, desc->world_view_inv16, sizeof(P.world_view_inv));
    P.width = w;
    P.height = h;
    P.packed_w = 8 * ((w + 7) / 8);   // MultiRenderer_host.cpp:80-81
    P.packed_h = 8 * ((h + 7) / 8);
    P.count = w * h;
    P.order = desc->order;
    P.render_mode = m;
    P.normal_mode = sc.preset.normal_mode;
    P.spp = sc.preset.spp ? sc.preset.spp : 1;
    P.aa_frame = desc->aa_frame;
    P.sample_id = desc->sample_id;
    P.n_lights = desc->n_lights;
    for (uint32_t i = 0; i < desc->n_lights; ++i) std::memcpy(&P.lights[i], &desc->lights[i], sizeof(ResolveLight));
    std::memcpy(P.background, desc->background, sizeof(P.background));
    P.highlight_mode = desc->highlight_mode;
    P.highlight_prim = desc->highlight_prim;
    P.highlight_geom = desc->highlight_geom;
    P.highlight_inst = desc->highlight_inst;
    const SceneView v = sc.make_view();
    P.geoms = v.geoms;
    P.vert_pos = v.vert_pos;
    P.vert_norm = v.vert_norm;
    auto dptr = [&](const char* n) -> const void* {
      auto it = sc.dev.find(n);
      return (it == sc.dev.end() || it->second.bytes() == 0) ? nullptr : it->second.ptr();
    };
    P.orig_indices = static_cast<const uint32_t*>(dptr("orig_indices"));
    P.inst_inv_t = static_cast<const float*>(dptr("inst_inv_t"));
    hipError_t e = launch_resolve(P, reinterpret_cast<const CRT_Hit*>(gbuffer), reinterpret_cast<lrt::float4*>(colour),
                                  reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "resolve launch");
    return LRT_OK;
  });
}

lrt_status lrt_tonemap_dev(const float* colour, uint64_t n, uint32_t sample_count, uint32_t* rgba8, void* stream) {
  return guarded([&]() -> lrt_status {
    if (n == 0) return LRT_OK;
    if (!colour || !rgba8) return fail(LRT_ERR_INVALID_ARG, "null buffer");
    if (n > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many pixels");
    if (sample_count == 0) return fail(LRT_ERR_INVALID_ARG, "sample_count must be positive");
    hipError_t e = lrt::launch_tonemap(reinterpret_cast<const lrt::float4*>(colour), uint32_t(n), sample_count, rgba8,
                                       reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "tonemap launch");
    return LRT_OK;
  });
}

lrt_status lrt_eye_rays_dev(const float* proj_inv16, const float* world_view_inv16, uint32_t width, uint32_t height,
                            const float* jitter, uint32_t order, float* pos_near, float* dir_far, void* stream) {
  return guarded([&]() -> lrt_status {
    if (!proj_inv16 || !world_view_inv16 || !pos_near || !dir_far) return fail(LRT_ERR_INVALID_ARG, "null argument");
    if (width == 0 || height == 0 || width > 65535 || height > 65535) return fail(LRT_ERR_INVALID_ARG, "bad viewport");
    if (order > 1) return fail(LRT_ERR_INVALID_ARG, "order must be 0 or 1");
    if (order == 1 && (width % 8 != 0 || height % 8 != 0))
      return fail(LRT_ERR_INVALID_ARG, "packed order needs width and height multiples of 8");
    lrt::EyeRayParams p{};
    std::memcpy(p.proj_inv, proj_inv16, sizeof(p.proj_inv));
    std::memcpy(p.world_view_inv, world_view_inv16, sizeof(p.world_view_inv));
    p.width = width;
    p.height = height;
    p.packed_w = 8 * ((width + 7) / 8);
    p.packed_h = 8 * ((height + 7) / 8);
    // spp = 1, sample 0 (MultiRenderer.cpp:163-166): offset (0.5, 0.5)
    p.d.x = 0.5f;
    p.d.y = 0.5f;
    p.order = order;
    hipError_t e = lrt::launch_eye_rays(p, reinterpret_cast<const lrt::float2*>(jitter),
                                        reinterpret_cast<lrt::float4*>(pos_near), reinterpret_cast<lrt::float4*>(dir_far),
                                        reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "eye ray launch");
    return LRT_OK;
  });
}

lrt_status lrt_set_launch_config(lrt_scene* s, int block, int lds_stack, int xcd_remap) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (block != 64 && block != 128 && block != 256) return fail(LRT_ERR_INVALID_ARG, "block must be 64, 128 or 256");
    if (lds_stack != 8 && lds_stack != 16 && lds_stack != 32)
      return fail(LRT_ERR_INVALID_ARG, "lds_stack must be 8, 16 or 32");
    s->s.block = block;
    s->s.lds_stack = lds_stack;
    s->s.xcd_remap = xcd_remap ? 1 : 0;
    return LRT_OK;
  });
}

lrt_status lrt_set_tail_config(lrt_scene* s, uint32_t budget, uint32_t tail_waves) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (tail_waves > 65536) return fail(LRT_ERR_INVALID_ARG, "tail_waves must be <= 65536");
    s->s.budget = budget;
    s->s.tail_waves = tail_waves ? tail_waves : 1024u;
    return LRT_OK;
  });
}

lrt_status lrt_set_ray_interleave(lrt_scene* s, uint32_t lanes_per_group) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (lanes_per_group > 64 || (lanes_per_group & (lanes_per_group - 1)) != 0)
      return fail(LRT_ERR_INVALID_ARG, "lanes_per_group must be 0 or a power of two up to 64");
    s->s.ray_groups = lanes_per_group;
    return LRT_OK;
  });
}

lrt_status lrt_set_persistent(lrt_scene* s, int on) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (on != 0 && on != 1) return fail(LRT_ERR_INVALID_ARG, "on must be 0 or 1");
    s->s.persistent = on;
    return LRT_OK;
  });
}

lrt_status lrt_set_leaf_postponement(lrt_scene* s, int mode) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (mode < -1 || mode > 1) return fail(LRT_ERR_INVALID_ARG, "mode must be -1, 0 or 1");
    s->s.leaf_postpone = mode;
    return LRT_OK;
  });
}

lrt_status lrt_autotune(lrt_scene* s, const float* pos_near, const float* dir_far, lrt_hit* hits, uint64_t n,
                        void* stream, lrt_tune_result* out) {
  return guarded([&]() -> lrt_status {
    if (!s) return fail(LRT_ERR_INVALID_ARG, "null scene");
    if (!s->s.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
    if (!pos_near || !dir_far || !hits || n == 0) return fail(LRT_ERR_INVALID_ARG, "autotune needs a non-empty ray batch");
    lrt::Scene& sc = s->s;
    const lrt::TuneConfig start{sc.block, sc.lds_stack, sc.xcd_remap, sc.budget, sc.tail_waves,
                                sc.postpone_triangles() ? 1 : 0, sc.ray_groups, sc.persistent};
    // the option only exists for triangle-only scenes
    const bool tune_postpone = sc.leaf_kinds() == lrt::LK_TRI;
    auto apply = [&](const lrt::TuneConfig& c) {
      sc.block = c.block;
      sc.lds_stack = c.lds_stack;
      sc.xcd_remap = c.xcd_remap;
      sc.budget = c.budget;
      sc.tail_waves = c.tail_waves;
      sc.ray_groups = c.ray_groups;
      sc.persistent = c.persistent;
      if (tune_postpone) sc.leaf_postpone = c.postpone;
    };
    lrt_status st = LRT_OK;
    auto launch = [&](const lrt::TuneConfig& c) -> hipError_t {
      apply(c);
      st = trace_common(s, pos_near, dir_far, hits, nullptr, n, stream);
      return st == LRT_OK ? hipSuccess : hipErrorLaunchFailure;
    };
    lrt::TuneResult r{};
    const int start_mode = sc.leaf_postpone;
    const hipError_t e = lrt::autotune(launch, reinterpret_cast<hipStream_t>(stream), start, tune_postpone, 3, &r);
    if (e != hipSuccess) {
      apply(start);
      sc.leaf_postpone = start_mode;
      return st != LRT_OK ? st : hip_fail(e, "autotune");
    }
    apply(r.best);
    if (out) {
      out->block = r.best.block;
      out->lds_stack = r.best.lds_stack;
      out->xcd_remap = r.best.xcd_remap;
      out->budget = r.best.budget;
      out->tail_waves = r.best.tail_waves;
      out->best_ms = r.best_ms;
      out->configs_tried = r.tried;
      out->leaf_postpone = r.best.postpone;
      out->ray_interleave = r.best.ray_groups;
      out->persistent = r.best.persistent;
    }
    return LRT_OK;
  });
}

lrt_status lrt_sbs_param_layout(lrt_scene* s, uint32_t geom_id, uint32_t* f_offset, uint32_t* n_dist,
                                uint64_t* n_params) {
  return guarded([&]() -> lrt_status {
    uint32_t off = 0, nd = 0;
    lrt_status st = diff_geometry(s, geom_id, &off, &nd);
    if (st != LRT_OK) return st;
    if (f_offset) *f_offset = off;
    if (n_dist) *n_dist = nd;
    if (n_params) *n_params = s->s.sbsDataF.size();
    return LRT_OK;
  });
}

lrt_status lrt_sbs_forward_dev(lrt_scene* s, uint32_t geom_id, const float* params, const float* pos_near,
                               const float* dir_far, uint64_t n, float* colour, void* stream) {
  return guarded([&]() -> lrt_status {
    if (n && !colour) return fail(LRT_ERR_INVALID_ARG, "null colour buffer");
    return sbs_diff_common(s, geom_id, params, pos_near, dir_far, n, colour, nullptr, nullptr, nullptr, nullptr,
                           stream);
  });
}

lrt_status lrt_sbs_backward_dev(lrt_scene* s, uint32_t geom_id, const float* params, const float* pos_near,
                                const float* dir_far, uint64_t n, const float* dL_dcolour, float* dL_dparams,
                                void* stream) {
  return guarded([&]() -> lrt_status {
    if (n && (!dL_dcolour || !dL_dparams)) return fail(LRT_ERR_INVALID_ARG, "null gradient buffer");
    return sbs_diff_common(s, geom_id, params, pos_near, dir_far, n, nullptr, dL_dcolour, nullptr, dL_dparams,
                           nullptr, stream);
  });
}

lrt_status lrt_sbs_l2_backward_dev(lrt_scene* s, uint32_t geom_id, const float* params, const float* pos_near,
                                   const float* dir_far, uint64_t n, const float* target, float* dL_dparams,
                                   float* loss, void* stream) {
  return guarded([&]() -> lrt_status {
    if (n && (!target || !dL_dparams)) return fail(LRT_ERR_INVALID_ARG, "null target or gradient buffer");
    return sbs_diff_common(s, geom_id, params, pos_near, dir_far, n, nullptr, nullptr, target, dL_dparams, loss,
                           stream);
  });
}

lrt_status lrt_sbs_render_l2_backward_dev(lrt_scene* s, uint32_t geom_id, const float* params, const float* pos_near,
                                          const float* dir_far, uint64_t n, const float* target, float* colour,
                                          float* dL_dparams, float* loss, void* stream) {
  return guarded([&]() -> lrt_status {
    if (n && (!target || !colour || !dL_dparams))
      return fail(LRT_ERR_INVALID_ARG, "null target, colour or gradient buffer");
    return sbs_diff_common(s, geom_id, params, pos_near, dir_far, n, colour, nullptr, target, dL_dparams, loss,
                           stream);
  });
}

lrt_status lrt_sbs_batch_create(const lrt_sbs_batch_item* items, uint32_t n_items, lrt_sbs_batch** out) {
  return guarded([&]() -> lrt_status {
    if (!out || (!items && n_items)) return fail(LRT_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (n_items > 65535) return fail(LRT_ERR_INVALID_ARG, "at most 65535 items per batch");
    auto b = std::make_unique<lrt_sbs_batch>();
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) return hip_fail(e, "no device");
    std::vector<uint32_t> chunk_begin{0};
    uint64_t total_bricks = 0;
    std::vector<size_t> mask_offset;
    size_t mask_words = 0;
    uint64_t max_zero4 = 1, max_bits = 1, max_bricks = 1;   // per-item extents of the set-up grids
    size_t max_flag_bytes = 4;
    for (uint32_t k = 0; k < n_items; ++k) {
      const lrt_sbs_batch_item& in = items[k];
      if (!in.scene) return fail(LRT_ERR_INVALID_ARG, "null scene in batch item");
      lrt::Scene& sc = in.scene->s;
      uint32_t f_offset = 0, n_dist = 0;
      lrt_status st = diff_geometry(in.scene, in.geom_id, &f_offset, &n_dist);
      if (st != LRT_OK) return st;
      if (!sc.committed) return fail(LRT_ERR_NOT_COMMITTED, "batch item scene is not committed");
      if (sc.sbsRoots.size() != 1) return fail(LRT_ERR_UNSUPPORTED, "batch items need scenes with exactly one brick set");
      if (!in.params || !in.grad || (in.n_rays && (!in.pos_near || !in.dir_far || !in.target)))
        return fail(LRT_ERR_INVALID_ARG, "null buffer in batch item");
      if (in.n_rays > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many rays in one item");
      lrt::SbsBatchItem it{};
      it.view = sc.make_view();
      it.view.sbs_data_f = in.params;
      it.view.cache_pairs = 0;
      it.view.cache_geom = 0xFFFFFFFFu;
      it.args.pos = reinterpret_cast<const lrt::float4*>(in.pos_near);
      it.args.dir = reinterpret_cast<const lrt::float4*>(in.dir_far);
      it.args.n = uint32_t(in.n_rays);
      it.args.params = in.params;
      it.args.geom_id = in.geom_id;
      it.args.f_offset = f_offset;
      it.args.n_dist = n_dist;
      it.args.colour = reinterpret_cast<lrt::float4*>(in.colour);
      it.args.target = reinterpret_cast<const lrt::float4*>(in.target);
      it.args.grad = in.grad;
      it.args.loss = in.loss;
      it.args.counters = sc.d_overflow ? sc.d_overflow + 1 : nullptr;
      it.n_bricks = sc.sbsMaskGeoms[0].n_bricks;
      it.brick_size = sc.sbsMaskGeoms[0].brick_size;
      it.n_params = sc.sbsDataF.size();
      const uint32_t wpb = (it.brick_size * it.brick_size * it.brick_size + 31u) / 32u;
      mask_offset.push_back(mask_words);
      mask_words += size_t(it.n_bricks) * (1 + wpb);
      max_zero4 = std::max<uint64_t>(max_zero4, (it.n_params + 3) / 4);
      max_bits = std::max<uint64_t>(max_bits, uint64_t(it.n_bricks) * wpb * 32u);
      max_bricks = std::max<uint64_t>(max_bricks, it.n_bricks);
      max_flag_bytes = std::max<size_t>(max_flag_bytes, size_t(lrt::sbs_mask_flag_words(it.brick_size)) * 4u);
      total_bricks += it.n_bricks;
      chunk_begin.push_back(chunk_begin.back() + (it.args.n + lrt::SBS_BATCH_CHUNK - 1) / lrt::SBS_BATCH_CHUNK);
      b->items.push_back(it);
      b->scene_gen.push_back(sc.generation);
      b->gen_at_create.push_back(*sc.generation);
    }
    e = b->d_masks.reserve(std::max<size_t>(mask_words, 1) * 4);
    if (e != hipSuccess) return hip_fail(e, "batch masks");
    for (uint32_t k = 0; k < n_items; ++k) {
      b->items[k].mask = static_cast<uint32_t*>(b->d_masks.ptr()) + mask_offset[k];
      b->items[k].view.sbs_mask = b->items[k].mask;
    }
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device)) != hipSuccess)
      return hip_fail(e, "device attribute");
    // resident blocks of every kernel variant (LDS stack 8/16/32 x voxel record off/on); the spill
    // covers the largest grid
    uint32_t max_grid = 1;
    for (int v = 0; v < 6; ++v) {
      int per_cu = 0;
      if ((e = lrt::sbs_batch_blocks_per_cu(8 << (v / 2), v % 2, &per_cu)) != hipSuccess)
        return hip_fail(e, "occupancy");
      b->resident[v] = uint32_t(std::max(1, cus * std::max(1, per_cu)));
      max_grid = std::max(max_grid, b->resident[v]);
    }
    const uint32_t grid = b->resident[2];   // LDS stack 16, hit brick marched again
    const size_t n1 = size_t(n_items) + 1;
    if ((e = b->d_items.upload(b->items.data(), b->items.size() * sizeof(lrt::SbsBatchItem))) != hipSuccess)
      return hip_fail(e, "batch table");
    if ((e = b->d_begins.upload(chunk_begin.data(), n1 * sizeof(uint32_t))) != hipSuccess)
      return hip_fail(e, "batch table");
    if ((e = b->d_spill.reserve(size_t(max_grid) * lrt::SBS_BATCH_BLOCK * (lrt::STACK_SIZE - 8) * 4)) != hipSuccess)
      return hip_fail(e, "batch spill");
    if ((e = b->d_counter.reserve(256)) != hipSuccess) return hip_fail(e, "batch counter");
    lrt::SbsBatchLaunch& L = b->L;
    L.items = static_cast<const lrt::SbsBatchItem*>(b->d_items.ptr());
    L.n_items = n_items;
    L.chunk_begin = static_cast<const uint32_t*>(b->d_begins.ptr());
    if (total_bricks > 0xFFFFFFFFull) return fail(LRT_ERR_INVALID_ARG, "too many bricks in one batch");
    L.total_bricks = uint32_t(total_bricks);
    L.total_chunks = chunk_begin.back();
    L.counter = static_cast<uint32_t*>(b->d_counter.ptr());
    L.spill = static_cast<uint32_t*>(b->d_spill.ptr());
    L.grid = grid;
    L.lds_stack = 16;
    L.record = 0;
    auto blocks = [](uint64_t work, uint64_t cap) { return uint32_t(std::min<uint64_t>(cap, (work + 255) / 256)); };
    L.zero_blocks = blocks(max_zero4, 1024);
    L.mask_blocks = blocks(max_bits, 4096);
    L.summary_blocks = blocks(max_bricks, 256);
    L.brick_blocks = uint32_t(std::min<uint64_t>(max_bricks, 4096));
    L.mask_flag_bytes = max_flag_bytes <= lrt::SBS_MASK_LDS_MAX ? uint32_t(max_flag_bytes) : 0u;
    *out = b.release();
    return LRT_OK;
  });
}

namespace {
// The plan may run on the current device and its scenes are unchanged since it was made.
lrt_status batch_usable(const lrt_sbs_batch* b) {
  int dev = -1;
  hipError_t ge = hipGetDevice(&dev);
  if (ge != hipSuccess) return hip_fail(ge, "no device");
  if (dev != b->device) return fail(LRT_ERR_INVALID_ARG, "batch plan was created on another device");
  for (size_t k = 0; k < b->scene_gen.size(); ++k) {
    const std::shared_ptr<uint64_t> g = b->scene_gen[k].lock();
    if (!g || *g != b->gen_at_create[k])
      return fail(LRT_ERR_INVALID_ARG, "batch plan is stale: item " + std::to_string(k) +
                                           "'s scene was changed, re-committed or destroyed since the plan was made");
  }
  return LRT_OK;
}
}  // namespace

lrt_status lrt_sbs_batch_autotune(lrt_sbs_batch* b, void* stream, lrt_sbs_batch_tune_result* out) {
  return guarded([&]() -> lrt_status {
    if (!b) return fail(LRT_ERR_INVALID_ARG, "null batch");
    lrt_status st = batch_usable(b);
    if (st != LRT_OK) return st;
    const hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const lrt::SbsBatchLaunch start = b->L;
    lrt::SbsBatchLaunch best = start;
    float best_ms = 1e30f;
    uint32_t tried = 0;
    // LDS stack 8 / 16 / 32 entries x voxel record off / on x 1/2, 3/4 or all of the resident blocks
    for (int v = 0; v < 6 && e == hipSuccess; ++v) {
      for (int quarters : {2, 3, 4}) {
        lrt::SbsBatchLaunch c = start;
        c.lds_stack = 8 << (v / 2);
        c.record = v % 2;
        c.grid = std::max<uint32_t>(1u, b->resident[v] * uint32_t(quarters) / 4u);
        float ms = 1e30f;
        if ((e = lrt::launch_sbs_batch(c, hs)) != hipSuccess) break;   // warm-up
        for (int r = 0; r < 3 && e == hipSuccess; ++r) {
          float t = 0.0f;
          if ((e = hipEventRecord(e0, hs)) != hipSuccess) break;
          if ((e = lrt::launch_sbs_batch(c, hs)) != hipSuccess) break;
          if ((e = hipEventRecord(e1, hs)) != hipSuccess) break;
          if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
          if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) break;
          ms = std::min(ms, t);
        }
        if (e != hipSuccess) break;
        ++tried;
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
      b->L = start;
      return hip_fail(e, "batch autotune");
    }
    b->L = best;
    if (out) {
      out->lds_stack = best.lds_stack;
      out->record = best.record;
      out->grid = best.grid;
      out->best_ms = best_ms;
      out->configs_tried = tried;
    }
    return LRT_OK;
  });
}

lrt_status lrt_sbs_batch_configure(lrt_sbs_batch* b, int lds_stack, int record, uint32_t grid) {
  return guarded([&]() -> lrt_status {
    if (!b) return fail(LRT_ERR_INVALID_ARG, "null batch");
    if (lds_stack != 8 && lds_stack != 16 && lds_stack != 32) return fail(LRT_ERR_INVALID_ARG, "lds_stack must be 8, 16 or 32");
    if (record != 0 && record != 1) return fail(LRT_ERR_INVALID_ARG, "record must be 0 or 1");
    const int v = (lds_stack == 8 ? 0 : lds_stack == 16 ? 2 : 4) + record;
    if (grid == 0 || grid > b->resident[v])
      return fail(LRT_ERR_INVALID_ARG, "grid must be 1.." + std::to_string(b->resident[v]) + " for this variant");
    b->L.lds_stack = lds_stack;
    b->L.record = record;
    b->L.grid = grid;
    return LRT_OK;
  });
}

lrt_status lrt_sbs_batch_run(lrt_sbs_batch* b, void* stream) {
  return guarded([&]() -> lrt_status {
    if (!b) return fail(LRT_ERR_INVALID_ARG, "null batch");
    lrt_status st = batch_usable(b);
    if (st != LRT_OK) return st;
    const hipError_t e = lrt::launch_sbs_batch(b->L, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "batch launch");
    return LRT_OK;
  });
}

lrt_status lrt_sbs_batch_destroy(lrt_sbs_batch* b) {
  return guarded([&]() -> lrt_status {
    delete b;
    return LRT_OK;
  });
}

lrt_status lrt_sbs_recompute_misses(lrt_scene* s, uint64_t* out, int reset) {
  return guarded([&]() -> lrt_status {
    if (!s || !out) return fail(LRT_ERR_INVALID_ARG, "null argument");
    *out = 0;
    if (!s->s.d_overflow) return LRT_OK;
    uint32_t v = 0;
    hipError_t e = hipMemcpy(&v, s->s.d_overflow + 1, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "counter read");
    *out = v;
    if (reset) {
      e = hipMemset(s->s.d_overflow + 1, 0, sizeof(uint32_t));
      if (e != hipSuccess) return hip_fail(e, "counter reset");
    }
    return LRT_OK;
  });
}

lrt_status lrt_stack_overflow_count(lrt_scene* s, uint64_t* out, int reset) {
  return guarded([&]() -> lrt_status {
    if (!s || !out) return fail(LRT_ERR_INVALID_ARG, "null argument");
    *out = 0;
    if (!s->s.d_overflow) return LRT_OK;
    uint32_t v = 0;
    hipError_t e = hipMemcpy(&v, s->s.d_overflow, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "read overflow counter");
    *out = v;
    if (reset) {
      e = hipMemset(s->s.d_overflow, 0, sizeof(uint32_t));
      if (e != hipSuccess) return hip_fail(e, "reset overflow counter");
    }
    return LRT_OK;
  });
}

lrt_status lrt_get_array(lrt_scene* s, const char* name, void* dst, size_t* bytes) {
  return guarded([&]() -> lrt_status {
    if (!s || !name || !bytes) return fail(LRT_ERR_INVALID_ARG, "null argument");
    const lrt::Scene& c = s->s;
    const std::string n = name;
    const void* src = nullptr;
    size_t sz = 0;
#define ARR(key, vec)                                   \
  if (n == key) { src = vec.data(); sz = vec.size() * sizeof(vec[0]); }
    ARR("nodes", c.nodePairs) ARR("tlas", c.tlas) ARR("instances", c.instances) ARR("geoms", c.geomData)
    ARR("geom_tags", c.geomTags) ARR("start_end", c.startEnd) ARR("prim_id_count", c.primIdCount)
    ARR("vert_pos", c.vertPos) ARR("vert_norm", c.vertNorm) ARR("indices", c.indices)
    ARR("prim_indices", c.primIndices) ARR("orig_nodes", c.origNodes) ARR("frame_nodes", c.frameNodes)
    ARR("frame_roots", c.frameRoots) ARR("frame_tex_nodes", c.frameTexNodes) ARR("frame_tex_roots", c.frameTexRoots) ARR("sbs_nodes", c.sbsNodes) ARR("sbs_data", c.sbsData)
    ARR("sbs_data_f", c.sbsDataF) ARR("sbs_roots", c.sbsRoots) ARR("sbs_headers", c.sbsHeaders)
    ARR("multi_nodes", c.multiNodes) ARR("multi_values", c.multiValues) ARR("orig_indices", c.origIndices)
    ARR("svs_nodes", c.svsNodes) ARR("svs_roots", c.svsRoots) ARR("tlas_parent", c.tlasParent)
    ARR("grid_data", c.gridData) ARR("grid_offsets", c.gridOffsets) ARR("grid_sizes", c.gridSizes)
    ARR("gprim_headers", c.gprimHeaders) ARR("gprim_points", c.gprimPoints)
    ARR("coct_data", c.coctData) ARR("coct_geoms", c.coctGeoms) ARR("coct_mods", c.coctMods)
    ARR("dag_nodes", c.dagNodes) ARR("dag_child_edges", c.dagChildEdges) ARR("dag_data_edges", c.dagDataEdges)
    ARR("dag_distances", c.dagDistances) ARR("dag_geoms", c.dagGeoms) ARR("dag_mods", c.dagMods)
    if (n == "rot_add") { src = lrt::rot_add_table().data(); sz = lrt::rot_add_table().size(); }
    ARR("scom2_data", c.scom2Data) ARR("scom2_headers", c.scom2Headers) ARR("scom2_mods", c.scom2Mods)
#undef ARR
    if (n == "preset") { src = &c.preset; sz = sizeof(c.preset); }
    std::vector<uint32_t> mask_copy;   // the committed occupancy masks live on the device only
    if (n == "sbs_mask") {
      if (!c.committed) return fail(LRT_ERR_NOT_COMMITTED, "scene is not committed");
      mask_copy.resize(c.sbsMaskWords);
      if (c.sbsMaskWords) {
        auto it = c.dev.find("sbs_mask");
        if (it == c.dev.end()) return fail(LRT_ERR_HIP, "no device mask");
        const hipError_t e = hipMemcpy(mask_copy.data(), it->second.ptr(), mask_copy.size() * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "mask copy");
      }
      src = mask_copy.data();
      sz = mask_copy.size() * 4;
    }
    if (!src && sz == 0) {
      static const char* known[] = {"nodes", "tlas", "instances", "geoms", "geom_tags", "start_end",
                                    "prim_id_count", "vert_pos", "vert_norm", "indices", "prim_indices",
                                    "orig_nodes", "frame_nodes", "frame_roots", "frame_tex_nodes", "frame_tex_roots",
                                    "sbs_nodes", "sbs_data",
                                    "sbs_data_f", "sbs_roots", "sbs_headers", "multi_nodes", "multi_values", "preset", "orig_indices", "svs_nodes", "svs_roots", "tlas_parent", "grid_data", "grid_offsets", "grid_sizes", "gprim_headers", "gprim_points", "coct_data", "coct_geoms", "coct_mods", "dag_nodes", "dag_child_edges", "dag_data_edges", "dag_distances", "dag_geoms", "dag_mods", "rot_add", "scom2_data", "scom2_headers", "scom2_mods",
                                    "sbs_mask"};
      bool ok = false;
      for (const char* k : known) ok = ok || n == k;
      if (!ok) return fail(LRT_ERR_INVALID_ARG, "unknown array '" + n + "'");
    }
    if (dst) {
      if (*bytes < sz) return fail(LRT_ERR_INVALID_ARG, "destination too small");
      if (sz) std::memcpy(dst, src, sz);
    }
    *bytes = sz;
    return LRT_OK;
  });
}

}  // extern "C"
