// extern "C" entry points declared in include/lrt_build.h.
#include "../../include/lrt_build.h"

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "gen.h"
#include "host_math.h"

namespace lrt {
void set_last_error(const std::string& msg);
}

namespace {

thread_local std::string g_gen_err;

template <class T>
lrt_status to_buffer(const std::vector<T>& v, lrt_buffer* out) {
  out->bytes = v.size() * sizeof(T);
  out->data = std::malloc(out->bytes ? out->bytes : 1);
  if (!out->data) return LRT_ERR_INVALID_ARG;
  if (out->bytes) std::memcpy(out->data, v.data(), out->bytes);
  return LRT_OK;
}

template <class F>
lrt_status guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_gen_err = e.what();
    lrt::set_last_error(g_gen_err);
    return LRT_ERR_INVALID_ARG;
  }
}

lrt_status mesh_out(std::vector<double>& v, const std::vector<uint32_t>& f, float amplitude, float frequency,
                    uint64_t seed, lrt_buffer* vpos, lrt_buffer* indices) {
  if (amplitude != 0.0f) lrt::displace_radial(v, amplitude, frequency, seed);
  std::vector<float> vf(v.begin(), v.end());
  lrt_status st = to_buffer(vf, vpos);
  if (st != LRT_OK) return st;
  return to_buffer(f, indices);
}

}  // namespace

extern "C" {

void lrt_buffer_free(lrt_buffer* b) {
  if (!b) return;
  std::free(b->data);
  b->data = nullptr;
  b->bytes = 0;
}

lrt_status lrt_camera_inverse(const float* eye3, const float* target3, const float* up3, float fovy_deg, float aspect,
                              float znear, float zfar, float* proj_inv16, float* world_view_inv16) {
  if (!eye3 || !target3 || !up3 || !proj_inv16 || !world_view_inv16) return LRT_ERR_INVALID_ARG;
  using namespace lrt;
  const float4x4 view = look_at(make3(eye3[0], eye3[1], eye3[2]), make3(target3[0], target3[1], target3[2]),
                                make3(up3[0], up3[1], up3[2]));
  const float4x4 proj = perspective_matrix(fovy_deg, aspect, znear, zfar);
  const float4x4 pi = inverse4x4(proj), vi = inverse4x4(view);
  std::memcpy(proj_inv16, pi.m, sizeof(pi.m));
  std::memcpy(world_view_inv16, vi.m, sizeof(vi.m));
  return LRT_OK;
}

lrt_status lrt_gen_icosphere(uint32_t subdiv, float amplitude, float frequency, uint64_t seed, lrt_buffer* vpos,
                             lrt_buffer* indices) {
  if (!vpos || !indices || subdiv > 10) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    std::vector<double> v;
    std::vector<uint32_t> f;
    lrt::make_icosphere(subdiv, v, f);
    return mesh_out(v, f, amplitude, frequency, seed, vpos, indices);
  });
}

lrt_status lrt_gen_geodesic(uint32_t n, float amplitude, float frequency, uint64_t seed, lrt_buffer* vpos,
                            lrt_buffer* indices) {
  if (!vpos || !indices || n == 0 || n > 2048) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    std::vector<double> v;
    std::vector<uint32_t> f;
    lrt::make_geodesic(n, v, f);
    return mesh_out(v, f, amplitude, frequency, seed, vpos, indices);
  });
}

lrt_status lrt_sdf_spheres_frame_octree(const float* spheres, size_t n_spheres, uint32_t depth, int threads,
                                        lrt_buffer* nodes) {
  if (!spheres || n_spheres == 0 || !nodes) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    lrt::GlobalOctree g;
    lrt::build_global_octree(lrt::SphereUnion{spheres, n_spheres}, depth, 1, 0, threads, g);
    std::vector<lrt::SdfFrameOctreeNode> frame;
    lrt::global_octree_to_frame_octree(g, frame);
    return to_buffer(frame, nodes);
  });
}

lrt_status lrt_sdf_spheres_multi_octree(const float* spheres, size_t n_spheres, uint32_t depth, int threads,
                                        lrt_buffer* nodes, lrt_buffer* values) {
  if (!spheres || n_spheres == 0 || !nodes || !values) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    lrt::GlobalOctree g;
    lrt::build_global_octree(lrt::SphereUnion{spheres, n_spheres}, depth, 1, 0, threads, g);
    std::vector<lrt::SdfMultiOctreeNode> n;
    std::vector<float> v;
    lrt::global_octree_to_multi_octree(g, n, v);
    lrt_status st = to_buffer(n, nodes);
    if (st != LRT_OK) return st;
    return to_buffer(v, values);
  });
}

lrt_status lrt_sdf_spheres_sbs(const float* spheres, size_t n_spheres, uint32_t depth, const void* header,
                               uint64_t colour_seed, int threads, lrt_buffer* nodes, lrt_buffer* values) {
  if (!spheres || n_spheres == 0 || !header || !nodes || !values) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    const lrt::SdfSBSHeader h = *static_cast<const lrt::SdfSBSHeader*>(header);
    if (h.bytes_per_value != 1 && h.bytes_per_value != 2 && h.bytes_per_value != 4) return LRT_ERR_INVALID_ARG;
    lrt::GlobalOctree g;
    lrt::build_global_octree(lrt::SphereUnion{spheres, n_spheres}, depth, h.brick_size, h.brick_pad, threads, g);
    std::vector<lrt::SdfSBSNode> n;
    std::vector<uint32_t> v;
    lrt::global_octree_to_sbs_dx(g, h, colour_seed, n, v);
    lrt_status st = to_buffer(n, nodes);
    if (st != LRT_OK) return st;
    return to_buffer(v, values);
  });
}

lrt_status lrt_sdf_spheres_sbs_indexed(const float* spheres, size_t n_spheres, uint32_t depth, uint32_t brick_size,
                                       uint64_t albedo_seed, int threads, lrt_buffer* nodes, lrt_buffer* values,
                                       lrt_buffer* values_f, uint32_t* n_distance_values) {
  if (!spheres || n_spheres == 0 || !nodes || !values || !values_f) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    lrt::GlobalOctree g;
    lrt::build_global_octree(lrt::SphereUnion{spheres, n_spheres}, depth, brick_size, 0, threads, g);
    std::vector<lrt::SdfSBSNode> n;
    std::vector<uint32_t> v;
    std::vector<float> vf;
    lrt::global_octree_to_sbs_indexed(g, albedo_seed, n, v, vf, n_distance_values);
    lrt_status st = to_buffer(n, nodes);
    if (st != LRT_OK) return st;
    st = to_buffer(v, values);
    if (st != LRT_OK) return st;
    return to_buffer(vf, values_f);
  });
}

lrt_status lrt_sdf_spheres_global_octree(const float* spheres, size_t n_spheres, uint32_t depth, uint32_t brick_size,
                                         uint32_t brick_pad, int threads, lrt_buffer* nodes, lrt_buffer* values) {
  if (!spheres || n_spheres == 0 || !nodes || !values) return LRT_ERR_INVALID_ARG;
  return guarded([&]() -> lrt_status {
    lrt::GlobalOctree g;
    lrt::build_global_octree(lrt::SphereUnion{spheres, n_spheres}, depth, brick_size, brick_pad, threads, g);
    std::vector<uint32_t> packed(g.nodes.size() * 3);
    for (size_t i = 0; i < g.nodes.size(); ++i) {
      packed[3 * i + 0] = g.nodes[i].type;
      packed[3 * i + 1] = g.nodes[i].val_off;
      packed[3 * i + 2] = g.nodes[i].offset;
    }
    lrt_status st = to_buffer(packed, nodes);
    if (st != LRT_OK) return st;
    return to_buffer(g.values, values);
  });
}

}  // extern "C"
