// Reference-format SDF scene files (the binary files the reference's save_sdf_* functions write):
// extern "C" loaders and savers declared in include/lrt_build.h. The loaders check every count
// against the file size and fail with a message instead of reading past the end.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lrt_build.h"
#include "lrt_types.h"

namespace lrt {
void set_last_error(const std::string& msg);
}

namespace {

struct File {
  std::FILE* f = nullptr;
  File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {
    if (!f) throw std::runtime_error(std::string("cannot open '") + (path ? path : "(null)") + "'");
  }
  ~File() {
    if (f) std::fclose(f);
  }
  long size() {
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return n;
  }
};

// Reads `bytes` bytes; `left` is what the file still holds from the current position.
void take(File& fl, void* dst, size_t bytes, size_t& left, const char* what) {
  if (bytes > left) throw std::runtime_error(std::string("file ends inside ") + what);
  if (bytes && std::fread(dst, 1, bytes, fl.f) != bytes) throw std::runtime_error(std::string("read error in ") + what);
  left -= bytes;
}

void put(File& fl, const void* src, size_t bytes) {
  if (bytes && std::fwrite(src, 1, bytes, fl.f) != bytes) throw std::runtime_error("write error");
}

lrt_status to_buf(const std::vector<char>& v, lrt_buffer* out) {
  out->bytes = v.size();
  out->data = std::malloc(v.size() ? v.size() : 1);
  if (!out->data) throw std::runtime_error("out of host memory");
  if (!v.empty()) std::memcpy(out->data, v.data(), v.size());
  return LRT_OK;
}

// u32 count then count records of `rec` bytes
std::vector<char> counted(File& fl, size_t rec, size_t& left, const char* what) {
  uint32_t n = 0;
  take(fl, &n, 4, left, what);
  if (size_t(n) * rec > left) throw std::runtime_error(std::string(what) + ": count " + std::to_string(n) +
                                                       " exceeds the file");
  std::vector<char> v(size_t(n) * rec);
  take(fl, v.data(), v.size(), left, what);
  return v;
}

template <class F>
lrt_status io_guarded(F&& f) {
  try {
    lrt::set_last_error("");
    return f();
  } catch (const std::exception& e) {
    lrt::set_last_error(e.what());
    return LRT_ERR_INVALID_ARG;
  }
}

void check_u32(size_t n, const char* what) {
  if (n > 0xFFFFFFFFull) throw std::runtime_error(std::string(what) + ": too many records for the file format");
}

}  // namespace

extern "C" {

lrt_status lrt_load_sdf_frame_octree(const char* path, lrt_buffer* nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes) throw std::runtime_error("null output buffer");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    std::vector<char> v = counted(fl, sizeof(lrt::SdfFrameOctreeNode), left, "frame-octree nodes");
    return to_buf(v, nodes);
  });
}

lrt_status lrt_save_sdf_frame_octree(const char* path, const void* nodes, size_t n_nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes && n_nodes) throw std::runtime_error("null nodes");
    check_u32(n_nodes, "frame-octree nodes");
    File fl(path, "wb");
    const uint32_t n = uint32_t(n_nodes);
    put(fl, &n, 4);
    put(fl, nodes, n_nodes * sizeof(lrt::SdfFrameOctreeNode));
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_frame_octree_tex(const char* path, lrt_buffer* nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes) throw std::runtime_error("null output buffer");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    std::vector<char> v = counted(fl, sizeof(lrt::SdfFrameOctreeTexNode), left, "textured frame-octree nodes");
    return to_buf(v, nodes);
  });
}

lrt_status lrt_save_sdf_frame_octree_tex(const char* path, const void* nodes, size_t n_nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes && n_nodes) throw std::runtime_error("null nodes");
    check_u32(n_nodes, "textured frame-octree nodes");
    File fl(path, "wb");
    const uint32_t n = uint32_t(n_nodes);
    put(fl, &n, 4);
    put(fl, nodes, n_nodes * sizeof(lrt::SdfFrameOctreeTexNode));
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_svs(const char* path, lrt_buffer* nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes) throw std::runtime_error("null output buffer");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    std::vector<char> v = counted(fl, sizeof(lrt::SdfSVSNode), left, "sparse-voxel-set nodes");
    return to_buf(v, nodes);
  });
}

lrt_status lrt_save_sdf_svs(const char* path, const void* nodes, size_t n_nodes) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes && n_nodes) throw std::runtime_error("null nodes");
    check_u32(n_nodes, "sparse-voxel-set nodes");
    File fl(path, "wb");
    const uint32_t n = uint32_t(n_nodes);
    put(fl, &n, 4);
    put(fl, nodes, n_nodes * sizeof(lrt::SdfSVSNode));
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_grid(const char* path, uint32_t size[3], lrt_buffer* data) {
  return io_guarded([&]() -> lrt_status {
    if (!size || !data) throw std::runtime_error("null output");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    take(fl, size, 12, left, "grid size");
    const uint64_t n = uint64_t(size[0]) * size[1] * size[2];
    if (n * 4 > left) throw std::runtime_error("grid of " + std::to_string(n) + " values exceeds the file");
    std::vector<char> v(size_t(n) * 4);
    take(fl, v.data(), v.size(), left, "grid values");
    return to_buf(v, data);
  });
}

lrt_status lrt_save_sdf_grid(const char* path, const uint32_t size[3], const float* data) {
  return io_guarded([&]() -> lrt_status {
    if (!size || !data) throw std::runtime_error("null argument");
    const uint64_t n = uint64_t(size[0]) * size[1] * size[2];
    File fl(path, "wb");
    put(fl, size, 12);
    put(fl, data, size_t(n) * 4);
    return LRT_OK;
  });
}

lrt_status lrt_load_coctree_v3(const char* path, uint32_t* version, uint32_t header[15], lrt_buffer* data) {
  return io_guarded([&]() -> lrt_status {
    if (!header || !data) throw std::runtime_error("null output");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    uint32_t ver = 0;
    take(fl, &ver, 4, left, "COctreeV3 version");
    if (version) *version = ver;
    take(fl, header, sizeof(lrt::COctreeV3Header), left, "COctreeV3 header");
    std::vector<char> v = counted(fl, 4, left, "COctreeV3 words");
    return to_buf(v, data);
  });
}

lrt_status lrt_save_coctree_v3(const char* path, const uint32_t header[15], const uint32_t* data, size_t size) {
  return io_guarded([&]() -> lrt_status {
    if (!header || (!data && size)) throw std::runtime_error("null argument");
    if (size > 0xFFFFFFFFull) throw std::runtime_error("too many words for a u32 count");
    const uint32_t ver = lrt::COCTREE_V3_VERSION, n = uint32_t(size);
    File fl(path, "wb");
    put(fl, &ver, 4);
    put(fl, header, sizeof(lrt::COctreeV3Header));
    put(fl, &n, 4);
    put(fl, data, size * 4);
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_dag(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* child_edges,
                            lrt_buffer* data_edges, lrt_buffer* distances) {
  return io_guarded([&]() -> lrt_status {
    if (!header || !nodes || !child_edges || !data_edges || !distances) throw std::runtime_error("null output");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    uint32_t counts[6];
    take(fl, counts, sizeof(counts), left, "SdfDAG counts");
    if (counts[4] || counts[5])
      throw std::runtime_error("SdfDAG files with data channels are not read (the intersection does not use them)");
    take(fl, header, sizeof(lrt::SdfDAGHeader), left, "SdfDAG header");
    const size_t rec[4] = {sizeof(lrt::SdfDAGNode), sizeof(lrt::SdfDAGChildEdge), sizeof(lrt::SdfDAGDataEdge), 4};
    lrt_buffer* out[4] = {nodes, child_edges, data_edges, distances};
    const char* what[4] = {"SdfDAG nodes", "SdfDAG child edges", "SdfDAG data edges", "SdfDAG distances"};
    std::vector<char> v[4];
    for (int i = 0; i < 4; ++i) {
      if (uint64_t(counts[i]) * rec[i] > left)
        throw std::runtime_error(std::string(what[i]) + ": count " + std::to_string(counts[i]) + " exceeds the file");
      v[i].resize(size_t(counts[i]) * rec[i]);
      take(fl, v[i].data(), v[i].size(), left, what[i]);
    }
    for (int i = 0; i < 4; ++i) to_buf(v[i], out[i]);
    return LRT_OK;
  });
}

lrt_status lrt_save_sdf_dag(const char* path, const void* header, const void* nodes, size_t n_nodes,
                            const void* child_edges, size_t n_child_edges, const void* data_edges, size_t n_data_edges,
                            const float* distances, size_t n_distances) {
  return io_guarded([&]() -> lrt_status {
    if (!header || (!nodes && n_nodes) || (!child_edges && n_child_edges) || (!data_edges && n_data_edges) ||
        (!distances && n_distances))
      throw std::runtime_error("null argument");
    if (n_nodes > 0xFFFFFFFFull || n_child_edges > 0xFFFFFFFFull || n_data_edges > 0xFFFFFFFFull ||
        n_distances > 0xFFFFFFFFull)
      throw std::runtime_error("too many records for a u32 count");
    const uint32_t counts[6] = {uint32_t(n_nodes), uint32_t(n_child_edges), uint32_t(n_data_edges),
                                uint32_t(n_distances), 0, 0};
    File fl(path, "wb");
    put(fl, counts, sizeof(counts));
    put(fl, header, sizeof(lrt::SdfDAGHeader));
    put(fl, nodes, n_nodes * sizeof(lrt::SdfDAGNode));
    put(fl, child_edges, n_child_edges * sizeof(lrt::SdfDAGChildEdge));
    put(fl, data_edges, n_data_edges * sizeof(lrt::SdfDAGDataEdge));
    put(fl, distances, n_distances * 4);
    return LRT_OK;
  });
}

lrt_status lrt_load_scom2(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* bricks) {
  return io_guarded([&]() -> lrt_status {
    if (!header || !nodes || !bricks) throw std::runtime_error("null output");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    uint32_t head[6];
    take(fl, head, 8, left, "SCom2 magic and version");
    if (head[0] != lrt::SCOM2_MAGIC_NUMBER)
      throw std::runtime_error("not an SCom2 file of the current format (no magic number; legacy files are not read)");
    if (head[1] != lrt::SCOM2_VERSION)
      throw std::runtime_error("SCom2 version " + std::to_string(head[1]) + " (the reference reads version 4 only)");
    take(fl, head + 2, 16, left, "SCom2 counts");
    if (head[4] || head[5])
      throw std::runtime_error("SCom2 files with data channels are not read (the intersection does not use them)");
    take(fl, header, sizeof(lrt::SCom2Header), left, "SCom2 header");
    if (uint64_t(head[2]) * 4 > left) throw std::runtime_error("SCom2 nodes exceed the file");
    std::vector<char> n(size_t(head[2]) * 4);
    take(fl, n.data(), n.size(), left, "SCom2 nodes");
    if (uint64_t(head[3]) * 4 > left) throw std::runtime_error("SCom2 bricks exceed the file");
    std::vector<char> b(size_t(head[3]) * 4);
    take(fl, b.data(), b.size(), left, "SCom2 bricks");
    to_buf(n, nodes);
    return to_buf(b, bricks);
  });
}

lrt_status lrt_save_scom2(const char* path, const void* header, const uint32_t* nodes, size_t n_nodes,
                          const uint32_t* bricks, size_t n_bricks) {
  return io_guarded([&]() -> lrt_status {
    if (!header || (!nodes && n_nodes) || (!bricks && n_bricks)) throw std::runtime_error("null argument");
    if (n_nodes > 0xFFFFFFFFull || n_bricks > 0xFFFFFFFFull) throw std::runtime_error("too many words for a u32 count");
    const uint32_t head[6] = {lrt::SCOM2_MAGIC_NUMBER, lrt::SCOM2_VERSION, uint32_t(n_nodes), uint32_t(n_bricks), 0, 0};
    File fl(path, "wb");
    put(fl, head, sizeof(head));
    put(fl, header, sizeof(lrt::SCom2Header));
    put(fl, nodes, n_nodes * 4);
    put(fl, bricks, n_bricks * 4);
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_sbs(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* values,
                            lrt_buffer* values_f) {
  return io_guarded([&]() -> lrt_status {
    if (!header || !nodes || !values || !values_f) throw std::runtime_error("null output");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    take(fl, header, sizeof(lrt::SdfSBSHeader), left, "brick-set header");
    std::vector<char> nd = counted(fl, sizeof(lrt::SdfSBSNode), left, "brick-set nodes");
    std::vector<char> vals = counted(fl, 4, left, "brick-set values");
    // files written before the fp32 block existed end here (the reference reads no count then)
    std::vector<char> vf;
    if (left > 0) vf = counted(fl, 4, left, "brick-set fp32 values");
    if (left != 0) throw std::runtime_error("brick-set file: " + std::to_string(left) + " bytes after the last block");
    to_buf(nd, nodes);
    to_buf(vals, values);
    return to_buf(vf, values_f);
  });
}

lrt_status lrt_save_sdf_sbs(const char* path, const void* header, const void* nodes, size_t n_nodes,
                            const uint32_t* values, size_t n_values, const float* values_f, size_t n_values_f) {
  return io_guarded([&]() -> lrt_status {
    if (!header || (!nodes && n_nodes) || (!values && n_values) || (!values_f && n_values_f))
      throw std::runtime_error("null input");
    check_u32(n_nodes, "brick-set nodes");
    check_u32(n_values, "brick-set values");
    check_u32(n_values_f, "brick-set fp32 values");
    File fl(path, "wb");
    put(fl, header, sizeof(lrt::SdfSBSHeader));
    uint32_t n = uint32_t(n_nodes);
    put(fl, &n, 4);
    put(fl, nodes, n_nodes * sizeof(lrt::SdfSBSNode));
    n = uint32_t(n_values);
    put(fl, &n, 4);
    put(fl, values, n_values * 4);
    n = uint32_t(n_values_f);
    put(fl, &n, 4);
    put(fl, values_f, n_values_f * 4);
    return LRT_OK;
  });
}

lrt_status lrt_load_sdf_multi_octree(const char* path, lrt_buffer* nodes, lrt_buffer* values) {
  return io_guarded([&]() -> lrt_status {
    if (!nodes || !values) throw std::runtime_error("null output buffer");
    File fl(path, "rb");
    size_t left = size_t(fl.size());
    uint32_t nn = 0, nv = 0;
    take(fl, &nn, 4, left, "multi-octree node count");
    take(fl, &nv, 4, left, "multi-octree value count");
    if (size_t(nn) * sizeof(lrt::SdfMultiOctreeNode) + size_t(nv) * 4 != left)
      throw std::runtime_error("multi-octree file: counts " + std::to_string(nn) + " / " + std::to_string(nv) +
                               " do not match the file size");
    std::vector<char> nd(size_t(nn) * sizeof(lrt::SdfMultiOctreeNode)), vals(size_t(nv) * 4);
    take(fl, nd.data(), nd.size(), left, "multi-octree nodes");
    take(fl, vals.data(), vals.size(), left, "multi-octree values");
    to_buf(nd, nodes);
    return to_buf(vals, values);
  });
}

lrt_status lrt_save_sdf_multi_octree(const char* path, const void* nodes, size_t n_nodes, const float* values,
                                     size_t n_values) {
  return io_guarded([&]() -> lrt_status {
    if ((!nodes && n_nodes) || (!values && n_values)) throw std::runtime_error("null input");
    check_u32(n_nodes, "multi-octree nodes");
    check_u32(n_values, "multi-octree values");
    File fl(path, "wb");
    uint32_t n = uint32_t(n_nodes), m = uint32_t(n_values);
    put(fl, &n, 4);
    put(fl, &m, 4);
    put(fl, nodes, n_nodes * sizeof(lrt::SdfMultiOctreeNode));
    put(fl, values, n_values * 4);
    return LRT_OK;
  });
}

}  // extern "C"
