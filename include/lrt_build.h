/*
 * lrt_build.h - C ABI of the scene-input builders that sit in front of the ray-query path:
 * the sparse-octree SDF builder and its converters (sdf/build, sdf/simple), the camera
 * matrices of MultiRenderer::UpdateCamera, and the synthetic meshes of the benchmark
 * workloads. Output arrays are library-allocated lrt_buffers released with lrt_buffer_free.
 */
#ifndef LRT_BUILD_H_
#define LRT_BUILD_H_

#include <stddef.h>
#include <stdint.h>

#include "lrt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lrt_buffer {
  void* data;
  size_t bytes;
} lrt_buffer;

/* Releases a buffer returned by this API (safe on a zeroed buffer). */
void lrt_buffer_free(lrt_buffer* b);

/* Camera of MultiRenderer::UpdateCamera (MultiRenderer/MultiRenderer_host.cpp:104-110) for
 * worldView = lookAt(eye, target, up) and proj = perspectiveMatrix(fovy_deg, aspect, znear, zfar)
 * (LiteMath.h:3315-3360, called as in render_CLI/offscreen_render.cpp:225-228): writes
 * inverse4x4(proj) and inverse4x4(worldView), column-major. */
lrt_status lrt_camera_inverse(const float* eye3, const float* target3, const float* up3, float fovy_deg, float aspect,
                              float znear, float zfar, float* proj_inv16, float* world_view_inv16);

/* Unit icosphere subdivided `subdiv` times (20*4^subdiv triangles), vertices displaced radially
 * by (1 + amplitude * noise(frequency * v)) with seeded 3-D value noise. vpos: float3 per
 * vertex, indices: uint32 triple per triangle. */
lrt_status lrt_gen_icosphere(uint32_t subdiv, float amplitude, float frequency, uint64_t seed, lrt_buffer* vpos,
                             lrt_buffer* indices);

/* Geodesic icosahedron of frequency n (20*n^2 triangles), displaced as above. */
lrt_status lrt_gen_geodesic(uint32_t n, float amplitude, float frequency, uint64_t seed, lrt_buffer* vpos,
                            lrt_buffer* indices);

/* The SDF generators below build the reference's global sparse octree first
 * (sdf_to_global_octree, sdf/build/sparse_octree_builder.cpp:106-264): threads >= 0 on the host
 * (0: all cores), threads < 0 on the current HIP device (same bytes; the converters that follow
 * run on the host). */

/* The global octree itself: nodes as uint32 triples (type, val_off, offset; global_octree.h
 * GlobalOctreeNode subset), values v^3 floats per node, v = brick_size + 2 brick_pad + 1. */
lrt_status lrt_sdf_spheres_global_octree(const float* spheres, size_t n_spheres, uint32_t depth, uint32_t brick_size,
                                         uint32_t brick_pad, int threads, lrt_buffer* nodes, lrt_buffer* values);

/* sdf_converter::create_sdf_frame_octree (sdf/simple/sdf_converter.cpp:265-273) for the union
 * of spheres (cx,cy,cz,r)[n_spheres]: sdf_to_global_octree with brick size 1, then
 * global_octree_to_frame_octree. nodes: SdfFrameOctreeNode array. */
lrt_status lrt_sdf_spheres_frame_octree(const float* spheres, size_t n_spheres, uint32_t depth, int threads,
                                        lrt_buffer* nodes);

/* sdf_to_global_octree (brick size 1) + global_octree_to_multi_octree (sdf/multi/sdf_multi.cpp:
 * 124-180) for the sphere union: nodes = SdfMultiOctreeNode array, values = 8 floats per node. */
lrt_status lrt_sdf_spheres_multi_octree(const float* spheres, size_t n_spheres, uint32_t depth, int threads,
                                        lrt_buffer* nodes, lrt_buffer* values);

/* sdf_converter::create_sdf_SBS (sdf/simple/sdf_converter.cpp:311-322) for the sphere union:
 * quantised DX / DX_UV16 / DX_RGB8 layouts. header: SdfSBSHeader (brick_size, brick_pad,
 * bytes_per_value, aux_data). colour_seed != 0 fills the 8 per-corner texture (u16|v16) or
 * colour (r8|g8|b8) words from a seeded smooth field; 0 leaves them zero, as the reference
 * converter does without a texture-coordinate channel. nodes: SdfSBSNode array, values: uint32. */
lrt_status lrt_sdf_spheres_sbs(const float* spheres, size_t n_spheres, uint32_t depth, const void* header,
                               uint64_t colour_seed, int threads, lrt_buffer* nodes, lrt_buffer* values);

/* Indexed fp32 sparse brick set (layout ID32F_IRGB32F) for the sphere union: lattice points
 * shared by bricks are stored once. values_f = [N distances][N albedo RGB] with a seeded
 * colour field; *n_distance_values = N. */
lrt_status lrt_sdf_spheres_sbs_indexed(const float* spheres, size_t n_spheres, uint32_t depth, uint32_t brick_size,
                                       uint64_t albedo_seed, int threads, lrt_buffer* nodes, lrt_buffer* values,
                                       lrt_buffer* values_f, uint32_t* n_distance_values);

/* Reference-format SDF scene files. Each loader checks the counts against the file size and fails
 * with LRT_ERR_INVALID_ARG and a message (the reference reads unchecked). Output arrays are
 * lrt_buffers; the scene structs (36-B SdfFrameOctreeNode, 16-B SdfSBSNode / SdfMultiOctreeNode,
 * 16-B SdfSBSHeader) are byte-identical to the reference's. */

/* load_sdf_frame_octree / save_sdf_frame_octree (sdf/simple/sdf_simple.cpp:49-66):
 * u32 count, then count SdfFrameOctreeNode. */
lrt_status lrt_load_sdf_frame_octree(const char* path, lrt_buffer* nodes);
lrt_status lrt_save_sdf_frame_octree(const char* path, const void* nodes, size_t n_nodes);

/* load_sdf_frame_octree_tex / save_sdf_frame_octree_tex (sdf/simple/sdf_simple.cpp:157-174):
 * u32 count, then count SdfFrameOctreeTexNode (104 B). */
lrt_status lrt_load_sdf_frame_octree_tex(const char* path, lrt_buffer* nodes);
lrt_status lrt_save_sdf_frame_octree_tex(const char* path, const void* nodes, size_t n_nodes);

/* load_sdf_grid / save_sdf_grid (sdf/simple/sdf_simple.cpp:20-36): 3 x u32 size, then
 * size.x * size.y * size.z floats (x fastest). */
lrt_status lrt_load_sdf_grid(const char* path, uint32_t size[3], lrt_buffer* data);
lrt_status lrt_save_sdf_grid(const char* path, const uint32_t size[3], const float* data);

/* load_coctree_v3 / save_coctree_v3 (sdf/scom/scom.cpp:86-111): u32 version (5 when written
 * here; another version is read all the same and returned in *version, as the reference only
 * warns), the 15-word COctreeV3Header, u32 count, then count u32 words. */
lrt_status lrt_load_coctree_v3(const char* path, uint32_t* version, uint32_t header[15], lrt_buffer* data);
lrt_status lrt_save_coctree_v3(const char* path, const uint32_t header[15], const uint32_t* data, size_t size);

/* load_sdf_DAG / save_sdf_DAG (sdf/scom2/scom_utils.cpp:5-73): u32 counts of nodes, child edges,
 * data edges, distances, point channels and voxel channels, the 64-B SdfDAGHeader, then the four
 * arrays (SdfDAGNode 24 B, SdfDAGChildEdge 8 B, SdfDAGDataEdge 16 B, float). Files with data
 * channels are refused (the ray query does not read them); the saver writes none. */
lrt_status lrt_load_sdf_dag(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* child_edges,
                            lrt_buffer* data_edges, lrt_buffer* distances);
lrt_status lrt_save_sdf_dag(const char* path, const void* header, const void* nodes, size_t n_nodes,
                            const void* child_edges, size_t n_child_edges, const void* data_edges, size_t n_data_edges,
                            const float* distances, size_t n_distances);

/* load_scom2 / save_scom2 (sdf/scom2/scom_utils.cpp:98-195): u32 magic 0xffffdefa, version 4,
 * counts of node words, brick words, voxel and point channels, the 216-B SCom2Header, then the
 * node and brick words. Legacy files without the magic number, other versions and files with
 * data channels are refused; the saver writes version 4 without channels. */
lrt_status lrt_load_scom2(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* bricks);
lrt_status lrt_save_scom2(const char* path, const void* header, const uint32_t* nodes, size_t n_nodes,
                          const uint32_t* bricks, size_t n_bricks);

/* load_sdf_SVS / save_sdf_SVS (sdf/simple/sdf_simple.cpp:69-94): u32 count, then count SdfSVSNode. */
lrt_status lrt_load_sdf_svs(const char* path, lrt_buffer* nodes);
lrt_status lrt_save_sdf_svs(const char* path, const void* nodes, size_t n_nodes);

/* load_sdf_SBS / save_sdf_SBS (sdf/simple/sdf_simple.cpp:96-142): SdfSBSHeader, u32 count +
 * SdfSBSNode[count], u32 count + uint32 values, u32 count + float values_f. A file that ends
 * after the uint32 values (written before the fp32 block existed, e.g. saves/test_10_sbs.bin)
 * loads with an empty values_f. header: 16 bytes written by the loader. */
lrt_status lrt_load_sdf_sbs(const char* path, void* header, lrt_buffer* nodes, lrt_buffer* values,
                            lrt_buffer* values_f);
lrt_status lrt_save_sdf_sbs(const char* path, const void* header, const void* nodes, size_t n_nodes,
                            const uint32_t* values, size_t n_values, const float* values_f, size_t n_values_f);

/* load_sdf_multi_octree / save_sdf_multi_octree (sdf/multi/sdf_multi.cpp:8-35): u32 node count,
 * u32 value count, SdfMultiOctreeNode[nodes], float[values]. */
lrt_status lrt_load_sdf_multi_octree(const char* path, lrt_buffer* nodes, lrt_buffer* values);
lrt_status lrt_save_sdf_multi_octree(const char* path, const void* nodes, size_t n_nodes, const float* values,
                                     size_t n_values);

#ifdef __cplusplus
}
#endif

#endif /* LRT_BUILD_H_ */
