// Host-visible launch interface of the gfx950 kernels (no torch types).
#pragma once
#include <hip/hip_runtime.h>
#include "scene_view.h"

namespace lrt {

// Queue of rays handed from pass 1 to the cooperative pass 2 (device memory).
struct TailQueue {
  uint32_t  count;   // rays queued by pass 1
  uint32_t  next;    // next queue entry pass 2 hands out
  uint32_t* list;    // capacity: rays of the launch
};

struct TraceArgs {
  const float4* pos;        // posAndNear per ray (device)
  const float4* dir;        // dirAndFar per ray (device)
  CRT_Hit*      out;        // closest hit per ray (device), may be null for any-hit
  uint32_t*     any_out;    // any-hit flag per ray (device), null for closest hit
  uint32_t      n;
  uint32_t*     overflow_count;  // device counter of rays that hit the stack limit (nullable)
  int           block;      // 64, 128 or 256
  int           lds_stack;  // 8, 16 or 32 stack entries kept in LDS
  uint32_t      xcd_remap;  // 1: contiguous ray range per XCD
  uint32_t*     counts;     // nullable: 6 counters per ray (RayCounts order), instrumented variant
  uint32_t*     spill;      // n * (STACK_SIZE - lds_stack) words: stack entries deeper than the LDS part
  uint32_t      budget;     // pass-1 steps per ray before it is queued for pass 2 (0: single pass)
  TailQueue*    queue;      // device queue (nullable when budget is 0)
  uint32_t      tail_waves; // persistent waves of pass 2 (0: default)
  uint32_t      postpone_tri; // 1: postpone triangle leaves (triangle-only scenes; pass 1)
  uint32_t      ray_groups;   // pass 1: lanes per group dealt to distant parts of the batch
                              // (power of two; 0 or 64: rays in thread order)
  uint32_t      persistent;   // 1: one persistent single-pass kernel with per-lane ray refill
  uint32_t*     pcount;       // its two counters (device, zero between calls)
};
// Rays beyond n that the work buffer's spill covers, so that a persistent grid (whole blocks of
// resident lanes, at most ceil(n / block) blocks) can index it by lane slot.
constexpr uint32_t SPILL_SLACK_RAYS = 256;

struct EyeRayParams {
  float    proj_inv[16];        // inverse4x4(proj), column-major
  float    world_view_inv[16];  // inverse4x4(worldView), column-major
  uint32_t width, height;
  uint32_t packed_w, packed_h;  // padded to multiples of 8 (MultiRenderer_host.cpp:80-81)
  float2   d;                   // sub-pixel offset when no jitter buffer is given
  uint32_t order;               // 0: row-major ray index, 1: SuperBlockIndex2DOpt order
};

// Differentiable brick-set rendering (one launch = one of: forward colour, backward with a
// given dL/dC, backward of the L2 loss against a target). The brick set uses an indexed fp32
// layout whose values_f slice [f_offset, f_offset + 4 n_dist) holds n_dist distances followed
// by n_dist RGB albedos; `params` replaces the scene's values_f for the launch.
struct SbsDiffArgs {
  const float4* pos;
  const float4* dir;
  uint32_t      n;
  const float*  params;       // device values_f (distances and albedo), scene-sized
  uint32_t      geom_id;      // the differentiable geometry (hits on others render 0, no gradient)
  uint32_t      f_offset;     // first values_f entry of the geometry
  uint32_t      n_dist;       // distance values of the geometry
  float4*       colour;       // forward: (r, g, b, hit) per ray
  const float4* dL_dcolour;   // backward: upstream gradient per ray (rgb used)
  const float4* target;       // fused L2 backward: target colour per ray
  float*        grad;         // backward: dL/dparams, accumulated (scene-sized)
  float*        loss;         // fused L2: sum over rays of |C - T|^2, accumulated (nullable)
  uint32_t*     counters;     // nullable: [0] += rays whose recomputed hit differed from the walk
  uint32_t*     spill;        // n * (STACK_SIZE - 16) words
  uint32_t      xcd_remap;
};

// One brick set's part of the occupancy mask: brick count, brick size, first mask word.
struct SbsMaskGeom {
  uint32_t sdf_id, n_bricks, brick_size, base;
};
// LDS bound of the one-wave-per-brick mask build (lattice flag bits; bricks up to 24^3 voxels)
constexpr size_t SBS_MASK_LDS_MAX = 2048;
// LDS words that build needs for bricks of size bs: one flag bit per lattice point, rounded up
// to whole waves.
inline uint32_t sbs_mask_flag_words(uint32_t bs) {
  const uint32_t np = (bs + 1) * (bs + 1) * (bs + 1);
  return (np + 63u) / 64u * 2u;
}
// Builds the occupancy masks of the listed brick sets from the view's current values.
hipError_t launch_sbs_masks(const SceneView& s, const SbsMaskGeom* geoms, uint32_t n_geoms, uint32_t* mask,
                            hipStream_t stream);

// One scene of a batched differentiable launch. Its view reads distances and albedo from the
// item's parameters and uses the item's own occupancy mask; the scene holds one brick set.
struct SbsBatchItem {
  SceneView   view;
  SbsDiffArgs args;      // pos, dir, n, params, geom_id, f_offset, n_dist, colour, target, grad, loss, counters
  uint32_t*   mask;      // the item's occupancy mask (1 + ceil(bs^3/32) words per brick)
  uint32_t    n_bricks, brick_size;
  uint64_t    n_params;  // gradient entries zeroed by every run
};

// A batched run: zero gradients and losses, build every item's occupancy mask from its
// parameters (set-up kernels on 2D grids, blockIdx.y = item), then one persistent kernel whose
// one-wave blocks take SBS_BATCH_CHUNK-ray chunks from a global counter (items in order) and run
// the forward colour and the L2 adjoint of each ray. Only kernel launches (no allocation, copy
// or memset), so a run can be captured in a graph.
struct SbsBatchLaunch {
  const SbsBatchItem* items;
  uint32_t n_items;
  const uint32_t* chunk_begin;   // [n_items + 1] prefix sums of ray chunks (SBS_BATCH_CHUNK rays)
  uint32_t total_bricks, total_chunks;
  uint32_t zero_blocks, mask_blocks, summary_blocks, brick_blocks;   // x extents of the set-up grids
  uint32_t mask_flag_bytes;      // LDS of the one-wave-per-brick mask build (0: bricks too large)
  uint32_t* counter;             // work counter (reset by the zero kernel)
  uint32_t* spill;               // grid * BLOCK * (STACK_SIZE - 8) words
  uint32_t grid;                 // persistent blocks (at most the resident blocks of the variant)
  int lds_stack;                 // stack entries in LDS: 8, 16 or 32
  int record;                    // 1: the walk records the hit voxel; 0: the hit brick is marched again
};
constexpr int SBS_BATCH_BLOCK = 64;
constexpr int SBS_BATCH_CHUNK = 64;   // rays per work item of the persistent batch kernel (one wave)
hipError_t launch_sbs_batch(const SbsBatchLaunch& L, hipStream_t stream);
// Resident blocks per CU of the persistent batch kernel variant (occupancy query).
hipError_t sbs_batch_blocks_per_cu(int lds_stack, int record, int* blocks);

// Light (core/IRenderer.h:72-78), 32 B
struct ResolveLight {
  float    space[3];   // direction (direct light, normalised) or position (point light)
  uint32_t type;       // LIGHT_TYPE_*
  float    color[3];
  uint32_t pad;
};
constexpr uint32_t RESOLVE_MAX_LIGHTS = 8;

// One g-buffer resolve launch (MultiRenderer::kernel1D_ResolveCommon / kernel1D_ResolveDebug).
struct ResolveParams {
  float    proj_inv[16], world_view_inv[16];   // camera of the eye rays (column-major)
  uint32_t width, height, packed_w, packed_h;
  uint32_t count;          // width * height pixels
  uint32_t order;          // 0: hit of pixel i at i; 1: at SuperBlockIndex2DOpt(x, y)
  uint32_t render_mode, normal_mode, spp, aa_frame, sample_id;
  uint32_t n_lights;
  ResolveLight lights[RESOLVE_MAX_LIGHTS];
  float    background[4];
  uint32_t highlight_mode, highlight_prim, highlight_geom, highlight_inst;
  const GeomData* geoms;
  const float4*   vert_pos;
  const float4*   vert_norm;
  const uint32_t* orig_indices;   // triangle indices in the caller's order (local vertex ids)
  const float*    inst_inv_t;     // per instance transpose(inverse(transform)), float4x4 layout
};
hipError_t launch_resolve(const ResolveParams& p, const CRT_Hit* gbuf, float4* out, hipStream_t stream);
// Tonemap: rgba8[i] = pack(255 * clamp(colour[i] / samples, 0, 1))
hipError_t launch_tonemap(const float4* colour, uint32_t n, uint32_t samples, uint32_t* out, hipStream_t stream);

// Refit of a TLAS in place: leaf boxes from the instances' boxes, every parent the union of its
// two children; parent[i] = node index of node i's parent (0xFFFFFFFF for the root's children);
// counters: n_nodes words, cleared by the launch.
hipError_t launch_tlas_refit(BVHNode* tlas, uint32_t n_nodes, const uint32_t* parent, const InstanceData* inst,
                             uint32_t* counters, hipStream_t stream);

hipError_t launch_trace(const SceneView& s, const TraceArgs& a, hipStream_t stream);
// launch_trace with the library's fixed per-device work array as queue and spill (a.queue and
// a.spill are ignored): at most LRT_NOALLOC_MAX_RAYS rays, no counters; `device` is current.
hipError_t launch_trace_noalloc(const SceneView& s, TraceArgs a, int device, hipStream_t stream);
hipError_t launch_sbs_diff(const SceneView& s, const SbsDiffArgs& a, hipStream_t stream);
hipError_t launch_eye_rays(const EyeRayParams& prm, const float2* jitter, float4* pos, float4* dir,
                           hipStream_t stream);

}  // namespace lrt
