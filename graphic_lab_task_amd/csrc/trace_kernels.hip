// Trace-call dispatch on the scene's leaf kinds, the queue reset, the allocation-free trace and
// the primary-ray generator.
#include <mutex>

#include "../../include/lrt.h"
#include "trace_launch.h"

namespace lrt {

// Work memory of the allocation-free trace: a 256-B header (tail queue at word 0, the persistent
// kernel's counters at word 16), the queue (one word per ray) and the deep-stack spill
// (STACK_SIZE - 8 words per ray, enough for every LDS stack depth, plus the persistent grid's
// slack), for at most LRT_NOALLOC_MAX_RAYS rays. Host code never uses this name as an address: each
// device's copy is resolved once with hipGetSymbolAddress and that pointer is passed on.
constexpr size_t NOALLOC_HEAD_WORDS = 64;
constexpr size_t NOALLOC_WORDS =
    NOALLOC_HEAD_WORDS + LRT_NOALLOC_MAX_RAYS + size_t(LRT_NOALLOC_MAX_RAYS + SPILL_SLACK_RAYS) * (STACK_SIZE - 8);
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

// Sets the queue header at the start of every allocation-free call (a kernel, not a memset, so a
// captured call records kernels only).
__global__ void k_queue_init(TailQueue* queue, uint32_t* list, uint32_t* pcount) {
  if (threadIdx.x == 0) {
    queue->count = 0;
    queue->next = 0;
    queue->list = list;
    pcount[0] = 0;
    pcount[1] = 0;
  }
}

// Clears the tail queue when a call's second pass could not be launched (a kernel rather than a
// memset, so that nothing but kernels enters a graph captured from the stream).
__global__ void k_queue_reset(TailQueue* queue) {
  if (threadIdx.x == 0) {
    queue->count = 0;
    queue->next = 0;
  }
}

// SuperBlockIndex2DOpt (utils/raytracing/render_common.h:225-242): 4x4 blocks in 2x2 super blocks.
__device__ __forceinline__ uint32_t super_block_index(uint32_t x, uint32_t y, uint32_t w) {
  const uint32_t local = (y & 3u) * 4 + (x & 3u);
  const uint32_t bx = x >> 2, by = y >> 2;
  const uint32_t localH = (by & 1u) * 2 + (bx & 1u);
  const uint32_t wbH = (w >> 2) >> 1;
  return ((bx >> 1) + (by >> 1) * wbH) * 64 + localH * 16 + local;
}

// MultiRenderer::initEyeRay (MultiRenderer/MultiRenderer.cpp:142-183) with the cutting plane
// disabled: EyeRayDirNormalized + transform_ray3f (utils/raytracing/render_common.h:25-55).
__global__ void k_eye_rays(EyeRayParams prm, const float2* __restrict__ jitter, float4* __restrict__ pos,
                           float4* __restrict__ dir) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= prm.width || y >= prm.height) return;
  const uint32_t pix = y * prm.width + x;
  float2 dd = prm.d;
  if (jitter) dd = jitter[pix];
  const float fx = (float(x) + dd.x) / float(prm.packed_w);
  const float fy = (float(y) + dd.y) / float(prm.packed_h);
  float4 p4 = make4(2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f);
  p4 = mat_mul4(prm.proj_inv, p4);
  const float w = p4.w;
  p4 = make4(p4.x / w, p4.y / w, p4.z / w, p4.w / w);
  float3 rd = normalize3(xyz(p4));
  const float3 zero = make3(0, 0, 0);
  const float3 o = matmul4x3(prm.world_view_inv, zero);
  const float3 o2 = matmul4x3(prm.world_view_inv, zero + 100.0f * rd);
  rd = normalize3(o2 - o);
  const uint32_t idx = prm.order == 1 ? super_block_index(x, y, prm.packed_w) : pix;
  pos[idx] = make4(o.x, o.y, o.z, 1e-9f);
  dir[idx] = make4(rd.x, rd.y, rd.z, 1e9f);
}

hipError_t launch_trace(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  const uint32_t k = s.leaf_kinds & ~(LK_SBS_SMOOTH | LK_TRICUBIC);
  hipError_t e;
  if (s.leaf_kinds & LK_TRICUBIC)
    e = k == LK_SBS ? launch_trace_sbs_tricubic(s, a, stream) : launch_trace_mixed_tricubic(s, a, stream);
  else if (s.leaf_kinds & LK_SBS_SMOOTH)
    e = k == LK_SBS ? launch_trace_sbs_smooth(s, a, stream) : launch_trace_mixed_smooth(s, a, stream);
  else if (k == LK_TRI)
    e = a.postpone_tri && !a.counts ? launch_trace_tri_postponed(s, a, stream) : launch_trace_tri(s, a, stream);
  else if (k == LK_FRAME)
    e = launch_trace_frame(s, a, stream);
  else if (k == LK_SBS)
    e = launch_trace_sbs(s, a, stream);
  else if (k == LK_MULTI)
    e = launch_trace_multi(s, a, stream);
  else if (k == LK_COCT)
    e = launch_trace_coct(s, a, stream);
  else if (k == LK_DAG)
    e = launch_trace_dag(s, a, stream);
  else if (k == LK_SCOM2)
    e = launch_trace_scom2(s, a, stream);
  else
    e = launch_trace_mixed(s, a, stream);
  // a call whose second pass did not launch must not leave queued rays to the next call
  if (e != hipSuccess && a.budget != 0 && a.queue != nullptr && a.counts == nullptr) {
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, stream, a.queue);
    (void)hipGetLastError();
  }
  return e;
}

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch(int device, uint32_t** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // the caller made `device` current; its copy of the array has its own address
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

hipError_t launch_trace_noalloc(const SceneView& s, TraceArgs a, int device, hipStream_t stream) {
  if (a.n > LRT_NOALLOC_MAX_RAYS || a.counts != nullptr) return hipErrorInvalidValue;
  uint32_t* base = nullptr;
  hipError_t e = noalloc_scratch(device, &base);
  if (e != hipSuccess) return e;
  a.queue = reinterpret_cast<TailQueue*>(base);
  a.pcount = base + 16;
  uint32_t* list = base + NOALLOC_HEAD_WORDS;
  a.spill = list + LRT_NOALLOC_MAX_RAYS;
  hipLaunchKernelGGL(k_queue_init, dim3(1), dim3(64), 0, stream, a.queue, list, a.pcount);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_trace(s, a, stream);
}

hipError_t launch_eye_rays(const EyeRayParams& prm, const float2* jitter, float4* pos, float4* dir,
                           hipStream_t stream) {
  dim3 block(16, 16);
  dim3 grid((prm.width + 15) / 16, (prm.height + 15) / 16);
  hipLaunchKernelGGL(k_eye_rays, grid, block, 0, stream, prm, jitter, pos, dir);
  return hipGetLastError();
}

}  // namespace lrt
