// G-buffer resolve and tonemap on the GPU: MultiRenderer::kernel1D_ResolveCommon /
// kernel1D_ResolveDebug / kernel1D_Tonemap (MultiRenderer/MultiRenderer_resolve.cpp:529-1005)
// over the CRT_Hit records a trace call wrote, one pixel per thread. Only +,-,*,/ and sqrt on
// the lighting paths (compiled without contraction, like the reference's float code), so the
// colours equal the CPU restatement's bit for bit; MULTI_RENDER_MODE_NORMAL calls powf.
#include "kernels.h"
#include "lrt_math.h"

namespace lrt {

namespace {

// MultiRenderer.h:270-284
__constant__ uint32_t c_palette[20] = {
    0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0, 0xfff032e6, 0xffd2f53c,
    0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8, 0xff800000, 0xffaaffc3, 0xff808000, 0xffffd8b1,
    0xff000080, 0xff808080};
__constant__ uint32_t c_lod_palette[20] = {
    0xff000000, 0xff000040, 0xff000080, 0xff0000c0, 0xff0040ff, 0xff0080ff, 0xff00c0ff, 0xff00ffff, 0xff00ffc0,
    0xff00ff80, 0xff00ff40, 0xff00ff00, 0xff00c000, 0xff008000, 0xff004000, 0xff000000, 0xffff0000, 0xffff0000,
    0xffff0000, 0xffff0000};
// HamiltonSequence of MultiRenderer::initEyeRay (MultiRenderer.cpp:145-161)
__constant__ float c_hamilton[32] = {0.500000f, 0.333333f, 0.250000f, 0.666667f, 0.750000f, 0.111111f, 0.125000f,
                                     0.444444f, 0.625000f, 0.777778f, 0.375000f, 0.222222f, 0.875000f, 0.555556f,
                                     0.062500f, 0.888889f, 0.562500f, 0.037037f, 0.312500f, 0.370370f, 0.812500f,
                                     0.703704f, 0.187500f, 0.148148f, 0.687500f, 0.481481f, 0.437500f, 0.814815f,
                                     0.937500f, 0.259259f, 0.031250f, 0.592593f};

struct C4 {
  float x, y, z, w;
};
__device__ __forceinline__ C4 c4(float x, float y, float z, float w) { return C4{x, y, z, w}; }
__device__ __forceinline__ C4 operator*(float a, C4 b) { return c4(a * b.x, a * b.y, a * b.z, a * b.w); }
__device__ __forceinline__ C4 operator+(C4 a, C4 b) { return c4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ C4 operator/(C4 a, float b) { return c4(a.x / b, a.y / b, a.z / b, a.w / b); }

// decode_RGBA8 (utils/raytracing/render_common.h:250-254)
__device__ __forceinline__ C4 decode_rgba8(uint32_t c) {
  return c4(float(c & 0xFF) * (1.0f / 255.0f), float((c >> 8) & 0xFF) * (1.0f / 255.0f),
            float((c >> 16) & 0xFF) * (1.0f / 255.0f), float((c >> 24) & 0xFF) * (1.0f / 255.0f));
}

// decode_normal (utils/raytracing/render_common.h:256-267)
__device__ __forceinline__ float3 decode_normal(float ex, float ey) {
  float3 v = make3(ex, ey, 1.0f - fabs_ref(ex) - fabs_ref(ey));
  if (v.z < 0) {
    const float vx = v.x;
    v.x = (1.0f - fabs_ref(v.y)) * ((v.x >= 0.0f) ? +1.0f : -1.0f);
    v.y = (1.0f - fabs_ref(vx)) * ((v.y >= 0.0f) ? +1.0f : -1.0f);
  }
  return normalize3(v);
}

// MultiRenderer::initEyeRay (MultiRenderer.cpp:142-183), cutting plane disabled: the direction
__device__ __forceinline__ void eye_ray(const ResolveParams& P, uint32_t x, uint32_t y, float3& o, float3& rd) {
  float2 d;
  if (P.aa_frame >= 2) {
    d.x = c_hamilton[2 * (P.aa_frame % 16)];
    d.y = c_hamilton[2 * (P.aa_frame % 16) + 1];
  } else {
    const uint32_t spp_sqrt = uint32_t(sqrtf(float(P.spp)));
    const float i_spp_sqrt = 1.0f / float(spp_sqrt);
    d.x = i_spp_sqrt * 0.5f;   // sample 0: float2(0 / spp_sqrt + 0.5, 0 % spp_sqrt + 0.5)
    d.y = i_spp_sqrt * 0.5f;
  }
  const float fx = (float(x) + d.x) / float(P.packed_w);
  const float fy = (float(y) + d.y) / float(P.packed_h);
  float4 p4 = make4(2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f);
  p4 = mat_mul4(P.proj_inv, p4);
  const float w = p4.w;
  p4 = make4(p4.x / w, p4.y / w, p4.z / w, p4.w / w);
  rd = normalize3(xyz(p4));
  const float3 zero = make3(0, 0, 0);
  o = matmul4x3(P.world_view_inv, zero);
  const float3 o2 = matmul4x3(P.world_view_inv, zero + 100.0f * rd);
  rd = normalize3(o2 - o);
}

// MultiRenderer::loadNormalTcFromGbuffer (MultiRenderer_resolve.cpp:27-75). Meshes carry no
// texture coordinates through this library's ABI, so tc is 0 for triangles.
__device__ __forceinline__ void load_normal_tc(const ResolveParams& P, const CRT_Hit& hit, float3 ray_dir,
                                               float3& norm, float2& tc) {
  const uint32_t type = hit.geomId >> SH_TYPE;
  const uint32_t geomId = hit.geomId & GEOM_ID_MASK;
  if (type == TYPE_MESH_TRIANGLE) {
    const uint2 go = P.geoms[geomId].offset;
    const uint32_t A = P.orig_indices[go.x + hit.primId * 3 + 0];
    const uint32_t B = P.orig_indices[go.x + hit.primId * 3 + 1];
    const uint32_t C = P.orig_indices[go.x + hit.primId * 3 + 2];
    float3 raw = make3(1, 0, 0);
    if (P.normal_mode == NORMAL_MODE_GEOMETRY) {
      const float3 pA = xyz(P.vert_pos[go.y + A]), pB = xyz(P.vert_pos[go.y + B]), pC = xyz(P.vert_pos[go.y + C]);
      raw = cross3(pB - pA, pC - pA);
    } else if (P.normal_mode == NORMAL_MODE_VERTEX) {
      const float4 nA = P.vert_norm[go.y + A], nB = P.vert_norm[go.y + B], nC = P.vert_norm[go.y + C];
      const float u = hit.coords[0], v = hit.coords[1];
      const float wa = 1.0f - u - v;
      // (1 - u - v) * normA + v * normB + u * normC, component by component, left to right
      raw = make3(wa * nA.x + v * nB.x + u * nC.x, wa * nA.y + v * nB.y + u * nC.y, wa * nA.z + v * nB.z + u * nC.z);
    }
    tc.x = 0.0f;
    tc.y = 0.0f;
    norm = normalize3(matmul4x3(P.inst_inv_t + 16 * size_t(hit.instId), raw));
    const float s = sign_ref(dot3(-1.0f * ray_dir, norm));
    norm = norm * s;
  } else {
    tc.x = hit.coords[0];
    tc.y = hit.coords[1];
    norm = decode_normal(hit.coords[2], hit.coords[3]);
  }
}

__device__ __forceinline__ float3 lights_lambert(const ResolveParams& P, float3 colour, float3 norm, float3 ray_pos,
                                                 float3 ray_dir, float t) {
  float3 fin = make3(0, 0, 0);
  for (uint32_t i = 0; i < P.n_lights; i++) {
    const ResolveLight L = P.lights[i];
    const float3 lc = make3(L.color[0], L.color[1], L.color[2]);
    const float3 ls = make3(L.space[0], L.space[1], L.space[2]);
    if (L.type == LIGHT_TYPE_DIRECT) {
      const float q = fmax_ref(0.0f, dot3(norm, ls));
      fin = fin + lc * colour * q;
    } else if (L.type == LIGHT_TYPE_POINT) {
      const float3 surf = ray_pos + t * ray_dir;
      float3 dir = ls - surf;
      const float l = length3(dir);
      dir = dir / l;
      const float q = fmax_ref(0.0f, dot3(norm, dir));
      fin = fin + lc * colour * q / (l * l);
    } else {
      fin = fin + lc * colour;
    }
  }
  return fin;
}

__global__ void k_resolve(ResolveParams P, const CRT_Hit* __restrict__ gbuf, float4* __restrict__ out) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= P.count) return;
  // the g-buffer of FillGbuffer is row-major; a trace over rays in the packed super-block order
  // (lrt_eye_rays_dev order 1) holds pixel (x, y) at super_block_index(x, y)
  uint32_t slot = tid;
  if (P.order == 1) {
    const uint32_t x = tid % P.width, y = tid / P.width;
    const uint32_t local = (y & 3u) * 4 + (x & 3u), bx = x >> 2, by = y >> 2;
    const uint32_t localH = (by & 1u) * 2 + (bx & 1u), wbH = (P.packed_w >> 2) >> 1;
    slot = ((bx >> 1) + (by >> 1) * wbH) * 64 + localH * 16 + local;
  }
  CRT_Hit hit;
  {
    const float4* h4 = reinterpret_cast<const float4*>(gbuf + slot);
    const float4 a = h4[0], b = h4[1];
    hit.t = a.x;
    hit.primId = __float_as_uint(a.y);
    hit.instId = __float_as_uint(a.z);
    hit.geomId = __float_as_uint(a.w);
    hit.coords[0] = b.x; hit.coords[1] = b.y; hit.coords[2] = b.z; hit.coords[3] = b.w;
  }
  C4 res = c4(1, 0, 1, 1);
  if (hit.primId == 0xFFFFFFFFu) {
    res = c4(P.background[0], P.background[1], P.background[2], P.aa_frame == 0 ? P.background[3] : -1.0f);
  } else {
    const uint32_t type = hit.geomId >> SH_TYPE;
    const uint32_t geomId = hit.geomId & GEOM_ID_MASK;
    float3 ray_pos, ray_dir;
    // the reference regenerates the ray from the packed width (initEyeRay(tidX % packed_w, ...))
    eye_ray(P, tid % P.packed_w, tid / P.packed_w, ray_pos, ray_dir);
    float3 norm = make3(1, 0, 0);
    float2 tc;
    tc.x = 0.0f;
    tc.y = 0.0f;
    load_normal_tc(P, hit, ray_dir, norm, tc);
    const uint32_t m = P.render_mode;
    if (m == MULTI_RENDER_MODE_LAMBERT_NO_TEX || m == MULTI_RENDER_MODE_GEOM || m == MULTI_RENDER_MODE_LOD ||
        m == MULTI_RENDER_MODE_LAMBERT) {
      // ResolveCommon (MultiRenderer_resolve.cpp:617-661); LAMBERT only for per-voxel colour types
      float3 colour = make3(0, 0, 1);
      if (m == MULTI_RENDER_MODE_LAMBERT_NO_TEX) {
        colour = make3(1, 1, 1);
      } else if (m == MULTI_RENDER_MODE_LOD) {
        const C4 c = decode_rgba8(c_lod_palette[hit.primId % 20u]);
        colour = make3(c.x, c.y, c.z);
      } else if (m == MULTI_RENDER_MODE_GEOM) {
        const C4 c = decode_rgba8(c_palette[geomId % 20u]);
        colour = make3(c.x, c.y, c.z);
      } else {
        // std::round: half away from zero, as roundf
        colour.x = roundf(hit.coords[0]) / 255.0f;
        colour.y = hit.coords[0] - floorf(hit.coords[0]);
        colour.z = roundf(hit.coords[1]) / 255.0f;
      }
      const float3 f = lights_lambert(P, colour, norm, ray_pos, ray_dir, hit.t);
      res = c4(f.x, f.y, f.z, 1.0f);
    } else {
      // ResolveDebug (MultiRenderer_resolve.cpp:808-1005)
      const float z = hit.t, z_near = 0.1f, z_far = 10.0f;
      switch (m) {
        case MULTI_RENDER_MODE_MASK: res = c4(1, 1, 1, 1); break;
        case MULTI_RENDER_MODE_DEPTH: {
          const float d = (1.0f / z - 1.0f / z_near) / (1.0f / z_far - 1.0f / z_near);
          res = c4(d, d, d, 1);
        } break;
        case MULTI_RENDER_MODE_LINEAR_DEPTH: {
          const float d = ((z - z_near) / (z_far - z_near));
          res = c4(d, d, d, 1.0f);
        } break;
        case MULTI_RENDER_MODE_INVERSE_LINEAR_DEPTH: {
          const float d = 1 - ((z - z_near) / (z_far - z_near));
          res = c4(d, d, d, 1);
        } break;
        case MULTI_RENDER_MODE_HSV_DEPTH: {
          const float3 ip = ray_pos + z * ray_dir;
          float zt = (ip.z + 1.f) * 2.f;
          zt = zt - float(int(zt));
          if (zt < 0.f) zt = 1.f + zt;
          const float hx = zt, hy = 1.f, hz = 1.f;
          float rc[4] = {hz, hz, hz, 1.f};
          const float vmin = (1.f - hy) * hz;
          const float hi = truncf(hx * 6.f);
          const float a = (hz - vmin) * (hx * 6.f - hi);   // std::modf: fractional part, exact
          uint32_t hii = uint32_t(hi);
          hii = hii <= 5u ? hii : 5u;
          rc[(2u + (hii >> 1)) % 3u] = vmin;
          if ((hii & 1u) != 0u)
            rc[hii >> 1] = hz - a;
          else
            rc[(1u + (hii >> 1)) % 3u] = vmin + a;
          res = c4(rc[0], rc[1], rc[2], 1.f);
        } break;
        case MULTI_RENDER_MODE_PRIMITIVE: res = decode_rgba8(c_palette[hit.primId % 20u]); break;
        case MULTI_RENDER_MODE_TYPE: res = decode_rgba8(c_palette[type % 20u]); break;
        case MULTI_RENDER_MODE_NORMAL: {
          const float vx = fabs_ref(norm.x), vy = fabs_ref(norm.y), vz = fabs_ref(norm.z);
          res = c4(0.25f, 0.25f, 0.25f, 0.5f) +
                0.5f * c4(powf(vx, 1.0f / 1.5f), powf(vy, 1.0f / 1.5f), powf(vz, 1.0f / 1.5f), 1.0f);
        } break;
        case MULTI_RENDER_MODE_BARYCENTRIC:
          res = c4(hit.coords[0], hit.coords[1], 1 - hit.coords[0] - hit.coords[1], 1);
          break;
        case MULTI_RENDER_MODE_ST_ITERATIONS:
          if (hit.primId == 0) {
            res = c4(0.5f, 0.5f, 0.5f, 1.0f);
          } else if (hit.primId <= 64) {
            const float q = clamp_ref(float(hit.primId) / 64, 0.0f, 1.0f);
            res = q * c4(0, 0, 1, 1) + (1.0f - q) * c4(0, 1, 0, 1);
          } else {
            const float q = clamp_ref(float(hit.primId - 64) / (256 - 64), 0.0f, 1.0f);
            res = q * c4(1, 0, 0, 1) + (1.0f - q) * c4(0, 0, 1, 1);
          }
          break;
        case MULTI_RENDER_MODE_TEX_COORDS: res = c4(tc.x, tc.y, 0, 1); break;
        case MULTI_RENDER_MODE_MO_VOXEL_COUNT:
          if (hit.primId == 0)
            res = c4(0.5, 0.5, 0.5, 1);
          else if (hit.primId == 1)
            res = c4(0.5, 1, 0.5, 1);
          else if (hit.primId == 2)
            res = c4(0.5, 0.5, 1, 1);
          else if (hit.primId <= 64) {
            const float q = clamp_ref(float(hit.primId) / 64, 0.0f, 1.0f);
            res = q * c4(1, 0, 0, 1) + (1 - q) * c4(0.5, 0.5, 1, 1);
          } else {
            res = c4(1, 0, 0, 1);
          }
          break;
        case MULTI_RENDER_MODE_MO_FLAGS: res = decode_rgba8(c_palette[hit.primId % 20u]); break;
        case MULTI_RENDER_MODE_INTERSECTION_COUNT:
          if (hit.instId == 0)
            res = c4(0.3, 0.3, 0.3, 1);
          else if (hit.instId == 1)
            res = c4(0, 100, 0, 255) / 255.0f;
          else if (hit.instId == 2)
            res = c4(0, 200, 0, 255) / 255.0f;
          else if (hit.instId == 3)
            res = c4(50, 250, 50, 255) / 255.0f;
          else if (hit.instId == 4)
            res = c4(150, 250, 150, 255) / 255.0f;
          else if (hit.instId == 5)
            res = c4(250, 250, 250, 255) / 255.0f;
          else if (hit.instId <= 20)
            res = c4(250, float(1000 - 50 * hit.instId) * 0.333f, float(1000 - 50 * hit.instId) * 0.333f, 255) / 255.0f;
          else
            res = c4(1, 0, fmin_ref(1.0f, 0.025f * float(hit.instId - 20)), 1);
          break;
        default: break;
      }
    }
    // applyHighlight (MultiRenderer_resolve.cpp:10-25)
    if (P.highlight_mode == 1 && P.highlight_prim == hit.primId && P.highlight_geom == hit.geomId &&
        P.highlight_inst == hit.instId)
      res = res + c4(0.5f, 0, 0, 0);
    res.w = P.aa_frame == 0 ? 1.0f : hit.t;
  }
  float4* o = out + tid;
  if (P.sample_id == 0) {
    *o = make4(res.x, res.y, res.z, res.w);
  } else {
    const float4 a = *o;
    *o = make4(a.x + res.x, a.y + res.y, a.z + res.z, a.w + res.w);
  }
}

// MultiRenderer::Tonemap (MultiRenderer_resolve.cpp:538-544)
__global__ void k_tonemap(const float4* __restrict__ colour, uint32_t n, uint32_t samples, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 c = colour[i];
  const float s = float(samples);
  const float4 f = make4(c.x / s, c.y / s, c.z / s, c.w / s);
  const uint32_t x = uint32_t(255 * clamp_ref(f.x, 0.0f, 1.0f)), y = uint32_t(255 * clamp_ref(f.y, 0.0f, 1.0f));
  const uint32_t z = uint32_t(255 * clamp_ref(f.z, 0.0f, 1.0f)), w = uint32_t(255 * clamp_ref(f.w, 0.0f, 1.0f));
  out[i] = (w << 24) | (z << 16) | (y << 8) | x;
}

}  // namespace

hipError_t launch_resolve(const ResolveParams& p, const CRT_Hit* gbuf, float4* out, hipStream_t stream) {
  if (p.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_resolve, dim3((p.count + 255) / 256), dim3(256), 0, stream, p, gbuf, out);
  return hipGetLastError();
}

hipError_t launch_tonemap(const float4* colour, uint32_t n, uint32_t samples, uint32_t* out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, stream, colour, n, samples, out);
  return hipGetLastError();
}

}  // namespace lrt
