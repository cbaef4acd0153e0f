// gfx950 device code for the closest-hit ray query: TLAS walk, fat-node BLAS walk and
// the leaf intersectors. One ray per lane. Arithmetic follows the reference's operation
// order (compiled with -ffp-contract=off); see the file:line citations per function.
#pragma once
#include <hip/hip_runtime.h>
#include "lrt_math.h"
#include "scene_view.h"
#include "scom2_unpack.h"

namespace lrt {

// ---------------------------------------------------------------------------------------
// Traversal stack: the first S entries live in LDS (entry-major, one column per lane, so
// lanes never conflict on a bank), deeper entries go to this ray's slot of a global spill
// buffer (only rays deeper than S touch it). Depth limit STACK_SIZE=80 as in
// raytrace_common.h:36; pushes beyond it are dropped and counted instead of writing out of
// bounds. The LDS column is an explicit address-space-3 pointer so that every access is a
// ds_read/ds_write rather than a flat access.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

template <int S, int BLOCK>
struct LaneStack {
  lds_u32* col;        // entry i at col[i * BLOCK]
  uint32_t* spill;     // entry i >= S at spill[i - S]
  __device__ __forceinline__ void put(int i, uint32_t v) {
    if (i < S) col[i * BLOCK] = v; else spill[i - S] = v;
  }
  __device__ __forceinline__ uint32_t get(int i) const {
    if (i < S) return col[i * BLOCK];
    return spill[i - S];
  }
};

// Per-ray traversal counters of the instrumented kernel variant (the reference's RayInfo
// counters, BVH2Common.cpp:3403-3571): fat nodes fetched, leaves entered, triangles tested,
// TLAS nodes fetched, instances entered, SBS voxels stepped.
struct RayCounts {
  uint32_t blas_nodes, leaves, prims, tlas_nodes, instances, sdf_voxels;
};
// Words per ray written by the instrumented variant: the 6 counters, then the device wall
// clock (100 MHz ticks) at ray start (low 32 bits) and the ray's elapsed ticks.
constexpr int RAY_COUNTERS = 8;

// The fields of a geometry record that the leaf intersectors read, loaded once per BLAS walk
// rather than once per leaf.
struct GeomRef {
  uint32_t type;
  uint2    offset;
};

__device__ __forceinline__ GeomRef geom_ref(const SceneView& s, uint32_t geomId) {
  const GeomData* g = s.geoms + geomId;
  return GeomRef{g->type, g->offset};
}

__device__ __forceinline__ BVHNodePair load_pair(const BVHNodePair* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1], c = q[2], d = q[3];
  BVHNodePair r;
  r.left.boxMin = make3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
  r.left.leftOffset = a.w;
  r.left.boxMax = make3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
  r.left.escapeIndex = b.w;
  r.right.boxMin = make3(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z));
  r.right.leftOffset = c.w;
  r.right.boxMax = make3(__uint_as_float(d.x), __uint_as_float(d.y), __uint_as_float(d.z));
  r.right.escapeIndex = d.w;
  return r;
}

typedef __attribute__((address_space(3))) BVHNodePair lds_pair;
typedef __attribute__((address_space(3))) const uint32_t lds_cu32;

__device__ __forceinline__ BVHNodePair load_pair_lds(const lds_pair* p) {
  lds_cu32* q = reinterpret_cast<lds_cu32*>(p);
  uint32_t w[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) w[k] = q[k];
  BVHNodePair r;
  r.left.boxMin = make3(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]));
  r.left.leftOffset = w[3];
  r.left.boxMax = make3(__uint_as_float(w[4]), __uint_as_float(w[5]), __uint_as_float(w[6]));
  r.left.escapeIndex = w[7];
  r.right.boxMin = make3(__uint_as_float(w[8]), __uint_as_float(w[9]), __uint_as_float(w[10]));
  r.right.leftOffset = w[11];
  r.right.boxMax = make3(__uint_as_float(w[12]), __uint_as_float(w[13]), __uint_as_float(w[14]));
  r.right.escapeIndex = w[15];
  return r;
}

__device__ __forceinline__ BVHNode load_node(const BVHNode* p) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], b = q[1];
  BVHNode r;
  r.boxMin = make3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
  r.leftOffset = a.w;
  r.boxMax = make3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z));
  r.escapeIndex = b.w;
  return r;
}

__device__ __forceinline__ void load_mat(const float4x4& m, float* out) {
  const float4* q = reinterpret_cast<const float4*>(&m);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float4 v = q[i];
    out[4 * i + 0] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
  }
}

// ---------------------------------------------------------------------------------------
// Triangles: BVHRT::IntersectAllTrianglesInLeaf (BVH2Common.cpp:1571-1635), reading the
// leaf-ordered vertex soup: triangle triId of geometry geomId is record off.x/3 + triId.
__device__ __forceinline__ void intersect_tris(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                               float tNear, uint32_t instId, uint32_t geomId, uint32_t a_start,
                                               uint32_t a_count, CRT_Hit& hit) {
  const uint32_t tri0 = g.offset.x / 3;
  for (uint32_t triId = a_start; triId < a_start + a_count; triId++) {
    const float4* q = s.tri_soup + size_t(tri0 + triId) * 3;
    const float4 a = q[0], b = q[1], c = q[2];
    const float3 A_pos = xyz(a);
    const float3 B_pos = xyz(b);
    const float3 C_pos = xyz(c);
    const float3 edge1 = B_pos - A_pos;
    const float3 edge2 = C_pos - A_pos;
    const float3 pvec = cross3(ray_dir, edge2);
    const float3 tvec = ray_pos - A_pos;
    const float3 qvec = cross3(tvec, edge1);
    const float invDet = 1.0f / dot3(edge1, pvec);
    const float v = dot3(tvec, pvec) * invDet;
    const float u = dot3(qvec, ray_dir) * invDet;
    const float t = dot3(edge2, qvec) * invDet;
    if (v >= -1e-6f && u >= -1e-6f && (u + v <= 1.0f + 1e-6f) && t > tNear && t < hit.t) {
      hit.t = t;
      hit.primId = triId;
      hit.instId = instId;
      hit.geomId = geomId | (TYPE_MESH_TRIANGLE << SH_TYPE);
      hit.coords[0] = u;
      hit.coords[1] = v;
      hit.coords[2] = 0.0f;
      hit.coords[3] = 0.0f;
    }
  }
}

// Normal of the final triangle hit (BVH2Common.cpp:1609-1631), computed once the closest hit
// is known: it depends only on the winning triangle and its barycentrics, which the hit record
// names (geomId, instId and primId before the m_primIndices remap; the render mode that
// rewrites instId, INTERSECTION_COUNT, takes no normal).
__device__ __forceinline__ void resolve_tri_normal(const SceneView& s, CRT_Hit& hit) {
  if (!preset_needs_normal(s.preset.render_mode)) return;
  const uint2 off = s.geoms[hit.geomId & GEOM_ID_MASK].offset;
  const float4* q = s.tri_soup + size_t(off.x / 3 + hit.primId) * 3;
  const float4 qa = q[0], qb = q[1], qc = q[2];
  const uint32_t A = __float_as_uint(qa.w), B = __float_as_uint(qb.w), C = __float_as_uint(qc.w);
  const float u = hit.coords[0], v = hit.coords[1];
  float3 n = make3(1.0f, 0.0f, 0.0f);
  if (s.preset.normal_mode == NORMAL_MODE_GEOMETRY) {
    const float3 A_pos = xyz(qa);
    const float3 edge1 = xyz(qb) - A_pos;
    const float3 edge2 = xyz(qc) - A_pos;
    n = cross3(edge1, edge2);
  } else if (s.preset.normal_mode == NORMAL_MODE_VERTEX) {
    const float4 nA = s.vert_norm[off.y + A], nB = s.vert_norm[off.y + B], nC = s.vert_norm[off.y + C];
    const float wA = 1.0f - u - v;
    n = make3(nA.x * wA + nB.x * v + u * nC.x, nA.y * wA + nB.y * v + u * nC.y, nA.z * wA + nB.z * v + u * nC.z);
  }
  float m[16];
  load_mat(s.instances[hit.instId].transformInvTransposed, m);
  n = normalize3(matmul4x3(m, n));
  const float2 e = encode_normal(n);
  hit.coords[2] = e.x;
  hit.coords[3] = e.y;
}

// ---------------------------------------------------------------------------------------
// SDF helpers
// BVHRT::eval_dist_trilinear (BVH2Common.cpp:126-136)
__device__ __forceinline__ float eval_dist_trilinear(const float v[8], float3 dp) {
  return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * v[0] +
         (1 - dp.x) * (1 - dp.y) * (dp.z) * v[1] +
         (1 - dp.x) * (dp.y) * (1 - dp.z) * v[2] +
         (1 - dp.x) * (dp.y) * (dp.z) * v[3] +
         (dp.x) * (1 - dp.y) * (1 - dp.z) * v[4] +
         (dp.x) * (1 - dp.y) * (dp.z) * v[5] +
         (dp.x) * (dp.y) * (1 - dp.z) * v[6] +
         (dp.x) * (dp.y) * (dp.z) * v[7];
}

// Gradient of the trilinear interpolant (same corner order as eval_dist_trilinear).
__device__ __forceinline__ float3 eval_grad_trilinear(const float v[8], float3 q) {
  const float x = q.x, y = q.y, z = q.z;
  const float gx = (1 - y) * (1 - z) * (v[4] - v[0]) + (1 - y) * z * (v[5] - v[1]) + y * (1 - z) * (v[6] - v[2]) +
                   y * z * (v[7] - v[3]);
  const float gy = (1 - x) * (1 - z) * (v[2] - v[0]) + (1 - x) * z * (v[3] - v[1]) + x * (1 - z) * (v[6] - v[4]) +
                   x * z * (v[7] - v[5]);
  const float gz = (1 - x) * (1 - y) * (v[1] - v[0]) + (1 - x) * y * (v[3] - v[2]) + x * (1 - y) * (v[5] - v[4]) +
                   x * y * (v[7] - v[6]);
  return make3(gx, gy, gz);
}

// BVHRT::LocalSurfaceIntersection (BVH2Common.cpp:345-725). Returns true when the hit record was
// updated; then *t_local (if given) receives the hit parameter in voxel units along ray_dir from
// start_q. `surface_node` (multi-octree, DAG and SCom2 nodes flagged as surface) makes the start
// sign follow the start distance, as SURFACE mode does. OCTREE_TYPES compiles in what only the
// multi-octree, DAG and SCom2 types use: the compact types' wider stopping distance on surface
// nodes and the channel render modes' hit-point quotients (the other callers stay free of it).
template <bool OCTREE_TYPES = false>
__device__ bool local_surface_intersection(const SceneView& s, uint32_t type, float3 ray_dir, uint32_t instId,
                                           uint32_t geomId, const float values[8], uint32_t primId, float d,
                                           float qNear, float qFar, float2 fNearFar, float3 start_q,
                                           CRT_Hit& hit, float* t_local = nullptr, bool surface_node = false) {
  // surface nodes of the compact formats stop at a distance scaled with the voxel (BVH2Common.cpp:340-351)
  const float EPS = (OCTREE_TYPES && surface_node && (type == TYPE_SDF_DAG || type == TYPE_SCOM2))
                        ? fmax_ref(1e-6f, s.preset.compact_sdf_eps * d) : 1e-6f;
  const float d_inv = 1.0f / d;
  float t = qNear;
  bool early_hit = false;
  bool hitp = false;
  unsigned iter = 0;
  float start_sign = 1;
  float start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
  if (surface_node || s.preset.representation_mode == REPRESENTATION_MODE_SURFACE) start_sign = sign_ref(start_dist);
  start_dist *= start_sign;
  const uint32_t mode = s.preset.sdf_node_intersect;

  if (start_dist <= EPS || mode == SDF_OCTREE_NODE_INTERSECT_BBOX) {
    early_hit = true;
    hitp = true;
  } else if (mode == SDF_OCTREE_NODE_INTERSECT_ST) {
    const unsigned ST_max_iters = 256;
    float dist = start_dist;
    while (t < qFar && dist > EPS && iter < ST_max_iters) {
      t += dist * d_inv;
      dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
      iter++;
    }
    hitp = (dist <= EPS);
  } else {
    const float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv,
                s011 = values[3] * d_inv, s100 = values[4] * d_inv, s101 = values[5] * d_inv,
                s110 = values[6] * d_inv, s111 = values[7] * d_inv;
    const float tmp = s101 - s001;
    const float k0 = s000;
    const float k1 = s100 - s000;
    const float k2 = s010 - s000;
    const float k3 = s110 - s010 - k1;
    const float k4 = k0 - s001;
    const float k5 = k1 - tmp;
    const float k6 = k2 - (s011 - s001);
    const float k7 = k3 - (s111 - s011 - tmp);
    const float3 o = start_q;
    const float3 d3 = ray_dir;
    const float m0 = o.x * o.y;
    const float m1 = d3.x * d3.y;
    const float m2 = o.x * d3.y + o.y * d3.x;
    const float m3 = k5 * o.z - k1;
    const float m4 = k6 * o.z - k2;
    const float m5 = k7 * o.z - k3;
    const float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
    const float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
    const float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
    const float c3 = k7 * m1 * d3.z;

    if (mode == SDF_OCTREE_NODE_INTERSECT_ANALYTIC) {
      float x1 = 1000, x2 = 1000, x3 = 1000;
      if (fabs_ref(c3) > 1e-5f) {
        const float a = c2 / c3, b = c1 / c3, c = c0 / c3;
        const float Q = (a * a - 3 * b) / 9;
        const float R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
        const float Q3 = Q * Q * Q;
        if (R * R < Q3) {
          const float PI = 3.14159265358979323846f;
          const float theta = acosf(R / sqrtf(Q3));
          x1 = -2 * sqrtf(Q) * cosf(theta / 3) - a / 3;
          x2 = -2 * sqrtf(Q) * cosf((theta + 2 * PI) / 3) - a / 3;
          x3 = -2 * sqrtf(Q) * cosf((theta - 2 * PI) / 3) - a / 3;
        } else {
          const float A = -sign_ref(R) * powf(fabs_ref(R) + sqrtf(R * R - Q3), 1.0f / 3.0f);
          const float B = fabs_ref(A) > EPS ? Q / A : 0;
          x1 = A + B - a / 3;
        }
      } else if (fabs_ref(c2) > 1e-7f) {
        const float a = c2, b = c1, c = c0;
        const float D = b * b - 4 * a * c;
        if (D > 0) {
          const float q = -0.5f * (b + sign_ref(b) * sqrtf(D));
          x1 = q / a;
          if (fabs_ref(q) > EPS) x2 = c / q;
        }
      } else if (fabs_ref(c1) > EPS) {
        x1 = -c0 / c1;
      }
      x1 = x1 < 0 ? 1000 : x1;
      x2 = x2 < 0 ? 1000 : x2;
      x3 = x3 < 0 ? 1000 : x3;
      t = fmin_ref(x1, fmin_ref(x2, x3));
      hitp = (t >= 0 && t <= qFar);
    } else if (mode == SDF_OCTREE_NODE_INTERSECT_NEWTON) {
      const float a = 3 * c3, b = 2 * c2, c = c1;
      const float t0 = 0;
      float t1 = qFar, t2 = qFar;
      const float t3 = qFar;
      const float D = b * b - 4 * a * c;
      if (D >= 0) {
        const float q = -0.5f * (b + sign_ref(b) * sqrtf(D));
        t1 = fabs_ref(a) > EPS ? q / a : t0;
        t2 = fabs_ref(q) > EPS ? c / q : qFar;
        const float min_t = fmin_ref(t1, t2);
        t2 = fmax_ref(t1, t2);
        t1 = min_t;
        t1 = clamp_ref(t1, t0, t3);
        t2 = clamp_ref(t2, t0, t3);
      }
      const bool s0 = c0 > 0;
      const bool s1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3))) > 0;
      const bool s2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3))) > 0;
      const bool s3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3))) > 0;
      float nwt_min = t0, nwt_max = t0;
      if (s0 != s1) { nwt_min = t0; nwt_max = t1; }
      else if (s1 != s2) { nwt_min = t1; nwt_max = t2; }
      else if (s2 != s3) { nwt_min = t2; nwt_max = t3; }
      if (nwt_min < nwt_max) {
        const unsigned max_iters = 10;
        unsigned nwt_iter = 0;
        float rtn = 0.5f * (nwt_min + nwt_max);
        float f = 1000;
        while (nwt_iter < max_iters && fabs_ref(f) >= EPS) {
          f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
          const float df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
          const float dx = f / (df + sign_ref(df) * 1e-9f);
          rtn -= dx;
          nwt_iter++;
        }
        t = rtn;
        hitp = (t >= 0 && t <= qFar && fabs_ref(f) < EPS);
      } else {
        hitp = false;
      }
    } else {  // SDF_OCTREE_NODE_INTERSECT_IT
      const unsigned IT_max_iters = 256;
      const float k = 2;
      float e = 0.1f * qFar;
      const float t_max = fabs_ref(c3) < EPS ? 1e6f : -c2 / (3 * c3);
      const float df_max = 3 * c3 * t_max * t_max + 2 * c2 * t_max + c1;
      float dist = start_dist;
      while (t < qFar && dist > EPS && iter < IT_max_iters) {
        const float df_1 = 3 * c3 * t * t + 2 * c2 * t + c1;
        const float df_2 = 3 * c3 * (t + e) * (t + e) + 2 * c2 * (t + e) + c1;
        float L = (t_max > t && t_max < t + e) ? fmax_ref(df_max, fmax_ref(df_1, df_2)) : fmax_ref(df_1, df_2);
        L = fmax_ref(L, EPS);
        const float st = fmin_ref((dist * d_inv) / L, e);
        t += st;
        e = k * st;
        dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
        iter++;
      }
      hitp = (dist <= EPS);
    }
  }

  const float tReal = fNearFar.x + d * t;
  if (t <= qFar && hitp && tReal < hit.t) {
    float3 norm = make3(0, 0, 1);
    if (preset_needs_normal(s.preset.render_mode)) {
      if (early_hit && mode == SDF_OCTREE_NODE_INTERSECT_BBOX) {
        norm = make3(0, 0, 0);
        const float3 v = make3(fmax_ref(start_q.x, 1 - start_q.x), fmax_ref(start_q.y, 1 - start_q.y),
                               fmax_ref(start_q.z, 1 - start_q.z));
        int dim = 0;
        if (v.y > v.x && v.y > v.z) dim = 1;
        else if (v.z > v.x && v.z > v.y) dim = 2;
        const float sq = dim == 0 ? start_q.x : (dim == 1 ? start_q.y : start_q.z);
        const float nv = sq < 0.5f ? -1.0f : 1.0f;
        if (dim == 0) norm.x = nv; else if (dim == 1) norm.y = nv; else norm.z = nv;
      } else {
        const float3 p0 = start_q + t * ray_dir;
        const float h = 0.001f;
        const float ddx = (eval_dist_trilinear(values, p0 + make3(h, 0, 0)) -
                           eval_dist_trilinear(values, p0 + make3(-h, 0, 0))) / (2 * h);
        const float ddy = (eval_dist_trilinear(values, p0 + make3(0, h, 0)) -
                           eval_dist_trilinear(values, p0 + make3(0, -h, 0))) / (2 * h);
        const float ddz = (eval_dist_trilinear(values, p0 + make3(0, 0, h)) -
                           eval_dist_trilinear(values, p0 + make3(0, 0, -h))) / (2 * h);
        norm = start_sign * make3(ddx, ddy, ddz);
      }
      float m[16];
      load_mat(s.instances[instId].transformInvTransposed, m);
      norm = normalize3(matmul4x3(m, norm));
    }
    const float2 en = encode_normal(norm);
    hit.t = tReal;
    hit.primId = primId;
    hit.instId = instId;
    hit.geomId = geomId | (type << SH_TYPE);
    hit.coords[0] = 0;
    hit.coords[1] = 0;
    hit.coords[2] = en.x;
    hit.coords[3] = en.y;
    if (s.preset.render_mode == MULTI_RENDER_MODE_ST_ITERATIONS) hit.primId = iter;
    // the channel render modes keep the voxel-local hit point (16 bits per axis) instead of the
    // texture coordinates (BVH2Common.cpp:714-723)
    if (OCTREE_TYPES && (type == TYPE_SDF_MULTI_OCTREE || type == TYPE_SDF_DAG || type == TYPE_SCOM2) &&
        (s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL || s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL_LAMBERT)) {
      const float3 dq = start_q + t * ray_dir;
      const int qx = int(clamp_ref(65535.0f * dq.x, 0.0f, 65535.0f));
      const int qy = int(clamp_ref(65535.0f * dq.y, 0.0f, 65535.0f));
      const int qz = int(clamp_ref(65535.0f * dq.z, 0.0f, 65535.0f));
      hit.coords[0] = __int_as_float((qy << 16) | qx);
      hit.coords[1] = __int_as_float(qz);
    }
    if (t_local) *t_local = t;
    return true;
  }
  return false;
}

// Texture coordinate c (0: u, 1: v) of a textured frame-octree hit: the corner pairs weighted
// trilinearly at the voxel-local hit point dp (BVH2Common.cpp:689-710, same order of operations).
__device__ __forceinline__ float frame_tex_interp(const float tc[16], float3 dp, int c) {
  const float x0 = 1 - dp.x, y0 = 1 - dp.y, z0 = 1 - dp.z;
  return x0 * y0 * z0 * tc[c] + x0 * y0 * dp.z * tc[2 + c] + x0 * dp.y * z0 * tc[4 + c] + x0 * dp.y * dp.z * tc[6 + c] +
         dp.x * y0 * z0 * tc[8 + c] + dp.x * y0 * dp.z * tc[10 + c] + dp.x * dp.y * z0 * tc[12 + c] +
         dp.x * dp.y * dp.z * tc[14 + c];
}

// Frame-octree leaf: BVHRT::OctreeNodeIntersect + RayNodeIntersection (BVH2Common.cpp:727-749,
// 141-188), reading the leaf's box, corner values and node index from one FrameLeaf record; with
// TEX (kernels taking LK_FRAME_TEX) a textured octree's hit also takes its texture coordinates
// (LocalSurfaceIntersection, :689-710).
template <bool TEX = false>
__device__ __forceinline__ void frame_node_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                                     float tNear, uint32_t instId, uint32_t geomId,
                                                     uint32_t bvhNodeId, CRT_Hit& hit) {
  const uint32_t type = g.type;
  // frame octrees and sparse voxel sets (BVH2Common.cpp:141-218): one leaf record per BLAS leaf
  if (type != TYPE_SDF_FRAME_OCTREE && type != TYPE_SDF_SVS && (!TEX || type != TYPE_SDF_FRAME_OCTREE_TEX)) return;
  const uint2 go = g.offset;   // y: first leaf record of this geometry
  const float4* q = reinterpret_cast<const float4*>(s.frame_leaves + go.y + bvhNodeId);
  const float4 r0 = q[0], r1 = q[1], r2 = q[2], r3 = q[3];
  const float3 min_pos = make3(r0.x, r0.y, r0.z);
  const float3 max_pos = make3(r0.w, r1.x, r1.y);
  const float values[8] = {r1.z, r1.w, r2.x, r2.y, r2.z, r2.w, r3.x, r3.y};
  const uint32_t primId = __float_as_uint(r3.z);
  const float3 size = max_pos - min_pos;
  const float2 fNearFar = ray_box2(ray_pos, safe_inverse(ray_dir), min_pos, max_pos);
  const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
  const float d = fmax_ref(size.x, fmax_ref(size.y, size.z));
  const float3 start_q = (start_pos - min_pos) / d;
  const float qFar = (fNearFar.y - fNearFar.x) / d;
  const float qNear = tNear > fNearFar.x ? (tNear - fNearFar.x) / d : 0.0f;
  if (qNear > 0.0f) return;
  if (!TEX) {
    local_surface_intersection(s, type, ray_dir, instId, geomId, values, primId, d, qNear, qFar, fNearFar, start_q, hit);
    return;
  }
  float t_local = 0.0f;
  if (local_surface_intersection(s, type, ray_dir, instId, geomId, values, primId, d, qNear, qFar, fNearFar, start_q, hit,
                                 &t_local) &&
      type == TYPE_SDF_FRAME_OCTREE_TEX) {
    const float4* tq = s.frame_tex + 4 * size_t(__float_as_uint(r3.w));
    const float4 a = tq[0], b = tq[1], c = tq[2], e = tq[3];
    const float tc[16] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, e.x, e.y, e.z, e.w};
    const float3 dp = start_q + t_local * ray_dir;
    hit.coords[0] = frame_tex_interp(tc, dp, 0);
    hit.coords[1] = frame_tex_interp(tc, dp, 1);
  }
}

// BVHRT::load_distance_values (BVH2Common.cpp:846-882); returns the minimum value.
__device__ __forceinline__ float sbs_load_values(const SceneView& s, const SdfSBSHeader& h, uint32_t nodeId,
                                                 float3 voxelPos, uint32_t v_size, float sz_inv, float values[8]) {
  float vmin = 1e6f;
  const uint32_t v_off = s.sbs_nodes[nodeId].data_offset;
  const uint32_t vx = uint32_t(voxelPos.x), vy = uint32_t(voxelPos.y), vz = uint32_t(voxelPos.z);
  if (h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F || h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t px = vx + ((i & 4) >> 2), py = vy + ((i & 2) >> 1), pz = vz + (i & 1);
      const uint32_t vId = (px + h.brick_pad) * v_size * v_size + (py + h.brick_pad) * v_size + (pz + h.brick_pad);
      values[i] = s.sbs_data_f[s.sbs_data[v_off + vId]];
      vmin = fmin_ref(vmin, values[i]);
    }
  } else {
    const uint32_t vals_per_int = 4 / h.bytes_per_value;
    const uint32_t bits = 8 * h.bytes_per_value;
    const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
    const float d_max = 1.73205081f * sz_inv;
    const float mult = 2 * d_max / float(max_val);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t px = vx + ((i & 4) >> 2), py = vy + ((i & 2) >> 1), pz = vz + (i & 1);
      const uint32_t vId = (px + h.brick_pad) * v_size * v_size + (py + h.brick_pad) * v_size + (pz + h.brick_pad);
      const uint32_t word = s.sbs_data[v_off + vId / vals_per_int];
      const uint32_t sh = bits * (vId % vals_per_int);
      values[i] = -d_max + mult * float((sh < 32 ? (word >> sh) : 0u) & max_val);
      vmin = fmin_ref(vmin, values[i]);
    }
  }
  return vmin;
}

// Value of lattice point (x, y, z) (corner coordinates 0..bs) of a brick, decoded exactly as
// sbs_load_values decodes it.
__device__ __forceinline__ float sbs_lattice_value(const SceneView& s, const SdfSBSHeader& h, uint32_t v_off,
                                                   uint32_t v_size, float sz_inv, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t vId = (x + h.brick_pad) * v_size * v_size + (y + h.brick_pad) * v_size + (z + h.brick_pad);
  if (h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F || h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN)
    return s.sbs_data_f[s.sbs_data[v_off + vId]];
  const uint32_t vals_per_int = 4 / h.bytes_per_value;
  const uint32_t bits = 8 * h.bytes_per_value;
  const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
  const float d_max = 1.73205081f * sz_inv;
  const float mult = 2 * d_max / float(max_val);
  const uint32_t word = s.sbs_data[v_off + vId / vals_per_int];
  const uint32_t sh = bits * (vId % vals_per_int);
  return -d_max + mult * float((sh < 32 ? (word >> sh) : 0u) & max_val);
}

// Occupancy mask of one brick, built by one wave (the block is one wave): each of the brick's
// (bs+1)^3 lattice values is decoded once (not once per voxel corner) and its "<= 0" flag kept as
// a bit in LDS; a voxel's bit is the OR of its 8 corner flags, which is the predicate "minimum of
// the 8 corner values <= 0" (that minimum skips NaN values, as the comparison does). Writes the
// summary word and the bs^3 bits to out[0 .. wpb].
__device__ __forceinline__ void sbs_mask_brick(const SceneView& s, const SdfSBSHeader& h, uint32_t root, uint32_t brick,
                                               uint32_t* __restrict__ out, uint32_t* flags) {
  const uint32_t bs = h.brick_size, n1 = bs + 1, np = n1 * n1 * n1, v3 = bs * bs * bs, wpb = (v3 + 31u) / 32u;
  const uint32_t node_id = brick + root;
  const uint32_t v_size = bs + 2 * h.brick_pad + 1;
  const float sz_inv = 2.0f / float(s.sbs_nodes[node_id].pos_z_lod_size & 0x0000FFFF);
  const uint32_t v_off = s.sbs_nodes[node_id].data_offset;
  const uint32_t lane = threadIdx.x;
  for (uint32_t base = 0; base < np; base += 64u) {
    const uint32_t q = base + lane;
    bool f = false;
    if (q < np) f = sbs_lattice_value(s, h, v_off, v_size, sz_inv, q / (n1 * n1), (q / n1) % n1, q % n1) <= 0.0f;
    const uint64_t m = __ballot(f);
    if ((lane & 31u) == 0u) flags[base / 32u + (lane >> 5)] = uint32_t(m >> (lane & 32u));
  }
  __syncthreads();
  bool any = false;
  for (uint32_t base = 0; base < v3; base += 64u) {
    const uint32_t v = base + lane;
    bool bit = false;
    if (v < v3) {
      const uint32_t vx = v / (bs * bs), vy = (v / bs) % bs, vz = v % bs;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t q = (vx + ((i >> 2) & 1)) * n1 * n1 + (vy + ((i >> 1) & 1)) * n1 + (vz + (i & 1));
        bit = bit || ((flags[q >> 5] >> (q & 31u)) & 1u) != 0u;
      }
    }
    const uint64_t m = __ballot(bit);
    any = any || m != 0;
    const uint32_t w = base / 32u + (lane >> 5);
    if ((lane & 31u) == 0u && w < wpb) out[1 + w] = uint32_t(m >> (lane & 32u));
  }
  if (lane == 0) out[0] = any ? 1u : 0u;
  __syncthreads();   // the flags are reused for the next brick
}

// Occupancy bit of voxel v (x-major, v < bs^3) of brick `brick`: the minimum of its 8 corner
// values, loaded as the march loads them, is <= 0.
__device__ __forceinline__ bool sbs_voxel_occupied(const SceneView& s, const SdfSBSHeader& h, uint32_t root,
                                                   uint32_t brick, uint32_t v) {
  const uint32_t bs = h.brick_size;
  const uint32_t node_id = brick + root;
  const uint32_t v_size = bs + 2 * h.brick_pad + 1;
  const float sz_inv = 2.0f / float(s.sbs_nodes[node_id].pos_z_lod_size & 0x0000FFFF);
  const float3 vp = make3(float(v / (bs * bs)), float((v / bs) % bs), float(v % bs));
  float values[8];
  return sbs_load_values(s, h, node_id, vp, v_size, sz_inv, values) <= 0.0f;
}

// The voxel whose surface produced a brick-set hit: the value indices of its 8 corners (into
// the scene's values_f, indexed layouts), the entry point in voxel units, and the hit
// parameter along the ray direction from there.
// tricubic_spline (BVH2Common.cpp:1526-1529): the Catmull-Rom cubic through p0..p3, at x in [0, 1]
// between p1 and p2 (same order of operations).
__device__ __forceinline__ float tricubic_spline(float p0, float p1, float p2, float p3, float x) {
  return p1 + 0.5f * x * (p2 - p0 + x * (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3 + x * (3.0f * (p1 - p2) + p3 - p0)));
}

// BVHRT::tricubicInterpolation (BVH2Common.cpp:1531-1569): g[16 z + 4 y + x] holds the 4^3 samples
// around the cell, dp the position in the cell; splines along x, then y, then z.
__device__ __forceinline__ float tricubic_interp(const float g[64], float3 dp) {
  float yz[16];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      yz[4 * j + k] = tricubic_spline(g[16 * k + 4 * j], g[16 * k + 4 * j + 1], g[16 * k + 4 * j + 2], g[16 * k + 4 * j + 3],
                                      dp.x);
  float z4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) z4[k] = tricubic_spline(yz[k], yz[4 + k], yz[8 + k], yz[12 + k], dp.y);
  return tricubic_spline(z4[0], z4[1], z4[2], z4[3], dp.z);
}

// BVHRT::load_tricubic_distance_values (BVH2Common.cpp:790-844): the 4^3 values from one voxel
// before the voxel's corner to two after, at 16 z + 4 y + x; returns their minimum. Needs
// brick_pad >= 1 (trace calls refuse tricubic brick sets without padding).
__device__ __forceinline__ float sbs_load_tricubic(const SceneView& s, const SdfSBSHeader& h, uint32_t nodeId,
                                                   float3 voxelPos, uint32_t v_size, float sz_inv, float values[64]) {
  float vmin = 1e6f;
  const uint32_t v_off = s.sbs_nodes[nodeId].data_offset;
  const uint32_t bx = uint32_t(voxelPos.x) + h.brick_pad - 1, by = uint32_t(voxelPos.y) + h.brick_pad - 1,
                 bz = uint32_t(voxelPos.z) + h.brick_pad - 1;
  const bool indexed = h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F || h.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN;
  const uint32_t vals_per_int = 4 / h.bytes_per_value;
  const uint32_t bits = 8 * h.bytes_per_value;
  const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
  const float d_max = 1.73205081f * sz_inv;
  const float mult = 2 * d_max / float(max_val);
  for (uint32_t x = 0; x < 4; x++)
    for (uint32_t y = 0; y < 4; y++)
#pragma unroll
      for (uint32_t z = 0; z < 4; z++) {
        const uint32_t vId = (bx + x) * v_size * v_size + (by + y) * v_size + (bz + z);
        float v;
        if (indexed) {
          v = s.sbs_data_f[s.sbs_data[v_off + vId]];
        } else {
          const uint32_t word = s.sbs_data[v_off + vId / vals_per_int];
          const uint32_t sh = bits * (vId % vals_per_int);
          v = -d_max + mult * float((sh < 32 ? (word >> sh) : 0u) & max_val);
        }
        values[16 * z + 4 * y + x] = v;
        vmin = fmin_ref(vmin, v);
      }
  return vmin;
}

// BVHRT::TricubicLocalSurfaceIntersection (BVH2Common.cpp:220-335): the start sample, then sphere
// tracing in ST mode (other modes hit only where the ray starts inside, or at once in BBOX mode);
// central-difference normal of the interpolant, turned by the instance's inverse transpose and
// then signed. Returns true when the hit record was updated.
__device__ bool tricubic_local_surface(const SceneView& s, uint32_t type, float3 ray_dir, uint32_t instId,
                                       uint32_t geomId, const float values[64], uint32_t primId, float d, float qNear,
                                       float qFar, float2 fNearFar, float3 start_q, CRT_Hit& hit) {
  const float EPS = 1e-6f;
  const float d_inv = 1.0f / d;
  float t = qNear;
  bool hitp = false;
  unsigned iter = 0;
  float start_sign = 1;
  float start_dist = tricubic_interp(values, start_q + t * ray_dir);
  if (s.preset.representation_mode == REPRESENTATION_MODE_SURFACE) start_sign = sign_ref(start_dist);
  start_dist *= start_sign;
  if (start_dist <= EPS || s.preset.sdf_node_intersect == SDF_OCTREE_NODE_INTERSECT_BBOX) {
    hitp = true;
  } else if (s.preset.sdf_node_intersect == SDF_OCTREE_NODE_INTERSECT_ST) {
    const unsigned ST_max_iters = 256;
    float dist = start_dist;
    while (t < qFar && dist > EPS && iter < ST_max_iters) {
      t += dist * d_inv;
      dist = tricubic_interp(values, start_q + t * ray_dir);
      iter++;
    }
    hitp = (dist <= EPS);
  }
  const float tReal = fNearFar.x + d * t;
  if (!(t <= qFar && hitp && tReal < hit.t)) return false;
  float3 norm = make3(0, 0, 1);
  if (preset_needs_normal(s.preset.render_mode)) {
    const float3 p0 = start_q + t * ray_dir;
    const float h = 0.001f;
    const float ddx = (tricubic_interp(values, make3(p0.x + h, p0.y, p0.z)) -
                       tricubic_interp(values, make3(p0.x + -h, p0.y, p0.z))) / (2 * h);
    const float ddy = (tricubic_interp(values, make3(p0.x, p0.y + h, p0.z)) -
                       tricubic_interp(values, make3(p0.x, p0.y + -h, p0.z))) / (2 * h);
    const float ddz = (tricubic_interp(values, make3(p0.x, p0.y, p0.z + h)) -
                       tricubic_interp(values, make3(p0.x, p0.y, p0.z + -h))) / (2 * h);
    float m[16];
    load_mat(s.instances[instId].transformInvTransposed, m);
    norm = start_sign * normalize3(matmul4x3(m, make3(ddx, ddy, ddz)));
  }
  const float2 en = encode_normal(norm);
  hit.t = tReal;
  hit.primId = primId;
  hit.instId = instId;
  hit.geomId = geomId | (type << SH_TYPE);
  hit.coords[0] = 0;
  hit.coords[1] = 0;
  hit.coords[2] = en.x;
  hit.coords[3] = en.y;
  if (s.preset.render_mode == MULTI_RENDER_MODE_ST_ITERATIONS) hit.primId = iter;
  return true;
}

// BVHRT::eval_distance_sdf_grid (BVH2Common.cpp:1442-1520): the grid spans the unit cube;
// trilinear, or with TRICUBIC the 4^3 samples around interior cells (trilinear next to the faces);
// outside the last cell the nearest sample is returned.
template <bool TRICUBIC = false>
__device__ __forceinline__ float sdf_grid_eval(const SceneView& s, uint32_t grid_id, float3 pos) {
  const uint32_t off = s.grid_offsets[grid_id];
  const uint32_t sx = s.grid_sizes[3 * grid_id], sy = s.grid_sizes[3 * grid_id + 1], sz = s.grid_sizes[3 * grid_id + 2];
  const float3 gsf = make3(float(sx - 1), float(sy - 1), float(sz - 1));
  float3 vox_f = gsf * ((pos - make3(-1, -1, -1)) / make3(2, 2, 2));
  vox_f = min3(max3(vox_f, make3(0.0f, 0.0f, 0.0f)), gsf - make3(1e-5f, 1e-5f, 1e-5f));
  const uint32_t ux = uint32_t(vox_f.x), uy = uint32_t(vox_f.y), uz = uint32_t(vox_f.z);
  const float3 dp = vox_f - make3(float(ux), float(uy), float(uz));
  const float* data = s.grid_data + off;
  float res = 0.0f;
  if (TRICUBIC && ux < sx - 2 && uy < sy - 2 && uz < sz - 2 && ux > 0 && uy > 0 && uz > 0) {
    float g[64];
    for (uint32_t x = 0; x < 4; ++x)
      for (uint32_t y = 0; y < 4; ++y)
#pragma unroll
        for (uint32_t z = 0; z < 4; ++z)
          g[16 * z + 4 * y + x] = data[(uz - 1 + z) * sx * sy + (uy - 1 + y) * sx + (ux - 1 + x)];
    res = tricubic_interp(g, dp);
  } else if (ux < sx - 1 && uy < sy - 1 && uz < sz - 1) {
    for (uint32_t i = 0; i < 2; i++)
      for (uint32_t j = 0; j < 2; j++)
        for (uint32_t k = 0; k < 2; k++) {
          const float qx = (1 - dp.x + float(i) * (2 * dp.x - 1));
          const float qy = (1 - dp.y + float(j) * (2 * dp.y - 1));
          const float qz = (1 - dp.z + float(k) * (2 * dp.z - 1));
          res += qx * qy * qz * data[(uz + k) * sx * sy + (uy + j) * sx + (ux + i)];
        }
  } else {
    res += data[uz * sx * sy + uy * sx + ux];
  }
  return res;
}

// BVHRT::box_intersects (BVH2Common.cpp:34-45)
__device__ __forceinline__ float2 box_intersects_ref(float3 min_pos, float3 max_pos, float3 origin, float3 dir) {
  const float3 safe_dir = make3(sign_ref(dir.x) * fmax_ref(1e-9f, fabs_ref(dir.x)),
                                sign_ref(dir.y) * fmax_ref(1e-9f, fabs_ref(dir.y)),
                                sign_ref(dir.z) * fmax_ref(1e-9f, fabs_ref(dir.z)));
  const float3 tMin = (min_pos - origin) / safe_dir;
  const float3 tMax = (max_pos - origin) / safe_dir;
  const float3 t1 = min3(tMin, tMax), t2 = max3(tMin, tMax);
  float2 r;
  r.x = fmax_ref(t1.x, fmax_ref(t1.y, t1.z));
  r.y = fmin_ref(t2.x, fmin_ref(t2.y, t2.z));
  return r;
}

// BVHRT::IntersectSDFGrid (BVH2Common.cpp:1114-1182): sphere tracing along the normalised
// direction from the grid box's entry (at most 1000 steps, EPS 1e-5), central-difference normal.
// TRICUBIC: the preset's tricubic interpolation.
template <bool TRICUBIC = false>
__device__ __forceinline__ void sdf_grid_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                                   float tNear, uint32_t instId, uint32_t geomId, CRT_Hit& hit) {
  const uint32_t sdfId = g.offset.x;
  const float dir_len = length3(ray_dir);
  const float3 dir = make3(ray_dir.x / dir_len, ray_dir.y / dir_len, ray_dir.z / dir_len);
  const float EPS = 1e-5f;
  const float2 tt = box_intersects_ref(make3(-1, -1, -1), make3(1, 1, 1), ray_pos, dir);
  float t = fmax_ref(tNear, tt.x);
  const float tFar = tt.y;
  if (t > tFar) return;
  uint32_t iter = 0;
  float d = sdf_grid_eval<TRICUBIC>(s, sdfId, ray_pos + t * dir);
  while (iter < 1000 && d > EPS && t < tFar) {
    t += d + EPS;
    d = sdf_grid_eval<TRICUBIC>(s, sdfId, ray_pos + t * dir);
    iter++;
  }
  if (d > EPS) return;
  const float3 p0 = ray_pos + t * dir;
  if (t > tNear && t < hit.t) {
    float3 norm = make3(0, 0, 1);
    if (preset_needs_normal(s.preset.render_mode)) {
      const float h = 0.001f;
      const float ddx = (sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(h, 0, 0)) - sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(-h, 0, 0))) / (2 * h);
      const float ddy = (sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(0, h, 0)) - sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(0, -h, 0))) / (2 * h);
      const float ddz = (sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(0, 0, h)) - sdf_grid_eval<TRICUBIC>(s, sdfId, p0 + make3(0, 0, -h))) / (2 * h);
      norm = normalize3(make3(ddx, ddy, ddz));
    }
    float m[16];
    load_mat(s.instances[instId].transformInvTransposed, m);
    const float2 e = encode_normal(normalize3(matmul4x3(m, norm)));
    hit.t = t;
    hit.primId = 0;
    hit.instId = instId;
    hit.geomId = geomId | (TYPE_SDF_GRID << SH_TYPE);
    hit.coords[0] = 0;
    hit.coords[1] = 0;
    hit.coords[2] = e.x;
    hit.coords[3] = e.y;
    if (s.preset.render_mode == MULTI_RENDER_MODE_ST_ITERATIONS) hit.primId = iter;
  }
}

// rayCapsuleIntersect (BVH2Common.cpp:1184-1223): (normal, distance), distance 2e15 on a miss.
__device__ __forceinline__ float4 ray_capsule(float3 ray_pos, float3 ray_dir, float3 pa, float3 pb, float ra) {
  const float3 ba = pb - pa, oa = ray_pos - pa, ob = ray_pos - pb;
  const float baba = dot3(ba, ba), bard = dot3(ba, ray_dir), baoa = dot3(ba, oa), rdoa = dot3(ray_dir, oa),
              oaoa = dot3(oa, oa);
  const float a = baba - bard * bard;
  float b = baba * rdoa - baoa * bard;
  float c = baba * oaoa - baoa * baoa - ra * ra * baba;
  float h = b * b - a * c;
  if (h >= 0.f) {
    float dist = (-b - std::sqrt(h)) / a;
    const float y = baoa + dist * bard;
    if (y > 0.f && y < baba) {
      const float3 n = normalize3(oa + dist * ray_dir - ba * y / baba);
      return make4(n.x, n.y, n.z, dist);
    }
    const float3 oc = (y <= 0.0f) ? oa : ob;
    b = dot3(ray_dir, oc);
    c = dot3(oc, oc) - ra * ra;
    h = b * b - c;
    if (h > 0.f) {
      dist = -b - std::sqrt(h);
      const float3 n = normalize3(ray_pos + dist * ray_dir + (oc - ray_pos));
      return make4(n.x, n.y, n.z, dist);
    }
  }
  return make4(2e15f, 2e15f, 2e15f, 2e15f);
}

// BVHRT::IntersectGraphicPrims (BVH2Common.cpp:1225-1440): one primitive per BLAS leaf — a
// sphere (point), an infinite cylinder (line), a capped cylinder (segment), a cylinder with a
// cone (arrow) or the 12 edge capsules of a box; colour from the header or the primitive.
__device__ __forceinline__ void gprim_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                                float tNear, uint32_t instId, uint32_t geomId, uint32_t a_start,
                                                CRT_Hit& hit) {
  const float EPS = 1e-5f, T_MAX = 1e15f;
  const GraphicsPrimHeader header = s.gprim_headers[g.offset.x];
  float3 color = make3(header.color[0], header.color[1], header.color[2]);
  const uint32_t ptype = header.prim_type;
  const uint32_t prim_len = graphics_prim_len(ptype);
  const uint32_t i = g.offset.y + prim_len * a_start;
  const bool has_custom_color = ptype >= GRAPH_PRIM_POINT_COLOR;
  float t = T_MAX;
  float3 norm = make3(0.f, 0.f, 0.f);
  if (ptype == GRAPH_PRIM_POINT || ptype == GRAPH_PRIM_POINT_COLOR) {
    const float4 point = s.gprim_points[i];
    if (has_custom_color) color = xyz(s.gprim_points[i + 1]);
    const float3 point3 = xyz(point);
    const float pt_radius = point.w;
    const float3 pmp = ray_pos - point3;
    const float b = 2.f * dot3(ray_dir, pmp);
    const float c = dot3(pmp, pmp) - pt_radius * pt_radius;
    float D = b * b - 4 * c;
    if (!(D >= EPS)) return;
    D = std::sqrt(D);
    t = -(b + D) * 0.5f;
    if (t < EPS) t = (-b + D) * 0.5f;
    norm = normalize3((ray_pos + t * ray_dir - point3) * 100.f);
  } else {
    const float4 point1 = s.gprim_points[i], point2 = s.gprim_points[i + 1];
    if (has_custom_color) color = xyz(s.gprim_points[i + 2]);
    const float3 point1_3 = xyz(point1);
    float3 point2_3 = xyz(point2);
    const bool arrow = ptype == GRAPH_PRIM_LINE_SEGMENT_DIR || ptype == GRAPH_PRIM_LINE_SEGMENT_DIR_COLOR;
    if (arrow) point2_3 = point2_3 * 0.71f + point1_3 * 0.29f;
    float ra = point1.w;
    if (ptype == GRAPH_PRIM_LINE || ptype == GRAPH_PRIM_LINE_SEGMENT || arrow || ptype == GRAPH_PRIM_LINE_COLOR ||
        ptype == GRAPH_PRIM_LINE_SEGMENT_COLOR) {
      const float3 oc = ray_pos - point1_3, ba = point2_3 - point1_3;
      const float baba = dot3(ba, ba), bard = dot3(ba, ray_dir), baoc = dot3(ba, oc);
      const float k2 = baba - (bard * bard);
      const float k1 = baba * dot3(oc, ray_dir) - baoc * bard;
      const float k0 = baba * dot3(oc, oc) - baoc * baoc - ra * ra * baba;
      const float D = k1 * k1 - k2 * k0;
      if (D > 0.f) {
        float dist = -(k1 + std::sqrt(D)) / k2;
        const float y = baoc + dist * bard;
        if (ptype == GRAPH_PRIM_LINE || (y > 0.f && y < baba)) {
          t = dist;
          norm = normalize3(oc + t * ray_dir - ba * y / baba);
        } else {
          dist = ((y < 0.f ? 0.f : baba) - baoc) / bard;
          if (fabs_ref(k1 + k2 * dist) < std::sqrt(D)) {
            t = dist;
            norm = normalize3(ba * sign_ref(y));
          }
        }
      }
    }
    if (arrow) {
      ra = length3(point2_3 - point1_3) * 0.15f;
      point2_3 = xyz(point2);
      const float3 pa = point2_3 * 0.7f + point1_3 * 0.3f;
      const float3 ba = point2_3 - pa, oa = ray_pos - pa;
      const float m0 = dot3(ba, ba), m1 = dot3(oa, ba), m2 = dot3(ray_dir, ba), m3 = dot3(ray_dir, oa),
                  m5 = dot3(oa, oa);
      if (m1 < 0.f) {
        const float3 tmp1 = oa * m2 - ray_dir * m1;
        if (dot3(tmp1, tmp1) < (ra * ra * m2 * m2)) {
          t = -m1 / m2;
          norm = normalize3(-1.0f * ba);
        }
      }
      const float rr = ra;
      const float hy = m0 + rr * rr;
      const float k2 = m0 * m0 - m2 * m2 * hy;
      const float k1 = m0 * m0 * m3 - m1 * m2 * hy + m0 * ra * (rr * m2);
      const float k0 = m0 * m0 * m5 - m1 * m1 * hy + m0 * ra * (rr * m1 * 2.0f - m0 * ra);
      const float D = k1 * k1 - k2 * k0;
      if (D > 0.f) {
        const float dist = -(k1 + std::sqrt(D)) / k2;
        const float y = m1 + dist * m2;
        if (y >= 0.f && y <= m0) {
          t = dist;
          norm = normalize3(m0 * (m0 * (oa + t * ray_dir) + rr * ba * ra) - ba * hy * y);
        }
      }
    }
    if (ptype == GRAPH_PRIM_BOX || ptype == GRAPH_PRIM_BOX_COLOR) {
      const float3 lo = min3(point2_3, point1_3), hi = max3(point2_3, point1_3);
      const float mm[3][2] = {{lo.x, hi.x}, {lo.y, hi.y}, {lo.z, hi.z}};
      const int verts[4] = {0, 3, 5, 6};
      for (int vi = 0; vi < 4; ++vi) {
        const int ax[3] = {int((verts[vi] & 4) != 0), int((verts[vi] & 2) != 0), int((verts[vi] & 1) != 0)};
        const float3 pa = make3(mm[0][ax[0]], mm[1][ax[1]], mm[2][ax[2]]);
        for (int axis = 0; axis < 3; ++axis) {
          float3 pb = pa;
          const float other = mm[axis][1 - ax[axis]];
          if (axis == 0) pb.x = other; else if (axis == 1) pb.y = other; else pb.z = other;
          const float4 nd = ray_capsule(ray_pos, ray_dir, pa, pb, ra);
          if (nd.w < t) {
            t = nd.w;
            norm = make3(nd.x, nd.y, nd.z);
          }
        }
      }
    }
  }
  if (t > tNear && t < hit.t) {
    const float2 e = encode_normal(norm);
    hit.t = t;
    hit.primId = a_start;
    hit.instId = instId;
    hit.geomId = geomId | (TYPE_GRAPHICS_PRIM << SH_TYPE);
    hit.coords[0] = color.x + color.y / 256.0f;
    hit.coords[1] = color.z;
    hit.coords[2] = e.x;
    hit.coords[3] = e.y;
  }
}

// eval_dist_trilinear_diff (BVH2Common.cpp:751-781): gradient of the trilinear of 8 values.
__device__ __forceinline__ float3 eval_dist_trilinear_diff(const float v[8], float3 dp) {
  const float ddx = -(1 - dp.y) * (1 - dp.z) * v[0] + -(1 - dp.y) * (dp.z) * v[1] + -(dp.y) * (1 - dp.z) * v[2] +
                    -(dp.y) * (dp.z) * v[3] + (1 - dp.y) * (1 - dp.z) * v[4] + (1 - dp.y) * (dp.z) * v[5] +
                    (dp.y) * (1 - dp.z) * v[6] + (dp.y) * (dp.z) * v[7];
  const float ddy = -(1 - dp.x) * (1 - dp.z) * v[0] + -(1 - dp.x) * (dp.z) * v[1] + (1 - dp.x) * (1 - dp.z) * v[2] +
                    (1 - dp.x) * (dp.z) * v[3] + -(dp.x) * (1 - dp.z) * v[4] + -(dp.x) * (dp.z) * v[5] +
                    (dp.x) * (1 - dp.z) * v[6] + (dp.x) * (dp.z) * v[7];
  const float ddz = -(1 - dp.x) * (1 - dp.y) * v[0] + (1 - dp.x) * (1 - dp.y) * v[1] + -(1 - dp.x) * (dp.y) * v[2] +
                    (1 - dp.x) * (dp.y) * v[3] + -(dp.x) * (1 - dp.y) * v[4] + (dp.x) * (1 - dp.y) * v[5] +
                    -(dp.x) * (dp.y) * v[6] + (dp.x) * (dp.y) * v[7];
  return make3(ddx, ddy, ddz);
}

// step (BVH2Common.cpp:784-788): 0 at edge0 rising linearly to 1 at edge1.
__device__ __forceinline__ float step_ref(float edge0, float edge1, float x) {
  return clamp_ref((x - edge0) / (edge1 - edge0), 0.0f, 1.0f);
}

// Smoothed normal of the brick-set layout with neighbour indices (ID32F_IRGB32F_IN;
// BVH2Common.cpp:958-1028): the gradients of the 8 voxels around the hit point (taken from the
// neighbouring bricks listed after the brick's colour indices where they leave the brick; a
// missing neighbour contributes zero) blended trilinearly by the hit point's offset from the
// voxel centres, normalised and encoded into coords[2..3]. As in the reference, the instance
// transform is not applied to it. Compiled only into the kernels selected by LK_SBS_SMOOTH.
__device__ __forceinline__ void sbs_smoothed_normal(const SceneView& s, const SdfSBSHeader& header, uint32_t nodeId,
                                                 uint32_t data_offset, float3 voxelPos, uint32_t v_size, float sz_inv,
                                                 float d, float3 start_q, float f_near, float3 ray_dir, CRT_Hit& hit) {
  const float beta = 0.5f;
  const float3 dp = start_q + ((hit.t - f_near) / d) * ray_dir;
  float3 normals[8];
  float values_n[8];
  for (int i = 0; i < 8; i++) normals[i] = make3(0, 0, 0);
  float3 nmq;
  nmq.x = dp.x < 0.5f ? step_ref(-beta, beta, dp.x) : step_ref(-beta, beta, dp.x - 1.0f);
  nmq.y = dp.y < 0.5f ? step_ref(-beta, beta, dp.y) : step_ref(-beta, beta, dp.y - 1.0f);
  nmq.z = dp.z < 0.5f ? step_ref(-beta, beta, dp.z) : step_ref(-beta, beta, dp.z - 1.0f);
  const int bs = int(header.brick_size);
  const int vx0 = dp.x < 0.5f ? int(voxelPos.x) - 1 : int(voxelPos.x);
  const int vy0 = dp.y < 0.5f ? int(voxelPos.y) - 1 : int(voxelPos.y);
  const int vz0 = dp.z < 0.5f ? int(voxelPos.z) - 1 : int(voxelPos.z);
  for (int i = 0; i < 8; i++) {
    const int vxi = vx0 + ((i & 4) >> 2), vyi = vy0 + ((i & 2) >> 1), vzi = vz0 + (i & 1);
    const int bx = vxi >= 0 ? (vxi < bs ? 0 : 1) : -1;
    const int by = vyi >= 0 ? (vyi < bs ? 0 : 1) : -1;
    const int bz = vzi >= 0 ? (vzi < bs ? 0 : 1) : -1;
    const uint32_t neighborId = uint32_t(9 * (bx + 1) + 3 * (by + 1) + (bz + 1));
    const float3 vpi = make3(float(vxi), float(vyi), float(vzi));
    const float3 dVoxelPos = vpi - voxelPos;
    uint32_t nnode;
    float3 nvp;
    if (neighborId == 9 + 3 + 1) {
      nnode = nodeId;
      nvp = vpi;
    } else {
      nnode = s.sbs_data[data_offset + v_size * v_size * v_size + 8 + neighborId];
      nvp = vpi - float(header.brick_size) * make3(float(bx), float(by), float(bz));
    }
    if (nnode != INVALID_IDX) {
      sbs_load_values(s, header, nnode, nvp, v_size, sz_inv, values_n);
      normals[i] = normalize3(eval_dist_trilinear_diff(values_n, dp - dVoxelPos));
    }
  }
  float3 sm = ((1 - nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[0];
  sm = sm + ((1 - nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[1];
  sm = sm + ((1 - nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[2];
  sm = sm + ((1 - nmq.x) * (nmq.y) * (nmq.z)) * normals[3];
  sm = sm + ((nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[4];
  sm = sm + ((nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[5];
  sm = sm + ((nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[6];
  sm = sm + ((nmq.x) * (nmq.y) * (nmq.z)) * normals[7];
  const float2 e = encode_normal(normalize3(sm));
  hit.coords[2] = e.x;
  hit.coords[3] = e.y;
}

struct SbsVoxelRec {
  uint32_t idx[8];
  float3 start_q;
  float t_local;
};

// Sparse-brick-set leaf: BVHRT::OctreeBrickIntersect (BVH2Common.cpp:884-1112), trilinear.
// Returns the number of voxels whose values were loaded (all voxels stepped without an
// occupancy mask). `rec` (indexed layouts only) records the hit voxel. SMOOTH compiles in the
// smoothed normals of the neighbour-index layout (applied when the preset asks for them).
// TRICUBIC takes the preset's tricubic interpolation instead: each voxel's 4^3 neighbourhood and
// the tricubic surface step; the occupancy mask (which knows only the 8 corners) is not used.
template <bool SMOOTH = false, bool TRICUBIC = false>
__device__ uint32_t sbs_brick_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir, float tNear,
                                        uint32_t instId, uint32_t geomId, uint32_t bvhNodeId, CRT_Hit& hit,
                                        SbsVoxelRec* rec = nullptr) {
  uint32_t steps = 0, loaded = 0;
  const uint32_t type = g.type;
  const uint32_t sdfId = g.offset.x;
  const uint32_t primId = bvhNodeId;
  const uint32_t nodeId = primId + s.sbs_roots[sdfId];
  const SdfSBSHeader header = s.sbs_headers[sdfId];
  const uint32_t v_size = header.brick_size + 2 * header.brick_pad + 1;
  const SdfSBSNode node = s.sbs_nodes[nodeId];
  const float px = float(node.pos_xy >> 16);
  const float py = float(node.pos_xy & 0x0000FFFF);
  const float pz = float(node.pos_z_lod_size >> 16);
  const float sz = float(node.pos_z_lod_size & 0x0000FFFF);
  const float sz_inv = 2.0f / sz;
  const float bs = float(header.brick_size);
  const float d = 2.0f / (sz * bs);
  const float3 brick_min_pos = make3(-1, -1, -1) + sz_inv * make3(px, py, pz);
  const float3 brick_max_pos = brick_min_pos + sz_inv * make3(1, 1, 1);
  const float3 inv = safe_inverse(ray_dir);
  float2 brick_fNearFar = ray_box2(ray_pos, inv, brick_min_pos, brick_max_pos);
  const float old_t = hit.t;
  const float scale = 0.5f * sz * bs;
  const uint32_t* bmask = nullptr;
  if (!TRICUBIC && s.sbs_mask) {
    const uint32_t bs_u = header.brick_size;
    bmask = s.sbs_mask + s.sbs_mask_base[sdfId] + primId * (1u + (bs_u * bs_u * bs_u + 31u) / 32u);
    if (bmask[0] == 0u) return 0;   // no occupied voxel: the march could not change the hit
    bmask += 1;
  }
  const float hi = bs - 1e-6f;
  // A ray crosses at most 3*brick_size-2 voxels of a brick; the cap only ends marches whose
  // 1e-6 advance is lost to rounding (the reference would loop forever on those).
  const uint32_t max_steps = 4 * header.brick_size + 8;
  while (brick_fNearFar.x < brick_fNearFar.y && hit.t == old_t && steps < max_steps) {
    steps++;
    const float3 hit_pos = ray_pos + brick_fNearFar.x * ray_dir;
    const float3 local_pos = (hit_pos - brick_min_pos) * scale;
    const float3 voxelPos = make3(floorf(clamp_ref(local_pos.x, 1e-6f, hi)), floorf(clamp_ref(local_pos.y, 1e-6f, hi)),
                                  floorf(clamp_ref(local_pos.z, 1e-6f, hi)));
    const float3 min_pos = brick_min_pos + d * voxelPos;
    const float3 max_pos = min_pos + d * make3(1, 1, 1);
    const float2 fNearFar = ray_box2(ray_pos, inv, min_pos, max_pos);
    if (TRICUBIC) {
      float tv[64];
      loaded++;
      const float vmin = sbs_load_tricubic(s, header, nodeId, voxelPos, v_size, sz_inv, tv);
      if (tNear < fNearFar.x && vmin <= 0.0f) {
        const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
        const float3 start_q = (start_pos - min_pos) * scale;
        const float qFar = (fNearFar.y - fNearFar.x) * scale;
        tricubic_local_surface(s, type, ray_dir, instId, geomId, tv, primId, d, 0.0f, qFar, fNearFar, start_q, hit);
      }
      brick_fNearFar.x += fmax_ref(0.0f, fNearFar.y - brick_fNearFar.x) + 1e-6f;
      continue;
    }
    bool occupied = true;
    if (bmask) {
      const uint32_t vlin = (uint32_t(voxelPos.x) * header.brick_size + uint32_t(voxelPos.y)) * header.brick_size +
                            uint32_t(voxelPos.z);
      occupied = ((bmask[vlin >> 5] >> (vlin & 31u)) & 1u) != 0u;
    }
    float values[8];
    float vmin = 1.0f;
    if (occupied) {
      loaded++;
      vmin = sbs_load_values(s, header, nodeId, voxelPos, v_size, sz_inv, values);
    }
    if (tNear < fNearFar.x && vmin <= 0.0f) {
      const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
      const float3 start_q = (start_pos - min_pos) * scale;
      const float qFar = (fNearFar.y - fNearFar.x) * scale;
      float t_local = 0.0f;
      const bool got = local_surface_intersection(s, type, ray_dir, instId, geomId, values, primId, d, 0.0f, qFar,
                                                  fNearFar, start_q, hit, rec ? &t_local : nullptr);
      if (SMOOTH && got && header.aux_data == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN &&
          s.preset.interpolation_mode == INTERPOLATION_MODE_TRILINEAR &&
          s.preset.normal_mode == NORMAL_MODE_SDF_SMOOTHED && preset_needs_normal(s.preset.render_mode))
        sbs_smoothed_normal(s, header, nodeId, node.data_offset, voxelPos, v_size, sz_inv, d, start_q, fNearFar.x,
                            ray_dir, hit);
      if (rec && got) {
        const uint32_t vx = uint32_t(voxelPos.x), vy = uint32_t(voxelPos.y), vz = uint32_t(voxelPos.z);
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint32_t vId = (vx + ((i & 4) >> 2) + header.brick_pad) * v_size * v_size +
                               (vy + ((i & 2) >> 1) + header.brick_pad) * v_size + (vz + (i & 1) + header.brick_pad);
          rec->idx[i] = s.sbs_data[node.data_offset + vId];
        }
        rec->start_q = start_q;
        rec->t_local = t_local;
      }
    }
    brick_fNearFar.x += fmax_ref(0.0f, fNearFar.y - brick_fNearFar.x) + 1e-6f;
  }
  if (hit.t < old_t) {
    const float3 pos = ray_pos + hit.t * ray_dir;
    const float3 dp = (pos - brick_min_pos) * (0.5f * sz);
    const uint32_t aux = header.aux_data;
    if (aux == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F || aux == SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN) {
      const uint32_t t_off = node.data_offset + v_size * v_size * v_size;
      float3 col = make3(0, 0, 0);
      const float w[8] = {(1 - dp.x) * (1 - dp.y) * (1 - dp.z), (1 - dp.x) * (1 - dp.y) * (dp.z),
                          (1 - dp.x) * (dp.y) * (1 - dp.z),     (1 - dp.x) * (dp.y) * (dp.z),
                          (dp.x) * (1 - dp.y) * (1 - dp.z),     (dp.x) * (1 - dp.y) * (dp.z),
                          (dp.x) * (dp.y) * (1 - dp.z),         (dp.x) * (dp.y) * (dp.z)};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t b = s.sbs_data[t_off + i];
        const float3 c = make3(s.sbs_data_f[b + 0], s.sbs_data_f[b + 1], s.sbs_data_f[b + 2]);
        col = (i == 0) ? w[0] * c : col + w[i] * c;
      }
      col = make3(clamp_ref(floorf(255.0f * col.x + 0.5f), 0.0f, 255.0f),
                  clamp_ref(floorf(255.0f * col.y + 0.5f), 0.0f, 255.0f),
                  clamp_ref(floorf(255.0f * col.z + 0.5f), 0.0f, 255.0f));
      hit.coords[0] = col.x + col.y / 256.0f;
      hit.coords[1] = col.z;
    } else if (aux == SDF_SBS_NODE_LAYOUT_DX_RGB8) {
      const uint32_t vals_per_int = 4 / header.bytes_per_value;
      const uint32_t t_off = node.data_offset + (v_size * v_size * v_size + vals_per_int - 1) / vals_per_int;
      const float w[8] = {(1 - dp.x) * (1 - dp.y) * (1 - dp.z), (1 - dp.x) * (1 - dp.y) * (dp.z),
                          (1 - dp.x) * (dp.y) * (1 - dp.z),     (1 - dp.x) * (dp.y) * (dp.z),
                          (dp.x) * (1 - dp.y) * (1 - dp.z),     (dp.x) * (1 - dp.y) * (dp.z),
                          (dp.x) * (dp.y) * (1 - dp.z),         (dp.x) * (dp.y) * (dp.z)};
      float3 col = make3(0, 0, 0);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t c = s.sbs_data[t_off + i];
        const float3 cf = make3(float((c >> 0) & 0xFF), float((c >> 8) & 0xFF), float((c >> 16) & 0xFF));
        col = (i == 0) ? w[0] * cf : col + w[i] * cf;
      }
      col = make3(clamp_ref(floorf(col.x + 0.5f), 0.0f, 255.0f), clamp_ref(floorf(col.y + 0.5f), 0.0f, 255.0f),
                  clamp_ref(floorf(col.z + 0.5f), 0.0f, 255.0f));
      hit.coords[0] = col.x + col.y / 256.0f;
      hit.coords[1] = col.z;
    } else if (aux == SDF_SBS_NODE_LAYOUT_DX_UV16) {
      const uint32_t vals_per_int = 4 / header.bytes_per_value;
      const uint32_t t_off = node.data_offset + (v_size * v_size * v_size + vals_per_int - 1) / vals_per_int;
      const float w[8] = {(1 - dp.x) * (1 - dp.y) * (1 - dp.z), (1 - dp.x) * (1 - dp.y) * (dp.z),
                          (1 - dp.x) * (dp.y) * (1 - dp.z),     (1 - dp.x) * (dp.y) * (dp.z),
                          (dp.x) * (1 - dp.y) * (1 - dp.z),     (dp.x) * (1 - dp.y) * (dp.z),
                          (dp.x) * (dp.y) * (1 - dp.z),         (dp.x) * (dp.y) * (dp.z)};
      float u = 0, v = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t c = s.sbs_data[t_off + i];
        const float tu = w[i] * (1.5259022e-5f * float(c >> 16));
        const float tv = w[i] * (1.5259022e-5f * float(c & 0xFFFF));
        u = (i == 0) ? tu : u + tu;
        v = (i == 0) ? tv : v + tv;
      }
      hit.coords[0] = u;
      hit.coords[1] = v;
    }
  }
  return loaded;
}

}  // namespace lrt

namespace lrt {

// ---- multi-octree (BVHRT::OctreeIntersect, MultiRenderer/BVH2Common.cpp:2176-2363) ----
// first_node / new_node of the parametric octree walk (utils/raytracing/intersection_common.h:58-93)
__device__ __forceinline__ int ot_first_node(float3 t0, float3 tm) {
  uint32_t answer = 0;
  if (t0.x > t0.y) {
    if (t0.x > t0.z) {
      if (t0.x > tm.y) answer |= 2;
      if (t0.x > tm.z) answer |= 1;
      return int(answer);
    }
  } else {
    if (t0.y > t0.z) {
      if (tm.x < t0.y) answer |= 4;
      if (t0.y > tm.z) answer |= 1;
      return int(answer);
    }
  }
  if (tm.x < t0.z) answer |= 4;
  if (tm.y < t0.z) answer |= 2;
  return int(answer);
}

__device__ __forceinline__ int ot_new_node(float txm, int x, float tym, int y, float tzm, int z) {
  return (txm < tym) ? (txm < tzm ? x : z) : (tym < tzm ? y : z);
}

// A stack entry of the octree walk, 8 bytes: the node and where it sits under its parent. The
// node's position is not stored: every entry's parent is an ancestor (or itself) of the node the
// walk expanded last (a depth-first stack holds deeper levels above shallower ones, and a level's
// entries all come from its most recent expansion), so it is that node's position shifted down
// to the parent's level, times two, plus the entry's octant bits.
struct OTEntry {
  uint32_t node_id;
  uint32_t meta;   // bits 0-2 octant (x<<2 | y<<1 | z), bits 3-6 level, bit 7 leaf with data (info > 0)
};
constexpr int OT_STACK = 32;
constexpr int OT_LDS = 16;   // entries kept in LDS when the caller provides room (2 words each)

// The walk's stack: entries below OT_LDS in the caller's LDS columns (entry k, word j at
// col[(2k + j) * BLOCK]) when `col` is set, the others in private memory, which only walks
// deeper than OT_LDS entries reach.
template <int BLOCK>
struct OTStack {
  lds_u32* col;
  OTEntry priv[OT_STACK];
  __device__ __forceinline__ OTEntry get(int k) const {
    if (col && k < OT_LDS) return OTEntry{col[(2 * k) * BLOCK], col[(2 * k + 1) * BLOCK]};
    return priv[k];
  }
  __device__ __forceinline__ void set(int k, const OTEntry& e) {
    if (col && k < OT_LDS) {
      col[(2 * k) * BLOCK] = e.node_id;
      col[(2 * k + 1) * BLOCK] = e.meta;
    } else {
      priv[k] = e;
    }
  }
};

// Walks the octree in front-to-back child order with mirrored coordinates (octant mask `a`)
// and stops at the first leaf whose surface is hit. `ot_col`: LDS columns for the top of the
// stack (OT_LDS x 2 words per lane, stride BLOCK; may be null: all in private memory).
// Returns the bytes read from the octree in 16-byte units (node records and voxel values).
template <int BLOCK = 64>
__device__ uint32_t multi_octree_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                           float tNear, uint32_t instId, uint32_t geomId, CRT_Hit& hit,
                                           lds_u32* ot_col = nullptr) {
  float3 pos_ray_pos = ray_pos;
  float3 pos_ray_dir = ray_dir;
  uint32_t a = 0;
  if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
  if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
  if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
  pos_ray_pos = -1.0f * pos_ray_pos;
  const float3 inv = safe_inverse(pos_ray_dir);
  const float3 t0_ = pos_ray_pos * inv - inv;
  const float3 t1_ = pos_ray_pos * inv + inv;
  const float3 l_ = t1_ - t0_;
  // nn_indices of the reference, packed as x | y << 4 | z << 8
  const uint32_t nn[8] = {0x124u, 0x835u, 0x386u, 0x887u, 0x568u, 0x878u, 0x788u, 0x888u};

  OTStack<BLOCK> stack;
  stack.col = ot_col;
  int top = 0;
  uint32_t units = 0;
  const float old_t = hit.t;
  const uint32_t type = g.type;
  const uint32_t root = g.offset.x;
  // position and level of the node expanded last (the root's until the first expansion)
  uint32_t cx = 0, cy = 0, cz = 0, cl = 0;
  stack.set(0, OTEntry{root, 0u});
  while (top >= 0) {
    const OTEntry e = stack.get(top);
    const uint32_t level = (e.meta >> 3) & 0xFu;
    uint32_t px = 0, py = 0, pz = 0;
    if (level > 0) {
      const uint32_t sh = cl - (level - 1);
      px = ((cx >> sh) << 1) | ((e.meta >> 2) & 1u);
      py = ((cy >> sh) << 1) | ((e.meta >> 1) & 1u);
      pz = ((cz >> sh) << 1) | (e.meta & 1u);
    }
    const uint32_t level_sz = 1u << level;
    float d = 1.0f / float(level_sz);
    const float3 p_f = make3(float(px), float(py), float(pz));
    const float3 t0 = t0_ + d * p_f * l_;
    const float3 t1 = t0_ + d * (p_f + 1.0f) * l_;
    const float t_out = fmin_ref(fmin_ref(t1.x, t1.y), t1.z);
    if (t_out < tNear) {
      top--;
      continue;
    }
    bool is_leaf = (e.meta & 0x80u) != 0;
    SdfMultiOctreeNode n;
    bool have_node = false;
    if (!is_leaf) {
      const float t = s.preset.fixed_lod > 0 ? 1.0f : fmax_ref(t0.x, fmax_ref(t0.y, t0.z));
      const float target_lod_size = (powf(2.0f, s.preset.level_of_detail) - 0.01f) / t;
      if (float(level_sz) > target_lod_size) {
        n = s.multi_nodes[e.node_id];
        have_node = true;
        units++;
        is_leaf = true;
        if (n.voxel_count == 0) {
          top--;
          continue;
        }
      }
    }
    if (is_leaf) {
      const float tmin = fmax_ref(t0.x, fmax_ref(t0.y, t0.z));
      const float tmax = fmin_ref(t1.x, fmin_ref(t1.y, t1.z));
      const uint32_t p_mask = level_sz - 1;
      const uint32_t rx = (a & 4) ? (~px & p_mask) : px;
      const uint32_t ry = (a & 2) ? (~py & p_mask) : py;
      const uint32_t rz = (a & 1) ? (~pz & p_mask) : pz;
      const uint32_t node_id = e.node_id;
      const float sz = 0.5f * float(level_sz);
      d = 1.0f / sz;
      const float3 min_pos = make3(-1, -1, -1) + d * make3(float(rx), float(ry), float(rz));
      float2 fNearFar;
      fNearFar.x = tmin;
      fNearFar.y = tmax;
      const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
      const float3 start_q = sz * (start_pos - min_pos);
      const float qNear = fmax_ref(0.0f, sz * (tNear - fNearFar.x));
      const float qFar = sz * (fNearFar.y - fNearFar.x);
      if (!have_node) n = s.multi_nodes[node_id];
      const uint32_t info = n.data_offset;
      units += 1 + 2 * n.voxel_count;
      const bool surface = (n.flags & MULTI_OCTREE_FLAG_SURFACE) != 0;
      const uint32_t primId = node_id - root;
      for (uint32_t v = 0; v < n.voxel_count; ++v) {
        float values[8];
#pragma unroll
        for (int i = 0; i < 8; i++) values[i] = s.multi_values[info + 8 * v + i];
        local_surface_intersection<true>(s, TYPE_SDF_MULTI_OCTREE, ray_dir, instId, geomId, values, primId, d, qNear, qFar,
                                   fNearFar, start_q, hit, nullptr, surface);
      }
      if (s.preset.render_mode == MULTI_RENDER_MODE_MO_VOXEL_COUNT) {
        if (hit.t < old_t) hit.primId = n.voxel_count;
      } else if (s.preset.render_mode == MULTI_RENDER_MODE_MO_FLAGS) {
        if (hit.t < old_t) hit.primId = n.flags;
      }
      top = (hit.t < old_t) ? -1 : top - 1;
    } else {
      OTEntry buf[4];
      int buf_top = 0;
      const float3 tm = 0.5f * (t0 + t1);
      int curr = ot_first_node(t0, tm);
      const uint32_t children = s.multi_nodes[e.node_id].children_offset;
      const uint32_t child_level = (level + 1) << 3;
      do {
        const uint32_t child = uint32_t(curr) ^ a;
        const SdfMultiOctreeNode cn = s.multi_nodes[children + child];
        units++;
        const bool child_is_node = cn.children_offset > 0;
        if ((cn.voxel_count > 0 || child_is_node) && buf_top < 4) {
          // a leaf entry of the reference carries info = its data offset and counts as a leaf
          // only when that is non-zero (info > 0); a node's info is 0
          const bool leaf = !child_is_node && cn.data_offset > 0;
          buf[buf_top].node_id = children + child;
          buf[buf_top].meta = uint32_t(curr) | child_level | (leaf ? 0x80u : 0u);
          buf_top++;
        }
        const uint32_t ni = nn[curr];
        curr = ot_new_node((curr & 4) ? t1.x : tm.x, int(ni & 0xF), (curr & 2) ? t1.y : tm.y, int((ni >> 4) & 0xF),
                           (curr & 1) ? t1.z : tm.z, int(ni >> 8));
      } while (curr < 8);
      cx = px;
      cy = py;
      cz = pz;
      cl = level;
      // children replace the current entry, the first one on top; entries past the stack are dropped
      for (int i = 0; i < buf_top; i++)
        if (top + i < OT_STACK) stack.set(top + i, buf[buf_top - i - 1]);
      top += buf_top - 1;
      if (top >= OT_STACK) top = OT_STACK - 1;
    }
  }
  return units;
}

}  // namespace lrt

namespace lrt {

// ---- compact octree COctreeV3 (BVHRT::COctreeIntersectV3, MultiRenderer/BVH2Common.cpp:1637-2173) ----

// COctreeV3_LoadDistanceValues* (BVH2Common.cpp:1637-1851): the 8 corner values of voxel
// (vx, vy, vz) of the brick at word `o`, decoded through the leaf's rotation (index modifiers M)
// and its additive shift. Returns the smallest value (grid leaves), -1 (a present voxel of a
// bit-pack or slice leaf) or 1e6 (an absent voxel). A voxel outside the padded brick (only the
// smoothed normal asks for one, and only without padding, where the reference reads outside
// the brick) counts as absent.
__device__ __forceinline__ float coct_load_values(const uint32_t* D, const int4* M, const CoctGeom& h,
                                                  uint32_t leaf_type, uint32_t o, int vx, int vy, int vz,
                                                  uint32_t v_size, float sz_inv, uint32_t code, float values[8]) {
  const int pad = int(h.brick_pad);
  const int p_size = int(h.brick_size + 2 * h.brick_pad);
  if (vx + pad < 0 || vy + pad < 0 || vz + pad < 0 || vx + pad >= p_size || vy + pad >= p_size || vz + pad >= p_size)
    return 1e6f;
  if (leaf_type < COCTREE_LEAF_TYPE_GRID || leaf_type > COCTREE_LEAF_TYPE_SLICES) return 1e6f;
  const uint32_t rotIdx = code & h.rot_mask;
  const int4 mv = M[rotIdx];
  const uint32_t bits = h.bits_per_value;
  const uint32_t vals_per_int = 32u / bits;
  const uint32_t max_val = bits == 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  const float add_transform =
      h.sim_compression > 0
          ? 2.0f * 1.73205081f * sz_inv * (2.0f * (float(code & h.add_mask) / float(h.add_mask)) - 1.0f)
          : 0.0f;
  uint32_t off_3, off_5;
  uint32_t off_1 = 0, off_2 = 0, slice_uints = 0;
  if (leaf_type == COCTREE_LEAF_TYPE_GRID) {
    off_3 = 0;
  } else {
    const uint32_t pfs = (uint32_t(p_size * p_size * p_size) + 31u) / 32u;
    const int4 mp = M[COCTREE_ROT_COUNT + rotIdx];
    const uint32_t pflag = uint32_t(mp.x * (vx + pad) + mp.y * (vy + pad) + mp.z * (vz + pad) + mp.w);
    if ((D[o + pflag / 32u] & (1u << (pflag % 32u))) == 0u) return 1e6f;
    off_1 = pfs;
    if (leaf_type == COCTREE_LEAF_TYPE_BIT_PACK) {
      off_3 = off_1 + (v_size * v_size * v_size + 31u) / 32u;
    } else {
      slice_uints = (v_size * v_size + 31u) / 32u;
      off_2 = off_1 + v_size * slice_uints;
      off_3 = off_2 + (v_size + 1u) / 2u;
    }
  }
  off_5 = off_3 + 2u + 8u * h.uv_size;
  const float min_val = -float(D[o + off_3]) / float(0xFFFFFFFFu) + add_transform;
  const float range = (float(D[o + off_3 + 1]) / float(0xFFFFFFFFu)) / float(max_val);
  float vmin = 1e6f;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int cx = vx + pad + ((i & 4) >> 2), cy = vy + pad + ((i & 2) >> 1), cz = vz + pad + (i & 1);
    const uint32_t lid = uint32_t(mv.x * cx + mv.y * cy + mv.z * cz + mv.w);
    uint32_t vId;
    if (leaf_type == COCTREE_LEAF_TYPE_GRID) {
      vId = lid;
    } else {
      uint32_t localId = lid, base = 0, w = off_1;
      if (leaf_type == COCTREE_LEAF_TYPE_SLICES) {
        const uint32_t sliceId = lid / (v_size * v_size);
        localId = lid % (v_size * v_size);
        w = off_1 + slice_uints * sliceId;
        base = (D[o + off_2 + sliceId / 2u] >> (16u * (sliceId % 2u))) & 0xFFFFu;
      }
      uint32_t b0 = D[o + w];
      b0 = localId > 31u ? b0 : (b0 & ((1u << localId) - 1u));
      const uint32_t b1 = localId > 32u ? (D[o + w + 1] & ((1u << ((localId - 32u) & 31u)) - 1u)) : 0u;
      vId = base + uint32_t(__popc(b0)) + uint32_t(__popc(b1));
    }
    const uint32_t dist = (D[o + off_5 + vId / vals_per_int] >> (bits * (vId % vals_per_int))) & max_val;
    values[i] = min_val + range * float(dist);
    vmin = fmin_ref(values[i], vmin);
  }
  return leaf_type == COCTREE_LEAF_TYPE_GRID ? vmin : -1.0f;
}

// COctreeV3_BrickIntersect (BVH2Common.cpp:1853-2003): marches the voxels of one leaf brick
// (position p at level size sz) along the ray; intersects voxels with surface like a frame-octree
// node (the hit reports TYPE_SDF_FRAME_OCTREE and the brick's word offset as primId), then the
// smoothed normal and the texture coordinates. The march ends after 4 bs + 8 voxels, which only
// cuts marches whose 1e-5 advance is lost to rounding (the reference would not end them).
__device__ __forceinline__ void coct_brick_intersect(const SceneView& s, const uint32_t* D, const int4* M,
                                                     const CoctGeom& h, float3 ray_pos, float3 ray_dir, float tNear,
                                                     uint32_t instId, uint32_t geomId, uint32_t o, float3 p, float sz,
                                                     uint32_t info, CRT_Hit& hit) {
  float values[8];
  const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
  const uint32_t leaf_type = info & COCTREE_LEAF_TYPE_MASK;
  const uint32_t code = info >> COCTREE_LEAF_TYPE_BITS;
  const float bs = float(h.brick_size);
  const float sz_inv = 2.0f / sz;
  const float d = 2.0f / (sz * bs);
  const float3 brick_min_pos = make3(-1, -1, -1) + sz_inv * p;
  const float3 brick_max_pos = brick_min_pos + sz_inv * make3(1, 1, 1);
  const float3 inv = safe_inverse(ray_dir);
  float2 bnf = ray_box2(ray_pos, inv, brick_min_pos, brick_max_pos);
  bnf.x = fmax_ref(tNear, bnf.x);
  const float old_t = hit.t;
  const float scale = 0.5f * sz * bs;
  const float hi = bs - 1e-6f;
  const uint32_t max_steps = 4 * h.brick_size + 8;
  uint32_t steps = 0;
  while (bnf.x < bnf.y && hit.t == old_t && steps < max_steps) {
    steps++;
    const float3 hit_pos = ray_pos + bnf.x * ray_dir;
    const float3 local_pos = (hit_pos - brick_min_pos) * scale;
    const float3 voxelPos = make3(floorf(clamp_ref(local_pos.x, 1e-6f, hi)), floorf(clamp_ref(local_pos.y, 1e-6f, hi)),
                                  floorf(clamp_ref(local_pos.z, 1e-6f, hi)));
    const float3 min_pos = brick_min_pos + d * voxelPos;
    const float3 max_pos = min_pos + d * make3(1, 1, 1);
    const int vx = int(voxelPos.x), vy = int(voxelPos.y), vz = int(voxelPos.z);
    const float vmin = coct_load_values(D, M, h, leaf_type, o, vx, vy, vz, v_size, sz_inv, code, values);
    const float2 fNearFar = ray_box2(ray_pos, inv, min_pos, max_pos);
    if (tNear < fNearFar.x && vmin <= 0.0f) {
      const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
      const float3 start_q = (start_pos - min_pos) * scale;
      const float qFar = (fNearFar.y - fNearFar.x) * scale;
      local_surface_intersection(s, TYPE_SDF_FRAME_OCTREE, ray_dir, instId, geomId, values, o, d, 0.0f, qFar, fNearFar,
                                 start_q, hit);
      if (s.preset.normal_mode == NORMAL_MODE_SDF_SMOOTHED && preset_needs_normal(s.preset.render_mode) &&
          hit.t < old_t) {
        const float beta = 0.5f;
        const float3 dp = start_q + ((hit.t - fNearFar.x) / d) * ray_dir;
        float3 normals[8];
        float values_n[8];
        float3 nmq;
        nmq.x = dp.x < 0.5f ? step_ref(-beta, beta, dp.x) : step_ref(-beta, beta, dp.x - 1.0f);
        nmq.y = dp.y < 0.5f ? step_ref(-beta, beta, dp.y) : step_ref(-beta, beta, dp.y - 1.0f);
        nmq.z = dp.z < 0.5f ? step_ref(-beta, beta, dp.z) : step_ref(-beta, beta, dp.z - 1.0f);
        const int x0 = dp.x < 0.5f ? vx - 1 : vx, y0 = dp.y < 0.5f ? vy - 1 : vy, z0 = dp.z < 0.5f ? vz - 1 : vz;
        for (int i = 0; i < 8; i++) {
          normals[i] = make3(0, 0, 0);
          const int xi = x0 + ((i & 4) >> 2), yi = y0 + ((i & 2) >> 1), zi = z0 + (i & 1);
          const float3 dVoxelPos = make3(float(xi), float(yi), float(zi)) - voxelPos;
          const float vmin_n = coct_load_values(D, M, h, leaf_type, o, xi, yi, zi, v_size, sz_inv, code, values_n);
          if (vmin_n <= 0.0f) normals[i] = normalize3(eval_dist_trilinear_diff(values_n, dp - dVoxelPos));
        }
        float3 sm = ((1 - nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[0];
        sm = sm + ((1 - nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[1];
        sm = sm + ((1 - nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[2];
        sm = sm + ((1 - nmq.x) * (nmq.y) * (nmq.z)) * normals[3];
        sm = sm + ((nmq.x) * (1 - nmq.y) * (1 - nmq.z)) * normals[4];
        sm = sm + ((nmq.x) * (1 - nmq.y) * (nmq.z)) * normals[5];
        sm = sm + ((nmq.x) * (nmq.y) * (1 - nmq.z)) * normals[6];
        sm = sm + ((nmq.x) * (nmq.y) * (nmq.z)) * normals[7];
        float m[16];
        load_mat(s.instances[instId].transformInvTransposed, m);
        const float2 e = encode_normal(normalize3(matmul4x3(m, sm)));
        hit.coords[2] = e.x;
        hit.coords[3] = e.y;
      }
    }
    bnf.x += fmax_ref(0.0f, fNearFar.y - bnf.x) + 1e-5f;
  }
  if (hit.t < old_t && h.uv_size > 0) {
    // texture coordinates sit where a slice leaf keeps them, whatever the leaf type
    const uint32_t p_size = h.brick_size + 2 * h.brick_pad;
    const uint32_t slice_uints = (v_size * v_size + 31u) / 32u;
    const uint32_t off_4 = (p_size * p_size * p_size + 31u) / 32u + v_size * slice_uints + (v_size + 1u) / 2u + 2u;
    const float3 pos = ray_pos + hit.t * ray_dir;
    const float3 dp = (pos - brick_min_pos) * (0.5f * sz);
    const float w[8] = {(1 - dp.x) * (1 - dp.y) * (1 - dp.z), (1 - dp.x) * (1 - dp.y) * (dp.z),
                        (1 - dp.x) * (dp.y) * (1 - dp.z),     (1 - dp.x) * (dp.y) * (dp.z),
                        (dp.x) * (1 - dp.y) * (1 - dp.z),     (dp.x) * (1 - dp.y) * (dp.z),
                        (dp.x) * (dp.y) * (1 - dp.z),         (dp.x) * (dp.y) * (dp.z)};
    float u = 0, v = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t c = D[o + off_4 + i];
      const float tu = w[i] * (1.5259022e-5f * float(c >> 16));
      const float tv = w[i] * (1.5259022e-5f * float(c & 0xFFFF));
      u = (i == 0) ? tu : u + tu;
      v = (i == 0) ? tv : v + tv;
    }
    hit.coords[0] = u;
    hit.coords[1] = v;
  }
}

// COctreeIntersectV3 (BVH2Common.cpp:2005-2173): front-to-back walk of the octree below the BLAS
// box `bvhNodeId` (its node, level size 2 / box width and position), with mirrored coordinates
// (octant mask a), optional LOD leaves and the first hitting leaf ending the walk. Stack entries
// are the reference's (node, info, packed position and level size); entries past 32 are dropped.
__device__ __forceinline__ void coct_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                            float tNear, uint32_t instId, uint32_t geomId, uint32_t bvhNodeId,
                                            CRT_Hit& hit) {
  const CoctGeom h = s.coct_geoms[g.offset.x];
  const uint32_t* D = s.coct_data + h.base;
  const int4* M = s.coct_mods + h.mods;
  float3 pos_ray_pos = ray_pos;
  float3 pos_ray_dir = ray_dir;
  uint32_t a = 0;
  if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
  if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
  if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
  const BVHNode box = load_node(s.orig_nodes + g.offset.y + bvhNodeId);
  const float3 min_pos = box.boxMin;
  const float3 max_pos = box.boxMax;
  const uint32_t start_sz = uint32_t(2.0f / (max_pos.x - min_pos.x));
  const float3 sp = float(start_sz) * 0.5f * (min_pos - make3(-1, -1, -1));
  const uint32_t spx = uint32_t(sp.x), spy = uint32_t(sp.y), spz = uint32_t(sp.z);
  uint32_t p_mask = start_sz - 1;
  const uint32_t sx = (a & 4) ? (~spx & p_mask) : spx;
  const uint32_t sy = (a & 2) ? (~spy & p_mask) : spy;
  const uint32_t sz0 = (a & 1) ? (~spz & p_mask) : spz;
  pos_ray_pos = -1.0f * pos_ray_pos;
  const float3 inv = safe_inverse(pos_ray_dir);
  const float3 t0_ = pos_ray_pos * inv - inv;
  const float3 t1_ = pos_ray_pos * inv + inv;
  const float3 l_ = t1_ - t0_;
  // nn_indices of the reference, packed as x | y << 4 | z << 8
  const uint32_t nn[8] = {0x124u, 0x835u, 0x386u, 0x887u, 0x568u, 0x878u, 0x788u, 0x888u};
  constexpr int COCT_STACK = 32;
  uint32_t st_node[COCT_STACK], st_info[COCT_STACK], st_pxy[COCT_STACK], st_pzs[COCT_STACK];
  int top = 0;
  const float old_t = hit.t;
  st_node[0] = box.leftOffset;
  st_info[0] = 0;
  st_pxy[0] = (sx << 16) | sy;
  st_pzs[0] = (sz0 << 16) | start_sz;
  while (top >= 0) {
    uint32_t nodeId = st_node[top], info = st_info[top];
    const uint32_t pxy = st_pxy[top], pzs = st_pzs[top];
    const uint32_t level_sz = pzs & 0xFFFFu;
    const uint32_t px = pxy >> 16, py = pxy & 0xFFFFu, pz = pzs >> 16;
    const float d = 1.0f / float(level_sz);
    const float3 p_f = make3(float(px), float(py), float(pz));
    const float3 t0 = t0_ + d * p_f * l_;
    const float3 t1 = t0_ + d * (p_f + 1.0f) * l_;
    const float t_out = fmin_ref(fmin_ref(t1.x, t1.y), t1.z);
    if (t_out < tNear) {
      top--;
      continue;
    }
    bool is_leaf = (info & COCTREE_LEAF_TYPE_MASK) != COCTREE_LEAF_TYPE_NOT_A_LEAF;
    if (h.lods > 0 && !is_leaf) {
      const float t = s.preset.fixed_lod > 0 ? 1.0f : fmax_ref(t0.x, fmax_ref(t0.y, t0.z));
      const float target_lod_size = (powf(2.0f, s.preset.level_of_detail) - 0.01f) / t;
      if (float(level_sz * h.brick_size) > target_lod_size) {
        const uint32_t lodLeafId = D[nodeId + 1];
        const uint32_t type = (D[nodeId] >> COCTREE_LOD_LEAF_TYPE_SHIFT) & COCTREE_LEAF_TYPE_MASK;
        const uint32_t infoOffset = nodeId + 1 + h.trans_off;
        info = (D[infoOffset] << COCTREE_LEAF_TYPE_BITS) | type;
        nodeId = lodLeafId;
        is_leaf = true;
        if (type == COCTREE_LEAF_TYPE_NOT_A_LEAF) {
          top--;
          continue;
        }
      }
    }
    if (is_leaf) {
      p_mask = level_sz - 1;
      const uint32_t rx = (a & 4) ? (~px & p_mask) : px;
      const uint32_t ry = (a & 2) ? (~py & p_mask) : py;
      const uint32_t rz = (a & 1) ? (~pz & p_mask) : pz;
      const uint32_t leafOffset = (nodeId & h.idx_mask) >> h.idx_sh;
      coct_brick_intersect(s, D, M, h, ray_pos, ray_dir, tNear, instId, geomId, leafOffset,
                           make3(float(rx), float(ry), float(rz)), float(level_sz), info, hit);
      if (hit.t < old_t) {
        top = -1;
        if (s.preset.render_mode == MULTI_RENDER_MODE_LOD) hit.primId = uint32_t(log2f(float(level_sz * h.brick_size)));
      } else {
        top--;
      }
    } else {
      uint32_t b_node[4], b_info[4], b_pxy[4], b_pzs[4];
      int buf_top = 0;
      const float3 tm = 0.5f * (t0 + t1);
      int curr = ot_first_node(t0, tm);
      const uint32_t childrenInfo = D[nodeId];
      do {
        const uint32_t childNode = uint32_t(curr) ^ a;
        if ((childrenInfo & (1u << childNode)) != 0u && buf_top < 4) {
          const uint32_t localChildPos = h.lods + uint32_t(__popc(childrenInfo & ((1u << childNode) - 1u)));
          const uint32_t childPos = 1 + h.uints_per_child_info * localChildPos;
          const uint32_t childType = (childrenInfo >> (8 + COCTREE_LEAF_TYPE_BITS * childNode)) & COCTREE_LEAF_TYPE_MASK;
          b_node[buf_top] = D[nodeId + childPos];
          b_info[buf_top] = (D[nodeId + childPos + h.trans_off] << COCTREE_LEAF_TYPE_BITS) | childType;
          b_pxy[buf_top] = (pxy << 1) | (((uint32_t(curr) & 4u) << 14) | ((uint32_t(curr) & 2u) >> 1));
          b_pzs[buf_top] = (pzs << 1) | ((uint32_t(curr) & 1u) << 16);
          buf_top++;
        }
        const uint32_t ni = nn[curr];
        curr = ot_new_node((curr & 4) ? t1.x : tm.x, int(ni & 0xF), (curr & 2) ? t1.y : tm.y, int((ni >> 4) & 0xF),
                           (curr & 1) ? t1.z : tm.z, int(ni >> 8));
      } while (curr < 8);
      for (int i = 0; i < buf_top; i++) {
        const int k = top + i;
        if (k < COCT_STACK) {
          st_node[k] = b_node[buf_top - i - 1];
          st_info[k] = b_info[buf_top - i - 1];
          st_pxy[k] = b_pxy[buf_top - i - 1];
          st_pzs[k] = b_pzs[buf_top - i - 1];
        }
      }
      top += buf_top - 1;
      if (top >= COCT_STACK) top = COCT_STACK - 1;
    }
  }
}

}  // namespace lrt

namespace lrt {

// m_SdfCompactOctreeRotTransforms[r] * (p, 1) (BVH2Common.cpp:2501, 2742; the matrices of
// octahedral_groups.cpp:37-101): rotation r turns the lattice [0, vt - 1]^3 about its centre. Read
// off the modifier of rotation r on a lattice of v points per axis: input axis i with coefficient
// +-v^k lands on output axis 2 - k, and a mirrored axis maps x to (vt - 1) - x. Each row of the
// reference's matrix has one non-zero term besides the translation, so the float result does not
// depend on the order of its sum.
__device__ __forceinline__ float3 rot_transform_point(int4 m, int v, float vt1, float3 p) {
  const float in[3] = {p.x, p.y, p.z};
  const int c[3] = {m.x, m.y, m.z};
  float out[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int a = c[i] < 0 ? -c[i] : c[i];
    const int o = a == 1 ? 2 : (a == v ? 1 : 0);
    out[o] = c[i] < 0 ? vt1 - in[i] : in[i];
  }
  return make3(out[0], out[1], out[2]);
}

// Channel render modes' quotients of a hit point dp turned by a rotation (BVH2Common.cpp:2492-2505,
// 2731-2747): clamp(65535 dp_rot) as (y << 16 | x, z).
__device__ __forceinline__ void channel_quotients_rotated(CRT_Hit& hit, int4 m, int v, float vt1, float3 dp) {
  float3 r = rot_transform_point(m, v, vt1, dp);
  r = make3(clamp_ref(65535.0f * r.x, 0.0f, 65535.0f), clamp_ref(65535.0f * r.y, 0.0f, 65535.0f),
            clamp_ref(65535.0f * r.z, 0.0f, 65535.0f));
  const int qx = int(r.x), qy = int(r.y), qz = int(r.z);
  hit.coords[0] = __int_as_float((qy << 16) | qx);
  hit.coords[1] = __int_as_float(qz);
}

// ---- octree DAG (BVHRT::DAGIntersect, MultiRenderer/BVH2Common.cpp:2365-2557) ----
// Depth-first walk from the root; each stack entry carries the cumulative rotation of its node
// (composed through the rotation table along the path). Children are pushed far to near by
// marching the node's child grid backwards along the ray, so the nearest ends on top; leaves
// (and inner nodes coarser than the level of detail) march their brick and intersect every
// surface the node lists. Stack pushes past 32 entries are dropped (the reference writes past
// its array); a brick march stops after 4 bs + 8 voxels and a child march after 256 cells per
// grid step, which only cuts marches whose advance is lost to rounding.
__device__ __forceinline__ void dag_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                              float tNear, uint32_t instId, uint32_t geomId, CRT_Hit& hit) {
  const DagGeom dg = s.dag_geoms[g.offset.y];
  const int4* Mb = s.dag_mods + dg.mods;                      // brick value lattice
  const int4* Mg = s.dag_mods + dg.mods + COCTREE_ROT_COUNT;  // node grid
  const float3 ray_dir_inv = safe_inverse(ray_dir);
  constexpr int DAG_STACK = 32;
  uint32_t st_node[DAG_STACK], st_info[DAG_STACK], st_pxy[DAG_STACK], st_pzs[DAG_STACK];
  int top = 0;
  const float old_t = hit.t;
  const uint32_t type = g.type;
  const uint32_t bs = dg.brick_size;
  const uint32_t v_size = dg.brick_size + 2 * dg.brick_pad + 1;
  const int v_sz = int(dg.node_grid_size);
  float2 fNearFar = float2{0.0f, 0.0f};
  st_node[0] = dg.root;
  st_info[0] = 0;
  st_pxy[0] = 0;
  st_pzs[0] = 1;
  while (top >= 0) {
    const uint32_t nodeId = st_node[top], info = st_info[top], pxy = st_pxy[top], pzs = st_pzs[top];
    top--;
    const SdfDAGNode node = s.dag_nodes[nodeId];
    const uint32_t level_sz = pzs & 0xFFFFu;
    const float level_sz_f = float(level_sz);
    const float3 p = make3(float(pxy >> 16), float(pxy & 0xFFFFu), float(pzs >> 16));
    const float sz_inv = 2.0f / level_sz_f;
    const float3 brick_min_pos = make3(-1, -1, -1) + sz_inv * p;
    const float3 brick_max_pos = brick_min_pos + sz_inv * make3(1, 1, 1);
    float2 bnf = ray_box2(ray_pos, ray_dir_inv, brick_min_pos, brick_max_pos);
    bnf.x = fmax_ref(tNear, bnf.x);
    if (bnf.x >= bnf.y) continue;
    bool is_leaf = node.children_edges_offset == 0;
    if (!is_leaf) {
      const float t = s.preset.fixed_lod > 0 ? 1.0f : bnf.x;
      const float target_lod_size = (powf(2.0f, s.preset.level_of_detail) - 0.01f) / t;
      is_leaf = float(level_sz) > target_lod_size;
    }
    if (is_leaf) {
      const uint32_t surface_count = node.voxel_count_flags & 0xFFFFu;
      if (surface_count == 0 || node.data_edges_offset == 0) continue;
      const float sz = level_sz_f;
      const float d = 2.0f / (sz * float(bs));
      const bool is_surface = ((node.voxel_count_flags >> 16) & 1u) != 0;
      const uint32_t primId = node.channels_edge.child_offset;
      const float scale = 0.5f * sz * float(bs);
      const float hi = float(bs) - 1e-6f;
      const uint32_t max_steps = 4 * bs + 8;
      uint32_t steps = 0;
      while (bnf.x < bnf.y && hit.t == old_t && steps < max_steps) {
        steps++;
        const float3 hit_pos = ray_pos + bnf.x * ray_dir;
        const float3 local_pos = (hit_pos - brick_min_pos) * scale;
        const float3 voxelPos = make3(floorf(clamp_ref(local_pos.x, 1e-6f, hi)), floorf(clamp_ref(local_pos.y, 1e-6f, hi)),
                                      floorf(clamp_ref(local_pos.z, 1e-6f, hi)));
        const float3 min_pos = brick_min_pos + d * voxelPos;
        const float3 max_pos = min_pos + d * make3(1, 1, 1);
        const int vx = int(voxelPos.x), vy = int(voxelPos.y), vz = int(voxelPos.z);
        for (uint32_t s_id = 0; s_id < surface_count && hit.t == old_t; ++s_id) {
          const SdfDAGDataEdge e = s.dag_data_edges[node.data_edges_offset + s_id];
          const int4 m = Mb[s.rot_add[info * COCTREE_ROT_COUNT + e.rotation_id]];
          float values[8];
          float vmin = 1.0f;
          float vmax = is_surface ? -1.0f : 1.0f;
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const int cx = vx + ((i & 4) >> 2), cy = vy + ((i & 2) >> 1), cz = vz + (i & 1);
            const uint32_t vId0 = uint32_t(m.x * cx + m.y * cy + m.z * cz + m.w);
            const float val = s.dag_distances[e.data_offset + vId0] + e.add;
            values[i] = val - s.preset.compact_sdf_eps * d;
            vmin = fmin_ref(vmin, val);
            vmax = fmax_ref(vmax, val);
          }
          fNearFar = ray_box2(ray_pos, safe_inverse(ray_dir), min_pos, max_pos);
          float3 start_q = make3(0.0f, 0.0f, 0.0f);
          if (fNearFar.x < fNearFar.y && vmin <= 0.0f && vmax >= 0.0f) {
            const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
            const float vox_sz = 0.5f * sz * float(bs);
            start_q = vox_sz * (start_pos - min_pos);
            const float qNear = fmax_ref(0.0f, vox_sz * (tNear - fNearFar.x));
            const float qFar = vox_sz * (fNearFar.y - fNearFar.x);
            local_surface_intersection<true>(s, type, ray_dir, instId, geomId, values, primId, d, qNear, qFar, fNearFar,
                                       start_q, hit, nullptr, is_surface);
          }
          // channel render modes: the quotients turned by the node's channel rotation, so that the
          // resolve reads the channel data as stored (BVH2Common.cpp:2492-2505)
          if (hit.t != old_t && (s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL ||
                                 s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL_LAMBERT)) {
            const uint32_t crot = s.rot_add[info * COCTREE_ROT_COUNT + node.channels_edge.rotation_id];
            const float tl = (hit.t - fNearFar.x) / d;
            channel_quotients_rotated(hit, Mb[crot], int(v_size), float(bs), start_q + tl * ray_dir);
          }
        }
        bnf.x += fmax_ref(0.0f, fNearFar.y - bnf.x) + 1e-5f;
        if (hit.t < old_t) top = -1;
      }
    } else {
      // children, pushed far to near by marching the node's grid backwards along the ray
      const float t_neg = bnf.y + tNear + 1;
      const float3 neg_pos = ray_pos + t_neg * ray_dir;
      const float3 neg_dir_inv = make3(-ray_dir_inv.x, -ray_dir_inv.y, -ray_dir_inv.z);
      const float3 neg_dir = make3(-ray_dir.x, -ray_dir.y, -ray_dir.z);
      float2 nbnf = float2{t_neg - bnf.y, t_neg - bnf.x};
      const float d = 2.0f / (level_sz_f * float(v_sz));
      const float hi = float(v_sz) - 1e-6f;
      const float scale = 0.5f * level_sz_f * float(v_sz);
      uint32_t steps = 0;
      // every step advances at least 0.01 of a cell, so a march ends within ~180 v_sz steps; the cap
      // only ends marches whose advance is lost to rounding
      while (nbnf.x < nbnf.y && steps < uint32_t(256 * v_sz)) {
        steps++;
        const float3 hit_pos = neg_pos + nbnf.x * neg_dir;
        const float3 local_pos = (hit_pos - brick_min_pos) * scale;
        const float3 voxelPos = make3(floorf(clamp_ref(local_pos.x, 1e-6f, hi)), floorf(clamp_ref(local_pos.y, 1e-6f, hi)),
                                      floorf(clamp_ref(local_pos.z, 1e-6f, hi)));
        const int4 m = Mg[info];
        const uint32_t voxId = uint32_t(m.x * int(voxelPos.x) + m.y * int(voxelPos.y) + m.z * int(voxelPos.z) + m.w);
        const float3 min_pos = brick_min_pos + d * voxelPos;
        const float3 max_pos = min_pos + d * make3(1, 1, 1);
        const float2 cnf = ray_box2(neg_pos, neg_dir_inv, min_pos, max_pos);
        const SdfDAGChildEdge edge = s.dag_child_edges[node.children_edges_offset + voxId];
        if (tNear < cnf.x && edge.child_offset > 0) {
          if (top + 1 < DAG_STACK) {
            top++;
            st_node[top] = edge.child_offset;
            st_info[top] = s.rot_add[info * COCTREE_ROT_COUNT + edge.rotation_id];
            st_pxy[top] = pxy * uint32_t(v_sz) + ((uint32_t(voxelPos.x) << 16) | uint32_t(voxelPos.y));
            st_pzs[top] = pzs * uint32_t(v_sz) + (uint32_t(voxelPos.z) << 16);
          }
        }
        nbnf.x += fmax_ref(0.0f, cnf.y - nbnf.x) + 0.01f * d;
      }
    }
  }
}

}  // namespace lrt

namespace lrt {

// ---- SCom2 tree (BVHRT::SCom2Intersect, MultiRenderer/BVH2Common.cpp:2560-2818) ----
// Front-to-back octree walk over the packed node heads with mirrored coordinates (octant mask a);
// each entry carries its node's cumulative rotation (8 bits) and, for leaves, the surface count
// and surface flag. A leaf marches its brick's voxels and intersects every surface link's brick
// (decoded through the composed rotation and the edge's shift); inner nodes coarser than the level
// of detail are taken as leaves through their own links. Stack pushes past 32 entries are dropped
// (the reference writes past its array); a brick march stops after 4 bs + 8 voxels, which only
// cuts marches whose advance is lost to rounding.
__device__ __forceinline__ void scom2_intersect(const SceneView& s, const GeomRef& g, float3 ray_pos, float3 ray_dir,
                                                float tNear, uint32_t instId, uint32_t geomId, CRT_Hit& hit) {
  const SCom2Header* H = s.scom2_headers + g.offset.y;
  const SCom2Header& h = *H;
  const uint32_t* D = s.scom2_data;
  const int4* M = s.scom2_mods + g.offset.y * (3 * COCTREE_ROT_COUNT);
  float3 pos_ray_pos = ray_pos;
  float3 pos_ray_dir = ray_dir;
  uint32_t a = 0;
  if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
  if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
  if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
  pos_ray_pos = -1.0f * pos_ray_pos;
  const float3 inv = safe_inverse(pos_ray_dir);
  const float3 t0_ = pos_ray_pos * inv - inv;
  const float3 t1_ = pos_ray_pos * inv + inv;
  const float3 l_ = t1_ - t0_;
  const uint32_t nn[8] = {0x124u, 0x835u, 0x386u, 0x887u, 0x568u, 0x878u, 0x788u, 0x888u};
  constexpr int SC_STACK = 32;
  uint32_t st_node[SC_STACK], st_info[SC_STACK], st_pxy[SC_STACK], st_pzs[SC_STACK];
  int top = 0;
  const uint32_t type = g.type;
  const uint32_t bs = h.brick_size;
  const uint32_t has_ch = h.has_channels;
  const uint32_t upl = h.uints_per_link, coo = h.child_offset_off, lo = h.links_offset;
  st_node[0] = h.nodes_arr_offset + h.root_node_off;
  st_info[0] = 0;
  st_pxy[0] = 0;
  st_pzs[0] = 1;
  while (top >= 0) {
    uint32_t nodeId = st_node[top];
    const uint32_t info = st_info[top], pxy = st_pxy[top], pzs = st_pzs[top];
    top--;
    const uint32_t level_sz = pzs & 0xFFFFu;
    const float level_sz_f = float(level_sz);
    const uint32_t px = pxy >> 16, py = pxy & 0xFFFFu, pz = pzs >> 16;
    float d = 1.0f / float(level_sz);
    const float3 p_f = make3(float(px), float(py), float(pz));
    const float3 t0 = t0_ + d * p_f * l_;
    const float3 t1 = t0_ + d * (p_f + 1.0f) * l_;
    if (fmin_ref(t1.x, fmin_ref(t1.y, t1.z)) < tNear) continue;
    bool is_leaf = info > 0xFFu;
    bool is_surface = info > 0xFFFFFFu;
    uint32_t surface_count = (info >> 8) & 0xFFFFu;
    if (!is_leaf) {
      const float t = s.preset.fixed_lod > 0 ? 1.0f : fmax_ref(t0.x, fmax_ref(t0.y, t0.z));
      const float target_lod_size = (powf(2.0f, s.preset.level_of_detail) - 0.01f) / t;
      is_leaf = float(level_sz) > target_lod_size;
      if (is_leaf) {
        const SCom2NodeHead node = scom2_node_head(h, D[nodeId], D[nodeId + lo - 1]);
        if (node.base_type == SCOM2_CHILD_EMPTY) continue;
        is_surface = (node.base_type & SCOM2_CHILD_TYPE_MASK) == SCOM2_CHILD_LEAF_SURFACE;
        nodeId = nodeId + lo;
        surface_count = node.base_links_end - has_ch;
      }
    }
    if (is_leaf) {
      const float sz_inv = 2.0f / level_sz_f;
      const uint32_t p_mask = level_sz - 1;
      const uint32_t rx = (a & 4) ? (~px & p_mask) : px;
      const uint32_t ry = (a & 2) ? (~py & p_mask) : py;
      const uint32_t rz = (a & 1) ? (~pz & p_mask) : pz;
      const float3 brick_min_pos = make3(-1, -1, -1) + sz_inv * make3(float(rx), float(ry), float(rz));
      float tmin = fmax_ref(t0.x, fmax_ref(t0.y, t0.z));
      const float tmax = fmin_ref(t1.x, fmin_ref(t1.y, t1.z));
      const uint32_t rotId = info & 0xFFu;
      const float max_val = scom2_max_sdf_val(level_sz_f);
      d = 2.0f / (level_sz_f * float(bs));
      const float old_t = hit.t;
      const float scale = 0.5f * level_sz_f * float(bs);
      const float hi = float(bs) - 1e-6f;
      uint32_t steps = 0;
      while (tmin < tmax && hit.t == old_t && steps < 4 * bs + 8) {
        steps++;
        const float3 hit_pos = ray_pos + tmin * ray_dir;
        const float3 local_pos = (hit_pos - brick_min_pos) * scale;
        const float3 voxelPos = make3(floorf(clamp_ref(local_pos.x, 1e-6f, hi)), floorf(clamp_ref(local_pos.y, 1e-6f, hi)),
                                      floorf(clamp_ref(local_pos.z, 1e-6f, hi)));
        const float3 min_pos = brick_min_pos + d * voxelPos;
        const float3 max_pos = min_pos + d * make3(1, 1, 1);
        const float2 fNearFar = ray_box2(ray_pos, safe_inverse(ray_dir), min_pos, max_pos);
        const int vx = int(voxelPos.x), vy = int(voxelPos.y), vz = int(voxelPos.z);
        for (uint32_t s_id = 0; fNearFar.x < fNearFar.y && s_id < surface_count && hit.t == old_t; ++s_id) {
          const uint32_t link_0 = nodeId + s_id * upl;
          const SCom2DataEdge de = scom2_data_edge(h, max_val, D[link_0], D[link_0 + coo]);
          const uint32_t offset = h.bricks_arr_offset + h.bricks_step * de.data_offset;
          const int4 m = M[s.rot_add[rotId * COCTREE_ROT_COUNT + de.rotation_id]];
          float values[8];
          float vmin = 1.0f;
          float vmax = is_surface ? -1.0f : 1.0f;
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const int cx = vx + ((i & 4) >> 2), cy = vy + ((i & 2) >> 1), cz = vz + (i & 1);
            const uint32_t vId0 = uint32_t(m.x * cx + m.y * cy + m.z * cz + m.w);
            const uint32_t p_val = D[offset + vId0 / h.values_per_uint];
            const uint32_t p_off = (vId0 % h.values_per_uint) * h.bits_per_value;
            const float val =
                max_val * (2.0f * float((p_val >> p_off) & h.value_mask) / float(h.value_mask) - 1.0f) + de.add;
            values[i] = val - s.preset.compact_sdf_eps * d;
            vmin = fmin_ref(vmin, val);
            vmax = fmax_ref(vmax, val);
          }
          float3 start_q = make3(0.0f, 0.0f, 0.0f);
          if (vmin <= 0.0f && vmax >= 0.0f) {
            const float3 start_pos = ray_pos + fNearFar.x * ray_dir;
            const float vox_sz = 0.5f * level_sz_f * float(bs);
            start_q = vox_sz * (start_pos - min_pos);
            const float qNear = fmax_ref(0.0f, vox_sz * (tNear - fNearFar.x));
            const float qFar = vox_sz * (fNearFar.y - fNearFar.x);
            local_surface_intersection<true>(s, type, ray_dir, instId, geomId, values, nodeId, d, qNear, qFar, fNearFar,
                                       start_q, hit, nullptr, is_surface);
          }
          // with channels the hit names the leaf's channel link (MULTI_RENDER_MODE_PRIMITIVE excepted);
          // the channel modes turn the quotients by the channel's rotation (BVH2Common.cpp:2731-2747)
          if (hit.t != old_t && has_ch > 0 && s.preset.render_mode != MULTI_RENDER_MODE_PRIMITIVE) {
            const uint32_t ch_off = nodeId + surface_count * upl;
            const SCom2ChildEdge ce = scom2_child_edge(h, D[ch_off], D[ch_off + coo]);
            hit.primId = ce.child_offset;
            if (s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL || s.preset.render_mode == MULTI_RENDER_MODE_CHANNEL_LAMBERT) {
              const uint32_t crot = s.rot_add[rotId * COCTREE_ROT_COUNT + ce.rotation_id];
              const float tl = (hit.t - fNearFar.x) / d;
              channel_quotients_rotated(hit, M[crot], int(bs + 1), float(bs), start_q + tl * ray_dir);
            }
          }
        }
        tmin += fmax_ref(0.0f, fNearFar.y - tmin) + 1e-5f;
        if (hit.t < old_t) {
          top = -1;
          if (s.preset.render_mode == MULTI_RENDER_MODE_LOD) hit.primId = uint32_t(log2f(float(level_sz * bs)));
        }
      }
    } else {
      const SCom2NodeHead node = scom2_node_head(h, D[nodeId], D[nodeId + lo - 1]);
      const uint32_t node1 = D[nodeId + h.references_offset];
      const uint32_t node2 = D[nodeId + lo - 1];
      uint32_t b_node[4], b_info[4], b_pxy[4], b_pzs[4];
      int buf_top = 0;
      const float3 tm = 0.5f * (t0 + t1);
      int curr = ot_first_node(t0, tm);
      const int4 mc = M[2 * COCTREE_ROT_COUNT + info];
      do {
        const uint32_t raw = uint32_t(curr) ^ a;
        const uint32_t voxId = uint32_t(mc.x * int((raw & 4) >> 2) + mc.y * int((raw & 2) >> 1) + mc.z * int(raw & 1) + mc.w);
        const uint32_t ch_type = (node.children_types >> (SCOM2_CHILD_TYPE_BITS * voxId)) & SCOM2_CHILD_TYPE_MASK;
        const uint32_t ch_start = voxId == 0 ? node.base_links_end : scom2_links_end(h, voxId - 1, node1, node2);
        const uint32_t ch_end = scom2_links_end(h, voxId, node1, node2);
        if (ch_type != SCOM2_CHILD_EMPTY && buf_top < 4) {
          b_pxy[buf_top] = (pxy << 1) | (((uint32_t(curr) & 4u) << 14) | ((uint32_t(curr) & 2u) >> 1));
          b_pzs[buf_top] = (pzs << 1) | ((uint32_t(curr) & 1u) << 16);
          const uint32_t link_off = nodeId + lo + ch_start * upl;
          if (ch_type == SCOM2_CHILD_NODE) {
            const SCom2ChildEdge ce = scom2_child_edge(h, D[link_off], D[link_off + coo]);
            b_node[buf_top] = h.nodes_arr_offset + ce.child_offset;
            b_info[buf_top] = s.rot_add[info * COCTREE_ROT_COUNT + ce.rotation_id];
          } else {
            const uint32_t surf_count = ch_end - ch_start - has_ch;
            b_node[buf_top] = link_off;
            b_info[buf_top] = (info & 0xFFu) | (surf_count << 8) | (ch_type == SCOM2_CHILD_LEAF_SURFACE ? (1u << 24) : 0u);
          }
          buf_top++;
        }
        const uint32_t ni = nn[curr];
        curr = ot_new_node((curr & 4) ? t1.x : tm.x, int(ni & 0xF), (curr & 2) ? t1.y : tm.y, int((ni >> 4) & 0xF),
                           (curr & 1) ? t1.z : tm.z, int(ni >> 8));
      } while (curr < 8);
      for (int i = 1; i <= buf_top; i++) {
        const int k = top + i;
        if (k < SC_STACK) {
          st_node[k] = b_node[buf_top - i];
          st_info[k] = b_info[buf_top - i];
          st_pxy[k] = b_pxy[buf_top - i];
          st_pzs[k] = b_pzs[buf_top - i];
        }
      }
      top += buf_top;
      if (top >= SC_STACK) top = SC_STACK - 1;
    }
  }
}

}  // namespace lrt

namespace lrt {

// LocalSurfaceIntersectionFast (BVH2Common.cpp:2820-2986): the analytic root of the trilinear
// cubic along the ray (or the box in BBOX mode); the normal from central differences, normalised
// without the instance transform; no iteration count or channel quotients.
__device__ __forceinline__ void local_surface_fast(const SceneView& s, uint32_t type, float3 ray_dir, uint32_t instId,
                                                   uint32_t geomId, const float values[8], uint32_t primId, float d,
                                                   float qNear, float qFar, float2 fNearFar, float3 start_q,
                                                   CRT_Hit& hit) {
  const float EPS = 1e-6f;
  const float d_inv = 1.0f / d;
  float t = qNear;
  bool hitp = false;
  const float start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
  if (start_dist <= EPS || s.preset.sdf_node_intersect == SDF_OCTREE_NODE_INTERSECT_BBOX) {
    hitp = true;
  } else {
    const float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv,
                s011 = values[3] * d_inv, s100 = values[4] * d_inv, s101 = values[5] * d_inv,
                s110 = values[6] * d_inv, s111 = values[7] * d_inv;
    const float tmp = s101 - s001;
    const float k0 = s000, k1 = s100 - s000, k2 = s010 - s000, k3 = s110 - s010 - k1, k4 = k0 - s001;
    const float k5 = k1 - tmp, k6 = k2 - (s011 - s001), k7 = k3 - (s111 - s011 - tmp);
    const float3 o = start_q, d3 = ray_dir;
    const float m0 = o.x * o.y, m1 = d3.x * d3.y, m2 = o.x * d3.y + o.y * d3.x;
    const float m3 = k5 * o.z - k1, m4 = k6 * o.z - k2, m5 = k7 * o.z - k3;
    const float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
    const float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
    const float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
    const float c3 = k7 * m1 * d3.z;
    float x1 = 1000, x2 = 1000, x3 = 1000;
    if (fabs_ref(c3) > 1e-5f) {
      const float a = c2 / c3, b = c1 / c3, c = c0 / c3;
      const float Q = (a * a - 3 * b) / 9;
      const float R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
      const float Q3 = Q * Q * Q;
      if (R * R < Q3) {
        const float PI = 3.14159265358979323846f;
        const float theta = acosf(R / sqrtf(Q3));
        x1 = -2 * sqrtf(Q) * cosf(theta / 3) - a / 3;
        x2 = -2 * sqrtf(Q) * cosf((theta + 2 * PI) / 3) - a / 3;
        x3 = -2 * sqrtf(Q) * cosf((theta - 2 * PI) / 3) - a / 3;
      } else {
        const float A = -sign_ref(R) * powf(fabs_ref(R) + sqrtf(R * R - Q3), 1.0f / 3.0f);
        const float B = fabs_ref(A) > EPS ? Q / A : 0;
        x1 = A + B - a / 3;
      }
    } else if (fabs_ref(c2) > 1e-7f) {
      const float a = c2, b = c1, c = c0;
      const float D = b * b - 4 * a * c;
      if (D > 0) {
        const float q = -0.5f * (b + sign_ref(b) * sqrtf(D));
        x1 = q / a;
        if (fabs_ref(q) > EPS) x2 = c / q;
      }
    } else if (fabs_ref(c1) > EPS) {
      x1 = -c0 / c1;
    }
    x1 = x1 < 0 ? 1000 : x1;
    x2 = x2 < 0 ? 1000 : x2;
    x3 = x3 < 0 ? 1000 : x3;
    t = fmin_ref(x1, fmin_ref(x2, x3));
    hitp = (t >= 0 && t <= qFar);
  }
  const float tReal = fNearFar.x + d * t;
  if (t <= qFar && hitp && tReal < hit.t) {
    const float3 p0 = start_q + t * ray_dir;
    const float h = 0.001f;
    const float ddx = (eval_dist_trilinear(values, p0 + make3(h, 0, 0)) - eval_dist_trilinear(values, p0 + make3(-h, 0, 0))) / (2 * h);
    const float ddy = (eval_dist_trilinear(values, p0 + make3(0, h, 0)) - eval_dist_trilinear(values, p0 + make3(0, -h, 0))) / (2 * h);
    const float ddz = (eval_dist_trilinear(values, p0 + make3(0, 0, h)) - eval_dist_trilinear(values, p0 + make3(0, 0, -h))) / (2 * h);
    const float2 en = encode_normal(normalize3(make3(ddx, ddy, ddz)));
    hit.t = tReal;
    hit.primId = primId;
    hit.instId = instId;
    hit.geomId = geomId | (type << SH_TYPE);
    hit.coords[0] = 0;
    hit.coords[1] = 0;
    hit.coords[2] = en.x;
    hit.coords[3] = en.y;
  }
}

// new_node(i, txm, tym, tzm) (utils/raytracing/intersection_common.h:87-93): the next child of a
// node after child i, 8 or more when the ray leaves the node
__device__ __forceinline__ uint32_t ot_new_node_i(uint32_t i, float txm, float tym, float tzm) {
  const uint32_t x = 4 + i, y = 2 + (i & 0x5) + ((i & 0x2) << 3), z = 1 + (i & 0x6) + ((i & 0x1) << 3);
  return (txm < tym) ? (txm < tzm ? x : z) : (tym < tzm ? y : z);
}

// SCom2IntersectFast (BVH2Common.cpp:2988-3231), taken instead of scom2_intersect for a scene whose
// only SCom2 tree is plain when it renders LAMBERT_NO_TEX: the walk keeps the current node's next
// child in its entry and steps one child at a time, pushing a node only while it has children
// left (16 entries; a push past them is dropped), and walks a leaf's 2^3 voxels front to back.
// Child offsets are taken as absolute word indices, as the reference does (the tree is the
// scene's only one, so its words start at 0).
__device__ __forceinline__ void scom2_intersect_fast(const SceneView& s, const GeomRef& g, float3 ray_pos,
                                                     float3 ray_dir, float tNear, uint32_t instId, uint32_t geomId,
                                                     CRT_Hit& hit) {
  const SCom2Header& h = s.scom2_headers[g.offset.y];
  const uint32_t* D = s.scom2_data;
  const int4* M = s.scom2_mods + g.offset.y * (3 * COCTREE_ROT_COUNT);
  const uint32_t type = g.type;
  float3 pos_ray_pos = ray_pos;
  float3 pos_ray_dir = ray_dir;
  uint32_t a = 0;
  if (ray_dir.x < 0) { pos_ray_pos.x *= -1; pos_ray_dir.x *= -1; a |= 4; }
  if (ray_dir.y < 0) { pos_ray_pos.y *= -1; pos_ray_dir.y *= -1; a |= 2; }
  if (ray_dir.z < 0) { pos_ray_pos.z *= -1; pos_ray_dir.z *= -1; a |= 1; }
  pos_ray_pos = -1.0f * pos_ray_pos;
  const float3 inv = safe_inverse(pos_ray_dir);
  const float3 t0_ = pos_ray_pos * inv - inv;
  const float3 t1_ = pos_ray_pos * inv + inv;
  const float3 l_ = t1_ - t0_;
  constexpr int FS = 16;
  uint32_t st_links[FS], st_info[FS], st_tr[FS], st_pxy[FS], st_pzs[FS];
  int top = 0;
  float d = 1.0f;
  float3 t0 = t0_;
  float3 tm = t0 + 0.5f * l_;
  uint32_t links, info, tr, pxy, pzs;
  {
    const uint32_t n0 = D[h.root_node_off];
    const SCom2NodeHead node = scom2_node_head(h, n0, n0);
    const uint32_t curr = uint32_t(ot_first_node(t0, tm));
    links = h.root_node_off + h.links_offset + (node.base_type == SCOM2_CHILD_EMPTY ? 0u : 1u);
    info = (node.children_types << 8) | curr;
    tr = 0;
    pxy = 0;
    pzs = 1;
  }
  st_links[0] = 0xFFFFFFFFu;
  st_info[0] = 0;
  st_tr[0] = 0;
  st_pxy[0] = 0;
  st_pzs[0] = 0;
  while (top >= 0) {
    const uint32_t currNode = info & 0x7u;
    const uint32_t raw = currNode ^ a;
    const int4 mc = M[2 * COCTREE_ROT_COUNT + tr];
    const uint32_t childN = uint32_t(mc.x * int((raw & 4) >> 2) + mc.y * int((raw & 2) >> 1) + mc.z * int(raw & 1) + mc.w);
    const uint32_t active = (((info & 0x00AAAA00u) >> 1) | (info & 0x00555500u)) >> 8;
    const uint32_t childLink = links + uint32_t(__popc(active & ((1u << (2 * childN)) - 1u)));
    const uint32_t childHasData = active & (1u << (2 * childN));
    const uint32_t childIsLeaf = info & (1u << (8 + 2 * childN));
    const uint32_t rotId = info >> 24;
    d = 1.0f / float(pzs & 0xFFFFu);
    const float3 pf2 = make3(float(pxy >> 16) + ((currNode & 4) ? 1.0f : 0.5f),
                             float(pxy & 0xFFFFu) + ((currNode & 2) ? 1.0f : 0.5f),
                             float(pzs >> 16) + ((currNode & 1) ? 1.0f : 0.5f));
    const float3 tl = t0_ + d * pf2 * l_;
    bool advance = false;
    if (fmin_ref(fmin_ref(tl.x, tl.y), tl.z) < tNear || childHasData == 0) {
      advance = true;
    } else if (childIsLeaf > 0) {
      const uint32_t level_sz = 2 * (pzs & 0xFFFFu);
      const uint32_t p0 = 2 * (pxy >> 16) + ((currNode & 4) >> 2);
      const uint32_t p1 = 2 * (pxy & 0xFFFFu) + ((currNode & 2) >> 1);
      const uint32_t p2 = 2 * (pzs >> 16) + (currNode & 1);
      d = 2.0f / float(level_sz);
      const float3 p_f = make3(float(p0), float(p1), float(p2));
      t0 = t0_ + 0.5f * d * p_f * l_;
      const float3 t1 = t0_ + 0.5f * d * (p_f + 1.0f) * l_;
      const float max_val = scom2_max_sdf_val(float(level_sz));
      const bool childIsSurface = (info & (1u << (8 + 2 * childN + 1))) != 0;
      const uint32_t link0 = D[childLink];
      const SCom2DataEdge de = scom2_data_edge(h, max_val, link0, link0);
      const uint32_t offset = h.bricks_arr_offset + h.bricks_step * de.data_offset;
      const int4 m = M[s.rot_add[rotId * COCTREE_ROT_COUNT + de.rotation_id]];
      const float3 vox_tm = 0.5f * (t0 + t1);
      uint32_t vox = uint32_t(ot_first_node(t0, vox_tm));
      while (vox < 8) {
        const float3 vox_tl = make3((vox & 4) ? t1.x : vox_tm.x, (vox & 2) ? t1.y : vox_tm.y, (vox & 1) ? t1.z : vox_tm.z);
        const uint32_t voxId = vox ^ a;
        float vmin = 1.0f;
        float vmax = childIsSurface ? -1.0f : 1.0f;
        float values[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int cx = int((voxId & 4) >> 2) + ((i & 4) >> 2), cy = int((voxId & 2) >> 1) + ((i & 2) >> 1),
                    cz = int(voxId & 1) + (i & 1);
          const uint32_t vId0 = uint32_t(m.x * cx + m.y * cy + m.z * cz + m.w);
          const uint32_t p_val = D[offset + vId0 / h.values_per_uint];
          const uint32_t p_off = (vId0 % h.values_per_uint) * h.bits_per_value;
          const float val = max_val * (2.0f * float((p_val >> p_off) & h.value_mask) / float(h.value_mask) - 1.0f) + de.add;
          values[i] = val - 0.5f * s.preset.compact_sdf_eps * d;
          vmin = fmin_ref(vmin, val);
          vmax = fmax_ref(vmax, val);
        }
        if (vmin <= 0.0f && vmax >= 0.0f) {
          const uint32_t p_mask = level_sz - 1;
          const uint32_t r0 = (a & 4) ? (~p0 & p_mask) : p0, r1 = (a & 2) ? (~p1 & p_mask) : p1,
                         r2 = (a & 1) ? (~p2 & p_mask) : p2;
          const float3 vox_t0 = make3((vox & 4) ? vox_tm.x : t0.x, (vox & 2) ? vox_tm.y : t0.y, (vox & 1) ? vox_tm.z : t0.z);
          const float3 vox_t1 = make3((vox & 4) ? t1.x : vox_tm.x, (vox & 2) ? t1.y : vox_tm.y, (vox & 1) ? t1.z : vox_tm.z);
          const float tmin = fmax_ref(vox_t0.x, fmax_ref(vox_t0.y, vox_t0.z));
          const float tmax = fmin_ref(vox_t1.x, fmin_ref(vox_t1.y, vox_t1.z));
          const float3 start_pos = ray_pos + tmin * ray_dir;
          const float3 min_pos = make3(-1, -1, -1) + d * make3(float(r0), float(r1), float(r2)) +
                                 0.5f * d * make3(float((voxId & 4) >> 2), float((voxId & 2) >> 1), float(voxId & 1));
          const float vox_sz = float(level_sz);
          const float3 start_q = vox_sz * (start_pos - min_pos);
          const float qNear = fmax_ref(0.0f, vox_sz * (tNear - tmin));
          const float qFar = vox_sz * (tmax - tmin);
          const float last_t = hit.t;
          local_surface_fast(s, type, ray_dir, instId, geomId, values, childLink, 0.5f * d, qNear, qFar,
                             float2{tmin, tmax}, start_q, hit);
          if (hit.t < last_t) return;
        }
        vox = ot_new_node_i(vox, vox_tl.x, vox_tl.y, vox_tl.z);
      }
      advance = true;
    } else {
      const uint32_t next = ot_new_node_i(currNode, tl.x, tl.y, tl.z);
      if (next < 8) {
        info = next | (info & 0xFFFFFF00u);
        if (top + 1 < FS) {
          ++top;
          st_links[top] = links;
          st_info[top] = info;
          st_tr[top] = tr;
          st_pxy[top] = pxy;
          st_pzs[top] = pzs;
        }
      }
      d = 0.5f / float(pzs & 0xFFFFu);
      const float3 p_f = make3(float(2 * (pxy >> 16) + ((currNode & 4) >> 2)), float(2 * (pxy & 0xFFFFu) + ((currNode & 2) >> 1)),
                               float(2 * (pzs >> 16) + (currNode & 1)));
      t0 = t0_ + d * p_f * l_;
      tm = t0 + 0.5f * d * l_;
      const uint32_t link0 = D[childLink];
      const SCom2ChildEdge ce = scom2_child_edge(h, link0, link0);
      const uint32_t n0 = D[ce.child_offset];
      const SCom2NodeHead node = scom2_node_head(h, n0, n0);
      const uint32_t nextNode = uint32_t(ot_first_node(t0, tm));
      const uint32_t rotIdx = s.rot_add[rotId * COCTREE_ROT_COUNT + ce.rotation_id];
      links = ce.child_offset + h.links_offset + (node.base_type == SCOM2_CHILD_EMPTY ? 0u : 1u);
      tr = s.rot_add[tr * COCTREE_ROT_COUNT + ce.rotation_id];
      info = (rotIdx << 24) | (node.children_types << 8) | nextNode;
      pxy = (pxy << 1) | (((currNode & 4) << 14) | ((currNode & 2) >> 1));
      pzs = (pzs << 1) | ((currNode & 1) << 16);
    }
    if (advance) {
      const uint32_t next = ot_new_node_i(currNode, tl.x, tl.y, tl.z);
      if (next >= 8) {
        links = st_links[top];
        info = st_info[top];
        tr = st_tr[top];
        pxy = st_pxy[top];
        pzs = st_pzs[top];
        top--;
      } else {
        info = next | (info & 0xFFFFFF00u);
      }
    }
  }
}

}  // namespace lrt
