// Scalar/vector helpers with the reference's exact operation order
// (LiteMath.h:133-138, 945-992, 1813-1819; raytrace_common.h:59-92;
// intersection_common.h:96-101). Every file that includes this header is
// compiled with -ffp-contract=off so that host and gfx950 code round
// identically: only IEEE +,-,*,/ and sqrt appear here.
#pragma once
#include "lrt_types.h"
#include <cmath>

namespace lrt {

// std::min / std::max semantics, including which operand a NaN comparison returns.
LRT_HD float fmin_ref(float a, float b) { return (b < a) ? b : a; }
LRT_HD float fmax_ref(float a, float b) { return (a < b) ? b : a; }
LRT_HD float fabs_ref(float a) { return std::fabs(a); }
LRT_HD float sign_ref(float x) { return x == 0.0f ? 0.0f : (x < 0.0f ? -1.0f : 1.0f); }
LRT_HD float clamp_ref(float u, float a, float b) { return fmin_ref(fmax_ref(a, u), b); }

LRT_HD float3 make3(float x, float y, float z) { float3 r; r.x = x; r.y = y; r.z = z; return r; }
LRT_HD float4 make4(float x, float y, float z, float w) { float4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
LRT_HD float3 xyz(const float4& a) { return make3(a.x, a.y, a.z); }
LRT_HD float3 operator+(float3 a, float3 b) { return make3(a.x + b.x, a.y + b.y, a.z + b.z); }
LRT_HD float3 operator-(float3 a, float3 b) { return make3(a.x - b.x, a.y - b.y, a.z - b.z); }
LRT_HD float3 operator*(float3 a, float3 b) { return make3(a.x * b.x, a.y * b.y, a.z * b.z); }
LRT_HD float3 operator/(float3 a, float3 b) { return make3(a.x / b.x, a.y / b.y, a.z / b.z); }
LRT_HD float3 operator*(float3 a, float k) { return make3(a.x * k, a.y * k, a.z * k); }
LRT_HD float3 operator*(float k, float3 a) { return make3(k * a.x, k * a.y, k * a.z); }
LRT_HD float3 operator/(float3 a, float k) { return make3(a.x / k, a.y / k, a.z / k); }
LRT_HD float3 operator+(float3 a, float k) { return make3(a.x + k, a.y + k, a.z + k); }
LRT_HD float3 operator-(float3 a, float k) { return make3(a.x - k, a.y - k, a.z - k); }
LRT_HD float3 min3(float3 a, float3 b) { return make3(fmin_ref(a.x, b.x), fmin_ref(a.y, b.y), fmin_ref(a.z, b.z)); }
LRT_HD float3 max3(float3 a, float3 b) { return make3(fmax_ref(a.x, b.x), fmax_ref(a.y, b.y), fmax_ref(a.z, b.z)); }

LRT_HD float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// LiteMath.h:987-992 (shuffle form expands to the textbook order below)
LRT_HD float3 cross3(float3 a, float3 b) {
  return make3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
LRT_HD float length3(float3 a) { return std::sqrt(dot3(a, a)); }
LRT_HD float3 normalize3(float3 a) { float lenInv = 1.0f / length3(a); return a * lenInv; }

// raytrace_common.h:59-67
LRT_HD float3 safe_inverse(float3 d) {
  const float ooeps = 1.0e-36f;
  float3 r;
  r.x = 1.0f / (fabs_ref(d.x) > ooeps ? d.x : ooeps * sign_ref(d.x));
  r.y = 1.0f / (fabs_ref(d.y) > ooeps ? d.y : ooeps * sign_ref(d.y));
  r.z = 1.0f / (fabs_ref(d.z) > ooeps ? d.z : ooeps * sign_ref(d.z));
  return r;
}

#ifdef __HIPCC__
// ray_box2 with the hardware min/max (v_min3/v_max3). It differs from ray_box2 only where a
// product is NaN (inf * 0: a zero direction component, or a NaN input) and in the sign of a zero
// result, which only ever meets comparisons. The walks use it for rays whose direction
// components are all finite and non-zero and whose origin is finite (see box_test_fast).
__device__ __forceinline__ float2 ray_box2_fast(float3 o, float3 inv, float3 bmin, float3 bmax) {
  const float lo  = inv.x * (bmin.x - o.x);
  const float hi  = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  float2 r;
  r.x = fmaxf(fmaxf(fminf(lo, hi), fminf(lo1, hi1)), fminf(lo2, hi2));
  r.y = fminf(fminf(fmaxf(lo, hi), fmaxf(lo1, hi1)), fmaxf(lo2, hi2));
  return r;
}

// True when ray_box2_fast gives the decisions of ray_box2 for this ray (see above).
__device__ __forceinline__ bool box_test_fast(float3 o, float3 d) {
  const float inf = __int_as_float(0x7f800000);
  return fabsf(d.x) > 0.0f && fabsf(d.y) > 0.0f && fabsf(d.z) > 0.0f && fabsf(d.x) < inf && fabsf(d.y) < inf &&
         fabsf(d.z) < inf && fabsf(o.x) < inf && fabsf(o.y) < inf && fabsf(o.z) < inf;
}
#endif

// raytrace_common.h:69-82; returns (tmin, tmax)
LRT_HD float2 ray_box2(float3 o, float3 inv, float3 bmin, float3 bmax) {
  const float lo  = inv.x * (bmin.x - o.x);
  const float hi  = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  const float tmin = fmax_ref(fmin_ref(lo, hi), fmin_ref(lo1, hi1));
  const float tmax = fmin_ref(fmax_ref(lo, hi), fmax_ref(lo1, hi1));
  float2 r;
  r.x = fmax_ref(tmin, fmin_ref(lo2, hi2));
  r.y = fmin_ref(tmax, fmax_ref(lo2, hi2));
  return r;
}

// LiteMath.h:1813-1819 column-major mat * vec4
LRT_HD float4 mat_mul4(const float* B, float4 V) {
  float4 r;
  r.x = V.x * B[0] + V.y * B[4] + V.z * B[8]  + V.w * B[12];
  r.y = V.x * B[1] + V.y * B[5] + V.z * B[9]  + V.w * B[13];
  r.z = V.x * B[2] + V.y * B[6] + V.z * B[10] + V.w * B[14];
  r.w = V.x * B[3] + V.y * B[7] + V.z * B[11] + V.w * B[15];
  return r;
}
// raytrace_common.h:84-92
LRT_HD float3 matmul4x3(const float* m, float3 v) { return xyz(mat_mul4(m, make4(v.x, v.y, v.z, 1.0f))); }
LRT_HD float3 matmul3x3(const float* m, float3 v) { return xyz(mat_mul4(m, make4(v.x, v.y, v.z, 0.0f))); }

// intersection_common.h:96-101 (octahedral normal encoding)
LRT_HD float2 encode_normal(float3 v) {
  const float k = 1.0f / (fabs_ref(v.x) + fabs_ref(v.y) + fabs_ref(v.z));
  float2 p; p.x = v.x * k; p.y = v.y * k;
  const float sx = (p.x >= 0.0f) ? 1.0f : -1.0f;
  const float sy = (p.y >= 0.0f) ? 1.0f : -1.0f;
  float2 r;
  if (v.z <= 0.0f) { r.x = (1.0f - fabs_ref(p.y)) * sx; r.y = (1.0f - fabs_ref(p.x)) * sy; }
  else r = p;
  return r;
}

}  // namespace lrt
