// Host-side matrix helpers with LiteMath's exact operation order (compiled with
// -ffp-contract=off): inverse4x4 (LiteMath.h:2070-2154), transpose (LiteMath.h:1986-1991),
// lookAt (LiteMath.h:3315-3340), perspectiveMatrix (LiteMath.h:3342-3360).
#pragma once
#include "lrt_math.h"

namespace lrt {

// element (row r, column c) of a column-major matrix
inline float& M(float4x4& m, int r, int c) { return m.m[c * 4 + r]; }
inline float M(const float4x4& m, int r, int c) { return m.m[c * 4 + r]; }

inline float4x4 identity4x4() {
  float4x4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
  return r;
}

inline float4x4 transpose4x4(const float4x4& a) {
  float4x4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) M(r, i, j) = M(a, j, i);
  return r;
}

inline float4x4 inverse4x4(const float4x4& m1) {
  float tmp[12];
  float4x4 m{};
  tmp[0] = M(m1, 2, 2) * M(m1, 3, 3);
  tmp[1] = M(m1, 2, 3) * M(m1, 3, 2);
  tmp[2] = M(m1, 2, 1) * M(m1, 3, 3);
  tmp[3] = M(m1, 2, 3) * M(m1, 3, 1);
  tmp[4] = M(m1, 2, 1) * M(m1, 3, 2);
  tmp[5] = M(m1, 2, 2) * M(m1, 3, 1);
  tmp[6] = M(m1, 2, 0) * M(m1, 3, 3);
  tmp[7] = M(m1, 2, 3) * M(m1, 3, 0);
  tmp[8] = M(m1, 2, 0) * M(m1, 3, 2);
  tmp[9] = M(m1, 2, 2) * M(m1, 3, 0);
  tmp[10] = M(m1, 2, 0) * M(m1, 3, 1);
  tmp[11] = M(m1, 2, 1) * M(m1, 3, 0);

  M(m, 0, 0) = tmp[0] * M(m1, 1, 1) + tmp[3] * M(m1, 1, 2) + tmp[4] * M(m1, 1, 3);
  M(m, 0, 0) -= tmp[1] * M(m1, 1, 1) + tmp[2] * M(m1, 1, 2) + tmp[5] * M(m1, 1, 3);
  M(m, 1, 0) = tmp[1] * M(m1, 1, 0) + tmp[6] * M(m1, 1, 2) + tmp[9] * M(m1, 1, 3);
  M(m, 1, 0) -= tmp[0] * M(m1, 1, 0) + tmp[7] * M(m1, 1, 2) + tmp[8] * M(m1, 1, 3);
  M(m, 2, 0) = tmp[2] * M(m1, 1, 0) + tmp[7] * M(m1, 1, 1) + tmp[10] * M(m1, 1, 3);
  M(m, 2, 0) -= tmp[3] * M(m1, 1, 0) + tmp[6] * M(m1, 1, 1) + tmp[11] * M(m1, 1, 3);
  M(m, 3, 0) = tmp[5] * M(m1, 1, 0) + tmp[8] * M(m1, 1, 1) + tmp[11] * M(m1, 1, 2);
  M(m, 3, 0) -= tmp[4] * M(m1, 1, 0) + tmp[9] * M(m1, 1, 1) + tmp[10] * M(m1, 1, 2);
  M(m, 0, 1) = tmp[1] * M(m1, 0, 1) + tmp[2] * M(m1, 0, 2) + tmp[5] * M(m1, 0, 3);
  M(m, 0, 1) -= tmp[0] * M(m1, 0, 1) + tmp[3] * M(m1, 0, 2) + tmp[4] * M(m1, 0, 3);
  M(m, 1, 1) = tmp[0] * M(m1, 0, 0) + tmp[7] * M(m1, 0, 2) + tmp[8] * M(m1, 0, 3);
  M(m, 1, 1) -= tmp[1] * M(m1, 0, 0) + tmp[6] * M(m1, 0, 2) + tmp[9] * M(m1, 0, 3);
  M(m, 2, 1) = tmp[3] * M(m1, 0, 0) + tmp[6] * M(m1, 0, 1) + tmp[11] * M(m1, 0, 3);
  M(m, 2, 1) -= tmp[2] * M(m1, 0, 0) + tmp[7] * M(m1, 0, 1) + tmp[10] * M(m1, 0, 3);
  M(m, 3, 1) = tmp[4] * M(m1, 0, 0) + tmp[9] * M(m1, 0, 1) + tmp[10] * M(m1, 0, 2);
  M(m, 3, 1) -= tmp[5] * M(m1, 0, 0) + tmp[8] * M(m1, 0, 1) + tmp[11] * M(m1, 0, 2);

  tmp[0] = M(m1, 0, 2) * M(m1, 1, 3);
  tmp[1] = M(m1, 0, 3) * M(m1, 1, 2);
  tmp[2] = M(m1, 0, 1) * M(m1, 1, 3);
  tmp[3] = M(m1, 0, 3) * M(m1, 1, 1);
  tmp[4] = M(m1, 0, 1) * M(m1, 1, 2);
  tmp[5] = M(m1, 0, 2) * M(m1, 1, 1);
  tmp[6] = M(m1, 0, 0) * M(m1, 1, 3);
  tmp[7] = M(m1, 0, 3) * M(m1, 1, 0);
  tmp[8] = M(m1, 0, 0) * M(m1, 1, 2);
  tmp[9] = M(m1, 0, 2) * M(m1, 1, 0);
  tmp[10] = M(m1, 0, 0) * M(m1, 1, 1);
  tmp[11] = M(m1, 0, 1) * M(m1, 1, 0);

  M(m, 0, 2) = tmp[0] * M(m1, 3, 1) + tmp[3] * M(m1, 3, 2) + tmp[4] * M(m1, 3, 3);
  M(m, 0, 2) -= tmp[1] * M(m1, 3, 1) + tmp[2] * M(m1, 3, 2) + tmp[5] * M(m1, 3, 3);
  M(m, 1, 2) = tmp[1] * M(m1, 3, 0) + tmp[6] * M(m1, 3, 2) + tmp[9] * M(m1, 3, 3);
  M(m, 1, 2) -= tmp[0] * M(m1, 3, 0) + tmp[7] * M(m1, 3, 2) + tmp[8] * M(m1, 3, 3);
  M(m, 2, 2) = tmp[2] * M(m1, 3, 0) + tmp[7] * M(m1, 3, 1) + tmp[10] * M(m1, 3, 3);
  M(m, 2, 2) -= tmp[3] * M(m1, 3, 0) + tmp[6] * M(m1, 3, 1) + tmp[11] * M(m1, 3, 3);
  M(m, 3, 2) = tmp[5] * M(m1, 3, 0) + tmp[8] * M(m1, 3, 1) + tmp[11] * M(m1, 3, 2);
  M(m, 3, 2) -= tmp[4] * M(m1, 3, 0) + tmp[9] * M(m1, 3, 1) + tmp[10] * M(m1, 3, 2);
  M(m, 0, 3) = tmp[2] * M(m1, 2, 2) + tmp[5] * M(m1, 2, 3) + tmp[1] * M(m1, 2, 1);
  M(m, 0, 3) -= tmp[4] * M(m1, 2, 3) + tmp[0] * M(m1, 2, 1) + tmp[3] * M(m1, 2, 2);
  M(m, 1, 3) = tmp[8] * M(m1, 2, 3) + tmp[0] * M(m1, 2, 0) + tmp[7] * M(m1, 2, 2);
  M(m, 1, 3) -= tmp[6] * M(m1, 2, 2) + tmp[9] * M(m1, 2, 3) + tmp[1] * M(m1, 2, 0);
  M(m, 2, 3) = tmp[6] * M(m1, 2, 1) + tmp[11] * M(m1, 2, 3) + tmp[3] * M(m1, 2, 0);
  M(m, 2, 3) -= tmp[10] * M(m1, 2, 3) + tmp[2] * M(m1, 2, 0) + tmp[7] * M(m1, 2, 1);
  M(m, 3, 3) = tmp[10] * M(m1, 2, 2) + tmp[4] * M(m1, 2, 0) + tmp[9] * M(m1, 2, 1);
  M(m, 3, 3) -= tmp[8] * M(m1, 2, 1) + tmp[11] * M(m1, 2, 2) + tmp[5] * M(m1, 2, 0);

  const float k = 1.0f / (M(m1, 0, 0) * M(m, 0, 0) + M(m1, 0, 1) * M(m, 1, 0) + M(m1, 0, 2) * M(m, 2, 0) +
                          M(m1, 0, 3) * M(m, 3, 0));
  for (int i = 0; i < 16; ++i) m.m[i] = m.m[i] * k;
  return m;
}

inline float4x4 look_at(float3 eye, float3 center, float3 up) {
  float3 z = normalize3(make3(eye.x - center.x, eye.y - center.y, eye.z - center.z));
  float3 y = up;
  float3 x = cross3(y, z);
  y = cross3(z, x);
  x = normalize3(x);
  y = normalize3(y);
  float4x4 r{};
  const float cols[16] = {x.x, y.x, z.x, 0.0f,
                          x.y, y.y, z.y, 0.0f,
                          x.z, y.z, z.z, 0.0f,
                          -x.x * eye.x - x.y * eye.y - x.z * eye.z,
                          -y.x * eye.x - y.y * eye.y - y.z * eye.z,
                          -z.x * eye.x - z.y * eye.y - z.z * eye.z, 1.0f};
  for (int i = 0; i < 16; ++i) r.m[i] = cols[i];
  return r;
}

inline float4x4 perspective_matrix(float fovy, float aspect, float zNear, float zFar) {
  const float ymax = zNear * tanf(fovy * 3.14159265358979323846f / 360.0f);
  const float xmax = ymax * aspect;
  const float left = -xmax, right = +xmax, bottom = -ymax, top = +ymax;
  const float temp = 2.0f * zNear;
  const float temp2 = right - left;
  const float temp3 = top - bottom;
  const float temp4 = zFar - zNear;
  float4x4 r{};
  const float cols[16] = {temp / temp2, 0.0f, 0.0f, 0.0f,
                          0.0f, temp / temp3, 0.0f, 0.0f,
                          (right + left) / temp2, (top + bottom) / temp3, (-zFar - zNear) / temp4, -1.0f,
                          0.0f, 0.0f, (-temp * zFar) / temp4, 0.0f};
  for (int i = 0; i < 16; ++i) r.m[i] = cols[i];
  return r;
}

}  // namespace lrt
