/* A plain C11 caller of include/lrt.h: the path a reference-side adapter takes (create, add
 * triangles, add an instance, commit, query one ray, trace a batch, update the instance,
 * re-commit, trace again, clear, destroy). Compiled with gcc against the header and linked to
 * liblrt_hip.so, so the header must be valid C and the library must export what it declares.
 * The -m gpu test calls lrt_c_caller_run through ctypes and checks the records. */
#include <stdint.h>
#include <string.h>

#include "lrt.h"

/* A quad [-1,1]^2 at z = 0 (two triangles), instanced at z = -2, then moved to z = -3.
 * Rays: an n_side x n_side grid of origins in [-0.9,0.9]^2 at z = 0, direction -z.
 * Writes (each lrt_hit[n_side^2]) `before` (instance at z = -2, batch over host buffers),
 * `after` (after lrt_update_instance and a re-commit), and `one` (lrt_query_one of ray 0 before
 * the update). Returns LRT_OK or the first failing status (message in lrt_last_error). */
int lrt_c_caller_run(uint32_t n_side, lrt_hit* before, lrt_hit* after, lrt_hit* one) {
  static const float verts[12] = {-1.f, -1.f, 0.f, 1.f, -1.f, 0.f, 1.f, 1.f, 0.f, -1.f, 1.f, 0.f};
  static const uint32_t tris[6] = {0, 1, 2, 0, 2, 3};
  float m[16];
  lrt_scene* s = 0;
  uint32_t gid = 0, iid = 0;
  lrt_status st;
  const uint32_t n = n_side * n_side;
  float pos[4 * 64 * 64], dir[4 * 64 * 64];
  uint32_t i;
  if (n_side == 0 || n_side > 64) return LRT_ERR_INVALID_ARG;
  for (i = 0; i < n; ++i) {
    const float fx = -0.9f + 1.8f * (float)(i % n_side) / (float)(n_side > 1 ? n_side - 1 : 1);
    const float fy = -0.9f + 1.8f * (float)(i / n_side) / (float)(n_side > 1 ? n_side - 1 : 1);
    pos[4 * i + 0] = fx; pos[4 * i + 1] = fy; pos[4 * i + 2] = 0.f; pos[4 * i + 3] = 1e-9f;
    dir[4 * i + 0] = 0.f; dir[4 * i + 1] = 0.f; dir[4 * i + 2] = -1.f; dir[4 * i + 3] = 1e9f;
  }
  memset(m, 0, sizeof(m));
  m[0] = m[5] = m[10] = m[15] = 1.f;
  m[14] = -2.f;   /* column-major: translation in elements 12..14 */
  if ((st = lrt_create("BVH2Common", "cbvh_embree2", "SuperTreeletAlignedMerged4", &s)) != LRT_OK) return st;
  if ((st = lrt_add_triangles(s, verts, 4, 12, 0, tris, 6, 0, &gid)) != LRT_OK) goto done;
  if ((st = lrt_add_instance(s, gid, m, &iid)) != LRT_OK) goto done;
  if ((st = lrt_commit(s, 0)) != LRT_OK) goto done;
  if ((st = lrt_query_one(s, pos, dir, one)) != LRT_OK) goto done;
  if ((st = lrt_trace_closest_host(s, pos, dir, before, n)) != LRT_OK) goto done;
  m[14] = -3.f;
  if ((st = lrt_update_instance(s, iid, m)) != LRT_OK) goto done;
  if ((st = lrt_commit(s, 0)) != LRT_OK) goto done;
  if ((st = lrt_trace_closest_host(s, pos, dir, after, n)) != LRT_OK) goto done;
  if ((st = lrt_clear_scene(s)) != LRT_OK) goto done;
  st = lrt_clear_geom(s);
done:
  lrt_destroy(s);
  return st;
}
