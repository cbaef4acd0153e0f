"""ORACLE - test infrastructure only.

CPU restatement of LiteRT's ray-query path (oracle/oracle.cpp), used as the checker of the
gfx950 library. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
it; the product never does.

Two builds of the same source (oracle/Makefile):
  "exact" (liboracle.so)     FMA contraction off, like the gfx950 kernels: bit-exact twin
  "fma"   (liboracle_fma.so) the reference's own flags (-O3, contraction on)

Function -> reference -> how it is pinned (tests/test_oracle_cpu.py unless stated):
  oracle_trace_closest  BVH2Common.cpp:3322-3392, 3239-3320, 47-124, 1571-1635
      pinned by brute force over all triangles (topology independence), by the closed-form
      axial-ray distance on an icosphere, by instance-transform invariance, by the
      stop-on-first-hit / miss-record special cases, and by a planted error (swapped
      barycentrics) that the brute-force check must catch.
  oracle_trace_closest  SDF leaves: BVH2Common.cpp:727-749, 141-167, 345-725, 884-1112
      pinned by the sphere-union closed form (|f(hit)| within a voxel-size bound, normal
      close to grad f), by cross-format agreement of frame octree vs. SBS of one octree, and
      by a root check of the analytic cubic in float64 (tests/test_oracle_sdf.py).
  oracle_trace_any      BVH2Common.cpp:3394-3399   pinned by AnyHit <=> closest primId != ~0
  oracle_brute_force_tris (ours)                    pinned by an analytic ray/plane case
  oracle_eye_rays       MultiRenderer.cpp:142-183   pinned by the pinhole closed form
      (direction of pixel centre from fov/aspect) and the SuperBlockIndex2DOpt permutation
      property (a bijection onto [0, w*h)).
  oracle_camera_inverse LiteMath.h:3315-3360, 2070-2154  pinned by M * inverse(M) = I in
      float64 and the closed-form perspective entries.
  oracle_trace_closest  COctreeV3 leaves: BVH2Common.cpp:1637-2173 (tests/test_coctree_cpu.py)
      loaders pinned by round trips of an independent encoder of sphere-union octrees (every leaf
      type, value width, padding, pack mode, rotation and additive shift) against the source
      values, presence flags against their definition; the walk by the identical hits of the
      three leaf types of one octree, by agreement with the frame octree of the same octree and
      the brick set of the same bricks, and by rotation/shift invariance; the rotation modifiers
      by the reference's own search (restated in numpy); the BLAS boxes by a restatement of
      add_border_nodes_rec; the reference's engine-scene COctreeV3 file by its structure and by
      hit points on the zero set of its decoded distances.
  oracle_trace_closest  textured frame octrees: BVH2Common.cpp:168-188, 689-710
      (tests/test_frame_tex_cpu.py) pinned by bit-identical hits with the frame octree of the same
      nodes and by the closed form of texture coordinates that are a linear map of position.
  oracle_trace_closest  tricubic interpolation: BVH2Common.cpp:220-335, 790-844, 884-955, 1478-1569
      (tests/test_tricubic_cpu.py) pinned by the spline's exact reproduction of quadratics (hits
      on the sphere |p-c| = r from samples of |p-c|^2 - r^2, radial normals, for brick sets and
      grids), the non-marching modes' entry-face hits and the representation modes' normal signs.
  oracle_coct_voxel     BVH2Common.cpp:1637-1851   the loaders alone (see above)
  oracle_local_surface  BVH2Common.cpp:345-725 (one voxel)  pinned by float64 root finding on
      the trilinear polynomial along the ray (first sign change), for every intersection mode
      (tests/test_oracle_sdf.py).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB_DIR overrides the build folder (the host-sanitizer build: `make -C oracle asan`)
LIB_DIR = os.environ.get("ORACLE_LIB_DIR") or os.path.join(_HERE, "lib")
_FILES = {"exact": "liboracle.so", "fma": "liboracle_fma.so"}

HIT_DTYPE = np.dtype([("t", "<f4"), ("primId", "<u4"), ("instId", "<u4"), ("geomId", "<u4"),
                      ("coords", "<f4", (4,))])
COUNTS_DTYPE = np.dtype([("blas_nodes", "<u4"), ("leaves", "<u4"), ("prims", "<u4"), ("tlas_nodes", "<u4"),
                         ("instances", "<u4"), ("sdf_voxels", "<u4")])

_ARRAYS = ["nodes", "tlas", "instances", "geoms", "geom_tags", "start_end", "prim_id_count", "vert_pos",
           "vert_norm", "indices", "prim_indices", "orig_nodes", "frame_nodes", "frame_roots", "frame_tex_nodes",
           "frame_tex_roots", "sbs_nodes",
           "sbs_data", "sbs_data_f", "sbs_roots", "sbs_headers", "multi_nodes", "multi_values", "svs_nodes", "svs_roots",
           "grid_data", "grid_offsets", "grid_sizes", "gprim_headers", "gprim_points", "coct_data", "coct_geoms",
           "coct_mods", "dag_nodes", "dag_child_edges", "dag_data_edges", "dag_distances", "dag_geoms", "dag_mods",
           "rot_add", "scom2_data", "scom2_headers", "scom2_mods"]


class OracleScene(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in _ARRAYS] + [("preset", ctypes.c_uint8 * 60), ("n_scom2", ctypes.c_uint32)]


_libs = {}


def available(flavor: str = "exact") -> bool:
    return os.path.exists(os.path.join(LIB_DIR, _FILES[flavor]))


def load(flavor: str = "exact") -> ctypes.CDLL:
    if flavor not in _libs:
        path = os.path.join(LIB_DIR, _FILES[flavor])
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} is missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp, u64, u32, i32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        L.oracle_flavor.restype = i32
        L.oracle_trace_closest.argtypes = [ctypes.POINTER(OracleScene), vp, vp, vp, u64, i32, vp]
        L.oracle_trace_closest.restype = u64
        L.oracle_trace_any.argtypes = [ctypes.POINTER(OracleScene), vp, vp, vp, u64, i32]
        L.oracle_trace_any.restype = None
        L.oracle_brute_force_tris.argtypes = [vp, vp, u64, vp, vp, u64, vp, vp, vp, i32]
        L.oracle_brute_force_tris.restype = None
        L.oracle_resolve.argtypes = [ctypes.POINTER(OracleScene), vp, vp, u32, vp, vp, u32, u32, u32, u32, vp, u32, vp,
                                     vp, vp]
        L.oracle_resolve.restype = None
        L.oracle_tonemap.argtypes = [vp, u64, u32, vp]
        L.oracle_tonemap.restype = None
        L.oracle_eye_rays.argtypes = [vp, vp, u32, u32, vp, u32, vp, vp]
        L.oracle_eye_rays.restype = None
        L.oracle_camera_inverse.argtypes = [vp, vp, vp, f32, f32, f32, f32, vp, vp]
        L.oracle_camera_inverse.restype = None
        L.oracle_inverse4x4.argtypes = [vp, vp]
        L.oracle_inverse4x4.restype = None
        L.oracle_local_surface.argtypes = [vp, vp, f32, f32, f32, f32, vp, vp, vp]
        L.oracle_local_surface.restype = i32
        L.oracle_coct_voxel.argtypes = [ctypes.POINTER(OracleScene), u32, u32, u32, i32, i32, i32, f32, u32, vp]
        L.oracle_coct_voxel.restype = f32
        _libs[flavor] = L
    return _libs[flavor]


def _p(a):
    return None if a is None else a.ctypes.data


class Scene:
    """Holds the host arrays of a committed scene and the OracleScene view over them."""

    def __init__(self, arrays: dict, preset: np.ndarray):
        self._keep = {}
        self.view = OracleScene()
        for name in _ARRAYS:
            a = arrays.get(name)
            if a is not None and a.size:
                a = np.ascontiguousarray(a)
                self._keep[name] = a
                setattr(self.view, name, a.ctypes.data)
            else:
                setattr(self.view, name, None)
        raw = np.ascontiguousarray(preset).tobytes()
        assert len(raw) == 60, "MultiRenderPreset is 60 bytes"
        ctypes.memmove(self.view.preset, raw, 60)
        hs = arrays.get("scom2_headers")
        self.view.n_scom2 = 0 if hs is None else int(np.asarray(hs).view(np.uint8).size // 216)

    @classmethod
    def from_scene(cls, scene) -> "Scene":
        """Copies every array out of a product SceneRT (its host-side mirror)."""
        return cls({n: scene.get_array(n) for n in _ARRAYS}, scene.get_array("preset"))


def trace_closest(scene: Scene, pos: np.ndarray, dir: np.ndarray, threads: int = 0, flavor: str = "exact",
                  counts: bool = False):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 4)
    dir = np.ascontiguousarray(dir, dtype=np.float32).reshape(-1, 4)
    n = pos.shape[0]
    hits = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros(n, COUNTS_DTYPE) if counts else None
    ov = load(flavor).oracle_trace_closest(ctypes.byref(scene.view), _p(pos), _p(dir), _p(hits), n, threads,
                                           _p(cnt))
    return (hits, int(ov), cnt) if counts else (hits, int(ov))


def trace_any(scene: Scene, pos, dir, threads: int = 0, flavor: str = "exact") -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 4)
    dir = np.ascontiguousarray(dir, dtype=np.float32).reshape(-1, 4)
    flags = np.zeros(pos.shape[0], np.uint32)
    load(flavor).oracle_trace_any(ctypes.byref(scene.view), _p(pos), _p(dir), _p(flags), pos.shape[0], threads)
    return flags


def brute_force_tris(verts: np.ndarray, tris: np.ndarray, pos, dir, threads: int = 0, flavor: str = "exact"):
    v4 = np.ones((verts.shape[0], 4), np.float32)
    v4[:, :3] = verts[:, :3]
    tris = np.ascontiguousarray(tris, dtype=np.uint32).reshape(-1)
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 4)
    dir = np.ascontiguousarray(dir, dtype=np.float32).reshape(-1, 4)
    n = pos.shape[0]
    t = np.zeros(n, np.float32)
    ids = np.zeros(n, np.uint32)
    t2 = np.zeros(n, np.float32)
    load(flavor).oracle_brute_force_tris(_p(v4), _p(tris), tris.size // 3, _p(pos), _p(dir), n, _p(t), _p(ids),
                                         _p(t2), threads)
    return t, ids, t2


def tlas_refit(tlas: np.ndarray, instances: np.ndarray) -> np.ndarray:
    """The TLAS with its topology kept and its boxes recomputed from the instances' boxes: a leaf
    takes its instance's (BVHRT::UpdateInstance, BVH2Common_host.cpp:1303-1335, computes them),
    an inner node the componentwise min/max of its two children at leftOffset and leftOffset + 1
    (the node layout of BuildBVH over boxes, cbvh.cpp:597-664). Pure numpy, children before
    parents by recursion from the two root children."""
    out = tlas.copy()
    leaf = (out["leftOffset"] & np.uint32(0x80000000)) != 0

    def rec(i):
        if leaf[i]:
            k = int(out["leftOffset"][i]) & 0x0FFFFFFF
            out["boxMin"][i] = instances["boxMin"][k][:3]
            out["boxMax"][i] = instances["boxMax"][k][:3]
        else:
            c = int(out["leftOffset"][i])
            rec(c)
            rec(c + 1)
            out["boxMin"][i] = np.minimum(out["boxMin"][c], out["boxMin"][c + 1])
            out["boxMax"][i] = np.maximum(out["boxMax"][c], out["boxMax"][c + 1])

    import sys
    sys.setrecursionlimit(max(sys.getrecursionlimit(), 10000))
    rec(0)
    if out.size > 1:
        rec(1)
    return out


def eye_rays(proj_inv, wv_inv, width, height, jitter=None, order=1, flavor: str = "exact"):
    pi = np.ascontiguousarray(proj_inv, dtype=np.float32)
    vi = np.ascontiguousarray(wv_inv, dtype=np.float32)
    jt = None if jitter is None else np.ascontiguousarray(jitter, dtype=np.float32)
    pos = np.zeros((width * height, 4), np.float32)
    dir = np.zeros((width * height, 4), np.float32)
    load(flavor).oracle_eye_rays(_p(pi), _p(vi), width, height, _p(jt), order, _p(pos), _p(dir))
    return pos, dir


def camera_inverse(eye, target, up, fovy, aspect, znear, zfar, flavor: str = "exact"):
    e, t, u = (np.ascontiguousarray(v, dtype=np.float32) for v in (eye, target, up))
    pi = np.zeros(16, np.float32)
    vi = np.zeros(16, np.float32)
    load(flavor).oracle_camera_inverse(_p(e), _p(t), _p(u), fovy, aspect, znear, zfar, _p(pi), _p(vi))
    return pi, vi


def inverse4x4(m16: np.ndarray, flavor: str = "exact") -> np.ndarray:
    m = np.ascontiguousarray(m16, dtype=np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    load(flavor).oracle_inverse4x4(_p(m), _p(out))
    return out


def local_surface(preset: np.ndarray, values, d: float, q_near: float, q_far: float, f_near: float, start_q, ray_dir,
                  flavor: str = "exact"):
    """One LocalSurfaceIntersection on a voxel of size d (identity instance); returns the hit record."""
    raw = np.ascontiguousarray(preset)
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(8)
    sq = np.ascontiguousarray(start_q, dtype=np.float32).reshape(3)
    rd = np.ascontiguousarray(ray_dir, dtype=np.float32).reshape(3)
    h = np.zeros((), HIT_DTYPE)
    load(flavor).oracle_local_surface(_p(raw), _p(v), d, q_near, q_far, f_near, _p(sq), _p(rd), _p(h))
    return h


def resolve(scene: Scene, orig_indices, gbuf, proj_inv, wv_inv, width, height, order=1, sample_id=0, aa_frame=0,
            lights=None, background=(0.0, 0.0, 0.0, 1.0), highlight=None, colour=None, flavor: str = "exact"):
    """MultiRenderer's g-buffer resolve (MultiRenderer_resolve.cpp:529-1005) on the host; lights:
    32-B Light records (uint8/structured array); returns float32 (width*height, 4)."""
    L = load(flavor)
    out = np.zeros((width * height, 4), np.float32) if colour is None else colour
    lt = np.ascontiguousarray(lights if lights is not None else np.zeros(0, np.uint8))
    hl = np.zeros(4, np.uint32)
    if highlight is not None:
        hl[0] = 1
        hl[1:] = highlight
    bg = np.ascontiguousarray(background, dtype=np.float32)
    oi = np.ascontiguousarray(orig_indices, dtype=np.uint32)
    g = np.ascontiguousarray(gbuf)
    pi = np.ascontiguousarray(proj_inv, dtype=np.float32)
    vi = np.ascontiguousarray(wv_inv, dtype=np.float32)
    L.oracle_resolve(ctypes.byref(scene.view), _p(oi), _p(g), order, _p(pi), _p(vi), width, height, sample_id, aa_frame,
                     _p(lt) if lt.size else None, int(lt.size if lt.dtype.itemsize == 32 else lt.nbytes // 32),
                     _p(bg), _p(hl), _p(out))
    return out


def tonemap(colour: np.ndarray, samples: int = 1, flavor: str = "exact") -> np.ndarray:
    c = np.ascontiguousarray(colour, dtype=np.float32)
    out = np.zeros(c.shape[0], np.uint32)
    load(flavor).oracle_tonemap(_p(c), c.shape[0], samples, _p(out))
    return out
