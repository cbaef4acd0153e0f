"""MI355X (gfx950) ray-query library: Python mirror of LiteRT's ISceneObject / BVHRT.

The compute path is the C-ABI shared library ``liblrt_hip.so`` built from ``csrc/``
(hand-written HIP kernels for gfx950 plus the host scene assembly). This module only binds
it with ctypes; it has no CPU fallback. If the library is missing, every entry point raises.

Reference interface mirrored here:
  ISceneObject  CrossRT/CrossRT.h:43-175
  BVHRT         MultiRenderer/BVH2Common.h:66-381
  CreateSceneRT MultiRenderer/BVH2Common_host.cpp:1594-1600
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LRT_LIBRARY overrides the library file (the host-sanitizer build: `make -C csrc asan`)
LIB_PATH = os.environ.get("LRT_LIBRARY") or os.path.join(_HERE, "liblrt_hip.so")

# ---- reference structs as numpy dtypes (byte-compatible, see csrc/lrt_types.h) ----
HIT_DTYPE = np.dtype([("t", "<f4"), ("primId", "<u4"), ("instId", "<u4"), ("geomId", "<u4"),
                      ("coords", "<f4", (4,))])                                   # CRT_Hit, 32 B
NODE_DTYPE = np.dtype([("boxMin", "<f4", (3,)), ("leftOffset", "<u4"), ("boxMax", "<f4", (3,)),
                       ("escapeIndex", "<u4")])                                    # BVHNode, 32 B
NODE_PAIR_DTYPE = np.dtype([("left", NODE_DTYPE), ("right", NODE_DTYPE)])         # BVHNodePair, 64 B
GEOM_DTYPE = np.dtype([("boxMin", "<f4", (4,)), ("boxMax", "<f4", (4,)), ("bvhOffset", "<u4"),
                       ("type", "<u4"), ("offset", "<u4", (2,)), ("hasChannels", "<u4"),
                       ("_pad", "<u4", (3,))])                                     # GeomData, 64 B
INSTANCE_DTYPE = np.dtype([("boxMin", "<f4", (4,)), ("boxMax", "<f4", (4,)), ("geomId", "<u4"),
                           ("_pad", "<u4", (7,)), ("transform", "<f4", (16,)),
                           ("transformInv", "<f4", (16,)),
                           ("transformInvTransposed", "<f4", (16,))])              # InstanceData, 256 B
FRAME_NODE_DTYPE = np.dtype([("values", "<f4", (8,)), ("offset", "<u4")])         # SdfFrameOctreeNode, 36 B
FRAME_TEX_NODE_DTYPE = np.dtype([("tex_coords", "<f4", (16,)), ("values", "<f4", (8,)), ("offset", "<u4"),
                                 ("material_id", "<u4")])                           # SdfFrameOctreeTexNode, 104 B
MULTI_NODE_DTYPE = np.dtype([("data_offset", "<u4"), ("children_offset", "<u4"), ("voxel_count", "<u4"),
                             ("flags", "<u4")])                                       # SdfMultiOctreeNode, 16 B
GPRIM_HEADER_DTYPE = np.dtype([("color", "<f4", (3,)), ("prim_type", "<u4")])   # GraphicsPrimHeader, 16 B
(GRAPH_PRIM_POINT, GRAPH_PRIM_LINE, GRAPH_PRIM_LINE_SEGMENT, GRAPH_PRIM_LINE_SEGMENT_DIR, GRAPH_PRIM_BOX,
 GRAPH_PRIM_POINT_COLOR, GRAPH_PRIM_LINE_COLOR, GRAPH_PRIM_LINE_SEGMENT_COLOR, GRAPH_PRIM_LINE_SEGMENT_DIR_COLOR,
 GRAPH_PRIM_BOX_COLOR) = range(10)
# COctreeV3Header (sdf/scom/scom_shared.h:24-44), 60 B, and its constants (scom_shared.h:4-16)
COCTREE_HEADER_DTYPE = np.dtype([(n, "<u4") for n in (
    "brick_size", "brick_pad", "bits_per_value", "uv_size", "sim_compression", "lods", "default_leaf_type",
    "fallback_leaf_type", "node_pack_mode", "uints_per_child_info", "idx_mask", "idx_sh", "trans_off", "rot_mask",
    "add_mask")])
COCTREE_NODE_PACK_MODE_DEFAULT, COCTREE_NODE_PACK_MODE_SIM_COMP_FULL, COCTREE_NODE_PACK_MODE_SIM_COMP_SMALL = 0, 1, 2
COCTREE_LEAF_TYPE_NOT_A_LEAF, COCTREE_LEAF_TYPE_GRID, COCTREE_LEAF_TYPE_BIT_PACK, COCTREE_LEAF_TYPE_SLICES = 0, 1, 2, 3
COCTREE_LOD_LEAF_TYPE_SHIFT = 30
COCTREE_V3_VERSION = 5
# per pack mode: uints_per_child_info, idx_mask, idx_sh, trans_off, rot_mask, add_mask (scom.h:27-61)
COCTREE_PACK_MODE_WORDS = {0: (1, 0xFFFFFFFF, 0, 0, 0x0, 0x0), 1: (2, 0xFFFFFFFF, 0, 1, 0x3F, 0x3FFFFF00),
                           2: (1, 0xFFFFC000, 14, 0, 0x3F, 0x3FC0)}
# SdfDAG records (sdf/scom2/scom_shared.h:77-119)
DAG_HEADER_DTYPE = np.dtype([("brick_size", "<u4"), ("brick_pad", "<u4"), ("node_grid_size", "<u4"), ("dim", "<u4"),
                             ("transform_subgroup", "<u4"), ("user_params", "<f4", (7,)), ("tex_id_off", "<i4"),
                             ("mat_id_off", "<i4"), ("all_float_tex_id_off", "<i4"), ("all_int_mat_id_off", "<i4")])
DAG_CHILD_EDGE_DTYPE = np.dtype([("child_offset", "<u4"), ("rotation_id", "<u4")])
DAG_DATA_EDGE_DTYPE = np.dtype([("data_offset", "<u4"), ("rotation_id", "<u4"), ("type_id", "<u4"), ("add", "<f4")])
DAG_NODE_DTYPE = np.dtype([("children_edges_offset", "<u4"), ("data_edges_offset", "<u4"),
                           ("channels_edge", DAG_CHILD_EDGE_DTYPE), ("voxel_count_flags", "<u4"), ("_pad", "<u4")])
DAG_GEOM_DTYPE = np.dtype([(n, "<u4") for n in ("root", "brick_size", "brick_pad", "node_grid_size", "mods", "_p0",
                                                 "_p1", "_p2")])


def dag_header(brick_size=1, brick_pad=0, node_grid_size=2) -> np.ndarray:
    """get_default_SdfDAGHeader (sdf/scom2/scom_shared.h:121-146) with the given brick shape."""
    h = np.zeros((), DAG_HEADER_DTYPE)
    h["brick_size"], h["brick_pad"], h["node_grid_size"], h["dim"] = brick_size, brick_pad, node_grid_size, 3
    h["tex_id_off"] = h["mat_id_off"] = h["all_float_tex_id_off"] = h["all_int_mat_id_off"] = -1
    return h


# SCom2Header (sdf/scom2/scom_shared.h:21-75), 216 B
SCOM2_HEADER_DTYPE = np.dtype([(n, "<u4") for n in (
    "brick_size", "v_size", "bits_per_value", "values_per_uint", "value_mask", "bitmask_len", "dimension",
    "child_rot_shift", "child_rot_mask", "child_add_shift", "child_add_mask", "child_offset_mask", "child_offset_off",
    "node_offset_mask", "uints_per_link", "unique_brick_prefix", "unique_brick_offset_mask", "children_types_shift",
    "children_types_mask", "base_reference_shift", "children_active_bits_shift", "children_active_bits_mask",
    "references_offset", "reference_bits", "reference_mask", "references_per_uint", "links_offset",
    "max_surface_count", "max_surface_count_per_leaf", "bricks_step", "bricks_arr_offset", "nodes_arr_offset",
    "root_node_off", "has_channels", "has_surfaces", "has_multi_nodes")] +
    [(n, "<i4") for n in ("tex_id_off", "mat_id_off", "all_float_tex_id_off", "all_int_mat_id_off")] +
    [("max_val", "<f4"), ("max_depth", "<u4"), ("user_params", "<f4", (7,)), ("_pad", "<u4", (5,))])

COCT_GEOM_DTYPE = np.dtype([(n, "<u4") for n in (
    "base", "mods", "size", "brick_size", "brick_pad", "bits_per_value", "uv_size", "sim_compression", "lods",
    "uints_per_child_info", "idx_mask", "idx_sh", "trans_off", "rot_mask", "add_mask", "_pad")])


def coctree_header(brick_size=1, brick_pad=0, bits_per_value=8, uv_size=0, sim_compression=0, lods=0,
                   node_pack_mode=COCTREE_NODE_PACK_MODE_DEFAULT, default_leaf_type=COCTREE_LEAF_TYPE_SLICES,
                   fallback_leaf_type=COCTREE_LEAF_TYPE_GRID) -> np.ndarray:
    """A COctreeV3Header with the words node_pack_mode fixes filled in (fill_coctree_v3_header,
    sdf/scom/scom.h:27-61)."""
    h = np.zeros((), COCTREE_HEADER_DTYPE)
    for k, v in dict(brick_size=brick_size, brick_pad=brick_pad, bits_per_value=bits_per_value, uv_size=uv_size,
                     sim_compression=sim_compression, lods=lods, default_leaf_type=default_leaf_type,
                     fallback_leaf_type=fallback_leaf_type, node_pack_mode=node_pack_mode).items():
        h[k] = v
    upc, idx_mask, idx_sh, trans_off, rot_mask, add_mask = COCTREE_PACK_MODE_WORDS[node_pack_mode]
    h["uints_per_child_info"], h["idx_mask"], h["idx_sh"] = upc, idx_mask, idx_sh
    h["trans_off"], h["rot_mask"], h["add_mask"] = trans_off, rot_mask, add_mask
    return h


SVS_NODE_DTYPE = np.dtype([("pos_xy", "<u4"), ("pos_z_lod_size", "<u4"), ("values", "<u4", (2,))])  # SdfSVSNode, 16 B
SBS_NODE_DTYPE = np.dtype([("pos_xy", "<u4"), ("pos_z_lod_size", "<u4"), ("data_offset", "<u4"),
                           ("_pad", "<u4")])                                       # SdfSBSNode, 16 B
SBS_HEADER_DTYPE = np.dtype([("brick_size", "<u4"), ("brick_pad", "<u4"), ("bytes_per_value", "<u4"),
                             ("aux_data", "<u4")])                                 # SdfSBSHeader, 16 B
PRESET_DTYPE = np.dtype([("render_mode", "<u4"), ("sdf_node_intersect", "<u4"), ("normal_mode", "<u4"),
                         ("ray_gen_mode", "<u4"), ("interpolation_mode", "<u4"),
                         ("representation_mode", "<u4"), ("spp", "<u4"), ("fixed_lod", "<u4"),
                         ("level_of_detail", "<f4"), ("compact_sdf_eps", "<f4"), ("LOD_distance", "<f4"),
                         ("channel_render_mode", "<u4"), ("channel_id", "<u4"), ("component_id", "<u4"),
                         ("CTF_id", "<u4")])                                       # MultiRenderPreset, 60 B

ARRAY_DTYPES = {
    "nodes": NODE_PAIR_DTYPE, "tlas": NODE_DTYPE, "instances": INSTANCE_DTYPE, "geoms": GEOM_DTYPE,
    "geom_tags": np.uint32, "start_end": np.uint32, "prim_id_count": np.uint32,
    "vert_pos": np.float32, "vert_norm": np.float32, "indices": np.uint32, "prim_indices": np.uint32,
    "orig_nodes": NODE_DTYPE, "frame_nodes": FRAME_NODE_DTYPE, "frame_roots": np.uint32,
    "frame_tex_nodes": FRAME_TEX_NODE_DTYPE, "frame_tex_roots": np.uint32,
    "sbs_nodes": SBS_NODE_DTYPE, "sbs_data": np.uint32, "sbs_data_f": np.float32, "sbs_roots": np.uint32,
    "sbs_headers": SBS_HEADER_DTYPE, "multi_nodes": MULTI_NODE_DTYPE, "multi_values": np.float32,
    "preset": PRESET_DTYPE, "sbs_mask": np.uint32, "orig_indices": np.uint32, "svs_nodes": SVS_NODE_DTYPE,
    "svs_roots": np.uint32, "tlas_parent": np.uint32, "grid_data": np.float32, "grid_offsets": np.uint32,
    "grid_sizes": np.uint32, "gprim_headers": GPRIM_HEADER_DTYPE, "gprim_points": np.float32,
    "coct_data": np.uint32, "coct_geoms": COCT_GEOM_DTYPE, "coct_mods": np.int32,
    "dag_nodes": DAG_NODE_DTYPE, "dag_child_edges": DAG_CHILD_EDGE_DTYPE, "dag_data_edges": DAG_DATA_EDGE_DTYPE,
    "dag_distances": np.float32, "dag_geoms": DAG_GEOM_DTYPE, "dag_mods": np.int32, "rot_add": np.uint8,
    "scom2_data": np.uint32, "scom2_headers": SCOM2_HEADER_DTYPE, "scom2_mods": np.int32,
}
LIGHT_DTYPE = np.dtype([("space", "<f4", (3,)), ("type", "<u4"), ("color", "<f4", (3,)), ("_pad", "<u4")])  # Light, 32 B
LIGHT_TYPE_DIRECT, LIGHT_TYPE_POINT, LIGHT_TYPE_AMBIENT = 0, 1, 2


def default_lights() -> np.ndarray:
    """MultiRenderer's lights (MultiRenderer_host.cpp:55-56): a direct light along normalize(1,1,1)
    of colour 2/3, an ambient light of 0 and one of 0.25 (create_*_light, core/IRenderer.h:80-93)."""
    L = np.zeros(3, LIGHT_DTYPE)
    inv = np.float32(1.0) / np.sqrt(np.float32(3.0))   # LiteMath normalize: v * (1 / length(v))
    L[0]["space"] = np.float32(1.0) * inv
    L[0]["type"], L[0]["color"] = LIGHT_TYPE_DIRECT, np.float32(2.0) / np.float32(3.0)
    L[1]["type"], L[1]["color"] = LIGHT_TYPE_AMBIENT, 0.0
    L[2]["type"], L[2]["color"] = LIGHT_TYPE_AMBIENT, 0.25
    return L

# enums (core/render_settings.h, BVH2Common.h:30-43, sdf_simple_shared.h:13-21)
TYPE_MESH_TRIANGLE, TYPE_SDF_FRAME_OCTREE, TYPE_SDF_SBS, TYPE_SDF_SBS_COL = 0, 3, 6, 11
TYPE_SDF_FRAME_OCTREE_TEX = 8
SH_TYPE, GEOM_ID_MASK = 24, 0x00FFFFFF
MISS = 0xFFFFFFFF
NOALLOC_MAX_RAYS = 65536   # LRT_NOALLOC_MAX_RAYS (include/lrt.h)
SDF_SBS_NODE_LAYOUT_DX = 1 << 24
SDF_SBS_NODE_LAYOUT_DX_UV16 = 2 << 24
SDF_SBS_NODE_LAYOUT_DX_RGB8 = 3 << 24
SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F = 4 << 24
SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN = 5 << 24
SDF_OCTREE_NODE_INTERSECT_ST, SDF_OCTREE_NODE_INTERSECT_ANALYTIC = 0, 1
SDF_OCTREE_NODE_INTERSECT_NEWTON, SDF_OCTREE_NODE_INTERSECT_BBOX, SDF_OCTREE_NODE_INTERSECT_IT = 2, 3, 4
MULTI_RENDER_MODE_MASK, MULTI_RENDER_MODE_LAMBERT_NO_TEX = 0, 1
NORMAL_MODE_GEOMETRY, NORMAL_MODE_VERTEX = 0, 1
BUILD_HIGH = 4


class LrtError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero lrt_status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"lrt error {code}: {msg}")
        self.code = code


class _TuneResult(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("lds_stack", ctypes.c_int32), ("xcd_remap", ctypes.c_int32),
                ("budget", ctypes.c_uint32), ("tail_waves", ctypes.c_uint32), ("best_ms", ctypes.c_float),
                ("configs_tried", ctypes.c_uint32), ("leaf_postpone", ctypes.c_int32),
                ("ray_interleave", ctypes.c_uint32), ("persistent", ctypes.c_int32)]


class _BatchTuneResult(ctypes.Structure):
    _fields_ = [("lds_stack", ctypes.c_int32), ("record", ctypes.c_int32), ("grid", ctypes.c_uint32),
                ("best_ms", ctypes.c_float), ("configs_tried", ctypes.c_uint32)]


class _CommitInfo(ctypes.Structure):
    """lrt_commit_info (include/lrt.h)."""
    _fields_ = [("kind", ctypes.c_uint32), ("gpu_builder", ctypes.c_uint32), ("blas_max_depth", ctypes.c_uint32),
                ("blas_leaves", ctypes.c_uint32), ("blas_merged_leaves", ctypes.c_uint32),
                ("tlas_nodes", ctypes.c_uint32), ("tlas_area_ratio", ctypes.c_float)]


COMMIT_FULL, COMMIT_REFIT, COMMIT_TLAS_REBUILD = 0, 1, 2


class _Buffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("bytes", ctypes.c_size_t)]


class _ResolveDesc(ctypes.Structure):
    """lrt_resolve_desc (include/lrt.h)."""
    _fields_ = [("proj_inv16", ctypes.c_void_p), ("world_view_inv16", ctypes.c_void_p), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("order", ctypes.c_uint32), ("sample_id", ctypes.c_uint32),
                ("aa_frame", ctypes.c_uint32), ("lights", ctypes.c_void_p), ("n_lights", ctypes.c_uint32),
                ("background", ctypes.c_float * 4), ("highlight_mode", ctypes.c_uint32),
                ("highlight_prim", ctypes.c_uint32), ("highlight_geom", ctypes.c_uint32),
                ("highlight_inst", ctypes.c_uint32)]


_lib_handle = None


def library_path() -> str:
    return LIB_PATH


def lib() -> ctypes.CDLL:
    """Loads liblrt_hip.so; raises if it has not been built (there is no fallback)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        try:
            # When torch is present its HIP runtime must be the one the library binds to, so that
            # device pointers and streams handed over from torch are valid in both.
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, sz, f32, cp = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t,
                                     ctypes.c_float, ctypes.c_char_p)
        i32 = ctypes.c_int
        pu32 = ctypes.POINTER(ctypes.c_uint32)
        sigs = {
            "lrt_last_error": ([], cp), "lrt_version": ([], cp),
            "lrt_create": ([cp, cp, cp, ctypes.POINTER(vp)], i32),
            "lrt_destroy": ([vp], i32), "lrt_clear_geom": ([vp], i32), "lrt_clear_scene": ([vp], i32),
            "lrt_add_triangles": ([vp, vp, sz, sz, vp, vp, sz, u32, pu32], i32),
            "lrt_add_sdf_frame_octree": ([vp, vp, sz, pu32], i32),
            "lrt_add_sdf_frame_octree_tex": ([vp, vp, sz, pu32], i32),
            "lrt_add_sdf_svs": ([vp, vp, sz, pu32], i32),
            "lrt_add_sdf_grid": ([vp, pu32, vp, pu32], i32),
            "lrt_add_graphics_prims": ([vp, vp, vp, sz, pu32], i32),
            "lrt_load_sdf_grid": ([cp, pu32, ctypes.POINTER(_Buffer)], i32),
            "lrt_add_coctree_v3": ([vp, vp, vp, sz, u32, pu32], i32),
            "lrt_add_sdf_dag": ([vp, vp, vp, sz, vp, sz, vp, sz, vp, sz, pu32], i32),
            "lrt_add_scom2": ([vp, vp, vp, sz, vp, sz, pu32], i32),
            "lrt_load_scom2": ([cp, vp, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_save_scom2": ([cp, vp, vp, sz, vp, sz], i32),
            "lrt_load_sdf_dag": ([cp, vp] + [ctypes.POINTER(_Buffer)] * 4, i32),
            "lrt_save_sdf_dag": ([cp, vp, vp, sz, vp, sz, vp, sz, vp, sz], i32),
            "lrt_load_coctree_v3": ([cp, pu32, vp, ctypes.POINTER(_Buffer)], i32),
            "lrt_save_coctree_v3": ([cp, vp, vp, sz], i32),
            "lrt_save_sdf_grid": ([cp, pu32, vp], i32),
            "lrt_load_sdf_svs": ([cp, ctypes.POINTER(_Buffer)], i32),
            "lrt_save_sdf_svs": ([cp, vp, sz], i32),
            "lrt_add_sdf_sbs": ([vp, vp, vp, sz, vp, sz, vp, sz, pu32], i32),
            "lrt_add_instance": ([vp, u32, vp, pu32], i32),
            "lrt_update_instance": ([vp, u32, vp], i32),
            "lrt_commit": ([vp, u32], i32), "lrt_set_preset": ([vp, vp], i32),
            "lrt_trace_closest_dev": ([vp, vp, vp, vp, u64, vp], i32),
            "lrt_trace_closest_host": ([vp, vp, vp, vp, u64], i32),
            "lrt_trace_closest_noalloc_dev": ([vp, vp, vp, vp, u64, vp], i32),
            "lrt_trace_any_dev": ([vp, vp, vp, vp, u64, vp], i32),
            "lrt_trace_closest_counted_dev": ([vp, vp, vp, vp, vp, u64, vp], i32),
            "lrt_query_one": ([vp, vp, vp, vp], i32),
            "lrt_eye_rays_dev": ([vp, vp, u32, u32, vp, u32, vp, vp, vp], i32),
            "lrt_set_launch_config": ([vp, i32, i32, i32], i32),
            "lrt_set_tail_config": ([vp, u32, u32], i32),
            "lrt_set_leaf_postponement": ([vp, i32], i32),
            "lrt_set_ray_interleave": ([vp, u32], i32),
            "lrt_set_persistent": ([vp, i32], i32),
            "lrt_set_tlas_update": ([vp, i32], i32),
            "lrt_get_commit_info": ([vp, ctypes.POINTER(_CommitInfo)], i32),
            "lrt_autotune": ([vp, vp, vp, vp, u64, vp, vp], i32),
            "lrt_sbs_param_layout": ([vp, u32, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u64)], i32),
            "lrt_sbs_forward_dev": ([vp, u32, vp, vp, vp, u64, vp, vp], i32),
            "lrt_sbs_backward_dev": ([vp, u32, vp, vp, vp, u64, vp, vp, vp], i32),
            "lrt_sbs_l2_backward_dev": ([vp, u32, vp, vp, vp, u64, vp, vp, vp, vp], i32),
            "lrt_sbs_render_l2_backward_dev": ([vp, u32, vp, vp, vp, u64, vp, vp, vp, vp, vp], i32),
            "lrt_sbs_recompute_misses": ([vp, ctypes.POINTER(u64), i32], i32),
            "lrt_sbs_batch_create": ([vp, u32, ctypes.POINTER(vp)], i32),
            "lrt_sbs_batch_run": ([vp, vp], i32),
            "lrt_sbs_batch_autotune": ([vp, vp, ctypes.POINTER(_BatchTuneResult)], i32),
            "lrt_sbs_batch_configure": ([vp, i32, i32, u32], i32),
            "lrt_sbs_batch_destroy": ([vp], i32),
            "lrt_stack_overflow_count": ([vp, ctypes.POINTER(u64), i32], i32),
            "lrt_get_array": ([vp, cp, vp, ctypes.POINTER(sz)], i32),
            "lrt_buffer_free": ([ctypes.POINTER(_Buffer)], None),
            "lrt_camera_inverse": ([vp, vp, vp, f32, f32, f32, f32, vp, vp], i32),
            "lrt_gen_icosphere": ([u32, f32, f32, u64, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_gen_geodesic": ([u32, f32, f32, u64, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_sdf_spheres_frame_octree": ([vp, sz, u32, i32, ctypes.POINTER(_Buffer)], i32),
            "lrt_sdf_spheres_global_octree": ([vp, sz, u32, u32, u32, i32, ctypes.POINTER(_Buffer),
                                               ctypes.POINTER(_Buffer)], i32),
            "lrt_add_sdf_multi_octree": ([vp, vp, sz, vp, sz, ctypes.POINTER(u32)], i32),
            "lrt_sdf_spheres_multi_octree": ([vp, sz, u32, i32, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_sdf_spheres_sbs": ([vp, sz, u32, vp, u64, i32, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_sdf_spheres_sbs_indexed": ([vp, sz, u32, u32, u64, i32, ctypes.POINTER(_Buffer),
                                             ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer), pu32], i32),
            "lrt_resolve_dev": ([vp, vp, ctypes.POINTER(_ResolveDesc), vp, vp], i32),
            "lrt_tonemap_dev": ([vp, u64, u32, vp, vp], i32),
            "lrt_load_sdf_frame_octree": ([cp, ctypes.POINTER(_Buffer)], i32),
            "lrt_save_sdf_frame_octree": ([cp, vp, sz], i32),
            "lrt_load_sdf_frame_octree_tex": ([cp, ctypes.POINTER(_Buffer)], i32),
            "lrt_save_sdf_frame_octree_tex": ([cp, vp, sz], i32),
            "lrt_load_sdf_sbs": ([cp, vp, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)],
                                 i32),
            "lrt_save_sdf_sbs": ([cp, vp, vp, sz, vp, sz, vp, sz], i32),
            "lrt_load_sdf_multi_octree": ([cp, ctypes.POINTER(_Buffer), ctypes.POINTER(_Buffer)], i32),
            "lrt_save_sdf_multi_octree": ([cp, vp, sz, vp, sz], i32),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib_handle = L
    return _lib_handle


def exported_symbols():
    return ["lrt_last_error", "lrt_version", "lrt_create", "lrt_destroy", "lrt_clear_geom", "lrt_clear_scene",
            "lrt_add_triangles", "lrt_add_sdf_frame_octree", "lrt_add_sdf_frame_octree_tex", "lrt_add_sdf_svs", "lrt_load_sdf_svs", "lrt_save_sdf_svs",
            "lrt_add_sdf_grid", "lrt_load_sdf_grid", "lrt_save_sdf_grid", "lrt_add_graphics_prims",
            "lrt_add_coctree_v3", "lrt_load_coctree_v3", "lrt_save_coctree_v3",
            "lrt_add_sdf_dag", "lrt_load_sdf_dag", "lrt_save_sdf_dag", "lrt_add_scom2", "lrt_load_scom2",
            "lrt_save_scom2", "lrt_add_sdf_sbs", "lrt_add_sdf_multi_octree", "lrt_sdf_spheres_multi_octree", "lrt_add_instance",
            "lrt_update_instance", "lrt_commit", "lrt_set_preset", "lrt_trace_closest_dev",
            "lrt_trace_closest_host", "lrt_trace_closest_noalloc_dev", "lrt_trace_any_dev", "lrt_trace_closest_counted_dev", "lrt_query_one", "lrt_eye_rays_dev",
            "lrt_set_launch_config", "lrt_set_tail_config", "lrt_set_leaf_postponement", "lrt_set_ray_interleave", "lrt_set_persistent",
            "lrt_set_tlas_update", "lrt_get_commit_info", "lrt_autotune", "lrt_sbs_param_layout",
            "lrt_sbs_forward_dev", "lrt_sbs_backward_dev", "lrt_sbs_l2_backward_dev", "lrt_sbs_render_l2_backward_dev", "lrt_sbs_recompute_misses",
            "lrt_sbs_batch_create", "lrt_sbs_batch_run", "lrt_sbs_batch_autotune", "lrt_sbs_batch_configure", "lrt_sbs_batch_destroy",
            "lrt_stack_overflow_count", "lrt_get_array", "lrt_buffer_free",
            "lrt_camera_inverse", "lrt_gen_icosphere", "lrt_gen_geodesic", "lrt_sdf_spheres_frame_octree",
            "lrt_sdf_spheres_global_octree", "lrt_sdf_spheres_sbs", "lrt_sdf_spheres_sbs_indexed", "lrt_resolve_dev", "lrt_tonemap_dev",
            "lrt_load_sdf_frame_octree",
            "lrt_save_sdf_frame_octree", "lrt_load_sdf_frame_octree_tex", "lrt_save_sdf_frame_octree_tex", "lrt_load_sdf_sbs", "lrt_save_sdf_sbs", "lrt_load_sdf_multi_octree",
            "lrt_save_sdf_multi_octree"]


def _check(code: int):
    if code != 0:
        raise LrtError(code, lib().lrt_last_error().decode())


def _ptr(a) -> Optional[int]:
    """Address of a numpy array (host) or a torch tensor (device or host)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported buffer type {type(a)}")


def _take(buf: _Buffer, dtype) -> np.ndarray:
    try:
        raw = ctypes.string_at(buf.data, buf.bytes) if buf.bytes else b""
        return np.frombuffer(raw, dtype=dtype).copy()
    finally:
        lib().lrt_buffer_free(ctypes.byref(buf))


def default_preset() -> np.ndarray:
    """getDefaultPreset (core/render_settings.h:138-160)."""
    p = np.zeros((), dtype=PRESET_DTYPE)
    p["render_mode"] = MULTI_RENDER_MODE_LAMBERT_NO_TEX
    p["sdf_node_intersect"] = SDF_OCTREE_NODE_INTERSECT_ANALYTIC
    p["normal_mode"] = NORMAL_MODE_VERTEX
    p["interpolation_mode"] = 0
    p["representation_mode"] = 1
    p["spp"] = 1
    p["fixed_lod"] = 1
    p["level_of_detail"] = 16.0
    p["compact_sdf_eps"] = 0.05
    p["LOD_distance"] = 10.0
    return p


class SceneRT:
    """BVHRT behind the C ABI: the ISceneObject methods plus the SDF geometry adders."""

    def __init__(self, impl_name="BVH2Common", build_name="cbvh_embree2", layout_name="SuperTreeletAlignedMerged4"):
        h = ctypes.c_void_p()
        _check(lib().lrt_create(impl_name.encode(), build_name.encode(), layout_name.encode(), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib_handle is not None:
            _lib_handle.lrt_destroy(h)
            self._h = None

    # ---- ISceneObject (CrossRT/CrossRT.h:43-175) ----
    def Name(self) -> str:
        return "BVH2Fat"

    def ClearGeom(self):
        _check(lib().lrt_clear_geom(self._h))

    def AddGeom_Triangles3f(self, vpos: np.ndarray, indices: np.ndarray, flags: int = BUILD_HIGH,
                            vnorm: Optional[np.ndarray] = None) -> int:
        vpos = np.ascontiguousarray(vpos, dtype=np.float32)
        if vpos.ndim != 2 or vpos.shape[1] not in (3, 4):
            raise ValueError("vpos must be (n, 3) or (n, 4)")
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        if vnorm is not None:
            vnorm = np.ascontiguousarray(vnorm, dtype=np.float32)
            if vnorm.shape != vpos.shape:
                raise ValueError("vnorm must have the shape of vpos")
        out = ctypes.c_uint32()
        _check(lib().lrt_add_triangles(self._h, _ptr(vpos), vpos.shape[0], vpos.shape[1] * 4, _ptr(vnorm),
                                       _ptr(idx), idx.size, flags, ctypes.byref(out)))
        return out.value

    def AddGeom_SdfFrameOctree(self, nodes: np.ndarray) -> int:
        nodes = np.ascontiguousarray(nodes, dtype=FRAME_NODE_DTYPE)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_frame_octree(self._h, _ptr(nodes), nodes.size, ctypes.byref(out)))
        return out.value

    def AddGeom_SdfFrameOctreeTex(self, nodes: np.ndarray) -> int:
        """BVHRT::AddGeom_SdfFrameOctreeTex (BVH2Common_host.cpp:697-731): a frame octree whose nodes
        carry 16 corner texture coordinates; hits report TYPE_SDF_FRAME_OCTREE_TEX and the
        coordinates interpolated at the hit point in coords[0..1]."""
        nodes = np.ascontiguousarray(nodes, dtype=FRAME_TEX_NODE_DTYPE)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_frame_octree_tex(self._h, _ptr(nodes), nodes.size, ctypes.byref(out)))
        return out.value

    def AddGeom_GraphicsPrim(self, prim_type: int, points: np.ndarray, color=(1.0, 1.0, 1.0)) -> int:
        """BVHRT::AddGeom_GraphicsPrim (BVH2Common_host.cpp:767-903): points (n, 4) float32 as the
        primitive type reads them; color is the header colour of the types without their own."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 4)
        h = np.zeros((), GPRIM_HEADER_DTYPE)
        h["color"], h["prim_type"] = color, prim_type
        out = ctypes.c_uint32()
        _check(lib().lrt_add_graphics_prims(self._h, _ptr(h), _ptr(pts), pts.shape[0], ctypes.byref(out)))
        return out.value

    def AddGeom_SdfGrid(self, data: np.ndarray) -> int:
        """BVHRT::AddGeom_SdfGrid (BVH2Common_host.cpp:466-499): data[z, y, x] distances over the
        unit cube (a 3-D array; x fastest in memory)."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        if data.ndim != 3:
            raise ValueError("grid data must be a 3-D array indexed [z, y, x]")
        size = np.array([data.shape[2], data.shape[1], data.shape[0]], np.uint32)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_grid(self._h, size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _ptr(data),
                                      ctypes.byref(out)))
        return out.value

    def AddGeom_COctreeV3(self, header: np.ndarray, data: np.ndarray, bvh_level: int = 0) -> int:
        """BVHRT::AddGeom_COctreeV3 (BVH2Common_host.cpp:915-966): a COctreeV3Header record
        (coctree_header()) and the octree's u32 words; bvh_level is BVHRT::preferredBVHLevel."""
        h = np.ascontiguousarray(header, dtype=COCTREE_HEADER_DTYPE).reshape(())
        d = np.ascontiguousarray(data, dtype=np.uint32).ravel()
        out = ctypes.c_uint32()
        _check(lib().lrt_add_coctree_v3(self._h, _ptr(h), _ptr(d), d.size, int(bvh_level), ctypes.byref(out)))
        return out.value

    def AddGeom_SdfDAG(self, header: np.ndarray, nodes: np.ndarray, child_edges: np.ndarray,
                       data_edges: np.ndarray, distances: np.ndarray) -> int:
        """BVHRT::AddGeom_SdfDAG (BVH2Common_host.cpp:1020-1131): an SdfDAG's header and arrays."""
        h = np.ascontiguousarray(header, dtype=DAG_HEADER_DTYPE).reshape(())
        n = np.ascontiguousarray(nodes, dtype=DAG_NODE_DTYPE)
        ce = np.ascontiguousarray(child_edges, dtype=DAG_CHILD_EDGE_DTYPE)
        de = np.ascontiguousarray(data_edges, dtype=DAG_DATA_EDGE_DTYPE)
        d = np.ascontiguousarray(distances, dtype=np.float32)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_dag(self._h, _ptr(h), _ptr(n), n.size, _ptr(ce), ce.size, _ptr(de), de.size, _ptr(d),
                                     d.size, ctypes.byref(out)))
        return out.value

    def AddGeom_SCom2(self, header: np.ndarray, nodes: np.ndarray, bricks: np.ndarray) -> int:
        """BVHRT::AddGeom_SCom2 (BVH2Common_host.cpp:1133-1191): an SCom2Tree's header, node and brick words."""
        h = np.ascontiguousarray(header, dtype=SCOM2_HEADER_DTYPE).reshape(())
        n = np.ascontiguousarray(nodes, dtype=np.uint32).ravel()
        b = np.ascontiguousarray(bricks, dtype=np.uint32).ravel()
        out = ctypes.c_uint32()
        _check(lib().lrt_add_scom2(self._h, _ptr(h), _ptr(n), n.size, _ptr(b), b.size, ctypes.byref(out)))
        return out.value

    def AddGeom_SdfSVS(self, nodes: np.ndarray) -> int:
        """BVHRT::AddGeom_SdfSVS (BVH2Common_host.cpp:536-597): SdfSVSNode array."""
        nodes = np.ascontiguousarray(nodes, dtype=SVS_NODE_DTYPE)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_svs(self._h, _ptr(nodes), nodes.size, ctypes.byref(out)))
        return out.value

    def AddGeom_SdfMultiOctree(self, nodes: np.ndarray, values: np.ndarray) -> int:
        n = np.ascontiguousarray(nodes, dtype=MULTI_NODE_DTYPE).reshape(-1)
        v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        gid = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_multi_octree(self._h, _ptr(n), n.size, _ptr(v), v.size, ctypes.byref(gid)))
        return gid.value

    def AddGeom_SdfSBS(self, header: np.ndarray, nodes: np.ndarray, values: np.ndarray,
                       values_f: Optional[np.ndarray] = None) -> int:
        header = np.ascontiguousarray(header, dtype=SBS_HEADER_DTYPE).reshape(())
        nodes = np.ascontiguousarray(nodes, dtype=SBS_NODE_DTYPE)
        values = np.ascontiguousarray(values, dtype=np.uint32)
        vf = None if values_f is None else np.ascontiguousarray(values_f, dtype=np.float32)
        out = ctypes.c_uint32()
        _check(lib().lrt_add_sdf_sbs(self._h, _ptr(header), _ptr(nodes), nodes.size, _ptr(values), values.size,
                                     _ptr(vf), 0 if vf is None else vf.size, ctypes.byref(out)))
        return out.value

    def ClearScene(self):
        _check(lib().lrt_clear_scene(self._h))

    def AddInstance(self, geom_id: int, matrix: Optional[np.ndarray] = None) -> int:
        """matrix: 4x4 (row-major numpy view of a column-major LiteMath float4x4 is transposed here)."""
        m = np.eye(4, dtype=np.float32) if matrix is None else np.asarray(matrix, dtype=np.float32)
        col = np.ascontiguousarray(m.T.reshape(16))
        out = ctypes.c_uint32()
        _check(lib().lrt_add_instance(self._h, geom_id, _ptr(col), ctypes.byref(out)))
        return out.value

    def UpdateInstance(self, inst_id: int, matrix: np.ndarray):
        col = np.ascontiguousarray(np.asarray(matrix, dtype=np.float32).T.reshape(16))
        _check(lib().lrt_update_instance(self._h, inst_id, _ptr(col)))

    def CommitScene(self, flags: int = BUILD_HIGH):
        _check(lib().lrt_commit(self._h, flags))

    def SetPreset(self, preset: np.ndarray):
        p = np.ascontiguousarray(preset, dtype=PRESET_DTYPE).reshape(())
        _check(lib().lrt_set_preset(self._h, _ptr(p)))

    def RayQuery_NearestHit(self, pos_near, dir_far) -> np.ndarray:
        p = np.ascontiguousarray(pos_near, dtype=np.float32).reshape(4)
        d = np.ascontiguousarray(dir_far, dtype=np.float32).reshape(4)
        h = np.zeros((), dtype=HIT_DTYPE)
        _check(lib().lrt_query_one(self._h, _ptr(p), _ptr(d), _ptr(h)))
        return h

    def RayQuery_AnyHit(self, pos_near, dir_far) -> bool:
        d = np.asarray(dir_far, dtype=np.float32).reshape(4).copy()
        d[3] *= -1.0
        return bool(self.RayQuery_NearestHit(pos_near, d)["geomId"] != MISS)

    # ---- batched trace ----
    def trace_closest(self, pos, dir, hits, stream: int = 0):
        """Device buffers (torch tensors): pos/dir float32 (n,4), hits uint8/int32 storage of n*32 bytes."""
        n = pos.numel() // 4
        _check(lib().lrt_trace_closest_dev(self._h, _ptr(pos), _ptr(dir), _ptr(hits), n, stream))

    def resolve(self, gbuffer, proj_inv, world_view_inv, width: int, height: int, colour, order: int = 1,
                sample_id: int = 0, lights: Optional[np.ndarray] = None, background=(0.0, 0.0, 0.0, 1.0),
                aa_frame: int = 0, highlight=None, stream: int = 0):
        """MultiRenderer's resolve of a traced g-buffer (device, width*height hits in the eye-ray
        `order`) into `colour` (device float32 (width*height, 4)); mode, normal mode and spp from
        the scene's preset."""
        L = np.ascontiguousarray(default_lights() if lights is None else lights, dtype=LIGHT_DTYPE)
        pi = np.ascontiguousarray(proj_inv, dtype=np.float32).reshape(16)
        vi = np.ascontiguousarray(world_view_inv, dtype=np.float32).reshape(16)
        d = _ResolveDesc()
        d.proj_inv16, d.world_view_inv16 = _ptr(pi), _ptr(vi)
        d.width, d.height, d.order, d.sample_id, d.aa_frame = width, height, order, sample_id, aa_frame
        d.lights, d.n_lights = (_ptr(L) if L.size else None), L.size
        d.background[:] = [float(x) for x in background]
        if highlight is not None:
            d.highlight_mode = 1
            d.highlight_prim, d.highlight_geom, d.highlight_inst = (int(x) for x in highlight)
        _check(lib().lrt_resolve_dev(self._h, _ptr(gbuffer), ctypes.byref(d), _ptr(colour), stream))

    def trace_closest_noalloc(self, pos, dir, hits, stream: int = 0):
        """trace_closest through the allocation-free entry point (at most NOALLOC_MAX_RAYS rays;
        overlapping calls on one GPU share its work array and must be serialised)."""
        n = pos.numel() // 4
        _check(lib().lrt_trace_closest_noalloc_dev(self._h, _ptr(pos), _ptr(dir), _ptr(hits), n, stream))

    def trace_closest_counted(self, pos, dir, hits, counters, stream: int = 0):
        """As trace_closest, plus 8 uint32 words per ray: 6 traversal counters, start tick, elapsed ticks."""
        n = pos.numel() // 4
        _check(lib().lrt_trace_closest_counted_dev(self._h, _ptr(pos), _ptr(dir), _ptr(hits), _ptr(counters), n,
                                                   stream))

    def autotune(self, pos, dir, hits, stream: int = 0) -> dict:
        """Times the launch configurations on this device ray batch and keeps the fastest."""
        n = pos.numel() // 4
        r = _TuneResult()
        _check(lib().lrt_autotune(self._h, _ptr(pos), _ptr(dir), _ptr(hits), n, stream, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in _TuneResult._fields_}

    # ---- differentiable sparse brick set ----
    def sbs_param_layout(self, geom_id: int):
        """(f_offset, n_dist, n_params): the geometry's slice of values_f holds n_dist distances
        then n_dist RGB albedos; parameter arrays span the scene's whole values_f (n_params)."""
        off, nd, npar = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().lrt_sbs_param_layout(self._h, geom_id, ctypes.byref(off), ctypes.byref(nd), ctypes.byref(npar)))
        return off.value, nd.value, npar.value

    def sbs_forward(self, geom_id, params, pos, dir, colour, stream: int = 0):
        _check(lib().lrt_sbs_forward_dev(self._h, geom_id, _ptr(params), _ptr(pos), _ptr(dir), pos.numel() // 4,
                                         _ptr(colour), stream))

    def sbs_backward(self, geom_id, params, pos, dir, dL_dcolour, grad, stream: int = 0):
        _check(lib().lrt_sbs_backward_dev(self._h, geom_id, _ptr(params), _ptr(pos), _ptr(dir), pos.numel() // 4,
                                          _ptr(dL_dcolour), _ptr(grad), stream))

    def sbs_l2_backward(self, geom_id, params, pos, dir, target, grad, loss=None, stream: int = 0):
        _check(lib().lrt_sbs_l2_backward_dev(self._h, geom_id, _ptr(params), _ptr(pos), _ptr(dir), pos.numel() // 4,
                                             _ptr(target), _ptr(grad), _ptr(loss), stream))

    def sbs_render_l2_backward(self, geom_id, params, pos, dir, target, colour, grad, loss=None, stream: int = 0):
        """Forward colour and the L2 adjoint (grad and loss accumulated) in one traversal per ray."""
        _check(lib().lrt_sbs_render_l2_backward_dev(self._h, geom_id, _ptr(params), _ptr(pos), _ptr(dir),
                                                    pos.numel() // 4, _ptr(target), _ptr(colour), _ptr(grad),
                                                    _ptr(loss), stream))

    def sbs_recompute_misses(self, reset: bool = False) -> int:
        v = ctypes.c_uint64()
        _check(lib().lrt_sbs_recompute_misses(self._h, ctypes.byref(v), int(reset)))
        return v.value

    def trace_any(self, pos, dir, flags, stream: int = 0):
        n = pos.numel() // 4
        _check(lib().lrt_trace_any_dev(self._h, _ptr(pos), _ptr(dir), _ptr(flags), n, stream))

    def trace_closest_host(self, pos: np.ndarray, dir: np.ndarray) -> np.ndarray:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 4)
        dir = np.ascontiguousarray(dir, dtype=np.float32).reshape(-1, 4)
        hits = np.zeros(pos.shape[0], dtype=HIT_DTYPE)
        _check(lib().lrt_trace_closest_host(self._h, _ptr(pos), _ptr(dir), _ptr(hits), pos.shape[0]))
        return hits

    def set_launch_config(self, block: int = 256, lds_stack: int = 16, xcd_remap: bool = True):
        _check(lib().lrt_set_launch_config(self._h, block, lds_stack, int(xcd_remap)))

    def set_tail_config(self, budget: int = 64, tail_waves: int = 1024):
        _check(lib().lrt_set_tail_config(self._h, budget, tail_waves))

    def set_ray_interleave(self, lanes_per_group: int = 32):
        """Pass-1 ray-to-lane assignment: groups of `lanes_per_group` lanes take rays from distant
        parts of the batch (0 or 64: thread order)."""
        _check(lib().lrt_set_ray_interleave(self._h, lanes_per_group))

    def set_persistent(self, on: bool = True):
        """One persistent single-pass kernel with per-lane ray refill (True) or the two-pass grid launch."""
        _check(lib().lrt_set_persistent(self._h, 1 if on else 0))

    def set_tlas_update(self, mode: int = 1):
        """Commits after UpdateInstance only: 1 refit the TLAS on the GPU, 0 rebuild it on the host."""
        _check(lib().lrt_set_tlas_update(self._h, mode))

    def commit_info(self) -> dict:
        """What the last commit did (COMMIT_FULL / COMMIT_REFIT / COMMIT_TLAS_REBUILD) and the last
        BLAS build's depth and leaf counts."""
        ci = _CommitInfo()
        _check(lib().lrt_get_commit_info(self._h, ctypes.byref(ci)))
        return {k: getattr(ci, k) for k, _ in _CommitInfo._fields_}

    def set_leaf_postponement(self, mode: int = -1):
        """Triangle leaf postponement of the one-ray-per-lane pass: -1 automatic (on for one
        instance), 0 off, 1 on."""
        _check(lib().lrt_set_leaf_postponement(self._h, mode))

    def stack_overflow_count(self, reset: bool = False) -> int:
        v = ctypes.c_uint64()
        _check(lib().lrt_stack_overflow_count(self._h, ctypes.byref(v), int(reset)))
        return v.value

    def get_array(self, name: str) -> np.ndarray:
        n = ctypes.c_size_t()
        _check(lib().lrt_get_array(self._h, name.encode(), None, ctypes.byref(n)))
        dt = np.dtype(ARRAY_DTYPES[name])
        out = np.zeros(n.value // dt.itemsize, dtype=dt)
        if n.value:
            _check(lib().lrt_get_array(self._h, name.encode(), _ptr(out), ctypes.byref(n)))
        return out


class _BatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("geom_id", ctypes.c_uint32), ("params", ctypes.c_void_p),
                ("grad", ctypes.c_void_p), ("loss", ctypes.c_void_p), ("pos_near", ctypes.c_void_p),
                ("dir_far", ctypes.c_void_p), ("target", ctypes.c_void_p), ("colour", ctypes.c_void_p),
                ("n_rays", ctypes.c_uint64)]


class SbsBatch:
    """Batched differentiable launch over several brick-set scenes (lrt_sbs_batch_*). `items` are
    dicts with scene, geom_id, params, grad, loss, pos, dir, target, colour (torch tensors; the
    plan keeps references to them). run() zeroes grads and losses and runs forward + L2 adjoint of
    every ray in one persistent kernel; it only launches kernels, so it can be graph-captured."""

    def __init__(self, items):
        arr = (_BatchItem * max(1, len(items)))()
        self._keep = items
        for k, it in enumerate(items):
            arr[k] = _BatchItem(it["scene"]._h, it["geom_id"], _ptr(it["params"]), _ptr(it["grad"]),
                                _ptr(it.get("loss")), _ptr(it["pos"]), _ptr(it["dir"]), _ptr(it["target"]),
                                _ptr(it.get("colour")), it["pos"].numel() // 4)
        h = ctypes.c_void_p()
        _check(lib().lrt_sbs_batch_create(arr, len(items), ctypes.byref(h)))
        self._h = h

    def run(self, stream: int = 0):
        _check(lib().lrt_sbs_batch_run(self._h, stream))

    def autotune(self, stream: int = 0) -> dict:
        """Times the batch kernel's shapes on this plan (overwriting its outputs) and keeps the fastest."""
        r = _BatchTuneResult()
        _check(lib().lrt_sbs_batch_autotune(self._h, stream, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in _BatchTuneResult._fields_}

    def configure(self, lds_stack: int, record: int, grid: int):
        """Sets the kernel variant (LDS stack entries, hit-voxel record) and persistent grid."""
        _check(lib().lrt_sbs_batch_configure(self._h, lds_stack, record, grid))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lrt_sbs_batch_destroy(self._h)
            self._h = None


def CreateSceneRT(impl_name="BVH2Common", build_name="cbvh_embree2", layout_name="SuperTreeletAlignedMerged4"):
    """CreateSceneRT (BVH2Common_host.cpp:1594-1600): None for an unknown implementation name."""
    try:
        return SceneRT(impl_name, build_name, layout_name)
    except LrtError as e:
        if e.code == 4:   # LRT_ERR_UNSUPPORTED
            return None
        raise


# ---- builders and generators (include/lrt_build.h) ----
def camera_inverse(eye, target, up, fovy_deg, aspect, znear, zfar):
    """(inverse4x4(proj), inverse4x4(worldView)) as column-major float32[16]."""
    e, t, u = (np.asarray(v, dtype=np.float32) for v in (eye, target, up))
    pi = np.zeros(16, np.float32)
    vi = np.zeros(16, np.float32)
    _check(lib().lrt_camera_inverse(_ptr(e), _ptr(t), _ptr(u), fovy_deg, aspect, znear, zfar, _ptr(pi), _ptr(vi)))
    return pi, vi


def eye_rays(proj_inv, world_view_inv, width, height, pos, dir, jitter=None, order=1, stream=0):
    """Primary rays of MultiRenderer::initEyeRay into device tensors pos/dir (n,4) float32."""
    pi = np.ascontiguousarray(proj_inv, dtype=np.float32)
    vi = np.ascontiguousarray(world_view_inv, dtype=np.float32)
    _check(lib().lrt_eye_rays_dev(_ptr(pi), _ptr(vi), width, height, _ptr(jitter), order, _ptr(pos), _ptr(dir),
                                  stream))


def gen_icosphere(subdiv, amplitude=0.0, frequency=4.0, seed=1):
    vb, ib = _Buffer(), _Buffer()
    _check(lib().lrt_gen_icosphere(subdiv, amplitude, frequency, seed, ctypes.byref(vb), ctypes.byref(ib)))
    return _take(vb, np.float32).reshape(-1, 3), _take(ib, np.uint32).reshape(-1, 3)


def gen_geodesic(n, amplitude=0.0, frequency=4.0, seed=1):
    vb, ib = _Buffer(), _Buffer()
    _check(lib().lrt_gen_geodesic(n, amplitude, frequency, seed, ctypes.byref(vb), ctypes.byref(ib)))
    return _take(vb, np.float32).reshape(-1, 3), _take(ib, np.uint32).reshape(-1, 3)


GPU_BUILD = -1   # `threads` value of the SDF generators that builds the global octree on the GPU


def sdf_spheres_global_octree(spheres: np.ndarray, depth: int, brick_size: int = 1, brick_pad: int = 0,
                              threads: int = 0):
    """(nodes (n, 3) uint32 [type, val_off, offset], values float32) of sdf_to_global_octree
    (sdf/build/sparse_octree_builder.cpp:106-264) for a sphere union; threads < 0: on the GPU."""
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    nb, vb = _Buffer(), _Buffer()
    _check(lib().lrt_sdf_spheres_global_octree(_ptr(s), s.shape[0], depth, brick_size, brick_pad, threads,
                                               ctypes.byref(nb), ctypes.byref(vb)))
    return _take(nb, np.uint32).reshape(-1, 3), _take(vb, np.float32)


def sdf_spheres_frame_octree(spheres: np.ndarray, depth: int, threads: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    nb = _Buffer()
    _check(lib().lrt_sdf_spheres_frame_octree(_ptr(s), s.shape[0], depth, threads, ctypes.byref(nb)))
    return _take(nb, FRAME_NODE_DTYPE)


def sdf_spheres_multi_octree(spheres: np.ndarray, depth: int, threads: int = 0):
    """(nodes, values) of the multi-octree of a sphere union (brick size 1)."""
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    nb, vb = _Buffer(), _Buffer()
    _check(lib().lrt_sdf_spheres_multi_octree(_ptr(s), s.shape[0], depth, threads, ctypes.byref(nb), ctypes.byref(vb)))
    return _take(nb, MULTI_NODE_DTYPE), _take(vb, np.float32)


def sdf_spheres_sbs(spheres: np.ndarray, depth: int, header: np.ndarray, threads: int = 0, colour_seed: int = 0):
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    h = np.ascontiguousarray(header, dtype=SBS_HEADER_DTYPE).reshape(())
    nb, vb = _Buffer(), _Buffer()
    _check(lib().lrt_sdf_spheres_sbs(_ptr(s), s.shape[0], depth, _ptr(h), colour_seed, threads, ctypes.byref(nb),
                                     ctypes.byref(vb)))
    return _take(nb, SBS_NODE_DTYPE), _take(vb, np.uint32)


def sdf_spheres_sbs_indexed(spheres: np.ndarray, depth: int, brick_size: int, albedo_seed: int = 0,
                            threads: int = 0):
    """Indexed fp32 brick set: (header, nodes, values, values_f, n_distance_values)."""
    s = np.ascontiguousarray(spheres, dtype=np.float32).reshape(-1, 4)
    nb, vb, fb = _Buffer(), _Buffer(), _Buffer()
    nd = ctypes.c_uint32()
    _check(lib().lrt_sdf_spheres_sbs_indexed(_ptr(s), s.shape[0], depth, brick_size, albedo_seed, threads,
                                             ctypes.byref(nb), ctypes.byref(vb), ctypes.byref(fb), ctypes.byref(nd)))
    h = np.zeros((), SBS_HEADER_DTYPE)
    h["brick_size"], h["brick_pad"], h["bytes_per_value"] = brick_size, 0, 4
    h["aux_data"] = SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F
    return h, _take(nb, SBS_NODE_DTYPE), _take(vb, np.uint32), _take(fb, np.float32), nd.value


def sbs_with_neighbors(header, nodes: np.ndarray, values: np.ndarray, values_f: np.ndarray):
    """SBS_ind_to_SBS_ind_with_neighbors (sdf/simple/sdf_converter_exotic.cpp:529-587): an indexed
    brick set (ID32F_IRGB32F) converted to the layout with neighbour indices (ID32F_IRGB32F_IN):
    every brick keeps its v_size^3 distance indices and 8 colour indices and gains 27 words, word
    9 (dx + 1) + 3 (dy + 1) + (dz + 1) holding the index of the brick of the same size at offset
    (dx, dy, dz) in the brick grid, INVALID_IDX (1 << 31) where there is none and for (0, 0, 0).
    The reference pairs bricks through shared corner value indices; bricks of one size on the grid
    share a corner exactly when their positions differ by at most one step on every axis, which is
    the rule applied here. Returns (header, nodes, values, values_f)."""
    h = np.array(header, dtype=SBS_HEADER_DTYPE).copy()
    if int(h["aux_data"]) & 0xFF000000 != SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F:
        raise ValueError("needs the ID32F_IRGB32F layout")
    v = int(h["brick_size"]) + 2 * int(h["brick_pad"]) + 1
    old = v * v * v + 8
    nodes = np.ascontiguousarray(nodes, dtype=SBS_NODE_DTYPE)
    values = np.ascontiguousarray(values, dtype=np.uint32)
    n = nodes.size
    x = (nodes["pos_xy"] >> 16).astype(np.int64)
    y = (nodes["pos_xy"] & 0xFFFF).astype(np.int64)
    z = (nodes["pos_z_lod_size"] >> 16).astype(np.int64)
    lod = (nodes["pos_z_lod_size"] & 0xFFFF).astype(np.int64)
    key = ((lod * 65536 + x) * 65536 + y) * 65536 + z
    order = np.argsort(key)
    skey = key[order]
    nbr = np.full((n, 27), 1 << 31, np.uint32)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                q = ((lod * 65536 + x + dx) * 65536 + y + dy) * 65536 + z + dz
                ok = (x + dx >= 0) & (y + dy >= 0) & (z + dz >= 0) & (x + dx < lod) & (y + dy < lod) & (z + dz < lod)
                pos = np.clip(np.searchsorted(skey, q), 0, n - 1)
                hit = ok & (skey[pos] == q)
                nbr[hit, 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1)] = order[pos[hit]].astype(np.uint32)
    out = np.zeros(n * (old + 27), np.uint32)
    base = nodes["data_offset"].astype(np.int64)
    blocks = values[base[:, None] + np.arange(old)[None, :]]
    out.reshape(n, old + 27)[:, :old] = blocks
    out.reshape(n, old + 27)[:, old:] = nbr
    nn = nodes.copy()
    nn["data_offset"] = (np.arange(n) * (old + 27)).astype(np.uint32)
    h["aux_data"] = SDF_SBS_NODE_LAYOUT_ID32F_IRGB32F_IN
    return h, nn, out, np.ascontiguousarray(values_f, dtype=np.float32)


# ---- reference-format SDF scene files (sdf/simple/sdf_simple.cpp, sdf/multi/sdf_multi.cpp) ----
def load_sdf_frame_octree(path: str) -> np.ndarray:
    """load_sdf_frame_octree (sdf/simple/sdf_simple.cpp:58-66): SdfFrameOctreeNode array."""
    b = _Buffer()
    _check(lib().lrt_load_sdf_frame_octree(os.fsencode(path), ctypes.byref(b)))
    return _take(b, FRAME_NODE_DTYPE)


def save_sdf_frame_octree(path: str, nodes: np.ndarray):
    n = np.ascontiguousarray(nodes, dtype=FRAME_NODE_DTYPE)
    _check(lib().lrt_save_sdf_frame_octree(os.fsencode(path), _ptr(n), n.size))


def load_sdf_frame_octree_tex(path: str) -> np.ndarray:
    """load_sdf_frame_octree_tex (sdf/simple/sdf_simple.cpp:166-174): SdfFrameOctreeTexNode array."""
    b = _Buffer()
    _check(lib().lrt_load_sdf_frame_octree_tex(os.fsencode(path), ctypes.byref(b)))
    return _take(b, FRAME_TEX_NODE_DTYPE)


def save_sdf_frame_octree_tex(path: str, nodes: np.ndarray):
    n = np.ascontiguousarray(nodes, dtype=FRAME_TEX_NODE_DTYPE)
    _check(lib().lrt_save_sdf_frame_octree_tex(os.fsencode(path), _ptr(n), n.size))


def load_sdf_grid(path: str) -> np.ndarray:
    """load_sdf_grid (sdf/simple/sdf_simple.cpp:29-36): distances as a [z, y, x] array."""
    size = np.zeros(3, np.uint32)
    b = _Buffer()
    _check(lib().lrt_load_sdf_grid(os.fsencode(path), size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                   ctypes.byref(b)))
    return _take(b, np.float32).reshape(int(size[2]), int(size[1]), int(size[0]))


def save_sdf_grid(path: str, data: np.ndarray):
    d = np.ascontiguousarray(data, dtype=np.float32)
    size = np.array([d.shape[2], d.shape[1], d.shape[0]], np.uint32)
    _check(lib().lrt_save_sdf_grid(os.fsencode(path), size.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _ptr(d)))


def load_coctree_v3(path: str):
    """load_coctree_v3 (sdf/scom/scom.cpp:97-111): (header, words, version); a version other than
    5 is read all the same (the reference only warns)."""
    h = np.zeros((), COCTREE_HEADER_DTYPE)
    ver = ctypes.c_uint32()
    b = _Buffer()
    _check(lib().lrt_load_coctree_v3(os.fsencode(path), ctypes.byref(ver), _ptr(h), ctypes.byref(b)))
    return h, _take(b, np.uint32), ver.value


def save_coctree_v3(path: str, header: np.ndarray, data: np.ndarray):
    """save_coctree_v3 (sdf/scom/scom.cpp:86-95), version 5."""
    h = np.ascontiguousarray(header, dtype=COCTREE_HEADER_DTYPE).reshape(())
    d = np.ascontiguousarray(data, dtype=np.uint32).ravel()
    _check(lib().lrt_save_coctree_v3(os.fsencode(path), _ptr(h), _ptr(d), d.size))


def load_sdf_dag(path: str):
    """load_sdf_DAG (sdf/scom2/scom_utils.cpp:37-73): (header, nodes, child_edges, data_edges, distances)."""
    h = np.zeros((), DAG_HEADER_DTYPE)
    bufs = [_Buffer() for _ in range(4)]
    _check(lib().lrt_load_sdf_dag(os.fsencode(path), _ptr(h), *[ctypes.byref(b) for b in bufs]))
    return (h, _take(bufs[0], DAG_NODE_DTYPE), _take(bufs[1], DAG_CHILD_EDGE_DTYPE), _take(bufs[2], DAG_DATA_EDGE_DTYPE),
            _take(bufs[3], np.float32))


def save_sdf_dag(path: str, header, nodes, child_edges, data_edges, distances):
    """save_sdf_DAG (sdf/scom2/scom_utils.cpp:5-35), without data channels."""
    h = np.ascontiguousarray(header, dtype=DAG_HEADER_DTYPE).reshape(())
    n = np.ascontiguousarray(nodes, dtype=DAG_NODE_DTYPE)
    ce = np.ascontiguousarray(child_edges, dtype=DAG_CHILD_EDGE_DTYPE)
    de = np.ascontiguousarray(data_edges, dtype=DAG_DATA_EDGE_DTYPE)
    d = np.ascontiguousarray(distances, dtype=np.float32)
    _check(lib().lrt_save_sdf_dag(os.fsencode(path), _ptr(h), _ptr(n), n.size, _ptr(ce), ce.size, _ptr(de), de.size,
                                  _ptr(d), d.size))


def load_scom2(path: str):
    """load_scom2 (sdf/scom2/scom_utils.cpp:153-195): (header, node words, brick words)."""
    h = np.zeros((), SCOM2_HEADER_DTYPE)
    nb, bb = _Buffer(), _Buffer()
    _check(lib().lrt_load_scom2(os.fsencode(path), _ptr(h), ctypes.byref(nb), ctypes.byref(bb)))
    return h, _take(nb, np.uint32), _take(bb, np.uint32)


def save_scom2(path: str, header, nodes, bricks):
    """save_scom2 (sdf/scom2/scom_utils.cpp:98-127), version 4 without channels."""
    h = np.ascontiguousarray(header, dtype=SCOM2_HEADER_DTYPE).reshape(())
    n = np.ascontiguousarray(nodes, dtype=np.uint32).ravel()
    b = np.ascontiguousarray(bricks, dtype=np.uint32).ravel()
    _check(lib().lrt_save_scom2(os.fsencode(path), _ptr(h), _ptr(n), n.size, _ptr(b), b.size))


def load_sdf_svs(path: str) -> np.ndarray:
    """load_sdf_SVS (sdf/simple/sdf_simple.cpp:85-94): SdfSVSNode array."""
    b = _Buffer()
    _check(lib().lrt_load_sdf_svs(os.fsencode(path), ctypes.byref(b)))
    return _take(b, SVS_NODE_DTYPE)


def save_sdf_svs(path: str, nodes: np.ndarray):
    n = np.ascontiguousarray(nodes, dtype=SVS_NODE_DTYPE)
    _check(lib().lrt_save_sdf_svs(os.fsencode(path), _ptr(n), n.size))


def load_sdf_sbs(path: str):
    """load_sdf_SBS (sdf/simple/sdf_simple.cpp:124-142): (header, nodes, values uint32, values_f float32)."""
    h = np.zeros((), SBS_HEADER_DTYPE)
    bn, bv, bf = _Buffer(), _Buffer(), _Buffer()
    _check(lib().lrt_load_sdf_sbs(os.fsencode(path), _ptr(h), ctypes.byref(bn), ctypes.byref(bv), ctypes.byref(bf)))
    return h, _take(bn, SBS_NODE_DTYPE), _take(bv, np.uint32), _take(bf, np.float32)


def save_sdf_sbs(path: str, header, nodes, values, values_f=None):
    h = np.ascontiguousarray(header, dtype=SBS_HEADER_DTYPE).reshape(())
    n = np.ascontiguousarray(nodes, dtype=SBS_NODE_DTYPE)
    v = np.ascontiguousarray(values, dtype=np.uint32)
    f = np.ascontiguousarray(values_f if values_f is not None else np.zeros(0), dtype=np.float32)
    _check(lib().lrt_save_sdf_sbs(os.fsencode(path), _ptr(h), _ptr(n), n.size, _ptr(v), v.size, _ptr(f), f.size))


def load_sdf_multi_octree(path: str):
    """load_sdf_multi_octree (sdf/multi/sdf_multi.cpp:21-35): (nodes, values float32)."""
    bn, bv = _Buffer(), _Buffer()
    _check(lib().lrt_load_sdf_multi_octree(os.fsencode(path), ctypes.byref(bn), ctypes.byref(bv)))
    return _take(bn, MULTI_NODE_DTYPE), _take(bv, np.float32)


def save_sdf_multi_octree(path: str, nodes, values):
    n = np.ascontiguousarray(nodes, dtype=MULTI_NODE_DTYPE)
    v = np.ascontiguousarray(values, dtype=np.float32)
    _check(lib().lrt_save_sdf_multi_octree(os.fsencode(path), _ptr(n), n.size, _ptr(v), v.size))


def tonemap(colour, rgba8, sample_count: int = 1, stream: int = 0):
    """MultiRenderer::kernel1D_Tonemap on device buffers: colour float32 (n, 4) -> rgba8 uint32/int32 (n)."""
    n = colour.numel() // 4
    _check(lib().lrt_tonemap_dev(_ptr(colour), n, sample_count, _ptr(rgba8), stream))
