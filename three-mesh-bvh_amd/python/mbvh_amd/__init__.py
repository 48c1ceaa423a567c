"""ctypes binding of libmbvh.so (the C ABI declared in include/mbvh.h).

The JavaScript package is the drop-in host API; this module gives Python callers (the benchmark,
the tests) the same entry points over plain numpy arrays and raw device pointers. It never falls
back to a CPU path: if the library or a GPU is missing, calls raise MbvhError.

A process that also uses torch must import torch first: torch bundles its own HIP runtime under the
same soname as /opt/rocm's, so whichever loads first serves both, and torch cannot run on the other.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

CENTER, AVERAGE, SAH = 0, 1, 2
FRONT_SIDE, BACK_SIDE, DOUBLE_SIDE, SIDE_PER_TRIANGLE = 0, 1, 2, 3
GRAD_ATOMIC, GRAD_DETERMINISTIC = 0, 1
NO_HIT = 0xFFFFFFFF
NOALLOC_MAX_QUERIES = 1 << 21  # MBVH_NOALLOC_MAX_QUERIES

_PKG_ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libmbvh.so")

EXPORTED_SYMBOLS = (
    "mbvh_abi_version", "mbvh_last_error", "mbvh_device_count",
    "mbvh_build", "mbvh_tree_info", "mbvh_tree_root_bytes", "mbvh_tree_copy_root",
    "mbvh_tree_copy_indirect", "mbvh_tree_free", "mbvh_build_gpu", "mbvh_build_gpu_d",
    "mbvh_ctx_create", "mbvh_ctx_destroy", "mbvh_ctx_synchronize", "mbvh_ctx_set_ray_lanes",
    "mbvh_ctx_set_point_lanes",
    "mbvh_mesh_upload", "mbvh_mesh_free", "mbvh_mesh_info", "mbvh_mesh_set_side_mask", "mbvh_mesh_refit",
    "mbvh_mesh_refit_d", "mbvh_mesh_refit_nodes", "mbvh_mesh_refit_nodes_d", "mbvh_mesh_copy_roots", "mbvh_intersects_sphere",
    "mbvh_raycast_first", "mbvh_raycast_first_d",
    "mbvh_closest_point", "mbvh_closest_point_d", "mbvh_closest_point_noalloc_d",
    "mbvh_raycast_first_f64", "mbvh_raycast_first_ff_d", "mbvh_closest_point_f64", "mbvh_closest_point_ff_d",
    "mbvh_raycast_f64", "mbvh_batch_raycast_first_local_f64",
    "mbvh_raycast", "mbvh_raycast_fetch", "mbvh_raycast_count_d", "mbvh_raycast_fill_d",
    "mbvh_order_workspace_bytes", "mbvh_order_points_d",
    "mbvh_batch_create", "mbvh_batch_free", "mbvh_batch_workspace_bytes", "mbvh_batch_raycast_first_d",
    "mbvh_batch_raycast_first", "mbvh_scene_create", "mbvh_scene_free", "mbvh_scene_raycast_first",
    "mbvh_scene_raycast_first_d", "mbvh_intersects_box", "mbvh_intersects_box_d",
    "mbvh_closest_point_to_geometry", "mbvh_closest_point_to_geometry_d",
    "mbvh_scene_closest_point", "mbvh_scene_closest_point_d", "mbvh_scene_webgpu_sizes", "mbvh_scene_export_webgpu", "mbvh_mesh_texture_sizes", "mbvh_mesh_export_textures",
    "mbvh_mesh_export_textures_d", "mbvh_mesh_export_nodes", "mbvh_mesh_export_nodes_d",
    "mbvh_mesh_import_nodes", "mbvh_mesh_import_textures",
    "mbvh_batch_autotune", "mbvh_batch_graph_create", "mbvh_batch_graph_launch", "mbvh_batch_graph_free",
    "mbvh_instance_rays_d", "mbvh_instance_hits_d",
    "mbvh_grad_workspace_bytes", "mbvh_raycast_first_bwd", "mbvh_raycast_first_bwd_d",
    "mbvh_closest_point_bwd", "mbvh_closest_point_bwd_d",
)


class MbvhError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"mbvh error {code}: {message}")
        self.code = code


class BuildOpts(C.Structure):
    _fields_ = [("strategy", C.c_int32), ("max_depth", C.c_int32), ("target_leaf_size", C.c_int32),
                ("indirect", C.c_int32), ("strict_leaf_size", C.c_int32)]


_lib = None


class BatchTuning(C.Structure):
    """mbvh_batch_tuning: lanes per ray (1|4), LDS stack entries (8|16), rays per work item,
    workgroups per CU, measured ms."""
    _fields_ = [("ray_lanes", C.c_uint32), ("stack_lds", C.c_uint32), ("tile_rays", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("ms", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def library() -> C.CDLL:
    """Loads libmbvh.so once. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MbvhError(-100, f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp, u64, u32, i32, f32, f64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32, C.c_float, C.c_double
    sig = {
        "mbvh_abi_version": (u32, []),
        "mbvh_last_error": (C.c_char_p, []),
        "mbvh_device_count": (C.c_int, []),
        "mbvh_build": (C.c_int, [vp, u64, u32, u32, vp, i32, u64, vp, u32, C.POINTER(BuildOpts), C.POINTER(vp)]),
        "mbvh_tree_info": (C.c_int, [vp, C.POINTER(u32), C.POINTER(i32), C.POINTER(u64)]),
        "mbvh_tree_root_bytes": (C.c_int, [vp, u32, C.POINTER(u64)]),
        "mbvh_tree_copy_root": (C.c_int, [vp, u32, vp, u64]),
        "mbvh_tree_copy_indirect": (C.c_int, [vp, vp, u64]),
        "mbvh_tree_free": (None, [vp]),
        "mbvh_build_gpu": (C.c_int, [vp, vp, u64, u32, u32, vp, i32, u64, vp, u32, C.POINTER(BuildOpts), C.POINTER(vp)]),
        "mbvh_build_gpu_d": (C.c_int, [vp, vp, u64, u32, u32, vp, i32, u64, vp, u32, C.POINTER(BuildOpts),
                                       vp, u64, C.POINTER(u64), u32, vp, u64, vp]),
        "mbvh_ctx_create": (C.c_int, [i32, C.POINTER(vp)]),
        "mbvh_ctx_destroy": (None, [vp]),
        "mbvh_ctx_synchronize": (C.c_int, [vp]),
        "mbvh_ctx_set_ray_lanes": (C.c_int, [vp, i32]),
        "mbvh_ctx_set_point_lanes": (C.c_int, [vp, i32]),
        "mbvh_mesh_upload": (C.c_int, [vp, C.POINTER(vp), C.POINTER(u64), u32, vp, i32, u64, vp, i32, u64,
                                       vp, u64, u32, u32, C.POINTER(vp)]),
        "mbvh_mesh_free": (None, [vp]),
        "mbvh_mesh_info": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u32)]),
        "mbvh_mesh_set_side_mask": (C.c_int, [vp, vp, vp, u64]),
        "mbvh_mesh_refit": (C.c_int, [vp, vp, vp, u64, u32, u32, vp, vp, u32]),
        "mbvh_mesh_refit_d": (C.c_int, [vp, vp, vp, u64, u32, u32, vp]),
        "mbvh_mesh_refit_nodes": (C.c_int, [vp, vp, vp, u64, u32, u32, vp, u64, vp, vp, u32]),
        "mbvh_mesh_refit_nodes_d": (C.c_int, [vp, vp, vp, u64, u32, u32, vp, u64, vp]),
        "mbvh_intersects_sphere": (C.c_int, [vp, vp, vp, vp, u64, vp]),
        "mbvh_mesh_copy_roots": (C.c_int, [vp, vp, vp, vp, u32]),
        "mbvh_raycast_first": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp, vp]),
        "mbvh_raycast_first_d": (C.c_int, [vp, vp, vp, vp, u64, vp, i32, f32, f32, vp, vp, vp, vp, vp]),
        "mbvh_closest_point": (C.c_int, [vp, vp, vp, u64, f32, f32, vp, vp, vp, vp]),
        "mbvh_closest_point_d": (C.c_int, [vp, vp, vp, u64, vp, f32, f32, vp, vp, vp, vp, vp]),
        "mbvh_closest_point_noalloc_d": (C.c_int, [vp, vp, vp, u64, f32, f32, vp, vp, vp, vp, vp]),
        "mbvh_raycast_first_f64": (C.c_int, [vp, vp, vp, vp, u64, i32, f64, f64, vp, vp, vp, vp]),
        "mbvh_raycast_first_ff_d": (C.c_int, [vp, vp, vp, vp, vp, vp, u64, vp, i32, f64, f64, vp, vp, vp, vp, vp]),
        "mbvh_closest_point_f64": (C.c_int, [vp, vp, vp, u64, f64, f64, vp, vp, vp, vp]),
        "mbvh_closest_point_ff_d": (C.c_int, [vp, vp, vp, vp, u64, vp, f64, f64, vp, vp, vp, vp, vp]),
        "mbvh_raycast_f64": (C.c_int, [vp, vp, vp, vp, u64, i32, f64, f64, vp]),
        "mbvh_batch_raycast_first_local_f64": (C.c_int, [vp, vp, vp, vp, vp, vp, u64, i32, vp, vp, vp, vp]),
        "mbvh_order_workspace_bytes": (C.c_int, [u64, C.POINTER(u64)]),
        "mbvh_order_points_d": (C.c_int, [vp, vp, u64, vp, vp, u64, vp]),
        "mbvh_batch_create": (C.c_int, [vp, C.POINTER(vp), u32, C.POINTER(vp)]),
        "mbvh_batch_free": (None, [vp]),
        "mbvh_batch_workspace_bytes": (C.c_int, [vp, u64, C.POINTER(u64)]),
        "mbvh_batch_raycast_first_d": (C.c_int, [vp, vp, vp, vp, vp, u32, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp,
                                                 vp, vp, u64, C.POINTER(BatchTuning), vp]),
        "mbvh_batch_raycast_first": (C.c_int, [vp, vp, vp, vp, vp, u32, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp,
                                               vp]),
        "mbvh_batch_autotune": (C.c_int, [vp, vp, vp, vp, vp, u32, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp, vp,
                                          vp, u64, C.POINTER(BatchTuning)]),
        "mbvh_batch_graph_create": (C.c_int, [vp, vp, vp, vp, vp, u32, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp,
                                              vp, vp, u64, C.POINTER(BatchTuning), C.POINTER(vp)]),
        "mbvh_scene_create": (C.c_int, [vp, C.POINTER(vp), u32, vp, vp, vp, vp, u32, C.POINTER(vp)]),
        "mbvh_scene_free": (None, [vp]),
        "mbvh_scene_raycast_first_d": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp, vp, vp, vp]),
        "mbvh_scene_raycast_first": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp, vp, vp, vp, vp]),
        "mbvh_scene_closest_point": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp]),
        "mbvh_scene_closest_point_d": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp]),
        "mbvh_scene_webgpu_sizes": (C.c_int, [vp, vp, u32, vp, vp, vp]),
        "mbvh_scene_export_webgpu": (C.c_int, [vp, vp, u32, vp, u64, vp, u64, vp, u64]),
        "mbvh_intersects_box": (C.c_int, [vp, vp, vp, vp, vp, u64, vp]),
        "mbvh_intersects_box_d": (C.c_int, [vp, vp, vp, vp, vp, u64, vp, vp]),
        "mbvh_closest_point_to_geometry": (C.c_int, [vp, vp, vp, u64, u32, u32, vp, u32, u64, vp, f64, f64, vp]),
        "mbvh_closest_point_to_geometry_d": (C.c_int, [vp, vp, vp, u64, u32, u32, vp, u32, u64, vp, f64, f64, vp, vp]),
        "mbvh_mesh_texture_sizes": (C.c_int, [vp, vp]),
        "mbvh_mesh_export_textures": (C.c_int, [vp, vp, vp, u32, u32, vp, vp, vp, vp]),
        "mbvh_mesh_export_textures_d": (C.c_int, [vp, vp, vp, u32, u32, vp, vp, vp, vp, vp]),
        "mbvh_mesh_export_nodes": (C.c_int, [vp, vp, vp, u64]),
        "mbvh_mesh_import_nodes": (C.c_int, [vp, vp, u64, u32, vp, i32, u64, vp, i32, u64, vp, u64, u32, u32,
                                             C.POINTER(vp)]),
        "mbvh_mesh_import_textures": (C.c_int, [vp, vp, vp, u64, vp, i32, u64, vp, i32, u64, vp, u64, u32, u32,
                                                C.POINTER(vp)]),
        "mbvh_mesh_export_nodes_d": (C.c_int, [vp, vp, vp, u64, vp]),
        "mbvh_batch_graph_launch": (C.c_int, [vp, vp]),
        "mbvh_batch_graph_free": (None, [vp]),
        "mbvh_instance_rays_d": (C.c_int, [vp, vp, u32, vp, vp, vp, u64, vp, vp, vp]),
        "mbvh_instance_hits_d": (C.c_int, [vp, vp, u32, vp, vp, u64, vp, vp, vp, vp]),
        "mbvh_grad_workspace_bytes": (C.c_int, [u64, i32, C.POINTER(u64)]),
        "mbvh_raycast_first_bwd": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]),
        "mbvh_raycast_first_bwd_d": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, u64,
                                               vp]),
        "mbvh_closest_point_bwd": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, i32]),
        "mbvh_closest_point_bwd_d": (C.c_int, [vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp, i32, vp, u64, vp]),
        "mbvh_raycast": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp]),
        "mbvh_raycast_fetch": (C.c_int, [vp, vp, vp, vp, vp, u64]),
        "mbvh_raycast_count_d": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp, vp]),
        "mbvh_raycast_fill_d": (C.c_int, [vp, vp, vp, vp, u64, i32, f32, f32, vp, u64, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MbvhError(rc, library().mbvh_last_error().decode("utf-8", "replace"))


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    return int(library().mbvh_device_count())


@dataclass
class Tree:
    """Reference-format build result: MeshBVH._roots and (indirect builds) the indirect buffer."""
    roots: List[np.ndarray]          # uint8 node buffers, 32 bytes per node
    indirect: Optional[np.ndarray]   # uint16 / uint32 or None


def build(positions: np.ndarray, index: Optional[np.ndarray], ranges: Optional[Sequence] = None,
          strategy: int = SAH, max_depth: int = 40, target_leaf_size: int = 10,
          indirect: bool = False, stride: int = 3, offset: int = 0, ctx: Optional["Context"] = None,
          strict_leaf_size: int = 0) -> Tree:
    """Builds the BVH. A non-indirect build permutes `index` IN PLACE (like the reference).

    positions: float32 array with `stride` floats per vertex. ranges: (offset, count) triangle
    ranges of the roots (defaults to every triangle). With a Context the build runs on its device
    (mbvh_build_gpu) and returns the same bytes.
    """
    lib = library()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    nverts = pos.size // stride
    if index is not None:
        if index.dtype not in (np.uint16, np.uint32) or not index.flags.c_contiguous:
            raise ValueError("index must be a C-contiguous uint16 or uint32 array")
        ntris = index.size // 3
    else:
        ntris = nverts // 3
    if ranges is None:
        ranges = [(0, ntris)]
    rr = np.asarray(ranges, dtype=np.uint32).reshape(-1)
    opts = BuildOpts(strategy, max_depth, target_leaf_size, 1 if indirect else 0, strict_leaf_size)
    tree = C.c_void_p()
    args = (_ptr(pos), nverts, stride, offset, _ptr(index), 0 if index is None else index.itemsize,
            0 if index is None else index.size, _ptr(rr), len(rr) // 2, C.byref(opts), C.byref(tree))
    if ctx is None:
        _check(lib.mbvh_build(*args))
    else:
        _check(lib.mbvh_build_gpu(ctx.handle, *args))
    try:
        nroots, ieb, icount = C.c_uint32(), C.c_int32(), C.c_uint64()
        _check(lib.mbvh_tree_info(tree, C.byref(nroots), C.byref(ieb), C.byref(icount)))
        roots = []
        for r in range(nroots.value):
            nbytes = C.c_uint64()
            _check(lib.mbvh_tree_root_bytes(tree, r, C.byref(nbytes)))
            buf = np.empty(nbytes.value, dtype=np.uint8)
            _check(lib.mbvh_tree_copy_root(tree, r, _ptr(buf), nbytes.value))
            roots.append(buf)
        ind = None
        if ieb.value:
            ind = np.empty(icount.value, dtype=np.uint16 if ieb.value == 2 else np.uint32)
            _check(lib.mbvh_tree_copy_indirect(tree, _ptr(ind), ind.nbytes))
        return Tree(roots, ind)
    finally:
        lib.mbvh_tree_free(tree)


class Context:
    """One HIP device and stream (mbvh_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = library()
        h = C.c_void_p()
        _check(self._lib.mbvh_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self._lib.mbvh_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        _check(self._lib.mbvh_ctx_synchronize(self.handle))

    def set_point_lanes(self, lanes: int) -> None:
        """closestPoint lanes per query: 0 automatic, 1 (binary tree) or 4 (4-wide tree)."""
        _check(self._lib.mbvh_ctx_set_point_lanes(self.handle, lanes))

    def set_ray_lanes(self, lanes: int) -> None:
        """raycastFirst lanes per ray: 0 automatic, 1 (binary tree) or 4 (4-wide tree)."""
        _check(self._lib.mbvh_ctx_set_ray_lanes(self.handle, lanes))

    def build_d(self, positions_ptr: int, vertex_count: int, index_ptr: int, index_elt: int, index_count: int,
                ranges: Sequence, nodes_ptr: int, nodes_capacity: int, strategy: int = SAH, max_depth: int = 40,
                target_leaf_size: int = 10, indirect: bool = False, indirect_ptr: int = 0,
                indirect_capacity: int = 0, stride: int = 3, offset: int = 0, stream: int = 0) -> List[int]:
        """mbvh_build_gpu_d: device positions / index (permuted in place); returns the node count of
        each tree written one after another at nodes_ptr."""
        rr = np.asarray(ranges, dtype=np.uint32).reshape(-1)
        n_trees = 1 if indirect else len(rr) // 2
        counts = (C.c_uint64 * n_trees)()
        opts = BuildOpts(strategy, max_depth, target_leaf_size, 1 if indirect else 0)
        _check(self._lib.mbvh_build_gpu_d(self.handle, positions_ptr, vertex_count, stride, offset, index_ptr or None,
                                          index_elt, index_count, _ptr(rr), len(rr) // 2, C.byref(opts), nodes_ptr,
                                          nodes_capacity, counts, n_trees, indirect_ptr or None, indirect_capacity,
                                          stream or None))
        return [int(c) for c in counts]

    def scene(self, meshes: Sequence["Mesh"], matrix_world: np.ndarray, matrix_world_inverse: np.ndarray,
              instance_geometry: np.ndarray, visible: Optional[np.ndarray] = None) -> "Scene":
        return Scene(self, meshes, matrix_world, matrix_world_inverse, instance_geometry, visible)

    def upload(self, tree: Tree, positions: np.ndarray, index: Optional[np.ndarray],
               stride: int = 3, offset: int = 0) -> "Mesh":
        return Mesh(self, tree, positions, index, stride, offset)

    def batch(self, meshes: Sequence["Mesh"]) -> "Batch":
        return Batch(self, meshes)

    def import_nodes(self, nodes: np.ndarray, positions: np.ndarray, index: Optional[np.ndarray],
                     geometry_offset: int = 0, indirect: Optional[np.ndarray] = None) -> "Mesh":
        """A device mesh from a WebGPU BVHNode buffer (uint32 words, 8 per node, the mesh's roots back
        to back; leaf offsets carrying `geometry_offset`), as mbvh_mesh_import_nodes."""
        words = np.ascontiguousarray(nodes)
        words = (words.view(np.uint32) if words.dtype == np.uint8 else words.astype(np.uint32)).reshape(-1)
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        h = C.c_void_p()
        _check(self._lib.mbvh_mesh_import_nodes(
            self.handle, _ptr(words), words.size // 8, geometry_offset,
            _ptr(index), 0 if index is None else index.itemsize, 0 if index is None else index.size,
            _ptr(indirect), 0 if indirect is None else indirect.itemsize, 0 if indirect is None else indirect.size,
            _ptr(pos), pos.size // 3, 3, 0, C.byref(h)))
        # the roots' extents, as the import found them (each subtree's span)
        w = words.reshape(-1, 8)
        sizes, start = [], 0
        while start < w.shape[0]:
            size = _subtree_span(w, start)
            sizes.append(size * 32)
            start += size
        return Mesh._adopt(self, h, sizes)

    def import_textures(self, bvh_bounds: np.ndarray, bvh_contents: np.ndarray, n_nodes: int, positions: np.ndarray,
                        index: Optional[np.ndarray], indirect: Optional[np.ndarray] = None) -> "Mesh":
        """A device mesh from MeshBVHUniformStruct's bvhBounds (float32, 4 per texel) and bvhContents
        (uint32, 2 per texel) of one root of n_nodes nodes, as mbvh_mesh_import_textures."""
        b = np.ascontiguousarray(bvh_bounds, np.float32).reshape(-1)
        c = np.ascontiguousarray(bvh_contents, np.uint32).reshape(-1)
        if b.size < 8 * n_nodes or c.size < 2 * n_nodes:
            raise ValueError("textures smaller than n_nodes nodes")
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        h = C.c_void_p()
        _check(self._lib.mbvh_mesh_import_textures(
            self.handle, _ptr(b), _ptr(c), n_nodes,
            _ptr(index), 0 if index is None else index.itemsize, 0 if index is None else index.size,
            _ptr(indirect), 0 if indirect is None else indirect.itemsize, 0 if indirect is None else indirect.size,
            _ptr(pos), pos.size // 3, 3, 0, C.byref(h)))
        return Mesh._adopt(self, h, [n_nodes * 32])

    def instance_rays_d(self, inverse_ptr: int, n_instances: int, origins_ptr: int, dirs_ptr: int, ray_inst_ptr: int,
                        n: int, out_origins_ptr: int, out_dirs_ptr: int, stream: int = 0) -> None:
        v = lambda x: C.c_void_p(x or None)
        _check(self._lib.mbvh_instance_rays_d(self.handle, v(inverse_ptr), n_instances, v(origins_ptr), v(dirs_ptr),
                                              v(ray_inst_ptr), n, v(out_origins_ptr), v(out_dirs_ptr), v(stream)))

    def instance_hits_d(self, world_ptr: int, n_instances: int, origins_ptr: int, ray_inst_ptr: int, n: int,
                        face_ptr: int, point_ptr: int, dist_ptr: int, stream: int = 0) -> None:
        v = lambda x: C.c_void_p(x or None)
        _check(self._lib.mbvh_instance_hits_d(self.handle, v(world_ptr), n_instances, v(origins_ptr), v(ray_inst_ptr),
                                              n, v(face_ptr), v(point_ptr), v(dist_ptr), v(stream)))

    @staticmethod
    def order_workspace_bytes(n: int) -> int:
        out = C.c_uint64()
        _check(library().mbvh_order_workspace_bytes(n, C.byref(out)))
        return out.value

    def order_points_d(self, points_ptr: int, n: int, order_ptr: int, workspace_ptr: int, workspace_bytes: int,
                       stream: int = 0) -> None:
        """Spatial (Morton bucket) visit order of n device float triples into order_ptr (n uint32)."""
        _check(self._lib.mbvh_order_points_d(self.handle, C.c_void_p(points_ptr), n, C.c_void_p(order_ptr),
                                             C.c_void_p(workspace_ptr), workspace_bytes, C.c_void_p(stream or None)))


# mbvh_geometry_hit
GEOMETRY_HIT = np.dtype([("distance", "f8"), ("point", "f8", 3), ("other_point", "f8", 3), ("other_distance", "f8"),
                         ("face_index", "u4"), ("other_face_index", "u4"), ("status", "i4"), ("pad", "u4")])


def _geometry_hit(r):
    if not np.isfinite(r["distance"]):
        return None
    return {"distance": float(r["distance"]), "point": r["point"].copy(), "face_index": int(r["face_index"]),
            "other_point": r["other_point"].copy(), "other_distance": float(r["other_distance"]),
            "other_face_index": int(r["other_face_index"])}


def _subtree_span(words: np.ndarray, start: int) -> int:
    """Nodes in the depth-first subtree at `start` of (n, 8) node words: the right spine's offsets
    plus its last leaf."""
    j, size = start, 1
    while (int(words[j, 7]) & 0xFFFF0000) != 0xFFFF0000:
        r = int(words[j, 6])
        j += r
        size += r
    return size


class Mesh:
    """Device-resident BVH (mbvh_mesh)."""

    @classmethod
    def _adopt(cls, ctx: "Context", handle, root_bytes) -> "Mesh":
        m = cls.__new__(cls)
        m.ctx = ctx
        m.handle = handle
        m.root_bytes = list(root_bytes)
        return m

    def __init__(self, ctx: Context, tree: Tree, positions: np.ndarray, index: Optional[np.ndarray],
                 stride: int = 3, offset: int = 0):
        lib = ctx._lib
        self.ctx = ctx
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        nroots = len(tree.roots)
        root_ptrs = (C.c_void_p * nroots)(*[r.ctypes.data for r in tree.roots])
        root_bytes = (C.c_uint64 * nroots)(*[r.nbytes for r in tree.roots])
        h = C.c_void_p()
        ind = tree.indirect
        _check(lib.mbvh_mesh_upload(
            ctx.handle, root_ptrs, root_bytes, nroots,
            _ptr(index), 0 if index is None else index.itemsize, 0 if index is None else index.size,
            _ptr(ind), 0 if ind is None else ind.itemsize, 0 if ind is None else ind.size,
            _ptr(pos), pos.size // stride, stride, offset, C.byref(h)))
        self.handle = h
        self.root_bytes = [r.nbytes for r in tree.roots]

    def refit(self, positions: np.ndarray, stride: int = 3, offset: int = 0, roots: bool = True, node_ids=None):
        """MeshBVH.refit(nodeIndices) on the device from new positions of the same vertices; returns the
        refit reference-format root buffers (uint8 arrays) when `roots` is true. node_ids: None (a full
        refit) or node indices counted over all roots (ids outside [0, 2^32) name no node)."""
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        n = len(self.root_bytes)
        outs = [np.empty(b, np.uint8) for b in self.root_bytes] if roots else None
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs]) if roots else None
        sizes = (C.c_uint64 * n)(*self.root_bytes)
        if node_ids is None:
            _check(self.ctx._lib.mbvh_mesh_refit(self.ctx.handle, self.handle, _ptr(pos), pos.size // stride, stride,
                                                 offset, ptrs, sizes, n))
        else:
            ids = np.asarray(list(node_ids), dtype=np.int64).reshape(-1)
            ids = np.ascontiguousarray(ids[(ids >= 0) & (ids < 2**32)].astype(np.uint32))
            keep = np.zeros(1, np.uint32) if ids.size == 0 else ids  # a valid pointer for an empty set
            _check(self.ctx._lib.mbvh_mesh_refit_nodes(self.ctx.handle, self.handle, _ptr(pos), pos.size // stride,
                                                       stride, offset, _ptr(keep), ids.size, ptrs, sizes, n))
        return outs

    def refit_d(self, positions_ptr: int, vertex_count: int, stride: int = 3, offset: int = 0,
                stream: int = 0) -> None:
        """Refit from device positions, enqueued on `stream` (no synchronisation)."""
        _check(self.ctx._lib.mbvh_mesh_refit_d(self.ctx.handle, self.handle, C.c_void_p(positions_ptr), vertex_count,
                                               stride, offset, C.c_void_p(stream or None)))

    def refit_nodes_d(self, positions_ptr: int, vertex_count: int, node_ids_ptr: int, n_ids: int, stride: int = 3,
                      offset: int = 0, stream: int = 0) -> None:
        """Partial refit from device positions and device node ids (uint32), enqueued on `stream`."""
        _check(self.ctx._lib.mbvh_mesh_refit_nodes_d(self.ctx.handle, self.handle, C.c_void_p(positions_ptr),
                                                     vertex_count, stride, offset, C.c_void_p(node_ids_ptr), n_ids,
                                                     C.c_void_p(stream or None)))

    def intersects_box(self, box_min: np.ndarray, box_max: np.ndarray, box_to_mesh: np.ndarray) -> np.ndarray:
        """Per box: whether any triangle intersects it (min / max in the box frame, 4x4 column-major
        box -> mesh matrices), as the reference's OrientedBox tests decide it."""
        mn = np.ascontiguousarray(box_min, np.float64).reshape(-1, 3)
        mx = np.ascontiguousarray(box_max, np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(box_to_mesh, np.float64).reshape(-1, 16)
        if mn.shape != mx.shape or m.shape[0] != mn.shape[0]:
            raise ValueError("inconsistent box inputs")
        out = np.empty(mn.shape[0], np.uint8)
        _check(self.ctx._lib.mbvh_intersects_box(self.ctx.handle, self.handle, _ptr(mn), _ptr(mx), _ptr(m),
                                                 mn.shape[0], _ptr(out)))
        return out.astype(bool)

    def closest_point_to_geometry(self, positions: np.ndarray, index: Optional[np.ndarray] = None,
                                  geometry_to_bvh: Optional[np.ndarray] = None, min_threshold: float = 0.0,
                                  max_threshold: float = float("inf")):
        """closestPointToGeometry: the closest pair between this mesh's triangles and another geometry's
        (positions (V, 3) float32, optional uint16 / uint32 index), the other placed in this mesh's frame
        by geometry_to_bvh (4x4 column-major, 16 numbers). Returns None when no pair is closer than
        max_threshold, else a dict: distance, point (on this mesh), face_index (triangle slot),
        other_point / other_distance (in the other geometry's frame), other_face_index."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = None if index is None else np.ascontiguousarray(index)
        if idx is not None and idx.dtype not in (np.uint16, np.uint32):
            raise ValueError("index must be uint16 or uint32")
        n_tris = (idx.size if idx is not None else pos.shape[0]) // 3
        m = np.eye(4).reshape(-1) if geometry_to_bvh is None else np.asarray(geometry_to_bvh, np.float64).reshape(-1)
        if m.size != 16:
            raise ValueError("geometry_to_bvh must hold 16 numbers")
        m = np.ascontiguousarray(m, np.float64)
        out = np.zeros(1, GEOMETRY_HIT)
        _check(self.ctx._lib.mbvh_closest_point_to_geometry(
            self.ctx.handle, self.handle, _ptr(pos), pos.shape[0], 3, 0, _ptr(idx), 0 if idx is None else idx.itemsize,
            n_tris, _ptr(m), float(min_threshold), float(max_threshold), _ptr(out)))
        return _geometry_hit(out[0])

    def texture_sizes(self):
        """Texture sides (bounds, contents, index, position) of MeshBVHUniformStruct for this mesh."""
        sides = np.zeros(4, np.uint32)
        _check(self.ctx._lib.mbvh_mesh_texture_sizes(self.handle, _ptr(sides)))
        return [int(x) for x in sides]

    def export_textures(self, positions: Optional[np.ndarray] = None, stride: int = 3, offset: int = 0):
        """MeshBVHUniformStruct texture data: bounds (side^2, 4) float32, contents (side^2, 2) uint32,
        index (side^2, 4) uint32, position (side^2, 4) float32 (when positions are given)."""
        sb, sc, si, sp = self.texture_sizes()
        out = {"bounds": np.empty((sb * sb, 4), np.float32), "contents": np.empty((sc * sc, 2), np.uint32),
               "index": np.empty((si * si, 4), np.uint32),
               "position": None if positions is None else np.empty((sp * sp, 4), np.float32)}
        pos = None if positions is None else np.ascontiguousarray(positions, np.float32).reshape(-1)
        _check(self.ctx._lib.mbvh_mesh_export_textures(self.ctx.handle, self.handle, _ptr(pos), stride, offset,
                                                       _ptr(out["bounds"]), _ptr(out["contents"]), _ptr(out["index"]),
                                                       _ptr(out["position"])))
        return out

    def export_nodes(self) -> np.ndarray:
        """The WebGPU BVHNode buffer of the mesh (roots one after another)."""
        n = sum(self.root_bytes)
        buf = np.empty(int(n), np.uint8)
        _check(self.ctx._lib.mbvh_mesh_export_nodes(self.ctx.handle, self.handle, _ptr(buf), buf.nbytes))
        return buf

    def copy_roots(self):
        """The mesh's current reference-format root buffers (uint8 arrays)."""
        n = len(self.root_bytes)
        outs = [np.empty(b, np.uint8) for b in self.root_bytes]
        ptrs = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        sizes = (C.c_uint64 * n)(*self.root_bytes)
        _check(self.ctx._lib.mbvh_mesh_copy_roots(self.ctx.handle, self.handle, ptrs, sizes, n))
        return outs

    @property
    def vertex_count(self) -> int:
        return self.info()["vertex_count"]

    def info(self):
        t, v, d = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _check(self.ctx._lib.mbvh_mesh_info(self.handle, C.byref(t), C.byref(v), C.byref(d)))
        return {"tri_count": t.value, "vertex_count": v.value, "max_depth": d.value}

    def free(self) -> None:
        if self.handle:
            self.ctx._lib.mbvh_mesh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_side_mask(self, mask: np.ndarray) -> None:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(self.ctx._lib.mbvh_mesh_set_side_mask(self.ctx.handle, self.handle, _ptr(m), m.size))

    # ---- host-pointer queries (synchronous) ----
    def raycast_first(self, origins: np.ndarray, directions: np.ndarray, side: int = FRONT_SIDE,
                      near: float = 0.0, far: float = float("inf")):
        """float64 rays go to mbvh_raycast_first_f64 (traced as the double rays), others as float32."""
        f64 = np.asarray(origins).dtype == np.float64 or np.asarray(directions).dtype == np.float64
        dt = np.float64 if f64 else np.float32
        o = np.ascontiguousarray(origins, dtype=dt).reshape(-1, 3)
        d = np.ascontiguousarray(directions, dtype=dt).reshape(-1, 3)
        if o.shape != d.shape:
            raise ValueError("origins and directions must have the same shape")
        n = o.shape[0]
        face = np.empty(n, np.uint32)
        dist = np.empty(n, np.float32)
        point = np.empty((n, 3), np.float32)
        bary = np.empty((n, 3), np.float32)
        fn = self.ctx._lib.mbvh_raycast_first_f64 if f64 else self.ctx._lib.mbvh_raycast_first
        _check(fn(self.ctx.handle, self.handle, _ptr(o), _ptr(d), n, side, near, far, _ptr(face), _ptr(dist),
                  _ptr(point), _ptr(bary)))
        return {"face": face, "distance": dist, "point": point, "barycoord": bary}

    def raycast(self, origins: np.ndarray, directions: np.ndarray, side: int = FRONT_SIDE,
                near: float = 0.0, far: float = float("inf")):
        """All hits: returns offsets (n+1) and packed per-hit arrays in the reference's order."""
        f64 = np.asarray(origins).dtype == np.float64 or np.asarray(directions).dtype == np.float64
        dt = np.float64 if f64 else np.float32
        o = np.ascontiguousarray(origins, dtype=dt).reshape(-1, 3)
        d = np.ascontiguousarray(directions, dtype=dt).reshape(-1, 3)
        n = o.shape[0]
        offsets = np.zeros(n + 1, np.uint64)
        lib = self.ctx._lib
        fn = lib.mbvh_raycast_f64 if f64 else lib.mbvh_raycast
        _check(fn(self.ctx.handle, self.handle, _ptr(o), _ptr(d), n, side, near, far, _ptr(offsets)))
        total = int(offsets[-1])
        face = np.empty(total, np.uint32)
        dist = np.empty(total, np.float32)
        point = np.empty((total, 3), np.float32)
        bary = np.empty((total, 3), np.float32)
        _check(lib.mbvh_raycast_fetch(self.ctx.handle, _ptr(face), _ptr(dist), _ptr(point), _ptr(bary), total))
        return {"offsets": offsets, "face": face, "distance": dist, "point": point, "barycoord": bary}

    def closest_point(self, points: np.ndarray, min_threshold: float = 0.0,
                      max_threshold: float = float("inf")):
        f64 = np.asarray(points).dtype == np.float64
        q = np.ascontiguousarray(points, dtype=np.float64 if f64 else np.float32).reshape(-1, 3)
        n = q.shape[0]
        face = np.empty(n, np.uint32)
        dist = np.empty(n, np.float32)
        point = np.empty((n, 3), np.float32)
        bary = np.empty((n, 3), np.float32)
        fn = self.ctx._lib.mbvh_closest_point_f64 if f64 else self.ctx._lib.mbvh_closest_point
        _check(fn(self.ctx.handle, self.handle, _ptr(q), n, min_threshold, max_threshold, _ptr(face), _ptr(dist),
                  _ptr(point), _ptr(bary)))
        return {"face": face, "distance": dist, "point": point, "barycoord": bary}

    def intersects_sphere(self, centers: np.ndarray, radii: np.ndarray) -> np.ndarray:
        """intersectsSphere for n spheres: a bool per sphere."""
        c = np.ascontiguousarray(centers, dtype=np.float32).reshape(-1, 3)
        r = np.ascontiguousarray(radii, dtype=np.float32).reshape(-1)
        if r.shape[0] != c.shape[0]:
            raise ValueError("one radius per centre")
        out = np.empty(c.shape[0], np.uint8)
        _check(self.ctx._lib.mbvh_intersects_sphere(self.ctx.handle, self.handle, _ptr(c), _ptr(r), c.shape[0],
                                                    _ptr(out)))
        return out.astype(bool)

    # ---- device-pointer queries (asynchronous on `stream`) ----
    def raycast_first_d(self, origins_ptr: int, dirs_ptr: int, n: int, face_ptr: int, dist_ptr: int = 0,
                        point_ptr: int = 0, bary_ptr: int = 0, side: int = FRONT_SIDE, near: float = 0.0,
                        far: float = float("inf"), stream: int = 0, order_ptr: int = 0) -> None:
        _check(self.ctx._lib.mbvh_raycast_first_d(
            self.ctx.handle, self.handle, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), n,
            C.c_void_p(order_ptr or None), side, near, far,
            C.c_void_p(face_ptr), C.c_void_p(dist_ptr or None), C.c_void_p(point_ptr or None),
            C.c_void_p(bary_ptr or None), C.c_void_p(stream or None)))

    def raycast_count_d(self, origins_ptr: int, dirs_ptr: int, n: int, counts_ptr: int, side: int = FRONT_SIDE,
                        near: float = 0.0, far: float = float("inf"), stream: int = 0) -> None:
        _check(self.ctx._lib.mbvh_raycast_count_d(
            self.ctx.handle, self.handle, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), n, side, near, far,
            C.c_void_p(counts_ptr), C.c_void_p(stream or None)))

    def raycast_fill_d(self, origins_ptr: int, dirs_ptr: int, n: int, offsets_ptr: int, capacity: int,
                       face_ptr: int, dist_ptr: int = 0, point_ptr: int = 0, bary_ptr: int = 0,
                       side: int = FRONT_SIDE, near: float = 0.0, far: float = float("inf"), stream: int = 0) -> None:
        _check(self.ctx._lib.mbvh_raycast_fill_d(
            self.ctx.handle, self.handle, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), n, side, near, far,
            C.c_void_p(offsets_ptr), capacity, C.c_void_p(face_ptr), C.c_void_p(dist_ptr or None),
            C.c_void_p(point_ptr or None), C.c_void_p(bary_ptr or None), C.c_void_p(stream or None)))

    def closest_point_d(self, points_ptr: int, n: int, face_ptr: int, dist_ptr: int = 0, point_ptr: int = 0,
                        bary_ptr: int = 0, min_threshold: float = 0.0, max_threshold: float = float("inf"),
                        stream: int = 0, order_ptr: int = 0) -> None:
        _check(self.ctx._lib.mbvh_closest_point_d(
            self.ctx.handle, self.handle, C.c_void_p(points_ptr), n, C.c_void_p(order_ptr or None),
            min_threshold, max_threshold,
            C.c_void_p(face_ptr), C.c_void_p(dist_ptr or None), C.c_void_p(point_ptr or None),
            C.c_void_p(bary_ptr or None), C.c_void_p(stream or None)))

    def raycast_first_ff_d(self, origins_ptr: int, origins_lo_ptr: int, dirs_ptr: int, dirs_lo_ptr: int, n: int,
                           face_ptr: int, dist_ptr: int = 0, point_ptr: int = 0, bary_ptr: int = 0,
                           side: int = FRONT_SIDE, near: float = 0.0, far: float = float("inf"), stream: int = 0,
                           order_ptr: int = 0) -> None:
        """mbvh_raycast_first_ff_d: double rays split by the caller (hi, lo float arrays on the device)."""
        v = lambda p: C.c_void_p(p or None)
        _check(self.ctx._lib.mbvh_raycast_first_ff_d(
            self.ctx.handle, self.handle, v(origins_ptr), v(origins_lo_ptr), v(dirs_ptr), v(dirs_lo_ptr), n,
            v(order_ptr), side, near, far, v(face_ptr), v(dist_ptr), v(point_ptr), v(bary_ptr), v(stream)))

    def closest_point_ff_d(self, points_ptr: int, points_lo_ptr: int, n: int, face_ptr: int, dist_ptr: int = 0,
                           point_ptr: int = 0, bary_ptr: int = 0, min_threshold: float = 0.0,
                           max_threshold: float = float("inf"), stream: int = 0, order_ptr: int = 0) -> None:
        """mbvh_closest_point_ff_d: double points split by the caller (hi, lo float arrays on the device)."""
        v = lambda p: C.c_void_p(p or None)
        _check(self.ctx._lib.mbvh_closest_point_ff_d(
            self.ctx.handle, self.handle, v(points_ptr), v(points_lo_ptr), n, v(order_ptr), min_threshold,
            max_threshold, v(face_ptr), v(dist_ptr), v(point_ptr), v(bary_ptr), v(stream)))

    def closest_point_noalloc_d(self, points_ptr: int, n: int, face_ptr: int, dist_ptr: int = 0, point_ptr: int = 0,
                                bary_ptr: int = 0, min_threshold: float = 0.0, max_threshold: float = float("inf"),
                                stream: int = 0) -> None:
        """mbvh_closest_point_noalloc_d: as closest_point_d, with the library's fixed per-device scratch
        (no device allocation; at most NOALLOC_MAX_QUERIES points; overlapping calls on one GPU race)."""
        _check(self.ctx._lib.mbvh_closest_point_noalloc_d(
            self.ctx.handle, self.handle, C.c_void_p(points_ptr), n, min_threshold, max_threshold,
            C.c_void_p(face_ptr), C.c_void_p(dist_ptr or None), C.c_void_p(point_ptr or None),
            C.c_void_p(bary_ptr or None), C.c_void_p(stream or None)))

    # ---- backward (host arrays) ----
    def raycast_first_bwd(self, directions: np.ndarray, fwd: dict, grad_distance: Optional[np.ndarray] = None,
                          grad_point: Optional[np.ndarray] = None, grad_barycoord: Optional[np.ndarray] = None,
                          mode: int = GRAD_DETERMINISTIC, vertices: bool = True) -> dict:
        """Gradients of a loss w.r.t. ray origins, directions and vertex positions, for the fixed hits in
        `fwd` (the dict raycast_first returned) and the upstream gradients of its outputs."""
        d = np.ascontiguousarray(directions, np.float32)
        n = d.shape[0]
        g_o = np.empty((n, 3), np.float32)
        g_d = np.empty((n, 3), np.float32)
        g_v = np.empty((self.vertex_count, 3), np.float32) if vertices else None
        f32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        _check(self.ctx._lib.mbvh_raycast_first_bwd(
            self.ctx.handle, self.handle, _ptr(d), n, _ptr(np.ascontiguousarray(fwd["face"], np.uint32)),
            _ptr(f32(fwd["distance"])), _ptr(f32(fwd["barycoord"])), _ptr(f32(grad_distance)),
            _ptr(f32(grad_point)), _ptr(f32(grad_barycoord)), _ptr(g_o), _ptr(g_d), _ptr(g_v), mode))
        return {"origins": g_o, "directions": g_d, "positions": g_v}

    def closest_point_bwd(self, points: np.ndarray, fwd: dict, grad_distance: np.ndarray,
                          mode: int = GRAD_DETERMINISTIC, vertices: bool = True) -> dict:
        """Gradients of a loss w.r.t. query points and vertex positions for the fixed closest points in
        `fwd` (the dict closest_point returned) and the upstream gradient of the distances."""
        q = np.ascontiguousarray(points, np.float32)
        n = q.shape[0]
        g_q = np.empty((n, 3), np.float32)
        g_v = np.empty((self.vertex_count, 3), np.float32) if vertices else None
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        _check(self.ctx._lib.mbvh_closest_point_bwd(
            self.ctx.handle, self.handle, _ptr(q), n, _ptr(np.ascontiguousarray(fwd["face"], np.uint32)),
            _ptr(f32(fwd["distance"])), _ptr(f32(fwd["point"])), _ptr(f32(fwd["barycoord"])),
            _ptr(f32(grad_distance)), _ptr(g_q), _ptr(g_v), mode))
        return {"points": g_q, "positions": g_v}

    # ---- backward (device pointers, asynchronous on `stream`) ----
    @staticmethod
    def grad_workspace_bytes(n: int, mode: int = GRAD_DETERMINISTIC) -> int:
        out = C.c_uint64()
        _check(library().mbvh_grad_workspace_bytes(n, mode, C.byref(out)))
        return out.value

    def raycast_first_bwd_d(self, dirs_ptr: int, n: int, face_ptr: int, dist_ptr: int, bary_ptr: int,
                            g_dist_ptr: int = 0, g_point_ptr: int = 0, g_bary_ptr: int = 0, g_origins_ptr: int = 0,
                            g_dirs_ptr: int = 0, g_pos_ptr: int = 0, mode: int = GRAD_DETERMINISTIC,
                            ws_ptr: int = 0, ws_bytes: int = 0, stream: int = 0) -> None:
        v = lambda x: C.c_void_p(x or None)
        _check(self.ctx._lib.mbvh_raycast_first_bwd_d(
            self.ctx.handle, self.handle, v(dirs_ptr), n, v(face_ptr), v(dist_ptr), v(bary_ptr), v(g_dist_ptr),
            v(g_point_ptr), v(g_bary_ptr), v(g_origins_ptr), v(g_dirs_ptr), v(g_pos_ptr), mode, v(ws_ptr), ws_bytes,
            v(stream)))

    def closest_point_bwd_d(self, points_ptr: int, n: int, face_ptr: int, dist_ptr: int, closest_ptr: int,
                            bary_ptr: int, g_dist_ptr: int, g_points_ptr: int = 0, g_pos_ptr: int = 0,
                            mode: int = GRAD_DETERMINISTIC, ws_ptr: int = 0, ws_bytes: int = 0,
                            stream: int = 0) -> None:
        v = lambda x: C.c_void_p(x or None)
        _check(self.ctx._lib.mbvh_closest_point_bwd_d(
            self.ctx.handle, self.handle, v(points_ptr), n, v(face_ptr), v(dist_ptr), v(closest_ptr), v(bary_ptr),
            v(g_dist_ptr), v(g_points_ptr), v(g_pos_ptr), mode, v(ws_ptr), ws_bytes, v(stream)))


class Batch:
    """Several uploaded meshes for one persistent batched launch (mbvh_batch)."""

    def __init__(self, ctx: Context, meshes: Sequence[Mesh]):
        self.ctx = ctx
        self.meshes = list(meshes)  # keeps the meshes alive
        arr = (C.c_void_p * len(self.meshes))(*[m.handle for m in self.meshes])
        h = C.c_void_p()
        _check(ctx._lib.mbvh_batch_create(ctx.handle, arr, len(self.meshes), C.byref(h)))
        self.handle = h

    def free(self) -> None:
        if self.handle:
            self.ctx._lib.mbvh_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def workspace_bytes(self, n: int) -> int:
        out = C.c_uint64()
        _check(self.ctx._lib.mbvh_batch_workspace_bytes(self.handle, n, C.byref(out)))
        return out.value

    def _args(self, p):
        v = lambda x: C.c_void_p(x or None)
        return [self.ctx.handle, self.handle, v(p["world"]), v(p["inverse"]), v(p["instance_geometry"]),
                p["n_instances"], v(p["origins"]), v(p["directions"]), v(p["ray_instance"]), p["n"],
                p.get("side", FRONT_SIDE), p.get("near", 0.0), p.get("far", float("inf")), v(p["face"]),
                v(p.get("distance")), v(p.get("point")), v(p.get("barycoord")), v(p["workspace"]),
                p["workspace_bytes"]]

    def raycast_first_d(self, p: dict, tuning: Optional[BatchTuning] = None, stream: int = 0) -> None:
        """p: device pointers world, inverse, instance_geometry, origins, directions, ray_instance, face,
        distance, point, barycoord, workspace (+ n_instances, n, workspace_bytes, side, near, far)."""
        _check(self.ctx._lib.mbvh_batch_raycast_first_d(*self._args(p), C.byref(tuning) if tuning else None,
                                                        C.c_void_p(stream or None)))

    def raycast_first(self, matrix_world: np.ndarray, matrix_world_inverse: np.ndarray,
                      instance_geometry: np.ndarray, origins: np.ndarray, directions: np.ndarray,
                      ray_instance: np.ndarray, side: int = FRONT_SIDE, near: float = 0.0,
                      far: float = float("inf")):
        """Host arrays: column-major 4x4 matrices per instance, the batch slot of each instance's
        mesh, world rays and the instance of each ray. Returns face / distance / point / barycoord."""
        mw = np.ascontiguousarray(matrix_world, np.float32).reshape(-1, 16)
        mi = np.ascontiguousarray(matrix_world_inverse, np.float32).reshape(-1, 16)
        ig = np.ascontiguousarray(instance_geometry, np.uint32).reshape(-1)
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        ri = np.ascontiguousarray(ray_instance, np.uint32).reshape(-1)
        if mw.shape != mi.shape or mw.shape[0] != ig.shape[0] or o.shape != d.shape or o.shape[0] != ri.shape[0]:
            raise ValueError("inconsistent batched ray inputs")
        n = o.shape[0]
        out = {"face": np.empty(n, np.uint32), "distance": np.empty(n, np.float32),
               "point": np.empty((n, 3), np.float32), "barycoord": np.empty((n, 3), np.float32)}
        _check(self.ctx._lib.mbvh_batch_raycast_first(
            self.ctx.handle, self.handle, _ptr(mw), _ptr(mi), _ptr(ig), mw.shape[0], _ptr(o), _ptr(d), _ptr(ri), n,
            side, near, far, _ptr(out["face"]), _ptr(out["distance"]), _ptr(out["point"]), _ptr(out["barycoord"])))
        return out

    def raycast_first_local(self, origins: np.ndarray, directions: np.ndarray, ray_mesh: np.ndarray,
                            windows: Optional[np.ndarray] = None, side: int = FRONT_SIDE):
        """mbvh_batch_raycast_first_local_f64: double rays already in their mesh's frame, ray_mesh naming
        the batch mesh of each ray, windows (n, 2) near / far per ray (None: 0, inf); local results."""
        o = np.ascontiguousarray(origins, np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float64).reshape(-1, 3)
        rm = np.ascontiguousarray(ray_mesh, np.uint32).reshape(-1)
        if o.shape != d.shape or o.shape[0] != rm.shape[0]:
            raise ValueError("inconsistent local ray inputs")
        w = None if windows is None else np.ascontiguousarray(windows, np.float64).reshape(-1, 2)
        n = o.shape[0]
        out = {"face": np.empty(n, np.uint32), "distance": np.empty(n, np.float32),
               "point": np.empty((n, 3), np.float32), "barycoord": np.empty((n, 3), np.float32)}
        _check(self.ctx._lib.mbvh_batch_raycast_first_local_f64(
            self.ctx.handle, self.handle, _ptr(o), _ptr(d), _ptr(w), _ptr(rm), n, side, _ptr(out["face"]),
            _ptr(out["distance"]), _ptr(out["point"]), _ptr(out["barycoord"])))
        return out

    def autotune(self, p: dict) -> BatchTuning:
        t = BatchTuning()
        _check(self.ctx._lib.mbvh_batch_autotune(*self._args(p), C.byref(t)))
        return t

    def graph(self, p: dict, tuning: Optional[BatchTuning] = None) -> "BatchGraph":
        h = C.c_void_p()
        _check(self.ctx._lib.mbvh_batch_graph_create(*self._args(p), C.byref(tuning) if tuning else None, C.byref(h)))
        return BatchGraph(self, h)


class Scene:
    """Instances of uploaded meshes under a top-level tree (mbvh_scene): closest hit over all
    instances in one launch."""

    def __init__(self, ctx: Context, meshes: Sequence["Mesh"], matrix_world, matrix_world_inverse,
                 instance_geometry, visible=None):
        self.ctx = ctx
        self.meshes = list(meshes)  # the meshes must outlive the scene
        arr = (C.c_void_p * len(self.meshes))(*[m.handle for m in self.meshes])
        mw = np.ascontiguousarray(matrix_world, np.float32).reshape(-1, 16)
        mi = np.ascontiguousarray(matrix_world_inverse, np.float32).reshape(-1, 16)
        ig = np.ascontiguousarray(instance_geometry, np.uint32).reshape(-1)
        vis = None if visible is None else np.ascontiguousarray(visible, np.uint8).reshape(-1)
        if mw.shape != mi.shape or mw.shape[0] != ig.shape[0] or (vis is not None and vis.shape[0] != ig.shape[0]):
            raise ValueError("inconsistent instance inputs")
        h = C.c_void_p()
        _check(ctx._lib.mbvh_scene_create(ctx.handle, arr, len(self.meshes), _ptr(mw), _ptr(mi), _ptr(ig),
                                          _ptr(vis), ig.shape[0], C.byref(h)))
        self.handle = h

    def free(self) -> None:
        if self.handle:
            self.ctx._lib.mbvh_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def raycast_first(self, origins: np.ndarray, directions: np.ndarray, side: int = FRONT_SIDE,
                      near: float = 0.0, far: float = float("inf")):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        if o.shape != d.shape:
            raise ValueError("origins and directions differ in shape")
        n = o.shape[0]
        out = {"face": np.empty(n, np.uint32), "instance": np.empty(n, np.uint32),
               "distance": np.empty(n, np.float32), "point": np.empty((n, 3), np.float32),
               "barycoord": np.empty((n, 3), np.float32)}
        _check(self.ctx._lib.mbvh_scene_raycast_first(
            self.ctx.handle, self.handle, _ptr(o), _ptr(d), n, side, near, far, _ptr(out["face"]),
            _ptr(out["instance"]), _ptr(out["distance"]), _ptr(out["point"]), _ptr(out["barycoord"])))
        return out

    def closest_point(self, points: np.ndarray):
        """Closest point over the instances (world points in, world results out): face (in the
        instance's geometry), instance, distance, point, barycoord."""
        q = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = q.shape[0]
        out = {"face": np.empty(n, np.uint32), "instance": np.empty(n, np.uint32),
               "distance": np.empty(n, np.float32), "point": np.empty((n, 3), np.float32),
               "barycoord": np.empty((n, 3), np.float32)}
        _check(self.ctx._lib.mbvh_scene_closest_point(
            self.ctx.handle, self.handle, _ptr(q), n, _ptr(out["face"]), _ptr(out["instance"]),
            _ptr(out["distance"]), _ptr(out["point"]), _ptr(out["barycoord"])))
        return out

    def export_webgpu(self, primitive_limit: int = 0):
        """The scene in the reference's WebGPU packing: {'nodes': uint32 (N, 8), 'transforms': uint32 (S, 36),
        'index': uint32 (3T,)} as mbvh_scene_export_webgpu writes them; primitive_limit > 0 cuts the roots
        into clusters of fewer triangles (0: whole roots)."""
        nn, nt, ni = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self.ctx._lib.mbvh_scene_webgpu_sizes(self.ctx.handle, self.handle, primitive_limit, C.byref(nn),
                                                     C.byref(nt), C.byref(ni)))
        nodes = np.empty((nn.value, 8), np.uint32)
        tr = np.empty((nt.value, 36), np.uint32)
        idx = np.empty(max(ni.value, 1), np.uint32)
        _check(self.ctx._lib.mbvh_scene_export_webgpu(self.ctx.handle, self.handle, primitive_limit, _ptr(nodes),
                                                      nodes.size, _ptr(tr), tr.size, _ptr(idx), ni.value))
        return {"nodes": nodes, "transforms": tr, "index": idx[:ni.value]}

    def closest_point_d(self, points_ptr: int, n: int, face_ptr: int, instance_ptr: int = 0, dist_ptr: int = 0,
                        point_ptr: int = 0, bary_ptr: int = 0, stream: int = 0, order_ptr: int = 0) -> None:
        """Device form; order_ptr: an optional visit order (n uint32, e.g. Context.order_points_d)."""
        v = lambda x: C.c_void_p(x or None)
        _check(self.ctx._lib.mbvh_scene_closest_point_d(
            self.ctx.handle, self.handle, v(points_ptr), n, v(order_ptr), v(face_ptr), v(instance_ptr), v(dist_ptr),
            v(point_ptr), v(bary_ptr), v(stream)))

    def raycast_first_d(self, origins_ptr: int, dirs_ptr: int, n: int, face_ptr: int, instance_ptr: int = 0,
                        dist_ptr: int = 0, point_ptr: int = 0, bary_ptr: int = 0, side: int = FRONT_SIDE,
                        near: float = 0.0, far: float = float("inf"), stream: int = 0) -> None:
        v = lambda x: C.c_void_p(x or None)
        _check(self.ctx._lib.mbvh_scene_raycast_first_d(
            self.ctx.handle, self.handle, v(origins_ptr), v(dirs_ptr), n, side, near, far, v(face_ptr),
            v(instance_ptr), v(dist_ptr), v(point_ptr), v(bary_ptr), v(stream)))


class BatchGraph:
    """A captured batched launch (mbvh_batch_graph)."""

    def __init__(self, batch: Batch, handle):
        self.batch = batch
        self.handle = handle

    def launch(self, stream: int = 0) -> None:
        _check(self.batch.ctx._lib.mbvh_batch_graph_launch(self.handle, C.c_void_p(stream or None)))

    def free(self) -> None:
        if self.handle:
            self.batch.ctx._lib.mbvh_batch_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
