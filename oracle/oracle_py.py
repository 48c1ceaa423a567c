"""ctypes wrapper of liboracle.so — the FP64 restatement of the reference query path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, where it is the checker (or the timed CPU baseline). The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
NO_HIT = 0xFFFFFFFF
FRONT, BACK, DOUBLE, PER_FACE = 0, 1, 2, 3


class _Mesh(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("vertex_count", C.c_uint64), ("stride", C.c_uint32), ("offset", C.c_uint32),
        ("index", C.c_void_p), ("index_elt_bytes", C.c_int32), ("index_count", C.c_uint64),
        ("indirect", C.c_void_p), ("indirect_elt_bytes", C.c_int32), ("indirect_count", C.c_uint64),
        ("roots", C.c_void_p), ("root_bytes", C.c_void_p), ("n_roots", C.c_uint32),
        ("face_mask", C.c_void_p),
    ]


class TreeStats(C.Structure):
    _fields_ = [("node_count", C.c_uint64), ("leaf_count", C.c_uint64), ("tri_count", C.c_uint64),
                ("depth_min", C.c_uint32), ("depth_max", C.c_uint32),
                ("leaf_tris_min", C.c_uint32), ("leaf_tris_max", C.c_uint32),
                ("bounds_failures", C.c_uint64), ("structure_failures", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        vp, u64, i32, f64, u32 = C.c_void_p, C.c_uint64, C.c_int32, C.c_double, C.c_uint32
        lib.orc_ray_triangle.argtypes = [vp, vp, vp, vp, vp, C.c_int, vp, vp]
        lib.orc_ray_triangle.restype = C.c_int
        lib.orc_closest_point_triangle.argtypes = [vp, vp, vp, vp, vp]
        lib.orc_closest_point_triangle.restype = None
        lib.orc_barycoord.argtypes = [vp, vp, vp, vp, vp]
        lib.orc_barycoord.restype = C.c_int
        lib.orc_raycast_first.argtypes = [vp, vp, vp, u64, C.c_int, f64, f64, C.c_int, C.c_int,
                                          vp, vp, vp, vp, vp, vp]
        lib.orc_raycast_first.restype = C.c_int
        lib.orc_raycast_all.argtypes = [vp, vp, vp, C.c_int, f64, f64, u64, vp, vp]
        lib.orc_raycast_all.restype = C.c_int64
        lib.orc_closest_point.argtypes = [vp, vp, u64, f64, f64, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        lib.orc_closest_point.restype = C.c_int
        lib.orc_ray_plane_distance.argtypes = [vp, u32, vp, vp]
        lib.orc_ray_plane_distance.restype = f64
        lib.orc_point_face_distance.argtypes = [vp, u32, vp]
        lib.orc_point_face_distance.restype = f64
        lib.orc_raycast_first_f64.argtypes = lib.orc_raycast_first.argtypes
        lib.orc_raycast_first_f64.restype = C.c_int
        lib.orc_closest_point_f64.argtypes = lib.orc_closest_point.argtypes
        lib.orc_closest_point_f64.restype = C.c_int
        lib.orc_raycast_all_f64.argtypes = lib.orc_raycast_all.argtypes
        lib.orc_raycast_all_f64.restype = C.c_int64
        lib.orc_ray_plane_distance_f64.argtypes = [vp, u32, vp, vp]
        lib.orc_ray_plane_distance_f64.restype = f64
        lib.orc_point_face_distance_f64.argtypes = [vp, u32, vp]
        lib.orc_point_face_distance_f64.restype = f64
        lib.orc_intersects_sphere.argtypes = [vp, vp, f64]
        lib.orc_intersects_sphere.restype = C.c_int
        lib.orc_intersects_box.argtypes = [vp, vp, vp, vp]
        lib.orc_intersects_box.restype = C.c_int
        lib.orc_box_triangle_sat.argtypes = [vp, vp, vp, vp, vp, vp]
        lib.orc_box_triangle_sat.restype = C.c_int
        lib.orc_box_triangle_clip.argtypes = [vp, vp, vp, vp, vp, vp]
        lib.orc_box_triangle_clip.restype = C.c_int
        lib.orc_scene_closest_point.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp, vp]
        lib.orc_scene_closest_point.restype = u32
        lib.orc_segment_segment.argtypes = [vp, vp, vp, vp, vp, vp]
        lib.orc_segment_segment.restype = None
        lib.orc_triangle_distance.argtypes = [vp, vp, vp, vp]
        lib.orc_triangle_distance.restype = f64
        lib.orc_closest_to_geometry.argtypes = [vp, vp, u64, u32, u32, vp, C.c_int, u64, vp, f64, C.c_int, vp, vp,
                                                vp, vp]
        lib.orc_closest_to_geometry.restype = u32
        lib.orc_refit_root.argtypes = [vp, vp, C.c_uint64]
        lib.orc_refit_root.restype = None
        lib.orc_refit_root_nodes.argtypes = [vp, vp, C.c_uint64, C.c_uint64, vp, C.c_uint64]
        lib.orc_refit_root_nodes.restype = None
        lib.orc_validate_root.argtypes = [vp, u32, C.POINTER(TreeStats)]
        lib.orc_validate_root.restype = C.c_int
        _lib = lib
    return _lib


def _is_f64(*arrays):
    return any(np.asarray(a).dtype == np.float64 for a in arrays)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _v3(x):
    return np.ascontiguousarray(x, dtype=np.float64).reshape(3)


def ray_triangle(o, d, a, b, c, backface_culling=True):
    t = np.zeros(1)
    p = np.zeros(3)
    arrs = [_v3(o), _v3(d), _v3(a), _v3(b), _v3(c)]
    ok = library().orc_ray_triangle(*[_p(x) for x in arrs], int(backface_culling), _p(t), _p(p))
    return (float(t[0]), p) if ok else None


def closest_point_triangle(p, a, b, c):
    out = np.zeros(3)
    arrs = [_v3(p), _v3(a), _v3(b), _v3(c)]
    library().orc_closest_point_triangle(*[_p(x) for x in arrs], _p(out))
    return out


def barycoord(p, a, b, c):
    out = np.zeros(3)
    arrs = [_v3(p), _v3(a), _v3(b), _v3(c)]
    ok = library().orc_barycoord(*[_p(x) for x in arrs], _p(out))
    return out if ok else None


class OracleMesh:
    """Holds the arrays of one reference-format BVH for the oracle."""

    def __init__(self, positions: np.ndarray, index: Optional[np.ndarray], roots: Sequence[np.ndarray],
                 indirect: Optional[np.ndarray] = None, face_mask: Optional[np.ndarray] = None,
                 stride: int = 3, offset: int = 0):
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        self.index = None if index is None else np.ascontiguousarray(index)
        self.indirect = None if indirect is None else np.ascontiguousarray(indirect)
        self.roots = [np.ascontiguousarray(r, dtype=np.uint8) for r in roots]
        self.face_mask = None if face_mask is None else np.ascontiguousarray(face_mask, dtype=np.uint8)
        self._root_ptrs = (C.c_void_p * len(self.roots))(*[r.ctypes.data for r in self.roots])
        self._root_bytes = (C.c_uint64 * len(self.roots))(*[r.nbytes for r in self.roots])
        m = _Mesh()
        m.positions = self.positions.ctypes.data
        m.vertex_count = self.positions.size // stride
        m.stride, m.offset = stride, offset
        if self.index is not None:
            m.index, m.index_elt_bytes, m.index_count = self.index.ctypes.data, self.index.itemsize, self.index.size
        if self.indirect is not None:
            m.indirect, m.indirect_elt_bytes = self.indirect.ctypes.data, self.indirect.itemsize
            m.indirect_count = self.indirect.size
        m.roots = C.cast(self._root_ptrs, C.c_void_p)
        m.root_bytes = C.cast(self._root_bytes, C.c_void_p)
        m.n_roots = len(self.roots)
        if self.face_mask is not None:
            m.face_mask = self.face_mask.ctypes.data
        self._m = m

    @property
    def ref(self):
        return C.byref(self._m)

    def raycast_first(self, origins, dirs, side=FRONT, near=0.0, far=float("inf"), brute=False,
                      threads: Optional[int] = None):
        """float64 rays are traced as given; other rays as float32."""
        f64 = _is_f64(origins, dirs)
        dt = np.float64 if f64 else np.float32
        o = np.ascontiguousarray(origins, dtype=dt).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=dt).reshape(-1, 3)
        n = o.shape[0]
        out = {"face": np.empty(n, np.uint32), "distance": np.empty(n), "point": np.empty((n, 3)),
               "barycoord": np.empty((n, 3)), "nodes": np.empty(n, np.uint32), "tris": np.empty(n, np.uint32)}
        fn = library().orc_raycast_first_f64 if f64 else library().orc_raycast_first
        fn(self.ref, _p(o), _p(d), n, side, near, far, int(brute),
                                    threads or os.cpu_count() or 1, _p(out["face"]), _p(out["distance"]),
                                    _p(out["point"]), _p(out["barycoord"]), _p(out["nodes"]), _p(out["tris"]))
        return out

    def raycast_all(self, origin, direction, side=FRONT, near=0.0, far=float("inf"), cap=4096):
        f64 = _is_f64(origin, direction)
        dt = np.float64 if f64 else np.float32
        o = np.ascontiguousarray(origin, dtype=dt).reshape(3)
        d = np.ascontiguousarray(direction, dtype=dt).reshape(3)
        face = np.empty(cap, np.uint32)
        dist = np.empty(cap)
        fn = library().orc_raycast_all_f64 if f64 else library().orc_raycast_all
        n = fn(self.ref, _p(o), _p(d), side, near, far, cap, _p(face), _p(dist))
        n = min(int(n), cap)
        return face[:n], dist[:n]

    def closest_point(self, points, min_threshold=0.0, max_threshold=float("inf"), brute=False,
                      threads: Optional[int] = None):
        f64 = _is_f64(points)
        q = np.ascontiguousarray(points, dtype=np.float64 if f64 else np.float32).reshape(-1, 3)
        n = q.shape[0]
        out = {"face": np.empty(n, np.uint32), "distance": np.empty(n), "point": np.empty((n, 3)),
               "nodes": np.empty(n, np.uint32), "tris": np.empty(n, np.uint32)}
        fn = library().orc_closest_point_f64 if f64 else library().orc_closest_point
        fn(self.ref, _p(q), n, min_threshold, max_threshold, int(brute),
                                    threads or os.cpu_count() or 1, _p(out["face"]), _p(out["distance"]),
                                    _p(out["point"]), _p(out["nodes"]), _p(out["tris"]))
        return out

    def ray_plane_distance(self, face, origin, direction):
        if _is_f64(origin, direction):
            o = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
            d = np.ascontiguousarray(direction, dtype=np.float64).reshape(3)
            return library().orc_ray_plane_distance_f64(self.ref, int(face), _p(o), _p(d))
        o = np.ascontiguousarray(origin, dtype=np.float32).reshape(3)
        d = np.ascontiguousarray(direction, dtype=np.float32).reshape(3)
        return library().orc_ray_plane_distance(self.ref, int(face), _p(o), _p(d))

    def point_face_distance(self, face, p):
        if _is_f64(p):
            return library().orc_point_face_distance_f64(self.ref, int(face),
                                                         _p(np.ascontiguousarray(p, dtype=np.float64).reshape(3)))
        q = np.ascontiguousarray(p, dtype=np.float32).reshape(3)
        return library().orc_point_face_distance(self.ref, int(face), _p(q))

    def intersects_sphere(self, centers, radii):
        """intersectsSphere per sphere (brute force, the reference's sphere-triangle test)."""
        c = np.ascontiguousarray(centers, dtype=np.float32).reshape(-1, 3)
        r = np.asarray(radii, dtype=np.float64).reshape(-1)
        return np.array([bool(library().orc_intersects_sphere(self.ref, _p(c[i]), float(r[i]))) for i in range(len(c))])

    def closest_to_geometry(self, positions, index=None, matrix=None, max_threshold=float("inf"), brute=False):
        """closestPointToGeometry: the closest pair over every mesh triangle and every triangle of the
        other geometry placed by `matrix` (16 numbers, column-major). Equal distances go to the lowest
        other triangle, then the lowest mesh slot. None when no pair is below max_threshold."""
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = None if index is None else np.ascontiguousarray(index)
        n = (idx.size if idx is not None else pos.shape[0]) // 3
        m = np.ascontiguousarray(np.eye(4).reshape(-1) if matrix is None else np.asarray(matrix, np.float64).reshape(-1),
                                 np.float64)
        slot = C.c_uint32()
        dist = C.c_double()
        p1 = np.zeros(3)
        p2 = np.zeros(3)
        j = library().orc_closest_to_geometry(self.ref, _p(pos), pos.shape[0], 3, 0, _p(idx),
                                              0 if idx is None else idx.itemsize, n, _p(m), float(max_threshold),
                                              int(brute), C.byref(slot), C.byref(dist), _p(p1), _p(p2))
        if j == NO_HIT:
            return None
        return {"distance": dist.value, "face_index": slot.value, "other_face_index": int(j), "point": p1,
                "point_in_mesh_frame": p2}

    def intersects_box(self, box_min, box_max, box_to_mesh):
        """intersectsBox per box (brute force, the reference's OrientedBox separating-axis test)."""
        mn = np.ascontiguousarray(box_min, np.float64).reshape(-1, 3)
        mx = np.ascontiguousarray(box_max, np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(box_to_mesh, np.float64).reshape(-1, 16)
        return np.array([bool(library().orc_intersects_box(self.ref, _p(mn[i]), _p(mx[i]), _p(m[i])))
                         for i in range(mn.shape[0])])

    def refit(self, node_ids=None):
        """The roots after MeshBVH.refit(nodeIndices) over this mesh's positions (new arrays; self.roots
        unchanged). node_ids: None (a full refit) or the node indices over all roots (refit.template.js
        numbers root r's nodes from the node count of the roots before it)."""
        total = sum(r.nbytes // 32 for r in self.roots)
        member = None
        if node_ids is not None:
            ids = np.asarray(list(node_ids), dtype=np.int64).reshape(-1)
            member = np.zeros(max(total, 1), np.uint8)
            member[ids[(ids >= 0) & (ids < total)]] = 1
        out = []
        base = 0
        for r in self.roots:
            nodes = r.copy()
            if member is None:
                library().orc_refit_root(self.ref, nodes.ctypes.data, nodes.nbytes)
            else:
                library().orc_refit_root_nodes(self.ref, nodes.ctypes.data, nodes.nbytes, base, member.ctypes.data,
                                               member.size)
            out.append(nodes)
            base += r.nbytes // 32
        return out

    def validate_root(self, root: int = 0):
        st = TreeStats()
        rc = library().orc_validate_root(self.ref, root, C.byref(st))
        return rc, st.as_dict()


# ---- shader data layouts (numpy restatements of the reference's packing) ------------------------------

def _ceil_sqrt(x: float) -> int:
    return int(np.ceil(np.sqrt(x)))


def uniform_textures(root: np.ndarray):
    """bvhToTextures (src/webgl/MeshBVHUniformStruct.js:129-210) for one root buffer: (bounds float32
    (side^2, 4), contents uint32 (side^2, 2)). Bounds side = 2 * ceil(sqrt(nodes / 2)): texel 2i holds
    the node's min and texel 2i + 1 its max (w = 0); contents side = ceil(sqrt(nodes)): a leaf is
    (LEAFNODE_MASK_32 | count, offset), an internal node (split axis, right child offset)."""
    buf = np.ascontiguousarray(root, np.uint8)
    n = buf.size // 32
    f = buf.view(np.float32).reshape(n, 8)
    u = buf.view(np.uint32).reshape(n, 8)
    h = buf.view(np.uint16).reshape(n, 16)
    bside = 2 * _ceil_sqrt(n / 2)
    cside = _ceil_sqrt(n)
    bounds = np.zeros((bside * bside, 4), np.float32)
    contents = np.zeros((cside * cside, 2), np.uint32)
    for i in range(n):
        bounds[2 * i, :3] = f[i, 0:3]
        bounds[2 * i + 1, :3] = f[i, 3:6]
        if h[i, 15] == 0xFFFF:  # IS_LEAF
            contents[i] = (0xFFFF0000 | int(h[i, 14]), u[i, 6])
        else:
            contents[i] = (u[i, 7], u[i, 6])
    return bounds, contents


def textures_to_root(bounds: np.ndarray, contents: np.ndarray, n_nodes: int) -> np.ndarray:
    """The inverse of uniform_textures: the 32-byte node buffer the two textures were made from."""
    out = np.zeros((n_nodes, 8), np.uint32)
    b = np.asarray(bounds, np.float32).reshape(-1, 4)
    c = np.asarray(contents, np.uint32).reshape(-1, 2)
    for i in range(n_nodes):
        out[i, 0:3] = b[2 * i, :3].view(np.uint32)
        out[i, 3:6] = b[2 * i + 1, :3].view(np.uint32)
        if (c[i, 0] & 0xFFFF0000) == 0xFFFF0000:
            out[i, 6] = c[i, 1]
            out[i, 7] = 0xFFFF0000 | (c[i, 0] & 0xFFFF)
        else:
            out[i, 6] = c[i, 1]
            out[i, 7] = c[i, 0]
    return out.view(np.uint8).reshape(-1)


def attribute_texture(values: np.ndarray, item_size: int) -> np.ndarray:
    """VertexAttributeTexture.updateFrom (src/webgl/VertexAttributeTexture.js:106-300) for a
    non-normalized attribute: side = ceil(sqrt(count)) or 1, item size 3 stored as RGBA with w = 1,
    the rest zero. Returns (side^2, 4 or item_size) in the attribute's element type."""
    v = np.asarray(values).reshape(-1, item_size)
    count = v.shape[0]
    side = _ceil_sqrt(count) or 1
    stride = 4 if item_size == 3 else item_size
    out = np.zeros((side * side, stride), v.dtype)
    out[:count, :item_size] = v
    if item_size == 3:
        out[:count, 3] = 1
    return out


def webgpu_nodes(roots: Sequence[np.ndarray], geometry_offset: int = 0) -> np.ndarray:
    """appendBVHSubtree (src/webgpu/utils/packBVHBufferUtils.js) of every root of a mesh, one after
    another: bounds copied (an empty root's min > max axis written as [1, -1]), leaf offsets rebased by
    geometry_offset, internal words copied."""
    parts = []
    for root in roots:
        buf = np.ascontiguousarray(root, np.uint8)
        n = buf.size // 32
        u = buf.view(np.uint32).reshape(n, 8).copy()
        f = u.view(np.float32)
        for c in range(3):
            if f[0, c] > f[0, c + 3]:
                f[0, c], f[0, c + 3] = 1.0, -1.0
        h = buf.view(np.uint16).reshape(n, 16)
        leaf = h[:, 15] == 0xFFFF
        u[leaf, 6] += np.uint32(geometry_offset)
        parts.append(u.view(np.uint8).reshape(-1))
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8)


def webgpu_nodes_to_roots(words: np.ndarray, geometry_offset: int = 0):
    """The inverse of webgpu_nodes: the reference root buffers packed back to back in a BVHNode buffer
    (each root's extent its subtree's span: the right spine's relative offsets plus one), leaf offsets
    less geometry_offset, an empty root's [1, -1] axes back to [inf, -inf]."""
    w = np.ascontiguousarray(words)
    w = (w.view(np.uint32) if w.dtype == np.uint8 else w.astype(np.uint32)).reshape(-1, 8)
    roots, start = [], 0
    while start < w.shape[0]:
        j, size = start, 1
        while (int(w[j, 7]) >> 16) != 0xFFFF:
            r = int(w[j, 6])
            j += r
            size += r
        u = w[start:start + size].copy()
        f = u.view(np.float32)
        for c in range(3):
            if f[0, c] > f[0, c + 3]:
                f[0, c], f[0, c + 3] = np.inf, -np.inf
        leaf = (u[:, 7] >> 16) == 0xFFFF
        u[leaf, 6] -= np.uint32(geometry_offset)
        roots.append(u.view(np.uint8).reshape(-1))
        start += size
    return roots


def box_triangle(box_min, box_max, box_to_mesh, a, b, c, clip=False) -> int:
    """One box against one triangle: OrientedBox's separating-axis test (clip=False) or the independent
    clipping formulation (clip=True; -1 for a singular matrix)."""
    f = library().orc_box_triangle_clip if clip else library().orc_box_triangle_sat
    arr = lambda v, n: np.ascontiguousarray(v, np.float64).reshape(n)
    return int(f(_p(arr(box_min, 3)), _p(arr(box_max, 3)), _p(arr(box_to_mesh, 16)), _p(arr(a, 3)), _p(arr(b, 3)),
                 _p(arr(c, 3))))


def triangle_distance(t1, t2):
    """ExtendedTriangle.distanceToTriangle of two triangles ((3, 3) each): (distance, point on t1, point on t2)."""
    a = np.ascontiguousarray(t1, np.float64).reshape(9)
    b = np.ascontiguousarray(t2, np.float64).reshape(9)
    o1 = np.zeros(3)
    o2 = np.zeros(3)
    d = library().orc_triangle_distance(_p(a), _p(b), _p(o1), _p(o2))
    return d, o1, o2


def segment_segment(s1, e1, s2, e2):
    """closestPointsSegmentToSegment: the closest points of two segments."""
    v = [np.ascontiguousarray(x, np.float64).reshape(3) for x in (s1, e1, s2, e2)]
    o1 = np.zeros(3)
    o2 = np.zeros(3)
    library().orc_segment_segment(*[_p(x) for x in v], _p(o1), _p(o2))
    return o1, o2


def scene_closest_point(meshes, world, geometry, visible, points):
    """Scene closest point by brute force (world space, double): per point (instance, face, distance,
    world point); equal distances to the lowest instance, then the lowest face."""
    refs = (C.c_void_p * len(meshes))(*[C.cast(C.pointer(m._m), C.c_void_p) for m in meshes])
    w = np.ascontiguousarray(world, np.float32).reshape(-1)
    g = np.ascontiguousarray(geometry, np.uint32)
    vis = None if visible is None else np.ascontiguousarray(visible, np.uint8)
    q = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    n = q.shape[0]
    inst = np.empty(n, np.uint32)
    face = np.empty(n, np.uint32)
    dist = np.empty(n)
    pts = np.empty((n, 3))
    f = C.c_uint32()
    d2 = C.c_double()
    for i in range(n):
        p = np.zeros(3)
        inst[i] = library().orc_scene_closest_point(refs, _p(w), _p(g), _p(vis), g.size, _p(q[i]), C.byref(f),
                                                    C.byref(d2), _p(p))
        face[i] = f.value
        dist[i] = np.sqrt(d2.value)
        pts[i] = p
    return {"instance": inst, "face": face, "distance": dist, "point": pts}


def _world_box(e, lb):
    """A local box through an affine column-major matrix: the 8 corners in double, rounded outward
    to float32 and widened by 2^-20 of their magnitude (the scene's top-level boxes)."""
    e = np.asarray(e, np.float64)
    corners = np.array([[lb[3 if c & 1 else 0], lb[4 if c & 2 else 1], lb[5 if c & 4 else 2]] for c in range(8)])
    out = []
    for c in corners:
        w = e[3] * c[0] + e[7] * c[1] + e[11] * c[2] + e[15]
        out.append([(e[k] * c[0] + e[4 + k] * c[1] + e[8 + k] * c[2] + e[12 + k]) / w for k in range(3)])
    out = np.array(out)
    mn, mx = out.min(axis=0), out.max(axis=0)
    pad = np.maximum(np.maximum(np.abs(mn), np.abs(mx)) * 9.5367431640625e-07, 1e-30)
    lo = np.nextafter((mn - pad).astype(np.float32), np.float32(-np.inf))
    hi = np.nextafter((mx + pad).astype(np.float32), np.float32(np.inf))
    return np.concatenate([lo, hi]).astype(np.float32)


def cluster_cut(nodes, primitive_limit):
    """ClusteredBVH's _traverseClusters (src/webgpu/ClusteredBVH.js:302-331) over one root's nodes
    ((N, 8) uint32, relative right child): depth first, left child first, a node is a cluster when it
    is a leaf or spans fewer than `primitive_limit` triangles (BVHTraversalHelper.getRangeStart /
    getRangeEnd, src/core/BVHTraversalHelper.js:41-67: the leftmost leaf's offset to the rightmost
    leaf's end); 0 means no limit (the root whole)."""
    def is_leaf(n):
        return (int(nodes[n, 7]) >> 16) == 0xFFFF

    def span(n):
        lo = hi = n
        while not is_leaf(lo):
            lo += 1
        while not is_leaf(hi):
            hi += int(nodes[hi, 6])
        return int(nodes[hi, 6]) + (int(nodes[hi, 7]) & 0xFFFF) - int(nodes[lo, 6])

    out = []

    def walk(n):
        if is_leaf(n) or primitive_limit == 0 or span(n) < primitive_limit:
            out.append(n)
            return
        walk(n + 1)
        walk(n + int(nodes[n, 6]))

    walk(0)
    return out


def subtree_node_count(nodes, n):
    """getSubtreeNodeCount (src/webgpu/utils/packBVHBufferUtils.js:142-161): the contiguous node span
    of the subtree at n."""
    if (int(nodes[n, 7]) >> 16) == 0xFFFF:
        return 1
    r = int(nodes[n, 6])
    return r + subtree_node_count(nodes, n + r)


def webgpu_scene(roots, slot_vertices, vertex_counts, world, geometry, visible, build, primitive_limit=0):
    """BVHComputeData.update's packing (src/webgpu/BVHComputeData.js:217-447) of ClusteredBVH's
    primitives (src/webgpu/ClusteredBVH.js:207-297): per instance and non-empty root, the root whole
    (primitive_limit 0) or its cluster cut (cluster_cut); one (instance, root, node) primitive per
    top-level leaf, the top level built by `build` (positions, index -> root buffer, permuted index)
    over the primitives' world boxes; appendBVHData (src/webgpu/utils/packBVHBufferUtils.js:44-131) for
    the top level, appendBVHSubtree (:216-280) for each (mesh, root, node) subtree the leaves name
    (first-named first; its contiguous span), transforms per slot (matrixWorld, its inverse, visible,
    padding; BVHComputeData.js:491-526) and the index per mesh rebased by its vertex base
    (appendIndexData, :293-322). roots[g]: the mesh's root buffers (uint8); slot_vertices[g]: its
    triangles' vertex ids in slot order (3 per slot). Returns nodes (N, 8) uint32, transforms (S, 36)
    uint32 with the inverse as float32 from numpy's double inverse, index uint32, and the primitive
    (instance, root, node) of each top-level slot."""
    prims, pos = [], []
    for i, g in enumerate(geometry):
        for r, rb in enumerate(roots[g]):
            b = rb.view(np.float32)[:6].astype(np.float64)
            if not (b[0] <= b[3] and b[1] <= b[4] and b[2] <= b[5]):
                continue
            rn = rb.view(np.uint32).reshape(-1, 8)
            for n in cluster_cut(rn, primitive_limit):
                wb = _world_box(world[i], rb.view(np.float32).reshape(-1, 8)[n, :6].astype(np.float64))
                prims.append((i, r, n))
                pos.append(np.concatenate([wb, wb[:3]]))
    np_ = len(prims)
    if np_ == 0:
        tl = np.zeros(8, np.uint32)
        tl.view(np.float32)[:3] = np.inf
        tl.view(np.float32)[3:6] = -np.inf
        tl[7] = 0xFFFF0000
        tlas, order = tl.reshape(1, 8), []
    else:
        pos = np.array(pos, np.float32).reshape(-1, 3)
        tb, idx = build(pos, np.arange(3 * np_, dtype=np.uint32))
        tlas = tb.view(np.uint32).reshape(-1, 8)
        order = [int(idx[3 * k]) // 3 for k in range(np_)]
    n_tlas = tlas.shape[0]
    sub, tri_base, vert_base, mesh_order, sub_order = {}, {}, {}, [], []
    nxt, ntri, nvert = n_tlas, 0, 0
    leaf_base = []
    for k in range(np_):
        i, r, n = prims[order[k]]
        g = int(geometry[i])
        if g not in tri_base:
            tri_base[g], vert_base[g] = ntri, nvert
            ntri += len(slot_vertices[g]) // 3
            nvert += int(vertex_counts[g])
            mesh_order.append(g)
        if (g, r, n) not in sub:
            size = subtree_node_count(roots[g][r].view(np.uint32).reshape(-1, 8), n)
            sub[(g, r, n)] = (nxt, size)
            nxt += size
            sub_order.append((g, r, n))
        leaf_base.append(sub[(g, r, n)][0])
    nodes = np.zeros((nxt, 8), np.uint32)

    def fix_root(dst, src):
        f = src.view(np.float32)
        df = dst.view(np.float32)
        for c in range(3):
            if f[c] > f[c + 3]:
                df[c], df[c + 3] = 1.0, -1.0

    for k in range(n_tlas):
        src = tlas[k]
        nodes[k, :6] = src[:6]
        if k == 0:
            fix_root(nodes[k], src)
        if (int(src[7]) >> 16) == 0xFFFF:
            cnt = int(src[7]) & 0xFFFF
            slot = 0 if cnt == 0 else int(src[6])
            nodes[k, 6] = 0 if cnt == 0 else leaf_base[slot]
            nodes[k, 7] = 0xFF000000 | (slot & 0x00FFFFFF)
        else:
            nodes[k, 6:8] = src[6:8]
    for g, r, n in sub_order:
        base, size = sub[(g, r, n)]
        src = roots[g][r].view(np.uint32).reshape(-1, 8)[n:n + size]
        for k in range(size):
            nodes[base + k, :6] = src[k, :6]
            if k == 0:
                fix_root(nodes[base + k], src[k])
            if (int(src[k, 7]) >> 16) == 0xFFFF:
                nodes[base + k, 6] = int(src[k, 6]) + tri_base[g]
                nodes[base + k, 7] = 0xFFFF0000 | (int(src[k, 7]) & 0xFFFF)
            else:
                nodes[base + k, 6:8] = src[k, 6:8]
    tr = np.zeros((max(np_, 1), 36), np.uint32)
    for k in range(np_):
        i = prims[order[k]][0]
        w = np.asarray(world[i], np.float32)
        tr[k, :16] = w.view(np.uint32)
        inv = np.linalg.inv(w.astype(np.float64).reshape(4, 4).T).T.reshape(-1).astype(np.float32)
        tr[k, 16:32] = inv.view(np.uint32)
        tr[k, 32] = 1 if (visible is None or visible[i]) else 0
    index = np.zeros(3 * ntri, np.uint32)
    for g in mesh_order:
        v = np.asarray(slot_vertices[g], np.uint32)
        index[3 * tri_base[g]:3 * tri_base[g] + v.size] = v + vert_base[g]
    return {"nodes": nodes, "transforms": tr, "index": index, "order": [prims[o] for o in order],
            "mesh_order": mesh_order, "vert_base": vert_base}
