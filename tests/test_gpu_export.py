"""Exports of an uploaded mesh to the reference's shader data layouts (mbvh_mesh_export_textures /
_nodes) through the C ABI: MeshBVHUniformStruct's bvhBounds / bvhContents / index / position
textures and the WebGPU BVHNode buffer equal the oracle's restatement of the reference's packing of
the same tree and geometry (uint16 and uint32 index, indirect, non-indexed, groups, after a refit);
the device form equals the host form. Imports (mbvh_mesh_import_nodes / _textures): the packed
buffers back to device meshes whose node buffers equal the original roots (and the oracle's inverse
restatement) and whose queries equal the original mesh's bit for bit."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G
import oracle_py as O
from helpers import build_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _index_triplets(g, idx, tree):
    if tree.indirect is None:
        return idx.reshape(-1, 3).astype(np.uint32)
    tri = tree.indirect.astype(np.int64)
    if idx is None:
        return (3 * tri[:, None] + np.arange(3)).astype(np.uint32)
    return idx.reshape(-1, 3)[tri].astype(np.uint32)


@pytest.mark.parametrize("name,make,indirect", [
    ("torusknot80k_u16", lambda: G.torus_knot(10, 3, 400, 100), False),
    ("sphere_u32", lambda: G.sphere(1, 300, 150), False),
    ("indirect_indexed", lambda: G.torus(1, 0.4, 30, 90), True),
    ("indirect_nonindexed", lambda: G.icosahedron(1, 6), True),
])
def test_textures_and_nodes(ctx, name, make, indirect):
    print(f"case {name}", flush=True)
    g, idx, tree = build_mesh(make(), indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    got = mesh.export_textures(g.positions)
    b, c = O.uniform_textures(tree.roots[0])
    assert np.array_equal(got["bounds"].view(np.uint32), b.view(np.uint32))
    assert np.array_equal(got["contents"], c)
    assert np.array_equal(got["index"], O.attribute_texture(_index_triplets(g, idx, tree), 3))
    assert np.array_equal(got["position"].view(np.uint32), O.attribute_texture(g.positions, 3).view(np.uint32))
    assert np.array_equal(mesh.export_nodes(), O.webgpu_nodes(tree.roots))
    sides = mesh.texture_sizes()
    assert [x * x for x in sides] == [b.shape[0], c.shape[0], got["index"].shape[0], got["position"].shape[0]]
    mesh.free()


def test_groups_refit_and_errors(ctx):
    g0 = G.torus_knot(1, 0.3, 200, 20)
    n = g0.tri_count
    g, idx, tree = build_mesh(g0, ranges=[(0, n // 3), (n // 3, n - n // 3)])
    mesh = ctx.upload(tree, g.positions, idx)
    assert np.array_equal(mesh.export_nodes(), O.webgpu_nodes(tree.roots))
    with pytest.raises(M.MbvhError):
        mesh.export_textures()  # multi-root BVHs are not supported by the uniform struct
    mesh.free()
    g, idx, tree = build_mesh(G.sphere(1, 64, 32))
    mesh = ctx.upload(tree, g.positions, idx)
    rng = np.random.default_rng(3)
    pos = (g.positions.reshape(-1, 3) + rng.normal(0, 0.05, (g.positions.shape[0], 3))).astype(np.float32)
    roots = mesh.refit(pos)
    got = mesh.export_textures(pos)
    b, c = O.uniform_textures(roots[0])
    assert np.array_equal(got["bounds"].view(np.uint32), b.view(np.uint32)) and np.array_equal(got["contents"], c)
    assert np.array_equal(mesh.export_nodes(), O.webgpu_nodes(roots))
    mesh.free()


def test_device_form(ctx):
    torch = pytest.importorskip("torch")
    g, idx, tree = build_mesh(G.torus_knot(1, 0.3, 300, 30))
    mesh = ctx.upload(tree, g.positions, idx)
    host = mesh.export_textures(g.positions)
    dev = torch.device("cuda:0")
    sb, sc, si, sp = mesh.texture_sizes()
    pos = torch.from_numpy(g.positions.reshape(-1)).to(dev)
    bounds = torch.full((sb * sb, 4), 7.0, device=dev)   # stale contents: padding must be rewritten
    contents = torch.full((sc * sc, 2), 7, dtype=torch.int32, device=dev)
    index = torch.full((si * si, 4), 7, dtype=torch.int32, device=dev)
    position = torch.full((sp * sp, 4), 7.0, device=dev)
    nodes = torch.empty(tree.roots[0].size, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    lib, v = ctx._lib, (lambda t: t.data_ptr())
    assert lib.mbvh_mesh_export_textures_d(ctx.handle, mesh.handle, v(pos), 3, 0, v(bounds), v(contents), v(index),
                                           v(position), None) == 0
    assert lib.mbvh_mesh_export_nodes_d(ctx.handle, mesh.handle, v(nodes), nodes.numel(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(bounds.cpu().numpy().view(np.uint32), host["bounds"].view(np.uint32))
    assert np.array_equal(contents.cpu().numpy().view(np.uint32), host["contents"])
    assert np.array_equal(index.cpu().numpy().view(np.uint32), host["index"])
    assert np.array_equal(position.cpu().numpy().view(np.uint32), host["position"].view(np.uint32))
    assert np.array_equal(nodes.cpu().numpy(), tree.roots[0])
    mesh.free()


def _same_queries(a, b, g):
    lo, hi = g.bounds()
    rng = np.random.default_rng(8)
    o = (lo + (hi - lo) * rng.random((2000, 3)) * 3 - (hi - lo)).astype(np.float32)
    t = (lo + (hi - lo) * rng.random((2000, 3))).astype(np.float32)
    d = (t - o) / np.linalg.norm(t - o, axis=1, keepdims=True)
    ra, rb, ca, cb = a.raycast_first(o, d), b.raycast_first(o, d), a.closest_point(t), b.closest_point(t)
    for k in ("face", "distance", "point", "barycoord"):  # bit for bit (misses: NaN points)
        assert np.array_equal(np.asarray(ra[k]).view(np.uint32), np.asarray(rb[k]).view(np.uint32)), k
        assert np.array_equal(np.asarray(ca[k]).view(np.uint32), np.asarray(cb[k]).view(np.uint32)), k


@pytest.mark.parametrize("name,make,indirect,ranges", [
    ("torusknot_u16", lambda: G.torus_knot(10, 3, 200, 40), False, None),
    ("sphere_u32_groups_empty", lambda: G.sphere(1, 300, 150), False, "empty"),
    ("indirect", lambda: G.torus(1, 0.4, 30, 90), True, None),
])
def test_import_nodes_and_textures(ctx, name, make, indirect, ranges):
    print(f"case {name}", flush=True)
    g0 = make()
    n = g0.tri_count
    rr = [(0, n // 3), (n // 3, 0), (n // 3, n - n // 3)] if ranges == "empty" else None
    g, idx, tree = build_mesh(g0, indirect=indirect, ranges=rr)
    mesh = ctx.upload(tree, g.positions, idx)
    for off in (0, 12345):
        words = O.webgpu_nodes(tree.roots, geometry_offset=off)
        back = O.webgpu_nodes_to_roots(words, off)
        assert len(back) == len(tree.roots)
        for a, b in zip(back, tree.roots):
            assert np.array_equal(a, np.asarray(b).view(np.uint8))
        imp = ctx.import_nodes(words, g.positions, idx, geometry_offset=off, indirect=tree.indirect)
        for a, b in zip(imp.copy_roots(), tree.roots):
            assert np.array_equal(a, np.asarray(b).view(np.uint8))
        _same_queries(mesh, imp, g)
        imp.free()
    if len(tree.roots) == 1:
        tex = mesh.export_textures(g.positions)
        nn = tree.roots[0].nbytes // 32
        assert np.array_equal(O.textures_to_root(tex["bounds"], tex["contents"], nn), np.asarray(tree.roots[0]).view(np.uint8))
        imp = ctx.import_textures(tex["bounds"], tex["contents"], nn, g.positions, idx, indirect=tree.indirect)
        assert np.array_equal(imp.copy_roots()[0], np.asarray(tree.roots[0]).view(np.uint8))
        _same_queries(mesh, imp, g)
        imp.free()
    mesh.free()


def test_import_rejects_bad_buffers(ctx):
    g, idx, tree = build_mesh(G.torus_knot(1, 0.3, 40, 8))
    words = O.webgpu_nodes(tree.roots).view(np.uint32).reshape(-1, 8).copy()
    bad = words.copy()
    bad[0, 6] = bad.shape[0] + 5  # the root's right child beyond the buffer
    with pytest.raises(M.MbvhError):
        ctx.import_nodes(bad, g.positions, idx)
    with pytest.raises(M.MbvhError):  # leaf offsets below the geometry offset
        ctx.import_nodes(words, g.positions, idx, geometry_offset=7)


# ---- a scene in the WebGPU packing (BVHComputeData.update with instances referenced whole) ----------

def _packed_first_hits(pk, positions, o, d):
    """The merged top-level / instance loop of the reference's WebGPU shapecast
    (src/webgpu/shapecastFns/getShapecastFn.js:97-217) with the first-hit tests, in float64 over the
    packed buffers: a top-level leaf (0xFF top byte) names its transform slot and subtree; the ray goes
    into the instance frame by the slot's inverse and the subtree's triangles come from the packed
    index. Returns the least world distance and the slot that reached it (double-sided)."""
    nodes, tr, index = pk["nodes"], pk["transforms"], pk["index"]
    nf = nodes.view(np.float32).astype(np.float64)

    def enters(n, oo, dd):
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / dd
            t0, t1 = (nf[n, 0:3] - oo) * inv, (nf[n, 3:6] - oo) * inv
        tmin, tmax = np.nanmax(np.minimum(t0, t1)), np.nanmin(np.maximum(t0, t1))
        return tmin <= tmax and tmax >= 0

    def tri_t(a, b, c, oo, dd):
        e1, e2 = b - a, c - a
        p = np.cross(dd, e2)
        det = e1 @ p
        if abs(det) < 1e-300:
            return np.inf
        tv = oo - a
        u = (tv @ p) / det
        q = np.cross(tv, e1)
        v = (dd @ q) / det
        t = (e2 @ q) / det
        return t if (u >= 0 and v >= 0 and u + v <= 1 and t >= 0) else np.inf

    best, best_slot = np.inf, -1
    stack = [0]
    while stack:
        n = stack.pop()
        if not enters(n, o, d):
            continue
        w7 = int(nodes[n, 7])
        if (w7 >> 24) == 0xFF:  # top-level leaf
            slot = w7 & 0xFFFFFF
            W = tr[slot, :16].view(np.float32).astype(np.float64).reshape(4, 4).T
            I = tr[slot, 16:32].view(np.float32).astype(np.float64).reshape(4, 4).T
            lo = I[:3, :3] @ o + I[:3, 3]
            ld = I[:3, :3] @ d
            sub = [int(nodes[n, 6])]
            while sub:
                m = sub.pop()
                if not enters(m, lo, ld):
                    continue
                x7 = int(nodes[m, 7])
                if (x7 >> 16) == 0xFFFF:
                    off, cnt = int(nodes[m, 6]), x7 & 0xFFFF
                    for t_ in range(off, off + cnt):
                        a, b, c = (positions[index[3 * t_ + j]].astype(np.float64) for j in range(3))
                        t = tri_t(a, b, c, lo, ld)
                        if np.isfinite(t):
                            wp = W[:3, :3] @ (lo + t * ld) + W[:3, 3]
                            dist = np.linalg.norm(wp - o)
                            if dist < best:
                                best, best_slot = dist, slot
                else:
                    sub += [m + int(nodes[m, 6]), m + 1]
        else:
            stack += [n + int(nodes[n, 6]), n + 1]
    return best, best_slot


def _check_cluster_cut(nodes, cut, limit):
    """A cluster cut of one root's nodes covers every triangle once, each cluster is a leaf or spans
    fewer than `limit` triangles, and its parent spans at least `limit` (the first such node on its path)."""
    parent, stack = {0: None}, [0]
    while stack:
        n = stack.pop()
        if (int(nodes[n, 7]) >> 16) != 0xFFFF:
            for c in (n + 1, n + int(nodes[n, 6])):
                parent[c] = n
                stack.append(c)

    def tris(n):
        if (int(nodes[n, 7]) >> 16) == 0xFFFF:
            return list(range(int(nodes[n, 6]), int(nodes[n, 6]) + (int(nodes[n, 7]) & 0xFFFF)))
        return tris(n + 1) + tris(n + int(nodes[n, 6]))

    covered = []
    for n in cut:
        t = tris(n)
        covered += t
        leaf = (int(nodes[n, 7]) >> 16) == 0xFFFF
        assert leaf or limit == 0 or len(t) < limit, (n, len(t))
        if limit and parent[n] is not None:
            assert len(tris(parent[n])) >= limit, n
        if limit == 0:
            assert n == 0
    assert sorted(covered) == tris(0)


@pytest.mark.parametrize("limit", [0, 64, 1])
def test_scene_webgpu_packing(ctx, limit):
    """mbvh_scene_export_webgpu == the oracle's restatement of the packing (top level with one
    (instance, root, cluster) per leaf, tagged leaves, each referenced subtree once with rebased leaves,
    transforms, rebased index) for plain, two-root and indirect meshes shared by several instances, one
    of them invisible, with roots whole (limit 0), cut into clusters of fewer than 64 triangles (the
    reference's limit) and cut at every leaf (limit 1); the cut is checked against its definition; and
    the merged loop over the packed buffers finds the scene raycast's hits."""
    geo = G.torus_knot(1.0, 0.3, 60, 10)
    geo.ensure_index()
    ntri = geo.index.size // 3
    built = [build_mesh(G.sphere(1.0, 24, 12)), build_mesh(geo, ranges=[(0, ntri // 2), (ntri // 2, ntri - ntri // 2)]),
             build_mesh(G.torus(1.0, 0.3, 12, 30), indirect=True)]
    ms = [ctx.upload(t, g.positions, idx) for g, idx, t in built]
    rng = np.random.default_rng(71)
    geometry = np.array([0, 1, 2, 1, 0, 2, 1], np.uint32)
    mats = []
    for k in range(len(geometry)):
        m = np.eye(4)
        a = rng.normal(size=3)
        c, s_ = np.cos(0.7 * k), np.sin(0.7 * k)
        m[:3, :3] = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]) * rng.uniform(0.6, 1.5)
        m[:3, 3] = a * 4
        mats.append(m)
    world = np.stack([m.T.reshape(-1) for m in mats]).astype(np.float32)
    inverse = np.stack([np.linalg.inv(m).T.reshape(-1) for m in mats]).astype(np.float32)
    visible = np.ones(len(geometry), np.uint8)
    visible[4] = 0
    sc = ctx.scene(ms, world, inverse, geometry, visible)
    got = sc.export_webgpu(limit)
    slot_vertices = [_index_triplets(g, idx, t).reshape(-1) for g, idx, t in built]
    for _, _, t in built:
        for rb in t.roots:
            _check_cluster_cut(rb.view(np.uint32).reshape(-1, 8), O.cluster_cut(rb.view(np.uint32).reshape(-1, 8), limit),
                               limit)

    def build(pos, idx):
        tree = M.build(pos, idx, strategy=M.SAH, target_leaf_size=1, max_depth=40, strict_leaf_size=1)
        return tree.roots[0], idx

    ref = O.webgpu_scene([t.roots for _, _, t in built], slot_vertices, [g.positions.size // 3 for g, _, _ in built],
                         world, geometry, visible, build, primitive_limit=limit)
    assert np.array_equal(got["nodes"], ref["nodes"])
    n_roots = sum(len(built[g][2].roots) for g in geometry)
    assert (got["transforms"].shape[0] == n_roots) == (limit == 0)
    assert np.array_equal(got["index"], ref["index"])
    assert np.array_equal(got["transforms"][:, :16], ref["transforms"][:, :16])
    assert np.array_equal(got["transforms"][:, 32:], ref["transforms"][:, 32:])  # visible flag, padding
    gi = got["transforms"][:, 16:32].view(np.float32)
    ri = ref["transforms"][:, 16:32].view(np.float32)
    assert np.all(np.abs(gi - ri) <= 2 * np.spacing(np.maximum(np.abs(gi), np.abs(ri))))  # inverses to an ulp
    # a top level of 2S - 1 nodes (one primitive per leaf) first, then each named subtree's span once,
    # and every top-level leaf enters a copy of its (mesh, root, node) subtree's boxes
    n_slots = got["transforms"].shape[0]
    spans = {(int(geometry[i]), r, n): O.subtree_node_count(built[int(geometry[i])][2].roots[r].view(np.uint32).reshape(-1, 8), n)
             for i, r, n in ref["order"]}
    assert got["nodes"].shape[0] == 2 * n_slots - 1 + sum(spans.values())
    leaf_of = {int(w) & 0xFFFFFF: j for j, w in enumerate(got["nodes"][:2 * n_slots - 1, 7]) if (int(w) >> 24) == 0xFF}
    assert sorted(leaf_of) == list(range(n_slots))
    for k, (i, r, n) in enumerate(ref["order"]):
        g = int(geometry[i])
        rn = built[g][2].roots[r].view(np.uint32).reshape(-1, 8)
        base, size = int(got["nodes"][leaf_of[k], 6]), spans[(g, r, n)]
        assert base >= 2 * n_slots - 1
        assert np.array_equal(got["nodes"][base + 1:base + size, :6], rn[n + 1:n + size, :6])
    # the packed buffers are usable: the merged loop finds the scene raycast's hits (visible instances)
    positions = np.concatenate([built[g][0].positions.reshape(-1, 3) for g in ref["mesh_order"]])
    sc_vis = ctx.scene(ms, world, inverse, geometry)  # the loop below ignores the visible flag
    n = 200
    o = rng.normal(0, 12, (n, 3)).astype(np.float32)
    tgt = world[rng.integers(0, len(geometry), n), 12:15] + rng.normal(0, 0.8, (n, 3)).astype(np.float32)
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1, keepdims=True)
    d = d.astype(np.float32)
    res = sc_vis.raycast_first(o, d, side=M.DOUBLE_SIDE)
    order = ref["order"]
    hits = 0
    for i in range(n):
        dist, slot = _packed_first_hits(ref, positions, o[i].astype(np.float64), d[i].astype(np.float64))
        if res["face"][i] == M.NO_HIT:
            assert not np.isfinite(dist), i
            continue
        hits += 1
        assert abs(dist - res["distance"][i]) <= 1e-5 * max(dist, 1.0), (i, dist, res["distance"][i])
        inst = order[slot][0]
        assert inst == res["instance"][i] or abs(dist - res["distance"][i]) <= 1e-6 * max(dist, 1.0), i
    assert hits > n // 4
    sc_vis.free()
    sc.free()
    for m in ms:
        m.free()
