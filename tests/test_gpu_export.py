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
