"""Exports of an uploaded mesh to the reference's shader data layouts (mbvh_mesh_export_textures /
_nodes) through the C ABI: MeshBVHUniformStruct's bvhBounds / bvhContents / index / position
textures and the WebGPU BVHNode buffer equal the oracle's restatement of the reference's packing of
the same tree and geometry (uint16 and uint32 index, indirect, non-indexed, groups, after a refit);
the device form equals the host form."""
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
