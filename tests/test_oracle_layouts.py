"""The oracle's restatements of the reference's shader data layouts (MeshBVHUniformStruct textures,
VertexAttributeTexture, the WebGPU BVHNode buffer), pinned by a hand-written three-node tree whose
texels are spelled out, by the texture-side closed forms, and by the round trip textures -> node
buffer on built trees."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G
import oracle_py as O


def _tiny_tree():
    # root (internal, split axis 1, right child 2 nodes on), leaf 0 (offset 0, 2 tris), leaf 1 (offset 2, 1 tri)
    f = np.zeros((3, 8), np.float32)
    f[0, :6] = [-1, -2, -3, 4, 5, 6]
    f[1, :6] = [-1, -2, -3, 0, 5, 1]
    f[2, :6] = [0, -1, -3, 4, 2, 6]
    u = f.view(np.uint32)
    u[0, 6], u[0, 7] = 2, 1
    u[1, 6], u[1, 7] = 0, 0xFFFF0000 | 2
    u[2, 6], u[2, 7] = 2, 0xFFFF0000 | 1
    return f.view(np.uint8).reshape(-1)


def test_tiny_tree_texels():
    bounds, contents = O.uniform_textures(_tiny_tree())
    assert bounds.shape == (4 * 4, 4) and contents.shape == (2 * 2, 2)   # sides 2*ceil(sqrt(1.5)) = 4, ceil(sqrt(3)) = 2
    assert bounds[0].tolist() == [-1, -2, -3, 0] and bounds[1].tolist() == [4, 5, 6, 0]
    assert bounds[2].tolist() == [-1, -2, -3, 0] and bounds[3].tolist() == [0, 5, 1, 0]
    assert bounds[4].tolist() == [0, -1, -3, 0] and bounds[5].tolist() == [4, 2, 6, 0]
    assert not bounds[6:].any()
    assert contents[0].tolist() == [1, 2]                   # split axis, right child offset
    assert contents[1].tolist() == [0xFFFF0002, 0]          # LEAFNODE_MASK_32 | count, offset
    assert contents[2].tolist() == [0xFFFF0001, 2]
    assert contents[3].tolist() == [0, 0]


@pytest.mark.parametrize("nodes,bside,cside", [(1, 2, 1), (3, 4, 2), (8, 4, 3), (9, 6, 3), (32, 8, 6), (33, 10, 6)])
def test_texture_sides(nodes, bside, cside):
    root = np.zeros(32 * nodes, np.uint8)
    root.view(np.uint16)[15::16] = 0xFFFF
    b, c = O.uniform_textures(root)
    assert b.shape[0] == bside * bside and c.shape[0] == cside * cside


def test_attribute_texture():
    v = np.arange(15, dtype=np.float32).reshape(5, 3)
    t = O.attribute_texture(v, 3)
    assert t.shape == (9, 4)                                 # ceil(sqrt(5)) = 3
    assert t[4].tolist() == [12, 13, 14, 1] and not t[5:].any()
    idx = np.arange(6, dtype=np.uint16)
    ti = O.attribute_texture(idx, 3)
    assert ti.dtype == np.uint16 and ti[1].tolist() == [3, 4, 5, 1]
    assert O.attribute_texture(np.zeros((0, 3), np.float32), 3).shape == (1, 4)


def test_round_trip_and_webgpu_nodes():
    g = G.torus_knot(10, 3, 60, 12)
    idx = g.index.copy()
    tree = M.build(g.positions, idx)
    root = tree.roots[0]
    b, c = O.uniform_textures(root)
    assert np.array_equal(O.textures_to_root(b, c, root.size // 32), root)
    nodes = O.webgpu_nodes(tree.roots)
    assert np.array_equal(nodes, root)                        # nothing to fix on a non-empty mesh
    empty = np.zeros(32, np.uint8)
    empty.view(np.float32)[:6] = [np.inf, np.inf, np.inf, -np.inf, -np.inf, -np.inf]
    empty.view(np.uint16)[15] = 0xFFFF
    fixed = O.webgpu_nodes([empty], geometry_offset=7)
    assert fixed.view(np.float32)[:6].tolist() == [1, 1, 1, -1, -1, -1]
    assert fixed.view(np.uint32)[6] == 7


def test_webgpu_nodes_round_trip():
    """webgpu_nodes_to_roots inverts webgpu_nodes: several roots (one empty, whose [inf, -inf] bounds
    the packing writes as [1, -1]) and a nonzero geometry offset."""
    from mbvh_amd import geometry as G
    from helpers import build_mesh
    g0 = G.sphere(1, 40, 20)
    n = g0.tri_count
    g, idx, tree = build_mesh(g0, ranges=[(0, n // 3), (n // 3, 0), (n // 3, n - n // 3)])
    assert len(tree.roots) == 3
    w = O.webgpu_nodes(tree.roots, geometry_offset=999)
    assert w.view(np.float32)[8 * (tree.roots[0].nbytes // 32):][:6].tolist() == [1, 1, 1, -1, -1, -1]
    back = O.webgpu_nodes_to_roots(w, 999)
    assert len(back) == 3
    for a, b in zip(back, tree.roots):
        assert np.array_equal(a, np.asarray(b).view(np.uint8))
