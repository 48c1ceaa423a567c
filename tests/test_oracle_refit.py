"""Pins the oracle's restatement of MeshBVH.refit (oracle.c orc_refit_root, refit.template.js): a
refit is tight and idempotent; after a deformation every leaf box is exactly the min / max of
its triangles' vertex coordinates and every internal box the union of its children (checked with
numpy, independently of the oracle), and the tree validates; only changed paths are rewritten (an
inconsistent internal box above unchanged leaves is kept, as the reference keeps it); -0 and +0 are
equal (a stored -0 survives). Partial refits (nodeIndices) against boxes worked out by hand on a
five-node tree and a two-root tree, and the reference's own case (every node listed = a full refit)."""
import numpy as np
import pytest

import mbvh_amd as M
import oracle_py as O
from mbvh_amd import geometry as G


def _tree(g):
    g.ensure_index()
    idx = g.index.copy()
    return idx, M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)


def _nodes(buf):
    f = buf.view(np.float32).reshape(-1, 8)
    u = buf.view(np.uint32).reshape(-1, 8)
    h = buf.view(np.uint16).reshape(-1, 16)
    return f, u, h


def _check_tight(buf, positions, index):
    """Leaf box = min / max over its triangles' vertices; internal box = union of the children."""
    f, u, h = _nodes(buf)
    pos = positions.reshape(-1, 3)
    tri = index.reshape(-1, 3)
    n = f.shape[0]
    for i in range(n):
        if h[i, 15] == 0xFFFF:
            off, cnt = int(u[i, 6]), int(h[i, 14])
            v = pos[tri[off:off + cnt].reshape(-1)]
            assert np.array_equal(f[i, :3], v.min(axis=0)) and np.array_equal(f[i, 3:6], v.max(axis=0)), i
        else:
            l, r = i + 1, i + int(u[i, 6])
            assert np.array_equal(f[i, :3], np.minimum(f[l, :3], f[r, :3])), i
            assert np.array_equal(f[i, 3:6], np.maximum(f[l, 3:6], f[r, 3:6])), i


def test_refit_is_tight_and_idempotent():
    g = G.torus_knot(10, 3, 120, 16)
    idx, tree = _tree(g)
    once = O.OracleMesh(g.positions, idx, tree.roots).refit()
    _check_tight(once[0], g.positions, idx)
    # the build stores float32 boxes rounded outwards from FP64 centre +- half extent, so they
    # contain the tight ones
    fb, _, _ = _nodes(tree.roots[0])
    ft, _, _ = _nodes(once[0])
    assert np.all(fb[:, :3] <= ft[:, :3]) and np.all(fb[:, 3:6] >= ft[:, 3:6])
    twice = O.OracleMesh(g.positions, idx, once).refit()
    assert np.array_equal(twice[0], once[0])  # unchanged positions: nothing changes


def test_refit_after_deformation_is_tight_and_valid():
    g = G.torus_knot(10, 3, 120, 16)
    idx, tree = _tree(g)
    rng = np.random.default_rng(4)
    pos = g.positions.reshape(-1, 3).copy()
    pos += rng.normal(0, 0.3, pos.shape).astype(np.float32)
    pos[rng.integers(0, len(pos), 20)] *= np.float32(3.0)  # a few vertices moved far out
    om = O.OracleMesh(pos, idx, tree.roots)
    new = om.refit()
    assert not np.array_equal(new[0], tree.roots[0])
    _check_tight(new[0], pos, idx)
    # only the boxes moved: offsets, counts, split axes and leaf flags are the original words
    fo, uo, ho = _nodes(tree.roots[0])
    fn, un, hn = _nodes(new[0])
    assert np.array_equal(uo[:, 6:], un[:, 6:])
    rc, st = O.OracleMesh(pos, idx, new).validate_root(0)
    assert rc == 0 and st["bounds_failures"] == 0


def test_refit_rewrites_changed_paths_only():
    g = G.torus_knot(10, 3, 60, 12)
    idx, tree = _tree(g)
    buf = O.OracleMesh(g.positions, idx, tree.roots).refit()[0]  # tight boxes first
    f, u, h = _nodes(buf)
    internal = [i for i in range(f.shape[0]) if h[i, 15] != 0xFFFF]
    k = internal[len(internal) // 2]
    f[k, 0] -= np.float32(5.0)  # an inconsistent (too large) internal box above unchanged leaves
    out = O.OracleMesh(g.positions, idx, [buf]).refit()[0]
    assert np.array_equal(out, buf)  # no leaf changed: nothing is recomputed, the box is kept
    # moving one vertex of a leaf under node k recomputes the path to the root, node k included
    fo, uo, ho = _nodes(out)
    lo, hi = k, k
    while h[hi, 15] != 0xFFFF:
        hi = hi + int(u[hi, 6])  # rightmost leaf of k's subtree
    off = int(u[hi, 6])
    pos = g.positions.reshape(-1, 3).copy()
    pos[idx.reshape(-1, 3)[off, 0], 0] = f[hi, 3] + np.float32(0.5)  # past the leaf's max.x: the leaf changes
    out2 = O.OracleMesh(pos, idx, [buf]).refit()[0]
    f2, _, _ = _nodes(out2)
    assert f2[k, 0] > f[k, 0]  # the too-large box was recomputed from its children
    assert np.array_equal(f2[0, :3], np.minimum(f2[1, :3], f2[int(u[0, 6]), :3]))


def test_refit_signed_zero_is_unchanged():
    pos = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([0, 1, 2], np.uint16)
    tree = M.build(pos, idx.copy(), strategy=M.SAH, target_leaf_size=10)
    buf = O.OracleMesh(pos, idx, tree.roots).refit()[0]
    f, _, _ = _nodes(buf)
    assert f[0, 2] == 0 and not np.signbit(f[0, 2])
    f[0, 2] = np.float32(-0.0)  # stored -0 for min.z; the vertices give +0
    out = O.OracleMesh(pos, idx, [buf]).refit()[0]
    assert np.signbit(_nodes(out)[0][0, 2])  # -0 == +0: the leaf is unchanged, -0 stays


# ---- refit(nodeIndices): refit.template.js:136-172 -------------------------------------------

def _node_bytes(specs):
    """Reference node buffer from (box6, kind, a, b): kind 'leaf' (a = offset, b = count) or 'inner'
    (a = right child offset, b = split axis)."""
    buf = np.zeros(32 * len(specs), np.uint8)
    f, u, h = _nodes(buf)
    for i, (box, kind, a, b) in enumerate(specs):
        f[i, :6] = box
        u[i, 6] = a
        if kind == "leaf":
            h[i, 14], h[i, 15] = b, 0xFFFF
        else:
            u[i, 7] = b
    return buf


# three triangles: one in z = 0 near x = 0, one in z = 0 near x = 2, one in z = 5
_TRIS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0],
                  [2, 0, 0], [3, 0, 0], [2, 1, 0],
                  [0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
_B0, _B1, _B2 = [0, 0, 0, 1, 1, 0], [2, 0, 0, 3, 1, 0], [0, 0, 5, 1, 1, 5]
# node 0 = (node 1 = (leaf 2: tri 0, leaf 3: tri 1), leaf 4: tri 2)
_FIVE = [([0, 0, 0, 3, 1, 5], "inner", 4, 2), ([0, 0, 0, 3, 1, 0], "inner", 2, 0),
         (_B0, "leaf", 0, 1), (_B1, "leaf", 1, 1), (_B2, "leaf", 2, 1)]


def _boxes(buf):
    return [list(map(float, r)) for r in _nodes(buf)[0][:, :6]]


@pytest.mark.parametrize("ids,expect", [
    # {1}: the root visits node 1 only; node 1 has neither child listed, so both are refit (forced);
    # leaf 4 keeps its old box and the root is the union of new node 1 and old leaf 4
    ({1}, [[0, 0, 0, 6, 2, 5], [0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0], _B2]),
    # {1, 3}: node 1 visits leaf 3 only; leaf 2 keeps its old box
    ({1, 3}, [[0, 0, 0, 6, 2, 5], [0, 0, 0, 6, 2, 0], _B0, [4, 0, 0, 6, 2, 0], _B2]),
    # {4}: the root visits leaf 4 only; node 1's subtree keeps its boxes
    ({4}, [[0, 0, 0, 3, 2, 10], [0, 0, 0, 3, 1, 0], _B0, _B1, [0, 0, 10, 2, 2, 10]]),
    # {2}, {0}, {}: neither of the root's children is listed, so everything is refit (forced)
    ({2}, [[0, 0, 0, 6, 2, 10], [0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0], [0, 0, 10, 2, 2, 10]]),
    ({0}, [[0, 0, 0, 6, 2, 10], [0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0], [0, 0, 10, 2, 2, 10]]),
    (set(), [[0, 0, 0, 6, 2, 10], [0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0], [0, 0, 10, 2, 2, 10]]),
    # ids that name no node never match
    ({-1, 99, 1}, [[0, 0, 0, 6, 2, 5], [0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0], _B2]),
])
def test_partial_refit_by_hand(ids, expect):
    idx = np.arange(9, dtype=np.uint16)
    out = O.OracleMesh(_TRIS * np.float32(2), idx, [_node_bytes(_FIVE)]).refit(ids)[0]
    assert _boxes(out) == [list(map(float, e)) for e in expect]
    full = O.OracleMesh(_TRIS * np.float32(2), idx, [_node_bytes(_FIVE)]).refit()[0]
    if not ids & {1, 3, 4}:
        assert np.array_equal(out, full)  # a set without the root's children is a full refit


def test_partial_refit_numbers_nodes_over_all_roots():
    """Root r's node j is id (nodes of the roots before r) + j, as shapecast reports them."""
    ra = _node_bytes([([0, 0, 0, 3, 1, 0], "inner", 2, 0), (_B0, "leaf", 0, 1), (_B1, "leaf", 1, 1)])
    rb = _node_bytes([(_B2, "leaf", 2, 1)])
    idx = np.arange(9, dtype=np.uint16)
    pos = _TRIS * np.float32(2)
    # {4}: no node of root A (ids 0-2) has a listed child, so root A is refit whole; root B (id 3) is a
    # leaf root, refit whatever the set holds
    a, b = O.OracleMesh(pos, idx, [ra, rb]).refit({4})
    assert _boxes(a) == [[0, 0, 0, 6, 2, 0], [0, 0, 0, 2, 2, 0], [4, 0, 0, 6, 2, 0]]
    assert _boxes(b) == [[0, 0, 10, 2, 2, 10]]
    # {2}: root A visits its right leaf only
    a, _ = O.OracleMesh(pos, idx, [ra, rb]).refit({2})
    assert _boxes(a) == [[0, 0, 0, 6, 2, 0], _B0, [4, 0, 0, 6, 2, 0]]


def test_partial_refit_with_every_node_is_a_full_refit():
    """The reference's own case (test/MeshBVH.test.js:340-367): every node id from a shapecast, then a
    scale by 2; the result is the full refit, tight everywhere."""
    g = G.torus_knot(10, 3, 120, 16)
    idx, tree = _tree(g)
    pos = (g.positions.reshape(-1, 3) * np.float32(2)).astype(np.float32)
    n = tree.roots[0].nbytes // 32
    part = O.OracleMesh(pos, idx, tree.roots).refit(range(n))[0]
    assert np.array_equal(part, O.OracleMesh(pos, idx, tree.roots).refit()[0])
    _check_tight(part, pos, idx)
