"""Pins the oracle's restatement of MeshBVH.refit (oracle.c orc_refit_root, refit.template.js): a
refit is tight and idempotent; after a deformation every leaf box is exactly the min / max of
its triangles' vertex coordinates and every internal box the union of its children (checked with
numpy, independently of the oracle), and the tree validates; only changed paths are rewritten (an
inconsistent internal box above unchanged leaves is kept, as the reference keeps it); -0 and +0 are
equal (a stored -0 survives)."""
import numpy as np

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
