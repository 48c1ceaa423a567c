"""GPU refit (MeshBVH.refit on the device) through the C ABI: the refit reference-format nodes equal
the oracle's refit byte for byte (deformed positions, an inconsistent internal box, groups,
indirect); queries after the refit equal queries on a fresh upload of the refit tree bit for bit and
match the FP64 oracle on the deformed geometry; a batch containing the mesh sees the refit."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, check_closest_parity, check_raycast_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _deform(positions, seed, scale=0.2):
    rng = np.random.default_rng(seed)
    p = positions.reshape(-1, 3).copy()
    p += rng.normal(0, scale, p.shape).astype(np.float32)
    p[rng.integers(0, len(p), 10)] *= np.float32(1.5)
    return p.astype(np.float32)


def _rays(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.standard_normal((n, 3)).astype(np.float32)
    o = (o / np.linalg.norm(o, axis=1, keepdims=True) * np.float32(2.5 * np.abs(hi - lo).max())).astype(np.float32)
    t = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = (t - o).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    return o, d


def _same(a, b):
    for k in ("face", "distance", "point", "barycoord"):
        assert np.array_equal(np.asarray(a[k]).view(np.uint32), np.asarray(b[k]).view(np.uint32)), k


@pytest.mark.parametrize("indirect", [False, True])
def test_refit_nodes_and_queries(ctx, indirect):
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 300, 40), indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    pos = _deform(g.positions, 1)
    roots = mesh.refit(pos)
    om = O.OracleMesh(pos, idx, tree.roots, indirect=tree.indirect)
    ref_roots = om.refit()
    for a, b in zip(roots, ref_roots):
        assert np.array_equal(a, b)
    # queries on the refit mesh == queries on a fresh upload of the refit tree, and parity
    fresh = ctx.upload(M.Tree(ref_roots, tree.indirect), pos, idx)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    o, d = _rays(lo, hi, 20000, 2)
    got = mesh.raycast_first(o, d)
    omn = O.OracleMesh(pos, idx, ref_roots, indirect=tree.indirect)
    check_raycast_parity(omn, o, d, got, omn.raycast_first(o, d), "refit rays")
    pts = W.grid_points(lo, hi, 24)
    gotc = mesh.closest_point(pts)
    check_closest_parity(omn, pts, gotc, omn.closest_point(pts), "refit points")
    _same(got, fresh.raycast_first(o, d))
    _same(gotc, fresh.closest_point(pts))
    fresh.free()
    mesh.free()


def test_refit_keeps_unchanged_paths(ctx):
    """An inconsistent internal box above unchanged leaves is kept; a changed leaf recomputes its path."""
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 120, 16))
    tight = O.OracleMesh(g.positions, idx, tree.roots).refit()
    buf = tight[0].copy()
    f = buf.view(np.float32).reshape(-1, 8)
    h = buf.view(np.uint16).reshape(-1, 16)
    k = [i for i in range(f.shape[0]) if h[i, 15] != 0xFFFF][5]
    f[k, 0] -= np.float32(3.0)
    mesh = ctx.upload(M.Tree([buf], None), g.positions, idx)
    same = mesh.refit(g.positions)
    assert np.array_equal(same[0], buf)
    pos = _deform(g.positions, 3, scale=0.05)
    got = mesh.refit(pos)
    assert np.array_equal(got[0], O.OracleMesh(pos, idx, [buf]).refit()[0])
    mesh.free()


def test_refit_groups_and_batch(ctx):
    """Several roots (geometry groups); a batch holding the mesh sees the refit geometry."""
    geo = G.torus_knot(10, 3, 200, 24)
    geo.ensure_index()
    ntri = geo.index.size // 3
    ranges = [(0, ntri // 3), (ntri // 3, ntri - ntri // 3)]
    g, idx, tree = build_mesh(geo, ranges=ranges)
    assert len(tree.roots) == 2
    mesh = ctx.upload(tree, g.positions, idx)
    batch = ctx.batch([mesh])
    pos = _deform(g.positions, 5)
    roots = mesh.refit(pos)
    for a, b in zip(roots, O.OracleMesh(pos, idx, tree.roots).refit()):
        assert np.array_equal(a, b)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    o, d = _rays(lo, hi, 5000, 6)
    ident = np.eye(4, dtype=np.float32).reshape(1, 16)
    res = batch.raycast_first(ident, ident, np.zeros(1, np.uint32), o, d, np.zeros(len(o), np.uint32))
    single = mesh.raycast_first(o, d)
    hit = single["face"] != M.NO_HIT
    assert np.array_equal(res["face"] != M.NO_HIT, hit)
    assert np.all(np.abs(res["distance"][hit] - single["distance"][hit]) <= 1e-5 * np.maximum(single["distance"][hit], 1))
    batch.free()
    mesh.free()


def test_refit_rejects_other_vertex_counts(ctx):
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 60, 8))
    mesh = ctx.upload(tree, g.positions, idx)
    with pytest.raises(M.MbvhError):
        mesh.refit(g.positions.reshape(-1, 3)[:-1])
    mesh.free()


def test_refit_device_form_in_an_optimisation_step(ctx):
    """Forward, backward (vertex gradients), a gradient step on the device positions, refit from
    those device positions: the refit tree equals the oracle's refit of the same positions, and the
    next forward pass matches the oracle on the new geometry and the host refit bit for bit."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 240, 32))
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    o, d = _rays(lo, hi, 8000, 9)
    fwd = mesh.raycast_first(o, d)
    gd = np.ones(len(o), np.float32)
    gv = mesh.raycast_first_bwd(d, fwd, grad_distance=gd, mode=M.GRAD_DETERMINISTIC)["positions"]
    pos_t = torch.from_numpy(g.positions.reshape(-1, 3).copy()).to(dev)
    pos_t -= 0.5 * torch.from_numpy(gv).to(dev)  # descend the summed hit distances
    pos_t *= 3.0  # and grow the mesh: the coordinate bound must follow
    torch.cuda.synchronize()
    mesh.refit_d(pos_t.data_ptr(), pos_t.shape[0])
    ctx.synchronize()
    pos = pos_t.cpu().numpy()
    roots = mesh.copy_roots()
    ref_roots = O.OracleMesh(pos, idx, tree.roots).refit()
    assert all(np.array_equal(a, b) for a, b in zip(roots, ref_roots))
    o2, d2 = _rays(pos.min(axis=0), pos.max(axis=0), 8000, 10)
    got = mesh.raycast_first(o2, d2)
    om = O.OracleMesh(pos, idx, ref_roots)
    check_raycast_parity(om, o2, d2, got, om.raycast_first(o2, d2), "refit_d rays")
    host = ctx.upload(tree, g.positions, idx)
    host.refit(pos)
    _same(got, host.raycast_first(o2, d2))
    host.free()
    mesh.free()


# ---- partial refit (refit(nodeIndices), refit.template.js:136-172) ----------------------------

def _partial_cases(total, rng):
    return {
        "sparse": rng.choice(total, max(1, total // 20), replace=False),
        "dense": rng.choice(total, total // 2, replace=False),
        "empty": np.zeros(0, np.int64),
        "all": np.arange(total),
        "foreign": np.array([total, total + 7, 2**32 - 1, 3]),  # ids past the tree name nothing
        "repeated": np.array([1, 1, 2, 2, 2]),
    }


@pytest.mark.parametrize("kind", ["plain", "indirect", "two_roots"])
def test_partial_refit_matches_oracle(ctx, kind):
    """Device partial refits of random node sets equal the oracle's byte for byte."""
    geo = G.torus_knot(10, 3, 200, 24)
    ranges = None
    if kind == "two_roots":
        geo.ensure_index()
        ntri = geo.index.size // 3
        ranges = [(0, ntri // 3), (ntri // 3, ntri - ntri // 3)]
    g, idx, tree = build_mesh(geo, ranges=ranges, indirect=kind == "indirect")
    total = sum(r.nbytes // 32 for r in tree.roots)
    rng = np.random.default_rng(21)
    for name, ids in _partial_cases(total, rng).items():
        mesh = ctx.upload(tree, g.positions, idx)
        pos = _deform(g.positions, 22)
        got = mesh.refit(pos, node_ids=ids)
        ref = O.OracleMesh(pos, idx, tree.roots, indirect=tree.indirect).refit(ids)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), name
        # the query layouts carry the same boxes as a fresh upload of those nodes
        fresh = ctx.upload(M.Tree(ref, tree.indirect), pos, idx)
        o, d = _rays(pos.min(axis=0), pos.max(axis=0), 3000, 23)
        _same(mesh.raycast_first(o, d), fresh.raycast_first(o, d))
        pts = W.grid_points(pos.min(axis=0), pos.max(axis=0), 12)
        _same(mesh.closest_point(pts), fresh.closest_point(pts))
        fresh.free()
        mesh.free()


def _leaves_and_parents(roots):
    """Per root: (base id, parent of each node (-1 at the root), leaf (node, offset, count) list)."""
    out, base = [], 0
    for r in roots:
        f = r.view(np.float32).reshape(-1, 8)
        u = r.view(np.uint32).reshape(-1, 8)
        h = r.view(np.uint16).reshape(-1, 16)
        n = f.shape[0]
        parent = np.full(n, -1, np.int64)
        leaves = []
        for i in range(n):
            if h[i, 15] == 0xFFFF:
                leaves.append((i, int(u[i, 6]), int(h[i, 14])))
            else:
                parent[i + 1] = i
                parent[i + int(u[i, 6])] = i
        out.append((base, parent, leaves))
        base += n
    return out


def test_partial_refit_of_changed_paths_is_a_full_refit(ctx):
    """Moving a few vertices and listing the leaves that use them with their ancestors (what a
    shapecast over the edited region reports) gives the full refit exactly; queries then match the
    FP64 oracle on the edited geometry."""
    geo = G.torus_knot(10, 3, 300, 40)
    geo.ensure_index()
    ntri = geo.index.size // 3
    g, idx, built = build_mesh(geo, ranges=[(0, ntri // 2), (ntri // 2, ntri - ntri // 2)])
    # tight boxes first (the build's are rounded outwards, so a full refit would rewrite every leaf)
    tree = M.Tree(O.OracleMesh(g.positions, idx, built.roots).refit(), None)
    mesh = ctx.upload(tree, g.positions, idx)
    rng = np.random.default_rng(31)
    pos = g.positions.reshape(-1, 3).copy()
    moved = rng.choice(len(pos), 6, replace=False)
    pos[moved] += rng.normal(0, 0.4, (len(moved), 3)).astype(np.float32)
    tri = idx.reshape(-1, 3)
    touched = np.isin(tri, moved).any(axis=1)
    ids = set()
    for base, parent, leaves in _leaves_and_parents(tree.roots):
        for node, off, cnt in leaves:
            if touched[off:off + cnt].any():
                j = node
                while j >= 0:
                    ids.add(base + j)
                    j = parent[j]
    full = O.OracleMesh(pos, idx, tree.roots).refit()
    part = O.OracleMesh(pos, idx, tree.roots).refit(ids)
    assert all(np.array_equal(a, b) for a, b in zip(part, full))
    got = mesh.refit(pos, node_ids=sorted(ids))
    assert all(np.array_equal(a, b) for a, b in zip(got, full))
    om = O.OracleMesh(pos, idx, full)
    o, d = _rays(pos.min(axis=0), pos.max(axis=0), 10000, 32)
    check_raycast_parity(om, o, d, mesh.raycast_first(o, d), om.raycast_first(o, d), "partial refit rays")
    pts = W.grid_points(pos.min(axis=0), pos.max(axis=0), 20)
    check_closest_parity(om, pts, mesh.closest_point(pts), om.closest_point(pts), "partial refit points")
    mesh.free()


def test_partial_refit_device_form(ctx):
    """Device positions and device node ids (mbvh_mesh_refit_nodes_d) == the host form."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 200, 24))
    total = tree.roots[0].nbytes // 32
    ids = np.random.default_rng(41).choice(total, total // 10, replace=False).astype(np.uint32)
    pos = _deform(g.positions, 42)
    a = ctx.upload(tree, g.positions, idx)
    b = ctx.upload(tree, g.positions, idx)
    pos_t = torch.from_numpy(pos).to(dev)
    ids_t = torch.from_numpy(ids.astype(np.int32)).to(dev)
    torch.cuda.synchronize()
    a.refit_nodes_d(pos_t.data_ptr(), pos_t.shape[0], ids_t.data_ptr(), ids.size)
    ctx.synchronize()
    want = b.refit(pos, node_ids=ids)
    assert all(np.array_equal(x, y) for x, y in zip(a.copy_roots(), want))
    assert all(np.array_equal(x, y) for x, y in zip(want, O.OracleMesh(pos, idx, tree.roots).refit(ids)))
    a.free()
    b.free()
