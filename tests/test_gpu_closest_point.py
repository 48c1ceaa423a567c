"""GPU parity of closestPointToPoint against the FP64 oracle (reference traversal), via the C ABI."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, check_closest_parity, point_faces_tied, threshold_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _run(ctx, geometry, points, min_t=0.0, max_t=float("inf"), label="", ranges=None, indirect=False):
    g, idx, tree = build_mesh(geometry, ranges=ranges, indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    got = mesh.closest_point(points, min_t, max_t)
    om = O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)
    ref = om.closest_point(points, min_t, max_t)
    stats = check_closest_parity(om, points, got, ref, label)
    print(label, stats, flush=True)
    mesh.free()
    return got, ref, stats, om


def test_c2_grid(ctx):
    """Config C2 geometry (TorusKnot(10,3,800,200)) on a 64^3 grid (full parity)."""
    w = W.c2_closest(res=64)
    got, ref, stats, om = _run(ctx, w.geometry, w.points, label="C2_64")
    # barycoords of the closest point reconstruct it
    assert np.all(np.abs(got["barycoord"].sum(axis=1) - 1) < 1e-5)


def test_sphere_kats(ctx):
    """ShapeCasts.test.js:715-752 restated on SphereGeometry(1, 200, 200)."""
    g = G.sphere(1, 200, 200)
    g2, idx, tree = build_mesh(g, strategy=M.CENTER)
    mesh = ctx.upload(tree, g2.positions, idx)
    r = mesh.closest_point(np.array([[0, 0, 0], [0, 1, 0]], np.float32))
    assert 1 - 1e-3 <= r["distance"][0] <= 1
    assert r["distance"][1] == 0
    rng = np.random.default_rng(3)
    v = rng.random((100, 3)) - 0.5
    length = rng.random(100) * 3
    v = (v / np.linalg.norm(v, axis=1, keepdims=True) * length[:, None]).astype(np.float32)
    r = mesh.closest_point(v)
    expected = np.abs(1 - np.linalg.norm(v.astype(np.float64), axis=1))
    assert np.all(np.abs(r["distance"] - expected) <= 1e-3)
    mesh.free()


def _faces_equal_up_to_rounding(om, pts, got, ref):
    """Indices where the faces differ and are not rounding ties (point_faces_tied)."""
    diff = np.nonzero(got["face"] != ref["face"])[0]
    return [int(i) for i in diff if not point_faces_tied(om, pts[i], got["face"][i], ref["face"][i])]


def test_thresholds(ctx):
    g = G.torus_knot(10, 3, 120, 30)
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 24)
    # maxThreshold prunes boxes; minThreshold stops at the first triangle under it
    got, ref, _, om = _run(ctx, g, pts, max_t=2.0, label="maxT")
    g2, idx, tree = build_mesh(g)
    mesh = ctx.upload(tree, g2.positions, idx)
    om2 = O.OracleMesh(g2.positions, idx, tree.roots)
    closest = om2.closest_point(pts)
    # with a minThreshold the result is the first face in the reference's visit order under it, or the
    # closest: the one-lane kernel follows that order exactly (children by exact box distance, the
    # threshold decided exactly), so the faces are the oracle's up to rounding ties, and exactly where
    # the early exit returned a face that is not the closest
    for mt in (1.0, 0.3, 2.5):
        got = mesh.closest_point(pts, mt, float("inf"))
        ref = om2.closest_point(pts, mt, float("inf"))
        under = ref["distance"] < mt
        early = under & (ref["distance"] > closest["distance"] * (1 + 1e-9))
        print(f"minThreshold {mt}: {under.mean():.3f} under it, {early.sum()} early exits before the closest",
              flush=True)
        assert under.mean() > 0.05 and early.sum() > 50
        assert np.array_equal(got["face"][early], ref["face"][early]), np.count_nonzero(got["face"][early] != ref["face"][early])
        bad = _faces_equal_up_to_rounding(om2, pts, got, ref)
        assert not bad, (mt, len(bad), bad[:5])
        assert np.all(got["distance"][under] < mt)
        assert np.all(np.abs(got["distance"] - ref["distance"]) <= 1e-5 * np.maximum(ref["distance"], 1))
    mesh.free()


@pytest.mark.parametrize("lanes", [1, 4])
def test_threshold_branches_by_hand(lanes):
    """The hand-built two-leaf mesh of the oracle's threshold test: each (minThreshold,
    maxThreshold) case gives the face and distance worked out by hand, float and double queries."""
    geo, cases = threshold_cases()
    g, idx, tree = build_mesh(geo, strategy=M.CENTER)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    c = M.Context(0)
    c.set_point_lanes(lanes)
    mesh = c.upload(tree, g.positions, idx)
    try:
        for dt in (np.float32, np.float64):
            p = np.zeros((64, 3), dt)
            for mn, mx, want in cases:
                got = mesh.closest_point(p, mn, mx)
                ref = om.closest_point(p[:1], mn, mx)
                assert np.all(got["face"] == ref["face"][0]), (dt, mn, mx, got["face"][0], ref["face"][0])
                if want is None:
                    assert np.all(np.isinf(got["distance"]))
                else:
                    assert np.all(got["distance"] == np.float32(want)), (dt, mn, mx, got["distance"][0])
    finally:
        mesh.free()
        c.close()


def test_groups_and_indirect(ctx):
    g = G.torus(1, 1, 40, 10)
    ntri = g.tri_count
    rng = np.random.default_rng(4)
    pts = (rng.random((4096, 3)) * 6 - 3).astype(np.float32)
    _run(ctx, g, pts, ranges=[(k * ntri // 5, ntri // 5) for k in range(5)], label="groups")
    _run(ctx, g, pts, ranges=[(0, ntri // 3), (ntri // 2, ntri // 4)], indirect=True, label="indirect")


def test_c2_full_grid_properties(ctx):
    """Full C2 grid (128^3): oracle parity on a strided sample, size-independent checks on all."""
    w = W.c2_closest(res=128)
    g, idx, tree = build_mesh(w.geometry)
    mesh = ctx.upload(tree, g.positions, idx)
    got = mesh.closest_point(w.points)
    assert np.all(got["face"] != M.NO_HIT)
    # distance equals |q - point| and the point lies in the reported face's plane span
    d = np.linalg.norm(w.points.astype(np.float64) - got["point"], axis=1)
    assert np.max(np.abs(d - got["distance"]) / np.maximum(got["distance"], 1)) < 2e-6
    om = O.OracleMesh(g.positions, idx, tree.roots)
    sample = np.arange(0, w.points.shape[0], 97)
    ref = om.closest_point(w.points[sample])
    sub = {k: v[sample] for k, v in got.items()}
    stats = check_closest_parity(om, w.points[sample], sub, ref, "C2_128_sample")
    print("C2_128_sample", stats, flush=True)
    mesh.free()


def test_degenerate_first_candidate(ctx):
    """A zero-area triangle (two coincident vertices) whose FP32 closest-point formula yields NaN is
    tested first; it must be ignored (reference: `distSq < best` is false for NaN) and the valid
    triangle reported. Regression test for the exact-comparison branch with no best yet."""
    a = [0.0, 0.0, 1.2246468525851679e-16]
    c = [0.009960341267287731, 0.007236612029373646, -0.15643446147441864]
    pos = np.array([a, a, c, [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    geo = G.Geometry(pos, np.array([0, 1, 2, 3, 4, 5], np.uint16))
    q = np.array([[-2.514983892440796, 0.6441349983215332, -0.7410805225372314]], np.float32)
    got, ref, stats, om = _run(ctx, geo, np.repeat(q, 64, axis=0), label="degenerate")
    assert np.all(got["face"] == ref["face"])


def _morton_buckets(pts):
    """Bucket of each point as the ordering pass computes it: cells of the finite bounds, 1..7 bits
    per axis (about 8 points per bucket), a NaN coordinate in cell 0 and an infinite one clamped to the
    end cell, bits interleaved x, y, z."""
    n = pts.shape[0]
    bits = 1
    while bits < 7 and (8 << (3 * bits)) < n:
        bits += 1
    cells = 1 << bits
    key = np.zeros(n, np.uint64)
    for k in range(3):
        v = pts[:, k]
        fin = v[np.isfinite(v)]
        lo, hi = np.float32(fin.min()), np.float32(fin.max())
        ext = np.float32(hi - lo)
        with np.errstate(invalid="ignore", divide="ignore"):
            t = (v - lo) / ext if ext > 0 else np.zeros(n, np.float32)
        t = np.where(np.isnan(t), np.float32(0), np.clip(t, 0, 1)).astype(np.float32)
        c = np.minimum((t * np.float32(cells)).astype(np.uint64), cells - 1)
        for b in range(bits):
            key |= ((c >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + k)
    return key


def test_query_order_changes_nothing(ctx):
    """The spatial visit order is a permutation, and results with it equal results without it bit for
    bit (each query's result depends on that query only)."""
    torch = pytest.importorskip("torch")
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 200, 40))
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 40)  # 64,000 points in row order
    rng = np.random.default_rng(3)
    pts = np.concatenate([pts, rng.standard_normal((1000, 3)).astype(np.float32) * 50,
                          np.array([[np.nan, 0, 0], [np.inf, 1, 1]], np.float32)])
    n = pts.shape[0]
    dev = torch.device("cuda:0")
    d_q = torch.from_numpy(pts).to(dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    wsb = M.Context.order_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    ctx.order_points_d(d_q.data_ptr(), n, order.data_ptr(), ws.data_ptr(), wsb)
    outs = []
    for optr in (0, order.data_ptr()):
        face = torch.empty(n, dtype=torch.int32, device=dev)
        dist = torch.empty(n, dtype=torch.float32, device=dev)
        pt = torch.empty((n, 3), dtype=torch.float32, device=dev)
        bc = torch.empty((n, 3), dtype=torch.float32, device=dev)
        mesh.closest_point_d(d_q.data_ptr(), n, face.data_ptr(), dist.data_ptr(), pt.data_ptr(), bc.data_ptr(),
                             order_ptr=optr)
        ctx.synchronize()
        outs.append([x.cpu().numpy().view(np.uint32) for x in (face, dist, pt, bc)])
    o = order.cpu().numpy().view(np.uint32)
    assert np.array_equal(np.sort(o), np.arange(n, dtype=np.uint32))
    assert not np.array_equal(o, np.arange(n, dtype=np.uint32))
    # the visit order is sorted by Morton bucket (same FP32 cell arithmetic on the host)
    keys = _morton_buckets(pts)
    assert np.all(np.diff(keys[o].astype(np.int64)) >= 0)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    # the host entry point (which orders internally at this size) gives the same results
    host = mesh.closest_point(pts)
    assert np.array_equal(host["face"], outs[0][0])
    assert np.array_equal(host["distance"].view(np.uint32), outs[0][1])
    mesh.free()


@pytest.mark.parametrize("lanes", [1, 4, 64])
def test_lanes_per_query(lanes):
    """One lane per query (binary tree), four lanes per query (4-wide tree) and packets of 64 queries
    walking the 4-wide tree together all meet the parity rule on the C2 mesh; with a minThreshold the
    call keeps one lane (the reference's early exit follows its own visit order)."""
    c = M.Context(0)
    c.set_point_lanes(lanes)
    try:
        w = W.c2_closest(res=48)
        _run(c, w.geometry, w.points, label=f"C2_48 lanes={lanes}")
        lo, hi = w.geometry.bounds()
        pts = W.grid_points(lo, hi, 16)
        _run(c, w.geometry, pts, min_t=0.5, label=f"minThreshold lanes={lanes}")
    finally:
        c.close()


def test_packets_equal_quads_bit_for_bit():
    """Packets of 64 queries return exactly the quad kernel's results (faces, distances, points,
    barycoords, bit for bit): grid and random points, double-precision points, a finite maxThreshold,
    several roots (groups), an indirect mesh, a ragged last packet, NaN and far points."""
    g = G.torus_knot(10, 3, 200, 40)
    lo, hi = g.bounds()
    rng = np.random.default_rng(11)
    grid = W.grid_points(lo, hi, 30)  # 27,000 points: the last packet is ragged
    rand = (lo + (hi - lo) * rng.random((20001, 3))).astype(np.float32)
    rand[::997] = np.nan
    rand[5] = [1e6, -1e6, 3e5]
    dbl = rand.astype(np.float64) + rng.standard_normal(rand.shape) * 1e-9
    ntri = g.tri_count
    meshes = [("plain", dict()), ("groups", dict(ranges=[(k * ntri // 3, ntri // 3) for k in range(3)])),
              ("indirect", dict(ranges=[(0, ntri // 2), (ntri // 2, ntri - ntri // 2)], indirect=True))]
    for name, kw in meshes:
        gg, idx, tree = build_mesh(g, **kw)
        out = {}
        for lanes in (4, 64):
            c = M.Context(0)
            c.set_point_lanes(lanes)
            mesh = c.upload(tree, gg.positions, idx)
            out[lanes] = [mesh.closest_point(grid), mesh.closest_point(rand), mesh.closest_point(dbl),
                          mesh.closest_point(rand, 0.0, 2.0)]
            mesh.free()
            c.close()
        for k, (a4, a64) in enumerate(zip(out[4], out[64])):
            for key in ("face", "distance", "point", "barycoord"):
                assert np.array_equal(a4[key].view(np.uint32) if a4[key].dtype != np.uint32 else a4[key],
                                      a64[key].view(np.uint32) if a64[key].dtype != np.uint32 else a64[key]), (name, k, key)
        print(name, "packets == quads on", [len(x["face"]) for x in out[64]], flush=True)


def test_packets_redo_overflowing_lists():
    """A cone of 16 triangles around one apex, queried from above: every triangle's closest point is
    the apex, so all 16 tie and overflow a packet's candidate list (8 entries); those queries are
    finished by the quad kernel and must equal its results and the oracle's face (the reference's
    first in visit order)."""
    k = 16
    ang = np.arange(k) * 2 * np.pi / k
    ring = np.stack([np.cos(ang), np.sin(ang), np.full(k, -0.5)], axis=1)
    pos = np.concatenate([np.zeros((1, 3)), ring]).astype(np.float32)
    tri = np.stack([np.zeros(k, int), 1 + np.arange(k), 1 + (np.arange(k) + 1) % k], axis=1)
    geo = G.Geometry(pos, tri.reshape(-1).astype(np.uint16))
    rng = np.random.default_rng(2)
    above = np.concatenate([rng.uniform(-0.05, 0.05, (5000, 2)), rng.uniform(0.5, 3.0, (5000, 1))], axis=1)
    pts = np.concatenate([above, rng.uniform(-2, 2, (3000, 3))]).astype(np.float32)
    g2, idx, tree = build_mesh(geo)
    om = O.OracleMesh(g2.positions, idx, tree.roots)
    ref = om.closest_point(pts)
    out = {}
    for lanes in (4, 64):
        c = M.Context(0)
        c.set_point_lanes(lanes)
        mesh = c.upload(tree, g2.positions, idx)
        out[lanes] = mesh.closest_point(pts)
        mesh.free()
        c.close()
    for key in ("face", "distance", "point", "barycoord"):
        assert np.array_equal(out[4][key], out[64][key]), key
    apex = np.all(out[64]["point"][:5000] == 0, axis=1)
    assert apex.mean() > 0.9
    assert np.array_equal(out[64]["face"][:5000][apex], ref["face"][:5000][apex])
