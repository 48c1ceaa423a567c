"""GPU parity of raycastFirst against the FP64 oracle (reference traversal), through the C ABI."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, check_raycast_parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _run(ctx, geometry, origins, dirs, side=M.FRONT_SIDE, near=0.0, far=float("inf"), label="",
         strategy=M.SAH, ranges=None, indirect=False):
    g, idx, tree = build_mesh(geometry, strategy=strategy, ranges=ranges, indirect=indirect)
    mesh = ctx.upload(tree, g.positions, idx)
    got = mesh.raycast_first(origins, dirs, side=side, near=near, far=far)
    om = O.OracleMesh(g.positions, idx, tree.roots, indirect=tree.indirect)
    ref = om.raycast_first(origins, dirs, side=side, near=near, far=far)
    stats = check_raycast_parity(om, origins, dirs, got, ref, label)
    print(label, stats, flush=True)
    mesh.free()
    return got, ref, stats


def test_c1_full_size(ctx):
    """Config C1 at full size: TorusKnot(10,3,400,100), 65,536 rays."""
    w = W.c1_raycast()
    got, ref, stats = _run(ctx, w.geometry, w.origins, w.directions, label="C1")
    assert stats["hits"] > 40000
    # every reported hit lies on the ray: point = origin + distance * dir (unit dirs)
    hit = got["face"] != M.NO_HIT
    recon = w.origins[hit] + got["distance"][hit, None] * w.directions[hit]
    assert np.max(np.abs(recon - got["point"][hit])) < 1e-4
    # barycoords sum to one
    assert np.max(np.abs(got["barycoord"][hit].sum(axis=1) - 1)) < 1e-5


@pytest.mark.parametrize("side", [M.FRONT_SIDE, M.BACK_SIDE, M.DOUBLE_SIDE])
def test_sides(ctx, side):
    g = G.torus_knot(10, 3, 120, 30)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(8192, 40.0, lo, hi, seed_origin=11, seed_target=12)
    _run(ctx, g, o, d, side=side, label=f"side{side}")


def test_near_far(ctx):
    g = G.torus_knot(10, 3, 120, 30)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(8192, 40.0, lo, hi, seed_origin=13, seed_target=14)
    _run(ctx, g, o, d, near=30.0, far=45.0, side=M.DOUBLE_SIDE, label="nearfar")


@pytest.mark.parametrize("strategy", [M.CENTER, M.AVERAGE, M.SAH])
def test_strategies_inside_origins(ctx, strategy):
    """Rays starting inside the mesh's bounds, in random directions (reference test style)."""
    g = G.torus(5, 5, 40, 10)
    rng = np.random.default_rng(5)
    o = (rng.random((4096, 3)) * 10).astype(np.float32)
    d = -o / np.linalg.norm(o, axis=1, keepdims=True)
    _run(ctx, g, o, d.astype(np.float32), side=M.DOUBLE_SIDE, strategy=strategy, label=f"strategy{strategy}")


def test_axis_aligned_rays(ctx):
    """Rays with zero direction components take the literal split-plane path."""
    g = G.torus(5, 5, 40, 10)
    rng = np.random.default_rng(6)
    n = 4096
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    for i in range(n):
        a = axis[i]
        o[i] = (rng.random(3) * 12 - 6).astype(np.float32)
        o[i, a] = -20.0 * sign[i]
        d[i, a] = sign[i]
    # a few rays exactly on coordinate planes through the origin
    o[:64, 0] = 0.0
    o[64:128, 1] = 0.0
    _run(ctx, g, o, d, side=M.DOUBLE_SIDE, label="axis")
    # the reference Raycaster test: origin (0,0,-10), dir (0,0,1) on TorusGeometry(5,5,40,10)
    _run(ctx, g, np.array([[0, 0, -10]], np.float32), np.array([[0, 0, 1]], np.float32), label="torus_axis")


def test_groups_multi_root(ctx):
    """Ten geometry groups → ten roots; closest hit across roots (strict <)."""
    g = G.torus(1, 1, 40, 10)
    ntri = g.tri_count
    ranges = [(k * ntri // 10, ntri // 10) for k in range(10)]
    rng = np.random.default_rng(7)
    o = (rng.random((4096, 3)) * 10).astype(np.float32)
    d = (-o / np.linalg.norm(o, axis=1, keepdims=True)).astype(np.float32)
    _run(ctx, g, o, d, ranges=ranges, side=M.DOUBLE_SIDE, label="groups")


def test_indirect(ctx):
    g = G.torus(1, 1, 40, 10)
    ntri = g.tri_count
    ranges = [(0, ntri // 3), (ntri // 2, ntri // 4)]  # a gap between the ranges
    rng = np.random.default_rng(8)
    o = (rng.random((4096, 3)) * 10).astype(np.float32)
    d = (-o / np.linalg.norm(o, axis=1, keepdims=True)).astype(np.float32)
    _run(ctx, g, o, d, ranges=ranges, indirect=True, side=M.DOUBLE_SIDE, label="indirect")
    # non-indexed indirect build
    g2 = G.icosahedron(1, 6)
    _run(ctx, g2, o, d, indirect=True, label="indirect_nonindexed")


def test_tiny_meshes(ctx):
    """Root is a leaf (<= 10 triangles): no node is bounds-tested."""
    rng = np.random.default_rng(9)
    o = (rng.random((2048, 3)) * 6 - 3).astype(np.float32)
    d = (-o / np.linalg.norm(o, axis=1, keepdims=True)).astype(np.float32)
    for geo in (G.box(), G.icosahedron(1, 0)):
        _run(ctx, geo, o, d, side=M.DOUBLE_SIDE, label="tiny")


def test_box_group_sides_kat(ctx):
    """ShapeCasts.test.js:931-979 restated: first hit through a box with per-face sides."""
    g = G.box()
    g2, idx, tree = build_mesh(g)
    mesh = ctx.upload(tree, g2.positions, idx)
    # materials [F, F, F, F, B, B] by group → per-face facing masks (1 front, 2 back)
    sides = [1, 1, 1, 1, 2, 2]
    face_mask = np.zeros(g2.tri_count, np.uint8)
    for (start, count, mat) in g2.groups:
        face_mask[start // 3:(start + count) // 3] = sides[mat]
    mesh.set_side_mask(face_mask)
    got = mesh.raycast_first(np.array([[0, 0.25, -10]], np.float32), np.array([[0, 0, 1]], np.float32),
                             side=M.SIDE_PER_TRIANGLE)
    assert got["face"][0] != M.NO_HIT
    assert got["point"][0, 2] == 0.5
    om = O.OracleMesh(g2.positions, idx, tree.roots, face_mask=face_mask)
    ref = om.raycast_first(np.array([[0, 0.25, -10]], np.float32), np.array([[0, 0, 1]], np.float32), side=O.PER_FACE)
    assert ref["face"][0] == got["face"][0]
    mesh.free()


def test_c3_sample(ctx):
    """Config C3 geometry (noise-displaced Icosahedron detail 200, ~808k tris), 131,072 rays."""
    w = W.c3_raycast(n_rays=1 << 17)
    _run(ctx, w.geometry, w.origins, w.directions, label="C3sample")


def test_ray_order_changes_nothing(ctx):
    """Rays traced in the spatial order of their origins give bit-identical results."""
    torch = pytest.importorskip("torch")
    w = W.c1_raycast(20000)
    g, idx, tree = build_mesh(w.geometry)
    mesh = ctx.upload(tree, g.positions, idx)
    n = w.origins.shape[0]
    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(w.origins).to(dev)
    d_d = torch.from_numpy(w.directions).to(dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    wsb = M.Context.order_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    ctx.order_points_d(d_o.data_ptr(), n, order.data_ptr(), ws.data_ptr(), wsb)
    outs = []
    for optr in (0, order.data_ptr()):
        face = torch.empty(n, dtype=torch.int32, device=dev)
        dist = torch.empty(n, dtype=torch.float32, device=dev)
        mesh.raycast_first_d(d_o.data_ptr(), d_d.data_ptr(), n, face.data_ptr(), dist.data_ptr(), order_ptr=optr)
        ctx.synchronize()
        outs.append((face.cpu().numpy().view(np.uint32), dist.cpu().numpy().view(np.uint32)))
    o = order.cpu().numpy().view(np.uint32)
    assert np.array_equal(np.sort(o), np.arange(n, dtype=np.uint32))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    mesh.free()


@pytest.mark.parametrize("lanes", [1, 4])
def test_lanes_per_ray(lanes):
    """One lane per ray on the binary tree and four lanes per ray on the 4-wide tree both meet the
    parity rule on C1 and on rays with zero direction components (which four-lane calls trace on
    the binary tree)."""
    c = M.Context(0)
    c.set_ray_lanes(lanes)
    try:
        w = W.c1_raycast()
        _run(c, w.geometry, w.origins, w.directions, label=f"C1 lanes={lanes}")
        g = G.torus_knot(10, 3, 120, 30)
        lo, hi = g.bounds()
        o, d = W.rays_toward_box(4096, 40.0, lo, hi, seed_origin=41, seed_target=42)
        d[::2, 1] = 0.0  # axis-parallel in y
        d[1::4, 2] = 0.0
        _run(c, g, o, d, label=f"zero components lanes={lanes}")
    finally:
        c.close()
