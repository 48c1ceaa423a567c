"""GPU parity of raycast (all hits) against the oracle's restatement of the reference traversal, via the
C ABI: same hits, same (left-first traversal) order, distances within tolerance."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_py as O
from helpers import build_mesh, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _compare(om, o, d, got, side, near=0.0, far=float("inf")):
    offs = got["offsets"].astype(np.int64)
    total = 0
    for i in range(o.shape[0]):
        rf, rd = om.raycast_all(o[i], d[i], side=side, near=near, far=far)
        gf = got["face"][offs[i]:offs[i + 1]]
        assert np.array_equal(gf, rf), (i, gf, rf)
        gd = got["distance"][offs[i]:offs[i + 1]]
        assert np.all(rel_err(gd, rd) <= 1e-5)
        total += len(rf)
    return total


@pytest.mark.parametrize("side", [M.FRONT_SIDE, M.DOUBLE_SIDE])
def test_torus_knot_all_hits(ctx, side):
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 120, 30))
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(2048, 40.0, lo, hi, seed_origin=21, seed_target=22)
    got = mesh.raycast(o, d, side=side)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    total = _compare(om, o, d, got, side)
    assert total > 1000
    # the nearest hit of each ray equals raycastFirst
    first = mesh.raycast_first(o, d, side=side)
    offs = got["offsets"].astype(np.int64)
    for i in range(o.shape[0]):
        seg = got["distance"][offs[i]:offs[i + 1]]
        if len(seg):
            assert seg.min() == first["distance"][i]
        else:
            assert first["face"][i] == M.NO_HIT
    mesh.free()


def test_torus_kat_and_near_far(ctx):
    """ShapeCasts.test.js:906-929: TorusGeometry(5,5,40,10), ray (0,0,-10) -> +z hits 10 triangles."""
    g, idx, tree = build_mesh(G.torus(5, 5, 40, 10), strategy=M.CENTER)
    mesh = ctx.upload(tree, g.positions, idx)
    got = mesh.raycast(np.array([[0, 0, -10]], np.float32), np.array([[0, 0, 1]], np.float32))
    assert int(got["offsets"][1]) == 10
    om = O.OracleMesh(g.positions, idx, tree.roots)
    rng = np.random.default_rng(9)
    o = (rng.random((1024, 3)) * 20 - 10).astype(np.float32)
    d = (-o / np.linalg.norm(o, axis=1, keepdims=True)).astype(np.float32)
    got = mesh.raycast(o, d, side=M.DOUBLE_SIDE, near=3.0, far=12.0)
    _compare(om, o, d, got, O.DOUBLE, near=3.0, far=12.0)
    mesh.free()


def test_device_two_pass_matches_host_form(ctx):
    torch = pytest.importorskip("torch")
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 64, 16))
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(1000, 40.0, lo, hi, seed_origin=23, seed_target=24)
    host = mesh.raycast(o, d, side=M.DOUBLE_SIDE)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    counts = torch.zeros(1000, dtype=torch.int32, device=dev)
    mesh.raycast_count_d(to.data_ptr(), td.data_ptr(), 1000, counts.data_ptr(), side=M.DOUBLE_SIDE)
    torch.cuda.synchronize()
    offs = torch.zeros(1001, dtype=torch.int64, device=dev)
    offs[1:] = torch.cumsum(counts.to(torch.int64), 0)
    total = int(offs[-1].item())
    face = torch.empty(total, dtype=torch.int32, device=dev)
    dist = torch.empty(total, dtype=torch.float32, device=dev)
    mesh.raycast_fill_d(to.data_ptr(), td.data_ptr(), 1000, offs.data_ptr(), total, face.data_ptr(), dist.data_ptr(),
                        side=M.DOUBLE_SIDE)
    torch.cuda.synchronize()
    assert np.array_equal(offs.cpu().numpy().astype(np.uint64), host["offsets"])
    assert np.array_equal(face.cpu().numpy().view(np.uint32), host["face"])
    assert np.array_equal(dist.cpu().numpy(), host["distance"])
    mesh.free()
