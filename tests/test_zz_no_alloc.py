"""The allocation-free closest point (mbvh_closest_point_noalloc_d) against mbvh_closest_point_d, through
the C ABI: the same results bit for bit, on two calls in a row (the second smaller), for the four-lane
path (no minThreshold) and the one-lane path (minThreshold > 0, traversal stack overflow in the fixed
scratch, a launch narrowed to fit it), and the documented size limit."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
from helpers import build_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    ctx = M.Context(0)
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 800, 200))
    mesh = ctx.upload(tree, g.positions, idx)
    yield torch, ctx, mesh, g
    mesh.free()
    ctx.close()


def _outputs(torch, n):
    dev = torch.device("cuda:0")
    return (torch.full((n,), 7, dtype=torch.int32, device=dev), torch.full((n,), 7.0, dtype=torch.float32, device=dev),
            torch.full((n, 3), 7.0, dtype=torch.float32, device=dev),
            torch.full((n, 3), 7.0, dtype=torch.float32, device=dev))


def _both(torch, ctx, mesh, pts, min_t=0.0, max_t=float("inf")):
    n = pts.shape[0]
    d_q = torch.from_numpy(pts).to("cuda:0")
    a = _outputs(torch, n)
    mesh.closest_point_noalloc_d(d_q.data_ptr(), n, *(x.data_ptr() for x in a), min_threshold=min_t,
                                 max_threshold=max_t)
    ctx.synchronize()
    got = [x.cpu().numpy().view(np.uint32) for x in a]
    b = _outputs(torch, n)
    mesh.closest_point_d(d_q.data_ptr(), n, *(x.data_ptr() for x in b), min_threshold=min_t, max_threshold=max_t)
    ctx.synchronize()
    ref = [x.cpu().numpy().view(np.uint32) for x in b]
    return got, ref


def _assert_same(got, ref, label):
    for name, a, b in zip(("face", "distance", "point", "barycoord"), got, ref):
        assert np.array_equal(a, b), f"{label}: {name} differs in {(a != b).sum()} entries"


@pytest.mark.parametrize("min_t", [0.0, 0.02])
def test_noalloc_equals_closest_point_d_on_two_calls(env, min_t):
    torch, ctx, mesh, g = env
    lo, hi = g.bounds()
    first = W.grid_points(lo, hi, 64 if min_t == 0.0 else 56)  # ordered inside (>= 4,096 queries)
    rng = np.random.default_rng(11)
    second = (lo + (hi - lo) * rng.random((3000, 3))).astype(np.float32)  # smaller, unordered
    for label, pts in (("first", first), ("second", second)):
        print(f"noalloc min_t={min_t} {label} call: {pts.shape[0]} points", flush=True)
        got, ref = _both(torch, ctx, mesh, pts, min_t=min_t)
        _assert_same(got, ref, f"min_t={min_t} {label}")
        assert np.all(got[0] != M.NO_HIT - 1), "stack overflow reported"
        assert np.mean(got[0] != M.NO_HIT) > 0.5


def test_noalloc_narrowed_launch_on_a_deep_tree(env):
    """A CENTER tree with one triangle per leaf is 23 levels deep: the one-lane path needs 8 overflow
    entries (64 B) per worker beyond its 16 LDS entries, and 300k workers would need more than the
    whole 32 MiB region less its 16 MiB of order buffers, so the call runs on fewer workers."""
    torch, ctx, _, g0 = env
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 800, 200), strategy=M.CENTER, leaf=1)
    mesh = ctx.upload(tree, g.positions, idx)
    depth = mesh.info()["max_depth"]
    lo, hi = g.bounds()
    first = W.grid_points(lo, hi, 67)
    need = depth + 1 - 16
    assert need * 8 * first.shape[0] > (32 - 16) << 20, "the case must not fit a full-width launch"
    rng = np.random.default_rng(12)
    second = (lo + (hi - lo) * rng.random((5000, 3))).astype(np.float32)
    for label, pts in (("first", first), ("second", second)):
        print(f"noalloc deep tree (depth {depth}) {label} call: {pts.shape[0]} points", flush=True)
        got, ref = _both(torch, ctx, mesh, pts, min_t=0.01)
        _assert_same(got, ref, f"deep {label}")
        assert np.all(got[0] != M.NO_HIT - 1), "stack overflow reported"
    mesh.free()


def test_noalloc_limit_and_empty(env):
    torch, ctx, mesh, g = env
    n = M.NOALLOC_MAX_QUERIES + 1
    d_q = torch.zeros((n, 3), dtype=torch.float32, device="cuda:0")
    out = _outputs(torch, n)
    with pytest.raises(M.MbvhError) as e:
        mesh.closest_point_noalloc_d(d_q.data_ptr(), n, *(x.data_ptr() for x in out))
    assert e.value.code == -6 and "allocation-free maximum" in str(e.value)
    mesh.closest_point_noalloc_d(0, 0, 0)  # an empty call enqueues nothing
    ctx.synchronize()
    assert torch.all(out[0] == 7)
