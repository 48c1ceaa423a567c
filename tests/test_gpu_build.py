"""Device BVH build (mbvh_build_gpu / mbvh_build_gpu_d) through the C ABI: the node buffers, the
permuted index and the indirect buffer equal the host build's (mbvh_build, the sequential
restatement of buildTree.js / splitUtils.js / sortUtils.js) byte for byte, and pass the oracle's
validateBounds. Cases: the raycast and closest-point workload meshes (uint16 and uint32 index),
triangle soups, groups with gaps between the ranges, indirect builds over ranges of indexed and
non-indexed geometry, every strategy, leaf sizes that reach the exhaustive SAH (fewer than 8
primitives), a shallow max depth, degenerate and extreme coordinates, and the error cases."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G
import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _soup(n, seed, spread=10.0, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    p = (c + rng.normal(0, size, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    idx = rng.permutation(3 * n).astype(np.uint32)
    pos = np.empty_like(p)
    pos[idx] = p  # shuffled vertex order; the index restores the triangles
    return G.Geometry(pos, idx)


def _both(ctx, g, **kw):
    """Host and device builds of the same input; returns (host tree, host index, device tree, device index)."""
    idx_h = None if g.index is None else g.index.copy()
    idx_d = None if g.index is None else g.index.copy()
    th = M.build(g.positions, idx_h, **kw)
    td = M.build(g.positions, idx_d, ctx=ctx, **kw)
    return th, idx_h, td, idx_d


def _check_same(g, th, idx_h, td, idx_d, validate=True):
    print(f"  roots {len(th.roots)} nodes {[r.size // 32 for r in th.roots]}", flush=True)
    assert len(th.roots) == len(td.roots)
    for r, (a, b) in enumerate(zip(th.roots, td.roots)):
        assert a.size == b.size, f"root {r}: {a.size // 32} host nodes vs {b.size // 32} device nodes"
        if not np.array_equal(a, b):
            rows = np.nonzero((a.reshape(-1, 32) != b.reshape(-1, 32)).any(axis=1))[0]
            raise AssertionError(f"root {r}: {rows.size} node records differ, first {rows[:5].tolist()}")
    if idx_h is not None:
        assert np.array_equal(idx_h, idx_d), "permuted index differs"
    if th.indirect is None:
        assert td.indirect is None
    else:
        assert td.indirect is not None and td.indirect.dtype == th.indirect.dtype
        assert np.array_equal(th.indirect, td.indirect), "indirect buffer differs"
    if validate:
        om = O.OracleMesh(g.positions, idx_d, td.roots, indirect=td.indirect)
        for r in range(len(td.roots)):
            rc, st = om.validate_root(r)
            assert rc == 0, st


@pytest.mark.parametrize("name,make", [
    ("torusknot80k_u16", lambda: G.torus_knot(10, 3, 400, 100)),
    ("torusknot320k_u32", lambda: G.torus_knot(10, 3, 800, 200)),
    ("soup50k", lambda: _soup(50000, 3)),
    ("icosphere_nonindexed", lambda: G.displaced_icosphere(40).ensure_index()),
])
def test_sah_workloads(ctx, name, make):
    print(f"case {name}", flush=True)
    g = make()
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH)
    _check_same(g, th, ih, td, idd)


@pytest.mark.parametrize("strategy", [M.CENTER, M.AVERAGE, M.SAH])
@pytest.mark.parametrize("leaf", [1, 4, 10, 40])
def test_strategies_and_leaf_sizes(ctx, strategy, leaf):
    print(f"case strategy {strategy} leaf {leaf}", flush=True)
    g = _soup(6000, 11 + leaf)
    th, ih, td, idd = _both(ctx, g, strategy=strategy, target_leaf_size=leaf)
    _check_same(g, th, ih, td, idd)


def test_big_average_node(ctx):
    # AVERAGE's sequential mean over a node handled by a whole workgroup
    print("case average 80k", flush=True)
    g = G.torus_knot(10, 3, 400, 100)
    th, ih, td, idd = _both(ctx, g, strategy=M.AVERAGE)
    _check_same(g, th, ih, td, idd)


@pytest.mark.parametrize("max_depth", [0, 3, 12])
def test_max_depth(ctx, max_depth):
    print(f"case max_depth {max_depth}", flush=True)
    g = _soup(20000, 5)
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH, max_depth=max_depth)
    _check_same(g, th, ih, td, idd)


def test_groups_with_gaps(ctx):
    print("case groups", flush=True)
    g = _soup(30000, 7)
    ranges = [(100, 4900), (5000, 1), (5001, 0), (6000, 12000), (20000, 9000)]
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH, ranges=ranges)
    _check_same(g, th, ih, td, idd, validate=False)  # the validator walks each root's own range
    om = O.OracleMesh(g.positions, idd, td.roots)
    for r in range(len(td.roots)):
        rc, _ = om.validate_root(r)
        assert rc == 0


@pytest.mark.parametrize("indexed", [True, False])
@pytest.mark.parametrize("small", [True, False])
def test_indirect(ctx, indexed, small):
    print(f"case indirect indexed={indexed} small={small}", flush=True)
    g = _soup(3000 if small else 90000, 9)
    if not indexed:
        g = G.Geometry(g.positions[g.index], None)
    n = g.tri_count
    ranges = [(0, n // 3), (n // 2, n // 4), (n - 100, 100)]
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH, ranges=ranges, indirect=True)
    assert th.indirect.dtype == (np.uint16 if small else np.uint32)
    _check_same(g, th, ih, td, idd, validate=False)


def test_degenerate_and_extreme(ctx):
    print("case degenerate", flush=True)
    # identical triangles: no split anywhere, one leaf
    p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (500, 1))
    g = G.Geometry(p, np.arange(1500, dtype=np.uint32))
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH, max_depth=40, target_leaf_size=10)
    _check_same(g, th, ih, td, idd, validate=False)
    # coincident centers on one axis, a flat sheet, huge and tiny magnitudes
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1, 1, (30000, 3)).astype(np.float32)
    pts[:10000, 2] = 0.0
    pts[10000:20000] *= np.float32(1e30)
    pts[20000:] *= np.float32(1e-30)
    g = G.Geometry(pts, rng.permutation(30000).astype(np.uint32))
    for strategy in (M.CENTER, M.SAH):
        th, ih, td, idd = _both(ctx, g, strategy=strategy)
        _check_same(g, th, ih, td, idd, validate=False)
    # a few infinite coordinates (NaN centers and costs follow the reference's comparisons)
    pts2 = pts.copy()
    pts2[[5, 600, 7000]] = np.float32(np.inf)
    pts2[[9001, 15000]] = np.float32(-np.inf)
    g = G.Geometry(pts2, g.index.copy())
    th, ih, td, idd = _both(ctx, g, strategy=M.SAH)
    _check_same(g, th, ih, td, idd, validate=False)


def test_errors(ctx):
    print("case errors", flush=True)
    g = _soup(70000, 4)
    for host in (True, False):
        with pytest.raises(M.MbvhError) as e:
            M.build(g.positions, g.index.copy(), max_depth=0, ctx=None if host else ctx)
        assert e.value.code == -6  # MBVH_ERR_LIMIT
    bad = g.index.copy()
    bad[1234] = g.positions.shape[0]
    msgs = []
    for host in (True, False):
        with pytest.raises(M.MbvhError) as e:
            M.build(g.positions, bad.copy(), ctx=None if host else ctx)
        assert e.value.code == -4
        msgs.append(str(e.value))
    assert "index[1234]" in msgs[0] and "index[1234]" in msgs[1]


def test_device_pointers(ctx):
    torch = pytest.importorskip("torch")
    print("case build_d", flush=True)
    g = G.torus_knot(10, 3, 400, 100)
    idx_h = g.index.copy()
    th = M.build(g.positions, idx_h)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(g.positions.reshape(-1)).to(dev)
    idx = torch.from_numpy(g.index.astype(np.int32).view(np.int32)).to(dev) if g.index.dtype == np.uint32 else \
        torch.from_numpy(g.index.view(np.int16)).to(dev)
    cap = (2 * g.tri_count + 1) * 32
    nodes = torch.zeros(cap // 4, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    counts = ctx.build_d(pos.data_ptr(), g.positions.shape[0], idx.data_ptr(), g.index.itemsize, g.index.size,
                         [(0, g.tri_count)], nodes.data_ptr(), cap)
    torch.cuda.synchronize()
    got = nodes.cpu().numpy().view(np.uint8)[:counts[0] * 32]
    assert np.array_equal(got, th.roots[0])
    got_idx = idx.cpu().numpy().view(g.index.dtype)
    assert np.array_equal(got_idx, idx_h)


def test_repeatable(ctx):
    # the device build allocates children with atomics; the bytes must not depend on that order
    print("case repeat", flush=True)
    g = _soup(40000, 21)
    outs = []
    for _ in range(3):
        idx = g.index.copy()
        t = M.build(g.positions, idx, ctx=ctx)
        outs.append((t.roots[0], idx))
    for a in outs[1:]:
        assert np.array_equal(a[0], outs[0][0]) and np.array_equal(a[1], outs[0][1])
