"""GPU batched multi-mesh raycastFirst (BatchedMesh instances, one persistent launch) through the C ABI:
identical results to the per-mesh plain loop (instance rays -> one raycastFirst launch per mesh ->
world hits), parity of those results with the FP64 oracle on the instance-local rays, identical
results for every tuning, and a captured HIP graph that replays correctly between other work."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import workloads as W
import oracle_py as O
from helpers import build_mesh, check_raycast_parity

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture(scope="module")
def case(ctx):
    w = W.c5_batched(n_geometries=16, n_rays=30000, tri_max=30000, seed=8)
    built = [build_mesh(g) for g in w.geometries]
    meshes = [ctx.upload(t, g.positions, idx) for g, idx, t in built]
    batch = ctx.batch(meshes)
    n = w.origins.shape[0]
    D = {"world": _t(w.matrix_world), "inverse": _t(w.matrix_world_inverse),
         "instance_geometry": _t(w.instance_geometry.view(np.int32)), "origins": _t(w.origins),
         "directions": _t(w.directions), "ray_instance": _t(w.ray_instance.view(np.int32))}
    wsb = batch.workspace_bytes(n)
    D["workspace"] = torch.empty(wsb, dtype=torch.uint8, device=DEV)
    torch.cuda.synchronize()
    yield w, built, meshes, batch, D, wsb
    batch.free()
    for m in meshes:
        m.free()


def _outputs(n):
    return {"face": torch.empty(n, dtype=torch.int32, device=DEV), "distance": torch.empty(n, device=DEV),
            "point": torch.empty((n, 3), device=DEV), "barycoord": torch.empty((n, 3), device=DEV)}


def _params(w, D, out, wsb, n=None):
    n = w.origins.shape[0] if n is None else n
    p = {k: v.data_ptr() for k, v in D.items()}
    p.update({k: v.data_ptr() for k, v in out.items()})
    p.update(n_instances=w.matrix_world.shape[0], n=n, workspace_bytes=wsb)
    return p


def _host(out):
    return {"face": out["face"].cpu().numpy().view(np.uint32), "distance": out["distance"].cpu().numpy(),
            "point": out["point"].cpu().numpy(), "barycoord": out["barycoord"].cpu().numpy()}


def _batched(w, batch, D, wsb, tuning=None):
    out = _outputs(w.origins.shape[0])
    batch.raycast_first_d(_params(w, D, out, wsb), tuning=tuning)
    torch.cuda.synchronize()
    return _host(out)


def _plain_loop(ctx, w, meshes, D):
    """Instance rays -> one raycast_first_d per mesh over that mesh's rays -> world hits."""
    n = w.origins.shape[0]
    lo = torch.empty((n, 3), device=DEV)
    ld = torch.empty((n, 3), device=DEV)
    ctx.instance_rays_d(D["inverse"].data_ptr(), w.matrix_world.shape[0], D["origins"].data_ptr(),
                        D["directions"].data_ptr(), D["ray_instance"].data_ptr(), n, lo.data_ptr(), ld.data_ptr())
    torch.cuda.synchronize()  # the library's calls run on the context's stream, torch's on its own
    geo = w.instance_geometry[w.ray_instance]
    order = np.argsort(geo, kind="stable")
    starts = np.concatenate([[0], np.cumsum(np.bincount(geo, minlength=len(meshes)))])
    idx = _t(order.astype(np.int64))
    slo, sld = lo[idx].contiguous(), ld[idx].contiguous()
    so = _outputs(n)
    for gi, mesh in enumerate(meshes):
        a, b = int(starts[gi]), int(starts[gi + 1])
        if a == b:
            continue
        mesh.raycast_first_d(slo[a:].data_ptr(), sld[a:].data_ptr(), b - a, so["face"][a:].data_ptr(),
                             so["distance"][a:].data_ptr(), so["point"][a:].data_ptr(),
                             so["barycoord"][a:].data_ptr())
    torch.cuda.synchronize()
    out = _outputs(n)
    for k in out:
        out[k][idx] = so[k]
    local = _host(out)
    torch.cuda.synchronize()
    ctx.instance_hits_d(D["world"].data_ptr(), w.matrix_world.shape[0], D["origins"].data_ptr(),
                        D["ray_instance"].data_ptr(), n, out["face"].data_ptr(), out["point"].data_ptr(),
                        out["distance"].data_ptr())
    torch.cuda.synchronize()
    return _host(out), local, lo.cpu().numpy(), ld.cpu().numpy()


def _same(a, b):
    for k in ("face", "distance", "point", "barycoord"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def _tie_equal(a, b):
    """Results of the one-lane and four-lane traversals: the same up to exactly equal hits, which
    may resolve to the other face at the same distance."""
    assert np.array_equal(a["face"] == M.NO_HIT, b["face"] == M.NO_HIT)
    hit = a["face"] != M.NO_HIT
    assert np.all(np.abs(a["distance"][hit] - b["distance"][hit]) <= 1e-6 * np.maximum(b["distance"][hit], 1))
    assert (a["face"] == b["face"]).mean() > 0.999


def test_batched_equals_plain_loop_and_oracle(ctx, case):
    w, built, meshes, batch, D, wsb = case
    got = _batched(w, batch, D, wsb)
    loop, local, lo, ld = _plain_loop(ctx, w, meshes, D)
    _same(got, loop)
    hit = got["face"] != M.NO_HIT
    assert 0.2 < hit.mean() < 1.0
    # the loop's per-mesh results against the FP64 oracle on the same local rays
    geo = w.instance_geometry[w.ray_instance]
    for gi, (g, idx, tree) in enumerate(built):
        sel = np.flatnonzero(geo == gi)
        if sel.size == 0:
            continue
        om = O.OracleMesh(g.positions, idx, tree.roots)
        ref = om.raycast_first(lo[sel], ld[sel])
        sub = {k: local[k][sel] for k in local}
        check_raycast_parity(om, lo[sel], ld[sel], sub, ref, f"mesh{gi}")
        # world point / distance of the hits, from the oracle's local hit in FP64
        h = (ref["face"] != O.NO_HIT) & (sub["face"] == ref["face"])
        mw = w.matrix_world.astype(np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)[w.ray_instance[sel][h]]
        pw = np.einsum("nij,nj->ni", mw[:, :3, :3], ref["point"][h]) + mw[:, :3, 3]
        dw = np.linalg.norm(pw - w.origins[sel][h].astype(np.float64), axis=1)
        assert np.all(np.abs(got["point"][sel][h] - pw) <= 1e-5 * np.maximum(np.abs(pw), 1))
        assert np.all(np.abs(got["distance"][sel][h] - dw) <= 1e-5 * np.maximum(dw, 1))


def test_c5_full_size(ctx):
    """Config C5 at its full size: 512 torus knots and spheres (triangle counts log-uniform in
    [100, 200,000]), one rigid instance each, 4,194,304 rays split by Zipf(1.1). The batched launch
    equals the per-mesh plain loop bit for bit; the loop's per-mesh results on 32 sampled rays of
    every mesh meet the parity rule against the FP64 oracle on the same local rays. The trees are
    built on the device (byte-identical to the host build)."""
    w = W.c5_batched()
    print("case C5 full: 512 meshes, 4M rays", flush=True)
    built = []
    for geo in w.geometries:
        g = geo.copy()
        g.ensure_index()
        idx = g.index.copy()
        built.append((g, idx, M.build(g.positions, idx, strategy=M.SAH, ctx=ctx)))
    meshes = [ctx.upload(t, g.positions, idx) for g, idx, t in built]
    batch = ctx.batch(meshes)
    n = w.origins.shape[0]
    assert len(meshes) == 512 and n == 1 << 22
    D = {"world": _t(w.matrix_world), "inverse": _t(w.matrix_world_inverse),
         "instance_geometry": _t(w.instance_geometry.view(np.int32)), "origins": _t(w.origins),
         "directions": _t(w.directions), "ray_instance": _t(w.ray_instance.view(np.int32))}
    wsb = batch.workspace_bytes(n)
    D["workspace"] = torch.empty(wsb, dtype=torch.uint8, device=DEV)
    torch.cuda.synchronize()
    got = _batched(w, batch, D, wsb)
    loop, local, lo, ld = _plain_loop(ctx, w, meshes, D)
    _same(got, loop)
    hit = got["face"] != M.NO_HIT
    print("C5 hit fraction", float(hit.mean()), flush=True)
    assert 0.2 < hit.mean() < 1.0
    geo = w.instance_geometry[w.ray_instance]
    checked = 0
    for gi, (g, idx, tree) in enumerate(built):
        sel = np.flatnonzero(geo == gi)
        if sel.size == 0:
            continue
        sel = sel[:: max(1, sel.size // 32)][:32]
        om = O.OracleMesh(g.positions, idx, tree.roots)
        ref = om.raycast_first(lo[sel], ld[sel], threads=1)
        check_raycast_parity(om, lo[sel], ld[sel], {k: local[k][sel] for k in local}, ref, f"C5 mesh{gi}")
        checked += sel.size
    print("C5 oracle-checked rays", checked, flush=True)
    assert checked > 10000
    batch.free()
    for m in meshes:
        m.free()


def test_every_tuning_gives_identical_results(ctx, case):
    w, built, meshes, batch, D, wsb = case
    base = _batched(w, batch, D, wsb)
    for lanes, sl, tile, bl in ((1, 8, 64, 1), (1, 16, 512, 8), (1, 8, 4096, 2), (4, 16, 128, 3), (4, 16, 64, 8)):
        got = _batched(w, batch, D, wsb, M.BatchTuning(lanes, sl, tile, bl, 0.0))
        if lanes == 4:
            _same(got, base)
        else:
            _tie_equal(got, base)
    out = _outputs(w.origins.shape[0])
    t = batch.autotune(_params(w, D, out, wsb))
    assert t.ray_lanes in (1, 4) and t.stack_lds in (8, 16) and t.tile_rays % 64 == 0
    assert 1 <= t.blocks_per_cu <= 8 and t.ms > 0
    tuned = _batched(w, batch, D, wsb)  # now with the tuned configuration
    _tie_equal(_host(out), base)
    if t.ray_lanes == 4:
        _same(tuned, base)
    else:
        _tie_equal(tuned, base)


def test_graph_replays_between_other_work(ctx, case):
    w, built, meshes, batch, D, wsb = case
    n = w.origins.shape[0]
    out = _outputs(n)
    graph = batch.graph(_params(w, D, out, wsb))
    base = _batched(w, batch, D, wsb)
    rng = np.random.default_rng(3)
    for rep in range(3):
        for v in out.values():
            v.fill_(-7)  # eager work on the outputs between replays
        junk = torch.randn(1 << 20, device=DEV).sum().item()  # and unrelated work
        torch.cuda.synchronize()  # torch's stream and the context's stream are not ordered
        graph.launch()
        torch.cuda.synchronize()
        _same(_host(out), base)
        assert np.isfinite(junk)
    # the graph reads its input buffers at replay time: new rays, same buffers
    saved = D["directions"].clone()
    perm = _t(rng.permutation(n).astype(np.int64))
    D["directions"].copy_(saved[perm])
    torch.cuda.synchronize()
    graph.launch()
    torch.cuda.synchronize()
    _same(_host(out), _batched(w, batch, D, wsb))
    D["directions"].copy_(saved)
    graph.free()


def test_graph_refused_after_stack_scratch_moves():
    """A captured graph holds the address of the context's traversal stack overflow region. Uploading
    a mesh deeper than that region covers makes the context reallocate it; a later replay of the old
    graph is refused with an error (and never launched)."""
    c = M.Context(0)
    try:
        w = W.c5_batched(n_geometries=4, n_rays=2000, tri_max=3000, seed=4)
        built = [build_mesh(g) for g in w.geometries]
        meshes = [c.upload(t, g.positions, idx) for g, idx, t in built]
        batch = c.batch(meshes)
        n = w.origins.shape[0]
        D = {"world": _t(w.matrix_world), "inverse": _t(w.matrix_world_inverse),
             "instance_geometry": _t(w.instance_geometry.view(np.int32)), "origins": _t(w.origins),
             "directions": _t(w.directions), "ray_instance": _t(w.ray_instance.view(np.int32))}
        wsb = batch.workspace_bytes(n)
        D["workspace"] = torch.empty(wsb, dtype=torch.uint8, device=DEV)
        out = _outputs(n)
        torch.cuda.synchronize()
        graph = batch.graph(_params(w, D, out, wsb))
        graph.launch()  # the scratch has not moved: replays
        c.synchronize()
        # 100 triangles at x = 2^i: CENTER splits peel one triangle per level, so the tree is as deep
        # as maxDepth allows; a depth of 64 needs 57 overflow entries per thread (56 are reserved)
        tris = []
        for i in range(100):
            x = float(2.0 ** i)
            tris += [[x, 0, 0], [x, 1, 0], [x, 0, 1]]
        pos = np.array(tris, np.float32)
        idx = np.arange(300, dtype=np.uint32)
        tree = M.build(pos, idx, strategy=M.CENTER, max_depth=70, target_leaf_size=1)
        deep = c.upload(tree, pos, idx)
        assert deep.info()["max_depth"] >= 64
        with pytest.raises(M.MbvhError) as e:
            graph.launch()
        assert e.value.code == -3 and "reallocated" in str(e.value)
        # a graph captured now replays, and matches the eager launch
        graph.free()
        graph = batch.graph(_params(w, D, out, wsb))
        graph.launch()
        c.synchronize()
        got = _host(out)
        _same(got, _batched(w, batch, D, wsb))
        graph.free()
        deep.free()
        batch.free()
        for m in meshes:
            m.free()
    finally:
        c.close()


def test_invalid_instances_miss(ctx, case):
    w, built, meshes, batch, D, wsb = case
    n = 2000
    bad = w.ray_instance[:n].copy()
    bad[::3] = 10_000  # no such instance
    D2 = dict(D)
    D2["ray_instance"] = _t(bad.view(np.int32))
    torch.cuda.synchronize()
    out = _outputs(n)
    batch.raycast_first_d(_params(w, D2, out, batch.workspace_bytes(n), n=n))
    torch.cuda.synchronize()
    got = _host(out)
    assert np.all(got["face"][::3] == M.NO_HIT) and np.all(np.isinf(got["distance"][::3]))
    ref = _batched(w, batch, D, wsb)
    keep = np.ones(n, bool)
    keep[::3] = False
    assert np.array_equal(got["face"][keep], ref["face"][:n][keep])
    with pytest.raises(M.MbvhError):
        batch.raycast_first_d(_params(w, D2, out, 16, n=n))  # workspace too small


def test_host_form_equals_device_form(ctx, case):
    w, built, meshes, batch, D, wsb = case
    got = batch.raycast_first(w.matrix_world, w.matrix_world_inverse, w.instance_geometry, w.origins,
                              w.directions, w.ray_instance)
    _same(got, _batched(w, batch, D, wsb))
    empty = batch.raycast_first(w.matrix_world, w.matrix_world_inverse, w.instance_geometry,
                                np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros(0, np.uint32))
    assert empty["face"].shape == (0,)
