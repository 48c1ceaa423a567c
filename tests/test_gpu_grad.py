"""GPU backward (adjoint) kernels through the C ABI, against the FP64 gradient oracle and FP64 central
finite differences: raycastFirst (dL/d origins, directions, vertex positions) and closestPoint
(dL/d query points, vertex positions), both vertex-gradient reductions (deterministic sorted
segmented sum, FP32 atomics).

Tolerances: every per-query gradient vector within 1e-3 relative of FP64 central finite differences
of the fixed-face forward map (the target for this path; the FP64 analytic gradient and those
differences agree to 1e-9 on these samples) and within 1e-4 of the FP64 analytic gradient, vertex
gradients within 1e-5 of the summed size of the query gradients they share (a vertex a
query barely weights is judged against that query's whole gradient: FP32 weights and sums carry
absolute rounding errors)."""
import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import geometry as G, workloads as W
import oracle_grad as OG
import oracle_py as O
from helpers import build_mesh, check_closest_parity, check_raycast_parity

pytestmark = pytest.mark.gpu

REL = 1e-3           # against FP64 central finite differences
REL_ANALYTIC = 1e-4  # against the FP64 analytic gradient


@pytest.fixture(scope="module")
def ctx():
    c = M.Context(0)
    yield c
    c.close()


def _rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-6)


def _vertex_check(got, ref, absum, label, count=None):
    err = np.abs(got - ref) - 1e-5 * absum - 1e-6
    v, k = np.unravel_index(np.argmax(err), err.shape)
    print(label, "vertex max excess", float(err.max()), "at", (int(v), int(k)), "got", float(got[v, k]),
          "ref", float(ref[v, k]), "scale", float(absum[v, k]), "contributions", None if count is None else int(count[v]))
    assert err.max() <= 0, (label, int(v), int(k))


@pytest.fixture(scope="module")
def ray_case(ctx):
    w = W.c3_raycast(1 << 16, detail=40)
    g, idx, tree = build_mesh(w.geometry)
    mesh = ctx.upload(tree, g.positions, idx)
    fwd = mesh.raycast_first(w.origins, w.directions)
    hit = fwd["face"] != M.NO_HIT
    rng = np.random.default_rng(11)
    n = w.origins.shape[0]
    up = {"grad_distance": rng.standard_normal(n).astype(np.float32),
          "grad_point": rng.standard_normal((n, 3)).astype(np.float32),
          "grad_barycoord": rng.standard_normal((n, 3)).astype(np.float32)}
    yield w, g, idx, mesh, fwd, hit, up
    mesh.free()


def test_raycast_bwd_vs_oracle_and_determinism(ray_case):
    w, g, idx, mesh, fwd, hit, up = ray_case
    assert hit.mean() > 0.3
    det = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_DETERMINISTIC, **up)
    det2 = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_DETERMINISTIC, **up)
    ato = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_ATOMIC, **up)
    for k in ("origins", "directions", "positions"):
        assert np.array_equal(det[k].view(np.uint32), det2[k].view(np.uint32)), k  # bitwise reproducible
    assert np.array_equal(det["origins"], ato["origins"]) and np.array_equal(det["directions"], ato["directions"])
    assert np.all(det["origins"][~hit] == 0) and np.all(det["directions"][~hit] == 0)

    f = fwd["face"][hit]
    a, b, c = OG.face_vertices(g.positions, idx, f)
    ref = OG.ray_grads(w.origins[hit], w.directions[hit], a, b, c, up["grad_distance"][hit],
                       up["grad_point"][hit], up["grad_barycoord"][hit])
    for k in ("origins", "directions"):
        e = _rel(det[k][hit], ref[k])
        print(k, "rel err p50/p99.9/max", np.quantile(e, [0.5, 0.999]), e.max())
        assert e.max() < REL_ANALYTIC, k
    # the saved forward values the backward reuses agree with the FP64 solve on the same face
    bref = np.stack([ref["w"], ref["u"], ref["v"]], axis=1)
    berr = np.abs(fwd["barycoord"][hit] - bref).max(axis=1)
    worst = int(np.argmax(berr))
    nrm = np.cross(b - a, c - a)
    cosv = np.abs(np.einsum("ij,ij->i", w.directions[hit], nrm)) / np.linalg.norm(nrm, axis=1)
    print("forward bary vs FP64 max", float(berr.max()), "ray", int(np.flatnonzero(hit)[worst]), "gpu",
          fwd["barycoord"][hit][worst], "fp64", bref[worst], "|cos|", float(cosv[worst]))
    ref_v, absum = OG.scatter_vertices(idx, f, ref["a"], ref["b"], ref["c"], g.positions.shape[0])
    count = np.bincount(idx.reshape(-1, 3)[f].ravel(), minlength=g.positions.shape[0])
    print("det vs atomic max |diff| / scale", float((np.abs(det["positions"] - ato["positions"]) /
                                                    np.maximum(absum, 1e-30)).max()))
    _vertex_check(ato["positions"], ref_v, absum, "ray atomic", count)
    _vertex_check(det["positions"], ref_v, absum, "ray det", count)
    # vertices no hit touches get exactly zero
    touched = np.zeros(g.positions.shape[0], bool)
    touched[idx.reshape(-1, 3)[f].ravel()] = True
    assert np.all(det["positions"][~touched] == 0)


def test_raycast_bwd_vs_finite_differences(ray_case):
    """1,024 sampled hits: FP64 central differences of the fixed-face forward map."""
    w, g, idx, mesh, fwd, hit, up = ray_case
    sel = np.flatnonzero(hit)[:: max(1, hit.sum() // 1024)][:1024]
    got = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_DETERMINISTIC, vertices=False, **up)
    o, d = w.origins[sel].astype(np.float64), w.directions[sel].astype(np.float64)
    a, b, c = OG.face_vertices(g.positions, idx, fwd["face"][sel])
    gd, gp, gb = (up[k][sel].astype(np.float64) for k in ("grad_distance", "grad_point", "grad_barycoord"))

    def loss(o, d, a, b, c):
        dist, point, bary = OG.ray_forward(o, d, a, b, c)
        return gd * dist + np.einsum("ij,ij->i", gp, point) + np.einsum("ij,ij->i", gb, bary)

    args = [o, d, a, b, c]
    for k, name in ((0, "origins"), (1, "directions")):
        fd = np.zeros_like(args[k])
        h = 1e-6
        for j in range(3):
            pl = [x.copy() for x in args]
            mi = [x.copy() for x in args]
            pl[k][:, j] += h
            mi[k][:, j] -= h
            fd[:, j] = (loss(*pl) - loss(*mi)) / (2 * h)
        e = _rel(got[name][sel], fd)
        print(name, "vs FD p50/p99/max", np.quantile(e, [0.5, 0.99]), e.max())
        assert e.max() < REL and np.median(e) < 1e-4, name


@pytest.fixture(scope="module")
def closest_case(ctx):
    g0 = G.displaced_icosphere(30)
    g, idx, tree = build_mesh(g0)
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 40)
    fwd = mesh.closest_point(pts)
    rng = np.random.default_rng(12)
    gd = rng.standard_normal(pts.shape[0]).astype(np.float32)
    yield g, idx, mesh, pts, fwd, gd
    mesh.free()


def test_closest_bwd_vs_oracle_and_determinism(closest_case):
    g, idx, mesh, pts, fwd, gd = closest_case
    det = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_DETERMINISTIC)
    det2 = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_DETERMINISTIC)
    ato = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_ATOMIC)
    for k in ("points", "positions"):
        assert np.array_equal(det[k].view(np.uint32), det2[k].view(np.uint32)), k
    assert np.array_equal(det["points"], ato["points"])
    a, b, c = OG.face_vertices(g.positions, idx, fwd["face"])
    ref = OG.closest_grads(pts, a, b, c, gd)
    e = _rel(det["points"], ref["points"])
    print("points rel err p50/p99.9/max", np.quantile(e, [0.5, 0.999]), e.max())
    assert e.max() < REL_ANALYTIC
    ref_v, absum = OG.scatter_vertices(idx, fwd["face"], ref["a"], ref["b"], ref["c"], g.positions.shape[0])
    _vertex_check(det["positions"], ref_v, absum, "closest det")
    _vertex_check(ato["positions"], ref_v, absum, "closest atomic")
    # gradient of the distance field has unit length
    nrm = np.linalg.norm(det["points"], axis=1) / np.abs(gd)
    assert np.all(np.abs(nrm - 1) < 1e-4)


def test_closest_bwd_vs_finite_differences(closest_case):
    g, idx, mesh, pts, fwd, gd = closest_case
    sel = np.arange(0, pts.shape[0], max(1, pts.shape[0] // 1024))[:1024]
    got = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_ATOMIC, vertices=False)
    q = pts[sel].astype(np.float64)
    a, b, c = OG.face_vertices(g.positions, idx, fwd["face"][sel])
    h = 1e-6
    fd = np.zeros_like(q)
    for j in range(3):
        qp, qm = q.copy(), q.copy()
        qp[:, j] += h
        qm[:, j] -= h
        fd[:, j] = gd[sel] * (OG.closest_distance(qp, a, b, c) - OG.closest_distance(qm, a, b, c)) / (2 * h)
    e = _rel(got["points"][sel], fd)
    print("points vs FD p50/p99/max", np.quantile(e, [0.5, 0.99]), e.max())
    assert e.max() < REL and np.median(e) < 1e-4


def test_bwd_edge_cases(ctx):
    """Misses give zero gradients; n = 0 zeroes the vertex gradients; an indirect mesh maps faces
    to its own slots; NULL outputs are skipped."""
    g, idx, tree = build_mesh(G.torus_knot(10, 3, 64, 16), indirect=True)
    mesh = ctx.upload(tree, g.positions, idx)
    lo, hi = g.bounds()
    o, d = W.rays_toward_box(4096, 40.0, lo, hi, seed_origin=5, seed_target=6)
    d[::7] = -d[::7]  # some rays point away: misses
    fwd = mesh.raycast_first(o, d)
    hit = fwd["face"] != M.NO_HIT
    assert 0.2 < hit.mean() < 0.95
    rng = np.random.default_rng(13)
    gdist = rng.standard_normal(4096).astype(np.float32)
    got = mesh.raycast_first_bwd(d, fwd, grad_distance=gdist, mode=M.GRAD_DETERMINISTIC)
    assert np.all(got["origins"][~hit] == 0)
    a, b, c = OG.face_vertices(g.positions, idx, fwd["face"][hit])
    ref = OG.ray_grads(o[hit], d[hit], a, b, c, gdist[hit])
    assert _rel(got["origins"][hit], ref["origins"]).max() < REL_ANALYTIC
    ref_v, absum = OG.scatter_vertices(idx, fwd["face"][hit], ref["a"], ref["b"], ref["c"], g.positions.shape[0])
    _vertex_check(got["positions"], ref_v, absum, "indirect")
    empty = {"face": np.zeros(0, np.uint32), "distance": np.zeros(0, np.float32),
             "barycoord": np.zeros((0, 3), np.float32)}
    z = mesh.raycast_first_bwd(np.zeros((0, 3), np.float32), empty, mode=M.GRAD_DETERMINISTIC)
    assert z["positions"].shape == (g.positions.shape[0], 3) and np.all(z["positions"] == 0)
    mesh.free()


def test_device_path_matches_host(ctx, closest_case):
    torch = pytest.importorskip("torch")
    g, idx, mesh, pts, fwd, gd = closest_case
    dev = torch.device("cuda:0")
    n = pts.shape[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_q, d_f = t(pts), t(fwd["face"].view(np.int32))
    d_dist, d_pt, d_bc, d_g = t(fwd["distance"]), t(fwd["point"]), t(fwd["barycoord"]), t(gd)
    g_q = torch.empty((n, 3), dtype=torch.float32, device=dev)
    g_v = torch.empty((mesh.vertex_count, 3), dtype=torch.float32, device=dev)
    wsb = M.Mesh.grad_workspace_bytes(n, M.GRAD_DETERMINISTIC)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()  # inputs are written on torch's stream, the call runs on the context's
    mesh.closest_point_bwd_d(d_q.data_ptr(), n, d_f.data_ptr(), d_dist.data_ptr(), d_pt.data_ptr(), d_bc.data_ptr(),
                             d_g.data_ptr(), g_q.data_ptr(), g_v.data_ptr(), M.GRAD_DETERMINISTIC, ws.data_ptr(), wsb)
    torch.cuda.synchronize()
    host = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_DETERMINISTIC)
    assert np.array_equal(g_q.cpu().numpy(), host["points"])
    assert np.array_equal(g_v.cpu().numpy(), host["positions"])
    with pytest.raises(M.MbvhError):
        mesh.closest_point_bwd_d(d_q.data_ptr(), n, d_f.data_ptr(), d_dist.data_ptr(), d_pt.data_ptr(),
                                 d_bc.data_ptr(), d_g.data_ptr(), g_q.data_ptr(), g_v.data_ptr(),
                                 M.GRAD_DETERMINISTIC, ws.data_ptr(), wsb - 1)


def _ray_fd_check(got, origins, directions, faces, g, idx, up, sel, label):
    """FP64 central differences of the fixed-face forward map on the rays `sel`."""
    o, d = origins[sel].astype(np.float64), directions[sel].astype(np.float64)
    a, b, c = OG.face_vertices(g.positions, idx, faces[sel])
    gd, gp, gb = (up[k][sel].astype(np.float64) for k in ("grad_distance", "grad_point", "grad_barycoord"))

    def loss(o, d, a, b, c):
        dist, point, bary = OG.ray_forward(o, d, a, b, c)
        return gd * dist + np.einsum("ij,ij->i", gp, point) + np.einsum("ij,ij->i", gb, bary)

    args = [o, d, a, b, c]
    for k, name in ((0, "origins"), (1, "directions")):
        fd = np.zeros_like(args[k])
        h = 1e-6
        for j in range(3):
            pl = [x.copy() for x in args]
            mi = [x.copy() for x in args]
            pl[k][:, j] += h
            mi[k][:, j] -= h
            fd[:, j] = (loss(*pl) - loss(*mi)) / (2 * h)
        e = _rel(got[name][sel], fd)
        print(label, name, "vs FD p50/max", float(np.median(e)), float(e.max()), flush=True)
        assert e.max() < REL, (label, name)


def test_c3_full_size(ctx):
    """Config C3 at its full size: the noise-displaced Icosahedron(1, 200) (808,020 triangles) and
    1,048,576 rays. Forward: every ray against the FP64 oracle's traversal. Backward (deterministic
    and atomic): every hit's origin / direction gradient against the FP64 analytic gradient, every
    vertex gradient against the FP64 sum of the analytic corner gradients, and 1,024 sampled hits
    against FP64 central differences."""
    w = W.c3_raycast()
    print("case C3 full: 808k triangles, 1M rays", flush=True)
    g, idx, tree = build_mesh(w.geometry)
    assert g.tri_count == 808020 and w.origins.shape[0] == 1 << 20
    mesh = ctx.upload(tree, g.positions, idx)
    fwd = mesh.raycast_first(w.origins, w.directions)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    ref_f = om.raycast_first(w.origins, w.directions)
    stats = check_raycast_parity(om, w.origins, w.directions, fwd, ref_f, "C3 full forward")
    print("C3 forward", stats, flush=True)
    hit = fwd["face"] != M.NO_HIT
    assert hit.mean() > 0.3
    rng = np.random.default_rng(21)
    n = w.origins.shape[0]
    up = {"grad_distance": rng.standard_normal(n).astype(np.float32),
          "grad_point": rng.standard_normal((n, 3)).astype(np.float32),
          "grad_barycoord": rng.standard_normal((n, 3)).astype(np.float32)}
    det = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_DETERMINISTIC, **up)
    ato = mesh.raycast_first_bwd(w.directions, fwd, mode=M.GRAD_ATOMIC, **up)
    assert np.array_equal(det["origins"], ato["origins"]) and np.array_equal(det["directions"], ato["directions"])
    assert np.all(det["origins"][~hit] == 0) and np.all(det["directions"][~hit] == 0)
    f = fwd["face"][hit]
    a, b, c = OG.face_vertices(g.positions, idx, f)
    ref = OG.ray_grads(w.origins[hit], w.directions[hit], a, b, c, up["grad_distance"][hit],
                       up["grad_point"][hit], up["grad_barycoord"][hit])
    for k in ("origins", "directions"):
        e = _rel(det[k][hit], ref[k])
        print("C3", k, "vs analytic p50/max", float(np.median(e)), float(e.max()), flush=True)
        assert e.max() < REL_ANALYTIC, k
    ref_v, absum = OG.scatter_vertices(idx, f, ref["a"], ref["b"], ref["c"], g.positions.shape[0])
    _vertex_check(det["positions"], ref_v, absum, "C3 det")
    _vertex_check(ato["positions"], ref_v, absum, "C3 atomic")
    sel = np.flatnonzero(hit)[:: max(1, hit.sum() // 1024)][:1024]
    _ray_fd_check(det, w.origins, w.directions, fwd["face"], g, idx, up, sel, "C3")
    mesh.free()


def test_c4_full_size(ctx):
    """Config C4 at its full size: the noise-displaced Icosahedron(1, 223) (1,003,520 triangles) and the
    256^3 grid over its box expanded 10% (16,777,216 points). Forward: a strided sample of 275k points
    against the FP64 oracle's traversal, and on every point: a face is found, distance = |q - point|,
    the barycoords sum to 1 and rebuild the point from the face's vertices. Backward (deterministic
    and atomic): every point's gradient against the FP64 analytic gradient of the distance (unit
    length), every vertex gradient against the FP64 sum of the corner gradients."""
    w = W.c4_closest(res=256)
    print("case C4 full: 1M triangles, 256^3 points", flush=True)
    g, idx, tree = build_mesh(w.geometry)
    assert g.tri_count == 1003520 and w.points.shape[0] == 256 ** 3
    mesh = ctx.upload(tree, g.positions, idx)
    pts = w.points
    fwd = mesh.closest_point(pts)
    n = pts.shape[0]
    assert np.all(fwd["face"] < g.tri_count)
    om = O.OracleMesh(g.positions, idx, tree.roots)
    sample = np.arange(0, n, 61)
    ref_s = om.closest_point(pts[sample])
    stats = check_closest_parity(om, pts[sample], {k: v[sample] for k, v in fwd.items()}, ref_s, "C4 sample")
    print("C4 forward sample", stats, flush=True)
    rng = np.random.default_rng(22)
    gd = rng.standard_normal(n).astype(np.float32)
    det = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_DETERMINISTIC)
    ato = mesh.closest_point_bwd(pts, fwd, gd, mode=M.GRAD_ATOMIC)
    assert np.array_equal(det["points"], ato["points"])
    ref_v = np.zeros((g.positions.shape[0], 3))
    absum = np.zeros_like(ref_v)
    worst = 0.0
    for s in range(0, n, 1 << 21):
        sl = slice(s, min(n, s + (1 << 21)))
        q = pts[sl].astype(np.float64)
        a, b, c = OG.face_vertices(g.positions, idx, fwd["face"][sl])
        # properties of the forward result on every point
        p = fwd["point"][sl].astype(np.float64)
        dd = np.linalg.norm(q - p, axis=1)
        assert np.max(np.abs(dd - fwd["distance"][sl]) / np.maximum(fwd["distance"][sl], 1)) < 2e-6
        bc = fwd["barycoord"][sl].astype(np.float64)
        assert np.max(np.abs(bc.sum(axis=1) - 1)) < 1e-5
        rebuilt = bc[:, :1] * a + bc[:, 1:2] * b + bc[:, 2:3] * c
        assert np.max(np.abs(rebuilt - p)) < 1e-5
        ref = OG.closest_grads(q, a, b, c, gd[sl])
        e = _rel(det["points"][sl], ref["points"])
        worst = max(worst, float(e.max()))
        rv, sc = OG.scatter_vertices(idx, fwd["face"][sl], ref["a"], ref["b"], ref["c"], g.positions.shape[0])
        ref_v += rv
        absum += sc
    print("C4 points vs analytic max", worst, flush=True)
    assert worst < REL_ANALYTIC
    nrm = np.linalg.norm(det["points"], axis=1) / np.abs(gd)
    assert np.all(np.abs(nrm - 1) < 1e-4)
    _vertex_check(det["positions"], ref_v, absum, "C4 det")
    _vertex_check(ato["positions"], ref_v, absum, "C4 atomic")
    mesh.free()
