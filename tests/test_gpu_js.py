"""GPU parity of the JavaScript API (MeshBVH / acceleratedRaycast via the Node-API addon) against the
FP64 oracle: the same geometry, rays and points go through Node, results come back here."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import mbvh_amd as M
from mbvh_amd import workloads as W
import oracle_py as O
from helpers import check_raycast_parity, check_closest_parity

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_js_api_matches_oracle(tmp_path):
    w = W.c1_raycast(8192)
    g = w.geometry
    lo, hi = g.bounds()
    pts = W.grid_points(lo, hi, 20)
    d = tmp_path
    g.positions.tofile(d / "positions.bin")
    g.index.tofile(d / "index.bin")
    w.origins.tofile(d / "origins.bin")
    w.directions.tofile(d / "dirs.bin")
    pts.tofile(d / "points.bin")
    rng = np.random.default_rng(31)
    g_dist = rng.standard_normal(w.origins.shape[0]).astype(np.float32)
    g_bary = rng.standard_normal((w.origins.shape[0], 3)).astype(np.float32)
    g_cp = rng.standard_normal(pts.shape[0]).astype(np.float32)
    g_dist.tofile(d / "g_dist.bin")
    g_bary.tofile(d / "g_bary.bin")
    g_cp.tofile(d / "g_cp.bin")
    (d / "manifest.json").write_text(json.dumps({"index_u32": g.index.dtype == np.uint32}))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "gpu_run.js"), str(d)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert '"asyncChecked":true' in r.stdout  # the asynchronous queries ran and matched
    assert '"microBatchChecked":' in r.stdout  # the micro-batched per-call queries ran and matched

    idx = np.fromfile(d / "bvh_index.bin", dtype=g.index.dtype)
    root = np.fromfile(d / "bvh_root0.bin", dtype=np.uint8)
    # the JS build (through the addon) equals the Python-ABI build of the same geometry
    idx_py = g.index.copy()
    tree_py = M.build(g.positions, idx_py, strategy=M.SAH, target_leaf_size=10)
    assert np.array_equal(idx, idx_py) and np.array_equal(root, tree_py.roots[0])

    om = O.OracleMesh(g.positions, idx, [root])
    got = {"face": np.fromfile(d / "rf_face.bin", np.uint32), "distance": np.fromfile(d / "rf_dist.bin", np.float32),
           "point": np.fromfile(d / "rf_point.bin", np.float32).reshape(-1, 3),
           "barycoord": np.fromfile(d / "rf_bary.bin", np.float32).reshape(-1, 3)}
    ref = om.raycast_first(w.origins, w.directions)
    print("js raycastFirst", check_raycast_parity(om, w.origins, w.directions, got, ref, "js_raycastFirst"))

    gotc = {"face": np.fromfile(d / "cp_face.bin", np.uint32), "distance": np.fromfile(d / "cp_dist.bin", np.float32),
            "point": np.fromfile(d / "cp_point.bin", np.float32).reshape(-1, 3)}
    refc = om.closest_point(pts)
    print("js closestPoint", check_closest_parity(om, pts, gotc, refc, "js_closest"))

    # raycastObject3D / acceleratedRaycast under rigid, scaled and windowed transforms: the GPU's first
    # hit for the double world ray against the FP64 oracle traced on the reference's own local ray
    # (double precision, written by the driver); faces equal up to ties, world distances to 1e-5
    lo3 = np.fromfile(d / "xf_origins.bin", np.float64).reshape(-1, 3)
    ld3 = np.fromfile(d / "xf_dirs.bin", np.float64).reshape(-1, 3)
    win = np.fromfile(d / "xf_windows.bin", np.float64).reshape(-1, 2)
    xface = np.fromfile(d / "xf_face.bin", np.uint32)
    xdist = np.fromfile(d / "xf_dist.bin", np.float64)
    xpoint = np.fromfile(d / "xf_point.bin", np.float64).reshape(-1, 3)
    xwo = np.fromfile(d / "xf_world_origins.bin", np.float64).reshape(-1, 3)
    mats = np.fromfile(d / "xf_matrices.bin", np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)
    nx = xface.shape[0] // mats.shape[0]
    assert nx * mats.shape[0] >= 10000
    bad = ties = hits = 0
    for ti in range(mats.shape[0]):
        sl = slice(ti * nx, (ti + 1) * nx)
        if np.all(np.isinf(win[sl, 1])) and np.all(win[sl, 0] == 0):
            ref = om.raycast_first(lo3[sl], ld3[sl])
        else:  # the window differs per ray (near / far over each ray's scale factor)
            parts = [om.raycast_first(lo3[sl][i:i + 1], ld3[sl][i:i + 1], near=win[sl][i, 0], far=win[sl][i, 1])
                     for i in range(nx)]
            ref = {k: np.concatenate([p[k] for p in parts]) for k in ("face", "distance", "point")}
        gf = xface[sl]
        assert np.array_equal(gf == M.NO_HIT, ref["face"] == O.NO_HIT), f"transform {ti}: hit / miss differs"
        for i in np.nonzero((gf != ref["face"]) & (gf != M.NO_HIT))[0]:
            dplane = om.ray_plane_distance(int(gf[i]), lo3[sl][i], ld3[sl][i])
            if abs(dplane - ref["distance"][i]) <= 1e-6 * max(ref["distance"][i], 1.0):
                ties += 1
            else:
                bad += 1
        h = gf != M.NO_HIT
        hits += int(h.sum())
        # the oracle's local hit point to world space, and its distance from the world origin
        pw = ref["point"][h] @ mats[ti, :3, :3].T + mats[ti, :3, 3]
        dw = np.linalg.norm(pw - xwo[sl][h], axis=1)
        assert np.all(np.abs(xdist[sl][h] - dw) <= 1e-5 * np.maximum(dw, 1)), f"transform {ti}: world distance"
        assert np.all(np.abs(xpoint[sl][h] - pw) <= 1e-5 * np.maximum(np.abs(pw), 1)), f"transform {ti}: world point"
    print("js acceleratedRaycast under transforms:", {"rays": int(xface.shape[0]), "hits": hits, "ties": ties})
    assert bad == 0 and hits > 3000

    # BatchedMesh firstHitOnly (one launch over the visible instances): the GPU's first hit after the
    # distance sort against the nearest of the FP64 oracle's per-instance hits on the reference's local
    # rays (equal world distances: the lower instance, as the stable sort keeps it)
    visible = np.fromfile(d / "bm_visible.bin", np.uint32)
    bo = np.fromfile(d / "bm_origins.bin", np.float64).reshape(len(visible), -1, 3)
    bd = np.fromfile(d / "bm_dirs.bin", np.float64).reshape(len(visible), -1, 3)
    bw = np.fromfile(d / "bm_world_origins.bin", np.float64).reshape(-1, 3)
    bm = np.fromfile(d / "bm_matrices.bin", np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)
    bface = np.fromfile(d / "bm_face.bin", np.uint32)
    binst = np.fromfile(d / "bm_inst.bin", np.uint32)
    bdist = np.fromfile(d / "bm_dist.bin", np.float64)
    nb = bface.shape[0]
    best_d = np.full(nb, np.inf)
    best_f = np.full(nb, O.NO_HIT, np.uint32)
    best_i = np.full(nb, O.NO_HIT, np.uint32)
    per = []
    for v in range(len(visible)):
        ref = om.raycast_first(bo[v], bd[v])
        per.append(ref)
        h = ref["face"] != O.NO_HIT
        pw = ref["point"] @ bm[v, :3, :3].T + bm[v, :3, 3]
        dw = np.where(h, np.linalg.norm(pw - bw, axis=1), np.inf)
        better = dw < best_d
        best_d[better], best_f[better], best_i[better] = dw[better], ref["face"][better], visible[v]
    assert np.array_equal(bface == M.NO_HIT, best_f == O.NO_HIT), "BatchedMesh hit / miss differs"
    bbad = bties = 0
    for i in np.nonzero((bface != best_f) | (binst != best_i))[0]:
        v = int(np.nonzero(visible == binst[i])[0][0])
        dplane = om.ray_plane_distance(int(bface[i]), bo[v][i], bd[v][i])
        pw = bo[v][i] + dplane * bd[v][i] / np.linalg.norm(bd[v][i])
        dw = np.linalg.norm(pw @ bm[v, :3, :3].T + bm[v, :3, 3] - bw[i])
        if abs(dw - best_d[i]) <= 1e-6 * max(best_d[i], 1.0):
            bties += 1
        else:
            bbad += 1
    h = bface != M.NO_HIT
    assert np.all(np.abs(bdist[h] - best_d[h]) <= 1e-5 * np.maximum(best_d[h], 1))
    print("js BatchedMesh firstHitOnly:", {"rays": nb, "hits": int(h.sum()), "ties": bties})
    assert bbad == 0 and h.sum() > 500

    # all hits, DoubleSide, in the reference's traversal order
    offs = np.fromfile(d / "ra_offsets.bin", np.float64).astype(np.int64)
    faces = np.fromfile(d / "ra_face.bin", np.uint32)
    dists = np.fromfile(d / "ra_dist.bin", np.float32)
    for i in range(len(offs) - 1):
        rf_, rd_ = om.raycast_all(w.origins[i], w.directions[i], side=O.DOUBLE)
        gf = faces[offs[i]:offs[i + 1]]
        gd = dists[offs[i]:offs[i + 1]]
        assert len(gf) == len(rf_), (i, gf, rf_)
        assert np.array_equal(gf, rf_), (i, gf, rf_)
        assert np.all(np.abs(gd - rd_) <= 1e-5 * np.maximum(rd_, 1))

    # forward / backward pair through the addon == the same backward through the C ABI here
    # (deterministic reduction: bitwise), whose values test_gpu_grad.py checks against FP64
    ctx = M.Context(0)
    mesh = ctx.upload(tree_py, g.positions, idx)
    fwd = mesh.raycast_first(w.origins, w.directions)
    ref = mesh.raycast_first_bwd(w.directions, fwd, grad_distance=g_dist, grad_barycoord=g_bary,
                                 mode=M.GRAD_DETERMINISTIC)
    assert np.array_equal(np.fromfile(d / "bw_origins.bin", np.float32).reshape(-1, 3), ref["origins"])
    assert np.array_equal(np.fromfile(d / "bw_dirs.bin", np.float32).reshape(-1, 3), ref["directions"])
    assert np.array_equal(np.fromfile(d / "bw_positions.bin", np.float32).reshape(-1, 3), ref["positions"])
    fwdc = mesh.closest_point(pts)
    refc = mesh.closest_point_bwd(pts, fwdc, g_cp, mode=M.GRAD_DETERMINISTIC)
    assert np.array_equal(np.fromfile(d / "bc_points.bin", np.float32).reshape(-1, 3), refc["points"])
    assert np.array_equal(np.fromfile(d / "bc_positions.bin", np.float32).reshape(-1, 3), refc["positions"])
    mesh.free()
    ctx.close()

    # intersectsSphere batch == the oracle's brute force with the reference's sphere-triangle test
    sr = np.fromfile(d / "sphere_radii.bin", np.float32)
    si = np.fromfile(d / "sphere_hit.bin", np.uint8).astype(bool)
    sub = np.arange(0, len(sr), 20)  # brute force over 80k triangles: a sample of the spheres
    assert np.array_equal(si[sub], om.intersects_sphere(pts[sub], sr[sub]))
    assert 0.3 < si.mean() < 0.7

    # MeshBVH.refit after a deformation: the refit nodes equal the oracle's refit, and raycasts on
    # the refit tree match the oracle on the deformed geometry
    pos2 = np.fromfile(d / "refit_positions.bin", np.float32)
    root2 = np.fromfile(d / "refit_root0.bin", np.uint8)
    assert not np.array_equal(pos2, g.positions.reshape(-1))
    assert np.array_equal(root2, O.OracleMesh(pos2, idx, [root]).refit()[0])
    om2 = O.OracleMesh(pos2, idx, [root2])
    got2 = {"face": np.fromfile(d / "refit_rf_face.bin", np.uint32),
            "distance": np.fromfile(d / "refit_rf_dist.bin", np.float32),
            "point": np.fromfile(d / "refit_rf_point.bin", np.float32).reshape(-1, 3),
            "barycoord": np.fromfile(d / "refit_rf_bary.bin", np.float32).reshape(-1, 3)}
    print("js refit raycastFirst", check_raycast_parity(om2, w.origins, w.directions, got2,
                                                        om2.raycast_first(w.origins, w.directions), "js_refit"))
    # MeshBVH.refit(Set of node ids) after a second edit: the oracle's partial refit of the same set
    # (the values that name no node are dropped on both sides)
    pos3 = np.fromfile(d / "partial_positions.bin", np.float32)
    root3 = np.fromfile(d / "partial_root0.bin", np.uint8)
    ids = {1} | set(range(0, root2.size // 32, 7))
    assert np.array_equal(root3, O.OracleMesh(pos3, idx, [root2]).refit(ids)[0])
    assert not np.array_equal(root3, O.OracleMesh(pos3, idx, [root2]).refit()[0])  # it was partial
