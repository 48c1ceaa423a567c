#!/usr/bin/env python3
"""Benchmark of the BVH query hot path on one MI355X.

Workload (one "step"): closestPointToPoint for every point of the 128^3 grid over the AABB (+10%)
of TorusKnot(10, 3, 800, 200) (320,000 triangles, SAH, targetLeafSize 10) — the closest-point
configuration of the benchmark spec. Inputs and outputs are resident in HBM before the timed
region; each step is one launch of the closest-point kernel over all 2,097,152 points.
Also reported in the same line: first-hit raycast throughput on TorusKnot(10, 3, 400, 100) with
65,536 rays (Mrays/s).

The JSON line carries:
  roofline      algorithmic bytes of the closest-point kernel (per query: 32 B of I/O + 32 B per
                node box read + 48 B per triangle tested (3 x u32 index + 3 x 12 B vertices), with
                the node / triangle counts of the reference traversal order measured by the FP64
                oracle on a sample) divided by the kernel's average duration (HIP events on the
                launch stream), against the 8 TB/s HBM peak.
  cpu_baseline  the reference side timed on this host: an FP64 JavaScript restatement of the
                reference's closestPointToPoint (oracle/js) in Node, a worker pool over every query
                (`value`) and one thread over a strided sample; the C oracle's rate beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "three-mesh-bvh_amd", "python"))

HBM_PEAK_GBS = 8000.0
IDX_BYTES = 4  # uint32 index for the 320k-triangle mesh


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-order", action="store_true",
                    help="trace queries in their given order (no spatial ordering pass in the step)")
    ap.add_argument("--cpu-sample", type=int, default=16, help="1/N stride of the CPU-baseline sample")
    ap.add_argument("--only-differentiable", action="store_true",
                    help="run only the C3 / C4 forward and backward comparison (for profiling)")
    ap.add_argument("--only-primary", action="store_true",
                    help="skip the differentiable (C3, C4) and batched (C5) measurements")
    ap.add_argument("--only-build", action="store_true",
                    help="time the device BVH build of the 1M-triangle icosphere alone (for profiling) and exit")
    ap.add_argument("--only-geometry", action="store_true",
                    help="time closestPointToGeometry on the closest-point mesh alone (for profiling) and exit")
    ap.add_argument("--only-scene", action="store_true",
                    help="time the scene queries (raycast and closest point over the C5 instances) alone (for "
                         "profiling) and exit")
    ap.add_argument("--c3-rays", type=int, default=1 << 20)
    ap.add_argument("--c4-grid", type=int, default=256)
    ap.add_argument("--c5-rays", type=int, default=1 << 22)
    ap.add_argument("--c5-geometries", type=int, default=512)
    ap.add_argument("--scene-rays", type=int, default=1 << 20)
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="also run one 256 MiB device copy, the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiling run")
    return ap.parse_args()


def timed_kernel(torch, launch, steps, warmup, stream, pre=None):
    """Runs `pre` (if any) then `launch` per step; returns the wall time of all steps and the per-step
    duration of `launch` alone from HIP events recorded on `stream` around it."""
    for _ in range(warmup):
        if pre:
            pre()
        launch()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        if pre:
            pre()
        a.record(stream)
        launch()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [a.elapsed_time(b) for a, b in evs]
    return wall, kern


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "closest_point_packet_torusknot320k_grid128.md")
SIMDS, CLOCK_GHZ = 1024, 2.4  # MI355X: 256 CUs x 4 SIMDs; one wave64 VALU op per SIMD per 4 cycles


def profiled_value(key):
    """A figure of the committed counter summary of the closest-point launch (`key: value` lines):
    traffic_bytes_per_launch (FETCH_SIZE and WRITE_SIZE of `rocprofv3 --pmc` runs, each scaled by
    what the same counter reads for a 256 MiB device copy) or valu_insts_per_launch (SQ_INSTS_VALU).
    None if the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def percall_latency(positions, index, rays, points, calls=2000):
    """Per-call latency of MeshBVH.raycastFirst / closestPointToPoint through the JS package, the
    Node-API addon and the C ABI (three-mesh-bvh_amd/bench/percall.js), one query per call. None
    when Node or the addon is absent."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    node = shutil.which("node")
    if node is None or not os.path.exists(os.path.join(ROOT, "three-mesh-bvh_amd", "lib", "mbvh.node")):
        return None
    d = tempfile.mkdtemp(prefix="mbvh_percall_")
    try:
        idx = np.ascontiguousarray(index)
        ib = 2 if idx.dtype == np.uint16 else 4
        np.ascontiguousarray(positions, np.float32).tofile(os.path.join(d, "positions.f32"))
        idx.tofile(os.path.join(d, "index.u16" if ib == 2 else "index.u32"))
        np.ascontiguousarray(rays, np.float32).tofile(os.path.join(d, "rays.f32"))
        np.ascontiguousarray(points, np.float32).tofile(os.path.join(d, "points.f32"))
        with open(os.path.join(d, "meta.json"), "w") as f:
            json.dump({"index_bytes": ib}, f)
        r = subprocess.run([node, os.path.join(ROOT, "three-mesh-bvh_amd", "bench", "percall.js"), d, str(calls)],
                           capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            log(f"per-call latency run failed: {r.stderr[-500:]}")
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def max_rel(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))) if a.size else 0.0


def barycoords64(p, a, b, c):
    """Barycentric weights (a, b, c) of points p on triangles (a, b, c), FP64 (three's getBarycoord)."""
    import numpy as np
    v0, v1, v2 = c - a, b - a, p - a
    d00, d01, d02 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v0 * v2).sum(1)
    d11, d12 = (v1 * v1).sum(1), (v1 * v2).sum(1)
    den = d00 * d11 - d01 * d01
    u = (d11 * d02 - d01 * d12) / den
    v = (d00 * d12 - d01 * d02) / den
    return np.stack([1 - u - v, v, u], axis=1)


def js_cpu_baseline(kind, positions, index, roots, queries, threads, single, **meta):
    """Times the FP64 JavaScript restatement of the reference query path (oracle/js) in Node on this
    host: one thread over `single` queries (the first rows of `queries`), and a worker pool of
    `threads` over all of them. Returns Node's JSON report plus the single-thread faces, or None when
    Node is absent."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    node = shutil.which("node")
    if node is None:
        return None
    d = tempfile.mkdtemp(prefix="mbvh_cpu_")
    try:
        idx = np.ascontiguousarray(index)
        ib = 2 if idx.dtype == np.uint16 else 4
        np.ascontiguousarray(positions, np.float32).tofile(os.path.join(d, "positions.f32"))
        idx.astype(np.uint16 if ib == 2 else np.uint32).tofile(os.path.join(d, "index.u16" if ib == 2 else "index.u32"))
        for i, r in enumerate(roots):
            with open(os.path.join(d, f"root{i}.bin"), "wb") as f:
                f.write(bytes(r))
        np.ascontiguousarray(queries, np.float32).tofile(os.path.join(d, "queries.f32"))
        n = queries.shape[0] if kind == "closest" else queries.shape[0] // 2
        enc = {k: ("inf" if v == float("inf") else v) for k, v in meta.items()}  # JSON has no infinity
        with open(os.path.join(d, "meta.json"), "w") as f:
            json.dump(dict(kind=kind, n=int(n), roots=len(roots), index_bytes=ib, **enc), f)
        r = subprocess.run([node, os.path.join(ROOT, "oracle", "js", "cpu_bench.js"), d, str(threads), str(single)],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            log(f"node CPU baseline failed: {r.stderr[-500:]}")
            return None
        rep = json.loads(r.stdout.strip().splitlines()[-1])
        rep["faces"] = np.fromfile(os.path.join(d, "faces_single.u32"), np.uint32)
        return rep
    finally:
        shutil.rmtree(d, ignore_errors=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def event_ms(torch, stream, fn, reps):
    """Average duration of fn() over reps runs, timed with HIP events on `stream` (after one warm-up)."""
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_differentiable(torch, M, W, ctx, stream, args):
    """C3 / C4: forward alone vs forward + backward (deterministic and atomic vertex reductions)."""
    import numpy as np
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    out = {}
    # C3: raycastFirst on the noise-displaced icosphere (detail 200), 1M rays
    log("C3 setup")
    w = W.c3_raycast(args.c3_rays, detail=200)
    g = w.geometry.copy()
    g.ensure_index()
    idx = g.index.copy()
    tree = M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)
    mesh = ctx.upload(tree, g.positions, idx)
    n = w.origins.shape[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    o, d = t(w.origins), t(w.directions)
    face = torch.empty(n, dtype=torch.int32, device=dev)
    dist, pt, bc = (torch.empty(s, device=dev) for s in (n, (n, 3), (n, 3)))
    rng = np.random.default_rng(17)
    g_dist, g_bary = t(rng.standard_normal(n).astype(np.float32)), t(rng.standard_normal((n, 3)).astype(np.float32))
    g_o, g_d = torch.empty((n, 3), device=dev), torch.empty((n, 3), device=dev)
    g_v = torch.empty((mesh.vertex_count, 3), device=dev)
    wsb = M.Mesh.grad_workspace_bytes(n, M.GRAD_DETERMINISTIC)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)

    def fwd():
        mesh.raycast_first_d(o.data_ptr(), d.data_ptr(), n, face.data_ptr(), dist.data_ptr(), pt.data_ptr(),
                             bc.data_ptr(), stream=sptr)

    def bwd(mode):
        mesh.raycast_first_bwd_d(d.data_ptr(), n, face.data_ptr(), dist.data_ptr(), bc.data_ptr(),
                                 g_dist.data_ptr(), 0, g_bary.data_ptr(), g_o.data_ptr(), g_d.data_ptr(),
                                 g_v.data_ptr(), mode, ws.data_ptr(), wsb, stream=sptr)

    f_ms = event_ms(torch, stream, fwd, 5)
    fb_det = event_ms(torch, stream, lambda: (fwd(), bwd(M.GRAD_DETERMINISTIC)), 5)
    fb_atom = event_ms(torch, stream, lambda: (fwd(), bwd(M.GRAD_ATOMIC)), 5)
    hits = int((face.cpu().numpy() >= 0).sum())
    out["differentiable_raycast"] = {
        "workload": f"raycastFirst displaced Icosahedron(1,200) {g.tri_count} tris, {n} rays, upstream "
                    "dL/ddistance and dL/dbarycoord ~ N(0,1); grads wrt origins, directions, vertex positions",
        "hits": hits, "fwd_ms": round(f_ms, 4), "fwd_bwd_ms": round(fb_det, 4),
        "fwd_bwd_over_fwd": round(fb_det / f_ms, 3), "fwd_bwd_ms_atomic": round(fb_atom, 4),
        "fwd_bwd_over_fwd_atomic": round(fb_atom / f_ms, 3), "fwd_mrays_per_s": round(n / f_ms / 1e3, 2)}
    log(f"C3 {out['differentiable_raycast']}")
    mesh.free()
    del g_v, ws
    # C4: closest point on the noise-displaced icosphere (detail 223), 256^3 grid
    log("C4 setup")
    w = W.c4_closest(args.c4_grid, detail=223)
    g = w.geometry.copy()
    g.ensure_index()
    idx = g.index.copy()
    tree = M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)
    mesh = ctx.upload(tree, g.positions, idx)
    n = w.points.shape[0]
    q = t(w.points)
    face = torch.empty(n, dtype=torch.int32, device=dev)
    dist, pt, bc = (torch.empty(s, device=dev) for s in (n, (n, 3), (n, 3)))
    order = torch.empty(n, dtype=torch.int32, device=dev)
    owsb = M.Context.order_workspace_bytes(n)
    ows = torch.empty(owsb, dtype=torch.uint8, device=dev)
    g_dist = t(rng.standard_normal(n).astype(np.float32))
    g_q = torch.empty((n, 3), device=dev)
    g_v = torch.empty((mesh.vertex_count, 3), device=dev)
    wsb = M.Mesh.grad_workspace_bytes(n, M.GRAD_DETERMINISTIC)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)

    def fwd_c():
        ctx.order_points_d(q.data_ptr(), n, order.data_ptr(), ows.data_ptr(), owsb, stream=sptr)
        mesh.closest_point_d(q.data_ptr(), n, face.data_ptr(), dist.data_ptr(), pt.data_ptr(), bc.data_ptr(),
                             stream=sptr, order_ptr=order.data_ptr())

    def bwd_c(mode):
        mesh.closest_point_bwd_d(q.data_ptr(), n, face.data_ptr(), dist.data_ptr(), pt.data_ptr(), bc.data_ptr(),
                                 g_dist.data_ptr(), g_q.data_ptr(), g_v.data_ptr(), mode, ws.data_ptr(), wsb,
                                 stream=sptr)

    ctx.set_point_lanes(4)
    f4_ms = event_ms(torch, stream, fwd_c, 2)
    ctx.set_point_lanes(64)  # grid queries: packets of 64 (identical results)
    f_ms = event_ms(torch, stream, fwd_c, 2)
    b_det = event_ms(torch, stream, lambda: bwd_c(M.GRAD_DETERMINISTIC), 3)
    b_atom = event_ms(torch, stream, lambda: bwd_c(M.GRAD_ATOMIC), 3)
    out["differentiable_closest"] = {
        "workload": f"closestPoint displaced Icosahedron(1,223) {g.tri_count} tris, {args.c4_grid}^3 grid, "
                    "upstream dL/ddistance ~ N(0,1); grads wrt query points, vertex positions",
        "fwd_ms": round(f_ms, 3), "fwd_ms_quad_mode": round(f4_ms, 3), "fwd_mode": "packets of 64 queries",
        "bwd_ms_deterministic": round(b_det, 3), "bwd_ms_atomic": round(b_atom, 3),
        "fwd_bwd_over_fwd": round((f_ms + b_det) / f_ms, 3),
        "fwd_bwd_over_fwd_atomic": round((f_ms + b_atom) / f_ms, 3), "fwd_mqueries_per_s": round(n / f_ms / 1e3, 2)}
    log(f"C4 {out['differentiable_closest']}")
    out["export_textures"] = bench_export(torch, ctx, stream, mesh, g, idx, tree)  # before the refit moves the mesh
    log(f"export {out['export_textures']}")
    out["refit"] = bench_refit(torch, stream, mesh, g, idx, tree)
    log(f"refit {out['refit']}")
    out["build"] = bench_build(torch, M, ctx, w.geometry, "the C4 mesh (displaced Icosahedron(1,223))")
    log(f"build {out['build']}")
    mesh.free()
    return out


def bench_refit(torch, stream, mesh, g, idx, tree):
    """MeshBVH.refit on the device on the C4 mesh after a vertex displacement: the device entry point
    (positions in HBM, kernels only, HIP events on the launch stream) and the host entry point
    (positions in, refit nodes out), against the reference's refit restated in JavaScript (one thread)."""
    import numpy as np
    p0 = g.positions.reshape(-1, 3)
    pos = [(p0 * (1.0 + 0.01 * np.sin(7.0 * p0[:, :1] + k))).astype(np.float32) for k in range(2)]
    first = mesh.refit(pos[0])  # from the uploaded tree: compared with the JavaScript refit below
    times = []
    for k in range(6):
        t0 = time.perf_counter()
        mesh.refit(pos[(k + 1) % 2])  # every call moves the vertices again
        times.append(time.perf_counter() - t0)
    dev = torch.device("cuda:0")
    pd = [torch.from_numpy(p).to(dev) for p in pos]
    torch.cuda.synchronize()
    step = [0]

    def refit_d():
        mesh.refit_d(pd[step[0] % 2].data_ptr(), p0.shape[0], stream=stream.cuda_stream)
        step[0] += 1

    d_ms = event_ms(torch, stream, refit_d, 10)
    res = {"workload": f"refit of the C4 mesh ({g.tri_count} tris, {p0.shape[0]} vertices, "
                       f"{tree.roots[0].nbytes // 32} nodes) after a vertex displacement",
           "gpu_device_ms": round(d_ms, 4),
           "gpu_host_api_ms": round(1000 * float(np.median(times)), 3),
           "note": "device: positions in HBM, every call moves the vertices; host: positions copied in, "
                   "refit nodes copied out, synchronised"}
    import shutil
    import subprocess
    import tempfile
    node = shutil.which("node")
    if node is not None:
        d = tempfile.mkdtemp(prefix="mbvh_refit_")
        try:
            pos[0].tofile(os.path.join(d, "positions.f32"))
            ib = 2 if idx.dtype == np.uint16 else 4
            idx.tofile(os.path.join(d, "index.u16" if ib == 2 else "index.u32"))
            for i, rt in enumerate(tree.roots):
                rt.tofile(os.path.join(d, f"root{i}.bin"))
            with open(os.path.join(d, "meta.json"), "w") as f:
                json.dump({"roots": len(tree.roots), "index_bytes": ib}, f)
            out = subprocess.run([node, os.path.join(ROOT, "oracle", "js", "refit_bench.js"), d], capture_output=True,
                                 text=True, timeout=600)
            if out.returncode == 0:
                rep = json.loads(out.stdout.strip().splitlines()[-1])
                js_roots = [np.fromfile(os.path.join(d, f"refit{i}.bin"), np.uint8) for i in range(len(tree.roots))]
                res["cpu_baseline"] = {"single_thread_ms": round(1000 * rep["seconds"], 3), "cores": 1, "kind": "port",
                                       "sample": f"FP64 JavaScript restatement of the reference refit (oracle/js) in "
                                                 f"Node {rep['node']}, the whole tree, median of 3"}
                res["nodes_equal_to_cpu"] = bool(all(np.array_equal(a, b) for a, b in zip(first, js_roots)))
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return res


def bench_geometry(torch, M, ctx, stream, mesh, g, idx, tree):
    """closestPointToGeometry on the closest-point mesh: one whole query per call (the device entry point,
    events on the launch stream) for a separated placement (a 16k-triangle sphere in the knot's hole) and
    an interpenetrating one (an 80k-triangle knot shifted through it), against the oracle's tree walk on
    one host thread over the same query; the answers are compared."""
    import numpy as np
    from mbvh_amd import geometry as G
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    om = O.OracleMesh(g.positions, idx, tree.roots)
    out = {}
    for name, other, off in (("separated", G.sphere(1, 128, 64), (0.0, 0.0, 0.0)),
                             ("interpenetrating", G.torus_knot(10, 3, 400, 100), (29.5, 0.0, 0.0))):
        other.ensure_index()
        oi = other.index.astype(np.uint32)
        m = np.eye(4)
        m[:3, 3] = off
        e = np.ascontiguousarray(m.T.reshape(-1))
        pos = torch.from_numpy(other.positions.reshape(-1).copy()).to(dev)
        di = torch.from_numpy(oi.view(np.int32)).to(dev)
        rec = torch.zeros(M.GEOMETRY_HIT.itemsize, dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()

        def query():
            rc = ctx._lib.mbvh_closest_point_to_geometry_d(ctx.handle, mesh.handle, pos.data_ptr(),
                                                           other.positions.shape[0], 3, 0, di.data_ptr(), 4,
                                                           other.tri_count, e.ctypes.data, 0.0, float("inf"),
                                                           rec.data_ptr(), sptr)
            if rc:
                raise RuntimeError(f"closestPointToGeometry failed ({rc})")

        d_ms = event_ms(torch, stream, query, 20)
        r = np.frombuffer(rec.cpu().numpy().tobytes(), M.GEOMETRY_HIT)[0]
        t0 = time.perf_counter()
        ref = om.closest_to_geometry(other.positions, oi, e)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        same = ref is not None and abs(float(r["distance"]) - ref["distance"]) <= 1e-9 * max(ref["distance"], 1.0)
        out[name] = {"other_tris": other.tri_count, "gpu_device_ms": round(d_ms, 4), "distance": float(r["distance"]),
                     "cpu_oracle_ms_1thread": round(cpu_ms, 2), "speedup_vs_cpu": round(cpu_ms / d_ms, 1),
                     "same_distance_as_oracle": bool(same),
                     "same_pair_as_oracle": bool(ref is not None and int(r["face_index"]) == ref["face_index"] and
                                                 int(r["other_face_index"]) == ref["other_face_index"])}
    out["workload"] = (f"closestPointToGeometry on the closest-point mesh ({g.tri_count} tris): a sphere (radius 1, "
                       f"16,128 tris) in the knot's hole, and TorusKnot(10,3,400,100) shifted by 29.5 in x")
    return out


def bench_boxes(torch, ctx, stream, mesh, g, idx, tree, n=1 << 20):
    """intersectsBox on the closest-point mesh: n random oriented boxes (the device entry point, events on
    the launch stream) against the JavaScript restatement of the reference's shapecast with OrientedBox
    tests on one thread over a sample; the sample's answers are compared."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    lo, hi = g.bounds()
    rng = np.random.default_rng(21)
    ext = np.linalg.norm(hi - lo)
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y),
                  2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x),
                  2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], axis=1).reshape(n, 3, 3)
    mats = np.zeros((n, 16))
    mats[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]] = R.reshape(n, 9)
    mats[:, 12:15] = lo + (hi - lo) * rng.random((n, 3))
    mats[:, 15] = 1.0
    half = ext * rng.uniform(0.002, 0.03, (n, 3))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)
    dmn, dmx, dm = t(-half), t(half), t(mats)
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def boxes_d():
        rc = ctx._lib.mbvh_intersects_box_d(ctx.handle, mesh.handle, dmn.data_ptr(), dmx.data_ptr(), dm.data_ptr(), n,
                                            out.data_ptr(), sptr)
        if rc:
            raise RuntimeError(f"intersectsBox failed ({rc})")

    d_ms = event_ms(torch, stream, boxes_d, 10)
    got = out.cpu().numpy()
    res = {"workload": f"intersectsBox: {n} random oriented boxes (half extents 0.2-3% of the mesh diagonal) on "
                       f"the closest-point mesh ({g.tri_count} tris)", "gpu_device_ms": round(d_ms, 3),
           "mboxes_per_s": round(n / d_ms / 1e3, 2), "hit_fraction": round(float((got == 1).mean()), 3)}
    import shutil
    import subprocess
    import tempfile
    node = shutil.which("node")
    if node is not None:
        d = tempfile.mkdtemp(prefix="mbvh_box_")
        try:
            m = 4096
            ib = 2 if idx.dtype == np.uint16 else 4
            g.positions.tofile(os.path.join(d, "positions.f32"))
            idx.tofile(os.path.join(d, "index.u16" if ib == 2 else "index.u32"))
            for i, rt in enumerate(tree.roots):
                rt.tofile(os.path.join(d, f"root{i}.bin"))
            (-half[:m]).tofile(os.path.join(d, "box_min.f64"))
            half[:m].tofile(os.path.join(d, "box_max.f64"))
            mats[:m].tofile(os.path.join(d, "box_matrix.f64"))
            with open(os.path.join(d, "meta.json"), "w") as f:
                json.dump({"roots": len(tree.roots), "index_bytes": ib, "boxes": m}, f)
            r = subprocess.run([node, os.path.join(ROOT, "oracle", "js", "box_bench.js"), d], capture_output=True,
                               text=True, timeout=600)
            if r.returncode == 0:
                rep = json.loads(r.stdout.strip().splitlines()[-1])
                js = np.fromfile(os.path.join(d, "box_hit.u8"), np.uint8)
                res["cpu_baseline"] = {"value": round(m / rep["seconds"] / 1e6, 4), "unit": "Mboxes/s", "cores": 1,
                                       "kind": "port",
                                       "sample": f"the first {m} boxes: JavaScript restatement of the reference's "
                                                 f"intersectsBox (shapecast + OrientedBox, oracle/js) in Node "
                                                 f"{rep['node']}, one thread, median of 3"}
                res["identical_to_cpu_sample"] = bool(np.array_equal(js, got[:m]))
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return res


def bench_export(torch, ctx, stream, mesh, g, idx, tree):
    """MeshBVHUniformStruct texture data (bounds, contents, index, position) of the C4 mesh written on
    the device from the resident mesh (events on the launch stream), against a JavaScript restatement
    of the reference's packing loops on one thread; the arrays are compared."""
    import numpy as np
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    sb, sc, si, sp = mesh.texture_sizes()
    pos = torch.from_numpy(g.positions.reshape(-1)).to(dev)
    outs = [torch.empty((sb * sb, 4), device=dev), torch.empty((sc * sc, 2), dtype=torch.int32, device=dev),
            torch.empty((si * si, 4), dtype=torch.int32, device=dev), torch.empty((sp * sp, 4), device=dev)]
    torch.cuda.synchronize()
    lib = ctx._lib

    def export_d():
        rc = lib.mbvh_mesh_export_textures_d(ctx.handle, mesh.handle, pos.data_ptr(), 3, 0, *[o.data_ptr() for o in outs],
                                             sptr)
        if rc:
            raise RuntimeError(f"export failed ({rc})")

    d_ms = event_ms(torch, stream, export_d, 10)
    moved = sum(o.numel() * 4 for o in outs) + tree.roots[0].nbytes + 12 * g.tri_count + g.positions.nbytes
    res = {"workload": f"MeshBVHUniformStruct textures of the C4 mesh ({tree.roots[0].nbytes // 32} nodes, "
                       f"{g.tri_count} tris, {g.positions.shape[0]} vertices)",
           "gpu_device_ms": round(d_ms, 4), "achieved_gb_per_s": round(moved / (d_ms / 1000) / 1e9, 1)}
    import shutil
    import subprocess
    import tempfile
    node = shutil.which("node")
    if node is not None:
        d = tempfile.mkdtemp(prefix="mbvh_export_")
        try:
            ib = 2 if idx.dtype == np.uint16 else 4
            tree.roots[0].tofile(os.path.join(d, "root0.bin"))
            idx.tofile(os.path.join(d, "index.u16" if ib == 2 else "index.u32"))
            g.positions.tofile(os.path.join(d, "positions.f32"))
            with open(os.path.join(d, "meta.json"), "w") as f:
                json.dump({"index_bytes": ib}, f)
            out = subprocess.run([node, os.path.join(ROOT, "oracle", "js", "export_bench.js"), d], capture_output=True,
                                 text=True, timeout=600)
            if out.returncode == 0:
                rep = json.loads(out.stdout.strip().splitlines()[-1])
                got = [o.cpu().numpy() for o in outs]
                ref_b = np.fromfile(os.path.join(d, "bounds.f32"), np.uint32)
                ref_c = np.fromfile(os.path.join(d, "contents.u32"), np.uint32)
                ref_i = np.fromfile(os.path.join(d, "index_tex.bin"), np.uint16 if ib == 2 else np.uint32)
                ref_p = np.fromfile(os.path.join(d, "position_tex.f32"), np.uint32)
                same = (np.array_equal(got[0].view(np.uint32).reshape(-1), ref_b)
                        and np.array_equal(got[1].view(np.uint32).reshape(-1), ref_c)
                        and np.array_equal(got[2].view(np.uint32).reshape(-1), ref_i.astype(np.uint32))
                        and np.array_equal(got[3].view(np.uint32).reshape(-1), ref_p))
                res["cpu_baseline"] = {"single_thread_ms": round(1000 * rep["seconds"], 3), "cores": 1, "kind": "port",
                                       "sample": f"JavaScript restatement of bvhToTextures and VertexAttributeTexture."
                                                 f"updateFrom (oracle/js) in Node {rep['node']}, median of 3"}
                res["identical_to_cpu"] = bool(same)
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return res


def bench_build(torch, M, ctx, g, label):
    """MeshBVH construction (SAH, leaf 10) on the device: the device-pointer entry point (positions and
    index in HBM, wall clock of the call, which waits for the build) and the host entry point (arrays
    copied in, tree and permuted index copied out), against the host builder (the reference's build
    restated in C++, one thread). Both device results are compared with the host build byte for byte."""
    import numpy as np
    g = g.copy()
    g.ensure_index()
    idx_h = g.index.copy()
    t0 = time.perf_counter()
    th = M.build(g.positions, idx_h)
    host_s = time.perf_counter() - t0
    api = []
    for _ in range(4):
        idx_d = g.index.copy()
        t0 = time.perf_counter()
        td = M.build(g.positions, idx_d, ctx=ctx)
        api.append(time.perf_counter() - t0)
    same = all(np.array_equal(a, b) for a, b in zip(th.roots, td.roots)) and np.array_equal(idx_h, idx_d)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(g.positions.reshape(-1)).to(dev)
    idx_view = g.index.view(np.int32) if g.index.dtype == np.uint32 else g.index.view(np.int16)
    idx0 = torch.from_numpy(idx_view).to(dev)
    idx = idx0.clone()
    cap = (2 * g.tri_count + 1) * 32
    nodes = torch.empty(cap, dtype=torch.uint8, device=dev)
    dev_s = []
    counts = None
    for _ in range(6):
        idx.copy_(idx0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        counts = ctx.build_d(pos.data_ptr(), g.positions.shape[0], idx.data_ptr(), g.index.itemsize, g.index.size,
                             [(0, g.tri_count)], nodes.data_ptr(), cap)
        dev_s.append(time.perf_counter() - t0)
    same_d = bool(np.array_equal(nodes[:counts[0] * 32].cpu().numpy(), th.roots[0])
                  and np.array_equal(idx.cpu().numpy().view(g.index.dtype), idx_h))
    dev_ms = 1000 * float(np.median(dev_s[1:]))
    return {"workload": f"SAH build (leaf 10, max depth 40) of {label}: {g.tri_count} tris, "
                        f"{th.roots[0].nbytes // 32} nodes",
            "gpu_device_ms": round(dev_ms, 3), "gpu_mtris_per_s": round(g.tri_count / dev_ms / 1e3, 1),
            "gpu_host_api_ms": round(1000 * float(np.median(api[1:])), 3),
            "identical_to_cpu": bool(same and same_d),
            "cpu_baseline": {"single_thread_ms": round(1000 * host_s, 1), "cores": 1, "kind": "port",
                             "sample": "the host builder (mbvh_build: the reference's buildTree / getOptimalSplit / "
                                       "partition restated in C++), the whole mesh, one run"},
            "speedup_vs_cpu": round(1000 * host_s / dev_ms, 1),
            "note": "device: positions and index in HBM, wall clock of mbvh_build_gpu_d (which waits for the build); "
                    "host: pageable arrays in, node buffers and permuted index out"}


def bench_batched(torch, M, W, ctx, stream, args):
    """C5: one persistent batched launch (captured in a HIP graph, autotuned) vs the per-mesh loop."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    log("C5 setup")
    w = W.c5_batched(args.c5_geometries, args.c5_rays)
    geos = [g.copy() for g in w.geometries]
    for g in geos:
        g.ensure_index()
    idxs = [g.index.copy() for g in geos]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        trees = list(ex.map(lambda k: M.build(geos[k].positions, idxs[k], strategy=M.SAH, target_leaf_size=10),
                            range(len(geos))))
    meshes = [ctx.upload(trees[k], geos[k].positions, idxs[k]) for k in range(len(geos))]
    batch = ctx.batch(meshes)
    n = w.origins.shape[0]
    n_inst = w.matrix_world.shape[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    D = {"world": t(w.matrix_world), "inverse": t(w.matrix_world_inverse),
         "instance_geometry": t(w.instance_geometry.view(np.int32)), "origins": t(w.origins),
         "directions": t(w.directions), "ray_instance": t(w.ray_instance.view(np.int32))}
    wsb = batch.workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    out = {"face": torch.empty(n, dtype=torch.int32, device=dev), "distance": torch.empty(n, device=dev),
           "point": torch.empty((n, 3), device=dev), "barycoord": torch.empty((n, 3), device=dev)}
    p = {k: v.data_ptr() for k, v in D.items()}
    p.update({k: v.data_ptr() for k, v in out.items()})
    p.update(workspace=ws.data_ptr(), workspace_bytes=wsb, n_instances=n_inst, n=n)
    log("C5 autotune")
    tuning = batch.autotune(p)
    graph = batch.graph(p)
    b_ms = event_ms(torch, stream, lambda: graph.launch(sptr), 10)
    e_ms = event_ms(torch, stream, lambda: batch.raycast_first_d(p, stream=sptr), 10)
    res_b = {k: v.clone() for k, v in out.items()}
    # the per-mesh plain loop: instance rays, one raycastFirst launch per mesh on its rays (grouped
    # once, outside the timed region), world hits
    lo, ld = torch.empty((n, 3), device=dev), torch.empty((n, 3), device=dev)
    geo = w.instance_geometry[w.ray_instance]
    order = np.argsort(geo, kind="stable")
    starts = np.concatenate([[0], np.cumsum(np.bincount(geo, minlength=len(meshes)))])
    gidx = t(order.astype(np.int64))
    so = {k: torch.empty_like(v) for k, v in out.items()}
    slo, sld = torch.empty_like(lo), torch.empty_like(ld)

    def loop():
        ctx.instance_rays_d(D["inverse"].data_ptr(), n_inst, D["origins"].data_ptr(), D["directions"].data_ptr(),
                            D["ray_instance"].data_ptr(), n, lo.data_ptr(), ld.data_ptr(), stream=sptr)
        torch.index_select(lo, 0, gidx, out=slo)
        torch.index_select(ld, 0, gidx, out=sld)
        for gi, mesh in enumerate(meshes):
            a, b = int(starts[gi]), int(starts[gi + 1])
            if a < b:
                mesh.raycast_first_d(slo[a:].data_ptr(), sld[a:].data_ptr(), b - a, so["face"][a:].data_ptr(),
                                     so["distance"][a:].data_ptr(), so["point"][a:].data_ptr(),
                                     so["barycoord"][a:].data_ptr(), stream=sptr)
        for k in out:
            out[k].index_copy_(0, gidx, so[k])
        ctx.instance_hits_d(D["world"].data_ptr(), n_inst, D["origins"].data_ptr(), D["ray_instance"].data_ptr(), n,
                            out["face"].data_ptr(), out["point"].data_ptr(), out["distance"].data_ptr(), stream=sptr)

    ctx.set_ray_lanes(tuning.ray_lanes)  # the loop traces with the same traversal as the tuned launch
    l_ms = event_ms(torch, stream, loop, 3)
    same = all(torch.equal(out[k].view(torch.int32), res_b[k].view(torch.int32)) for k in out)
    ctx.set_ray_lanes(0)  # and the fastest plain loop: per-mesh automatic lanes (equal up to ties)
    la_ms = event_ms(torch, stream, loop, 3)
    tris = int(sum(g.tri_count for g in geos))
    r = {"workload": f"{len(meshes)} meshes (torus knots / spheres, {tris} tris, 100..200k each), one rigid "
                     f"instance each, {n} rays split by Zipf(1.1)",
         "batched_graph_ms": round(b_ms, 3), "batched_eager_ms": round(e_ms, 3), "plain_loop_ms": round(l_ms, 3),
         "speedup_vs_loop": round(l_ms / b_ms, 2), "plain_loop_auto_lanes_ms": round(la_ms, 3),
         "speedup_vs_fastest_loop": round(la_ms / b_ms, 2), "batched_mrays_per_s": round(n / b_ms / 1e3, 2),
         "identical_to_loop": bool(same), "tuning": tuning.as_dict()}
    log(f"C5 {r}")
    graph.free()
    scene = bench_scene(torch, M, W, ctx, stream, args, meshes, batch, geos, idxs, trees, w)
    log(f"scene {scene}")
    batch.free()
    for m in meshes:
        m.free()
    return {"batched_raycast_first": r, "scene_raycast_first": scene}


def bench_scene_points(torch, M, ctx, stream, args, sc, geos, idxs, trees, w, n=1 << 20):
    """Scene closest point (world space, one launch) for n world points near random instances of the C5
    scene, against the FP64 oracle's closest point run once per instance on a sample (the instances are
    rigid, so each instance's local answer is its world answer); the sample's distances are compared."""
    import numpy as np
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(29)
    tr = w.matrix_world[:, 12:15].astype(np.float64)
    pts = (tr[rng.integers(0, tr.shape[0], n)] + rng.normal(0, 1.5, (n, 3))).astype(np.float32)
    q = torch.from_numpy(pts).to(dev)
    face = torch.empty(n, dtype=torch.int32, device=dev)
    inst = torch.empty(n, dtype=torch.int32, device=dev)
    dist = torch.empty(n, device=dev)
    torch.cuda.synchronize()

    order = torch.empty(n, dtype=torch.int32, device=dev)
    owsb = M.Context.order_workspace_bytes(n)
    ows = torch.empty(owsb, dtype=torch.uint8, device=dev)

    def run():  # the spatial visit order, then the query (as the host entry point does)
        ctx.order_points_d(q.data_ptr(), n, order.data_ptr(), ows.data_ptr(), owsb, stream=stream.cuda_stream)
        sc.closest_point_d(q.data_ptr(), n, face.data_ptr(), inst.data_ptr(), dist.data_ptr(), stream=stream.cuda_stream,
                           order_ptr=order.data_ptr())

    ms_gpu = event_ms(torch, stream, run, 10)
    res = {"workload": f"closest point over the {w.matrix_world.shape[0]} instances of the same scene, {n} world "
                       "points within a few units of random instances", "scene_ms": round(ms_gpu, 3),
           "mqueries_per_s": round(n / ms_gpu / 1e3, 2)}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        oms = [O.OracleMesh(geos[k].positions, idxs[k], trees[k].roots) for k in range(len(geos))]
        m = 64
        threads = min(16, os.cpu_count() or 1)
        best = np.full(m, np.inf)
        t0 = time.perf_counter()
        for i in range(w.matrix_world.shape[0]):
            mi = w.matrix_world_inverse[i].astype(np.float64).reshape(4, 4).T
            mw = w.matrix_world[i].astype(np.float64).reshape(4, 4).T
            lp = (pts[:m] @ mi[:3, :3].T + mi[:3, 3]).astype(np.float32)
            r = oms[w.instance_geometry[i]].closest_point(lp, threads=threads)
            wp = r["point"] @ mw[:3, :3].T + mw[:3, 3]
            best = np.minimum(best, np.linalg.norm(wp - pts[:m].astype(np.float64), axis=1))
        cpu_s = time.perf_counter() - t0
        gd = dist[:m].cpu().numpy().astype(np.float64)
        res["cpu_baseline"] = {"value": round(m / cpu_s / 1e6, 5), "unit": "Mqueries/s", "cores": threads,
                               "kind": "port", "sample": f"{m} points x {w.matrix_world.shape[0]} instances: the FP64 "
                               "C oracle's closestPointToPoint once per instance, the least world distance"}
        res["sample_max_rel_diff_vs_oracle"] = float(np.max(np.abs(gd - best) / np.maximum(best, 1.0)))
    return res


def bench_scene(torch, M, W, ctx, stream, args, meshes, batch, geos, idxs, trees, w):
    """Scene raycast (closest hit over all instances, one launch over a top-level tree) on the C5 meshes
    and placements, against the batched launch on every (ray, instance) pair reduced to the closest
    (what a loop over the instances computes), and against the FP64 oracle's per-instance loop on the
    host (the reference's per-instance traversal, one call per instance)."""
    import numpy as np
    dev = torch.device("cuda:0")
    sptr = stream.cuda_stream
    rng = np.random.default_rng(17)
    n = args.scene_rays
    o, d = W.rays_through_instances(geos, w.matrix_world, w.instance_geometry, n, rng)
    sc = ctx.scene(meshes, w.matrix_world, w.matrix_world_inverse, w.instance_geometry)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    do, dd = t(o), t(d)
    face = torch.empty(n, dtype=torch.int32, device=dev)
    inst = torch.empty(n, dtype=torch.int32, device=dev)
    dist = torch.empty(n, device=dev)
    point = torch.empty((n, 3), device=dev)
    bary = torch.empty((n, 3), device=dev)
    torch.cuda.synchronize()

    def scene_d():
        sc.raycast_first_d(do.data_ptr(), dd.data_ptr(), n, face.data_ptr(), inst.data_ptr(), dist.data_ptr(),
                           point.data_ptr(), bary.data_ptr(), stream=sptr)

    s_ms = event_ms(torch, stream, scene_d, 10)
    # every (ray, instance) pair of a subsample through the batched launch, then the closest per ray
    n_inst = w.matrix_world.shape[0]
    m = max(1, min(n, (1 << 21) // n_inst))
    ri = np.repeat(np.arange(n_inst, dtype=np.uint32), m)
    P = {"world": t(w.matrix_world), "inverse": t(w.matrix_world_inverse),
         "instance_geometry": t(w.instance_geometry.view(np.int32)), "origins": t(np.tile(o[:m], (n_inst, 1))),
         "directions": t(np.tile(d[:m], (n_inst, 1))), "ray_instance": t(ri.view(np.int32))}
    npair = n_inst * m
    wsb = batch.workspace_bytes(npair)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    pout = {"face": torch.empty(npair, dtype=torch.int32, device=dev), "distance": torch.empty(npair, device=dev),
            "point": torch.empty((npair, 3), device=dev), "barycoord": torch.empty((npair, 3), device=dev)}
    p = {k: v.data_ptr() for k, v in P.items()}
    p.update({k: v.data_ptr() for k, v in pout.items()})
    p.update(workspace=ws.data_ptr(), workspace_bytes=wsb, n_instances=n_inst, n=npair)
    sel = {}

    def pairwise():
        batch.raycast_first_d(p, stream=sptr)
        dm = pout["distance"].view(n_inst, m)
        miss = pout["face"].view(n_inst, m) == -1
        dm = torch.where(miss, torch.full_like(dm, float("inf")), dm)
        sel["best"] = torch.argmin(dm, dim=0)  # first minimum: the lowest instance on ties
        sel["dm"] = dm

    pw_ms = event_ms(torch, stream, pairwise, 3)
    best = sel["best"].cpu().numpy()
    fm = pout["face"].view(n_inst, m).cpu().numpy().view(np.uint32)
    dmn = sel["dm"].cpu().numpy()
    cols = np.arange(m)
    pf = fm[best, cols]
    hit = pf != M.NO_HIT
    sf = face[:m].cpu().numpy().view(np.uint32)
    si = inst[:m].cpu().numpy().view(np.uint32)
    sdist = dist[:m].cpu().numpy()
    same = bool(np.array_equal(sf, np.where(hit, pf, M.NO_HIT)) and np.array_equal(si[hit], best[hit].astype(np.uint32))
                and np.array_equal(sdist[hit].view(np.uint32), dmn[best, cols][hit].astype(np.float32).view(np.uint32)))
    pw_per_ray_ms = pw_ms / m
    res = {"workload": f"closest first hit over {n_inst} instances (the C5 meshes and placements, {len(meshes)} "
                       f"meshes, {sum(g.tri_count for g in geos)} tris), {n} world rays from outside the scene towards "
                       "random instances",
           "scene_ms": round(s_ms, 3), "scene_mrays_per_s": round(n / s_ms / 1e3, 2),
           "hit_fraction": round(float((face.cpu().numpy().view(np.uint32) != M.NO_HIT).mean()), 3),
           "pairwise_ms_for_all_rays": round(pw_per_ray_ms * n, 1),
           "speedup_vs_pairwise": round(pw_per_ray_ms * n / s_ms, 1),
           "pairwise_sample_rays": m, "identical_to_pairwise": same,
           "note": "pairwise = the batched launch on every (ray, instance) pair of the first rays, then the "
                   "closest per ray (torch argmin), scaled to all rays"}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        oms = [O.OracleMesh(geos[k].positions, idxs[k], trees[k].roots) for k in range(len(geos))]
        ms = min(n, 256)
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        for i in range(n_inst):
            mi = w.matrix_world_inverse[i].astype(np.float64).reshape(4, 4).T
            lo = (o[:ms] @ mi[:3, :3].T + mi[:3, 3]).astype(np.float32)
            ld = (d[:ms] @ mi[:3, :3].T).astype(np.float32)
            oms[w.instance_geometry[i]].raycast_first(lo, ld, threads=threads)
        cpu_s = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(ms / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"{ms} rays x {n_inst} instances: the FP64 C oracle's restatement of the "
                                         "reference raycastFirst traversal, one call per instance over the sample "
                                         "(the per-instance loop of acceleratedBatchedMeshRaycast)"}
    res["closest_point"] = bench_scene_points(torch, M, ctx, stream, args, sc, geos, idxs, trees, w)
    sc.free()
    return res


def bench_host_api(mesh, mesh1, w, w1, n, nr):
    """Host-pointer entry points (the JS path): pageable host arrays in, results back, PCIe and the
    query ordering included."""
    import numpy as np
    mesh.closest_point(w.points)
    t0 = time.perf_counter()
    for _ in range(3):
        mesh.closest_point(w.points)
    host_cp = n / ((time.perf_counter() - t0) / 3) / 1e6
    mesh1.raycast_first(w1.origins, w1.directions)
    t0 = time.perf_counter()
    for _ in range(10):
        mesh1.raycast_first(w1.origins, w1.directions)
    host_rc = nr / ((time.perf_counter() - t0) / 10) / 1e6
    radii = np.full(n, 0.5, np.float32)  # about half of the grid points reach the surface
    mesh.intersects_sphere(w.points, radii)
    t0 = time.perf_counter()
    hit_frac = float(mesh.intersects_sphere(w.points, radii).mean())
    host_sp = n / (time.perf_counter() - t0) / 1e6
    host_api = {"closest_point_mqueries_per_s": round(host_cp, 2), "raycast_first_mrays_per_s": round(host_rc, 2),
                "intersects_sphere_mspheres_per_s": round(host_sp, 2), "intersects_sphere_hit_fraction": round(hit_frac, 3),
                "note": "host-pointer entry points (the JS path): pageable host arrays in and out, copies over "
                        "PCIe and the query ordering included"}
    return host_api


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md but not launched)"}))
        sys.exit(2)
    import numpy as np
    import torch
    import mbvh_amd as M
    from mbvh_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # A dedicated (default-priority) stream: kernels and timing events share it. torch's default
    # stream has handle 0, which the C ABI reads as "the context's own stream".
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream

    if args.only_geometry:
        from mbvh_amd import geometry as G
        ctx = M.Context(0)
        g = G.torus_knot(10, 3, 800, 200).ensure_index()
        idx = g.index.copy()
        tree = M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)
        mesh = ctx.upload(tree, g.positions, idx)
        print(json.dumps({"closest_point_to_geometry": bench_geometry(torch, M, ctx, stream, mesh, g, idx, tree)}),
              flush=True)
        return

    if args.only_differentiable:
        ctx = M.Context(0)
        print(json.dumps(bench_differentiable(torch, M, W, ctx, stream, args)), flush=True)
        return

    if args.only_scene:
        from concurrent.futures import ThreadPoolExecutor
        ctx = M.Context(0)
        w = W.c5_batched(args.c5_geometries, 1 << 10)
        geos = [g.copy() for g in w.geometries]
        for g in geos:
            g.ensure_index()
        idxs = [g.index.copy() for g in geos]
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            trees = list(ex.map(lambda k: M.build(geos[k].positions, idxs[k], strategy=M.SAH, target_leaf_size=10),
                                range(len(geos))))
        meshes = [ctx.upload(trees[k], geos[k].positions, idxs[k]) for k in range(len(geos))]
        batch = ctx.batch(meshes)
        print(json.dumps({"scene": bench_scene(torch, M, W, ctx, stream, args, meshes, batch, geos, idxs, trees, w)}),
              flush=True)
        return

    if args.only_build:
        from mbvh_amd import geometry as G
        ctx = M.Context(0)
        print(json.dumps({"build": bench_build(torch, M, ctx, G.displaced_icosphere(223),
                                               "the C4 mesh (displaced Icosahedron(1,223))")}), flush=True)
        return

    # ---------------- closest point (primary) ----------------
    w = W.c2_closest(args.grid)
    g = w.geometry
    idx = g.index.copy()
    tree = M.build(g.positions, idx, strategy=M.SAH, target_leaf_size=10)
    ctx = M.Context(0)
    ctx.upload(tree, g.positions, idx).free()  # first upload: allocator and module warm-up
    t_up = time.perf_counter()
    mesh = ctx.upload(tree, g.positions, idx)
    ctx.synchronize()
    upload_ms = (time.perf_counter() - t_up) * 1e3
    # the grid's queries are coherent: packets of 64 neighbouring queries walk the tree together
    ctx.set_point_lanes(64)
    n = w.points.shape[0]
    d_q = torch.from_numpy(w.points).to(dev)
    d_face = torch.empty(n, dtype=torch.int32, device=dev)
    d_dist = torch.empty(n, dtype=torch.float32, device=dev)
    d_pt = torch.empty((n, 3), dtype=torch.float32, device=dev)
    d_bc = torch.empty((n, 3), dtype=torch.float32, device=dev)
    # query order (Morton buckets), recomputed inside every step
    d_order = torch.empty(n, dtype=torch.int32, device=dev)
    ws_bytes = M.Context.order_workspace_bytes(n)
    d_ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def order_cp():
        if not args.no_order:
            ctx.order_points_d(d_q.data_ptr(), n, d_order.data_ptr(), d_ws.data_ptr(), ws_bytes, stream=sptr)

    def launch_cp():
        mesh.closest_point_d(d_q.data_ptr(), n, d_face.data_ptr(), d_dist.data_ptr(), d_pt.data_ptr(),
                             d_bc.data_ptr(), stream=sptr, order_ptr=0 if args.no_order else d_order.data_ptr())

    wall, kern = timed_kernel(torch, launch_cp, args.steps, args.warmup, stream, pre=order_cp)
    ms_per_step = wall * 1000.0 / args.steps
    value = n / (wall / args.steps) / 1e6
    kern_ms = float(np.mean(kern))
    # the same launch with four lanes per query (the default for scattered queries): identical results
    packet_out = [x.cpu().numpy().view(np.uint32).copy() for x in (d_face, d_dist, d_pt, d_bc)]
    ctx.set_point_lanes(4)
    _, kern4 = timed_kernel(torch, launch_cp, max(3, args.steps // 5), 2, stream, pre=order_cp)
    quad_same = all(np.array_equal(a, x.cpu().numpy().view(np.uint32))
                    for a, x in zip(packet_out, (d_face, d_dist, d_pt, d_bc)))
    ctx.set_point_lanes(64)
    launch_cp()
    torch.cuda.synchronize()

    # ---------------- raycastFirst (secondary) ----------------
    w1 = W.c1_raycast()
    g1 = w1.geometry
    idx1 = g1.index.copy()
    tree1 = M.build(g1.positions, idx1, strategy=M.SAH, target_leaf_size=10)
    t_up = time.perf_counter()
    mesh1 = ctx.upload(tree1, g1.positions, idx1)
    ctx.synchronize()
    upload1_ms = (time.perf_counter() - t_up) * 1e3
    nr = w1.origins.shape[0]
    d_o = torch.from_numpy(w1.origins).to(dev)
    d_d = torch.from_numpy(w1.directions).to(dev)
    r_face = torch.empty(nr, dtype=torch.int32, device=dev)
    r_dist = torch.empty(nr, dtype=torch.float32, device=dev)
    r_pt = torch.empty((nr, 3), dtype=torch.float32, device=dev)
    r_bc = torch.empty((nr, 3), dtype=torch.float32, device=dev)

    # rays in their given order: origin-ordering does not make these incoherent rays (random
    # origins, random targets) share paths, and the extra pass costs more than it saves
    def launch_rc():
        mesh1.raycast_first_d(d_o.data_ptr(), d_d.data_ptr(), nr, r_face.data_ptr(), r_dist.data_ptr(),
                              r_pt.data_ptr(), r_bc.data_ptr(), stream=sptr)

    wall1, kern1 = timed_kernel(torch, launch_rc, max(args.steps, 20) * 4, args.warmup, stream)
    rc_steps = max(args.steps, 20) * 4
    mrays = nr / (wall1 / rc_steps) / 1e6
    mrays_kernel = nr / (float(np.mean(kern1)) / 1000.0) / 1e6

    # ---------------- host-pointer entry points (PCIe included) ----------------
    # what the JS API sees: pageable host arrays in, results back in host arrays, ordering included
    # (skipped with --only-primary, whose profiles then hold the primary launches alone)
    host_api = None
    if not args.only_primary:
        host_api = bench_host_api(mesh, mesh1, w, w1, n, nr)

    # ---------------- oracle counts + CPU baseline ----------------
    # ---------------- oracle counts + CPU baseline ----------------
    cpu = None
    roof = None
    extra = {}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O

        om = O.OracleMesh(g.positions, idx, tree.roots)
        sample = w.points[:: args.cpu_sample]
        threads = min(16, os.cpu_count() or 1)  # this host's CPU share for one GPU
        t0 = time.perf_counter()
        ref = om.closest_point(sample, threads=threads)
        cpu_s = time.perf_counter() - t0
        nodes = float(ref["nodes"].mean())
        tris = float(ref["tris"].mean())
        bq = 32.0 + 32.0 * nodes + (3 * IDX_BYTES + 36) * tris
        achieved = n * bq / (kern_ms / 1000.0) / 1e9
        traffic = profiled_value("traffic_bytes_per_launch")
        valu = profiled_value("valu_insts_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": round(traffic) if traffic else None,
                "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                "algorithmic_bytes_per_launch": round(n * bq),
                "bytes_per_query": round(bq, 1), "nodes_per_query": round(nodes, 2),
                "tris_per_query": round(tris, 2), "kernel_ms": round(kern_ms, 4),
                "kernel": "k_closest_point_packet"}
        if valu:
            # the measured limiter: VALU issue (HBM traffic is a small fraction of the algorithmic
            # bytes: node and triangle reads hit in the caches)
            valu_ms = valu * 4 / (SIMDS * CLOCK_GHZ * 1e9) * 1e3
            roof["limiter"] = "VALU issue"
            roof["valu_insts_per_launch"] = round(valu)
            roof["valu_issue_ms"] = round(valu_ms, 3)
            roof["valu_issue_frac"] = round(valu_ms / kern_ms, 3)
        # error of every result against the FP64 oracle on the sample (faces: the parity rule's
        # rounding ties aside; point and barycoord where the faces agree)
        gsel = slice(None, None, args.cpu_sample)
        gface = d_face.cpu().numpy().view(np.uint32)[gsel]
        gdist = d_dist.cpu().numpy()[gsel]
        gpt = d_pt.cpu().numpy()[gsel]
        gbc = d_bc.cpu().numpy()[gsel]
        eq = gface == ref["face"]
        tri = idx.reshape(-1, 3).astype(np.int64)[ref["face"][eq].astype(np.int64)]
        pos64 = g.positions.reshape(-1, 3).astype(np.float64)
        rbc = barycoords64(ref["point"][eq], pos64[tri[:, 0]], pos64[tri[:, 1]], pos64[tri[:, 2]])
        errors_c2 = {"queries": int(sample.shape[0]), "face_equal": round(float(eq.mean()), 6),
                     "max_distance_rel_err": max_rel(gdist, ref["distance"]),
                     "max_point_rel_err": max_rel(gpt[eq], ref["point"][eq]),
                     "max_barycoord_abs_err": float(np.max(np.abs(gbc[eq] - rbc))) if eq.any() else 0.0,
                     "note": "against the FP64 C oracle on every cpu-sample query; point / barycoord where the "
                             "faces agree (elsewhere the two faces are at the same distance to rounding)"}
        oracle_rate = sample.shape[0] / cpu_s / 1e6
        # the reference side in its own language: the JS restatement in Node, single thread over a
        # strided sample, then a worker pool over every query
        single = w.points[:: max(1, n // 32768)][:32768]
        log("CPU baseline (Node)")
        js = js_cpu_baseline("closest", g.positions, idx, tree.roots, np.concatenate([single, w.points]), threads,
                             single.shape[0], min_threshold=0.0, max_threshold=float("inf"), side=0, near=0.0,
                             far=float("inf"))
        if js is not None:
            gf = d_face.cpu().numpy().view(np.uint32)[:: max(1, n // 32768)][:32768]
            # where the faces differ, both must be at the same FP64 distance (an exact tie, e.g. a
            # shared vertex: the reference keeps the first in its visit order)
            tie_ok = np.ones(single.shape[0], bool)
            for i in np.flatnonzero(js["faces"] != gf):
                da = om.point_face_distance(int(js["faces"][i]), single[i])
                db = om.point_face_distance(int(gf[i]), single[i])
                tie_ok[i] = abs(da - db) <= 1e-6 * max(da, 1e-30)
            cpu = {"value": round(js["pool"]["queries_per_s"] / 1e6, 4),
                   "unit": "Mqueries/s", "cores": threads, "kind": "port",
                   "sample": f"FP64 JavaScript restatement of the reference closestPointToPoint (oracle/js) in "
                             f"Node {js['node']} on {js['cpu']}: worker pool of {threads} over all {n} grid points "
                             f"and the {single.shape[0]} single-thread ones, median of 3 rounds after a prewarm",
                   "single_thread_mqueries_per_s": round(js["single"]["queries_per_s"] / 1e6, 5),
                   "single_thread_sample": f"{single.shape[0]} points (every {max(1, n // 32768)}th)",
                   "face_agreement_with_gpu": round(float((js["faces"] == gf).mean()), 6),
                   "face_or_exact_tie_agreement_with_gpu": round(float(tie_ok.mean()), 6),
                   "oracle_c_mqueries_per_s": round(oracle_rate, 4)}
        else:
            cpu = {"value": round(oracle_rate, 4), "unit": "Mqueries/s", "cores": threads, "kind": "port",
                   "sample": f"every {args.cpu_sample}th grid point ({sample.shape[0]} queries), FP64 C oracle "
                             f"restatement of the reference traversal, {threads} threads (Node absent)"}
        # raycast parity spot check + counts for the secondary line
        om1 = O.OracleMesh(g1.positions, idx1, tree1.roots)
        r1 = om1.raycast_first(w1.origins, w1.directions, threads=threads)
        gf = r_face.cpu().numpy().view(np.uint32)
        extra["raycast_face_agreement"] = float((gf == r1["face"]).mean())
        hit = (gf == r1["face"]) & (r1["face"] != O.NO_HIT)
        extra["errors_vs_oracle"] = {
            "rays": int(nr), "hits": int(hit.sum()),
            "max_distance_rel_err": max_rel(r_dist.cpu().numpy()[hit], r1["distance"][hit]),
            "max_point_rel_err": max_rel(r_pt.cpu().numpy()[hit], r1["point"][hit]),
            "max_barycoord_abs_err": float(np.max(np.abs(r_bc.cpu().numpy()[hit] - r1["barycoord"][hit])))}
        extra["raycast_nodes_per_ray"] = round(float(r1["nodes"].mean()), 2)
        extra["raycast_tris_per_ray"] = round(float(r1["tris"].mean()), 2)
        # algorithmic bytes per ray: 56 B of I/O, 32 B per node box, 42 B per triangle (3 x u16 + 36 B)
        b_ray = 56.0 + 32.0 * float(r1["nodes"].mean()) + 42.0 * float(r1["tris"].mean())
        extra["raycast_bytes_per_ray"] = round(b_ray, 1)
        extra["raycast_achieved_gbs"] = round(nr * b_ray / (float(np.mean(kern1)) / 1000.0) / 1e9, 1)
        js1 = js_cpu_baseline("raycast", g1.positions, idx1, tree1.roots, np.concatenate([w1.origins, w1.directions]),
                              threads, nr, min_threshold=0.0, max_threshold=float("inf"), side=0, near=0.0,
                              far=float("inf"))
        if js1 is not None:
            extra["cpu_baseline"] = {
                "single_thread_mrays_per_s": round(js1["single"]["queries_per_s"] / 1e6, 5),
                "pool_mrays_per_s": round(js1["pool"]["queries_per_s"] / 1e6, 4), "cores": threads,
                "kind": "port", "sample": f"all {nr} rays, FP64 JavaScript restatement of the reference "
                                          f"raycastFirst (oracle/js) in Node {js1['node']}",
                "face_agreement_with_gpu": round(float((js1["faces"] == gf).mean()), 6)}
    line = {
        "metric": "Mrays/s raycastFirst, Mqueries/s closestPoint, fwd+bwd/fwd ratio, 1 MI355X",
        "value": round(value, 3),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"closestPointToPoint TorusKnot(10,3,800,200) 320k tris SAH leaf10, "
                               f"{args.grid}^3 grid over AABB+10%", "queries": n,
                   "query_order": "given" if args.no_order else "Morton buckets (in the step)",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "closest_point": {"mode": "packets of 64 queries (mbvh_ctx_set_point_lanes 64)",
                          "kernel_ms": round(kern_ms, 4),
                          "quad_mode_kernel_ms": round(float(np.mean(kern4)), 4),
                          "quad_results_identical": quad_same,
                          "errors_vs_oracle": errors_c2 if not args.no_cpu_baseline else None},
        "upload": {"closest_mesh_ms": round(upload_ms, 3), "raycast_mesh_ms": round(upload1_ms, 3),
                   "note": "mbvh_mesh_upload wall time: reference node buffers, index and positions in; the "
                           "binary-wide and 4-wide layouts derived on the host and copied to HBM"},
        "host_api": host_api,
        "raycast_first": {"workload": "TorusKnot(10,3,400,100) 80k tris SAH leaf10, 65,536 rays from r=40",
                          "mrays_per_s": round(mrays, 2), "mrays_per_s_kernel_only": round(mrays_kernel, 2),
                          "kernel_ms": round(float(np.mean(kern1)), 4), **extra},
    }
    if not args.only_primary:
        rng = np.random.default_rng(5)
        lo1, hi1 = g1.bounds()
        pts1 = (lo1 + (hi1 - lo1) * rng.random((4096, 3))).astype(np.float32)
        pc = percall_latency(g1.positions, idx1, np.concatenate([w1.origins[:4096], w1.directions[:4096]]), pts1)
        if pc is not None:
            js_single = (extra.get("cpu_baseline") or {}).get("single_thread_mrays_per_s")
            line["per_call"] = {
                "workload": "TorusKnot(10,3,400,100) 80k tris: one C1 ray per MeshBVH.raycastFirst call, one point "
                            "in the mesh box per closestPointToPoint call, from Node through the addon",
                **pc,
                "reference_js_us_per_ray_single_thread": round(1.0 / js_single, 3) if js_single else None,
                "note": "each call is a synchronous device round trip; the batch methods amortise it"}
            log(f"per-call {line['per_call']}")
        line["intersects_box"] = bench_boxes(torch, ctx, stream, mesh, g, idx, tree)
        log(f"boxes {line['intersects_box']}")
        line["closest_point_to_geometry"] = bench_geometry(torch, M, ctx, stream, mesh, g, idx, tree)
        log(f"geometry {line['closest_point_to_geometry']}")
        line["build_closest_mesh"] = bench_build(torch, M, ctx, g, "the closest-point mesh TorusKnot(10,3,800,200)")
        log(f"build {line['build_closest_mesh']}")
        line.update(bench_differentiable(torch, M, W, ctx, stream, args))
        line.update(bench_batched(torch, M, W, ctx, stream, args))
    print(json.dumps(line), flush=True)
    if args.traffic_calibration:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    mesh.free()
    mesh1.free()
    ctx.close()


if __name__ == "__main__":
    main()
