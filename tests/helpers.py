"""Shared test helpers: building meshes, running the oracle, and the parity rules.

Parity rules (stated once, used by every GPU parity test):
  face index   equal to the oracle's, except ties. Rays: the oracle's FP64 distance along the ray to
               the GPU's face's plane is within TIE_REL of the oracle's best distance. Points:
               rounding ties only (point_tie_band): the two faces' FP64 squared distances differ by
               less than the kernel's double-float error bound for the two plus the FP64 oracle's own
               rounding, about 1e-12 of the squared distance or less.
  distance, point, barycoord
               |gpu - oracle| <= VALUE_REL * max(|oracle|, 1) per component (barycoords only where
               the faces are equal).
"""
from __future__ import annotations

import numpy as np

import mbvh_amd as M
import oracle_py as O

TIE_REL = 1e-6
VALUE_REL = 1e-5


def build_mesh(geometry, strategy=M.SAH, ranges=None, indirect=False, max_depth=40, leaf=10):
    g = geometry.copy()
    if not indirect:
        g.ensure_index()
    idx = None if g.index is None else g.index.copy()
    tree = M.build(g.positions, idx, ranges=ranges, strategy=strategy, indirect=indirect,
                   max_depth=max_depth, target_leaf_size=leaf)
    return g, idx, tree


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1.0)


def check_raycast_parity(om, origins, dirs, got, ref, label=""):
    """Returns a dict of statistics; asserts the parity rules."""
    gf, rf = got["face"], ref["face"]
    both_miss = (gf == M.NO_HIT) & (rf == O.NO_HIT)
    same = gf == rf
    diff_idx = np.nonzero(~same)[0]
    ties = 0
    bad = []
    for i in diff_idx:
        if gf[i] == M.NO_HIT or rf[i] == O.NO_HIT:
            bad.append(i)
            continue
        dplane = om.ray_plane_distance(int(gf[i]), origins[i], dirs[i])
        if abs(dplane - ref["distance"][i]) <= TIE_REL * max(ref["distance"][i], 1.0):
            ties += 1
        else:
            bad.append(i)
    assert not bad, (f"{label}: {len(bad)} face mismatches beyond ties, e.g. ray {bad[0]}: gpu {gf[bad[0]]} "
                     f"oracle {rf[bad[0]]} dist gpu {got['distance'][bad[0]]} oracle {ref['distance'][bad[0]]}")
    hit = (gf != M.NO_HIT) & (rf != O.NO_HIT)
    stats = {"rays": len(gf), "hits": int(hit.sum()), "misses": int(both_miss.sum()), "ties": ties,
             "face_equal": float(same.mean())}
    if hit.any():
        de = rel_err(got["distance"][hit], ref["distance"][hit])
        pe = rel_err(got["point"][hit], ref["point"][hit])
        stats["max_dist_err"] = float(de.max())
        stats["max_point_err"] = float(pe.max())
        assert de.max() <= VALUE_REL, f"{label}: distance error {de.max():.3e}"
        assert pe.max() <= VALUE_REL, f"{label}: point error {pe.max():.3e}"
        eq = hit & same
        if eq.any():
            be = rel_err(got["barycoord"][eq], ref["barycoord"][eq])
            stats["max_bary_err"] = float(be.max())
            assert be.max() <= VALUE_REL, f"{label}: barycoord error {be.max():.3e}"
    miss_ok = np.all(np.isinf(got["distance"][gf == M.NO_HIT]))
    assert miss_ok, f"{label}: a miss did not report distance +inf"
    return stats


def _face_vertices(om, face):
    f = int(face)
    v = om.index[3 * f:3 * f + 3].astype(np.int64) if om.index is not None else np.arange(3 * f, 3 * f + 3)
    return om.positions.reshape(-1, 3)[v].astype(np.float64)


def point_tie_band(om, p, face_a, face_b):
    """The largest difference of two faces' squared distances from p that is a rounding tie.

    The kernel decides d_a^2 < d_b^2 exactly unless the difference is within the sum of its two
    double-float error bounds, 2^-43 (d R + d^2) each (R: the L1 sizes of the triangle's edges and of
    p minus a vertex, taken here over all three vertices), and then keeps the face the reference's
    traversal meets first; the FP64 oracle (like the reference) rebuilds the closest point in absolute
    coordinates, so its own rounding reaches 2^-48 (C + d) d (C: the largest coordinate magnitude).
    Below the sum of the three the oracle's choice is set by its rounding, not by the geometry."""
    p = np.asarray(p, np.float64).reshape(3)
    band = 0.0
    cmax = float(np.abs(p).max())
    for f in (face_a, face_b):
        v = _face_vertices(om, f)
        d = om.point_face_distance(int(f), p)
        r = float(np.abs(v - p).sum() + np.abs(v[1] - v[0]).sum() + np.abs(v[2] - v[0]).sum()
                  + np.abs(v[2] - v[1]).sum())
        band += 2.0 ** -43 * (d * r + d * d)
        cmax = max(cmax, float(np.abs(v).max()))
    d = om.point_face_distance(int(face_b), p)
    return band + 2.0 ** -48 * (cmax + d) * d


def point_faces_tied(om, p, face_a, face_b):
    """Whether the two faces' FP64 squared distances from p differ by no more than point_tie_band."""
    da = om.point_face_distance(int(face_a), p)
    db = om.point_face_distance(int(face_b), p)
    return abs(da * da - db * db) <= point_tie_band(om, p, face_a, face_b)


def check_closest_parity(om, points, got, ref, label=""):
    gf, rf = got["face"], ref["face"]
    same = gf == rf
    ties = 0
    bad = []
    for i in np.nonzero(~same)[0]:
        if gf[i] == M.NO_HIT or rf[i] == O.NO_HIT:
            bad.append(i)
            continue
        if point_faces_tied(om, points[i], gf[i], rf[i]):
            ties += 1
        else:
            bad.append(i)
    assert not bad, (f"{label}: {len(bad)} face mismatches beyond ties, e.g. query {bad[0]}: gpu {gf[bad[0]]} "
                     f"oracle {rf[bad[0]]} dist gpu {got['distance'][bad[0]]} oracle {ref['distance'][bad[0]]}")
    found = rf != O.NO_HIT
    stats = {"queries": len(gf), "found": int(found.sum()), "ties": ties, "face_equal": float(same.mean())}
    if found.any():
        de = rel_err(got["distance"][found], ref["distance"][found])
        pe = rel_err(got["point"][found], ref["point"][found])
        stats["max_dist_err"] = float(de.max())
        stats["max_point_err"] = float(pe.max())
        assert de.max() <= VALUE_REL, f"{label}: distance error {de.max():.3e}"
        eqf = found & same
        if eqf.any():
            pe2 = rel_err(got["point"][eqf], ref["point"][eqf])
            assert pe2.max() <= VALUE_REL, f"{label}: point error {pe2.max():.3e}"
    return stats


def threshold_cases():
    """A hand-built two-leaf mesh whose closestPointToPoint answers from the origin follow from
    closestPointToPoint.js:6-78 and shapecast.js:41-214 by hand.

    Twelve triangles in planes x = c, each centred on the x axis (distance |c| from the origin) unless
    stated. Leaf B (x > 0.5): 0.95, 1.5, 2, 3, 4 and a sliver near (0.5, 3, 0) that pulls B's box to
    x = 0.5 (box distance 0.5, faces all 0.95 or farther). Leaf A (x < 0): -0.6, -1, -1.5, -2, -2.5, -3
    (box distance 0.6). Split at the centroid centre, x = 0.5 (CENTER). B is visited first (smaller
    box distance); A only while its box distance 0.6 is under the best distance and maxThreshold.
    Returns (geometry, [(minThreshold, maxThreshold, expected distance or None)])."""
    from mbvh_amd import geometry as G

    s = 0.3
    tris = []
    for c in (0.95, 1.5, 2.0, 3.0, 4.0, -0.6, -1.0, -1.5, -2.0, -2.5, -3.0):
        tris.append([[c, -s, -s], [c, s, -s], [c, 0.0, s]])
    tris.insert(5, [[0.5, 3.0, 0.0], [0.6, 3.2, 0.0], [0.5, 3.0, 0.2]])
    pos = np.array(tris, np.float32).reshape(-1, 3)
    geo = G.Geometry(pos, np.arange(pos.shape[0], dtype=np.uint16))
    f = lambda x: float(np.float32(x))  # noqa: E731
    cases = [
        (0.0, float("inf"), f(0.6)),   # the closest (leaf A, entered after B: 0.6 < best 0.95)
        (1.0, float("inf"), f(0.95)),  # early exit in leaf B: the first face under 1.0, not the closest
        (0.9, float("inf"), f(0.6)),   # B has none under 0.9: A is entered and 0.6 ends the query
        (0.7, 0.7, f(0.6)),            # min = max ("is any face within 0.7"): A's 0.6
        (0.0, 0.7, f(0.6)),            # A's box (0.6) under maxThreshold and under best (0.95)
        (0.0, 0.55, f(0.95)),          # A pruned by maxThreshold; B's 0.95 kept though beyond it
        (0.0, 0.45, None),             # both boxes pruned: no result
        (0.96, 0.55, f(0.95)),         # min exit inside B
    ]
    return geo, cases
