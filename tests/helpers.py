"""Shared test helpers: building meshes, running the oracle, and the parity rules.

Parity rules (stated once, used by every GPU parity test):
  face index   equal to the oracle's, except ties: the oracle's FP64 distance to the GPU's face
               (ray: along the ray to the face's plane; point: to the face) is within TIE_REL of the
               oracle's best distance.
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


def check_closest_parity(om, points, got, ref, label=""):
    gf, rf = got["face"], ref["face"]
    same = gf == rf
    ties = 0
    bad = []
    for i in np.nonzero(~same)[0]:
        if gf[i] == M.NO_HIT or rf[i] == O.NO_HIT:
            bad.append(i)
            continue
        dface = om.point_face_distance(int(gf[i]), points[i])
        if abs(dface - ref["distance"][i]) <= TIE_REL * max(ref["distance"][i], 1.0):
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
