"""Synthetic geometry in three.js's BufferGeometry layout (float32 positions, uint16/uint32 index).

three.js is a peer dependency of the reference and is not available here, so the generators the
benchmark configurations and tests name are written from three's published formulas
(TorusKnotGeometry, TorusGeometry, SphereGeometry, IcosahedronGeometry/PolyhedronGeometry,
BoxGeometry). Positions are computed in float64 and rounded to float32 the way a
Float32BufferAttribute stores them; index arrays are uint32 when any index is >= 65535, as
BufferGeometry.setIndex does.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class Geometry:
    positions: np.ndarray                 # (V, 3) float32
    index: Optional[np.ndarray]           # (3T,) uint16/uint32 or None (non-indexed)
    groups: List[Tuple[int, int, int]] = field(default_factory=list)  # (start, count, materialIndex)

    @property
    def tri_count(self) -> int:
        return (self.index.size if self.index is not None else self.positions.shape[0]) // 3

    def copy(self) -> "Geometry":
        return Geometry(self.positions.copy(), None if self.index is None else self.index.copy(),
                        list(self.groups))

    def ensure_index(self) -> "Geometry":
        """ensureIndex (src/core/build/geometryUtils.js:30-47): identity index if none."""
        if self.index is None:
            n = self.positions.shape[0]
            self.index = np.arange(n, dtype=np.uint32 if n > 65535 else np.uint16)
        return self

    def bounds(self) -> Tuple[np.ndarray, np.ndarray]:
        return self.positions.min(axis=0).astype(np.float64), self.positions.max(axis=0).astype(np.float64)


def _index_array(idx: np.ndarray) -> np.ndarray:
    return idx.astype(np.uint32 if idx.size and idx.max() >= 65535 else np.uint16)


def torus_knot(radius=1.0, tube=0.4, tubular_segments=64, radial_segments=8, p=2, q=3) -> Geometry:
    """TorusKnotGeometry."""
    ts, rs = int(tubular_segments), int(radial_segments)

    def curve(u):
        cu, su = np.cos(u), np.sin(u)
        qu = q / p * u
        cs = np.cos(qu)
        return np.stack([radius * (2 + cs) * 0.5 * cu, radius * (2 + cs) * su * 0.5,
                         radius * np.sin(qu) * 0.5], axis=-1)

    i = np.arange(ts + 1, dtype=np.float64)
    u = i / ts * p * np.pi * 2
    P1, P2 = curve(u), curve(u + 0.01)
    T = P2 - P1
    N = P2 + P1
    B = np.cross(T, N)
    N = np.cross(B, T)
    B = B / np.linalg.norm(B, axis=1, keepdims=True)
    N = N / np.linalg.norm(N, axis=1, keepdims=True)
    v = np.arange(rs + 1, dtype=np.float64) / rs * np.pi * 2
    cx = -tube * np.cos(v)
    cy = tube * np.sin(v)
    verts = P1[:, None, :] + (cx[None, :, None] * N[:, None, :] + cy[None, :, None] * B[:, None, :])
    positions = verts.reshape(-1, 3).astype(np.float32)

    j = np.arange(1, ts + 1)[:, None]
    k = np.arange(1, rs + 1)[None, :]
    a = (rs + 1) * (j - 1) + (k - 1)
    b = (rs + 1) * j + (k - 1)
    c = (rs + 1) * j + k
    d = (rs + 1) * (j - 1) + k
    idx = np.stack([a, b, d, b, c, d], axis=-1).reshape(-1)
    return Geometry(positions, _index_array(idx))


def torus(radius=1.0, tube=0.4, radial_segments=12, tubular_segments=48, arc=2 * np.pi) -> Geometry:
    """TorusGeometry."""
    rs, ts = int(radial_segments), int(tubular_segments)
    jj, ii = np.meshgrid(np.arange(rs + 1), np.arange(ts + 1), indexing="ij")
    u = ii / ts * arc
    v = jj / rs * np.pi * 2
    x = (radius + tube * np.cos(v)) * np.cos(u)
    y = (radius + tube * np.cos(v)) * np.sin(u)
    z = tube * np.sin(v)
    positions = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
    j = np.arange(1, rs + 1)[:, None]
    i = np.arange(1, ts + 1)[None, :]
    a = (ts + 1) * j + i - 1
    b = (ts + 1) * (j - 1) + i - 1
    c = (ts + 1) * (j - 1) + i
    d = (ts + 1) * j + i
    idx = np.stack([a, b, d, b, c, d], axis=-1).reshape(-1)
    return Geometry(positions, _index_array(idx))


def sphere(radius=1.0, width_segments=32, height_segments=16) -> Geometry:
    """SphereGeometry (full sphere)."""
    ws, hs = max(3, int(width_segments)), max(2, int(height_segments))
    iy, ix = np.meshgrid(np.arange(hs + 1), np.arange(ws + 1), indexing="ij")
    u = ix / ws
    v = iy / hs
    x = -radius * np.cos(u * np.pi * 2) * np.sin(v * np.pi)
    y = radius * np.cos(v * np.pi)
    z = radius * np.sin(u * np.pi * 2) * np.sin(v * np.pi)
    positions = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
    grid = np.arange((hs + 1) * (ws + 1)).reshape(hs + 1, ws + 1)
    idx = []
    for y_ in range(hs):
        a = grid[y_, 1:]
        b = grid[y_, :-1]
        c = grid[y_ + 1, :-1]
        d = grid[y_ + 1, 1:]
        parts = []
        if y_ != 0:
            parts.append(np.stack([a, b, d], axis=-1))
        if y_ != hs - 1:
            parts.append(np.stack([b, c, d], axis=-1))
        # three pushes (a, b, d) then (b, c, d) per quad, quad by quad
        if len(parts) == 2:
            idx.append(np.concatenate(parts, axis=-1).reshape(-1))
        else:
            idx.append(parts[0].reshape(-1))
    return Geometry(positions, _index_array(np.concatenate(idx)))


_ICO_T = (1 + np.sqrt(5)) / 2
_ICO_V = np.array([-1, _ICO_T, 0, 1, _ICO_T, 0, -1, -_ICO_T, 0, 1, -_ICO_T, 0, 0, -1, _ICO_T, 0, 1, _ICO_T,
                   0, -1, -_ICO_T, 0, 1, -_ICO_T, _ICO_T, 0, -1, _ICO_T, 0, 1, -_ICO_T, 0, -1, -_ICO_T, 0, 1],
                  dtype=np.float64).reshape(-1, 3)
_ICO_F = np.array([0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                   3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]).reshape(-1, 3)


def icosahedron(radius=1.0, detail=0) -> Geometry:
    """IcosahedronGeometry → PolyhedronGeometry.subdivide / applyRadius (non-indexed)."""
    cols = int(detail) + 1
    # Triangle corner grid coordinates (i, k) following subdivideFace's emission order.
    tri_ij = []
    for i in range(cols):
        for j in range(2 * (cols - i) - 1):
            k = j // 2
            if j % 2 == 0:
                tri_ij.append(((i, k + 1), (i + 1, k), (i, k)))
            else:
                tri_ij.append(((i, k + 1), (i + 1, k + 1), (i + 1, k)))
    tri_ij = np.array(tri_ij, dtype=np.int64)  # (ntri, 3, 2)
    I = tri_ij[..., 0].astype(np.float64)
    J = tri_ij[..., 1].astype(np.float64)
    rows = cols - I
    out = []
    for f in _ICO_F:
        a, b, c = _ICO_V[f[0]], _ICO_V[f[1]], _ICO_V[f[2]]
        t = (I / cols)[..., None]
        aj = a + (c - a) * t          # a.clone().lerp(c, i / cols)
        bj = b + (c - b) * t
        with np.errstate(invalid="ignore", divide="ignore"):
            s = np.where(rows > 0, J / np.where(rows > 0, rows, 1), 0.0)[..., None]
        v = aj + (bj - aj) * s        # aj.clone().lerp(bj, j / rows)
        apex = (J == 0) & (I == cols)  # v[cols][0] is aj itself
        v = np.where(apex[..., None], aj, v)
        out.append(v.reshape(-1, 3))
    verts = np.concatenate(out, axis=0)
    length = np.sqrt((verts ** 2).sum(axis=1, keepdims=True))
    verts = verts * (1.0 / np.where(length == 0, 1.0, length)) * radius
    return Geometry(verts.astype(np.float32), None)


def box(width=1.0, height=1.0, depth=1.0) -> Geometry:
    """BoxGeometry(1 segment per side) with one group per face (materialIndex 0..5)."""
    positions, index, groups = [], [], []
    nverts = 0

    def plane(u, v, w, udir, vdir, wdt, hgt, dpt, mat):
        nonlocal nverts
        start = len(index)
        for iy in range(2):
            y = iy * hgt - hgt / 2
            for ix in range(2):
                x = ix * wdt - wdt / 2
                vec = [0.0, 0.0, 0.0]
                vec[u] = x * udir
                vec[v] = y * vdir
                vec[w] = dpt / 2
                positions.append(vec)
        a, b, c, d = nverts + 0, nverts + 2, nverts + 3, nverts + 1
        index.extend([a, b, d, b, c, d])
        groups.append((start, 6, mat))
        nverts += 4

    X, Y, Z = 0, 1, 2
    plane(Z, Y, X, -1, -1, depth, height, width, 0)
    plane(Z, Y, X, 1, -1, depth, height, -width, 1)
    plane(X, Z, Y, 1, 1, width, depth, height, 2)
    plane(X, Z, Y, 1, -1, width, depth, -height, 3)
    plane(X, Y, Z, 1, -1, width, height, depth, 4)
    plane(X, Y, Z, -1, -1, width, height, -depth, 5)
    return Geometry(np.array(positions, dtype=np.float32), np.array(index, dtype=np.uint16), groups)


def value_noise(p: np.ndarray, seed: int, frequency: float = 4.0) -> np.ndarray:
    """Seeded trilinear value noise in [-1, 1], a pure function of position."""
    x = p.astype(np.float64) * frequency
    base = np.floor(x)
    f = x - base
    f = f * f * (3 - 2 * f)
    bi = base.astype(np.int64)

    def lattice(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        h = h & 0xFFFFFFFF
        h ^= h >> 16
        h = (h * 0x7FEB352D) & 0xFFFFFFFF
        h ^= h >> 15
        h = (h * 0x846CA68B) & 0xFFFFFFFF
        h ^= h >> 16
        return h.astype(np.float64) / 4294967295.0 * 2 - 1

    out = 0.0
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                out = out + wx * wy * wz * lattice(bi[:, 0] + dx, bi[:, 1] + dy, bi[:, 2] + dz)
    return out


def displaced_icosphere(detail: int, amplitude: float = 0.05, seed: int = 7) -> Geometry:
    """Icosahedron(r=1, detail) with every vertex moved along its radius by amplitude*noise(p).
    Noise depends only on position, so coincident (non-indexed) vertices stay welded."""
    g = icosahedron(1.0, detail)
    p = g.positions.astype(np.float64)
    n = value_noise(p, seed)
    r = np.sqrt((p ** 2).sum(axis=1, keepdims=True))
    q = p + (amplitude * n)[:, None] * p / np.where(r == 0, 1, r)
    g.positions = q.astype(np.float32)
    return g
