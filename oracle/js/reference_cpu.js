'use strict';
// FP64 JavaScript restatement of the reference's CPU query path, for timing the reference side
// beside the GPU. TEST INFRASTRUCTURE ONLY: run by bench.py's CPU-baseline leg and by tests/, never
// by the product.
//
// It follows the same files as oracle/oracle.c, in JavaScript doubles (JS never fuses multiply-adds,
// so every operation rounds as the reference's does):
//   raycastFirst   src/core/cast/raycastFirst.template.js:19-103 (ordered recursion, split-plane
//                  shortcut), intersectsNodeBounds src/core/utils/intersectUtils.js:2-78,
//                  intersectClosestTri src/core/utils/iterationUtils.template.js:24-53, three's
//                  Ray.intersectTriangle (GeometricTools IntrRay3Triangle3) as called by
//                  src/utils/ThreeRayIntersectUtilities.js:21-47, roots combined with strict <
//                  (src/core/MeshBVH.js:558-577)
//   refit          src/core/cast/refit.template.js:4-196 (nodeIndices = null)
//   intersectsBox  src/core/MeshBVH.js:780-792, OrientedBox src/math/OrientedBox.js:79-260
//   closestPoint   src/core/cast/closestPointToPoint.js:6-78 over shapecastTraverse
//                  (src/core/cast/shapecast.js:41-214): children by box distance^2, right first only
//                  if strictly closer, score < best^2 and < maxThreshold^2, strict < triangle update,
//                  minThreshold exit; three's Triangle.closestPointToPoint (Ericson RTCD 5.1.5)
// Node buffers are the reference's 32-byte format (src/core/Constants.js:61-66): float32 bounds,
// u32[6] right offset or first triangle, u32[7] split axis, u16[14] count, u16[15] 0xFFFF for leaves.
// The reference allocates result objects and Vector3 temporaries per call; this restatement does
// not, so it is at least as fast as the reference's own code on the same machine.

const FRONT = 0, BACK = 1, DOUBLE = 2;

function makeMesh(positions, index, roots) {
  return {
    pos: positions,
    index,
    roots: roots.map(buf => ({ f32: new Float32Array(buf), u32: new Uint32Array(buf), u16: new Uint16Array(buf) })),
  };
}

// ---- ray -------------------------------------------------------------------------------------
function boundsHit(f, n, r) {
  const b = 8 * n;
  let tmin, tmax, tymin, tymax, tzmin, tzmax;
  if (r.ix >= 0) { tmin = (f[b] - r.ox) * r.ix; tmax = (f[b + 3] - r.ox) * r.ix; } else { tmin = (f[b + 3] - r.ox) * r.ix; tmax = (f[b] - r.ox) * r.ix; }
  if (r.iy >= 0) { tymin = (f[b + 1] - r.oy) * r.iy; tymax = (f[b + 4] - r.oy) * r.iy; } else { tymin = (f[b + 4] - r.oy) * r.iy; tymax = (f[b + 1] - r.oy) * r.iy; }
  if (tmin > tymax || tymin > tmax) return false;
  if (tymin > tmin || tmin !== tmin) tmin = tymin;
  if (tymax < tmax || tmax !== tmax) tmax = tymax;
  if (r.iz >= 0) { tzmin = (f[b + 2] - r.oz) * r.iz; tzmax = (f[b + 5] - r.oz) * r.iz; } else { tzmin = (f[b + 5] - r.oz) * r.iz; tzmax = (f[b + 2] - r.oz) * r.iz; }
  if (tmin > tzmax || tzmin > tmax) return false;
  if (tzmin > tmin || tmin !== tmin) tmin = tzmin;
  if (tzmax < tmax || tmax !== tmax) tmax = tzmax;
  return tmin <= r.far && tmax >= r.near;
}

// Ray / triangle; on a hit fills r.hx, r.hy, r.hz (the point) and returns the distance, else -1.
function rayTri(r, ax, ay, az, bx, by, bz, cx, cy, cz, cull) {
  const e1x = bx - ax, e1y = by - ay, e1z = bz - az, e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
  const nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  let ddn = r.dx * nx + r.dy * ny + r.dz * nz;
  let sign;
  if (ddn > 0) { if (cull) return -1; sign = 1; } else if (ddn < 0) { sign = -1; ddn = -ddn; } else return -1;
  const qx = r.ox - ax, qy = r.oy - ay, qz = r.oz - az;
  const t1x = qy * e2z - qz * e2y, t1y = qz * e2x - qx * e2z, t1z = qx * e2y - qy * e2x;
  const e0 = sign * (r.dx * t1x + r.dy * t1y + r.dz * t1z);
  if (e0 < 0) return -1;
  const t2x = e1y * qz - e1z * qy, t2y = e1z * qx - e1x * qz, t2z = e1x * qy - e1y * qx;
  const e1 = sign * (r.dx * t2x + r.dy * t2y + r.dz * t2z);
  if (e1 < 0) return -1;
  if (e0 + e1 > ddn) return -1;
  const qdn = -sign * (qx * nx + qy * ny + qz * nz);
  if (qdn < 0) return -1;
  const t = qdn / ddn;
  r.hx = r.ox + r.dx * t; r.hy = r.oy + r.dy * t; r.hz = r.oz + r.dz * t;
  const dx = r.ox - r.hx, dy = r.oy - r.hy, dz = r.oz - r.hz;
  const dist = Math.sqrt(dx * dx + dy * dy + dz * dz);
  return dist < r.near || dist > r.far ? -1 : dist;
}

function leafFirst(m, r, off, cnt) {
  const p = m.pos, ix = m.index;
  let best = -1, bestDist = Infinity, px = 0, py = 0, pz = 0;
  for (let s = off; s < off + cnt; s++) {
    const a = 3 * ix[3 * s], b = 3 * ix[3 * s + 1], c = 3 * ix[3 * s + 2];
    let d;
    if (r.side === BACK) d = rayTri(r, p[c], p[c + 1], p[c + 2], p[b], p[b + 1], p[b + 2], p[a], p[a + 1], p[a + 2], true);
    else d = rayTri(r, p[a], p[a + 1], p[a + 2], p[b], p[b + 1], p[b + 2], p[c], p[c + 1], p[c + 2], r.side === FRONT);
    if (d >= 0 && d < bestDist) { best = s; bestDist = d; px = r.hx; py = r.hy; pz = r.hz; }
  }
  if (best < 0) return null;
  return { face: best, dist: bestDist, p: [px, py, pz] };
}

function firstNode(m, R, r, n) {
  if (R.u16[16 * n + 15] === 0xFFFF) return leafFirst(m, r, R.u32[8 * n + 6], R.u16[16 * n + 14]);
  const axis = R.u32[8 * n + 7];
  const d = axis === 0 ? r.dx : axis === 1 ? r.dy : r.dz;
  const ltr = d >= 0;
  const left = n + 1, right = n + R.u32[8 * n + 6];
  const c1 = ltr ? left : right, c2 = ltr ? right : left;
  const h1 = boundsHit(R.f32, c1, r) ? firstNode(m, R, r, c1) : null;
  if (h1) {
    const pv = h1.p[axis];
    const outside = ltr ? pv <= R.f32[8 * c2 + axis] : pv >= R.f32[8 * c2 + axis + 3];
    if (outside) return h1;
  }
  const h2 = boundsHit(R.f32, c2, r) ? firstNode(m, R, r, c2) : null;
  if (h1 && h2) return h1.dist <= h2.dist ? h1 : h2;
  return h1 || h2;
}

function raycastFirst(m, ox, oy, oz, dx, dy, dz, side, near, far) {
  const r = { ox, oy, oz, dx, dy, dz, ix: 1 / dx, iy: 1 / dy, iz: 1 / dz, side, near, far, hx: 0, hy: 0, hz: 0 };
  let best = null;
  for (const R of m.roots) {
    const h = firstNode(m, R, r, 0);
    if (h && (!best || h.dist < best.dist)) best = h;
  }
  return best;
}

// ---- closest point -----------------------------------------------------------------------------
function boxScore(f, n, px, py, pz) {
  const b = 8 * n;
  const x = Math.max(f[b], Math.min(f[b + 3], px)) - px;
  const y = Math.max(f[b + 1], Math.min(f[b + 4], py)) - py;
  const z = Math.max(f[b + 2], Math.min(f[b + 5], pz)) - pz;
  return x * x + y * y + z * z;
}

// Squared distance from (px, py, pz) to the triangle's closest point (Ericson's regions in order).
function triDistSq(px, py, pz, ax, ay, az, bx, by, bz, cx, cy, cz) {
  const abx = bx - ax, aby = by - ay, abz = bz - az, acx = cx - ax, acy = cy - ay, acz = cz - az;
  const apx = px - ax, apy = py - ay, apz = pz - az;
  const d1 = abx * apx + aby * apy + abz * apz, d2 = acx * apx + acy * apy + acz * apz;
  let qx, qy, qz;
  if (d1 <= 0 && d2 <= 0) { qx = ax; qy = ay; qz = az; } else {
    const bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const d3 = abx * bpx + aby * bpy + abz * bpz, d4 = acx * bpx + acy * bpy + acz * bpz;
    const vc = d1 * d4 - d3 * d2;
    if (d3 >= 0 && d4 <= d3) { qx = bx; qy = by; qz = bz; } else if (vc <= 0 && d1 >= 0 && d3 <= 0) {
      const v = d1 / (d1 - d3);
      qx = ax + abx * v; qy = ay + aby * v; qz = az + abz * v;
    } else {
      const cpx = px - cx, cpy = py - cy, cpz = pz - cz;
      const d5 = abx * cpx + aby * cpy + abz * cpz, d6 = acx * cpx + acy * cpy + acz * cpz;
      const vb = d5 * d2 - d1 * d6;
      if (d6 >= 0 && d5 <= d6) { qx = cx; qy = cy; qz = cz; } else if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const w = d2 / (d2 - d6);
        qx = ax + acx * w; qy = ay + acy * w; qz = az + acz * w;
      } else {
        const va = d3 * d6 - d5 * d4;
        if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
          const w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
          qx = bx + (cx - bx) * w; qy = by + (cy - by) * w; qz = bz + (cz - bz) * w;
        } else {
          const denom = 1 / (va + vb + vc);
          const v = vb * denom, w = vc * denom;
          qx = (ax + abx * v) + acx * w; qy = (ay + aby * v) + acy * w; qz = (az + abz * v) + acz * w;
        }
      }
    }
  }
  const x = px - qx, y = py - qy, z = pz - qz;
  return x * x + y * y + z * z;
}

function pointLeaf(m, c, off, cnt) {
  const p = m.pos, ix = m.index;
  for (let s = off; s < off + cnt; s++) {
    const a = 3 * ix[3 * s], b = 3 * ix[3 * s + 1], cc = 3 * ix[3 * s + 2];
    const d2 = triDistSq(c.px, c.py, c.pz, p[a], p[a + 1], p[a + 2], p[b], p[b + 1], p[b + 2], p[cc], p[cc + 1], p[cc + 2]);
    if (d2 < c.best) { c.best = d2; c.face = s; }
    if (d2 < c.min2) return true;
  }
  return false;
}

function pointNode(m, R, c, n) {
  if (R.u16[16 * n + 15] === 0xFFFF) return pointLeaf(m, c, R.u32[8 * n + 6], R.u16[16 * n + 14]);
  let c1 = n + 1, c2 = n + R.u32[8 * n + 6];
  let s1 = boxScore(R.f32, c1, c.px, c.py, c.pz), s2 = boxScore(R.f32, c2, c.px, c.py, c.pz);
  if (s2 < s1) { const t = c1; c1 = c2; c2 = t; const u = s1; s1 = s2; s2 = u; }
  if (s1 < c.best && s1 < c.max2 && pointNode(m, R, c, c1)) return true;
  if (s2 < c.best && s2 < c.max2 && pointNode(m, R, c, c2)) return true;
  return false;
}

function closestPoint(m, px, py, pz, minT, maxT) {
  const c = { px, py, pz, best: Infinity, face: -1, min2: minT * minT, max2: maxT * maxT };
  for (const R of m.roots) if (pointNode(m, R, c, 0)) break;
  return c.face < 0 ? null : { face: c.face, distSq: c.best };
}

// ---- refit (src/core/cast/refit.template.js:4-196, nodeIndices = null) ---------------------------
// Rewrites the boxes of every root of `m` in place from its positions: a leaf's box is the min / max
// of its triangles' vertices (strict < / >, slot order), written only if a component differs; an
// internal node is the union of its children, recomputed only when a child changed.
function refitNode(m, R, n) {
  const f = R.f32, b = 8 * n;
  if (R.u16[16 * n + 15] === 0xFFFF) {
    const off = R.u32[b + 6], cnt = R.u16[16 * n + 14];
    const p = m.pos, ix = m.index;
    let minx = Infinity, miny = Infinity, minz = Infinity, maxx = -Infinity, maxy = -Infinity, maxz = -Infinity;
    for (let i = 3 * off, l = 3 * (off + cnt); i < l; i++) {
      const v = 3 * ix[i];
      const x = p[v], y = p[v + 1], z = p[v + 2];
      if (x < minx) minx = x;
      if (x > maxx) maxx = x;
      if (y < miny) miny = y;
      if (y > maxy) maxy = y;
      if (z < minz) minz = z;
      if (z > maxz) maxz = z;
    }
    if (f[b] !== minx || f[b + 1] !== miny || f[b + 2] !== minz || f[b + 3] !== maxx || f[b + 4] !== maxy || f[b + 5] !== maxz) {
      f[b] = minx; f[b + 1] = miny; f[b + 2] = minz; f[b + 3] = maxx; f[b + 4] = maxy; f[b + 5] = maxz;
      return true;
    }
    return false;
  }
  const left = n + 1, right = n + R.u32[b + 6];
  const lc = refitNode(m, R, left);
  const rc = refitNode(m, R, right);
  if (lc || rc) {
    const l = 8 * left, r = 8 * right;
    for (let i = 0; i < 3; i++) {
      f[b + i] = f[l + i] < f[r + i] ? f[l + i] : f[r + i];
      f[b + i + 3] = f[l + i + 3] > f[r + i + 3] ? f[l + i + 3] : f[r + i + 3];
    }
  }
  return lc || rc;
}

function refit(m) {
  for (const R of m.roots) refitNode(m, R, 0);
}

// ---- intersectsBox (src/core/MeshBVH.js:780-792 over shapecast) ---------------------------------
// OrientedBox.update / intersectsBox / intersectsTriangle (src/math/OrientedBox.js:79-260) with
// SeparatingAxisBounds and ExtendedTriangle.update (src/math/ExtendedTriangle.js:47-75); node boxes by
// the six-axis test, triangles by the full test, depth first, the first intersecting triangle ends it.
function sepPoints(ax, pts, out) {
  let mn = Infinity, mx = -Infinity;
  for (let i = 0; i < pts.length; i += 3) {
    const v = ax[0] * pts[i] + ax[1] * pts[i + 1] + ax[2] * pts[i + 2];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  out[0] = mn; out[1] = mx;
  return out;
}
const sep = (a, b) => a[0] > b[1] || b[0] > a[1];

function makeObb(bmin, bmax, e) {
  const p = new Float64Array(24);
  for (let x = 0; x <= 1; x++) for (let y = 0; y <= 1; y++) for (let z = 0; z <= 1; z++) {
    const i = x | (y << 1) | (z << 2);
    const vx = x ? bmax[0] : bmin[0], vy = y ? bmax[1] : bmin[1], vz = z ? bmax[2] : bmin[2];
    const w = 1 / (e[3] * vx + e[7] * vy + e[11] * vz + e[15]);
    p[3 * i] = (e[0] * vx + e[4] * vy + e[8] * vz + e[12]) * w;
    p[3 * i + 1] = (e[1] * vx + e[5] * vy + e[9] * vz + e[13]) * w;
    p[3 * i + 2] = (e[2] * vx + e[6] * vy + e[10] * vz + e[14]) * w;
  }
  const ax = [], sb = [], ab = [];
  for (let i = 0; i < 3; i++) {
    const j = 3 * (1 << i);
    const a = [p[0] - p[j], p[1] - p[j + 1], p[2] - p[j + 2]];
    ax.push(a);
    sb.push(sepPoints(a, p, [0, 0]));
    let mn = Infinity, mx = -Infinity;
    for (let k = 0; k < 8; k++) { const v = p[3 * k + i]; mn = v < mn ? v : mn; mx = v > mx ? v : mx; }
    ab.push([mn, mx]);
  }
  return { p, ax, sb, ab, t: [0, 0], u: [0, 0] };
}

function obbBox(o, f, n) {
  const b = 8 * n;
  for (let k = 0; k < 3; k++) if (o.ab[k][0] > f[b + 3 + k] || f[b + k] > o.ab[k][1]) return false;
  for (let i = 0; i < 3; i++) {
    const ax = o.ax[i];
    let mn = Infinity, mx = -Infinity;
    for (let x = 0; x <= 1; x++) for (let y = 0; y <= 1; y++) for (let z = 0; z <= 1; z++) {
      const px = f[b] * x + f[b + 3] * (1 - x), py = f[b + 1] * y + f[b + 4] * (1 - y), pz = f[b + 2] * z + f[b + 5] * (1 - z);
      const v = ax[0] * px + ax[1] * py + ax[2] * pz;
      mn = Math.min(v, mn);
      mx = Math.max(v, mx);
    }
    if (o.sb[i][0] > mx || mn > o.sb[i][1]) return false;
  }
  return true;
}

function obbTri(o, tri) {
  const a = [tri[0], tri[1], tri[2]], b = [tri[3], tri[4], tri[5]], c = [tri[6], tri[7], tri[8]];
  const cb = [c[0] - b[0], c[1] - b[1], c[2] - b[2]], ab = [a[0] - b[0], a[1] - b[1], a[2] - b[2]];
  let n = [cb[1] * ab[2] - cb[2] * ab[1], cb[2] * ab[0] - cb[0] * ab[2], cb[0] * ab[1] - cb[1] * ab[0]];
  const l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (l2 > 0) { const s = 1 / Math.sqrt(l2); n = [n[0] * s, n[1] * s, n[2] * s]; } else n = [0, 0, 0];
  const tax = [n, ab, [b[0] - c[0], b[1] - c[1], b[2] - c[2]], [c[0] - a[0], c[1] - a[1], c[2] - a[2]]];
  for (let i = 0; i < 3; i++) if (sep(o.sb[i], sepPoints(o.ax[i], tri, o.t))) return false;
  for (let i = 0; i < 3; i++) if (sep(sepPoints(tax[i], tri, o.t), sepPoints(tax[i], o.p, o.u))) return false;
  for (let i = 0; i < 3; i++) {
    const s1 = o.ax[i];
    for (let i2 = 0; i2 < 4; i2++) {
      const s2 = tax[i2];
      const x = [s1[1] * s2[2] - s1[2] * s2[1], s1[2] * s2[0] - s1[0] * s2[2], s1[0] * s2[1] - s1[1] * s2[0]];
      if (sep(sepPoints(x, tri, o.t), sepPoints(x, o.p, o.u))) return false;
    }
  }
  return true;
}

function boxNode(m, R, o, n, tri) {
  if (R.u16[16 * n + 15] === 0xFFFF) {
    const off = R.u32[8 * n + 6], cnt = R.u16[16 * n + 14], p = m.pos, ix = m.index;
    for (let s = off; s < off + cnt; s++) {
      for (let k = 0; k < 3; k++) {
        const v = 3 * ix[3 * s + k];
        tri[3 * k] = p[v]; tri[3 * k + 1] = p[v + 1]; tri[3 * k + 2] = p[v + 2];
      }
      if (obbTri(o, tri)) return true;
    }
    return false;
  }
  const c1 = n + 1, c2 = n + R.u32[8 * n + 6];
  if (obbBox(o, R.f32, c1) && boxNode(m, R, o, c1, tri)) return true;
  if (obbBox(o, R.f32, c2) && boxNode(m, R, o, c2, tri)) return true;
  return false;
}

function intersectsBox(m, bmin, bmax, e) {
  const o = makeObb(bmin, bmax, e);
  const tri = new Float64Array(9);
  for (const R of m.roots) if (obbBox(o, R.f32, 0) && boxNode(m, R, o, 0, tri)) return true;
  return false;
}

module.exports = { FRONT, BACK, DOUBLE, makeMesh, raycastFirst, closestPoint, refit, intersectsBox };
