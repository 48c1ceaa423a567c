'use strict';
// getTriangleHitPointInfo (src/utils/TriangleUtilities.js:65-147): face vertices, material index,
// normal, barycoord and uv of a point on a triangle (for closestPointToPoint results). Host-side
// bookkeeping over the geometry's arrays; Triangle.getBarycoord is restated from three.js.
const M = require('./math.js');

function barycoord(p, a, b, c) {
  const v0 = [c[0] - a[0], c[1] - a[1], c[2] - a[2]];
  const v1 = [b[0] - a[0], b[1] - a[1], b[2] - a[2]];
  const v2 = [p[0] - a[0], p[1] - a[1], p[2] - a[2]];
  const dot = (x, y) => x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
  const d00 = dot(v0, v0), d01 = dot(v0, v1), d02 = dot(v0, v2), d11 = dot(v1, v1), d12 = dot(v1, v2);
  const denom = d00 * d11 - d01 * d01;
  if (denom === 0) return null;
  const inv = 1 / denom;
  const u = (d11 * d02 - d01 * d12) * inv;
  const v = (d00 * d12 - d01 * d02) * inv;
  return [1 - u - v, v, u];
}

function getTriangleHitPointInfo(point, geometry, triangleIndex, target) {
  const indices = geometry.index.array;
  const positions = geometry.attributes.position;
  const uvs = geometry.attributes.uv;
  const a = indices[triangleIndex * 3], b = indices[triangleIndex * 3 + 1], c = indices[triangleIndex * 3 + 2];
  const get = i => [positions.getX(i), positions.getY(i), positions.getZ(i)];
  const va = get(a), vb = get(b), vc = get(c);
  let materialIndex = 0;
  const first = triangleIndex * 3;
  for (const g of geometry.groups || []) {
    if (first >= g.start && first < g.start + g.count) {
      materialIndex = g.materialIndex;
      break;
    }
  }
  const bc = barycoord([point.x, point.y, point.z], va, vb, vc) || [0, 0, 0];
  const barycoordVec = target && target.barycoord ? target.barycoord : M.makeVec3(point, 0, 0, 0);
  barycoordVec.set(bc[0], bc[1], bc[2]);
  let uv = null;
  if (uvs) {
    const ux = uvs.getX(a) * bc[0] + uvs.getX(b) * bc[1] + uvs.getX(c) * bc[2];
    const uy = uvs.getY(a) * bc[0] + uvs.getY(b) * bc[1] + uvs.getY(c) * bc[2];
    if (target && target.uv) {
      uv = target.uv;
      uv.x = ux;
      uv.y = uy;
    } else {
      uv = M.makeVec2(null, ux, uy);
    }
  }
  const n = M.triangleNormal(va, vb, vc);
  const normal = target && target.face && target.face.normal ? target.face.normal : M.makeVec3(point, 0, 0, 0);
  normal.set(n[0], n[1], n[2]);
  const face = target && target.face ? target.face : {};
  face.a = a;
  face.b = b;
  face.c = c;
  face.materialIndex = materialIndex;
  face.normal = normal;
  const out = target || {};
  out.face = face;
  out.uv = uv;
  out.barycoord = barycoordVec;
  return out;
}

module.exports = { getTriangleHitPointInfo };
