'use strict';
// Minimal vector / matrix helpers so the package works with three.js objects (duck-typed: any
// {x, y, z} vector and any matrix with a column-major `elements` array) without depending on three.

class Vec3 {
  constructor(x = 0, y = 0, z = 0) {
    this.x = x;
    this.y = y;
    this.z = z;
    this.isVector3 = true;
  }

  set(x, y, z) {
    this.x = x;
    this.y = y;
    this.z = z;
    return this;
  }

  copy(v) {
    return this.set(v.x, v.y, v.z);
  }

  clone() {
    return new Vec3(this.x, this.y, this.z);
  }

  distanceTo(v) {
    const dx = this.x - v.x, dy = this.y - v.y, dz = this.z - v.z;
    return Math.sqrt(dx * dx + dy * dy + dz * dz);
  }

  dot(v) {
    return this.x * v.x + this.y * v.y + this.z * v.z;
  }

  length() {
    return Math.sqrt(this.dot(this));
  }

  multiplyScalar(s) {
    return this.set(this.x * s, this.y * s, this.z * s);
  }

  applyMatrix4(m) {
    const p = applyToPoint(m.elements || m, this);
    return this.set(p[0], p[1], p[2]);
  }
}

// A new vector of the same kind as `like` (a three.js Vector3 when the caller uses three).
function makeVec3(like, x, y, z) {
  if (like && like.isVector3 && typeof like.constructor === 'function' && like.constructor !== Object) {
    const v = new like.constructor(x, y, z);
    if (v.x === x && v.y === y && v.z === z) return v;
  }
  return new Vec3(x, y, z);
}

function makeVec2(like, x, y) {
  if (like && like.isVector2 && typeof like.constructor === 'function') return new like.constructor(x, y);
  return { x, y, isVector2: true };
}

// Matrix4.invert (cofactor expansion); the zero matrix when singular, like three.js.
function invert(e) {
  const n11 = e[0], n21 = e[1], n31 = e[2], n41 = e[3];
  const n12 = e[4], n22 = e[5], n32 = e[6], n42 = e[7];
  const n13 = e[8], n23 = e[9], n33 = e[10], n43 = e[11];
  const n14 = e[12], n24 = e[13], n34 = e[14], n44 = e[15];
  const t11 = n23 * n34 * n42 - n24 * n33 * n42 + n24 * n32 * n43 - n22 * n34 * n43 - n23 * n32 * n44 + n22 * n33 * n44;
  const t12 = n14 * n33 * n42 - n13 * n34 * n42 - n14 * n32 * n43 + n12 * n34 * n43 + n13 * n32 * n44 - n12 * n33 * n44;
  const t13 = n13 * n24 * n42 - n14 * n23 * n42 + n14 * n22 * n43 - n12 * n24 * n43 - n13 * n22 * n44 + n12 * n23 * n44;
  const t14 = n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 + n13 * n22 * n34 - n12 * n23 * n34;
  const det = n11 * t11 + n21 * t12 + n31 * t13 + n41 * t14;
  const out = new Array(16).fill(0);
  if (det === 0) return out;
  const d = 1 / det;
  out[0] = t11 * d;
  out[1] = (n24 * n33 * n41 - n23 * n34 * n41 - n24 * n31 * n43 + n21 * n34 * n43 + n23 * n31 * n44 - n21 * n33 * n44) * d;
  out[2] = (n22 * n34 * n41 - n24 * n32 * n41 + n24 * n31 * n42 - n21 * n34 * n42 - n22 * n31 * n44 + n21 * n32 * n44) * d;
  out[3] = (n23 * n32 * n41 - n22 * n33 * n41 - n23 * n31 * n42 + n21 * n33 * n42 + n22 * n31 * n43 - n21 * n32 * n43) * d;
  out[4] = t12 * d;
  out[5] = (n13 * n34 * n41 - n14 * n33 * n41 + n14 * n31 * n43 - n11 * n34 * n43 - n13 * n31 * n44 + n11 * n33 * n44) * d;
  out[6] = (n14 * n32 * n41 - n12 * n34 * n41 - n14 * n31 * n42 + n11 * n34 * n42 + n12 * n31 * n44 - n11 * n32 * n44) * d;
  out[7] = (n12 * n33 * n41 - n13 * n32 * n41 + n13 * n31 * n42 - n11 * n33 * n42 - n12 * n31 * n43 + n11 * n32 * n43) * d;
  out[8] = t13 * d;
  out[9] = (n14 * n23 * n41 - n13 * n24 * n41 - n14 * n21 * n43 + n11 * n24 * n43 + n13 * n21 * n44 - n11 * n23 * n44) * d;
  out[10] = (n12 * n24 * n41 - n14 * n22 * n41 + n14 * n21 * n42 - n11 * n24 * n42 - n12 * n21 * n44 + n11 * n22 * n44) * d;
  out[11] = (n13 * n22 * n41 - n12 * n23 * n41 - n13 * n21 * n42 + n11 * n23 * n42 + n12 * n21 * n43 - n11 * n22 * n43) * d;
  out[12] = t14 * d;
  out[13] = (n13 * n24 * n31 - n14 * n23 * n31 + n14 * n21 * n33 - n11 * n24 * n33 - n13 * n21 * n34 + n11 * n23 * n34) * d;
  out[14] = (n14 * n22 * n31 - n12 * n24 * n31 - n14 * n21 * n32 + n11 * n24 * n32 + n12 * n21 * n34 - n11 * n22 * n34) * d;
  out[15] = (n12 * n23 * n31 - n13 * n22 * n31 + n13 * n21 * n32 - n11 * n23 * n32 - n12 * n21 * n33 + n11 * n22 * n33) * d;
  return out;
}

function multiply(a, b) {
  const out = new Array(16);
  for (let c = 0; c < 4; c++) {
    for (let r = 0; r < 4; r++) {
      out[c * 4 + r] = a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] + a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3];
    }
  }
  return out;
}

// Vector3.applyMatrix4 (with the projective divide).
function applyToPoint(e, v) {
  const x = v.x, y = v.y, z = v.z;
  const w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
  return [(e[0] * x + e[4] * y + e[8] * z + e[12]) * w, (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
    (e[2] * x + e[6] * y + e[10] * z + e[14]) * w];
}

// Vector3.length: the square root of the sum of squares (not Math.hypot, which rounds differently).
function length3(x, y, z) {
  return Math.sqrt(x * x + y * y + z * z);
}

// Vector3.transformDirection: upper 3x3, then normalize (Vector3.divideScalar multiplies by the
// reciprocal of the length, or of 1 for a zero vector).
function transformDirection(e, v) {
  const x = v.x, y = v.y, z = v.z;
  const rx = e[0] * x + e[4] * y + e[8] * z;
  const ry = e[1] * x + e[5] * y + e[9] * z;
  const rz = e[2] * x + e[6] * y + e[10] * z;
  const inv = 1 / (length3(rx, ry, rz) || 1);
  return [rx * inv, ry * inv, rz * inv];
}

// Vector3.setFromMatrixScale: lengths of the basis columns.
function matrixScale(e) {
  return [length3(e[0], e[1], e[2]), length3(e[4], e[5], e[6]), length3(e[8], e[9], e[10])];
}

// Triangle.getNormal: normalize((c - b) x (a - b)); zero for a degenerate triangle.
function triangleNormal(a, b, c) {
  const ux = c[0] - b[0], uy = c[1] - b[1], uz = c[2] - b[2];
  const vx = a[0] - b[0], vy = a[1] - b[1], vz = a[2] - b[2];
  let x = uy * vz - uz * vy, y = uz * vx - ux * vz, z = ux * vy - uy * vx;
  const l2 = x * x + y * y + z * z;
  if (l2 > 0) {
    const s = 1 / Math.sqrt(l2);
    x *= s;
    y *= s;
    z *= s;
  } else {
    x = y = z = 0;
  }
  return [x, y, z];
}

module.exports = {
  Vec3, makeVec3, makeVec2, invert, multiply, applyToPoint, transformDirection, matrixScale, triangleNormal, length3,
};
