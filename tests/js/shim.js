'use strict';
// Minimal stand-ins for the three.js classes the API is used with (three.js is not installed
// here). They carry only what the package reads: BufferAttribute accessors, geometry groups /
// draw range / index, Object3D.matrixWorld, Raycaster.ray / near / far / firstHitOnly.

class BufferAttribute {
  constructor(array, itemSize, normalized = false) {
    this.array = array;
    this.itemSize = itemSize;
    this.count = array.length / itemSize;
    this.normalized = normalized;
    this.isBufferAttribute = true;
  }
  getX(i) { return this.array[i * this.itemSize]; }
  getY(i) { return this.array[i * this.itemSize + 1]; }
  getZ(i) { return this.array[i * this.itemSize + 2]; }
}

class BufferGeometry {
  constructor() {
    this.isBufferGeometry = true;
    this.attributes = {};
    this.index = null;
    this.groups = [];
    this.drawRange = { start: 0, count: Infinity };
    this.boundingBox = null;
  }
  setIndex(attr) { this.index = attr; return this; }
  setAttribute(name, attr) { this.attributes[name] = attr; return this; }
  addGroup(start, count, materialIndex = 0) { this.groups.push({ start, count, materialIndex }); }
  clearGroups() { this.groups = []; }
}

function geometryFrom(positions, index) {
  const g = new BufferGeometry();
  g.setAttribute('position', new BufferAttribute(positions, 3));
  if (index) g.setIndex(new BufferAttribute(index, 1));
  return g;
}

class Vector3 {
  constructor(x = 0, y = 0, z = 0) { this.x = x; this.y = y; this.z = z; this.isVector3 = true; }
  set(x, y, z) { this.x = x; this.y = y; this.z = z; return this; }
  copy(v) { return this.set(v.x, v.y, v.z); }
  clone() { return new Vector3(this.x, this.y, this.z); }
  distanceTo(v) { return Math.hypot(this.x - v.x, this.y - v.y, this.z - v.z); }
  multiplyScalar(s) { return this.set(this.x * s, this.y * s, this.z * s); }
  normalize() { const l = Math.hypot(this.x, this.y, this.z) || 1; return this.multiplyScalar(1 / l); }
}

// Column-major 4x4 from translation, unit quaternion and scale (Matrix4.compose).
function compose(p, q, s) {
  const [x, y, z, w] = q;
  const x2 = x + x, y2 = y + y, z2 = z + z;
  const xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
  const wx = w * x2, wy = w * y2, wz = w * z2;
  return { elements: [
    (1 - (yy + zz)) * s[0], (xy + wz) * s[0], (xz - wy) * s[0], 0,
    (xy - wz) * s[1], (1 - (xx + zz)) * s[1], (yz + wx) * s[1], 0,
    (xz + wy) * s[2], (yz - wx) * s[2], (1 - (xx + yy)) * s[2], 0,
    p[0], p[1], p[2], 1,
  ] };
}

class Mesh {
  constructor(geometry, material) {
    this.geometry = geometry;
    this.material = material;
    this.matrixWorld = { elements: [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1] };
  }
}

class Raycaster {
  constructor(origin, direction) {
    this.ray = { origin: new Vector3(...origin), direction: new Vector3(...direction) };
    this.near = 0;
    this.far = Infinity;
    this.firstHitOnly = false;
  }
  intersectObject(object) {
    const hits = [];
    object.raycast(this, hits);
    hits.sort((a, b) => a.distance - b.distance);
    return hits;
  }
}

module.exports = { BufferAttribute, BufferGeometry, geometryFrom, Vector3, compose, Mesh, Raycaster };
