'use strict';
// three.js extension functions (src/utils/ExtensionUtilities.js): install with
//   Mesh.prototype.raycast = acceleratedRaycast;
//   BufferGeometry.prototype.computeBoundsTree = computeBoundsTree;
//   BatchedMesh.prototype.raycast = acceleratedRaycast;
//   BatchedMesh.prototype.computeBoundsTree = computeBatchedBoundsTree;
// Objects are duck-typed; a mesh without a bounds tree falls back to `fallbackRaycast` (three's own
// Mesh.prototype.raycast when the caller registers it with setFallbackRaycast).
const { MeshBVH } = require('./MeshBVH.js');
const M = require('./math.js');
const C = require('./Constants.js');
const native = require('./native.js');

let fallback = null;

// Registers the brute-force raycast used for objects without a bounds tree (three's
// Mesh.prototype.raycast / BatchedMesh.prototype.raycast, captured before patching).
function setFallbackRaycast(fn) {
  fallback = fn;
}

function runFallback(self, raycaster, intersects) {
  if (!fallback) throw new Error('BVH: Fallback raycast function not found.');
  fallback.call(self, raycaster, intersects);
}

// acceleratedRaycast (src/utils/ExtensionUtilities.js:37-85)
function acceleratedRaycast(raycaster, intersects) {
  if (this.isBatchedMesh) {
    acceleratedBatchedMeshRaycast.call(this, raycaster, intersects);
  } else if (this.geometry && this.geometry.boundsTree) {
    this.geometry.boundsTree.raycastObject3D(this, raycaster, intersects);
  } else {
    runFallback(this, raycaster, intersects);
  }
}

function sideOf(material) {
  const s = material && material.side !== undefined ? material.side : C.FrontSide;
  return s === C.BackSide ? C.BackSide : (s === C.DoubleSide ? C.DoubleSide : C.FrontSide);
}

// The batched device launch over all of a BatchedMesh's bounds trees, rebuilt when a tree or its
// device copy changes.
function batchFor(self) {
  const trees = self.boundsTrees;
  const slot = new Array(trees.length).fill(-1);
  const handles = [];
  for (let g = 0; g < trees.length; g++) {
    if (!trees[g]) continue;
    slot[g] = handles.length;
    handles.push(trees[g]._mesh());
  }
  const cache = self._mbvhBatch;
  if (cache && cache.handles.length === handles.length && cache.handles.every((h, k) => h === handles[k])) {
    return { batch: cache.batch, slot };
  }
  releaseBatch(self);
  if (!handles.length) return null;
  const batch = native.load().batchCreate(native.getDevice(), handles);
  self._mbvhBatch = { batch, handles };
  return { batch, slot };
}

function releaseBatch(self) {
  if (self._mbvhBatch) native.load().freeBatch(self._mbvhBatch.batch);
  self._mbvhBatch = null;
}

// firstHitOnly on a BatchedMesh with one material: every visible instance in ONE device launch
// instead of one raycastFirst call per instance. Each instance's local ray is raycastObject3D's
// (src/core/MeshBVH.js:456-468) computed here in double precision exactly as the reference computes
// it (matrixWorld * instanceMatrix, Matrix4.invert, Ray.applyMatrix4, near / far over the scale
// factor); the launch traces those double rays (mbvh_batch_raycast_first_local_f64) and each hit goes
// back to world space as convertRaycastIntersect does (src/utils/GeometryRayIntersectUtilities.js:3-17).
function batchedFirstHits(self, raycaster, intersects) {
  const b = batchFor(self);
  if (!b) return;
  const drawInfo = self._drawInfo || self._instanceInfo;
  const matrixWorld = self.matrixWorld.elements;
  const tmp = self.matrixWorld.clone ? self.matrixWorld.clone() : { elements: new Array(16) };
  const ro = raycaster.ray.origin, rd = raycaster.ray.direction;
  const inst = [];
  for (let i = 0, l = drawInfo.length; i < l; i++) {
    if (!self.getVisibleAt(i)) continue;
    const g = drawInfo[i].geometryIndex;
    if (!self.boundsTrees[g]) continue;
    self.getMatrixAt(i, tmp);
    const world = M.multiply(matrixWorld, tmp.elements);
    const inv = M.invert(world);
    const lo = M.applyToPoint(inv, ro);
    const ld = M.transformDirection(inv, rd);
    const s = M.matrixScale(world);
    const scaleFactor = M.length3(ld[0] * s[0], ld[1] * s[1], ld[2] * s[2]);
    inst.push({ i, g, world, lo, ld, near: raycaster.near / scaleFactor, far: raycaster.far / scaleFactor });
  }
  const n = inst.length;
  if (!n) return;
  const o = new Float64Array(3 * n), d = new Float64Array(3 * n), win = new Float64Array(2 * n);
  const mesh = new Uint32Array(n);
  for (let k = 0; k < n; k++) {
    const it = inst[k];
    o.set(it.lo, 3 * k);
    d.set(it.ld, 3 * k);
    win[2 * k] = it.near;
    win[2 * k + 1] = it.far;
    mesh[k] = b.slot[it.g];
  }
  const face = new Uint32Array(n), dist = new Float32Array(n), pt = new Float32Array(3 * n), bc = new Float32Array(3 * n);
  native.load().batchRaycastFirstLocal(b.batch, o, d, win, mesh, sideOf(self.material), face, dist, pt, bc);
  for (let k = 0; k < n; k++) {
    if (face[k] === C.NO_HIT) continue;
    const it = inst[k];
    const tree = self.boundsTrees[it.g];
    const localRay = { origin: M.makeVec3(ro, it.lo[0], it.lo[1], it.lo[2]),
      direction: M.makeVec3(rd, it.ld[0], it.ld[1], it.ld[2]) };
    const hit = tree._makeHit(localRay, self.material, face[k], dist[k], [pt[3 * k], pt[3 * k + 1], pt[3 * k + 2]],
      [bc[3 * k], bc[3 * k + 1], bc[3 * k + 2]]);
    // convertRaycastIntersect: the point to world space, the distance from the world ray's origin
    const pw = M.applyToPoint(it.world, hit.point);
    hit.point.set(pw[0], pw[1], pw[2]);
    hit.distance = hit.point.distanceTo(ro);
    hit.object = self;
    hit.batchId = it.i;
    intersects.push(hit);
  }
}

// acceleratedBatchedMeshRaycast (src/utils/ExtensionUtilities.js:87-161): each visible instance is
// raycast against its geometry's BVH with matrixWorld * instanceMatrix; hits carry batchId.
function acceleratedBatchedMeshRaycast(raycaster, intersects) {
  if (!this.boundsTrees) {
    runFallback(this, raycaster, intersects);
    return;
  }
  if (raycaster.firstHitOnly === true && !Array.isArray(this.material)) {
    batchedFirstHits(this, raycaster, intersects);
    return;
  }
  const drawInfo = this._drawInfo || this._instanceInfo;
  const matrixWorld = this.matrixWorld.elements;
  const batchHits = [];
  const tmp = this.matrixWorld.clone ? this.matrixWorld.clone() : { elements: new Array(16) };
  for (let i = 0, l = drawInfo.length; i < l; i++) {
    if (!this.getVisibleAt(i)) continue;
    const geometryId = drawInfo[i].geometryIndex;
    const tree = this.boundsTrees[geometryId];
    if (!tree) continue;
    this.getMatrixAt(i, tmp);
    const world = M.multiply(matrixWorld, tmp.elements);
    const proxy = { material: this.material, matrixWorld: { elements: world } };
    tree.raycastObject3D(proxy, raycaster, batchHits);
    for (const hit of batchHits) {
      hit.object = this;
      hit.batchId = i;
      intersects.push(hit);
    }
    batchHits.length = 0;
  }
}

// computeBoundsTree (src/utils/ExtensionUtilities.js:176-182)
function computeBoundsTree(options = {}) {
  const Type = options.type || MeshBVH;
  this.boundsTree = new Type(this, options);
  return this.boundsTree;
}

function disposeBoundsTree() {
  if (this.boundsTree && this.boundsTree.dispose) this.boundsTree.dispose();
  this.boundsTree = null;
}

// computeBatchedBoundsTree (src/utils/ExtensionUtilities.js:216-268): one BVH per geometry over the
// shared geometry, restricted to that geometry's draw range.
function computeBatchedBoundsTree(index = -1, options = {}) {
  options = Object.assign({}, options, { range: null });
  const drawRanges = this._drawRanges || this._geometryInfo;
  const geometryCount = this._geometryCount;
  if (!this.boundsTrees) this.boundsTrees = new Array(geometryCount).fill(null);
  const trees = this.boundsTrees;
  while (trees.length < geometryCount) trees.push(null);
  if (index < 0) {
    for (let i = 0; i < geometryCount; i++) {
      options.range = drawRanges[i];
      trees[i] = new MeshBVH(this.geometry, options);
    }
    return trees;
  }
  if (index < drawRanges.length) {
    options.range = drawRanges[index];
    trees[index] = new MeshBVH(this.geometry, options);
  }
  return trees[index] || null;
}

function disposeBatchedBoundsTree(index = -1) {
  releaseBatch(this);
  if (index < 0) {
    this.boundsTrees.forEach(t => t && t.dispose && t.dispose());
    this.boundsTrees.fill(null);
  } else if (index < this.boundsTrees.length) {
    const t = this.boundsTrees[index];
    if (t && t.dispose) t.dispose();
    this.boundsTrees[index] = null;
  }
}

module.exports = {
  acceleratedRaycast, computeBoundsTree, disposeBoundsTree, computeBatchedBoundsTree, disposeBatchedBoundsTree,
  setFallbackRaycast,
};
