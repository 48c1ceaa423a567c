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
// (mbvh_batch_raycast_first), instead of one raycastFirst call per instance. The frame change is
// raycastObject3D's, done on the device in FP32 from the float32 matrices (the per-instance path
// does it in FP64 and rounds the local ray), so hits agree with that path up to that rounding.
function batchedFirstHits(self, raycaster, intersects) {
  const b = batchFor(self);
  if (!b) return;
  const drawInfo = self._drawInfo || self._instanceInfo;
  const matrixWorld = self.matrixWorld.elements;
  const tmp = self.matrixWorld.clone ? self.matrixWorld.clone() : { elements: new Array(16) };
  const ids = [];
  const mats = [];
  for (let i = 0, l = drawInfo.length; i < l; i++) {
    if (!self.getVisibleAt(i)) continue;
    const g = drawInfo[i].geometryIndex;
    if (!self.boundsTrees[g]) continue;
    self.getMatrixAt(i, tmp);
    const world = M.multiply(matrixWorld, tmp.elements);
    ids.push(i);
    mats.push({ g, world, inv: M.invert(world) });
  }
  const n = ids.length;
  if (!n) return;
  const mw = new Float32Array(16 * n), mi = new Float32Array(16 * n), ig = new Uint32Array(n);
  const o = new Float32Array(3 * n), d = new Float32Array(3 * n), ri = new Uint32Array(n);
  const ro = raycaster.ray.origin, rd = raycaster.ray.direction;
  for (let k = 0; k < n; k++) {
    mw.set(mats[k].world, 16 * k);
    mi.set(mats[k].inv, 16 * k);
    ig[k] = b.slot[mats[k].g];
    o[3 * k] = ro.x; o[3 * k + 1] = ro.y; o[3 * k + 2] = ro.z;
    d[3 * k] = rd.x; d[3 * k + 1] = rd.y; d[3 * k + 2] = rd.z;
    ri[k] = k;
  }
  const face = new Uint32Array(n), dist = new Float32Array(n), pt = new Float32Array(3 * n), bc = new Float32Array(3 * n);
  native.load().batchRaycastFirst(b.batch, mw, mi, ig, o, d, ri, sideOf(self.material), raycaster.near, raycaster.far,
    face, dist, pt, bc);
  for (let k = 0; k < n; k++) {
    if (face[k] === C.NO_HIT) continue;
    const tree = self.boundsTrees[mats[k].g];
    // face data in the instance's frame (three's intersection keeps face / normals local)
    const lo = M.applyToPoint(mats[k].inv, ro);
    const ld = M.transformDirection(mats[k].inv, rd);
    const localRay = { origin: M.makeVec3(ro, lo[0], lo[1], lo[2]), direction: M.makeVec3(rd, ld[0], ld[1], ld[2]) };
    const hit = tree._makeHit(localRay, self.material, face[k], dist[k], [pt[3 * k], pt[3 * k + 1], pt[3 * k + 2]],
      [bc[3 * k], bc[3 * k + 1], bc[3 * k + 2]]);
    hit.object = self;
    hit.batchId = ids[k];
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
