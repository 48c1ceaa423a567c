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

// acceleratedBatchedMeshRaycast (src/utils/ExtensionUtilities.js:87-161): each visible instance is
// raycast against its geometry's BVH with matrixWorld * instanceMatrix; hits carry batchId.
function acceleratedBatchedMeshRaycast(raycaster, intersects) {
  if (!this.boundsTrees) {
    runFallback(this, raycaster, intersects);
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
