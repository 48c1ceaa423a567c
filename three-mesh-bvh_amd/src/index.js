'use strict';
// Public entry point: the query half of three-mesh-bvh's API (src/index.js of the reference).
const C = require('./Constants.js');
const { MeshBVH } = require('./MeshBVH.js');
const { SceneBVH } = require('./SceneBVH.js');
const { MeshBVHUniformStruct, bvhNodeBuffer } = require('./MeshBVHUniformStruct.js');
const ext = require('./ExtensionUtilities.js');
const { getTriangleHitPointInfo } = require('./TriangleUtilities.js');
const native = require('./native.js');
const { Vec3 } = require('./math.js');

module.exports = {
  MeshBVH,
  SceneBVH,
  MeshBVHUniformStruct,
  bvhNodeBuffer,
  CENTER: C.CENTER,
  AVERAGE: C.AVERAGE,
  SAH: C.SAH,
  NOT_INTERSECTED: C.NOT_INTERSECTED,
  INTERSECTED: C.INTERSECTED,
  CONTAINED: C.CONTAINED,
  SKIP_GENERATION: C.SKIP_GENERATION,
  FrontSide: C.FrontSide,
  BackSide: C.BackSide,
  DoubleSide: C.DoubleSide,
  acceleratedRaycast: ext.acceleratedRaycast,
  computeBoundsTree: ext.computeBoundsTree,
  disposeBoundsTree: ext.disposeBoundsTree,
  computeBatchedBoundsTree: ext.computeBatchedBoundsTree,
  disposeBatchedBoundsTree: ext.disposeBatchedBoundsTree,
  setFallbackRaycast: ext.setFallbackRaycast,
  getTriangleHitPointInfo,
  setDevice: native.setDevice,
  getDevice: native.getDevice,
  deviceCount: () => native.load().deviceCount(),
  // Lanes per query on the current device (process-wide; no counterpart in the reference):
  // closest point 0 (automatic: four), 1, 4, or 64 (packets of 64 neighbouring queries sharing one
  // traversal: grids and SDF sampling, the same answers bit for bit); raycastFirst 0, 1 or 4.
  setQueryLanes: ({ closestPoint, raycastFirst } = {}) =>
    native.load().setQueryLanes(native.getDevice(), closestPoint, raycastFirst),
  Vec3,
};
