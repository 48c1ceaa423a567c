'use strict';
// Scene queries: the closest first hit over many instances of MeshBVHs in ONE device launch — the
// top level of the reference's ClusteredBVH (src/webgpu/ClusteredBVH.js:18-325, traversed as in
// src/webgpu/shapecastFns/getShapecastFn.js:97-217) with one top-level primitive per instance. The
// answer is the hit three's Raycaster puts first among the per-instance first hits of
// acceleratedBatchedMeshRaycast (src/utils/ExtensionUtilities.js:87-161): the smallest world
// distance, equal distances to the lower instance index.
const C = require('./Constants.js');
const M = require('./math.js');
const native = require('./native.js');

const elementsOf = m => (m && m.elements ? m.elements : m);

class SceneBVH {
  // instances: [{ bvh: MeshBVH, matrixWorld: Matrix4 | 16 numbers (column-major), visible?: bool }]
  // The BVHs must stay alive; re-create the scene after a refit or after instances move.
  constructor(instances) {
    if (!Array.isArray(instances)) throw new Error('SceneBVH: instances must be an array');
    const trees = [];
    const slot = new Map();
    const n = instances.length;
    const world = new Float32Array(16 * n), inverse = new Float32Array(16 * n);
    const geometry = new Uint32Array(n), visible = new Uint8Array(n);
    instances.forEach((inst, i) => {
      if (!inst || !inst.bvh) throw new Error(`SceneBVH: instance ${i} has no bvh`);
      if (!slot.has(inst.bvh)) {
        slot.set(inst.bvh, trees.length);
        trees.push(inst.bvh);
      }
      const e = elementsOf(inst.matrixWorld) || [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1];
      world.set(e, 16 * i);
      inverse.set(M.invert(e), 16 * i);
      geometry[i] = slot.get(inst.bvh);
      visible[i] = inst.visible === false ? 0 : 1;
    });
    if (!trees.length) throw new Error('SceneBVH: no instances');
    this.instances = instances;
    this.trees = trees;
    this._world = world;
    this._inverse = inverse;
    this._handles = trees.map(t => t._mesh());
    this._scene = native.load().sceneCreate(native.getDevice(), this._handles, world, inverse, geometry, visible);
  }

  // The instances of a BatchedMesh with bounds trees (computeBatchedBoundsTree): instance i is
  // drawn geometry drawInfo[i].geometryIndex with matrixWorld * instanceMatrix, visible as
  // getVisibleAt(i); results carry batchId = i.
  static fromBatchedMesh(batchedMesh) {
    const drawInfo = batchedMesh._drawInfo || batchedMesh._instanceInfo;
    const trees = batchedMesh.boundsTrees || [];
    const mw = batchedMesh.matrixWorld.elements;
    const tmp = batchedMesh.matrixWorld.clone ? batchedMesh.matrixWorld.clone() : { elements: new Array(16) };
    const list = [];
    for (let i = 0; i < drawInfo.length; i++) {
      const tree = trees[drawInfo[i].geometryIndex];
      batchedMesh.getMatrixAt(i, tmp);
      list.push({ bvh: tree, matrixWorld: M.multiply(mw, tmp.elements), visible: !!tree && batchedMesh.getVisibleAt(i) });
    }
    // instances without a tree keep their index (as batchId) but are never hit
    const any = list.find(x => x.bvh);
    if (!any) throw new Error('SceneBVH.fromBatchedMesh: no bounds trees');
    list.forEach(x => { if (!x.bvh) x.bvh = any.bvh; });
    const scene = new SceneBVH(list);
    scene.object = batchedMesh;
    return scene;
  }

  // Many world rays at once (SoA Float32Arrays of 3n): faceIndex (in the hit instance's geometry,
  // 0xFFFFFFFF on a miss), instance, world distance and point, barycoord.
  raycastFirstBatch(origins, directions, options = {}) {
    const n = origins.length / 3;
    const out = {
      faceIndex: new Uint32Array(n), instance: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    const side = options.side === undefined ? C.FrontSide : options.side;
    native.load().sceneRaycastFirst(this._scene, origins, directions, side, options.near || 0,
      options.far === undefined ? Infinity : options.far, out.faceIndex, out.instance, out.distance, out.point,
      out.barycoord);
    return out;
  }

  // One ray (three's Ray): the closest hit as an intersection object (the instance's face data,
  // world point and distance, `instance`, and `batchId` for a BatchedMesh scene), or null.
  raycastFirst(ray, materialOrSide = C.FrontSide, near = 0, far = Infinity) {
    const side = typeof materialOrSide === 'number' ? materialOrSide
      : (materialOrSide && materialOrSide.side !== undefined ? materialOrSide.side : C.FrontSide);
    const o = new Float32Array([ray.origin.x, ray.origin.y, ray.origin.z]);
    const d = new Float32Array([ray.direction.x, ray.direction.y, ray.direction.z]);
    const r = this.raycastFirstBatch(o, d, { side, near, far });
    if (r.faceIndex[0] === C.NO_HIT) return null;
    const i = r.instance[0];
    const inst = this.instances[i];
    const inv = this._inverse.subarray(16 * i, 16 * i + 16);
    const lo = M.applyToPoint(inv, ray.origin);
    const ld = M.transformDirection(inv, ray.direction);
    const localRay = {
      origin: M.makeVec3(ray.origin, lo[0], lo[1], lo[2]), direction: M.makeVec3(ray.direction, ld[0], ld[1], ld[2]),
    };
    const hit = inst.bvh._makeHit(localRay, materialOrSide, r.faceIndex[0], r.distance[0],
      [r.point[0], r.point[1], r.point[2]], [r.barycoord[0], r.barycoord[1], r.barycoord[2]]);
    hit.instance = i;
    if (this.object) {
      hit.object = this.object;
      hit.batchId = i;
    }
    return hit;
  }

  // Closest point over the instances in world space (the reference's bvh_ClosestPointToPoint over a
  // ClusteredBVH, src/webgpu/shapecastFns/getClosestPointToPointFn.js:20-151): world points in (a
  // Float32Array of 3n); faceIndex (in the instance's geometry), instance, world distance and point,
  // barycoord out. Equal distances go to the lowest instance, then the lowest face.
  closestPointToPointBatch(points) {
    const q = points instanceof Float32Array ? points : Float32Array.from(points);
    const n = q.length / 3;
    const out = {
      faceIndex: new Uint32Array(n), instance: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    native.load().sceneClosestPoint(this._scene, q, out.faceIndex, out.instance, out.distance, out.point,
      out.barycoord);
    return out;
  }

  // One point (three's Vector3): { point, distance, faceIndex, instance } (and batchId for a
  // BatchedMesh scene), or null for a scene without visible instances.
  closestPointToPoint(point, target = {}) {
    const r = this.closestPointToPointBatch(new Float32Array([point.x, point.y, point.z]));
    if (r.faceIndex[0] === C.NO_HIT) return null;
    target.point = M.makeVec3(target.point || point, r.point[0], r.point[1], r.point[2]);
    target.distance = r.distance[0];
    target.faceIndex = r.faceIndex[0];
    target.instance = r.instance[0];
    if (this.object) target.batchId = r.instance[0];
    return target;
  }

  // The scene in the reference's WebGPU packing (BVHComputeData.update over ClusteredBVH's primitives,
  // src/webgpu/BVHComputeData.js:217-447, src/webgpu/ClusteredBVH.js:207-331): { nodes, transforms,
  // index } Uint32Arrays — BVHNode words (top level first: a leaf holds its subtree's node offset and
  // 0xFF000000 | transform slot; then each referenced subtree once, leaves rebased into the index), 36
  // words per transform slot (matrixWorld, its inverse, visible, padding) and the packed index (meshes
  // in the order the top level first names them, each rebased by the vertex count of the meshes before
  // it). primitiveLimit as BVHComputeData chooses it: a BatchedMesh's instances, and scenes of fewer
  // than three meshes, are referenced whole (Infinity); other scenes cut every root into clusters of
  // fewer than 64 triangles.
  exportWebGPU({ primitiveLimit } = {}) {
    if (primitiveLimit === undefined) primitiveLimit = this.object || this.instances.length < 3 ? Infinity : 64;
    return native.load().sceneExportWebGPU(this._scene, primitiveLimit);
  }

  dispose() {
    if (this._scene) native.load().freeScene(this._scene);
    this._scene = null;
  }
}

module.exports = { SceneBVH };
