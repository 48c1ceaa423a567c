'use strict';
// MeshBVH with the reference's query API (src/core/MeshBVH.js, src/index.d.ts:151-167), backed by
// the MI355X kernels: the BVH is built on the host in the reference's packed format, uploaded to
// HBM once (on first query), and every query runs on the GPU. Per-call methods take one ray or
// point; the *Batch methods take packed Float32Arrays of many.
const C = require('./Constants.js');
const native = require('./native.js');
const M = require('./math.js');
const GU = require('./geometryUtils.js');

const FLOATS_PER_VEC = 3;

function getSide(materialOrSide) {
  return materialOrSide && materialOrSide.isMaterial ? materialOrSide.side : materialOrSide;
}

// Facing mask for one material side: bit 0 front-facing hits, bit 1 back-facing hits
// (checkIntersection, src/utils/ThreeRayIntersectUtilities.js:21-47; an undefined side behaves as
// FrontSide there because only BackSide / DoubleSide are tested).
function sideMask(side) {
  if (side === C.BackSide) return 2;
  if (side === C.DoubleSide) return 3;
  return 1;
}

// Query arrays: a Float32Array is traced as FP32 values; a Float64Array (or a plain array of JS
// numbers) as the doubles it holds, every decision the reference would make for those doubles.
function queryArray(a) {
  return a instanceof Float32Array || a instanceof Float64Array ? a : Float64Array.from(a);
}
function queryArrays(a, b) {
  const x = queryArray(a), y = queryArray(b);
  if (x.constructor === y.constructor) return [x, y];
  return [Float64Array.from(x), Float64Array.from(y)];
}

function f32OrNull(a) {
  if (a === undefined || a === null) return null;
  return a instanceof Float32Array ? a : Float32Array.from(a);
}

class MeshBVH {
  constructor(geometry, options = {}) {
    if (!geometry || !geometry.isBufferGeometry) {
      throw new Error('BVH: Only BufferGeometries are supported.');
    } else if (geometry.index && geometry.index.isInterleavedBufferAttribute) {
      throw new Error('BVH: InterleavedBufferAttribute is not supported for the index attribute.');
    }
    if (options.useSharedArrayBuffer && typeof SharedArrayBuffer === 'undefined') {
      throw new Error('BVH: SharedArrayBuffer is not available.');
    }
    if (options.maxLeafTris) {
      options = Object.assign({}, options, { targetLeafSize: options.maxLeafTris });
    }
    this.geometry = geometry;
    this._roots = null;
    this._indirectBuffer = null;
    this._gpu = null;        // native mesh handle (uploaded on first query)
    this._maskKey = null;    // material-array signature of the uploaded facing masks
    options = Object.assign({}, C.DEFAULT_OPTIONS, options);
    if (!options[C.SKIP_GENERATION]) this.init(options);
  }

  get indirect() {
    return !!this._indirectBuffer;
  }

  get resolveTriangleIndex() {
    return this.indirect ? i => this._indirectBuffer[i] : i => i;
  }

  get primitiveStride() {
    return 3;
  }

  init(options) {
    const geometry = this.geometry;
    let ranges;
    if (options.indirect) {
      ranges = GU.getRootPrimitiveRanges(geometry, options.range, 3);
    } else {
      GU.ensureIndex(geometry, options);
      ranges = GU.getRootPrimitiveRanges(geometry, options.range, 3);
    }
    const flat = new Uint32Array(2 * ranges.length);
    ranges.forEach((r, i) => {
      flat[2 * i] = r.offset;
      flat[2 * i + 1] = r.count;
    });
    const pos = GU.positionLayout(geometry);
    const index = geometry.index ? geometry.index.array : null;
    // gpuBuild: the same tree built on the current device (byte-for-byte the host build's)
    const out = native.load().build(pos.array, pos.stride, pos.offset, index, flat, {
      strategy: options.strategy,
      maxDepth: options.maxDepth,
      targetLeafSize: options.targetLeafSize,
      indirect: !!options.indirect,
      device: options.gpuBuild ? native.getDevice() : -1,
    });
    let roots = out.roots;
    if (options.useSharedArrayBuffer) {
      roots = roots.map(r => {
        const s = new SharedArrayBuffer(r.byteLength);
        new Uint8Array(s).set(new Uint8Array(r));
        return s;
      });
    }
    this._roots = roots;
    this._indirectBuffer = out.indirect;
    if (geometry.index && !options.indirect) geometry.index.needsUpdate = true;
    if (!geometry.boundingBox && options.setBoundingBox) {
      const bb = this.getBoundingBox({ min: new M.Vec3(), max: new M.Vec3() });
      geometry.boundingBox = bb;
    }
  }

  // ---- serialization (src/core/MeshBVH.js:112-232) ------------------------------------------

  static serialize(bvh, options = {}) {
    options = Object.assign({ cloneBuffers: true }, options);
    const geometry = bvh.geometry;
    const indexAttr = geometry.index;
    const result = { version: 1, roots: null, index: null, indirectBuffer: null };
    if (options.cloneBuffers) {
      result.roots = bvh._roots.map(r => r.slice(0));
      result.index = indexAttr ? indexAttr.array.slice() : null;
      result.indirectBuffer = bvh._indirectBuffer ? bvh._indirectBuffer.slice() : null;
    } else {
      result.roots = bvh._roots;
      result.index = indexAttr ? indexAttr.array : null;
      result.indirectBuffer = bvh._indirectBuffer;
    }
    return result;
  }

  static deserialize(data, geometry, options = {}) {
    options = Object.assign({ setIndex: true, indirect: Boolean(data.indirectBuffer) }, options);
    const { index, roots, indirectBuffer } = data;
    if (!data.version) {
      // version 0 stored absolute uint32 offsets of right children; convert to node distances
      for (const root of roots) {
        const u32 = new Uint32Array(root);
        const u16 = new Uint16Array(root);
        for (let node = 0, l = root.byteLength / C.BYTES_PER_NODE; node < l; node++) {
          const n32 = C.UINT32_PER_NODE * node;
          if (u16[2 * n32 + 15] !== C.IS_LEAFNODE_FLAG) u32[n32 + 6] = u32[n32 + 6] / C.UINT32_PER_NODE - node;
        }
      }
    }
    const bvh = new MeshBVH(geometry, Object.assign({}, options, { [C.SKIP_GENERATION]: true }));
    bvh._roots = roots;
    bvh._indirectBuffer = indirectBuffer || null;
    if (options.setIndex) {
      const indexAttr = geometry.index;
      if (!indexAttr) {
        const attr = GU.makeIndexAttribute(geometry, index);
        if (typeof geometry.setIndex === 'function') geometry.setIndex(attr);
        else geometry.index = attr;
      } else if (indexAttr.array !== index) {
        indexAttr.array.set(index);
        indexAttr.needsUpdate = true;
      }
    }
    return bvh;
  }

  getBoundingBox(target) {
    let mnx = Infinity, mny = Infinity, mnz = Infinity, mxx = -Infinity, mxy = -Infinity, mxz = -Infinity;
    for (const root of this._roots) {
      const f = new Float32Array(root, 0, 6);
      mnx = Math.min(mnx, f[0]);
      mny = Math.min(mny, f[1]);
      mnz = Math.min(mnz, f[2]);
      mxx = Math.max(mxx, f[3]);
      mxy = Math.max(mxy, f[4]);
      mxz = Math.max(mxz, f[5]);
    }
    if (target.min && typeof target.min.set === 'function') {
      target.min.set(mnx, mny, mnz);
      target.max.set(mxx, mxy, mxz);
    } else {
      target.min = new M.Vec3(mnx, mny, mnz);
      target.max = new M.Vec3(mxx, mxy, mxz);
    }
    return target;
  }

  // ---- device residency ------------------------------------------------------------------------

  // Uploads the BVH, index and positions to the GPU (done automatically on the first query).
  upload() {
    const addon = native.load();
    if (this._gpu) addon.freeMesh(this._gpu);
    const geo = this.geometry;
    const pos = GU.positionLayout(geo);
    this._gpu = addon.uploadMesh(native.getDevice(), this._roots, geo.index ? geo.index.array : null,
      this._indirectBuffer, pos.array, pos.stride, pos.offset);
    this._maskKey = null;
    return this;
  }

  // refit (src/core/MeshBVH.js:504-508 over src/core/cast/refit.template.js): the node boxes are
  // recomputed on the GPU from the geometry's current positions with the reference's rules (a leaf
  // is the min / max of its triangles' vertices and is rewritten only if it differs, an internal
  // node is the union of its children, recomputed only if a child changed). The refit nodes are
  // written back into this._roots, so serialize() and later uploads see them, and the device copy
  // is updated in place. nodeIndices (a Set or an Array of node indices over all roots, as shapecast
  // reports them) makes it a partial refit (refit.template.js:136-172): below each root only the
  // listed children are visited, or both children, whole, when neither is listed; values that are
  // not node indices match nothing, as in the reference's Set.has.
  refit(nodeIndices = null) {
    let ids = null;
    if (nodeIndices) {
      const list = [];
      for (const v of nodeIndices) if (Number.isInteger(v) && v >= 0 && v < 4294967296) list.push(v);
      ids = Uint32Array.from(list);
    }
    const pos = GU.positionLayout(this.geometry);
    native.load().refitMesh(this._mesh(), pos.array, pos.stride, pos.offset, this._roots, ids);
  }

  // Releases the device copy now (otherwise released when the BVH is garbage collected).
  dispose() {
    if (this._gpu) native.load().freeMesh(this._gpu);
    this._gpu = null;
  }

  _mesh() {
    if (!this._gpu) this.upload();
    return this._gpu;
  }

  // Side argument for the kernels; material arrays become per-face facing masks from the groups
  // (intersectTri's group loop, src/utils/ThreeRayIntersectUtilities.js:171-202).
  // The mask is written into the mesh's triangle rows on the device, so it cannot change while an
  // asynchronous query that reads it is in flight (raycastFirstBatchAsync waits for those instead).
  _maskKeyFor(materialOrSide) {
    const groups = this.geometry.groups || [];
    const sides = groups.map(g => sideMask(getSide(materialOrSide[g.materialIndex])));
    return { groups, sides, key: groups.map((g, i) => `${g.start}:${g.count}:${sides[i]}`).join(',') };
  }

  _sideFor(materialOrSide) {
    if (!Array.isArray(materialOrSide)) {
      const side = getSide(materialOrSide);
      return side === C.BackSide ? C.BackSide : (side === C.DoubleSide ? C.DoubleSide : C.FrontSide);
    }
    const { groups, sides, key } = this._maskKeyFor(materialOrSide);
    if (key !== this._maskKey) {
      if (this._maskInFlight > 0) {
        throw new Error('BVH: an asynchronous raycast with other materials is in flight on this BVH; await it first');
      }
      const faceCount = (this.geometry.index ? this.geometry.index.count : this.geometry.attributes.position.count) / 3;
      const mask = new Uint8Array(faceCount);
      groups.forEach((g, i) => {
        const end = Math.min(faceCount, (g.start + g.count) / 3);
        for (let f = Math.floor(g.start / 3); f < end; f++) mask[f] |= sides[i];
      });
      native.load().setSideMask(this._mesh(), mask);
      this._maskKey = key;
    }
    return 3;  // MBVH_SIDE_PER_TRIANGLE
  }

  // ---- batch queries ---------------------------------------------------------------------------

  // First hit for n rays: origins / directions are arrays of length 3n (Float32Array: FP32 rays;
  // Float64Array or plain arrays: double rays, traced as the reference traces them). Returns packed
  // results; faceIndex is 0xFFFFFFFF and distance Infinity for a miss.
  raycastFirstBatch(origins, directions, options = {}) {
    const [o, d] = queryArrays(origins, directions);
    const n = o.length / FLOATS_PER_VEC;
    const side = this._sideFor(options.side !== undefined ? options.side : (options.material !== undefined ? options.material : C.FrontSide));
    const out = {
      faceIndex: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    native.load().raycastFirst(this._mesh(), o, d, side, options.near !== undefined ? options.near : 0,
      options.far !== undefined ? options.far : Infinity, out.faceIndex, out.distance, out.point, out.barycoord);
    return out;
  }

  // All hits for n rays, in the reference's traversal order: hits of ray i are entries
  // offsets[i] .. offsets[i + 1] - 1 of the packed arrays.
  raycastBatch(origins, directions, options = {}) {
    const [o, d] = queryArrays(origins, directions);
    const side = this._sideFor(options.side !== undefined ? options.side : (options.material !== undefined ? options.material : C.FrontSide));
    return native.load().raycast(this._mesh(), o, d, side, options.near !== undefined ? options.near : 0,
      options.far !== undefined ? options.far : Infinity);
  }

  // Closest point for n query points (length 3n; Float32Array: FP32 points, Float64Array or plain
  // arrays: double points).
  closestPointToPointBatch(points, options = {}) {
    const q = queryArray(points);
    const n = q.length / FLOATS_PER_VEC;
    const out = {
      faceIndex: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    native.load().closestPoint(this._mesh(), q, options.minThreshold || 0,
      options.maxThreshold !== undefined ? options.maxThreshold : Infinity, out.faceIndex, out.distance, out.point,
      out.barycoord);
    return out;
  }

  // The same two batch queries without blocking the event loop: the device call runs on the
  // addon's query thread and the Promise resolves to the same packed results. The input arrays must not
  // change until it settles; the BVH cannot be disposed meanwhile (dispose() throws).
  raycastFirstBatchAsync(origins, directions, options = {}) {
    const mos = options.side !== undefined ? options.side : (options.material !== undefined ? options.material : C.FrontSide);
    if (Array.isArray(mos) && this._maskInFlight > 0 && this._maskKeyFor(mos).key !== this._maskKey) {
      // another material array's mask is in use: start once the queries reading it have settled
      return this._maskSettled.then(() => this.raycastFirstBatchAsync(origins, directions, options));
    }
    const [o, d] = queryArrays(origins, directions);
    const n = o.length / FLOATS_PER_VEC;
    const side = this._sideFor(mos);
    const out = {
      faceIndex: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    const p = native.load().raycastFirstAsync(this._mesh(), o, d, side, options.near !== undefined ? options.near : 0,
      options.far !== undefined ? options.far : Infinity, out.faceIndex, out.distance, out.point, out.barycoord)
      .then(() => out);
    if (side === 3) {  // reads the per-triangle mask
      this._maskInFlight = (this._maskInFlight || 0) + 1;
      const settled = p.then(() => undefined, () => undefined).then(() => { this._maskInFlight--; });
      this._maskSettled = this._maskSettled ? Promise.all([this._maskSettled, settled]) : settled;
    }
    return p;
  }

  closestPointToPointBatchAsync(points, options = {}) {
    const q = queryArray(points);
    const n = q.length / FLOATS_PER_VEC;
    const out = {
      faceIndex: new Uint32Array(n), distance: new Float32Array(n),
      point: new Float32Array(3 * n), barycoord: new Float32Array(3 * n),
    };
    return native.load().closestPointAsync(this._mesh(), q, options.minThreshold || 0,
      options.maxThreshold !== undefined ? options.maxThreshold : Infinity, out.faceIndex, out.distance, out.point,
      out.barycoord).then(() => out);
  }

  // intersectsSphere for n spheres: centers a Float32Array of length 3n, radii one of length n.
  // Returns a Uint8Array (1: some triangle within the radius, inclusive).
  // intersectsBox (src/core/MeshBVH.js:780-792): box = { min, max } in its own frame, boxToMesh its
  // Matrix4 (or 16 numbers, column-major) into the mesh frame. Decided on the GPU in double with the
  // reference's separating-axis tests.
  intersectsBox(box, boxToMesh) {
    const e = boxToMesh && boxToMesh.elements ? boxToMesh.elements : boxToMesh;
    return this.intersectsBoxBatch(new Float64Array([box.min.x, box.min.y, box.min.z]),
      new Float64Array([box.max.x, box.max.y, box.max.z]), Float64Array.from(e))[0] === 1;
  }

  // Many boxes: Float64Arrays of 3n mins, 3n maxs and 16n matrices -> Uint8Array (1 intersects).
  intersectsBoxBatch(mins, maxs, matrices) {
    const out = new Uint8Array(mins.length / 3);
    native.load().intersectsBox(this._mesh(), Float64Array.from(mins), Float64Array.from(maxs),
      Float64Array.from(matrices), out);
    return out;
  }

  // closestPointToGeometry (src/core/MeshBVH.js:840-853): the closest pair of triangles between this
  // BVH's geometry and otherGeometry placed by geometryToBvh (Matrix4 or 16 numbers). target1 gets the
  // point on this geometry, distance and faceIndex (the triangle slot the reference reports); target2
  // (unless null) the point on otherGeometry in its own frame, its distance and faceIndex. Returns
  // target1, or null when no pair is closer than maxThreshold. The pair is the closest overall (equal
  // distances: the lowest other triangle, then the lowest slot); the reference may stop earlier at any
  // pair under minThreshold, and the closest pair is under it whenever such a pair exists.
  closestPointToGeometry(otherGeometry, geometryToBvh, target1 = {}, target2 = {}, minThreshold = 0,
    maxThreshold = Infinity) {
    const e = geometryToBvh && geometryToBvh.elements ? geometryToBvh.elements : geometryToBvh;
    const pos = GU.positionLayout(otherGeometry);
    const index = otherGeometry.index ? otherGeometry.index.array : null;
    const triCount = Math.floor((otherGeometry.index ? otherGeometry.index.count
      : otherGeometry.attributes.position.count) / 3);
    const out = new Float64Array(10);
    const found = native.load().closestPointToGeometry(this._mesh(), pos.array, pos.stride, pos.offset,
      index, triCount, Float64Array.from(e), minThreshold, maxThreshold, out);
    if (!found) return null;
    const vec = (t, x, y, z) => {
      if (t.point && t.point.set) t.point.set(x, y, z);
      else t.point = M.makeVec3(t.point, x, y, z);
    };
    vec(target1, out[1], out[2], out[3]);
    target1.distance = out[0];
    target1.faceIndex = out[8];
    if (target2) {
      vec(target2, out[4], out[5], out[6]);
      target2.distance = out[7];
      target2.faceIndex = out[9];
    }
    return target1;
  }

  intersectsSphereBatch(centers, radii) {
    const c = centers instanceof Float32Array ? centers : Float32Array.from(centers);
    const r = radii instanceof Float32Array ? radii : Float32Array.from(radii);
    const out = new Uint8Array(r.length);
    native.load().intersectsSphere(this._mesh(), c, r, out);
    return out;
  }

  // ---- differentiable queries: autograd-style forward / backward pairs ---------------------------
  // (no counterpart in the reference). forward() runs the batch query and returns its outputs
  // plus `saved`, the hit data the backward reuses (it never re-traverses). backward(saved, grads)
  // takes the upstream gradients of a scalar loss with respect to the forward outputs and returns
  // the gradients with respect to the inputs and to the vertex positions (one x, y, z triple per
  // vertex of the position attribute). options.deterministic (default true) selects the sorted
  // segmented vertex reduction (bitwise reproducible) instead of FP32 atomics;
  // options.positions === false skips vertex gradients.

  raycastFirstForward(origins, directions, options = {}) {
    const d = directions instanceof Float32Array ? directions.slice() : Float32Array.from(directions);
    const out = this.raycastFirstBatch(origins, d, options);
    out.saved = { kind: 'raycastFirst', directions: d, faceIndex: out.faceIndex, distance: out.distance,
      barycoord: out.barycoord };
    return out;
  }

  // grads: { distance?, point?, barycoord? } (Float32Arrays of n, 3n, 3n; missing = 0).
  raycastFirstBackward(saved, grads = {}, options = {}) {
    if (!saved || saved.kind !== 'raycastFirst') throw new Error('raycastFirstBackward: `saved` must come from raycastFirstForward');
    const n = saved.faceIndex.length;
    const addon = native.load();
    const mesh = this._mesh();
    const out = {
      origins: new Float32Array(3 * n), directions: new Float32Array(3 * n),
      positions: options.positions === false ? null : new Float32Array(3 * addon.meshInfo(mesh).vertexCount),
    };
    addon.raycastFirstBackward(mesh, saved.directions, saved.faceIndex, saved.distance, saved.barycoord,
      f32OrNull(grads.distance), f32OrNull(grads.point), f32OrNull(grads.barycoord),
      out.origins, out.directions, out.positions, options.deterministic === false ? 0 : 1);
    return out;
  }

  closestPointToPointForward(points, options = {}) {
    const q = points instanceof Float32Array ? points.slice() : Float32Array.from(points);
    const out = this.closestPointToPointBatch(q, options);
    out.saved = { kind: 'closestPointToPoint', points: q, faceIndex: out.faceIndex, distance: out.distance,
      point: out.point, barycoord: out.barycoord };
    return out;
  }

  // grads: { distance } (Float32Array of n).
  closestPointToPointBackward(saved, grads = {}, options = {}) {
    if (!saved || saved.kind !== 'closestPointToPoint') {
      throw new Error('closestPointToPointBackward: `saved` must come from closestPointToPointForward');
    }
    const n = saved.faceIndex.length;
    const addon = native.load();
    const mesh = this._mesh();
    const gd = f32OrNull(grads.distance);
    if (!gd || gd.length !== n) throw new Error('closestPointToPointBackward: grads.distance must have one entry per point');
    const out = {
      points: new Float32Array(3 * n),
      positions: options.positions === false ? null : new Float32Array(3 * addon.meshInfo(mesh).vertexCount),
    };
    addon.closestPointBackward(mesh, saved.points, saved.faceIndex, saved.distance, saved.point, saved.barycoord,
      gd, out.points, out.positions, options.deterministic === false ? 0 : 1);
    return out;
  }

  // ---- the reference's per-call API ------------------------------------------------------------

  raycastFirst(ray, materialOrSide = C.FrontSide, near = 0, far = Infinity) {
    // the ray's JS numbers as they are (the reference traces them in double precision)
    const o = new Float64Array([ray.origin.x, ray.origin.y, ray.origin.z]);
    const d = new Float64Array([ray.direction.x, ray.direction.y, ray.direction.z]);
    const r = this.raycastFirstBatch(o, d, { side: materialOrSide, near, far });
    if (r.faceIndex[0] === C.NO_HIT) return null;
    return this._makeHit(ray, materialOrSide, r.faceIndex[0], r.distance[0],
      [r.point[0], r.point[1], r.point[2]], [r.barycoord[0], r.barycoord[1], r.barycoord[2]]);
  }

  // All hits of one ray (src/core/MeshBVH.js:532-545). With a material array a triangle yields
  // one hit per group whose material admits it, in group order, as intersectTri does.
  raycast(ray, materialOrSide = C.FrontSide, near = 0, far = Infinity) {
    const o = new Float64Array([ray.origin.x, ray.origin.y, ray.origin.z]);
    const d = new Float64Array([ray.direction.x, ray.direction.y, ray.direction.z]);
    const r = this.raycastBatch(o, d, { side: materialOrSide, near, far });
    const hits = [];
    for (let k = 0; k < r.faceIndex.length; k++) {
      const p = [r.point[3 * k], r.point[3 * k + 1], r.point[3 * k + 2]];
      const b = [r.barycoord[3 * k], r.barycoord[3 * k + 1], r.barycoord[3 * k + 2]];
      if (!Array.isArray(materialOrSide)) {
        hits.push(this._makeHit(ray, materialOrSide, r.faceIndex[k], r.distance[k], p, b));
        continue;
      }
      const first = this._makeHit(ray, materialOrSide, r.faceIndex[k], r.distance[k], p, b);
      const back = first.face.normal.x * ray.direction.x + first.face.normal.y * ray.direction.y +
        first.face.normal.z * ray.direction.z > 0;
      const start = r.faceIndex[k] * 3;
      for (const g of this.geometry.groups || []) {
        if (start >= g.start && start < g.start + g.count &&
          (sideMask(getSide(materialOrSide[g.materialIndex])) & (back ? 2 : 1))) {
          const h = this._makeHit(ray, materialOrSide, r.faceIndex[k], r.distance[k], p, b);
          h.face.materialIndex = g.materialIndex;
          hits.push(h);
        }
      }
    }
    return hits;
  }

  // intersectsSphere (src/core/MeshBVH.js:800-809): whether any triangle lies within the sphere.
  intersectsSphere(sphere) {
    const c = sphere.center;
    return this.intersectsSphereBatch(new Float32Array([c.x, c.y, c.z]), new Float32Array([sphere.radius]))[0] === 1;
  }

  closestPointToPoint(point, target = {}, minThreshold = 0, maxThreshold = Infinity) {
    const q = new Float64Array([point.x, point.y, point.z]);
    const r = this.closestPointToPointBatch(q, { minThreshold, maxThreshold });
    if (r.faceIndex[0] === C.NO_HIT) return null;
    if (!target.point) target.point = M.makeVec3(point, r.point[0], r.point[1], r.point[2]);
    else target.point.set(r.point[0], r.point[1], r.point[2]);
    target.distance = r.distance[0];
    target.faceIndex = r.faceIndex[0];
    return target;
  }

  // ---- micro-batched per-call queries -----------------------------------------------------------
  // raycastFirstAsync / closestPointToPointAsync take what raycastFirst / closestPointToPoint take
  // and return Promises of what those return. Calls made in one turn of the event loop with the same
  // material (or side) and near / far, or the same thresholds, are answered by ONE asynchronous batch
  // launch when the turn ends (setImmediate), so a loop of per-call queries costs one device round
  // trip instead of one each. The tree is read when the batch runs: a refit later in the same turn
  // applies to the calls queued before it.
  raycastFirstAsync(ray, materialOrSide = C.FrontSide, near = 0, far = Infinity) {
    const { origin: o, direction: d } = ray;
    return this._enqueue('ray', materialOrSide, near, far, [o.x, o.y, o.z], [d.x, d.y, d.z], ray);
  }

  closestPointToPointAsync(point, target = {}, minThreshold = 0, maxThreshold = Infinity) {
    return this._enqueue('point', null, minThreshold, maxThreshold, [point.x, point.y, point.z], null,
      { point, target });
  }

  _enqueue(kind, materialOrSide, a, b, first, second, item) {
    return new Promise((resolve, reject) => {
      if (!this._queued) this._queued = [];
      let g = this._queued.find(x => x.kind === kind && x.materialOrSide === materialOrSide &&
        Object.is(x.a, a) && Object.is(x.b, b));
      if (!g) {
        g = { kind, materialOrSide, a, b, first: [], second: [], items: [] };
        this._queued.push(g);
      }
      g.first.push(first[0], first[1], first[2]);
      if (second) g.second.push(second[0], second[1], second[2]);
      g.items.push({ item, resolve, reject });
      if (this._queued.length === 1 && g.items.length === 1) {
        const later = typeof setImmediate === 'function' ? setImmediate : f => setTimeout(f, 0);
        later(() => this._flushQueued());
      }
    });
  }

  _flushQueued() {
    const groups = this._queued || [];
    this._queued = null;
    for (const g of groups) {
      let p;
      try {
        p = g.kind === 'ray'
          ? this.raycastFirstBatchAsync(Float64Array.from(g.first), Float64Array.from(g.second),
            { side: g.materialOrSide, near: g.a, far: g.b })
          : this.closestPointToPointBatchAsync(Float64Array.from(g.first), { minThreshold: g.a, maxThreshold: g.b });
      } catch (e) {
        p = Promise.reject(e);
      }
      p.then((r) => {
        g.items.forEach(({ item, resolve }, k) => {
          if (r.faceIndex[k] === C.NO_HIT) {
            resolve(null);
            return;
          }
          const pt = [r.point[3 * k], r.point[3 * k + 1], r.point[3 * k + 2]];
          if (g.kind === 'ray') {
            resolve(this._makeHit(item, g.materialOrSide, r.faceIndex[k], r.distance[k], pt,
              [r.barycoord[3 * k], r.barycoord[3 * k + 1], r.barycoord[3 * k + 2]]));
            return;
          }
          const { point, target } = item;
          if (!target.point) target.point = M.makeVec3(point, pt[0], pt[1], pt[2]);
          else target.point.set(pt[0], pt[1], pt[2]);
          target.distance = r.distance[k];
          target.faceIndex = r.faceIndex[k];
          resolve(target);
        });
      }, (e) => g.items.forEach(({ reject }) => reject(e)));
    }
  }

  // raycastObject3D (src/core/MeshBVH.js:447-494): world ray to local, query, hits back to world.
  raycastObject3D(object, raycaster, intersects = []) {
    const { material } = object;
    if (material === undefined) return undefined;
    const mw = object.matrixWorld.elements;
    const inv = M.invert(mw);
    const lo = M.applyToPoint(inv, raycaster.ray.origin);
    const ld = M.transformDirection(inv, raycaster.ray.direction);
    const s = M.matrixScale(mw);
    const scaleFactor = M.length3(ld[0] * s[0], ld[1] * s[1], ld[2] * s[2]);
    const near = raycaster.near / scaleFactor;
    const far = raycaster.far / scaleFactor;
    const localRay = {
      origin: M.makeVec3(raycaster.ray.origin, lo[0], lo[1], lo[2]),
      direction: M.makeVec3(raycaster.ray.direction, ld[0], ld[1], ld[2]),
    };
    const convert = hit => {
      // convertRaycastIntersect (src/utils/GeometryRayIntersectUtilities.js:3-17)
      const p = M.applyToPoint(mw, hit.point);
      hit.point.set(p[0], p[1], p[2]);
      hit.distance = hit.point.distanceTo(raycaster.ray.origin);
      hit.object = object;
      return hit;
    };
    if (raycaster.firstHitOnly === true) {
      const hit = this.raycastFirst(localRay, material, near, far);
      if (hit) intersects.push(convert(hit));
    } else {
      const hits = this.raycast(localRay, material, near, far);
      for (const hit of hits) intersects.push(convert(hit));
    }
    return intersects;
  }

  // Builds the three.js intersection object (checkBufferGeometryIntersection,
  // src/utils/ThreeRayIntersectUtilities.js:49-145) from the kernel's face, point and barycoord.
  _makeHit(ray, materialOrSide, faceIndex, distance, p, bary) {
    const geo = this.geometry;
    const idx = geo.index ? geo.index.array : null;
    const a = idx ? idx[3 * faceIndex] : 3 * faceIndex;
    const b = idx ? idx[3 * faceIndex + 1] : 3 * faceIndex + 1;
    const c = idx ? idx[3 * faceIndex + 2] : 3 * faceIndex + 2;
    const posAttr = geo.attributes.position;
    const va = [posAttr.getX(a), posAttr.getY(a), posAttr.getZ(a)];
    const vb = [posAttr.getX(b), posAttr.getY(b), posAttr.getZ(b)];
    const vc = [posAttr.getX(c), posAttr.getY(c), posAttr.getZ(c)];
    const like = ray.origin;
    const hit = { distance, point: M.makeVec3(like, p[0], p[1], p[2]) };
    const interp = (attr, Make) => {
      const n = attr.itemSize;
      const get = (i, k) => (k === 0 ? attr.getX(i) : k === 1 ? attr.getY(i) : attr.getZ(i));
      const v = [];
      for (let k = 0; k < n && k < 3; k++) v.push(get(a, k) * bary[0] + get(b, k) * bary[1] + get(c, k) * bary[2]);
      return Make(v);
    };
    const attrs = geo.attributes;
    if (attrs.uv) hit.uv = interp(attrs.uv, v => M.makeVec2(null, v[0], v[1]));
    if (attrs.uv1) hit.uv1 = interp(attrs.uv1, v => M.makeVec2(null, v[0], v[1]));
    if (attrs.normal) {
      const nrm = interp(attrs.normal, v => M.makeVec3(like, v[0], v[1], v[2]));
      if (nrm.x * ray.direction.x + nrm.y * ray.direction.y + nrm.z * ray.direction.z > 0) nrm.multiplyScalar(-1);
      hit.normal = nrm;
    }
    const n = M.triangleNormal(va, vb, vc);
    let materialIndex = 0;
    if (Array.isArray(materialOrSide)) {
      // first group holding the face whose material admits the hit's facing
      const e1 = [vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]];
      const e2 = [vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]];
      const nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
      const back = nx * ray.direction.x + ny * ray.direction.y + nz * ray.direction.z > 0;
      const first = faceIndex * 3;
      for (const g of geo.groups || []) {
        if (first >= g.start && first < g.start + g.count &&
          (sideMask(getSide(materialOrSide[g.materialIndex])) & (back ? 2 : 1))) {
          materialIndex = g.materialIndex;
          break;
        }
      }
    }
    hit.face = { a, b, c, normal: M.makeVec3(like, n[0], n[1], n[2]), materialIndex };
    hit.faceIndex = faceIndex;
    hit.barycoord = M.makeVec3(like, bary[0], bary[1], bary[2]);
    return hit;
  }
}

module.exports = { MeshBVH };
