'use strict';
// Host-side checks of the JavaScript package (no GPU needed): build, roots from groups, indirect,
// serialize / deserialize (incl. the version-0 fixup), API surface, and the error a query raises
// without a HIP device. Exits non-zero on the first failure.
const assert = require('assert');
const path = require('path');
const api = require(path.join(__dirname, '..', '..', 'three-mesh-bvh_amd', 'src', 'index.js'));
const { geometryFrom, BufferAttribute } = require('./shim.js');

function torus(radius, tube, radialSegments, tubularSegments) {
  const pos = [], idx = [];
  for (let j = 0; j <= radialSegments; j++) {
    for (let i = 0; i <= tubularSegments; i++) {
      const u = i / tubularSegments * Math.PI * 2, v = j / radialSegments * Math.PI * 2;
      pos.push((radius + tube * Math.cos(v)) * Math.cos(u), (radius + tube * Math.cos(v)) * Math.sin(u), tube * Math.sin(v));
    }
  }
  for (let j = 1; j <= radialSegments; j++) {
    for (let i = 1; i <= tubularSegments; i++) {
      const a = (tubularSegments + 1) * j + i - 1, b = (tubularSegments + 1) * (j - 1) + i - 1;
      const c = (tubularSegments + 1) * (j - 1) + i, d = (tubularSegments + 1) * j + i;
      idx.push(a, b, d, b, c, d);
    }
  }
  return geometryFrom(new Float32Array(pos), new Uint16Array(idx));
}

const checks = [];
function check(name, fn) { checks.push([name, fn]); }

check('exports the query API', () => {
  for (const k of ['MeshBVH', 'acceleratedRaycast', 'computeBoundsTree', 'disposeBoundsTree',
    'computeBatchedBoundsTree', 'disposeBatchedBoundsTree', 'getTriangleHitPointInfo', 'SAH', 'CENTER', 'AVERAGE',
    'setQueryLanes']) {
    assert.ok(api[k] !== undefined, k);
  }
  for (const m of ['raycast', 'raycastFirst', 'closestPointToPoint', 'raycastObject3D', 'raycastFirstBatch', 'refit', 'intersectsSphere', 'intersectsSphereBatch',
    'raycastBatch', 'closestPointToPointBatch', 'getBoundingBox', 'raycastFirstBatchAsync',
    'closestPointToPointBatchAsync', 'raycastFirstAsync', 'closestPointToPointAsync']) {
    assert.strictEqual(typeof api.MeshBVH.prototype[m], 'function', m);
  }
  assert.strictEqual(typeof api.MeshBVH.serialize, 'function');
  assert.strictEqual(typeof api.MeshBVH.deserialize, 'function');
});

check('build permutes the index in place and sets the bounding box', () => {
  const g = torus(1, 0.4, 12, 48);
  const before = Array.from(g.index.array);
  const bvh = new api.MeshBVH(g, { strategy: api.SAH, maxLeafTris: 10 });
  assert.strictEqual(bvh._roots.length, 1);
  assert.ok(bvh._roots[0].byteLength % 32 === 0 && bvh._roots[0].byteLength > 32);
  const tri = a => { const t = []; for (let i = 0; i < a.length; i += 3) t.push(`${a[i]},${a[i + 1]},${a[i + 2]}`); return t.sort(); };
  assert.deepStrictEqual(tri(Array.from(g.index.array)), tri(before));
  assert.ok(g.boundingBox.min.x <= -1.4 && g.boundingBox.max.x >= 1.4);
});

check('groups become roots; gaps are excluded', () => {
  const g = torus(1, 0.4, 12, 48);
  const n = g.index.count;
  g.addGroup(0, n / 2, 0);
  g.addGroup(n / 2 + 30, n / 2 - 30, 1);
  const bvh = new api.MeshBVH(g);
  assert.strictEqual(bvh._roots.length, 2);
});

check('indirect build leaves the index untouched', () => {
  const g = torus(1, 0.4, 12, 48);
  const before = Array.from(g.index.array);
  const bvh = new api.MeshBVH(g, { indirect: true });
  assert.ok(bvh.indirect);
  assert.deepStrictEqual(Array.from(g.index.array), before);
  assert.strictEqual(bvh._indirectBuffer.length, before.length / 3);
  assert.strictEqual(bvh.resolveTriangleIndex(0), bvh._indirectBuffer[0]);
});

check('non-indexed geometry gets an identity index', () => {
  const pos = new Float32Array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 1, 0, 1, 1]);
  const g = geometryFrom(pos, null);
  new api.MeshBVH(g);
  assert.ok(g.index && g.index.array.length === 6);
});

check('serialize / deserialize round trip', () => {
  const g = torus(1, 0.4, 12, 48);
  const bvh = new api.MeshBVH(g, { strategy: api.SAH });
  const data = api.MeshBVH.serialize(bvh);
  assert.strictEqual(data.version, 1);
  const g2 = torus(1, 0.4, 12, 48);
  const bvh2 = api.MeshBVH.deserialize(data, g2);
  assert.deepStrictEqual(new Uint8Array(bvh2._roots[0]), new Uint8Array(bvh._roots[0]));
  assert.deepStrictEqual(Array.from(g2.index.array), Array.from(g.index.array));
  const shared = api.MeshBVH.serialize(bvh, { cloneBuffers: false });
  assert.strictEqual(shared.roots[0], bvh._roots[0]);
});

check('version-0 data (absolute child offsets) is fixed up', () => {
  const g = torus(1, 0.4, 12, 48);
  const bvh = new api.MeshBVH(g, { strategy: api.SAH });
  const data = api.MeshBVH.serialize(bvh);
  const v0 = data.roots[0].slice(0);
  const u32 = new Uint32Array(v0), u16 = new Uint16Array(v0);
  for (let node = 0; node < v0.byteLength / 32; node++) {
    if (u16[16 * node + 15] !== 0xFFFF) u32[8 * node + 6] = (u32[8 * node + 6] + node) * 8;  // absolute u32 index
  }
  const bvh2 = api.MeshBVH.deserialize({ roots: [v0], index: data.index, indirectBuffer: null }, torus(1, 0.4, 12, 48));
  assert.deepStrictEqual(new Uint8Array(bvh2._roots[0]), new Uint8Array(bvh._roots[0]));
});

check('rejects non-BufferGeometry input', () => {
  assert.throws(() => new api.MeshBVH({}), /Only BufferGeometries/);
});

check('queries fail loudly without a HIP device', () => {
  if (api.deviceCount() > 0) return;
  const bvh = new api.MeshBVH(torus(1, 0.4, 12, 48));
  assert.throws(() => bvh.raycastFirst({ origin: { x: 0, y: 0, z: 5 }, direction: { x: 0, y: 0, z: -1 } }),
    /no HIP device/);
});

check('positions through accessors for non-float attributes', () => {
  const g = torus(1, 0.4, 6, 12);
  const f = g.attributes.position.array;
  const q = new Int16Array(f.length);
  for (let i = 0; i < f.length; i++) q[i] = Math.round(f[i] * 1000);
  const attr = new BufferAttribute(q, 3, true);
  attr.getX = i => q[3 * i] / 1000; attr.getY = i => q[3 * i + 1] / 1000; attr.getZ = i => q[3 * i + 2] / 1000;
  g.setAttribute('position', attr);
  const bvh = new api.MeshBVH(g);
  assert.ok(bvh._roots[0].byteLength > 32);
});

check('getTriangleHitPointInfo: face, material group, normal, barycoord, uv (closed form)', () => {
  // two right triangles; the second one in material group 2
  const pos = new Float32Array([0, 0, 0, 2, 0, 0, 0, 2, 0, 0, 0, 1, 0, 2, 1, 2, 0, 1]);
  const g = geometryFrom(pos, new Uint16Array([0, 1, 2, 3, 4, 5]));
  g.setAttribute('uv', new BufferAttribute(new Float32Array([0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1]), 2));
  g.addGroup(0, 3, 0);
  g.addGroup(3, 3, 2);
  const near = (x, y) => Math.abs(x - y) < 1e-12;
  // p = 0.5 a + 0.25 b + 0.25 c on triangle 0: normal (b - a) x (c - a) = +z
  const r0 = api.getTriangleHitPointInfo({ x: 0.5, y: 0.5, z: 0 }, g, 0, null);
  assert.deepStrictEqual([r0.face.a, r0.face.b, r0.face.c, r0.face.materialIndex], [0, 1, 2, 0]);
  assert.ok(near(r0.barycoord.x, 0.5) && near(r0.barycoord.y, 0.25) && near(r0.barycoord.z, 0.25));
  assert.ok(near(r0.face.normal.x, 0) && near(r0.face.normal.y, 0) && near(r0.face.normal.z, 1));
  assert.ok(near(r0.uv.x, 0.25) && near(r0.uv.y, 0.25));
  // vertex c of triangle 1 (reversed winding: normal -z), material 2, written into `target`
  const target = { face: { normal: { x: 0, y: 0, z: 0, set(x, y, z) { this.x = x; this.y = y; this.z = z; } } } };
  const r1 = api.getTriangleHitPointInfo({ x: 2, y: 0, z: 1 }, g, 1, target);
  assert.strictEqual(r1, target);
  assert.deepStrictEqual([r1.face.a, r1.face.b, r1.face.c, r1.face.materialIndex], [3, 4, 5, 2]);
  assert.ok(near(r1.barycoord.x, 0) && near(r1.barycoord.y, 0) && near(r1.barycoord.z, 1));
  assert.ok(near(r1.face.normal.z, -1) && near(r1.uv.x, 0) && near(r1.uv.y, 1));
});

// The micro-batch queue over stand-in batch methods (the device ones need a GPU): the calls of one
// turn become one batch per (side, near, far) or threshold group, answers go back in call order,
// a failed launch rejects every call of its group, and the next turn starts a new batch.
check('micro-batched per-call queries group the calls of one turn', async () => {
  const bvh = new api.MeshBVH(torus(1, 0.4, 12, 48));
  const launches = [];
  const result = (n, miss) => ({
    faceIndex: Uint32Array.from({ length: n }, (_, k) => (miss(k) ? 0xFFFFFFFF : 7)),
    distance: Float32Array.from({ length: n }, (_, k) => k + 0.5),
    point: new Float32Array(3 * n).fill(0.25), barycoord: new Float32Array(3 * n).fill(1 / 3),
  });
  bvh.raycastFirstBatchAsync = (o, d, opts) => {
    launches.push({ kind: 'ray', n: o.length / 3, o: Array.from(o), opts });
    if (opts.far === 13) return Promise.reject(new Error('launch failed'));
    return Promise.resolve(result(o.length / 3, k => o[3 * k] < 0));
  };
  bvh.closestPointToPointBatchAsync = (q, opts) => {
    launches.push({ kind: 'point', n: q.length / 3, opts });
    return Promise.resolve(result(q.length / 3, () => false));
  };
  const ray = x => ({ origin: { x, y: 0, z: 0 }, direction: { x: 0, y: 0, z: 1 } });
  const calls = [bvh.raycastFirstAsync(ray(1)), bvh.raycastFirstAsync(ray(-1), api.DoubleSide),
    bvh.raycastFirstAsync(ray(2)), bvh.raycastFirstAsync(ray(-3)), bvh.raycastFirstAsync(ray(4), api.FrontSide, 0, 13),
    bvh.closestPointToPointAsync({ x: 1, y: 2, z: 3 }), bvh.closestPointToPointAsync({ x: 1, y: 2, z: 3 }, {}, 0, 2)];
  assert.strictEqual(launches.length, 0);  // nothing launches before the turn ends
  const settled = await Promise.allSettled(calls);
  assert.deepStrictEqual(launches.map(l => [l.kind, l.n]), [['ray', 3], ['ray', 1], ['ray', 1], ['point', 1], ['point', 1]]);
  assert.deepStrictEqual(launches[0].o, [1, 0, 0, 2, 0, 0, -3, 0, 0]);  // call order within the group
  assert.strictEqual(launches[1].opts.side, api.DoubleSide);
  assert.strictEqual(launches[4].opts.maxThreshold, 2);
  const v = settled.map(x => x.value);
  assert.ok(v[0].faceIndex === 7 && v[0].distance === 0.5 && v[2].distance === 1.5 && v[3] === null && v[1] === null);
  assert.ok(settled[4].status === 'rejected' && /launch failed/.test(settled[4].reason.message));
  assert.ok(v[5].faceIndex === 7 && v[5].point.x === 0.25 && v[6].distance === 0.5);
  await bvh.raycastFirstAsync(ray(1));
  assert.strictEqual(launches.length, 6);  // a later turn: a batch of its own
});

// Per-group side masks live in the mesh's triangle rows: an asynchronous raycast with another
// material array waits for the ones in flight, and a synchronous one that would rewrite the mask
// under them throws (stand-in addon functions; the device ones need a GPU).
check('material-array masks wait for the asynchronous raycasts reading them', async () => {
  const g = torus(1, 0.4, 12, 48);
  const n = g.index.count;
  g.addGroup(0, n / 2, 0);
  g.addGroup(n / 2, n / 2, 1);
  const bvh = new api.MeshBVH(g);
  bvh._mesh = () => ({});
  const addon = require('../../three-mesh-bvh_amd/src/native.js').load();
  const saved = { launch: addon.raycastFirstAsync, mask: addon.setSideMask };
  const events = [];
  let pending = [];
  addon.setSideMask = (mesh, mask) => events.push(`mask ${mask[0]}`);
  addon.raycastFirstAsync = (mesh, o, d, side) => {
    events.push(`launch ${side}`);
    return new Promise(resolve => pending.push(resolve));
  };
  try {
    const mat = side => ({ isMaterial: true, side });
    const front = mat(api.FrontSide), back = mat(api.BackSide), both = mat(api.DoubleSide);
    const o = new Float32Array(3), d = new Float32Array([0, 0, 1]);
    const a = bvh.raycastFirstBatchAsync(o, d, { material: [front, back] });
    const b = bvh.raycastFirstBatchAsync(o, d, { material: [back, front] });  // another mask: waits
    const c = bvh.raycastFirstBatchAsync(o, d, { material: [front, back] });  // the same mask: runs now
    bvh.raycastFirstBatchAsync(o, d, { side: api.DoubleSide }).catch(() => {});  // no mask: runs now
    assert.throws(() => bvh.raycastFirstBatch(o, d, { material: [both, both] }), /in flight/);
    assert.deepStrictEqual(events, ['mask 1', 'launch 3', 'launch 3', `launch ${api.DoubleSide}`]);
    pending.forEach(r => r());
    pending = [];
    await Promise.all([a, c]);
    for (let i = 0; i < 4 && events.length < 6; i++) await new Promise(r => setImmediate(r));
    assert.deepStrictEqual(events.slice(4), ['mask 2', 'launch 3']);
    pending.forEach(r => r());
    await b;
  } finally {
    addon.raycastFirstAsync = saved.launch;
    addon.setSideMask = saved.mask;
  }
});

(async () => {
  let failed = 0;
  for (const [name, fn] of checks) {
    try {
      await fn();
      console.log(`ok   ${name}`);
    } catch (e) {
      failed++;
      console.log(`FAIL ${name}: ${e && e.stack ? e.stack : e}`);
    }
  }
  console.log(`${checks.length - failed}/${checks.length} passed`);
  process.exit(failed ? 1 : 0);
})();
