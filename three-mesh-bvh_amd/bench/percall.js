'use strict';
// Per-call latency of the drop-in path: one ray per MeshBVH.raycastFirst call and one point per
// closestPointToPoint call, as acceleratedRaycast and a three.js app issue them (the JS numbers go
// through the Node-API addon and the C ABI to the GPU and back, synchronously). Reads a geometry
// and queries from <dir> (positions.f32, index.u16 or index.u32, rays.f32 = origins then
// directions, points.f32, meta.json {index_bytes}) and prints one JSON line. Also times the same
// per-call queries through the micro-batched Promise forms (raycastFirstAsync, closestPointToPointAsync).
const fs = require('fs');
const path = require('path');
const api = require(path.join(__dirname, '..', 'src', 'index.js'));

const dir = process.argv[2];
const calls = Number(process.argv[3] || 2000);
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
function read(name, Type) {
  const b = fs.readFileSync(path.join(dir, name));
  return new Type(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
const positions = read('positions.f32', Float32Array);
const index = meta.index_bytes === 2 ? read('index.u16', Uint16Array) : read('index.u32', Uint32Array);
const rays = read('rays.f32', Float32Array);
const points = read('points.f32', Float32Array);
const nRays = rays.length / 6;
const nPts = points.length / 3;

// the parts of a three.js BufferGeometry the package reads
const attr = (array, itemSize) => ({ array, itemSize, count: array.length / itemSize, isBufferAttribute: true,
  getX(i) { return array[i * itemSize]; }, getY(i) { return array[i * itemSize + 1]; },
  getZ(i) { return array[i * itemSize + 2]; } });
const geometry = { isBufferGeometry: true, attributes: { position: attr(positions, 3) }, index: attr(index, 1),
  groups: [], drawRange: { start: 0, count: Infinity }, boundingBox: null,
  setIndex(a) { this.index = a; return this; } };

const t0 = process.hrtime.bigint();
const bvh = new api.MeshBVH(geometry, { strategy: api.SAH, maxLeafTris: 10 });
bvh.upload();
const buildMs = Number(process.hrtime.bigint() - t0) / 1e6;

const v = (x, y, z) => ({ x, y, z });
function timeCalls(n, fn) {
  for (let i = 0; i < 100; i++) fn(i);  // warm-up (first upload, JIT)
  const per = [];
  for (let rep = 0; rep < 5; rep++) {
    const a = process.hrtime.bigint();
    for (let i = 0; i < n; i++) fn(i);
    per.push(Number(process.hrtime.bigint() - a) / 1e3 / n);
  }
  per.sort((x, y) => x - y);
  return per[2];  // median of 5 rounds, microseconds per call
}
let hits = 0, found = 0;
const rayUs = timeCalls(calls, (i) => {
  const k = i % nRays;
  const ray = { origin: v(rays[3 * k], rays[3 * k + 1], rays[3 * k + 2]),
    direction: v(rays[3 * (nRays + k)], rays[3 * (nRays + k) + 1], rays[3 * (nRays + k) + 2]) };
  if (bvh.raycastFirst(ray)) hits++;
});
const ptUs = timeCalls(calls, (i) => {
  const k = i % nPts;
  if (bvh.closestPointToPoint(v(points[3 * k], points[3 * k + 1], points[3 * k + 2]))) found++;
});
// the same calls batched: what one launch of `calls` queries costs per query
const o = new Float32Array(3 * calls), d = new Float32Array(3 * calls), q = new Float32Array(3 * calls);
for (let i = 0; i < calls; i++) {
  const k = i % nRays, j = i % nPts;
  for (let c = 0; c < 3; c++) {
    o[3 * i + c] = rays[3 * k + c];
    d[3 * i + c] = rays[3 * (nRays + k) + c];
    q[3 * i + c] = points[3 * j + c];
  }
}
const batchRayUs = timeCalls(1, () => bvh.raycastFirstBatch(o, d)) / calls;
const batchPtUs = timeCalls(1, () => bvh.closestPointToPointBatch(q)) / calls;

// the same per-call calls through the micro-batched Promise forms: `calls` calls in one turn of the
// event loop, awaited together (one launch), median of 5 rounds, microseconds per call
async function timeQueued(make) {
  await Promise.all(Array.from({ length: 100 }, (_, i) => make(i)));  // warm-up
  const per = [];
  for (let rep = 0; rep < 5; rep++) {
    const a = process.hrtime.bigint();
    await Promise.all(Array.from({ length: calls }, (_, i) => make(i)));
    per.push(Number(process.hrtime.bigint() - a) / 1e3 / calls);
  }
  per.sort((x, y) => x - y);
  return per[2];
}
(async () => {
  const queuedRayUs = await timeQueued((i) => {
    const k = i % nRays;
    return bvh.raycastFirstAsync({ origin: v(rays[3 * k], rays[3 * k + 1], rays[3 * k + 2]),
      direction: v(rays[3 * (nRays + k)], rays[3 * (nRays + k) + 1], rays[3 * (nRays + k) + 2]) });
  });
  const queuedPtUs = await timeQueued((i) => {
    const k = i % nPts;
    return bvh.closestPointToPointAsync(v(points[3 * k], points[3 * k + 1], points[3 * k + 2]));
  });
  console.log(JSON.stringify({ node: process.version, calls, build_and_upload_ms: buildMs,
    raycast_first_us_per_call: rayUs, closest_point_us_per_call: ptUs,
    raycast_first_batch_us_per_query: batchRayUs, closest_point_batch_us_per_query: batchPtUs,
    raycast_first_queued_us_per_call: queuedRayUs, closest_point_queued_us_per_call: queuedPtUs,
    hit_fraction: hits / (100 + 5 * calls), found_fraction: found / (100 + 5 * calls) }));
})().catch((e) => {
  console.error(e);
  process.exit(1);
});
