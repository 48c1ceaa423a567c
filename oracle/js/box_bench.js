'use strict';
// Times the JavaScript restatement of the reference's intersectsBox (reference_cpu.js) over a set of
// oriented boxes on one thread, median of 3 after a prewarm, and writes the answers. TEST
// INFRASTRUCTURE ONLY: run by bench.py's CPU-baseline leg and by tests/.
//   node box_bench.js <dir>
// <dir>: meta.json {"roots": k, "index_bytes": 2 | 4, "boxes": n}, positions.f32, index.u16 | index.u32,
// root<i>.bin, box_min.f64 (3n), box_max.f64 (3n), box_matrix.f64 (16n). Writes box_hit.u8.
const fs = require('fs');
const path = require('path');
const { performance } = require('perf_hooks');
const R = require('./reference_cpu.js');

const dir = process.argv[2];
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
const rd = f => { const b = fs.readFileSync(path.join(dir, f)); return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); };
const pos = new Float32Array(rd('positions.f32'));
const index = meta.index_bytes === 2 ? new Uint16Array(rd('index.u16')) : new Uint32Array(rd('index.u32'));
const roots = Array.from({ length: meta.roots }, (_, i) => rd(`root${i}.bin`));
const mn = new Float64Array(rd('box_min.f64')), mx = new Float64Array(rd('box_max.f64'));
const mt = new Float64Array(rd('box_matrix.f64'));
const mesh = R.makeMesh(pos, index, roots);
const n = meta.boxes;
const out = new Uint8Array(n);
const times = [];
for (let k = 0; k < 4; k++) {
  const t0 = performance.now();
  for (let i = 0; i < n; i++) {
    out[i] = R.intersectsBox(mesh, mn.subarray(3 * i, 3 * i + 3), mx.subarray(3 * i, 3 * i + 3),
      mt.subarray(16 * i, 16 * i + 16)) ? 1 : 0;
  }
  if (k > 0) times.push((performance.now() - t0) / 1000);
}
fs.writeFileSync(path.join(dir, 'box_hit.u8'), Buffer.from(out.buffer));
times.sort((a, b) => a - b);
console.log(JSON.stringify({ node: process.version, seconds: times[1] }));
