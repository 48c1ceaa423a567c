'use strict';
// Times the JavaScript restatement of the reference's refit (reference_cpu.js) on one thread - the
// reference's refit is a sequential recursion - median of 3 after a prewarm, each run on a fresh
// copy of the roots. TEST INFRASTRUCTURE ONLY: run by bench.py's CPU-baseline leg.
//   node refit_bench.js <dir>
// <dir>: meta.json {"roots": k, "index_bytes": 2 | 4}, positions.f32 (the new positions),
// index.u16 | index.u32, root<i>.bin (the roots before the refit). Writes refit<i>.bin.
const fs = require('fs');
const path = require('path');
const { performance } = require('perf_hooks');
const R = require('./reference_cpu.js');

const dir = process.argv[2];
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
const rd = f => { const b = fs.readFileSync(path.join(dir, f)); return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); };
const pos = new Float32Array(rd('positions.f32'));
const index = meta.index_bytes === 2 ? new Uint16Array(rd('index.u16')) : new Uint32Array(rd('index.u32'));
const orig = Array.from({ length: meta.roots }, (_, i) => rd(`root${i}.bin`));
const times = [];
let roots = null;
for (let k = 0; k < 4; k++) {
  roots = orig.map(b => b.slice(0));
  const mesh = R.makeMesh(pos, index, roots);
  const t0 = performance.now();
  R.refit(mesh);
  if (k > 0) times.push((performance.now() - t0) / 1000);
}
roots.forEach((b, i) => fs.writeFileSync(path.join(dir, `refit${i}.bin`), Buffer.from(b)));
times.sort((a, b) => a - b);
console.log(JSON.stringify({ node: process.version, seconds: times[1] }));
