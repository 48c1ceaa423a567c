'use strict';
// Times the JavaScript restatement of the reference query path (reference_cpu.js) on this host:
// one thread, then a worker_threads pool sharing the mesh and queries through SharedArrayBuffer,
// queries split into equal contiguous chunks. Protocol (as the reference's benchmark/lib/bench.js):
// prewarm runs, then the median of timed runs, wall clock from performance.now().
// TEST INFRASTRUCTURE ONLY: run by bench.py's CPU-baseline leg.
//
//   node cpu_bench.js <dir> <threads> [single_queries]
// <dir>/meta.json: {"kind": "closest" | "raycast", "n": queries, "roots": k, "index_bytes": 2 | 4,
//                   "min_threshold", "max_threshold", "side", "near", "far"} ("inf" for infinity)
// <dir>/positions.f32, index.u16 | index.u32, root<i>.bin, queries.f32 (points, or origins then
// directions). Writes <dir>/faces_single.u32 (first single_queries results) and prints one JSON line.
const fs = require('fs');
const os = require('os');
const path = require('path');
const { performance } = require('perf_hooks');
const { Worker, isMainThread, parentPort, workerData } = require('worker_threads');
const R = require('./reference_cpu.js');

function shared(buf) {
  const sab = new SharedArrayBuffer(buf.byteLength);
  new Uint8Array(sab).set(new Uint8Array(buf.buffer, buf.byteOffset, buf.byteLength));
  return sab;
}

function runRange(data, lo, hi, out) {
  const meta = data.meta;
  const pos = new Float32Array(data.positions);
  const index = meta.index_bytes === 2 ? new Uint16Array(data.index) : new Uint32Array(data.index);
  const mesh = R.makeMesh(pos, index, data.roots);
  const q = new Float32Array(data.queries);
  const n = meta.n;
  if (meta.kind === 'closest') {
    for (let i = lo; i < hi; i++) {
      const h = R.closestPoint(mesh, q[3 * i], q[3 * i + 1], q[3 * i + 2], meta.min_threshold, meta.max_threshold);
      if (out) out[i - lo] = h ? h.face : 0xFFFFFFFF;
    }
  } else {
    for (let i = lo; i < hi; i++) {
      const h = R.raycastFirst(mesh, q[3 * i], q[3 * i + 1], q[3 * i + 2], q[3 * (n + i)], q[3 * (n + i) + 1],
        q[3 * (n + i) + 2], meta.side, meta.near, meta.far);
      if (out) out[i - lo] = h ? h.face : 0xFFFFFFFF;
    }
  }
}

if (!isMainThread) {
  parentPort.on('message', msg => {
    runRange(workerData, msg.lo, msg.hi, null);
    parentPort.postMessage('done');
  });
} else {
  const dir = process.argv[2];
  const threads = Math.max(1, parseInt(process.argv[3] || '1', 10));
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  for (const k of Object.keys(meta)) if (meta[k] === 'inf') meta[k] = Infinity;  // JSON has no infinity
  const rd = f => fs.readFileSync(path.join(dir, f));
  const data = {
    meta,
    positions: shared(rd('positions.f32')),
    index: shared(rd(meta.index_bytes === 2 ? 'index.u16' : 'index.u32')),
    roots: Array.from({ length: meta.roots }, (_, i) => shared(rd(`root${i}.bin`))),
    queries: shared(rd('queries.f32')),
  };
  const median = a => a.slice().sort((x, y) => x - y)[a.length >> 1];
  const single = Math.min(meta.n, parseInt(process.argv[4] || String(meta.n), 10));
  // one thread over the first `single` queries
  const faces = new Uint32Array(single);
  runRange(data, 0, Math.min(single, 256), faces);  // prewarm (JIT)
  const st = [];
  for (let k = 0; k < 3; k++) {
    const t0 = performance.now();
    runRange(data, 0, single, faces);
    st.push((performance.now() - t0) / 1000);
  }
  fs.writeFileSync(path.join(dir, 'faces_single.u32'), Buffer.from(faces.buffer));
  // worker pool over all n queries
  const workers = Array.from({ length: threads }, () => new Worker(__filename, { workerData: data }));
  const chunk = Math.ceil(meta.n / threads);
  const round = () => new Promise(resolve => {
    let left = threads;
    workers.forEach((w, k) => {
      w.once('message', () => { if (--left === 0) resolve(); });
      w.postMessage({ lo: Math.min(meta.n, k * chunk), hi: Math.min(meta.n, (k + 1) * chunk) });
    });
  });
  (async () => {
    await round();  // prewarm
    const pt = [];
    for (let k = 0; k < 3; k++) {
      const t0 = performance.now();
      await round();
      pt.push((performance.now() - t0) / 1000);
    }
    workers.forEach(w => w.terminate());
    const s1 = median(st), sp = median(pt);
    console.log(JSON.stringify({
      node: process.version, cpu: (os.cpus()[0] || {}).model || 'unknown', host_cpus: os.cpus().length,
      single: { queries: single, seconds: s1, queries_per_s: single / s1 },
      pool: { threads, queries: meta.n, seconds: sp, queries_per_s: meta.n / sp },
    }));
  })();
}
