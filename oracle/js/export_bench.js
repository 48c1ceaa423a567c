'use strict';
// Times a JavaScript restatement of the reference's shader-data packing for one BVH on one thread:
// bvhToTextures (src/webgl/MeshBVHUniformStruct.js:129-210) and VertexAttributeTexture.updateFrom
// (src/webgl/VertexAttributeTexture.js:106-300) for the index (item size 3) and position attributes;
// median of 3 after a prewarm. Writes the four arrays for comparison. TEST INFRASTRUCTURE ONLY: run
// by bench.py's CPU-baseline leg.
//   node export_bench.js <dir>
// <dir>: meta.json {"index_bytes": 2 | 4}, root0.bin, index.u16 | index.u32 (permuted), positions.f32.
const fs = require('fs');
const path = require('path');
const { performance } = require('perf_hooks');

const dir = process.argv[2];
const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
const rd = f => { const b = fs.readFileSync(path.join(dir, f)); return b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); };
const root = rd('root0.bin');
const index = meta.index_bytes === 2 ? new Uint16Array(rd('index.u16')) : new Uint32Array(rd('index.u32'));
const positions = new Float32Array(rd('positions.f32'));

function bvhToTextures(root) {
  const u16 = new Uint16Array(root), u32 = new Uint32Array(root), f32 = new Float32Array(root);
  const nodeCount = root.byteLength / 32;
  const bSide = 2 * Math.ceil(Math.sqrt(nodeCount / 2));
  const bounds = new Float32Array(4 * bSide * bSide);
  const cSide = Math.ceil(Math.sqrt(nodeCount));
  const contents = new Uint32Array(2 * cSide * cSide);
  for (let i = 0; i < nodeCount; i++) {
    const n32 = 8 * i, n16 = 2 * n32;
    for (let b = 0; b < 3; b++) {
      bounds[8 * i + b] = f32[n32 + b];
      bounds[8 * i + 4 + b] = f32[n32 + 3 + b];
    }
    if (u16[n16 + 15] === 0xFFFF) {
      contents[2 * i] = (0xFFFF0000 | u16[n16 + 14]) >>> 0;
      contents[2 * i + 1] = u32[n32 + 6];
    } else {
      contents[2 * i] = u32[n32 + 7];
      contents[2 * i + 1] = u32[n32 + 6];
    }
  }
  return { bounds, contents };
}

function attributeTexture(array, itemSize, Out) {
  const count = array.length / itemSize;
  const side = Math.ceil(Math.sqrt(count)) || 1;
  const stride = itemSize === 3 ? 4 : itemSize;
  const data = new Out(stride * side * side);
  for (let i = 0; i < count; i++) {
    const ii = stride * i;
    for (let c = 0; c < itemSize; c++) data[ii + c] = array[itemSize * i + c];
    if (itemSize === 3) data[ii + 3] = 1;
  }
  return data;
}

const times = [];
let out = null;
for (let k = 0; k < 4; k++) {
  const t0 = performance.now();
  const t = bvhToTextures(root);
  const idx = attributeTexture(index, 3, index.constructor);
  const pos = attributeTexture(positions, 3, Float32Array);
  if (k > 0) times.push((performance.now() - t0) / 1000);
  out = { ...t, idx, pos };
}
fs.writeFileSync(path.join(dir, 'bounds.f32'), Buffer.from(out.bounds.buffer));
fs.writeFileSync(path.join(dir, 'contents.u32'), Buffer.from(out.contents.buffer));
fs.writeFileSync(path.join(dir, 'index_tex.bin'), Buffer.from(out.idx.buffer));
fs.writeFileSync(path.join(dir, 'position_tex.f32'), Buffer.from(out.pos.buffer));
times.sort((a, b) => a - b);
console.log(JSON.stringify({ node: process.version, seconds: times[1] }));
