'use strict';
// Geometry bookkeeping the build needs, restated from the reference:
//   ensureIndex            src/core/build/geometryUtils.js:30-47
//   getRootPrimitiveRanges src/core/build/geometryUtils.js:60-159
// Geometries are duck-typed three.js BufferGeometry objects: `index` (a BufferAttribute or null),
// `attributes.position`, `groups`, `drawRange`.

function getVertexCount(geo) {
  return geo.index ? geo.index.count : geo.attributes.position.count;
}

// A BufferAttribute-shaped wrapper; when the geometry came from three.js, three's own
// BufferAttribute class is used so renderers keep working.
function makeIndexAttribute(geo, array) {
  const pos = geo.attributes.position;
  let Ctor = pos && pos.constructor;
  while (Ctor && Ctor !== Object) {
    if (Ctor.name === 'BufferAttribute') return new Ctor(array, 1, false);
    Ctor = Object.getPrototypeOf(Ctor);
  }
  return {
    array,
    itemSize: 1,
    count: array.length,
    normalized: false,
    isBufferAttribute: true,
    getX(i) {
      return this.array[i];
    },
  };
}

function ensureIndex(geo, options) {
  if (geo.index) return;
  const vertexCount = geo.attributes.position.count;
  const Buffer = options.useSharedArrayBuffer ? SharedArrayBuffer : ArrayBuffer;
  const index = vertexCount > 65535 ? new Uint32Array(new Buffer(4 * vertexCount)) : new Uint16Array(new Buffer(2 * vertexCount));
  for (let i = 0; i < vertexCount; i++) index[i] = i;
  const attr = makeIndexAttribute(geo, index);
  if (typeof geo.setIndex === 'function') geo.setIndex(attr);
  else geo.index = attr;
}

function getFullPrimitiveRange(geo, range, stride) {
  const primitiveCount = getVertexCount(geo) / stride;
  const drawRange = range ? range : (geo.drawRange || { start: 0, count: Infinity });
  const start = drawRange.start / stride;
  const end = (drawRange.start + drawRange.count) / stride;
  const offset = Math.max(0, start);
  const count = Math.min(primitiveCount, end) - offset;
  return { offset: Math.floor(offset), count: Math.floor(count) };
}

// Mutually exclusive (offset, count) triangle ranges covered by groups within the draw range; each
// becomes its own root so no triangle is ever moved across a group boundary.
function getRootPrimitiveRanges(geo, range, stride = 3) {
  const drawRange = getFullPrimitiveRange(geo, range, stride);
  const groups = (geo.groups || []).map(g => ({ offset: g.start / stride, count: g.count / stride }));
  if (!groups.length) return [drawRange];

  const ranges = [];
  const drawStart = drawRange.offset;
  const drawEnd = drawRange.offset + drawRange.count;
  const primitiveCount = getVertexCount(geo) / stride;
  const events = [];
  for (const group of groups) {
    const groupStart = group.offset;
    const groupCount = isFinite(group.count) ? group.count : primitiveCount - group.offset;
    const groupEnd = group.offset + groupCount;
    if (groupStart < drawEnd && groupEnd > drawStart) {
      events.push({ pos: Math.max(drawStart, groupStart), isStart: true });
      events.push({ pos: Math.min(drawEnd, groupEnd), isStart: false });
    }
  }
  // the reference's comparator orders equal positions with its `type === 'end'` test, which is
  // never true, so equal-position events keep no particular order; counting makes that harmless
  events.sort((a, b) => (a.pos !== b.pos ? a.pos - b.pos : (a.type === 'end' ? -1 : 1)));
  let active = 0;
  let lastPos = null;
  for (const event of events) {
    if (active !== 0 && event.pos !== lastPos) ranges.push({ offset: lastPos, count: event.pos - lastPos });
    active += event.isStart ? 1 : -1;
    lastPos = event.pos;
  }
  return ranges;
}

// Position data as (Float32Array, stride, offset) without copying when possible.
function positionLayout(geo) {
  const attr = geo.attributes.position;
  if (attr.isInterleavedBufferAttribute) {
    const arr = attr.data.array;
    if (arr instanceof Float32Array && !attr.normalized) return { array: arr, stride: attr.data.stride, offset: attr.offset || 0 };
  } else if (attr.array instanceof Float32Array && !attr.normalized && (attr.itemSize || 3) >= 3) {
    return { array: attr.array, stride: attr.itemSize || 3, offset: 0 };
  }
  // normalized or non-float data: unpack through the accessors, as the reference does
  const n = attr.count;
  const out = new Float32Array(3 * n);
  for (let i = 0; i < n; i++) {
    out[3 * i] = attr.getX(i);
    out[3 * i + 1] = attr.getY(i);
    out[3 * i + 2] = attr.getZ(i);
  }
  return { array: out, stride: 3, offset: 0 };
}

module.exports = { getVertexCount, ensureIndex, getRootPrimitiveRanges, positionLayout, makeIndexAttribute };
