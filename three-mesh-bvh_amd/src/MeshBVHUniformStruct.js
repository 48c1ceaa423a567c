'use strict';
// The reference's shader data for a BVH (src/webgl/MeshBVHUniformStruct.js, src/webgl/
// VertexAttributeTexture.js), produced on the device from the uploaded mesh: bvhBounds (RGBA32F, two
// texels per node), bvhContents (RG32UI, one texel per node), index (RGBA unsigned texels of the
// permuted / dereferenced index, the index array's element type) and position (RGBA32F). three.js is
// not a dependency of this package, so each texture is a plain object with the DataTexture fields a
// renderer reads (image.data / width / height, format, type, internalFormat); copy them into
// THREE.DataTexture objects, or hand `image.data` to the GPU directly.
const native = require('./native.js');
const GU = require('./geometryUtils.js');

function texture(data, side, format, type, internalFormat) {
  return { image: { data, width: side, height: side }, format, type, internalFormat, minFilter: 'NearestFilter',
    magFilter: 'NearestFilter', generateMipmaps: false };
}

class MeshBVHUniformStruct {
  constructor() {
    this.index = null;
    this.position = null;
    this.bvhBounds = null;
    this.bvhContents = null;
  }

  // updateFrom (src/webgl/MeshBVHUniformStruct.js:55-93)
  updateFrom(bvh) {
    if (bvh._roots.length !== 1) throw new Error('MeshBVHUniformStruct: Multi-root BVHs not supported.');
    const addon = native.load();
    const mesh = bvh._mesh();
    const geometry = bvh.geometry;
    const [sb, sc, si, sp] = addon.textureSizes(mesh);
    const pos = GU.positionLayout(geometry);
    const bounds = new Float32Array(4 * sb * sb), contents = new Uint32Array(2 * sc * sc);
    const index32 = new Uint32Array(4 * si * si), position = new Float32Array(4 * sp * sp);
    addon.exportTextures(mesh, pos.array, pos.stride, pos.offset, bounds, contents, index32, position);
    this.bvhBounds = texture(bounds, sb, 'RGBAFormat', 'FloatType', 'RGBA32F');
    this.bvhContents = texture(contents, sc, 'RGIntegerFormat', 'UnsignedIntType', 'RG32UI');
    this.position = texture(position, sp, 'RGBAFormat', 'FloatType', 'RGBA32F');

    // the index attribute the reference packs: the geometry index, or for an indirect BVH a copy of it
    // (a zeroed index array without one) whose first triangles are dereferenced through the indirect buffer
    const src = geometry.index ? geometry.index.array : null;
    const vertexCount = geometry.index ? geometry.index.count : geometry.attributes.position.count;
    const Arr = src ? src.constructor : (vertexCount > 65535 ? Uint32Array : Uint16Array);
    const tris = vertexCount / 3;
    const side = Math.ceil(Math.sqrt(tris)) || 1;
    const data = new Arr(4 * side * side);
    const done = bvh._indirectBuffer ? bvh._indirectBuffer.length : tris;
    data.set(index32.subarray(0, 4 * done));
    for (let t = done; t < tris; t++) {
      for (let c = 0; c < 3; c++) data[4 * t + c] = src ? src[3 * t + c] : 0;
      data[4 * t + 3] = 1;
    }
    const bits = Arr.BYTES_PER_ELEMENT * 8;
    this.index = texture(data, side, 'RGBAIntegerFormat', bits === 16 ? 'UnsignedShortType' : 'UnsignedIntType',
      `RGBA${bits}UI`);
  }

  dispose() {
    this.index = this.position = this.bvhBounds = this.bvhContents = null;
  }
}

// The WebGPU BVHNode storage buffer of one BVH (bvhNodeStruct src/webgpu/tsl/structs.js:3-14, written as
// appendBVHSubtree src/webgpu/utils/packBVHBufferUtils.js does for a mesh): 8 words per node, roots one
// after another, an empty root's bounds as [1, -1].
function bvhNodeBuffer(bvh) {
  const words = bvh._roots.reduce((n, r) => n + r.byteLength, 0) / 4;
  const out = new Uint32Array(words);
  native.load().exportNodes(bvh._mesh(), out);
  return out;
}

module.exports = { MeshBVHUniformStruct, bvhNodeBuffer };
