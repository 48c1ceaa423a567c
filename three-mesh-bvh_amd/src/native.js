'use strict';
// Loads the Node-API addon (lib/mbvh.node → lib/libmbvh.so). There is no CPU fallback: without the
// addon or a HIP device, the query methods throw.
const path = require('path');

let addon = null;
let device = process.env.MBVH_DEVICE !== undefined ? parseInt(process.env.MBVH_DEVICE, 10) : 0;

function load() {
  if (addon) return addon;
  const file = path.join(__dirname, '..', 'lib', 'mbvh.node');
  try {
    addon = require(file);
  } catch (e) {
    throw new Error(`three-mesh-bvh-amd: cannot load the native addon ${file} (${e.message}). ` +
      'Build it with `make -C three-mesh-bvh_amd/csrc`.');
  }
  return addon;
}

function getDevice() {
  return device;
}

// Selects the HIP device used by meshes uploaded afterwards.
function setDevice(d) {
  if (!Number.isInteger(d) || d < 0) throw new Error('setDevice: device must be a non-negative integer');
  device = d;
}

module.exports = { load, getDevice, setDevice };
