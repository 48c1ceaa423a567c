'use strict';
// Same values as the reference (src/core/Constants.js) and three.js's side constants.

const CENTER = 0;
const AVERAGE = 1;
const SAH = 2;

const NOT_INTERSECTED = 0;
const INTERSECTED = 1;
const CONTAINED = 2;

const FrontSide = 0;
const BackSide = 1;
const DoubleSide = 2;

const BYTES_PER_NODE = 32;
const UINT32_PER_NODE = 8;
const IS_LEAFNODE_FLAG = 0xFFFF;

const SKIP_GENERATION = Symbol('SKIP_GENERATION');

const DEFAULT_OPTIONS = {
  strategy: CENTER,
  maxDepth: 40,
  targetLeafSize: 10,
  useSharedArrayBuffer: false,
  setBoundingBox: true,
  onProgress: null,
  indirect: false,
  verbose: true,
  range: null,
  gpuBuild: false,  // build on the device (this package's option; the tree is the same)
  [SKIP_GENERATION]: false,
};

// Face index the native layer reports for a miss.
const NO_HIT = 0xFFFFFFFF;

module.exports = {
  CENTER, AVERAGE, SAH, NOT_INTERSECTED, INTERSECTED, CONTAINED, FrontSide, BackSide, DoubleSide,
  BYTES_PER_NODE, UINT32_PER_NODE, IS_LEAFNODE_FLAG, SKIP_GENERATION, DEFAULT_OPTIONS, NO_HIT,
};
