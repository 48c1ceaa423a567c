// Host-side internals shared by the builder (g++) and the device runtime (hipcc).
#pragma once

#include <cstdint>
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/mbvh.h"

namespace mbvh {

// Reference node format (src/core/Constants.js:61-66): 32 bytes per node, leaf flag in the top
// uint16 of the last word.
constexpr uint32_t kBytesPerNode = 32;
constexpr uint32_t kU32PerNode = kBytesPerNode / 4;
constexpr uint16_t kLeafFlag = 0xFFFF;

constexpr uint32_t kTreeMagic = 0x45455254u;   // "TREE"
constexpr uint32_t kCtxMagic = 0x58544343u;    // "CCTX"
constexpr uint32_t kMeshMagic = 0x4853454Du;   // "MESH"

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const char* fmt, ...);
void clear_error();

// Argument checks shared by the host and the device builds (messages prefixed with `fn`); the
// index values are checked only when `check_index_values` (the device build checks them itself).
int check_build_args(const char* fn, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                     uint32_t position_offset, const void* index, int32_t index_elt_bytes, uint64_t index_count,
                     const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                     bool check_index_values);

// Read-only view of one reference-format node buffer.
struct NodeView {
  const uint8_t* bytes;
  uint64_t n_nodes;

  const float* bounds(uint64_t node) const {
    return reinterpret_cast<const float*>(bytes + node * kBytesPerNode);
  }
  uint32_t word(uint64_t node, int w) const {
    uint32_t v;
    __builtin_memcpy(&v, bytes + node * kBytesPerNode + 4 * w, 4);
    return v;
  }
  uint16_t half(uint64_t node, int h) const {
    uint16_t v;
    __builtin_memcpy(&v, bytes + node * kBytesPerNode + 2 * h, 2);
    return v;
  }
  bool is_leaf(uint64_t node) const { return half(node, 15) == kLeafFlag; }
  uint32_t leaf_offset(uint64_t node) const { return word(node, 6); }
  uint32_t leaf_count(uint64_t node) const { return half(node, 14); }
  uint64_t right_child(uint64_t node) const { return node + word(node, 6); }
  uint32_t split_axis(uint64_t node) const { return word(node, 7); }
};

}  // namespace mbvh

struct mbvh_tree {
  uint32_t magic = mbvh::kTreeMagic;
  std::vector<std::vector<uint8_t>> roots;  // reference-format node buffers, one per root
  int32_t indirect_elt_bytes = 0;           // 0, 2 or 4
  std::vector<uint8_t> indirect;            // indirect buffer bytes (indirect builds only)
  uint64_t indirect_count = 0;
};
