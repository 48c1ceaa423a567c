// Refit: every node box of an uploaded tree recomputed from new vertex positions, on the device.
//
// Semantics of the reference's refit with nodeIndices = null (src/core/cast/refit.template.js:4-196,
// src/core/MeshBVH.js:504-508): a leaf's box becomes the min / max over its triangles' vertices,
// visited in slot order with strict < and > (so the first of equal values, e.g. -0 before +0, is
// kept), and is written, and marks the leaf changed, only if some component differs (!==) from the
// stored one; an internal node is recomputed as the union of its children (min: left < right ?
// left : right, max: left > right ? left : right) exactly when a child changed, and then counts as
// changed itself.
//
// Device form: one thread per leaf walks up through parent links. Each node has an arrival counter
// (cleared by k_refit_clear, which also resets the coordinate bound that k_refit_bound recomputes);
// the second child to arrive computes the parent (release fence before
// the counter increment, acquire fence after it), so every internal node is decided once, after
// both children. A node's box is written to the device copy of the reference nodes and, where the
// node has one, to its place in the parent's wide node; k_refit_four then copies the changed wide
// boxes into the 4-wide tree, and k_refit_tris re-gathers the triangle corners (their .w words -
// face id, facing mask, vertex id - are kept).
#include "mbvh_kernels.h"

namespace mbvh {
namespace {

constexpr int kRefitBlock = 256;

__global__ __launch_bounds__(kRefitBlock) void k_refit_clear(uint32_t* counters, uint32_t* changed, uint64_t n,
                                                              float* coord_bound) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  if (blockIdx.x == 0 && threadIdx.x == 0) *coord_bound = 0.0f;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < n; i += stride) {
    counters[i] = 0u;
    changed[i] = 0u;
  }
}

// The mesh's coordinate bound: max |coordinate| over the vertices (NaN ignored; an infinite one
// makes the bound infinite), through one atomic max per wave on the float bits (monotonic for
// non-negative floats).
__global__ __launch_bounds__(kRefitBlock) void k_refit_bound(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  float m = 0.0f;
  for (uint64_t v = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; v < a.n_vertices; v += stride) {
    const float* p = a.positions + v * a.stride + a.offset;
    for (int k = 0; k < 3; k++) {
      const float x = fabsf(p[k]);
      if (x > m) m = x;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_xor(m, off, 64);
    if (o > m) m = o;
  }
  if ((threadIdx.x & 63u) == 0u) atomicMax(reinterpret_cast<unsigned int*>(a.coord_bound), __float_as_uint(m));
}

__global__ __launch_bounds__(kRefitBlock) void k_refit_tris(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < 3 * a.n_slots; i += stride) {
    const float* p = a.positions + static_cast<uint64_t>(a.vidx[i]) * a.stride + a.offset;
    float4 r = a.tris[i];
    r.x = p[0];
    r.y = p[1];
    r.z = p[2];
    a.tris[i] = r;
  }
}

// Box of node i into the reference copy (floats 0..5 of the node) and its wide-node place.
__device__ __forceinline__ void put_box(const RefitArgs& a, uint64_t i, const float* b) {
  float* dst = reinterpret_cast<float*>(a.ref_nodes + 8 * i);
  for (int k = 0; k < 6; k++) dst[k] = b[k];
  const int32_t w = a.wide_off[i];
  if (w >= 0) {
    float* wd = a.wide + w;
    for (int k = 0; k < 6; k++) wd[k] = b[k];
  }
}

__global__ __launch_bounds__(kRefitBlock) void k_refit_up(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t li = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; li < a.n_leaves; li += stride) {
    uint64_t node = a.leaves[li];
    const uint32_t* w = a.ref_nodes + 8 * node;
    const uint32_t off = w[6];
    const uint32_t cnt = w[7] & 0xFFFFu;  // u16[14]: count (u16[15] is the leaf flag)
    const float inf = __builtin_huge_valf();
    float b[6] = {inf, inf, inf, -inf, -inf, -inf};
    for (uint64_t c = 3ull * off; c < 3ull * (static_cast<uint64_t>(off) + cnt); c++) {
      const float* p = a.positions + static_cast<uint64_t>(a.vidx[c]) * a.stride + a.offset;
      for (int k = 0; k < 3; k++) {
        if (p[k] < b[k]) b[k] = p[k];
        if (p[k] > b[3 + k]) b[3 + k] = p[k];
      }
    }
    const float* old = reinterpret_cast<const float*>(w);
    bool changed = false;
    for (int k = 0; k < 6; k++) changed = changed || !(old[k] == b[k]);
    if (changed) put_box(a, node, b);
    a.changed[node] = changed ? 1u : 0u;
    // walk up: the second child to arrive at a parent decides it
    while (true) {
      const int32_t parent = a.parent[node];
      if (parent < 0) break;
      __threadfence();  // this subtree's boxes and flags before the arrival
      if (atomicAdd(&a.counters[parent], 1u) == 0u) break;  // the sibling is not done yet
      __threadfence();  // the sibling's boxes and flags after its arrival
      const uint64_t left = static_cast<uint64_t>(parent) + 1;
      const uint64_t right = static_cast<uint64_t>(parent) + a.ref_nodes[8 * parent + 6];
      const bool any = a.changed[left] != 0u || a.changed[right] != 0u;
      if (any) {
        const float* lb = reinterpret_cast<const float*>(a.ref_nodes + 8 * left);
        const float* rb = reinterpret_cast<const float*>(a.ref_nodes + 8 * right);
        float u[6];
        for (int k = 0; k < 3; k++) {
          u[k] = lb[k] < rb[k] ? lb[k] : rb[k];
          u[3 + k] = lb[3 + k] > rb[3 + k] ? lb[3 + k] : rb[3 + k];
        }
        put_box(a, static_cast<uint64_t>(parent), u);
      }
      a.changed[parent] = any ? 1u : 0u;
      node = static_cast<uint64_t>(parent);
    }
  }
}

// 4-wide child slots: six floats each, copied from the wide box they were made from.
__global__ __launch_bounds__(kRefitBlock) void k_refit_four(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < a.n_four; i += stride) {
    const uint2 m = a.four_map[i];  // (float offset in the 4-wide tree, float offset in the wide nodes)
    for (int k = 0; k < 6; k++) a.four[m.x + k] = a.wide[m.y + k];
  }
}

uint32_t blocks_for(uint64_t n) {
  const uint64_t b = (n + kRefitBlock - 1) / kRefitBlock;
  return static_cast<uint32_t>(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

hipError_t launch_refit(const RefitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_refit_clear, dim3(blocks_for(a.n_nodes)), dim3(kRefitBlock), 0, s, a.counters, a.changed,
                     a.n_nodes, a.coord_bound);
  hipLaunchKernelGGL(k_refit_bound, dim3(blocks_for(a.n_vertices) < 1024 ? blocks_for(a.n_vertices) : 1024),
                     dim3(kRefitBlock), 0, s, a);
  if (a.n_slots) hipLaunchKernelGGL(k_refit_tris, dim3(blocks_for(3 * a.n_slots)), dim3(kRefitBlock), 0, s, a);
  if (a.n_leaves) hipLaunchKernelGGL(k_refit_up, dim3(blocks_for(a.n_leaves)), dim3(kRefitBlock), 0, s, a);
  if (a.n_four) hipLaunchKernelGGL(k_refit_four, dim3(blocks_for(a.n_four)), dim3(kRefitBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace mbvh
