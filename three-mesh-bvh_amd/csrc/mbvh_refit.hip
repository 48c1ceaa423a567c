// Refit: every node box of an uploaded tree recomputed from new vertex positions, on the device; and
// the exports of the tree to the reference's shader data layouts.
//
// Semantics of the reference's refit with nodeIndices = null (src/core/cast/refit.template.js:4-196,
// src/core/MeshBVH.js:504-508): a leaf's box becomes the min / max over its triangles' vertices,
// visited in slot order with strict < and > (so the first of equal values, e.g. -0 before +0, is
// kept), and is written, and marks the leaf changed, only if some component differs (!==) from the
// stored one; an internal node is recomputed as the union of its children (min: left < right ?
// left : right, max: left > right ? left : right) exactly when a child changed, and then counts as
// changed itself.
//
// Device form, level by level: one launch decides every leaf, then one launch per depth, deepest
// first, decides the internal nodes of that depth from their children (decided by an earlier
// launch on the same stream, so no fences or counters are needed). A node's box is written to the
// device copy of the reference nodes and, where the node has one, to its place in the parent's wide
// node; k_refit_four then copies the wide boxes into the 4-wide tree, and k_refit_tris re-gathers
// the triangle corners (their .w words - face id, facing mask, vertex id - are kept). The
// coordinate bound the query kernels scale their error bounds by is recomputed from the positions.
//
// A partial refit (nodeIndices, refit.template.js:136-172) first marks the nodes the reference's
// traversal visits, top down, one launch per depth: every root is visited; a visited internal node
// that is not forced visits the children in the set, or, when neither is, both, forced; a forced
// node forces its children. The per-node word holds the marks until the bottom-up launches replace
// them by the changed flag; a node not visited keeps its box and counts as unchanged. The triangle
// corners and the coordinate bound are refreshed everywhere, as the queries read the new positions.
#include "mbvh_kernels.h"

namespace mbvh {
namespace {

constexpr int kRefitBlock = 256;

// Per-node marks of a partial refit (the changed flag, written bottom up, is bit 0).
constexpr uint32_t kRefitChanged = 1u, kRefitVisit = 2u, kRefitMember = 4u, kRefitForce = 8u;

__global__ __launch_bounds__(64) void k_refit_reset_bound(float* coord_bound) {
  if (threadIdx.x == 0) *coord_bound = 0.0f;
}

// The mesh's coordinate bound: max |coordinate| over the vertices (NaN ignored; an infinite one
// makes the bound infinite): a wave and then a workgroup reduction, and one atomic max per
// workgroup on the float bits (monotonic for non-negative floats; atomics on one address serialise,
// so the grid is kept small).
constexpr uint32_t kBoundBlocks = 256;

__global__ __launch_bounds__(kRefitBlock) void k_refit_bound(RefitArgs a) {
  __shared__ float wave_max[kRefitBlock / 64];
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
  if ((threadIdx.x & 63u) == 0u) wave_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kRefitBlock / 64; w++)
      if (wave_max[w] > m) m = wave_max[w];
    atomicMax(reinterpret_cast<unsigned int*>(a.coord_bound), __float_as_uint(m));
  }
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

__global__ __launch_bounds__(kRefitBlock) void k_refit_leaves(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  const float inf = __builtin_huge_valf();
  for (uint64_t li = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; li < a.n_leaves; li += stride) {
    const uint64_t node = a.leaves[li];
    if (a.node_ids && !(a.changed[node] & kRefitVisit)) {
      a.changed[node] = 0u;
      continue;
    }
    const uint32_t* w = a.ref_nodes + 8 * node;
    const uint32_t off = w[6];
    const uint32_t cnt = w[7] & 0xFFFFu;  // u16[14]: count (u16[15] is the leaf flag)
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
    a.changed[node] = changed ? kRefitChanged : 0u;
  }
}

// The internal nodes of one depth, whose children were decided by earlier launches.
__global__ __launch_bounds__(kRefitBlock) void k_refit_level(RefitArgs a, const uint32_t* nodes, uint64_t count) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < count; i += stride) {
    const uint64_t node = nodes[i];
    if (a.node_ids && !(a.changed[node] & kRefitVisit)) {
      a.changed[node] = 0u;
      continue;
    }
    const uint64_t left = node + 1;
    const uint64_t right = node + a.ref_nodes[8 * node + 6];
    const bool any = ((a.changed[left] | a.changed[right]) & kRefitChanged) != 0u;
    if (any) {
      const float* lb = reinterpret_cast<const float*>(a.ref_nodes + 8 * left);
      const float* rb = reinterpret_cast<const float*>(a.ref_nodes + 8 * right);
      float u[6];
      for (int k = 0; k < 3; k++) {
        u[k] = lb[k] < rb[k] ? lb[k] : rb[k];
        u[3 + k] = lb[3 + k] > rb[3 + k] ? lb[3 + k] : rb[3 + k];
      }
      put_box(a, node, u);
    }
    a.changed[node] = any ? kRefitChanged : 0u;
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

__global__ __launch_bounds__(kRefitBlock) void k_refit_mark_clear(uint32_t* marks, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < n; i += stride) marks[i] = 0u;
}

// Members of the set (plain stores of one value, so repeated ids are harmless).
__global__ __launch_bounds__(kRefitBlock) void k_refit_mark_ids(RefitArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < a.n_ids; i += stride) {
    const uint32_t id = a.node_ids[i];
    if (id < a.n_nodes) a.changed[id] = kRefitMember;
  }
}

__global__ __launch_bounds__(64) void k_refit_mark_roots(RefitArgs a) {
  for (uint32_t r = threadIdx.x; r < a.n_roots; r += 64) a.changed[a.root_first[r]] |= kRefitVisit;
}

// The children of the visited internal nodes of one depth (each node writes only its own two
// children, whose member bits it reads first).
__global__ __launch_bounds__(kRefitBlock) void k_refit_mark_level(RefitArgs a, const uint32_t* nodes, uint64_t count) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < count; i += stride) {
    const uint64_t node = nodes[i];
    const uint32_t st = a.changed[node];
    if (!(st & kRefitVisit)) continue;
    const uint64_t left = node + 1;
    const uint64_t right = node + a.ref_nodes[8 * node + 6];
    const uint32_t ml = a.changed[left], mr = a.changed[right];
    const bool force = (st & kRefitForce) || !((ml | mr) & kRefitMember);
    if (force) {
      a.changed[left] = ml | kRefitVisit | kRefitForce;
      a.changed[right] = mr | kRefitVisit | kRefitForce;
    } else {
      if (ml & kRefitMember) a.changed[left] = ml | kRefitVisit;
      if (mr & kRefitMember) a.changed[right] = mr | kRefitVisit;
    }
  }
}

uint32_t blocks_for(uint64_t n) {
  const uint64_t b = (n + kRefitBlock - 1) / kRefitBlock;
  return static_cast<uint32_t>(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// MeshBVHUniformStruct's textures (src/webgl/MeshBVHUniformStruct.js:129-210, VertexAttributeTexture
// updateFrom src/webgl/VertexAttributeTexture.js:106-300): one thread per texel of the largest
// texture; texels past the data are zero (the reference's arrays start zeroed).
__global__ __launch_bounds__(kRefitBlock) void k_export_textures(ExportArgs a) {
  uint64_t most = a.bounds_texels > a.contents_texels ? a.bounds_texels : a.contents_texels;
  if (a.index && a.index_texels > most) most = a.index_texels;
  if (a.position && a.position_texels > most) most = a.position_texels;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; t < most; t += stride) {
    if (t < a.bounds_texels) {  // texel 2i: min.xyz, 0; texel 2i + 1: max.xyz, 0
      const uint64_t node = t / 2, half = t % 2;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (node < a.n_nodes) {
        const float* b = reinterpret_cast<const float*>(a.nodes + 8 * node) + 3 * half;
        v = make_float4(b[0], b[1], b[2], 0.f);
      }
      reinterpret_cast<float4*>(a.bounds)[t] = v;
    }
    if (t < a.contents_texels) {  // leaf: (0xFFFF0000 | count, offset); internal: (split axis, right offset)
      uint2 v = make_uint2(0u, 0u);
      if (t < a.n_nodes) {
        const uint32_t* w = a.nodes + 8 * t;
        const bool leaf = (w[7] >> 16) == 0xFFFFu;
        v = leaf ? make_uint2(0xFFFF0000u | (w[7] & 0xFFFFu), w[6]) : make_uint2(w[7], w[6]);
      }
      reinterpret_cast<uint2*>(a.contents)[t] = v;
    }
    if (a.index && t < a.index_texels) {
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (t < a.n_tris) v = make_uint4(a.vidx[3 * t], a.vidx[3 * t + 1], a.vidx[3 * t + 2], 1u);
      reinterpret_cast<uint4*>(a.index)[t] = v;
    }
    if (a.position && t < a.position_texels) {
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (t < a.n_vertices) {
        const float* p = a.positions + t * a.stride + a.offset;
        v = make_float4(p[0], p[1], p[2], 1.0f);
      }
      reinterpret_cast<float4*>(a.position)[t] = v;
    }
  }
}

__global__ __launch_bounds__(kRefitBlock) void k_export_nodes(const uint32_t* nodes, uint64_t n_nodes,
                                                               const uint64_t* root_first, uint32_t n_roots,
                                                               uint32_t* out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kRefitBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kRefitBlock + threadIdx.x; i < n_nodes; i += stride) {
    uint32_t w[8];
    for (int k = 0; k < 8; k++) w[k] = nodes[8 * i + k];
    bool is_root = false;
    for (uint32_t r = 0; r < n_roots; r++) is_root = is_root || root_first[r] == i;
    if (is_root) {
      for (int c = 0; c < 3; c++) {
        if (__uint_as_float(w[c]) > __uint_as_float(w[c + 3])) {
          w[c] = __float_as_uint(1.0f);
          w[c + 3] = __float_as_uint(-1.0f);
        }
      }
    }
    for (int k = 0; k < 8; k++) out[8 * i + k] = w[k];
  }
}

}  // namespace

hipError_t launch_export_textures(const ExportArgs& a, hipStream_t s) {
  uint64_t most = a.bounds_texels > a.contents_texels ? a.bounds_texels : a.contents_texels;
  if (a.index && a.index_texels > most) most = a.index_texels;
  if (a.position && a.position_texels > most) most = a.position_texels;
  if (most) hipLaunchKernelGGL(k_export_textures, dim3(blocks_for(most)), dim3(kRefitBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_export_nodes(const uint32_t* nodes, uint64_t n_nodes, const uint64_t* root_first, uint32_t n_roots,
                               uint32_t* out, hipStream_t s) {
  if (n_nodes)
    hipLaunchKernelGGL(k_export_nodes, dim3(blocks_for(n_nodes)), dim3(kRefitBlock), 0, s, nodes, n_nodes, root_first,
                       n_roots, out);
  return hipGetLastError();
}

hipError_t launch_refit(const RefitArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_refit_reset_bound, dim3(1), dim3(64), 0, s, a.coord_bound);
  hipLaunchKernelGGL(k_refit_bound, dim3(blocks_for(a.n_vertices) < kBoundBlocks ? blocks_for(a.n_vertices) : kBoundBlocks),
                     dim3(kRefitBlock), 0, s, a);
  if (a.n_slots) hipLaunchKernelGGL(k_refit_tris, dim3(blocks_for(3 * a.n_slots)), dim3(kRefitBlock), 0, s, a);
  if (a.node_ids) {
    hipLaunchKernelGGL(k_refit_mark_clear, dim3(blocks_for(a.n_nodes)), dim3(kRefitBlock), 0, s, a.changed, a.n_nodes);
    if (a.n_ids) hipLaunchKernelGGL(k_refit_mark_ids, dim3(blocks_for(a.n_ids)), dim3(kRefitBlock), 0, s, a);
    hipLaunchKernelGGL(k_refit_mark_roots, dim3(1), dim3(64), 0, s, a);
    for (uint32_t d = 0; d < a.n_levels; d++) {
      const uint64_t count = a.level_off[d + 1] - a.level_off[d];
      if (count)
        hipLaunchKernelGGL(k_refit_mark_level, dim3(blocks_for(count)), dim3(kRefitBlock), 0, s, a,
                           a.levels + a.level_off[d], count);
    }
  }
  if (a.n_leaves) hipLaunchKernelGGL(k_refit_leaves, dim3(blocks_for(a.n_leaves)), dim3(kRefitBlock), 0, s, a);
  for (uint32_t d = a.n_levels; d-- > 0;) {
    const uint64_t count = a.level_off[d + 1] - a.level_off[d];
    if (count)
      hipLaunchKernelGGL(k_refit_level, dim3(blocks_for(count)), dim3(kRefitBlock), 0, s, a, a.levels + a.level_off[d],
                         count);
  }
  if (a.n_four) hipLaunchKernelGGL(k_refit_four, dim3(blocks_for(a.n_four)), dim3(kRefitBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace mbvh
