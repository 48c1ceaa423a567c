// Spatial ordering of queries: a bucket sort of query points by their Morton cell.
//
// A wave runs 64 queries in lock step, so its cost is that of the union of the lanes' traversal
// paths. Queries that arrive in an arbitrary order (a grid in row order, rays in generation order)
// put far-apart points in one wave. Reordering by Morton cell makes a wave's lanes spatial
// neighbours, which walk nearly the same nodes. Results are written at each query's own index,
// so the order changes speed only, never values.
//
// Passes (all graph-capturable: no memset, no allocation):
//   init     bounds <- (+inf, -inf), bucket counts <- 0
//   bounds   min / max of the finite coordinates (wave reduction, one atomic per wave and axis)
//   hist     bucket of each query (kept) and count per Morton bucket
//   scan     exclusive prefix sum of the counts within chunks of kScanChunk buckets (one
//            workgroup per chunk), then of the chunk totals (one workgroup)
//   scatter  order[chunk base + cursor[bucket]++] = query index, from the kept buckets
// Within a bucket the order depends on atomic timing; it carries no meaning.
#include <algorithm>

#include "mbvh_kernels.h"

namespace mbvh {
namespace {

constexpr int kOrderBlock = 256;
constexpr int kScanBlock = 256;
constexpr int kScanPer = 16;                         // buckets per thread of a chunk
constexpr uint32_t kScanChunk = kScanBlock * kScanPer;  // buckets per chunk
constexpr uint32_t kMaxChunks = 1024;                 // chunk totals are scanned by one workgroup

// Total order on float bits as unsigned integers (negative values reversed).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Spreads the low 10 bits of v to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t bucket_of(const float* p, const uint32_t* bounds, uint32_t axis_bits) {
  const uint32_t cells = 1u << axis_bits;
  uint32_t c[3];
  for (int k = 0; k < 3; k++) {
    const float lo = from_ordered_bits(bounds[k]);
    const float hi = from_ordered_bits(bounds[3 + k]);
    const float ext = hi - lo;
    float t = (ext > 0.0f) ? (p[k] - lo) / ext : 0.0f;
    t = (t == t) ? fminf(fmaxf(t, 0.0f), 1.0f) : 0.0f;  // NaN goes to cell 0, +-inf to an end cell
    c[k] = min(static_cast<uint32_t>(t * static_cast<float>(cells)), cells - 1u);
  }
  return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
}

__global__ __launch_bounds__(kOrderBlock) void k_order_init(uint32_t* bounds, uint32_t* counts, uint32_t n_buckets) {
  const uint32_t t = blockIdx.x * kOrderBlock + threadIdx.x;
  if (t < 3) bounds[t] = 0xFFFFFFFFu;
  else if (t < 6) bounds[t] = 0u;
  for (uint32_t b = t; b < n_buckets; b += gridDim.x * kOrderBlock) counts[b] = 0u;
}

__global__ __launch_bounds__(kOrderBlock) void k_order_bounds(const float* p, uint64_t n, uint32_t* bounds) {
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kOrderBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x; i < n; i += stride) {
    for (int k = 0; k < 3; k++) {
      const float v = p[3 * i + k];
      if (v - v == 0.0f) {  // finite
        const uint32_t o = ordered_bits(v);
        mn[k] = min(mn[k], o);
        mx[k] = max(mx[k], o);
      }
    }
  }
  for (int k = 0; k < 3; k++) {
    for (int off = 32; off > 0; off >>= 1) {
      mn[k] = min(mn[k], static_cast<uint32_t>(__shfl_xor(static_cast<int>(mn[k]), off)));
      mx[k] = max(mx[k], static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx[k]), off)));
    }
  }
  // block reduction in LDS, then one atomic per block and axis (atomics on one line serialise)
  __shared__ uint32_t red[kOrderBlock / 64][6];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0)
    for (int k = 0; k < 3; k++) {
      red[w][k] = mn[k];
      red[w][3 + k] = mx[k];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    uint32_t v = red[0][threadIdx.x];
    for (int j = 1; j < kOrderBlock / 64; j++)
      v = threadIdx.x < 3 ? min(v, red[j][threadIdx.x]) : max(v, red[j][threadIdx.x]);
    if (threadIdx.x < 3)
      atomicMin(&bounds[threadIdx.x], v);
    else
      atomicMax(&bounds[threadIdx.x], v);
  }
}

__global__ __launch_bounds__(kOrderBlock) void k_order_hist(const float* p, uint64_t n, const uint32_t* bounds,
                                                            uint32_t axis_bits, uint32_t* counts, uint32_t* keys) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kOrderBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t b = bucket_of(p + 3 * i, bounds, axis_bits);
    keys[i] = b;
    atomicAdd(&counts[b], 1u);
  }
}

// Exclusive prefix sum over the kScanBlock threads of a workgroup (block-wide, LDS for the waves).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_sums[kScanBlock / 64];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), off, 64));
    if (lane >= static_cast<uint32_t>(off)) x += y;
  }
  if (lane == 63u) wave_sums[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t k = 0; k < kScanBlock / 64; k++) {
    if (k < w) before += wave_sums[k];
    all += wave_sums[k];
  }
  *total = all;
  return before + x - v;
}

// Exclusive scan of each chunk of kScanChunk counts in place (thread t: a contiguous run of
// kScanPer buckets); the chunk's total goes to chunk_sums[chunk].
__global__ __launch_bounds__(kScanBlock) void k_order_scan_chunks(uint32_t* counts, uint32_t n_buckets,
                                                                  uint32_t* chunk_sums) {
  const uint32_t b0 = blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  uint32_t c[kScanPer];
  uint32_t s = 0;
  for (int k = 0; k < kScanPer; k++) {
    c[k] = b0 + k < n_buckets ? counts[b0 + k] : 0u;
    s += c[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, &total);
  for (int k = 0; k < kScanPer; k++) {
    if (b0 + k < n_buckets) counts[b0 + k] = run;
    run += c[k];
  }
  if (threadIdx.x == 0) chunk_sums[blockIdx.x] = total;
}

// Exclusive scan of the chunk totals in place (one workgroup, n_chunks <= kMaxChunks).
__global__ __launch_bounds__(kScanBlock) void k_order_scan_sums(uint32_t* chunk_sums, uint32_t n_chunks) {
  constexpr uint32_t kPer = kMaxChunks / kScanBlock;
  const uint32_t b0 = threadIdx.x * kPer;
  uint32_t c[kPer];
  uint32_t s = 0;
  for (uint32_t k = 0; k < kPer; k++) {
    c[k] = b0 + k < n_chunks ? chunk_sums[b0 + k] : 0u;
    s += c[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, &total);
  for (uint32_t k = 0; k < kPer; k++) {
    if (b0 + k < n_chunks) chunk_sums[b0 + k] = run;
    run += c[k];
  }
}
// Every bucket's cursor range holds exactly the queries counted into it, so pos < n.
__global__ __launch_bounds__(kOrderBlock) void k_order_scatter(const uint32_t* keys, uint64_t n, uint32_t* cursor,
                                                               const uint32_t* chunk_base, uint32_t* order) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kOrderBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t key = keys[i];
    const uint32_t pos = chunk_base[key / kScanChunk] + atomicAdd(&cursor[key], 1u);
    order[pos] = static_cast<uint32_t>(i);
  }
}

// Sort key of a ray in the by-mesh order: the mesh of its instance, or n_meshes when invalid.
__global__ __launch_bounds__(kOrderBlock) void k_mesh_keys(const uint32_t* ray_instance, const uint32_t* instance_geometry,
                                                           uint32_t n_instances, uint64_t n, uint32_t n_meshes,
                                                           uint32_t* keys, uint32_t* vals) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kOrderBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kOrderBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t inst = ray_instance[i];
    // without an instance table the ray names its mesh directly
    uint32_t g = !instance_geometry ? inst : (inst < n_instances ? instance_geometry[inst] : n_meshes);
    if (g > n_meshes) g = n_meshes;
    keys[i] = g;
    vals[i] = static_cast<uint32_t>(i);
  }
}

}  // namespace

uint64_t order_by_mesh_workspace_bytes(uint64_t n) { return radix_sort_workspace_bytes(n); }

hipError_t launch_order_by_mesh(const uint32_t* ray_instance, const uint32_t* instance_geometry, uint32_t n_instances,
                                uint64_t n, uint32_t n_meshes, void* workspace, const uint32_t** order, hipStream_t s) {
  uint32_t *keys, *vals, *sk, *sv;
  radix_sort_inputs(workspace, n, &keys, &vals);
  if (n) {
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(4096, (n + kOrderBlock - 1) / kOrderBlock));
    hipLaunchKernelGGL(k_mesh_keys, dim3(blocks), dim3(kOrderBlock), 0, s, ray_instance, instance_geometry, n_instances,
                       n, n_meshes, keys, vals);
  }
  const uint32_t bits = 32u - static_cast<uint32_t>(__builtin_clz(n_meshes | 1u));
  const hipError_t e = launch_radix_sort(workspace, n, bits, &sk, &sv, s);
  *order = sv;
  return e;
}

uint32_t order_axis_bits(uint64_t n) {
  // about 8 queries per bucket, 1..7 bits per axis (at most 2^21 buckets)
  uint32_t bits = 1;
  while (bits < 7 && (uint64_t{8} << (3 * bits)) < n) bits++;
  return bits;
}

// header (bounds) | bucket counts | chunk totals | per-query bucket
uint64_t order_workspace_bytes(uint64_t n) {
  return kOrderHeaderBytes + (uint64_t{4} << (3 * order_axis_bits(n))) + 4 * kMaxChunks + 4 * n;
}

hipError_t launch_order_points(const float* points, uint64_t n, uint32_t* order, void* workspace, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t axis_bits = order_axis_bits(n);
  const uint32_t n_buckets = 1u << (3 * axis_bits);
  uint32_t* bounds = static_cast<uint32_t*>(workspace);
  uint32_t* counts = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + kOrderHeaderBytes);
  uint32_t* chunk_sums = counts + n_buckets;
  uint32_t* keys = chunk_sums + kMaxChunks;
  const uint32_t n_chunks = (n_buckets + kScanChunk - 1) / kScanChunk;
  static_assert((uint64_t{1} << 21) / kScanChunk <= kMaxChunks, "chunk totals fit one workgroup");
  const uint32_t init_blocks = std::max<uint32_t>(1, std::min<uint32_t>(1024, (n_buckets + kOrderBlock - 1) / kOrderBlock));
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(4096, (n + kOrderBlock - 1) / kOrderBlock));
  hipLaunchKernelGGL(k_order_init, dim3(init_blocks), dim3(kOrderBlock), 0, s, bounds, counts, n_buckets);
  hipLaunchKernelGGL(k_order_bounds, dim3(std::min<uint32_t>(blocks, 512)), dim3(kOrderBlock), 0, s, points, n, bounds);
  hipLaunchKernelGGL(k_order_hist, dim3(blocks), dim3(kOrderBlock), 0, s, points, n, bounds, axis_bits, counts, keys);
  hipLaunchKernelGGL(k_order_scan_chunks, dim3(n_chunks), dim3(kScanBlock), 0, s, counts, n_buckets, chunk_sums);
  hipLaunchKernelGGL(k_order_scan_sums, dim3(1), dim3(kScanBlock), 0, s, chunk_sums, n_chunks);
  hipLaunchKernelGGL(k_order_scatter, dim3(blocks), dim3(kOrderBlock), 0, s, keys, n, counts, chunk_sums, order);
  return hipGetLastError();
}

}  // namespace mbvh
