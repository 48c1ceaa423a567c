// Backward (adjoint) kernels of raycastFirst and closestPoint, and the vertex-gradient reductions.
//
// The discrete result of the forward query (face, and the hit's barycentrics / distance / point)
// is held fixed: the backward reads the saved forward outputs and differentiates the continuous
// map from (ray or query point, triangle vertices) to the outputs, without re-traversing.
//
// raycastFirst. The hit solves  o + t d = a + u (b - a) + v (c - a),  i.e. M x = s with
// M = [-d | e1 | e2], x = (t, u, v), s = o - a, e1 = b - a, e2 = c - a. Outputs:
// distance = t |d|, point = o + t d, barycoord = (1 - u - v, u, v). With upstream gradients
// (g_dist, g_point, g_bary), the gradient of the loss with respect to x is
//   g_t = g_dist |d| + g_point . d,  g_u = g_bary.y - g_bary.x,  g_v = g_bary.z - g_bary.x,
// and lambda = M^-T (g_t, g_u, g_v) = (g_t (e1 x e2) + g_u (e2 x -d) + g_v (-d x e1)) / det,
// det = -d . (e1 x e2) (rows of M^-1 by the cofactor formula). Then
//   dL/do = lambda + g_point
//   dL/dd = t lambda + t g_point + g_dist t d / |d|
//   dL/da = -w lambda, dL/db = -u lambda, dL/dc = -v lambda     (w = 1 - u - v).
// closestPoint. D(q, a, b, c) = min over the triangle of |q - p|; at the minimiser p* with
// barycentrics (w, u, v) and n = (q - p*) / D (envelope theorem; the feasible set of
// barycentrics does not depend on the vertices):
//   dL/dq = g_dist n,  dL/da = -g_dist w n,  dL/db = -g_dist u n,  dL/dc = -g_dist v n.
// D = 0 (query on the surface) has no gradient; zero is returned there.
//
// Vertex gradients go to the geometry's vertices (the position attribute's entries, packed
// x, y, z per vertex), summed over all queries. Two reductions:
//   atomic         FP32 atomic adds (fast; the summation order, hence the rounding, varies)
//   deterministic  every contribution is keyed by vertex id; a stable LSD radix sort orders them
//                  by (vertex, query, corner); a segmented sum over fixed-size chunks of the
//                  sorted list, then across chunk boundaries in chunk order, gives bitwise
//                  identical results on every run.
#include "mbvh_closest_ff.h"
#include "mbvh_kernels.h"

namespace mbvh {
namespace {

constexpr int kGradBlock = 256;
constexpr uint32_t kNoFace = 0xFFFFFFFFu;

__device__ __forceinline__ f3 add3(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 scale3(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 load3(const float* p, uint64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void store3(float* p, uint64_t i, f3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

// Slot of a face (identity unless the mesh was built indirect); kNoFace if not in the mesh.
__device__ __forceinline__ uint32_t slot_of(const GradMesh& M, uint32_t face) {
  if (face >= M.face_limit) return kNoFace;
  return M.face_slot ? M.face_slot[face] : face;
}

// Vertex-gradient contribution e of a query: an atomic add, or a keyed entry for the
// deterministic reduction. Vertex `sink.sentinel` means "no contribution".
template <bool kAtomic>
__device__ __forceinline__ void put(const VertexSink& sink, uint64_t e, uint32_t vertex, f3 g) {
  if (kAtomic) {
    if (vertex != sink.sentinel) {
      unsafeAtomicAdd(&sink.grad[3ull * vertex + 0], g.x);
      unsafeAtomicAdd(&sink.grad[3ull * vertex + 1], g.y);
      unsafeAtomicAdd(&sink.grad[3ull * vertex + 2], g.z);
    }
  } else {
    sink.keys[e] = vertex;
    sink.vals[e] = static_cast<uint32_t>(e);
    store3(sink.contrib, e, g);
  }
}

template <bool kAtomic>
__global__ __launch_bounds__(kGradBlock) void k_raycast_first_bwd(RaycastBwdArgs A) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kGradBlock + threadIdx.x;
  if (i >= A.n) return;
  const uint32_t face = A.face[i];
  const uint32_t slot = face == kNoFace ? kNoFace : slot_of(A.mesh, face);
  f3 d_o = mk(0.0f, 0.0f, 0.0f), d_d = d_o, ga = d_o, gb = d_o, gc = d_o;
  uint32_t v0 = A.sink.sentinel, v1 = v0, v2 = v0;
  if (slot != kNoFace) {
    const float4 pa = A.mesh.tris[3ull * slot], pb = A.mesh.tris[3ull * slot + 1], pc = A.mesh.tris[3ull * slot + 2];
    const f3 a = mk(pa.x, pa.y, pa.z), b = mk(pb.x, pb.y, pb.z), c = mk(pc.x, pc.y, pc.z);
    v0 = A.mesh.vidx[3ull * slot];
    v1 = A.mesh.vidx[3ull * slot + 1];
    v2 = A.mesh.vidx[3ull * slot + 2];
    const f3 d = load3(A.directions, i);
    const f3 bc = load3(A.barycoord, i);  // (w, u, v)
    const float dlen = sqrtf(dot(d, d));
    const float t = A.distance[i] / dlen;
    const float g_dist = A.grad_distance ? A.grad_distance[i] : 0.0f;
    const f3 g_pt = A.grad_point ? load3(A.grad_point, i) : mk(0.0f, 0.0f, 0.0f);
    const f3 g_bc = A.grad_barycoord ? load3(A.grad_barycoord, i) : mk(0.0f, 0.0f, 0.0f);
    // upstream of (t, u, v) in double-float too: the three terms of lambda's numerator can cancel
    const ff dlen_ff = ff_sqrt(ff_add(ff_add(two_prod(d.x, d.x), two_prod(d.y, d.y)), two_prod(d.z, d.z)));
    const ff g_t = ff_add(ff_mulf(dlen_ff, g_dist),
                          ff_add(ff_add(two_prod(g_pt.x, d.x), two_prod(g_pt.y, d.y)), two_prod(g_pt.z, d.z)));
    const ff g_u = two_sum(g_bc.y, -g_bc.x);
    const ff g_v = two_sum(g_bc.z, -g_bc.x);
    // lambda in double-float: edges of a small triangle far from the origin lose most of their
    // FP32 bits to cancellation, and det loses more on grazing rays; the exact edge differences
    // keep lambda at FP32 accuracy.
    const ff3 e1 = ff3_diff(b, a), e2 = ff3_diff(c, a);
    const ff3 nd = {ff_from(-d.x), ff_from(-d.y), ff_from(-d.z)};
    const ff3 n = ff3_cross(e1, e2);
    const ff det = ff3_dot(nd, n);
    const ff3 c2 = ff3_cross(e2, nd), c3 = ff3_cross(nd, e1);
    const auto comp = [&](ff nk, ff c2k, ff c3k) {
      return ff_to_float(ff_div(ff_add(ff_add(ff_mul(nk, g_t), ff_mul(c2k, g_u)), ff_mul(c3k, g_v)), det));
    };
    const f3 lam = mk(comp(n.x, c2.x, c3.x), comp(n.y, c2.y, c3.y), comp(n.z, c2.z, c3.z));
    d_o = add3(lam, g_pt);
    d_d = add3(scale3(add3(lam, g_pt), t), scale3(d, g_dist * t / dlen));
    ga = scale3(lam, -bc.x);
    gb = scale3(lam, -bc.y);
    gc = scale3(lam, -bc.z);
  }
  if (A.grad_origins) store3(A.grad_origins, i, d_o);
  if (A.grad_directions) store3(A.grad_directions, i, d_d);
  if (A.want_vertices) {
    put<kAtomic>(A.sink, 3 * i + 0, v0, ga);
    put<kAtomic>(A.sink, 3 * i + 1, v1, gb);
    put<kAtomic>(A.sink, 3 * i + 2, v2, gc);
  }
}

template <bool kAtomic>
__global__ __launch_bounds__(kGradBlock) void k_closest_point_bwd(ClosestBwdArgs A) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kGradBlock + threadIdx.x;
  if (i >= A.n) return;
  const uint32_t face = A.face[i];
  const uint32_t slot = face == kNoFace ? kNoFace : slot_of(A.mesh, face);
  f3 d_q = mk(0.0f, 0.0f, 0.0f), ga = d_q, gb = d_q, gc = d_q;
  uint32_t v0 = A.sink.sentinel, v1 = v0, v2 = v0;
  if (slot != kNoFace) {
    v0 = A.mesh.vidx[3ull * slot];
    v1 = A.mesh.vidx[3ull * slot + 1];
    v2 = A.mesh.vidx[3ull * slot + 2];
    if (A.distance[i] > 0.0f) {
      // q - p in double-float from the saved face: an FP32 closest point loses the direction's
      // bits to cancellation when the query is near the surface
      const float4 pa = A.mesh.tris[3ull * slot], pb = A.mesh.tris[3ull * slot + 1], pc = A.mesh.tris[3ull * slot + 2];
      const f3 q = load3(A.points, i);
      const ClosestFF cf = closest_on_triangle_ff<false>(q, mk(0.0f, 0.0f, 0.0f), mk(pa.x, pa.y, pa.z), mk(pb.x, pb.y, pb.z),
                                                          mk(pc.x, pc.y, pc.z));
      const ff D = ff_sqrt(cf.dist2);
      if (D.hi > 0.0f) {
        const float g = A.grad_distance[i];
        const f3 nrm = mk(ff_to_float(ff_div(cf.qp.x, D)), ff_to_float(ff_div(cf.qp.y, D)),
                          ff_to_float(ff_div(cf.qp.z, D)));
        d_q = scale3(nrm, g);
        ga = scale3(nrm, -g * ff_to_float(cf.bw));
        gb = scale3(nrm, -g * ff_to_float(cf.bu));
        gc = scale3(nrm, -g * ff_to_float(cf.bv));
      }
    }
  }
  if (A.grad_points) store3(A.grad_points, i, d_q);
  if (A.want_vertices) {
    put<kAtomic>(A.sink, 3 * i + 0, v0, ga);
    put<kAtomic>(A.sink, 3 * i + 1, v1, gb);
    put<kAtomic>(A.sink, 3 * i + 2, v2, gc);
  }
}

__global__ __launch_bounds__(kGradBlock) void k_fill_zero(float* p, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kGradBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kGradBlock + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// ---- stable LSD radix sort of (key, value) pairs, 8-bit digits ---------------------------------
// Pass = hist (per-tile digit counts, digit-major) -> digit totals -> per-digit row scans ->
// scatter (stable ranks inside a tile from wave ballots). Elements are visited in index order
// (round r, thread t -> tile base + 256 r + t), so equal digits keep their order.
constexpr int kSortBlock = 256;
constexpr int kSortRounds = 16;
constexpr uint32_t kSortTile = kSortBlock * kSortRounds;

__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const uint32_t* keys, uint64_t n, uint32_t shift,
                                                         uint32_t n_tiles, uint32_t* hist) {
  __shared__ uint32_t h[256];
  const uint32_t tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  for (int r = 0; r < kSortRounds; r++) {
    const uint64_t i = base + static_cast<uint64_t>(r) * kSortBlock + tid;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(tid) * n_tiles + blockIdx.x] = h[tid];
}

// One block per digit: total count of that digit over all tiles.
__global__ __launch_bounds__(kSortBlock) void k_sort_totals(const uint32_t* hist, uint32_t n_tiles, uint32_t* totals) {
  __shared__ uint32_t part[kSortBlock];
  const uint32_t* row = hist + static_cast<uint64_t>(blockIdx.x) * n_tiles;
  uint32_t s = 0;
  for (uint32_t j = threadIdx.x; j < n_tiles; j += kSortBlock) s += row[j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t off = kSortBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = part[0];
}

// One block per digit: exclusive scan of the digit's row (tile order), offset by the counts of
// all smaller digits. hist becomes the first output position of (digit, tile).
__global__ __launch_bounds__(kSortBlock) void k_sort_scan(uint32_t* hist, uint32_t n_tiles, const uint32_t* totals) {
  __shared__ uint32_t part[kSortBlock];
  __shared__ uint32_t base_sh;
  const uint32_t d = blockIdx.x, tid = threadIdx.x;
  if (tid == 0) {
    uint32_t b = 0;
    for (uint32_t k = 0; k < d; k++) b += totals[k];
    base_sh = b;
  }
  uint32_t* row = hist + static_cast<uint64_t>(d) * n_tiles;
  const uint32_t per = (n_tiles + kSortBlock - 1) / kSortBlock;
  const uint32_t j0 = min(tid * per, n_tiles), j1 = min(j0 + per, n_tiles);
  uint32_t s = 0;
  for (uint32_t j = j0; j < j1; j++) s += row[j];
  part[tid] = s;
  __syncthreads();
  for (uint32_t off = 1; off < kSortBlock; off <<= 1) {
    const uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = base_sh + part[tid] - s;
  for (uint32_t j = j0; j < j1; j++) {
    const uint32_t c = row[j];
    row[j] = run;
    run += c;
  }
}

__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(const uint32_t* keys_in, const uint32_t* vals_in,
                                                            uint64_t n, uint32_t shift, uint32_t n_tiles,
                                                            const uint32_t* offsets, uint32_t* keys_out,
                                                            uint32_t* vals_out) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[kSortBlock / 64][256];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  run[tid] = offsets[static_cast<uint64_t>(tid) * n_tiles + blockIdx.x];
  for (int k = 0; k < kSortBlock / 64; k++) wc[k][tid] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < kSortRounds; r++) {
    const uint64_t i = base + static_cast<uint64_t>(r) * kSortBlock + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? vals_in[i] : 0u;
    const uint32_t dg = (key >> shift) & 255u;
    uint64_t m = __ballot(valid);
    for (int b = 0; b < 8; b++) {
      const bool bit = (dg >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(m & below));
    if (valid && rank == 0) wc[w][dg] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (valid) {
      uint32_t pos = run[dg] + rank;
      for (uint32_t k = 0; k < w; k++) pos += wc[k][dg];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int k = 0; k < kSortBlock / 64; k++) {
      add += wc[k][tid];
      wc[k][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }
}

// ---- segmented sum of the sorted contributions --------------------------------------------------
constexpr uint32_t kSegChunk = 16;

// One thread per chunk of kSegChunk sorted entries. Runs (equal keys) that start and end in the
// chunk are final; the chunk's last run, if it continues, leaves its partial sum in tail[c]; its
// first run, if it began earlier, leaves its partial sum in head[c].
__global__ __launch_bounds__(kGradBlock) void k_seg_chunks(const uint32_t* keys, const uint32_t* vals,
                                                          const float* contrib, uint64_t n, uint32_t sentinel,
                                                          float* grad, float4* head, float4* tail) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kGradBlock + threadIdx.x;
  const uint64_t i0 = c * kSegChunk;
  if (i0 >= n) return;
  const uint64_t i1 = min(i0 + kSegChunk, n);
  uint64_t i = i0;
  while (i < i1) {
    const uint32_t k = keys[i];
    f3 s = mk(0.0f, 0.0f, 0.0f);
    uint64_t j = i;
    while (j < i1 && keys[j] == k) {
      s = add3(s, load3(contrib, vals[j]));
      j++;
    }
    const bool starts = i > i0 || i0 == 0 || keys[i0 - 1] != k;
    const bool ends = j < i1 || j == n || keys[j] != k;
    if (k != sentinel) {
      if (starts && ends)
        store3(grad, k, s);
      else if (starts)
        tail[c] = make_float4(s.x, s.y, s.z, 0.0f);
      else
        head[c] = make_float4(s.x, s.y, s.z, 0.0f);
    }
    i = j;
  }
}

// One thread per chunk: if the chunk's last run starts in it and continues, add the following
// chunks' head partials in chunk order until the run ends.
__global__ __launch_bounds__(kGradBlock) void k_seg_finish(const uint32_t* keys, uint64_t n, uint32_t sentinel,
                                                          const float4* head, const float4* tail, float* grad) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kGradBlock + threadIdx.x;
  const uint64_t i0 = c * kSegChunk;
  if (i0 >= n) return;
  const uint64_t i1 = min(i0 + kSegChunk, n);
  if (i1 >= n) return;
  const uint32_t k = keys[i1 - 1];
  if (k == sentinel || keys[i1] != k) return;            // last run ends here (or carries nothing)
  if (keys[i0] == k && i0 > 0 && keys[i0 - 1] == k) return;  // it began in an earlier chunk
  float4 t = tail[c];
  f3 s = mk(t.x, t.y, t.z);
  for (uint64_t j = c + 1;; j++) {
    const float4 h = head[j];
    s = add3(s, mk(h.x, h.y, h.z));
    const uint64_t e = (j + 1) * kSegChunk;
    if (e >= n || keys[e] != k) break;
  }
  store3(grad, k, s);
}

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t{255}; }

struct GradLayout {
  uint64_t m, n_tiles, n_chunks;
  uint64_t keys0, vals0, keys1, vals1, contrib, hist, totals, head, tail, total;
};

GradLayout grad_layout(uint64_t n_queries) {
  GradLayout L{};
  L.m = 3 * n_queries;
  L.n_tiles = (L.m + kSortTile - 1) / kSortTile;
  L.n_chunks = (L.m + kSegChunk - 1) / kSegChunk;
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t at = o;
    o += align256(bytes);
    return at;
  };
  L.keys0 = take(4 * L.m);
  L.vals0 = take(4 * L.m);
  L.keys1 = take(4 * L.m);
  L.vals1 = take(4 * L.m);
  L.contrib = take(12 * L.m);
  L.hist = take(4 * 256 * L.n_tiles);
  L.totals = take(4 * 256);
  L.head = take(16 * L.n_chunks);
  L.tail = take(16 * L.n_chunks);
  L.total = o;
  return L;
}

uint32_t blocks_for(uint64_t n, uint32_t per_block, uint32_t cap = 1u << 20) {
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(cap, (n + per_block - 1) / per_block)));
}

// Sorts the keyed contributions and sums them per vertex into grad (zeroed first).
hipError_t reduce_deterministic(const GradLayout& L, char* ws, uint32_t sentinel, float* grad, uint64_t n_vertices,
                                hipStream_t s) {
  uint32_t* keys[2] = {reinterpret_cast<uint32_t*>(ws + L.keys0), reinterpret_cast<uint32_t*>(ws + L.keys1)};
  uint32_t* vals[2] = {reinterpret_cast<uint32_t*>(ws + L.vals0), reinterpret_cast<uint32_t*>(ws + L.vals1)};
  uint32_t* hist = reinterpret_cast<uint32_t*>(ws + L.hist);
  uint32_t* totals = reinterpret_cast<uint32_t*>(ws + L.totals);
  const uint32_t bits = 32u - static_cast<uint32_t>(__builtin_clz(sentinel | 1u));
  const uint32_t n_tiles = static_cast<uint32_t>(L.n_tiles);
  int cur = 0;
  for (uint32_t shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_sort_hist, dim3(n_tiles), dim3(kSortBlock), 0, s, keys[cur], L.m, shift, n_tiles, hist);
    hipLaunchKernelGGL(k_sort_totals, dim3(256), dim3(kSortBlock), 0, s, hist, n_tiles, totals);
    hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kSortBlock), 0, s, hist, n_tiles, totals);
    hipLaunchKernelGGL(k_sort_scatter, dim3(n_tiles), dim3(kSortBlock), 0, s, keys[cur], vals[cur], L.m, shift,
                       n_tiles, hist, keys[cur ^ 1], vals[cur ^ 1]);
    cur ^= 1;
  }
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(3 * n_vertices, kGradBlock, 8192)), dim3(kGradBlock), 0, s, grad,
                     3 * n_vertices);
  const float* contrib = reinterpret_cast<const float*>(ws + L.contrib);
  float4* head = reinterpret_cast<float4*>(ws + L.head);
  float4* tail = reinterpret_cast<float4*>(ws + L.tail);
  const uint32_t cb = blocks_for(L.n_chunks, kGradBlock);
  hipLaunchKernelGGL(k_seg_chunks, dim3(cb), dim3(kGradBlock), 0, s, keys[cur], vals[cur], contrib, L.m, sentinel,
                     grad, head, tail);
  hipLaunchKernelGGL(k_seg_finish, dim3(cb), dim3(kGradBlock), 0, s, keys[cur], L.m, sentinel, head, tail, grad);
  return hipGetLastError();
}

}  // namespace

// ---- generic stable radix sort (also orders the rays of a batched launch by mesh) ----------------
namespace {
struct SortLayout {
  uint64_t n_tiles, keys0, vals0, keys1, vals1, hist, totals, total;
};
SortLayout sort_layout(uint64_t n) {
  SortLayout L{};
  L.n_tiles = (n + kSortTile - 1) / kSortTile;
  uint64_t o = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t at = o;
    o += align256(bytes);
    return at;
  };
  L.keys0 = take(4 * n);
  L.vals0 = take(4 * n);
  L.keys1 = take(4 * n);
  L.vals1 = take(4 * n);
  L.hist = take(4 * 256 * std::max<uint64_t>(L.n_tiles, 1));
  L.totals = take(4 * 256);
  L.total = o;
  return L;
}
}  // namespace

uint64_t radix_sort_workspace_bytes(uint64_t n) { return sort_layout(n).total; }

void radix_sort_inputs(void* ws, uint64_t n, uint32_t** keys, uint32_t** vals) {
  const SortLayout L = sort_layout(n);
  *keys = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + L.keys0);
  *vals = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + L.vals0);
}

hipError_t launch_radix_sort(void* ws, uint64_t n, uint32_t bits, uint32_t** sorted_keys, uint32_t** sorted_vals,
                             hipStream_t s) {
  const SortLayout L = sort_layout(n);
  char* w = static_cast<char*>(ws);
  uint32_t* keys[2] = {reinterpret_cast<uint32_t*>(w + L.keys0), reinterpret_cast<uint32_t*>(w + L.keys1)};
  uint32_t* vals[2] = {reinterpret_cast<uint32_t*>(w + L.vals0), reinterpret_cast<uint32_t*>(w + L.vals1)};
  uint32_t* hist = reinterpret_cast<uint32_t*>(w + L.hist);
  uint32_t* totals = reinterpret_cast<uint32_t*>(w + L.totals);
  const uint32_t n_tiles = static_cast<uint32_t>(L.n_tiles);
  int cur = 0;
  if (n > 0) {
    for (uint32_t shift = 0; shift < bits; shift += 8) {
      hipLaunchKernelGGL(k_sort_hist, dim3(n_tiles), dim3(kSortBlock), 0, s, keys[cur], n, shift, n_tiles, hist);
      hipLaunchKernelGGL(k_sort_totals, dim3(256), dim3(kSortBlock), 0, s, hist, n_tiles, totals);
      hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(kSortBlock), 0, s, hist, n_tiles, totals);
      hipLaunchKernelGGL(k_sort_scatter, dim3(n_tiles), dim3(kSortBlock), 0, s, keys[cur], vals[cur], n, shift,
                         n_tiles, hist, keys[cur ^ 1], vals[cur ^ 1]);
      cur ^= 1;
    }
  }
  *sorted_keys = keys[cur];
  *sorted_vals = vals[cur];
  return hipGetLastError();
}

uint64_t grad_workspace_bytes(uint64_t n_queries, int32_t mode) {
  return mode == kGradDeterministic ? grad_layout(n_queries).total : 0;
}

template <typename Args, typename Kernel>
static hipError_t launch_bwd(Args a, Kernel k_atomic, Kernel k_det, int32_t mode, float* grad_positions,
                             uint64_t n_vertices, void* workspace, hipStream_t s) {
  const uint32_t blocks = blocks_for(a.n, kGradBlock, 0xFFFFFFFFu);
  a.want_vertices = grad_positions != nullptr;
  a.sink.sentinel = static_cast<uint32_t>(n_vertices);
  if (a.n == 0) {
    if (a.want_vertices)
      hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(3 * n_vertices, kGradBlock, 8192)), dim3(kGradBlock), 0, s,
                         grad_positions, 3 * n_vertices);
    return hipGetLastError();
  }
  if (!a.want_vertices || mode == kGradAtomic) {
    if (a.want_vertices) {
      a.sink.grad = grad_positions;
      hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(3 * n_vertices, kGradBlock, 8192)), dim3(kGradBlock), 0, s,
                         grad_positions, 3 * n_vertices);
    }
    hipLaunchKernelGGL(k_atomic, dim3(blocks), dim3(kGradBlock), 0, s, a);
    return hipGetLastError();
  }
  const GradLayout L = grad_layout(a.n);
  char* ws = static_cast<char*>(workspace);
  a.sink.keys = reinterpret_cast<uint32_t*>(ws + L.keys0);
  a.sink.vals = reinterpret_cast<uint32_t*>(ws + L.vals0);
  a.sink.contrib = reinterpret_cast<float*>(ws + L.contrib);
  hipLaunchKernelGGL(k_det, dim3(blocks), dim3(kGradBlock), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return reduce_deterministic(L, ws, a.sink.sentinel, grad_positions, n_vertices, s);
}

hipError_t launch_raycast_first_bwd(const RaycastBwdArgs& a, int32_t mode, float* grad_positions,
                                    uint64_t n_vertices, void* workspace, hipStream_t s) {
  if (a.n == 0 && !grad_positions) return hipSuccess;
  return launch_bwd(a, k_raycast_first_bwd<true>, k_raycast_first_bwd<false>, mode, grad_positions, n_vertices,
                    workspace, s);
}

hipError_t launch_closest_point_bwd(const ClosestBwdArgs& a, int32_t mode, float* grad_positions,
                                    uint64_t n_vertices, void* workspace, hipStream_t s) {
  if (a.n == 0 && !grad_positions) return hipSuccess;
  return launch_bwd(a, k_closest_point_bwd<true>, k_closest_point_bwd<false>, mode, grad_positions, n_vertices,
                    workspace, s);
}

}  // namespace mbvh
