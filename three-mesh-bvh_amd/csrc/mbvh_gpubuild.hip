// Device BVH construction in the reference's packed node format: the same bytes as the host builder
// (mbvh_build.cpp), which follows the reference's build step for step:
//   triangle bounds   MeshBVH.computePrimitiveBounds   src/core/MeshBVH.js:337-436
//   node bounds       getBounds                        src/core/build/computeBoundsUtils.js:4-68
//   split choice      getOptimalSplit                  src/core/build/splitUtils.js:21-307
//   partition         partition                        src/core/build/sortUtils.js:4-62
//   recursion         buildTree                        src/core/build/buildTree.js:10-121
//   packing           populateBuffer                   src/core/build/buildUtils.js:35-91
//
// Level-synchronous: the nodes of one depth are split together, one wave per node of at most
// kBuildBigCount triangles and, for the bigger nodes, a sequence of launches that each spread every
// big node's triangles over all workgroups in chunks (bins, split choice, counts, ranks, swaps,
// children). Children are allocated in pairs at the next level, so node ids are grouped by depth.
// Packing then counts subtree sizes deepest level first, hands out depth-first slots from the roots
// down and writes every node record.
//
// Why the results are identical to the sequential build:
//  - Arithmetic is IEEE double wherever the reference computes with JS numbers, rounded to float
//    where it stores into a Float32Array; this file is compiled without contraction, so no fused
//    multiply-add changes a rounding, and double division is correctly rounded on the device.
//  - A node box is the first (in primitive order) of the minimal / maximal values: the reduction
//    keys are (order-preserving bits of the float value, position), with -0 and +0 given the same
//    value bits, so ties resolve to the lowest position as the sequential strict < / > do; the value
//    is then re-read at that position (its sign of zero included). Rounding to float is monotonic,
//    so the minimum of the rounded values is the rounded minimum of the doubles, and a nonzero
//    difference of two floats never rounds to zero.
//  - SAH bin boxes only feed surface areas and costs, where the sign of a zero bound cannot change a
//    value, so bins accumulate with plain atomic min / max on the same keys. Costs of all 93
//    candidates are computed independently and the winner is the first minimal one below the leaf
//    cost, which is what the sequential `cost < bestCost` scan keeps.
//  - The Hoare partition (sortUtils.js) swaps the i-th element at or above the plane counted from
//    the left with the i-th element below it counted from the right, for every i up to the number of
//    misplaced elements; the device computes both ranks with prefix counts and swaps the same pairs.
#include "../../include/mbvh.h"
#include "mbvh_kernels.h"

#pragma clang fp contract(off)

namespace mbvh {
namespace {

constexpr int kBins = 32;                           // splitUtils.js:4
constexpr double kTraversalCost = 1.0;              // Constants.js:57
constexpr double kPrimitiveIntersectCost = 1.25;    // Constants.js:56
constexpr double kFloat32Epsilon = 5.9604644775390625e-08;  // 2^-24, Constants.js:71
constexpr uint32_t kLeafMark = 0xFFFFFFFFu;
constexpr uint32_t kBinWords = 3 * kBins * 7;       // per axis and bin: count, 3 min keys, 3 max keys
constexpr uint32_t kKeyPosInf = 0xFF800000u;        // key of +inf
constexpr uint32_t kKeyNegInf = 0x007FFFFFu;        // key of -inf
constexpr uint32_t kBuildBlock = 256;
constexpr uint32_t kWideChunk = kBuildChunk;

enum : uint32_t { kErrLeafLimit = 1u, kErrIndex = 2u };

// ---- keys ---------------------------------------------------------------------------------------

// Order-preserving unsigned key of a float; -0 and +0 share the key of +0.
__device__ __forceinline__ uint32_t fkey(float v) {
  uint32_t b = __float_as_uint(v);
  if ((b << 1) == 0u) b = 0u;
  return (b >> 31) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  return __uint_as_float((k >> 31) ? (k & 0x7FFFFFFFu) : ~k);
}

// Running bound keys of a set of primitives: mn = box min xyz, center min xyz; mx = box max xyz,
// center max xyz. A min key carries the position in its low word, a max key its complement, so the
// lowest position wins ties either way.
struct BoundKeys {
  unsigned long long mn[6];
  unsigned long long mx[6];
};

__device__ __forceinline__ void keys_init(BoundKeys& k) {
  for (int i = 0; i < 6; i++) {
    k.mn[i] = ~0ull;
    k.mx[i] = 0ull;
  }
}

__device__ __forceinline__ void key_min(unsigned long long& k, float v, uint32_t q) {
  if (v == v) {
    const unsigned long long c = (static_cast<unsigned long long>(fkey(v)) << 32) | q;
    k = c < k ? c : k;
  }
}
__device__ __forceinline__ void key_max(unsigned long long& k, float v, uint32_t q) {
  if (v == v) {
    const unsigned long long c = (static_cast<unsigned long long>(fkey(v)) << 32) | (0xFFFFFFFFu - q);
    k = c > k ? c : k;
  }
}

__device__ __forceinline__ float lo_of(float c, float h) {
  return static_cast<float>(static_cast<double>(c) - static_cast<double>(h));
}
__device__ __forceinline__ float hi_of(float c, float h) {
  return static_cast<float>(static_cast<double>(c) + static_cast<double>(h));
}

__device__ __forceinline__ void keys_add(BoundKeys& k, const float* c, const float* h, uint32_t q) {
  for (int d = 0; d < 3; d++) {
    key_min(k.mn[d], lo_of(c[d], h[d]), q);
    key_max(k.mx[d], hi_of(c[d], h[d]), q);
    key_min(k.mn[3 + d], c[d], q);
    key_max(k.mx[3 + d], c[d], q);
  }
}

// Value of bound key j (0..5 min keys, 6..11 max keys) re-read at the winning position.
__device__ float key_value(const GpuBuildArgs& A, int j, unsigned long long key) {
  const bool is_min = j < 6;
  const int k = is_min ? j : j - 6;
  if (is_min ? key == ~0ull : key == 0ull) return is_min ? __builtin_huge_valf() : -__builtin_huge_valf();
  const uint32_t low = static_cast<uint32_t>(key);
  const uint32_t q = is_min ? low : 0xFFFFFFFFu - low;
  const int d = k % 3;
  const float c = A.prim[(2 * d) * A.span + q];
  if (k >= 3) return c;
  const float h = A.prim[(2 * d + 1) * A.span + q];
  return is_min ? lo_of(c, h) : hi_of(c, h);
}

__device__ __forceinline__ float pick3(const float* v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__device__ __forceinline__ void load_prim(const GpuBuildArgs& A, uint32_t q, float* c, float* h) {
  for (int d = 0; d < 3; d++) {
    c[d] = A.prim[(2 * d) * A.span + q];
    h[d] = A.prim[(2 * d + 1) * A.span + q];
  }
}

// ---- reference box arithmetic (ArrayBoxUtilities.js) ----------------------------------------------

__device__ __forceinline__ double surface_area(const float* b) {  // ArrayBoxUtilities.js:100-108
  const double d0 = static_cast<double>(b[3]) - b[0];
  const double d1 = static_cast<double>(b[4]) - b[1];
  const double d2 = static_cast<double>(b[5]) - b[2];
  return 2 * (d0 * d1 + d1 * d2 + d2 * d0);
}

__device__ int longest_edge(const float* b) {  // ArrayBoxUtilities.js:22-41
  int idx = -1;
  double best = -__builtin_huge_val();
  for (int i = 0; i < 3; i++) {
    const double dist = static_cast<double>(b[i + 3]) - b[i];
    if (dist > best) {
      best = dist;
      idx = i;
    }
  }
  return idx;
}

// JavaScript `~~x` (ToInt32), splitUtils.js:207.
__device__ int32_t js_to_int32(double x) {
  if (!isfinite(x)) return 0;
  const double t = trunc(x);
  if (t > -2147483648.0 && t < 2147483648.0) return static_cast<int32_t>(t);
  double m = fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return static_cast<int32_t>(static_cast<uint32_t>(m));
}

__device__ __forceinline__ void grow(const GpuBuildArgs& A, uint32_t q, float* b) {  // ArrayBoxUtilities.js:73-97
  for (int d = 0; d < 3; d++) {
    const double c = A.prim[(2 * d) * A.span + q], h = A.prim[(2 * d + 1) * A.span + q];
    const double lo = c - h, hi = c + h;
    if (lo < b[d]) b[d] = static_cast<float>(lo);
    if (hi > b[d + 3]) b[d + 3] = static_cast<float>(hi);
  }
}

__device__ __forceinline__ void make_empty(float* b) {
  b[0] = b[1] = b[2] = __builtin_huge_valf();
  b[3] = b[4] = b[5] = -__builtin_huge_valf();
}

// Split choices made by one thread: CENTER, AVERAGE and the SAH over every primitive center when
// a node holds fewer than 8 primitives (splitUtils.js:21-42, 63-167, 285-307).
struct SerialWs {  // LDS state of the exhaustive SAH: one candidate per primitive
  double cand[kBins / 4];
  uint32_t count[kBins / 4];
  float lcache[kBins / 4][6];
  float rcache[kBins / 4][6];
  int order[kBins / 4];
};

__device__ void serial_split(const GpuBuildArgs& A, SerialWs& ws, const float* box, const float* cen, uint32_t off,
                             uint32_t cnt, int* axis_out, double* pos_out) {
  int axis = -1;
  double pos = 0.0;
  if (A.strategy == MBVH_CENTER) {
    axis = longest_edge(cen);
    if (axis != -1) pos = (static_cast<double>(cen[axis]) + cen[axis + 3]) / 2;
  } else if (A.strategy == MBVH_AVERAGE) {
    axis = longest_edge(box);
    if (axis != -1) {
      double sum = 0;
      for (uint32_t i = off; i < off + cnt; i++) sum += A.prim[(2 * axis) * A.span + i];
      pos = sum / static_cast<double>(cnt);
    }
  } else {
    const double root_area = surface_area(box);
    double best_cost = kPrimitiveIntersectCost * static_cast<double>(cnt);
    const double dcount = static_cast<double>(cnt);
    double* cand = ws.cand;
    uint32_t* count = ws.count;
    float(*lcache)[6] = ws.lcache;
    float(*rcache)[6] = ws.rcache;
    int* order = ws.order;
    const int n = static_cast<int>(cnt);
    for (int a = 0; a < 3; a++) {
      for (int b = 0; b < n; b++) {
        cand[b] = A.prim[(2 * a) * A.span + off + b];
        count[b] = 0;
        make_empty(lcache[b]);
        make_empty(rcache[b]);
        order[b] = b;
      }
      for (int i = 1; i < n; i++) {  // insertion sort by candidate
        const int key = order[i];
        int j = i - 1;
        while (j >= 0 && cand[order[j]] > cand[key]) {
          order[j + 1] = order[j];
          j--;
        }
        order[j + 1] = key;
      }
      int split_count = 0;
      for (int b = 0; b < n; b++) {
        if (split_count > 0 && cand[order[split_count - 1]] == cand[order[b]]) continue;
        order[split_count++] = order[b];
      }
      for (uint32_t i = off; i < off + cnt; i++) {
        const double center = A.prim[(2 * a) * A.span + i];
        for (int b = 0; b < split_count; b++) {
          const int o = order[b];
          if (center >= cand[o]) {
            grow(A, i, rcache[o]);
          } else {
            grow(A, i, lcache[o]);
            count[o]++;
          }
        }
      }
      for (int b = 0; b < split_count; b++) {
        const int o = order[b];
        const double lc = static_cast<double>(count[o]);
        const double rc = dcount - lc;
        const double lp = count[o] != 0 ? surface_area(lcache[o]) / root_area : 0.0;
        const double rp = rc != 0 ? surface_area(rcache[o]) / root_area : 0.0;
        const double cost = kTraversalCost + kPrimitiveIntersectCost * (lp * lc + rp * rc);
        if (cost < best_cost) {
          axis = a;
          best_cost = cost;
          pos = cand[o];
        }
      }
    }
  }
  *axis_out = axis;
  *pos_out = pos;
}

// ---- wave and workgroup primitives ----------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
  const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), m, 64);
  const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), m, 64);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void keys_wave_reduce(BoundKeys& k) {
  for (int o = 32; o > 0; o >>= 1) {
    for (int i = 0; i < 6; i++) {
      unsigned long long t = shfl_xor_u64(k.mn[i], o);
      k.mn[i] = t < k.mn[i] ? t : k.mn[i];
      t = shfl_xor_u64(k.mx[i], o);
      k.mx[i] = t > k.mx[i] ? t : k.mx[i];
    }
  }
}

constexpr int kBlockWaves = kBuildBlock / 64;

// Sum over a workgroup of kBuildBlock threads (every thread gets it).
__device__ uint32_t block_sum(uint32_t v, uint32_t* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63u) == 0u) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int i = 0; i < kBlockWaves; i++) t += sh[i];
  __syncthreads();
  return t;
}

// Both children's keys reduced over the workgroup and merged into dst[0..23] with 64-bit atomics.
__device__ void block_keys_to(BoundKeys& kl, BoundKeys& kr, unsigned long long (*sh)[24], unsigned long long* dst) {
  keys_wave_reduce(kl);
  keys_wave_reduce(kr);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0u) {
    for (int i = 0; i < 6; i++) {
      sh[w][i] = kl.mn[i];
      sh[w][6 + i] = kl.mx[i];
      sh[w][12 + i] = kr.mn[i];
      sh[w][18 + i] = kr.mx[i];
    }
  }
  __syncthreads();
  if (threadIdx.x < 24) {
    const int t = threadIdx.x;
    const bool is_min = (t % 12) < 6;
    unsigned long long v = sh[0][t];
    for (int i = 1; i < kBlockWaves; i++) {
      const unsigned long long o = sh[i][t];
      v = is_min ? (o < v ? o : v) : (o > v ? o : v);
    }
    if (is_min) {
      if (v != ~0ull) atomicMin(dst + t, v);
    } else if (v != 0ull) {
      atomicMax(dst + t, v);
    }
  }
  __syncthreads();
}

// ---- split choice ---------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t bin_init(uint32_t word) {
  return word == 0 ? 0u : (word <= 3 ? kKeyPosInf : kKeyNegInf);
}

struct BinFrame {  // splitUtils.js:180-186: bins start at the centroid box's min, 32 equal widths
  double left[3];
  double width[3];
};

__device__ __forceinline__ BinFrame bin_frame(const float* cen) {
  BinFrame f;
  for (int a = 0; a < 3; a++) {
    f.left[a] = cen[a];
    f.width[a] = (static_cast<double>(cen[a + 3]) - f.left[a]) / kBins;
  }
  return f;
}

// Adds a primitive to its bin on every axis (splitUtils.js:197-213): count and box keys.
__device__ __forceinline__ void bin_prim(uint32_t* bins, const BinFrame& f, const float* c, const float* h) {
  float lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = lo_of(c[d], h[d]);
    hi[d] = hi_of(c[d], h[d]);
  }
  for (int a = 0; a < 3; a++) {
    int32_t idx = js_to_int32((static_cast<double>(c[a]) - f.left[a]) / f.width[a]);
    if (idx >= kBins) idx = kBins - 1;
    if (idx < 0) idx = 0;
    uint32_t* e = bins + (a * kBins + idx) * 7;
    atomicAdd(e, 1u);
    for (int d = 0; d < 3; d++) {
      if (lo[d] == lo[d]) atomicMin(e + 1 + d, fkey(lo[d]));
      if (hi[d] == hi[d]) atomicMax(e + 4 + d, fkey(hi[d]));
    }
  }
}

// The first cheapest of the 93 binned SAH candidates (splitUtils.js:215-284), by one whole wave;
// every lane returns the choice (axis -1: no candidate beats the leaf cost).
__device__ void sah_select(const uint32_t* bins, const float* box, const float* cen, uint32_t cnt, int lane,
                           int* axis_out, double* pos_out) {
  const BinFrame f = bin_frame(cen);
  const double root_area = surface_area(box);
  const double dcount = static_cast<double>(cnt);
  const double leaf_cost = kPrimitiveIntersectCost * dcount;
  double best = __builtin_huge_val();
  uint32_t best_id = 0xFFFFFFFFu;
  for (int rep = 0; rep < 2; rep++) {
    const uint32_t cid = lane + 64 * rep;
    if (cid >= 3 * kBins || (cid % kBins) == kBins - 1) continue;
    const int a = cid / kBins, i = cid % kBins;
    uint32_t lc = 0;
    uint32_t lk[6] = {kKeyPosInf, kKeyPosInf, kKeyPosInf, kKeyNegInf, kKeyNegInf, kKeyNegInf};
    uint32_t rk[6] = {kKeyPosInf, kKeyPosInf, kKeyPosInf, kKeyNegInf, kKeyNegInf, kKeyNegInf};
    for (int b = 0; b < kBins; b++) {
      const uint32_t* e = bins + (a * kBins + b) * 7;
      const bool to_left = b <= i;
      if (to_left) lc += e[0];
      for (int d = 0; d < 3; d++) {
        const uint32_t mn = e[1 + d], mx = e[4 + d];
        if (to_left) {
          lk[d] = mn < lk[d] ? mn : lk[d];
          lk[3 + d] = mx > lk[3 + d] ? mx : lk[3 + d];
        } else {
          rk[d] = mn < rk[d] ? mn : rk[d];
          rk[3 + d] = mx > rk[3 + d] ? mx : rk[3 + d];
        }
      }
    }
    float lb[6], rb[6];
    for (int d = 0; d < 6; d++) {
      lb[d] = key_float(lk[d]);
      rb[d] = key_float(rk[d]);
    }
    const double lcd = static_cast<double>(lc);
    const double rcd = dcount - lcd;
    const double lp = lc != 0 ? surface_area(lb) / root_area : 0.0;
    const double rp = rcd != 0 ? surface_area(rb) / root_area : 0.0;
    const double cost = kTraversalCost + kPrimitiveIntersectCost * (lp * lcd + rp * rcd);
    if (cost < leaf_cost && (cost < best || (cost == best && cid < best_id))) {
      best = cost;
      best_id = cid;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const uint32_t oi = __shfl_xor(best_id, o, 64);
    if (ob < best || (ob == best && oi < best_id)) {
      best = ob;
      best_id = oi;
    }
  }
  if (best_id == 0xFFFFFFFFu) {
    *axis_out = -1;
    *pos_out = 0.0;
  } else {
    const int a = best_id / kBins, i = best_id % kBins;
    const double left = a == 0 ? f.left[0] : (a == 1 ? f.left[1] : f.left[2]);
    const double width = a == 0 ? f.width[0] : (a == 1 ? f.width[1] : f.width[2]);
    *axis_out = a;
    *pos_out = left + width + i * width;
  }
}

// ---- node bookkeeping -----------------------------------------------------------------------------

__device__ __forceinline__ uint32_t n_lvl(const GpuBuildArgs& A) { return static_cast<uint32_t>(A.max_depth) + 2; }

__device__ __forceinline__ bool is_big(const GpuBuildArgs& A, uint32_t cnt, uint32_t depth) {
  return cnt > kBuildBigCount && cnt > static_cast<uint32_t>(A.leaf_size) && depth < static_cast<uint32_t>(A.max_depth);
}

__device__ void make_leaf(const GpuBuildArgs& A, uint32_t node, uint32_t off, uint32_t cnt) {
  A.ninfo[node] = make_uint4(off, cnt, kLeafMark, 0u);
  if (cnt > 0xFFFFu) {
    atomicOr(&A.err[0], static_cast<unsigned long long>(kErrLeafLimit));
    atomicMax(&A.err[1], static_cast<unsigned long long>(cnt));
  }
}

// A node's outcome at its level position p (node id - level start): split[p] = (elements left of the
// plane, axis), (0, 0) for a leaf; bound[24 p + j] = the children's bounds (left child j = 0..11,
// right 12..23: box min xyz, box max xyz, center-box min xyz, center-box max xyz). Children ids are
// handed out afterwards in level order (k_level_finish), so no counter is shared between nodes.
__device__ __forceinline__ void put_bound(const GpuBuildArgs& A, uint32_t p, int j, unsigned long long key) {
  const int jj = j % 12;                    // keys: box min, center min (0..5), box max, center max (6..11)
  const bool is_min = jj < 6;
  const int k = is_min ? jj : jj - 6;       // 0..2 box, 3..5 centers
  const int slot = (k < 3 ? 0 : 6) + (is_min ? 0 : 3) + k % 3;
  A.tbound[24ull * p + 12 * (j / 12) + slot] = key_value(A, jj, key);
}

// ---- small nodes: one wave per node ---------------------------------------------------------------

struct WaveLds {
  uint32_t bins[kBinWords];
  unsigned long long keys[24];
};

// Leaf at level position p.
__device__ __forceinline__ void leaf_at(const GpuBuildArgs& A, uint32_t node, uint32_t p, uint32_t off, uint32_t cnt) {
  make_leaf(A, node, off, cnt);
  A.tsplit[p] = make_uint2(0u, 0u);
}

__device__ void split_small(const GpuBuildArgs& A, WaveLds& L, uint32_t node, uint32_t p, uint32_t depth, int lane) {
  const uint4 info = A.ninfo[node];
  const uint32_t off = info.x, cnt = info.y;
  if (cnt <= static_cast<uint32_t>(A.leaf_size) || depth >= static_cast<uint32_t>(A.max_depth)) {
    if (lane == 0) leaf_at(A, node, p, off, cnt);
    return;
  }
  float box[6], cen[6];
  for (int i = 0; i < 6; i++) {
    box[i] = A.nbox[6ull * node + i];
    cen[i] = A.ncen[6ull * node + i];
  }
  int axis = -1;
  double pos = 0.0;
  if (A.strategy == MBVH_SAH && cnt >= kBins / 4) {
    for (uint32_t i = lane; i < kBinWords; i += 64) L.bins[i] = bin_init(i % 7);
    wave_sync();
    const BinFrame f = bin_frame(cen);
    for (uint32_t q = off + lane; q < off + cnt; q += 64) {
      float c[3], h[3];
      load_prim(A, q, c, h);
      bin_prim(L.bins, f, c, h);
    }
    wave_sync();
    sah_select(L.bins, box, cen, cnt, lane, &axis, &pos);
  } else {
    if (lane == 0) serial_split(A, *reinterpret_cast<SerialWs*>(L.bins), box, cen, off, cnt, &axis, &pos);
    axis = __shfl(axis, 0, 64);
    pos = __shfl(pos, 0, 64);
  }
  if (axis < 0) {
    if (lane == 0) leaf_at(A, node, p, off, cnt);
    return;
  }

  const float* cax = A.prim + (2ull * axis) * A.span;
  uint32_t n_left = 0;
  for (uint32_t q0 = off; q0 < off + cnt; q0 += 64) {
    const uint32_t q = q0 + lane;
    n_left += __popcll(__ballot(q < off + cnt && static_cast<double>(cax[q]) < pos));
  }
  if (n_left == 0 || n_left == cnt) {
    if (lane == 0) leaf_at(A, node, p, off, cnt);
    return;
  }
  const uint32_t split = off + n_left;

  // Ranks of the misplaced elements; bounds of the elements that stay.
  BoundKeys kl, kr;
  keys_init(kl);
  keys_init(kr);
  uint32_t misplaced = 0, run = 0;
  for (uint32_t q0 = off; q0 < off + cnt; q0 += 64) {
    const uint32_t q = q0 + lane;
    const bool valid = q < off + cnt;
    float c[3] = {0.f, 0.f, 0.f}, h[3] = {0.f, 0.f, 0.f};
    if (valid) load_prim(A, q, c, h);
    const bool f = valid && static_cast<double>(pick3(c, axis)) < pos;
    const unsigned long long ball = __ballot(f);
    const uint32_t below_q = run + lanes_below(ball);  // elements below the plane in [off, q)
    run += __popcll(ball);
    if (valid) {
      if (q < split) {
        if (!f) {
          A.pair_l[off + (q - off) - below_q] = q;
          misplaced++;
        } else {
          keys_add(kl, c, h, q);
        }
      } else {
        if (f)
          A.pair_r[off + below_q] = q;
        else
          keys_add(kr, c, h, q);
      }
    }
  }
  const uint32_t m = wave_sum(misplaced);
  wave_sync();

  // Swap pair i: the i-th misplaced element from the left with the i-th from the right.
  for (uint32_t i = lane; i < m; i += 64) {
    const uint32_t pl = A.pair_l[off + i];
    const uint32_t pr = A.pair_r[off + n_left - 1 - i];
    float cl[3], hl[3], cr[3], hr[3];
    load_prim(A, pl, cl, hl);
    load_prim(A, pr, cr, hr);
    const uint32_t tl = A.perm[pl], tr = A.perm[pr];
    for (int d = 0; d < 3; d++) {
      A.prim[(2 * d) * A.span + pl] = cr[d];
      A.prim[(2 * d + 1) * A.span + pl] = hr[d];
      A.prim[(2 * d) * A.span + pr] = cl[d];
      A.prim[(2 * d + 1) * A.span + pr] = hl[d];
    }
    A.perm[pl] = tr;
    A.perm[pr] = tl;
    keys_add(kr, cl, hl, pr);
    keys_add(kl, cr, hr, pl);
  }
  keys_wave_reduce(kl);
  keys_wave_reduce(kr);
  if (lane == 0) {
    for (int i = 0; i < 6; i++) {
      L.keys[i] = kl.mn[i];
      L.keys[6 + i] = kl.mx[i];
      L.keys[12 + i] = kr.mn[i];
      L.keys[18 + i] = kr.mx[i];
    }
  }
  wave_sync();  // also orders the swaps before the bounds re-read them
  if (lane == 0) A.tsplit[p] = make_uint2(n_left, static_cast<uint32_t>(axis));
  if (lane < 24) put_bound(A, p, lane, L.keys[lane]);
}

// ---- kernels: set-up ------------------------------------------------------------------------------

__global__ __launch_bounds__(kBuildBlock) void k_build_init(GpuBuildArgs A) {
  const uint32_t L = n_lvl(A);
  for (uint32_t i = threadIdx.x; i < 4 * L; i += kBuildBlock) A.lvl[i] = 0u;
  for (uint32_t i = threadIdx.x; i < 3; i += kBuildBlock) A.err[i] = i == 2 ? ~0ull : 0ull;
  __syncthreads();
  if (threadIdx.x == 0) A.lvl[L] = A.n_roots;  // level 0 count
}

__global__ __launch_bounds__(kBuildBlock) void k_build_root_keys(GpuBuildArgs A) {
  const uint64_t n = 12ull * A.n_roots;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * kBuildBlock)
    A.rkeys[i] = (i % 12) < 6 ? ~0ull : 0ull;
}

// Out-of-range index entries (the host builder's check, first offending entry reported).
__global__ __launch_bounds__(kBuildBlock) void k_build_check_index(GpuBuildArgs A) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBuildBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; i < A.index_count; i += stride) {
    const uint64_t v = A.index_elt == 2 ? static_cast<const uint16_t*>(A.index)[i] : static_cast<const uint32_t*>(A.index)[i];
    if (v >= A.n_vertices) {
      atomicOr(&A.err[0], static_cast<unsigned long long>(kErrIndex));
      atomicMin(&A.err[2], static_cast<unsigned long long>(i));
    }
  }
}

// Triangle at each position and its (center, half extent) per axis, MeshBVH.js:337-436.
__global__ __launch_bounds__(kBuildBlock) void k_build_prims(GpuBuildArgs A) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBuildBlock;
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; q < A.span; q += stride) {
    uint32_t lo = 0, hi = A.n_src;  // last source range starting at or before q
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (A.src[2 * mid + 1] <= q) lo = mid; else hi = mid;
    }
    const uint64_t tri = A.src[2 * lo] + (q - A.src[2 * lo + 1]);
    A.perm[q] = static_cast<uint32_t>(tri);
    uint64_t v[3] = {3 * tri, 3 * tri + 1, 3 * tri + 2};
    bool ok = true;
    for (int k = 0; k < 3; k++) {
      if (A.index) {
        if (v[k] >= A.index_count) {
          ok = false;
          continue;
        }
        v[k] = A.index_elt == 2 ? static_cast<const uint16_t*>(A.index)[v[k]] : static_cast<const uint32_t*>(A.index)[v[k]];
      }
      if (v[k] >= A.n_vertices) ok = false;
    }
    for (int el = 0; el < 3; el++) {
      float center = 0.f, half_ext = 0.f;
      if (ok) {
        const double a = A.positions[v[0] * A.stride + A.offset + el];
        const double b = A.positions[v[1] * A.stride + A.offset + el];
        const double c = A.positions[v[2] * A.stride + A.offset + el];
        double mn = a, mx = a;
        if (b < mn) mn = b;
        if (c < mn) mn = c;
        if (b > mx) mx = b;
        if (c > mx) mx = c;
        const double half = (mx - mn) / 2;
        center = static_cast<float>(mn + half);
        half_ext = static_cast<float>(half + (fabs(mn) + half) * kFloat32Epsilon);
      }
      A.prim[(2 * el) * A.span + q] = center;
      A.prim[(2 * el + 1) * A.span + q] = half_ext;
    }
    if (!ok) atomicOr(&A.err[0], static_cast<unsigned long long>(kErrIndex));
  }
}

// Root bounds (getBounds over each root range): workgroup reductions, then 64-bit atomics.
__global__ __launch_bounds__(kBuildBlock) void k_build_root_bounds(GpuBuildArgs A) {
  __shared__ unsigned long long part[kBlockWaves][12];
  for (uint32_t r = blockIdx.y; r < A.n_roots; r += gridDim.y) {
    const uint32_t off = A.roots[2 * r], cnt = A.roots[2 * r + 1];
    BoundKeys k;
    keys_init(k);
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; i < cnt; i += static_cast<uint64_t>(gridDim.x) * kBuildBlock) {
      float c[3], h[3];
      const uint32_t q = off + static_cast<uint32_t>(i);
      load_prim(A, q, c, h);
      keys_add(k, c, h, q);
    }
    keys_wave_reduce(k);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
      for (int i = 0; i < 6; i++) {
        part[w][i] = k.mn[i];
        part[w][6 + i] = k.mx[i];
      }
    }
    __syncthreads();
    if (threadIdx.x < 12) {
      unsigned long long v = part[0][threadIdx.x];
      for (int i = 1; i < kBlockWaves; i++) {
        const unsigned long long o = part[i][threadIdx.x];
        v = threadIdx.x < 6 ? (o < v ? o : v) : (o > v ? o : v);
      }
      if (threadIdx.x < 6) {
        if (v != ~0ull) atomicMin(&A.rkeys[12ull * r + threadIdx.x], v);
      } else if (v != 0ull) {
        atomicMax(&A.rkeys[12ull * r + threadIdx.x], v);
      }
    }
    __syncthreads();
  }
}

// Root nodes (ids 0..n_roots-1) from their keys.
__global__ __launch_bounds__(kBuildBlock) void k_build_roots(GpuBuildArgs A) {
  for (uint32_t r = blockIdx.x * kBuildBlock + threadIdx.x; r < A.n_roots; r += gridDim.x * kBuildBlock) {
    const uint32_t off = A.roots[2 * r], cnt = A.roots[2 * r + 1];
    for (int j = 0; j < 12; j++) {
      const float v = key_value(A, j, A.rkeys[12ull * r + j]);
      const int k = j < 6 ? j : j - 6;
      const int d = k % 3 + (j < 6 ? 0 : 3);
      if (k < 3)
        A.nbox[6ull * r + d] = v;
      else
        A.ncen[6ull * r + d] = v;
    }
    A.ninfo[r] = make_uint4(off, cnt, 0u, 0u);
    A.nroot[r] = r;
    if (is_big(A, cnt, 0)) {
      const uint32_t at = atomicAdd(&A.lvl[3 * n_lvl(A)], 1u);
      A.big[at] = r;
    }
  }
}

// ---- kernels: one level ---------------------------------------------------------------------------

// Nodes of at most kBuildBigCount primitives: one wave per node.
__global__ __launch_bounds__(kBuildBlock) void k_build_level_small(GpuBuildArgs A, uint32_t depth) {
  __shared__ WaveLds lds[kBlockWaves];
  const uint32_t L = n_lvl(A);
  const uint32_t start = A.lvl[depth], count = A.lvl[L + depth];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    A.lvl[depth + 1] = start + count;                            // next level's first id
    A.lvl[2 * L + depth + 1] = A.lvl[2 * L + depth] + A.lvl[3 * L + depth];  // next big-list start
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * kBlockWaves + w; i < count; i += gridDim.x * kBlockWaves) {
    const uint32_t node = start + i;
    if (is_big(A, A.ninfo[node].y, depth)) continue;
    wave_sync();
    split_small(A, lds[w], node, i, depth, lane);
  }
}

// Nodes of more than kBuildBigCount primitives ("wide" nodes, the big-node list of the level) are
// split by all workgroups together: the node's primitives are cut into chunks of kWideChunk, and
// each step below is one launch over every chunk (or every node) of the level.
struct WideLevel {
  uint32_t start;  // first big-list entry of the level
  uint32_t n;      // big nodes in the level
};

__device__ __forceinline__ WideLevel wide_level(const GpuBuildArgs& A, uint32_t depth) {
  const uint32_t L = n_lvl(A);
  return {A.lvl[2 * L + depth], A.lvl[3 * L + depth]};
}

// Big node holding chunk c (the last node whose first chunk is at or before c).
__device__ __forceinline__ uint32_t chunk_node(const GpuBuildArgs& A, uint32_t n, uint32_t c) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (A.wbase[mid] <= c) lo = mid; else hi = mid;
  }
  return lo;
}

// Chunk counts and first chunks of the level's big nodes; clears their bins, keys and pair counts.
__global__ __launch_bounds__(1024) void k_wide_plan(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t wsum[16];
  const WideLevel lv = wide_level(A, depth);
  if (lv.n == 0) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t running = 0;
  for (uint32_t b = 0; b < lv.n; b += 1024) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t c = i < lv.n ? (A.ninfo[A.big[lv.start + i]].y + kWideChunk - 1) / kWideChunk : 0u;
    uint32_t v = c;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(v, o, 64);
      if (lane >= o) v += t;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < 16; k++) {
      if (k < w) before += wsum[k];
      total += wsum[k];
    }
    if (i < lv.n) A.wbase[i] = running + before + v - c;
    running += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) A.wbase[lv.n] = running;
  for (uint64_t i = threadIdx.x; i < static_cast<uint64_t>(lv.n) * kBinWords; i += 1024)
    A.wbins[i] = bin_init(static_cast<uint32_t>(i % 7));
  for (uint64_t i = threadIdx.x; i < 24ull * lv.n; i += 1024) A.wkeys[i] = (i % 12) < 6 ? ~0ull : 0ull;
  for (uint32_t i = threadIdx.x; i < lv.n; i += 1024) A.wm[i] = 0u;
}

// SAH bins of every chunk, merged into the node's bins.
__global__ __launch_bounds__(kBuildBlock) void k_wide_bin(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t bins[kBlockWaves][kBinWords];
  const WideLevel lv = wide_level(A, depth);
  if (lv.n == 0) return;
  const uint32_t total = A.wbase[lv.n];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint32_t i = chunk_node(A, lv.n, c);
    const uint32_t node = A.big[lv.start + i];
    const uint4 info = A.ninfo[node];
    const uint32_t first = info.x + (c - A.wbase[i]) * kWideChunk, end = info.x + info.y;
    for (uint32_t t = lane; t < kBinWords; t += 64) bins[w][t] = bin_init(t % 7);
    float cen[6];
    for (int k = 0; k < 6; k++) cen[k] = A.ncen[6ull * node + k];
    const BinFrame f = bin_frame(cen);
    wave_sync();
    for (uint32_t r = 0; r < kWideChunk; r += kBuildBlock) {
      const uint32_t q = first + r + threadIdx.x;
      if (q < end) {
        float cc[3], hh[3];
        load_prim(A, q, cc, hh);
        bin_prim(bins[w], f, cc, hh);
      }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kBinWords; t += kBuildBlock) {
      const uint32_t word = t % 7;
      uint32_t v = bins[0][t];
      for (int k = 1; k < kBlockWaves; k++) {
        const uint32_t o = bins[k][t];
        v = word == 0 ? v + o : (word <= 3 ? (o < v ? o : v) : (o > v ? o : v));
      }
      uint32_t* dst = A.wbins + static_cast<uint64_t>(i) * kBinWords + t;
      if (word == 0) {
        if (v) atomicAdd(dst, v);
      } else if (word <= 3) {
        if (v != kKeyPosInf) atomicMin(dst, v);
      } else if (v != kKeyNegInf) {
        atomicMax(dst, v);
      }
    }
    __syncthreads();
  }
}

// Split choice of every big node, one wave per node; a node no candidate improves becomes a leaf.
__global__ __launch_bounds__(kBuildBlock) void k_wide_split(GpuBuildArgs A, uint32_t depth) {
  __shared__ SerialWs ws[kBlockWaves];
  const WideLevel lv = wide_level(A, depth);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * kBlockWaves + w; i < lv.n; i += gridDim.x * kBlockWaves) {
    const uint32_t node = A.big[lv.start + i];
    const uint4 info = A.ninfo[node];
    float box[6], cen[6];
    for (int k = 0; k < 6; k++) {
      box[k] = A.nbox[6ull * node + k];
      cen[k] = A.ncen[6ull * node + k];
    }
    int axis = -1;
    double pos = 0.0;
    if (A.strategy == MBVH_SAH) {
      sah_select(A.wbins + static_cast<uint64_t>(i) * kBinWords, box, cen, info.y, lane, &axis, &pos);
    } else {
      if (lane == 0) serial_split(A, ws[w], box, cen, info.x, info.y, &axis, &pos);
      axis = __shfl(axis, 0, 64);
      pos = __shfl(pos, 0, 64);
    }
    if (lane == 0) {
      A.waxis[i] = axis;
      A.wpos[i] = pos;
      if (axis < 0) make_leaf(A, node, info.x, info.y);
    }
  }
}

// Elements below the plane in every chunk.
__global__ __launch_bounds__(kBuildBlock) void k_wide_count(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t sh[kBlockWaves];
  const WideLevel lv = wide_level(A, depth);
  if (lv.n == 0) return;
  const uint32_t total = A.wbase[lv.n];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint32_t i = chunk_node(A, lv.n, c);
    const int axis = A.waxis[i];
    if (axis < 0) continue;
    const double pos = A.wpos[i];
    const uint4 info = A.ninfo[A.big[lv.start + i]];
    const uint32_t first = info.x + (c - A.wbase[i]) * kWideChunk, end = info.x + info.y;
    const float* cax = A.prim + (2ull * axis) * A.span;
    uint32_t below = 0;
    for (uint32_t r = 0; r < kWideChunk; r += kBuildBlock) {
      const uint32_t q = first + r + threadIdx.x;
      below += (q < end && static_cast<double>(cax[q]) < pos) ? 1u : 0u;
    }
    below = block_sum(below, sh);
    if (threadIdx.x == 0) A.wcnt[c] = below;
  }
}

// Per node: below-plane elements before each chunk, and the split position; a node whose plane
// leaves one side empty becomes a leaf.
__global__ __launch_bounds__(kBuildBlock) void k_wide_scan(GpuBuildArgs A, uint32_t depth) {
  const WideLevel lv = wide_level(A, depth);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t i = blockIdx.x * kBlockWaves + w; i < lv.n; i += gridDim.x * kBlockWaves) {
    if (A.waxis[i] < 0) continue;
    const uint32_t c0 = A.wbase[i], c1 = A.wbase[i + 1];
    uint32_t running = 0;
    for (uint32_t cb = c0; cb < c1; cb += 64) {
      const uint32_t c = cb + lane;
      const uint32_t v = c < c1 ? A.wcnt[c] : 0u;
      uint32_t incl = v;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
      }
      if (c < c1) A.wpre[c] = running + incl - v;
      running += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
      const uint32_t node = A.big[lv.start + i];
      const uint4 info = A.ninfo[node];
      A.wnl[i] = running;
      if (running == 0 || running == info.y) {
        make_leaf(A, node, info.x, info.y);
        A.waxis[i] = -1;
      }
    }
  }
}

// Ranks of the misplaced elements of every chunk (pair lists) and the bounds of the elements that stay.
__global__ __launch_bounds__(kBuildBlock) void k_wide_rank(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t sh[kBlockWaves];
  __shared__ unsigned long long shk[kBlockWaves][24];
  const WideLevel lv = wide_level(A, depth);
  if (lv.n == 0) return;
  const uint32_t total = A.wbase[lv.n];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint32_t i = chunk_node(A, lv.n, c);
    const int axis = A.waxis[i];
    if (axis < 0) continue;
    const double pos = A.wpos[i];
    const uint4 info = A.ninfo[A.big[lv.start + i]];
    const uint32_t off = info.x, end = info.x + info.y;
    const uint32_t n_left = A.wnl[i], split = off + n_left;
    const uint32_t first = off + (c - A.wbase[i]) * kWideChunk;
    uint32_t run = A.wpre[c], misplaced = 0;
    BoundKeys kl, kr;
    keys_init(kl);
    keys_init(kr);
    for (uint32_t r = 0; r < kWideChunk; r += kBuildBlock) {
      const uint32_t q = first + r + threadIdx.x;
      const bool valid = q < end;
      float cc[3] = {0.f, 0.f, 0.f}, hh[3] = {0.f, 0.f, 0.f};
      if (valid) load_prim(A, q, cc, hh);
      const bool f = valid && static_cast<double>(pick3(cc, axis)) < pos;
      const unsigned long long ball = __ballot(f);
      if (lane == 0) sh[w] = __popcll(ball);
      __syncthreads();
      uint32_t before = 0, all = 0;
      for (int k = 0; k < kBlockWaves; k++) {
        if (k < w) before += sh[k];
        all += sh[k];
      }
      const uint32_t below_q = run + before + lanes_below(ball);
      run += all;
      if (valid) {
        if (q < split) {
          if (!f) {
            A.pair_l[off + (q - off) - below_q] = q;
            misplaced++;
          } else {
            keys_add(kl, cc, hh, q);
          }
        } else {
          if (f)
            A.pair_r[off + below_q] = q;
          else
            keys_add(kr, cc, hh, q);
        }
      }
      __syncthreads();
    }
    misplaced = block_sum(misplaced, sh);
    if (threadIdx.x == 0 && misplaced) atomicAdd(&A.wm[i], misplaced);
    block_keys_to(kl, kr, shk, A.wkeys + 24ull * i);
  }
}

// Swaps of the pairs (pair chunk j of a node = pairs [j * kWideChunk, (j + 1) * kWideChunk)).
__global__ __launch_bounds__(kBuildBlock) void k_wide_swap(GpuBuildArgs A, uint32_t depth) {
  __shared__ unsigned long long shk[kBlockWaves][24];
  const WideLevel lv = wide_level(A, depth);
  if (lv.n == 0) return;
  const uint32_t total = A.wbase[lv.n];
  for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
    const uint32_t i = chunk_node(A, lv.n, c);
    if (A.waxis[i] < 0) continue;
    const uint32_t m = A.wm[i];
    const uint32_t p0 = (c - A.wbase[i]) * kWideChunk;
    if (p0 >= m) continue;
    const uint4 info = A.ninfo[A.big[lv.start + i]];
    const uint32_t off = info.x, n_left = A.wnl[i];
    BoundKeys kl, kr;
    keys_init(kl);
    keys_init(kr);
    for (uint32_t r = 0; r < kWideChunk; r += kBuildBlock) {
      const uint32_t p = p0 + r + threadIdx.x;
      if (p < m) {
        const uint32_t pl = A.pair_l[off + p];
        const uint32_t pr = A.pair_r[off + n_left - 1 - p];
        float cl[3], hl[3], cr[3], hr[3];
        load_prim(A, pl, cl, hl);
        load_prim(A, pr, cr, hr);
        const uint32_t tl = A.perm[pl], tr = A.perm[pr];
        for (int d = 0; d < 3; d++) {
          A.prim[(2 * d) * A.span + pl] = cr[d];
          A.prim[(2 * d + 1) * A.span + pl] = hr[d];
          A.prim[(2 * d) * A.span + pr] = cl[d];
          A.prim[(2 * d + 1) * A.span + pr] = hl[d];
        }
        A.perm[pl] = tr;
        A.perm[pr] = tl;
        keys_add(kr, cl, hl, pr);
        keys_add(kl, cr, hr, pl);
      }
    }
    block_keys_to(kl, kr, shk, A.wkeys + 24ull * i);
  }
}

// Outcome of every big node at its level position (leaves were written by the split or scan step).
__global__ __launch_bounds__(kBuildBlock) void k_wide_finish(GpuBuildArgs A, uint32_t depth) {
  const WideLevel lv = wide_level(A, depth);
  const uint32_t start = A.lvl[depth];
  for (uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x; i < lv.n; i += gridDim.x * kBuildBlock) {
    const int axis = A.waxis[i];
    const uint32_t p = A.big[lv.start + i] - start;
    if (axis < 0) {
      A.tsplit[p] = make_uint2(0u, 0u);
      continue;
    }
    A.tsplit[p] = make_uint2(A.wnl[i], static_cast<uint32_t>(axis));
    for (int j = 0; j < 24; j++) put_bound(A, p, j, A.wkeys[24ull * i + j]);
  }
}

// Split nodes per tile of kLevelTile level positions.
constexpr uint32_t kLevelTile = 1024;

__global__ __launch_bounds__(kLevelTile) void k_level_tiles(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t sh[kLevelTile / 64];
  const uint32_t count = A.lvl[n_lvl(A) + depth];
  const uint32_t tiles = (count + kLevelTile - 1) / kLevelTile;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t p = t * kLevelTile + threadIdx.x;
    uint32_t v = (p < count && A.tsplit[p].x != 0u) ? 1u : 0u;
    v = wave_sum(v);
    if ((threadIdx.x & 63u) == 0u) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int k = 0; k < kLevelTile / 64; k++) s += sh[k];
      A.ttile[t] = s;
    }
    __syncthreads();
  }
}

// Children of the level's split nodes, in level order (buildTree.js:97-117): ids at the next level,
// ranges, tree, bounds, and the next level's big-node list.
__global__ __launch_bounds__(kLevelTile) void k_level_finish(GpuBuildArgs A, uint32_t depth) {
  __shared__ uint32_t sh[kLevelTile / 64];
  __shared__ uint32_t base_sh;
  const uint32_t L = n_lvl(A);
  const uint32_t start = A.lvl[depth], count = A.lvl[L + depth];
  const uint32_t next_start = start + count;
  const uint32_t big_next = A.lvl[2 * L + depth] + A.lvl[3 * L + depth];
  const uint32_t tiles = (count + kLevelTile - 1) / kLevelTile;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    // splits in the earlier tiles (and, for block 0, the level total: the next level's size)
    uint32_t before = 0, all = 0;
    for (uint32_t k = threadIdx.x; k < tiles; k += kLevelTile) {
      const uint32_t v = A.ttile[k];
      all += v;
      if (k < t) before += v;
    }
    before = wave_sum(before);
    all = wave_sum(all);
    if (lane == 0) sh[w] = before;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
      for (int k = 0; k < kLevelTile / 64; k++) s += sh[k];
      base_sh = s;
    }
    __syncthreads();
    if (t == 0 && lane == 0) sh[w] = all;
    __syncthreads();
    if (t == 0 && threadIdx.x == 0) {
      uint32_t s = 0;
      for (int k = 0; k < kLevelTile / 64; k++) s += sh[k];
      A.lvl[L + depth + 1] = 2 * s;
    }
    const uint32_t p = t * kLevelTile + threadIdx.x;
    const uint2 sp = p < count ? A.tsplit[p] : make_uint2(0u, 0u);
    const bool split = sp.x != 0u;
    const unsigned long long ball = __ballot(split);
    __syncthreads();
    if (lane == 0) sh[w] = __popcll(ball);
    __syncthreads();
    uint32_t rank = base_sh + lanes_below(ball);
    for (int k = 0; k < w; k++) rank += sh[k];
    if (split) {
      const uint32_t node = start + p;
      const uint4 info = A.ninfo[node];
      const uint32_t first = next_start + 2 * rank;
      const uint32_t n_left = sp.x;
      A.ninfo[node] = make_uint4(info.x, info.y, first, sp.y);
      A.ninfo[first] = make_uint4(info.x, n_left, 0u, 0u);
      A.ninfo[first + 1] = make_uint4(info.x + n_left, info.y - n_left, 0u, 0u);
      const uint32_t tree = A.nroot[node];
      A.nroot[first] = tree;
      A.nroot[first + 1] = tree;
      for (int side = 0; side < 2; side++) {
        const float* b = A.tbound + 24ull * p + 12 * side;
        for (int k = 0; k < 6; k++) {
          A.nbox[6ull * (first + side) + k] = b[k];
          A.ncen[6ull * (first + side) + k] = b[6 + k];
        }
        const uint32_t c = side == 0 ? n_left : info.y - n_left;
        if (is_big(A, c, depth + 1)) {
          const uint32_t at = atomicAdd(&A.lvl[3 * L + depth + 1], 1u);  // big nodes are few
          A.big[big_next + at] = first + side;
        }
      }
    }
    __syncthreads();
  }
}

// ---- packing ------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBuildBlock) void k_pack_sizes(GpuBuildArgs A, uint32_t depth) {
  const uint32_t L = static_cast<uint32_t>(A.max_depth) + 2;
  const uint32_t start = A.lvl[depth], count = A.lvl[L + depth];
  for (uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x; i < count; i += gridDim.x * kBuildBlock) {
    const uint32_t n = start + i;
    const uint4 info = A.ninfo[n];
    A.nsize[n] = info.z == kLeafMark ? 1u : 1u + A.nsize[info.z] + A.nsize[info.z + 1];
  }
}

__global__ void k_pack_roots(GpuBuildArgs A) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t at = 0;
    for (uint32_t r = 0; r < A.n_roots; r++) {
      A.nslot[r] = at;
      A.root_nodes[r] = A.nsize[r];
      at += A.nsize[r];
    }
  }
}

__global__ __launch_bounds__(kBuildBlock) void k_pack_slots(GpuBuildArgs A, uint32_t depth) {
  const uint32_t L = static_cast<uint32_t>(A.max_depth) + 2;
  const uint32_t start = A.lvl[depth], count = A.lvl[L + depth];
  for (uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x; i < count; i += gridDim.x * kBuildBlock) {
    const uint32_t n = start + i;
    const uint4 info = A.ninfo[n];
    if (info.z == kLeafMark) continue;
    const uint32_t s = A.nslot[n];
    A.nslot[info.z] = s + 1;
    A.nslot[info.z + 1] = s + 1 + A.nsize[info.z];
  }
}

// Node records (buildUtils.js:35-91): bounds, then the right child's distance and the split axis, or
// the first triangle, the count and the leaf flag.
__global__ __launch_bounds__(kBuildBlock) void k_pack_write(GpuBuildArgs A, uint64_t total) {
  for (uint64_t n = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; n < total; n += static_cast<uint64_t>(gridDim.x) * kBuildBlock) {
    const uint4 info = A.ninfo[n];
    uint32_t* rec = A.packed + 8ull * A.nslot[n];
    for (int k = 0; k < 6; k++) rec[k] = __float_as_uint(A.nbox[6 * n + k]);
    if (info.z == kLeafMark) {
      rec[6] = static_cast<uint32_t>(A.base + info.x);
      rec[7] = (info.y & 0xFFFFu) | 0xFFFF0000u;
    } else {
      rec[6] = 1u + A.nsize[info.z];
      rec[7] = info.w;
    }
  }
}

// The primitive buffer in build order: triangle triplets of the index, or the indirect buffer.
__global__ __launch_bounds__(kBuildBlock) void k_pack_prims(GpuBuildArgs A) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBuildBlock;
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * kBuildBlock + threadIdx.x; q < A.span; q += stride) {
    const uint32_t t = A.perm[q];
    if (A.indirect) {
      if (A.out_elt == 2)
        static_cast<uint16_t*>(A.prims_out)[q] = static_cast<uint16_t>(t);
      else
        static_cast<uint32_t*>(A.prims_out)[q] = t;
    } else if (A.index_elt == 2) {
      const uint16_t* in = static_cast<const uint16_t*>(A.index);
      uint16_t* out = static_cast<uint16_t*>(A.prims_out);
      for (int k = 0; k < 3; k++) out[3 * q + k] = in[3ull * t + k];
    } else {
      const uint32_t* in = static_cast<const uint32_t*>(A.index);
      uint32_t* out = static_cast<uint32_t*>(A.prims_out);
      for (int k = 0; k < 3; k++) out[3 * q + k] = in[3ull * t + k];
    }
  }
}

uint32_t grid_for(uint64_t n, uint32_t cap) {
  const uint64_t b = (n + kBuildBlock - 1) / kBuildBlock;
  return static_cast<uint32_t>(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

hipError_t launch_build_init(const GpuBuildArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(kBuildBlock), 0, s, a);
  hipLaunchKernelGGL(k_build_root_keys, dim3(grid_for(12ull * a.n_roots, 1024)), dim3(kBuildBlock), 0, s, a);
  if (a.index && a.index_count)
    hipLaunchKernelGGL(k_build_check_index, dim3(grid_for(a.index_count, 4096)), dim3(kBuildBlock), 0, s, a);
  if (a.span) hipLaunchKernelGGL(k_build_prims, dim3(grid_for(a.span, 8192)), dim3(kBuildBlock), 0, s, a);
  const uint32_t gx = grid_for(a.span / 16 + 1, 1024);
  const uint32_t gy = a.n_roots < 1024 ? a.n_roots : 1024;
  hipLaunchKernelGGL(k_build_root_bounds, dim3(gx, gy), dim3(kBuildBlock), 0, s, a);
  hipLaunchKernelGGL(k_build_roots, dim3(grid_for(a.n_roots, 1024)), dim3(kBuildBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_build_level(const GpuBuildArgs& a, uint32_t depth, bool wide, hipStream_t s) {
  hipLaunchKernelGGL(k_build_level_small, dim3(2048), dim3(kBuildBlock), 0, s, a, depth);
  if (wide) {
    hipLaunchKernelGGL(k_wide_plan, dim3(1), dim3(1024), 0, s, a, depth);
    if (a.strategy == MBVH_SAH) hipLaunchKernelGGL(k_wide_bin, dim3(1024), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_split, dim3(256), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_count, dim3(1024), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_scan, dim3(256), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_rank, dim3(1024), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_swap, dim3(1024), dim3(kBuildBlock), 0, s, a, depth);
    hipLaunchKernelGGL(k_wide_finish, dim3(16), dim3(kBuildBlock), 0, s, a, depth);
  }
  hipLaunchKernelGGL(k_level_tiles, dim3(256), dim3(kLevelTile), 0, s, a, depth);
  hipLaunchKernelGGL(k_level_finish, dim3(256), dim3(kLevelTile), 0, s, a, depth);
  return hipGetLastError();
}

hipError_t launch_build_pack(const GpuBuildArgs& a, uint32_t n_levels, uint64_t total_nodes, hipStream_t s) {
  for (uint32_t d = n_levels; d-- > 0;)
    hipLaunchKernelGGL(k_pack_sizes, dim3(1024), dim3(kBuildBlock), 0, s, a, d);
  hipLaunchKernelGGL(k_pack_roots, dim3(1), dim3(64), 0, s, a);
  for (uint32_t d = 0; d < n_levels; d++) hipLaunchKernelGGL(k_pack_slots, dim3(1024), dim3(kBuildBlock), 0, s, a, d);
  if (total_nodes) hipLaunchKernelGGL(k_pack_write, dim3(grid_for(total_nodes, 8192)), dim3(kBuildBlock), 0, s, a, total_nodes);
  if (a.span) hipLaunchKernelGGL(k_pack_prims, dim3(grid_for(a.span, 8192)), dim3(kBuildBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace mbvh
