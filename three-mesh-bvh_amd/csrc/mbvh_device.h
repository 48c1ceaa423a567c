// Device-side data layout and FP32 helpers for the traversal kernels.
//
// HBM layout of one uploaded mesh (derived once from the reference buffers at upload):
//   nodes    WideNode[]  one 64-byte record per INTERNAL node of the reference tree, holding both
//                        children's float32 boxes (copied bit-exactly from the reference nodes),
//                        the two child references and the node's split axis. One 64 B load per
//                        visited node replaces the reference's two 32 B child reads.
//   roots    uint32[]    child reference of each root (a wide node, or a leaf for tiny roots).
//   bigleaf  uint2[]     (offset, count) of leaves that do not fit the inline reference encoding.
//   tris     float4[3]   per BVH slot (the reference's permuted triangle order):
//                          [a.x a.y a.z faceId] [b.x b.y b.z facingMask] [c.x c.y c.z vertexA]
//                        faceId is the reference faceIndex (slot, or indirect[slot]).
//   tri_vidx uint32[3]   per slot: the three vertex indices (for vertex-gradient scatter).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mbvh {

// ---- child / stack references --------------------------------------------------------------
// bit 31 clear: index of an internal WideNode.
// bit 31 set, bits 26..30 = c in 1..31: leaf with `c` triangles starting at slot bits 0..25.
// bit 31 set, bits 26..30 = 0: leaf described by bigleaf[bits 0..25].
constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefCountShift = 26;
constexpr uint32_t kRefCountMask = 0x1Fu;
constexpr uint32_t kRefOffsetMask = (1u << kRefCountShift) - 1u;
constexpr uint32_t kMaxInlineCount = 31;

struct alignas(16) WideNode {
  float4 b0;  // L.min.x L.min.y L.min.z L.max.x
  float4 b1;  // L.max.y L.max.z R.min.x R.min.y
  float4 b2;  // R.min.z R.max.x R.max.y R.max.z
  uint4 meta; // refL, refR, split axis, first slot of the right subtree (the left one's slots are below it)
};
static_assert(sizeof(WideNode) == 64, "wide node is one 64-byte record");

// Per-lane traversal stack: LDS for the first kStackLds entries, then a global overflow region
// (indexed [depth - kStackLds][thread]) sized at upload from the deepest tree.
constexpr int kStackLds = 16;
constexpr int kBlock = 256;
// LDS stack entries per quad (four lanes tracing one query on the 4-wide tree). A closest-point
// descent enters every child until a first bound exists (three pushes per 4-wide level), so the
// stack holds a few levels' worth in LDS: 48 x 8 B x 64 quads = 24 KB per block, which with the
// 8 KB candidate lists still lets four blocks (the VGPR limit) share a CU's LDS.
constexpr int kQuadLdsStack = 48;

// ---- small float3 helpers -------------------------------------------------------------------
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float maxabs(f3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }
__device__ __forceinline__ float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---- double-float ("float-float") arithmetic -------------------------------------------------
// An unevaluated sum hi + lo of two FP32 numbers carries ~48 significant bits. Used only for the
// rare near-degenerate triangle decisions and for the one winning triangle per query, so that the
// reported hit, distance and barycentrics agree with an FP64 evaluation while every GPU operation
// stays FP32. Error-free transformations: TwoSum (Knuth, TAOCP vol. 2 §4.2.2), FastTwoSum
// (Dekker 1971), TwoProduct via fused multiply-add; composite add/mul/div follow the double-double
// algorithms of Hida, Li & Bailey, "Library for double-double and quad-double arithmetic" (2001).
struct ff {
  float hi, lo;
};

// Error-free transformations are only exact if every add / subtract rounds on its own: a
// fused multiply-add formed across them (contraction) silently drops the low word's meaning.
#define MBVH_NO_CONTRACT _Pragma("clang fp contract(off)")

__device__ __forceinline__ ff ff_from(float a) {
  MBVH_NO_CONTRACT
  return {a, 0.0f};
}
__device__ __forceinline__ ff two_sum(float a, float b) {
  MBVH_NO_CONTRACT
  const float s = a + b;
  const float bb = s - a;
  const float e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__device__ __forceinline__ ff fast_two_sum(float a, float b) {
  MBVH_NO_CONTRACT
  const float s = a + b;
  const float e = b - (s - a);
  return {s, e};
}
__device__ __forceinline__ ff two_prod(float a, float b) {
  MBVH_NO_CONTRACT
  const float p = a * b;
  const float e = fmaf(a, b, -p);
  return {p, e};
}
__device__ __forceinline__ ff ff_add(ff a, ff b) {
  MBVH_NO_CONTRACT
  ff s = two_sum(a.hi, b.hi);
  ff t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ ff ff_neg(ff a) {
  MBVH_NO_CONTRACT
  return {-a.hi, -a.lo};
}
__device__ __forceinline__ ff ff_sub(ff a, ff b) {
  MBVH_NO_CONTRACT
  return ff_add(a, ff_neg(b));
}
__device__ __forceinline__ ff ff_mul(ff a, ff b) {
  MBVH_NO_CONTRACT
  ff p = two_prod(a.hi, b.hi);
  p.lo = fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo));
  return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ ff ff_mulf(ff a, float b) {
  MBVH_NO_CONTRACT
  ff p = two_prod(a.hi, b);
  p.lo = fmaf(a.lo, b, p.lo);
  return fast_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ ff ff_div(ff a, ff b) {
  MBVH_NO_CONTRACT
  const float q1 = a.hi / b.hi;
  ff r = ff_sub(a, ff_mulf(b, q1));
  const float q2 = r.hi / b.hi;
  r = ff_sub(r, ff_mulf(b, q2));
  const float q3 = r.hi / b.hi;
  ff q = fast_two_sum(q1, q2);
  return ff_add(q, ff_from(q3));
}
__device__ __forceinline__ ff ff_sqrt(ff a) {
  MBVH_NO_CONTRACT
  if (a.hi <= 0.0f) return {0.0f, 0.0f};
  const float x = rsqrtf(a.hi);
  const float ax = a.hi * x;
  const ff ax2 = two_prod(ax, ax);
  const float corr = (ff_sub(a, ax2)).hi * (x * 0.5f);
  return two_sum(ax, corr);
}
__device__ __forceinline__ float ff_to_float(ff a) {
  MBVH_NO_CONTRACT
  return a.hi + a.lo;
}
__device__ __forceinline__ int ff_sign(ff a) {
  MBVH_NO_CONTRACT
  return a.hi > 0.0f ? 1 : (a.hi < 0.0f ? -1 : (a.lo > 0.0f ? 1 : (a.lo < 0.0f ? -1 : 0)));
}

struct ff3 {
  ff x, y, z;
};
__device__ __forceinline__ ff3 ff3_diff(f3 a, f3 b) {
  MBVH_NO_CONTRACT
  // exact a - b
  return {two_sum(a.x, -b.x), two_sum(a.y, -b.y), two_sum(a.z, -b.z)};
}
__device__ __forceinline__ ff3 ff3_cross(ff3 a, ff3 b) {
  MBVH_NO_CONTRACT
  return {ff_sub(ff_mul(a.y, b.z), ff_mul(a.z, b.y)), ff_sub(ff_mul(a.z, b.x), ff_mul(a.x, b.z)),
          ff_sub(ff_mul(a.x, b.y), ff_mul(a.y, b.x))};
}
__device__ __forceinline__ ff ff3_dot(ff3 a, ff3 b) {
  MBVH_NO_CONTRACT
  return ff_add(ff_add(ff_mul(a.x, b.x), ff_mul(a.y, b.y)), ff_mul(a.z, b.z));
}
__device__ __forceinline__ ff ff3_dotf(f3 a, ff3 b) {
  MBVH_NO_CONTRACT
  return ff_add(ff_add(ff_mulf(b.x, a.x), ff_mulf(b.y, a.y)), ff_mulf(b.z, a.z));
}

}  // namespace mbvh
