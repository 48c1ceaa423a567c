// Wave64 traversal kernels for gfx950: first-hit raycast and closest point.
//
// One query per lane. Each lane walks the uploaded tree with a private stack whose first
// kStackLds entries live in LDS (one column per lane, so a wave's 64 lanes touch 64 distinct
// banks) and whose deeper entries spill to a global overflow region sized from the tree depth at
// upload. Traversal is "while-while": lanes at internal nodes step together, then lanes at
// leaves test triangles together.
//
// Exactness. All traversal arithmetic is FP32. Every FP32 decision that could differ from the
// exact one is bounded: triangle tests fall back to double-float (mbvh_device.h) inside a rigorous
// error band, and candidate hits are compared through error intervals. Candidates whose interval
// overlaps the best upper bound go to a short per-lane list (LDS); subtrees are pruned against
// that bound; the list is resolved in double-float once, after traversal, when all lanes of the
// wave do it together. The resolution keeps the first candidate in visit order among exactly
// equal minima, which is the reference's strict-< rule.
#include "mbvh_kernels.h"
#include "mbvh_closest_ff.h"

namespace mbvh {
namespace {

constexpr float kU = 5.9604644775390625e-08f;  // unit roundoff 2^-24
constexpr float kSlabEps = 8.0f * kU;          // widening of FP32 slab intervals (>= 3 roundings)
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kOverflow = 0xFFFFFFFEu;    // reported face when a stack would overflow
constexpr uint32_t kPruneFlag = 0x80000000u;
constexpr uint32_t kDone = 0x7FFFFFFFu;        // "no more work" traversal reference
constexpr int kCand = 4;                       // near-tie candidates per lane (LDS)
constexpr float kInf = __builtin_huge_valf();

// Per-lane LDS columns: stack entries (ref, aux) and candidates (slot, lower bound bits).
__shared__ uint2 sh_stack[kStackLds * kBlock];
__shared__ uint2 sh_cand[kCand * kBlock];

// Stack entries through address-space-qualified pointers: LDS accesses compile to ds_* and
// overflow accesses to global_* instructions. (Through generic pointers the compiler merges the
// two branches of a push or pop into one flat access, which waits on both memory counters.) The
// overflow load carries the nontemporal hint so it stays a distinct global load.
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(1))) uint64_t glb_u64;
__device__ __forceinline__ uint64_t pack_entry(uint2 v) { return static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32); }
__device__ __forceinline__ uint2 unpack_entry(uint64_t v) {
  return make_uint2(static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32));
}
__device__ __forceinline__ void lds_store(uint2* base, uint32_t i, uint2 v) { ((lds_u64*)base)[i] = pack_entry(v); }
__device__ __forceinline__ uint2 lds_load(uint2* base, uint32_t i) { return unpack_entry(((lds_u64*)base)[i]); }
__device__ __forceinline__ void ovf_store(uint2* base, uint64_t i, uint2 v) { ((glb_u64*)base)[i] = pack_entry(v); }
__device__ __forceinline__ uint2 ovf_load(uint2* base, uint64_t i) {
  return unpack_entry(__builtin_nontemporal_load((glb_u64*)base + i));
}

// Same as Stack, with the LDS part in a kernel-local array of kL entries per lane.
template <int kL>
struct StackL {
  uint2* lds;
  uint2* ovf;
  uint64_t ovf_stride;
  uint64_t gtid;
  uint32_t lane;
  int capacity;  // kL + overflow entries
  int sp;

  __device__ __forceinline__ bool push(uint32_t ref, uint32_t aux) {
    if (sp >= capacity) return false;
    if (sp < kL)
      lds_store(lds, sp * kBlock + lane, make_uint2(ref, aux));
    else
      ovf_store(ovf, static_cast<uint64_t>(sp - kL) * ovf_stride + gtid, make_uint2(ref, aux));
    sp++;
    return true;
  }
  __device__ __forceinline__ uint2 pop() {
    sp--;
    if (sp < kL) return lds_load(lds, sp * kBlock + lane);
    return ovf_load(ovf, static_cast<uint64_t>(sp - kL) * ovf_stride + gtid);
  }
};

__device__ __forceinline__ float widen_lo(float t) { return t * (t > 0.0f ? 1.0f - kSlabEps : 1.0f + kSlabEps); }
__device__ __forceinline__ float widen_hi(float t) { return t * (t > 0.0f ? 1.0f + kSlabEps : 1.0f - kSlabEps); }

struct Stack {
  uint2* ovf;
  uint64_t ovf_stride;
  uint64_t gtid;
  uint32_t lane;
  int capacity;  // kStackLds + overflow entries
  int sp;

  __device__ __forceinline__ bool push(uint32_t ref, uint32_t aux) {
    if (sp >= capacity) return false;
    if (sp < kStackLds)
      lds_store(sh_stack, sp * kBlock + lane, make_uint2(ref, aux));
    else
      ovf_store(ovf, static_cast<uint64_t>(sp - kStackLds) * ovf_stride + gtid, make_uint2(ref, aux));
    sp++;
    return true;
  }
  __device__ __forceinline__ uint2 pop() {
    sp--;
    if (sp < kStackLds) return lds_load(sh_stack, sp * kBlock + lane);
    return ovf_load(ovf, static_cast<uint64_t>(sp - kStackLds) * ovf_stride + gtid);
  }
};

__device__ __forceinline__ void leaf_range(uint32_t ref, const uint2* bigleaf, uint32_t& off,
                                           uint32_t& cnt) {
  const uint32_t c = (ref >> kRefCountShift) & kRefCountMask;
  if (c != 0) {
    off = ref & kRefOffsetMask;
    cnt = c;
  } else {
    const uint2 bl = bigleaf[ref & kRefOffsetMask];
    off = bl.x;
    cnt = bl.y;
  }
}

struct Tri {
  float4 p0, p1, p2;  // [a faceId] [b facingMask] [c vertexA]
  __device__ __forceinline__ f3 a() const { return mk(p0.x, p0.y, p0.z); }
  __device__ __forceinline__ f3 b() const { return mk(p1.x, p1.y, p1.z); }
  __device__ __forceinline__ f3 c() const { return mk(p2.x, p2.y, p2.z); }
  __device__ __forceinline__ uint32_t face() const { return __float_as_uint(p0.w); }
  __device__ __forceinline__ uint32_t mask() const { return __float_as_uint(p1.w); }
};
__device__ __forceinline__ Tri load_tri(const float4* tris, uint32_t slot) {
  const float4* p = tris + 3ull * slot;  // one address, the three rows at immediate offsets
  Tri t;
  t.p0 = p[0];
  t.p1 = p[1];
  t.p2 = p[2];
  return t;
}

// ------------------------------------------------------------------------------------------
// Ray / box. Same decision as intersectsNodeBounds (src/core/utils/intersectUtils.js:2-78):
// per-axis entry/exit ordered by the sign of 1/dir, NaN entries (0 * inf on a face through the
// origin) ignored, hit iff max(entries) <= min(exits) && entry <= far && exit >= near. The FP32
// interval is widened by a few ulps so a box the exact ray touches is never rejected.
// ------------------------------------------------------------------------------------------
struct RayBox {
  float ox, oy, oz;
  float ix, iy, iz;  // 1 / direction
  bool sx, sy, sz;   // 1/dir >= 0 (reference's `invdir >= 0` selection)
  float wa;          // absolute widening for the ray's low parts (0 for an FP32 ray)
};

// A ray given as hi + lo float pairs (a double split as hi = fl(x), lo = fl(x - hi)): the FP32
// kernels trace the hi parts with every error bound widened by the lo parts, and the exact steps use
// both. |lo| <= ulp(hi) / 2, so dividing by the direction changes each plane distance by at most
// |ol / d| (1 + 2u) plus a relative u (inside the relative widening); zero direction components
// have zero lo parts.
__device__ __forceinline__ float slab_lo_widening(f3 ol, float ix, float iy, float iz) {
  const float ax = ol.x == 0.0f ? 0.0f : fabsf(ol.x * ix);
  const float ay = ol.y == 0.0f ? 0.0f : fabsf(ol.y * iy);
  const float az = ol.z == 0.0f ? 0.0f : fabsf(ol.z * iz);
  return fmaxf(fmaxf(ax, ay), az) * (1.0f + 4.0f * kU);
}

__device__ __forceinline__ bool slab(const RayBox& r, float mnx, float mny, float mnz, float mxx,
                                     float mxy, float mxz, float near_t, float far_t, float& tmin) {
  const float tx0 = ((r.sx ? mnx : mxx) - r.ox) * r.ix;
  const float tx1 = ((r.sx ? mxx : mnx) - r.ox) * r.ix;
  const float ty0 = ((r.sy ? mny : mxy) - r.oy) * r.iy;
  const float ty1 = ((r.sy ? mxy : mny) - r.oy) * r.iy;
  const float tz0 = ((r.sz ? mnz : mxz) - r.oz) * r.iz;
  const float tz1 = ((r.sz ? mxz : mnz) - r.oz) * r.iz;
  const float t0 = widen_lo(fmaxf(fmaxf(tx0, ty0), tz0)) - r.wa;
  const float t1 = widen_hi(fminf(fminf(tx1, ty1), tz1)) + r.wa;
  tmin = t0;
  return (t0 <= t1) && (t0 <= far_t) && (t1 >= near_t);
}

// The quad traversal's box test, for rays with no zero and no subnormal direction component (all
// of 1/dir finite): one FMA per slab plane, b * (1/d) + (-o/d), with the interval widened per axis
// by an absolute bound instead of the relative ulps above. With S = (B + |o|) |1/d| per axis
// (B = largest |coordinate| of the mesh, so |b| <= B), each computed plane distance is within
// 5 u S of (b - o) / d: u S from -o * (1/d), 2 u S from 1/d and the FMA, u S from folding the
// widening into the offset, u S from the FMA's own rounding. The widening is 8 u S.
// ------------------------------------------------------------------------------------------
struct RayBox4 {
  float ix, iy, iz;  // 1 / direction
  float lx, ly, lz;  // -o / d - E (entry offsets)
  float hx, hy, hz;  // -o / d + E (exit offsets)
  bool sx, sy, sz;   // 1/dir >= 0
};

// With low parts (ol, dl) the exact plane distance moves by at most (|ol| + S |dl|) |1/d| (1 + 2u)
// more per axis, which is added to that axis's widening (zero for an FP32 ray).
template <bool kLo>
__device__ __forceinline__ RayBox4 ray_box4(f3 o, f3 d, f3 ol, f3 dl, float coord_max) {
  RayBox4 r;
  r.ix = 1.0f / d.x;
  r.iy = 1.0f / d.y;
  r.iz = 1.0f / d.z;
  const float ox = -o.x * r.ix, oy = -o.y * r.iy, oz = -o.z * r.iz;
  float ex = 8.0f * kU * ((coord_max + fabsf(o.x)) * fabsf(r.ix));
  float ey = 8.0f * kU * ((coord_max + fabsf(o.y)) * fabsf(r.iy));
  float ez = 8.0f * kU * ((coord_max + fabsf(o.z)) * fabsf(r.iz));
  if (kLo) {
    ex += (fabsf(ol.x) + (coord_max + fabsf(o.x)) * fabsf(r.ix) * fabsf(dl.x)) * fabsf(r.ix) * (1.0f + 4.0f * kU);
    ey += (fabsf(ol.y) + (coord_max + fabsf(o.y)) * fabsf(r.iy) * fabsf(dl.y)) * fabsf(r.iy) * (1.0f + 4.0f * kU);
    ez += (fabsf(ol.z) + (coord_max + fabsf(o.z)) * fabsf(r.iz) * fabsf(dl.z)) * fabsf(r.iz) * (1.0f + 4.0f * kU);
  }
  r.lx = ox - ex;
  r.ly = oy - ey;
  r.lz = oz - ez;
  r.hx = ox + ex;
  r.hy = oy + ey;
  r.hz = oz + ez;
  r.sx = r.ix >= 0.0f;
  r.sy = r.iy >= 0.0f;
  r.sz = r.iz >= 0.0f;
  return r;
}

// Rays the quad traversal does not take: a zero direction component (the reference's split-plane
// shortcut is then needed literally; see trace_first), or slab constants that are not finite.
template <bool kLo>
__device__ __forceinline__ bool literal_ray(f3 o, f3 d, f3 ol, f3 dl, float coord_max) {
  if (d.x == 0.0f || d.y == 0.0f || d.z == 0.0f) return true;
  const RayBox4 r = ray_box4<kLo>(o, d, ol, dl, coord_max);
  const float m = fabsf(r.ix) + fabsf(r.iy) + fabsf(r.iz) + fabsf(r.lx) + fabsf(r.ly) + fabsf(r.lz) + fabsf(r.hx) +
                  fabsf(r.hy) + fabsf(r.hz);
  return !(m < kInf);
}

__device__ __forceinline__ bool slab4(const RayBox4& r, float4 c0, float4 c1, float near_t, float far_t,
                                      float& tmin) {
  // child box: c0 = (min.x, min.y, min.z, max.x), c1 = (max.y, max.z, ref, 0)
  const float t0 = fmaxf(fmaxf(fmaf(r.sx ? c0.x : c0.w, r.ix, r.lx), fmaf(r.sy ? c0.y : c1.x, r.iy, r.ly)),
                         fmaf(r.sz ? c0.z : c1.y, r.iz, r.lz));
  const float t1 = fminf(fminf(fmaf(r.sx ? c0.w : c0.x, r.ix, r.hx), fmaf(r.sy ? c1.x : c0.y, r.iy, r.hy)),
                         fmaf(r.sz ? c1.y : c0.z, r.iz, r.hz));
  tmin = t0;
  return (t0 <= t1) & (t0 <= far_t) & (t1 >= near_t);
}

// ------------------------------------------------------------------------------------------
// Ray / triangle: three's Ray.intersectTriangle (GeometricTools IntrRay3Triangle3) as called by
// checkIntersection (src/utils/ThreeRayIntersectUtilities.js:21-47). With N = (b-a)x(c-a),
// Q = o - a and DdN = d.N: front-facing iff DdN < 0; hit iff the signed edge terms
// d.(QxE2), d.(E1xQ), DdN - both, and -Q.N (all sign-folded by DdN) are >= 0, then
// t = QdN / |DdN| and |o + t d - o| in [near, far].
// `mask` bit 0 admits front-facing hits, bit 1 back-facing ones (FrontSide 1, BackSide 2,
// DoubleSide 3; BackSide's swapped winding accepts exactly the back-facing hits).
// ------------------------------------------------------------------------------------------
struct TriTerms {
  ff ddn_abs, e0, e1, qdn;
  bool back;
  bool ok;
};

// The ray's direction in double-float (hi + lo; the lo part is 0 for an FP32 ray).
template <bool kLo>
__device__ __forceinline__ ff3 ray_dir_ff(f3 d, f3 dl) {
  if (kLo) return {two_sum(d.x, dl.x), two_sum(d.y, dl.y), two_sum(d.z, dl.z)};
  return {ff_from(d.x), ff_from(d.y), ff_from(d.z)};
}

// The exact sign terms for the ray o + ol + t (d + dl).
template <bool kLo>
__device__ TriTerms tri_terms_exact(f3 o, f3 d, f3 ol, f3 dl, f3 a, f3 b, f3 c) {
  TriTerms r;
  const ff3 E1 = ff3_diff(b, a), E2 = ff3_diff(c, a);
  ff3 Q = ff3_diff(o, a);
  if (kLo) Q = {ff_add(Q.x, ff_from(ol.x)), ff_add(Q.y, ff_from(ol.y)), ff_add(Q.z, ff_from(ol.z))};
  const ff3 N = ff3_cross(E1, E2);
  ff ddn, e0, e1;
  if (kLo) {
    const ff3 D = ray_dir_ff<true>(d, dl);
    ddn = ff3_dot(D, N);
    e0 = ff3_dot(D, ff3_cross(Q, E2));
    e1 = ff3_dot(D, ff3_cross(E1, Q));
  } else {
    ddn = ff3_dotf(d, N);
    e0 = ff3_dotf(d, ff3_cross(Q, E2));
    e1 = ff3_dotf(d, ff3_cross(E1, Q));
  }
  const int sg = ff_sign(ddn);
  r.ok = sg != 0;
  r.back = sg > 0;
  const bool neg = !r.back;  // sign = -1 for front-facing
  r.ddn_abs = neg ? ff_neg(ddn) : ddn;
  ff qn = ff3_dot(Q, N);
  r.e0 = neg ? ff_neg(e0) : e0;
  r.e1 = neg ? ff_neg(e1) : e1;
  r.qdn = neg ? qn : ff_neg(qn);  // QdN = -sign * Q.N
  return r;
}

// The near / far window of a query in double-float (lo parts 0 for FP32 bounds).
struct Window {
  float near_d, near_lo, far_d, far_lo;
};
__device__ __forceinline__ Window window_of(float near_d, float near_lo, float far_d, float far_lo) {
  return {near_d, near_lo, far_d, far_lo};
}

// |d + dl| in double-float.
template <bool kLo>
__device__ __forceinline__ ff ray_len_ff(f3 d, f3 dl) {
  if (kLo) {
    const ff3 D = ray_dir_ff<true>(d, dl);
    return ff_sqrt(ff3_dot(D, D));
  }
  return ff_sqrt(ff_add(ff_add(two_prod(d.x, d.x), two_prod(d.y, d.y)), two_prod(d.z, d.z)));
}

// Exact hit test; on a hit returns true and the ray parameter in double-float. The distance
// t |d| is compared with near / far in double-float.
// Arguments are passed by value so callers keep them in registers.
template <bool kLo>
__device__ __noinline__ bool hit_exact(f3 o, f3 d, f3 ol, f3 dl, f3 a, f3 b, f3 c, uint32_t mask, Window win,
                                       ff& t_out) {
  const TriTerms r = tri_terms_exact<kLo>(o, d, ol, dl, a, b, c);
  if (!r.ok) return false;
  if (!(mask & (r.back ? 2u : 1u))) return false;
  if (ff_sign(r.e0) < 0) return false;
  if (ff_sign(r.e1) < 0) return false;
  if (ff_sign(ff_sub(r.ddn_abs, ff_add(r.e0, r.e1))) < 0) return false;
  if (ff_sign(r.qdn) < 0) return false;
  const ff t = ff_div(r.qdn, r.ddn_abs);
  if (win.near_d > 0.0f || win.far_d < kInf) {
    const ff dist = ff_mul(t, ray_len_ff<kLo>(d, dl));
    if (win.near_d > 0.0f && ff_sign(ff_sub(dist, {win.near_d, win.near_lo})) < 0) return false;
    if (win.far_d < kInf && ff_sign(ff_sub(dist, {win.far_d, win.far_lo})) > 0) return false;
  }
  t_out = t;
  return true;
}

enum : int { kMiss = 0, kHit = 1, kUnsure = 2 };

// Per-ray constants of the FP32 test's error bounds. For an FP32 ray: kd = 32 u |d|_1. With low
// parts (|ol|_inf = w, |dl|_1 = e) each sign term moves by at most: d.N by e |N|_inf <= 2 e mE^2;
// the edge terms by 2 mE (e mQ + w (|d|_1 + e)); Q.N by 3 w |N|_inf <= 6 w mE^2 (the factors
// 2.0001 / 6.0001 cover the FP32 roundings of the bounds themselves).
struct RayBand {
  float kd, a, c6, dm;
};
template <bool kLo>
__device__ __forceinline__ RayBand ray_band(f3 ol, f3 dl, float dl1) {
  RayBand b;
  b.kd = 32.0f * kU * dl1;
  b.a = 0.0f;
  b.c6 = 0.0f;
  b.dm = 0.0f;
  if (kLo) {
    const float w = maxabs(ol), e = fabsf(dl.x) + fabsf(dl.y) + fabsf(dl.z);
    b.kd += 2.0001f * e;
    b.a = 2.0001f * w * (dl1 + e);
    b.c6 = 6.0001f * w;
    b.dm = e;
  }
  return b;
}

// FP32 test. kHit comes with t and a bound et on its error; kUnsure means the FP32 decision lies
// inside its error band and must be made exactly.
template <bool kLo>
__device__ __forceinline__ int hit_fast(f3 o, f3 d, float dlen, const RayBand& bd, const Tri& tr,
                                        uint32_t mask, float near_d, float far_d, float& t_out,
                                        float& et_out) {
  const f3 a = tr.a(), b = tr.b(), c = tr.c();
  const f3 e1 = sub(b, a), e2 = sub(c, a);
  const f3 n = cross(e1, e2);
  const float ddn = dot(d, n);
  const float mE = fmaxf(maxabs(e1), maxabs(e2));
  const float bN = bd.kd * mE * mE;
  if (!(fabsf(ddn) > bN)) return kUnsure;
  const bool back = ddn > 0.0f;
  if (!(mask & (back ? 2u : 1u))) return kMiss;
  const float s = back ? 1.0f : -1.0f;
  const float DdN = fabsf(ddn);
  const f3 q = sub(o, a);
  const float mQ = maxabs(q);
  const float bE = kLo ? (bd.kd * mQ + bd.a) * mE : bd.kd * mQ * mE;
  const float e0 = s * dot(d, cross(q, e2));
  if (e0 < -bE) return kMiss;
  const float f1 = s * dot(d, cross(e1, q));
  if (f1 < -bE) return kMiss;
  const float g = DdN - e0 - f1;
  const float bS = bN + 2.0f * bE + 4.0f * kU * (DdN + fabsf(e0) + fabsf(f1));
  if (g < -bS) return kMiss;
  const float qdn = -s * dot(q, n);
  const float bQ = kLo ? (64.0f * kU * mQ + bd.c6) * mE * mE : 64.0f * kU * mQ * mE * mE;
  if (qdn < -bQ) return kMiss;
  if (!(e0 > bE && f1 > bE && g > bS && qdn > bQ)) return kUnsure;
  const float t = qdn / DdN;
  const float et = t * (bQ / qdn + bN / DdN + 4.0f * kU);
  const float dist = t * dlen;
  if (near_d > 0.0f || far_d < kInf) {
    // |o + t d - o| against the window: undecided inside the distance's error band
    const float ed = (8.0f * kU * dist + et * dlen + (kLo ? t * bd.dm : 0.0f)) * (1.0f + 4.0f * kU);
    if (dist + ed < near_d || dist - ed > far_d) return kMiss;
    if (dist - ed < near_d || dist + ed > far_d) return kUnsure;
  }
  t_out = t;
  et_out = et;
  return kHit;
}

// ------------------------------------------------------------------------------------------
// Candidate list (per lane, in visit order). `lb` is a lower bound of the candidate's exact key
// (ray parameter or squared distance); -inf marks a candidate whose hit/miss is undecided. A
// candidate whose lb exceeds the best upper bound U can never win and is dropped.
// ------------------------------------------------------------------------------------------
struct Cands {
  uint32_t lane;
  int n;
  __device__ __forceinline__ void put(int i, uint32_t slot, float lb) {
    sh_cand[i * kBlock + lane] = make_uint2(slot, __float_as_uint(lb));
  }
  __device__ __forceinline__ uint2 get(int i) const { return sh_cand[i * kBlock + lane]; }
  // Drops candidates with lb > U, keeping order. Returns true if a slot is free afterwards.
  __device__ __forceinline__ bool compact(float U) {
    int m = 0;
    for (int i = 0; i < n; i++) {
      const uint2 e = get(i);
      if (__uint_as_float(e.y) <= U || !(__uint_as_float(e.y) == __uint_as_float(e.y))) {
        if (m != i) put(m, e.x, __uint_as_float(e.y));
        m++;
      }
    }
    n = m;
    return n < kCand;
  }
};

// ------------------------------------------------------------------------------------------
// Ties. Exactly equal results (a point whose closest feature is a vertex shared by several
// triangles) compute to identical values; results equal in exact arithmetic but computed along
// different paths (a shared edge seen from its two triangles) agree to about 2^-44. Values within
// 2^-40 of each other are ties, and a tie goes to the triangle the reference meets first. Both
// reference traversals are depth-first over the binary tree, roots in order, a leaf's triangles in
// slot order, and keep the first of equal results (strict < within leaves and across roots,
// raycastFirst.template.js:91-97 `<=` preferring the near child). At the deepest node whose subtree
// holds both slots, raycastFirst enters the child on the side the ray comes from (dir[splitAxis] >= 0:
// left, raycastFirst.template.js:46-48) and closestPointToPoint the child at the smaller box
// distance^2, the right one only if strictly smaller (src/core/cast/shapecast.js:76-92).
// ------------------------------------------------------------------------------------------
// A resolution's winner: its slot (kNoHit if none), exact value and that value's error bound.
struct Winner {
  uint32_t w;
  ff v;
  float err;
};

// Two hit distances are tied when within 2^-40 of each other, relatively.
__device__ __forceinline__ bool near_tie(ff a, ff b) {
  const ff d = ff_sub(a, b);
  return fabsf(d.hi) <= 9.094947017729282e-13f * fmaxf(fabsf(a.hi), fabsf(b.hi));  // 2^-40
}
// Two squared point distances are tied when their difference is within the sum of their
// evaluations' error bounds (ClosestFF::err): neither can then be shown the smaller.
__device__ __forceinline__ bool point_tie(ff a, float ea, ff b, float eb) {
  return fabsf(ff_sub(a, b).hi) <= (ea + eb) * (1.0f + 2.0f * kU);
}

// What the tie order reads of a mesh.
struct TieTree {
  const WideNode* nodes;
  const uint32_t* roots;
  const uint2* root_slots;
  uint32_t n_roots;
};
__device__ __forceinline__ TieTree tie_tree(const MeshView& M) { return {M.nodes, M.roots, M.root_slots, M.n_roots}; }

__device__ __forceinline__ uint32_t root_of(const TieTree& T, uint32_t s) {
  uint32_t r = 0;
  while (r < T.n_roots && !(s >= T.root_slots[r].x && s < T.root_slots[r].y)) r++;
  return r;
}

// The deepest wide node whose subtree holds slots a and b of root r, or kDone when they share a
// leaf. Each step descends one level (children follow their parents in the node array).
__device__ __forceinline__ uint32_t split_node(const TieTree& T, uint32_t r, uint32_t a, uint32_t b) {
  uint32_t ref = T.roots[r];
  while (!(ref & kRefLeaf)) {
    const uint4 meta = T.nodes[ref].meta;
    const bool ra = a >= meta.w, rb = b >= meta.w;
    if (ra != rb) return ref;
    ref = ra ? meta.y : meta.x;
  }
  return kDone;
}

// Whether the reference's raycastFirst meets slot a before slot b (a != b) on direction d.
__device__ __noinline__ bool ray_first_of(TieTree T, float dx, float dy, float dz, uint32_t a, uint32_t b) {
  const uint32_t ra = root_of(T, a), rb = root_of(T, b);
  if (ra != rb) return ra < rb;
  if (ra >= T.n_roots) return a < b;  // a slot outside every root's range: no tree order to follow
  const uint32_t n = split_node(T, ra, a, b);
  if (n == kDone) return a < b;
  const uint4 meta = T.nodes[n].meta;
  const float dax = meta.z == 0 ? dx : (meta.z == 1 ? dy : dz);
  return (a < meta.w) == (dax >= 0.0f);
}

// Squared distance from q + ql to a box in double-float (the clamp is exact; the differences are
// exact, the squares and sum carry double-float roundings).
template <bool kLo>
__device__ __forceinline__ ff box_score_ff(f3 q, f3 ql, float mnx, float mny, float mnz, float mxx, float mxy,
                                           float mxz) {
  auto axis = [](float p, float pl, float mn, float mx) {
    ff g = {0.0f, 0.0f};
    if (p < mn || (kLo && p == mn && pl < 0.0f)) {
      g = two_sum(mn, -p);
      if (kLo) g = ff_sub(g, ff_from(pl));
    } else if (p > mx || (kLo && p == mx && pl > 0.0f)) {
      g = two_sum(p, -mx);
      if (kLo) g = ff_add(g, ff_from(pl));
    }
    return ff_mul(g, g);
  };
  return ff_add(ff_add(axis(q.x, ql.x, mnx, mxx), axis(q.y, ql.y, mny, mxy)), axis(q.z, ql.z, mnz, mxz));
}

// Whether the reference's closestPointToPoint meets slot a before slot b (a != b) from q + ql.
template <bool kLo>
__device__ __noinline__ bool point_first_of(TieTree T, f3 q, f3 ql, uint32_t a, uint32_t b) {
  const uint32_t ra = root_of(T, a), rb = root_of(T, b);
  if (ra != rb) return ra < rb;
  if (ra >= T.n_roots) return a < b;  // a slot outside every root's range: no tree order to follow
  const uint32_t n = split_node(T, ra, a, b);
  if (n == kDone) return a < b;
  const WideNode nd = T.nodes[n];
  const ff sl = box_score_ff<kLo>(q, ql, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
  const ff sr = box_score_ff<kLo>(q, ql, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
  const bool right_first = ff_sign(ff_sub(sr, sl)) < 0;
  return (a < nd.meta.w) != right_first;
}

// Exact resolution for the raycast: the smallest exact t; ties (near_tie) go to the reference's
// first. kRefOrder: the candidates are not in the reference's order (four lanes per ray), so ties
// are ordered by ray_first_of; otherwise the list order is the reference's and the earlier wins.
template <bool kLo, bool kRefOrder>
__device__ __noinline__ Winner resolve_ray(Cands cl, float U, f3 o, f3 d, f3 ol, f3 dl, const float4* tris,
                                           TieTree T, uint32_t side_mask, Window win_d) {
  uint32_t win = kNoHit;
  ff best_t = {kInf, 0.0f};
  for (int i = 0; i < cl.n; i++) {
    const uint2 e = cl.get(i);
    const float lb = __uint_as_float(e.y);
    if (lb > U) continue;
    const Tri tr = load_tri(tris, e.x);
    const uint32_t mask = side_mask ? side_mask : tr.mask();
    ff t;
    if (!hit_exact<kLo>(o, d, ol, dl, tr.a(), tr.b(), tr.c(), mask, win_d, t)) continue;
    bool take = win == kNoHit;
    if (!take) {
      if (near_tie(t, best_t))
        take = kRefOrder && e.x != win && ray_first_of(T, d.x, d.y, d.z, e.x, win);
      else
        take = ff_sign(ff_sub(t, best_t)) < 0;
    }
    if (take) {
      win = e.x;
      best_t = t;
    }
  }
  return {win, best_t, 0.0f};
}

// ------------------------------------------------------------------------------------------
// First hit. Traversal order and tie rules of _raycastFirst
// (src/core/cast/raycastFirst.template.js:19-103): at each node the child on the side the ray
// comes from (dir[splitAxis] >= 0 ⇒ left) is finished first; inside a leaf the first closest
// triangle wins (strict <, iterationUtils.template.js:24-53); across children and roots the
// earlier-visited hit wins ties. The reference has no distance pruning; this kernel skips
// subtrees whose entry distance already exceeds the best hit's upper bound, which cannot change
// that result. Its split-plane shortcut (raycastFirst.template.js:69-84) is implied by the
// distance pruning whenever dir[splitAxis] != 0. Rays with a zero direction component are traced
// without distance pruning and with that shortcut applied literally: on such an axis the near
// child's hit point lies in the ray's own plane, so the shortcut fires iff the near subtree found
// any hit and origin[axis] <= far.min[axis].
// ------------------------------------------------------------------------------------------
struct FirstHit {
  uint32_t face;  // MBVH_NO_HIT, or kOverflow when the stack would overflow
  float dist;
  f3 p;
  f3 bary;
};

// The reported values of the winning triangle, re-evaluated in double-float: distance t |d|, point
// o + t d and barycoord (the exact terms of three's Ray.intersectTriangle).
template <bool kLo>
__device__ __forceinline__ FirstHit finish_hit(const float4* tris, uint32_t w, f3 o, f3 d, f3 ol, f3 dl) {
  const Tri tr = load_tri(tris, w);
  const TriTerms r = tri_terms_exact<kLo>(o, d, ol, dl, tr.a(), tr.b(), tr.c());
  const ff t = ff_div(r.qdn, r.ddn_abs);
  const ff u = ff_div(r.e0, r.ddn_abs);
  const ff v = ff_div(r.e1, r.ddn_abs);
  const ff wgt = ff_sub(ff_sub(ff_from(1.0f), u), v);
  FirstHit h;
  h.face = tr.face();
  h.dist = ff_to_float(ff_mul(t, ray_len_ff<kLo>(d, dl)));
  if (kLo) {
    const ff3 D = ray_dir_ff<true>(d, dl);
    h.p = mk(ff_to_float(ff_add(two_sum(o.x, ol.x), ff_mul(t, D.x))),
             ff_to_float(ff_add(two_sum(o.y, ol.y), ff_mul(t, D.y))),
             ff_to_float(ff_add(two_sum(o.z, ol.z), ff_mul(t, D.z))));
  } else {
    h.p = mk(ff_to_float(ff_add(ff_from(o.x), ff_mulf(t, d.x))), ff_to_float(ff_add(ff_from(o.y), ff_mulf(t, d.y))),
             ff_to_float(ff_add(ff_from(o.z), ff_mulf(t, d.z))));
  }
  h.bary = mk(ff_to_float(wgt), ff_to_float(u), ff_to_float(v));
  return h;
}

__device__ __forceinline__ FirstHit no_hit() {
  FirstHit h;
  h.face = kNoHit;
  h.dist = kInf;
  const float qnan = __int_as_float(0x7fc00000);
  h.p = mk(qnan, qnan, qnan);
  h.bary = h.p;
  return h;
}

// One ray (o + ol) + t (d + dl) against one mesh; `st` is the lane's traversal stack (any LDS depth).
// kLo = false: an FP32 ray (ol, dl unused).
template <bool kLo, typename StackT>
__device__ __forceinline__ FirstHit trace_first(const MeshView& M, StackT& st, uint32_t lane, f3 o, f3 d, f3 ol,
                                                f3 dl, uint32_t side_mask, Window win) {
  RayBox rb;
  rb.ox = o.x;
  rb.oy = o.y;
  rb.oz = o.z;
  rb.ix = 1.0f / d.x;
  rb.iy = 1.0f / d.y;
  rb.iz = 1.0f / d.z;
  rb.sx = rb.ix >= 0.0f;
  rb.sy = rb.iy >= 0.0f;
  rb.sz = rb.iz >= 0.0f;
  rb.wa = kLo ? slab_lo_widening(ol, rb.ix, rb.iy, rb.iz) : 0.0f;
  const float near_dist = win.near_d, far_dist = win.far_d;
  const float dlen = sqrtf(dot(d, d));
  const RayBand band = ray_band<kLo>(ol, dl, fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
  const bool literal = (d.x == 0.0f) || (d.y == 0.0f) || (d.z == 0.0f);

  float U = kInf;  // upper bound of the best hit's t
  uint32_t hits = 0;
  bool overflow = false;
  Cands cl;
  cl.lane = lane;
  cl.n = 0;

  for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
    uint32_t ref = M.roots[r];
    st.sp = 0;
    while (true) {
      // ---- internal nodes ----
      while (!(ref & kRefLeaf) && ref != kDone) {
        const WideNode nd = M.nodes[ref];
        float tl, tr;
        bool hl = slab(rb, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y, near_dist, far_dist, tl);
        bool hr = slab(rb, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w, near_dist, far_dist, tr);
        if (!literal) {
          hl = hl && (tl <= U);
          hr = hr && (tr <= U);
        }
        const uint32_t axis = nd.meta.z;
        const float daxis = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
        const bool l2r = daxis >= 0.0f;  // raycastFirst.template.js:46-48
        const uint32_t c1 = l2r ? nd.meta.x : nd.meta.y;
        const uint32_t c2 = l2r ? nd.meta.y : nd.meta.x;
        const bool h1 = l2r ? hl : hr;
        const bool h2 = l2r ? hr : hl;
        if (h1 && h2) {
          uint32_t aux;
          if (!literal) {
            aux = __float_as_uint(l2r ? tr : tl);
          } else {
            // daxis == +-0 means l2r, so the far child is the right one.
            const float oaxis = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
            const float fmin = axis == 0 ? nd.b1.z : (axis == 1 ? nd.b1.w : nd.b2.x);
            bool below = oaxis <= fmin;
            if (kLo && oaxis == fmin) below = (axis == 0 ? ol.x : (axis == 1 ? ol.y : ol.z)) <= 0.0f;
            const bool prunable = (daxis == 0.0f) && below;
            aux = (hits & ~kPruneFlag) | (prunable ? kPruneFlag : 0u);
          }
          if (!st.push(c2, aux)) {
            overflow = true;
            ref = kDone;
            break;
          }
          ref = c1;
        } else if (h1) {
          ref = c1;
        } else if (h2) {
          ref = c2;
        } else {
          ref = kDone;
          while (st.sp > 0) {
            const uint2 e = st.pop();
            if (!literal ? (__uint_as_float(e.y) <= U)
                         : !((e.y & kPruneFlag) && hits > (e.y & ~kPruneFlag))) {
              ref = e.x;
              break;
            }
          }
        }
      }
      // ---- leaves ----
      while (ref & kRefLeaf) {
        uint32_t off, cnt;
        leaf_range(ref, M.bigleaf, off, cnt);
        Tri next;
        if (cnt != 0) next = load_tri(M.tris, off);  // an empty leaf may sit at the slot count
        for (uint32_t s = off; s < off + cnt; s++) {
          const Tri tr = next;
          if (s + 1 < off + cnt) next = load_tri(M.tris, s + 1);
          const uint32_t mask = side_mask ? side_mask : tr.mask();
          float t = 0.0f, et = 0.0f;
          int res = hit_fast<kLo>(o, d, dlen, band, tr, mask, near_dist, far_dist, t, et);
          if (res == kUnsure && literal) {
            // the literal shortcut counts hits, so an undecided hit is decided now
            ff te;
            res = hit_exact<kLo>(o, d, ol, dl, tr.a(), tr.b(), tr.c(), mask, win, te) ? kHit : kMiss;
            t = ff_to_float(te);
            et = 2.0f * kU * t;
          }
          if (res == kMiss) continue;
          if (res == kHit) {
            hits++;
            const float lb = t - et;
            if (lb > U) continue;  // certainly behind a known hit
            U = fminf(U, t + et);
          }
          if (cl.n == kCand && !cl.compact(U)) {
            // full of live candidates: settle them now and keep the winner
            const Winner r = resolve_ray<kLo, false>(cl, U, o, d, ol, dl, M.tris, tie_tree(M), side_mask, win);
            const uint32_t w = r.w;
            cl.n = 0;
            if (w != kNoHit) {
              const float tw = ff_to_float(r.v);
              U = fminf(U, tw * (1.0f + 4.0f * kU));
              cl.put(0, w, tw * (1.0f - 4.0f * kU));
              cl.n = 1;
            }
          }
          cl.put(cl.n++, s, res == kHit ? t - et : -kInf);
        }
        ref = kDone;
        while (st.sp > 0) {
          const uint2 e = st.pop();
          if (!literal ? (__uint_as_float(e.y) <= U)
                       : !((e.y & kPruneFlag) && hits > (e.y & ~kPruneFlag))) {
            ref = e.x;
            break;
          }
        }
      }
      if (ref == kDone) break;
    }
  }

  FirstHit h = no_hit();
  if (overflow) {
    h.face = kOverflow;
  } else if (cl.n > 0) {
    const uint32_t w = resolve_ray<kLo, false>(cl, U, o, d, ol, dl, M.tris, tie_tree(M), side_mask, win).w;
    if (w != kNoHit) h = finish_hit<kLo>(M.tris, w, o, d, ol, dl);
  }
  return h;
}

// ------------------------------------------------------------------------------------------
// First hit with four lanes per ray on the 4-wide tree (a "quad" of lanes traces one ray).
// Node step: lane q tests child q of the current node; the quad exchanges (entry t, ref), sorts
// the hit children by t, continues with the nearest and pushes the others (the pushed t is
// re-checked against the best bound on pop). Leaf step: lane q tests triangles q, q+4, ... with
// its own candidate list; the bound is shared after each leaf. After traversal each lane
// resolves its candidates exactly and the quad keeps the smallest exact t. Any visit order gives
// the nearest hit; exactly equal hits may resolve to a different face than the reference's order,
// which ties allow. Rays with a zero direction component need the reference's split-plane
// shortcut literally, which is defined on the binary tree: they take the binary traversal, run
// identically by the four lanes.
// Sixteen rays per wave instead of 64, and four times the waves for a given ray count: a launch
// of few rays (latency-bound) gets more waves to overlap and shorter paths per wave.
// ------------------------------------------------------------------------------------------
constexpr int kQuadLds = kQuadLdsStack;

// Stack shared by the four lanes of a quad: every lane pushes and pops the same entries.
struct QuadStack {
  uint2* lds;
  uint2* ovf;
  uint64_t ovf_stride;  // quads in the launch
  uint64_t gquad;
  uint32_t qid;         // quad within the block
  int capacity;
  int sp;

  // entry `pos`, written or read by one lane
  __device__ __forceinline__ void put(int pos, uint32_t ref, uint32_t aux) {
    if (pos < kQuadLds)
      lds_store(lds, pos * (kBlock / 4) + qid, make_uint2(ref, aux));
    else
      ovf_store(ovf, static_cast<uint64_t>(pos - kQuadLds) * ovf_stride + gquad, make_uint2(ref, aux));
  }
  __device__ __forceinline__ uint2 get(int pos) const {
    if (pos < kQuadLds) return lds_load(lds, pos * (kBlock / 4) + qid);
    return ovf_load(ovf, static_cast<uint64_t>(pos - kQuadLds) * ovf_stride + gquad);
  }
  // the same entry pushed / popped by all four lanes (binary traversal run by a quad)
  __device__ __forceinline__ bool push(uint32_t ref, uint32_t aux) {
    if (sp >= capacity) return false;
    put(sp++, ref, aux);
    return true;
  }
  __device__ __forceinline__ uint2 pop() { return get(--sp); }
};

// Quad exchanges as DPP quad permutations (VALU operand modifiers) rather than LDS permutes.
// quad_perm selects, for each lane of a group of four, the source lane within the group.
template <int kCtrl>
__device__ __forceinline__ uint32_t quad_dpp(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, 0xF, 0xF, false));
}
template <int kCtrl>
__device__ __forceinline__ float quad_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}
template <int kLane, typename T>
__device__ __forceinline__ T quad_bcast(T v) {  // value of lane kLane of the quad
  return quad_dpp<kLane | (kLane << 2) | (kLane << 4) | (kLane << 6)>(v);
}
template <typename T>
__device__ __forceinline__ T quad_xor1(T v) {  // lanes 1 0 3 2
  return quad_dpp<0xB1>(v);
}
template <typename T>
__device__ __forceinline__ T quad_xor2(T v) {  // lanes 2 3 0 1
  return quad_dpp<0x4E>(v);
}

// One node step of a quad: lane q holds child q of the node, `in` saying whether it is entered,
// with its key (entry t or box score, a lower bound) and ref. The entered children are ranked by
// their key (the order only steers the search, any order gives the same result): the key's bits,
// made monotonic as an unsigned integer, with the lane in the two low bits so that all four are
// distinct. The rank-0 child is returned to every lane (kDone if none is entered) and each other
// entered child is written by its own lane onto the quad stack, farthest deepest, so that pops come
// nearest first; the stack keeps the exact key. Sets `overflow` (and returns kDone) when the stack
// cannot take them.
__device__ __forceinline__ uint32_t quad_children(QuadStack& st, uint32_t lane, bool in, float key, uint32_t cref,
                                                  bool& overflow) {
  const uint32_t q = lane & 3u;
  const uint32_t kb = __float_as_uint(key);
  const uint32_t mono = kb ^ ((kb >> 31) ? 0xFFFFFFFFu : 0x80000000u);  // float order as unsigned order
  const uint32_t kq = in ? ((mono & ~3u) | q) : 0xFFFFFFFFu;           // not entered: ranks last
  const uint32_t k1 = quad_xor1(kq), k2 = quad_xor2(kq), k3 = quad_dpp<0x1B>(kq);  // lanes q^1, q^2, q^3
  const int rank = static_cast<int>(k1 < kq) + static_cast<int>(k2 < kq) + static_cast<int>(k3 < kq);
  const uint64_t entered = __ballot(in);
  const int nh = __popcll((entered >> ((lane & 63u) & ~3u)) & 0xFull);
  if (nh == 0) return kDone;
  uint32_t first = (in && rank == 0) ? cref : 0u;  // child refs are never 0 (node 0 is a root)
  first |= quad_xor1(first);
  first |= quad_xor2(first);
  if (st.sp + nh - 1 > st.capacity) {
    overflow = true;
    return kDone;
  }
  if (in && rank > 0) st.put(st.sp + nh - 1 - rank, cref, __float_as_uint(key));
  st.sp += nh - 1;
  return first;
}

// Whether another lane's exact result (o, ow) beats this lane's (b, w): a smaller value, or a tie
// (`tie`, a symmetric test of the two) that the reference meets first. Antisymmetric, so the lanes
// of a quad agree.
template <typename FirstOf>
__device__ __forceinline__ bool beats(ff o, uint32_t ow, ff b, uint32_t w, bool tie, FirstOf first_of) {
  if (ow == kNoHit || ow == w) return false;
  if (w == kNoHit) return true;
  if (tie) return first_of(ow, w);
  return ff_sign(ff_sub(o, b)) < 0;
}

template <bool kLo>
__device__ __forceinline__ FirstHit trace_first4(const MeshView& M, QuadStack& st, uint32_t lane, f3 o, f3 d, f3 ol,
                                                 f3 dl, uint32_t side_mask, Window win) {
  const uint32_t q = lane & 3u;
  const RayBox4 rb = ray_box4<kLo>(o, d, ol, dl, *M.coord_max);
  const float near_dist = win.near_d, far_dist = win.far_d;
  const float dlen = sqrtf(dot(d, d));
  const RayBand band = ray_band<kLo>(ol, dl, fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
  const TieTree T = tie_tree(M);
  float U = kInf;  // quad-wide upper bound of the best hit's t (after each leaf)
  bool overflow = false;
  Cands cl;
  cl.lane = lane;
  cl.n = 0;

  for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
    uint32_t ref = M.roots4[r];
    st.sp = 0;
    while (true) {
      // ---- internal nodes (quad-uniform) ----
      while (!(ref & kRefLeaf) && ref != kDone) {
        const float4 c0 = M.nodes4[8ull * ref + 2 * q];
        const float4 c1 = M.nodes4[8ull * ref + 2 * q + 1];
        const uint32_t cref = __float_as_uint(c1.z);
        float t;
        const bool box = slab4(rb, c0, c1, near_dist, far_dist, t);
        ref = quad_children(st, lane, (cref != kDone) & box & (t <= U), t, cref, overflow);
        if (overflow) break;
        if (ref == kDone) {
          while (st.sp > 0) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) <= U) {
              ref = e.x;
              break;
            }
          }
        }
      }
      // ---- leaves: the quad splits the triangles ----
      while (ref & kRefLeaf) {
        uint32_t off, cnt;
        leaf_range(ref, M.bigleaf, off, cnt);
        float Ul = U;
        for (uint32_t s = off + q; s < off + cnt; s += 4) {
          const Tri tr = load_tri(M.tris, s);
          const uint32_t mask = side_mask ? side_mask : tr.mask();
          float t = 0.0f, et = 0.0f;
          const int res = hit_fast<kLo>(o, d, dlen, band, tr, mask, near_dist, far_dist, t, et);
          if (res == kMiss) continue;
          if (res == kHit) {
            if (t - et > Ul) continue;  // certainly behind a known hit
            Ul = fminf(Ul, t + et);
          }
          if (cl.n == kCand && !cl.compact(Ul)) {
            const Winner r = resolve_ray<kLo, true>(cl, Ul, o, d, ol, dl, M.tris, T, side_mask, win);
            const uint32_t w = r.w;
            cl.n = 0;
            if (w != kNoHit) {
              const float tw = ff_to_float(r.v);
              Ul = fminf(Ul, tw * (1.0f + 4.0f * kU));
              cl.put(0, w, tw * (1.0f - 4.0f * kU));
              cl.n = 1;
            }
          }
          cl.put(cl.n++, s, res == kHit ? t - et : -kInf);
        }
        Ul = fminf(Ul, quad_xor1(Ul));
        U = fminf(Ul, quad_xor2(Ul));
        ref = kDone;
        while (st.sp > 0) {
          const uint2 e = st.pop();
          if (__uint_as_float(e.y) <= U) {
            ref = e.x;
            break;
          }
        }
      }
      if (ref == kDone) break;
    }
  }

  FirstHit h = no_hit();
  if (overflow) {
    h.face = kOverflow;
    return h;
  }
  Winner r = {kNoHit, {kInf, 0.0f}, 0.0f};
  if (cl.n > 0) r = resolve_ray<kLo, true>(cl, U, o, d, ol, dl, M.tris, T, side_mask, win);
  ff bt = r.w == kNoHit ? ff{kInf, 0.0f} : r.v;
  uint32_t w = r.w;
  // the quad's winner: smallest exact t, ties to the reference's first
  for (int m = 1; m <= 2; m <<= 1) {
    const float oh = m == 1 ? quad_xor1(bt.hi) : quad_xor2(bt.hi);
    const float olo = m == 1 ? quad_xor1(bt.lo) : quad_xor2(bt.lo);
    const uint32_t ow = m == 1 ? quad_xor1(w) : quad_xor2(w);
    const ff ot = {oh, olo};
    if (beats(ot, ow, bt, w, near_tie(ot, bt),
              [&](uint32_t a, uint32_t b) { return ray_first_of(T, d.x, d.y, d.z, a, b); })) {
      bt = ot;
      w = ow;
    }
  }
  if (w != kNoHit) h = finish_hit<kLo>(M.tris, w, o, d, ol, dl);
  return h;
}

// The triple at index i of a packed float array (zero when the array is NULL).
__device__ __forceinline__ f3 load3(const float* p, uint64_t i) {
  return p ? mk(p[3 * i + 0], p[3 * i + 1], p[3 * i + 2]) : mk(0.0f, 0.0f, 0.0f);
}

template <bool kLo>
__global__ __launch_bounds__(kBlock, 4) void k_raycast_first4(RaycastFirstArgs A) {
  __shared__ uint2 qstack[kQuadLds * (kBlock / 4)];
  const uint32_t lane = threadIdx.x;
  const uint32_t qid = lane >> 2;
  const uint64_t gquad = static_cast<uint64_t>(blockIdx.x) * (kBlock / 4) + qid;
  const uint64_t nquads = static_cast<uint64_t>(gridDim.x) * (kBlock / 4);
  QuadStack st;
  st.lds = qstack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = nquads;
  st.gquad = gquad;
  st.qid = qid;
  st.capacity = kQuadLds + static_cast<int>(A.stack.ovf_entries);
  for (uint64_t visit = gquad; visit < A.n; visit += nquads) {
    const uint64_t ray = A.order ? A.order[visit] : visit;
    if (ray >= A.n) continue;
    const f3 o = mk(A.origins[3 * ray + 0], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
    const f3 d = mk(A.directions[3 * ray + 0], A.directions[3 * ray + 1], A.directions[3 * ray + 2]);
    const f3 ol = kLo ? load3(A.origins_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const f3 dl = kLo ? load3(A.directions_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const Window win = window_of(A.near_dist, A.near_lo, A.far_dist, A.far_lo);
    const bool literal = literal_ray<kLo>(o, d, ol, dl, *A.mesh.coord_max);
    const FirstHit h = literal ? trace_first<kLo>(A.mesh, st, lane, o, d, ol, dl, A.side_mask, win)
                               : trace_first4<kLo>(A.mesh, st, lane, o, d, ol, dl, A.side_mask, win);
    if ((lane & 3u) == 0) {
      A.out_face[ray] = h.face;
      if (A.out_dist) A.out_dist[ray] = h.dist;
      if (A.out_point) {
        A.out_point[3 * ray + 0] = h.p.x;
        A.out_point[3 * ray + 1] = h.p.y;
        A.out_point[3 * ray + 2] = h.p.z;
      }
      if (A.out_bary) {
        A.out_bary[3 * ray + 0] = h.bary.x;
        A.out_bary[3 * ray + 1] = h.bary.y;
        A.out_bary[3 * ray + 2] = h.bary.z;
      }
    }
  }
}

// (kBlock, 4): at most 128 VGPRs, so four waves fit on a SIMD (the LDS use allows four blocks per CU)
template <bool kLo>
__global__ __launch_bounds__(kBlock, 4) void k_raycast_first(RaycastFirstArgs A) {
  const uint32_t lane = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + lane;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kBlock;

  Stack st;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = lane;
  st.capacity = kStackLds + static_cast<int>(A.stack.ovf_entries);

  for (uint64_t visit = gtid; visit < A.n; visit += gstride) {
    const uint64_t ray = A.order ? A.order[visit] : visit;
    if (ray >= A.n) continue;  // not a permutation of 0..n-1: skip rather than read out of bounds
    const f3 o = mk(A.origins[3 * ray + 0], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
    const f3 d = mk(A.directions[3 * ray + 0], A.directions[3 * ray + 1], A.directions[3 * ray + 2]);
    const f3 ol = kLo ? load3(A.origins_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const f3 dl = kLo ? load3(A.directions_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const FirstHit h = trace_first<kLo>(A.mesh, st, lane, o, d, ol, dl, A.side_mask,
                                        window_of(A.near_dist, A.near_lo, A.far_dist, A.far_lo));
    A.out_face[ray] = h.face;
    if (A.out_dist) A.out_dist[ray] = h.dist;
    if (A.out_point) {
      A.out_point[3 * ray + 0] = h.p.x;
      A.out_point[3 * ray + 1] = h.p.y;
      A.out_point[3 * ray + 2] = h.p.z;
    }
    if (A.out_bary) {
      A.out_bary[3 * ray + 0] = h.bary.x;
      A.out_bary[3 * ray + 1] = h.bary.y;
      A.out_bary[3 * ray + 2] = h.bary.z;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Instances (BatchedMesh). MeshBVH.raycastObject3D (src/core/MeshBVH.js:447-494): the world ray
// is brought into the instance's frame with three's Ray.applyMatrix4(matrixWorld^-1) (origin by
// Vector3.applyMatrix4, direction by Vector3.transformDirection, i.e. normalised), near/far are
// divided by |local direction * world scale|, and the hit point goes back with
// Vector3.applyMatrix4(matrixWorld), its distance measured from the world origin
// (convertRaycastIntersect, src/utils/GeometryRayIntersectUtilities.js:3-17). Matrices are
// column-major 4x4 (three's Matrix4.elements). The arithmetic is fixed (no contraction) so that
// every kernel computing an instance ray gets the same bits.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ f3 apply_matrix4(const float* e, f3 v) {
  MBVH_NO_CONTRACT
  const float w = 1.0f / (e[3] * v.x + e[7] * v.y + e[11] * v.z + e[15]);
  return mk((e[0] * v.x + e[4] * v.y + e[8] * v.z + e[12]) * w, (e[1] * v.x + e[5] * v.y + e[9] * v.z + e[13]) * w,
            (e[2] * v.x + e[6] * v.y + e[10] * v.z + e[14]) * w);
}
__device__ __forceinline__ f3 transform_direction(const float* e, f3 v) {
  MBVH_NO_CONTRACT
  const f3 r = mk(e[0] * v.x + e[4] * v.y + e[8] * v.z, e[1] * v.x + e[5] * v.y + e[9] * v.z,
                  e[2] * v.x + e[6] * v.y + e[10] * v.z);
  const float len = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z);
  const float inv = 1.0f / (len != 0.0f ? len : 1.0f);
  return mk(r.x * inv, r.y * inv, r.z * inv);
}
// |local direction * column lengths of matrixWorld| (raycastObject3D's scaleFactor)
__device__ __forceinline__ float instance_scale(const float* e, f3 dl) {
  MBVH_NO_CONTRACT
  const float sx = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const float sy = sqrtf(e[4] * e[4] + e[5] * e[5] + e[6] * e[6]);
  const float sz = sqrtf(e[8] * e[8] + e[9] * e[9] + e[10] * e[10]);
  const f3 v = mk(dl.x * sx, dl.y * sy, dl.z * sz);
  return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
}
__device__ __forceinline__ float world_distance(f3 pw, f3 o) {
  MBVH_NO_CONTRACT
  const f3 dv = sub(pw, o);
  return sqrtf(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
}

__global__ __launch_bounds__(kBlock) void k_instance_rays(InstanceRayArgs A) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= A.n) return;
  const uint32_t inst = A.ray_instance[i];
  const float qnan = __int_as_float(0x7fc00000);
  f3 o = mk(qnan, qnan, qnan), d = o;
  if (inst < A.n_instances) {
    const float* inv = A.inverse + 16ull * inst;
    o = apply_matrix4(inv, mk(A.origins[3 * i], A.origins[3 * i + 1], A.origins[3 * i + 2]));
    d = transform_direction(inv, mk(A.directions[3 * i], A.directions[3 * i + 1], A.directions[3 * i + 2]));
  }
  A.out_origins[3 * i] = o.x;
  A.out_origins[3 * i + 1] = o.y;
  A.out_origins[3 * i + 2] = o.z;
  A.out_directions[3 * i] = d.x;
  A.out_directions[3 * i + 1] = d.y;
  A.out_directions[3 * i + 2] = d.z;
}

__global__ __launch_bounds__(kBlock) void k_instance_hits(InstanceHitArgs A) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= A.n) return;
  const uint32_t inst = A.ray_instance[i];
  if (A.face[i] >= kOverflow || inst >= A.n_instances) return;
  const f3 pw = apply_matrix4(A.world + 16ull * inst, mk(A.point[3 * i], A.point[3 * i + 1], A.point[3 * i + 2]));
  A.point[3 * i] = pw.x;
  A.point[3 * i + 1] = pw.y;
  A.point[3 * i + 2] = pw.z;
  A.distance[i] = world_distance(pw, mk(A.origins[3 * i], A.origins[3 * i + 1], A.origins[3 * i + 2]));
}

__global__ void k_batch_reset(uint32_t* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// Batched first hit over many meshes: a persistent grid pulls tiles of `tile` rays (in the given
// visit order, grouped by mesh) from an atomic counter until none are left; each lane brings its
// ray into its instance's frame, traces it against that instance's mesh and converts the hit to
// world space.
// kLo (local mode, A.inverse == NULL): rays already in their mesh's frame as hi + lo pairs, ray_instance
// naming the mesh, a window per ray, results left in the mesh's frame.
template <int kL, bool kLo>
__global__ __launch_bounds__(kBlock, 4) void k_batch_raycast_first(BatchArgs A) {
  __shared__ uint2 lds_stack[kL * kBlock];
  const uint32_t tid = threadIdx.x, lane64 = tid & 63u;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + tid;
  StackL<kL> st;
  st.lds = lds_stack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = tid;
  st.capacity = kL + static_cast<int>(A.stack.ovf_entries);
  const float qnan = __int_as_float(0x7fc00000);
  while (true) {
    uint32_t t = 0;
    if (lane64 == 0) t = atomicAdd(A.counter, 1u);
    t = __shfl(t, 0, 64);
    if (t >= A.n_tiles) break;  // every wave leaves once the counter passes the last tile
    const uint64_t base = static_cast<uint64_t>(t) * A.tile;
    for (uint32_t k = lane64; k < A.tile; k += 64) {
      const uint64_t visit = base + k;
      if (visit >= A.n) break;
      const uint64_t ray = A.order ? A.order[visit] : visit;
      if (ray >= A.n) continue;
      const uint32_t inst = A.ray_instance[ray];
      const uint32_t geom = kLo ? inst : (inst < A.n_instances ? A.instance_geometry[inst] : 0xFFFFFFFFu);
      FirstHit h;
      h.face = kNoHit;
      h.dist = kInf;
      h.p = mk(qnan, qnan, qnan);
      h.bary = h.p;
      if (kLo && geom < A.n_meshes) {
        const float* wn = A.windows + 4 * ray;
        h = trace_first<true>(A.meshes[geom], st, tid, load3(A.origins, ray), load3(A.directions, ray),
                              load3(A.origins_lo, ray), load3(A.directions_lo, ray), A.side_mask,
                              window_of(wn[0], wn[1], wn[2], wn[3]));
      } else if (!kLo && geom < A.n_meshes) {
        const f3 ow = mk(A.origins[3 * ray], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
        const float* inv = A.inverse + 16ull * inst;
        const f3 o = apply_matrix4(inv, ow);
        const f3 d = transform_direction(inv, mk(A.directions[3 * ray], A.directions[3 * ray + 1], A.directions[3 * ray + 2]));
        const float sc = instance_scale(A.world + 16ull * inst, d);
        const f3 z = mk(0.0f, 0.0f, 0.0f);
        h = trace_first<false>(A.meshes[geom], st, tid, o, d, z, z, A.side_mask,
                               window_of(A.near_dist / sc, 0.0f, A.far_dist / sc, 0.0f));
        if (h.face < kOverflow) {
          // matrices and the world origin are re-read here rather than kept live across the trace
          const f3 ow2 = mk(A.origins[3 * ray], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
          h.p = apply_matrix4(A.world + 16ull * inst, h.p);
          h.dist = world_distance(h.p, ow2);
        }
      }
      A.out_face[ray] = h.face;
      if (A.out_dist) A.out_dist[ray] = h.dist;
      if (A.out_point) {
        A.out_point[3 * ray + 0] = h.p.x;
        A.out_point[3 * ray + 1] = h.p.y;
        A.out_point[3 * ray + 2] = h.p.z;
      }
      if (A.out_bary) {
        A.out_bary[3 * ray + 0] = h.bary.x;
        A.out_bary[3 * ray + 1] = h.bary.y;
        A.out_bary[3 * ray + 2] = h.bary.z;
      }
    }
  }
}

// Same work queue with four lanes per ray (quads, the 4-wide trees): a wave traces 16 rays at a
// time; rays with a zero direction component take the binary traversal (see trace_first4).
template <bool kLo>
__global__ __launch_bounds__(kBlock, 4) void k_batch_raycast_first4(BatchArgs A) {
  __shared__ uint2 qstack[kQuadLds * (kBlock / 4)];
  const uint32_t tid = threadIdx.x, lane64 = tid & 63u, q = tid & 3u, wq = lane64 >> 2;
  const uint32_t qid = tid >> 2;
  QuadStack st;
  st.lds = qstack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = static_cast<uint64_t>(gridDim.x) * (kBlock / 4);
  st.gquad = static_cast<uint64_t>(blockIdx.x) * (kBlock / 4) + qid;
  st.qid = qid;
  st.capacity = kQuadLds + static_cast<int>(A.stack.ovf_entries);
  const float qnan = __int_as_float(0x7fc00000);
  while (true) {
    uint32_t t = 0;
    if (lane64 == 0) t = atomicAdd(A.counter, 1u);
    t = __shfl(t, 0, 64);
    if (t >= A.n_tiles) break;  // every wave leaves once the counter passes the last tile
    const uint64_t base = static_cast<uint64_t>(t) * A.tile;
    for (uint32_t k = wq; k < A.tile; k += 16) {
      const uint64_t visit = base + k;
      if (visit >= A.n) break;
      const uint64_t ray = A.order ? A.order[visit] : visit;
      if (ray >= A.n) continue;
      const uint32_t inst = A.ray_instance[ray];
      const uint32_t geom = kLo ? inst : (inst < A.n_instances ? A.instance_geometry[inst] : 0xFFFFFFFFu);
      FirstHit h;
      h.face = kNoHit;
      h.dist = kInf;
      h.p = mk(qnan, qnan, qnan);
      h.bary = h.p;
      if (kLo && geom < A.n_meshes) {
        const MeshView& M = A.meshes[geom];
        const f3 o = load3(A.origins, ray), d = load3(A.directions, ray);
        const f3 ol = load3(A.origins_lo, ray), dl = load3(A.directions_lo, ray);
        const float* wn = A.windows + 4 * ray;
        const Window win = window_of(wn[0], wn[1], wn[2], wn[3]);
        h = literal_ray<true>(o, d, ol, dl, *M.coord_max) ? trace_first<true>(M, st, tid, o, d, ol, dl, A.side_mask, win)
                                                         : trace_first4<true>(M, st, tid, o, d, ol, dl, A.side_mask, win);
      } else if (!kLo && geom < A.n_meshes) {
        const f3 ow = mk(A.origins[3 * ray], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
        const float* inv = A.inverse + 16ull * inst;
        const f3 o = apply_matrix4(inv, ow);
        const f3 d = transform_direction(inv, mk(A.directions[3 * ray], A.directions[3 * ray + 1], A.directions[3 * ray + 2]));
        const float sc = instance_scale(A.world + 16ull * inst, d);
        const MeshView& M = A.meshes[geom];
        const f3 z = mk(0.0f, 0.0f, 0.0f);
        const Window win = window_of(A.near_dist / sc, 0.0f, A.far_dist / sc, 0.0f);
        const bool literal = literal_ray<false>(o, d, z, z, *M.coord_max);
        h = literal ? trace_first<false>(M, st, tid, o, d, z, z, A.side_mask, win)
                    : trace_first4<false>(M, st, tid, o, d, z, z, A.side_mask, win);
        if (h.face < kOverflow) {
          const f3 ow2 = mk(A.origins[3 * ray], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
          h.p = apply_matrix4(A.world + 16ull * inst, h.p);
          h.dist = world_distance(h.p, ow2);
        }
      }
      if (q == 0) {
        A.out_face[ray] = h.face;
        if (A.out_dist) A.out_dist[ray] = h.dist;
        if (A.out_point) {
          A.out_point[3 * ray + 0] = h.p.x;
          A.out_point[3 * ray + 1] = h.p.y;
          A.out_point[3 * ray + 2] = h.p.z;
        }
        if (A.out_bary) {
          A.out_bary[3 * ray + 0] = h.bary.x;
          A.out_bary[3 * ray + 1] = h.bary.y;
          A.out_bary[3 * ray + 2] = h.bary.z;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// All hits. _raycast (src/core/cast/raycast.template.js:16-53): left child before right child,
// both bounds-tested against [near, far], no distance pruning; inside a leaf every triangle in
// order (intersectTris, iterationUtils.template.js:4-22). The count pass and the fill pass walk
// the identical order, so hit k of a ray lands at offsets[ray] + k in the reference's order.
// Every accepted hit is re-evaluated in double-float for its reported values.
// ------------------------------------------------------------------------------------------
template <bool kFill, bool kLo>
__global__ __launch_bounds__(kBlock) void k_raycast_all(RaycastAllArgs A) {
  const uint32_t lane = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + lane;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kBlock;
  const MeshView M = A.mesh;

  Stack st;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = lane;
  st.capacity = kStackLds + static_cast<int>(A.stack.ovf_entries);

  for (uint64_t ray = gtid; ray < A.n; ray += gstride) {
    const f3 o = mk(A.origins[3 * ray + 0], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
    const f3 d = mk(A.directions[3 * ray + 0], A.directions[3 * ray + 1], A.directions[3 * ray + 2]);
    const f3 ol = kLo ? load3(A.origins_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const f3 dl = kLo ? load3(A.directions_lo, ray) : mk(0.0f, 0.0f, 0.0f);
    const Window win = window_of(A.near_dist, A.near_lo, A.far_dist, A.far_lo);
    RayBox rb;
    rb.ox = o.x;
    rb.oy = o.y;
    rb.oz = o.z;
    rb.ix = 1.0f / d.x;
    rb.iy = 1.0f / d.y;
    rb.iz = 1.0f / d.z;
    rb.sx = rb.ix >= 0.0f;
    rb.sy = rb.iy >= 0.0f;
    rb.sz = rb.iz >= 0.0f;
    rb.wa = kLo ? slab_lo_widening(ol, rb.ix, rb.iy, rb.iz) : 0.0f;
    const float dlen = sqrtf(dot(d, d));
    const RayBand band = ray_band<kLo>(ol, dl, fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
    uint32_t count = 0;
    uint64_t base = 0;
    if (kFill) base = A.offsets[ray];
    bool overflow = false;

    for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
      uint32_t ref = M.roots[r];
      st.sp = 0;
      while (true) {
        while (!(ref & kRefLeaf) && ref != kDone) {
          const WideNode nd = M.nodes[ref];
          float tl, tr;
          const bool hl = slab(rb, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y, win.near_d, win.far_d, tl);
          const bool hr = slab(rb, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w, win.near_d, win.far_d, tr);
          if (hl && hr) {
            if (!st.push(nd.meta.y, 0u)) {
              overflow = true;
              ref = kDone;
              break;
            }
            ref = nd.meta.x;
          } else if (hl) {
            ref = nd.meta.x;
          } else if (hr) {
            ref = nd.meta.y;
          } else {
            ref = st.sp > 0 ? st.pop().x : kDone;
          }
        }
        while (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          for (uint32_t s = off; s < off + cnt; s++) {
            const Tri tr = load_tri(M.tris, s);
            const uint32_t mask = A.side_mask ? A.side_mask : tr.mask();
            float t = 0.0f, et = 0.0f;
            int res = hit_fast<kLo>(o, d, dlen, band, tr, mask, A.near_dist, A.far_dist, t, et);
            if (res == kUnsure) {
              ff te;
              res = hit_exact<kLo>(o, d, ol, dl, tr.a(), tr.b(), tr.c(), mask, win, te) ? kHit : kMiss;
            }
            if (res != kHit) continue;
            if (kFill) {
              const uint64_t k = base + count;
              if (k < A.capacity) {
                const FirstHit h = finish_hit<kLo>(M.tris, s, o, d, ol, dl);
                A.out_face[k] = h.face;
                if (A.out_dist) A.out_dist[k] = h.dist;
                if (A.out_point) {
                  A.out_point[3 * k + 0] = h.p.x;
                  A.out_point[3 * k + 1] = h.p.y;
                  A.out_point[3 * k + 2] = h.p.z;
                }
                if (A.out_bary) {
                  A.out_bary[3 * k + 0] = h.bary.x;
                  A.out_bary[3 * k + 1] = h.bary.y;
                  A.out_bary[3 * k + 2] = h.bary.z;
                }
              }
            }
            count++;
          }
          ref = st.sp > 0 ? st.pop().x : kDone;
        }
        if (ref == kDone) break;
      }
    }
    if (!kFill) A.counts[ray] = overflow ? kOverflow : count;
  }
}

// ------------------------------------------------------------------------------------------
// Closest point on a triangle: Ericson, Real-Time Collision Detection §5.1.5 (the algorithm of
// three's Triangle.closestPointToPoint), Voronoi regions tested in the same order. This FP32
// version only feeds error-bounded comparisons (the decisions are made in double-float), so its
// divisions use the hardware reciprocal (1 ulp) and the face region clamps the three sub-areas at
// 0: every branch returns a convex combination of a, b, c with weights valid to a few ulps.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float rcp_approx(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c) {
  const f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const f3 bp = sub(p, b);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 * rcp_approx(d1 - d3);
    return mk(a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v);
  }
  const f3 cp = sub(p, c);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 * rcp_approx(d2 - d6);
    return mk(a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) * rcp_approx((d4 - d3) + (d5 - d6));
    const f3 bc = sub(c, b);
    return mk(b.x + bc.x * w, b.y + bc.y * w, b.z + bc.z * w);
  }
  const float ua = fmaxf(va, 0.0f), ub = fmaxf(vb, 0.0f), uc = fmaxf(vc, 0.0f);
  const float denom = rcp_approx(ua + ub + uc);  // all three 0: inf, then NaN, decided exactly
  const float v = ub * denom, w = uc * denom;
  return mk(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

// An upper bound of sqrt(x) for x >= 0 (hardware square root, 1 ulp, raised by 4 ulps).
__device__ __forceinline__ float sqrt_up(float x) { return __builtin_amdgcn_sqrtf(x) * (1.0f + 4.0f * kU); }

__device__ __forceinline__ float box_score(f3 q, float mnx, float mny, float mnz, float mxx,
                                           float mxy, float mxz) {
  // distance^2 from q to the box (three Vector3.clamp then distanceToSquared); the median of
  // (q, min, max) is the clamp for min <= max
  const float cx = __builtin_amdgcn_fmed3f(q.x, mnx, mxx);
  const float cy = __builtin_amdgcn_fmed3f(q.y, mny, mxy);
  const float cz = __builtin_amdgcn_fmed3f(q.z, mnz, mxz);
  const float dx = cx - q.x, dy = cy - q.y, dz = cz - q.z;
  return dx * dx + dy * dy + dz * dz;
}

// Exact resolution for the closest point: the smallest exact squared distance from q + ql; ties
// (point_tie) go to the reference's first (kRefOrder: by point_first_of, the candidates coming from
// four lanes; otherwise the list is in the reference's order and the earlier wins).
template <bool kLo, bool kRefOrder>
__device__ __noinline__ Winner resolve_point(Cands cl, float U, f3 q, f3 ql, const float4* tris, TieTree T) {
  uint32_t win = kNoHit;
  ff best_d2 = {kInf, 0.0f};
  float best_err = 0.0f;
  for (int i = 0; i < cl.n; i++) {
    const uint2 e = cl.get(i);
    if (__uint_as_float(e.y) > U) continue;
    const Tri tr = load_tri(tris, e.x);
    const ClosestFF cf = closest_on_triangle_ff_inline<kLo>(q, ql, tr.a(), tr.b(), tr.c());
    const ff d2 = cf.dist2;
    if (!(d2.hi == d2.hi)) continue;  // degenerate: never wins (reference: `NaN < best` is false)
    bool take = win == kNoHit;
    if (!take) {
      if (point_tie(d2, cf.err, best_d2, best_err))
        take = kRefOrder && e.x != win && point_first_of<kLo>(T, q, ql, e.x, win);
      else
        take = ff_sign(ff_sub(d2, best_d2)) < 0;
    }
    if (take) {
      win = e.x;
      best_d2 = d2;
      best_err = cf.err;
    }
  }
  return {win, best_d2, best_err};
}

// Exact decisions against a threshold given as hi + lo floats: whether the squared distance of
// triangle `tr` from q + ql is below t^2, and whether a box's squared distance is below t^2. The
// FP32 value x (within a relative error rel, plus e2 absolute for a triangle) decides when it is
// clear of the band; otherwise double-float does.
__device__ __forceinline__ ff thr_sq(float t, float tl) {
  const ff T = two_sum(t, tl);
  return ff_mul(T, T);
}
template <bool kLo>
__device__ __noinline__ bool tri_below(f3 q, f3 ql, const Tri& tr, ff t2) {
  const ff d2 = closest_on_triangle_ff<kLo>(q, ql, tr.a(), tr.b(), tr.c()).dist2;
  return d2.hi == d2.hi && ff_sign(ff_sub(d2, t2)) < 0;  // NaN (degenerate) is never below
}
template <bool kLo>
__device__ __noinline__ bool box_below(f3 q, f3 ql, float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                       ff t2) {
  return ff_sign(ff_sub(box_score_ff<kLo>(q, ql, mnx, mny, mnz, mxx, mxy, mxz), t2)) < 0;
}

// ------------------------------------------------------------------------------------------
// Closest point. Traversal of closestPointToPoint over shapecast
// (src/core/cast/closestPointToPoint.js:6-78, src/core/cast/shapecast.js:41-214): children
// ordered by squared box distance (right first only if strictly closer), a child is entered iff
// its score is below the best distance^2 (here its upper bound) and below maxThreshold^2 (the far
// child's score is re-checked against the best found meanwhile), root nodes are never
// bounds-tested, triangles update the best on strict <, and the first triangle closer than
// minThreshold ends the whole query (all roots).
// ------------------------------------------------------------------------------------------
// Closest point with four lanes per query on the 4-wide tree (used when minThreshold is 0: the
// reference's minThreshold exit returns the FIRST triangle closer than it in its own visit order,
// which only the binary traversal reproduces). Node step: lane q scores child q; the quad sorts
// the entered children by score, continues with the nearest and pushes the others (scores
// re-checked on pop). Leaf step: lane q tests triangles q, q+4, ...; the bound is shared after
// each leaf. Each lane resolves its candidates exactly; the quad keeps the smallest exact squared
// distance (exactly equal distances are ties).
// The closest point's reported values: distance, point and barycoord in double-float.
template <bool kLo>
__device__ __forceinline__ void finish_point(const float4* tris, uint32_t w, f3 q, f3 ql, uint32_t& face, float& dist,
                                             float& px, float& py, float& pz, float& bw, float& bu, float& bv) {
  const Tri tr = load_tri(tris, w);
  const ClosestFF cf = closest_on_triangle_ff<kLo>(q, ql, tr.a(), tr.b(), tr.c());
  face = tr.face();
  dist = ff_to_float(ff_sqrt(cf.dist2));
  px = ff_to_float(cf.point.x);
  py = ff_to_float(cf.point.y);
  pz = ff_to_float(cf.point.z);
  bw = ff_to_float(cf.bw);
  bu = ff_to_float(cf.bu);
  bv = ff_to_float(cf.bv);
}

// The absolute error bound of an FP32 point-to-triangle distance for the query q (+ ql): FP32
// roundings (8u of the magnitudes), and for a double query twice its distance to q (so that bounds
// measured from q also bound the distances from q + ql, and the boxes pruned from q stay pruned).
template <bool kLo>
__device__ __forceinline__ float query_error(f3 q, f3 ql, float coord_max) {
  float e = 8.0f * kU * (maxabs(q) + coord_max);
  if (kLo) e += 2.0f * (fabsf(ql.x) + fabsf(ql.y) + fabsf(ql.z)) * (1.0f + 4.0f * kU);
  return e;
}

// Whether a box (FP32 squared distance s from q) is entered under maxThreshold^2 (t2): the FP32
// value decides when it is clear of its error band, double-float otherwise. Double queries always
// decide in double-float (their FP32 score is measured from q, not from q + ql).
template <bool kLo>
__device__ __forceinline__ bool below_max(bool finite, float s, ff t2, f3 q, f3 ql, float mnx, float mny, float mnz,
                                          float mxx, float mxy, float mxz) {
  if (!finite) return true;
  if (!kLo) {
    if (s * (1.0f + 8.0f * kU) < t2.hi * (1.0f - 4.0f * kU)) return true;
    if (s * (1.0f - 8.0f * kU) > t2.hi * (1.0f + 4.0f * kU)) return false;
  }
  return box_below<kLo>(q, ql, mnx, mny, mnz, mxx, mxy, mxz, t2);
}

template <bool kLo>
__global__ __launch_bounds__(kBlock, 4) void k_closest_point4(ClosestPointArgs A) {
  __shared__ uint2 qstack[kQuadLds * (kBlock / 4)];
  const uint32_t lane = threadIdx.x, q4 = lane & 3u, wq = (lane & 63u) >> 2;
  const uint32_t qid = lane >> 2;
  const MeshView M = A.mesh;
  const TieTree T = tie_tree(M);
  QuadStack st;
  st.lds = qstack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = static_cast<uint64_t>(gridDim.x) * (kBlock / 4);
  st.gquad = static_cast<uint64_t>(blockIdx.x) * (kBlock / 4) + qid;
  st.qid = qid;
  st.capacity = kQuadLds + static_cast<int>(A.stack.ovf_entries);
  const bool max_finite = A.max_threshold < kInf;
  const ff max_t2 = max_finite ? thr_sq(A.max_threshold, A.max_lo) : ff{kInf, 0.0f};
  constexpr float kScoreSlack = 1.0f - 4.0f * kU;
  const float qnan = __int_as_float(0x7fc00000);
  const uint64_t nv = A.n_dev ? static_cast<uint64_t>(*A.n_dev) : A.n;  // queries to visit

  while (true) {
    uint32_t tile = 0;
    if ((lane & 63u) == 0) tile = atomicAdd(A.queue, 1u);
    tile = __shfl(tile, 0, 64);
    if (static_cast<uint64_t>(tile) * 16 >= nv) break;  // every wave reaches this exit
    const uint64_t visit = static_cast<uint64_t>(tile) * 16 + wq;
    if (visit >= nv) continue;
    const uint64_t qi = A.order ? A.order[visit] : visit;
    if (qi >= A.n) continue;
    const f3 q = mk(A.points[3 * qi + 0], A.points[3 * qi + 1], A.points[3 * qi + 2]);
    const f3 ql = kLo ? load3(A.points_lo, qi) : mk(0.0f, 0.0f, 0.0f);
    const float e_abs = query_error<kLo>(q, ql, *M.coord_max);
    float U = kInf;  // quad-wide upper bound of the best squared distance (after each leaf)
    bool overflow = false;
    Cands cl;
    cl.lane = lane;
    cl.n = 0;
    for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
      uint32_t ref = M.roots4[r];
      st.sp = 0;
      while (true) {
        while (!(ref & kRefLeaf) && ref != kDone) {
          const float4 c0 = M.nodes4[8ull * ref + 2 * q4];
          const float4 c1 = M.nodes4[8ull * ref + 2 * q4 + 1];
          const uint32_t cref = __float_as_uint(c1.z);
          const float raw = box_score(q, c0.x, c0.y, c0.z, c0.w, c1.x, c1.y);
          const float sc = raw * kScoreSlack;
          bool in = (cref != kDone) & (sc <= U);
          if (in && max_finite) in = below_max<kLo>(true, raw, max_t2, q, ql, c0.x, c0.y, c0.z, c0.w, c1.x, c1.y);
          ref = quad_children(st, lane, in, sc, cref, overflow);
          if (overflow) break;
          if (ref == kDone) {
            while (st.sp > 0) {
              const uint2 e = st.pop();
              if (__uint_as_float(e.y) <= U) {  // maxThreshold was decided when it was pushed
                ref = e.x;
                break;
              }
            }
          }
        }
        while (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          float Ul = U;
          for (uint32_t s = off + q4; s < off + cnt; s += 4) {
            const Tri tr = load_tri(M.tris, s);
            // the triangle's own box first: its squared distance (clamp exact, three roundings per
            // axis and two in the sum, covered by 8 ulps) bounds the triangle's from below, so a
            // box beyond the upper bound Ul rules the triangle out without the full test
            const float sb = box_score(q, fminf(fminf(tr.p0.x, tr.p1.x), tr.p2.x), fminf(fminf(tr.p0.y, tr.p1.y), tr.p2.y),
                                       fminf(fminf(tr.p0.z, tr.p1.z), tr.p2.z), fmaxf(fmaxf(tr.p0.x, tr.p1.x), tr.p2.x),
                                       fmaxf(fmaxf(tr.p0.y, tr.p1.y), tr.p2.y), fmaxf(fmaxf(tr.p0.z, tr.p1.z), tr.p2.z));
            if (sb * (1.0f - 8.0f * kU) > Ul) continue;
            const f3 cp = closest_on_triangle(q, tr.a(), tr.b(), tr.c());
            const f3 dv = sub(q, cp);
            const float d2 = dot(dv, dv);
            float lb = -kInf;  // FP32 failed (degenerate triangle, NaN): decide exactly at the end
            if (d2 == d2) {
              const float e2 = (2.0f * sqrt_up(d2) + e_abs) * e_abs + 4.0f * kU * d2;
              lb = d2 - e2;
              if (lb > Ul) continue;
              // every candidate that may tie stays listed: ties go to the reference's first
              Ul = fminf(Ul, d2 + e2);
            }
            if (cl.n == kCand && !cl.compact(Ul)) {
              const Winner r = resolve_point<kLo, true>(cl, Ul, q, ql, M.tris, T);
              const uint32_t w = r.w;
              cl.n = 0;
              if (w != kNoHit) {
                const float dw = ff_to_float(r.v);
                Ul = fminf(Ul, dw * (1.0f + 4.0f * kU));
                cl.put(0, w, dw * (1.0f - 4.0f * kU));
                cl.n = 1;
              }
            }
            cl.put(cl.n++, s, lb);
          }
          Ul = fminf(Ul, quad_xor1(Ul));
          U = fminf(Ul, quad_xor2(Ul));
          ref = kDone;
          while (st.sp > 0) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) <= U) {
              ref = e.x;
              break;
            }
          }
        }
        if (ref == kDone) break;
      }
    }
    Winner r = {kNoHit, {kInf, 0.0f}, 0.0f};
    if (!overflow && cl.n > 0) r = resolve_point<kLo, true>(cl, U, q, ql, M.tris, T);
    ff bd = r.w == kNoHit ? ff{kInf, 0.0f} : r.v;
    float be = r.w == kNoHit ? 0.0f : r.err;
    uint32_t w = r.w;
    // the quad's winner: smallest exact squared distance, ties to the reference's first
    for (int m = 1; m <= 2; m <<= 1) {
      const float oh = m == 1 ? quad_xor1(bd.hi) : quad_xor2(bd.hi);
      const float olo = m == 1 ? quad_xor1(bd.lo) : quad_xor2(bd.lo);
      const float oe = m == 1 ? quad_xor1(be) : quad_xor2(be);
      const uint32_t ow = m == 1 ? quad_xor1(w) : quad_xor2(w);
      const ff od = {oh, olo};
      if (beats(od, ow, bd, w, point_tie(od, oe, bd, be),
                [&](uint32_t a, uint32_t b) { return point_first_of<kLo>(T, q, ql, a, b); })) {
        bd = od;
        be = oe;
        w = ow;
      }
    }
    uint32_t out_face = overflow ? kOverflow : kNoHit;
    float out_dist = kInf;
    float px = qnan, py = qnan, pz = qnan, bw = qnan, bu = qnan, bv = qnan;
    if (!overflow && w != kNoHit && q4 == 0) finish_point<kLo>(M.tris, w, q, ql, out_face, out_dist, px, py, pz, bw, bu, bv);
    if (q4 == 0) {
      A.out_face[qi] = out_face;
      if (A.out_dist) A.out_dist[qi] = out_dist;
      if (A.out_point) {
        A.out_point[3 * qi + 0] = px;
        A.out_point[3 * qi + 1] = py;
        A.out_point[3 * qi + 2] = pz;
      }
      if (A.out_bary) {
        A.out_bary[3 * qi + 0] = bw;
        A.out_bary[3 * qi + 1] = bu;
        A.out_bary[3 * qi + 2] = bv;
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Closest point for coherent queries (grid samples, Morton-ordered point sets): a packet of 64
// queries, one per lane of a wave, walks the 4-wide tree together. A child is entered when any
// lane's query may need it (its squared box distance under that lane's upper bound and under
// maxThreshold^2); node records and triangles are read at wave-uniform addresses (scalar loads),
// so the node and leaf steps run with every lane busy instead of one quad in three. Each lane keeps
// its own bound and candidate list and resolves them exactly as the quad kernel does, so the
// results are bit for bit those of k_closest_point4: the packet's order only steers the search.
// Children are ordered by the scores of one lane near the middle of the packet. The stack (the
// children's own 32-byte records, re-scored per lane on pop) lives in LDS; the host uses this
// kernel only for trees whose 4-wide stack fits it.
// ------------------------------------------------------------------------------------------
constexpr int kPacketStack = static_cast<int>(kPacketStackEntries);
constexpr uint32_t kPacketRedoMark = 0xFFFFFFFFu;   // cand_meta: the query went to the redo list
constexpr uint32_t kPacketStackMark = 0xFFFFFFFEu;  // cand_meta: the packet's stack overflowed
constexpr uint32_t kPoolEnd = 0xFFFFFFFFu;          // end of a query's chain in the pool

__device__ __forceinline__ float4 pick4(uint32_t k, float4 a, float4 b, float4 c, float4 d) {
  return k == 0 ? a : (k == 1 ? b : (k == 2 ? c : d));
}

typedef __attribute__((address_space(4))) const float ConstF;  // read-only for the launch: scalar loads
__device__ __forceinline__ float4 cload4(ConstF* p) { return make_float4(p[0], p[1], p[2], p[3]); }

template <bool kLo>
__global__ __launch_bounds__(kBlock, 8) void k_closest_point_packet(ClosestPointArgs A) {
  __shared__ float4 pstack[(kBlock / 64) * kPacketStack * 2];
  const uint32_t lane = threadIdx.x, wl = lane & 63u;
  float4* const stk = pstack + (lane >> 6) * kPacketStack * 2;
  const MeshView M = A.mesh;
  ConstF* const nodes4 = (ConstF*)(M.nodes4);
  ConstF* const tris = (ConstF*)(M.tris);
  const bool max_finite = A.max_threshold < kInf;
  const ff max_t2 = max_finite ? thr_sq(A.max_threshold, A.max_lo) : ff{kInf, 0.0f};
  constexpr float kScoreSlack = 1.0f - 4.0f * kU;
  const float qnan = __int_as_float(0x7fc00000);

  while (true) {
    uint32_t tile = 0;
    if (wl == 0) tile = atomicAdd(A.queue, 1u);
    tile = __builtin_amdgcn_readfirstlane(__shfl(tile, 0, 64));
    if (tile >= A.tiles || (A.tile0 + tile) * 64 >= A.n) break;  // every wave reaches this exit
    const uint64_t visit = (A.tile0 + tile) * 64 + wl;
    uint64_t qi = visit;
    if (visit < A.n && A.order) qi = A.order[visit];
    const bool active = visit < A.n && qi < A.n;
    f3 q = mk(qnan, qnan, qnan), ql = mk(0.0f, 0.0f, 0.0f);
    if (active) {
      q = mk(A.points[3 * qi + 0], A.points[3 * qi + 1], A.points[3 * qi + 2]);
      if (kLo) ql = load3(A.points_lo, qi);
    }
    const float e_abs = query_error<kLo>(q, ql, *M.coord_max);
    float U = active ? kInf : -1.0f;  // this lane's upper bound of its best squared distance
    const uint64_t act = __ballot(active);
    if (act == 0) continue;
    // the lane whose scores order the children: lane 32 if its query is live, else the first live one
    const uint32_t leader = (act >> 32) & 1u ? 32u : static_cast<uint32_t>(__builtin_ctzll(act));
    // this query's candidates: a kCand-entry list in LDS, compacted against the bound when full; a
    // candidate that finds it full of live ones goes onto the query's chain in the global pool
    Cands cl;
    cl.lane = lane;
    cl.n = 0;
    uint32_t head = kPoolEnd;
    bool redo = false, overflow = false;
    for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
      uint32_t ref = __builtin_amdgcn_readfirstlane(M.roots4[r]);
      int sp = 0;
      // the current node's own box record (a root has none: roots are never bounds-tested)
      float4 cb0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), cb1 = cb0;
      bool has_box = false;
      while (ref != kDone) {
        if (!(ref & kRefLeaf)) {
          // ---- node step: the four children, scored by every lane ----
          // all four records are loaded before any is scored (one scalar round trip per node, not
          // one per child), and the maxThreshold decisions, which may call out of line, come after
          ref = __builtin_amdgcn_readfirstlane(ref);
          ConstF* const nd = nodes4 + 32ull * ref;
          uint32_t keys[4];
          uint64_t need_any = 0;
          float4 rec0[4], rec1[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            rec0[k] = cload4(nd + 8 * k);
            rec1[k] = cload4(nd + 8 * k + 4);
          }
          float raw[4];
          bool in[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const uint32_t cref = __float_as_uint(rec1[k].z);
            raw[k] = box_score(q, rec0[k].x, rec0[k].y, rec0[k].z, rec0[k].w, rec1[k].x, rec1[k].y);
            in[k] = cref != kDone && raw[k] * kScoreSlack <= U;
          }
          if (max_finite) {
#pragma unroll
            for (int k = 0; k < 4; k++)
              if (in[k])
                in[k] = below_max<kLo>(true, raw[k], max_t2, q, ql, rec0[k].x, rec0[k].y, rec0[k].z, rec0[k].w,
                                       rec1[k].x, rec1[k].y);
          }
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const uint64_t m = __ballot(in[k]);
            need_any |= m ? (1ull << k) : 0ull;
            // the leader's score (bits of a non-negative float order as unsigned) with the child's index
            const uint32_t kb = __builtin_amdgcn_readlane(__float_as_uint(raw[k]), leader);
            keys[k] = m ? ((kb & ~3u) | static_cast<uint32_t>(k)) : 0xFFFFFFFFu;
          }
          // sort the (at most four) entered children nearest first
#define MBVH_CSWAP(i, j)                  \
  {                                       \
    const uint32_t lo_ = min(keys[i], keys[j]); \
    keys[j] = max(keys[i], keys[j]);      \
    keys[i] = lo_;                        \
  }
          MBVH_CSWAP(0, 1) MBVH_CSWAP(2, 3) MBVH_CSWAP(0, 2) MBVH_CSWAP(1, 3) MBVH_CSWAP(1, 2)
#undef MBVH_CSWAP
          const int nh = __builtin_popcount(static_cast<uint32_t>(need_any));
          if (nh == 0) {
            ref = kDone;
          } else {
            if (sp + nh - 1 > kPacketStack) {
              overflow = true;
              break;
            }
            // push the farther entered children (farthest deepest), continue with the nearest
            for (int i = nh - 1; i >= 1; i--) {
              const uint32_t k = keys[i] & 3u;
              if (wl == 0) {
                stk[2 * sp] = pick4(k, rec0[0], rec0[1], rec0[2], rec0[3]);
                stk[2 * sp + 1] = pick4(k, rec1[0], rec1[1], rec1[2], rec1[3]);
              }
              sp++;
            }
            cb0 = pick4(keys[0] & 3u, rec0[0], rec0[1], rec0[2], rec0[3]);
            cb1 = pick4(keys[0] & 3u, rec1[0], rec1[1], rec1[2], rec1[3]);
            has_box = true;
            ref = __float_as_uint(cb1.z);
          }
        } else {
          // ---- leaf step: every lane tests its query against each triangle ----
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          // the packet entered this leaf for some lane; a lane whose maxThreshold excludes the leaf's
          // box (and so, the boxes nesting, the leaf) must not record its triangles
          bool reach = active;
          if (reach && max_finite && has_box) {
            const float raw = box_score(q, cb0.x, cb0.y, cb0.z, cb0.w, cb1.x, cb1.y);
            reach = below_max<kLo>(true, raw, max_t2, q, ql, cb0.x, cb0.y, cb0.z, cb0.w, cb1.x, cb1.y);
          }
          // the leaf (and so its range) is the packet's: the walk stays in scalar registers, and the
          // next triangle's scalar loads are issued before this one is tested so that their latency
          // overlaps the test instead of stalling the wave at the top of every iteration
          off = __builtin_amdgcn_readfirstlane(off);
          cnt = __builtin_amdgcn_readfirstlane(cnt);
          ConstF* tp = tris + 12ull * off;
          Tri nx;
          if (cnt) {
            nx.p0 = cload4(tp);
            nx.p1 = cload4(tp + 4);
            nx.p2 = cload4(tp + 8);
          }
          for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t s = off + k;
            const Tri tr = nx;
            tp += 12;
            if (k + 1 < cnt) {
              nx.p0 = cload4(tp);
              nx.p1 = cload4(tp + 4);
              nx.p2 = cload4(tp + 8);
            }
            const float sb = box_score(q, fminf(fminf(tr.p0.x, tr.p1.x), tr.p2.x), fminf(fminf(tr.p0.y, tr.p1.y), tr.p2.y),
                                       fminf(fminf(tr.p0.z, tr.p1.z), tr.p2.z), fmaxf(fmaxf(tr.p0.x, tr.p1.x), tr.p2.x),
                                       fmaxf(fmaxf(tr.p0.y, tr.p1.y), tr.p2.y), fmaxf(fmaxf(tr.p0.z, tr.p1.z), tr.p2.z));
            const bool near = reach && sb * (1.0f - 8.0f * kU) <= U;
            if (__ballot(near) == 0) continue;
            if (near) {
              const f3 cp = closest_on_triangle(q, tr.a(), tr.b(), tr.c());
              const f3 dv = sub(q, cp);
              const float d2 = dot(dv, dv);
              float lb = -kInf;  // FP32 failed (degenerate triangle, NaN): decide exactly at the end
              bool add = true;
              if (d2 == d2) {
                const float e2 = (2.0f * sqrt_up(d2) + e_abs) * e_abs + 4.0f * kU * d2;
                lb = d2 - e2;
                add = lb <= U;
                if (add) U = fminf(U, d2 + e2);
              }
              if (add) {
                if (cl.n < kCand || cl.compact(U)) {
                  cl.put(cl.n++, s, lb);
                } else {
                  const uint32_t at = atomicAdd(A.pool_count, 1u);
                  if (at < A.pool_cap) {
                    A.pool[at] = make_uint4(s, __float_as_uint(lb), head, 0u);
                    head = at;
                  } else {
                    redo = true;  // the pool ran out: the quad kernel finishes this query
                    U = -1.0f;
                  }
                }
              }
            }
          }
          ref = kDone;
        }
        // ---- pop: the nearest pushed child some lane still needs ----
        while (ref == kDone && sp > 0) {
          sp--;
          const float4 e0 = stk[2 * sp], e1 = stk[2 * sp + 1];
          const float raw = box_score(q, e0.x, e0.y, e0.z, e0.w, e1.x, e1.y);
          bool in = raw * kScoreSlack <= U;
          if (in && max_finite) in = below_max<kLo>(true, raw, max_t2, q, ql, e0.x, e0.y, e0.z, e0.w, e1.x, e1.y);
          if (__ballot(in)) {
            cb0 = e0;
            cb1 = e1;
            has_box = true;
            ref = __builtin_amdgcn_readfirstlane(__float_as_uint(e1.z));
          }
        }
      }
    }
    if (active) {
      const uint64_t v = visit - A.tile0 * 64;
      for (int c = 0; c < cl.n; c++) A.cand[v * kPacketCand + c] = cl.get(c);
      A.cand_meta[v] = make_uint4(overflow ? kPacketStackMark : (redo ? kPacketRedoMark : static_cast<uint32_t>(cl.n)),
                                  __float_as_uint(U), head, 0u);
      if (redo) A.redo[atomicAdd(A.redo_count, 1u)] = static_cast<uint32_t>(qi);
    }
  }
}

// The exact resolution of the packets' candidate lists: one lane per query, the smallest exact
// squared distance from q + ql, ties (point_tie) to the triangle the reference meets first; the
// reported distance, point and barycoord in double-float.
template <bool kLo>
__global__ __launch_bounds__(kBlock) void k_closest_point_resolve(ClosestPointArgs A) {
  const uint64_t v = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= A.tiles * 64) return;
  const uint64_t visit = A.tile0 * 64 + v;
  if (visit >= A.n) return;
  const uint64_t qi = A.order ? A.order[visit] : visit;
  if (qi >= A.n) return;
  const uint4 meta = A.cand_meta[v];
  if (meta.x == kPacketRedoMark) return;  // finished by the quad kernel
  const MeshView M = A.mesh;
  const TieTree T = tie_tree(M);
  const f3 q = mk(A.points[3 * qi + 0], A.points[3 * qi + 1], A.points[3 * qi + 2]);
  const f3 ql = kLo ? load3(A.points_lo, qi) : mk(0.0f, 0.0f, 0.0f);
  const float U = __uint_as_float(meta.y);
  const float qnan = __int_as_float(0x7fc00000);
  uint32_t out_face = meta.x == kPacketStackMark ? kOverflow : kNoHit;
  float out_dist = kInf;
  float px = qnan, py = qnan, pz = qnan, bw = qnan, bu = qnan, bv = qnan;
  if (meta.x != kPacketStackMark) {
    const uint2* const row = A.cand + v * kPacketCand;
    uint32_t win = kNoHit;
    ff best_d2 = {kInf, 0.0f};
    float best_err = 0.0f;
    uint32_t link = meta.z;
    // the listed entries, then the pool chain
    for (uint32_t c = 0; c < meta.x || link != kPoolEnd; c++) {
      uint2 e;
      if (c < meta.x) {
        e = row[c];
      } else {
        const uint4 pe = A.pool[link];
        e = make_uint2(pe.x, pe.y);
        link = pe.z;
      }
      if (__uint_as_float(e.y) > U) continue;
      const Tri tr = load_tri(M.tris, e.x);
      const ClosestFF cf = closest_on_triangle_ff_inline<kLo>(q, ql, tr.a(), tr.b(), tr.c());
      if (!(cf.dist2.hi == cf.dist2.hi)) continue;  // degenerate: never wins
      bool take = win == kNoHit;
      if (!take) {
        if (point_tie(cf.dist2, cf.err, best_d2, best_err))
          take = e.x != win && point_first_of<kLo>(T, q, ql, e.x, win);
        else
          take = ff_sign(ff_sub(cf.dist2, best_d2)) < 0;
      }
      if (take) {
        win = e.x;
        best_d2 = cf.dist2;
        best_err = cf.err;
      }
    }
    if (win != kNoHit) finish_point<kLo>(M.tris, win, q, ql, out_face, out_dist, px, py, pz, bw, bu, bv);
  }
  A.out_face[qi] = out_face;
  if (A.out_dist) A.out_dist[qi] = out_dist;
  if (A.out_point) {
    A.out_point[3 * qi + 0] = px;
    A.out_point[3 * qi + 1] = py;
    A.out_point[3 * qi + 2] = pz;
  }
  if (A.out_bary) {
    A.out_bary[3 * qi + 0] = bw;
    A.out_bary[3 * qi + 1] = bu;
    A.out_bary[3 * qi + 2] = bv;
  }
}



// One lane per query on the binary tree in the reference's own visit order (minThreshold's early
// exit returns the first triangle under it in that order): children ordered by their squared box
// distances (decided in double-float when the FP32 values are within their error band of each
// other), maxThreshold decided exactly, and the first triangle whose exact squared distance is below
// minThreshold^2 ends the query.
template <bool kLo>
__global__ __launch_bounds__(kBlock) void k_closest_point(ClosestPointArgs A) {
  const uint32_t lane = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + lane;
  const MeshView M = A.mesh;
  const TieTree T = tie_tree(M);

  Stack st;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = lane;
  st.capacity = kStackLds + static_cast<int>(A.stack.ovf_entries);

  const bool min_pos = A.min_threshold > 0.0f;
  const ff min_t2 = min_pos ? thr_sq(A.min_threshold, A.min_lo) : ff{0.0f, 0.0f};
  const bool max_finite = A.max_threshold < kInf;
  const ff max_t2 = max_finite ? thr_sq(A.max_threshold, A.max_lo) : ff{kInf, 0.0f};
  constexpr float kScoreSlack = 1.0f - 4.0f * kU;

  // Work queue: each wave takes the next 64 queries until none are left (query costs vary a lot,
  // so a dynamic assignment keeps the tail short). Every wave reaches the exit test.
  while (true) {
    uint32_t tile = 0;
    if ((lane & 63u) == 0) tile = atomicAdd(A.queue, 1u);
    tile = __shfl(tile, 0, 64);
    if (static_cast<uint64_t>(tile) * 64 >= A.n) break;
    const uint64_t visit = static_cast<uint64_t>(tile) * 64 + (lane & 63u);
    if (visit >= A.n) continue;
    const uint64_t qi = A.order ? A.order[visit] : visit;
    if (qi >= A.n) continue;  // not a permutation of 0..n-1: skip rather than read out of bounds
    const f3 q = mk(A.points[3 * qi + 0], A.points[3 * qi + 1], A.points[3 * qi + 2]);
    const f3 ql = kLo ? load3(A.points_lo, qi) : mk(0.0f, 0.0f, 0.0f);
    const float e_abs = query_error<kLo>(q, ql, *M.coord_max);
    const float w1 = kLo ? fabsf(ql.x) + fabsf(ql.y) + fabsf(ql.z) : 0.0f;
    float U = kInf;       // upper bound of the best squared distance
    bool done = false, overflow = false;
    Cands cl;
    cl.lane = lane;
    cl.n = 0;

    for (uint32_t r = 0; r < M.n_roots && !done; r++) {
      uint32_t ref = M.roots[r];
      st.sp = 0;
      while (true) {
        // ---- internal nodes ----
        while (!(ref & kRefLeaf) && ref != kDone) {
          const WideNode nd = M.nodes[ref];
          const float sl = box_score(q, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
          const float sr = box_score(q, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
          // right first only if strictly closer, decided exactly when the FP32 scores are too close
          // (each within 6u of its exact value; a double query's shift moves each by (2 sqrt(s) + w) w)
          bool swap = sr < sl;
          const float smax = fmaxf(sl, sr);
          const float band = 12.0f * kU * smax + (kLo ? (2.0f * sqrt_up(smax) + w1) * w1 * 2.0f : 0.0f);
          if (!(fabsf(sr - sl) > band))
            swap = ff_sign(ff_sub(box_score_ff<kLo>(q, ql, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w),
                                  box_score_ff<kLo>(q, ql, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y))) < 0;
          const uint32_t c1 = swap ? nd.meta.y : nd.meta.x;
          const uint32_t c2 = swap ? nd.meta.x : nd.meta.y;
          const float s1 = (swap ? sr : sl) * kScoreSlack;
          const float s2 = (swap ? sl : sr) * kScoreSlack;
          bool in1 = s1 <= U, in2 = s2 <= U;
          if (in1 && max_finite)
            in1 = swap ? below_max<kLo>(true, sr, max_t2, q, ql, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w)
                       : below_max<kLo>(true, sl, max_t2, q, ql, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
          if (in2 && max_finite)
            in2 = swap ? below_max<kLo>(true, sl, max_t2, q, ql, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y)
                       : below_max<kLo>(true, sr, max_t2, q, ql, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
          if (in1 && in2) {
            if (!st.push(c2, __float_as_uint(s2))) {
              overflow = done = true;
              ref = kDone;
              break;
            }
            ref = c1;
          } else if (in1 || in2) {
            ref = in1 ? c1 : c2;
          } else {
            ref = kDone;
            while (st.sp > 0) {
              const uint2 e = st.pop();
              if (__uint_as_float(e.y) <= U) {  // maxThreshold was decided when it was pushed
                ref = e.x;
                break;
              }
            }
          }
        }
        // ---- leaves ----
        while (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          Tri next;
          if (cnt != 0) next = load_tri(M.tris, off);  // an empty leaf may sit at the slot count
          for (uint32_t s = off; s < off + cnt; s++) {
            const Tri tr = next;
            if (s + 1 < off + cnt) next = load_tri(M.tris, s + 1);
            const f3 cp = closest_on_triangle(q, tr.a(), tr.b(), tr.c());
            const f3 dv = sub(q, cp);
            const float d2 = dot(dv, dv);
            float lb, e2 = 0.0f;
            bool add;
            if (!(d2 == d2)) {
              add = true;  // FP32 failed (degenerate triangle): decide exactly at the end
              lb = -kInf;
            } else {
              e2 = (2.0f * sqrt_up(d2) + e_abs) * e_abs + 4.0f * kU * d2;
              lb = d2 - e2;
              // every candidate that may tie stays listed (equal FP32 values do not make exact ties)
              add = lb <= U;
              if (add) U = fminf(U, d2 + e2);
            }
            if (add) {
              if (cl.n == kCand && !cl.compact(U)) {
                const Winner r = resolve_point<kLo, false>(cl, U, q, ql, M.tris, T);
                const uint32_t w = r.w;
                cl.n = 0;
                if (w != kNoHit) {
                  const float dw = ff_to_float(r.v);
                  U = fminf(U, dw * (1.0f + 4.0f * kU));
                  cl.put(0, w, dw * (1.0f - 4.0f * kU));
                  cl.n = 1;
                }
              }
              cl.put(cl.n++, s, lb);
            }
            // minThreshold: the first triangle whose exact squared distance is below it ends the
            // query (a degenerate triangle's NaN never is)
            if (min_pos && d2 == d2) {
              bool under;
              if (d2 + e2 < min_t2.hi * (1.0f - 4.0f * kU))
                under = true;
              else if (d2 - e2 > min_t2.hi * (1.0f + 4.0f * kU))
                under = false;
              else
                under = tri_below<kLo>(q, ql, tr, min_t2);
              if (under) {
                done = true;
                break;
              }
            }
          }
          ref = kDone;
          if (done) break;
          while (st.sp > 0) {
            const uint2 e = st.pop();
            if (__uint_as_float(e.y) <= U) {
              ref = e.x;
              break;
            }
          }
        }
        if (ref == kDone) break;
      }
    }

    uint32_t out_face = kNoHit;
    float out_dist = kInf;
    float px = __int_as_float(0x7fc00000), py = px, pz = px;
    float bw = px, bu = px, bv = px;
    if (overflow) {
      out_face = kOverflow;
    } else if (cl.n > 0) {
      const uint32_t w = resolve_point<kLo, false>(cl, U, q, ql, M.tris, T).w;
      if (w != kNoHit) finish_point<kLo>(M.tris, w, q, ql, out_face, out_dist, px, py, pz, bw, bu, bv);
    }
    A.out_face[qi] = out_face;
    if (A.out_dist) A.out_dist[qi] = out_dist;
    if (A.out_point) {
      A.out_point[3 * qi + 0] = px;
      A.out_point[3 * qi + 1] = py;
      A.out_point[3 * qi + 2] = pz;
    }
    if (A.out_bary) {
      A.out_bary[3 * qi + 0] = bw;
      A.out_bary[3 * qi + 1] = bu;
      A.out_bary[3 * qi + 2] = bv;
    }
  }
}

// ---- scenes: closest first hit over instances (TLAS over instance world boxes) ------------------

constexpr int kSceneStack = 25;  // TLAS stack entries per lane: the TLAS is at most 24 levels deep
constexpr uint32_t kTlasStack = 24;  // the wave-per-point kernel's top-level stack (one far child per level)

// Entry distance of a world ray into a box of the TLAS (near plane respected).
__device__ __forceinline__ bool box_entry(f3 o, f3 inv, const float* b, float near_dist, float& t_enter) {
  const float tx0 = (b[0] - o.x) * inv.x, tx1 = (b[3] - o.x) * inv.x;
  const float ty0 = (b[1] - o.y) * inv.y, ty1 = (b[4] - o.y) * inv.y;
  const float tz0 = (b[2] - o.z) * inv.z, tz1 = (b[5] - o.z) * inv.z;
  const float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
  const float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
  t_enter = tmin;
  return tmin <= tmax && tmax >= near_dist;
}

// One lane per ray. The TLAS (reference node format over the instances' world boxes, expanded
// outward when built) is walked near child first and pruned by the best world distance so far
// (with a 2^-16 relative margin: box distances and hit distances round differently); each
// instance leaf brings the ray into the instance's frame and traces its mesh exactly as the batched
// per-instance launch does (near / far scaled by the instance, hit point and distance back in world
// space), so every instance's hit is that launch's hit. The closest world distance wins; equal
// distances keep the lowest instance index, the order a stable sort of the reference's
// per-instance hits (acceleratedBatchedMeshRaycast) gives.
__global__ __launch_bounds__(kBlock) void k_scene_raycast_first(SceneArgs A) {
  __shared__ uint2 lds_stack[8 * kBlock];
  __shared__ uint32_t tstack[kSceneStack * kBlock];
  const uint32_t tid = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + tid;
  StackL<8> st;
  st.lds = lds_stack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = tid;
  st.capacity = 8 + static_cast<int>(A.stack.ovf_entries);
  const float qnan = __int_as_float(0x7fc00000);
  const float margin = 1.0f + 1.0f / 65536.0f;
  for (uint64_t ray = gtid; ray < A.n; ray += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const f3 ow = mk(A.origins[3 * ray], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
    const f3 dw = mk(A.directions[3 * ray], A.directions[3 * ray + 1], A.directions[3 * ray + 2]);
    const f3 iw = mk(1.0f / dw.x, 1.0f / dw.y, 1.0f / dw.z);
    FirstHit best;
    best.face = kNoHit;
    best.dist = kInf;
    best.p = mk(qnan, qnan, qnan);
    best.bary = best.p;
    uint32_t best_inst = kNoHit;
    bool overflow = false;
    int sp = 1;
    tstack[tid] = 0u;
    while (sp > 0 && !overflow) {
      sp--;
      const uint32_t node = tstack[sp * kBlock + tid];
      const uint32_t* nw = A.tlas + 8ull * node;
      float te;
      if (!box_entry(ow, iw, reinterpret_cast<const float*>(nw), A.near_dist, te)) continue;
      if (te > fminf(best.dist, A.far_dist) * margin + 1e-30f) continue;
      const uint32_t w7 = nw[7];
      if ((w7 >> 16) == 0xFFFFu) {
        const uint32_t off = nw[6], cnt = w7 & 0xFFFFu;
        for (uint32_t s = off; s < off + cnt; s++) {
          const uint32_t inst = A.tlas_inst[s];
          const uint32_t geom = A.instance_geometry[inst];
          const float* inv = A.inverse + 16ull * inst;
          const f3 o = apply_matrix4(inv, ow);
          const f3 d = transform_direction(inv, dw);
          const float sc = instance_scale(A.world + 16ull * inst, d);
          const f3 z = mk(0.0f, 0.0f, 0.0f);
          FirstHit h = trace_first<false>(A.meshes[geom], st, tid, o, d, z, z, A.side_mask,
                                          window_of(A.near_dist / sc, 0.0f, A.far_dist / sc, 0.0f));
          if (h.face == kOverflow) {
            overflow = true;
            break;
          }
          if (h.face != kNoHit) {
            h.p = apply_matrix4(A.world + 16ull * inst, h.p);
            h.dist = world_distance(h.p, ow);
            if (h.dist < best.dist || (h.dist == best.dist && inst < best_inst)) {
              best = h;
              best_inst = inst;
            }
          }
        }
      } else {
        if (sp + 2 > kSceneStack) {
          overflow = true;
          break;
        }
        const uint32_t left = node + 1, right = node + nw[6];
        const bool l2r = comp(dw, static_cast<int>(w7)) >= 0.0f;
        tstack[sp * kBlock + tid] = l2r ? right : left;  // far child below the near one
        tstack[(sp + 1) * kBlock + tid] = l2r ? left : right;
        sp += 2;
      }
    }
    if (overflow) {
      best.face = kOverflow;
      best_inst = kNoHit;
    }
    A.out_face[ray] = best.face;
    if (A.out_instance) A.out_instance[ray] = best_inst;
    if (A.out_dist) A.out_dist[ray] = best.dist;
    if (A.out_point) {
      A.out_point[3 * ray + 0] = best.p.x;
      A.out_point[3 * ray + 1] = best.p.y;
      A.out_point[3 * ray + 2] = best.p.z;
    }
    if (A.out_bary) {
      A.out_bary[3 * ray + 0] = best.bary.x;
      A.out_bary[3 * ray + 1] = best.bary.y;
      A.out_bary[3 * ray + 2] = best.bary.z;
    }
  }
}

// ---- intersectsBox: an oriented box against the mesh ----------------------------------------------
// The reference decides it in JS numbers (IEEE double) with separating-axis tests
// (OrientedBox.update / intersectsBox / intersectsTriangle src/math/OrientedBox.js:79-260,
// SeparatingAxisBounds src/math/SeparatingAxisBounds.js, ExtendedTriangle.update
// src/math/ExtendedTriangle.js:47-75 and three's Vector3 / Triangle arithmetic). The same expressions
// are evaluated here in double, in the same order and without contraction, so every projection and
// every comparison has the reference's value. The node test only uses six axes, so it never rejects a
// node that holds an intersecting triangle; the answer is whether any triangle intersects.

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) {
  MBVH_NO_CONTRACT
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ double ddot(D3 a, D3 b) {  // Vector3.dot
  MBVH_NO_CONTRACT
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {  // Vector3.crossVectors
  MBVH_NO_CONTRACT
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dcomp(D3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
// JavaScript Math.min / Math.max: NaN if either operand is NaN; -0 below +0.
__device__ __forceinline__ double js_min(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == b) return __builtin_signbit(a) ? a : b;
  return a < b ? a : b;
}
__device__ __forceinline__ double js_max(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == b) return __builtin_signbit(a) ? b : a;
  return a > b ? a : b;
}

struct SepBounds {  // SeparatingAxisBounds
  double mn, mx;
};
// setFromPoints: min / max of the projections with the ternaries (a NaN projection is skipped)
__device__ __forceinline__ SepBounds sep_points(D3 axis, const D3* pts, int n) {
  SepBounds b{__builtin_huge_val(), -__builtin_huge_val()};
  for (int i = 0; i < n; i++) {
    const double v = ddot(axis, pts[i]);
    b.mn = v < b.mn ? v : b.mn;
    b.mx = v > b.mx ? v : b.mx;
  }
  return b;
}
__device__ __forceinline__ bool separated(SepBounds a, SepBounds b) { return a.mn > b.mx || b.mn > a.mx; }

struct Obb {  // OrientedBox after update()
  D3 p[8];
  D3 ax[3];
  SepBounds sb[3];
  SepBounds ab[3];  // alignedSatBounds
};

__device__ void obb_update(const double* bmin, const double* bmax, const double* e, Obb& o) {
  MBVH_NO_CONTRACT
  for (int x = 0; x <= 1; x++) {
    for (int y = 0; y <= 1; y++) {
      for (int z = 0; z <= 1; z++) {
        const int i = x | (y << 1) | (z << 2);
        const double vx = x ? bmax[0] : bmin[0], vy = y ? bmax[1] : bmin[1], vz = z ? bmax[2] : bmin[2];
        const double w = 1 / (e[3] * vx + e[7] * vy + e[11] * vz + e[15]);  // Vector3.applyMatrix4
        o.p[i] = d3((e[0] * vx + e[4] * vy + e[8] * vz + e[12]) * w, (e[1] * vx + e[5] * vy + e[9] * vz + e[13]) * w,
                    (e[2] * vx + e[6] * vy + e[10] * vz + e[14]) * w);
      }
    }
  }
  for (int i = 0; i < 3; i++) {
    o.ax[i] = dsub(o.p[0], o.p[1 << i]);
    o.sb[i] = sep_points(o.ax[i], o.p, 8);
  }
  for (int k = 0; k < 3; k++) {  // setFromPointsField
    SepBounds b{__builtin_huge_val(), -__builtin_huge_val()};
    for (int i = 0; i < 8; i++) {
      const double v = dcomp(o.p[i], k);
      b.mn = v < b.mn ? v : b.mn;
      b.mx = v > b.mx ? v : b.mx;
    }
    o.ab[k] = b;
  }
}

// OrientedBox.intersectsBox against a node box (float bounds as JS numbers).
__device__ bool obb_box(const Obb& o, const double* mn, const double* mx) {
  MBVH_NO_CONTRACT
  for (int k = 0; k < 3; k++)
    if (separated(o.ab[k], SepBounds{mn[k], mx[k]})) return false;
  for (int i = 0; i < 3; i++) {
    SepBounds b{__builtin_huge_val(), -__builtin_huge_val()};  // setFromBox: corners, Math.min / Math.max
    for (int x = 0; x <= 1; x++) {
      for (int y = 0; y <= 1; y++) {
        for (int z = 0; z <= 1; z++) {
          const D3 q = d3(mn[0] * x + mx[0] * (1 - x), mn[1] * y + mx[1] * (1 - y), mn[2] * z + mx[2] * (1 - z));
          const double v = ddot(o.ax[i], q);
          b.mn = js_min(v, b.mn);
          b.mx = js_max(v, b.mx);
        }
      }
    }
    if (separated(o.sb[i], b)) return false;
  }
  return true;
}

// OrientedBox.intersectsTriangle with the triangle's ExtendedTriangle.update axes: the normal
// (Triangle.getNormal) and the edge vectors a - b, b - c, c - a.
__device__ bool obb_triangle(const Obb& o, D3 a, D3 b, D3 c) {
  MBVH_NO_CONTRACT
  const D3 pts[3] = {a, b, c};
  D3 tax[4];
  D3 nrm = dcross(dsub(c, b), dsub(a, b));
  const double len2 = nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z;
  if (len2 > 0) {
    const double s = 1 / sqrt(len2);
    nrm = d3(nrm.x * s, nrm.y * s, nrm.z * s);
  } else {
    nrm = d3(0, 0, 0);
  }
  tax[0] = nrm;
  tax[1] = dsub(a, b);
  tax[2] = dsub(b, c);
  tax[3] = dsub(c, a);
  for (int i = 0; i < 3; i++)
    if (separated(o.sb[i], sep_points(o.ax[i], pts, 3))) return false;
  for (int i = 0; i < 3; i++)
    if (separated(sep_points(tax[i], pts, 3), sep_points(tax[i], o.p, 8))) return false;
  for (int i = 0; i < 3; i++) {
    for (int i2 = 0; i2 < 4; i2++) {
      const D3 axis = dcross(o.ax[i], tax[i2]);
      if (separated(sep_points(axis, pts, 3), sep_points(axis, o.p, 8))) return false;
    }
  }
  return true;
}

__device__ __forceinline__ void node_box_d(float a, float b, float c, float d, float e, float f, double* mn, double* mx) {
  mn[0] = a;
  mn[1] = b;
  mn[2] = c;
  mx[0] = d;
  mx[1] = e;
  mx[2] = f;
}

// One lane per box: depth-first any-hit traversal of the binary tree (node boxes by the six-axis test,
// triangles by the full test), stopping at the first intersecting triangle.
__global__ __launch_bounds__(kBlock) void k_intersects_box(BoxArgs A) {
  __shared__ uint2 lds_stack[8 * kBlock];
  const uint32_t tid = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + tid;
  const MeshView M = A.mesh;
  StackL<8> st;
  st.lds = lds_stack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = tid;
  st.capacity = 8 + static_cast<int>(A.stack.ovf_entries);
  for (uint64_t q = gtid; q < A.n; q += static_cast<uint64_t>(gridDim.x) * kBlock) {
    Obb o;
    obb_update(A.box_min + 3 * q, A.box_max + 3 * q, A.matrices + 16 * q, o);
    uint8_t hit = 0;
    bool overflow = false;
    for (uint32_t r = 0; r < M.n_roots && !hit && !overflow; r++) {
      st.sp = 0;
      uint32_t ref = M.roots[r];
      while (true) {
        if (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          for (uint32_t s = off; s < off + cnt && !hit; s++) {
            const Tri t = load_tri(M.tris, s);
            if (obb_triangle(o, d3(t.p0.x, t.p0.y, t.p0.z), d3(t.p1.x, t.p1.y, t.p1.z), d3(t.p2.x, t.p2.y, t.p2.z)))
              hit = 1;
          }
          if (hit) break;
          ref = kDone;
        } else if (ref != kDone) {
          const WideNode nd = M.nodes[ref];
          double mn[3], mx[3];
          node_box_d(nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y, mn, mx);
          const bool hl = obb_box(o, mn, mx);
          node_box_d(nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w, mn, mx);
          const bool hr = obb_box(o, mn, mx);
          if (hl && hr) {
            if (!st.push(nd.meta.y, 0u)) {
              overflow = true;
              break;
            }
            ref = nd.meta.x;
          } else {
            ref = hl ? nd.meta.x : (hr ? nd.meta.y : kDone);
          }
          continue;
        }
        if (st.sp == 0) break;
        ref = st.pop().x;
      }
    }
    A.out[q] = overflow ? 2 : hit;
  }
}

// ---- closestPointToGeometry: the closest pair of triangles between the mesh and another geometry ----
// (src/core/cast/closestPointToGeometry.template.js:15-271 with ExtendedTriangle.distanceToTriangle
// src/math/ExtendedTriangle.js:774-860). In double: 0 when an edge of one triangle pierces the other
// (intersecting triangles; coplanar overlaps give 0 below), else the least of the six vertex-to-triangle
// distances and the nine edge-to-edge distances, as distanceToTriangle takes it.

__device__ __forceinline__ D3 dadd(D3 a, D3 b) {
  MBVH_NO_CONTRACT
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ D3 dscale(D3 a, double s) {
  MBVH_NO_CONTRACT
  return {a.x * s, a.y * s, a.z * s};
}
__device__ __forceinline__ double ddist2(D3 a, D3 b) {
  const D3 d = dsub(a, b);
  return ddot(d, d);
}

// Closest point of triangle abc to p (Ericson, Real-Time Collision Detection 5.1.5; three's
// Triangle.closestPointToPoint).
__device__ D3 closest_on_tri_d(D3 p, D3 a, D3 b, D3 c) {
  MBVH_NO_CONTRACT
  const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
  const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
  if (d1 <= 0 && d2 <= 0) return a;
  const D3 bp = dsub(p, b);
  const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return dadd(a, dscale(ab, d1 / (d1 - d3)));
  const D3 cp = dsub(p, c);
  const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return dadd(a, dscale(ac, d2 / (d2 - d6)));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0)
    return dadd(b, dscale(dsub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  const double denom = 1 / (va + vb + vc);
  return dadd(dadd(a, dscale(ab, vb * denom)), dscale(ac, vc * denom));
}

// Closest points of segments p1q1 and p2q2 (Ericson 5.1.9, clamped).
__device__ void seg_seg_d(D3 p1, D3 q1, D3 p2, D3 q2, D3& c1, D3& c2) {
  MBVH_NO_CONTRACT
  const D3 d1 = dsub(q1, p1), d2 = dsub(q2, p2), r = dsub(p1, p2);
  const double a = ddot(d1, d1), e = ddot(d2, d2), f = ddot(d2, r);
  double s, t;
  if (a <= 0 && e <= 0) {
    s = t = 0;
  } else if (a <= 0) {
    s = 0;
    t = fmin(fmax(f / e, 0.0), 1.0);
  } else {
    const double c = ddot(d1, r);
    if (e <= 0) {
      t = 0;
      s = fmin(fmax(-c / a, 0.0), 1.0);
    } else {
      const double b = ddot(d1, d2), den = a * e - b * b;
      s = den != 0 ? fmin(fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
      t = (b * s + f) / e;
      if (t < 0) {
        t = 0;
        s = fmin(fmax(-c / a, 0.0), 1.0);
      } else if (t > 1) {
        t = 1;
        s = fmin(fmax((b - c) / a, 0.0), 1.0);
      }
    }
  }
  c1 = dadd(p1, dscale(d1, s));
  c2 = dadd(p2, dscale(d2, t));
}

__device__ __forceinline__ double orient3(D3 a, D3 b, D3 c, D3 d) {  // (b - a) x (c - a) . (d - a)
  return ddot(dcross(dsub(b, a), dsub(c, a)), dsub(d, a));
}

// Whether segment pq crosses triangle abc at a single point (strictly opposite sides, or one end on the
// plane; both ends on the plane is left to the coplanar distances), writing that point.
__device__ bool seg_pierces_d(D3 p, D3 q, D3 a, D3 b, D3 c, D3& x) {
  MBVH_NO_CONTRACT
  const double sp = orient3(a, b, c, p), sq = orient3(a, b, c, q);
  if ((sp > 0 && sq > 0) || (sp < 0 && sq < 0) || (sp == 0 && sq == 0)) return false;
  const double u = orient3(p, q, a, b), v = orient3(p, q, b, c), w = orient3(p, q, c, a);
  if (!((u >= 0 && v >= 0 && w >= 0) || (u <= 0 && v <= 0 && w <= 0))) return false;
  x = dadd(p, dscale(dsub(q, p), sp / (sp - sq)));
  return true;
}

// distanceToTriangle: the distance and the closest points on t1 (a1 b1 c1) and t2. Intersecting
// triangles give 0 and, as the reference's intersection segment does, the middle of the segment
// where they cross (its ends are where edges of one meet the other).
__device__ double tri_tri_d(const D3* t1, const D3* t2, D3& o1, D3& o2) {
  MBVH_NO_CONTRACT
  const D3 line = dcross(dcross(dsub(t1[1], t1[0]), dsub(t1[2], t1[0])), dcross(dsub(t2[1], t2[0]), dsub(t2[2], t2[0])));
  bool found = false;
  double lo = 0, hi = 0;
  D3 plo = d3(0, 0, 0), phi = d3(0, 0, 0);
  for (int k = 0; k < 6; k++) {
    const D3* e = k < 3 ? t1 : t2;
    const D3* t = k < 3 ? t2 : t1;
    const int i = k % 3;
    D3 x;
    if (!seg_pierces_d(e[i], e[(i + 1) % 3], t[0], t[1], t[2], x)) continue;
    const double sx = ddot(x, line);
    if (!found || sx < lo) {
      lo = sx;
      plo = x;
    }
    if (!found || sx > hi) {
      hi = sx;
      phi = x;
    }
    found = true;
  }
  if (found) {
    o1 = o2 = dscale(dadd(plo, phi), 0.5);
    return 0.0;
  }
  double best = __builtin_huge_val();
  for (int i = 0; i < 3; i++) {
    D3 q = closest_on_tri_d(t2[i], t1[0], t1[1], t1[2]);
    double d = ddist2(t2[i], q);
    if (d < best) {
      best = d;
      o1 = q;
      o2 = t2[i];
    }
    q = closest_on_tri_d(t1[i], t2[0], t2[1], t2[2]);
    d = ddist2(t1[i], q);
    if (d < best) {
      best = d;
      o1 = t1[i];
      o2 = q;
    }
  }
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) {
      D3 c1, c2;
      seg_seg_d(t1[i], t1[(i + 1) % 3], t2[j], t2[(j + 1) % 3], c1, c2);
      const double d = ddist2(c1, c2);
      if (d < best) {
        best = d;
        o1 = c1;
        o2 = c2;
      }
    }
  }
  return sqrt(best);
}

// three's Vector3.applyMatrix4 in double (column-major elements, divided by w).
__device__ __forceinline__ D3 apply_matrix4_d(const double* e, double x, double y, double z) {
  MBVH_NO_CONTRACT
  const double w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
  return d3((e[0] * x + e[4] * y + e[8] * z + e[12]) * w, (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
            (e[2] * x + e[6] * y + e[10] * z + e[14]) * w);
}

// Triangle j of the other geometry in the mesh frame; false when an index entry is out of range.
__device__ __forceinline__ bool other_triangle(const GeometryArgs& A, uint64_t j, D3* t2) {
  for (int k = 0; k < 3; k++) {
    uint64_t v = 3 * j + k;
    if (A.other_index)
      v = A.other_elt == 2 ? static_cast<const uint16_t*>(A.other_index)[v] : static_cast<const uint32_t*>(A.other_index)[v];
    if (v >= A.other_vertices) return false;
    const float* p = A.other_positions + v * A.other_stride + A.other_offset;
    t2[k] = apply_matrix4_d(A.matrix, p[0], p[1], p[2]);
  }
  return true;
}

// A lower bound of the computed distance between a triangle inside the double box (mn, mx) and one
// inside the float node box: the boxes' gap, less a margin for the roundings of both computations
// (1e-12 of the largest coordinate involved), never negative.
__device__ __forceinline__ double box_box_lb(const double* mn, const double* mx, double mag, float a, float b, float c,
                                             float d, float e, float f) {
  MBVH_NO_CONTRACT
  const double lo[3] = {a, b, c}, hi[3] = {d, e, f};
  double s = 0, m = mag;
  for (int k = 0; k < 3; k++) {
    const double g = fmax(fmax(lo[k] - mx[k], mn[k] - hi[k]), 0.0);
    s += g * g;
    m = fmax(m, fmax(fabs(lo[k]), fabs(hi[k])));
  }
  return fmax(sqrt(s) * (1 - 1e-12) - 1e-12 * m, 0.0);
}

__device__ __forceinline__ double shared_best(const unsigned long long* bits) {
  return __longlong_as_double(static_cast<long long>(__hip_atomic_load(bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}

// One lane per triangle of the other geometry (brought into the mesh frame by `matrix`, three's
// applyMatrix4): a depth-first traversal of the mesh, near child first. A node is skipped when its lower
// bound exceeds the lane's best distance or the best of all lanes (an atomic minimum on the distance's
// bits), or reaches maxThreshold. Equal bounds are still visited and equal distances go to the lower
// triangle slot, so a lane whose closest distance is the overall minimum finds it, and the lowest slot
// at it, whatever the other lanes have published by then: the reduction over lanes is deterministic.
__global__ __launch_bounds__(kBlock) void k_closest_to_geometry(GeometryArgs A) {
  __shared__ uint2 lds_stack[8 * kBlock];
  const uint32_t tid = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + tid;
  const MeshView M = A.mesh;
  StackL<8> st;
  st.lds = lds_stack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = tid;
  st.capacity = 8 + static_cast<int>(A.stack.ovf_entries);
  const double cap = A.max_threshold;
  for (uint64_t j = gtid; j < A.n_other; j += static_cast<uint64_t>(gridDim.x) * kBlock) {
    D3 t2[3];
    double mn[3] = {__builtin_huge_val(), __builtin_huge_val(), __builtin_huge_val()};
    double mx[3] = {-__builtin_huge_val(), -__builtin_huge_val(), -__builtin_huge_val()};
    if (!other_triangle(A, j, t2)) {
      atomicOr(A.status, 2u);
      A.lane_dist[j] = __builtin_huge_val();
      A.lane_face[j] = kNoHit;
      continue;
    }
    for (int k = 0; k < 3; k++) {
      mn[0] = fmin(mn[0], t2[k].x);
      mn[1] = fmin(mn[1], t2[k].y);
      mn[2] = fmin(mn[2], t2[k].z);
      mx[0] = fmax(mx[0], t2[k].x);
      mx[1] = fmax(mx[1], t2[k].y);
      mx[2] = fmax(mx[2], t2[k].z);
    }
    const double mag = fmax(fmax(fmax(fabs(mn[0]), fabs(mn[1])), fmax(fabs(mn[2]), fabs(mx[0]))), fmax(fabs(mx[1]), fabs(mx[2])));
    double best = __builtin_huge_val();
    uint32_t best_face = kNoHit;
    bool overflow = false;
    for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
      st.sp = 0;
      uint32_t ref = M.roots[r];
      while (true) {
        if (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          const double leaf_bound = fmin(best, shared_best(A.best_bits));
          for (uint32_t s = off; s < off + cnt; s++) {
            const Tri t = load_tri(M.tris, s);
            // the triangle's own box first: the full test only where the pair could still win or tie
            const double tlb = box_box_lb(mn, mx, mag, fminf(fminf(t.p0.x, t.p1.x), t.p2.x),
                                          fminf(fminf(t.p0.y, t.p1.y), t.p2.y), fminf(fminf(t.p0.z, t.p1.z), t.p2.z),
                                          fmaxf(fmaxf(t.p0.x, t.p1.x), t.p2.x), fmaxf(fmaxf(t.p0.y, t.p1.y), t.p2.y),
                                          fmaxf(fmaxf(t.p0.z, t.p1.z), t.p2.z));
            if (tlb > fmin(best, leaf_bound) || tlb >= cap) continue;
            const D3 t1[3] = {d3(t.p0.x, t.p0.y, t.p0.z), d3(t.p1.x, t.p1.y, t.p1.z), d3(t.p2.x, t.p2.y, t.p2.z)};
            D3 o1, o2;
            const double dist = tri_tri_d(t1, t2, o1, o2);
            if (dist < cap && (dist < best || (dist == best && s < best_face))) {
              if (dist < best) atomicMin(A.best_bits, static_cast<unsigned long long>(__double_as_longlong(dist)));
              best = dist;
              best_face = s;
            }
          }
        } else {
          const double bound = fmin(best, shared_best(A.best_bits));
          const WideNode nd = M.nodes[ref];
          const double dl = box_box_lb(mn, mx, mag, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
          const double dr = box_box_lb(mn, mx, mag, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
          const bool hl = dl <= bound && dl < cap, hr = dr <= bound && dr < cap;
          if (hl && hr) {
            const bool left_first = dl <= dr;
            // the upper word of a non-negative double rounds it down: still a lower bound when popped
            if (!st.push(left_first ? nd.meta.y : nd.meta.x,
                         static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(left_first ? dr : dl)) >> 32))) {
              overflow = true;
              break;
            }
            ref = left_first ? nd.meta.x : nd.meta.y;
            continue;
          }
          if (hl || hr) {
            ref = hl ? nd.meta.x : nd.meta.y;
            continue;
          }
        }
        ref = kDone;
        while (st.sp > 0) {
          const uint2 en = st.pop();
          const double lb = __longlong_as_double(static_cast<long long>(static_cast<uint64_t>(en.y) << 32));
          if (lb <= fmin(best, shared_best(A.best_bits))) {
            ref = en.x;
            break;
          }
        }
        if (ref == kDone) break;
      }
    }
    if (overflow) atomicOr(A.status, 1u);
    A.lane_dist[j] = best;
    A.lane_face[j] = best_face;
  }
}

// A first bound for every lane: each other triangle follows the nearer child down to one leaf per root
// and publishes its closest pair there. Always a real pair's distance, so never below the overall
// minimum: the traversal that follows finds the same answer, with less to visit.
__global__ __launch_bounds__(kBlock) void k_geometry_seed(GeometryArgs A) {
  const MeshView M = A.mesh;
  for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < A.n_other;
       j += static_cast<uint64_t>(gridDim.x) * kBlock) {
    D3 t2[3];
    if (!other_triangle(A, j, t2)) continue;
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
      const double a = k == 0 ? t2[0].x : (k == 1 ? t2[0].y : t2[0].z);
      const double b = k == 0 ? t2[1].x : (k == 1 ? t2[1].y : t2[1].z);
      const double c = k == 0 ? t2[2].x : (k == 1 ? t2[2].y : t2[2].z);
      mn[k] = fmin(fmin(a, b), c);
      mx[k] = fmax(fmax(a, b), c);
    }
    double best = __builtin_huge_val();
    for (uint32_t r = 0; r < M.n_roots; r++) {
      uint32_t ref = M.roots[r];
      while (!(ref & kRefLeaf)) {
        const WideNode nd = M.nodes[ref];
        const double dl = box_box_lb(mn, mx, 0.0, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
        const double dr = box_box_lb(mn, mx, 0.0, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
        ref = dl <= dr ? nd.meta.x : nd.meta.y;
      }
      uint32_t off, cnt;
      leaf_range(ref, M.bigleaf, off, cnt);
      for (uint32_t s = off; s < off + cnt; s++) {
        const Tri t = load_tri(M.tris, s);
        const D3 t1[3] = {d3(t.p0.x, t.p0.y, t.p0.z), d3(t.p1.x, t.p1.y, t.p1.z), d3(t.p2.x, t.p2.y, t.p2.z)};
        D3 o1, o2;
        best = fmin(best, tri_tri_d(t1, t2, o1, o2));
      }
    }
    if (best < A.max_threshold) atomicMin(A.best_bits, static_cast<unsigned long long>(__double_as_longlong(best)));
  }
}

__global__ void k_geometry_clear(GeometryArgs A) {
  *A.best_bits = static_cast<unsigned long long>(__double_as_longlong(__builtin_huge_val()));
  *A.winner = kNoHit;
  *A.status = 0;
}

// The lowest other triangle whose closest distance is the overall minimum.
__global__ __launch_bounds__(kBlock) void k_geometry_pick(GeometryArgs A) {
  const double best = __longlong_as_double(static_cast<long long>(*A.best_bits));
  for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; j < A.n_other;
       j += static_cast<uint64_t>(gridDim.x) * kBlock)
    if (A.lane_face[j] != kNoHit && A.lane_dist[j] == best) atomicMin(A.winner, static_cast<uint32_t>(j));
}

// The answer record: the winning pair's closest points (recomputed by the same code, so the same
// values), the other point and the reference's target2.distance in the other geometry's frame.
__global__ void k_geometry_gather(GeometryArgs A) {
  MBVH_NO_CONTRACT
  GeometryHit h{};
  h.status = static_cast<int32_t>(*A.status);
  const uint32_t j = *A.winner;
  const double nan = __builtin_nan("");
  if (j == kNoHit || h.status) {
    h.distance = __builtin_huge_val();
    h.face_index = h.other_face_index = kNoHit;
    for (int k = 0; k < 3; k++) h.point[k] = h.other_point[k] = nan;
    h.other_distance = __builtin_huge_val();
  } else {
    const uint32_t s = A.lane_face[j];
    D3 t2[3];
    other_triangle(A, j, t2);
    const Tri t = load_tri(A.mesh.tris, s);
    const D3 t1[3] = {d3(t.p0.x, t.p0.y, t.p0.z), d3(t.p1.x, t.p1.y, t.p1.z), d3(t.p2.x, t.p2.y, t.p2.z)};
    D3 o1, o2;
    h.distance = tri_tri_d(t1, t2, o1, o2);
    h.face_index = s;
    h.other_face_index = j;
    h.point[0] = o1.x;
    h.point[1] = o1.y;
    h.point[2] = o1.z;
    const D3 q1 = apply_matrix4_d(A.inverse, o1.x, o1.y, o1.z), q2 = apply_matrix4_d(A.inverse, o2.x, o2.y, o2.z);
    h.other_point[0] = q2.x;
    h.other_point[1] = q2.y;
    h.other_point[2] = q2.z;
    h.other_distance = sqrt(ddist2(q1, q2));
  }
  *A.out = h;
}

// ---- scenes: closest point over instances ------------------------------------------------------------
// The reference's scene closest point (bvh_ClosestPointToPoint, src/webgpu/shapecastFns/
// getClosestPointToPointFn.js:20-151 over ClusteredBVH's top level) measures in world space: node
// boxes of an instance's tree are taken to world boxes through its matrix, triangles are taken to
// world space, and the least world distance wins. Here the triangles and distances are in double;
// the boxes are FP32 lower bounds with margins for their roundings.

// Squared distance from p to the world box of a local box under W (local -> world, column-major): the
// box's centre through W, its half extents through |W|, both widened by 2^-16 of the magnitudes
// involved; a lower bound of the squared distance to anything W puts inside it.
__device__ __forceinline__ double world_box_d2_lb(const float* W, float px, float py, float pz, float mnx, float mny,
                                                  float mnz, float mxx, float mxy, float mxz) {
  const float cx = 0.5f * (mnx + mxx), cy = 0.5f * (mny + mxy), cz = 0.5f * (mnz + mxz);
  const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);
  const float ax = fabsf(cx) + hx, ay = fabsf(cy) + hy, az = fabsf(cz) + hz;
  float s = 0.0f;
  for (int i = 0; i < 3; i++) {
    const float w0 = W[i], w1 = W[4 + i], w2 = W[8 + i], t = W[12 + i];
    const float wc = w0 * cx + w1 * cy + w2 * cz + t;
    const float wh = fabsf(w0) * hx + fabsf(w1) * hy + fabsf(w2) * hz;
    const float m = (fabsf(w0) * ax + fabsf(w1) * ay + fabsf(w2) * az + fabsf(t)) * (1.0f / 65536.0f);
    const float q = i == 0 ? px : (i == 1 ? py : pz);
    const float g = fmaxf(fabsf(q - wc) - (wh + m), 0.0f);
    s += g * g;
  }
  return static_cast<double>(s) * (1.0 - 1.0 / 65536.0);
}

// The same for a world box (the top level), widened by 2^-16 of its coordinates.
__device__ __forceinline__ double tlas_box_d2_lb(const float* b, float px, float py, float pz) {
  float s = 0.0f;
  for (int i = 0; i < 3; i++) {
    const float q = i == 0 ? px : (i == 1 ? py : pz);
    const float e = fmaxf(fabsf(b[i]), fabsf(b[3 + i])) * (1.0f / 65536.0f);
    const float g = fmaxf(fmaxf(b[i] - e - q, q - b[3 + i] - e), 0.0f);
    s += g * g;
  }
  return static_cast<double>(s) * (1.0 - 1.0 / 65536.0);
}

__device__ __forceinline__ D3 world_vertex(const double* Wd, float4 v) { return apply_matrix4_d(Wd, v.x, v.y, v.z); }
// The same from the float matrix, each entry widened to double where it is used (same values).
__device__ __forceinline__ D3 world_vertex_f(const float* W, float4 v) {
  MBVH_NO_CONTRACT
  const double x = v.x, y = v.y, z = v.z;
  const double w = 1 / (static_cast<double>(W[3]) * x + static_cast<double>(W[7]) * y + static_cast<double>(W[11]) * z +
                        static_cast<double>(W[15]));
  return d3((static_cast<double>(W[0]) * x + static_cast<double>(W[4]) * y + static_cast<double>(W[8]) * z +
             static_cast<double>(W[12])) * w,
            (static_cast<double>(W[1]) * x + static_cast<double>(W[5]) * y + static_cast<double>(W[9]) * z +
             static_cast<double>(W[13])) * w,
            (static_cast<double>(W[2]) * x + static_cast<double>(W[6]) * y + static_cast<double>(W[10]) * z +
             static_cast<double>(W[14])) * w);
}

// The point in an instance's frame, for box bounds measured there: for an affine W, the world
// distance to anything W puts in a box is at least sigma * (the local distance to the box), sigma the
// least singular value of W's 3x3 part; s2 = a lower bound of sigma^2 (Gershgorin on W^T W, shrunk for
// its FP32 roundings; just under 1 for a rotation), or 0 when that bound is not positive (then world
// boxes are used). d[k] bounds the FP32 error of each local coordinate; I is the float rounding of
// the double-precision inverse of the affine W, computed and residual-checked at scene creation (a
// caller-supplied inverse is never used for bounds).
struct LocalFrame {
  float l[3], d[3], s2;
  float s2u;  // an upper bound of the largest singular value squared (Gershgorin, grown for the roundings)
};
__device__ __forceinline__ LocalFrame local_frame(const float* W, const float* I, float px, float py, float pz) {
  LocalFrame f;
  for (int k = 0; k < 3; k++) {
    const float a = I[k], b = I[4 + k], c = I[8 + k], t = I[12 + k];
    f.l[k] = a * px + b * py + c * pz + t;
    f.d[k] = (fabsf(a * px) + fabsf(b * py) + fabsf(c * pz) + fabsf(t)) * (1.0f / 65536.0f);
  }
  float g[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      g[i][j] = W[4 * i] * W[4 * j] + W[4 * i + 1] * W[4 * j + 1] + W[4 * i + 2] * W[4 * j + 2];
  // each entry's FP32 error is a few ulps of |col i| |col j|: 2^-20 of that per off-diagonal entry
  float s2 = kInf, s2u = 0.0f;
  for (int i = 0; i < 3; i++) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    const float off = fabsf(g[i][j]) + fabsf(g[i][k]) +
                      (sqrtf(g[i][i] * g[j][j]) + sqrtf(g[i][i] * g[k][k])) * (1.0f / 1048576.0f);
    s2 = fminf(s2, g[i][i] * (1.0f - 1.0f / 65536.0f) - off * (1.0f + 1.0f / 65536.0f));
    s2u = fmaxf(s2u, (g[i][i] + off) * (1.0f + 1.0f / 65536.0f));
  }
  f.s2 = s2 > 0.0f ? s2 : 0.0f;
  f.s2u = s2u;
  return f;
}

// A lower bound of the squared world distance to anything W puts in the local box (frame f with s2 > 0).
__device__ __forceinline__ double local_box_d2_lb(const LocalFrame& f, float mnx, float mny, float mnz, float mxx,
                                                  float mxy, float mxz) {
  const float gx = fmaxf(fmaxf(mnx - f.l[0], f.l[0] - mxx) - f.d[0], 0.0f);
  const float gy = fmaxf(fmaxf(mny - f.l[1], f.l[1] - mxy) - f.d[1], 0.0f);
  const float gz = fmaxf(fmaxf(mnz - f.l[2], f.l[2] - mxz) - f.d[2], 0.0f);
  return static_cast<double>((gx * gx + gy * gy + gz * gz) * f.s2) * (1.0 - 1.0 / 16384.0);
}

struct ScenePointBest {
  double d2;
  uint32_t inst, face, slot;
  D3 p;
};

// One instance's tree walked near child first in world space; updates `b` (least squared distance,
// then lowest instance, then lowest face). Returns false on a stack overflow.
__device__ bool scene_point_instance(const MeshView& M, const float* W, const float* I, bool frame_ok, uint32_t inst,
                                     D3 p, float px, float py, float pz, StackL<8>& st, ScenePointBest& b) {
  double Wd[16];
  for (int k = 0; k < 16; k++) Wd[k] = W[k];
  const LocalFrame lf = local_frame(W, I, px, py, pz);
  const bool local = frame_ok && lf.s2 > 0.0f;
  auto box_lb = [&](float a, float bb, float c, float d, float e, float g) {
    return local ? local_box_d2_lb(lf, a, bb, c, d, e, g) : world_box_d2_lb(W, px, py, pz, a, bb, c, d, e, g);
  };
  // the FP32 local point and the absolute error of an FP32 point-to-triangle distance from it: the
  // distance's roundings (8u of the magnitudes) and the local coordinates' own error bounds
  const f3 lq = mk(lf.l[0], lf.l[1], lf.l[2]);
  const float tri_err = 8.0f * kU * (maxabs(lq) + *M.coord_max) + (lf.d[0] + lf.d[1] + lf.d[2]);
  const int base = st.sp;  // the caller's entries (the top level's) stay below
  for (uint32_t r = 0; r < M.n_roots; r++) {
    st.sp = base;
    uint32_t ref = M.roots[r];
    while (true) {
      if (ref & kRefLeaf) {
        uint32_t off, cnt;
        leaf_range(ref, M.bigleaf, off, cnt);
        for (uint32_t s = off; s < off + cnt; s++) {
          const Tri t = load_tri(M.tris, s);
          // the triangle's own box through the matrix first: a lower bound, so only a triangle that
          // could still win or tie gets the double-precision step
          if (box_lb(fminf(fminf(t.p0.x, t.p1.x), t.p2.x), fminf(fminf(t.p0.y, t.p1.y), t.p2.y),
                     fminf(fminf(t.p0.z, t.p1.z), t.p2.z), fmaxf(fmaxf(t.p0.x, t.p1.x), t.p2.x),
                     fmaxf(fmaxf(t.p0.y, t.p1.y), t.p2.y), fmaxf(fmaxf(t.p0.z, t.p1.z), t.p2.z)) > b.d2)
            continue;
          if (local) {
            // then the FP32 distance in the instance frame: sigma * (local distance less its error)
            // bounds the world distance from below (a NaN bound never skips)
            const f3 c = closest_on_triangle(lq, t.a(), t.b(), t.c());
            const float ex = lq.x - c.x, ey = lq.y - c.y, ez = lq.z - c.z;
            const float dl = __builtin_amdgcn_sqrtf(ex * ex + ey * ey + ez * ez);
            const float lb = dl - (tri_err + dl * (4.0f * kU));
            if (lb > 0.0f && static_cast<double>(lb * lb * lf.s2) * (1.0 - 1.0 / 16384.0) > b.d2) continue;
          }
          const D3 q = closest_on_tri_d(p, world_vertex(Wd, t.p0), world_vertex(Wd, t.p1), world_vertex(Wd, t.p2));
          const double d2 = ddist2(p, q);
          const uint32_t face = t.face();
          if (d2 < b.d2 || (d2 == b.d2 && (inst < b.inst || (inst == b.inst && face < b.face)))) {
            b.d2 = d2;
            b.inst = inst;
            b.face = face;
            b.slot = s;
            b.p = q;
          }
        }
      } else {
        const WideNode nd = M.nodes[ref];
        const double dl = box_lb(nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
        const double dr = box_lb(nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
        const bool hl = dl <= b.d2, hr = dr <= b.d2;
        if (hl && hr) {
          const bool left_first = dl <= dr;
          // the upper word of a non-negative double rounds it down: still a lower bound when popped
          if (!st.push(left_first ? nd.meta.y : nd.meta.x,
                       static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(left_first ? dr : dl)) >> 32)))
            return false;
          ref = left_first ? nd.meta.x : nd.meta.y;
          continue;
        }
        if (hl || hr) {
          ref = hl ? nd.meta.x : nd.meta.y;
          continue;
        }
      }
      ref = kDone;
      while (st.sp > base) {
        const uint2 e = st.pop();
        if (__longlong_as_double(static_cast<long long>(static_cast<uint64_t>(e.y) << 32)) <= b.d2) {
          ref = e.x;
          break;
        }
      }
      if (ref == kDone) break;
    }
  }
  return true;
}

// One lane per world point: the top level near child first (pruned by the best squared distance,
// equal bounds still visited so ties resolve the same way every time), each instance leaf walks
// that instance's tree. The top level's far children and the instance walks' share one stack (its
// LDS part small, so more waves fit), the instance walks above. Outputs: face (in the instance's
// geometry), instance, world distance, world point and the barycoords of the point in the world
// triangle (those of the local triangle too).
__global__ __launch_bounds__(kBlock) void k_scene_closest_point(SceneArgs A) {
  __shared__ uint2 lds_stack[8 * kBlock];
  const uint32_t tid = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + tid;
  StackL<8> st;
  st.lds = lds_stack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = tid;
  st.capacity = 8 + static_cast<int>(A.stack.ovf_entries);
  const float qnan = __int_as_float(0x7fc00000);
  for (uint64_t k = gtid; k < A.n; k += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t i = A.order ? A.order[k] : k;  // neighbouring lanes take neighbouring points
    const float px = A.origins[3 * i], py = A.origins[3 * i + 1], pz = A.origins[3 * i + 2];
    const D3 p = d3(px, py, pz);
    ScenePointBest b;
    b.d2 = __builtin_huge_val();
    b.inst = b.face = b.slot = kNoHit;
    b.p = d3(0, 0, 0);
    bool overflow = false;
    st.sp = 0;
    uint32_t node = 0;  // the root is always entered (an empty scene's is an empty leaf)
    while (!overflow) {
      const uint32_t* nw = A.tlas + 8ull * node;
      const uint32_t w7 = nw[7];
      if ((w7 >> 16) == 0xFFFFu) {
        const uint32_t off = nw[6], cnt = w7 & 0xFFFFu;
        for (uint32_t s = off; s < off + cnt && !overflow; s++) {
          const uint32_t inst = A.tlas_inst[s];
          overflow = !scene_point_instance(A.meshes[A.instance_geometry[inst]], A.world + 16ull * inst,
                                           A.frame + 16ull * inst, A.frame_ok[inst] != 0, inst, p, px, py, pz, st, b);
        }
        if (overflow) break;
      } else {
        const uint32_t left = node + 1, right = node + nw[6];
        const double dl = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * left), px, py, pz);
        const double dr = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * right), px, py, pz);
        const bool hl = dl <= b.d2, hr = dr <= b.d2;
        if (hl && hr) {
          const bool left_first = dl <= dr;
          // the upper word of a non-negative double rounds it down: still a lower bound when popped
          if (!st.push(left_first ? right : left,
                       static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(left_first ? dr : dl)) >> 32))) {
            overflow = true;
            break;
          }
          node = left_first ? left : right;
          continue;
        }
        if (hl || hr) {
          node = hl ? left : right;
          continue;
        }
      }
      node = kDone;
      while (st.sp > 0) {
        const uint2 e = st.pop();
        if (__longlong_as_double(static_cast<long long>(static_cast<uint64_t>(e.y) << 32)) <= b.d2) {
          node = e.x;
          break;
        }
      }
      if (node == kDone) break;
    }
    uint32_t face = overflow ? kOverflow : b.face;
    float dist = kInf, wx = qnan, wy = qnan, wz = qnan, bw = qnan, bu = qnan, bv = qnan;
    if (!overflow && b.inst != kNoHit) {
      // three's Triangle.getBarycoord of the point in the winning world triangle
      const MeshView& M = A.meshes[A.instance_geometry[b.inst]];
      double Wd[16];
      for (int k = 0; k < 16; k++) Wd[k] = A.world[16ull * b.inst + k];
      const Tri t = load_tri(M.tris, b.slot);
      const D3 a = world_vertex(Wd, t.p0), bb = world_vertex(Wd, t.p1), c = world_vertex(Wd, t.p2);
      const D3 v0 = dsub(c, a), v1 = dsub(bb, a), v2 = dsub(b.p, a);
      const double d00 = ddot(v0, v0), d01 = ddot(v0, v1), d02 = ddot(v0, v2), d11 = ddot(v1, v1), d12 = ddot(v1, v2);
      const double den = d00 * d11 - d01 * d01;
      if (den != 0) {
        const double inv = 1 / den, u = (d11 * d02 - d01 * d12) * inv, v = (d00 * d12 - d01 * d02) * inv;
        bw = static_cast<float>(1 - u - v);
        bu = static_cast<float>(v);
        bv = static_cast<float>(u);
      }
      dist = static_cast<float>(sqrt(b.d2));
      wx = static_cast<float>(b.p.x);
      wy = static_cast<float>(b.p.y);
      wz = static_cast<float>(b.p.z);
    }
    A.out_face[i] = face;
    if (A.out_instance) A.out_instance[i] = overflow ? kNoHit : b.inst;
    if (A.out_dist) A.out_dist[i] = dist;
    if (A.out_point) {
      A.out_point[3 * i] = wx;
      A.out_point[3 * i + 1] = wy;
      A.out_point[3 * i + 2] = wz;
    }
    if (A.out_bary) {
      A.out_bary[3 * i] = bw;
      A.out_bary[3 * i + 1] = bu;
      A.out_bary[3 * i + 2] = bv;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Scene closest point with four lanes per point (the mesh kernel's quad scheme): the quad walks the
// top level together (one far child pushed per level), and each instance's 4-wide tree with one
// child per lane, ranked and pushed by quad_children; a leaf's triangles are split over the lanes.
// The bounds and the decisions are those of the one-lane kernel (local-frame FP32 lower bounds,
// FP64 world distances, least (squared distance, instance, face) wins), so the results are the same
// bit for bit: a pruned child or triangle has a certified lower bound above a distance some
// triangle reaches. The best is reduced over the quad after every leaf, so the node steps see one
// bound.
// ------------------------------------------------------------------------------------------
struct ScenePick {
  double d2;
  uint32_t inst, face, slot;
};

__device__ __forceinline__ bool scene_pick_less(double d2, uint32_t inst, uint32_t face, const ScenePick& b) {
  return d2 < b.d2 || (d2 == b.d2 && (inst < b.inst || (inst == b.inst && face < b.face)));
}

template <int kCtrl>
__device__ __forceinline__ double quad_dpp_d(double v) {
  const uint64_t u = static_cast<uint64_t>(__double_as_longlong(v));
  const uint32_t lo = quad_dpp<kCtrl>(static_cast<uint32_t>(u)), hi = quad_dpp<kCtrl>(static_cast<uint32_t>(u >> 32));
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}

// The least pick of the quad, in every lane (the order is total, so the lanes agree).
__device__ __forceinline__ void quad_min_pick(ScenePick& b) {
  ScenePick o;
  o.d2 = quad_dpp_d<0xB1>(b.d2);
  o.inst = quad_xor1(b.inst);
  o.face = quad_xor1(b.face);
  o.slot = quad_xor1(b.slot);
  if (scene_pick_less(o.d2, o.inst, o.face, b)) b = o;
  o.d2 = quad_dpp_d<0x4E>(b.d2);
  o.inst = quad_xor2(b.inst);
  o.face = quad_xor2(b.face);
  o.slot = quad_xor2(b.slot);
  if (scene_pick_less(o.d2, o.inst, o.face, b)) b = o;
}

// Pops the nearest stacked entry whose lower bound (a float rounded down) is within the best.
__device__ __forceinline__ uint32_t quad_pop_within(QuadStack& st, int base, double best) {
  while (st.sp > base) {
    const uint2 e = st.pop();
    if (static_cast<double>(__uint_as_float(e.y)) <= best) return e.x;
  }
  return kDone;
}

// One instance's 4-wide tree walked by a quad in world space; updates `b` (quad-uniform on return).
// Returns false on a stack overflow.
__device__ __forceinline__ bool scene_point_instance4(const MeshView& M, const float* W, const float* I, bool frame_ok,
                                                      uint32_t inst, float px, float py, float pz, QuadStack& st,
                                                      uint32_t lane, ScenePick& b, float& ubw, uint32_t& steps,
                                                      uint32_t budget) {
  const uint32_t q4 = lane & 3u;
  const LocalFrame lf = local_frame(W, I, px, py, pz);
  const bool local = frame_ok && lf.s2 > 0.0f;
  auto box_lb = [&](float a, float bb, float c, float d, float e, float g) {
    return local ? local_box_d2_lb(lf, a, bb, c, d, e, g) : world_box_d2_lb(W, px, py, pz, a, bb, c, d, e, g);
  };
  const f3 lq = mk(lf.l[0], lf.l[1], lf.l[2]);
  const float tri_err = 8.0f * kU * (maxabs(lq) + *M.coord_max) + (lf.d[0] + lf.d[1] + lf.d[2]);
  const int base = st.sp;  // the top level's entries stay below
  bool overflow = false;
  for (uint32_t r = 0; r < M.n_roots; r++) {
    st.sp = base;
    uint32_t ref = M.roots4[r];
    while (true) {
      while (!(ref & kRefLeaf) && ref != kDone) {
        const float4 c0 = M.nodes4[8ull * ref + 2 * q4];
        const float4 c1 = M.nodes4[8ull * ref + 2 * q4 + 1];
        const uint32_t cref = __float_as_uint(c1.z);
        bool in = cref != kDone;
        float key = kInf;
        if (in) {
          const double lb = box_lb(c0.x, c0.y, c0.z, c0.w, c1.x, c1.y);
          in = lb <= fmin(b.d2, static_cast<double>(ubw));
          key = __double2float_rd(lb);  // a lower bound still, for the test when popped
        }
        ref = quad_children(st, lane, in, key, cref, overflow);
        if (overflow) return false;
        if (ref == kDone) ref = quad_pop_within(st, base, fmin(b.d2, static_cast<double>(ubw)));
      }
      while (ref & kRefLeaf) {
        uint32_t off, cnt;
        leaf_range(ref, M.bigleaf, off, cnt);
        for (uint32_t s = off + q4; s < off + cnt; s += 4) {
          const Tri t = load_tri(M.tris, s);
          if (box_lb(fminf(fminf(t.p0.x, t.p1.x), t.p2.x), fminf(fminf(t.p0.y, t.p1.y), t.p2.y),
                     fminf(fminf(t.p0.z, t.p1.z), t.p2.z), fmaxf(fmaxf(t.p0.x, t.p1.x), t.p2.x),
                     fmaxf(fmaxf(t.p0.y, t.p1.y), t.p2.y), fmaxf(fmaxf(t.p0.z, t.p1.z), t.p2.z)) >
              fmin(b.d2, static_cast<double>(ubw)))
            continue;
          if (local) {
            // sigma_min * (local distance less its error) <= world distance <= sigma_max * (local
            // distance plus its error): the upper bound joins ubw (a NaN leaves it), and a triangle
            // whose lower bound is above min(best, ubw) can neither win nor tie
            const f3 c = closest_on_triangle(lq, t.a(), t.b(), t.c());
            const float ex = lq.x - c.x, ey = lq.y - c.y, ez = lq.z - c.z;
            const float dl = __builtin_amdgcn_sqrtf(ex * ex + ey * ey + ez * ez);
            const float e = tri_err + dl * (4.0f * kU);
            const float hi = dl + e;
            ubw = fminf(ubw, hi * hi * lf.s2u * (1.0f + 1.0f / 16384.0f));
            const float lb = dl - e;
            if (lb > 0.0f && static_cast<double>(lb * lb * lf.s2) * (1.0 - 1.0 / 16384.0) >
                                 fmin(b.d2, static_cast<double>(ubw)))
              continue;
          }
          // the matrix and the point widened again here rather than held in registers through the walk
          const D3 p = d3(px, py, pz);
          const D3 qd = closest_on_tri_d(p, world_vertex_f(W, t.p0), world_vertex_f(W, t.p1), world_vertex_f(W, t.p2));
          const double d2 = ddist2(p, qd);
          const uint32_t face = t.face();
          if (scene_pick_less(d2, inst, face, b)) {
            b.d2 = d2;
            b.inst = inst;
            b.face = face;
            b.slot = s;
          }
        }
        quad_min_pick(b);
        ubw = fminf(ubw, quad_xor1(ubw));
        ubw = fminf(ubw, quad_xor2(ubw));
        // past the step budget the walk ends here (quad-uniform) and the point goes to a wave of its own
        ref = ++steps > budget ? kDone : quad_pop_within(st, base, fmin(b.d2, static_cast<double>(ubw)));
      }
      if (ref == kDone) break;
    }
    if (steps > budget) break;
  }
  return !overflow;
}

__device__ __forceinline__ void scene_point_out(const SceneArgs& A, uint64_t i, bool overflow, ScenePick b, D3 p);

__global__ __launch_bounds__(kBlock, 4) void k_scene_closest_point4(SceneArgs A) {
  __shared__ uint2 qstack[kQuadLds * (kBlock / 4)];
  const uint32_t lane = threadIdx.x, q4 = lane & 3u, wq = (lane & 63u) >> 2, qid = lane >> 2;
  QuadStack st;
  st.lds = qstack;
  st.ovf = A.stack.ovf;
  st.ovf_stride = static_cast<uint64_t>(gridDim.x) * (kBlock / 4);
  st.gquad = static_cast<uint64_t>(blockIdx.x) * (kBlock / 4) + qid;
  st.qid = qid;
  st.capacity = kQuadLds + static_cast<int>(A.stack.ovf_entries);
  while (true) {
    uint32_t tile = 0;
    if ((lane & 63u) == 0) tile = atomicAdd(A.queue, 1u);
    tile = __shfl(tile, 0, 64);
    if (static_cast<uint64_t>(tile) * 16 >= A.n) break;  // every wave reaches this exit
    const uint64_t visit = static_cast<uint64_t>(tile) * 16 + wq;
    if (visit >= A.n) continue;
    const uint64_t i = A.order ? A.order[visit] : visit;
    const float px = A.origins[3 * i], py = A.origins[3 * i + 1], pz = A.origins[3 * i + 2];
    ScenePick b = {__builtin_huge_val(), kNoHit, kNoHit, kNoHit};
    float ubw = kInf;  // an upper bound of the least world squared distance, from FP32 distances
    bool overflow = false;
    uint32_t steps = 0;
    const uint32_t budget = A.heavy ? A.heavy_budget : 0xFFFFFFFFu;
    st.sp = 0;
    uint32_t node = 0;  // the root is always entered (an empty scene's is an empty leaf)
    while (true) {
      const uint32_t* nw = A.tlas + 8ull * node;
      const uint32_t w7 = nw[7];
      if ((w7 >> 16) == 0xFFFFu) {
        const uint32_t off = nw[6], cnt = w7 & 0xFFFFu;
        for (uint32_t s = off; s < off + cnt && !overflow && steps <= budget; s++) {
          const uint32_t inst = A.tlas_inst[s];
          overflow = !scene_point_instance4(A.meshes[A.instance_geometry[inst]], A.world + 16ull * inst,
                                            A.frame + 16ull * inst, A.frame_ok[inst] != 0, inst, px, py, pz, st,
                                            lane, b, ubw, steps, budget);
        }
        if (overflow || steps > budget) break;
      } else {
        const uint32_t left = node + 1, right = node + nw[6];
        const double dl = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * left), px, py, pz);
        const double dr = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * right), px, py, pz);
        const double bnd = fmin(b.d2, static_cast<double>(ubw));
        const bool hl = dl <= bnd, hr = dr <= bnd;
        if (hl && hr) {
          const bool left_first = dl <= dr;
          if (!st.push(left_first ? right : left, __float_as_uint(__double2float_rd(left_first ? dr : dl)))) {
            overflow = true;
            break;
          }
          node = left_first ? left : right;
          continue;
        }
        if (hl || hr) {
          node = hl ? left : right;
          continue;
        }
      }
      node = quad_pop_within(st, 0, fmin(b.d2, static_cast<double>(ubw)));
      if (node == kDone) break;
    }
    if (q4 == 0) {
      if (!overflow && steps > budget)
        A.heavy[atomicAdd(A.heavy_count, 1u)] = static_cast<uint32_t>(i);
      else
        scene_point_out(A, i, overflow, b, d3(px, py, pz));
    }
  }
}

// ------------------------------------------------------------------------------------------
// Scene closest point, one wave per point, for the points the four-lane kernel left (those whose
// walks are long: a point near the centre of a closed mesh is close to every triangle). The wave
// walks the top level together; in each instance it keeps an LDS frontier of 4-wide nodes: every
// round the lanes take the top 64 entries (those whose bound is still within the best), expand
// nodes (children within the best pushed at offsets from a prefix sum over the wave) or test a
// leaf's triangles, and the best is reduced over the wave. The frontier, taken 64 at a time from
// the top, holds at most 192 entries per level of the tree plus 256. Same bounds, same decisions
// and the same order of picks as the one-lane kernel: the same results bit for bit.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void wave_min_pick(ScenePick& b) {
  for (int m = 1; m < 64; m <<= 1) {
    ScenePick o;
    o.d2 = __shfl_xor(b.d2, m, 64);
    o.inst = __shfl_xor(b.inst, m, 64);
    o.face = __shfl_xor(b.face, m, 64);
    o.slot = __shfl_xor(b.slot, m, 64);
    if (scene_pick_less(o.d2, o.inst, o.face, b)) b = o;
  }
}

__device__ __forceinline__ float wave_min_f(float v) {
  for (int m = 1; m < 64; m <<= 1) v = fminf(v, __shfl_xor(v, m, 64));
  return v;
}

// Exclusive prefix sum of c over the wave; *total receives the sum.
__device__ __forceinline__ uint32_t wave_prefix(uint32_t c, uint32_t lane, uint32_t* total) {
  uint32_t x = c;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= static_cast<uint32_t>(d)) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - c;
}

__global__ __launch_bounds__(64) void k_scene_closest_point_wide(SceneArgs A) {
  __shared__ uint2 front[kWideFrontier];
  __shared__ uint2 top[kTlasStack];
  const uint32_t lane = threadIdx.x;
  const uint32_t nh = *A.heavy_count;
  for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
    const uint64_t i = A.heavy[h];
    const float px = A.origins[3 * i], py = A.origins[3 * i + 1], pz = A.origins[3 * i + 2];
    const D3 p = d3(px, py, pz);
    ScenePick b = {__builtin_huge_val(), kNoHit, kNoHit, kNoHit};
    float ubw = kInf;
    bool overflow = false;
    int tsp = 0;
    uint32_t node = 0;
    while (true) {  // the top level, the same in every lane
      const uint32_t* nw = A.tlas + 8ull * node;
      const uint32_t w7 = nw[7];
      if ((w7 >> 16) == 0xFFFFu) {
        const uint32_t off = nw[6], cnt = w7 & 0xFFFFu;
        for (uint32_t ts = off; ts < off + cnt && !overflow; ts++) {
          const uint32_t inst = A.tlas_inst[ts];
          const MeshView& M = A.meshes[A.instance_geometry[inst]];
          const float* W = A.world + 16ull * inst;
          const LocalFrame lf = local_frame(W, A.frame + 16ull * inst, px, py, pz);
          const bool local = A.frame_ok[inst] != 0 && lf.s2 > 0.0f;
          const f3 lq = mk(lf.l[0], lf.l[1], lf.l[2]);
          const float tri_err = 8.0f * kU * (maxabs(lq) + *M.coord_max) + (lf.d[0] + lf.d[1] + lf.d[2]);
          for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
            if (lane == 0) front[0] = make_uint2(M.roots4[r], 0u);
            __syncthreads();
            uint32_t fc = 1;
            while (fc > 0) {
              const uint32_t take = fc < 64u ? fc : 64u;
              const uint32_t fb = fc - take;
              const uint2 e = lane < take ? front[fb + lane] : make_uint2(kDone, 0u);
              __syncthreads();  // every lane has read its entry before the pushes overwrite them
              const double bnd = fmin(b.d2, static_cast<double>(ubw));
              uint32_t kids[4];
              float keys[4];
              uint32_t nk = 0;
              if (e.x != kDone && static_cast<double>(__uint_as_float(e.y)) <= bnd) {
                if (!(e.x & kRefLeaf)) {
                  for (uint32_t k = 0; k < 4; k++) {
                    const float4 c0 = M.nodes4[8ull * e.x + 2 * k];
                    const float4 c1 = M.nodes4[8ull * e.x + 2 * k + 1];
                    const uint32_t cref = __float_as_uint(c1.z);
                    if (cref == kDone) continue;
                    const double lb = local ? local_box_d2_lb(lf, c0.x, c0.y, c0.z, c0.w, c1.x, c1.y)
                                            : world_box_d2_lb(W, px, py, pz, c0.x, c0.y, c0.z, c0.w, c1.x, c1.y);
                    if (lb <= bnd) {
                      kids[nk] = cref;
                      keys[nk] = __double2float_rd(lb);
                      nk++;
                    }
                  }
                } else {
                  uint32_t off2, cnt2;
                  leaf_range(e.x, M.bigleaf, off2, cnt2);
                  for (uint32_t s = off2; s < off2 + cnt2; s++) {
                    const Tri t = load_tri(M.tris, s);
                    const float mnx = fminf(fminf(t.p0.x, t.p1.x), t.p2.x), mny = fminf(fminf(t.p0.y, t.p1.y), t.p2.y);
                    const float mnz = fminf(fminf(t.p0.z, t.p1.z), t.p2.z), mxx = fmaxf(fmaxf(t.p0.x, t.p1.x), t.p2.x);
                    const float mxy = fmaxf(fmaxf(t.p0.y, t.p1.y), t.p2.y), mxz = fmaxf(fmaxf(t.p0.z, t.p1.z), t.p2.z);
                    const double tb = local ? local_box_d2_lb(lf, mnx, mny, mnz, mxx, mxy, mxz)
                                            : world_box_d2_lb(W, px, py, pz, mnx, mny, mnz, mxx, mxy, mxz);
                    if (tb > fmin(b.d2, static_cast<double>(ubw))) continue;
                    if (local) {
                      const f3 c = closest_on_triangle(lq, t.a(), t.b(), t.c());
                      const float ex = lq.x - c.x, ey = lq.y - c.y, ez = lq.z - c.z;
                      const float dl = __builtin_amdgcn_sqrtf(ex * ex + ey * ey + ez * ez);
                      const float er = tri_err + dl * (4.0f * kU);
                      const float hi = dl + er;
                      ubw = fminf(ubw, hi * hi * lf.s2u * (1.0f + 1.0f / 16384.0f));
                      const float lb = dl - er;
                      if (lb > 0.0f && static_cast<double>(lb * lb * lf.s2) * (1.0 - 1.0 / 16384.0) >
                                           fmin(b.d2, static_cast<double>(ubw)))
                        continue;
                    }
                    const D3 qd = closest_on_tri_d(p, world_vertex_f(W, t.p0), world_vertex_f(W, t.p1),
                                                   world_vertex_f(W, t.p2));
                    const double d2 = ddist2(p, qd);
                    const uint32_t face = t.face();
                    if (scene_pick_less(d2, inst, face, b)) {
                      b.d2 = d2;
                      b.inst = inst;
                      b.face = face;
                      b.slot = s;
                    }
                  }
                }
              }
              uint32_t total;
              const uint32_t at = fb + wave_prefix(nk, lane, &total);
              if (fb + total > kWideFrontier) {
                overflow = true;  // wave-uniform (total and fb are)
                break;
              }
              for (uint32_t k = 0; k < nk; k++) front[at + k] = make_uint2(kids[k], __float_as_uint(keys[k]));
              __syncthreads();
              fc = fb + total;
              wave_min_pick(b);
              ubw = wave_min_f(ubw);
            }
          }
        }
        if (overflow) break;
      } else {
        const uint32_t left = node + 1, right = node + nw[6];
        const double dl = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * left), px, py, pz);
        const double dr = tlas_box_d2_lb(reinterpret_cast<const float*>(A.tlas + 8ull * right), px, py, pz);
        const double bnd = fmin(b.d2, static_cast<double>(ubw));
        const bool hl = dl <= bnd, hr = dr <= bnd;
        if (hl && hr) {
          const bool left_first = dl <= dr;
          if (tsp >= static_cast<int>(kTlasStack)) {
            overflow = true;
            break;
          }
          if (lane == 0) top[tsp] = make_uint2(left_first ? right : left, __float_as_uint(__double2float_rd(left_first ? dr : dl)));
          tsp++;
          node = left_first ? left : right;
          __syncthreads();
          continue;
        }
        if (hl || hr) {
          node = hl ? left : right;
          continue;
        }
      }
      node = kDone;
      const double bnd = fmin(b.d2, static_cast<double>(ubw));
      while (tsp > 0) {
        const uint2 e = top[--tsp];
        if (static_cast<double>(__uint_as_float(e.y)) <= bnd) {
          node = e.x;
          break;
        }
      }
      if (node == kDone) break;
    }
    if (lane == 0) scene_point_out(A, i, overflow, b, p);
    __syncthreads();  // the next point's entries overwrite this one's
  }
}

// Outputs of the four-lane kernel for point i (one lane of the quad).
__device__ __forceinline__ void scene_point_out(const SceneArgs& A, uint64_t i, bool overflow, ScenePick b, D3 p) {
  const float qnan = __int_as_float(0x7fc00000);
  {
    uint32_t face = overflow ? kOverflow : b.face;
    float dist = kInf, wx = qnan, wy = qnan, wz = qnan, bw = qnan, bu = qnan, bv = qnan;
    if (!overflow && b.inst != kNoHit) {
      // the winner's world point again (the same double-precision steps), then three's
      // Triangle.getBarycoord of the point in the world triangle
      const MeshView& M = A.meshes[A.instance_geometry[b.inst]];
      double Wd[16];
      for (int k = 0; k < 16; k++) Wd[k] = A.world[16ull * b.inst + k];
      const Tri t = load_tri(M.tris, b.slot);
      const D3 a = world_vertex(Wd, t.p0), bb = world_vertex(Wd, t.p1), c = world_vertex(Wd, t.p2);
      const D3 wp = closest_on_tri_d(p, a, bb, c);
      const D3 v0 = dsub(c, a), v1 = dsub(bb, a), v2 = dsub(wp, a);
      const double d00 = ddot(v0, v0), d01 = ddot(v0, v1), d02 = ddot(v0, v2), d11 = ddot(v1, v1), d12 = ddot(v1, v2);
      const double den = d00 * d11 - d01 * d01;
      if (den != 0) {
        const double inv = 1 / den, u = (d11 * d02 - d01 * d12) * inv, v = (d00 * d12 - d01 * d02) * inv;
        bw = static_cast<float>(1 - u - v);
        bu = static_cast<float>(v);
        bv = static_cast<float>(u);
      }
      dist = static_cast<float>(sqrt(b.d2));
      wx = static_cast<float>(wp.x);
      wy = static_cast<float>(wp.y);
      wz = static_cast<float>(wp.z);
    }
    A.out_face[i] = face;
    if (A.out_instance) A.out_instance[i] = overflow ? kNoHit : b.inst;
    if (A.out_dist) A.out_dist[i] = dist;
    if (A.out_point) {
      A.out_point[3 * i] = wx;
      A.out_point[3 * i + 1] = wy;
      A.out_point[3 * i + 2] = wz;
    }
    if (A.out_bary) {
      A.out_bary[3 * i] = bw;
      A.out_bary[3 * i + 1] = bu;
      A.out_bary[3 * i + 2] = bv;
    }
  }
}

__global__ void k_apply_side_mask(float4* tris, uint64_t n_slots, const uint8_t* mask) {
  const uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  const uint32_t face = __float_as_uint(tris[3 * s].w);
  tris[3 * s + 1].w = __uint_as_float(static_cast<uint32_t>(mask[face]) & 3u);
}

}  // namespace

uint64_t launch_threads(uint64_t n, uint64_t max_threads) {
  uint64_t t = (n + kBlock - 1) / kBlock * kBlock;
  if (t > max_threads) t = max_threads / kBlock * kBlock;
  if (t == 0) t = kBlock;
  return t;
}

hipError_t launch_raycast_first4(const RaycastFirstArgs& a, uint64_t quads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 grid(static_cast<uint32_t>(quads / (kBlock / 4)));
  if (a.origins_lo || a.directions_lo)
    hipLaunchKernelGGL(k_raycast_first4<true>, grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_raycast_first4<false>, grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_raycast_first(const RaycastFirstArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 grid(static_cast<uint32_t>(threads / kBlock));
  if (a.origins_lo || a.directions_lo)
    hipLaunchKernelGGL(k_raycast_first<true>, grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_raycast_first<false>, grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_instance_rays(const InstanceRayArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_instance_rays, dim3(static_cast<uint32_t>((a.n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_instance_hits(const InstanceHitArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_instance_hits, dim3(static_cast<uint32_t>((a.n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_batch_raycast_first(const BatchArgs& a, uint32_t stack_lds, uint32_t lanes, uint32_t blocks,
                                      hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.counter);
  if (a.n == 0) return hipGetLastError();
  const bool local = a.inverse == nullptr;  // rays already in their mesh's frame (hi + lo pairs)
  if (lanes == 4 && local)
    hipLaunchKernelGGL(k_batch_raycast_first4<true>, dim3(blocks), dim3(kBlock), 0, s, a);
  else if (lanes == 4)
    hipLaunchKernelGGL(k_batch_raycast_first4<false>, dim3(blocks), dim3(kBlock), 0, s, a);
  else if (stack_lds == 8 && local)
    hipLaunchKernelGGL((k_batch_raycast_first<8, true>), dim3(blocks), dim3(kBlock), 0, s, a);
  else if (stack_lds == 8)
    hipLaunchKernelGGL((k_batch_raycast_first<8, false>), dim3(blocks), dim3(kBlock), 0, s, a);
  else if (local)
    hipLaunchKernelGGL((k_batch_raycast_first<16, true>), dim3(blocks), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_batch_raycast_first<16, false>), dim3(blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scene_raycast_first(const SceneArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scene_raycast_first, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_intersects_box(const BoxArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_intersects_box, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scene_closest_point(const SceneArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scene_closest_point, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scene_closest_point4(const SceneArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.queue);
  if (a.heavy) hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.heavy_count);
  hipLaunchKernelGGL(k_scene_closest_point4, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scene_closest_point_wide(const SceneArgs& a, uint32_t waves, hipStream_t s) {
  if (a.n == 0 || !a.heavy) return hipSuccess;
  hipLaunchKernelGGL(k_scene_closest_point_wide, dim3(waves), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_closest_to_geometry(const GeometryArgs& a, uint64_t threads, hipStream_t s) {
  hipLaunchKernelGGL(k_geometry_clear, dim3(1), dim3(1), 0, s, a);
  if (a.n_other) {
    const dim3 grid(static_cast<uint32_t>(threads / kBlock));
    hipLaunchKernelGGL(k_geometry_seed, grid, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_closest_to_geometry, grid, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_geometry_pick, grid, dim3(kBlock), 0, s, a);
  }
  hipLaunchKernelGGL(k_geometry_gather, dim3(1), dim3(1), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_closest_point4(const ClosestPointArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.queue);
  const dim3 grid(static_cast<uint32_t>(threads / kBlock));
  if (a.points_lo)
    hipLaunchKernelGGL(k_closest_point4<true>, grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_closest_point4<false>, grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

// intersectsSphere (src/core/MeshBVH.js:800-809, sphereIntersectTriangle src/math/MathUtilities.js:160-204):
// any triangle within `radius` of the centre, the comparison inclusive. The closest triangle (found
// within the largest radius by the closest-point kernels) decides it: its squared distance in
// double-float against radius^2 (exact as two_prod of the float radius).
__global__ __launch_bounds__(kBlock) void k_sphere_decide(SphereArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < a.n; i += stride) {
    const uint32_t f = a.face[i];
    const float r = a.radii[i];
    uint8_t res = 0;
    if (f == kOverflow) {
      res = 2;
    } else if (f != kNoHit && r >= 0.0f && f < a.mesh.face_limit) {
      const uint32_t slot = a.mesh.face_slot ? a.mesh.face_slot[f] : f;
      if (slot != kNoHit) {
        const Tri tr = load_tri(a.mesh.tris, slot);
        const f3 c = mk(a.centers[3 * i + 0], a.centers[3 * i + 1], a.centers[3 * i + 2]);
        const ClosestFF cf = closest_on_triangle_ff<false>(c, mk(0.0f, 0.0f, 0.0f), tr.a(), tr.b(), tr.c());
        res = (cf.dist2.hi == cf.dist2.hi) && ff_sign(ff_sub(cf.dist2, two_prod(r, r))) <= 0 ? 1 : 0;
      }
    }
    a.out[i] = res;
  }
}

hipError_t launch_sphere_decide(const SphereArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const uint64_t blocks = (a.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_sphere_decide, dim3(static_cast<uint32_t>(blocks < 16384 ? blocks : 16384)), dim3(kBlock), 0, s,
                     a);
  return hipGetLastError();
}

hipError_t launch_closest_point_packet(const ClosestPointArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0 || a.tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.queue);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.redo_count);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.pool_count);
  const dim3 grid(static_cast<uint32_t>(threads / kBlock));
  const dim3 rgrid(static_cast<uint32_t>((a.tiles * 64 + kBlock - 1) / kBlock));
  if (a.points_lo) {
    hipLaunchKernelGGL(k_closest_point_packet<true>, grid, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_closest_point_resolve<true>, rgrid, dim3(kBlock), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_closest_point_packet<false>, grid, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_closest_point_resolve<false>, rgrid, dim3(kBlock), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_closest_point(const ClosestPointArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.queue);
  const dim3 grid(static_cast<uint32_t>(threads / kBlock));
  if (a.points_lo)
    hipLaunchKernelGGL(k_closest_point<true>, grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_closest_point<false>, grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_raycast_all(const RaycastAllArgs& a, bool fill, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  const dim3 grid(static_cast<uint32_t>(threads / kBlock)), block(kBlock);
  const bool lo = a.origins_lo || a.directions_lo;
  if (fill && lo)
    hipLaunchKernelGGL((k_raycast_all<true, true>), grid, block, 0, s, a);
  else if (fill)
    hipLaunchKernelGGL((k_raycast_all<true, false>), grid, block, 0, s, a);
  else if (lo)
    hipLaunchKernelGGL((k_raycast_all<false, true>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_raycast_all<false, false>), grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_apply_side_mask(float4* tris, uint64_t n_slots, const uint8_t* mask, hipStream_t s) {
  if (n_slots == 0) return hipSuccess;
  const uint32_t blocks = static_cast<uint32_t>((n_slots + 255) / 256);
  hipLaunchKernelGGL(k_apply_side_mask, dim3(blocks), dim3(256), 0, s, tris, n_slots, mask);
  return hipGetLastError();
}

}  // namespace mbvh
