// Wave64 traversal kernels for gfx950: first-hit raycast and closest point.
//
// One query per lane. Each lane walks the uploaded tree with a private stack whose first
// kStackLds entries live in LDS (column per lane, so the 64 lanes of a wave hit 64 distinct banks)
// and whose deeper entries spill to a global overflow region sized from the tree depth at upload.
// All arithmetic is FP32; triangles whose FP32 decision lies inside a rigorous error bound are
// re-decided in double-float (see mbvh_device.h), and the winning triangle of each query is
// re-evaluated in double-float for the reported distance / point / barycoord.
#include "mbvh_kernels.h"

namespace mbvh {
namespace {

constexpr float kU = 5.9604644775390625e-08f;  // unit roundoff 2^-24
constexpr float kSlabEps = 8.0f * kU;          // widening of FP32 slab intervals (>= 3 roundings)
constexpr uint32_t kNoHit = 0xFFFFFFFFu;
constexpr uint32_t kPruneFlag = 0x80000000u;

__device__ __forceinline__ float lerp_bound_lo(float t) { return t * (t > 0.0f ? 1.0f - kSlabEps : 1.0f + kSlabEps); }
__device__ __forceinline__ float lerp_bound_hi(float t) { return t * (t > 0.0f ? 1.0f + kSlabEps : 1.0f - kSlabEps); }

// Stack with an LDS part and a global overflow part.
struct LaneStack {
  uint32_t* s_ref;
  uint32_t* s_aux;
  uint2* ovf;
  uint64_t ovf_stride;
  uint64_t gtid;
  uint32_t lane;
  uint32_t capacity;  // kStackLds + overflow entries
  int sp;

  __device__ __forceinline__ bool push(uint32_t ref, uint32_t aux) {
    if (sp >= static_cast<int>(capacity)) return false;
    if (sp < kStackLds) {
      s_ref[sp * kBlock + lane] = ref;
      s_aux[sp * kBlock + lane] = aux;
    } else {
      ovf[static_cast<uint64_t>(sp - kStackLds) * ovf_stride + gtid] = make_uint2(ref, aux);
    }
    sp++;
    return true;
  }
  __device__ __forceinline__ void pop(uint32_t& ref, uint32_t& aux) {
    sp--;
    if (sp < kStackLds) {
      ref = s_ref[sp * kBlock + lane];
      aux = s_aux[sp * kBlock + lane];
    } else {
      const uint2 e = ovf[static_cast<uint64_t>(sp - kStackLds) * ovf_stride + gtid];
      ref = e.x;
      aux = e.y;
    }
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

__device__ __forceinline__ void load_tri(const float4* tris, uint32_t slot, f3& a, f3& b, f3& c,
                                         uint32_t& face, uint32_t& mask) {
  const float4 p0 = tris[3u * slot + 0u];
  const float4 p1 = tris[3u * slot + 1u];
  const float4 p2 = tris[3u * slot + 2u];
  a = mk(p0.x, p0.y, p0.z);
  b = mk(p1.x, p1.y, p1.z);
  c = mk(p2.x, p2.y, p2.z);
  face = __float_as_uint(p0.w);
  mask = __float_as_uint(p1.w);
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
};

__device__ __forceinline__ bool slab(const RayBox& r, float mnx, float mny, float mnz, float mxx,
                                     float mxy, float mxz, float near_t, float far_t, float& tmin) {
  const float tx0 = ((r.sx ? mnx : mxx) - r.ox) * r.ix;
  const float tx1 = ((r.sx ? mxx : mnx) - r.ox) * r.ix;
  const float ty0 = ((r.sy ? mny : mxy) - r.oy) * r.iy;
  const float ty1 = ((r.sy ? mxy : mny) - r.oy) * r.iy;
  const float tz0 = ((r.sz ? mnz : mxz) - r.oz) * r.iz;
  const float tz1 = ((r.sz ? mxz : mnz) - r.oz) * r.iz;
  const float t0 = lerp_bound_lo(fmaxf(fmaxf(tx0, ty0), tz0));
  const float t1 = lerp_bound_hi(fminf(fminf(tx1, ty1), tz1));
  tmin = t0;
  return (t0 <= t1) && (t0 <= far_t) && (t1 >= near_t);
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

__device__ TriTerms tri_terms_exact(f3 o, f3 d, f3 a, f3 b, f3 c) {
  TriTerms r;
  const ff3 E1 = ff3_diff(b, a), E2 = ff3_diff(c, a), Q = ff3_diff(o, a);
  const ff3 N = ff3_cross(E1, E2);
  const ff ddn = ff3_dotf(d, N);
  const int sg = ff_sign(ddn);
  r.ok = sg != 0;
  r.back = sg > 0;
  const bool neg = !r.back;  // sign = -1 for front-facing
  r.ddn_abs = neg ? ff_neg(ddn) : ddn;
  ff e0 = ff3_dotf(d, ff3_cross(Q, E2));
  ff e1 = ff3_dotf(d, ff3_cross(E1, Q));
  ff qn = ff3_dot(Q, N);
  r.e0 = neg ? ff_neg(e0) : e0;
  r.e1 = neg ? ff_neg(e1) : e1;
  r.qdn = neg ? qn : ff_neg(qn);  // QdN = -sign * Q.N
  return r;
}

__device__ __noinline__ bool hit_triangle_exact(f3 o, f3 d, float dlen, f3 a, f3 b, f3 c,
                                                uint32_t mask, float near_d, float far_d,
                                                float& t_out, float& et_out) {
  const TriTerms r = tri_terms_exact(o, d, a, b, c);
  if (!r.ok) return false;
  if (!(mask & (r.back ? 2u : 1u))) return false;
  if (ff_sign(r.e0) < 0) return false;
  if (ff_sign(r.e1) < 0) return false;
  if (ff_sign(ff_sub(r.ddn_abs, ff_add(r.e0, r.e1))) < 0) return false;
  if (ff_sign(r.qdn) < 0) return false;
  const float t = ff_to_float(ff_div(r.qdn, r.ddn_abs));
  const float dist = t * dlen;
  if (dist < near_d || dist > far_d) return false;
  t_out = t;
  et_out = 2.0f * kU * t;
  return true;
}

// t of a triangle the ray is known to hit, in double-float (for near-tie comparisons).
__device__ __noinline__ ff exact_hit_t(f3 o, f3 d, const float4* tris, uint32_t slot) {
  const float4 p0 = tris[3u * slot + 0u];
  const float4 p1 = tris[3u * slot + 1u];
  const float4 p2 = tris[3u * slot + 2u];
  const TriTerms r = tri_terms_exact(o, d, mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z));
  return ff_div(r.qdn, r.ddn_abs);
}

// On a hit, t_out is the FP32 ray parameter and et_out a bound on its error.
__device__ __forceinline__ bool hit_triangle(f3 o, f3 d, float dl1, float dlen, f3 a, f3 b, f3 c,
                                             uint32_t mask, float near_d, float far_d,
                                             float& t_out, float& et_out) {
  const f3 e1 = sub(b, a), e2 = sub(c, a);
  const f3 n = cross(e1, e2);
  const float ddn = dot(d, n);
  const float mE = fmaxf(maxabs(e1), maxabs(e2));
  const float bN = 32.0f * kU * dl1 * mE * mE;
  if (!(fabsf(ddn) > bN)) return hit_triangle_exact(o, d, dlen, a, b, c, mask, near_d, far_d, t_out, et_out);
  const bool back = ddn > 0.0f;
  if (!(mask & (back ? 2u : 1u))) return false;
  const float s = back ? 1.0f : -1.0f;
  const float DdN = fabsf(ddn);
  const f3 q = sub(o, a);
  const float mQ = maxabs(q);
  const float bE = 32.0f * kU * dl1 * mQ * mE;
  const float e0 = s * dot(d, cross(q, e2));
  if (e0 < -bE) return false;
  const float f1 = s * dot(d, cross(e1, q));
  if (f1 < -bE) return false;
  const float g = DdN - e0 - f1;
  const float bS = bN + 2.0f * bE + 4.0f * kU * (DdN + fabsf(e0) + fabsf(f1));
  if (g < -bS) return false;
  const float qdn = -s * dot(q, n);
  const float bQ = 64.0f * kU * mQ * mE * mE;
  if (qdn < -bQ) return false;
  if (!(e0 > bE && f1 > bE && g > bS && qdn > bQ))
    return hit_triangle_exact(o, d, dlen, a, b, c, mask, near_d, far_d, t_out, et_out);
  const float t = qdn / DdN;
  const float dist = t * dlen;
  if (dist < near_d || dist > far_d) return false;
  t_out = t;
  et_out = t * (bQ / qdn + bN / DdN + 4.0f * kU);
  return true;
}

// ------------------------------------------------------------------------------------------
// First hit. Traversal order and tie rules of _raycastFirst
// (src/core/cast/raycastFirst.template.js:19-103): at each node the child on the side the ray
// comes from (dir[splitAxis] >= 0 ⇒ left) is finished first; inside a leaf the first closest
// triangle wins (strict <, iterationUtils.template.js:24-53); across children and roots the
// earlier-visited hit wins ties. The reference has no distance pruning; this kernel skips
// subtrees whose entry distance already exceeds the best hit, which cannot change that result.
// Its split-plane shortcut (raycastFirst.template.js:69-84) is implied by the distance pruning
// whenever dir[splitAxis] != 0. Rays with a zero direction component are traced without distance
// pruning and with that shortcut applied literally: on such an axis the near child's hit point
// lies in the ray's own plane, so the shortcut fires iff the near subtree found any hit and
// origin[axis] <= far.min[axis].
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_raycast_first(RaycastFirstArgs A) {
  __shared__ uint32_t s_ref[kStackLds * kBlock];
  __shared__ uint32_t s_aux[kStackLds * kBlock];

  const uint32_t lane = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + lane;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kBlock;

  LaneStack st;
  st.s_ref = s_ref;
  st.s_aux = s_aux;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = lane;
  st.capacity = kStackLds + A.stack.ovf_entries;

  const MeshView M = A.mesh;

  for (uint64_t ray = gtid; ray < A.n; ray += gstride) {
    const f3 o = mk(A.origins[3 * ray + 0], A.origins[3 * ray + 1], A.origins[3 * ray + 2]);
    const f3 d = mk(A.directions[3 * ray + 0], A.directions[3 * ray + 1], A.directions[3 * ray + 2]);
    RayBox rb;
    rb.ox = o.x; rb.oy = o.y; rb.oz = o.z;
    rb.ix = 1.0f / d.x; rb.iy = 1.0f / d.y; rb.iz = 1.0f / d.z;
    rb.sx = rb.ix >= 0.0f; rb.sy = rb.iy >= 0.0f; rb.sz = rb.iz >= 0.0f;
    const float dlen = sqrtf(dot(d, d));
    const float dl1 = fabsf(d.x) + fabsf(d.y) + fabsf(d.z);
    const bool literal = (d.x == 0.0f) || (d.y == 0.0f) || (d.z == 0.0f);

    float best_t = __int_as_float(0x7f800000);  // +inf
    float best_et = 0.0f;                       // error bound of best_t
    ff best_ff = {0.0f, 0.0f};                  // best_t in double-float when best_ff_ok
    bool best_ff_ok = false;
    uint32_t best_slot = kNoHit;
    uint32_t hits = 0;
    bool overflow = false;

    for (uint32_t r = 0; r < M.n_roots && !overflow; r++) {
      uint32_t ref = M.roots[r];
      st.sp = 0;
      while (true) {
        if (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          for (uint32_t s = off; s < off + cnt; s++) {
            f3 a, b, c;
            uint32_t face, tmask;
            load_tri(M.tris, s, a, b, c, face, tmask);
            const uint32_t mask = A.side_mask ? A.side_mask : tmask;
            float t, et;
            if (hit_triangle(o, d, dl1, dlen, a, b, c, mask, A.near_dist, A.far_dist, t, et)) {
              hits++;
              // Strict < as in intersectClosestTri; when the FP32 values are within their error
              // bands the order is decided in double-float.
              bool better;
              bool exact = false;
              ff t_ff;
              if (t + et < best_t - best_et) {
                better = true;
              } else if (t - et > best_t + best_et) {
                better = false;
              } else if (best_slot == kNoHit) {
                better = true;  // any hit beats none (reached only when t or et is not finite)
              } else {
                t_ff = exact_hit_t(o, d, M.tris, s);
                if (!best_ff_ok) best_ff = exact_hit_t(o, d, M.tris, best_slot);
                better = ff_sign(ff_sub(t_ff, best_ff)) < 0;
                exact = true;
                best_ff_ok = true;
              }
              if (better) {
                best_slot = s;
                best_ff_ok = exact;
                if (exact) {
                  best_ff = t_ff;
                  best_t = ff_to_float(t_ff);
                  best_et = fmaxf(et, 2.0f * kU * best_t);
                } else {
                  best_t = t;
                  best_et = et;
                }
              }
            }
          }
        } else {
          const WideNode nd = M.nodes[ref];
          float tl, tr;
          bool hl = slab(rb, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y, A.near_dist, A.far_dist, tl);
          bool hr = slab(rb, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w, A.near_dist, A.far_dist, tr);
          const float t_reach = best_t + best_et;  // no hit in a box entered beyond this can win
          if (!literal) {
            hl = hl && (tl <= t_reach);
            hr = hr && (tr <= t_reach);
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
              // Far child is the right one whenever daxis is +-0 (l2r holds).
              const float oaxis = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
              const float fmin = axis == 0 ? nd.b1.z : (axis == 1 ? nd.b1.w : nd.b2.x);
              const bool prunable = (daxis == 0.0f) && (oaxis <= fmin);
              aux = (hits & ~kPruneFlag) | (prunable ? kPruneFlag : 0u);
            }
            if (!st.push(c2, aux)) {
              overflow = true;
              break;
            }
            ref = c1;
            continue;
          }
          if (h1) {
            ref = c1;
            continue;
          }
          if (h2) {
            ref = c2;
            continue;
          }
        }
        // pop the next pending subtree that can still matter
        bool found = false;
        while (st.sp > 0) {
          uint32_t aux;
          st.pop(ref, aux);
          if (!literal) {
            if (__uint_as_float(aux) <= best_t + best_et) {
              found = true;
              break;
            }
          } else {
            if ((aux & kPruneFlag) && hits > (aux & ~kPruneFlag)) continue;
            found = true;
            break;
          }
        }
        if (!found) break;
      }
    }

    uint32_t out_face = kNoHit;
    float out_dist = __int_as_float(0x7f800000);
    float px = __int_as_float(0x7fc00000), py = px, pz = px;
    float bw = px, bu = px, bv = px;
    if (overflow) {
      // A tree deeper than the stack sized at upload; cannot happen for a validated mesh.
      out_face = kNoHit - 1u;
    } else if (best_slot != kNoHit) {
      f3 a, b, c;
      uint32_t face, tmask;
      load_tri(M.tris, best_slot, a, b, c, face, tmask);
      const TriTerms r = tri_terms_exact(o, d, a, b, c);
      const ff t = ff_div(r.qdn, r.ddn_abs);
      const ff u = ff_div(r.e0, r.ddn_abs);
      const ff v = ff_div(r.e1, r.ddn_abs);
      const ff w = ff_sub(ff_sub(ff_from(1.0f), u), v);
      const ff dl = ff_sqrt(ff_add(ff_add(two_prod(d.x, d.x), two_prod(d.y, d.y)), two_prod(d.z, d.z)));
      out_face = face;
      out_dist = ff_to_float(ff_mul(t, dl));
      px = ff_to_float(ff_add(ff_from(o.x), ff_mulf(t, d.x)));
      py = ff_to_float(ff_add(ff_from(o.y), ff_mulf(t, d.y)));
      pz = ff_to_float(ff_add(ff_from(o.z), ff_mulf(t, d.z)));
      bw = ff_to_float(w);
      bu = ff_to_float(u);
      bv = ff_to_float(v);
    }
    A.out_face[ray] = out_face;
    if (A.out_dist) A.out_dist[ray] = out_dist;
    if (A.out_point) {
      A.out_point[3 * ray + 0] = px;
      A.out_point[3 * ray + 1] = py;
      A.out_point[3 * ray + 2] = pz;
    }
    if (A.out_bary) {
      A.out_bary[3 * ray + 0] = bw;
      A.out_bary[3 * ray + 1] = bu;
      A.out_bary[3 * ray + 2] = bv;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Closest point on a triangle: Ericson, Real-Time Collision Detection §5.1.5 (the algorithm of
// three's Triangle.closestPointToPoint), Voronoi regions tested in the same order.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ f3 closest_on_triangle(f3 p, f3 a, f3 b, f3 c) {
  const f3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const f3 bp = sub(p, b);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    return mk(a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v);
  }
  const f3 cp = sub(p, c);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    return mk(a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const f3 bc = sub(c, b);
    return mk(b.x + bc.x * w, b.y + bc.y * w, b.z + bc.z * w);
  }
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom, w = vc * denom;
  return mk(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

// Same regions in double-float; returns the point, its barycoord (weights of a, b, c) and the
// squared distance.
struct ClosestFF {
  ff3 point;
  ff bw, bu, bv;
  ff dist2;
};

__device__ __forceinline__ ff ff_dot3(const ff3& x, const ff3& y) { return ff3_dot(x, y); }
__device__ __forceinline__ bool ff_le0(ff x) { return ff_sign(x) <= 0; }
__device__ __forceinline__ bool ff_ge0(ff x) { return ff_sign(x) >= 0; }

__device__ __noinline__ ClosestFF closest_on_triangle_ff(f3 p, f3 a, f3 b, f3 c) {
  const ff3 ab = ff3_diff(b, a), ac = ff3_diff(c, a), ap = ff3_diff(p, a);
  const ff one = ff_from(1.0f), zero = ff_from(0.0f);
  const ff d1 = ff_dot3(ab, ap), d2 = ff_dot3(ac, ap);
  ClosestFF r;
  int region;  // 0 face, 1 a, 2 b, 3 c, 4 ab, 5 ac, 6 bc
  ff s = zero;  // edge parameter
  ff vb = zero, vc = zero, va = zero;
  const ff3 bp = ff3_diff(p, b);
  const ff d3 = ff_dot3(ab, bp), d4 = ff_dot3(ac, bp);
  const ff3 cpv = ff3_diff(p, c);
  const ff d5 = ff_dot3(ab, cpv), d6 = ff_dot3(ac, cpv);
  if (ff_le0(d1) && ff_le0(d2)) {
    region = 1;
  } else if (ff_ge0(d3) && ff_le0(ff_sub(d4, d3))) {
    region = 2;
  } else {
    vc = ff_sub(ff_mul(d1, d4), ff_mul(d3, d2));
    if (ff_le0(vc) && ff_ge0(d1) && ff_le0(d3)) {
      region = 4;
      s = ff_div(d1, ff_sub(d1, d3));
    } else if (ff_ge0(d6) && ff_le0(ff_sub(d5, d6))) {
      region = 3;
    } else {
      vb = ff_sub(ff_mul(d5, d2), ff_mul(d1, d6));
      if (ff_le0(vb) && ff_ge0(d2) && ff_le0(d6)) {
        region = 5;
        s = ff_div(d2, ff_sub(d2, d6));
      } else {
        va = ff_sub(ff_mul(d3, d6), ff_mul(d5, d4));
        const ff d43 = ff_sub(d4, d3), d56 = ff_sub(d5, d6);
        if (ff_le0(va) && ff_ge0(d43) && ff_ge0(d56)) {
          region = 6;
          s = ff_div(d43, ff_add(d43, d56));
        } else {
          region = 0;
        }
      }
    }
  }
  ff bw = zero, bu = zero, bv = zero;
  switch (region) {
    case 1: bw = one; break;
    case 2: bu = one; break;
    case 3: bv = one; break;
    case 4: bu = s; bw = ff_sub(one, s); break;
    case 5: bv = s; bw = ff_sub(one, s); break;
    case 6: bv = s; bu = ff_sub(one, s); break;
    default: {
      const ff sum = ff_add(ff_add(va, vb), vc);
      bu = ff_div(vb, sum);
      bv = ff_div(vc, sum);
      bw = ff_sub(ff_sub(one, bu), bv);
    }
  }
  // point = a + u*(b - a) + v*(c - a) (vertex regions return the vertex itself, exactly)
  if (region == 1) {
    r.point = {ff_from(a.x), ff_from(a.y), ff_from(a.z)};
  } else if (region == 2) {
    r.point = {ff_from(b.x), ff_from(b.y), ff_from(b.z)};
  } else if (region == 3) {
    r.point = {ff_from(c.x), ff_from(c.y), ff_from(c.z)};
  } else if (region == 6) {
    const ff3 bc = ff3_diff(c, b);
    r.point = {ff_add(ff_from(b.x), ff_mul(bc.x, s)), ff_add(ff_from(b.y), ff_mul(bc.y, s)),
               ff_add(ff_from(b.z), ff_mul(bc.z, s))};
  } else {
    r.point = {ff_add(ff_add(ff_from(a.x), ff_mul(ab.x, bu)), ff_mul(ac.x, bv)),
               ff_add(ff_add(ff_from(a.y), ff_mul(ab.y, bu)), ff_mul(ac.y, bv)),
               ff_add(ff_add(ff_from(a.z), ff_mul(ab.z, bu)), ff_mul(ac.z, bv))};
  }
  const ff3 diff = {ff_sub(ff_from(p.x), r.point.x), ff_sub(ff_from(p.y), r.point.y),
                    ff_sub(ff_from(p.z), r.point.z)};
  r.dist2 = ff3_dot(diff, diff);
  r.bw = bw;
  r.bu = bu;
  r.bv = bv;
  return r;
}

__device__ __forceinline__ float box_score(f3 q, float mnx, float mny, float mnz, float mxx,
                                           float mxy, float mxz) {
  // distance^2 from q to the box (three Vector3.clamp then distanceToSquared)
  const float cx = fmaxf(mnx, fminf(mxx, q.x));
  const float cy = fmaxf(mny, fminf(mxy, q.y));
  const float cz = fmaxf(mnz, fminf(mxz, q.z));
  const float dx = cx - q.x, dy = cy - q.y, dz = cz - q.z;
  return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------------------------------------------------
// Closest point. Traversal of closestPointToPoint over shapecast
// (src/core/cast/closestPointToPoint.js:6-78, src/core/cast/shapecast.js:41-214): children
// ordered by squared box distance (right first only if strictly closer), a child is entered iff
// score < best^2 and score < maxThreshold^2 (the far child's score is re-checked against the
// best found meanwhile), root nodes are never bounds-tested, triangles update the best on strict
// <, and the first triangle closer than minThreshold ends the whole query (all roots).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_closest_point(ClosestPointArgs A) {
  __shared__ uint32_t s_ref[kStackLds * kBlock];
  __shared__ uint32_t s_aux[kStackLds * kBlock];

  const uint32_t lane = threadIdx.x;
  const uint64_t gtid = static_cast<uint64_t>(blockIdx.x) * kBlock + lane;
  const uint64_t gstride = static_cast<uint64_t>(gridDim.x) * kBlock;

  LaneStack st;
  st.s_ref = s_ref;
  st.s_aux = s_aux;
  st.ovf = A.stack.ovf;
  st.ovf_stride = A.stack.ovf_stride;
  st.gtid = gtid;
  st.lane = lane;
  st.capacity = kStackLds + A.stack.ovf_entries;

  const MeshView M = A.mesh;
  const float min2 = A.min_threshold * A.min_threshold;
  const float max2 = A.max_threshold * A.max_threshold;
  // Scores are compared with a tiny relative slack so FP32 rounding never skips a box the exact
  // score would enter.
  constexpr float kScoreSlack = 1.0f - 4.0f * kU;

  for (uint64_t qi = gtid; qi < A.n; qi += gstride) {
    const f3 q = mk(A.points[3 * qi + 0], A.points[3 * qi + 1], A.points[3 * qi + 2]);
    // Absolute error bound of an FP32 point-to-triangle distance for this query.
    const float e_abs = 16.0f * kU * (maxabs(q) + M.coord_max);
    float best = __int_as_float(0x7f800000);
    float best_e = 0.0f;
    ff best_ff = {0.0f, 0.0f};
    bool best_ff_ok = false;
    uint32_t best_slot = kNoHit;
    bool done = false;
    bool overflow = false;

    for (uint32_t r = 0; r < M.n_roots && !done; r++) {
      uint32_t ref = M.roots[r];
      st.sp = 0;
      while (true) {
        if (ref & kRefLeaf) {
          uint32_t off, cnt;
          leaf_range(ref, M.bigleaf, off, cnt);
          for (uint32_t s = off; s < off + cnt; s++) {
            f3 a, b, c;
            uint32_t face, tmask;
            load_tri(M.tris, s, a, b, c, face, tmask);
            const f3 cp = closest_on_triangle(q, a, b, c);
            const f3 dv = sub(q, cp);
            const float d2 = dot(dv, dv);
            const float e2 = (2.0f * sqrtf(d2) + e_abs) * e_abs + 4.0f * kU * d2;
            // Strict < as in closestPointToPoint; near-equal candidates are ordered in
            // double-float (exact ties, e.g. a shared vertex, keep the first as the reference).
            bool better;
            bool exact = false;
            ff d2_ff;
            if (d2 + e2 < best - best_e) {
              better = true;
            } else if (d2 - e2 > best + best_e) {
              better = false;
            } else {
              // Also reached when d2 is NaN (a degenerate triangle in FP32): the double-float
              // value decides, and a NaN there never wins, as in the reference's `distSq <`.
              d2_ff = closest_on_triangle_ff(q, a, b, c).dist2;
              if (best_slot == kNoHit) {
                better = d2_ff.hi < best;  // best is +inf here
              } else {
                if (!best_ff_ok) {
                  f3 ba, bb, bc;
                  uint32_t bface, bmask;
                  load_tri(M.tris, best_slot, ba, bb, bc, bface, bmask);
                  best_ff = closest_on_triangle_ff(q, ba, bb, bc).dist2;
                  best_ff_ok = true;
                }
                better = ff_sign(ff_sub(d2_ff, best_ff)) < 0;
              }
              exact = true;
            }
            if (better) {
              best_slot = s;
              best_ff_ok = exact;
              if (exact) {
                best_ff = d2_ff;
                best = ff_to_float(d2_ff);
                best_e = (2.0f * sqrtf(best) + e_abs) * e_abs + 4.0f * kU * best;
              } else {
                best = d2;
                best_e = e2;
              }
            }
            if (d2 < min2) {
              done = true;
              break;
            }
          }
          if (done) break;
        } else {
          const WideNode nd = M.nodes[ref];
          const float sl = box_score(q, nd.b0.x, nd.b0.y, nd.b0.z, nd.b0.w, nd.b1.x, nd.b1.y);
          const float sr = box_score(q, nd.b1.z, nd.b1.w, nd.b2.x, nd.b2.y, nd.b2.z, nd.b2.w);
          const bool swap = sr < sl;
          const uint32_t c1 = swap ? nd.meta.y : nd.meta.x;
          const uint32_t c2 = swap ? nd.meta.x : nd.meta.y;
          const float s1 = (swap ? sr : sl) * kScoreSlack;
          const float s2 = (swap ? sl : sr) * kScoreSlack;
          const float reach = best + best_e;  // a box scoring beyond this holds no better triangle
          const bool in1 = s1 < reach && s1 < max2;
          const bool in2 = s2 < reach && s2 < max2;
          if (in1) {
            if (in2) {
              if (!st.push(c2, __float_as_uint(s2))) {
                overflow = true;
                done = true;
                break;
              }
            }
            ref = c1;
            continue;
          }
          // s2 >= s1, so if the near child is out the far one is too.
        }
        bool found = false;
        while (st.sp > 0) {
          uint32_t aux;
          st.pop(ref, aux);
          const float sc = __uint_as_float(aux);
          if (sc < best + best_e && sc < max2) {
            found = true;
            break;
          }
        }
        if (!found) break;
      }
    }

    uint32_t out_face = kNoHit;
    float out_dist = __int_as_float(0x7f800000);
    float px = __int_as_float(0x7fc00000), py = px, pz = px;
    float bw = px, bu = px, bv = px;
    if (overflow) {
      out_face = kNoHit - 1u;
    } else if (best_slot != kNoHit) {
      f3 a, b, c;
      uint32_t face, tmask;
      load_tri(M.tris, best_slot, a, b, c, face, tmask);
      const ClosestFF cf = closest_on_triangle_ff(q, a, b, c);
      out_face = face;
      out_dist = ff_to_float(ff_sqrt(cf.dist2));
      px = ff_to_float(cf.point.x);
      py = ff_to_float(cf.point.y);
      pz = ff_to_float(cf.point.z);
      bw = ff_to_float(cf.bw);
      bu = ff_to_float(cf.bu);
      bv = ff_to_float(cf.bv);
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

hipError_t launch_raycast_first(const RaycastFirstArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_raycast_first, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_closest_point(const ClosestPointArgs& a, uint64_t threads, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_closest_point, dim3(static_cast<uint32_t>(threads / kBlock)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_apply_side_mask(float4* tris, uint64_t n_slots, const uint8_t* mask, hipStream_t s) {
  if (n_slots == 0) return hipSuccess;
  const uint32_t blocks = static_cast<uint32_t>((n_slots + 255) / 256);
  hipLaunchKernelGGL(k_apply_side_mask, dim3(blocks), dim3(256), 0, s, tris, n_slots, mask);
  return hipGetLastError();
}

}  // namespace mbvh
