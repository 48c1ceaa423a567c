// Closest point on a triangle in double-float (shared by the query and backward kernels).
// Ericson, Real-Time Collision Detection §5.1.5 (the algorithm of three's
// Triangle.closestPointToPoint): Voronoi regions tested in the same order, every sign decided on
// double-float values, the point and barycoord in double-float.
#pragma once

#include "mbvh_device.h"

namespace mbvh {
namespace {

// The closest point, its barycoord (weights of a, b, c), the vector from it to the query, the
// squared distance and an absolute bound on that squared distance's double-float error.
struct ClosestFF {
  ff3 point;
  ff bw, bu, bv;
  ff3 qp;  // query - point
  ff dist2;
  float err;
};

__device__ __forceinline__ bool ff_le0(ff x) { return ff_sign(x) <= 0; }
__device__ __forceinline__ bool ff_ge0(ff x) { return ff_sign(x) >= 0; }

// The query is p + pl (pl: the lo part of a double-precision point, unused when kLo is false).
// query - point is formed from the region's base vertex (p - a, p - b or p - c, exact up to the lo
// part) minus the edge vectors times the barycentric weights, never from the absolute point, so
// its rounding scales with the triangle and the distance, not with the coordinates. An error in
// the weights moves the point within the region's face or edge, perpendicular to query - point,
// and changes the squared distance only to second order; the first-order error is the roundings
// of the products and sums: each component within 2^-45 (|e1| + |e2|) + 2^-46 (|base| + |qp|),
// which bounds the squared distance's error by 2^-43 (d R + d^2), R = |base|_1 + |e1|_1 + |e2|_1.
template <bool kLo>
__device__ __forceinline__ ClosestFF closest_on_triangle_ff_inline(f3 p, f3 pl, f3 a, f3 b, f3 c) {
  auto diff = [&](f3 v) {  // p + pl - v
    ff3 r = ff3_diff(p, v);
    if (kLo) r = {ff_add(r.x, ff_from(pl.x)), ff_add(r.y, ff_from(pl.y)), ff_add(r.z, ff_from(pl.z))};
    return r;
  };
  const ff3 ab = ff3_diff(b, a), ac = ff3_diff(c, a), ap = diff(a);
  const ff one = ff_from(1.0f), zero = ff_from(0.0f);
  const ff d1 = ff3_dot(ab, ap), d2 = ff3_dot(ac, ap);
  ClosestFF r;
  int region;  // 0 face, 1 a, 2 b, 3 c, 4 ab, 5 ac, 6 bc
  ff s = zero;  // edge parameter
  ff vb = zero, vc = zero, va = zero;
  const ff3 bp = diff(b);
  const ff d3 = ff3_dot(ab, bp), d4 = ff3_dot(ac, bp);
  const ff3 cpv = diff(c);
  const ff d5 = ff3_dot(ab, cpv), d6 = ff3_dot(ac, cpv);
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
  // query - point from the region's base vertex: base - e1 * w1 - e2 * w2
  const ff3 bc = ff3_diff(c, b);
  const ff3 base = region == 2 ? bp : (region == 3 ? cpv : (region == 6 ? bp : ap));
  const ff3 e1 = region == 6 ? bc : ab;
  const ff w1 = region == 6 ? bv : bu;
  const ff w2 = region == 6 ? zero : bv;
  const bool vertex = region >= 1 && region <= 3;
  ff3 qp = base;
  if (!vertex) {
    qp = {ff_sub(qp.x, ff_mul(e1.x, w1)), ff_sub(qp.y, ff_mul(e1.y, w1)), ff_sub(qp.z, ff_mul(e1.z, w1))};
    if (region == 0 || region == 5)
      qp = {ff_sub(qp.x, ff_mul(ac.x, w2)), ff_sub(qp.y, ff_mul(ac.y, w2)), ff_sub(qp.z, ff_mul(ac.z, w2))};
  }
  // the point itself (reported): the vertex exactly, else a + u (b - a) + v (c - a) / b + v (c - b)
  if (region == 1) {
    r.point = {ff_from(a.x), ff_from(a.y), ff_from(a.z)};
  } else if (region == 2) {
    r.point = {ff_from(b.x), ff_from(b.y), ff_from(b.z)};
  } else if (region == 3) {
    r.point = {ff_from(c.x), ff_from(c.y), ff_from(c.z)};
  } else if (region == 6) {
    r.point = {ff_add(ff_from(b.x), ff_mul(bc.x, s)), ff_add(ff_from(b.y), ff_mul(bc.y, s)),
               ff_add(ff_from(b.z), ff_mul(bc.z, s))};
  } else {
    r.point = {ff_add(ff_add(ff_from(a.x), ff_mul(ab.x, bu)), ff_mul(ac.x, bv)),
               ff_add(ff_add(ff_from(a.y), ff_mul(ab.y, bu)), ff_mul(ac.y, bv)),
               ff_add(ff_add(ff_from(a.z), ff_mul(ab.z, bu)), ff_mul(ac.z, bv))};
  }
  r.qp = qp;
  r.dist2 = ff3_dot(qp, qp);
  auto l1 = [](const ff3& v) { return fabsf(v.x.hi) + fabsf(v.y.hi) + fabsf(v.z.hi); };
  float R = l1(base);
  if (!vertex) R += l1(e1) + (region == 0 || region == 5 ? l1(ac) : 0.0f);
  const float dd = fabsf(r.dist2.hi);
  // 2^-43 (d R + d^2), its own float roundings covered by the factor 1.001
  r.err = 1.1368683772161603e-13f * 1.001f * (sqrtf(dd) * (R * 1.001f) + dd);
  r.bw = bw;
  r.bu = bu;
  r.bv = bv;
  return r;
}

// The same, as one out-of-line function: callers that reach it from deep inside a traversal keep
// their register budget; a kernel whose loop is mostly this call takes the inline form.
template <bool kLo>
__device__ __noinline__ ClosestFF closest_on_triangle_ff(f3 p, f3 pl, f3 a, f3 b, f3 c) {
  return closest_on_triangle_ff_inline<kLo>(p, pl, a, b, c);
}

}  // namespace
}  // namespace mbvh
