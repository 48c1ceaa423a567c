// Closest point on a triangle in double-float (shared by the query and backward kernels).
// Ericson, Real-Time Collision Detection §5.1.5 (the algorithm of three's
// Triangle.closestPointToPoint): Voronoi regions tested in the same order, every sign decided on
// double-float values, the point and barycoord in double-float.
#pragma once

#include "mbvh_device.h"

namespace mbvh {
namespace {

// The closest point, its barycoord (weights of a, b, c) and the squared distance.
struct ClosestFF {
  ff3 point;
  ff bw, bu, bv;
  ff dist2;
};

__device__ __forceinline__ bool ff_le0(ff x) { return ff_sign(x) <= 0; }
__device__ __forceinline__ bool ff_ge0(ff x) { return ff_sign(x) >= 0; }

__device__ __noinline__ ClosestFF closest_on_triangle_ff(f3 p, f3 a, f3 b, f3 c) {
  const ff3 ab = ff3_diff(b, a), ac = ff3_diff(c, a), ap = ff3_diff(p, a);
  const ff one = ff_from(1.0f), zero = ff_from(0.0f);
  const ff d1 = ff3_dot(ab, ap), d2 = ff3_dot(ac, ap);
  ClosestFF r;
  int region;  // 0 face, 1 a, 2 b, 3 c, 4 ab, 5 ac, 6 bc
  ff s = zero;  // edge parameter
  ff vb = zero, vc = zero, va = zero;
  const ff3 bp = ff3_diff(p, b);
  const ff d3 = ff3_dot(ab, bp), d4 = ff3_dot(ac, bp);
  const ff3 cpv = ff3_diff(p, c);
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

}  // namespace
}  // namespace mbvh
