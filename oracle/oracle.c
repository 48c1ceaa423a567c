/*
 * oracle.c — FP64 CPU restatement of the three-mesh-bvh query path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker,
 * never by the product. See oracle.h for what each function restates.
 *
 * Built with -ffp-contract=off -fno-fast-math so every operation rounds exactly as the JS
 * reference's IEEE double arithmetic does (JS never fuses multiply-adds).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NO_HIT 0xFFFFFFFFu

/* ------------------------------------------------------------------------------------------ */
/* three.js math restated (Vector3 operation order preserved)                                 */
/* ------------------------------------------------------------------------------------------ */

static void v_sub(const double* a, const double* b, double* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static double v_dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v_cross(const double* a, const double* b, double* o) { /* Vector3.crossVectors */
  const double ax = a[0], ay = a[1], az = a[2], bx = b[0], by = b[1], bz = b[2];
  o[0] = ay * bz - az * by;
  o[1] = az * bx - ax * bz;
  o[2] = ax * by - ay * bx;
}
static double v_dist_sq(const double* a, const double* b) { /* Vector3.distanceToSquared */
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

/* Ray.intersectTriangle (GeometricTools IntrRay3Triangle3, as published in three.js). */
int orc_ray_triangle(const double o[3], const double d[3], const double a[3], const double b[3],
                     const double c[3], int backface_culling, double* t, double point[3]) {
  double edge1[3], edge2[3], normal[3], diff[3], tmp[3];
  v_sub(b, a, edge1);
  v_sub(c, a, edge2);
  v_cross(edge1, edge2, normal);
  double DdN = v_dot(d, normal);
  double sign;
  if (DdN > 0) {
    if (backface_culling) return 0;
    sign = 1;
  } else if (DdN < 0) {
    sign = -1;
    DdN = -DdN;
  } else {
    return 0;
  }
  v_sub(o, a, diff);
  v_cross(diff, edge2, tmp);
  const double DdQxE2 = sign * v_dot(d, tmp);
  if (DdQxE2 < 0) return 0;
  v_cross(edge1, diff, tmp);
  const double DdE1xQ = sign * v_dot(d, tmp);
  if (DdE1xQ < 0) return 0;
  if (DdQxE2 + DdE1xQ > DdN) return 0;
  const double QdN = -sign * v_dot(diff, normal);
  if (QdN < 0) return 0;
  const double tt = QdN / DdN;
  *t = tt;
  /* Ray.at: origin + direction * t */
  point[0] = o[0] + d[0] * tt;
  point[1] = o[1] + d[1] * tt;
  point[2] = o[2] + d[2] * tt;
  return 1;
}

/* Triangle.closestPointToPoint: Ericson, Real-Time Collision Detection §5.1.5. */
void orc_closest_point_triangle(const double p[3], const double a[3], const double b[3],
                                const double c[3], double out[3]) {
  double vab[3], vac[3], vap[3], vbp[3], vcp[3];
  v_sub(b, a, vab);
  v_sub(c, a, vac);
  v_sub(p, a, vap);
  const double d1 = v_dot(vab, vap), d2 = v_dot(vac, vap);
  if (d1 <= 0 && d2 <= 0) {
    memcpy(out, a, sizeof(double) * 3);
    return;
  }
  v_sub(p, b, vbp);
  const double d3 = v_dot(vab, vbp), d4 = v_dot(vac, vbp);
  if (d3 >= 0 && d4 <= d3) {
    memcpy(out, b, sizeof(double) * 3);
    return;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    const double v = d1 / (d1 - d3);
    for (int k = 0; k < 3; k++) out[k] = a[k] + vab[k] * v;
    return;
  }
  v_sub(p, c, vcp);
  const double d5 = v_dot(vab, vcp), d6 = v_dot(vac, vcp);
  if (d6 >= 0 && d5 <= d6) {
    memcpy(out, c, sizeof(double) * 3);
    return;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    const double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; k++) out[k] = a[k] + vac[k] * w;
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    double vbc[3];
    v_sub(c, b, vbc);
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; k++) out[k] = b[k] + vbc[k] * w;
    return;
  }
  const double denom = 1 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  for (int k = 0; k < 3; k++) out[k] = (a[k] + vab[k] * v) + vac[k] * w;
}

/* Triangle.getBarycoord */
int orc_barycoord(const double p[3], const double a[3], const double b[3], const double c[3],
                  double out[3]) {
  double v0[3], v1[3], v2[3];
  v_sub(c, a, v0);
  v_sub(b, a, v1);
  v_sub(p, a, v2);
  const double dot00 = v_dot(v0, v0), dot01 = v_dot(v0, v1), dot02 = v_dot(v0, v2);
  const double dot11 = v_dot(v1, v1), dot12 = v_dot(v1, v2);
  const double denom = dot00 * dot11 - dot01 * dot01;
  if (denom == 0) {
    out[0] = out[1] = out[2] = 0;
    return 0;
  }
  const double inv = 1 / denom;
  const double u = (dot11 * dot02 - dot01 * dot12) * inv;
  const double v = (dot00 * dot12 - dot01 * dot02) * inv;
  out[0] = 1 - u - v;
  out[1] = v;
  out[2] = u;
  return 1;
}

/* ------------------------------------------------------------------------------------------ */
/* Mesh access                                                                                */
/* ------------------------------------------------------------------------------------------ */

static uint64_t rd(const void* p, int eb, uint64_t i) {
  return eb == 2 ? ((const uint16_t*)p)[i] : ((const uint32_t*)p)[i];
}
static uint32_t face_of_slot(const orc_mesh* m, uint64_t slot) {
  return (uint32_t)(m->indirect ? rd(m->indirect, m->indirect_elt_bytes, slot) : slot);
}
static void face_vertices(const orc_mesh* m, uint32_t face, double a[3], double b[3], double c[3]) {
  uint64_t v[3] = {3ull * face, 3ull * face + 1, 3ull * face + 2};
  if (m->index)
    for (int k = 0; k < 3; k++) v[k] = rd(m->index, m->index_elt_bytes, v[k]);
  double* out[3] = {a, b, c};
  for (int k = 0; k < 3; k++) {
    const float* p = m->positions + v[k] * m->stride + m->offset;
    out[k][0] = p[0];
    out[k][1] = p[1];
    out[k][2] = p[2];
  }
}

/* Reference node accessors (src/core/utils/nodeBufferUtils.js:3-47) */
typedef struct {
  const uint8_t* bytes;
  uint64_t n;
} NodeBuf;
static const float* nb_bounds(NodeBuf nb, uint64_t i) { return (const float*)(nb.bytes + 32 * i); }
static uint32_t nb_u32(NodeBuf nb, uint64_t i, int w) {
  uint32_t v;
  memcpy(&v, nb.bytes + 32 * i + 4 * w, 4);
  return v;
}
static uint16_t nb_u16(NodeBuf nb, uint64_t i, int h) {
  uint16_t v;
  memcpy(&v, nb.bytes + 32 * i + 2 * h, 2);
  return v;
}
static int nb_leaf(NodeBuf nb, uint64_t i) { return nb_u16(nb, i, 15) == 0xFFFF; }
static NodeBuf root_buf(const orc_mesh* m, uint32_t r) {
  NodeBuf nb = {(const uint8_t*)m->roots[r], m->root_bytes[r] / 32};
  return nb;
}

/* ------------------------------------------------------------------------------------------ */
/* Raycast                                                                                    */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  double o[3], d[3];
  int side;
  double near_d, far_d;
  const orc_mesh* m;
  uint32_t nodes, tris;
} RayCtx;

typedef struct {
  int found;
  uint32_t face;
  double dist;
  double point[3];
} Hit;

/* intersectsNodeBounds (src/core/utils/intersectUtils.js:2-78) */
static int node_bounds_hit(const float* b, const RayCtx* r) {
  double tmin, tmax, tymin, tymax, tzmin, tzmax;
  const double invdirx = 1 / r->d[0], invdiry = 1 / r->d[1], invdirz = 1 / r->d[2];
  const double ox = r->o[0], oy = r->o[1], oz = r->o[2];
  const double minx = b[0], maxx = b[3], miny = b[1], maxy = b[4], minz = b[2], maxz = b[5];
  if (invdirx >= 0) {
    tmin = (minx - ox) * invdirx;
    tmax = (maxx - ox) * invdirx;
  } else {
    tmin = (maxx - ox) * invdirx;
    tmax = (minx - ox) * invdirx;
  }
  if (invdiry >= 0) {
    tymin = (miny - oy) * invdiry;
    tymax = (maxy - oy) * invdiry;
  } else {
    tymin = (maxy - oy) * invdiry;
    tymax = (miny - oy) * invdiry;
  }
  if ((tmin > tymax) || (tymin > tmax)) return 0;
  if (tymin > tmin || isnan(tmin)) tmin = tymin;
  if (tymax < tmax || isnan(tmax)) tmax = tymax;
  if (invdirz >= 0) {
    tzmin = (minz - oz) * invdirz;
    tzmax = (maxz - oz) * invdirz;
  } else {
    tzmin = (maxz - oz) * invdirz;
    tzmax = (minz - oz) * invdirz;
  }
  if ((tmin > tzmax) || (tzmin > tmax)) return 0;
  if (tzmin > tmin || tmin != tmin) tmin = tzmin;
  if (tzmax < tmax || tmax != tmax) tmax = tzmax;
  return tmin <= r->far_d && tmax >= r->near_d;
}

static int side_of_face(const RayCtx* r, uint32_t face) {
  if (r->side != ORC_PER_FACE) return r->side;
  const uint8_t mk = r->m->face_mask ? r->m->face_mask[face] & 3 : 3;
  if (mk == 1) return ORC_FRONT;
  if (mk == 2) return ORC_BACK;
  if (mk == 3) return ORC_DOUBLE;
  return -1;
}

/* intersectTri → checkBufferGeometryIntersection → checkIntersection
 * (src/utils/ThreeRayIntersectUtilities.js:21-47,49-145,154-230) */
static int intersect_face(RayCtx* r, uint32_t face, Hit* h) {
  r->tris++;
  const int side = side_of_face(r, face);
  if (side < 0) return 0;
  double a[3], b[3], c[3], t, p[3];
  face_vertices(r->m, face, a, b, c);
  int ok;
  if (side == ORC_BACK)
    ok = orc_ray_triangle(r->o, r->d, c, b, a, 1, &t, p);
  else
    ok = orc_ray_triangle(r->o, r->d, a, b, c, side != ORC_DOUBLE, &t, p);
  if (!ok) return 0;
  const double dist = sqrt(v_dist_sq(r->o, p)); /* ray.origin.distanceTo(point) */
  if (dist < r->near_d || dist > r->far_d) return 0;
  h->found = 1;
  h->face = face;
  h->dist = dist;
  memcpy(h->point, p, sizeof(p));
  return 1;
}

/* intersectClosestTri (src/core/utils/iterationUtils.template.js:24-53) */
static Hit closest_in_leaf(RayCtx* r, uint64_t off, uint64_t cnt) {
  Hit best;
  best.found = 0;
  double dist = INFINITY;
  for (uint64_t s = off; s < off + cnt; s++) {
    Hit h;
    if (intersect_face(r, face_of_slot(r->m, s), &h) && h.dist < dist) {
      best = h;
      dist = h.dist;
    }
  }
  return best;
}

/* _raycastFirst (src/core/cast/raycastFirst.template.js:19-103) */
static Hit raycast_first_node(RayCtx* r, NodeBuf nb, uint64_t node) {
  Hit none;
  none.found = 0;
  if (nb_leaf(nb, node)) return closest_in_leaf(r, nb_u32(nb, node, 6), nb_u16(nb, node, 14));
  const uint32_t axis = nb_u32(nb, node, 7);
  const int left_to_right = r->d[axis] >= 0;
  const uint64_t left = node + 1, right = node + nb_u32(nb, node, 6);
  const uint64_t c1 = left_to_right ? left : right, c2 = left_to_right ? right : left;
  r->nodes++;
  Hit r1 = none;
  if (node_bounds_hit(nb_bounds(nb, c1), r)) r1 = raycast_first_node(r, nb, c1);
  if (r1.found) {
    const double p = r1.point[axis];
    const float* b2 = nb_bounds(nb, c2);
    const int outside = left_to_right ? p <= b2[axis] : p >= b2[axis + 3];
    if (outside) return r1;
  }
  r->nodes++;
  Hit r2 = none;
  if (node_bounds_hit(nb_bounds(nb, c2), r)) r2 = raycast_first_node(r, nb, c2);
  if (r1.found && r2.found) return r1.dist <= r2.dist ? r1 : r2;
  if (r1.found) return r1;
  return r2;
}

static Hit raycast_first_one(RayCtx* r, int brute) {
  const orc_mesh* m = r->m;
  Hit best;
  best.found = 0;
  for (uint32_t i = 0; i < m->n_roots; i++) {
    NodeBuf nb = root_buf(m, i);
    Hit h;
    if (brute) {
      /* every slot of this root's range, in slot order */
      uint64_t lo = UINT64_MAX, hi = 0;
      for (uint64_t k = 0; k < nb.n; k++) {
        if (!nb_leaf(nb, k)) continue;
        const uint64_t off = nb_u32(nb, k, 6), cnt = nb_u16(nb, k, 14);
        if (off < lo) lo = off;
        if (off + cnt > hi) hi = off + cnt;
      }
      h = lo < hi ? closest_in_leaf(r, lo, hi - lo) : best;
      if (lo >= hi) h.found = 0;
    } else {
      h = raycast_first_node(r, nb, 0);
    }
    /* MeshBVH.raycastFirst: keep a root's result only if strictly closer (MeshBVH.js:567) */
    if (h.found && (!best.found || h.dist < best.dist)) best = h;
  }
  return best;
}

/* _raycast (src/core/cast/raycast.template.js:16-53) */
typedef struct {
  uint64_t count, cap;
  uint32_t* face;
  double* dist;
} HitList;
static void raycast_all_node(RayCtx* r, NodeBuf nb, uint64_t node, HitList* out) {
  if (nb_leaf(nb, node)) {
    const uint64_t off = nb_u32(nb, node, 6), cnt = nb_u16(nb, node, 14);
    for (uint64_t s = off; s < off + cnt; s++) {
      Hit h;
      if (intersect_face(r, face_of_slot(r->m, s), &h)) {
        if (out->count < out->cap) {
          out->face[out->count] = h.face;
          out->dist[out->count] = h.dist;
        }
        out->count++;
      }
    }
    return;
  }
  const uint64_t left = node + 1, right = node + nb_u32(nb, node, 6);
  r->nodes++;
  if (node_bounds_hit(nb_bounds(nb, left), r)) raycast_all_node(r, nb, left, out);
  r->nodes++;
  if (node_bounds_hit(nb_bounds(nb, right), r)) raycast_all_node(r, nb, right, out);
}

int64_t orc_raycast_all_f64(const orc_mesh* m, const double origin[3], const double dir[3], int side,
                            double near_d, double far_d, uint64_t cap, uint32_t* face, double* dist) {
  RayCtx r;
  memset(&r, 0, sizeof(r));
  for (int k = 0; k < 3; k++) {
    r.o[k] = origin[k];
    r.d[k] = dir[k];
  }
  r.side = side;
  r.near_d = near_d;
  r.far_d = far_d;
  r.m = m;
  HitList out = {0, cap, face, dist};
  for (uint32_t i = 0; i < m->n_roots; i++) raycast_all_node(&r, root_buf(m, i), 0, &out);
  return (int64_t)out.count;
}

int64_t orc_raycast_all(const orc_mesh* m, const float origin[3], const float dir[3], int side,
                        double near_d, double far_d, uint64_t cap, uint32_t* face, double* dist) {
  RayCtx r;
  memset(&r, 0, sizeof(r));
  for (int k = 0; k < 3; k++) {
    r.o[k] = origin[k];
    r.d[k] = dir[k];
  }
  r.side = side;
  r.near_d = near_d;
  r.far_d = far_d;
  r.m = m;
  HitList out = {0, cap, face, dist};
  for (uint32_t i = 0; i < m->n_roots; i++) raycast_all_node(&r, root_buf(m, i), 0, &out);
  return (int64_t)out.count;
}

/* ------------------------------------------------------------------------------------------ */
/* Closest point                                                                              */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  double p[3];
  double best_sq, min_sq, max_sq;
  uint32_t best_face;
  double best_point[3];
  const orc_mesh* m;
  uint32_t nodes, tris;
} PointCtx;

static double box_score(const float* b, const double p[3]) { /* clamp + distanceToSquared */
  double t[3];
  for (int k = 0; k < 3; k++) t[k] = fmax((double)b[k], fmin((double)b[k + 3], p[k]));
  return v_dist_sq(t, p);
}

/* closestPointToPoint's intersectsTriangle callback (closestPointToPoint.js:39-61) */
static int point_leaf(PointCtx* c, uint64_t off, uint64_t cnt) {
  for (uint64_t s = off; s < off + cnt; s++) {
    const uint32_t face = face_of_slot(c->m, s);
    double a[3], b[3], cc[3], cp[3];
    face_vertices(c->m, face, a, b, cc);
    c->tris++;
    orc_closest_point_triangle(c->p, a, b, cc, cp);
    const double d2 = v_dist_sq(c->p, cp);
    if (d2 < c->best_sq) {
      memcpy(c->best_point, cp, sizeof(cp));
      c->best_sq = d2;
      c->best_face = face;
    }
    if (d2 < c->min_sq) return 1;
  }
  return 0;
}

/* shapecastTraverse (src/core/cast/shapecast.js:41-214) with the closest-point callbacks */
static int point_node(PointCtx* c, NodeBuf nb, uint64_t node) {
  if (nb_leaf(nb, node)) return point_leaf(c, nb_u32(nb, node, 6), nb_u16(nb, node, 14));
  const uint64_t left = node + 1, right = node + nb_u32(nb, node, 6);
  uint64_t c1 = left, c2 = right;
  double s1 = box_score(nb_bounds(nb, left), c->p);
  double s2 = box_score(nb_bounds(nb, right), c->p);
  c->nodes += 2;
  if (s2 < s1) {
    c1 = right;
    c2 = left;
    const double t = s1;
    s1 = s2;
    s2 = t;
  }
  if (s1 < c->best_sq && s1 < c->max_sq)
    if (point_node(c, nb, c1)) return 1;
  if (s2 < c->best_sq && s2 < c->max_sq)
    if (point_node(c, nb, c2)) return 1;
  return 0;
}

static void closest_point_one(PointCtx* c, int brute) {
  const orc_mesh* m = c->m;
  for (uint32_t i = 0; i < m->n_roots; i++) {
    NodeBuf nb = root_buf(m, i);
    int stop;
    if (brute) {
      uint64_t lo = UINT64_MAX, hi = 0;
      for (uint64_t k = 0; k < nb.n; k++) {
        if (!nb_leaf(nb, k)) continue;
        const uint64_t off = nb_u32(nb, k, 6), cnt = nb_u16(nb, k, 14);
        if (off < lo) lo = off;
        if (off + cnt > hi) hi = off + cnt;
      }
      stop = lo < hi ? point_leaf(c, lo, hi - lo) : 0;
    } else {
      stop = point_node(c, nb, 0);
    }
    if (stop) break;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Threaded drivers                                                                           */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  const orc_mesh* m;
  int kind; /* 0 raycast first, 1 closest point */
  const float *a, *b;
  const double *ad, *bd; /* double-precision inputs (used instead of a, b when set) */
  uint64_t n;
  int side, brute;
  double p0, p1;
  uint32_t* face;
  double *dist, *point, *bary;
  uint32_t *nodes, *tris;
  uint64_t begin, end;
} Job;

static void run_ray(const Job* j, uint64_t i) {
  RayCtx r;
  memset(&r, 0, sizeof(r));
  for (int k = 0; k < 3; k++) {
    r.o[k] = j->ad ? j->ad[3 * i + k] : j->a[3 * i + k];
    r.d[k] = j->bd ? j->bd[3 * i + k] : j->b[3 * i + k];
  }
  r.side = j->side;
  r.near_d = j->p0;
  r.far_d = j->p1;
  r.m = j->m;
  Hit h = raycast_first_one(&r, j->brute);
  j->face[i] = h.found ? h.face : NO_HIT;
  if (j->dist) j->dist[i] = h.found ? h.dist : INFINITY;
  if (j->point)
    for (int k = 0; k < 3; k++) j->point[3 * i + k] = h.found ? h.point[k] : NAN;
  if (j->bary) {
    double bc[3] = {NAN, NAN, NAN};
    if (h.found) {
      double a[3], b[3], c[3];
      face_vertices(j->m, h.face, a, b, c);
      orc_barycoord(h.point, a, b, c, bc);
    }
    for (int k = 0; k < 3; k++) j->bary[3 * i + k] = bc[k];
  }
  if (j->nodes) j->nodes[i] = r.nodes;
  if (j->tris) j->tris[i] = r.tris;
}

static void run_point(const Job* j, uint64_t i) {
  PointCtx c;
  memset(&c, 0, sizeof(c));
  for (int k = 0; k < 3; k++) c.p[k] = j->ad ? j->ad[3 * i + k] : j->a[3 * i + k];
  c.best_sq = INFINITY;
  c.min_sq = j->p0 * j->p0;
  c.max_sq = j->p1 * j->p1;
  c.best_face = NO_HIT;
  c.m = j->m;
  closest_point_one(&c, j->brute);
  const int found = c.best_sq != INFINITY;
  j->face[i] = found ? c.best_face : NO_HIT;
  if (j->dist) j->dist[i] = found ? sqrt(c.best_sq) : INFINITY;
  if (j->point)
    for (int k = 0; k < 3; k++) j->point[3 * i + k] = found ? c.best_point[k] : NAN;
  if (j->nodes) j->nodes[i] = c.nodes;
  if (j->tris) j->tris[i] = c.tris;
}

static void* worker(void* arg) {
  const Job* j = (const Job*)arg;
  for (uint64_t i = j->begin; i < j->end; i++) {
    if (j->kind == 0)
      run_ray(j, i);
    else
      run_point(j, i);
  }
  return NULL;
}

static int run_jobs(Job proto, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if ((uint64_t)nthreads > proto.n) nthreads = proto.n ? (int)proto.n : 1;
  Job* jobs = (Job*)calloc((size_t)nthreads, sizeof(Job));
  pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
  if (!jobs || !th) {
    free(jobs);
    free(th);
    return -1;
  }
  /* interleaved chunks of 256 queries keep the threads balanced */
  const uint64_t chunk = 256;
  const uint64_t nchunks = (proto.n + chunk - 1) / chunk;
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = proto;
    jobs[t].begin = nchunks * (uint64_t)t / (uint64_t)nthreads * chunk;
    jobs[t].end = nchunks * (uint64_t)(t + 1) / (uint64_t)nthreads * chunk;
    if (jobs[t].end > proto.n) jobs[t].end = proto.n;
  }
  for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, worker, &jobs[t]);
  worker(&jobs[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
  free(jobs);
  free(th);
  return 0;
}

int orc_raycast_first(const orc_mesh* m, const float* origins, const float* dirs, uint64_t n,
                      int side, double near_d, double far_d, int brute, int nthreads,
                      uint32_t* face, double* dist, double* point, double* bary,
                      uint32_t* nodes_read, uint32_t* tris_tested) {
  Job j;
  memset(&j, 0, sizeof(j));
  j.m = m;
  j.kind = 0;
  j.a = origins;
  j.b = dirs;
  j.n = n;
  j.side = side;
  j.brute = brute;
  j.p0 = near_d;
  j.p1 = far_d;
  j.face = face;
  j.dist = dist;
  j.point = point;
  j.bary = bary;
  j.nodes = nodes_read;
  j.tris = tris_tested;
  return run_jobs(j, nthreads);
}

int orc_closest_point(const orc_mesh* m, const float* points, uint64_t n, double min_t,
                      double max_t, int brute, int nthreads, uint32_t* face, double* dist,
                      double* point, uint32_t* nodes_read, uint32_t* tris_tested) {
  Job j;
  memset(&j, 0, sizeof(j));
  j.m = m;
  j.kind = 1;
  j.a = points;
  j.n = n;
  j.brute = brute;
  j.p0 = min_t;
  j.p1 = max_t;
  j.face = face;
  j.dist = dist;
  j.point = point;
  j.nodes = nodes_read;
  j.tris = tris_tested;
  return run_jobs(j, nthreads);
}

int orc_raycast_first_f64(const orc_mesh* m, const double* origins, const double* dirs, uint64_t n, int side,
                          double near_d, double far_d, int brute, int nthreads, uint32_t* face, double* dist,
                          double* point, double* bary, uint32_t* nodes_read, uint32_t* tris_tested) {
  Job j;
  memset(&j, 0, sizeof(j));
  j.m = m;
  j.kind = 0;
  j.ad = origins;
  j.bd = dirs;
  j.n = n;
  j.side = side;
  j.brute = brute;
  j.p0 = near_d;
  j.p1 = far_d;
  j.face = face;
  j.dist = dist;
  j.point = point;
  j.bary = bary;
  j.nodes = nodes_read;
  j.tris = tris_tested;
  return run_jobs(j, nthreads);
}

int orc_closest_point_f64(const orc_mesh* m, const double* points, uint64_t n, double min_t, double max_t, int brute,
                          int nthreads, uint32_t* face, double* dist, double* point, uint32_t* nodes_read,
                          uint32_t* tris_tested) {
  Job j;
  memset(&j, 0, sizeof(j));
  j.m = m;
  j.kind = 1;
  j.ad = points;
  j.n = n;
  j.brute = brute;
  j.p0 = min_t;
  j.p1 = max_t;
  j.face = face;
  j.dist = dist;
  j.point = point;
  j.nodes = nodes_read;
  j.tris = tris_tested;
  return run_jobs(j, nthreads);
}

double orc_ray_plane_distance_f64(const orc_mesh* m, uint32_t face, const double o[3], const double d[3]) {
  double a[3], b[3], c[3], e1[3], e2[3], n[3], q[3];
  face_vertices(m, face, a, b, c);
  v_sub(b, a, e1);
  v_sub(c, a, e2);
  v_cross(e1, e2, n);
  const double ddn = v_dot(d, n);
  if (ddn == 0) return INFINITY;
  v_sub(a, o, q);
  const double t = v_dot(q, n) / ddn;
  return fabs(t) * sqrt(v_dot(d, d));
}

double orc_point_face_distance_f64(const orc_mesh* m, uint32_t face, const double q[3]) {
  double a[3], b[3], c[3], cp[3];
  face_vertices(m, face, a, b, c);
  orc_closest_point_triangle(q, a, b, c, cp);
  return sqrt(v_dist_sq(q, cp));
}

double orc_ray_plane_distance(const orc_mesh* m, uint32_t face, const float origin[3],
                              const float dir[3]) {
  double a[3], b[3], c[3], e1[3], e2[3], n[3], q[3];
  face_vertices(m, face, a, b, c);
  v_sub(b, a, e1);
  v_sub(c, a, e2);
  v_cross(e1, e2, n);
  const double o[3] = {origin[0], origin[1], origin[2]}, d[3] = {dir[0], dir[1], dir[2]};
  const double ddn = v_dot(d, n);
  if (ddn == 0) return INFINITY;
  v_sub(a, o, q);
  const double t = v_dot(q, n) / ddn;
  return fabs(t) * sqrt(v_dot(d, d));
}

double orc_point_face_distance(const orc_mesh* m, uint32_t face, const float p[3]) {
  double a[3], b[3], c[3], cp[3];
  const double q[3] = {p[0], p[1], p[2]};
  face_vertices(m, face, a, b, c);
  orc_closest_point_triangle(q, a, b, c, cp);
  return sqrt(v_dist_sq(q, cp));
}

/* ------------------------------------------------------------------------------------------ */
/* validateBounds (src/debug/Debug.js:207-267) + structural checks                            */
/* ------------------------------------------------------------------------------------------ */

static int box_contains(const float* outer, const float* inner) { /* Box3.containsBox */
  return outer[0] <= inner[0] && inner[3] <= outer[3] && outer[1] <= inner[1] && inner[4] <= outer[4] &&
         outer[2] <= inner[2] && inner[5] <= outer[5];
}

int orc_validate_root(const orc_mesh* m, uint32_t root, orc_tree_stats* st) {
  memset(st, 0, sizeof(*st));
  if (root >= m->n_roots) return -1;
  NodeBuf nb = root_buf(m, root);
  if (nb.n == 0) {
    st->structure_failures++;
    return 1;
  }
  const uint64_t n_slots = m->indirect ? m->indirect_count : (m->index ? m->index_count / 3 : m->vertex_count / 3);
  uint8_t* seen = (uint8_t*)calloc(n_slots ? n_slots : 1, 1);
  uint64_t* stack = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (nb.n + 1));
  uint64_t sp = 0, lo = UINT64_MAX, hi = 0;
  st->depth_min = UINT32_MAX;
  st->leaf_tris_min = UINT32_MAX;
  stack[sp++] = 0;
  stack[sp++] = 0;
  uint64_t visited = 0;
  while (sp) {
    const uint64_t depth = stack[--sp];
    const uint64_t node = stack[--sp];
    if (node >= nb.n || ++visited > nb.n) {
      st->structure_failures++;
      break;
    }
    st->node_count++;
    const float* box = nb_bounds(nb, node);
    if (nb_leaf(nb, node)) {
      const uint64_t off = nb_u32(nb, node, 6), cnt = nb_u16(nb, node, 14);
      st->leaf_count++;
      st->tri_count += cnt;
      if (depth < st->depth_min) st->depth_min = (uint32_t)depth;
      if (depth > st->depth_max) st->depth_max = (uint32_t)depth;
      if (cnt < st->leaf_tris_min) st->leaf_tris_min = (uint32_t)cnt;
      if (cnt > st->leaf_tris_max) st->leaf_tris_max = (uint32_t)cnt;
      if (off + cnt > n_slots) {
        st->structure_failures++;
        continue;
      }
      if (off < lo) lo = off;
      if (off + cnt > hi) hi = off + cnt;
      float tb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
      for (uint64_t s = off; s < off + cnt; s++) {
        if (seen[s]) st->structure_failures++;
        seen[s] = 1;
        double a[3], b[3], c[3];
        face_vertices(m, face_of_slot(m, s), a, b, c);
        for (int k = 0; k < 3; k++) {
          const double mn = fmin(a[k], fmin(b[k], c[k])), mx = fmax(a[k], fmax(b[k], c[k]));
          if (mn < tb[k]) tb[k] = (float)mn;
          if (mx > tb[k + 3]) tb[k + 3] = (float)mx;
        }
      }
      if (cnt && !box_contains(box, tb)) st->bounds_failures++;
    } else {
      const uint64_t left = node + 1, right = node + nb_u32(nb, node, 6);
      const uint32_t axis = nb_u32(nb, node, 7);
      if (axis > 2 || right <= node || right >= nb.n || left >= nb.n) {
        st->structure_failures++;
        continue;
      }
      if (!box_contains(box, nb_bounds(nb, left))) st->bounds_failures++;
      if (!box_contains(box, nb_bounds(nb, right))) st->bounds_failures++;
      stack[sp++] = right;
      stack[sp++] = depth + 1;
      stack[sp++] = left;
      stack[sp++] = depth + 1;
    }
  }
  /* the root's leaves tile one contiguous slot range */
  if (lo < hi)
    for (uint64_t s = lo; s < hi; s++)
      if (!seen[s]) st->structure_failures++;
  if (visited != nb.n) st->structure_failures++;
  free(seen);
  free(stack);
  return (st->bounds_failures || st->structure_failures) ? 1 : 0;
}

/* ------------------------------------------------------------------------------------------ */
/* refit (src/core/cast/refit.template.js:4-196)                                               */
/* ------------------------------------------------------------------------------------------ */

/* The node set of a partial refit: member[id] != 0 for the ids in nodeIndices, an id being the
 * node's index over all roots (this root's first node is `base`); NULL for nodeIndices = null. */
typedef struct {
  const uint8_t* member;
  uint64_t n_member;
  uint64_t base;
} refit_set;

static int refit_has(const refit_set* set, uint64_t node) {
  const uint64_t id = set->base + node;
  return id < set->n_member && set->member[id] != 0;
}

/* One node of _traverse: a leaf's box is the min / max over its triangles' vertices in slot order
 * (strict < and >), written only if a component differs (!==); an internal node is the union of its
 * children, recomputed only when a child changed. With a node set (refit.template.js:136-172), an
 * internal node not forced visits the children in the set, or, when neither is, both, forced (and a
 * forced node forces its children); a child not visited counts as unchanged. Returns whether the
 * node changed. */
static int refit_node(const orc_mesh* m, uint8_t* nodes, uint64_t n_nodes, uint64_t node, int depth,
                      const refit_set* set, int force) {
  if (node >= n_nodes || depth > 4096) return 0;
  float* box = (float*)(nodes + 32 * node);
  uint16_t flag, count;
  uint32_t word6;
  memcpy(&flag, nodes + 32 * node + 30, 2);
  memcpy(&count, nodes + 32 * node + 28, 2);
  memcpy(&word6, nodes + 32 * node + 24, 4);
  if (flag == 0xFFFF) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t s = word6; s < (uint64_t)word6 + count; s++) {
      const uint32_t face = face_of_slot(m, s);
      for (int j = 0; j < 3; j++) {
        uint64_t v = 3ull * face + j;
        if (m->index) v = rd(m->index, m->index_elt_bytes, v);
        const float* p = m->positions + v * m->stride + m->offset;
        for (int k = 0; k < 3; k++) {
          const double x = p[k];
          if (x < mn[k]) mn[k] = x;
          if (x > mx[k]) mx[k] = x;
        }
      }
    }
    int changed = 0;
    for (int k = 0; k < 3; k++) changed = changed || !((double)box[k] == mn[k]) || !((double)box[k + 3] == mx[k]);
    if (changed)
      for (int k = 0; k < 3; k++) {
        box[k] = (float)mn[k];
        box[k + 3] = (float)mx[k];
      }
    return changed;
  }
  const uint64_t left = node + 1, right = node + word6;
  int visit_left = 1, visit_right = 1, force_children = force;
  if (set && !force) {
    const int has_left = refit_has(set, left), has_right = refit_has(set, right);
    force_children = !has_left && !has_right;
    visit_left = force_children || has_left;
    visit_right = force_children || has_right;
  }
  const int lc = visit_left ? refit_node(m, nodes, n_nodes, left, depth + 1, set, force_children) : 0;
  const int rc = visit_right ? refit_node(m, nodes, n_nodes, right, depth + 1, set, force_children) : 0;
  if (lc || rc) {
    const float* lb = (const float*)(nodes + 32 * left);
    const float* rb = (const float*)(nodes + 32 * right);
    for (int k = 0; k < 3; k++) {
      box[k] = lb[k] < rb[k] ? lb[k] : rb[k];
      box[k + 3] = lb[k + 3] > rb[k + 3] ? lb[k + 3] : rb[k + 3];
    }
  }
  return lc || rc;
}

void orc_refit_root(const orc_mesh* m, uint8_t* nodes, uint64_t n_bytes) {
  if (n_bytes >= 32) refit_node(m, nodes, n_bytes / 32, 0, 0, NULL, 0);
}

void orc_refit_root_nodes(const orc_mesh* m, uint8_t* nodes, uint64_t n_bytes, uint64_t base, const uint8_t* member,
                          uint64_t n_member) {
  const refit_set set = {member, n_member, base};
  if (n_bytes >= 32) refit_node(m, nodes, n_bytes / 32, 0, 0, member ? &set : NULL, 0);
}

/* ------------------------------------------------------------------------------------------ */
/* intersectsSphere: brute force over the triangles with sphereIntersectTriangle              */
/* (src/math/MathUtilities.js:160-204)                                                        */
/* ------------------------------------------------------------------------------------------ */

/* Line3.closestPointToPoint(point, clampToLine = true) */
static void segment_closest(const double s[3], const double e[3], const double p[3], double out[3]) {
  double se[3], sp[3];
  v_sub(e, s, se);
  v_sub(p, s, sp);
  double t = v_dot(se, sp) / v_dot(se, se);
  t = fmax(0.0, fmin(1.0, t)); /* MathUtils.clamp: Math.max(min, Math.min(max, t)) */
  if (t != t) {
    out[0] = out[1] = out[2] = NAN; /* degenerate segment: 0 / 0 */
    return;
  }
  for (int k = 0; k < 3; k++) out[k] = s[k] + se[k] * t;
}

static int sphere_triangle(const double c[3], double r, const double a[3], const double b[3], const double cc[3]) {
  double q[3];
  const double* ends[3][2] = {{a, b}, {a, cc}, {b, cc}};
  for (int k = 0; k < 3; k++) {
    segment_closest(ends[k][0], ends[k][1], c, q);
    if (sqrt(v_dist_sq(q, c)) <= r) return 1;
  }
  /* Triangle.getPlane: setFromCoplanarPoints(a, b, c), normal = normalize((c - b) x (a - b)) */
  double cb[3], ab[3], n[3];
  v_sub(cc, b, cb);
  v_sub(a, b, ab);
  v_cross(cb, ab, n);
  const double len = sqrt(v_dot(n, n));
  for (int k = 0; k < 3; k++) n[k] /= len;
  const double constant = -v_dot(a, n);
  const double dist = v_dot(n, c) + constant;
  if (fabs(dist) <= r) {
    double pp[3], bc[3];
    for (int k = 0; k < 3; k++) pp[k] = c[k] - n[k] * dist; /* projectPoint */
    if (orc_barycoord(pp, a, b, cc, bc) && bc[0] >= 0 && bc[1] >= 0 && bc[0] + bc[1] <= 1) return 1;
  }
  return 0;
}

int orc_intersects_sphere(const orc_mesh* m, const float center[3], double radius) {
  const uint64_t n_tri = m->index ? m->index_count / 3 : m->vertex_count / 3;
  const double c[3] = {center[0], center[1], center[2]};
  for (uint64_t f = 0; f < n_tri; f++) {
    double a[3], b[3], cc[3];
    face_vertices(m, (uint32_t)f, a, b, cc);
    if (sphere_triangle(c, radius, a, b, cc)) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* intersectsBox (src/core/MeshBVH.js:780-792): whether any triangle intersects an oriented    */
/* box, decided as OrientedBox.intersectsTriangle does (src/math/OrientedBox.js:190-260 after   */
/* update, :79-130) with SeparatingAxisBounds (src/math/SeparatingAxisBounds.js) and the        */
/* ExtendedTriangle.update axes (src/math/ExtendedTriangle.js:47-75), in double, in the         */
/* reference's operation order (three's Vector3.applyMatrix4 / dot / crossVectors and           */
/* Triangle.getNormal). Brute force over the triangles.                                          */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  double mn, mx;
} sab;

static sab sab_points(const double ax[3], const double (*p)[3], int n) {
  sab b = {INFINITY, -INFINITY};
  for (int i = 0; i < n; i++) {
    const double v = ax[0] * p[i][0] + ax[1] * p[i][1] + ax[2] * p[i][2];
    b.mn = v < b.mn ? v : b.mn;
    b.mx = v > b.mx ? v : b.mx;
  }
  return b;
}

static int sab_separated(sab a, sab b) { return a.mn > b.mx || b.mn > a.mx; }

static void obb_corners(const double bmin[3], const double bmax[3], const double e[16], double p[8][3], double ax[3][3],
                        sab sb[3]) {
  for (int x = 0; x <= 1; x++)
    for (int y = 0; y <= 1; y++)
      for (int z = 0; z <= 1; z++) {
        const int i = x | (y << 1) | (z << 2);
        const double vx = x ? bmax[0] : bmin[0], vy = y ? bmax[1] : bmin[1], vz = z ? bmax[2] : bmin[2];
        const double w = 1 / (e[3] * vx + e[7] * vy + e[11] * vz + e[15]);
        p[i][0] = (e[0] * vx + e[4] * vy + e[8] * vz + e[12]) * w;
        p[i][1] = (e[1] * vx + e[5] * vy + e[9] * vz + e[13]) * w;
        p[i][2] = (e[2] * vx + e[6] * vy + e[10] * vz + e[14]) * w;
      }
  for (int i = 0; i < 3; i++) {
    for (int k = 0; k < 3; k++) ax[i][k] = p[0][k] - p[1 << i][k];
    sb[i] = sab_points(ax[i], (const double(*)[3])p, 8);
  }
}

static void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

int orc_box_triangle_sat(const double bmin[3], const double bmax[3], const double e[16], const double a[3],
                         const double b[3], const double c[3]) {
  double p[8][3], ax[3][3];
  sab sb[3];
  obb_corners(bmin, bmax, e, p, ax, sb);
  const double pts[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
  double tax[4][3], cb[3], ab[3];
  for (int k = 0; k < 3; k++) {
    cb[k] = c[k] - b[k];
    ab[k] = a[k] - b[k];
  }
  cross3(cb, ab, tax[0]);
  const double l2 = tax[0][0] * tax[0][0] + tax[0][1] * tax[0][1] + tax[0][2] * tax[0][2];
  if (l2 > 0) {
    const double s = 1 / sqrt(l2);
    for (int k = 0; k < 3; k++) tax[0][k] = tax[0][k] * s;
  } else {
    tax[0][0] = tax[0][1] = tax[0][2] = 0;
  }
  for (int k = 0; k < 3; k++) {
    tax[1][k] = a[k] - b[k];
    tax[2][k] = b[k] - c[k];
    tax[3][k] = c[k] - a[k];
  }
  for (int i = 0; i < 3; i++)
    if (sab_separated(sb[i], sab_points(ax[i], pts, 3))) return 0;
  for (int i = 0; i < 3; i++)
    if (sab_separated(sab_points(tax[i], pts, 3), sab_points(tax[i], (const double(*)[3])p, 8))) return 0;
  for (int i = 0; i < 3; i++)
    for (int i2 = 0; i2 < 4; i2++) {
      double axis[3];
      cross3(ax[i], tax[i2], axis);
      if (sab_separated(sab_points(axis, pts, 3), sab_points(axis, (const double(*)[3])p, 8))) return 0;
    }
  return 1;
}

int orc_intersects_box(const orc_mesh* m, const double bmin[3], const double bmax[3], const double e[16]) {
  const uint64_t n_tri = m->index ? m->index_count / 3 : m->vertex_count / 3;
  for (uint64_t f = 0; f < n_tri; f++) {
    double a[3], b[3], c[3];
    face_vertices(m, (uint32_t)f, a, b, c);
    if (orc_box_triangle_sat(bmin, bmax, e, a, b, c)) return 1;
  }
  return 0;
}

/* An independent formulation for checking the one above: the triangle brought into the box frame by
   the inverse of an affine e (cofactors, double), then clipped against the six slabs
   (Sutherland-Hodgman, points on a plane kept); intersecting iff something is left. */
int orc_box_triangle_clip(const double bmin[3], const double bmax[3], const double e[16], const double a[3],
                          const double b[3], const double c[3]) {
  const double m00 = e[0], m01 = e[4], m02 = e[8], m10 = e[1], m11 = e[5], m12 = e[9], m20 = e[2], m21 = e[6],
               m22 = e[10];
  const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
  if (det == 0) return -1;
  const double inv[3][3] = {{(m11 * m22 - m12 * m21) / det, (m02 * m21 - m01 * m22) / det, (m01 * m12 - m02 * m11) / det},
                            {(m12 * m20 - m10 * m22) / det, (m00 * m22 - m02 * m20) / det, (m02 * m10 - m00 * m12) / det},
                            {(m10 * m21 - m11 * m20) / det, (m01 * m20 - m00 * m21) / det, (m00 * m11 - m01 * m10) / det}};
  double poly[16][3], tmp[16][3];
  int n = 3;
  const double* src[3] = {a, b, c};
  for (int v = 0; v < 3; v++) {
    const double d[3] = {src[v][0] - e[12], src[v][1] - e[13], src[v][2] - e[14]};
    for (int k = 0; k < 3; k++) poly[v][k] = inv[k][0] * d[0] + inv[k][1] * d[1] + inv[k][2] * d[2];
  }
  for (int plane = 0; plane < 6 && n > 0; plane++) {
    const int k = plane % 3;
    const int keep_above = plane < 3;  /* x >= min, then x <= max */
    const double lim = keep_above ? bmin[k] : bmax[k];
    int m = 0;
    for (int i = 0; i < n; i++) {
      const double* P = poly[i];
      const double* Q = poly[(i + 1) % n];
      const double dp = keep_above ? P[k] - lim : lim - P[k];
      const double dq = keep_above ? Q[k] - lim : lim - Q[k];
      if (dp >= 0) {
        for (int j = 0; j < 3; j++) tmp[m][j] = P[j];
        m++;
      }
      if ((dp >= 0) != (dq >= 0)) {
        const double t = dp / (dp - dq);
        for (int j = 0; j < 3; j++) tmp[m][j] = P[j] + (Q[j] - P[j]) * t;
        m++;
      }
    }
    n = m;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < 3; j++) poly[i][j] = tmp[i][j];
  }
  return n > 0;
}

/* ------------------------------------------------------------------------------------------ */
/* closestPointToGeometry                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* Line3.at: (end - start) * t + start */
static void line_at(const double s[3], const double e[3], double t, double o[3]) {
  for (int k = 0; k < 3; k++) o[k] = (e[k] - s[k]) * t + s[k];
}
/* Line3.closestPointToPoint(p, clamp = true) */
static void line_closest(const double s[3], const double e[3], const double p[3], double o[3]) {
  double sp[3], se[3];
  v_sub(p, s, sp);
  v_sub(e, s, se);
  double t = v_dot(se, sp) / v_dot(se, se);
  if (!isnan(t)) t = fmax(0.0, fmin(1.0, t)); /* MathUtils.clamp (Math.max / Math.min keep a NaN) */
  line_at(s, e, t, o);
}

/* closestPointsSegmentToSegment (src/math/MathUtilities.js:58-150, after MathGeoLib's
   LineSegment::ClosestPoint) with closestPointLineToLine (MathUtilities.js:3-56). */
void orc_segment_segment(const double s1[3], const double e1[3], const double s2[3], const double e2[3], double o1[3],
                         double o2[3]) {
  double v02[3], v10[3], v32[3];
  v_sub(s1, s2, v02);
  v_sub(e1, s1, v10);
  v_sub(e2, s2, v32);
  const double d0232 = v_dot(v02, v32), d3210 = v_dot(v32, v10), d3232 = v_dot(v32, v32), d0210 = v_dot(v02, v10),
               d1010 = v_dot(v10, v10);
  const double denom = d1010 * d3232 - d3210 * d3210;
  const double d = denom != 0 ? (d0232 * d3210 - d0210 * d3232) / denom : 0;
  const double d2 = (d0232 + d * d3210) / d3232;
  const int in1 = d >= 0 && d <= 1, in2 = d2 >= 0 && d2 <= 1;
  if (in1 && in2) {
    line_at(s1, e1, d, o1);
    line_at(s2, e2, d2, o2);
  } else if (in1) {
    memcpy(o2, d2 < 0 ? s2 : e2, sizeof(double) * 3);
    line_closest(s1, e1, o2, o1);
  } else if (in2) {
    memcpy(o1, d < 0 ? s1 : e1, sizeof(double) * 3);
    line_closest(s2, e2, o1, o2);
  } else {
    const double* p = d < 0 ? s1 : e1;
    const double* p2 = d2 < 0 ? s2 : e2;
    double c1[3], c2[3];
    line_closest(s1, e1, p2, c1);
    line_closest(s2, e2, p, c2);
    if (v_dist_sq(c1, p2) <= v_dist_sq(c2, p)) {
      memcpy(o1, c1, sizeof(c1));
      memcpy(o2, p2, sizeof(double) * 3);
    } else {
      memcpy(o1, p, sizeof(double) * 3);
      memcpy(o2, c2, sizeof(c2));
    }
  }
}

/* Where segment pq meets triangle abc, by the parametric solve p + t (q - p) = a + u (b - a) + v (c - a)
   (Cramer's rule as in Moller-Trumbore); 0 for a segment parallel to the triangle's plane. */
static int segment_meets_triangle(const double p[3], const double q[3], const double a[3], const double b[3],
                                  const double c[3], double x[3]) {
  double e1[3], e2[3], d[3], h[3], s[3], qv[3];
  v_sub(b, a, e1);
  v_sub(c, a, e2);
  v_sub(q, p, d);
  v_cross(d, e2, h);
  const double det = v_dot(e1, h);
  if (det == 0) return 0;
  const double f = 1 / det;
  v_sub(p, a, s);
  const double u = f * v_dot(s, h);
  if (u < 0 || u > 1) return 0;
  v_cross(s, e1, qv);
  const double v = f * v_dot(d, qv);
  if (v < 0 || u + v > 1) return 0;
  const double t = f * v_dot(e2, qv);
  if (t < 0 || t > 1) return 0;
  for (int k = 0; k < 3; k++) x[k] = p[k] + d[k] * t;
  return 1;
}

/* ExtendedTriangle.distanceToTriangle (src/math/ExtendedTriangle.js:774-860): 0 when the triangles
   intersect (the points: the middle of the intersection segment, whose ends are where edges of one
   triangle meet the other), else the least of the vertex-to-triangle and edge-to-edge distances in
   the reference's order, strict improvements only. */
double orc_triangle_distance(const double* t1, const double* t2, double o1[3], double o2[3]) {
  const double* A[3] = {t1, t1 + 3, t1 + 6};
  const double* B[3] = {t2, t2 + 3, t2 + 6};
  double n1[3], n2[3], ea[3], eb[3], line[3];
  v_sub(A[1], A[0], ea);
  v_sub(A[2], A[0], eb);
  v_cross(ea, eb, n1);
  v_sub(B[1], B[0], ea);
  v_sub(B[2], B[0], eb);
  v_cross(ea, eb, n2);
  v_cross(n1, n2, line);
  int found = 0;
  double lo = INFINITY, hi = -INFINITY, plo[3] = {0, 0, 0}, phi[3] = {0, 0, 0};
  for (int side = 0; side < 2; side++) {
    const double* const* E = side ? B : A;
    const double* const* T = side ? A : B;
    for (int k = 0; k < 3; k++) {
      double x[3];
      if (!segment_meets_triangle(E[k], E[(k + 1) % 3], T[0], T[1], T[2], x)) continue;
      const double s = v_dot(x, line);
      if (!found || s < lo) {
        lo = s;
        memcpy(plo, x, sizeof(x));
      }
      if (!found || s > hi) {
        hi = s;
        memcpy(phi, x, sizeof(x));
      }
      found = 1;
    }
  }
  if (found) {
    for (int k = 0; k < 3; k++) o1[k] = o2[k] = (plo[k] + phi[k]) * 0.5;
    return 0.0;
  }
  double best = INFINITY, q[3];
  for (int i = 0; i < 3; i++) {
    orc_closest_point_triangle(B[i], A[0], A[1], A[2], q);
    double d = v_dist_sq(B[i], q);
    if (d < best) {
      best = d;
      memcpy(o1, q, sizeof(q));
      memcpy(o2, B[i], sizeof(q));
    }
    orc_closest_point_triangle(A[i], B[0], B[1], B[2], q);
    d = v_dist_sq(A[i], q);
    if (d < best) {
      best = d;
      memcpy(o1, A[i], sizeof(q));
      memcpy(o2, q, sizeof(q));
    }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double c1[3], c2[3];
      orc_segment_segment(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3], c1, c2);
      const double d = v_dist_sq(c1, c2);
      if (d < best) {
        best = d;
        memcpy(o1, c1, sizeof(c1));
        memcpy(o2, c2, sizeof(c2));
      }
    }
  return sqrt(best);
}

/* Vector3.applyMatrix4 */
static void apply_m4(const double e[16], const double v[3], double o[3]) {
  const double x = v[0], y = v[1], z = v[2];
  const double w = 1 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
  o[0] = (e[0] * x + e[4] * y + e[8] * z + e[12]) * w;
  o[1] = (e[1] * x + e[5] * y + e[9] * z + e[13]) * w;
  o[2] = (e[2] * x + e[6] * y + e[10] * z + e[14]) * w;
}

typedef struct {
  const orc_mesh* m;
  const double* t2;  /* 9 doubles */
  double mn[3], mx[3];
  double best;      /* this other triangle's best */
  double bound;     /* the best over earlier other triangles */
  double max_t;
  uint32_t best_slot;
  double o1[3], o2[3];
} GeoCtx;

/* A lower bound of the distance from the triangle's box to a node box (the gap less 1e-12 of the
   largest coordinate, as roundings of the triangle distance can go below the exact gap). */
static double geo_box_lb(const GeoCtx* g, const float* b) {
  double s = 0, m = 0;
  for (int k = 0; k < 3; k++) {
    const double gap = fmax(fmax((double)b[k] - g->mx[k], g->mn[k] - (double)b[k + 3]), 0.0);
    s += gap * gap;
    m = fmax(m, fmax(fmax(fabs((double)b[k]), fabs((double)b[k + 3])), fmax(fabs(g->mn[k]), fabs(g->mx[k]))));
  }
  return fmax(sqrt(s) * (1 - 1e-12) - 1e-12 * m, 0.0);
}

static void geo_leaf(GeoCtx* g, uint64_t off, uint64_t cnt) {
  for (uint64_t s = off; s < off + cnt; s++) {
    double t1[9], o1[3], o2[3];
    face_vertices(g->m, face_of_slot(g->m, s), t1, t1 + 3, t1 + 6);
    const double d = orc_triangle_distance(t1, g->t2, o1, o2);
    if (d < g->max_t && (d < g->best || (d == g->best && s < g->best_slot))) {
      g->best = d;
      g->best_slot = (uint32_t)s;
      memcpy(g->o1, o1, sizeof(o1));
      memcpy(g->o2, o2, sizeof(o2));
    }
  }
}

static void geo_node(GeoCtx* g, NodeBuf nb, uint64_t node) {
  if (nb_leaf(nb, node)) {
    geo_leaf(g, nb_u32(nb, node, 6), nb_u16(nb, node, 14));
    return;
  }
  const uint64_t kids[2] = {node + 1, node + nb_u32(nb, node, 6)};
  double s[2] = {geo_box_lb(g, nb_bounds(nb, kids[0])), geo_box_lb(g, nb_bounds(nb, kids[1]))};
  const int first = s[1] < s[0];
  for (int i = 0; i < 2; i++) {
    const int c = i ? 1 - first : first;
    if (s[c] <= fmin(g->best, g->bound) && s[c] < g->max_t) geo_node(g, nb, kids[c]);
  }
}

/* closestPointToGeometry (src/core/cast/closestPointToGeometry.template.js:15-271): the closest pair
   over every triangle of the other geometry (brought into the mesh frame by e, applyMatrix4) and every
   triangle of the mesh; pairs at or beyond max_t do not count. Equal distances: the lowest other
   triangle, then the lowest mesh slot. brute != 0 tries every pair; otherwise each other triangle
   walks the tree, skipping nodes whose box lower bound exceeds the best so far. Returns the other
   triangle (NO_HIT when no pair counts) and writes the slot, distance and points (mesh frame). */
uint32_t orc_closest_to_geometry(const orc_mesh* m, const float* positions, uint64_t n_vertices, uint32_t stride,
                                 uint32_t offset, const void* index, int elt, uint64_t n_tris, const double e[16],
                                 double max_t, int brute, uint32_t* slot, double* dist, double p1[3], double p2[3]) {
  (void)n_vertices;
  uint32_t win = NO_HIT;
  double best = INFINITY;
  *slot = NO_HIT;
  *dist = INFINITY;
  for (uint64_t j = 0; j < n_tris; j++) {
    double t2[9];
    for (int k = 0; k < 3; k++) {
      const uint64_t v = index ? rd(index, elt, 3 * j + k) : 3 * j + k;
      const float* p = positions + v * stride + offset;
      const double src[3] = {p[0], p[1], p[2]};
      apply_m4(e, src, t2 + 3 * k);
    }
    GeoCtx g;
    memset(&g, 0, sizeof(g));
    g.m = m;
    g.t2 = t2;
    g.best = INFINITY;
    g.bound = best;
    g.max_t = max_t;
    g.best_slot = NO_HIT;
    for (int k = 0; k < 3; k++) {
      g.mn[k] = fmin(fmin(t2[k], t2[3 + k]), t2[6 + k]);
      g.mx[k] = fmax(fmax(t2[k], t2[3 + k]), t2[6 + k]);
    }
    for (uint32_t r = 0; r < m->n_roots; r++) {
      NodeBuf nb = root_buf(m, r);
      if (brute) {
        for (uint64_t k = 0; k < nb.n; k++)
          if (nb_leaf(nb, k)) geo_leaf(&g, nb_u32(nb, k, 6), nb_u16(nb, k, 14));
      } else {
        geo_node(&g, nb, 0);
      }
    }
    if (g.best_slot != NO_HIT && g.best < best) {
      best = g.best;
      win = (uint32_t)j;
      *slot = g.best_slot;
      *dist = g.best;
      memcpy(p1, g.o1, sizeof(g.o1));
      memcpy(p2, g.o2, sizeof(g.o2));
    }
  }
  return win;
}

/* The scene closest point (bvh_ClosestPointToPoint over a ClusteredBVH,
   src/webgpu/shapecastFns/getClosestPointToPointFn.js:20-151) by brute force: every triangle of every
   visible instance taken to world space (applyMatrix4 of the float matrix in double), three's
   Triangle.closestPointToPoint, the least squared world distance; equal distances to the lowest
   instance, then the lowest face. meshes[g] is instance geometry g; world: 16 floats per instance.
   Returns the instance (NO_HIT when none) and writes the face, squared distance and world point. */
uint32_t orc_scene_closest_point(const orc_mesh* const* meshes, const float* world, const uint32_t* geometry,
                                 const uint8_t* visible, uint32_t n_instances, const float p[3], uint32_t* face,
                                 double* d2_out, double point[3]) {
  const double q[3] = {p[0], p[1], p[2]};
  double best = INFINITY;
  uint32_t win = NO_HIT;
  *face = NO_HIT;
  for (uint32_t i = 0; i < n_instances; i++) {
    if (visible && !visible[i]) continue;
    const orc_mesh* m = meshes[geometry[i]];
    double e[16];
    for (int k = 0; k < 16; k++) e[k] = world[16 * (uint64_t)i + k];
    const uint64_t n_tri = m->index ? m->index_count / 3 : m->vertex_count / 3;
    for (uint64_t f = 0; f < n_tri; f++) {
      double a[3], b[3], c[3], wa[3], wb[3], wc[3], cp[3];
      face_vertices(m, (uint32_t)f, a, b, c);
      apply_m4(e, a, wa);
      apply_m4(e, b, wb);
      apply_m4(e, c, wc);
      orc_closest_point_triangle(q, wa, wb, wc, cp);
      const double d2 = v_dist_sq(q, cp);
      if (d2 < best) {
        best = d2;
        win = i;
        *face = (uint32_t)f;
        memcpy(point, cp, sizeof(cp));
      }
    }
  }
  *d2_out = best;
  return win;
}
