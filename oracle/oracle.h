/*
 * oracle.h — FP64 CPU restatement of the three-mesh-bvh query path. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the tests, the smoke check and the benchmark's CPU baseline use.
 * Nothing in the product (libmbvh.so, mbvh.node, the JS package) may load or call it.
 *
 * It restates, in IEEE double and in the reference's operation order:
 *   - three.js Ray.intersectTriangle (GeometricTools IntrRay3Triangle3), Triangle.getBarycoord,
 *     Triangle.closestPointToPoint (Ericson, Real-Time Collision Detection §5.1.5), as called by
 *     src/utils/ThreeRayIntersectUtilities.js:21-145 and src/core/cast/closestPointToPoint.js:39-61;
 *     three@0.185.0 is not vendored in the reference, so these follow the published algorithms;
 *   - the reference traversals over its packed node buffers:
 *       raycastFirst   src/core/MeshBVH.js:558-577, src/core/cast/raycastFirst.template.js:9-103,
 *                      src/core/utils/intersectUtils.js:2-78, iterationUtils.template.js:24-53
 *       raycast        src/core/MeshBVH.js:532-545, src/core/cast/raycast.template.js:7-53
 *       closest point  src/core/MeshBVH.js:874-884, src/core/cast/closestPointToPoint.js:6-78,
 *                      src/core/BVH.js:342-413, src/core/cast/shapecast.js:41-214
 *   - brute force over every triangle (BVH-independent ground truth for closest hits/points);
 *   - validateBounds (src/debug/Debug.js:207-267) and tree statistics (getBVHExtremes).
 * Every query function also counts, per query, the node bounds it read and the triangles it
 * tested in the reference traversal order (the roofline's algorithmic-bytes numerator).
 */
#ifndef MBVH_ORACLE_H
#define MBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh {
  const float* positions;      /* vertex v axis k at positions[v*stride + offset + k] */
  uint64_t vertex_count;
  uint32_t stride, offset;
  const void* index;           /* NULL for non-indexed geometry */
  int32_t index_elt_bytes;     /* 2 or 4 */
  uint64_t index_count;
  const void* indirect;        /* NULL unless built with `indirect` */
  int32_t indirect_elt_bytes;
  uint64_t indirect_count;
  const void* const* roots;    /* reference-format node buffers */
  const uint64_t* root_bytes;
  uint32_t n_roots;
  const uint8_t* face_mask;    /* optional per-face facing mask (bit0 front, bit1 back) */
} orc_mesh;

/* three.js side constants */
enum { ORC_FRONT = 0, ORC_BACK = 1, ORC_DOUBLE = 2, ORC_PER_FACE = 3 };

/* Ray.intersectTriangle: returns 1 on hit and writes t and the point. */
int orc_ray_triangle(const double o[3], const double d[3], const double a[3], const double b[3],
                     const double c[3], int backface_culling, double* t, double point[3]);
/* Triangle.closestPointToPoint */
void orc_closest_point_triangle(const double p[3], const double a[3], const double b[3],
                                const double c[3], double out[3]);
/* Triangle.getBarycoord; returns 0 for a degenerate triangle (target zeroed). */
int orc_barycoord(const double p[3], const double a[3], const double b[3], const double c[3],
                  double out[3]);

/* First hit per ray (reference traversal, or brute force when brute != 0). Outputs may be NULL
 * except face. Miss: face 0xFFFFFFFF, dist +inf. */
int orc_raycast_first(const orc_mesh* m, const float* origins, const float* dirs, uint64_t n,
                      int side, double near_d, double far_d, int brute, int nthreads,
                      uint32_t* face, double* dist, double* point, double* bary,
                      uint32_t* nodes_read, uint32_t* tris_tested);

/* All hits of one ray in reference order (left-first traversal, leaf order). Writes at most
 * `cap` hits and returns the total hit count. */
int64_t orc_raycast_all(const orc_mesh* m, const float origin[3], const float dir[3], int side,
                        double near_d, double far_d, uint64_t cap, uint32_t* face, double* dist);

/* Closest point per query (reference traversal, or brute force when brute != 0). */
int orc_closest_point(const orc_mesh* m, const float* points, uint64_t n, double min_t,
                      double max_t, int brute, int nthreads, uint32_t* face, double* dist,
                      double* point, uint32_t* nodes_read, uint32_t* tris_tested);

/* The same for double-precision rays / points (three's Ray and Vector3 hold JS numbers). */
int orc_raycast_first_f64(const orc_mesh* m, const double* origins, const double* dirs, uint64_t n, int side,
                          double near_d, double far_d, int brute, int nthreads, uint32_t* face, double* dist,
                          double* point, double* bary, uint32_t* nodes_read, uint32_t* tris_tested);
int orc_closest_point_f64(const orc_mesh* m, const double* points, uint64_t n, double min_t, double max_t, int brute,
                          int nthreads, uint32_t* face, double* dist, double* point, uint32_t* nodes_read,
                          uint32_t* tris_tested);
int64_t orc_raycast_all_f64(const orc_mesh* m, const double origin[3], const double dir[3], int side,
                            double near_d, double far_d, uint64_t cap, uint32_t* face, double* dist);
double orc_ray_plane_distance_f64(const orc_mesh* m, uint32_t face, const double o[3], const double d[3]);
double orc_point_face_distance_f64(const orc_mesh* m, uint32_t face, const double q[3]);

/* FP64 distance from the ray origin to triangle `face` along the ray's line (the plane hit, no
 * edge test); used to decide whether two faces tie for a ray. Returns +inf if parallel. */
double orc_ray_plane_distance(const orc_mesh* m, uint32_t face, const float origin[3],
                              const float dir[3]);
/* FP64 distance from p to triangle `face`. */
double orc_point_face_distance(const orc_mesh* m, uint32_t face, const float p[3]);

typedef struct orc_tree_stats {
  uint64_t node_count, leaf_count, tri_count;
  uint32_t depth_min, depth_max;
  uint32_t leaf_tris_min, leaf_tris_max;
  uint64_t bounds_failures;   /* validateBounds failures */
  uint64_t structure_failures;/* malformed offsets / axes / coverage */
} orc_tree_stats;

/* Validates one root: node layout, child containment, leaf ⊇ triangles, every slot of
 * [first, first+count) of that root in exactly one leaf. Returns 0 when all checks pass. */
int orc_validate_root(const orc_mesh* m, uint32_t root, orc_tree_stats* stats);

/* MeshBVH.refit with nodeIndices = null (src/core/cast/refit.template.js:4-196): rewrites the boxes
 * of one root's node buffer `nodes` (n_bytes, modified in place) from the mesh's positions, index
 * and indirect buffer. */
void orc_refit_root(const orc_mesh* m, uint8_t* nodes, uint64_t n_bytes);
/* refit(nodeIndices) for one root whose first node has index `base` over all roots: member[id] != 0
 * (id < n_member) for the ids in the set; member NULL is nodeIndices = null. */
void orc_refit_root_nodes(const orc_mesh* m, uint8_t* nodes, uint64_t n_bytes, uint64_t base, const uint8_t* member,
                          uint64_t n_member);

/* MeshBVH.intersectsSphere by brute force over the geometry's triangles with the reference's
 * sphereIntersectTriangle (src/math/MathUtilities.js:160-204): segment closest points, then the
 * plane distance and projected-point containment. */
int orc_intersects_sphere(const orc_mesh* m, const float center[3], double radius);
/* intersectsBox: any triangle intersecting the box (min / max in its frame, e = box -> mesh, 16 doubles
   column-major), OrientedBox's separating-axis test in double (brute force). */
int orc_intersects_box(const orc_mesh* m, const double bmin[3], const double bmax[3], const double e[16]);
int orc_box_triangle_sat(const double bmin[3], const double bmax[3], const double e[16], const double a[3],
                         const double b[3], const double c[3]);
/* The same question by clipping the triangle to the box in the box frame (affine e; -1 if singular). */
int orc_box_triangle_clip(const double bmin[3], const double bmax[3], const double e[16], const double a[3],
                          const double b[3], const double c[3]);

/* Scene closest point by brute force over every visible instance's world triangles. */
uint32_t orc_scene_closest_point(const orc_mesh* const* meshes, const float* world, const uint32_t* geometry,
                                 const uint8_t* visible, uint32_t n_instances, const float p[3], uint32_t* face,
                                 double* d2_out, double point[3]);
/* closestPointsSegmentToSegment (MathGeoLib's form, as the reference restates it) */
void orc_segment_segment(const double s1[3], const double e1[3], const double s2[3], const double e2[3], double o1[3],
                         double o2[3]);
/* ExtendedTriangle.distanceToTriangle of two triangles (9 doubles each); writes the closest points. */
double orc_triangle_distance(const double* t1, const double* t2, double o1[3], double o2[3]);
/* closestPointToGeometry over the mesh and another geometry placed by e (16 doubles, column-major). */
uint32_t orc_closest_to_geometry(const orc_mesh* m, const float* positions, uint64_t n_vertices, uint32_t stride,
                                 uint32_t offset, const void* index, int elt, uint64_t n_tris, const double e[16],
                                 double max_t, int brute, uint32_t* slot, double* dist, double p1[3], double p2[3]);

#ifdef __cplusplus
}
#endif

#endif
