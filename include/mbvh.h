/*
 * mbvh.h — C ABI of the MI355X-native three-mesh-bvh query engine (libmbvh.so).
 *
 * The reference library (three-mesh-bvh 0.9.14) has no FFI: its boundary is the JavaScript
 * method API of MeshBVH. Each entry point below is the native half of one of those methods and
 * cites the reference code it replaces. The Node-API addon (csrc/napi/mbvh_napi.c) and the
 * Python ctypes binding are the only callers; both marshal typed arrays into these signatures.
 *
 * Conventions
 *   - Plain pointers and sizes only. Every buffer is borrowed for the duration of the call.
 *   - Functions return 0 (MBVH_OK) or a negative mbvh_status. mbvh_last_error() describes the
 *     most recent failure on the calling thread.
 *   - "_d" entry points take DEVICE pointers and enqueue asynchronously on `stream` (a
 *     hipStream_t, or NULL for the context's own stream). They allocate nothing and never
 *     synchronise, so a caller may capture them in a HIP graph. The other query entry points
 *     take HOST pointers and are synchronous.
 *   - All query kernels of one context share that context's work-queue counter and traversal
 *     stack overflow region. The "_d" calls of one context must therefore be serialised: issue
 *     them on one stream (or order the streams with events). Two "_d" calls of one context that
 *     run at the same time on different streams overwrite each other's spilled stack entries.
 *     Use one context per stream for concurrent queries.
 *   - Vectors are packed float triples (x, y, z) per element.
 *   - A miss is reported as face index MBVH_NO_HIT and distance +Infinity.
 */
#ifndef MBVH_H
#define MBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBVH_ABI_VERSION 3u
#define MBVH_NO_HIT 0xFFFFFFFFu

typedef enum mbvh_status {
  MBVH_OK = 0,
  MBVH_ERR_NO_DEVICE = -1,   /* no HIP device, or the device index is out of range */
  MBVH_ERR_OOM = -2,         /* device or host allocation failed */
  MBVH_ERR_BAD_HANDLE = -3,  /* NULL or foreign context / mesh / tree / batch handle */
  MBVH_ERR_BAD_ARGS = -4,    /* invalid sizes, strides, ranges or option values */
  MBVH_ERR_HIP = -5,         /* a HIP runtime call failed (message has the HIP error) */
  MBVH_ERR_LIMIT = -6        /* input exceeds a documented limit of this build */
} mbvh_status;

/* three.js side constants (FrontSide = 0, BackSide = 1, DoubleSide = 2). */
typedef enum mbvh_side {
  MBVH_FRONT_SIDE = 0,
  MBVH_BACK_SIDE = 1,
  MBVH_DOUBLE_SIDE = 2,
  MBVH_SIDE_PER_TRIANGLE = 3 /* use the per-triangle facing mask set by mbvh_mesh_set_side_mask */
} mbvh_side;

/* Split strategies, same values as src/core/Constants.js:9,19,29 (CENTER, AVERAGE, SAH). */
typedef enum mbvh_strategy { MBVH_CENTER = 0, MBVH_AVERAGE = 1, MBVH_SAH = 2 } mbvh_strategy;

typedef struct mbvh_ctx mbvh_ctx;     /* one HIP device + one stream + scratch buffers */
typedef struct mbvh_tree mbvh_tree;   /* host-side build result (reference node buffers) */
typedef struct mbvh_mesh mbvh_mesh;   /* device-resident BVH + triangles of one MeshBVH */
typedef struct mbvh_batch mbvh_batch; /* several meshes concatenated for one batched launch */
typedef struct mbvh_scene mbvh_scene; /* instances of meshes under a top-level tree (TLAS) */

/* ------------------------------------------------------------------------------------------ */
/* Library                                                                                    */
/* ------------------------------------------------------------------------------------------ */

uint32_t mbvh_abi_version(void);
/* Message for the most recent failing call on this thread ("" if none). */
const char* mbvh_last_error(void);
/* Number of visible HIP devices (0 without a GPU; never fails). */
int mbvh_device_count(void);

/* ------------------------------------------------------------------------------------------ */
/* Host build (no device needed)                                                              */
/* ------------------------------------------------------------------------------------------ */

typedef struct mbvh_build_opts {
  int32_t strategy;         /* mbvh_strategy; reference default CENTER (Constants.js:75) */
  int32_t max_depth;        /* default 40 (Constants.js:76) */
  int32_t target_leaf_size; /* default 10 (Constants.js:77); `maxLeafTris` alias maps here */
  int32_t indirect;         /* 0: permute `index` in place; 1: build an indirect buffer */
  int32_t strict_leaf_size; /* 0: none; else no leaf holds more primitives (unless max_depth stops it):
                               an over-full node is split even against the heuristic, at the median
                               of its order on its box's longest axis when the heuristic finds no
                               split (the _strictLeafSize of src/core/build/buildTree.js:59-94, which
                               ClusteredBVH sets to 1); host build only */
} mbvh_build_opts;

/*
 * Builds the packed BVH in the reference's 32-byte node format.
 * Replaces buildPackedTree (src/core/build/buildTree.js:123-152) together with
 * MeshBVH.computePrimitiveBounds (src/core/MeshBVH.js:337-436), getBounds
 * (src/core/build/computeBoundsUtils.js:4-68), getOptimalSplit (src/core/build/splitUtils.js:21-294),
 * partition (src/core/build/sortUtils.js:4-62) and populateBuffer (src/core/build/buildUtils.js:20-91).
 *
 * positions       float32 vertex data; vertex v, axis k is positions[v*stride + offset + k]
 * index           geometry index (elt bytes 2 or 4), or NULL for a non-indexed geometry (only
 *                 allowed when opts->indirect). Non-indirect builds permute its triangle triplets
 *                 IN PLACE, as the reference does through bvh.primitiveBuffer (GeometryBVH.js:86-89).
 * root_ranges     n_roots (offset, count) pairs in triangles, as getRootPrimitiveRanges
 *                 (src/core/build/geometryUtils.js:87-159) returns them. With opts->indirect the
 *                 ranges only select the triangles listed in the generated indirect buffer and a
 *                 single root covers it (generateIndirectBuffer GeometryBVH.js:9-35,
 *                 getRootRanges GeometryBVH.js:169-181).
 */
int mbvh_build(const float* positions, uint64_t vertex_count, uint32_t position_stride,
               uint32_t position_offset, void* index, int32_t index_elt_bytes, uint64_t index_count,
               const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
               mbvh_tree** out_tree);
/* Number of roots (= MeshBVH._roots.length), indirect buffer element size (0 when none) and count. */
int mbvh_tree_info(const mbvh_tree* tree, uint32_t* n_roots, int32_t* indirect_elt_bytes,
                   uint64_t* indirect_count);
int mbvh_tree_root_bytes(const mbvh_tree* tree, uint32_t root, uint64_t* bytes);
int mbvh_tree_copy_root(const mbvh_tree* tree, uint32_t root, void* dst, uint64_t dst_bytes);
int mbvh_tree_copy_indirect(const mbvh_tree* tree, void* dst, uint64_t dst_bytes);
void mbvh_tree_free(mbvh_tree* tree);

/* ------------------------------------------------------------------------------------------ */
/* Device context and upload-once meshes                                                      */
/* ------------------------------------------------------------------------------------------ */

int mbvh_ctx_create(int32_t device, mbvh_ctx** out_ctx);
void mbvh_ctx_destroy(mbvh_ctx* ctx);
/* raycastFirst traversal width: 1 lane per ray on the binary tree, 4 lanes per ray on a 4-wide
 * tree built at upload, or 0 (default): 4 lanes when all rays of a call fit in one resident round
 * of the device, else 1. Results are the same up to exactly equal hits (ties). */
int mbvh_ctx_set_ray_lanes(mbvh_ctx* ctx, int32_t lanes);
/* closestPoint traversal width: 1 lane per query on the binary tree, 4 lanes per query on the
 * 4-wide tree, or 64: packets of 64 queries (consecutive in the spatial visit order) walking the
 * 4-wide tree together, one lane each, node records and triangles read once per packet (faster for
 * coherent queries such as grid samples, slower for scattered ones; trees too deep for its stack
 * use 4); 0 (default) = 4. Widths 4 and 64 give the same results bit for bit. Calls with a
 * nonzero minThreshold always use 1 lane: the reference returns the first triangle under
 * minThreshold in its own (binary) visit order. */
int mbvh_ctx_set_point_lanes(mbvh_ctx* ctx, int32_t lanes);
/* Waits for all work the context enqueued on its own stream. */
int mbvh_ctx_synchronize(mbvh_ctx* ctx);

/* ------------------------------------------------------------------------------------------ */
/* Device build                                                                               */
/* ------------------------------------------------------------------------------------------ */

/*
 * mbvh_build on the device: the same arguments, the same tree handle and byte-for-byte the same
 * node buffers, permuted index (in place, host memory) or indirect buffer. Replaces the same
 * reference functions as mbvh_build (buildPackedTree src/core/build/buildTree.js:123-152 with
 * computePrimitiveBounds src/core/MeshBVH.js:337-436, getBounds, getOptimalSplit, partition and
 * populateBuffer). Inputs are copied to the device, the result back; the call returns when done.
 * Limit: 2 x triangles + roots below 2^32. Errors as mbvh_build; a leaf over 65535 triangles is
 * reported with the largest such count.
 */
int mbvh_build_gpu(mbvh_ctx* ctx, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                   uint32_t position_offset, void* index, int32_t index_elt_bytes, uint64_t index_count,
                   const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                   mbvh_tree** out_tree);
/*
 * Device-pointer form: positions and index in device memory (the index is permuted in place on the
 * device); root_ranges and opts stay on the host. The trees are written one after another to
 * nodes_out (capacity at least (2 x triangles + roots) x 32 bytes), their node counts to the host
 * array tree_nodes; an indirect build writes its indirect buffer (2-byte entries when the last
 * range ends at or below 65536, else 4) to indirect_out. Waits for the build before returning.
 */
int mbvh_build_gpu_d(mbvh_ctx* ctx, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                     uint32_t position_offset, void* index, int32_t index_elt_bytes, uint64_t index_count,
                     const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                     void* nodes_out, uint64_t nodes_capacity, uint64_t* tree_nodes, uint32_t n_tree_slots,
                     void* indirect_out, uint64_t indirect_capacity, void* stream);

/*
 * Uploads one MeshBVH (its serialized form, MeshBVH.serialize src/core/MeshBVH.js:112-145:
 * roots, index, indirectBuffer) plus its float32 positions to HBM, once. The library derives its
 * traversal layout from those buffers on upload; later queries read only device memory.
 * index may be NULL only when an indirect buffer is given (non-indexed indirect build).
 */
int mbvh_mesh_upload(mbvh_ctx* ctx, const void* const* roots, const uint64_t* root_bytes,
                     uint32_t n_roots, const void* index, int32_t index_elt_bytes,
                     uint64_t index_count, const void* indirect, int32_t indirect_elt_bytes,
                     uint64_t indirect_count, const float* positions, uint64_t vertex_count,
                     uint32_t position_stride, uint32_t position_offset, mbvh_mesh** out_mesh);
void mbvh_mesh_free(mbvh_mesh* mesh);
/* Triangles (BVH slots) and vertices held by a mesh, and the deepest root's node depth. */
int mbvh_mesh_info(const mbvh_mesh* mesh, uint64_t* tri_count, uint64_t* vertex_count,
                   uint32_t* max_depth);

/*
 * Per-triangle facing mask for material arrays with geometry groups (intersectTri's group loop,
 * src/utils/ThreeRayIntersectUtilities.js:171-202). mask[f] for geometry triangle f (the
 * reference faceIndex): bit 0 = front-facing hits allowed, bit 1 = back-facing hits allowed.
 * Used by queries called with side = MBVH_SIDE_PER_TRIANGLE.
 */
int mbvh_mesh_set_side_mask(mbvh_ctx* ctx, mbvh_mesh* mesh, const uint8_t* mask, uint64_t n);

/*
 * Refit (MeshBVH.refit with nodeIndices = null, src/core/MeshBVH.js:504-508 over
 * src/core/cast/refit.template.js:4-196): recomputes the node boxes of the uploaded tree from new
 * positions of the same vertices (same count and layout as at upload), on the device and with the
 * reference's rules - a leaf box is the min / max of its triangles' vertices and is rewritten only
 * if it differs; an internal node is the union of its children, recomputed only if a child changed.
 * The device copies of the triangles and of every node layout are updated, so later queries see the
 * new geometry. If roots_out is not NULL, it has n_roots entries (the mesh's root count) and
 * roots_out[r] (root_bytes[r] bytes, the size of root r at upload; NULL entries skipped) receives
 * the refit reference-format node buffer of root r.
 * Synchronous. Queries on this mesh must not run concurrently with a refit.
 */
int mbvh_mesh_refit(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                    uint32_t position_stride, uint32_t position_offset, void* const* roots_out,
                    const uint64_t* root_bytes, uint32_t n_roots);
/* Device form of mbvh_mesh_refit: positions in device memory (for example updated by an optimiser
 * from the vertex gradients), enqueued on `stream`, no copy back, no synchronisation, no allocation.
 * Queries launched after it on the same stream see the refit geometry. */
int mbvh_mesh_refit_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                      uint32_t position_stride, uint32_t position_offset, void* stream);
/*
 * Partial refit (MeshBVH.refit(nodeIndices), src/core/MeshBVH.js:504-508 over
 * src/core/cast/refit.template.js:4-196): node_ids holds n_ids node indices counted over all roots
 * (root r's first node is the node count of the roots before it; the numbers shapecast reports,
 * src/core/BVH.js:394-409). The refit visits the nodes the reference visits for that set: every
 * root; below a visited node, the children in the set, or both (and then their whole subtrees) when
 * neither is. Nodes not visited keep their boxes and count as unchanged. Ids that name no node
 * match nothing; repeated ids are harmless. node_ids NULL is a full refit (nodeIndices = null).
 * The triangles' device copies take the new positions in any case. Otherwise as mbvh_mesh_refit;
 * the _d form takes device positions and device ids, as mbvh_mesh_refit_d.
 */
int mbvh_mesh_refit_nodes(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                          uint32_t position_stride, uint32_t position_offset, const uint32_t* node_ids, uint64_t n_ids,
                          void* const* roots_out, const uint64_t* root_bytes, uint32_t n_roots);
int mbvh_mesh_refit_nodes_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                            uint32_t position_stride, uint32_t position_offset, const uint32_t* node_ids, uint64_t n_ids,
                            void* stream);
/* Copies the mesh's current reference-format node buffers (as refit) into roots_out[r]
 * (root_bytes[r] bytes each, n_roots entries; NULL entries skipped). Synchronous. */
int mbvh_mesh_copy_roots(mbvh_ctx* ctx, const mbvh_mesh* mesh, void* const* roots_out, const uint64_t* root_bytes,
                         uint32_t n_roots);

/* ------------------------------------------------------------------------------------------ */
/* Queries                                                                                    */
/* ------------------------------------------------------------------------------------------ */

/*
 * First hit per ray. Replaces MeshBVH.raycastFirst (src/core/MeshBVH.js:558-577) →
 * raycastFirst (src/core/cast/raycastFirst.template.js:9-103) → intersectClosestTri
 * (src/core/utils/iterationUtils.template.js:24-53) → intersectTri / checkIntersection
 * (src/utils/ThreeRayIntersectUtilities.js:21-47,154-230), with three's Ray.intersectTriangle.
 * Results per ray: face index (reference faceIndex), distance |point - origin|, point
 * (origin + t*dir) and barycoord (weights of face a, b, c as Triangle.getBarycoord). Any of
 * distance / point / barycoord may be NULL.
 */
int mbvh_raycast_first(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                       const float* directions, uint64_t n, int32_t side, float near_dist,
                       float far_dist, uint32_t* face_index, float* distance, float* point,
                       float* barycoord);
/* `order` (device, may be NULL): a permutation of 0..n-1 giving the order in which the rays are
 * traced, e.g. from mbvh_order_points_d on the origins. Results still go to each ray's own index.
 * Entries >= n are skipped. */
int mbvh_raycast_first_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                         const float* directions, uint64_t n, const uint32_t* order, int32_t side, float near_dist,
                         float far_dist, uint32_t* face_index, float* distance, float* point,
                         float* barycoord, void* stream);
/*
 * Double-precision rays, as three.js holds them (its Ray and the local ray raycastObject3D builds,
 * src/core/MeshBVH.js:456-468, are JS numbers), traced as the reference traces them: every decision
 * (box, triangle, near / far, which of two equal hits) is the one for the double ray, not for its
 * float rounding. Each double x is split as hi = fl(x), lo = fl(x - hi); the kernels trace hi with
 * error bounds widened by lo and decide in double-float with both (all GPU arithmetic stays FP32).
 * Results as mbvh_raycast_first (float outputs). Magnitudes beyond the float range count as
 * infinite, and below 2^-149 as zero.
 */
int mbvh_raycast_first_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* origins, const double* directions,
                           uint64_t n, int32_t side, double near_dist, double far_dist, uint32_t* face_index,
                           float* distance, float* point, float* barycoord);
/* Device form: the rays already split (origins / directions the hi parts, *_lo the lo parts, both
 * device arrays, lo = fl(x - hi)); `order` as mbvh_raycast_first_d. */
int mbvh_raycast_first_ff_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins, const float* origins_lo,
                            const float* directions, const float* directions_lo, uint64_t n, const uint32_t* order,
                            int32_t side, double near_dist, double far_dist, uint32_t* face_index, float* distance,
                            float* point, float* barycoord, void* stream);

/*
 * intersectsSphere (MeshBVH.intersectsSphere src/core/MeshBVH.js:800-809 with
 * sphereIntersectTriangle src/math/MathUtilities.js:160-204) for n spheres: out[i] = 1 when some
 * triangle lies within radii[i] of centers[i] (inclusive), else 0. Decided from the closest
 * triangle's squared distance in double-float against radius^2. Synchronous.
 */
int mbvh_intersects_sphere(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* centers, const float* radii,
                           uint64_t n, uint8_t* out);
/*
 * intersectsBox (src/core/MeshBVH.js:780-792): whether any triangle intersects the oriented box
 * (box_min, box_max: 3 doubles per box in the box's frame; box_to_mesh: 16 doubles per box,
 * column-major, the box frame -> the mesh frame). The separating-axis tests of OrientedBox
 * (src/math/OrientedBox.js:79-260) and ExtendedTriangle (src/math/ExtendedTriangle.js:47-75) are
 * evaluated in double in the reference's order, so the answers are the reference's. out: 1 / 0 per
 * box. Host-pointer form (synchronous) and device-pointer form (enqueues on `stream`; out 2 marks a
 * traversal stack overflow).
 */
int mbvh_intersects_box(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* box_min, const double* box_max,
                        const double* box_to_mesh, uint64_t n, uint8_t* out);
int mbvh_intersects_box_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* box_min, const double* box_max,
                          const double* box_to_mesh, uint64_t n, uint8_t* out, void* stream);

/*
 * closestPointToGeometry (MeshBVH.closestPointToGeometry src/core/MeshBVH.js:840-853,
 * src/core/cast/closestPointToGeometry.template.js:15-271 with ExtendedTriangle.distanceToTriangle
 * src/math/ExtendedTriangle.js:774-860): the closest pair of triangles between the mesh and another
 * geometry (float positions with `stride` floats between vertices and x at `offset`; an optional
 * 2- or 4-byte index, NULL for a triangle soup; n_tris triangles) placed in the mesh frame by
 * geometry_to_bvh (16 doubles, column-major, three's applyMatrix4). Distances in double. The answer
 * is the closest pair overall (equal distances: the lowest other triangle, then the lowest mesh
 * slot); a pair counts only when its distance is below max_threshold. min_threshold is accepted for
 * the reference's signature: the reference stops at the first pair under it, and the overall
 * closest pair is under it whenever that one is. `face_index` is the mesh triangle slot (the
 * reference's closestDistanceTriIndex), `other_point` / `other_distance` are target2's point and
 * distance in the other geometry's frame. No pair: distance +inf, faces 0xFFFFFFFF.
 * Host-pointer form (synchronous) and device-pointer form (positions, index and `out` on the device;
 * enqueues on `stream`; out->status 1 marks a traversal stack overflow, 2 an index entry beyond
 * n_vertices). The context's workspace is shared: device-pointer calls on one context must not overlap.
 */
typedef struct mbvh_geometry_hit {
  double distance;
  double point[3];        /* on the mesh triangle, mesh frame (target1.point) */
  double other_point[3];  /* on the other triangle, the other geometry's frame (target2.point) */
  double other_distance;  /* target2.distance */
  uint32_t face_index;    /* target1.faceIndex */
  uint32_t other_face_index; /* target2.faceIndex */
  int32_t status;
  uint32_t pad;
} mbvh_geometry_hit;
int mbvh_closest_point_to_geometry(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions, uint64_t n_vertices,
                                   uint32_t stride, uint32_t offset, const void* index, uint32_t index_elt,
                                   uint64_t n_tris, const double* geometry_to_bvh, double min_threshold,
                                   double max_threshold, mbvh_geometry_hit* out);
int mbvh_closest_point_to_geometry_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions,
                                     uint64_t n_vertices, uint32_t stride, uint32_t offset, const void* index,
                                     uint32_t index_elt, uint64_t n_tris, const double* geometry_to_bvh,
                                     double min_threshold, double max_threshold, mbvh_geometry_hit* out, void* stream);

/*
 * All hits per ray. Replaces MeshBVH.raycast (src/core/MeshBVH.js:532-545) → raycast
 * (src/core/cast/raycast.template.js:7-53) → intersectTris (src/core/utils/iterationUtils.template.js:4-22).
 * The hits of ray i are entries offsets[i] .. offsets[i+1]-1, in the reference's traversal order
 * (unsorted, as the reference returns them; three's Raycaster sorts by distance afterwards).
 * Host form: mbvh_raycast writes offsets (n + 1 entries, offsets[n] = total hits) and keeps the
 * hits on the device until the next mbvh_raycast on this context; mbvh_raycast_fetch copies them.
 */
int mbvh_raycast(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins, const float* directions,
                 uint64_t n, int32_t side, float near_dist, float far_dist, uint64_t* offsets);
int mbvh_raycast_fetch(mbvh_ctx* ctx, uint32_t* face_index, float* distance, float* point,
                       float* barycoord, uint64_t total);
/* mbvh_raycast for double-precision rays (as mbvh_raycast_first_f64); fetch with mbvh_raycast_fetch. */
int mbvh_raycast_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* origins, const double* directions,
                     uint64_t n, int32_t side, double near_dist, double far_dist, uint64_t* offsets);
/* Device form in two passes: hits per ray, then the fill at caller-computed offsets (the exclusive
 * prefix sum of the counts); `capacity` bounds the output arrays. */
int mbvh_raycast_count_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                         const float* directions, uint64_t n, int32_t side, float near_dist,
                         float far_dist, uint32_t* counts, void* stream);
int mbvh_raycast_fill_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                        const float* directions, uint64_t n, int32_t side, float near_dist,
                        float far_dist, const uint64_t* offsets, uint64_t capacity,
                        uint32_t* face_index, float* distance, float* point, float* barycoord,
                        void* stream);

/*
 * Closest point on the mesh per query point. Replaces MeshBVH.closestPointToPoint
 * (src/core/MeshBVH.js:874-884) → closestPointToPoint (src/core/cast/closestPointToPoint.js:6-78)
 * → shapecast (src/core/BVH.js:342-413, src/core/cast/shapecast.js:12-214) with three's
 * Triangle.closestPointToPoint. Results: face index, distance, closest point and the barycoord of
 * that point in its face (getTriangleHitPointInfo, src/utils/TriangleUtilities.js:65-147).
 * Nothing found → MBVH_NO_HIT / +Infinity. distance / point / barycoord may be NULL.
 */
int mbvh_closest_point(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                       float min_threshold, float max_threshold, uint32_t* face_index,
                       float* distance, float* closest, float* barycoord);
/* `order` (device, may be NULL): a permutation of 0..n-1 giving the order in which the points are
 * processed, e.g. from mbvh_order_points_d. Results still go to each point's own index. Entries
 * >= n are skipped. */
int mbvh_closest_point_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                         const uint32_t* order, float min_threshold, float max_threshold, uint32_t* face_index,
                         float* distance, float* closest, float* barycoord, void* stream);
/* Double-precision query points and thresholds (three's Vector3 are JS numbers), decided as the
 * reference decides them for the double point (split and traced as mbvh_raycast_first_f64). */
int mbvh_closest_point_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* points, uint64_t n,
                           double min_threshold, double max_threshold, uint32_t* face_index, float* distance,
                           float* closest, float* barycoord);
/* Device form: points split as hi (points) + lo (points_lo) device arrays. */
int mbvh_closest_point_ff_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, const float* points_lo,
                            uint64_t n, const uint32_t* order, double min_threshold, double max_threshold,
                            uint32_t* face_index, float* distance, float* closest, float* barycoord, void* stream);

/* Size of the per-device scratch the allocation-free entry point below uses, and its query limit. */
#define MBVH_NOALLOC_SCRATCH_BYTES (32ull << 20)
#define MBVH_NOALLOC_MAX_QUERIES (1ull << 21)
/*
 * Allocation-free closest point. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as mbvh_closest_point_d (device pointers, enqueued on
 * `stream`, NULL = the context's stream), except that the query order is computed inside (the host
 * entry point's spatial order, from 4,096 queries on) and that the work-queue counter, the order,
 * its workspace and the traversal stack overflow live in a fixed MBVH_NOALLOC_SCRATCH_BYTES region
 * of device memory that the library reserves once per GPU for the life of the process. The call
 * makes no device allocation and no library call that allocates.
 *   - n > MBVH_NOALLOC_MAX_QUERIES is rejected with MBVH_ERR_LIMIT (nothing is enqueued).
 *   - A tree too deep for the stack region even at one workgroup is rejected with MBVH_ERR_LIMIT;
 *     a deep tree otherwise runs on fewer workers (same results, lower throughput).
 *   - Every call on one GPU shares that GPU's region: calls that overlap in time on one GPU (two
 *     streams, two contexts, two threads) race and give undefined results. Serialise them.
 */
int mbvh_closest_point_noalloc_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                 float min_threshold, float max_threshold, uint32_t* face_index,
                                 float* distance, float* closest, float* barycoord, void* stream);

/* ------------------------------------------------------------------------------------------ */
/* Instances and the batched multi-mesh launch (BatchedMesh)                                  */
/* ------------------------------------------------------------------------------------------ */
/*
 * acceleratedRaycast on a BatchedMesh (acceleratedBatchedMeshRaycast,
 * src/utils/ExtensionUtilities.js:87-160) raycasts each visible instance through
 * MeshBVH.raycastObject3D (src/core/MeshBVH.js:447-494): the world ray goes to the instance's
 * frame with matrixWorld^-1 (three's Ray.applyMatrix4: origin by Vector3.applyMatrix4,
 * direction by transformDirection, normalised), near / far are divided by
 * |local direction * world scale|, raycastFirst runs on the instance's geometry BVH, and the hit
 * comes back with matrixWorld (convertRaycastIntersect, src/utils/GeometryRayIntersectUtilities.js:3-17:
 * world point, distance from the world origin). Here each ray names its instance
 * (ray_instance[i]); instance k uses mesh instance_geometry[k] of the batch and the column-major
 * 4x4 matrices matrix_world + 16k and matrix_world_inverse + 16k (three's Matrix4.elements, the
 * inverse computed by the caller as raycastObject3D does). Results per ray: face index in that
 * mesh, world distance, world point, barycoord. A ray whose instance or mesh index is out of
 * range is a miss.
 *
 * mbvh_batch_raycast_first_d is ONE persistent launch over all meshes: rays are grouped by mesh
 * (counting sort), then a fixed grid pulls tiles of rays from an atomic work counter until none
 * remain. It allocates nothing and uses no memset; the device workspace holds the order.
 * Concurrent launches of one batch race on its work counter: serialise them (one stream).
 * tuning may be NULL (the batch's autotuned values, or defaults before mbvh_batch_autotune).
 */
typedef struct mbvh_batch_tuning {
  uint32_t ray_lanes;     /* lanes per ray: 1 (binary trees) or 4 (4-wide trees) */
  uint32_t stack_lds;     /* LDS traversal-stack entries per lane (1 lane per ray): 8 or 16 */
  uint32_t tile_rays;     /* rays per work item: a multiple of 64, at most 4096 */
  uint32_t blocks_per_cu; /* persistent workgroups of 256 lanes per compute unit: 1..8 */
  float ms;               /* mbvh_batch_autotune: measured time of the chosen configuration */
} mbvh_batch_tuning;
typedef struct mbvh_batch_graph mbvh_batch_graph;

/* A batch references meshes uploaded on ctx; they must outlive it. */
int mbvh_batch_create(mbvh_ctx* ctx, const mbvh_mesh* const* meshes, uint32_t n_meshes, mbvh_batch** out_batch);
void mbvh_batch_free(mbvh_batch* batch);
int mbvh_batch_workspace_bytes(const mbvh_batch* batch, uint64_t n_rays, uint64_t* bytes);
int mbvh_batch_raycast_first_d(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                               const float* matrix_world_inverse, const uint32_t* instance_geometry,
                               uint32_t n_instances, const float* origins, const float* directions,
                               const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                               float far_dist, uint32_t* face_index, float* distance, float* point,
                               float* barycoord, void* workspace, uint64_t workspace_bytes,
                               const mbvh_batch_tuning* tuning, void* stream);
/* Host-pointer form of mbvh_batch_raycast_first_d (the batch's tuning, the context's stream and
 * scratch): stages the inputs, launches, copies the results back and synchronises. */
int mbvh_batch_raycast_first(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                             const float* matrix_world_inverse, const uint32_t* instance_geometry,
                             uint32_t n_instances, const float* origins, const float* directions,
                             const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist, float far_dist,
                             uint32_t* face_index, float* distance, float* point, float* barycoord);
/*
 * The batched launch on rays already brought into their mesh's frame in double precision: what
 * acceleratedBatchedMeshRaycast (src/utils/ExtensionUtilities.js:87-161) hands each instance's
 * raycastObject3D (src/core/MeshBVH.js:456-468: the world ray through matrixWorld^-1 and near / far
 * divided by the scale, all JS numbers). One ray per (ray, instance) pair: ray_mesh[i] names the batch
 * mesh, windows[2i], windows[2i + 1] its near / far in local units (windows NULL: 0 and +Infinity).
 * Traced as mbvh_raycast_first_f64 (every decision for the double ray) in one launch; results stay
 * in the mesh's frame (the caller converts them as convertRaycastIntersect does). A ray_mesh value
 * outside the batch misses. Host pointers, synchronous.
 */
int mbvh_batch_raycast_first_local_f64(mbvh_ctx* ctx, mbvh_batch* batch, const double* origins,
                                       const double* directions, const double* windows, const uint32_t* ray_mesh,
                                       uint64_t n, int32_t side, uint32_t* face_index, float* distance, float* point,
                                       float* barycoord);
/* ------------------------------------------------------------------------------------------ */
/* Export to the reference's shader data layouts                                              */
/* ------------------------------------------------------------------------------------------ */

/*
 * Texture sides (width = height) of MeshBVHUniformStruct's textures for this mesh, in texels:
 * sides[0] bvhBounds (2 * ceil(sqrt(nodes / 2))), sides[1] bvhContents (ceil(sqrt(nodes))),
 * sides[2] the index texture (ceil(sqrt(triangles)), at least 1), sides[3] the position texture
 * (ceil(sqrt(vertices)), at least 1); src/webgl/MeshBVHUniformStruct.js:146-150,
 * src/webgl/VertexAttributeTexture.js:251.
 */
int mbvh_mesh_texture_sizes(const mbvh_mesh* mesh, uint32_t* sides);
/*
 * MeshBVHUniformStruct.updateFrom's texture data (bvhToTextures src/webgl/MeshBVHUniformStruct.js:129-210
 * and VertexAttributeTexture.updateFrom src/webgl/VertexAttributeTexture.js:106-300) written from the
 * device copy of a single-root mesh (multi-root meshes are rejected, as the reference does):
 * bounds = RGBA32F texels, two per node (min.xyz, 0 / max.xyz, 0); contents = RG32UI texels
 * ((0xFFFF0000 | count, offset) for a leaf, (split axis, right child distance) otherwise); index
 * (optional) = RGBA texels (a, b, c, 1) of the permuted or, for an indirect mesh, dereferenced index, as
 * 32-bit values (the reference keeps the index array's element type); position (written when
 * positions is given) = RGBA32F texels (x, y, z, 1). Texels past the data are zero. Sizes from
 * mbvh_mesh_texture_sizes. Device pointers; enqueues on `stream`.
 */
int mbvh_mesh_export_textures_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint32_t position_stride,
                                uint32_t position_offset, float* bounds, uint32_t* contents, uint32_t* index,
                                float* position, void* stream);
/* Host-pointer form (synchronous). */
int mbvh_mesh_export_textures(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint32_t position_stride,
                              uint32_t position_offset, float* bounds, uint32_t* contents, uint32_t* index,
                              float* position);
/*
 * The WebGPU BVHNode storage layout of one mesh (bvhNodeStruct src/webgpu/tsl/structs.js:3-14, as
 * appendBVHSubtree src/webgpu/utils/packBVHBufferUtils.js writes a mesh's roots): 8 words per node,
 * roots one after another; a root whose bounds are empty (min > max) gets [1, -1] on that axis.
 * The node buffers are the device copy of the mesh (refits included). Device pointer form enqueues
 * on `stream`; the host form is synchronous.
 */
int mbvh_mesh_export_nodes_d(mbvh_ctx* ctx, mbvh_mesh* mesh, uint32_t* nodes, uint64_t capacity_bytes, void* stream);
int mbvh_mesh_export_nodes(mbvh_ctx* ctx, mbvh_mesh* mesh, uint32_t* nodes, uint64_t capacity_bytes);

/*
 * Imports into a device mesh (no counterpart in the reference, which only packs; the inverses of the
 * two packings above). mbvh_mesh_import_nodes reads a WebGPU BVHNode buffer of n_nodes 8-word
 * records holding a mesh's root subtrees back to back (appendBVHSubtree,
 * src/webgpu/utils/packBVHBufferUtils.js:216-280): each root's extent is its subtree's span, leaf
 * triangle offsets lose `geometry_offset`, and a root axis written [1, -1] for empty bounds becomes
 * [+inf, -inf] again. mbvh_mesh_import_textures reads MeshBVHUniformStruct's bvhBounds (RGBA32F,
 * node i's min at texel 2i, max at 2i + 1) and bvhContents (RG32UI: a leaf (0xFFFF0000 | count,
 * offset), an internal node (split axis, right child offset)) of one root
 * (src/webgl/MeshBVHUniformStruct.js:129-210). The index / indirect / position arguments are those
 * of mbvh_mesh_upload. Host arrays; the conversion runs on the host, then the mesh is uploaded.
 */
int mbvh_mesh_import_nodes(mbvh_ctx* ctx, const uint32_t* nodes, uint64_t n_nodes, uint32_t geometry_offset,
                           const void* index, int32_t index_elt_bytes, uint64_t index_count, const void* indirect,
                           int32_t indirect_elt_bytes, uint64_t indirect_count, const float* positions,
                           uint64_t vertex_count, uint32_t position_stride, uint32_t position_offset,
                           mbvh_mesh** out_mesh);
int mbvh_mesh_import_textures(mbvh_ctx* ctx, const float* bvh_bounds, const uint32_t* bvh_contents, uint64_t n_nodes,
                              const void* index, int32_t index_elt_bytes, uint64_t index_count, const void* indirect,
                              int32_t indirect_elt_bytes, uint64_t indirect_count, const float* positions,
                              uint64_t vertex_count, uint32_t position_stride, uint32_t position_offset,
                              mbvh_mesh** out_mesh);

/* ------------------------------------------------------------------------------------------ */
/* Scenes: closest hit over all instances in one launch                                      */
/* ------------------------------------------------------------------------------------------ */

/*
 * A scene of instances (mesh meshes[instance_geometry[i]] placed by matrix_world[i], 16 floats,
 * column-major as three's Matrix4.elements, with its inverse) under a top-level tree over the
 * instances' world boxes, built at creation (the ClusteredBVH top level, src/webgpu/ClusteredBVH.js:18-325,
 * with one primitive per instance; its traversal src/webgpu/shapecastFns/getShapecastFn.js:97-217).
 * visible (NULL: all) excludes instances as getVisibleAt does (src/utils/ExtensionUtilities.js:112-116).
 * The world boxes come from the meshes' current root boxes: re-create the scene after a refit or
 * after instances move. The meshes must outlive the scene.
 * Every matrix_world must be affine (last row 0 0 0 1, as Matrix4.compose makes it): others are
 * rejected with MBVH_ERR_BAD_ARGS. matrix_world_inverse is used as given to bring rays into each
 * instance's frame (as raycastObject3D inverts matrixWorld itself); the closest point's pruning
 * bounds never rely on it: the scene computes its own inverse in double and uses world boxes for an
 * instance whose inverse fails a residual check.
 */
int mbvh_scene_create(mbvh_ctx* ctx, const mbvh_mesh* const* meshes, uint32_t n_meshes, const float* matrix_world,
                      const float* matrix_world_inverse, const uint32_t* instance_geometry, const uint8_t* visible,
                      uint32_t n_instances, mbvh_scene** out_scene);
void mbvh_scene_free(mbvh_scene* scene);
/*
 * Closest first hit of each world ray over every visible instance: the result the reference's
 * acceleratedBatchedMeshRaycast (src/utils/ExtensionUtilities.js:87-161) with firstHitOnly puts first
 * after three's Raycaster sorts the per-instance hits by distance (equal distances: the lower instance
 * index). face_index is the face in the hit instance's mesh (MBVH_NO_HIT on a miss), instance the
 * instance index; distance and point in world space; barycoord as raycastFirst. near / far in world
 * units (scaled into each instance's frame as raycastObject3D does, src/core/MeshBVH.js:447-494).
 * Device pointers; instance, distance, point, barycoord may be NULL. Enqueues on `stream`.
 */
int mbvh_scene_raycast_first_d(mbvh_ctx* ctx, const mbvh_scene* scene, const float* origins, const float* directions,
                               uint64_t n, int32_t side, float near_dist, float far_dist, uint32_t* face_index,
                               uint32_t* instance, float* distance, float* point, float* barycoord, void* stream);
/*
 * Closest point over the instances of a scene: the reference's bvh_ClosestPointToPoint over a
 * ClusteredBVH (src/webgpu/shapecastFns/getClosestPointToPointFn.js:20-151, traversed as in
 * src/webgpu/shapecastFns/getShapecastFn.js:97-217) — node boxes and triangles taken to world space
 * by each instance's matrix, the least world distance wins — for n world points (3n floats). Triangles
 * and distances in double; equal distances go to the lowest instance, then the lowest face. Outputs
 * per point: face_index (in the instance's geometry), instance, world distance, world point and
 * barycoord (NULL outputs other than face_index are skipped). Host-pointer form (synchronous; it
 * orders the points spatially itself) and device-pointer form (enqueues on `stream`; `order`: NULL or
 * a visit order of the n points such as mbvh_order_points_d writes, which lets the lanes of a wave
 * share instances and tree paths; face 0xFFFFFFFE marks a traversal stack overflow). Points whose
 * walks are long finish in a pass of one wave each, through a list the context allocates once: in
 * calls of up to 2^20 points.
 */
int mbvh_scene_closest_point(mbvh_ctx* ctx, const mbvh_scene* scene, const float* points, uint64_t n,
                             uint32_t* face_index, uint32_t* instance, float* distance, float* point,
                             float* barycoord);
int mbvh_scene_closest_point_d(mbvh_ctx* ctx, const mbvh_scene* scene, const float* points, uint64_t n,
                               const uint32_t* order, uint32_t* face_index, uint32_t* instance, float* distance,
                               float* point, float* barycoord, void* stream);
/*
 * The scene in the reference's WebGPU packing (BVHComputeData.update, src/webgpu/BVHComputeData.js:217-447,
 * over ClusteredBVH's primitives, src/webgpu/ClusteredBVH.js:207-331). `primitive_limit` 0 references
 * every instance's roots whole (what ClusteredBVH does for instanced and batched objects and for
 * scenes of fewer than three meshes); otherwise each root is cut into clusters as _traverseClusters
 * does (depth first, left first, the first node on each path that is a leaf or spans fewer than
 * `primitive_limit` triangles; the reference uses 64). `nodes` (8 words per BVHNode,
 * src/webgpu/tsl/structs.js:9-14): a top level with one (instance, root, cluster) primitive per leaf,
 * each leaf holding its subtree's absolute node offset and 0xFF000000 | its transform slot
 * (appendBVHData, src/webgpu/utils/packBVHBufferUtils.js:44-131), then each referenced
 * (mesh, root, cluster) subtree once, its contiguous node span (getSubtreeNodeCount, :142-161) with
 * leaf offsets rebased by the mesh's triangle base (appendBVHSubtree, :216-280); `transforms` (36
 * words per slot: matrixWorld, its inverse, visible, 3 padding words; writeTransformData,
 * BVHComputeData.js:491-526); `index` (3 words per triangle: each mesh's triangles in slot order,
 * rebased by the mesh's vertex base, meshes in the order the top level first names them; an
 * attribute buffer packed in that mesh order, each mesh's vertices whole, matches it). Instances whose
 * roots hold no triangles are left out; invisible instances stay, with visible 0. The top level is
 * this library's SAH build over the primitives' world boxes. Fails with MBVH_ERR_LIMIT when two
 * primitives with identical boxes would share a leaf. Host arrays; sizes in words from
 * mbvh_scene_webgpu_sizes (same `primitive_limit`); NULL outputs are skipped.
 */
int mbvh_scene_webgpu_sizes(mbvh_ctx* ctx, const mbvh_scene* scene, uint32_t primitive_limit, uint64_t* n_nodes,
                            uint64_t* n_transforms, uint64_t* n_index);
int mbvh_scene_export_webgpu(mbvh_ctx* ctx, const mbvh_scene* scene, uint32_t primitive_limit, uint32_t* nodes,
                             uint64_t nodes_capacity, uint32_t* transforms, uint64_t transforms_capacity,
                             uint32_t* index, uint64_t index_capacity);
/* Host-pointer form: stages the inputs, launches, copies the results back and synchronises. */
int mbvh_scene_raycast_first(mbvh_ctx* ctx, const mbvh_scene* scene, const float* origins, const float* directions,
                             uint64_t n, int32_t side, float near_dist, float far_dist, uint32_t* face_index,
                             uint32_t* instance, float* distance, float* point, float* barycoord);

/* Times every candidate (ray_lanes x stack_lds x tile_rays x blocks_per_cu) on these inputs, keeps the fastest
 * in the batch (used when tuning is NULL) and returns it. Synchronous; outputs hold the last run. */
int mbvh_batch_autotune(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                        const float* matrix_world_inverse, const uint32_t* instance_geometry,
                        uint32_t n_instances, const float* origins, const float* directions,
                        const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist, float far_dist,
                        uint32_t* face_index, float* distance, float* point, float* barycoord, void* workspace,
                        uint64_t workspace_bytes, mbvh_batch_tuning* out_tuning);
/* Captures the whole batched launch (ordering + traversal) into a HIP graph for these buffers;
 * mbvh_batch_graph_launch replays it on `stream` (NULL: the context's stream). The graph holds the
 * addresses of these buffers, of the batch and of the context's traversal stack region: free the
 * graph before the batch, its meshes or the buffers. When a later upload, batch or scene makes the
 * context grow its stack region (a deeper tree), the graph is stale and mbvh_batch_graph_launch
 * returns MBVH_ERR_BAD_HANDLE without launching; capture it again. */
int mbvh_batch_graph_create(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                            const float* matrix_world_inverse, const uint32_t* instance_geometry,
                            uint32_t n_instances, const float* origins, const float* directions,
                            const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                            float far_dist, uint32_t* face_index, float* distance, float* point,
                            float* barycoord, void* workspace, uint64_t workspace_bytes,
                            const mbvh_batch_tuning* tuning, mbvh_batch_graph** out_graph);
int mbvh_batch_graph_launch(mbvh_batch_graph* graph, void* stream);
void mbvh_batch_graph_free(mbvh_batch_graph* graph);
/* The two halves of raycastObject3D's frame change, for callers that loop over meshes
 * themselves: world rays -> instance-local rays, and local hits (face, point) -> world point and
 * distance in place (misses untouched). Same arithmetic as the batched launch, bit for bit. */
int mbvh_instance_rays_d(mbvh_ctx* ctx, const float* matrix_world_inverse, uint32_t n_instances,
                         const float* origins, const float* directions, const uint32_t* ray_instance, uint64_t n,
                         float* local_origins, float* local_directions, void* stream);
int mbvh_instance_hits_d(mbvh_ctx* ctx, const float* matrix_world, uint32_t n_instances, const float* origins,
                         const uint32_t* ray_instance, uint64_t n, const uint32_t* face_index, float* point,
                         float* distance, void* stream);

/* ------------------------------------------------------------------------------------------ */
/* Backward (adjoint) of the queries                                                          */
/* ------------------------------------------------------------------------------------------ */
/*
 * No counterpart in the reference (which has no derivatives). For a fixed forward result (the
 * saved face index, distance, barycoord and closest point of each query), the backward returns
 * the gradients of a scalar loss L with respect to the query inputs and to the mesh's vertex
 * positions, given the upstream gradients of L with respect to the forward outputs.
 * Vertex gradients: vertex_count x 3 floats (the vertices the geometry index refers to, packed
 * x, y, z whatever the position stride), overwritten. Misses contribute nothing; a closest point
 * at distance 0 has no gradient (zero).
 * mode MBVH_GRAD_ATOMIC sums vertex contributions with FP32 atomics (run-to-run rounding may
 * differ); MBVH_GRAD_DETERMINISTIC sorts them by vertex (stable radix sort) and sums each
 * vertex's contributions in (query, corner) order, bitwise reproducible. The deterministic mode
 * needs a device workspace of mbvh_grad_workspace_bytes() bytes (256-byte aligned).
 */
#define MBVH_GRAD_ATOMIC 0
#define MBVH_GRAD_DETERMINISTIC 1
int mbvh_grad_workspace_bytes(uint64_t n, int32_t mode, uint64_t* bytes);
/* Upstream: grad_distance (n), grad_point (3n), grad_barycoord (3n), each may be NULL (= 0).
 * Outputs: grad_origins (3n), grad_directions (3n), grad_positions (3 x vertex_count), each
 * may be NULL (not computed). */
int mbvh_raycast_first_bwd_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* directions, uint64_t n,
                             const uint32_t* face_index, const float* distance, const float* barycoord,
                             const float* grad_distance, const float* grad_point, const float* grad_barycoord,
                             float* grad_origins, float* grad_directions, float* grad_positions, int32_t mode,
                             void* workspace, uint64_t workspace_bytes, void* stream);
int mbvh_raycast_first_bwd(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* directions, uint64_t n,
                           const uint32_t* face_index, const float* distance, const float* barycoord,
                           const float* grad_distance, const float* grad_point, const float* grad_barycoord,
                           float* grad_origins, float* grad_directions, float* grad_positions, int32_t mode);
/* Upstream: grad_distance (n). Outputs: grad_points (3n), grad_positions (3 x vertex_count). */
int mbvh_closest_point_bwd_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                             const uint32_t* face_index, const float* distance, const float* closest,
                             const float* barycoord, const float* grad_distance, float* grad_points,
                             float* grad_positions, int32_t mode, void* workspace, uint64_t workspace_bytes,
                             void* stream);
int mbvh_closest_point_bwd(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                           const uint32_t* face_index, const float* distance, const float* closest,
                           const float* barycoord, const float* grad_distance, float* grad_points,
                           float* grad_positions, int32_t mode);

/*
 * Spatial query order (no counterpart in the reference, which answers one query per call). A wave
 * of 64 lanes costs the union of its lanes' traversal paths, so neighbouring queries should share
 * a wave: mbvh_order_points_d writes a permutation `order` (n entries) that groups the n packed
 * float triples by Morton cell of their bounding box. It needs a device workspace of
 * mbvh_order_workspace_bytes(n) bytes (256-byte aligned), allocates nothing and uses no memset,
 * so it may be captured in a HIP graph. The host query entry points order their queries this way
 * by themselves. n must be below 2^32.
 */
int mbvh_order_workspace_bytes(uint64_t n, uint64_t* bytes);
int mbvh_order_points_d(mbvh_ctx* ctx, const float* points, uint64_t n, uint32_t* order, void* workspace,
                        uint64_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MBVH_H */
