// Kernel argument blocks and host-side launchers (implemented in mbvh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "mbvh_device.h"

namespace mbvh {

struct MeshView {
  const WideNode* nodes;
  const uint32_t* roots;
  const uint2* root_slots;  // slot range (first, end) of each root (ties across roots: the earlier root)
  const float4* nodes4;   // 4-wide tree: 8 float4 per node (4 children x {min.xyz max.x}{max.yz ref 0})
  const uint32_t* roots4; // 4-wide root reference per root
  uint32_t max_stack4;    // stack entries a quad traversal can need (4-wide or binary)
  const uint2* bigleaf;
  const float4* tris;  // 3 per slot
  uint32_t n_roots;
  uint32_t max_stack;  // stack entries a traversal of this mesh can need
  const float* coord_max;  // device: an upper bound of |vertex coordinate| (scales FP32 error bounds;
                           // rewritten by a refit)
};

struct StackScratch {
  uint2* ovf;            // overflow entries beyond the LDS part
  uint64_t ovf_stride;   // threads in the launch (row pitch of the overflow region)
  uint32_t ovf_entries;  // overflow entries available per thread
};

struct RaycastFirstArgs {
  MeshView mesh;
  StackScratch stack;
  const float* origins;
  const float* directions;
  // double-precision rays as hi + lo pairs: the lo parts (both NULL for FP32 rays), lo = fl(x - hi)
  const float* origins_lo;
  const float* directions_lo;
  const uint32_t* order;  // visit order of the rays (NULL: 0..n-1); results go to each ray's own index
  uint64_t n;
  uint32_t side_mask;  // 1 front, 2 back, 3 both; 0 = per-triangle masks stored in the mesh
  float near_dist;
  float far_dist;
  float near_lo;  // lo parts of near / far (0 for FP32 bounds)
  float far_lo;
  uint32_t* out_face;
  float* out_dist;
  float* out_point;
  float* out_bary;
};

struct RaycastAllArgs {
  MeshView mesh;
  StackScratch stack;
  const float* origins;
  const float* directions;
  const float* origins_lo;     // lo parts (NULL for FP32 rays), as RaycastFirstArgs
  const float* directions_lo;
  uint64_t n;
  uint32_t side_mask;
  float near_dist;
  float far_dist;
  float near_lo;
  float far_lo;
  uint32_t* counts;          // count pass: hits per ray
  const uint64_t* offsets;   // fill pass: first output slot of each ray
  uint32_t* out_face;
  float* out_dist;
  float* out_point;
  float* out_bary;
  uint64_t capacity;         // fill pass: output slots available
};

struct ClosestPointArgs {
  MeshView mesh;
  StackScratch stack;
  const float* points;
  const float* points_lo;  // lo parts of double-precision points (NULL for FP32 points)
  const uint32_t* order;  // visit order of the points (NULL: 0..n-1); results go to each point's own index
  uint32_t* queue;        // work-queue counter (reset by the launch)
  uint64_t n;
  float min_threshold;
  float max_threshold;
  float min_lo;  // lo parts of the thresholds (0 for FP32 thresholds)
  float max_lo;
  uint32_t* out_face;
  float* out_dist;
  float* out_point;
  float* out_bary;
  // packet kernels only: the launch covers packets [tile0, tile0 + tiles) of 64 queries in visit
  // order. cand: kPacketCand candidates (slot, FP32 lower bound bits) per query of the launch;
  // cand_meta: per query (listed count or a mark, FP32 upper bound bits, head of its further
  // candidates in pool or kPoolEnd, 0); pool: further candidates (slot, lower bound bits, next, 0),
  // pool_count of pool_cap used; redo / redo_count: queries left when the pool ran out, finished by
  // the quad kernel
  uint64_t tile0;
  uint64_t tiles;
  uint2* cand;
  uint4* cand_meta;
  uint4* pool;
  uint32_t* pool_count;
  uint32_t pool_cap;
  uint32_t* redo;
  uint32_t* redo_count;
  // quad kernel: when set, the number of queries to visit (device-side, e.g. a redo list's length);
  // n stays the number of points (the bound of the indices in order)
  const uint32_t* n_dev;
};

// ---- refit ------------------------------------------------------------------------------------
struct RefitArgs {
  uint32_t* ref_nodes;        // the reference-format nodes of all roots, 8 words each
  const int32_t* wide_off;    // float offset of the node's box in the wide nodes (-1: none)
  const uint32_t* leaves;     // leaf node indices
  uint64_t n_nodes;
  uint64_t n_leaves;
  uint32_t* changed;          // per node: box recomputed
  const uint32_t* levels;     // internal nodes grouped by depth
  const uint64_t* level_off;  // host: levels[level_off[d] .. level_off[d + 1]) are at depth d
  uint32_t n_levels;
  float* wide;                // wide nodes as floats (16 per node)
  float* four;                // 4-wide nodes as floats
  const uint2* four_map;      // (float offset in the 4-wide tree, float offset in the wide nodes)
  uint64_t n_four;
  float4* tris;
  const uint32_t* vidx;
  uint64_t n_slots;
  const float* positions;     // new vertex positions (device)
  uint64_t n_vertices;
  uint32_t stride;
  uint32_t offset;
  float* coord_bound;         // the mesh's coordinate bound, recomputed from the positions
  // partial refit (nodeIndices): node_ids != nullptr restricts the refit to the nodes the reference's
  // traversal visits for that set; ids count nodes over all roots, ids >= n_nodes match nothing
  const uint32_t* node_ids;
  uint64_t n_ids;
  const uint64_t* root_first;  // first node of each root
  uint32_t n_roots;
};
hipError_t launch_refit(const RefitArgs& a, hipStream_t s);

// ---- device build -------------------------------------------------------------------------------
// Positions q in [0, span) of the primitive buffer being built; a non-indirect build's buffer is the
// triangle range [base, base + span) of the index, an indirect build's is the generated indirect
// buffer (base 0).
struct GpuBuildArgs {
  const float* positions;      // device vertex data; vertex v, axis k at positions[v * stride + offset + k]
  uint64_t n_vertices;
  uint32_t stride;
  uint32_t offset;
  const void* index;           // device geometry index (NULL: non-indexed, triangle t = vertices 3t..3t+2)
  int32_t index_elt;           // 2 or 4
  uint64_t index_count;
  const uint32_t* src;         // n_src (first triangle, first position) pairs: position q of source range s
  uint32_t n_src;              // holds triangle src[2s] + (q - src[2s + 1])
  const uint32_t* roots;       // n_roots (first position, count) pairs of the trees to build
  uint32_t n_roots;
  uint64_t base;               // added to positions in leaf offsets and index writes
  uint64_t span;
  int32_t strategy;
  int32_t max_depth;
  int32_t leaf_size;
  // workspace
  float* prim;                 // 6 planes of span floats: center x, half x, center y, half y, center z, half z
  uint32_t* perm;              // triangle at each position
  uint32_t* pair_l;            // partition scratch: misplaced left / right elements by rank
  uint32_t* pair_r;
  float* nbox;                 // 6 per node
  float* ncen;                 // 6 per node: box of the triangle centers
  uint4* ninfo;                // (first position, count, left child or leaf mark, split axis)
  uint32_t* nroot;             // tree of each node
  uint32_t* nsize;             // nodes in the subtree
  uint32_t* nslot;             // packed slot
  uint64_t max_nodes;
  uint32_t* lvl;               // 4 x (max_depth + 2): level start, level count, big-list start, big-list count
  uint32_t* big;               // nodes split by a whole workgroup
  unsigned long long* rkeys;   // 12 per root: bound keys of the root ranges
  // big nodes of the level being split (index i = position in the level's big-node list)
  uint32_t* wbins;             // SAH bins, 3 x 32 x 7 words per node
  int32_t* waxis;              // split axis (-1: made a leaf)
  double* wpos;                // split plane
  uint32_t* wbase;             // first chunk of each node (+ the level's chunk total)
  uint32_t* wcnt;              // per chunk: elements below the plane
  uint32_t* wpre;              // per chunk: elements below the plane in the node's earlier chunks
  uint32_t* wnl;               // elements below the plane
  uint32_t* wm;                // misplaced pairs
  unsigned long long* wkeys;   // 24 per node: bound keys of both children
  // outcome of every node of the level being split (index = node id - level start)
  uint2* tsplit;               // (elements left of the plane, axis); (0, 0) for a leaf
  float* tbound;               // 24 per node: both children's box and center box
  uint32_t* ttile;             // split nodes per tile of 1024 positions
  unsigned long long* err;     // [0] error flags, [1] largest leaf count over the limit, [2] first bad index entry
  // outputs
  uint32_t* packed;            // 8 words per node, trees one after another
  uint32_t* root_nodes;        // nodes per tree
  void* prims_out;             // permuted triangle triplets (3 * span entries, index_elt bytes each) or the
  int32_t out_elt;             // indirect buffer (span entries of out_elt bytes)
  int32_t indirect;
};
constexpr uint32_t kBuildBigCount = 2048;  // nodes with more triangles are split by all workgroups together
constexpr uint32_t kBuildChunk = 1024;     // big-node primitives per workgroup pass
// Primitive bounds, root nodes and the level-0 big-node list.
hipError_t launch_build_init(const GpuBuildArgs& a, hipStream_t s);
// Splits the nodes of one depth; `wide` launches the big-node steps (needed while the level has big nodes).
hipError_t launch_build_level(const GpuBuildArgs& a, uint32_t depth, bool wide, hipStream_t s);
// Packs levels [0, n_levels) into `packed` (depth-first, left child first) and writes the primitives out.
hipError_t launch_build_pack(const GpuBuildArgs& a, uint32_t n_levels, uint64_t total_nodes, hipStream_t s);

// ---- instances and the batched multi-mesh launch -------------------------------------------------
struct InstanceRayArgs {
  const float* inverse;          // 16 floats per instance (world -> local, column-major)
  uint32_t n_instances;
  const float* origins;          // world rays
  const float* directions;
  const uint32_t* ray_instance;
  uint64_t n;
  float* out_origins;            // local rays (NaN for an invalid instance)
  float* out_directions;
};

struct InstanceHitArgs {
  const float* world;            // 16 floats per instance (local -> world)
  uint32_t n_instances;
  const float* origins;          // world ray origins
  const uint32_t* ray_instance;
  uint64_t n;
  const uint32_t* face;          // hits (misses are left alone)
  float* point;                  // local in, world out
  float* distance;               // world distance out
};

struct BatchArgs {
  const MeshView* meshes;
  uint32_t n_meshes;
  const float* world;
  const float* inverse;
  const uint32_t* instance_geometry;
  uint32_t n_instances;
  const float* origins;
  const float* directions;
  const uint32_t* ray_instance;
  const uint32_t* order;         // visit order (rays grouped by mesh), may be NULL
  uint64_t n;
  uint32_t side_mask;
  float near_dist;
  float far_dist;
  // Local mode (inverse == NULL): the rays are already in their mesh's frame as hi + lo pairs
  // (origins / directions the hi parts, these the lo parts), ray_instance[i] names the MESH, each
  // ray has its own window (4 floats: near, near lo, far, far lo) and results stay local.
  const float* origins_lo;
  const float* directions_lo;
  const float* windows;
  uint32_t* out_face;
  float* out_dist;
  float* out_point;
  float* out_bary;
  uint32_t* counter;             // work-queue counter (reset by the launch)
  uint32_t tile;                 // rays per work item
  uint64_t n_tiles;
  StackScratch stack;
};

// Scene: closest first hit over instances, one launch (TLAS over the instances' world boxes).
struct SceneArgs {
  const MeshView* meshes;
  uint32_t n_meshes;
  const uint32_t* tlas;            // reference-format nodes (8 words) over the instances' world boxes
  const uint32_t* tlas_inst;       // instance at each TLAS primitive slot
  const float* world;              // 16 floats per instance (local -> world, column-major)
  const float* inverse;            // 16 floats per instance (world -> local)
  // closest point's box bounds in the instance frame: 16 floats per instance, the float rounding of
  // the double-precision inverse of `world` computed at scene creation, and whether that inverse
  // passed its residual check (0: world boxes only)
  const float* frame;
  const uint8_t* frame_ok;
  const uint32_t* instance_geometry;
  const float* origins;            // world rays
  const float* directions;
  uint64_t n;
  uint32_t side_mask;
  float near_dist;
  float far_dist;
  uint32_t* out_face;
  uint32_t* out_instance;          // instance of the hit (MBVH_NO_HIT on a miss)
  float* out_dist;                 // world distance
  float* out_point;                // world point
  float* out_bary;
  StackScratch stack;
  const uint32_t* order;           // closest point: visit order of the points (NULL: 0..n-1)
  uint32_t* queue;                 // closest point, four lanes: work-queue counter
  // closest point, four lanes: a point still unfinished after `heavy_budget` leaf steps is left to
  // the wave-per-point pass, its index appended to `heavy` (count in heavy_count); heavy NULL: none
  uint32_t* heavy;
  uint32_t* heavy_count;
  uint32_t heavy_budget;
};
hipError_t launch_scene_raycast_first(const SceneArgs& a, uint64_t threads, hipStream_t s);
// Closest point over the instances: `origins` holds the world points (directions, side, near / far unused).
hipError_t launch_scene_closest_point(const SceneArgs& a, uint64_t threads, hipStream_t s);
// The same with four lanes per point (threads = 4 per point, a multiple of the block; a work queue).
hipError_t launch_scene_closest_point4(const SceneArgs& a, uint64_t threads, hipStream_t s);
// The points left by the four-lane kernel (a.heavy, *a.heavy_count of them), one wave each over an
// LDS frontier of kWideFrontier entries; `waves` persistent waves. The instance trees' 4-wide depth
// must fit that frontier: 64 * (max_stack4 - 1) + 256 <= kWideFrontier.
constexpr uint32_t kWideFrontier = 4096;
hipError_t launch_scene_closest_point_wide(const SceneArgs& a, uint32_t waves, hipStream_t s);

// ---- intersectsBox ----------------------------------------------------------------------------------
struct BoxArgs {
  MeshView mesh;
  StackScratch stack;
  const double* box_min;    // 3 per box
  const double* box_max;    // 3 per box
  const double* matrices;   // 16 per box: box -> mesh frame (column-major)
  uint64_t n;
  uint8_t* out;             // 1 intersects, 0 not, 2 stack overflow
};
hipError_t launch_intersects_box(const BoxArgs& a, uint64_t threads, hipStream_t s);

// ---- closestPointToGeometry --------------------------------------------------------------------------
// The answer record (the layout of mbvh_geometry_hit in the C ABI header).
struct GeometryHit {
  double distance;        // +inf: no pair closer than maxThreshold
  double point[3];        // on the mesh triangle, mesh frame
  double other_point[3];  // on the other triangle, the other geometry's frame
  double other_distance;  // |inverse(point) - other_point|
  uint32_t face_index;    // mesh triangle slot
  uint32_t other_face_index;
  int32_t status;         // 0, 1 traversal stack overflow, 2 an index entry out of range
  uint32_t pad;
};

struct GeometryArgs {
  MeshView mesh;
  StackScratch stack;
  const float* other_positions;   // the other geometry (device)
  uint64_t other_vertices;
  uint32_t other_stride;          // floats between vertices
  uint32_t other_offset;          // float offset of x
  const void* other_index;        // NULL: non-indexed
  int32_t other_elt;              // 2 or 4 bytes
  uint64_t n_other;               // triangles of the other geometry
  double matrix[16];              // other geometry frame -> mesh frame (column-major)
  double inverse[16];             // mesh frame -> other geometry frame
  double max_threshold;
  unsigned long long* best_bits;  // workspace: the best distance of all lanes (bits of a non-negative double)
  uint32_t* winner;               // workspace: the winning other triangle
  uint32_t* status;               // workspace: 1 stack overflow, 2 index out of range
  double* lane_dist;              // workspace, per other triangle: its closest distance found
  uint32_t* lane_face;            // workspace, per other triangle: the mesh slot of that pair
  GeometryHit* out;               // device
};
// Enqueues the kernels of one query (clear, seed bound, per-triangle traversal, pick, gather).
hipError_t launch_closest_to_geometry(const GeometryArgs& a, uint64_t threads, hipStream_t s);

// ---- export to the reference's shader data layouts ------------------------------------------------
struct ExportArgs {
  const uint32_t* nodes;     // reference-format nodes of one root (8 words each)
  uint64_t n_nodes;
  float* bounds;             // MeshBVHUniformStruct bvhBounds: 2 RGBA32F texels per node, side^2 texels
  uint64_t bounds_texels;
  uint32_t* contents;        // bvhContents: 1 RG32UI texel per node
  uint64_t contents_texels;
  const uint32_t* vidx;      // 3 vertex ids per triangle slot (the permuted / dereferenced index)
  uint64_t n_tris;
  uint32_t* index;           // index texture: RGBA texels (a, b, c, 1) per triangle
  uint64_t index_texels;
  const float* positions;    // optional: vertex positions, vertex v at positions[v * stride + offset]
  uint32_t stride;
  uint32_t offset;
  uint64_t n_vertices;
  float* position;           // position texture: RGBA32F texels (x, y, z, 1) per vertex
  uint64_t position_texels;
};
hipError_t launch_export_textures(const ExportArgs& a, hipStream_t s);
// WebGPU BVHNode buffer of one mesh (appendBVHSubtree, src/webgpu/utils/packBVHBufferUtils.js): the
// nodes of every root in order, an empty root's inverted bounds written as [1, -1].
hipError_t launch_export_nodes(const uint32_t* nodes, uint64_t n_nodes, const uint64_t* root_first, uint32_t n_roots,
                               uint32_t* out, hipStream_t s);

hipError_t launch_instance_rays(const InstanceRayArgs& a, hipStream_t s);
hipError_t launch_instance_hits(const InstanceHitArgs& a, hipStream_t s);
// lanes 1: one lane per ray (stack_lds 8 or 16 LDS stack entries per lane); lanes 4: four lanes
// per ray on the 4-wide trees; blocks: persistent grid size
hipError_t launch_batch_raycast_first(const BatchArgs& a, uint32_t stack_lds, uint32_t lanes, uint32_t blocks,
                                      hipStream_t s);
// Visit order grouping rays by the mesh of their instance (stable radix sort on the mesh id; a
// hot mesh does not serialise on one counter); workspace of order_by_mesh_workspace_bytes(n)
// bytes. *order points into the workspace.
uint64_t order_by_mesh_workspace_bytes(uint64_t n);
hipError_t launch_order_by_mesh(const uint32_t* ray_instance, const uint32_t* instance_geometry, uint32_t n_instances,
                                uint64_t n, uint32_t n_meshes, void* workspace, const uint32_t** order, hipStream_t s);

// Stable LSD radix sort (8-bit digits) of n (key, value) pairs by the low `bits` key bits. The
// caller fills the input arrays radix_sort_inputs() names inside the workspace; the sorted arrays
// are returned (also inside the workspace).
uint64_t radix_sort_workspace_bytes(uint64_t n);
void radix_sort_inputs(void* ws, uint64_t n, uint32_t** keys, uint32_t** vals);
hipError_t launch_radix_sort(void* ws, uint64_t n, uint32_t bits, uint32_t** sorted_keys, uint32_t** sorted_vals,
                             hipStream_t s);

// ---- backward ----------------------------------------------------------------------------------
constexpr int32_t kGradAtomic = 0;
constexpr int32_t kGradDeterministic = 1;

struct GradMesh {
  const float4* tris;         // 3 per slot (vertex positions in .xyz)
  const uint32_t* vidx;       // 3 vertex ids per slot
  const uint32_t* face_slot;  // face -> slot for indirect meshes (NULL: slot = face)
  uint64_t face_limit;        // faces are < face_limit
};

// ---- sphere intersection --------------------------------------------------------------------------
struct SphereArgs {
  GradMesh mesh;            // triangles and the face -> slot map
  const float* centers;     // 3 per sphere
  const float* radii;
  const uint32_t* face;     // closest face within the largest radius (MBVH_NO_HIT: none)
  uint64_t n;
  uint8_t* out;             // 1 intersects, 0 not, 2 the closest-point search overflowed
};
hipError_t launch_sphere_decide(const SphereArgs& a, hipStream_t s);

struct VertexSink {
  float* grad;        // atomic mode: vertex gradients, 3 floats per vertex
  uint32_t* keys;     // deterministic mode: vertex id of each contribution (3 per query)
  uint32_t* vals;     // contribution index
  float* contrib;     // 3 floats per contribution
  uint32_t sentinel;  // vertex id meaning "no contribution" (= vertex count)
};

struct RaycastBwdArgs {
  GradMesh mesh;
  const float* directions;
  const uint32_t* face;         // saved forward outputs
  const float* distance;
  const float* barycoord;
  const float* grad_distance;   // upstream gradients (each may be NULL = 0)
  const float* grad_point;
  const float* grad_barycoord;
  float* grad_origins;          // outputs (each may be NULL)
  float* grad_directions;
  uint64_t n;
  bool want_vertices;
  VertexSink sink;
};

struct ClosestBwdArgs {
  GradMesh mesh;
  const float* points;
  const uint32_t* face;         // saved forward outputs
  const float* distance;
  const float* closest;
  const float* barycoord;
  const float* grad_distance;   // upstream gradient
  float* grad_points;           // output (may be NULL)
  uint64_t n;
  bool want_vertices;
  VertexSink sink;
};

// Device workspace the backward needs for n queries (0 for the atomic mode), 256-byte aligned.
uint64_t grad_workspace_bytes(uint64_t n_queries, int32_t mode);
// grad_positions (NULL: no vertex gradients) receives n_vertices x 3 floats, overwritten.
hipError_t launch_raycast_first_bwd(const RaycastBwdArgs& a, int32_t mode, float* grad_positions,
                                    uint64_t n_vertices, void* workspace, hipStream_t s);
hipError_t launch_closest_point_bwd(const ClosestBwdArgs& a, int32_t mode, float* grad_positions,
                                    uint64_t n_vertices, void* workspace, hipStream_t s);

// Threads a launch uses for n queries: one query per thread, capped at `max_threads` (the
// launch then strides over the remaining queries).
uint64_t launch_threads(uint64_t n, uint64_t max_threads);

hipError_t launch_raycast_first(const RaycastFirstArgs& a, uint64_t threads, hipStream_t s);
// four lanes per ray on the 4-wide tree; `quads` = rays traced at once (a multiple of 64)
hipError_t launch_raycast_first4(const RaycastFirstArgs& a, uint64_t quads, hipStream_t s);
hipError_t launch_closest_point(const ClosestPointArgs& a, uint64_t threads, hipStream_t s);
// four lanes per query on the 4-wide tree (minThreshold must be 0); `threads` lanes resident
hipError_t launch_closest_point4(const ClosestPointArgs& a, uint64_t threads, hipStream_t s);
// One wave of 64 queries per packet walking the 4-wide tree together (coherent queries); the tree's
// 4-wide stack must fit kPacketStackEntries. The traversal lists each query's candidates, a second
// launch resolves them exactly; queries whose candidates overflow the list go to a redo list that
// the caller finishes with launch_closest_point4 (n_dev = redo_count, order = redo).
constexpr uint32_t kPacketStackEntries = 64;
constexpr uint32_t kPacketCand = 4;  // the LDS list per query (kCand)
// queries per packet launch: the buffers (48 bytes per query, plus the pool) are sized for it. Each
// launch ends with its slowest packets, so fewer, larger launches finish sooner.
constexpr uint64_t kPacketChunk = uint64_t{1} << 24;
constexpr uint32_t kPacketPool = 1u << 23;  // further candidates per launch
hipError_t launch_closest_point_packet(const ClosestPointArgs& a, uint64_t threads, hipStream_t s);
// fill == false: count hits per ray into a.counts; fill == true: write them at a.offsets.
hipError_t launch_raycast_all(const RaycastAllArgs& a, bool fill, uint64_t threads, hipStream_t s);
// Spatial order of n packed float triples (Morton bucket sort); workspace of
// order_workspace_bytes(n) bytes, 256-byte aligned.
constexpr uint64_t kOrderHeaderBytes = 256;
uint32_t order_axis_bits(uint64_t n);
uint64_t order_workspace_bytes(uint64_t n);
hipError_t launch_order_points(const float* points, uint64_t n, uint32_t* order, void* workspace, hipStream_t s);
hipError_t launch_apply_side_mask(float4* tris, uint64_t n_slots, const uint8_t* mask,
                                  hipStream_t s);

}  // namespace mbvh
