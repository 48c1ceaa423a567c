// C ABI runtime: device context, upload-once meshes and query dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <map>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "mbvh_common.h"
#include "mbvh_kernels.h"

using namespace mbvh;

namespace {

constexpr uint64_t kMaxThreadsPerCU = 2048;  // 32 waves of 64 lanes
// host calls of at most this many queries use the context's pinned block: 12 input and 8 output
// words per query (origins, directions and their lo parts; face, distance, point, barycoord)
constexpr uint64_t kSmallQueries = 256;
constexpr size_t kSmallBytes = kSmallQueries * 20 * sizeof(float);

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

// Capacity of the context's list of queries left to a wave each: a launch of more queries runs
// without that pass.
constexpr uint64_t kHeavyCap = uint64_t{1} << 20;

struct mbvh_ctx {
  uint32_t magic = kCtxMagic;
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t max_threads = 0;     // resident-thread cap for one launch
  DeviceBuffer stack_ovf;       // traversal stack overflow region
  DeviceBuffer queue;           // work-queue counter of the single-mesh query kernels
  int32_t ray_lanes = 0;        // raycastFirst lanes per ray: 0 automatic, 1, or 4
  int32_t point_lanes = 0;      // closestPoint lanes per query: 0 automatic, 1, or 4
  uint32_t stack_ovf_entries = 0;
  uint64_t stack_gen = 0;       // bumped whenever stack_ovf moves (captured graphs hold its address)
  DeviceBuffer staging[8];      // host-API input / output staging
  DeviceBuffer order;           // host-API query order
  DeviceBuffer order_ws;        // host-API ordering workspace
  DeviceBuffer heavy;           // queries left to a wave each (kHeavyCap entries, allocated with the context)
  DeviceBuffer hits[5];         // host-API all-hits results: offsets, face, dist, point, bary
  DeviceBuffer grad[11];        // host-API backward staging
  DeviceBuffer grad_ws;         // host-API backward workspace
  DeviceBuffer build_ws;        // device build workspace
  DeviceBuffer geo_ws;          // closestPointToGeometry workspace (per other triangle + answer)
  DeviceBuffer lo_stage[3];     // host-API lo parts of double-precision inputs (origins / points, directions, windows)
  // small host calls (one ray or point from a per-call API): inputs and outputs in one pinned host
  // block the kernels read and write in place (no staging copies); host and device addresses
  void* small_host = nullptr;
  void* small_dev = nullptr;
  DeviceBuffer packet[4];       // packet closest point, per chunk: candidate lists, per-query metadata, counters + redo list, pool
  uint64_t hits_total = 0;
  std::mutex lock;              // host entry points on one context are serialised
};

// Device memory of the allocation-free closest-point entry point (mbvh_closest_point_noalloc_d):
// work-queue counter, query order, ordering workspace and traversal stack overflow. Every GPU the
// library runs on holds its own copy until the process ends; host code must never pass this name
// to a kernel or a copy (it is a host-side stand-in), only the per-device address looked up below.
static __device__ __align__(256) unsigned char scratch[MBVH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxScratchDevices] = {};  // device address of `scratch`, per device

// Looks up (once per device, with that device current) the device address of `scratch`.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return MBVH_OK;  // the entry point reports it
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (g_scratch_ptr[device]) return MBVH_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(MBVH_ERR_HIP, "scratch lookup on device %d failed (%s)", device, hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return MBVH_OK;
}

void* scratch_of(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  return g_scratch_ptr[device];
}

}  // namespace

struct mbvh_batch {
  uint32_t magic = 0x48435442u;  // "BTCH"
  mbvh_ctx* ctx = nullptr;
  uint32_t n_meshes = 0;
  MeshView* d_meshes = nullptr;
  uint32_t* d_counter = nullptr;  // work-queue counter
  uint32_t max_stack = 0;
  uint32_t max_stack4 = 0;
  mbvh_batch_tuning tuning{4, 16, 256, 4, 0.0f};
};

struct mbvh_batch_graph {
  uint32_t magic = 0x48505247u;  // "GRPH"
  mbvh_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t stack_gen = 0;  // the context's stack_gen at capture
};

struct mbvh_mesh {
  uint32_t magic = kMeshMagic;
  mbvh_ctx* ctx = nullptr;
  WideNode* d_nodes = nullptr;
  uint32_t* d_roots = nullptr;
  uint2* d_root_slots = nullptr;    // slot range (first, end) of each root
  uint2* d_bigleaf = nullptr;
  float4* d_tris = nullptr;
  uint32_t* d_vidx = nullptr;
  uint32_t* d_face_slot = nullptr;  // indirect meshes: face -> slot (0xFFFFFFFF if absent)
  float4* d_nodes4 = nullptr;       // 4-wide tree
  uint32_t* d_roots4 = nullptr;
  uint32_t max_stack4 = 0;
  uint64_t n_nodes = 0;
  uint32_t n_roots = 0;
  uint64_t n_slots = 0;
  uint64_t vertex_count = 0;
  uint64_t face_limit = 0;  // face ids are < face_limit
  uint32_t max_depth = 0;   // deepest reference node depth over all roots
  uint32_t max_stack = 0;   // traversal stack entries needed
  float* d_coord_max = nullptr;  // device: largest |vertex coordinate| (after a refit: over all vertices)
  // refit: the reference-format nodes of all roots and how their boxes map onto the device layouts
  uint32_t* d_ref_nodes = nullptr;   // 8 words per node, roots concatenated
  uint32_t* d_ref_levels = nullptr;  // reachable internal nodes, grouped by depth
  int32_t* d_ref_wide = nullptr;     // float offset of the node's box in the wide nodes (-1: none)
  uint32_t* d_ref_leaves = nullptr;  // leaf nodes
  uint32_t* d_ref_scratch = nullptr; // per node: changed flag
  uint2* d_four_map = nullptr;       // 4-wide child slot <- wide box (float offsets)
  uint64_t* d_root_first = nullptr;  // first reference node of each root (the exports)
  uint64_t n_ref_nodes = 0, n_ref_leaves = 0, n_four_map = 0;
  std::vector<uint64_t> root_base;   // first node of each root
  std::vector<uint64_t> root_nodes;  // nodes of each root
  std::vector<uint64_t> level_off;   // d_ref_levels[level_off[d] .. level_off[d + 1]) are at depth d
};

namespace {

#define MBVH_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess)                                                                 \
      return set_error(_e == hipErrorOutOfMemory ? MBVH_ERR_OOM : MBVH_ERR_HIP, "%s: %s (%s)", \
                       __func__, #call, hipGetErrorString(_e));                           \
  } while (0)

bool valid_ctx(const mbvh_ctx* c) { return c && c->magic == kCtxMagic; }
bool valid_mesh(const mbvh_mesh* m) { return m && m->magic == kMeshMagic; }

int ensure_buffer(DeviceBuffer& b, size_t bytes) {
  if (b.bytes >= bytes) return MBVH_OK;
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return set_error(MBVH_ERR_OOM, "device allocation of %zu bytes failed (%s)", want,
                     hipGetErrorString(e));
  }
  b.bytes = want;
  return MBVH_OK;
}

// Per-thread stack overflow entries of a context: enough for a stack of max_stack entries with
// the smallest LDS part any kernel uses (8). Sized for a 64-deep stack on first need, so the region
// (whose address launches and captured graphs hold) moves only for deeper trees.
constexpr uint32_t kMinLdsStack = 8;
constexpr uint32_t kOvfEntriesDefault = 64 - kMinLdsStack;

int reserve_stack(mbvh_ctx* ctx, uint32_t max_stack) {
  if (max_stack <= kMinLdsStack) return MBVH_OK;
  const uint32_t need = max_stack - kMinLdsStack;
  if (need <= ctx->stack_ovf_entries) return MBVH_OK;
  const uint32_t entries = std::max(need, kOvfEntriesDefault);
  if (ctx->stack_ovf.ptr) {
    const hipError_t e = hipDeviceSynchronize();  // earlier launches may still use the old region
    if (e != hipSuccess) return set_error(MBVH_ERR_HIP, "stack scratch: %s", hipGetErrorString(e));
  }
  ctx->stack_gen++;  // a graph captured before this point holds the old address
  const int rc = ensure_buffer(ctx->stack_ovf, static_cast<size_t>(entries) * ctx->max_threads * sizeof(uint2));
  if (rc) {
    ctx->stack_ovf_entries = 0;
    return rc;
  }
  ctx->stack_ovf_entries = entries;
  return MBVH_OK;
}

uint32_t inline_leaf_ref(uint64_t off, uint64_t cnt, std::vector<uint2>& bigleaf) {
  if (cnt >= 1 && cnt <= kMaxInlineCount && off <= kRefOffsetMask)
    return kRefLeaf | (static_cast<uint32_t>(cnt) << kRefCountShift) | static_cast<uint32_t>(off);
  bigleaf.push_back(make_uint2(static_cast<uint32_t>(off), static_cast<uint32_t>(cnt)));
  return kRefLeaf | static_cast<uint32_t>(bigleaf.size() - 1);
}

// Where each reference node's box lives in the device layouts (for refit); node indices count over
// all roots, `base` being this root's first.
struct RefitMaps {
  uint64_t base = 0;
  std::vector<int32_t> wide_off;  // float offset of the node's box in the wide nodes, -1: none
  std::vector<std::vector<uint32_t>> internal_by_depth;  // reachable internal nodes per depth
  std::vector<uint32_t> leaves;   // reachable leaves
  std::vector<uint2> four_map;    // (float offset of a 4-wide child slot, float offset of its wide box)
};

// Converts one reference root into wide nodes. Validates every index it follows so a corrupt or
// foreign buffer is rejected on the host instead of steering a kernel out of bounds.
int convert_root(const NodeView& v, uint64_t n_slots, std::vector<WideNode>& wide,
                 std::vector<uint2>& bigleaf, uint32_t* root_ref, uint32_t* max_depth, RefitMaps& maps) {
  if (v.n_nodes == 0) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: empty root buffer");
  auto leaf_ok = [&](uint64_t node) {
    const uint64_t off = v.leaf_offset(node), cnt = v.leaf_count(node);
    return off + cnt <= n_slots;
  };
  struct Item {
    uint64_t node;
    uint32_t depth;
    int64_t parent_wide;  // -1 for the root
    int side;             // 0 left, 1 right
  };
  std::vector<Item> todo;
  todo.push_back({0, 0, -1, 0});
  if (maps.base + v.n_nodes > 0x7FFFFFFFull) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many nodes");
  maps.wide_off.resize(maps.base + v.n_nodes, -1);
  uint64_t visited = 0;
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    if (++visited > v.n_nodes)
      return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: node buffer contains a cycle");
    *max_depth = std::max(*max_depth, it.depth);
    const uint64_t g = maps.base + it.node;
    if (it.parent_wide >= 0) maps.wide_off[g] = static_cast<int32_t>(16 * it.parent_wide + 6 * it.side);
    uint32_t ref;
    if (v.is_leaf(it.node)) {
      if (!leaf_ok(it.node))
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: leaf %llu references triangles past the end",
                         static_cast<unsigned long long>(it.node));
      maps.leaves.push_back(static_cast<uint32_t>(g));
      ref = inline_leaf_ref(v.leaf_offset(it.node), v.leaf_count(it.node), bigleaf);
    } else {
      if (maps.internal_by_depth.size() <= it.depth) maps.internal_by_depth.resize(it.depth + 1);
      maps.internal_by_depth[it.depth].push_back(static_cast<uint32_t>(g));
      const uint64_t left = it.node + 1, right = v.right_child(it.node);
      const uint32_t axis = v.split_axis(it.node);
      if (axis > 2 || left >= v.n_nodes || right >= v.n_nodes || right <= it.node)
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: malformed internal node %llu",
                         static_cast<unsigned long long>(it.node));
      if (wide.size() >= 0x7FFFFFFull)  // float offsets of wide boxes fit 31 bits
        return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many nodes");
      const uint32_t w = static_cast<uint32_t>(wide.size());
      WideNode wn;
      const float* L = v.bounds(left);
      const float* R = v.bounds(right);
      wn.b0 = make_float4(L[0], L[1], L[2], L[3]);
      wn.b1 = make_float4(L[4], L[5], R[0], R[1]);
      wn.b2 = make_float4(R[2], R[3], R[4], R[5]);
      wn.meta = make_uint4(0, 0, axis, 0);
      wide.push_back(wn);
      ref = w;
      todo.push_back({right, it.depth + 1, static_cast<int64_t>(w), 1});
      todo.push_back({left, it.depth + 1, static_cast<int64_t>(w), 0});
    }
    if (it.parent_wide < 0) {
      *root_ref = ref;
    } else if (it.side == 0) {
      wide[it.parent_wide].meta.x = ref;
    } else {
      wide[it.parent_wide].meta.y = ref;
    }
  }
  return MBVH_OK;
}

// Slot ranges: the reference partitions a node's triangles into [left | right] and lays leaves out
// depth-first, so each subtree covers a contiguous run of slots. Stores in meta.w of every wide node
// the first slot of its right subtree, and returns each root's slot range (first, end). Wide nodes
// are in preorder (children after parents), so one backward pass sees children first.
void assign_split_slots(std::vector<WideNode>& wide, const std::vector<uint2>& bigleaf,
                        const std::vector<uint32_t>& root_refs, std::vector<uint2>& root_slots) {
  auto leaf = [&](uint32_t ref, uint32_t& off, uint32_t& cnt) {
    const uint32_t c = (ref >> kRefCountShift) & kRefCountMask;
    if (c) {
      off = ref & kRefOffsetMask;
      cnt = c;
    } else {
      off = bigleaf[ref & kRefOffsetMask].x;
      cnt = bigleaf[ref & kRefOffsetMask].y;
    }
  };
  std::vector<uint32_t> first(wide.size()), end(wide.size());
  auto range = [&](uint32_t ref, uint32_t& lo, uint32_t& hi) {
    if (ref & kRefLeaf) {
      uint32_t off, cnt;
      leaf(ref, off, cnt);
      lo = off;
      hi = off + cnt;
    } else {
      lo = first[ref];
      hi = end[ref];
    }
  };
  for (size_t w = wide.size(); w-- > 0;) {
    uint32_t llo, lhi, rlo, rhi;
    range(wide[w].meta.x, llo, lhi);
    range(wide[w].meta.y, rlo, rhi);
    wide[w].meta.w = rlo;
    first[w] = std::min(llo, rlo);
    end[w] = std::max(lhi, rhi);
  }
  root_slots.resize(root_refs.size());
  for (size_t r = 0; r < root_refs.size(); r++) range(root_refs[r], root_slots[r].x, root_slots[r].y);
}

// 4-wide tree from the wide binary nodes: a 4-wide node holds the children of a binary node's
// children (a child that is a leaf stays a child), so it replaces two binary levels. Child k of
// node n is nodes4[8n + 2k] = {min.x, min.y, min.z, max.x}, nodes4[8n + 2k + 1] = {max.y, max.z,
// ref, 0} with the reference's float32 bounds bit for bit; unused children have ref kDone4 and an
// empty box. Leaf references are the binary tree's.
constexpr uint32_t kDone4 = 0x7FFFFFFFu;

int build_bvh4(const std::vector<WideNode>& wide, uint32_t root_ref, std::vector<float4>& nodes4,
               uint32_t* root4, uint32_t* depth4, std::vector<uint2>& four_map) {
  struct Item {
    uint32_t bref;    // binary (wide) node
    int64_t parent4;  // -1: the root
    int slot;
    uint32_t depth;
  };
  if (root_ref & kRefLeaf) {
    *root4 = root_ref;
    return MBVH_OK;
  }
  std::vector<Item> todo{{root_ref, -1, 0, 1}};
  auto as_float = [](uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    if (nodes4.size() / 8 >= 0x7FFFFFFEull) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many nodes");
    const uint32_t n4 = static_cast<uint32_t>(nodes4.size() / 8);
    *depth4 = std::max(*depth4, it.depth);
    if (it.parent4 < 0)
      *root4 = n4;
    else
      nodes4[8 * it.parent4 + 2 * it.slot + 1].z = as_float(n4);
    struct Child {
      float box[6];
      uint32_t ref;
      uint32_t wide_off;  // float offset of the box in the wide nodes
    };
    Child ch[4];
    int nc = 0;
    const WideNode& w = wide[it.bref];
    const float lb[6] = {w.b0.x, w.b0.y, w.b0.z, w.b0.w, w.b1.x, w.b1.y};
    const float rb[6] = {w.b1.z, w.b1.w, w.b2.x, w.b2.y, w.b2.z, w.b2.w};
    const uint32_t refs[2] = {w.meta.x, w.meta.y};
    const float* boxes[2] = {lb, rb};
    for (int side = 0; side < 2; side++) {
      const uint32_t c = refs[side];
      if (c & kRefLeaf) {
        std::memcpy(ch[nc].box, boxes[side], sizeof(ch[nc].box));
        ch[nc].wide_off = 16 * it.bref + 6 * side;
        ch[nc++].ref = c;
      } else {
        const WideNode& g = wide[c];
        const float gl[6] = {g.b0.x, g.b0.y, g.b0.z, g.b0.w, g.b1.x, g.b1.y};
        const float gr[6] = {g.b1.z, g.b1.w, g.b2.x, g.b2.y, g.b2.z, g.b2.w};
        std::memcpy(ch[nc].box, gl, sizeof(gl));
        ch[nc].wide_off = 16 * c;
        ch[nc++].ref = g.meta.x;
        std::memcpy(ch[nc].box, gr, sizeof(gr));
        ch[nc].wide_off = 16 * c + 6;
        ch[nc++].ref = g.meta.y;
      }
    }
    const float inf = std::numeric_limits<float>::infinity();
    for (int k = 0; k < 4; k++) {
      if (k < nc) {
        const Child& c = ch[k];
        if (4 * nodes4.size() > 0xFFFFFFF0ull) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many nodes");
        four_map.push_back(make_uint2(static_cast<uint32_t>(4 * nodes4.size()), c.wide_off));
        nodes4.push_back(make_float4(c.box[0], c.box[1], c.box[2], c.box[3]));
        nodes4.push_back(make_float4(c.box[4], c.box[5], as_float(c.ref), 0.0f));
      } else {
        nodes4.push_back(make_float4(inf, inf, inf, -inf));
        nodes4.push_back(make_float4(-inf, -inf, as_float(kDone4), 0.0f));
      }
    }
    for (int k = 0; k < nc; k++)
      if (!(ch[k].ref & kRefLeaf)) todo.push_back({ch[k].ref, static_cast<int64_t>(n4), k, it.depth + 1});
  }
  return MBVH_OK;
}

uint64_t read_u(const void* p, int elt_bytes, uint64_t i) {
  if (elt_bytes == 2) return static_cast<const uint16_t*>(p)[i];
  return static_cast<const uint32_t*>(p)[i];
}

void free_mesh_buffers(mbvh_mesh* m) {
  if (m->d_nodes) (void)hipFree(m->d_nodes);
  if (m->d_roots) (void)hipFree(m->d_roots);
  if (m->d_root_slots) (void)hipFree(m->d_root_slots);
  if (m->d_bigleaf) (void)hipFree(m->d_bigleaf);
  if (m->d_tris) (void)hipFree(m->d_tris);
  if (m->d_vidx) (void)hipFree(m->d_vidx);
  if (m->d_face_slot) (void)hipFree(m->d_face_slot);
  if (m->d_nodes4) (void)hipFree(m->d_nodes4);
  if (m->d_roots4) (void)hipFree(m->d_roots4);
  for (void* p : {static_cast<void*>(m->d_ref_nodes), static_cast<void*>(m->d_ref_levels),
                  static_cast<void*>(m->d_ref_wide), static_cast<void*>(m->d_ref_leaves),
                  static_cast<void*>(m->d_ref_scratch), static_cast<void*>(m->d_four_map),
                  static_cast<void*>(m->d_coord_max), static_cast<void*>(m->d_root_first)})
    if (p) (void)hipFree(p);
  m->d_ref_nodes = nullptr;
  m->d_ref_levels = nullptr;
  m->d_ref_wide = nullptr;
  m->d_ref_leaves = nullptr;
  m->d_ref_scratch = nullptr;
  m->d_four_map = nullptr;
  m->d_coord_max = nullptr;
  m->d_root_first = nullptr;
  m->d_face_slot = nullptr;
  m->d_nodes4 = nullptr;
  m->d_roots4 = nullptr;
  m->d_nodes = nullptr;
  m->d_roots = nullptr;
  m->d_bigleaf = nullptr;
  m->d_tris = nullptr;
  m->d_vidx = nullptr;
}

MeshView view_of(const mbvh_mesh* m) {
  MeshView v;
  v.nodes = m->d_nodes;
  v.roots = m->d_roots;
  v.root_slots = m->d_root_slots;
  v.nodes4 = m->d_nodes4;
  v.roots4 = m->d_roots4;
  v.max_stack4 = m->max_stack4;
  v.bigleaf = m->d_bigleaf;
  v.tris = m->d_tris;
  v.n_roots = m->n_roots;
  v.max_stack = m->max_stack;
  v.coord_max = m->d_coord_max;
  return v;
}

GradMesh grad_view_of(const mbvh_mesh* m) {
  GradMesh g;
  g.tris = m->d_tris;
  g.vidx = m->d_vidx;
  g.face_slot = m->d_face_slot;
  g.face_limit = m->face_limit;
  return g;
}

uint32_t tris_face(const std::vector<float4>& tris, uint64_t slot) {
  uint32_t f;
  std::memcpy(&f, &tris[3 * slot].w, 4);
  return f;
}

StackScratch stack_of(const mbvh_ctx* c, uint64_t threads) {
  StackScratch s;
  s.ovf = static_cast<uint2*>(c->stack_ovf.ptr);
  s.ovf_stride = threads;
  s.ovf_entries = c->stack_ovf_entries;
  return s;
}

uint32_t side_mask_of(int32_t side) {
  switch (side) {
    case MBVH_FRONT_SIDE: return 1u;
    case MBVH_BACK_SIDE: return 2u;
    case MBVH_DOUBLE_SIDE: return 3u;
    default: return 0u;  // per-triangle
  }
}

int check_query(const char* fn, const mbvh_ctx* ctx, const mbvh_mesh* mesh) {
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!valid_mesh(mesh)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid mesh handle", fn);
  if (mesh->ctx != ctx) return set_error(MBVH_ERR_BAD_HANDLE, "%s: mesh belongs to another context", fn);
  return MBVH_OK;
}

int set_device(const mbvh_ctx* ctx) {
  MBVH_HIP(hipSetDevice(ctx->device));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

extern "C" int mbvh_ctx_create(int32_t device, mbvh_ctx** out_ctx) {
  clear_error();
  if (!out_ctx) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_ctx_create: out_ctx is NULL");
  *out_ctx = nullptr;
  const int n = mbvh_device_count();
  if (n <= 0) return set_error(MBVH_ERR_NO_DEVICE, "mbvh_ctx_create: no HIP device is visible");
  if (device < 0 || device >= n)
    return set_error(MBVH_ERR_NO_DEVICE, "mbvh_ctx_create: device %d out of range (%d visible)", device, n);
  MBVH_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MBVH_HIP(hipGetDeviceProperties(&prop, device));
  if (const int rc = scratch_init(device)) return rc;
  mbvh_ctx* c = new (std::nothrow) mbvh_ctx();
  if (!c) return set_error(MBVH_ERR_OOM, "mbvh_ctx_create: out of host memory");
  c->device = device;
  c->max_threads = static_cast<uint64_t>(prop.multiProcessorCount) * kMaxThreadsPerCU;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(MBVH_ERR_HIP, "mbvh_ctx_create: hipStreamCreate failed (%s)", hipGetErrorString(e));
  }
  if (ensure_buffer(c->queue, 256)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return MBVH_ERR_OOM;
  }
  // the list of queries left to a wave each (long walks), allocated once so that the device-pointer
  // entry points allocate nothing; without it those passes are off
  if (ensure_buffer(c->heavy, kHeavyCap * sizeof(uint32_t))) c->heavy = DeviceBuffer{};
  // the pinned block of small calls (without it, small calls take the staging path)
  if (hipHostMalloc(&c->small_host, kSmallBytes, hipHostMallocMapped) == hipSuccess) {
    if (hipHostGetDevicePointer(&c->small_dev, c->small_host, 0) != hipSuccess) {
      (void)hipHostFree(c->small_host);
      c->small_host = c->small_dev = nullptr;
    }
  } else {
    c->small_host = nullptr;
  }
  *out_ctx = c;
  return MBVH_OK;
}

extern "C" void mbvh_ctx_destroy(mbvh_ctx* ctx) {
  if (!valid_ctx(ctx)) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stack_ovf.ptr) (void)hipFree(ctx->stack_ovf.ptr);
  if (ctx->queue.ptr) (void)hipFree(ctx->queue.ptr);
  for (auto& b : ctx->staging)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto& b : ctx->hits)
    if (b.ptr) (void)hipFree(b.ptr);
  if (ctx->order.ptr) (void)hipFree(ctx->order.ptr);
  for (auto& b : ctx->grad)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto& b : ctx->lo_stage)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto& b : ctx->packet)
    if (b.ptr) (void)hipFree(b.ptr);
  if (ctx->small_host) (void)hipHostFree(ctx->small_host);
  if (ctx->grad_ws.ptr) (void)hipFree(ctx->grad_ws.ptr);
  if (ctx->order_ws.ptr) (void)hipFree(ctx->order_ws.ptr);
  if (ctx->heavy.ptr) (void)hipFree(ctx->heavy.ptr);
  if (ctx->build_ws.ptr) (void)hipFree(ctx->build_ws.ptr);
  if (ctx->geo_ws.ptr) (void)hipFree(ctx->geo_ws.ptr);
  (void)hipStreamDestroy(ctx->stream);
  ctx->magic = 0;
  delete ctx;
}

extern "C" int mbvh_ctx_set_ray_lanes(mbvh_ctx* ctx, int32_t lanes) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_ctx_set_ray_lanes: invalid context handle");
  if (lanes != 0 && lanes != 1 && lanes != 4)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_ctx_set_ray_lanes: lanes must be 0 (automatic), 1 or 4");
  std::lock_guard<std::mutex> guard(ctx->lock);  // host calls in flight on other threads read it
  ctx->ray_lanes = lanes;
  return MBVH_OK;
}

extern "C" int mbvh_ctx_set_point_lanes(mbvh_ctx* ctx, int32_t lanes) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_ctx_set_point_lanes: invalid context handle");
  if (lanes != 0 && lanes != 1 && lanes != 4 && lanes != 64)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_ctx_set_point_lanes: lanes must be 0 (automatic), 1, 4 or 64");
  std::lock_guard<std::mutex> guard(ctx->lock);  // host calls in flight on other threads read it
  if (lanes == 64) {
    // the packets' buffers are allocated here, so that queries allocate nothing
    MBVH_HIP(hipSetDevice(ctx->device));
    if (int rc = ensure_buffer(ctx->packet[0], kPacketChunk * kPacketCand * sizeof(uint2))) return rc;
    if (int rc = ensure_buffer(ctx->packet[1], kPacketChunk * sizeof(uint4))) return rc;
    if (int rc = ensure_buffer(ctx->packet[2], 512 + kPacketChunk * sizeof(uint32_t))) return rc;
    if (int rc = ensure_buffer(ctx->packet[3], size_t{kPacketPool} * sizeof(uint4))) return rc;
  }
  ctx->point_lanes = lanes;
  return MBVH_OK;
}

extern "C" int mbvh_ctx_synchronize(mbvh_ctx* ctx) {
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_ctx_synchronize: invalid context handle");
  int rc = set_device(ctx);
  if (rc) return rc;
  MBVH_HIP(hipStreamSynchronize(ctx->stream));
  return MBVH_OK;
}

extern "C" int mbvh_mesh_upload(mbvh_ctx* ctx, const void* const* roots, const uint64_t* root_bytes,
                                uint32_t n_roots, const void* index, int32_t index_elt_bytes,
                                uint64_t index_count, const void* indirect, int32_t indirect_elt_bytes,
                                uint64_t indirect_count, const float* positions, uint64_t vertex_count,
                                uint32_t position_stride, uint32_t position_offset, mbvh_mesh** out_mesh) {
  clear_error();
  if (!out_mesh) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: out_mesh is NULL");
  *out_mesh = nullptr;
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_mesh_upload: invalid context handle");
  if (n_roots == 0 || !roots || !root_bytes)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: at least one root is required");
  if (index && index_elt_bytes != 2 && index_elt_bytes != 4)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: index element size must be 2 or 4");
  if (index && index_count % 3 != 0)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: index count is not a multiple of 3");
  if (indirect && indirect_elt_bytes != 2 && indirect_elt_bytes != 4)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: indirect element size must be 2 or 4");
  if (!index && !indirect)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: an index or an indirect buffer is required");
  if (position_stride < position_offset + 3)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: position stride too small");
  if (!positions && vertex_count)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: positions is NULL");

  const uint64_t geo_tris = index ? index_count / 3 : vertex_count / 3;
  const uint64_t n_slots = indirect ? indirect_count : geo_tris;

  // Reference trees → wide nodes.
  std::vector<WideNode> wide;
  std::vector<uint2> bigleaf;
  std::vector<uint32_t> root_refs(n_roots);
  uint32_t max_depth = 0;
  RefitMaps maps;
  std::vector<uint64_t> root_base(n_roots), root_nodes(n_roots);
  for (uint32_t r = 0; r < n_roots; r++) {
    if (!roots[r] || root_bytes[r] % kBytesPerNode != 0)
      return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: root %u size is not a multiple of 32 bytes", r);
    NodeView v{static_cast<const uint8_t*>(roots[r]), root_bytes[r] / kBytesPerNode};
    root_base[r] = maps.base;
    root_nodes[r] = v.n_nodes;
    int rc = convert_root(v, n_slots, wide, bigleaf, &root_refs[r], &max_depth, maps);
    if (rc) return rc;
    maps.base += v.n_nodes;
  }
  if (bigleaf.size() > kRefOffsetMask) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many large leaves");
  std::vector<uint2> root_slots;
  assign_split_slots(wide, bigleaf, root_refs, root_slots);
  std::vector<float4> nodes4;
  std::vector<uint32_t> roots4(n_roots);
  uint32_t depth4 = 0;
  for (uint32_t r = 0; r < n_roots; r++) {
    const int rc4 = build_bvh4(wide, root_refs[r], nodes4, &roots4[r], &depth4, maps.four_map);
    if (rc4) return rc4;
  }

  // Per-slot triangle records in BVH order.
  std::vector<float4> tris;
  std::vector<uint32_t> vidx;
  try {
    tris.resize(3 * n_slots);
    vidx.resize(3 * n_slots);
  } catch (...) {
    return set_error(MBVH_ERR_OOM, "mbvh_mesh_upload: out of host memory");
  }
  uint64_t face_limit = 0;
  float coord_max = 0.0f;
  for (uint64_t s = 0; s < n_slots; s++) {
    const uint64_t face = indirect ? read_u(indirect, indirect_elt_bytes, s) : s;
    if (face >= geo_tris)
      return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: indirect[%llu] is out of range",
                       static_cast<unsigned long long>(s));
    if (face > 0xFFFFFFFEull) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: face index exceeds 32 bits");
    face_limit = std::max(face_limit, face + 1);
    uint64_t v[3] = {3 * face, 3 * face + 1, 3 * face + 2};
    if (index)
      for (int k = 0; k < 3; k++) v[k] = read_u(index, index_elt_bytes, v[k]);
    float4 rec[3];
    for (int k = 0; k < 3; k++) {
      if (v[k] >= vertex_count)
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: face %llu references a vertex past the end",
                         static_cast<unsigned long long>(face));
      const float* p = positions + v[k] * position_stride + position_offset;
      rec[k] = make_float4(p[0], p[1], p[2], 0.0f);
      coord_max = std::max(coord_max, std::max(std::fabs(p[0]), std::max(std::fabs(p[1]), std::fabs(p[2]))));
      vidx[3 * s + k] = static_cast<uint32_t>(v[k]);
    }
    uint32_t face32 = static_cast<uint32_t>(face), mask_all = 3u, va = static_cast<uint32_t>(v[0]);
    std::memcpy(&rec[0].w, &face32, 4);
    std::memcpy(&rec[1].w, &mask_all, 4);
    std::memcpy(&rec[2].w, &va, 4);
    tris[3 * s + 0] = rec[0];
    tris[3 * s + 1] = rec[1];
    tris[3 * s + 2] = rec[2];
  }

  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;

  // Stack: one pending entry per internal ancestor at most.
  const uint32_t max_stack = max_depth + 1;
  // a quad pushes up to three children per 4-wide level, and runs binary traversals for rays with a
  // zero direction component
  const uint32_t max_stack4 = std::max(max_stack, 3 * depth4 + 1);
  if ((rc = reserve_stack(ctx, max_stack4))) return rc;

  mbvh_mesh* m = new (std::nothrow) mbvh_mesh();
  if (!m) return set_error(MBVH_ERR_OOM, "mbvh_mesh_upload: out of host memory");
  m->ctx = ctx;
  m->n_nodes = wide.size();
  m->n_roots = n_roots;
  m->n_slots = n_slots;
  m->vertex_count = vertex_count;
  m->face_limit = face_limit;
  m->max_depth = max_depth;
  m->max_stack = max_stack;
  m->max_stack4 = max_stack4;
  m->n_ref_nodes = maps.base;
  m->n_ref_leaves = maps.leaves.size();
  m->n_four_map = maps.four_map.size();
  m->root_base = root_base;
  m->root_nodes = root_nodes;

  auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
      *dst = nullptr;
      return set_error(MBVH_ERR_OOM, "mbvh_mesh_upload: device allocation of %zu bytes failed", bytes);
    }
    if (bytes && src) {
      e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) return set_error(MBVH_ERR_HIP, "mbvh_mesh_upload: copy failed (%s)", hipGetErrorString(e));
    }
    return MBVH_OK;
  };
  rc = upload(reinterpret_cast<void**>(&m->d_nodes), wide.data(), wide.size() * sizeof(WideNode));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_roots), root_refs.data(), root_refs.size() * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_root_slots), root_slots.data(), root_slots.size() * sizeof(uint2));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_bigleaf), bigleaf.data(), bigleaf.size() * sizeof(uint2));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_tris), tris.data(), tris.size() * sizeof(float4));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_vidx), vidx.data(), vidx.size() * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_nodes4), nodes4.data(), nodes4.size() * sizeof(float4));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_roots4), roots4.data(), roots4.size() * sizeof(uint32_t));
  // refit tables: the reference nodes as uploaded, parents, box places, leaves, arrival scratch
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_ref_nodes), nullptr, maps.base * kBytesPerNode);
  for (uint32_t r = 0; !rc && r < n_roots; r++) {
    const hipError_t e = hipMemcpy(reinterpret_cast<uint8_t*>(m->d_ref_nodes) + root_base[r] * kBytesPerNode, roots[r],
                                   root_bytes[r], hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = set_error(MBVH_ERR_HIP, "mbvh_mesh_upload: copy failed (%s)", hipGetErrorString(e));
  }
  std::vector<uint32_t> levels;
  m->level_off.push_back(0);
  for (const auto& lv : maps.internal_by_depth) {
    levels.insert(levels.end(), lv.begin(), lv.end());
    m->level_off.push_back(levels.size());
  }
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_ref_levels), levels.data(), levels.size() * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_ref_wide), maps.wide_off.data(), maps.wide_off.size() * sizeof(int32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_ref_leaves), maps.leaves.data(), maps.leaves.size() * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_ref_scratch), nullptr, maps.base * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_four_map), maps.four_map.data(), maps.four_map.size() * sizeof(uint2));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_coord_max), &coord_max, sizeof(float));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_root_first), root_base.data(), root_base.size() * sizeof(uint64_t));
  if (!rc && indirect) {
    std::vector<uint32_t> face_slot(face_limit, 0xFFFFFFFFu);
    for (uint64_t sl = 0; sl < n_slots; sl++) face_slot[tris_face(tris, sl)] = static_cast<uint32_t>(sl);
    rc = upload(reinterpret_cast<void**>(&m->d_face_slot), face_slot.data(), face_slot.size() * sizeof(uint32_t));
  }
  if (rc) {
    free_mesh_buffers(m);
    m->magic = 0;
    delete m;
    return rc;
  }
  *out_mesh = m;
  return MBVH_OK;
}

extern "C" void mbvh_mesh_free(mbvh_mesh* mesh) {
  if (!valid_mesh(mesh)) return;
  if (valid_ctx(mesh->ctx)) {
    std::lock_guard<std::mutex> guard(mesh->ctx->lock);
    (void)hipSetDevice(mesh->ctx->device);
    (void)hipStreamSynchronize(mesh->ctx->stream);
    free_mesh_buffers(mesh);
  }
  mesh->magic = 0;
  delete mesh;
}

extern "C" int mbvh_mesh_info(const mbvh_mesh* mesh, uint64_t* tri_count, uint64_t* vertex_count,
                              uint32_t* max_depth) {
  if (!valid_mesh(mesh)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_mesh_info: invalid mesh handle");
  if (tri_count) *tri_count = mesh->n_slots;
  if (vertex_count) *vertex_count = mesh->vertex_count;
  if (max_depth) *max_depth = mesh->max_depth;
  return MBVH_OK;
}

extern "C" int mbvh_mesh_set_side_mask(mbvh_ctx* ctx, mbvh_mesh* mesh, const uint8_t* mask, uint64_t n) {
  clear_error();
  int rc = check_query("mbvh_mesh_set_side_mask", ctx, mesh);
  if (rc) return rc;
  if (!mask || n < mesh->face_limit)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_set_side_mask: need a mask entry for each of %llu faces",
                     static_cast<unsigned long long>(mesh->face_limit));
  std::lock_guard<std::mutex> guard(ctx->lock);
  rc = set_device(ctx);
  if (rc) return rc;
  rc = ensure_buffer(ctx->staging[7], n);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[7].ptr, mask, n, hipMemcpyHostToDevice, ctx->stream));
  MBVH_HIP(launch_apply_side_mask(mesh->d_tris, mesh->n_slots, static_cast<const uint8_t*>(ctx->staging[7].ptr), ctx->stream));
  MBVH_HIP(hipStreamSynchronize(ctx->stream));
  return MBVH_OK;
}

namespace {

int check_refit(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                uint32_t position_stride, uint32_t position_offset) {
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (vertex_count != mesh->vertex_count)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: %llu vertices given, the mesh was uploaded with %llu", fn,
                     static_cast<unsigned long long>(vertex_count), static_cast<unsigned long long>(mesh->vertex_count));
  if (position_stride < position_offset + 3) return set_error(MBVH_ERR_BAD_ARGS, "%s: position stride too small", fn);
  if (!positions && vertex_count) return set_error(MBVH_ERR_BAD_ARGS, "%s: positions is NULL", fn);
  return MBVH_OK;
}

int check_roots_out(const char* fn, const mbvh_mesh* mesh, void* const* roots_out, const uint64_t* root_bytes,
                    uint32_t n_roots) {
  if (!roots_out) return MBVH_OK;
  if (!root_bytes) return set_error(MBVH_ERR_BAD_ARGS, "%s: root sizes are NULL", fn);
  if (n_roots != mesh->n_roots)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: %u root buffers given, the mesh has %u roots", fn, n_roots, mesh->n_roots);
  for (uint32_t r = 0; r < mesh->n_roots; r++)
    if (roots_out[r] && root_bytes[r] != mesh->root_nodes[r] * kBytesPerNode)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: root %u buffer holds %llu bytes, the tree %llu", fn, r,
                       static_cast<unsigned long long>(root_bytes[r]),
                       static_cast<unsigned long long>(mesh->root_nodes[r] * kBytesPerNode));
  return MBVH_OK;
}

int enqueue_refit(mbvh_mesh* mesh, const float* d_positions, uint64_t vertex_count, uint32_t stride, uint32_t offset,
                  hipStream_t s, const uint32_t* d_node_ids = nullptr, uint64_t n_ids = 0) {
  RefitArgs a{};
  a.ref_nodes = mesh->d_ref_nodes;
  a.wide_off = mesh->d_ref_wide;
  a.leaves = mesh->d_ref_leaves;
  a.n_nodes = mesh->n_ref_nodes;
  a.n_leaves = mesh->n_ref_leaves;
  a.changed = mesh->d_ref_scratch;
  a.levels = mesh->d_ref_levels;
  a.level_off = mesh->level_off.data();
  a.n_levels = static_cast<uint32_t>(mesh->level_off.size() - 1);
  a.wide = reinterpret_cast<float*>(mesh->d_nodes);
  a.four = reinterpret_cast<float*>(mesh->d_nodes4);
  a.four_map = mesh->d_four_map;
  a.n_four = mesh->n_four_map;
  a.tris = mesh->d_tris;
  a.vidx = mesh->d_vidx;
  a.n_slots = mesh->n_slots;
  a.positions = d_positions;
  a.n_vertices = vertex_count;
  a.stride = stride;
  a.offset = offset;
  a.coord_bound = mesh->d_coord_max;
  a.node_ids = d_node_ids;
  a.n_ids = n_ids;
  a.root_first = mesh->d_root_first;
  a.n_roots = mesh->n_roots;
  MBVH_HIP(launch_refit(a, s));
  return MBVH_OK;
}

int enqueue_copy_roots(const mbvh_mesh* mesh, void* const* roots_out, const uint64_t* root_bytes, hipStream_t s) {
  if (!roots_out) return MBVH_OK;
  for (uint32_t r = 0; r < mesh->n_roots; r++)
    if (roots_out[r])
      MBVH_HIP(hipMemcpyAsync(roots_out[r], reinterpret_cast<uint8_t*>(mesh->d_ref_nodes) + mesh->root_base[r] * kBytesPerNode,
                              root_bytes[r], hipMemcpyDeviceToHost, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_mesh_refit(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                               uint32_t position_stride, uint32_t position_offset, void* const* roots_out,
                               const uint64_t* root_bytes, uint32_t n_roots) {
  clear_error();
  const char* fn = "mbvh_mesh_refit";
  int rc = check_refit(fn, ctx, mesh, positions, vertex_count, position_stride, position_offset);
  if (!rc) rc = check_roots_out(fn, mesh, roots_out, root_bytes, n_roots);
  if (rc) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  const size_t bytes = vertex_count ? ((vertex_count - 1) * position_stride + position_offset + 3) * sizeof(float) : 0;
  if ((rc = ensure_buffer(ctx->staging[0], bytes))) return rc;
  if (bytes) MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, positions, bytes, hipMemcpyHostToDevice, s));
  if ((rc = enqueue_refit(mesh, static_cast<const float*>(ctx->staging[0].ptr), vertex_count, position_stride,
                          position_offset, s)) ||
      (rc = enqueue_copy_roots(mesh, roots_out, root_bytes, s)))
    return rc;
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

extern "C" int mbvh_mesh_refit_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                                 uint32_t position_stride, uint32_t position_offset, void* stream) {
  clear_error();
  const char* fn = "mbvh_mesh_refit_d";
  int rc = check_refit(fn, ctx, mesh, positions, vertex_count, position_stride, position_offset);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  return enqueue_refit(mesh, positions, vertex_count, position_stride, position_offset,
                       stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

// Partial refit: node ids as shapecast reports them (a node's index over all roots). A NULL id list
// is a full refit (nodeIndices = null); an empty one visits what the reference visits for an empty
// set (which is everything, as neither of a root's children is listed).
extern "C" int mbvh_mesh_refit_nodes(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                                     uint32_t position_stride, uint32_t position_offset, const uint32_t* node_ids,
                                     uint64_t n_ids, void* const* roots_out, const uint64_t* root_bytes,
                                     uint32_t n_roots) {
  clear_error();
  const char* fn = "mbvh_mesh_refit_nodes";
  int rc = check_refit(fn, ctx, mesh, positions, vertex_count, position_stride, position_offset);
  if (!rc) rc = check_roots_out(fn, mesh, roots_out, root_bytes, n_roots);
  if (rc) return rc;
  if (!node_ids && n_ids) return set_error(MBVH_ERR_BAD_ARGS, "%s: node_ids is NULL", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  const size_t bytes = vertex_count ? ((vertex_count - 1) * position_stride + position_offset + 3) * sizeof(float) : 0;
  if ((rc = ensure_buffer(ctx->staging[0], bytes))) return rc;
  if (bytes) MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, positions, bytes, hipMemcpyHostToDevice, s));
  const uint32_t* d_ids = nullptr;
  if (node_ids) {
    if ((rc = ensure_buffer(ctx->staging[1], (n_ids ? n_ids : 1) * sizeof(uint32_t)))) return rc;
    if (n_ids) MBVH_HIP(hipMemcpyAsync(ctx->staging[1].ptr, node_ids, n_ids * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    d_ids = static_cast<const uint32_t*>(ctx->staging[1].ptr);
  }
  if ((rc = enqueue_refit(mesh, static_cast<const float*>(ctx->staging[0].ptr), vertex_count, position_stride,
                          position_offset, s, d_ids, n_ids)) ||
      (rc = enqueue_copy_roots(mesh, roots_out, root_bytes, s)))
    return rc;
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

extern "C" int mbvh_mesh_refit_nodes_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions, uint64_t vertex_count,
                                       uint32_t position_stride, uint32_t position_offset, const uint32_t* node_ids,
                                       uint64_t n_ids, void* stream) {
  clear_error();
  const char* fn = "mbvh_mesh_refit_nodes_d";
  int rc = check_refit(fn, ctx, mesh, positions, vertex_count, position_stride, position_offset);
  if (rc) return rc;
  if (!node_ids && n_ids) return set_error(MBVH_ERR_BAD_ARGS, "%s: node_ids is NULL", fn);
  if ((rc = set_device(ctx))) return rc;
  return enqueue_refit(mesh, positions, vertex_count, position_stride, position_offset,
                       stream ? static_cast<hipStream_t>(stream) : ctx->stream, node_ids, n_ids);
}

// ---- imports: the node words of the WebGPU / WebGL packings back to reference node buffers ----

namespace {
constexpr uint32_t kLeafWord = 0xFFFF0000u;  // IS_LEAF: the upper half-word of word 7 set

int import_roots(const char* fn, std::vector<std::vector<uint32_t>>& roots, mbvh_ctx* ctx, const void* index,
                 int32_t index_elt_bytes, uint64_t index_count, const void* indirect, int32_t indirect_elt_bytes,
                 uint64_t indirect_count, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                 uint32_t position_offset, mbvh_mesh** out_mesh) {
  if (roots.empty()) return set_error(MBVH_ERR_BAD_ARGS, "%s: no nodes", fn);
  std::vector<const void*> ptrs;
  std::vector<uint64_t> bytes;
  for (const auto& r : roots) {
    ptrs.push_back(r.data());
    bytes.push_back(r.size() * sizeof(uint32_t));
  }
  return mbvh_mesh_upload(ctx, ptrs.data(), bytes.data(), static_cast<uint32_t>(roots.size()), index, index_elt_bytes,
                          index_count, indirect, indirect_elt_bytes, indirect_count, positions, vertex_count,
                          position_stride, position_offset, out_mesh);
}
}  // namespace

extern "C" int mbvh_mesh_import_nodes(mbvh_ctx* ctx, const uint32_t* nodes, uint64_t n_nodes, uint32_t geometry_offset,
                                      const void* index, int32_t index_elt_bytes, uint64_t index_count,
                                      const void* indirect, int32_t indirect_elt_bytes, uint64_t indirect_count,
                                      const float* positions, uint64_t vertex_count, uint32_t position_stride,
                                      uint32_t position_offset, mbvh_mesh** out_mesh) {
  clear_error();
  const char* fn = "mbvh_mesh_import_nodes";
  if (!out_mesh) return set_error(MBVH_ERR_BAD_ARGS, "%s: out_mesh is NULL", fn);
  *out_mesh = nullptr;
  if (!nodes || n_nodes == 0) return set_error(MBVH_ERR_BAD_ARGS, "%s: no nodes", fn);
  std::vector<std::vector<uint32_t>> roots;
  uint64_t start = 0;
  while (start < n_nodes) {
    // a subtree's extent: the right spine's offsets plus its last leaf
    uint64_t j = start, size = 1;
    while ((nodes[8 * j + 7] & kLeafWord) != kLeafWord) {
      const uint64_t r = nodes[8 * j + 6];
      if (r == 0 || j + r >= n_nodes)
        return set_error(MBVH_ERR_BAD_ARGS, "%s: node %llu's right child %llu is outside the %llu nodes", fn,
                         static_cast<unsigned long long>(j), static_cast<unsigned long long>(j + r),
                         static_cast<unsigned long long>(n_nodes));
      j += r;
      size += r;
    }
    std::vector<uint32_t> root(nodes + 8 * start, nodes + 8 * (start + size));
    float b[6];
    std::memcpy(b, root.data(), sizeof(b));
    for (int c = 0; c < 3; c++)
      if (b[c] > b[c + 3]) {  // the packing's [1, -1] for an empty root
        b[c] = std::numeric_limits<float>::infinity();
        b[c + 3] = -std::numeric_limits<float>::infinity();
      }
    std::memcpy(root.data(), b, sizeof(b));
    for (uint64_t k = 0; k < size; k++) {
      uint32_t* w = root.data() + 8 * k;
      if ((w[7] & kLeafWord) == kLeafWord) {
        if (w[6] < geometry_offset)
          return set_error(MBVH_ERR_BAD_ARGS, "%s: node %llu's triangle offset %u is below the geometry offset %u", fn,
                           static_cast<unsigned long long>(start + k), w[6], geometry_offset);
        w[6] -= geometry_offset;
      }
    }
    roots.push_back(std::move(root));
    start += size;
  }
  return import_roots(fn, roots, ctx, index, index_elt_bytes, index_count, indirect, indirect_elt_bytes, indirect_count,
                      positions, vertex_count, position_stride, position_offset, out_mesh);
}

extern "C" int mbvh_mesh_import_textures(mbvh_ctx* ctx, const float* bvh_bounds, const uint32_t* bvh_contents,
                                         uint64_t n_nodes, const void* index, int32_t index_elt_bytes,
                                         uint64_t index_count, const void* indirect, int32_t indirect_elt_bytes,
                                         uint64_t indirect_count, const float* positions, uint64_t vertex_count,
                                         uint32_t position_stride, uint32_t position_offset, mbvh_mesh** out_mesh) {
  clear_error();
  const char* fn = "mbvh_mesh_import_textures";
  if (!out_mesh) return set_error(MBVH_ERR_BAD_ARGS, "%s: out_mesh is NULL", fn);
  *out_mesh = nullptr;
  if (!bvh_bounds || !bvh_contents || n_nodes == 0) return set_error(MBVH_ERR_BAD_ARGS, "%s: no nodes", fn);
  std::vector<std::vector<uint32_t>> roots(1, std::vector<uint32_t>(8 * n_nodes));
  uint32_t* w = roots[0].data();
  for (uint64_t i = 0; i < n_nodes; i++, w += 8) {
    std::memcpy(w, bvh_bounds + 8 * i, 3 * sizeof(float));         // texel 2i: min.xyz
    std::memcpy(w + 3, bvh_bounds + 8 * i + 4, 3 * sizeof(float));  // texel 2i + 1: max.xyz
    const uint32_t c0 = bvh_contents[2 * i], c1 = bvh_contents[2 * i + 1];
    w[6] = c1;
    w[7] = (c0 & kLeafWord) == kLeafWord ? (kLeafWord | (c0 & 0xFFFFu)) : c0;
  }
  return import_roots(fn, roots, ctx, index, index_elt_bytes, index_count, indirect, indirect_elt_bytes, indirect_count,
                      positions, vertex_count, position_stride, position_offset, out_mesh);
}

extern "C" int mbvh_mesh_copy_roots(mbvh_ctx* ctx, const mbvh_mesh* mesh, void* const* roots_out,
                                    const uint64_t* root_bytes, uint32_t n_roots) {
  clear_error();
  const char* fn = "mbvh_mesh_copy_roots";
  int rc = check_query(fn, ctx, mesh);
  if (!rc && !roots_out) rc = set_error(MBVH_ERR_BAD_ARGS, "%s: roots_out is NULL", fn);
  if (!rc) rc = check_roots_out(fn, mesh, roots_out, root_bytes, n_roots);
  if (rc) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  MBVH_HIP(hipDeviceSynchronize());  // a refit may be in flight on another stream
  if ((rc = enqueue_copy_roots(mesh, roots_out, root_bytes, ctx->stream))) return rc;
  MBVH_HIP(hipStreamSynchronize(ctx->stream));
  return MBVH_OK;
}

// ---------------------------------------------------------------------------------------------
// Queries
// ---------------------------------------------------------------------------------------------

namespace {

// Double-precision inputs travel as float pairs hi = fl(x), lo = fl(x - hi): the kernels trace the hi
// parts with error bounds widened by the lo parts and decide in double-float with both. These are
// the lo parts (device arrays of origins / points and directions, and the lo parts of near / far or
// of the thresholds); all NULL / 0 for FP32 inputs.
struct LoInputs {
  const float* a = nullptr;
  const float* b = nullptr;
  float near_lo = 0.0f, far_lo = 0.0f;
};

void split_f64(double x, float& hi, float& lo) {
  hi = static_cast<float>(x);
  lo = std::isfinite(x) && std::isfinite(hi) ? static_cast<float>(x - static_cast<double>(hi)) : 0.0f;
}
void split_f64(const double* x, size_t n, std::vector<float>& hi, std::vector<float>& lo) {
  hi.resize(n);
  lo.resize(n);
  for (size_t i = 0; i < n; i++) split_f64(x[i], hi[i], lo[i]);
}

int raycast_first_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* o, const float* d,
                       uint64_t n, const uint32_t* order, int32_t side, float near_dist, float far_dist, uint32_t* face,
                       float* dist, float* point, float* bary, hipStream_t s, const LoInputs& lo = LoInputs{}) {
  if (side < MBVH_FRONT_SIDE || side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "raycast_first: invalid side %d", side);
  if (n && (!o || !d || !face)) return set_error(MBVH_ERR_BAD_ARGS, "raycast_first: NULL ray or face buffer");
  // four lanes per ray (4-wide tree) when all rays fit in one resident round: a small launch is
  // latency-bound and gains from shorter paths and more waves; big launches keep one lane per ray
  const bool quads = ctx->ray_lanes == 4 || (ctx->ray_lanes == 0 && 4 * n <= ctx->max_threads);
  if (quads) {
    if (mesh->max_stack4 > kQuadLdsStack + ctx->stack_ovf_entries)
      return set_error(MBVH_ERR_LIMIT, "raycast_first: stack scratch smaller than the tree depth");
  } else if (mesh->max_stack > kStackLds + ctx->stack_ovf_entries) {
    return set_error(MBVH_ERR_LIMIT, "raycast_first: stack scratch smaller than the tree depth");
  }
  const uint64_t threads = quads ? launch_threads(4 * n, ctx->max_threads) : launch_threads(n, ctx->max_threads);
  RaycastFirstArgs a{};
  a.mesh = view_of(mesh);
  a.stack = stack_of(ctx, quads ? threads / 4 : threads);
  a.origins = o;
  a.directions = d;
  a.origins_lo = lo.a;
  a.directions_lo = lo.b;
  a.order = order;
  a.n = n;
  a.side_mask = side_mask_of(side);
  a.near_dist = near_dist;
  a.far_dist = far_dist;
  a.near_lo = lo.near_lo;
  a.far_lo = lo.far_lo;
  a.out_face = face;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  if (quads)
    MBVH_HIP(launch_raycast_first4(a, threads / 4, s));
  else
    MBVH_HIP(launch_raycast_first(a, threads, s));
  return MBVH_OK;
}

// A caller-provided region for a query launch's work-queue counter (256 bytes) and traversal stack
// overflow, used instead of the context's buffers. The launch is narrowed (the kernels are
// persistent: fewer workers pull the same tiles) until every worker's overflow rows fit.
struct FixedScratch {
  uint32_t* queue;
  uint2* ovf;
  uint64_t ovf_bytes;
};

int closest_point_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* q, uint64_t n,
                       const uint32_t* order, float min_t,
                       float max_t, uint32_t* face, float* dist, float* point, float* bary, hipStream_t s,
                       const FixedScratch* fixed = nullptr, const LoInputs& lo = LoInputs{}) {
  if (n && (!q || !face)) return set_error(MBVH_ERR_BAD_ARGS, "closest_point: NULL point or face buffer");
  // packets and quads only without a minThreshold (its early exit follows the binary order); packets
  // only where the tree's 4-wide stack fits the packet kernel's
  const bool packet = !fixed && min_t == 0.0f && ctx->point_lanes == 64 && mesh->max_stack4 <= kPacketStackEntries &&
                      ctx->packet[3].ptr && mesh->max_stack4 <= kQuadLdsStack + ctx->stack_ovf_entries;
  const bool quads = min_t == 0.0f && !packet && (ctx->point_lanes == 4 || ctx->point_lanes == 0 || ctx->point_lanes == 64);
  uint64_t threads = quads ? launch_threads(4 * n, ctx->max_threads) : launch_threads(n, ctx->max_threads);
  ClosestPointArgs a{};
  if (fixed) {
    const uint32_t need = packet ? 0
                          : quads ? (mesh->max_stack4 > kQuadLdsStack ? mesh->max_stack4 - kQuadLdsStack : 0)
                                  : (mesh->max_stack > kStackLds ? mesh->max_stack - kStackLds : 0);
    const uint64_t per_thread = quads ? 4 : 1;  // threads per stack (one stack per quad)
    if (need) {
      const uint64_t stacks = fixed->ovf_bytes / (uint64_t{need} * sizeof(uint2));
      const uint64_t fit = stacks * per_thread / kBlock * kBlock;
      if (fit == 0)
        return set_error(MBVH_ERR_LIMIT, "closest_point: the tree (stack of %u entries) is too deep for %llu bytes of "
                         "stack scratch", quads ? mesh->max_stack4 : mesh->max_stack,
                         static_cast<unsigned long long>(fixed->ovf_bytes));
      threads = std::min(threads, fit);
    }
    a.stack.ovf = fixed->ovf;
    a.stack.ovf_stride = threads / per_thread;
    a.stack.ovf_entries = need;
    a.queue = fixed->queue;
  } else {
    if (!packet && (quads ? mesh->max_stack4 > kQuadLdsStack + ctx->stack_ovf_entries
                          : mesh->max_stack > kStackLds + ctx->stack_ovf_entries))
      return set_error(MBVH_ERR_LIMIT, "closest_point: stack scratch smaller than the tree depth");
    a.stack = stack_of(ctx, quads ? threads / 4 : threads);
    a.queue = static_cast<uint32_t*>(ctx->queue.ptr);
  }
  a.mesh = view_of(mesh);
  a.points = q;
  a.points_lo = lo.a;
  a.min_lo = lo.near_lo;
  a.max_lo = lo.far_lo;
  a.order = order;
  a.n = n;
  a.min_threshold = min_t;
  a.max_threshold = max_t;
  a.out_face = face;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  if (packet) {
    // per chunk: the packets (lists), the exact resolution, then the quad kernel on the queries whose
    // lists overflowed (a device-side count: nothing to do when it is 0)
    a.cand = static_cast<uint2*>(ctx->packet[0].ptr);
    a.cand_meta = static_cast<uint4*>(ctx->packet[1].ptr);
    a.redo_count = static_cast<uint32_t*>(ctx->packet[2].ptr);
    a.pool_count = a.redo_count + 64;  // its own 256-byte line
    a.redo = a.redo_count + 128;
    a.pool = static_cast<uint4*>(ctx->packet[3].ptr);
    a.pool_cap = kPacketPool;
    const uint64_t tiles_total = (n + 63) / 64, chunk_tiles = kPacketChunk / 64;
    for (uint64_t t0 = 0; t0 < tiles_total; t0 += chunk_tiles) {
      a.tile0 = t0;
      a.tiles = std::min(chunk_tiles, tiles_total - t0);
      MBVH_HIP(launch_closest_point_packet(a, launch_threads(a.tiles * 64, ctx->max_threads), s));
      ClosestPointArgs r = a;
      // rare (only when the candidate pool runs out): a small grid, whose waves leave at once when
      // the redo count is 0
      const uint64_t rt = launch_threads(4 * std::min<uint64_t>(a.tiles * 64, 4096), ctx->max_threads);
      r.stack = stack_of(ctx, rt / 4);
      r.order = a.redo;
      r.n_dev = a.redo_count;
      MBVH_HIP(launch_closest_point4(r, rt, s));
    }
  } else if (quads)
    MBVH_HIP(launch_closest_point4(a, threads, s));
  else
    MBVH_HIP(launch_closest_point(a, threads, s));
  return MBVH_OK;
}

// Host entry points order queries spatially from this many on (below it a launch has few waves and
// the ordering passes cost more than they save).
constexpr uint64_t kOrderMin = 4096;

// Orders the n packed triples at `p` (device) into ctx->order; returns the order or NULL (small n).
int host_order(mbvh_ctx* ctx, const float* p, uint64_t n, const uint32_t** out) {
  *out = nullptr;
  if (n < kOrderMin || n >= (uint64_t{1} << 32)) return MBVH_OK;
  int rc;
  if ((rc = ensure_buffer(ctx->order, n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->order_ws, order_workspace_bytes(n))))
    return rc;
  MBVH_HIP(launch_order_points(p, n, static_cast<uint32_t*>(ctx->order.ptr), ctx->order_ws.ptr, ctx->stream));
  *out = static_cast<const uint32_t*>(ctx->order.ptr);
  return MBVH_OK;
}

int check_no_overflow(const uint32_t* face, uint64_t n, const char* fn) {
  for (uint64_t i = 0; i < n; i++)
    if (face[i] == MBVH_NO_HIT - 1u)
      return set_error(MBVH_ERR_LIMIT, "%s: traversal stack overflow on query %llu", fn,
                       static_cast<unsigned long long>(i));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_raycast_first_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                                    const float* directions, uint64_t n, const uint32_t* order, int32_t side, float near_dist,
                                    float far_dist, uint32_t* face_index, float* distance, float* point,
                                    float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_first_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_first_impl(ctx, mesh, origins, directions, n, order, side, near_dist, far_dist, face_index,
                            distance, point, barycoord, s);
}

namespace {

// Host-pointer raycastFirst: stages the rays (and the lo parts of double rays), orders them, launches,
// copies the results back and synchronises.
int host_raycast_first(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                       const float* directions, const float* origins_lo, const float* directions_lo, uint64_t n,
                       int32_t side, float near_dist, float far_dist, float near_lo, float far_lo,
                       uint32_t* face_index, float* distance, float* point, float* barycoord) {
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if (n > 0 && n <= kSmallQueries && ctx->small_host) {
    // one pinned block: the kernels read the rays and write the results there, no copies
    float* h = static_cast<float*>(ctx->small_host);
    float* d = static_cast<float*>(ctx->small_dev);
    const uint64_t k = 3 * n;
    std::memcpy(h, origins, v3);
    std::memcpy(h + k, directions, v3);
    LoInputs lo;
    lo.near_lo = near_lo;
    lo.far_lo = far_lo;
    if (origins_lo) {
      std::memcpy(h + 2 * k, origins_lo, v3);
      std::memcpy(h + 3 * k, directions_lo, v3);
      lo.a = d + 2 * k;
      lo.b = d + 3 * k;
    }
    float* out = d + 4 * k;  // face n, distance n, point 3n, barycoord 3n
    rc = raycast_first_impl(ctx, mesh, d, d + k, n, nullptr, side, near_dist, far_dist,
                            reinterpret_cast<uint32_t*>(out), out + n, out + 2 * n, out + 5 * n, ctx->stream, lo);
    if (rc) return rc;
    MBVH_HIP(hipStreamSynchronize(ctx->stream));
    const float* ho = h + 4 * k;
    std::memcpy(face_index, ho, n * sizeof(uint32_t));
    if (distance) std::memcpy(distance, ho + n, n * sizeof(float));
    if (point) std::memcpy(point, ho + 2 * n, v3);
    if (barycoord) std::memcpy(barycoord, ho + 5 * n, v3);
    return check_no_overflow(face_index, n, fn);
  }
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[1], v3)) ||
      (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->staging[3], n * sizeof(float))) || (rc = ensure_buffer(ctx->staging[4], v3)) ||
      (rc = ensure_buffer(ctx->staging[5], v3)))
    return rc;
  hipStream_t s = ctx->stream;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, origins, v3, hipMemcpyHostToDevice, s));
  MBVH_HIP(hipMemcpyAsync(ctx->staging[1].ptr, directions, v3, hipMemcpyHostToDevice, s));
  LoInputs lo;
  lo.near_lo = near_lo;
  lo.far_lo = far_lo;
  if (origins_lo) {
    if ((rc = ensure_buffer(ctx->lo_stage[0], v3)) || (rc = ensure_buffer(ctx->lo_stage[1], v3))) return rc;
    MBVH_HIP(hipMemcpyAsync(ctx->lo_stage[0].ptr, origins_lo, v3, hipMemcpyHostToDevice, s));
    MBVH_HIP(hipMemcpyAsync(ctx->lo_stage[1].ptr, directions_lo, v3, hipMemcpyHostToDevice, s));
    lo.a = static_cast<const float*>(ctx->lo_stage[0].ptr);
    lo.b = static_cast<const float*>(ctx->lo_stage[1].ptr);
  }
  const uint32_t* order = nullptr;
  if ((rc = host_order(ctx, static_cast<float*>(ctx->staging[0].ptr), n, &order))) return rc;
  rc = raycast_first_impl(ctx, mesh, static_cast<float*>(ctx->staging[0].ptr),
                          static_cast<float*>(ctx->staging[1].ptr), n, order, side, near_dist, far_dist,
                          static_cast<uint32_t*>(ctx->staging[2].ptr), static_cast<float*>(ctx->staging[3].ptr),
                          static_cast<float*>(ctx->staging[4].ptr), static_cast<float*>(ctx->staging[5].ptr), s, lo);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, ctx->staging[3].ptr, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, ctx->staging[4].ptr, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, ctx->staging[5].ptr, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, fn);
}

}  // namespace

extern "C" int mbvh_raycast_first(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                                  const float* directions, uint64_t n, int32_t side, float near_dist,
                                  float far_dist, uint32_t* face_index, float* distance, float* point,
                                  float* barycoord) {
  clear_error();
  int rc = check_query("mbvh_raycast_first", ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!origins || !directions || !face_index)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast_first: NULL ray or face buffer");
  return host_raycast_first("mbvh_raycast_first", ctx, mesh, origins, directions, nullptr, nullptr, n, side, near_dist,
                            far_dist, 0.0f, 0.0f, face_index, distance, point, barycoord);
}

extern "C" int mbvh_raycast_first_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* origins,
                                      const double* directions, uint64_t n, int32_t side, double near_dist,
                                      double far_dist, uint32_t* face_index, float* distance, float* point,
                                      float* barycoord) {
  clear_error();
  const char* fn = "mbvh_raycast_first_f64";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!origins || !directions || !face_index) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL ray or face buffer", fn);
  std::vector<float> oh, ol, dh, dl;
  try {
    split_f64(origins, 3 * n, oh, ol);
    split_f64(directions, 3 * n, dh, dl);
  } catch (...) {
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  float nh, nl, fh, fl;
  split_f64(near_dist, nh, nl);
  split_f64(far_dist, fh, fl);
  return host_raycast_first(fn, ctx, mesh, oh.data(), dh.data(), ol.data(), dl.data(), n, side, nh, fh, nl, fl,
                            face_index, distance, point, barycoord);
}

extern "C" int mbvh_raycast_first_ff_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                                       const float* origins_lo, const float* directions, const float* directions_lo,
                                       uint64_t n, const uint32_t* order, int32_t side, double near_dist,
                                       double far_dist, uint32_t* face_index, float* distance, float* point,
                                       float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_first_ff_d", ctx, mesh);
  if (rc) return rc;
  if (n && (!origins_lo || !directions_lo))
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast_first_ff_d: NULL lo parts");
  rc = set_device(ctx);
  if (rc) return rc;
  LoInputs lo;
  lo.a = origins_lo;
  lo.b = directions_lo;
  float nh, fh;
  split_f64(near_dist, nh, lo.near_lo);
  split_f64(far_dist, fh, lo.far_lo);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_first_impl(ctx, mesh, origins, directions, n, order, side, nh, fh, face_index, distance, point,
                            barycoord, s, lo);
}

extern "C" int mbvh_closest_point_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                    const uint32_t* order, float min_threshold, float max_threshold, uint32_t* face_index,
                                    float* distance, float* closest, float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_closest_point_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return closest_point_impl(ctx, mesh, points, n, order, min_threshold, max_threshold, face_index, distance,
                            closest, barycoord, s);
}

namespace {

// Regions of `scratch`, at fixed offsets: queue counter | order (4 B per query) | ordering
// workspace for the largest n | traversal stack overflow (the rest).
struct NoAllocLayout {
  uint64_t order_off, ws_off, ovf_off;
};

NoAllocLayout noalloc_layout() {
  auto up = [](uint64_t x) { return (x + 255) / 256 * 256; };
  NoAllocLayout l;
  l.order_off = 256;
  l.ws_off = up(l.order_off + 4 * uint64_t{MBVH_NOALLOC_MAX_QUERIES});
  l.ovf_off = up(l.ws_off + order_workspace_bytes(MBVH_NOALLOC_MAX_QUERIES));
  return l;
}

}  // namespace

extern "C" int mbvh_closest_point_noalloc_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                            float min_threshold, float max_threshold, uint32_t* face_index,
                                            float* distance, float* closest, float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_closest_point_noalloc_d", ctx, mesh);
  if (rc) return rc;
  if (n > MBVH_NOALLOC_MAX_QUERIES)
    return set_error(MBVH_ERR_LIMIT, "mbvh_closest_point_noalloc_d: %llu queries exceed the allocation-free maximum "
                     "of %llu", static_cast<unsigned long long>(n),
                     static_cast<unsigned long long>(MBVH_NOALLOC_MAX_QUERIES));
  if (n && (!points || !face_index))
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_closest_point_noalloc_d: NULL point or face buffer");
  char* base = static_cast<char*>(scratch_of(ctx->device));
  if (!base)
    return set_error(MBVH_ERR_NO_DEVICE, "mbvh_closest_point_noalloc_d: no scratch on device %d", ctx->device);
  const NoAllocLayout l = noalloc_layout();
  if (l.ovf_off >= MBVH_NOALLOC_SCRATCH_BYTES)
    return set_error(MBVH_ERR_LIMIT, "mbvh_closest_point_noalloc_d: scratch layout exceeds its size");
  if ((rc = set_device(ctx))) return rc;
  if (n == 0) return MBVH_OK;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  // the host entry points' spatial order (results are written at each query's own index)
  const uint32_t* order = nullptr;
  if (n >= kOrderMin) {
    uint32_t* o = reinterpret_cast<uint32_t*>(base + l.order_off);
    MBVH_HIP(launch_order_points(points, n, o, base + l.ws_off, s));
    order = o;
  }
  FixedScratch fx;
  fx.queue = reinterpret_cast<uint32_t*>(base);
  fx.ovf = reinterpret_cast<uint2*>(base + l.ovf_off);
  fx.ovf_bytes = MBVH_NOALLOC_SCRATCH_BYTES - l.ovf_off;
  return closest_point_impl(ctx, mesh, points, n, order, min_threshold, max_threshold, face_index, distance,
                            closest, barycoord, s, &fx);
}

namespace {

// Host-pointer closest point: stages the points (and the lo parts of double points), orders them,
// launches, copies the results back and synchronises.
int host_closest_point(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, const float* points_lo,
                       uint64_t n, float min_threshold, float max_threshold, float min_lo, float max_lo,
                       uint32_t* face_index, float* distance, float* closest, float* barycoord) {
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if (n > 0 && n <= kSmallQueries && ctx->small_host) {
    // one pinned block: the kernels read the points and write the results there, no copies
    float* h = static_cast<float*>(ctx->small_host);
    float* d = static_cast<float*>(ctx->small_dev);
    const uint64_t k = 3 * n;
    std::memcpy(h, points, v3);
    LoInputs lo;
    lo.near_lo = min_lo;
    lo.far_lo = max_lo;
    if (points_lo) {
      std::memcpy(h + k, points_lo, v3);
      lo.a = d + k;
    }
    float* out = d + 4 * k;  // face n, distance n, point 3n, barycoord 3n
    rc = closest_point_impl(ctx, mesh, d, n, nullptr, min_threshold, max_threshold, reinterpret_cast<uint32_t*>(out),
                            out + n, out + 2 * n, out + 5 * n, ctx->stream, nullptr, lo);
    if (rc) return rc;
    MBVH_HIP(hipStreamSynchronize(ctx->stream));
    const float* ho = h + 4 * k;
    std::memcpy(face_index, ho, n * sizeof(uint32_t));
    if (distance) std::memcpy(distance, ho + n, n * sizeof(float));
    if (closest) std::memcpy(closest, ho + 2 * n, v3);
    if (barycoord) std::memcpy(barycoord, ho + 5 * n, v3);
    return check_no_overflow(face_index, n, fn);
  }
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->staging[3], n * sizeof(float))) || (rc = ensure_buffer(ctx->staging[4], v3)) ||
      (rc = ensure_buffer(ctx->staging[5], v3)))
    return rc;
  hipStream_t s = ctx->stream;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, points, v3, hipMemcpyHostToDevice, s));
  LoInputs lo;
  lo.near_lo = min_lo;
  lo.far_lo = max_lo;
  if (points_lo) {
    if ((rc = ensure_buffer(ctx->lo_stage[0], v3))) return rc;
    MBVH_HIP(hipMemcpyAsync(ctx->lo_stage[0].ptr, points_lo, v3, hipMemcpyHostToDevice, s));
    lo.a = static_cast<const float*>(ctx->lo_stage[0].ptr);
  }
  const uint32_t* order = nullptr;
  if ((rc = host_order(ctx, static_cast<float*>(ctx->staging[0].ptr), n, &order))) return rc;
  rc = closest_point_impl(ctx, mesh, static_cast<float*>(ctx->staging[0].ptr), n, order, min_threshold, max_threshold,
                          static_cast<uint32_t*>(ctx->staging[2].ptr), static_cast<float*>(ctx->staging[3].ptr),
                          static_cast<float*>(ctx->staging[4].ptr), static_cast<float*>(ctx->staging[5].ptr), s,
                          nullptr, lo);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, ctx->staging[3].ptr, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (closest) MBVH_HIP(hipMemcpyAsync(closest, ctx->staging[4].ptr, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, ctx->staging[5].ptr, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, fn);
}

}  // namespace

extern "C" int mbvh_closest_point(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                  float min_threshold, float max_threshold, uint32_t* face_index,
                                  float* distance, float* closest, float* barycoord) {
  clear_error();
  int rc = check_query("mbvh_closest_point", ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!points || !face_index) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_closest_point: NULL point or face buffer");
  return host_closest_point("mbvh_closest_point", ctx, mesh, points, nullptr, n, min_threshold, max_threshold, 0.0f,
                            0.0f, face_index, distance, closest, barycoord);
}

extern "C" int mbvh_closest_point_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* points, uint64_t n,
                                      double min_threshold, double max_threshold, uint32_t* face_index,
                                      float* distance, float* closest, float* barycoord) {
  clear_error();
  const char* fn = "mbvh_closest_point_f64";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!points || !face_index) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL point or face buffer", fn);
  std::vector<float> ph, pl;
  try {
    split_f64(points, 3 * n, ph, pl);
  } catch (...) {
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  float mh, ml, xh, xl;
  split_f64(min_threshold, mh, ml);
  split_f64(max_threshold, xh, xl);
  return host_closest_point(fn, ctx, mesh, ph.data(), pl.data(), n, mh, xh, ml, xl, face_index, distance, closest,
                            barycoord);
}

extern "C" int mbvh_closest_point_ff_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, const float* points_lo,
                                       uint64_t n, const uint32_t* order, double min_threshold, double max_threshold,
                                       uint32_t* face_index, float* distance, float* closest, float* barycoord,
                                       void* stream) {
  clear_error();
  int rc = check_query("mbvh_closest_point_ff_d", ctx, mesh);
  if (rc) return rc;
  if (n && !points_lo) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_closest_point_ff_d: NULL lo parts");
  rc = set_device(ctx);
  if (rc) return rc;
  LoInputs lo;
  lo.a = points_lo;
  float mh, xh;
  split_f64(min_threshold, mh, lo.near_lo);
  split_f64(max_threshold, xh, lo.far_lo);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return closest_point_impl(ctx, mesh, points, n, order, mh, xh, face_index, distance, closest, barycoord, s, nullptr,
                            lo);
}

extern "C" int mbvh_intersects_sphere(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* centers, const float* radii,
                                     uint64_t n, uint8_t* out) {
  clear_error();
  const char* fn = "mbvh_intersects_sphere";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!centers || !radii || !out) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL centre, radius or output buffer", fn);
  // the closest-point search runs with the largest radius (slightly enlarged: its pruning compares
  // rounded squares) as maxThreshold; the exact decision is per sphere
  float r_max = -1.0f;
  for (uint64_t i = 0; i < n; i++)
    if (radii[i] > r_max) r_max = radii[i];
  if (!(r_max >= 0.0f)) {
    std::memset(out, 0, n);
    return MBVH_OK;
  }
  const float max_t = r_max * (1.0f + 1.0e-6f) + std::numeric_limits<float>::min();
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[1], n * sizeof(float))) ||
      (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) || (rc = ensure_buffer(ctx->staging[7], n)))
    return rc;
  hipStream_t s = ctx->stream;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, centers, v3, hipMemcpyHostToDevice, s));
  MBVH_HIP(hipMemcpyAsync(ctx->staging[1].ptr, radii, n * sizeof(float), hipMemcpyHostToDevice, s));
  const uint32_t* order = nullptr;
  if ((rc = host_order(ctx, static_cast<float*>(ctx->staging[0].ptr), n, &order))) return rc;
  if ((rc = closest_point_impl(ctx, mesh, static_cast<float*>(ctx->staging[0].ptr), n, order, 0.0f, max_t,
                               static_cast<uint32_t*>(ctx->staging[2].ptr), nullptr, nullptr, nullptr, s)))
    return rc;
  SphereArgs a{};
  a.mesh = grad_view_of(mesh);
  a.centers = static_cast<const float*>(ctx->staging[0].ptr);
  a.radii = static_cast<const float*>(ctx->staging[1].ptr);
  a.face = static_cast<const uint32_t*>(ctx->staging[2].ptr);
  a.n = n;
  a.out = static_cast<uint8_t*>(ctx->staging[7].ptr);
  MBVH_HIP(launch_sphere_decide(a, s));
  MBVH_HIP(hipMemcpyAsync(out, ctx->staging[7].ptr, n, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < n; i++)
    if (out[i] == 2)
      return set_error(MBVH_ERR_LIMIT, "%s: traversal stack overflow on sphere %llu", fn, static_cast<unsigned long long>(i));
  return MBVH_OK;
}

// ---------------------------------------------------------------------------------------------
// All hits
// ---------------------------------------------------------------------------------------------

namespace {

int raycast_all_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* o, const float* d, uint64_t n,
                     int32_t side, float near_dist, float far_dist, uint32_t* counts, const uint64_t* offsets,
                     uint64_t capacity, uint32_t* face, float* dist, float* point, float* bary, bool fill,
                     hipStream_t s, const LoInputs& lo = LoInputs{}) {
  if (side < MBVH_FRONT_SIDE || side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "raycast: invalid side %d", side);
  if (n && (!o || !d)) return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL ray buffer");
  if (n && !fill && !counts) return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL counts buffer");
  if (n && fill && (!offsets || (capacity && !face)))
    return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL offsets or face buffer");
  if (mesh->max_stack > kStackLds + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "raycast: stack scratch smaller than the tree depth");
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  RaycastAllArgs a{};
  a.mesh = view_of(mesh);
  a.stack = stack_of(ctx, threads);
  a.origins = o;
  a.directions = d;
  a.origins_lo = lo.a;
  a.directions_lo = lo.b;
  a.n = n;
  a.side_mask = side_mask_of(side);
  a.near_dist = near_dist;
  a.far_dist = far_dist;
  a.near_lo = lo.near_lo;
  a.far_lo = lo.far_lo;
  a.counts = counts;
  a.offsets = offsets;
  a.out_face = face;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  a.capacity = capacity;
  MBVH_HIP(launch_raycast_all(a, fill, threads, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_raycast_count_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                                    const float* directions, uint64_t n, int32_t side, float near_dist,
                                    float far_dist, uint32_t* counts, void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_count_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_all_impl(ctx, mesh, origins, directions, n, side, near_dist, far_dist, counts, nullptr, 0,
                          nullptr, nullptr, nullptr, nullptr, false, s);
}

extern "C" int mbvh_raycast_fill_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins,
                                   const float* directions, uint64_t n, int32_t side, float near_dist,
                                   float far_dist, const uint64_t* offsets, uint64_t capacity,
                                   uint32_t* face_index, float* distance, float* point, float* barycoord,
                                   void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_fill_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_all_impl(ctx, mesh, origins, directions, n, side, near_dist, far_dist, nullptr, offsets,
                          capacity, face_index, distance, point, barycoord, true, s);
}

namespace {

// Host-pointer all-hits raycast: a count pass, the offsets on the host, a fill pass into the
// context's hit buffers (mbvh_raycast_fetch copies them out).
int host_raycast_all(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins, const float* directions,
                     const float* origins_lo, const float* directions_lo, uint64_t n, int32_t side, float near_dist,
                     float far_dist, float near_lo, float far_lo, uint64_t* offsets) {
  std::lock_guard<std::mutex> guard(ctx->lock);
  ctx->hits_total = 0;
  offsets[0] = 0;
  if (n == 0) return MBVH_OK;
  int rc = set_device(ctx);
  if (rc) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[1], v3)) ||
      (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->hits[0], (n + 1) * sizeof(uint64_t))))
    return rc;
  hipStream_t s = ctx->stream;
  const float* dO = static_cast<float*>(ctx->staging[0].ptr);
  const float* dD = static_cast<float*>(ctx->staging[1].ptr);
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, origins, v3, hipMemcpyHostToDevice, s));
  MBVH_HIP(hipMemcpyAsync(ctx->staging[1].ptr, directions, v3, hipMemcpyHostToDevice, s));
  LoInputs lo;
  lo.near_lo = near_lo;
  lo.far_lo = far_lo;
  if (origins_lo) {
    if ((rc = ensure_buffer(ctx->lo_stage[0], v3)) || (rc = ensure_buffer(ctx->lo_stage[1], v3))) return rc;
    MBVH_HIP(hipMemcpyAsync(ctx->lo_stage[0].ptr, origins_lo, v3, hipMemcpyHostToDevice, s));
    MBVH_HIP(hipMemcpyAsync(ctx->lo_stage[1].ptr, directions_lo, v3, hipMemcpyHostToDevice, s));
    lo.a = static_cast<const float*>(ctx->lo_stage[0].ptr);
    lo.b = static_cast<const float*>(ctx->lo_stage[1].ptr);
  }
  rc = raycast_all_impl(ctx, mesh, dO, dD, n, side, near_dist, far_dist, static_cast<uint32_t*>(ctx->staging[2].ptr),
                        nullptr, 0, nullptr, nullptr, nullptr, nullptr, false, s, lo);
  if (rc) return rc;
  std::vector<uint32_t> counts(n);
  MBVH_HIP(hipMemcpyAsync(counts.data(), ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (counts[i] == MBVH_NO_HIT - 1u)
      return set_error(MBVH_ERR_LIMIT, "%s: traversal stack overflow on ray %llu", fn, static_cast<unsigned long long>(i));
    offsets[i] = total;
    total += counts[i];
  }
  offsets[n] = total;
  const size_t t = static_cast<size_t>(total);
  if ((rc = ensure_buffer(ctx->hits[1], t * sizeof(uint32_t))) || (rc = ensure_buffer(ctx->hits[2], t * sizeof(float))) ||
      (rc = ensure_buffer(ctx->hits[3], t * 3 * sizeof(float))) || (rc = ensure_buffer(ctx->hits[4], t * 3 * sizeof(float))))
    return rc;
  MBVH_HIP(hipMemcpyAsync(ctx->hits[0].ptr, offsets, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  rc = raycast_all_impl(ctx, mesh, dO, dD, n, side, near_dist, far_dist, nullptr,
                        static_cast<const uint64_t*>(ctx->hits[0].ptr), total, static_cast<uint32_t*>(ctx->hits[1].ptr),
                        static_cast<float*>(ctx->hits[2].ptr), static_cast<float*>(ctx->hits[3].ptr),
                        static_cast<float*>(ctx->hits[4].ptr), true, s, lo);
  if (rc) return rc;
  MBVH_HIP(hipStreamSynchronize(s));
  ctx->hits_total = total;
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_raycast(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins, const float* directions,
                            uint64_t n, int32_t side, float near_dist, float far_dist, uint64_t* offsets) {
  clear_error();
  int rc = check_query("mbvh_raycast", ctx, mesh);
  if (rc) return rc;
  if (!offsets) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast: offsets is NULL");
  if (n && (!origins || !directions)) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast: NULL ray buffer");
  return host_raycast_all("mbvh_raycast", ctx, mesh, origins, directions, nullptr, nullptr, n, side, near_dist,
                          far_dist, 0.0f, 0.0f, offsets);
}

extern "C" int mbvh_raycast_f64(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* origins, const double* directions,
                                uint64_t n, int32_t side, double near_dist, double far_dist, uint64_t* offsets) {
  clear_error();
  const char* fn = "mbvh_raycast_f64";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (!offsets) return set_error(MBVH_ERR_BAD_ARGS, "%s: offsets is NULL", fn);
  if (n && (!origins || !directions)) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL ray buffer", fn);
  std::vector<float> oh, ol, dh, dl;
  try {
    split_f64(origins, 3 * n, oh, ol);
    split_f64(directions, 3 * n, dh, dl);
  } catch (...) {
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  float nh, nl, fh, fl;
  split_f64(near_dist, nh, nl);
  split_f64(far_dist, fh, fl);
  return host_raycast_all(fn, ctx, mesh, oh.data(), dh.data(), ol.data(), dl.data(), n, side, nh, fh, nl, fl, offsets);
}

extern "C" int mbvh_raycast_fetch(mbvh_ctx* ctx, uint32_t* face_index, float* distance, float* point,
                                  float* barycoord, uint64_t total) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_raycast_fetch: invalid context handle");
  std::lock_guard<std::mutex> guard(ctx->lock);
  if (total != ctx->hits_total)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast_fetch: %llu hits are held, not %llu",
                     static_cast<unsigned long long>(ctx->hits_total), static_cast<unsigned long long>(total));
  if (total == 0) return MBVH_OK;
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const size_t t = static_cast<size_t>(total);
  if (face_index) MBVH_HIP(hipMemcpyAsync(face_index, ctx->hits[1].ptr, t * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, ctx->hits[2].ptr, t * sizeof(float), hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, ctx->hits[3].ptr, t * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, ctx->hits[4].ptr, t * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

// ---------------------------------------------------------------------------------------------
// Query ordering
// ---------------------------------------------------------------------------------------------

extern "C" int mbvh_order_workspace_bytes(uint64_t n, uint64_t* bytes) {
  clear_error();
  if (!bytes) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_order_workspace_bytes: NULL output");
  *bytes = order_workspace_bytes(n);
  return MBVH_OK;
}

extern "C" int mbvh_order_points_d(mbvh_ctx* ctx, const float* points, uint64_t n, uint32_t* order,
                                   void* workspace, uint64_t workspace_bytes, void* stream) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_order_points_d: invalid context");
  if (n == 0) return MBVH_OK;
  if (n >= (uint64_t{1} << 32)) return set_error(MBVH_ERR_LIMIT, "mbvh_order_points_d: n must be below 2^32");
  if (!points || !order || !workspace)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_order_points_d: NULL points, order or workspace");
  if (workspace_bytes < order_workspace_bytes(n))
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_order_points_d: workspace of %llu bytes, %llu needed",
                     static_cast<unsigned long long>(workspace_bytes),
                     static_cast<unsigned long long>(order_workspace_bytes(n)));
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_order_points_d: workspace must be 256-byte aligned");
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  MBVH_HIP(launch_order_points(points, n, order, workspace, s));
  return MBVH_OK;
}

// ---------------------------------------------------------------------------------------------
// Backward
// ---------------------------------------------------------------------------------------------

namespace {

int check_mode(const char* fn, int32_t mode) {
  if (mode != MBVH_GRAD_ATOMIC && mode != MBVH_GRAD_DETERMINISTIC)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: mode must be MBVH_GRAD_ATOMIC or MBVH_GRAD_DETERMINISTIC", fn);
  return MBVH_OK;
}

int check_workspace(const char* fn, uint64_t n, int32_t mode, const float* grad_positions, const void* ws,
                    uint64_t ws_bytes) {
  if (!grad_positions || mode != MBVH_GRAD_DETERMINISTIC || n == 0) return MBVH_OK;
  const uint64_t need = grad_workspace_bytes(n, mode);
  if (!ws || ws_bytes < need)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: deterministic mode needs a workspace of %llu bytes", fn,
                     static_cast<unsigned long long>(need));
  if (reinterpret_cast<uintptr_t>(ws) % 256 != 0)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: workspace must be 256-byte aligned", fn);
  return MBVH_OK;
}

int raycast_bwd_impl(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* directions, uint64_t n,
                     const uint32_t* face, const float* distance, const float* bary, const float* g_dist,
                     const float* g_point, const float* g_bary, float* g_origins, float* g_dirs, float* g_pos,
                     int32_t mode, void* ws, uint64_t ws_bytes, hipStream_t s) {
  int rc = check_mode(fn, mode);
  if (rc) return rc;
  if (n && (!directions || !face || !distance || !bary))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL direction or saved forward buffer", fn);
  if (mesh->vertex_count >= 0xFFFFFFFFull) return set_error(MBVH_ERR_LIMIT, "%s: too many vertices", fn);
  if (n >= (uint64_t{1} << 32) / 3) return set_error(MBVH_ERR_LIMIT, "%s: too many queries", fn);
  if ((rc = check_workspace(fn, n, mode, g_pos, ws, ws_bytes))) return rc;
  RaycastBwdArgs a{};
  a.mesh = grad_view_of(mesh);
  a.directions = directions;
  a.face = face;
  a.distance = distance;
  a.barycoord = bary;
  a.grad_distance = g_dist;
  a.grad_point = g_point;
  a.grad_barycoord = g_bary;
  a.grad_origins = g_origins;
  a.grad_directions = g_dirs;
  a.n = n;
  MBVH_HIP(launch_raycast_first_bwd(a, mode, g_pos, mesh->vertex_count, ws, s));
  return MBVH_OK;
}

int closest_bwd_impl(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                     const uint32_t* face, const float* distance, const float* closest, const float* bary,
                     const float* g_dist, float* g_points, float* g_pos, int32_t mode, void* ws,
                     uint64_t ws_bytes, hipStream_t s) {
  (void)ctx;
  int rc = check_mode(fn, mode);
  if (rc) return rc;
  if (n && (!points || !face || !distance || !closest || !bary || !g_dist))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL point, saved forward or upstream buffer", fn);
  if (mesh->vertex_count >= 0xFFFFFFFFull) return set_error(MBVH_ERR_LIMIT, "%s: too many vertices", fn);
  if (n >= (uint64_t{1} << 32) / 3) return set_error(MBVH_ERR_LIMIT, "%s: too many queries", fn);
  if ((rc = check_workspace(fn, n, mode, g_pos, ws, ws_bytes))) return rc;
  ClosestBwdArgs a{};
  a.mesh = grad_view_of(mesh);
  a.points = points;
  a.face = face;
  a.distance = distance;
  a.closest = closest;
  a.barycoord = bary;
  a.grad_distance = g_dist;
  a.grad_points = g_points;
  a.n = n;
  MBVH_HIP(launch_closest_point_bwd(a, mode, g_pos, mesh->vertex_count, ws, s));
  return MBVH_OK;
}

// Copies `bytes` of host data into staging buffer b (or returns NULL for a NULL source).
int stage_in(mbvh_ctx* ctx, DeviceBuffer& b, const void* src, size_t bytes, void** out) {
  *out = nullptr;
  if (!src) return MBVH_OK;
  int rc = ensure_buffer(b, bytes);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  *out = b.ptr;
  return MBVH_OK;
}

int stage_out(DeviceBuffer& b, const void* dst, size_t bytes, void** out) {
  *out = nullptr;
  if (!dst) return MBVH_OK;
  int rc = ensure_buffer(b, bytes);
  if (rc) return rc;
  *out = b.ptr;
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_grad_workspace_bytes(uint64_t n, int32_t mode, uint64_t* bytes) {
  clear_error();
  if (!bytes) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_grad_workspace_bytes: NULL output");
  int rc = check_mode("mbvh_grad_workspace_bytes", mode);
  if (rc) return rc;
  *bytes = grad_workspace_bytes(n, mode);
  return MBVH_OK;
}

extern "C" int mbvh_raycast_first_bwd_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* directions, uint64_t n,
                                        const uint32_t* face_index, const float* distance, const float* barycoord,
                                        const float* grad_distance, const float* grad_point,
                                        const float* grad_barycoord, float* grad_origins, float* grad_directions,
                                        float* grad_positions, int32_t mode, void* workspace,
                                        uint64_t workspace_bytes, void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_first_bwd_d", ctx, mesh);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_bwd_impl("mbvh_raycast_first_bwd_d", ctx, mesh, directions, n, face_index, distance, barycoord,
                          grad_distance, grad_point, grad_barycoord, grad_origins, grad_directions, grad_positions,
                          mode, workspace, workspace_bytes, s);
}

extern "C" int mbvh_raycast_first_bwd(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* directions, uint64_t n,
                                      const uint32_t* face_index, const float* distance, const float* barycoord,
                                      const float* grad_distance, const float* grad_point,
                                      const float* grad_barycoord, float* grad_origins, float* grad_directions,
                                      float* grad_positions, int32_t mode) {
  clear_error();
  const char* fn = "mbvh_raycast_first_bwd";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if ((rc = check_mode(fn, mode))) return rc;
  if (n && (!directions || !face_index || !distance || !barycoord))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL direction or saved forward buffer", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  const size_t v3 = n * 3 * sizeof(float), s1 = n * sizeof(float), vp = mesh->vertex_count * 3 * sizeof(float);
  void *dD, *dF, *dT, *dB, *gT, *gP, *gB, *oO, *oD, *oV;
  if ((rc = stage_in(ctx, ctx->grad[0], directions, v3, &dD)) || (rc = stage_in(ctx, ctx->grad[1], face_index, s1, &dF)) ||
      (rc = stage_in(ctx, ctx->grad[2], distance, s1, &dT)) || (rc = stage_in(ctx, ctx->grad[3], barycoord, v3, &dB)) ||
      (rc = stage_in(ctx, ctx->grad[4], grad_distance, s1, &gT)) ||
      (rc = stage_in(ctx, ctx->grad[5], grad_point, v3, &gP)) ||
      (rc = stage_in(ctx, ctx->grad[6], grad_barycoord, v3, &gB)) ||
      (rc = stage_out(ctx->grad[7], grad_origins, v3, &oO)) || (rc = stage_out(ctx->grad[8], grad_directions, v3, &oD)) ||
      (rc = stage_out(ctx->grad[9], grad_positions, vp, &oV)))
    return rc;
  const uint64_t wsb = grad_positions ? grad_workspace_bytes(n, mode) : 0;
  if (wsb && (rc = ensure_buffer(ctx->grad_ws, wsb))) return rc;
  rc = raycast_bwd_impl(fn, ctx, mesh, static_cast<float*>(dD), n, static_cast<uint32_t*>(dF), static_cast<float*>(dT),
                        static_cast<float*>(dB), static_cast<float*>(gT), static_cast<float*>(gP),
                        static_cast<float*>(gB), static_cast<float*>(oO), static_cast<float*>(oD),
                        static_cast<float*>(oV), mode, ctx->grad_ws.ptr, ctx->grad_ws.bytes, ctx->stream);
  if (rc) return rc;
  if (oO) MBVH_HIP(hipMemcpyAsync(grad_origins, oO, v3, hipMemcpyDeviceToHost, ctx->stream));
  if (oD) MBVH_HIP(hipMemcpyAsync(grad_directions, oD, v3, hipMemcpyDeviceToHost, ctx->stream));
  if (oV) MBVH_HIP(hipMemcpyAsync(grad_positions, oV, vp, hipMemcpyDeviceToHost, ctx->stream));
  MBVH_HIP(hipStreamSynchronize(ctx->stream));
  return MBVH_OK;
}

extern "C" int mbvh_closest_point_bwd_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                        const uint32_t* face_index, const float* distance, const float* closest,
                                        const float* barycoord, const float* grad_distance, float* grad_points,
                                        float* grad_positions, int32_t mode, void* workspace,
                                        uint64_t workspace_bytes, void* stream) {
  clear_error();
  int rc = check_query("mbvh_closest_point_bwd_d", ctx, mesh);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return closest_bwd_impl("mbvh_closest_point_bwd_d", ctx, mesh, points, n, face_index, distance, closest, barycoord,
                          grad_distance, grad_points, grad_positions, mode, workspace, workspace_bytes, s);
}

extern "C" int mbvh_closest_point_bwd(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                      const uint32_t* face_index, const float* distance, const float* closest,
                                      const float* barycoord, const float* grad_distance, float* grad_points,
                                      float* grad_positions, int32_t mode) {
  clear_error();
  const char* fn = "mbvh_closest_point_bwd";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if ((rc = check_mode(fn, mode))) return rc;
  if (n && (!points || !face_index || !distance || !closest || !barycoord || !grad_distance))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL point, saved forward or upstream buffer", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  const size_t v3 = n * 3 * sizeof(float), s1 = n * sizeof(float), vp = mesh->vertex_count * 3 * sizeof(float);
  void *dQ, *dF, *dT, *dC, *dB, *gT, *oQ, *oV;
  if ((rc = stage_in(ctx, ctx->grad[0], points, v3, &dQ)) || (rc = stage_in(ctx, ctx->grad[1], face_index, s1, &dF)) ||
      (rc = stage_in(ctx, ctx->grad[2], distance, s1, &dT)) || (rc = stage_in(ctx, ctx->grad[3], closest, v3, &dC)) ||
      (rc = stage_in(ctx, ctx->grad[4], barycoord, v3, &dB)) ||
      (rc = stage_in(ctx, ctx->grad[5], grad_distance, s1, &gT)) ||
      (rc = stage_out(ctx->grad[7], grad_points, v3, &oQ)) || (rc = stage_out(ctx->grad[9], grad_positions, vp, &oV)))
    return rc;
  const uint64_t wsb = grad_positions ? grad_workspace_bytes(n, mode) : 0;
  if (wsb && (rc = ensure_buffer(ctx->grad_ws, wsb))) return rc;
  rc = closest_bwd_impl(fn, ctx, mesh, static_cast<float*>(dQ), n, static_cast<uint32_t*>(dF), static_cast<float*>(dT),
                        static_cast<float*>(dC), static_cast<float*>(dB), static_cast<float*>(gT),
                        static_cast<float*>(oQ), static_cast<float*>(oV), mode, ctx->grad_ws.ptr, ctx->grad_ws.bytes,
                        ctx->stream);
  if (rc) return rc;
  if (oQ) MBVH_HIP(hipMemcpyAsync(grad_points, oQ, v3, hipMemcpyDeviceToHost, ctx->stream));
  if (oV) MBVH_HIP(hipMemcpyAsync(grad_positions, oV, vp, hipMemcpyDeviceToHost, ctx->stream));
  MBVH_HIP(hipStreamSynchronize(ctx->stream));
  return MBVH_OK;
}

// ---------------------------------------------------------------------------------------------
// Instances and batches
// ---------------------------------------------------------------------------------------------

namespace {

constexpr uint32_t kBatchMagic = 0x48435442u;
constexpr uint32_t kGraphMagic = 0x48505247u;

bool valid_batch(const mbvh_batch* b) { return b && b->magic == kBatchMagic; }

uint64_t batch_ws_bytes(const mbvh_batch* b, uint64_t n) {
  (void)b;
  return order_by_mesh_workspace_bytes(n);
}

int check_tuning(const char* fn, const mbvh_batch_tuning& t) {
  if ((t.ray_lanes != 1 && t.ray_lanes != 4) || (t.stack_lds != 8 && t.stack_lds != 16) || t.tile_rays == 0 ||
      t.tile_rays % 64 != 0 || t.tile_rays > 4096 || t.blocks_per_cu < 1 || t.blocks_per_cu > 8)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: invalid tuning (ray_lanes 1|4, stack_lds 8|16, tile_rays 64..4096 "
                     "step 64, blocks_per_cu 1..8)", fn);
  return MBVH_OK;
}

struct BatchCall {
  const float* world;
  const float* inverse;
  const uint32_t* inst_geom;
  uint32_t n_inst;
  const float* origins;
  const float* dirs;
  const uint32_t* ray_inst;
  uint64_t n;
  int32_t side;
  float near_d, far_d;
  uint32_t* face;
  float* dist;
  float* point;
  float* bary;
  void* ws;
  uint64_t ws_bytes;
  // local mode (windows != NULL): rays in their mesh's frame as hi + lo pairs, ray_inst naming the mesh,
  // a window (near, near lo, far, far lo) per ray; no matrices or instance table
  const float* origins_lo;
  const float* dirs_lo;
  const float* windows;
};

int check_batch_call(const char* fn, const mbvh_ctx* ctx, const mbvh_batch* b, const BatchCall& c) {
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!valid_batch(b) || b->ctx != ctx) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid batch handle", fn);
  if (c.side < MBVH_FRONT_SIDE || c.side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: invalid side %d", fn, c.side);
  if (c.n >= (uint64_t{1} << 32)) return set_error(MBVH_ERR_LIMIT, "%s: n must be below 2^32", fn);
  if (c.n == 0) return MBVH_OK;
  const bool local = c.windows != nullptr;
  if ((!local && (!c.world || !c.inverse || !c.inst_geom)) || !c.origins || !c.dirs || !c.ray_inst || !c.face ||
      (local && (!c.origins_lo || !c.dirs_lo)))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL matrix, instance, ray or face buffer", fn);
  const uint64_t need = batch_ws_bytes(b, c.n);
  if (!c.ws || c.ws_bytes < need)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: workspace of %llu bytes needed", fn, static_cast<unsigned long long>(need));
  if (reinterpret_cast<uintptr_t>(c.ws) % 256 != 0)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: workspace must be 256-byte aligned", fn);
  return MBVH_OK;
}

int enqueue_batch(mbvh_ctx* ctx, mbvh_batch* b, const BatchCall& c, const mbvh_batch_tuning& t, hipStream_t s) {
  const bool local = c.windows != nullptr;
  const uint32_t* order = nullptr;
  if (c.n)
    MBVH_HIP(launch_order_by_mesh(c.ray_inst, local ? nullptr : c.inst_geom, c.n_inst, c.n, b->n_meshes, c.ws, &order, s));
  int device_cus = static_cast<int>(ctx->max_threads / kMaxThreadsPerCU);
  const uint64_t tiles = (c.n + t.tile_rays - 1) / t.tile_rays;
  uint64_t blocks = static_cast<uint64_t>(device_cus) * t.blocks_per_cu;
  const uint64_t useful = (tiles + kBlock / 64 - 1) / (kBlock / 64);  // one wave per tile at most
  if (t.ray_lanes == 4 && b->max_stack4 > kQuadLdsStack + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "batch: stack scratch smaller than the tree depth");
  if (blocks > useful) blocks = std::max<uint64_t>(useful, 1);
  BatchArgs a{};
  a.meshes = b->d_meshes;
  a.n_meshes = b->n_meshes;
  a.world = local ? nullptr : c.world;
  a.inverse = local ? nullptr : c.inverse;
  a.instance_geometry = local ? nullptr : c.inst_geom;
  a.origins_lo = c.origins_lo;
  a.directions_lo = c.dirs_lo;
  a.windows = c.windows;
  a.n_instances = c.n_inst;
  a.origins = c.origins;
  a.directions = c.dirs;
  a.ray_instance = c.ray_inst;
  a.order = c.n ? order : nullptr;
  a.n = c.n;
  a.side_mask = side_mask_of(c.side);
  a.near_dist = c.near_d;
  a.far_dist = c.far_d;
  a.out_face = c.face;
  a.out_dist = c.dist;
  a.out_point = c.point;
  a.out_bary = c.bary;
  a.counter = b->d_counter;
  a.tile = t.tile_rays;
  a.n_tiles = tiles;
  a.stack = stack_of(ctx, t.ray_lanes == 4 ? blocks * kBlock / 4 : blocks * kBlock);
  MBVH_HIP(launch_batch_raycast_first(a, t.stack_lds, t.ray_lanes, static_cast<uint32_t>(blocks), s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_batch_create(mbvh_ctx* ctx, const mbvh_mesh* const* meshes, uint32_t n_meshes,
                                 mbvh_batch** out_batch) {
  clear_error();
  if (!out_batch) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_batch_create: out_batch is NULL");
  *out_batch = nullptr;
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_batch_create: invalid context handle");
  if (n_meshes == 0 || !meshes) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_batch_create: no meshes");
  std::vector<MeshView> views(n_meshes);
  uint32_t max_stack = 0, max_stack4 = 0;
  for (uint32_t i = 0; i < n_meshes; i++) {
    if (!valid_mesh(meshes[i]) || meshes[i]->ctx != ctx)
      return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_batch_create: mesh %u is not a mesh of this context", i);
    views[i] = view_of(meshes[i]);
    max_stack = std::max(max_stack, meshes[i]->max_stack);
    max_stack4 = std::max(max_stack4, meshes[i]->max_stack4);
  }
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = reserve_stack(ctx, std::max(max_stack, max_stack4)))) return rc;
  mbvh_batch* b = new (std::nothrow) mbvh_batch();
  if (!b) return set_error(MBVH_ERR_OOM, "mbvh_batch_create: out of host memory");
  b->ctx = ctx;
  b->n_meshes = n_meshes;
  b->max_stack = max_stack;
  b->max_stack4 = max_stack4;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_meshes), views.size() * sizeof(MeshView));
  if (e == hipSuccess) e = hipMemcpy(b->d_meshes, views.data(), views.size() * sizeof(MeshView), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_counter), 256);
  if (e != hipSuccess) {
    if (b->d_meshes) (void)hipFree(b->d_meshes);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
    return set_error(MBVH_ERR_OOM, "mbvh_batch_create: device allocation failed (%s)", hipGetErrorString(e));
  }
  *out_batch = b;
  return MBVH_OK;
}

extern "C" void mbvh_batch_free(mbvh_batch* batch) {
  if (!valid_batch(batch)) return;
  if (valid_ctx(batch->ctx)) (void)hipSetDevice(batch->ctx->device);
  (void)hipDeviceSynchronize();
  (void)hipFree(batch->d_meshes);
  (void)hipFree(batch->d_counter);
  batch->magic = 0;
  delete batch;
}

extern "C" int mbvh_batch_workspace_bytes(const mbvh_batch* batch, uint64_t n_rays, uint64_t* bytes) {
  clear_error();
  if (!valid_batch(batch)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_batch_workspace_bytes: invalid batch handle");
  if (!bytes) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_batch_workspace_bytes: NULL output");
  *bytes = batch_ws_bytes(batch, n_rays);
  return MBVH_OK;
}

#define MBVH_BATCH_CALL                                                                          \
  BatchCall c{matrix_world, matrix_world_inverse, instance_geometry, n_instances, origins, directions, \
              ray_instance, n, side, near_dist, far_dist, face_index, distance, point, barycoord,  \
              workspace, workspace_bytes}

extern "C" int mbvh_batch_raycast_first_d(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                                          const float* matrix_world_inverse, const uint32_t* instance_geometry,
                                          uint32_t n_instances, const float* origins, const float* directions,
                                          const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                                          float far_dist, uint32_t* face_index, float* distance, float* point,
                                          float* barycoord, void* workspace, uint64_t workspace_bytes,
                                          const mbvh_batch_tuning* tuning, void* stream) {
  clear_error();
  const char* fn = "mbvh_batch_raycast_first_d";
  MBVH_BATCH_CALL;
  int rc = check_batch_call(fn, ctx, batch, c);
  if (rc) return rc;
  const mbvh_batch_tuning t = tuning ? *tuning : batch->tuning;
  if ((rc = check_tuning(fn, t))) return rc;
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return enqueue_batch(ctx, batch, c, t, s);
}

extern "C" int mbvh_batch_raycast_first(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                                        const float* matrix_world_inverse, const uint32_t* instance_geometry,
                                        uint32_t n_instances, const float* origins, const float* directions,
                                        const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                                        float far_dist, uint32_t* face_index, float* distance, float* point,
                                        float* barycoord) {
  clear_error();
  const char* fn = "mbvh_batch_raycast_first";
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!valid_batch(batch) || batch->ctx != ctx) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid batch handle", fn);
  if (n == 0) return MBVH_OK;
  if (!matrix_world || !matrix_world_inverse || !instance_geometry || !origins || !directions || !ray_instance ||
      !face_index)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL matrix, instance, ray or face buffer", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const size_t v3 = n * 3 * sizeof(float), m16 = static_cast<size_t>(n_instances) * 16 * sizeof(float);
  void *dW, *dI, *dG, *dO, *dD, *dR, *oF, *oT, *oP, *oB;
  if ((rc = stage_in(ctx, ctx->grad[0], matrix_world, m16, &dW)) ||
      (rc = stage_in(ctx, ctx->grad[1], matrix_world_inverse, m16, &dI)) ||
      (rc = stage_in(ctx, ctx->grad[2], instance_geometry, n_instances * sizeof(uint32_t), &dG)) ||
      (rc = stage_in(ctx, ctx->grad[3], origins, v3, &dO)) || (rc = stage_in(ctx, ctx->grad[4], directions, v3, &dD)) ||
      (rc = stage_in(ctx, ctx->grad[5], ray_instance, n * sizeof(uint32_t), &dR)) ||
      (rc = stage_out(ctx->grad[6], face_index, n * sizeof(uint32_t), &oF)) ||
      (rc = stage_out(ctx->grad[7], distance, n * sizeof(float), &oT)) ||
      (rc = stage_out(ctx->grad[8], point, v3, &oP)) || (rc = stage_out(ctx->grad[9], barycoord, v3, &oB)) ||
      (rc = ensure_buffer(ctx->grad_ws, batch_ws_bytes(batch, n))))
    return rc;
  const BatchCall c{static_cast<float*>(dW), static_cast<float*>(dI), static_cast<uint32_t*>(dG), n_instances,
                    static_cast<float*>(dO), static_cast<float*>(dD), static_cast<uint32_t*>(dR), n, side, near_dist,
                    far_dist, static_cast<uint32_t*>(oF), static_cast<float*>(oT), static_cast<float*>(oP),
                    static_cast<float*>(oB), ctx->grad_ws.ptr, ctx->grad_ws.bytes};
  if ((rc = check_batch_call(fn, ctx, batch, c))) return rc;
  if ((rc = enqueue_batch(ctx, batch, c, batch->tuning, s))) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, oF, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, oT, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, oP, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, oB, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

extern "C" int mbvh_batch_raycast_first_local_f64(mbvh_ctx* ctx, mbvh_batch* batch, const double* origins,
                                                  const double* directions, const double* windows,
                                                  const uint32_t* ray_mesh, uint64_t n, int32_t side,
                                                  uint32_t* face_index, float* distance, float* point,
                                                  float* barycoord) {
  clear_error();
  const char* fn = "mbvh_batch_raycast_first_local_f64";
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!valid_batch(batch) || batch->ctx != ctx) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid batch handle", fn);
  if (n == 0) return MBVH_OK;
  if (!origins || !directions || !ray_mesh || !face_index)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL ray, mesh or face buffer", fn);
  std::vector<float> oh, ol, dh, dl, win(4 * n);
  try {
    split_f64(origins, 3 * n, oh, ol);
    split_f64(directions, 3 * n, dh, dl);
  } catch (...) {
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  for (uint64_t i = 0; i < n; i++) {
    split_f64(windows ? windows[2 * i] : 0.0, win[4 * i], win[4 * i + 1]);
    split_f64(windows ? windows[2 * i + 1] : std::numeric_limits<double>::infinity(), win[4 * i + 2], win[4 * i + 3]);
  }
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const size_t v3 = n * 3 * sizeof(float);
  void *dO, *dD, *dOl, *dDl, *dWn, *dR, *oF, *oT, *oP, *oB;
  if ((rc = stage_in(ctx, ctx->grad[3], oh.data(), v3, &dO)) || (rc = stage_in(ctx, ctx->grad[4], dh.data(), v3, &dD)) ||
      (rc = stage_in(ctx, ctx->lo_stage[0], ol.data(), v3, &dOl)) ||
      (rc = stage_in(ctx, ctx->lo_stage[1], dl.data(), v3, &dDl)) ||
      (rc = stage_in(ctx, ctx->lo_stage[2], win.data(), win.size() * sizeof(float), &dWn)) ||
      (rc = stage_in(ctx, ctx->grad[5], ray_mesh, n * sizeof(uint32_t), &dR)) ||
      (rc = stage_out(ctx->grad[6], face_index, n * sizeof(uint32_t), &oF)) ||
      (rc = stage_out(ctx->grad[7], distance, n * sizeof(float), &oT)) ||
      (rc = stage_out(ctx->grad[8], point, v3, &oP)) || (rc = stage_out(ctx->grad[9], barycoord, v3, &oB)) ||
      (rc = ensure_buffer(ctx->grad_ws, batch_ws_bytes(batch, n))))
    return rc;
  BatchCall c{nullptr, nullptr, nullptr, 0, static_cast<float*>(dO), static_cast<float*>(dD),
              static_cast<uint32_t*>(dR), n, side, 0.0f, std::numeric_limits<float>::infinity(), static_cast<uint32_t*>(oF),
              static_cast<float*>(oT),
              static_cast<float*>(oP), static_cast<float*>(oB), ctx->grad_ws.ptr, ctx->grad_ws.bytes,
              static_cast<float*>(dOl), static_cast<float*>(dDl), static_cast<float*>(dWn)};
  if ((rc = check_batch_call(fn, ctx, batch, c))) return rc;
  if ((rc = enqueue_batch(ctx, batch, c, batch->tuning, s))) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, oF, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, oT, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, oP, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, oB, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, fn);
}

extern "C" int mbvh_batch_autotune(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                                   const float* matrix_world_inverse, const uint32_t* instance_geometry,
                                   uint32_t n_instances, const float* origins, const float* directions,
                                   const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                                   float far_dist, uint32_t* face_index, float* distance, float* point,
                                   float* barycoord, void* workspace, uint64_t workspace_bytes,
                                   mbvh_batch_tuning* out_tuning) {
  clear_error();
  const char* fn = "mbvh_batch_autotune";
  MBVH_BATCH_CALL;
  int rc = check_batch_call(fn, ctx, batch, c);
  if (rc) return rc;
  if (!out_tuning) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL output", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  hipEvent_t e0, e1;
  MBVH_HIP(hipEventCreate(&e0));
  MBVH_HIP(hipEventCreate(&e1));
  mbvh_batch_tuning best{4, 16, 256, 4, 3.4e38f};
  struct Shape {
    uint32_t lanes, stack;
  };
  static const Shape kShapes[] = {{1, 8}, {1, 16}, {4, 16}};
  static const uint32_t kTiles[] = {64, 128, 256, 512};
  static const uint32_t kBlocks[] = {1, 2, 4, 8};
  for (const Shape& sh : kShapes)
    for (uint32_t tl : kTiles)
      for (uint32_t bl : kBlocks) {
        const mbvh_batch_tuning t{sh.lanes, sh.stack, tl, bl, 0.0f};
        if ((rc = enqueue_batch(ctx, batch, c, t, s))) break;  // warm-up
        float total = 0.0f;
        for (int rep = 0; rep < 3 && !rc; rep++) {
          float ms = 0.0f;
          if (hipEventRecord(e0, s) != hipSuccess || (rc = enqueue_batch(ctx, batch, c, t, s)) ||
              hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            if (!rc) rc = set_error(MBVH_ERR_HIP, "%s: timing failed", fn);
            break;
          }
          total += ms;
        }
        if (rc) break;
        if (total / 3.0f < best.ms) best = mbvh_batch_tuning{sh.lanes, sh.stack, tl, bl, total / 3.0f};
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  batch->tuning = best;
  *out_tuning = best;
  return MBVH_OK;
}

extern "C" int mbvh_batch_graph_create(mbvh_ctx* ctx, mbvh_batch* batch, const float* matrix_world,
                                       const float* matrix_world_inverse, const uint32_t* instance_geometry,
                                       uint32_t n_instances, const float* origins, const float* directions,
                                       const uint32_t* ray_instance, uint64_t n, int32_t side, float near_dist,
                                       float far_dist, uint32_t* face_index, float* distance, float* point,
                                       float* barycoord, void* workspace, uint64_t workspace_bytes,
                                       const mbvh_batch_tuning* tuning, mbvh_batch_graph** out_graph) {
  clear_error();
  const char* fn = "mbvh_batch_graph_create";
  if (!out_graph) return set_error(MBVH_ERR_BAD_ARGS, "%s: out_graph is NULL", fn);
  *out_graph = nullptr;
  MBVH_BATCH_CALL;
  int rc = check_batch_call(fn, ctx, batch, c);
  if (rc) return rc;
  const mbvh_batch_tuning t = tuning ? *tuning : batch->tuning;
  if ((rc = check_tuning(fn, t))) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t cs;
  MBVH_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = enqueue_batch(ctx, batch, c, t, cs);
    e = hipStreamEndCapture(cs, &graph);
  }
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return set_error(MBVH_ERR_HIP, "%s: capture failed (%s)", fn, hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return set_error(MBVH_ERR_HIP, "%s: instantiate failed (%s)", fn, hipGetErrorString(e));
  }
  mbvh_batch_graph* g = new (std::nothrow) mbvh_batch_graph();
  if (!g) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  g->ctx = ctx;
  g->graph = graph;
  g->exec = exec;
  g->stack_gen = ctx->stack_gen;
  *out_graph = g;
  return MBVH_OK;
}

extern "C" int mbvh_batch_graph_launch(mbvh_batch_graph* graph, void* stream) {
  clear_error();
  if (!graph || graph->magic != kGraphMagic || !valid_ctx(graph->ctx))
    return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_batch_graph_launch: invalid graph handle");
  if (graph->stack_gen != graph->ctx->stack_gen)
    return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_batch_graph_launch: the context's traversal stack scratch was "
                     "reallocated after capture (a deeper mesh, batch or scene was created); free the graph and "
                     "capture it again");
  int rc = set_device(graph->ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : graph->ctx->stream;
  MBVH_HIP(hipGraphLaunch(graph->exec, s));
  return MBVH_OK;
}

extern "C" void mbvh_batch_graph_free(mbvh_batch_graph* graph) {
  if (!graph || graph->magic != kGraphMagic) return;
  (void)hipSetDevice(graph->ctx->device);
  (void)hipDeviceSynchronize();
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  graph->magic = 0;
  delete graph;
}

extern "C" int mbvh_instance_rays_d(mbvh_ctx* ctx, const float* matrix_world_inverse, uint32_t n_instances,
                                    const float* origins, const float* directions, const uint32_t* ray_instance,
                                    uint64_t n, float* local_origins, float* local_directions, void* stream) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_instance_rays_d: invalid context handle");
  if (n == 0) return MBVH_OK;
  if (!matrix_world_inverse || !origins || !directions || !ray_instance || !local_origins || !local_directions)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_instance_rays_d: NULL buffer");
  int rc = set_device(ctx);
  if (rc) return rc;
  InstanceRayArgs a{matrix_world_inverse, n_instances, origins, directions, ray_instance, n, local_origins,
                    local_directions};
  MBVH_HIP(launch_instance_rays(a, stream ? static_cast<hipStream_t>(stream) : ctx->stream));
  return MBVH_OK;
}

extern "C" int mbvh_instance_hits_d(mbvh_ctx* ctx, const float* matrix_world, uint32_t n_instances,
                                    const float* origins, const uint32_t* ray_instance, uint64_t n,
                                    const uint32_t* face_index, float* point, float* distance, void* stream) {
  clear_error();
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_instance_hits_d: invalid context handle");
  if (n == 0) return MBVH_OK;
  if (!matrix_world || !origins || !ray_instance || !face_index || !point || !distance)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_instance_hits_d: NULL buffer");
  int rc = set_device(ctx);
  if (rc) return rc;
  InstanceHitArgs a{matrix_world, n_instances, origins, ray_instance, n, face_index, point, distance};
  MBVH_HIP(launch_instance_hits(a, stream ? static_cast<hipStream_t>(stream) : ctx->stream));
  return MBVH_OK;
}

// ---- device build ----------------------------------------------------------------------------------

namespace {

// Layout of one build: the primitive buffer (positions [0, span), triangle base + q of the index, or
// the generated indirect buffer) and the trees over it (buildTree.js:128-140, GeometryBVH.js:9-35,
// 169-181).
struct BuildPlan {
  bool indirect = false;
  uint64_t base = 0, span = 0;
  std::vector<uint32_t> src;    // (first triangle, first position) of each source range
  std::vector<uint32_t> roots;  // (first position, count) of each tree
  int32_t out_elt = 4;
};

int plan_build(const char* fn, const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
               int32_t index_elt, BuildPlan* p) {
  const uint64_t last_end = static_cast<uint64_t>(root_ranges[2 * (n_roots - 1)]) + root_ranges[2 * (n_roots - 1) + 1];
  p->indirect = opts->indirect != 0;
  if (p->indirect) {
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_roots; r++) {
      p->src.push_back(root_ranges[2 * r]);
      p->src.push_back(static_cast<uint32_t>(total));
      total += root_ranges[2 * r + 1];
    }
    p->base = 0;
    p->span = total;
    p->roots = {0u, static_cast<uint32_t>(total)};
    p->out_elt = last_end > 65536 ? 4 : 2;
  } else {
    p->base = root_ranges[0];
    p->span = last_end - p->base;
    p->src = {static_cast<uint32_t>(p->base), 0u};
    for (uint32_t r = 0; r < n_roots; r++) {
      p->roots.push_back(static_cast<uint32_t>(root_ranges[2 * r] - p->base));
      p->roots.push_back(root_ranges[2 * r + 1]);
    }
    p->out_elt = index_elt;
  }
  const uint64_t n_trees = p->roots.size() / 2;
  if (2 * p->span + n_trees + 2 > 0xFFFFFFFFull || last_end > 0xFFFFFFFFull)
    return set_error(MBVH_ERR_LIMIT, "%s: %llu triangles exceed the device build's 32-bit node and position ids", fn,
                     static_cast<unsigned long long>(p->span));
  return MBVH_OK;
}

struct BuildOut {
  const uint32_t* packed = nullptr;   // device, trees one after another
  std::vector<uint64_t> tree_nodes;   // nodes per tree
  const void* prims = nullptr;        // device: permuted triplets or the indirect buffer
  uint64_t prims_bytes = 0;
};

size_t carve(size_t& at, size_t bytes) {
  const size_t here = at;
  at = (at + bytes + 255) & ~static_cast<size_t>(255);
  return here;
}

// Runs the device build on the context's stream (inputs already on the device) and waits for it.
int gpu_build(const char* fn, mbvh_ctx* ctx, const float* d_positions, uint64_t vertex_count, uint32_t stride,
              uint32_t offset, const void* d_index, int32_t index_elt, uint64_t index_count, const BuildPlan& p,
              const mbvh_build_opts* opts, hipStream_t s, BuildOut* out) {
  const uint32_t n_trees = static_cast<uint32_t>(p.roots.size() / 2);
  const uint64_t max_nodes = 2 * p.span + n_trees + 2;
  const uint32_t n_lvl = static_cast<uint32_t>(opts->max_depth) + 2;
  const uint64_t prims_bytes = p.indirect ? p.span * p.out_elt : 3 * p.span * index_elt;
  size_t at = 0;
  const size_t o_prim = carve(at, 6 * p.span * sizeof(float));
  const size_t o_perm = carve(at, p.span * 4), o_pl = carve(at, p.span * 4), o_pr = carve(at, p.span * 4);
  const size_t o_box = carve(at, 6 * max_nodes * 4), o_cen = carve(at, 6 * max_nodes * 4);
  const size_t o_info = carve(at, max_nodes * 16), o_root = carve(at, max_nodes * 4);
  const size_t o_size = carve(at, max_nodes * 4), o_slot = carve(at, max_nodes * 4);
  const size_t o_lvl = carve(at, 4ull * n_lvl * 4), o_big = carve(at, max_nodes * 4);
  const size_t o_rkeys = carve(at, 12ull * n_trees * 8), o_err = carve(at, 4 * 8);
  const uint64_t max_big = p.span / kBuildBigCount + n_trees + 1;
  const uint64_t max_chunks = p.span / kBuildChunk + max_big + 1;
  const size_t o_wbins = carve(at, max_big * 3 * 32 * 7 * 4), o_waxis = carve(at, max_big * 4);
  const size_t o_wpos = carve(at, max_big * 8), o_wbase = carve(at, (max_big + 1) * 4);
  const size_t o_wcnt = carve(at, max_chunks * 4), o_wpre = carve(at, max_chunks * 4);
  const size_t o_wnl = carve(at, max_big * 4), o_wm = carve(at, max_big * 4), o_wkeys = carve(at, max_big * 24 * 8);
  const uint64_t max_width = p.span + n_trees + 1;  // nodes of one level (disjoint ranges, empty roots aside)
  const size_t o_tsplit = carve(at, max_width * 8), o_tbound = carve(at, max_width * 24 * 4);
  const size_t o_ttile = carve(at, (max_width / 1024 + 2) * 4);
  const size_t o_packed = carve(at, max_nodes * kBytesPerNode), o_rn = carve(at, n_trees * 4ull);
  const size_t o_out = carve(at, prims_bytes), o_src = carve(at, p.src.size() * 4), o_roots = carve(at, p.roots.size() * 4);
  int rc = ensure_buffer(ctx->build_ws, at);
  if (rc) return rc;
  uint8_t* ws = static_cast<uint8_t*>(ctx->build_ws.ptr);
  GpuBuildArgs a{};
  a.positions = d_positions;
  a.n_vertices = vertex_count;
  a.stride = stride;
  a.offset = offset;
  a.index = d_index;
  a.index_elt = index_elt;
  a.index_count = d_index ? index_count : 0;
  a.src = reinterpret_cast<const uint32_t*>(ws + o_src);
  a.n_src = static_cast<uint32_t>(p.src.size() / 2);
  a.roots = reinterpret_cast<const uint32_t*>(ws + o_roots);
  a.n_roots = n_trees;
  a.base = p.base;
  a.span = p.span;
  a.strategy = opts->strategy;
  a.max_depth = opts->max_depth;
  a.leaf_size = opts->target_leaf_size;
  a.prim = reinterpret_cast<float*>(ws + o_prim);
  a.perm = reinterpret_cast<uint32_t*>(ws + o_perm);
  a.pair_l = reinterpret_cast<uint32_t*>(ws + o_pl);
  a.pair_r = reinterpret_cast<uint32_t*>(ws + o_pr);
  a.nbox = reinterpret_cast<float*>(ws + o_box);
  a.ncen = reinterpret_cast<float*>(ws + o_cen);
  a.ninfo = reinterpret_cast<uint4*>(ws + o_info);
  a.nroot = reinterpret_cast<uint32_t*>(ws + o_root);
  a.nsize = reinterpret_cast<uint32_t*>(ws + o_size);
  a.nslot = reinterpret_cast<uint32_t*>(ws + o_slot);
  a.max_nodes = max_nodes;
  a.lvl = reinterpret_cast<uint32_t*>(ws + o_lvl);
  a.big = reinterpret_cast<uint32_t*>(ws + o_big);
  a.rkeys = reinterpret_cast<unsigned long long*>(ws + o_rkeys);
  a.err = reinterpret_cast<unsigned long long*>(ws + o_err);
  a.wbins = reinterpret_cast<uint32_t*>(ws + o_wbins);
  a.waxis = reinterpret_cast<int32_t*>(ws + o_waxis);
  a.wpos = reinterpret_cast<double*>(ws + o_wpos);
  a.wbase = reinterpret_cast<uint32_t*>(ws + o_wbase);
  a.wcnt = reinterpret_cast<uint32_t*>(ws + o_wcnt);
  a.wpre = reinterpret_cast<uint32_t*>(ws + o_wpre);
  a.wnl = reinterpret_cast<uint32_t*>(ws + o_wnl);
  a.wm = reinterpret_cast<uint32_t*>(ws + o_wm);
  a.wkeys = reinterpret_cast<unsigned long long*>(ws + o_wkeys);
  a.tsplit = reinterpret_cast<uint2*>(ws + o_tsplit);
  a.tbound = reinterpret_cast<float*>(ws + o_tbound);
  a.ttile = reinterpret_cast<uint32_t*>(ws + o_ttile);
  a.packed = reinterpret_cast<uint32_t*>(ws + o_packed);
  a.root_nodes = reinterpret_cast<uint32_t*>(ws + o_rn);
  a.prims_out = ws + o_out;
  a.out_elt = p.out_elt;
  a.indirect = p.indirect ? 1 : 0;
  MBVH_HIP(hipMemcpyAsync(ws + o_src, p.src.data(), p.src.size() * 4, hipMemcpyHostToDevice, s));
  MBVH_HIP(hipMemcpyAsync(ws + o_roots, p.roots.data(), p.roots.size() * 4, hipMemcpyHostToDevice, s));
  MBVH_HIP(launch_build_init(a, s));
  // Level by level. While a level has big nodes the host reads the next level's counts after it
  // (the big-node launches are needed only then); afterwards every fourth level, to stop at the
  // first empty one.
  bool wide = true;
  uint32_t counts[2] = {0, 0};
  for (uint32_t d = 0; d <= static_cast<uint32_t>(opts->max_depth); d++) {
    MBVH_HIP(launch_build_level(a, d, wide, s));
    if (wide || d % 4 == 3) {
      MBVH_HIP(hipMemcpyAsync(&counts[0], a.lvl + n_lvl + d + 1, 4, hipMemcpyDeviceToHost, s));
      MBVH_HIP(hipMemcpyAsync(&counts[1], a.lvl + 3 * n_lvl + d + 1, 4, hipMemcpyDeviceToHost, s));
      MBVH_HIP(hipStreamSynchronize(s));
      if (counts[0] == 0) break;
      wide = counts[1] != 0;
    }
  }
  std::vector<uint32_t> lvl(4ull * n_lvl);
  unsigned long long err[3] = {0, 0, 0};
  MBVH_HIP(hipMemcpyAsync(lvl.data(), a.lvl, lvl.size() * 4, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipMemcpyAsync(err, a.err, sizeof(err), hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  if (err[0] & 2u) {
    if (err[2] != ~0ull)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: index[%llu] is out of range", fn, err[2]);
    return set_error(MBVH_ERR_BAD_ARGS, "%s: a triangle references a vertex past the end", fn);
  }
  if (err[0] & 1u)
    return set_error(MBVH_ERR_LIMIT,
                     "%s: a leaf would hold %llu triangles; the node format stores at most 65535 (raise max_depth)", fn,
                     err[1]);
  uint32_t levels = 0;
  for (uint32_t d = 0; d + 1 < n_lvl; d++)
    if (lvl[n_lvl + d]) levels = d + 1;
  const uint64_t total = levels ? static_cast<uint64_t>(lvl[levels - 1]) + lvl[n_lvl + levels - 1] : 0;
  MBVH_HIP(launch_build_pack(a, levels, total, s));
  std::vector<uint32_t> rn(n_trees);
  MBVH_HIP(hipMemcpyAsync(rn.data(), a.root_nodes, n_trees * 4ull, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  out->packed = a.packed;
  out->tree_nodes.assign(rn.begin(), rn.end());
  out->prims = a.prims_out;
  out->prims_bytes = prims_bytes;
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_build_gpu(mbvh_ctx* ctx, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                              uint32_t position_offset, void* index, int32_t index_elt_bytes, uint64_t index_count,
                              const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                              mbvh_tree** out_tree) {
  clear_error();
  const char* fn = "mbvh_build_gpu";
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!out_tree) return set_error(MBVH_ERR_BAD_ARGS, "%s: out_tree is NULL", fn);
  *out_tree = nullptr;
  int rc = check_build_args(fn, positions, vertex_count, position_stride, position_offset, index, index_elt_bytes,
                            index_count, root_ranges, n_roots, opts, false);
  if (!rc && opts->strict_leaf_size != 0)
    rc = set_error(MBVH_ERR_BAD_ARGS, "%s: strict_leaf_size is a host build option", fn);
  BuildPlan plan;
  if (!rc) rc = plan_build(fn, root_ranges, n_roots, opts, index_elt_bytes, &plan);
  if (rc) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  const size_t pos_bytes = vertex_count ? ((vertex_count - 1) * position_stride + position_offset + 3) * sizeof(float) : 0;
  const size_t idx_bytes = index ? index_count * static_cast<size_t>(index_elt_bytes) : 0;
  void* d_pos = nullptr;
  void* d_idx = nullptr;
  if ((rc = stage_in(ctx, ctx->staging[0], pos_bytes ? positions : nullptr, pos_bytes, &d_pos)) ||
      (rc = stage_in(ctx, ctx->staging[1], idx_bytes ? index : nullptr, idx_bytes, &d_idx)))
    return rc;
  BuildOut res;
  if ((rc = gpu_build(fn, ctx, static_cast<const float*>(d_pos), vertex_count, position_stride, position_offset, d_idx,
                      index_elt_bytes, index_count, plan, opts, s, &res)))
    return rc;
  mbvh_tree* tree = new (std::nothrow) mbvh_tree();
  if (!tree) return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  try {
    tree->roots.resize(res.tree_nodes.size());
    for (size_t r = 0; r < res.tree_nodes.size(); r++) tree->roots[r].resize(res.tree_nodes[r] * kBytesPerNode);
    if (plan.indirect) {
      tree->indirect_elt_bytes = plan.out_elt;
      tree->indirect_count = plan.span;
      tree->indirect.resize(res.prims_bytes);
    }
  } catch (...) {
    delete tree;
    return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  }
  uint64_t node_at = 0;
  for (size_t r = 0; r < tree->roots.size(); r++) {
    const size_t bytes = tree->roots[r].size();
    hipError_t e = hipMemcpyAsync(tree->roots[r].data(), res.packed + node_at * kU32PerNode, bytes, hipMemcpyDeviceToHost, s);
    node_at += res.tree_nodes[r];
    if (e != hipSuccess) {
      delete tree;
      return set_error(MBVH_ERR_HIP, "%s: node copy failed (%s)", fn, hipGetErrorString(e));
    }
  }
  hipError_t e = hipSuccess;
  if (plan.indirect) {
    if (res.prims_bytes) e = hipMemcpyAsync(tree->indirect.data(), res.prims, res.prims_bytes, hipMemcpyDeviceToHost, s);
  } else if (res.prims_bytes) {
    e = hipMemcpyAsync(static_cast<uint8_t*>(index) + 3 * plan.base * index_elt_bytes, res.prims, res.prims_bytes,
                       hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete tree;
    return set_error(MBVH_ERR_HIP, "%s: result copy failed (%s)", fn, hipGetErrorString(e));
  }
  *out_tree = tree;
  return MBVH_OK;
}

extern "C" int mbvh_build_gpu_d(mbvh_ctx* ctx, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                                uint32_t position_offset, void* index, int32_t index_elt_bytes, uint64_t index_count,
                                const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                                void* nodes_out, uint64_t nodes_capacity, uint64_t* tree_nodes, uint32_t n_tree_slots,
                                void* indirect_out, uint64_t indirect_capacity, void* stream) {
  clear_error();
  const char* fn = "mbvh_build_gpu_d";
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  int rc = check_build_args(fn, positions, vertex_count, position_stride, position_offset, index, index_elt_bytes,
                            index_count, root_ranges, n_roots, opts, false);
  if (!rc && opts->strict_leaf_size != 0)
    rc = set_error(MBVH_ERR_BAD_ARGS, "%s: strict_leaf_size is a host build option", fn);
  BuildPlan plan;
  if (!rc) rc = plan_build(fn, root_ranges, n_roots, opts, index_elt_bytes, &plan);
  if (rc) return rc;
  const uint32_t n_trees = static_cast<uint32_t>(plan.roots.size() / 2);
  if (!tree_nodes || n_tree_slots < n_trees)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: tree_nodes must hold %u counts", fn, n_trees);
  const uint64_t need_nodes = (2 * plan.span + n_trees) * kBytesPerNode;
  if (!nodes_out || nodes_capacity < need_nodes)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: nodes_out must hold %llu bytes (2 x triangles + trees nodes)", fn,
                     static_cast<unsigned long long>(need_nodes));
  if (plan.indirect && (!indirect_out || indirect_capacity < plan.span * plan.out_elt))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: indirect_out must hold %llu bytes", fn,
                     static_cast<unsigned long long>(plan.span * plan.out_elt));
  std::lock_guard<std::mutex> guard(ctx->lock);  // the build workspace belongs to the context
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  BuildOut res;
  if ((rc = gpu_build(fn, ctx, positions, vertex_count, position_stride, position_offset, index, index_elt_bytes,
                      index_count, plan, opts, s, &res)))
    return rc;
  uint64_t total = 0;
  for (uint32_t r = 0; r < n_trees; r++) {
    tree_nodes[r] = res.tree_nodes[r];
    total += res.tree_nodes[r];
  }
  if (total) MBVH_HIP(hipMemcpyAsync(nodes_out, res.packed, total * kBytesPerNode, hipMemcpyDeviceToDevice, s));
  if (res.prims_bytes) {
    void* dst = plan.indirect ? indirect_out : static_cast<uint8_t*>(index) + 3 * plan.base * index_elt_bytes;
    MBVH_HIP(hipMemcpyAsync(dst, res.prims, res.prims_bytes, hipMemcpyDeviceToDevice, s));
  }
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

// ---- scenes ------------------------------------------------------------------------------------------

// Depth bound of a scene's top-level tree (the scene raycast's per-lane stack holds that many levels).
constexpr uint32_t kTlasMaxDepth = 24;
// Scene closest point: leaf steps a quad takes on one point before leaving it to a wave of its own,
// and the persistent waves of that pass.
constexpr uint32_t kSceneHeavyBudget = 512;
constexpr uint32_t kSceneWideWaves = 2048;

struct mbvh_scene {
  uint32_t magic = 0x4E454353u;  // "SCEN"
  mbvh_ctx* ctx = nullptr;
  uint32_t n_meshes = 0;
  uint32_t n_instances = 0;
  MeshView* d_meshes = nullptr;
  float* d_world = nullptr;
  float* d_inverse = nullptr;
  float* d_frame = nullptr;        // the library's own inverse of each matrixWorld (closest point bounds)
  uint8_t* d_frame_ok = nullptr;   // that inverse passed its residual check
  uint32_t* d_geom = nullptr;
  uint32_t* d_tlas = nullptr;
  uint32_t* d_tlas_inst = nullptr;
  uint32_t max_stack = 0;         // stack entries of the deepest instance walk (raycast)
  uint32_t max_stack_points = 0;  // closest point: the top level's far children below them
  uint32_t max_stack_points4 = 0;  // the same for the four-lane kernel (4-wide instance trees)
  bool wide_ok = false;            // every instance tree's 4-wide walk fits the wave-per-point frontier
  // host copies for the WebGPU packing
  std::vector<const mbvh_mesh*> meshes;
  std::vector<float> world_h, frame_h;  // 16 per instance: matrixWorld, the library's own inverse
  std::vector<uint32_t> geom_h;
  std::vector<uint8_t> visible_h;
};

namespace {

bool valid_scene(const mbvh_scene* s) { return s && s->magic == 0x4E454353u; }

// The inverse of an affine column-major 4x4 (local -> world) in double, rounded to float, and whether
// W times that rounded inverse is the identity to 2^-20 in every entry of the 3x3 part and of the
// translation relative to the magnitudes involved (false for singular or badly conditioned W).
bool frame_inverse(const float* e, float* out) {
  const double a = e[0], b = e[4], c = e[8], d = e[1], f = e[5], g = e[9], h = e[2], i = e[6], j = e[10];
  const double c0 = f * j - g * i, c1 = g * h - d * j, c2 = d * i - f * h;
  const double det = a * c0 + b * c1 + c * c2;
  const double m[9] = {c0 / det, (c * i - b * j) / det, (b * g - c * f) / det,
                       c1 / det, (a * j - c * h) / det, (c * d - a * g) / det,
                       c2 / det, (b * h - a * i) / det, (a * f - b * d) / det};  // row-major inverse of the 3x3
  const double t[3] = {e[12], e[13], e[14]};
  float r[16] = {0};
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) r[4 * col + row] = static_cast<float>(m[3 * row + col]);
    r[12 + row] = static_cast<float>(-(m[3 * row] * t[0] + m[3 * row + 1] * t[1] + m[3 * row + 2] * t[2]));
  }
  r[15] = 1.0f;
  std::memcpy(out, r, sizeof(r));
  if (!(std::isfinite(det) && det != 0.0)) return false;
  // residual W * R - I over the 3x3 part, and the translation W * (R t) + t_W - 0
  for (int row = 0; row < 3; row++) {
    double scale = 0.0, tr = e[12 + row], trs = std::fabs(e[12 + row]);
    for (int col = 0; col < 3; col++) {
      double s = 0.0, sa = 0.0;
      for (int k = 0; k < 3; k++) {
        s += static_cast<double>(e[4 * k + row]) * r[4 * col + k];
        sa += std::fabs(static_cast<double>(e[4 * k + row]) * r[4 * col + k]);
      }
      if (!(std::fabs(s - (row == col ? 1.0 : 0.0)) <= 9.5367431640625e-07 * std::max(sa, 1e-300))) return false;
      scale = std::max(scale, sa);
      tr += static_cast<double>(e[4 * col + row]) * r[12 + col];
      trs += std::fabs(static_cast<double>(e[4 * col + row]) * r[12 + col]);
    }
    if (!(std::fabs(tr) <= 9.5367431640625e-07 * std::max(trs, 1e-300)) || !(scale > 0.0)) return false;
  }
  return true;
}

void free_scene_buffers(mbvh_scene* sc) {
  void* bufs[] = {sc->d_meshes, sc->d_world, sc->d_inverse, sc->d_frame, sc->d_frame_ok, sc->d_geom, sc->d_tlas,
                  sc->d_tlas_inst};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
}

// Union of a mesh's root boxes (the device copy of its reference nodes).
// World box of a local box (min xyz, max xyz in double) under an affine matrixWorld: the 8 corners
// through the matrix in double, rounded outward to float and widened by 2^-20 of their magnitude, so
// that no hit rounds outside it. False if it is not finite.
bool world_box_of(const float* e, const double* lb, float* fb) {
  double mn[3] = {std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                  std::numeric_limits<double>::infinity()};
  double mx[3] = {-mn[0], -mn[1], -mn[2]};
  for (int c = 0; c < 8; c++) {
    const double x = lb[(c & 1) ? 3 : 0], y = lb[(c & 2) ? 4 : 1], z = lb[(c & 4) ? 5 : 2];
    const double w = e[3] * x + e[7] * y + e[11] * z + e[15];
    for (int k = 0; k < 3; k++) {
      const double v = (e[k] * x + e[4 + k] * y + e[8 + k] * z + e[12 + k]) / w;
      mn[k] = std::min(mn[k], v);
      mx[k] = std::max(mx[k], v);
    }
  }
  bool finite = true;
  for (int k = 0; k < 3; k++) {
    const double pad = std::max(std::max(std::fabs(mn[k]), std::fabs(mx[k])) * 9.5367431640625e-07, 1e-30);
    fb[k] = std::nextafter(static_cast<float>(mn[k] - pad), -std::numeric_limits<float>::infinity());
    fb[k + 3] = std::nextafter(static_cast<float>(mx[k] + pad), std::numeric_limits<float>::infinity());
    finite = finite && std::isfinite(mn[k]) && std::isfinite(mx[k]);
  }
  return finite;
}

int mesh_local_box(const mbvh_mesh* m, double* box) {
  for (int k = 0; k < 3; k++) {
    box[k] = std::numeric_limits<double>::infinity();
    box[k + 3] = -std::numeric_limits<double>::infinity();
  }
  for (uint32_t r = 0; r < m->root_base.size(); r++) {
    float b[6];
    MBVH_HIP(hipMemcpy(b, m->d_ref_nodes + 8 * m->root_base[r], sizeof(b), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) {
      if (b[k] <= b[k + 3]) {
        box[k] = std::min(box[k], static_cast<double>(b[k]));
        box[k + 3] = std::max(box[k + 3], static_cast<double>(b[k + 3]));
      }
    }
  }
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_scene_create(mbvh_ctx* ctx, const mbvh_mesh* const* meshes, uint32_t n_meshes,
                                 const float* matrix_world, const float* matrix_world_inverse,
                                 const uint32_t* instance_geometry, const uint8_t* visible, uint32_t n_instances,
                                 mbvh_scene** out_scene) {
  clear_error();
  const char* fn = "mbvh_scene_create";
  if (!out_scene) return set_error(MBVH_ERR_BAD_ARGS, "%s: out_scene is NULL", fn);
  *out_scene = nullptr;
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (n_meshes == 0 || !meshes) return set_error(MBVH_ERR_BAD_ARGS, "%s: no meshes", fn);
  if (n_instances && (!matrix_world || !matrix_world_inverse || !instance_geometry))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL matrices or instance geometry", fn);
  if (n_instances >= (1u << 24)) return set_error(MBVH_ERR_LIMIT, "%s: at most 2^24 - 1 instances", fn);
  std::vector<MeshView> views(n_meshes);
  uint32_t max_stack = 0, max_stack4 = 0;
  for (uint32_t i = 0; i < n_meshes; i++) {
    if (!valid_mesh(meshes[i]) || meshes[i]->ctx != ctx)
      return set_error(MBVH_ERR_BAD_HANDLE, "%s: mesh %u is not a mesh of this context", fn, i);
    views[i] = view_of(meshes[i]);
    max_stack = std::max(max_stack, meshes[i]->max_stack);
    max_stack4 = std::max(max_stack4, meshes[i]->max_stack4);
  }
  for (uint32_t i = 0; i < n_instances; i++)
    if (instance_geometry[i] >= n_meshes)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: instance %u names mesh %u of %u", fn, i, instance_geometry[i], n_meshes);
  // matrixWorld of an Object3D is affine (Matrix4.compose); the world boxes and bounds rely on it
  std::vector<float> frame(std::max<size_t>(16ull * n_instances, 16), 0.0f);
  std::vector<uint8_t> frame_ok(std::max<uint32_t>(n_instances, 1), 0);
  for (uint32_t i = 0; i < n_instances; i++) {
    const float* e = matrix_world + 16ull * i;
    if (!(e[3] == 0.0f && e[7] == 0.0f && e[11] == 0.0f && e[15] == 1.0f))
      return set_error(MBVH_ERR_BAD_ARGS, "%s: matrixWorld of instance %u is not affine (its last row must be 0 0 0 1)",
                       fn, i);
    frame_ok[i] = frame_inverse(e, &frame[16ull * i]) ? 1 : 0;
  }
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<double> local(6ull * n_meshes);
  for (uint32_t i = 0; i < n_meshes; i++)
    if ((rc = mesh_local_box(meshes[i], &local[6ull * i]))) return rc;

  // World boxes of the listed instances (8 transformed corners, rounded outward and widened by
  // 2^-20 of their magnitude so that no hit rounds outside its instance's box); one build
  // "triangle" per instance spanning its box.
  std::vector<uint32_t> listed;
  std::vector<float> prim_pos;
  for (uint32_t i = 0; i < n_instances; i++) {
    if (visible && !visible[i]) continue;
    const double* lb = &local[6ull * instance_geometry[i]];
    if (!(lb[0] <= lb[3] && lb[1] <= lb[4] && lb[2] <= lb[5])) continue;  // no triangles
    float fb[6];
    if (!world_box_of(matrix_world + 16ull * i, lb, fb))
      return set_error(MBVH_ERR_BAD_ARGS, "%s: instance %u has a non-finite world box", fn, i);
    listed.push_back(i);
    const float v[9] = {fb[0], fb[1], fb[2], fb[3], fb[4], fb[5], fb[0], fb[1], fb[2]};
    prim_pos.insert(prim_pos.end(), v, v + 9);
  }
  const uint32_t n_listed = static_cast<uint32_t>(listed.size());
  std::vector<uint32_t> index(3ull * n_listed);
  for (uint32_t k = 0; k < 3 * n_listed; k++) index[k] = k;
  std::vector<uint8_t> nodes(kBytesPerNode, 0);
  if (n_listed == 0) {  // one empty leaf (its empty box is never entered)
    const float empty[6] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                            std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                            -std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity()};
    const uint32_t leaf_word = 0xFFFF0000u;
    std::memcpy(nodes.data(), empty, sizeof(empty));
    std::memcpy(nodes.data() + 28, &leaf_word, 4);
  } else {
    const uint32_t range[2] = {0u, n_listed};
    const mbvh_build_opts opts{MBVH_SAH, kTlasMaxDepth, 2, 0};  // depth bound = the kernels' TLAS stacks
    mbvh_tree* tlas = nullptr;
    rc = mbvh_build(prim_pos.data(), 3ull * n_listed, 3, 0, index.data(), 4, index.size(), range, 1, &opts, &tlas);
    if (rc) return rc;
    nodes = tlas->roots[0];
    mbvh_tree_free(tlas);
  }
  std::vector<uint32_t> slot_inst(std::max<uint32_t>(n_listed, 1), 0u);
  for (uint32_t s = 0; s < n_listed; s++) slot_inst[s] = listed[index[3ull * s] / 3];

  // closest point keeps the top level's far children (at most one per level) below the instance walks
  const uint32_t max_stack_points = max_stack + kTlasMaxDepth;
  const uint32_t max_stack_points4 = max_stack4 + kTlasMaxDepth;
  if ((rc = reserve_stack(ctx, std::max(max_stack_points, max_stack_points4)))) return rc;
  mbvh_scene* sc = new (std::nothrow) mbvh_scene();
  if (!sc) return set_error(MBVH_ERR_OOM, "%s: out of host memory", fn);
  sc->ctx = ctx;
  sc->n_meshes = n_meshes;
  sc->n_instances = n_instances;
  sc->max_stack = max_stack;
  sc->max_stack_points = max_stack_points;
  sc->max_stack_points4 = max_stack_points4;
  sc->wide_ok = 64ull * (max_stack4 > 0 ? max_stack4 - 1 : 0) + 256 <= kWideFrontier;
  sc->meshes.assign(meshes, meshes + n_meshes);
  sc->world_h.assign(matrix_world, matrix_world + 16ull * n_instances);
  sc->frame_h.assign(frame.begin(), frame.begin() + 16ull * n_instances);
  sc->geom_h.assign(instance_geometry, instance_geometry + n_instances);
  sc->visible_h.assign(n_instances, 1);
  if (visible)
    for (uint32_t i = 0; i < n_instances; i++) sc->visible_h[i] = visible[i] ? 1 : 0;
  const size_t mat_bytes = std::max<size_t>(16ull * n_instances * sizeof(float), 64);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&sc->d_meshes), views.size() * sizeof(MeshView));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_world), mat_bytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_inverse), mat_bytes);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_frame), frame.size() * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_frame_ok), frame_ok.size());
  if (e == hipSuccess) e = hipMemcpy(sc->d_frame, frame.data(), frame.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(sc->d_frame_ok, frame_ok.data(), frame_ok.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_geom), std::max<size_t>(4ull * n_instances, 4));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_tlas), nodes.size());
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&sc->d_tlas_inst), slot_inst.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(sc->d_meshes, views.data(), views.size() * sizeof(MeshView), hipMemcpyHostToDevice);
  if (e == hipSuccess && n_instances) {
    e = hipMemcpy(sc->d_world, matrix_world, 16ull * n_instances * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(sc->d_inverse, matrix_world_inverse, 16ull * n_instances * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sc->d_geom, instance_geometry, 4ull * n_instances, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMemcpy(sc->d_tlas, nodes.data(), nodes.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(sc->d_tlas_inst, slot_inst.data(), slot_inst.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_scene_buffers(sc);
    delete sc;
    return set_error(MBVH_ERR_OOM, "%s: device allocation or copy failed (%s)", fn, hipGetErrorString(e));
  }
  *out_scene = sc;
  return MBVH_OK;
}

// ---- a scene in the reference's WebGPU packing -------------------------------------------------
// BVHComputeData.update (src/webgpu/BVHComputeData.js:217-447) over ClusteredBVH's primitives
// (src/webgpu/ClusteredBVH.js:207-297): per instance and root, either the whole root (primitive
// limit 0, what ClusteredBVH does for instanced and batched objects and for scenes of fewer than three
// meshes) or the cluster cut of _traverseClusters (:302-331: depth first, left child first, the first
// node on each path that is a leaf or spans fewer triangles than the limit); one primitive per
// top-level leaf (_strictLeafSize 1); the top level first (appendBVHData,
// src/webgpu/utils/packBVHBufferUtils.js:44-131: a leaf holds the absolute node offset of its
// subtree and 0xFF000000 | transform slot), then each (mesh, root, node) subtree once, in the order
// the leaves first name it, its contiguous node span copied (getSubtreeNodeCount, :142-161;
// appendBVHSubtree, :216-280: leaf offsets rebased by the mesh's triangle base); transforms per slot as writeTransformData
// (BVHComputeData.js:491-526: matrixWorld, its inverse, visible, three padding words); the index
// buffer per mesh in the same order, rebased by the mesh's vertex base (appendIndexData,
// packBVHBufferUtils.js:293-322, with the mesh's triangles in slot order). The top level is this
// library's SAH build over the primitives' world boxes, so its shape is this build's, not ClusteredBVH's.
namespace {

struct WebgpuPacking {
  std::vector<uint32_t> nodes;       // 8 words per node
  std::vector<uint32_t> transforms;  // 36 words per slot
  std::vector<uint32_t> index;
};

// Node span of the subtree at `n` (relative right child in word 6) and the triangles it covers.
uint64_t subtree_span(const uint32_t* nodes, uint64_t n) {
  uint64_t span = 1;
  while ((nodes[8 * n + 7] >> 16) != 0xFFFFu) {
    span += nodes[8 * n + 6];
    n += nodes[8 * n + 6];
  }
  return span;
}

uint32_t subtree_triangles(const uint32_t* nodes, uint64_t n) {
  uint64_t lo = n, hi = n;
  while ((nodes[8 * lo + 7] >> 16) != 0xFFFFu) lo += 1;
  while ((nodes[8 * hi + 7] >> 16) != 0xFFFFu) hi += nodes[8 * hi + 6];
  return nodes[8 * hi + 6] + (nodes[8 * hi + 7] & 0xFFFFu) - nodes[8 * lo + 6];
}

struct WebgpuPrim {
  uint32_t instance, root, node;  // node: index within the root's buffer
};

int pack_scene_webgpu(const mbvh_scene* sc, uint32_t primitive_limit, WebgpuPacking& out, const char* fn) {
  struct Host {
    bool ready = false;
    std::vector<uint32_t> nodes, vidx;
  };
  std::vector<Host> hm(sc->meshes.size());
  auto fetch = [&](uint32_t g) -> int {
    if (hm[g].ready) return MBVH_OK;
    const mbvh_mesh* m = sc->meshes[g];
    hm[g].nodes.resize(8 * m->n_ref_nodes);
    hm[g].vidx.resize(3 * m->n_slots);
    MBVH_HIP(hipMemcpy(hm[g].nodes.data(), m->d_ref_nodes, hm[g].nodes.size() * 4, hipMemcpyDeviceToHost));
    if (m->n_slots) MBVH_HIP(hipMemcpy(hm[g].vidx.data(), m->d_vidx, hm[g].vidx.size() * 4, hipMemcpyDeviceToHost));
    hm[g].ready = true;
    return MBVH_OK;
  };
  // primitives: (instance, root, cluster node) of every instance's non-empty roots
  std::vector<WebgpuPrim> prims;
  std::vector<float> prim_pos;
  std::vector<uint64_t> walk;
  int rc;
  for (uint32_t i = 0; i < sc->n_instances; i++) {
    const uint32_t g = sc->geom_h[i];
    if ((rc = fetch(g))) return rc;
    const mbvh_mesh* m = sc->meshes[g];
    for (uint32_t r = 0; r < m->root_base.size(); r++) {
      const uint32_t* rn = hm[g].nodes.data() + 8 * m->root_base[r];
      const float* rb = reinterpret_cast<const float*>(rn);
      if (!(rb[0] <= rb[3] && rb[1] <= rb[4] && rb[2] <= rb[5])) continue;  // no triangles
      walk.assign(1, 0);
      while (!walk.empty()) {
        const uint64_t n = walk.back();
        walk.pop_back();
        const bool leaf = (rn[8 * n + 7] >> 16) == 0xFFFFu;
        if (!leaf && primitive_limit != 0 && subtree_triangles(rn, n) >= primitive_limit) {
          walk.push_back(n + rn[8 * n + 6]);  // right below left: the left subtree comes first
          walk.push_back(n + 1);
          continue;
        }
        const float* b = rb + 8 * n;
        const double lb[6] = {b[0], b[1], b[2], b[3], b[4], b[5]};
        float fb[6];
        if (!world_box_of(sc->world_h.data() + 16ull * i, lb, fb))
          return set_error(MBVH_ERR_BAD_ARGS, "%s: instance %u has a non-finite world box", fn, i);
        prims.push_back({i, r, static_cast<uint32_t>(n)});
        const float v[9] = {fb[0], fb[1], fb[2], fb[3], fb[4], fb[5], fb[0], fb[1], fb[2]};
        prim_pos.insert(prim_pos.end(), v, v + 9);
      }
    }
  }
  const uint64_t np = prims.size();
  if (np >= (uint64_t{1} << 24)) return set_error(MBVH_ERR_LIMIT, "%s: at most 2^24 - 1 primitives", fn);
  std::vector<uint8_t> tlas(kBytesPerNode, 0);
  std::vector<uint32_t> order;  // primitive of each top-level slot
  if (np == 0) {  // one empty leaf
    const float empty[6] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                            std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                            -std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity()};
    const uint32_t leaf_word = 0xFFFF0000u;
    std::memcpy(tlas.data(), empty, sizeof(empty));
    std::memcpy(tlas.data() + 28, &leaf_word, 4);
  } else {
    std::vector<uint32_t> idx(3 * np);
    for (uint64_t k = 0; k < 3 * np; k++) idx[k] = static_cast<uint32_t>(k);
    const uint32_t range[2] = {0u, static_cast<uint32_t>(np)};
    const mbvh_build_opts opts{MBVH_SAH, 40, 1, 0, 1};  // one primitive per leaf (_strictLeafSize 1)
    mbvh_tree* t = nullptr;
    if ((rc = mbvh_build(prim_pos.data(), 3 * np, 3, 0, idx.data(), 4, idx.size(), range, 1, &opts, &t))) return rc;
    tlas = t->roots[0];
    mbvh_tree_free(t);
    order.resize(np);
    for (uint64_t k = 0; k < np; k++) order[k] = idx[3 * k] / 3;
  }
  const uint64_t n_tlas = tlas.size() / kBytesPerNode;
  const uint32_t* tw = reinterpret_cast<const uint32_t*>(tlas.data());
  for (uint64_t k = 0; k < n_tlas && np; k++)
    if ((tw[8 * k + 7] >> 16) == 0xFFFFu && (tw[8 * k + 7] & 0xFFFFu) != 1u)
      return set_error(MBVH_ERR_LIMIT, "%s: primitives with identical world boxes share a top-level leaf", fn);
  // subtrees and geometry in the order the top-level slots first name them
  struct Subtree {
    uint32_t mesh, root, node;
    uint64_t base, span;
  };
  std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint64_t> sub_of;  // (mesh, root, node) -> entry
  std::vector<Subtree> subs;
  std::vector<int64_t> tri_base(sc->meshes.size(), -1), vert_base(sc->meshes.size(), -1);
  std::vector<uint32_t> mesh_order;
  uint64_t next_node = n_tlas, next_tri = 0, next_vert = 0;
  std::vector<uint64_t> leaf_base(np);
  for (uint64_t k = 0; k < np; k++) {
    const WebgpuPrim& p = prims[order[k]];
    const uint32_t g = sc->geom_h[p.instance];
    const mbvh_mesh* m = sc->meshes[g];
    if (tri_base[g] < 0) {
      tri_base[g] = static_cast<int64_t>(next_tri);
      vert_base[g] = static_cast<int64_t>(next_vert);
      next_tri += m->n_slots;
      next_vert += m->vertex_count;
      mesh_order.push_back(g);
    }
    const auto key = std::make_tuple(g, p.root, p.node);
    auto it = sub_of.find(key);
    if (it == sub_of.end()) {
      const uint64_t span = subtree_span(hm[g].nodes.data() + 8 * m->root_base[p.root], p.node);
      it = sub_of.emplace(key, subs.size()).first;
      subs.push_back({g, p.root, p.node, next_node, span});
      next_node += span;
    }
    leaf_base[k] = subs[it->second].base;
  }
  if (next_node >= (uint64_t{1} << 32)) return set_error(MBVH_ERR_LIMIT, "%s: more than 2^32 nodes", fn);
  out.nodes.assign(8 * next_node, 0u);
  // the top level (appendBVHData)
  for (uint64_t k = 0; k < n_tlas; k++) {
    uint32_t* dst = out.nodes.data() + 8 * k;
    const uint32_t* src = tw + 8 * k;
    for (int c = 0; c < 6; c++) dst[c] = src[c];
    if (k == 0) {
      const float* bf = reinterpret_cast<const float*>(src);
      for (int c = 0; c < 3; c++)
        if (bf[c] > bf[c + 3]) {
          const float one = 1.0f, mone = -1.0f;
          std::memcpy(&dst[c], &one, 4);
          std::memcpy(&dst[c + 3], &mone, 4);
        }
    }
    if ((src[7] >> 16) == 0xFFFFu) {
      const uint32_t cnt = src[7] & 0xFFFFu;
      const uint32_t slot = cnt == 0 ? 0u : src[6];  // the slot is the leaf's primitive offset
      dst[6] = cnt == 0 ? 0u : static_cast<uint32_t>(leaf_base[slot]);
      dst[7] = 0xFF000000u | (slot & 0x00FFFFFFu);
    } else {
      dst[6] = src[6];
      dst[7] = src[7];
    }
  }
  // the subtrees (appendBVHSubtree)
  for (const Subtree& sb : subs) {
    const uint32_t g = sb.mesh;
    const mbvh_mesh* m = sc->meshes[g];
    const uint32_t* src0 = hm[g].nodes.data() + 8 * (m->root_base[sb.root] + sb.node);
    uint32_t* dst0 = out.nodes.data() + 8 * sb.base;
    for (uint64_t k = 0; k < sb.span; k++) {
      const uint32_t* src = src0 + 8 * k;
      uint32_t* dst = dst0 + 8 * k;
      for (int c = 0; c < 6; c++) dst[c] = src[c];
      if (k == 0) {
        const float* bf = reinterpret_cast<const float*>(src);
        for (int c = 0; c < 3; c++)
          if (bf[c] > bf[c + 3]) {
            const float one = 1.0f, mone = -1.0f;
            std::memcpy(&dst[c], &one, 4);
            std::memcpy(&dst[c + 3], &mone, 4);
          }
      }
      if ((src[7] >> 16) == 0xFFFFu) {
        dst[6] = src[6] + static_cast<uint32_t>(tri_base[g]);
        dst[7] = 0xFFFF0000u | (src[7] & 0xFFFFu);
      } else {
        dst[6] = src[6];
        dst[7] = src[7];
      }
    }
  }
  // transforms, one per slot
  out.transforms.assign(36 * std::max<uint64_t>(np, 1), 0u);
  for (uint64_t k = 0; k < np; k++) {
    const uint32_t inst = prims[order[k]].instance;
    std::memcpy(out.transforms.data() + 36 * k, sc->world_h.data() + 16ull * inst, 64);
    std::memcpy(out.transforms.data() + 36 * k + 16, sc->frame_h.data() + 16ull * inst, 64);
    out.transforms[36 * k + 32] = sc->visible_h[inst];
  }
  // the index, per mesh
  out.index.assign(3 * next_tri, 0u);
  for (uint32_t g : mesh_order) {
    const uint64_t base = 3 * static_cast<uint64_t>(tri_base[g]);
    for (uint64_t k = 0; k < hm[g].vidx.size(); k++)
      out.index[base + k] = hm[g].vidx[k] + static_cast<uint32_t>(vert_base[g]);
  }
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_scene_webgpu_sizes(mbvh_ctx* ctx, const mbvh_scene* scene, uint32_t primitive_limit,
                                       uint64_t* n_nodes, uint64_t* n_transforms, uint64_t* n_index) {
  clear_error();
  const char* fn = "mbvh_scene_webgpu_sizes";
  if (!valid_ctx(ctx) || !valid_scene(scene) || scene->ctx != ctx)
    return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context or scene handle", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  WebgpuPacking pk;
  if ((rc = pack_scene_webgpu(scene, primitive_limit, pk, fn))) return rc;
  if (n_nodes) *n_nodes = pk.nodes.size() / 8;
  if (n_transforms) *n_transforms = pk.transforms.size() / 36;
  if (n_index) *n_index = pk.index.size();
  return MBVH_OK;
}

extern "C" int mbvh_scene_export_webgpu(mbvh_ctx* ctx, const mbvh_scene* scene, uint32_t primitive_limit,
                                        uint32_t* nodes, uint64_t nodes_capacity,
                                        uint32_t* transforms, uint64_t transforms_capacity, uint32_t* index,
                                        uint64_t index_capacity) {
  clear_error();
  const char* fn = "mbvh_scene_export_webgpu";
  if (!valid_ctx(ctx) || !valid_scene(scene) || scene->ctx != ctx)
    return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context or scene handle", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  int rc = set_device(ctx);
  if (rc) return rc;
  WebgpuPacking pk;
  if ((rc = pack_scene_webgpu(scene, primitive_limit, pk, fn))) return rc;
  if ((nodes && nodes_capacity < pk.nodes.size()) || (transforms && transforms_capacity < pk.transforms.size()) ||
      (index && index_capacity < pk.index.size()))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: an output holds fewer words than mbvh_scene_webgpu_sizes gives", fn);
  if (nodes) std::memcpy(nodes, pk.nodes.data(), pk.nodes.size() * 4);
  if (transforms) std::memcpy(transforms, pk.transforms.data(), pk.transforms.size() * 4);
  if (index && !pk.index.empty()) std::memcpy(index, pk.index.data(), pk.index.size() * 4);
  return MBVH_OK;
}

extern "C" void mbvh_scene_free(mbvh_scene* scene) {
  if (!valid_scene(scene)) return;
  if (valid_ctx(scene->ctx)) {
    (void)hipSetDevice(scene->ctx->device);
    (void)hipStreamSynchronize(scene->ctx->stream);
  }
  free_scene_buffers(scene);
  scene->magic = 0;
  delete scene;
}

namespace {

int check_scene_call(const char* fn, const mbvh_ctx* ctx, const mbvh_scene* sc, int32_t side, uint64_t n,
                     const void* o, const void* d, const void* face) {
  if (!valid_ctx(ctx)) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid context handle", fn);
  if (!valid_scene(sc) || sc->ctx != ctx) return set_error(MBVH_ERR_BAD_HANDLE, "%s: invalid scene handle", fn);
  if (side < MBVH_FRONT_SIDE || side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: invalid side %d", fn, side);
  if (n && (!o || !d || !face)) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL ray or face buffer", fn);
  return MBVH_OK;
}

int enqueue_scene(mbvh_ctx* ctx, const mbvh_scene* sc, const float* o, const float* d, uint64_t n, int32_t side,
                  float near_d, float far_d, uint32_t* face, uint32_t* inst, float* dist, float* point, float* bary,
                  hipStream_t s) {
  if (sc->max_stack > 8 + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "scene: stack scratch smaller than the tree depth");
  SceneArgs a{};
  a.meshes = sc->d_meshes;
  a.n_meshes = sc->n_meshes;
  a.tlas = sc->d_tlas;
  a.tlas_inst = sc->d_tlas_inst;
  a.world = sc->d_world;
  a.inverse = sc->d_inverse;
  a.frame = sc->d_frame;
  a.frame_ok = sc->d_frame_ok;
  a.instance_geometry = sc->d_geom;
  a.origins = o;
  a.directions = d;
  a.n = n;
  a.side_mask = side_mask_of(side);
  a.near_dist = near_d;
  a.far_dist = far_d;
  a.out_face = face;
  a.out_instance = inst;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  a.stack = stack_of(ctx, threads);
  MBVH_HIP(launch_scene_raycast_first(a, threads, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_scene_raycast_first_d(mbvh_ctx* ctx, const mbvh_scene* scene, const float* origins,
                                          const float* directions, uint64_t n, int32_t side, float near_dist,
                                          float far_dist, uint32_t* face_index, uint32_t* instance, float* distance,
                                          float* point, float* barycoord, void* stream) {
  clear_error();
  const char* fn = "mbvh_scene_raycast_first_d";
  int rc = check_scene_call(fn, ctx, scene, side, n, origins, directions, face_index);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  return enqueue_scene(ctx, scene, origins, directions, n, side, near_dist, far_dist, face_index, instance, distance,
                       point, barycoord, stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_scene_raycast_first(mbvh_ctx* ctx, const mbvh_scene* scene, const float* origins,
                                        const float* directions, uint64_t n, int32_t side, float near_dist,
                                        float far_dist, uint32_t* face_index, uint32_t* instance, float* distance,
                                        float* point, float* barycoord) {
  clear_error();
  const char* fn = "mbvh_scene_raycast_first";
  int rc = check_scene_call(fn, ctx, scene, side, n, origins, directions, face_index);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  const size_t v3 = 3 * n * sizeof(float), s1 = n * sizeof(float);
  void *dO, *dD, *dF, *dI = nullptr, *dT = nullptr, *dP = nullptr, *dB = nullptr;
  if ((rc = stage_in(ctx, ctx->staging[0], origins, v3, &dO)) || (rc = stage_in(ctx, ctx->staging[1], directions, v3, &dD)) ||
      (rc = stage_out(ctx->staging[2], face_index, s1, &dF)) || (rc = stage_out(ctx->staging[3], instance, s1, &dI)) ||
      (rc = stage_out(ctx->staging[4], distance, s1, &dT)) || (rc = stage_out(ctx->staging[5], point, v3, &dP)) ||
      (rc = stage_out(ctx->staging[6], barycoord, v3, &dB)))
    return rc;
  if ((rc = enqueue_scene(ctx, scene, static_cast<const float*>(dO), static_cast<const float*>(dD), n, side, near_dist,
                          far_dist, static_cast<uint32_t*>(dF), static_cast<uint32_t*>(dI), static_cast<float*>(dT),
                          static_cast<float*>(dP), static_cast<float*>(dB), s)))
    return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, dF, s1, hipMemcpyDeviceToHost, s));
  if (instance) MBVH_HIP(hipMemcpyAsync(instance, dI, s1, hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, dT, s1, hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, dP, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, dB, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, fn);
}

// ---- scenes: closest point over instances -------------------------------------------------------------
namespace {

int enqueue_scene_points(mbvh_ctx* ctx, const mbvh_scene* sc, const float* points, uint64_t n, const uint32_t* order,
                         uint32_t* face, uint32_t* inst, float* dist, float* point, float* bary, hipStream_t s) {
  const bool quads = ctx->point_lanes != 1;  // four lanes per point unless one was asked for
  if (quads ? sc->max_stack_points4 > static_cast<uint32_t>(kQuadLdsStack) + ctx->stack_ovf_entries
            : sc->max_stack_points > 8 + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "scene: stack scratch smaller than the tree depth");
  SceneArgs a{};
  a.meshes = sc->d_meshes;
  a.n_meshes = sc->n_meshes;
  a.tlas = sc->d_tlas;
  a.tlas_inst = sc->d_tlas_inst;
  a.world = sc->d_world;
  a.inverse = sc->d_inverse;
  a.frame = sc->d_frame;
  a.frame_ok = sc->d_frame_ok;
  a.instance_geometry = sc->d_geom;
  a.origins = points;
  a.n = n;
  a.out_face = face;
  a.out_instance = inst;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  a.order = order;  // a spatial order lets the lanes of a wave share instances and tree paths
  if (quads) {
    const uint64_t threads = launch_threads(4 * n, ctx->max_threads);
    a.stack = stack_of(ctx, threads / 4);
    a.queue = static_cast<uint32_t*>(ctx->queue.ptr);
    // points whose walks run past kSceneHeavyBudget leaf steps go to one wave each (their walks can
    // be thousands of steps, which one quad would take alone while the launch waits for it)
    if (sc->wide_ok && ctx->heavy.ptr && n <= kHeavyCap) {
      a.heavy = static_cast<uint32_t*>(ctx->heavy.ptr);
      a.heavy_count = static_cast<uint32_t*>(ctx->queue.ptr) + 32;  // its own line of the counter block
      a.heavy_budget = kSceneHeavyBudget;
    }
    MBVH_HIP(launch_scene_closest_point4(a, threads, s));
    if (a.heavy) MBVH_HIP(launch_scene_closest_point_wide(a, kSceneWideWaves, s));
  } else {
    const uint64_t threads = launch_threads(n, ctx->max_threads);
    a.stack = stack_of(ctx, threads);
    MBVH_HIP(launch_scene_closest_point(a, threads, s));
  }
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_scene_closest_point_d(mbvh_ctx* ctx, const mbvh_scene* scene, const float* points, uint64_t n,
                                          const uint32_t* order, uint32_t* face_index, uint32_t* instance,
                                          float* distance, float* point, float* barycoord, void* stream) {
  clear_error();
  const char* fn = "mbvh_scene_closest_point_d";
  int rc = check_scene_call(fn, ctx, scene, MBVH_FRONT_SIDE, n, points, points, face_index);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  return enqueue_scene_points(ctx, scene, points, n, order, face_index, instance, distance, point, barycoord,
                              stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_scene_closest_point(mbvh_ctx* ctx, const mbvh_scene* scene, const float* points, uint64_t n,
                                        uint32_t* face_index, uint32_t* instance, float* distance, float* point,
                                        float* barycoord) {
  clear_error();
  const char* fn = "mbvh_scene_closest_point";
  int rc = check_scene_call(fn, ctx, scene, MBVH_FRONT_SIDE, n, points, points, face_index);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  const size_t v3 = 3 * n * sizeof(float), s1 = n * sizeof(float);
  void *dQ, *dF, *dI = nullptr, *dT = nullptr, *dP = nullptr, *dB = nullptr;
  if ((rc = stage_in(ctx, ctx->staging[0], points, v3, &dQ)) || (rc = stage_out(ctx->staging[2], face_index, s1, &dF)) ||
      (rc = stage_out(ctx->staging[3], instance, s1, &dI)) || (rc = stage_out(ctx->staging[4], distance, s1, &dT)) ||
      (rc = stage_out(ctx->staging[5], point, v3, &dP)) || (rc = stage_out(ctx->staging[6], barycoord, v3, &dB)))
    return rc;
  const uint32_t* order = nullptr;
  if ((rc = host_order(ctx, static_cast<const float*>(dQ), n, &order))) return rc;
  if ((rc = enqueue_scene_points(ctx, scene, static_cast<const float*>(dQ), n, order, static_cast<uint32_t*>(dF),
                                 static_cast<uint32_t*>(dI), static_cast<float*>(dT), static_cast<float*>(dP),
                                 static_cast<float*>(dB), s)))
    return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, dF, s1, hipMemcpyDeviceToHost, s));
  if (instance) MBVH_HIP(hipMemcpyAsync(instance, dI, s1, hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, dT, s1, hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, dP, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, dB, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, fn);
}

// ---- export to the reference's shader data layouts ---------------------------------------------------

namespace {

uint32_t ceil_sqrt(double x) { return static_cast<uint32_t>(std::ceil(std::sqrt(x))); }

// Texture sides: MeshBVHUniformStruct.js:146-150 (bounds 2 * ceil(sqrt(nodes / 2)), contents
// ceil(sqrt(nodes))), VertexAttributeTexture.js:251 (ceil(sqrt(count)) || 1).
void texture_sides(const mbvh_mesh* m, uint32_t* sides) {
  const double nodes = m->root_nodes.empty() ? 0.0 : static_cast<double>(m->root_nodes[0]);
  sides[0] = 2 * ceil_sqrt(nodes / 2);
  sides[1] = ceil_sqrt(nodes);
  sides[2] = std::max<uint32_t>(ceil_sqrt(static_cast<double>(m->n_slots)), 1u);
  sides[3] = std::max<uint32_t>(ceil_sqrt(static_cast<double>(m->vertex_count)), 1u);
}

int enqueue_export_textures(mbvh_mesh* m, const float* d_positions, uint32_t stride, uint32_t offset, float* bounds,
                            uint32_t* contents, uint32_t* index, float* position, hipStream_t s) {
  uint32_t sides[4];
  texture_sides(m, sides);
  ExportArgs a{};
  a.nodes = m->d_ref_nodes;
  a.n_nodes = m->root_nodes[0];
  a.bounds = bounds;
  a.bounds_texels = static_cast<uint64_t>(sides[0]) * sides[0];
  a.contents = contents;
  a.contents_texels = static_cast<uint64_t>(sides[1]) * sides[1];
  a.vidx = m->d_vidx;
  a.n_tris = m->n_slots;
  a.index = index;
  a.index_texels = static_cast<uint64_t>(sides[2]) * sides[2];
  a.positions = d_positions;
  a.stride = stride;
  a.offset = offset;
  a.n_vertices = m->vertex_count;
  a.position = d_positions ? position : nullptr;
  a.position_texels = static_cast<uint64_t>(sides[3]) * sides[3];
  MBVH_HIP(launch_export_textures(a, s));
  return MBVH_OK;
}

int check_export(const char* fn, mbvh_ctx* ctx, const mbvh_mesh* mesh, const void* bounds, const void* contents,
                 const void* positions, const void* position, uint32_t stride, uint32_t offset) {
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (mesh->root_nodes.size() != 1)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: Multi-root BVHs not supported (%zu roots)", fn, mesh->root_nodes.size());
  if (!bounds || !contents) return set_error(MBVH_ERR_BAD_ARGS, "%s: bounds and contents are required", fn);
  if (positions && (!position || stride < offset + 3))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: positions need a position texture and a stride >= offset + 3", fn);
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_mesh_texture_sizes(const mbvh_mesh* mesh, uint32_t* sides) {
  clear_error();
  if (!valid_mesh(mesh)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_mesh_texture_sizes: invalid mesh handle");
  if (!sides) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_texture_sizes: sides is NULL");
  texture_sides(mesh, sides);
  return MBVH_OK;
}

extern "C" int mbvh_mesh_export_textures_d(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions,
                                           uint32_t position_stride, uint32_t position_offset, float* bounds,
                                           uint32_t* contents, uint32_t* index, float* position, void* stream) {
  clear_error();
  const char* fn = "mbvh_mesh_export_textures_d";
  int rc = check_export(fn, ctx, mesh, bounds, contents, positions, position, position_stride, position_offset);
  if (rc) return rc;
  if ((rc = set_device(ctx))) return rc;
  return enqueue_export_textures(mesh, positions, position_stride, position_offset, bounds, contents, index, position,
                                 stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_mesh_export_textures(mbvh_ctx* ctx, mbvh_mesh* mesh, const float* positions,
                                         uint32_t position_stride, uint32_t position_offset, float* bounds,
                                         uint32_t* contents, uint32_t* index, float* position) {
  clear_error();
  const char* fn = "mbvh_mesh_export_textures";
  int rc = check_export(fn, ctx, mesh, bounds, contents, positions, position, position_stride, position_offset);
  if (rc) return rc;
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  uint32_t sides[4];
  texture_sides(mesh, sides);
  const size_t nb = 16ull * sides[0] * sides[0], nc = 8ull * sides[1] * sides[1];
  const size_t ni = 16ull * sides[2] * sides[2], np = 16ull * sides[3] * sides[3];
  const size_t pos_bytes = mesh->vertex_count ? ((mesh->vertex_count - 1) * position_stride + position_offset + 3) * 4 : 0;
  void *dB, *dC, *dI = nullptr, *dPin = nullptr, *dP = nullptr;
  if ((rc = stage_out(ctx->staging[2], bounds, nb, &dB)) || (rc = stage_out(ctx->staging[3], contents, nc, &dC)) ||
      (rc = stage_out(ctx->staging[4], index, ni, &dI)) ||
      (rc = stage_in(ctx, ctx->staging[0], pos_bytes ? positions : nullptr, pos_bytes, &dPin)) ||
      (rc = stage_out(ctx->staging[5], dPin ? position : nullptr, np, &dP)))
    return rc;
  if ((rc = enqueue_export_textures(mesh, static_cast<const float*>(dPin), position_stride, position_offset,
                                    static_cast<float*>(dB), static_cast<uint32_t*>(dC), static_cast<uint32_t*>(dI),
                                    static_cast<float*>(dP), s)))
    return rc;
  MBVH_HIP(hipMemcpyAsync(bounds, dB, nb, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipMemcpyAsync(contents, dC, nc, hipMemcpyDeviceToHost, s));
  if (index) MBVH_HIP(hipMemcpyAsync(index, dI, ni, hipMemcpyDeviceToHost, s));
  if (dP) MBVH_HIP(hipMemcpyAsync(position, dP, np, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

namespace {

int enqueue_export_nodes(mbvh_mesh* m, uint32_t* out, hipStream_t s) {
  MBVH_HIP(launch_export_nodes(m->d_ref_nodes, m->n_ref_nodes, m->d_root_first,
                               static_cast<uint32_t>(m->root_base.size()), out, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_mesh_export_nodes_d(mbvh_ctx* ctx, mbvh_mesh* mesh, uint32_t* nodes, uint64_t capacity_bytes,
                                        void* stream) {
  clear_error();
  const char* fn = "mbvh_mesh_export_nodes_d";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (!nodes || capacity_bytes < mesh->n_ref_nodes * kBytesPerNode)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: nodes must hold %llu bytes", fn,
                     static_cast<unsigned long long>(mesh->n_ref_nodes * kBytesPerNode));
  if ((rc = set_device(ctx))) return rc;
  return enqueue_export_nodes(mesh, nodes, stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_mesh_export_nodes(mbvh_ctx* ctx, mbvh_mesh* mesh, uint32_t* nodes, uint64_t capacity_bytes) {
  clear_error();
  const char* fn = "mbvh_mesh_export_nodes";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  const size_t bytes = mesh->n_ref_nodes * kBytesPerNode;
  if (!nodes || capacity_bytes < bytes)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: nodes must hold %llu bytes", fn, static_cast<unsigned long long>(bytes));
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  hipStream_t s = ctx->stream;
  void* d = nullptr;
  if ((rc = stage_out(ctx->staging[6], nodes, bytes, &d))) return rc;
  if ((rc = enqueue_export_nodes(mesh, static_cast<uint32_t*>(d), s))) return rc;
  MBVH_HIP(hipMemcpyAsync(nodes, d, bytes, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return MBVH_OK;
}

// ---- intersectsBox ---------------------------------------------------------------------------------

namespace {

int enqueue_boxes(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* mn, const double* mx, const double* mats,
                  uint64_t n, uint8_t* out, hipStream_t s) {
  if (mesh->max_stack > 8 + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "intersectsBox: stack scratch smaller than the tree depth");
  BoxArgs a{};
  a.mesh = view_of(mesh);
  a.box_min = mn;
  a.box_max = mx;
  a.matrices = mats;
  a.n = n;
  a.out = out;
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  a.stack = stack_of(ctx, threads);
  MBVH_HIP(launch_intersects_box(a, threads, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_intersects_box_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* box_min, const double* box_max,
                                     const double* box_to_mesh, uint64_t n, uint8_t* out, void* stream) {
  clear_error();
  const char* fn = "mbvh_intersects_box_d";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (n && (!box_min || !box_max || !box_to_mesh || !out))
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL box, matrix or output buffer", fn);
  if ((rc = set_device(ctx))) return rc;
  if ((rc = reserve_stack(ctx, mesh->max_stack))) return rc;
  return enqueue_boxes(ctx, mesh, box_min, box_max, box_to_mesh, n, out,
                       stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_intersects_box(mbvh_ctx* ctx, const mbvh_mesh* mesh, const double* box_min, const double* box_max,
                                   const double* box_to_mesh, uint64_t n, uint8_t* out) {
  clear_error();
  const char* fn = "mbvh_intersects_box";
  int rc = check_query(fn, ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!box_min || !box_max || !box_to_mesh || !out)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL box, matrix or output buffer", fn);
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  if ((rc = reserve_stack(ctx, mesh->max_stack))) return rc;
  hipStream_t s = ctx->stream;
  void *dmn, *dmx, *dm, *dout;
  if ((rc = stage_in(ctx, ctx->staging[0], box_min, 3 * n * sizeof(double), &dmn)) ||
      (rc = stage_in(ctx, ctx->staging[1], box_max, 3 * n * sizeof(double), &dmx)) ||
      (rc = stage_in(ctx, ctx->staging[2], box_to_mesh, 16 * n * sizeof(double), &dm)) ||
      (rc = stage_out(ctx->staging[7], out, n, &dout)))
    return rc;
  if ((rc = enqueue_boxes(ctx, mesh, static_cast<const double*>(dmn), static_cast<const double*>(dmx),
                          static_cast<const double*>(dm), n, static_cast<uint8_t*>(dout), s)))
    return rc;
  MBVH_HIP(hipMemcpyAsync(out, dout, n, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < n; i++)
    if (out[i] == 2)
      return set_error(MBVH_ERR_LIMIT, "%s: traversal stack overflow on box %llu", fn, static_cast<unsigned long long>(i));
  return MBVH_OK;
}


// ---- closestPointToGeometry ------------------------------------------------------------------------
namespace {

static_assert(sizeof(GeometryHit) == sizeof(mbvh_geometry_hit), "mbvh_geometry_hit layout");

// three's Matrix4.invert (the cofactor expansion; a zero determinant gives the zero matrix), in double.
void invert4(const double* te, double* out) {
  const double n11 = te[0], n21 = te[1], n31 = te[2], n41 = te[3];
  const double n12 = te[4], n22 = te[5], n32 = te[6], n42 = te[7];
  const double n13 = te[8], n23 = te[9], n33 = te[10], n43 = te[11];
  const double n14 = te[12], n24 = te[13], n34 = te[14], n44 = te[15];
  const double t11 = n23 * n34 * n42 - n24 * n33 * n42 + n24 * n32 * n43 - n22 * n34 * n43 - n23 * n32 * n44 + n22 * n33 * n44;
  const double t12 = n14 * n33 * n42 - n13 * n34 * n42 - n14 * n32 * n43 + n12 * n34 * n43 + n13 * n32 * n44 - n12 * n33 * n44;
  const double t13 = n13 * n24 * n42 - n14 * n23 * n42 + n14 * n22 * n43 - n12 * n24 * n43 - n13 * n22 * n44 + n12 * n23 * n44;
  const double t14 = n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 + n13 * n22 * n34 - n12 * n23 * n34;
  const double det = n11 * t11 + n21 * t12 + n31 * t13 + n41 * t14;
  if (det == 0) {
    for (int i = 0; i < 16; i++) out[i] = 0;
    return;
  }
  const double d = 1 / det;
  out[0] = t11 * d;
  out[1] = (n24 * n33 * n41 - n23 * n34 * n41 - n24 * n31 * n43 + n21 * n34 * n43 + n23 * n31 * n44 - n21 * n33 * n44) * d;
  out[2] = (n22 * n34 * n41 - n24 * n32 * n41 + n24 * n31 * n42 - n21 * n34 * n42 - n22 * n31 * n44 + n21 * n32 * n44) * d;
  out[3] = (n23 * n32 * n41 - n22 * n33 * n41 - n23 * n31 * n42 + n21 * n33 * n42 + n22 * n31 * n43 - n21 * n32 * n43) * d;
  out[4] = t12 * d;
  out[5] = (n13 * n34 * n41 - n14 * n33 * n41 + n14 * n31 * n43 - n11 * n34 * n43 - n13 * n31 * n44 + n11 * n33 * n44) * d;
  out[6] = (n14 * n32 * n41 - n12 * n34 * n41 - n14 * n31 * n42 + n11 * n34 * n42 + n12 * n31 * n44 - n11 * n32 * n44) * d;
  out[7] = (n12 * n33 * n41 - n13 * n32 * n41 + n13 * n31 * n42 - n11 * n33 * n42 - n12 * n31 * n43 + n11 * n32 * n43) * d;
  out[8] = t13 * d;
  out[9] = (n14 * n23 * n41 - n13 * n24 * n41 - n14 * n21 * n43 + n11 * n24 * n43 + n13 * n21 * n44 - n11 * n23 * n44) * d;
  out[10] = (n12 * n24 * n41 - n14 * n22 * n41 + n14 * n21 * n42 - n11 * n24 * n42 - n12 * n21 * n44 + n11 * n22 * n44) * d;
  out[11] = (n13 * n22 * n41 - n12 * n23 * n41 - n13 * n21 * n42 + n11 * n23 * n42 + n12 * n21 * n43 - n11 * n22 * n43) * d;
  out[12] = t14 * d;
  out[13] = (n13 * n24 * n31 - n14 * n23 * n31 + n14 * n21 * n33 - n11 * n24 * n33 - n13 * n21 * n34 + n11 * n23 * n34) * d;
  out[14] = (n14 * n22 * n31 - n12 * n24 * n31 - n14 * n21 * n32 + n11 * n24 * n32 + n12 * n21 * n34 - n11 * n22 * n34) * d;
  out[15] = (n12 * n23 * n31 - n13 * n22 * n31 + n13 * n21 * n32 - n11 * n23 * n32 - n12 * n21 * n33 + n11 * n22 * n33) * d;
}

int check_geometry(const char* fn, const float* positions, uint64_t n_vertices, uint32_t stride, uint32_t offset,
                   const void* index, uint32_t index_elt, uint64_t n_tris, const double* matrix, void* out) {
  if (!matrix || !out) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL matrix or output", fn);
  if (n_tris && !positions) return set_error(MBVH_ERR_BAD_ARGS, "%s: NULL positions", fn);
  if (stride < 3 || offset > stride - 3) return set_error(MBVH_ERR_BAD_ARGS, "%s: bad stride %u / offset %u", fn, stride, offset);
  if (index && index_elt != 2 && index_elt != 4)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: index element size must be 2 or 4", fn);
  if (n_tris >= 0xFFFFFFFFull) return set_error(MBVH_ERR_LIMIT, "%s: too many triangles", fn);
  if (!index && 3 * n_tris > n_vertices)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: %llu triangles need more than %llu vertices", fn,
                     static_cast<unsigned long long>(n_tris), static_cast<unsigned long long>(n_vertices));
  return MBVH_OK;
}

int enqueue_geometry(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions, uint64_t n_vertices, uint32_t stride,
                     uint32_t offset, const void* index, uint32_t index_elt, uint64_t n_tris, const double* matrix,
                     double max_threshold, mbvh_geometry_hit* out, hipStream_t s) {
  if (mesh->max_stack > 8 + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "closestPointToGeometry: stack scratch smaller than the tree depth");
  const size_t per = sizeof(double) + sizeof(uint32_t);
  int rc = ensure_buffer(ctx->geo_ws, 256 + n_tris * per);
  if (rc) return rc;
  uint8_t* ws = static_cast<uint8_t*>(ctx->geo_ws.ptr);
  GeometryArgs a{};
  a.mesh = view_of(mesh);
  a.other_positions = positions;
  a.other_vertices = n_vertices;
  a.other_stride = stride;
  a.other_offset = offset;
  a.other_index = index;
  a.other_elt = static_cast<int32_t>(index_elt);
  a.n_other = n_tris;
  for (int i = 0; i < 16; i++) a.matrix[i] = matrix[i];
  invert4(matrix, a.inverse);
  a.max_threshold = max_threshold;
  a.best_bits = reinterpret_cast<unsigned long long*>(ws);
  a.winner = reinterpret_cast<uint32_t*>(ws + 8);
  a.status = reinterpret_cast<uint32_t*>(ws + 12);
  a.lane_dist = reinterpret_cast<double*>(ws + 256);
  a.lane_face = reinterpret_cast<uint32_t*>(ws + 256 + n_tris * sizeof(double));
  a.out = reinterpret_cast<GeometryHit*>(out);
  const uint64_t threads = launch_threads(n_tris ? n_tris : 1, ctx->max_threads);
  a.stack = stack_of(ctx, threads);
  MBVH_HIP(launch_closest_to_geometry(a, threads, s));
  return MBVH_OK;
}

}  // namespace

extern "C" int mbvh_closest_point_to_geometry_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions,
                                                uint64_t n_vertices, uint32_t stride, uint32_t offset, const void* index,
                                                uint32_t index_elt, uint64_t n_tris, const double* geometry_to_bvh,
                                                double min_threshold, double max_threshold, mbvh_geometry_hit* out,
                                                void* stream) {
  clear_error();
  const char* fn = "mbvh_closest_point_to_geometry_d";
  (void)min_threshold;
  int rc = check_query(fn, ctx, mesh);
  if (rc || (rc = check_geometry(fn, positions, n_vertices, stride, offset, index, index_elt, n_tris, geometry_to_bvh, out)))
    return rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = reserve_stack(ctx, mesh->max_stack))) return rc;
  return enqueue_geometry(ctx, mesh, positions, n_vertices, stride, offset, index, index_elt, n_tris, geometry_to_bvh,
                          max_threshold, out, stream ? static_cast<hipStream_t>(stream) : ctx->stream);
}

extern "C" int mbvh_closest_point_to_geometry(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* positions,
                                              uint64_t n_vertices, uint32_t stride, uint32_t offset, const void* index,
                                              uint32_t index_elt, uint64_t n_tris, const double* geometry_to_bvh,
                                              double min_threshold, double max_threshold, mbvh_geometry_hit* out) {
  clear_error();
  const char* fn = "mbvh_closest_point_to_geometry";
  (void)min_threshold;
  int rc = check_query(fn, ctx, mesh);
  if (rc || (rc = check_geometry(fn, positions, n_vertices, stride, offset, index, index_elt, n_tris, geometry_to_bvh, out)))
    return rc;
  if (index)
    for (uint64_t i = 0; i < 3 * n_tris; i++) {
      const uint64_t v = index_elt == 2 ? static_cast<const uint16_t*>(index)[i] : static_cast<const uint32_t*>(index)[i];
      if (v >= n_vertices)
        return set_error(MBVH_ERR_BAD_ARGS, "%s: index[%llu] = %llu is not below the %llu vertices", fn,
                         static_cast<unsigned long long>(i), static_cast<unsigned long long>(v),
                         static_cast<unsigned long long>(n_vertices));
    }
  std::lock_guard<std::mutex> guard(ctx->lock);
  if ((rc = set_device(ctx))) return rc;
  if ((rc = reserve_stack(ctx, mesh->max_stack))) return rc;
  hipStream_t s = ctx->stream;
  const uint64_t n_floats = n_vertices ? (n_vertices - 1) * stride + offset + 3 : 0;
  void *dp = nullptr, *di = nullptr, *dout = nullptr;
  if ((rc = stage_in(ctx, ctx->staging[0], positions, n_floats * sizeof(float), &dp)) ||
      (index && (rc = stage_in(ctx, ctx->staging[1], index, 3 * n_tris * index_elt, &di))) ||
      (rc = stage_out(ctx->staging[7], out, sizeof(mbvh_geometry_hit), &dout)))
    return rc;
  if ((rc = enqueue_geometry(ctx, mesh, static_cast<const float*>(dp), n_vertices, stride, offset, di, index_elt, n_tris,
                             geometry_to_bvh, max_threshold, static_cast<mbvh_geometry_hit*>(dout), s)))
    return rc;
  MBVH_HIP(hipMemcpyAsync(out, dout, sizeof(mbvh_geometry_hit), hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  if (out->status == 1) return set_error(MBVH_ERR_LIMIT, "%s: traversal stack overflow", fn);
  if (out->status == 2) return set_error(MBVH_ERR_BAD_ARGS, "%s: an index entry is out of range", fn);
  return MBVH_OK;
}
