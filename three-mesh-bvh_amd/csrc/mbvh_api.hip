// C ABI runtime: device context, upload-once meshes and query dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "mbvh_common.h"
#include "mbvh_kernels.h"

using namespace mbvh;

namespace {

constexpr uint64_t kMaxThreadsPerCU = 2048;  // 32 waves of 64 lanes

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct mbvh_ctx {
  uint32_t magic = kCtxMagic;
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t max_threads = 0;     // resident-thread cap for one launch
  DeviceBuffer stack_ovf;       // traversal stack overflow region
  uint32_t stack_ovf_entries = 0;
  DeviceBuffer staging[8];      // host-API input / output staging
  DeviceBuffer hits[5];         // host-API all-hits results: offsets, face, dist, point, bary
  uint64_t hits_total = 0;
  std::mutex lock;              // host entry points on one context are serialised
};

struct mbvh_mesh {
  uint32_t magic = kMeshMagic;
  mbvh_ctx* ctx = nullptr;
  WideNode* d_nodes = nullptr;
  uint32_t* d_roots = nullptr;
  uint2* d_bigleaf = nullptr;
  float4* d_tris = nullptr;
  uint32_t* d_vidx = nullptr;
  uint64_t n_nodes = 0;
  uint32_t n_roots = 0;
  uint64_t n_slots = 0;
  uint64_t vertex_count = 0;
  uint64_t face_limit = 0;  // face ids are < face_limit
  uint32_t max_depth = 0;   // deepest reference node depth over all roots
  uint32_t max_stack = 0;   // traversal stack entries needed
  float coord_max = 0.0f;   // largest |vertex coordinate| over the mesh's triangles
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

uint32_t inline_leaf_ref(uint64_t off, uint64_t cnt, std::vector<uint2>& bigleaf) {
  if (cnt >= 1 && cnt <= kMaxInlineCount && off <= kRefOffsetMask)
    return kRefLeaf | (static_cast<uint32_t>(cnt) << kRefCountShift) | static_cast<uint32_t>(off);
  bigleaf.push_back(make_uint2(static_cast<uint32_t>(off), static_cast<uint32_t>(cnt)));
  return kRefLeaf | static_cast<uint32_t>(bigleaf.size() - 1);
}

// Converts one reference root into wide nodes. Validates every index it follows so a corrupt or
// foreign buffer is rejected on the host instead of steering a kernel out of bounds.
int convert_root(const NodeView& v, uint64_t n_slots, std::vector<WideNode>& wide,
                 std::vector<uint2>& bigleaf, uint32_t* root_ref, uint32_t* max_depth) {
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
  uint64_t visited = 0;
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    if (++visited > v.n_nodes)
      return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: node buffer contains a cycle");
    *max_depth = std::max(*max_depth, it.depth);
    uint32_t ref;
    if (v.is_leaf(it.node)) {
      if (!leaf_ok(it.node))
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: leaf %llu references triangles past the end",
                         static_cast<unsigned long long>(it.node));
      ref = inline_leaf_ref(v.leaf_offset(it.node), v.leaf_count(it.node), bigleaf);
    } else {
      const uint64_t left = it.node + 1, right = v.right_child(it.node);
      const uint32_t axis = v.split_axis(it.node);
      if (axis > 2 || left >= v.n_nodes || right >= v.n_nodes || right <= it.node)
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: malformed internal node %llu",
                         static_cast<unsigned long long>(it.node));
      if (wide.size() >= 0x7FFFFFFFull)
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

uint64_t read_u(const void* p, int elt_bytes, uint64_t i) {
  if (elt_bytes == 2) return static_cast<const uint16_t*>(p)[i];
  return static_cast<const uint32_t*>(p)[i];
}

void free_mesh_buffers(mbvh_mesh* m) {
  if (m->d_nodes) (void)hipFree(m->d_nodes);
  if (m->d_roots) (void)hipFree(m->d_roots);
  if (m->d_bigleaf) (void)hipFree(m->d_bigleaf);
  if (m->d_tris) (void)hipFree(m->d_tris);
  if (m->d_vidx) (void)hipFree(m->d_vidx);
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
  v.bigleaf = m->d_bigleaf;
  v.tris = m->d_tris;
  v.n_roots = m->n_roots;
  v.max_stack = m->max_stack;
  v.coord_max = m->coord_max;
  return v;
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
  mbvh_ctx* c = new (std::nothrow) mbvh_ctx();
  if (!c) return set_error(MBVH_ERR_OOM, "mbvh_ctx_create: out of host memory");
  c->device = device;
  c->max_threads = static_cast<uint64_t>(prop.multiProcessorCount) * kMaxThreadsPerCU;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(MBVH_ERR_HIP, "mbvh_ctx_create: hipStreamCreate failed (%s)", hipGetErrorString(e));
  }
  *out_ctx = c;
  return MBVH_OK;
}

extern "C" void mbvh_ctx_destroy(mbvh_ctx* ctx) {
  if (!valid_ctx(ctx)) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stack_ovf.ptr) (void)hipFree(ctx->stack_ovf.ptr);
  for (auto& b : ctx->staging)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto& b : ctx->hits)
    if (b.ptr) (void)hipFree(b.ptr);
  (void)hipStreamDestroy(ctx->stream);
  ctx->magic = 0;
  delete ctx;
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
  for (uint32_t r = 0; r < n_roots; r++) {
    if (!roots[r] || root_bytes[r] % kBytesPerNode != 0)
      return set_error(MBVH_ERR_BAD_ARGS, "mbvh_mesh_upload: root %u size is not a multiple of 32 bytes", r);
    NodeView v{static_cast<const uint8_t*>(roots[r]), root_bytes[r] / kBytesPerNode};
    int rc = convert_root(v, n_slots, wide, bigleaf, &root_refs[r], &max_depth);
    if (rc) return rc;
  }
  if (bigleaf.size() > kRefOffsetMask) return set_error(MBVH_ERR_LIMIT, "mbvh_mesh_upload: too many large leaves");

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
  if (max_stack > static_cast<uint32_t>(kStackLds)) {
    const uint32_t need = max_stack - kStackLds;
    if (need > ctx->stack_ovf_entries) {
      rc = ensure_buffer(ctx->stack_ovf, static_cast<size_t>(need) * ctx->max_threads * sizeof(uint2));
      if (rc) return rc;
      ctx->stack_ovf_entries = need;
    }
  }

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
  m->coord_max = coord_max;

  auto upload = [&](void** dst, const void* src, size_t bytes) -> int {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
      *dst = nullptr;
      return set_error(MBVH_ERR_OOM, "mbvh_mesh_upload: device allocation of %zu bytes failed", bytes);
    }
    if (bytes) {
      e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) return set_error(MBVH_ERR_HIP, "mbvh_mesh_upload: copy failed (%s)", hipGetErrorString(e));
    }
    return MBVH_OK;
  };
  rc = upload(reinterpret_cast<void**>(&m->d_nodes), wide.data(), wide.size() * sizeof(WideNode));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_roots), root_refs.data(), root_refs.size() * sizeof(uint32_t));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_bigleaf), bigleaf.data(), bigleaf.size() * sizeof(uint2));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_tris), tris.data(), tris.size() * sizeof(float4));
  if (!rc) rc = upload(reinterpret_cast<void**>(&m->d_vidx), vidx.data(), vidx.size() * sizeof(uint32_t));
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

// ---------------------------------------------------------------------------------------------
// Queries
// ---------------------------------------------------------------------------------------------

namespace {

int raycast_first_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* o, const float* d,
                       uint64_t n, int32_t side, float near_dist, float far_dist, uint32_t* face,
                       float* dist, float* point, float* bary, hipStream_t s) {
  if (side < MBVH_FRONT_SIDE || side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "raycast_first: invalid side %d", side);
  if (n && (!o || !d || !face)) return set_error(MBVH_ERR_BAD_ARGS, "raycast_first: NULL ray or face buffer");
  if (mesh->max_stack > kStackLds + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "raycast_first: stack scratch smaller than the tree depth");
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  RaycastFirstArgs a;
  a.mesh = view_of(mesh);
  a.stack = stack_of(ctx, threads);
  a.origins = o;
  a.directions = d;
  a.n = n;
  a.side_mask = side_mask_of(side);
  a.near_dist = near_dist;
  a.far_dist = far_dist;
  a.out_face = face;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  MBVH_HIP(launch_raycast_first(a, threads, s));
  return MBVH_OK;
}

int closest_point_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* q, uint64_t n, float min_t,
                       float max_t, uint32_t* face, float* dist, float* point, float* bary, hipStream_t s) {
  if (n && (!q || !face)) return set_error(MBVH_ERR_BAD_ARGS, "closest_point: NULL point or face buffer");
  if (mesh->max_stack > kStackLds + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "closest_point: stack scratch smaller than the tree depth");
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  ClosestPointArgs a;
  a.mesh = view_of(mesh);
  a.stack = stack_of(ctx, threads);
  a.points = q;
  a.n = n;
  a.min_threshold = min_t;
  a.max_threshold = max_t;
  a.out_face = face;
  a.out_dist = dist;
  a.out_point = point;
  a.out_bary = bary;
  MBVH_HIP(launch_closest_point(a, threads, s));
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
                                    const float* directions, uint64_t n, int32_t side, float near_dist,
                                    float far_dist, uint32_t* face_index, float* distance, float* point,
                                    float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_raycast_first_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return raycast_first_impl(ctx, mesh, origins, directions, n, side, near_dist, far_dist, face_index,
                            distance, point, barycoord, s);
}

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
  std::lock_guard<std::mutex> guard(ctx->lock);
  rc = set_device(ctx);
  if (rc) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[1], v3)) ||
      (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->staging[3], n * sizeof(float))) || (rc = ensure_buffer(ctx->staging[4], v3)) ||
      (rc = ensure_buffer(ctx->staging[5], v3)))
    return rc;
  hipStream_t s = ctx->stream;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, origins, v3, hipMemcpyHostToDevice, s));
  MBVH_HIP(hipMemcpyAsync(ctx->staging[1].ptr, directions, v3, hipMemcpyHostToDevice, s));
  rc = raycast_first_impl(ctx, mesh, static_cast<float*>(ctx->staging[0].ptr),
                          static_cast<float*>(ctx->staging[1].ptr), n, side, near_dist, far_dist,
                          static_cast<uint32_t*>(ctx->staging[2].ptr), static_cast<float*>(ctx->staging[3].ptr),
                          static_cast<float*>(ctx->staging[4].ptr), static_cast<float*>(ctx->staging[5].ptr), s);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, ctx->staging[3].ptr, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (point) MBVH_HIP(hipMemcpyAsync(point, ctx->staging[4].ptr, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, ctx->staging[5].ptr, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, "mbvh_raycast_first");
}

extern "C" int mbvh_closest_point_d(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                    float min_threshold, float max_threshold, uint32_t* face_index,
                                    float* distance, float* closest, float* barycoord, void* stream) {
  clear_error();
  int rc = check_query("mbvh_closest_point_d", ctx, mesh);
  if (rc) return rc;
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return closest_point_impl(ctx, mesh, points, n, min_threshold, max_threshold, face_index, distance,
                            closest, barycoord, s);
}

extern "C" int mbvh_closest_point(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* points, uint64_t n,
                                  float min_threshold, float max_threshold, uint32_t* face_index,
                                  float* distance, float* closest, float* barycoord) {
  clear_error();
  int rc = check_query("mbvh_closest_point", ctx, mesh);
  if (rc) return rc;
  if (n == 0) return MBVH_OK;
  if (!points || !face_index) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_closest_point: NULL point or face buffer");
  std::lock_guard<std::mutex> guard(ctx->lock);
  rc = set_device(ctx);
  if (rc) return rc;
  const size_t v3 = n * 3 * sizeof(float);
  if ((rc = ensure_buffer(ctx->staging[0], v3)) || (rc = ensure_buffer(ctx->staging[2], n * sizeof(uint32_t))) ||
      (rc = ensure_buffer(ctx->staging[3], n * sizeof(float))) || (rc = ensure_buffer(ctx->staging[4], v3)) ||
      (rc = ensure_buffer(ctx->staging[5], v3)))
    return rc;
  hipStream_t s = ctx->stream;
  MBVH_HIP(hipMemcpyAsync(ctx->staging[0].ptr, points, v3, hipMemcpyHostToDevice, s));
  rc = closest_point_impl(ctx, mesh, static_cast<float*>(ctx->staging[0].ptr), n, min_threshold, max_threshold,
                          static_cast<uint32_t*>(ctx->staging[2].ptr), static_cast<float*>(ctx->staging[3].ptr),
                          static_cast<float*>(ctx->staging[4].ptr), static_cast<float*>(ctx->staging[5].ptr), s);
  if (rc) return rc;
  MBVH_HIP(hipMemcpyAsync(face_index, ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (distance) MBVH_HIP(hipMemcpyAsync(distance, ctx->staging[3].ptr, n * sizeof(float), hipMemcpyDeviceToHost, s));
  if (closest) MBVH_HIP(hipMemcpyAsync(closest, ctx->staging[4].ptr, v3, hipMemcpyDeviceToHost, s));
  if (barycoord) MBVH_HIP(hipMemcpyAsync(barycoord, ctx->staging[5].ptr, v3, hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  return check_no_overflow(face_index, n, "mbvh_closest_point");
}

// ---------------------------------------------------------------------------------------------
// All hits
// ---------------------------------------------------------------------------------------------

namespace {

int raycast_all_impl(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* o, const float* d, uint64_t n,
                     int32_t side, float near_dist, float far_dist, uint32_t* counts, const uint64_t* offsets,
                     uint64_t capacity, uint32_t* face, float* dist, float* point, float* bary, bool fill,
                     hipStream_t s) {
  if (side < MBVH_FRONT_SIDE || side > MBVH_SIDE_PER_TRIANGLE)
    return set_error(MBVH_ERR_BAD_ARGS, "raycast: invalid side %d", side);
  if (n && (!o || !d)) return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL ray buffer");
  if (n && !fill && !counts) return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL counts buffer");
  if (n && fill && (!offsets || (capacity && !face)))
    return set_error(MBVH_ERR_BAD_ARGS, "raycast: NULL offsets or face buffer");
  if (mesh->max_stack > kStackLds + ctx->stack_ovf_entries)
    return set_error(MBVH_ERR_LIMIT, "raycast: stack scratch smaller than the tree depth");
  const uint64_t threads = launch_threads(n, ctx->max_threads);
  RaycastAllArgs a;
  a.mesh = view_of(mesh);
  a.stack = stack_of(ctx, threads);
  a.origins = o;
  a.directions = d;
  a.n = n;
  a.side_mask = side_mask_of(side);
  a.near_dist = near_dist;
  a.far_dist = far_dist;
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

extern "C" int mbvh_raycast(mbvh_ctx* ctx, const mbvh_mesh* mesh, const float* origins, const float* directions,
                            uint64_t n, int32_t side, float near_dist, float far_dist, uint64_t* offsets) {
  clear_error();
  int rc = check_query("mbvh_raycast", ctx, mesh);
  if (rc) return rc;
  if (!offsets) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast: offsets is NULL");
  std::lock_guard<std::mutex> guard(ctx->lock);
  ctx->hits_total = 0;
  offsets[0] = 0;
  if (n == 0) return MBVH_OK;
  if (!origins || !directions) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_raycast: NULL ray buffer");
  rc = set_device(ctx);
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
  rc = raycast_all_impl(ctx, mesh, dO, dD, n, side, near_dist, far_dist, static_cast<uint32_t*>(ctx->staging[2].ptr),
                        nullptr, 0, nullptr, nullptr, nullptr, nullptr, false, s);
  if (rc) return rc;
  std::vector<uint32_t> counts(n);
  MBVH_HIP(hipMemcpyAsync(counts.data(), ctx->staging[2].ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  MBVH_HIP(hipStreamSynchronize(s));
  uint64_t total = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (counts[i] == MBVH_NO_HIT - 1u)
      return set_error(MBVH_ERR_LIMIT, "mbvh_raycast: traversal stack overflow on ray %llu",
                       static_cast<unsigned long long>(i));
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
                        static_cast<float*>(ctx->hits[4].ptr), true, s);
  if (rc) return rc;
  MBVH_HIP(hipStreamSynchronize(s));
  ctx->hits_total = total;
  return MBVH_OK;
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
