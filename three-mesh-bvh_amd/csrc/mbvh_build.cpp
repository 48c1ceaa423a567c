// Host-side BVH construction in the reference's packed node format.
//
// This produces byte-for-byte the buffers MeshBVH builds (`_roots`, and the permuted index or the
// indirect buffer) so they can be serialized, handed back to JavaScript and uploaded to the GPU
// as-is. The steps and their floating-point storage points follow the reference:
//   triangle bounds      MeshBVH.computePrimitiveBounds   src/core/MeshBVH.js:337-436
//   node bounds          getBounds                        src/core/build/computeBoundsUtils.js:4-68
//   split choice         getOptimalSplit (CENTER/AVERAGE/SAH, 32 bins, exhaustive below 8 prims)
//                                                         src/core/build/splitUtils.js:21-307
//   partition            Hoare partition of the primitive buffer + bounds
//                                                         src/core/build/sortUtils.js:4-62
//   recursion            buildTree / splitNode            src/core/build/buildTree.js:10-121
//   packing              countNodes / populateBuffer      src/core/build/buildUtils.js:6-91
// Arithmetic is IEEE double where the reference uses JS numbers and float32 wherever it stores
// into a Float32Array, which is what makes the output bit-identical. This file is compiled with
// -ffp-contract=off so no fused multiply-add changes a rounding.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>

#include "mbvh_common.h"

namespace mbvh {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

}  // namespace mbvh

extern "C" const char* mbvh_last_error(void) { return mbvh::g_last_error.c_str(); }
extern "C" uint32_t mbvh_abi_version(void) { return MBVH_ABI_VERSION; }

namespace mbvh {
namespace {

constexpr int kBinCount = 32;                      // splitUtils.js:4
constexpr double kPrimitiveIntersectCost = 1.25;   // Constants.js:56
constexpr double kTraversalCost = 1.0;             // Constants.js:57
constexpr double kFloat32Epsilon = 5.9604644775390625e-08;  // 2^-24, Constants.js:71
constexpr double kInf = std::numeric_limits<double>::infinity();

// JavaScript `~~x` (ToInt32) for the bin index computation in splitUtils.js:207.
int32_t js_to_int32(double x) {
  if (!std::isfinite(x)) return 0;
  double t = std::trunc(x);
  double m = std::fmod(t, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  uint32_t u = static_cast<uint32_t>(m);
  return static_cast<int32_t>(u);
}

// Box helpers over float[6] = min xyz, max xyz (ArrayBoxUtilities.js).
void make_empty(float* b) {
  b[0] = b[1] = b[2] = std::numeric_limits<float>::infinity();
  b[3] = b[4] = b[5] = -std::numeric_limits<float>::infinity();
}

double surface_area(const float* b) {  // ArrayBoxUtilities.js:100-108
  const double d0 = static_cast<double>(b[3]) - b[0];
  const double d1 = static_cast<double>(b[4]) - b[1];
  const double d2 = static_cast<double>(b[5]) - b[2];
  return 2 * (d0 * d1 + d1 * d2 + d2 * d0);
}

int longest_edge(const float* b) {  // ArrayBoxUtilities.js:22-41
  int idx = -1;
  double best = -kInf;
  for (int i = 0; i < 3; i++) {
    const double dist = static_cast<double>(b[i + 3]) - b[i];
    if (dist > best) {
      best = dist;
      idx = i;
    }
  }
  return idx;
}

void union_into(const float* a, const float* b, float* out) {  // ArrayBoxUtilities.js:51-70
  for (int d = 0; d < 3; d++) {
    out[d] = a[d] < b[d] ? a[d] : b[d];
    out[d + 3] = a[d + 3] > b[d + 3] ? a[d + 3] : b[d + 3];
  }
}

// Primitive buffer whose triangle records are swapped by the partition: the geometry index
// (3 entries per triangle) or the indirect buffer (1 entry per triangle).
struct PrimitiveBuffer {
  uint8_t* data = nullptr;
  int elt_bytes = 4;
  uint32_t stride = 3;

  void swap_records(uint64_t i, uint64_t j) {
    const size_t rec = static_cast<size_t>(elt_bytes) * stride;
    uint8_t tmp[12];
    std::memcpy(tmp, data + i * rec, rec);
    std::memcpy(data + i * rec, data + j * rec, rec);
    std::memcpy(data + j * rec, tmp, rec);
  }
};

struct BuildNode {
  float box[6];
  uint64_t offset = 0;
  uint64_t count = 0;
  int32_t left = -1;
  int32_t right = -1;
  int32_t axis = -1;
  bool leaf = false;
};

struct SahBin {
  double candidate = 0;
  uint64_t count = 0;
  float bounds[6];
  float left_cache[6];
  float right_cache[6];
};

class TreeBuilder {
 public:
  TreeBuilder(const mbvh_build_opts& opts, PrimitiveBuffer prims, std::vector<float>& prim_bounds,
              uint64_t bounds_base)
      : opts_(opts), prims_(prims), pb_(prim_bounds), base_(bounds_base) {}

  // Builds one root over triangles [offset, offset + count) and returns it packed.
  int build_root(uint64_t offset, uint64_t count, std::vector<uint8_t>* out) {
    nodes_.clear();
    nodes_.reserve(static_cast<size_t>(count / 2 + 2));
    nodes_.emplace_back();
    float centroid[6];
    node_bounds(offset, count, nodes_[0].box, centroid);
    int rc = split(0, offset, count, centroid, 0);
    if (rc != MBVH_OK) return rc;
    return pack(out);
  }

 private:
  // Reads primitive record i as (center, half-extent) per axis.
  float pcenter(uint64_t i, int axis) const { return pb_[(i - base_) * 6 + 2 * axis]; }
  float phalf(uint64_t i, int axis) const { return pb_[(i - base_) * 6 + 2 * axis + 1]; }

  // computeBoundsUtils.js:4-68: union of primitive boxes and of their centers.
  void node_bounds(uint64_t offset, uint64_t count, float* box, float* centroid) const {
    double mn[3] = {kInf, kInf, kInf}, mx[3] = {-kInf, -kInf, -kInf};
    double cmn[3] = {kInf, kInf, kInf}, cmx[3] = {-kInf, -kInf, -kInf};
    for (uint64_t i = offset; i < offset + count; i++) {
      for (int k = 0; k < 3; k++) {
        const double c = pcenter(i, k), h = phalf(i, k);
        const double lo = c - h, hi = c + h;
        if (lo < mn[k]) mn[k] = lo;
        if (hi > mx[k]) mx[k] = hi;
        if (c < cmn[k]) cmn[k] = c;
        if (c > cmx[k]) cmx[k] = c;
      }
    }
    for (int k = 0; k < 3; k++) {
      box[k] = static_cast<float>(mn[k]);
      box[k + 3] = static_cast<float>(mx[k]);
      centroid[k] = static_cast<float>(cmn[k]);
      centroid[k + 3] = static_cast<float>(cmx[k]);
    }
  }

  // ArrayBoxUtilities.js:73-97: grows a float32 box by primitive i.
  void grow(uint64_t i, float* b) const {
    for (int d = 0; d < 3; d++) {
      const double c = pcenter(i, d), h = phalf(i, d);
      const double lo = c - h, hi = c + h;
      if (lo < b[d]) b[d] = static_cast<float>(lo);
      if (hi > b[d + 3]) b[d + 3] = static_cast<float>(hi);
    }
  }

  struct Split {
    int axis;
    double pos;
  };

  Split choose_split(const float* box, const float* centroid, uint64_t offset,
                     uint64_t count) {
    Split s{-1, 0.0};
    if (opts_.strategy == MBVH_CENTER) {
      s.axis = longest_edge(centroid);
      if (s.axis != -1)
        s.pos = (static_cast<double>(centroid[s.axis]) + centroid[s.axis + 3]) / 2;
    } else if (opts_.strategy == MBVH_AVERAGE) {
      s.axis = longest_edge(box);
      if (s.axis != -1) {
        double sum = 0;
        for (uint64_t i = offset; i < offset + count; i++) sum += pcenter(i, s.axis);
        s.pos = sum / static_cast<double>(count);
      }
    } else {
      sah_split(box, centroid, offset, count, &s);
    }
    return s;
  }

  // SAH split search, splitUtils.js:43-284.
  void sah_split(const float* box, const float* centroid, uint64_t offset, uint64_t count,
                 Split* s) {
    const double root_area = surface_area(box);
    double best_cost = kPrimitiveIntersectCost * static_cast<double>(count);
    const double dcount = static_cast<double>(count);

    for (int a = 0; a < 3; a++) {
      const double axis_left = centroid[a];
      const double axis_right = centroid[a + 3];
      const double axis_length = axis_right - axis_left;
      const double bin_width = axis_length / kBinCount;

      if (count < kBinCount / 4) {
        // Fewer primitives than 8: every primitive center is a candidate plane.
        int n = static_cast<int>(count);
        for (int b = 0; b < n; b++) {
          SahBin& bin = bins_[b];
          bin.candidate = pcenter(offset + b, a);
          bin.count = 0;
          make_empty(bin.bounds);
          make_empty(bin.left_cache);
          make_empty(bin.right_cache);
          grow(offset + b, bin.bounds);
        }
        // Sort by candidate; bins with equal candidates carry identical state, so ordering among
        // them does not matter, and the duplicates are dropped right after.
        SahBin* order[kBinCount / 4];
        for (int b = 0; b < n; b++) order[b] = &bins_[b];
        for (int i = 1; i < n; i++) {
          SahBin* key = order[i];
          int j = i - 1;
          while (j >= 0 && order[j]->candidate > key->candidate) {
            order[j + 1] = order[j];
            j--;
          }
          order[j + 1] = key;
        }
        int split_count = 0;
        for (int b = 0; b < n; b++) {
          if (split_count > 0 && order[split_count - 1]->candidate == order[b]->candidate) continue;
          order[split_count++] = order[b];
        }
        for (uint64_t i = offset; i < offset + count; i++) {
          const double center = pcenter(i, a);
          for (int b = 0; b < split_count; b++) {
            SahBin& bin = *order[b];
            if (center >= bin.candidate) {
              grow(i, bin.right_cache);
            } else {
              grow(i, bin.left_cache);
              bin.count++;
            }
          }
        }
        for (int b = 0; b < split_count; b++) {
          const SahBin& bin = *order[b];
          const double lc = static_cast<double>(bin.count);
          const double rc = dcount - lc;
          const double lp = bin.count != 0 ? surface_area(bin.left_cache) / root_area : 0.0;
          const double rp = rc != 0 ? surface_area(bin.right_cache) / root_area : 0.0;
          const double cost = kTraversalCost + kPrimitiveIntersectCost * (lp * lc + rp * rc);
          if (cost < best_cost) {
            s->axis = a;
            best_cost = cost;
            s->pos = bin.candidate;
          }
        }
      } else {
        for (int i = 0; i < kBinCount; i++) {
          SahBin& bin = bins_[i];
          bin.count = 0;
          bin.candidate = axis_left + bin_width + i * bin_width;
          make_empty(bin.bounds);
        }
        for (uint64_t i = offset; i < offset + count; i++) {
          const double rel = static_cast<double>(pcenter(i, a)) - axis_left;
          int32_t idx = js_to_int32(rel / bin_width);
          if (idx >= kBinCount) idx = kBinCount - 1;
          // rel >= 0 always holds (centers are bounded below by the centroid box); the guard
          // keeps a NaN-poisoned input from indexing outside the bins.
          if (idx < 0) idx = 0;
          bins_[idx].count++;
          grow(i, bins_[idx].bounds);
        }
        std::memcpy(bins_[kBinCount - 1].right_cache, bins_[kBinCount - 1].bounds, sizeof(float) * 6);
        for (int i = kBinCount - 2; i >= 0; i--)
          union_into(bins_[i].bounds, bins_[i + 1].right_cache, bins_[i].right_cache);

        uint64_t left_count = 0;
        for (int i = 0; i < kBinCount - 1; i++) {
          const SahBin& bin = bins_[i];
          if (bin.count != 0) {
            if (left_count == 0)
              std::memcpy(left_bounds_, bin.bounds, sizeof(float) * 6);
            else
              union_into(bin.bounds, left_bounds_, left_bounds_);
          }
          left_count += bin.count;
          const double lc = static_cast<double>(left_count);
          const double rc = dcount - lc;
          const double lp = left_count != 0 ? surface_area(left_bounds_) / root_area : 0.0;
          const double rp = rc != 0 ? surface_area(bins_[i + 1].right_cache) / root_area : 0.0;
          const double cost = kTraversalCost + kPrimitiveIntersectCost * (lp * lc + rp * rc);
          if (cost < best_cost) {
            s->axis = a;
            best_cost = cost;
            s->pos = bin.candidate;
          }
        }
      }
    }
  }

  // sortUtils.js:4-62. Centers equal to the plane go right.
  uint64_t partition(uint64_t offset, uint64_t count, const Split& s) {
    int64_t left = static_cast<int64_t>(offset);
    int64_t right = static_cast<int64_t>(offset + count) - 1;
    while (true) {
      while (left <= right && static_cast<double>(pcenter(left, s.axis)) < s.pos) left++;
      while (left <= right && static_cast<double>(pcenter(right, s.axis)) >= s.pos) right--;
      if (left < right) {
        prims_.swap_records(left, right);
        float* l = &pb_[(left - base_) * 6];
        float* r = &pb_[(right - base_) * 6];
        for (int i = 0; i < 6; i++) {
          const float t = l[i];
          l[i] = r[i];
          r[i] = t;
        }
        left++;
        right--;
      } else {
        return static_cast<uint64_t>(left);
      }
    }
  }

  // buildTree.js:44-119.
  int split(int32_t node, uint64_t offset, uint64_t count, const float* centroid, int depth) {
    // _strictLeafSize (buildTree.js:59-94): an over-full node is split whatever the heuristic says
    const bool must_split = opts_.strict_leaf_size > 0 && count > static_cast<uint64_t>(opts_.strict_leaf_size);
    if ((count <= static_cast<uint64_t>(opts_.target_leaf_size) && !must_split) || depth >= opts_.max_depth) {
      return make_leaf(node, offset, count);
    }
    Split s = choose_split(nodes_[node].box, centroid, offset, count);
    uint64_t split_at = 0;
    if (s.axis != -1) split_at = partition(offset, count, s);
    if (s.axis == -1 || split_at == offset || split_at == offset + count) {
      if (!must_split) return make_leaf(node, offset, count);
      // no usable split: the median of the current order, on the node box's longest axis
      s.axis = std::max(0, longest_edge(nodes_[node].box));
      split_at = offset + std::max<uint64_t>(1, count / 2);
    }
    nodes_[node].axis = s.axis;

    const int32_t left = static_cast<int32_t>(nodes_.size());
    nodes_.emplace_back();
    nodes_[node].left = left;
    float child_centroid[6];
    node_bounds(offset, split_at - offset, nodes_[left].box, child_centroid);
    int rc = split(left, offset, split_at - offset, child_centroid, depth + 1);
    if (rc != MBVH_OK) return rc;

    const int32_t right = static_cast<int32_t>(nodes_.size());
    nodes_.emplace_back();
    nodes_[node].right = right;
    node_bounds(split_at, offset + count - split_at, nodes_[right].box, child_centroid);
    return split(right, split_at, offset + count - split_at, child_centroid, depth + 1);
  }

  int make_leaf(int32_t node, uint64_t offset, uint64_t count) {
    if (count > 0xFFFF)
      return set_error(MBVH_ERR_LIMIT,
                       "mbvh_build: a leaf would hold %llu triangles; the node format stores at "
                       "most 65535 (raise max_depth)",
                       static_cast<unsigned long long>(count));
    if (offset > 0xFFFFFFFFull)
      return set_error(MBVH_ERR_LIMIT, "mbvh_build: leaf offset exceeds 32 bits");
    nodes_[node].leaf = true;
    nodes_[node].offset = offset;
    nodes_[node].count = count;
    return MBVH_OK;
  }

  // buildUtils.js:35-91: depth-first, left child immediately after its parent, right child
  // addressed by its node distance from the parent.
  int pack(std::vector<uint8_t>* out) {
    out->assign(nodes_.size() * kBytesPerNode, 0);
    uint64_t next = 0;
    // Pre-order slots (left subtree first), then one pass writing every record.
    std::vector<uint64_t> slot_of(nodes_.size(), 0);
    std::vector<int32_t> order;
    order.reserve(nodes_.size());
    std::vector<int32_t> dfs{0};
    while (!dfs.empty()) {
      const int32_t n = dfs.back();
      dfs.pop_back();
      slot_of[n] = next++;
      order.push_back(n);
      if (!nodes_[n].leaf) {
        dfs.push_back(nodes_[n].right);
        dfs.push_back(nodes_[n].left);
      }
    }
    for (int32_t n : order) {
      const BuildNode& bn = nodes_[n];
      uint8_t* rec = out->data() + slot_of[n] * kBytesPerNode;
      std::memcpy(rec, bn.box, sizeof(float) * 6);
      if (bn.leaf) {
        const uint32_t off = static_cast<uint32_t>(bn.offset);
        const uint16_t cnt = static_cast<uint16_t>(bn.count);
        const uint16_t flag = kLeafFlag;
        std::memcpy(rec + 24, &off, 4);
        std::memcpy(rec + 28, &cnt, 2);
        std::memcpy(rec + 30, &flag, 2);
      } else {
        const uint64_t rel = slot_of[bn.right] - slot_of[n];
        if (rel > 0xFFFFFFFFull)
          return set_error(MBVH_ERR_LIMIT,
                           "mbvh_build: cannot store relative child node offset greater than 32 bits");
        const uint32_t rel32 = static_cast<uint32_t>(rel);
        const uint32_t axis = static_cast<uint32_t>(bn.axis);
        std::memcpy(rec + 24, &rel32, 4);
        std::memcpy(rec + 28, &axis, 4);
      }
    }
    return MBVH_OK;
  }

  const mbvh_build_opts opts_;
  PrimitiveBuffer prims_;
  std::vector<float>& pb_;
  uint64_t base_;
  std::vector<BuildNode> nodes_;
  SahBin bins_[kBinCount];
  float left_bounds_[6] = {0, 0, 0, 0, 0, 0};
};

uint64_t read_index(const void* index, int elt_bytes, uint64_t i) {
  if (elt_bytes == 2) return static_cast<const uint16_t*>(index)[i];
  return static_cast<const uint32_t*>(index)[i];
}

}  // namespace
}  // namespace mbvh

using namespace mbvh;

namespace mbvh {

int check_build_args(const char* fn, const float* positions, uint64_t vertex_count, uint32_t position_stride,
                     uint32_t position_offset, const void* index, int32_t index_elt_bytes, uint64_t index_count,
                     const uint32_t* root_ranges, uint32_t n_roots, const mbvh_build_opts* opts,
                     bool check_index_values) {
  if (!opts) return set_error(MBVH_ERR_BAD_ARGS, "%s: opts is NULL", fn);
  if (!positions && vertex_count > 0)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: positions is NULL", fn);
  if (position_stride < position_offset + 3)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: position stride %u too small for offset %u", fn,
                     position_stride, position_offset);
  if (opts->strategy < MBVH_CENTER || opts->strategy > MBVH_SAH)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: invalid build strategy value %d", fn, opts->strategy);
  if (opts->max_depth < 0 || opts->target_leaf_size < 0 || opts->strict_leaf_size < 0)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: negative max_depth / target_leaf_size / strict_leaf_size", fn);
  if (index) {
    if (index_elt_bytes != 2 && index_elt_bytes != 4)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: index element size must be 2 or 4", fn);
    if (index_count % 3 != 0)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: index count %llu is not a multiple of 3", fn,
                       static_cast<unsigned long long>(index_count));
  } else if (!opts->indirect) {
    return set_error(MBVH_ERR_BAD_ARGS, "%s: a non-indirect build needs an index (ensureIndex)", fn);
  }
  if (n_roots == 0 || !root_ranges)
    return set_error(MBVH_ERR_BAD_ARGS, "%s: at least one root range is required", fn);

  const uint64_t tri_total = index ? index_count / 3 : vertex_count / 3;
  for (uint32_t r = 0; r < n_roots; r++) {
    const uint64_t off = root_ranges[2 * r], cnt = root_ranges[2 * r + 1];
    if (off + cnt > tri_total)
      return set_error(MBVH_ERR_BAD_ARGS, "%s: root range %u [%llu, +%llu) exceeds %llu triangles", fn,
                       r, static_cast<unsigned long long>(off), static_cast<unsigned long long>(cnt),
                       static_cast<unsigned long long>(tri_total));
    if (r > 0 && off < static_cast<uint64_t>(root_ranges[2 * r - 2]) + root_ranges[2 * r - 1])
      return set_error(MBVH_ERR_BAD_ARGS, "%s: root ranges must be sorted and disjoint", fn);
  }
  if (index && check_index_values) {
    for (uint64_t i = 0; i < index_count; i++)
      if (read_index(index, index_elt_bytes, i) >= vertex_count)
        return set_error(MBVH_ERR_BAD_ARGS, "%s: index[%llu] is out of range", fn,
                         static_cast<unsigned long long>(i));
  }
  return MBVH_OK;
}

}  // namespace mbvh

extern "C" int mbvh_build(const float* positions, uint64_t vertex_count, uint32_t position_stride,
                          uint32_t position_offset, void* index, int32_t index_elt_bytes,
                          uint64_t index_count, const uint32_t* root_ranges, uint32_t n_roots,
                          const mbvh_build_opts* opts, mbvh_tree** out_tree) {
  clear_error();
  if (!out_tree) return set_error(MBVH_ERR_BAD_ARGS, "mbvh_build: out_tree is NULL");
  *out_tree = nullptr;
  int rc = check_build_args("mbvh_build", positions, vertex_count, position_stride, position_offset, index,
                            index_elt_bytes, index_count, root_ranges, n_roots, opts, true);
  if (rc) return rc;

  mbvh_tree* tree = new (std::nothrow) mbvh_tree();
  if (!tree) return set_error(MBVH_ERR_OOM, "mbvh_build: out of host memory");

  // Ranges the roots cover. Indirect builds list the selected triangles in one buffer
  // (generateIndirectBuffer, GeometryBVH.js:9-35) and build a single root over it.
  std::vector<std::pair<uint64_t, uint64_t>> ranges;
  PrimitiveBuffer prims;
  if (opts->indirect) {
    const uint64_t last_end =
        static_cast<uint64_t>(root_ranges[2 * (n_roots - 1)]) + root_ranges[2 * (n_roots - 1) + 1];
    uint64_t total = 0;
    for (uint32_t r = 0; r < n_roots; r++) total += root_ranges[2 * r + 1];
    tree->indirect_elt_bytes = last_end > 65536 ? 4 : 2;
    tree->indirect_count = total;
    tree->indirect.assign(total * tree->indirect_elt_bytes, 0);
    uint64_t at = 0;
    for (uint32_t r = 0; r < n_roots; r++) {
      for (uint64_t i = 0; i < root_ranges[2 * r + 1]; i++, at++) {
        const uint64_t v = root_ranges[2 * r] + i;
        if (tree->indirect_elt_bytes == 2) {
          const uint16_t v16 = static_cast<uint16_t>(v);
          std::memcpy(tree->indirect.data() + at * 2, &v16, 2);
        } else {
          const uint32_t v32 = static_cast<uint32_t>(v);
          std::memcpy(tree->indirect.data() + at * 4, &v32, 4);
        }
      }
    }
    ranges.push_back({0, total});
    prims.data = tree->indirect.data();
    prims.elt_bytes = tree->indirect_elt_bytes;
    prims.stride = 1;
  } else {
    for (uint32_t r = 0; r < n_roots; r++) ranges.push_back({root_ranges[2 * r], root_ranges[2 * r + 1]});
    prims.data = static_cast<uint8_t*>(index);
    prims.elt_bytes = index_elt_bytes;
    prims.stride = 3;
  }

  // buildTree.js:128-140: primitive bounds over the full span of the ranges.
  const uint64_t base = ranges.front().first;
  const uint64_t span = ranges.back().first + ranges.back().second - base;
  std::vector<float> pb;
  try {
    pb.assign(span * 6, 0.0f);
  } catch (...) {
    delete tree;
    return set_error(MBVH_ERR_OOM, "mbvh_build: out of host memory for triangle bounds");
  }
  for (uint64_t i = base; i < base + span; i++) {
    uint64_t tri = i;
    if (opts->indirect) {
      tri = tree->indirect_elt_bytes == 2
                ? static_cast<uint64_t>(reinterpret_cast<const uint16_t*>(tree->indirect.data())[i])
                : static_cast<uint64_t>(reinterpret_cast<const uint32_t*>(tree->indirect.data())[i]);
    }
    uint64_t v[3] = {3 * tri, 3 * tri + 1, 3 * tri + 2};
    if (index)
      for (int k = 0; k < 3; k++) v[k] = read_index(index, index_elt_bytes, v[k]);
    for (int k = 0; k < 3; k++) {
      if (v[k] >= vertex_count) {
        delete tree;
        return set_error(MBVH_ERR_BAD_ARGS, "mbvh_build: triangle %llu references a vertex past the end",
                         static_cast<unsigned long long>(tri));
      }
    }
    for (int el = 0; el < 3; el++) {
      const double a = positions[v[0] * position_stride + position_offset + el];
      const double b = positions[v[1] * position_stride + position_offset + el];
      const double c = positions[v[2] * position_stride + position_offset + el];
      double mn = a, mx = a;
      if (b < mn) mn = b;
      if (c < mn) mn = c;
      if (b > mx) mx = b;
      if (c > mx) mx = c;
      const double half = (mx - mn) / 2;
      pb[(i - base) * 6 + 2 * el] = static_cast<float>(mn + half);
      pb[(i - base) * 6 + 2 * el + 1] =
          static_cast<float>(half + (std::fabs(mn) + half) * kFloat32Epsilon);
    }
  }

  TreeBuilder builder(*opts, prims, pb, base);
  tree->roots.resize(ranges.size());
  for (size_t r = 0; r < ranges.size(); r++) {
    int rc = builder.build_root(ranges[r].first, ranges[r].second, &tree->roots[r]);
    if (rc != MBVH_OK) {
      delete tree;
      return rc;
    }
  }
  *out_tree = tree;
  return MBVH_OK;
}

static bool valid_tree(const mbvh_tree* t) { return t && t->magic == kTreeMagic; }

extern "C" int mbvh_tree_info(const mbvh_tree* tree, uint32_t* n_roots, int32_t* indirect_elt_bytes,
                              uint64_t* indirect_count) {
  if (!valid_tree(tree)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_tree_info: invalid tree handle");
  if (n_roots) *n_roots = static_cast<uint32_t>(tree->roots.size());
  if (indirect_elt_bytes) *indirect_elt_bytes = tree->indirect_elt_bytes;
  if (indirect_count) *indirect_count = tree->indirect_count;
  return MBVH_OK;
}

extern "C" int mbvh_tree_root_bytes(const mbvh_tree* tree, uint32_t root, uint64_t* bytes) {
  if (!valid_tree(tree)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_tree_root_bytes: invalid tree handle");
  if (root >= tree->roots.size() || !bytes)
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_tree_root_bytes: root %u out of range", root);
  *bytes = tree->roots[root].size();
  return MBVH_OK;
}

extern "C" int mbvh_tree_copy_root(const mbvh_tree* tree, uint32_t root, void* dst, uint64_t dst_bytes) {
  if (!valid_tree(tree)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_tree_copy_root: invalid tree handle");
  if (root >= tree->roots.size())
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_tree_copy_root: root %u out of range", root);
  if (!dst || dst_bytes < tree->roots[root].size())
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_tree_copy_root: destination too small");
  std::memcpy(dst, tree->roots[root].data(), tree->roots[root].size());
  return MBVH_OK;
}

extern "C" int mbvh_tree_copy_indirect(const mbvh_tree* tree, void* dst, uint64_t dst_bytes) {
  if (!valid_tree(tree)) return set_error(MBVH_ERR_BAD_HANDLE, "mbvh_tree_copy_indirect: invalid tree handle");
  if (tree->indirect.empty()) return MBVH_OK;
  if (!dst || dst_bytes < tree->indirect.size())
    return set_error(MBVH_ERR_BAD_ARGS, "mbvh_tree_copy_indirect: destination too small");
  std::memcpy(dst, tree->indirect.data(), tree->indirect.size());
  return MBVH_OK;
}

extern "C" void mbvh_tree_free(mbvh_tree* tree) {
  if (valid_tree(tree)) {
    tree->magic = 0;
    delete tree;
  }
}
