// Kernel argument blocks and host-side launchers (implemented in mbvh_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "mbvh_device.h"

namespace mbvh {

struct MeshView {
  const WideNode* nodes;
  const uint32_t* roots;
  const uint2* bigleaf;
  const float4* tris;  // 3 per slot
  uint32_t n_roots;
  uint32_t max_stack;  // stack entries a traversal of this mesh can need
  float coord_max;     // largest |vertex coordinate| (scales FP32 error bounds)
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
  uint64_t n;
  uint32_t side_mask;  // 1 front, 2 back, 3 both; 0 = per-triangle masks stored in the mesh
  float near_dist;
  float far_dist;
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
  uint64_t n;
  uint32_t side_mask;
  float near_dist;
  float far_dist;
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
  uint64_t n;
  float min_threshold;
  float max_threshold;
  uint32_t* out_face;
  float* out_dist;
  float* out_point;
  float* out_bary;
};

// Threads a launch uses for n queries: one query per thread, capped at `max_threads` (the
// launch then strides over the remaining queries).
uint64_t launch_threads(uint64_t n, uint64_t max_threads);

hipError_t launch_raycast_first(const RaycastFirstArgs& a, uint64_t threads, hipStream_t s);
hipError_t launch_closest_point(const ClosestPointArgs& a, uint64_t threads, hipStream_t s);
// fill == false: count hits per ray into a.counts; fill == true: write them at a.offsets.
hipError_t launch_raycast_all(const RaycastAllArgs& a, bool fill, uint64_t threads, hipStream_t s);
hipError_t launch_apply_side_mask(float4* tris, uint64_t n_slots, const uint8_t* mask,
                                  hipStream_t s);

}  // namespace mbvh
