// Host-side declarations shared by the olhip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "common.h"

// Error carrying the C-ABI status class (OLHIP_ERR_*).
struct OlError {
  int code;
  std::string message;
};

#define OL_HIP_CHECK(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess)                                                         \
      throw OlError{-4, std::string(#expr) + ": " + hipGetErrorString(_e)};       \
  } while (0)

// Device buffer owned by RAII.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  void swap(DevBuf& o) noexcept {
    std::swap(ptr, o.ptr);
    std::swap(bytes, o.bytes);
  }
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  // Grows (never shrinks) to at least `n` bytes; contents are not preserved.
  void reserve(size_t n) {
    if (n <= bytes && ptr) return;
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
      ptr = nullptr;
      throw OlError{-3, "hipMalloc of " + std::to_string(n) + " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = n;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---- BVH ------------------------------------------------------------------------
// Input: world-space vertices and indexed triangles (original order).
struct BvhInput {
  const float4* positions;  // device, [V] (w ignored)
  const uint4* tri_vidx;    // device, [T] (i0, i1, i2, material)
  uint32_t tri_count;
  uint32_t leaf_size;       // 1..16
  // Optional host copy of the primitives' boxes (min xyz, max xyz per
  // primitive): the binary hierarchy is then built on the host by binned SAH
  // instead of Morton codes (small top levels), the boxes phase as usual.
  const float* host_boxes = nullptr;
  bool isolate_large = false;  // sah_topology's `isolate` (top levels over instances)
};

// Binary LBVH topology of a build (Morton order and Karras hierarchy), kept
// so that a refit can recompute boxes for moved vertices without re-sorting.
struct BvhTopology {
  uint32_t tri_count = 0, leaf_size = 0;
  DevBuf sorted;                  // [T] original triangle of each leaf position
  DevBuf left, right;             // [T-1] children (leaf refs | OL_LEAF_BIT)
  DevBuf rfirst, rlast;           // [T-1] leaf range of each internal node
  DevBuf parent_int, parent_leaf; // parents of internal nodes / leaves
};

// Output of a build.  All arrays are device-resident and owned by the caller.
struct BvhOutput {
  DevBuf nodes;      // WideNode[node_count]
  DevBuf tri;        // float4[3 * T]: (v0, orig id bits), e1, e2 in leaf order
  // float4[4 * T] shading records in leaf order (triangle hierarchies of
  // shaded scenes): (n0, material bits), (n1, triangle id bits), (n2, 0),
  // (e1 x e2, 0), so a hit's surface is one 64-byte read.
  DevBuf shade;
  // float4[T] per-triangle base colour and roughness in leaf order (scenes
  // with per-triangle material parameters; the metallic factor is word 11 of
  // the shading record).
  DevBuf face;
  uint32_t node_count = 0;
  uint32_t depth = 0;      // wide levels (root = level 1)
  double build_ms = 0.0;
  std::unique_ptr<BvhTopology> topology;  // kept when the build asked for it (refit)
};

// Optional debug capture of intermediate build arrays (bit-exact checks).
struct BvhDebug {
  std::vector<uint32_t> morton;       // [T] per original triangle
  std::vector<uint32_t> sorted_index; // [T] original index of sorted position
  std::vector<uint32_t> left, right;  // [T-1] binary children (leaf refs | 0x80000000)
  std::vector<float> node_box;        // [T-1][8] binary internal boxes (min.xyzw, max.xyzw)
};

// Host binned-SAH binary hierarchy over n >= 2 primitive boxes (min xyz, max
// xyz each), one primitive per leaf, in the Karras arrays' conventions: leaf
// order, children (leaf refs | OL_LEAF_BIT), leaf ranges, parents.  isolate:
// also price splitting off the largest box alone (top levels over instances).
void sah_topology(const float* box, uint32_t n, std::vector<uint32_t>& order, std::vector<uint32_t>& left,
                  std::vector<uint32_t>& right, std::vector<uint32_t>& rfirst, std::vector<uint32_t>& rlast,
                  std::vector<uint32_t>& parent_int, std::vector<uint32_t>& parent_leaf, bool isolate = false);

// keep_topology: retain the binary topology in out.topology for bvh_refit.
void bvh_build(const BvhInput& in, BvhOutput& out, hipStream_t stream, BvhDebug* debug, bool keep_topology = false);
// Recomputes the boxes, the wide nodes and the leaf triangles of a kept
// topology for new vertex positions (same triangles, same indices), as an
// acceleration-structure update does: the leaf order stays, boxes follow the
// geometry.  in.tri_count and in.leaf_size must match the build.
void bvh_refit(const BvhInput& in, BvhOutput& out, hipStream_t stream);

// ---- scans ---------------------------------------------------------------------
// Exclusive prefix sum of `n` uint32 values in place; writes the total to
// `total` (device pointer, may be null).  Scratch is allocated internally.
void scan_exclusive_u32(uint32_t* data, uint32_t n, uint32_t* total, hipStream_t stream);

// Streaming device-to-device copy of n16 float4 records (bandwidth probe and
// hardware-counter calibration).
void stream_copy(const float4* src, float4* dst, uint64_t n16, int cu_count, hipStream_t s);

// ---- device-visible scene view ---------------------------------------------------
// One placement of a bottom-level BVH in a two-level scene.
struct InstanceDev {
  float4 w2o[3];        // world -> object affine rows: o = (dot(row.xyz, p) + row.w)
  float4 nrm[3];        // object -> world normal matrix rows (inverse transpose of the linear part)
  uint32_t blas;        // bottom-level structure
  uint32_t material;    // material record index
  uint32_t tri_offset;  // first triangle of this instance in packed (reference) triangle order
  uint32_t root;        // child word of `blas` where traversal starts (0: its root node)
  float4 sphere;        // world bounding sphere: centre xyz, radius (+inf: none)
};

// Bottom-level BVH of one mesh resource, object space.
struct BlasDev {
  const WideNode* nodes;
  const float4* tri;      // leaf-order (v0, local id bits), e1, e2
  const uint4* vidx;      // local triangle -> (i0, i1, i2, unused), indices into `normals`
  const float4* normals;  // object-space vertex normals of this resource
  const float4* shade;    // leaf-order shading records (BvhOutput::shade; material bits 0)
  uint32_t node_count, tri_count, pad0, pad1;
};

struct DScene {
  const WideNode* nodes;
  const float4* tri;        // leaf-order traversal triangles
  const uint4* tri_vidx;    // original order (i0, i1, i2, material)
  const float4* normals;    // [V]
  const float4* shade;      // leaf-order shading records (BvhOutput::shade)
  const float4* materials;  // [M * 12]
  const float4* analytic;   // [P * 4]
  const float4* area;       // [L * 5]
  uint32_t tri_count, node_count, material_count, analytic_count, area_count;
  float area_weight;
  uint32_t env_valid;       // environment light present (analytic type 3)
  float env_r, env_g, env_b;  // color * intensity
  // Two-level scenes: `nodes` is the top-level BVH over instances (leaf
  // records hold the instance index in v0.w), `tri_vidx` and `normals` are
  // unused.
  uint32_t instanced, instance_count;
  const InstanceDev* inst;
  const BlasDev* blas;
  // Emitter selection guide (null without area lights): area_guide_size
  // buckets (a power of two, so bucket j = floor(selection * size) is exact),
  // entry j the first emitter whose cumulative weight reaches j / size
  // (clamped to the last), so a selection in bucket j lies between entries j
  // and j + 1.
  const uint32_t* area_guide;
  uint32_t area_guide_size;
  // Per-triangle material parameters (null without): base colour and
  // roughness per leaf slot (the metallic factor is the shading record's
  // word 11), and per original triangle (the reference estimator's colours).
  const float4* face;
  const float4* face_orig;
  // Textures (null / -1 without): the texture word array, binding records
  // (2 float4 each), per-vertex uv0.xy uv1.xy and world tangents, per-triangle
  // uv0 density and world positions (flat scenes only).
  const uint32_t* tex_words;
  const float4* bindings;
  const float4* uvs;
  const float4* tangents;
  const float* uv_density;
  const float4* positions;
  uint32_t textured;       // some material has texture bindings
  int32_t env_texture;     // image environment's texture, or -1
  float env_rotation, env_log_range;
  float env_color_r, env_color_g, env_color_b, env_intensity;
};
// Guide buckets: the power of two at or above twice the emitter count, at
// least kAreaGuideMin and at most kAreaGuideMax, so that a bucket holds
// about half an emitter and the bisection inside it rarely takes a step.
constexpr uint32_t kAreaGuideMin = 4096, kAreaGuideMax = 1u << 22;
