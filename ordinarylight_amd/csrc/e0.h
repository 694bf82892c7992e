// Reference-estimator (E0) mode: parameter blocks and host launchers.
#pragma once

#include "internal.h"

struct E0Params {
  DScene sc;                  // triangles via tri_vidx + positions; BVH when not brute force
  const float4* positions;    // world-space vertices
  float pos[3], fwd[3], right[3], up[3];  // float32 camera basis (unit, unscaled)
  double k_x, k_y;            // perspective: aspect*tan(fov/2), tan(fov/2); ortho: aspect*h/2, h/2; pano: radians
  uint32_t projection;        // 0 perspective, 1 orthographic, 2 panoramic
  uint32_t width, height;
  uint32_t brute_force;       // index-order scan (exact tie semantics)
  uint32_t env_mode;          // 0 gradient sky, 1 constant environment, 2 image environment
  float env_color[3];
  float env_intensity;        // float32(intensity)
  double env_rotation;
  const float* env_image;     // [env_h][env_w][3]
  uint32_t env_w, env_h;
};

// Device buffers of one frame; o/d/T/rad element types depend on projection.
struct E0State {
  void *o, *d, *T, *rad;
  float* acc;          // [N][3] float32 accumulation over samples
  uint32_t* active;    // active ray ids (ascending)
  uint32_t* flags;     // per active entry: hit?
  uint32_t* scan;      // exclusive scan of flags (+ total)
  float* ht;           // per active entry: hit distance (float32)
  int* htri;           // per active entry: triangle or -1
  uint32_t* hitlist;   // compacted hit ray ids (ascending)
  float* hit_t;
  int* hit_tri;
  uint32_t* ctr;       // [1] stack overflow flag
};

void e0_camera(const E0Params& p, const E0State& s, const float* jx, const float* jy, hipStream_t st);
void e0_trace(int cus, int stack, const E0Params& p, const E0State& s, uint32_t n_active, hipStream_t st);
uint32_t e0_compact(const E0State& s, uint32_t n_active, hipStream_t st);
void e0_shade(const E0Params& p, const E0State& s, uint32_t n_hit, const float* gauss, hipStream_t st);
void e0_accumulate(const E0Params& p, const E0State& s, hipStream_t st);
void e0_finish(const E0Params& p, const E0State& s, uint32_t samples, float4* out, hipStream_t st);
