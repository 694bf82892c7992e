// Kernel parameter blocks of the GI wavefront (host and device).
#pragma once

#include "internal.h"

// Queue counters are striped: each queue (rays, shadow groups, shadow
// entries) is cut into kStripes segments of equal capacity, each with its own
// counter, so that slot reservations spread over kStripes addresses instead of
// serialising on one.  A producer wave appends to the segment of stripe
// (input index / 256) % kStripes; since every input yields at most one
// continuation and one group, a segment of ceil(capacity / (256 * kStripes)) *
// 256 slots never overflows.  Consumers enumerate the live entries densely
// through the inclusive prefix of the kStripes counts.
constexpr uint32_t kStripes = 32;
constexpr uint32_t kStripeUnit = 256;  // producer block size: inputs per stripe step
constexpr uint32_t CTR_Q0 = 0, CTR_Q1 = kStripes, CTR_G = 2 * kStripes, CTR_E = 3 * kStripes;
// CTR_W: work cursors of the refilling closest-hit kernel, one per stripe of
// its input (rays [s * n / kStripes, (s + 1) * n / kStripes)).
// CTR_WS: the same for the shadow kernel's groups.
constexpr uint32_t CTR_W = 4 * kStripes, CTR_WS = 5 * kStripes;
constexpr uint32_t CTR_OVERFLOW = 6 * kStripes, CTR_STACK = 6 * kStripes + 1, CTR_WORDS = 6 * kStripes + 8;
OL_HD constexpr uint32_t ctr_queue(uint32_t q) { return q ? CTR_Q1 : CTR_Q0; }

// Per-path state, structure of arrays indexed by path id
// p = sample * (W * H) + pixel (wavefront/_core.py:19-105 hot path state).
struct GiState {
  float4* T;    // throughput rgb, bounce index
  float4* L;    // radiance rgb, previous BSDF pdf
  uint4* M;     // pixel, frame << 8 | sample, rng state, flags | medium depth << 8
  float4* stk;  // 4 x float4 medium IOR stack per path (null without transmission)
  // Ray queues (ping-pong), structure of arrays.
  float4* qo[2];     // origin, tmin
  float4* qd[2];     // direction, tmax
  uint32_t* qp[2];   // path id
  float4* hit;       // (t, u, v, leaf slot bits) per input-queue entry
  // Shadow work: one group per shaded hit with next-event estimation.
  float4* gT;        // throughput rgb, NEE selection probability
  uint4* gI;         // path, first entry, entry count, area count | env count << 8 | mode << 16
  float4* eo;        // entry origin, tmin
  float4* ed;        // entry direction, tmax
  float4* ec;        // entry contribution rgb, kind (0 point, 1 area, 2 environment)
  uint32_t* ctr;     // CTR_WORDS: per-stripe counts of queue 0, queue 1, groups, entries; overflow flags
  unsigned long long* stats;  // [0] primary, [1] extension, [2] shadow rays, [3] shaded hits
  unsigned long long* work;   // optional: [0] closest nodes, [1] closest tris, [2] shadow nodes, [3] shadow tris
  uint32_t qseg, gseg, eseg;  // slots per stripe segment
};

struct GiParams {
  DScene sc;
  float cam_o[3], cam_f[3], cam_r[3], cam_u[3];
  uint32_t projection;
  uint32_t width, height, spp, frame;  // spp = samples in this chunk
  uint32_t sample_base;                // first sample index of the chunk
  uint32_t max_bounces;
  uint32_t point_count;       // analytic lights scanned for NEE (<= 64)
  uint32_t area_count;        // emissive triangles
  uint32_t secondary_area_samples;
  uint32_t environment_samples;
  uint32_t roulette_start;
  float roulette_min_survival;
  float nee_probability;
  uint32_t unified;
  uint32_t two_pass;          // NEE may emit more than one shadow ray
};

// Host launchers (gi.hip).
struct GiLaunch {
  int cu_count;
  int stack_depth;   // total traversal stack entries needed (3 * BVH depth + 1)
  int lds_stack;     // entries kept in LDS: 8, 12, 16, 24, 32
  int trace_block;   // 64, 128, 256
  int scheduler;     // closest hit: 0 one ray per lane per grid-stride trip, 1 lanes refilled from 64-ray chunks
  uint32_t* spill;   // global spill columns for entries beyond lds_stack
  uint64_t spill_words;
};

// Resident grid of the traversal kernels and the spill words it needs.
uint64_t trace_spill_words(const GiLaunch& l);

void gi_generate(const GiParams& p, const GiState& s, hipStream_t st);
void gi_bounce_prologue(const GiState& s, uint32_t bounce, hipStream_t st);
void gi_trace_closest(const GiLaunch& l, const DScene& sc, const GiState& s, uint32_t in, hipStream_t st);
void gi_shade(const GiLaunch& l, const GiParams& p, const GiState& s, uint32_t in, hipStream_t st);
void gi_trace_shadow(const GiLaunch& l, const DScene& sc, const GiState& s, hipStream_t st);
void gi_accumulate(const GiParams& p, const GiState& s, double4* acc, bool first, hipStream_t st);
void gi_final(const double4* acc, uint32_t np, uint32_t samples, float4* out, hipStream_t st);
void gi_reset(const GiState& s, bool all, hipStream_t st);
void gi_finish(const GiState& s, hipStream_t st);

// Debug / products: closest hit of caller-supplied rays.
void trace_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                uint32_t* overflow, unsigned long long* work, hipStream_t st);
int stack_depth_for(uint32_t bvh_depth);
