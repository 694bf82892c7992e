// Batched GI wavefront: per-job parameters, the device job table, the shared
// path/queue state and the host launchers (gi.hip).
//
// Every render is a job: a scene, a camera, an image size and a range of
// samples.  Resident jobs share one pool of path slots and one set of ray,
// hit and shadow queues; every kernel serves all resident jobs at once, and a
// job may join at any step (continuous batching).  A single frame is the
// one-job case.
#pragma once

#include "internal.h"

// Per-job estimator parameters (the reference's camera block, tile constants
// and shade push constants: targets/vulkan/core.py:4452-4456,
// wavefront_shade_candidate.glsl:482-498).
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
  // ReSTIR DI at primary hits: candidates (0 = off), spatial neighbours, radius (pixels),
  // pairwise spatial merges, whether the previous pass's reservoirs exist,
  // and the represented-sample limit.
  uint32_t restir, restir_spatial, restir_radius, restir_pairwise, restir_history;
  float restir_limit;
};

// ---- batching limits and queue layout ------------------------------------------
// A job's queues are cut into up to kJobStripes segments with one counter
// each, so that slot reservations (one atomic per wave) spread over many
// addresses: MI355X serialises device-scope atomics on one address.  A
// (job, segment) pair is a run; consumers enumerate the runs' live entries in
// 64-entry chunks through a prefix over all runs, so each chunk lies in one
// job and a wave that takes a chunk is job-uniform.
constexpr uint32_t kMaxJobs = 64;
constexpr uint32_t kWorkWords = 8;
constexpr uint32_t kJobStripes = 32;
constexpr uint32_t kRuns = kMaxJobs * kJobStripes;
constexpr uint32_t kChunk = 64;
// Work cursors of the refilling traversal kernels: block b takes chunks from
// stripe b % kWorkStripes, and workgroups are dealt to the eight XCDs round
// robin, so with eight stripes each XCD's blocks share one dynamic pool (C2
// traversal 5.59 ms per frame with 32 stripes, 4.93 with 16, 4.77 with 8).
constexpr uint32_t kWorkStripes = 8;

// Counter block (uint32 words).  Ray queues (C_Q0/1) alternate by step
// parity; so do the shadow groups and entries (C_G0/1, C_E0/1), because a
// step's shadow rays are traced in the next step, beside its closest hits.
constexpr uint32_t C_Q0 = 0, C_Q1 = kRuns, C_G0 = 2 * kRuns, C_G1 = 3 * kRuns, C_E0 = 4 * kRuns, C_E1 = 5 * kRuns;
// Chunk prefixes are kept separately for jobs of flat scenes (C_RP, C_GP)
// and of two-level scenes (C_RP2, C_GP2), so each traversal kernel variant
// sees one kind only.
// A plan: the exclusive prefix of chunk counts over the runs (kRuns + 1
// words, the last the total), then a bucket index that takes a chunk to its
// run in a few steps: the chunks are cut into at most kPlanBuckets buckets
// of 2^shift chunks, and bucket b names the run holding its first chunk.
constexpr uint32_t kPlanBuckets = 1024;
constexpr uint32_t P_SHIFT = kRuns + 1, P_BUCKETS = kRuns + 2, P_LAST = kRuns + 3, P_INDEX = kRuns + 4;
constexpr uint32_t kPlanWords = P_INDEX + kPlanBuckets;
constexpr uint32_t C_RP = 6 * kRuns;             // chunk plan of the input ray queue (flat)
constexpr uint32_t C_GP = C_RP + kPlanWords;     // chunk plan of the shadow groups (flat)
constexpr uint32_t C_RP2 = C_GP + kPlanWords;    // the same for two-level scenes
constexpr uint32_t C_GP2 = C_RP2 + kPlanWords;
// Work cursors of the refilling traversal kernels, one per 256-byte line:
// every wave's chunk claim is an atomic on its stripe's cursor, and cursors
// sharing a line serialise there (C2 traversal 6.12 -> 5.60 ms per frame,
// small single jobs far more).
constexpr uint32_t kCursorStride = 64;
constexpr uint32_t C_W = ((C_GP2 + kPlanWords + 63) / 64) * 64;  // closest-hit cursors (flat)
constexpr uint32_t C_WS = C_W + kWorkStripes * kCursorStride;   // shadow cursors (flat)
constexpr uint32_t C_W2 = C_WS + kWorkStripes * kCursorStride;  // closest-hit cursors (two-level)
constexpr uint32_t C_WS2 = C_W2 + kWorkStripes * kCursorStride; // shadow cursors (two-level)
constexpr uint32_t C_OVERFLOW = C_WS2 + kWorkStripes * kCursorStride, C_STACK = C_OVERFLOW + 1;
constexpr uint32_t C_REPORT = C_OVERFLOW + 8;    // kMaxJobs x 4: rays traced, shadow rays, live paths, shaded hits
constexpr uint32_t C_WORDS = C_REPORT + 4 * kMaxJobs;
OL_HD constexpr uint32_t c_queue(uint32_t q) { return q ? C_Q1 : C_Q0; }
OL_HD constexpr uint32_t c_groups(uint32_t p) { return p ? C_G1 : C_G0; }
OL_HD constexpr uint32_t c_entries(uint32_t p) { return p ? C_E1 : C_E0; }
// Word of run r (job * kJobStripes + stripe) in a run-counter block,
// stripe-major: the stripes of one job, which the waves of its chunks
// reserve queue space from with atomics, lie in different 256-byte lines.
OL_HD constexpr uint32_t run_word(uint32_t r) { return (r % kJobStripes) * kMaxJobs + r / kJobStripes; }

// Device job record (read wave-uniformly by the kernels).
struct JobDev {
  GiParams P;
  uint32_t base;     // first path slot of the job's region; path = base + sample * W * H + pixel
  uint32_t n_paths;  // spp * W * H
  uint32_t stripes;  // queue segments, 1..kJobStripes
  uint32_t seg;      // slots per segment: segment s holds slots base + s * seg ...
  uint32_t ebase;    // first shadow-entry slot
  uint32_t eseg;     // shadow-entry slots per segment
  uint32_t append;   // continuations / shadow groups a segment accepts per step (<= seg)
  uint32_t pad;
  // Primary-hit products of sample 0 (null unless requested): per pixel
  // (t or +inf, geometric normal) and the original triangle index (or
  // OL_NONE) (wavefront_intersect.comp:317-356, targets/vulkan/core.py:4500-4536).
  float4* geo;
  uint32_t* prim;
  // ReSTIR DI (null unless enabled): this pass's and the previous pass's
  // stored reservoirs and primary surfaces (2 x float4 per pixel: shading
  // normal and primary distance, or -1 without a ReSTIR hit; material id bits).
  uint4* res_out;
  const uint4* res_in;
  float4* surf_out;
  const float4* surf_in;
};

// Accumulation of one finished sample chunk into its job's float64 sums.
struct AccItem {
  uint32_t base, np, spp, first_chunk;
  uint32_t final_samples;  // > 0: also write the image divided by this count
  uint32_t pad;
  double4* acc;
  float4* image;
  // Variance product (null unless requested): per pixel float64 sums of the
  // float32 sample luminance and its square; with final_samples, the
  // unbiased variance (render_wavefront_outputs, api.py:1408-1432).
  double2* lum;
  float* variance;
};

// Per-step work lists written by the host before each step.
struct StepList {
  uint32_t n_gen, n_acc;
  // Job slots generated this step (bit k: slot k): their paths are at bounce
  // 0 with unit throughput and no radiance, which the generate kernel leaves
  // unwritten and the shading kernels take without loading.
  unsigned long long fresh;
  uint32_t gen_job[kMaxJobs];          // job slots whose primary rays are generated this step
  uint32_t gen_prefix[kMaxJobs + 1];   // exclusive prefix of their n_paths
  uint32_t acc_prefix[kMaxJobs + 1];   // exclusive prefix of the accumulation items' pixel counts
  AccItem acc[kMaxJobs];
};

// Device pointers of an engine (kernel argument).
struct Wave {
  const JobDev* jobs;
  const StepList* list;
  float4* T;    // throughput rgb, bounce index
  float4* L;    // radiance rgb, previous BSDF pdf
  uint4* M;     // pixel, frame << 8 | sample, rng state, flags | medium depth << 8
  float4* stk;  // 4 x float4 medium IOR stack per path (null when no resident job transmits)
  // Textured scenes (null when no resident job samples textures): per input
  // queue slot, the hit's material after its textures (base colour and
  // roughness, emission and metallic, transmission and resolved occlusion)
  // and its shading normal; per path, the ray cone (width, spread) of the
  // path's next ray.
  float4* tex;
  float2* cone;
  float4* qo[2];
  float4* qd[2];
  uint32_t* qp[2];
  float4* hit;  // (t, u, v, leaf slot bits) per ray-queue slot
  uint32_t* hit_inst;  // instance of the hit (two-level scenes) per ray-queue slot
  float4* gT;   // shadow group: throughput rgb, NEE selection probability
  uint4* gI;    // path, first entry, entry count, area count | env count << 8 | mode << 16
  float4* eo;   // shadow entry: origin, tmin
  float4* ed;   // direction, tmax
  float4* ec;   // contribution rgb, kind (0 point, 1 area, 2 environment)
  // ReSTIR primary samples (null until a ReSTIR job is admitted), 3 x float4
  // per path: origin and valid flag, direction and tmax, contribution and
  // the path's random state after the reservoir pass.
  float4* rsamp;
  uint32_t* ctr;
  // Optional work counters (kWorkWords): closest nodes, closest tris, shadow
  // nodes, shadow tris, closest instance entries, shadow instance entries,
  // shaded hits, area-light records read by NEE.
  unsigned long long* work;
};

// Traversal launch shape.
struct GiLaunch {
  int cu_count;
  int stack_depth;   // total traversal stack entries needed (3 * BVH depth + 1)
  int lds_stack;     // entries kept in LDS: 8, 12, 16, 24, 32, 48
  int trace_block;   // 64, 128, 256
  uint32_t* spill;   // global spill columns for entries beyond lds_stack
  uint64_t spill_words;
};

uint64_t trace_spill_words(const GiLaunch& l);
int stack_depth_for(uint32_t bvh_depth);

// One step of the batched wavefront: accumulate finished chunks, generate
// newly admitted jobs, then trace every resident path once together with
// the shadow rays the previous step's shading emitted, shade, and report
// per-job counts.  `in` is the parity of the input ray queue and of the
// shadow groups this step's shading emits.
// `kinds` says which resident work a step holds (STEP_* bits); the launches
// for absent kinds are left out (their chunk totals stay zero).
// STEP_RESTIR: a ReSTIR job's chunk is at its primary hits (generated this
// step).  STEP_TWO_PASS: some job evaluates more than one next-event sample
// per hit (point lights beside the environment, or per-domain NEE); shading
// without it, ReSTIR or textures runs a variant compiled without those
// paths.
constexpr uint32_t STEP_FLAT = 1u, STEP_INSTANCED = 2u, STEP_RESTIR = 4u, STEP_TEXTURED = 8u, STEP_TWO_PASS = 16u;
constexpr uint32_t kStepKinds = 31u;
// Fills b.shade from the hierarchy's leaf triangles, the triangles' vertex
// indices (material in .w when `material`) and vertex normals.
// face_in / metal_in (original order, optional): per-triangle base colour
// and roughness, and metallic factor, placed in leaf order (b.face, record
// word 11).
void build_shade_records(BvhOutput& b, const uint4* vidx, const float4* normals, uint32_t count, bool material,
                         hipStream_t st, const float4* face_in = nullptr, const float* metal_in = nullptr);
void wave_accumulate(const Wave& w, int cu_count, hipStream_t st);
// rays = false: queue counts and medium stacks only (the hybrid primary
// kernel computes the camera rays itself).
void wave_generate(const Wave& w, uint32_t in, int cu_count, hipStream_t st, bool rays = true);
void wave_trace(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, hipStream_t st);
// `gen`: the step generates (its step list holds the fresh slots).
void wave_shade(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, bool gen, hipStream_t st);
// Hybrid primary step: plans the input queue and traces and shades it in one
// kernel (every ray a fresh camera ray of a flat, untextured scene without
// ReSTIR).
void wave_primary(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, hipStream_t st);
void wave_epilogue(const Wave& w, uint32_t in, hipStream_t st);
void wave_clear(const Wave& w, hipStream_t st);

// Closest hit of caller-supplied rays (primary-hit products, debugging).
// Primary rays of a pixel tile (row-major) for the camera, size and frame in P.
void camera_rays(const GiParams& P, uint32_t x0, uint32_t y0, uint32_t tw, uint32_t th, uint32_t sample, float4* o,
                 float4* d, hipStream_t st);
void trace_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                uint32_t* overflow, unsigned long long* work, hipStream_t st);
