/*
 * olhip -- C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * The reference exposes no native interface: its renderer boundary is the
 * Python RendererProtocol (ordinarylight/renderers/base.py:41-71), driven by
 * Renderer._render_now (ordinarylight/renderer/_core.py:757-766).  This ABI
 * is what the drop-in renderers.hip implementation binds through ctypes;
 * each entry point names the reference step it replaces.
 *
 * Conventions: every function returns OLHIP_OK (0) or a negative status
 * whose class mirrors the reference's exceptions (ValueError, RuntimeError,
 * MemoryError); olhip_last_error() describes the last failure of the calling
 * thread.  Pointers are plain host pointers; nothing here takes ownership of
 * caller memory.  One context owns one HIP stream on one device; calls on
 * one context must not overlap.
 */
#ifndef OLHIP_H
#define OLHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OLHIP_OK 0
#define OLHIP_ERR_VALUE (-1)   /* invalid argument: ValueError/TypeError */
#define OLHIP_ERR_RUNTIME (-2) /* unsupported or invalid state: RuntimeError */
#define OLHIP_ERR_MEMORY (-3)  /* device allocation failed: MemoryError */
#define OLHIP_ERR_DEVICE (-4)  /* HIP runtime error */

typedef struct olhip_ctx olhip_ctx;
typedef struct olhip_scene olhip_scene;

/* Description of the last error on this thread ("" if none). */
const char* olhip_last_error(void);
/* ABI version (major * 100 + minor). */
int olhip_version(void);
/* Number of visible HIP devices. */
int olhip_device_count(int* count);

/* Renderer device state; replaces the Vulkan device/queue setup of
 * _VulkanGlobalIlluminationEngine (targets/vulkan/api.py:844-917). */
int olhip_create(int device, olhip_ctx** out);
int olhip_destroy(olhip_ctx* ctx);

/* Scene in world space, indexed triangles in the reference's packed
 * triangle order (Scene.render_triangles, scene/_core.py:2755-2762).
 * Materials use the 12 x vec4 record of Scene.triangle_material_data
 * (scene/_core.py:2816-2924), analytic lights the 4 x vec4 record of
 * analytic_light_data (:2578-2631), area lights the 5 x vec4 record of
 * emissive_triangle_data (:2637-2690). */
typedef struct olhip_scene_desc {
  uint32_t vertex_count;
  const float* positions;      /* [vertex_count][3] */
  const float* normals;        /* [vertex_count][3] unit shading normals */
  uint32_t triangle_count;
  const uint32_t* indices;     /* [triangle_count][3] */
  const uint32_t* material_ids;/* [triangle_count] index into materials */
  uint32_t material_count;
  const float* materials;      /* [material_count][12][4] */
  uint32_t analytic_light_count;
  const float* analytic_lights;/* [analytic_light_count][4][4] */
  uint32_t area_light_count;
  const float* area_lights;    /* [area_light_count][5][4] */
  float area_light_weight;     /* Scene.emissive_light_weight */
  uint32_t leaf_size;          /* BVH leaf size 1..16, 0 = default (4) */
  uint32_t builder;            /* OLHIP_BUILDER_LBVH (GPU Morton LBVH, default) or
                                  OLHIP_BUILDER_SAH (host binned-SAH hierarchy,
                                  GPU boxes and wide collapse) */
  /* Textures (all NULL / 0 without): the packed word array of
   * Scene.texture_data (scene/_core.py:3005-3043: count, 8 metadata words per
   * texture, sRGB and linear RGBA8 mip chains), binding records of
   * texture_binding_data (:2984-3003), and for textured materials the
   * per-vertex inputs of triangle_attribute_data (:2042-2069), indexed.
   * Materials with bindings carry 1.0 in record word [8].w (a padding word
   * of the reference record) and their transmission binding in [5].w.  An
   * image environment's record holds (texture, rotation, log range) in row 3
   * (analytic_light_data, :2613-2629). */
  uint32_t texture_word_count;
  const uint32_t* texture_words;
  uint32_t binding_count;
  const float* bindings;       /* [binding_count][2][4] */
  const float* uvs;            /* [vertex_count][4] uv0.xy, uv1.xy (textured materials) */
  const float* tangents;       /* [vertex_count][4] world tangent, handedness */
  const float* uv_density;     /* [triangle_count] sqrt(uv0 area / world area) */
  /* Per-triangle material parameters (NULL without): base colour rgb,
   * roughness and metallic factor of each triangle, replacing those words of
   * its material record -- the per-triangle records of
   * Scene.triangle_material_data (scene/_core.py:2816-2924) stored as the
   * five words that differ between the triangles of a mesh. */
  const float* face_params;    /* [triangle_count][5] */
} olhip_scene_desc;

#define OLHIP_BUILDER_LBVH 0u
#define OLHIP_BUILDER_SAH 1u

/* Two-level scene: object-space mesh resources, each with its own
 * bottom-level BVH, placed by instances under affine transforms; a top-level
 * BVH over the instances' world bounds.  Replaces the BLAS/TLAS build of
 * VulkanSceneResources (targets/vulkan/scene.py:1132-1205, 1322-1561).
 * Triangle order (ids, tie-breaks, products) is the flattened packed order:
 * instance k's triangles start at instance_triangle_offset[k]. */
typedef struct olhip_instanced_desc {
  uint32_t resource_count;
  const uint32_t* vertex_offsets;      /* [resource_count + 1] into positions / normals */
  const uint32_t* triangle_offsets;    /* [resource_count + 1] into indices */
  const float* positions;              /* object space [vertex_offsets[R]][3] */
  const float* normals;                /* object space unit normals, same layout */
  const uint32_t* indices;             /* [triangle_offsets[R]][3], local to the resource */
  uint32_t instance_count;
  const uint32_t* instance_resource;   /* [I] */
  const uint32_t* instance_material;   /* [I] index into materials */
  const uint32_t* instance_triangle_offset; /* [I] */
  const uint32_t* instance_id;         /* [I] scene instance ids */
  const float* world_to_object;        /* [I][3][4] affine rows */
  const float* normal_matrix;          /* [I][3][3] inverse transpose of the linear part */
  const float* determinant;            /* [I] determinant of the linear part */
  const float* world_bounds;           /* [I][6] min xyz, max xyz (rounded outward) */
  uint32_t material_count;
  const float* materials;              /* [material_count][12][4] */
  uint32_t analytic_light_count;
  const float* analytic_lights;
  uint32_t area_light_count;
  const float* area_lights;            /* world-space emissive triangles [L][5][4] */
  float area_light_weight;
  uint32_t leaf_size;                  /* bottom-level leaf size 1..16, 0 = 4 */
  uint32_t builder;                    /* bottom-level hierarchy: OLHIP_BUILDER_LBVH or
                                          OLHIP_BUILDER_SAH; the top level is always
                                          host binned SAH */
  const float* world_spheres;          /* [I][4] centre xyz, radius of a sphere containing
                                          the instance, or NULL; rays whose line misses
                                          it skip the instance */
  uint32_t texture_word_count;         /* an image environment's texture words (as in
                                          olhip_scene_desc), or 0; materials with texture
                                          bindings need a flat scene */
  const uint32_t* texture_words;
} olhip_instanced_desc;

int olhip_scene_create_instanced(olhip_ctx* ctx, const olhip_instanced_desc* desc, olhip_scene** out);

/* In-place update of a resident scene with the same triangles (vertex
 * indices; for two-level scenes the same resources and each instance's
 * resource): material ids may change; new vertex positions and normals (optional: NULL positions
 * keeps them) refit the BVH over the kept leaf order, as the reference's
 * acceleration-structure update (targets/vulkan/scene.py:1217-1275); new
 * instance transforms rebuild the top level; materials and lights are
 * replaced.  Replaces _try_update_window_scene (targets/vulkan/core.py
 * :4696-4870).  A different structure is rejected (OLHIP_ERR_VALUE).  Jobs
 * that use the scene must not be queued or running during an update. */
int olhip_scene_update(olhip_ctx* ctx, olhip_scene* scene, const olhip_scene_desc* desc);

/* Host-only: the binned-SAH binary hierarchy a two-level scene's top level is
 * built from (no counterpart in the reference, whose top level is built by
 * the driver): centroid-binned splits, and the largest box split off alone
 * where that prices lower.  boxes [n][6] (min xyz, max xyz), n >= 2; order [n] is the
 * primitive of each leaf position; internal node i (0 = root, n - 1 nodes)
 * has children left/right (internal index, or leaf position | 0x80000000)
 * and leaf range [first, last]. */
int olhip_sah_topology(const float* boxes, uint32_t n, uint32_t* order, uint32_t* left, uint32_t* right,
                       uint32_t* first, uint32_t* last);
int olhip_scene_update_instanced(olhip_ctx* ctx, olhip_scene* scene, const olhip_instanced_desc* desc);

typedef struct olhip_scene_info {
  uint32_t triangle_count;
  uint32_t node_count;
  uint32_t bvh_depth;
  uint32_t stack_depth;
  double build_ms;
  uint64_t device_bytes;
} olhip_scene_info;

/* Upload + GPU LBVH build; replaces VulkanSceneResources._build_scene
 * (targets/vulkan/scene.py:1322-1561). */
int olhip_scene_create(olhip_ctx* ctx, const olhip_scene_desc* desc, olhip_scene** out);
int olhip_scene_destroy(olhip_scene* scene);
int olhip_scene_info_get(const olhip_scene* scene, olhip_scene_info* out);

/* Debug: copy the built BVH.  nodes: node_count * 64 bytes; tri:
 * triangle_count * 48 bytes (leaf order).  Either pointer may be NULL. */
int olhip_scene_bvh(const olhip_scene* scene, void* nodes, void* tri);

/* Camera basis as _camera_vectors (targets/vulkan/core.py:103-121):
 * unit forward, right/up pre-scaled by the projection half extent. */
typedef struct olhip_camera {
  float position[3];
  float forward[3];
  float right[3];
  float up[3];
  uint32_t projection;  /* 0 perspective, 1 orthographic, 2 panoramic */
} olhip_camera;

/* GI estimator settings: the WavefrontShadeSettings push constants
 * (wavefront/shading.py:66-114) of the split path. */
typedef struct olhip_gi_settings {
  uint32_t max_bounces;               /* 1..64 */
  uint32_t secondary_area_samples;    /* >= 1 */
  uint32_t environment_samples;       /* 0..4 */
  uint32_t russian_roulette_start;    /* 0 = off */
  float russian_roulette_min_survival;
  float secondary_nee_probability;    /* (0, 1] */
  uint32_t unified_secondary_nee;     /* 0 or 1 */
  uint32_t path_capacity;             /* resident paths per chunk, 0 = default */
  /* ReSTIR DI for area lights at primary hits (wavefront_restir.glsl,
   * wavefront_primary_impl.glsl:1539-1700; settings api.py:60-75):
   * candidates per pixel (0 = off, <= 64), spatial neighbours (0 = no
   * spatial reuse, <= 8) and their pixel radius (<= 32), the represented-
   * sample limit of reused reservoirs (1..64), and pairwise (1) or canonical
   * (0) spatial merges.  The samples of such a job run as successive
   * one-sample passes, each reusing the previous pass's reservoirs as the
   * reference reuses the previous frame's. */
  uint32_t restir_candidates;
  uint32_t restir_spatial_neighbors;
  uint32_t restir_spatial_radius;
  uint32_t restir_history_limit;
  uint32_t restir_pairwise_mis;
} olhip_gi_settings;

typedef struct olhip_stats {
  uint64_t primary_rays;
  uint64_t extension_rays;
  uint64_t shadow_rays;
  uint64_t queue_overflow;
  double total_ms;         /* device time of the render (events) */
  double trace_ms;         /* traversal launches (when profiling): each traces a step's
                              closest hits and the previous step's shadow rays */
  double shade_ms;
  double shadow_ms;        /* 0: shadow rays are traced within the traversal launches */
  uint32_t trace_launches;
  uint32_t shade_launches;
  uint32_t shadow_launches; /* 0, as shadow_ms */
  uint32_t chunks;
  /* Work counters (only with olhip_set_work_counters(ctx, 1)). */
  uint64_t closest_nodes;  /* 64-byte nodes visited by closest-hit rays */
  uint64_t closest_tris;   /* 48-byte triangles tested by closest-hit rays */
  uint64_t shadow_nodes;
  uint64_t shadow_tris;
  uint64_t closest_instances; /* two-level scenes: instances entered by closest-hit rays */
  uint64_t shadow_instances;  /* and by shadow rays */
  uint64_t shaded_hits;       /* hits whose shading record and material were read */
  uint64_t area_samples;      /* area-light records read by next-event estimation */
} olhip_stats;

/* One GI frame; replaces _VulkanGlobalIlluminationEngine.render_frame /
 * render_wavefront_outputs (targets/vulkan/api.py:1233-1451).  out_rgba is
 * caller-owned float32 [height][width][4] (linear HDR, alpha 1), or NULL to
 * keep the image on the device (olhip_frame_download).  frame_index is the
 * renderer's sequence index (stored as float32 like the reference camera
 * block).  stats may be NULL. */
int olhip_render_gi(olhip_ctx* ctx, const olhip_scene* scene, const olhip_camera* camera, uint32_t width,
                    uint32_t height, uint32_t samples, uint32_t frame_index, const olhip_gi_settings* settings,
                    float* out_rgba, olhip_stats* stats);
/* ---- batched render engine (continuous batching) ----
 * Many jobs of different scenes, sizes and sample counts share one pool of
 * path slots and one set of ray/shadow queues; each step advances every
 * resident path by one bounce.  Jobs are admitted first-come first-served
 * when a slot range for their next sample chunk is free (admission by
 * memory), so a job may join while others are in flight.  A job's image is
 * bit-identical to olhip_render_gi of the same job: each path depends only
 * on (pixel, sample, frame) and the float64 pixel sums run in sample order.
 * No counterpart in the reference (its executor renders one frame at a time,
 * targets/vulkan/core.py:2338-3215). */
typedef struct olhip_engine olhip_engine;

typedef struct olhip_engine_config {
  uint64_t path_slots;    /* resident path pool; 0 = 2^25 */
  uint32_t entry_factor;  /* shadow-entry slots per path slot; 0 = 2 */
  int32_t lds_stack;      /* traversal stack entries in LDS; 0 = context setting */
  int32_t trace_block;    /* traversal block size; 0 = context setting */
  int32_t use_graph;      /* replay each step as a captured hipGraph (1) or launch directly (0) */
  /* Failure-path testing: most continuations and shadow groups one queue
   * segment accepts per step (0 = the segment's size, which cannot
   * overflow).  A smaller value makes shading drop paths; the jobs resident
   * in that step then fail with "ray queue overflow". */
  uint32_t append_limit;
  /* Execution strategy of a step whose every ray is a fresh camera ray of a
   * flat, untextured scene without ReSTIR: 0 traces and shades in separate
   * kernels (wavefront), 1 runs both in one kernel (hybrid: the
   * coherent primary bounce inline, later bounces through the queues).  Both
   * give the same images bit for bit. */
  uint32_t fused_primary;
} olhip_engine_config;

/* Named products beside the color image (render_wavefront_outputs,
 * targets/vulkan/api.py:1252-1451): sample 0's primary hit (depth = ray
 * distance or +inf, unflipped geometric normal, original triangle index or
 * 0xFFFFFFFF) and the unbiased per-pixel luminance variance over samples. */
#define OLHIP_PRODUCT_PRIMARY 1u
#define OLHIP_PRODUCT_VARIANCE 2u

typedef struct olhip_job_desc {
  const olhip_scene* scene;
  olhip_camera camera;
  uint32_t width, height, samples, frame_index;
  olhip_gi_settings settings;  /* path_capacity: most paths per sample chunk, 0 = no limit */
  uint32_t products;           /* OLHIP_PRODUCT_* bits */
  uint32_t pad;
} olhip_job_desc;

enum {
  OLHIP_JOB_UNKNOWN = -1,
  OLHIP_JOB_QUEUED = 0,
  OLHIP_JOB_RUNNING = 1,
  OLHIP_JOB_DONE = 2,
  OLHIP_JOB_FAILED = 3
};

typedef struct olhip_job_stats {
  uint64_t primary_rays, extension_rays, shadow_rays;
  uint32_t steps, chunks;
  double submit_ms, admit_ms, done_ms;  /* engine clock: ms since the engine was created */
} olhip_job_stats;

typedef struct olhip_engine_stats {
  uint64_t steps, jobs_done, rays, queue_overflow;
  double busy_ms;  /* host time spent inside steps */
  uint32_t resident_jobs, queued_jobs, max_resident_jobs;
  uint32_t fused_steps;  /* steps run as one traversal-and-shading kernel (hybrid strategy) */
  uint64_t path_slots, entry_slots;
  double device_ms;  /* stream time of the retired steps, first launch to report copy */
} olhip_engine_stats;

int olhip_engine_create(olhip_ctx* ctx, const olhip_engine_config* config, olhip_engine** out);
int olhip_engine_destroy(olhip_engine* engine);
/* Queue a job (the scene must outlive it).  Thread-safe. */
int olhip_engine_submit(olhip_engine* engine, const olhip_job_desc* job, uint64_t* job_id);
/* Run one step; *pending = jobs not yet done.  Not with a running worker. */
int olhip_engine_step(olhip_engine* engine, uint32_t* pending);
/* Step until every submitted job is done.  Not with a running worker. */
int olhip_engine_run(olhip_engine* engine);
/* Job state (OLHIP_JOB_*).  Thread-safe. */
int olhip_engine_poll(olhip_engine* engine, uint64_t job_id, int* state);
/* Wait for a job (timeout_s < 0: no limit); *state = final or current state. */
int olhip_engine_wait(olhip_engine* engine, uint64_t job_id, double timeout_s, int* state);
/* Copy a done job's float32 [H][W][4] image (out_rgba may be NULL) and
 * statistics, then release it. */
int olhip_engine_result(olhip_engine* engine, uint64_t job_id, float* out_rgba, olhip_job_stats* stats);
/* Copy a done job's products before olhip_engine_result: depth [H][W],
 * normal [H][W][3], primitive [H][W] (needs OLHIP_PRODUCT_PRIMARY) and
 * variance [H][W] (needs OLHIP_PRODUCT_VARIANCE); any pointer may be NULL. */
int olhip_engine_products(olhip_engine* engine, uint64_t job_id, float* depth, float* normal, uint32_t* primitive,
                          float* variance);
/* Background worker that steps whenever work is queued or resident. */
int olhip_engine_start(olhip_engine* engine);
int olhip_engine_stop(olhip_engine* engine);
int olhip_engine_get_stats(olhip_engine* engine, olhip_engine_stats* out);

/* olhip_render_gi with named products: out_rgba (may be NULL), depth,
 * normal, primitive, variance as for olhip_engine_products (NULL = not
 * requested). */
int olhip_render_products(olhip_ctx* ctx, const olhip_scene* scene, const olhip_camera* camera, uint32_t width,
                          uint32_t height, uint32_t samples, uint32_t frame_index, const olhip_gi_settings* settings,
                          float* out_rgba, float* depth, float* normal, uint32_t* primitive, float* variance,
                          olhip_stats* stats);

/* Streaming copy of `bytes` (read + write) repeated `repeats` times on the
 * context's stream; reports milliseconds per copy.  Used to measure
 * achievable HBM bandwidth and to calibrate hardware byte counters. */
int olhip_bandwidth_probe(olhip_ctx* ctx, uint64_t bytes, uint32_t repeats, double* ms_per_copy);
/* Copy the last frame of this context to host memory. */
int olhip_frame_download(olhip_ctx* ctx, float* out_rgba);
/* Enable per-kernel event timing inside olhip_render_gi (0 or 1). */
int olhip_set_profiling(olhip_ctx* ctx, int enabled);
/* Count traversal work (nodes visited, triangles tested) in olhip_stats. */
int olhip_set_work_counters(olhip_ctx* ctx, int enabled);
/* Traversal launch shape: stack entries kept in LDS per lane (the rest of
 * the 3 * depth + 1 entries spill to a global per-lane column) and the block
 * size.  Supported: (8|12|24|32|48, 128) and (16, 64|128|256).  Results do
 * not depend on these settings. */
int olhip_set_traversal(olhip_ctx* ctx, int lds_entries, int block);
/* Execution strategy of the single-frame entry points' primary step
 * (olhip_engine_config.fused_primary; 0 = wavefront, 1 = hybrid).  Replaces
 * the reference's wavefront_execution_strategy choice
 * (targets/vulkan/api.py:158, :611-616; core.py:886) for the two strategies
 * built here. */
int olhip_set_fused_primary(olhip_ctx* ctx, int enabled);

/* Output formats of olhip_frame_output. */
#define OLHIP_OUTPUT_RGBA8 0u
#define OLHIP_OUTPUT_NV12 1u
#define OLHIP_OUTPUT_P010 2u

/* Display output of the last frame rendered with olhip_render_gi (kept on the
 * device whether or not it was copied out); replaces the presentation passes
 * wavefront_tone_map.comp, rgba_to_nv12.comp and hdr_to_p010.comp behind
 * render_gpu_frame (targets/vulkan/api.py:1074-1115).  Colour is
 * sRGB(ACES(max(hdr, 0) * exposure)).
 *   RGBA8: uint8 [height][width][4], alpha 255; size must equal the frame's.
 *   NV12:  luma plane (pitch = width bytes) then interleaved CbCr plane of
 *          the 2x2 averages of the RGBA8 pixels, BT.709 studio range.
 *   P010:  the same with 10-bit values in the high bits of 16-bit words
 *          (pitch = 2 * width bytes), from the half-float HDR image scaled
 *          to width x height by nearest neighbour.
 * NV12 and P010 need width % 4 == 0 and even height.  out is host memory, or
 * device memory of ctx's device when out_on_device is 1 (a GPU-resident
 * frame, no host copy). */
int olhip_frame_output(olhip_ctx* ctx, uint32_t format, float exposure, uint32_t width, uint32_t height, void* out,
                       uint32_t out_on_device);

/* Primary rays of a pixel tile, exactly as a render generates them for
 * (pixel, frame_index, sample) (wavefront_generate.comp:86-167): origins /
 * directions [tile_height][tile_width][4] (xyz, tmin / tmax), row-major from
 * (x0, y0).  With olhip_trace this is the GPU pick of
 * targets/vulkan/api.py:936-987 (a 1x1 tile at frame 0, sample 0). */
int olhip_camera_rays(olhip_ctx* ctx, const olhip_camera* cam, uint32_t width, uint32_t height, uint32_t frame_index,
                      uint32_t sample, uint32_t x0, uint32_t y0, uint32_t tile_width, uint32_t tile_height,
                      float* origins, float* directions);

/* Closest hit of caller rays: origins/directions [n][4] (xyz, tmin / tmax).
 * hits [n][4]: t, u, v, original triangle index as uint32 bits (0xFFFFFFFF
 * on a miss).  Primary-hit products use this (core.py:4500-4536). */
int olhip_trace(olhip_ctx* ctx, const olhip_scene* scene, const float* origins, const float* directions, uint32_t n,
                float* hits);

/* olhip_trace for rays already in device memory, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  d_origins, d_directions, d_hits: device float4[n] with
 * olhip_trace's layout; d_status: one device word, cleared by the call and
 * set non-zero when a traversal stack overflowed (the hits are then not
 * valid).  stream: a hipStream_t, or NULL for the context's stream; the
 * call only enqueues work on it.  The traversal stack's spill lives in a
 * fixed per-device array of OLHIP_TRACE_SCRATCH_BYTES: scenes whose
 * hierarchy needs more spill than that for one workgroup are refused
 * (OLHIP_ERR_VALUE).  Every call on a GPU shares that array, so calls that
 * overlap on one GPU (other streams, contexts or threads) race. */
#define OLHIP_TRACE_SCRATCH_BYTES (16u << 20)
int olhip_trace_device(olhip_ctx* ctx, const olhip_scene* scene, const float* d_origins, const float* d_directions,
                       uint32_t n, float* d_hits, uint32_t* d_status, void* stream);

/* ---- reference-estimator mode (renderers/reference/path_tracer.py) ----
 * The host owns the PCG64 stream; the device runs the per-bounce work.
 * Sequence per frame: e0_begin; per sample { e0_sample(jx, jy);
 * while (e0_trace(&n) and n > 0) e0_shade(deviates[n][3]); e0_sample_end };
 * e0_end(out). */
typedef struct olhip_e0_camera {
  float position[3];
  float forward[3];   /* unit, float32 as the reference computes it */
  float right[3];
  float up[3];
  double k_x, k_y;    /* see olhip_e0_begin */
  uint32_t projection;
} olhip_e0_camera;

typedef struct olhip_e0_env {
  uint32_t mode;      /* 0 gradient sky, 1 constant, 2 image */
  float color[3];
  float intensity;
  double rotation;
  const float* image; /* [h][w][3] when mode == 2 */
  uint32_t width, height;
} olhip_e0_env;

/* k_x, k_y: perspective aspect*tan(fov/2), tan(fov/2); orthographic
 * aspect*size/2, size/2; panoramic radians(hfov), radians(vfov).
 * brute_force = 1 scans triangles in index order (exact tie semantics). */
int olhip_e0_begin(olhip_ctx* ctx, const olhip_scene* scene, const olhip_e0_camera* camera, uint32_t width,
                   uint32_t height, const olhip_e0_env* env, uint32_t brute_force);
int olhip_e0_sample(olhip_ctx* ctx, const float* jitter_x, const float* jitter_y);
/* Traces the active rays, adds miss radiance, compacts the hits in ray
 * order; *n_hit receives how many normal triples the next shade needs. */
int olhip_e0_trace(olhip_ctx* ctx, uint32_t* n_hit, uint32_t* n_active);
/* deviates: float32 [n_hit][3] (rng.normal(...).astype(float32)). */
int olhip_e0_shade(olhip_ctx* ctx, const float* deviates, uint32_t n_hit);
int olhip_e0_sample_end(olhip_ctx* ctx);
/* out: float32 [height][width][4], alpha 0. */
int olhip_e0_end(olhip_ctx* ctx, uint32_t samples, float* out_rgba);
/* Primary-hit triangle of the last e0_trace call for ray i of the active
 * list (debug): writes n_active ints (-1 on a miss). */
int olhip_e0_last_hits(olhip_ctx* ctx, int32_t* tri, uint32_t n_active);

#ifdef __cplusplus
}
#endif

#endif /* OLHIP_H */
