// Library handles and the batched render engine (host side).
//
// The engine keeps a pool of path slots, shadow-entry slots and kMaxJobs
// device job records.  Jobs are admitted first-come first-served when a
// contiguous slot range for their next sample chunk is free (admission by
// memory), advance together one wavefront step at a time, and release
// their ranges when their paths have terminated.  A step is a fixed
// sequence of launches, captured once per step shape as a hipGraph and
// replayed.  Steps are pipelined: the next step is queued on the stream
// before the host reads the previous step's report, so report handling and
// launch overlap device work.  A worker thread can drive steps while other
// threads submit and wait (serving).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/olhip.h"
#include "e0.h"
#include "gi.h"
#include "internal.h"

struct olhip_scene {
  olhip_ctx* ctx = nullptr;
  DevBuf positions, normals, vidx, materials, analytic, area, area_guide;
  DevBuf tex_words, bindings, uvs, tangents, uv_density;  // textured scenes
  DevBuf face_orig, face_metal;  // per-triangle material parameters, original order
  BvhOutput bvh;  // the scene's BVH, or the top-level BVH over instances
  // Two-level scenes: bottom-level BVHs per resource, instance and BLAS tables.
  std::vector<std::unique_ptr<BvhOutput>> blas;
  DevBuf tlas_pos, tlas_vidx, inst_table, blas_table;
  uint32_t instance_count = 0, tlas_depth = 0, blas_depth = 0;
  DScene view{};
  uint32_t tri_count = 0, vertex_count = 0;
  bool has_transmission = false;
  uint32_t point_count = 0;
  std::vector<float> materials_host;
  // Fingerprint of the triangle structure (indices, material ids, resource
  // and instance layout): updates must keep it.
  uint64_t structure_hash = 0;
  uint32_t leaf_size = 0;
  std::vector<uint32_t> vertex_offsets, triangle_offsets;  // two-level: per resource
};

// Shadow entries one shaded hit of this job may emit (1 for unified NEE
// without point lights).
uint32_t job_entries_per_hit(const olhip_scene* sc, const olhip_gi_settings& s);

// First-fit allocator of [0, size) ranges.
class RangeAllocator {
 public:
  void reset(uint64_t size);
  // Start of a free range of `n`, or UINT64_MAX.
  uint64_t allocate(uint64_t n);
  void release(uint64_t begin, uint64_t n);
  uint64_t largest_free() const;
  uint64_t size() const { return size_; }

 private:
  std::map<uint64_t, uint64_t> free_;  // begin -> length
  uint64_t size_ = 0;
};

class Engine {
 public:
  Engine(olhip_ctx* ctx, const olhip_engine_config& cfg);
  ~Engine();

  uint64_t submit(const olhip_job_desc& job);
  // Queues the next step when there is work for it (keeping at most two in
  // flight), then waits for the oldest queued step and applies its report.
  // Returns the number of jobs not yet done (queued, resident or
  // accumulating) plus the steps still in flight.  Callers hold no lock; the
  // engine serialises steps.
  uint32_t step();
  void run_until_idle();
  int state(uint64_t id);
  // Copies the result (optional) and statistics, then forgets the job.
  void take_result(uint64_t id, float* out_rgba, olhip_job_stats* stats, bool keep = false);
  // Copies a done job's products (any pointer may be null).
  void products(uint64_t id, float* depth, float* normal, uint32_t* primitive, float* variance);
  void engine_stats(olhip_engine_stats* out);
  // Device image of a done job (valid until take_result).
  const float4* device_image(uint64_t id);
  // Exchanges a done job's image buffer with `buf` (no copy); the job keeps
  // the previous contents of `buf` as its image.
  void swap_image(uint64_t id, DevBuf& buf);

  void start_worker();
  void stop_worker();
  // Blocks until the job is done or failed, or the timeout (seconds, < 0 =
  // none) passes.  Returns the job state.
  int wait(uint64_t id, double timeout_s);

  olhip_ctx* ctx() const { return ctx_; }
  const olhip_engine_config& config() const { return cfg_; }
  // Per-step kernel timing (events around the traversal launch, closest hits
  // and the previous step's shadow rays, and around shading).
  void set_profiling(bool on) { profiling_ = on; }
  void set_work_counters(bool on);
  void read_work(unsigned long long out[kWorkWords]);
  double trace_ms = 0, shade_ms = 0;
  uint32_t trace_launches = 0;

 private:
  // A sample chunk of a job: resident in a device job slot while its paths
  // are traced, then finished until its accumulation is queued, in sample
  // order, together with the release of its slot ranges.
  struct Chunk {
    int slot = -1;
    uint32_t spp = 0, base_sample = 0;
    uint64_t region = 0, span = 0, ebase = 0, espan = 0;
    uint32_t steps_in_chunk = 0;  // steps whose reports were applied
    uint32_t steps_launched = 0;  // steps queued so far
    uint64_t admit_seq = 0;       // step that generates the chunk
  };
  struct Job {
    uint64_t id = 0;
    olhip_job_desc desc{};
    uint32_t per_hit = 1;
    bool transmits = false;
    // ReSTIR passes are admitted one per step while earlier passes are still
    // resident (a pass only reads the previous pass's reservoirs, written in
    // that pass's first step); other jobs keep one chunk resident at a time.
    bool stagger = false;
    bool buffers = false;       // accumulation and product buffers allocated
    uint32_t next_sample = 0;   // first sample not yet admitted
    uint32_t acc_next = 0;      // first sample whose accumulation is not yet queued
    std::vector<Chunk> resident, finished;
    uint32_t passed = 0;        // admission rounds in which this job did not fit
    int status = OLHIP_JOB_QUEUED;
    DevBuf acc, image, geo, prim, lum, variance, res, surf;
    olhip_job_stats stats{};
    std::string error;
  };

  void ensure_buffers();
  // A step's shape: queue parity, STEP_* kinds of resident work, and whether
  // it accumulates or generates (bits 1.. of the key).  Absent work is not
  // launched; each shape is captured once as a hipGraph and replayed.
  static uint32_t step_key(uint32_t in, uint32_t kinds, bool acc, bool gen, bool fused) {
    return in | (kinds << 1) | (acc ? kKeyAcc : 0u) | (gen ? kKeyGen : 0u) | (fused ? kKeyFused : 0u);
  }
  static constexpr uint32_t kKeyAcc = 64u, kKeyGen = 128u, kKeyFused = 256u;
  hipGraphExec_t graph_for(uint32_t key);
  void clear_graphs();
  void launch_step(uint32_t key, bool timed, int buf);
  // A step queued on the stream whose report has not been applied yet.
  struct InFlight {
    uint64_t seq = 0;
    bool traced = false;  // traversal ran (otherwise accumulation only)
    bool timed = false;
    int buf = 0;          // staging, report and event set of the step
    std::vector<uint64_t> finals;  // jobs whose last chunk this step accumulates
  };
  // Builds and queues the next step; false when there is nothing to do.
  bool launch_next();
  void retire_front();
  // Waits for the device before buffers that queued steps use are replaced.
  void quiesce();
  bool admit(Job& j, int slot);
  // Per-job device buffers come from, and return to, a pool of buffers of
  // forgotten jobs: consecutive frames of one size then allocate nothing.
  void take_buffer(DevBuf& dst, size_t bytes);
  void keep_buffers(Job& j);
  Chunk* chunk_in_slot(Job& j, int slot);
  void drop_chunks(Job& j);
  GiLaunch make_launch();
  double now_ms() const;

  olhip_ctx* ctx_;
  olhip_engine_config cfg_;
  std::mutex mu_;                 // jobs, queue, counters
  std::mutex step_mu_;            // one step at a time
  std::condition_variable cv_;    // job state changes / new work
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs_;
  std::deque<uint64_t> queue_;    // waiting for admission (FIFO)
  std::vector<uint64_t> slot_job_;  // kMaxJobs: job id per device slot (0 = free)
  std::vector<uint64_t> acc_jobs_;  // jobs with finished chunks not yet accumulated
  static constexpr size_t kSpareBytes = size_t(4) << 30;  // the context's pool of spare job buffers
  RangeAllocator paths_, entries_;
  uint64_t next_id_ = 1;
  uint32_t parity_ = 0;           // input queue of the next step
  uint64_t steps_ = 0;
  uint64_t seq_ = 0;              // sequence number of the next step queued
  std::deque<InFlight> inflight_;
  std::atomic<uint32_t> inflight_count_{0};  // inflight_.size() for readers without step_mu_
  uint64_t overflow_checked_from_ = 0;       // first step whose stack-overflow flag is still unhandled
  int max_depth_ = 1;
  bool stk_needed_ = false;
  bool restir_needed_ = false;
  bool tex_needed_ = false;

  // Device state.
  DevBuf T_, L_, M_, stk_, rsamp_, tex_, cone_, qo_[2], qd_[2], qp_[2], hit_, hit_inst_, gT_, gI_, eo_, ed_, ec_, ctr_, work_, jobs_dev_,
      list_dev_, spill_;
  uint64_t slots_ = 0, eslots_ = 0;
  Wave wave_{};
  GiLaunch launch_{};
  bool counting_ = false;
  bool profiling_ = false;
  std::unordered_map<uint32_t, hipGraphExec_t> graphs_;  // by step_key
  bool graph_valid_ = false;
  // Job table (host copy; the device copy is refreshed when it changes) and
  // pinned staging, one set per in-flight step.
  std::vector<JobDev> jobs_state_;
  bool jobs_dirty_ = true;
  JobDev* jobs_host_[2] = {nullptr, nullptr};
  StepList* list_host_[2] = {nullptr, nullptr};
  uint32_t* report_host_[2] = {nullptr, nullptr};
  hipEvent_t done_[2] = {nullptr, nullptr};
  hipEvent_t begin_[2] = {nullptr, nullptr};  // step start on the stream (device-time accounting)
  hipEvent_t prof_ev_[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double t0_ms_ = 0;
  // Engine totals.
  olhip_engine_stats totals_{};

  std::thread worker_;
  std::atomic<bool> stop_{false};
  bool worker_running_ = false;
};

struct olhip_engine {
  std::unique_ptr<Engine> engine;
};

struct olhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int profiling = 0;
  int counting = 0;
  int lds_stack = 16;
  int trace_block = 128;
  int fused_primary = 0;  // olhip_set_fused_primary
  // Device buffers of forgotten jobs, shared by the context's engines (an
  // engine created later reuses what earlier ones released, so a new server
  // allocates nothing for jobs of sizes already seen), kept under
  // Engine::kSpareBytes.  Declared before the engines it outlives.
  std::mutex spare_mu;
  std::vector<DevBuf> spare;
  size_t spare_bytes = 0;
  // Engine of the single-frame entry point (created on first use).
  std::unique_ptr<Engine> frame_engine;
  uint64_t frame_engine_slots = 0;
  uint32_t frame_engine_entries = 0;
  DevBuf image;  // last single frame (olhip_frame_download, olhip_frame_output)
  uint32_t image_pixels = 0, image_width = 0, image_height = 0;
  DevBuf spill;  // olhip_trace
  // Device address of the fixed traversal spill array (olhip_trace_device).
  uint32_t* trace_scratch = nullptr;
  // Reference-estimator frame.
  struct E0Frame {
    DevBuf o, d, T, rad, acc, active, flags, scan, ht, htri, hitlist, hit_t, hit_tri, jx, jy, gauss, env, ctr, out;
    E0Params params{};
    E0State state{};
    const olhip_scene* scene = nullptr;
    uint32_t n_active = 0, n_hit = 0;
    int stack = 32;
    bool open = false;
  } e0;
};
