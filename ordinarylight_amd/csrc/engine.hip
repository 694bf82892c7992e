// Batched render engine: admission, sample chunking, pipelined step launch
// (direct or captured hipGraph), per-job reports and the serving worker
// thread.

#include <algorithm>
#include <chrono>
#include <cstring>

#include "engine.h"

// ---- RangeAllocator -------------------------------------------------------------------
void RangeAllocator::reset(uint64_t size) {
  free_.clear();
  size_ = size;
  if (size) free_[0] = size;
}

uint64_t RangeAllocator::allocate(uint64_t n) {
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < n) continue;
    const uint64_t begin = it->first, len = it->second;
    free_.erase(it);
    if (len > n) free_[begin + n] = len - n;
    return begin;
  }
  return UINT64_MAX;
}

void RangeAllocator::release(uint64_t begin, uint64_t n) {
  if (n == 0) return;
  auto next = free_.lower_bound(begin);
  uint64_t b = begin, len = n;
  if (next != free_.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == b) {
      b = prev->first;
      len += prev->second;
      free_.erase(prev);
    }
  }
  next = free_.lower_bound(begin);
  if (next != free_.end() && b + len == next->first) {
    len += next->second;
    free_.erase(next);
  }
  free_[b] = len;
}

uint64_t RangeAllocator::largest_free() const {
  uint64_t m = 0;
  for (const auto& kv : free_) m = std::max(m, kv.second);
  return m;
}

// ---- helpers ----------------------------------------------------------------------------
namespace {

uint32_t stripes_for(uint64_t n) {
  return (uint32_t)std::min<uint64_t>(kJobStripes, std::max<uint64_t>(1, (n + 65535) / 65536));
}

// Slots per segment: a job's shade chunks number at most n / 64 + stripes,
// so (ceil(n / (64 * stripes)) + 1) * 64 slots per segment cannot overflow.
uint32_t seg_for(uint64_t n, uint32_t stripes) {
  return (uint32_t)(((n + (uint64_t)kChunk * stripes - 1) / ((uint64_t)kChunk * stripes) + 1) * kChunk);
}

}  // namespace

uint32_t job_entries_per_hit(const olhip_scene* sc, const olhip_gi_settings& s) {
  const bool two_pass = sc->point_count > (sc->view.env_valid ? 1u : 0u) || !s.unified_secondary_nee;
  uint32_t n = 1;
  if (two_pass)
    n = sc->point_count +
        (s.unified_secondary_nee ? 1u : (std::min<uint32_t>(std::max<uint32_t>(s.secondary_area_samples, 1u), 16u) + 4u));
  // ReSTIR primary hits: point lights, the reservoir's sample and up to four
  // environment samples.
  if (s.restir_candidates) n = std::max(n, sc->point_count + 1u + std::min<uint32_t>(s.environment_samples, 4u));
  return n;
}

namespace {

void require_arg(bool ok, const char* message) {
  if (!ok) throw OlError{OLHIP_ERR_VALUE, message};
}

template <class T>
T* pinned(size_t count) {
  void* p = nullptr;
  OL_HIP_CHECK(hipHostMalloc(&p, sizeof(T) * count, hipHostMallocDefault));
  std::memset(p, 0, sizeof(T) * count);
  return static_cast<T*>(p);
}

}  // namespace

// ---- Engine -------------------------------------------------------------------------------
Engine::Engine(olhip_ctx* ctx, const olhip_engine_config& cfg) : ctx_(ctx), cfg_(cfg) {
  if (cfg_.path_slots == 0) cfg_.path_slots = 1ull << 25;
  if (cfg_.entry_factor == 0) cfg_.entry_factor = 2;
  if (cfg_.lds_stack == 0) cfg_.lds_stack = ctx->lds_stack;
  if (cfg_.trace_block == 0) cfg_.trace_block = ctx->trace_block;
  require_arg(cfg_.path_slots >= 4096 && cfg_.path_slots < (1ull << 32), "path_slots must be in [4096, 2^32)");
  require_arg(cfg_.entry_factor <= 128 && cfg_.path_slots * cfg_.entry_factor < (1ull << 32),
              "entry_factor * path_slots must be below 2^32");
  slots_ = cfg_.path_slots;
  eslots_ = cfg_.path_slots * cfg_.entry_factor;
  OL_HIP_CHECK(hipSetDevice(ctx_->device));
  slot_job_.assign(kMaxJobs, 0);
  paths_.reset(slots_);
  entries_.reset(eslots_);
  jobs_state_.assign(kMaxJobs, JobDev{});
  for (int b = 0; b < 2; ++b) {
    jobs_host_[b] = pinned<JobDev>(kMaxJobs);
    list_host_[b] = pinned<StepList>(1);
    report_host_[b] = pinned<uint32_t>(C_WORDS - C_OVERFLOW);
    OL_HIP_CHECK(hipEventCreate(&done_[b]));
    OL_HIP_CHECK(hipEventCreate(&begin_[b]));
  }
  t0_ms_ = 0;
  t0_ms_ = now_ms();
  ctr_.reserve(4ull * C_WORDS);
  work_.reserve(64);
  jobs_dev_.reserve(sizeof(JobDev) * kMaxJobs);
  list_dev_.reserve(sizeof(StepList));
  ensure_buffers();
  launch_ = make_launch();
  wave_clear(wave_, ctx_->stream);
  OL_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  totals_.path_slots = slots_;
  totals_.entry_slots = eslots_;
}

Engine::~Engine() {
  stop_worker();
  (void)hipSetDevice(ctx_->device);
  (void)hipStreamSynchronize(ctx_->stream);
  // Jobs never taken leave their buffers to the context's pool.
  for (auto& kv : jobs_) keep_buffers(*kv.second);
  for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  for (int b = 0; b < 2; ++b) {
    if (done_[b]) (void)hipEventDestroy(done_[b]);
    if (begin_[b]) (void)hipEventDestroy(begin_[b]);
    for (auto& x : prof_ev_[b])
      if (x) (void)hipEventDestroy(x);
    if (jobs_host_[b]) (void)hipHostFree(jobs_host_[b]);
    if (list_host_[b]) (void)hipHostFree(list_host_[b]);
    if (report_host_[b]) (void)hipHostFree(report_host_[b]);
  }
}

void Engine::quiesce() {
  if (!inflight_.empty()) OL_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

double Engine::now_ms() const {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count() - t0_ms_;
}

void Engine::ensure_buffers() {
  quiesce();
  T_.reserve(16 * slots_);
  L_.reserve(16 * slots_);
  M_.reserve(16 * slots_);
  for (int q = 0; q < 2; ++q) {
    qo_[q].reserve(16 * slots_);
    qd_[q].reserve(16 * slots_);
    qp_[q].reserve(4 * slots_);
  }
  hit_.reserve(16 * slots_);
  hit_inst_.reserve(4 * slots_);
  gT_.reserve(16 * slots_);
  gI_.reserve(16 * slots_);
  eo_.reserve(16 * eslots_);
  ed_.reserve(16 * eslots_);
  ec_.reserve(16 * eslots_);
  if (stk_needed_) stk_.reserve(64 * slots_);
  if (restir_needed_) rsamp_.reserve(48 * slots_);
  if (tex_needed_) {
    tex_.reserve(64 * slots_);
    cone_.reserve(8 * slots_);
  }
  Wave w{};
  w.jobs = jobs_dev_.as<JobDev>();
  w.list = list_dev_.as<StepList>();
  w.T = T_.as<float4>();
  w.L = L_.as<float4>();
  w.M = M_.as<uint4>();
  w.stk = stk_needed_ ? stk_.as<float4>() : nullptr;
  w.rsamp = restir_needed_ ? rsamp_.as<float4>() : nullptr;
  w.tex = tex_needed_ ? tex_.as<float4>() : nullptr;
  w.cone = tex_needed_ ? cone_.as<float2>() : nullptr;
  for (int q = 0; q < 2; ++q) {
    w.qo[q] = qo_[q].as<float4>();
    w.qd[q] = qd_[q].as<float4>();
    w.qp[q] = qp_[q].as<uint32_t>();
  }
  w.hit = hit_.as<float4>();
  w.hit_inst = hit_inst_.as<uint32_t>();
  w.gT = gT_.as<float4>();
  w.gI = gI_.as<uint4>();
  w.eo = eo_.as<float4>();
  w.ed = ed_.as<float4>();
  w.ec = ec_.as<float4>();
  w.ctr = ctr_.as<uint32_t>();
  w.work = counting_ ? work_.as<unsigned long long>() : nullptr;
  wave_ = w;
  graph_valid_ = false;
}

GiLaunch Engine::make_launch() {
  quiesce();
  GiLaunch l{};
  l.cu_count = ctx_->cu_count;
  l.stack_depth = (int)(3u * (uint32_t)std::max(max_depth_, 1) + 1u);
  l.lds_stack = cfg_.lds_stack;
  l.trace_block = cfg_.trace_block;
  const uint64_t words = trace_spill_words(l);
  if (words) spill_.reserve(4 * words);
  l.spill = words ? spill_.as<uint32_t>() : nullptr;
  l.spill_words = words;
  graph_valid_ = false;
  return l;
}

void Engine::set_work_counters(bool on) {
  std::lock_guard<std::mutex> g(step_mu_);
  if (on == counting_) return;
  counting_ = on;
  OL_HIP_CHECK(hipMemsetAsync(work_.ptr, 0, 8 * kWorkWords, ctx_->stream));
  ensure_buffers();
}

void Engine::read_work(unsigned long long out[kWorkWords]) {
  OL_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  OL_HIP_CHECK(hipMemcpy(out, work_.ptr, 8 * kWorkWords, hipMemcpyDeviceToHost));
}

uint64_t Engine::submit(const olhip_job_desc& d) {
  require_arg(d.scene != nullptr, "job scene must not be NULL");
  require_arg(d.scene->ctx == ctx_, "scene belongs to another context");
  require_arg(d.width >= 1 && d.height >= 1, "width and height must be positive");
  require_arg((uint64_t)d.width * d.height <= (1ull << 28), "image too large");
  require_arg(d.samples >= 1 && d.samples <= 65536, "samples must be between 1 and 65536");
  const olhip_gi_settings& s = d.settings;
  require_arg(s.max_bounces >= 1 && s.max_bounces <= 64, "max_bounces must be between 1 and 64");
  require_arg(d.camera.projection <= 2, "unknown projection");
  require_arg(s.secondary_nee_probability > 0.0f && s.secondary_nee_probability <= 1.0f,
              "secondary_nee_probability must be in (0, 1]");
  require_arg(s.russian_roulette_min_survival >= 0.0f && s.russian_roulette_min_survival <= 1.0f,
              "russian_roulette_min_survival must be in [0, 1]");
  require_arg(s.environment_samples <= 4, "environment_samples must be between 0 and 4");
  require_arg((d.products & ~(OLHIP_PRODUCT_PRIMARY | OLHIP_PRODUCT_VARIANCE)) == 0, "unknown product bits");
  if (s.restir_candidates) {
    require_arg(s.restir_candidates <= 64, "restir_candidates must be at most 64");
    require_arg(s.restir_spatial_neighbors <= 8, "restir_spatial_neighbors must be at most 8");
    require_arg(s.restir_spatial_neighbors == 0 || (s.restir_spatial_radius >= 1 && s.restir_spatial_radius <= 32),
                "restir_spatial_radius must be between 1 and 32");
    require_arg(s.restir_history_limit >= 1 && s.restir_history_limit <= 64,
                "restir_history_limit must be between 1 and 64");
    require_arg(s.restir_pairwise_mis <= 1, "restir_pairwise_mis must be 0 or 1");
  }
  auto j = std::make_unique<Job>();
  j->desc = d;
  j->per_hit = job_entries_per_hit(d.scene, s);
  j->transmits = d.scene->has_transmission;
  j->stagger = s.restir_candidates > 0;
  // One sample chunk must fit an empty engine.
  const uint64_t n = (uint64_t)d.width * d.height;
  const uint32_t S = stripes_for(n);
  const uint64_t span = (uint64_t)S * seg_for(n, S);
  if (span > slots_ || span * j->per_hit > eslots_)
    throw OlError{OLHIP_ERR_MEMORY, "one sample of this job (" + std::to_string(n) +
                                        " paths) exceeds the engine's path or shadow-entry slots"};
  if (stripes_for(n) > kJobStripes) throw OlError{OLHIP_ERR_VALUE, "job too large"};
  std::lock_guard<std::mutex> g(mu_);
  j->id = next_id_++;
  j->stats.submit_ms = now_ms();
  const uint64_t id = j->id;
  queue_.push_back(id);
  jobs_[id] = std::move(j);
  cv_.notify_all();
  return id;
}

// Admits the job's next sample chunk into device slot `slot` if a slot range
// is free.  Called with mu_ held.
bool Engine::admit(Job& j, int slot) {
  const olhip_job_desc& d = j.desc;
  const uint64_t np = (uint64_t)d.width * d.height;
  uint64_t spp = d.samples - j.next_sample;
  if (d.settings.path_capacity) spp = std::min<uint64_t>(spp, std::max<uint64_t>(1, d.settings.path_capacity / np));
  // ReSTIR: one sample per pass, each reusing the previous pass's reservoirs.
  if (d.settings.restir_candidates) spp = std::min<uint64_t>(spp, 1);
  const uint64_t free_p = paths_.largest_free(), free_e = entries_.largest_free();
  while (spp >= 1) {
    const uint64_t n = spp * np;
    const uint32_t S = stripes_for(n);
    const uint32_t seg = seg_for(n, S);
    const uint64_t span = (uint64_t)S * seg, espan = span * j.per_hit;
    if (span <= free_p && espan <= free_e) {
      const uint64_t base = paths_.allocate(span);
      const uint64_t ebase = entries_.allocate(espan);
      if (base == UINT64_MAX || ebase == UINT64_MAX) throw OlError{OLHIP_ERR_RUNTIME, "slot allocator inconsistency"};
      Chunk c;
      c.slot = slot;
      c.region = base;
      c.span = span;
      c.ebase = ebase;
      c.espan = espan;
      c.spp = (uint32_t)spp;
      c.base_sample = j.next_sample;
      c.admit_seq = seq_;
      j.next_sample += (uint32_t)spp;
      j.resident.push_back(c);
      jobs_dirty_ = true;
      JobDev& J = jobs_state_[slot];
      std::memset(&J, 0, sizeof(J));
      GiParams& P = J.P;
      P.sc = d.scene->view;
      for (int k = 0; k < 3; ++k) {
        P.cam_o[k] = d.camera.position[k];
        P.cam_f[k] = d.camera.forward[k];
        P.cam_r[k] = d.camera.right[k];
        P.cam_u[k] = d.camera.up[k];
      }
      P.projection = d.camera.projection;
      P.width = d.width;
      P.height = d.height;
      P.spp = (uint32_t)spp;
      // The reference stores the frame index as float32 in the camera block.
      P.frame = (uint32_t)((float)d.frame_index + 0.5f);
      P.sample_base = c.base_sample;
      P.max_bounces = d.settings.max_bounces;
      P.point_count = d.scene->point_count;
      P.area_count = d.scene->view.area_count;
      P.secondary_area_samples = d.settings.secondary_area_samples;
      P.environment_samples = d.settings.environment_samples;
      P.roulette_start = d.settings.russian_roulette_start;
      P.roulette_min_survival = d.settings.russian_roulette_min_survival;
      P.nee_probability = d.settings.secondary_nee_probability;
      P.unified = d.settings.unified_secondary_nee ? 1u : 0u;
      const bool two_pass = d.scene->point_count > (d.scene->view.env_valid ? 1u : 0u) || !P.unified;
      P.two_pass = two_pass ? 1u : 0u;
      J.base = (uint32_t)base;
      J.n_paths = (uint32_t)n;
      J.stripes = S;
      J.seg = seg;
      J.ebase = (uint32_t)ebase;
      J.eseg = seg * j.per_hit;
      J.append = cfg_.append_limit ? std::min<uint32_t>(seg, cfg_.append_limit) : seg;
      if (!j.buffers) {
        take_buffer(j.acc, sizeof(double4) * np);
        take_buffer(j.image, sizeof(float4) * np);
        if (d.products & OLHIP_PRODUCT_PRIMARY) {
          take_buffer(j.geo, sizeof(float4) * np);
          take_buffer(j.prim, sizeof(uint32_t) * np);
        }
        if (d.products & OLHIP_PRODUCT_VARIANCE) {
          take_buffer(j.lum, sizeof(double2) * np);
          take_buffer(j.variance, sizeof(float) * np);
        }
        j.stats.admit_ms = now_ms();
      }
      if (d.settings.restir_candidates) {
        if (!j.buffers) {
          take_buffer(j.res, sizeof(uint4) * 2 * np);
          take_buffer(j.surf, sizeof(float4) * 4 * np);
        }
        const uint32_t pass = c.base_sample;
        P.restir = d.settings.restir_candidates;
        P.restir_spatial = d.settings.restir_spatial_neighbors;
        P.restir_radius = d.settings.restir_spatial_radius;
        P.restir_pairwise = d.settings.restir_pairwise_mis;
        P.restir_limit = (float)d.settings.restir_history_limit;
        P.restir_history = pass > 0 ? 1u : 0u;
        J.res_out = j.res.as<uint4>() + (pass & 1u) * np;
        J.res_in = j.res.as<uint4>() + ((pass + 1u) & 1u) * np;
        J.surf_out = j.surf.as<float4>() + (pass & 1u) * 2 * np;
        J.surf_in = j.surf.as<float4>() + ((pass + 1u) & 1u) * 2 * np;
      }
      j.buffers = true;
      J.geo = (d.products & OLHIP_PRODUCT_PRIMARY) ? j.geo.as<float4>() : nullptr;
      J.prim = (d.products & OLHIP_PRODUCT_PRIMARY) ? j.prim.as<uint32_t>() : nullptr;
      j.status = OLHIP_JOB_RUNNING;
      ++j.stats.chunks;
      slot_job_[slot] = j.id;
      if (j.transmits && !stk_needed_) {
        stk_needed_ = true;
        ensure_buffers();
      }
      if (d.settings.restir_candidates && !restir_needed_) {
        restir_needed_ = true;
        ensure_buffers();
      }
      if (d.scene->view.textured && !tex_needed_) {
        tex_needed_ = true;
        ensure_buffers();
      }
      const int depth = (int)d.scene->bvh.depth;
      if (depth > max_depth_) {
        max_depth_ = depth;
        launch_ = make_launch();
      }
      return true;
    }
    // Shrink towards what the largest free range can hold.
    const uint64_t fit = std::min(free_p * spp / std::max<uint64_t>(span, 1), free_e * spp / std::max<uint64_t>(espan, 1));
    spp = std::min(spp - 1, fit);
  }
  return false;
}

// The smallest spare buffer of the context's pool that holds `bytes` (its
// contents are not preserved), else a new allocation.  Called with mu_ held.
void Engine::take_buffer(DevBuf& dst, size_t bytes) {
  {
    std::lock_guard<std::mutex> g(ctx_->spare_mu);
    std::vector<DevBuf>& spare = ctx_->spare;
    size_t best = spare.size();
    for (size_t k = 0; k < spare.size(); ++k)
      if (spare[k].bytes >= bytes && (best == spare.size() || spare[k].bytes < spare[best].bytes)) best = k;
    if (best != spare.size()) {
      ctx_->spare_bytes -= spare[best].bytes;
      dst = std::move(spare[best]);
      spare.erase(spare.begin() + best);
      return;
    }
  }
  dst.reserve(bytes);
}

// Keeps a forgotten job's buffers in the context's pool for later jobs,
// evicting the oldest beyond kSpareBytes.  Called once the job's last step
// has completed (no queued work uses the buffers).
void Engine::keep_buffers(Job& j) {
  std::lock_guard<std::mutex> g(ctx_->spare_mu);
  for (DevBuf* b : {&j.acc, &j.image, &j.geo, &j.prim, &j.lum, &j.variance, &j.res, &j.surf}) {
    if (!b->ptr) continue;
    ctx_->spare_bytes += b->bytes;
    ctx_->spare.push_back(std::move(*b));
  }
  while (ctx_->spare_bytes > kSpareBytes && !ctx_->spare.empty()) {
    ctx_->spare_bytes -= ctx_->spare.front().bytes;
    ctx_->spare.erase(ctx_->spare.begin());
  }
}

Engine::Chunk* Engine::chunk_in_slot(Job& j, int slot) {
  for (Chunk& c : j.resident)
    if (c.slot == slot) return &c;
  throw OlError{OLHIP_ERR_RUNTIME, "engine slot table inconsistency"};
}

// Releases the slot ranges of a failed job's chunks and forgets them.
void Engine::drop_chunks(Job& j) {
  for (const std::vector<Chunk>* v : {&j.resident, &j.finished})
    for (const Chunk& c : *v) {
      paths_.release(c.region, c.span);
      entries_.release(c.ebase, c.espan);
      if (c.slot >= 0) slot_job_[c.slot] = 0;
    }
  j.resident.clear();
  j.finished.clear();
  queue_.erase(std::remove(queue_.begin(), queue_.end(), j.id), queue_.end());
  acc_jobs_.erase(std::remove(acc_jobs_.begin(), acc_jobs_.end(), j.id), acc_jobs_.end());
}

void Engine::launch_step(uint32_t key, bool timed, int buf) {
  hipStream_t s = ctx_->stream;
  const uint32_t in = key & 1u, kinds = (key >> 1) & kStepKinds;
  const bool gen = (key & kKeyGen) != 0u;
  if (key & kKeyAcc) wave_accumulate(wave_, ctx_->cu_count, s);
  if (gen) wave_generate(wave_, in, ctx_->cu_count, s, (key & kKeyFused) == 0u);
  if (key & kKeyFused) {
    // Traversal and shading in one kernel, timed as traversal.
    hipEvent_t* ev = prof_ev_[buf];
    if (timed) {
      for (int k = 0; k < 3; ++k)
        if (!ev[k]) OL_HIP_CHECK(hipEventCreate(&ev[k]));
      OL_HIP_CHECK(hipEventRecord(ev[0], s));
    }
    wave_primary(launch_, wave_, in, kinds, s);
    if (timed) {
      OL_HIP_CHECK(hipEventRecord(ev[1], s));
      OL_HIP_CHECK(hipEventRecord(ev[2], s));
    }
  } else if (!timed) {
    wave_trace(launch_, wave_, in, kinds, s);
    wave_shade(launch_, wave_, in, kinds, gen, s);
  } else {
    // Read when the step is retired (no synchronisation inside the step).
    hipEvent_t* ev = prof_ev_[buf];
    for (int k = 0; k < 3; ++k)
      if (!ev[k]) OL_HIP_CHECK(hipEventCreate(&ev[k]));
    OL_HIP_CHECK(hipEventRecord(ev[0], s));
    wave_trace(launch_, wave_, in, kinds, s);
    OL_HIP_CHECK(hipEventRecord(ev[1], s));
    wave_shade(launch_, wave_, in, kinds, gen, s);
    OL_HIP_CHECK(hipEventRecord(ev[2], s));
  }
  wave_epilogue(wave_, in, s);
}

void Engine::clear_graphs() {
  quiesce();
  for (auto& kv : graphs_) OL_HIP_CHECK(hipGraphExecDestroy(kv.second));
  graphs_.clear();
  graph_valid_ = true;
}

hipGraphExec_t Engine::graph_for(uint32_t key) {
  if (!graph_valid_) clear_graphs();
  auto it = graphs_.find(key);
  if (it != graphs_.end()) return it->second;
  hipStream_t s = ctx_->stream;
  hipGraph_t g = nullptr;
  quiesce();
  OL_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    launch_step(key, false, 0);
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  OL_HIP_CHECK(hipStreamEndCapture(s, &g));
  hipGraphExec_t exec = nullptr;
  OL_HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
  OL_HIP_CHECK(hipGraphDestroy(g));
  graphs_[key] = exec;
  return exec;
}

uint32_t Engine::step() {
  std::lock_guard<std::mutex> step_guard(step_mu_);
  OL_HIP_CHECK(hipSetDevice(ctx_->device));
  const double t_begin = now_ms();
  try {
    // One step stays queued behind the one the host waits for, so the report
    // handling and the next launch overlap device work.  Slots of chunks
    // that finish are then reused one step later than a step built after
    // every report would reuse them; keeping the device fed is worth more
    // (on the mixed 192-job stream at 1.0x and 1.5x its batched capacity:
    // device busy 0.95 -> 0.99 of the wall time, 3.9 % and 1.5 % more jobs/s,
    // p99 no higher).
    constexpr size_t depth = 2;
    while (inflight_.size() < depth && launch_next()) {
    }
    if (!inflight_.empty()) retire_front();
  } catch (...) {
    // The queued steps' reports can no longer be matched to jobs.
    (void)hipStreamSynchronize(ctx_->stream);
    inflight_.clear();
    inflight_count_ = 0;
    throw;
  }
  inflight_count_ = (uint32_t)inflight_.size();
  uint32_t pending = 0;
  {
    std::lock_guard<std::mutex> g(mu_);
    totals_.steps = steps_;
    totals_.busy_ms += now_ms() - t_begin;
    for (const auto& kv : jobs_)
      if (kv.second->status == OLHIP_JOB_QUEUED || kv.second->status == OLHIP_JOB_RUNNING) ++pending;
    cv_.notify_all();
  }
  return pending + (uint32_t)inflight_.size();
}

// Builds the next step from the state the applied reports leave: admission,
// the chunks to accumulate, and the resident chunks that may still hold live
// paths.  A chunk's paths end by its max_bounces-th step, so steps queued
// ahead of a report are only issued for chunks that can still need them; a
// step that finds no paths of a chunk (they all ended earlier) costs only
// empty launches.
bool Engine::launch_next() {
  const int b = (int)(seq_ & 1u);
  StepList& Ls = *list_host_[b];
  bool traced = false, all_fresh = true;
  uint32_t kinds = 0;
  InFlight f;
  f.seq = seq_;
  f.buf = b;
  {
    std::lock_guard<std::mutex> g(mu_);
    // Accumulation: for each job, the finished chunk next in sample order,
    // at most one per job and step (two items of one job would race on its
    // pixel sums).  Its slot ranges are released here, before admission: a
    // chunk generated in this step may reuse them, since accumulation runs
    // first.
    Ls.n_acc = 0;
    Ls.acc_prefix[0] = 0;
    for (auto it = acc_jobs_.begin(); it != acc_jobs_.end();) {
      Job& j = *jobs_[*it];
      auto c = std::find_if(j.finished.begin(), j.finished.end(),
                            [&](const Chunk& x) { return x.base_sample == j.acc_next; });
      if (c == j.finished.end() || Ls.n_acc == kMaxJobs) {
        ++it;
        continue;
      }
      const uint64_t np = (uint64_t)j.desc.width * j.desc.height;
      AccItem a{};
      a.base = (uint32_t)c->region;
      a.np = (uint32_t)np;
      a.spp = c->spp;
      a.first_chunk = c->base_sample == 0 ? 1u : 0u;
      const bool last = c->base_sample + c->spp >= j.desc.samples;
      a.final_samples = last ? j.desc.samples : 0u;
      a.acc = j.acc.as<double4>();
      a.image = j.image.as<float4>();
      a.lum = (j.desc.products & OLHIP_PRODUCT_VARIANCE) ? j.lum.as<double2>() : nullptr;
      a.variance = (j.desc.products & OLHIP_PRODUCT_VARIANCE) ? j.variance.as<float>() : nullptr;
      Ls.acc[Ls.n_acc] = a;
      Ls.acc_prefix[Ls.n_acc + 1] = Ls.acc_prefix[Ls.n_acc] + a.np;
      ++Ls.n_acc;
      paths_.release(c->region, c->span);
      entries_.release(c->ebase, c->espan);
      j.acc_next += c->spp;
      j.finished.erase(c);
      if (last) f.finals.push_back(j.id);
      if (j.finished.empty()) it = acc_jobs_.erase(it);
      else ++it;
    }
    Ls.n_gen = 0;
    Ls.gen_prefix[0] = 0;
    Ls.fresh = 0ull;
    // Admission: first-fit over the queue; a job passed over for 8 steps
    // blocks the ones behind it so it cannot starve.  A staggered job stays
    // queued until its last pass is admitted, one pass per step.
    for (auto it = queue_.begin(); it != queue_.end();) {
      int slot = -1;
      for (uint32_t k = 0; k < kMaxJobs; ++k)
        if (slot_job_[k] == 0) {
          slot = (int)k;
          break;
        }
      if (slot < 0) break;
      Job& j = *jobs_[*it];
      if (admit(j, slot)) {
        Ls.gen_job[Ls.n_gen] = (uint32_t)slot;
        Ls.fresh |= 1ull << slot;
        Ls.gen_prefix[Ls.n_gen + 1] = Ls.gen_prefix[Ls.n_gen] + jobs_state_[slot].n_paths;
        ++Ls.n_gen;
        if (j.stagger && j.next_sample < j.desc.samples) ++it;
        else it = queue_.erase(it);
        continue;
      }
      // A staggered job waiting for memory its own passes hold does not
      // block the jobs behind it.
      if (j.resident.empty() && ++j.passed > 8) break;
      ++it;
    }
    for (uint32_t k = 0; k < kMaxJobs; ++k) {
      if (slot_job_[k] == 0) continue;
      Job& j = *jobs_[slot_job_[k]];
      Chunk& c = *chunk_in_slot(j, (int)k);
      if (c.steps_launched >= j.desc.settings.max_bounces) continue;
      ++c.steps_launched;
      traced = true;
      if (!((Ls.fresh >> k) & 1ull)) all_fresh = false;
      kinds |= j.desc.scene->view.instanced ? STEP_INSTANCED : STEP_FLAT;
      // The reservoir pass and its replay in shading run at primary hits
      // only: the step that generated the chunk.
      if (j.desc.settings.restir_candidates && ((Ls.fresh >> k) & 1ull)) kinds |= STEP_RESTIR;
      if (j.desc.scene->view.textured) kinds |= STEP_TEXTURED;
      if (jobs_state_[k].P.two_pass) kinds |= STEP_TWO_PASS;
    }
    totals_.resident_jobs = 0;
    for (uint32_t k = 0; k < kMaxJobs; ++k) totals_.resident_jobs += slot_job_[k] != 0;
    totals_.max_resident_jobs = std::max(totals_.max_resident_jobs, totals_.resident_jobs);
    totals_.queued_jobs = (uint32_t)queue_.size();
  }
  if (!traced && Ls.n_acc == 0) return false;
  hipStream_t s = ctx_->stream;
  OL_HIP_CHECK(hipEventRecord(begin_[b], s));
  if (jobs_dirty_) {
    std::memcpy(jobs_host_[b], jobs_state_.data(), sizeof(JobDev) * kMaxJobs);
    OL_HIP_CHECK(hipMemcpyAsync(jobs_dev_.ptr, jobs_host_[b], sizeof(JobDev) * kMaxJobs, hipMemcpyHostToDevice, s));
    jobs_dirty_ = false;
  }
  // Only accumulation and generation read the step list.
  if (Ls.n_acc || Ls.n_gen)
    OL_HIP_CHECK(hipMemcpyAsync(list_dev_.ptr, list_host_[b], sizeof(StepList), hipMemcpyHostToDevice, s));
  const uint32_t in = parity_;
  f.traced = traced;
  f.timed = traced && profiling_;
  // Hybrid strategy: a step of fresh camera rays only, of flat untextured
  // scenes without ReSTIR, runs traversal and shading as one kernel.
  const bool fused = cfg_.fused_primary != 0u && traced && all_fresh && (kinds & ~STEP_TWO_PASS) == STEP_FLAT;
  if (fused) {
    std::lock_guard<std::mutex> g(mu_);
    ++totals_.fused_steps;
  }
  const uint32_t key = step_key(in, kinds, Ls.n_acc > 0, Ls.n_gen > 0, fused);
  if (!traced) {
    // Only accumulation left: no traversal launches.
    wave_accumulate(wave_, ctx_->cu_count, s);
  } else if (profiling_ || !cfg_.use_graph || seq_ == 0) {
    launch_step(key, profiling_, b);
  } else {
    OL_HIP_CHECK(hipGraphLaunch(graph_for(key), s));
  }
  OL_HIP_CHECK(hipGetLastError());
  OL_HIP_CHECK(hipMemcpyAsync(report_host_[b], ctr_.as<uint32_t>() + C_OVERFLOW, 4ull * (C_WORDS - C_OVERFLOW),
                              hipMemcpyDeviceToHost, s));
  OL_HIP_CHECK(hipEventRecord(done_[b], s));
  if (traced) parity_ ^= 1u;
  ++seq_;
  inflight_.push_back(std::move(f));
  return true;
}

// Waits for the oldest queued step and applies its report.  Reports are
// matched to the chunks that were resident when the step was built: a slot
// released and re-admitted since then belongs to a later chunk, whose
// admission step is later than the report's.
void Engine::retire_front() {
  InFlight f = std::move(inflight_.front());
  inflight_.pop_front();
  OL_HIP_CHECK(hipEventSynchronize(done_[f.buf]));
  float step_ms = 0.0f;
  OL_HIP_CHECK(hipEventElapsedTime(&step_ms, begin_[f.buf], done_[f.buf]));
  if (f.timed) {
    float a = 0, b = 0;
    OL_HIP_CHECK(hipEventElapsedTime(&a, prof_ev_[f.buf][0], prof_ev_[f.buf][1]));
    OL_HIP_CHECK(hipEventElapsedTime(&b, prof_ev_[f.buf][1], prof_ev_[f.buf][2]));
    trace_ms += a;
    shade_ms += b;
    ++trace_launches;
  }
  bool reset_overflow = false;
  {
    std::lock_guard<std::mutex> g(mu_);
    const double t = now_ms();
    totals_.device_ms += step_ms;
    for (uint64_t id : f.finals) {
      auto it = jobs_.find(id);
      if (it == jobs_.end() || it->second->status == OLHIP_JOB_FAILED) continue;
      it->second->status = OLHIP_JOB_DONE;
      it->second->stats.done_ms = t;
      ++totals_.jobs_done;
    }
    if (f.traced) {
      ++steps_;
      const uint32_t* report = report_host_[f.buf];
      const uint32_t* rep = report + (C_REPORT - C_OVERFLOW);
      // The overflow flags are sticky until reset; a reset queued after an
      // overflow follows the steps already in flight, whose reports still
      // show it.  Those steps' jobs were failed with the first report.
      const bool fresh = f.seq >= overflow_checked_from_;
      const bool stack_overflow = report[C_STACK - C_OVERFLOW] != 0 && fresh;
      const bool queue_overflow = report[0] != 0 && fresh;
      if (queue_overflow) totals_.queue_overflow += report[0];
      for (uint32_t k = 0; k < kMaxJobs; ++k) {
        if (slot_job_[k] == 0) continue;
        Job& j = *jobs_[slot_job_[k]];
        Chunk& c = *chunk_in_slot(j, (int)k);
        // A traversal stack overflow, or a shading step that had to drop
        // paths or shadow rays for want of queue space, fails every resident
        // chunk's job (the dropped work is not attributed to a job),
        // including chunks of the step queued behind this one.
        if (stack_overflow || queue_overflow) {
          j.status = OLHIP_JOB_FAILED;
          j.error = stack_overflow ? "traversal stack overflow" : "ray queue overflow: shading dropped paths";
          drop_chunks(j);
          continue;
        }
        if (c.admit_seq > f.seq) continue;
        const uint32_t rays = rep[4 * k], shadow = rep[4 * k + 1], live = rep[4 * k + 2];
        if (c.steps_in_chunk == 0) j.stats.primary_rays += rays;
        else j.stats.extension_rays += rays;
        j.stats.shadow_rays += shadow;
        totals_.rays += (uint64_t)rays + shadow;
        ++c.steps_in_chunk;
        ++j.stats.steps;
        if (live == 0) {
          // Chunk finished: its device slot is free now; its slot ranges are
          // released when its accumulation is queued, in sample order.
          Chunk done = c;
          done.slot = -1;
          j.resident.erase(j.resident.begin() + (&c - j.resident.data()));
          slot_job_[k] = 0;
          if (j.finished.empty()) acc_jobs_.push_back(j.id);
          j.finished.push_back(done);
          if (!j.stagger && j.next_sample < j.desc.samples) queue_.push_front(j.id);
        }
      }
      if (stack_overflow || queue_overflow) {
        reset_overflow = true;
        overflow_checked_from_ = seq_;
      }
    }
  }
  if (reset_overflow) {
    // Every resident job has failed.  Once the steps in flight have run, the
    // counter block returns to its initial state: the failed jobs' queue and
    // group counts (which a job admitted later into the same slot would
    // otherwise read as its own) and both flags.
    OL_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    wave_clear(wave_, ctx_->stream);
    OL_HIP_CHECK(hipGetLastError());
    OL_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
  }
}

void Engine::run_until_idle() {
  if (worker_running_) throw OlError{OLHIP_ERR_RUNTIME, "the engine's worker thread is running"};
  while (step() > 0) {
  }
}

int Engine::state(uint64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = jobs_.find(id);
  return it == jobs_.end() ? OLHIP_JOB_UNKNOWN : it->second->status;
}

int Engine::wait(uint64_t id, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto done = [&] {
    auto it = jobs_.find(id);
    return it == jobs_.end() || it->second->status == OLHIP_JOB_DONE || it->second->status == OLHIP_JOB_FAILED;
  };
  if (timeout_s < 0) cv_.wait(lk, done);
  else cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  auto it = jobs_.find(id);
  return it == jobs_.end() ? OLHIP_JOB_UNKNOWN : it->second->status;
}

void Engine::swap_image(uint64_t id, DevBuf& buf) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end() || it->second->status != OLHIP_JOB_DONE) throw OlError{OLHIP_ERR_RUNTIME, "job is not done"};
  it->second->image.swap(buf);
}

const float4* Engine::device_image(uint64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end() || it->second->status != OLHIP_JOB_DONE) return nullptr;
  return it->second->image.as<float4>();
}

void Engine::take_result(uint64_t id, float* out_rgba, olhip_job_stats* stats, bool keep) {
  std::unique_ptr<Job> j;
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) throw OlError{OLHIP_ERR_VALUE, "unknown job id"};
    if (it->second->status == OLHIP_JOB_FAILED) {
      std::string e = it->second->error;
      jobs_.erase(it);
      throw OlError{OLHIP_ERR_RUNTIME, "job failed: " + e};
    }
    if (it->second->status != OLHIP_JOB_DONE) throw OlError{OLHIP_ERR_RUNTIME, "job is not done"};
    if (stats) *stats = it->second->stats;
    if (keep) {
      if (out_rgba) {
        OL_HIP_CHECK(hipSetDevice(ctx_->device));
        OL_HIP_CHECK(hipMemcpy(out_rgba, it->second->image.ptr,
                               sizeof(float4) * it->second->desc.width * it->second->desc.height,
                               hipMemcpyDeviceToHost));
      }
      return;
    }
    j = std::move(it->second);
    jobs_.erase(it);
  }
  OL_HIP_CHECK(hipSetDevice(ctx_->device));
  if (out_rgba)
    OL_HIP_CHECK(hipMemcpy(out_rgba, j->image.ptr, sizeof(float4) * j->desc.width * j->desc.height,
                           hipMemcpyDeviceToHost));
  std::lock_guard<std::mutex> g(mu_);
  keep_buffers(*j);
}

void Engine::products(uint64_t id, float* depth, float* normal, uint32_t* primitive, float* variance) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = jobs_.find(id);
  if (it == jobs_.end()) throw OlError{OLHIP_ERR_VALUE, "unknown job id"};
  const Job& j = *it->second;
  if (j.status != OLHIP_JOB_DONE) throw OlError{OLHIP_ERR_RUNTIME, "job is not done"};
  const size_t np = (size_t)j.desc.width * j.desc.height;
  OL_HIP_CHECK(hipSetDevice(ctx_->device));
  if (depth || normal) {
    if (!(j.desc.products & OLHIP_PRODUCT_PRIMARY)) throw OlError{OLHIP_ERR_VALUE, "primary-hit products not requested"};
    std::vector<float4> g4(np);
    OL_HIP_CHECK(hipMemcpy(g4.data(), j.geo.ptr, sizeof(float4) * np, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; ++i) {
      if (depth) depth[i] = g4[i].x;
      if (normal) {
        normal[3 * i] = g4[i].y;
        normal[3 * i + 1] = g4[i].z;
        normal[3 * i + 2] = g4[i].w;
      }
    }
  }
  if (primitive) {
    if (!(j.desc.products & OLHIP_PRODUCT_PRIMARY)) throw OlError{OLHIP_ERR_VALUE, "primary-hit products not requested"};
    OL_HIP_CHECK(hipMemcpy(primitive, j.prim.ptr, sizeof(uint32_t) * np, hipMemcpyDeviceToHost));
  }
  if (variance) {
    if (!(j.desc.products & OLHIP_PRODUCT_VARIANCE)) throw OlError{OLHIP_ERR_VALUE, "variance not requested"};
    OL_HIP_CHECK(hipMemcpy(variance, j.variance.ptr, sizeof(float) * np, hipMemcpyDeviceToHost));
  }
}

void Engine::engine_stats(olhip_engine_stats* out) {
  std::lock_guard<std::mutex> g(mu_);
  *out = totals_;
}

void Engine::start_worker() {
  if (worker_running_) return;
  stop_ = false;
  worker_running_ = true;
  worker_ = std::thread([this] {
    (void)hipSetDevice(ctx_->device);
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] {
          if (stop_) return true;
          if (!queue_.empty() || !acc_jobs_.empty() || inflight_count_ > 0) return true;
          for (uint32_t k = 0; k < kMaxJobs; ++k)
            if (slot_job_[k]) return true;
          return false;
        });
        if (stop_) break;
      }
      try {
        step();
      } catch (const OlError& e) {
        std::lock_guard<std::mutex> g(mu_);
        for (auto& kv : jobs_)
          if (kv.second->status == OLHIP_JOB_QUEUED || kv.second->status == OLHIP_JOB_RUNNING) {
            kv.second->status = OLHIP_JOB_FAILED;
            kv.second->error = e.message;
          }
        queue_.clear();
        acc_jobs_.clear();
        std::fill(slot_job_.begin(), slot_job_.end(), 0);
        cv_.notify_all();
        break;
      }
    }
  });
}

void Engine::stop_worker() {
  if (!worker_running_) return;
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
    cv_.notify_all();
  }
  worker_.join();
  worker_running_ = false;
}
