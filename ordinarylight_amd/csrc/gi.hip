// Batched GI wavefront kernels: per step, accumulate finished chunks ->
// generate admitted jobs -> closest hit -> shade -> shadow rays -> report.
//
// The estimator is the reference's production split-path shade stage
// (shaders/wavefront_shade_candidate.glsl main :3106-3493) for built-in PBR
// materials without textures, custom programs or volumes; the device
// functions it uses are in shade.h.
//
// MI355X structure:
// * All resident jobs share the path slots and queues; every kernel serves
//   every job, so small jobs fill the GPU together (continuous batching).
// * Queues are cut into (job, segment) runs with one counter each; a plan
//   kernel turns the run counts into a prefix of 64-entry chunks, and a
//   wave takes whole chunks, so its job is uniform and the job's parameters
//   are scalar loads.
// * The next-event-estimation shadow rays are deferred to their own any-hit
//   kernel; each hit's visible contributions are summed in the reference's
//   order and added before the next bounce, so each path's radiance is the
//   same sequence of float operations as the inline form.
// * The traversal kernels keep their grid resident: waves take chunks from
//   striped cursors, refill lanes whose ray has finished, and choose per step
//   between advancing lanes at inner nodes and lanes at leaves.
// No host synchronisation happens inside a step; the host reads the per-job
// report after it.

#include <algorithm>

#include "gi.h"
#include "shade.h"
#include "restir.h"
#include "trace.h"

namespace {

using namespace olshade;

// Idle lanes of a wave that trigger a refill from the chunk pool: flat
// traversal (measured best at 16 of 64) and two-level traversal, whose rays
// run longer (8: C3 671 -> 680 Mrays/s, while 8 cost C2 0.1 ms per frame).
constexpr uint32_t kRefillMin = 16;
constexpr uint32_t kRefillMin2 = 8;

// ---- chunk plan ----------------------------------------------------------------------
// Up to four plans in one launch, one block each: block b computes the
// exclusive prefix of ceil(count / kChunk) over the runs of counter block
// counts[b] whose jobs' scenes are flat (kind 0) or two-level (kind 1), into
// out[b] (out[b][kRuns] = total), and the plan's bucket index (P_SHIFT ..
// P_INDEX).  Offsets are counter-block words.
struct PlanSet {
  uint32_t counts[4], out[4], kind[4];
};

__global__ void __launch_bounds__(1024) k_plan(uint32_t* __restrict__ ctr, const JobDev* __restrict__ jobs, PlanSet ps) {
  static_assert(kRuns == 2 * 1024 && kJobStripes % 2 == 0, "two runs of one job per thread");
  __shared__ uint32_t wave_tot[16], wave_last[16], shift_s;
  const uint32_t t = threadIdx.x;
  const uint32_t* __restrict__ counts = ctr + ps.counts[blockIdx.x];
  uint32_t* __restrict__ out = ctr + ps.out[blockIdx.x];
  const uint32_t kind = ps.kind[blockIdx.x];
  const bool mine = (jobs[(2 * t) / kJobStripes].P.sc.instanced != 0u) == (kind != 0u);
  const uint32_t a = mine ? (counts[run_word(2 * t)] + kChunk - 1u) / kChunk : 0u;
  const uint32_t b = mine ? (counts[run_word(2 * t + 1)] + kChunk - 1u) / kChunk : 0u;
  const uint32_t sum = a + b;
  uint32_t incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(incl, (unsigned)off, 64);
    if ((int)lane_id() >= off) incl += o;
  }
  // The last run holding chunks bounds the consumers' run search.
  uint32_t last = b ? 2u * t + 1u : (a ? 2u * t : 0u);
  for (int off = 32; off; off >>= 1) last = max(last, (uint32_t)__shfl_xor((int)last, off, 64));
  const uint32_t wave = t / 64u;
  if (lane_id() == 63u) wave_tot[wave] = incl;
  if (lane_id() == 0u) wave_last[wave] = last;
  __syncthreads();
  uint32_t base = 0, total = 0;
  for (uint32_t w = 0; w < 16u; ++w) {
    if (w < wave) base += wave_tot[w];
    total += wave_tot[w];
  }
  const uint32_t excl = base + incl - sum;
  out[2 * t] = excl;
  out[2 * t + 1] = excl + a;
  if (t == 0) {
    uint32_t shift = 0;
    while (((total + (1u << shift) - 1u) >> shift) > kPlanBuckets) ++shift;
    uint32_t top = 0;
    for (uint32_t w = 0; w < 16u; ++w) top = max(top, wave_last[w]);
    out[kRuns] = total;
    out[P_SHIFT] = shift;
    out[P_BUCKETS] = (total + (1u << shift) - 1u) >> shift;
    out[P_LAST] = top;
    shift_s = shift;
  }
  __syncthreads();
  // Bucket k starts at chunk k << shift; the run whose chunks hold that
  // start names it (runs without chunks name none).
  const uint32_t shift = shift_s;
  const uint32_t firsts[2] = {excl, excl + a}, ns[2] = {a, b};
  for (int h = 0; h < 2; ++h) {
    if (ns[h] == 0u) continue;
    const uint32_t p = firsts[h];
    for (uint32_t k = (p + (1u << shift) - 1u) >> shift; k <= (p + ns[h] - 1u) >> shift; ++k) out[P_INDEX + k] = 2u * t + (uint32_t)h;
  }
}

// A plan's bucket parameters, read once per kernel.
struct PlanHead {
  uint32_t shift, buckets, last;
};
OL_DEV PlanHead plan_head(const uint32_t* __restrict__ prefix) {
  return PlanHead{prefix[P_SHIFT], prefix[P_BUCKETS], prefix[P_LAST]};
}

// Run holding chunk c: the largest r with prefix[r] <= c.  It lies between
// the run of c's bucket and the run of the next bucket (or the last run
// holding chunks): a bisection over that range, usually a step or two.
OL_DEV uint32_t find_run(const uint32_t* __restrict__ prefix, const PlanHead& h, uint32_t c) {
  const uint32_t k = c >> h.shift;
  uint32_t lo = prefix[P_INDEX + k];
  uint32_t hi = k + 1u < h.buckets ? prefix[P_INDEX + k + 1u] : h.last;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (prefix[mid] <= c) lo = mid;
    else hi = mid - 1u;
  }
  return lo;
}

// A chunk located in its run: job slot, segment and the first queue slot.
struct ChunkLoc {
  uint32_t job, stripe, first, count;  // count: live entries in this chunk (1..64)
  uint32_t chunk_in_job;               // chunk index counted from the job's first run
};

OL_DEV ChunkLoc locate_chunk(const uint32_t* __restrict__ prefix, const PlanHead& top, const uint32_t* __restrict__ counts,
                             const JobDev* __restrict__ jobs, uint32_t c) {
  ChunkLoc k;
  const uint32_t r = find_run(prefix, top, c);
  k.job = r / kJobStripes;
  k.stripe = r % kJobStripes;
  const uint32_t off = (c - prefix[r]) * kChunk;
  k.count = min(kChunk, counts[run_word(r)] - off);
  k.first = jobs[k.job].base + k.stripe * jobs[k.job].seg + off;
  k.chunk_in_job = c - prefix[k.job * kJobStripes];
  return k;
}

OL_DEV uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

OL_DEV uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---- accumulate + generate (host lists) ------------------------------------------------------
// Index of the list item owning global thread index g (prefix[n] = total):
// the last k < n with prefix[k] <= g.  The wave's first active lane (the
// lowest g) binary-searches with wave-uniform loads; every lane then steps
// forward from that item, which already holds it unless the wave straddles
// an item boundary.
OL_DEV uint32_t list_item(const uint32_t* prefix, uint32_t n, uint32_t g) {
  const uint32_t g0 = uniform(g);
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= g0) lo = mid;
    else hi = mid;
  }
  uint32_t k = lo;
  while (k + 1u < n && prefix[k + 1u] <= g) ++k;
  return k;
}

// Adds the float32 radiance of a finished chunk's samples to its job's
// float64 pixel sums in sample order, then writes the image after the last
// chunk (render_wavefront_outputs, api.py:1404-1422).
__global__ void __launch_bounds__(256) k_accumulate(Wave W) {
  const StepList& Ls = *W.list;
  const uint32_t n = Ls.n_acc;
  if (n == 0) return;
  const uint32_t total = Ls.acc_prefix[n];
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const uint32_t k = list_item(Ls.acc_prefix, n, g);
    const AccItem& it = Ls.acc[k];
    const uint32_t pixel = g - Ls.acc_prefix[k];
    double4 a = it.first_chunk ? make_double4(0.0, 0.0, 0.0, 0.0) : it.acc[pixel];
    for (uint32_t s = 0; s < it.spp; ++s) {
      const float4 l = W.L[it.base + s * it.np + pixel];
      if (it.first_chunk && s == 0) {
        a = make_double4((double)l.x, (double)l.y, (double)l.z, 1.0);
      } else {
        a.x += (double)l.x;
        a.y += (double)l.y;
        a.z += (double)l.z;
        a.w += 1.0;
      }
    }
    // The running sums are read only by the job's later chunks.
    const bool only_chunk = it.first_chunk && it.final_samples;
    if (!only_chunk) it.acc[pixel] = a;
    if (it.final_samples) {
      const double ns = (double)it.final_samples;
      it.image[pixel] = make_float4((float)(a.x / ns), (float)(a.y / ns), (float)(a.z / ns), (float)(a.w / ns));
    }
    if (it.lum) {
      // Luminance of each float32 sample (float32 products and sums, as the
      // reference's NumPy expression), summed with its square in float64.
      double2 m = it.first_chunk ? make_double2(0.0, 0.0) : it.lum[pixel];
      for (uint32_t s = 0; s < it.spp; ++s) {
        const float4 l = W.L[it.base + s * it.np + pixel];
        const double y = (double)((l.x * 0.2126f + l.y * 0.7152f) + l.z * 0.0722f);
        if (it.first_chunk && s == 0) {
          m = make_double2(y, y * y);
        } else {
          m.x += y;
          m.y += y * y;
        }
      }
      if (!only_chunk) it.lum[pixel] = m;
      if (it.final_samples) {
        const double n = (double)it.final_samples;
        double v = 0.0;
        if (it.final_samples > 1) v = fmax((m.y - m.x * m.x / n) / (n - 1.0), 0.0);
        it.variance[pixel] = (float)v;
      }
    }
  }
}

// Initial random state of the camera path of (pixel, sample).
OL_DEV uint32_t primary_seed(uint32_t frame, uint32_t pixel, uint32_t sample) {
  return wave_hash((pixel ^ wave_hash(frame)) ^ wave_hash(sample + 1u));
}

// Camera ray of one (pixel, sample) with the hash jitter of
// wavefront_generate.comp:115-149; rng is the path's initial state.
OL_DEV void primary_ray(const GiParams& P, uint32_t px, uint32_t py, uint32_t sample, f3& org, f3& dir,
                        uint32_t& rng) {
  const uint32_t Wd = P.width, H = P.height;
  const uint32_t pixel = py * Wd + px;
  rng = primary_seed(P.frame, pixel, sample);
  rng = wave_hash(rng);
  const float jx = u32_to_unit(rng);
  rng = wave_hash(rng);
  const float jy = u32_to_unit(rng);
  const float ndx = (((float)px + jx) / (float)Wd) * 2.0f - 1.0f;
  const float ndy = (((float)py + jy) / (float)H) * 2.0f - 1.0f;
  const float aspect = (float)Wd / (float)H;
  org = mk3(P.cam_o[0], P.cam_o[1], P.cam_o[2]);
  const f3 fw = mk3(P.cam_f[0], P.cam_f[1], P.cam_f[2]);
  const f3 rt = mk3(P.cam_r[0], P.cam_r[1], P.cam_r[2]);
  const f3 up = mk3(P.cam_u[0], P.cam_u[1], P.cam_u[2]);
  if (P.projection == 1u) {
    org = (org + ((ndx * aspect) * rt)) - (ndy * up);
    dir = normalize3(fw);
  } else if (P.projection == 2u) {
    float yaw = ndx * length3(rt);
    float pitch = (-ndy) * length3(up);
    float sy, cy, sp, cp;
    ol_sincos(yaw, &sy, &cy);
    ol_sincos(pitch, &sp, &cp);
    dir = normalize3((((normalize3(fw) * cp) * cy) + ((normalize3(rt) * cp) * sy)) + (normalize3(up) * sp));
  } else {
    dir = normalize3((fw + ((ndx * aspect) * rt)) - (ndy * up));
  }
}

// Path state (M) of a camera path before its first shading: pixel,
// frame << 8 | sample, the random state primary_ray leaves after its two
// jitter draws, active at medium depth 1.  Generated slots store no copy
// (like their unit throughput and zero radiance): the shading kernels rebuild
// it from the path's offset in its chunk, local sample * pixels + pixel.
OL_DEV uint4 fresh_path_state(const GiParams& P, uint32_t rel) {
  const uint32_t np = P.width * P.height;
  const uint32_t local = rel / np, pixel = rel - local * np;
  const uint32_t sample = P.sample_base + local;
  const uint32_t rng = wave_hash(wave_hash(primary_seed(P.frame, pixel, sample)));
  return make_uint4(pixel, (P.frame << 8) | (sample & 255u), rng, F_ACTIVE | (1u << 8));
}

// Primary rays of a pixel tile [x0, x0 + tw) x [y0, y0 + th), row-major
// (picking and tests): origin/tmin, direction/tmax as the queues hold them.
__global__ void __launch_bounds__(256) k_camera_rays(GiParams P, uint32_t x0, uint32_t y0, uint32_t tw, uint32_t th,
                                                     uint32_t sample, float4* o, float4* d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tw * th) return;
  const uint32_t ly = i / tw, lx = i - ly * tw;
  f3 org, dir;
  uint32_t rng;
  primary_ray(P, x0 + lx, y0 + ly, sample, org, dir, rng);
  o[i] = mk4(org, 0.001f);
  d[i] = mk4(dir, 1e30f);
}

// The camera ray of a job's i-th generated queue slot: slots follow 8x8
// pixel tiles (sample-major), so a wave's rays are coherent.
OL_DEV void generated_ray(const JobDev& J, uint32_t i, f3& org, f3& dir, uint32_t& path) {
  const GiParams& P = J.P;
  const uint32_t Wd = P.width, H = P.height, np = Wd * H;
  // i -> (sample, 8x8 tile, position in tile) in tile order.
  const uint32_t local = i / np, rem = i - local * np;
  const uint32_t ty = rem / (8u * Wd);            // tile row
  const uint32_t th = min(8u, H - 8u * ty);       // rows in this tile row
  const uint32_t in_row = rem - ty * 8u * Wd;     // index inside the tile row
  const uint32_t tx = in_row / (8u * th);         // full-height tiles to the left are 8 wide
  const uint32_t tw = min(8u, Wd - 8u * tx);
  const uint32_t in_tile = in_row - tx * 8u * th;
  const uint32_t ly = in_tile / tw, lx = in_tile - ly * tw;
  const uint32_t px = tx * 8u + lx, py = ty * 8u + ly;
  const uint32_t sample = P.sample_base + local;
  path = J.base + local * np + py * Wd + px;
  uint32_t rng;
  primary_ray(P, px, py, sample, org, dir, rng);
}

// Camera rays and path initialisation of the listed jobs
// (wavefront_generate.comp:86-167): the job's queue segments are filled in
// order (slot base + i for the i-th ray in tile order).  With `rays` 0 only
// the queue counts and medium stacks are written: the hybrid primary kernel
// computes each slot's ray itself.
__global__ void __launch_bounds__(256) k_generate(Wave W, uint32_t in, uint32_t rays) {
  const StepList& Ls = *W.list;
  const uint32_t n = Ls.n_gen;
  if (n == 0) return;
  const uint32_t total = Ls.gen_prefix[n];
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const uint32_t k = list_item(Ls.gen_prefix, n, g);
    const uint32_t job = Ls.gen_job[k];
    const JobDev& J = W.jobs[job];
    const uint32_t i = g - Ls.gen_prefix[k];
    if (i == 0) {
      for (uint32_t s = 0; s < J.stripes; ++s) {
        const uint32_t lo = s * J.seg;
        W.ctr[c_queue(in) + run_word(job * kJobStripes + s)] = J.n_paths > lo ? min(J.seg, J.n_paths - lo) : 0u;
      }
    }
    // Without rays, the medium stacks of the same set of paths, in slot order.
    uint32_t path = J.base + i;
    if (rays) {
      f3 org, dir;
      generated_ray(J, i, org, dir, path);
      const uint32_t slot = J.base + i;
      W.qo[in][slot] = mk4(org, 0.001f);
      W.qd[in][slot] = mk4(dir, 1e30f);
      W.qp[in][slot] = path;
    }
    // T = (1, 1, 1, bounce 0), L = 0 and M = fresh_path_state are implied
    // for a fresh slot (StepList::fresh) and first written by shade.
    if (W.stk) W.stk[4u * path] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ---- traversal ------------------------------------------------------------------------------
template <int LSTACK, int BLOCK>
OL_DEV TravStack<LSTACK> lane_stack(uint32_t* lds, uint32_t* spill, uint32_t spill_cap, uint32_t* overflow) {
  TravStack<LSTACK> st;
  st.lds = (lds_word*)lds + threadIdx.x;
  st.lds_stride = BLOCK;
  st.spill = spill ? (global_word*)spill + (blockIdx.x * BLOCK + threadIdx.x) : nullptr;
  st.spill_stride = gridDim.x * BLOCK;
  st.spill_cap = spill ? spill_cap : 0u;
  st.overflow = overflow;
  st.sp = 0;
  return st;
}

// Wave-uniform pool of rays not yet started: `n` entries from queue slot
// `slot` of one chunk, all in one job.
struct Pool {
  uint32_t slot, n;
  const WideNode* nodes;
  const float4* tri;
  uint32_t node_count;
  bool exhausted;
};

// Takes the next chunk of this block's work stripe when the pool is empty.
OL_DEV void pool_refill(Pool& p, const Wave& W, const uint32_t* prefix, const PlanHead& top, const uint32_t* counts,
                        uint32_t* cursor, uint32_t c_begin, uint32_t c_end) {
  if (p.n != 0 || p.exhausted) return;
  uint32_t c = 0;
  if (lane_id() == 0) c = atomicAdd(cursor, 1u);
  c = uniform(__shfl(c, 0, 64));
  if (c >= c_end - c_begin) {
    p.exhausted = true;
    return;
  }
  const ChunkLoc k = locate_chunk(prefix, top, counts, W.jobs, c_begin + c);
  const DScene& sc = W.jobs[uniform(k.job)].P.sc;
  p.slot = uniform(k.first);
  p.n = uniform(k.count);
  p.nodes = sc.nodes;
  p.tri = sc.tri;
  p.node_count = sc.node_count;
}

// Closest hit for the input ray queue: refilled lanes, split inner / leaf steps.
template <int LSTACK>
OL_DEV void closest_pass(const Wave& W, uint32_t in, TravStack<LSTACK>& st) {
  const uint32_t* prefix = W.ctr + C_RP;
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t chunks = prefix[kRuns];
  if (chunks == 0) return;
  const PlanHead top = plan_head(prefix);
  const uint32_t ws = blockIdx.x % kWorkStripes;
  const uint32_t c_begin = (uint32_t)((uint64_t)chunks * ws / kWorkStripes);
  const uint32_t c_end = (uint32_t)((uint64_t)chunks * (ws + 1u) / kWorkStripes);
  uint32_t* cursor = &W.ctr[C_W + ws * kCursorStride];
  const float4* __restrict__ qo = W.qo[in];
  const float4* __restrict__ qd = W.qd[in];
  Pool pool{0u, 0u, nullptr, nullptr, 0u, false};
  ClosestRay q;
  q.node = OL_NONE;
  const WideNode* nodes = nullptr;
  const float4* tri = nullptr;
  uint32_t slot = OL_NONE;
  bool busy = false;
  uint32_t nv = 0, nt = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin) {
      pool_refill(pool, W, prefix, top, counts, cursor, c_begin, c_end);
      const uint32_t take = min(n_idle, pool.n);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          slot = pool.slot + rank;
          const float4 o = qo[slot], d = qd[slot];
          closest_begin(q, mk3(o), mk3(d), o.w, d.w);
          if (pool.node_count == 0) {
            W.hit[slot] = make_float4(d.w, 0.0f, 0.0f, __uint_as_float(OL_NONE));
          } else {
            nodes = pool.nodes;
            tri = pool.tri;
            st.sp = 0;
            busy = true;
          }
        }
      }
      pool.slot += take;
      pool.n -= take;
    }
    const bool at_leaf = busy && item_is_leaf(q.node);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t inner = __ballot(busy && !at_leaf);
    if ((leafy | inner) == 0) {
      if (pool.exhausted && pool.n == 0) break;
      continue;
    }
    if (inner == 0 || 3u * (uint32_t)__popcll(leafy) >= (uint32_t)__popcll(leafy | inner)) {
      if (at_leaf) closest_leaf<LSTACK>(tri, q, st, nt);
    } else if (busy && !at_leaf) {
      closest_inner<LSTACK>(nodes, q, st, nv);
    }
    if (busy && q.node == OL_NONE) {
      W.hit[slot] = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.slot));
      busy = false;
    }
  }
  if (W.work && (nv | nt)) {
    atomicAdd(&W.work[0], (unsigned long long)nv);
    atomicAdd(&W.work[1], (unsigned long long)nt);
  }
}

// Adds one visible shadow entry's contribution.
OL_DEV void shadow_add(uint32_t mode, float4 c, f3& direct, f3& area_direct, f3& env_direct) {
  const uint32_t kind = (uint32_t)c.w;
  if (mode == 0u || kind == 0u) direct = direct + rgb(c);
  else if (kind == 1u) area_direct = area_direct + rgb(c);
  else env_direct = env_direct + rgb(c);
}

// Per-domain averages, then L += T * direct / p_nee.
OL_DEV void shadow_resolve(const Wave& W, uint32_t g, f3 direct, f3 area_direct, f3 env_direct) {
  const float4 gT = W.gT[g];
  const uint4 gI = W.gI[g];
  const uint32_t path = gI.x, area_n = gI.w & 255u, env_n = (gI.w >> 8) & 255u, mode = gI.w >> 16;
  if (mode != 0u) {
    if (area_n > 0u) direct = direct + area_direct / (float)area_n;
    if (env_n > 0u) direct = direct + env_direct / (float)env_n;
  }
  float4 L = W.L[path];
  f3 add = (rgb(gT) * direct) / fmaxf(gT.w, 1e-6f);
  W.L[path] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, L.w);
}

// Shadow groups the previous step's shading emitted (parity in ^ 1): a lane
// owns one group and traces its entries in order (any hit), summing the
// visible contributions as the inline form does.
template <int LSTACK>
OL_DEV void shadow_pass(const Wave& W, uint32_t in, TravStack<LSTACK>& st) {
  const uint32_t* prefix = W.ctr + C_GP;
  const uint32_t* counts = W.ctr + c_groups(in ^ 1u);
  const uint32_t chunks = prefix[kRuns];
  if (chunks == 0) return;
  const PlanHead top = plan_head(prefix);
  const uint32_t ws = blockIdx.x % kWorkStripes;
  const uint32_t c_begin = (uint32_t)((uint64_t)chunks * ws / kWorkStripes);
  const uint32_t c_end = (uint32_t)((uint64_t)chunks * (ws + 1u) / kWorkStripes);
  uint32_t* cursor = &W.ctr[C_WS + ws * kCursorStride];
  Pool pool{0u, 0u, nullptr, nullptr, 0u, false};
  AnyRay q;
  q.node = OL_NONE;
  q.occluded = false;
  const WideNode* nodes = nullptr;
  const float4* tri = nullptr;
  uint32_t g = 0, e = 0, e_end = 0, mode = 0;
  f3 direct = mk3(0.0f, 0.0f, 0.0f), area_direct = direct, env_direct = direct;
  bool busy = false;
  uint32_t nv = 0, nt = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin) {
      pool_refill(pool, W, prefix, top, counts, cursor, c_begin, c_end);
      const uint32_t take = min(n_idle, pool.n);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          g = pool.slot + rank;
          const uint4 gI = W.gI[g];
          e = gI.y;
          e_end = gI.y + gI.z;
          mode = gI.w >> 16;
          direct = mk3(0.0f, 0.0f, 0.0f);
          area_direct = direct;
          env_direct = direct;
          if (pool.node_count == 0) {
            // Nothing can occlude: every entry is visible.
            for (; e < e_end; ++e) shadow_add(mode, W.ec[e], direct, area_direct, env_direct);
            shadow_resolve(W, g, direct, area_direct, env_direct);
          } else if (e < e_end) {
            const float4 o = W.eo[e], d = W.ed[e];
            any_begin(q, mk3(o), mk3(d), o.w, d.w);
            nodes = pool.nodes;
            tri = pool.tri;
            st.sp = 0;
            busy = true;
          } else {
            shadow_resolve(W, g, direct, area_direct, env_direct);
          }
        }
      }
      pool.slot += take;
      pool.n -= take;
    }
    const bool at_leaf = busy && item_is_leaf(q.node);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t inner = __ballot(busy && !at_leaf);
    if ((leafy | inner) == 0) {
      if (pool.exhausted && pool.n == 0) break;
      continue;
    }
    if (inner == 0 || 3u * (uint32_t)__popcll(leafy) >= (uint32_t)__popcll(leafy | inner)) {
      if (at_leaf) any_leaf<LSTACK>(tri, q, st, nt);
    } else if (busy && !at_leaf) {
      any_inner<LSTACK>(nodes, q, st, nv);
    }
    if (busy && q.node == OL_NONE) {
      if (!q.occluded) shadow_add(mode, W.ec[e], direct, area_direct, env_direct);
      ++e;
      if (e < e_end) {
        const float4 o = W.eo[e], d = W.ed[e];
        any_begin(q, mk3(o), mk3(d), o.w, d.w);
        st.sp = 0;
      } else {
        shadow_resolve(W, g, direct, area_direct, env_direct);
        busy = false;
      }
    }
  }
  if (W.work && (nv | nt)) {
    atomicAdd(&W.work[2], (unsigned long long)nv);
    atomicAdd(&W.work[3], (unsigned long long)nt);
  }
}

// Two-level pool: rays of one chunk, all of one job (hence one scene view).
struct Pool2 {
  uint32_t slot, n, job;
  bool exhausted;
};

OL_DEV void pool2_refill(Pool2& p, const Wave& W, const uint32_t* prefix, const PlanHead& top, const uint32_t* counts,
                         uint32_t* cursor, uint32_t c_begin, uint32_t c_end) {
  if (p.n != 0 || p.exhausted) return;
  uint32_t c = 0;
  if (lane_id() == 0) c = atomicAdd(cursor, 1u);
  c = uniform(__shfl(c, 0, 64));
  if (c >= c_end - c_begin) {
    p.exhausted = true;
    return;
  }
  const ChunkLoc k = locate_chunk(prefix, top, counts, W.jobs, c_begin + c);
  p.slot = uniform(k.first);
  p.n = uniform(k.count);
  p.job = uniform(k.job);
}

// Closest hit in two-level scenes (top-level BVH over instances): the
// scheduling of closest_pass with the instance-entering inner step.
template <int LSTACK>
OL_DEV void closest2_pass(const Wave& W, uint32_t in, TravStack<LSTACK>& st) {
  const uint32_t* prefix = W.ctr + C_RP2;
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t chunks = prefix[kRuns];
  if (chunks == 0) return;
  const PlanHead top = plan_head(prefix);
  const uint32_t ws = blockIdx.x % kWorkStripes;
  const uint32_t c_begin = (uint32_t)((uint64_t)chunks * ws / kWorkStripes);
  const uint32_t c_end = (uint32_t)((uint64_t)chunks * (ws + 1u) / kWorkStripes);
  uint32_t* cursor = &W.ctr[C_W2 + ws * kCursorStride];
  const float4* __restrict__ qo = W.qo[in];
  const float4* __restrict__ qd = W.qd[in];
  Pool2 pool{0u, 0u, 0u, false};
  Closest2 q;
  q.node = OL_NONE;
  q.inst = OL_NONE;
  uint32_t job = 0, slot = OL_NONE;
  bool busy = false;
  uint32_t nv = 0, nt = 0, ni = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin2) {
      pool2_refill(pool, W, prefix, top, counts, cursor, c_begin, c_end);
      const uint32_t take = min(n_idle, pool.n);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          slot = pool.slot + rank;
          job = pool.job;
          const float4 o = qo[slot], d = qd[slot];
          closest2_begin(q, mk3(o), mk3(d), o.w, d.w);
          if (W.jobs[job].P.sc.node_count == 0) {
            W.hit[slot] = make_float4(d.w, 0.0f, 0.0f, __uint_as_float(OL_NONE));
          } else {
            two_level_start(st);
            busy = true;
          }
        }
      }
      pool.slot += take;
      pool.n -= take;
    }
    const bool at_leaf = busy && closest2_at_leaf(q);
    const bool at_entry = busy && closest2_at_entry(q);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t entries = __ballot(at_entry);
    const uint64_t inner = __ballot(busy && !at_leaf && !at_entry);
    const uint32_t n_busy = (uint32_t)__popcll(leafy | entries | inner);
    if (n_busy == 0) {
      if (pool.exhausted && pool.n == 0) break;
      continue;
    }
    // Leaf, entry or node step: a class runs when it holds a quarter of the
    // busy lanes or nothing else is left (leaves first, then entries).  A
    // quarter measured faster than a third or a fifth for both classes on
    // the instanced sphere field (the flat kernels keep a third).
    if (leafy && ((entries | inner) == 0 || 4u * (uint32_t)__popcll(leafy) >= n_busy)) {
      if (at_leaf) closest2_leaf<LSTACK>(q, st, nt);
    } else if (entries && (inner == 0 || 4u * (uint32_t)__popcll(entries) >= n_busy)) {
      if (at_entry) closest2_enter<LSTACK>(W.jobs[job].P.sc, q, st, ni);
    } else if (inner) {
      if (busy && !at_leaf && !at_entry) closest2_node<LSTACK>(W.jobs[job].P.sc, q, st, nv);
    } else {
      if (at_leaf) closest2_leaf<LSTACK>(q, st, nt);
    }
    if (busy && q.node == OL_NONE) {
      W.hit[slot] = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.slot));
      W.hit_inst[slot] = q.best_inst;
      busy = false;
    }
  }
  if (W.work && (nv | nt | ni)) {
    atomicAdd(&W.work[0], (unsigned long long)nv);
    atomicAdd(&W.work[1], (unsigned long long)nt);
    atomicAdd(&W.work[4], (unsigned long long)ni);
  }
}

// Shadow groups in two-level scenes.
template <int LSTACK>
OL_DEV void shadow2_pass(const Wave& W, uint32_t in, TravStack<LSTACK>& st) {
  const uint32_t* prefix = W.ctr + C_GP2;
  const uint32_t* counts = W.ctr + c_groups(in ^ 1u);
  const uint32_t chunks = prefix[kRuns];
  if (chunks == 0) return;
  const PlanHead top = plan_head(prefix);
  const uint32_t ws = blockIdx.x % kWorkStripes;
  const uint32_t c_begin = (uint32_t)((uint64_t)chunks * ws / kWorkStripes);
  const uint32_t c_end = (uint32_t)((uint64_t)chunks * (ws + 1u) / kWorkStripes);
  uint32_t* cursor = &W.ctr[C_WS2 + ws * kCursorStride];
  Pool2 pool{0u, 0u, 0u, false};
  Any2 q;
  q.node = OL_NONE;
  q.inst = OL_NONE;
  q.occluded = false;
  uint32_t job = 0, g = 0, e = 0, e_end = 0, mode = 0;
  f3 direct = mk3(0.0f, 0.0f, 0.0f), area_direct = direct, env_direct = direct;
  bool busy = false;
  uint32_t nv = 0, nt = 0, ni = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin2) {
      pool2_refill(pool, W, prefix, top, counts, cursor, c_begin, c_end);
      const uint32_t take = min(n_idle, pool.n);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          g = pool.slot + rank;
          job = pool.job;
          const uint4 gI = W.gI[g];
          e = gI.y;
          e_end = gI.y + gI.z;
          mode = gI.w >> 16;
          direct = mk3(0.0f, 0.0f, 0.0f);
          area_direct = direct;
          env_direct = direct;
          if (W.jobs[job].P.sc.node_count == 0) {
            for (; e < e_end; ++e) shadow_add(mode, W.ec[e], direct, area_direct, env_direct);
            shadow_resolve(W, g, direct, area_direct, env_direct);
          } else if (e < e_end) {
            const float4 o = W.eo[e], d = W.ed[e];
            any2_begin(q, mk3(o), mk3(d), o.w, d.w);
            two_level_start(st);
            busy = true;
          } else {
            shadow_resolve(W, g, direct, area_direct, env_direct);
          }
        }
      }
      pool.slot += take;
      pool.n -= take;
    }
    const bool at_leaf = busy && any2_at_leaf(q);
    const bool at_entry = busy && any2_at_entry(q);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t entries = __ballot(at_entry);
    const uint64_t inner = __ballot(busy && !at_leaf && !at_entry);
    const uint32_t n_busy = (uint32_t)__popcll(leafy | entries | inner);
    if (n_busy == 0) {
      if (pool.exhausted && pool.n == 0) break;
      continue;
    }
    if (leafy && ((entries | inner) == 0 || 4u * (uint32_t)__popcll(leafy) >= n_busy)) {
      if (at_leaf) any2_leaf<LSTACK>(q, st, nt);
    } else if (entries && (inner == 0 || 4u * (uint32_t)__popcll(entries) >= n_busy)) {
      if (at_entry) any2_enter<LSTACK>(W.jobs[job].P.sc, q, st, ni);
    } else if (inner) {
      if (busy && !at_leaf && !at_entry) any2_node<LSTACK>(W.jobs[job].P.sc, q, st, nv);
    } else {
      if (at_leaf) any2_leaf<LSTACK>(q, st, nt);
    }
    if (busy && q.node == OL_NONE) {
      if (!q.occluded) shadow_add(mode, W.ec[e], direct, area_direct, env_direct);
      ++e;
      if (e < e_end) {
        const float4 o = W.eo[e], d = W.ed[e];
        any2_begin(q, mk3(o), mk3(d), o.w, d.w);
        two_level_start(st);
      } else {
        shadow_resolve(W, g, direct, area_direct, env_direct);
        busy = false;
      }
    }
  }
  if (W.work && (nv | nt | ni)) {
    atomicAdd(&W.work[2], (unsigned long long)nv);
    atomicAdd(&W.work[3], (unsigned long long)nt);
    atomicAdd(&W.work[5], (unsigned long long)ni);
  }
}

// A step's traversal: closest hits of the input ray queue, then the shadow
// rays of the previous step's shading.  Waves move on to shadow rays as the
// closest-hit chunks of their work stripe run out, so each pass's last few
// long rays overlap the other pass's work instead of ending a launch alone.
// Neither pass writes what the other reads: closest hit writes hit records,
// shadow rays add to path radiance, which only the shading that follows
// reads.
// Six waves per SIMD (80 VGPRs): the closest-hit pass alone needs no more,
// and the shadow pass was measured no slower held to it.
template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6)))
k_trace(Wave W, uint32_t in, uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &W.ctr[C_STACK]);
  closest_pass<LSTACK>(W, in, st);
  shadow_pass<LSTACK>(W, in, st);
}

// The same for jobs of two-level scenes.  Five waves per SIMD (96 VGPRs, a
// few registers of the shadow pass spilled) instead of the four that the
// shadow pass's 111 VGPRs allow: dependent node and instance loads need the
// extra wave more than the spilled registers cost (instanced sphere field:
// traversal 73 -> 65.5 ms per frame; six waves spill 104 bytes per lane and
// are slower than four).
template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_trace2(Wave W, uint32_t in, uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &W.ctr[C_STACK]);
  closest2_pass<LSTACK>(W, in, st);
  shadow2_pass<LSTACK>(W, in, st);
}

template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_rays(DScene sc, const float4* o4, const float4* d4, uint32_t n,
                                                      float4* hits, uint32_t* overflow, unsigned long long* work,
                                                      uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, overflow);
  uint32_t nv = 0, nt = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    float4 o = o4[i], d = d4[i];
    TraceHit h = ol_trace_closest<LSTACK>(sc.nodes, sc.tri, sc.node_count, mk3(o), mk3(d), o.w, d.w, st, nv, nt);
    hits[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.slot == OL_NONE ? OL_NONE : h.orig));
  }
  if (work && (nv | nt)) {
    atomicAdd(&work[0], (unsigned long long)nv);
    atomicAdd(&work[1], (unsigned long long)nt);
  }
}

template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_rays2(DScene sc, const float4* o4, const float4* d4, uint32_t n,
                                                       float4* hits, uint32_t* overflow, unsigned long long* work,
                                                       uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, overflow);
  uint32_t nv = 0, nt = 0, ni = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    const float4 o = o4[i], d = d4[i];
    Closest2 q;
    closest2_begin(q, mk3(o), mk3(d), o.w, d.w);
    two_level_start(st);
    if (sc.node_count == 0) q.node = OL_NONE;
    while (q.node != OL_NONE) {
      if (closest2_at_leaf(q)) closest2_leaf<LSTACK>(q, st, nt);
      else closest2_inner<LSTACK>(sc, q, st, nv, ni);
    }
    hits[i] = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.orig));
  }
  if (work && (nv | nt)) {
    atomicAdd(&work[0], (unsigned long long)nv);
    atomicAdd(&work[1], (unsigned long long)nt);
  }
}

// ---- ReSTIR DI at primary hits -----------------------------------------------------------------
// Pixel of a world point through the job's camera (reprojectRestir,
// wavefront_primary_impl.glsl:350-365, with the current camera as the
// previous one: passes of one job share it).
OL_DEV bool restir_reproject(const GiParams& P, f3 p, int& px, int& py) {
  const f3 off = p - mk3(P.cam_o[0], P.cam_o[1], P.cam_o[2]);
  const f3 fw = mk3(P.cam_f[0], P.cam_f[1], P.cam_f[2]);
  const f3 rt = mk3(P.cam_r[0], P.cam_r[1], P.cam_r[2]);
  const f3 up = mk3(P.cam_u[0], P.cam_u[1], P.cam_u[2]);
  const float depth = dot3(off, fw);
  const float scale = length3(up);
  const float aspect = (float)P.width / (float)P.height;
  if (depth <= 0.0001f || scale <= 0.0001f) return false;
  const float ps = P.projection == 1u ? 1.0f : depth;
  const float nx = dot3(off, normalize3(rt)) / ((ps * aspect) * scale);
  const float ny = (-dot3(off, normalize3(up))) / (ps * scale);
  const float fx = rintf(((nx * 0.5f + 0.5f) * (float)P.width) - 0.5f);
  const float fy = rintf(((ny * 0.5f + 0.5f) * (float)P.height) - 0.5f);
  if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)P.width && fy < (float)P.height)) return false;
  px = (int)fx;
  py = (int)fy;
  return true;
}

// restirPreviousWorldPosition (:367-378): the point at `dist` along the
// history pixel's centre ray.
OL_DEV f3 restir_previous_position(const GiParams& P, int px, int py, float dist) {
  const float ndx = (((float)px + 0.5f) / (float)P.width) * 2.0f - 1.0f;
  const float ndy = (((float)py + 0.5f) / (float)P.height) * 2.0f - 1.0f;
  const float aspect = (float)P.width / (float)P.height;
  const f3 org = mk3(P.cam_o[0], P.cam_o[1], P.cam_o[2]);
  const f3 fw = mk3(P.cam_f[0], P.cam_f[1], P.cam_f[2]);
  const f3 rt = mk3(P.cam_r[0], P.cam_r[1], P.cam_r[2]);
  const f3 up = mk3(P.cam_u[0], P.cam_u[1], P.cam_u[2]);
  if (P.projection == 1u) return ((org + rt * (ndx * aspect)) - up * ndy) + normalize3(fw) * dist;
  return org + normalize3((fw + rt * (ndx * aspect)) - up * ndy) * dist;
}

// restirHistorySurfaceCompatible (:382-393) against the stored surface of
// the history pixel: (shading normal, primary distance or -1), material id.
OL_DEV bool restir_surface_matches(const GiParams& P, const float4* s, int hx, int hy, f3 p, f3 n, float dist,
                                   uint32_t mat_id, bool temporal_center) {
  const float4 a = s[0];
  const uint32_t old_mat = __float_as_uint(s[1].x);
  const f3 delta = restir_previous_position(P, hx, hy, a.w) - p;
  const float tol = fmaxf(0.03f, dist * 0.01f);
  const float err = temporal_center ? length3(delta) : fabsf(dot3(delta, normalize3(n)));
  const float agree = dot3(normalize3(rgb(a)), normalize3(n));
  return a.w >= 0.0f && err <= tol && agree > 0.9f && old_mat == mat_id;
}

// restirSpatialOffset (:319-331)
OL_DEV void restir_offset(uint32_t index, int radius, int& dx, int& dy) {
  const int ox[8] = {1, -1, 0, 0, 1, -1, 1, -1}, oy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
  dx = ox[index & 7u] * radius;
  dy = oy[index & 7u] * radius;
}

// Candidates, then merges of the previous pass's reservoirs at the
// reprojected pixel (temporal) and its spatial neighbours (pairwise), as the
// primary ReSTIR loop (:1539-1700); stores this pass's reservoir and returns
// the selected light point as one shadow entry weighted by the reservoir's
// normalisation.  Incompatible history surfaces are skipped.
OL_DEV void restir_primary(const GiParams& P, const JobDev& J, uint32_t& rs, f3 hit, f3 n, const Mat& mat, f3 incoming,
                           float dist, uint32_t pixel, uint32_t mat_id, NeeSample& out) {
  const DScene& sc = P.sc;
  const uint32_t count = P.restir;
  Reservoir r = empty_reservoir();
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t light;
    float b1, b2, target;
    candidate_generate(sc, rs, hit, n, incoming, mat, k, count, light, b1, b2, target);
    rs = pcg_step(rs);
    reservoir_update(r, light, b1, b2, target, target, 1.0f, pcg_value(rs));
  }
  int px, py;
  if (P.restir_history && restir_reproject(P, hit, px, py)) {
    const uint32_t spatial = P.restir_spatial;
    // Pass s plays the role of the reference's frame frame_index + s.
    const uint32_t rotation = wave_hash(pixel ^ (P.frame + P.sample_base)) & 7u;
    for (uint32_t ri = 0; ri <= spatial; ++ri) {
      int hx = px, hy = py;
      if (ri > 0u) {
        int dx, dy;
        restir_offset((ri - 1u) + rotation, (int)P.restir_radius, dx, dy);
        hx += dx;
        hy += dy;
      }
      const bool inside = hx >= 0 && hy >= 0 && hx < (int)P.width && hy < (int)P.height;
      Reservoir h = empty_reservoir();
      bool present = false, valid = false;
      if (inside) {
        const uint32_t hidx = (uint32_t)hy * P.width + (uint32_t)hx;
        h = reservoir_load(J.res_in[hidx]);
        present = h.light != OL_NONE;
        if (present) valid = restir_surface_matches(P, J.surf_in + 2u * hidx, hx, hy, hit, n, dist, mat_id, ri == 0u);
      }
      if (valid) {
        const float remaining = fmaxf(P.restir_limit - half_hi(r.tc), 0.0f);
        h = reservoir_limit(h, spatial > 0u ? fminf(remaining, 1.0f) : remaining);
        if (half_hi(h.tc) <= 0.0f) h = empty_reservoir();
      } else {
        h = empty_reservoir();
      }
      if (h.light != OL_NONE) {
        f3 dir;
        float d;
        const float ht = candidate_target(candidate_contribution(sc, h.light, half_lo(h.bary), half_hi(h.bary), hit, n,
                                                                 incoming, mat, count, dir, d));
        rs = pcg_step(rs);
        const float u = pcg_value(rs);
        if (ri == 0u) reservoir_merge(r, h, ht, u);
        else if (P.restir_pairwise) reservoir_merge_pairwise(r, h, ht, half_lo(h.tc), u);
        else reservoir_merge_canonical(r, h, ht, u);
      } else if (ri == 0u && !present) {
        break;
      }
    }
  }
  J.res_out[pixel] = reservoir_store(r);
  J.surf_out[2u * pixel] = mk4(n, dist);
  J.surf_out[2u * pixel + 1u] = make_float4(__uint_as_float(mat_id), 0.0f, 0.0f, 0.0f);
  out.valid = false;
  out.kind = 0u;
  if (r.light == OL_NONE) return;
  f3 dir;
  float d;
  const f3 c = candidate_contribution(sc, r.light, half_lo(r.bary), half_hi(r.bary), hit, n, incoming, mat, count, dir,
                                      d);
  if (d <= 0.0f) return;
  out.origin = hit + n * 0.002f;
  out.dir = dir;
  out.tmax = fmaxf(d - 0.004f, 0.001f);
  out.c = c * reservoir_normalization(r);
  out.valid = true;
}

// Next-event samples of a ReSTIR primary hit in the reference's order: point
// lights, the reservoir's light point (`chosen`, counted like a point light),
// then min(environment_samples, 4) environment samples drawn from `rs`.
template <class Emit>
OL_DEV void restir_nee(const GiParams& P, const NeeSample& chosen, uint32_t& rs, f3 hit, f3 n, const Mat& m,
                       f3 incoming, uint32_t& env_n, Emit&& emit) {
  for (uint32_t li = 0; li < 64u; ++li) {
    if (li >= P.point_count) break;
    NeeSample s = point_sample(P.sc, li, hit, n, m, incoming);
    if (s.valid) emit(s);
  }
  if (chosen.valid) emit(chosen);
  const uint32_t ec = P.environment_samples < 4u ? P.environment_samples : 4u;
  env_n = ec;
  for (uint32_t k = 0; k < 4u; ++k) {
    if (k >= ec) break;
    NeeSample s = env_sample(P.sc, rs, hit, n, m, incoming, ec, 1.0f);
    if (s.valid) emit(s);
  }
}

// Shading record of each leaf triangle (BvhOutput::shade): the vertex
// normals, material, original triangle id and edge cross product that
// resolve_surface needs, gathered once per hierarchy instead of per hit.
__global__ void __launch_bounds__(256) k_shade_records(const float4* __restrict__ tri, const uint4* __restrict__ vidx,
                                                       const float4* __restrict__ normals, uint32_t count,
                                                       uint32_t material, float4* __restrict__ rec,
                                                       const float4* __restrict__ face_in,
                                                       const float* __restrict__ metal_in, float4* __restrict__ face) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  const uint32_t id = __float_as_uint(tri[3u * s].w);
  const uint4 vi = vidx[id];
  const f3 cw = cross3(mk3(tri[3u * s + 1u]), mk3(tri[3u * s + 2u]));
  rec[4u * s] = mk4(rgb(normals[vi.x]), __uint_as_float(material ? vi.w : 0u));
  rec[4u * s + 1u] = mk4(rgb(normals[vi.y]), __uint_as_float(id));
  rec[4u * s + 2u] = mk4(rgb(normals[vi.z]), face_in ? metal_in[id] : 0.0f);
  rec[4u * s + 3u] = mk4(cw, 0.0f);
  if (face_in) face[s] = face_in[id];
}

// A hit's per-triangle material parameters (scenes with them): base colour,
// roughness and metallic factor replace the record's.
OL_DEV void apply_face(const DScene& sc, uint32_t slot, Mat& mat) {
  mat.br = sc.face[slot];
  mat.em.w = sc.shade[4u * slot + 2u].w;
}

// Hit reconstruction (shadeLoadHit) and surface (shadeResolveSurface,
// wave_shade_candidate.glsl:957-997, :1170-1215): position, geometric
// normal, oriented shading normal, the world-space edge cross product, the
// original triangle, its material and the entering flag.
OL_DEV void resolve_surface(const Wave& W, const DScene& sc, uint32_t i, uint32_t slot, float4 ro, float4 rd, float4 hv,
                            f3& hitp, f3& gn, f3& nrm, f3& cw, uint32_t& prim, uint32_t& mat_id, bool& entering) {
  const float t = hv.x, bu = hv.y, bv = hv.z;
  f3 n0, n1, n2;
  if (!sc.instanced) {
    const float4* r = sc.shade + 4u * slot;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    n0 = rgb(r0);
    n1 = rgb(r1);
    n2 = rgb(r2);
    cw = rgb(r3);
    mat_id = __float_as_uint(r0.w);
    prim = __float_as_uint(r1.w);
  } else {
    // Object-space triangle of the instance: normals and the edge cross
    // product go to world space through the normal matrix; the cross
    // product of world edges is det(A) * N * (e1 x e2).
    const InstanceDev& I = sc.inst[W.hit_inst[i]];
    const float4* r = sc.blas[I.blas].shade + 4u * slot;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    prim = I.tri_offset + __float_as_uint(r1.w);
    n0 = normalize3(xform_vector(I.nrm, rgb(r0)));
    n1 = normalize3(xform_vector(I.nrm, rgb(r1)));
    n2 = normalize3(xform_vector(I.nrm, rgb(r2)));
    cw = xform_vector(I.nrm, rgb(r3)) * I.nrm[0].w;
    mat_id = I.material;
  }
  hitp = mk3(ro.x + t * rd.x, ro.y + t * rd.y, ro.z + t * rd.z);
  gn = normalize3(cw);
  const float wx = (1.0f - bu) - bv, wy = bu, wz = bv;
  f3 sn = normalize3(((n0 * wx) + (n1 * wy)) + (n2 * wz));
  if (dot3(sn, gn) < 0.0f) sn = -sn;
  entering = dot3(rgb(rd), gn) < 0.0f;
  nrm = entering ? sn : -sn;
}

// ---- textured surfaces ------------------------------------------------------------------------
struct f2 {
  float x, y;
};
OL_DEV f2 mk2(float x, float y) { return f2{x, y}; }

// sqrt(uv area / world area) of a triangle (shadeTriangleUvDensity, :872-879).
OL_DEV float uv_density(f3 a, f3 b, f3 c, f2 ua, f2 ub, f2 uc) {
  const f2 e1 = mk2(ub.x - ua.x, ub.y - ua.y), e2 = mk2(uc.x - ua.x, uc.y - ua.y);
  const float uv_area = fabsf(e1.x * e2.y - e1.y * e2.x);
  const float world_area = length3(cross3(b - a, c - a));
  return sqrtf(uv_area / fmaxf(world_area, 1e-8f));
}

// Tangent frame from a triangle's uv1 gradients (shadeTriangleTangent, :889-913).
OL_DEV float4 triangle_tangent(f3 a, f3 b, f3 c, f2 ua, f2 ub, f2 uc, f3 n) {
  const f3 ea = b - a, eb = c - a;
  const f2 da = mk2(ub.x - ua.x, ub.y - ua.y), db = mk2(uc.x - ua.x, uc.y - ua.y);
  const float det = da.x * db.y - da.y * db.x;
  if (fabsf(det) < 1e-8f) {
    f3 fb = cross3(n, mk3(0.0f, 1.0f, 0.0f));
    if (fabsf(n.z) < 0.999f) fb = cross3(n, mk3(0.0f, 0.0f, 1.0f));
    return mk4(normalize3(fb), 1.0f);
  }
  const float inv = 1.0f / det;
  const f3 t = normalize3((ea * db.y - eb * da.y) * inv);
  const f3 bt = normalize3((eb * da.x - ea * db.x) * inv);
  return mk4(t, dot3(cross3(n, t), bt) < 0.0f ? -1.0f : 1.0f);
}

// Material textures and normal map of a hit on a flat scene
// (shadeResolveSurface's textured branch, :971-993; shadeApplyMaterialTextures
// :915-934; shadeApplyNormalTexture :936-955).  `sn` is the shading normal
// on the geometric side; it returns the mapped one, also on that side.
OL_DEV void texture_surface(const DScene& sc, uint32_t mat_id, uint32_t prim, float bu, float bv, f3 gn,
                            float cone_width, Mat& mat, f3& sn) {
  const float4 ti = sc.materials[12u * mat_id + 4u];
  const uint4 vi = sc.tri_vidx[prim];
  const float wx = (1.0f - bu) - bv, wy = bu, wz = bv;
  const float4 qa = sc.uvs[vi.x], qb = sc.uvs[vi.y], qc = sc.uvs[vi.z];
  const float u0 = ((qa.x * wx) + (qb.x * wy)) + (qc.x * wz);
  const float v0 = ((qa.y * wx) + (qb.y * wy)) + (qc.y * wz);
  float u1 = 0.0f, v1 = 0.0f, fp1 = 0.0f;
  const float fp0 = cone_width * sc.uv_density[prim];
  const f3 a = rgb(sc.positions[vi.x]), b = rgb(sc.positions[vi.y]), c = rgb(sc.positions[vi.z]);
  const f2 a1 = mk2(qa.z, qa.w), b1 = mk2(qb.z, qb.w), c1 = mk2(qc.z, qc.w);
  if (mat.id.z - floorf(mat.id.z) > 0.125f) {
    u1 = ((qa.z * wx) + (qb.z * wy)) + (qc.z * wz);
    v1 = ((qa.w * wx) + (qb.w * wy)) + (qc.w * wz);
    fp1 = cone_width * uv_density(a, b, c, a1, b1, c1);
  }
  const float tr = sample_material_texture(sc, mat.tp.w, u0, v0, u1, v1, false, fp0, fp1).x;
  mat.at.w = mat.at.w * tr;
  const float4 base = sample_material_texture(sc, ti.x, u0, v0, u1, v1, true, fp0, fp1);
  mat.br.x = mat.br.x * base.x;
  mat.br.y = mat.br.y * base.y;
  mat.br.z = mat.br.z * base.z;
  const float4 emit = sample_material_texture(sc, ti.z, u0, v0, u1, v1, true, fp0, fp1);
  mat.em.x = mat.em.x * emit.x;
  mat.em.y = mat.em.y * emit.y;
  mat.em.z = mat.em.z * emit.z;
  if (!(mat.at.w > 0.001f)) {
    const float4 mr = sample_material_texture(sc, ti.y, u0, v0, u1, v1, false, fp0, fp1);
    mat.br.w = mat.br.w * mr.y;
    mat.em.w = mat.em.w * mr.z;
    const float occ = sample_material_texture(sc, mat.tp.y, u0, v0, u1, v1, false, fp0, fp1).x;
    mat.tp.w = mixf(1.0f, occ, mat.tp.z);
  }
  // Interpolated vertex tangents, or the uv1 triangle frame when the normal
  // map reads uv1.
  const float4 ta = sc.tangents[vi.x], tb = sc.tangents[vi.y], tc = sc.tangents[vi.z];
  float4 tan = make_float4(((ta.x * wx) + (tb.x * wy)) + (tc.x * wz), ((ta.y * wx) + (tb.y * wy)) + (tc.y * wz),
                           ((ta.z * wx) + (tb.z * wy)) + (tc.z * wz), ((ta.w * wx) + (tb.w * wy)) + (tc.w * wz));
  const int nb = (int)ti.w;
  if (nb >= 0 && sc.bindings[2 * nb].w > 0.5f) tan = triangle_tangent(a, b, c, a1, b1, c1, sn);
  if (nb >= 0) {
    const float4 s4 = sample_material_texture(sc, ti.w, u0, v0, u1, v1, false, fp0, fp1);
    f3 tn = mk3(s4.x * 2.0f - 1.0f, s4.y * 2.0f - 1.0f, s4.z * 2.0f - 1.0f);
    tn.x = tn.x * mat.tp.x;
    tn.y = tn.y * mat.tp.x;
    tn = normalize3(tn);
    const f3 td = rgb(tan);
    const f3 t = normalize3(td - sn * dot3(sn, td));
    const f3 bt = cross3(sn, t) * tan.w;
    const float4 r0 = sc.bindings[2 * nb], r1 = sc.bindings[2 * nb + 1];
    const float cs = r0.y, sns = r0.z;
    const float ix = 1.0f / r1.z, iy = 1.0f / r1.w;
    const f3 tt = normalize3(t * (cs * ix) + bt * ((-sns) * iy));
    const f3 tbn = normalize3(t * (sns * ix) + bt * (cs * iy));
    sn = normalize3(((tt * tn.x) + (tbn * tn.y)) + (sn * tn.z));
  }
  if (dot3(sn, gn) < 0.0f) sn = -sn;
}

// Texture pass of the jobs whose scenes have textured materials, before the
// ReSTIR pass and shade: for every hit, the material after its textures and
// the shading normal (per input queue slot), and the cone of the path's next
// ray (shadeRayCone :1217-1222 at the hit; the spread widened by the
// roughness after an opaque scatter, :1551).  Primary rays start with a zero
// cone, as wavefront_generate.comp:153-155 leaves it.
__global__ void __launch_bounds__(256) k_texture(Wave W, uint32_t in, uint32_t gen) {
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t flat_chunks = W.ctr[C_RP + kRuns];
  const PlanHead top = plan_head(W.ctr + C_RP);
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t c = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; c < flat_chunks; c += waves) {
    const ChunkLoc k = locate_chunk(W.ctr + C_RP, top, counts, W.jobs, uniform(c));
    const uint32_t job = uniform(k.job);
    const JobDev& J = W.jobs[job];
    const DScene& sc = J.P.sc;
    if (!sc.textured) continue;
    if (lane_id() >= k.count) continue;
    const uint32_t i = k.first + lane_id();
    const float4 hv = W.hit[i];
    const uint32_t slot = __float_as_uint(hv.w);
    if (slot == OL_NONE) continue;
    const uint32_t path = W.qp[in][i];
    const bool fresh = gen && ((W.list->fresh >> job) & 1ull);
    const float2 cone = fresh ? make_float2(0.0f, 0.0f) : W.cone[path];
    const float4 ro = W.qo[in][i], rd = W.qd[in][i];
    f3 hitp, gn, nrm, cw;
    uint32_t prim, mat_id;
    bool entering;
    resolve_surface(W, sc, i, slot, ro, rd, hv, hitp, gn, nrm, cw, prim, mat_id, entering);
    f3 sn = entering ? nrm : -nrm;
    Mat mat = load_mat(sc.materials, mat_id);
    if (sc.face) apply_face(sc, slot, mat);
    const float width = cone.x + hv.x * cone.y;
    if (sc.materials[12u * mat_id + 8u].w == 1.0f) texture_surface(sc, mat_id, prim, hv.y, hv.z, gn, width, mat, sn);
    float4* out = W.tex + 4u * i;
    out[0] = mat.br;
    out[1] = mat.em;
    out[2] = make_float4(mat.at.w, mat.tp.w, 0.0f, 0.0f);
    out[3] = mk4(sn, 0.0f);
    W.cone[path] = make_float2(width, mat.at.w > 0.001f ? cone.y : cone.y + mat.br.w * 0.25f);
  }
}

// A textured job's hit: the material after its textures (from k_texture).
OL_DEV void textured_material(const Wave& W, uint32_t i, Mat& mat) {
  const float4* r = W.tex + 4u * i;
  mat.br = r[0];
  mat.em = r[1];
  const float4 a = r[2];
  mat.at.w = a.x;
  mat.tp.w = a.y;
}

// ReSTIR pass at primary hits, before shade: for each bounce-0 path of a
// ReSTIR job, clears the pixel's stored reservoir and surface, and where
// shade will sample lights (an opaque hit below the bounce limit, chosen for
// next-event estimation) runs the reservoir pass with the path's random
// state and leaves the selected light sample and the advanced state for
// shade.  Kept out of k_shade so that its registers do not lower shade's
// occupancy for every other job.  Four waves per SIMD (128 VGPRs, a few
// dwords spilled) instead of the three that 161 VGPRs allow: the pass waits
// on dependent light loads (instanced sphere field: 13.73 -> 13.48 ms of
// shading and reservoir passes per frame).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_restir(Wave W, uint32_t in, uint32_t gen) {
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t flat_chunks = W.ctr[C_RP + kRuns];
  const uint32_t chunks = flat_chunks + W.ctr[C_RP2 + kRuns];
  const PlanHead top = plan_head(W.ctr + C_RP), top2 = plan_head(W.ctr + C_RP2);
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t c = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; c < chunks; c += waves) {
    const uint32_t cu = uniform(c);
    const ChunkLoc k = cu < flat_chunks ? locate_chunk(W.ctr + C_RP, top, counts, W.jobs, cu)
                                        : locate_chunk(W.ctr + C_RP2, top2, counts, W.jobs, cu - flat_chunks);
    const JobDev& J = W.jobs[uniform(k.job)];
    const GiParams& P = J.P;
    if (!P.restir) continue;
    // Primary hits only: the chunks generated this step.
    if (!gen || !((W.list->fresh >> uniform(k.job)) & 1ull)) continue;
    if (lane_id() >= k.count) continue;
    const uint32_t i = k.first + lane_id();
    const uint32_t path = W.qp[in][i];
    const float4 T = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    const uint4 M = fresh_path_state(P, path - J.base);
    J.res_out[M.x] = reservoir_store(empty_reservoir());
    J.surf_out[2u * M.x] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    J.surf_out[2u * M.x + 1u] = make_float4(__uint_as_float(OL_NONE), 0.0f, 0.0f, 0.0f);
    const float4 hv = W.hit[i];
    const uint32_t slot = __float_as_uint(hv.w);
    if (fmaxf(T.x, fmaxf(T.y, T.z)) < 0.0001f || slot == OL_NONE || 1u >= P.max_bounces) continue;
    const DScene& sc = P.sc;
    const float4 ro = W.qo[in][i], rd = W.qd[in][i];
    f3 hitp, gn, nrm, cw;
    uint32_t prim, mat_id;
    bool entering;
    resolve_surface(W, sc, i, slot, ro, rd, hv, hitp, gn, nrm, cw, prim, mat_id, entering);
    Mat mat = load_mat(sc.materials, mat_id);
    if (sc.face) apply_face(sc, slot, mat);
    if (sc.textured) {
      textured_material(W, i, mat);
      const f3 sn = rgb(W.tex[4u * i + 3u]);
      nrm = entering ? sn : -sn;
    }
    if (mat.at.w > 0.001f) continue;
    if (!select_secondary_nee(clampf(P.nee_probability, 1e-6f, 1.0f), M.x, M.y, 1u)) continue;
    uint32_t rs = M.z;
    NeeSample s;
    restir_primary(P, J, rs, hitp, nrm, mat, rgb(rd), hv.x, M.x, mat_id, s);
    float4* out = W.rsamp + 3u * path;
    out[0] = mk4(s.origin, s.valid ? 1.0f : 0.0f);
    out[1] = mk4(s.dir, s.tmax);
    out[2] = mk4(s.c, __uint_as_float(rs));
  }
}

// ---- shade -----------------------------------------------------------------------------------
// One wave per 64-ray chunk (grid-stride over chunks), so the job and its
// parameters are wave-uniform.  Outputs go to the job's segment
// (chunk index within the job) % stripes: a job's chunks number at most
// n_paths / 64 + stripes, so a segment of (ceil(n_paths / (64 * stripes)) + 1)
// * 64 slots cannot overflow.
// Four waves per SIMD: the register budget drops from 153 to 128 VGPRs with
// a few spilled dwords, which measured 7 % faster than three waves.
// Shading of one chunk of the input queue by one wave: lanes below k.count
// hold entry k.first + lane, with its path slot, ray (origin / tmin,
// direction / tmax) and hit record (t, u, v, triangle slot bits).  kFull =
// false compiles out the ReSTIR primary hit, textured materials and the
// multi-sample next-event replay, for steps whose jobs use none of them.
template <bool kFull>
OL_DEV void shade_chunk(const Wave& W, uint32_t in, uint32_t gen, const ChunkLoc& k, uint32_t path_in, float4 ro,
                        float4 rd, float4 hv) {
  const uint32_t out = in ^ 1u;
  {
    const uint32_t job = uniform(k.job);
    const JobDev& J = W.jobs[job];
    const GiParams& P = J.P;
    const DScene& sc = P.sc;
    const uint32_t so = uniform(k.chunk_in_job) % J.stripes;
    const uint32_t run = job * kJobStripes + so;
    const bool live = lane_id() < k.count;
    const uint32_t i = k.first + lane_id();
    bool cont = false, nee = false, used_restir = false;
    float4 co = make_float4(0, 0, 0, 0), cd = co;
    uint32_t path = 0, nee_count = 0, area_n = 0, env_n = 0, nee_rng = 0;
    f3 nee_T = mk3(0, 0, 0), hitp = nee_T, nrm = nee_T, incoming = nee_T;
    float nee_p = 1.0f;
    uint32_t nee_mat = 0;  // material of the hit, reloaded to replay its samples
    NeeSample single;
    single.valid = false;
    if (live) {
      path = path_in;
      const bool fresh = gen && ((W.list->fresh >> job) & 1ull);
      float4 T = fresh ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : W.T[path];
      float4 L = fresh ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : W.L[path];
      uint4 M = fresh ? fresh_path_state(P, path - J.base) : W.M[path];
      incoming = rgb(rd);
      const uint32_t slot = __float_as_uint(hv.w);
      do {
        if (fmaxf(T.x, fmaxf(T.y, T.z)) < 0.0001f) {
          M.w &= ~F_ACTIVE;
          break;
        }
        // Sample 0's primary hit for the named products.
        const bool record_primary = J.geo != nullptr && T.w == 0.0f && P.sample_base == 0u &&
                                    path - J.base < P.width * P.height;
        if (slot == OL_NONE) {
          if (record_primary) {
            J.geo[M.x] = make_float4(__builtin_huge_valf(), 0.0f, 0.0f, 0.0f);
            J.prim[M.x] = OL_NONE;
          }
          // shadeResolveEnvironmentMiss
          f3 env = environment_radiance(sc, incoming);
          float mis = 1.0f;
          if ((M.w & F_PREV_DIFFUSE) && P.environment_samples > 0u) {
            float pdf = L.w;
            float lpdf = pdf * (float)P.environment_samples;
            if (M.w & F_PREV_UNIFIED) lpdf = pdf * (1.0f - unified_area_probability(P));
            mis = power_heuristic(pdf, lpdf);
          }
          f3 add = (rgb(T) * env) * mis;
          L.x = L.x + add.x;
          L.y = L.y + add.y;
          L.z = L.z + add.z;
          M.w &= ~F_ACTIVE;
          break;
        }
        // Hit reconstruction (shadeLoadHit) and surface (shadeResolveSurface).
        const float t = hv.x;
        f3 gn, cw;
        uint32_t prim, mat_id;
        bool entering;
        resolve_surface(W, sc, i, slot, ro, rd, hv, hitp, gn, nrm, cw, prim, mat_id, entering);
        if (W.work) count_active(&W.work[6]);
        if (record_primary) {
          J.geo[M.x] = make_float4(t, gn.x, gn.y, gn.z);
          J.prim[M.x] = prim;
        }
        Mat mat = load_mat(sc.materials, mat_id);
        if (sc.face) apply_face(sc, slot, mat);
        if (kFull && sc.textured) {
          textured_material(W, i, mat);
          const f3 sn = rgb(W.tex[4u * i + 3u]);
          nrm = entering ? sn : -sn;
        }
        nee_mat = mat_id;
        uint32_t rs = M.z;
        uint32_t depth = (M.w >> 8) > 1u ? (M.w >> 8) : 1u;
        const uint32_t bounce = (uint32_t)T.w;
        // shadeEmissionContribution
        if (entering || mat.id.w > 0.5f) {
          float emis = 1.0f;
          f3 e = rgb(mat.em);
          if ((M.w & F_PREV_DIFFUSE) && dot3(e, e) > 0.0f) {
            float area = 0.5f * length3(cw);
            float raw = dot3(gn, -incoming);
            float lc = fmaxf(raw, 0.0f);
            if (mat.id.w > 0.5f) lc = fabsf(raw);
            float lum = dot3(e, mk3(0.2126f, 0.7152f, 0.0722f));
            float sel = (area * lum) / fmaxf(sc.area_weight, 1e-6f);
            float lpdf = ((sel * t) * t) / fmaxf(lc * area, 1e-6f);
            float sampled = lpdf * (float)(P.secondary_area_samples > 1u ? P.secondary_area_samples : 1u);
            if (M.w & F_PREV_UNIFIED) sampled = lpdf * unified_area_probability(P);
            emis = power_heuristic(L.w, sampled);
          }
          f3 add = (rgb(T) * e) * emis;
          L.x = L.x + add.x;
          L.y = L.y + add.y;
          L.z = L.z + add.z;
        }
        const uint32_t next_bounce = bounce + 1u;
        if (next_bounce >= P.max_bounces) {
          T.w = (float)next_bounce;
          M.w &= ~F_ACTIVE;
          break;
        }
        float transmission = mat.at.w;
        f3 next_dir;
        float bsdf_pdf = 0.0f;
        if (transmission > 0.001f) {
          // shadeTransmitPath
          float4* st = W.stk + 4u * path;
          float cur = stack_get(st, depth - 1u);
          float target = 1.0f;
          if (entering) target = fmaxf(mat.id.x, 1.0001f);
          else if (depth > 1u) target = stack_get(st, depth - 2u);
          f3 dir = refract3(incoming, nrm, cur / target);
          if (dot3(dir, dir) < 0.01f) {
            dir = reflect3(incoming, nrm);
          } else if (entering && depth < 16u) {
            stack_set(st, depth, target);
            depth = depth + 1u;
          } else if (!entering && depth > 1u) {
            depth = depth - 1u;
          }
          float optical = entering ? 0.0f : t;
          if (mat.a1.z > 0.5f) optical = mat.a1.w;
          float expo = optical / fmaxf(mat.id.y, 1e-6f);
          f3 absorb = mk3(powf(fmaxf(mat.at.x, 1e-6f), expo), powf(fmaxf(mat.at.y, 1e-6f), expo),
                          powf(fmaxf(mat.at.z, 1e-6f), expo));
          f3 tint = mix3(mk3(1.0f, 1.0f, 1.0f), rgb(mat.br), mat.a1.z);
          f3 nt = ((rgb(T) * absorb) * tint) * transmission;
          T.x = nt.x;
          T.y = nt.y;
          T.z = nt.z;
          next_dir = dir;
        } else {
          const float nee_prob = clampf(P.nee_probability, 1e-6f, 1.0f);
          if (select_secondary_nee(nee_prob, M.x, M.y, next_bounce)) {
            nee = true;
            nee_T = rgb(T);
            nee_p = nee_prob;
            nee_rng = rs;
            if (kFull && P.restir && bounce == 0u) {
              // Primary hit: point lights, the ReSTIR area-light sample (from
              // k_restir, with the random state it advanced to), then the
              // environment samples (wavefront_primary_impl.glsl:1527-1785).
              const float4* rsp = W.rsamp + 3u * path;
              const float4 ra = rsp[0], rb = rsp[1], rc = rsp[2];
              single.origin = rgb(ra);
              single.valid = ra.w != 0.0f;
              single.dir = rgb(rb);
              single.tmax = rb.w;
              single.c = rgb(rc);
              single.kind = 0u;
              rs = __float_as_uint(rc.w);
              nee_rng = rs;
              restir_nee(P, single, rs, hitp, nrm, mat, incoming, env_n, [&](const NeeSample&) { ++nee_count; });
              used_restir = true;
            } else if (!kFull || !P.two_pass) {
              // At most one sample: evaluate once and keep it in registers.
              for_each_nee(P, rs, hitp, nrm, mat, incoming, area_n, env_n, [&](const NeeSample& s) {
                single = s;
                nee_count = 1u;
              });
              if (W.work && area_n) atomicAdd(&W.work[7], (unsigned long long)area_n);
            } else {
              for_each_nee(P, rs, hitp, nrm, mat, incoming, area_n, env_n, [&](const NeeSample&) { ++nee_count; });
              if (W.work && area_n) atomicAdd(&W.work[7], (unsigned long long)area_n);
            }
          }
          // shadeScatterOpaquePath / shadeSamplePbr
          const f3 view = -incoming;
          const float prob = spec_probability(mat);
          rs = pcg_step(rs);
          const bool specular = pcg_value(rs) < prob;
          rs = pcg_step(rs);
          float u1 = pcg_value(rs);
          rs = pcg_step(rs);
          float u2 = pcg_value(rs);
          f3 o = cosine_hemisphere(nrm, u1, u2);
          if (specular) o = reflect3(incoming, ggx_half(nrm, mat.br.w, u1, u2));
          if (dot3(nrm, o) <= 0.0f) {
            rs = pcg_step(rs);
            u1 = pcg_value(rs);
            rs = pcg_step(rs);
            u2 = pcg_value(rs);
            o = cosine_hemisphere(nrm, u1, u2);
          }
          o = normalize3(o);
          const float pdf = fmaxf(pdf_pbr(mat, nrm, view, o), 1e-6f);
          f3 w = fdiv3(eval_pbr(mat, nrm, view, o) * fmaxf(dot3(nrm, o), 0.0f), pdf);
          w = w * mixf(mat.tp.w, 1.0f, mat.em.w);
          T.x = T.x * w.x;
          T.y = T.y * w.y;
          T.z = T.z * w.z;
          next_dir = o;
          bsdf_pdf = pdf;
        }
        // shadeApplyRussianRoulette
        if (!(P.roulette_start == 0u || next_bounce < P.roulette_start || transmission > 0.001f)) {
          float survival = clampf(fmaxf(T.x, fmaxf(T.y, T.z)), P.roulette_min_survival, 0.95f);
          rs = pcg_step(rs);
          if (pcg_value(rs) >= survival) {
            M.w &= ~F_ACTIVE;
            M.z = rs;
            break;
          }
          T.x = T.x / survival;
          T.y = T.y / survival;
          T.z = T.z / survival;
          M.z = rs;
        }
        // shadeBuildContinuation
        const f3 dir = normalize3(next_dir);
        const uint32_t capture = M.w & F_CAPTURE;
        T.w = (float)next_bounce;
        M.w = (F_ACTIVE | (depth << 8)) | capture;
        if (transmission <= 0.001f) {
          M.w |= F_PREV_DIFFUSE;
          // After ReSTIR the next hit weighs emission against per-sample
          // area NEE (continuation flags of wavefront_primary_impl.glsl:1837).
          if (P.unified && !used_restir) M.w |= F_PREV_UNIFIED;
          L.w = bsdf_pdf;
        }
        M.z = rs;
        cont = true;
        co = mk4(hitp + dir * 0.002f, 0.001f);
        cd = mk4(dir, 1e30f);
      } while (false);
      // Throughput and state are read only by the path's next shading.
      if (cont) {
        W.T[path] = T;
        W.M[path] = M;
      }
      W.L[path] = L;
    }
    // Continuation slot in the job's output segment, the shadow group's slot
    // and its entries' slots (three atomics per wave, issued together).
    const bool has_group = nee && nee_count > 0u;
    const Reserve3 rv = wave_reserve3(&W.ctr[c_queue(out) + run_word(run)], cont,
                                      &W.ctr[c_groups(in) + run_word(run)], has_group,
                                      &W.ctr[c_entries(in) + run_word(run)], has_group ? nee_count : 0u);
    const uint32_t cl = rv.a;
    if (cont) {
      if (cl < J.append) {
        const uint32_t cslot = J.base + so * J.seg + cl;
        W.qo[out][cslot] = co;
        W.qd[out][cslot] = cd;
        W.qp[out][cslot] = path;
      } else {
        atomicAdd(&W.ctr[C_OVERFLOW], 1u);
        W.M[path].w &= ~F_ACTIVE;
      }
    }
    // Shadow group + entries.
    const uint32_t gl = rv.b, el = rv.n;
    if (has_group) {
      if (gl < J.append && el + nee_count <= J.eseg) {
        const uint32_t gslot = J.base + so * J.seg + gl, eslot = J.ebase + so * J.eseg + el;
        W.gT[gslot] = mk4(nee_T, nee_p);
        W.gI[gslot] = make_uint4(path, eslot, nee_count,
                                 area_n | (env_n << 8) | ((P.unified && !used_restir ? 0u : 1u) << 16));
        if (kFull && used_restir) {
          uint32_t kk = 0;
          uint32_t rs2 = nee_rng;
          uint32_t en;
          Mat mat = load_mat(sc.materials, nee_mat);
          if (sc.face) apply_face(sc, __float_as_uint(hv.w), mat);
          if (sc.textured) textured_material(W, i, mat);
          restir_nee(P, single, rs2, hitp, nrm, mat, incoming, en, [&](const NeeSample& s) {
            W.eo[eslot + kk] = mk4(s.origin, 0.001f);
            W.ed[eslot + kk] = mk4(s.dir, s.tmax);
            W.ec[eslot + kk] = mk4(s.c, (float)s.kind);
            ++kk;
          });
        } else if (!kFull || !P.two_pass) {
          W.eo[eslot] = mk4(single.origin, 0.001f);
          W.ed[eslot] = mk4(single.dir, single.tmax);
          W.ec[eslot] = mk4(single.c, (float)single.kind);
        } else {
          uint32_t kk = 0, an, en;
          uint32_t rs2 = nee_rng;
          Mat mat = load_mat(sc.materials, nee_mat);
          if (sc.face) apply_face(sc, __float_as_uint(hv.w), mat);
          if (sc.textured) textured_material(W, i, mat);
          for_each_nee(P, rs2, hitp, nrm, mat, incoming, an, en, [&](const NeeSample& s) {
            W.eo[eslot + kk] = mk4(s.origin, 0.001f);
            W.ed[eslot + kk] = mk4(s.dir, s.tmax);
            W.ec[eslot + kk] = mk4(s.c, (float)s.kind);
            ++kk;
          });
        }
      } else {
        atomicAdd(&W.ctr[C_OVERFLOW], 1u);
      }
    }
  }
}

template <bool kFull>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_shade(Wave W, uint32_t in,
                                                                                       uint32_t gen) {
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t flat_chunks = W.ctr[C_RP + kRuns];
  const uint32_t chunks = flat_chunks + W.ctr[C_RP2 + kRuns];
  const PlanHead top = plan_head(W.ctr + C_RP), top2 = plan_head(W.ctr + C_RP2);
  const uint32_t waves = gridDim.x * (blockDim.x / 64u);
  for (uint32_t c = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; c < chunks; c += waves) {
    const uint32_t cu = uniform(c);
    const ChunkLoc k = cu < flat_chunks ? locate_chunk(W.ctr + C_RP, top, counts, W.jobs, cu)
                                        : locate_chunk(W.ctr + C_RP2, top2, counts, W.jobs, cu - flat_chunks);
    const uint32_t i = k.first + lane_id();
    uint32_t path = 0;
    float4 ro = make_float4(0, 0, 0, 0), rd = ro, hv = ro;
    if (lane_id() < k.count) {
      path = W.qp[in][i];
      ro = W.qo[in][i];
      rd = W.qd[in][i];
      hv = W.hit[i];
    }
    shade_chunk<kFull>(W, in, gen, k, path, ro, rd, hv);
  }
}

// ---- fused primary step ------------------------------------------------------------------------
// The camera rays of a step whose every chunk is fresh (flat, untextured
// scenes without ReSTIR), traced and shaded by one kernel: each wave takes a
// chunk, traverses its rays to their closest hits (one ray per lane, no
// refill; camera rays of an 8x8 tile are coherent) and shades them in
// registers, so the hit records never go through memory and the step runs
// one kernel instead of a traversal and a shading launch
// (processPrimaryPixel, wavefront_primary_impl.glsl:1097).  Hits and shading
// are those of the split kernels, bit for bit.
template <int LSTACK, int BLOCK, bool kFull>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_primary(Wave W, uint32_t in, uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &W.ctr[C_STACK]);
  const uint32_t* counts = W.ctr + c_queue(in);
  const uint32_t chunks = W.ctr[C_RP + kRuns];
  const PlanHead top = plan_head(W.ctr + C_RP);
  const uint32_t waves = gridDim.x * (BLOCK / 64u);
  uint32_t nv = 0, nt = 0;
  for (uint32_t c = blockIdx.x * (BLOCK / 64u) + threadIdx.x / 64u; c < chunks; c += waves) {
    const ChunkLoc k = locate_chunk(W.ctr + C_RP, top, counts, W.jobs, uniform(c));
    const DScene& sc = W.jobs[uniform(k.job)].P.sc;
    const uint32_t i = k.first + lane_id();
    uint32_t path = 0;
    float4 ro = make_float4(0, 0, 0, 0), rd = ro, hv = ro;
    if (lane_id() < k.count) {
      f3 org, dir;
      generated_ray(W.jobs[uniform(k.job)], i - W.jobs[uniform(k.job)].base, org, dir, path);
      ro = mk4(org, 0.001f);
      rd = mk4(dir, 1e30f);
      hv = make_float4(rd.w, 0.0f, 0.0f, __uint_as_float(OL_NONE));
      if (sc.node_count != 0u) {
        ClosestRay q;
        closest_begin(q, mk3(ro), mk3(rd), ro.w, rd.w);
        st.sp = 0;
        while (!closest_step<LSTACK>(sc.nodes, sc.tri, q, st, nv, nt)) {
        }
        hv = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.slot));
      }
    }
    shade_chunk<kFull>(W, in, 1u, k, path, ro, rd, hv);
  }
  if (W.work && (nv | nt)) {
    atomicAdd(&W.work[0], (unsigned long long)nv);
    atomicAdd(&W.work[1], (unsigned long long)nt);
  }
}

// ---- step epilogue -----------------------------------------------------------------------------
// Per-job report: rays traced from the input queue, shadow rays traced (the
// previous step's entries), work left (paths in the output queue plus the
// shadow groups this step's shading emitted, which the next step traces)
// and shaded hits (those groups); then a reset of every consumed counter,
// so the input queue and the traced groups can serve the next step's output.
__global__ void __launch_bounds__(1024) k_epilogue(Wave W, uint32_t in) {
  __shared__ uint32_t rep[4 * kMaxJobs];
  const uint32_t t = threadIdx.x;
  if (t < 4 * kMaxJobs) rep[t] = 0u;
  __syncthreads();
  for (uint32_t r = t; r < kRuns; r += blockDim.x) {
    const uint32_t job = r / kJobStripes;
    const uint32_t a = W.ctr[c_queue(in) + run_word(r)], e = W.ctr[c_entries(in ^ 1u) + run_word(r)];
    uint32_t l = W.ctr[c_queue(in ^ 1u) + run_word(r)], g = W.ctr[c_groups(in) + run_word(r)];
    // A reservation past the segment's capacity was counted but not written
    // (shade flagged the overflow): the next step enumerates written entries only.
    const uint32_t cap = W.jobs[job].append;
    if (l > cap) {
      l = cap;
      W.ctr[c_queue(in ^ 1u) + run_word(r)] = cap;
    }
    if (g > cap) {
      g = cap;
      W.ctr[c_groups(in) + run_word(r)] = cap;
    }
    if (a) atomicAdd(&rep[4 * job + 0], a);
    if (e) atomicAdd(&rep[4 * job + 1], e);
    if (l + g) atomicAdd(&rep[4 * job + 2], l + g);
    if (g) atomicAdd(&rep[4 * job + 3], g);
    W.ctr[c_queue(in) + run_word(r)] = 0u;
    W.ctr[c_groups(in ^ 1u) + run_word(r)] = 0u;
    W.ctr[c_entries(in ^ 1u) + run_word(r)] = 0u;
  }
  if (t < 4 * kWorkStripes) W.ctr[C_W + t * kCursorStride] = 0u;
  // Chunk totals of the four plans: a plan left out of the next step (no
  // resident job of its kind) must read as empty.
  if (t == 0) {
    W.ctr[C_RP + kRuns] = 0u;
    W.ctr[C_GP + kRuns] = 0u;
    W.ctr[C_RP2 + kRuns] = 0u;
    W.ctr[C_GP2 + kRuns] = 0u;
  }
  __syncthreads();
  if (t < 4 * kMaxJobs) W.ctr[C_REPORT + t] = rep[t];
}

// Zeroes the whole counter block (engine start).
__global__ void __launch_bounds__(1024) k_clear(Wave W) {
  for (uint32_t i = threadIdx.x; i < C_WORDS; i += blockDim.x) W.ctr[i] = 0u;
  if (W.work && threadIdx.x < kWorkWords) W.work[threadIdx.x] = 0ull;
}

// ---- launch helpers ------------------------------------------------------------------------------
template <class K>
int resident_blocks(K kernel, int block) {
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0) != hipSuccess || occ < 1) occ = 1;
  return occ;
}

uint32_t spill_cap_of(const GiLaunch& l) {
  return l.stack_depth > l.lds_stack ? (uint32_t)(l.stack_depth - l.lds_stack) : 0u;
}

void check_spill(const GiLaunch& l, uint64_t threads) {
  if ((uint64_t)spill_cap_of(l) * threads > l.spill_words)
    throw OlError{-2, "traversal spill buffer too small for the launch"};
}

// Block b of a refilling traversal launch serves work stripe b % kWorkStripes;
// a grid of fewer blocks would leave whole stripes untraced.
void check_stripes(dim3 grid) {
  if (grid.x < kWorkStripes) throw OlError{-2, "traversal grid smaller than the number of work stripes"};
}

template <int LSTACK, int BLOCK>
uint64_t spill_words_for(const GiLaunch& l) {
  int occ = std::max(std::max(resident_blocks(k_trace<LSTACK, BLOCK>, BLOCK), resident_blocks(k_trace2<LSTACK, BLOCK>, BLOCK)),
                     std::max(resident_blocks(k_trace_rays<LSTACK, BLOCK>, BLOCK),
                              resident_blocks(k_trace_rays2<LSTACK, BLOCK>, BLOCK)));
  occ = std::max(occ, std::max(resident_blocks(k_primary<LSTACK, BLOCK, true>, BLOCK),
                               resident_blocks(k_primary<LSTACK, BLOCK, false>, BLOCK)));
  return (uint64_t)spill_cap_of(l) * (uint64_t)l.cu_count * (uint64_t)occ * BLOCK;
}

template <int LSTACK, int BLOCK>
void launch_trace(const GiLaunch& l, const Wave& w, uint32_t in, hipStream_t st) {
  static int occ = resident_blocks(k_trace<LSTACK, BLOCK>, BLOCK);
  dim3 grid((unsigned)(l.cu_count * occ));
  check_stripes(grid);
  check_spill(l, (uint64_t)grid.x * BLOCK);
  hipLaunchKernelGGL((k_trace<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, w, in, l.spill, spill_cap_of(l));
}

template <int LSTACK, int BLOCK>
void launch_trace2(const GiLaunch& l, const Wave& w, uint32_t in, hipStream_t st) {
  static int occ = resident_blocks(k_trace2<LSTACK, BLOCK>, BLOCK);
  dim3 grid((unsigned)(l.cu_count * occ));
  check_stripes(grid);
  check_spill(l, (uint64_t)grid.x * BLOCK);
  hipLaunchKernelGGL((k_trace2<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, w, in, l.spill, spill_cap_of(l));
}

template <int LSTACK, int BLOCK>
void launch_primary(const GiLaunch& l, const Wave& w, uint32_t in, bool full, hipStream_t st) {
  static int occ_full = resident_blocks(k_primary<LSTACK, BLOCK, true>, BLOCK);
  static int occ_one = resident_blocks(k_primary<LSTACK, BLOCK, false>, BLOCK);
  dim3 grid((unsigned)(l.cu_count * (full ? occ_full : occ_one)));
  check_spill(l, (uint64_t)grid.x * BLOCK);
  if (full)
    hipLaunchKernelGGL((k_primary<LSTACK, BLOCK, true>), grid, dim3(BLOCK), 0, st, w, in, l.spill, spill_cap_of(l));
  else
    hipLaunchKernelGGL((k_primary<LSTACK, BLOCK, false>), grid, dim3(BLOCK), 0, st, w, in, l.spill, spill_cap_of(l));
}

template <int LSTACK, int BLOCK>
void launch_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                 uint32_t* overflow, unsigned long long* work, hipStream_t st) {
  static int occ = std::min(resident_blocks(k_trace_rays<LSTACK, BLOCK>, BLOCK),
                            resident_blocks(k_trace_rays2<LSTACK, BLOCK>, BLOCK));
  uint32_t need = (n + BLOCK - 1) / BLOCK;
  uint32_t cap = (uint32_t)(l.cu_count * occ);
  // The grid strides over the rays, so a smaller spill buffer only bounds
  // the lanes in flight.
  if (spill_cap_of(l)) cap = (uint32_t)std::min<uint64_t>(cap, l.spill_words / ((uint64_t)spill_cap_of(l) * BLOCK));
  if (cap == 0) throw OlError{-2, "traversal spill buffer too small for one block"};
  dim3 grid(need < cap ? (need ? need : 1) : cap);
  check_spill(l, (uint64_t)grid.x * BLOCK);
  if (sc.instanced)
    hipLaunchKernelGGL((k_trace_rays2<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, o, d, n, hits, overflow, work,
                       l.spill, spill_cap_of(l));
  else
    hipLaunchKernelGGL((k_trace_rays<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, o, d, n, hits, overflow, work,
                       l.spill, spill_cap_of(l));
}

// (LDS entries, block size) variants the launch autotuner chooses from.
#define OL_STACK_DISPATCH(FN, ...)                                                  \
  switch (l.lds_stack * 1000 + l.trace_block) {                                     \
    case 8128: return FN<8, 128>(__VA_ARGS__);                                      \
    case 12128: return FN<12, 128>(__VA_ARGS__);                                    \
    case 16064: return FN<16, 64>(__VA_ARGS__);                                     \
    case 16128: return FN<16, 128>(__VA_ARGS__);                                    \
    case 16256: return FN<16, 256>(__VA_ARGS__);                                    \
    case 24128: return FN<24, 128>(__VA_ARGS__);                                    \
    case 32128: return FN<32, 128>(__VA_ARGS__);                                    \
    case 48128: return FN<48, 128>(__VA_ARGS__);                                    \
    default: throw OlError{-1, "unsupported traversal stack / block configuration"}; \
  }

}  // namespace

int stack_depth_for(uint32_t bvh_depth) {
  uint32_t need = 3u * bvh_depth + 1u;
  const int depths[] = {16, 24, 32, 48, 64, 96, 128};
  for (int d : depths)
    if ((uint32_t)d >= need) return d;
  return -1;
}

uint64_t trace_spill_words(const GiLaunch& l) { OL_STACK_DISPATCH(spill_words_for, l); }

void wave_clear(const Wave& w, hipStream_t st) { hipLaunchKernelGGL(k_clear, dim3(1), dim3(1024), 0, st, w); }

void build_shade_records(BvhOutput& b, const uint4* vidx, const float4* normals, uint32_t count, bool material,
                         hipStream_t st, const float4* face_in, const float* metal_in) {
  if (count == 0) return;
  b.shade.reserve(sizeof(float4) * 4ull * count);
  if (face_in) b.face.reserve(sizeof(float4) * count);
  hipLaunchKernelGGL(k_shade_records, dim3((count + 255u) / 256u), dim3(256), 0, st, b.tri.as<float4>(), vidx, normals,
                     count, material ? 1u : 0u, b.shade.as<float4>(), face_in, metal_in,
                     face_in ? b.face.as<float4>() : nullptr);
  OL_HIP_CHECK(hipGetLastError());
}

void wave_accumulate(const Wave& w, int cu_count, hipStream_t st) {
  hipLaunchKernelGGL(k_accumulate, dim3((unsigned)(cu_count * 4)), dim3(256), 0, st, w);
}

void wave_generate(const Wave& w, uint32_t in, int cu_count, hipStream_t st, bool rays) {
  hipLaunchKernelGGL(k_generate, dim3((unsigned)(cu_count * 8)), dim3(256), 0, st, w, in, rays ? 1u : 0u);
}

void wave_trace(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, hipStream_t st) {
  // One launch plans the input ray queue and the previous step's shadow
  // groups for each scene kind present.
  PlanSet ps{};
  uint32_t n = 0;
  auto add = [&](uint32_t counts, uint32_t out, uint32_t kind) {
    ps.counts[n] = counts;
    ps.out[n] = out;
    ps.kind[n] = kind;
    ++n;
  };
  if (kinds & STEP_FLAT) {
    add(c_queue(in), C_RP, 0u);
    add(c_groups(in ^ 1u), C_GP, 0u);
  }
  if (kinds & STEP_INSTANCED) {
    add(c_queue(in), C_RP2, 1u);
    add(c_groups(in ^ 1u), C_GP2, 1u);
  }
  if (n == 0) return;
  hipLaunchKernelGGL(k_plan, dim3(n), dim3(1024), 0, st, w.ctr, w.jobs, ps);
  if (kinds & STEP_FLAT) [&] { OL_STACK_DISPATCH(launch_trace, l, w, in, st); }();
  if (kinds & STEP_INSTANCED) [&] { OL_STACK_DISPATCH(launch_trace2, l, w, in, st); }();
}

void wave_primary(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, hipStream_t st) {
  PlanSet ps{};
  ps.counts[0] = c_queue(in);
  ps.out[0] = C_RP;
  ps.kind[0] = 0u;
  ps.counts[1] = c_groups(in ^ 1u);
  ps.out[1] = C_GP;
  ps.kind[1] = 0u;
  hipLaunchKernelGGL(k_plan, dim3(2), dim3(1024), 0, st, w.ctr, w.jobs, ps);
  OL_STACK_DISPATCH(launch_primary, l, w, in, (kinds & STEP_TWO_PASS) != 0u, st);
}

void wave_shade(const GiLaunch& l, const Wave& w, uint32_t in, uint32_t kinds, bool gen, hipStream_t st) {
  if (w.tex && (kinds & STEP_TEXTURED)) {
    static int tocc = resident_blocks(k_texture, 256);
    hipLaunchKernelGGL(k_texture, dim3((unsigned)(l.cu_count * tocc)), dim3(256), 0, st, w, in, gen ? 1u : 0u);
  }
  if (w.rsamp && (kinds & STEP_RESTIR) && gen) {
    static int rocc = resident_blocks(k_restir, 256);
    hipLaunchKernelGGL(k_restir, dim3((unsigned)(l.cu_count * rocc)), dim3(256), 0, st, w, in, 1u);
  }
  if (kinds & (STEP_RESTIR | STEP_TEXTURED | STEP_TWO_PASS)) {
    static int occ = resident_blocks(k_shade<true>, 256);
    hipLaunchKernelGGL(k_shade<true>, dim3((unsigned)(l.cu_count * occ)), dim3(256), 0, st, w, in, gen ? 1u : 0u);
  } else {
    static int occ = resident_blocks(k_shade<false>, 256);
    hipLaunchKernelGGL(k_shade<false>, dim3((unsigned)(l.cu_count * occ)), dim3(256), 0, st, w, in, gen ? 1u : 0u);
  }
}

void wave_epilogue(const Wave& w, uint32_t in, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue, dim3(1), dim3(1024), 0, st, w, in);
}

void camera_rays(const GiParams& P, uint32_t x0, uint32_t y0, uint32_t tw, uint32_t th, uint32_t sample, float4* o,
                 float4* d, hipStream_t st) {
  const uint32_t n = tw * th;
  hipLaunchKernelGGL(k_camera_rays, dim3((n + 255u) / 256u), dim3(256), 0, st, P, x0, y0, tw, th, sample, o, d);
}

void trace_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                uint32_t* overflow, unsigned long long* work, hipStream_t st) {
  OL_STACK_DISPATCH(launch_rays, l, sc, o, d, n, hits, overflow, work, st);
}
