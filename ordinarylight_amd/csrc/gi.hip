// GI wavefront kernels: generate -> [trace closest -> shade -> trace shadow]
// x bounces -> resolve.
//
// The estimator is the reference's production split-path shade stage
// (shaders/wavefront_shade_candidate.glsl main :3106-3493) for built-in PBR
// materials without textures, custom programs or volumes: emission with MIS,
// environment miss with MIS, point/spot/directional lights, unified or
// per-domain area/environment next-event estimation, PBR/GGX BSDF sampling,
// deterministic transmission with a 16-deep IOR stack, Russian roulette and
// continuation.  Every expression keeps the GLSL evaluation order.
//
// MI355X structure: the shadow rays of next-event estimation are deferred to
// their own any-hit kernel instead of being traced inside shade; each hit's
// visible contributions are summed in the reference's order and added before
// the next bounce, so the radiance of every path is the same sequence of
// float operations as the inline form.  Queue slots are reserved with one
// atomic per wave (ballot + mbcnt) on one of 32 striped counters.  Counts
// stay on the device; kernels use
// grid-stride loops sized to the resident grid, so no host synchronisation
// happens between bounces.

#include <algorithm>

#include "gi.h"
#include "trace.h"

namespace {

constexpr uint32_t F_ACTIVE = 1u, F_PREV_DIFFUSE = 2u, F_PREV_UNIFIED = 4u, F_CAPTURE = 8u;

// ---- materials ------------------------------------------------------------------------
struct Mat {
  float4 br;  // base colour, roughness
  float4 em;  // emission, metallic
  float4 at;  // attenuation colour, transmission
  float4 id;  // ior, attenuation distance, program, two-sided emission
  float4 tp;  // normal scale, occlusion binding, occlusion strength, resolved occlusion factor
  float4 a0;  // clearcoat, clearcoat roughness, sheen roughness, anisotropy
  float4 a1;  // subsurface, radius, thin walled, thickness
  float4 sh;  // sheen colour
  float4 ss;  // subsurface colour
};

OL_DEV Mat load_mat(const float4* m, uint32_t id) {
  const float4* r = m + 12u * id;
  Mat x;
  x.br = r[0];
  x.em = r[1];
  x.at = r[2];
  x.id = r[3];
  x.tp = r[5];
  x.a0 = r[6];
  x.a1 = r[7];
  x.sh = r[8];
  x.ss = r[9];
  return x;
}

OL_DEV f3 rgb(float4 v) { return mk3(v.x, v.y, v.z); }
OL_DEV float pow5(float x) { return ((x * x) * (x * x)) * x; }

OL_DEV f3 cosine_hemisphere(f3 n, float u, float v) {
  float radius = sqrtf(u);
  float phi = OL_TWO_PI * v;
  f3 tangent = cross3(n, mk3(0.0f, 1.0f, 0.0f));
  if (fabsf(n.z) < 0.999f) tangent = cross3(n, mk3(0.0f, 0.0f, 1.0f));
  tangent = normalize3(tangent);
  f3 bitangent = cross3(n, tangent);
  float sp, cp;
  ol_sincos(phi, &sp, &cp);
  return normalize3(((tangent * (radius * cp)) + (bitangent * (radius * sp))) + (n * sqrtf(fmaxf(0.0f, 1.0f - u))));
}

OL_DEV f3 fresnel_schlick(f3 f0, float c) {
  float k = pow5(1.0f - clampf(c, 0.0f, 1.0f));
  return f0 + (mk3(1.0f, 1.0f, 1.0f) - f0) * k;
}

OL_DEV float ggx_d(float nh, float roughness) {
  float alpha = fmaxf(roughness * roughness, 0.0009f);
  float a2 = alpha * alpha;
  float den = ((nh * nh) * (a2 - 1.0f)) + 1.0f;
  return a2 / fmaxf((OL_PI * den) * den, 1e-6f);
}

OL_DEV float ggx_g1(float nd, float roughness) {
  float alpha = fmaxf(roughness * roughness, 0.0009f);
  float a2 = alpha * alpha;
  return (2.0f * nd) / fmaxf(nd + sqrtf(a2 + ((1.0f - a2) * nd) * nd), 1e-6f);
}

OL_DEV f3 f0_of(const Mat& m) { return mix3(mk3(0.04f, 0.04f, 0.04f), rgb(m.br), m.em.w); }

OL_DEV float spec_probability(const Mat& m) {
  f3 f0 = f0_of(m);
  return clampf(fmaxf(f0.x, fmaxf(f0.y, f0.z)) + m.a0.x * 0.15f, 0.1f, 0.95f);
}

// shadeEvaluatePbrLobes (:1438-1468); returns diffuse + specular.
OL_DEV f3 eval_pbr(const Mat& m, f3 n, f3 view, f3 out) {
  float nv = fmaxf(dot3(n, view), 0.0f);
  float nl = fmaxf(dot3(n, out), 0.0f);
  if (nv <= 0.0f || nl <= 0.0f) return mk3(0.0f, 0.0f, 0.0f);
  f3 h = normalize3(view + out);
  float nh = fmaxf(dot3(n, h), 0.0f);
  float vh = fmaxf(dot3(view, h), 0.0f);
  float metallic = m.em.w;
  f3 base = rgb(m.br);
  f3 fr = fresnel_schlick(f0_of(m), vh);
  float eff_r = m.br.w * (1.0f - 0.25f * fabsf(clampf(m.a0.w, -1.0f, 1.0f)));
  float D = ggx_d(nh, eff_r);
  float G = ggx_g1(nv, eff_r) * ggx_g1(nl, eff_r);
  float den = fmaxf((4.0f * nv) * nl, 1e-6f);
  f3 spec = ((fr * D) * G) / den;
  f3 diffuse = (((mk3(1.0f, 1.0f, 1.0f) - fr) * (1.0f - metallic)) * base) / OL_PI;
  float subsurface = clampf(m.a1.x, 0.0f, 1.0f);
  diffuse = mix3(diffuse, diffuse * rgb(m.ss), subsurface);
  float sv = pow5(1.0f - vh);
  f3 sheen = rgb(m.sh) * ((1.0f - clampf(m.a0.z, 0.0f, 1.0f)) * sv);
  float clearcoat = clampf(m.a0.x, 0.0f, 1.0f);
  float coat_r = clampf(m.a0.y, 0.02f, 1.0f);
  float cD = ggx_d(nh, coat_r);
  float cG = ggx_g1(nv, coat_r) * ggx_g1(nl, coat_r);
  float cF = 0.04f + 0.96f * sv;
  float coat = (((clearcoat * cD) * cG) * cF) / den;
  float layer = 1.0f - clearcoat * cF;
  f3 d = diffuse * layer;
  f3 s = ((spec + sheen) * layer) + mk3(coat, coat, coat);
  return d + s;
}

OL_DEV float pdf_pbr(const Mat& m, f3 n, f3 view, f3 out) {
  float nl = fmaxf(dot3(n, out), 0.0f);
  if (nl <= 0.0f) return 0.0f;
  f3 h = normalize3(view + out);
  float nh = fmaxf(dot3(n, h), 0.0f);
  float vh = fmaxf(dot3(view, h), 1e-6f);
  float sp = (ggx_d(nh, m.br.w) * nh) / (4.0f * vh);
  float dp = nl / OL_PI;
  return mixf(dp, sp, spec_probability(m));
}

OL_DEV f3 ggx_half(f3 n, float roughness, float u, float v) {
  float alpha = fmaxf(roughness * roughness, 0.0009f);
  float a2 = alpha * alpha;
  float phi = OL_TWO_PI * v;
  float c = sqrtf((1.0f - u) / fmaxf(1.0f + (a2 - 1.0f) * u, 1e-6f));
  float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
  f3 tangent = cross3(n, mk3(0.0f, 1.0f, 0.0f));
  if (fabsf(n.z) < 0.999f) tangent = cross3(n, mk3(0.0f, 0.0f, 1.0f));
  tangent = normalize3(tangent);
  f3 bitangent = cross3(n, tangent);
  float sp, cp;
  ol_sincos(phi, &sp, &cp);
  return normalize3((((tangent * s) * cp) + ((bitangent * s) * sp)) + (n * c));
}

OL_DEV float power_heuristic(float a, float b) {
  float a2 = a * a, b2 = b * b;
  return a2 / fmaxf(a2 + b2, 1e-6f);
}

OL_DEV f3 analytic_environment(f3 d) {
  float sky = fmaxf(d.y, 0.0f);
  return mix3(mk3(0.018f, 0.022f, 0.032f), mk3(0.32f, 0.46f, 0.72f), sky);
}

OL_DEV f3 environment_radiance(const DScene& sc, f3 d) {
  if (!sc.env_valid) return analytic_environment(d);
  return mk3(sc.env_r, sc.env_g, sc.env_b);
}

OL_DEV float unified_area_probability(const GiParams& P) {
  if (P.area_count == 0u) return 0.0f;
  if (P.environment_samples == 0u) return 1.0f;
  float w = sqrtf(fmaxf(P.sc.area_weight, 1e-6f));
  return w / (w + 1.0f);
}

OL_DEV bool select_secondary_nee(float probability, uint32_t pixel, uint32_t frame_sample, uint32_t bounce) {
  if (probability >= 0.999999f) return true;
  uint32_t frame = frame_sample >> 8, sample = frame_sample & 255u;
  uint32_t scramble = wave_hash((pixel ^ wave_hash(bounce + 1u)) ^ wave_hash(sample + 1u));
  uint32_t sequence = bit_reverse32(frame) ^ scramble;
  float selector = ((float)sequence + 0.5f) * (1.0f / 4294967296.0f);
  return selector < probability;
}

// ---- next-event samples ---------------------------------------------------------------
struct NeeSample {
  f3 origin, dir;
  float tmax;
  f3 c;        // unoccluded contribution (lobe contribution of the reference)
  uint32_t kind;
  bool valid;
};

OL_DEV NeeSample point_sample(const DScene& sc, uint32_t li, f3 hit, f3 n, const Mat& m, f3 incoming) {
  NeeSample s;
  s.valid = false;
  s.kind = 0;
  const float4* L = sc.analytic + 4u * li;
  float4 pt = L[0], dr = L[1], ci = L[2], sp = L[3];
  int type = (int)(pt.w + 0.5f);
  if (type == 3) return s;
  f3 dir = mk3(0.0f, 0.0f, 0.0f);
  float dist2 = 1.0f, dist = 10000.0f, att = 1.0f;
  if (type == 1) {
    dir = -normalize3(rgb(dr));
  } else {
    f3 off = rgb(pt) - hit;
    dist2 = fmaxf(dot3(off, off), 1e-6f);
    dist = sqrtf(dist2);
    dir = off / dist;
    if (dr.w > 0.0f && dist > dr.w) return s;
    att = 1.0f / dist2;
    if (type == 2) {
      float cone = dot3(normalize3(rgb(dr)), -dir);
      float amount = clampf((cone - sp.y) / fmaxf(sp.x - sp.y, 1e-6f), 0.0f, 1.0f);
      float spot = (amount * amount) * (3.0f - 2.0f * amount);
      if (spot <= 0.0f) return s;
      att = att * spot;
    }
  }
  float cosine = fmaxf(dot3(n, dir), 0.0f);
  if (cosine <= 0.0f) return s;
  s.origin = hit + n * 0.002f;
  s.tmax = fmaxf(dist - 0.004f, 0.001f);
  if (type == 1) s.tmax = 10000.0f;
  f3 incident = (rgb(ci) * ci.w) * att;
  s.dir = dir;
  // shadePointLightContribution with unit volume transmittance.
  s.c = ((eval_pbr(m, n, -incoming, dir) * incident) * mk3(1.0f, 1.0f, 1.0f)) * cosine;
  s.valid = true;
  return s;
}

OL_DEV uint32_t select_area_light(const DScene& sc, float selection) {
  uint32_t lower = 0, upper = sc.area_count - 1u;
  for (int step = 0; step < 32; ++step) {
    if (lower >= upper) break;
    uint32_t middle = lower + (upper - lower) / 2u;
    if (selection <= sc.area[5u * middle + 4u].x) upper = middle; else lower = middle + 1u;
  }
  return lower;
}

OL_DEV NeeSample area_sample(const DScene& sc, uint32_t& rs, f3 hit, f3 n, const Mat& m, f3 incoming,
                             uint32_t sample_index, uint32_t sample_count, float technique) {
  NeeSample s;
  s.valid = false;
  s.kind = 1;
  if (sc.area_count == 0u) return s;
  rs = pcg_step(rs);
  float selection = ((float)sample_index + pcg_value(rs)) / (float)sample_count;
  uint32_t li = select_area_light(sc, selection);
  const float4* A = sc.area + 5u * li;
  f3 a = rgb(A[0]), b = rgb(A[1]), c = rgb(A[2]);
  float4 ea = A[3], dist = A[4];
  rs = pcg_step(rs);
  float ru = sqrtf(pcg_value(rs));
  rs = pcg_step(rs);
  float vv = pcg_value(rs);
  f3 lp = (((1.0f - ru) * a) + ((ru * (1.0f - vv)) * b)) + ((ru * vv) * c);
  f3 off = lp - hit;
  float d2 = dot3(off, off);
  float dl = sqrtf(d2);
  f3 dir = off / fmaxf(dl, 1e-6f);
  float sc_cos = fmaxf(dot3(n, dir), 0.0f);
  f3 ln = normalize3(cross3(b - a, c - a));
  float raw = dot3(ln, -dir);
  float lc = fmaxf(raw, 0.0f);
  if (dist.z > 0.5f) lc = fabsf(raw);
  if (sc_cos <= 0.0f || lc <= 1e-6f) return s;
  float light_pdf = (dist.y * d2) / fmaxf(lc * ea.w, 1e-6f);
  float eff = light_pdf * fmaxf(technique, 1e-6f);
  s.origin = hit + n * 0.002f;
  s.dir = dir;
  s.tmax = fmaxf(dl - 0.004f, 0.001f);
  // shadeAreaLightContribution with unit volume transmittance.
  f3 view = -incoming;
  float bp = pdf_pbr(m, n, view, dir);
  float mis = power_heuristic(eff * (float)sample_count, bp);
  s.c = ((((eval_pbr(m, n, view, dir) * rgb(ea)) * sc_cos) * mis) * mk3(1.0f, 1.0f, 1.0f)) / fmaxf(eff, 1e-6f);
  s.valid = true;
  return s;
}

OL_DEV NeeSample env_sample(const DScene& sc, uint32_t& rs, f3 hit, f3 n, const Mat& m, f3 incoming,
                            uint32_t sample_count, float technique) {
  NeeSample s;
  s.kind = 2;
  rs = pcg_step(rs);
  float u = pcg_value(rs);
  rs = pcg_step(rs);
  float v = pcg_value(rs);
  f3 dir = cosine_hemisphere(n, u, v);
  float cosine = fmaxf(dot3(n, dir), 0.0f);
  float pdf = cosine / OL_PI;
  float eff = pdf * fmaxf(technique, 1e-6f);
  s.origin = hit + n * 0.002f;
  s.dir = dir;
  s.tmax = 1e30f;
  s.valid = cosine > 0.0f;
  if (!s.valid) return s;
  // shadeEnvironmentContribution with unit volume transmittance.
  float pdf2 = cosine / OL_PI;
  float mis = power_heuristic(eff * (float)sample_count, pdf2);
  float mw = mixf(m.tp.w, 1.0f, m.em.w);
  s.c = (((((eval_pbr(m, n, -incoming, dir) * environment_radiance(sc, dir)) * cosine) * mis) * mk3(1.0f, 1.0f, 1.0f)) /
         fmaxf(eff, 1e-6f)) * mw;
  return s;
}

// Visits every next-event sample of one hit in the reference's order and
// calls emit(sample) for the valid ones.  Returns counts for the divisors.
template <class Emit>
OL_DEV void for_each_nee(const GiParams& P, uint32_t& rs, f3 hit, f3 n, const Mat& m, f3 incoming, uint32_t& area_n,
                         uint32_t& env_n, Emit&& emit) {
  const DScene& sc = P.sc;
  for (uint32_t li = 0; li < 64u; ++li) {
    if (li >= P.point_count) break;
    NeeSample s = point_sample(sc, li, hit, n, m, incoming);
    if (s.valid) emit(s);
  }
  area_n = 0;
  env_n = 0;
  if (P.unified) {
    bool area_on = P.area_count > 0u, env_on = P.environment_samples > 0u;
    if (!area_on && !env_on) return;
    float aw = 0.0f;
    if (area_on) aw = sqrtf(fmaxf(sc.area_weight, 1e-6f)) * (float)(P.secondary_area_samples > 1u ? P.secondary_area_samples : 1u);
    float ew = 0.0f;
    if (env_on) ew = (float)(P.environment_samples < 4u ? P.environment_samples : 4u);
    float ap = aw / fmaxf(aw + ew, 1e-6f);
    rs = pcg_step(rs);
    bool area_sel = pcg_value(rs) < ap;
    if (area_sel) {
      NeeSample s = area_sample(sc, rs, hit, n, m, incoming, 0u, 1u, ap);
      if (s.valid) emit(s);
    } else {
      NeeSample s = env_sample(sc, rs, hit, n, m, incoming, 1u, 1.0f - ap);
      if (s.valid) emit(s);
    }
    return;
  }
  uint32_t ac = P.secondary_area_samples < 1u ? 1u : (P.secondary_area_samples > 16u ? 16u : P.secondary_area_samples);
  area_n = ac;
  for (uint32_t k = 0; k < 16u; ++k) {
    if (k >= ac) break;
    NeeSample s = area_sample(sc, rs, hit, n, m, incoming, k, ac, 1.0f);
    if (s.valid) emit(s);
  }
  uint32_t ec = P.environment_samples < 4u ? P.environment_samples : 4u;
  env_n = ec;
  for (uint32_t k = 0; k < 4u; ++k) {
    if (k >= ec) break;
    NeeSample s = env_sample(sc, rs, hit, n, m, incoming, ec, 1.0f);
    if (s.valid) emit(s);
  }
}

// ---- medium stack -----------------------------------------------------------------------------
OL_DEV float stack_get(const float4* st, uint32_t i) {
  float4 v = st[i >> 2];
  uint32_t k = i & 3u;
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
OL_DEV void stack_set(float4* st, uint32_t i, float value) {
  float4 v = st[i >> 2];
  uint32_t k = i & 3u;
  if (k == 0) v.x = value; else if (k == 1) v.y = value; else if (k == 2) v.z = value; else v.w = value;
  st[i >> 2] = v;
}

// GLSL refract(I, N, eta).
OL_DEV f3 refract3(f3 I, f3 N, float eta) {
  float ndi = dot3(N, I);
  float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
  if (k < 0.0f) return mk3(0.0f, 0.0f, 0.0f);
  return eta * I - (eta * ndi + sqrtf(k)) * N;
}

// ---- striped queues ---------------------------------------------------------------------------
// Inclusive prefix of the kStripes counts at `counts` into `incl` (LDS).
// Every thread of the block calls it; wave 0 does the scan.
OL_DEV void stripe_prefix(const uint32_t* counts, uint32_t* incl) {
  if (threadIdx.x < 64u) {
    uint32_t v = threadIdx.x < kStripes ? counts[threadIdx.x] : 0u;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t o = __shfl_up(v, (unsigned)off, 64);
      if ((int)threadIdx.x >= off) v += o;
    }
    if (threadIdx.x < kStripes) incl[threadIdx.x] = v;
  }
  __syncthreads();
}

// Slot of the j-th live entry (j < incl[kStripes - 1]): stripe s holds
// entries [incl[s-1], incl[s]) at slots s * seg + 0, 1, ...
OL_DEV uint32_t stripe_slot(const uint32_t* incl, uint32_t seg, uint32_t j) {
  uint32_t s = 0;
#pragma unroll
  for (uint32_t step = kStripes / 2u; step; step >>= 1)
    if (incl[s + step - 1u] <= j) s += step;
  return s * seg + (j - (s ? incl[s - 1u] : 0u));
}

OL_DEV uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---- kernels ----------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_generate(GiParams P, GiState S) {
  const uint32_t W = P.width, H = P.height;
  const uint32_t tiles_x = (W + 7u) / 8u, tiles_y = (H + 7u) / 8u;
  const uint32_t per_sample = tiles_x * tiles_y * 64u;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t local = g / per_sample;
  const uint32_t rem = g - local * per_sample;
  const uint32_t sample = P.sample_base + local;
  const uint32_t tile = rem >> 6, lane = rem & 63u;
  const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  const uint32_t lx = lane & 7u, ly = lane >> 3;
  const uint32_t px = tx * 8u + lx, py = ty * 8u + ly;
  // The primary queue is dense: all entries in stripe 0 (slot = index).
  if (g == 0) S.ctr[CTR_Q0] = P.spp * W * H;
  if (!(local < P.spp && px < W && py < H)) return;
  // Queue slot in 8x8 tile order without atomics: full tile rows above,
  // full-width tiles to the left in this tile row, row-major inside the tile.
  const uint32_t th = min(8u, H - 8u * ty), tw = min(8u, W - 8u * tx);
  const uint32_t slot = local * (W * H) + 8u * ty * W + 8u * tx * th + ly * tw + lx;
  const uint32_t pixel = py * W + px;
  const uint32_t path = local * (W * H) + pixel;
  uint32_t rng = wave_hash((pixel ^ wave_hash(P.frame)) ^ wave_hash(sample + 1u));
  rng = wave_hash(rng);
  const float jx = u32_to_unit(rng);
  rng = wave_hash(rng);
  const float jy = u32_to_unit(rng);
  const float ndx = (((float)px + jx) / (float)W) * 2.0f - 1.0f;
  const float ndy = (((float)py + jy) / (float)H) * 2.0f - 1.0f;
  const float aspect = (float)W / (float)H;
  f3 org = mk3(P.cam_o[0], P.cam_o[1], P.cam_o[2]);
  const f3 fw = mk3(P.cam_f[0], P.cam_f[1], P.cam_f[2]);
  const f3 rt = mk3(P.cam_r[0], P.cam_r[1], P.cam_r[2]);
  const f3 up = mk3(P.cam_u[0], P.cam_u[1], P.cam_u[2]);
  f3 dir;
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
  S.qo[0][slot] = mk4(org, 0.001f);
  S.qd[0][slot] = mk4(dir, 1e30f);
  S.qp[0][slot] = path;
  S.T[path] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  S.L[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  S.M[path] = make_uint4(pixel, (P.frame << 8) | (sample & 255u), rng, 257u);
  if (S.stk) S.stk[4u * path] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
}

// One wave: adds the input queue and the previous bounce's shadow entries to
// the ray statistics, then empties the output queue and the shadow queues.
__global__ void __launch_bounds__(64) k_prologue(GiState S, uint32_t bounce) {
  const uint32_t in = bounce & 1u, out = in ^ 1u, k = threadIdx.x;
  const uint32_t q = wave_sum(k < kStripes ? S.ctr[ctr_queue(in) + k] : 0u);
  const uint32_t e = wave_sum(k < kStripes ? S.ctr[CTR_E + k] : 0u);
  if (k == 0) {
    S.stats[bounce == 0 ? 0 : 1] += q;
    S.stats[2] += e;
  }
  if (k < kStripes) {
    S.ctr[ctr_queue(out) + k] = 0u;
    S.ctr[CTR_G + k] = 0u;
    S.ctr[CTR_E + k] = 0u;
    S.ctr[CTR_W + k] = 0u;
    S.ctr[CTR_WS + k] = 0u;
  }
}

__global__ void __launch_bounds__(64) k_finish(GiState S) {
  const uint32_t k = threadIdx.x;
  const uint32_t e = wave_sum(k < kStripes ? S.ctr[CTR_E + k] : 0u);
  if (k == 0) S.stats[2] += e;
  if (k < kStripes) S.ctr[CTR_E + k] = 0u;
}

// Whole render: counters, overflow flags, statistics.  Per chunk
// (`all` = 0): the queue counters only; overflow flags accumulate.
__global__ void __launch_bounds__(256) k_reset(GiState S, uint32_t all) {
  const uint32_t k = threadIdx.x;
  if (k < (all ? CTR_WORDS : CTR_OVERFLOW)) S.ctr[k] = 0u;
  if (all && k < 4) S.stats[k] = 0;
  if (all && S.work && k < 4) S.work[k] = 0;
}

// Lane-private traversal stack: LDS column + global spill column.
template <int LSTACK, int BLOCK>
OL_DEV TravStack<LSTACK> lane_stack(uint32_t* lds, uint32_t* spill, uint32_t spill_cap, uint32_t* overflow) {
  TravStack<LSTACK> st;
  st.lds = lds + threadIdx.x;
  st.lds_stride = BLOCK;
  st.spill = spill ? spill + (blockIdx.x * BLOCK + threadIdx.x) : nullptr;
  st.spill_stride = gridDim.x * BLOCK;
  st.spill_cap = spill ? spill_cap : 0u;
  st.overflow = overflow;
  st.sp = 0;
  return st;
}

template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_closest(DScene sc, GiState S, uint32_t in, uint32_t* spill,
                                                         uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  __shared__ uint32_t incl[kStripes];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &S.ctr[CTR_STACK]);
  stripe_prefix(S.ctr + ctr_queue(in), incl);
  const uint32_t count = incl[kStripes - 1u];
  const float4* __restrict__ qo = S.qo[in];
  const float4* __restrict__ qd = S.qd[in];
  uint32_t nv = 0, nt = 0;
  for (uint32_t j = blockIdx.x * BLOCK + threadIdx.x; j < count; j += gridDim.x * BLOCK) {
    const uint32_t i = stripe_slot(incl, S.qseg, j);
    float4 o = qo[i], d = qd[i];
    TraceHit h = ol_trace_closest<LSTACK>(sc.nodes, sc.tri, sc.node_count, mk3(o), mk3(d), o.w, d.w, st, nv, nt);
    S.hit[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.slot));
  }
  if (S.work && (nv | nt)) {
    atomicAdd(&S.work[0], (unsigned long long)nv);
    atomicAdd(&S.work[1], (unsigned long long)nt);
  }
}

// Closest hit with lane refill.  Block b works on stripe b % kStripes of the
// dense input; its waves take 64-ray chunks of that stripe from a global
// cursor (one atomic per chunk) and hand rays to lanes whose traversal has
// finished whenever at least kRefillMin lanes are idle, so a wave no longer
// waits for its slowest ray before starting new ones.  Results are the same
// as k_trace_closest's: each ray's traversal is independent of scheduling.
constexpr uint32_t kChunk = 64, kRefillMin = 16;

template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_closest_pool(DScene sc, GiState S, uint32_t in, uint32_t* spill,
                                                              uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  __shared__ uint32_t incl[kStripes];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &S.ctr[CTR_STACK]);
  stripe_prefix(S.ctr + ctr_queue(in), incl);
  const uint32_t count = incl[kStripes - 1u];
  const uint32_t ws = blockIdx.x % kStripes;
  const uint32_t w_begin = (uint32_t)((uint64_t)count * ws / kStripes);
  const uint32_t w_end = (uint32_t)((uint64_t)count * (ws + 1u) / kStripes);
  uint32_t* cursor = &S.ctr[CTR_W + ws];
  const float4* __restrict__ qo = S.qo[in];
  const float4* __restrict__ qd = S.qd[in];
  uint32_t pool = 0, pool_end = 0;  // wave-uniform: dense rays [pool, pool_end) not yet started
  bool exhausted = false;            // wave-uniform: the stripe has no chunks left
  ClosestRay q;
  uint32_t slot = OL_NONE;
  bool busy = false;
  uint32_t nv = 0, nt = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin) {
      if (pool == pool_end && !exhausted) {
        uint32_t c = 0;
        if (lane_id() == 0) c = atomicAdd(cursor, kChunk);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        const uint32_t remaining = w_end - w_begin;
        if (c >= remaining) {
          exhausted = true;
        } else {
          pool = w_begin + c;
          pool_end = w_begin + min(c + kChunk, remaining);
        }
      }
      const uint32_t take = min(n_idle, pool_end - pool);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          slot = stripe_slot(incl, S.qseg, pool + rank);
          const float4 o = qo[slot], d = qd[slot];
          closest_begin(q, mk3(o), mk3(d), o.w, d.w);
          if (sc.node_count == 0) {
            S.hit[slot] = make_float4(d.w, 0.0f, 0.0f, __uint_as_float(OL_NONE));
          } else {
            st.sp = 0;
            busy = true;
          }
        }
      }
      pool += take;
    }
    if (__ballot(busy) == 0) {
      if (exhausted && pool == pool_end) break;
      continue;
    }
    if (busy && closest_step<LSTACK>(sc.nodes, sc.tri, q, st, nv, nt)) {
      S.hit[slot] = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.slot));
      busy = false;
    }
  }
  if (S.work && (nv | nt)) {
    atomicAdd(&S.work[0], (unsigned long long)nv);
    atomicAdd(&S.work[1], (unsigned long long)nt);
  }
}

// Lane refill as in k_trace_closest_pool, with split inner / leaf steps: a
// wave advances its lanes at leaves when at least a third of its busy lanes
// hold one (or none can take an inner step), its lanes at inner nodes
// otherwise.
template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_closest_split(DScene sc, GiState S, uint32_t in, uint32_t* spill,
                                                               uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  __shared__ uint32_t incl[kStripes];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &S.ctr[CTR_STACK]);
  stripe_prefix(S.ctr + ctr_queue(in), incl);
  const uint32_t count = incl[kStripes - 1u];
  const uint32_t ws = blockIdx.x % kStripes;
  const uint32_t w_begin = (uint32_t)((uint64_t)count * ws / kStripes);
  const uint32_t w_end = (uint32_t)((uint64_t)count * (ws + 1u) / kStripes);
  uint32_t* cursor = &S.ctr[CTR_W + ws];
  const float4* __restrict__ qo = S.qo[in];
  const float4* __restrict__ qd = S.qd[in];
  uint32_t pool = 0, pool_end = 0;
  bool exhausted = false;
  ClosestRay q;
  q.node = OL_NONE;
  uint32_t slot = OL_NONE;
  bool busy = false;
  uint32_t nv = 0, nt = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin) {
      if (pool == pool_end && !exhausted) {
        uint32_t c = 0;
        if (lane_id() == 0) c = atomicAdd(cursor, kChunk);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        const uint32_t remaining = w_end - w_begin;
        if (c >= remaining) {
          exhausted = true;
        } else {
          pool = w_begin + c;
          pool_end = w_begin + min(c + kChunk, remaining);
        }
      }
      const uint32_t take = min(n_idle, pool_end - pool);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          slot = stripe_slot(incl, S.qseg, pool + rank);
          const float4 o = qo[slot], d = qd[slot];
          closest_begin(q, mk3(o), mk3(d), o.w, d.w);
          if (sc.node_count == 0) {
            S.hit[slot] = make_float4(d.w, 0.0f, 0.0f, __uint_as_float(OL_NONE));
          } else {
            st.sp = 0;
            busy = true;
          }
        }
      }
      pool += take;
    }
    const bool at_leaf = busy && item_is_leaf(q.node);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t inner = __ballot(busy && !at_leaf);
    if ((leafy | inner) == 0) {
      if (exhausted && pool == pool_end) break;
      continue;
    }
    if (inner == 0 || 3u * (uint32_t)__popcll(leafy) >= (uint32_t)__popcll(leafy | inner)) {
      if (at_leaf) closest_leaf<LSTACK>(sc.tri, q, st, nt);
    } else if (busy && !at_leaf) {
      closest_inner<LSTACK>(sc.nodes, q, st, nv);
    }
    if (busy && q.node == OL_NONE) {
      S.hit[slot] = make_float4(q.best.t, q.best.u, q.best.v, __uint_as_float(q.best.slot));
      busy = false;
    }
  }
  if (S.work && (nv | nt)) {
    atomicAdd(&S.work[0], (unsigned long long)nv);
    atomicAdd(&S.work[1], (unsigned long long)nt);
  }
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
__global__ void __launch_bounds__(BLOCK) k_trace_shadow(DScene sc, GiState S, uint32_t* spill, uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  __shared__ uint32_t incl[kStripes];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &S.ctr[CTR_STACK]);
  stripe_prefix(S.ctr + CTR_G, incl);
  const uint32_t count = incl[kStripes - 1u];
  uint32_t nv = 0, nt = 0;
  for (uint32_t j = blockIdx.x * BLOCK + threadIdx.x; j < count; j += gridDim.x * BLOCK) {
    const uint32_t g = stripe_slot(incl, S.gseg, j);
    const float4 gT = S.gT[g];
    const uint4 gI = S.gI[g];
    const uint32_t path = gI.x, first = gI.y, n = gI.z;
    const uint32_t area_n = gI.w & 255u, env_n = (gI.w >> 8) & 255u, mode = gI.w >> 16;
    f3 direct = mk3(0.0f, 0.0f, 0.0f), area_direct = direct, env_direct = direct;
    for (uint32_t k = 0; k < n; ++k) {
      const float4 o = S.eo[first + k], d = S.ed[first + k], c = S.ec[first + k];
      if (ol_trace_any<LSTACK>(sc.nodes, sc.tri, sc.node_count, mk3(o), mk3(d), o.w, d.w, st, nv, nt)) continue;
      const uint32_t kind = (uint32_t)c.w;
      if (mode == 0u || kind == 0u) direct = direct + rgb(c);
      else if (kind == 1u) area_direct = area_direct + rgb(c);
      else env_direct = env_direct + rgb(c);
    }
    if (mode != 0u) {
      direct = direct + area_direct / (float)area_n;
      if (env_n > 0u) direct = direct + env_direct / (float)env_n;
    }
    float4 L = S.L[path];
    f3 add = (rgb(gT) * direct) / fmaxf(gT.w, 1e-6f);
    S.L[path] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, L.w);
  }
  if (S.work && (nv | nt)) {
    atomicAdd(&S.work[2], (unsigned long long)nv);
    atomicAdd(&S.work[3], (unsigned long long)nt);
  }
}

// Adds one visible shadow entry's contribution (the summation order of
// k_trace_shadow).
OL_DEV void shadow_add(uint32_t mode, float4 c, f3& direct, f3& area_direct, f3& env_direct) {
  const uint32_t kind = (uint32_t)c.w;
  if (mode == 0u || kind == 0u) direct = direct + rgb(c);
  else if (kind == 1u) area_direct = area_direct + rgb(c);
  else env_direct = env_direct + rgb(c);
}

OL_DEV void shadow_resolve(const GiState& S, uint32_t g, f3 direct, f3 area_direct, f3 env_direct) {
  const float4 gT = S.gT[g];
  const uint4 gI = S.gI[g];
  const uint32_t path = gI.x, area_n = gI.w & 255u, env_n = (gI.w >> 8) & 255u, mode = gI.w >> 16;
  if (mode != 0u) {
    direct = direct + area_direct / (float)area_n;
    if (env_n > 0u) direct = direct + env_direct / (float)env_n;
  }
  float4 L = S.L[path];
  f3 add = (rgb(gT) * direct) / fmaxf(gT.w, 1e-6f);
  S.L[path] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, L.w);
}

// Shadow groups with lane refill and split inner / leaf steps (the
// scheduling of k_trace_closest_split): a lane owns one group and traces its
// entries in order, accumulating visible contributions as k_trace_shadow does.
template <int LSTACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_shadow_split(DScene sc, GiState S, uint32_t* spill,
                                                              uint32_t spill_cap) {
  __shared__ uint32_t lds[LSTACK * BLOCK];
  __shared__ uint32_t incl[kStripes];
  TravStack<LSTACK> st = lane_stack<LSTACK, BLOCK>(lds, spill, spill_cap, &S.ctr[CTR_STACK]);
  stripe_prefix(S.ctr + CTR_G, incl);
  const uint32_t count = incl[kStripes - 1u];
  const uint32_t ws = blockIdx.x % kStripes;
  const uint32_t w_begin = (uint32_t)((uint64_t)count * ws / kStripes);
  const uint32_t w_end = (uint32_t)((uint64_t)count * (ws + 1u) / kStripes);
  uint32_t* cursor = &S.ctr[CTR_WS + ws];
  uint32_t pool = 0, pool_end = 0;
  bool exhausted = false;
  AnyRay q;
  q.node = OL_NONE;
  q.occluded = false;
  uint32_t g = 0, e = 0, e_end = 0, mode = 0;
  f3 direct = mk3(0.0f, 0.0f, 0.0f), area_direct = direct, env_direct = direct;
  bool busy = false;
  uint32_t nv = 0, nt = 0;
  while (true) {
    const uint64_t idle = __ballot(!busy);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefillMin) {
      if (pool == pool_end && !exhausted) {
        uint32_t c = 0;
        if (lane_id() == 0) c = atomicAdd(cursor, kChunk);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        const uint32_t remaining = w_end - w_begin;
        if (c >= remaining) {
          exhausted = true;
        } else {
          pool = w_begin + c;
          pool_end = w_begin + min(c + kChunk, remaining);
        }
      }
      const uint32_t take = min(n_idle, pool_end - pool);
      if (!busy) {
        const uint32_t rank = lanes_below(idle);
        if (rank < take) {
          g = stripe_slot(incl, S.gseg, pool + rank);
          const uint4 gI = S.gI[g];
          e = gI.y;
          e_end = gI.y + gI.z;
          mode = gI.w >> 16;
          direct = mk3(0.0f, 0.0f, 0.0f);
          area_direct = direct;
          env_direct = direct;
          if (sc.node_count == 0) {
            // Nothing can occlude: every entry is visible.
            for (; e < e_end; ++e) shadow_add(mode, S.ec[e], direct, area_direct, env_direct);
            shadow_resolve(S, g, direct, area_direct, env_direct);
          } else if (e < e_end) {
            const float4 o = S.eo[e], d = S.ed[e];
            any_begin(q, mk3(o), mk3(d), o.w, d.w);
            st.sp = 0;
            busy = true;
          } else {
            shadow_resolve(S, g, direct, area_direct, env_direct);
          }
        }
      }
      pool += take;
    }
    const bool at_leaf = busy && item_is_leaf(q.node);
    const uint64_t leafy = __ballot(at_leaf);
    const uint64_t inner = __ballot(busy && !at_leaf);
    if ((leafy | inner) == 0) {
      if (exhausted && pool == pool_end) break;
      continue;
    }
    if (inner == 0 || 3u * (uint32_t)__popcll(leafy) >= (uint32_t)__popcll(leafy | inner)) {
      if (at_leaf) any_leaf<LSTACK>(sc.tri, q, st, nt);
    } else if (busy && !at_leaf) {
      any_inner<LSTACK>(sc.nodes, q, st, nv);
    }
    if (busy && q.node == OL_NONE) {
      if (!q.occluded) shadow_add(mode, S.ec[e], direct, area_direct, env_direct);
      ++e;
      if (e < e_end) {
        const float4 o = S.eo[e], d = S.ed[e];
        any_begin(q, mk3(o), mk3(d), o.w, d.w);
        st.sp = 0;
      } else {
        shadow_resolve(S, g, direct, area_direct, env_direct);
        busy = false;
      }
    }
  }
  if (S.work && (nv | nt)) {
    atomicAdd(&S.work[2], (unsigned long long)nv);
    atomicAdd(&S.work[3], (unsigned long long)nt);
  }
}

struct ShadeOut {
  bool cont;
  float4 co, cd;
};

// Launched with blocks of kStripeUnit threads, so the output stripe
// (j / kStripeUnit) % kStripes of dense input j is uniform per block.
__global__ void __launch_bounds__(kStripeUnit) k_shade(GiParams P, GiState S, uint32_t in) {
  const DScene& sc = P.sc;
  __shared__ uint32_t incl[kStripes];
  stripe_prefix(S.ctr + ctr_queue(in), incl);
  const uint32_t count = incl[kStripes - 1u];
  const uint32_t out = in ^ 1u;
  const uint32_t stride = gridDim.x * blockDim.x;
  // Every lane runs the same number of loop trips so the wave-wide slot
  // reservations below are reached together.
  const uint32_t trips = (count + stride - 1u) / stride;
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t trip = 0; trip < trips; ++trip, j += stride) {
    const bool live = j < count;
    const uint32_t so = (j / kStripeUnit) % kStripes;
    bool cont = false, nee = false;
    float4 co = make_float4(0, 0, 0, 0), cd = co;
    uint32_t path = 0, nee_count = 0, area_n = 0, env_n = 0, nee_rng = 0;
    f3 nee_T = mk3(0, 0, 0), hitp = nee_T, nrm = nee_T, incoming = nee_T;
    float nee_p = 1.0f;
    Mat mat;
    NeeSample single;
    single.valid = false;
    if (live) {
      const uint32_t i = stripe_slot(incl, S.qseg, j);
      path = S.qp[in][i];
      const float4 ro = S.qo[in][i], rd = S.qd[in][i], hv = S.hit[i];
      float4 T = S.T[path], L = S.L[path];
      uint4 M = S.M[path];
      incoming = rgb(rd);
      const uint32_t slot = __float_as_uint(hv.w);
      do {
        if (fmaxf(T.x, fmaxf(T.y, T.z)) < 0.0001f) {
          M.w &= ~F_ACTIVE;
          break;
        }
        if (slot == OL_NONE) {
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
        const float t = hv.x, bu = hv.y, bv = hv.z;
        const float4 ta = sc.tri[3 * slot], te1 = sc.tri[3 * slot + 1], te2 = sc.tri[3 * slot + 2];
        const uint32_t prim = __float_as_uint(ta.w);
        hitp = mk3(ro.x + t * rd.x, ro.y + t * rd.y, ro.z + t * rd.z);
        const f3 gn = normalize3(cross3(mk3(te1), mk3(te2)));
        const uint4 vi = sc.tri_vidx[prim];
        const float wx = (1.0f - bu) - bv, wy = bu, wz = bv;
        const f3 n0 = rgb(sc.normals[vi.x]), n1 = rgb(sc.normals[vi.y]), n2 = rgb(sc.normals[vi.z]);
        f3 sn = normalize3(((n0 * wx) + (n1 * wy)) + (n2 * wz));
        if (dot3(sn, gn) < 0.0f) sn = -sn;
        const bool entering = dot3(incoming, gn) < 0.0f;
        nrm = entering ? sn : -sn;
        mat = load_mat(sc.materials, vi.w);
        uint32_t rs = M.z;
        uint32_t depth = (M.w >> 8) > 1u ? (M.w >> 8) : 1u;
        const uint32_t bounce = (uint32_t)T.w;
        // shadeEmissionContribution
        if (entering || mat.id.w > 0.5f) {
          float emis = 1.0f;
          f3 e = rgb(mat.em);
          if ((M.w & F_PREV_DIFFUSE) && dot3(e, e) > 0.0f) {
            float area = 0.5f * length3(cross3(mk3(te1), mk3(te2)));
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
          float4* st = S.stk + 4u * path;
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
            if (!P.two_pass) {
              // At most one sample: evaluate once and keep it in registers.
              for_each_nee(P, rs, hitp, nrm, mat, incoming, area_n, env_n, [&](const NeeSample& s) {
                single = s;
                nee_count = 1u;
              });
            } else {
              for_each_nee(P, rs, hitp, nrm, mat, incoming, area_n, env_n, [&](const NeeSample&) { ++nee_count; });
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
          f3 w = (eval_pbr(mat, nrm, view, o) * fmaxf(dot3(nrm, o), 0.0f)) / pdf;
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
          if (P.unified) M.w |= F_PREV_UNIFIED;
          L.w = bsdf_pdf;
        }
        M.z = rs;
        cont = true;
        co = mk4(hitp + dir * 0.002f, 0.001f);
        cd = mk4(dir, 1e30f);
      } while (false);
      S.T[path] = T;
      S.L[path] = L;
      S.M[path] = M;
    }
    // Continuation slot in this block's stripe (one atomic per wave).
    const uint32_t cl = wave_reserve(&S.ctr[ctr_queue(out) + so], cont);
    if (cont) {
      if (cl < S.qseg) {
        const uint32_t cslot = so * S.qseg + cl;
        S.qo[out][cslot] = co;
        S.qd[out][cslot] = cd;
        S.qp[out][cslot] = path;
      } else {
        atomicAdd(&S.ctr[CTR_OVERFLOW], 1u);
        S.M[path].w &= ~F_ACTIVE;
      }
    }
    // Shadow group + entries.
    const bool has_group = nee && nee_count > 0u;
    const uint32_t gl = wave_reserve(&S.ctr[CTR_G + so], has_group);
    const uint32_t el = wave_reserve_n(&S.ctr[CTR_E + so], has_group ? nee_count : 0u);
    if (has_group) {
      if (gl < S.gseg && el + nee_count <= S.eseg) {
        const uint32_t gslot = so * S.gseg + gl, eslot = so * S.eseg + el;
        S.gT[gslot] = mk4(nee_T, nee_p);
        S.gI[gslot] = make_uint4(path, eslot, nee_count, area_n | (env_n << 8) | ((P.unified ? 0u : 1u) << 16));
        if (!P.two_pass) {
          S.eo[eslot] = mk4(single.origin, 0.001f);
          S.ed[eslot] = mk4(single.dir, single.tmax);
          S.ec[eslot] = mk4(single.c, (float)single.kind);
        } else {
          uint32_t k = 0, an, en;
          uint32_t rs2 = nee_rng;
          for_each_nee(P, rs2, hitp, nrm, mat, incoming, an, en, [&](const NeeSample& s) {
            S.eo[eslot + k] = mk4(s.origin, 0.001f);
            S.ed[eslot + k] = mk4(s.dir, s.tmax);
            S.ec[eslot + k] = mk4(s.c, (float)s.kind);
            ++k;
          });
        }
      } else {
        atomicAdd(&S.ctr[CTR_OVERFLOW], 1u);
      }
    }
  }
}

// Adds the float32 radiance of this chunk's samples to the float64 pixel
// sums in sample order (render_wavefront_outputs, api.py:1404-1407).
__global__ void __launch_bounds__(256) k_accumulate(GiParams P, GiState S, double4* acc, uint32_t first_chunk) {
  const uint32_t np = P.width * P.height;
  const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= np) return;
  double4 a = first_chunk ? make_double4(0.0, 0.0, 0.0, 0.0) : acc[pixel];
  for (uint32_t s = 0; s < P.spp; ++s) {
    const float4 l = S.L[(size_t)s * np + pixel];
    if (first_chunk && s == 0) {
      a = make_double4((double)l.x, (double)l.y, (double)l.z, 1.0);
    } else {
      a.x += (double)l.x;
      a.y += (double)l.y;
      a.z += (double)l.z;
      a.w += 1.0;
    }
  }
  acc[pixel] = a;
}

__global__ void __launch_bounds__(256) k_final(const double4* acc, uint32_t np, uint32_t samples, float4* out) {
  const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
  if (pixel >= np) return;
  const double n = (double)samples;
  const double4 a = acc[pixel];
  out[pixel] = make_float4((float)(a.x / n), (float)(a.y / n), (float)(a.z / n), (float)(a.w / n));
}

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

template <int LSTACK, int BLOCK>
uint64_t spill_words_for(const GiLaunch& l) {
  int occ = std::max(std::max(std::max(resident_blocks(k_trace_closest<LSTACK, BLOCK>, BLOCK),
                                       resident_blocks(k_trace_closest_split<LSTACK, BLOCK>, BLOCK)),
                              resident_blocks(k_trace_closest_pool<LSTACK, BLOCK>, BLOCK)),
                     std::max(std::max(resident_blocks(k_trace_shadow<LSTACK, BLOCK>, BLOCK),
                                       resident_blocks(k_trace_shadow_split<LSTACK, BLOCK>, BLOCK)),
                              resident_blocks(k_trace_rays<LSTACK, BLOCK>, BLOCK)));
  return (uint64_t)spill_cap_of(l) * (uint64_t)l.cu_count * (uint64_t)occ * BLOCK;
}

template <int LSTACK, int BLOCK>
void launch_closest(const GiLaunch& l, const DScene& sc, const GiState& s, uint32_t in, hipStream_t st) {
  if (l.scheduler == 2) {
    static int occ = resident_blocks(k_trace_closest_split<LSTACK, BLOCK>, BLOCK);
    dim3 grid((unsigned)(l.cu_count * occ));
    check_spill(l, (uint64_t)grid.x * BLOCK);
    hipLaunchKernelGGL((k_trace_closest_split<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, s, in, l.spill,
                       spill_cap_of(l));
    return;
  }
  if (l.scheduler == 1) {
    static int occ = resident_blocks(k_trace_closest_pool<LSTACK, BLOCK>, BLOCK);
    dim3 grid((unsigned)(l.cu_count * occ));
    check_spill(l, (uint64_t)grid.x * BLOCK);
    hipLaunchKernelGGL((k_trace_closest_pool<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, s, in, l.spill,
                       spill_cap_of(l));
    return;
  }
  static int occ = resident_blocks(k_trace_closest<LSTACK, BLOCK>, BLOCK);
  dim3 grid((unsigned)(l.cu_count * occ));
  check_spill(l, (uint64_t)grid.x * BLOCK);
  hipLaunchKernelGGL((k_trace_closest<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, s, in, l.spill, spill_cap_of(l));
}

template <int LSTACK, int BLOCK>
void launch_shadow(const GiLaunch& l, const DScene& sc, const GiState& s, hipStream_t st) {
  if (l.scheduler == 2) {
    static int occ2 = resident_blocks(k_trace_shadow_split<LSTACK, BLOCK>, BLOCK);
    dim3 grid((unsigned)(l.cu_count * occ2));
    check_spill(l, (uint64_t)grid.x * BLOCK);
    hipLaunchKernelGGL((k_trace_shadow_split<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, s, l.spill,
                       spill_cap_of(l));
    return;
  }
  static int occ = resident_blocks(k_trace_shadow<LSTACK, BLOCK>, BLOCK);
  dim3 grid((unsigned)(l.cu_count * occ));
  check_spill(l, (uint64_t)grid.x * BLOCK);
  hipLaunchKernelGGL((k_trace_shadow<LSTACK, BLOCK>), grid, dim3(BLOCK), 0, st, sc, s, l.spill, spill_cap_of(l));
}

template <int LSTACK, int BLOCK>
void launch_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                 uint32_t* overflow, unsigned long long* work, hipStream_t st) {
  static int occ = resident_blocks(k_trace_rays<LSTACK, BLOCK>, BLOCK);
  uint32_t need = (n + BLOCK - 1) / BLOCK;
  uint32_t cap = (uint32_t)(l.cu_count * occ);
  dim3 grid(need < cap ? (need ? need : 1) : cap);
  check_spill(l, (uint64_t)grid.x * BLOCK);
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

void gi_reset(const GiState& s, bool all, hipStream_t st) {
  static_assert(CTR_WORDS <= 256, "counter block larger than the reset kernel");
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, st, s, all ? 1u : 0u);
}

void gi_finish(const GiState& s, hipStream_t st) { hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, s); }

void gi_generate(const GiParams& p, const GiState& s, hipStream_t st) {
  const uint32_t tiles = ((p.width + 7u) / 8u) * ((p.height + 7u) / 8u);
  const uint64_t threads = (uint64_t)tiles * 64u * p.spp;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)((threads + 255u) / 256u)), dim3(256), 0, st, p, s);
}

void gi_bounce_prologue(const GiState& s, uint32_t bounce, hipStream_t st) {
  hipLaunchKernelGGL(k_prologue, dim3(1), dim3(64), 0, st, s, bounce);
}

void gi_trace_closest(const GiLaunch& l, const DScene& sc, const GiState& s, uint32_t in, hipStream_t st) {
  OL_STACK_DISPATCH(launch_closest, l, sc, s, in, st);
}

void gi_trace_shadow(const GiLaunch& l, const DScene& sc, const GiState& s, hipStream_t st) {
  OL_STACK_DISPATCH(launch_shadow, l, sc, s, st);
}

void trace_rays(const GiLaunch& l, const DScene& sc, const float4* o, const float4* d, uint32_t n, float4* hits,
                uint32_t* overflow, unsigned long long* work, hipStream_t st) {
  OL_STACK_DISPATCH(launch_rays, l, sc, o, d, n, hits, overflow, work, st);
}

void gi_shade(const GiLaunch& l, const GiParams& p, const GiState& s, uint32_t in, hipStream_t st) {
  static int occ = resident_blocks(k_shade, (int)kStripeUnit);
  hipLaunchKernelGGL(k_shade, dim3((unsigned)(l.cu_count * occ)), dim3(kStripeUnit), 0, st, p, s, in);
}

void gi_accumulate(const GiParams& p, const GiState& s, double4* acc, bool first, hipStream_t st) {
  const uint32_t np = p.width * p.height;
  hipLaunchKernelGGL(k_accumulate, dim3((np + 255u) / 256u), dim3(256), 0, st, p, s, acc, first ? 1u : 0u);
}

void gi_final(const double4* acc, uint32_t np, uint32_t samples, float4* out, hipStream_t st) {
  hipLaunchKernelGGL(k_final, dim3((np + 255u) / 256u), dim3(256), 0, st, acc, np, samples, out);
}
