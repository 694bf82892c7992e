// Device functions of the GI estimator shared by the wavefront kernels:
// material records, PBR/GGX BSDF, next-event estimation (point, spot,
// directional, area and environment lights), the IOR medium stack and
// refraction.  Each function cites the shader function it restates
// (shaders/wavefront_shade_candidate.glsl); every expression keeps the GLSL
// evaluation order.
#pragma once

#include "gi.h"
#include "texture.h"
#include "trace.h"

namespace olshade {

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

// Division inside the BSDF and light-sample weights: a * rcp(b), within 2
// ulps as GLSL allows its divisions (2.5 ulp), instead of a correctly
// rounded sequence.  Sampled directions and the decisions drawn against
// random numbers keep exact divisions.
OL_DEV float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
OL_DEV f3 fdiv3(f3 a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  return mk3(a.x * r, a.y * r, a.z * r);
}

OL_DEV f3 fresnel_schlick(f3 f0, float c) {
  float k = pow5(1.0f - clampf(c, 0.0f, 1.0f));
  return f0 + (mk3(1.0f, 1.0f, 1.0f) - f0) * k;
}

OL_DEV float ggx_d(float nh, float roughness) {
  float alpha = fmaxf(roughness * roughness, 0.0009f);
  float a2 = alpha * alpha;
  float den = ((nh * nh) * (a2 - 1.0f)) + 1.0f;
  return fdiv(a2, fmaxf((OL_PI * den) * den, 1e-6f));
}

OL_DEV float ggx_g1(float nd, float roughness) {
  float alpha = fmaxf(roughness * roughness, 0.0009f);
  float a2 = alpha * alpha;
  return fdiv(2.0f * nd, fmaxf(nd + sqrtf(a2 + ((1.0f - a2) * nd) * nd), 1e-6f));
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
  f3 spec = fdiv3((fr * D) * G, den);
  f3 diffuse = fdiv3(((mk3(1.0f, 1.0f, 1.0f) - fr) * (1.0f - metallic)) * base, OL_PI);
  float subsurface = clampf(m.a1.x, 0.0f, 1.0f);
  diffuse = mix3(diffuse, diffuse * rgb(m.ss), subsurface);
  float sv = pow5(1.0f - vh);
  f3 sheen = rgb(m.sh) * ((1.0f - clampf(m.a0.z, 0.0f, 1.0f)) * sv);
  float clearcoat = clampf(m.a0.x, 0.0f, 1.0f);
  float coat_r = clampf(m.a0.y, 0.02f, 1.0f);
  float cD = ggx_d(nh, coat_r);
  float cG = ggx_g1(nv, coat_r) * ggx_g1(nl, coat_r);
  float cF = 0.04f + 0.96f * sv;
  float coat = fdiv(((clearcoat * cD) * cG) * cF, den);
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
  float sp = fdiv(ggx_d(nh, m.br.w) * nh, 4.0f * vh);
  float dp = fdiv(nl, OL_PI);
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
  return fdiv(a2, fmaxf(a2 + b2, 1e-6f));
}

OL_DEV f3 analytic_environment(f3 d) {
  float sky = fmaxf(d.y, 0.0f);
  return mix3(mk3(0.018f, 0.022f, 0.032f), mk3(0.32f, 0.46f, 0.72f), sky);
}

OL_DEV f3 environment_radiance(const DScene& sc, f3 d) {
  if (!sc.env_valid) return analytic_environment(d);
  if (sc.env_texture >= 0) return environment_image(sc, d);
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

// First emitter whose cumulative weight reaches the selection (the last
// if none does), by bisection of the CDF (shadeSelectAreaLight,
// wavefront_shade_candidate.glsl:1621-1645).  The guide narrows the range
// to the selection's bucket first; the bisection answer does not depend on
// the starting range as long as it contains that answer.
OL_DEV uint32_t select_area_light(const DScene& sc, float selection) {
  uint32_t lower = 0, upper = sc.area_count - 1u;
  if (sc.area_guide) {
    const uint32_t b = min((uint32_t)(selection * (float)sc.area_guide_size), sc.area_guide_size - 1u);
    lower = sc.area_guide[b];
    upper = sc.area_guide[b + 1u];
  }
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
  float light_pdf = fdiv(dist.y * d2, fmaxf(lc * ea.w, 1e-6f));
  float eff = light_pdf * fmaxf(technique, 1e-6f);
  s.origin = hit + n * 0.002f;
  s.dir = dir;
  s.tmax = fmaxf(dl - 0.004f, 0.001f);
  // shadeAreaLightContribution with unit volume transmittance.
  f3 view = -incoming;
  float bp = pdf_pbr(m, n, view, dir);
  float mis = power_heuristic(eff * (float)sample_count, bp);
  s.c = fdiv3((((eval_pbr(m, n, view, dir) * rgb(ea)) * sc_cos) * mis) * mk3(1.0f, 1.0f, 1.0f), fmaxf(eff, 1e-6f));
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
  float pdf = fdiv(cosine, OL_PI);
  float eff = pdf * fmaxf(technique, 1e-6f);
  s.origin = hit + n * 0.002f;
  s.dir = dir;
  s.tmax = 1e30f;
  s.valid = cosine > 0.0f;
  if (!s.valid) return s;
  // shadeEnvironmentContribution with unit volume transmittance.
  float pdf2 = fdiv(cosine, OL_PI);
  float mis = power_heuristic(eff * (float)sample_count, pdf2);
  float mw = mixf(m.tp.w, 1.0f, m.em.w);
  s.c = fdiv3((((eval_pbr(m, n, -incoming, dir) * environment_radiance(sc, dir)) * cosine) * mis) *
                  mk3(1.0f, 1.0f, 1.0f), fmaxf(eff, 1e-6f)) * mw;
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


}  // namespace olshade
