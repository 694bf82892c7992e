// ReSTIR DI at primary hits: reservoirs of area-light samples.
//
// Restates the reference's reservoir arithmetic (shaders/wavefront_restir.glsl
// :51-180) and area-light candidates (shaders/transport_v1/lighting.glsl
// :441-552, helpers in shaders/ordinaryshade_primary.glsl:672-702, 825-828):
// * a reservoir is (light index, barycentrics as half2, weight sum,
//   (target, represented samples) as half2);
// * candidates pick a light by stratified inversion of the emission CDF and a
//   uniform point on it (sqrt(u) barycentrics); the target is the luminance
//   of the MIS-weighted next-event contribution, and the RIS weight is the
//   target itself (the contribution is already divided by the source pdf);
// * stored reservoirs pack (light | count << 25, half2 barycentrics,
//   half2 (weight sum, target));
// * the radiance estimate is contribution * visibility * weight_sum /
//   (count * target).
#pragma once

#include <hip/hip_fp16.h>

#include "shade.h"

namespace olshade {

struct Reservoir {
  uint32_t light;  // OL_NONE: empty
  uint32_t bary;   // half2 barycentrics (b, c weights)
  float w_sum;
  uint32_t tc;     // half2 (target, represented samples)
};

OL_DEV uint32_t pack_half2(float a, float b) {
  const uint32_t lo = __half_as_ushort(__float2half_rn(a));
  const uint32_t hi = __half_as_ushort(__float2half_rn(b));
  return lo | (hi << 16);
}
OL_DEV float half_lo(uint32_t v) { return __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu))); }
OL_DEV float half_hi(uint32_t v) { return __half2float(__ushort_as_half((unsigned short)(v >> 16))); }

OL_DEV Reservoir empty_reservoir() { return Reservoir{OL_NONE, 0u, 0.0f, 0u}; }

// updateDirectLightReservoir
OL_DEV bool reservoir_update(Reservoir& r, uint32_t light, float b1, float b2, float target, float weight,
                             float represented, float random) {
  weight = fmaxf(weight, 0.0f);
  const float w_sum = r.w_sum + weight;
  float t = half_lo(r.tc);
  const float count = half_hi(r.tc) + represented;
  const bool selected = weight > 0.0f && (r.light == OL_NONE || random * w_sum < weight);
  if (selected) {
    r.light = light;
    r.bary = pack_half2(fminf(fmaxf(b1, 0.0f), 1.0f), fminf(fmaxf(b2, 0.0f), 1.0f));
    t = fminf(fmaxf(target, 0.0f), 65504.0f);
  }
  r.w_sum = w_sum;
  r.tc = pack_half2(t, count);
  return selected;
}

// mergeDirectLightReservoir (temporal reuse)
OL_DEV bool reservoir_merge(Reservoir& dst, const Reservoir& src, float target_here, float random) {
  const float st = half_lo(src.tc), sc = half_hi(src.tc);
  if (src.light == OL_NONE || sc <= 0.0f) return false;
  float weight = 0.0f;
  if (st > 0.0f) weight = (fmaxf(target_here, 0.0f) * src.w_sum) / st;
  return reservoir_update(dst, src.light, half_lo(src.bary), half_hi(src.bary), target_here, weight, sc, random);
}

// mergeCanonicalDirectLightReservoir (spatial reuse without pairwise MIS)
OL_DEV bool reservoir_merge_canonical(Reservoir& dst, const Reservoir& src, float target_here, float random) {
  const float st = half_lo(src.tc), sc = half_hi(src.tc);
  if (src.light == OL_NONE || st <= 0.0f || sc <= 0.0f) return false;
  const float weight = (fmaxf(target_here, 0.0f) * src.w_sum) / (st * sc);
  return reservoir_update(dst, src.light, half_lo(src.bary), half_hi(src.bary), target_here, weight, 1.0f, random);
}

// mergePairwiseDirectLightReservoir (spatial reuse)
OL_DEV bool reservoir_merge_pairwise(Reservoir& dst, const Reservoir& src, float target_here, float target_there,
                                     float random) {
  const float st = half_lo(src.tc), sc = half_hi(src.tc);
  if (src.light == OL_NONE || st <= 0.0f || sc <= 0.0f) return false;
  const float total = fmaxf(target_here, 0.0f) + fmaxf(target_there, 0.0f);
  float balance = 0.0f;
  if (total > 0.0f) balance = (2.0f * fmaxf(target_there, 0.0f)) / total;
  const float weight = ((fmaxf(target_here, 0.0f) * src.w_sum) / (st * sc)) * balance;
  return reservoir_update(dst, src.light, half_lo(src.bary), half_hi(src.bary), target_here, weight, 1.0f, random);
}

// directLightReservoirNormalization
OL_DEV float reservoir_normalization(const Reservoir& r) {
  const float t = half_lo(r.tc), c = half_hi(r.tc);
  if (r.light == OL_NONE || t <= 0.0f || c <= 0.0f) return 0.0f;
  return r.w_sum / (c * t);
}

// limitDirectLightReservoir
OL_DEV Reservoir reservoir_limit(Reservoir r, float maximum) {
  const float t = half_lo(r.tc), c = half_hi(r.tc);
  maximum = fmaxf(maximum, 0.0f);
  if (c > maximum && c > 0.0f) {
    const float scale = maximum / c;
    r.w_sum = r.w_sum * scale;
    r.tc = pack_half2(t, maximum);
  }
  return r;
}

// storeCurrentDirectLightReservoir / unpackStoredDirectLightReservoir:
// (light | count << 25, half2 barycentrics, half2 (weight sum, target)).
OL_DEV uint4 reservoir_store(const Reservoir& r) {
  const float t = half_lo(r.tc), c = half_hi(r.tc);
  uint32_t light = 33554431u;
  if (r.light != OL_NONE) light = min(r.light, 33554430u);
  const uint32_t count = (uint32_t)fminf(fmaxf(rintf(c), 0.0f), 127.0f);
  const float weight = fminf(fmaxf(r.w_sum, 0.0f), 65504.0f);
  return make_uint4(light | (count << 25), r.bary, pack_half2(weight, t), 0u);
}
OL_DEV Reservoir reservoir_load(uint4 w) {
  uint32_t light = w.x & 33554431u;
  const uint32_t count = w.x >> 25;
  if (light == 33554431u) light = OL_NONE;
  return Reservoir{light, w.y, half_lo(w.z), pack_half2(half_hi(w.z), (float)count)};
}

// evaluateAreaLightCandidate: MIS-weighted next-event contribution of the
// light point at `bary` on area light `light` (technique probability 1).
OL_DEV f3 candidate_contribution(const DScene& sc, uint32_t light, float b1, float b2, f3 hit, f3 n, f3 incoming,
                                 const Mat& m, uint32_t sample_count, f3& dir, float& dist) {
  dir = n;
  dist = 0.0f;
  if (light >= sc.area_count) return mk3(0.0f, 0.0f, 0.0f);
  const float4* L = sc.area + 5u * light;
  const f3 a = rgb(L[0]), b = rgb(L[1]), c = rgb(L[2]);
  const float aw = fmaxf((1.0f - b1) - b2, 0.0f);
  const f3 p = ((a * aw) + (b * b1)) + (c * b2);
  const f3 off = p - hit;
  const float d2 = dot3(off, off);
  dist = sqrtf(d2);
  dir = off / fmaxf(dist, 1e-6f);
  const float cs = fmaxf(dot3(n, dir), 0.0f);
  const f3 ln = normalize3(cross3(b - a, c - a));
  const float raw = dot3(ln, -dir);
  const float cl = L[4].z > 0.5f ? fabsf(raw) : fmaxf(raw, 0.0f);
  if (cs <= 0.0f || cl <= 1e-6f) return mk3(0.0f, 0.0f, 0.0f);
  const float eff = ((L[4].y * d2) / fmaxf(cl * L[3].w, 1e-6f)) * fmaxf(1.0f, 1e-6f);
  const float bp = pdf_pbr(m, n, -incoming, dir);
  const float first = eff * (float)sample_count;
  const float fs = first * first;
  const float mis = fs / fmaxf(fs + bp * bp, 1e-6f);
  return ((((eval_pbr(m, n, -incoming, dir) * rgb(L[3])) * cs) * mis) * 1.0f) / fmaxf(eff, 1e-6f);
}

OL_DEV float candidate_target(f3 c) { return fmaxf(dot3(c, mk3(0.2126f, 0.7152f, 0.0722f)), 0.0f); }

// generateAreaLightCandidate: stratified selection (index + u) / count over
// the emission CDF, sqrt(u) barycentrics; the path's PCG stream supplies u.
OL_DEV void candidate_generate(const DScene& sc, uint32_t& rs, f3 hit, f3 n, f3 incoming, const Mat& m, uint32_t k,
                               uint32_t count, uint32_t& light, float& b1, float& b2, float& target) {
  light = OL_NONE;
  b1 = b2 = 0.0f;
  target = 0.0f;
  if (sc.area_count == 0u) return;
  rs = pcg_step(rs);
  const float selection = ((float)k + pcg_value(rs)) / (float)count;
  const uint32_t lower = select_area_light(sc, selection);
  rs = pcg_step(rs);
  const float root_u = sqrtf(pcg_value(rs));
  rs = pcg_step(rs);
  const float v = pcg_value(rs);
  light = lower;
  b1 = root_u * (1.0f - v);
  b2 = root_u * v;
  f3 dir;
  float dist;
  target = candidate_target(candidate_contribution(sc, light, b1, b2, hit, n, incoming, m, count, dir, dist));
}

}  // namespace olshade
