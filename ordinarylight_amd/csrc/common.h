// Shared device helpers for the olhip kernels (gfx950, wave64).
//
// All kernels are compiled with -ffp-contract=off: every float expression
// below rounds exactly as written, left to right, so the oracle (a separate
// C restatement) reproduces each value bit for bit.  Where a fused
// multiply-add is wanted (BVH box decode, slab tests) it is spelled out with
// __builtin_fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OL_DEV __device__ __forceinline__
#define OL_HD __host__ __device__ __forceinline__

constexpr uint32_t OL_NONE = 0xFFFFFFFFu;
constexpr float OL_PI = 3.14159265359f;       // constant of the reference GLSL
constexpr float OL_TWO_PI = 6.28318530718f;   // constant of the reference GLSL

struct f3 {
  float x, y, z;
};

OL_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
OL_HD f3 mk3(float4 v) { return f3{v.x, v.y, v.z}; }
OL_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
OL_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
OL_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
OL_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
OL_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
OL_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
OL_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
OL_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
OL_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OL_HD f3 cross3(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
OL_HD float length3(f3 a) { return sqrtf(dot3(a, a)); }
// normalize(v) := v * (1 / sqrt(dot(v, v))), correctly rounded sqrt and divide.
OL_HD f3 normalize3(f3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }
OL_HD f3 max3(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
OL_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
OL_HD float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
// GLSL mix(a, b, t) = a * (1 - t) + b * t.
OL_HD float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
OL_HD f3 mix3(f3 a, f3 b, float t) {
  return f3{a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t};
}
OL_HD f3 reflect3(f3 d, f3 n) { return d - (2.0f * dot3(n, d)) * n; }
OL_HD float4 mk4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// -- deterministic sine/cosine -------------------------------------------
// Cephes single-precision sinf/cosf method (S. Moshier, public domain): reduce
// by pi/4 octants with a three-part Cody-Waite constant, then the Cephes
// minimax polynomials on [-pi/4, pi/4].  Used instead of the device library
// so that the oracle, which implements the same published method, agrees
// bit for bit.  Valid for |x| < 8192 (all callers pass |x| <= 2*pi).
OL_HD void ol_sincos(float x, float* s_out, float* c_out) {
  float sign_s = 1.0f;
  if (x < 0.0f) { x = -x; sign_s = -1.0f; }
  const float FOPI = 1.27323954473516f;  // 4/pi
  int j = (int)(x * FOPI);
  float y = (float)j;
  if (j & 1) { j += 1; y += 1.0f; }
  j &= 7;
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
  float z = ((x - y * DP1) - y * DP2) - y * DP3;
  float zz = z * z;
  float ps = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
  float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz - 0.5f * zz + 1.0f;
  float sign_c = 1.0f;
  if (j > 3) { j -= 4; sign_s = -sign_s; sign_c = -sign_c; }
  if (j > 1) sign_c = -sign_c;
  float s, c;
  if (j == 1 || j == 2) { s = pc; c = ps; } else { s = ps; c = pc; }
  *s_out = sign_s * s;
  *c_out = sign_c * c;
}

// -- hashes and random numbers ---------------------------------------------
// waveHash: wavefront_generate.comp:71-79 (also shadeSecondaryNeeHash).
OL_HD uint32_t wave_hash(uint32_t v) {
  v = v ^ (v >> 16);
  v = v * 2146121005u;
  v = v ^ (v >> 15);
  v = v * 2221713035u;
  v = v ^ (v >> 16);
  return v;
}
// float(uint) rounds to nearest; the product with 2^-32 is exact (may be 1.0).
OL_HD float u32_to_unit(uint32_t v) { return (float)v * (1.0f / 4294967296.0f); }
// rayQueryRandomState / rayQueryRandomValue: PCG LCG step and RXS-M-XS
// output (Jarzynski & Olano 2020), wavefront_shade_candidate.glsl:661-671.
OL_HD uint32_t pcg_step(uint32_t s) { return s * 747796405u + 2891336453u; }
OL_HD float pcg_value(uint32_t s) {
  uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
  w = (w >> 22) ^ w;
  return u32_to_unit(w);
}
OL_HD uint32_t bit_reverse32(uint32_t v) {
  v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
  v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
  v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
  v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
  return (v >> 16) | (v << 16);
}

// -- wave64 helpers ------------------------------------------------------------
OL_DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
OL_DEV uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// Reserve one slot per lane whose `want` is true with one atomic per wave.
// Must be reached by every active lane of the wave.
OL_DEV uint32_t wave_reserve(uint32_t* counter, bool want) {
  uint64_t mask = __ballot(want);
  uint64_t active = __ballot(true);
  uint32_t leader = (uint32_t)__ffsll((unsigned long long)active) - 1u;
  uint32_t base = 0;
  if (lane_id() == leader && mask != 0) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader, 64);
  return base + lanes_below(mask);
}
// Shading's three reservations at once: one slot per lane with `a` (counter
// ca), one per lane with `b` (cb) and `n` consecutive slots per lane (cn).
// The leader issues the three atomics back to back, so their round trips
// overlap instead of following each other.  Must be reached by every active
// lane of the wave.
struct Reserve3 {
  uint32_t a, b, n;
};
OL_DEV Reserve3 wave_reserve3(uint32_t* ca, bool a, uint32_t* cb, bool b, uint32_t* cn, uint32_t n) {
  const uint64_t ma = __ballot(a), mb = __ballot(b);
  uint32_t incl = n;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t v = __shfl_up(incl, (unsigned)off, 64);
    if ((int)lane_id() >= off) incl += v;
  }
  const uint64_t active = __ballot(true);
  const uint32_t last = 63u - (uint32_t)__clzll((long long)active);
  const uint32_t total = __shfl(incl, (int)last, 64);
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)active) - 1u;
  uint32_t base_a = 0, base_b = 0, base_n = 0;
  if (lane_id() == leader) {
    if (ma != 0) base_a = atomicAdd(ca, (uint32_t)__popcll(ma));
    if (mb != 0) base_b = atomicAdd(cb, (uint32_t)__popcll(mb));
    if (total != 0) base_n = atomicAdd(cn, total);
  }
  Reserve3 r;
  r.a = __shfl(base_a, (int)leader, 64) + lanes_below(ma);
  r.b = __shfl(base_b, (int)leader, 64) + lanes_below(mb);
  r.n = __shfl(base_n, (int)leader, 64) + incl - n;
  return r;
}
// Adds the number of active lanes to a 64-bit counter with one atomic per
// wave (work counters).
OL_DEV void count_active(unsigned long long* counter) {
  const uint64_t active = __ballot(true);
  if (lanes_below(active) == 0u) atomicAdd(counter, (unsigned long long)__popcll(active));
}
// Reserve `n` consecutive slots per lane (wave prefix sum, one atomic per wave).
OL_DEV uint32_t wave_reserve_n(uint32_t* counter, uint32_t n) {
  uint32_t incl = n;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t v = __shfl_up(incl, (unsigned)off, 64);
    if ((int)lane_id() >= off) incl += v;
  }
  uint64_t active = __ballot(true);
  uint32_t last = 63u - (uint32_t)__clzll((long long)active);
  uint32_t total = __shfl(incl, (int)last, 64);
  uint32_t leader = (uint32_t)__ffsll((unsigned long long)active) - 1u;
  uint32_t base = 0;
  if (lane_id() == leader && total != 0) base = atomicAdd(counter, total);
  base = __shfl(base, (int)leader, 64);
  return base + incl - n;
}

// -- scene records ---------------------------------------------------------------
// 4-wide BVH node, 64 bytes (one cache-line-aligned record, four dwordx4 loads).
// Child boxes are 8-bit offsets from `origin` in units of 2^(e-127) per axis,
// rounded outward at build time against the exact float decode below.
struct alignas(64) WideNode {
  float ox, oy, oz;
  uint32_t exps;      // ex | ey << 8 | ez << 16 (biased by 127)
  uint32_t qlo[3];    // byte c = child c, per axis x/y/z
  uint32_t qhi[3];
  uint32_t child[4];  // OL_NONE | leaf: 0x80000000 | (count-1) << 27 | first | inner index
  uint32_t pad[2];
};
static_assert(sizeof(WideNode) == 64, "WideNode must be 64 bytes");

constexpr uint32_t OL_LEAF_BIT = 0x80000000u;
constexpr uint32_t OL_LEAF_MAX = 16u;
constexpr uint32_t OL_LEAF_FIRST_MASK = (1u << 27) - 1u;

OL_HD bool child_is_leaf(uint32_t c) { return (c & OL_LEAF_BIT) != 0u && c != OL_NONE; }
OL_HD uint32_t leaf_first(uint32_t c) { return c & OL_LEAF_FIRST_MASK; }
OL_HD uint32_t leaf_count(uint32_t c) { return ((c >> 27) & 15u) + 1u; }
OL_HD float exp_scale(uint32_t biased) {
  union { uint32_t u; float f; } v;
  v.u = biased << 23;
  return v.f;
}
