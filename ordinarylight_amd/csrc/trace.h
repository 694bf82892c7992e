// Software ray traversal of the 64-byte 4-wide BVH (closest hit, any hit).
//
// One lane per ray (wave64).  The traversal stack is two-level: the top
// LSTACK entries live in LDS, one column per lane (entry k of lane t at
// lds[k * BLOCK + t], bank-conflict free), deeper entries spill to a per-lane
// column of a global buffer.  The total depth is chosen on the host from the
// BVH's level count (a 4-wide level pushes at most three entries), so it
// cannot overflow; the small LDS part keeps occupancy bounded by registers,
// not LDS.
//
// Closest-hit semantics are those of a ray query over opaque triangles
// (native_intersection.glsl:44-183): candidates with tmin <= t <= tmax; the
// nearest wins and equal distances resolve to the lowest original triangle
// index, which makes the result independent of traversal order and equal to
// a brute-force scan (the oracle's check).
#pragma once

#include "common.h"

struct TraceHit {
  float t, u, v;
  uint32_t slot;  // leaf-order triangle index, OL_NONE on a miss
  uint32_t orig;  // original triangle index
};

// Moller-Trumbore in float32, evaluated exactly as written (no contraction).
OL_DEV bool ol_tri_test(f3 o, f3 d, float4 a, float4 e1, float4 e2, float& t, float& u, float& v) {
  f3 E1 = mk3(e1), E2 = mk3(e2);
  f3 p = cross3(d, E2);
  float det = dot3(E1, p);
  if (det == 0.0f) return false;
  float inv = 1.0f / det;
  f3 s = o - mk3(a);
  u = dot3(s, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  f3 q = cross3(s, E1);
  v = dot3(d, q) * inv;
  if (!(v >= 0.0f && u + v <= 1.0f)) return false;
  t = dot3(E2, q) * inv;
  return true;
}

struct RayPre {
  f3 o, d, inv;
};

OL_DEV RayPre ray_pre(f3 o, f3 d) {
  auto safe = [](float x) { return fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x); };
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = mk3(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
  return r;
}

// Slab test of the four quantised child boxes.  Each plane is decoded
// exactly as the build rounded it (origin + q * 2^e, one fma, the product is
// exact), then t = (plane - o) * inv: two roundings relative to t, so
// widening tfar by 2^-20 keeps every box that contains a candidate hit.
// Returns a 4-bit hit mask and entry distances.
OL_DEV uint32_t ol_node_hits(const uint4 w0, const uint4 w1, const uint4 w2, const uint4 w3, const RayPre& r,
                             float tmin, float tmax, float tnear[4], uint32_t child[4]) {
  const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
  const float sx = exp_scale(w0.w & 255u), sy = exp_scale((w0.w >> 8) & 255u), sz = exp_scale((w0.w >> 16) & 255u);
  const uint32_t qlx = w1.x, qly = w1.y, qlz = w1.z, qhx = w1.w, qhy = w2.x, qhz = w2.y;
  child[0] = w2.z;
  child[1] = w2.w;
  child[2] = w3.x;
  child[3] = w3.y;
  uint32_t mask = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int sh = 8 * c;
    float lx = (__builtin_fmaf((float)((qlx >> sh) & 255u), sx, ox) - r.o.x) * r.inv.x;
    float hx = (__builtin_fmaf((float)((qhx >> sh) & 255u), sx, ox) - r.o.x) * r.inv.x;
    float ly = (__builtin_fmaf((float)((qly >> sh) & 255u), sy, oy) - r.o.y) * r.inv.y;
    float hy = (__builtin_fmaf((float)((qhy >> sh) & 255u), sy, oy) - r.o.y) * r.inv.y;
    float lz = (__builtin_fmaf((float)((qlz >> sh) & 255u), sz, oz) - r.o.z) * r.inv.z;
    float hz = (__builtin_fmaf((float)((qhz >> sh) & 255u), sz, oz) - r.o.z) * r.inv.z;
    float tn = fmaxf(fmaxf(fminf(lx, hx), fminf(ly, hy)), fmaxf(fminf(lz, hz), tmin));
    float tf = fminf(fminf(fmaxf(lx, hx), fmaxf(ly, hy)), fminf(fmaxf(lz, hz), tmax));
    tnear[c] = tn;
    bool hit = child[c] != OL_NONE && tn <= tf * 1.00000095367431640625f;
    mask |= hit ? (1u << c) : 0u;
  }
  return mask;
}

// Culling distance for a current best hit at t: boxes entered beyond it are
// skipped.  The 2^-16 relative slack absorbs the rounding of the
// Moller-Trumbore distance itself, so a triangle that ties the best hit
// (equal float t, lower index) is never culled.
OL_DEV float cull_distance(float t) { return t * 1.0000152587890625f; }

OL_DEV void load_node(const WideNode* nodes, uint32_t index, uint4& w0, uint4& w1, uint4& w2, uint4& w3) {
  const uint4* p = reinterpret_cast<const uint4*>(nodes + index);
  w0 = p[0];
  w1 = p[1];
  w2 = p[2];
  w3 = p[3];
}

template <int LSTACK>
struct TravStack {
  uint32_t* lds;       // this lane's LDS column
  uint32_t lds_stride; // block size
  uint32_t* spill;     // this lane's global column (may be null when spill_cap == 0)
  uint32_t spill_stride;
  uint32_t spill_cap;
  uint32_t* overflow;
  uint32_t sp;

  OL_DEV void push(uint32_t v) {
    if (sp < (uint32_t)LSTACK) {
      lds[sp * lds_stride] = v;
    } else if (sp - LSTACK < spill_cap) {
      spill[(sp - LSTACK) * spill_stride] = v;
    } else {
      atomicOr(overflow, 1u);
      return;
    }
    ++sp;
  }
  OL_DEV uint32_t pop() {
    --sp;
    return sp < (uint32_t)LSTACK ? lds[sp * lds_stride] : spill[(sp - LSTACK) * spill_stride];
  }
  OL_DEV bool empty() const { return sp == 0; }
};

// Closest hit.
// nv / nt accumulate nodes visited and triangles tested (work counters).
template <int LSTACK>
OL_DEV TraceHit ol_trace_closest(const WideNode* __restrict__ nodes, const float4* __restrict__ tri,
                                 uint32_t node_count, f3 o, f3 d, float tmin, float tmax, TravStack<LSTACK>& st,
                                 uint32_t& nv, uint32_t& nt) {
  TraceHit best{tmax, 0.0f, 0.0f, OL_NONE, OL_NONE};
  if (node_count == 0) return best;
  const RayPre r = ray_pre(o, d);
  st.sp = 0;
  uint32_t node = 0;
  while (true) {
    uint4 w0, w1, w2, w3;
    load_node(nodes, node, w0, w1, w2, w3);
    ++nv;
    float tn[4];
    uint32_t ch[4];
    // Cull against the current best distance widened like the slab test.
    uint32_t mask = ol_node_hits(w0, w1, w2, w3, r, tmin, cull_distance(best.t), tn, ch);
    // Leaves first (order does not matter for the lexicographic minimum).
    uint32_t inner_mask = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!(mask & (1u << c))) continue;
      if (!child_is_leaf(ch[c])) {
        inner_mask |= 1u << c;
        continue;
      }
      const uint32_t first = leaf_first(ch[c]), cnt = leaf_count(ch[c]);
      nt += cnt;
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t slot = first + k;
        const float4 a = tri[3 * slot], e1 = tri[3 * slot + 1], e2 = tri[3 * slot + 2];
        float t, u, v;
        if (ol_tri_test(o, d, a, e1, e2, t, u, v) && t >= tmin) {
          const uint32_t orig = __float_as_uint(a.w);
          if (t < best.t || (t == best.t && orig < best.orig)) best = TraceHit{t, u, v, slot, orig};
        }
      }
    }
    // Inner children nearest first: push the others far to near.
    float dist[4];
    uint32_t idx[4];
    int n = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (inner_mask & (1u << c)) {
        dist[n] = tn[c];
        idx[n] = ch[c];
        ++n;
      }
    }
    if (n == 0) {
      if (st.empty()) break;
      node = st.pop();
      continue;
    }
    // Insertion sort of up to four (distance, node) pairs, ascending.
    for (int i = 1; i < n; ++i) {
      float dv = dist[i];
      uint32_t iv = idx[i];
      int j = i - 1;
      while (j >= 0 && dist[j] > dv) {
        dist[j + 1] = dist[j];
        idx[j + 1] = idx[j];
        --j;
      }
      dist[j + 1] = dv;
      idx[j + 1] = iv;
    }
    for (int i = n - 1; i >= 1; --i) st.push(idx[i]);
    node = idx[0];
  }
  return best;
}

// Any hit in [tmin, tmax]: returns true as soon as one triangle qualifies.
template <int LSTACK>
OL_DEV bool ol_trace_any(const WideNode* __restrict__ nodes, const float4* __restrict__ tri, uint32_t node_count,
                         f3 o, f3 d, float tmin, float tmax, TravStack<LSTACK>& st, uint32_t& nv, uint32_t& nt) {
  if (node_count == 0) return false;
  const RayPre r = ray_pre(o, d);
  const float tcull = cull_distance(tmax);
  st.sp = 0;
  uint32_t node = 0;
  while (true) {
    uint4 w0, w1, w2, w3;
    load_node(nodes, node, w0, w1, w2, w3);
    ++nv;
    float tn[4];
    uint32_t ch[4];
    uint32_t mask = ol_node_hits(w0, w1, w2, w3, r, tmin, tcull, tn, ch);
    uint32_t next = OL_NONE;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (!(mask & (1u << c))) continue;
      if (child_is_leaf(ch[c])) {
        const uint32_t first = leaf_first(ch[c]), cnt = leaf_count(ch[c]);
        for (uint32_t k = 0; k < cnt; ++k) {
          const uint32_t slot = first + k;
          const float4 a = tri[3 * slot], e1 = tri[3 * slot + 1], e2 = tri[3 * slot + 2];
          float t, u, v;
          ++nt;
          if (ol_tri_test(o, d, a, e1, e2, t, u, v) && t >= tmin && t <= tmax) return true;
        }
      } else if (next == OL_NONE) {
        next = ch[c];
      } else {
        st.push(ch[c]);
      }
    }
    if (next != OL_NONE) {
      node = next;
      continue;
    }
    if (st.empty()) return false;
    node = st.pop();
  }
}
