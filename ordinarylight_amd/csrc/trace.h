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
  uint32_t neg;  // bit a set when the direction's axis-a component is negative
};

OL_DEV RayPre ray_pre(f3 o, f3 d) {
  auto safe = [](float x) { return fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x); };
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = mk3(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
  r.neg = (r.inv.x < 0.0f ? 1u : 0u) | (r.inv.y < 0.0f ? 2u : 0u) | (r.inv.z < 0.0f ? 4u : 0u);
  return r;
}

// The same constants for a ray entering an instance, with reciprocals from
// v_rcp_f32 (within 1 ulp) instead of correctly rounded divisions: the node
// boxes' build-time growth (2^-17 of the largest root coordinate) is far
// wider than the extra half ulp, and the triangle test does not use them,
// so hits do not change.
OL_DEV RayPre ray_pre_fast(f3 o, f3 d) {
  auto safe = [](float x) { return fabsf(x) > 1e-30f ? x : copysignf(1e-30f, x); };
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = mk3(__builtin_amdgcn_rcpf(safe(d.x)), __builtin_amdgcn_rcpf(safe(d.y)), __builtin_amdgcn_rcpf(safe(d.z)));
  r.neg = (r.inv.x < 0.0f ? 1u : 0u) | (r.inv.y < 0.0f ? 2u : 0u) | (r.inv.z < 0.0f ? 4u : 0u);
  return r;
}

OL_DEV float byte_f(uint32_t w, int c) { return (float)((w >> (8 * c)) & 255u); }

// Slab test of the four quantised child boxes of a node.
//
// Plane q of axis a lies at origin_a + q * 2^e_a; its ray distance is taken
// as fma(q, 2^e_a * inv_a, (origin_a - o_a) * inv_a), three float32 roundings
// of quantities no larger than the scene extent plus the origin distance.
// The near plane of each axis is the low plane when the direction is
// positive and the high plane otherwise, so each child costs six fma and
// four min/max.  Every box was grown at build time by 2^-17 of the largest
// scene coordinate, far more (in space) than these roundings, so a box that
// contains a candidate hit is never culled.
//
// Returns one sort key per child: the entry distance's bits with the low two
// bits replaced by the child's position, or OL_NONE when the child is empty
// or missed.  Entry distances are >= tmin >= 0, so unsigned key order is
// distance order.
OL_DEV void ol_node_keys(const uint4 w0, const uint4 w1, const uint4 w2, const RayPre& r, float tmin, float tmax,
                         uint32_t key[4], const uint32_t child[4]) {
  const float ax = exp_scale(w0.w & 255u) * r.inv.x, bx = (__uint_as_float(w0.x) - r.o.x) * r.inv.x;
  const float ay = exp_scale((w0.w >> 8) & 255u) * r.inv.y, by = (__uint_as_float(w0.y) - r.o.y) * r.inv.y;
  const float az = exp_scale((w0.w >> 16) & 255u) * r.inv.z, bz = (__uint_as_float(w0.z) - r.o.z) * r.inv.z;
  const uint32_t nx = (r.neg & 1u) ? w1.w : w1.x, fx = (r.neg & 1u) ? w1.x : w1.w;
  const uint32_t ny = (r.neg & 2u) ? w2.x : w1.y, fy = (r.neg & 2u) ? w1.y : w2.x;
  const uint32_t nz = (r.neg & 4u) ? w2.y : w1.z, fz = (r.neg & 4u) ? w1.z : w2.y;
  const float t0 = fmaxf(tmin, 0.0f);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float tnx = __builtin_fmaf(byte_f(nx, c), ax, bx), tfx = __builtin_fmaf(byte_f(fx, c), ax, bx);
    const float tny = __builtin_fmaf(byte_f(ny, c), ay, by), tfy = __builtin_fmaf(byte_f(fy, c), ay, by);
    const float tnz = __builtin_fmaf(byte_f(nz, c), az, bz), tfz = __builtin_fmaf(byte_f(fz, c), az, bz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, t0));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
    const bool hit = child[c] != OL_NONE && tn <= tf;
    key[c] = hit ? ((__float_as_uint(tn) & ~3u) | (uint32_t)c) : OL_NONE;
  }
}

OL_DEV void key_swap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}
// Ascending sort of four keys (five compare-exchanges).
OL_DEV void sort4(uint32_t k[4]) {
  key_swap(k[0], k[1]);
  key_swap(k[2], k[3]);
  key_swap(k[0], k[2]);
  key_swap(k[1], k[3]);
  key_swap(k[1], k[2]);
}
OL_DEV float key_distance(uint32_t key) { return __uint_as_float(key & ~3u); }

// Ascending sort of four keys carrying their child words along (five
// compare-exchanges of pairs), so no child needs selecting by key afterwards.
OL_DEV void pair_swap(uint32_t& ka, uint32_t& kb, uint32_t& ca, uint32_t& cb) {
  const bool s = kb < ka;
  const uint32_t k0 = s ? kb : ka, k1 = s ? ka : kb, c0 = s ? cb : ca, c1 = s ? ca : cb;
  ka = k0;
  kb = k1;
  ca = c0;
  cb = c1;
}
OL_DEV void sort4_pairs(uint32_t k[4], uint32_t c[4]) {
  pair_swap(k[0], k[1], c[0], c[1]);
  pair_swap(k[2], k[3], c[2], c[3]);
  pair_swap(k[0], k[2], c[0], c[2]);
  pair_swap(k[1], k[3], c[1], c[3]);
  pair_swap(k[1], k[2], c[1], c[2]);
}

// Culling distance for a current best hit at t: boxes entered beyond it are
// skipped.  The 2^-16 relative slack absorbs the rounding of the
// Moller-Trumbore distance itself, so a triangle that ties the best hit
// (equal float t, lower index) is never culled.
OL_DEV float cull_distance(float t) { return t * 1.0000152587890625f; }

// Pointers with their address space spelled out, so that stack, node and
// triangle accesses compile to LDS (ds_*) and global (global_*) instructions
// rather than flat ones, which also wait on the other memory's counter.
typedef __attribute__((address_space(3))) uint32_t lds_word;
typedef __attribute__((address_space(1))) uint32_t global_word;
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4_t global_u32x4;
typedef __attribute__((address_space(1))) const f32x4_t global_f32x4;

OL_DEV void load_node(const WideNode* nodes, uint32_t index, uint4& w0, uint4& w1, uint4& w2, uint4& w3) {
  global_u32x4* p = (global_u32x4*)(nodes + index);
  const u32x4_t a = p[0], b = p[1], c = p[2], d = p[3];
  w0 = make_uint4(a.x, a.y, a.z, a.w);
  w1 = make_uint4(b.x, b.y, b.z, b.w);
  w2 = make_uint4(c.x, c.y, c.z, c.w);
  w3 = make_uint4(d.x, d.y, d.z, d.w);
}

OL_DEV void load_tri(const float4* tri, uint32_t slot, float4& a, float4& e1, float4& e2) {
  global_f32x4* p = (global_f32x4*)(tri + 3u * slot);
  const f32x4_t x = p[0], y = p[1], z = p[2];
  a = make_float4(x.x, x.y, x.z, x.w);
  e1 = make_float4(y.x, y.y, y.z, y.w);
  e2 = make_float4(z.x, z.y, z.z, z.w);
}

template <int LSTACK>
struct TravStack {
  lds_word* lds;        // this lane's LDS column
  uint32_t lds_stride;  // block size
  global_word* spill;   // this lane's global column (may be null when spill_cap == 0)
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
  // Pushes m (0..3) entries: c3 (m = 3), c2 (m >= 2), c1 (m >= 1), in that
  // order, so c1 ends on top.  When three entries fit in LDS the lane
  // writes three words unconditionally (those above the new top are
  // scratch) instead of branching per entry.
  OL_DEV void push_upto3(uint32_t c1, uint32_t c2, uint32_t c3, uint32_t m) {
    if (sp + 3u <= (uint32_t)LSTACK) {
      const uint32_t v0 = m >= 3u ? c3 : (m == 2u ? c2 : c1);
      const uint32_t v1 = m >= 3u ? c2 : c1;
      lds[sp * lds_stride] = v0;
      lds[(sp + 1u) * lds_stride] = v1;
      lds[(sp + 2u) * lds_stride] = c1;
      sp += m;
    } else {
      if (m >= 3u) push(c3);
      if (m >= 2u) push(c2);
      if (m >= 1u) push(c1);
    }
  }
  OL_DEV uint32_t pop() {
    --sp;
    if (sp < (uint32_t)LSTACK) return lds[sp * lds_stride];
    return spill[(sp - LSTACK) * spill_stride];
  }
  OL_DEV bool empty() const { return sp == 0; }
  // The bottom entry (always in LDS), for the two-level traversal's
  // held-back instance.
  OL_DEV uint32_t bottom() const { return lds[0]; }
  OL_DEV void set_bottom(uint32_t v) { lds[0] = v; }
};

// Children of a node in child-word order.
OL_DEV void node_children(const uint4 w2, const uint4 w3, uint32_t ch[4]) {
  ch[0] = w2.z;
  ch[1] = w2.w;
  ch[2] = w3.x;
  ch[3] = w3.y;
}

// Closest-hit traversal state of one lane, advanced one node per step.
struct ClosestRay {
  RayPre r;
  float tmin;
  TraceHit best;
  uint32_t node;  // next node to visit
};

OL_DEV void closest_begin(ClosestRay& q, f3 o, f3 d, float tmin, float tmax) {
  q.r = ray_pre(o, d);
  q.tmin = tmin;
  q.best = TraceHit{tmax, 0.0f, 0.0f, OL_NONE, OL_NONE};
  q.node = 0;
}

// Visits q.node: tests its leaf children (near to far), continues with the
// nearest inner child and pushes the others far to near, dropping those
// that start beyond the culling distance.  Returns true when the traversal
// has finished (nothing left to visit).  nv / nt count nodes visited and
// triangles tested (work counters).
template <int LSTACK>
OL_DEV bool closest_step(const WideNode* __restrict__ nodes, const float4* __restrict__ tri, ClosestRay& q,
                         TravStack<LSTACK>& st, uint32_t& nv, uint32_t& nt) {
  uint4 w0, w1, w2, w3;
  load_node(nodes, q.node, w0, w1, w2, w3);
  ++nv;
  uint32_t ch[4], key[4];
  node_children(w2, w3, ch);
  // Cull against the current best distance widened like the slab test.
  ol_node_keys(w0, w1, w2, q.r, q.tmin, cull_distance(q.best.t), key, ch);
  sort4(key);
  // The lexicographic minimum does not depend on the order; nearer leaves
  // shrink the culling distance sooner.
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (key[k] == OL_NONE) continue;
    const uint32_t c = ch[key[k] & 3u];
    if (!child_is_leaf(c)) continue;
    const uint32_t first = leaf_first(c), cnt = leaf_count(c);
    nt += cnt;
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t slot = first + j;
      float4 a, e1, e2;
      load_tri(tri, slot, a, e1, e2);
      float t, u, v;
      if (ol_tri_test(q.r.o, q.r.d, a, e1, e2, t, u, v) && t >= q.tmin) {
        const uint32_t orig = __float_as_uint(a.w);
        if (t < q.best.t || (t == q.best.t && orig < q.best.orig)) q.best = TraceHit{t, u, v, slot, orig};
      }
    }
  }
  const float cull = cull_distance(q.best.t);
  uint32_t next = OL_NONE;
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    if (key[k] == OL_NONE) continue;
    const uint32_t c = ch[key[k] & 3u];
    if (child_is_leaf(c) || key_distance(key[k]) > cull) continue;
    if (next != OL_NONE) st.push(next);
    next = c;
  }
  if (next != OL_NONE) {
    q.node = next;
    return false;
  }
  if (st.empty()) return true;
  q.node = st.pop();
  return false;
}

// Any hit ends at the first occluder found, so its children need no
// distance order: they are taken in build order (the first hit child
// continues), which has measured faster for shadow rays than near to far.
// The keys become the hit children's positions, ascending, misses last.
OL_DEV void build_order(uint32_t key[4], uint32_t ch[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) key[k] = key[k] == OL_NONE ? OL_NONE : (uint32_t)k;
  sort4_pairs(key, ch);
}

// With the hit children first and in visiting order, keys and child words
// sorted together (sort4_pairs or build_order): continues with the first
// and pushes the others in reverse, so the second is on top; with no child
// hit, takes the next stack entry.
template <int LSTACK>
OL_DEV void descend(TravStack<LSTACK>& st, const uint32_t ch[4], const uint32_t key[4], uint32_t& node) {
  if (key[0] == OL_NONE) {
    node = st.empty() ? OL_NONE : st.pop();
    return;
  }
  const uint32_t m = (key[1] != OL_NONE ? 1u : 0u) + (key[2] != OL_NONE ? 1u : 0u) + (key[3] != OL_NONE ? 1u : 0u);
  st.push_upto3(ch[1], ch[2], ch[3], m);
  node = ch[0];
}

// Split-step form used by the refilling scheduler.  q.node holds the next
// item to visit: an inner node index, a leaf word (OL_LEAF_BIT set), or
// OL_NONE once the traversal has finished.  A wave chooses per step whether
// its lanes at inner nodes or its lanes at leaves advance, so the triangle
// tests of a leaf run with many lanes active instead of stalling the box
// tests of the others.
OL_DEV bool item_is_leaf(uint32_t item) { return child_is_leaf(item); }

template <int LSTACK>
OL_DEV void closest_next(ClosestRay& q, TravStack<LSTACK>& st) { q.node = st.empty() ? OL_NONE : st.pop(); }

// Inner step: continue with the nearest hit child (inner or leaf) and push the
// others far to near.
template <int LSTACK>
OL_DEV void closest_inner(const WideNode* __restrict__ nodes, ClosestRay& q, TravStack<LSTACK>& st, uint32_t& nv) {
  uint4 w0, w1, w2, w3;
  load_node(nodes, q.node, w0, w1, w2, w3);
  ++nv;
  uint32_t ch[4], key[4];
  node_children(w2, w3, ch);
  ol_node_keys(w0, w1, w2, q.r, q.tmin, cull_distance(q.best.t), key, ch);
  sort4_pairs(key, ch);
  descend(st, ch, key, q.node);
}

// Leaf step: test the leaf's triangles, then take the next stack entry.
template <int LSTACK>
OL_DEV void closest_leaf(const float4* __restrict__ tri, ClosestRay& q, TravStack<LSTACK>& st, uint32_t& nt) {
  const uint32_t first = leaf_first(q.node), cnt = leaf_count(q.node);
  nt += cnt;
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t slot = first + j;
    float4 a, e1, e2;
    load_tri(tri, slot, a, e1, e2);
    float t, u, v;
    if (ol_tri_test(q.r.o, q.r.d, a, e1, e2, t, u, v) && t >= q.tmin) {
      const uint32_t orig = __float_as_uint(a.w);
      if (t < q.best.t || (t == q.best.t && orig < q.best.orig)) q.best = TraceHit{t, u, v, slot, orig};
    }
  }
  closest_next(q, st);
}

// Any-hit traversal state of one lane in split-step form (see above).
// q.node is OL_NONE once finished; occluded tells whether a triangle in
// [tmin, tmax] was found.
struct AnyRay {
  RayPre r;
  float tmin, tmax, tcull;
  uint32_t node;
  bool occluded;
};

OL_DEV void any_begin(AnyRay& q, f3 o, f3 d, float tmin, float tmax) {
  q.r = ray_pre(o, d);
  q.tmin = tmin;
  q.tmax = tmax;
  q.tcull = cull_distance(tmax);
  q.node = 0;
  q.occluded = false;
}

template <int LSTACK>
OL_DEV void any_next(AnyRay& q, TravStack<LSTACK>& st) { q.node = st.empty() ? OL_NONE : st.pop(); }

template <int LSTACK>
OL_DEV void any_inner(const WideNode* __restrict__ nodes, AnyRay& q, TravStack<LSTACK>& st, uint32_t& nv) {
  uint4 w0, w1, w2, w3;
  load_node(nodes, q.node, w0, w1, w2, w3);
  ++nv;
  uint32_t ch[4], key[4];
  node_children(w2, w3, ch);
  ol_node_keys(w0, w1, w2, q.r, q.tmin, q.tcull, key, ch);
  build_order(key, ch);
  descend(st, ch, key, q.node);
}

template <int LSTACK>
OL_DEV void any_leaf(const float4* __restrict__ tri, AnyRay& q, TravStack<LSTACK>& st, uint32_t& nt) {
  const uint32_t first = leaf_first(q.node), cnt = leaf_count(q.node);
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t slot = first + j;
    float4 a, e1, e2;
    load_tri(tri, slot, a, e1, e2);
    float t, u, v;
    ++nt;
    if (ol_tri_test(q.r.o, q.r.d, a, e1, e2, t, u, v) && t >= q.tmin && t <= q.tmax) {
      q.occluded = true;
      q.node = OL_NONE;
      return;
    }
  }
  any_next(q, st);
}

// Closest hit of one ray, run to completion.
template <int LSTACK>
OL_DEV TraceHit ol_trace_closest(const WideNode* __restrict__ nodes, const float4* __restrict__ tri,
                                 uint32_t node_count, f3 o, f3 d, float tmin, float tmax, TravStack<LSTACK>& st,
                                 uint32_t& nv, uint32_t& nt) {
  ClosestRay q;
  closest_begin(q, o, d, tmin, tmax);
  if (node_count == 0) return q.best;
  st.sp = 0;
  while (!closest_step<LSTACK>(nodes, tri, q, st, nv, nt)) {
  }
  return q.best;
}

// ---- two-level traversal (top-level BVH over instances of bottom-level BVHs) -----------
// The stack is shared by both levels: entering an instance pushes
// OL_SENTINEL, and popping it returns to the top level.  The ray is
// transformed into object space without renormalising the direction, so
// distances along it are the world distances and hits of different
// instances compare directly.  Ties resolve to the lowest packed-order
// triangle index (instance triangle offset + local index), the order of the
// flattened scene.
//
// An instance flagged "later" (the sign bit of its sphere radius; set on the
// host for instances whose box spans a large part of the scene) is not
// entered when the top-level traversal reaches it, but once everything else
// has been visited, with the best hit of the rest of the scene known, so its
// nodes beyond that hit are culled (any hit: an occluder found elsewhere ends
// the ray without entering it).  A two-level traversal starts with OL_NONE
// pushed at the bottom of the stack (two_level_start), which ends it when
// popped; holding an instance back overwrites that bottom entry with the
// instance's leaf word, so it is popped last, at no cost to the other steps.
// One instance per ray is held back (the bottom entry reads OL_NONE only
// while the slot is free); another reached later is entered at once.  The
// visiting order does not change the result.
constexpr uint32_t OL_SENTINEL = 0x7FFFFFFFu;

OL_DEV bool instance_later(const float4& sphere) { return (__float_as_uint(sphere.w) >> 31) != 0u; }

template <int LSTACK>
OL_DEV void two_level_start(TravStack<LSTACK>& st) {
  st.sp = 0;
  st.push(OL_NONE);
}

// Holds the instance at leaf word `word` back when the slot is free.
template <int LSTACK>
OL_DEV bool hold_back(TravStack<LSTACK>& st, uint32_t word) {
  if (st.sp == 0 || st.bottom() != OL_NONE) return false;
  st.set_bottom(word);
  return true;
}

OL_DEV f3 xform_point(const float4* rows, f3 p) {
  return mk3(((rows[0].x * p.x + rows[0].y * p.y) + rows[0].z * p.z) + rows[0].w,
             ((rows[1].x * p.x + rows[1].y * p.y) + rows[1].z * p.z) + rows[1].w,
             ((rows[2].x * p.x + rows[2].y * p.y) + rows[2].z * p.z) + rows[2].w);
}
OL_DEV f3 xform_vector(const float4* rows, f3 v) {
  return mk3((rows[0].x * v.x + rows[0].y * v.y) + rows[0].z * v.z,
             (rows[1].x * v.x + rows[1].y * v.y) + rows[1].z * v.z,
             (rows[2].x * v.x + rows[2].y * v.y) + rows[2].z * v.z);
}

OL_DEV float max_abs3(f3 a) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z)); }

// Whether the ray's line passes within an instance's world bounding sphere
// (centre xyz, radius w); an instance whose sphere it misses cannot be hit.
// v is the offset of the centre from the computed closest point of the line:
// its rounding (a few ulps of the operands) and the closest point's shift
// along the line (which only lengthens v, by ~1e-7 of the magnitudes) stay
// far inside the 1e-5 allowance, so a reachable instance is never skipped.
OL_DEV bool sphere_reachable(const RayPre& r, float4 s) {
  const f3 c = mk3(s.x, s.y, s.z);
  const f3 oc = c - r.o;
  const float tc = dot3(oc, r.d) * __builtin_amdgcn_rcpf(dot3(r.d, r.d));
  const f3 v = oc - r.d * tc;
  const float scale = fmaxf(fmaxf(max_abs3(r.o), max_abs3(c)), fabsf(tc) * max_abs3(r.d));
  const float reach = fabsf(s.w) + 1e-5f * scale;
  return dot3(v, v) <= reach * reach;
}

// Whether the ray can meet an instance's world bounding sphere at a
// parameter t <= tmax: the line passes within the sphere (sphere_reachable)
// and its entry into the sphere, grown by the same allowance, lies no later
// than tmax plus a further allowance of 1e-4 of the magnitudes.  Every point
// of the instance lies in the sphere, so an instance failing the test holds
// no hit at t <= tmax: a closest-hit ray skips it once a nearer hit is known.
// The divisions and square roots use the hardware approximations (within an
// ulp or two): the allowances above are orders of magnitude wider.  The
// radius's sign bit is the "later" flag, not part of the radius.
OL_DEV bool sphere_before(const RayPre& r, float4 s, float tmax) {
  const f3 c = mk3(s.x, s.y, s.z);
  const f3 oc = c - r.o;
  const float dd = dot3(r.d, r.d);
  const float tc = dot3(oc, r.d) * __builtin_amdgcn_rcpf(dd);
  const f3 v = oc - r.d * tc;
  const float scale = fmaxf(fmaxf(max_abs3(r.o), max_abs3(c)), fabsf(tc) * max_abs3(r.d));
  const float radius = fabsf(s.w);
  const float reach = radius + 1e-5f * scale;
  const float h2 = reach * reach - dot3(v, v);
  if (!(h2 >= 0.0f)) return false;
  const float inv_len = __builtin_amdgcn_rsqf(dd);
  const float t_in = tc - __builtin_amdgcn_sqrtf(h2) * inv_len;
  return t_in <= tmax + 1e-4f * (radius + scale) * inv_len;
}

struct Closest2 {
  RayPre w;                // world ray
  RayPre r;                // current ray: world at the top level, object space inside an instance
  float tmin;
  TraceHit best;           // best.orig = packed-order triangle index
  uint32_t best_inst;
  uint32_t node;           // next item
  uint32_t inst;           // current instance, OL_NONE at the top level
  uint32_t offset;         // current instance's triangle offset
  const WideNode* bnodes;  // current bottom-level BVH
  const float4* btri;
};

OL_DEV void closest2_begin(Closest2& q, f3 o, f3 d, float tmin, float tmax) {
  q.w = ray_pre(o, d);
  q.r = q.w;
  q.tmin = tmin;
  q.best = TraceHit{tmax, 0.0f, 0.0f, OL_NONE, OL_NONE};
  q.best_inst = OL_NONE;
  q.node = 0;
  q.inst = OL_NONE;
  q.offset = 0;
  q.bnodes = nullptr;
  q.btri = nullptr;
}

// Leaf step: only bottom-level leaves (triangle tests in object space).
OL_DEV bool closest2_at_leaf(const Closest2& q) { return q.inst != OL_NONE && child_is_leaf(q.node); }

template <int LSTACK>
OL_DEV void closest2_next(Closest2& q, TravStack<LSTACK>& st) { q.node = st.empty() ? OL_NONE : st.pop(); }

// Entry step: a top-level leaf names an instance (an entry of one); the ray
// enters it, in object space, unless its bounding sphere starts beyond the
// best hit.  Kept apart from node steps, so that a wave runs the entry code
// only for lanes that all enter (it is several times a node test's cost).
OL_DEV bool closest2_at_entry(const Closest2& q) { return q.inst == OL_NONE && child_is_leaf(q.node); }

template <int LSTACK>
OL_DEV void closest2_enter(const DScene& sc, Closest2& q, TravStack<LSTACK>& st, uint32_t& ni) {
  const uint32_t k = leaf_first(q.node);  // the instance table is in top-level leaf order
  const InstanceDev& I = sc.inst[k];
  if (!sphere_before(q.w, I.sphere, q.best.t) || (instance_later(I.sphere) && hold_back(st, q.node))) {
    closest2_next(q, st);
    return;
  }
  const BlasDev& B = sc.blas[I.blas];
  if (B.node_count == 0) {
    closest2_next(q, st);
    return;
  }
  ++ni;
  st.push(OL_SENTINEL);
  q.inst = k;
  q.offset = I.tri_offset;
  q.bnodes = B.nodes;
  q.btri = B.tri;
  q.r = ray_pre_fast(xform_point(I.w2o, q.w.o), xform_vector(I.w2o, q.w.d));
  q.node = I.root;
}

// Node step: leave an instance (sentinel) or test a node's boxes at the
// current level.
template <int LSTACK>
OL_DEV void closest2_node(const DScene& sc, Closest2& q, TravStack<LSTACK>& st, uint32_t& nv) {
  if (q.node == OL_SENTINEL) {
    q.inst = OL_NONE;
    q.r = q.w;
    closest2_next(q, st);
    return;
  }
  const WideNode* nodes = q.inst == OL_NONE ? sc.nodes : q.bnodes;
  uint4 w0, w1, w2, w3;
  load_node(nodes, q.node, w0, w1, w2, w3);
  ++nv;
  uint32_t ch[4], key[4];
  node_children(w2, w3, ch);
  ol_node_keys(w0, w1, w2, q.r, q.tmin, cull_distance(q.best.t), key, ch);
  sort4_pairs(key, ch);
  descend(st, ch, key, q.node);
}

// Entry or node step, whichever the lane is at.
template <int LSTACK>
OL_DEV void closest2_inner(const DScene& sc, Closest2& q, TravStack<LSTACK>& st, uint32_t& nv, uint32_t& ni) {
  if (closest2_at_entry(q)) closest2_enter<LSTACK>(sc, q, st, ni);
  else closest2_node<LSTACK>(sc, q, st, nv);
}

template <int LSTACK>
OL_DEV void closest2_leaf(Closest2& q, TravStack<LSTACK>& st, uint32_t& nt) {
  const uint32_t first = leaf_first(q.node), cnt = leaf_count(q.node);
  nt += cnt;
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t slot = first + j;
    float4 a, e1, e2;
    load_tri(q.btri, slot, a, e1, e2);
    float t, u, v;
    if (ol_tri_test(q.r.o, q.r.d, a, e1, e2, t, u, v) && t >= q.tmin) {
      const uint32_t orig = q.offset + __float_as_uint(a.w);
      if (t < q.best.t || (t == q.best.t && orig < q.best.orig)) {
        q.best = TraceHit{t, u, v, slot, orig};
        q.best_inst = q.inst;
      }
    }
  }
  closest2_next(q, st);
}

struct Any2 {
  RayPre w, r;
  float tmin, tmax, tcull;
  uint32_t node, inst;
  const WideNode* bnodes;
  const float4* btri;
  bool occluded;
};

OL_DEV void any2_begin(Any2& q, f3 o, f3 d, float tmin, float tmax) {
  q.w = ray_pre(o, d);
  q.r = q.w;
  q.tmin = tmin;
  q.tmax = tmax;
  q.tcull = cull_distance(tmax);
  q.node = 0;
  q.inst = OL_NONE;
  q.bnodes = nullptr;
  q.btri = nullptr;
  q.occluded = false;
}

OL_DEV bool any2_at_leaf(const Any2& q) { return q.inst != OL_NONE && child_is_leaf(q.node); }

template <int LSTACK>
OL_DEV void any2_next(Any2& q, TravStack<LSTACK>& st) { q.node = st.empty() ? OL_NONE : st.pop(); }

OL_DEV bool any2_at_entry(const Any2& q) { return q.inst == OL_NONE && child_is_leaf(q.node); }

template <int LSTACK>
OL_DEV void any2_enter(const DScene& sc, Any2& q, TravStack<LSTACK>& st, uint32_t& ni) {
  const uint32_t k = leaf_first(q.node);  // the instance table is in top-level leaf order
  const InstanceDev& I = sc.inst[k];
  if (!sphere_before(q.w, I.sphere, q.tmax) || (instance_later(I.sphere) && hold_back(st, q.node))) {
    any2_next(q, st);
    return;
  }
  const BlasDev& B = sc.blas[I.blas];
  if (B.node_count == 0) {
    any2_next(q, st);
    return;
  }
  ++ni;
  st.push(OL_SENTINEL);
  q.inst = k;
  q.bnodes = B.nodes;
  q.btri = B.tri;
  q.r = ray_pre_fast(xform_point(I.w2o, q.w.o), xform_vector(I.w2o, q.w.d));
  q.node = I.root;
}

template <int LSTACK>
OL_DEV void any2_node(const DScene& sc, Any2& q, TravStack<LSTACK>& st, uint32_t& nv) {
  if (q.node == OL_SENTINEL) {
    q.inst = OL_NONE;
    q.r = q.w;
    any2_next(q, st);
    return;
  }
  const WideNode* nodes = q.inst == OL_NONE ? sc.nodes : q.bnodes;
  uint4 w0, w1, w2, w3;
  load_node(nodes, q.node, w0, w1, w2, w3);
  ++nv;
  uint32_t ch[4], key[4];
  node_children(w2, w3, ch);
  ol_node_keys(w0, w1, w2, q.r, q.tmin, q.tcull, key, ch);
  build_order(key, ch);
  descend(st, ch, key, q.node);
}

template <int LSTACK>
OL_DEV void any2_inner(const DScene& sc, Any2& q, TravStack<LSTACK>& st, uint32_t& nv, uint32_t& ni) {
  if (any2_at_entry(q)) any2_enter<LSTACK>(sc, q, st, ni);
  else any2_node<LSTACK>(sc, q, st, nv);
}

template <int LSTACK>
OL_DEV void any2_leaf(Any2& q, TravStack<LSTACK>& st, uint32_t& nt) {
  const uint32_t first = leaf_first(q.node), cnt = leaf_count(q.node);
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t slot = first + j;
    float4 a, e1, e2;
    load_tri(q.btri, slot, a, e1, e2);
    float t, u, v;
    ++nt;
    if (ol_tri_test(q.r.o, q.r.d, a, e1, e2, t, u, v) && t >= q.tmin && t <= q.tmax) {
      q.occluded = true;
      q.node = OL_NONE;
      return;
    }
  }
  any2_next(q, st);
}
