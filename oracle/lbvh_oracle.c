/*
 * LBVH oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Sequential CPU build of the algorithm the GPU runs in
 * ordinarylight_amd/csrc/bvh_build.hip, written independently of it (no
 * shared headers): per-triangle bounds, 30-bit Morton codes of the
 * centroids against the centroid bounds, a stable sort by code, the Karras
 * (2012) binary radix tree with index-augmented keys, bottom-up boxes and a
 * breadth-first collapse into 64-byte 4-wide nodes with 8-bit outward-rounded
 * child boxes, each first grown by 2^-17 of the largest root-box coordinate.
 * All steps are integer work, float min/max/compare or single correctly
 * rounded float operations, so the GPU output must match this byte for byte
 * (tests/test_gpu_parity.py::test_bvh_bitexact).
 *
 * The reference uses driver-built hardware BVHs (targets/vulkan/scene.py
 * :1132-1194); there is nothing in it to pin this against except the
 * traversal results, which the GI and brute-force checks pin.  The build
 * itself is pinned here by construction and by tests/test_oracle_lbvh.py
 * (Karras invariants, box containment, exact leaf coverage).
 *
 * Compile: gcc -O2 -ffp-contract=off -shared -fPIC (see oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define LEAF_BIT 0x80000000u
#define NONE 0xFFFFFFFFu

typedef struct {
  float mn[3], mx[3];
} Box;

static uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

static uint32_t q10(float c, float lo, float ext) {
  float n = ext > 0.0f ? (c - lo) / ext : 0.5f;
  float q = fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
  return (uint32_t)q;
}

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static int delta(const uint32_t* codes, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  if (codes[i] != codes[j]) return clz32(codes[i] ^ codes[j]);
  return 32 + clz32((uint32_t)i ^ (uint32_t)j);
}

static float scale_of(uint32_t biased) {
  union {
    uint32_t u;
    float f;
  } v;
  v.u = biased << 23;
  return v.f;
}

static uint32_t exponent_for(float lo, float hi) {
  float ext = hi - lo;
  int e = -100;
  if (ext > 0.0f) {
    union {
      float f;
      uint32_t u;
    } s;
    s.f = ext / 255.0f;
    int E = (int)((s.u >> 23) & 0xFFu);
    e = E == 0 ? -100 : E - 126;
  }
  if (e < -100) e = -100;
  while (e < 127 && lo + 255.0f * scale_of((uint32_t)(e + 127)) < hi) ++e;
  return (uint32_t)(e + 127);
}

static uint32_t quant_down(float lo, float sc, float v) {
  float f = floorf((v - lo) / sc);
  int q = f < 0.0f ? 0 : (f > 255.0f ? 255 : (int)f);
  while (q > 0 && lo + (float)q * sc > v) --q;
  return (uint32_t)q;
}

static uint32_t quant_up(float lo, float sc, float v) {
  float f = ceilf((v - lo) / sc);
  int q = f < 0.0f ? 0 : (f > 255.0f ? 255 : (int)f);
  while (q < 255 && lo + (float)q * sc < v) ++q;
  return (uint32_t)q;
}

typedef struct {
  int n;
  uint32_t leaf_size;
  const uint32_t *left, *right, *first, *last;
  const Box *leaf_box, *node_box;
} Tree;

static uint32_t ref_count(const Tree* t, uint32_t r) { return (r & LEAF_BIT) ? 1u : t->last[r] - t->first[r] + 1u; }
static uint32_t ref_first(const Tree* t, uint32_t r) { return (r & LEAF_BIT) ? (r & ~LEAF_BIT) : t->first[r]; }
static const Box* ref_box(const Tree* t, uint32_t r) {
  return (r & LEAF_BIT) ? &t->leaf_box[r & ~LEAF_BIT] : &t->node_box[r];
}
static int ref_leafy(const Tree* t, uint32_t r) { return (r & LEAF_BIT) || ref_count(t, r) <= t->leaf_size; }
static float ref_area(const Tree* t, uint32_t r) {
  const Box* b = ref_box(t, r);
  float dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
  return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

static int plan(const Tree* t, uint32_t r, uint32_t kids[4]) {
  if (ref_leafy(t, r)) {
    kids[0] = r;
    return 1;
  }
  int nk = 2;
  kids[0] = t->left[r];
  kids[1] = t->right[r];
  while (nk < 4) {
    int best = -1;
    float best_area = -1.0f;
    for (int k = 0; k < nk; ++k) {
      if (ref_leafy(t, kids[k])) continue;
      float a = ref_area(t, kids[k]);
      if (a > best_area) {
        best_area = a;
        best = k;
      }
    }
    if (best < 0) break;
    uint32_t x = kids[best];
    for (int k = nk; k > best + 1; --k) kids[k] = kids[k - 1];
    kids[best] = t->left[x];
    kids[best + 1] = t->right[x];
    ++nk;
  }
  return nk;
}

/* Outputs (caller-allocated):
 *   morton[T], sorted[T], left[T-1], right[T-1], node_box[(T-1)*8],
 *   nodes[(T+1)*64 bytes], tri[T*12 floats];  node_count, depth.
 * Returns 0 on success, -1 on allocation failure. */
int ol_lbvh_build(const float* positions, const uint32_t* indices, uint32_t T, uint32_t leaf_size, uint32_t* morton,
                  uint32_t* sorted, uint32_t* left_out, uint32_t* right_out, float* box_out, uint8_t* nodes_out,
                  uint32_t* node_count, uint32_t* depth_out, float* tri_out) {
  if (leaf_size == 0) leaf_size = 4;
  if (leaf_size > 16) leaf_size = 16;
  *node_count = 0;
  *depth_out = 0;
  if (T == 0) return 0;
  Box* tb = (Box*)malloc(sizeof(Box) * T);
  uint32_t* tmpk = (uint32_t*)malloc(4ull * T);
  uint32_t* tmpv = (uint32_t*)malloc(4ull * T);
  uint32_t* codes = (uint32_t*)malloc(4ull * T);
  if (!tb || !tmpk || !tmpv || !codes) return -1;
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = 0; i < T; ++i) {
    const float* a = positions + 3ull * indices[3ull * i];
    const float* b = positions + 3ull * indices[3ull * i + 1];
    const float* c = positions + 3ull * indices[3ull * i + 2];
    for (int k = 0; k < 3; ++k) {
      tb[i].mn[k] = fminf(fminf(a[k], b[k]), c[k]);
      tb[i].mx[k] = fmaxf(fmaxf(a[k], b[k]), c[k]);
      float cc = (tb[i].mn[k] + tb[i].mx[k]) * 0.5f;
      clo[k] = fminf(clo[k], cc);
      chi[k] = fmaxf(chi[k], cc);
    }
  }
  for (uint32_t i = 0; i < T; ++i) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = q10((tb[i].mn[k] + tb[i].mx[k]) * 0.5f, clo[k], chi[k] - clo[k]);
    morton[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  }
  /* Stable LSD counting sort, 8-bit digits. */
  memcpy(codes, morton, 4ull * T);
  for (uint32_t i = 0; i < T; ++i) sorted[i] = i;
  for (int shift = 0; shift < 32; shift += 8) {
    uint32_t count[257] = {0};
    for (uint32_t i = 0; i < T; ++i) count[((codes[i] >> shift) & 255u) + 1]++;
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (uint32_t i = 0; i < T; ++i) {
      uint32_t d = (codes[i] >> shift) & 255u;
      uint32_t p = count[d]++;
      tmpk[p] = codes[i];
      tmpv[p] = sorted[i];
    }
    memcpy(codes, tmpk, 4ull * T);
    memcpy(sorted, tmpv, 4ull * T);
  }
  const int n = (int)T;
  const int ni = n > 1 ? n - 1 : 1;
  uint32_t* L = (uint32_t*)malloc(4ull * ni);
  uint32_t* R = (uint32_t*)malloc(4ull * ni);
  uint32_t* F = (uint32_t*)malloc(4ull * ni);
  uint32_t* La = (uint32_t*)malloc(4ull * ni);
  Box* nb = (Box*)malloc(sizeof(Box) * ni);
  Box* lb = (Box*)malloc(sizeof(Box) * T);
  if (!L || !R || !F || !La || !nb || !lb) return -1;
  for (int k = 0; k < n; ++k) lb[k] = tb[sorted[k]];
  for (int i = 0; i < n - 1; ++i) {
    int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(codes, n, i, j);
    int s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    L[i] = (lo == gamma) ? ((uint32_t)gamma | LEAF_BIT) : (uint32_t)gamma;
    R[i] = (hi == gamma + 1) ? ((uint32_t)(gamma + 1) | LEAF_BIT) : (uint32_t)(gamma + 1);
    F[i] = (uint32_t)lo;
    La[i] = (uint32_t)hi;
  }
  /* Boxes bottom-up by an explicit post-order walk from the root. */
  if (n > 1) {
    uint32_t* stack = (uint32_t*)malloc(8ull * ni + 16);
    uint8_t* seen = (uint8_t*)calloc((size_t)ni, 1);
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      uint32_t x = stack[sp - 1];
      if (!seen[x]) {
        seen[x] = 1;
        if (!(L[x] & LEAF_BIT)) stack[sp++] = L[x];
        if (!(R[x] & LEAF_BIT)) stack[sp++] = R[x];
        continue;
      }
      --sp;
      const Box* a = (L[x] & LEAF_BIT) ? &lb[L[x] & ~LEAF_BIT] : &nb[L[x]];
      const Box* b = (R[x] & LEAF_BIT) ? &lb[R[x] & ~LEAF_BIT] : &nb[R[x]];
      for (int k = 0; k < 3; ++k) {
        nb[x].mn[k] = fminf(a->mn[k], b->mn[k]);
        nb[x].mx[k] = fmaxf(a->mx[k], b->mx[k]);
      }
    }
    free(stack);
    free(seen);
    for (int i = 0; i < n - 1; ++i) {
      left_out[i] = L[i];
      right_out[i] = R[i];
      float* o = box_out + 8ull * i;
      o[0] = nb[i].mn[0];
      o[1] = nb[i].mn[1];
      o[2] = nb[i].mn[2];
      o[3] = 0.0f;
      o[4] = nb[i].mx[0];
      o[5] = nb[i].mx[1];
      o[6] = nb[i].mx[2];
      o[7] = 0.0f;
    }
  }
  Tree tr = {n, leaf_size, L, R, F, La, lb, nb};
  /* Breadth-first collapse. */
  uint32_t* cur = (uint32_t*)malloc(4ull * (T + 1));
  uint32_t* nxt = (uint32_t*)malloc(4ull * (T + 1));
  if (!cur || !nxt) return -1;
  cur[0] = n > 1 ? 0u : LEAF_BIT;
  /* Box growth: 2^-17 of the largest root-box coordinate on every side. */
  float pad = 0.0f;
  {
    const Box* rb = ref_box(&tr, cur[0]);
    float m = 0.0f;
    for (int a = 0; a < 3; ++a) m = fmaxf(m, fmaxf(fabsf(rb->mn[a]), fabsf(rb->mx[a])));
    pad = m * 0x1p-17f;
  }
  uint32_t level_count = 1, level_base = 0, depth = 0;
  while (level_count > 0) {
    ++depth;
    uint32_t next_base = level_base + level_count, next_count = 0;
    for (uint32_t k = 0; k < level_count; ++k) {
      uint32_t kids[4];
      int nk = plan(&tr, cur[k], kids);
      float nmn[3] = {INFINITY, INFINITY, INFINITY}, nmx[3] = {-INFINITY, -INFINITY, -INFINITY};
      float cmn[4][3], cmx[4][3];
      for (int c = 0; c < nk; ++c) {
        const Box* b = ref_box(&tr, kids[c]);
        for (int a = 0; a < 3; ++a) {
          cmn[c][a] = b->mn[a] - pad;
          cmx[c][a] = b->mx[a] + pad;
          nmn[a] = fminf(nmn[a], cmn[c][a]);
          nmx[a] = fmaxf(nmx[a], cmx[c][a]);
        }
      }
      uint32_t words[16];
      memset(words, 0, sizeof(words));
      memcpy(&words[0], &nmn[0], 4);
      memcpy(&words[1], &nmn[1], 4);
      memcpy(&words[2], &nmn[2], 4);
      uint32_t e[3], ql[3] = {0, 0, 0}, qh[3] = {0, 0, 0};
      for (int a = 0; a < 3; ++a) e[a] = exponent_for(nmn[a], nmx[a]);
      words[3] = e[0] | (e[1] << 8) | (e[2] << 16);
      for (int c = 0; c < 4; ++c) {
        if (c >= nk) {
          words[10 + c] = NONE;
          continue;
        }
        for (int a = 0; a < 3; ++a) {
          float sc = scale_of(e[a]);
          ql[a] |= quant_down(nmn[a], sc, cmn[c][a]) << (8 * c);
          qh[a] |= quant_up(nmn[a], sc, cmx[c][a]) << (8 * c);
        }
        if (ref_leafy(&tr, kids[c])) {
          words[10 + c] = LEAF_BIT | ((ref_count(&tr, kids[c]) - 1u) << 27) | ref_first(&tr, kids[c]);
        } else {
          words[10 + c] = next_base + next_count;
          nxt[next_count++] = kids[c];
        }
      }
      words[4] = ql[0];
      words[5] = ql[1];
      words[6] = ql[2];
      words[7] = qh[0];
      words[8] = qh[1];
      words[9] = qh[2];
      memcpy(nodes_out + 64ull * (level_base + k), words, 64);
    }
    level_base = next_base;
    level_count = next_count;
    uint32_t* sw = cur;
    cur = nxt;
    nxt = sw;
  }
  *node_count = level_base;
  *depth_out = depth;
  for (uint32_t k = 0; k < T; ++k) {
    uint32_t o = sorted[k];
    const float* a = positions + 3ull * indices[3ull * o];
    const float* b = positions + 3ull * indices[3ull * o + 1];
    const float* c = positions + 3ull * indices[3ull * o + 2];
    float* r = tri_out + 12ull * k;
    r[0] = a[0];
    r[1] = a[1];
    r[2] = a[2];
    memcpy(&r[3], &o, 4);
    r[4] = b[0] - a[0];
    r[5] = b[1] - a[1];
    r[6] = b[2] - a[2];
    r[7] = 0.0f;
    r[8] = c[0] - a[0];
    r[9] = c[1] - a[1];
    r[10] = c[2] - a[2];
    r[11] = 0.0f;
  }
  free(tb);
  free(tmpk);
  free(tmpv);
  free(codes);
  free(L);
  free(R);
  free(F);
  free(La);
  free(nb);
  free(lb);
  free(cur);
  free(nxt);
  return 0;
}
