/*
 * GI estimator oracle (O-E1) -- TEST INFRASTRUCTURE ONLY.
 *
 * Sequential C restatement of the reference's split-path GI shade stage,
 * ordinarylight/shaders/wavefront_shade_candidate.glsl main (:3106-3493),
 * with its ray generation (wavefront_generate.comp:86-167) and host
 * accumulation (targets/vulkan/api.py:1322-1422).  One path at a time, the
 * next-event shadow rays traced inline exactly where the GLSL traces them
 * (nativeOccluded, :3026), over this file's own BVH (binary, median split,
 * double-precision boxes) -- nothing is shared with the HIP sources.
 *
 * Compiled twice (oracle/Makefile): OL_REAL=float is the parity oracle of
 * the GPU (same float operation order, -ffp-contract=off); OL_REAL=double
 * is the FP64 evaluation of the same estimator for convergence checks.
 *
 * Built-in untextured PBR materials only.  Deliberate deviation shared with
 * the product: the occlusion factor mix(texture_parameters.w, 1, metallic)
 * (:1542, :1957) reads the value the packer resolves (1.0 for untextured
 * materials) instead of the transmission-texture binding slot (-1).
 *
 * Functions follow (file:line in wavefront_shade_candidate.glsl):
 *   waveCosineHemisphere :630, rayQueryRandomState/Value :661-671,
 *   shadeResolveSurface :957, shadeResolveEnvironmentMiss :1231,
 *   shadeTransmitPath :1264, shadeApplyRussianRoulette :1335,
 *   shadeBuildContinuation :1356, PBR lobes/pdf/sampling :1406-1552,
 *   point lights :1554-1619, area lights :1621-1733, unified domain
 *   :1735-1754/:1903-1928, secondary NEE selection :1756-1777, emission
 *   :1779-1809, environment :1811-1894, environment NEE :1930-1959.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OL_REAL
#define OL_REAL float
#endif
#ifndef OL_SUFFIX
#define OL_SUFFIX _f32
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define API(name) CAT(name, OL_SUFFIX)

typedef OL_REAL R;

/* Planted mistakes for the tests' mutation checks: a build with
 * -DOL_MUTANT=k (oracle/Makefile, libol_oracle_m<k>.so) breaks exactly one
 * term, so that the closed-form checks can show they catch it.  The default
 * build (OL_MUTANT 0) has none.
 *   1 refraction eta inverted          7 exit pops to air, not the outer medium
 *   2 point light without 1/d^2        8 area-light pdf without d^2
 *   3 point light without cos          9 environment NEE without its MIS weight
 *   4 roulette survivors not 1/p      10 spot cone linear instead of smoothstep
 *   5 Beer absorption on the wrong    11 emission hit without its MIS weight
 *     segment                         12 GGX specular lobe without D
 *   6 total internal reflection passes straight on
 *  13 bilinear filtering without the half-texel shift
 *  14 sRGB texels not decoded to linear
 *  15 mip level chosen without the binding's uv scale
 *  16 normal map without the tangent frame's handedness */
#ifndef OL_MUTANT
#define OL_MUTANT 0
#endif
#define OL_MUT(k) (OL_MUTANT == (k))

/* Float32 Moller-Trumbore accepts points a few ulps of the scene scale outside
 * the exact triangle; node boxes are padded by this fraction of the largest
 * coordinate so that the exact box test never culls such a hit. */
#ifndef OL_BOX_PAD
#define OL_BOX_PAD 1e-5
#endif
typedef struct {
  R x, y, z;
} V;

static inline V v3(R x, R y, R z) {
  V r = {x, y, z};
  return r;
}
static inline V vadd(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V vsub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V vneg(V a) { return v3(-a.x, -a.y, -a.z); }
static inline V vmul(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V vs(V a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V sv(R s, V a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V vdiv(V a, R s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline R dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V cross(V a, V b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline R rsqrt_(R x) { return sizeof(R) == 4 ? (R)sqrtf((float)x) : (R)sqrt((double)x); }
static inline R len(V a) { return rsqrt_(dot(a, a)); }
static inline V norm(V a) { return vs(a, (R)1 / rsqrt_(dot(a, a))); }
static inline R rmax(R a, R b) { return a > b ? a : (b != b ? a : b); }
static inline R rmin(R a, R b) { return a < b ? a : (b != b ? a : b); }
static inline R clampr(R v, R lo, R hi) { return rmin(rmax(v, lo), hi); }
static inline R mixr(R a, R b, R t) { return a * ((R)1 - t) + b * t; }
static inline V mixv(V a, V b, R t) {
  return v3(a.x * ((R)1 - t) + b.x * t, a.y * ((R)1 - t) + b.y * t, a.z * ((R)1 - t) + b.z * t);
}
static inline V reflect_(V d, V n) { return vsub(d, sv((R)2 * dot(n, d), n)); }
static inline R pow5(R x) { return ((x * x) * (x * x)) * x; }
static inline R absr(R x) { return x < 0 ? -x : x; }

#define PI_R ((R)3.14159265359)
#define TWO_PI_R ((R)6.28318530718)

/* Cephes sinf/cosf (S. Moshier): octant reduction with a three-part pi/4,
 * minimax polynomials on [-pi/4, pi/4].  The double build uses libm. */
static void sincos_r(R xr, R* so, R* co) {
  if (sizeof(R) == 8) {
    *so = (R)sin((double)xr);
    *co = (R)cos((double)xr);
    return;
  }
  float x = (float)xr, ss = 1.0f, cs = 1.0f;
  if (x < 0.0f) {
    x = -x;
    ss = -1.0f;
  }
  int j = (int)(x * 1.27323954473516f);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  if (j > 3) {
    j -= 4;
    ss = -ss;
    cs = -cs;
  }
  if (j > 1) cs = -cs;
  float z = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
  float zz = z * z;
  float ps = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
  float pc = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz - 0.5f * zz + 1.0f;
  float s = (j == 1 || j == 2) ? pc : ps;
  float c = (j == 1 || j == 2) ? ps : pc;
  *so = (R)(ss * s);
  *co = (R)(cs * c);
}

static uint32_t hash32(uint32_t v) {
  v ^= v >> 16;
  v *= 2146121005u;
  v ^= v >> 15;
  v *= 2221713035u;
  v ^= v >> 16;
  return v;
}
static R unit(uint32_t w) { return (R)w * (R)(1.0 / 4294967296.0); }
static uint32_t pcg(uint32_t s) { return s * 747796405u + 2891336453u; }

/* Rounding-size perturbation (for the tests' mismatch classification, not
 * part of the estimator): with a nonzero seed every uniform sample moves
 * `ulps` units in the last place, up or down by a hash of its state and the
 * seed (staying in [0, 1)).  Pixels whose value changes under such moves sit
 * on a near-tie decision (a selection threshold, a triangle edge, a
 * closest-hit tie), where a rounding difference between two float32
 * evaluations of the same path can send it elsewhere. */
static uint32_t g_jitter = 0u, g_jitter_ulps = 1u, g_jitter_t_ulps = 0u;
void API(olgi_set_jitter)(uint32_t seed, uint32_t ulps, uint32_t t_ulps) {
  g_jitter = seed;
  g_jitter_ulps = ulps;
  g_jitter_t_ulps = t_ulps;
}

/* The same for path-tracing hit distances: `t_ulps` units in the last place
 * of t (the size of the rounding difference between intersecting in world
 * space and in an instance's object space). */
static R jitter_t(R t, uint32_t salt) {
  if (g_jitter == 0u || g_jitter_t_ulps == 0u) return t;
  const R up = sizeof(R) == sizeof(float) ? (R)nextafterf((float)t, 3.0e38f) : (R)nextafter((double)t, 1.0e300);
  const R step = (up - t) * (R)g_jitter_t_ulps;
  return (hash32(salt ^ g_jitter) & 1u) ? t + step : t - step;
}

static R pcg_unit(uint32_t s) {
  uint32_t w = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
  w = (w >> 22) ^ w;
  R u = unit(w);
  if (g_jitter != 0u && g_jitter_ulps != 0u) {
    const R up = sizeof(R) == sizeof(float) ? (R)nextafterf((float)u, 1.0f) : (R)nextafter((double)u, 1.0);
    const R step = (up - u) * (R)g_jitter_ulps;
    const R moved = (hash32(s ^ g_jitter) & 1u) ? u + step : u - step;
    u = moved < (R)0.0 ? (R)0.0 : (moved >= (R)1.0 ? u : moved);
  }
  return u;
}
static uint32_t brev(uint32_t v) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= ((v >> i) & 1u) << (31 - i);
  return r;
}

/* ---- scene + BVH2 ------------------------------------------------------ */
typedef struct {
  double mn[3], mx[3];
  uint32_t left, right; /* leaf: right = count | 0x80000000, left = first */
} BNode;

typedef struct Scene_ {
  uint32_t V, T, M, P, L;
  float* pos;      /* V*3 */
  float* nrm;      /* V*3 */
  uint32_t* idx;   /* T*3 */
  uint32_t* mid;   /* T */
  float* mat;      /* M*48 */
  float* ana;      /* P*16 */
  float* area;     /* L*20 */
  float weight;
  int env_valid;
  float env[3];
  BNode* nodes;
  uint32_t node_count;
  uint32_t* order; /* leaf order -> triangle */
  double pad;      /* node boxes are grown by this much on every side */
  /* Textures (olgi_scene_textures; NULL / 0 without): packed texture words,
   * binding records (8 floats each), per-vertex uv0.xy uv1.xy and world
   * tangents, per-triangle uv0 density; the image environment (texture,
   * rotation, log range from the environment record's row 3). */
  uint32_t* words;
  uint32_t nwords;
  float* bind;
  uint32_t B;
  float* uvs;
  float* tans;
  float* dens;
  int env_tex;
  float env_rot, env_log, env_col[3], env_int;
  /* Primitives that are boxes (6 floats: min xyz, max xyz) instead of
   * triangles: the top level of a two-level scene. */
  float* boxes;
  /* Two-level scenes (olgi_scene_new_instanced): I instances of R object-
   * space resource scenes (positions, normals, indices, their own BVH); the
   * instance arrays hold world-to-object rows (12 floats), normal-matrix
   * rows (9), determinant, resource, material and first global triangle; a
   * BVH over the instances' world boxes (`top`).  Rays are intersected in
   * each instance's object space with the device's arithmetic (the ray is
   * transformed, not the triangles), and ids are global: the instance's
   * first triangle plus the local id.  T counts the instanced triangles. */
  uint32_t I, R_count;
  struct Scene_* res;
  struct Scene_* top;
  float *w2o, *nmx, *det;
  uint32_t *ires, *imat, *ioff;
} Scene;

static void tri_box(const Scene* s, uint32_t t, double mn[3], double mx[3]) {
  if (s->boxes) {
    for (int k = 0; k < 3; ++k) {
      mn[k] = s->boxes[6ull * t + k];
      mx[k] = s->boxes[6ull * t + 3 + k];
    }
    return;
  }
  for (int k = 0; k < 3; ++k) {
    mn[k] = INFINITY;
    mx[k] = -INFINITY;
  }
  for (int c = 0; c < 3; ++c) {
    const float* p = s->pos + 3ull * s->idx[3ull * t + c];
    for (int k = 0; k < 3; ++k) {
      if (p[k] < mn[k]) mn[k] = p[k];
      if (p[k] > mx[k]) mx[k] = p[k];
    }
  }
}

static const Scene* g_sort_scene;
static int g_sort_axis;
static double* g_cent;
static int cmp_cent(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  double cx = g_cent[3ull * x + g_sort_axis], cy = g_cent[3ull * y + g_sort_axis];
  if (cx < cy) return -1;
  if (cx > cy) return 1;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static uint32_t build_rec(Scene* s, uint32_t first, uint32_t count) {
  uint32_t id = s->node_count++;
  BNode* n = &s->nodes[id];
  for (int k = 0; k < 3; ++k) {
    n->mn[k] = INFINITY;
    n->mx[k] = -INFINITY;
  }
  double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = first; i < first + count; ++i) {
    double a[3], b[3];
    tri_box(s, s->order[i], a, b);
    for (int k = 0; k < 3; ++k) {
      if (a[k] < n->mn[k]) n->mn[k] = a[k];
      if (b[k] > n->mx[k]) n->mx[k] = b[k];
      double c = g_cent[3ull * s->order[i] + k];
      if (c < cmn[k]) cmn[k] = c;
      if (c > cmx[k]) cmx[k] = c;
    }
  }
  for (int k = 0; k < 3; ++k) {
    n->mn[k] -= s->pad;
    n->mx[k] += s->pad;
  }
  if (count <= 4) {
    n->left = first;
    n->right = count | 0x80000000u;
    return id;
  }
  int axis = 0;
  double ext = cmx[0] - cmn[0];
  for (int k = 1; k < 3; ++k)
    if (cmx[k] - cmn[k] > ext) {
      ext = cmx[k] - cmn[k];
      axis = k;
    }
  g_sort_axis = axis;
  qsort(s->order + first, count, sizeof(uint32_t), cmp_cent);
  uint32_t half = count / 2;
  uint32_t l = build_rec(s, first, half);
  uint32_t r = build_rec(s, first + half, count - half);
  s->nodes[id].left = l;
  s->nodes[id].right = r;
  return id;
}

/* Median-split BVH2 over the scene's T primitives (triangles or boxes). */
static void build_bvh(Scene* s) {
  const uint32_t T = s->T;
  s->nodes = (BNode*)malloc(sizeof(BNode) * (2ull * T + 2));
  s->order = (uint32_t*)malloc(4ull * T + 4);
  g_cent = (double*)malloc(24ull * T + 8);
  double scale = 0.0;
  for (uint32_t t = 0; t < T; ++t) {
    double a[3], b[3];
    tri_box(s, t, a, b);
    for (int k = 0; k < 3; ++k) {
      g_cent[3ull * t + k] = 0.5 * (a[k] + b[k]);
      if (s->boxes) scale = fmax(scale, fmax(fabs(a[k]), fabs(b[k])));
    }
    s->order[t] = t;
  }
  if (!s->boxes)
    for (uint64_t i = 0; i < 3ull * s->V; ++i) scale = fmax(scale, fabs((double)s->pos[i]));
  s->pad = OL_BOX_PAD * (1.0 + scale);
  g_sort_scene = s;
  if (T) build_rec(s, 0, T);
  free(g_cent);
  g_cent = NULL;
}

void* API(olgi_scene_new)(const float* pos, const float* nrm, uint32_t V, const uint32_t* idx, const uint32_t* mid,
                          uint32_t T, const float* mat, uint32_t M, const float* ana, uint32_t P, const float* area,
                          uint32_t L, float weight) {
  Scene* s = (Scene*)calloc(1, sizeof(Scene));
  s->V = V;
  s->T = T;
  s->M = M;
  s->P = P;
  s->L = L;
  s->pos = (float*)malloc(12ull * V + 4);
  s->nrm = (float*)malloc(12ull * V + 4);
  s->idx = (uint32_t*)malloc(12ull * T + 4);
  s->mid = (uint32_t*)malloc(4ull * T + 4);
  s->mat = (float*)malloc(192ull * M + 4);
  s->ana = (float*)malloc(64ull * P + 4);
  s->area = (float*)malloc(80ull * L + 4);
  memcpy(s->pos, pos, 12ull * V);
  memcpy(s->nrm, nrm, 12ull * V);
  memcpy(s->idx, idx, 12ull * T);
  memcpy(s->mid, mid, 4ull * T);
  memcpy(s->mat, mat, 192ull * M);
  if (P) memcpy(s->ana, ana, 64ull * P);
  if (L) memcpy(s->area, area, 80ull * L);
  s->weight = weight;
  s->env_tex = -1;
  for (uint32_t l = 0; l < P; ++l) {
    const float* a = s->ana + 16ull * l;
    if ((int)(a[3] + 0.5f) == 3) {
      s->env_valid = 1;
      s->env[0] = (1.0f * a[8]) * a[11];
      s->env[1] = (1.0f * a[9]) * a[11];
      s->env[2] = (1.0f * a[10]) * a[11];
      s->env_col[0] = a[8];
      s->env_col[1] = a[9];
      s->env_col[2] = a[10];
      s->env_int = a[11];
      s->env_tex = a[12] >= 0.0f ? (int)a[12] : -1;
      s->env_rot = a[13];
      s->env_log = a[14];
      break;
    }
  }
  build_bvh(s);
  return s;
}

/* A two-level scene: R object-space resources (vertices pos/nrm split by
 * voff[R + 1], local indices split by toff[R + 1]) placed by I instances
 * (resource, material, first global triangle, world-to-object rows, normal
 * matrix rows, determinant, world box min/max), with materials and lights
 * as olgi_scene_new takes them (emitters as world-space records). */
void* API(olgi_scene_new_instanced)(const float* pos, const float* nrm, const uint32_t* voff, const uint32_t* idx,
                                    const uint32_t* toff, uint32_t R_count, uint32_t I, const uint32_t* ires,
                                    const uint32_t* imat, const uint32_t* ioff, const float* w2o, const float* nmx,
                                    const float* det, const float* wbox, const float* mat, uint32_t M,
                                    const float* ana, uint32_t P, const float* area, uint32_t L, float weight) {
  static const float none[4] = {0, 0, 0, 0};
  static const uint32_t none_u[4] = {0, 0, 0, 0};
  Scene* s = (Scene*)API(olgi_scene_new)(none, none, 0, none_u, none_u, 0, mat, M, ana, P, area, L, weight);
  s->R_count = R_count;
  s->res = (Scene*)calloc(R_count ? R_count : 1, sizeof(Scene));
  for (uint32_t r = 0; r < R_count; ++r) {
    Scene* q = &s->res[r];
    q->V = voff[r + 1] - voff[r];
    q->T = toff[r + 1] - toff[r];
    q->pos = (float*)malloc(12ull * q->V + 4);
    q->nrm = (float*)malloc(12ull * q->V + 4);
    q->idx = (uint32_t*)malloc(12ull * q->T + 4);
    memcpy(q->pos, pos + 3ull * voff[r], 12ull * q->V);
    memcpy(q->nrm, nrm + 3ull * voff[r], 12ull * q->V);
    memcpy(q->idx, idx + 3ull * toff[r], 12ull * q->T);
    build_bvh(q);
  }
  s->I = I;
  s->ires = (uint32_t*)malloc(4ull * I + 4);
  s->imat = (uint32_t*)malloc(4ull * I + 4);
  s->ioff = (uint32_t*)malloc(4ull * I + 4);
  s->w2o = (float*)malloc(48ull * I + 4);
  s->nmx = (float*)malloc(36ull * I + 4);
  s->det = (float*)malloc(4ull * I + 4);
  memcpy(s->ires, ires, 4ull * I);
  memcpy(s->imat, imat, 4ull * I);
  memcpy(s->ioff, ioff, 4ull * I);
  memcpy(s->w2o, w2o, 48ull * I);
  memcpy(s->nmx, nmx, 36ull * I);
  memcpy(s->det, det, 4ull * I);
  s->T = I ? ioff[I - 1] + s->res[ires[I - 1]].T : 0u;
  s->top = (Scene*)calloc(1, sizeof(Scene));
  s->top->T = I;
  s->top->boxes = (float*)malloc(24ull * I + 4);
  memcpy(s->top->boxes, wbox, 24ull * I);
  build_bvh(s->top);
  return s;
}

void API(olgi_scene_free)(void* h) {
  Scene* s = (Scene*)h;
  if (!s) return;
  free(s->pos);
  free(s->nrm);
  free(s->idx);
  free(s->mid);
  free(s->mat);
  free(s->ana);
  free(s->area);
  free(s->nodes);
  free(s->order);
  free(s->words);
  free(s->bind);
  free(s->uvs);
  free(s->tans);
  free(s->dens);
  free(s->boxes);
  if (s->top) API(olgi_scene_free)(s->top);
  for (uint32_t r = 0; r < s->R_count; ++r) {
    Scene* q = &s->res[r];
    free(q->pos);
    free(q->nrm);
    free(q->idx);
    free(q->nodes);
    free(q->order);
  }
  free(s->res);
  free(s->w2o);
  free(s->nmx);
  free(s->det);
  free(s->ires);
  free(s->imat);
  free(s->ioff);
  free(s);
}

/* Attaches texture data to a scene (copied).  uvs / tangents / density may
 * be NULL when no material has texture bindings. */
void API(olgi_scene_textures)(void* h, const uint32_t* words, uint32_t nwords, const float* bind, uint32_t B,
                              const float* uvs, const float* tans, const float* dens) {
  Scene* s = (Scene*)h;
  if (nwords) {
    s->words = (uint32_t*)malloc(4ull * nwords);
    memcpy(s->words, words, 4ull * nwords);
    s->nwords = nwords;
  }
  if (B) {
    s->bind = (float*)malloc(32ull * B);
    memcpy(s->bind, bind, 32ull * B);
    s->B = B;
  }
  if (uvs) {
    s->uvs = (float*)malloc(16ull * s->V + 4);
    s->tans = (float*)malloc(16ull * s->V + 4);
    s->dens = (float*)malloc(4ull * s->T + 4);
    memcpy(s->uvs, uvs, 16ull * s->V);
    memcpy(s->tans, tans, 16ull * s->V);
    memcpy(s->dens, dens, 4ull * s->T);
  }
}

/* Moller-Trumbore in R, expressions as written. */
static int tri_hit(const Scene* s, uint32_t t, V o, V d, R* tt, R* uu, R* vv) {
  const float* a = s->pos + 3ull * s->idx[3ull * t];
  const float* b = s->pos + 3ull * s->idx[3ull * t + 1];
  const float* c = s->pos + 3ull * s->idx[3ull * t + 2];
  V A = v3(a[0], a[1], a[2]);
  V E1 = v3((R)(b[0] - a[0]), (R)(b[1] - a[1]), (R)(b[2] - a[2]));
  V E2 = v3((R)(c[0] - a[0]), (R)(c[1] - a[1]), (R)(c[2] - a[2]));
  V p = cross(d, E2);
  R det = dot(E1, p);
  if (det == (R)0) return 0;
  R inv = (R)1 / det;
  V sv_ = vsub(o, A);
  R u = dot(sv_, p) * inv;
  if (!(u >= (R)0 && u <= (R)1)) return 0;
  V q = cross(sv_, E1);
  R v = dot(d, q) * inv;
  if (!(v >= (R)0 && u + v <= (R)1)) return 0;
  *tt = dot(E2, q) * inv;
  *uu = u;
  *vv = v;
  return 1;
}

static int box_hit(const BNode* n, const double o[3], const double inv[3], double tmin, double tmax) {
  double t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double a = (n->mn[k] - o[k]) * inv[k], b = (n->mx[k] - o[k]) * inv[k];
    if (a > b) {
      double x = a;
      a = b;
      b = x;
    }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
  }
  return t0 <= t1 * (1.0 + 1e-9) + 1e-9;
}

/* Closest hit in [tmin, tmax]: lowest t, ties to the lowest triangle index. */
static int closest1(const Scene* s, V o, V d, R tmin, R tmax, R* bt, R* bu, R* bv, uint32_t* bi) {
  *bt = tmax;
  *bi = 0xFFFFFFFFu;
  if (!s->T) return 0;
  double od[3] = {o.x, o.y, o.z}, inv[3];
  double dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (fabs(dd[k]) > 1e-300 ? dd[k] : copysign(1e-300, dd[k]));
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode* n = &s->nodes[stack[--sp]];
    if (!box_hit(n, od, inv, (double)tmin - 1e-9, (double)*bt * (1.0 + 1e-5) + 1e-9)) continue;
    if (n->right & 0x80000000u) {
      uint32_t cnt = n->right & 0x7FFFFFFFu;
      for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t t = s->order[n->left + k];
        R tt, uu, vv;
        if (tri_hit(s, t, o, d, &tt, &uu, &vv) && tt >= tmin &&
            (tt < *bt || (tt == *bt && t < *bi))) {
          *bt = tt;
          *bu = uu;
          *bv = vv;
          *bi = t;
        }
      }
    } else {
      stack[sp++] = n->left;
      stack[sp++] = n->right;
    }
  }
  return *bi != 0xFFFFFFFFu;
}

static _Thread_local uint64_t g_shadow_rays;

static int occluded1(const Scene* s, V o, V d, R tmin, R tmax) {
  if (!s->T) return 0;
  double od[3] = {o.x, o.y, o.z}, inv[3], dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (fabs(dd[k]) > 1e-300 ? dd[k] : copysign(1e-300, dd[k]));
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode* n = &s->nodes[stack[--sp]];
    if (!box_hit(n, od, inv, (double)tmin - 1e-9, (double)tmax * (1.0 + 1e-5) + 1e-9)) continue;
    if (n->right & 0x80000000u) {
      uint32_t cnt = n->right & 0x7FFFFFFFu;
      for (uint32_t k = 0; k < cnt; ++k) {
        R tt, uu, vv;
        if (tri_hit(s, s->order[n->left + k], o, d, &tt, &uu, &vv) && tt >= tmin && tt <= tmax) return 1;
      }
    } else {
      stack[sp++] = n->left;
      stack[sp++] = n->right;
    }
  }
  return 0;
}

/* Instance transforms, each component summed left to right as the device
 * does: world -> object point and vector (rows of 4: x y z translation),
 * object -> world normal (rows of 3). */
static V xf_point(const float* r, V p) {
  return v3((((R)r[0] * p.x + (R)r[1] * p.y) + (R)r[2] * p.z) + (R)r[3],
            (((R)r[4] * p.x + (R)r[5] * p.y) + (R)r[6] * p.z) + (R)r[7],
            (((R)r[8] * p.x + (R)r[9] * p.y) + (R)r[10] * p.z) + (R)r[11]);
}
static V xf_vector(const float* r, V v) {
  return v3(((R)r[0] * v.x + (R)r[1] * v.y) + (R)r[2] * v.z, ((R)r[4] * v.x + (R)r[5] * v.y) + (R)r[6] * v.z,
            ((R)r[8] * v.x + (R)r[9] * v.y) + (R)r[10] * v.z);
}
static V xf_normal(const float* m, V v) {
  return v3(((R)m[0] * v.x + (R)m[1] * v.y) + (R)m[2] * v.z, ((R)m[3] * v.x + (R)m[4] * v.y) + (R)m[5] * v.z,
            ((R)m[6] * v.x + (R)m[7] * v.y) + (R)m[8] * v.z);
}

/* Two-level closest hit: the top level's boxes in world space, each reached
 * instance's resource in its object space (the direction is not
 * renormalised, so t stays the world distance); ties to the lowest global
 * triangle id.  *inst = the hit's instance. */
static int closest2(const Scene* s, V o, V d, R tmin, R tmax, R* bt, R* bu, R* bv, uint32_t* bi, uint32_t* inst) {
  const Scene* tl = s->top;
  *bt = tmax;
  *bi = 0xFFFFFFFFu;
  *inst = 0xFFFFFFFFu;
  double od[3] = {o.x, o.y, o.z}, inv[3], dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (fabs(dd[k]) > 1e-300 ? dd[k] : copysign(1e-300, dd[k]));
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode* n = &tl->nodes[stack[--sp]];
    if (!box_hit(n, od, inv, (double)tmin - 1e-9, (double)*bt * (1.0 + 1e-5) + 1e-9)) continue;
    if (n->right & 0x80000000u) {
      const uint32_t cnt = n->right & 0x7FFFFFFFu;
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t i = tl->order[n->left + k];
        const float* w = s->w2o + 12ull * i;
        R tt, uu, vv;
        uint32_t local;
        if (!closest1(&s->res[s->ires[i]], xf_point(w, o), xf_vector(w, d), tmin, *bt, &tt, &uu, &vv, &local))
          continue;
        const uint32_t id = s->ioff[i] + local;
        if (tt < *bt || (tt == *bt && id < *bi)) {
          *bt = tt;
          *bu = uu;
          *bv = vv;
          *bi = id;
          *inst = i;
        }
      }
    } else {
      stack[sp++] = n->left;
      stack[sp++] = n->right;
    }
  }
  return *bi != 0xFFFFFFFFu;
}

static int occluded2(const Scene* s, V o, V d, R tmin, R tmax) {
  const Scene* tl = s->top;
  double od[3] = {o.x, o.y, o.z}, inv[3], dd[3] = {d.x, d.y, d.z};
  for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (fabs(dd[k]) > 1e-300 ? dd[k] : copysign(1e-300, dd[k]));
  uint32_t stack[256];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode* n = &tl->nodes[stack[--sp]];
    if (!box_hit(n, od, inv, (double)tmin - 1e-9, (double)tmax * (1.0 + 1e-5) + 1e-9)) continue;
    if (n->right & 0x80000000u) {
      const uint32_t cnt = n->right & 0x7FFFFFFFu;
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t i = tl->order[n->left + k];
        const float* w = s->w2o + 12ull * i;
        if (occluded1(&s->res[s->ires[i]], xf_point(w, o), xf_vector(w, d), tmin, tmax)) return 1;
      }
    } else {
      stack[sp++] = n->left;
      stack[sp++] = n->right;
    }
  }
  return 0;
}

static int closest(const Scene* s, V o, V d, R tmin, R tmax, R* bt, R* bu, R* bv, uint32_t* bi, uint32_t* inst) {
  if (s->I) return closest2(s, o, d, tmin, tmax, bt, bu, bv, bi, inst);
  *inst = 0xFFFFFFFFu;
  return closest1(s, o, d, tmin, tmax, bt, bu, bv, bi);
}

static int occluded(const Scene* s, V o, V d, R tmin, R tmax) {
  ++g_shadow_rays;
  if (s->I) return occluded2(s, o, d, tmin, tmax);
  return occluded1(s, o, d, tmin, tmax);
}

/* A hit's edge cross product (world space), interpolated shading normal
 * (not normalised) and material: from the flat triangle, or from the
 * instance's object-space triangle through its normal matrix and
 * determinant (the device's shading-record path). */
static void hit_frame(const Scene* s, uint32_t prim, uint32_t inst, R bu, R bv, V* cw, V* sn, uint32_t* mat) {
  const Scene* g = s;
  uint32_t t = prim;
  if (s->I) {
    g = &s->res[s->ires[inst]];
    t = prim - s->ioff[inst];
  }
  const uint32_t* tv = g->idx + 3ull * t;
  const float *pa = g->pos + 3ull * tv[0], *pb = g->pos + 3ull * tv[1], *pc = g->pos + 3ull * tv[2];
  V e1 = v3((R)(pb[0] - pa[0]), (R)(pb[1] - pa[1]), (R)(pb[2] - pa[2]));
  V e2 = v3((R)(pc[0] - pa[0]), (R)(pc[1] - pa[1]), (R)(pc[2] - pa[2]));
  const float *q0 = g->nrm + 3ull * tv[0], *q1 = g->nrm + 3ull * tv[1], *q2 = g->nrm + 3ull * tv[2];
  V n0 = v3(q0[0], q0[1], q0[2]), n1 = v3(q1[0], q1[1], q1[2]), n2 = v3(q2[0], q2[1], q2[2]);
  *cw = cross(e1, e2);
  if (s->I) {
    const float* m = s->nmx + 9ull * inst;
    *cw = vs(xf_normal(m, *cw), (R)s->det[inst]);
    n0 = norm(xf_normal(m, n0));
    n1 = norm(xf_normal(m, n1));
    n2 = norm(xf_normal(m, n2));
    *mat = s->imat[inst];
  } else {
    *mat = s->mid[prim];
  }
  const R wx = ((R)1 - bu) - bv, wy = bu, wz = bv;
  *sn = vadd(vadd(vs(n0, wx), vs(n1, wy)), vs(n2, wz));
}

/* ---- materials and BSDF ----------------------------------------------------- */
typedef struct {
  V base;
  R rough, metal, trans, ior, adist, two_sided, occl;
  V atten;
  R cc, ccr, sheen_r, aniso, sub, thin, thick;
  V sheen, subc;
} Mat;

static Mat load_mat(const Scene* s, uint32_t m) {
  const float* r = s->mat + 48ull * m;
  Mat x;
  x.base = v3(r[0], r[1], r[2]);
  x.rough = r[3];
  x.metal = r[7];
  x.atten = v3(r[8], r[9], r[10]);
  x.trans = r[11];
  x.ior = r[12];
  x.adist = r[13];
  x.two_sided = r[15];
  x.occl = r[23];
  x.cc = r[24];
  x.ccr = r[25];
  x.sheen_r = r[26];
  x.aniso = r[27];
  x.sub = r[28];
  x.thin = r[30];
  x.thick = r[31];
  x.sheen = v3(r[32], r[33], r[34]);
  x.subc = v3(r[36], r[37], r[38]);
  return x;
}
static V emission_of(const Scene* s, uint32_t m) {
  const float* r = s->mat + 48ull * m;
  return v3(r[4], r[5], r[6]);
}

static R ggx_d(R nh, R rough) {
  R alpha = rmax(rough * rough, (R)0.0009);
  R a2 = alpha * alpha;
  R den = ((nh * nh) * (a2 - (R)1)) + (R)1;
  return a2 / rmax((PI_R * den) * den, (R)1e-6);
}
static R ggx_g(R nd, R rough) {
  R alpha = rmax(rough * rough, (R)0.0009);
  R a2 = alpha * alpha;
  return ((R)2 * nd) / rmax(nd + rsqrt_(a2 + (((R)1 - a2) * nd) * nd), (R)1e-6);
}
static V f0_of(const Mat* m) { return mixv(v3((R)0.04, (R)0.04, (R)0.04), m->base, m->metal); }
static R spec_prob(const Mat* m) {
  V f0 = f0_of(m);
  return clampr(rmax(f0.x, rmax(f0.y, f0.z)) + m->cc * (R)0.15, (R)0.1, (R)0.95);
}

static V eval_pbr(const Mat* m, V n, V view, V out) {
  R nv = rmax(dot(n, view), (R)0), nl = rmax(dot(n, out), (R)0);
  if (nv <= (R)0 || nl <= (R)0) return v3(0, 0, 0);
  V h = norm(vadd(view, out));
  R nh = rmax(dot(n, h), (R)0), vh = rmax(dot(view, h), (R)0);
  R k = pow5((R)1 - clampr(vh, (R)0, (R)1));
  V f0 = f0_of(m);
  V fr = vadd(f0, vs(vsub(v3(1, 1, 1), f0), k));
  R er = m->rough * ((R)1 - (R)0.25 * absr(clampr(m->aniso, (R)-1, (R)1)));
  R D = OL_MUT(12) ? (R)1 : ggx_d(nh, er);
  R G = ggx_g(nv, er) * ggx_g(nl, er);
  R den = rmax(((R)4 * nv) * nl, (R)1e-6);
  V spec = vdiv(vs(vs(fr, D), G), den);
  V diff = vdiv(vmul(vs(vsub(v3(1, 1, 1), fr), (R)1 - m->metal), m->base), PI_R);
  diff = mixv(diff, vmul(diff, m->subc), clampr(m->sub, (R)0, (R)1));
  R sv5 = pow5((R)1 - vh);
  V sheen = vs(m->sheen, ((R)1 - clampr(m->sheen_r, (R)0, (R)1)) * sv5);
  R cc = clampr(m->cc, (R)0, (R)1), cr = clampr(m->ccr, (R)0.02, (R)1);
  R cD = ggx_d(nh, cr), cG = ggx_g(nv, cr) * ggx_g(nl, cr);
  R cF = (R)0.04 + (R)0.96 * sv5;
  R coat = (((cc * cD) * cG) * cF) / den;
  R layer = (R)1 - cc * cF;
  V dl = vs(diff, layer);
  V sl = vadd(vs(vadd(spec, sheen), layer), v3(coat, coat, coat));
  return vadd(dl, sl);
}

static R pdf_pbr(const Mat* m, V n, V view, V out) {
  R nl = rmax(dot(n, out), (R)0);
  if (nl <= (R)0) return (R)0;
  V h = norm(vadd(view, out));
  R nh = rmax(dot(n, h), (R)0), vh = rmax(dot(view, h), (R)1e-6);
  R sp = (ggx_d(nh, m->rough) * nh) / ((R)4 * vh);
  R dp = nl / PI_R;
  return mixr(dp, sp, spec_prob(m));
}

static V tangent_of(V n) {
  V t = cross(n, v3(0, 1, 0));
  if (absr(n.z) < (R)0.999) t = cross(n, v3(0, 0, 1));
  return norm(t);
}
static V cosine_dir(V n, R u, R v) {
  R radius = rsqrt_(u), phi = TWO_PI_R * v, sp, cp;
  V t = tangent_of(n), b = cross(n, t);
  sincos_r(phi, &sp, &cp);
  return norm(vadd(vadd(vs(t, radius * cp), vs(b, radius * sp)), vs(n, rsqrt_(rmax((R)0, (R)1 - u)))));
}
static V ggx_half(V n, R rough, R u, R v) {
  R alpha = rmax(rough * rough, (R)0.0009), a2 = alpha * alpha, phi = TWO_PI_R * v;
  R c = rsqrt_(((R)1 - u) / rmax((R)1 + (a2 - (R)1) * u, (R)1e-6));
  R s = rsqrt_(rmax((R)0, (R)1 - c * c));
  V t = tangent_of(n), b = cross(n, t);
  R sp, cp;
  sincos_r(phi, &sp, &cp);
  return norm(vadd(vadd(vs(vs(t, s), cp), vs(vs(b, s), sp)), vs(n, c)));
}
static R power_h(R a, R b) {
  R a2 = a * a, b2 = b * b;
  return a2 / rmax(a2 + b2, (R)1e-6);
}
/* ---- texture sampling (wavefront_shade_candidate.glsl:678-881) ----------- */
static R r_floor(R x) { return sizeof(R) == 4 ? (R)floorf((float)x) : (R)floor((double)x); }
static R r_pow(R x, R y) { return sizeof(R) == 4 ? (R)powf((float)x, (float)y) : (R)pow((double)x, (double)y); }
static R r_log2(R x) { return sizeof(R) == 4 ? (R)log2f((float)x) : (R)log2((double)x); }
static R r_exp2(R x) { return sizeof(R) == 4 ? (R)exp2f((float)x) : (R)exp2((double)x); }
static R r_atan2(R y, R x) { return sizeof(R) == 4 ? (R)atan2f((float)y, (float)x) : (R)atan2((double)y, (double)x); }
static R r_acos(R x) { return sizeof(R) == 4 ? (R)acosf((float)x) : (R)acos((double)x); }

typedef struct {
  R x, y, z, w;
} V4;
static V4 v4mix(V4 a, V4 b, R t) {
  V4 r = {mixr(a.x, b.x, t), mixr(a.y, b.y, t), mixr(a.z, b.z, t), mixr(a.w, b.w, t)};
  return r;
}
/* shadeSrgbChannel :678 */
static R srgb_ch(R v) { return v <= (R)0.04045 ? v / (R)12.92 : r_pow((v + (R)0.055) / (R)1.055, (R)2.4); }
/* shadeWrapTextureCoordinate :687 */
static R wrap_c(R v, uint32_t mode) {
  if (mode == 1u) return clampr(v, (R)0, (R)0.99999994f);
  if (mode == 2u) {
    R period = v - r_floor(v / (R)2) * (R)2;
    if (period < (R)0) period = period + (R)2;
    return period <= (R)1 ? period : (R)2 - period;
  }
  return v - r_floor(v);
}
/* shadeWrapTextureIndex :708 */
static int wrap_i(int v, int size, uint32_t mode) {
  if (mode == 1u) return v < 0 ? 0 : (v > size - 1 ? size - 1 : v);
  int period = mode == 2u ? size * 2 : size;
  int w = v % period;
  if (w < 0) w += period;
  if (mode == 2u && w >= size) return (period - 1) - w;
  return w;
}
/* shadeDecodeTextureTexel / shadeFetchTextureTexel :730-747 */
static V4 fetch(const Scene* s, uint32_t off, int w, int h, int x, int y, uint32_t wx, uint32_t wy, int srgb) {
  uint32_t pk = s->words[off + (uint32_t)(wrap_i(y, h, wy) * w + wrap_i(x, w, wx))];
  V4 c = {(R)(float)(pk & 255u) / (R)255, (R)(float)((pk >> 8) & 255u) / (R)255,
          (R)(float)((pk >> 16) & 255u) / (R)255, (R)(float)(pk >> 24) / (R)255};
  if (srgb && !OL_MUT(14)) {
    c.x = srgb_ch(c.x);
    c.y = srgb_ch(c.y);
    c.z = srgb_ch(c.z);
  }
  return c;
}
/* shadeTextureMipOffset / shadeSampleTextureLevel :749-782 */
static V4 sample_lv(const Scene* s, uint32_t base, int bw, int bh, uint32_t level, R u, R v, uint32_t wx, uint32_t wy,
                    int srgb, int linear) {
  uint32_t off = base;
  int w = bw, h = bh;
  for (uint32_t k = 0; k < level && k < 32u; ++k) {
    off += (uint32_t)(w * h);
    w = (w + 1) / 2 > 1 ? (w + 1) / 2 : 1;
    h = (h + 1) / 2 > 1 ? (h + 1) / 2 : 1;
  }
  R wu = wrap_c(u, wx), wv = wrap_c(v, wy);
  if (!linear) return fetch(s, off, w, h, (int)r_floor(wu * (R)w), (int)r_floor(wv * (R)h), wx, wy, srgb);
  R half = OL_MUT(13) ? (R)0 : (R)0.5;
  R tx = wu * (R)w - half, ty = wv * (R)h - half;
  int x0 = (int)r_floor(tx), y0 = (int)r_floor(ty);
  R fx = tx - r_floor(tx), fy = ty - r_floor(ty);
  V4 top = v4mix(fetch(s, off, w, h, x0, y0, wx, wy, srgb), fetch(s, off, w, h, x0 + 1, y0, wx, wy, srgb), fx);
  V4 bot = v4mix(fetch(s, off, w, h, x0, y0 + 1, wx, wy, srgb), fetch(s, off, w, h, x0 + 1, y0 + 1, wx, wy, srgb), fx);
  return v4mix(top, bot, fy);
}
/* shadeSampleMaterialTexture :784-827 */
static V4 sample_mat_tex(const Scene* s, R binding, R u0, R v0, R u1, R v1, int srgb, R fp0, R fp1) {
  V4 white = {1, 1, 1, 1};
  int b = (int)binding;
  if (b < 0) return white;
  const float* r = s->bind + 8ull * (uint32_t)b;
  int t = (int)r[0];
  if (t < 0 || (uint32_t)t >= s->words[0]) return white;
  int use1 = r[3] > 0.5f;
  R u = use1 ? u1 : u0, v = use1 ? v1 : v0, fp = use1 ? fp1 : fp0;
  R cs = r[1], sn = r[2];
  R su = u * (R)r[6], sv_ = v * (R)r[7];
  u = (R)r[4] + (cs * su - sn * sv_);
  v = (R)r[5] + (sn * su + cs * sv_);
  if (!OL_MUT(15)) fp = fp * rmax(absr((R)r[6]), absr((R)r[7]));
  uint32_t meta = 1u + (uint32_t)t * 8u;
  uint32_t off = s->words[meta + (srgb ? 0u : 1u)];
  int w = (int)s->words[meta + 2u], h = (int)s->words[meta + 3u];
  uint32_t flags = s->words[meta + 4u], levels = s->words[meta + 5u];
  R lod = clampr(r_log2(rmax(fp * (R)(w > h ? w : h), (R)1)), (R)0, (R)((levels > 1u ? levels : 1u) - 1u));
  uint32_t lower = (uint32_t)r_floor(lod);
  int linear = (flags & 16u) != 0u;
  V4 first = sample_lv(s, off, w, h, lower, u, v, flags & 3u, (flags >> 2) & 3u, srgb, linear);
  if (!linear || lower + 1u >= levels) return first;
  V4 second = sample_lv(s, off, w, h, lower + 1u, u, v, flags & 3u, (flags >> 2) & 3u, srgb, 1);
  return v4mix(first, second, lod - r_floor(lod));
}

static V env_rad(const Scene* s, V d) {
  if (!s->env_valid) return mixv(v3((R)0.018, (R)0.022, (R)0.032), v3((R)0.32, (R)0.46, (R)0.72), rmax(d.y, (R)0));
  if (s->env_tex >= 0 && s->words && (uint32_t)s->env_tex < s->words[0]) {
    /* shadeEnvironmentUv, shadeSamplePackedSceneTexture, shadeDecodeEnvironmentRadiance :1835-1894 */
    R lon = r_atan2(d.z, d.x) + (R)s->env_rot;
    R x = lon * (R)0.15915494309189535 + (R)0.5;
    R u = x - r_floor(x);
    R v = r_acos(clampr(d.y, (R)-1, (R)1)) * (R)0.3183098861837907;
    uint32_t meta = 1u + (uint32_t)s->env_tex * 8u, flags = s->words[meta + 4u];
    V4 c = sample_lv(s, s->words[meta + 1u], (int)s->words[meta + 2u], (int)s->words[meta + 3u], 0u, u, v, flags & 3u,
                     (flags >> 2) & 3u, 0, (flags & 16u) != 0u);
    V rad = v3(r_exp2(c.x * (R)s->env_log) - (R)1, r_exp2(c.y * (R)s->env_log) - (R)1,
               r_exp2(c.z * (R)s->env_log) - (R)1);
    return vs(vmul(rad, v3(s->env_col[0], s->env_col[1], s->env_col[2])), (R)s->env_int);
  }
  return v3(s->env[0], s->env[1], s->env[2]);
}

typedef struct {
  uint32_t max_bounces, sec_area, env_samples, rr_start;
  float rr_min, nee_prob;
  uint32_t unified;
} Settings;

static R unified_prob(const Scene* s, const Settings* st) {
  if (s->L == 0) return (R)0;
  if (st->env_samples == 0) return (R)1;
  R w = rsqrt_(rmax((R)s->weight, (R)1e-6));
  return w / (w + (R)1);
}

/* Shadow-tested NEE sample contribution (0 when invalid or occluded). */
static int area_sample(const Scene* s, uint32_t* rs, V hit, V n, const Mat* m, V inc, uint32_t si, uint32_t sc,
                       R tp, V* out) {
  if (s->L == 0) return 0;
  *rs = pcg(*rs);
  R sel = ((R)si + pcg_unit(*rs)) / (R)sc;
  uint32_t lo = 0, hi = s->L - 1;
  for (int k = 0; k < 32 && lo < hi; ++k) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (sel <= (R)s->area[20ull * mid + 16]) hi = mid; else lo = mid + 1;
  }
  const float* A = s->area + 20ull * lo;
  V a = v3(A[0], A[1], A[2]), b = v3(A[4], A[5], A[6]), c = v3(A[8], A[9], A[10]);
  *rs = pcg(*rs);
  R ru = rsqrt_(pcg_unit(*rs));
  *rs = pcg(*rs);
  R vv = pcg_unit(*rs);
  V lp = vadd(vadd(sv((R)1 - ru, a), sv(ru * ((R)1 - vv), b)), sv(ru * vv, c));
  V off = vsub(lp, hit);
  R d2 = dot(off, off), dl = rsqrt_(d2);
  V dir = vdiv(off, rmax(dl, (R)1e-6));
  R scos = rmax(dot(n, dir), (R)0);
  V ln = norm(cross(vsub(b, a), vsub(c, a)));
  R raw = dot(ln, vneg(dir));
  R lc = rmax(raw, (R)0);
  if (A[18] > 0.5f) lc = absr(raw);
  if (scos <= (R)0 || lc <= (R)1e-6) return 0;
  R lpdf = ((R)A[17] * (OL_MUT(8) ? (R)1 : d2)) / rmax(lc * (R)A[15], (R)1e-6);
  R eff = lpdf * rmax(tp, (R)1e-6);
  if (occluded(s, vadd(hit, vs(n, (R)0.002)), dir, (R)0.001, rmax(dl - (R)0.004, (R)0.001))) return 0;
  V view = vneg(inc);
  R bp = pdf_pbr(m, n, view, dir);
  R mis = power_h(eff * (R)sc, bp);
  *out = vdiv(vs(vs(vs(vmul(eval_pbr(m, n, view, dir), v3(A[12], A[13], A[14])), scos), mis), (R)1),
              rmax(eff, (R)1e-6));
  return 1;
}

static int env_sample(const Scene* s, uint32_t* rs, V hit, V n, const Mat* m, V inc, uint32_t sc, R tp, V* out) {
  *rs = pcg(*rs);
  R u = pcg_unit(*rs);
  *rs = pcg(*rs);
  R v = pcg_unit(*rs);
  V dir = cosine_dir(n, u, v);
  R c = rmax(dot(n, dir), (R)0);
  R pdf = c / PI_R, eff = pdf * rmax(tp, (R)1e-6);
  if (!(c > (R)0)) return 0;
  if (occluded(s, vadd(hit, vs(n, (R)0.002)), dir, (R)0.001, (R)1e30)) return 0;
  R mis = OL_MUT(9) ? (R)1 : power_h(eff * (R)sc, c / PI_R);
  R mw = mixr(m->occl, (R)1, m->metal);
  *out = vs(vdiv(vs(vs(vs(vmul(eval_pbr(m, n, vneg(inc), dir), env_rad(s, dir)), c), mis), (R)1), rmax(eff, (R)1e-6)),
            mw);
  return 1;
}

static int point_sample(const Scene* s, uint32_t li, V hit, V n, const Mat* m, V inc, V* out) {
  const float* L = s->ana + 16ull * li;
  int type = (int)(L[3] + 0.5f);
  if (type == 3) return 0;
  V dir;
  R d2 = 1, dist = 10000, att = 1;
  if (type == 1) {
    dir = vneg(norm(v3(L[4], L[5], L[6])));
  } else {
    V off = vsub(v3(L[0], L[1], L[2]), hit);
    d2 = rmax(dot(off, off), (R)1e-6);
    dist = rsqrt_(d2);
    dir = vdiv(off, dist);
    if (L[7] > 0.0f && dist > (R)L[7]) return 0;
    att = OL_MUT(2) ? (R)1 : (R)1 / d2;
    if (type == 2) {
      R cone = dot(norm(v3(L[4], L[5], L[6])), vneg(dir));
      R amount = clampr((cone - (R)L[13]) / rmax((R)L[12] - (R)L[13], (R)1e-6), (R)0, (R)1);
      R spot = OL_MUT(10) ? amount : (amount * amount) * ((R)3 - (R)2 * amount);
      if (spot <= (R)0) return 0;
      att = att * spot;
    }
  }
  R c = rmax(dot(n, dir), (R)0);
  if (c <= (R)0) return 0;
  R sd = rmax(dist - (R)0.004, (R)0.001);
  if (type == 1) sd = 10000;
  if (occluded(s, vadd(hit, vs(n, (R)0.002)), dir, (R)0.001, sd)) return 0;
  V incident = vs(vs(v3(L[8], L[9], L[10]), (R)L[11]), att);
  *out = vs(vmul(vmul(eval_pbr(m, n, vneg(inc), dir), incident), v3(1, 1, 1)), OL_MUT(3) ? (R)1 : c);
  return 1;
}

/* ---- ReSTIR DI at primary hits ----------------------------------------------
 * Reservoir arithmetic of wavefront_restir.glsl:51-180 (update :51, merge
 * :68, canonical :83, pairwise :94, normalisation :124, limit :134, stored
 * form :150-180), candidates of transport_v1/lighting.glsl:441-552 and the
 * primary reuse loop of wavefront_primary_impl.glsl:1539-1700 (reprojection
 * :350-365, history position :367-378, surface test :382-393, spatial
 * offsets :319-331).  Halves are IEEE binary16 with round-to-nearest-even,
 * as packHalf2x16. */
#define OL_NONE_ID 0xFFFFFFFFu

uint16_t API(olgi_f2h)(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
  const uint32_t e = ax >> 23, m = (ax & 0x7FFFFFu) | 0x800000u;
  if (e > 142u) return (uint16_t)(sign | 0x7C00u);
  if (e < 113u) { /* half subnormal: round(value * 2^24) */
    if (e < 102u) return (uint16_t)sign;
    const uint32_t sh = 126u - e, q = m >> sh, rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    return (uint16_t)(sign | (q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u)));
  }
  uint32_t hbits = ((e - 112u) << 10) | ((ax & 0x7FFFFFu) >> 13);
  const uint32_t rem = ax & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (hbits & 1u))) hbits++;
  return (uint16_t)(sign | hbits);
}
float API(olgi_h2f)(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
  float v;
  if (e == 0u) v = ldexpf((float)m, -24);
  else if (e == 31u) v = m ? NAN : INFINITY;
  else v = ldexpf((float)(m | 1024u), (int)e - 25);
  uint32_t b;
  memcpy(&b, &v, 4);
  b |= sign;
  memcpy(&v, &b, 4);
  return v;
}
static uint32_t pack_h2(R a, R b) {
  return (uint32_t)API(olgi_f2h)((float)a) | ((uint32_t)API(olgi_f2h)((float)b) << 16);
}
static R h_lo(uint32_t v) { return (R)API(olgi_h2f)((uint16_t)(v & 0xFFFFu)); }
static R h_hi(uint32_t v) { return (R)API(olgi_h2f)((uint16_t)(v >> 16)); }

typedef struct {
  uint32_t light, bary; /* light (OL_NONE_ID: empty), half2 barycentrics */
  R w_sum;
  uint32_t tc;          /* half2 (target, represented samples) */
} Rsv;

static Rsv rsv_empty(void) {
  Rsv r = {OL_NONE_ID, 0u, (R)0, 0u};
  return r;
}
static int rsv_update(Rsv* r, uint32_t light, R b1, R b2, R target, R weight, R represented, R random) {
  weight = rmax(weight, (R)0);
  const R w_sum = r->w_sum + weight;
  R t = h_lo(r->tc);
  const R count = h_hi(r->tc) + represented;
  const int sel = weight > (R)0 && (r->light == OL_NONE_ID || random * w_sum < weight);
  if (sel) {
    r->light = light;
    r->bary = pack_h2(clampr(b1, 0, 1), clampr(b2, 0, 1));
    t = rmin(rmax(target, (R)0), (R)65504);
  }
  r->w_sum = w_sum;
  r->tc = pack_h2(t, count);
  return sel;
}
static void rsv_merge(Rsv* d, const Rsv* src, R target_here, R random) {
  const R st = h_lo(src->tc), sc = h_hi(src->tc);
  if (src->light == OL_NONE_ID || sc <= (R)0) return;
  R w = 0;
  if (st > (R)0) w = (rmax(target_here, (R)0) * src->w_sum) / st;
  rsv_update(d, src->light, h_lo(src->bary), h_hi(src->bary), target_here, w, sc, random);
}
static void rsv_merge_canonical(Rsv* d, const Rsv* src, R target_here, R random) {
  const R st = h_lo(src->tc), sc = h_hi(src->tc);
  if (src->light == OL_NONE_ID || st <= (R)0 || sc <= (R)0) return;
  const R w = (rmax(target_here, (R)0) * src->w_sum) / (st * sc);
  rsv_update(d, src->light, h_lo(src->bary), h_hi(src->bary), target_here, w, 1, random);
}
static void rsv_merge_pairwise(Rsv* d, const Rsv* src, R target_here, R target_there, R random) {
  const R st = h_lo(src->tc), sc = h_hi(src->tc);
  if (src->light == OL_NONE_ID || st <= (R)0 || sc <= (R)0) return;
  const R total = rmax(target_here, (R)0) + rmax(target_there, (R)0);
  R balance = 0;
  if (total > (R)0) balance = ((R)2 * rmax(target_there, (R)0)) / total;
  const R w = ((rmax(target_here, (R)0) * src->w_sum) / (st * sc)) * balance;
  rsv_update(d, src->light, h_lo(src->bary), h_hi(src->bary), target_here, w, 1, random);
}
static R rsv_normalization(const Rsv* r) {
  const R t = h_lo(r->tc), c = h_hi(r->tc);
  if (r->light == OL_NONE_ID || t <= (R)0 || c <= (R)0) return 0;
  return r->w_sum / (c * t);
}
static Rsv rsv_limit(Rsv r, R maximum) {
  const R t = h_lo(r.tc), c = h_hi(r.tc);
  maximum = rmax(maximum, (R)0);
  if (c > maximum && c > (R)0) {
    const R scale = maximum / c;
    r.w_sum = r.w_sum * scale;
    r.tc = pack_h2(t, maximum);
  }
  return r;
}
/* Stored form: light | count << 25, half2 barycentrics, half2 (w_sum, target). */
static void rsv_store(const Rsv* r, uint32_t out[4]) {
  const R t = h_lo(r->tc), c = h_hi(r->tc);
  uint32_t light = 33554431u;
  if (r->light != OL_NONE_ID) light = r->light < 33554430u ? r->light : 33554430u;
  const R rc = sizeof(R) == 4 ? (R)rintf((float)c) : (R)rint((double)c);
  const uint32_t count = (uint32_t)rmin(rmax(rc, (R)0), (R)127);
  const R w = rmin(rmax(r->w_sum, (R)0), (R)65504);
  out[0] = light | (count << 25);
  out[1] = r->bary;
  out[2] = pack_h2(w, t);
  out[3] = 0u;
}
static Rsv rsv_load(const uint32_t w[4]) {
  uint32_t light = w[0] & 33554431u;
  const uint32_t count = w[0] >> 25;
  if (light == 33554431u) light = OL_NONE_ID;
  Rsv r = {light, w[1], h_lo(w[2]), pack_h2(h_hi(w[2]), (R)count)};
  return r;
}

/* Reservoir entry points for the tests: a reservoir in register form is
 * (light, half2 bary, w_sum float bits, half2 (target, count)). */
static Rsv rsv_from(const uint32_t w[4]) {
  float f;
  memcpy(&f, &w[2], 4);
  Rsv r = {w[0], w[1], (R)f, w[3]};
  return r;
}
static void rsv_to(const Rsv* r, uint32_t w[4]) {
  const float f = (float)r->w_sum;
  w[0] = r->light;
  w[1] = r->bary;
  memcpy(&w[2], &f, 4);
  w[3] = r->tc;
}
/* n updates from an empty reservoir; selected[i] = 1 when candidate i was taken. */
void API(olgi_rsv_updates)(uint32_t n, const uint32_t* light, const float* b1, const float* b2, const float* target,
                           const float* weight, const float* represented, const float* random, uint32_t out[4],
                           uint8_t* selected) {
  Rsv r = rsv_empty();
  for (uint32_t i = 0; i < n; ++i) {
    const int s = rsv_update(&r, light[i], b1[i], b2[i], target[i], weight[i], represented[i], random[i]);
    if (selected) selected[i] = (uint8_t)s;
  }
  rsv_to(&r, out);
}
/* mode 0: temporal merge, 1: pairwise, 2: canonical; 3: limit(src, target_here). */
void API(olgi_rsv_merge)(uint32_t dst[4], const uint32_t src[4], float target_here, float target_there, float random,
                         uint32_t mode) {
  Rsv d = rsv_from(dst), s = rsv_from(src);
  if (mode == 0u) rsv_merge(&d, &s, target_here, random);
  else if (mode == 1u) rsv_merge_pairwise(&d, &s, target_here, target_there, random);
  else if (mode == 2u) rsv_merge_canonical(&d, &s, target_here, random);
  else d = rsv_limit(s, target_here);
  rsv_to(&d, dst);
}
float API(olgi_rsv_normalization)(const uint32_t r[4]) {
  Rsv x = rsv_from(r);
  return (float)rsv_normalization(&x);
}
void API(olgi_rsv_store)(const uint32_t r[4], uint32_t stored[4]) {
  Rsv x = rsv_from(r);
  rsv_store(&x, stored);
}
void API(olgi_rsv_load)(const uint32_t stored[4], uint32_t r[4]) {
  Rsv x = rsv_load(stored);
  rsv_to(&x, r);
}

/* evaluateAreaLightCandidate: MIS-weighted NEE contribution of the point at
 * (b1, b2) on area light `li` (technique probability 1). */
static V cand_contribution(const Scene* s, uint32_t li, R b1, R b2, V hit, V n, V inc, const Mat* m,
                           uint32_t count, V* dir, R* dist) {
  *dir = n;
  *dist = 0;
  if (li >= s->L) return v3(0, 0, 0);
  const float* A = s->area + 20ull * li;
  V a = v3(A[0], A[1], A[2]), b = v3(A[4], A[5], A[6]), c = v3(A[8], A[9], A[10]);
  const R aw = rmax(((R)1 - b1) - b2, (R)0);
  V p = vadd(vadd(vs(a, aw), vs(b, b1)), vs(c, b2));
  V off = vsub(p, hit);
  const R d2 = dot(off, off);
  *dist = rsqrt_(d2);
  *dir = vdiv(off, rmax(*dist, (R)1e-6));
  const R cs = rmax(dot(n, *dir), (R)0);
  V ln = norm(cross(vsub(b, a), vsub(c, a)));
  const R raw = dot(ln, vneg(*dir));
  const R cl = A[18] > 0.5f ? absr(raw) : rmax(raw, (R)0);
  if (cs <= (R)0 || cl <= (R)1e-6) return v3(0, 0, 0);
  const R eff = (((R)A[17] * d2) / rmax(cl * (R)A[15], (R)1e-6)) * rmax((R)1, (R)1e-6);
  const R bp = pdf_pbr(m, n, vneg(inc), *dir);
  const R first = eff * (R)count, fs = first * first;
  const R mis = fs / rmax(fs + bp * bp, (R)1e-6);
  return vdiv(vs(vs(vs(vmul(eval_pbr(m, n, vneg(inc), *dir), v3(A[12], A[13], A[14])), cs), mis), (R)1),
              rmax(eff, (R)1e-6));
}
static R cand_target(V c) { return rmax(dot(c, v3((R)0.2126, (R)0.7152, (R)0.0722)), (R)0); }

/* ReSTIR settings and the reservoir/surface buffers of one pass.  Reservoirs
 * are 4 words per pixel; surfaces 8 floats per pixel (shading normal, primary
 * distance or -1; material id bits). */
typedef struct {
  uint32_t candidates, spatial, radius, pairwise, history;
  R limit;
  const float* cam;
  uint32_t proj, W, H, frame, pass;
  const uint32_t* res_in;
  const float* surf_in;
  uint32_t* res_out;
  float* surf_out;
} Restir;

static int restir_reproject(const Restir* rc, V p, int* px, int* py) {
  const float* c = rc->cam;
  V off = vsub(p, v3(c[0], c[1], c[2]));
  V fw = v3(c[3], c[4], c[5]), rt = v3(c[6], c[7], c[8]), up = v3(c[9], c[10], c[11]);
  const R depth = dot(off, fw), scale = len(up), aspect = (R)rc->W / (R)rc->H;
  if (depth <= (R)0.0001 || scale <= (R)0.0001) return 0;
  const R ps = rc->proj == 1u ? (R)1 : depth;
  const R nx = dot(off, norm(rt)) / ((ps * aspect) * scale);
  const R ny = (-dot(off, norm(up))) / (ps * scale);
  R fx = ((nx * (R)0.5 + (R)0.5) * (R)rc->W) - (R)0.5;
  R fy = ((ny * (R)0.5 + (R)0.5) * (R)rc->H) - (R)0.5;
  fx = sizeof(R) == 4 ? (R)rintf((float)fx) : (R)rint((double)fx);
  fy = sizeof(R) == 4 ? (R)rintf((float)fy) : (R)rint((double)fy);
  if (!(fx >= (R)0 && fy >= (R)0 && fx < (R)rc->W && fy < (R)rc->H)) return 0;
  *px = (int)fx;
  *py = (int)fy;
  return 1;
}
static V restir_previous_position(const Restir* rc, int px, int py, R dist) {
  const float* c = rc->cam;
  const R ndx = (((R)px + (R)0.5) / (R)rc->W) * (R)2 - (R)1;
  const R ndy = (((R)py + (R)0.5) / (R)rc->H) * (R)2 - (R)1;
  const R aspect = (R)rc->W / (R)rc->H;
  V org = v3(c[0], c[1], c[2]), fw = v3(c[3], c[4], c[5]), rt = v3(c[6], c[7], c[8]), up = v3(c[9], c[10], c[11]);
  if (rc->proj == 1u) return vadd(vsub(vadd(org, vs(rt, ndx * aspect)), vs(up, ndy)), vs(norm(fw), dist));
  return vadd(org, vs(norm(vsub(vadd(fw, vs(rt, ndx * aspect)), vs(up, ndy))), dist));
}
static int restir_surface_matches(const Restir* rc, int hx, int hy, V p, V n, R dist, uint32_t mat_id, int center) {
  const float* sf = rc->surf_in + 8ull * ((uint64_t)hy * rc->W + (uint64_t)hx);
  uint32_t old_mat;
  memcpy(&old_mat, sf + 4, 4);
  V delta = vsub(restir_previous_position(rc, hx, hy, (R)sf[3]), p);
  const R tol = rmax((R)0.03, dist * (R)0.01);
  const R err = center ? len(delta) : absr(dot(delta, norm(n)));
  const R agree = dot(norm(v3(sf[0], sf[1], sf[2])), norm(n));
  return sf[3] >= 0.0f && err <= tol && agree > (R)0.9 && old_mat == mat_id;
}
static void restir_clear_pixel(const Restir* rc, uint32_t pixel) {
  Rsv e = rsv_empty();
  rsv_store(&e, rc->res_out + 4ull * pixel);
  float* sf = rc->surf_out + 8ull * pixel;
  const uint32_t none = OL_NONE_ID;
  sf[0] = sf[1] = sf[2] = 0.0f;
  sf[3] = -1.0f;
  memcpy(sf + 4, &none, 4);
  sf[5] = sf[6] = sf[7] = 0.0f;
}

/* Candidates, temporal and spatial merges of the previous pass's reservoirs,
 * store; returns 1 with the selected point's shadow ray and its contribution
 * times the reservoir normalisation (unoccluded) when a light was selected. */
static int restir_primary(const Scene* s, const Restir* rc, uint32_t* rs, V hit, V n, const Mat* m, V inc, R dist,
                          uint32_t pixel, uint32_t mat_id, V* c_out, V* dir_out, R* tmax_out) {
  const uint32_t count = rc->candidates;
  Rsv r = rsv_empty();
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t light = OL_NONE_ID;
    R b1 = 0, b2 = 0, target = 0;
    if (s->L > 0) {
      *rs = pcg(*rs);
      const R sel = ((R)k + pcg_unit(*rs)) / (R)count;
      uint32_t lo = 0, hi = s->L - 1;
      for (int st = 0; st < 32 && lo < hi; ++st) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (sel <= (R)s->area[20ull * mid + 16]) hi = mid; else lo = mid + 1;
      }
      *rs = pcg(*rs);
      const R ru = rsqrt_(pcg_unit(*rs));
      *rs = pcg(*rs);
      const R vv = pcg_unit(*rs);
      light = lo;
      b1 = ru * ((R)1 - vv);
      b2 = ru * vv;
      V d;
      R dd;
      target = cand_target(cand_contribution(s, light, b1, b2, hit, n, inc, m, count, &d, &dd));
    }
    *rs = pcg(*rs);
    rsv_update(&r, light, b1, b2, target, target, 1, pcg_unit(*rs));
  }
  int px, py;
  if (rc->history && restir_reproject(rc, hit, &px, &py)) {
    const uint32_t rot = hash32(pixel ^ (rc->frame + rc->pass)) & 7u;
    static const int ox[8] = {1, -1, 0, 0, 1, -1, 1, -1}, oy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
    for (uint32_t ri = 0; ri <= rc->spatial; ++ri) {
      int hx = px, hy = py;
      if (ri > 0u) {
        const uint32_t idx = ((ri - 1u) + rot) & 7u;
        hx += ox[idx] * (int)rc->radius;
        hy += oy[idx] * (int)rc->radius;
      }
      const int inside = hx >= 0 && hy >= 0 && hx < (int)rc->W && hy < (int)rc->H;
      Rsv h = rsv_empty();
      int present = 0, valid = 0;
      if (inside) {
        const uint64_t hidx = (uint64_t)hy * rc->W + (uint64_t)hx;
        h = rsv_load(rc->res_in + 4ull * hidx);
        present = h.light != OL_NONE_ID;
        if (present) valid = restir_surface_matches(rc, hx, hy, hit, n, dist, mat_id, ri == 0u);
      }
      if (valid) {
        const R remaining = rmax(rc->limit - h_hi(r.tc), (R)0);
        h = rsv_limit(h, rc->spatial > 0u ? rmin(remaining, (R)1) : remaining);
        if (h_hi(h.tc) <= (R)0) h = rsv_empty();
      } else {
        h = rsv_empty();
      }
      if (h.light != OL_NONE_ID) {
        V d;
        R dd;
        const R ht = cand_target(cand_contribution(s, h.light, h_lo(h.bary), h_hi(h.bary), hit, n, inc, m, count, &d, &dd));
        *rs = pcg(*rs);
        const R u = pcg_unit(*rs);
        if (ri == 0u) rsv_merge(&r, &h, ht, u);
        else if (rc->pairwise) rsv_merge_pairwise(&r, &h, ht, h_lo(h.tc), u);
        else rsv_merge_canonical(&r, &h, ht, u);
      } else if (ri == 0u && !present) {
        break;
      }
    }
  }
  rsv_store(&r, rc->res_out + 4ull * pixel);
  float* sf = rc->surf_out + 8ull * pixel;
  sf[0] = (float)n.x;
  sf[1] = (float)n.y;
  sf[2] = (float)n.z;
  sf[3] = (float)dist;
  memcpy(sf + 4, &mat_id, 4);
  sf[5] = sf[6] = sf[7] = 0.0f;
  if (r.light == OL_NONE_ID) return 0;
  V d;
  R dd;
  V c = cand_contribution(s, r.light, h_lo(r.bary), h_hi(r.bary), hit, n, inc, m, count, &d, &dd);
  if (dd <= (R)0) return 0;
  *c_out = vs(c, rsv_normalization(&r));
  *dir_out = d;
  *tmax_out = rmax(dd - (R)0.004, (R)0.001);
  return 1;
}

/* Probes of the sampler for the closed-form tests. */
void API(olgi_texture_sample)(void* h, float binding, float u0, float v0, float u1, float v1, int srgb, float fp0,
                              float fp1, float out[4]) {
  V4 c = sample_mat_tex((const Scene*)h, binding, u0, v0, u1, v1, srgb, fp0, fp1);
  out[0] = (float)c.x;
  out[1] = (float)c.y;
  out[2] = (float)c.z;
  out[3] = (float)c.w;
}
void API(olgi_env_radiance)(void* h, const float d[3], float out[3]) {
  V c = env_rad((const Scene*)h, v3(d[0], d[1], d[2]));
  out[0] = (float)c.x;
  out[1] = (float)c.y;
  out[2] = (float)c.z;
}

/* ---- textured surfaces (shadeResolveSurface :957-997) -------------------- */
static int mat_textured(const Scene* s, uint32_t m) { return s->mat[48ull * m + 35] == 1.0f; }

/* shadeTriangleUvDensity :872-879 */
static R uv_dens(V a, V b, V c, R ua, R va, R ub, R vb, R uc, R vc) {
  R fx = ub - ua, fy = vb - va, sx = uc - ua, sy = vc - va;
  R uv_area = absr(fx * sy - fy * sx);
  R world = len(cross(vsub(b, a), vsub(c, a)));
  return rsqrt_(uv_area / rmax(world, (R)1e-8));
}

/* Material textures and normal map of a hit (shadeApplyMaterialTextures
 * :915-934, shadeTriangleTangent :889-913, shadeApplyNormalTexture :936-955):
 * updates m (base colour, roughness, metallic, transmission, occlusion), the
 * emission *em and the shading normal *sn (on the geometric side). */
static void tex_surface(const Scene* s, uint32_t mid, uint32_t prim, R bu, R bv, V gn, R cone_w, Mat* m, V* em,
                        V* sn) {
  const float* rec = s->mat + 48ull * mid;
  const uint32_t* tv = s->idx + 3ull * prim;
  R wx = ((R)1 - bu) - bv, wy = bu, wz = bv;
  const float *qa = s->uvs + 4ull * tv[0], *qb = s->uvs + 4ull * tv[1], *qc = s->uvs + 4ull * tv[2];
  R u0 = (((R)qa[0] * wx) + ((R)qb[0] * wy)) + ((R)qc[0] * wz);
  R v0 = (((R)qa[1] * wx) + ((R)qb[1] * wy)) + ((R)qc[1] * wz);
  R u1 = 0, v1 = 0, fp1 = 0;
  R fp0 = cone_w * (R)s->dens[prim];
  const float *pa = s->pos + 3ull * tv[0], *pb = s->pos + 3ull * tv[1], *pc = s->pos + 3ull * tv[2];
  V A = v3(pa[0], pa[1], pa[2]), B = v3(pb[0], pb[1], pb[2]), C = v3(pc[0], pc[1], pc[2]);
  R idz = rec[14];
  if (idz - r_floor(idz) > (R)0.125) {
    u1 = (((R)qa[2] * wx) + ((R)qb[2] * wy)) + ((R)qc[2] * wz);
    v1 = (((R)qa[3] * wx) + ((R)qb[3] * wy)) + ((R)qc[3] * wz);
    fp1 = cone_w * uv_dens(A, B, C, qa[2], qa[3], qb[2], qb[3], qc[2], qc[3]);
  }
  m->trans = m->trans * sample_mat_tex(s, rec[23], u0, v0, u1, v1, 0, fp0, fp1).x;
  V4 base = sample_mat_tex(s, rec[16], u0, v0, u1, v1, 1, fp0, fp1);
  m->base = vmul(m->base, v3(base.x, base.y, base.z));
  V4 emit = sample_mat_tex(s, rec[18], u0, v0, u1, v1, 1, fp0, fp1);
  *em = vmul(*em, v3(emit.x, emit.y, emit.z));
  if (!(m->trans > (R)0.001)) {
    V4 mr = sample_mat_tex(s, rec[17], u0, v0, u1, v1, 0, fp0, fp1);
    m->rough = m->rough * mr.y;
    m->metal = m->metal * mr.z;
    R occ = sample_mat_tex(s, rec[21], u0, v0, u1, v1, 0, fp0, fp1).x;
    m->occl = mixr((R)1, occ, (R)rec[22]);
  }
  const float *ta = s->tans + 4ull * tv[0], *tb = s->tans + 4ull * tv[1], *tc = s->tans + 4ull * tv[2];
  V tan = v3((((R)ta[0] * wx) + ((R)tb[0] * wy)) + ((R)tc[0] * wz), (((R)ta[1] * wx) + ((R)tb[1] * wy)) + ((R)tc[1] * wz),
             (((R)ta[2] * wx) + ((R)tb[2] * wy)) + ((R)tc[2] * wz));
  R hand = (((R)ta[3] * wx) + ((R)tb[3] * wy)) + ((R)tc[3] * wz);
  int nb = (int)rec[19];
  V n = *sn;
  if (nb >= 0 && s->bind[8ull * (uint32_t)nb + 3] > 0.5f) {
    R dax = (R)qb[2] - (R)qa[2], day = (R)qb[3] - (R)qa[3], dbx = (R)qc[2] - (R)qa[2], dby = (R)qc[3] - (R)qa[3];
    V ea = vsub(B, A), eb = vsub(C, A);
    R det = dax * dby - day * dbx;
    if (absr(det) < (R)1e-8) {
      V fb = cross(n, v3(0, 1, 0));
      if (absr(n.z) < (R)0.999) fb = cross(n, v3(0, 0, 1));
      tan = norm(fb);
      hand = 1;
    } else {
      R inv = (R)1 / det;
      V t = norm(vs(vsub(vs(ea, dby), vs(eb, day)), inv));
      V bt = norm(vs(vsub(vs(eb, dax), vs(ea, dbx)), inv));
      tan = t;
      hand = dot(cross(n, t), bt) < (R)0 ? (R)-1 : (R)1;
    }
  }
  if (nb >= 0) {
    V4 t4 = sample_mat_tex(s, rec[19], u0, v0, u1, v1, 0, fp0, fp1);
    V tn = v3(t4.x * (R)2 - (R)1, t4.y * (R)2 - (R)1, t4.z * (R)2 - (R)1);
    tn.x = tn.x * (R)rec[20];
    tn.y = tn.y * (R)rec[20];
    tn = norm(tn);
    V t = norm(vsub(tan, vs(n, dot(n, tan))));
    V bt = vs(cross(n, t), OL_MUT(16) ? (R)1 : hand);
    const float* br = s->bind + 8ull * (uint32_t)nb;
    R cs = br[1], snn = br[2];
    R ix = (R)1 / (R)br[6], iy = (R)1 / (R)br[7];
    V tt = norm(vadd(vs(t, cs * ix), vs(bt, (-snn) * iy)));
    V tbn = norm(vadd(vs(t, snn * ix), vs(bt, cs * iy)));
    n = norm(vadd(vadd(vs(tt, tn.x), vs(tbn, tn.y)), vs(n, tn.z)));
  }
  if (dot(n, gn) < (R)0) n = vneg(n);
  *sn = n;
}

/* Probe for the tests: the material and shading normal a hit on triangle
 * `prim` at barycentrics (bu, bv) shades with, for a ray arriving along
 * `dir` with cone width `cone_w` at the hit. out: shading normal xyz, base
 * rgb, roughness, metallic, transmission, occlusion, emission rgb. */
void API(olgi_textured_surface)(void* h, uint32_t prim, float bu, float bv, const float dir[3], float cone_w,
                                float out[14]) {
  const Scene* s = (const Scene*)h;
  const uint32_t* tv = s->idx + 3ull * prim;
  const float *pa = s->pos + 3ull * tv[0], *pb = s->pos + 3ull * tv[1], *pc = s->pos + 3ull * tv[2];
  V e1 = v3((R)(pb[0] - pa[0]), (R)(pb[1] - pa[1]), (R)(pb[2] - pa[2]));
  V e2 = v3((R)(pc[0] - pa[0]), (R)(pc[1] - pa[1]), (R)(pc[2] - pa[2]));
  V gn = norm(cross(e1, e2));
  R wx = ((R)1 - bu) - bv, wy = bu, wz = bv;
  const float *n0 = s->nrm + 3ull * tv[0], *n1 = s->nrm + 3ull * tv[1], *n2 = s->nrm + 3ull * tv[2];
  V sn = norm(vadd(vadd(vs(v3(n0[0], n0[1], n0[2]), wx), vs(v3(n1[0], n1[1], n1[2]), wy)),
                   vs(v3(n2[0], n2[1], n2[2]), wz)));
  if (dot(sn, gn) < (R)0) sn = vneg(sn);
  (void)dir;
  Mat m = load_mat(s, s->mid[prim]);
  V em = emission_of(s, s->mid[prim]);
  if (s->uvs && mat_textured(s, s->mid[prim])) tex_surface(s, s->mid[prim], prim, bu, bv, gn, cone_w, &m, &em, &sn);
  float vals[14] = {(float)sn.x, (float)sn.y, (float)sn.z, (float)m.base.x, (float)m.base.y, (float)m.base.z,
                    (float)m.rough, (float)m.metal, (float)m.trans, (float)m.occl, (float)em.x, (float)em.y,
                    (float)em.z, 0.0f};
  memcpy(out, vals, sizeof(vals));
}

static R medium_get(const R st[16], uint32_t i) { return st[i]; }

/* One path; returns radiance (rgb) as floats of R.  counts[3] += rays. */
static void trace_path(const Scene* s, const Settings* st, const Restir* rc, uint32_t pixel, uint32_t frame,
                       uint32_t sample, V org, V dir, R out[3], uint64_t counts[3]) {
  V T = v3(1, 1, 1), Lr = v3(0, 0, 0);
  R prev_pdf = 0;
  uint32_t flags = 257u, bounce = 0;
  uint32_t rng;
  R stack[16];
  stack[0] = 1;
  {
    uint32_t r = hash32((pixel ^ hash32(frame)) ^ hash32(sample + 1u));
    r = hash32(r);
    r = hash32(r);
    rng = r;
  }
  V o = org, d = dir;
  R tmin = (R)0.001, tmax = (R)1e30;
  R cone_w = 0, cone_s = 0; /* primary rays carry a zero cone (wavefront_generate.comp:153-155) */
  const uint32_t frame_sample = (frame << 8) | (sample & 255u);
  counts[0]++;
  /* This pass's history for the pixel: empty unless ReSTIR runs below. */
  if (rc) restir_clear_pixel(rc, pixel);
  for (uint32_t iter = 0; iter < st->max_bounces; ++iter) {
    R t = 0, bu = 0, bv = 0;
    uint32_t prim, inst;
    int hit = closest(s, o, d, tmin, tmax, &t, &bu, &bv, &prim, &inst);
    if (hit) t = jitter_t(t, hash32(pixel ^ hash32(frame_sample)) ^ iter);
    V inc = d;
    if (rmax(T.x, rmax(T.y, T.z)) < (R)0.0001) break;
    if (!hit) {
      V env = env_rad(s, inc);
      R mis = 1;
      if ((flags & 2u) && st->env_samples > 0) {
        R lp = prev_pdf * (R)st->env_samples;
        if (flags & 4u) lp = prev_pdf * ((R)1 - unified_prob(s, st));
        mis = power_h(prev_pdf, lp);
      }
      Lr = vadd(Lr, vs(vmul(T, env), mis));
      break;
    }
    V hp = vadd(o, sv(t, d));
    V cw, sn_sum;
    uint32_t mat_id;
    hit_frame(s, prim, inst, bu, bv, &cw, &sn_sum, &mat_id);
    V gn = norm(cw);
    V sn = norm(sn_sum);
    if (dot(sn, gn) < (R)0) sn = vneg(sn);
    int entering = dot(inc, gn) < (R)0;
    V n = entering ? sn : vneg(sn);
    Mat m = load_mat(s, mat_id);
    V em = emission_of(s, mat_id);
    /* The ray cone at the hit (shadeRayCone :1217-1222). */
    R cone_hit = cone_w + t * cone_s;
    if (s->uvs && mat_textured(s, mat_id)) {
      tex_surface(s, mat_id, prim, bu, bv, gn, cone_hit, &m, &em, &sn);
      n = entering ? sn : vneg(sn);
    }
    uint32_t depth = (flags >> 8) > 1u ? (flags >> 8) : 1u;
    uint32_t rs = rng;
    if (entering || m.two_sided > (R)0.5) {
      R emis = 1;
      if ((flags & 2u) && dot(em, em) > (R)0) {
        R area = (R)0.5 * len(cw);
        R raw = dot(gn, vneg(inc));
        R lc = rmax(raw, (R)0);
        if (m.two_sided > (R)0.5) lc = absr(raw);
        R lum = dot(em, v3((R)0.2126, (R)0.7152, (R)0.0722));
        R sel = (area * lum) / rmax((R)s->weight, (R)1e-6);
        R lp = ((sel * t) * t) / rmax(lc * area, (R)1e-6);
        R sp = lp * (R)(st->sec_area > 1u ? st->sec_area : 1u);
        if (flags & 4u) sp = lp * unified_prob(s, st);
        emis = OL_MUT(11) ? (R)1 : power_h(prev_pdf, sp);
      }
      Lr = vadd(Lr, vs(vmul(T, em), emis));
    }
    uint32_t nb = bounce + 1u;
    if (nb >= st->max_bounces) break;
    R trans = m.trans, bsdf_pdf = 0;
    int used_restir = 0;
    V nd;
    if (trans > (R)0.001) {
      R cur = medium_get(stack, depth - 1u), target = 1;
      if (entering) target = rmax(m.ior, (R)1.0001);
      else if (depth > 1u && !OL_MUT(7)) target = medium_get(stack, depth - 2u);
      R eta = OL_MUT(1) ? target / cur : cur / target, ndi = dot(n, inc);
      R k = (R)1 - eta * eta * ((R)1 - ndi * ndi);
      V rd = k < (R)0 ? v3(0, 0, 0) : vsub(sv(eta, inc), sv(eta * ndi + rsqrt_(k), n));
      if (dot(rd, rd) < (R)0.01) {
        rd = OL_MUT(6) ? inc : reflect_(inc, n);
      } else if (entering && depth < 16u) {
        stack[depth] = target;
        depth++;
      } else if (!entering && depth > 1u) {
        depth--;
      }
      R optical = (entering != OL_MUT(5)) ? (R)0 : t;
      if (m.thin > (R)0.5) optical = m.thick;
      R ex = optical / rmax(m.adist, (R)1e-6);
      V ab;
      if (sizeof(R) == 4) {
        ab = v3((R)powf((float)rmax(m.atten.x, (R)1e-6), (float)ex), (R)powf((float)rmax(m.atten.y, (R)1e-6), (float)ex),
                (R)powf((float)rmax(m.atten.z, (R)1e-6), (float)ex));
      } else {
        ab = v3((R)pow((double)rmax(m.atten.x, (R)1e-6), (double)ex), (R)pow((double)rmax(m.atten.y, (R)1e-6), (double)ex),
                (R)pow((double)rmax(m.atten.z, (R)1e-6), (double)ex));
      }
      V tint = mixv(v3(1, 1, 1), m.base, m.thin);
      T = vs(vmul(vmul(T, ab), tint), trans);
      nd = rd;
    } else {
      R np = clampr((R)st->nee_prob, (R)1e-6, (R)1);
      int do_nee = 1;
      if (!(np >= (R)0.999999)) {
        uint32_t scr = hash32((pixel ^ hash32(nb + 1u)) ^ hash32((frame_sample & 255u) + 1u));
        uint32_t seq = brev(frame_sample >> 8) ^ scr;
        do_nee = ((R)seq + (R)0.5) * (R)(1.0 / 4294967296.0) < np;
      }
      if (do_nee) {
        V direct = v3(0, 0, 0), c;
        for (uint32_t li = 0; li < 64u && li < s->P; ++li) {
          if (point_sample(s, li, hp, n, &m, inc, &c)) direct = vadd(direct, c);
        }
        if (rc && bounce == 0u) {
          /* Primary hit: the ReSTIR area-light sample, then environment
           * samples (wavefront_primary_impl.glsl:1527-1785). */
          V rdir;
          R rtmax;
          used_restir = 1;
          if (restir_primary(s, rc, &rs, hp, n, &m, inc, t, pixel, mat_id, &c, &rdir, &rtmax) &&
              !occluded(s, vadd(hp, vs(n, (R)0.002)), rdir, (R)0.001, rtmax))
            direct = vadd(direct, c);
          uint32_t ec = st->env_samples < 4u ? st->env_samples : 4u;
          V ed = v3(0, 0, 0);
          for (uint32_t k = 0; k < ec; ++k)
            if (env_sample(s, &rs, hp, n, &m, inc, ec, 1, &c)) ed = vadd(ed, c);
          if (ec > 0) direct = vadd(direct, vdiv(ed, (R)ec));
        } else if (st->unified) {
          int area_on = s->L > 0, env_on = st->env_samples > 0;
          if (area_on || env_on) {
            R aw = 0, ew = 0;
            if (area_on) aw = rsqrt_(rmax((R)s->weight, (R)1e-6)) * (R)(st->sec_area > 1u ? st->sec_area : 1u);
            if (env_on) ew = (R)(st->env_samples < 4u ? st->env_samples : 4u);
            R ap = aw / rmax(aw + ew, (R)1e-6);
            rs = pcg(rs);
            if (pcg_unit(rs) < ap) {
              if (area_sample(s, &rs, hp, n, &m, inc, 0, 1, ap, &c)) direct = vadd(direct, c);
            } else {
              if (env_sample(s, &rs, hp, n, &m, inc, 1, (R)1 - ap, &c)) direct = vadd(direct, c);
            }
          }
        } else {
          uint32_t ac = st->sec_area < 1u ? 1u : (st->sec_area > 16u ? 16u : st->sec_area);
          V ad = v3(0, 0, 0), ed = v3(0, 0, 0);
          for (uint32_t k = 0; k < ac; ++k)
            if (area_sample(s, &rs, hp, n, &m, inc, k, ac, 1, &c)) ad = vadd(ad, c);
          direct = vadd(direct, vdiv(ad, (R)ac));
          uint32_t ec = st->env_samples < 4u ? st->env_samples : 4u;
          for (uint32_t k = 0; k < ec; ++k)
            if (env_sample(s, &rs, hp, n, &m, inc, ec, 1, &c)) ed = vadd(ed, c);
          if (ec > 0) direct = vadd(direct, vdiv(ed, (R)ec));
        }
        Lr = vadd(Lr, vdiv(vmul(T, direct), rmax(np, (R)1e-6)));
      }
      V view = vneg(inc);
      R prob = spec_prob(&m);
      rs = pcg(rs);
      int spec = pcg_unit(rs) < prob;
      rs = pcg(rs);
      R u1 = pcg_unit(rs);
      rs = pcg(rs);
      R u2 = pcg_unit(rs);
      V ov = cosine_dir(n, u1, u2);
      if (spec) ov = reflect_(inc, ggx_half(n, m.rough, u1, u2));
      if (dot(n, ov) <= (R)0) {
        rs = pcg(rs);
        u1 = pcg_unit(rs);
        rs = pcg(rs);
        u2 = pcg_unit(rs);
        ov = cosine_dir(n, u1, u2);
      }
      ov = norm(ov);
      R pdf = rmax(pdf_pbr(&m, n, view, ov), (R)1e-6);
      V w = vdiv(vs(eval_pbr(&m, n, view, ov), rmax(dot(n, ov), (R)0)), pdf);
      w = vs(w, mixr(m.occl, (R)1, m.metal));
      T = vmul(T, w);
      nd = ov;
      bsdf_pdf = pdf;
    }
    if (!(st->rr_start == 0 || nb < st->rr_start || trans > (R)0.001)) {
      R surv = clampr(rmax(T.x, rmax(T.y, T.z)), (R)st->rr_min, (R)0.95);
      rs = pcg(rs);
      if (pcg_unit(rs) >= surv) break;
      if (!OL_MUT(4)) T = vdiv(T, surv);
    }
    d = norm(nd);
    bounce = nb;
    /* The continuation's cone: the width at this hit, the spread widened by
     * the roughness after an opaque scatter (:1551, :1373). */
    cone_w = cone_hit;
    if (!(trans > (R)0.001)) cone_s = cone_s + m.rough * (R)0.25;
    flags = 1u | (depth << 8);
    if (trans <= (R)0.001) {
      flags |= 2u;
      if (st->unified && !used_restir) flags |= 4u;
      prev_pdf = bsdf_pdf;
    }
    rng = rs;
    o = vadd(hp, vs(d, (R)0.002));
    tmin = (R)0.001;
    tmax = (R)1e30;
    counts[1]++;
  }
  out[0] = Lr.x;
  out[1] = Lr.y;
  out[2] = Lr.z;
}

/* Camera ray of wavefront_generate.comp for one (pixel, sample). */
static void camera_ray(const float cam[12], uint32_t proj, uint32_t W, uint32_t H, uint32_t px, uint32_t py,
                       uint32_t frame, uint32_t sample, V* o, V* d) {
  uint32_t pixel = py * W + px;
  uint32_t r = hash32((pixel ^ hash32(frame)) ^ hash32(sample + 1u));
  r = hash32(r);
  R jx = unit(r);
  r = hash32(r);
  R jy = unit(r);
  R ndx = (((R)px + jx) / (R)W) * (R)2 - (R)1;
  R ndy = (((R)py + jy) / (R)H) * (R)2 - (R)1;
  R aspect = (R)W / (R)H;
  V org = v3(cam[0], cam[1], cam[2]), f = v3(cam[3], cam[4], cam[5]), rt = v3(cam[6], cam[7], cam[8]),
    up = v3(cam[9], cam[10], cam[11]);
  if (proj == 1) {
    *o = vsub(vadd(org, sv(ndx * aspect, rt)), sv(ndy, up));
    *d = norm(f);
  } else if (proj == 2) {
    R yaw = ndx * len(rt), pitch = (-ndy) * len(up), sy, cy, sp, cp;
    sincos_r(yaw, &sy, &cy);
    sincos_r(pitch, &sp, &cp);
    *o = org;
    *d = norm(vadd(vadd(vs(vs(norm(f), cp), cy), vs(vs(norm(rt), cp), sy)), vs(norm(up), sp)));
  } else {
    *o = org;
    *d = norm(vsub(vadd(f, sv(ndx * aspect, rt)), sv(ndy, up)));
  }
}

/* Camera rays of a pixel tile, row-major: o/d [n][4] (xyz, tmin / tmax). */
void API(olgi_camera_rays)(const float cam[12], uint32_t proj, uint32_t W, uint32_t H, uint32_t frame,
                           uint32_t sample, uint32_t x0, uint32_t y0, uint32_t tw, uint32_t th, float* o, float* d) {
  for (uint32_t ly = 0; ly < th; ++ly)
    for (uint32_t lx = 0; lx < tw; ++lx) {
      V oo, dd;
      camera_ray(cam, proj, W, H, x0 + lx, y0 + ly, frame, sample, &oo, &dd);
      const uint64_t i = (uint64_t)ly * tw + lx;
      o[4 * i] = (float)oo.x;
      o[4 * i + 1] = (float)oo.y;
      o[4 * i + 2] = (float)oo.z;
      o[4 * i + 3] = 0.001f;
      d[4 * i] = (float)dd.x;
      d[4 * i + 1] = (float)dd.y;
      d[4 * i + 2] = (float)dd.z;
      d[4 * i + 3] = 1e30f;
    }
}

/* Renders rows [row0, row1) into out (W*H*4 floats, full image layout).
 * settings: max_bounces, sec_area, env_samples, rr_start, rr_min, nee_prob,
 * unified.  frame is the float32-rounded frame index.  counts[3] receives
 * primary/extension ray counts (shadow rays are not counted here). */
int API(olgi_render)(void* h, const float cam[12], uint32_t proj, uint32_t W, uint32_t H, uint32_t spp, uint32_t frame,
                     const uint32_t iset[4], const float fset[2], uint32_t unified, uint32_t row0, uint32_t row1,
                     float* out, uint64_t* counts_out) {
  const Scene* s = (const Scene*)h;
  Settings st = {iset[0], iset[1], iset[2], iset[3], fset[0], fset[1], unified};
  uint64_t c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c0, c1, c2)
  for (int64_t py = (int64_t)row0; py < (int64_t)row1; ++py) {
    uint64_t counts[3] = {0, 0, 0};
    g_shadow_rays = 0;
    for (uint32_t px = 0; px < W; ++px) {
      double acc[3] = {0, 0, 0};
      for (uint32_t smp = 0; smp < spp; ++smp) {
        V o, d;
        camera_ray(cam, proj, W, H, px, (uint32_t)py, frame, smp, &o, &d);
        R rad[3];
        trace_path(s, &st, NULL, (uint32_t)py * W + px, frame, smp, o, d, rad, counts);
        /* float32 radiance per sample, summed in float64 (api.py:1404-1407) */
        for (int k = 0; k < 3; ++k) acc[k] += (double)(float)rad[k];
      }
      float* p = out + 4ull * ((uint64_t)py * W + px);
      for (int k = 0; k < 3; ++k) p[k] = (float)(acc[k] / (double)spp);
      p[3] = (float)((double)spp / (double)spp);
    }
    c0 += counts[0];
    c1 += counts[1];
    c2 += g_shadow_rays;
  }
  if (counts_out) {
    counts_out[0] = c0;
    counts_out[1] = c1;
    counts_out[2] = c2;
  }
  return 0;
}

/* Named products (targets/vulkan/api.py:1252-1451, core.py:4500-4536):
 * sample 0's primary ray -> closest hit: depth = t (or +inf), geometric
 * normal normalize(cross(b - a, c - a)) unflipped (wavefront_intersect.comp
 * :317-356), triangle index (or 0xFFFFFFFF); and the unbiased per-pixel
 * variance of the float32 sample luminance 0.2126 r + 0.7152 g + 0.0722 b
 * (float32 arithmetic, float64 sums).  Any output may be NULL. */
int API(olgi_render_products)(void* h, const float cam[12], uint32_t proj, uint32_t W, uint32_t H, uint32_t spp,
                              uint32_t frame, const uint32_t iset[4], const float fset[2], uint32_t unified,
                              float* depth, float* normal, uint32_t* prim, float* variance) {
  const Scene* s = (const Scene*)h;
  Settings st = {iset[0], iset[1], iset[2], iset[3], fset[0], fset[1], unified};
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t py = 0; py < (int64_t)H; ++py) {
    uint64_t counts[3] = {0, 0, 0};
    for (uint32_t px = 0; px < W; ++px) {
      const uint64_t pix = (uint64_t)py * W + px;
      if (depth || normal || prim) {
        V o, d;
        camera_ray(cam, proj, W, H, px, (uint32_t)py, frame, 0, &o, &d);
        R t, u, v;
        uint32_t tri, inst;
        const int hit = closest(s, o, d, (R)0.001, (R)1e30, &t, &u, &v, &tri, &inst);
        V n = v3(0, 0, 0);
        if (hit) {
          V cw, sn_sum;
          uint32_t mat_id;
          hit_frame(s, tri, inst, u, v, &cw, &sn_sum, &mat_id);
          n = norm(cw);
        }
        if (depth) depth[pix] = hit ? (float)t : INFINITY;
        if (normal) {
          normal[3 * pix] = (float)n.x;
          normal[3 * pix + 1] = (float)n.y;
          normal[3 * pix + 2] = (float)n.z;
        }
        if (prim) prim[pix] = hit ? tri : 0xFFFFFFFFu;
      }
      if (variance) {
        double sum = 0.0, sq = 0.0;
        for (uint32_t smp = 0; smp < spp; ++smp) {
          V o, d;
          camera_ray(cam, proj, W, H, px, (uint32_t)py, frame, smp, &o, &d);
          R rad[3];
          trace_path(s, &st, NULL, (uint32_t)pix, frame, smp, o, d, rad, counts);
          const float y = ((float)rad[0] * 0.2126f + (float)rad[1] * 0.7152f) + (float)rad[2] * 0.0722f;
          sum += (double)y;
          sq += (double)y * (double)y;
        }
        double var = 0.0;
        if (spp > 1) {
          var = (sq - sum * sum / (double)spp) / (double)(spp - 1);
          if (var < 0.0) var = 0.0;
        }
        variance[pix] = (float)var;
      }
    }
  }
  return 0;
}

/* ReSTIR DI render: samples run as successive passes over the whole image,
 * pass s reusing pass s - 1's reservoirs (double-buffered).  rset:
 * candidates, spatial neighbours (0 = none), radius, pairwise (0/1),
 * history limit.  Other arguments as olgi_render (whole image). */
int API(olgi_render_restir)(void* h, const float cam[12], uint32_t proj, uint32_t W, uint32_t H, uint32_t spp,
                            uint32_t frame, const uint32_t iset[4], const float fset[2], uint32_t unified,
                            const uint32_t rset[5], float* out, uint64_t* counts_out) {
  const Scene* s = (const Scene*)h;
  Settings st = {iset[0], iset[1], iset[2], iset[3], fset[0], fset[1], unified};
  const uint64_t np = (uint64_t)W * H;
  uint32_t* res = (uint32_t*)calloc(8 * np, sizeof(uint32_t));
  float* surf = (float*)calloc(16 * np, sizeof(float));
  double* acc = (double*)calloc(3 * np, sizeof(double));
  if (!res || !surf || !acc) {
    free(res);
    free(surf);
    free(acc);
    return -1;
  }
  uint64_t c0 = 0, c1 = 0, c2 = 0;
  for (uint32_t pass = 0; pass < spp; ++pass) {
    Restir rc;
    rc.candidates = rset[0];
    rc.spatial = rset[1];
    rc.radius = rset[2];
    rc.pairwise = rset[3];
    rc.limit = (R)rset[4];
    rc.history = pass > 0u;
    rc.cam = cam;
    rc.proj = proj;
    rc.W = W;
    rc.H = H;
    rc.frame = frame;
    rc.pass = pass;
    rc.res_out = res + 4 * np * (pass & 1u);
    rc.res_in = res + 4 * np * ((pass + 1u) & 1u);
    rc.surf_out = surf + 8 * np * (pass & 1u);
    rc.surf_in = surf + 8 * np * ((pass + 1u) & 1u);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c0, c1, c2)
    for (int64_t py = 0; py < (int64_t)H; ++py) {
      uint64_t counts[3] = {0, 0, 0};
      g_shadow_rays = 0;
      for (uint32_t px = 0; px < W; ++px) {
        V o, d;
        camera_ray(cam, proj, W, H, px, (uint32_t)py, frame, pass, &o, &d);
        R rad[3];
        const uint64_t pix = (uint64_t)py * W + px;
        trace_path(s, &st, &rc, (uint32_t)pix, frame, pass, o, d, rad, counts);
        for (int k = 0; k < 3; ++k) acc[3 * pix + k] += (double)(float)rad[k];
      }
      c0 += counts[0];
      c1 += counts[1];
      c2 += g_shadow_rays;
    }
  }
  for (uint64_t i = 0; i < np; ++i) {
    for (int k = 0; k < 3; ++k) out[4 * i + k] = (float)(acc[3 * i + k] / (double)spp);
    out[4 * i + 3] = 1.0f;
  }
  free(res);
  free(surf);
  free(acc);
  if (counts_out) {
    counts_out[0] = c0;
    counts_out[1] = c1;
    counts_out[2] = c2;
  }
  return 0;
}

/* Closest hits for caller rays: o/d [n][4] (xyz + tmin/tmax); hits [n][4]. */
int API(olgi_trace)(void* h, const float* o, const float* d, uint32_t n, float* hits) {
  const Scene* s = (const Scene*)h;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    R t, u = 0, v = 0;
    uint32_t prim, inst;
    V oo = v3(o[4 * i], o[4 * i + 1], o[4 * i + 2]), dd = v3(d[4 * i], d[4 * i + 1], d[4 * i + 2]);
    int hit = closest(s, oo, dd, (R)o[4 * i + 3], (R)d[4 * i + 3], &t, &u, &v, &prim, &inst);
    hits[4 * i] = (float)t;
    hits[4 * i + 1] = (float)u;
    hits[4 * i + 2] = (float)v;
    uint32_t id = hit ? prim : 0xFFFFFFFFu;
    memcpy(&hits[4 * i + 3], &id, 4);
  }
  return 0;
}

/* Known-answer helpers for tests. */
uint32_t API(olgi_wave_hash)(uint32_t v) { return hash32(v); }
uint32_t API(olgi_pcg_step)(uint32_t s) { return pcg(s); }
float API(olgi_pcg_value)(uint32_t s) { return (float)pcg_unit(s); }
void API(olgi_sincos)(float x, float* s, float* c) {
  R a, b;
  sincos_r((R)x, &a, &b);
  *s = (float)a;
  *c = (float)b;
}

/* BSDF hooks for property tests: material record (48 floats), unit normal,
 * view and outgoing directions (3 floats each). */
void API(olgi_eval_pbr)(const float* rec, const float* n, const float* view, const float* out, float* f_out,
                        float* pdf_out) {
  Scene tmp;
  memset(&tmp, 0, sizeof(tmp));
  tmp.mat = (float*)rec;
  Mat m = load_mat(&tmp, 0);
  V N = v3(n[0], n[1], n[2]), Vw = v3(view[0], view[1], view[2]), O = v3(out[0], out[1], out[2]);
  V f = eval_pbr(&m, N, Vw, O);
  f_out[0] = (float)f.x;
  f_out[1] = (float)f.y;
  f_out[2] = (float)f.z;
  *pdf_out = (float)pdf_pbr(&m, N, Vw, O);
}

/* One BSDF sample (shadeSamplePbr): outgoing direction, weight, pdf. */
uint32_t API(olgi_sample_pbr)(const float* rec, const float* n, const float* incoming, uint32_t rs, float* out_dir,
                              float* weight, float* pdf_out) {
  Scene tmp;
  memset(&tmp, 0, sizeof(tmp));
  tmp.mat = (float*)rec;
  Mat m = load_mat(&tmp, 0);
  V N = v3(n[0], n[1], n[2]), I = v3(incoming[0], incoming[1], incoming[2]);
  V view = vneg(I);
  R prob = spec_prob(&m);
  rs = pcg(rs);
  int spec = pcg_unit(rs) < prob;
  rs = pcg(rs);
  R u1 = pcg_unit(rs);
  rs = pcg(rs);
  R u2 = pcg_unit(rs);
  V ov = cosine_dir(N, u1, u2);
  if (spec) ov = reflect_(I, ggx_half(N, m.rough, u1, u2));
  if (dot(N, ov) <= (R)0) {
    rs = pcg(rs);
    u1 = pcg_unit(rs);
    rs = pcg(rs);
    u2 = pcg_unit(rs);
    ov = cosine_dir(N, u1, u2);
  }
  ov = norm(ov);
  R pdf = rmax(pdf_pbr(&m, N, view, ov), (R)1e-6);
  V w = vdiv(vs(eval_pbr(&m, N, view, ov), rmax(dot(N, ov), (R)0)), pdf);
  w = vs(w, mixr(m.occl, (R)1, m.metal));
  out_dir[0] = (float)ov.x;
  out_dir[1] = (float)ov.y;
  out_dir[2] = (float)ov.z;
  weight[0] = (float)w.x;
  weight[1] = (float)w.y;
  weight[2] = (float)w.z;
  *pdf_out = (float)pdf;
  return rs;
}
