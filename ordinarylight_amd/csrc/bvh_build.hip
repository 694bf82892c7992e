// GPU LBVH build: 30-bit Morton codes -> stable LDS radix sort -> Karras
// binary radix tree -> bottom-up boxes -> breadth-first collapse into the
// 64-byte 4-wide node format of common.h.
//
// Every step is integer work or exact float min/max/compare, so the result is
// bit-identical to the sequential CPU build of the same algorithm in
// oracle/lbvh_oracle.c (checked by tests/test_bvh_gpu.py).  Node numbering is
// breadth-first with per-level exclusive scans (no atomics decide an index).
//
// Replaces the driver acceleration-structure builds of the reference
// (targets/vulkan/scene.py:1132-1194, :1480-1561).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

constexpr int kBlock = 256;

// ---- float <-> ordered int for atomic min/max of finite floats ----------------
__device__ __forceinline__ int float_to_ordered(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
OL_HD float ordered_to_float(int i) {
  int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
  union { int i; float f; } v;
  v.i = b;
  return v.f;
}

OL_HD uint32_t expand_bits10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

OL_HD uint32_t quantize10(float c, float lo, float ext) {
  float n = ext > 0.0f ? (c - lo) / ext : 0.5f;
  float q = fminf(fmaxf(n * 1024.0f, 0.0f), 1023.0f);
  return (uint32_t)q;
}

// ---- 1. triangle bounds + centroid bounds -----------------------------------------
__global__ void __launch_bounds__(kBlock) k_tri_bounds(const float4* __restrict__ pos, const uint4* __restrict__ vidx,
                                                       uint32_t n, float4* __restrict__ tmin, float4* __restrict__ tmax,
                                                       int* __restrict__ cbounds /* 6 ordered ints */) {
  __shared__ int red[6][kBlock / 64];
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    uint4 t = vidx[i];
    float4 a = pos[t.x], b = pos[t.y], c = pos[t.z];
    float4 mn = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z), 0.0f);
    float4 mx = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z), 0.0f);
    tmin[i] = mn;
    tmax[i] = mx;
    cmin[0] = cmax[0] = (mn.x + mx.x) * 0.5f;
    cmin[1] = cmax[1] = (mn.y + mx.y) * 0.5f;
    cmin[2] = cmax[2] = (mn.z + mx.z) * 0.5f;
  }
  int v[6] = {float_to_ordered(cmin[0]), float_to_ordered(cmin[1]), float_to_ordered(cmin[2]),
              float_to_ordered(cmax[0]), float_to_ordered(cmax[1]), float_to_ordered(cmax[2])};
  for (int off = 32; off >= 1; off >>= 1) {
    for (int k = 0; k < 3; ++k) v[k] = min(v[k], __shfl_xor(v[k], off, 64));
    for (int k = 3; k < 6; ++k) v[k] = max(v[k], __shfl_xor(v[k], off, 64));
  }
  uint32_t wave = threadIdx.x / 64;
  if (lane_id() == 0)
    for (int k = 0; k < 6; ++k) red[k][wave] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; ++k) {
      int r = red[k][0];
      for (int w = 1; w < kBlock / 64; ++w) r = k < 3 ? min(r, red[k][w]) : max(r, red[k][w]);
      if (k < 3) atomicMin(&cbounds[k], r); else atomicMax(&cbounds[k], r);
    }
  }
}

__global__ void k_init_bounds(int* cb) {
  if (threadIdx.x < 3) cb[threadIdx.x] = 0x7FFFFFFF;
  else if (threadIdx.x < 6) cb[threadIdx.x] = (int)0x80000000;
}

// ---- 2. Morton codes ---------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_morton(const float4* __restrict__ tmin, const float4* __restrict__ tmax,
                                                   uint32_t n, const int* __restrict__ cb,
                                                   uint32_t* __restrict__ codes, uint32_t* __restrict__ values) {
  uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float lo[3] = {ordered_to_float(cb[0]), ordered_to_float(cb[1]), ordered_to_float(cb[2])};
  float hi[3] = {ordered_to_float(cb[3]), ordered_to_float(cb[4]), ordered_to_float(cb[5])};
  float4 mn = tmin[i], mx = tmax[i];
  float c[3] = {(mn.x + mx.x) * 0.5f, (mn.y + mx.y) * 0.5f, (mn.z + mx.z) * 0.5f};
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) q[k] = quantize10(c[k], lo[k], hi[k] - lo[k]);
  codes[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  values[i] = i;
}

// ---- 3. stable LSD radix sort, 8-bit digits, 4096-key tiles --------------------------
constexpr int kSortItems = 16;
constexpr int kTile = kBlock * kSortItems;

__global__ void __launch_bounds__(kBlock) k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                      uint32_t nblocks, uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  uint32_t base = blockIdx.x * kTile;
  for (int r = 0; r < kSortItems; ++r) {
    uint32_t i = base + r * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter(const uint32_t* __restrict__ keys_in,
                                                         const uint32_t* __restrict__ vals_in, uint32_t n, int shift,
                                                         uint32_t nblocks, const uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t digit_base[256];
  __shared__ uint32_t wave_cnt[kBlock / 64][256];
  const uint32_t tid = threadIdx.x, wave = tid / 64;
  digit_base[tid] = offsets[tid * nblocks + blockIdx.x];
  for (int w = 0; w < kBlock / 64; ++w) wave_cnt[w][tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTile;
  for (int r = 0; r < kSortItems; ++r) {
    uint32_t i = base + r * kBlock + tid;
    bool valid = i < n;
    uint32_t key = valid ? keys_in[i] : 0u;
    uint32_t val = valid ? vals_in[i] : 0u;
    uint32_t d = (key >> shift) & 255u;
    // Lanes of this wave holding the same digit (valid lanes only).
    uint64_t same = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      uint64_t bits = __ballot((d >> b) & 1u);
      same &= ((d >> b) & 1u) ? bits : ~bits;
    }
    uint32_t rank = lanes_below(same);
    if (valid && rank == 0) wave_cnt[wave][d] = (uint32_t)__popcll(same);
    __syncthreads();
    uint32_t pos = 0;
    if (valid) {
      pos = digit_base[d] + rank;
      for (uint32_t w = 0; w < wave; ++w) pos += wave_cnt[w][d];
    }
    __syncthreads();
    // Advance the per-digit running base by this round's counts; clear counters.
    uint32_t add = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      add += wave_cnt[w][tid];
      wave_cnt[w][tid] = 0;
    }
    digit_base[tid] += add;
    __syncthreads();
    if (valid) {
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
  }
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                      hipStream_t s, int bits) {
  if (n <= 1) return;
  uint32_t nblocks = (n + kTile - 1) / kTile;
  DevBuf counts;
  counts.reserve(sizeof(uint32_t) * 256u * nblocks);
  uint32_t *src_k = keys, *src_v = vals, *dst_k = keys_tmp, *dst_v = vals_tmp;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(kBlock), 0, s, src_k, n, shift, nblocks, counts.as<uint32_t>());
    scan_exclusive_u32(counts.as<uint32_t>(), 256u * nblocks, nullptr, s);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(kBlock), 0, s, src_k, src_v, n, shift, nblocks,
                       counts.as<uint32_t>(), dst_k, dst_v);
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  if (src_k != keys) {
    OL_HIP_CHECK(hipMemcpyAsync(keys, src_k, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    OL_HIP_CHECK(hipMemcpyAsync(vals, src_v, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
  }
  OL_HIP_CHECK(hipStreamSynchronize(s));
}

// ---- 4. Karras binary radix tree ---------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes, int n, int i, int j) {
  if (j < 0 || j > n - 1) return -1;
  uint32_t a = codes[i], b = codes[j];
  if (a != b) return __clz((int)(a ^ b));
  return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ void __launch_bounds__(kBlock) k_karras(const uint32_t* __restrict__ codes, int n, uint32_t* __restrict__ left,
                                                   uint32_t* __restrict__ right, uint32_t* __restrict__ rfirst,
                                                   uint32_t* __restrict__ rlast, uint32_t* __restrict__ parent_int,
                                                   uint32_t* __restrict__ parent_leaf) {
  int i = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (i >= n - 1) return;
  int d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta(codes, n, i, i - d);
  int lmax = 2;
  while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta(codes, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + min(d, 0);
  int lo = min(i, j), hi = max(i, j);
  uint32_t lc = (lo == gamma) ? ((uint32_t)gamma | OL_LEAF_BIT) : (uint32_t)gamma;
  uint32_t rc = (hi == gamma + 1) ? ((uint32_t)(gamma + 1) | OL_LEAF_BIT) : (uint32_t)(gamma + 1);
  left[i] = lc;
  right[i] = rc;
  rfirst[i] = (uint32_t)lo;
  rlast[i] = (uint32_t)hi;
  if (lc & OL_LEAF_BIT) parent_leaf[lc & ~OL_LEAF_BIT] = (uint32_t)i; else parent_int[lc] = (uint32_t)i;
  if (rc & OL_LEAF_BIT) parent_leaf[rc & ~OL_LEAF_BIT] = (uint32_t)i; else parent_int[rc] = (uint32_t)i;
  if (i == 0) parent_int[0] = OL_NONE;
}

// ---- 5. bottom-up boxes -------------------------------------------------------------------
__device__ __forceinline__ void ref_box(uint32_t ref, const float4* lmin, const float4* lmax, const float4* bmin,
                                        const float4* bmax, float4& mn, float4& mx) {
  if (ref & OL_LEAF_BIT) {
    mn = lmin[ref & ~OL_LEAF_BIT];
    mx = lmax[ref & ~OL_LEAF_BIT];
  } else {
    mn = bmin[ref];
    mx = bmax[ref];
  }
}

__global__ void __launch_bounds__(kBlock) k_leaf_boxes(const uint32_t* __restrict__ sorted, uint32_t n,
                                                       const float4* __restrict__ tmin, const float4* __restrict__ tmax,
                                                       float4* __restrict__ lmin, float4* __restrict__ lmax) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  uint32_t o = sorted[k];
  lmin[k] = tmin[o];
  lmax[k] = tmax[o];
}

__global__ void __launch_bounds__(kBlock) k_bottom_up(uint32_t n, const uint32_t* __restrict__ left,
                                                      const uint32_t* __restrict__ right,
                                                      const uint32_t* __restrict__ parent_int,
                                                      const uint32_t* __restrict__ parent_leaf,
                                                      const float4* __restrict__ lmin, const float4* __restrict__ lmax,
                                                      float4* bmin, float4* bmax, uint32_t* flags) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  uint32_t node = parent_leaf[k];
  while (node != OL_NONE) {
    // Release this subtree's box, acquire the sibling's (device scope).
    uint32_t old = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;  // the sibling subtree finishes this node
    float4 amn, amx, bmn, bmx;
    ref_box(left[node], lmin, lmax, bmin, bmax, amn, amx);
    ref_box(right[node], lmin, lmax, bmin, bmax, bmn, bmx);
    float4 mn = make_float4(fminf(amn.x, bmn.x), fminf(amn.y, bmn.y), fminf(amn.z, bmn.z), 0.0f);
    float4 mx = make_float4(fmaxf(amx.x, bmx.x), fmaxf(amx.y, bmx.y), fmaxf(amx.z, bmx.z), 0.0f);
    bmin[node] = mn;
    bmax[node] = mx;
    node = parent_int[node];
  }
}

// ---- 6. breadth-first collapse to 4-wide nodes --------------------------------------------
struct TreeView {
  const uint32_t *left, *right, *rfirst, *rlast;
  const float4 *lmin, *lmax, *bmin, *bmax;
  uint32_t leaf_size;
  uint32_t root;  // binary root reference
};

// Float32 ray/triangle tests accept points a few ulps of the coordinate
// magnitude outside the exact triangle (vertex- and edge-grazing rays).
// Every wide-node box is grown by 2^-17 of the largest root-box coordinate
// on each side before quantisation, so no such hit is culled.
__device__ __forceinline__ float box_pad(const TreeView& t) {
  float4 mn, mx;
  ref_box(t.root, t.lmin, t.lmax, t.bmin, t.bmax, mn, mx);
  const float m = fmaxf(fmaxf(fmaxf(fabsf(mn.x), fabsf(mn.y)), fmaxf(fabsf(mn.z), fabsf(mx.x))),
                        fmaxf(fabsf(mx.y), fabsf(mx.z)));
  return m * 0x1p-17f;
}

__device__ __forceinline__ uint32_t ref_count(const TreeView& t, uint32_t ref) {
  return (ref & OL_LEAF_BIT) ? 1u : t.rlast[ref] - t.rfirst[ref] + 1u;
}
__device__ __forceinline__ uint32_t ref_first(const TreeView& t, uint32_t ref) {
  return (ref & OL_LEAF_BIT) ? (ref & ~OL_LEAF_BIT) : t.rfirst[ref];
}
__device__ __forceinline__ float ref_area(const TreeView& t, uint32_t ref) {
  float4 mn, mx;
  ref_box(ref, t.lmin, t.lmax, t.bmin, t.bmax, mn, mx);
  float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
  return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}
__device__ __forceinline__ bool ref_is_wide_leaf(const TreeView& t, uint32_t ref) {
  return (ref & OL_LEAF_BIT) || ref_count(t, ref) <= t.leaf_size;
}

// Greedy opening: repeatedly replace the largest-area child that is not a
// wide leaf by its two children, until four children or nothing to open.
__device__ int plan_children(const TreeView& t, uint32_t ref, uint32_t kids[4]) {
  if (ref_is_wide_leaf(t, ref)) {
    kids[0] = ref;
    return 1;
  }
  int nk = 2;
  kids[0] = t.left[ref];
  kids[1] = t.right[ref];
  while (nk < 4) {
    int best = -1;
    float best_area = -1.0f;
    for (int k = 0; k < nk; ++k) {
      if (ref_is_wide_leaf(t, kids[k])) continue;
      float a = ref_area(t, kids[k]);
      if (a > best_area) { best_area = a; best = k; }
    }
    if (best < 0) break;
    uint32_t r = kids[best];
    for (int k = nk; k > best + 1; --k) kids[k] = kids[k - 1];
    kids[best] = t.left[r];
    kids[best + 1] = t.right[r];
    ++nk;
  }
  return nk;
}

__global__ void __launch_bounds__(kBlock) k_collapse_plan(TreeView t, const uint32_t* __restrict__ front, uint32_t count,
                                                          uint32_t* __restrict__ inner_count) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  uint32_t kids[4];
  int nk = plan_children(t, front[k], kids);
  uint32_t c = 0;
  for (int i = 0; i < nk; ++i) c += ref_is_wide_leaf(t, kids[i]) ? 0u : 1u;
  inner_count[k] = c;
}

// Quantise one axis of up to four child boxes against the node box [lo, hi].
OL_HD uint32_t quant_exponent(float lo, float hi) {
  float ext = hi - lo;
  int e = -100;
  if (ext > 0.0f) {
    union { float f; uint32_t u; } s;
    s.f = ext / 255.0f;
    int E = (int)((s.u >> 23) & 0xFFu);
    e = E == 0 ? -100 : E - 126;
  }
  if (e < -100) e = -100;
  while (e < 127 && lo + 255.0f * exp_scale((uint32_t)(e + 127)) < hi) ++e;
  return (uint32_t)(e + 127);
}
OL_HD uint32_t quant_lo(float lo, float scale, float v) {
  float f = floorf((v - lo) / scale);
  int q = f < 0.0f ? 0 : (f > 255.0f ? 255 : (int)f);
  while (q > 0 && lo + (float)q * scale > v) --q;
  return (uint32_t)q;
}
OL_HD uint32_t quant_hi(float lo, float scale, float v) {
  float f = ceilf((v - lo) / scale);
  int q = f < 0.0f ? 0 : (f > 255.0f ? 255 : (int)f);
  while (q < 255 && lo + (float)q * scale < v) ++q;
  return (uint32_t)q;
}

__global__ void __launch_bounds__(kBlock) k_collapse_emit(TreeView t, const uint32_t* __restrict__ front, uint32_t count,
                                                          uint32_t level_base, uint32_t next_base,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ next_front, WideNode* __restrict__ nodes) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  uint32_t kids[4];
  int nk = plan_children(t, front[k], kids);
  float4 cmn[4], cmx[4];
  float4 nmn = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), nmx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  const float pad = box_pad(t);
  for (int i = 0; i < nk; ++i) {
    ref_box(kids[i], t.lmin, t.lmax, t.bmin, t.bmax, cmn[i], cmx[i]);
    cmn[i] = make_float4(cmn[i].x - pad, cmn[i].y - pad, cmn[i].z - pad, 0.0f);
    cmx[i] = make_float4(cmx[i].x + pad, cmx[i].y + pad, cmx[i].z + pad, 0.0f);
    nmn = make_float4(fminf(nmn.x, cmn[i].x), fminf(nmn.y, cmn[i].y), fminf(nmn.z, cmn[i].z), 0.0f);
    nmx = make_float4(fmaxf(nmx.x, cmx[i].x), fmaxf(nmx.y, cmx[i].y), fmaxf(nmx.z, cmx[i].z), 0.0f);
  }
  WideNode node;
  node.ox = nmn.x;
  node.oy = nmn.y;
  node.oz = nmn.z;
  uint32_t ex = quant_exponent(nmn.x, nmx.x), ey = quant_exponent(nmn.y, nmx.y), ez = quant_exponent(nmn.z, nmx.z);
  node.exps = ex | (ey << 8) | (ez << 16);
  float sx = exp_scale(ex), sy = exp_scale(ey), sz = exp_scale(ez);
  uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
  uint32_t r = 0;
  for (int i = 0; i < 4; ++i) {
    if (i >= nk) {
      node.child[i] = OL_NONE;
      continue;
    }
    qlo[0] |= quant_lo(nmn.x, sx, cmn[i].x) << (8 * i);
    qlo[1] |= quant_lo(nmn.y, sy, cmn[i].y) << (8 * i);
    qlo[2] |= quant_lo(nmn.z, sz, cmn[i].z) << (8 * i);
    qhi[0] |= quant_hi(nmn.x, sx, cmx[i].x) << (8 * i);
    qhi[1] |= quant_hi(nmn.y, sy, cmx[i].y) << (8 * i);
    qhi[2] |= quant_hi(nmn.z, sz, cmx[i].z) << (8 * i);
    uint32_t ref = kids[i];
    if (ref_is_wide_leaf(t, ref)) {
      node.child[i] = OL_LEAF_BIT | ((ref_count(t, ref) - 1u) << 27) | ref_first(t, ref);
    } else {
      uint32_t slot = offsets[k] + r++;
      node.child[i] = next_base + slot;
      next_front[slot] = ref;
    }
  }
  for (int a = 0; a < 3; ++a) {
    node.qlo[a] = qlo[a];
    node.qhi[a] = qhi[a];
  }
  node.pad[0] = 0;
  node.pad[1] = 0;
  nodes[level_base + k] = node;
}

// ---- 7. leaf-order traversal triangles ----------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_reorder(const uint32_t* __restrict__ sorted, uint32_t n,
                                                    const float4* __restrict__ pos, const uint4* __restrict__ vidx,
                                                    float4* __restrict__ tri) {
  uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  uint32_t o = sorted[k];
  uint4 t = vidx[o];
  float4 a = pos[t.x], b = pos[t.y], c = pos[t.z];
  tri[3 * k + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(o));
  tri[3 * k + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
  tri[3 * k + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
}

inline dim3 grid_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

}  // namespace

namespace {

// Boxes of every leaf and binary node for the current vertex positions, the
// wide collapse, and the leaf-order triangles (steps 5-7 of the build).
void build_boxes(const BvhInput& in, const BvhTopology& tp, BvhOutput& out, hipStream_t s, BvhDebug* debug) {
  const uint32_t n = in.tri_count;
  const uint32_t leaf_size = tp.leaf_size;
  DevBuf tmin, tmax, cb;
  tmin.reserve(sizeof(float4) * n);
  tmax.reserve(sizeof(float4) * n);
  cb.reserve(sizeof(int) * 8);
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, cb.as<int>());
  hipLaunchKernelGGL(k_tri_bounds, grid_for(n), dim3(kBlock), 0, s, in.positions, in.tri_vidx, n, tmin.as<float4>(),
                     tmax.as<float4>(), cb.as<int>());
  const uint32_t ni = n > 1 ? n - 1 : 1;
  DevBuf lmin, lmax, bmin, bmax, flags;
  lmin.reserve(sizeof(float4) * n);
  lmax.reserve(sizeof(float4) * n);
  bmin.reserve(sizeof(float4) * ni);
  bmax.reserve(sizeof(float4) * ni);
  flags.reserve(4ull * ni);
  hipLaunchKernelGGL(k_leaf_boxes, grid_for(n), dim3(kBlock), 0, s, tp.sorted.as<uint32_t>(), n, tmin.as<float4>(),
                     tmax.as<float4>(), lmin.as<float4>(), lmax.as<float4>());
  if (n > 1) {
    OL_HIP_CHECK(hipMemsetAsync(flags.ptr, 0, 4ull * ni, s));
    hipLaunchKernelGGL(k_bottom_up, grid_for(n), dim3(kBlock), 0, s, n, tp.left.as<uint32_t>(),
                       tp.right.as<uint32_t>(), tp.parent_int.as<uint32_t>(), tp.parent_leaf.as<uint32_t>(),
                       lmin.as<float4>(), lmax.as<float4>(), bmin.as<float4>(), bmax.as<float4>(),
                       flags.as<uint32_t>());
  }
  if (debug) {
    OL_HIP_CHECK(hipStreamSynchronize(s));
    debug->sorted_index.resize(n);
    OL_HIP_CHECK(hipMemcpy(debug->sorted_index.data(), tp.sorted.ptr, 4ull * n, hipMemcpyDeviceToHost));
    if (n > 1) {
      debug->left.resize(n - 1);
      debug->right.resize(n - 1);
      debug->node_box.resize(8ull * (n - 1));
      OL_HIP_CHECK(hipMemcpy(debug->left.data(), tp.left.ptr, 4ull * (n - 1), hipMemcpyDeviceToHost));
      OL_HIP_CHECK(hipMemcpy(debug->right.data(), tp.right.ptr, 4ull * (n - 1), hipMemcpyDeviceToHost));
      std::vector<float4> a(n - 1), b(n - 1);
      OL_HIP_CHECK(hipMemcpy(a.data(), bmin.ptr, sizeof(float4) * (n - 1), hipMemcpyDeviceToHost));
      OL_HIP_CHECK(hipMemcpy(b.data(), bmax.ptr, sizeof(float4) * (n - 1), hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i + 1 < n; ++i) {
        float* d = &debug->node_box[8ull * i];
        d[0] = a[i].x; d[1] = a[i].y; d[2] = a[i].z; d[3] = a[i].w;
        d[4] = b[i].x; d[5] = b[i].y; d[6] = b[i].z; d[7] = b[i].w;
      }
    }
  }

  // Wide collapse, one breadth-first level per iteration.
  TreeView tv{tp.left.as<uint32_t>(), tp.right.as<uint32_t>(), tp.rfirst.as<uint32_t>(), tp.rlast.as<uint32_t>(),
              lmin.as<float4>(), lmax.as<float4>(), bmin.as<float4>(), bmax.as<float4>(), leaf_size,
              n > 1 ? 0u : OL_LEAF_BIT};
  out.nodes.reserve(sizeof(WideNode) * (size_t)(n + 1));
  DevBuf front_a, front_b, icount;
  front_a.reserve(4ull * (n + 1));
  front_b.reserve(4ull * (n + 1));
  icount.reserve(4ull * (n + 2));
  uint32_t root = n > 1 ? 0u : OL_LEAF_BIT;
  OL_HIP_CHECK(hipMemcpyAsync(front_a.ptr, &root, 4, hipMemcpyHostToDevice, s));
  uint32_t level_count = 1, level_base = 0, depth = 0;
  uint32_t* cur = front_a.as<uint32_t>();
  uint32_t* nxt = front_b.as<uint32_t>();
  while (level_count > 0) {
    ++depth;
    hipLaunchKernelGGL(k_collapse_plan, grid_for(level_count), dim3(kBlock), 0, s, tv, cur, level_count,
                       icount.as<uint32_t>());
    scan_exclusive_u32(icount.as<uint32_t>(), level_count, icount.as<uint32_t>() + level_count, s);
    uint32_t next_count = 0;
    OL_HIP_CHECK(hipMemcpyAsync(&next_count, icount.as<uint32_t>() + level_count, 4, hipMemcpyDeviceToHost, s));
    OL_HIP_CHECK(hipStreamSynchronize(s));
    uint32_t next_base = level_base + level_count;
    hipLaunchKernelGGL(k_collapse_emit, grid_for(level_count), dim3(kBlock), 0, s, tv, cur, level_count, level_base,
                       next_base, icount.as<uint32_t>(), nxt, out.nodes.as<WideNode>());
    level_base = next_base;
    level_count = next_count;
    std::swap(cur, nxt);
  }
  out.node_count = level_base;
  out.depth = depth;

  out.tri.reserve(sizeof(float4) * 3ull * n);
  hipLaunchKernelGGL(k_reorder, grid_for(n), dim3(kBlock), 0, s, tp.sorted.as<uint32_t>(), n, in.positions,
                     in.tri_vidx, out.tri.as<float4>());
  OL_HIP_CHECK(hipGetLastError());
  OL_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

// Binned SAH over primitive boxes, one primitive per leaf: a top-down split
// of [b, e) at the best of 16 centroid bins on each axis (surface area x
// count), the middle when every centroid coincides.  With `isolate`, one more
// candidate is priced the same way: the primitive of largest box alone
// against the rest.  Centroid bins cannot separate a box that spans much of
// its node from the small boxes around its centre, so without it every node
// on such a primitive's path takes its size.  Internal nodes are numbered in
// creation order from the root (0); leaves are the positions of the final
// order, so every node covers a contiguous leaf range, as in the Karras
// hierarchy.
void sah_topology(const float* box, uint32_t n, std::vector<uint32_t>& order, std::vector<uint32_t>& left,
                  std::vector<uint32_t>& right, std::vector<uint32_t>& rfirst, std::vector<uint32_t>& rlast,
                  std::vector<uint32_t>& parent_int, std::vector<uint32_t>& parent_leaf, bool isolate) {
  constexpr int kBins = 16;
  order.resize(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  const uint32_t ni = n - 1;
  left.assign(ni, 0);
  right.assign(ni, 0);
  rfirst.assign(ni, 0);
  rlast.assign(ni, 0);
  parent_int.assign(ni, OL_NONE);
  parent_leaf.assign(n, OL_NONE);
  auto centroid = [&](uint32_t p, int a) { return 0.5f * (box[6 * p + a] + box[6 * p + 3 + a]); };
  auto area = [](const float* mn, const float* mx) {
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return dx * dy + dy * dz + dz * dx;
  };
  // Node ids follow the tree's pre-order, so they do not depend on the
  // order in which subtrees are built: the node over [b, e) has id X, its
  // left subtree (m leaves) the ids X + 1 .. X + m - 1, its right subtree
  // those after.  Subtrees below the first levels are built in parallel.
  struct Task { uint32_t b, e, node; };
  auto split = [&](const Task& t, std::vector<Task>& out) {
    float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = t.b; k < t.e; ++k)
      for (int a = 0; a < 3; ++a) {
        const float c = centroid(order[k], a);
        cmin[a] = std::min(cmin[a], c);
        cmax[a] = std::max(cmax[a], c);
      }
    float best_cost = INFINITY;
    int best_axis = -1, best_bin = 0;
    for (int a = 0; a < 3; ++a) {
      const float ext = cmax[a] - cmin[a];
      if (!(ext > 0.0f)) continue;
      float bmn[kBins][3], bmx[kBins][3];
      uint32_t cnt[kBins] = {0};
      for (int i = 0; i < kBins; ++i)
        for (int k = 0; k < 3; ++k) { bmn[i][k] = INFINITY; bmx[i][k] = -INFINITY; }
      for (uint32_t k = t.b; k < t.e; ++k) {
        const uint32_t p = order[k];
        const int bin = std::min(kBins - 1, (int)((centroid(p, a) - cmin[a]) / ext * kBins));
        ++cnt[bin];
        for (int q = 0; q < 3; ++q) {
          bmn[bin][q] = std::min(bmn[bin][q], box[6 * p + q]);
          bmx[bin][q] = std::max(bmx[bin][q], box[6 * p + 3 + q]);
        }
      }
      // Sweep: cost of splitting after bin i.
      float rmn[kBins][3], rmx[kBins][3];
      uint32_t rc[kBins];
      float amn[3] = {INFINITY, INFINITY, INFINITY}, amx[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t acc = 0;
      for (int i = kBins - 1; i >= 0; --i) {
        for (int q = 0; q < 3; ++q) { amn[q] = std::min(amn[q], bmn[i][q]); amx[q] = std::max(amx[q], bmx[i][q]); }
        acc += cnt[i];
        for (int q = 0; q < 3; ++q) { rmn[i][q] = amn[q]; rmx[i][q] = amx[q]; }
        rc[i] = acc;
      }
      float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t lc = 0;
      for (int i = 0; i + 1 < kBins; ++i) {
        for (int q = 0; q < 3; ++q) { lmn[q] = std::min(lmn[q], bmn[i][q]); lmx[q] = std::max(lmx[q], bmx[i][q]); }
        lc += cnt[i];
        if (lc == 0 || rc[i + 1] == 0) continue;
        const float cost = area(lmn, lmx) * (float)lc + area(rmn[i + 1], rmx[i + 1]) * (float)rc[i + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = i; }
      }
    }
    // The largest box alone against the rest (when centroid bins gave a
    // candidate: boxes of one centroid keep the middle split).
    uint32_t big = OL_NONE;
    if (isolate && best_axis >= 0 && t.e - t.b > 2) {
      float big_area = -1.0f;
      for (uint32_t k = t.b; k < t.e; ++k) {
        const float a = area(box + 6 * order[k], box + 6 * order[k] + 3);
        if (a > big_area) { big_area = a; big = k; }
      }
      float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t k = t.b; k < t.e; ++k) {
        if (k == big) continue;
        const float* pb = box + 6 * order[k];
        for (int q = 0; q < 3; ++q) { mn[q] = std::min(mn[q], pb[q]); mx[q] = std::max(mx[q], pb[3 + q]); }
      }
      const float cost = big_area + area(mn, mx) * (float)(t.e - t.b - 1);
      if (!(cost < best_cost)) big = OL_NONE;
    }
    uint32_t mid;
    if (big != OL_NONE) {
      std::swap(order[big], order[t.e - 1]);
      mid = t.e - 1;
    } else if (best_axis < 0) {
      mid = t.b + (t.e - t.b) / 2;
    } else {
      const int a = best_axis;
      const float ext = cmax[a] - cmin[a];
      auto it = std::partition(order.begin() + t.b, order.begin() + t.e, [&](uint32_t p) {
        return std::min(kBins - 1, (int)((centroid(p, a) - cmin[a]) / ext * kBins)) <= best_bin;
      });
      mid = (uint32_t)(it - order.begin());
      if (mid == t.b || mid == t.e) mid = t.b + (t.e - t.b) / 2;
    }
    rfirst[t.node] = t.b;
    rlast[t.node] = t.e - 1;
    const uint32_t halves[2][2] = {{t.b, mid}, {mid, t.e}};
    const uint32_t ids[2] = {t.node + 1u, t.node + (mid - t.b)};
    for (int h = 0; h < 2; ++h) {
      const uint32_t b = halves[h][0], e = halves[h][1];
      uint32_t ref;
      if (e - b == 1) {
        ref = b | OL_LEAF_BIT;
        parent_leaf[b] = t.node;
      } else {
        ref = ids[h];
        parent_int[ref] = t.node;
        out.push_back({b, e, ref});
      }
      (h == 0 ? left : right)[t.node] = ref;
    }
  };
  auto build_subtree = [&](Task root) {
    std::vector<Task> stack{root};
    while (!stack.empty()) {
      const Task t = stack.back();
      stack.pop_back();
      split(t, stack);
    }
  };
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<Task> front{{0, n, 0}};
  // Split breadth-first until there is work for every thread.
  while (!front.empty() && front.size() < 4 * threads) {
    std::vector<Task> next;
    for (const Task& t : front) {
      if (t.e - t.b < 4096) next.push_back(t);
      else split(t, next);
    }
    if (next.size() == front.size()) {
      front.swap(next);
      break;
    }
    front.swap(next);
  }
  if (threads == 1 || front.size() < 2) {
    for (const Task& t : front) build_subtree(t);
    return;
  }
  std::atomic<size_t> cursor{0};
  std::vector<std::thread> pool;
  for (unsigned k = 0; k < threads; ++k)
    pool.emplace_back([&] {
      for (size_t i = cursor++; i < front.size(); i = cursor++) build_subtree(front[i]);
    });
  for (auto& th : pool) th.join();
}

namespace {

template <class T>
void upload_vec(DevBuf& d, const std::vector<T>& v, hipStream_t s) {
  d.reserve(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) OL_HIP_CHECK(hipMemcpyAsync(d.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
}

}  // namespace

void bvh_build(const BvhInput& in, BvhOutput& out, hipStream_t s, BvhDebug* debug, bool keep_topology) {
  auto t0 = std::chrono::steady_clock::now();
  const uint32_t n = in.tri_count;
  out.topology.reset();
  if (n == 0) {
    out.node_count = 0;
    out.depth = 0;
    out.nodes.reserve(64);
    out.tri.reserve(48);
    return;
  }
  if (n > OL_LEAF_FIRST_MASK) throw OlError{-1, "too many triangles for one BVH (max 134217727)"};
  auto tp = std::make_unique<BvhTopology>();
  tp->tri_count = n;
  tp->leaf_size = in.leaf_size == 0 ? 4u : (in.leaf_size > OL_LEAF_MAX ? OL_LEAF_MAX : in.leaf_size);
  if (in.host_boxes && n > 1) {
    std::vector<uint32_t> order, l, r, f, la, pi, pl;
    sah_topology(in.host_boxes, n, order, l, r, f, la, pi, pl, in.isolate_large);
    upload_vec(tp->sorted, order, s);
    upload_vec(tp->left, l, s);
    upload_vec(tp->right, r, s);
    upload_vec(tp->rfirst, f, s);
    upload_vec(tp->rlast, la, s);
    upload_vec(tp->parent_int, pi, s);
    upload_vec(tp->parent_leaf, pl, s);
    OL_HIP_CHECK(hipStreamSynchronize(s));
    build_boxes(in, *tp, out, s, debug);
    if (keep_topology) out.topology = std::move(tp);
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return;
  }

  // Steps 1-4: bounds, Morton codes, radix sort, Karras hierarchy.
  DevBuf tmin, tmax, cb, codes, codes_tmp, vals_tmp;
  tmin.reserve(sizeof(float4) * n);
  tmax.reserve(sizeof(float4) * n);
  cb.reserve(sizeof(int) * 8);
  codes.reserve(sizeof(uint32_t) * n);
  tp->sorted.reserve(sizeof(uint32_t) * n);
  codes_tmp.reserve(sizeof(uint32_t) * n);
  vals_tmp.reserve(sizeof(uint32_t) * n);
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, cb.as<int>());
  hipLaunchKernelGGL(k_tri_bounds, grid_for(n), dim3(kBlock), 0, s, in.positions, in.tri_vidx, n, tmin.as<float4>(),
                     tmax.as<float4>(), cb.as<int>());
  hipLaunchKernelGGL(k_morton, grid_for(n), dim3(kBlock), 0, s, tmin.as<float4>(), tmax.as<float4>(), n, cb.as<int>(),
                     codes.as<uint32_t>(), tp->sorted.as<uint32_t>());
  if (debug) {
    OL_HIP_CHECK(hipStreamSynchronize(s));
    debug->morton.resize(n);
    OL_HIP_CHECK(hipMemcpy(debug->morton.data(), codes.ptr, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  }
  radix_sort_pairs(codes.as<uint32_t>(), tp->sorted.as<uint32_t>(), codes_tmp.as<uint32_t>(),
                   vals_tmp.as<uint32_t>(), n, s, 32);
  codes_tmp.release();
  vals_tmp.release();
  tmin.release();
  tmax.release();
  const uint32_t ni = n > 1 ? n - 1 : 1;
  tp->left.reserve(4ull * ni);
  tp->right.reserve(4ull * ni);
  tp->rfirst.reserve(4ull * ni);
  tp->rlast.reserve(4ull * ni);
  tp->parent_int.reserve(4ull * ni);
  tp->parent_leaf.reserve(4ull * n);
  if (n > 1)
    hipLaunchKernelGGL(k_karras, grid_for(n - 1), dim3(kBlock), 0, s, codes.as<uint32_t>(), (int)n,
                       tp->left.as<uint32_t>(), tp->right.as<uint32_t>(), tp->rfirst.as<uint32_t>(),
                       tp->rlast.as<uint32_t>(), tp->parent_int.as<uint32_t>(), tp->parent_leaf.as<uint32_t>());
  // Steps 5-7.
  build_boxes(in, *tp, out, s, debug);
  if (keep_topology) out.topology = std::move(tp);
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void bvh_refit(const BvhInput& in, BvhOutput& out, hipStream_t s) {
  auto t0 = std::chrono::steady_clock::now();
  if (in.tri_count == 0) return;
  if (!out.topology) throw OlError{-1, "the BVH was built without a kept topology"};
  const BvhTopology& tp = *out.topology;
  if (tp.tri_count != in.tri_count) throw OlError{-1, "refit needs the triangle count of the build"};
  build_boxes(in, tp, out, s, nullptr);
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}
