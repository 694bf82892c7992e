// Reference-estimator mode: the diffuse path tracer of
// ordinarylight/renderers/reference/path_tracer.py:31-211 on the GPU.
//
// That estimator consumes one sequential PCG64 stream (camera jitter, then a
// block of normal deviates per bounce for the rays that hit, in ascending
// ray order), so the host draws the stream and this file supplies the
// per-bounce kernels: camera rays, intersection (+ miss radiance), an
// order-preserving compaction of the hit rays, and the hit update.
//
// Precision follows the reference array for array: with a perspective or
// panoramic camera origins/throughput/radiance are float32 and directions
// float64; with an orthographic camera origins/throughput/radiance are
// float64 and directions float32.  The kernels are templated on those two
// element types and promote exactly where NumPy promotes.

#include "e0.h"
#include "trace.h"

namespace {

template <class A, class B>
using Pr = decltype(A() * B());

template <class T>
struct V3 {
  T x, y, z;
};

template <class OT, class DT>
struct E0Ptr {
  OT* o;     // [N][3]
  DT* d;     // [N][3]
  OT* T;     // [N][3]
  OT* rad;   // [N][3]
};

template <class OT, class DT>
OL_DEV E0Ptr<OT, DT> ptrs(const E0State& s) {
  return E0Ptr<OT, DT>{static_cast<OT*>(s.o), static_cast<DT*>(s.d), static_cast<OT*>(s.T), static_cast<OT*>(s.rad)};
}

// ---- camera ---------------------------------------------------------------------------------
template <class OT, class DT>
__global__ void __launch_bounds__(256) k_e0_camera(E0Params P, E0State S, const float* jx, const float* jy) {
  const uint32_t n = P.width * P.height;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  E0Ptr<OT, DT> A = ptrs<OT, DT>(S);
  const uint32_t row = i / P.width, col = i - row * P.width;
  // int64 + float32 -> float64 grid (path_tracer.py:120-121).
  const double ndx = (((double)col + (double)jx[i]) / (double)P.width) * 2.0 - 1.0;
  const double ndy = 1.0 - (((double)row + (double)jy[i]) / (double)P.height) * 2.0;
  const double fx = P.fwd[0], fy = P.fwd[1], fz = P.fwd[2];
  const double rx = P.right[0], ry = P.right[1], rz = P.right[2];
  const double ux = P.up[0], uy = P.up[1], uz = P.up[2];
  if (P.projection == 1u) {
    // Orthographic: float64 origins, float32 unit forward directions.
    const double a = ndx * P.k_x, b = ndy * P.k_y;
    A.o[3 * i + 0] = (OT)(((double)P.pos[0] + a * rx) + b * ux);
    A.o[3 * i + 1] = (OT)(((double)P.pos[1] + a * ry) + b * uy);
    A.o[3 * i + 2] = (OT)(((double)P.pos[2] + a * rz) + b * uz);
    const float f0 = P.fwd[0], f1 = P.fwd[1], f2 = P.fwd[2];
    const float len = sqrtf((f0 * f0 + f1 * f1) + f2 * f2);
    A.d[3 * i + 0] = (DT)(f0 / len);
    A.d[3 * i + 1] = (DT)(f1 / len);
    A.d[3 * i + 2] = (DT)(f2 / len);
  } else {
    double dx, dy, dz;
    if (P.projection == 2u) {
      const double yaw = (ndx * P.k_x) * 0.5, pitch = (ndy * P.k_y) * 0.5;
      const double cp = cos(pitch), sp = sin(pitch), cy = cos(yaw), sy = sin(yaw);
      const double a = cp * cy, b = cp * sy;
      dx = (a * fx + b * rx) + sp * ux;
      dy = (a * fy + b * ry) + sp * uy;
      dz = (a * fz + b * rz) + sp * uz;
    } else {
      const double a = ndx * P.k_x, b = ndy * P.k_y;
      dx = (fx + a * rx) + b * ux;
      dy = (fy + a * ry) + b * uy;
      dz = (fz + a * rz) + b * uz;
    }
    const double len = sqrt((dx * dx + dy * dy) + dz * dz);
    A.d[3 * i + 0] = (DT)(dx / len);
    A.d[3 * i + 1] = (DT)(dy / len);
    A.d[3 * i + 2] = (DT)(dz / len);
    A.o[3 * i + 0] = (OT)P.pos[0];
    A.o[3 * i + 1] = (OT)P.pos[1];
    A.o[3 * i + 2] = (OT)P.pos[2];
  }
  for (int k = 0; k < 3; ++k) {
    A.T[3 * i + k] = (OT)1;
    A.rad[3 * i + k] = (OT)0;
  }
  S.active[i] = i;
}

// ---- triangle test with the reference's dtypes -----------------------------------------------------
template <class OT, class DT>
struct E0Hit {
  Pr<OT, DT> t;
  bool hit;
};

template <class OT, class DT>
OL_DEV E0Hit<OT, DT> e0_tri(const V3<OT>& o, const V3<DT>& d, const float4 a, const float4 e1, const float4 e2) {
  using U = Pr<OT, DT>;
  E0Hit<OT, DT> r;
  r.hit = false;
  r.t = (U)0;
  const DT px = d.y * (DT)e2.z - d.z * (DT)e2.y;
  const DT py = d.z * (DT)e2.x - d.x * (DT)e2.z;
  const DT pz = d.x * (DT)e2.y - d.y * (DT)e2.x;
  const DT det = (px * (DT)e1.x + py * (DT)e1.y) + pz * (DT)e1.z;
  const bool valid = (det < (DT)0 ? -det : det) > (DT)1e-7;
  if (!valid) return r;
  const DT inv = (DT)1 / det;
  const OT tx = o.x - (OT)a.x, ty = o.y - (OT)a.y, tz = o.z - (OT)a.z;
  const U u = (((U)tx * (U)px + (U)ty * (U)py) + (U)tz * (U)pz) * (U)inv;
  const OT qx = ty * (OT)e1.z - tz * (OT)e1.y;
  const OT qy = tz * (OT)e1.x - tx * (OT)e1.z;
  const OT qz = tx * (OT)e1.y - ty * (OT)e1.x;
  const U v = (((U)d.x * (U)qx + (U)d.y * (U)qy) + (U)d.z * (U)qz) * (U)inv;
  const OT qe = (qx * (OT)e2.x + qy * (OT)e2.y) + qz * (OT)e2.z;
  const U t = (U)qe * (U)inv;
  r.t = t;
  r.hit = (u >= (U)0) && (v >= (U)0) && (u + v <= (U)1) && (t > (U)1e-4);
  return r;
}

// Miss radiance (path_tracer.py:184-202), added with the reference's promotions.
template <class OT, class DT>
OL_DEV void e0_sky(const E0Params& P, const V3<DT>& d, OT* T, OT* rad) {
  if (P.env_mode == 0u) {
    // Gradient: blend in the direction dtype, sky in float64.
    DT blend = (DT)0.5 * (d.y + (DT)1);
    blend = blend < (DT)0 ? (DT)0 : (blend > (DT)1 ? (DT)1 : blend);
    const double ob = (double)((DT)1 - blend), bb = (double)blend;
    const double sky[3] = {ob * 0.8 + bb * 0.2, ob * 0.8 + bb * 0.4, ob * 0.8 + bb * 0.9};
    for (int k = 0; k < 3; ++k) rad[k] = (OT)((double)rad[k] + (double)T[k] * sky[k]);
    return;
  }
  float base[3] = {1.0f, 1.0f, 1.0f};
  if (P.env_mode == 2u) {
    // Nearest texel of the equirectangular image.
    DT lon = atan2(d.z, d.x);
    lon = lon + (DT)P.env_rotation;
    DT uu = lon / (DT)(2.0 * 3.141592653589793) + (DT)0.5;
    DT m = fmod(uu, (DT)1);
    if (m != (DT)0 && m < (DT)0) m += (DT)1;
    if (m == (DT)0) m = (DT)0;
    DT cy = d.y < (DT)-1 ? (DT)-1 : (d.y > (DT)1 ? (DT)1 : d.y);
    DT vv = acos(cy) / (DT)3.141592653589793;
    long long x = (long long)(m * (DT)P.env_w);
    long long y = (long long)(vv * (DT)P.env_h);
    if (x > (long long)P.env_w - 1) x = P.env_w - 1;
    if (y > (long long)P.env_h - 1) y = P.env_h - 1;
    const float* px = P.env_image + 3ull * ((size_t)y * P.env_w + (size_t)x);
    base[0] = px[0];
    base[1] = px[1];
    base[2] = px[2];
  }
  for (int k = 0; k < 3; ++k) {
    const float sky = (base[k] * P.env_color[k]) * P.env_intensity;
    rad[k] = rad[k] + T[k] * (OT)sky;
  }
}

template <int STACK, int BLOCK, class OT, class DT>
__global__ void __launch_bounds__(BLOCK) k_e0_trace(E0Params P, E0State S, uint32_t n_active) {
  __shared__ uint32_t stack[STACK * BLOCK];
  using U = Pr<OT, DT>;
  E0Ptr<OT, DT> A = ptrs<OT, DT>(S);
  for (uint32_t j = blockIdx.x * BLOCK + threadIdx.x; j < n_active; j += gridDim.x * BLOCK) {
    const uint32_t r = S.active[j];
    const V3<OT> o{A.o[3 * r], A.o[3 * r + 1], A.o[3 * r + 2]};
    const V3<DT> d{A.d[3 * r], A.d[3 * r + 1], A.d[3 * r + 2]};
    float best_t = INFINITY;
    int best = -1;
    if (P.brute_force) {
      // Index order with the float32 running best (path_tracer.py:159-176).
      for (uint32_t k = 0; k < P.sc.tri_count; ++k) {
        const uint4 vi = P.sc.tri_vidx[k];
        const float4 a = P.positions[vi.x], b = P.positions[vi.y], c = P.positions[vi.z];
        const float4 e1 = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
        const float4 e2 = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
        E0Hit<OT, DT> h = e0_tri<OT, DT>(o, d, a, e1, e2);
        if (h.hit && h.t < (U)best_t) {
          best_t = (float)h.t;
          best = (int)k;
        }
      }
    } else if (P.sc.node_count > 0) {
      // BVH: lexicographic minimum of (t, index), boxes tested in float32.
      const f3 of = mk3((float)o.x, (float)o.y, (float)o.z);
      const f3 df = mk3((float)d.x, (float)d.y, (float)d.z);
      const RayPre rp = ray_pre(of, df);
      U bt = (U)INFINITY;
      uint32_t bi = OL_NONE;
      uint32_t* col = stack + threadIdx.x;
      uint32_t sp = 0, node = 0;
      while (true) {
        uint4 w0, w1, w2, w3;
        load_node(P.sc.nodes, node, w0, w1, w2, w3);
        uint32_t ch[4], key[4];
        node_children(w2, w3, ch);
        const float cull = bt < (U)3.0e38 ? cull_distance((float)bt) + 1e-6f : INFINITY;
        ol_node_keys(w0, w1, w2, rp, 0.0f, cull, key, ch);
        uint32_t next = OL_NONE;
        for (int c = 0; c < 4; ++c) {
          if (key[c] == OL_NONE) continue;
          if (child_is_leaf(ch[c])) {
            for (uint32_t k = 0; k < leaf_count(ch[c]); ++k) {
              const uint32_t slot = leaf_first(ch[c]) + k;
              const float4 a = P.sc.tri[3 * slot], e1 = P.sc.tri[3 * slot + 1], e2 = P.sc.tri[3 * slot + 2];
              E0Hit<OT, DT> h = e0_tri<OT, DT>(o, d, a, e1, e2);
              const uint32_t orig = __float_as_uint(a.w);
              if (h.hit && (h.t < bt || (h.t == bt && orig < bi))) {
                bt = h.t;
                bi = orig;
              }
            }
          } else if (next == OL_NONE) {
            next = ch[c];
          } else if (sp < (uint32_t)STACK) {
            col[(sp++) * BLOCK] = ch[c];
          } else {
            atomicOr(&S.ctr[1], 1u);
          }
        }
        if (next != OL_NONE) {
          node = next;
          continue;
        }
        if (sp == 0) break;
        node = col[(--sp) * BLOCK];
      }
      if (bi != OL_NONE) {
        best_t = (float)bt;
        best = (int)bi;
      }
    }
    S.ht[j] = best_t;
    S.htri[j] = best;
    S.flags[j] = best >= 0 ? 1u : 0u;
    if (best < 0) e0_sky<OT, DT>(P, d, A.T + 3 * r, A.rad + 3 * r);
  }
}

__global__ void __launch_bounds__(256) k_e0_compact(E0State S, uint32_t n_active) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_active) return;
  if (S.htri[j] < 0) return;
  const uint32_t k = S.scan[j];
  S.hitlist[k] = S.active[j];
  S.hit_t[k] = S.ht[j];
  S.hit_tri[k] = S.htri[j];
}

template <class OT, class DT>
__global__ void __launch_bounds__(256) k_e0_shade(E0Params P, E0State S, uint32_t n_hit, const float* gauss) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_hit) return;
  E0Ptr<OT, DT> A = ptrs<OT, DT>(S);
  const uint32_t r = S.hitlist[k];
  const uint32_t tri = (uint32_t)S.hit_tri[k];
  const float t = S.hit_t[k];
  const uint4 vi = P.sc.tri_vidx[tri];
  const float4 a = P.positions[vi.x], b = P.positions[vi.y], c = P.positions[vi.z];
  // points = origins + (directions * t): the product in promote(DT, float32),
  // the sum in promote(OT, that).
  using DTt = Pr<DT, float>;
  using U2 = Pr<OT, DTt>;
  U2 pt[3];
  for (int q = 0; q < 3; ++q) pt[q] = (U2)A.o[3 * r + q] + (U2)((DTt)A.d[3 * r + q] * (DTt)t);
  // Unit geometric normal in float32, flipped to face the incoming ray.
  const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
  const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
  float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const float nl = sqrtf((nx * nx + ny * ny) + nz * nz);
  nx = nx / nl;
  ny = ny / nl;
  nz = nz / nl;
  const Pr<float, DT> facing = ((Pr<float, DT>)nx * (Pr<float, DT>)A.d[3 * r] + (Pr<float, DT>)ny * (Pr<float, DT>)A.d[3 * r + 1]) +
                               (Pr<float, DT>)nz * (Pr<float, DT>)A.d[3 * r + 2];
  if (facing > (Pr<float, DT>)0) {
    nx = nx * -1.0f;
    ny = ny * -1.0f;
    nz = nz * -1.0f;
  }
  const float4 em = P.sc.materials[12u * vi.w + 1u];
  const float4 co = P.sc.face_orig ? P.sc.face_orig[tri] : P.sc.materials[12u * vi.w];
  const float emis[3] = {em.x, em.y, em.z}, col[3] = {co.x, co.y, co.z};
  const float nrm[3] = {nx, ny, nz};
  for (int q = 0; q < 3; ++q) {
    A.rad[3 * r + q] = A.rad[3 * r + q] + A.T[3 * r + q] * (OT)emis[q];
    A.T[3 * r + q] = A.T[3 * r + q] * (OT)col[q];
    const float off = nrm[q] * 9.99999974737875163555e-05f;  // float32(1e-4)
    A.o[3 * r + q] = (OT)(pt[q] + (U2)off);
  }
  // Cosine-weighted continuation from the host-drawn deviates (float32).
  float sx = gauss[3 * k], sy = gauss[3 * k + 1], sz = gauss[3 * k + 2];
  const float sl = sqrtf((sx * sx + sy * sy) + sz * sz);
  sx = sx / sl;
  sy = sy / sl;
  sz = sz / sl;
  if ((sx * nx + sy * ny) + sz * nz < 0.0f) {
    sx = sx * -1.0f;
    sy = sy * -1.0f;
    sz = sz * -1.0f;
  }
  float dx = nx + sx, dy = ny + sy, dz = nz + sz;
  const float dl = sqrtf((dx * dx + dy * dy) + dz * dz);
  A.d[3 * r] = (DT)(dx / dl);
  A.d[3 * r + 1] = (DT)(dy / dl);
  A.d[3 * r + 2] = (DT)(dz / dl);
}

template <class OT, class DT>
__global__ void __launch_bounds__(256) k_e0_accumulate(E0Params P, E0State S) {
  const uint32_t n = 3u * P.width * P.height;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const OT* rad = static_cast<const OT*>(S.rad);
  S.acc[i] = (float)((Pr<float, OT>)S.acc[i] + (Pr<float, OT>)rad[i]);
}

__global__ void __launch_bounds__(256) k_e0_finish(E0Params P, E0State S, uint32_t samples, float4* out) {
  const uint32_t n = P.width * P.height;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float s = (float)samples;
  out[i] = make_float4(S.acc[3 * i] / s, S.acc[3 * i + 1] / s, S.acc[3 * i + 2] / s, 0.0f);
}

inline dim3 g256(uint32_t n) { return dim3((n + 255u) / 256u); }

template <int STACK, int BLOCK, class OT, class DT>
void launch_e0_trace(int cus, const E0Params& p, const E0State& s, uint32_t n, hipStream_t st) {
  static int occ = -1;
  if (occ < 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_e0_trace<STACK, BLOCK, OT, DT>, BLOCK, 0) != hipSuccess || occ < 1) occ = 1;
  }
  uint32_t need = (n + BLOCK - 1) / BLOCK;
  uint32_t cap = (uint32_t)(cus * occ);
  hipLaunchKernelGGL((k_e0_trace<STACK, BLOCK, OT, DT>), dim3(need < cap ? need : cap), dim3(BLOCK), 0, st, p, s, n);
}

template <class OT, class DT>
void e0_trace_typed(int cus, int stack, const E0Params& p, const E0State& s, uint32_t n, hipStream_t st) {
  switch (stack) {
    case 16: launch_e0_trace<16, 128, OT, DT>(cus, p, s, n, st); break;
    case 24: launch_e0_trace<24, 128, OT, DT>(cus, p, s, n, st); break;
    case 32: launch_e0_trace<32, 128, OT, DT>(cus, p, s, n, st); break;
    case 48: launch_e0_trace<48, 128, OT, DT>(cus, p, s, n, st); break;
    case 64: launch_e0_trace<64, 64, OT, DT>(cus, p, s, n, st); break;
    case 96: launch_e0_trace<96, 64, OT, DT>(cus, p, s, n, st); break;
    case 128: launch_e0_trace<128, 64, OT, DT>(cus, p, s, n, st); break;
    default: throw OlError{-2, "unsupported traversal stack depth"};
  }
}

}  // namespace

void e0_camera(const E0Params& p, const E0State& s, const float* jx, const float* jy, hipStream_t st) {
  const uint32_t n = p.width * p.height;
  if (p.projection == 1u)
    hipLaunchKernelGGL((k_e0_camera<double, float>), g256(n), dim3(256), 0, st, p, s, jx, jy);
  else
    hipLaunchKernelGGL((k_e0_camera<float, double>), g256(n), dim3(256), 0, st, p, s, jx, jy);
}

void e0_trace(int cus, int stack, const E0Params& p, const E0State& s, uint32_t n_active, hipStream_t st) {
  if (n_active == 0) return;
  if (p.projection == 1u) e0_trace_typed<double, float>(cus, stack, p, s, n_active, st);
  else e0_trace_typed<float, double>(cus, stack, p, s, n_active, st);
}

uint32_t e0_compact(const E0State& s, uint32_t n_active, hipStream_t st) {
  if (n_active == 0) return 0;
  OL_HIP_CHECK(hipMemcpyAsync(s.scan, s.flags, 4ull * n_active, hipMemcpyDeviceToDevice, st));
  scan_exclusive_u32(s.scan, n_active, s.scan + n_active, st);
  hipLaunchKernelGGL(k_e0_compact, g256(n_active), dim3(256), 0, st, s, n_active);
  uint32_t n_hit = 0;
  OL_HIP_CHECK(hipMemcpyAsync(&n_hit, s.scan + n_active, 4, hipMemcpyDeviceToHost, st));
  OL_HIP_CHECK(hipStreamSynchronize(st));
  return n_hit;
}

void e0_shade(const E0Params& p, const E0State& s, uint32_t n_hit, const float* gauss, hipStream_t st) {
  if (n_hit == 0) return;
  if (p.projection == 1u)
    hipLaunchKernelGGL((k_e0_shade<double, float>), g256(n_hit), dim3(256), 0, st, p, s, n_hit, gauss);
  else
    hipLaunchKernelGGL((k_e0_shade<float, double>), g256(n_hit), dim3(256), 0, st, p, s, n_hit, gauss);
}

void e0_accumulate(const E0Params& p, const E0State& s, hipStream_t st) {
  const uint32_t n = 3u * p.width * p.height;
  if (p.projection == 1u)
    hipLaunchKernelGGL((k_e0_accumulate<double, float>), g256(n), dim3(256), 0, st, p, s);
  else
    hipLaunchKernelGGL((k_e0_accumulate<float, double>), g256(n), dim3(256), 0, st, p, s);
}

void e0_finish(const E0Params& p, const E0State& s, uint32_t samples, float4* out, hipStream_t st) {
  hipLaunchKernelGGL(k_e0_finish, g256(p.width * p.height), dim3(256), 0, st, p, s, samples, out);
}
