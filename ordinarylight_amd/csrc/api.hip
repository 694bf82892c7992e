// C ABI of the olhip library (include/olhip.h): context, scene upload and
// BVH build, GI frames, closest-hit queries and the reference-estimator mode.

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/olhip.h"
#include "e0.h"
#include "engine.h"
#include "gi.h"
#include "internal.h"
#include "output.h"

// Traversal-stack spill of olhip_trace_device, which allocates nothing: one
// array per device for the life of the process.  Host code never uses this
// name as an address; it uses the per-device address that
// trace_scratch_address() takes once per device with that device current.
static __device__ __align__(256) uint32_t scratch[OLHIP_TRACE_SCRATCH_BYTES / 4u];

namespace {

std::mutex g_scratch_mu;
std::vector<uint32_t*> g_scratch_ptr;  // by device index; null until taken

// Device address of `scratch` on `device` (which must be current).
uint32_t* trace_scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr.size() <= (size_t)device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    OL_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw OlError{OLHIP_ERR_DEVICE, "trace scratch array has no device address"};
    g_scratch_ptr[device] = static_cast<uint32_t*>(ptr);
  }
  return g_scratch_ptr[device];
}

__global__ void k_clear_word(uint32_t* word) { *word = 0u; }

thread_local std::string g_last_error;

int fail(int code, const std::string& message) {
  g_last_error = message;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return OLHIP_OK;
  } catch (const OlError& e) {
    return fail(e.code, e.message);
  } catch (const std::bad_alloc&) {
    return fail(OLHIP_ERR_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(OLHIP_ERR_RUNTIME, e.what());
  }
}

void require(bool ok, const std::string& message) {
  if (!ok) throw OlError{OLHIP_ERR_VALUE, message};
}

template <class T>
void upload(DevBuf& buf, const T* host, size_t count, hipStream_t s) {
  buf.reserve(sizeof(T) * std::max<size_t>(count, 1));
  if (count) OL_HIP_CHECK(hipMemcpyAsync(buf.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

struct Event {
  hipEvent_t e = nullptr;
  Event() { OL_HIP_CHECK(hipEventCreate(&e)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
};

float elapsed(const Event& a, const Event& b) {
  float ms = 0.0f;
  OL_HIP_CHECK(hipEventElapsedTime(&ms, a.e, b.e));
  return ms;
}

constexpr uint64_t kFnvBasis = 1469598103934665603ull;
// Emitter selection guide (DScene::area_guide): entry j is the first emitter
// whose cumulative weight (record word 5 * i + 4 .x) is >= j / size, clamped
// to the last emitter; entry `size` closes the last bucket.  j / size is
// exact in float32 for a power-of-two size.
void upload_area_guide(olhip_scene* sc, const float* area_lights, uint32_t area_count, hipStream_t s) {
  if (area_count == 0) {
    sc->view.area_guide = nullptr;
    sc->view.area_guide_size = 0;
    return;
  }
  uint32_t size = kAreaGuideMin;
  while (size < kAreaGuideMax && size < 2ull * area_count) size <<= 1;
  std::vector<uint32_t> guide(size + 1u);
  uint32_t i = 0;
  for (uint32_t j = 0; j <= size; ++j) {
    const float x = (float)j / (float)size;
    while (i + 1u < area_count && !(x <= area_lights[20ull * i + 16])) ++i;
    guide[j] = i;
  }
  upload(sc->area_guide, guide.data(), guide.size(), s);
  OL_HIP_CHECK(hipStreamSynchronize(s));
  sc->view.area_guide = sc->area_guide.as<uint32_t>();
  sc->view.area_guide_size = size;
}

uint64_t fnv1a(const void* data, size_t bytes, uint64_t h) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < bytes; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

bool material_textured(const float* record) { return record[8 * 4 + 3] == 1.0f; }

// Checks a texture word array (header, metadata, both mip chains inside the
// array) so that every fetch the sampler can make stays in bounds.
void validate_texture_words(const uint32_t* w, uint32_t n) {
  require(w != nullptr && n >= 1, "texture words missing");
  const uint64_t count = w[0];
  require(1 + 8 * count <= n, "texture word header exceeds the array");
  for (uint64_t k = 0; k < count; ++k) {
    const uint32_t* m = w + 1 + 8 * k;
    const uint64_t width = m[2], height = m[3], levels = m[5];
    require(width >= 1 && height >= 1 && width <= 65536 && height <= 65536, "texture size out of range");
    uint64_t need = 0, lw = width, lh = height, expect = 1;
    while (lw > 1 || lh > 1) {
      lw = std::max<uint64_t>((lw + 1) / 2, 1);
      lh = std::max<uint64_t>((lh + 1) / 2, 1);
      ++expect;
    }
    require(levels == expect, "texture mip level count does not match its size");
    lw = width;
    lh = height;
    for (uint64_t l = 0; l < levels; ++l) {
      need += lw * lh;
      lw = std::max<uint64_t>((lw + 1) / 2, 1);
      lh = std::max<uint64_t>((lh + 1) / 2, 1);
    }
    require((uint64_t)m[0] + need <= n && (uint64_t)m[1] + need <= n, "texture mip chain exceeds the word array");
  }
}

// Texture words (a flat scene's material textures and an image environment,
// or a two-level scene's image environment) and, when some material has
// bindings, the bindings and per-vertex inputs.  Returns the texture count.
uint32_t upload_textures(olhip_scene* sc, uint32_t word_count, const uint32_t* words, uint32_t material_count,
                         const float* materials, const olhip_scene_desc* flat, hipStream_t s) {
  DScene& v = sc->view;
  uint32_t tex_count = 0;
  bool textured = false;
  for (uint32_t m = 0; m < material_count; ++m) textured = textured || material_textured(materials + 48ull * m);
  if (word_count) {
    validate_texture_words(words, word_count);
    tex_count = words[0];
    upload(sc->tex_words, words, word_count, s);
  }
  v.tex_words = word_count ? sc->tex_words.as<uint32_t>() : nullptr;
  v.textured = textured ? 1u : 0u;
  v.bindings = nullptr;
  v.uvs = v.tangents = nullptr;
  v.uv_density = nullptr;
  v.positions = sc->positions.as<float4>();
  if (!textured) return tex_count;
  require(flat != nullptr, "materials with texture bindings need a flat scene");
  const olhip_scene_desc* d = flat;
  require(d->binding_count >= 1 && d->bindings, "texture bindings missing");
  require(d->uvs && d->tangents && d->uv_density, "textured materials need uvs, tangents and uv density");
  for (uint32_t b = 0; b < d->binding_count; ++b) {
    const float t = d->bindings[8ull * b];
    require(t < 0.0f || (uint32_t)t < tex_count, "binding texture index out of range");
  }
  for (uint32_t m = 0; m < material_count; ++m) {
    const float* r = materials + 48ull * m;
    if (!material_textured(r)) continue;
    for (const float b : {r[16], r[17], r[18], r[19], r[21], r[23]})
      require(b < 0.0f || (uint32_t)b < d->binding_count, "material texture binding out of range");
  }
  upload(sc->bindings, reinterpret_cast<const float4*>(d->bindings), 2ull * d->binding_count, s);
  upload(sc->uvs, reinterpret_cast<const float4*>(d->uvs), d->vertex_count, s);
  upload(sc->tangents, reinterpret_cast<const float4*>(d->tangents), d->vertex_count, s);
  upload(sc->uv_density, d->uv_density, d->triangle_count, s);
  v.bindings = sc->bindings.as<float4>();
  v.uvs = sc->uvs.as<float4>();
  v.tangents = sc->tangents.as<float4>();
  v.uv_density = sc->uv_density.as<float>();
  return tex_count;
}

// Per-triangle material parameters of a flat scene (original order on the
// device: base colour and roughness as float4, metallic as float), or none.
void upload_faces(olhip_scene* sc, const olhip_scene_desc* d, hipStream_t s) {
  if (!d->face_params) {
    sc->face_orig.release();
    sc->face_metal.release();
    return;
  }
  const uint32_t T = d->triangle_count;
  std::vector<float4> f(T);
  std::vector<float> m(T);
  for (uint32_t t = 0; t < T; ++t) {
    const float* q = d->face_params + 5ull * t;
    for (int k = 0; k < 5; ++k) require(std::isfinite(q[k]), "per-triangle material parameters must be finite");
    f[t] = make_float4(q[0], q[1], q[2], q[3]);
    m[t] = q[4];
  }
  upload(sc->face_orig, f.data(), f.size(), s);
  upload(sc->face_metal, m.data(), m.size(), s);
}

// Environment light of the analytic records (type 3): constant radiance, or
// an image when row 3 names a texture of `texture_count`.
void set_environment(DScene& v, const float* analytic, uint32_t count, uint32_t texture_count) {
  v.env_valid = 0;
  v.env_texture = -1;
  for (uint32_t l = 0; l < count; ++l) {
    const float* L = analytic + 16ull * l;
    if ((int)(L[3] + 0.5f) != 3) continue;
    v.env_valid = 1;
    v.env_r = (1.0f * L[8]) * L[11];
    v.env_g = (1.0f * L[9]) * L[11];
    v.env_b = (1.0f * L[10]) * L[11];
    v.env_color_r = L[8];
    v.env_color_g = L[9];
    v.env_color_b = L[10];
    v.env_intensity = L[11];
    if (L[12] >= 0.0f) {
      require((uint32_t)L[12] < texture_count, "environment texture index out of range");
      v.env_texture = (int32_t)L[12];
      v.env_rotation = L[13];
      v.env_log_range = L[14];
    }
    return;
  }
}

}  // namespace

extern "C" {

const char* olhip_last_error(void) { return g_last_error.c_str(); }

int olhip_version(void) { return 100; }

int olhip_device_count(int* count) {
  return guarded([&] {
    require(count != nullptr, "count must not be NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int olhip_create(int device, olhip_ctx** out) {
  return guarded([&] {
    require(out != nullptr, "out must not be NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw OlError{OLHIP_ERR_RUNTIME, "no HIP device is available"};
    require(device >= 0 && device < n, "device index out of range");
    OL_HIP_CHECK(hipSetDevice(device));
    auto ctx = std::make_unique<olhip_ctx>();
    ctx->device = device;
    OL_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    OL_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    ctx->cu_count = prop.multiProcessorCount;
    ctx->trace_scratch = trace_scratch_address(device);
    *out = ctx.release();
  });
}

int olhip_destroy(olhip_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->frame_engine.reset();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int olhip_set_profiling(olhip_ctx* ctx, int enabled) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    ctx->profiling = enabled ? 1 : 0;
  });
}

int olhip_set_work_counters(olhip_ctx* ctx, int enabled) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    ctx->counting = enabled ? 1 : 0;
  });
}

int olhip_set_fused_primary(olhip_ctx* ctx, int enabled) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    const int on = enabled ? 1 : 0;
    if (ctx->fused_primary != on) ctx->frame_engine.reset();
    ctx->fused_primary = on;
  });
}

int olhip_set_traversal(olhip_ctx* ctx, int lds_entries, int block) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    const int pairs[][2] = {{8, 128}, {12, 128}, {16, 64}, {16, 128}, {16, 256}, {24, 128}, {32, 128}, {48, 128}};
    bool ok = false;
    for (auto& p : pairs) ok = ok || (p[0] == lds_entries && p[1] == block);
    require(ok, "unsupported (LDS stack entries, block size): use (8|12|24|32|48, 128) or (16, 64|128|256)");
    if (ctx->lds_stack != lds_entries || ctx->trace_block != block) ctx->frame_engine.reset();
    ctx->lds_stack = lds_entries;
    ctx->trace_block = block;
  });
}

int olhip_scene_create(olhip_ctx* ctx, const olhip_scene_desc* d, olhip_scene** out) {
  return guarded([&] {
    require(ctx && d && out, "ctx, desc and out must not be NULL");
    require(d->triangle_count == 0 || (d->positions && d->indices && d->material_ids), "geometry arrays missing");
    require(d->vertex_count == 0 || d->normals, "normals missing");
    require(d->material_count > 0 && d->materials, "at least one material is required");
    require(d->analytic_light_count <= 64, "at most 64 analytic lights are supported");
    require(d->analytic_light_count == 0 || d->analytic_lights, "analytic lights missing");
    require(d->area_light_count == 0 || d->area_lights, "area lights missing");
    require(d->leaf_size <= 16, "leaf_size must be 0..16");
    require(d->builder <= OLHIP_BUILDER_SAH, "unknown BVH builder");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    auto sc = std::make_unique<olhip_scene>();
    sc->ctx = ctx;
    sc->tri_count = d->triangle_count;
    sc->vertex_count = d->vertex_count;
    // Host-side validation and packing to 16-byte records.
    std::vector<float4> pos(d->vertex_count), nrm(d->vertex_count);
    for (uint32_t i = 0; i < d->vertex_count; ++i) {
      pos[i] = make_float4(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2], 0.0f);
      nrm[i] = make_float4(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2], 0.0f);
      require(std::isfinite(pos[i].x) && std::isfinite(pos[i].y) && std::isfinite(pos[i].z),
              "vertex positions must be finite");
    }
    std::vector<uint4> vidx(d->triangle_count);
    for (uint32_t t = 0; t < d->triangle_count; ++t) {
      uint32_t a = d->indices[3 * t], b = d->indices[3 * t + 1], c = d->indices[3 * t + 2];
      require(a < d->vertex_count && b < d->vertex_count && c < d->vertex_count, "triangle index out of range");
      require(d->material_ids[t] < d->material_count, "material id out of range");
      vidx[t] = make_uint4(a, b, c, d->material_ids[t]);
    }
    sc->materials_host.assign(d->materials, d->materials + 48ull * d->material_count);
    for (uint32_t m = 0; m < d->material_count; ++m)
      if (d->materials[48ull * m + 11] > 0.001f) sc->has_transmission = true;
    upload(sc->positions, pos.data(), pos.size(), s);
    upload(sc->normals, nrm.data(), nrm.size(), s);
    upload(sc->vidx, vidx.data(), vidx.size(), s);
    upload(sc->materials, reinterpret_cast<const float4*>(d->materials), 12ull * d->material_count, s);
    std::vector<float4> analytic(4ull * std::max<uint32_t>(d->analytic_light_count, 1), make_float4(0, 0, 0, 0));
    if (d->analytic_light_count)
      std::memcpy(analytic.data(), d->analytic_lights, sizeof(float) * 16ull * d->analytic_light_count);
    upload(sc->analytic, analytic.data(), analytic.size(), s);
    std::vector<float4> area(5ull * std::max<uint32_t>(d->area_light_count, 1), make_float4(0, 0, 0, 0));
    if (d->area_light_count) std::memcpy(area.data(), d->area_lights, sizeof(float) * 20ull * d->area_light_count);
    upload(sc->area, area.data(), area.size(), s);
    upload_area_guide(sc.get(), d->area_lights, d->area_light_count, s);
    OL_HIP_CHECK(hipStreamSynchronize(s));
    sc->structure_hash = fnv1a(d->indices, 12ull * d->triangle_count, kFnvBasis);
    sc->leaf_size = d->leaf_size;
    std::vector<float> boxes;
    if (d->builder == OLHIP_BUILDER_SAH && d->triangle_count > 1) {
      boxes.resize(6ull * d->triangle_count);
      for (uint32_t t = 0; t < d->triangle_count; ++t) {
        const float4 a = pos[vidx[t].x], b = pos[vidx[t].y], c = pos[vidx[t].z];
        float* o = &boxes[6ull * t];
        o[0] = std::min(a.x, std::min(b.x, c.x));
        o[1] = std::min(a.y, std::min(b.y, c.y));
        o[2] = std::min(a.z, std::min(b.z, c.z));
        o[3] = std::max(a.x, std::max(b.x, c.x));
        o[4] = std::max(a.y, std::max(b.y, c.y));
        o[5] = std::max(a.z, std::max(b.z, c.z));
      }
    }
    BvhInput in{sc->positions.as<float4>(), sc->vidx.as<uint4>(), d->triangle_count, d->leaf_size,
                boxes.empty() ? nullptr : boxes.data()};
    bvh_build(in, sc->bvh, s, nullptr, true);
    upload_faces(sc.get(), d, s);
    build_shade_records(sc->bvh, sc->vidx.as<uint4>(), sc->normals.as<float4>(), d->triangle_count, true, s,
                        d->face_params ? sc->face_orig.as<float4>() : nullptr,
                        d->face_params ? sc->face_metal.as<float>() : nullptr);
    DScene& v = sc->view;
    v.nodes = sc->bvh.nodes.as<WideNode>();
    v.tri = sc->bvh.tri.as<float4>();
    v.shade = sc->bvh.shade.as<float4>();
    v.face = d->face_params ? sc->bvh.face.as<float4>() : nullptr;
    v.face_orig = d->face_params ? sc->face_orig.as<float4>() : nullptr;
    v.tri_vidx = sc->vidx.as<uint4>();
    v.normals = sc->normals.as<float4>();
    v.materials = sc->materials.as<float4>();
    v.analytic = sc->analytic.as<float4>();
    v.area = sc->area.as<float4>();
    v.tri_count = d->triangle_count;
    v.node_count = sc->bvh.node_count;
    v.material_count = d->material_count;
    v.analytic_count = d->analytic_light_count;
    v.area_count = d->area_light_count;
    v.area_weight = d->area_light_weight;
    const uint32_t tex_count = upload_textures(sc.get(), d->texture_word_count, d->texture_words, d->material_count,
                                               d->materials, d, s);
    set_environment(v, d->analytic_lights, d->analytic_light_count, tex_count);
    OL_HIP_CHECK(hipStreamSynchronize(s));
    sc->point_count = d->analytic_light_count;
    *out = sc.release();
  });
}

namespace {

// Lights, materials and environment of either scene kind.
void upload_shading(olhip_scene* sc, uint32_t material_count, const float* materials, uint32_t analytic_count,
                    const float* analytic_lights, uint32_t area_count, const float* area_lights, float area_weight,
                    hipStream_t s) {
  require(material_count > 0 && materials, "at least one material is required");
  require(analytic_count <= 64, "at most 64 analytic lights are supported");
  require(analytic_count == 0 || analytic_lights, "analytic lights missing");
  require(area_count == 0 || area_lights, "area lights missing");
  sc->materials_host.assign(materials, materials + 48ull * material_count);
  for (uint32_t m = 0; m < material_count; ++m)
    if (materials[48ull * m + 11] > 0.001f) sc->has_transmission = true;
  upload(sc->materials, reinterpret_cast<const float4*>(materials), 12ull * material_count, s);
  std::vector<float4> analytic(4ull * std::max<uint32_t>(analytic_count, 1), make_float4(0, 0, 0, 0));
  if (analytic_count) std::memcpy(analytic.data(), analytic_lights, sizeof(float) * 16ull * analytic_count);
  upload(sc->analytic, analytic.data(), analytic.size(), s);
  std::vector<float4> area(5ull * std::max<uint32_t>(area_count, 1), make_float4(0, 0, 0, 0));
  if (area_count) std::memcpy(area.data(), area_lights, sizeof(float) * 20ull * area_count);
  upload(sc->area, area.data(), area.size(), s);
  upload_area_guide(sc, area_lights, area_count, s);
  DScene& v = sc->view;
  v.materials = sc->materials.as<float4>();
  v.analytic = sc->analytic.as<float4>();
  v.area = sc->area.as<float4>();
  v.material_count = material_count;
  v.analytic_count = analytic_count;
  v.area_count = area_count;
  v.area_weight = area_weight;
  sc->point_count = analytic_count;
}

}  // namespace

namespace {

uint64_t instanced_hash(const olhip_instanced_desc* d, const std::vector<uint4>& vidx) {
  const uint32_t R = d->resource_count, I = d->instance_count;
  uint64_t h = fnv1a(vidx.data(), sizeof(uint4) * vidx.size(), kFnvBasis);
  h = fnv1a(d->vertex_offsets, 4ull * (R + 1), h);
  h = fnv1a(d->triangle_offsets, 4ull * (R + 1), h);
  h = fnv1a(&I, 4, h);
  h = fnv1a(d->instance_resource, 4ull * I, h);
  h = fnv1a(d->instance_triangle_offset, 4ull * I, h);
  return h;
}

// An instance whose box has at least this fraction of the scene box's
// surface area (and is not the only one) is entered after the rest of the
// scene: a ray reaches such a box at its entry into the scene, before any
// other instance has given it a best hit, and would otherwise traverse it
// along its whole length.
constexpr double kLaterArea = 0.25;
// Top levels price splitting off their largest box (sah_topology).
constexpr bool kIsolateLarge = true;

double box_area(const float* b) {
  const double dx = (double)b[3] - b[0], dy = (double)b[4] - b[1], dz = (double)b[5] - b[2];
  return dx * dy + dy * dz + dz * dx;
}

// Instance records and the top-level BVH over the instances' world bounds
// (a box is the bounds of the degenerate triangle (min, max, min)); sets the
// scene view.  The bottom levels must be built.
void build_top_level(olhip_scene* sc, const olhip_instanced_desc* d, hipStream_t s) {
  const uint32_t R = d->resource_count, I = d->instance_count;
  std::vector<InstanceDev> inst(I);
  std::vector<float4> tpos(2ull * I);
  std::vector<uint4> tvidx(I);
  uint64_t tri_total = 0;
  float scene_box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t k = 0; k < I; ++k)
    for (int a = 0; a < 3; ++a) {
      scene_box[a] = std::min(scene_box[a], d->world_bounds[6ull * k + a]);
      scene_box[3 + a] = std::max(scene_box[3 + a], d->world_bounds[6ull * k + 3 + a]);
    }
  const double later_area = kLaterArea * box_area(scene_box);
  for (uint32_t k = 0; k < I; ++k) {
    require(d->instance_resource[k] < R, "instance resource out of range");
    require(d->instance_material[k] < d->material_count, "instance material out of range");
    InstanceDev& x = inst[k];
    for (int row = 0; row < 3; ++row) {
      const float* w = d->world_to_object + 12ull * k + 4 * row;
      const float* n = d->normal_matrix + 9ull * k + 3 * row;
      x.w2o[row] = make_float4(w[0], w[1], w[2], w[3]);
      x.nrm[row] = make_float4(n[0], n[1], n[2], 0.0f);
    }
    x.nrm[0].w = d->determinant[k];
    x.blas = d->instance_resource[k];
    x.material = d->instance_material[k];
    x.tri_offset = d->instance_triangle_offset[k];
    x.root = 0u;
    const float* b = d->world_bounds + 6ull * k;
    for (int a = 0; a < 6; ++a) require(std::isfinite(b[a]), "instance bounds must be finite");
    x.sphere = make_float4(0.0f, 0.0f, 0.0f, INFINITY);
    if (d->world_spheres) {
      const float* sp = d->world_spheres + 4ull * k;
      require(std::isfinite(sp[0]) && std::isfinite(sp[1]) && std::isfinite(sp[2]) && sp[3] >= 0.0f,
              "instance spheres need a finite centre and a radius >= 0");
      x.sphere = make_float4(sp[0], sp[1], sp[2], sp[3]);
    }
    // The "later" flag is the sign bit of the sphere radius (traversal reads |r|).
    if (I > 1 && box_area(b) >= later_area) x.sphere.w = std::copysign(x.sphere.w, -1.0f);
    tpos[2 * k] = make_float4(b[0], b[1], b[2], 0.0f);
    tpos[2 * k + 1] = make_float4(b[3], b[4], b[5], 0.0f);
    tvidx[k] = make_uint4(2 * k, 2 * k + 1, 2 * k, 0u);
    tri_total += d->triangle_offsets[x.blas + 1] - d->triangle_offsets[x.blas];
  }
  upload(sc->tlas_pos, tpos.data(), tpos.size(), s);
  upload(sc->tlas_vidx, tvidx.data(), tvidx.size(), s);
  OL_HIP_CHECK(hipStreamSynchronize(s));
  BvhInput tin{sc->tlas_pos.as<float4>(), sc->tlas_vidx.as<uint4>(), I, 1u, d->world_bounds, kIsolateLarge};
  bvh_build(tin, sc->bvh, s, nullptr);
  // The instance table in top-level leaf order: a top-level leaf's first
  // slot is then its instance's index, and entering an instance needs no
  // read of the leaf record.
  {
    std::vector<float4> leaves(3ull * I);
    OL_HIP_CHECK(hipMemcpy(leaves.data(), sc->bvh.tri.ptr, sizeof(float4) * leaves.size(), hipMemcpyDeviceToHost));
    std::vector<InstanceDev> ordered(I);
    for (uint32_t j = 0; j < I; ++j) {
      uint32_t k;
      std::memcpy(&k, &leaves[3ull * j].w, sizeof(k));
      require(k < I, "top-level leaf names no instance");
      ordered[j] = inst[k];
    }
    upload(sc->inst_table, ordered.data(), ordered.size(), s);
    OL_HIP_CHECK(hipStreamSynchronize(s));
  }
  sc->tlas_depth = sc->bvh.depth;
  // Stack sizing: both levels plus the sentinel.
  sc->bvh.depth = sc->tlas_depth + sc->blas_depth + 1u;
  sc->instance_count = I;
  sc->tri_count = (uint32_t)std::min<uint64_t>(tri_total, 0xFFFFFFFFull);
  DScene& v = sc->view;
  v.nodes = sc->bvh.nodes.as<WideNode>();
  v.tri = sc->bvh.tri.as<float4>();
  v.tri_vidx = nullptr;
  v.normals = nullptr;
  v.tri_count = sc->tri_count;
  v.node_count = sc->bvh.node_count;
  v.instanced = 1u;
  v.instance_count = I;
  v.inst = sc->inst_table.as<InstanceDev>();
  v.blas = sc->blas_table.as<BlasDev>();
}

}  // namespace

int olhip_scene_create_instanced(olhip_ctx* ctx, const olhip_instanced_desc* d, olhip_scene** out) {
  return guarded([&] {
    require(ctx && d && out, "ctx, desc and out must not be NULL");
    require(d->resource_count >= 1 && d->vertex_offsets && d->triangle_offsets, "resources missing");
    require(d->instance_count >= 1 && d->instance_resource && d->instance_material && d->instance_triangle_offset &&
                d->instance_id && d->world_to_object && d->normal_matrix && d->determinant && d->world_bounds,
            "instance arrays missing");
    require(d->leaf_size <= 16, "leaf_size must be 0..16");
    require(d->builder <= OLHIP_BUILDER_SAH, "unknown BVH builder");
    const uint32_t R = d->resource_count, V = d->vertex_offsets[R], T = d->triangle_offsets[R];
    require(V == 0 || (d->positions && d->normals), "geometry arrays missing");
    require(T == 0 || d->indices, "indices missing");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    auto sc = std::make_unique<olhip_scene>();
    sc->ctx = ctx;
    sc->vertex_count = V;
    std::vector<float4> pos(V), nrm(V);
    for (uint32_t i = 0; i < V; ++i) {
      pos[i] = make_float4(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2], 0.0f);
      nrm[i] = make_float4(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2], 0.0f);
      require(std::isfinite(pos[i].x) && std::isfinite(pos[i].y) && std::isfinite(pos[i].z),
              "vertex positions must be finite");
    }
    std::vector<uint4> vidx(T);
    for (uint32_t r = 0; r < R; ++r) {
      require(d->vertex_offsets[r] <= d->vertex_offsets[r + 1] && d->triangle_offsets[r] <= d->triangle_offsets[r + 1],
              "resource offsets must be non-decreasing");
      const uint32_t nv = d->vertex_offsets[r + 1] - d->vertex_offsets[r];
      for (uint32_t t = d->triangle_offsets[r]; t < d->triangle_offsets[r + 1]; ++t) {
        const uint32_t a = d->indices[3 * t], b = d->indices[3 * t + 1], c = d->indices[3 * t + 2];
        require(a < nv && b < nv && c < nv, "triangle index out of range of its resource");
        vidx[t] = make_uint4(a, b, c, 0u);
      }
    }
    upload(sc->positions, pos.data(), pos.size(), s);
    upload(sc->normals, nrm.data(), nrm.size(), s);
    upload(sc->vidx, vidx.data(), vidx.size(), s);
    upload_shading(sc.get(), d->material_count, d->materials, d->analytic_light_count, d->analytic_lights,
                   d->area_light_count, d->area_lights, d->area_light_weight, s);
    set_environment(sc->view, d->analytic_lights, d->analytic_light_count,
                    upload_textures(sc.get(), d->texture_word_count, d->texture_words, d->material_count,
                                    d->materials, nullptr, s));
    sc->leaf_size = d->leaf_size;
    sc->vertex_offsets.assign(d->vertex_offsets, d->vertex_offsets + R + 1);
    sc->triangle_offsets.assign(d->triangle_offsets, d->triangle_offsets + R + 1);
    sc->structure_hash = instanced_hash(d, vidx);
    OL_HIP_CHECK(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    // Bottom level: one BVH per resource, object space.
    std::vector<BlasDev> blas(R);
    uint32_t blas_depth = 0;
    for (uint32_t r = 0; r < R; ++r) {
      auto out_r = std::make_unique<BvhOutput>();
      const uint32_t vo = d->vertex_offsets[r], to = d->triangle_offsets[r];
      const uint32_t nt = d->triangle_offsets[r + 1] - to;
      if (nt) {
        std::vector<float> boxes;
        if (d->builder == OLHIP_BUILDER_SAH && nt > 1) {
          boxes.resize(6ull * nt);
          for (uint32_t t = 0; t < nt; ++t) {
            const uint4 tv = vidx[to + t];
            const float4 a = pos[vo + tv.x], b = pos[vo + tv.y], c = pos[vo + tv.z];
            float* o = &boxes[6ull * t];
            o[0] = std::min(a.x, std::min(b.x, c.x));
            o[1] = std::min(a.y, std::min(b.y, c.y));
            o[2] = std::min(a.z, std::min(b.z, c.z));
            o[3] = std::max(a.x, std::max(b.x, c.x));
            o[4] = std::max(a.y, std::max(b.y, c.y));
            o[5] = std::max(a.z, std::max(b.z, c.z));
          }
        }
        BvhInput in{sc->positions.as<float4>() + vo, sc->vidx.as<uint4>() + to, nt, d->leaf_size,
                    boxes.empty() ? nullptr : boxes.data()};
        bvh_build(in, *out_r, s, nullptr, true);
        build_shade_records(*out_r, sc->vidx.as<uint4>() + to, sc->normals.as<float4>() + vo, nt, false, s);
      }
      blas[r] = BlasDev{out_r->nodes.as<WideNode>(), out_r->tri.as<float4>(), sc->vidx.as<uint4>() + to,
                        sc->normals.as<float4>() + vo, out_r->shade.as<float4>(), out_r->node_count, nt, 0u, 0u};
      blas_depth = std::max(blas_depth, out_r->depth);
      sc->blas.push_back(std::move(out_r));
    }
    upload(sc->blas_table, blas.data(), blas.size(), s);
    sc->blas_depth = blas_depth;
    build_top_level(sc.get(), d, s);
    sc->bvh.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = sc.release();
  });
}

int olhip_scene_update(olhip_ctx* ctx, olhip_scene* sc, const olhip_scene_desc* d) {
  return guarded([&] {
    require(ctx && sc && d, "ctx, scene and desc must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    require(!sc->view.instanced, "two-level scenes are updated with olhip_scene_update_instanced");
    require(d->vertex_count == sc->vertex_count && d->triangle_count == sc->tri_count,
            "an update keeps the vertex and triangle counts (create a new scene instead)");
    require(d->triangle_count == 0 || (d->indices && d->material_ids), "indices and material ids are required");
    require(d->material_count > 0 && d->materials, "at least one material is required");
    std::vector<uint4> vidx(d->triangle_count);
    for (uint32_t t = 0; t < d->triangle_count; ++t) {
      require(d->material_ids[t] < d->material_count, "material id out of range");
      vidx[t] = make_uint4(d->indices[3 * t], d->indices[3 * t + 1], d->indices[3 * t + 2], d->material_ids[t]);
    }
    require(fnv1a(d->indices, 12ull * d->triangle_count, kFnvBasis) == sc->structure_hash,
            "an update keeps every triangle's vertex indices (create a new scene instead)");
    std::vector<float4> pos, nrm;
    if (d->positions) {
      require(d->vertex_count == 0 || d->normals, "normals missing");
      pos.resize(d->vertex_count);
      nrm.resize(d->vertex_count);
      for (uint32_t i = 0; i < d->vertex_count; ++i) {
        pos[i] = make_float4(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2], 0.0f);
        nrm[i] = make_float4(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2], 0.0f);
        require(std::isfinite(pos[i].x) && std::isfinite(pos[i].y) && std::isfinite(pos[i].z),
                "vertex positions must be finite");
      }
    }
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // Work already queued on the old contents finishes first.
    OL_HIP_CHECK(hipStreamSynchronize(s));
    upload(sc->vidx, vidx.data(), vidx.size(), s);  // material ids may change
    if (d->positions) {
      upload(sc->positions, pos.data(), pos.size(), s);
      upload(sc->normals, nrm.data(), nrm.size(), s);
      OL_HIP_CHECK(hipStreamSynchronize(s));
      BvhInput in{sc->positions.as<float4>(), sc->vidx.as<uint4>(), d->triangle_count, sc->leaf_size};
      bvh_refit(in, sc->bvh, s);
      DScene& v = sc->view;
      v.nodes = sc->bvh.nodes.as<WideNode>();
      v.tri = sc->bvh.tri.as<float4>();
      v.normals = sc->normals.as<float4>();
      v.node_count = sc->bvh.node_count;
    }
    // Normals, material ids or per-triangle parameters may have changed.
    upload_faces(sc, d, s);
    build_shade_records(sc->bvh, sc->vidx.as<uint4>(), sc->normals.as<float4>(), d->triangle_count, true, s,
                        d->face_params ? sc->face_orig.as<float4>() : nullptr,
                        d->face_params ? sc->face_metal.as<float>() : nullptr);
    sc->view.shade = sc->bvh.shade.as<float4>();
    sc->view.face = d->face_params ? sc->bvh.face.as<float4>() : nullptr;
    sc->view.face_orig = d->face_params ? sc->face_orig.as<float4>() : nullptr;
    sc->has_transmission = false;
    upload_shading(sc, d->material_count, d->materials, d->analytic_light_count, d->analytic_lights,
                   d->area_light_count, d->area_lights, d->area_light_weight, s);
    set_environment(sc->view, d->analytic_lights, d->analytic_light_count,
                    upload_textures(sc, d->texture_word_count, d->texture_words, d->material_count, d->materials,
                                    d, s));
    sc->point_count = d->analytic_light_count;
    OL_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int olhip_scene_update_instanced(olhip_ctx* ctx, olhip_scene* sc, const olhip_instanced_desc* d) {
  return guarded([&] {
    require(ctx && sc && d, "ctx, scene and desc must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    require(sc->view.instanced, "flat scenes are updated with olhip_scene_update");
    const uint32_t R = d->resource_count;
    require(R == sc->blas.size() && d->vertex_offsets && d->triangle_offsets,
            "an update keeps the mesh resources (create a new scene instead)");
    require(d->instance_count == sc->instance_count && d->instance_resource && d->instance_material &&
                d->instance_triangle_offset && d->instance_id && d->world_to_object && d->normal_matrix &&
                d->determinant && d->world_bounds,
            "an update keeps the instance count (create a new scene instead)");
    for (uint32_t r = 0; r <= R; ++r)
      require(d->vertex_offsets[r] == sc->vertex_offsets[r] && d->triangle_offsets[r] == sc->triangle_offsets[r],
              "an update keeps the resources' vertex and triangle counts");
    const uint32_t V = d->vertex_offsets[R], T = d->triangle_offsets[R];
    require(T == 0 || d->indices, "indices missing");
    std::vector<uint4> vidx(T);
    for (uint32_t t = 0; t < T; ++t) vidx[t] = make_uint4(d->indices[3 * t], d->indices[3 * t + 1], d->indices[3 * t + 2], 0u);
    require(instanced_hash(d, vidx) == sc->structure_hash,
            "an update keeps the triangles and each instance's resource (create a new scene instead)");
    std::vector<float4> pos, nrm;
    if (d->positions) {
      require(V == 0 || d->normals, "normals missing");
      pos.resize(V);
      nrm.resize(V);
      for (uint32_t i = 0; i < V; ++i) {
        pos[i] = make_float4(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2], 0.0f);
        nrm[i] = make_float4(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2], 0.0f);
        require(std::isfinite(pos[i].x) && std::isfinite(pos[i].y) && std::isfinite(pos[i].z),
                "vertex positions must be finite");
      }
    }
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    OL_HIP_CHECK(hipStreamSynchronize(s));
    if (d->positions) {
      upload(sc->positions, pos.data(), pos.size(), s);
      upload(sc->normals, nrm.data(), nrm.size(), s);
      OL_HIP_CHECK(hipStreamSynchronize(s));
      std::vector<BlasDev> blas(R);
      uint32_t blas_depth = 0;
      for (uint32_t r = 0; r < R; ++r) {
        const uint32_t vo = d->vertex_offsets[r], to = d->triangle_offsets[r];
        const uint32_t nt = d->triangle_offsets[r + 1] - to;
        BvhOutput& b = *sc->blas[r];
        if (nt) {
          BvhInput in{sc->positions.as<float4>() + vo, sc->vidx.as<uint4>() + to, nt, sc->leaf_size};
          bvh_refit(in, b, s);
          build_shade_records(b, sc->vidx.as<uint4>() + to, sc->normals.as<float4>() + vo, nt, false, s);
        }
        blas[r] = BlasDev{b.nodes.as<WideNode>(), b.tri.as<float4>(), sc->vidx.as<uint4>() + to,
                          sc->normals.as<float4>() + vo, b.shade.as<float4>(), b.node_count, nt, 0u, 0u};
        blas_depth = std::max(blas_depth, b.depth);
      }
      upload(sc->blas_table, blas.data(), blas.size(), s);
      sc->blas_depth = blas_depth;
    }
    sc->has_transmission = false;
    upload_shading(sc, d->material_count, d->materials, d->analytic_light_count, d->analytic_lights,
                   d->area_light_count, d->area_lights, d->area_light_weight, s);
    set_environment(sc->view, d->analytic_lights, d->analytic_light_count,
                    upload_textures(sc, d->texture_word_count, d->texture_words, d->material_count, d->materials,
                                    nullptr, s));
    sc->point_count = d->analytic_light_count;
    build_top_level(sc, d, s);
  });
}

int olhip_sah_topology(const float* boxes, uint32_t n, uint32_t* order, uint32_t* left, uint32_t* right,
                       uint32_t* first, uint32_t* last) {
  return guarded([&] {
    require(boxes && order && left && right && first && last, "arguments must not be NULL");
    require(n >= 2, "at least two primitives are required");
    std::vector<uint32_t> o, l, r, f, la, pi, pl;
    sah_topology(boxes, n, o, l, r, f, la, pi, pl, kIsolateLarge);
    std::copy(o.begin(), o.end(), order);
    std::copy(l.begin(), l.end(), left);
    std::copy(r.begin(), r.end(), right);
    std::copy(f.begin(), f.end(), first);
    std::copy(la.begin(), la.end(), last);
  });
}

int olhip_scene_destroy(olhip_scene* scene) {
  return guarded([&] {
    if (!scene) return;
    (void)hipSetDevice(scene->ctx->device);
    (void)hipStreamSynchronize(scene->ctx->stream);
    delete scene;
  });
}

int olhip_scene_info_get(const olhip_scene* scene, olhip_scene_info* out) {
  return guarded([&] {
    require(scene && out, "scene and out must not be NULL");
    out->triangle_count = scene->tri_count;
    out->node_count = scene->bvh.node_count;
    out->bvh_depth = scene->bvh.depth;
    out->stack_depth = (uint32_t)std::max(0, stack_depth_for(scene->bvh.depth));
    out->build_ms = scene->bvh.build_ms;
    out->device_bytes = scene->positions.bytes + scene->normals.bytes + scene->vidx.bytes + scene->materials.bytes +
                        scene->analytic.bytes + scene->area.bytes + scene->bvh.nodes.bytes + scene->bvh.tri.bytes;
  });
}

int olhip_scene_bvh(const olhip_scene* scene, void* nodes, void* tri) {
  return guarded([&] {
    require(scene != nullptr, "scene must not be NULL");
    OL_HIP_CHECK(hipSetDevice(scene->ctx->device));
    OL_HIP_CHECK(hipStreamSynchronize(scene->ctx->stream));
    if (nodes && scene->bvh.node_count)
      OL_HIP_CHECK(hipMemcpy(nodes, scene->bvh.nodes.ptr, 64ull * scene->bvh.node_count, hipMemcpyDeviceToHost));
    if (tri && scene->tri_count)
      OL_HIP_CHECK(hipMemcpy(tri, scene->bvh.tri.ptr, 48ull * scene->tri_count, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"

namespace {

// Fixed-size stack of the reference-estimator traversal (LDS only).
int choose_stack(const olhip_ctx* ctx, const olhip_scene* sc) {
  (void)ctx;
  int need = stack_depth_for(std::max<uint32_t>(sc->bvh.depth, 1));
  if (need < 0) throw OlError{OLHIP_ERR_RUNTIME, "BVH too deep for the traversal stack"};
  return need;
}

// Two-level traversal stack: LDS part from the context, spill for the rest.
GiLaunch make_launch(olhip_ctx* ctx, const olhip_scene* sc) {
  GiLaunch l{};
  l.cu_count = ctx->cu_count;
  l.stack_depth = (int)(3u * std::max<uint32_t>(sc->bvh.depth, 1) + 1u);
  l.lds_stack = ctx->lds_stack;
  l.trace_block = ctx->trace_block;
  uint64_t words = trace_spill_words(l);
  if (words) ctx->spill.reserve(4 * words);
  l.spill = words ? ctx->spill.as<uint32_t>() : nullptr;
  l.spill_words = words;
  return l;
}

}  // namespace

extern "C" {

int olhip_render_gi(olhip_ctx* ctx, const olhip_scene* sc, const olhip_camera* cam, uint32_t width, uint32_t height,
                    uint32_t samples, uint32_t frame_index, const olhip_gi_settings* set, float* out_rgba,
                    olhip_stats* stats_out) {
  return olhip_render_products(ctx, sc, cam, width, height, samples, frame_index, set, out_rgba, nullptr, nullptr,
                               nullptr, nullptr, stats_out);
}

int olhip_render_products(olhip_ctx* ctx, const olhip_scene* sc, const olhip_camera* cam, uint32_t width,
                          uint32_t height, uint32_t samples, uint32_t frame_index, const olhip_gi_settings* set,
                          float* out_rgba, float* depth, float* normal, uint32_t* primitive, float* variance,
                          olhip_stats* stats_out) {
  return guarded([&] {
    require(ctx && sc && cam && set, "ctx, scene, camera and settings must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    require(width >= 1 && height >= 1, "width and height must be positive");
    require((uint64_t)width * height <= (1ull << 28), "image too large");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    // A one-job engine sized for this frame (grown when a larger frame comes).
    const uint64_t np = (uint64_t)width * height;
    const uint32_t per_hit = std::max<uint32_t>(2, job_entries_per_hit(sc, *set));
    // Shadow entries are bounded to 2^28 (12 GB): many NEE samples per hit
    // shrink the sample chunk instead.
    const uint64_t capacity = std::min<uint64_t>(set->path_capacity ? set->path_capacity : (1ull << 25),
                                                 (1ull << 28) / per_hit);
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(samples, capacity / np));
    // A chunk's queue segments round its span up by at most two 64-path
    // chunks per segment.  ReSTIR passes are resident together, one span
    // each; other jobs keep one chunk of `chunk` samples resident.
    const uint64_t pad = 2ull * kChunk * kJobStripes;
    const uint64_t need = std::max<uint64_t>(4096, set->restir_candidates ? chunk * (np + pad) : chunk * np + pad);
    if (!ctx->frame_engine || ctx->frame_engine_slots < need || ctx->frame_engine_entries < per_hit) {
      ctx->frame_engine.reset();
      olhip_engine_config cfg{};
      cfg.path_slots = need;
      cfg.entry_factor = per_hit;
      cfg.use_graph = 1;
      cfg.fused_primary = (uint32_t)ctx->fused_primary;
      ctx->frame_engine = std::make_unique<Engine>(ctx, cfg);
      ctx->frame_engine_slots = need;
      ctx->frame_engine_entries = per_hit;
    }
    Engine& eng = *ctx->frame_engine;
    eng.set_profiling(ctx->profiling != 0);
    eng.set_work_counters(ctx->counting != 0);
    eng.trace_ms = eng.shade_ms = 0.0;
    eng.trace_launches = 0;
    olhip_job_desc d{};
    d.scene = sc;
    d.camera = *cam;
    d.width = width;
    d.height = height;
    d.samples = samples;
    d.frame_index = frame_index;
    d.settings = *set;
    d.settings.path_capacity = (uint32_t)std::min<uint64_t>(capacity, 0xFFFFFFFFull);
    d.products = ((depth || normal || primitive) ? OLHIP_PRODUCT_PRIMARY : 0u) | (variance ? OLHIP_PRODUCT_VARIANCE : 0u);
    unsigned long long work0[kWorkWords] = {};
    if (ctx->counting) eng.read_work(work0);
    olhip_engine_stats es0{};
    eng.engine_stats(&es0);
    Event e_start, e_end;
    OL_HIP_CHECK(hipEventRecord(e_start.e, ctx->stream));
    const uint64_t id = eng.submit(d);
    eng.run_until_idle();
    OL_HIP_CHECK(hipEventRecord(e_end.e, ctx->stream));
    OL_HIP_CHECK(hipEventSynchronize(e_end.e));
    olhip_job_stats js{};
    if (eng.state(id) != OLHIP_JOB_DONE) {
      eng.take_result(id, nullptr, nullptr);  // throws with the failure reason
      throw OlError{OLHIP_ERR_RUNTIME, "render did not complete"};
    }
    // The frame stays in HBM for olhip_frame_output / olhip_frame_download:
    // the job's image buffer becomes the context's (no copy).
    eng.swap_image(id, ctx->image);
    ctx->image_pixels = (uint32_t)np;
    ctx->image_width = width;
    ctx->image_height = height;
    if (d.products) eng.products(id, depth, normal, primitive, variance);
    if (out_rgba) OL_HIP_CHECK(hipMemcpy(out_rgba, ctx->image.ptr, sizeof(float4) * np, hipMemcpyDeviceToHost));
    eng.take_result(id, nullptr, &js);
    if (stats_out) {
      olhip_stats& o = *stats_out;
      std::memset(&o, 0, sizeof(o));
      o.primary_rays = js.primary_rays;
      o.extension_rays = js.extension_rays;
      o.shadow_rays = js.shadow_rays;
      olhip_engine_stats es{};
      eng.engine_stats(&es);
      o.queue_overflow = es.queue_overflow - es0.queue_overflow;
      o.total_ms = elapsed(e_start, e_end);
      o.trace_ms = eng.trace_ms;
      o.shade_ms = eng.shade_ms;
      o.shadow_ms = 0.0;
      o.trace_launches = eng.trace_launches;
      o.shade_launches = eng.trace_launches;
      o.shadow_launches = 0;
      o.chunks = js.chunks;
      if (ctx->counting) {
        unsigned long long w[kWorkWords];
        eng.read_work(w);
        o.closest_nodes = w[0] - work0[0];
        o.closest_tris = w[1] - work0[1];
        o.shadow_nodes = w[2] - work0[2];
        o.shadow_tris = w[3] - work0[3];
        o.closest_instances = w[4] - work0[4];
        o.shadow_instances = w[5] - work0[5];
        o.shaded_hits = w[6] - work0[6];
        o.area_samples = w[7] - work0[7];
      }
    }
  });
}

// ---- batched engine ------------------------------------------------------------------------------
int olhip_engine_create(olhip_ctx* ctx, const olhip_engine_config* config, olhip_engine** out) {
  return guarded([&] {
    require(ctx && out, "ctx and out must not be NULL");
    olhip_engine_config cfg{};
    if (config) cfg = *config;
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    auto e = std::make_unique<olhip_engine>();
    e->engine = std::make_unique<Engine>(ctx, cfg);
    *out = e.release();
  });
}

int olhip_engine_destroy(olhip_engine* engine) {
  return guarded([&] { delete engine; });
}

int olhip_engine_submit(olhip_engine* engine, const olhip_job_desc* job, uint64_t* job_id) {
  return guarded([&] {
    require(engine && job && job_id, "arguments must not be NULL");
    *job_id = engine->engine->submit(*job);
  });
}

int olhip_engine_step(olhip_engine* engine, uint32_t* pending) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    const uint32_t p = engine->engine->step();
    if (pending) *pending = p;
  });
}

int olhip_engine_run(olhip_engine* engine) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    engine->engine->run_until_idle();
  });
}

int olhip_engine_poll(olhip_engine* engine, uint64_t job_id, int* state) {
  return guarded([&] {
    require(engine && state, "arguments must not be NULL");
    *state = engine->engine->state(job_id);
  });
}

int olhip_engine_wait(olhip_engine* engine, uint64_t job_id, double timeout_s, int* state) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    const int st = engine->engine->wait(job_id, timeout_s);
    if (state) *state = st;
  });
}

int olhip_engine_result(olhip_engine* engine, uint64_t job_id, float* out_rgba, olhip_job_stats* stats) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    engine->engine->take_result(job_id, out_rgba, stats);
  });
}

int olhip_engine_products(olhip_engine* engine, uint64_t job_id, float* depth, float* normal, uint32_t* primitive,
                          float* variance) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    engine->engine->products(job_id, depth, normal, primitive, variance);
  });
}

int olhip_engine_start(olhip_engine* engine) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    engine->engine->start_worker();
  });
}

int olhip_engine_stop(olhip_engine* engine) {
  return guarded([&] {
    require(engine != nullptr, "engine must not be NULL");
    engine->engine->stop_worker();
  });
}

int olhip_engine_get_stats(olhip_engine* engine, olhip_engine_stats* out) {
  return guarded([&] {
    require(engine && out, "arguments must not be NULL");
    engine->engine->engine_stats(out);
  });
}

int olhip_bandwidth_probe(olhip_ctx* ctx, uint64_t bytes, uint32_t repeats, double* ms_per_copy) {
  return guarded([&] {
    require(ctx && ms_per_copy, "arguments must not be NULL");
    require(bytes >= 16 && bytes <= (1ull << 36) && repeats >= 1, "bytes must be in [16, 2^36], repeats >= 1");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint64_t n16 = bytes / 16;
    DevBuf a, b;
    a.reserve(16 * n16);
    b.reserve(16 * n16);
    OL_HIP_CHECK(hipMemsetAsync(a.ptr, 0, 16 * n16, s));
    stream_copy(a.as<float4>(), b.as<float4>(), n16, ctx->cu_count, s);
    Event e0, e1;
    OL_HIP_CHECK(hipEventRecord(e0.e, s));
    for (uint32_t r = 0; r < repeats; ++r) stream_copy(a.as<float4>(), b.as<float4>(), n16, ctx->cu_count, s);
    OL_HIP_CHECK(hipEventRecord(e1.e, s));
    OL_HIP_CHECK(hipEventSynchronize(e1.e));
    *ms_per_copy = elapsed(e0, e1) / repeats;
  });
}

int olhip_frame_download(olhip_ctx* ctx, float* out_rgba) {
  return guarded([&] {
    require(ctx && out_rgba, "ctx and out must not be NULL");
    require(ctx->image_pixels > 0, "no frame has been rendered");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    OL_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    OL_HIP_CHECK(hipMemcpy(out_rgba, ctx->image.ptr, sizeof(float4) * ctx->image_pixels, hipMemcpyDeviceToHost));
  });
}

int olhip_frame_output(olhip_ctx* ctx, uint32_t format, float exposure, uint32_t width, uint32_t height, void* out,
                       uint32_t out_on_device) {
  return guarded([&] {
    require(ctx && out, "ctx and out must not be NULL");
    require(ctx->image_pixels > 0, "no frame has been rendered");
    require(format <= OLHIP_OUTPUT_P010, "unknown output format");
    require(std::isfinite(exposure) && exposure > 0.0f, "exposure must be positive");
    require(width >= 1 && height >= 1, "width and height must be positive");
    if (format != OLHIP_OUTPUT_P010)
      require(width == ctx->image_width && height == ctx->image_height, "output size must equal the frame size");
    if (format != OLHIP_OUTPUT_RGBA8)
      require(width % 4 == 0 && height % 2 == 0, "NV12/P010 output requires width divisible by 4 and even height");
    require((uint64_t)width * height <= (1ull << 28), "output too large");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t bytes = output_bytes(format, width, height);
    DevBuf staging;
    uint32_t* dst = static_cast<uint32_t*>(out);
    if (!out_on_device) {
      staging.reserve(bytes);
      dst = staging.as<uint32_t>();
    }
    launch_output(format, ctx->image.as<float4>(), ctx->image_width, ctx->image_height, width, height, exposure, dst,
                  s);
    OL_HIP_CHECK(hipGetLastError());
    if (!out_on_device) OL_HIP_CHECK(hipMemcpyAsync(out, staging.ptr, bytes, hipMemcpyDeviceToHost, s));
    OL_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int olhip_camera_rays(olhip_ctx* ctx, const olhip_camera* cam, uint32_t width, uint32_t height, uint32_t frame_index,
                      uint32_t sample, uint32_t x0, uint32_t y0, uint32_t tile_width, uint32_t tile_height,
                      float* origins, float* directions) {
  return guarded([&] {
    require(ctx && cam && origins && directions, "arguments must not be NULL");
    require(width >= 1 && height >= 1, "width and height must be positive");
    require(cam->projection <= 2, "unknown projection");
    require(tile_width >= 1 && tile_height >= 1, "tile must not be empty");
    require((uint64_t)x0 + tile_width <= width && (uint64_t)y0 + tile_height <= height, "tile outside the image");
    require((uint64_t)tile_width * tile_height <= (1ull << 26), "tile too large");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    GiParams P;
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < 3; ++k) {
      P.cam_o[k] = cam->position[k];
      P.cam_f[k] = cam->forward[k];
      P.cam_r[k] = cam->right[k];
      P.cam_u[k] = cam->up[k];
    }
    P.projection = cam->projection;
    P.width = width;
    P.height = height;
    P.frame = (uint32_t)((float)frame_index + 0.5f);
    const uint32_t n = tile_width * tile_height;
    DevBuf o, d;
    o.reserve(sizeof(float4) * n);
    d.reserve(sizeof(float4) * n);
    camera_rays(P, x0, y0, tile_width, tile_height, sample, o.as<float4>(), d.as<float4>(), s);
    OL_HIP_CHECK(hipGetLastError());
    OL_HIP_CHECK(hipMemcpyAsync(origins, o.ptr, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
    OL_HIP_CHECK(hipMemcpyAsync(directions, d.ptr, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
    OL_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int olhip_trace(olhip_ctx* ctx, const olhip_scene* sc, const float* origins, const float* directions, uint32_t n,
                float* hits) {
  return guarded([&] {
    require(ctx && sc && origins && directions && hits, "arguments must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    if (n == 0) return;
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevBuf o, d, h, ov;
    upload(o, reinterpret_cast<const float4*>(origins), n, s);
    upload(d, reinterpret_cast<const float4*>(directions), n, s);
    h.reserve(sizeof(float4) * n);
    ov.reserve(4);
    OL_HIP_CHECK(hipMemsetAsync(ov.ptr, 0, 4, s));
    GiLaunch l = make_launch(ctx, sc);
    trace_rays(l, sc->view, o.as<float4>(), d.as<float4>(), n, h.as<float4>(), ov.as<uint32_t>(), nullptr, s);
    OL_HIP_CHECK(hipGetLastError());
    OL_HIP_CHECK(hipStreamSynchronize(s));
    uint32_t overflow = 0;
    OL_HIP_CHECK(hipMemcpy(&overflow, ov.ptr, 4, hipMemcpyDeviceToHost));
    if (overflow) throw OlError{OLHIP_ERR_RUNTIME, "traversal stack overflow"};
    OL_HIP_CHECK(hipMemcpy(hits, h.ptr, sizeof(float4) * n, hipMemcpyDeviceToHost));
  });
}

int olhip_trace_device(olhip_ctx* ctx, const olhip_scene* sc, const float* d_origins, const float* d_directions,
                       uint32_t n, float* d_hits, uint32_t* d_status, void* stream) {
  return guarded([&] {
    require(ctx && sc && d_origins && d_directions && d_hits && d_status, "arguments must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    require(ctx->trace_scratch != nullptr, "context has no trace scratch array");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    // The launch shape of make_launch, with the spill columns in the fixed
    // array; the grid is capped so that its lanes' columns fit there.
    GiLaunch l{};
    l.cu_count = ctx->cu_count;
    l.stack_depth = (int)(3u * std::max<uint32_t>(sc->bvh.depth, 1) + 1u);
    l.lds_stack = ctx->lds_stack;
    l.trace_block = ctx->trace_block;
    l.spill = ctx->trace_scratch;
    l.spill_words = OLHIP_TRACE_SCRATCH_BYTES / 4u;
    const uint64_t per_block =
        (uint64_t)std::max(l.stack_depth - l.lds_stack, 0) * (uint64_t)std::max(l.trace_block, 1);
    require(per_block <= l.spill_words,
            "hierarchy too deep for the fixed trace scratch array (" + std::to_string(4 * per_block) +
                " bytes of stack spill per workgroup, at most " + std::to_string(OLHIP_TRACE_SCRATCH_BYTES) + ")");
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(1), 0, s, d_status);
    OL_HIP_CHECK(hipGetLastError());
    if (n == 0) return;
    trace_rays(l, sc->view, reinterpret_cast<const float4*>(d_origins), reinterpret_cast<const float4*>(d_directions),
               n, reinterpret_cast<float4*>(d_hits), d_status, nullptr, s);
    OL_HIP_CHECK(hipGetLastError());
  });
}

// ---- reference-estimator mode -------------------------------------------------------------------
int olhip_e0_begin(olhip_ctx* ctx, const olhip_scene* sc, const olhip_e0_camera* cam, uint32_t width,
                   uint32_t height, const olhip_e0_env* env, uint32_t brute_force) {
  return guarded([&] {
    require(ctx && sc && cam && env, "arguments must not be NULL");
    require(sc->ctx == ctx, "scene belongs to another context");
    require(width >= 1 && height >= 1, "width and height must be positive");
    require((uint64_t)width * height <= (1ull << 27), "image too large");
    require(cam->projection <= 2, "unknown projection");
    require(env->mode <= 2, "unknown environment mode");
    require(env->mode != 2 || (env->image && env->width && env->height), "environment image missing");
    require(!sc->view.instanced, "the reference estimator takes flat scenes");
    OL_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    auto& f = ctx->e0;
    const uint32_t n = width * height;
    const size_t ot = cam->projection == 1 ? 8 : 4, dt = cam->projection == 1 ? 4 : 8;
    f.o.reserve(3 * ot * n);
    f.d.reserve(3 * dt * n);
    f.T.reserve(3 * ot * n);
    f.rad.reserve(3 * ot * n);
    f.acc.reserve(12ull * n);
    f.active.reserve(4ull * n);
    f.flags.reserve(4ull * n);
    f.scan.reserve(4ull * (n + 1));
    f.ht.reserve(4ull * n);
    f.htri.reserve(4ull * n);
    f.hitlist.reserve(4ull * n);
    f.hit_t.reserve(4ull * n);
    f.hit_tri.reserve(4ull * n);
    f.jx.reserve(4ull * n);
    f.jy.reserve(4ull * n);
    f.gauss.reserve(12ull * n);
    f.ctr.reserve(16);
    f.out.reserve(16ull * n);
    OL_HIP_CHECK(hipMemsetAsync(f.acc.ptr, 0, 12ull * n, s));
    OL_HIP_CHECK(hipMemsetAsync(f.ctr.ptr, 0, 16, s));
    E0Params& p = f.params;
    p = E0Params{};
    p.sc = sc->view;
    p.positions = sc->positions.as<float4>();
    for (int k = 0; k < 3; ++k) {
      p.pos[k] = cam->position[k];
      p.fwd[k] = cam->forward[k];
      p.right[k] = cam->right[k];
      p.up[k] = cam->up[k];
    }
    p.k_x = cam->k_x;
    p.k_y = cam->k_y;
    p.projection = cam->projection;
    p.width = width;
    p.height = height;
    p.brute_force = (brute_force || sc->bvh.node_count == 0) ? 1u : 0u;
    p.env_mode = env->mode;
    for (int k = 0; k < 3; ++k) p.env_color[k] = env->color[k];
    p.env_intensity = env->intensity;
    p.env_rotation = env->rotation;
    if (env->mode == 2) {
      upload(f.env, env->image, 3ull * env->width * env->height, s);
      p.env_image = f.env.as<float>();
      p.env_w = env->width;
      p.env_h = env->height;
    }
    E0State& st = f.state;
    st.o = f.o.ptr;
    st.d = f.d.ptr;
    st.T = f.T.ptr;
    st.rad = f.rad.ptr;
    st.acc = f.acc.as<float>();
    st.active = f.active.as<uint32_t>();
    st.flags = f.flags.as<uint32_t>();
    st.scan = f.scan.as<uint32_t>();
    st.ht = f.ht.as<float>();
    st.htri = f.htri.as<int>();
    st.hitlist = f.hitlist.as<uint32_t>();
    st.hit_t = f.hit_t.as<float>();
    st.hit_tri = f.hit_tri.as<int>();
    st.ctr = f.ctr.as<uint32_t>();
    f.scene = sc;
    f.stack = choose_stack(ctx, sc);
    f.n_active = 0;
    f.n_hit = 0;
    f.open = true;
  });
}

int olhip_e0_sample(olhip_ctx* ctx, const float* jx, const float* jy) {
  return guarded([&] {
    require(ctx && jx && jy, "arguments must not be NULL");
    auto& f = ctx->e0;
    if (!f.open) throw OlError{OLHIP_ERR_RUNTIME, "olhip_e0_begin has not been called"};
    hipStream_t s = ctx->stream;
    const uint32_t n = f.params.width * f.params.height;
    OL_HIP_CHECK(hipMemcpyAsync(f.jx.ptr, jx, 4ull * n, hipMemcpyHostToDevice, s));
    OL_HIP_CHECK(hipMemcpyAsync(f.jy.ptr, jy, 4ull * n, hipMemcpyHostToDevice, s));
    e0_camera(f.params, f.state, f.jx.as<float>(), f.jy.as<float>(), s);
    OL_HIP_CHECK(hipGetLastError());
    f.n_active = n;
  });
}

int olhip_e0_trace(olhip_ctx* ctx, uint32_t* n_hit, uint32_t* n_active) {
  return guarded([&] {
    require(ctx && n_hit, "arguments must not be NULL");
    auto& f = ctx->e0;
    if (!f.open) throw OlError{OLHIP_ERR_RUNTIME, "olhip_e0_begin has not been called"};
    hipStream_t s = ctx->stream;
    if (n_active) *n_active = f.n_active;
    e0_trace(ctx->cu_count, f.stack, f.params, f.state, f.n_active, s);
    OL_HIP_CHECK(hipGetLastError());
    f.n_hit = e0_compact(f.state, f.n_active, s);
    uint32_t ctr[4];
    OL_HIP_CHECK(hipMemcpy(ctr, f.ctr.ptr, 16, hipMemcpyDeviceToHost));
    if (ctr[1]) throw OlError{OLHIP_ERR_RUNTIME, "traversal stack overflow"};
    *n_hit = f.n_hit;
  });
}

int olhip_e0_last_hits(olhip_ctx* ctx, int32_t* tri, uint32_t n_active) {
  return guarded([&] {
    require(ctx && tri, "arguments must not be NULL");
    auto& f = ctx->e0;
    require(n_active <= f.params.width * f.params.height, "n_active too large");
    OL_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (n_active) OL_HIP_CHECK(hipMemcpy(tri, f.htri.ptr, 4ull * n_active, hipMemcpyDeviceToHost));
  });
}

int olhip_e0_shade(olhip_ctx* ctx, const float* deviates, uint32_t n_hit) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    auto& f = ctx->e0;
    if (!f.open) throw OlError{OLHIP_ERR_RUNTIME, "olhip_e0_begin has not been called"};
    require(n_hit == f.n_hit, "deviate count does not match the last trace");
    require(n_hit == 0 || deviates, "deviates must not be NULL");
    hipStream_t s = ctx->stream;
    if (n_hit) {
      OL_HIP_CHECK(hipMemcpyAsync(f.gauss.ptr, deviates, 12ull * n_hit, hipMemcpyHostToDevice, s));
      e0_shade(f.params, f.state, n_hit, f.gauss.as<float>(), s);
      OL_HIP_CHECK(hipGetLastError());
      // The hit rays form the next active list, already in ascending order.
      OL_HIP_CHECK(hipMemcpyAsync(f.active.ptr, f.hitlist.ptr, 4ull * n_hit, hipMemcpyDeviceToDevice, s));
    }
    f.n_active = n_hit;
  });
}

int olhip_e0_sample_end(olhip_ctx* ctx) {
  return guarded([&] {
    require(ctx != nullptr, "ctx must not be NULL");
    auto& f = ctx->e0;
    if (!f.open) throw OlError{OLHIP_ERR_RUNTIME, "olhip_e0_begin has not been called"};
    e0_accumulate(f.params, f.state, ctx->stream);
    OL_HIP_CHECK(hipGetLastError());
    f.n_active = 0;
  });
}

int olhip_e0_end(olhip_ctx* ctx, uint32_t samples, float* out_rgba) {
  return guarded([&] {
    require(ctx && out_rgba, "arguments must not be NULL");
    require(samples >= 1, "samples must be positive");
    auto& f = ctx->e0;
    if (!f.open) throw OlError{OLHIP_ERR_RUNTIME, "olhip_e0_begin has not been called"};
    hipStream_t s = ctx->stream;
    e0_finish(f.params, f.state, samples, f.out.as<float4>(), s);
    OL_HIP_CHECK(hipGetLastError());
    OL_HIP_CHECK(hipStreamSynchronize(s));
    OL_HIP_CHECK(hipMemcpy(out_rgba, f.out.ptr, 16ull * f.params.width * f.params.height, hipMemcpyDeviceToHost));
    f.open = false;
  });
}

}  // extern "C"
