// Output stage of a frame kept on the device: tone-mapped RGBA8 and the
// pitch-linear NV12 / P010 video formats.
//
// Restates wavefront_tone_map.comp (exposure, ACES approximation, sRGB
// encode, packUnorm4x8), rgba_to_nv12.comp (BT.709 studio-range luma of each
// RGBA8 pixel, chroma of each 2x2 average) and hdr_to_p010.comp (the same on
// the half-float HDR image, 10-bit values in the high bits of 16-bit words,
// nearest-neighbour source scaling).  One thread handles one pixel (RGBA8) or
// one 4x2 block (NV12, P010), so every luma word and chroma word is written
// by exactly one thread.  Rounding to integers is round-half-even (rintf).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.h"
#include "output.h"

namespace {

OL_DEV float aces(float c) {
  return clampf((c * ((2.51f * c) + 0.03f)) / ((c * ((2.43f * c) + 0.59f)) + 0.14f), 0.0f, 1.0f);
}
OL_DEV float srgb_encode(float c) {
  const float lower = c * 12.92f;
  const float upper = (1.055f * powf(c, 1.0f / 2.4f)) - 0.055f;
  return c <= 0.0031308f ? lower : upper;
}
OL_DEV f3 display(float4 hdr, float exposure) {
  return mk3(srgb_encode(aces(fmaxf(hdr.x, 0.0f) * exposure)), srgb_encode(aces(fmaxf(hdr.y, 0.0f) * exposure)),
             srgb_encode(aces(fmaxf(hdr.z, 0.0f) * exposure)));
}
OL_DEV uint32_t unorm8(float v) { return (uint32_t)rintf(clampf(v, 0.0f, 1.0f) * 255.0f); }
OL_DEV uint32_t rgba8_of(float4 hdr, float exposure) {
  const f3 c = display(hdr, exposure);
  return ((unorm8(c.x) | (unorm8(c.y) << 8)) | (unorm8(c.z) << 16)) | (255u << 24);
}
// An RGBA8 pixel as the unorm floats an image load returns.
OL_DEV f3 rgb_of_rgba8(uint32_t w) {
  return mk3((float)(w & 255u) / 255.0f, (float)((w >> 8) & 255u) / 255.0f, (float)((w >> 16) & 255u) / 255.0f);
}
OL_DEV uint32_t byte_value(float v) { return (uint32_t)clampf(rintf(v), 0.0f, 255.0f); }
OL_DEV uint32_t ten_bit_value(float v) { return (uint32_t)clampf(rintf(v), 0.0f, 1023.0f); }
OL_DEV float luma709(f3 c) { return dot3(c, mk3(0.2126f, 0.7152f, 0.0722f)); }
OL_DEV float cb709(f3 c) { return dot3(c, mk3(-0.114572f, -0.385428f, 0.5f)); }
OL_DEV float cr709(f3 c) { return dot3(c, mk3(0.5f, -0.454153f, -0.045847f)); }
OL_DEV uint32_t pack4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  return ((a | (b << 8)) | (c << 16)) | (d << 24);
}

__global__ void __launch_bounds__(256) k_rgba8(const float4* __restrict__ hdr, uint32_t n, float exposure,
                                               uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = rgba8_of(hdr[i], exposure);
}

// One 4x2 block of an NV12 frame of the same size as the HDR frame.
__global__ void __launch_bounds__(256) k_nv12(const float4* __restrict__ hdr, uint32_t W, uint32_t H, uint32_t pitch,
                                              float exposure, uint32_t* __restrict__ words) {
  const uint32_t bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y;
  const uint32_t x = bx * 4u, y = by * 2u;
  if (x >= W || y >= H) return;
  f3 top[4], bottom[4];
  uint32_t ly[4], lb[4];
  for (int k = 0; k < 4; ++k) {
    const uint32_t sx = min(x + (uint32_t)k, W - 1u);
    const uint32_t y1 = min(y + 1u, H - 1u);
    top[k] = rgb_of_rgba8(rgba8_of(hdr[(size_t)y * W + sx], exposure));
    bottom[k] = rgb_of_rgba8(rgba8_of(hdr[(size_t)y1 * W + sx], exposure));
    ly[k] = byte_value(16.0f + 219.0f * luma709(top[k]));
    lb[k] = byte_value(16.0f + 219.0f * luma709(bottom[k]));
  }
  const uint32_t wpr = pitch / 4u;
  words[y * wpr + x / 4u] = pack4(ly[0], ly[1], ly[2], ly[3]);
  if (y + 1u < H) words[(y + 1u) * wpr + x / 4u] = pack4(lb[0], lb[1], lb[2], lb[3]);
  const f3 first = 0.25f * (((top[0] + top[1]) + bottom[0]) + bottom[1]);
  const f3 second = 0.25f * (((top[2] + top[3]) + bottom[2]) + bottom[3]);
  const uint32_t u0 = byte_value(128.0f + 224.0f * cb709(first)), v0 = byte_value(128.0f + 224.0f * cr709(first));
  const uint32_t u1 = byte_value(128.0f + 224.0f * cb709(second)), v1 = byte_value(128.0f + 224.0f * cr709(second));
  words[(pitch * H) / 4u + (y / 2u) * wpr + x / 4u] = pack4(u0, v0, u1, v1);
}

OL_DEV f3 p010_color(const float4* hdr, uint32_t SW, uint32_t SH, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                     float exposure) {
  const int sx = min(max((int)((x * SW) / W), 0), (int)SW - 1);
  const int sy = min(max((int)((y * SH) / H), 0), (int)SH - 1);
  float4 v = hdr[(size_t)sy * SW + (uint32_t)sx];
  // The presented HDR image holds half floats.
  v.x = __half2float(__float2half_rn(v.x));
  v.y = __half2float(__float2half_rn(v.y));
  v.z = __half2float(__float2half_rn(v.z));
  return display(v, exposure);
}

// One 4x2 block of a P010 frame of W x H scaled from the SW x SH HDR frame.
__global__ void __launch_bounds__(256) k_p010(const float4* __restrict__ hdr, uint32_t SW, uint32_t SH, uint32_t W,
                                              uint32_t H, uint32_t pitch, float exposure,
                                              uint32_t* __restrict__ words) {
  const uint32_t bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y;
  const uint32_t x = bx * 4u, y = by * 2u;
  if (x >= W || y >= H) return;
  f3 top[4], bottom[4];
  uint32_t ly[4], lb[4];
  for (int k = 0; k < 4; ++k) {
    top[k] = p010_color(hdr, SW, SH, x + (uint32_t)k, y, W, H, exposure);
    bottom[k] = p010_color(hdr, SW, SH, x + (uint32_t)k, y + 1u, W, H, exposure);
    ly[k] = ten_bit_value(64.0f + 876.0f * luma709(top[k])) << 6;
    lb[k] = ten_bit_value(64.0f + 876.0f * luma709(bottom[k])) << 6;
  }
  const uint32_t wpr = pitch / 4u;
  uint32_t word = y * wpr + x / 2u;
  words[word] = ly[0] | (ly[1] << 16);
  words[word + 1u] = ly[2] | (ly[3] << 16);
  if (y + 1u < H) {
    word = (y + 1u) * wpr + x / 2u;
    words[word] = lb[0] | (lb[1] << 16);
    words[word + 1u] = lb[2] | (lb[3] << 16);
  }
  const f3 first = 0.25f * (((top[0] + top[1]) + bottom[0]) + bottom[1]);
  const f3 second = 0.25f * (((top[2] + top[3]) + bottom[2]) + bottom[3]);
  const uint32_t c = (pitch * H) / 4u + (y / 2u) * wpr + x / 2u;
  words[c] = (ten_bit_value(512.0f + 896.0f * cb709(first)) << 6) |
             ((ten_bit_value(512.0f + 896.0f * cr709(first)) << 6) << 16);
  words[c + 1u] = (ten_bit_value(512.0f + 896.0f * cb709(second)) << 6) |
                  ((ten_bit_value(512.0f + 896.0f * cr709(second)) << 6) << 16);
}

}  // namespace

size_t output_bytes(uint32_t format, uint32_t width, uint32_t height) {
  const size_t n = (size_t)width * height;
  if (format == kOutputRgba8) return 4 * n;
  if (format == kOutputNv12) return n + n / 2;
  return 2 * n + n;  // P010: 16-bit luma plane + half-size interleaved chroma plane
}

void launch_output(uint32_t format, const float4* hdr, uint32_t src_w, uint32_t src_h, uint32_t width,
                   uint32_t height, float exposure, uint32_t* out, hipStream_t st) {
  if (format == kOutputRgba8) {
    const uint32_t n = width * height;
    hipLaunchKernelGGL(k_rgba8, dim3((n + 255u) / 256u), dim3(256), 0, st, hdr, n, exposure, out);
  } else if (format == kOutputNv12) {
    const dim3 grid((width / 4u + 255u) / 256u, height / 2u);
    hipLaunchKernelGGL(k_nv12, grid, dim3(256), 0, st, hdr, width, height, width, exposure, out);
  } else {
    const dim3 grid((width / 4u + 255u) / 256u, height / 2u);
    hipLaunchKernelGGL(k_p010, grid, dim3(256), 0, st, hdr, src_w, src_h, width, height, 2u * width, exposure, out);
  }
}
