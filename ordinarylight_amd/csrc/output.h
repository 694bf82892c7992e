// Output stage launchers (output.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

constexpr uint32_t kOutputRgba8 = 0, kOutputNv12 = 1, kOutputP010 = 2;

// Bytes of a width x height frame in `format` (NV12 pitch = width bytes, P010
// pitch = 2 * width bytes; the chroma plane follows the luma plane).
size_t output_bytes(uint32_t format, uint32_t width, uint32_t height);

// Writes the frame: RGBA8 and NV12 take a width x height HDR image
// (src_w == width, src_h == height); P010 scales src_w x src_h to width x
// height by nearest neighbour.  NV12 and P010 need width % 4 == 0 and even
// height.
void launch_output(uint32_t format, const float4* hdr, uint32_t src_w, uint32_t src_h, uint32_t width,
                   uint32_t height, float exposure, uint32_t* out, hipStream_t st);
