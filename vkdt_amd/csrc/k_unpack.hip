// Bit-packed raw frames to 16-bit samples on the device (Magic Lantern raw
// video: 10, 12 or 14 bits per sample in 16-bit little-endian words, each
// word read from its most significant bit).  The host loop it replaces
// (reference i-mlv/video_mlv.c:262-276) reads a 32-bit window per sample
// and rotates; here each lane makes four samples from the words that hold
// them.  HBM-bound: bits/8 bytes read and 2 bytes written per sample.
#include "launch.h"

#include <vector>

namespace {

__device__ inline uint32_t up_word(const uint8_t *p, uint64_t w, uint64_t bytes)
{ // 16-bit little-endian word w, zero past the data
  const uint64_t b = 2 * w;
  return b + 1 < bytes ? (uint32_t)p[b] | ((uint32_t)p[b + 1] << 8) : 0u;
}

__global__ void __launch_bounds__(256) k_unpack_msb16(const uint8_t *packed, uint64_t bytes, uint16_t *out,
                                                      uint32_t n, int bits)
{
  const uint64_t i0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if(i0 >= n) return;
  const uint32_t mask = (1u << bits) - 1;
  uint16_t v[4];
  for(int j = 0; j < 4; j++)
  {
    const uint64_t bit = (i0 + j) * (uint64_t)bits, w = bit >> 4;
    const int s = (int)(bit & 15);
    const uint32_t win = (up_word(packed, w, bytes) << 16) | up_word(packed, w + 1, bytes);
    v[j] = (uint16_t)((win >> (32 - s - bits)) & mask);
  }
  if(i0 + 3 < n && (((uintptr_t)(out + i0)) & 7) == 0)
    *(uint2 *)(out + i0) = make_uint2(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16));
  else
    for(int j = 0; j < 4 && i0 + j < n; j++) out[i0 + j] = v[j];
}

} // namespace

extern "C" int vkdt_hip_unpack_msb16(const uint8_t *packed, size_t bytes, void *out, uint32_t n, int bits, void *stream)
{
  if(bits < 1 || bits > 16) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "unpack: %d bits per sample", bits);
  if(n == 0) return VKDT_HIP_OK;
  if(!out || !packed) return vkdt_set_error(VKDT_HIP_INCOMPLETE, "unpack: no buffer");
  uint8_t *dev = nullptr;
  if(int e = hip_fail(hipMalloc((void **)&dev, bytes + 8), "unpack staging")) return e;
  hipStream_t s = (hipStream_t)stream;
  int err = hip_fail(hipMemcpyAsync(dev, packed, bytes, hipMemcpyHostToDevice, s), "unpack upload");
  if(!err)
  {
    const uint64_t groups = ((uint64_t)n + 1023) / 1024;
    hipLaunchKernelGGL(k_unpack_msb16, dim3((uint32_t)groups), dim3(256), 0, s, (const uint8_t *)dev, (uint64_t)bytes,
                       (uint16_t *)out, n, bits);
    err = launch_post("unpack");
  }
  if(!err) err = hip_fail(hipStreamSynchronize(s), "unpack sync");
  (void)hipFree(dev);
  return err;
}
