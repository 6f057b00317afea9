// translate.hip -- camera-following atlas translation as one integer remap.
//
// Replaces the two vkCmdCopyImage steps of IrradianceProbeTranslationPass.cpp:56-186
// (atlas -> temp, then per probe layer a region shifted by one 10x10 tile back
// into the atlas).  Per texel (X, Y) of layer plane i (height 10*probes_z):
//   inside the destination region of the shifted copy  -> read the source texel
//   (X + 10*sx, Y + 10*sz), sx = sign(dx), sz = sign(dz)
//   outside it (the entry edge)                          -> keep the old texel
// Done out of place into the second atlas buffer: 2 x atlas bytes of HBM
// traffic instead of the reference's 4 x.
#include "common.hpp"
#include "kernels.hpp"

namespace ane {

__global__ __launch_bounds__(256) void k_translate(int W, int H, int plane_h, int sx, int sz,
                                                   const float4* src, float4* dst)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)W * H) return;
  const int Y = (int)(i / W), X = (int)(i - (int64_t)Y * W);
  const int yy = Y % plane_h;
  // destination region of the copy (IrradianceProbeTranslationPass.cpp:137-157)
  const bool in_x = (sx > 0) ? (X < W - 10) : (sx < 0 ? X >= 10 : true);
  const bool in_y = (sz > 0) ? (yy < plane_h - 10) : (sz < 0 ? yy >= 10 : true);
  int64_t s = i;
  if (in_x && in_y) s = (int64_t)(Y + 10 * sz) * W + (X + 10 * sx);
  dst[i] = src[s];
}

void launch_translate(const GridK& g, const float4* src, float4* dst, int dx, int dz, hipStream_t s)
{
  const int64_t n = (int64_t)g.atlas_w * g.atlas_h;
  const int sx = dx > 0 ? 1 : (dx < 0 ? -1 : 0);
  const int sz = dz > 0 ? 1 : (dz < 0 ? -1 : 0);
  hipLaunchKernelGGL(k_translate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.atlas_w,
                     g.atlas_h, 10 * g.nz, sx, sz, src, dst);
}

// ---------------------------------------------------------------------------
// RGBA16F form of the atlas, the format the reference stores a baked tile in
// (VulkanRenderer.cpp:2705-2770: R16G16B16A16_SFLOAT image copied to a buffer):
// IEEE binary16 per channel, round to nearest even, texel order unchanged.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_atlas_to_half(const float4* src, uint2* dst, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = src[i];
  const _Float16 h[4] = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
  uint16_t b[4];
  __builtin_memcpy(b, h, 8);
  dst[i] = make_uint2((uint32_t)b[0] | ((uint32_t)b[1] << 16), (uint32_t)b[2] | ((uint32_t)b[3] << 16));
}

__global__ __launch_bounds__(256) void k_half_to_atlas(const uint2* src, float4* dst, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint2 p = src[i];
  const uint16_t b[4] = {(uint16_t)(p.x & 0xFFFFu), (uint16_t)(p.x >> 16), (uint16_t)(p.y & 0xFFFFu),
                         (uint16_t)(p.y >> 16)};
  _Float16 h[4];
  __builtin_memcpy(h, b, 8);
  dst[i] = make_float4((float)h[0], (float)h[1], (float)h[2], (float)h[3]);
}

void launch_atlas_to_half(const float4* src, uint2* dst, int64_t texels, hipStream_t s)
{
  if (texels <= 0) return;
  hipLaunchKernelGGL(k_atlas_to_half, dim3((unsigned)((texels + 255) / 256)), dim3(256), 0, s, src, dst,
                     texels);
}

void launch_half_to_atlas(const uint2* src, float4* dst, int64_t texels, hipStream_t s)
{
  if (texels <= 0) return;
  hipLaunchKernelGGL(k_half_to_atlas, dim3((unsigned)((texels + 255) / 256)), dim3(256), 0, s, src, dst,
                     texels);
}

}  // namespace ane
