// conv.hip -- probe convolution + hysteresis blend on gfx950.
//
// Replaces probe_conv.comp (dispatched at IrradianceProbeConvolvePass.cpp:109-143).
// The reference loops, per atlas texel, over the N rays of its probe
// (probe_conv.comp:49-61).  The cosine weights W[t][r] = max(0, d_r . o_t) only
// depend on the texel direction and the ray direction, which every probe
// shares, so the whole pass is one contraction
//     E[t][p][c] = sum_r Wn[t][r] * L[p][c][r],   Wn = W / sum_r W
// (M = 64 texels, K = N rays, columns = probes x 3 channels), followed by the
// FP32 hysteresis epilogue (probe_conv.comp:63-89) written straight into the
// atlas texels of each probe:
//   k_conv_fp32  plain FP32 register-tiled kernel (the drop-in default)
//   k_conv_mfma  BF16 (v_mfma_f32_16x16x32_bf16) or FP8-e4m3
//                (v_mfma_f32_16x16x32_fp8_fp8) operands, FP32 accumulate;
//                FP8 carries a per-texel weight scale and a per-probe radiance scale
#include "common.hpp"
#include "kernels.hpp"

namespace ane {

// ---------------------------------------------------------------------------
// epilogue: hysteresis blend / bake accumulation of one texel
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 blend_texel(float4 old, float r, float g, float b, float wsum,
                                              float h, bool bake)
{
  float4 o;
  if (bake) {
    // probe_conv.comp:67-79 (r, g, b are the unnormalized sums sum W L)
    float tw = old.w + wsum;
    float x = r + old.x * old.w, y = g + old.y * old.w, z = b + old.z * old.w;
    if (tw > 1e-6f) {
      x = x / tw;
      y = y / tw;
      z = z / tw;
    }
    o = make_float4(x, y, z, tw);
  } else {
    // probe_conv.comp:81-86 (alpha blends the unnormalized weight sum)
    float omh = 1.0f - h;
    o = make_float4(r * omh + old.x * h, g * omh + old.y * h, b * omh + old.z * h,
                    wsum * omh + old.w * h);
  }
  return o;
}

__device__ __forceinline__ int64_t texel_index(const GridK& g, int q, int t)
{
  const int px = q % g.nx, pz = (q / g.nx) % g.nz, py = q / (g.nx * g.nz);
  const int64_t X = 10 * px + 1 + (t & 7);
  const int64_t Y = 10 * ((int64_t)py * g.nz + pz) + 1 + (t >> 3);
  return Y * g.atlas_w + X;
}

// ---------------------------------------------------------------------------
// plain FP32: block = 32 probes x 64 texels, thread = 8 texels (one tile row) x 1 probe x 3
// ---------------------------------------------------------------------------
static constexpr int kPB = 32;
static constexpr int kKC = 32;

__global__ __launch_bounds__(256) void k_conv_fp32(GridK g, const float* rad, const float* w32,
                                                   const float* wsum, float4* atlas, int nprobes,
                                                   int probe_base, int32_t bake)
{
  __shared__ float Ws[kKC][64];
  __shared__ float Ls[kPB][3 * kKC + 1];
  const int n = g.rays;
  const int tid = threadIdx.x;
  const int tg = tid & 7;   // tile row (8 texels)
  const int pl = tid >> 3;  // probe within block
  const int p0 = blockIdx.x * kPB;
  float acc[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j][0] = acc[j][1] = acc[j][2] = 0.0f;
  for (int k0 = 0; k0 < n; k0 += kKC) {
    // W chunk: 64 texels x 32 rays, stored transposed
    for (int e = tid; e < 64 * kKC; e += 256) {
      int t = e / kKC, k = e % kKC;
      Ws[k][t] = (k0 + k < n) ? w32[(int64_t)t * n + k0 + k] : 0.0f;
    }
    // L chunk: 32 probes x 3 channels x 32 rays
    for (int e = tid; e < kPB * 3 * kKC; e += 256) {
      int row = e / kKC, k = e % kKC;
      int p = row / 3, c = row % 3;
      float v = 0.0f;
      if (p0 + p < nprobes && k0 + k < n) v = rad[((int64_t)(p0 + p) * 3 + c) * n + k0 + k];
      Ls[p][c * kKC + k] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kKC; ++k) {
      const float4 wa = *reinterpret_cast<const float4*>(&Ws[k][tg * 8]);
      const float4 wb = *reinterpret_cast<const float4*>(&Ws[k][tg * 8 + 4]);
      const float l0 = Ls[pl][k], l1 = Ls[pl][kKC + k], l2 = Ls[pl][2 * kKC + k];
      const float w[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        acc[j][0] = fmaf(w[j], l0, acc[j][0]);
        acc[j][1] = fmaf(w[j], l1, acc[j][1]);
        acc[j][2] = fmaf(w[j], l2, acc[j][2]);
      }
    }
    __syncthreads();
  }
  const int pq = p0 + pl;
  if (pq >= nprobes) return;
  const int q = probe_base + pq;
  const float h = g.hysteresis;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int t = tg * 8 + j;
    const int64_t ti = texel_index(g, q, t);
    const float4 old = atlas[ti];
    atlas[ti] = blend_texel(old, acc[j][0], acc[j][1], acc[j][2], wsum[t], h, bake != 0);
  }
}

// ---------------------------------------------------------------------------
// MFMA: one wave = 64 texels (4 row tiles) x 16 probes x 3 channels.
//   A (weights)   16 x 32 per tile, lane l: row l%16, k = 8*(l/16) .. +8
//   B (radiance)  32 x 16 per tile, lane l: column (probe) l%16, same k range
//   C             16 x 16, lane l: column l%16, rows 4*(l/16) + i, i = 0..3
// The weights of a K chunk are staged in LDS once per block.
// ---------------------------------------------------------------------------
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

static constexpr int kWaves = 4;
static constexpr int kKChunk = 256;  // rays of A staged per LDS fill

template <int PREC>
__global__ __launch_bounds__(64 * kWaves) void k_conv_mfma(GridK g, const void* rad,
                                                          const float* rad_scale, const void* wq,
                                                          const float* wscale, const float* wsum,
                                                          float4* atlas, int nprobes,
                                                          int probe_base, int32_t bake)
{
  constexpr int EB = (PREC == ANE_PRECISION_BF16) ? 2 : 1;  // bytes per element
  __shared__ __attribute__((aligned(16))) uint8_t As[64 * kKChunk * EB];
  const int n = g.rays;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int pw = (blockIdx.x * kWaves + wave) * 16;  // first probe of the wave
  const int pcol = pw + col;
  const bool pvalid = pcol < nprobes;
  const uint8_t* radb = (const uint8_t*)rad;
  const int64_t rowB = (int64_t)3 * n * EB;  // bytes per probe
  f32x4 acc[4][3];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[rt][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int k0 = 0; k0 < n; k0 += kKChunk) {
    const int kc = min(kKChunk, n - k0);
    // stage A[64][kc] (row stride kKChunk) with 16-byte copies
    const int vec_per_row = kc * EB / 16;
    for (int e = tid; e < 64 * vec_per_row; e += 64 * kWaves) {
      int t = e / vec_per_row, v = e % vec_per_row;
      const uint4* src = reinterpret_cast<const uint4*>((const uint8_t*)wq + ((int64_t)t * n + k0) * EB) + v;
      *reinterpret_cast<uint4*>(&As[(t * kKChunk) * EB + v * 16]) = *src;
    }
    __syncthreads();
    for (int kk = 0; kk < kc; kk += 32) {
      const int kl = kk + 8 * kq;  // this lane's k offset within the chunk
      if (PREC == ANE_PRECISION_BF16) {
        bf16x8 a[4], b[3];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
          a[rt] = *reinterpret_cast<const bf16x8*>(&As[((rt * 16 + col) * kKChunk + kl) * 2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (pvalid)
            b[c] = *reinterpret_cast<const bf16x8*>(radb + pcol * rowB + ((int64_t)c * n + k0 + kl) * 2);
          else
            b[c] = bf16x8{};
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            acc[rt][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], b[c], acc[rt][c], 0, 0, 0);
      } else {
        long a[4], b[3];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
          a[rt] = *reinterpret_cast<const long*>(&As[(rt * 16 + col) * kKChunk + kl]);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          b[c] = pvalid ? *reinterpret_cast<const long*>(radb + pcol * rowB + (int64_t)c * n + k0 + kl)
                        : 0l;
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            acc[rt][c] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a[rt], b[c], acc[rt][c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (!pvalid) return;
  const int q = probe_base + pcol;
  const float h = g.hysteresis;
  const float ps = (PREC == ANE_PRECISION_FP8) ? rad_scale[pcol] : 1.0f;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = rt * 16 + 4 * kq + i;
      // per-texel weight scale (FP8 dequantization) x weight-sum correction
      const float s = (PREC == ANE_PRECISION_FP8) ? wscale[t] * ps : wscale[64 + t];
      const int64_t ti = texel_index(g, q, t);
      const float4 old = atlas[ti];
      atlas[ti] = blend_texel(old, acc[rt][0][i] * s, acc[rt][1][i] * s, acc[rt][2][i] * s, wsum[t],
                              h, bake != 0);
    }
}

void launch_convolve(const GridK& g, const FrameK& f, int precision, const void* rad,
                     const float* rad_scale, const float* w32, const float* wsum,
                     const uint16_t* wbf16, const uint8_t* wfp8, const float* wfp8_scale,
                     float* atlas, int num_probes, int probe_base, bool bake, hipStream_t s)
{
  (void)f;
  if (num_probes <= 0) return;
  float4* a4 = reinterpret_cast<float4*>(atlas);
  if (precision == ANE_PRECISION_FP32) {
    const unsigned blocks = (unsigned)((num_probes + kPB - 1) / kPB);
    hipLaunchKernelGGL(k_conv_fp32, dim3(blocks), dim3(256), 0, s, g, (const float*)rad, w32, wsum,
                       a4, num_probes, probe_base, (int32_t)bake);
    return;
  }
  const unsigned blocks = (unsigned)((num_probes + 16 * kWaves - 1) / (16 * kWaves));
  if (precision == ANE_PRECISION_BF16)
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_BF16>, dim3(blocks), dim3(64 * kWaves), 0, s, g,
                       rad, rad_scale, (const void*)wbf16, wfp8_scale, wsum, a4,
                       num_probes, probe_base, (int32_t)bake);
  else
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_FP8>, dim3(blocks), dim3(64 * kWaves), 0, s, g,
                       rad, rad_scale, (const void*)wfp8, wfp8_scale, wsum, a4, num_probes,
                       probe_base, (int32_t)bake);
}

// ---------------------------------------------------------------------------
// radiance upload/download in the context precision (one block per probe)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_quantize(int prec, const float* src, void* dst,
                                                  float* scale, int n)
{
  __shared__ float red[256];
  const int p = blockIdx.x;
  const float* s = src + (int64_t)p * 3 * n;
  if (prec == ANE_PRECISION_BF16) {
    uint16_t* d = (uint16_t*)dst + (int64_t)p * 3 * n;
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f32_to_bf16(s[i]);
    return;
  }
  float mx = 0.0f;
  for (int i = threadIdx.x; i < 3 * n; i += 256) mx = fmaxf(mx, fabsf(s[i]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  const float sc = red[0] > 0.0f ? red[0] / 448.0f : 1.0f;
  uint8_t* d = (uint8_t*)dst + (int64_t)p * 3 * n;
  for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f32_to_e4m3(s[i] / sc);
  if (threadIdx.x == 0) scale[p] = sc;
}

__global__ __launch_bounds__(256) void k_dequantize(int prec, const void* src, const float* scale,
                                                    float* dst, int n)
{
  const int p = blockIdx.x;
  float* d = dst + (int64_t)p * 3 * n;
  if (prec == ANE_PRECISION_BF16) {
    const uint16_t* s = (const uint16_t*)src + (int64_t)p * 3 * n;
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = bf16_to_f32(s[i]);
  } else {
    const uint8_t* s = (const uint8_t*)src + (int64_t)p * 3 * n;
    const float sc = scale[p];
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = e4m3_to_f32(s[i]) * sc;
  }
}

void launch_quantize(int precision, const float* src, void* dst, float* scale, int num_probes,
                     int n, hipStream_t s)
{
  if (num_probes <= 0) return;
  hipLaunchKernelGGL(k_quantize, dim3(num_probes), dim3(256), 0, s, precision, src, dst, scale, n);
}

void launch_dequantize(int precision, const void* src, const float* scale, float* dst,
                       int num_probes, int n, hipStream_t s)
{
  if (num_probes <= 0) return;
  hipLaunchKernelGGL(k_dequantize, dim3(num_probes), dim3(256), 0, s, precision, src, scale, dst, n);
}

}  // namespace ane
