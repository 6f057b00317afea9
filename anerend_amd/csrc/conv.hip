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
//   k_conv_mfma  FP32 (v_mfma_f32_16x16x4_f32, the plain drop-in path), BF16
//                (v_mfma_f32_16x16x32_bf16) or FP8-e4m3 (v_mfma_f32_16x16x32_fp8_fp8)
//                operands, FP32 accumulate; FP8 carries a per-texel weight scale
//                and an FP32 radiance scale per probe, channel and 32 rays
//   k_conv_fp32  FP32 register-tiled kernel for ray counts that are not a
//                multiple of 32
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
// MFMA convolution (FP32, BF16, FP8 operands, FP32 accumulate).
// Block = 16 probes (one MFMA column tile) x 64 texels x 3 channels, 4 waves.
// The 4 waves split the rays: wave w takes the 32-ray steps s = w, w + 4, ...,
// loading the weight (A) and radiance (B) fragments of CH steps at a time, and
// accumulates 4 row tiles x 3 channels.
// The partial sums meet in LDS; then all 256 threads run the hysteresis
// epilogue with consecutive threads on consecutive texels of a tile row (each
// probe's row is 128 contiguous bytes of the atlas).
//   16x16x32 BF16 / FP8: lane l holds A[row l%16][k = 8(l/16) .. +7] and
//     B[same k][col l%16]; C rows 4(l/16) + i, col l%16
//   16x16x4 FP32: the same 8 k per lane, one per instruction: instruction e
//     contracts k = 8q + e over the lane groups q = 0..3 (a permutation of the
//     32 rays of the step, the same for A and B)
// ---------------------------------------------------------------------------
// 32-ray steps a wave loads at once (16-bit / FP8 and FP32 operands).  One step
// keeps the registers low enough for 4 waves per SIMD; 2 and 4 measured 3-8% slower.
#ifndef ANE_CONV_CH
#define ANE_CONV_CH 1
#endif
#ifndef ANE_CONV_CH32
#define ANE_CONV_CH32 1
#endif
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) int i32x4;

template <int PREC>
struct ConvFrag;
template <>
struct ConvFrag<ANE_PRECISION_FP32> {
  static constexpr int EB = 4, CH = ANE_CONV_CH32;
  float v[8];
  __device__ __forceinline__ void load(const uint8_t* p)
  {
    const float4 x = *reinterpret_cast<const float4*>(p), y = *reinterpret_cast<const float4*>(p + 16);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w; v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
  }
  __device__ __forceinline__ void zero()
  {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = 0.0f;
  }
  static __device__ __forceinline__ f32x4 mma(const ConvFrag& a, const ConvFrag& b, f32x4 c)
  {
#pragma unroll
    for (int e = 0; e < 8; ++e) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v[e], b.v[e], c, 0, 0, 0);
    return c;
  }
};
template <>
struct ConvFrag<ANE_PRECISION_BF16> {
  static constexpr int EB = 2, CH = ANE_CONV_CH;
  bf16x8 v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *reinterpret_cast<const bf16x8*>(p); }
  __device__ __forceinline__ void zero() { v = bf16x8{}; }
  static __device__ __forceinline__ f32x4 mma(const ConvFrag& a, const ConvFrag& b, f32x4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, c, 0, 0, 0);
  }
};
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
template <>
struct ConvFrag<ANE_PRECISION_FP16> {
  static constexpr int EB = 2, CH = ANE_CONV_CH;
  f16x8 v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *reinterpret_cast<const f16x8*>(p); }
  __device__ __forceinline__ void zero() { v = f16x8{}; }
  static __device__ __forceinline__ f32x4 mma(const ConvFrag& a, const ConvFrag& b, f32x4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a.v, b.v, c, 0, 0, 0);
  }
};
template <>
struct ConvFrag<ANE_PRECISION_FP8> {
  static constexpr int EB = 1, CH = ANE_CONV_CH;
  long v;
  __device__ __forceinline__ void load(const uint8_t* p) { v = *reinterpret_cast<const long*>(p); }
  __device__ __forceinline__ void zero() { v = 0l; }
  static __device__ __forceinline__ f32x4 mma(const ConvFrag& a, const ConvFrag& b, f32x4 c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a.v, b.v, c, 0, 0, 0);
  }
};

static constexpr int kConvProbes = 16;
static constexpr int kConvWaves = 4;
static constexpr int kPartStride = 17;  // texel stride of the LDS partials (probe-minor, padded)

// LDS of the partial sums: waves 2 and 3 hand theirs to waves 0 and 1, which
// add them and leave two partials for the epilogue (26 KB, 6 blocks per CU)
typedef float ConvPart[2][3][64 * kPartStride];

// One block's 16 probes (see above); blk is the block's index within its update.
template <int PREC>
__device__ __forceinline__ void conv_block(const GridK& g, const void* rad, const float* rad_scale,
                                           const void* wq, const float* wscale, const float* wsum,
                                           float4* atlas, int nprobes, int probe_base, int32_t bake, int blk,
                                           ConvPart& part)
{
  using F = ConvFrag<PREC>;
  constexpr int EB = F::EB, CH = F::CH;
  const int n = g.rays, S = n / 32;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int p0 = blk * kConvProbes;
  const bool pvalid = p0 + col < nprobes;
  const uint8_t* radp = (const uint8_t*)rad + (int64_t)(pvalid ? p0 + col : 0) * 3 * n * EB;
  // FP8: the scale of each (channel, 32-ray step) of this lane's probe
  const float* rsp = rad_scale + (int64_t)(pvalid ? p0 + col : 0) * 3 * S;
  const uint8_t* wp = (const uint8_t*)wq;
  f32x4 acc[4][3];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[rt][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int s0 = wave; s0 < S; s0 += kConvWaves * CH) {
    F a[CH][4], b[CH][3];
    float bs[CH][3];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int s = s0 + kConvWaves * j;
      const int64_t k = 32 * s + 8 * kq;
      if (PREC == ANE_PRECISION_FP8) {
#pragma unroll
        for (int c = 0; c < 3; ++c) bs[j][c] = (s < S && pvalid) ? rsp[c * S + s] : 0.0f;
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        if (s < S) a[j][rt].load(wp + ((int64_t)(rt * 16 + col) * n + k) * EB);
        else a[j][rt].zero();
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (s < S && pvalid) b[j][c].load(radp + ((int64_t)c * n + k) * EB);
        else b[j][c].zero();
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (PREC == ANE_PRECISION_FP8) {
            // each 32-ray step carries its own radiance scale: a fresh product, scaled in
            const f32x4 t = F::mma(a[j][rt], b[j][c], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            acc[rt][c] += t * bs[j][c];
          } else {
            acc[rt][c] = F::mma(a[j][rt], b[j][c], acc[rt][c]);
          }
        }
  }
  // partial sums, part[w][c][t * stride + probe]: waves 2, 3 store theirs, waves
  // 0, 1 add them to their own and store the pair sums
  if (wave >= 2) {
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = rt * 16 + 4 * kq + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) part[wave - 2][c][t * kPartStride + col] = acc[rt][c][i];
      }
  }
  __syncthreads();
  if (wave < 2) {
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = rt * 16 + 4 * kq + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float* q = &part[wave][c][t * kPartStride + col];
          *q = acc[rt][c][i] + *q;
        }
      }
  }
  __syncthreads();
  const float h = g.hysteresis;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int item = tid + 256 * j;  // (ty, probe, tx), tx fastest
    const int tx = item & 7, p = (item >> 3) & 15, ty = item >> 7;
    if (p0 + p >= nprobes) continue;
    const int t = ty * 8 + tx;
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = part[0][c][t * kPartStride + p] + part[1][c][t * kPartStride + p];
    // per-texel weight scale (FP8 dequantization) or weight-sum correction (BF16)
    const float sc = (PREC == ANE_PRECISION_FP8) ? wscale[t]
                     : (PREC == ANE_PRECISION_BF16 || PREC == ANE_PRECISION_FP16) ? wscale[64 + t] : 1.0f;
    const int64_t ti = texel_index(g, probe_base + p0 + p, t);
    const float4 old = atlas[ti];
    atlas[ti] = blend_texel(old, r[0] * sc, r[1] * sc, r[2] * sc, wsum[t], h, bake != 0);
  }
}

template <int PREC>
__global__ __launch_bounds__(64 * kConvWaves) void k_conv_mfma(GridK g, const void* rad,
                                                              const float* rad_scale, const void* wq,
                                                              const float* wscale, const float* wsum,
                                                              float4* atlas, int nprobes,
                                                              int probe_base, int32_t bake)
{
  __shared__ ConvPart part;
  conv_block<PREC>(g, rad, rad_scale, wq, wscale, wsum, atlas, nprobes, probe_base, bake, blockIdx.x, part);
}

// The convolutions of many updates in one launch (the request loop's fused
// stage): block b runs block b - block_begin of the last job whose block_begin <= b.
template <int PREC>
__global__ __launch_bounds__(64 * kConvWaves) void k_conv_jobs(const ConvJob* jobs, int njobs)
{
  __shared__ ConvPart part;
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  const ConvJob& j = jobs[lo];
  conv_block<PREC>(j.g, j.rad, j.rad_scale, j.wq, j.wscale, j.wsum, j.atlas, j.nprobes, j.probe_base, j.bake,
                   b - j.block_begin, part);
}

int conv_job_blocks(int num_probes) { return (num_probes + kConvProbes - 1) / kConvProbes; }

void launch_conv_jobs(int precision, const ConvJob* jobs, int njobs, int total_blocks, hipStream_t s)
{
  if (njobs <= 0 || total_blocks <= 0) return;
  const dim3 grid((unsigned)total_blocks), block(64 * kConvWaves);
  if (precision == ANE_PRECISION_FP32)
    hipLaunchKernelGGL(k_conv_jobs<ANE_PRECISION_FP32>, grid, block, 0, s, jobs, njobs);
  else if (precision == ANE_PRECISION_FP16)
    hipLaunchKernelGGL(k_conv_jobs<ANE_PRECISION_FP16>, grid, block, 0, s, jobs, njobs);
  else if (precision == ANE_PRECISION_BF16)
    hipLaunchKernelGGL(k_conv_jobs<ANE_PRECISION_BF16>, grid, block, 0, s, jobs, njobs);
  else
    hipLaunchKernelGGL(k_conv_jobs<ANE_PRECISION_FP8>, grid, block, 0, s, jobs, njobs);
}

void launch_convolve(const GridK& g, const FrameK& f, int precision, const void* rad,
                     const float* rad_scale, const float* w32, const float* wsum,
                     const uint16_t* wbf16, const uint8_t* wfp8, const float* wfp8_scale,
                     float* atlas, int num_probes, int probe_base, bool bake, hipStream_t s)
{
  (void)f;
  if (num_probes <= 0) return;
  float4* a4 = reinterpret_cast<float4*>(atlas);
  const unsigned blocks = (unsigned)((num_probes + kConvProbes - 1) / kConvProbes);
  if (precision == ANE_PRECISION_FP32) {
    if (g.rays % 32 != 0) {  // the register-tiled kernel takes any ray count
      const unsigned b32 = (unsigned)((num_probes + kPB - 1) / kPB);
      hipLaunchKernelGGL(k_conv_fp32, dim3(b32), dim3(256), 0, s, g, (const float*)rad, w32, wsum,
                         a4, num_probes, probe_base, (int32_t)bake);
      return;
    }
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_FP32>, dim3(blocks), dim3(64 * kConvWaves), 0, s, g, rad,
                       rad_scale, (const void*)w32, wfp8_scale, wsum, a4, num_probes, probe_base,
                       (int32_t)bake);
  } else if (precision == ANE_PRECISION_FP16) {
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_FP16>, dim3(blocks), dim3(64 * kConvWaves), 0, s, g,
                       rad, rad_scale, (const void*)wbf16, wfp8_scale, wsum, a4,
                       num_probes, probe_base, (int32_t)bake);
  } else if (precision == ANE_PRECISION_BF16) {
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_BF16>, dim3(blocks), dim3(64 * kConvWaves), 0, s, g,
                       rad, rad_scale, (const void*)wbf16, wfp8_scale, wsum, a4,
                       num_probes, probe_base, (int32_t)bake);
  } else {
    hipLaunchKernelGGL(k_conv_mfma<ANE_PRECISION_FP8>, dim3(blocks), dim3(64 * kConvWaves), 0, s, g,
                       rad, rad_scale, (const void*)wfp8, wfp8_scale, wsum, a4, num_probes,
                       probe_base, (int32_t)bake);
  }
}

// ---------------------------------------------------------------------------
// radiance upload/download in the context precision (one block per probe)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_quantize(int prec, const float* src, void* dst,
                                                  float* scale, int n)
{
  __shared__ float red[3 * 4096 / 32];  // block factors for up to 4096 rays
  const int p = blockIdx.x;
  const float* s = src + (int64_t)p * 3 * n;
  if (prec == ANE_PRECISION_BF16 || prec == ANE_PRECISION_FP16) {
    uint16_t* d = (uint16_t*)dst + (int64_t)p * 3 * n;
    if (prec == ANE_PRECISION_BF16)
      for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f32_to_bf16(s[i]);
    else
      for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f32_to_f16_sat(s[i]);
    return;
  }
  // one scale per channel and 32 rays (n % 32 == 0), as the shading stores them
  float* se = red;  // [3 n / 32] quantization factors
  const int nb = 3 * (n / 32);
  float* sco = scale + (int64_t)p * nb;
  for (int b = threadIdx.x; b < nb; b += 256) {
    float m = 0.0f;
    for (int k = 0; k < 32; ++k) m = fmaxf(m, fabsf(s[32 * b + k]));
    se[b] = fp8_block_inv(m);
    sco[b] = fp8_block_scale(m);
  }
  __syncthreads();
  uint8_t* d = (uint8_t*)dst + (int64_t)p * 3 * n;
  for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f32_to_e4m3(s[i] * se[i >> 5]);
}

__global__ __launch_bounds__(256) void k_dequantize(int prec, const void* src, const float* scale,
                                                    float* dst, int n)
{
  const int p = blockIdx.x;
  float* d = dst + (int64_t)p * 3 * n;
  if (prec == ANE_PRECISION_BF16) {
    const uint16_t* s = (const uint16_t*)src + (int64_t)p * 3 * n;
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = bf16_to_f32(s[i]);
  } else if (prec == ANE_PRECISION_FP16) {
    const uint16_t* s = (const uint16_t*)src + (int64_t)p * 3 * n;
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = f16_to_f32(s[i]);
  } else {
    const uint8_t* s = (const uint8_t*)src + (int64_t)p * 3 * n;
    const float* se = scale + (int64_t)p * 3 * (n / 32);
    for (int i = threadIdx.x; i < 3 * n; i += 256) d[i] = e4m3_to_f32(s[i]) * se[i >> 5];
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

// ---------------------------------------------------------------------------
// Validation: the irradiance of a convolution against a float64 evaluation of
// the reference's sum (probe_conv.comp:49-61) over FP32 radiance, reduced on
// the GPU.  One block per probe: the probe's radiance and the ray directions
// are staged in LDS; thread t < 64 evaluates texel t in float64:
//   W = max(0, d_r . o_t) (kept if > 1e-6), E64 = sum W L / sum W,
// recovers the convolution's E from the blended atlas, E = (new - h old)/(1 - h),
// and folds |E - E64| / max(|E64|, floor) over its 3 channels into the block's
// max and sum of squares, then one atomic each (max as the bit pattern of a
// non-negative double).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_conv_err64(GridK g, const float* dirs, const double* texel_dirs,
                                                   const float* rad, const float4* atlas_new,
                                                   const float4* atlas_old, double h, double floor_,
                                                   int probe_base, unsigned long long* out_max,
                                                   double* out_sumsq)
{
  extern __shared__ float sm[];  // [3][N] radiance, then [N][3] directions
  const int n = g.rays;
  const int ql = blockIdx.x;
  const float* L = rad + (int64_t)ql * 3 * n;
  for (int i = threadIdx.x; i < 3 * n; i += blockDim.x) {
    sm[i] = L[i];
    sm[3 * n + i] = dirs[i];
  }
  __syncthreads();
  const int t = threadIdx.x;
  const double ox = texel_dirs[3 * t], oy = texel_dirs[3 * t + 1], oz = texel_dirs[3 * t + 2];
  double acc[3] = {0.0, 0.0, 0.0}, ws = 0.0;
  for (int r = 0; r < n; ++r) {
    const float* d = sm + 3 * n + 3 * r;
    double w = (double)d[0] * ox + (double)d[1] * oy + (double)d[2] * oz;
    if (w > 1e-6) {
      acc[0] += w * (double)sm[r];
      acc[1] += w * (double)sm[n + r];
      acc[2] += w * (double)sm[2 * n + r];
      ws += w;
    }
  }
  const int64_t ti = texel_index(g, probe_base + ql, t);
  const float4 a = atlas_new[ti];
  const float4 o = atlas_old ? atlas_old[ti] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const double e[3] = {((double)a.x - h * (double)o.x) / (1.0 - h), ((double)a.y - h * (double)o.y) / (1.0 - h),
                       ((double)a.z - h * (double)o.z) / (1.0 - h)};
  double mx = 0.0, sq = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double ref = ws > 1e-6 ? acc[c] / ws : acc[c];
    const double den = fabs(ref) > floor_ ? fabs(ref) : floor_;
    const double rel = fabs(e[c] - ref) / den;
    mx = rel > mx ? rel : mx;
    sq += rel * rel;
  }
  // wave reduction (the block is one wave of 64)
  for (int off = 32; off > 0; off >>= 1) {
    const double m2 = __shfl_xor(mx, off, 64);
    const double s2 = __shfl_xor(sq, off, 64);
    mx = m2 > mx ? m2 : mx;
    sq += s2;
  }
  if (t == 0) {
    atomicMax(out_max, (unsigned long long)__double_as_longlong(mx));
    atomicAdd(out_sumsq, sq);
  }
}

void launch_conv_err64(const GridK& g, const float* dirs, const double* texel_dirs, const float* rad,
                       const float* atlas_new, const float* atlas_old, double h, double floor_,
                       int num_probes, int probe_base, unsigned long long* out_max, double* out_sumsq,
                       hipStream_t s)
{
  if (num_probes <= 0) return;
  const size_t lds = sizeof(float) * 6 * (size_t)g.rays;  // up to 96 KB at 4096 rays
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_conv_err64,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_conv_err64, dim3(num_probes), dim3(64), lds, s, g, dirs, texel_dirs, rad,
                     reinterpret_cast<const float4*>(atlas_new), reinterpret_cast<const float4*>(atlas_old),
                     h, floor_, probe_base, out_max, out_sumsq);
}

}  // namespace ane
