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
// atlas texels of each probe.  Operands FP32 (v_mfma_f32_16x16x4_f32, the plain
// drop-in path), BF16 / FP16 (v_mfma_f32_16x16x32_bf16 / _f16) or FP8-e4m3
// (v_mfma_f32_16x16x32_fp8_fp8), FP32 accumulate; FP8 carries a per-texel weight
// scale and an FP32 radiance scale per probe, channel and 32 rays.
//   k_conv_stream  a whole update: 64 probes per block, the weights in LDS, each
//                  wave streams its 16 probes over all rays
//   k_conv_mfma    16 probes per block, 4 waves splitting the rays: weights too
//                  large for LDS (FP32 beyond 640 rays, 16-bit beyond 1280)
//   k_conv_jobs    the same blocks over the updates of a serving batch
//   k_conv_fp32    FP32 register-tiled kernel for ray counts that are not a
//                  multiple of 32
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

// ---------------------------------------------------------------------------
// Streaming MFMA convolution of a whole update.  Block = 4 waves x 16 probes.
// The weights of every ray are staged once per block in LDS, in fragment order:
// entry (s, rt, lane) holds the 8 k values A[rt*16 + lane%16][32 s + 8 (lane/16) ..]
// that lane contracts in step s, so a wave's fragment read is 64 consecutive
// lanes of LDS.  Each wave then contracts its own 16 probes over all rays (no
// reduction across waves).  Its radiance is a ring of D steps: step s + D is
// loaded while step s multiplies.  The loop body has no branches: D divides the
// step count and the last D steps are peeled off (they load no radiance), so
// the ring stays in the same registers from one iteration to the next.
// Memory operations complete in issue order for the wait counters, so the
// prologue issues the radiance ring, then the weights; writing the weights to
// LDS waits for both, and step 0 needs its ring anyway.  The atlas texels of the
// epilogue are loaded after that, during the contraction.  The epilogue blends
// straight from the accumulator layout: lane (col, kq) holds texels
// rt*16 + 4kq + i, i = 0..3, 4 consecutive texels (64 contiguous atlas bytes)
// of tile row 2rt + kq/2 of probe col; its per-texel factors come from LDS.
// ---------------------------------------------------------------------------
#ifndef ANE_CONV_STREAM
#define ANE_CONV_STREAM 1
#endif
// Radiance ring depth in 32-ray steps, per operand size (a sweep of 1, 2, 4 and 8:
// deeper rings hold more registers; these were the fastest on the production shape).
#ifndef ANE_CONV_DEPTH16
#define ANE_CONV_DEPTH16 2
#endif
#ifndef ANE_CONV_DEPTH8
#define ANE_CONV_DEPTH8 4
#endif
#ifndef ANE_CONV_DEPTH32
#define ANE_CONV_DEPTH32 4
#endif
static constexpr int kStreamProbes = 64;
static constexpr int kStreamLdsMax = 160 * 1024;
static constexpr int kStreamFactorBytes = 2 * 64 * 4;  // per-texel scale and weight sum, after the weights
static constexpr int kStreamStage = 16;                // weight entries per thread in flight at once

// One block's 64 probes (see above); blk is the block's index within its update.
template <int PREC, int D>
__device__ __forceinline__ void conv_stream_block(const GridK& g, const void* rad, const float* rad_scale,
                                                  const void* wq, const float* wscale, const float* wsum,
                                                  float4* atlas, int nprobes, int probe_base, int32_t bake,
                                                  int blk)
{
  using F = ConvFrag<PREC>;
  constexpr int EB = F::EB, FB = 8 * EB;
  extern __shared__ __align__(16) uint8_t wl[];
  const int n = g.rays, S = n / 32;  // S % D == 0
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int pw = blk * kStreamProbes + wave * 16;  // the wave's first probe
  const bool live = pw + col < nprobes;
  // lanes past the last probe read the last probe: their columns are never stored
  const int pc = live ? pw + col : nprobes - 1;
  const uint8_t* radp = (const uint8_t*)rad + ((int64_t)pc * 3 * n + 8 * kq) * EB;
  const float* rsp = (PREC == ANE_PRECISION_FP8) ? rad_scale + (int64_t)pc * 3 * S : nullptr;

  F b[D][3];
  float bs[D][3];
  auto load_b = [&](int s, int j) {
#pragma unroll
    for (int c = 0; c < 3; ++c) b[j][c].load(radp + ((int64_t)c * n + 32 * s) * EB);
    if constexpr (PREC == ANE_PRECISION_FP8) {
#pragma unroll
      for (int c = 0; c < 3; ++c) bs[j][c] = rsp[c * S + s];
    }
  };
#pragma unroll
  for (int j = 0; j < D; ++j) load_b(j, j);

  // weights to LDS in fragment order, kStreamStage entries per thread in flight.
  // Entries past the last step load and store the last step's entry of the same
  // (row tile, lane) again: identical bytes, so no branch is needed.
  const uint8_t* wp = (const uint8_t*)wq;
  float* fac = reinterpret_cast<float*>(wl + (size_t)64 * n * EB);
  for (int m0 = 0; m0 < S; m0 += kStreamStage) {
    uint4 x[kStreamStage][FB == 32 ? 2 : 1];
#pragma unroll
    for (int m = 0; m < kStreamStage; ++m) {
      const int s = min(m0 + m, S - 1), l = tid & 63, rt = tid >> 6;
      const uint8_t* src = wp + ((int64_t)(rt * 16 + (l & 15)) * n + 32 * s + 8 * (l >> 4)) * EB;
      if constexpr (FB == 32) {
        x[m][0] = reinterpret_cast<const uint4*>(src)[0];
        x[m][1] = reinterpret_cast<const uint4*>(src)[1];
      } else if constexpr (FB == 16) {
        x[m][0] = *reinterpret_cast<const uint4*>(src);
      } else {
        const uint2 y = *reinterpret_cast<const uint2*>(src);
        x[m][0] = make_uint4(y.x, y.y, 0u, 0u);
      }
    }
    // the per-texel factors (the weight-sum correction or FP8 weight scale, and the weight sum)
    const int tf = tid & 63;
    const float f_sc = (PREC == ANE_PRECISION_FP8) ? wscale[tf]
                       : (PREC == ANE_PRECISION_BF16 || PREC == ANE_PRECISION_FP16) ? wscale[64 + tf] : 1.0f;
    const float f_ws = wsum[tf];
#pragma unroll
    for (int m = 0; m < kStreamStage; ++m) {
      const size_t e = (size_t)min(m0 + m, S - 1) * 256 + tid;
      if constexpr (FB == 32) {
        reinterpret_cast<uint4*>(wl + e * FB)[0] = x[m][0];
        reinterpret_cast<uint4*>(wl + e * FB)[1] = x[m][1];
      } else if constexpr (FB == 16) {
        *reinterpret_cast<uint4*>(wl + e * FB) = x[m][0];
      } else {
        *reinterpret_cast<uint2*>(wl + e * FB) = make_uint2(x[m][0].x, x[m][0].y);
      }
    }
    if (m0 == 0 && tid < 64) {
      fac[tid] = f_sc;
      fac[64 + tid] = f_ws;
    }
  }
  __syncthreads();
  if (pw >= nprobes) return;  // no barrier follows

  // atlas texels of the epilogue, in flight during the contraction
  const int q = probe_base + pc;
  const int px = q % g.nx, pz = (q / g.nx) % g.nz, py = q / (g.nx * g.nz);
  const int64_t tile = ((int64_t)10 * ((int64_t)py * g.nz + pz) + 1) * g.atlas_w + 10 * px + 1 + 4 * (kq & 1);
  float4 old[4][4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const int64_t row = tile + (int64_t)(2 * rt + (kq >> 1)) * g.atlas_w;
#pragma unroll
    for (int i = 0; i < 4; ++i) old[rt][i] = atlas[row + i];
  }

  f32x4 acc[4][3];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[rt][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  auto load_a = [&](int s, F (&a)[4]) {
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) a[rt].load(wl + ((size_t)(s * 4 + rt) * 64 + lane) * FB);
  };
  // step j of the ring: contract the step's radiance b[j] with the weights a
  auto contract = [&](const F (&a)[4], int j) {
    if constexpr (PREC == ANE_PRECISION_FP8) {
      // each 32-ray step carries its own radiance scale: all 12 products first,
      // then they are scaled into the accumulators
      f32x4 t[4][3];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 3; ++c) t[rt][c] = F::mma(a[rt], b[j][c], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[rt][c] += t[rt][c] * bs[j][c];
    } else {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[rt][c] = F::mma(a[rt], b[j][c], acc[rt][c]);
    }
  };
  // the weights of step s + 1 are read from LDS while step s multiplies; the
  // last D steps load no more radiance, so nothing is in flight at the epilogue
  F a[4];
  load_a(0, a);
  for (int s0 = 0; s0 < S - D; s0 += D) {
#pragma unroll
    for (int j = 0; j < D; ++j) {
      F an[4];
      load_a(s0 + j + 1, an);
      contract(a, j);
      load_b(s0 + j + D, j);
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) a[rt] = an[rt];
    }
  }
#pragma unroll
  for (int j = 0; j < D; ++j) {
    F an[4];
    if (j + 1 < D) load_a(S - D + j + 1, an);
    contract(a, j);
    if (j + 1 < D) {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) a[rt] = an[rt];
    }
  }
  if (!live) return;
  const float h = g.hysteresis;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const int64_t row = tile + (int64_t)(2 * rt + (kq >> 1)) * g.atlas_w;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = rt * 16 + 4 * kq + i;
      const float sc = fac[t];
      atlas[row + i] = blend_texel(old[rt][i], acc[rt][0][i] * sc, acc[rt][1][i] * sc, acc[rt][2][i] * sc,
                                   fac[64 + t], h, bake != 0);
    }
  }
}

template <int PREC, int D>
__global__ __launch_bounds__(256) void k_conv_stream(GridK g, const void* rad, const float* rad_scale,
                                                     const void* wq, const float* wscale, const float* wsum,
                                                     float4* atlas, int nprobes, int probe_base, int32_t bake)
{
  conv_stream_block<PREC, D>(g, rad, rad_scale, wq, wscale, wsum, atlas, nprobes, probe_base, bake, blockIdx.x);
}

// The streaming convolutions of many updates in one launch (the request loop's
// fused stage): block b runs block b - block_begin of the last job whose
// block_begin <= b; the dynamic LDS fits the jobs' largest weight table.
template <int PREC, int D>
__global__ __launch_bounds__(256) void k_conv_stream_jobs(const ConvJob* jobs, int njobs)
{
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  const ConvJob& j = jobs[lo];
  conv_stream_block<PREC, D>(j.g, j.rad, j.rad_scale, j.wq, j.wscale, j.wsum, j.atlas, j.nprobes, j.probe_base,
                             j.bake, b - j.block_begin);
}

static int conv_operand_bytes(int precision)
{
  return precision == ANE_PRECISION_FP32 ? 4 : precision == ANE_PRECISION_FP8 ? 1 : 2;
}

static size_t conv_stream_lds(int precision, int rays)
{
  return (size_t)64 * rays * conv_operand_bytes(precision) + kStreamFactorBytes;
}

bool conv_stream_fits(int precision, int rays)
{
  return ANE_CONV_STREAM && rays > 0 && rays % 32 == 0 && conv_stream_lds(precision, rays) <= (size_t)kStreamLdsMax;
}

// The ring depth: the largest power of two up to the precision's depth that
// divides the step count S (all of them, for a job launch).  The results do not
// depend on it: every step is accumulated in order whatever the depth.
template <int PREC, int D>
static void launch_stream_d(int S, bool jobs, const GridK& g, const void* rad, const float* rad_scale,
                            const void* wq, const float* wscale, const float* wsum, float4* atlas, int num_probes,
                            int probe_base, bool bake, const ConvJob* job_table, int njobs, int blocks, size_t lds,
                            hipStream_t s)
{
  if constexpr (D > 1) {
    if (S % D != 0) {
      launch_stream_d<PREC, D / 2>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes, probe_base,
                                   bake, job_table, njobs, blocks, lds, s);
      return;
    }
  }
  static bool lds_set = false;
  if (!lds_set) {
    (void)hipFuncSetAttribute((const void*)k_conv_stream<PREC, D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              kStreamLdsMax);
    (void)hipFuncSetAttribute((const void*)k_conv_stream_jobs<PREC, D>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, kStreamLdsMax);
    lds_set = true;
  }
  if (jobs) {
    hipLaunchKernelGGL((k_conv_stream_jobs<PREC, D>), dim3((unsigned)blocks), dim3(256), lds, s, job_table, njobs);
  } else {
    hipLaunchKernelGGL((k_conv_stream<PREC, D>), dim3((unsigned)blocks), dim3(256), lds, s, g, rad, rad_scale, wq,
                       wscale, wsum, atlas, num_probes, probe_base, (int32_t)bake);
  }
}

template <int PREC>
static void launch_conv_stream(int S, bool jobs, const GridK& g, const void* rad, const float* rad_scale,
                               const void* wq, const float* wscale, const float* wsum, float4* atlas,
                               int num_probes, int probe_base, bool bake, const ConvJob* job_table, int njobs,
                               int blocks, size_t lds, hipStream_t s)
{
  constexpr int D = PREC == ANE_PRECISION_FP32 ? ANE_CONV_DEPTH32
                    : PREC == ANE_PRECISION_FP8 ? ANE_CONV_DEPTH8 : ANE_CONV_DEPTH16;
  launch_stream_d<PREC, D>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes, probe_base, bake,
                           job_table, njobs, blocks, lds, s);
}

static void launch_conv_stream_any(int precision, int S, bool jobs, const GridK& g, const void* rad,
                                   const float* rad_scale, const void* wq, const float* wscale, const float* wsum,
                                   float4* atlas, int num_probes, int probe_base, bool bake,
                                   const ConvJob* job_table, int njobs, int blocks, size_t lds, hipStream_t s)
{
  if (precision == ANE_PRECISION_FP32)
    launch_conv_stream<ANE_PRECISION_FP32>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes,
                                           probe_base, bake, job_table, njobs, blocks, lds, s);
  else if (precision == ANE_PRECISION_FP16)
    launch_conv_stream<ANE_PRECISION_FP16>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes,
                                           probe_base, bake, job_table, njobs, blocks, lds, s);
  else if (precision == ANE_PRECISION_BF16)
    launch_conv_stream<ANE_PRECISION_BF16>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes,
                                           probe_base, bake, job_table, njobs, blocks, lds, s);
  else
    launch_conv_stream<ANE_PRECISION_FP8>(S, jobs, g, rad, rad_scale, wq, wscale, wsum, atlas, num_probes,
                                          probe_base, bake, job_table, njobs, blocks, lds, s);
}

int conv_job_blocks(int precision, int rays, int num_probes)
{
  const int per = conv_stream_fits(precision, rays) ? kStreamProbes : kConvProbes;
  return (num_probes + per - 1) / per;
}

void launch_conv_jobs(int precision, const ConvJob* jobs, int njobs, int total_blocks, bool stream, int max_rays,
                      int rays_gcd, hipStream_t s)
{
  if (njobs <= 0 || total_blocks <= 0) return;
  if (stream) {
    launch_conv_stream_any(precision, rays_gcd / 32, true, GridK{}, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, 0, 0, false, jobs, njobs, total_blocks, conv_stream_lds(precision, max_rays), s);
    return;
  }
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
  if (conv_stream_fits(precision, g.rays)) {
    const void* wq = precision == ANE_PRECISION_FP32 ? (const void*)w32
                     : precision == ANE_PRECISION_FP8 ? (const void*)wfp8 : (const void*)wbf16;
    const int blocks = (num_probes + kStreamProbes - 1) / kStreamProbes;
    launch_conv_stream_any(precision, g.rays / 32, false, g, rad, rad_scale, wq, wfp8_scale, wsum, a4, num_probes,
                           probe_base, bake, nullptr, 0, blocks, conv_stream_lds(precision, g.rays), s);
    return;
  }
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
