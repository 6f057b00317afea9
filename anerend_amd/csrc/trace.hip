// trace.hip -- probe ray traversal and hit shading on gfx950.
//
// Replaces vkCmdTraceRaysKHR of IrradianceProbeRT (IrradianceProbeRayTracingPass.cpp:95-165)
// and its shaders irradiance_probe_update.{rgen,rchit,rmiss}:
//   k_frame_setup    rotated Fibonacci directions (rgen:35-46,86-88), their ray setup
//                    (reciprocals, watertight shear), the sun ray setup and the 64 x N
//                    cosine-lobe weight matrix of the convolution (probe_conv.comp:37-61)
//   k_trace          persistent wave-wide BVH traversal, one ray per lane: each wave
//                    claims chunks of consecutive (direction-coherent) rays and refills
//                    finished lanes by ballot between traversal steps; short stack in
//                    LDS with a global spill area; leaves tested in wave-wide batches.
//                    A probe ray's closest hit (rgen:90-106) is followed in the same
//                    lane by its any-hit shadow ray from the hit position (rchit:30-66)
//                    through the hierarchy, unless the context has a current sun grid
//                    (sungrid.hip).  Q = true reads the scene's quantized 32-byte nodes
//                    (large scenes), else the 64-byte nodes; caller rays may carry a
//                    per-ray t_max
//   k_probe_shadow_grid  with a sun grid: the shadow rays of the probe hits against
//                    the triangles of their cells, the candidates of a wave's 64 rays
//                    spread over its lanes
//   k_shade          one block per probe: attribute fetch, Cook-Torrance direct light,
//                    indirect light from the previous atlas, emissive, sky on miss
//                    (rchit:24-110, rmiss:17-21, pbr_light.glsl); radiance stored in the
//                    context precision (FP8: a scale per probe, channel and 32 rays)
//   k_gi_lookup      the probe term of deferred_tiled.comp:264-298 per G-buffer texel
//   k_shadow_*       RT hard shadows of a depth buffer (shadow.rgen), through the
//                    scene's sun grid (k_shadow_mask_grid) or caller-ray traversal
//   k_mirror_*       mirror reflections (mirror_reflections.rgen/.rchit/.rmiss)
//   k_light_*        point-light shadow moments (light_shadow.rgen/.rchit/.rmiss and
//                    light_shadow_sum.comp)
#include "common.hpp"
#include "kernels.hpp"

namespace ane {

#ifndef ANE_TRACE_BLOCK
#define ANE_TRACE_BLOCK 128
#endif
#ifndef ANE_TRACE_STACK
#define ANE_TRACE_STACK 16
#endif
#ifndef ANE_TRACE_CHUNK
#define ANE_TRACE_CHUNK 256  // probe rays claimed by a wave at a time (64-byte nodes)
#endif
#ifndef ANE_TRACE_CHUNK_RAYS
#define ANE_TRACE_CHUNK_RAYS 128  // caller rays claimed by a wave at a time (64-byte nodes)
#endif
#ifndef ANE_TRACE_REFILL
#define ANE_TRACE_REFILL 32
#endif
#ifndef ANE_TRACE_UNROLL
#define ANE_TRACE_UNROLL 5  // traversal steps per refill check
#endif
#ifndef ANE_TRACE_WAVES
#define ANE_TRACE_WAVES 7  // waves per SIMD the traversal is register-limited to
#endif
#ifndef ANE_TRACE_LEAFBATCH
#define ANE_TRACE_LEAFBATCH 8  // waiting leaves that trigger a wave-wide leaf batch (0: no batching)
#endif
#ifndef ANE_SHADE_WAVES
#define ANE_SHADE_WAVES 1  // waves per SIMD the shading kernels are register-limited to (1: compiler's choice)
#endif
#ifndef ANE_SHADE_TB
#define ANE_SHADE_TB 512  // threads per shading block for probes of 512 rays or more (256: 0.357 against 0.348 ms)
#endif
#ifndef ANE_SHADE_UNROLL
#define ANE_SHADE_UNROLL 1  // probe corners of sampleProbe unrolled together (register use vs loads in flight)
#endif
#ifndef ANE_TRACE_LEAFBATCH_RAYS
#define ANE_TRACE_LEAFBATCH_RAYS 0  // the same for caller rays
#endif
#ifndef ANE_TRACE_SPEC
#define ANE_TRACE_SPEC 0  // 1: probe rays keep visiting internal nodes while a leaf waits
#endif
static constexpr int kTB = ANE_TRACE_BLOCK;      // threads per traversal block
static constexpr int kSS = ANE_TRACE_STACK;      // LDS stack entries per lane (ring, any size)
// Every claim is one device-scope atomic on the launch's counter, which all 8 XCDs
// share: with 128-ray chunks the 8.4M probe rays of an update made 65k claims and the
// trace waited on them (terrain1m trace 1.028 ms; 256: 0.895, 384: 0.906, 512: 0.979,
// 64: 1.69).  Caller rays (1-2M per pass) keep 128: larger chunks unbalance their
// launch's end (mirror pass 0.78 -> 0.87 ms at 256).
static constexpr int kChunk = ANE_TRACE_CHUNK;
static constexpr int kChunkRays = ANE_TRACE_CHUNK_RAYS;
static constexpr int kRefill = ANE_TRACE_REFILL; // idle lanes that trigger a refill
// The quantized-node traversal serves scenes whose nodes miss the caches: its
// lanes wait on memory rather than on VALU issue, so it refills earlier and claims
// larger chunks (4.04M-triangle scene: trace 10.0 -> 9.0 ms against 32 / 128), and
// checks for refills every 7 steps (9.09 -> 8.98 ms against 5); chunks of 384 / 512
// measured 8.98 / 9.60 ms against 8.84
#ifndef ANE_TRACE_CHUNK_Q
#define ANE_TRACE_CHUNK_Q 256
#endif
#ifndef ANE_TRACE_REFILL_Q
#define ANE_TRACE_REFILL_Q 16
#endif
#ifndef ANE_TRACE_UNROLL_Q
#define ANE_TRACE_UNROLL_Q 7  // steps per refill check (quantized nodes)
#endif
#ifndef ANE_TRACE_LEAFBATCH_Q
#define ANE_TRACE_LEAFBATCH_Q ANE_TRACE_LEAFBATCH
#endif
static constexpr int kChunkQ = ANE_TRACE_CHUNK_Q;
static constexpr int kRefillQ = ANE_TRACE_REFILL_Q;
static constexpr int kLeafBatchProbe = ANE_TRACE_LEAFBATCH;
static constexpr int kLeafBatchRays = ANE_TRACE_LEAFBATCH_RAYS;
static constexpr int kShadeUnroll = ANE_SHADE_UNROLL;

bool trace_grid_supported() { return true; }

int trace_stack_lds() { return kSS; }
int trace_block_size() { return kTB; }

// Direction-only part of the ray setup (RayPre without the origin), 48 bytes:
//   a = (d.x, d.y, d.z, inv.x)  b = (inv.y, inv.z, sx, sy)  c = (sz, kx | ky << 2 | kz << 4, 0, 0)
__device__ __forceinline__ void store_dirpre(float4* out, const RayPre& r)
{
  out[0] = make_float4(r.d.x, r.d.y, r.d.z, r.inv.x);
  out[1] = make_float4(r.inv.y, r.inv.z, r.sx, r.sy);
  out[2] = make_float4(r.sz, __int_as_float(r.kx | (r.ky << 2) | (r.kz << 4)), 0.0f, 0.0f);
}

__device__ __forceinline__ void load_dirpre(RayPre& r, const float4* p, f3 o)
{
  const float4 a = p[0], b = p[1], c = p[2];
  r.o = o;
  r.d = mk(a.x, a.y, a.z);
  r.inv = mk(a.w, b.x, b.y);
  r.sx = b.z;
  r.sy = b.w;
  r.sz = c.x;
  const int k = __float_as_int(c.y);
  r.kx = k & 3;
  r.ky = (k >> 2) & 3;
  r.kz = (k >> 4) & 3;
}

// ---------------------------------------------------------------------------
// frame setup: rotated directions, their ray setup, the sun ray setup and the
// convolution weights (one block per texel)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frame_setup(GridK g, FrameK f, Mat3 rot, f3 sun,
                                                     const float* base, const uint32_t* order,
                                                     float* dirs, float4* dirpre,
                                                     float4* probe_org, float* w32, float* wsum,
                                                     uint16_t* wbf16, uint8_t* wfp8,
                                                     float* wfp8_scale, int32_t bake, int32_t f16)
{
  const int t = blockIdx.x;  // texel 0..63
  // probe origins of the whole grid (rgen:60-84), read by ray generation and shading
  const int P = g.nx * g.ny * g.nz;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < P; q += gridDim.x * blockDim.x) {
    const int px = q % g.nx, pz = (q / g.nx) % g.nz, py = q / (g.nx * g.nz);
    const f3 o = probe_origin(g, f, px, py, pz);
    probe_org[q] = make_float4(o.x, o.y, o.z, 0.0f);
  }
  const int n = g.rays;
  const f3 o = texel_dir(t);
  __shared__ float red[256];
  float s = 0.0f, mx = 0.0f;
  if (t == 0) {
    // ray setup in processing order: entry j belongs to ray order[j]
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      RayPre rp;
      ray_setup(rp, mk(0.0f, 0.0f, 0.0f), rot3(rot.m, ld3(base + 3 * order[j])));
      store_dirpre(dirpre + 3 * j, rp);
    }
  }
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    f3 d = rot3(rot.m, ld3(base + 3 * r));
    if (t == 0) {
      dirs[3 * r + 0] = d.x;
      dirs[3 * r + 1] = d.y;
      dirs[3 * r + 2] = d.z;
    }
    float w = fmaxf(0.0f, dot(d, o));
    if (!(w > 1e-6f)) w = 0.0f;
    w32[(int64_t)t * n + r] = w;
    s += w;
  }
  if (t == 1 && threadIdx.x == 0) {
    RayPre rp;
    ray_setup(rp, mk(0.0f, 0.0f, 0.0f), sun);
    store_dirpre(dirpre + 3 * n, rp);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  const float sum = red[0];
  __syncthreads();
  // normalized weights (E = sum W L / sum W); bake mode keeps them raw
  const bool norm = !bake && sum > 1e-6f;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    float w = w32[(int64_t)t * n + r];
    if (norm) w = w / sum;
    w32[(int64_t)t * n + r] = w;
    // the 16-bit weights: BF16, or FP16 for the FP16 precision
    wbf16[(int64_t)t * n + r] = f16 ? f32_to_f16_sat(w) : f32_to_bf16(w);
    mx = fmaxf(mx, w);
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  const float wmax = red[0];
  __syncthreads();
  const float sc = wmax > 0.0f ? wmax / 448.0f : 1.0f;
  // quantized weights and the sums that undo their normalization error: the
  // epilogue scales texel t by (sum of FP32 weights) / (sum of quantized weights)
  float s32 = 0.0f, sb = 0.0f, s8 = 0.0f;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const float w = w32[(int64_t)t * n + r];
    const uint8_t q = f32_to_e4m3(w / sc);
    wfp8[(int64_t)t * n + r] = q;
    s32 += w;
    sb += f16 ? f16_to_f32(wbf16[(int64_t)t * n + r]) : bf16_to_f32(wbf16[(int64_t)t * n + r]);
    s8 += e4m3_to_f32(q) * sc;
  }
  __shared__ float red3[3][256];
  red3[0][threadIdx.x] = s32;
  red3[1][threadIdx.x] = sb;
  red3[2][threadIdx.x] = s8;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k)
      for (int c = 0; c < 3; ++c) red3[c][threadIdx.x] += red3[c][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    wsum[t] = sum;
    const float S = red3[0][0], SB = red3[1][0], S8 = red3[2][0];
    wfp8_scale[t] = S8 > 0.0f ? sc * (S / S8) : sc;  // FP8: dequantization x correction
    wfp8_scale[64 + t] = SB > 0.0f ? S / SB : 1.0f;  // BF16 / FP16: correction
  }
}

void launch_frame_setup(const GridK& g, const FrameK& f, const float* base_dirs,
                        const uint32_t* order, float* dirs, float4* dirpre, float4* probe_org, float* w32, float* wsum, uint16_t* wbf16,
                        uint8_t* wfp8, float* wfp8_scale, bool bake, bool f16, hipStream_t s)
{
  Mat3 rot;
  for (int i = 0; i < 16; ++i) rot.m[i] = f.rot[i];
  const f3 sun = normalize(mk(-f.light_dir[0], -f.light_dir[1], -f.light_dir[2]));
  hipLaunchKernelGGL(k_frame_setup, dim3(64), dim3(256), 0, s, g, f, rot, sun, base_dirs, order,
                     dirs, dirpre, probe_org, w32, wsum, wbf16, wfp8, wfp8_scale, (int32_t)bake, (int32_t)f16);
}

// Division of a ray index x < nrays by n as (x * m) >> k, m = ceil(2^k / n).
// With e = m n - 2^k < n, x m / 2^k = x / n + x e / (n 2^k) and the floor is
// exact when x e / 2^k < 1, which k = bit_length(nrays * n) guarantees
// (x e < nrays n <= 2^k).  No overflow: x m < nrays (2^k / n + 1)
// <= 2 nrays^2 + nrays < 2^63 for nrays < 2^31.
RayDiv ray_div_magic(int64_t nrays, int n)
{
  const uint64_t bound = (uint64_t)nrays * (uint64_t)n;
  int k = 0;
  while (k < 63 && (1ull << k) <= bound) ++k;
  RayDiv d;
  d.shift = k;
  d.magic = ((1ull << k) + (uint64_t)n - 1) / (uint64_t)n;
  return d;
}

// ---------------------------------------------------------------------------
// hit position (rchit:30-42): object-space barycentric interpolation, then the
// instance transform.  Bit-exact with the oracle (it is the shadow-ray origin).
// ---------------------------------------------------------------------------
struct HitGeom {
  uint32_t inst;
  uint32_t v[3];  // absolute vertex indices
};

// Instance, absolute vertex indices (rchit:27-38) and world position of a hit
// on triangle id, from the triangle's 64-byte record written by k_flatten.
__device__ __forceinline__ f3 hit_geom_pos(const SceneK& sc, int32_t id, float u, float v,
                                           HitGeom& h)
{
  const float4* r = sc.tri_rec + 4 * (int64_t)id;
  const float4 a = r[0], b = r[1], c = r[2], d = r[3];
  h.inst = __float_as_uint(d.x);
  h.v[0] = __float_as_uint(a.w);
  h.v[1] = __float_as_uint(b.w);
  h.v[2] = __float_as_uint(c.w);
  const f3 bw = mk(1.0f - u - v, u, v);
  const f3 pos = add(add(muls(mk(a.x, a.y, a.z), bw.x), muls(mk(b.x, b.y, b.z), bw.y)),
                     muls(mk(c.x, c.y, c.z), bw.z));
  return xform_point(sc.inst[h.inst].transform, pos);
}

// ---------------------------------------------------------------------------
// traversal
//   RAYS = false: probe rays.  A lane traces its primary ray (closest hit,
//     rgen:90-106); when it hits, the same lane continues at once with the
//     shadow ray from the hit position toward the sun (any hit, rchit:44-66),
//     except with a sun grid in use: then k_probe_shadow_grid runs them next.
//   RAYS = true: caller rays, closest hit or (ANY) any hit.
// ---------------------------------------------------------------------------
template <bool ANY, bool RAYS, bool COUNT, bool Q>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(ANE_TRACE_WAVES))) void k_trace(TraceArgs a)
{
  constexpr int32_t kDone = 0x7FFFFFFF;  // node sentinel: the lane's traversal has ended
  constexpr int32_t kPop = 0x7FFFFFFE;   // step result: continue from the stack
  constexpr int32_t kWait = 0x7FFFFFFD;  // stack empty, the last leaf still waits for its test
  // (internal nodes are 0 .. ntri - 2 < kWait; leaves are ~position < 0)
  // Leaf batching: a lane that reaches a leaf parks it in `pend`, pops its next
  // node and waits; the triangle tests of all parked leaves run together once
  // kLeafBatch lanes wait or no lane can go on, so that a wave runs the leaf
  // code once per batch instead of in nearly every step.  The parked leaf is
  // tested before the popped node is visited, so visits, hits and counts are
  // those of the one-step-at-a-time traversal.  SPEC lanes instead keep
  // visiting internal nodes with the older best distance (more visits, the same
  // closest hit).
  constexpr bool SPEC = ANE_TRACE_SPEC && !RAYS && !COUNT;
  constexpr int kLeafBatch = RAYS ? kLeafBatchRays : (Q ? ANE_TRACE_LEAFBATCH_Q : kLeafBatchProbe);
  constexpr int kUnroll = Q ? ANE_TRACE_UNROLL_Q : ANE_TRACE_UNROLL;
  // with a sun grid in use, the probe shadow rays are k_probe_shadow_grid's
  // (launched after this kernel): the lanes here end at the primary hit
  const bool sgrid = !RAYS && a.sg.list != nullptr && a.sg.params->valid != 0u;
  __shared__ int32_t stk[kSS][kTB];
  if (RAYS && a.skip_if_set && *a.skip_if_set != 0u) return;  // a sun grid answered these rays
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int gl = blockIdx.x * kTB + tid;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const SceneK& sc = a.sc;
  const float tmin = RAYS ? a.t_min : a.g.t_min;
  const float tmax0 = RAYS ? a.t_max : a.g.t_max;
  const int n = a.g.rays;
  const uint32_t nrays = (uint32_t)a.nrays;
  const int32_t root = (sc.ntri >= 2) ? 0 : (sc.ntri == 1 ? ~0 : kDone);
  // a tree traversal visits < 2 * ntri nodes; a caller's step limit ends rays earlier (as faults)
  const uint32_t budget = a.step_limit ? a.step_limit : 2u * (uint32_t)sc.ntri + 8u;

  int32_t ray = -1;  // active ray (processing index), -1 idle
  bool exhausted = false;
  bool sph = ANY;    // the lane's ray is an any-hit (shadow) ray
  uint32_t cbeg = 0, cend = 0;  // wave-uniform: remaining part of the claimed chunk
  RayPre r;
  float tbest = 0.0f, bu = 0.0f, bv = 0.0f;
  int32_t bid = -1, node = kDone, sp = 0, out = 0;
  int32_t pend = 0;  // parked leaf (~position < 0), 0: none
  uint32_t steps_left = 0;
  unsigned long long cn = 0, ct = 0, sn = 0, st = 0;

  for (;;) {
    // ---- finish ended traversals; a probe ray that hit continues as its shadow ray ----
    if (ray >= 0 && node == kDone) {
      if (RAYS || !sph) {
        ane_hit h;
        h.tri = bid;
        h.t = (bid >= 0) ? tbest : 0.0f;
        h.u = (bid >= 0) ? bu : 0.0f;
        h.v = (bid >= 0) ? bv : 0.0f;
        a.hits[out] = h;
      }
      if (!RAYS && !sph && bid >= 0 && !sgrid) {
        // shadow ray from the hit position toward the sun (rchit:44-66), same lane
        HitGeom hg;
        const f3 p = hit_geom_pos(sc, bid, bu, bv, hg);
        load_dirpre(r, a.dirpre + 3 * n, p);
        sph = true;
        tbest = tmax0;
        bid = -1;
        sp = 0;
        steps_left = budget;
        node = root;
      } else {
        if (!RAYS && !sgrid) a.shadow[out] = sph ? (uint8_t)(bid >= 0 ? 1 : 0) : (uint8_t)255;
        ray = -1;
      }
    }
    // ---- refill idle lanes (between steps) ----
    const bool idle = (ray < 0) && !exhausted;
    const uint64_t im = __ballot(idle);
    const uint64_t am = __ballot(ray >= 0);
    const uint32_t need = (uint32_t)__popcll(im);
    constexpr uint32_t chunk = (uint32_t)(Q ? kChunkQ : (RAYS ? kChunkRays : kChunk));
    if (need != 0u && (am == 0ull || need >= (uint32_t)(Q ? kRefillQ : kRefill))) {
      if (cbeg >= cend) {  // claim the next chunk (one atomic per wave and chunk)
        uint32_t base = 0;
        if (lane == __ffsll((unsigned long long)im) - 1) base = atomicAdd(a.work, chunk);
        base = __shfl(base, __ffsll((unsigned long long)im) - 1, 64);
        cbeg = base;
        cend = base + chunk < nrays ? base + chunk : nrays;
      }
      const uint32_t avail = cend > cbeg ? cend - cbeg : 0u;
      const uint32_t rank = (uint32_t)__popcll(im & lt);
      if (idle) {
        if (avail == 0u) {
          exhausted = true;
        } else if (rank < avail) {
          const uint32_t gidx = cbeg + rank;
          bool none = false;
          if (RAYS) {
            out = (int32_t)gidx;
            const f3 d = ld3(a.ray_dir + 3 * (int64_t)gidx);
            // a zero or non-finite direction is no ray (a NaN direction would pass
            // every slab test and visit the whole hierarchy)
            none = (d.x == 0.0f && d.y == 0.0f && d.z == 0.0f) || !__builtin_isfinite(d.x) ||
                   !__builtin_isfinite(d.y) || !__builtin_isfinite(d.z);
            ray_setup(r, ld3(a.ray_orig + 3 * (int64_t)gidx), d);
          } else {
            // independent loads: processing-order ray setup, probe origin, output slot
            const uint32_t ql = (uint32_t)(((uint64_t)gidx * a.div_magic) >> a.div_shift);
            const uint32_t j = gidx - ql * (uint32_t)n;
            const float4 po = a.probe_org[a.probe_base + (int)ql];
            load_dirpre(r, a.dirpre + 3 * j, mk(po.x, po.y, po.z));
            out = (int32_t)ql * n + (int)a.order[j];
            sph = false;
          }
          tbest = tmax0;
          if (RAYS && a.ray_tmax) tbest = a.ray_tmax[gidx];
          bid = -1;
          sp = 0;
          steps_left = budget;
          ray = (int32_t)gidx;
          node = none ? kDone : root;
        }
      }
      cbeg += need < avail ? need : avail;
    }
    if (am == 0ull && __ballot(ray >= 0) == 0ull) {
      if (__ballot(!exhausted) == 0ull) break;
      continue;
    }
    // ---- traversal steps: only node, sp, the best hit and the budget change ----
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
    const bool go = ray >= 0 && node != kDone &&
                    (kLeafBatch == 0 || (node != kWait && (SPEC ? !(node < 0 && pend != 0) : pend == 0)));
    if (go) {
      int32_t next = kPop;
      // guards: a malformed hierarchy ends the ray and raises the error flag
      // instead of reading out of bounds or looping forever
      // internal node index < ntri - 1, leaf position ~node < ntri
      const bool isnode = node >= 0;
      const uint32_t idx = isnode ? (uint32_t)node : (uint32_t)~node;
      const uint32_t lim = isnode ? (uint32_t)sc.ntri - 1u : (uint32_t)sc.ntri;
      const bool bad = (steps_left == 0u) | (idx >= lim);
      --steps_left;
      if (bad) {
        *a.overflow = 1u;
        next = kDone;
        pend = 0;
      } else if (isnode) {
        if (COUNT) {
          if (sph) ++sn; else ++cn;
        }
        float4 na, nb, nc;
        int32_t L, R;
        if (Q) {
          // quantized node: decode the two child boxes (layout next to NodeBuild)
          const uint4* nq = sc.qnodes + 2 * (int64_t)node;
          const uint4 q0 = nq[0], q1 = nq[1];
          const float ox = __uint_as_float(q0.x), oy = __uint_as_float(q0.y), oz = __uint_as_float(q0.z);
          const float sx = qnode_step(q0.w & 0xFFu), sy = qnode_step((q0.w >> 8) & 0xFFu);
          const float sz = qnode_step((q0.w >> 16) & 0xFFu);
          na = make_float4(qnode_bound(ox, q1.x & 0xFFu, sx), qnode_bound(ox, q1.x >> 24, sx),
                           qnode_bound(oy, (q1.x >> 8) & 0xFFu, sy), qnode_bound(oy, q1.y & 0xFFu, sy));
          nb = make_float4(qnode_bound(ox, (q1.y >> 16) & 0xFFu, sx), qnode_bound(ox, (q1.z >> 8) & 0xFFu, sx),
                           qnode_bound(oy, q1.y >> 24, sy), qnode_bound(oy, (q1.z >> 16) & 0xFFu, sy));
          nc = make_float4(qnode_bound(oz, (q1.x >> 16) & 0xFFu, sz), qnode_bound(oz, (q1.y >> 8) & 0xFFu, sz),
                           qnode_bound(oz, q1.z & 0xFFu, sz), qnode_bound(oz, q1.z >> 24, sz));
          const int32_t g = (int32_t)q1.w;
          L = (q0.w >> 24) & 1u ? ~g : g;
          R = (q0.w >> 25) & 1u ? ~(g + 1) : g + 1;
        } else {
          const float4* nd = sc.nodes + 4 * (int64_t)node;
          na = nd[0];
          nb = nd[1];
          nc = nd[2];
          const float4 nn = nd[3];
          L = __float_as_int(nn.x);
          R = __float_as_int(nn.y);
        }
        float tn0, tn1;
        const bool b0 = box_test(r, na.x, na.y, na.z, na.w, nc.x, nc.y, tmin, tbest, tn0);
        const bool b1 = box_test(r, nb.x, nb.y, nb.z, nb.w, nc.z, nc.w, tmin, tbest, tn1);
        const bool sw = b0 && b1 && tn1 < tn0;
        next = (b0 || b1) ? ((b0 && !sw) ? L : R) : kPop;
        if (b0 && b1) {
          // push the far child (LDS ring; the oldest entry spills when full)
          const int32_t farc = sw ? L : R;
          const int slot = (int)((uint32_t)sp % (uint32_t)kSS);
          if (sp >= kSS) {
            const int si = sp - kSS;
            if (si < a.spill_depth) {
              a.spill[(int64_t)si * a.total_lanes + gl] = stk[slot][tid];
            } else {
              *a.overflow = 1u;
              next = kDone;
            }
          }
          stk[slot][tid] = farc;
          ++sp;
        }
      } else if (kLeafBatch > 0) {
        pend = node;  // the leaf waits for the next batch
      } else {
        if (COUNT) {
          if (sph) ++st; else ++ct;
        }
        const int64_t p = ~node;
        const float4 v0 = sc.tris[3 * p], v1 = sc.tris[3 * p + 1], v2 = sc.tris[3 * p + 2];
        float t, u, v;
        if (tri_test(r, mk(v0.x, v0.y, v0.z), mk(v1.x, v1.y, v1.z), mk(v2.x, v2.y, v2.z), tmin,
                     tbest, t, u, v)) {
          const int32_t id = __float_as_int(v0.w);
          if (sph || t < tbest || (t == tbest && id < bid)) {
            bid = id;
            tbest = t;
            bu = u;
            bv = v;
          }
          if (sph) next = kDone;  // any hit ends a shadow ray
        }
      }
      if (next == kPop) {
        if (sp == 0) {
          next = (kLeafBatch > 0 && pend != 0) ? kWait : kDone;
        } else {
          --sp;
          const int slot = (int)((uint32_t)sp % (uint32_t)kSS);
          next = stk[slot][tid];
          if (sp >= kSS) stk[slot][tid] = a.spill[(int64_t)(sp - kSS) * a.total_lanes + gl];
        }
      }
      node = next;
    }
    if (kLeafBatch > 0) {
      const uint64_t pm = __ballot(pend != 0);
      if (pm != 0ull) {
        const bool can = ray >= 0 && node != kDone && node != kWait &&
                         (SPEC ? !(node < 0 && pend != 0) : pend == 0);
        if (__popcll(pm) >= kLeafBatch || __ballot(can) == 0ull) {
          if (pend != 0) {
            if (COUNT) {
              if (sph) ++st; else ++ct;
            }
            const int64_t p = ~pend;
            const float4 v0 = sc.tris[3 * p], v1 = sc.tris[3 * p + 1], v2 = sc.tris[3 * p + 2];
            float t, u, v;
            bool ended = node == kWait;  // nothing left to visit
            if (tri_test(r, mk(v0.x, v0.y, v0.z), mk(v1.x, v1.y, v1.z), mk(v2.x, v2.y, v2.z), tmin,
                         tbest, t, u, v)) {
              const int32_t id = __float_as_int(v0.w);
              if (sph || t < tbest || (t == tbest && id < bid)) {
                bid = id;
                tbest = t;
                bu = u;
                bv = v;
              }
              if (sph) ended = true;  // any hit ends a shadow ray
            }
            if (ended) node = kDone;
            pend = 0;
          }
        }
      }
    }
    }
  }
  if (COUNT) {
    if (RAYS) {  // caller rays: one ray kind, two counters
      atomicAdd(&a.stats[0], cn + sn);
      atomicAdd(&a.stats[1], ct + st);
    } else {
      atomicAdd(&a.stats[0], cn);
      atomicAdd(&a.stats[1], ct);
      atomicAdd(&a.stats[2], sn);
      atomicAdd(&a.stats[3], st);
    }
  }
}

// ---------------------------------------------------------------------------
// Probe shadow rays through the sun grid (rchit:44-66), after the traversal of
// the primary rays: the shadow ray of every hit tests the triangles of its
// sun-grid cell in list order, and any hit shadows it.
// A wave takes 64 consecutive output slots and spreads the candidate tests of
// their rays over all its lanes: the candidates of the 64 rays (ray after ray,
// list order within a ray) are tested 64 at a time, so every lane runs a test
// instead of walking its own ray's list.  Per ray the earliest event in list
// order is kept in LDS (key 2 pos for a hit, 2 pos + 1 for a list entry out of
// range; 2 count + 1 for none), and the tests of candidates after a known event
// are skipped.
// The outcome is that of the in-order walk that stops at its first hit: the ray
// is shadowed iff its first event is a hit, after pos + 1 tests (pos without
// a hit).  Every shadow flag of the launch is written here (255: no hit).
// ---------------------------------------------------------------------------
#ifndef ANE_SHADOW_U
#define ANE_SHADOW_U 2  // 1, 2, 4 and 8 swept: 238, 230, 250 and 383 µs on terrain1m
#endif
static constexpr int kShadowU = ANE_SHADOW_U;  // candidates per lane and round

// The candidates of a wave's 64 rays spread over its lanes: lane l's ray (origin
// p, direction of r, list [b, b + c) of the sun grid) has its candidates tested
// with the watertight any-hit test, 64 at a time, ray after ray and in list order
// within a ray (kShadowU per lane and round, every load of a round issued before
// any test).  key[l] (this wave's 64 LDS slots, set to 2 c + 1 by the caller)
// ends as the earliest event of lane l's list in list order: 2 pos for a hit,
// 2 pos + 1 for an entry out of range when STOP_ON_BAD (which also raises
// *overflow; otherwise such entries are skipped).  Tests after a known event are
// skipped.  All lanes of the wave must call it (shuffles).
template <bool STOP_ON_BAD>
__device__ __forceinline__ void grid_wave_events(const SunGridK& sg, const float4* tris, int32_t ntri, RayPre r,
                                                 f3 p, uint32_t b, uint32_t c, float tmin, float tmax,
                                                 uint32_t* key, uint32_t* overflow)
{
  const int lane = threadIdx.x & 63;
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  const uint32_t excl = incl - c;
  constexpr int U = kShadowU;
  for (uint32_t t0 = 0; t0 < total; t0 += 64u * U) {
    uint32_t pos[U], e[U];
    int jj[U];
    f3 o[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t t = t0 + 64u * u + (uint32_t)lane;
      const uint32_t tc = t < total ? t : total - 1u;
      // the ray of candidate tc: the first lane whose inclusive count exceeds it
      int j = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const uint32_t v = __shfl(incl, j + step - 1, 64);
        if (v <= tc) j += step;
      }
      const uint32_t ej = __shfl(excl, j, 64), bj = __shfl(b, j, 64);
      o[u] = mk(__shfl(p.x, j, 64), __shfl(p.y, j, 64), __shfl(p.z, j, 64));
      jj[u] = t < total ? j : -1;
      pos[u] = tc - ej;
      e[u] = sg.list[bj + pos[u]];
    }
    float4 v0[U], v1[U], v2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t q = e[u] < (uint32_t)ntri ? (int64_t)e[u] : 0;
      v0[u] = tris[3 * q];
      v1[u] = tris[3 * q + 1];
      v2[u] = tris[3 * q + 2];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (jj[u] >= 0 && 2u * pos[u] < key[jj[u]]) {
        if (e[u] >= (uint32_t)ntri) {
          if (STOP_ON_BAD) {
            *overflow = 1u;
            atomicMin(&key[jj[u]], 2u * pos[u] + 1u);
          }
        } else {
          float tt, uu, vv;
          r.o = o[u];
          if (tri_test(r, mk(v0[u].x, v0[u].y, v0[u].z), mk(v1[u].x, v1[u].y, v1[u].z),
                       mk(v2[u].x, v2[u].y, v2[u].z), tmin, tmax, tt, uu, vv))
            atomicMin(&key[jj[u]], 2u * pos[u]);
        }
      }
    }
  }
}

#ifndef ANE_SHADOW_TB
#define ANE_SHADOW_TB 512  // threads (output slots) per block of the probe shadow kernel
#endif
static constexpr int kShadowTB = ANE_SHADOW_TB;

template <bool COUNT>
__device__ __forceinline__ void probe_shadow_group(const TraceArgs& a, int64_t first)
{
  __shared__ uint32_t key[kShadowTB];
  const SunGridParams* gq = a.sg.params;
  const int tid = threadIdx.x, lane = tid & 63, wb = tid & ~63;
  const int64_t i = first + tid;
  const bool in = i < (int64_t)a.nrays;
  const SceneK& sc = a.sc;
  ane_hit h;
  h.tri = -1;
  h.u = h.v = 0.0f;
  if (in) h = a.hits[i];
  const bool has = in && h.tri >= 0;
  f3 p = mk(0.0f, 0.0f, 0.0f);
  uint32_t b = 0u, c = 0u;
  if (has) {
    HitGeom hg;
    p = hit_geom_pos(sc, h.tri, h.u, h.v, hg);
    const float X = comp(p, gq->kx) - gq->sx * comp(p, gq->kz);
    const float Y = comp(p, gq->ky) - gq->sy * comp(p, gq->kz);
    const int64_t cell =
        (int64_t)sun_cell(Y, gq->y0, gq->ich, gq->gy) * gq->gx + sun_cell(X, gq->x0, gq->icw, gq->gx);
    b = a.sg.start[cell];
    c = a.sg.start[cell + 1] - b;
  }
  key[tid] = 2u * c + 1u;
  RayPre r;
  load_dirpre(r, a.dirpre + 3 * a.g.rays, p);  // the sun direction's ray setup
  grid_wave_events<true>(a.sg, sc.tris, sc.ntri, r, p, b, c, a.g.t_min, a.g.t_max, key + wb, a.overflow);
  const uint32_t k = key[tid];
  if (in) a.shadow[i] = has ? (uint8_t)((k & 1u) ? 0 : 1) : (uint8_t)255;
  if (COUNT) {
    // one cell lookup per hit; the tests of the in-order walk; one atomic per block
    __shared__ unsigned long long part[2][kShadowTB / 64];
    unsigned long long sn = has ? 1ull : 0ull;
    unsigned long long st = has ? (unsigned long long)((k & 1u) ? (k >> 1) : (k >> 1) + 1u) : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      sn += __shfl_xor(sn, d, 64);
      st += __shfl_xor(st, d, 64);
    }
    if (lane == 0) {
      part[0][tid >> 6] = sn;
      part[1][tid >> 6] = st;
    }
    __syncthreads();
    if (tid == 0) {
      unsigned long long t0 = 0ull, t1 = 0ull;
      for (int w = 0; w < kShadowTB / 64; ++w) {
        t0 += part[0][w];
        t1 += part[1][w];
      }
      atomicAdd(&a.stats[2], t0);
      atomicAdd(&a.stats[3], t1);
    }
  }
}

template <bool COUNT>
__global__ __launch_bounds__(kShadowTB) void k_probe_shadow_grid(TraceArgs a)
{
  const SunGridParams* gq = a.sg.params;
  if (gq->valid == 0u) return;  // no grid in use: the traversal ran the shadow rays
  probe_shadow_group<COUNT>(a, (int64_t)blockIdx.x * kShadowTB);
}

static int g_lanes = 0;

int trace_lanes_per_launch()
{
  if (g_lanes) return g_lanes;
  int dev = 0, cus = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trace<false, false, false, false>, kTB, 0);
  if (cus <= 0) cus = 256;
  if (per <= 0) per = 8;
  g_lanes = cus * per * kTB;
  return g_lanes;
}

void launch_trace(const TraceArgs& a, bool any_hit, bool rays, bool count, hipStream_t s)
{
  // a persistent grid of every resident block, or fewer when the caller caps the
  // lanes (a.total_lanes > 0: the spill area is sized for that many)
  const int full = trace_lanes_per_launch() / kTB;
  const int want = a.total_lanes > 0 ? a.total_lanes / kTB : full;
  const int blocks = want < 1 ? 1 : (want < full ? want : full);
  TraceArgs b = a;
  b.total_lanes = blocks * kTB;
#define ANE_TR(S, R, C)                                                                             \
  do {                                                                                              \
    if (b.sc.qnodes) hipLaunchKernelGGL((k_trace<S, R, C, true>), dim3(blocks), dim3(kTB), 0, s, b);  \
    else hipLaunchKernelGGL((k_trace<S, R, C, false>), dim3(blocks), dim3(kTB), 0, s, b);             \
  } while (0)
  if (rays) {
    if (any_hit) { if (count) ANE_TR(true, true, true); else ANE_TR(true, true, false); }
    else { if (count) ANE_TR(false, true, true); else ANE_TR(false, true, false); }
  } else {
    if (count) ANE_TR(false, false, true); else ANE_TR(false, false, false);
  }
#undef ANE_TR
}

void launch_probe_shadow_grid(const TraceArgs& a, bool count, hipStream_t s)
{
  if (!a.sg.list || a.nrays <= 0) return;
  // a block per 512 output slots (64 / 128 / 256 / 1024: 0.235 / 0.226 / 0.217 / 0.217 ms
  // against 0.202 on terrain1m; blocks looping over several groups measured 1.5-4% slower:
  // their waves' chains of dependent loads then run one group after another)
  const dim3 grid((unsigned)(((int64_t)a.nrays + kShadowTB - 1) / kShadowTB));
  if (count) hipLaunchKernelGGL(k_probe_shadow_grid<true>, grid, dim3(kShadowTB), 0, s, a);
  else hipLaunchKernelGGL(k_probe_shadow_grid<false>, grid, dim3(kShadowTB), 0, s, a);
}

// ---------------------------------------------------------------------------
// shading
// ---------------------------------------------------------------------------
static constexpr float kPi = 3.14159265359f;  // pbr_light.glsl:8

// Hardware reciprocal, reciprocal square root and square root (1 ulp) for the
// shading's continuous terms.  Radiance carries a relative tolerance; no branch
// of the shading depends on these values, and the hit positions (shadow-ray
// origins) keep correctly rounded arithmetic.  The view and half vectors of
// calcLight are NOT among them: GGX's 1 + NdotH^2 (a^2 - 1) cancels near the
// specular peak (a^2 = 0.0016 at roughness 0.2), which turns one ulp of NdotH
// into 1.5e-4 of radiance, so V and H keep the correctly rounded normalize.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// toLinear (bindless.glsl:240-247) with the power as hardware exp2(2.4 * log2 x)
// (about 1e-6 relative on the branch's range x >= 0.09): a continuous term of
// the hit color
__device__ __forceinline__ float to_linear_hw(float c)
{
  return (c < 0.04045f) ? c / 12.92f
                        : __builtin_amdgcn_exp2f(2.4f * __builtin_amdgcn_logf((c + 0.055f) / 1.055f));
}

__device__ __forceinline__ float distribution_ggx(f3 N, f3 H, float rough)
{
  float a = rough * rough;
  float a2 = a * a;
  float ndoth = fmaxf(dot(N, H), 0.0f);
  float ndoth2 = ndoth * ndoth;
  float num = a2 + 0.001f;
  float den = ndoth2 * (a2 - 1.0f) + 1.0f;
  den = kPi * den * den;
  return num * hw_rcp(den);
}

__device__ __forceinline__ float schlick_ggx(float ndotv, float rough)
{
  float r = rough + 1.0f;
  float k = (r * r) * 0.125f;  // / 8 exactly
  return ndotv * hw_rcp(ndotv * (1.0f - k) + k);
}

__device__ __forceinline__ float pow5_term(float c)
{
  const float x = clampf(1.0f - c, 0.0f, 1.0f);
  const float x2 = x * x;
  return x2 * x2 * x;
}

// calcLight, pbr_light.glsl:98-166 (directional, specular on); V = normalize(view - pos)
__device__ f3 calc_light(const FrameK& f, f3 N, f3 albedo, float metallic, float rough, f3 V)
{
  f3 light_color = muls(mk(1.0f, 1.0f, 1.0f), f.sun_intensity);
  f3 F0 = add(muls(mk(0.04f, 0.04f, 0.04f), 1.0f - metallic), muls(albedo, metallic));
  f3 L = mk(f.sun_l[0], f.sun_l[1], f.sun_l[2]);  // normalize(-lightDir), once per frame
  f3 H = normalize(add(V, L));  // correctly rounded: NdotH feeds GGX's cancellation
  f3 radiance = muls(light_color, 1.0f);
  float ndf = distribution_ggx(N, H, rough);
  float ndotv_ = fmaxf(dot(N, V), 0.0f);
  float ndotl_ = fmaxf(dot(N, L), 0.0f);
  float G = schlick_ggx(ndotl_, rough) * schlick_ggx(ndotv_, rough);
  float p5 = pow5_term(fmaxf(dot(H, V), 0.0f));
  f3 F = add(F0, muls(sub(mk(1.0f, 1.0f, 1.0f), F0), p5));
  f3 kD = muls(sub(mk(1.0f, 1.0f, 1.0f), F), 1.0f - metallic);
  float ndotl = fmaxf(dot(N, L), 0.0f);
  float ndotv = fmaxf(dot(N, V), 0.0f);
  f3 numer = muls(F, ndf * G);
  float denom = 4.0f * ndotv * ndotl + 0.0001f;
  f3 spec = muls(numer, hw_rcp(denom));
  spec = mk(clampf(spec.x, 0.0f, 1.0f), clampf(spec.y, 0.0f, 1.0f), clampf(spec.z, 0.0f, 1.0f));
  f3 lo = add(muls(mul(kD, albedo), 1.0f / kPi), spec);
  return muls(mul(lo, radiance), ndotl);
}

__device__ __forceinline__ int imod(int a, int m)
{
  int r = a % m;
  return r < 0 ? r + m : r;
}

// a mod m for REPEAT addressing.  Tile texel coordinates lie in [-1, 3m) (the
// tiles one past the grid edge wrap once), which a conditional add and two
// conditional subtractions reduce; anything else (a non-finite coordinate)
// takes the general modulo.
__device__ __forceinline__ int wrap_repeat(int a, int m)
{
  int r = a < 0 ? a + m : a;
  r = r >= m ? r - m : r;
  r = r >= m ? r - m : r;
  if (r < 0 || r >= m) r = imod(a, m);
  return r;
}

// an atlas texel: FP32 RGBA (the live atlas) or RGBA16F widened exactly (a baked tile)
__device__ __forceinline__ float4 texel(const float4* a, int64_t i) { return a[i]; }
__device__ __forceinline__ float h2f(uint32_t b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b); }
__device__ __forceinline__ float4 texel(const uint2* a, int64_t i)
{
  const uint2 t = a[i];
  return make_float4(h2f(t.x & 0xFFFFu), h2f(t.x >> 16), h2f(t.y & 0xFFFFu), h2f(t.y >> 16));
}

// texture() with LINEAR filtering and REPEAT addressing (PipelineUtil.h:98-112),
// binary32 weights.
template <class T>
__device__ f3 bilinear_repeat(const T* atlas, int W, int H, float u, float v)
{
  float x = u * (float)W - 0.5f;
  float y = v * (float)H - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  int x0 = wrap_repeat((int)fx, W), x1 = wrap_repeat((int)fx + 1, W);
  int y0 = wrap_repeat((int)fy, H), y1 = wrap_repeat((int)fy + 1, H);
  float4 t00 = texel(atlas, (int64_t)y0 * W + x0);
  float4 t10 = texel(atlas, (int64_t)y0 * W + x1);
  float4 t01 = texel(atlas, (int64_t)y1 * W + x0);
  float4 t11 = texel(atlas, (int64_t)y1 * W + x1);
  float r0 = t00.x * (1.0f - ax) + t10.x * ax, r1 = t01.x * (1.0f - ax) + t11.x * ax;
  float g0 = t00.y * (1.0f - ax) + t10.y * ax, g1 = t01.y * (1.0f - ax) + t11.y * ax;
  float b0 = t00.z * (1.0f - ax) + t10.z * ax, b1 = t01.z * (1.0f - ax) + t11.z * ax;
  return mk(r0 * (1.0f - ay) + r1 * ay, g0 * (1.0f - ay) + g1 * ay, b0 * (1.0f - ay) + b1 * ay);
}

// sampleSingleProbe, pbr_light.glsl:168-188: the texture coordinate of the
// normal's octahedral position inside the tile of probe (ix, iy, iz) is
// start + (end - start) * o, with the tile's (start, end - start) from the
// grid's table (the same binary32 operations, done once per grid).  The 8
// probes of sampleProbe take 2 distinct columns and 4 distinct rows.
__device__ __forceinline__ float tile_coord(float2 t, float o) { return t.x + t.y * o; }

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// worldPosToTileInfo (bindless.glsl:291-308) + tileInfoBuffer[i].ddgiAtlasTex: the
// baked atlas of the tile holding pos within the camera's window, or -1.  The
// window has radius ANE_TILE_RADIUS on both sides (the reference's shader checks
// only the upper side of a radius-10 window against a radius-2 buffer, Q14).
__device__ __forceinline__ int tile_lookup(const FrameK& f, f3 pos, int& tx, int& tz)
{
  const float ts = (float)f.tile_size;
  tx = (int)floorf(pos.x / ts);
  tz = (int)floorf(pos.z / ts);
  const int rx = tx - f.tile_cam[0], rz = tz - f.tile_cam[1];
  if (rx < -ANE_TILE_RADIUS || rx > ANE_TILE_RADIUS || rz < -ANE_TILE_RADIUS || rz > ANE_TILE_RADIUS)
    return -1;
  return f.tile_win[(rz + ANE_TILE_RADIUS) * (2 * ANE_TILE_RADIUS + 1) + rx + ANE_TILE_RADIUS];
}

// sampleProbe, pbr_light.glsl:240-363.  A baked tile at pos places the probes on
// the tile's grid (:251-275); TILE_TEX then also samples the tile's own atlas
// (the deferred lookup, deferred_tiled.comp:276-283), otherwise `atlas` (the
// closest-hit shaders pass the live atlas, rchit:99-106).
template <bool TILE_TEX>
__device__ f3 sample_probe(const GridK& g, const FrameK& f, const float4* atlas, f3 pos, f3 n)
{
  int px, py, pz;
  f3 base, step;
  int tx = 0, tz = 0;
  const int ti = f.tile_count > 0 ? tile_lookup(f, pos, tx, tz) : -1;
  if (ti >= 0 || (f.flags & ANE_FLAG_BAKE)) {
    // a baked tile's grid (:258-275), else the tile being baked (:276-293)
    const bool tiled = ti >= 0;
    const int ts = tiled ? f.tile_size : f.bake_tile[2];
    const int ix = tiled ? tx : f.bake_tile[0], iz = tiled ? tz : f.bake_tile[1];
    f3 toff = mk((float)ix * (float)ts, 0.0f, (float)iz * (float)ts);
    step = mk((float)ts / (float)g.nx, 1.0f, (float)ts / (float)g.nz);
    f3 pit = sub(pos, toff);
    px = clampi((int)floorf(pit.x / step.x), 0, g.nx - 1);
    py = clampi((int)floorf(pit.y / step.y), 0, g.ny - 1);
    pz = clampi((int)floorf(pit.z / step.z), 0, g.nz - 1);
    base = add(mul(mk((float)px, (float)py, (float)pz), step), toff);
  } else {
    step = mk(g.step[0], g.step[1], g.step[2]);
    f3 off = cam_offset(g, f.cam);
    px = clampi((int)floorf((pos.x - off.x) / step.x), 0, g.nx - 1);
    py = clampi((int)floorf(pos.y / step.y), 0, g.ny - 1);
    pz = clampi((int)floorf((pos.z - off.z) / step.z), 0, g.nz - 1);
    base = add(mul(mk((float)px, (float)py, (float)pz), step), off);
  }
  f3 alpha = mk(clampf((pos.x - base.x) / step.x, 0.0f, 1.0f),
                clampf((pos.y - base.y) / step.y, 0.0f, 1.0f),
                clampf((pos.z - base.z) / step.z, 0.0f, 1.0f));
  // octahedral position of the normal, shared by the 8 lookups
  float octx, octy;
  oct_encode(normalize(n), octx, octy);
  octx = (octx + 1.0f) * 0.5f;
  octy = (octy + 1.0f) * 0.5f;
  const int W = g.atlas_w, H = g.atlas_h;
  float tu[2], tv[4];
#pragma unroll
  for (int dx = 0; dx < 2; ++dx) tu[dx] = tile_coord(g.tile_u[px + dx], octx);
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
      tv[2 * dy + dz] = tile_coord(g.tile_v[(py + dy) * (g.nz + 1) + pz + dz], octy);
  // corner order of pbr_light.glsl:318-355, packed as ox | oy << 1 | oz << 2
  const int corners = (0) | (7 << 3) | (3 << 6) | (1 << 9) | (6 << 12) | (2 << 15) | (4 << 18) |
                      (5 << 21);
  float sumw = 0.0f;
  f3 sumi = mk(0.0f, 0.0f, 0.0f);
#pragma unroll kShadeUnroll
  for (int k = 0; k < 8; ++k) {
    const int c = (corners >> (3 * k)) & 7;
    const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
    f3 ppos = add(base, mul(mk((float)ox, (float)oy, (float)oz), step));
    f3 tri = mk(ox ? alpha.x : 1.0f - alpha.x, oy ? alpha.y : 1.0f - alpha.y,
                oz ? alpha.z : 1.0f - alpha.z);
    const f3 dv = sub(ppos, pos);
    float x = fmaxf(0.0001f, (dot(dv, n) * hw_rsqrt(dot(dv, dv)) + 1.0f) * 0.5f);
    float w = x * x + 0.2f;
    w = fmaxf(0.000001f, w);
    if (w < 0.2f) w = w * (w * w * (1.0f / (0.2f * 0.2f)));
    w = w * (tri.x * tri.y * tri.z);
    const float tvs = oy ? (oz ? tv[3] : tv[2]) : (oz ? tv[1] : tv[0]);
    f3 s = (TILE_TEX && ti >= 0)
               ? bilinear_repeat(g.tile_atlas + (int64_t)ti * W * H, W, H, ox ? tu[1] : tu[0], tvs)
               : bilinear_repeat(atlas, W, H, ox ? tu[1] : tu[0], tvs);
    s = mk(hw_sqrt(s.x), hw_sqrt(s.y), hw_sqrt(s.z));
    sumi = add(sumi, muls(s, w));
    sumw = sumw + w;
  }
  f3 net = muls(sumi, hw_rcp(sumw));
  net = mul(net, net);
  return muls(muls(net, 0.5f), 3.1415926f);
}

// calcIndirectDiffuseLight, pbr_light.glsl:365-388
template <bool TILE_TEX = false>
__device__ f3 calc_indirect(const GridK& g, const FrameK& f, const float4* atlas, f3 N, f3 albedo,
                            float metallic, float rough, f3 pos, f3 V)
{
  f3 F0 = add(muls(mk(0.04f, 0.04f, 0.04f), 1.0f - metallic), muls(albedo, metallic));
  float r = rough * rough;
  float p5 = pow5_term(fmaxf(dot(N, V), 0.0f));
  float omr = 1.0f - r;
  f3 kS = mk(F0.x + (fmaxf(omr, F0.x) - F0.x) * p5, F0.y + (fmaxf(omr, F0.y) - F0.y) * p5,
             F0.z + (fmaxf(omr, F0.z) - F0.z) * p5);
  f3 kD = muls(sub(mk(1.0f, 1.0f, 1.0f), kS), 1.0f - metallic);
  f3 irr = sample_probe<TILE_TEX>(g, f, atlas, pos, N);
  f3 diffuse = mul(irr, albedo);
  return mul(kD, diffuse);
}

// rchit:24-110 (+ rmiss:17-21 for misses).  MIRROR: mirror_reflections.rchit,
// which transforms the normal by mat3(ObjectToWorld) (:74) and gates the
// indirect term on ANE_FLAG_DDGI (:107).  The world hit position goes to *hit_pos.
template <bool MIRROR>
__device__ f3 shade_ray(const SceneK& sc, const GridK& g, const FrameK& f, const float4* atlas,
                        f3 origin, ane_hit h, uint8_t shadowed, f3* hit_pos = nullptr)
{
  if (h.tri < 0) return mk(f.sky[0], f.sky[1], f.sky[2]);
  // instance, vertices and position (rchit:27-42): the same arithmetic that gave
  // the shadow-ray origin in the traversal
  HitGeom hg;
  const f3 pos = hit_geom_pos(sc, h.tri, h.u, h.v, hg);
  if (hit_pos) *hit_pos = pos;
  const ane_instance& in = sc.inst[hg.inst];
  const ane_material& mat = sc.mats[mat_of(in.custom_index_and_mask)];
  const bool lit = shadowed == 0;
  float bw[3] = {1.0f - h.u - h.v, h.u, h.v};
  f3 nrm = mk(0.0f, 0.0f, 0.0f), col = mk(0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float4* vp = reinterpret_cast<const float4*>(sc.vertices + (int64_t)hg.v[j] * 24);
    float4 d0 = vp[0], d1 = vp[1];
    float d2x = sc.vertices[(int64_t)hg.v[j] * 24 + 8];
    f3 nj = mk(d1.z, d1.w, d2x);
    f3 cj = mk(d0.w, d1.x, d1.y);
    nrm = add(nrm, muls(nj, bw[j]));
    col = add(col, muls(cj, bw[j]));
  }
  nrm = normalize(nrm);
  f3 nw;
  if (MIRROR) {
    // mat3(ObjectToWorld) * n over the row-major 3x4 instance transform
    const float* t = in.transform;
    nw = mk(t[0] * nrm.x + t[1] * nrm.y + t[2] * nrm.z, t[4] * nrm.x + t[5] * nrm.y + t[6] * nrm.z,
            t[8] * nrm.x + t[9] * nrm.y + t[10] * nrm.z);
  } else {
    // n * WorldToObject: row vector times the inverse transform
    const float* wi = sc.inv_xform + 12 * (int64_t)hg.inst;
    nw = mk(nrm.x * wi[0] + nrm.y * wi[4] + nrm.z * wi[8],
            nrm.x * wi[1] + nrm.y * wi[5] + nrm.z * wi[9],
            nrm.x * wi[2] + nrm.y * wi[6] + nrm.z * wi[10]);
  }
  nrm = normalize(nw);
  col = mk(to_linear_hw(col.x), to_linear_hw(col.y), to_linear_hw(col.z));
  f3 albedo = col;
  if (mat.base_col_fac[0] + mat.base_col_fac[1] + mat.base_col_fac[2] > 0.1f)
    albedo = mk(mat.base_col_fac[0], mat.base_col_fac[1], mat.base_col_fac[2]);
  const float metallic = mat.met_rough[0];
  const float rough = mat.met_rough[1];
  f3 emissive = muls(mk(mat.emissive[0], mat.emissive[1], mat.emissive[2]), mat.emissive[3]);
  f3 acc = mk(0.0f, 0.0f, 0.0f);
  const f3 V = normalize(sub(origin, pos));  // view vector of both terms (pbr_light.glsl:101,369)
  if (lit) acc = calc_light(f, nrm, albedo, metallic, rough, V);
  if (f.flags & (MIRROR ? ANE_FLAG_DDGI : ANE_FLAG_DDGI_MULTI))
    acc = add(acc, calc_indirect(g, f, atlas, nrm, albedo, metallic, rough, pos, V));
  return add(acc, emissive);
}

// One probe's rays (block-level): shading of every ray, stored in the context
// precision.  The block's shared arrays are passed in (sL: FP8 staging [3][N]).
template <int PREC, int TB>
__device__ __forceinline__ void shade_probe(const SceneK& sc, const GridK& g, const FrameK& f,
                                            const float4* probe_org, const ane_hit* hits,
                                            const uint8_t* shadow, const float4* atlas, void* rad,
                                            float* rad_scale, int probe_base, int ql, float* sL, float* red,
                                            uint16_t* lst, uint32_t* cnt, uint32_t* cur)
{
  // the probe's rays grouped as lit hits, shadowed hits, misses, so that a wave
  // runs one kind of work (results are stored per ray, the order is immaterial)
  const int q = probe_base + ql;
  const int n = g.rays;
  const float4 po = probe_org[q];
  const f3 origin = mk(po.x, po.y, po.z);
  const int32_t* tri = reinterpret_cast<const int32_t*>(hits + (int64_t)ql * n);
  const uint8_t* shp = shadow + (int64_t)ql * n;
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0u;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += TB) {
    const int c = tri[4 * i] < 0 ? 2 : (shp[i] == 0 ? 0 : 1);
    atomicAdd(&cnt[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    cur[0] = 0u;
    cur[1] = cnt[0];
    cur[2] = cnt[0] + cnt[1];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += TB) {
    const int c = tri[4 * i] < 0 ? 2 : (shp[i] == 0 ? 0 : 1);
    lst[atomicAdd(&cur[c], 1u)] = (uint16_t)i;
  }
  __syncthreads();
  float mx = 0.0f;
  for (int j = threadIdx.x; j < n; j += TB) {
    const int i = lst[j];
    const int64_t idx = (int64_t)ql * n + i;
    const ane_hit h = hits[idx];
    const uint8_t sh = (h.tri >= 0) ? shadow[idx] : (uint8_t)255;
    const f3 L = shade_ray<false>(sc, g, f, atlas, origin, h, sh);
    if (PREC == ANE_PRECISION_FP32) {
      float* o = (float*)rad + (int64_t)ql * 3 * n;
      o[i] = L.x;
      o[n + i] = L.y;
      o[2 * n + i] = L.z;
    } else if (PREC == ANE_PRECISION_BF16) {
      uint16_t* o = (uint16_t*)rad + (int64_t)ql * 3 * n;
      o[i] = f32_to_bf16(L.x);
      o[n + i] = f32_to_bf16(L.y);
      o[2 * n + i] = f32_to_bf16(L.z);
    } else if (PREC == ANE_PRECISION_FP16) {
      uint16_t* o = (uint16_t*)rad + (int64_t)ql * 3 * n;
      o[i] = f32_to_f16_sat(L.x);
      o[n + i] = f32_to_f16_sat(L.y);
      o[2 * n + i] = f32_to_f16_sat(L.z);
    } else {
      sL[i] = L.x;
      sL[n + i] = L.y;
      sL[2 * n + i] = L.z;
      mx = fmaxf(mx, fmaxf(fmaxf(fabsf(L.x), fabsf(L.y)), fabsf(L.z)));
    }
  }
  if (PREC == ANE_PRECISION_FP8) {
    // one scale per channel and 32 rays (n % 32 == 0): sL[c n + r] is in block
    // (c n + r) / 32 = c (n / 32) + r / 32
    float* se = sL + 3 * n;  // [3 n / 32] quantization factors, after the staged radiance
    float* sco = rad_scale + (int64_t)ql * 3 * (n / 32);
    __syncthreads();
    for (int b = threadIdx.x; b < 3 * (n / 32); b += TB) {
      float m = 0.0f;
      for (int k = 0; k < 32; ++k) m = fmaxf(m, fabsf(sL[32 * b + k]));
      se[b] = fp8_block_inv(m);
      sco[b] = fp8_block_scale(m);
    }
    __syncthreads();
    uint8_t* o = (uint8_t*)rad + (int64_t)ql * 3 * n;
    for (int i = threadIdx.x; i < 3 * n; i += TB) o[i] = f32_to_e4m3(sL[i] * se[i >> 5]);
    (void)mx;
  }
}

template <int PREC, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(ANE_SHADE_WAVES))) void k_shade(SceneK sc, GridK g, FrameK f, const float4* probe_org,
                                              const ane_hit* hits, const uint8_t* shadow, const float4* atlas, void* rad,
                                              float* rad_scale, int probe_base)
{
  extern __shared__ float sL[];  // FP8: [3][N] staged radiance, then [3][N / 32] block factors
  __shared__ float red[TB];
  __shared__ uint16_t lst[4096];
  __shared__ uint32_t cnt[3], cur[3];
  shade_probe<PREC, TB>(sc, g, f, probe_org, hits, shadow, atlas, rad, rad_scale, probe_base, blockIdx.x, sL,
                        red, lst, cnt, cur);
}

// The probes of many updates in one launch (the request loop's fused stage): block
// b shades probe b - block_begin of the last job whose block_begin <= b.
// dynamic LDS of the FP8 shading: staged radiance [3][N] and block factors [3][N / 32]
static size_t fp8_shade_smem(int rays) { return sizeof(float) * (3 * (size_t)rays + 3 * (size_t)((rays + 31) / 32)); }

template <int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ANE_SHADE_WAVES))) void k_shade_jobs(const ShadeJob* jobs, int njobs)
{
  extern __shared__ float sL[];
  __shared__ float red[256];
  __shared__ uint16_t lst[4096];
  __shared__ uint32_t cnt[3], cur[3];
  const int b = blockIdx.x;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  const ShadeJob& j = jobs[lo];
  shade_probe<PREC, 256>(j.sc, j.g, j.f, j.probe_org, j.hits, j.shadow, j.atlas, j.rad, j.rad_scale, j.probe_base,
                         b - j.block_begin, sL, red, lst, cnt, cur);
}

void launch_shade_jobs(int precision, const ShadeJob* jobs, int njobs, int total_blocks, int max_rays,
                       hipStream_t s)
{
  if (njobs <= 0 || total_blocks <= 0) return;
  const dim3 grid((unsigned)total_blocks), block(256);
  if (precision == ANE_PRECISION_FP32)
    hipLaunchKernelGGL(k_shade_jobs<ANE_PRECISION_FP32>, grid, block, 0, s, jobs, njobs);
  else if (precision == ANE_PRECISION_FP16)
    hipLaunchKernelGGL(k_shade_jobs<ANE_PRECISION_FP16>, grid, block, 0, s, jobs, njobs);
  else if (precision == ANE_PRECISION_BF16)
    hipLaunchKernelGGL(k_shade_jobs<ANE_PRECISION_BF16>, grid, block, 0, s, jobs, njobs);
  else
    hipLaunchKernelGGL(k_shade_jobs<ANE_PRECISION_FP8>, grid, block, fp8_shade_smem(max_rays), s, jobs, njobs);
}

template <int TB>
static void shade_launch(const SceneK& sc, const GridK& g, const FrameK& f, const float4* probe_org,
                         const ane_hit* hits,
                         const uint8_t* shadow,
                         const float4* a4, int precision, void* rad, float* rad_scale,
                         int num_probes, int probe_base, hipStream_t s)
{
  if (precision == ANE_PRECISION_FP32)
    hipLaunchKernelGGL((k_shade<ANE_PRECISION_FP32, TB>), dim3(num_probes), dim3(TB), 0, s, sc, g, f,
                       probe_org, hits, shadow, a4, rad, rad_scale, probe_base);
  else if (precision == ANE_PRECISION_FP16)
    hipLaunchKernelGGL((k_shade<ANE_PRECISION_FP16, TB>), dim3(num_probes), dim3(TB), 0, s, sc, g, f,
                       probe_org, hits, shadow, a4, rad, rad_scale, probe_base);
  else if (precision == ANE_PRECISION_BF16)
    hipLaunchKernelGGL((k_shade<ANE_PRECISION_BF16, TB>), dim3(num_probes), dim3(TB), 0, s, sc, g, f,
                       probe_org, hits, shadow, a4, rad, rad_scale, probe_base);
  else
    hipLaunchKernelGGL((k_shade<ANE_PRECISION_FP8, TB>), dim3(num_probes), dim3(TB),
                       fp8_shade_smem(g.rays), s, sc, g, f, probe_org, hits, shadow, a4, rad,
                       rad_scale, probe_base);
}

void launch_shade(const SceneK& sc, const GridK& g, const FrameK& f, const float4* probe_org,
                  const ane_hit* hits,
                  const uint8_t* shadow, const float* atlas, int precision, void* rad,
                  float* rad_scale, int num_probes, int probe_base, hipStream_t s)
{
  if (num_probes <= 0) return;
  const float4* a4 = reinterpret_cast<const float4*>(atlas);
  if (g.rays >= 512)
    shade_launch<ANE_SHADE_TB>(sc, g, f, probe_org, hits, shadow, a4, precision, rad, rad_scale,
                               num_probes, probe_base, s);
  else
    shade_launch<256>(sc, g, f, probe_org, hits, shadow, a4, precision, rad, rad_scale,
                      num_probes, probe_base, s);
}

// ---------------------------------------------------------------------------
// deferred GI lookup (deferred_tiled.comp:264-298): ambient light of G-buffer
// texels from the atlas, or the baked atlas of the texel's tile, through the
// shading's calcIndirectDiffuseLight
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ANE_SHADE_WAVES))) void k_gi_lookup(GridK g, FrameK f, const float4* atlas,
                                                   const float4* gbuf, int64_t n, float4* out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = gbuf[3 * i], b = gbuf[3 * i + 1], c = gbuf[3 * i + 2];
  const f3 pos = mk(a.x, a.y, a.z), N = mk(b.x, b.y, b.z), albedo = mk(c.x, c.y, c.z);
  const float sh = clampf(a.w, 0.3f, 1.0f);
  f3 amb;
  // a baked atlas for the texel's tile (deferred_tiled.comp:265-268) switches the
  // lookup on without ANE_FLAG_DDGI and lifts the distance cut-off
  int tx, tz;
  const bool baked = f.tile_count > 0 && tile_lookup(f, pos, tx, tz) >= 0;
  if ((f.flags & ANE_FLAG_DDGI) || baked) {
    const f3 cam = mk(f.cam[0], f.cam[1], f.cam[2]);
    const f3 d = sub(cam, pos);
    const float dist = sqrtf(dot(d, d));
    if (!baked && dist > (float)g.nx * g.step[0] / 2.0f) {
      amb = muls(mul(muls(mk(0.3f, 0.3f, 0.3f), f.sky_intensity), albedo), sh);
    } else {
      const f3 V = normalize(sub(cam, pos));
      amb = calc_indirect<true>(g, f, atlas, N, albedo, b.w, c.w, pos, V);
    }
  } else {
    amb = muls(mul(mk(0.3f, 0.3f, 0.3f), albedo), sh);
  }
  out[i] = make_float4(amb.x, amb.y, amb.z, 0.0f);
}

void launch_gi_lookup(const GridK& g, const FrameK& f, const float* atlas, const void* gbuf, int64_t n,
                      float* out, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gi_lookup, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, f,
                     reinterpret_cast<const float4*>(atlas), reinterpret_cast<const float4*>(gbuf), n,
                     reinterpret_cast<float4*>(out));
}

// ---------------------------------------------------------------------------
// RT hard shadows (shadow.rgen:29-82): rays from the depth buffer's world
// positions toward the sun, then visibility from the any-hit results
// ---------------------------------------------------------------------------
struct Mat4 {
  float m[16];  // column-major
};

// world position of pixel i of a W-wide depth buffer: pixelToUv
// (helpers.glsl:1-4), then calcWorldPos (shadow.rgen:20-27,
// mirror_reflections.rgen:25-33)
__device__ __forceinline__ f3 pixel_world_pos(int64_t i, int W, int H, float d, const Mat4& ivp)
{
  const int y = (int)(i / W), x = (int)(i - (int64_t)y * W);
  const float u = (2.0f * (float)x + 1.0f) / (2.0f * (float)W);
  const float v = (2.0f * (float)y + 1.0f) / (2.0f * (float)H);
  const float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
  const float* m = ivp.m;
  const float px = m[0] * cx + m[4] * cy + m[8] * d + m[12];
  const float py = m[1] * cx + m[5] * cy + m[9] * d + m[13];
  const float pz = m[2] * cx + m[6] * cy + m[10] * d + m[14];
  const float pw = m[3] * cx + m[7] * cy + m[11] * d + m[15];
  return mk(px / pw, py / pw, pz / pw);
}

__global__ __launch_bounds__(256) void k_shadow_rays(const float* depth, int W, int H, Mat4 ivp, f3 dir,
                                                     float* orig, float* dirs, const uint32_t* skip_if_set)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)W * H) return;
  if (skip_if_set && *skip_if_set != 0u) return;
  const float d = depth[i];
  const f3 p = pixel_world_pos(i, W, H, d, ivp);
  orig[3 * i + 0] = p.x;
  orig[3 * i + 1] = p.y;
  orig[3 * i + 2] = p.z;
  // sky pixels get a zero direction, which the traversal takes as no ray
  const bool sky = d == 1.0f;
  dirs[3 * i + 0] = sky ? 0.0f : dir.x;
  dirs[3 * i + 1] = sky ? 0.0f : dir.y;
  dirs[3 * i + 2] = sky ? 0.0f : dir.z;
}

__global__ __launch_bounds__(256) void k_shadow_resolve(const float* depth, const ane_hit* hits, int64_t n,
                                                        float* vis, const uint32_t* skip_if_set)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (skip_if_set && *skip_if_set != 0u) return;
  vis[i] = (depth[i] == 1.0f || hits[i].tri < 0) ? 1.0f : 0.0f;
}

// ---------------------------------------------------------------------------
// point-light shadow moments (LightShadowRayTracingPass.cpp:17-100 with
// light_shadow.rgen/.rchit/.rmiss; LightShadowSumPass.cpp:16-87 with
// light_shadow_sum.comp; constants of light_shadow_helpers.glsl:1-4)
// ---------------------------------------------------------------------------
static constexpr int kLightSide = 32;  // LIGHT_DIR_DIST_PIX_SIZE: ray y * 32 + x at pixel (x, y)

// light_shadow.rgen:40-81: ray r of light l leaves the light's position along
// Fibonacci direction r over [0.1, range]; a disabled light (color.w < 0.5) casts
// no ray (zero direction).  Slot j of a light traces ray order[j], so that the
// lanes of a wave follow neighbouring directions.
__global__ __launch_bounds__(256) void k_light_rays(const float* lights, int n, const float* dirs,
                                                    const uint32_t* order, float* orig, float* rdir,
                                                    float* tmax)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * kLightRays) return;
  const int l = (int)(i / kLightRays), r = (int)order[i % kLightRays];
  const float* L = lights + 8 * (int64_t)l;
  const bool on = !(L[7] < 0.5f);  // rgen:54: disabled when color.w < 0.5
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    orig[3 * i + k] = L[k];
    rdir[3 * i + k] = on ? dirs[3 * r + k] : 0.0f;
  }
  tmax[i] = L[3];
}

// light_shadow.rchit:15-18 (hit distance), light_shadow.rmiss:12-15 (the range),
// light_shadow.rgen:54-58 (disabled light: 0), stored at rgen:43-49, 84
__global__ __launch_bounds__(256) void k_light_dist(const float* lights, int n, const uint32_t* order,
                                                    const ane_hit* hits, float* dist, float* dist_img)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * kLightRays) return;
  const int l = (int)(i / kLightRays), r = (int)order[i % kLightRays];
  const float* L = lights + 8 * (int64_t)l;
  const float d = L[7] < 0.5f ? 0.0f : (hits[i].tri >= 0 ? hits[i].t : L[3]);
  dist[(int64_t)l * kLightRays + r] = d;
  if (dist_img)
    dist_img[(int64_t)(r / kLightSide) * kLightSide * n + (int64_t)kLightSide * l + r % kLightSide] = d;
}

// light_shadow_sum.comp:16-71: one thread per texel of one light's 8x8 tile.  The
// shader visits the 1024 rays x outer, y inner and adds those whose weight
// pow(max(0, dot(dir, octDir)), 50) is above 0.001 (63-69 of them per texel).
// lw / lidx [k][texel] list exactly those (weight, ray) pairs of each texel in that
// order, padded with weight 0, so the sums are the shader's, operation for
// operation, whatever the launch shape.  Texel (tx, ty) of light l lands at
// (10 l + 1 + tx, ty) of the RG32F image; the border columns are not written.
__global__ __launch_bounds__(64) void k_light_moments(const float* lights, int n, const float* dist,
                                                      const float* lw, const int32_t* lidx, int kmax,
                                                      float* mom)
{
  __shared__ float ds[kLightRays];
  const int l = blockIdx.x, t = threadIdx.x;
  for (int k = t; k < kLightRays; k += 64) ds[k] = dist[(int64_t)l * kLightRays + k];
  __syncthreads();
  float ar = 0.0f, ag = 0.0f, aw = 0.0f;
  if (!(lights[8 * (int64_t)l + 7] < 0.5f)) {  // sum.comp:35
#pragma unroll 8
    for (int k = 0; k < kmax; ++k) {
      const float w = lw[k * 64 + t];
      if (w > 0.001f) {
        const float d = ds[lidx[k * 64 + t]];
        ar += w * d;
        ag += w * d * d;
        aw += w;
      }
    }
    if (aw > 1e-6f) {
      ar = ar / aw;
      ag = ag / aw;
    }
  }
  float* o = mom + 2 * ((int64_t)(t >> 3) * (10 * (int64_t)n) + 10 * (int64_t)l + 1 + (t & 7));
  o[0] = ar;
  o[1] = ag;
}

void launch_light_rays(const float* lights, int n, const float* dirs, const uint32_t* order, float* orig,
                       float* rdir, float* tmax, hipStream_t s)
{
  const int64_t m = (int64_t)n * kLightRays;
  if (m <= 0) return;
  hipLaunchKernelGGL(k_light_rays, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, lights, n, dirs,
                     order, orig, rdir, tmax);
}

void launch_light_dist(const float* lights, int n, const uint32_t* order, const ane_hit* hits, float* dist,
                       float* dist_img, hipStream_t s)
{
  const int64_t m = (int64_t)n * kLightRays;
  if (m <= 0) return;
  hipLaunchKernelGGL(k_light_dist, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, lights, n, order,
                     hits, dist, dist_img);
}

void launch_light_moments(const float* lights, int n, const float* dist, const float* lw,
                          const int32_t* lidx, int kmax, float* mom, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_light_moments, dim3((unsigned)n), dim3(64), 0, s, lights, n, dist, lw, lidx, kmax,
                     mom);
}

void launch_shadow_rays(const float* depth, int W, int H, const float* ivp16, const float* light3,
                        float* orig, float* dirs, hipStream_t s, const uint32_t* skip_if_set)
{
  Mat4 m;
  for (int k = 0; k < 16; ++k) m.m[k] = ivp16[k];
  const f3 dir = mk(-light3[0], -light3[1], -light3[2]);
  const int64_t n = (int64_t)W * H;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_shadow_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, depth, W, H, m, dir,
                     orig, dirs, skip_if_set);
}

void launch_shadow_resolve(const float* depth, const ane_hit* hits, int64_t n, float* vis, hipStream_t s,
                           const uint32_t* skip_if_set)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_shadow_resolve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, depth, hits, n, vis,
                     skip_if_set);
}

// RT hard shadows through a sun grid built for the rays' direction: a thread per
// pixel sets up its ray (the world position and ray setup of k_shadow_rays and the
// trace) and looks up its cell; the candidates of a wave's 64 pixels are then
// spread over its lanes (grid_wave_events), each in the watertight test: any hit
// is the any-hit answer of the traversal.  Nothing unless the grid is valid.
__global__ __launch_bounds__(256) void k_shadow_mask_grid(const float* depth, int W, int H, Mat4 ivp, f3 dir,
                                                          SunGridK sg, SceneK sc, float* vis)
{
  __shared__ uint32_t key[256];
  const SunGridParams g = *sg.params;
  if (!g.valid) return;
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  const bool in = i < (int64_t)W * H;
  const bool dir_ok = !(dir.x == 0.0f && dir.y == 0.0f && dir.z == 0.0f) && __builtin_isfinite(dir.x) &&
                      __builtin_isfinite(dir.y) && __builtin_isfinite(dir.z);
  const float d = in ? depth[i] : 1.0f;
  const bool ray = in && d != 1.0f && dir_ok;
  f3 p = mk(0.0f, 0.0f, 0.0f);
  uint32_t b = 0u, c = 0u;
  if (ray) {
    p = pixel_world_pos(i, W, H, d, ivp);
    const float X = comp(p, g.kx) - g.sx * comp(p, g.kz);
    const float Y = comp(p, g.ky) - g.sy * comp(p, g.kz);
    const int64_t cell = (int64_t)sun_cell(Y, g.y0, g.ich, g.gy) * g.gx + sun_cell(X, g.x0, g.icw, g.gx);
    b = sg.start[cell];
    c = sg.start[cell + 1] - b;
  }
  key[tid] = 2u * c + 1u;
  RayPre r;
  ray_setup(r, p, dir_ok ? dir : mk(0.0f, 0.0f, 1.0f));
  grid_wave_events<false>(sg, sc.tris, sc.ntri, r, p, b, c, 0.1f, 50.0f, key + (tid & ~63), nullptr);
  if (in) vis[i] = (ray && !(key[tid] & 1u)) ? 0.0f : 1.0f;  // shadow.rgen:66-67
}

// any hit of caller rays along the grid's direction (hits: tri -1 on a miss)
__global__ __launch_bounds__(256) void k_anyhit_grid(const float* orig, const float* dirs, int64_t n, SunGridK sg,
                                                     SceneK sc, float tmin, float tmax, ane_hit* hits)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SunGridParams g = *sg.params;
  if (!g.valid) return;
  const f3 o = ld3(orig + 3 * i), d = ld3(dirs + 3 * i);
  ane_hit h;
  h.tri = -1;
  h.t = h.u = h.v = 0.0f;
  const bool none = (d.x == 0.0f && d.y == 0.0f && d.z == 0.0f) || !__builtin_isfinite(d.x) ||
                    !__builtin_isfinite(d.y) || !__builtin_isfinite(d.z);
  if (!none) {
    RayPre r;
    ray_setup(r, o, d);
    (void)grid_any_hit(sg, g, sc.tris, sc.ntri, r, o, tmin, tmax, &h);
  }
  hits[i] = h;
}

void launch_shadow_mask_grid(const float* depth, int W, int H, const float* ivp16, const float* light3,
                             const SunGridK& sg, const SceneK& sc, float* vis, hipStream_t s)
{
  Mat4 m;
  for (int k = 0; k < 16; ++k) m.m[k] = ivp16[k];
  const f3 dir = mk(-light3[0], -light3[1], -light3[2]);
  const int64_t n = (int64_t)W * H;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_shadow_mask_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, depth, W, H, m, dir,
                     sg, sc, vis);
}

void launch_anyhit_grid(const float* orig, const float* dirs, int64_t n, const SunGridK& sg, const SceneK& sc,
                        float tmin, float tmax, ane_hit* hits, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_anyhit_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, orig, dirs, n, sg, sc,
                     tmin, tmax, hits);
}

// ---------------------------------------------------------------------------
// mirror reflections (mirror_reflections.rgen/.rchit/.rmiss): a closest-hit
// ray per depth-buffer pixel along the view vector reflected about the
// G-buffer normal, a shadow ray from each hit toward the sun, then the hit
// shading of the probe rays with the mirror variant's normal and DDGI gate
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mirror_rays(const float* depth, const float4* normal4, int W, int H,
                                                     Mat4 ivp, f3 cam, float* orig, float* dirs)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)W * H) return;
  const float d = depth[i];
  const f3 p = pixel_world_pos(i, W, H, d, ivp);
  orig[3 * i + 0] = p.x;
  orig[3 * i + 1] = p.y;
  orig[3 * i + 2] = p.z;
  f3 r = mk(0.0f, 0.0f, 0.0f);  // sky pixels: zero direction, no ray
  if (d != 1.0f) {
    const float4 nq = normal4[i];
    const f3 n = normalize(mk(nq.x, nq.y, nq.z));
    const f3 v = normalize(sub(p, cam));
    r = sub(v, muls(n, 2.0f * dot(n, v)));  // reflect(v, n)
  }
  dirs[3 * i + 0] = r.x;
  dirs[3 * i + 1] = r.y;
  dirs[3 * i + 2] = r.z;
}

// the shadow ray of every reflection hit (mirror_reflections.rchit:44-66)
__global__ __launch_bounds__(256) void k_mirror_shadow_rays(SceneK sc, const ane_hit* hits, int64_t n, f3 sun,
                                                            float* orig, float* dirs)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ane_hit h = hits[i];
  f3 p = mk(0.0f, 0.0f, 0.0f), d = mk(0.0f, 0.0f, 0.0f);
  if (h.tri >= 0) {
    HitGeom hg;
    p = hit_geom_pos(sc, h.tri, h.u, h.v, hg);
    d = sun;
  }
  orig[3 * i + 0] = p.x;
  orig[3 * i + 1] = p.y;
  orig[3 * i + 2] = p.z;
  dirs[3 * i + 0] = d.x;
  dirs[3 * i + 1] = d.y;
  dirs[3 * i + 2] = d.z;
}

// out = (irradiance, distance to the hit); a miss: (sky, 1e6); a sky pixel: 0
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ANE_SHADE_WAVES))) void k_mirror_shade(SceneK sc, GridK g, FrameK f, const float4* atlas,
                                                      const float* depth, const float* orig,
                                                      const ane_hit* hits, const ane_hit* shadow_hits,
                                                      int64_t n, float4* out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (depth[i] == 1.0f) {
    out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const ane_hit h = hits[i];
  if (h.tri < 0) {
    out[i] = make_float4(f.sky[0], f.sky[1], f.sky[2], 1e6f);
    return;
  }
  const f3 o = mk(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
  const uint8_t shadowed = shadow_hits[i].tri >= 0 ? 1 : 0;
  f3 pos;
  const f3 L = shade_ray<true>(sc, g, f, atlas, o, h, shadowed, &pos);
  const f3 dd = sub(pos, o);
  out[i] = make_float4(L.x, L.y, L.z, sqrtf(dot(dd, dd)));
}

void launch_mirror_rays(const float* depth, const float* normal4, int W, int H, const float* ivp16,
                        const float* cam3, float* orig, float* dirs, hipStream_t s)
{
  Mat4 m;
  for (int k = 0; k < 16; ++k) m.m[k] = ivp16[k];
  const int64_t n = (int64_t)W * H;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, depth,
                     reinterpret_cast<const float4*>(normal4), W, H, m, mk(cam3[0], cam3[1], cam3[2]), orig,
                     dirs);
}

void launch_mirror_shadow_rays(const SceneK& sc, const ane_hit* hits, int64_t n, const float* sun3, float* orig,
                               float* dirs, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror_shadow_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sc, hits, n,
                     mk(sun3[0], sun3[1], sun3[2]), orig, dirs);
}

void launch_mirror_shade(const SceneK& sc, const GridK& g, const FrameK& f, const float* atlas,
                         const float* depth, const float* orig, const ane_hit* hits, const ane_hit* shadow_hits,
                         int64_t n, float* out, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror_shade, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sc, g, f,
                     reinterpret_cast<const float4*>(atlas), depth, orig, hits, shadow_hits, n,
                     reinterpret_cast<float4*>(out));
}

}  // namespace ane
