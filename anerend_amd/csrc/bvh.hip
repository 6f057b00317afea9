// bvh.hip -- LBVH build and refit on gfx950.
//
// Replaces the driver's acceleration-structure build
// (VulkanRenderer.cpp:1267-1417 BLAS, UpdateTLASPass.cpp:184-228 TLAS,
// UpdateBlasPass.cpp:131-138 refit).  Instances are flattened to world-space
// triangles and one LBVH (Karras 2012) is built over them:
//   flatten -> centroid bounds -> 63-bit Morton keys -> stable LSD radix sort
//   (8 x 8-bit passes over (key, id)) -> Karras hierarchy -> triangle reorder
//   -> bottom-up child boxes (atomic arrival counters) -> depth.
// Every step is integer or exact min/max arithmetic on binary32 values
// computed with the same float contract as the oracle, so the topology and
// the boxes are bit-exact with the CPU build.
#include "common.hpp"
#include "kernels.hpp"

#include <vector>

namespace ane {

static constexpr int kMaxLevels = 128;  // > the 96-level bound of a 63-bit-key LBVH
static constexpr int kTopLevels = 10;   // depth < 10 holds at most 512 nodes per level
static constexpr int kBlock = 256;
static constexpr int kItems = 8;
static constexpr int kTile = kBlock * kItems;

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2o(float f)
{
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float o2f(uint32_t o)
{
  uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  return __uint_as_float(b);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

// ---------------------------------------------------------------------------
// flatten: world-space triangles in input order + centroid bounds.  The first
// build finds each triangle's instance, mesh, indices and vertices and writes its
// 64-byte record (object-space positions, vertex indices, instance); rebuilds and
// refits keep every instance's mesh (checked on the host) and read that record.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_flatten(SceneK sc, float* world, uint32_t* tri_inst,
                                                 float4* tri_rec, uint32_t* part,
                                                 int32_t write_inst)
{
  __shared__ uint32_t sb[6];
  if (threadIdx.x < 6) sb[threadIdx.x] = (threadIdx.x < 3) ? 0xFFFFFFFFu : 0u;
  __syncthreads();
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t cmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, cmax[3] = {0u, 0u, 0u};
  if (t < sc.ntri) {
    int lo;
    f3 po[3];
    float4* r = tri_rec + 4 * (int64_t)t;
    if (write_inst) {
      // instance k: last k with inst_tri_off[k] <= t
      lo = 0;
      int hi = sc.ninst - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sc.inst_tri_off[mid] <= (uint32_t)t) lo = mid; else hi = mid - 1;
      }
      const ane_mesh& m = sc.meshes[mesh_of(sc.inst[lo].custom_index_and_mask)];
      uint32_t prim = (uint32_t)t - sc.inst_tri_off[lo];
      uint32_t vi[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        uint32_t idx = sc.indices[m.index_offset + 3 * prim + j];
        vi[j] = m.vertex_offset + idx;
        po[j] = ld3(sc.vertices + (int64_t)(m.vertex_offset + idx) * 24);
      }
      tri_inst[t] = (uint32_t)lo;
      // one 64-byte record per triangle: what a hit needs, in one cache line
#pragma unroll
      for (int j = 0; j < 3; ++j) r[j] = make_float4(po[j].x, po[j].y, po[j].z, __uint_as_float(vi[j]));
      r[3] = make_float4(__uint_as_float((uint32_t)lo), 0.0f, 0.0f, 0.0f);
    } else {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float4 q = r[j];
        po[j] = mk(q.x, q.y, q.z);
      }
      lo = (int)__float_as_uint(r[3].x);
    }
    const ane_instance& in = sc.inst[lo];
    float w[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      f3 q = xform_point(in.transform, po[j]);
      w[3 * j + 0] = q.x;
      w[3 * j + 1] = q.y;
      w[3 * j + 2] = q.z;
    }
    float* dst = world + (int64_t)9 * t;
#pragma unroll
    for (int j = 0; j < 9; ++j) dst[j] = w[j];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float l = fminf(fminf(w[a], w[3 + a]), w[6 + a]);
      float h = fmaxf(fmaxf(w[a], w[3 + a]), w[6 + a]);
      float c = (l + h) * 0.5f;
      cmin[a] = f2o(c);
      cmax[a] = cmin[a];
    }
  }
  // wave-wide min / max first: one LDS atomic per wave and bound, not one per thread
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cmin[a] = min(cmin[a], (uint32_t)__shfl_xor((int)cmin[a], off, 64));
      cmax[a] = max(cmax[a], (uint32_t)__shfl_xor((int)cmax[a], off, 64));
    }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&sb[a], cmin[a]);
      atomicMax(&sb[3 + a], cmax[a]);
    }
  }
  __syncthreads();
  // one partial per block (k_bounds_reduce combines them): a device-wide atomic per
  // block on the same six words would serialize the whole launch
  if (part && threadIdx.x < 6) part[threadIdx.x * gridDim.x + blockIdx.x] = sb[threadIdx.x];
}

// centroid bounds of the scene from the per-block partials of k_flatten: min of
// part[0..2][*], max of part[3..5][*] (order-preserving uint encoding)
__global__ __launch_bounds__(256) void k_bounds_reduce(const uint32_t* part, int nblocks, uint32_t* cbounds)
{
  __shared__ uint32_t red[6][256];
  const int t = threadIdx.x;
  uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  for (int i = t; i < nblocks; i += 256)
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const uint32_t x = part[a * nblocks + i];
      v[a] = a < 3 ? min(v[a], x) : max(v[a], x);
    }
#pragma unroll
  for (int a = 0; a < 6; ++a) red[a][t] = v[a];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
#pragma unroll
      for (int a = 0; a < 6; ++a)
        red[a][t] = a < 3 ? min(red[a][t], red[a][t + s]) : max(red[a][t], red[a][t + s]);
    __syncthreads();
  }
  if (t < 6) cbounds[t] = red[t][0];
}

// ---------------------------------------------------------------------------
// Morton keys (contract: quantization of the centroid in a cube of side
// max extent, 21 bits per axis, x most significant)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__device__ __forceinline__ uint32_t quant21(float c, float cmin, float e)
{
  float q = (c - cmin) / e * 2097152.0f;
  q = fminf(fmaxf(q, 0.0f), 2097151.0f);
  return (uint32_t)q;
}

__global__ __launch_bounds__(256) void k_morton(const float* world, int n, const uint32_t* cbounds,
                                                uint64_t* keys, uint32_t* ids)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float cmin[3], cmax[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    cmin[a] = o2f(cbounds[a]);
    cmax[a] = o2f(cbounds[3 + a]);
  }
  float e = fmaxf(fmaxf(cmax[0] - cmin[0], cmax[1] - cmin[1]), cmax[2] - cmin[2]);
  if (!(e > 0.0f)) e = 1.0f;
  const float* w = world + (int64_t)9 * t;
  uint64_t q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float l = fminf(fminf(w[a], w[3 + a]), w[6 + a]);
    float h = fmaxf(fmaxf(w[a], w[3 + a]), w[6 + a]);
    float c = (l + h) * 0.5f;
    q[a] = quant21(c, cmin[a], e);
  }
  keys[t] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
  ids[t] = (uint32_t)t;
}

// ---------------------------------------------------------------------------
// exclusive scan (uint32), three-kernel block scan, recursive over levels
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
  __shared__ uint32_t ws[kBlock / 64];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) wbase += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  *total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_block(uint32_t* data, int64_t n, uint32_t* sums)
{
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    v[i] = (base + i < n) ? data[base + i] : 0u;
    s += v[i];
  }
  uint32_t total;
  uint32_t pre = block_exclusive_scan(s, &total);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i < n) data[base + i] = pre;
    pre += v[i];
  }
  if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t* data, int64_t n, const uint32_t* sums)
{
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  uint32_t add = sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; ++i)
    if (base + i < n) data[base + i] += add;
}

// scratch must hold scan_scratch_words(n) words
int64_t scan_scratch_words(int64_t n)
{
  int64_t w = 0;
  while (n > kTile) {
    n = (n + kTile - 1) / kTile;
    w += n;
  }
  return w + 1;
}

void exclusive_scan(uint32_t* data, int64_t n, uint32_t* scratch, hipStream_t s)
{
  int64_t blocks = (n + kTile - 1) / kTile;
  if (blocks <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kBlock), 0, s, data, n, (uint32_t*)nullptr);
    return;
  }
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)blocks), dim3(kBlock), 0, s, data, n, scratch);
  exclusive_scan(scratch, blocks, scratch + blocks, s);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)blocks), dim3(kBlock), 0, s, data, n,
                     (const uint32_t*)scratch);
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (uint64 key, uint32 value), 8-bit digits
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* keys, int64_t n, int shift,
                                                      uint32_t* hist, int64_t ntiles)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    int64_t idx = base + i * kBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// One block per digit d: exclusive scan of hist[d][0..ntiles) in place (the
// digit's count in the tiles before each tile) and the digit's total.
__global__ __launch_bounds__(kBlock) void k_radix_digit_scan(uint32_t* hist, int64_t ntiles, uint32_t* totals)
{
  uint32_t* h = hist + (int64_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  for (int64_t base = 0; base < ntiles; base += kTile) {
    const int64_t i0 = base + (int64_t)threadIdx.x * kItems;
    uint32_t v[kItems], s = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      v[i] = (i0 + i < ntiles) ? h[i0 + i] : 0u;
      s += v[i];
    }
    uint32_t total;
    uint32_t pre = carry + block_exclusive_scan(s, &total);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if (i0 + i < ntiles) h[i0 + i] = pre;
      pre += v[i];
    }
    carry += total;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one tile: digit d of tile t goes to (digits below d in all
// tiles) + (digit d in the tiles before t) + (rank within the tile).  The tile is
// first ordered by digit in LDS (ranks by ballots, as a stable counting sort), so
// that consecutive threads write consecutive addresses of each digit's run.
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* kin, const uint32_t* vin,
                                                         uint64_t* kout, uint32_t* vout, int64_t n,
                                                         int shift, const uint32_t* offs,
                                                         const uint32_t* totals, int64_t ntiles)
{
  __shared__ uint32_t run[256];    // next tile-local position of each digit
  __shared__ uint32_t gbase[256];  // global position of the tile's digit d, minus its local start
  __shared__ uint32_t wcnt[kBlock / 64][256];
  __shared__ uint64_t skey[kTile];
  __shared__ uint32_t sval[kTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * kTile;
  const int tn = (int)min((int64_t)kTile, n - base);
  uint64_t key[kItems];
  uint32_t val[kItems];
  run[threadIdx.x] = 0u;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const int64_t idx = base + r * kBlock + threadIdx.x;
    key[r] = idx < n ? kin[idx] : 0ull;
    val[r] = idx < n ? vin[idx] : 0u;
    if (idx < n) atomicAdd(&run[(uint32_t)(key[r] >> shift) & 255u], 1u);
  }
  __syncthreads();
  uint32_t tot;
  const uint32_t lstart = block_exclusive_scan(run[threadIdx.x], &tot);
  const uint32_t below = block_exclusive_scan(totals[threadIdx.x], &tot);
  gbase[threadIdx.x] = below + offs[(int64_t)threadIdx.x * ntiles + blockIdx.x] - lstart;
  run[threadIdx.x] = lstart;
  for (int r = 0; r < kItems; ++r) {
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const bool valid = base + r * kBlock + threadIdx.x < n;
    const uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
    uint64_t mask = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      mask &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(mask & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (valid) {
      uint32_t q = run[d];
      for (int w = 0; w < wave; ++w) q += wcnt[w][d];
      q += rank;
      skey[q] = key[r];
      sval[q] = val[r];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) add += wcnt[w][threadIdx.x];
    run[threadIdx.x] += add;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < tn; q += kBlock) {
    const uint64_t k = skey[q];
    const uint32_t dst = gbase[(uint32_t)(k >> shift) & 255u] + (uint32_t)q;
    if (dst < (uint64_t)n) {
      kout[dst] = k;
      vout[dst] = sval[q];
    }
  }
}

int64_t radix_scratch_words(int64_t n)
{
  int64_t ntiles = (n + kTile - 1) / kTile;
  return 256 * ntiles + 256;  // tile histograms, digit totals
}

// Sorts (keys, vals) in place using (k2, v2) as ping-pong buffers (8 passes:
// the result ends in keys/vals).  scratch: radix_scratch_words(n) words.
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* k2, uint32_t* v2, int64_t n,
                      uint32_t* scratch, hipStream_t s)
{
  if (n <= 1) return;
  int64_t ntiles = (n + kTile - 1) / kTile;
  uint32_t* hist = scratch;
  uint32_t* totals = scratch + 256 * ntiles;
  uint64_t* ka = keys;
  uint32_t* va = vals;
  uint64_t* kb = k2;
  uint32_t* vb = v2;
  for (int pass = 0; pass < 8; ++pass) {
    int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kBlock), 0, s, (const uint64_t*)ka,
                       n, shift, hist, ntiles);
    hipLaunchKernelGGL(k_radix_digit_scan, dim3(256), dim3(kBlock), 0, s, hist, ntiles, totals);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)ntiles), dim3(kBlock), 0, s,
                       (const uint64_t*)ka, (const uint32_t*)va, kb, vb, n, shift,
                       (const uint32_t*)hist, (const uint32_t*)totals, ntiles);
    uint64_t* tk = ka; ka = kb; kb = tk;
    uint32_t* tv = va; va = vb; vb = tv;
  }
  // 8 passes: data is back in (keys, vals)
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint64_t* k, int n, int i, int j)
{
  if (j < 0 || j >= n) return -1;
  uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(a ^ b);
}

__global__ __launch_bounds__(256) void k_karras(const uint64_t* k, int n, NodeBuild* nodes,
                                                int32_t* leaf_parent, int4* topo)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (kdelta(k, n, i, i + 1) - kdelta(k, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta(k, n, i, i - d);
  int lmax = 2;
  while (kdelta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(k, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta(k, n, i, j);
  int sp = 0;
  int div = 2;
  int t = (l + div - 1) / div;
  for (;;) {
    if (kdelta(k, n, i, i + (sp + t) * d) > dnode) sp += t;
    if (t <= 1) break;
    div *= 2;
    t = (l + div - 1) / div;
  }
  int gamma = i + sp * d + (d < 0 ? d : 0);
  int lo = min(i, j), hi = max(i, j);
  int L = (lo == gamma) ? ~gamma : gamma;
  int R = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  nodes[i].d.x = L;
  nodes[i].d.y = R;
  // compact topology for the fit: children, parent, and j, the other end of the
  // node's leaf range [min(i, j), max(i, j)]
  topo[i].x = L;
  topo[i].y = R;
  topo[i].w = j;
  if (i == 0) {
    nodes[0].d.z = -1;
    topo[0].z = -1;
  }
  if (L < 0) leaf_parent[~L] = i; else { nodes[L].d.z = i; topo[L].z = i; }
  if (R < 0) leaf_parent[~R] = i; else { nodes[R].d.z = i; topo[R].z = i; }
}

// ---------------------------------------------------------------------------
// triangle reorder and bottom-up boxes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reorder(const float* world, const uint32_t* ids, int n,
                                                 float4* tris)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t id = ids[p];
  const float* w = world + (int64_t)9 * id;
  tris[3 * (int64_t)p + 0] = make_float4(w[0], w[1], w[2], __uint_as_float(id));
  tris[3 * (int64_t)p + 1] = make_float4(w[3], w[4], w[5], 0.0f);
  tris[3 * (int64_t)p + 2] = make_float4(w[6], w[7], w[8], 0.0f);
}

// Refit: the triangles in leaf order straight from their 64-byte records and the
// moved instances (the flatten's transform, written at the triangle's sorted
// position), without the input-order world copy a build sorts from.
__global__ __launch_bounds__(256) void k_refit_tris(SceneK sc, const uint32_t* ids, int n, float4* tris)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t id = ids[p];
  const float4* r = sc.tri_rec + 4 * (int64_t)id;
  const float4 a = r[0], b = r[1], c = r[2], d = r[3];
  const float* m = sc.inst[__float_as_uint(d.x)].transform;
  const f3 w0 = xform_point(m, mk(a.x, a.y, a.z));
  const f3 w1 = xform_point(m, mk(b.x, b.y, b.z));
  const f3 w2 = xform_point(m, mk(c.x, c.y, c.z));
  tris[3 * (int64_t)p + 0] = make_float4(w0.x, w0.y, w0.z, __uint_as_float(id));
  tris[3 * (int64_t)p + 1] = make_float4(w1.x, w1.y, w1.z, 0.0f);
  tris[3 * (int64_t)p + 2] = make_float4(w2.x, w2.y, w2.z, 0.0f);
}

// The input-order world triangles from the leaf-order ones (for inspection after a refit).
__global__ __launch_bounds__(256) void k_world_from_tris(const float4* tris, int n, float* world)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float4 v0 = tris[3 * (int64_t)p], v1 = tris[3 * (int64_t)p + 1], v2 = tris[3 * (int64_t)p + 2];
  float* w = world + 9 * (int64_t)__float_as_uint(v0.w);
  w[0] = v0.x; w[1] = v0.y; w[2] = v0.z;
  w[3] = v1.x; w[4] = v1.y; w[5] = v1.z;
  w[6] = v2.x; w[7] = v2.y; w[8] = v2.z;
}

// Box of one child: a leaf's triangle box or an internal node's stored box.
__device__ __forceinline__ void child_box(const float4* tris, const float* node_box, int c, float* lo,
                                          float* hi)
{
  if (c < 0) {
    const int64_t p = ~c;
    float4 v0 = tris[3 * p], v1 = tris[3 * p + 1], v2 = tris[3 * p + 2];
    lo[0] = fminf(fminf(v0.x, v1.x), v2.x); lo[1] = fminf(fminf(v0.y, v1.y), v2.y);
    lo[2] = fminf(fminf(v0.z, v1.z), v2.z);
    hi[0] = fmaxf(fmaxf(v0.x, v1.x), v2.x); hi[1] = fmaxf(fmaxf(v0.y, v1.y), v2.y);
    hi[2] = fmaxf(fmaxf(v0.z, v1.z), v2.z);
  } else {
    const float* b = node_box + 6 * (int64_t)c;
    lo[0] = b[0]; lo[1] = b[1]; lo[2] = b[2]; hi[0] = b[3]; hi[1] = b[4]; hi[2] = b[5];
  }
}


// Quantized nodes (layout next to NodeBuild): one thread per internal node, from
// the fitted child boxes.  The same arithmetic restated on the CPU gives the same
// bytes, so a traversal over these nodes visits what the CPU's does.
__device__ int quant_exp(double extent)
{
  int e = -126;
  if (extent > 0.0) {
    // frexp's exponent k (extent = m 2^k, m in [1/2, 1)) from the bits: extent is the
    // difference of two binary32 values, so a normal binary64
    const int k = (int)((__double_as_longlong(extent) >> 52) & 0x7FF) - 1022;  // extent < 2^k
    e = k - 9;                // 254 * 2^(k-9) < 2^k
    while (ldexp(254.0, e) < extent) ++e;
    if (e < -126) e = -126;
  }
  return e;
}

// The quantized record of one node from its two child boxes (lo, hi per axis).
__device__ __forceinline__ void quantize_node(const float* l0, const float* h0, const float* l1,
                                              const float* h1, int L, int R, uint4* q2)
{
  uint32_t by[12], eb = 0u, ob[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float o = l0[a] < l1[a] ? l0[a] : l1[a];
    const float hi = h0[a] > h1[a] ? h0[a] : h1[a];
    const int e = quant_exp((double)hi - (double)o);
    // division by the power-of-two step 2^e, exact as a scaling by 2^-e
    by[0 + a] = (uint32_t)floor(ldexp((double)l0[a] - (double)o, -e));
    by[3 + a] = (uint32_t)ceil(ldexp((double)h0[a] - (double)o, -e));
    by[6 + a] = (uint32_t)floor(ldexp((double)l1[a] - (double)o, -e));
    by[9 + a] = (uint32_t)ceil(ldexp((double)h1[a] - (double)o, -e));
    eb |= (uint32_t)(e + 128) << (8 * a);
    ob[a] = __float_as_uint(o);
  }
  eb |= (L < 0 ? 1u : 0u) << 24 | (R < 0 ? 1u : 0u) << 25;
  const uint32_t g = (uint32_t)(L >= 0 ? L : ~L);
  q2[0] = make_uint4(ob[0], ob[1], ob[2], eb);
  q2[1] = make_uint4(by[0] | by[1] << 8 | by[2] << 16 | by[3] << 24,
                     by[4] | by[5] << 8 | by[6] << 16 | by[7] << 24,
                     by[8] | by[9] << 8 | by[10] << 16 | by[11] << 24, g);
}

__global__ __launch_bounds__(256) void k_quantize_nodes(const NodeBuild* nodes, int n, uint4* q)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const NodeBuild nd = nodes[i];
  const float l0[3] = {nd.a.x, nd.a.z, nd.c.x}, h0[3] = {nd.a.y, nd.a.w, nd.c.y};
  const float l1[3] = {nd.b.x, nd.b.z, nd.c.z}, h1[3] = {nd.b.y, nd.b.w, nd.c.w};
  quantize_node(l0, h0, l1, h1, nd.d.x, nd.d.y, q + 2 * (int64_t)i);
}

void launch_quantize_nodes(const NodeBuild* nodes, int n, uint4* q, hipStream_t s)
{
  if (n < 2) return;
  hipLaunchKernelGGL(k_quantize_nodes, dim3((unsigned)((n - 1 + 255) / 256)), dim3(256), 0, s, nodes, n, q);
}

// Fit one internal node from its (already final) children: child boxes into
// the traversal record (and its quantized record when qn is set), the union
// into node_box.
// A crossing node i with children (L, R): its 64-byte record's boxes, its
// quantized record (qn non-null) and its own box.
__device__ __forceinline__ void fit_cross(const float4* tris, NodeBuild* nodes, float* node_box, uint4* qn, int i,
                                          int L, int R)
{
  NodeBuild* nd = &nodes[i];
  float l0[3], h0[3], l1[3], h1[3];
  child_box(tris, node_box, L, l0, h0);
  child_box(tris, node_box, R, l1, h1);
  nd->a = make_float4(l0[0], h0[0], l0[1], h0[1]);
  nd->b = make_float4(l1[0], h1[0], l1[1], h1[1]);
  nd->c = make_float4(l0[2], h0[2], l1[2], h1[2]);
  if (qn) quantize_node(l0, h0, l1, h1, L, R, qn + 2 * (int64_t)i);
  float* nb = node_box + 6 * (int64_t)i;
  nb[0] = fminf(l0[0], l1[0]); nb[1] = fminf(l0[1], l1[1]); nb[2] = fminf(l0[2], l1[2]);
  nb[3] = fmaxf(h0[0], h1[0]); nb[4] = fmaxf(h0[1], h1[1]); nb[5] = fmaxf(h0[2], h1[2]);
}

// ---------------------------------------------------------------------------
// Treelet fit.  A Karras node covers the leaves [min(i, j), max(i, j)] (i is one
// end of its range).  The nodes covering at most kSpan leaves whose parent covers
// more are the roots of treelets, which partition the hierarchy below the "top"
// nodes (more than kSpan leaves, a fraction of a percent of all).  A workgroup per
// treelet finishes it from LDS: the topology of the node indices in its leaf
// range, its leaf boxes, the boxes it fits and arrival counters; each leaf walks
// up, the first of the two arrivals at a node stops, the second fits it
// (workgroup-scope ordering).  A walk ends at the first node whose range leaves
// the treelet's.  The top nodes follow, deepest level first.  (Cutting the leaves
// into fixed spans of kSpan instead leaves every ancestor of a span boundary to
// the level pass: 7x the nodes and twice the levels.)
// ---------------------------------------------------------------------------
static constexpr int kSpan = 1024;
static constexpr size_t kSpanLds = (size_t)kSpan * (sizeof(int4) + 12 * sizeof(float) + sizeof(uint32_t) + 1);

__device__ __forceinline__ bool node_inside(int i, int j, int s, int e)
{
  return min(i, j) >= s && max(i, j) < e;
}

__device__ __forceinline__ int node_leaves(int i, int j) { return (i > j ? i - j : j - i) + 1; }

__device__ __forceinline__ void span_child_box(const float* lb, const float* ib, int c, int s, float* lo,
                                               float* hi)
{
  const float* b = c < 0 ? lb + (~c - s) : ib + (c - s);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = b[a * kSpan];
    hi[a] = b[(3 + a) * kSpan];
  }
}

// One thread per leaf of the span.  FROM_REC (refits): the leaf-order triangles
// are first made from the triangle records and the moved instances (the
// flatten's transform) and written to tris.
template <bool FROM_REC>
__global__ __launch_bounds__(kSpan) void k_fit_span(SceneK sc, const uint32_t* ids, float4* tris, NodeBuild* nodes,
                                                    float* node_box, uint4* qn, const int4* topo,
                                                    const int32_t* leaf_parent, const int2* units, int n,
                                                    uint8_t* uheight, int full)
{
  extern __shared__ int4 span_lds[];
  int4* nd = span_lds;                                       // [kSpan] topology of nodes s..
  float* lb = reinterpret_cast<float*>(nd + kSpan);          // [6][kSpan] leaf boxes
  float* ib = lb + 6 * kSpan;                                // [6][kSpan] boxes fitted here
  uint32_t* cnt = reinterpret_cast<uint32_t*>(ib + 6 * kSpan);  // [kSpan] arrivals
  uint8_t* hg = reinterpret_cast<uint8_t*>(cnt + kSpan);      // [kSpan] heights of the fitted nodes
  const int2 unit = units[blockIdx.x];  // the treelet's leaf range [x, y], at most kSpan leaves
  const int s = unit.x;
  const int e = unit.y + 1;
  const int ni = min(e, n - 1) - s;  // node indices in the range (the treelet's and maybe an ancestor)
  const int leaf = s + (int)threadIdx.x;
  int p = leaf < e ? leaf_parent[leaf] : -1;
  for (int k = threadIdx.x; k < kSpan; k += blockDim.x) cnt[k] = 0u;
  for (int k = threadIdx.x; k < ni; k += blockDim.x) nd[k] = topo[s + k];
  for (int k = threadIdx.x; k < e - s; k += blockDim.x) {
    const int64_t p = s + k;
    float4 v0, v1, v2;
    if (FROM_REC) {
      const uint32_t id = ids[p];
      const float4* r = sc.tri_rec + 4 * (int64_t)id;
      const float4 ra = r[0], rb = r[1], rc = r[2], rd = r[3];
      const float* m = sc.inst[__float_as_uint(rd.x)].transform;
      const f3 w0 = xform_point(m, mk(ra.x, ra.y, ra.z));
      const f3 w1 = xform_point(m, mk(rb.x, rb.y, rb.z));
      const f3 w2 = xform_point(m, mk(rc.x, rc.y, rc.z));
      v0 = make_float4(w0.x, w0.y, w0.z, __uint_as_float(id));
      v1 = make_float4(w1.x, w1.y, w1.z, 0.0f);
      v2 = make_float4(w2.x, w2.y, w2.z, 0.0f);
      tris[3 * p] = v0;
      tris[3 * p + 1] = v1;
      tris[3 * p + 2] = v2;
    } else {
      v0 = tris[3 * p];
      v1 = tris[3 * p + 1];
      v2 = tris[3 * p + 2];
    }
    lb[0 * kSpan + k] = fminf(fminf(v0.x, v1.x), v2.x);
    lb[1 * kSpan + k] = fminf(fminf(v0.y, v1.y), v2.y);
    lb[2 * kSpan + k] = fminf(fminf(v0.z, v1.z), v2.z);
    lb[3 * kSpan + k] = fmaxf(fmaxf(v0.x, v1.x), v2.x);
    lb[4 * kSpan + k] = fmaxf(fmaxf(v0.y, v1.y), v2.y);
    lb[5 * kSpan + k] = fmaxf(fmaxf(v0.z, v1.z), v2.z);
  }
  __syncthreads();
  {
    while (p >= s && p < s + ni) {
      const int4 d = nd[p - s];
      if (!node_inside(p, d.w, s, e)) break;  // crossing: fitted later
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      const uint32_t old = atomicAdd(&cnt[p - s], 1u);
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (old == 0u) break;  // the sibling subtree is not finished yet
      float l0[3], h0[3], l1[3], h1[3];
      span_child_box(lb, ib, d.x, s, l0, h0);
      span_child_box(lb, ib, d.y, s, l1, h1);
      if (full) {  // the 64-byte record's boxes (a scene that traverses quantized records skips them)
        NodeBuild* q = &nodes[p];
        q->a = make_float4(l0[0], h0[0], l0[1], h0[1]);
        q->b = make_float4(l1[0], h1[0], l1[1], h1[1]);
        q->c = make_float4(l0[2], h0[2], l1[2], h1[2]);
      }
      float u[6];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        u[a] = fminf(l0[a], l1[a]);
        u[3 + a] = fmaxf(h0[a], h1[a]);
      }
#pragma unroll
      for (int a = 0; a < 6; ++a) ib[a * kSpan + (p - s)] = u[a];
      const uint8_t h = (uint8_t)(1 + max(d.x < 0 ? 0 : (int)hg[d.x - s], d.y < 0 ? 0 : (int)hg[d.y - s]));
      hg[p - s] = h;
      // the treelet root: its box goes to global memory for its top parent, its height
      // for the depth of the tree
      const int up = d.z;
      if (!(up >= s && up < s + ni && node_inside(up, nd[up - s].w, s, e))) {
        float* nb = node_box + 6 * (int64_t)p;
#pragma unroll
        for (int a = 0; a < 6; ++a) nb[a] = u[a];
        if (uheight) uheight[blockIdx.x] = h;
      }
      p = up;  // -1 past the root
    }
  }
  if (qn) {
    // the quantized records of the treelet's nodes, all at once from the boxes in
    // LDS once every walk has ended (off the walks' serial path)
    __syncthreads();
    for (int k = threadIdx.x; k < ni; k += blockDim.x) {
      const int4 d = nd[k];
      if (!node_inside(s + k, d.w, s, e)) continue;
      float l0[3], h0[3], l1[3], h1[3];
      span_child_box(lb, ib, d.x, s, l0, h0);
      span_child_box(lb, ib, d.y, s, l1, h1);
      quantize_node(l0, h0, l1, h1, d.x, d.y, qn + 2 * (int64_t)(s + k));
    }
  }
}

// Treelet roots (flag[i], i < m), top nodes (tflag[i], i < m) and the leaves
// hanging directly from a top node (sflag[k], k < n); the last slot of each stays
// 0 so that the exclusive scans end in the totals.
__global__ __launch_bounds__(256) void k_unit_flags(const int4* topo, const int32_t* leaf_parent, int n,
                                                    uint32_t* flag, uint32_t* sflag, uint32_t* tflag)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = n - 1;
  if (k < m) {
    const int i = (int)k;
    const int4 t = topo[i];
    const bool fits = node_leaves(i, t.w) <= kSpan;
    const bool parent_top = t.z < 0 || node_leaves(t.z, topo[t.z].w) > kSpan;
    flag[i] = (fits && parent_top) ? 1u : 0u;
    tflag[i] = fits ? 0u : 1u;
  } else if (k == m) {
    flag[m] = 0u;
    tflag[m] = 0u;
  }
  if (k < n) {
    const int p = leaf_parent[k];
    sflag[k] = node_leaves(p, topo[p].w) > kSpan ? 1u : 0u;
  } else if (k == n) {
    sflag[n] = 0u;
  }
}

__global__ __launch_bounds__(256) void k_unit_scatter(const int4* topo, int n, const uint32_t* pos,
                                                      const uint32_t* spos, const uint32_t* tpos, int2* units,
                                                      int32_t* unit_root, int32_t* stray, int32_t* top)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n - 1) {
    const int i = (int)k;
    if (pos[k + 1] != pos[k]) {
      const int j = topo[i].w;
      units[pos[k]] = make_int2(min(i, j), max(i, j));
      unit_root[pos[k]] = i;
    }
    if (tpos[k + 1] != tpos[k]) top[tpos[k]] = i;
  }
  if (k < n && spos[k + 1] != spos[k]) stray[spos[k]] = (int32_t)k;
}

// Depth of every top node (its ancestors are all top nodes): the walk to the root.
__global__ __launch_bounds__(256) void k_top_depth(const int32_t* top, const uint32_t* ntop, const int4* topo,
                                                   uint8_t* ndepth)
{
  const uint32_t nt = *ntop;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) {
    const int i = top[k];
    int d = 0, p = topo[i].z;
    while (p >= 0 && d < kMaxLevels - 1) {
      ++d;
      p = topo[p].z;
    }
    ndepth[i] = (uint8_t)d;
  }
}

// The top nodes in depth order (a counting sort in one workgroup; the order
// within a level is immaterial): cross, and out[d] = the start of depth d in it
// for d <= kMaxLevels, out[kMaxLevels + 1] = the number of depth levels.
__global__ __launch_bounds__(1024) void k_top_sort(const int32_t* top, const uint32_t* ntop, const uint8_t* ndepth,
                                                   int32_t* cross, uint32_t* out)
{
  __shared__ uint32_t cnt[kMaxLevels], cur[kMaxLevels];
  __shared__ uint32_t maxd;
  const uint32_t nt = *ntop;
  if (threadIdx.x < kMaxLevels) cnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) maxd = 0u;
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) {
    const uint32_t d = ndepth[top[k]];
    atomicAdd(&cnt[d], 1u);
    atomicMax(&maxd, d + 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0u;
    for (int d = 0; d < kMaxLevels; ++d) {
      cur[d] = run;
      out[d] = run;
      run += cnt[d];
    }
    out[kMaxLevels] = run;
    out[kMaxLevels + 1] = maxd;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) {
    const int i = top[k];
    cross[atomicAdd(&cur[ndepth[i]], 1u)] = i;
  }
}

// Levels of the tree (internal nodes on the longest root-leaf path): over the
// treelets, the depth of the root plus the treelet's height; over the top nodes,
// their depth + 1.
__global__ __launch_bounds__(256) void k_tree_depth(const int32_t* unit_root, const uint8_t* uheight, int nunits,
                                                    const int32_t* top, const uint32_t* ntop, const int4* topo,
                                                    const uint8_t* ndepth, uint32_t* out)
{
  const uint32_t nt = *ntop;
  uint32_t best = 0u;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (int64_t)nunits + nt;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k < nunits) {
      const int p = topo[unit_root[k]].z;
      best = max(best, (p < 0 ? 0u : (uint32_t)ndepth[p] + 1u) + (uint32_t)uheight[k]);
    } else {
      best = max(best, (uint32_t)ndepth[top[k - nunits]] + 1u);
    }
  }
  atomicMax(out, best);
}

// Refit: the leaf-order triangles of the leaves outside every treelet.
__global__ __launch_bounds__(256) void k_refit_stray(SceneK sc, const uint32_t* ids, const int32_t* stray, int nstray,
                                                     float4* tris)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nstray) return;
  const int64_t p = stray[k];
  const uint32_t id = ids[p];
  const float4* r = sc.tri_rec + 4 * (int64_t)id;
  const float4 ra = r[0], rb = r[1], rc = r[2], rd = r[3];
  const float* mt = sc.inst[__float_as_uint(rd.x)].transform;
  const f3 w0 = xform_point(mt, mk(ra.x, ra.y, ra.z));
  const f3 w1 = xform_point(mt, mk(rb.x, rb.y, rb.z));
  const f3 w2 = xform_point(mt, mk(rc.x, rc.y, rc.z));
  tris[3 * p] = make_float4(w0.x, w0.y, w0.z, __uint_as_float(id));
  tris[3 * p + 1] = make_float4(w1.x, w1.y, w1.z, 0.0f);
  tris[3 * p + 2] = make_float4(w2.x, w2.y, w2.z, 0.0f);
}

// One depth level of the crossing nodes: every child is a leaf, inside a span
// (fitted by k_fit_span) or a crossing node of a deeper, finished level.  kids:
// the children of order[k] (gathered once per build), so that a level's loads
// wait on one round trip (the child boxes), not three.
__global__ __launch_bounds__(256) void k_fit_level(const float4* tris, NodeBuild* nodes,
                                                   float* node_box, uint4* qn, const int32_t* order, const int2* kids,
                                                   int begin, int count)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int2 c = kids[begin + k];
  fit_cross(tris, nodes, node_box, qn, order[begin + k], c.x, c.y);
}

// the children of the crossing nodes, in their order
__global__ __launch_bounds__(256) void k_cross_kids(const int32_t* order, const int4* topo, int total, int2* kids)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int4 t = topo[order[k]];
  kids[k] = make_int2(t.x, t.y);
}

// The top levels (at most 2^d nodes at depth d) in one workgroup, deepest first.
struct TopLevels {
  int32_t begin[kTopLevels];
  int32_t count[kTopLevels];
  int32_t num;
};

__global__ __launch_bounds__(1024) void k_fit_top(const float4* tris, NodeBuild* nodes,
                                                  float* node_box, uint4* qn, const int32_t* order, const int2* kids,
                                                  TopLevels t)
{
  for (int l = t.num - 1; l >= 0; --l) {
    for (int k = threadIdx.x; k < t.count[l]; k += blockDim.x) {
      const int2 c = kids[t.begin[l] + k];
      fit_cross(tris, nodes, node_box, qn, order[t.begin[l] + k], c.x, c.y);
    }
    __threadfence_block();
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int flatten_partial_words(int ntri) { return 6 * (int)blocks_for(ntri > 0 ? ntri : 1, 256); }

void launch_flatten(const SceneK& sc, float* world, uint32_t* tri_inst, float4* tri_rec, uint32_t* part,
                    uint32_t* cbounds, bool write_inst, hipStream_t s)
{
  if (sc.ntri <= 0) return;
  const int nb = (int)blocks_for(sc.ntri, 256);
  hipLaunchKernelGGL(k_flatten, dim3(nb), dim3(256), 0, s, sc, world, tri_inst, tri_rec,
                     cbounds ? part : (uint32_t*)nullptr, (int32_t)write_inst);
  if (cbounds) hipLaunchKernelGGL(k_bounds_reduce, dim3(1), dim3(256), 0, s, (const uint32_t*)part, nb, cbounds);
}

void launch_morton(const float* world, int n, const uint32_t* cbounds, uint64_t* keys,
                   uint32_t* ids, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, s, world, n, cbounds, keys,
                     ids);
}

void launch_karras(const uint64_t* keys, int n, NodeBuild* nodes, int32_t* leaf_parent, int4* topo,
                   hipStream_t s)
{
  if (n < 2) return;
  hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, keys, n, nodes,
                     leaf_parent, topo);
}

void launch_refit_tris(const SceneK& sc, const uint32_t* ids, int n, float4* tris, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_refit_tris, dim3(blocks_for(n, 256)), dim3(256), 0, s, sc, ids, n, tris);
}

void launch_world_from_tris(const float4* tris, int n, float* world, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_world_from_tris, dim3(blocks_for(n, 256)), dim3(256), 0, s, tris, n, world);
}

void launch_reorder(const float* world, const uint32_t* ids, int n, float4* tris, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_reorder, dim3(blocks_for(n, 256)), dim3(256), 0, s, world, ids, n, tris);
}

int launch_fit_plan(const int4* topo, const int32_t* leaf_parent, int n, uint32_t* scratch, uint32_t* uscratch,
                    uint32_t* info, uint8_t* ndepth, int32_t* cross, int32_t* top, int2* units, int32_t* unit_root,
                    int32_t* stray, std::vector<int32_t>* cross_begin, int* nunits, int* nstray, hipStream_t s)
{
  cross_begin->clear();
  *nunits = 0;
  *nstray = 0;
  if (n < 2) return 0;
  const int64_t m = n - 1;
  uint32_t* upos = uscratch;              // m + 1 words
  uint32_t* spos = uscratch + n;          // n + 1 words
  uint32_t* tpos = uscratch + 2 * n + 1;  // m + 1 words
  hipLaunchKernelGGL(k_unit_flags, dim3(blocks_for((int64_t)n + 1, 256)), dim3(256), 0, s, topo, leaf_parent, n,
                     upos, spos, tpos);
  exclusive_scan(upos, m + 1, scratch, s);
  exclusive_scan(spos, (int64_t)n + 1, scratch, s);
  exclusive_scan(tpos, m + 1, scratch, s);
  hipLaunchKernelGGL(k_unit_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, topo, n, (const uint32_t*)upos,
                     (const uint32_t*)spos, (const uint32_t*)tpos, units, unit_root, stray, top);
  const uint32_t* ntop = tpos + m;
  hipLaunchKernelGGL(k_top_depth, dim3(1024), dim3(256), 0, s, (const int32_t*)top, ntop, topo, ndepth);
  hipLaunchKernelGGL(k_top_sort, dim3(1), dim3(1024), 0, s, (const int32_t*)top, ntop, (const uint8_t*)ndepth,
                     cross, info);
  // info: [0, kMaxLevels] top-level offsets, [kMaxLevels + 1] levels, then the treelet
  // and stray-leaf counts
  if (hipMemcpyAsync(info + kMaxLevels + 2, upos + m, 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipMemcpyAsync(info + kMaxLevels + 3, spos + n, 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return -1;
  uint32_t h[kMaxLevels + 4];
  if (hipMemcpyAsync(h, info, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
  if (hipStreamSynchronize(s) != hipSuccess) return -1;
  const int levels = (int)h[kMaxLevels + 1];
  for (int d = 0; d <= levels && d <= kMaxLevels; ++d) cross_begin->push_back((int32_t)h[d]);
  if (cross_begin->size() < 2) cross_begin->push_back(0);
  *nunits = (int)h[kMaxLevels + 2];
  *nstray = (int)h[kMaxLevels + 3];
  return 0;
}

int launch_tree_depth(const int4* topo, int n, const int32_t* unit_root, const uint8_t* uheight, int nunits,
                      const int32_t* top, const uint32_t* uscratch, const uint8_t* ndepth, uint32_t* out,
                      hipStream_t s)
{
  if (n < 2) return 0;
  if (hipMemsetAsync(out, 0, 4, s) != hipSuccess) return -1;
  const uint32_t* ntop = uscratch + 2 * n + 1 + (n - 1);
  hipLaunchKernelGGL(k_tree_depth, dim3(256), dim3(256), 0, s, unit_root, uheight, nunits, top, ntop, topo, ndepth,
                     out);
  return 0;
}

void launch_fit(float4* tris, int n, NodeBuild* nodes, float* node_box, uint4* qnodes, const int4* topo,
                const int32_t* leaf_parent, const int2* units, int nunits, const int32_t* stray, int nstray,
                const int32_t* cross, const int2* cross_kids, const std::vector<int32_t>& cross_begin, const SceneK* refit,
                const uint32_t* ids, uint8_t* uheight, hipStream_t s, bool full_boxes)
{
  if (n < 2 || cross_begin.size() < 2) {
    if (refit && n == 1) launch_refit_tris(*refit, ids, n, tris, s);
    return;
  }
  static bool lds_set = false;
  if (!lds_set) {
    (void)hipFuncSetAttribute((const void*)k_fit_span<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kSpanLds);
    (void)hipFuncSetAttribute((const void*)k_fit_span<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kSpanLds);
    lds_set = true;
  }
  const SceneK none{};
  if (refit && nstray > 0)
    hipLaunchKernelGGL(k_refit_stray, dim3(blocks_for(nstray, 256)), dim3(256), 0, s, *refit, ids, stray, nstray,
                       tris);
  if (nunits > 0) {
    if (refit)
      hipLaunchKernelGGL(k_fit_span<true>, dim3((unsigned)nunits), dim3(kSpan), kSpanLds, s, *refit, ids, tris,
                         nodes, node_box, qnodes, topo, leaf_parent, units, n, uheight, full_boxes ? 1 : 0);
    else
      hipLaunchKernelGGL(k_fit_span<false>, dim3((unsigned)nunits), dim3(kSpan), kSpanLds, s, none,
                         (const uint32_t*)nullptr, tris, nodes, node_box, qnodes, topo, leaf_parent, units, n,
                         uheight, full_boxes ? 1 : 0);
  }
  // the crossing nodes, deepest level first: a launch per level below the top
  // ones, then the top levels in one workgroup
  const int levels = (int)cross_begin.size() - 1;
  for (int d = levels - 1; d >= kTopLevels; --d) {
    const int cnt = cross_begin[d + 1] - cross_begin[d];
    if (cnt > 0)
      hipLaunchKernelGGL(k_fit_level, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, (const float4*)tris, nodes,
                         node_box, qnodes, cross, cross_kids, cross_begin[d], cnt);
  }
  TopLevels t{};
  t.num = levels < kTopLevels ? levels : kTopLevels;
  int top = 0;
  for (int d = 0; d < t.num; ++d) {
    t.begin[d] = cross_begin[d];
    t.count[d] = cross_begin[d + 1] - cross_begin[d];
    top += t.count[d];
  }
  if (top > 0)
    hipLaunchKernelGGL(k_fit_top, dim3(1), dim3(1024), 0, s, (const float4*)tris, nodes, node_box, qnodes, cross,
                       cross_kids, t);
}

void launch_cross_kids(const int32_t* cross, const int4* topo, int total, int2* kids, hipStream_t s)
{
  if (total > 0)
    hipLaunchKernelGGL(k_cross_kids, dim3(blocks_for(total, 256)), dim3(256), 0, s, cross, topo, total, kids);
}


}  // namespace ane
