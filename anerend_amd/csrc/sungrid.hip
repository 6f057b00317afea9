// sungrid.hip -- the candidate triangles of the probe shadow rays of one frame.
//
// Every shadow ray of IrradianceProbeRT goes from a hit toward the sun
// (rchit:44-66, direction normalize(-lightDir)), so all of them are parallel.
// Projected along that direction a ray is a single point of the plane, and a
// triangle can stop it only if the triangle's projection contains that point.
// The grid bins the projected triangles (their bounding rectangles, widened by
// a margin far above the rounding of the projection and of the watertight test)
// into cells; a shadow ray then runs the same watertight any-hit test on the
// triangles of its cell instead of traversing the hierarchy.  Any of them hitting
// within [t_min, t_max] is exactly the any-hit answer of the traversal, which is
// conservative too, so the shadow decisions are unchanged.
//
// Projection: with kz the largest axis of the direction d and (kx, ky) the other
// two (the watertight test's permutation), P(p) = (p[kx] - sx p[kz], p[ky] - sy p[kz])
// with sx = d[kx] / d[kz], sy = d[ky] / d[kz]; P is constant along d.
//
// Build (on the stream, no host synchronization): projected bounds and the
// largest coordinate magnitude (margin), grid placement, per-cell counts,
// exclusive scan, scatter.  The grid is marked invalid (the traversal then
// handles shadow rays) when the entries exceed the capacity, a triangle covers
// more than kMaxCellsPerTri cells, or a vertex is not finite.
#include "common.hpp"
#include "kernels.hpp"

namespace ane {

static constexpr int kMaxCellsPerTri = 1 << 16;
#ifndef ANE_SUNGRID_MAX_CANDIDATES
#define ANE_SUNGRID_MAX_CANDIDATES 32.0f  // estimated candidates per shadow ray above which traversal is kept
#endif

// ordered encoding of floats for unsigned atomicMin/atomicMax
__device__ __forceinline__ uint32_t f2ord(float f)
{
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o)
{
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ void sun_project(const SunGridParams& g, f3 p, float& X, float& Y)
{
  X = comp(p, g.kx) - g.sx * comp(p, g.kz);
  Y = comp(p, g.ky) - g.sy * comp(p, g.kz);
}

static constexpr int kBoundBlocks = 256;

// per block, [6][kBoundBlocks]: min X, min Y, max X, max Y, max |coordinate|,
// non-finite seen (ordered encodings); k_sun_params reduces the blocks
__global__ __launch_bounds__(256) void k_sun_bounds(const float4* tris, int32_t ntri, SunGridParams g,
                                                    uint32_t* part)
{
  __shared__ uint32_t red[6][256];
  uint32_t mnx = 0xFFFFFFFFu, mny = 0xFFFFFFFFu, mxx = 0u, mxy = 0u, mr = 0u, bad = 0u;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)ntri * 3;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float4 v = tris[i];
    const f3 p = mk(v.x, v.y, v.z);
    if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) {
      bad = 1u;
      continue;
    }
    float X, Y;
    sun_project(g, p, X, Y);
    mnx = min(mnx, f2ord(X));
    mny = min(mny, f2ord(Y));
    mxx = max(mxx, f2ord(X));
    mxy = max(mxy, f2ord(Y));
    mr = max(mr, f2ord(fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fmaxf(fabsf(X), fabsf(Y))))));
  }
  const int t = threadIdx.x;
  red[0][t] = mnx; red[1][t] = mny; red[2][t] = mxx; red[3][t] = mxy; red[4][t] = mr; red[5][t] = bad;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = min(red[0][t], red[0][t + s]);
      red[1][t] = min(red[1][t], red[1][t + s]);
      red[2][t] = max(red[2][t], red[2][t + s]);
      red[3][t] = max(red[3][t], red[3][t + s]);
      red[4][t] = max(red[4][t], red[4][t + s]);
      red[5][t] = max(red[5][t], red[5][t + s]);
    }
    __syncthreads();
  }
  if (t < 6) part[t * kBoundBlocks + blockIdx.x] = red[t][0];
}

// placement: margin 2e-5 of the largest coordinate magnitude (the rounding of
// the projection and of the watertight test is below 1e-6 of it); cells of the
// widened bounds
__global__ __launch_bounds__(kBoundBlocks) void k_sun_params(SunGridParams g, const uint32_t* part,
                                                            SunGridParams* out)
{
  __shared__ uint32_t red[6][kBoundBlocks];
  const int t = threadIdx.x;
  for (int k = 0; k < 6; ++k) red[k][t] = part[k * kBoundBlocks + t];
  __syncthreads();
  for (int s = kBoundBlocks / 2; s > 0; s >>= 1) {
    if (t < s) {
      red[0][t] = min(red[0][t], red[0][t + s]);
      red[1][t] = min(red[1][t], red[1][t + s]);
      for (int k = 2; k < 6; ++k) red[k][t] = max(red[k][t], red[k][t + s]);
    }
    __syncthreads();
  }
  if (t != 0) return;
  const uint32_t acc[6] = {red[0][0], red[1][0], red[2][0], red[3][0], red[4][0], red[5][0]};
  SunGridParams r = g;
  const bool empty = acc[0] == 0xFFFFFFFFu;
  const float R = empty ? 0.0f : ord2f(acc[4]);
  r.margin = 2e-5f * R + 1e-30f;
  const float x0 = empty ? 0.0f : ord2f(acc[0]) - r.margin;
  const float y0 = empty ? 0.0f : ord2f(acc[1]) - r.margin;
  const float x1 = empty ? 1.0f : ord2f(acc[2]) + r.margin;
  const float y1 = empty ? 1.0f : ord2f(acc[3]) + r.margin;
  r.x0 = x0;
  r.y0 = y0;
  r.icw = (x1 > x0) ? (float)g.gx / (x1 - x0) : 0.0f;
  r.ich = (y1 > y0) ? (float)g.gy / (y1 - y0) : 0.0f;
  r.valid = (acc[5] == 0u && isfinite(r.icw) && isfinite(r.ich)) ? 1u : 0u;
  *out = r;
}

struct CellRange {
  int cx0, cx1, cy0, cy1;
  float px[3], py[3];  // projected vertices
};

// cells of triangle p's widened projected rectangle (inclusive)
__device__ __forceinline__ CellRange sun_tri_cells(const SunGridParams& g, const float4* tris, int64_t p)
{
  CellRange c;
  float xl = 0.0f, xh = 0.0f, yl = 0.0f, yh = 0.0f;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float4 v = tris[3 * p + j];
    float X, Y;
    sun_project(g, mk(v.x, v.y, v.z), X, Y);
    c.px[j] = X;
    c.py[j] = Y;
    xl = j ? fminf(xl, X) : X;
    xh = j ? fmaxf(xh, X) : X;
    yl = j ? fminf(yl, Y) : Y;
    yh = j ? fmaxf(yh, Y) : Y;
  }
  c.cx0 = sun_cell(xl - g.margin, g.x0, g.icw, g.gx);
  c.cx1 = sun_cell(xh + g.margin, g.x0, g.icw, g.gx);
  c.cy0 = sun_cell(yl - g.margin, g.y0, g.ich, g.gy);
  c.cy1 = sun_cell(yh + g.margin, g.y0, g.ich, g.gy);
  return c;
}

// Whether the projected triangle can come within the margin of cell (x, y):
// separating axes are the triangle's edge normals (the rectangle's own axes hold
// by the cell range).  The cell rectangle is widened by the margin, and border
// cells reach to infinity on their outer side (points outside the grid clamp to
// them); the tolerance margin * (|ex| + |ey|) exceeds the rounding of the dot
// products.  Degenerate edges separate nothing.
__device__ __forceinline__ bool sun_tri_overlaps_cell(const SunGridParams& g, const CellRange& c, int x, int y)
{
  const float big = 3.0e38f;
  const float cw = 1.0f / g.icw, ch = 1.0f / g.ich;
  const float rx0 = x == 0 ? -big : g.x0 + (float)x * cw - g.margin;
  const float rx1 = x == g.gx - 1 ? big : g.x0 + (float)(x + 1) * cw + g.margin;
  const float ry0 = y == 0 ? -big : g.y0 + (float)y * ch - g.margin;
  const float ry1 = y == g.gy - 1 ? big : g.y0 + (float)(y + 1) * ch + g.margin;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int k = j == 2 ? 0 : j + 1, o = j == 0 ? 2 : j - 1;
    const float ex = c.px[k] - c.px[j], ey = c.py[k] - c.py[j];
    const float nx = -ey, ny = ex;
    const float side = nx * (c.px[o] - c.px[j]) + ny * (c.py[o] - c.py[j]);
    if (!(side != 0.0f)) continue;  // degenerate (or NaN): no separation
    const float s = side > 0.0f ? 1.0f : -1.0f;
    // the rectangle corner furthest toward the triangle's side of this edge
    const float cx = (s * nx > 0.0f) ? rx1 : rx0, cy = (s * ny > 0.0f) ? ry1 : ry0;
    const float reach = s * (nx * (cx - c.px[j]) + ny * (cy - c.py[j]));
    if (reach < -g.margin * (fabsf(ex) + fabsf(ey))) return false;
  }
  return true;
}

// Every (triangle, cell of its rectangle) pair of a wave's 64 triangles, spread
// over the wave's lanes 64 at a time: f(ok, j, cj, x, y) runs in every lane of
// the wave, round after round; ok says whether the lane has a pair in the round,
// that of lane j's triangle cj and cell (x, y).  The pairs of one triangle take
// consecutive lanes (j does not decrease over the lanes that have pairs, which
// come first).  Lanes whose `has` is false have no pairs.  Neighbouring lanes
// mostly take neighbouring cells of one triangle, so the per-cell atomics of a
// wave rarely meet.
template <class F>
__device__ __forceinline__ void sun_wave_pairs(const CellRange& c, bool has, F&& f)
{
  const int lane = threadIdx.x & 63;
  const int w = has ? c.cx1 - c.cx0 + 1 : 0;
  const uint32_t r = has ? (uint32_t)w * (uint32_t)(c.cy1 - c.cy0 + 1) : 0u;
  uint32_t incl = r;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  const uint32_t total = __shfl(incl, 63, 64);
  const uint32_t excl = incl - r;
  for (uint32_t t0 = 0; t0 < total; t0 += 64u) {
    const uint32_t t = t0 + (uint32_t)lane;
    const uint32_t tc = t < total ? t : total - 1u;
    int j = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const uint32_t v = __shfl(incl, j + step - 1, 64);
      if (v <= tc) j += step;
    }
    CellRange cj;
    cj.cx0 = __shfl(c.cx0, j, 64);
    cj.cy0 = __shfl(c.cy0, j, 64);
    cj.cx1 = __shfl(c.cx1, j, 64);
    cj.cy1 = __shfl(c.cy1, j, 64);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cj.px[k] = __shfl(c.px[k], j, 64);
      cj.py[k] = __shfl(c.py[k], j, 64);
    }
    const uint32_t q = tc - __shfl(excl, j, 64);
    const uint32_t wj = (uint32_t)(cj.cx1 - cj.cx0 + 1);
    const int y = cj.cy0 + (int)(q / wj), x = cj.cx0 + (int)(q % wj);
    f(t < total, j, cj, x, y);
  }
}

// Per-block aggregation of the per-cell atomics.  The triangles are in leaf
// (Morton) order, so a block's 256 triangles cover a compact window of cells,
// each cell several times.  When the window's bounding rectangle has at most
// kWin cells, the block counts its pairs in LDS and adds one global atomic per
// touched cell; otherwise every pair takes its own global atomic.
#ifndef ANE_SUNGRID_WIN
#define ANE_SUNGRID_WIN 1024  // 512 / 2048 swept: build 0.330 / 0.244 ms against 0.245 on 1M triangles
#endif
static constexpr int kWin = ANE_SUNGRID_WIN;

struct CellWindow {
  int x0, y0, w;  // first cell and width of the rectangle
  int n;          // cells in it, or 0: the block does not aggregate
};

// The block's window over the cell ranges of its triangles (has: the thread's
// triangle takes part).  All threads of the block call it.
__device__ __forceinline__ CellWindow sun_block_window(const CellRange& c, bool has)
{
  __shared__ int red[4][4];
  int v[4] = {has ? c.cx0 : INT_MAX, has ? c.cy0 : INT_MAX, has ? -c.cx1 : INT_MAX, has ? -c.cy1 : INT_MAX};
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = min(v[k], __shfl_xor(v[k], d, 64));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) red[k][wv] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = min(min(red[k][0], red[k][1]), min(red[k][2], red[k][3]));
  __syncthreads();  // red is reused by the next call
  CellWindow w{0, 0, 1, 0};
  if (v[0] == INT_MAX) return w;  // no triangle in the block
  const int64_t ww = (int64_t)(-v[2]) - v[0] + 1, wh = (int64_t)(-v[3]) - v[1] + 1;
  if (ww * wh > kWin) return w;
  w.x0 = v[0];
  w.y0 = v[1];
  w.w = (int)ww;
  w.n = (int)(ww * wh);
  return w;
}

// the validity flag as one value for the whole block (k_sun_count clears it
// while other blocks read it)
__device__ __forceinline__ bool sun_block_valid(const SunGridParams* gp)
{
  __shared__ uint32_t v;
  if (threadIdx.x == 0) v = gp->valid;
  __syncthreads();
  return v != 0u;
}

// An aggregating block also keeps its window's counts in tab (kWin per block,
// at most 256 per cell) for the scatter.
__global__ __launch_bounds__(256) void k_sun_count(const float4* tris, int32_t ntri, SunGridParams* gp,
                                                   uint32_t* cnt, uint16_t* tab)
{
  __shared__ uint32_t wc[kWin];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (!sun_block_valid(gp)) return;
  const SunGridParams g = *gp;
  CellRange c{};
  bool has = p < ntri;
  if (has) {
    c = sun_tri_cells(g, tris, p);
    const int64_t n = (int64_t)(c.cx1 - c.cx0 + 1) * (c.cy1 - c.cy0 + 1);
    if (n > kMaxCellsPerTri) {
      gp->valid = 0u;  // a triangle too large for the grid: the traversal takes the shadow rays
      has = false;
    }
  }
  const CellWindow win = sun_block_window(c, has);
  for (int i = threadIdx.x; i < win.n; i += 256) wc[i] = 0u;
  __syncthreads();
  sun_wave_pairs(c, has, [&](bool ok, int, const CellRange& cj, int x, int y) {
    if (ok && sun_tri_overlaps_cell(g, cj, x, y)) {
      if (win.n) atomicAdd(&wc[(y - win.y0) * win.w + (x - win.x0)], 1u);
      else atomicAdd(&cnt[(int64_t)y * g.gx + x], 1u);
    }
  });
  __syncthreads();
  uint16_t* bt = tab + (int64_t)blockIdx.x * kWin;
  for (int i = threadIdx.x; i < win.n; i += 256) {
    bt[i] = (uint16_t)wc[i];
    if (wc[i]) atomicAdd(&cnt[(int64_t)(win.y0 + i / win.w) * g.gx + win.x0 + i % win.w], wc[i]);
  }
}

// after the scan: start[g.gx * g.gy] entries in total; the cursors begin at the starts
__global__ __launch_bounds__(256) void k_sun_cursor(const uint32_t* start, int64_t ncell, uint32_t cap,
                                                    SunGridParams* gp, uint32_t* cur)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) cur[i] = start[i];
  if (i == 0 && start[ncell] > cap) gp->valid = 0u;
}

// The scatter of the candidate lists, and the candidate estimate from the same
// pairs: a shadow ray starts on a triangle and tests the candidates of the cell
// its origin falls in, one of the triangle's cells; per triangle the mean count
// of its cells (integers: the sums are exact in any order), then the mean over
// the triangles, summed per block here and over the blocks by k_sun_decide.
// An aggregating block (the same window as in k_sun_count, whose per-cell counts
// it reads back) reserves each cell's run with one global atomic on the cell's
// cursor, then places its pairs in the runs from LDS cursors; the others take a
// cursor atomic per pair.  Both reserve from the same cursors, so the runs never
// overlap.
__global__ __launch_bounds__(256) void k_sun_scatter(const float4* tris, int32_t ntri, const SunGridParams* gp,
                                                     const uint32_t* start, uint32_t* cur, uint32_t* list,
                                                     uint32_t cap, const uint16_t* tab, float* part)
{
  __shared__ float ts[256], tk[256];  // per lane's triangle: the counts of its cells, and their number
  __shared__ float rs[256], rn[256];
  __shared__ uint32_t wcur[kWin];  // an aggregating block's cursor in each window cell's run
  __shared__ float wtot[kWin];     // and the cells' entry counts (for the estimate)
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int t = threadIdx.x, wb = t & ~63;
  const SunGridParams g = *gp;
  if (!g.valid) {
    if (t == 0) part[2 * blockIdx.x] = part[2 * blockIdx.x + 1] = 0.0f;
    return;
  }
  const bool has = p < ntri;
  CellRange c{};
  if (has) c = sun_tri_cells(g, tris, p);
  ts[t] = 0.0f;
  tk[t] = 0.0f;
  const CellWindow win = sun_block_window(c, has);
  if (win.n) {
    const uint16_t* bt = tab + (int64_t)blockIdx.x * kWin;
    for (int i = t; i < win.n; i += 256) {
      const uint32_t n = bt[i];
      if (n) {
        const int64_t cell = (int64_t)(win.y0 + i / win.w) * g.gx + win.x0 + i % win.w;
        wcur[i] = atomicAdd(&cur[cell], n);
        wtot[i] = (float)(start[cell + 1] - start[cell]);
      }
    }
    __syncthreads();
  }
  const int64_t p0 = p - (t & 63);  // the wave's first triangle
  const int lane = t & 63;
  sun_wave_pairs(c, has, [&](bool ok, int j, const CellRange& cj, int x, int y) {
    float sc = 0.0f, sk = 0.0f;
    if (ok && sun_tri_overlaps_cell(g, cj, x, y)) {
      const int64_t cell = (int64_t)y * g.gx + x;
      const int l = (y - win.y0) * win.w + (x - win.x0);
      const uint32_t k = win.n ? atomicAdd(&wcur[l], 1u) : atomicAdd(&cur[cell], 1u);
      if (k < cap) list[k] = (uint32_t)(p0 + j);
      sc = win.n ? wtot[l] : (float)(start[cell + 1] - start[cell]);
      sk = 1.0f;
    }
    // per triangle, the sums over its lanes (a segmented scan: its lanes are
    // consecutive), added by its last lane
    const int key = ok ? j : -1;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const float uc = __shfl_up(sc, d, 64), uk = __shfl_up(sk, d, 64);
      const int kk = __shfl_up(key, d, 64);
      if (lane >= d && kk == key) {
        sc += uc;
        sk += uk;
      }
    }
    const int next = __shfl_down(key, 1, 64);
    if (ok && (lane == 63 || next != key)) {
      ts[wb + j] += sc;
      tk[wb + j] += sk;
    }
  });
  const float k = tk[t];
  rs[t] = (has && k > 0.0f) ? ts[t] / k : 0.0f;
  rn[t] = (has && k > 0.0f) ? 1.0f : 0.0f;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      rs[t] += rs[t + h];
      rn[t] += rn[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    part[2 * blockIdx.x] = rs[0];
    part[2 * blockIdx.x + 1] = rn[0];
  }
}

__global__ __launch_bounds__(256) void k_sun_decide(const float* part, int nblocks, SunGridParams* gp)
{
  __shared__ float rs[256], rn[256];
  const int t = threadIdx.x;
  float s = 0.0f, n = 0.0f;
  for (int b = t; b < nblocks; b += 256) {
    s += part[2 * b];
    n += part[2 * b + 1];
  }
  rs[t] = s;
  rn[t] = n;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      rs[t] += rs[t + h];
      rn[t] += rn[t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    const float est = rn[0] > 0.0f ? rs[0] / rn[0] : 0.0f;
    gp->est = est;
    if (!(est <= gp->max_est)) gp->valid = 0u;
  }
}

void sun_grid_setup(SunGridParams& g, const float* dir3, int gx, int gy)
{
  // the watertight test's axis permutation and shear for the direction (ray_setup)
  RayPre r;
  ray_setup(r, mk(0.0f, 0.0f, 0.0f), mk(dir3[0], dir3[1], dir3[2]));
  g.kx = r.kx;
  g.ky = r.ky;
  g.kz = r.kz;
  g.sx = r.sx;
  g.sy = r.sy;
  g.gx = gx;
  g.gy = gy;
  g.x0 = g.y0 = g.icw = g.ich = g.margin = 0.0f;
  g.valid = 0u;
  g.est = 0.0f;
  g.max_est = ANE_SUNGRID_MAX_CANDIDATES;
}

void launch_sun_grid_build(const SceneK& sc, const SunGridParams& g, SunGridParams* d_params, uint32_t* acc,
                           uint32_t* cnt, uint32_t* cur, uint32_t* list, uint32_t cap, uint32_t* scan_scratch,
                           float* est, hipStream_t s)
{
  const int64_t ncell = (int64_t)g.gx * g.gy;
  fill_u32(cnt, 0u, ncell + 1, s);
  hipLaunchKernelGGL(k_sun_bounds, dim3(kBoundBlocks), dim3(256), 0, s, sc.tris, sc.ntri, g, acc);
  hipLaunchKernelGGL(k_sun_params, dim3(1), dim3(kBoundBlocks), 0, s, g, (const uint32_t*)acc, d_params);
  const unsigned tb = (unsigned)((sc.ntri + 255) / 256);
  uint16_t* tab = reinterpret_cast<uint16_t*>(est + 2 * (int64_t)tb);  // after the estimate partials
  if (sc.ntri > 0)
    hipLaunchKernelGGL(k_sun_count, dim3(tb), dim3(256), 0, s, sc.tris, sc.ntri, d_params, cnt, tab);
  exclusive_scan(cnt, ncell + 1, scan_scratch, s);
  hipLaunchKernelGGL(k_sun_cursor, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                     (const uint32_t*)cnt, ncell, cap, d_params, cur);
  if (sc.ntri > 0) {
    hipLaunchKernelGGL(k_sun_scatter, dim3(tb), dim3(256), 0, s, sc.tris, sc.ntri,
                       (const SunGridParams*)d_params, (const uint32_t*)cnt, cur, list, cap, (const uint16_t*)tab,
                       est);
    hipLaunchKernelGGL(k_sun_decide, dim3(1), dim3(256), 0, s, (const float*)est, (int)tb, d_params);
  }
}

int64_t sun_grid_scratch_words(int32_t ntri)
{
  const int64_t nb = (std::max(ntri, 1) + 255) / 256;
  return 2 * nb + nb * (kWin / 2);  // estimate partials, then the blocks' window counts (16-bit)
}

}  // namespace ane
