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

__global__ __launch_bounds__(256) void k_sun_count(const float4* tris, int32_t ntri, SunGridParams* gp,
                                                   uint32_t* cnt)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const SunGridParams g = *gp;
  if (p >= ntri || !g.valid) return;
  const CellRange c = sun_tri_cells(g, tris, p);
  const int64_t n = (int64_t)(c.cx1 - c.cx0 + 1) * (c.cy1 - c.cy0 + 1);
  if (n > kMaxCellsPerTri) {
    gp->valid = 0u;  // a triangle too large for the grid: the traversal takes the shadow rays
    return;
  }
  for (int y = c.cy0; y <= c.cy1; ++y)
    for (int x = c.cx0; x <= c.cx1; ++x)
      if (sun_tri_overlaps_cell(g, c, x, y)) atomicAdd(&cnt[(int64_t)y * g.gx + x], 1u);
}

// after the scan: start[g.gx * g.gy] entries in total; the cursors begin at the starts
__global__ __launch_bounds__(256) void k_sun_cursor(const uint32_t* start, int64_t ncell, uint32_t cap,
                                                    SunGridParams* gp, uint32_t* cur)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncell) cur[i] = start[i];
  if (i == 0 && start[ncell] > cap) gp->valid = 0u;
}

__global__ __launch_bounds__(256) void k_sun_scatter(const float4* tris, int32_t ntri, const SunGridParams* gp,
                                                     uint32_t* cur, uint32_t* list, uint32_t cap)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const SunGridParams g = *gp;
  if (p >= ntri || !g.valid) return;
  const CellRange c = sun_tri_cells(g, tris, p);
  for (int y = c.cy0; y <= c.cy1; ++y)
    for (int x = c.cx0; x <= c.cx1; ++x)
      if (sun_tri_overlaps_cell(g, c, x, y)) {
        const uint32_t k = atomicAdd(&cur[(int64_t)y * g.gx + x], 1u);
        if (k < cap) list[k] = (uint32_t)p;
      }
}

static constexpr int kEstBlocks = 2048;

// Candidate estimate: a shadow ray starts on a triangle and tests the
// candidates of the cell its origin falls in, one of the triangle's cells; per
// triangle the mean count of its cells, then the mean over the triangles.
__global__ __launch_bounds__(256) void k_sun_estimate(const float4* tris, int32_t ntri, const SunGridParams* gp,
                                                      const uint32_t* start, float* part)
{
  __shared__ float rs[256], rn[256];
  const SunGridParams g = *gp;
  float sum = 0.0f, num = 0.0f;
  if (g.valid) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ntri;
         p += (int64_t)gridDim.x * blockDim.x) {
      const CellRange c = sun_tri_cells(g, tris, p);
      float s = 0.0f, k = 0.0f;
      for (int y = c.cy0; y <= c.cy1; ++y)
        for (int x = c.cx0; x <= c.cx1; ++x)
          if (sun_tri_overlaps_cell(g, c, x, y)) {
            const int64_t cell = (int64_t)y * g.gx + x;
            s += (float)(start[cell + 1] - start[cell]);
            k += 1.0f;
          }
      if (k > 0.0f) {
        sum += s / k;
        num += 1.0f;
      }
    }
  }
  const int t = threadIdx.x;
  rs[t] = sum;
  rn[t] = num;
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

__global__ __launch_bounds__(256) void k_sun_decide(const float* part, SunGridParams* gp)
{
  __shared__ float rs[256], rn[256];
  const int t = threadIdx.x;
  float s = 0.0f, n = 0.0f;
  for (int b = t; b < kEstBlocks; b += 256) {
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
  if (sc.ntri > 0)
    hipLaunchKernelGGL(k_sun_count, dim3(tb), dim3(256), 0, s, sc.tris, sc.ntri, d_params, cnt);
  exclusive_scan(cnt, ncell + 1, scan_scratch, s);
  hipLaunchKernelGGL(k_sun_cursor, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, s,
                     (const uint32_t*)cnt, ncell, cap, d_params, cur);
  hipLaunchKernelGGL(k_sun_estimate, dim3(kEstBlocks), dim3(256), 0, s, sc.tris, sc.ntri,
                     (const SunGridParams*)d_params, (const uint32_t*)cnt, est);
  hipLaunchKernelGGL(k_sun_decide, dim3(1), dim3(256), 0, s, (const float*)est, d_params);
  if (sc.ntri > 0)
    hipLaunchKernelGGL(k_sun_scatter, dim3(tb), dim3(256), 0, s, sc.tris, sc.ntri,
                       (const SunGridParams*)d_params, cur, list, cap);
}

}  // namespace ane
