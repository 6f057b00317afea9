// kernels.hpp -- host-side launchers of the gfx950 kernels (internal).
#pragma once

#include <vector>

#include "common.hpp"

namespace ane {

// bvh.hip
void fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
int64_t scan_scratch_words(int64_t n);
void exclusive_scan(uint32_t* data, int64_t n, uint32_t* scratch, hipStream_t s);
int64_t radix_scratch_words(int64_t n);
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* k2, uint32_t* v2, int64_t n,
                      uint32_t* scratch, hipStream_t s);
// quantized traversal nodes [n-1][2] from the fitted nodes (layout next to NodeBuild)
void launch_quantize_nodes(const NodeBuild* nodes, int n, uint4* q, hipStream_t s);
// world-space triangles; with cbounds (a rebuild), also the centroid bounds through
// per-block partials in part (flatten_partial_words(ntri) words)
int flatten_partial_words(int ntri);
void launch_flatten(const SceneK& sc, float* world, uint32_t* tri_inst, float4* tri_rec, uint32_t* part,
                    uint32_t* cbounds, bool write_inst, hipStream_t s);
void launch_morton(const float* world, int n, const uint32_t* cbounds, uint64_t* keys,
                   uint32_t* ids, hipStream_t s);
// topo[i] = (left, right, parent, j): node i covers the leaves [min(i, j), max(i, j)]
void launch_karras(const uint64_t* keys, int n, NodeBuild* nodes, int32_t* leaf_parent, int4* topo,
                   hipStream_t s);
void launch_reorder(const float* world, const uint32_t* ids, int n, float4* tris, hipStream_t s);
// refit: leaf-order triangles from the triangle records and the instances (no world copy)
void launch_refit_tris(const SceneK& sc, const uint32_t* ids, int n, float4* tris, hipStream_t s);
// input-order world triangles [n][9] from the leaf-order ones
void launch_world_from_tris(const float4* tris, int n, float* world, hipStream_t s);
// The plan of a fit, once per topology (syncs the stream once): the treelets
// (nodes of at most 1024 leaves under a node of more, the "top" nodes) as leaf
// ranges and roots, the leaves directly under a top node, the top nodes in depth
// order (cross) with the start of each depth level (cross_begin, one past the
// last level at the end) and their depths (ndepth, by node).  scratch:
// radix_scratch_words(n) words, uscratch: 3 n + 3 words, info: 136 words; cross,
// top, units, unit_root: n - 1 entries, stray: n entries.
int launch_fit_plan(const int4* topo, const int32_t* leaf_parent, int n, uint32_t* scratch, uint32_t* uscratch,
                    uint32_t* info, uint8_t* ndepth, int32_t* cross, int32_t* top, int2* units, int32_t* unit_root,
                    int32_t* stray, std::vector<int32_t>* cross_begin, int* nunits, int* nstray, hipStream_t s);
// After a build's fit: the number of internal levels on the longest root-leaf path
// into *out (device word), from the treelet heights and the top-node depths.
int launch_tree_depth(const int4* topo, int n, const int32_t* unit_root, const uint8_t* uheight, int nunits,
                      const int32_t* top, const uint32_t* uscratch, const uint8_t* ndepth, uint32_t* out,
                      hipStream_t s);
// Bottom-up boxes: one launch for every treelet (LDS-local walks), then the top
// nodes level by level (no inter-workgroup fences).
// refit (refits): the leaf-order triangles are first made from the triangle
// records and instances of that scene view (ids: the leaf order), else tris is
// read as it is.
// uheight (nullable): the height of each treelet's root, by treelet.
// qnodes (nullable): the quantized records are written too, from the same child
// boxes (the bytes launch_quantize_nodes makes from the fitted nodes).
// full_boxes false: the treelet nodes' boxes in the 64-byte records are left
// as they were (a scene that traverses the quantized records; the top nodes'
// are written either way).
void launch_fit(float4* tris, int n, NodeBuild* nodes, float* node_box, uint4* qnodes, const int4* topo,
                const int32_t* leaf_parent, const int2* units, int nunits, const int32_t* stray, int nstray,
                const int32_t* cross, const int2* cross_kids, const std::vector<int32_t>& cross_begin, const SceneK* refit,
                const uint32_t* ids, uint8_t* uheight, hipStream_t s, bool full_boxes = true);
// cross_kids[k]: the children of crossing node cross[k], for the total crossing
// nodes of the fit plan (once per build)
void launch_cross_kids(const int32_t* cross, const int4* topo, int total, int2* kids, hipStream_t s);
// trace.hip
// Sun grid of the probe shadow rays (sungrid.hip): projected along the sun
// direction, cell (cx, cy) of a point = (sun_cell(X, x0, icw, gx), sun_cell(Y, y0, ich, gy))
struct SunGridParams {
  float x0, y0, icw, ich;  // grid origin and inverse cell sizes
  float sx, sy, margin;    // shear of the direction; widening of the triangle rectangles
  int32_t kx, ky, kz;      // axis permutation (the watertight test's)
  int32_t gx, gy;          // cells
  uint32_t valid;          // 0: the traversal takes the shadow rays
  float est;               // mean candidates of a ray from a triangle (over triangles)
  float max_est;           // above it the grid is not used (traversal is cheaper)
};
struct SunGridK {
  const SunGridParams* params;  // device
  const uint32_t* start;        // [gx * gy + 1] first candidate of each cell
  const uint32_t* list;         // candidates (sorted triangle positions); null: no grid
};
__host__ __device__ __forceinline__ int sun_cell(float v, float o, float inv, int n)
{
  float f = floorf((v - o) * inv);
  f = fminf(fmaxf(f, 0.0f), (float)(n - 1));  // also NaN -> 0
  return (int)f;
}
void sun_grid_setup(SunGridParams& g, const float* dir3, int gx, int gy);

// Any hit of the ray r (origin o, the grid's direction) among the triangles of
// its cell: the watertight test in list order, the first hit ends it.
__device__ __forceinline__ bool grid_any_hit(const SunGridK& sg, const SunGridParams& g, const float4* tris,
                                             int32_t ntri, const RayPre& r, f3 o, float tmin, float tmax,
                                             ane_hit* h)
{
  const float X = comp(o, g.kx) - g.sx * comp(o, g.kz);
  const float Y = comp(o, g.ky) - g.sy * comp(o, g.kz);
  const int64_t cell = (int64_t)sun_cell(Y, g.y0, g.ich, g.gy) * g.gx + sun_cell(X, g.x0, g.icw, g.gx);
  const uint32_t b = sg.start[cell], e = sg.start[cell + 1];
  for (uint32_t k = b; k < e; ++k) {
    const uint32_t p = sg.list[k];
    if (p >= (uint32_t)ntri) continue;
    const float4 v0 = tris[3 * (int64_t)p], v1 = tris[3 * (int64_t)p + 1], v2 = tris[3 * (int64_t)p + 2];
    float t, u, v;
    if (tri_test(r, mk(v0.x, v0.y, v0.z), mk(v1.x, v1.y, v1.z), mk(v2.x, v2.y, v2.z), tmin, tmax, t, u, v)) {
      h->tri = __float_as_int(v0.w);
      h->t = t;
      h->u = u;
      h->v = v;
      return true;
    }
  }
  return false;
}
// rays sharing the grid's direction (zero or non-finite direction: no ray); kernels
// do nothing unless the grid is valid
void launch_shadow_mask_grid(const float* depth, int W, int H, const float* ivp16, const float* light3,
                             const SunGridK& sg, const SceneK& sc, float* vis, hipStream_t s);
void launch_anyhit_grid(const float* orig, const float* dirs, int64_t n, const SunGridK& sg, const SceneK& sc,
                        float tmin, float tmax, ane_hit* hits, hipStream_t s);
// est: [sun_grid_scratch_words(ntri)] build scratch (partial sums of the candidate
// estimate, per-block cell counts)
int64_t sun_grid_scratch_words(int32_t ntri);
void launch_sun_grid_build(const SceneK& sc, const SunGridParams& g, SunGridParams* d_params, uint32_t* acc,
                           uint32_t* cnt, uint32_t* cur, uint32_t* list, uint32_t cap, uint32_t* scan_scratch,
                           float* est, hipStream_t s);
// probe traces can take a sun grid (k_probe_shadow_grid runs their shadow rays)
bool trace_grid_supported();

struct TraceArgs {
  SceneK sc;
  GridK g;
  FrameK f;
  const float* dirs;      // rotated directions [N][3]
  const float4* dirpre;   // ray setup in processing order [N + 1][3] (entry N: the sun)
  const uint32_t* order;  // processing order of ray indices [N]
  const float4* probe_org; // probe origins of the grid
  uint64_t div_magic;      // ray index -> probe: (gidx * div_magic) >> div_shift == gidx / N
  int32_t div_shift;
  ane_hit* hits;          // out: closest hit per ray
  uint8_t* shadow;        // out (probe rays): 0 lit, 1 shadowed, 255 miss
  uint32_t* work;         // dynamic ray counter (zeroed before launch)
  unsigned long long* stats; // counting kernels: nodes, triangles [, shadow nodes, shadow triangles]
  uint32_t* overflow;     // set to 1 when a ray ends without a result: a stack would exceed the
                          // spill capacity, the step limit ran out or the hierarchy is malformed
  uint32_t step_limit;    // node visits allowed per ray; 0: 2 * ntri + 8 (never reached by a valid tree)
  int32_t* spill;         // [spill_depth][total_lanes]
  int32_t spill_depth;
  int32_t total_lanes;     // lanes of the launch (set by launch_trace; a caller's value > 0 caps it)
  int32_t nrays;          // probes_in_range * N
  int32_t probe_base;     // first probe of the range
  float sun[3];           // normalize(-light_dir)
  // arbitrary-ray mode (origins/dirs given per ray)
  const float* ray_orig;
  const float* ray_dir;
  float t_min, t_max;
  const float* ray_tmax;   // caller rays: per-ray t_max (null: t_max for every ray)
  SunGridK sg;             // probe rays: candidates of the shadow rays (list null: traversal)
  const uint32_t* skip_if_set; // caller rays: the launch does nothing when *skip_if_set != 0
};

// A probe update's shading and convolution as job records, so that the request
// loop can run the probes of many updates in one launch per stage.  block_begin:
// the job's first block in the fused launch (jobs sorted by it).
struct ShadeJob {
  SceneK sc;
  GridK g;
  FrameK f;
  const float4* probe_org;
  const ane_hit* hits;
  const uint8_t* shadow;
  const float4* atlas;
  void* rad;
  float* rad_scale;
  int32_t probe_base;
  int32_t block_begin;  // one block per probe
};
struct ConvJob {
  GridK g;
  const void* rad;
  const float* rad_scale;
  const void* wq;       // weights in the job's precision
  const float* wscale;
  const float* wsum;
  float4* atlas;
  int32_t nprobes, probe_base, bake;
  int32_t block_begin;  // conv_job_blocks(precision, g.rays, nprobes) blocks
};
// max_rays: the largest ray count among the jobs (FP8 stages a probe in LDS)
void launch_shade_jobs(int precision, const ShadeJob* jobs, int njobs, int total_blocks, int max_rays,
                       hipStream_t s);
// Whether a convolution of this precision and ray count takes the streaming
// kernel (its weights fit LDS); the request loop launches such jobs apart from
// the others so that every update runs the kernel its own launch would.
bool conv_stream_fits(int precision, int rays);
int conv_job_blocks(int precision, int rays, int num_probes);
// stream: every job fits the streaming kernel; max_rays: their largest ray
// count (LDS size); rays_gcd: the greatest common divisor of their ray counts
void launch_conv_jobs(int precision, const ConvJob* jobs, int njobs, int total_blocks, bool stream, int max_rays,
                      int rays_gcd, hipStream_t s);

// With a sun grid view in a.sg (probe rays): the shadow rays of the hits of the
// preceding launch_trace (k_probe_shadow_grid; it does nothing when the grid is
// not in use on the device, in which case the trace ran them).
void launch_probe_shadow_grid(const TraceArgs& a, bool count, hipStream_t s);
int trace_lanes_per_launch();
int trace_stack_lds();
int trace_block_size();
// rays = false: probe rays (closest hit, then the shadow ray of every hit);
// rays = true: caller rays, closest or any hit
void launch_trace(const TraceArgs& a, bool any_hit, bool rays, bool count, hipStream_t s);
void launch_frame_setup(const GridK& g, const FrameK& f, const float* base_dirs,
                        const uint32_t* order, float* dirs, float4* dirpre, float4* probe_org, float* w32, float* wsum, uint16_t* wbf16,
                        uint8_t* wfp8, float* wfp8_scale, bool bake, bool f16, hipStream_t s);
struct RayDiv {
  uint64_t magic;
  int32_t shift;
};
// exact x / n for 0 <= x < nrays (nrays < 2^31, n >= 1): 2^shift > nrays * n
RayDiv ray_div_magic(int64_t nrays, int n);
void launch_shade(const SceneK& sc, const GridK& g, const FrameK& f, const float4* probe_org,
                  const ane_hit* hits,
                  const uint8_t* shadow, const float* atlas, int precision, void* rad,
                  float* rad_scale, int num_probes, int probe_base, hipStream_t s);

// conv.hip
// FP64 validation of a convolution (see conv.hip): out_max (the bits of a
// non-negative double) and out_sumsq accumulate over the probes
void launch_conv_err64(const GridK& g, const float* dirs, const double* texel_dirs, const float* rad,
                       const float* atlas_new, const float* atlas_old, double h, double floor_,
                       int num_probes, int probe_base, unsigned long long* out_max, double* out_sumsq,
                       hipStream_t s);
void launch_convolve(const GridK& g, const FrameK& f, int precision, const void* rad,
                     const float* rad_scale, const float* w32, const float* wsum,
                     const uint16_t* wbf16, const uint8_t* wfp8, const float* wfp8_scale,
                     float* atlas, int num_probes, int probe_base, bool bake, hipStream_t s);
void launch_quantize(int precision, const float* src, void* dst, float* scale, int num_probes,
                     int n, hipStream_t s);
void launch_dequantize(int precision, const void* src, const float* scale, float* dst,
                       int num_probes, int n, hipStream_t s);

// deferred GI lookup: gbuf = n x 3 float4 (ane_gbuffer_texel), out = n x float4
void launch_gi_lookup(const GridK& g, const FrameK& f, const float* atlas, const void* gbuf, int64_t n,
                      float* out, hipStream_t s);

// RT hard shadows: per-pixel rays (orig, dirs: W*H*3 floats) and visibility
void launch_shadow_rays(const float* depth, int W, int H, const float* ivp16, const float* light3,
                        float* orig, float* dirs, hipStream_t s, const uint32_t* skip_if_set = nullptr);
// point-light shadow moments: lights = n x 8 floats {x, y, z, range, r, g, b, enabled};
// 1024 rays per light; lw / lidx [kmax][64 texels] = each texel's (weight, ray) pairs
// with weight above 0.001 in the summation order, padded with weight 0
static constexpr int kLightRays = 1024;
// order[j]: the ray traced in slot j of each light (a permutation of 0..1023)
void launch_light_rays(const float* lights, int n, const float* dirs, const uint32_t* order, float* orig,
                       float* rdir, float* tmax, hipStream_t s);
void launch_light_dist(const float* lights, int n, const uint32_t* order, const ane_hit* hits, float* dist,
                       float* dist_img, hipStream_t s);
void launch_light_moments(const float* lights, int n, const float* dist, const float* lw,
                          const int32_t* lidx, int kmax, float* mom, hipStream_t s);
void launch_shadow_resolve(const float* depth, const ane_hit* hits, int64_t n, float* vis, hipStream_t s,
                           const uint32_t* skip_if_set = nullptr);
// mirror reflections: reflection rays, the shadow rays of their hits, shading
void launch_mirror_rays(const float* depth, const float* normal4, int W, int H, const float* ivp16,
                        const float* cam3, float* orig, float* dirs, hipStream_t s);
void launch_mirror_shadow_rays(const SceneK& sc, const ane_hit* hits, int64_t n, const float* sun3, float* orig,
                               float* dirs, hipStream_t s);
void launch_mirror_shade(const SceneK& sc, const GridK& g, const FrameK& f, const float* atlas,
                         const float* depth, const float* orig, const ane_hit* hits, const ane_hit* shadow_hits,
                         int64_t n, float* out, hipStream_t s);

// translate.hip
void launch_translate(const GridK& g, const float4* src, float4* dst, int dx, int dz,
                      hipStream_t s);
// atlas <-> RGBA16F (binary16 per channel, round to nearest even), texels float4 <-> uint2
void launch_atlas_to_half(const float4* src, uint2* dst, int64_t texels, hipStream_t s);
void launch_half_to_atlas(const uint2* src, float4* dst, int64_t texels, hipStream_t s);

// numeric formats (software RNE conversions shared by shade and quantize)
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float x)
{
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h)
{
  uint32_t b = (uint32_t)h << 16;
  float x;
  __builtin_memcpy(&x, &b, 4);
  return x;
}

// IEEE binary16, round to nearest even (the hardware conversion), saturating at 65504
__device__ __forceinline__ uint16_t f32_to_f16_sat(float x)
{
  x = fminf(fmaxf(x, -65504.0f), 65504.0f);
  return __builtin_bit_cast(uint16_t, (_Float16)x);
}

__device__ __forceinline__ float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// FP8 radiance block scales: one FP32 scale per probe, channel and 32 consecutive
// rays, block maximum / 448, so the largest value of a block (and every value of
// a constant block, such as rays that all see the sky) is stored exactly.
__device__ __forceinline__ float fp8_block_scale(float mx) { return mx > 0.0f ? mx / 448.0f : 1.0f; }
__device__ __forceinline__ float fp8_block_inv(float mx) { return mx > 0.0f ? 448.0f / mx : 1.0f; }

// OCP FP8 E4M3 (e4m3fn: bias 7, max 448, no infinities), round to nearest even,
// saturating at 448.
__host__ __device__ __forceinline__ uint8_t f32_to_e4m3(float x)
{
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  uint32_t s = (b >> 24) & 0x80u;
  uint32_t a = b & 0x7FFFFFFFu;
  if (a >= 0x43E00000u) return (uint8_t)(s | 0x7Eu);  // >= 448
  if (a < 0x3C800000u) {                               // < 2^-6: subnormal m * 2^-9
    float ax;
    __builtin_memcpy(&ax, &a, 4);
    float m = rintf(ax * 512.0f);
    return (uint8_t)(s | (uint32_t)m);
  }
  uint32_t r = a + 0x7FFFFu + ((a >> 20) & 1u);
  uint32_t e = (r >> 23) - 127u + 7u;
  uint32_t m = (r >> 20) & 7u;
  uint32_t code = (e << 3) | m;
  if (code > 0x7Eu) code = 0x7Eu;
  return (uint8_t)(s | code);
}

__host__ __device__ __forceinline__ float e4m3_to_f32(uint8_t q)
{
  uint32_t s = (q & 0x80u) ? 1u : 0u;
  uint32_t e = (q >> 3) & 15u, m = q & 7u;
  float v = (e == 0) ? (float)m * 0.001953125f : ldexpf(1.0f + (float)m * 0.125f, (int)e - 7);
  return s ? -v : v;
}

}  // namespace ane
