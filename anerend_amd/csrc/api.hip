// api.hip -- C-ABI of libane_ddgi (host side): scene build/refit, probe-grid
// contexts, the three probe passes, state access and diagnostics.
// The entry points and the reference calls they replace are listed in
// include/ane_ddgi.h.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <algorithm>
#include <utility>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace ane {

static thread_local std::string g_err;

int fail(int status, const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return status;
}

int fail_hip(hipError_t e, const char* expr, const char* file, int line)
{
  (void)file;
  int st = (e == hipErrorOutOfMemory) ? ANE_ERR_OUT_OF_MEMORY : ANE_ERR_DEVICE;
  return fail(st, "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), "libane_ddgi", line);
}

template <typename T>
static int dalloc(T** p, int64_t count, int64_t* bytes)
{
  *p = nullptr;
  size_t b = (size_t)(count > 0 ? count : 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? ANE_ERR_OUT_OF_MEMORY : ANE_ERR_DEVICE,
                "hipMalloc of %zu bytes failed: %s", b, hipGetErrorString(e));
  }
  if (bytes) *bytes += (int64_t)b;
  return ANE_OK;
}

template <typename T>
static void dfree(T*& p)
{
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

#define ANE_TRY(expr)          \
  do {                         \
    int _s = (expr);           \
    if (_s != ANE_OK) return _s; \
  } while (0)

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Affine inverse of a 3x4 row-major transform: cofactors in binary64, rounded once.
static void invert_3x4(const float* m, float* out)
{
  double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  double det = a * A + b * B + c * C;
  double inv[9];
  if (det == 0.0) det = 1e-30;
  inv[0] = A / det; inv[1] = -(b * i - c * h) / det; inv[2] = (b * f - c * e) / det;
  inv[3] = B / det; inv[4] = (a * i - c * g) / det;  inv[5] = -(a * f - c * d) / det;
  inv[6] = C / det; inv[7] = -(a * h - b * g) / det; inv[8] = (a * e - b * d) / det;
  double t0 = m[3], t1 = m[7], t2 = m[11];
  for (int r = 0; r < 3; ++r) {
    out[4 * r + 0] = (float)inv[3 * r + 0];
    out[4 * r + 1] = (float)inv[3 * r + 1];
    out[4 * r + 2] = (float)inv[3 * r + 2];
    out[4 * r + 3] = (float)(-(inv[3 * r + 0] * t0 + inv[3 * r + 1] * t1 + inv[3 * r + 2] * t2));
  }
}

// Fibonacci lattice, rgen:35-46.  Contract: binary64 libm evaluation, binary64
// normalization, one rounding to binary32.
static void fibonacci_dirs(int n, float* out)
{
  const double golden = 1.6180339887, pi = 3.14159265359;
  for (int i = 0; i < n; ++i) {
    double theta = 2.0 * pi * (double)i / golden;
    double phi = acos(1.0 - 2.0 * ((double)i + 0.5) / (double)n);
    double x = cos(theta) * sin(phi), y = sin(theta) * sin(phi), z = cos(phi);
    double l = sqrt(x * x + y * y + z * z);
    out[3 * i + 0] = (float)(x / l);
    out[3 * i + 1] = (float)(y / l);
    out[3 * i + 2] = (float)(z / l);
  }
}

// Weights of light_shadow_sum.comp:45-61, wt[ray][texel] for the 1024 rays of a
// light (light_shadow_helpers.glsl:1) and the 64 texels of its octahedral tile.
// The direction image is RGBA8_UNORM (LightShadowRayTracingPass.cpp:53): the
// rgen stores (dir + 1) * 0.5 (light_shadow.rgen:83) and the sum reads it back as
// c / 255 * 2 - 1.  Contract: clamp to [0, 1], times 255, round to nearest even.
// pow(x, 50) is the binary32 product chain x^2, x^4, x^8, x^16, x^32, then
// (x^32 * x^16) * x^2 (GLSL pow precision is implementation-defined).  Weights not
// above 0.001 are stored as 0 (sum.comp:59).  The rounded directions are not
// renormalized, so a weight can exceed 1, as in the shader.
static void light_weights(const float* dirs, float* wt)
{
  for (int r = 0; r < kLightRays; ++r) {
    float q[3];
    for (int k = 0; k < 3; ++k) {
      float v = (dirs[3 * r + k] + 1.0f) * 0.5f;
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      q[k] = std::rint(v * 255.0f) / 255.0f * 2.0f - 1.0f;
    }
    for (int t = 0; t < 64; ++t) {
      const f3 o = texel_dir(t);
      float c = q[0] * o.x + q[1] * o.y + q[2] * o.z;
      c = c > 0.0f ? c : 0.0f;
      const float x2 = c * c, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8, x32 = x16 * x16;
      const float w = (x32 * x16) * x2;
      wt[r * 64 + t] = w > 0.001f ? w : 0.0f;
    }
  }
}

// Processing order of the N ray indices: sorted by the 2-D Morton code of the
// octahedral encoding of their (unrotated) direction, so that the 64 lanes of a
// wave trace a compact cone of directions from one probe.  Every probe shares
// the directions and a rotation preserves their neighbourhoods.  Only the
// order in which rays are processed changes; results are per ray.
static void ray_order(int n, const float* dirs, std::vector<uint32_t>* out)
{
  std::vector<std::pair<uint32_t, uint32_t>> k((size_t)n);
  for (int i = 0; i < n; ++i) {
    float ox, oy;
    oct_encode(ld3(dirs + 3 * i), ox, oy);
    uint32_t qx = (uint32_t)fminf(fmaxf((ox + 1.0f) * 512.0f, 0.0f), 1023.0f);
    uint32_t qy = (uint32_t)fminf(fmaxf((oy + 1.0f) * 512.0f, 0.0f), 1023.0f);
    uint32_t m = 0;
    for (int b = 0; b < 10; ++b) m |= (((qx >> b) & 1u) << (2 * b + 1)) | (((qy >> b) & 1u) << (2 * b));
    k[i] = {m, (uint32_t)i};
  }
  std::sort(k.begin(), k.end());
  out->resize((size_t)n);
  for (int i = 0; i < n; ++i) (*out)[i] = k[i].second;
}

static float to_linear_host(float c)
{
  return (c < 0.04045f) ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// Ring of hipEvent pairs per stage: recording never waits on the device; a
// slot is harvested (elapsed time accumulated) before it is reused or when
// statistics are read.
struct StageTimer {
  static constexpr int kSlots = 32;
  hipEvent_t s[kSlots] = {};
  hipEvent_t e[kSlots] = {};
  bool pending[kSlots] = {};
  int next = 0, cur = -1;
  bool off = false;  // no events (while an update is captured into a graph)
  double acc_ms = 0.0;
  uint64_t count = 0;
  float last_ms = 0.0f;

  int init()
  {
    for (int i = 0; i < kSlots; ++i) {
      ANE_CHECK_HIP(hipEventCreate(&s[i]));
      ANE_CHECK_HIP(hipEventCreate(&e[i]));
    }
    return 0;
  }
  void destroy()
  {
    for (int i = 0; i < kSlots; ++i) {
      if (s[i]) (void)hipEventDestroy(s[i]);
      if (e[i]) (void)hipEventDestroy(e[i]);
      s[i] = e[i] = nullptr;
    }
  }
  int harvest(int i)
  {
    if (!pending[i]) return 0;
    ANE_CHECK_HIP(hipEventSynchronize(e[i]));
    float t = 0.0f;
    ANE_CHECK_HIP(hipEventElapsedTime(&t, s[i], e[i]));
    acc_ms += t;
    ++count;
    last_ms = t;
    pending[i] = false;
    return 0;
  }
  int harvest_all()
  {
    for (int k = 0; k < kSlots; ++k) {
      int i = (next + k) % kSlots;  // oldest first
      int st = harvest(i);
      if (st) return st;
    }
    return 0;
  }
  int begin(hipStream_t st)
  {
    if (off) return 0;
    cur = next;
    int r = harvest(cur);
    if (r) return r;
    ANE_CHECK_HIP(hipEventRecord(s[cur], st));
    return 0;
  }
  int end(hipStream_t st)
  {
    if (off) return 0;
    ANE_CHECK_HIP(hipEventRecord(e[cur], st));
    pending[cur] = true;
    next = (cur + 1) % kSlots;
    return 0;
  }
  void reset()
  {
    acc_ms = 0.0;
    count = 0;
  }
};

enum { kStSetup = 0, kStTrace, kStSunGrid, kStShade, kStConv, kStTrans, kStShadow, kStages };
static_assert(kStages == sizeof(((ane_ddgi_stats*)nullptr)->mean_ms) / sizeof(float),
              "one stats entry per stage timer");

}  // namespace ane

using namespace ane;

// ===========================================================================
// sun grids (sungrid.hip): the triangles of a scene binned along one direction,
// for any-hit rays that all share it, and the key they were built for.  A key
// is first only remembered; the second request with it builds the grid on the
// request's stream, so a direction or scene that changes with every call keeps
// the traversal and never pays a build.  One exception: when a grid of the same
// scene version (another direction) was built and found valid, and the request
// has at least ANE_SUNGRID_EAGER_RAYS rays per triangle (a build costs about as
// much as the traversal of 5 probe rays per triangle saves), the first request
// with a new direction builds: a sun that moves every frame over a static scene.
// ===========================================================================
#ifndef ANE_SUNGRID_DENSITY
#define ANE_SUNGRID_DENSITY 1.0  // cells per triangle
#endif
#ifndef ANE_SUNGRID_CAPACITY
#define ANE_SUNGRID_CAPACITY 16  // entries per triangle before the grid falls back to traversal
#endif
#ifndef ANE_SUNGRID_EAGER
#define ANE_SUNGRID_EAGER 0  // 1: build on the first request with a key
#endif
#ifndef ANE_SUNGRID_EAGER_RAYS
#define ANE_SUNGRID_EAGER_RAYS 6  // rays per triangle from which a new direction alone builds at once
#endif

struct SunGridCache {
  SunGridParams* params = nullptr;
  uint32_t* acc = nullptr;    // [6][256] per-block bounds
  uint32_t* start = nullptr;  // [cells + 1] counts, scanned into starts
  uint32_t* cur = nullptr;    // [cells] scatter cursors
  uint32_t* list = nullptr;   // [cap] candidates
  uint32_t* scan = nullptr;   // scan scratch
  float* est = nullptr;       // build scratch, sun_grid_scratch_words(ntri)
  int64_t est_words = 0;
  int g = 0;
  uint32_t cap = 0;
  int64_t bytes = 0;          // device bytes of the buffers above
  struct Key {
    const void* scene = nullptr;
    uint64_t version = 0;
    uint32_t dir[3] = {0, 0, 0};
    bool set = false;
    bool operator==(const Key& o) const
    {
      return set && o.set && scene == o.scene && version == o.version && dir[0] == o.dir[0] &&
             dir[1] == o.dir[1] && dir[2] == o.dir[2];
    }
  } built, seen;
  // the validity flag of a build, read back without blocking: copied to pinned
  // memory behind an event (not under stream capture) and taken once the event
  // has completed.  `known` is the build whose flag last_valid holds; a newer
  // build of the same scene version is predicted to be valid alike.
  uint32_t* h_valid = nullptr;
  hipEvent_t ev = nullptr;
  bool pending = false;
  Key pend, known;
  int last_valid = -1;

  void release()
  {
    dfree(params); dfree(acc); dfree(start); dfree(cur); dfree(list); dfree(scan); dfree(est);
    g = 0;
    cap = 0;
    est_words = 0;
    bytes = 0;
    built = Key{};
    pending = false;
    known = Key{};
    last_valid = -1;
  }
  // release() and the host-side read-back state (teardown)
  void destroy()
  {
    release();
    if (ev) (void)hipEventDestroy(ev);
    if (h_valid) (void)hipHostFree(h_valid);
    ev = nullptr;
    h_valid = nullptr;
  }
  void poll()
  {
    if (!pending) return;
    const hipError_t q = hipEventQuery(ev);
    if (q == hipErrorNotReady) {
      (void)hipGetLastError();  // not an error: the copy has not landed yet
      return;
    }
    pending = false;
    if (q == hipSuccess) {
      last_valid = *reinterpret_cast<volatile uint32_t*>(h_valid) ? 1 : 0;
      known = pend;
    }
  }
  // device buffers for a scene of ntri triangles: about ANE_SUNGRID_DENSITY cells
  // per triangle (at most 4096 x 4096), ANE_SUNGRID_CAPACITY entries per triangle
  int alloc(int32_t ntri)
  {
    int gg = (int)std::ceil(std::sqrt((double)std::max(ntri, 1) * ANE_SUNGRID_DENSITY));
    gg = std::min(std::max(gg, 16), 4096);
    const int64_t capl = std::min<int64_t>((int64_t)ANE_SUNGRID_CAPACITY * ntri + 1024, (int64_t)1 << 30);
    const int64_t ew = sun_grid_scratch_words(ntri);
    if (gg <= g && capl <= (int64_t)cap && ew <= est_words) return ANE_OK;
    release();
    const int64_t ncell = (int64_t)gg * gg;
    ANE_TRY(dalloc(&params, 1, &bytes));
    ANE_TRY(dalloc(&acc, 6 * 256, &bytes));
    ANE_TRY(dalloc(&start, ncell + 1, &bytes));
    ANE_TRY(dalloc(&cur, ncell, &bytes));
    ANE_TRY(dalloc(&list, capl, &bytes));
    ANE_TRY(dalloc(&scan, (ncell + 1) / 1024 + 256, &bytes));
    ANE_TRY(dalloc(&est, ew, &bytes));
    est_words = ew;
    g = gg;
    cap = (uint32_t)capl;
    return ANE_OK;
  }
  // *use: the grid for (scene, version, dir) is current after the call.  A
  // direction with a zero (or non-finite) component never gets one: its rays
  // can lie exactly in a box face, where the traversal's slab test (zero
  // components take the reciprocal of +-1e-30) culls boxes whose triangles the
  // watertight test would touch on an edge, so a grid's answer could differ
  // from the traversal's there.
  // timer (nullable) brackets a build's launches on the stream
  // rays: the rays of the request (the eager-build rule)
  int update(const SceneK& view, int32_t ntri, const void* scene, uint64_t version, const float* dir3,
             int64_t rays, hipStream_t st, bool* use, StageTimer* timer = nullptr)
  {
    *use = false;
    for (int k = 0; k < 3; ++k)
      if (!(dir3[k] != 0.0f) || !std::isfinite(dir3[k])) return ANE_OK;
    Key k;
    k.scene = scene;
    k.version = version;
    std::memcpy(k.dir, dir3, sizeof(k.dir));
    k.set = true;
    if (built == k) {
      *use = true;
      return ANE_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ANE_CHECK_HIP(hipStreamIsCapturing(st, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (!capturing) poll();  // no event queries under capture
    const bool eager = known.set && known.scene == scene && known.version == version && last_valid == 1 &&
                       rays >= (int64_t)ANE_SUNGRID_EAGER_RAYS * ntri;
    if (!ANE_SUNGRID_EAGER && !eager && !(seen == k)) {
      seen = k;
      return ANE_OK;
    }
    ANE_TRY(alloc(ntri));
    SunGridParams p;
    sun_grid_setup(p, dir3, g, g);
    if (timer) ANE_TRY(timer->begin(st));
    launch_sun_grid_build(view, p, params, acc, start, cur, list, cap, scan, est, st);
    if (timer) ANE_TRY(timer->end(st));
    ANE_CHECK_HIP(hipGetLastError());
    built = k;
    if (!capturing) {
      if (!h_valid) ANE_CHECK_HIP(hipHostMalloc((void**)&h_valid, sizeof(uint32_t), hipHostMallocDefault));
      if (!ev) ANE_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      ANE_CHECK_HIP(hipMemcpyAsync(h_valid, &params->valid, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      ANE_CHECK_HIP(hipEventRecord(ev, st));
      pend = k;
      pending = true;
    }
    *use = true;
    return ANE_OK;
  }
  SunGridK view() const { return SunGridK{params, start, list}; }
  // device address of the validity flag (the traversal launches skip when it is set)
  const uint32_t* valid_flag() const { return &params->valid; }
};

// ===========================================================================
// scene
// ===========================================================================
struct ane_scene {
  int device = 0;
  int64_t nvert = 0, nidx = 0;
  int32_t nmesh = 0, nmat = 0, ninst = 0, ntri = 0;
  std::vector<ane_instance> inst;
  std::vector<ane_mesh> meshes;
  float* d_vertices = nullptr;
  uint32_t* d_indices = nullptr;
  ane_mesh* d_meshes = nullptr;
  ane_material* d_mats = nullptr;
  ane_instance* d_inst = nullptr;
  float* d_inv = nullptr;
  uint32_t* d_tri_inst = nullptr;
  float4* d_tri_rec = nullptr;
  uint32_t* d_inst_off = nullptr;
  float* d_world = nullptr;
  uint64_t* d_keys = nullptr;
  uint64_t* d_keys2 = nullptr;
  uint32_t* d_ids = nullptr;
  uint32_t* d_ids2 = nullptr;
  NodeBuild* d_nodes = nullptr;
  int32_t* d_leaf_parent = nullptr;
  float4* d_tris = nullptr;
  float* d_node_box = nullptr;
  uint32_t* d_cbounds = nullptr;
  uint32_t* d_cpart = nullptr;  // per-block centroid-bound partials of a build
  uint4* d_qnodes = nullptr;    // quantized traversal nodes (allocated when first used)
  bool boxes_stale = false;     // the last fit left the 64-byte records' boxes out (quantized traversal)
  int32_t node_format = ANE_NODES_AUTO;
  uint32_t* d_scratch = nullptr;
  uint32_t* d_info = nullptr;   // fit plan read back at a build: level offsets, counts, tree depth
  int4* d_topo = nullptr;       // children, parent and leaf-range end of every internal node
  int32_t* d_cross = nullptr;   // top nodes (more leaves than a fit treelet holds), by depth
  int2* d_cross_kids = nullptr; // their children, in the same order (cross_kids_cap entries)
  int64_t cross_kids_cap = 0;
  int32_t* d_top = nullptr;     // top nodes, by index
  uint8_t* d_ndepth = nullptr;  // depth of the top nodes
  int2* d_units = nullptr;      // leaf ranges of the fit treelets
  int32_t* d_unit_root = nullptr;
  uint8_t* d_uheight = nullptr; // height of each treelet's root
  int32_t* d_stray = nullptr;   // leaves directly under a top node
  uint32_t* d_uscratch = nullptr;
  int32_t nunits = 0, nstray = 0;
  std::vector<int32_t> cross_begin;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int32_t depth = 0;
  bool world_stale = false;  // d_world predates the last refit (rebuilt from d_tris on inspection)
  uint64_t version = 0;  // changes with every build or refit (keys derived data such as sun grids)
  mutable float build_ms = 0.0f;
  mutable bool build_ms_pending = false;  // refit timing not yet read back
  // pinned double-buffered staging of instance records and inverse transforms:
  // an upload waits only for the copy that used its slot two uploads ago
  ane_instance* h_inst[2] = {nullptr, nullptr};
  float* h_inv[2] = {nullptr, nullptr};
  hipEvent_t ev_up[2] = {nullptr, nullptr};
  bool up_pending[2] = {false, false};
  int up_slot = 0;
  int64_t bytes = 0;
  SceneK view{};
  SunGridCache sgc;  // the shadow mask's rays along -light_dir
  // shadow-mask workspace (grows to the largest image seen)
  int64_t shadow_px = 0;
  float* d_sh_orig = nullptr;
  float* d_sh_dir = nullptr;
  ane_hit* d_sh_hits = nullptr;
  uint32_t* d_sh_work = nullptr;
  unsigned long long* d_sh_stats = nullptr;
  uint32_t* d_sh_ov = nullptr;
  int32_t* d_sh_spill = nullptr;
  int64_t sh_spill_words = 0;
  // point-light shadow workspace (grows to the most lights seen); the sum's
  // weights and the 1024 ray directions are made once per scene
  int32_t lt_cap = 0;
  float* d_lt_lights = nullptr;
  float* d_lt_orig = nullptr;
  float* d_lt_dir = nullptr;
  float* d_lt_tmax = nullptr;
  ane_hit* d_lt_hits = nullptr;
  float* d_lt_dist = nullptr;
  float* d_lt_dirs = nullptr;
  float* d_lt_wt = nullptr;       // [lt_kmax][64] weights of each texel's rays, summation order
  int32_t* d_lt_idx = nullptr;    // [lt_kmax][64] their ray indices (lt_kmax 0: not set up yet)
  uint32_t* d_lt_order = nullptr; // [1024] ray traced in each slot (direction-coherent order)
  int32_t lt_kmax = 0;
  uint32_t* d_lt_work = nullptr;
  unsigned long long* d_lt_stats = nullptr;
  uint32_t* d_lt_ov = nullptr;
  int32_t* d_lt_spill = nullptr;
  int64_t lt_spill_words = 0;
};

static void scene_free(ane_scene* s)
{
  dfree(s->d_vertices); dfree(s->d_indices); dfree(s->d_meshes); dfree(s->d_mats);
  dfree(s->d_inst); dfree(s->d_inv); dfree(s->d_tri_inst); dfree(s->d_tri_rec); dfree(s->d_inst_off);
  dfree(s->d_world); dfree(s->d_keys); dfree(s->d_keys2); dfree(s->d_ids); dfree(s->d_ids2);
  dfree(s->d_nodes); dfree(s->d_leaf_parent); dfree(s->d_tris);
  dfree(s->d_node_box); dfree(s->d_cbounds); dfree(s->d_cpart); dfree(s->d_qnodes); dfree(s->d_scratch);
  dfree(s->d_info); dfree(s->d_topo); dfree(s->d_cross); dfree(s->d_cross_kids); dfree(s->d_top); dfree(s->d_ndepth); dfree(s->d_units);
  dfree(s->d_unit_root); dfree(s->d_uheight); dfree(s->d_stray); dfree(s->d_uscratch);
  dfree(s->d_sh_orig); dfree(s->d_sh_dir); dfree(s->d_sh_hits); dfree(s->d_sh_work);
  dfree(s->d_sh_stats); dfree(s->d_sh_ov); dfree(s->d_sh_spill);
  dfree(s->d_lt_lights); dfree(s->d_lt_orig); dfree(s->d_lt_dir); dfree(s->d_lt_tmax); dfree(s->d_lt_hits);
  dfree(s->d_lt_dist); dfree(s->d_lt_dirs); dfree(s->d_lt_wt); dfree(s->d_lt_idx); dfree(s->d_lt_order); dfree(s->d_lt_work);
  dfree(s->d_lt_stats);
  dfree(s->d_lt_ov); dfree(s->d_lt_spill);
  s->sgc.destroy();
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  for (int k = 0; k < 2; ++k) {
    if (s->up_pending[k]) (void)hipEventSynchronize(s->ev_up[k]);
    if (s->ev_up[k]) (void)hipEventDestroy(s->ev_up[k]);
    if (s->h_inst[k]) (void)hipHostFree(s->h_inst[k]);
    if (s->h_inv[k]) (void)hipHostFree(s->h_inv[k]);
  }
}

static int validate_instances(const ane_scene* s, const ane_instance* inst, int32_t ninst)
{
  for (int32_t k = 0; k < ninst; ++k) {
    uint32_t m = mesh_of(inst[k].custom_index_and_mask);
    uint32_t mt = mat_of(inst[k].custom_index_and_mask);
    if ((int32_t)m >= s->nmesh)
      return fail(ANE_ERR_INVALID_ARG, "instance %d references mesh %u of %d", k, m, s->nmesh);
    if ((int32_t)mt >= s->nmat)
      return fail(ANE_ERR_INVALID_ARG, "instance %d references material %u of %d", k, mt, s->nmat);
    for (int j = 0; j < 12; ++j)
      if (!std::isfinite(inst[k].transform[j]))
        return fail(ANE_ERR_INVALID_ARG, "instance %d has a non-finite transform", k);
  }
  return ANE_OK;
}

static int upload_instances(ane_scene* s, const ane_instance* inst, hipStream_t st)
{
  s->inst.assign(inst, inst + s->ninst);
  const size_t cap = (size_t)(s->ninst > 0 ? s->ninst : 1);
  const int k = s->up_slot;
  if (!s->h_inst[k]) {
    ANE_CHECK_HIP(hipHostMalloc((void**)&s->h_inst[k], sizeof(ane_instance) * cap, hipHostMallocDefault));
    ANE_CHECK_HIP(hipHostMalloc((void**)&s->h_inv[k], sizeof(float) * 12 * cap, hipHostMallocDefault));
    ANE_CHECK_HIP(hipEventCreateWithFlags(&s->ev_up[k], hipEventDisableTiming));
  }
  if (s->up_pending[k]) {
    ANE_CHECK_HIP(hipEventSynchronize(s->ev_up[k]));
    s->up_pending[k] = false;
  }
  std::memcpy(s->h_inst[k], inst, sizeof(ane_instance) * s->ninst);
  for (int32_t i = 0; i < s->ninst; ++i) invert_3x4(inst[i].transform, s->h_inv[k] + 12 * i);
  ANE_CHECK_HIP(hipMemcpyAsync(s->d_inst, s->h_inst[k], sizeof(ane_instance) * s->ninst,
                               hipMemcpyHostToDevice, st));
  ANE_CHECK_HIP(hipMemcpyAsync(s->d_inv, s->h_inv[k], sizeof(float) * 12 * s->ninst,
                               hipMemcpyHostToDevice, st));
  ANE_CHECK_HIP(hipEventRecord(s->ev_up[k], st));
  s->up_pending[k] = true;
  s->up_slot = k ^ 1;
  return ANE_OK;
}

// The traversal layout after a build or refit: quantized 32-byte nodes when asked
// for, or by default for scenes of at least ANE_NODES_QUANTIZED_MIN_TRIS triangles,
// whose 64-byte nodes outgrow the Infinity Cache; else the 64-byte nodes.
static bool scene_wants_qnodes(const ane_scene* s)
{
  if (s->ntri < 2) return false;
  if (s->node_format == ANE_NODES_QUANTIZED) return true;
  return s->node_format == ANE_NODES_AUTO && (int64_t)s->ntri >= ANE_NODES_QUANTIZED_MIN_TRIS;
}

// The quantized records a fit should write (allocated on first use), or null
// when the scene traverses the full nodes.
static int scene_fit_qnodes(ane_scene* s, uint4** out)
{
  *out = nullptr;
  if (!scene_wants_qnodes(s)) return ANE_OK;
  if (!s->d_qnodes) ANE_TRY(dalloc(&s->d_qnodes, 2 * (int64_t)(s->ntri - 1), &s->bytes));
  *out = s->d_qnodes;
  return ANE_OK;
}

// The 64-byte records' boxes, when the last fit left them out: a fit from the
// current leaf-order triangles that writes them (the traversal of quantized
// records does not read them; a change of node format and the BVH read-back do).
static int scene_full_boxes(ane_scene* s, hipStream_t st)
{
  if (!s->boxes_stale) return ANE_OK;
  launch_fit(s->d_tris, s->ntri, s->d_nodes, s->d_node_box, nullptr, s->d_topo, s->d_leaf_parent, s->d_units,
             s->nunits, s->d_stray, s->nstray, s->d_cross, s->d_cross_kids, s->cross_begin, nullptr, nullptr, nullptr, st, true);
  ANE_CHECK_HIP(hipGetLastError());
  s->boxes_stale = false;
  return ANE_OK;
}

// fitted: the fit that just ran wrote the quantized records (scene_fit_qnodes)
static int scene_apply_node_format(ane_scene* s, hipStream_t st, bool fitted = false)
{
  if (!fitted) ANE_TRY(scene_full_boxes(s, st));
  if (!scene_wants_qnodes(s)) {
    s->view.qnodes = nullptr;
    return ANE_OK;
  }
  if (!s->d_qnodes) ANE_TRY(dalloc(&s->d_qnodes, 2 * (int64_t)(s->ntri - 1), &s->bytes));
  if (!fitted) launch_quantize_nodes(s->d_nodes, s->ntri, s->d_qnodes, st);
  s->view.qnodes = s->d_qnodes;
  return ANE_OK;
}

// full LBVH build from the uploaded instances
static int scene_build(ane_scene* s, hipStream_t st, bool write_inst)
{
  ++s->version;
  ANE_CHECK_HIP(hipEventRecord(s->ev0, st));
  const int n = s->ntri;
  launch_flatten(s->view, s->d_world, s->d_tri_inst, s->d_tri_rec, s->d_cpart, s->d_cbounds, write_inst, st);
  s->world_stale = false;
  launch_morton(s->d_world, n, s->d_cbounds, s->d_keys, s->d_ids, st);
  radix_sort_pairs(s->d_keys, s->d_ids, s->d_keys2, s->d_ids2, n, s->d_scratch, st);
  launch_karras(s->d_keys, n, s->d_nodes, s->d_leaf_parent, s->d_topo, st);
  launch_reorder(s->d_world, s->d_ids, n, s->d_tris, st);
  if (launch_fit_plan(s->d_topo, s->d_leaf_parent, n, s->d_scratch, s->d_uscratch, s->d_info, s->d_ndepth,
                      s->d_cross, s->d_top, s->d_units, s->d_unit_root, s->d_stray, &s->cross_begin, &s->nunits,
                      &s->nstray, st) != 0)
    return fail(ANE_ERR_DEVICE, "LBVH fit plan failed");
  const int64_t ncross = s->cross_begin.empty() ? 0 : s->cross_begin.back();
  if (ncross > s->cross_kids_cap) {  // the plan synchronized the stream
    dfree(s->d_cross_kids);
    s->bytes -= s->cross_kids_cap * (int64_t)sizeof(int2);
    s->cross_kids_cap = 0;
    ANE_TRY(dalloc(&s->d_cross_kids, ncross, &s->bytes));
    s->cross_kids_cap = ncross;
  }
  launch_cross_kids(s->d_cross, s->d_topo, (int)ncross, s->d_cross_kids, st);
  uint4* qn = nullptr;
  ANE_TRY(scene_fit_qnodes(s, &qn));
  launch_fit(s->d_tris, n, s->d_nodes, s->d_node_box, qn, s->d_topo, s->d_leaf_parent, s->d_units, s->nunits,
             s->d_stray, s->nstray, s->d_cross, s->d_cross_kids, s->cross_begin, nullptr, nullptr, s->d_uheight, st, qn == nullptr);
  s->boxes_stale = qn != nullptr;
  ANE_TRY(scene_apply_node_format(s, st, qn != nullptr));
  uint32_t* d_depth = s->d_info + 132;
  if (launch_tree_depth(s->d_topo, n, s->d_unit_root, s->d_uheight, s->nunits, s->d_top, s->d_uscratch,
                        s->d_ndepth, d_depth, st) != 0)
    return fail(ANE_ERR_DEVICE, "LBVH depth failed");
  ANE_CHECK_HIP(hipGetLastError());
  ANE_CHECK_HIP(hipEventRecord(s->ev1, st));
  uint32_t depth = 0;
  if (n >= 2) ANE_CHECK_HIP(hipMemcpyAsync(&depth, d_depth, 4, hipMemcpyDeviceToHost, st));
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  // the deepest internal nodes have two leaf children: leaf depth = number of internal levels
  s->depth = (int32_t)depth;
  ANE_CHECK_HIP(hipEventElapsedTime(&s->build_ms, s->ev0, s->ev1));
  s->build_ms_pending = false;
  return ANE_OK;
}

static int scene_refit_dev(ane_scene* s, hipStream_t st)
{
  ++s->version;
  ANE_CHECK_HIP(hipEventRecord(s->ev0, st));
  const int n = s->ntri;
  // a refit keeps the Morton order: the fit makes the leaf-order triangles straight
  // from the triangle records (no centroid bounds, no input-order world copy)
  const SceneK view = s->view;
  uint4* qn = nullptr;
  ANE_TRY(scene_fit_qnodes(s, &qn));
  launch_fit(s->d_tris, n, s->d_nodes, s->d_node_box, qn, s->d_topo, s->d_leaf_parent, s->d_units, s->nunits,
             s->d_stray, s->nstray, s->d_cross, s->d_cross_kids, s->cross_begin, &view, s->d_ids, nullptr, st, qn == nullptr);
  s->boxes_stale = qn != nullptr;
  s->world_stale = true;
  ANE_TRY(scene_apply_node_format(s, st, qn != nullptr));
  ANE_CHECK_HIP(hipGetLastError());
  ANE_CHECK_HIP(hipEventRecord(s->ev1, st));
  // the refit stays asynchronous; its device time is read when it is asked for
  s->build_ms_pending = true;
  return ANE_OK;
}

extern "C" int ane_scene_create(const float* vertices, int64_t num_vertices, const uint32_t* indices,
                                int64_t num_indices, const ane_mesh* meshes, int32_t num_meshes,
                                const ane_material* materials, int32_t num_materials,
                                const ane_instance* instances, int32_t num_instances, void* stream,
                                ane_scene** out)
{
  if (!out) return fail(ANE_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (num_vertices < 0 || num_indices < 0 || num_meshes < 0 || num_materials < 0 || num_instances < 0)
    return fail(ANE_ERR_INVALID_ARG, "negative count");
  if ((num_vertices && !vertices) || (num_indices && !indices) || (num_meshes && !meshes) ||
      (num_materials && !materials) || (num_instances && !instances))
    return fail(ANE_ERR_INVALID_ARG, "NULL array with a non-zero count");
  if (num_meshes > (1 << 14)) return fail(ANE_ERR_INVALID_ARG, "at most 16384 meshes (14-bit mesh index)");
  if (num_materials > (1 << 10)) return fail(ANE_ERR_INVALID_ARG, "at most 1024 materials (10-bit material index)");
  for (int32_t m = 0; m < num_meshes; ++m) {
    const ane_mesh& me = meshes[m];
    if (me.reserved != 0)
      return fail(ANE_ERR_INVALID_ARG, "mesh %d: reserved field must be 0 (a GPUMeshInfo array goes to "
                  "ane_scene_create_mesh_info)", m);
    if (me.index_count % 3 != 0)
      return fail(ANE_ERR_INVALID_ARG, "mesh %d index_count %u is not a multiple of 3", m, me.index_count);
    if ((int64_t)me.index_offset + me.index_count > num_indices)
      return fail(ANE_ERR_INVALID_ARG, "mesh %d indices exceed the index array", m);
    for (uint32_t i = 0; i < me.index_count; ++i)
      if ((int64_t)me.vertex_offset + indices[me.index_offset + i] >= num_vertices)
        return fail(ANE_ERR_INVALID_ARG, "mesh %d index %u out of the vertex array", m, i);
  }
  for (int32_t k = 0; k < num_materials; ++k)
    for (int j = 0; j < 4; ++j)
      if (materials[k].bindless_indices[j] != -1)
        return fail(ANE_ERR_UNSUPPORTED, "material %d uses a texture (bindless index != -1)", k);
  ane_scene* s = new ane_scene();
  s->nvert = num_vertices; s->nidx = num_indices; s->nmesh = num_meshes; s->nmat = num_materials;
  s->ninst = num_instances;
  s->meshes.assign(meshes, meshes + num_meshes);
  int st = validate_instances(s, instances, num_instances);
  if (st != ANE_OK) { delete s; return st; }
  std::vector<uint32_t> off((size_t)num_instances + 1);
  int64_t nt = 0;
  for (int32_t k = 0; k < num_instances; ++k) {
    off[k] = (uint32_t)nt;
    nt += meshes[mesh_of(instances[k].custom_index_and_mask)].index_count / 3;
    if (nt >= (int64_t)1 << 31) { delete s; return fail(ANE_ERR_INVALID_ARG, "more than 2^31 triangles"); }
  }
  off[num_instances] = (uint32_t)nt;
  s->ntri = (int32_t)nt;
  (void)hipGetDevice(&s->device);
  hipStream_t sm = S(stream);
  const int64_t n = s->ntri;
#define ANE_A(p, c) do { st = dalloc(&s->p, (c), &s->bytes); if (st) { scene_free(s); delete s; return st; } } while (0)
  ANE_A(d_vertices, num_vertices * 24);
  ANE_A(d_indices, num_indices);
  ANE_A(d_meshes, num_meshes);
  ANE_A(d_mats, num_materials);
  ANE_A(d_inst, num_instances);
  ANE_A(d_inv, (int64_t)num_instances * 12);
  ANE_A(d_tri_inst, n);
  ANE_A(d_tri_rec, (int64_t)n * 4);
  ANE_A(d_inst_off, (int64_t)num_instances + 1);
  ANE_A(d_world, n * 9);
  ANE_A(d_keys, n);
  ANE_A(d_keys2, n);
  ANE_A(d_ids, n);
  ANE_A(d_ids2, n);
  ANE_A(d_nodes, n > 1 ? n - 1 : 1);
  ANE_A(d_leaf_parent, n);
  ANE_A(d_tris, n * 3);
  ANE_A(d_node_box, (n > 1 ? n - 1 : 1) * 6);
  ANE_A(d_cbounds, 6);
  ANE_A(d_cpart, flatten_partial_words((int)n));
  ANE_A(d_scratch, radix_scratch_words(n > 0 ? n : 1));
  ANE_A(d_info, 136);
  ANE_A(d_topo, n > 1 ? n - 1 : 1);
  ANE_A(d_cross, n > 1 ? n - 1 : 1);
  ANE_A(d_top, n > 1 ? n - 1 : 1);
  ANE_A(d_ndepth, n > 1 ? n - 1 : 1);
  ANE_A(d_units, n > 1 ? n - 1 : 1);
  ANE_A(d_unit_root, n > 1 ? n - 1 : 1);
  ANE_A(d_uheight, n > 1 ? n - 1 : 1);
  ANE_A(d_stray, n > 0 ? n : 1);
  ANE_A(d_uscratch, 3 * n + 3);
#undef ANE_A
  auto bail = [&](int code) { scene_free(s); delete s; return code; };
  hipError_t e = hipSuccess;
  if ((e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess)
    return bail(fail_hip(e, "hipEventCreate", __FILE__, __LINE__));
  if ((e = hipMemcpyAsync(s->d_vertices, vertices, sizeof(float) * 24 * num_vertices, hipMemcpyHostToDevice, sm)) ||
      (e = hipMemcpyAsync(s->d_indices, indices, sizeof(uint32_t) * num_indices, hipMemcpyHostToDevice, sm)) ||
      (e = hipMemcpyAsync(s->d_meshes, meshes, sizeof(ane_mesh) * num_meshes, hipMemcpyHostToDevice, sm)) ||
      (e = hipMemcpyAsync(s->d_mats, materials, sizeof(ane_material) * num_materials, hipMemcpyHostToDevice, sm)) ||
      (e = hipMemcpyAsync(s->d_inst_off, off.data(), sizeof(uint32_t) * off.size(), hipMemcpyHostToDevice, sm)))
    return bail(fail_hip(e, "hipMemcpyAsync", __FILE__, __LINE__));
  // nodes start zeroed so that unused fields are deterministic
  if ((e = hipMemsetAsync(s->d_nodes, 0, sizeof(NodeBuild) * (n > 1 ? n - 1 : 1), sm)))
    return bail(fail_hip(e, "hipMemsetAsync", __FILE__, __LINE__));
  if (n == 1) {
    int32_t m1 = -1;
    if ((e = hipMemcpyAsync(s->d_leaf_parent, &m1, 4, hipMemcpyHostToDevice, sm)))
      return bail(fail_hip(e, "hipMemcpyAsync", __FILE__, __LINE__));
  }
  st = upload_instances(s, instances, sm);
  if (st) return bail(st);
  SceneK& v = s->view;
  v.vertices = s->d_vertices; v.indices = s->d_indices; v.meshes = s->d_meshes; v.mats = s->d_mats;
  v.inst = s->d_inst; v.inv_xform = s->d_inv; v.tri_inst = s->d_tri_inst; v.tri_rec = s->d_tri_rec; v.inst_tri_off = s->d_inst_off;
  v.nodes = reinterpret_cast<const float4*>(s->d_nodes); v.tris = s->d_tris;
  v.ntri = s->ntri; v.ninst = s->ninst;
  st = scene_build(s, sm, true);
  if (st) return bail(st);
  *out = s;
  return ANE_OK;
}

extern "C" int ane_scene_create_mesh_info(const float* vertices, int64_t num_vertices,
                                          const uint32_t* indices, int64_t num_indices,
                                          const ane_gpu_mesh_info* meshes,
                                          const uint32_t* mesh_index_counts, int32_t num_meshes,
                                          const ane_material* materials, int32_t num_materials,
                                          const ane_instance* instances, int32_t num_instances,
                                          void* stream, ane_scene** out)
{
  if (num_meshes < 0) return fail(ANE_ERR_INVALID_ARG, "negative count");
  if (num_meshes && (!meshes || !mesh_index_counts))
    return fail(ANE_ERR_INVALID_ARG, "NULL mesh array or index counts with a non-zero count");
  if (num_meshes > (1 << 14)) return fail(ANE_ERR_INVALID_ARG, "at most 16384 meshes (14-bit mesh index)");
  std::vector<ane_mesh> m((size_t)num_meshes);
  for (int32_t k = 0; k < num_meshes; ++k) {
    memcpy(m[k].min_pos, meshes[k].min_pos, sizeof(m[k].min_pos));
    memcpy(m[k].max_pos, meshes[k].max_pos, sizeof(m[k].max_pos));
    m[k].vertex_offset = meshes[k].vertex_offset;
    m[k].index_offset = meshes[k].index_offset;
    m[k].index_count = mesh_index_counts[k];
    m[k].reserved = 0;
  }
  return ane_scene_create(vertices, num_vertices, indices, num_indices, m.data(), num_meshes, materials,
                          num_materials, instances, num_instances, stream, out);
}

static int check_same_meshes(const ane_scene* s, const ane_instance* inst, int32_t n)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "scene is NULL");
  if (n != s->ninst) return fail(ANE_ERR_INVALID_ARG, "instance count %d != %d", n, s->ninst);
  if (n && !inst) return fail(ANE_ERR_INVALID_ARG, "instances is NULL");
  for (int32_t k = 0; k < n; ++k)
    if (mesh_of(inst[k].custom_index_and_mask) != mesh_of(s->inst[k].custom_index_and_mask))
      return fail(ANE_ERR_INVALID_ARG, "instance %d changed its mesh", k);
  return validate_instances(s, inst, n);
}

extern "C" int ane_scene_refit(ane_scene* s, const ane_instance* inst, int32_t n, void* stream)
{
  ANE_TRY(check_same_meshes(s, inst, n));
  ANE_TRY(upload_instances(s, inst, S(stream)));
  return scene_refit_dev(s, S(stream));
}

extern "C" int ane_scene_rebuild(ane_scene* s, const ane_instance* inst, int32_t n, void* stream)
{
  ANE_TRY(check_same_meshes(s, inst, n));
  ANE_TRY(upload_instances(s, inst, S(stream)));
  return scene_build(s, S(stream), false);
}

extern "C" int ane_scene_info_get(const ane_scene* s, ane_scene_info* out)
{
  if (!s || !out) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  out->num_triangles = s->ntri;
  out->num_instances = s->ninst;
  out->bvh_depth = s->depth;
  out->device_bytes = s->bytes;
  if (s->build_ms_pending) {
    ANE_CHECK_HIP(hipEventSynchronize(s->ev1));
    ANE_CHECK_HIP(hipEventElapsedTime(&s->build_ms, s->ev0, s->ev1));
    s->build_ms_pending = false;
  }
  out->build_ms = s->build_ms;
  out->node_format = s->view.qnodes ? ANE_NODES_QUANTIZED : ANE_NODES_FULL;
  return ANE_OK;
}

extern "C" int ane_scene_set_node_format(ane_scene* s, int32_t format, void* stream)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "NULL scene");
  if (format != ANE_NODES_AUTO && format != ANE_NODES_FULL && format != ANE_NODES_QUANTIZED)
    return fail(ANE_ERR_INVALID_ARG, "unknown node format");
  s->node_format = format;
  hipStream_t st = S(stream);
  // work already queued on the scene's streams must not see the nodes change under it,
  // and work queued afterwards on any stream must see them complete
  ANE_CHECK_HIP(hipDeviceSynchronize());
  ANE_TRY(scene_apply_node_format(s, st));
  ANE_CHECK_HIP(hipGetLastError());
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  return ANE_OK;
}

extern "C" int ane_scene_destroy(ane_scene* s)
{
  if (!s) return ANE_OK;
  (void)hipDeviceSynchronize();
  scene_free(s);
  delete s;
  return ANE_OK;
}

extern "C" int ane_scene_debug_bvh(const ane_scene* s, uint64_t* keys, uint32_t* ids, int32_t* left,
                                   int32_t* right, int32_t* parent, float* boxes)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "scene is NULL");
  const int64_t n = s->ntri;
  // the boxes come from the 64-byte records, refitted first if the last fit left
  // them out (derived data: the scene is unchanged otherwise)
  if (boxes) ANE_TRY(scene_full_boxes(const_cast<ane_scene*>(s), nullptr));
  ANE_CHECK_HIP(hipDeviceSynchronize());
  if (keys) ANE_CHECK_HIP(hipMemcpy(keys, s->d_keys, 8 * n, hipMemcpyDeviceToHost));
  if (ids) ANE_CHECK_HIP(hipMemcpy(ids, s->d_ids, 4 * n, hipMemcpyDeviceToHost));
  if (n < 2) return ANE_OK;
  std::vector<NodeBuild> nodes((size_t)(n - 1));
  ANE_CHECK_HIP(hipMemcpy(nodes.data(), s->d_nodes, sizeof(NodeBuild) * (n - 1), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n - 1; ++i) {
    if (left) left[i] = nodes[i].d.x;
    if (right) right[i] = nodes[i].d.y;
    if (parent) parent[i] = nodes[i].d.z;
  }
  // a node's box is the union of its two child boxes in the traversal record
  if (boxes)
    for (int64_t i = 0; i < n - 1; ++i) {
      const NodeBuild& q = nodes[i];
      float* b = boxes + 6 * i;
      b[0] = std::fmin(q.a.x, q.b.x); b[1] = std::fmin(q.a.z, q.b.z); b[2] = std::fmin(q.c.x, q.c.z);
      b[3] = std::fmax(q.a.y, q.b.y); b[4] = std::fmax(q.a.w, q.b.w); b[5] = std::fmax(q.c.y, q.c.w);
    }
  return ANE_OK;
}

extern "C" int ane_scene_debug_world_tris(const ane_scene* s, float* tris9)
{
  if (!s || !tris9) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  ANE_CHECK_HIP(hipDeviceSynchronize());
  if (s->world_stale) {
    launch_world_from_tris(s->d_tris, s->ntri, s->d_world, nullptr);
    ANE_CHECK_HIP(hipGetLastError());
    ANE_CHECK_HIP(hipDeviceSynchronize());
  }
  ANE_CHECK_HIP(hipMemcpy(tris9, s->d_world, sizeof(float) * 9 * (size_t)s->ntri, hipMemcpyDeviceToHost));
  return ANE_OK;
}

// ===========================================================================
// probe-grid context
// ===========================================================================
struct ane_ddgi_ctx {
  ane_ddgi_config cfg{};
  GridK g{};
  int device = 0;
  int64_t P = 0, N = 0, texels = 0;
  float* d_atlas[2] = {nullptr, nullptr};
  int cur = 0;
  float* d_base_dirs = nullptr;
  float* d_dirs = nullptr;
  float4* d_dirpre = nullptr;
  float4* d_probe_org = nullptr;
  float* d_w32 = nullptr;
  float* d_wsum = nullptr;
  uint16_t* d_wbf16 = nullptr;
  uint8_t* d_wfp8 = nullptr;
  float* d_wfp8_scale = nullptr;
  uint32_t* d_order = nullptr;
  ane_hit* d_hits = nullptr;
  uint8_t* d_shadow = nullptr;
  void* d_rad = nullptr;
  float* d_rad_scale = nullptr;
  float* d_stage = nullptr;  // FP32 staging for radiance get/set
  float2* d_tiles = nullptr;  // tile texture bounds: nx + 1 columns, then (ny + 1) * (nz + 1) rows
  // baked tile atlases (ane_ddgi_tiles_set): RGBA16F texels per tile, tile indices
  uint2* d_baked = nullptr;
  std::vector<int32_t> baked_xz;
  int32_t baked_size = 32;
  int64_t baked_bytes = 0;
  uint32_t* d_work = nullptr;
  unsigned long long* d_stats = nullptr;
  uint32_t* d_overflow = nullptr;  // sticky traversal fault of the probe traces (ane_ddgi_check)
  uint32_t step_limit = 0;          // ane_ddgi_set_traversal_limit
  int32_t* d_spill = nullptr;
  int64_t spill_words = 0;
  StageTimer timers[kStages];
  int32_t last_cell[2] = {0, 0};
  int32_t rad_begin = -1, rad_count = 0;  // layers held in the radiance buffer
  bool rad_bake = false;
  int32_t last_begin = 0, last_count = 0;
  int counting = 0;  // 1: counting kernels, shadow rays traverse; 2: counting kernels, sun grid
  SunGridCache sgc;  // probe and mirror shadow rays (normalize(-light_dir))
  ane_ddgi_stats stats{};
  int64_t bytes = 0;
  hipGraphExec_t gexec = nullptr;  // the last captured update (request loop)
  // mirror-reflection workspace, sized on first use: [0] reflection rays,
  // [1] their shadow rays
  float* d_mr_orig[2] = {nullptr, nullptr};
  float* d_mr_dir[2] = {nullptr, nullptr};
  ane_hit* d_mr_hits[2] = {nullptr, nullptr};
  uint32_t* d_mr_work = nullptr;             // [2] ray counters
  unsigned long long* d_mr_stats = nullptr;  // [4]
  uint32_t* d_mr_ov = nullptr;
  int32_t* d_mr_spill = nullptr;
  int64_t mr_px = 0, mr_spill_words = 0;
};

static void ctx_free(ane_ddgi_ctx* c)
{
  dfree(c->d_atlas[0]); dfree(c->d_atlas[1]); dfree(c->d_base_dirs); dfree(c->d_dirs);
  dfree(c->d_dirpre); dfree(c->d_probe_org); 
  dfree(c->d_w32); dfree(c->d_wsum); dfree(c->d_wbf16); dfree(c->d_wfp8); dfree(c->d_wfp8_scale);
  dfree(c->d_order); dfree(c->d_hits); dfree(c->d_shadow); dfree(c->d_rad); dfree(c->d_rad_scale);
  dfree(c->d_stage); dfree(c->d_work); dfree(c->d_stats); dfree(c->d_overflow); dfree(c->d_spill);
  dfree(c->d_tiles);
  dfree(c->d_baked);
  c->sgc.destroy();
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  for (int k = 0; k < 2; ++k) { dfree(c->d_mr_orig[k]); dfree(c->d_mr_dir[k]); dfree(c->d_mr_hits[k]); }
  dfree(c->d_mr_work); dfree(c->d_mr_stats); dfree(c->d_mr_ov); dfree(c->d_mr_spill);
  for (auto& t : c->timers) t.destroy();
}

static int validate_config(const ane_ddgi_config* c)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "config is NULL");
  if (c->probes_x <= 0 || c->probes_y <= 0 || c->probes_z <= 0)
    return fail(ANE_ERR_INVALID_ARG, "probe grid %d x %d x %d must be positive", c->probes_x, c->probes_y, c->probes_z);
  int64_t P = (int64_t)c->probes_x * c->probes_y * c->probes_z;
  if (P > (1 << 22)) return fail(ANE_ERR_INVALID_ARG, "too many probes (%lld)", (long long)P);
  if ((int64_t)10 * c->probes_z * c->probes_y > (1 << 24) || 10 * (int64_t)c->probes_x > (1 << 24))
    return fail(ANE_ERR_INVALID_ARG, "atlas too large");
  if (c->rays_per_probe <= 0 || c->rays_per_probe > 4096)
    return fail(ANE_ERR_INVALID_ARG, "rays_per_probe %d outside 1..4096", c->rays_per_probe);
  if (P * c->rays_per_probe > ((int64_t)1 << 31) - 1)
    return fail(ANE_ERR_INVALID_ARG, "probes x rays exceeds 2^31");
  for (int i = 0; i < 3; ++i)
    if (!(c->probe_step[i] > 0.0f) || !std::isfinite(c->probe_step[i]))
      return fail(ANE_ERR_INVALID_ARG, "probe_step[%d] must be positive", i);
  if (!(c->hysteresis >= 0.0f && c->hysteresis <= 1.0f))
    return fail(ANE_ERR_INVALID_ARG, "hysteresis must be in [0, 1]");
  if (!(c->t_min >= 0.0f) || !(c->t_max > c->t_min))
    return fail(ANE_ERR_INVALID_ARG, "need 0 <= t_min < t_max");
  if (c->precision < ANE_PRECISION_FP32 || c->precision > ANE_PRECISION_FP16)
    return fail(ANE_ERR_INVALID_ARG, "unknown precision %d", c->precision);
  if (c->precision != ANE_PRECISION_FP32 && c->rays_per_probe % 32 != 0)
    return fail(ANE_ERR_UNSUPPORTED, "BF16/FP16/FP8 convolution needs rays_per_probe %% 32 == 0");
  if (c->quirks != 0) return fail(ANE_ERR_UNSUPPORTED, "quirks must be 0");
  return ANE_OK;
}

extern "C" int ane_ddgi_create(const ane_ddgi_config* cfg, ane_ddgi_ctx** out)
{
  if (!out) return fail(ANE_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  ANE_TRY(validate_config(cfg));
  ane_ddgi_ctx* c = new ane_ddgi_ctx();
  c->cfg = *cfg;
  (void)hipGetDevice(&c->device);
  GridK& g = c->g;
  g.nx = cfg->probes_x; g.ny = cfg->probes_y; g.nz = cfg->probes_z;
  for (int i = 0; i < 3; ++i) g.step[i] = cfg->probe_step[i];
  g.rays = cfg->rays_per_probe; g.hysteresis = cfg->hysteresis;
  g.t_min = cfg->t_min; g.t_max = cfg->t_max;
  g.atlas_w = 10 * g.nx; g.atlas_h = 10 * g.nz * g.ny;
  c->P = (int64_t)g.nx * g.ny * g.nz;
  c->N = g.rays;
  c->texels = (int64_t)g.atlas_w * g.atlas_h;
  int st = ANE_OK;
  const int64_t PN = c->P * c->N;
  const int esz = cfg->precision == ANE_PRECISION_FP32 ? 4 : (cfg->precision == ANE_PRECISION_FP8 ? 1 : 2);
#define ANE_A(p, cnt) do { st = dalloc(&c->p, (cnt), &c->bytes); if (st) { ctx_free(c); delete c; return st; } } while (0)
  ANE_A(d_atlas[0], c->texels * 4);
  ANE_A(d_atlas[1], c->texels * 4);
  ANE_A(d_base_dirs, c->N * 3);
  ANE_A(d_dirs, c->N * 3);
  ANE_A(d_dirpre, (c->N + 1) * 3);
  ANE_A(d_probe_org, c->P);
  ANE_A(d_w32, 64 * c->N);
  ANE_A(d_wsum, 64);
  ANE_A(d_wbf16, 64 * c->N);
  ANE_A(d_wfp8, 64 * c->N);
  ANE_A(d_wfp8_scale, 128);  // [64] FP8 texel scales, [64] BF16 weight-sum corrections
  ANE_A(d_order, c->N);
  ANE_A(d_hits, PN);
  ANE_A(d_shadow, PN);
  {
    uint8_t* r = nullptr;
    st = dalloc(&r, PN * 3 * esz, &c->bytes);
    if (st) { ctx_free(c); delete c; return st; }
    c->d_rad = r;
  }
  // FP8 radiance scales: one float per (probe, channel, 32 rays)
  ANE_A(d_rad_scale, c->P * 3 * ((c->N + 31) / 32));
  ANE_A(d_work, 4);
  ANE_A(d_stats, 4);
  ANE_A(d_overflow, 1);
  ANE_A(d_tiles, (int64_t)(g.nx + 1) + (int64_t)(g.ny + 1) * (g.nz + 1));
#undef ANE_A
  hipError_t e;
  for (auto& t : c->timers)
    if ((st = t.init()) != ANE_OK) {
      ctx_free(c); delete c; return st;
    }
  {
    // tile texture bounds, the operations of sampleSingleProbe (pbr_light.glsl:177-183)
    // in binary32: start = (2 * s0 + 1) / (2 * extent), end for s0 + 7
    std::vector<float2> tiles((size_t)(g.nx + 1) + (size_t)(g.ny + 1) * (g.nz + 1));
    auto bounds = [](int s0, int extent) {
      const float su = (2.0f * (float)s0 + 1.0f) / (2.0f * (float)extent);
      const float eu = (2.0f * (float)(s0 + 7) + 1.0f) / (2.0f * (float)extent);
      return make_float2(su, eu - su);
    };
    for (int x = 0; x <= g.nx; ++x) tiles[(size_t)x] = bounds(10 * x + 1, g.atlas_w);
    for (int y = 0; y <= g.ny; ++y)
      for (int z = 0; z <= g.nz; ++z)
        tiles[(size_t)(g.nx + 1) + (size_t)y * (g.nz + 1) + z] = bounds(10 * g.nz * y + 10 * z + 1, g.atlas_h);
    e = hipMemcpy(c->d_tiles, tiles.data(), tiles.size() * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      ctx_free(c); delete c;
      return fail(ANE_ERR_DEVICE, "tile table upload: %s", hipGetErrorString(e));
    }
    g.tile_u = c->d_tiles;
    g.tile_v = c->d_tiles + (g.nx + 1);
  }
  // atlas cleared to (0, 0, 0, 1): IrradianceProbeConvolvePass.cpp:20-24
  std::vector<float> clear((size_t)c->texels * 4);
  for (int64_t i = 0; i < c->texels; ++i) {
    clear[4 * i + 0] = 0.0f; clear[4 * i + 1] = 0.0f; clear[4 * i + 2] = 0.0f; clear[4 * i + 3] = 1.0f;
  }
  std::vector<float> dirs((size_t)c->N * 3);
  fibonacci_dirs((int)c->N, dirs.data());
  std::vector<uint32_t> order;
  ray_order((int)c->N, dirs.data(), &order);
  if ((e = hipMemcpy(c->d_atlas[0], clear.data(), clear.size() * 4, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(c->d_atlas[1], clear.data(), clear.size() * 4, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(c->d_base_dirs, dirs.data(), dirs.size() * 4, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(c->d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice)) ||
      (e = hipMemset(c->d_overflow, 0, 4))) {
    st = fail_hip(e, "hipMemcpy", __FILE__, __LINE__);
    ctx_free(c); delete c; return st;
  }
  *out = c;
  return ANE_OK;
}

extern "C" int ane_ddgi_destroy(ane_ddgi_ctx* c)
{
  if (!c) return ANE_OK;
  (void)hipDeviceSynchronize();
  ctx_free(c);
  delete c;
  return ANE_OK;
}

static int frame_to_k(const ane_ddgi_ctx* c, const ane_ddgi_frame* f, FrameK* k)
{
  if (!f) return fail(ANE_ERR_INVALID_ARG, "frame is NULL");
  if (f->layer_count <= 0 || f->layer_begin < 0 || f->layer_begin + f->layer_count > c->g.ny)
    return fail(ANE_ERR_INVALID_ARG, "layer range [%d, %d) outside 0..%d", f->layer_begin,
                f->layer_begin + f->layer_count, c->g.ny);
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(f->camera_pos[i]) || !std::isfinite(f->light_dir[i]))
      return fail(ANE_ERR_INVALID_ARG, "non-finite camera or light");
  if ((f->flags & ANE_FLAG_BAKE) && f->bake_tile[2] <= 0)
    return fail(ANE_ERR_INVALID_ARG, "bake mode needs a positive tile size");
  for (int i = 0; i < 4; ++i) {
    k->cam[i] = f->camera_pos[i];
    k->light_dir[i] = f->light_dir[i];
    k->bake_tile[i] = f->bake_tile[i];
  }
  // rmiss:20 toLinear(skyIntensity * skyColor)
  for (int i = 0; i < 3; ++i) k->sky[i] = to_linear_host(f->sky_intensity * f->sky_color[i]);
  k->sun_intensity = f->sun_intensity;
  k->flags = f->flags;
  k->layer_begin = f->layer_begin;
  k->layer_count = f->layer_count;
  for (int i = 0; i < 16; ++i) k->rot[i] = f->rotation[i];
  const f3 L = normalize(mk(-f->light_dir[0], -f->light_dir[1], -f->light_dir[2]));
  k->sun_l[0] = L.x;
  k->sun_l[1] = L.y;
  k->sun_l[2] = L.z;
  k->sky_intensity = f->sky_intensity;
  // the camera's window of baked tiles (prefillGPUTileInfoBuffer, VulkanRenderer.cpp:3698-3738)
  const int R = ANE_TILE_RADIUS, w = 2 * ANE_TILE_RADIUS + 1;
  k->tile_size = c->baked_size;
  k->tile_count = 0;
  for (int i = 0; i < w * w; ++i) k->tile_win[i] = -1;
  const float cx = std::floor(f->camera_pos[0] / (float)c->baked_size);
  const float cz = std::floor(f->camera_pos[2] / (float)c->baked_size);
  if (!c->baked_xz.empty() && (std::fabs(cx) > 1e9f || std::fabs(cz) > 1e9f))
    return fail(ANE_ERR_INVALID_ARG, "camera too far out for the tile index");
  k->tile_cam[0] = c->baked_xz.empty() ? 0 : (int32_t)cx;
  k->tile_cam[1] = c->baked_xz.empty() ? 0 : (int32_t)cz;
  for (size_t t = 0; t < c->baked_xz.size() / 2; ++t) {
    const int64_t rx = (int64_t)c->baked_xz[2 * t] - k->tile_cam[0];
    const int64_t rz = (int64_t)c->baked_xz[2 * t + 1] - k->tile_cam[1];
    if (rx < -R || rx > R || rz < -R || rz > R) continue;
    k->tile_win[(rz + R) * w + rx + R] = (int32_t)t;
    ++k->tile_count;
  }
  return ANE_OK;
}

static int ensure_spill(ane_ddgi_ctx* c, int32_t depth)
{
  const int64_t lanes = trace_lanes_per_launch();
  const int64_t d = depth - trace_stack_lds() + 1;
  const int64_t need = d > 0 ? d * lanes : 0;
  if (need <= c->spill_words) return ANE_OK;
  dfree(c->d_spill);
  c->spill_words = 0;
  ANE_TRY(dalloc(&c->d_spill, need, nullptr));
  c->spill_words = need;
  return ANE_OK;
}

static bool sun_grid_wanted(const ane_ddgi_ctx* c) { return c->counting != 1 && trace_grid_supported(); }

// Device bytes the context holds: its fixed buffers, baked tiles, the traversal
// spill, the sun grid, and the mirror workspace if a mirror pass ran.
static int64_t ctx_device_bytes(const ane_ddgi_ctx* c)
{
  return c->bytes + c->spill_words * 4 + c->sgc.bytes + c->mr_px * 2 * (12 + 12 + (int64_t)sizeof(ane_hit)) +
         c->mr_spill_words * 4;
}

namespace ane {
// Allocates ahead of time what the first update of scene `s` would allocate on
// the context (traversal spill sized for the scene's depth, sun-grid buffers
// sized for its triangles) and returns the context's device bytes with them, so
// that an HBM budget charged at admission stays exact.
int ddgi_prepare(ane_ddgi_ctx* c, const ane_scene* s, int64_t* bytes)
{
  if (!c || !s || !bytes) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  ANE_TRY(ensure_spill(c, s->depth));
  if (trace_grid_supported()) ANE_TRY(c->sgc.alloc(s->ntri));
  *bytes = ctx_device_bytes(c);
  return ANE_OK;
}
int64_t ddgi_device_bytes(const ane_ddgi_ctx* c) { return ctx_device_bytes(c); }
}  // namespace ane

static TraceArgs trace_args(ane_ddgi_ctx* c, const ane_scene* s, const FrameK& fk)
{
  TraceArgs a{};
  a.sc = s->view;
  a.g = c->g;
  a.f = fk;
  a.dirs = c->d_dirs;
  a.dirpre = c->d_dirpre;
  a.probe_org = c->d_probe_org;
  const RayDiv rd = ray_div_magic(c->P * c->N, (int)c->N);
  a.div_magic = rd.magic;
  a.div_shift = rd.shift;
  a.order = c->d_order;
  a.hits = c->d_hits;
  a.shadow = c->d_shadow;
  a.stats = c->d_stats;
  a.overflow = c->d_overflow;
  a.step_limit = c->step_limit;
  a.spill = c->d_spill;
  const int64_t lanes = trace_lanes_per_launch();
  a.spill_depth = (int32_t)(lanes ? c->spill_words / lanes : 0);
  a.nrays = (int32_t)((int64_t)fk.layer_count * c->g.nx * c->g.nz * c->N);
  a.probe_base = fk.layer_begin * c->g.nx * c->g.nz;
  f3 sun = normalize(mk(-fk.light_dir[0], -fk.light_dir[1], -fk.light_dir[2]));
  a.sun[0] = sun.x; a.sun[1] = sun.y; a.sun[2] = sun.z;
  return a;
}

static int do_frame_setup(ane_ddgi_ctx* c, const FrameK& fk, bool bake, hipStream_t st)
{
  launch_frame_setup(c->g, fk, c->d_base_dirs, c->d_order, c->d_dirs, c->d_dirpre, c->d_probe_org, c->d_w32,
                     c->d_wsum, c->d_wbf16, c->d_wfp8, c->d_wfp8_scale, bake,
                     c->cfg.precision == ANE_PRECISION_FP16, st);
  ANE_CHECK_HIP(hipGetLastError());
  return ANE_OK;
}

extern "C" int ane_ddgi_trace(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, void* stream)
{
  if (!c || !s) return fail(ANE_ERR_INVALID_ARG, "NULL context or scene");
  FrameK fk;
  ANE_TRY(frame_to_k(c, f, &fk));
  ANE_TRY(ensure_spill(c, s->depth));
  hipStream_t st = S(stream);
  const bool bake = (fk.flags & ANE_FLAG_BAKE) != 0;
  const int nprobes = fk.layer_count * c->g.nx * c->g.nz;
  ANE_TRY(c->timers[kStSetup].begin(st));
  ANE_TRY(do_frame_setup(c, fk, bake, st));
  fill_u32(c->d_work, 0u, 4, st);
  if (c->counting) fill_u32(reinterpret_cast<uint32_t*>(c->d_stats), 0u, 8, st);
  ANE_TRY(c->timers[kStSetup].end(st));
  TraceArgs a = trace_args(c, s, fk);
  a.work = c->d_work;
  if (sun_grid_wanted(c)) {
    bool use = false;
    ANE_TRY(c->sgc.update(s->view, s->ntri, s, s->version, a.sun, a.nrays, st, &use, &c->timers[kStSunGrid]));
    if (use) a.sg = c->sgc.view();
  }
  // primary rays in one persistent launch (with the shadow rays of their hits
  // unless a sun grid is in use), then the sun-grid shadow rays
  ANE_TRY(c->timers[kStTrace].begin(st));
  launch_trace(a, false, false, c->counting, st);
  ANE_TRY(c->timers[kStTrace].end(st));
  if (a.sg.list) {
    ANE_TRY(c->timers[kStShadow].begin(st));
    launch_probe_shadow_grid(a, c->counting, st);
    ANE_TRY(c->timers[kStShadow].end(st));
  }
  ANE_TRY(c->timers[kStShade].begin(st));
  launch_shade(s->view, c->g, fk, c->d_probe_org, c->d_hits, c->d_shadow,
               c->d_atlas[c->cur], c->cfg.precision, c->d_rad, c->d_rad_scale, nprobes,
               a.probe_base, st);
  ANE_CHECK_HIP(hipGetLastError());
  ANE_TRY(c->timers[kStShade].end(st));
  c->rad_begin = fk.layer_begin;
  c->rad_count = fk.layer_count;
  c->rad_bake = bake;
  c->last_begin = fk.layer_begin;
  c->last_count = fk.layer_count;
  c->stats.rays = (uint64_t)nprobes * c->N;
  return ANE_OK;
}

extern "C" int ane_ddgi_convolve(ane_ddgi_ctx* c, const ane_ddgi_frame* f, void* stream)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  FrameK fk;
  ANE_TRY(frame_to_k(c, f, &fk));
  if (c->rad_begin != fk.layer_begin || c->rad_count != fk.layer_count)
    return fail(ANE_ERR_NOT_READY, "no radiance for layers [%d, %d): trace or set them first",
                fk.layer_begin, fk.layer_begin + fk.layer_count);
  const bool bake = (fk.flags & ANE_FLAG_BAKE) != 0;
  if (bake != c->rad_bake)
    return fail(ANE_ERR_INVALID_ARG, "bake flag differs from the one the radiance was produced with");
  hipStream_t st = S(stream);
  const int nprobes = fk.layer_count * c->g.nx * c->g.nz;
  const int probe_base = fk.layer_begin * c->g.nx * c->g.nz;
  ANE_TRY(c->timers[kStConv].begin(st));
  launch_convolve(c->g, fk, c->cfg.precision, c->d_rad, c->d_rad_scale, c->d_w32, c->d_wsum,
                  c->d_wbf16, c->d_wfp8, c->d_wfp8_scale, c->d_atlas[c->cur], nprobes, probe_base,
                  bake, st);
  ANE_CHECK_HIP(hipGetLastError());
  ANE_TRY(c->timers[kStConv].end(st));
  return ANE_OK;
}

extern "C" int ane_ddgi_update(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, void* stream)
{
  ANE_TRY(ane_ddgi_trace(c, s, f, stream));
  return ane_ddgi_convolve(c, f, stream);
}

namespace ane {
// The update of ane_ddgi_update captured from `st` into a graph and launched on
// it (used by the request loop).  The context keeps the executable graph and
// updates it in place from each new capture (kernel arguments change with the
// frame; the node structure changes only with the layer range or the bake flag,
// which re-instantiates it).  Validation and allocation happen before the
// capture; stage timers record nothing meanwhile, so the graph holds kernel
// nodes only (counters are cleared by a fill kernel).  Counting contexts update
// directly.
// The front of an update for the request loop's fused stages: frame setup, sun
// grid and trace of (c, s, f) on `st`, as ane_ddgi_trace does them; the
// shading and convolution that ane_ddgi_update would launch next are returned
// as job records (run later in one launch per stage with other updates' jobs,
// on the same stream).  *precision: the jobs' precision.
int ddgi_update_front(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, hipStream_t st,
                      ShadeJob* sj, ConvJob* cj, int* precision)
{
  if (!c || !s || !sj || !cj || !precision) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  FrameK fk;
  ANE_TRY(frame_to_k(c, f, &fk));
  ANE_TRY(ensure_spill(c, s->depth));
  const bool bake = (fk.flags & ANE_FLAG_BAKE) != 0;
  const int nprobes = fk.layer_count * c->g.nx * c->g.nz;
  ANE_TRY(do_frame_setup(c, fk, bake, st));
  fill_u32(c->d_work, 0u, 4, st);
  if (c->counting) fill_u32(reinterpret_cast<uint32_t*>(c->d_stats), 0u, 8, st);
  TraceArgs a = trace_args(c, s, fk);
  a.work = c->d_work;
  if (sun_grid_wanted(c)) {
    bool use = false;
    ANE_TRY(c->sgc.update(s->view, s->ntri, s, s->version, a.sun, a.nrays, st, &use));
    if (use) a.sg = c->sgc.view();
  }
  launch_trace(a, false, false, c->counting, st);
  launch_probe_shadow_grid(a, c->counting, st);
  ANE_CHECK_HIP(hipGetLastError());
  *sj = ShadeJob{};
  sj->sc = s->view;
  sj->g = c->g;
  sj->f = fk;
  sj->probe_org = c->d_probe_org;
  sj->hits = c->d_hits;
  sj->shadow = c->d_shadow;
  sj->atlas = reinterpret_cast<const float4*>(c->d_atlas[c->cur]);
  sj->rad = c->d_rad;
  sj->rad_scale = c->d_rad_scale;
  sj->probe_base = a.probe_base;
  *cj = ConvJob{};
  cj->g = c->g;
  cj->rad = c->d_rad;
  cj->rad_scale = c->d_rad_scale;
  const int prec = c->cfg.precision;
  cj->wq = prec == ANE_PRECISION_FP32 ? (const void*)c->d_w32
           : prec == ANE_PRECISION_FP8 ? (const void*)c->d_wfp8 : (const void*)c->d_wbf16;
  cj->wscale = c->d_wfp8_scale;
  cj->wsum = c->d_wsum;
  cj->atlas = reinterpret_cast<float4*>(c->d_atlas[c->cur]);
  cj->nprobes = nprobes;
  cj->probe_base = a.probe_base;
  cj->bake = bake ? 1 : 0;
  *precision = prec;
  c->rad_begin = fk.layer_begin;
  c->rad_count = fk.layer_count;
  c->rad_bake = bake;
  c->last_begin = fk.layer_begin;
  c->last_count = fk.layer_count;
  c->stats.rays = (uint64_t)nprobes * c->N;
  return ANE_OK;
}

// whether the context's convolution can run as a fused job (MFMA kernels need
// a ray count that is a multiple of 32)
bool ddgi_fusable(const ane_ddgi_ctx* c) { return c && c->N % 32 == 0; }

int ddgi_update_graph(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, hipStream_t st)
{
  if (!c || !s) return fail(ANE_ERR_INVALID_ARG, "NULL context or scene");
  FrameK fk;
  ANE_TRY(frame_to_k(c, f, &fk));
  ANE_TRY(ensure_spill(c, s->depth));
  if (sun_grid_wanted(c)) ANE_TRY(c->sgc.alloc(s->ntri));  // allocation outside the capture
  if (c->counting) return ane_ddgi_update(c, s, f, st);
  for (auto& t : c->timers) t.off = true;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    for (auto& t : c->timers) t.off = false;
    return fail(ANE_ERR_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  }
  const int r = ane_ddgi_update(c, s, f, st);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(st, &g);
  for (auto& t : c->timers) t.off = false;
  if (r != ANE_OK || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return r != ANE_OK ? r : fail(ANE_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  if (c->gexec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult how = hipGraphExecUpdateError;
    if (hipGraphExecUpdate(c->gexec, g, &bad, &how) != hipSuccess || how != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();  // a structural change: instantiate anew
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
  }
  if (!c->gexec) e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipGraphLaunch(c->gexec, st);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(ANE_ERR_DEVICE, "graph launch: %s", hipGetErrorString(e));
  return ANE_OK;
}
}  // namespace ane

extern "C" int ane_ddgi_translate(ane_ddgi_ctx* c, const float* cam, int32_t* shifted, void* stream)
{
  if (!c || !cam) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (shifted) *shifted = 0;
  if (!std::isfinite(cam[0]) || !std::isfinite(cam[2])) return fail(ANE_ERR_INVALID_ARG, "non-finite camera");
  // IrradianceProbeTranslationPass.cpp:60-76 (probe cell of the camera, delta to the last one)
  const int ix = (int)(cam[0] / c->g.step[0]);
  const int iz = (int)(cam[2] / c->g.step[2]);
  const int dx = ix - c->last_cell[0], dz = iz - c->last_cell[1];
  if (dx == 0 && dz == 0) return ANE_OK;
  c->last_cell[0] = ix;
  c->last_cell[1] = iz;
  hipStream_t st = S(stream);
  ANE_TRY(c->timers[kStTrans].begin(st));
  launch_translate(c->g, reinterpret_cast<const float4*>(c->d_atlas[c->cur]),
                   reinterpret_cast<float4*>(c->d_atlas[1 - c->cur]), dx, dz, st);
  ANE_CHECK_HIP(hipGetLastError());
  ANE_TRY(c->timers[kStTrans].end(st));
  c->cur = 1 - c->cur;
  if (shifted) *shifted = 1;
  return ANE_OK;
}

// The spare atlas buffer (the translation target) stages the binary16 texels.
extern "C" int ane_ddgi_atlas_export_rgba16f(ane_ddgi_ctx* c, uint16_t* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  hipStream_t st = S(stream);
  uint2* stage = reinterpret_cast<uint2*>(c->d_atlas[c->cur ^ 1]);
  launch_atlas_to_half(reinterpret_cast<const float4*>(c->d_atlas[c->cur]), stage, c->texels, st);
  ANE_CHECK_HIP(hipGetLastError());
  ANE_CHECK_HIP(hipMemcpyAsync(host, stage, sizeof(uint2) * c->texels, hipMemcpyDeviceToHost, st));
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  return ANE_OK;
}

extern "C" int ane_ddgi_atlas_import_rgba16f(ane_ddgi_ctx* c, const uint16_t* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  hipStream_t st = S(stream);
  uint2* stage = reinterpret_cast<uint2*>(c->d_atlas[c->cur ^ 1]);
  ANE_CHECK_HIP(hipMemcpyAsync(stage, host, sizeof(uint2) * c->texels, hipMemcpyHostToDevice, st));
  launch_half_to_atlas(stage, reinterpret_cast<float4*>(c->d_atlas[c->cur]), c->texels, st);
  ANE_CHECK_HIP(hipGetLastError());
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  return ANE_OK;
}

// Directions of the 8 x 8 tile texel centres in float64, t = ty * 8 + tx
// (probe_conv.comp:37-46 with octDecode, octahedron_helpers.glsl:25-34: y and z swapped).
static void texel_dirs_f64(double* out)
{
  for (int ty = 0; ty < 8; ++ty)
    for (int tx = 0; tx < 8; ++tx) {
      double ox = (2.0 * tx + 1.0) / 16.0 * 2.0 - 1.0, oy = (2.0 * ty + 1.0) / 16.0 * 2.0 - 1.0;
      double z = 1.0 - std::fabs(ox) - std::fabs(oy);
      if (z < 0.0) {
        const double nx = (1.0 - std::fabs(oy)) * (ox >= 0.0 ? 1.0 : -1.0);
        const double ny = (1.0 - std::fabs(ox)) * (oy >= 0.0 ? 1.0 : -1.0);
        ox = nx;
        oy = ny;
      }
      const double l = std::sqrt(ox * ox + z * z + oy * oy);
      double* o = out + 3 * (ty * 8 + tx);
      o[0] = ox / l;
      o[1] = z / l;
      o[2] = oy / l;
    }
}

extern "C" int ane_ddgi_conv_error_fp64(ane_ddgi_ctx* conv, const ane_ddgi_ctx* ref, double rel_floor,
                                        double* max_rel, double* rms_rel, void* stream)
{
  if (!conv || !ref || !max_rel || !rms_rel) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (ref->cfg.precision != ANE_PRECISION_FP32)
    return fail(ANE_ERR_INVALID_ARG, "the reference context must hold FP32 radiance");
  if (ref->g.nx != conv->g.nx || ref->g.ny != conv->g.ny || ref->g.nz != conv->g.nz || ref->N != conv->N)
    return fail(ANE_ERR_INVALID_ARG, "contexts differ in grid or ray count");
  if (ref->rad_count <= 0 || ref->rad_begin != conv->rad_begin || ref->rad_count != conv->rad_count)
    return fail(ANE_ERR_NOT_READY, "the contexts hold radiance of different layers");
  if (ref->rad_bake || conv->rad_bake) return fail(ANE_ERR_UNSUPPORTED, "bake-mode blends are not checked");
  if (!(rel_floor > 0.0)) return fail(ANE_ERR_INVALID_ARG, "rel_floor must be positive");
  hipStream_t st = S(stream);
  double td[64 * 3];
  texel_dirs_f64(td);
  double* d_td = nullptr;
  double* d_out = nullptr;  // [0] max (bits), [1] sum of squares
  ANE_TRY(dalloc(&d_td, 64 * 3, nullptr));
  int rc = dalloc(&d_out, 2, nullptr);
  if (rc) { dfree(d_td); return rc; }
  const double zero[2] = {0.0, 0.0};
  hipError_t e;
  const int nprobes = ref->rad_count * ref->g.nx * ref->g.nz;
  const double h = conv->cfg.hysteresis;
  if (!(e = hipMemcpyAsync(d_td, td, sizeof(td), hipMemcpyHostToDevice, st)) &&
      !(e = hipMemcpyAsync(d_out, zero, sizeof(zero), hipMemcpyHostToDevice, st))) {
    launch_conv_err64(ref->g, ref->d_dirs, d_td, static_cast<const float*>(ref->d_rad), conv->d_atlas[conv->cur],
                      h != 0.0 ? ref->d_atlas[ref->cur] : nullptr, h, rel_floor, nprobes,
                      ref->rad_begin * ref->g.nx * ref->g.nz,
                      reinterpret_cast<unsigned long long*>(d_out), d_out + 1, st);
    double out[2] = {0.0, 0.0};
    if (!(e = hipGetLastError()) && !(e = hipMemcpyAsync(out, d_out, sizeof(out), hipMemcpyDeviceToHost, st)) &&
        !(e = hipStreamSynchronize(st))) {
      *max_rel = out[0];  // the bits of a non-negative double
      *rms_rel = std::sqrt(out[1] / (3.0 * 64.0 * nprobes));
    }
  }
  dfree(d_td);
  dfree(d_out);
  if (e) return fail_hip(e, "conv_error_fp64", __FILE__, __LINE__);
  return ANE_OK;
}

extern "C" int ane_ddgi_tiles_set(ane_ddgi_ctx* c, const int32_t* tile_xz, const uint16_t* rgba16f,
                                  int32_t num_tiles, int32_t tile_size, void* stream)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  if (num_tiles < 0 || num_tiles > 4096) return fail(ANE_ERR_INVALID_ARG, "num_tiles must be in [0, 4096]");
  if (tile_size <= 0) return fail(ANE_ERR_INVALID_ARG, "tile_size must be positive");
  if (num_tiles && (!tile_xz || !rgba16f)) return fail(ANE_ERR_INVALID_ARG, "NULL tiles with a non-zero count");
  for (int32_t a = 0; a < num_tiles; ++a)
    for (int32_t b = 0; b < a; ++b)
      if (tile_xz[2 * a] == tile_xz[2 * b] && tile_xz[2 * a + 1] == tile_xz[2 * b + 1])
        return fail(ANE_ERR_INVALID_ARG, "tile (%d, %d) given twice", tile_xz[2 * a], tile_xz[2 * a + 1]);
  hipStream_t st = S(stream);
  // the previous set may still be read by queued work on this stream
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  dfree(c->d_baked);
  c->bytes -= c->baked_bytes;
  c->baked_bytes = 0;
  c->baked_xz.clear();
  c->g.tile_atlas = nullptr;
  c->baked_size = tile_size;
  if (num_tiles == 0) return ANE_OK;
  ANE_TRY(dalloc(&c->d_baked, c->texels * num_tiles, &c->baked_bytes));
  c->bytes += c->baked_bytes;
  ANE_CHECK_HIP(hipMemcpyAsync(c->d_baked, rgba16f, sizeof(uint2) * c->texels * num_tiles,
                               hipMemcpyHostToDevice, st));
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  c->baked_xz.assign(tile_xz, tile_xz + 2 * num_tiles);
  c->g.tile_atlas = c->d_baked;
  return ANE_OK;
}

extern "C" int ane_ddgi_gi_lookup(ane_ddgi_ctx* c, const ane_ddgi_frame* f, const ane_gbuffer_texel* gbuf,
                                  int64_t n, float* out_rgba, void* stream)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  if (n < 0) return fail(ANE_ERR_INVALID_ARG, "negative texel count");
  if (n > 0 && (!gbuf || !out_rgba)) return fail(ANE_ERR_INVALID_ARG, "NULL G-buffer or output");
  if (!f) return fail(ANE_ERR_INVALID_ARG, "frame is NULL");
  FrameK fk;
  ane_ddgi_frame full = *f;
  full.layer_begin = 0;  // the lookup reads the whole atlas; the layer range is not used
  full.layer_count = c->g.ny;
  ANE_TRY(frame_to_k(c, &full, &fk));
  launch_gi_lookup(c->g, fk, c->d_atlas[c->cur], gbuf, n, out_rgba, S(stream));
  ANE_CHECK_HIP(hipGetLastError());
  return ANE_OK;
}

extern "C" int ane_ddgi_mirror_reflections(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f,
                                           const float* depth, const float* normal4, int32_t width,
                                           int32_t height, const float* ivp16, float* out_rgba,
                                           uint64_t* counters4, void* stream)
{
  if (!c || !s || !f || !ivp16) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (width < 0 || height < 0) return fail(ANE_ERR_INVALID_ARG, "negative image size");
  const int64_t n = (int64_t)width * height;
  if (n == 0) return ANE_OK;
  if (!depth || !normal4 || !out_rgba) return fail(ANE_ERR_INVALID_ARG, "NULL depth, normal or output");
  if (n >= ((int64_t)1 << 31)) return fail(ANE_ERR_INVALID_ARG, "image too large");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(ivp16[k])) return fail(ANE_ERR_INVALID_ARG, "non-finite inverse view-projection");
  if (c->device != s->device) return fail(ANE_ERR_INVALID_ARG, "context and scene on different devices");
  FrameK fk;
  ane_ddgi_frame full = *f;
  full.layer_begin = 0;  // the hit shading reads the whole atlas; the layer range is not used
  full.layer_count = c->g.ny;
  ANE_TRY(frame_to_k(c, &full, &fk));
  hipStream_t st = S(stream);
  if (n > c->mr_px) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 2; ++k) { dfree(c->d_mr_orig[k]); dfree(c->d_mr_dir[k]); dfree(c->d_mr_hits[k]); }
    c->mr_px = 0;
    for (int k = 0; k < 2; ++k) {
      ANE_TRY(dalloc(&c->d_mr_orig[k], 3 * n, nullptr));
      ANE_TRY(dalloc(&c->d_mr_dir[k], 3 * n, nullptr));
      ANE_TRY(dalloc(&c->d_mr_hits[k], n, nullptr));
    }
    c->mr_px = n;
  }
  if (!c->d_mr_work || !c->d_mr_stats || !c->d_mr_ov) {  // all three, or redo them
    dfree(c->d_mr_work); dfree(c->d_mr_stats); dfree(c->d_mr_ov);
    ANE_TRY(dalloc(&c->d_mr_work, 2, nullptr));
    ANE_TRY(dalloc(&c->d_mr_stats, 4, nullptr));
    ANE_TRY(dalloc(&c->d_mr_ov, 1, nullptr));
  }
  const int64_t lanes = trace_lanes_per_launch();
  const int64_t sd = s->depth - trace_stack_lds() + 1 > 0 ? s->depth - trace_stack_lds() + 1 : 0;
  if (sd * lanes > c->mr_spill_words) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    dfree(c->d_mr_spill);
    c->mr_spill_words = 0;
    ANE_TRY(dalloc(&c->d_mr_spill, sd * lanes, nullptr));
    c->mr_spill_words = sd * lanes;
  }
  fill_u32(c->d_mr_work, 0u, 2, st);
  fill_u32(reinterpret_cast<uint32_t*>(c->d_mr_stats), 0u, 8, st);
  fill_u32(c->d_mr_ov, 0u, 1, st);
  // mirror_reflections.rgen:57-79: reflection rays over [0.1, 50], closest hit
  launch_mirror_rays(depth, normal4, width, height, ivp16, f->camera_pos, c->d_mr_orig[0], c->d_mr_dir[0], st);
  TraceArgs a{};
  a.sc = s->view;
  a.g.rays = 1;
  a.overflow = c->d_mr_ov;
  a.spill = c->d_mr_spill;
  a.spill_depth = (int32_t)sd;
  a.nrays = (int32_t)n;
  a.t_min = 0.1f;
  a.t_max = 50.0f;
  a.hits = c->d_mr_hits[0];
  a.work = c->d_mr_work;
  a.stats = c->d_mr_stats;
  a.ray_orig = c->d_mr_orig[0];
  a.ray_dir = c->d_mr_dir[0];
  launch_trace(a, false, true, counters4 != nullptr, st);
  // mirror_reflections.rchit:44-66: shadow rays from the hits, any hit over [0.1, 50]
  const f3 sun = normalize(mk(-fk.light_dir[0], -fk.light_dir[1], -fk.light_dir[2]));
  const float sun3[3] = {sun.x, sun.y, sun.z};
  launch_mirror_shadow_rays(s->view, c->d_mr_hits[0], n, sun3, c->d_mr_orig[1], c->d_mr_dir[1], st);
  const uint32_t* skip = nullptr;
  if (!counters4 && sun_grid_wanted(c)) {
    bool use = false;
    ANE_TRY(c->sgc.update(s->view, s->ntri, s, s->version, sun3, n, st, &use));
    if (use) {
      launch_anyhit_grid(c->d_mr_orig[1], c->d_mr_dir[1], n, c->sgc.view(), s->view, 0.1f, 50.0f, c->d_mr_hits[1],
                         st);
      skip = c->sgc.valid_flag();
    }
  }
  a.skip_if_set = skip;
  a.hits = c->d_mr_hits[1];
  a.work = c->d_mr_work + 1;
  a.stats = c->d_mr_stats + 2;
  a.ray_orig = c->d_mr_orig[1];
  a.ray_dir = c->d_mr_dir[1];
  launch_trace(a, true, true, counters4 != nullptr, st);
  launch_mirror_shade(s->view, c->g, fk, c->d_atlas[c->cur], depth, c->d_mr_orig[0], c->d_mr_hits[0],
                      c->d_mr_hits[1], n, out_rgba, st);
  ANE_CHECK_HIP(hipGetLastError());
  if (counters4) {
    unsigned long long c4[4] = {0, 0, 0, 0};
    uint32_t ov = 0;
    ANE_CHECK_HIP(hipMemcpyAsync(c4, c->d_mr_stats, 32, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipMemcpyAsync(&ov, c->d_mr_ov, 4, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 4; ++k) counters4[k] = c4[k];
    if (ov) return fail(ANE_ERR_DEVICE, "traversal stack overflow");
  }
  return ANE_OK;
}

extern "C" int ane_ddgi_translation_state(ane_ddgi_ctx* c, int32_t* cell, int32_t set)
{
  if (!c || !cell) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (set) { c->last_cell[0] = cell[0]; c->last_cell[1] = cell[1]; }
  else { cell[0] = c->last_cell[0]; cell[1] = c->last_cell[1]; }
  return ANE_OK;
}

extern "C" int ane_ddgi_atlas_get(ane_ddgi_ctx* c, float* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  ANE_CHECK_HIP(hipMemcpyAsync(host, c->d_atlas[c->cur], sizeof(float) * 4 * c->texels, hipMemcpyDeviceToHost, S(stream)));
  ANE_CHECK_HIP(hipStreamSynchronize(S(stream)));
  return ANE_OK;
}

extern "C" int ane_ddgi_atlas_set(ane_ddgi_ctx* c, const float* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  ANE_CHECK_HIP(hipMemcpyAsync(c->d_atlas[c->cur], host, sizeof(float) * 4 * c->texels, hipMemcpyHostToDevice, S(stream)));
  ANE_CHECK_HIP(hipStreamSynchronize(S(stream)));
  return ANE_OK;
}

extern "C" int ane_ddgi_atlas_device(ane_ddgi_ctx* c, float** dev, int64_t* w, int64_t* h)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  if (dev) *dev = c->d_atlas[c->cur];
  if (w) *w = c->g.atlas_w;
  if (h) *h = c->g.atlas_h;
  return ANE_OK;
}

static int ensure_stage(ane_ddgi_ctx* c)
{
  if (c->d_stage) return ANE_OK;
  return dalloc(&c->d_stage, c->P * 3 * c->N, &c->bytes);
}

extern "C" int ane_ddgi_radiance_set(ane_ddgi_ctx* c, const ane_ddgi_frame* f, const float* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  FrameK fk;
  ANE_TRY(frame_to_k(c, f, &fk));
  hipStream_t st = S(stream);
  const int nprobes = fk.layer_count * c->g.nx * c->g.nz;
  const int64_t cnt = (int64_t)nprobes * 3 * c->N;
  const bool bake = (fk.flags & ANE_FLAG_BAKE) != 0;
  ANE_TRY(do_frame_setup(c, fk, bake, st));
  if (c->cfg.precision == ANE_PRECISION_FP32) {
    ANE_CHECK_HIP(hipMemcpyAsync(c->d_rad, host, sizeof(float) * cnt, hipMemcpyHostToDevice, st));
  } else {
    ANE_TRY(ensure_stage(c));
    ANE_CHECK_HIP(hipMemcpyAsync(c->d_stage, host, sizeof(float) * cnt, hipMemcpyHostToDevice, st));
    launch_quantize(c->cfg.precision, c->d_stage, c->d_rad, c->d_rad_scale, nprobes, (int)c->N, st);
    ANE_CHECK_HIP(hipGetLastError());
  }
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  c->rad_begin = fk.layer_begin;
  c->rad_count = fk.layer_count;
  c->rad_bake = bake;
  return ANE_OK;
}

extern "C" int ane_ddgi_radiance_get(ane_ddgi_ctx* c, float* host, void* stream)
{
  if (!c || !host) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (c->rad_begin < 0) return fail(ANE_ERR_NOT_READY, "no radiance yet");
  hipStream_t st = S(stream);
  const int nprobes = c->rad_count * c->g.nx * c->g.nz;
  const int64_t cnt = (int64_t)nprobes * 3 * c->N;
  if (c->cfg.precision == ANE_PRECISION_FP32) {
    ANE_CHECK_HIP(hipMemcpyAsync(host, c->d_rad, sizeof(float) * cnt, hipMemcpyDeviceToHost, st));
  } else {
    ANE_TRY(ensure_stage(c));
    launch_dequantize(c->cfg.precision, c->d_rad, c->d_rad_scale, c->d_stage, nprobes, (int)c->N, st);
    ANE_CHECK_HIP(hipGetLastError());
    ANE_CHECK_HIP(hipMemcpyAsync(host, c->d_stage, sizeof(float) * cnt, hipMemcpyDeviceToHost, st));
  }
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  return ANE_OK;
}

extern "C" int ane_ddgi_debug_rays(ane_ddgi_ctx* c, ane_hit* hits, uint8_t* shadowed, void* stream)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  const int64_t cnt = (int64_t)c->last_count * c->g.nx * c->g.nz * c->N;
  hipStream_t st = S(stream);
  if (hits) ANE_CHECK_HIP(hipMemcpyAsync(hits, c->d_hits, sizeof(ane_hit) * cnt, hipMemcpyDeviceToHost, st));
  if (shadowed) ANE_CHECK_HIP(hipMemcpyAsync(shadowed, c->d_shadow, cnt, hipMemcpyDeviceToHost, st));
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  return ANE_OK;
}

extern "C" int ane_ddgi_set_counting(ane_ddgi_ctx* c, int32_t enable)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  if (enable < 0 || enable > 2) return fail(ANE_ERR_INVALID_ARG, "counting mode %d outside 0..2", enable);
  c->counting = enable;
  return ANE_OK;
}

extern "C" int ane_ddgi_stats_get(const ane_ddgi_ctx* cc, ane_ddgi_stats* out)
{
  if (!cc || !out) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  ane_ddgi_ctx* c = const_cast<ane_ddgi_ctx*>(cc);
  ane_ddgi_stats s = c->stats;
  s.device_bytes = ctx_device_bytes(c);
  for (int k = 0; k < kStages; ++k) {
    StageTimer& t = c->timers[k];
    ANE_TRY(t.harvest_all());
    s.last_ms[k] = t.last_ms;
    s.mean_ms[k] = t.count ? (float)(t.acc_ms / (double)t.count) : 0.0f;
    s.calls[k] = t.count;
  }
  (void)hipGetLastError();
  if (c->counting) {
    unsigned long long v[4] = {0, 0, 0, 0};
    ANE_CHECK_HIP(hipMemcpy(v, c->d_stats, sizeof(v), hipMemcpyDeviceToHost));
    s.primary_nodes = v[0]; s.primary_tris = v[1]; s.shadow_nodes = v[2]; s.shadow_tris = v[3];
  }
  s.sun_grid_state = 0;
  s.sun_grid_candidates = 0.0f;
  if (c->sgc.built.set) {
    SunGridParams p;
    ANE_CHECK_HIP(hipMemcpy(&p, c->sgc.params, sizeof(p), hipMemcpyDeviceToHost));
    s.sun_grid_state = p.valid ? 1 : 2;
    s.sun_grid_candidates = p.est;
  }
  uint32_t ov = 0;
  ANE_CHECK_HIP(hipMemcpy(&ov, c->d_overflow, 4, hipMemcpyDeviceToHost));
  s.traversal_fault = ov ? 1u : 0u;
  *out = s;
  if (ov) return fail(ANE_ERR_DEVICE, "traversal fault: some probe rays ended without a result "
                      "(stack over its spill capacity, step limit or malformed hierarchy)");
  return ANE_OK;
}

// reads and clears the context's traversal fault flags (probe traces and mirror rays)
static int take_faults(ane_ddgi_ctx* c, hipStream_t st, uint32_t* probe, uint32_t* mirror)
{
  uint32_t v[2] = {0, 0};
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  ANE_CHECK_HIP(hipMemcpy(&v[0], c->d_overflow, 4, hipMemcpyDeviceToHost));
  if (c->d_mr_ov) ANE_CHECK_HIP(hipMemcpy(&v[1], c->d_mr_ov, 4, hipMemcpyDeviceToHost));
  ANE_CHECK_HIP(hipMemset(c->d_overflow, 0, 4));
  if (c->d_mr_ov) ANE_CHECK_HIP(hipMemset(c->d_mr_ov, 0, 4));
  *probe = v[0];
  *mirror = v[1];
  return ANE_OK;
}

extern "C" int ane_ddgi_stats_reset(ane_ddgi_ctx* c)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  for (auto& t : c->timers) {
    ANE_TRY(t.harvest_all());
    t.reset();
  }
  uint32_t p = 0, m = 0;
  return take_faults(c, nullptr, &p, &m);
}

extern "C" int ane_ddgi_check(ane_ddgi_ctx* c, void* stream)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  uint32_t p = 0, m = 0;
  ANE_TRY(take_faults(c, S(stream), &p, &m));
  if (p || m)
    return fail(ANE_ERR_DEVICE, "traversal fault in the %s: some rays ended without a result (stack over "
                "its spill capacity, step limit or malformed hierarchy)",
                p && m ? "probe traces and mirror rays" : (p ? "probe traces" : "mirror rays"));
  return ANE_OK;
}

extern "C" int ane_ddgi_set_traversal_limit(ane_ddgi_ctx* c, uint32_t max_node_visits)
{
  if (!c) return fail(ANE_ERR_INVALID_ARG, "NULL context");
  c->step_limit = max_node_visits;
  return ANE_OK;
}

extern "C" int ane_scene_shadow_mask(ane_scene* s, const float* depth, int32_t width, int32_t height,
                                     const float* ivp16, const float* light3, float* out_vis,
                                     uint64_t* counters2, void* stream)
{
  if (!s || !ivp16 || !light3) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (width < 0 || height < 0) return fail(ANE_ERR_INVALID_ARG, "negative image size");
  const int64_t n = (int64_t)width * height;
  if (n == 0) return ANE_OK;
  if (!depth || !out_vis) return fail(ANE_ERR_INVALID_ARG, "NULL depth or output");
  if (n >= ((int64_t)1 << 31)) return fail(ANE_ERR_INVALID_ARG, "image too large");
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(ivp16[k])) return fail(ANE_ERR_INVALID_ARG, "non-finite inverse view-projection");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(light3[k])) return fail(ANE_ERR_INVALID_ARG, "non-finite light direction");
  hipStream_t st = S(stream);
  if (n > s->shadow_px) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    dfree(s->d_sh_orig); dfree(s->d_sh_dir); dfree(s->d_sh_hits);
    s->shadow_px = 0;
    ANE_TRY(dalloc(&s->d_sh_orig, 3 * n, nullptr));
    ANE_TRY(dalloc(&s->d_sh_dir, 3 * n, nullptr));
    ANE_TRY(dalloc(&s->d_sh_hits, n, nullptr));
    s->shadow_px = n;
  }
  if (!s->d_sh_work || !s->d_sh_stats || !s->d_sh_ov) {  // all three, or redo them
    dfree(s->d_sh_work); dfree(s->d_sh_stats); dfree(s->d_sh_ov);
    ANE_TRY(dalloc(&s->d_sh_work, 1, nullptr));
    ANE_TRY(dalloc(&s->d_sh_stats, 2, nullptr));
    ANE_TRY(dalloc(&s->d_sh_ov, 1, nullptr));
  }
  const int64_t lanes = trace_lanes_per_launch();
  const int64_t sd = s->depth - trace_stack_lds() + 1 > 0 ? s->depth - trace_stack_lds() + 1 : 0;
  if (sd * lanes > s->sh_spill_words) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    dfree(s->d_sh_spill);
    s->sh_spill_words = 0;
    ANE_TRY(dalloc(&s->d_sh_spill, sd * lanes, nullptr));
    s->sh_spill_words = sd * lanes;
  }
  // without counters the rays go through the scene's grid along -light_dir once it
  // is current; the traversal launches below then do nothing (device flag)
  const uint32_t* skip = nullptr;
  if (!counters2) {
    const float dir[3] = {-light3[0], -light3[1], -light3[2]};
    bool use = false;
    ANE_TRY(s->sgc.update(s->view, s->ntri, s, s->version, dir, (int64_t)width * height, st, &use));
    if (use) {
      launch_shadow_mask_grid(depth, width, height, ivp16, light3, s->sgc.view(), s->view, out_vis, st);
      skip = s->sgc.valid_flag();
    }
  }
  launch_shadow_rays(depth, width, height, ivp16, light3, s->d_sh_orig, s->d_sh_dir, st, skip);
  fill_u32(s->d_sh_work, 0u, 1, st);
  fill_u32(reinterpret_cast<uint32_t*>(s->d_sh_stats), 0u, 4, st);
  fill_u32(s->d_sh_ov, 0u, 1, st);
  TraceArgs a{};
  a.sc = s->view;
  a.g.rays = 1;
  a.hits = s->d_sh_hits;
  a.work = s->d_sh_work;
  a.stats = s->d_sh_stats;
  a.overflow = s->d_sh_ov;
  a.spill = s->d_sh_spill;
  a.spill_depth = (int32_t)sd;
  a.nrays = (int32_t)n;
  a.ray_orig = s->d_sh_orig;
  a.ray_dir = s->d_sh_dir;
  a.t_min = 0.1f;  // shadow.rgen:66-67
  a.t_max = 50.0f;
  a.skip_if_set = skip;
  launch_trace(a, true, true, counters2 != nullptr, st);
  launch_shadow_resolve(depth, s->d_sh_hits, n, out_vis, st, skip);
  ANE_CHECK_HIP(hipGetLastError());
  if (counters2) {
    unsigned long long c2[2] = {0, 0};
    uint32_t ov = 0;
    ANE_CHECK_HIP(hipMemcpyAsync(c2, s->d_sh_stats, 16, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipMemcpyAsync(&ov, s->d_sh_ov, 4, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    counters2[0] = c2[0];
    counters2[1] = c2[1];
    if (ov) return fail(ANE_ERR_DEVICE, "traversal stack overflow");
  }
  return ANE_OK;
}

extern "C" int ane_scene_check(ane_scene* s, void* stream)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "NULL scene");
  uint32_t v[2] = {0, 0};
  ANE_CHECK_HIP(hipStreamSynchronize(S(stream)));
  if (s->d_sh_ov) ANE_CHECK_HIP(hipMemcpy(&v[0], s->d_sh_ov, 4, hipMemcpyDeviceToHost));
  if (s->d_lt_ov) ANE_CHECK_HIP(hipMemcpy(&v[1], s->d_lt_ov, 4, hipMemcpyDeviceToHost));
  if (s->d_sh_ov) ANE_CHECK_HIP(hipMemset(s->d_sh_ov, 0, 4));
  if (s->d_lt_ov) ANE_CHECK_HIP(hipMemset(s->d_lt_ov, 0, 4));
  if (v[0] || v[1])
    return fail(ANE_ERR_DEVICE, "traversal fault in the last %s: some rays ended without a result",
                v[0] && v[1] ? "shadow mask and light shadows" : (v[0] ? "shadow mask" : "light shadows"));
  return ANE_OK;
}

extern "C" int ane_scene_light_shadows(ane_scene* s, const ane_light* lights, int32_t num_lights,
                                       float* dist_img, float* moments_img, uint64_t* counters2,
                                       void* stream)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "NULL scene");
  if (num_lights < 0 || num_lights > ANE_MAX_LIGHTS)
    return fail(ANE_ERR_INVALID_ARG, "num_lights must be in [0, 1024]");
  if (num_lights == 0) return ANE_OK;
  if (!lights || !moments_img) return fail(ANE_ERR_INVALID_ARG, "NULL lights or moments image");
  for (int32_t l = 0; l < num_lights; ++l) {
    if (lights[l].color[3] < 0.5f) continue;  // disabled; the shaders' test, so NaN counts as enabled
    for (int k = 0; k < 4; ++k)
      if (!std::isfinite(lights[l].world_pos[k]))
        return fail(ANE_ERR_INVALID_ARG, "non-finite position or range of an enabled light");
  }
  hipStream_t st = S(stream);
  const int64_t m = (int64_t)num_lights * kLightRays;
  if (s->lt_kmax == 0) {  // set last: a setup that failed part-way is redone from scratch
    dfree(s->d_lt_order); dfree(s->d_lt_dirs); dfree(s->d_lt_wt); dfree(s->d_lt_idx);
    dfree(s->d_lt_work); dfree(s->d_lt_stats); dfree(s->d_lt_ov);
    std::vector<float> dirs((size_t)kLightRays * 3), wt((size_t)kLightRays * 64);
    fibonacci_dirs(kLightRays, dirs.data());  // worldDirFromFibonacciIndex, light_shadow.rgen:27-38
    light_weights(dirs.data(), wt.data());
    // each texel's rays with a weight above 0.001, in the shader's order (x outer,
    // y inner over the 32 x 32 rays, light_shadow_sum.comp:50-52)
    std::vector<std::vector<std::pair<float, int32_t>>> list(64);
    for (int x = 0; x < 32; ++x)
      for (int y = 0; y < 32; ++y) {
        const int idx = y * 32 + x;
        for (int t = 0; t < 64; ++t)
          if (wt[(size_t)idx * 64 + t] > 0.001f) list[t].push_back({wt[(size_t)idx * 64 + t], idx});
      }
    int kmax = 1;
    for (const auto& e : list) kmax = std::max(kmax, (int)e.size());
    std::vector<float> lw((size_t)kmax * 64, 0.0f);
    std::vector<int32_t> li((size_t)kmax * 64, 0);
    for (int t = 0; t < 64; ++t)
      for (size_t k = 0; k < list[t].size(); ++k) {
        lw[k * 64 + t] = list[t][k].first;
        li[k * 64 + t] = list[t][k].second;
      }
    std::vector<uint32_t> order;
    ray_order(kLightRays, dirs.data(), &order);
    ANE_TRY(dalloc(&s->d_lt_order, kLightRays, nullptr));
    ANE_CHECK_HIP(hipMemcpy(s->d_lt_order, order.data(), sizeof(uint32_t) * kLightRays, hipMemcpyHostToDevice));
    ANE_TRY(dalloc(&s->d_lt_dirs, 3 * kLightRays, nullptr));
    ANE_TRY(dalloc(&s->d_lt_wt, 64 * (int64_t)kmax, nullptr));
    ANE_TRY(dalloc(&s->d_lt_idx, 64 * (int64_t)kmax, nullptr));
    ANE_CHECK_HIP(hipMemcpy(s->d_lt_dirs, dirs.data(), sizeof(float) * dirs.size(), hipMemcpyHostToDevice));
    ANE_CHECK_HIP(hipMemcpy(s->d_lt_wt, lw.data(), sizeof(float) * lw.size(), hipMemcpyHostToDevice));
    ANE_CHECK_HIP(hipMemcpy(s->d_lt_idx, li.data(), sizeof(int32_t) * li.size(), hipMemcpyHostToDevice));
    ANE_TRY(dalloc(&s->d_lt_work, 1, nullptr));
    ANE_TRY(dalloc(&s->d_lt_stats, 2, nullptr));
    ANE_TRY(dalloc(&s->d_lt_ov, 1, nullptr));
    s->lt_kmax = kmax;
  }
  if (num_lights > s->lt_cap) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    dfree(s->d_lt_lights); dfree(s->d_lt_orig); dfree(s->d_lt_dir); dfree(s->d_lt_tmax);
    dfree(s->d_lt_hits); dfree(s->d_lt_dist);
    s->lt_cap = 0;
    ANE_TRY(dalloc(&s->d_lt_lights, 8 * (int64_t)num_lights, nullptr));
    ANE_TRY(dalloc(&s->d_lt_orig, 3 * m, nullptr));
    ANE_TRY(dalloc(&s->d_lt_dir, 3 * m, nullptr));
    ANE_TRY(dalloc(&s->d_lt_tmax, m, nullptr));
    ANE_TRY(dalloc(&s->d_lt_hits, m, nullptr));
    ANE_TRY(dalloc(&s->d_lt_dist, m, nullptr));
    s->lt_cap = num_lights;
  }
  const int64_t lanes = trace_lanes_per_launch();
  const int64_t sd = s->depth - trace_stack_lds() + 1 > 0 ? s->depth - trace_stack_lds() + 1 : 0;
  if (sd * lanes > s->lt_spill_words) {
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    dfree(s->d_lt_spill);
    s->lt_spill_words = 0;
    ANE_TRY(dalloc(&s->d_lt_spill, sd * lanes, nullptr));
    s->lt_spill_words = sd * lanes;
  }
  // the light records are read on the device (LightBuffer, bindless.glsl:129-131)
  ANE_CHECK_HIP(hipMemcpyAsync(s->d_lt_lights, lights, sizeof(ane_light) * (size_t)num_lights,
                               hipMemcpyHostToDevice, st));
  launch_light_rays(s->d_lt_lights, num_lights, s->d_lt_dirs, s->d_lt_order, s->d_lt_orig, s->d_lt_dir,
                    s->d_lt_tmax, st);
  fill_u32(s->d_lt_work, 0u, 1, st);
  fill_u32(reinterpret_cast<uint32_t*>(s->d_lt_stats), 0u, 4, st);
  fill_u32(s->d_lt_ov, 0u, 1, st);
  TraceArgs a{};
  a.sc = s->view;
  a.g.rays = 1;
  a.hits = s->d_lt_hits;
  a.work = s->d_lt_work;
  a.stats = s->d_lt_stats;
  a.overflow = s->d_lt_ov;
  a.spill = s->d_lt_spill;
  a.spill_depth = (int32_t)sd;
  a.nrays = (int32_t)m;
  a.ray_orig = s->d_lt_orig;
  a.ray_dir = s->d_lt_dir;
  a.t_min = 0.1f;          // light_shadow.rgen:65
  a.t_max = 0.0f;          // per ray: the light's range (rgen:66)
  a.ray_tmax = s->d_lt_tmax;
  launch_trace(a, false, true, counters2 != nullptr, st);  // closest hit: rchit reports gl_HitTEXT
  launch_light_dist(s->d_lt_lights, num_lights, s->d_lt_order, s->d_lt_hits, s->d_lt_dist, dist_img, st);
  launch_light_moments(s->d_lt_lights, num_lights, s->d_lt_dist, s->d_lt_wt, s->d_lt_idx, s->lt_kmax,
                       moments_img, st);
  ANE_CHECK_HIP(hipGetLastError());
  if (counters2) {
    unsigned long long c2[2] = {0, 0};
    uint32_t ov = 0;
    ANE_CHECK_HIP(hipMemcpyAsync(c2, s->d_lt_stats, 16, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipMemcpyAsync(&ov, s->d_lt_ov, 4, hipMemcpyDeviceToHost, st));
    ANE_CHECK_HIP(hipStreamSynchronize(st));
    counters2[0] = c2[0];
    counters2[1] = c2[1];
    if (ov) return fail(ANE_ERR_DEVICE, "traversal stack overflow");
  }
  return ANE_OK;
}

extern "C" int ane_scene_trace_rays(const ane_scene* s, const float* origins, const float* dirs,
                                    int64_t num_rays, float t_min, float t_max, int32_t any_hit,
                                    ane_hit* hits, uint64_t* counters2, void* stream)
{
  if (!s || !origins || !dirs || !hits) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (num_rays <= 0) return ANE_OK;
  if (num_rays >= ((int64_t)1 << 31)) return fail(ANE_ERR_INVALID_ARG, "too many rays");
  if (!(t_min >= 0.0f) || !(t_max > t_min)) return fail(ANE_ERR_INVALID_ARG, "need 0 <= t_min < t_max");
  hipStream_t st = S(stream);
  float *d_o = nullptr, *d_d = nullptr;
  ane_hit* d_h = nullptr;
  uint32_t* d_work = nullptr;
  unsigned long long* d_stats = nullptr;
  uint32_t* d_ov = nullptr;
  int32_t* d_spill = nullptr;
  const int64_t lanes = trace_lanes_per_launch();
  const int64_t sd = s->depth - trace_stack_lds() + 1 > 0 ? s->depth - trace_stack_lds() + 1 : 0;
  int st_ = ANE_OK;
  auto cleanup = [&]() { dfree(d_o); dfree(d_d); dfree(d_h); dfree(d_work); dfree(d_stats); dfree(d_ov); dfree(d_spill); };
  if ((st_ = dalloc(&d_o, num_rays * 3, nullptr)) || (st_ = dalloc(&d_d, num_rays * 3, nullptr)) ||
      (st_ = dalloc(&d_h, num_rays, nullptr)) || (st_ = dalloc(&d_work, 1, nullptr)) ||
      (st_ = dalloc(&d_stats, 2, nullptr)) || (st_ = dalloc(&d_ov, 1, nullptr)) ||
      (st_ = dalloc(&d_spill, sd * lanes, nullptr))) {
    cleanup();
    return st_;
  }
  hipError_t e;
  if ((e = hipMemcpyAsync(d_o, origins, 12 * num_rays, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(d_d, dirs, 12 * num_rays, hipMemcpyHostToDevice, st)) ||
      (e = hipMemsetAsync(d_stats, 0, 16, st)) || (e = hipMemsetAsync(d_ov, 0, 4, st))) {
    cleanup();
    return fail_hip(e, "upload", __FILE__, __LINE__);
  }
  fill_u32(d_work, 0u, 1, st);
  TraceArgs a{};
  a.sc = s->view;
  a.g.t_min = t_min;
  a.g.t_max = t_max;
  a.g.rays = 1;
  a.hits = d_h;
  a.work = d_work;
  a.stats = d_stats;
  a.overflow = d_ov;
  a.spill = d_spill;
  a.spill_depth = (int32_t)sd;
  a.nrays = (int32_t)num_rays;
  a.ray_orig = d_o;
  a.ray_dir = d_d;
  a.t_min = t_min;
  a.t_max = t_max;
  launch_trace(a, any_hit != 0, true, counters2 != nullptr, st);
  unsigned long long c2[2] = {0, 0};
  uint32_t ov = 0;
  if ((e = hipGetLastError()) ||
      (e = hipMemcpyAsync(hits, d_h, sizeof(ane_hit) * num_rays, hipMemcpyDeviceToHost, st)) ||
      (e = hipMemcpyAsync(c2, d_stats, 16, hipMemcpyDeviceToHost, st)) ||
      (e = hipMemcpyAsync(&ov, d_ov, 4, hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st))) {
    cleanup();
    return fail_hip(e, "trace", __FILE__, __LINE__);
  }
  cleanup();
  if (ov) return fail(ANE_ERR_DEVICE, "traversal stack exceeded its spill capacity");
  if (counters2) { counters2[0] = c2[0]; counters2[1] = c2[1]; }
  return ANE_OK;
}

// ===========================================================================
// library
// ===========================================================================
// ===========================================================================
// Allocation-free ray query (ane_scene_trace_rays_noalloc).  The traversal's
// scratch (ray counter, work counters, fault flag, stack spill) lives in one
// file-scope device array per GPU, so the call allocates nothing.  Host code
// never uses the array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device with that device current.
// ===========================================================================
static constexpr int64_t kScratchBytes = (int64_t)16 << 20;  // per GPU, held until the process ends
static constexpr int64_t kScratchSpillWord = 64;              // spill area starts at byte 256
static __device__ __align__(256) uint32_t scratch[kScratchBytes / 4];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kMaxDevices] = {};

static int scratch_for_current_device(uint32_t** out)
{
  int dev = 0;
  ANE_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ANE_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ANE_ERR_DEVICE, "hipGetSymbolAddress of the traversal scratch failed: %s",
                  hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ANE_OK;
}

extern "C" int ane_scene_trace_rays_noalloc(const ane_scene* s, const float* origins, const float* dirs,
                                            int64_t num_rays, float t_min, float t_max, int32_t any_hit,
                                            ane_hit* hits, uint64_t* counters2, void* stream)
{
  if (!s || !origins || !dirs || !hits) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  if (num_rays <= 0) return ANE_OK;
  if (num_rays >= ((int64_t)1 << 31)) return fail(ANE_ERR_INVALID_ARG, "too many rays");
  if (!(t_min >= 0.0f) || !(t_max > t_min)) return fail(ANE_ERR_INVALID_ARG, "need 0 <= t_min < t_max");
  uint32_t* sp = nullptr;
  ANE_TRY(scratch_for_current_device(&sp));
  // lanes of the persistent launch: all resident blocks, fewer if the hierarchy's
  // depth needs more spilled stack than the scratch holds, at least one block per CU
  const int64_t tb = trace_block_size();
  const int64_t full = trace_lanes_per_launch();
  const int64_t depth = s->depth - trace_stack_lds() + 1 > 0 ? s->depth - trace_stack_lds() + 1 : 0;
  const int64_t cap_words = kScratchBytes / 4 - kScratchSpillWord;
  int64_t lanes = full;
  if (depth > 0 && depth * lanes > cap_words) lanes = (cap_words / depth) / tb * tb;
  int cus = 0, dev = 0;
  ANE_CHECK_HIP(hipGetDevice(&dev));
  ANE_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (lanes < (int64_t)cus * tb)
    return fail(ANE_ERR_UNSUPPORTED, "hierarchy depth %d needs %lld bytes of traversal stack per %lld lanes, "
                "more than the %lld-byte scratch", s->depth, (long long)(depth * 4), (long long)((int64_t)cus * tb),
                (long long)kScratchBytes);
  hipStream_t st = S(stream);
  // scratch words 0 ray counter, 2-5 the two 64-bit work counters, 6 the fault flag:
  // cleared by a kernel (the array keeps the previous call's contents)
  fill_u32(sp, 0u, 8, st);
  TraceArgs a{};
  a.sc = s->view;
  a.g.t_min = t_min;
  a.g.t_max = t_max;
  a.g.rays = 1;
  a.hits = hits;
  a.work = sp;
  a.stats = reinterpret_cast<unsigned long long*>(sp + 2);
  a.overflow = sp + 6;
  a.spill = reinterpret_cast<int32_t*>(sp + kScratchSpillWord);
  a.spill_depth = (int32_t)depth;
  a.total_lanes = (int32_t)lanes;
  a.nrays = (int32_t)num_rays;
  a.ray_orig = origins;
  a.ray_dir = dirs;
  a.t_min = t_min;
  a.t_max = t_max;
  launch_trace(a, any_hit != 0, true, counters2 != nullptr, st);
  ANE_CHECK_HIP(hipGetLastError());
  uint32_t ov = 0;
  unsigned long long c2[2] = {0, 0};
  ANE_CHECK_HIP(hipMemcpyAsync(&ov, sp + 6, 4, hipMemcpyDeviceToHost, st));
  if (counters2) ANE_CHECK_HIP(hipMemcpyAsync(c2, sp + 2, 16, hipMemcpyDeviceToHost, st));
  ANE_CHECK_HIP(hipStreamSynchronize(st));
  if (ov) return fail(ANE_ERR_DEVICE, "traversal stack exceeded its spill capacity");
  if (counters2) { counters2[0] = c2[0]; counters2[1] = c2[1]; }
  return ANE_OK;
}

extern "C" int ane_abi_version(void) { return ANE_DDGI_ABI_VERSION; }
extern "C" const char* ane_last_error(void) { return g_err.c_str(); }

extern "C" const char* ane_status_string(int s)
{
  switch (s) {
    case ANE_OK: return "ok";
    case ANE_ERR_INVALID_ARG: return "invalid argument";
    case ANE_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ANE_ERR_ADMISSION_REFUSED: return "admission refused";
    case ANE_ERR_DEVICE: return "device error";
    case ANE_ERR_NOT_READY: return "not ready";
    case ANE_ERR_UNSUPPORTED: return "unsupported";
    case ANE_ERR_CANCELLED: return "cancelled";
  }
  return "unknown status";
}

extern "C" int ane_device(int32_t* device)
{
  if (!device) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  int d = 0;
  ANE_CHECK_HIP(hipGetDevice(&d));
  *device = d;
  return ANE_OK;
}

extern "C" int ane_stream_synchronize(void* stream)
{
  ANE_CHECK_HIP(hipStreamSynchronize(S(stream)));
  return ANE_OK;
}

extern "C" int ane_ddgi_ray_directions(int32_t n, const float* rot16, float* out)
{
  if (n <= 0 || !rot16 || !out) return fail(ANE_ERR_INVALID_ARG, "bad arguments");
  std::vector<float> base((size_t)n * 3);
  fibonacci_dirs(n, base.data());
  for (int i = 0; i < n; ++i) {
    f3 d = rot3(rot16, ld3(base.data() + 3 * i));
    out[3 * i + 0] = d.x;
    out[3 * i + 1] = d.y;
    out[3 * i + 2] = d.z;
  }
  return ANE_OK;
}
