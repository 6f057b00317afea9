/*
 * ddgi_oracle.h -- CPU oracle for the DDGI probe update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (anerend_amd, libane_ddgi.so) never links, loads or calls it.
 *
 * It is a plain-C restatement of AneRend's probe-update GLSL and host passes
 * (paths relative to the reference checkout):
 *   assets/shaders/irradiance_probe_update.rgen   ray generation, probe placement
 *   assets/shaders/irradiance_probe_update.rchit  hit shading, shadow ray
 *   assets/shaders/irradiance_probe_update.rmiss  sky / shadow miss
 *   assets/shaders/pbr_light.glsl                 BRDF, sampleProbe, indirect light
 *   assets/shaders/probe_conv.comp                convolution + hysteresis
 *   assets/shaders/octahedron_helpers.glsl        octEncode / octDecode
 *   src/anerend/render/passes/IrradianceProbeTranslationPass.cpp  atlas shift
 * and of the acceleration structure the Vulkan driver builds, replaced by an
 * explicit LBVH contract (Karras 2012) and the watertight ray/triangle test of
 * Woop, Benthin and Wald (JCGT 2013).
 *
 * Every function documents the reference file:line it follows.  Where the
 * reference is implementation-defined (GLSL builtins, hardware filtering, the
 * driver's BVH) the contract chosen here is stated next to the code.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   ora_fibonacci_dirs     closed form: unit length, equally spaced z, sum ~ 0
 *   ora_probe_origin       closed form: integer grid, trunc-toward-zero camera cell
 *   ora_oct_encode/decode  round trip, texel centres, seam cases
 *   ora_scene_*            (LBVH) invariants: every primitive in exactly one leaf,
 *                          parents enclose children, keys sorted, depth bound;
 *                          traversal == brute force over all triangles
 *   ora_trace              brute force (ora_trace_bruteforce), analytic
 *                          plane/box hits
 *   ora_shade_hit          Cook-Torrance closed forms at N = L = V, at V and L
 *                          symmetric about N (H = N, cosines 0.6) and at cosines
 *                          (N.H, N.V, N.L, H.V) = (0.7 sqrt2, 0.6, 0.8, 1/sqrt2);
 *                          indirect Fresnel-roughness at N.V = 0.6; shadowed,
 *                          sun-below-horizon, emissive and miss cases
 *   ora_sample_probe,      constant atlas (E = c*pi/2); alpha = 0 reads the base
 *   ora_gi_lookup          probe's own tile; grid edges read the REPEAT-wrapped
 *                          tile; alpha = 1/2 on x, y, z and at a cell centre gives
 *                          the sqrt-space blend of distinct tiles with the wrap
 *                          weights of cosines 0, +-1/sqrt2, +-1; a linear ramp per
 *                          tile fixes the texel addressing (start 10 i + 1, end
 *                          start + 7, octEncode's y/z swap); Fresnel-roughness
 *                          with roughness^2 at N.V < 1.  Eleven planted mistakes
 *                          (ORA_MUTANT, ddgi_oracle.c) each fail one of these.
 *   ora_convolve_blend     float64 GEMM, uniform radiance => E = L, fixed point,
 *                          geometric series over k updates
 *   ora_translate          independent per-texel restatement of the two copies
 *   ora_scene_set_node32   decoded boxes contain the exact child boxes and stay within
 *                          one step of them; traversal with them == brute force
 *   ora_light_shadows      analytic distances from a light inside a closed box,
 *                          all-miss case (mean = range), disabled lights,
 *                          float64 weights and moments from the distances
 *
 * All arithmetic is IEEE binary32 with contraction off (-ffp-contract=off).
 */
#ifndef DDGI_ORACLE_H
#define DDGI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Oracle-only quirk bit of ora_config.quirks (the library takes quirks = 0):
 * emulate the reference's RGBA16F images (Q6).  Ray radiance (ProbeRaysIrrTex),
 * ray directions as the convolution reads them (ProbeRaysDirTex) and every atlas
 * texel the convolution writes (ProbeRaysConvTex) are rounded to IEEE binary16,
 * round to nearest even (IrradianceProbeRayTracingPass.cpp:42-68,
 * IrradianceProbeConvolvePass.cpp:91; rgen:108-109, probe_conv.comp:49-50,89).
 * Tracing keeps the binary32 direction, as the rgen does.  Hardware filtering at
 * reduced precision is not modelled.  With this bit the oracle measures how far
 * the reference's own storage moves the result from the FP32 contract. */
#define ORA_QUIRK_FP16 (1u << 0)

/* scene_ubo.glsl:5,6,12 */
#define ORA_FLAG_DDGI (1u << 4)
#define ORA_FLAG_DDGI_MULTI (1u << 5)
#define ORA_FLAG_BAKE (1u << 11)

/* Probe-update configuration; same field layout as ane_ddgi_config. */
typedef struct ora_config {
  int32_t probes_x, probes_y, probes_z; /* probe_helpers.glsl:1-2 */
  float probe_step[3];                  /* probe_helpers.glsl:3 */
  int32_t rays_per_probe;               /* probe_helpers.glsl:4 */
  float hysteresis;                     /* probe_conv.comp:85 */
  float t_min, t_max;                   /* irradiance_probe_update.rgen:91-92 */
  int32_t precision;                    /* ignored by the oracle (always fp32) */
  uint32_t quirks;
} ora_config;

/* Per-update constants; same field layout as ane_ddgi_frame. */
typedef struct ora_frame {
  float camera_pos[4];   /* scene_ubo.glsl:22 */
  float light_dir[4];    /* scene_ubo.glsl:24, direction the light travels */
  float sky_color[4];    /* scene_ubo.glsl:26 */
  float sun_intensity;   /* scene_ubo.glsl:31 */
  float sky_intensity;   /* scene_ubo.glsl:32 */
  uint32_t flags;        /* scene_ubo.glsl:34 */
  int32_t layer_begin;   /* push constant probeLayer, rgen:28 */
  int32_t layer_count;
  float rotation[16];    /* push constant randomRot (column-major mat4), rgen:27 */
  int32_t bake_tile[4];  /* scene_ubo.glsl:25 bakeTileInfo */
} ora_frame;

/* Input records.  Byte layouts follow the reference's GPU structs. */
typedef struct ora_mesh { /* GpuBuffers.h:38-44 (index count in place of blasRef) */
  float min_pos[4];
  float max_pos[4];
  uint32_t vertex_offset;
  uint32_t index_offset;
  uint32_t index_count;
  uint32_t reserved;
} ora_mesh;

typedef struct ora_material { /* GpuBuffers.h:29-36 */
  float base_col_fac[4];
  float emissive[4];
  int32_t bindless_indices[4];
  float met_rough[4];
} ora_material;

typedef struct ora_instance { /* tlas_update.comp:10-21 (VkAccelerationStructureInstanceKHR) */
  float transform[12];            /* 3x4 row-major object-to-world */
  uint32_t custom_index_and_mask; /* material bits 0-9, mesh bits 10-23, mask 24-31 */
  uint32_t sbt_offset_and_flags;
  uint64_t reserved;
} ora_instance;

/* Packed vertex: 24 words (bindless.glsl:60-68, Vertex.h:7-19). */
#define ORA_VERTEX_WORDS 24

typedef struct ora_hit {
  int32_t tri;   /* global triangle id (instance-major input order), -1 = miss */
  float t, u, v; /* u, v = barycentric weights of vertices 1 and 2 */
} ora_hit;

typedef struct ora_scene ora_scene;

/* ---- directions, placement, octahedral mapping ---- */
void ora_fibonacci_dirs(int32_t n, float* out_xyz);
void ora_rotate_dirs(int32_t n, const float* rot16, const float* in_xyz, float* out_xyz);
void ora_probe_origin(const ora_config* cfg, const ora_frame* fr, int32_t px, int32_t py,
                      int32_t pz, float* out_xyz);
void ora_oct_encode(const float* v3, float* out2);
void ora_oct_decode(const float* o2, float* out3);
void ora_texel_dirs(float* out_64x3);
/* W[64][n] (cosine weights, zero where <= 1e-6) and row sums; probe_conv.comp:49-61 */
void ora_conv_weights(int32_t n, const float* dirs_xyz, float* w, float* wsum);

/* ---- scene + LBVH ---- */
ora_scene* ora_scene_create(const float* vertices, int64_t nvert, const uint32_t* indices,
                            int64_t nidx, const ora_mesh* meshes, int32_t nmesh,
                            const ora_material* mats, int32_t nmat, const ora_instance* inst,
                            int32_t ninst);
void ora_scene_destroy(ora_scene* s);
int ora_scene_refit(ora_scene* s, const ora_instance* inst, int32_t ninst);
int64_t ora_scene_num_tris(const ora_scene* s);
int32_t ora_scene_max_depth(const ora_scene* s);
/* world-space triangles in input order: [ntri][9] */
void ora_scene_world_tris(const ora_scene* s, float* out);
/* LBVH export: keys/ids in sorted order, internal children (>= 0 internal,
 * < 0 leaf = ~sorted_position), parents, internal node boxes [n-1][6] (lo, hi) */
/* RT hard shadows (shadow.rgen:29-82): visibility (1 lit, 0 shadowed) of each
 * pixel of a w x h depth buffer; ivp = column-major inverse view-projection */
void ora_shadow_mask(const ora_scene* s, const float* depth, int32_t w, int32_t h, const float* ivp,
                     const float* light, float* out, uint64_t* counters);
/* Point-light shadow moments (light_shadow.rgen/.rchit/.rmiss, light_shadow_sum.comp):
 * lights = n x 8 floats {x, y, z, range, r, g, b, enabled (>= 0.5)}; 1024 rays per light.
 * dist_img (nullable): R32F image 32 x (32 n), ray y * 32 + x of light l at (32 l + x, y).
 * moments_img: RG32F image 8 x (10 n); texel (tx, ty) of light l at (10 l + 1 + tx, ty);
 * the border columns are not written.  counters as ora_trace. */
#define ORA_LIGHT_RAYS 1024
#define ORA_LIGHT_SIDE 32
void ora_light_shadows(const ora_scene* s, const float* lights, int32_t n, float* dist_img,
                       float* moments_img, uint64_t* counters);
/* the sum's weights: w[t][r] = pow(max(0, dot(dir_r, octDir_t)), 50), 0 unless > 0.001,
 * with dir_r rounded through RGBA8_UNORM; [64][1024] */
void ora_light_shadow_weights(float* w);
/* Quantized 32-byte nodes, as the library stores them for large scenes (8-bit child
 * bounds from the node box's lower corner in power-of-two steps, rounded outward):
 * on = 1 makes the traversal test the decoded boxes (refits keep them current).
 * qboxes: [ntri-1][12] decoded boxes {child0 lo, hi, child1 lo, hi}. */
void ora_scene_set_node32(ora_scene* s, int32_t on);
void ora_scene_qboxes(const ora_scene* s, float* out);
void ora_scene_bvh(const ora_scene* s, uint64_t* keys, uint32_t* sorted_ids, int32_t* left,
                   int32_t* right, int32_t* parent, float* node_boxes);

/* ---- traversal ---- */
/* counters (nullable): [0] internal nodes visited, [1] triangles tested */
void ora_trace(const ora_scene* s, const float* orig, const float* dir, int64_t nrays,
               float tmin, float tmax, int32_t any_hit, ora_hit* out, uint64_t* counters);
void ora_trace_bruteforce(const ora_scene* s, const float* orig, const float* dir,
                          int64_t nrays, float tmin, float tmax, int32_t any_hit, ora_hit* out);

/* ---- shading (rchit/rmiss) ---- */
void ora_sky_radiance(const ora_frame* fr, float* out3);
/* deferred GI lookup (deferred_tiled.comp:264-298): gbuf n x 12 floats
 * (pos, shadow, normal, metallic, albedo, roughness) -> out n x 4 (rgb, 0) */
void ora_gi_lookup(const ora_config* c, const ora_frame* f, const float* atlas, const float* gbuf,
                   int64_t n, float* out);
/* Mirror reflections (mirror_reflections.rgen/.rchit/.rmiss): out4 per pixel =
 * (irradiance, hit distance); normal4 = 4 floats per pixel (xyz used). */
void ora_mirror_reflections(const ora_scene* s, const ora_config* c, const ora_frame* f,
                            const float* atlas, const float* depth, const float* normal4, int32_t w,
                            int32_t h, const float* ivp, float* out4, uint64_t* counters4);
void ora_shade_hit(const ora_scene* s, const ora_config* cfg, const ora_frame* fr,
                   const float* atlas, const float* origin, const ora_hit* hit, float* out3,
                   int32_t* shadowed, uint64_t* counters);
void ora_sample_probe(const ora_config* cfg, const ora_frame* fr, const float* atlas,
                      const float* pos, const float* normal, float* out3);

/* Baked atlases of world tiles (Scene::setDDGIAtlas, Scene.cpp:409-438): tile_xz
 * [n][2] tile indices (x, z), atlases [n][H][W][4] binary32 (the RGBA16F texels
 * widened), tile_size in metres (TILE_SIZE, bindless.glsl:5).  Replaces the set
 * (n = 0 clears).  Global state of the oracle, read by sampleProbe and the
 * deferred lookup; window radius 2 around the camera tile (see ddgi_oracle.c). */
void ora_tiles_set(const ora_config* cfg, const int32_t* tile_xz, const float* atlases, int32_t n,
                   int32_t tile_size);

/* wall-clock ms of the last ora_update: [0] trace + shadow + shade, [1] convolve + blend */
void ora_last_update_stage_ms(double* out2);

/* binary32 -> binary16 (round to nearest, ties to even; >= 65520 -> inf) -> binary32 */
void ora_f16_round(const float* in, float* out, int64_t n);

/* ---- convolution + hysteresis, translation ---- */
/* radiance: [P_layers][3][n] for the probes of layers [layer_begin, +layer_count) */
void ora_convolve_blend(const ora_config* cfg, const ora_frame* fr, const float* dirs_xyz,
                        const float* radiance, float* atlas);
/* last_idx: in/out camera cell (the pass keeps it as _lastProbeIdx); returns 1 if shifted */
int ora_translate(const ora_config* cfg, const float* camera_pos, int32_t* last_idx,
                  float* atlas);

/* ---- one full update: trace + shade (reads atlas) then convolve + blend ----
 * hits/shadowed/radiance are optional outputs for the traced layers,
 * ray-major [probe][ray] (radiance [probe][3][ray]). counters[4]: primary
 * nodes, primary tris, shadow nodes, shadow tris. */
void ora_update(const ora_scene* s, const ora_config* cfg, const ora_frame* fr, float* atlas,
                ora_hit* hits, uint8_t* shadowed, float* radiance, uint64_t* counters);

/* atlas helpers */
int64_t ora_atlas_width(const ora_config* cfg);
int64_t ora_atlas_height(const ora_config* cfg);
void ora_atlas_clear(const ora_config* cfg, float* atlas);

int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
