/*
 * ane_ddgi.h -- C-ABI of the MI355X (gfx950) DDGI probe-update engine.
 *
 * Drop-in for AneRend's probe update (paths relative to the reference checkout):
 *
 *   reference pass / call                                     replaced by
 *   --------------------------------------------------------  -------------------------
 *   IrradianceProbeTrans exe: 2x vkCmdCopyImage               ane_ddgi_translate
 *     (src/anerend/render/passes/IrradianceProbeTranslationPass.cpp:56-186)
 *   IrradianceProbeRT exe: vkCmdTraceRaysKHR                  ane_ddgi_trace
 *     (IrradianceProbeRayTracingPass.cpp:95-165, shaders irradiance_probe_update.*)
 *   IrradianceProbeConvolve exe: vkCmdDispatch                ane_ddgi_convolve
 *     (IrradianceProbeConvolvePass.cpp:109-143, shader probe_conv.comp)
 *   BLASUpdate / TLASUpdate: vkCmdBuildAccelerationStructures ane_scene_create / ane_scene_refit
 *     (VulkanRenderer.cpp:1267-1417, UpdateBlasPass.cpp:16-209, UpdateTLASPass.cpp:35-249)
 *   ProbeRaysConvTex resource (RGBA16F atlas)                 ane_ddgi_atlas_get / _set / _device
 *     (IrradianceProbeConvolvePass.cpp:82-99)
 *
 * Conventions
 *   - Every function returns an ane_status; nothing throws across the ABI.
 *     The message of the last failure is ane_last_error() (thread-local).
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  All
 *     device work of a call is ordered on it; host-buffer calls synchronize it.
 *   - Host arrays are copied during the call; no pointer is retained.
 *   - One ane_ddgi_ctx per probe grid; a context is not thread-safe (the
 *     reference's pass objects are not either).  Scenes are read-only during
 *     traces and may be shared by several contexts.
 *
 * Data layouts
 *   vertices   [nvert][24] words, the 96-byte PackedVertex (bindless.glsl:60-68)
 *   atlas      [10*probes_z*probes_y][10*probes_x][4] FP32 RGBA, the reference's
 *              ProbeRaysConvTex texel layout (probe (x,y,z) interior at
 *              (10x+1, 10*probes_z*y + 10z + 1), 1-texel borders)
 *   probes     p = (y*probes_z + z)*probes_x + x  (layer-major)
 *   radiance   [probe][3][ray] (channel-planar per probe)
 */
#ifndef ANE_DDGI_H
#define ANE_DDGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANE_DDGI_ABI_VERSION 1

typedef enum ane_status {
  ANE_OK = 0,
  ANE_ERR_INVALID_ARG = 1,
  ANE_ERR_OUT_OF_MEMORY = 2,
  ANE_ERR_ADMISSION_REFUSED = 3,
  ANE_ERR_DEVICE = 4,
  ANE_ERR_NOT_READY = 5,
  ANE_ERR_UNSUPPORTED = 6,
  ANE_ERR_CANCELLED = 7
} ane_status;

/* scene_ubo.glsl:1-12 flag bits honoured by the probe update */
#define ANE_FLAG_DDGI (1u << 4)
#define ANE_FLAG_DDGI_MULTI (1u << 5) /* indirect light from the previous atlas (rchit:98) */
#define ANE_FLAG_BAKE (1u << 11)      /* bake placement + accumulate blend (rgen:61, probe_conv.comp:67) */

typedef enum ane_precision {
  ANE_PRECISION_FP32 = 0, /* plain path: FP32 radiance, FP32 convolution */
  ANE_PRECISION_BF16 = 1, /* BF16 radiance and weights, MFMA, FP32 accumulate */
  ANE_PRECISION_FP8 = 2,  /* FP8-e4m3 radiance (FP32 scale per probe, channel and 32 rays) and weights,
                              MFMA, FP32 accumulate */
  ANE_PRECISION_FP16 = 3  /* FP16 radiance (saturating at 65504) and weights, MFMA, FP32 accumulate:
                             the 16-bit mode whose 10-bit mantissa keeps correlated rounding (runs
                             of nearly equal radiance: sky, lit planes) within 2e-3 of FP64 */
} ane_precision;

/* Probe grid constants: probe_helpers.glsl:1-7 (compile-time in the reference),
 * probe_conv.comp:85 (hysteresis), irradiance_probe_update.rgen:91-92 (t range).
 * The octahedral tile is fixed at 8x8 texels + 1-texel border. */
typedef struct ane_ddgi_config {
  int32_t probes_x, probes_y, probes_z; /* NUM_PROBES_PER_PLANE (x, z), NUM_PROBE_PLANES (y) */
  float probe_step[3];                  /* PROBE_STEP */
  int32_t rays_per_probe;               /* NUM_RAYS_PER_PROBE, 1..4096 (need not be square) */
  float hysteresis;                     /* 0.96 in the reference */
  float t_min, t_max;                   /* 0.1, 50.0 in the reference */
  int32_t precision;                    /* ane_precision */
  uint32_t quirks;                      /* reserved, must be 0 */
} ane_ddgi_config;

/* Per-update constants: the GPUSceneData fields the path reads
 * (GpuBuffers.h:125-148, filled at VulkanRenderer.cpp:1201-1243) and the
 * ray-generation push constants (IrradianceProbeRayTracingPass.cpp:14-17). */
typedef struct ane_ddgi_frame {
  float camera_pos[4];
  float light_dir[4]; /* direction the sun light travels; the sun is at -light_dir */
  float sky_color[4];
  float sun_intensity;
  float sky_intensity;
  uint32_t flags;      /* ANE_FLAG_* */
  int32_t layer_begin; /* first probe layer (y) to update (push constant probeLayer) */
  int32_t layer_count; /* number of layers; probes_y = full update */
  float rotation[16];  /* randomRot, column-major mat4; the upper 3x3 rotates ray directions */
  int32_t bake_tile[4];/* bakeTileInfo: tile x, tile z, tile size (bake mode) */
} ane_ddgi_frame;

/* This library's own mesh record: the fields of GPUMeshInfo (GpuBuffers.h:38-44)
 * that the path reads, with the mesh's index count (InternalMesh::_numIndices,
 * internal/InternalMesh.h:12, used as primitiveCount at VulkanRenderer.cpp:1313)
 * where GPUMeshInfo keeps its 64-bit _blasRef.  It is NOT GPUMeshInfo: an array the
 * reference filled goes to ane_scene_create_mesh_info with the counts beside it. */
typedef struct ane_mesh {
  float min_pos[4];
  float max_pos[4];
  uint32_t vertex_offset;
  uint32_t index_offset;
  uint32_t index_count;
  uint32_t reserved;   /* must be 0 */
} ane_mesh;

/* GPUMeshInfo exactly as the reference fills it (GpuBuffers.h:38-44, written by
 * VulkanRenderer.cpp:3486-3493), 48 bytes.  blas_ref (the Vulkan BLAS device
 * address) means nothing here and is ignored. */
typedef struct ane_gpu_mesh_info {
  float min_pos[4];
  float max_pos[4];
  uint32_t vertex_offset;
  uint32_t index_offset;
  uint64_t blas_ref;
} ane_gpu_mesh_info;

typedef struct ane_material { /* GPUMaterialInfo (GpuBuffers.h:29-36); textures unsupported (-1) */
  float base_col_fac[4];
  float emissive[4];
  int32_t bindless_indices[4];
  float met_rough[4];
} ane_material;

typedef struct ane_instance { /* VkAccelerationStructureInstanceKHR as written by tlas_update.comp:94-97 */
  float transform[12];            /* 3x4 row-major object-to-world */
  uint32_t custom_index_and_mask; /* material bits 0-9, mesh bits 10-23, mask bits 24-31 */
  uint32_t sbt_offset_and_flags;
  uint64_t reserved;
} ane_instance;

typedef struct ane_hit {
  int32_t tri;   /* global triangle id (instance-major input order), -1 = miss */
  float t, u, v; /* hit distance; barycentric weights of vertices 1 and 2 */
} ane_hit;

/* A point light as the ray-tracing passes read it: GPULight (GpuBuffers.h:69-72,
 * bindless.glsl:21-25), 32 bytes. */
#define ANE_MAX_LIGHTS 1024 /* MAX_NUM_LIGHTS, light_shadow_helpers.glsl:4 */
typedef struct ane_light {
  float world_pos[4]; /* xyz position, w range */
  float color[4];     /* rgb colour, w enabled (>= 0.5) */
} ane_light;

typedef struct ane_scene_info {
  int64_t num_triangles;
  int32_t num_instances;
  int32_t bvh_depth;      /* max root-to-leaf edge count of the LBVH */
  int64_t device_bytes;   /* HBM held by the scene */
  float build_ms;         /* last build or refit, device time */
  int32_t node_format;    /* traversal layout in use: ANE_NODES_FULL or ANE_NODES_QUANTIZED */
} ane_scene_info;

/* Traversal node layouts (ane_scene_set_node_format). */
#define ANE_NODES_AUTO 0      /* quantized from ANE_NODES_QUANTIZED_MIN_TRIS triangles up (default) */
#define ANE_NODES_FULL 1      /* 64 bytes: binary32 child boxes */
#define ANE_NODES_QUANTIZED 2 /* 32 bytes: 8-bit child bounds, rounded outward */
#define ANE_NODES_QUANTIZED_MIN_TRIS 2000000

typedef struct ane_ddgi_stats {
  /* device time per stage from hipEvent pairs, ms: the last call and the mean
   * over all calls since the last ane_ddgi_stats_reset.  Stage order:
   * setup, trace (primary rays and, in the same launch, the shadow ray of
   * every hit unless a sun grid serves them), sun grid (the build of the shadow
   * rays' candidate lists, only in the updates that build one: a new scene
   * version or sun), shade, convolve, translate, shadow (the shadow rays of the
   * hits against their sun-grid cells, in the updates that have a grid). */
  float last_ms[7];
  float mean_ms[7];
  uint64_t calls[7];
  /* traversal work of the last trace (only while counting is enabled) */
  uint64_t primary_nodes, primary_tris, shadow_nodes, shadow_tris;
  uint64_t rays;          /* primary rays of the last trace */
  int64_t device_bytes;   /* HBM held by the context */
  /* the context's sun grid of the probe shadow rays: 0 none built, 1 in use,
   * 2 built but not used (over its capacity, or more estimated candidates per
   * ray than traversal pays for); the estimate of its last build */
  int32_t sun_grid_state;
  float sun_grid_candidates;
  /* 1 when a probe ray of a trace since the last ane_ddgi_check / _stats_reset
   * ended without a result (stack over its spill capacity, the step limit of
   * ane_ddgi_set_traversal_limit, or a malformed hierarchy); such a ray reads as
   * a miss.  ane_ddgi_stats_get then returns ANE_ERR_DEVICE with *out filled. */
  uint32_t traversal_fault;
} ane_ddgi_stats;

/* One G-buffer texel for the deferred GI lookup (deferred_tiled.comp:264-298 inputs):
 * world position, the sun shadow factor of the pixel, world normal, metallic,
 * albedo (linear), roughness. */
typedef struct ane_gbuffer_texel {
  float pos[3];
  float shadow;
  float normal[3];
  float metallic;
  float albedo[3];
  float roughness;
} ane_gbuffer_texel;

typedef struct ane_ddgi_ctx ane_ddgi_ctx;
typedef struct ane_scene ane_scene;

/* ---- library ---- */
int ane_abi_version(void);
const char* ane_last_error(void);
const char* ane_status_string(int status);
/* Number of the HIP device used by this thread (hipGetDevice). */
int ane_device(int32_t* device);
/* Wait for all work queued on `stream` (hipStreamSynchronize). */
int ane_stream_synchronize(void* stream);

/* ---- scene: replaces BLAS/TLAS build (VulkanRenderer.cpp:1267-1417, UpdateTLASPass.cpp:184-228) ----
 * Uploads the geometry, flattens every instance to world-space triangles and
 * builds the LBVH on the GPU (Morton codes, radix sort, Karras hierarchy,
 * bottom-up boxes). */
int ane_scene_create(const float* vertices, int64_t num_vertices, const uint32_t* indices,
                     int64_t num_indices, const ane_mesh* meshes, int32_t num_meshes,
                     const ane_material* materials, int32_t num_materials,
                     const ane_instance* instances, int32_t num_instances, void* stream,
                     ane_scene** out);
/* ane_scene_create for the reference's own buffers: `meshes` is its GPUMeshInfo
 * array, taken as it is (blas_ref ignored), and mesh_index_counts[m] is mesh m's
 * index count (InternalMesh::_numIndices, the BLAS primitiveCount * 3 of
 * VulkanRenderer.cpp:1313), which GPUMeshInfo does not carry.  Same checks, same
 * scene as ane_scene_create with ane_mesh{min, max, offsets, count, 0}. */
int ane_scene_create_mesh_info(const float* vertices, int64_t num_vertices,
                               const uint32_t* indices, int64_t num_indices,
                               const ane_gpu_mesh_info* meshes, const uint32_t* mesh_index_counts,
                               int32_t num_meshes, const ane_material* materials,
                               int32_t num_materials, const ane_instance* instances,
                               int32_t num_instances, void* stream, ane_scene** out);
/* New instance transforms, same topology (UpdateBlasPass.cpp:131-138 MODE_UPDATE).
 * Instances must keep their mesh; `instances` is a host array. */
int ane_scene_refit(ane_scene* scene, const ane_instance* instances, int32_t num_instances,
                    void* stream);
/* Full rebuild (sort + hierarchy) after new transforms (UpdateTLASPass.cpp:184-228 BUILD). */
int ane_scene_rebuild(ane_scene* scene, const ane_instance* instances, int32_t num_instances,
                      void* stream);
int ane_scene_info_get(const ane_scene* scene, ane_scene_info* out);
/* Waits for `stream`, then reads and clears the traversal fault flags of the
 * scene's last shadow-mask and light-shadow passes (those passes report a fault
 * by themselves only when their counters are requested): ANE_ERR_DEVICE if a
 * ray ended without a result, else ANE_OK. */
int ane_scene_check(ane_scene* scene, void* stream);
int ane_scene_destroy(ane_scene* scene);

/* ---- probe grid context ---- */
int ane_ddgi_create(const ane_ddgi_config* config, ane_ddgi_ctx** out);
int ane_ddgi_destroy(ane_ddgi_ctx* ctx);

/* IrradianceProbeTrans (IrradianceProbeTranslationPass.cpp:56-186): if the camera
 * cell (trunc(cam.x / step.x), trunc(cam.z / step.z)) differs from the last one
 * the atlas tiles shift by one probe toward the motion (entry edge keeps its old
 * texels).  *shifted (nullable) receives 1 if the atlas moved.  The last cell
 * starts at (0, 0) like _lastProbeIdx. */
int ane_ddgi_translate(ane_ddgi_ctx* ctx, const float* camera_pos3, int32_t* shifted,
                       void* stream);
/* IrradianceProbeRT: trace probe rays of layers [layer_begin, +layer_count),
 * shade hits (shadow ray, direct + indirect + emissive) and store radiance.
 * Shadow rays all point at the sun: from the second trace that sees the same
 * scene (no refit or rebuild in between) and the same sun direction, the context
 * keeps a grid of the triangles projected along the sun and a shadow ray tests
 * the triangles of its cell instead of traversing; the decisions are identical.
 * The grid takes about 28 bytes per triangle plus 4 per cell (one cell per
 * triangle) and is rebuilt on the stream when that scene or sun recurs. */
int ane_ddgi_trace(ane_ddgi_ctx* ctx, const ane_scene* scene, const ane_ddgi_frame* frame,
                   void* stream);
/* IrradianceProbeConvolve: cosine-weighted octahedral convolution of the stored
 * radiance and hysteresis blend (or bake accumulation) into the atlas. */
int ane_ddgi_convolve(ane_ddgi_ctx* ctx, const ane_ddgi_frame* frame, void* stream);
/* trace then convolve. */
int ane_ddgi_update(ane_ddgi_ctx* ctx, const ane_scene* scene, const ane_ddgi_frame* frame,
                    void* stream);

/* Deferred GI lookup (deferred_tiled.comp:264-298 with calcIndirectDiffuseLight,
 * pbr_light.glsl:365-388): ambient light of n G-buffer texels from the current
 * atlas.  With ANE_FLAG_DDGI set in frame->flags: texels farther from the
 * camera than probes_x * probe_step.x / 2 get skyIntensity * 0.3 * albedo *
 * clamp(shadow, 0.3, 1), the others kD * sampleProbe(pos, normal) * albedo;
 * without it 0.3 * albedo * clamp(shadow, 0.3, 1).  gbuf and out_rgba
 * (n x float4, w = 0) are DEVICE pointers, ordered on `stream`. */
int ane_ddgi_gi_lookup(ane_ddgi_ctx* ctx, const ane_ddgi_frame* frame, const ane_gbuffer_texel* gbuf,
                       int64_t n, float* out_rgba, void* stream);

/* Mirror reflections (replaces the MirrorReflections pass's vkCmdTraceRaysKHR,
 * mirror_reflections.rgen:35-82 + .rchit:24-119 + .rmiss:17-22).  Per pixel of a
 * width x height depth buffer: the world position (inv_view_proj16, column-major,
 * as ane_scene_shadow_mask), the normal normalize(normal4[pixel].xyz), a
 * closest-hit ray over [0.1, 50] along reflect(normalize(pos - camera_pos),
 * normal); at a hit, a shadow ray toward -light_dir and the probe hit shading with
 * the normal taken through mat3(ObjectToWorld) and the indirect term from the
 * context's atlas when frame->flags has ANE_FLAG_DDGI.  out_rgba = (irradiance,
 * distance to the hit); a miss gives (sky radiance, 1e6), a sky pixel (depth 1)
 * gives 0.  A zero normal gives a non-finite direction, which is no ray (a miss).  depth, normal4 (4 floats per pixel) and out_rgba are DEVICE pointers,
 * ordered on `stream`.  counters4 (host, nullable): node visits and triangle
 * tests of the reflection rays, then of the shadow rays (the stream is then
 * synchronized). */
int ane_ddgi_mirror_reflections(ane_ddgi_ctx* ctx, const ane_scene* scene, const ane_ddgi_frame* frame,
                                const float* depth, const float* normal4, int32_t width, int32_t height,
                                const float* inv_view_proj16, float* out_rgba, uint64_t* counters4,
                                void* stream);

/* ---- probe state (ProbeRaysConvTex) ---- */
int ane_ddgi_atlas_get(ane_ddgi_ctx* ctx, float* host_rgba, void* stream);
int ane_ddgi_atlas_set(ane_ddgi_ctx* ctx, const float* host_rgba, void* stream);
int ane_ddgi_atlas_device(ane_ddgi_ctx* ctx, float** device_rgba, int64_t* width,
                          int64_t* height);
int ane_ddgi_translation_state(ane_ddgi_ctx* ctx, int32_t* last_cell2, int32_t set);
/* The atlas in the format the reference stores a baked tile
 * (VulkanRenderer.cpp:2705-2770 downloadDDGIAtlas, asset::Texture RGBA16F_SFLOAT):
 * width 10*probes_x, height 10*probes_z*probes_y, 4 x IEEE binary16 per texel,
 * row-major; export rounds to nearest even (values beyond 65504 become inf),
 * import widens exactly. */
int ane_ddgi_atlas_export_rgba16f(ane_ddgi_ctx* ctx, uint16_t* host_rgba16f, void* stream);
int ane_ddgi_atlas_import_rgba16f(ane_ddgi_ctx* ctx, const uint16_t* host_rgba16f, void* stream);

/* ---- radiance (ProbeRaysIrrTex) in the context's precision ----
 * set: host FP32 [probe][3][ray] for the layers of `frame`, quantized on upload. */
int ane_ddgi_radiance_set(ane_ddgi_ctx* ctx, const ane_ddgi_frame* frame, const float* host,
                          void* stream);
/* get: FP32 [probe][3][ray] (dequantized) for the layers of the last trace. */
int ane_ddgi_radiance_get(ane_ddgi_ctx* ctx, float* host, void* stream);

/* ---- parity / diagnostics ---- */
/* Per-ray hits [probe][ray] and shadow flags (0 lit, 1 shadowed, 255 miss) of the last trace. */
int ane_ddgi_debug_rays(ane_ddgi_ctx* ctx, ane_hit* hits, uint8_t* shadowed, void* stream);
/* Count traversal work in the next traces (slower kernels).  mode 0: off;
 * 1: every ray traverses the hierarchy (node visits, triangle tests); 2: the
 * shadow rays take the sun grid when it is current, and their counters hold
 * cell lookups and candidate triangle tests instead. */
int ane_ddgi_set_counting(ane_ddgi_ctx* ctx, int32_t mode);
/* Fills *out; returns ANE_ERR_DEVICE (with *out filled) while a traversal
 * fault is pending. */
int ane_ddgi_stats_get(const ane_ddgi_ctx* ctx, ane_ddgi_stats* out);
/* Resets the stage timers and clears the traversal fault flags. */
int ane_ddgi_stats_reset(ane_ddgi_ctx* ctx);
/* Baked DDGI atlases of world tiles (Scene::setDDGIAtlas, Scene.cpp:409-438; the
 * per-frame GPUTileInfo window, GpuBuffers.h:74-76, VulkanRenderer.cpp:3698-3738).
 * tile_xz [n][2] = tile indices (x, z) of tiles tile_size metres wide (TILE_SIZE,
 * bindless.glsl:5: 32); rgba16f [n][H][W][4] = each tile's baked atlas, the
 * context's atlas layout in IEEE binary16 (what ane_ddgi_atlas_export_rgba16f and
 * downloadDDGIAtlas, VulkanRenderer.cpp:2705-2770, produce).  Replaces the set
 * (n = 0 clears); tiles must be distinct.  From the next call on, a shading point
 * in a tile within 2 tiles of the camera's (both axes; floor(pos / tile_size))
 * takes its probes from that tile's grid (pbr_light.glsl:251-275): the deferred
 * lookup samples the tile's atlas, the probe and mirror rays the live atlas (as
 * the reference's closest-hit shaders do).  Waits for `stream` first. */
int ane_ddgi_tiles_set(ane_ddgi_ctx* ctx, const int32_t* tile_xz, const uint16_t* rgba16f,
                       int32_t num_tiles, int32_t tile_size, void* stream);
/* Validation in FP64 on the GPU (no counterpart in the reference): the error of
 * conv's last convolution against a float64 evaluation of the reference's sum
 * (probe_conv.comp:49-61) over the FP32 radiance `ref` holds.  Use: give both
 * contexts the same grid and atlas, trace the same frame in ref (ane_ddgi_trace,
 * or set its radiance) and update conv with it (any precision).  Per texel and
 * channel, E = (conv atlas - h * ref atlas) / (1 - h) with conv's hysteresis h,
 * E64 = sum W L / sum W in float64 (W = max(0, d . o) over the FP32 directions,
 * kept above 1e-6), rel = |E - E64| / max(|E64|, rel_floor); out: the max and
 * the root mean square of rel over the layers ref holds.  Waits for `stream`. */
int ane_ddgi_conv_error_fp64(ane_ddgi_ctx* conv, const ane_ddgi_ctx* ref, double rel_floor,
                             double* max_rel, double* rms_rel, void* stream);
/* Waits for `stream`, then reads and clears the context's traversal fault flags
 * (probe traces and the last mirror pass): ANE_ERR_DEVICE if any ray since the
 * last check ended without a result, else ANE_OK.  The reference has no such
 * error path (its traversal is the driver's); the probe passes themselves stay
 * asynchronous, so this is the call that surfaces it. */
int ane_ddgi_check(ane_ddgi_ctx* ctx, void* stream);
/* Caps the node visits of each probe ray (0, the default: 2 * triangles + 8,
 * which a valid hierarchy never reaches).  A ray at the cap ends as a miss and
 * raises the traversal fault: a latency guard for serving loops. */
int ane_ddgi_set_traversal_limit(ane_ddgi_ctx* ctx, uint32_t max_node_visits);
/* LBVH export in sorted order: keys, triangle ids, internal children (>= 0
 * internal, < 0 leaf ~position), parents, internal boxes [n-1][6] (lo, hi). */
int ane_scene_debug_bvh(const ane_scene* scene, uint64_t* keys, uint32_t* ids, int32_t* left,
                        int32_t* right, int32_t* parent, float* boxes);
int ane_scene_debug_world_tris(const ane_scene* scene, float* tris9);
/* Trace arbitrary rays (host arrays) against a scene: closest or any hit.  A
 * zero or non-finite direction is no ray (a miss, no traversal). */
int ane_scene_trace_rays(const ane_scene* scene, const float* origins, const float* dirs,
                         int64_t num_rays, float t_min, float t_max, int32_t any_hit,
                         ane_hit* hits, uint64_t* counters2, void* stream);

/* ane_scene_trace_rays without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc; it synchronizes the stream, so it is
 * not meant for graph capture).  origins, dirs
 * ([num_rays][3] float) and hits are DEVICE arrays the caller owns; counters2 is
 * a nullable host array.  The traversal's scratch (ray counter, counters, fault
 * flag and spilled stack) is one 16 MiB device array per GPU, held until the
 * process ends: calls that overlap on one GPU race on it.  Scenes whose hierarchy
 * is too deep for that stack at one block per CU (depth above about 140) are
 * refused with ANE_ERR_UNSUPPORTED.  Waits for `stream` before returning. */
int ane_scene_trace_rays_noalloc(const ane_scene* scene, const float* origins, const float* dirs,
                                 int64_t num_rays, float t_min, float t_max, int32_t any_hit,
                                 ane_hit* hits, uint64_t* counters2, void* stream);

/* RT hard shadows (ShadowRayTracingPass, shadow.rgen:29-82): for every pixel of a
 * width x height depth buffer, world position = (invViewProj * (uv * 2 - 1, depth, 1)).xyz / w
 * with uv = pixelToUv(pixel) (helpers.glsl:1-4), then an any-hit ray from it along
 * -light_dir (not normalized, as the reference) over [0.1, 50]: visibility 0 if
 * something is hit, else 1; depth == 1 (sky) gives 1 without a ray.
 * inv_view_proj16 is column-major; depth and out_visibility are DEVICE pointers
 * (width * height floats), ordered on `stream`.  counters2 (host, nullable)
 * receives node visits and triangle tests (the stream is then synchronized).  The
 * rays, hits and sun grid live on the scene: calls on one scene must be ordered on
 * one stream or must not overlap. */
int ane_scene_shadow_mask(ane_scene* scene, const float* depth, int32_t width, int32_t height,
                          const float* inv_view_proj16, const float* light_dir3, float* out_visibility,
                          uint64_t* counters2, void* stream);

/* Point-light shadow moments (replaces LightShadowRT, LightShadowRayTracingPass.cpp:72-99
 * with light_shadow.rgen:40-86, .rchit:29-32, .rmiss:12-15; and LightShadowSum,
 * LightShadowSumPass.cpp:65-87 with light_shadow_sum.comp:16-71).  Each of the
 * num_lights (<= 1024) lights casts 1024 rays from its position along the
 * unrotated Fibonacci directions (rgen:27-38) over [0.1, range]; a ray's distance is
 * its closest hit, or the range on a miss, or 0 for a disabled light (color[3] < 0.5;
 * as in the shaders, a NaN color[3] is enabled).
 *   dist_img     (DEVICE, nullable) R32F 32 x (32 num_lights) image:
 *                ray y * 32 + x of light l at pixel (32 l + x, y);
 *   moments_img  (DEVICE) RG32F 8 x (10 num_lights) image: for texel (tx, ty) of
 *                light l's 8x8 octahedral tile, at pixel (10 l + 1 + tx, ty), the
 *                mean and mean square distance weighted by pow(max(0, dot(dir,
 *                octDir)), 50) over the rays whose weight is above 0.001, dir being
 *                rounded through RGBA8 as the reference's direction image; the border
 *                columns 10 l and 10 l + 9 are not written.
 * The lights are read from HOST memory during the call; the images are ordered on
 * `stream`.  counters2 (host, nullable) receives node visits and triangle tests
 * (the stream is then synchronized).  Enabled lights need a finite position and
 * range.  Like ane_scene_shadow_mask, the call uses workspace kept on the scene:
 * calls on one scene must be ordered on one stream or must not overlap. */
int ane_scene_light_shadows(ane_scene* scene, const ane_light* lights, int32_t num_lights,
                            float* dist_img, float* moments_img, uint64_t* counters2, void* stream);

/* Traversal layout of a scene's hierarchy (no counterpart in the reference, whose
 * acceleration structures are the driver's).  ANE_NODES_FULL keeps 64-byte nodes
 * with binary32 child boxes.  ANE_NODES_QUANTIZED stores 32-byte nodes: each child
 * bound is an 8-bit offset from the node box's lower corner in a power-of-two step,
 * rounded outward, so decoded boxes contain the exact ones.  That halves the node
 * bytes for about 2-3% more node visits, which pays once the nodes no longer fit
 * the caches.  ANE_NODES_AUTO (the default) picks quantized nodes from
 * ANE_NODES_QUANTIZED_MIN_TRIS triangles up.  Hits are identical in every layout;
 * traversal counters count the visits of the layout in use.  The call waits for
 * the device to be idle, re-encodes the nodes on `stream` and returns when they are
 * complete; refits and rebuilds re-encode them on their own stream. */
int ane_scene_set_node_format(ane_scene* scene, int32_t format, void* stream);

/* ---- host-only helpers (no device work) ---- */
/* Fibonacci-lattice ray directions rotated by mat3(rot16) (rgen:35-46,86-88). */
int ane_ddgi_ray_directions(int32_t num_rays, const float* rot16, float* out_xyz);

#ifdef __cplusplus
}
#endif
#endif /* ANE_DDGI_H */
