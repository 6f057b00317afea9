// ddgi_frames -- runs the probe passes frame by frame through the frame graph,
// the way the reference's render loop drives them (VulkanRenderer.cpp:4471-4475:
// IrradianceProbeTrans, then IrradianceProbeRT, then IrradianceProbeConvolve).
//
//   ddgi_frames SCENE OUTDIR [--probes X Y Z] [--step SX SY SZ] [--rays R]
//               [--frames F] [--precision fp32|bf16|fp8] [--seed S] [--full]
//               [--camera X Y Z] [--velocity VX VY VZ] [--hysteresis H]
//
// SCENE is a scene file (layout in load_scene).  The camera starts at --camera
// and moves by --velocity each frame.  Writes OUTDIR/frames.bin (the
// ane_ddgi_frame of every traced frame, in order), OUTDIR/atlas.f32 (final FP32
// RGBA atlas) and prints one line per frame with the pass timings.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ddgi_passes.hpp"

namespace {

struct SceneFile {
  std::vector<float> vertices;
  std::vector<uint32_t> indices;
  std::vector<ane_mesh> meshes;
  std::vector<ane_material> materials;
  std::vector<ane_instance> instances;
};

template <class T>
bool read_n(FILE* fp, std::vector<T>& v, int64_t n)
{
  v.resize((size_t)n);
  return n == 0 || std::fread(v.data(), sizeof(T), (size_t)n, fp) == (size_t)n;
}

// magic "ANESCN01", int64 num_vertices, int64 num_indices, int32 num_meshes,
// int32 num_materials, int32 num_instances, int32 0, then the arrays in that
// order: 24 floats per vertex, uint32 indices, ane_mesh, ane_material, ane_instance.
bool load_scene(const char* path, SceneFile& s)
{
  FILE* fp = std::fopen(path, "rb");
  if (!fp) {
    std::fprintf(stderr, "cannot open %s: %s\n", path, std::strerror(errno));
    return false;
  }
  char magic[8];
  int64_t nv = 0, ni = 0;
  int32_t c[4] = {0, 0, 0, 0};
  bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, "ANESCN01", 8) == 0 &&
            std::fread(&nv, 8, 1, fp) == 1 && std::fread(&ni, 8, 1, fp) == 1 &&
            std::fread(c, 4, 4, fp) == 4 && nv >= 0 && ni >= 0 && c[0] >= 0 && c[1] >= 0 &&
            c[2] >= 0;
  ok = ok && read_n(fp, s.vertices, nv * 24) && read_n(fp, s.indices, ni) &&
       read_n(fp, s.meshes, c[0]) && read_n(fp, s.materials, c[1]) &&
       read_n(fp, s.instances, c[2]);
  std::fclose(fp);
  if (!ok) std::fprintf(stderr, "%s: not a scene file or truncated\n", path);
  return ok;
}

bool write_file(const std::string& path, const void* p, size_t n)
{
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) {
    std::fprintf(stderr, "cannot write %s: %s\n", path.c_str(), std::strerror(errno));
    return false;
  }
  const bool ok = std::fwrite(p, 1, n, fp) == n;
  return std::fclose(fp) == 0 && ok;
}

int usage()
{
  std::fprintf(stderr,
               "usage: ddgi_frames SCENE OUTDIR [--probes X Y Z] [--step SX SY SZ] [--rays R]\n"
               "       [--frames F] [--precision fp32|bf16|fp8] [--seed S] [--full]\n"
               "       [--camera X Y Z] [--velocity VX VY VZ] [--hysteresis H]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc < 3) return usage();
  const char* scene_path = argv[1];
  const std::string outdir = argv[2];
  ane_ddgi_config cfg{};
  cfg.probes_x = 8; cfg.probes_y = 8; cfg.probes_z = 8;
  cfg.probe_step[0] = 1.0f; cfg.probe_step[1] = 2.0f; cfg.probe_step[2] = 1.0f;
  cfg.rays_per_probe = 256;
  cfg.hysteresis = 0.96f;
  cfg.t_min = 0.1f;
  cfg.t_max = 50.0f;
  cfg.precision = ANE_PRECISION_FP32;
  int frames = 8;
  uint64_t seed = 1;
  bool full = false;
  float cam[3] = {0.5f, 1.5f, 0.5f}, vel[3] = {0.0f, 0.0f, 0.0f};
  for (int i = 3; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int k) { return i + k < argc; };
    if (a == "--probes" && need(3)) {
      cfg.probes_x = std::atoi(argv[++i]); cfg.probes_y = std::atoi(argv[++i]); cfg.probes_z = std::atoi(argv[++i]);
    } else if (a == "--step" && need(3)) {
      for (int k = 0; k < 3; ++k) cfg.probe_step[k] = std::strtof(argv[++i], nullptr);
    } else if (a == "--rays" && need(1)) {
      cfg.rays_per_probe = std::atoi(argv[++i]);
    } else if (a == "--frames" && need(1)) {
      frames = std::atoi(argv[++i]);
    } else if (a == "--hysteresis" && need(1)) {
      cfg.hysteresis = std::strtof(argv[++i], nullptr);
    } else if (a == "--precision" && need(1)) {
      const std::string p = argv[++i];
      if (p == "fp32") cfg.precision = ANE_PRECISION_FP32;
      else if (p == "bf16") cfg.precision = ANE_PRECISION_BF16;
      else if (p == "fp8") cfg.precision = ANE_PRECISION_FP8;
      else return usage();
    } else if (a == "--seed" && need(1)) {
      seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (a == "--full") {
      full = true;
    } else if (a == "--camera" && need(3)) {
      for (int k = 0; k < 3; ++k) cam[k] = std::strtof(argv[++i], nullptr);
    } else if (a == "--velocity" && need(3)) {
      for (int k = 0; k < 3; ++k) vel[k] = std::strtof(argv[++i], nullptr);
    } else {
      return usage();
    }
  }

  SceneFile sf;
  if (!load_scene(scene_path, sf)) return 1;
  // The renderer holds GPUMeshInfo records (GpuBuffers.h:38-44, filled at
  // VulkanRenderer.cpp:3486-3493, _blasRef = the BLAS device address) and the
  // index counts on the host (InternalMesh::_numIndices): the scene is created
  // from those, as the reference would hand them over.
  std::vector<ane_gpu_mesh_info> mesh_info(sf.meshes.size());
  std::vector<uint32_t> index_counts(sf.meshes.size());
  for (size_t m = 0; m < sf.meshes.size(); ++m) {
    std::memcpy(mesh_info[m].min_pos, sf.meshes[m].min_pos, sizeof(mesh_info[m].min_pos));
    std::memcpy(mesh_info[m].max_pos, sf.meshes[m].max_pos, sizeof(mesh_info[m].max_pos));
    mesh_info[m].vertex_offset = sf.meshes[m].vertex_offset;
    mesh_info[m].index_offset = sf.meshes[m].index_offset;
    mesh_info[m].blas_ref = 0x7f3a00000000ull + 0x10000ull * m;
    index_counts[m] = sf.meshes[m].index_count;
  }
  ane_scene* scene = nullptr;
  int st = ane_scene_create_mesh_info(sf.vertices.data(), (int64_t)(sf.vertices.size() / 24),
                                      sf.indices.data(), (int64_t)sf.indices.size(), mesh_info.data(),
                                      index_counts.data(), (int32_t)sf.meshes.size(), sf.materials.data(),
                                      (int32_t)sf.materials.size(), sf.instances.data(),
                                      (int32_t)sf.instances.size(), nullptr, &scene);
  if (st != ANE_OK) {
    std::fprintf(stderr, "ane_scene_create_mesh_info: %s (%s)\n", ane_status_string(st), ane_last_error());
    return 1;
  }
  ane_ddgi_ctx* grid = nullptr;
  st = ane_ddgi_create(&cfg, &grid);
  if (st != ANE_OK) {
    std::fprintf(stderr, "ane_ddgi_create: %s (%s)\n", ane_status_string(st), ane_last_error());
    ane_scene_destroy(scene);
    return 1;
  }
  ane_scene_info info{};
  ane_scene_info_get(scene, &info);
  std::printf("[scene] %lld triangles, %d instances, bvh depth %d, build %.3f ms\n",
              (long long)info.num_triangles, info.num_instances, info.bvh_depth, info.build_ms);

  ddgi::RenderContext rc(grid, cfg, scene);
  rc.setCameraPosition(cam[0], cam[1], cam[2]);
  rc.setLightDir(-0.3f, -0.8f, -0.5f);
  ddgi::FrameGraphBuilder fgb;
  ddgi::IrradianceProbeTranslationPass trans;
  ddgi::IrradianceProbeRayTracingPass rt(seed, full);
  ddgi::IrradianceProbeConvolvePass conv;
  ddgi::RenderPass* passes[] = {&trans, &rt, &conv};
  for (ddgi::RenderPass* p : passes) {
    p->init(&rc);
    p->registerToGraph(fgb, &rc);
  }

  std::vector<ane_ddgi_frame> log;
  int rc_status = 0;
  for (int fi = 0; fi < frames; ++fi) {
    fgb.executeGraph(&rc);
    if (rc.lastStatus != ANE_OK) {
      rc_status = 1;
      break;
    }
    ane_stream_synchronize(rc.stream());
    ane_ddgi_stats s{};
    ane_ddgi_stats_get(grid, &s);
    log.push_back(rc.lastTracedFrame);
    std::printf("[frame %d] layer %d cam (%.3f %.3f %.3f) trace %.3f shade %.3f conv %.3f trans %.3f ms\n",
                fi, rc.lastTracedFrame.layer_begin, rc.getCameraPosition()[0], rc.getCameraPosition()[1],
                rc.getCameraPosition()[2], s.last_ms[1], s.last_ms[3], s.last_ms[4], s.last_ms[5]);
    const float* c = rc.getCameraPosition();
    rc.setCameraPosition(c[0] + vel[0], c[1] + vel[1], c[2] + vel[2]);
  }

  const size_t atlas_floats = (size_t)100 * cfg.probes_x * cfg.probes_y * cfg.probes_z * 4;
  std::vector<float> atlas(atlas_floats);
  if (rc_status == 0) {
    st = ane_ddgi_atlas_get(grid, atlas.data(), nullptr);
    if (st != ANE_OK) {
      std::fprintf(stderr, "ane_ddgi_atlas_get: %s (%s)\n", ane_status_string(st), ane_last_error());
      rc_status = 1;
    }
  }
  if (rc_status == 0 &&
      !(write_file(outdir + "/frames.bin", log.data(), log.size() * sizeof(ane_ddgi_frame)) &&
        write_file(outdir + "/atlas.f32", atlas.data(), atlas.size() * sizeof(float))))
    rc_status = 1;
  ane_ddgi_destroy(grid);
  ane_scene_destroy(scene);
  return rc_status;
}
