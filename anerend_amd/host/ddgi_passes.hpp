// ddgi_passes.hpp -- C++ host side of the probe update above the C-ABI.
//
// Mirrors the reference's plugin interface for this path: RenderPass with
// registerToGraph (passes/RenderPass.h:14-35), FrameGraphBuilder's named exe
// callbacks run in submission order (FrameGraphBuilder.h:150-185), the
// RenderContext queries the passes make (RenderContext.h:85-188) and the three
// passes with their state rules:
//   IrradianceProbeTranslationPass  camera-cell shift (IrradianceProbeTranslationPass.cpp:56-186)
//   IrradianceProbeRayTracingPass   layer round robin + random rotation (IrradianceProbeRayTracingPass.cpp:95-165)
//   IrradianceProbeConvolvePass     convolution of the traced layer (IrradianceProbeConvolvePass.cpp:109-143)
// Each exe calls libane_ddgi where the reference recorded vkCmdCopyImage /
// vkCmdTraceRaysKHR / vkCmdDispatch.  Errors follow the reference's behaviour
// (the exe logs and returns); the last status is kept for the caller.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../include/ane_ddgi.h"

namespace ddgi {

struct RenderOptions {  // RenderOptions.h:7-28 (the fields the probe passes read)
  bool raytracingEnabled = true;
  bool ddgiEnabled = true;
  bool multiBounceDdgiEnabled = true;
  float sunIntensity = 5.0f;
  float skyIntensity = 1.0f;
  float skyColor[3] = {0.2f, 0.5f, 0.7f};
};

class RenderContext {
 public:
  RenderContext(ane_ddgi_ctx* grid, const ane_ddgi_config& cfg, const ane_scene* scene)
      : grid_(grid), cfg_(cfg), scene_(scene) {}

  uint32_t getNumIrradianceProbesXZ() const { return (uint32_t)cfg_.probes_x; }
  uint32_t getNumIrradianceProbesY() const { return (uint32_t)cfg_.probes_y; }
  RenderOptions& getRenderOptions() { return options_; }
  const float* getCameraPosition() const { return camera_; }
  void setCameraPosition(float x, float y, float z) { camera_[0] = x; camera_[1] = y; camera_[2] = z; }
  void setLightDir(float x, float y, float z) { light_[0] = x; light_[1] = y; light_[2] = z; }
  bool isBaking() const { return baking_; }
  void setBaking(bool b, int tileX = 0, int tileZ = 0, int tileSize = 32)
  {
    baking_ = b;
    bakeTile_[0] = tileX; bakeTile_[1] = tileZ; bakeTile_[2] = tileSize;
  }
  void setBlackboardValueInt(const std::string& k, int v) { blackboard_[k] = v; }
  int blackboardValueInt(const std::string& k) const
  {
    auto it = blackboard_.find(k);
    return it == blackboard_.end() ? 0 : it->second;
  }

  // Scene UBO subset + push constants (VulkanRenderer.cpp:1201-1243)
  ane_ddgi_frame sceneFrame(int layerBegin, int layerCount, const float rot16[16]) const;

  ane_ddgi_ctx* grid() const { return grid_; }
  const ane_scene* scene() const { return scene_; }
  void* stream() const { return stream_; }
  void setStream(void* s) { stream_ = s; }
  ane_ddgi_frame lastTracedFrame{};
  int lastStatus = ANE_OK;

 private:
  ane_ddgi_ctx* grid_;
  ane_ddgi_config cfg_;
  const ane_scene* scene_;
  RenderOptions options_;
  float camera_[3] = {0.0f, 0.0f, 0.0f};
  float light_[3] = {0.0f, -1.0f, 0.0f};
  bool baking_ = false;
  int bakeTile_[3] = {0, 0, 32};
  std::map<std::string, int> blackboard_;
  void* stream_ = nullptr;
};

class FrameGraphBuilder {
 public:
  using Exe = std::function<void(RenderContext*)>;
  bool registerRenderPass(const std::string& name);
  bool registerRenderPassExe(const std::string& name, Exe fn);
  void executeGraph(RenderContext* rc);
  const std::vector<std::string>& order() const { return order_; }

 private:
  std::vector<std::string> order_;
  std::map<std::string, Exe> exes_;
};

class RenderPass {
 public:
  virtual ~RenderPass() = default;
  virtual bool init(RenderContext*) { return true; }
  virtual void registerToGraph(FrameGraphBuilder&, RenderContext* rc) = 0;
  virtual void cleanup(RenderContext*) {}
};

class IrradianceProbeTranslationPass final : public RenderPass {
 public:
  void registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc) override;
};

class IrradianceProbeRayTracingPass final : public RenderPass {
 public:
  explicit IrradianceProbeRayTracingPass(uint64_t seed = 0, bool fullGrid = false)
      : rng_(seed), fullGrid_(fullGrid) {}
  void registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc) override;
  // glm::rotate(mat4(1), radians(U[0,360)), axis in {x, y, z}), column-major
  void randomRotation(float out16[16]);

 private:
  uint32_t currentLayer_ = 0;
  std::mt19937_64 rng_;  // seeded in place of std::random_device (explicit input)
  bool fullGrid_;
};

class IrradianceProbeConvolvePass final : public RenderPass {
 public:
  void registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc) override;
};

}  // namespace ddgi
