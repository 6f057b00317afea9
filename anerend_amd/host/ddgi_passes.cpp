// ddgi_passes.cpp -- the three probe passes over the C-ABI (see ddgi_passes.hpp).
#include "ddgi_passes.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace ddgi {

ane_ddgi_frame RenderContext::sceneFrame(int layerBegin, int layerCount, const float rot16[16]) const
{
  ane_ddgi_frame f{};
  for (int i = 0; i < 3; ++i) {
    f.camera_pos[i] = camera_[i];
    f.light_dir[i] = light_[i];
    f.sky_color[i] = options_.skyColor[i];
  }
  f.camera_pos[3] = 1.0f;
  f.sky_color[3] = 1.0f;
  f.sun_intensity = options_.sunIntensity;
  f.sky_intensity = options_.skyIntensity;
  uint32_t flags = 0;
  if (options_.ddgiEnabled) flags |= ANE_FLAG_DDGI;
  if (options_.multiBounceDdgiEnabled) flags |= ANE_FLAG_DDGI_MULTI;
  if (baking_) flags |= ANE_FLAG_BAKE;
  f.flags = flags;
  f.layer_begin = layerBegin;
  f.layer_count = layerCount;
  std::memcpy(f.rotation, rot16, sizeof(f.rotation));
  f.bake_tile[0] = bakeTile_[0];
  f.bake_tile[1] = bakeTile_[1];
  f.bake_tile[2] = bakeTile_[2];
  return f;
}

bool FrameGraphBuilder::registerRenderPass(const std::string& name)
{
  if (exes_.count(name)) return false;
  order_.push_back(name);
  exes_[name] = nullptr;
  return true;
}

bool FrameGraphBuilder::registerRenderPassExe(const std::string& name, Exe fn)
{
  auto it = exes_.find(name);
  if (it == exes_.end()) {
    std::printf("Could not find render pass %s when registering exe\n", name.c_str());  // FrameGraphBuilder.cpp:184-187
    return false;
  }
  it->second = std::move(fn);
  return true;
}

void FrameGraphBuilder::executeGraph(RenderContext* rc)
{
  for (const auto& name : order_) {
    const Exe& fn = exes_[name];
    if (fn) fn(rc);
  }
}

static void report(RenderContext* rc, const char* pass, int st)
{
  rc->lastStatus = st;
  if (st != ANE_OK) std::printf("%s: %s (%s)\n", pass, ane_status_string(st), ane_last_error());
}

void IrradianceProbeTranslationPass::registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc)
{
  (void)rc;
  fgb.registerRenderPass("IrradianceProbeTrans");
  fgb.registerRenderPassExe("IrradianceProbeTrans", [](RenderContext* rc) {
    if (!rc->getRenderOptions().ddgiEnabled) return;
    if (rc->isBaking()) return;
    int32_t shifted = 0;
    report(rc, "IrradianceProbeTrans",
           ane_ddgi_translate(rc->grid(), rc->getCameraPosition(), &shifted, rc->stream()));
  });
}

void IrradianceProbeRayTracingPass::randomRotation(float m[16])
{
  // axis first, then the angle (IrradianceProbeRayTracingPass.cpp:114-133)
  std::uniform_real_distribution<double> angle(0.0, 360.0);
  std::uniform_int_distribution<int> axis(0, 2);
  const int ax = axis(rng_);
  const float a = (float)angle(rng_) * 0.0174532925199432957692f;  // glm::radians
  const float c = std::cos(a), s = std::sin(a);
  float v[3] = {0.0f, 0.0f, 0.0f};
  v[ax] = 1.0f;
  float t[3] = {(1.0f - c) * v[0], (1.0f - c) * v[1], (1.0f - c) * v[2]};
  std::memset(m, 0, 16 * sizeof(float));
  m[0] = c + t[0] * v[0];        m[1] = t[0] * v[1] + s * v[2]; m[2] = t[0] * v[2] - s * v[1];
  m[4] = t[1] * v[0] - s * v[2]; m[5] = c + t[1] * v[1];        m[6] = t[1] * v[2] + s * v[0];
  m[8] = t[2] * v[0] + s * v[1]; m[9] = t[2] * v[1] - s * v[0]; m[10] = c + t[2] * v[2];
  m[15] = 1.0f;
}

void IrradianceProbeRayTracingPass::registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc)
{
  fgb.registerRenderPass("IrradianceProbeRT");
  const uint32_t ny = rc->getNumIrradianceProbesY();
  currentLayer_ = 0;
  fgb.registerRenderPassExe("IrradianceProbeRT", [this, ny](RenderContext* rc) {
    if (!rc->getRenderOptions().raytracingEnabled) return;
    if (!rc->getRenderOptions().ddgiEnabled) return;
    const uint32_t layer = currentLayer_;
    float rot[16];
    randomRotation(rot);
    const ane_ddgi_frame f = fullGrid_ ? rc->sceneFrame(0, (int)ny, rot) : rc->sceneFrame((int)layer, 1, rot);
    report(rc, "IrradianceProbeRT", ane_ddgi_trace(rc->grid(), rc->scene(), &f, rc->stream()));
    rc->lastTracedFrame = f;
    rc->setBlackboardValueInt("ProbeLayer", (int)layer);
    ++currentLayer_;
    if (currentLayer_ >= ny) currentLayer_ = 0;
  });
}

void IrradianceProbeConvolvePass::registerToGraph(FrameGraphBuilder& fgb, RenderContext* rc)
{
  (void)rc;
  fgb.registerRenderPass("IrradianceProbeConvolve");
  fgb.registerRenderPassExe("IrradianceProbeConvolve", [](RenderContext* rc) {
    if (!rc->getRenderOptions().raytracingEnabled) return;
    if (!rc->getRenderOptions().ddgiEnabled) return;
    ane_ddgi_frame f = rc->lastTracedFrame;
    // the layer comes from the blackboard, as the push constant does (IrradianceProbeConvolvePass.cpp:125)
    if (f.layer_count == 1) f.layer_begin = rc->blackboardValueInt("ProbeLayer");
    report(rc, "IrradianceProbeConvolve", ane_ddgi_convolve(rc->grid(), &f, rc->stream()));
  });
}

}  // namespace ddgi
