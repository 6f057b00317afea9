// serve.hip -- persistent request loop for probe updates of many resident scenes.
//
// No counterpart in the reference (it updates one scene per frame from its
// render loop, IrradianceProbeRayTracingPass.cpp:95-165).  A session is a
// resident scene (its LBVH) plus a probe-grid context (its atlas and buffers);
// sessions are admitted against an HBM budget.  Requests (one probe update of
// one session with its frame constants) are queued from any host thread; a
// dispatcher thread admits them continuously: whenever a slot is free and the
// request's session has no update on the device, the request is launched on
// the least-loaded of the server's HIP streams, so work of different sessions
// overlaps on the device while each session's updates stay in submission order.
// Completion is observed through one event per request.  A session's results
// depend only on its own sequence of frames, so a request gives the same atlas
// alone and in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ane_serve.h"
#include "kernels.hpp"

namespace ane {
int fail(int status, const char* fmt, ...);
int ddgi_update_front(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, hipStream_t st,
                      ShadeJob* sj, ConvJob* cj, int* precision);
bool ddgi_fusable(const ane_ddgi_ctx* c);
int ddgi_update_graph(ane_ddgi_ctx* c, const ane_scene* s, const ane_ddgi_frame* f, hipStream_t st);
int ddgi_prepare(ane_ddgi_ctx* c, const ane_scene* s, int64_t* bytes);
int64_t ddgi_device_bytes(const ane_ddgi_ctx* c);
}
using ane::fail;

namespace {

double now_ms()
{
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Session {
  const ane_scene* scene = nullptr;
  ane_ddgi_ctx* ctx = nullptr;
  int64_t bytes = 0;        // the context's device bytes (the scene is charged once, per scene)
  int64_t rays = 0;         // primary rays of one full update
  int32_t probes_y = 0;
  bool busy = false;        // an update of this session is on the device
  uint64_t submitted = 0, finished = 0;
};

struct Request {
  int32_t session = -1;
  ane_ddgi_frame frame{};
  double t_submit = 0.0, t_launch = 0.0, t_done = 0.0;
  int status = ANE_OK;
  int stream = -1;
  hipEvent_t done = nullptr;
  bool launched = false, finished = false;
  int64_t rays = 0;
  int slot = -1;  // fused stages: the job-table slot of the request's batch
};

}  // namespace

struct ane_serve {
  ane_serve_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int32_t, Session> sessions;
  struct SceneUse {
    int refs = 0;
    int64_t bytes = 0;
  };
  std::unordered_map<const ane_scene*, SceneUse> scenes;
  std::unordered_map<uint64_t, Request> reqs;
  std::deque<uint64_t> queue;        // submitted, not launched (FIFO)
  std::vector<uint64_t> inflight;    // launched, not finished
  std::vector<hipStream_t> streams;
  std::vector<int> stream_load;
  std::vector<hipEvent_t> event_pool;
  int64_t used_bytes = 0;
  int32_t next_session = 0;
  uint64_t next_ticket = 1;
  bool stop = false;
  int device = 0;
  std::thread worker;
  // counters
  uint64_t completed = 0, failed = 0;
  int64_t rays_done = 0;
  int32_t peak_inflight = 0;
  // fused stages: job tables of the batches on the device, one slot per batch
  // (pinned host staging + device copy), reused once every request of its batch
  // has finished
  int slot_jobs = 0;                 // jobs per slot = the largest batch
  std::vector<int> slot_refs;        // unfinished requests per slot
  ane::ShadeJob* h_shade = nullptr;  // [slots][slot_jobs]
  ane::ConvJob* h_conv = nullptr;
  ane::ShadeJob* d_shade = nullptr;
  ane::ConvJob* d_conv = nullptr;
  std::vector<hipEvent_t> front_ev;  // fused stages: end of each request's front on its stream

  void run();
  int retire_locked(double t);
  struct Launch {
    uint64_t ticket;
    ane_ddgi_ctx* ctx;
    const ane_scene* scene;
    ane_ddgi_frame frame;
    hipStream_t stream;
    hipEvent_t done;
  };
  void launch_fused(std::vector<Launch>& batch, int slot, std::vector<std::pair<uint64_t, int>>& launched);
};

// One batch: each update's front (setup, sun grid, trace) on the server's streams in
// turn, starting with the batch's stream, so that the traces overlap; the batch's
// stream waits for them, then shades the probes of every update in one launch and
// convolves them in one launch per precision, from job tables copied to the
// batch's slot.
void ane_serve::launch_fused(std::vector<Launch>& batch, int slot,
                             std::vector<std::pair<uint64_t, int>>& launched)
{
  hipStream_t st = batch[0].stream;
  const int ns = (int)streams.size();
  int si = 0;
  while (si < ns && streams[si] != st) ++si;
  while (front_ev.size() < batch.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) break;
    front_ev.push_back(e);
  }
  const bool fork = ns > 1 && front_ev.size() >= batch.size();
  ane::ShadeJob* hs = h_shade + (size_t)slot * slot_jobs;
  ane::ConvJob* hc = h_conv + (size_t)slot * slot_jobs;
  ane::ShadeJob* ds = d_shade + (size_t)slot * slot_jobs;
  ane::ConvJob* dc = d_conv + (size_t)slot * slot_jobs;
  std::vector<int> status(batch.size(), ANE_OK), prec(batch.size(), -1);
  std::vector<ane::ShadeJob> sj(batch.size());
  std::vector<ane::ConvJob> cj(batch.size());
  for (size_t i = 0; i < batch.size(); ++i) {
    Launch& b = batch[i];
    hipStream_t fs = fork ? streams[(si + (int)i) % ns] : st;
    if (!ane::ddgi_fusable(b.ctx)) {
      status[i] = ane_ddgi_update(b.ctx, b.scene, &b.frame, fs);  // launched on its own
    } else {
      status[i] = ane::ddgi_update_front(b.ctx, b.scene, &b.frame, fs, &sj[i], &cj[i], &prec[i]);
      if (status[i] != ANE_OK) prec[i] = -1;
    }
    // the batch's stream (the fused stages, every request's completion event) waits for the front
    if (fs != st && (hipEventRecord(front_ev[i], fs) != hipSuccess || hipStreamWaitEvent(st, front_ev[i], 0) != hipSuccess)) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(fs);
    }
  }
  // job tables grouped by precision, each group contiguous in the slot; within a
  // group the convolutions that take the streaming kernel come first (one launch
  // each kind, so that every update runs the kernel its own launch would)
  struct Group {
    int first = 0, count = 0, shade_blocks = 0, max_rays = 0;
    int conv_count[2] = {0, 0}, conv_blocks[2] = {0, 0}, conv_max_rays[2] = {0, 0}, conv_gcd[2] = {0, 0};
  } grp[4];
  int k = 0;
  for (int p = 0; p < 4; ++p) {
    grp[p].first = k;
    for (int kind = 0; kind < 2; ++kind) {
      for (size_t i = 0; i < batch.size(); ++i) {
        if (prec[i] != p) continue;
        const int rays = cj[i].g.rays;
        if ((ane::conv_stream_fits(p, rays) ? 0 : 1) != kind) continue;
        sj[i].block_begin = grp[p].shade_blocks;
        cj[i].block_begin = grp[p].conv_blocks[kind];
        grp[p].shade_blocks += cj[i].nprobes;
        grp[p].conv_blocks[kind] += ane::conv_job_blocks(p, rays, cj[i].nprobes);
        grp[p].max_rays = std::max(grp[p].max_rays, (int)sj[i].g.rays);
        grp[p].conv_max_rays[kind] = std::max(grp[p].conv_max_rays[kind], rays);
        grp[p].conv_gcd[kind] = std::gcd(grp[p].conv_gcd[kind], rays);
        hs[k] = sj[i];
        hc[k] = cj[i];
        ++k;
        ++grp[p].count;
        ++grp[p].conv_count[kind];
      }
    }
  }
  hipError_t e = hipSuccess;
  if (k > 0) {
    if ((e = hipMemcpyAsync(ds, hs, sizeof(ane::ShadeJob) * k, hipMemcpyHostToDevice, st)) == hipSuccess)
      e = hipMemcpyAsync(dc, hc, sizeof(ane::ConvJob) * k, hipMemcpyHostToDevice, st);
    for (int p = 0; p < 4 && e == hipSuccess; ++p)
      ane::launch_shade_jobs(p, ds + grp[p].first, grp[p].count, grp[p].shade_blocks, grp[p].max_rays, st);
    for (int p = 0; p < 4 && e == hipSuccess; ++p) {
      ane::launch_conv_jobs(p, dc + grp[p].first, grp[p].conv_count[0], grp[p].conv_blocks[0], true,
                            grp[p].conv_max_rays[0], grp[p].conv_gcd[0], st);
      ane::launch_conv_jobs(p, dc + grp[p].first + grp[p].conv_count[0], grp[p].conv_count[1],
                            grp[p].conv_blocks[1], false, grp[p].conv_max_rays[1], grp[p].conv_gcd[1], st);
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  for (size_t i = 0; i < batch.size(); ++i) {
    int sti = status[i];
    if (sti == ANE_OK && prec[i] >= 0 && e != hipSuccess) sti = ANE_ERR_DEVICE;
    if (sti == ANE_OK && hipEventRecord(batch[i].done, st) != hipSuccess) sti = ANE_ERR_DEVICE;
    launched.push_back({batch[i].ticket, sti});
  }
  (void)slot;
}

int ane_serve::retire_locked(double t)
{
  int n = 0;
  for (size_t i = 0; i < inflight.size();) {
    Request& r = reqs[inflight[i]];
    const hipError_t q = hipEventQuery(r.done);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    r.finished = true;
    r.t_done = t;
    if (q != hipSuccess) {
      r.status = ANE_ERR_DEVICE;
      ++failed;
    } else {
      ++completed;
      rays_done += r.rays;
    }
    event_pool.push_back(r.done);
    r.done = nullptr;
    stream_load[r.stream] -= 1;
    if (r.slot >= 0) slot_refs[r.slot] -= 1;
    Session& s = sessions[r.session];
    s.busy = false;
    s.finished += 1;
    inflight[i] = inflight.back();
    inflight.pop_back();
    ++n;
  }
  return n;
}

void ane_serve::run()
{
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  std::vector<Launch> batch;
  for (;;) {
    if (retire_locked(now_ms()) > 0) cv_done.notify_all();
    // admit: oldest first, one update per session on the device at a time
    batch.clear();
    int slots = cfg.max_inflight - (int)inflight.size();
    // fused stages: a batch needs a free job-table slot and holds at most slot_jobs
    // requests (capping a batch at half the in-flight requests, so that two batches
    // overlap, measured 8% lower burst throughput)
    int slot = -1;
    if (cfg.fuse_stages) {
      for (int k = 0; k < (int)slot_refs.size(); ++k)
        if (slot_refs[k] == 0) { slot = k; break; }
      if (slot < 0) slots = 0;
      slots = std::min(slots, slot_jobs);
    }
    int batch_stream = 0;  // fused stages: the whole batch on the least-loaded stream
    for (int k = 1; k < (int)streams.size(); ++k)
      if (stream_load[k] < stream_load[batch_stream]) batch_stream = k;
    if (slots > 0 && !queue.empty()) {
      std::vector<int32_t> taken;
      for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < slots;) {
        Request& r = reqs[*it];
        Session& s = sessions[r.session];
        if (s.busy || std::find(taken.begin(), taken.end(), r.session) != taken.end()) {
          ++it;
          continue;
        }
        int best = batch_stream;
        if (!cfg.fuse_stages)
          for (int k = 0; k < (int)streams.size(); ++k)
            if (stream_load[k] < stream_load[best]) best = k;
        if (event_pool.empty()) {
          hipEvent_t e = nullptr;
          if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) break;
          event_pool.push_back(e);
        }
        r.done = event_pool.back();
        event_pool.pop_back();
        r.stream = best;
        r.slot = cfg.fuse_stages ? slot : -1;
        if (r.slot >= 0) slot_refs[r.slot] += 1;
        stream_load[best] += 1;
        s.busy = true;
        taken.push_back(r.session);
        batch.push_back({*it, s.ctx, s.scene, r.frame, streams[best], r.done});
        it = queue.erase(it);
      }
    }
    if (!batch.empty()) {
      // launch outside the lock: submitters keep queueing meanwhile
      lk.unlock();
      std::vector<std::pair<uint64_t, int>> launched;
      launched.reserve(batch.size());
      if (cfg.fuse_stages) {
        launch_fused(batch, slot, launched);
      } else {
        for (Launch& b : batch) {
          int st = cfg.capture_graphs ? ane::ddgi_update_graph(b.ctx, b.scene, &b.frame, b.stream)
                                      : ane_ddgi_update(b.ctx, b.scene, &b.frame, b.stream);
          if (st == ANE_OK && hipEventRecord(b.done, b.stream) != hipSuccess) st = ANE_ERR_DEVICE;
          launched.push_back({b.ticket, st});
        }
      }
      lk.lock();
      const double t = now_ms();
      for (auto& l : launched) {
        Request& r = reqs[l.first];
        r.launched = true;
        r.t_launch = t;
        r.status = l.second;
        if (l.second == ANE_OK) {
          inflight.push_back(l.first);
        } else {
          // failed before reaching the device: finish now, free the slot
          r.finished = true;
          r.t_done = t;
          ++failed;
          event_pool.push_back(r.done);
          r.done = nullptr;
          stream_load[r.stream] -= 1;
          if (r.slot >= 0) slot_refs[r.slot] -= 1;
          Session& s = sessions[r.session];
          s.busy = false;
          s.finished += 1;
        }
      }
      peak_inflight = std::max(peak_inflight, (int32_t)inflight.size());
      cv_done.notify_all();
      continue;
    }
    if (stop && queue.empty() && inflight.empty()) break;
    if (inflight.empty()) {
      cv_work.wait(lk);  // nothing on the device: sleep until a submission or close
    } else {
      // completions are polled; a submission wakes the loop at once
      cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
  cv_done.notify_all();
}

extern "C" int ane_serve_open(const ane_serve_config* cfg, ane_serve** out)
{
  if (!cfg || !out) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (cfg->hbm_budget_bytes <= 0) return fail(ANE_ERR_INVALID_ARG, "hbm_budget_bytes must be positive");
  if (cfg->max_inflight < 1 || cfg->max_inflight > 256)
    return fail(ANE_ERR_INVALID_ARG, "max_inflight must be in 1..256");
  if (cfg->capture_graphs != 0 && cfg->capture_graphs != 1)
    return fail(ANE_ERR_INVALID_ARG, "capture_graphs must be 0 or 1");
  if (cfg->fuse_stages != 0 && cfg->fuse_stages != 1)
    return fail(ANE_ERR_INVALID_ARG, "fuse_stages must be 0 or 1");
  if (cfg->num_streams < 1 || cfg->num_streams > 16)
    return fail(ANE_ERR_INVALID_ARG, "num_streams must be in 1..16");
  ane_serve* s = new ane_serve();
  s->cfg = *cfg;
  (void)hipGetDevice(&s->device);
  if (cfg->fuse_stages) {
    s->slot_jobs = std::min(cfg->max_inflight, 64);
    const size_t n = (size_t)cfg->max_inflight * s->slot_jobs;
    if (hipHostMalloc((void**)&s->h_shade, sizeof(ane::ShadeJob) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->h_conv, sizeof(ane::ConvJob) * n, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&s->d_shade, sizeof(ane::ShadeJob) * n) != hipSuccess ||
        hipMalloc((void**)&s->d_conv, sizeof(ane::ConvJob) * n) != hipSuccess) {
      (void)hipGetLastError();
      if (s->h_shade) (void)hipHostFree(s->h_shade);
      if (s->h_conv) (void)hipHostFree(s->h_conv);
      if (s->d_shade) (void)hipFree(s->d_shade);
      if (s->d_conv) (void)hipFree(s->d_conv);
      delete s;
      return fail(ANE_ERR_OUT_OF_MEMORY, "job tables of the fused stages");
    }
    s->slot_refs.assign(cfg->max_inflight, 0);
  }
  for (int k = 0; k < cfg->num_streams; ++k) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
      delete s;
      return fail(ANE_ERR_DEVICE, "hipStreamCreate failed");
    }
    s->streams.push_back(st);
    s->stream_load.push_back(0);
  }
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return ANE_OK;
}

extern "C" int ane_serve_add_session(ane_serve* s, const ane_scene* scene,
                                     const ane_ddgi_config* config, int32_t* session)
{
  if (!s || !scene || !config || !session) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  *session = -1;
  ane_scene_info si{};
  int st = ane_scene_info_get(scene, &si);
  if (st != ANE_OK) return st;
  ane_ddgi_ctx* ctx = nullptr;
  {
    // refuse before allocating when even the scene alone does not fit
    std::lock_guard<std::mutex> g(s->mu);
    const int64_t scene_bytes = s->scenes.count(scene) ? 0 : si.device_bytes;
    if (s->used_bytes + scene_bytes > s->cfg.hbm_budget_bytes)
      return fail(ANE_ERR_ADMISSION_REFUSED, "HBM budget: %lld of %lld bytes in use, scene needs %lld",
                  (long long)s->used_bytes, (long long)s->cfg.hbm_budget_bytes, (long long)scene_bytes);
  }
  st = ane_ddgi_create(config, &ctx);
  if (st != ANE_OK) return st;
  // the spill and sun-grid buffers the first update would allocate are allocated
  // now and charged with the rest (the budget covers every byte the session holds)
  int64_t ctx_bytes = 0;
  st = ane::ddgi_prepare(ctx, scene, &ctx_bytes);
  if (st != ANE_OK) {
    ane_ddgi_destroy(ctx);
    return st;
  }
  std::lock_guard<std::mutex> g(s->mu);
  const int64_t scene_bytes = s->scenes.count(scene) ? 0 : si.device_bytes;
  const int64_t need = ctx_bytes + scene_bytes;
  if (s->used_bytes + need > s->cfg.hbm_budget_bytes) {
    ane_ddgi_destroy(ctx);
    return fail(ANE_ERR_ADMISSION_REFUSED, "HBM budget: %lld of %lld bytes in use, session needs %lld",
                (long long)s->used_bytes, (long long)s->cfg.hbm_budget_bytes, (long long)need);
  }
  Session ses;
  ses.scene = scene;
  ses.ctx = ctx;
  ses.bytes = ctx_bytes;
  ses.rays = (int64_t)config->probes_x * config->probes_y * config->probes_z * config->rays_per_probe;
  ses.probes_y = config->probes_y;
  const int32_t id = s->next_session++;
  s->sessions[id] = ses;
  ane_serve::SceneUse& su = s->scenes[scene];
  su.refs += 1;
  su.bytes = si.device_bytes;
  s->used_bytes += need;
  *session = id;
  return ANE_OK;
}

// Waits until every submitted request of the session has finished.
static void wait_session_idle(ane_serve* s, std::unique_lock<std::mutex>& lk, int32_t id)
{
  s->cv_done.wait(lk, [&] {
    auto it = s->sessions.find(id);
    return it == s->sessions.end() || it->second.finished == it->second.submitted;
  });
}

extern "C" int ane_serve_remove_session(ane_serve* s, int32_t session)
{
  if (!s) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->sessions.count(session)) return fail(ANE_ERR_INVALID_ARG, "unknown session %d", session);
  wait_session_idle(s, lk, session);
  const Session ses = s->sessions[session];
  s->sessions.erase(session);
  s->used_bytes -= ses.bytes;
  auto sc = s->scenes.find(ses.scene);
  if (--sc->second.refs == 0) {
    s->used_bytes -= sc->second.bytes;
    s->scenes.erase(sc);
  }
  lk.unlock();
  return ane_ddgi_destroy(ses.ctx);
}

extern "C" int ane_serve_submit(ane_serve* s, int32_t session, const ane_ddgi_frame* frame,
                                uint64_t* ticket)
{
  if (!s || !frame || !ticket) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) return fail(ANE_ERR_CANCELLED, "server is closing");
  auto it = s->sessions.find(session);
  if (it == s->sessions.end()) return fail(ANE_ERR_INVALID_ARG, "unknown session %d", session);
  if (frame->layer_begin < 0 || frame->layer_count < 1 ||
      frame->layer_begin + frame->layer_count > it->second.probes_y)
    return fail(ANE_ERR_INVALID_ARG, "layers [%d, +%d) outside the grid of session %d",
                frame->layer_begin, frame->layer_count, session);
  const uint64_t t = s->next_ticket++;
  Request r;
  r.session = session;
  r.frame = *frame;
  r.t_submit = now_ms();
  r.rays = it->second.rays / it->second.probes_y * frame->layer_count;
  s->reqs[t] = r;
  s->queue.push_back(t);
  it->second.submitted += 1;
  *ticket = t;
  s->cv_work.notify_one();
  return ANE_OK;
}

extern "C" int ane_serve_poll(ane_serve* s, uint64_t ticket, int32_t timeout_ms,
                              ane_serve_result* out)
{
  if (!s || !out) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] {
    auto it = s->reqs.find(ticket);
    return it == s->reqs.end() || it->second.finished;
  };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, ready);
  else if (timeout_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(ANE_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
  const Request& r = it->second;
  if (!r.finished) return ANE_ERR_NOT_READY;
  out->status = r.status;
  out->session = r.session;
  out->queue_ms = r.t_launch - r.t_submit;
  out->latency_ms = r.t_done - r.t_submit;
  out->rays = r.rays;
  s->reqs.erase(it);
  return ANE_OK;
}

extern "C" int ane_serve_session_atlas(ane_serve* s, int32_t session, float* host_rgba)
{
  if (!s || !host_rgba) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->sessions.find(session);
  if (it == s->sessions.end()) return fail(ANE_ERR_INVALID_ARG, "unknown session %d", session);
  wait_session_idle(s, lk, session);
  ane_ddgi_ctx* ctx = s->sessions[session].ctx;
  lk.unlock();
  return ane_ddgi_atlas_get(ctx, host_rgba, nullptr);
}

extern "C" int ane_serve_stats_get(ane_serve* s, ane_serve_stats* out)
{
  if (!s || !out) return fail(ANE_ERR_INVALID_ARG, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  out->sessions = (int32_t)s->sessions.size();
  out->queued = (int32_t)s->queue.size();
  out->inflight = (int32_t)s->inflight.size();
  out->peak_inflight = s->peak_inflight;
  out->completed = s->completed;
  out->failed = s->failed;
  out->rays_done = s->rays_done;
  out->used_bytes = s->used_bytes;
  out->budget_bytes = s->cfg.hbm_budget_bytes;
  // what the sessions hold now: equal to used_bytes unless an update allocated
  // beyond what admission charged
  int64_t held = 0;
  for (const auto& kv : s->sessions) held += ane::ddgi_device_bytes(kv.second.ctx);
  for (const auto& kv : s->scenes) {
    ane_scene_info si{};
    if (ane_scene_info_get(kv.first, &si) == ANE_OK) held += si.device_bytes;
  }
  out->held_bytes = held;
  return ANE_OK;
}

extern "C" int ane_serve_close(ane_serve* s)
{
  if (!s) return ANE_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();  // drains the queue and the device work
  for (auto& kv : s->sessions) ane_ddgi_destroy(kv.second.ctx);
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->front_ev) (void)hipEventDestroy(e);
  for (auto& kv : s->reqs)
    if (kv.second.done) (void)hipEventDestroy(kv.second.done);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  if (s->h_shade) (void)hipHostFree(s->h_shade);
  if (s->h_conv) (void)hipHostFree(s->h_conv);
  if (s->d_shade) (void)hipFree(s->d_shade);
  if (s->d_conv) (void)hipFree(s->d_conv);
  delete s;
  return ANE_OK;
}
