// firex_server.hip — continuous batching of many resident Fire-X plumes on
// one MI355X (C ABI: include/firex_serve.h).
//
// Requests queue in arrival order.  A pump admits every queued request that
// fits the free slots and the HBM byte budget (first fit, so a large head
// request does not block smaller ones behind it), then issues steps for every
// resident plume, round-robin, up to a small in-flight limit per plume, each
// plume bound to one stream of a pool so that plumes overlap on the device.
// Completion is observed with HIP events, never by blocking; a plume that
// finishes its steps leaves (its memory is freed unless the caller keeps its
// result) and the next pump admits waiting requests into the freed space.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/firex_serve.h"
#include "firex_host.h"
#include "firex_solver_internal.h"

using namespace fx;

namespace {

using Clock = std::chrono::steady_clock;

struct Inflight {
  hipEvent_t stop;
  double issued;  // server clock, seconds
};

struct Request {
  int64_t id = 0;
  firex_request req;
  int state = FIREX_REQUEST_QUEUED;
  int32_t steps_issued = 0, steps_done = 0;
  double submit = 0, admit = 0, finish = 0;
  firex_solver* solver = nullptr;
  uint64_t bytes = 0;
  uint32_t cells = 0;
  bool batched = false;  // batch mode: a small plume on a batch stream
  std::deque<Inflight> inflight;
  std::vector<float> latency_ms;
};

double percentile(std::vector<float> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  double pos = q * (double)(v.size() - 1);
  size_t i = (size_t)pos;
  double f = pos - (double)i;
  if (i + 1 >= v.size()) return v.back();
  return v[i] * (1 - f) + v[i + 1] * f;
}

}  // namespace

struct firex_server {
  firex_server_config cfg;
  std::vector<hipStream_t> streams;
  hipEvent_t epoch = nullptr;   // device-clock anchor
  double epoch_seconds = 0;     // server clock at the anchor
  Clock::time_point t0;
  std::map<int64_t, Request> requests;
  std::deque<int64_t> queue;
  std::vector<int64_t> active;
  uint64_t resident = 0;
  int64_t next_id = 1;
  int next_stream = 0, next_batch_stream = 0;
  std::vector<hipEvent_t> pool;
  std::vector<float> latency_ms;
  int64_t steps_completed = 0;
  double cell_updates = 0;
  int32_t done = 0, failed = 0;
  int64_t grouped_steps = 0, group_steps = 0;
};

namespace {

double now(const firex_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

hipEvent_t take_event(firex_server* s) {
  if (!s->pool.empty()) {
    hipEvent_t e = s->pool.back();
    s->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Server-clock time at which a completed event fired.
double event_time(firex_server* s, hipEvent_t e) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, s->epoch, e) != hipSuccess) return now(s);
  return s->epoch_seconds + (double)ms * 1e-3;
}

uint64_t batch_max_cells(const firex_server* s) {
  return s->cfg.batch_max_cells > 0 ? (uint64_t)s->cfg.batch_max_cells : (uint64_t)4 << 20;
}

int batch_streams(const firex_server* s) {
  const int n = (int)s->streams.size();
  const int b = s->cfg.batch_streams > 0 ? s->cfg.batch_streams : n / 2;
  return b < 1 ? 1 : b > n ? n : b;
}

bool batched_request(const firex_server* s, const firex_request& r) {
  return s->cfg.batch_plumes && (uint64_t)r.dims[0] * (uint64_t)r.dims[1] * (uint64_t)r.dims[2] <= batch_max_cells(s);
}

firex_config plume_config(const firex_server* s, const firex_request& r) {
  firex_config c;
  firex_config_init(&c);
  firex_config_from_tier(&c, "uhd");  // offline tier: cinematic flag, dt 1/60
  memcpy(c.dims, r.dims, sizeof c.dims);
  c.pressure_iterations = r.pressure_iterations;
  c.correction_iterations = r.correction_iterations;
  c.species_storage = r.species_storage;
  c.heavy_mode = FIREX_HEAVY_OFF;
  c.plan = s->cfg.plan;
  c.device = s->cfg.device;
  c.capture_graphs = s->cfg.capture_graphs;
  c.jacobi_autotune = s->cfg.jacobi_autotune;
  if (batched_request(s, r)) {
    // one block shape and depth for every plume, so that plumes can share
    // launches: the 16-wave ring shape at depth 4 where the grid allows it,
    // with 8-plane z chunks.  A plume alone on a small grid runs fastest on
    // chunks of 2-3 planes (more blocks for the GPU), but in a group the
    // other plumes supply the blocks and short chunks only add halo planes:
    // on the small-grid stream, groups served 3.10 G cell-updates/s with
    // 8-plane chunks, 2.94 G with 16 and 2.07 G with the single-plume plans.
    c.jacobi_sweeps_per_launch = 4;
    c.jacobi_ring_waves = 16;
    c.jacobi_chunk_z = 8;
  }
  return c;
}

void retire(firex_server* s, Request& r) {
  if (r.solver && !r.req.keep_result) {
    firex_destroy(r.solver);
    r.solver = nullptr;
    s->resident -= r.bytes;
  }
}

// Folds finished steps into the statistics; returns the number harvested.
int harvest(firex_server* s) {
  int n = 0;
  for (size_t i = 0; i < s->active.size();) {
    Request& r = s->requests[s->active[i]];
    while (!r.inflight.empty() && hipEventQuery(r.inflight.front().stop) == hipSuccess) {
      const Inflight f = r.inflight.front();
      r.inflight.pop_front();
      double t = event_time(s, f.stop);
      float lat = (float)((t - f.issued) * 1e3);
      if (lat < 0) lat = 0;
      r.latency_ms.push_back(lat);
      s->latency_ms.push_back(lat);
      s->pool.push_back(f.stop);
      r.steps_done += 1;
      s->steps_completed += 1;
      s->cell_updates += (double)r.cells;
      r.finish = t;
      ++n;
    }
    if (r.steps_done >= r.req.steps && r.inflight.empty()) {
      r.state = FIREX_REQUEST_DONE;
      s->done += 1;
      retire(s, r);
      s->active.erase(s->active.begin() + (long)i);
    } else {
      ++i;
    }
  }
  return n;
}

int admit(firex_server* s) {
  int admitted = 0;
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    if ((int)s->active.size() >= s->cfg.max_active) break;
    Request& r = s->requests[*it];
    if (s->resident + r.bytes > s->cfg.memory_budget_bytes) {
      ++it;  // first fit: a smaller request behind may still fit
      continue;
    }
    firex_config c = plume_config(s, r.req);
    // batch mode: small plumes on the first batch_streams streams, the
    // others on the rest (all of them when the batch streams are all)
    r.batched = batched_request(s, r.req);
    const int nb = s->cfg.batch_plumes ? batch_streams(s) : 0, ns = (int)s->streams.size();
    const int slot = r.batched ? s->next_batch_stream++ % nb
                   : nb < ns   ? nb + s->next_stream++ % (ns - nb)
                               : s->next_stream++ % ns;
    hipStream_t stream = s->streams[(size_t)slot];
    int rc = solver_create(&c, stream, &r.solver);
    if (rc == FIREX_OK && r.req.emitter_count > 0)
      rc = firex_set_emitters(r.solver, r.req.emitters, r.req.emitter_count, r.req.jitter, r.req.velocity_noise,
                              r.req.seed);
    if (rc != FIREX_OK) {
      if (r.solver) firex_destroy(r.solver);
      r.solver = nullptr;
      if (rc == FIREX_ERR_OUT_OF_MEMORY && !s->active.empty()) {
        ++it;  // retry once memory frees up
        continue;
      }
      r.state = FIREX_REQUEST_FAILED;
      s->failed += 1;
      it = s->queue.erase(it);
      continue;
    }
    r.state = FIREX_REQUEST_ACTIVE;
    r.admit = now(s);
    s->resident += r.bytes;
    s->active.push_back(r.id);
    it = s->queue.erase(it);
    ++admitted;
  }
  return admitted;
}

// Records the completion event of a step just issued for r.
int note_issued(firex_server* s, Request& r, double t) {
  hipEvent_t stop = take_event(s);
  if (!stop) return set_error(FIREX_ERR_DEVICE, "hipEventCreate failed");
  if (hipEventRecord(stop, solver_stream(r.solver)) != hipSuccess) {
    s->pool.push_back(stop);
    return set_error(FIREX_ERR_DEVICE, "hipEventRecord failed for a step's completion event");
  }
  r.inflight.push_back({stop, t});
  r.steps_issued += 1;
  return FIREX_OK;
}

// Batch mode: small plumes with one group key and one dt step in rounds, every
// plume of the group with steps left joining each round (in groups of up to
// kStepGroupMax), and a round is issued while the group's most loaded plume
// is below the in-flight limit.  Issuing per plume instead would let the
// plumes drift out of phase (each refilling as its own step completes) and
// split the group into ever smaller launches.  Large plumes and those
// without a key step alone.
int issue_batched(firex_server* s, size_t limit) {
  int issued = 0;
  for (;;) {
    std::vector<std::pair<uint64_t, Request*>> left;
    for (int64_t id : s->active) {
      Request& r = s->requests[id];
      if (r.steps_issued < r.req.steps) left.push_back({r.batched ? solver_group_key(r.solver) : 0, &r});
    }
    // stable: plumes keep their admission order within a group
    std::stable_sort(left.begin(), left.end(), [](const auto& a, const auto& b) {
      return a.first != b.first ? a.first < b.first : a.second->req.dt < b.second->req.dt;
    });
    bool any = false;
    for (size_t i = 0; i < left.size();) {
      size_t j = i + 1;
      if (left[i].first != 0)
        while (j < left.size() && left[j].first == left[i].first && left[j].second->req.dt == left[i].second->req.dt)
          ++j;
      size_t busiest = 0;
      for (size_t k = i; k < j; ++k) busiest = std::max(busiest, left[k].second->inflight.size());
      if (busiest >= limit) {
        i = j;
        continue;
      }
      for (size_t c = i; c < j; c += (size_t)kStepGroupMax) {
        const size_t e = std::min(j, c + (size_t)kStepGroupMax);
        firex_solver* group[kStepGroupMax];
        for (size_t k = c; k < e; ++k) group[k - c] = left[k].second->solver;
        const double t = now(s);
        const double dt = left[c].second->req.dt;
        const int rc = e - c == 1 ? solver_step_async(group[0], dt) : solver_step_group(group, (int32_t)(e - c), dt);
        if (rc != FIREX_OK) return -1;  // the solver's message stays in firex_last_error()
        if (e - c > 1) {
          s->group_steps += 1;
          s->grouped_steps += (int64_t)(e - c);
        }
        for (size_t k = c; k < e; ++k) {
          if (note_issued(s, *left[k].second, t) != FIREX_OK) return -1;
          ++issued;
        }
      }
      any = true;
      i = j;
    }
    if (!any) return issued;
  }
}

int issue(firex_server* s) {
  int issued = 0;
  // One step in flight per plume by default: on the mixed serving stream a
  // second one doubled the step latency (each step also queued behind its
  // plume's previous one) without raising throughput.
  const size_t limit = (size_t)(s->cfg.max_inflight_steps > 0 ? s->cfg.max_inflight_steps : 1);
  if (s->cfg.batch_plumes) return issue_batched(s, limit);
  for (int64_t id : s->active) {
    Request& r = s->requests[id];
    while (r.inflight.size() < limit && r.steps_issued < r.req.steps) {
      hipEvent_t stop = take_event(s);
      if (!stop) {
        set_error(FIREX_ERR_DEVICE, "hipEventCreate failed");
        return -1;
      }
      double t = now(s);
      int rc = solver_step_async(r.solver, r.req.dt);
      if (rc != FIREX_OK) {
        s->pool.push_back(stop);
        return -1;  // the solver's message stays in firex_last_error()
      }
      if (hipEventRecord(stop, solver_stream(r.solver)) != hipSuccess) {
        s->pool.push_back(stop);
        set_error(FIREX_ERR_DEVICE, "hipEventRecord failed for a step's completion event");
        return -1;
      }
      r.inflight.push_back({stop, t});
      r.steps_issued += 1;
      ++issued;
    }
  }
  return issued;
}

}  // namespace

extern "C" {

int firex_server_create(const firex_server_config* cfg, firex_server** out) {
  clear_error();
  if (!cfg || !out) return set_error(FIREX_ERR_INVALID, "config and out must be non-null");
  *out = nullptr;
  if (cfg->max_active < 1) return set_error(FIREX_ERR_INVALID, "max_active must be at least 1");
  if (cfg->streams < 1 || cfg->streams > 8) return set_error(FIREX_ERR_INVALID, "streams must be in [1, 8]");
  if (cfg->plan < 0 || cfg->plan > 1) return set_error(FIREX_ERR_INVALID, "unknown plan");
  if (cfg->batch_max_cells < 0 || cfg->batch_streams < 0 || cfg->batch_streams > cfg->streams)
    return set_error(FIREX_ERR_INVALID, "batch_max_cells must be non-negative and batch_streams in [0, streams]");
  if (firex_device_count() <= cfg->device || cfg->device < 0)
    return set_error(FIREX_ERR_DEVICE, "No HIP device is available for the Fire-X server.");
  if (hipSetDevice(cfg->device) != hipSuccess) return set_error(FIREX_ERR_DEVICE, "hipSetDevice failed");
  firex_server* s = new firex_server();
  s->cfg = *cfg;
  if (s->cfg.memory_budget_bytes == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    s->cfg.memory_budget_bytes = (uint64_t)free_b / 4 * 3;
  }
  for (int i = 0; i < cfg->streams; ++i) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
      delete s;
      return set_error(FIREX_ERR_DEVICE, "hipStreamCreate failed");
    }
    s->streams.push_back(st);
  }
  s->t0 = Clock::now();
  if (hipEventCreate(&s->epoch) != hipSuccess || hipEventRecord(s->epoch, s->streams[0]) != hipSuccess ||
      hipEventSynchronize(s->epoch) != hipSuccess) {
    if (s->epoch) (void)hipEventDestroy(s->epoch);
    for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
    delete s;
    return set_error(FIREX_ERR_DEVICE, "could not anchor the device clock");
  }
  s->epoch_seconds = now(s);
  *out = s;
  return FIREX_OK;
}

void firex_server_destroy(firex_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->cfg.device);  // teardown: best effort
  for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);  // teardown: best effort
  for (auto& kv : s->requests) {
    for (auto& f : kv.second.inflight) (void)hipEventDestroy(f.stop);
    if (kv.second.solver) firex_destroy(kv.second.solver);
  }
  for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
  if (s->epoch) (void)hipEventDestroy(s->epoch);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  delete s;
}

int firex_server_submit(firex_server* s, const firex_request* req, int64_t* id) {
  if (!s || !req || !id) return set_error(FIREX_ERR_INVALID, "server, request and id must be non-null");
  if (req->steps < 0 || !(req->dt >= 0)) return set_error(FIREX_ERR_INVALID, "steps and dt must be non-negative");
  if (req->emitter_count < 0 || req->emitter_count > 8) return set_error(FIREX_ERR_INVALID, "between 0 and 8 emitters");
  Request r;
  r.id = s->next_id++;
  r.req = *req;
  firex_config c = plume_config(s, *req);
  bool heavy, f16p, compact;
  int rc = solver_plan(&c, &heavy, &f16p, &compact, &r.bytes);
  if (rc) return rc;
  if (r.bytes > s->cfg.memory_budget_bytes)
    return set_error(FIREX_ERR_BUDGET, "The request needs " + std::to_string(r.bytes) +
                                           " bytes, above the server's " + std::to_string(s->cfg.memory_budget_bytes) +
                                           " byte budget.");
  r.cells = (uint32_t)((uint64_t)req->dims[0] * req->dims[1] * req->dims[2]);
  r.submit = now(s);
  *id = r.id;
  s->queue.push_back(r.id);
  s->requests.emplace(r.id, std::move(r));
  return FIREX_OK;
}

int firex_server_pump(firex_server* s, int32_t* busy) {
  if (!s) return set_error(FIREX_ERR_INVALID, "server is null");
  if (hipSetDevice(s->cfg.device) != hipSuccess) return set_error(FIREX_ERR_DEVICE, "hipSetDevice failed");
  harvest(s);
  admit(s);
  if (issue(s) < 0) return FIREX_ERR_DEVICE;
  if (busy) *busy = (!s->queue.empty() || !s->active.empty()) ? 1 : 0;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(FIREX_ERR_DEVICE, std::string("server step: ") + hipGetErrorString(e));
  return FIREX_OK;
}

int firex_server_drain(firex_server* s) {
  if (!s) return set_error(FIREX_ERR_INVALID, "server is null");
  if (hipSetDevice(s->cfg.device) != hipSuccess) return set_error(FIREX_ERR_DEVICE, "hipSetDevice failed");
  for (;;) {
    int h = harvest(s);
    int a = admit(s);
    int i = issue(s);
    if (i < 0) return FIREX_ERR_DEVICE;
    if (s->queue.empty() && s->active.empty()) break;
    if (h == 0 && a == 0 && i == 0) {
      if (s->active.empty()) {
        // nothing resident yet nothing admissible: the queue can never run
        return set_error(FIREX_ERR_BUDGET, "queued requests cannot be admitted");
      }
      // wait for the oldest step in flight instead of spinning
      for (int64_t id : s->active) {
        Request& r = s->requests[id];
        if (!r.inflight.empty()) {
          if (hipEventSynchronize(r.inflight.front().stop) != hipSuccess)
            return set_error(FIREX_ERR_DEVICE, "waiting for a step in flight failed");
          break;
        }
      }
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(FIREX_ERR_DEVICE, std::string("server drain: ") + hipGetErrorString(e));
  return FIREX_OK;
}

int firex_server_request_status(firex_server* s, int64_t id, firex_request_status* out) {
  if (!s || !out) return set_error(FIREX_ERR_INVALID, "server and out must be non-null");
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_error(FIREX_ERR_INVALID, "unknown request id");
  const Request& r = it->second;
  memset(out, 0, sizeof *out);
  out->state = r.state;
  out->steps_done = r.steps_done;
  out->submit_time = r.submit;
  out->admit_time = r.admit;
  out->finish_time = r.finish;
  out->latency_ms_p50 = percentile(r.latency_ms, 0.50);
  out->latency_ms_p99 = percentile(r.latency_ms, 0.99);
  out->latency_ms_max = r.latency_ms.empty() ? 0.0 : *std::max_element(r.latency_ms.begin(), r.latency_ms.end());
  out->device_bytes = r.bytes;
  return FIREX_OK;
}

int firex_server_get_stats(firex_server* s, firex_server_stats* out) {
  if (!s || !out) return set_error(FIREX_ERR_INVALID, "server and out must be non-null");
  memset(out, 0, sizeof *out);
  out->queued = (int32_t)s->queue.size();
  out->active = (int32_t)s->active.size();
  out->done = s->done;
  out->failed = s->failed;
  out->resident_bytes = s->resident;
  out->memory_budget_bytes = s->cfg.memory_budget_bytes;
  out->steps_completed = s->steps_completed;
  out->cell_updates = s->cell_updates;
  out->elapsed_seconds = now(s);
  out->latency_ms_p50 = percentile(s->latency_ms, 0.50);
  out->latency_ms_p99 = percentile(s->latency_ms, 0.99);
  out->latency_ms_max = s->latency_ms.empty() ? 0.0 : *std::max_element(s->latency_ms.begin(), s->latency_ms.end());
  out->grouped_steps = s->grouped_steps;
  out->group_steps = s->group_steps;
  return FIREX_OK;
}

int firex_server_read_field(firex_server* s, int64_t id, int32_t field, void* dst, size_t bytes, int32_t dtype) {
  if (!s) return set_error(FIREX_ERR_INVALID, "server is null");
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_error(FIREX_ERR_INVALID, "unknown request id");
  Request& r = it->second;
  if (r.state != FIREX_REQUEST_DONE || !r.solver)
    return set_error(FIREX_ERR_INVALID, "the request is not finished or its result was not kept");
  return firex_read_field(r.solver, field, dst, bytes, dtype);
}

int firex_server_release(firex_server* s, int64_t id) {
  if (!s) return set_error(FIREX_ERR_INVALID, "server is null");
  auto it = s->requests.find(id);
  if (it == s->requests.end()) return set_error(FIREX_ERR_INVALID, "unknown request id");
  Request& r = it->second;
  if (r.state == FIREX_REQUEST_QUEUED || r.state == FIREX_REQUEST_ACTIVE)
    return set_error(FIREX_ERR_INVALID, "the request is still queued or running");
  if (r.solver) {
    firex_destroy(r.solver);
    r.solver = nullptr;
    s->resident -= r.bytes;
  }
  if (r.state == FIREX_REQUEST_DONE) r.state = FIREX_REQUEST_RELEASED;
  return FIREX_OK;
}

double firex_server_clock(firex_server* s) { return s ? now(s) : 0.0; }

}  // extern "C"
