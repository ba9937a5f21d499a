// firex_solver.hip — the C ABI solver object: device buffers of one plume,
// the step graph of FireXEngine.encodeSimulationStep
// (src/sim/firex/FireXEngine.ts:2115-2276) and field transfer.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/firex.h"
#include "firex_host.h"
#include "firex_jacobi.h"
#include "firex_kernels.h"
#include "firex_solver_internal.h"

using namespace fx;

struct firex_solver {
  firex_config cfg;
  int32_t dims[3];
  uint32_t cells = 0;
  bool heavy = false;      // >= 64 Mi cell path
  bool f16p = false;       // f16 pressure storage
  bool compact = false;    // compact f32 heavy fallback (shared pressure, cold start)
  Params params;
  EmitterConfig emitters;
  FixedStepClock clock;
  double simulation_time = 0;
  uint64_t frame_index = 0;
  int species_index = 0, pressure_index = 0;
  int flow_a = 0;          // buffer holding the reference's flow[0]
  bool has_liquid = false;
  bool lost = false;
  float4* flow[2] = {nullptr, nullptr};
  void* species[2] = {nullptr, nullptr};
  void* pressure[2] = {nullptr, nullptr};
  void* corr[2] = {nullptr, nullptr};
  float* divergence = nullptr;
  float* reaction = nullptr;
  uint32_t* liquid = nullptr;
  double* scratch = nullptr;
  float* simtab = nullptr;  // simulate's per-step coordinate tables
  void* packed = nullptr;   // render volumes of firex_pack_fields (24 B/cell), on first use
  // SPH droplets: two particle buffers (2 x float4 per particle), density
  // records, a pinned staging buffer for the host-written initial state and
  // the event that marks its last upload.
  float4* particles[2] = {nullptr, nullptr};
  float4* particle_density = nullptr;
  float* particle_host = nullptr;
  hipEvent_t particle_upload = nullptr;
  int particle_index = 0;
  bool liquid_may_contain_water = false;  // FireXEngine.ts:399
  bool particle_field_active = false;     // FireXEngine.ts:400
  // Volume rendering: the 64^3 rgba8 detail volume (uploaded on first use)
  // and the rgba16f scene image of the last firex_render_volume.
  uint32_t* detail = nullptr;
  void* scene = nullptr;
  uint64_t scene_bytes = 0;
  int32_t scene_width = 0, scene_height = 0;
  // presentation of the last firex_render_frame: bloom (rgba16f, quarter
  // size) and the rgba8 frame
  void* post = nullptr;
  uint64_t post_bytes = 0;
  bool has_frame = false;
  // droplet streak rasterizer scratch (sprites, boxes, bins, lists) and the
  // render constants of the last firex_render_volume
  void* sprite_work = nullptr;
  uint64_t sprite_work_size = 0;
  float* bricks = nullptr;  // occupancy bricks of the media volume
  FxRender last_render;
  std::vector<void*> allocations;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  uint64_t bytes = 0;
  JacobiPlan jacobi;
  float jacobi_tuned_us = 0;  // autotuner's time per sweep (0: not tuned)
  int corr_index = 0;      // buffer holding the last residual solution
  // The divergence field holds the residual solve's right-hand side (the
  // divergence of the primary-projected flow, FireXEngine.ts:2257) until the
  // divergence view's refresh in firex_pack_fields replaces it with the final
  // flow's divergence; the residual norm is defined only while this holds.
  bool divergence_is_rhs = true;
  // HIP-event timing of launches (firex_set_pass_timing)
  struct TimedLaunch {
    int pass;
    hipEvent_t start, stop;
    int launches;
  };
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<TimedLaunch> pending;
  firex_pass_timing totals;
  // Captured rest-of-step graphs, one per buffer-index state.
  struct RestGraph {
    int flow_a, pressure_index;
    int post_flow_a, post_pressure_index, post_corr_index;
    hipGraphExec_t exec;
  };
  std::vector<RestGraph> graphs;
  uint64_t graph_key = 0;
};

namespace {

int hip_fail(firex_solver* s, hipError_t e, const char* what) {
  if (s && (e == hipErrorIllegalAddress || e == hipErrorLaunchFailure || e == hipErrorAssert ||
            e == hipErrorECCNotCorrectable || e == hipErrorLaunchTimeOut))
    s->lost = true;
  return set_error(s && s->lost ? FIREX_ERR_DEVICE_LOST : FIREX_ERR_DEVICE,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define FX_HIP(s, call, what)                         \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(s, e_, what); \
  } while (0)

size_t species_word(int storage) { return storage == FIREX_STORAGE_F32 ? 16 : 8; }

int check_usable(firex_solver* s) {
  if (!s) return set_error(FIREX_ERR_INVALID, "solver is null");
  if (s->lost) return set_error(FIREX_ERR_DEVICE_LOST, "The Fire-X HIP device was lost.");
  return FIREX_OK;
}

// Launch errors are reported at the next call that checks them.
int check_launch(firex_solver* s, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(s, e, what);
  return FIREX_OK;
}

void free_all(firex_solver* s) {
  for (void* p : s->allocations) (void)hipFree(p);  // teardown: best effort
  s->allocations.clear();
  if (s->particle_host) (void)hipHostFree(s->particle_host);
  s->particle_host = nullptr;
  if (s->particle_upload) (void)hipEventDestroy(s->particle_upload);
  s->particle_upload = nullptr;
}

int alloc(firex_solver* s, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes < 256 ? 256 : bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FIREX_ERR_OUT_OF_MEMORY,
                     std::string("The selected Fire-X grid could not be allocated (") + what + ", " +
                         std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
  }
  s->allocations.push_back(*p);
  return FIREX_OK;
}

// writeParticleInitialState, FireXEngine.ts:1976-2034: the host computes the
// state (float64, stored as f32) and uploads it to both particle buffers.
int write_particle_state(firex_solver* s, bool active) {
  const int32_t n = s->cfg.particle_count;
  if (n > 0) {
    // the staging buffer is rewritten only after its last upload finished
    FX_HIP(s, hipEventSynchronize(s->particle_upload), "particle state upload");
    particle_initial_state(s->params, n, active, s->particle_host);
    const size_t bytes = (size_t)n * 32;
    for (int k = 0; k < 2; ++k)
      FX_HIP(s, hipMemcpyAsync(s->particles[k], s->particle_host, bytes, hipMemcpyHostToDevice, s->stream),
             "particle state upload");
    FX_HIP(s, hipEventRecord(s->particle_upload, s->stream), "particle state upload");
  }
  s->particle_index = 0;
  s->particle_field_active = active;
  return FIREX_OK;
}

// The liquid field is allocated on first use (droplets or a write).
int ensure_liquid(firex_solver* s) {
  if (s->liquid) return FIREX_OK;
  const uint64_t bytes = (uint64_t)s->cells * 4;
  if (s->cfg.memory_budget_bytes && s->bytes + bytes > s->cfg.memory_budget_bytes)
    return set_error(FIREX_ERR_BUDGET, "the liquid field (" + std::to_string(bytes) +
                                           " bytes) does not fit the memory budget");
  int rc = alloc(s, (void**)&s->liquid, bytes, "liquid");
  if (rc) return rc;
  s->bytes += bytes;
  launch_fill_u32(s->liquid, s->cells, 0, s->stream);
  return FIREX_OK;
}

// initializeFields, FireXEngine.ts:1933-1974: clear every buffer, then
// flow[0] = (0, 0, 0, 300), species[0] = (0, oxygenRate, 0, 0), and write the
// particle state for the current water setting.
int initialize_fields(firex_solver* s) {
  const uint64_t c = s->cells;
  size_t pw = s->f16p ? 2 : 4;
  for (int k = 0; k < 2; ++k) {
    launch_fill_u32(s->flow[k], c * 4, 0, s->stream);
    launch_fill_u32(s->species[k], c * species_word(s->cfg.species_storage) / 4, 0, s->stream);
    launch_fill_u32(s->pressure[k], (c * pw + 3) / 4, 0, s->stream);
    if (!s->compact) launch_fill_u32(s->corr[k], (c * pw + 3) / 4, 0, s->stream);
  }
  launch_fill_u32(s->divergence, c, 0, s->stream);
  launch_fill_u32(s->reaction, c, 0, s->stream);
  if (s->liquid) launch_fill_u32(s->liquid, c, 0, s->stream);
  if (s->particle_density) launch_fill_u32(s->particle_density, (uint64_t)s->cfg.particle_count * 4, 0, s->stream);
  int prc = write_particle_state(s, water_active(s->params));
  if (prc) return prc;
  s->liquid_may_contain_water = false;
  s->flow_a = 0;
  s->corr_index = 0;
  s->divergence_is_rhs = true;  // zero correction against a zero divergence
  launch_initialize(s->cfg.species_storage, s->flow[0], s->species[0], s->cells, 300.0f, (float)s->params.oxygenRate,
                    s->stream);
  return check_launch(s, "field initialization");
}

hipEvent_t take_event(firex_solver* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Folds the finished event pairs into the totals (waits for the stream).
int collect_timing(firex_solver* s) {
  if (s->pending.empty()) return FIREX_OK;
  FX_HIP(s, hipStreamSynchronize(s->stream), "timing synchronize");
  for (const auto& t : s->pending) {
    float ms = 0;
    // a pair whose time cannot be read counts neither its time nor its launches
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      s->totals.milliseconds[t.pass] += ms;
      s->totals.launches[t.pass] += t.launches;
    }
    s->event_pool.push_back(t.start);
    s->event_pool.push_back(t.stop);
  }
  s->pending.clear();
  return FIREX_OK;
}

// Runs launch() (which enqueues `launches` kernels back to back) between one
// pair of timing events when pass timing is on.  A group of consecutive
// launches shares one pair: events between kernels would add their own
// gaps to what is measured.
template <typename F>
void timed(firex_solver* s, int pass, F&& launch, int launches = 1) {
  if (!s->timing) {
    launch();
    return;
  }
  hipEvent_t a = take_event(s), b = take_event(s);
  if (!a || !b) {
    if (a) s->event_pool.push_back(a);
    if (b) s->event_pool.push_back(b);
    launch();
    return;
  }
  const bool started = hipEventRecord(a, s->stream) == hipSuccess;
  launch();
  if (!started || hipEventRecord(b, s->stream) != hipSuccess) {
    s->event_pool.push_back(a);  // untimed: the launches still ran
    s->event_pool.push_back(b);
    return;
  }
  s->pending.push_back({pass, a, b, launches});
}

float4* flow0(firex_solver* s) { return s->flow[s->flow_a]; }
float4* flow1(firex_solver* s) { return s->flow[1 - s->flow_a]; }

// Jacobi sweeps on a buffer pair; returns the index of the result buffer.
int sweeps(firex_solver* s, int f16, void* const bufs[2], int start, bool zero_start, int count, const FxStep& st,
           bool tb) {
  if (!tb) {
    int cur = start;
    timed(
        s, FIREX_PASS_JACOBI,
        [&] {
          for (int i = 0; i < count; ++i) {
            launch_jacobi(f16, bufs[cur], bufs[1 - cur], s->divergence, st, s->stream);
            cur = 1 - cur;
          }
        },
        count);
    s->totals.jacobi_sweeps += s->timing ? count : 0;
    return cur;
  }
  // count sweeps as launches of up to K sweeps, alternating buffers; only
  // the first reads the zero field instead of bufs[start] when zero_start.
  int result = start;
  timed(
      s, FIREX_PASS_JACOBI,
      [&] {
        bool zero = zero_start;
        for (int left = count; left > 0;) {
          const int k = left < s->jacobi.sweeps_per_launch ? left : s->jacobi.sweeps_per_launch;
          launch_jacobi_tb(s->jacobi, f16, JacobiLaunch{k, result, zero}, bufs, s->divergence, st, s->stream);
          result = 1 - result;
          zero = false;
          left -= k;
        }
      },
      (count + s->jacobi.sweeps_per_launch - 1) / s->jacobi.sweeps_per_launch);
  s->totals.jacobi_sweeps += s->timing ? count : 0;
  return result;
}

void fill(firex_solver* s, void* p, uint64_t words) {
  timed(s, FIREX_PASS_FILL_COPY, [&] { launch_fill_u32(p, words, 0, s->stream); });
}

// simulate, the first pass of encodeSimulationStep (FireXEngine.ts:2172-2173).
void encode_simulate(firex_solver* s, const FxStep& st) {
  SimulateArgs a{flow0(s), s->species[s->species_index], flow1(s), s->species[1 - s->species_index], s->reaction,
                 s->liquid, s->simtab, st};
  timed(s, FIREX_PASS_SIMULATE, [&] { launch_simulate(s->cfg.species_storage, a, s->stream); });
  s->species_index = 1 - s->species_index;
}

// The rest of the step (FireXEngine.ts:2174-2275): divergence, both Jacobi
// solves, both projections, clears and copyFlow.  Its launches depend only
// on the grid metric and on the buffer-index state (flow_a, pressure_index),
// never on the time, so it can be captured once per state and replayed.
void encode_rest(firex_solver* s, const FxStep& st) {
  hipStream_t q = s->stream;
  const bool fused = s->cfg.plan == FIREX_PLAN_FUSED;
  const int N = s->cfg.pressure_iterations, M = s->cfg.correction_iterations;
  const uint64_t pwords = ((uint64_t)s->cells * (s->f16p ? 2 : 4) + 3) / 4;
  const int f16 = s->f16p ? 1 : 0;

  timed(s, FIREX_PASS_DIVERGENCE, [&] { launch_divergence(flow1(s), s->divergence, st, q); });

  if (s->heavy && s->compact) {
    // compact f32 fallback, FireXEngine.ts:2214-2246: shared fields, cleared
    // before each solve (no warm start).  In the fused plan the first launch
    // of a solve reads the cleared field as zeros without reading it, and
    // every launch writes its whole output field, so the clears are only
    // needed where a solve has no sweep; copyFlow becomes a buffer swap as in
    // the other fused paths.
    if (!fused || N == 0) {
      fill(s, s->pressure[0], pwords);
      fill(s, s->pressure[1], pwords);
    }
    s->pressure_index = sweeps(s, 0, s->pressure, 0, fused, N, st, fused);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(0, flow1(s), s->pressure[s->pressure_index], flow0(s), st, q); });
    timed(s, FIREX_PASS_DIVERGENCE, [&] { launch_divergence(flow0(s), s->divergence, st, q); });
    if (!fused || M == 0) {
      fill(s, s->pressure[0], pwords);
      fill(s, s->pressure[1], pwords);
    }
    s->corr_index = sweeps(s, 0, s->pressure, 0, fused, M, st, fused);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(0, flow0(s), s->pressure[s->corr_index], flow1(s), st, q); });
    if (fused)
      s->flow_a = 1 - s->flow_a;
    else
      timed(s, FIREX_PASS_FILL_COPY, [&] { launch_copy_u128(flow0(s), flow1(s), s->cells, q); });
  } else if (fused) {
    // The same passes, merged: temporally blocked sweeps, the residual
    // solve started from its known zero field, and copyFlow replaced by
    // swapping the flow buffers.  (Projection and the residual divergence
    // stay two kernels: a fused kernel must recompute the neighbours'
    // projected velocity and measured slower than the pair.)
    s->pressure_index = sweeps(s, f16, s->pressure, s->pressure_index, false, N, st, true);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(f16, flow1(s), s->pressure[s->pressure_index], flow0(s), st, q); });
    timed(s, FIREX_PASS_DIVERGENCE, [&] { launch_divergence(flow0(s), s->divergence, st, q); });
    if (M == 0) {
      fill(s, s->corr[0], pwords);
      fill(s, s->corr[1], pwords);
    }
    s->corr_index = sweeps(s, f16, s->corr, 0, true, M, st, true);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(f16, flow0(s), s->corr[s->corr_index], flow1(s), st, q); });
    s->flow_a = 1 - s->flow_a;
  } else {
    s->pressure_index = sweeps(s, f16, s->pressure, s->pressure_index, false, N, st, false);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(f16, flow1(s), s->pressure[s->pressure_index], flow0(s), st, q); });
    timed(s, FIREX_PASS_DIVERGENCE, [&] { launch_divergence(flow0(s), s->divergence, st, q); });
    fill(s, s->corr[0], pwords);
    fill(s, s->corr[1], pwords);
    s->corr_index = sweeps(s, f16, s->corr, 0, false, M, st, false);
    timed(s, FIREX_PASS_PROJECT, [&] { launch_project(f16, flow0(s), s->corr[s->corr_index], flow1(s), st, q); });
    timed(s, FIREX_PASS_FILL_COPY, [&] { launch_copy_u128(flow0(s), flow1(s), s->cells, q); });
  }
}

// Key of the values the captured launches bake in (metric and stencil).
uint64_t graph_key(const FxStep& st) {
  uint64_t h = 0x9e3779b97f4a7c15ull;
  const float v[] = {st.mx, st.my, st.mz, st.wx, st.wy, st.wz, st.jacobi_inv};
  for (float f : v) h = (h ^ fxm_float_to_bits(f)) * 0x100000001b3ull;
  return h;
}

void drop_graphs(firex_solver* s) {
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.exec);  // best effort
  s->graphs.clear();
}

// Replays (capturing on first use) the rest of the step as one graph launch.
// Returns false when capture is unavailable; the caller then runs it eagerly.
bool launch_rest_graph(firex_solver* s, const FxStep& st) {
  const uint64_t key = graph_key(st);
  if (key != s->graph_key) {
    drop_graphs(s);
    s->graph_key = key;
  }
  for (const auto& g : s->graphs) {
    if (g.flow_a == s->flow_a && g.pressure_index == s->pressure_index) {
      if (hipGraphLaunch(g.exec, s->stream) != hipSuccess) {
        (void)hipGetLastError();  // state untouched: the caller runs the step eagerly
        return false;
      }
      s->flow_a = g.post_flow_a;
      s->pressure_index = g.post_pressure_index;
      s->corr_index = g.post_corr_index;
      return true;
    }
  }
  firex_solver::RestGraph g;
  g.flow_a = s->flow_a;
  g.pressure_index = s->pressure_index;
  const int corr_before = s->corr_index;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  encode_rest(s, st);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);  // the executable graph, if any, is independent of it
  if (e != hipSuccess) {
    (void)hipGetLastError();
    s->flow_a = g.flow_a;  // nothing ran: restore the state for the eager path
    s->pressure_index = g.pressure_index;
    s->corr_index = corr_before;
    return false;
  }
  g.post_flow_a = s->flow_a;
  g.post_pressure_index = s->pressure_index;
  g.post_corr_index = s->corr_index;
  s->graphs.push_back(g);
  if (hipGraphLaunch(g.exec, s->stream) != hipSuccess) {
    // The capture advanced the buffer indices but nothing ran: put them back
    // so the caller's eager fallback starts from this step's real state.
    (void)hipGetLastError();
    s->flow_a = g.flow_a;
    s->pressure_index = g.pressure_index;
    s->corr_index = corr_before;
    return false;
  }
  return true;
}

// The water branch at the head of encodeSimulationStep
// (FireXEngine.ts:2158-2171): SPH density, integration into the other
// particle buffer and deposition into the cleared liquid field while water is
// on; once it is off, one clear of the liquid field and an inactive particle
// state.
int droplet_passes(firex_solver* s, const float u[FXU_COUNT]) {
  if (water_active(s->params)) {
    int rc = s->particle_field_active ? FIREX_OK : write_particle_state(s, true);
    if (!rc) rc = ensure_liquid(s);
    if (rc) return rc;
    fill(s, s->liquid, s->cells);
    const FxDroplets d = make_droplets(u, s->dims);
    const int pi = s->particle_index;
    timed(
        s, FIREX_PASS_DROPLETS,
        [&] {
          launch_sph_density(s->particles[pi], s->particle_density, d, s->stream);
          launch_sph_integrate(s->particles[pi], s->particles[1 - pi], s->particle_density, flow0(s), d, s->stream);
          launch_liquid_deposit(s->particles[1 - pi], s->particles[pi], s->liquid, d, s->stream);
        },
        3);
    s->particle_index = 1 - pi;
    s->liquid_may_contain_water = true;
    s->has_liquid = true;
  } else if (s->liquid_may_contain_water || s->particle_field_active) {
    if (s->liquid) fill(s, s->liquid, s->cells);
    int rc = write_particle_state(s, false);
    if (rc) return rc;
    s->liquid_may_contain_water = false;
    s->has_liquid = false;
  }
  return FIREX_OK;
}

// One solver step with the given uniform block (encodeSimulationStep,
// FireXEngine.ts:2158-2275).
int encode_step(firex_solver* s, const float u[FXU_COUNT]) {
  int drc = droplet_passes(s, u);
  if (drc) return drc;
  const FxStep st = make_step(u, s->dims, s->emitters, s->frame_index, s->has_liquid);
  encode_simulate(s, st);
  if (s->cfg.capture_graphs && !s->timing) {
    if (!launch_rest_graph(s, st)) {
      s->cfg.capture_graphs = 0;  // capture unavailable: stay eager from now on
      drop_graphs(s);
      encode_rest(s, st);
    }
  } else {
    encode_rest(s, st);
  }
  s->divergence_is_rhs = true;
  if (s->timing) {
    s->totals.steps += 1;
    if (s->pending.size() > 8192) {
      int rc = collect_timing(s);
      if (rc) return rc;
    }
  }
  return check_launch(s, "solver step");
}

int step_with_dt(firex_solver* s, double dt) {
  float u[FXU_COUNT];
  write_uniforms(s->params, s->dims, dt, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                 s->cfg.cinematic != 0, u, s->cfg.particle_count);
  int rc = encode_step(s, u);
  if (rc) return rc;
  s->simulation_time += dt;
  s->frame_index += 1;
  return FIREX_OK;
}

int correction_buffer(firex_solver* s, void** p) {
  *p = s->compact ? s->pressure[s->corr_index] : s->corr[s->corr_index];
  return FIREX_OK;
}

// Where a readable field lives and how it is stored.
struct FieldView {
  const void* ptr = nullptr;  // null: a liquid field never written (all zero)
  uint64_t records = 0;       // cells, or particles for the per-particle fields
  int comps = 1;
  int kind = 0;               // FIREX_KIND_*
};

uint64_t field_element_bytes(int kind) {
  return kind == FIREX_KIND_U8 ? 1 : kind == FIREX_KIND_F16 || kind == FIREX_KIND_BF16 ? 2 : 4;
}

uint64_t field_stored_bytes(const FieldView& v) { return v.records * (uint64_t)v.comps * field_element_bytes(v.kind); }

// Bytes of the presentation buffers for a w x h frame: the rgba8 frame, then
// the quarter-size rgba16f bloom.
uint64_t post_frame_bytes(int32_t w, int32_t h) { return (uint64_t)w * (uint64_t)h * 4; }
uint64_t post_bytes_for(int32_t w, int32_t h) {
  const uint64_t bw = (uint64_t)std::max((w + 3) / 4, 1), bh = (uint64_t)std::max((h + 3) / 4, 1);
  return (post_frame_bytes(w, h) + 255) / 256 * 256 + bw * bh * 8;
}

int field_view(firex_solver* s, int32_t field, FieldView* v) {
  *v = FieldView();
  v->records = s->cells;
  switch (field) {
    case FIREX_FIELD_FLOW: v->ptr = flow0(s); v->comps = 4; break;
    case FIREX_FIELD_FLOW_PREDICTED:
      // After a step the reference's flow[1] equals flow[0] (copyFlow); the
      // fused plan swaps instead of copying, so both read the same buffer.
      v->ptr = s->cfg.plan == FIREX_PLAN_FUSED && s->frame_index > 0 ? flow0(s) : flow1(s);
      v->comps = 4;
      break;
    case FIREX_FIELD_SPECIES:
      v->ptr = s->species[s->species_index];
      v->comps = 4;
      v->kind = s->cfg.species_storage;
      break;
    case FIREX_FIELD_PRESSURE: v->ptr = s->pressure[s->pressure_index]; v->kind = s->f16p ? 1 : 0; break;
    case FIREX_FIELD_CORRECTION: {
      void* p;
      correction_buffer(s, &p);
      v->ptr = p;
      v->kind = s->f16p ? 1 : 0;
      break;
    }
    case FIREX_FIELD_DIVERGENCE: v->ptr = s->divergence; break;
    case FIREX_FIELD_REACTION: v->ptr = s->reaction; break;
    case FIREX_FIELD_LIQUID: v->ptr = s->liquid; v->kind = FIREX_KIND_U32; break;
    case FIREX_FIELD_PACKED_FLOW: case FIREX_FIELD_PACKED_CHEMISTRY: case FIREX_FIELD_PACKED_MEDIA:
      if (!s->packed) return set_error(FIREX_ERR_INVALID, "the render volumes are filled by firex_pack_fields");
      v->ptr = (const uint8_t*)s->packed + (size_t)(field - FIREX_FIELD_PACKED_FLOW) * s->cells * 8;
      v->comps = 4;
      v->kind = FIREX_KIND_F16;
      break;
    case FIREX_FIELD_PARTICLES:
      v->ptr = s->particles[s->particle_index];
      v->records = (uint64_t)s->cfg.particle_count;
      v->comps = 8;
      break;
    case FIREX_FIELD_PARTICLE_DENSITY:
      v->ptr = s->particle_density;
      v->records = (uint64_t)s->cfg.particle_count;
      v->comps = 4;
      break;
    case FIREX_FIELD_FRAME:
      if (!s->has_frame) return set_error(FIREX_ERR_INVALID, "the frame is written by firex_render_frame");
      v->ptr = s->post;
      v->records = (uint64_t)s->scene_width * (uint64_t)s->scene_height;
      v->comps = 4;
      v->kind = FIREX_KIND_U8;
      break;
    case FIREX_FIELD_SCENE:
      if (!s->scene) return set_error(FIREX_ERR_INVALID, "the scene image is written by firex_render_volume");
      v->ptr = s->scene;
      v->records = (uint64_t)s->scene_width * (uint64_t)s->scene_height;
      v->comps = 4;
      v->kind = FIREX_KIND_F16;
      break;
    default: return set_error(FIREX_ERR_INVALID, "unknown field");
  }
  return FIREX_OK;
}

}  // namespace

extern "C" int firex_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

namespace fx {

// Validates a configuration; fills the path flags and the device bytes.
int solver_plan(const firex_config* cfg, bool* heavy_out, bool* f16p_out, bool* compact_out, uint64_t* bytes_out) {
  if (!cfg) return set_error(FIREX_ERR_INVALID, "config is null");
  for (int i = 0; i < 3; ++i) {
    if (cfg->dims[i] < 2 || cfg->dims[i] > 1024)
      return set_error(FIREX_ERR_UNSUPPORTED, "each grid axis must be in [2, 1024]");
  }
  if (cfg->pressure_iterations < 0 || cfg->correction_iterations < 0)
    return set_error(FIREX_ERR_INVALID, "iteration counts must be non-negative");
  if (!(cfg->time_step > 0) || !isfinite(cfg->time_step))
    return set_error(FIREX_ERR_INVALID, "Fixed step must be finite and positive.");
  if (cfg->maximum_substeps <= 0) return set_error(FIREX_ERR_INVALID, "Maximum step count must be a positive integer.");
  if (cfg->species_storage < 0 || cfg->species_storage > 2) return set_error(FIREX_ERR_INVALID, "unknown species storage");
  if (cfg->heavy_mode < 0 || cfg->heavy_mode > 3) return set_error(FIREX_ERR_INVALID, "unknown heavy mode");
  if (cfg->plan < 0 || cfg->plan > 1) return set_error(FIREX_ERR_INVALID, "unknown plan");
  if (cfg->particle_count < 0 || cfg->particle_count > (1 << 20))
    return set_error(FIREX_ERR_INVALID, "particle count must be in [0, 2^20]");
  if (cfg->jacobi_tile_height != 0 && cfg->jacobi_tile_height != 32 && cfg->jacobi_tile_height != 48 &&
      cfg->jacobi_tile_height != 64)
    return set_error(FIREX_ERR_INVALID, "jacobi tile height must be 0, 32, 48 or 64");
  if (cfg->jacobi_ring_waves != 0 && cfg->jacobi_ring_waves != 8 && cfg->jacobi_ring_waves != kRingWideWaves)
    return set_error(FIREX_ERR_INVALID, "jacobi ring waves must be 0 (tuned), 8 or 16");
  if (cfg->jacobi_ring_waves == kRingWideWaves && cfg->jacobi_tile_height != 0 && cfg->jacobi_tile_height != 32)
    return set_error(FIREX_ERR_INVALID, "16 ring waves make 32-row tiles only");
  if (cfg->jacobi_prefetch < 0 || cfg->jacobi_prefetch > 2)
    return set_error(FIREX_ERR_INVALID, "jacobi prefetch must be 0 (tuned), 1 (off) or 2 (on)");

  bool heavy = cfg->heavy_mode == FIREX_HEAVY_AUTO ? heavy_grid(cfg->dims)
                                                     : (cfg->heavy_mode == FIREX_HEAVY_F16 || cfg->heavy_mode == FIREX_HEAVY_F32);
  bool f16p = heavy && cfg->heavy_mode != FIREX_HEAVY_F32;
  bool compact = heavy && !f16p;
  uint64_t need = device_bytes(cfg->dims, cfg->species_storage, f16p, compact, false) +
                  particle_device_bytes(cfg->particle_count);
  if (cfg->memory_budget_bytes && need > cfg->memory_budget_bytes)
    return set_error(FIREX_ERR_BUDGET, "Transactional rebuild peak requires " + std::to_string(need) +
                                           " bytes, above the selected " + std::to_string(cfg->memory_budget_bytes) +
                                           " byte guard.");
  if (heavy_out) *heavy_out = heavy;
  if (f16p_out) *f16p_out = f16p;
  if (compact_out) *compact_out = compact;
  if (bytes_out) *bytes_out = need;
  return FIREX_OK;
}

// Creates a plume.  With a non-null stream the solver issues its work there
// and neither creates nor destroys a stream of its own, and the initial
// field reset is left in flight on that stream.
int solver_create(const firex_config* cfg, hipStream_t stream, firex_solver** out) {
  clear_error();
  if (!cfg || !out) return set_error(FIREX_ERR_INVALID, "config and out must be non-null");
  *out = nullptr;
  bool heavy = false, f16p = false, compact = false;
  uint64_t need = 0;
  int prc = solver_plan(cfg, &heavy, &f16p, &compact, &need);
  if (prc) return prc;

  int ndev = firex_device_count();
  if (ndev <= 0) return set_error(FIREX_ERR_DEVICE, "No HIP device is available for the Fire-X solver.");
  if (cfg->device < 0 || cfg->device >= ndev) return set_error(FIREX_ERR_INVALID, "device ordinal out of range");

  firex_solver* s = new firex_solver();
  s->cfg = *cfg;
  memcpy(s->dims, cfg->dims, sizeof s->dims);
  s->cells = (uint32_t)((uint64_t)cfg->dims[0] * cfg->dims[1] * cfg->dims[2]);
  s->heavy = heavy;
  s->f16p = f16p;
  s->compact = compact;
  s->clock.step_seconds = cfg->time_step;
  s->clock.maximum_steps = cfg->maximum_substeps;
  s->bytes = need;

  hipError_t e = hipSetDevice(cfg->device);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(nullptr, e, "hipSetDevice");
  }
  s->own_stream = stream == nullptr;
  if (stream) {
    s->stream = stream;
  } else {
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete s;
      return hip_fail(nullptr, e, "hipStreamCreate");
    }
  }
  const uint64_t c = s->cells;
  // Rounded up to whole 32-bit words: the fill kernel clears words
  // (the reference rounds the same way, FireXEngine.ts:1560).
  const size_t pbytes = (c * (f16p ? 2 : 4) + 3) / 4 * 4;
  int rc = FIREX_OK;
  for (int k = 0; k < 2 && !rc; ++k) {
    rc = alloc(s, (void**)&s->flow[k], c * 16, "flow");
    if (!rc) rc = alloc(s, &s->species[k], c * species_word(cfg->species_storage), "species");
    if (!rc) rc = alloc(s, &s->pressure[k], pbytes, "pressure");
    if (!rc && !compact) rc = alloc(s, &s->corr[k], pbytes, "residual pressure");
  }
  if (!rc) rc = alloc(s, (void**)&s->divergence, c * 4, "divergence");
  if (!rc) rc = alloc(s, (void**)&s->reaction, c * 4, "reaction");
  if (!rc) rc = alloc(s, (void**)&s->scratch, kReduceScratchBytes, "reduction scratch");
  if (!rc) rc = alloc(s, (void**)&s->simtab, sim_table_bytes(s->dims), "simulate tables");
  if (!rc && cfg->particle_count > 0) {
    const size_t n = (size_t)cfg->particle_count;
    rc = alloc(s, (void**)&s->particles[0], n * 32, "particles");
    if (!rc) rc = alloc(s, (void**)&s->particles[1], n * 32, "particles");
    if (!rc) rc = alloc(s, (void**)&s->particle_density, n * 16, "particle density");
    if (!rc && hipHostMalloc((void**)&s->particle_host, n * 32, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      s->particle_host = nullptr;
      rc = set_error(FIREX_ERR_OUT_OF_MEMORY, "pinned particle staging buffer");
    }
    if (!rc) {
      e = hipEventCreateWithFlags(&s->particle_upload, hipEventDisableTiming);
      if (e != hipSuccess) rc = hip_fail(s, e, "particle upload event");
    }
  }
  if (!rc) {
    if (cfg->jacobi_autotune && cfg->plan == FIREX_PLAN_FUSED) {
      // Times candidate plans on the solver's own pressure and divergence
      // fields, which initialize_fields() clears afterwards.
      void* bufs[2] = {s->pressure[0], s->pressure[1]};
      e = (hipError_t)tune_jacobi_plan(s->dims, f16p ? 1 : 0, cfg->jacobi_sweeps_per_launch, cfg->jacobi_chunk_z,
                                       cfg->jacobi_kernel, cfg->jacobi_tile_height, cfg->jacobi_prefetch,
                                       cfg->jacobi_ring_waves, bufs,
                                       s->divergence, s->stream, &s->jacobi, &s->jacobi_tuned_us);
      if (e != hipSuccess) rc = hip_fail(s, e, "Jacobi plan autotuning");
    } else {
      s->jacobi = make_jacobi_plan(s->dims, cfg->jacobi_sweeps_per_launch, cfg->jacobi_chunk_z, cfg->jacobi_kernel,
                                   cfg->jacobi_tile_height, cfg->jacobi_prefetch, cfg->jacobi_ring_waves);
    }
    if (!rc) rc = initialize_fields(s);
  }
  if (!rc && s->own_stream) {
    e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = hip_fail(s, e, "field initialization");
  }
  if (rc) {
    std::string msg = firex_last_error();
    if (!s->own_stream) (void)hipStreamSynchronize(s->stream);  // cleanup after a failure: best effort
    free_all(s);
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return set_error(rc, msg);
  }
  *out = s;
  return FIREX_OK;
}

// Enqueues one step of dt seconds without waiting.
int solver_step_async(firex_solver* s, double dt) {
  int rc = check_usable(s);
  if (rc) return rc;
  return step_with_dt(s, dt);
}

uint64_t solver_group_key(firex_solver* s) {
  if (!s || s->lost || s->timing || s->heavy || s->cfg.plan != FIREX_PLAN_FUSED || !jacobi_plan_batchable(s->jacobi))
    return 0;
  const uint64_t v[] = {(uint64_t)(uintptr_t)s->stream,           (uint64_t)s->cfg.device,
                        (uint64_t)s->f16p,                         (uint64_t)s->cfg.pressure_iterations,
                        (uint64_t)s->cfg.correction_iterations,    (uint64_t)s->jacobi.sweeps_per_launch};
  uint64_t h = 0xcbf29ce484222325ull;
  for (uint64_t x : v) h = (h ^ x) * 0x100000001b3ull;
  return h | 1;
}

namespace {

bool same_group(const firex_solver* a, const firex_solver* b) {
  return a->stream == b->stream && a->cfg.device == b->cfg.device && a->f16p == b->f16p &&
         a->cfg.pressure_iterations == b->cfg.pressure_iterations &&
         a->cfg.correction_iterations == b->cfg.correction_iterations &&
         a->jacobi.sweeps_per_launch == b->jacobi.sweeps_per_launch;
}

// `count` sweeps on every plume of a group, in batched ring launches of up
// to K sweeps; corr picks the residual pair.  idx[i] is plume i's source
// buffer on entry and its result buffer on return.
void group_sweeps(firex_solver* const* g, int n, bool corr, const FxStep* st, int count, bool zero_start, int* idx) {
  const JacobiPlan& plan = g[0]->jacobi;
  bool zero = zero_start;
  for (int left = count; left > 0;) {
    const int k = left < plan.sweeps_per_launch ? left : plan.sweeps_per_launch;
    RingBatch b;
    b.count = n;
    for (int i = 0; i < n; ++i) {
      const firex_solver* s = g[i];
      void* const* bufs = corr ? s->corr : s->pressure;
      RingBatchEntry& e = b.entry[i];
      e.pin = bufs[idx[i]];
      e.pout = bufs[1 - idx[i]];
      e.div = s->divergence;
      e.nx = s->dims[0];
      e.ny = s->dims[1];
      e.nz = s->dims[2];
      e.wx = st[i].wx;
      e.wy = st[i].wy;
      e.wz = st[i].wz;
      e.inv = st[i].jacobi_inv;
      e.chunk_z = s->jacobi.chunk_z;
      e.zero_start = zero ? 1 : 0;
      idx[i] = 1 - idx[i];
    }
    launch_jacobi_ring_batch(plan, g[0]->f16p ? 1 : 0, k, b, g[0]->stream);
    zero = false;
    left -= k;
  }
}

}  // namespace

// The fused branch of encode_rest over a group, with its Jacobi launches
// batched (FireXEngine.ts:2158-2275 for each plume).
int solver_step_group(firex_solver* const* g, int32_t n, double dt) {
  static_assert(kStepGroupMax == kRingBatchMax, "a group's plumes share one ring launch");
  if (!g || n < 1 || n > kStepGroupMax) return set_error(FIREX_ERR_INVALID, "a group holds 1 to 16 plumes");
  if (!isfinite(dt) || dt < 0) return set_error(FIREX_ERR_INVALID, "step dt must be finite and non-negative");
  for (int i = 0; i < n; ++i) {
    int rc = check_usable(g[i]);
    if (rc) return rc;
    if (solver_group_key(g[i]) == 0 || !same_group(g[0], g[i]))
      return set_error(FIREX_ERR_INVALID, "the plumes of a group must share a stream, pressure storage, iteration "
                                          "counts and a batched ring Jacobi plan");
    for (int j = 0; j < i; ++j)
      if (g[j] == g[i]) return set_error(FIREX_ERR_INVALID, "a plume appears twice in the group");
  }
  FX_HIP(g[0], hipSetDevice(g[0]->cfg.device), "hipSetDevice");
  if (n == 1) return step_with_dt(g[0], dt);
  const hipStream_t q = g[0]->stream;
  const int N = g[0]->cfg.pressure_iterations, M = g[0]->cfg.correction_iterations;
  const int f16 = g[0]->f16p ? 1 : 0;
  FxStep st[kRingBatchMax];
  int idx[kRingBatchMax];
  for (int i = 0; i < n; ++i) {
    firex_solver* s = g[i];
    float u[FXU_COUNT];
    write_uniforms(s->params, s->dims, dt, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                   s->cfg.cinematic != 0, u, s->cfg.particle_count);
    int rc = droplet_passes(s, u);
    if (rc) return rc;
    st[i] = make_step(u, s->dims, s->emitters, s->frame_index, s->has_liquid);
    encode_simulate(s, st[i]);
    launch_divergence(flow1(s), s->divergence, st[i], q);
    idx[i] = s->pressure_index;
  }
  group_sweeps(g, n, false, st, N, false, idx);
  for (int i = 0; i < n; ++i) {
    firex_solver* s = g[i];
    s->pressure_index = idx[i];
    launch_project(f16, flow1(s), s->pressure[s->pressure_index], flow0(s), st[i], q);
    launch_divergence(flow0(s), s->divergence, st[i], q);
    if (M == 0) {
      const uint64_t pwords = ((uint64_t)s->cells * (s->f16p ? 2 : 4) + 3) / 4;
      launch_fill_u32(s->corr[0], pwords, 0, q);
      launch_fill_u32(s->corr[1], pwords, 0, q);
    }
    idx[i] = 0;
  }
  group_sweeps(g, n, true, st, M, true, idx);
  for (int i = 0; i < n; ++i) {
    firex_solver* s = g[i];
    s->corr_index = idx[i];
    launch_project(f16, flow0(s), s->corr[s->corr_index], flow1(s), st[i], q);
    s->flow_a = 1 - s->flow_a;
    s->divergence_is_rhs = true;
    s->simulation_time += dt;
    s->frame_index += 1;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    for (int i = 1; i < n; ++i) hip_fail(g[i], e, "group step");
    return hip_fail(g[0], e, "group step");
  }
  return FIREX_OK;
}

hipStream_t solver_stream(firex_solver* s) { return s->stream; }
uint64_t solver_bytes(firex_solver* s) { return s->bytes; }
uint32_t solver_cells(firex_solver* s) { return s->cells; }

}  // namespace fx

extern "C" {

int firex_create(const firex_config* cfg, firex_solver** out) { return fx::solver_create(cfg, nullptr, out); }

void firex_destroy(firex_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->cfg.device);  // teardown: best effort
  if (!s->lost) (void)hipStreamSynchronize(s->stream);
  for (const auto& t : s->pending) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  drop_graphs(s);
  free_all(s);
  if (s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

namespace {

// formatFireXBytes (src/sim/firex/gridConfiguration.ts:386-399): KiB below
// 1 MiB, else MiB, GiB or TiB with two decimals (one from 10 up), trailing
// zeros dropped (formatDecimal, :372-375).
std::string format_bytes(double bytes) {
  auto dec = [](double v, int digits) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.*f", digits, v);
    std::string t = buf;
    const size_t dot = t.find('.');
    if (dot != std::string::npos) {
      size_t end = t.size();
      while (end > dot + 1 && t[end - 1] == '0') --end;
      if (end == dot + 1) end = dot;
      t.resize(end);
    }
    return t;
  };
  if (!(bytes >= 0)) return "\u2014";
  const double mib = bytes / (1024.0 * 1024.0);
  if (mib < 1) return dec(bytes / 1024.0, 1) + " KiB";
  const double gib = mib / 1024.0;
  if (gib >= 1024) {
    const double tib = gib / 1024.0;
    return dec(tib, tib >= 10 ? 1 : 2) + " TiB";
  }
  if (gib >= 1) return dec(gib, gib >= 10 ? 1 : 2) + " GiB";
  return dec(mib, mib >= 10 ? 1 : 2) + " MiB";
}

uint64_t estimated_bytes(const int32_t dims[3], int32_t particles) {
  firex_memory_estimate e;
  firex_estimate_memory(dims, particles, &e);
  return (uint64_t)e.total_bytes;
}

bool same_buffer_set(const firex_solver* s, const firex_config& c, bool heavy, bool f16p, bool compact) {
  const firex_config& o = s->cfg;
  return c.dims[0] == o.dims[0] && c.dims[1] == o.dims[1] && c.dims[2] == o.dims[2] &&
         c.particle_count == o.particle_count && c.species_storage == o.species_storage && heavy == s->heavy &&
         f16p == s->f16p && compact == s->compact && c.plan == o.plan &&
         c.jacobi_sweeps_per_launch == o.jacobi_sweeps_per_launch && c.jacobi_chunk_z == o.jacobi_chunk_z &&
         c.jacobi_kernel == o.jacobi_kernel && c.jacobi_autotune == o.jacobi_autotune &&
         c.jacobi_tile_height == o.jacobi_tile_height && c.jacobi_prefetch == o.jacobi_prefetch &&
         c.jacobi_ring_waves == o.jacobi_ring_waves;
}

}  // namespace

int firex_reconfigure(firex_solver* s, const firex_config* config, uint64_t memory_guard_bytes, int32_t* rebuilt) {
  clear_error();
  int rc = check_usable(s);
  if (rc) return rc;
  if (!config) return set_error(FIREX_ERR_INVALID, "config is null");
  if (!s->own_stream) return set_error(FIREX_ERR_INVALID, "a plume on a serving stream cannot be reconfigured");
  if (config->device != s->cfg.device) return set_error(FIREX_ERR_INVALID, "the device of a solver cannot change");
  firex_config next = *config;
  next.memory_budget_bytes = 0;  // the transactional guard below replaces the creation guard
  bool heavy = false, f16p = false, compact = false;
  uint64_t need = 0;
  rc = solver_plan(&next, &heavy, &f16p, &compact, &need);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  FX_HIP(s, hipStreamSynchronize(s->stream), "reconfigure");
  rc = collect_timing(s);
  if (rc) return rc;
  if (same_buffer_set(s, next, heavy, f16p, compact)) {
    // The buffer set stays (FireXEngine.ts:814-820, 870-880 keep matching
    // resources): only the per-step settings change, the fields are kept.
    if (next.pressure_iterations != s->cfg.pressure_iterations ||
        next.correction_iterations != s->cfg.correction_iterations || next.capture_graphs != s->cfg.capture_graphs)
      drop_graphs(s);
    s->cfg.pressure_iterations = next.pressure_iterations;
    s->cfg.correction_iterations = next.correction_iterations;
    s->cfg.time_step = next.time_step;
    s->cfg.maximum_substeps = next.maximum_substeps;
    s->cfg.cinematic = next.cinematic;
    s->cfg.ray_steps = next.ray_steps;
    s->cfg.capture_graphs = next.capture_graphs;
    s->cfg.memory_budget_bytes = config->memory_budget_bytes;
    s->clock = FixedStepClock();  // a new FixedStepClock for the configuration
    s->clock.step_seconds = next.time_step;
    s->clock.maximum_steps = next.maximum_substeps;
    if (rebuilt) *rebuilt = 0;
    return FIREX_OK;
  }
  // assertGridSupported (FireXEngine.ts:2823-2853): the estimated peak of the
  // current and the replacement set must fit the guard.
  const uint64_t current = estimated_bytes(s->dims, s->cfg.particle_count);
  const uint64_t replacement = estimated_bytes(next.dims, next.particle_count);
  if (memory_guard_bytes && current + replacement > memory_guard_bytes)
    return set_error(FIREX_ERR_BUDGET, "Transactional rebuild peak requires about " +
                                           format_bytes((double)(current + replacement)) + " (" +
                                           format_bytes((double)current) + " current + " +
                                           format_bytes((double)replacement) + " replacement), above the selected " +
                                           format_bytes((double)memory_guard_bytes) + " guard.");
  // rebuildResourcesSafely (FireXEngine.ts:1891-1915): build the replacement
  // while the current set stays live; any failure leaves the solver as it was.
  firex_solver* fresh = nullptr;
  rc = solver_create(&next, nullptr, &fresh);
  if (rc) return rc;
  fresh->cfg.memory_budget_bytes = config->memory_budget_bytes;
  // installResourceSet (:1917-1931): the parameters carry over, the clock,
  // time and indices start again and the fields are initialised with the
  // current parameters.
  fresh->params = s->params;
  fresh->emitters = s->emitters;
  rc = initialize_fields(fresh);
  if (!rc) {
    const hipError_t e = hipStreamSynchronize(fresh->stream);
    if (e != hipSuccess) rc = hip_fail(fresh, e, "field initialization");
  }
  if (rc) {
    const std::string msg = firex_last_error();
    firex_destroy(fresh);
    return set_error(rc, msg);
  }
  std::swap(*s, *fresh);  // the handle now holds the new set
  firex_destroy(fresh);   // and the old one is freed
  if (rebuilt) *rebuilt = 1;
  return FIREX_OK;
}

int firex_reset(firex_solver* s) {
  int rc = check_usable(s);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  s->clock.reset();
  s->simulation_time = 0;
  s->frame_index = 0;
  s->species_index = 0;
  s->pressure_index = 0;
  return initialize_fields(s);
}

int firex_set_parameter(firex_solver* s, const char* name, double value) {
  if (!s || !name) return set_error(FIREX_ERR_INVALID, "solver and name must be non-null");
  params_set(s->params, name, value);  // unknown keys are ignored, as in the reference
  return FIREX_OK;
}

int firex_set_parameter_bool(firex_solver* s, const char* name, int32_t value) {
  if (!s || !name) return set_error(FIREX_ERR_INVALID, "solver and name must be non-null");
  params_set_bool(s->params, name, value != 0);
  return FIREX_OK;
}

int firex_get_parameter(firex_solver* s, const char* name, double* value) {
  if (!s || !name || !value) return set_error(FIREX_ERR_INVALID, "solver, name and value must be non-null");
  if (!params_get(s->params, name, value)) return set_error(FIREX_ERR_INVALID, std::string("unknown parameter: ") + name);
  return FIREX_OK;
}

int firex_set_emitters(firex_solver* s, const firex_emitter* list, int32_t count, float jitter, float velocity_noise,
                       uint64_t seed) {
  if (!s) return set_error(FIREX_ERR_INVALID, "solver is null");
  if (count < 0 || count > FX_MAX_EMITTERS) return set_error(FIREX_ERR_INVALID, "between 0 and 8 emitters");
  if (count > 0 && !list) return set_error(FIREX_ERR_INVALID, "emitter list is null");
  EmitterConfig e;
  e.count = count;
  for (int i = 0; i < count; ++i) {
    if (!(list[i].radius > 0) || !isfinite(list[i].radius) || !isfinite(list[i].amplitude))
      return set_error(FIREX_ERR_INVALID, "emitter radius must be positive and values finite");
    e.data[i][0] = list[i].center[0];
    e.data[i][1] = list[i].center[1];
    e.data[i][2] = list[i].center[2];
    e.data[i][3] = list[i].radius;
    e.data[i][4] = list[i].amplitude;
  }
  e.jitter = jitter;
  e.velocity_noise = velocity_noise;
  e.seed = seed;
  s->emitters = e;
  return FIREX_OK;
}

int firex_step(firex_solver* s, double dt, int32_t count) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!isfinite(dt) || dt < 0) return set_error(FIREX_ERR_INVALID, "step dt must be finite and non-negative");
  if (count < 0) return set_error(FIREX_ERR_INVALID, "count must be non-negative");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  for (int32_t i = 0; i < count; ++i) {
    rc = step_with_dt(s, dt);
    if (rc) return rc;
  }
  return FIREX_OK;
}

int firex_render_offline_frame(firex_solver* s, double delta, const volatile int32_t* abort_flag) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!isfinite(delta) || delta < 0)
    return set_error(FIREX_ERR_INVALID, "Offline frame delta must be finite and non-negative.");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  s->clock.reset();
  int32_t count = 0;
  rc = firex_split_offline_delta(delta, s->cfg.time_step, nullptr, 0, &count);
  if (rc) return rc;
  const double step = count > 0 ? delta / (double)count : 0.0;
  for (int32_t i = 0; i < count; ++i) {
    if (abort_flag && *abort_flag) return set_error(FIREX_ERR_ABORTED, "The Fire-X offline render was aborted.");
    rc = step_with_dt(s, step);
    if (rc) return rc;
  }
  return firex_sync(s);
}

int firex_frame(firex_solver* s, double delta, int32_t* steps_taken) {
  int rc = check_usable(s);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  int32_t n = s->clock.advance(delta);
  if (steps_taken) *steps_taken = n;
  if (n == 0) return FIREX_OK;
  // One uniform write covers every step of the frame (FireXEngine.ts:1089-1097).
  float u[FXU_COUNT];
  write_uniforms(s->params, s->dims, s->cfg.time_step, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                 s->cfg.cinematic != 0, u, s->cfg.particle_count);
  for (int32_t i = 0; i < n; ++i) {
    rc = encode_step(s, u);
    if (rc) return rc;
    s->simulation_time += s->cfg.time_step;
    s->frame_index += 1;
  }
  return FIREX_OK;
}

int firex_sync(firex_solver* s) {
  int rc = check_usable(s);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  FX_HIP(s, hipStreamSynchronize(s->stream), "synchronize");
  return FIREX_OK;
}

int firex_read_field(firex_solver* s, int32_t field, void* dst, size_t bytes, int32_t dtype) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!dst) return set_error(FIREX_ERR_INVALID, "dst is null");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  FieldView v;
  rc = field_view(s, field, &v);
  if (rc) return rc;
  const void* src = v.ptr;
  const int kind = v.kind;
  const uint64_t n = v.records * v.comps;
  const size_t sbytes = (size_t)field_stored_bytes(v);
  if (dtype == FIREX_DTYPE_RAW) {
    if (bytes != sbytes) return set_error(FIREX_ERR_INVALID, "byte count does not match the field's stored size");
    FX_HIP(s, hipStreamSynchronize(s->stream), "synchronize before read");
    if (sbytes == 0) return FIREX_OK;
    if (!src) {
      memset(dst, 0, bytes);
      return FIREX_OK;
    }
    FX_HIP(s, hipMemcpy(dst, src, sbytes, hipMemcpyDeviceToHost), "field readback");
    return FIREX_OK;
  }
  size_t elem = dtype == FIREX_DTYPE_F64 ? 8 : 4;
  if (dtype != FIREX_DTYPE_F32 && dtype != FIREX_DTYPE_F64 && dtype != FIREX_DTYPE_U32)
    return set_error(FIREX_ERR_INVALID, "unknown dtype");
  const bool integral = kind == FIREX_KIND_U32 || kind == FIREX_KIND_U8;
  if (integral != (dtype == FIREX_DTYPE_U32) && !(integral && dtype == FIREX_DTYPE_F64))
    return set_error(FIREX_ERR_INVALID, "dtype does not match the field");
  if (bytes != n * elem) return set_error(FIREX_ERR_INVALID, "byte count does not match the field size");
  FX_HIP(s, hipStreamSynchronize(s->stream), "synchronize before read");
  if (n == 0) return FIREX_OK;
  if (!src) {
    memset(dst, 0, bytes);  // no liquid field was ever written: it is zero
    return FIREX_OK;
  }
  std::vector<uint8_t> staging;
  void* host = dst;
  if (!(kind == 0 && dtype == FIREX_DTYPE_F32) && !(kind == 3 && dtype == FIREX_DTYPE_U32)) {
    staging.resize(sbytes);
    host = staging.data();
  }
  FX_HIP(s, hipMemcpy(host, src, sbytes, hipMemcpyDeviceToHost), "field readback");
  if (host == dst) return FIREX_OK;
  for (uint64_t i = 0; i < n; ++i) {
    double x;
    if (kind == 0) x = ((const float*)host)[i];
    else if (kind == 1) x = (double)(float)((const _Float16*)host)[i];
    else if (kind == 2) x = fx_bf16_bits_to_float(((const uint16_t*)host)[i]);
    else if (kind == FIREX_KIND_U8) x = (double)((const uint8_t*)host)[i];
    else x = (double)((const uint32_t*)host)[i];
    if (dtype == FIREX_DTYPE_F64) ((double*)dst)[i] = x;
    else if (dtype == FIREX_DTYPE_U32) ((uint32_t*)dst)[i] = (uint32_t)x;
    else ((float*)dst)[i] = (float)x;
  }
  return FIREX_OK;
}

int firex_get_field_layout(firex_solver* s, int32_t field, firex_field_layout* out) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!out) return set_error(FIREX_ERR_INVALID, "out is null");
  FieldView v;
  rc = field_view(s, field, &v);
  if (rc) return rc;
  out->kind = v.kind;
  out->components = v.comps;
  out->records = (int64_t)v.records;
  out->stored_bytes = (int64_t)field_stored_bytes(v);
  return FIREX_OK;
}

int firex_field_checksum(firex_solver* s, int32_t field, uint64_t* out) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!out) return set_error(FIREX_ERR_INVALID, "out is null");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  FieldView v;
  rc = field_view(s, field, &v);
  if (rc) return rc;
  uint64_t* result = (uint64_t*)(s->scratch + kReducePartials);
  launch_checksum(v.ptr, field_stored_bytes(v), (uint64_t*)s->scratch, result, s->stream);
  rc = check_launch(s, "field checksum");
  if (rc) return rc;
  uint64_t h = 0;
  FX_HIP(s, hipMemcpyAsync(&h, result, sizeof h, hipMemcpyDeviceToHost, s->stream), "checksum readback");
  FX_HIP(s, hipStreamSynchronize(s->stream), "field checksum");
  *out = h;
  return FIREX_OK;
}

int firex_write_field(firex_solver* s, int32_t field, const void* srcp, size_t bytes, int32_t dtype) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!srcp) return set_error(FIREX_ERR_INVALID, "src is null");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  uint64_t c = s->cells;
  void* dst = nullptr;
  int comps = 1, kind = 0;
  switch (field) {
    case FIREX_FIELD_PARTICLES:
      // both buffers, as writeParticleInitialState writes them; the written
      // state is the live one (the next water step advances it)
      dst = s->particles[0];
      c = (uint64_t)s->cfg.particle_count;
      comps = 8;
      break;
    case FIREX_FIELD_FLOW: dst = flow0(s); comps = 4; break;
    case FIREX_FIELD_FLOW_PREDICTED: dst = flow1(s); comps = 4; break;
    case FIREX_FIELD_SPECIES: dst = s->species[s->species_index]; comps = 4; kind = s->cfg.species_storage; break;
    case FIREX_FIELD_PRESSURE: dst = s->pressure[s->pressure_index]; kind = s->f16p ? 1 : 0; break;
    case FIREX_FIELD_DIVERGENCE: dst = s->divergence; s->divergence_is_rhs = true; break;
    case FIREX_FIELD_REACTION: dst = s->reaction; break;
    case FIREX_FIELD_LIQUID:
      kind = 3;
      rc = ensure_liquid(s);
      if (rc) return rc;
      dst = s->liquid;
      s->has_liquid = true;
      break;
    default: return set_error(FIREX_ERR_INVALID, "field is not writable");
  }
  if (dtype != FIREX_DTYPE_F32 && dtype != FIREX_DTYPE_F64 && dtype != FIREX_DTYPE_U32)
    return set_error(FIREX_ERR_INVALID, "unknown dtype");
  if (dtype == FIREX_DTYPE_U32 && kind != 3) return set_error(FIREX_ERR_INVALID, "dtype does not match the field");
  const uint64_t n = c * comps;
  size_t elem = dtype == FIREX_DTYPE_F64 ? 8 : 4;
  if (bytes != n * elem) return set_error(FIREX_ERR_INVALID, "byte count does not match the field size");
  const size_t sbytes = n * (kind == 1 || kind == 2 ? 2 : 4);
  std::vector<uint8_t> staging(sbytes);
  for (uint64_t i = 0; i < n; ++i) {
    double v = dtype == FIREX_DTYPE_F64 ? ((const double*)srcp)[i]
             : dtype == FIREX_DTYPE_U32 ? (double)((const uint32_t*)srcp)[i]
                                         : (double)((const float*)srcp)[i];
    if (kind == 0) ((float*)staging.data())[i] = (float)v;
    else if (kind == 1) ((_Float16*)staging.data())[i] = (_Float16)(float)v;
    else if (kind == 2) ((uint16_t*)staging.data())[i] = fx_float_to_bf16_bits((float)v);
    else ((uint32_t*)staging.data())[i] = dtype == FIREX_DTYPE_U32 ? ((const uint32_t*)srcp)[i] : (uint32_t)v;
  }
  FX_HIP(s, hipStreamSynchronize(s->stream), "synchronize before write");
  if (n == 0) return FIREX_OK;
  FX_HIP(s, hipMemcpy(dst, staging.data(), sbytes, hipMemcpyHostToDevice), "field upload");
  if (field == FIREX_FIELD_PARTICLES) {
    FX_HIP(s, hipMemcpy(s->particles[1], staging.data(), sbytes, hipMemcpyHostToDevice), "field upload");
    s->particle_index = 0;
    s->particle_field_active = true;
  }
  return FIREX_OK;
}

int firex_pack_fields(firex_solver* s) {
  int rc = check_usable(s);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  if (!s->packed) {
    const uint64_t bytes = (uint64_t)s->cells * 24;
    if (s->cfg.memory_budget_bytes && s->bytes + bytes > s->cfg.memory_budget_bytes)
      return set_error(FIREX_ERR_BUDGET, "the render volumes (" + std::to_string(bytes) +
                                             " bytes) do not fit the memory budget");
    rc = alloc(s, &s->packed, bytes, "render volumes");
    if (rc) return rc;
    s->bytes += bytes;
  }
  float u[FXU_COUNT];
  write_uniforms(s->params, s->dims, 0.0, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                 s->cfg.cinematic != 0, u, s->cfg.particle_count);
  const FxStep st = make_step(u, s->dims, s->emitters, s->frame_index, s->has_liquid);
  // u32(clamp(round(controls.x), 0, 7)) == 7u, qualityShaders.ts:55
  const float view = fxm_clamp(nearbyintf(u[FXU_VIEW_MODE]), 0.0f, 7.0f);
  const int divergence_view = (uint32_t)view == 7u ? 1 : 0;
  if (divergence_view) {
    timed(s, FIREX_PASS_DIVERGENCE, [&] { launch_divergence(flow0(s), s->divergence, st, s->stream); });
    s->divergence_is_rhs = false;
  }
  timed(s, FIREX_PASS_PACK, [&] {
    launch_pack_fields(s->cfg.species_storage, flow0(s), s->species[s->species_index], s->reaction, s->liquid,
                       s->divergence, s->packed, st, divergence_view, s->stream);
  });
  return check_launch(s, "packFields");
}

// The droplet streaks over target (rgba16f scene, or the rgba8 frame with
// unorm8), when the particle field is live (particleFieldActive,
// FireXEngine.ts:2636-2639, 2686-2689).
static int draw_streaks(firex_solver* s, const FxRender& r, void* target, int unorm8) {
  const int n = s->cfg.particle_count;
  if (!s->particle_field_active || n <= 0 || !s->particles[0]) return FIREX_OK;
  const uint64_t need = sprite_work_bytes(n);
  if (!s->sprite_work) {
    if (s->cfg.memory_budget_bytes && s->bytes + need > s->cfg.memory_budget_bytes)
      return set_error(FIREX_ERR_BUDGET, "the droplet streak buffers (" + std::to_string(need) +
                                             " bytes) do not fit the memory budget");
    int rc = alloc(s, &s->sprite_work, need, "droplet streak buffers");
    if (rc) return rc;
    s->sprite_work_size = need;
    s->bytes += need;
  }
  timed(s, FIREX_PASS_STREAKS, [&] {
    launch_droplet_streaks(s->particles[s->particle_index], n, r, target, unorm8, s->sprite_work, s->stream);
  });
  return check_launch(s, "droplet streaks");
}

int firex_render_volume(firex_solver* s, const firex_render_config* rc) {
  int rc0 = check_usable(s);
  if (rc0) return rc0;
  if (!rc) return set_error(FIREX_ERR_INVALID, "render config is null");
  if (rc->width < 1 || rc->height < 1 || rc->width > 16384 || rc->height > 16384)
    return set_error(FIREX_ERR_INVALID, "render width and height must be in [1, 16384]");
  const int32_t steps = rc->ray_steps > 0 ? rc->ray_steps : s->cfg.ray_steps;
  if (steps < 1) return set_error(FIREX_ERR_INVALID, "ray steps must be positive");
  int rc1 = firex_pack_fields(s);
  if (rc1) return rc1;
  if (!s->detail) {
    const uint64_t bytes = (uint64_t)kFxDetailSize * kFxDetailSize * kFxDetailSize * 4;
    std::vector<uint8_t> host(bytes);
    periodic_detail_volume(kFxDetailSize, host.data());
    rc1 = alloc(s, (void**)&s->detail, bytes, "detail volume");
    if (rc1) return rc1;
    s->bytes += bytes;
    FX_HIP(s, hipMemcpyAsync(s->detail, host.data(), bytes, hipMemcpyHostToDevice, s->stream), "detail upload");
    FX_HIP(s, hipStreamSynchronize(s->stream), "detail upload");
  }
  const uint64_t image_bytes = (uint64_t)rc->width * (uint64_t)rc->height * 8;
  if (image_bytes > s->scene_bytes) {
    if (s->cfg.memory_budget_bytes && s->bytes + image_bytes - s->scene_bytes > s->cfg.memory_budget_bytes)
      return set_error(FIREX_ERR_BUDGET, "the scene image (" + std::to_string(image_bytes) +
                                             " bytes) does not fit the memory budget");
    void* p = nullptr;
    rc1 = alloc(s, &p, image_bytes, "scene image");
    if (rc1) return rc1;
    if (s->scene) {
      FX_HIP(s, hipStreamSynchronize(s->stream), "scene image resize");
      s->allocations.erase(std::find(s->allocations.begin(), s->allocations.end(), s->scene));
      (void)hipFree(s->scene);  // no longer referenced: a failed free only leaks it
    }
    s->bytes += image_bytes - s->scene_bytes;
    s->scene = p;
    s->scene_bytes = image_bytes;
  }
  s->scene_width = rc->width;
  s->scene_height = rc->height;
  s->has_frame = false;
  float u[FXU_COUNT];
  write_uniforms(s->params, s->dims, 0.0, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                 s->cfg.cinematic != 0, u, s->cfg.particle_count, rc->width, rc->height, steps);
  const FxRender r = make_render(u, s->dims, rc->width, rc->height);
  s->last_render = r;
  if (!s->bricks) {
    const uint64_t bytes = (uint64_t)r.bnx * r.bny * r.bnz * 4;
    rc1 = alloc(s, (void**)&s->bricks, bytes, "occupancy bricks");
    if (rc1) return rc1;
    s->bytes += bytes;
  }
  // the beauty view of the cinematic path skips empty space
  const bool skipping = s->cfg.cinematic && (uint32_t)fxm_clamp(fxm_rint(r.view_mode), 0.0f, 7.0f) == 0u;
  timed(s, FIREX_PASS_RENDER, [&] {
    if (skipping) launch_render_bricks(s->packed, s->bricks, r, s->stream);
    launch_render_volume(s->packed, s->detail, s->bricks, s->scene, r, s->stream);
  });
  rc1 = check_launch(s, "volume render");
  if (rc1) return rc1;
  // the cinematic scene pass draws the droplets over the volume; the
  // standard path draws them into the canvas (firex_render_frame)
  return s->cfg.cinematic ? draw_streaks(s, r, s->scene, 0) : FIREX_OK;
}

int firex_render_frame(firex_solver* s, const firex_render_config* rc) {
  int r0 = firex_render_volume(s, rc);
  if (r0) return r0;
  const uint64_t need = post_bytes_for(rc->width, rc->height);
  if (need > s->post_bytes) {
    if (s->cfg.memory_budget_bytes && s->bytes + need - s->post_bytes > s->cfg.memory_budget_bytes)
      return set_error(FIREX_ERR_BUDGET, "the presentation buffers (" + std::to_string(need) +
                                             " bytes) do not fit the memory budget");
    void* p = nullptr;
    r0 = alloc(s, &p, need, "presentation buffers");
    if (r0) return r0;
    if (s->post) {
      FX_HIP(s, hipStreamSynchronize(s->stream), "presentation resize");
      s->allocations.erase(std::find(s->allocations.begin(), s->allocations.end(), s->post));
      (void)hipFree(s->post);  // no longer referenced: a failed free only leaks it
    }
    s->bytes += need - s->post_bytes;
    s->post = p;
    s->post_bytes = need;
  }
  uint8_t* base = (uint8_t*)s->post;
  void* bloom = base + (post_frame_bytes(rc->width, rc->height) + 255) / 256 * 256;
  const float exposure = (float)s->params.exposure;
  timed(s, FIREX_PASS_PRESENT, [&] {
    launch_render_post(s->scene, bloom, rc->width, rc->height, exposure, s->cfg.cinematic != 0 ? 1 : 0,
                       (uint32_t*)base, s->stream);
  });
  r0 = check_launch(s, "presentation");
  if (r0) return r0;
  if (!s->cfg.cinematic) {
    r0 = draw_streaks(s, s->last_render, base, 1);
    if (r0) return r0;
  }
  s->has_frame = true;
  return FIREX_OK;
}

// Enqueues reduction `which` on the solver's stream: block partials into
// `partial` (at most kReducePartials doubles), the FP64 sum into *result.
static int enqueue_reduction(firex_solver* s, int32_t which, double* partial, double* result) {
  if (which < FIREX_REDUCE_SUM_FUEL || which > FIREX_REDUCE_FLOW_DIVERGENCE_L2)
    return set_error(FIREX_ERR_INVALID, "unknown reduction");
  if (which == FIREX_REDUCE_RESIDUAL_L2 && !s->divergence_is_rhs)
    return set_error(FIREX_ERR_INVALID,
                     "the residual's right-hand side is gone: the divergence view (firex_pack_fields with viewMode 7) "
                     "replaced the divergence field with the final flow's divergence; step again first");
  float u[FXU_COUNT];
  write_uniforms(s->params, s->dims, 0.0, s->simulation_time, (double)s->frame_index, s->cfg.time_step,
                 s->cfg.cinematic != 0, u, s->cfg.particle_count);
  const FxStep st = make_step(u, s->dims, s->emitters, s->frame_index, s->has_liquid);
  switch (which) {
    case FIREX_REDUCE_SUM_FUEL: case FIREX_REDUCE_SUM_OXYGEN: case FIREX_REDUCE_SUM_PRODUCTS: case FIREX_REDUCE_SUM_VAPOR:
      launch_reduce_field(s->cfg.species_storage, 0, which - FIREX_REDUCE_SUM_FUEL, s->species[s->species_index], s->cells,
                          partial, result, s->stream);
      break;
    case FIREX_REDUCE_SUM_TEMPERATURE:
      launch_reduce_field(s->cfg.species_storage, 1, 3, flow0(s), s->cells, partial, result, s->stream);
      break;
    case FIREX_REDUCE_SUM_REACTION:
      launch_reduce_field(s->cfg.species_storage, 2, 0, s->reaction, s->cells, partial, result, s->stream);
      break;
    case FIREX_REDUCE_RESIDUAL_L2: {
      void* p;
      correction_buffer(s, &p);
      launch_reduce_residual(s->f16p ? 1 : 0, p, s->divergence, st, partial, result, s->stream);
      break;
    }
    default:  // FIREX_REDUCE_FLOW_DIVERGENCE_L2
      launch_reduce_flow_divergence(flow0(s), st, partial, result, s->stream);
      break;
  }
  return check_launch(s, "reduction");
}

int firex_reduce(firex_solver* s, int32_t which, double* out) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!out) return set_error(FIREX_ERR_INVALID, "out is null");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  double* result = s->scratch + kReducePartials;
  rc = enqueue_reduction(s, which, s->scratch, result);
  if (rc) return rc;
  double v = 0;
  FX_HIP(s, hipMemcpyAsync(&v, result, sizeof v, hipMemcpyDeviceToHost, s->stream), "reduction readback");
  FX_HIP(s, hipStreamSynchronize(s->stream), "reduction");
  if (which == FIREX_REDUCE_RESIDUAL_L2 || which == FIREX_REDUCE_FLOW_DIVERGENCE_L2) v = sqrt(v);
  *out = v;
  return FIREX_OK;
}

// ------------------------------------------------ allocation-free reduction
// firex_reduce_noalloc's scratch: the block partials of an FP64 reduction (at
// most kReducePartials doubles, 8 KiB), a file-scope device array that each
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol): its address on each device comes from hipGetSymbolAddress,
// once per device, and that pointer is what the kernels receive.
namespace fx_noalloc {
__device__ __align__(256) double scratch[kReducePartials];
constexpr int kMaxDevices = 64;
std::mutex lock;
double* device_scratch[kMaxDevices] = {};
}  // namespace fx_noalloc

// The scratch array's address on device `dev`, which must be current.
static int noalloc_scratch(int dev, double** out) {
  if (dev < 0 || dev >= fx_noalloc::kMaxDevices)
    return set_error(FIREX_ERR_INVALID, "device index beyond the allocation-free scratch table");
  std::lock_guard<std::mutex> hold(fx_noalloc::lock);
  if (!fx_noalloc::device_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(fx_noalloc::scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(FIREX_ERR_DEVICE, std::string("hipGetSymbolAddress of the reduction scratch failed: ") +
                                             hipGetErrorString(e));
    }
    fx_noalloc::device_scratch[dev] = static_cast<double*>(p);
  }
  *out = fx_noalloc::device_scratch[dev];
  return FIREX_OK;
}

int firex_reduce_noalloc(firex_solver* s, int32_t which, double* device_out) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!device_out) return set_error(FIREX_ERR_INVALID, "device_out is null");
  if (reduce_block_count(s->cells) > kReducePartials)
    return set_error(FIREX_ERR_INVALID, "grid of " + std::to_string(s->cells) + " cells needs more than " +
                                            std::to_string(kReducePartials) + " reduction partials");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  double* partial = nullptr;
  rc = noalloc_scratch(s->cfg.device, &partial);
  if (rc) return rc;
  rc = enqueue_reduction(s, which, partial, device_out);
  if (rc) return rc;
  if (which == FIREX_REDUCE_RESIDUAL_L2 || which == FIREX_REDUCE_FLOW_DIVERGENCE_L2)
    launch_sqrt_f64(device_out, s->stream);
  return check_launch(s, "allocation-free reduction");
}

int firex_set_pass_timing(firex_solver* s, int32_t enabled) {
  int rc = check_usable(s);
  if (rc) return rc;
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  rc = collect_timing(s);
  if (rc) return rc;
  memset(&s->totals, 0, sizeof s->totals);
  s->timing = enabled != 0;
  return FIREX_OK;
}

int firex_get_pass_timing(firex_solver* s, firex_pass_timing* out) {
  int rc = check_usable(s);
  if (rc) return rc;
  if (!out) return set_error(FIREX_ERR_INVALID, "out is null");
  FX_HIP(s, hipSetDevice(s->cfg.device), "hipSetDevice");
  rc = collect_timing(s);
  if (rc) return rc;
  *out = s->totals;
  return FIREX_OK;
}

int firex_get_info(firex_solver* s, firex_info* info) {
  if (!s || !info) return set_error(FIREX_ERR_INVALID, "solver and info must be non-null");
  memset(info, 0, sizeof *info);
  memcpy(info->dims, s->dims, sizeof info->dims);
  info->cell_count = s->cells;
  info->simulation_time = s->simulation_time;
  info->frame_index = (int64_t)s->frame_index;
  info->species_index = s->species_index;
  info->pressure_index = s->pressure_index;
  info->heavy = s->heavy;
  info->f16_pressure = s->f16p;
  info->species_storage = s->cfg.species_storage;
  info->plan = s->cfg.plan;
  info->device_bytes = s->bytes;
  info->jacobi_kernel = s->jacobi.kernel;
  info->jacobi_sweeps_per_launch = s->jacobi.sweeps_per_launch;
  info->jacobi_chunk_z = s->jacobi.chunk_z;
  info->jacobi_prefetch = s->jacobi.prefetch;
  info->jacobi_tuned_us_per_sweep = s->jacobi_tuned_us;
  info->jacobi_tile_height = s->jacobi.rows * s->jacobi.waves;
  info->jacobi_ring_waves = s->jacobi.waves;
  info->particle_count = s->cfg.particle_count;
  return FIREX_OK;
}

}  // extern "C"
