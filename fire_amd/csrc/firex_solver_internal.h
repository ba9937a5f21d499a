// firex_solver_internal.h — solver hooks used by the serving layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/firex.h"

namespace fx {

// Validates a configuration and reports its device bytes (no allocation).
int solver_plan(const firex_config* cfg, bool* heavy, bool* f16p, bool* compact, uint64_t* bytes);
// Creates a plume issuing its work on `stream` (null: a stream of its own).
int solver_create(const firex_config* cfg, hipStream_t stream, firex_solver** out);
// Enqueues one step of dt seconds without waiting.
int solver_step_async(firex_solver* s, double dt);
// Key of the groups a plume can step in with solver_step_group (plumes on one
// stream and device, with f32 or f16 pressure alike, equal iteration counts,
// the FUSED plan and the 16-wave ring Jacobi plan of one depth); 0: the plume
// steps alone.
uint64_t solver_group_key(firex_solver* s);
constexpr int kStepGroupMax = 16;
// One step of dt seconds for each of `count` (<= kStepGroupMax) distinct plumes with
// equal nonzero group keys.  Simulate, divergence and projection run per
// plume; each Jacobi launch of the step is one launch over every plume's
// tiles.  Each plume's fields are those of its own step.
int solver_step_group(firex_solver* const* group, int32_t count, double dt);
hipStream_t solver_stream(firex_solver* s);
uint64_t solver_bytes(firex_solver* s);
uint32_t solver_cells(firex_solver* s);

}  // namespace fx
