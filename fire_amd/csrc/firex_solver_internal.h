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
hipStream_t solver_stream(firex_solver* s);
uint64_t solver_bytes(firex_solver* s);
uint32_t solver_cells(firex_solver* s);

}  // namespace fx
