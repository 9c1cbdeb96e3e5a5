// Shared helpers of the MI355X plugin model: the tile -> clover_mi355x chunk
// binding and the profiler synchronisation the reference's GPU models do.
#pragma once

#include "clover_mi355x.h"
#include "definitions.h"

namespace mi355x {

// The chunk over tile `tile`'s buffers, wrapped on first use (clv_chunk_wrap).
// One chunk and one tile only: every halo is the external reflective halo.
clv_chunk *chunk(global_variables &globals, int tile);
// When the reference profiler is on, wait for the tile's queued kernels so the
// driver's host timers measure them (as src/hip/update_halo.cpp:635-661 does).
void sync_if_profiling(global_variables &globals, int tile);
// Destroys the chunk bindings (finalise).
void release_chunks();

inline void check(clv_status st, const char *what) { clover::check_clv(st, what); }
[[noreturn]] void unsupported(const char *what);

}  // namespace mi355x
