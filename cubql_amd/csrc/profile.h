// profile.h -- optional per-kernel timing with HIP events.
//
// Counterpart of the reference's compile-time CUBQL_PROFILE timers
// (builder/cuda/sm_builder.h:24-66), but event based: when enabled, host
// code brackets selected launches with events on the launch stream and, after
// its own synchronisation point, accumulates elapsed time and algorithmic
// bytes per kernel name.  Off by default; costs nothing when off.
#pragma once
#include "common.h"

namespace cubql {
bool prof_on();
// Records a start event on s; returns a slot (or -1 when profiling is off).
int prof_begin(hipStream_t s);
// Records the end event for `slot` and names the kernel.
void prof_end(int slot, hipStream_t s, const char *name);
// Sets the algorithmic bytes and work units (prims, queries...) of a slot.
void prof_set(int slot, double bytes, uint64_t units);
// After the stream has been synchronised: accumulate all pending slots of
// this thread (slots with zero units are dropped as empty launches).
void prof_flush();
}  // namespace cubql
