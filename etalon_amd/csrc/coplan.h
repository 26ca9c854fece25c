// What joint launches of a forward plan and an adjoint plan (coplan.hip) need from
// each: their prepared state and tile lists.  Internal to libetalon_hip.so.
#pragma once

#include "common.h"
#include "fused.h"

struct EtlPlan;
struct EtlAdjPlan;

namespace etl {

struct FwdInfo {
    int precision, temporal, trace_chunk, launch, lane, vec, band;
    const void* args;  // FusedArgs<R>[n_sims] on the device
    TileList tl;
    int* counter;
    int parity;
    unsigned gen;  // layout generation
};

struct AdjInfo {
    int precision, temporal, launch, band, lane;
    const void* args;  // AdjTileArgs<R>[n_sims] on the device
    TileList tl;
    int* step;
    int parity;
    unsigned gen;  // layout generation
};

// Bring a forward plan up to date for steps [n, n + count) (layout, tables, argument
// arrays, probe rows from probe_row0) and set its step counter on its lane — or, with
// counter_out, return the three counter values to set instead.
int fwd_prepare(EtlPlan* p, int64_t n, int64_t count, int probe_row0,
                int* counter_out = nullptr);
void fwd_info(const EtlPlan* p, FwdInfo* out);
void fwd_set_parity(EtlPlan* p, int parity);
// After a trace chunk of `rows` steps: probe copies, then the DFT fold (or a counter
// bump), on stream s.
int fwd_after_chunk(EtlPlan* p, int rows, hipStream_t s);
// Save (restore = 0) or load (1) checkpoint slot `seg` of every simulation, on the
// plan's lane, into/from the buffers of the current parity.
int fwd_checkpoint(EtlPlan* p, int seg, int restore);

// Bring an adjoint plan up to date for reverse steps n_first + count - 1 .. n_first
// (history rows from hist_row0) and set its step record on its lane — or, with
// step_out, return the two values to set instead.
int adj_prepare(EtlAdjPlan* p, int64_t n_first, int count, int hist_row0,
                int* step_out = nullptr);
void adj_info(const EtlAdjPlan* p, AdjInfo* out);
void adj_set_parity(EtlAdjPlan* p, int parity);

}  // namespace etl
