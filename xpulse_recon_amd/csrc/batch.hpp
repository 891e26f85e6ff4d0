// Batched steps of several reconstructions (library-internal interface
// between the HIP kernels and the job server).
//
// One step is a table of work items grouped by phase.  Per phase, one
// persistent kernel walks every item's work units (forward bricks, backward
// row blocks, elementwise chunks) through an atomic counter, so a step runs
// the sub-iteration pieces of many jobs of different sizes in nine launches.
// The launch sequence of a step (table copy, counter reset, convert, the
// three forwards, error, the three backwards, fill) is captured once into a
// HIP graph per table slot and replayed; two slots let the host write step
// k+1 while step k runs.
#pragma once

#include <cstdint>

#include "../../include/xpr_hip.h"

namespace xpr_batch {

// The forward and the backward have one phase per volume storage type (F32,
// F16, BF16), each walked by its own kernel: one kernel holding the three
// paths would need the registers of all three and spill.
enum Phase : uint32_t {
    kConvert = 0, kForward = 1, kError = 2, kBackward = 3, kFill = 4,
    kForwardF16 = 5, kForwardBF16 = 6, kBackwardF16 = 7, kBackwardBF16 = 8
};
constexpr int kPhases = 9;
inline Phase forward_phase(int32_t dt)
{
    return dt == XPR_F16 ? kForwardF16 : dt == XPR_BF16 ? kForwardBF16 : kForward;
}
inline Phase backward_phase(int32_t dt)
{
    return dt == XPR_F16 ? kBackwardF16 : dt == XPR_BF16 ? kBackwardBF16 : kBackward;
}
constexpr int kMaxItems = 32;              // items per phase and step
constexpr uint32_t kElemUnit = 1u << 16;   // elements per elementwise work unit

struct Item {
    const void *src;         // convert: source volume; error: measured images
    void *dst;               // convert: destination; forward/backward: volume slab;
                             // error: error samples; fill: sums
    void *sums;              // forward, error: u64 Q.22 detector sums
    const void *err;         // backward: error samples
    const float *pd;         // forward, backward: ProjData of the sub-iteration
    xpr_proj_args a;         // forward, backward (slab in vmin/vmax; line_offset 0)
    uint64_t n;              // convert, error, fill: elements
    uint64_t value;          // fill: value
    float weight;            // error: sub-iteration weight
    int32_t dt_src;          // convert: source dtype; error: image dtype
    int32_t dt_dst;          // convert: destination; forward/backward: volume; error: error
    int32_t dt_aux;          // backward: error dtype
    uint32_t first, units;   // work units of the item (prefix within the phase, count)
};

struct Table {
    uint32_t count[kPhases];   // items per phase
    uint32_t total[kPhases];   // work units per phase
    Item items[kPhases][kMaxItems];
};

// work units of an item
uint32_t forward_units(const xpr_proj_args &a);
uint32_t backward_units(const xpr_proj_args &a);
inline uint32_t elementwise_units(uint64_t n) { return (uint32_t)((n + kElemUnit - 1) / kElemUnit); }

struct Engine;
// Creates the engine on the calling thread's device: stream, two pinned host
// tables, two device tables and counters, two captured graphs.
int engine_create(Engine **e);
Table *engine_table(Engine *e, int slot);       // pinned host table of a slot
// Replays the slot's graph (it copies the slot's host table first) and
// records the slot's events around it.
int engine_launch(Engine *e, int slot);
// Waits for the slot's last replay; *ms receives its device time.
int engine_wait(Engine *e, int slot, float *ms);
// Waits for everything enqueued on the engine's stream (both slots), whatever
// its outcome: after an error, no replay still uses a job's buffers.
void engine_sync(Engine *e);
void engine_destroy(Engine *e);

} // namespace xpr_batch
