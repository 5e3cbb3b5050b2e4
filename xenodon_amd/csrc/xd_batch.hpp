// Persistent ray-tile-queue kernel shared by the batched form (xd_batch.hip)
// and the served form (xd_serve.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xd_runtime.hpp"

namespace xd {

// One output of a launch: its kernel parameters and wave-tile geometry.
struct BatchView {
    RenderParams p;
    uint32_t traversal;
    uint32_t tw_shift;  // log2 of the tile width
    uint32_t tiles_x;
    uint32_t rpl;       // rays per lane: rows of wave tiles queued and grabbed together
};

constexpr uint32_t kQueues = 8;                 // one tile range per XCD, per traversal class
constexpr uint32_t kClasses = 5;                // one persistent kernel per traversal
constexpr uint32_t kMaxBatchViews = 1u << 12;   // a tile entry is (view << 20 | tile)
constexpr uint32_t kMaxTilesPerView = 1u << 20;

uint32_t tw_shift_of(uint32_t tile_w);

// The per-class tiles per grab the kernels take, packed 6 bits per class:
// tiles_per_grab x rays_per_lane[c] (null: 1), so that a grab holds whole
// groups of rays_per_lane rows.
uint32_t pack_grabs(uint32_t tiles_per_grab, const uint32_t rays_per_lane[kClasses]);

// Fills B for one output (frame pixels at out with row pitch `pitch`) and
// returns its number of wave tiles in *ntiles.  The LDS stack stride
// B.p.lds_slots is left at the view's own depth; the launch sets it.
void make_batch_view(const xd_volume* vol, int traversal, const xd_camera* cam, const xd_uniforms* u, uint32_t* out,
                     uint32_t pitch, uint32_t tile_w, uint32_t rays_per_lane, BatchView& B, uint64_t* ntiles);

// Stack slots a view needs in LDS (0 for traversals without a stack).
uint32_t batch_view_slots(const BatchView& B);

// Macro-cell bitmasks of at most this many words are staged in LDS by the
// DDA class kernel, one copy per wave (4 KiB).
constexpr uint32_t kWaveMacroWords = 1024;

// Dynamic LDS bytes of the persistent kernel of class `traversal` (kClasses:
// the all-class kernel) whose deepest stack has `slots` entries.
size_t batch_lds_bytes(int traversal, uint32_t slots);

// Resident blocks of traversal class `traversal`'s persistent kernel (or, for
// traversal == kClasses, of the all-class kernel) on the whole device (a
// multiple of kQueues) for this LDS size and register budget.
int batch_grid(int device, int traversal, uint32_t slots, bool w8, uint32_t* grid);

// Launch shape: entry c < kClasses is traversal c's own kernel, entry
// kClasses the kernel that drains every class's tiles.  grid[c] = 0 skips
// entry c; slots[c] is the LDS stack stride that every view drained by that
// kernel carries in p.lds_slots.
struct BatchClasses {
    uint32_t grid[kClasses + 1];
    uint32_t slots[kClasses + 1];
};

// Enqueues the counter reset and the present kernels of cls on s.  qbeg holds
// kClasses * kQueues + 1 range starts (class c's ranges are entries
// c*kQueues .. c*kQueues+kQueues); counters holds kClasses * kQueues.
// grabs: tiles per grab of each class (pack_grabs).  mix: the all-class
// kernel's waves start on different traversal classes.
hipError_t launch_batch(const BatchView* views, const uint32_t* tiles, const uint32_t* qbeg, uint32_t* counters,
                        uint32_t grabs, const BatchClasses& cls, bool w8, bool mix, hipStream_t s);

}  // namespace xd
