// Device-tuned parameter table (C ABI: bg_tune_load, bg_tune_lookup).
//
// Replaces the reference's fixed heuristics Msm.pick_c / Msm.pick_fd
// (bend2/src/msm.bend:37-44) and Ntt.fd (bend2/src/ntt.bend:382-385) with values measured
// on the device by `bench.py --mode tune` and stored in a text table next to the library
// (tune_gfx950.txt).  The planners ask the table first; a size the table does not hold
// falls back to the cost model.  Environment overrides (BG16_MSM_C_G1 ...) still win.
//
// Table lines (whitespace separated, '#' starts a comment):
//   msm <group 1|2> <windowed|fixed> <log2 of the batch's total points, in half steps (18, 18.5,
//       ...); a batch matches the entry nearest its log2> c=<bits> L=<entries per accumulation
//       thread, target> m=<buckets per level-1 group> K=<level-2 block>
//   ntt <log2 n> tile=<log2 elements per workgroup, 9..11> radix=<2|4>
// Any key=value may be left out (0 = the planner's own choice).
#pragma once
#include <stdint.h>

namespace bg {

struct MsmTune {
  int c = 0, L = 0, m = 0, K = 0;
};
struct NttTune {
  int tile_log2 = 0, radix = 0;
};

// Entry for a batch of MSMs of this group and kind with total_n points; false when the
// table has none (or no table is loaded).
bool tune_msm(int group, bool fixed, uint64_t total_n, MsmTune& out);
bool tune_ntt(int log2n, NttTune& out);

// Incremented by every bg_tune_load: launch sequences recorded for reuse (the prover's
// graphs) are only valid for the table they were planned with.
uint64_t tune_generation();

}  // namespace bg
