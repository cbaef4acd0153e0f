// Launch-configuration autotuner for the trace kernels. Every configuration returns identical
// hit records, so tuning on the caller's own rays is safe: it only changes the time.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

namespace lrt {

struct TuneConfig {
  int block, lds_stack, xcd_remap;
  uint32_t budget, tail_waves;
  int postpone;          // triangle leaf postponement (0/1)
  uint32_t ray_groups;   // pass-1 ray interleaving: lanes per group (0: off)
  int persistent;        // 1: persistent single-pass kernel with per-lane ray refill
};

struct TuneResult {
  TuneConfig best;
  float best_ms;
  uint32_t tried;
};

// Two stages, each candidate timed with HIP events on `stream` (one warm-up launch, then the
// minimum of `reps`): (1) block size x LDS stack entries x ray interleaving (off, groups of 8
// or 32 lanes) with the single-pass walk, and block size x LDS stack entries as the persistent
// kernel, (2) step budget x tail waves of the two-pass scheme on the three best grid shapes from (1); both stages also try triangle leaf postponement off and on when
// `tune_postpone`; (3) the four fastest distinct candidates timed again in turns over seven
// rounds, the lowest median time winning (best_ms is that median).
hipError_t autotune(const std::function<hipError_t(const TuneConfig&)>& launch, hipStream_t stream,
                    const TuneConfig& start, bool tune_postpone, int reps, TuneResult* out);

}  // namespace lrt
