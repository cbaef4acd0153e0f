#include "tune.h"

#include <vector>

namespace lrt {

namespace {

hipError_t time_one(const std::function<hipError_t(const TuneConfig&)>& launch, hipStream_t stream,
                    const TuneConfig& c, int reps, hipEvent_t e0, hipEvent_t e1, float* ms) {
  hipError_t e = launch(c);   // warm-up
  if (e != hipSuccess) return e;
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    if ((e = hipEventRecord(e0, stream)) != hipSuccess) return e;
    if ((e = launch(c)) != hipSuccess) return e;
    if ((e = hipEventRecord(e1, stream)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
    float t = 0.0f;
    if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return e;
    if (t < best) best = t;
  }
  *ms = best;
  return hipSuccess;
}

}  // namespace

hipError_t autotune(const std::function<hipError_t(const TuneConfig&)>& launch, hipStream_t stream,
                    const TuneConfig& start, bool tune_postpone, int reps, TuneResult* out) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  TuneResult r{start, 1e30f, 0};
  auto consider = [&](const TuneConfig& c) -> hipError_t {
    float ms = 0.0f;
    hipError_t err = time_one(launch, stream, c, reps, e0, e1, &ms);
    if (err != hipSuccess) return err;
    r.tried++;
    if (ms < r.best_ms) {
      r.best_ms = ms;
      r.best = c;
    }
    return hipSuccess;
  };
  const int n_postpone = tune_postpone ? 2 : 1;
  for (int pp = 0; pp < n_postpone && e == hipSuccess; ++pp) {
    for (int block : {64, 128, 256}) {
      for (int lds : {8, 16, 32}) {
        TuneConfig c = start;
        c.block = block;
        c.lds_stack = lds;
        c.budget = 0;
        if (tune_postpone) c.postpone = pp;
        if ((e = consider(c)) != hipSuccess) break;
      }
      if (e != hipSuccess) break;
    }
  }
  // the two-pass scheme changes which rays the postponed walk holds up, so stage 2 tries both
  // leaf orders again
  const TuneConfig shape = r.best;
  for (int pp = 0; pp < n_postpone && e == hipSuccess; ++pp) {
    for (uint32_t budget : {32u, 48u, 64u, 96u, 128u, 192u}) {
      for (uint32_t waves : {1024u, 2048u}) {
        TuneConfig c = shape;
        c.budget = budget;
        c.tail_waves = waves;
        if (tune_postpone) c.postpone = pp;
        if ((e = consider(c)) != hipSuccess) break;
      }
      if (e != hipSuccess) break;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e == hipSuccess) *out = r;
  return e;
}

}  // namespace lrt
