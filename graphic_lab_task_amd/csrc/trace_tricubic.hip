// Trace kernels with the preset's tricubic interpolation compiled in (brick sets and dense grids):
// leaf kinds LK_SBS and the mixed set, each with LK_TRICUBIC.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_sbs_tricubic(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_SBS | LK_TRICUBIC>(s, a, stream);
}

hipError_t launch_trace_mixed_tricubic(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_TRI | LK_FRAME | LK_SBS | LK_MULTI | LK_GRID | LK_GPRIM | LK_COCT | LK_DAG | LK_SCOM2 | LK_FRAME_TEX | LK_TRICUBIC>(s, a, stream);
}

}  // namespace lrt
