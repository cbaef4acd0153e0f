// Trace kernels with the smoothed normals of brick sets with neighbour indices compiled in:
// leaf kinds LK_SBS and the mixed set, each with LK_SBS_SMOOTH.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_sbs_smooth(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_SBS | LK_SBS_SMOOTH>(s, a, stream);
}

hipError_t launch_trace_mixed_smooth(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_TRI | LK_FRAME | LK_SBS | LK_MULTI | LK_GRID | LK_GPRIM | LK_COCT | LK_DAG | LK_SCOM2 | LK_FRAME_TEX | LK_SBS_SMOOTH>(s, a, stream);
}

}  // namespace lrt
