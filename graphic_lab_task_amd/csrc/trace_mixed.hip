// Trace kernels instantiated for leaf kinds LK_TRI | LK_FRAME | LK_SBS | LK_MULTI | LK_GRID | LK_GPRIM | LK_COCT | LK_DAG | LK_SCOM2 | LK_FRAME_TEX.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_mixed(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_TRI | LK_FRAME | LK_SBS | LK_MULTI | LK_GRID | LK_GPRIM | LK_COCT | LK_DAG | LK_SCOM2 | LK_FRAME_TEX>(s, a, stream);
}

}  // namespace lrt
