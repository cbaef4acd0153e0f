// Trace kernels instantiated for leaf kinds LK_TRI | LK_TRI_POSTPONE.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_tri_postponed(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_TRI | LK_TRI_POSTPONE>(s, a, stream);
}

}  // namespace lrt
