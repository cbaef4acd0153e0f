// Trace kernels instantiated for leaf kinds LK_FRAME.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_frame(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_FRAME>(s, a, stream);
}

}  // namespace lrt
