// Trace kernels instantiated for leaf kinds LK_TRI.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_tri(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_TRI>(s, a, stream);
}

}  // namespace lrt
