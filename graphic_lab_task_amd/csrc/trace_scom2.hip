// Trace kernels instantiated for leaf kinds LK_SCOM2.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_scom2(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_SCOM2>(s, a, stream);
}

}  // namespace lrt
