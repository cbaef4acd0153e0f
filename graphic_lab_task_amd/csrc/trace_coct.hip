// Trace kernels instantiated for leaf kinds LK_COCT.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_coct(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_COCT>(s, a, stream);
}

}  // namespace lrt
