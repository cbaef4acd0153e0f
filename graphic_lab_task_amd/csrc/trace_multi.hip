// Trace kernels instantiated for leaf kinds LK_MULTI.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_multi(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_MULTI>(s, a, stream);
}

}  // namespace lrt
