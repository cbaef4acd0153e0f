// Trace kernels instantiated for leaf kinds LK_SBS.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_sbs(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_SBS>(s, a, stream);
}

}  // namespace lrt
