// Trace kernels instantiated for leaf kinds LK_DAG.
#include "trace_launch.h"

namespace lrt {

hipError_t launch_trace_dag(const SceneView& s, const TraceArgs& a, hipStream_t stream) {
  return launch_kinds<LK_DAG>(s, a, stream);
}

}  // namespace lrt
