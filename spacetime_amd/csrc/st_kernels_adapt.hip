// Adaptive Tsit5 (ORD 45/46) instantiations of the frame and FP64 queue
// kernels, compiled in their own translation unit.
#include <hip/hip_runtime.h>

#include "st_kernels.h"

namespace st {

hipError_t launch_tiles_adapt(int fp64, bool kerr, bool vol, bool route, const DeviceScene* D, const TraceParams& P,
                              const FrameRef& F, int col0, int row0, int cols, int rows, int tw, hipStream_t s) {
    return launch_tiles_impl<true>(fp64, kerr, vol, route, D, P, F, col0, row0, cols, rows, tw, s);
}

hipError_t launch_queue64_adapt(bool kerr, bool vol, const DeviceScene* D, const TraceParams& P, const FrameRef& F0,
                                const FrameRef* frames_dev, size_t capacity, hipStream_t s) {
    return launch_queue64_impl<true>(kerr, vol, D, P, F0, frames_dev, capacity, s);
}

}  // namespace st
