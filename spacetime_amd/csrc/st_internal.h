// Internal declarations shared by the kernels and the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>

namespace st {

struct DeviceScene;
struct TraceParams;
struct FrameRef;

hipError_t launch_tiles(int fp64, bool kerr, bool vol, bool route, const DeviceScene& D, const TraceParams& P,
                        const FrameRef& F, int col0, int row0, int cols, int rows, int tw, hipStream_t s);
hipError_t launch_queue64(bool kerr, bool vol, const DeviceScene& D, const TraceParams& P, const FrameRef& F0,
                          const FrameRef* frames_dev, int nblocks, hipStream_t s);
hipError_t launch_reset(unsigned* qcount, unsigned long long* counters, unsigned* work_next, hipStream_t s);
hipError_t launch_batch(bool kerr, bool vol, bool route, const DeviceScene& D, const TraceParams& P,
                        const FrameRef* frames_dev, const int* tile_base_dev, int nframes, int total_items, int tw,
                        unsigned* work_next, int nblocks, hipStream_t s);

}  // namespace st
