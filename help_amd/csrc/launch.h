// Kernel launch helper shared by every translation unit of the backend.
//
// A profiled region (ProfScope in backend.hip) arms a pair of events; the
// launches inside the region then carry them as launch events, so the first
// launch's start and the last launch's end timestamps bracket the region.
// That is the kernels' own execution span, as a kernel trace reports it,
// without the marker packets that separate hipEventRecord calls add.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace help {

struct RegionEvents {
    hipEvent_t start = nullptr;   // taken by the region's first launch
    hipEvent_t stop = nullptr;    // taken by every launch; the last one stands
    bool used = false;            // some launch took the events
};

inline thread_local RegionEvents g_region_events;

template <typename... P, typename... A>
inline void launch_kernel(void (*kern)(P...), dim3 grid, dim3 block, uint32_t shmem, hipStream_t st,
                          A... args) {
    RegionEvents& e = g_region_events;
    if (e.stop) {
        hipExtLaunchKernelGGL(kern, grid, block, shmem, st, e.used ? nullptr : e.start, e.stop, 0,
                              static_cast<P>(args)...);
        e.used = true;
    } else {
        kern<<<grid, block, shmem, st>>>(static_cast<P>(args)...);
    }
}

}  // namespace help
