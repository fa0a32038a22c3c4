// Frame timing with HIP events for the autotuners (host code of the .hip
// translation units).
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace yx {

// Mean time of `reps` frames after one untimed frame, on `stream`.
template <class Frame>
inline int time_frames(Frame frame, int reps, hipStream_t stream, float* ms) {
    int rc = frame();
    if (rc != YX_OK) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return fail(YX_ERR_HIP, "hipEventCreate");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(YX_ERR_HIP, "hipEventCreate");
    }
    hipError_t he = hipEventRecord(e0, stream);
    for (int i = 0; i < reps && rc == YX_OK && he == hipSuccess; ++i) rc = frame();
    if (he == hipSuccess) he = hipEventRecord(e1, stream);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float total = 0.0f;
    if (he == hipSuccess) he = hipEventElapsedTime(&total, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != YX_OK) return rc;
    if (he != hipSuccess) return fail(YX_ERR_HIP, "timing: %s", hipGetErrorString(he));
    *ms = total / (float)reps;
    return YX_OK;
}

}  // namespace yx
