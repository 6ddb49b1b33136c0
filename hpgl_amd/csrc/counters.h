// counters.h - statistics counters written by the device kernels.
//
// The counter buffer holds kCounterSlots lines of kCounterStride 64-bit
// words (one 64-byte line each).  A wavefront counts its lanes with a
// ballot and adds the total to the line of its workgroup index, so the
// adds of concurrent wavefronts are spread over many addresses instead of
// serialising on one; the host sums the lines.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace hpgl_amd {

__device__ inline unsigned long long *counter_line(unsigned long long *c) {
    return c + (size_t)(blockIdx.x & (kCounterSlots - 1)) * kCounterStride;
}

// Adds the number of lanes of the wavefront with `pred` to counter `which`.
// Every lane of the wavefront must call it.
__device__ inline void wave_count(unsigned long long *c, int which, bool pred) {
    const unsigned long long n = __popcll(__ballot(pred));
    if (c && n != 0 && (threadIdx.x & 63) == 0) atomicAdd(counter_line(c) + which, n);
}

}  // namespace hpgl_amd
