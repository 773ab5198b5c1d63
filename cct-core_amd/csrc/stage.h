// Device-side staging helper shared by the MFMA kernels (included by the .hip
// sources only).
#pragma once

#include "internal.h"

namespace cogconv {

// Stages IU rows of up to two 64-lane passes into LDS: every load first (from
// in-bounds offsets off[u] + col, with no branch between the loads), then the
// masked stores.  Row u goes to lds + dst[u] + lane*pitch (second pass: lane
// + 64); a row that is not live (dst[u] < 0) or a lane past `width` stores
// into the 128-slot scratch `junk` instead of branching.
template <bool TWO, int IU>
__device__ __forceinline__ void stage_rows(const float* __restrict__ src, const int (&off)[IU], const bool (&on)[IU],
                                           const int (&dst)[IU], float* lds, float* junk, int lane, int width,
                                           int pitch, int col0, int col1, bool in0, bool in1) {
    float v0[IU], v1[IU];
#pragma unroll
    for (int u = 0; u < IU; u++) {
        v0[u] = src[off[u] + col0];
        if (TWO) v1[u] = src[off[u] + col1];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
#pragma unroll
    for (int u = 0; u < IU; u++) {
        const bool live = dst[u] >= 0;  // wave-uniform
        float* d0 = (live && lane < width) ? lds + dst[u] + lane * pitch : junk + lane;
        *d0 = (on[u] && in0) ? v0[u] : 0.f;
        if (TWO) {
            float* d1 = (live && lane + 64 < width) ? lds + dst[u] + (lane + 64) * pitch : junk + 64 + lane;
            *d1 = (on[u] && in1) ? v1[u] : 0.f;
        }
    }
}

}  // namespace cogconv
