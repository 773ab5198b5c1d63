// Device-side staging helper shared by the MFMA kernels (included by the .hip
// sources only).
#pragma once

#include "internal.h"

namespace cogconv {

// Raw (unswizzled) buffer resource over `bytes` bytes at `base`.  Word 3 =
// 0x00020000 selects the 32-bit data format that untyped buffer_load_dword /
// _dwordx4 use.  Loads through it take a per-lane byte offset (VGPR) and a
// wave-uniform byte offset (SGPR), so a staging row costs no vector address
// arithmetic; offsets at or past `bytes` read 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float buffer_load_f32(__amdgpu_buffer_rsrc_t r, int vbytes, int sbytes) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vbytes, sbytes, 0));
}

// Stages IU rows of up to two 64-lane passes into LDS: every load first (from
// in-bounds element offsets off[u] + col of buffer `src`, with no branch
// between the loads), then the masked stores.  Row u goes to lds + dst[u] + lane*pitch (second pass: lane
// + 64); a row that is not live (dst[u] < 0) or a lane past `width` stores
// into the 128-slot scratch `junk` instead of branching.
template <bool TWO, int IU>
__device__ __forceinline__ void stage_rows(__amdgpu_buffer_rsrc_t src, const int (&off)[IU], const bool (&on)[IU],
                                           const int (&dst)[IU], float* lds, float* junk, int lane, int width,
                                           int pitch, int col0, int col1, bool in0, bool in1) {
    float v0[IU], v1[IU];
#pragma unroll
    for (int u = 0; u < IU; u++) {
        v0[u] = buffer_load_f32(src, 4 * col0, 4 * off[u]);
        if (TWO) v1[u] = buffer_load_f32(src, 4 * col1, 4 * off[u]);
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
