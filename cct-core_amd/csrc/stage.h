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

// Cache policy of the loads that stage input fields in the BF16 / FP16 frame
// kernels and the FP32 FilterAdjoint: non-temporal (nt), so that the streamed
// field does not push out of the L2 what every block re-reads (the packed
// filter bank: c4 BF16 7x7 92 -> 83 us).  The FP32 frame kernels (MFMA-bound)
// keep cached loads: with nt their halo re-reads go to HBM (c4 9x9: 204 ->
// 326 MB a launch) for no gain in time.
constexpr int LOAD_NT = 2;

__device__ __forceinline__ float buffer_load_f32_nt(__amdgpu_buffer_rsrc_t r, int vbytes, int sbytes) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vbytes, sbytes, LOAD_NT));
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

// Out-of-range lane byte offset: bit 31 set puts it past every staged buffer
// (each under 2^31 bytes), so the buffer unit's range check returns 0.
constexpr int STAGE_OOB = (int)0x80000000u;

// Byte offset of stored column x of a staged tile with the border applied, or
// STAGE_OOB where the column is off the field (BorderZero and the like).
__device__ __forceinline__ int stage_col_offset(const Geom& g, int x) {
    bool in = (unsigned)x < (unsigned)g.W;
    if (g.border == COGCONV_BORDER_CYCLIC) {
        x %= g.W;
        x += x < 0 ? g.W : 0;
        in = true;
    } else {
        in = in || g.border == COGCONV_BORDER_CLAMP;
        x = min(max(x, 0), g.W - 1);
    }
    return in ? 4 * x : STAGE_OOB;
}

// Row staging with the row data computed on the fly: row(u) gives the row's
// wave-uniform byte offset and whether it is in the field; dst(u) its LDS
// element offset (< 0: not live, store to `junk`).  vcol0 / vcol1 are the
// lanes' byte offsets of their columns (STAGE_OOB where the lane has no
// column).  Rows off the field load through the out-of-range offset, so the
// stores need no mask, and no per-row arrays are held across the loads (they
// would spill scalar registers).
template <bool TWO, int IU, class RowFn, class DstFn>
__device__ __forceinline__ void stage_rows_fn(__amdgpu_buffer_rsrc_t src, RowFn row, DstFn dst, float* lds, float* junk,
                                              int lane, int width, int pitch, int vcol0, int vcol1) {
    float v0[IU], v1[IU];
#pragma unroll
    for (int u = 0; u < IU; u++) {
        int soff;
        bool on;
        row(u, soff, on);
        v0[u] = buffer_load_f32(src, on ? vcol0 : STAGE_OOB, soff);
        if (TWO) v1[u] = buffer_load_f32(src, on ? vcol1 : STAGE_OOB, soff);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
#pragma unroll
    for (int u = 0; u < IU; u++) {
        const int d = dst(u);
        const bool live = d >= 0;  // wave-uniform
        float* d0 = (live && lane < width) ? lds + d + lane * pitch : junk + lane;
        *d0 = v0[u];
        if (TWO) {
            float* d1 = (live && lane + 64 < width) ? lds + d + (lane + 64) * pitch : junk + 64 + lane;
            *d1 = v1[u];
        }
    }
}

// The two halves of stage_rows_fn, for callers that put the loads of several
// row sets in flight before the first store: load_rows_fn issues the loads of
// IU rows (row(u) as above) into v0 / v1, store_rows_fn stores them (dst(u)
// as above).  A __builtin_amdgcn_sched_barrier(0) between the last load and
// the first store keeps the loads together.
//
// The second pass's load of a row comes first, so the one- and two-pass forms
// both end with a load into v0: a caller that picks the form at run time then
// joins on one array element (a join on v0 / v1 elements would keep the
// arrays in scratch memory).
template <bool TWO, int IU, class RowFn>
__device__ __forceinline__ void load_rows_fn(__amdgpu_buffer_rsrc_t src, RowFn row, int vcol0, int vcol1,
                                             float (&v0)[IU], float (&v1)[IU]) {
#pragma unroll
    for (int u = 0; u < IU; u++) {
        int soff;
        bool on;
        row(u, soff, on);
        if (TWO) v1[u] = buffer_load_f32_nt(src, on ? vcol1 : STAGE_OOB, soff);
        v0[u] = buffer_load_f32_nt(src, on ? vcol0 : STAGE_OOB, soff);
    }
}

template <bool TWO, int IU, class DstFn>
__device__ __forceinline__ void store_rows_fn(DstFn dst, float* lds, float* junk, int lane, int width, int pitch,
                                              const float (&v0)[IU], const float (&v1)[IU]) {
#pragma unroll
    for (int u = 0; u < IU; u++) {
        const int d = dst(u);
        const bool live = d >= 0;  // wave-uniform
        float* d0 = (live && lane < width) ? lds + d + lane * pitch : junk + lane;
        *d0 = v0[u];
        if (TWO) {
            float* d1 = (live && lane + 64 < width) ? lds + d + (lane + 64) * pitch : junk + 64 + lane;
            *d1 = v1[u];
        }
    }
}

}  // namespace cogconv
