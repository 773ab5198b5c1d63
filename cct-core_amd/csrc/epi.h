// Elementwise epilogue applied to every output element at its store: the
// device side of cogconv_plan_set_epilogue (see Epi in internal.h).
//
// The ops run in order on the value in registers, each as one IEEE float
// operation (or one call of the device math library for exp / log / sqrt /
// tanh), with contraction off, so that a fused store gives bit for bit what
// the same ops give when they run as their own pass over the stored output
// (k_elementwise).
#pragma once

#include "internal.h"

namespace cogconv {

__device__ __forceinline__ float epi_operand(const Epi& e, int slot, int plane, int point) {
    const size_t at = (size_t)(e.srcPlanes[slot] == 1 ? 0 : plane) * e.srcPoints[slot] +
                      (e.srcPoints[slot] == 1 ? 0 : point);
    return e.src[slot][at];
}

__device__ __forceinline__ float epi_eval(const Epi& e, float v, int plane, int point) {
#pragma clang fp contract(off)
    for (int i = 0; i < e.n; i++) {
        const int op = e.op[i];
        float a = e.val[i];
        if (op >= COGCONV_EPI_ADD_FIELD) a = epi_operand(e, e.slot[i], plane, point);
        switch (op) {
        case COGCONV_EPI_ADD_CONST:
        case COGCONV_EPI_ADD_FIELD: v = v + a; break;
        case COGCONV_EPI_SUB_CONST:
        case COGCONV_EPI_SUB_FIELD: v = v - a; break;
        case COGCONV_EPI_RSUB_CONST:
        case COGCONV_EPI_RSUB_FIELD: v = a - v; break;
        case COGCONV_EPI_MUL_CONST:
        case COGCONV_EPI_MUL_FIELD: v = v * a; break;
        case COGCONV_EPI_DIV_CONST:
        case COGCONV_EPI_DIV_FIELD: v = v / a; break;
        case COGCONV_EPI_RDIV_CONST:
        case COGCONV_EPI_RDIV_FIELD: v = a / v; break;
        case COGCONV_EPI_MAX_CONST:
        case COGCONV_EPI_MAX_FIELD: v = fmaxf(v, a); break;
        case COGCONV_EPI_MIN_CONST:
        case COGCONV_EPI_MIN_FIELD: v = fminf(v, a); break;
        case COGCONV_EPI_NEG: v = -v; break;
        case COGCONV_EPI_ABS: v = fabsf(v); break;
        case COGCONV_EPI_SQ: v = v * v; break;
        case COGCONV_EPI_SQRT: v = sqrtf(v); break;
        case COGCONV_EPI_EXP: v = expf(v); break;
        case COGCONV_EPI_LOG: v = logf(v); break;
        case COGCONV_EPI_TANH: v = tanhf(v); break;
        case COGCONV_EPI_SIGNUM: v = v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); break;
        default: break;
        }
    }
    return v;
}

// Four adjacent output elements (16-byte aligned) as one store: non-temporal
// (streamed to memory, no cache line kept) when no epilogue reads them back.
__device__ __forceinline__ void store_out4(float* p, float a, float b, float c, float d, const Epi& e) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = {a, b, c, d};
    if (e.n == 0)
        __builtin_nontemporal_store(v, reinterpret_cast<f4*>(p));
    else
        *reinterpret_cast<f4*>(p) = v;
}

// The value stored for output element (plane, point): v itself without an epilogue.
__device__ __forceinline__ float epi(const Epi& e, float v, int plane, int point) {
    return e.n == 0 ? v : epi_eval(e, v, plane, point);
}

// Re-applies the epilogue to elements this thread has just stored (kernels
// whose accumulator arrays must stay in registers store them plainly first).
// The fence orders the thread's earlier stores before these loads.
__device__ __forceinline__ void epi_rewrite4(const Epi& e, float* p, int plane, int point) {
    __threadfence();
    float4 v = *reinterpret_cast<const float4*>(p);
    v.x = epi_eval(e, v.x, plane, point);
    v.y = epi_eval(e, v.y, plane, point + 1);
    v.z = epi_eval(e, v.z, plane, point + 2);
    v.w = epi_eval(e, v.w, plane, point + 3);
    *reinterpret_cast<float4*>(p) = v;
}

__device__ __forceinline__ void epi_rewrite1(const Epi& e, float* p, int plane, int point) {
    __threadfence();
    *p = epi_eval(e, *p, plane, point);
}

}  // namespace cogconv
