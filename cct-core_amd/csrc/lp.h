// MFMA input types of the BF16 / FP16 / split-BF16 kernels (device code,
// included by the .hip sources only).
#pragma once

#include "internal.h"

namespace cogconv {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

template <int PREC>
struct LpType;
template <>
struct LpType<COGCONV_BF16> {
    typedef __bf16 T;
    typedef bf16x8 V;
};
template <>
struct LpType<COGCONV_SPLIT_BF16> {
    typedef __bf16 T;
    typedef bf16x8 V;
};
template <>
struct LpType<COGCONV_FP16> {
    typedef _Float16 T;
    typedef halfx8 V;
};

template <int PREC>
__device__ __forceinline__ floatx4 lp_mfma(typename LpType<PREC>::V a, typename LpType<PREC>::V b, floatx4 c) {
    if constexpr (PREC == COGCONV_FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace cogconv
