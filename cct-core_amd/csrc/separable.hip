// Separable convolution of a 2-D ScalarField: a 1-D row filter, then a 1-D
// column filter (CogFunctionAPI.convolveSeparableInternal,
// src/main/scala/cogx/api/CogFunctionAPI.scala:1277-1316, which builds
// ConvolveRows2DHyperKernel then ConvolveColumns2DHyperKernel,
// $HK/ConvolveRows2DHyperKernel.scala:50-125 and ConvolveColumns2DHyperKernel.scala).
//
//   mid(r, c) = sum_i  X(r, c - hr + i) * tr[i]        (border on the column)
//   out(r, c) = sum_j mid(r - hc + j, c) * tc[j]        (border on the row of mid)
//
// with t = the filter for cross-correlation, reversed for convolution, and
// hr = (FR - 1) / 2, hc = (FC - 1) / 2; each sum runs in tap order in float.
// Both passes run in one kernel: a block stages its (32 + FC - 1) x (64 + FR - 1)
// input tile with the border applied to both coordinates, row-filters every
// staged row into LDS, then column-filters that.  A border applied to the rows
// of the staged input equals the border applied to the rows of mid (Zero: rows
// of zeros filter to zero; Clamp and Cyclic: the filtered row is that of the
// clamped or wrapped row), so the result is the two-pass one, element for
// element, without the intermediate field in HBM.
#include "epi.h"
#include "internal.h"

namespace cogconv {

constexpr int SEP_TR = 32;        // output rows per block
constexpr int SEP_TC = 64;        // output columns per block
constexpr int SEP_MAX_TAPS = 63;  // per filter

__device__ __forceinline__ int sep_coord(int v, int n, int border, bool* in) {
    if (border == COGCONV_BORDER_CLAMP) return min(max(v, 0), n - 1);
    if (border == COGCONV_BORDER_CYCLIC) {
        v %= n;
        return v < 0 ? v + n : v;
    }
    *in = *in && v >= 0 && v < n;
    return min(max(v, 0), n - 1);
}

__global__ __launch_bounds__(256) void k_separable(const float* __restrict__ X, const float* __restrict__ RF,
                                                   const float* __restrict__ CF, float* __restrict__ O, int H, int W,
                                                   int FR, int FC, int border, int flip, int tilesC, Epi e) {
    extern __shared__ float sep_smem[];
    const int SR = SEP_TR + FC - 1;  // staged rows
    const int SC = SEP_TC + FR - 1;  // staged columns
    const int SCP = SC | 1;          // odd pitch
    float* sIn = sep_smem;                    // [SR][SCP]
    float* sMid = sIn + (size_t)SR * SCP;     // [SR][SEP_TC + 1]
    float* sRf = sMid + (size_t)SR * (SEP_TC + 1);
    float* sCf = sRf + SEP_MAX_TAPS;
    const int MP = SEP_TC + 1;
    const int plane = blockIdx.y;
    const int tr = blockIdx.x / tilesC, tc = blockIdx.x - tr * tilesC;
    const int r0 = tr * SEP_TR, c0 = tc * SEP_TC;
    const int hr = (FR - 1) / 2, hc = (FC - 1) / 2;
    const float* x = X + (size_t)plane * H * W;
    const int tid = threadIdx.x;
    for (int i = tid; i < FR; i += 256) sRf[i] = RF[flip ? FR - 1 - i : i];
    for (int j = tid; j < FC; j += 256) sCf[j] = CF[flip ? FC - 1 - j : j];
    for (int q = tid; q < SR * SC; q += 256) {
        const int rr = q / SC, cc = q - rr * SC;
        bool in = true;
        const int y = sep_coord(r0 - hc + rr, H, border, &in);
        const int xx = sep_coord(c0 - hr + cc, W, border, &in);
        sIn[rr * SCP + cc] = in ? x[(size_t)y * W + xx] : 0.f;
    }
    __syncthreads();
    // Row pass: every staged row, the block's 64 output columns.
    for (int q = tid; q < SR * SEP_TC; q += 256) {
        const int rr = q >> 6, cc = q & 63;
        const float* src = sIn + rr * SCP + cc;
        float s = 0.f;
        for (int i = 0; i < FR; i++) s += src[i] * sRf[i];
        sMid[rr * MP + cc] = s;
    }
    __syncthreads();
    // Column pass: thread (ry, cc) computes rows ry, ry + 4, .. of column cc.
    const int cc = tid & 63, ry = tid >> 6;
    const int col = c0 + cc;
    if (col >= W) return;
    for (int r = ry; r < SEP_TR; r += 4) {
        const int row = r0 + r;
        if (row >= H) break;
        const float* src = sMid + r * MP + cc;
        float s = 0.f;
        for (int j = 0; j < FC; j++) s += src[j * MP] * sCf[j];
        O[((size_t)plane * H + row) * W + col] = epi(e, s, plane, row * W + col);
    }
}

size_t separable_lds_bytes(int FR, int FC) {
    const int SR = SEP_TR + FC - 1, SCP = (SEP_TC + FR - 1) | 1;
    return ((size_t)SR * SCP + (size_t)SR * (SEP_TC + 1) + 2 * SEP_MAX_TAPS) * sizeof(float);
}

}  // namespace cogconv

using namespace cogconv;

extern "C" int cogconv_separable(const float* in, const cogconv_field_type* type, const float* row_filter,
                                 int32_t row_taps, const float* column_filter, int32_t column_taps, int32_t border,
                                 int32_t orientation, float* out, const cogconv_epilogue* epi,
                                 const float* const* operands, cogconv_stream stream) {
    clear_error();
    if (!in || !type || !row_filter || !column_filter || !out) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    // convolveSeparableInternal's checks (CogFunctionAPI.scala:1284-1303).
    if (type->dims != 2) return fail(COGCONV_ERR_ILLEGAL_OP, "field must be 2D");
    if (type->tensor_order != 0 || type->planes != 1) return fail(COGCONV_ERR_ILLEGAL_OP, "field must be scalar");
    if (row_taps < 1 || row_taps % 2 != 1) return fail(COGCONV_ERR_ILLEGAL_OP, "row filter must be an odd size");
    if (column_taps < 1 || column_taps % 2 != 1)
        return fail(COGCONV_ERR_ILLEGAL_OP, "column filter must be an odd size");
    if (border != COGCONV_BORDER_CLAMP && border != COGCONV_BORDER_ZERO && border != COGCONV_BORDER_CYCLIC)
        return fail(COGCONV_ERR_ILLEGAL_OP, "unsupported border policy %d", border);
    if (orientation != COGCONV_CROSS_CORRELATION && orientation != COGCONV_CONVOLUTION)
        return fail(COGCONV_ERR_INVALID_ARG, "bad orientation %d", orientation);
    if (row_taps > SEP_MAX_TAPS || column_taps > SEP_MAX_TAPS)
        return fail(COGCONV_ERR_UNSUPPORTED, "separable filters of at most %d taps", SEP_MAX_TAPS);
    if (type->rows < 1 || type->cols < 1) return fail(COGCONV_ERR_INVALID_ARG, "empty field");
    Epi e;
    int n = 0;
    if (int st = epi_build(*type, epi, &e, &n)) return st;
    if (int st = epi_bind(&e, n, operands)) return st;
    const int tilesR = (type->rows + SEP_TR - 1) / SEP_TR, tilesC = (type->cols + SEP_TC - 1) / SEP_TC;
    const long blocks = (long)tilesR * tilesC;
    if (blocks > 0x7fffffffL) return fail(COGCONV_ERR_UNSUPPORTED, "field too large");
    hipLaunchKernelGGL(k_separable, dim3((unsigned)blocks, type->planes), dim3(256),
                       separable_lds_bytes(row_taps, column_taps), reinterpret_cast<hipStream_t>(stream), in,
                       row_filter, column_filter, out, type->rows, type->cols, row_taps, column_taps, border,
                       orientation == COGCONV_CONVOLUTION ? 1 : 0, tilesC, e);
    return check_hip(hipGetLastError(), "separable launch");
}
