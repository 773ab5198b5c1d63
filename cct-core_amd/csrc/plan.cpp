// ConvolveOp planning: legality checks, result FieldType, geometry, kernel choice.
//
// Host-only code.  Mirrors, for the ConvolveOp path only:
//   BinaryOperator.apply checks   src/main/scala/cogx/compiler/parser/syntaxtree/BinaryOperator.scala:65-93
//   outputFieldType               src/main/scala/cogx/compiler/codegenerator/opencl/hyperkernels/ConvolveHyperKernel.scala:1297-1450
//   convolveDirectly checks       src/main/scala/cogx/compiler/codegenerator/opencl/generator/DynamicConvolutionGenerator.scala:484-539
//   Up/Down(1) -> NoSampling      src/main/scala/cogx/api/CogFunctionAPI.scala:1140-1144
//   kernel routing                ConvolveHyperKernel.scala:1003-1148 (re-decided for gfx950, see choose_kernel)
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

using namespace cogconv;

namespace {

bool is_fa(int mode) {
    return mode == COGCONV_FILTER_ADJOINT || mode == COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM;
}

int border_extent(int border, int u, int f) {
    if (border == COGCONV_BORDER_VALID) return u - f + 1;
    if (border == COGCONV_BORDER_FULL) return u + f - 1;
    return u;
}

const char* mode_name(int m) {
    static const char* names[] = {"PlaneByPlane", "ProjectFrame", "ProjectFrameBlockReduceSum", "BackProjectFrame",
                                  "BackProjectFrameBlockReduceSum", "FilterAdjoint", "FilterAdjointBlockReduceSum"};
    return (m >= 0 && m < 7) ? names[m] : "?";
}

// Validates `d`, fills the result type and (optionally) the geometry and plane map.
int analyse(const cogconv_desc* d, cogconv_field_type* out, Geom* gout, PlaneMap* mout) {
    if (!d || !out) return fail(COGCONV_ERR_INVALID_ARG, "null descriptor or result pointer");
    if (d->mode < 0 || d->mode > 6) return fail(COGCONV_ERR_INVALID_ARG, "bad vector mode %d", d->mode);
    if (d->border < 0 || d->border > 4) return fail(COGCONV_ERR_INVALID_ARG, "bad border policy %d", d->border);
    if (d->orientation < 0 || d->orientation > 1)
        return fail(COGCONV_ERR_INVALID_ARG, "bad filter orientation %d", d->orientation);
    if (d->sampling < 0 || d->sampling > 2) return fail(COGCONV_ERR_INVALID_ARG, "bad sampling policy %d", d->sampling);
    if (d->step < 1) return fail(COGCONV_ERR_INVALID_ARG, "sampling step must be >= 1, got %d", d->step);
    if (d->batch < 1) return fail(COGCONV_ERR_INVALID_ARG, "batchSize must be >= 1, got %d", d->batch);
    if (d->precision < 0 || d->precision > 3) return fail(COGCONV_ERR_INVALID_ARG, "bad precision %d", d->precision);
    const cogconv_field_type& a = d->in0;
    const cogconv_field_type& f = d->in1;
    if (a.rows < 1 || a.cols < 1 || a.planes < 1 || f.rows < 1 || f.cols < 1 || f.planes < 1)
        return fail(COGCONV_ERR_INVALID_ARG, "field sizes must be positive");
    if (a.tensor_order < 0 || a.tensor_order > 1 || f.tensor_order < 0 || f.tensor_order > 1)
        return fail(COGCONV_ERR_ILLEGAL_OP, "Frame convolve operations not supported on MatrixFields");
    // BinaryOperator.scala:65-93
    if (a.dims < 1 || a.dims > 2)
        return fail(COGCONV_ERR_ILLEGAL_OP, "convolution not supported on field of dimension %d", a.dims);
    if (a.dims != f.dims)
        return fail(COGCONV_ERR_ILLEGAL_OP, "1st field's dimension %d must equal 2nd field's dimension %d", a.dims,
                    f.dims);
    if (a.dims == 1 && (a.rows != 1 || f.rows != 1))
        return fail(COGCONV_ERR_INVALID_ARG, "1-D fields must have rows == 1");
    const int sampling = (d->step == 1) ? COGCONV_NO_SAMPLING : d->sampling;
    const int step = d->step;
    if (a.planes % d->batch)
        return fail(COGCONV_ERR_ILLEGAL_OP, "Image depth %d must be a multiple of the batchSize %d", a.planes,
                    d->batch);
    const int ppi = a.planes / d->batch;
    const int T = f.planes;
    switch (d->mode) {
    case COGCONV_PROJECT_FRAME:
    case COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM:
    case COGCONV_BACK_PROJECT_FRAME:
    case COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM:
        if (T % ppi)
            return fail(COGCONV_ERR_ILLEGAL_OP, "Filter depth %d must be a multiple of the single image depth %d", T,
                        ppi);
        break;
    case COGCONV_FILTER_ADJOINT:
    case COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM:
        if (T % d->batch)
            return fail(COGCONV_ERR_ILLEGAL_OP, "Filter tensorshape points %d must be a multiple of the batchSize %d",
                        T, d->batch);
        break;
    default:  // PlaneByPlane: ConvolveHyperKernel.scala:1378-1386
        if (!(a.tensor_order == 0 || f.tensor_order == 0 || a.planes == f.planes))
            return fail(COGCONV_ERR_ILLEGAL_OP, "PlaneByPlane needs a scalar operand or equal tensor shapes (%d vs %d)",
                        a.planes, f.planes);
        break;
    }
    // DynamicConvolutionGenerator.scala:500-506
    if (d->border != COGCONV_BORDER_VALID && d->border != COGCONV_BORDER_FULL) {
        if ((f.cols % 2) == 0 || (a.dims == 2 && (f.rows % 2) == 0))
            return fail(COGCONV_ERR_ILLEGAL_OP, "filter must be of odd shape for border policy %d (got %dx%d)",
                        d->border, f.rows, f.cols);
    }
    // L512-526
    if (sampling == COGCONV_DOWNSAMPLE_OUTPUT) {
        int cr = 0, cc = 0;
        if (d->border == COGCONV_BORDER_VALID) { cr = f.rows - 1; cc = f.cols - 1; }
        if (d->border == COGCONV_BORDER_FULL) { cr = 1 - f.rows; cc = 1 - f.cols; }
        if ((a.cols - cc) % step || (a.dims == 2 && (a.rows - cr) % step))
            return fail(COGCONV_ERR_ILLEGAL_OP,
                        "The size of the input for convolution must be divisible by the downsampling step size "
                        "after border contraction.");
    }
    // L531-536
    if (sampling == COGCONV_UPSAMPLE_INPUT && d->border == COGCONV_BORDER_CLAMP)
        return fail(COGCONV_ERR_ILLEGAL_OP, "Integrated upsampling with BorderClamp convolution not supported");
    // FilterAdjoint up-sampling applies to the 2nd operand and is only defined
    // for BorderValid (ConvolveHyperKernel.scala:1090-1098 routes it away from
    // the standard kernel); other borders have no consistent definition.
    if (sampling == COGCONV_UPSAMPLE_INPUT && is_fa(d->mode) && d->border != COGCONV_BORDER_VALID)
        return fail(COGCONV_ERR_ILLEGAL_OP, "FilterAdjoint with UpsampleInputConvolution requires BorderValid");

    Geom g;
    std::memset(&g, 0, sizeof g);
    g.H = a.rows; g.W = a.cols; g.FH = f.rows; g.FW = f.cols;
    g.upAr = g.upAc = g.upFr = g.upFc = g.dnR = g.dnC = 1;
    const bool two = a.dims == 2;
    if (sampling == COGCONV_UPSAMPLE_INPUT) {
        if (is_fa(d->mode)) { g.upFc = step; g.upFr = two ? step : 1; }
        else { g.upAc = step; g.upAr = two ? step : 1; }
    } else if (sampling == COGCONV_DOWNSAMPLE_OUTPUT) {
        g.dnC = step; g.dnR = two ? step : 1;
    }
    g.UR = g.H * g.upAr; g.UC = g.W * g.upAc;
    g.FHe = g.FH * g.upFr; g.FWe = g.FW * g.upFc;
    int br = border_extent(d->border, g.UR, g.FHe);
    int bc = border_extent(d->border, g.UC, g.FWe);
    if (!two) br = 1;
    if (br < 1 || bc < 1)
        return fail(COGCONV_ERR_ILLEGAL_OP, "filter (%dx%d) larger than input (%dx%d) for BorderValid", g.FHe, g.FWe,
                    g.UR, g.UC);
    g.OH = (br + g.dnR - 1) / g.dnR;
    g.OW = (bc + g.dnC - 1) / g.dnC;
    if (d->border == COGCONV_BORDER_FULL) { g.hR = g.FHe - 1; g.hC = g.FWe - 1; }
    else if (d->border == COGCONV_BORDER_VALID) { g.hR = 0; g.hC = 0; }
    else { g.hR = (g.FHe - 1) / 2; g.hC = (g.FWe - 1) / 2; }
    if (!two) g.hR = 0;
    g.border = d->border;
    g.flip = d->orientation == COGCONV_CONVOLUTION;

    int planes = 0, order = 1;
    switch (d->mode) {
    case COGCONV_PROJECT_FRAME:
    case COGCONV_BACK_PROJECT_FRAME:
        planes = T * d->batch;
        order = 1;
        break;
    case COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM:
    case COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM:
        planes = d->batch * (T / ppi);
        order = (planes == 1 && a.planes <= 4) ? 0 : 1;  // ConvolveHyperKernel.scala:1393-1410
        break;
    case COGCONV_FILTER_ADJOINT:
        planes = ppi * (T / d->batch) * d->batch;
        order = planes > 1;
        break;
    case COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM:
        planes = ppi * (T / d->batch);
        order = planes > 1;
        break;
    default:
        if (a.tensor_order == 0) { planes = f.planes; order = f.tensor_order; }
        else { planes = a.planes; order = a.tensor_order; }
        break;
    }
    out->dims = a.dims;
    out->rows = g.OH;
    out->cols = g.OW;
    out->planes = planes;
    out->tensor_order = order;

    if (gout) *gout = g;
    if (mout) {
        PlaneMap m;
        m.mode = d->mode;
        m.B = d->batch;
        m.N = ppi;
        m.T = T;
        m.PA = a.planes;
        m.PFn = f.planes;
        m.nOut = planes;
        switch (d->mode) {
        case COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM:
        case COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM: m.R = ppi; break;
        case COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM: m.R = d->batch; break;
        default: m.R = 1; break;
        }
        *mout = m;
    }
    return COGCONV_OK;
}

DirectCfg choose_direct(const Geom& g, int rgMax) {
    DirectCfg c;
    c.RG = rgMax;
    c.FCR = std::min(g.FHe, 16);
    c.FCC = std::min(g.FWe, 32);
    const size_t budget = 48 * 1024;
    for (;;) {
        const int rows = (4 * c.RG - 1) * g.dnR + c.FCR;
        const int cols = 63 * g.dnC + c.FCC;
        const size_t bytes = (size_t)(rows * cols + c.FCR * c.FCC) * sizeof(float);
        if (bytes <= budget || (c.RG == 1 && c.FCR == 1 && c.FCC == 1)) {
            c.ldsRows = rows;
            c.ldsCols = cols;
            c.lds_bytes = bytes;
            break;
        }
        if (c.RG > 1) c.RG /= 2;
        else if (c.FCR > 1) c.FCR = (c.FCR + 1) / 2;
        else c.FCC = (c.FCC + 1) / 2;
    }
    c.tilesR = (g.OH + 4 * c.RG - 1) / (4 * c.RG);
    c.tilesC = (g.OW + 63) / 64;
    return c;
}

// FP32 MFMA implicit GEMM for the frame block-reduce modes: square odd-or-1
// filters up to 9x9 on 2-D fields without input up-sampling.
bool igemm_frame_eligible(const cogconv_desc* d, const Geom& g) {
    if (d->mode != COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM && d->mode != COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM)
        return false;
    if (d->in0.dims != 2 || g.FH != g.FW || g.upAr != 1 || g.upAc != 1 || g.dnR != g.dnC) return false;
    if (31 * g.dnC + g.FW > 128) return false;  // staged tile rows are at most two 64-lane passes
    if (7 * g.dnR + g.FH > 64) return false;    // one lane per staged tile row in the row table
    if ((long)d->in0.planes * g.H * g.W >= (1L << 31)) return false;  // 32-bit staging offsets
    if ((long)(d->in0.planes / d->batch) * g.H * g.W * 4 >= (1L << 31)) return false;  // buffer byte offsets
    return igemm_f32_supported(g.FH);
}

// mbOnly / rcOnly (0 = any) restrict the search to one output block or plane chunk.
IgemmCfg choose_igemm_frame(const Geom& g, const PlaneMap& m, int mbOnly = 0, int rcOnly = 0) {
    const size_t limit = 80 * 1024;  // two blocks per CU (160 KB of LDS)
    IgemmCfg c;
    std::memset(&c, 0, sizeof c);
    c.variant = IV_F32_FRAME;
    c.Mout = m.T / m.N;
    c.tilesR = (g.OH + 7) / 8;
    c.tilesC = (g.OW + 31) / 32;
    c.tileRows = 7 * g.dnR + g.FH;
    c.tileCols = 31 * g.dnC + g.FW;
    int ps = c.tileRows * c.tileCols;
    ps += ((16 - ps % 32) + 32) % 32;  // plane stride = 16 mod 32 words: the four k-lane groups hit different banks
    c.planeStride = ps;
    const int taps = g.FH * g.FW;
    const int maxRC = std::min(32, ((m.R + 7) / 8) * 8);
    // Largest output block (MB) and plane chunk (RC, a multiple of 8: two
    // plane groups of whole MFMA k-quads) whose staging fits in LDS.
    for (int MB = c.Mout <= 16 ? 1 : (c.Mout <= 32 ? 2 : 4); MB >= 1; MB = MB == 4 ? 2 : MB - 1) {
        if (mbOnly && MB != mbOnly) {
            if (MB == 1) break;
            continue;
        }
        int as = taps * 16 * MB;
        as += ((16 - as % 32) + 32) % 32;
        // Of the chunks that fit, the one staging the fewest plane slots in all
        // (ceil(R / RC) * RC: 32 planes go as 16 + 16, not 24 + 8), then the
        // larger chunk.
        int best = 0;
        size_t bestBytes = 0;
        long bestSlots = 0;
        for (int rc = maxRC; rc >= 8; rc -= 8) {
            if (rcOnly && rc != rcOnly) continue;
            const size_t stage = ((size_t)rc * (ps + as) + 128) * sizeof(float);  // + scratch slots
            const size_t xch = igemm_f32_exchange_floats(MB) * sizeof(float);
            const size_t bytes = std::max(stage, xch);
            if (bytes > limit) continue;
            const long slots = (long)((m.R + rc - 1) / rc) * rc;
            if (!best || slots < bestSlots) {
                best = rc;
                bestBytes = bytes;
                bestSlots = slots;
            }
        }
        if (best) {
            c.MB = MB;
            c.RC = best;
            c.aStride = as;
            c.lds_bytes = bestBytes;
            c.mTiles = (c.Mout + 16 * MB - 1) / (16 * MB);
            return c;
        }
        if (MB == 1) break;
    }
    c.MB = 0;  // does not fit: caller falls back
    return c;
}

// FP32 MFMA split-reduction GEMM for FilterAdjointBlockReduceSum, BorderValid.
// cbw (0 = by size) fixes the column-group width; partials (0 = by size) is
// the target number of blocks, each of which writes one partial result.
bool choose_fa_igemm(const cogconv_desc* d, const Geom& g, const PlaneMap& m, FaCfg* out, int cbw = 0,
                     long partials = 0, int cw = 0, int nw = 0) {
    if (d->mode != COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM || d->border != COGCONV_BORDER_VALID || d->in0.dims != 2)
        return false;
    FaCfg c;
    std::memset(&c, 0, sizeof c);
    c.K = m.T / m.B;
    if (c.K > 64) return false;
    c.MB = (c.K + 15) / 16;
    const int P = g.OH * g.OW;
    const long ncols = (long)m.N * P;
    // The kernel splits output columns j < N*OH*OW into (plane, row, column)
    // with fast_div, which is exact below 2^22.
    if (ncols >= (1L << 22)) return false;
    const long colBlocks = (ncols + 15) / 16;
    c.CBW = cbw ? cbw : (colBlocks <= 3 ? 3 : (colBlocks <= 6 ? 6 : 9));
    c.colGroups = (int)((colBlocks + c.CBW - 1) / c.CBW);
    c.nCntMax = std::min(m.N, (16 * c.CBW + P - 1) / P + 1);
    // Tap columns in chunks whose staged x rows fit two 64-lane passes.
    if ((g.OW - 1) * g.dnC > 64) return false;
    // cw > 0 asks for narrower chunks (more items, so more blocks).
    c.colChunks = cw > 0 ? (g.FWe + cw - 1) / cw : 1;
    for (;;) {
        c.CW = (g.FWe + c.colChunks - 1) / c.colChunks;
        c.DWp = (c.CW + 3) / 4 * 4;
        c.xCols = (g.OW - 1) * g.dnC + c.DWp;
        if (c.xCols <= 128) break;
        c.colChunks++;
    }
    c.CW = c.DWp;  // chunk starts on multiples of 4; the last chunk is masked at FWe
    c.colChunks = (g.FWe + c.CW - 1) / c.CW;
    c.MTp = c.MB == 1 ? 17 : (c.MB == 4 ? 81 : 49);  // 16*MB padded to 17 mod 32 (k_fa_igemm_f32's MTP)
    if ((long)m.B * m.N * g.H * g.W >= (1L << 31) || (long)m.B * c.K * g.FH * g.FW >= (1L << 31)) return false;
    if ((long)m.N * g.H * g.W * 4 >= (1L << 31) || (long)c.K * g.FH * g.FW * 4 >= (1L << 31)) return false;
    // nw = 0 (heuristic): 4 waves, or 8 when the grid is one block per CU
    // (decided below, once the split count is known).
    if (nw != 0 && nw != 4 && nw != 8) return false;
    c.NW = nw ? nw : 4;
    // Cross-wave reduction buffer: two parked sums per wave quad (8 waves with
    // MB < 4: four, each of all MB row blocks).
    const size_t red = (size_t)(c.NW == 8 && c.MB != 4 ? 4 : 2) * c.MB * c.CBW * 4 * 64 * sizeof(float);
    // Tap rows per item: as many as fit 80 KB (two blocks per CU); but at
    // least 4, one per wave, when the tap rows allow it and 160 KB holds them.
    // x rows pitched at 4 mod 32 words and planes at 16 mod 32: the B-operand
    // reads of a half-wave (output columns (n, R, C) x two tap-column lanes)
    // then spread over the banks instead of piling onto a few.
    c.xColsP = c.xCols + ((4 - c.xCols % 32) + 32) % 32;
    auto plane_pitch = [&](int xRows) {
        const int pp = xRows * c.xColsP;
        return pp + ((16 - pp % 32) + 32) % 32;
    };
    auto lds_for = [&](int rr) {
        const int xRows = rr + (g.OH - 1) * g.dnR;
        const size_t bytes =
            ((size_t)c.MTp * rr * c.DWp + (size_t)c.nCntMax * plane_pitch(xRows) + 128) * sizeof(float);
        return std::max(bytes, red);
    };
    c.RR = 0;
    for (int rr = std::min(g.FHe, 8); rr >= 1; rr--) {
        if (lds_for(rr) <= 80 * 1024) { c.RR = rr; break; }
    }
    if (c.RR < 4 && g.FHe >= 4 && lds_for(4) <= 160 * 1024) c.RR = 4;
    if (c.RR == 0) return false;
    c.xRows = c.RR + (g.OH - 1) * g.dnR;
    c.xPlaneP = plane_pitch(c.xRows);
    // dy rows, x rows and 128 scratch slots for masked-off stores
    c.lds_bytes = ((size_t)c.MTp * c.RR * c.DWp + (size_t)c.nCntMax * c.xPlaneP + 128) * sizeof(float);
    c.lds_bytes = std::max(c.lds_bytes, red);
    c.rowChunks = (g.FHe + c.RR - 1) / c.RR;
    const long items = (long)m.B * c.colChunks * c.rowChunks;
    if (partials <= 0) partials = items * c.colGroups >= 8192 ? 512 : 256;  // blocks: 1-2 per CU
    const long want = (partials + c.colGroups - 1) / c.colGroups;
    c.splits = (int)std::max(1L, std::min(items, want));
    if (items * (c.splits + 1) >= (1L << 31)) return false;  // the kernel's 32-bit item arithmetic
    // 8 waves, two per SIMD, so one wave's MFMAs overlap the other's LDS
    // reads (c2 FilterAdjoint 20.1 -> 18.4 us, c3 495 -> 436 us).
    FaCfg c8;
    if (nw == 0 && choose_fa_igemm(d, g, m, &c8, cbw, partials, cw, 8) &&
        c8.splits == c.splits && c8.RR == c.RR && c8.CW == c.CW) {
        *out = c8;
        return true;
    }
    *out = c;
    return true;
}

// BF16 / FP16 / split-BF16 MFMA split-reduction GEMM for
// FilterAdjointBlockReduceSum, BorderValid (k_fa_igemm_lp).
bool choose_fa_igemm_lp(const cogconv_desc* d, const Geom& g, const PlaneMap& m, FaCfg* out, int cbw = 0,
                        long partials = 0, int cw = 0, int nw = 0) {
    if (d->mode != COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM || d->border != COGCONV_BORDER_VALID || d->in0.dims != 2)
        return false;
    FaCfg c;
    std::memset(&c, 0, sizeof c);
    c.K = m.T / m.B;
    if (c.K > 64) return false;
    c.MB = (c.K + 15) / 16;
    // 8 waves (the default: two per SIMD; c2 BF16 25.6 -> 23.9 us, c3 414 ->
    // 379 us) or 4.
    if (nw != 0 && nw != 4 && nw != 8) return false;
    c.NW = nw ? nw : 8;
    const int P = g.OH * g.OW;
    const long ncols = (long)m.N * P;
    const long colBlocks = (ncols + 15) / 16;
    c.CBW = cbw ? cbw : (colBlocks <= 3 ? 3 : (colBlocks <= 6 ? 6 : 9));
    c.colGroups = (int)((colBlocks + c.CBW - 1) / c.CBW);
    // The split-BF16 8-wave kernel with three row blocks and nine column
    // blocks (all accumulators in every wave) would spill registers.
    if (c.NW == 8 && d->precision == COGCONV_SPLIT_BF16 && c.MB == 3 && c.CBW == 9) {
        if (nw == 8) return false;
        c.NW = 4;
    }
    c.nCntMax = std::min(m.N, (16 * c.CBW + P - 1) / P + 1);
    // Tap columns per chunk: 32, 64 or 128 (whole 32-deep MFMA slices; 8
    // columns per staging lane).  The heuristic stops at 64 (32 for the
    // two-term split-BF16 staging): wider chunks stage more x copies per item
    // (c3 BF16 FilterAdjoint 530 us at 128, 379 us at 64; c3 split-BF16
    // 872 us at 64, 593 us at 32).
    const int widest = g.FWe <= 32 ? 32 : (g.FWe <= 64 ? 64 : 128);
    c.CW = std::min(widest, d->precision == COGCONV_SPLIT_BF16 ? 32 : 64);
    if (cw) {
        if (cw > widest) return false;  // no wider than the dy rows need
        c.CW = cw;
    }
    c.DWp = c.CW;
    c.colChunks = (g.FWe + c.CW - 1) / c.CW;
    c.MTp = 16 * c.MB;
    if ((long)m.B * m.N * g.H * g.W >= (1L << 31) || (long)m.B * c.K * g.FH * g.FW >= (1L << 31)) return false;
    if ((long)m.N * g.H * g.W * 4 >= (1L << 31) || (long)c.K * g.FH * g.FW * 4 >= (1L << 31)) return false;
    auto lds_for = [&](int rr) {
        FaCfg t = c;
        t.RR = rr;
        t.xRows = rr + (g.OH - 1) * g.dnR;
        return fa_igemm_lp_lds_bytes(t, g.OW, d->precision);
    };
    c.RR = 0;
    for (int rr = std::min(g.FHe, 8); rr >= 1; rr--) {
        if (lds_for(rr) <= 80 * 1024) { c.RR = rr; break; }
    }
    if (c.RR < 4 && g.FHe >= 4 && lds_for(4) <= 160 * 1024) c.RR = 4;
    if (c.RR == 0 && lds_for(1) <= 160 * 1024) c.RR = 1;
    if (c.RR == 0) return false;
    c.xRows = c.RR + (g.OH - 1) * g.dnR;
    if ((long)m.N * c.xRows * g.OW >= (1L << 22)) return false;  // staging row index (fast_div range)
    c.lds_bytes = lds_for(c.RR);
    c.rowChunks = (g.FHe + c.RR - 1) / c.RR;
    const long items = (long)m.B * c.colChunks * c.rowChunks;
    if (partials <= 0) partials = items * c.colGroups >= 8192 ? 512 : 256;
    long want = (partials + c.colGroups - 1) / c.colGroups;
    c.splits = (int)std::max(1L, std::min(items, want));
    if (items * (c.splits + 1) >= (1L << 31)) return false;  // the kernel's 32-bit item arithmetic
    *out = c;
    return true;
}

// Row-streaming BF16 / FP16 FilterAdjointBlockReduceSum (k_fa_rows_lp):
// BorderValid, unit steps, cross-correlation, a square 1x1, 3x3 or 5x5 dW.
// (nb, kh, kbw) pick the tile and waves (0: from K and N); rowSplits (0: from
// the wave count) splits each image's dy rows so that the grid fills the CUs.
bool choose_fa_rows_lp(const cogconv_desc* d, const Geom& g, const PlaneMap& m, FaCfg* out, int nb = 0, int kh = 0,
                       int kbw = 0, int rowSplits = 0) {
    if (d->mode != COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM || d->border != COGCONV_BORDER_VALID || d->in0.dims != 2)
        return false;
    if (d->precision != COGCONV_BF16 && d->precision != COGCONV_FP16) return false;
    if (g.flip || g.upFr != 1 || g.upFc != 1 || g.upAr != 1 || g.upAc != 1 || g.dnR != 1 || g.dnC != 1) return false;
    if (g.OH != g.OW || (g.OH != 1 && g.OH != 3 && g.OH != 5)) return false;
    FaCfg c;
    std::memset(&c, 0, sizeof c);
    c.rows = 1;
    c.K = m.T / m.B;
    const int N = m.N;
    if (!nb) {
        // Whole 64 x 64 tiles (8 waves) read every dy and x row once; narrower
        // tiles for fewer planes.
        if (N > 32 && c.K > 32 && g.OH <= 3) { nb = 4; kh = 2; kbw = 2; }
        else if (N > 16) { nb = 2; kh = 2; kbw = (g.OH == 5 || c.K <= 32) ? 1 : 2; }
        else { nb = 1; kh = 4; kbw = 1; }
    }
    // The tiles the launcher instantiates; 5x5 dW keeps one dy block a wave
    // (25 accumulators).
    const int sel = nb * 100 + kh * 10 + kbw;
    if (sel != 222 && sel != 221 && sel != 142 && sel != 141 && sel != 422 && sel != 421 && sel != 414) return false;
    if (g.OH == 5 && kbw != 1) return false;
    c.nb = nb;
    c.kh = kh;
    c.kbw = kbw;
    const int KT = kh * kbw * 16, NT = nb * 16;
    c.nTiles = (N + NT - 1) / NT;
    c.kTiles = (c.K + KT - 1) / KT;
    c.chunks = (g.FW + 31) / 32;
    // Byte offsets of the staged samples (up to 40 pixels past a row start,
    // and the tile's padding planes) stay below 2^31 inside one image.
    if ((long)(c.K + KT) * g.FH * g.FW * 4 >= (1L << 31) || (long)(N + NT) * g.H * g.W * 4 >= (1L << 31)) return false;
    const long perRows = (long)c.nTiles * c.kTiles * m.B * c.chunks;  // blocks without row splits
    if (!rowSplits) {
        // About two waves a SIMD over the 256 CUs.
        rowSplits = 1;
        while (perRows * rowSplits * nb * kh < 2048 && rowSplits * 2 <= g.FH && rowSplits < 64) rowSplits *= 2;
    }
    if (rowSplits < 1 || rowSplits > g.FH) return false;
    if (perRows * rowSplits > 0x7fffffffL || (long)m.B * c.chunks * rowSplits > 0x7fffffffL) return false;
    c.rowSplits = rowSplits;
    c.rowsPer = (g.FH + rowSplits - 1) / rowSplits;
    c.splits = m.B * c.chunks * rowSplits;
    c.lds_bytes = fa_rows_lp_lds_bytes(c, g.OH);
    *out = c;
    return true;
}

// BF16 / FP16 / split-BF16 MFMA implicit GEMM for the frame block-reduce modes.
bool choose_igemm_lp(const cogconv_desc* d, const Geom& g, const PlaneMap& m, IgemmCfg* out, int mbOnly = 0,
                     bool ring = false) {
    if (d->mode != COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM && d->mode != COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM)
        return false;
    if (d->in0.dims != 2 || g.FH != g.FW || !igemm_lp_supported(g.FH) || g.upAr != 1 || g.upAc != 1 ||
        g.dnR != g.dnC)
        return false;
    if (31 * g.dnC + g.FW > 64) return false;  // one 64-lane pass per staged row
    if (7 * g.dnR + g.FH > 64) return false;   // one lane per staged tile row in the row table
    if ((long)d->in0.planes * g.H * g.W >= (1L << 31)) return false;
    if ((long)(d->in0.planes / d->batch) * g.H * g.W * 4 >= (1L << 31)) return false;  // buffer byte offsets
    // Staging offsets reach 8 planes past the image's last plane on top of the
    // 2^31 out-of-range marker.
    if ((long)(m.N + 8) * g.H * g.W * 4 >= (1L << 31)) return false;
    IgemmCfg c;
    std::memset(&c, 0, sizeof c);
    c.Mout = m.T / m.N;
    c.tilesR = (g.OH + 7) / 8;
    c.tilesC = (g.OW + 31) / 32;
    c.tileRows = 7 * g.dnR + g.FH;
    c.tileCols = 31 * g.dnC + g.FW;
    c.RC = (m.R + 31) / 32 * 32;  // planes padded to whole 32-deep MFMA slices
    // 16-byte staging loads of 4-column groups: groups aligned to 4 columns are
    // wholly inside or outside a row when W % 4 == 0, so borders that read 0
    // outside the field need no per-column remap; unit steps keep the staged
    // columns contiguous.  (The launch also checks the input's alignment.)
    // Clamp and Cyclic tiles stage their columns outside the field in a second
    // pass.  A wave's passes (VP of IP item slots) must cover its TRp items.
    c.vec4 = g.dnR == 1 && g.dnC == 1 && g.W % 4 == 0 && g.FH <= 9;
    if (c.vec4) {
        const int s4 = ((-g.hC) % 4 + 4) % 4;  // every tile starts at 32 tc - hC
        const int ng = (s4 + c.tileCols + 3) / 4;
        const int ip = 64 / ng, vp = (g.FH + 12) / 6;
        if (vp * ip < c.tileRows) c.vec4 = 0;
    }
    // The largest output block (MB) whose filter slice fits in LDS in at most two
    // groups of tap rows (beside the input tile); with mbOnly, that block with
    // as many tap rows a group as fit.  A larger MB reads each staged pixel for
    // more MFMAs and stages the input tile once for more outputs.
    for (int MB = c.Mout <= 16 ? 1 : (c.Mout <= 32 ? 2 : 4); MB >= 1; MB = MB == 4 ? 2 : MB - 1) {
        c.MB = MB;
        c.mTiles = (c.Mout + 16 * MB - 1) / (16 * MB);
        int trg = 0;
        c.ring = ring ? 1 : 0;
        for (int t = ring ? 1 : g.FH; t >= 1; t--) {
            c.trg = t;
            if (igemm_lp_lds_bytes(c, g.FH, d->precision) <= 160 * 1024) {
                trg = t;
                break;
            }
        }
        const bool fits = trg > 0 && (mbOnly || ring || MB == 1 || 2 * trg >= g.FH);
        if (fits && (!mbOnly || MB == mbOnly) && igemm_lp_workspace_bytes(c, g.FH, d->precision) < (1UL << 31)) {
            c.trg = trg;
            c.lds_bytes = igemm_lp_lds_bytes(c, g.FH, d->precision);
            *out = c;
            return true;
        }
        if (MB == 1) break;
    }
    return false;
}

SmallCfg choose_small(const Geom& g, const PlaneMap& m, long blocks = 4096) {
    SmallCfg c;
    c.qchunks = (g.OH * g.OW + 15) / 16;
    const long rowsTotal = (long)m.R * g.FHe;
    // Aim for a few thousand blocks so that all 256 CUs are busy, with at least
    // ~8 filter rows of reduction per block.
    long want = (blocks + (long)m.nOut * c.qchunks - 1) / ((long)m.nOut * c.qchunks);
    long maxs = std::max(1L, rowsTotal / 8);
    c.splits = (int)std::max(1L, std::min(want, maxs));
    return c;
}

// One candidate kernel configuration of a plan.
struct Variant {
    int kind;
    DirectCfg d;
    SmallCfg s;
    IgemmCfg i;
    FaCfg f;
    PhaseSet ps;
    FftCfg fft;
    size_t workspace;
};

// The FFT route (DynamicConvolutionGenerator.scala:109-148, convolveBy2DFFT
// L322-476): PlaneByPlane on 2-D fields, BorderZero / Clamp / Cyclic, no
// sampling, a filter no larger than the field.  The transform covers the
// field plus twice the filter's half-width in each dimension (the span a
// circular convolution needs to equal the bordered one; FFTConvolutionUtils.scala:50-97),
// rounded up to a 2/3/5/7-smooth length; a Cyclic field of such a length is
// transformed as it is.
bool choose_fft(const cogconv_desc* d, const Geom& g, const PlaneMap& m, FftCfg* out) {
    if (d->mode != COGCONV_PLANE_BY_PLANE || d->in0.dims != 2 || d->in1.dims != 2) return false;
    if (g.border != COGCONV_BORDER_ZERO && g.border != COGCONV_BORDER_CLAMP && g.border != COGCONV_BORDER_CYCLIC)
        return false;
    if (g.upAr != 1 || g.upAc != 1 || g.upFr != 1 || g.upFc != 1 || g.dnR != 1 || g.dnC != 1) return false;
    if (g.FH > g.H || g.FW > g.W || (g.FH % 2) == 0 || (g.FW % 2) == 0) return false;
    FftCfg c;
    const int hr = (g.FH - 1) / 2, hc = (g.FW - 1) / 2;
    c.Er = fft_good_length(g.H + 2 * hr);
    c.Ec = fft_good_length(g.W + 2 * hc);
    if (g.border == COGCONV_BORDER_CYCLIC && fft_good_length(g.H) == g.H) c.Er = g.H;
    if (g.border == COGCONV_BORDER_CYCLIC && fft_good_length(g.W) == g.W) c.Ec = g.W;
    if ((long)c.Er * c.Ec >= (1L << 30)) return false;
    c.PA = m.PA;
    c.PK = m.PFn;
    c.P = m.nOut;
    *out = c;
    return true;
}

// FP32 implicit GEMM for the frame block-reduce modes on 2x up-sampled input:
// the zero-inserted input makes output pixel (2R+pr, 2C+pc) see only the taps
// i = i0 + 2a, j = j0 + 2b whose sample is stored (i0 = (h - pr) mod 2), so
// each of the four output phases is a dense convolution of the stored samples
// with a ceil/floor(F/2)-tap sub-filter, at input row R + a + (pr - h + i0)/2.
bool choose_igemm_up(const cogconv_desc* d, const Geom& g, const PlaneMap& m, PhaseSet* out) {
    if (d->mode != COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM && d->mode != COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM)
        return false;
    if (d->in0.dims != 2 || g.upAr != 2 || g.upAc != 2 || g.dnR != 1 || g.dnC != 1 || g.FH != g.FW || g.FH < 3 ||
        g.FH > 9 || (g.FH & 1) == 0)
        return false;
    if ((long)d->in0.planes * g.H * g.W >= (1L << 31)) return false;
    if ((long)(d->in0.planes / d->batch) * g.H * g.W * 4 >= (1L << 31)) return false;
    PhaseSet ps;
    std::memset(&ps, 0, sizeof ps);
    for (int pr = 0; pr < 2; pr++)
        for (int pc = 0; pc < 2; pc++) {
            const int OHp = (g.OH - pr + 1) / 2, OWp = (g.OW - pc + 1) / 2;
            if (OHp <= 0 || OWp <= 0) continue;
            const int i0 = ((g.hR - pr) % 2 + 2) % 2, j0 = ((g.hC - pc) % 2 + 2) % 2;
            const int FR = (g.FH - i0 + 1) / 2, FC = (g.FW - j0 + 1) / 2;
            if (!igemm_f32_supported_rc(FR, FC)) return false;
            Geom q = g;
            q.FH = q.FHe = FR;
            q.FW = q.FWe = FC;
            q.upAr = q.upAc = q.upFr = q.upFc = 1;
            q.hR = -((pr - g.hR + i0) / 2);
            q.hC = -((pc - g.hC + j0) / 2);
            q.OH = OHp;
            q.OW = OWp;
            q.flip = 0;
            const IgemmCfg c = choose_igemm_frame(q, m);
            if (c.MB == 0) return false;
            Phase f;
            f.i0 = i0;
            f.j0 = j0;
            f.step = 2;
            f.FHf = g.FH;
            f.FWf = g.FW;
            f.flip = g.flip;
            f.oRow0 = pr;
            f.oCol0 = pc;
            f.oStride = 2;
            f.OHf = g.OH;
            f.OWf = g.OW;
            ps.g[ps.n] = q;
            ps.c[ps.n] = c;
            ps.p[ps.n] = f;
            ps.n++;
        }
    if (ps.n == 0) return false;
    *out = ps;
    return true;
}

// All four phases of a 2x up-sampled plan in one launch (k_igemm_f32_up2q):
// the window of stored samples a tile of 8 x 32 phase pixels needs for every
// phase, and the output block (MB) and plane chunk (RC) whose staging (the
// window and the whole filter) fits in `limit` bytes of LDS.
bool choose_igemm_up2q(const Geom& g, const PlaneMap& m, const PhaseSet& ps, UpFused* out, int mbOnly = 0,
                       int rcOnly = 0, size_t limit = 80 * 1024, bool two = false) {
    if (ps.n != 4 || g.FH != g.FW || !igemm_f32_up2q_supported(g.FH) || (g.hR & 1) != (g.hC & 1)) return false;
    UpFused u;
    std::memset(&u, 0, sizeof u);
    u.hp = g.hR & 1;
    const int F = g.FH;
    int hr[2], hc[2];
    for (int p = 0; p < 2; p++) {
        hr[p] = ps.g[2 * p].hR;
        hc[p] = ps.g[p].hC;
        u.OHp[p] = ps.g[2 * p].OH;
        u.OWp[p] = ps.g[p].OW;
    }
    for (int pr = 0; pr < 2; pr++)
        for (int pc = 0; pc < 2; pc++) {
            const Geom& q = ps.g[2 * pr + pc];
            if (q.FH != up2_taps(F, u.hp, pr) || q.FW != up2_taps(F, u.hp, pc) || q.hR != hr[pr] || q.hC != hc[pc])
                return false;
        }
    u.hU = std::max(hr[0], hr[1]);
    u.wU = std::max(hc[0], hc[1]);
    int UR = 0, UC = 0;
    for (int p = 0; p < 2; p++) {
        u.dr[p] = u.hU - hr[p];
        u.dc[p] = u.wU - hc[p];
        UR = std::max(UR, u.dr[p] + up2_taps(F, u.hp, p));
        UC = std::max(UC, u.dc[p] + up2_taps(F, u.hp, p));
    }
    IgemmCfg& c = u.c;
    c.variant = IV_F32_FRAME;
    c.Mout = m.T / m.N;
    c.tilesR = (std::max(u.OHp[0], u.OHp[1]) + 7) / 8;
    c.tilesC = (std::max(u.OWp[0], u.OWp[1]) + 31) / 32;
    c.tileRows = 7 + UR;
    c.tileCols = 31 + UC;
    if (c.tileCols > 64) return false;
    int pstr = c.tileRows * c.tileCols;
    pstr += ((16 - pstr % 32) + 32) % 32;  // 16 mod 32 words: the four k-lane groups hit different banks
    c.planeStride = pstr;
    const int maxRC = std::min(32, ((m.R + 7) / 8) * 8);
    if (two) limit = std::min(limit, (size_t)80 * 1024);
    for (int MB = c.Mout <= 16 || two ? 1 : 2; MB >= 1; MB--) {
        if (mbOnly && MB != mbOnly) continue;
        int as = F * F * 16 * MB;
        as += ((16 - as % 32) + 32) % 32;
        int best = 0;
        long bestSlots = 0;
        size_t bestBytes = 0;
        for (int rc = maxRC; rc >= 4; rc -= 4) {
            if (rcOnly && rc != rcOnly) continue;
            const size_t bytes = ((size_t)rc * (pstr + as) + 128) * sizeof(float);
            if (bytes > limit) continue;
            const long slots = (long)((m.R + rc - 1) / rc) * rc;
            if (!best || slots < bestSlots) {
                best = rc;
                bestSlots = slots;
                bestBytes = bytes;
            }
        }
        if (best) {
            c.MB = MB;
            c.RC = best;
            c.aStride = as;
            c.lds_bytes = bestBytes;
            c.mTiles = (c.Mout + 16 * MB - 1) / (16 * MB);
            u.on = 1;
            u.two = two ? 1 : 0;
            *out = u;
            return true;
        }
    }
    return false;
}

// The BF16 / FP16 form of the fused up-sampled launch (k_up2q_lp): the same
// window as the FP32 form, 32-plane chunks, the whole filter slice of a chunk
// in LDS beside the window (one block a CU).
bool choose_up2q_lp(const cogconv_desc* d, const Geom& g, const PlaneMap& m, const PhaseSet& ps, UpFused* out,
                    int mbOnly) {
    if (d->precision != COGCONV_BF16 && d->precision != COGCONV_FP16) return false;
    UpFused u;
    if (!choose_igemm_up2q(g, m, ps, &u, 1, 8, 160 * 1024)) return false;  // the window and phase offsets
    if ((long)(m.N + 8) * g.H * g.W * 4 >= (1L << 31)) return false;  // staging offsets 8 planes past the image
    if (u.c.tileRows > 64) return false;                                  // one lane per window row
    IgemmCfg& c = u.c;
    c.RC = (m.R + 31) / 32 * 32;
    for (int MB = c.Mout <= 16 ? 1 : 2; MB >= 1; MB--) {
        if (mbOnly && MB != mbOnly) continue;
        c.MB = MB;
        c.mTiles = (c.Mout + 16 * MB - 1) / (16 * MB);
        c.aStride = 0;
        c.lds_bytes = up2q_lp_lds_bytes(c, g.FH);
        if (c.lds_bytes <= 160 * 1024 && up2q_lp_workspace_bytes(c, g.FH) < (1UL << 31)) {
            *out = u;
            return true;
        }
    }
    return false;
}

void describe_into(const cogconv_plan_s* plan, char* buf, size_t len);

// The candidate configurations for `desc`, the heuristic choice first.  The
// list depends on the descriptor only, so an index into it is a stable
// autotuning result (the tuner also keeps the description to check it).
void enumerate_variants(const cogconv_desc* desc, const Geom& g, const PlaneMap& m, std::vector<Variant>& out) {
    out.clear();
    const long outPts = (long)g.OH * g.OW;
    std::vector<std::string> seen;
    auto add = [&](int kind, const DirectCfg* dc, const SmallCfg* sc, const IgemmCfg* ic, const FaCfg* fc) {
        Variant v;
        std::memset(&v, 0, sizeof v);
        v.kind = kind;
        if (dc) v.d = *dc;
        if (sc) v.s = *sc;
        if (ic) v.i = *ic;
        if (fc) v.f = *fc;
        switch (kind) {
        case KK_IGEMM_LP: v.workspace = igemm_lp_workspace_bytes(v.i, g.FH, desc->precision); break;
        case KK_FA_IGEMM:
        case KK_FA_IGEMM_LP: v.workspace = (size_t)v.f.splits * v.f.K * m.N * outPts * sizeof(float); break;
        case KK_SMALL_OUTPUT: v.workspace = (size_t)v.s.splits * m.nOut * outPts * sizeof(float); break;
        default: v.workspace = 0; break;
        }
        cogconv_plan_s tmp;
        std::memset(&tmp, 0, sizeof tmp);
        tmp.desc = *desc;
        tmp.g = g;
        tmp.m = m;
        tmp.kind = v.kind;
        tmp.dcfg = v.d;
        tmp.scfg = v.s;
        tmp.icfg = v.i;
        tmp.fcfg = v.f;
        tmp.ps = v.ps;
        char name[192];
        describe_into(&tmp, name, sizeof name);
        for (const std::string& n : seen)
            if (n == name) return;
        seen.emplace_back(name);
        out.push_back(v);
    };
    auto add_up = [&](const PhaseSet& ps) {
        Variant v;
        std::memset(&v, 0, sizeof v);
        v.kind = KK_IGEMM_UP;
        v.ps = ps;
        v.workspace = 0;
        out.push_back(v);
    };
    auto add_fft = [&](const FftCfg& c) {
        Variant v;
        std::memset(&v, 0, sizeof v);
        v.kind = KK_FFT;
        v.fft = c;
        v.workspace = fft_workspace_bytes(c);
        out.push_back(v);
    };
    if (desc->precision != COGCONV_FP32) {
        // 2x up-sampled frame modes: the fused four-phase launch in BF16 / FP16.
        PhaseSet up;
        if (choose_igemm_up(desc, g, m, &up)) {
            bool any = false;
            for (int mb : {2, 1}) {
                PhaseSet q = up;
                if (!choose_up2q_lp(desc, g, m, up, &q.f, mb)) continue;
                bool dup = false;
                for (const Variant& v : out) dup = dup || (v.kind == KK_IGEMM_UP && v.ps.f.c.MB == q.f.c.MB);
                if (dup) continue;
                Variant v;
                std::memset(&v, 0, sizeof v);
                v.kind = KK_IGEMM_UP;
                v.ps = q;
                v.workspace = up2q_lp_workspace_bytes(q.f.c, g.FH);
                out.push_back(v);
                any = true;
            }
            if (any) return;
        }
        FaCfg flp, frl;
        const bool rows = choose_fa_rows_lp(desc, g, m, &frl);
        if (rows) {
            add(KK_FA_IGEMM_LP, nullptr, nullptr, nullptr, &frl);
            static const int tiles[][3] = {{4, 2, 2}, {4, 1, 4}, {4, 2, 1}, {2, 2, 2}, {2, 2, 1}, {1, 4, 2}, {1, 4, 1}};
            for (const auto& t : tiles)
                for (int rsp : {1, 2, 4, 8}) {
                    FaCfg c2;
                    if (choose_fa_rows_lp(desc, g, m, &c2, t[0], t[1], t[2], rsp) &&
                        (c2.nb != frl.nb || c2.kh != frl.kh || c2.kbw != frl.kbw || c2.rowSplits != frl.rowSplits))
                        add(KK_FA_IGEMM_LP, nullptr, nullptr, nullptr, &c2);
                }
        }
        if (choose_fa_igemm_lp(desc, g, m, &flp)) {
            add(KK_FA_IGEMM_LP, nullptr, nullptr, nullptr, &flp);
            for (int cw : {0, 128, 64, 32})
                for (int cbw : {3, 6, 9})
                    for (long parts : {256L, 512L})
                        for (int nw : {4, 8}) {
                            FaCfg c2;
                            if (choose_fa_igemm_lp(desc, g, m, &c2, cbw, parts, cw, nw))
                                add(KK_FA_IGEMM_LP, nullptr, nullptr, nullptr, &c2);
                        }
            return;
        }
        if (rows) return;
        IgemmCfg lp;
        if (!choose_igemm_lp(desc, g, m, &lp)) return;
        add(KK_IGEMM_LP, nullptr, nullptr, &lp, nullptr);
        const int mbMax = lp.MB;
        for (int mb : {4, 2, 1}) {
            if (mb < mbMax && choose_igemm_lp(desc, g, m, &lp, mb)) add(KK_IGEMM_LP, nullptr, nullptr, &lp, nullptr);
        }
        // One tap row of the filter at a time, double-buffered: a smaller LDS
        // footprint (two blocks a CU for 7x7 / 9x9 at MB <= 2) and the next
        // row's loads behind the current row's MFMAs.
        for (int mb : {4, 2, 1}) {
            if (mb <= mbMax && g.FH > 1 && desc->precision != COGCONV_SPLIT_BF16 &&
                choose_igemm_lp(desc, g, m, &lp, mb, true))
                add(KK_IGEMM_LP, nullptr, nullptr, &lp, nullptr);
        }
        return;
    }
    const bool small_legal = outPts <= 4096;
    const bool small_out = g.border == COGCONV_BORDER_VALID && is_fa(desc->mode) && outPts <= 256 &&
                           (long)g.FHe * g.FWe >= 64;
    const IgemmCfg ig = igemm_frame_eligible(desc, g) ? choose_igemm_frame(g, m) : IgemmCfg{};
    const bool igemm_ok = ig.MB > 0;
    FaCfg fa;
    const bool fa_ok = choose_fa_igemm(desc, g, m, &fa);
    const SmallCfg sm = choose_small(g, m);
    const DirectCfg dr = choose_direct(g, 4);
    DirectCfg drt[3];
    const bool rt_ok = direct_rt_supported(g);
    int rt_first = 0;
    if (rt_ok) {
        for (int i = 0; i < 3; i++) direct_rt_config(g, 4 >> i, drt[i]);
        // The largest tile that still gives two blocks per CU.
        while (rt_first < 2 && (long)drt[rt_first].tilesR * drt[rt_first].tilesC * m.nOut < 512) rt_first++;
    }
    PhaseSet up;
    const bool up_ok = choose_igemm_up(desc, g, m, &up);
    FftCfg fc;
    const bool fft_ok = choose_fft(desc, g, m, &fc);
    // The FFT route first for filters the register-tiled kernel does not take and
    // that are at least 31 x 31 (below that the direct kernels measure faster);
    // the autotuner times both routes.
    const bool fft_first = fft_ok && !rt_ok && (long)g.FH * g.FW >= 31L * 31;
    // Heuristic default.
    // The fused up-sampled launch first: all four phases from one staged window,
    // one output block and two blocks a CU (2x 7x7, 32 -> 32 planes of 256 x 256:
    // 0.49 ms, against 0.54 ms for two output blocks at one block a CU and
    // 0.74 ms for the four phase launches).
    PhaseSet upq = up;
    const bool upq_ok = up_ok && (choose_igemm_up2q(g, m, up, &upq.f, 1, 0, 80 * 1024, true) ||
                                  choose_igemm_up2q(g, m, up, &upq.f));
    if (fft_first) add_fft(fc);
    else if (upq_ok) add_up(upq);
    else if (up_ok) add_up(up);
    else if (igemm_ok) add(KK_IGEMM, nullptr, nullptr, &ig, nullptr);
    else if (fa_ok) add(KK_FA_IGEMM, nullptr, nullptr, nullptr, &fa);
    else if (small_out) add(KK_SMALL_OUTPUT, nullptr, &sm, nullptr, nullptr);
    else if (rt_ok) add(KK_DIRECT_RT, &drt[rt_first], nullptr, nullptr, nullptr);
    else add(KK_DIRECT, &dr, nullptr, nullptr, nullptr);
    // Alternatives.
    if (igemm_ok) {
        for (int mb : {4, 2, 1})
            for (int rc : {32, 16, 8}) {
                const IgemmCfg c = choose_igemm_frame(g, m, mb, rc);
                if (c.MB > 0) add(KK_IGEMM, nullptr, nullptr, &c, nullptr);
            }
    }
    if (fa_ok) {
        for (int cw : {0, 32, 16})
            for (int cbw : {3, 6, 9})
                for (int parts : {128, 256, 512, 1024})
                    for (int nw : {4, 8}) {
                        FaCfg c;
                        if (choose_fa_igemm(desc, g, m, &c, cbw, parts, cw, nw))
                            add(KK_FA_IGEMM, nullptr, nullptr, nullptr, &c);
                    }
    }
    if (small_legal && (long)m.R * g.FHe * g.FWe >= 64) {
        for (long blocks : {2048L, 4096L, 8192L}) {
            const SmallCfg c = choose_small(g, m, blocks);
            add(KK_SMALL_OUTPUT, nullptr, &c, nullptr, nullptr);
        }
    }
    if (rt_ok)
        for (const DirectCfg& c : drt) add(KK_DIRECT_RT, &c, nullptr, nullptr, nullptr);
    for (int rg : {4, 2, 1}) {
        const DirectCfg c = choose_direct(g, rg);
        add(KK_DIRECT, &c, nullptr, nullptr, nullptr);
    }
    if (upq_ok) {
        add_up(up);
        for (int two : {0, 1})
            for (size_t limit : {(size_t)80 * 1024, (size_t)160 * 1024})
                for (int mb : {2, 1})
                    for (int rc : {32, 16, 8}) {
                        PhaseSet q = up;
                        if (!choose_igemm_up2q(g, m, up, &q.f, mb, rc, limit, two)) continue;
                        bool dup = false;
                        for (const Variant& v : out)
                            dup = dup || (v.kind == KK_IGEMM_UP && v.ps.f.on && v.ps.f.c.MB == q.f.c.MB &&
                                          v.ps.f.c.RC == q.f.c.RC && v.ps.f.two == q.f.two);
                        if (!dup) add_up(q);
                    }
    }
    if (fft_ok && !fft_first) add_fft(fc);
}

bool kind_matches_algo(int kind, int algo) {
    switch (algo) {
    case COGCONV_ALGO_DIRECT: return kind == KK_DIRECT || kind == KK_DIRECT_RT || kind == KK_FFT;
    case COGCONV_ALGO_SMALL_OUTPUT: return kind == KK_SMALL_OUTPUT;
    case COGCONV_ALGO_IMPLICIT_GEMM:
        return kind == KK_IGEMM || kind == KK_FA_IGEMM || kind == KK_IGEMM_LP || kind == KK_FA_IGEMM_LP ||
               kind == KK_IGEMM_UP;
    default: return true;
    }
}

// Builds the plan for variant `vi` of `desc` (vi < 0: by algorithm hint,
// then by the autotuning cache, then the heuristic default).
int make_plan(const cogconv_desc* desc, int vi, cogconv_plan* plan) {
    if (!plan) return fail(COGCONV_ERR_INVALID_ARG, "null plan pointer");
    *plan = nullptr;
    cogconv_field_type out;
    Geom g;
    PlaneMap m;
    int e = analyse(desc, &out, &g, &m);
    if (e) return e;
    std::vector<Variant> vars;
    enumerate_variants(desc, g, m, vars);
    if (vars.empty())
        return fail(COGCONV_ERR_UNSUPPORTED,
                    "precision %d is offered for ProjectFrame/BackProjectFrame with block reduction on 2-D fields with "
                    "a square 1/3/5/7/9 filter and a 32-column tile that fits one wavefront, without input "
                    "up-sampling or (BF16 / FP16, odd filters from 3) with 2x up-sampling, and for FilterAdjoint "
                    "with block reduction, BorderValid and at most 64 planes per image in dy; not for this %s",
                    desc->precision, mode_name(desc->mode));
    const long outPts = (long)g.OH * g.OW;
    int pick = -1;
    if (vi >= 0) {
        if (vi >= (int)vars.size())
            return fail(COGCONV_ERR_INVALID_ARG, "variant %d out of range (%d variants)", vi, (int)vars.size());
        pick = vi;
    } else if (desc->algo != COGCONV_ALGO_AUTO) {
        if (desc->algo == COGCONV_ALGO_SMALL_OUTPUT && outPts > 4096)
            return fail(COGCONV_ERR_UNSUPPORTED, "small-output kernel needs an output plane of at most 4096 points");
        for (int i = 0; i < (int)vars.size() && pick < 0; i++)
            if (kind_matches_algo(vars[i].kind, desc->algo)) pick = i;
        if (pick < 0)
            return fail(COGCONV_ERR_UNSUPPORTED,
                        "implicit GEMM needs a block-reduce frame mode with a square 1/3/5/7/9 filter on 2-D fields "
                        "and no input up-sampling, or FilterAdjoint with block reduction, BorderValid and at most 64 "
                        "planes per image in dy");
    } else {
        pick = 0;
        int tuned;
        std::string tunedName;
        if (tune_cache_lookup(desc, &tuned, &tunedName) && tuned < (int)vars.size()) {
            // Accept the cached index only if it still names the same configuration.
            cogconv_plan_s tmp;
            std::memset(&tmp, 0, sizeof tmp);
            tmp.desc = *desc;
            tmp.g = g;
            tmp.m = m;
            tmp.kind = vars[tuned].kind;
            tmp.dcfg = vars[tuned].d;
            tmp.scfg = vars[tuned].s;
            tmp.icfg = vars[tuned].i;
            tmp.fcfg = vars[tuned].f;
            tmp.ps = vars[tuned].ps;
            tmp.fft = vars[tuned].fft;
            char name[192];
            describe_into(&tmp, name, sizeof name);
            if (tunedName == name) pick = tuned;
        }
    }
    cogconv_plan p = new (std::nothrow) cogconv_plan_s;
    if (!p) return fail(COGCONV_ERR_OOM, "host allocation failed");
    std::memset(p, 0, sizeof *p);
    p->desc = *desc;
    p->out = out;
    p->g = g;
    p->m = m;
    const Variant& v = vars[pick];
    p->kind = v.kind;
    p->dcfg = v.d;
    p->scfg = v.s;
    p->icfg = v.i;
    p->fcfg = v.f;
    p->ps = v.ps;
    p->fft = v.fft;
    p->workspace = v.workspace;
    p->variant = pick;
    if (p->kind == KK_FFT && !(p->fft_state = fft_state_new())) {
        delete p;
        return fail(COGCONV_ERR_OOM, "host allocation failed");
    }
    // Algorithmic work: every multiply-add whose tap multiplies a stored
    // (non-inserted) sample.  Up-sampled operand 1 leaves FH*FW/(upAr*upAc)
    // such taps per output on average; FilterAdjoint up-sampling of operand 2
    // leaves its FH*FW stored taps.
    double taps = (double)g.FH * g.FW / ((double)g.upAr * g.upAc);
    p->flops = 2.0 * m.nOut * (double)m.R * (double)outPts * taps;
    p->bytes = 4.0 * ((double)desc->in0.planes * desc->in0.rows * desc->in0.cols +
                      (double)desc->in1.planes * desc->in1.rows * desc->in1.cols + (double)m.nOut * outPts);
    *plan = p;
    return COGCONV_OK;
}

}  // namespace

int cogconv::plan_variant_count(const cogconv_desc* desc, int* count) {
    cogconv_field_type out;
    Geom g;
    PlaneMap m;
    int e = analyse(desc, &out, &g, &m);
    if (e) return e;
    std::vector<Variant> vars;
    enumerate_variants(desc, g, m, vars);
    *count = (int)vars.size();
    return COGCONV_OK;
}

extern "C" int cogconv_output_type(const cogconv_desc* desc, cogconv_field_type* out) {
    clear_error();
    return analyse(desc, out, nullptr, nullptr);
}

extern "C" int cogconv_plan_create(const cogconv_desc* desc, cogconv_plan* plan) {
    clear_error();
    return make_plan(desc, -1, plan);
}

extern "C" int cogconv_plan_variant_count(const cogconv_desc* desc, int32_t* count) {
    clear_error();
    if (!count) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    int n = 0;
    const int e = plan_variant_count(desc, &n);
    if (e) return e;
    *count = n;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_create_variant(const cogconv_desc* desc, int32_t variant, cogconv_plan* plan) {
    clear_error();
    if (variant < 0) return fail(COGCONV_ERR_INVALID_ARG, "negative variant index");
    return make_plan(desc, variant, plan);
}

extern "C" int cogconv_plan_variant(cogconv_plan plan, int32_t* variant) {
    clear_error();
    if (!plan || !variant) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *variant = plan->variant;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_destroy(cogconv_plan plan) {
    clear_error();
    if (plan) fft_release(plan->fft_state);
    delete plan;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_output_type(cogconv_plan plan, cogconv_field_type* out) {
    clear_error();
    if (!plan || !out) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *out = plan->out;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_workspace_bytes(cogconv_plan plan, size_t* bytes) {
    clear_error();
    if (!plan || !bytes) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *bytes = plan->workspace;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_flops(cogconv_plan plan, double* flops) {
    clear_error();
    if (!plan || !flops) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *flops = plan->flops;
    return COGCONV_OK;
}

extern "C" int cogconv_plan_bytes(cogconv_plan plan, double* bytes) {
    clear_error();
    if (!plan || !bytes) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *bytes = plan->bytes;
    return COGCONV_OK;
}

namespace {
void describe_into(const cogconv_plan_s* plan, char* buf, size_t len) {
    const Geom& g = plan->g;
    switch (plan->kind) {
    case KK_DIRECT:
        snprintf(buf, len, "direct %s out=%dx%dx%d tile=%dx64 fchunk=%dx%d lds=%zu", mode_name(plan->desc.mode),
                 plan->m.nOut, g.OH, g.OW, 4 * plan->dcfg.RG, plan->dcfg.FCR, plan->dcfg.FCC, plan->dcfg.lds_bytes);
        break;
    case KK_DIRECT_RT:
        snprintf(buf, len, "direct_rt %s out=%dx%dx%d tile=%dx128 thread=%dx4 filter=%dx%d lds=%zu",
                 mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW, 8 * plan->dcfg.RG, plan->dcfg.RG, g.FH, g.FW,
                 plan->dcfg.lds_bytes);
        break;
    case KK_SMALL_OUTPUT:
        snprintf(buf, len, "small_output %s out=%dx%dx%d splits=%d", mode_name(plan->desc.mode), plan->m.nOut, g.OH,
                 g.OW, plan->scfg.splits);
        break;
    case KK_IGEMM_LP: {
        static const char* pn[] = {"fp32", "bf16", "fp16", "split_bf16"};
        snprintf(buf, len, "igemm_%s %s out=%dx%dx%d MB=%d tile=8x32 threads=256%s taprows=%d%s lds=%zu",
                 pn[plan->desc.precision], mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW, plan->icfg.MB,
                 plan->icfg.vec4 ? " ld16" : "", plan->icfg.trg, plan->icfg.ring ? " ring" : "", plan->icfg.lds_bytes);
        break;
    }
    case KK_IGEMM_UP: {
        if (plan->ps.f.on) {
            static const char* pn[] = {"f32", "bf16", "fp16", "split_bf16"};
            const IgemmCfg& c = plan->ps.f.c;
            snprintf(buf, len, "igemm_%s_up2 %s out=%dx%dx%d phases=4 fused window=%dx%d MB=%d RC=%d%s lds=%zu",
                     pn[plan->desc.precision], mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW, c.tileRows,
                     c.tileCols, c.MB, c.RC, plan->ps.f.two ? " 2/CU" : "", c.lds_bytes);
            break;
        }
        int n = snprintf(buf, len, "igemm_f32_up2 %s out=%dx%dx%d phases=%d", mode_name(plan->desc.mode), plan->m.nOut,
                         g.OH, g.OW, plan->ps.n);
        for (int i = 0; i < plan->ps.n && n > 0 && (size_t)n < len; i++)
            n += snprintf(buf + n, len - n, " %dx%d:MB=%d,RC=%d", plan->ps.g[i].FH, plan->ps.g[i].FW,
                          plan->ps.c[i].MB, plan->ps.c[i].RC);
        break;
    }
    case KK_FA_IGEMM_LP: {
        static const char* pn[] = {"fp32", "bf16", "fp16", "split_bf16"};
        if (plan->fcfg.rows) {
            const FaCfg& f = plan->fcfg;
            snprintf(buf, len, "fa_rows_%s %s out=%dx%dx%d tile=%dx%d waves=%d chunks=%d rowsplits=%d splits=%d lds=%zu",
                     pn[plan->desc.precision], mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW,
                     f.kh * f.kbw * 16, f.nb * 16, f.nb * f.kh, f.chunks, f.rowSplits, f.splits, f.lds_bytes);
            break;
        }
        snprintf(buf, len, "fa_igemm_%s %s out=%dx%dx%d MB=%d CBW=%d RR=%d CW=%d groups=%d splits=%d waves=%d lds=%zu",
                 pn[plan->desc.precision], mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW, plan->fcfg.MB,
                 plan->fcfg.CBW, plan->fcfg.RR, plan->fcfg.CW, plan->fcfg.colGroups, plan->fcfg.splits,
                 plan->fcfg.NW, plan->fcfg.lds_bytes);
        break;
    }
    case KK_FFT:
        snprintf(buf, len, "fft %s out=%dx%dx%d filter=%dx%d transform=%dx%d", mode_name(plan->desc.mode),
                 plan->m.nOut, g.OH, g.OW, g.FH, g.FW, plan->fft.Er, plan->fft.Ec);
        break;
    case KK_FA_IGEMM:
        snprintf(buf, len, "fa_igemm_f32 %s out=%dx%dx%d MB=%d CBW=%d RR=%d CW=%d groups=%d splits=%d waves=%d lds=%zu",
                 mode_name(plan->desc.mode), plan->m.nOut, g.OH, g.OW, plan->fcfg.MB, plan->fcfg.CBW, plan->fcfg.RR,
                 plan->fcfg.CW, plan->fcfg.colGroups, plan->fcfg.splits, plan->fcfg.NW, plan->fcfg.lds_bytes);
        break;
    default:
        snprintf(buf, len, "igemm_f32 %s out=%dx%dx%d MB=%d RC=%d tile=8x32 threads=512 lds=%zu", mode_name(plan->desc.mode),
                 plan->m.nOut, g.OH, g.OW, plan->icfg.MB, plan->icfg.RC, plan->icfg.lds_bytes);
        break;
    }
}
}  // namespace

extern "C" int cogconv_plan_describe(cogconv_plan plan, char* buf, size_t len) {
    clear_error();
    if (!plan || !buf || len == 0) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    describe_into(plan, buf, len);
    return COGCONV_OK;
}

extern "C" int cogconv_plan_pack_operand(cogconv_plan plan, const float* in1, void* workspace, cogconv_stream stream,
                                         int32_t* packed) {
    clear_error();
    if (!plan || !packed) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *packed = 0;
    if (plan->kind != KK_IGEMM_LP) return COGCONV_OK;
    if (!in1 || !workspace) return fail(COGCONV_ERR_INVALID_ARG, "null operand or workspace");
    const hipError_t e = launch_igemm_lp(plan->desc.precision, plan->g, plan->m, plan->icfg, nullptr, in1, nullptr,
                                         workspace, reinterpret_cast<hipStream_t>(stream), LP_PACK_ONLY);
    if (int st = check_hip(e, "operand pack")) return st;
    *packed = 1;
    return COGCONV_OK;
}

std::string cogconv_plan_pack_key(const cogconv_plan_s* plan) {
    if (!plan || plan->kind != KK_IGEMM_LP) return std::string();
    // What k_pack_filters_lp reads: precision, mode, orientation, taps, planes
    // and the padded output / input-plane counts of the kernel configuration.
    const int key[] = {plan->desc.precision, plan->m.mode, plan->g.flip, plan->g.FH, plan->g.FW, plan->m.N,
                       plan->m.T, plan->m.R, plan->icfg.Mout, plan->icfg.mTiles * 16 * plan->icfg.MB, plan->icfg.RC};
    return std::string(reinterpret_cast<const char*>(key), sizeof key);
}

extern "C" int cogconv_launch_packed(cogconv_plan plan, const float* in0, float* out, void* workspace,
                                     cogconv_stream stream) {
    clear_error();
    if (!plan || !in0 || !out || !workspace) return fail(COGCONV_ERR_INVALID_ARG, "null plan or buffer");
    if (plan->kind != KK_IGEMM_LP) return fail(COGCONV_ERR_INVALID_ARG, "plan does not pack operand 2");
    const hipError_t e = launch_igemm_lp(plan->desc.precision, plan->g, plan->m, plan->icfg, in0, nullptr, out,
                                         workspace, reinterpret_cast<hipStream_t>(stream), LP_RUN_ONLY);
    return check_hip(e, "convolution launch");
}

namespace {

// Every kernel of the plan's launch sequence, with epilogue `ep` on the final stores.
int launch_all(cogconv_plan plan, const float* in0, const float* in1, float* out, void* workspace, hipStream_t s,
               const Epi& ep) {
    hipError_t e = hipSuccess;
    switch (plan->kind) {
    case KK_DIRECT:
        e = launch_direct(plan->g, plan->m, plan->dcfg, in0, in1, out, s, ep);
        break;
    case KK_DIRECT_RT:
        e = launch_direct_rt(plan->g, plan->m, plan->dcfg, in0, in1, out, s, ep);
        break;
    case KK_SMALL_OUTPUT:
        e = launch_small_output(plan->g, plan->m, plan->scfg, in0, in1, out, static_cast<float*>(workspace), s, ep);
        break;
    case KK_IGEMM:
        e = launch_igemm_f32(plan->g, plan->m, plan->icfg, in0, in1, out, s, ep);
        break;
    case KK_FA_IGEMM:
        e = launch_fa_igemm_f32(plan->g, plan->m, plan->fcfg, in0, in1, out, static_cast<float*>(workspace), s, ep);
        break;
    case KK_IGEMM_LP:
        e = launch_igemm_lp(plan->desc.precision, plan->g, plan->m, plan->icfg, in0, in1, out, workspace, s,
                            LP_PACK_AND_RUN, ep);
        break;
    case KK_IGEMM_UP:
        if (plan->ps.f.on && plan->desc.precision != COGCONV_FP32) {
            e = launch_up2q_lp(plan->desc.precision, plan->g, plan->m, plan->ps.f, in0, in1, out, workspace, s, ep);
            break;
        }
        if (plan->ps.f.on) {
            e = launch_igemm_f32_up2q(plan->g, plan->m, plan->ps.f, in0, in1, out, s, ep);
            break;
        }
        for (int i = 0; i < plan->ps.n && e == hipSuccess; i++)
            e = launch_igemm_f32_phase(plan->ps.g[i], plan->m, plan->ps.c[i], plan->ps.p[i], in0, in1, out, s, ep);
        break;
    case KK_FA_IGEMM_LP:
        if (plan->fcfg.rows) {
            e = launch_fa_rows_lp(plan->desc.precision, plan->g, plan->m, plan->fcfg, in0, in1, out,
                                  static_cast<float*>(workspace), s, ep);
        } else {
            e = launch_fa_igemm_lp(plan->desc.precision, plan->g, plan->m, plan->fcfg, in0, in1, out,
                                   static_cast<float*>(workspace), s, ep);
        }
        break;
    case KK_FFT:
        return launch_fft(plan->g, plan->fft, &plan->fft_state, in0, in1, out, workspace, s, ep);
    default:
        return fail(COGCONV_ERR_RUNTIME, "internal: unknown kernel kind %d", plan->kind);
    }
    return check_hip(e, "convolution launch");
}

// The plan's epilogue with its field operands bound to `operands`.
int bind_epilogue(const cogconv_plan_s* plan, const float* const* operands, Epi* ep) {
    *ep = plan->epi;
    return epi_bind(ep, plan->epi_operands, operands);
}

}  // namespace

extern "C" int cogconv_launch(cogconv_plan plan, const float* in0, const float* in1, float* out, void* workspace,
                              cogconv_stream stream) {
    clear_error();
    if (!plan || !in0 || !in1 || !out) return fail(COGCONV_ERR_INVALID_ARG, "null plan or buffer");
    if (plan->workspace && !workspace)
        return fail(COGCONV_ERR_INVALID_ARG, "plan needs %zu bytes of workspace", plan->workspace);
    if (plan->epi_operands)
        return fail(COGCONV_ERR_INVALID_ARG, "the plan's epilogue reads field operands: use cogconv_launch_epilogue");
    return launch_all(plan, in0, in1, out, workspace, reinterpret_cast<hipStream_t>(stream), plan->epi);
}

bool cogconv::plan_table_capable(const cogconv_plan_s* plan) {
    return plan->kind == KK_IGEMM_LP || (plan->kind == KK_IGEMM && plan->m.mode != COGCONV_PLANE_BY_PLANE);
}

int cogconv::launch_table(cogconv_plan plan, const cogconv_image_io* device_table, bool x16, const float* in1,
                          void* workspace, bool packed, hipStream_t s) {
    hipError_t e;
    if (plan->kind == KK_IGEMM_LP)
        e = launch_igemm_lp(plan->desc.precision, plan->g, plan->m, plan->icfg, nullptr, in1, nullptr, workspace, s,
                            packed ? LP_RUN_ONLY : LP_PACK_AND_RUN, plan->epi, device_table, x16);
    else
        e = launch_igemm_f32(plan->g, plan->m, plan->icfg, nullptr, in1, nullptr, s, plan->epi, device_table);
    return check_hip(e, "table launch");
}

extern "C" int cogconv_plan_table_supported(cogconv_plan plan, int32_t* supported) {
    clear_error();
    if (!plan || !supported) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *supported = plan_table_capable(plan) ? 1 : 0;
    return COGCONV_OK;
}

extern "C" int cogconv_launch_table(cogconv_plan plan, const cogconv_image_io* host_table,
                                    cogconv_image_io* device_table, const float* in1, void* workspace, int32_t packed,
                                    cogconv_stream stream) {
    clear_error();
    if (!plan || !host_table || !device_table) return fail(COGCONV_ERR_INVALID_ARG, "null plan or table");
    if (!plan_table_capable(plan))
        return fail(COGCONV_ERR_UNSUPPORTED, "table launches are offered for the MFMA frame plans (no up-sampling)");
    if (packed && plan->kind != KK_IGEMM_LP) return fail(COGCONV_ERR_INVALID_ARG, "plan does not pack operand 2");
    if (!packed && !in1) return fail(COGCONV_ERR_INVALID_ARG, "null operand 2");
    if ((plan->workspace || packed) && !workspace)
        return fail(COGCONV_ERR_INVALID_ARG, "plan needs %zu bytes of workspace", plan->workspace);
    if (plan->epi_operands)
        return fail(COGCONV_ERR_INVALID_ARG, "table launches take no epilogue field operands");
    bool x16 = true;
    for (int b = 0; b < plan->m.B; b++) {
        if (!host_table[b].in0 || !host_table[b].out)
            return fail(COGCONV_ERR_INVALID_ARG, "null buffer in table entry %d", b);
        x16 = x16 && (reinterpret_cast<uintptr_t>(host_table[b].in0) & 15) == 0;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (int st = check_hip(hipMemcpyAsync(device_table, host_table, sizeof(cogconv_image_io) * plan->m.B,
                                          hipMemcpyHostToDevice, s),
                           "table copy"))
        return st;
    return launch_table(plan, device_table, x16, in1, workspace, packed != 0, s);
}

extern "C" int cogconv_launch_epilogue(cogconv_plan plan, const float* in0, const float* in1, float* out,
                                       void* workspace, const float* const* operands, cogconv_stream stream) {
    clear_error();
    if (!plan || !in0 || !in1 || !out) return fail(COGCONV_ERR_INVALID_ARG, "null plan or buffer");
    if (plan->workspace && !workspace)
        return fail(COGCONV_ERR_INVALID_ARG, "plan needs %zu bytes of workspace", plan->workspace);
    Epi ep;
    if (int st = bind_epilogue(plan, operands, &ep)) return st;
    return launch_all(plan, in0, in1, out, workspace, reinterpret_cast<hipStream_t>(stream), ep);
}

namespace {

bool epi_op_known(int op) {
    return (op >= COGCONV_EPI_ADD_CONST && op <= COGCONV_EPI_MIN_CONST) ||
           (op >= COGCONV_EPI_NEG && op <= COGCONV_EPI_SIGNUM) ||
           (op >= COGCONV_EPI_ADD_FIELD && op <= COGCONV_EPI_MIN_FIELD);
}

// Checks `in` against a field of type `t` and builds the device form.  An
// operand must combine with `t` to give `t` under ResultType.scala:29-58: the
// same type, a 0-D field of its tensor shape, a tensor-order-0 field of its
// field shape, or a 0-D scalar.
}  // namespace

int cogconv::epi_build(const cogconv_field_type& t, const cogconv_epilogue* in, Epi* out, int* operands) {
    std::memset(out, 0, sizeof *out);
    *operands = 0;
    if (!in || in->n_ops == 0) return COGCONV_OK;
    if (in->n_ops < 0 || in->n_ops > COGCONV_EPI_MAX_OPS)
        return fail(COGCONV_ERR_INVALID_ARG, "epilogue has %d ops (at most %d)", in->n_ops, COGCONV_EPI_MAX_OPS);
    if (in->n_operands < 0 || in->n_operands > COGCONV_EPI_MAX_OPERANDS)
        return fail(COGCONV_ERR_INVALID_ARG, "epilogue has %d operands (at most %d)", in->n_operands,
                    COGCONV_EPI_MAX_OPERANDS);
    const int points = t.rows * t.cols;
    for (int s = 0; s < in->n_operands; s++) {
        const cogconv_field_type& f = in->operands[s];
        const bool zeroD = f.rows * f.cols == 1 && f.dims == 0;
        const bool sameField = f.dims == t.dims && f.rows == t.rows && f.cols == t.cols;
        const bool sameTensor = f.planes == t.planes && f.tensor_order == t.tensor_order;
        const bool scalar = f.tensor_order == 0 && f.planes == 1;
        if (!((sameField && sameTensor) || (zeroD && sameTensor) || (sameField && scalar) || (zeroD && scalar)))
            return fail(COGCONV_ERR_ILLEGAL_OP,
                        "epilogue operand %d (%dD %dx%d, %d planes) does not combine with the %dD %dx%d x %d-plane "
                        "output into the output's type",
                        s, f.dims, f.rows, f.cols, f.planes, t.dims, t.rows, t.cols, t.planes);
        out->srcPlanes[s] = scalar ? 1 : t.planes;
        out->srcPoints[s] = zeroD ? 1 : points;
    }
    for (int i = 0; i < in->n_ops; i++) {
        const cogconv_epilogue_op& o = in->ops[i];
        if (!epi_op_known(o.op)) return fail(COGCONV_ERR_INVALID_ARG, "unknown epilogue op %d", o.op);
        if (o.op >= COGCONV_EPI_ADD_FIELD && (o.operand < 0 || o.operand >= in->n_operands))
            return fail(COGCONV_ERR_INVALID_ARG, "epilogue op %d names operand %d of %d", i, o.operand,
                        in->n_operands);
        out->op[i] = o.op;
        out->val[i] = o.value;
        out->slot[i] = o.op >= COGCONV_EPI_ADD_FIELD ? o.operand : 0;
    }
    out->n = in->n_ops;
    *operands = in->n_operands;
    return COGCONV_OK;
}

int cogconv::epi_bind(Epi* e, int n, const float* const* operands) {
    if (n && !operands) return fail(COGCONV_ERR_INVALID_ARG, "the epilogue reads %d field operand(s)", n);
    for (int i = 0; i < n; i++) {
        if (!operands[i]) return fail(COGCONV_ERR_INVALID_ARG, "null epilogue operand %d", i);
        e->src[i] = operands[i];
    }
    return COGCONV_OK;
}

namespace {
}  // namespace

extern "C" int cogconv_plan_set_epilogue(cogconv_plan plan, const cogconv_epilogue* epi) {
    clear_error();
    if (!plan) return fail(COGCONV_ERR_INVALID_ARG, "null plan");
    Epi e;
    int n = 0;
    if (int st = epi_build(plan->out, epi, &e, &n)) return st;
    plan->epi = e;
    plan->epi_operands = n;
    return COGCONV_OK;
}

extern "C" int cogconv_elementwise(const float* in, float* out, const cogconv_field_type* type,
                                   const cogconv_epilogue* epi, const float* const* operands,
                                   cogconv_stream stream) {
    clear_error();
    if (!in || !out || !type) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    if (type->rows < 1 || type->cols < 1 || type->planes < 1) return fail(COGCONV_ERR_INVALID_ARG, "empty field");
    Epi e;
    int n = 0;
    if (int st = epi_build(*type, epi, &e, &n)) return st;
    if (int st = epi_bind(&e, n, operands)) return st;
    return check_hip(launch_elementwise(in, out, (int64_t)type->rows * type->cols, type->planes, e,
                                        reinterpret_cast<hipStream_t>(stream)),
                     "elementwise launch");
}
