// Precision variants of FilterAdjointBlockReduceSum with BorderValid (the
// weight gradient of a layer): BF16 / FP16 MFMA with FP32 accumulation, and
// split-BF16 (three-term compensated products).  Same contraction as
// fa_igemm_f32.hip:
//
//   dW[k*N + n](R, C) = sum_b sum_{i,j} tap_{b,k}(i, j) * X_{b*N + n}(R*dn + i, C*dn + j)
//
// (ConvolveHyperKernel.scala:1310-1336; ConvolveToSmallFieldPipelinedTiledHyperKernel.scala:45-127.)
//
// gfx950 mapping: v_mfma_f32_16x16x32_{bf16,f16}.  The 32-deep k slice is 32
// consecutive tap columns of one tap row; lane l holds A[l%16][8*(l/16) .. +7]
// (dy plane l%16, eight tap columns) and B[8*(l/16) .. +7][l%16] (the x
// samples under those taps for output column l%16).  Both are 16-byte LDS
// reads: dy rows are staged [tap row][plane][column]; x rows are staged once
// per output column C, shifted left by C*dn, so that the x samples of output
// (n, R, C) under tap column t sit at column t of copy C.  Operands arrive in
// FP32 and are rounded while staged.  A block (NW = 4 wave64, or 8 when MB =
// 4) owns a column group of CBG 16-column blocks and all MB row blocks; wave w
// reduces tap rows w%4, w%4+4, .. of each item (image, tap-column chunk,
// tap-row chunk), and with 8 waves the two wave quads (one wave of each per
// SIMD) own two row blocks each; the waves' sums meet in LDS and one partial
// dW per block is written, added afterwards in a fixed order
// (launch_reduce_partials).
#include "lp.h"
#include "stage.h"

namespace cogconv {


// Lane byte offset with bit 31 set: past every staged buffer (each is under
// 2^31 bytes), so the buffer unit's range check answers the load with 0.
constexpr int FA_OOB = (int)0x80000000u;

template <int MB, int CBG, int PREC, int NW>
__global__ __launch_bounds__(64 * NW) void k_fa_igemm_lp(Geom g, PlaneMap m, FaCfg c, const float* __restrict__ X,
                                                     const float* __restrict__ DY, float* __restrict__ ws) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr bool SPLIT = PREC == COGCONV_SPLIT_BF16;
    constexpr int TERMS = SPLIT ? 2 : 1;
    constexpr int MT = 16 * MB;
    extern __shared__ __attribute__((aligned(16))) unsigned char fa_lp_smem[];
    const int OW = g.OW;
    const int DP = c.DWp + 8;  // LDS row pitch in elements (DWp is 32, 64 or 128)
    T* sD = reinterpret_cast<T*>(fa_lp_smem);                  // [TERMS][RR][MT][DP]
    const size_t dTerm = (size_t)c.RR * MT * DP;
    T* sX = sD + TERMS * dTerm;                                  // [TERMS][nCntMax][xRows][OW][DP]
    const size_t xTerm = (size_t)c.nCntMax * c.xRows * OW * DP;
    T* sJunk = sX + TERMS * xTerm;                               // [TERMS][64 lanes][8]
    const int FHW = g.FH * g.FW;
    const int HWx = g.H * g.W;

    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int cg = bid % c.colGroups;
    const int s = bid / c.colGroups;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g4 = lane >> 4, l16 = lane & 15;
    const int N = m.N, K = c.K;
    const int P = g.OH * OW;
    const int Ncols = N * P;
    const int colBase = cg * CBG * 16;
    const int n_lo = colBase / P;
    const int n_hi = min(N - 1, (min(colBase + CBG * 16, Ncols) - 1) / P);
    const int nCnt = n_hi - n_lo + 1;

    // B-operand offset of each column block: copy C of x row R*dn (+ tap row).
    // One column per thread into LDS (free until the first item's barrier),
    // then each lane reads the CBG columns it feeds.
    {
        int* colTab = reinterpret_cast<int*>(fa_lp_smem);
        if (tid < CBG * 16) {
            const int j = colBase + tid;
            int off = 0;
            if (j < Ncols) {
                const int n = fast_div(j, P, 1.0f / P), rem = j - n * P;
                const int R = fast_div(rem, OW, 1.0f / OW), C = rem - R * OW;
                off = (((n - n_lo) * c.xRows + R * g.dnR) * OW + C) * DP;
            }
            colTab[tid] = off;
        }
        __syncthreads();
    }
    int xoff[CBG];
#pragma unroll
    for (int q = 0; q < CBG; q++) xoff[q] = reinterpret_cast<const int*>(fa_lp_smem)[q * 16 + l16] + 8 * g4;

    // 8 waves: with MB = 4 the two wave quads own two row blocks each
    // (MSPLIT); otherwise every wave accumulates all MB row blocks and the
    // waves split tap rows and 32-column MFMA slices between them.
    constexpr bool MSPLIT = NW == 8 && MB == 4;
    constexpr int MW = MSPLIT ? MB / 2 : MB;  // row blocks a wave accumulates
    const int mb0 = MSPLIT ? (wave >> 2) * MW : 0;
    floatx4 acc[MW][CBG];
#pragma unroll
    for (int i = 0; i < MW; i++)
#pragma unroll
        for (int q = 0; q < CBG; q++) acc[i][q] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Staging lanes: LPR lanes per staged row, each holding 8 consecutive
    // columns.  dy: RPP consecutive planes per wave pass, lane row myRow.  x:
    // one row per pass, lane (copy C = q*RPP + lane/LPR, column group) for
    // pass q of QP per row.  Loads go through buffer resources with a
    // wave-uniform row offset and lane offsets fixed per tap-column chunk;
    // masked samples use an out-of-range lane offset and read 0.
    const int LPR = c.DWp >> 3;
    const int RPP = 64 / LPR;  // dy rows per wave pass; also copies per x pass
    const int myRow = lane / LPR, cgrp = lane - myRow * LPR;
    const int QP = (OW + RPP - 1) / RPP;  // x passes per staged row
    int dv[8];   // dy lane byte offsets (plane myRow of the pass, tap column), or FA_OOB
    int xv[8];   // x lane byte offsets (copy C of pass 0, column), or FA_OOB
    int jChunk = -1, j0 = 0;
    // Rows need no per-sample mask when every staged dy plane is real (K fills
    // the 16-plane blocks) and no tap row is off the up-sampling lattice; x
    // rows past the field's last row are never read by a valid output.
    const bool dMask = K != MT || g.upFr > 1;

    // 32-bit item arithmetic: the plan keeps items * (splits + 1) below 2^31.
    const int perImage = c.colChunks * c.rowChunks;
    const int items = m.B * perImage;
    const int it0 = items * s / c.splits, it1 = items * (s + 1) / c.splits;
    for (int it = it0; it < it1; it++) {
        const int b = it / perImage;
        const int rest = it - b * perImage;
        const int cc = rest / c.rowChunks;
        const int i0 = (rest - cc * c.rowChunks) * c.RR;
        const int rr = min(c.RR, g.FHe - i0);
        if (cc != jChunk) {  // wave-uniform
            jChunk = cc;
            j0 = cc * c.CW;
            const int jEnd = min(j0 + c.CW, g.FWe);
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int jj = j0 + 8 * cgrp + e;
                int tj = g.flip ? g.FWe - 1 - jj : jj;
                bool on = jj < jEnd;
                if (g.upFc > 1) {
                    on = on && (tj % g.upFc) == 0;
                    tj /= g.upFc;
                }
                dv[e] = on ? 4 * (myRow * FHW + min(max(tj, 0), g.FW - 1)) : FA_OOB;
                const int xc = j0 + myRow * g.dnC + 8 * cgrp + e;  // copy C = myRow in pass q = 0
                xv[e] = (myRow < OW && xc < g.W) ? 4 * xc : FA_OOB;
            }
        }
        __syncthreads();
        // dy rows (il, k), k fastest; a pass's RPP rows share il (RPP divides
        // MT).  Per batch of DUD passes: a per-lane table (lane d = the wave's
        // d-th pass) of the pass's byte offset (flip and up-sampling lattice
        // applied), LDS row base (-1: not live), planes left (K - k0) and
        // tap-row mask, read back per pass with readlane / a ballot bit; then
        // every load, then the rounding stores.
        // Passes per batch: 64 / 96 samples in flight per lane (half that for the
        // two-term split, whose rounding doubles the registers a sample holds).
        constexpr int DUD = (SPLIT ? 16 : 32) / NW, DUX = (SPLIT ? 24 : 48) / NW;
        const int dRows = MT * rr;
        const __amdgpu_buffer_rsrc_t dyr = buffer_rsrc(DY + (size_t)b * K * FHW, (uint32_t)K * FHW * 4u);
        for (int pb = wave * RPP; pb < dRows; pb += NW * RPP * DUD) {
            int tSoff, tDst, tKl;
            uint64_t tOk;
            {
                const int p0 = pb + NW * RPP * lane;
                const int il = p0 / MT, k0 = p0 - il * MT;
                int ti = i0 + il;
                if (g.flip) ti = g.FHe - 1 - ti;
                bool on = lane < DUD && p0 < dRows && il < rr;
                if (g.upFr > 1) {
                    on = on && (ti % g.upFr) == 0;
                    ti /= g.upFr;
                }
                tSoff = 4 * (min(k0, K - 1) * FHW + min(max(ti, 0), g.FH - 1) * g.FW);
                tDst = (lane < DUD && p0 < dRows) ? (il * MT + k0) * DP : -1;
                tKl = K - k0;
                tOk = __ballot(on);
            }
            float v[DUD][8];
#pragma unroll
            for (int d = 0; d < DUD; d++) {
                const int soff = __builtin_amdgcn_readlane(tSoff, d);
                if (dMask) {  // wave-uniform, decided per block
                    // Without the mask K == MT, so k0 + myRow < K for every lane.
                    const bool rowOn = ((tOk >> d) & 1) && myRow < __builtin_amdgcn_readlane(tKl, d);
#pragma unroll
                    for (int e = 0; e < 8; e++) v[d][e] = buffer_load_f32(dyr, rowOn ? dv[e] : FA_OOB, soff);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; e++) v[d][e] = buffer_load_f32(dyr, dv[e], soff);
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
#pragma unroll
            for (int d = 0; d < DUD; d++) {
                V hv, lv;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const T h = (T)v[d][e];
                    hv[e] = h;
                    if (SPLIT) lv[e] = (T)(v[d][e] - (float)h);
                }
                const int uni = __builtin_amdgcn_readlane(tDst, d);
                const bool live = uni >= 0;
                const int dst = uni + myRow * DP + 8 * cgrp;
                *reinterpret_cast<V*>(live ? sD + dst : sJunk + 8 * lane) = hv;
                if (SPLIT) *reinterpret_cast<V*>(live ? sD + dTerm + dst : sJunk + 512 + 8 * lane) = lv;
            }
        }
        // x rows (nl, r): one staged row per pass q of QP; lane (copy C, column
        // group) reads x columns j0 + C*dn + 8*cgrp .. +7 of row i0 + r of
        // plane n_lo + nl, copy C holding the row shifted left by C*dn.  The
        // passes of a batch come from a per-lane table as above.
        const int xPasses = nCnt * c.xRows * QP;
        const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + ((size_t)b * N + n_lo) * HWx, (uint32_t)nCnt * HWx * 4u);
        for (int tb = wave; tb < xPasses; tb += NW * DUX) {
            int tSoff, tDst, tQ;
            {
                const int t = tb + NW * lane;
                const int u = t / QP, q = t - u * QP;
                const int nl = u / c.xRows, r = u - nl * c.xRows;
                tSoff = 4 * (min(nl, nCnt - 1) * HWx + min(i0 + r, g.H - 1) * g.W);
                tDst = (lane < DUX && t < xPasses) ? (nl * c.xRows + r) * OW * DP : -1;
                tQ = q;
            }
            float v[DUX][8];
            int cU[DUX];
#pragma unroll
            for (int d = 0; d < DUX; d++) {
                const int soff = __builtin_amdgcn_readlane(tSoff, d);
                if (QP == 1) {  // wave-uniform: the lane offsets of the chunk apply as they are
                    cU[d] = myRow;
#pragma unroll
                    for (int e = 0; e < 8; e++) v[d][e] = buffer_load_f32(xr, xv[e], soff);
                } else {
                    const int C = __builtin_amdgcn_readlane(tQ, d) * RPP + myRow;
                    cU[d] = C;
                    const int colStart = j0 + C * g.dnC + 8 * cgrp;
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        v[d][e] = buffer_load_f32(xr, (C < OW && colStart + e < g.W) ? 4 * (colStart + e) : FA_OOB,
                                                  soff);
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
#pragma unroll
            for (int d = 0; d < DUX; d++) {
                V hv, lv;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const T h = (T)v[d][e];
                    hv[e] = h;
                    if (SPLIT) lv[e] = (T)(v[d][e] - (float)h);
                }
                const int uni = __builtin_amdgcn_readlane(tDst, d);
                const bool live = uni >= 0 && cU[d] < OW;
                const int dst = uni + cU[d] * DP + 8 * cgrp;
                *reinterpret_cast<V*>(live ? sX + dst : sJunk + 8 * lane) = hv;
                if (SPLIT) *reinterpret_cast<V*>(live ? sX + xTerm + dst : sJunk + 512 + 8 * lane) = lv;
            }
        }
        __syncthreads();
        // Row lanes RW x column parts: 4 x 1 (4 waves, or an MSPLIT quad);
        // for 8 waves otherwise 4 x 2 when the chunk has two or more 32-column
        // slices, else 8 x 1.
        constexpr int QW = MSPLIT ? 4 : NW;
        const int RW = (QW == 8 && c.DWp < 64) ? 8 : 4;
        const int wq = MSPLIT ? (wave & 3) : wave;
        const int parts = QW / RW;
        const int il0 = wq % RW, part = wq / RW;
        const int jq = (c.DWp / 32 + parts - 1) / parts * 32;
        const int jb0 = part * jq, jb1 = min(c.DWp, jb0 + jq);
        for (int il = il0; il < rr; il += RW) {
            const T* dRow = sD + ((size_t)il * MT + l16) * DP + 8 * g4;
            const T* xRow = sX + (size_t)il * OW * DP;
            for (int jb = jb0; jb < jb1; jb += 32) {
                V av[MW], bv[CBG];
#pragma unroll
                for (int mb = 0; mb < MW; mb++) av[mb] = *reinterpret_cast<const V*>(dRow + (mb0 + mb) * 16 * DP + jb);
#pragma unroll
                for (int q = 0; q < CBG; q++) bv[q] = *reinterpret_cast<const V*>(xRow + xoff[q] + jb);
                if (SPLIT) {
                    V al[MW], bl[CBG];
#pragma unroll
                    for (int mb = 0; mb < MW; mb++)
                        al[mb] = *reinterpret_cast<const V*>(dRow + dTerm + (mb0 + mb) * 16 * DP + jb);
#pragma unroll
                    for (int q = 0; q < CBG; q++) bl[q] = *reinterpret_cast<const V*>(xRow + xTerm + xoff[q] + jb);
#pragma unroll
                    for (int mb = 0; mb < MW; mb++)
#pragma unroll
                        for (int q = 0; q < CBG; q++) {
                            acc[mb][q] = lp_mfma<PREC>(al[mb], bv[q], acc[mb][q]);
                            acc[mb][q] = lp_mfma<PREC>(av[mb], bl[q], acc[mb][q]);
                        }
                }
#pragma unroll
                for (int mb = 0; mb < MW; mb++)
#pragma unroll
                    for (int q = 0; q < CBG; q++) acc[mb][q] = lp_mfma<PREC>(av[mb], bv[q], acc[mb][q]);
            }
        }
    }
    // The waves' sums meet in LDS in a fixed order, then every wave writes a
    // share of the block's partial dW, ws[s][k][j].
    // D[i][j] of a 16x16 MFMA block lives in lane (j + 16*(i/4)), element i%4.
    __syncthreads();
    float* out = ws + (size_t)s * K * Ncols;
    if constexpr (NW == 4 || MSPLIT) {
        // Per wave quad (all four waves, or each quad of an MSPLIT block), over
        // its waves r = wave % 4, (r0 + r2) + (r1 + r3): waves 2 and 3 park
        // their sums, waves 0 and 1 add them and park the results, then wave
        // r writes element v = r of every accumulator.
        constexpr int E = MW * CBG * 4;
        const int r = wave & 3;
        float* red = reinterpret_cast<float*>(fa_lp_smem) + (size_t)(wave >> 2) * 2 * E * 64;
        if (r >= 2) {
#pragma unroll
            for (int mb = 0; mb < MW; mb++)
#pragma unroll
                for (int q = 0; q < CBG; q++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        red[(size_t)((r - 2) * E + (mb * CBG + q) * 4 + v) * 64 + lane] = acc[mb][q][v];
        }
        __syncthreads();
        if (r < 2) {
#pragma unroll
            for (int mb = 0; mb < MW; mb++)
#pragma unroll
                for (int q = 0; q < CBG; q++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        float* p = red + (size_t)(r * E + (mb * CBG + q) * 4 + v) * 64 + lane;
                        *p = acc[mb][q][v] + *p;
                    }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < MW * CBG; t++) {
            const int mb = t / CBG, q = t - mb * CBG;
            const int e = 4 * t + r;
            const float val = red[(size_t)e * 64 + lane] + red[(size_t)(E + e) * 64 + lane];
            const int j = colBase + q * 16 + l16;
            const int k = (mb0 + mb) * 16 + 4 * g4 + r;
            if (j < Ncols && k < K) out[(size_t)k * Ncols + j] = val;
        }
    } else {
        // ((w0 + w4) + (w2 + w6)) + ((w1 + w5) + (w3 + w7)): waves 4..7 park
        // whole accumulators in slots 0..3 (16 bytes a lane), waves 0..3 add
        // theirs into slot w, then the eight waves write tiles t = w, w + 8, ...
        constexpr int T4 = MW * CBG;
        floatx4* red4 = reinterpret_cast<floatx4*>(fa_lp_smem);
        if (wave >= 4) {
#pragma unroll
            for (int t = 0; t < T4; t++) red4[(size_t)((wave - 4) * T4 + t) * 64 + lane] = acc[t / CBG][t % CBG];
        }
        __syncthreads();
        if (wave < 4) {
#pragma unroll
            for (int t = 0; t < T4; t++) {
                floatx4* p = red4 + (size_t)(wave * T4 + t) * 64 + lane;
                *p = acc[t / CBG][t % CBG] + *p;
            }
        }
        __syncthreads();
        for (int t = wave; t < T4; t += NW) {
            const floatx4* p = red4 + (size_t)t * 64 + lane;
            const floatx4 val = (p[0] + p[(size_t)2 * T4 * 64]) + (p[(size_t)T4 * 64] + p[(size_t)3 * T4 * 64]);
            const int mb = t / CBG, q = t - mb * CBG;
            const int j = colBase + q * 16 + l16;
            if (j < Ncols) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int k = mb * 16 + 4 * g4 + v;
                    if (k < K) out[(size_t)k * Ncols + j] = val[v];
                }
            }
        }
    }
}

size_t fa_igemm_lp_lds_bytes(const FaCfg& c, int OW, int precision) {
    const int terms = precision == COGCONV_SPLIT_BF16 ? 2 : 1;
    const size_t dp = (size_t)c.DWp + 8;
    const size_t stage = (size_t)terms * 2 *
                         ((size_t)c.RR * 16 * c.MB * dp + (size_t)c.nCntMax * c.xRows * OW * dp + 64 * 8);
    // Cross-wave sums: two parked sums per wave quad; four (of all MB row
    // blocks) for 8 waves with MB < 4.
    const size_t red = (size_t)(c.NW == 8 && c.MB != 4 ? 4 : 2) * c.MB * c.CBW * 4 * 64 * sizeof(float);
    return stage > red ? stage : red;
}

template <int MB, int PREC, int NW>
static void fa_lp_launch3(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy, float* ws,
                          hipStream_t s, dim3 grid) {
    const dim3 block(64 * NW);
    switch (c.CBW) {
    case 3: hipLaunchKernelGGL((k_fa_igemm_lp<MB, 3, PREC, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    case 6: hipLaunchKernelGGL((k_fa_igemm_lp<MB, 6, PREC, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    default: hipLaunchKernelGGL((k_fa_igemm_lp<MB, 9, PREC, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    }
}

template <int MB, int PREC>
static hipError_t fa_lp_launch2(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                                float* ws, hipStream_t s, dim3 grid) {
    if (c.NW == 8) fa_lp_launch3<MB, PREC, 8>(g, m, c, x, dy, ws, s, grid);
    else fa_lp_launch3<MB, PREC, 4>(g, m, c, x, dy, ws, s, grid);
    return hipGetLastError();
}

template <int PREC>
static hipError_t fa_lp_launch1(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                                float* ws, hipStream_t s, dim3 grid) {
    switch (c.MB) {
    case 1: return fa_lp_launch2<1, PREC>(g, m, c, x, dy, ws, s, grid);
    case 2: return fa_lp_launch2<2, PREC>(g, m, c, x, dy, ws, s, grid);
    case 3: return fa_lp_launch2<3, PREC>(g, m, c, x, dy, ws, s, grid);
    default: return fa_lp_launch2<4, PREC>(g, m, c, x, dy, ws, s, grid);
    }
}

hipError_t launch_fa_igemm_lp(int precision, const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x,
                              const float* dy, float* out, float* ws, hipStream_t s, const Epi& ep) {
    const long blocks = (long)c.colGroups * c.splits;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks);
    hipError_t e;
    switch (precision) {
    case COGCONV_BF16: e = fa_lp_launch1<COGCONV_BF16>(g, m, c, x, dy, ws, s, grid); break;
    case COGCONV_FP16: e = fa_lp_launch1<COGCONV_FP16>(g, m, c, x, dy, ws, s, grid); break;
    case COGCONV_SPLIT_BF16: e = fa_lp_launch1<COGCONV_SPLIT_BF16>(g, m, c, x, dy, ws, s, grid); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    const long total = (long)c.K * m.N * g.OH * g.OW;
    return launch_reduce_partials(ws, out, c.splits, total, (long)g.OH * g.OW, s, ep);
}

}  // namespace cogconv
