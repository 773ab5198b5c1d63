// FP32 implicit GEMM on MFMA for FilterAdjointBlockReduceSum with BorderValid
// (the weight gradient of a layer):
//
//   dW[k*N + n](R, C) = sum_b sum_{i,j} tap_{b,k}(i, j) * X_{b*N + n}(R*dn + i, C*dn + j)
//
// tap = dy (cross-correlation) or dy flipped (convolution), zero-inserted when
// up-sampled (ConvolveHyperKernel.scala:1310-1336; ConvolveToSmallFieldPipelinedTiledHyperKernel.scala:45-127).
// As a GEMM: M = K (dy planes per image), columns j = (n, R, C) (N*OH*OW of
// them, which is the output layout), and the reduction runs over
// (image b, tap row i, tap column j) -- tens of thousands of terms.
//
// gfx950 mapping: v_mfma_f32_16x16x4_f32; each MFMA consumes four consecutive
// tap columns of one tap row (the four k-lanes, lane/16).  A block (NW = 4 or
// 8 wave64) owns a column group of CBG 16-column blocks and all MB row blocks.  It
// stages RR tap rows of all K dy planes of one image (stored k-fastest with a
// row pitch of 17 mod 32 words, spreading both the staging stores and the
// A-operand reads over the banks)
// plus the x rows they touch, for the input planes of its column group; staging
// runs one row per wave pass with lanes across columns.  Wave w reduces tap
// rows w%4, w%4+4, .. for every tile of the group; with 8 waves, waves w and
// w+4 (one SIMD) take the two halves of the row blocks (even MB) or of the tap
// columns (odd MB), so one wave's MFMAs run while the other waits on LDS.  The
// waves' sums meet in LDS.  Blocks split the (image, tap-row chunk) space; each writes one partial
// dW, and the partials are added in a fixed order by a two-level reduction, so
// results do not depend on scheduling.
#include "epi.h"
#include "stage.h"

namespace cogconv {

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int MB, int CBG, int NW>
__global__ __launch_bounds__(64 * NW) void k_fa_igemm_f32(Geom g, PlaneMap m, FaCfg c, const float* __restrict__ X,
                                                      const float* __restrict__ DY, float* __restrict__ ws) {
    // Staged dy rows are k-fastest with a pitch of 16*MB padded to 17 mod 32
    // words: the staging stores (lanes along tap columns, stride MTP) hit 32
    // different banks, and the A-operand reads ((jb+g4)*MTP + l16) of a
    // half-wave hit at most one bank twice.
    constexpr int MTP = MB == 1 ? 17 : (MB == 4 ? 81 : 49);
    extern __shared__ float smem[];
    const int MTp = MTP;
    float* sD = smem;                                  // [RR][DWp][MTP]
    float* sX = smem + (size_t)c.RR * c.DWp * MTP;     // [nCnt][xRows][xColsP], plane pitch xPlaneP
    float* sJunk = sX + (size_t)c.nCntMax * c.xPlaneP;  // 128 scratch slots for masked-off stores
    const int FHW = g.FH * g.FW;                       // 32-bit offsets: the plan checks the sizes
    const int HWx = g.H * g.W;

    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int cg = bid % c.colGroups;
    const int s = bid / c.colGroups;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g4 = lane >> 4, l16 = lane & 15;
    const int N = m.N, K = c.K;
    // Per-lane column offsets (two 64-lane passes at most) of the staged dy
    // rows (flip and up-sampling lattice applied) and x rows; they depend on
    // the item's tap-column chunk and are set at the start of each item.
    const bool dTwo = c.DWp > 64, xTwo = c.xCols > 64;
    int dCol0 = 0, dCol1 = 0, xCol0 = 0, xCol1 = 0;
    bool dIn0 = false, dIn1 = false, xIn0 = false, xIn1 = false;
    int jChunk = -1;  // tap-column chunk the offsets are set for
    const int P = g.OH * g.OW;
    const int Ncols = N * P;
    const int colBase = cg * CBG * 16;
    const int n_lo = colBase / P;
    const int n_hi = min(N - 1, (min(colBase + CBG * 16, Ncols) - 1) / P);
    const int nCnt = n_hi - n_lo + 1;
    const int xPlane = c.xPlaneP;

    // Staged-x offset of each of the group's CBG*16 output columns (n, R, C):
    // one column per thread into LDS (the staging area is free until the first
    // item's barrier), then each lane reads the CBG columns it feeds.
    {
        int* colTab = reinterpret_cast<int*>(smem);
        if (tid < CBG * 16) {
            const int j = colBase + tid;
            int off = 0;
            if (j < Ncols) {
                const int n = fast_div(j, P, 1.0f / P), rem = j - n * P;
                const int R = fast_div(rem, g.OW, 1.0f / g.OW), C = rem - R * g.OW;
                off = (n - n_lo) * xPlane + R * g.dnR * c.xColsP + C * g.dnC;
            }
            colTab[tid] = off;
        }
        __syncthreads();
    }
    int xoff[CBG];
#pragma unroll
    for (int q = 0; q < CBG; q++) xoff[q] = reinterpret_cast<const int*>(smem)[q * 16 + l16] + g4;

    // With 8 waves and MB = 4, the two wave quads (waves 0-3, 4-7; one wave
    // of each per SIMD) own halves of the row blocks; otherwise they take the
    // two halves of the tap columns (one row block per wave would double the
    // B-operand LDS reads per MFMA).
    constexpr bool MSPLIT = NW == 8 && MB == 4;
    constexpr int MW = MSPLIT ? MB / 2 : MB;  // row blocks a wave accumulates
    const int mb0 = MSPLIT ? (wave >> 2) * MW : 0;
    floatx4 acc[MW][CBG];
#pragma unroll
    for (int i = 0; i < MW; i++)
#pragma unroll
        for (int q = 0; q < CBG; q++) acc[i][q] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Items: (image b, tap-column chunk, tap-row chunk), row chunks fastest.
    // 32-bit item arithmetic: the plan keeps items * (splits + 1) below 2^31.
    const int perImage = c.colChunks * c.rowChunks;
    const int items = m.B * perImage;
    const int it0 = items * s / c.splits, it1 = items * (s + 1) / c.splits;
    constexpr int UD = 64 * MB / NW, UX = 96 / NW;  // UD: a wave's dy rows when RR = 4
    constexpr int MT = 16 * MB;                     // dy planes the MFMAs read (K padded with zeros)
    const int xRowsAll = nCnt * c.xRows;            // staged x rows (plane, row)

    // The current item: image, first tap row, tap-row count.  set_item also
    // sets the per-lane column offsets when the tap-column chunk changes.
    int b = 0, i0 = 0, rr = 0;
    auto set_item = [&](int it) __attribute__((always_inline)) {
        b = it / perImage;
        const int rest = it - b * perImage;
        const int cc = rest / c.rowChunks;
        i0 = (rest - cc * c.rowChunks) * c.RR;
        rr = min(c.RR, g.FHe - i0);
        if (cc != jChunk) {  // wave-uniform
            jChunk = cc;
            const int j0 = cc * c.CW;
            const int jEnd = min(j0 + c.CW, g.FWe);
            auto col = [&](int jj, int& dc, bool& di, int& xc, bool& xi) {
                int tj = g.flip ? g.FWe - 1 - jj : jj;
                bool on = jj < jEnd;
                if (g.upFc > 1) { on = on && (tj % g.upFc) == 0; tj /= g.upFc; }
                dc = min(max(tj, 0), g.FW - 1);
                di = on;
                xc = min(jj, g.W - 1);
                xi = jj < g.W;
            };
            col(j0 + lane, dCol0, dIn0, xCol0, xIn0);
            col(j0 + lane + 64, dCol1, dIn1, xCol1, xIn1);
        }
    };

    // Staging of the current item's dy tap rows and x rows: one row per wave
    // pass, lanes across columns (at most 128).  Per batch of UD dy rows and
    // UX x rows (rows w, w+NW, ... of each set): the rows' offsets and masks
    // (per-lane tables), then all loads of both sets, from in-bounds offsets
    // or the out-of-range offset, with no branch between them, then the LDS
    // stores; rows past the end and lanes past the row width store into a
    // scratch slot instead of branching.  One batch covers a block's rows for
    // the common shapes, so staging costs one memory round trip.
    int dDst = -1, xDst = -1;
    auto load_batch = [&](int bd, int bx, float(&dv0)[UD], float(&dv1)[UD], float(&xv0)[UX],
                          float(&xv1)[UX]) __attribute__((always_inline)) {
        const int dRows = MT * rr;  // (il, k); the pitch padding beyond MT is never read
        const __amdgpu_buffer_rsrc_t dyr = buffer_rsrc(DY + (size_t)b * K * FHW, (uint32_t)K * FHW * 4u);
        const int vd0 = dIn0 ? 4 * dCol0 : STAGE_OOB, vd1 = dIn1 ? 4 * dCol1 : STAGE_OOB;
        const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + ((size_t)b * N + n_lo) * HWx, (uint32_t)nCnt * HWx * 4u);
        const int vx0 = xIn0 ? 4 * xCol0 : STAGE_OOB, vx1 = xIn1 ? 4 * xCol1 : STAGE_OOB;
        // Per-lane row tables of the batch (lane u = the wave's u-th dy row
        // and u-th x row): byte offsets (flip, up-sampling lattice and clamps
        // applied), LDS destinations (-1: not live) and in-field masks,
        // computed in parallel across lanes and read back per row with
        // readlane / a ballot bit.
        int dSoff, xSoff;
        uint64_t dMask, xMask;
        {
            const int row = bd + NW * lane;
            const int il = row / MT, k = row - il * MT;
            int ti = i0 + il;
            if (g.flip) ti = g.FHe - 1 - ti;
            bool on = lane < UD && row < dRows && k < K && il < rr;
            if (g.upFr > 1) {
                on = on && (ti % g.upFr) == 0;
                ti /= g.upFr;
            }
            dSoff = 4 * (min(k, K - 1) * FHW + min(max(ti, 0), g.FH - 1) * g.FW);
            dDst = (lane < UD && row < dRows) ? il * c.DWp * MTP + k : -1;
            dMask = __ballot(on);
        }
        {
            const int row = bx + NW * lane;
            const int nl = row / c.xRows, lr = row - nl * c.xRows;
            const int gy = i0 + lr;
            const bool live = lane < UX && row < xRowsAll;
            xSoff = 4 * (min(nl, nCnt - 1) * HWx + min(gy, g.H - 1) * g.W);
            xDst = live ? nl * xPlane + lr * c.xColsP : -1;
            xMask = __ballot(live && gy < g.H);
        }
        auto drow = [&](int u, int& soff, bool& on) {
            soff = __builtin_amdgcn_readlane(dSoff, u);
            on = (dMask >> u) & 1;
        };
        auto xrow = [&](int u, int& soff, bool& on) {
            soff = __builtin_amdgcn_readlane(xSoff, u);
            on = (xMask >> u) & 1;
        };
        if (dTwo) load_rows_fn<true, UD>(dyr, drow, vd0, vd1, dv0, dv1);
        else load_rows_fn<false, UD>(dyr, drow, vd0, vd1, dv0, dv1);
        if (xTwo) load_rows_fn<true, UX>(xr, xrow, vx0, vx1, xv0, xv1);
        else load_rows_fn<false, UX>(xr, xrow, vx0, vx1, xv0, xv1);
    };
    auto store_batch = [&](float* bufD, float* bufX, const float(&dv0)[UD], const float(&dv1)[UD],
                           const float(&xv0)[UX], const float(&xv1)[UX]) __attribute__((always_inline)) {
        auto ddst = [&](int u) { return __builtin_amdgcn_readlane(dDst, u); };
        auto xdst = [&](int u) { return __builtin_amdgcn_readlane(xDst, u); };
        if (dTwo) store_rows_fn<true, UD>(ddst, bufD, sJunk, lane, c.DWp, MTP, dv0, dv1);
        else store_rows_fn<false, UD>(ddst, bufD, sJunk, lane, c.DWp, MTP, dv0, dv1);
        if (xTwo) store_rows_fn<true, UX>(xdst, bufX, sJunk, lane, c.xCols, 1, xv0, xv1);
        else store_rows_fn<false, UX>(xdst, bufX, sJunk, lane, c.xCols, 1, xv0, xv1);
    };

    // MFMAs of a staged item with rrI tap rows.  The waves sharing a set of
    // row blocks (all NW, or a quad with MSPLIT) form RW tap-row lanes (RW =
    // 4, 2 or 1 as RR allows) times column parts: wave (il0, part) reduces
    // tap rows il0, il0 + RW, ... over tap columns [jb0, jb1).
    const int RW = c.RR >= 4 ? 4 : (c.RR >= 2 ? 2 : 1);
    const int lgRW = RW == 4 ? 2 : (RW == 2 ? 1 : 0);
    constexpr int QW = MSPLIT ? 4 : NW;
    const int wq = MSPLIT ? (wave & 3) : wave;
    const int parts = QW >> lgRW;
    const int il0 = wq & (RW - 1), part = wq >> lgRW;
    const int jq = (c.DWp / 4 + parts - 1) / parts * 4;
    const int jb0 = part * jq, jb1 = min(c.DWp, jb0 + jq);
    auto mfma_item = [&](const float* bufD, const float* bufX, int rrI) __attribute__((always_inline)) {
        for (int il = il0; il < rrI; il += RW) {
            const float* dRow = bufD + ((size_t)il * c.DWp + g4) * MTp + l16;
            const float* xRow = bufX + il * c.xColsP;
            for (int jb = jb0; jb < jb1; jb += 4) {
                float av[MW];
#pragma unroll
                for (int mb = 0; mb < MW; mb++) av[mb] = dRow[(size_t)jb * MTp + (mb0 + mb) * 16];
                float bv[CBG];
#pragma unroll
                for (int q = 0; q < CBG; q++) bv[q] = xRow[xoff[q] + jb];
#pragma unroll
                for (int mb = 0; mb < MW; mb++)
#pragma unroll
                    for (int q = 0; q < CBG; q++)
                        acc[mb][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mb], bv[q], acc[mb][q], 0, 0, 0);
            }
        }
    };

    for (int it = it0; it < it1; it++) {
        set_item(it);
        __syncthreads();  // the previous item's MFMAs are done with the staging area
        const int dRows = MT * rr;
        for (int bd = wave, bx = wave; bd < dRows || bx < xRowsAll; bd += NW * UD, bx += NW * UX) {
            float dv0[UD], dv1[UD], xv0[UX], xv1[UX];
            load_batch(bd, bx, dv0, dv1, xv0, xv1);
            __builtin_amdgcn_sched_barrier(0);  // every load of the batch ahead of the first store
            store_batch(sD, sX, dv0, dv1, xv0, xv1);
        }
        __syncthreads();
        mfma_item(sD, sX, rr);
    }
    // The waves' sums meet in LDS in a fixed order and every wave then writes
    // a share of the block's partial dW, ws[s][k][j].
    // D[i][j] of a 16x16 MFMA block lives in lane (j + 16*(i/4)), element i%4.
    constexpr int E = MW * CBG * 4;  // accumulator floats per lane
    __syncthreads();
    float* red = smem;
    float* out = ws + (size_t)s * K * Ncols;
    if constexpr (NW == 4 || MSPLIT) {
        // Per wave quad, over its waves r = wave % 4, (r0 + r2) + (r1 + r3):
        // waves 2 and 3 park their sums, waves 0 and 1 add them and park the
        // results, then wave r writes element v = r of every accumulator.
        const int r = wave & 3;
        float* rq = red + (size_t)(wave >> 2) * 2 * E * 64;
        if (r >= 2) {
#pragma unroll
            for (int mb = 0; mb < MW; mb++)
#pragma unroll
                for (int q = 0; q < CBG; q++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        rq[(size_t)((r - 2) * E + (mb * CBG + q) * 4 + v) * 64 + lane] = acc[mb][q][v];
        }
        __syncthreads();
        if (r < 2) {
#pragma unroll
            for (int mb = 0; mb < MW; mb++)
#pragma unroll
                for (int q = 0; q < CBG; q++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        float* p = rq + (size_t)(r * E + (mb * CBG + q) * 4 + v) * 64 + lane;
                        *p = acc[mb][q][v] + *p;
                    }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < MW * CBG; t++) {
            const int mb = t / CBG, q = t - mb * CBG;
            const int e = 4 * t + r;
            const float val = rq[(size_t)e * 64 + lane] + rq[(size_t)(E + e) * 64 + lane];
            const int j = colBase + q * 16 + l16;
            const int k = (mb0 + mb) * 16 + 4 * g4 + r;
            if (j < Ncols && k < K) out[(size_t)k * Ncols + j] = val;
        }
    } else {
        // ((w0 + w4) + (w2 + w6)) + ((w1 + w5) + (w3 + w7)): waves 4..7 park
        // their sums in slots 0..3, waves 0..3 add theirs into slot w, then
        // the eight waves write tiles t = w, w + 8, ...  An accumulator is
        // parked whole (slot, tile, lane: 16 bytes a lane), so every LDS
        // access is one 16-byte access a lane.
        constexpr int T = MW * CBG;
        floatx4* red4 = reinterpret_cast<floatx4*>(smem);
        if (wave >= 4) {
#pragma unroll
            for (int t = 0; t < T; t++) red4[(size_t)((wave - 4) * T + t) * 64 + lane] = acc[t / CBG][t % CBG];
        }
        __syncthreads();
        if (wave < 4) {
#pragma unroll
            for (int t = 0; t < T; t++) {
                floatx4* p = red4 + (size_t)(wave * T + t) * 64 + lane;
                *p = acc[t / CBG][t % CBG] + *p;
            }
        }
        __syncthreads();
        for (int t = wave; t < T; t += NW) {
            const floatx4* p = red4 + (size_t)t * 64 + lane;
            const floatx4 val = (p[0] + p[(size_t)2 * T * 64]) + (p[(size_t)T * 64] + p[(size_t)3 * T * 64]);
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

// Sum `splits` partial arrays of `total` floats in a fixed order, in one
// launch.  A block owns OPB (64 or 32) consecutive outputs and 16 waves; with
// OPB = 32 each wave's two half-waves take different partials, so a block has
// G = 16 * 64 / OPB partial groups.  Group gi adds a contiguous run of
// ceil(splits / G) partials (loads batched 16 or 8 at a time), and the G group sums
// are then added in group order.  The order depends on `splits` and OPB only,
// so results do not depend on scheduling.  OPB = 32 doubles the blocks when
// the outputs are few (a c2 weight gradient has 4608).
constexpr int RED_W = 16;  // waves per reduction block

template <int OPB>
__global__ __launch_bounds__(64 * RED_W) void k_reduce_partials(const float* __restrict__ ws, float* __restrict__ out,
                                                                  int splits, long total, long points, Epi e) {
    constexpr int G = RED_W * 64 / OPB;
    __shared__ float part[G][OPB];
    const int o = threadIdx.x % OPB;
    const int gi = threadIdx.x / OPB;
    const long idx = (long)blockIdx.x * OPB + o;
    const long safe = idx < total ? idx : total - 1;
    const int per = (splits + G - 1) / G;
    const int s0 = gi * per;
    const int s1 = min(splits, s0 + per);
    constexpr int LB = OPB == 32 ? 8 : 16;  // loads in flight per lane
    float acc = 0.f;
    for (int sb = s0; sb < s1; sb += LB) {
        float v[LB];
#pragma unroll
        for (int i = 0; i < LB; i++) v[i] = ws[(size_t)min(sb + i, splits - 1) * total + safe];
#pragma unroll
        for (int i = 0; i < LB; i++)
            if (sb + i < s1) acc += v[i];
    }
    part[gi][o] = acc;
    __syncthreads();
    if (gi == 0 && idx < total) {
        float t = part[0][o];
#pragma unroll
        for (int k = 1; k < G; k++) t += part[k][o];
        const long p = idx / points;
        out[idx] = epi(e, t, (int)p, (int)(idx - p * points));
    }
}

hipError_t launch_reduce_partials(const float* ws, float* out, int splits, long total, long points, hipStream_t s,
                                  const Epi& e) {
    if (total <= 0) return hipErrorInvalidConfiguration;
    if (total < 64L * 256) {
        const long blocks = (total + 31) / 32;
        hipLaunchKernelGGL(k_reduce_partials<32>, dim3((unsigned)blocks), dim3(64 * RED_W), 0, s, ws, out, splits,
                           total, points, e);
    } else {
        const long blocks = (total + 63) / 64;
        if (blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
        hipLaunchKernelGGL(k_reduce_partials<64>, dim3((unsigned)blocks), dim3(64 * RED_W), 0, s, ws, out, splits,
                           total, points, e);
    }
    return hipGetLastError();
}

template <int MB, int NW>
static void launch_mbw(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy, float* ws,
                       hipStream_t s, dim3 grid) {
    const dim3 block(64 * NW);
    switch (c.CBW) {
    case 3: hipLaunchKernelGGL((k_fa_igemm_f32<MB, 3, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    case 6: hipLaunchKernelGGL((k_fa_igemm_f32<MB, 6, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    default: hipLaunchKernelGGL((k_fa_igemm_f32<MB, 9, NW>), grid, block, c.lds_bytes, s, g, m, c, x, dy, ws); break;
    }
}

template <int MB>
static hipError_t launch_mb(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                            float* ws, hipStream_t s, dim3 grid) {
    if (c.NW == 8) launch_mbw<MB, 8>(g, m, c, x, dy, ws, s, grid);
    else launch_mbw<MB, 4>(g, m, c, x, dy, ws, s, grid);
    return hipGetLastError();
}

hipError_t launch_fa_igemm_f32(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                               float* out, float* ws, hipStream_t s, const Epi& ep) {
    const long blocks = (long)c.colGroups * c.splits;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks);
    hipError_t e;
    switch (c.MB) {
    case 1: e = launch_mb<1>(g, m, c, x, dy, ws, s, grid); break;
    case 2: e = launch_mb<2>(g, m, c, x, dy, ws, s, grid); break;
    case 3: e = launch_mb<3>(g, m, c, x, dy, ws, s, grid); break;
    default: e = launch_mb<4>(g, m, c, x, dy, ws, s, grid); break;
    }
    if (e != hipSuccess) return e;
    const long total = (long)c.K * m.N * g.OH * g.OW;
    return launch_reduce_partials(ws, out, c.splits, total, (long)g.OH * g.OW, s, ep);
}

}  // namespace cogconv
