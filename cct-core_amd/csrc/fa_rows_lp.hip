// Row-streaming BF16 / FP16 FilterAdjointBlockReduceSum for BorderValid,
// unit steps, cross-correlation orientation (the weight gradient of a
// stride-1 layer):
//
//   dW[k*N + n](i, j) = sum_b sum_{y, x} dy_{b*K + k}(y, x) * X_{b*N + n}(y + i, x + j)
//
// (ConvolveHyperKernel.scala:1310-1336 for the contraction and output shape;
// ConvolveToSmallFieldPipelinedTiledHyperKernel.scala:45-127 for the
// reference's own split reduction of this case.)
//
// GEMM view: rows = the K dy planes, columns = (n, i, j), reduction over the
// pixels (b, y, x).  v_mfma_f32_16x16x32_{bf16,f16}: the 32-deep slice is 32
// consecutive pixels x0 .. x0 + 31 of one dy row y.  A = dy (lane l: plane
// l%16, pixels 8(l/16) .. +7), B = the x samples under tap (i, j) (lane l:
// plane l%16, pixels 8(l/16) + j .. +7 of x row y + i); D[k][n] leaves lane
// l with planes k = 4(l/16) .. +3 of column n = l%16.
//
// A block owns a 32-pixel column chunk of one image, a run of dy rows and a
// tile of (k, n) planes, and walks down the rows one a step: each step stages
// the next dy row and the x row entering the window (a ring of OD + 1 x rows)
// while the current row's MFMAs run, so each sample is loaded once per block.
// The tap-column shift j is applied in registers: a lane reads its 8 pixels
// and the next 8 (two 16-byte LDS reads per x row) and forms the window at
// offset j with byte-align instructions (odd j) or a choice of registers
// (even j); no shifted copies are staged.
//
// A block has NB x KH waves: wave w owns x-plane block w % NB and dy-plane
// blocks (w / NB) * KBW .. +KBW-1 of the tile (a 64 x 64 tile with 8 waves
// reads each dy and x row once) and writes its accumulators straight into the
// split's partial,
// ws[s][k][n][i][j]; the partials are summed afterwards in split order
// (launch_reduce_partials), so the result does not depend on timing.
#include "lp.h"
#include "stage.h"

namespace cogconv {

namespace {

constexpr int FR_OOB = (int)0x80000000u;  // lane offset past every buffer: the load reads 0
constexpr int FR_XS = 5;                   // 8-pixel slots of a staged x row (32 + OD - 1 <= 40 pixels)
constexpr int FR_PD = 4;                   // steps of row loads in flight (register queue stages)

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Four consecutive floats at lane byte offset `vo` (4-byte aligned: the
// 16-byte load needs no more here).  The range check is per dword, so a
// piece that runs past the end of the buffer keeps the pixels before it.
__device__ __forceinline__ void load4(__amdgpu_buffer_rsrc_t r, int vo, float (&v)[4]) {
    const float4 q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, 0, 0));
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}

// Four rounded elements, as the 8-byte half of an 8-element LDS slot.
template <int PREC>
__device__ __forceinline__ void store4(typename LpType<PREC>::T* dst, const float (&v)[4]) {
    typedef typename LpType<PREC>::T T;
    typedef T T4 __attribute__((ext_vector_type(4)));
    *reinterpret_cast<T4*>(dst) = T4{(T)v[0], (T)v[1], (T)v[2], (T)v[3]};
}

// Elements j .. j + 7 of the 16-element run cur ++ nxt (16-bit elements).
template <int J, typename V>
__device__ __forceinline__ V window(V cur, V nxt) {
    if constexpr (J == 0) {
        return cur;
    } else {
        const u32x4 c = __builtin_bit_cast(u32x4, cur), n = __builtin_bit_cast(u32x4, nxt);
        const unsigned w[8] = {c.x, c.y, c.z, c.w, n.x, n.y, n.z, n.w};
        u32x4 o;
        constexpr int q = J / 2;
        if constexpr (J % 2 == 0) {
            o = u32x4{w[q], w[q + 1], w[q + 2], w[q + 3]};
        } else {
            o = u32x4{__builtin_amdgcn_alignbyte(w[q + 1], w[q], 2), __builtin_amdgcn_alignbyte(w[q + 2], w[q + 1], 2),
                      __builtin_amdgcn_alignbyte(w[q + 3], w[q + 2], 2),
                      __builtin_amdgcn_alignbyte(w[q + 4], w[q + 3], 2)};
        }
        return __builtin_bit_cast(V, o);
    }
}

}  // namespace

template <int PREC, int OD, int KBW, int NB, int KH>
__global__ __launch_bounds__(64 * NB * KH) void k_fa_rows_lp(Geom g, PlaneMap m, FaCfg c, const float* __restrict__ X,
                                                             const float* __restrict__ DY, float* __restrict__ ws) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr int NTH = 64 * NB * KH;  // threads: NB x-plane blocks by KH groups of KBW dy-plane blocks
    constexpr int KT = KH * KBW * 16;  // dy planes of the tile
    constexpr int NT = NB * 16;        // x planes of the tile
    constexpr int RB = OD + 1;         // x rows in the ring
    constexpr int P = OD * OD;
    extern __shared__ __attribute__((aligned(16))) unsigned char fr_smem[];
    T* sD = reinterpret_cast<T*>(fr_smem);  // [2][4][KT][8]: dy rows, double-buffered
    T* sX = sD + 2 * 4 * KT * 8;            // [RB][FR_XS][NT][8]: x row ring
    T* sJunk = sX + RB * FR_XS * NT * 8;    // [8]: where lanes without a piece store theirs

    const int tiles = c.nTiles * c.kTiles;
    const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tile = bid % tiles, s = bid / tiles;  // s: the partial this block writes
    const int nt = tile % c.nTiles, kt = tile / c.nTiles;
    const int rs = s % c.rowSplits, bc = s / c.rowSplits;
    const int ch = bc % c.chunks, b = bc / c.chunks;
    const int K = c.K, N = m.N;
    const int dyH = g.FH, dyW = g.FW, H = g.H, W = g.W;
    const int x0 = ch * 32;
    const int yBeg = rs * c.rowsPer, yEnd = min(dyH, yBeg + c.rowsPer);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g4 = lane >> 4, l16 = lane & 15;
    const int nb = wave % NB, kh = wave / NB;

    const __amdgpu_buffer_rsrc_t dyr = buffer_rsrc(DY + (size_t)b * K * dyH * dyW, (uint32_t)K * dyH * dyW * 4u);
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + (size_t)b * N * H * W, (uint32_t)N * H * W * 4u);

    // Staging pieces: 4 pixels (16 bytes) a lane, piece fastest, so the lanes
    // of one plane's row segment sit side by side and each cache line of the
    // segment is requested by one load instruction (lines that several waves
    // request while the first fill is pending stall the vector L1).  dy:
    // piece h (pixels x0 + 4h .. +3) of plane k, 8 KT pieces; x: piece h of
    // plane n of the entering row, 2 FR_XS NT pieces; piece u = tid + NTH d.
    // The whole offset is the lane offset, so the buffer's range check covers
    // it (a load past the image reads 0); pixels past the row are zeroed
    // after the load (the loads run on into the next row).  A piece lands in
    // half h % 2 of the 8-pixel LDS slot h / 2.
    constexpr int DP = (8 * KT + NTH - 1) / NTH;
    constexpr int XP = (2 * FR_XS * NT + NTH - 1) / NTH;
    bool dUnit[DP], xUnit[XP];
    int dPix[DP], dVo[DP], dDst[DP], xPix[XP], xVo[XP], xDst[XP];
#pragma unroll
    for (int d = 0; d < DP; d++) {
        const int u = tid + NTH * d;
        const int h = u % 8, dk = u / 8;
        const int dPlane = kt * KT + dk;
        dUnit[d] = u < 8 * KT;
        dPix[d] = x0 + 4 * h;
        dVo[d] = dUnit[d] && dPlane < K ? 4 * (dPlane * dyH * dyW + dPix[d]) : FR_OOB;
        dDst[d] = ((h / 2) * KT + dk) * 8 + 4 * (h % 2);  // + buffer * 4 * KT * 8
    }
#pragma unroll
    for (int d = 0; d < XP; d++) {
        const int u = tid + NTH * d;
        const int h = u % (2 * FR_XS), xn = u / (2 * FR_XS);
        const int xPlane = nt * NT + xn;
        xUnit[d] = u < 2 * FR_XS * NT;
        xPix[d] = x0 + 4 * h;
        xVo[d] = xUnit[d] && xPlane < N ? 4 * (xPlane * H * W + xPix[d]) : FR_OOB;
        xDst[d] = ((h / 2) * NT + xn) * 8 + 4 * (h % 2);  // + ring row * FR_XS * NT * 8
    }

    // The queue's loads stay in flight across steps only if nothing between
    // a load and its store branches or reads the loaded registers (the
    // compiler's wait counts are merged over every path into a point, and any
    // such path drains the queue).  So: pixels past the row are masked at the
    // store; a row past the split (on == false) is loaded at an offset past
    // every buffer (it reads 0 without touching memory); every lane stores,
    // lanes without a piece into sJunk.
    auto row_off = [&](int y, int pitch, bool on) { return on ? 4u * (unsigned)(y * pitch) : (unsigned)FR_OOB; };
    auto load_dy = [&](int y, bool on, float (&v)[DP][4]) {
        const unsigned ro = row_off(y, dyW, on);
#pragma unroll
        for (int d = 0; d < DP; d++) load4(dyr, (int)((unsigned)dVo[d] + ro), v[d]);
    };
    auto store_dy = [&](int buf, const float (&v)[DP][4]) {
#pragma unroll
        for (int d = 0; d < DP; d++) {
            float m[4];
#pragma unroll
            for (int e = 0; e < 4; e++) m[e] = dPix[d] + e < dyW ? v[d][e] : 0.f;
            store4<PREC>(dUnit[d] ? sD + buf * 4 * KT * 8 + dDst[d] : sJunk, m);
        }
    };
    auto load_x = [&](int y, bool on, float (&v)[XP][4]) {
        const unsigned ro = row_off(y, W, on);
#pragma unroll
        for (int d = 0; d < XP; d++) load4(xr, (int)((unsigned)xVo[d] + ro), v[d]);
    };
    auto store_x = [&](int row, const float (&v)[XP][4]) {
#pragma unroll
        for (int d = 0; d < XP; d++) {
            float m[4];
#pragma unroll
            for (int e = 0; e < 4; e++) m[e] = xPix[d] + e < W ? v[d][e] : 0.f;
            store4<PREC>(xUnit[d] ? sX + row * FR_XS * NT * 8 + xDst[d] : sJunk, m);
        }
    };

    floatx4 acc[KBW][OD][OD];
#pragma unroll
    for (int kb = 0; kb < KBW; kb++)
#pragma unroll
        for (int i = 0; i < OD; i++)
#pragma unroll
            for (int j = 0; j < OD; j++) acc[kb][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Rows y0 .. y1 - 1, streamed down with a fresh prologue.
    auto stream = [&](const int y0, const int y1) {
    if (y0 < y1) {  // block-uniform
        // Prologue: x rows y0 .. y0 + OD - 1 into ring rows (y0 + i) % RB, dy row y0 into buffer 0.
        for (int i = 0; i < OD; i++) {
            float v[XP][4];
            load_x(y0 + i, true, v);
            store_x((y0 + i) % RB, v);
        }
        {
            float v[DP][4];
            load_dy(y0, true, v);
            store_dy(0, v);
        }
        // Register queue of FR_PD stages: the rows stored at the end of step t
        // (dy row y + 1, x row y + OD) were loaded FR_PD - 1 steps earlier into
        // stage (t + 1) % FR_PD, so that many steps of MFMAs cover their memory
        // latency.  Steps run in groups of FR_PD (stage indices are constants).
        float qd[FR_PD][DP][4], qx[FR_PD][XP][4];
#pragma unroll
        for (int k = 1; k < FR_PD; k++) {
            load_dy(y0 + k, y0 + k < y1, qd[k]);
            load_x(y0 + k + OD - 1, y0 + k < y1, qx[k]);
        }
        __syncthreads();
        int ring0 = y0 % RB;  // ring row of x row y
        for (int yb = y0; yb < y1; yb += FR_PD) {
#pragma unroll
            for (int u = 0; u < FR_PD; u++) {
                const int y = yb + u;
                if (y >= y1) break;  // block-uniform
                const int t = y - y0;
                // Stage u (last stored at the end of the previous step) takes the
                // rows for the end of step t + FR_PD - 1.
                load_dy(y + FR_PD, y + FR_PD < y1, qd[u]);
                load_x(y + FR_PD + OD - 1, y + FR_PD < y1, qx[u]);
                const T* dBuf = sD + (t & 1) * 4 * KT * 8;
                V a[KBW];
#pragma unroll
                for (int kb = 0; kb < KBW; kb++)
                    a[kb] = *reinterpret_cast<const V*>(dBuf + (g4 * KT + (kh * KBW + kb) * 16 + l16) * 8);
#pragma unroll
                for (int i = 0; i < OD; i++) {
                    int rr = ring0 + i;
                    rr -= rr >= RB ? RB : 0;
                    const T* xRow = sX + (rr * FR_XS * NT + nb * 16 + l16) * 8;
                    const V cur = *reinterpret_cast<const V*>(xRow + g4 * NT * 8);
                    const V nxt = *reinterpret_cast<const V*>(xRow + (g4 + 1) * NT * 8);
                    auto tap = [&](auto jc) {
                        constexpr int J = decltype(jc)::value;
                        const V bj = window<J>(cur, nxt);
#pragma unroll
                        for (int kb = 0; kb < KBW; kb++) acc[kb][i][J] = lp_mfma<PREC>(a[kb], bj, acc[kb][i][J]);
                    };
                    tap(std::integral_constant<int, 0>{});
                    if constexpr (OD > 1) {
                        tap(std::integral_constant<int, 1>{});
                        tap(std::integral_constant<int, 2>{});
                    }
                    if constexpr (OD > 3) {
                        tap(std::integral_constant<int, 3>{});
                        tap(std::integral_constant<int, 4>{});
                    }
                }
                {
                    // dy buffer (t + 1) & 1 and the ring row of x row y - 1 were
                    // last read by the previous step, which every wave finished
                    // before that step's barrier.  (After the last row these
                    // stores write rows that no step reads.)
                    const int nx = (u + 1) % FR_PD;  // a constant once the group loop is unrolled
                    store_dy((t + 1) & 1, qd[nx]);
                    int rw = ring0 + OD;
                    rw -= rw >= RB ? RB : 0;
                    store_x(rw, qx[nx]);
                }
                __syncthreads();
                ring0 = ring0 + 1 == RB ? 0 : ring0 + 1;
            }
        }
    }
    };
    // The blocks of one image start their walk at a row of their own, 11 b
    // modulo the run, and wrap: blocks that run together then read rows that
    // lie apart in memory (plane strides are often multiples of a large power
    // of two, so rows at one position of every plane share their low address
    // bits).  The last step of a walk ends with a barrier, so the second
    // walk's prologue may overwrite the LDS rows at once.
    const int run = yEnd - yBeg;
    const int off = run > 0 ? (b * 11) % run : 0;
    stream(yBeg + off, yEnd);
    if (off) stream(yBeg, yBeg + off);
    // The split's partial: ws[s][k][n][i][j] (every (k, n, i, j) of the tile,
    // zeros for a block without rows).
    const size_t total = (size_t)K * N * P;
    float* out = ws + (size_t)s * total;
    const int n = nt * NT + nb * 16 + l16;
#pragma unroll
    for (int kb = 0; kb < KBW; kb++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = kt * KT + (kh * KBW + kb) * 16 + 4 * g4 + v;
            if (k < K && n < N) {
                float* o = out + ((size_t)k * N + n) * P;
#pragma unroll
                for (int i = 0; i < OD; i++)
#pragma unroll
                    for (int j = 0; j < OD; j++) o[i * OD + j] = acc[kb][i][j][v];
            }
        }
}

size_t fa_rows_lp_lds_bytes(const FaCfg& c, int OD) {
    const int KT = c.kh * c.kbw * 16, NT = c.nb * 16;
    return (size_t)2 * 4 * KT * 16 + (size_t)(OD + 1) * FR_XS * NT * 16 + 16;  // + the junk slot
}

// Tiles offered (x-plane blocks NB, dy-plane block groups KH, dy-plane blocks a wave KBW).
template <int PREC, int OD>
static void fa_rows_launch2(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                            float* ws, hipStream_t s, dim3 grid) {
    const int sel = c.nb * 100 + c.kh * 10 + c.kbw;
    const dim3 block(64 * c.nb * c.kh);
    // (5x5 dW: one dy block a wave, as the planner requires.)
#define FR_CASE(NB_, KH_, KBW_)                                                                                  \
    case NB_ * 100 + KH_ * 10 + KBW_:                                                                           \
        if constexpr (OD != 5 || KBW_ == 1)                                                                     \
            hipLaunchKernelGGL((k_fa_rows_lp<PREC, OD, KBW_, NB_, KH_>), grid, block, c.lds_bytes, s, g, m, c, x, dy, \
                               ws);                                                                              \
        break;
    switch (sel) {
        FR_CASE(2, 2, 2)
        FR_CASE(2, 2, 1)
        FR_CASE(1, 4, 2)
        FR_CASE(1, 4, 1)
        FR_CASE(4, 2, 2)
        FR_CASE(4, 2, 1)
        FR_CASE(4, 1, 4)
    default: break;
    }
#undef FR_CASE
}

template <int PREC>
static hipError_t fa_rows_launch1(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                                  float* ws, hipStream_t s, dim3 grid) {
    switch (g.OH) {
    case 1: fa_rows_launch2<PREC, 1>(g, m, c, x, dy, ws, s, grid); break;
    case 3: fa_rows_launch2<PREC, 3>(g, m, c, x, dy, ws, s, grid); break;
    case 5: fa_rows_launch2<PREC, 5>(g, m, c, x, dy, ws, s, grid); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fa_rows_lp(int precision, const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x,
                             const float* dy, float* out, float* ws, hipStream_t s, const Epi& ep) {
    const long blocks = (long)c.nTiles * c.kTiles * c.splits;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks);
    hipError_t e;
    switch (precision) {
    case COGCONV_BF16: e = fa_rows_launch1<COGCONV_BF16>(g, m, c, x, dy, ws, s, grid); break;
    case COGCONV_FP16: e = fa_rows_launch1<COGCONV_FP16>(g, m, c, x, dy, ws, s, grid); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    const long total = (long)c.K * m.N * g.OH * g.OW;
    return launch_reduce_partials(ws, out, c.splits, total, (long)g.OH * g.OW, s, ep);
}

}  // namespace cogconv
