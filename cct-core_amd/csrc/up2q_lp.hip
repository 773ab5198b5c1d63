// 2x up-sampled ProjectFrame / BackProjectFrame with block reduction in BF16 /
// FP16 (FP32 accumulation): the four output phases in one launch, as the FP32
// k_igemm_f32_up2q does, on v_mfma_f32_16x16x32_{bf16,f16}.
//
// Output pixel (2R + pr, 2C + pc) of the zero-inserted input is a dense
// convolution of the stored samples with the sub-filter of taps
// (i0(pr) + 2a, j0(pc) + 2b), i0(p) = (h - p) mod 2 (ConvolveHyperKernel.scala:
// 205-233 and 404-459 read the same zero-inserted input).  A block stages, per
// 32-plane chunk, one window of stored samples (TRp x TCp = (8 + UR - 1) x
// (32 + UC - 1) pixels for 8 x 32 phase pixels) rounded to the MFMA type, and
// the chunk's whole packed filter slice (k_pack_filters_lp's layout, taps in
// the kernel's orientation); every phase reads its taps from them.
//
// LDS holds both operands quarter-major, as the BF16 frame kernel does: the
// window [4 quarters][npix][8 planes], npix padded to 16 slots, and the filter
// [taps][4][MT][8].  Waves w % 4 own phase-pixel rows 2(w%4), 2(w%4)+1; w / 4
// is the row phase.  A wave's MFMAs take the pixels as A and the filters as B,
// so lane (g4, l16) of block (mb, pb) holds output plane 16 mb + l16 at phase
// pixels 4 g4 .. 4 g4 + 3, for column phase 0 in acc[0] and 1 in acc[1]: eight
// adjacent output columns, two 16-byte stores.
#include "epi.h"
#include "lp.h"
#include "stage.h"

namespace cogconv {

constexpr int UQ_RC = 32;  // planes per staged chunk (one MFMA k slice)
constexpr int UQ_OOB = (int)0x80000000u;

template <int F, int HP, int PR, int MB, int PREC>
__device__ __forceinline__ void up2q_lp_chunk(const typename LpType<PREC>::T* sB, const typename LpType<PREC>::T* sA,
                                              int npix, int TCp, int dr, int dc0, int dc1, const int (&boff)[4],
                                              int g4, int l16, floatx4 (&acc)[2][MB][4]) {
    typedef typename LpType<PREC>::V V;
    constexpr int MT = 16 * MB;
    constexpr int FRp = up2_taps(F, HP, PR), FC0 = up2_taps(F, HP, 0), FC1 = up2_taps(F, HP, 1);
    constexpr int I0 = (HP ^ PR) & 1;
    const auto* ab = sA + ((size_t)g4 * MT + l16) * 8;
    // Two output blocks: one sub-filter row at a time (the operand loads the
    // compiler hoists over a whole unrolled sub-filter would spill).
    constexpr int UNROLL = MB == 2 ? 1 : FRp;
#pragma unroll UNROLL
    for (int a = 0; a < FRp; a++) {
        const auto* bRow = sB + ((size_t)g4 * npix + (a + dr) * TCp) * 8;
#pragma unroll
        for (int pc = 0; pc < 2; pc++) {
            const int J0 = (HP ^ pc) & 1;
            const auto* bq = bRow + (pc ? dc1 : dc0) * 8;
#pragma unroll
            for (int bc = 0; bc < (pc ? FC1 : FC0); bc++) {
                const int t = (I0 + 2 * a) * F + J0 + 2 * bc;  // the tap in the filter's own order
                V av[MB], bv[4];
#pragma unroll
                for (int mb = 0; mb < MB; mb++)
                    av[mb] = *reinterpret_cast<const V*>(ab + ((size_t)t * 4 * MT + mb * 16) * 8);
#pragma unroll
                for (int pb = 0; pb < 4; pb++) bv[pb] = *reinterpret_cast<const V*>(bq + boff[pb] + bc * 8);
#pragma unroll
                for (int mb = 0; mb < MB; mb++)
#pragma unroll
                    for (int pb = 0; pb < 4; pb++) acc[pc][mb][pb] = lp_mfma<PREC>(bv[pb], av[mb], acc[pc][mb][pb]);
            }
        }
    }
}

template <int F, int HP, int MB, int PREC>
__global__ __launch_bounds__(512) void k_up2q_lp(Geom g, PlaneMap m, UpFused u, const float* __restrict__ X,
                                                 const typename LpType<PREC>::T* __restrict__ Wp,
                                                 float* __restrict__ O, Epi e) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr int TAPS = F * F;
    constexpr int MT = 16 * MB;
    const IgemmCfg& c = u.c;
    extern __shared__ __attribute__((aligned(16))) unsigned char uq_smem[];
    const int TRp = c.tileRows, TCp = c.tileCols;
    const int npix = (TRp * TCp + 15) & ~15;  // quarter stride padded to 16 slots (LDS banks)
    T* sB = reinterpret_cast<T*>(uq_smem);     // [4][npix][8]
    T* sA = sB + (size_t)4 * npix * 8;         // [TAPS][4][MT][8]
    T* sJunk = sA + (size_t)TAPS * 4 * MT * 8;  // 64 lanes x 8 elements for lanes past the window

    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int mt = bid % c.mTiles;  // output-plane blocks of one tile: adjacent indices
    bid /= c.mTiles;
    const int tc = bid % c.tilesC;
    bid /= c.tilesC;
    const int tr = bid % c.tilesR;
    const int b = bid / c.tilesR;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rw = wave & 3, pr = wave >> 2;
    const int g4 = lane >> 4, l16 = lane & 15;
    const int Mout = c.Mout;
    const int m0 = mt * MT;
    const int y0 = tr * 8 - u.hU;
    const int x0 = tc * 32 - u.wU;
    const int HW = g.H * g.W;
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + (size_t)b * m.N * HW, (uint32_t)m.N * HW * 4u);
    const int Mpad = c.mTiles * MT;
    const __amdgpu_buffer_rsrc_t wr = buffer_rsrc(Wp, (uint32_t)((size_t)TAPS * Mpad * c.RC * 2));
    // The lane's window column (the window is at most 64 columns wide) and
    // each window row's offset, with the border applied; planes past the
    // image's read 0 (past the buffer's range).
    const int colByte = lane < TCp ? stage_col_offset(g, x0 + lane) : UQ_OOB;
    int yOffLane;
    uint64_t yMask;
    {
        int y = y0 + lane;
        bool in = (unsigned)y < (unsigned)g.H;
        if (g.border == COGCONV_BORDER_CYCLIC) {
            y %= g.H;
            y += y < 0 ? g.H : 0;
            in = true;
        } else {
            in = in || g.border == COGCONV_BORDER_CLAMP;
            y = min(max(y, 0), g.H - 1);
        }
        yOffLane = y * g.W;
        yMask = __ballot(in);
    }
    const int planeBytes = 4 * HW;

    floatx4 acc[2][MB][4];
#pragma unroll
    for (int pc = 0; pc < 2; pc++)
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[pc][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    int boff[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) boff[pb] = ((2 * rw + (pb >> 1)) * TCp + (pb & 1) * 16 + l16) * 8;

    const int nitems = 4 * TRp;  // (8-plane quarter, window row)
    constexpr int IU = 3;        // items a wave has in flight
    const int units = TAPS * 4 * MT;
    for (int r0 = 0; r0 < c.RC; r0 += UQ_RC) {
        __syncthreads();
        // Window: items (quarter q, row ty), one per wave pass; a lane loads
        // the eight planes of its column and writes one 16-byte slot.
        for (int base = wave; base < nitems; base += 8 * IU) {
            float v[IU][8];
#pragma unroll
            for (int k = 0; k < IU; k++) {
                const int it = base + 8 * k;
                const int q = it / TRp, ty = it - q * TRp;
                const int tyc = min(ty, 63);
                // A quarter that starts past the image's planes reads nothing
                // (its offsets would pass the buffer's range by up to 31 planes).
                const bool rowOn = it < nitems && ((yMask >> tyc) & 1) && r0 + 8 * q < m.N;
                const int rowByte = 4 * __builtin_amdgcn_readlane(yOffLane, tyc);
                int vo = rowOn ? colByte + planeBytes * (r0 + 8 * q) : UQ_OOB;
                vo = colByte == UQ_OOB ? UQ_OOB : vo;
#pragma unroll
                for (int p = 0; p < 8; p++) {
                    v[k][p] = buffer_load_f32_nt(xr, vo, rowByte);
                    vo += planeBytes;
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // every load ahead of the first store
#pragma unroll
            for (int k = 0; k < IU; k++) {
                const int it = base + 8 * k;
                if (it < nitems) {  // wave-uniform
                    const int q = it / TRp, ty = it - q * TRp;
                    V h;
#pragma unroll
                    for (int p = 0; p < 8; p++) h[p] = (T)v[k][p];
                    T* d = lane < TCp ? sB + ((size_t)q * npix + ty * TCp + lane) * 8 : sJunk + 8 * lane;
                    *reinterpret_cast<V*>(d) = h;
                }
            }
        }
        // Filter slice of the chunk: 16-byte units (tap, quarter, output
        // plane), output plane fastest, copied from the packed bank.
        const int chunkOff = (r0 / UQ_RC) * TAPS * 4 * Mpad * 16;  // bytes
        for (int u0 = tid; u0 < units; u0 += 512 * 4) {
            V w[4];
            int dst[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int uu = u0 + 512 * k;
                const int mm = uu % MT, tq = uu / MT;  // tq = tap * 4 + quarter
                const bool ok = uu < units;
                const int off = ok ? ((tq * Mpad + m0 + mm) * 16) : UQ_OOB;
                w[k] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(wr, off, chunkOff, 0));
                dst[k] = ok ? uu * 8 : -1;
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (dst[k] >= 0) *reinterpret_cast<V*>(sA + dst[k]) = w[k];
        }
        __syncthreads();
        if (pr == 0)
            up2q_lp_chunk<F, HP, 0, MB, PREC>(sB, sA, npix, TCp, u.dr[0], u.dc[0], u.dc[1], boff, g4, l16, acc);
        else
            up2q_lp_chunk<F, HP, 1, MB, PREC>(sB, sA, npix, TCp, u.dr[1], u.dc[0], u.dc[1], boff, g4, l16, acc);
    }
    // Epilogue (as in k_igemm_f32_up2q): output columns 2C .. 2C + 7 of row 2R + pr.
    const bool vec = (g.OW & 3) == 0 && (reinterpret_cast<uintptr_t>(O) & 15) == 0;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
        const int mm = m0 + mb * 16 + l16;
#pragma unroll
        for (int pb = 0; pb < 4; pb++) {
            const int R = tr * 8 + 2 * rw + (pb >> 1);
            const int C = tc * 32 + (pb & 1) * 16 + 4 * g4;
            if (mm >= Mout || R >= u.OHp[pr] || C >= u.OWp[0]) continue;
            const int orow = 2 * R + pr, ocol = 2 * C;
            float* dst = O + ((size_t)(b * Mout + mm) * g.OH + orow) * g.OW + ocol;
            const floatx4 s0 = acc[0][mb][pb], s1 = acc[1][mb][pb];
            if (vec && ocol + 8 <= g.OW) {
                store_out4(dst, s0[0], s1[0], s0[1], s1[1], e);
                store_out4(dst + 4, s0[2], s1[2], s0[3], s1[3], e);
            } else {
#pragma unroll
                for (int v = 0; v < 8; v++)
                    if (ocol + v < g.OW) dst[v] = (v & 1) ? s1[v >> 1] : s0[v >> 1];
            }
        }
    }
    if (e.n) {
        for (int t = 0; t < MB * 4; t++) {
            const int mb = t >> 2, pb = t & 3;
            const int mm = m0 + mb * 16 + l16;
            const int R = tr * 8 + 2 * rw + (pb >> 1);
            const int C = tc * 32 + (pb & 1) * 16 + 4 * g4;
            if (mm >= Mout || R >= u.OHp[pr] || C >= u.OWp[0]) continue;
            const int orow = 2 * R + pr, ocol = 2 * C;
            float* dst = O + ((size_t)(b * Mout + mm) * g.OH + orow) * g.OW + ocol;
            const int pt = orow * g.OW + ocol;
            if (vec && ocol + 8 <= g.OW) {
                epi_rewrite4(e, dst, b * Mout + mm, pt);
                epi_rewrite4(e, dst + 4, b * Mout + mm, pt + 4);
            } else {
                for (int v = 0; v < 8 && ocol + v < g.OW; v++) epi_rewrite1(e, dst + v, b * Mout + mm, pt + v);
            }
        }
    }
}

template <int F, int HP, int PREC>
static hipError_t launch_uq_mb(const Geom& g, const PlaneMap& m, const UpFused& u, const float* x, const void* wp, float* out,
                        hipStream_t s, dim3 grid, const Epi& e) {
    typedef typename LpType<PREC>::T T;
    if (u.c.MB == 1)
        hipLaunchKernelGGL((k_up2q_lp<F, HP, 1, PREC>), grid, dim3(512), u.c.lds_bytes, s, g, m, u, x,
                           static_cast<const T*>(wp), out, e);
    else
        hipLaunchKernelGGL((k_up2q_lp<F, HP, 2, PREC>), grid, dim3(512), u.c.lds_bytes, s, g, m, u, x,
                           static_cast<const T*>(wp), out, e);
    return hipGetLastError();
}

template <int PREC>
static hipError_t launch_uq(const Geom& g, const PlaneMap& m, const UpFused& u, const float* x, const void* wp, float* out,
                     hipStream_t s, dim3 grid, const Epi& e) {
    switch (g.FH * 2 + (u.hp & 1)) {
    case 6: return launch_uq_mb<3, 0, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 7: return launch_uq_mb<3, 1, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 10: return launch_uq_mb<5, 0, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 11: return launch_uq_mb<5, 1, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 14: return launch_uq_mb<7, 0, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 15: return launch_uq_mb<7, 1, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 18: return launch_uq_mb<9, 0, PREC>(g, m, u, x, wp, out, s, grid, e);
    case 19: return launch_uq_mb<9, 1, PREC>(g, m, u, x, wp, out, s, grid, e);
    default: return hipErrorInvalidValue;
    }
}

size_t up2q_lp_lds_bytes(const IgemmCfg& c, int F) {
    const size_t npix = ((size_t)c.tileRows * c.tileCols + 15) & ~(size_t)15;
    return (npix * 4 * 8 + (size_t)F * F * 4 * 16 * c.MB * 8 + 64 * 8) * 2;
}

size_t up2q_lp_workspace_bytes(const IgemmCfg& c, int F) { return (size_t)F * F * c.mTiles * 16 * c.MB * c.RC * 2; }

hipError_t launch_up2q_lp(int precision, const Geom& g, const PlaneMap& m, const UpFused& u, const float* x,
                          const float* w, float* out, void* ws, hipStream_t s, const Epi& e) {
    const IgemmCfg& c = u.c;
    const long blocks = (long)c.tilesC * c.tilesR * c.mTiles * m.B;
    if (blocks <= 0 || blocks > 0x7fffffffL || c.tileCols > 64 || c.tileRows > 64 || (c.MB != 1 && c.MB != 2) ||
        !ws)
        return hipErrorInvalidConfiguration;
    // The filter bank in the kernels' packed layout (the BF16 frame kernel's pack).
    hipError_t err = launch_pack_filters_lp(precision, g, m, c.Mout, c.mTiles * 16 * c.MB, c.RC, w, ws, s);
    if (err != hipSuccess) return err;
    const dim3 grid((unsigned)blocks);
    if (precision == COGCONV_FP16) return launch_uq<COGCONV_FP16>(g, m, u, x, ws, out, s, grid, e);
    if (precision == COGCONV_BF16) return launch_uq<COGCONV_BF16>(g, m, u, x, ws, out, s, grid, e);
    return hipErrorInvalidValue;
}

}  // namespace cogconv
