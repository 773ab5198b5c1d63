// Precision variants of the frame block-reduce modes: BF16 / FP16 MFMA
// implicit GEMM with FP32 accumulation, and split-BF16 (three-term compensated
// products).  Same contraction as igemm_f32.hip:
//
//   out[b*Mout + m](y, x) = sum_{r < R} sum_{t < F*F} A[m][r][t] * X_{b*N + r}(y*dn - h + ti, x*dn - h + tj)
//
// with A from the filter bank (ProjectFrame: W[m*N + r], BackProjectFrame:
// W[r*L + m], flipped for the convolution orientation;
// ConvolveHyperKernel.scala:153-167, 327-332) and borders applied while the
// input tile is staged (ConvolveHyperKernel.scala:235-317).  Operands arrive
// in FP32 (the Field storage type) and are rounded to the MFMA input type
// while they are staged.
//
// gfx950 mapping: v_mfma_f32_16x16x32_{bf16,f16}.  Lane l of an MFMA holds
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16]: the 32-deep k slice is
// 32 consecutive input planes at one tap, so LDS holds both operands
// planes-fastest ("channels last"): a lane's eight k values are one 16-byte
// read.  A block (4 wave64) owns an 8 x 32 pixel tile and 16*MB output planes;
// wave w owns pixel rows 2w, 2w+1.  Input planes are staged 32 at a time.  The
// MFMAs take the pixels as A and the filters as B, so a lane's accumulator
// holds four adjacent pixels of one output plane (direct 16-byte stores).
//
// Split-BF16: x = xh + xl with xh = bf16(x), xl = bf16(x - xh) (same for the
// filter); x*w ~ xh*wh + xh*wl + xl*wh, each term on the MFMA, all into the
// same FP32 accumulator.  BF16 has FP32's exponent range, so the split needs
// no rescaling except for values below ~2^-118 whose low part underflows.
#include "epi.h"
#include "lp.h"
#include "stage.h"

namespace cogconv {

constexpr int LP_RC = 32;   // planes per staged chunk (one MFMA k slice)
// Lane byte offset with bit 31 set: past every staged buffer (each is under
// 2^31 bytes), so the buffer unit's range check answers the load with 0.
constexpr int LP_OOB = (int)0x80000000u;

// Filter pre-pack: Wp[term][chunk][t][q][m][e] in the MFMA input type, for
// input plane r = 32 chunk + 8 q + e (zeros past the planes), output m
// padded to Mpad (zeros), taps flipped for the convolution orientation:
// the order in which a block stages a tap group into LDS, so that staging is
// a contiguous copy.  term 0 = rounded value, term 1 (split only) = rounded
// residual.
template <int PREC>
__global__ void k_pack_filters_lp(Geom g, PlaneMap m, int Mout, int Mpad, int Rpad, const float* __restrict__ Wt,
                                  typename LpType<PREC>::T* __restrict__ Wp) {
    typedef typename LpType<PREC>::T T;
    const int taps = g.FH * g.FW;
    const long total = (long)taps * Mpad * Rpad;
    const bool bpf = m.mode == COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM;
    const int L = m.T / m.N;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int e8 = (int)(idx & 7);
        long rest = idx >> 3;
        const int mm = (int)(rest % Mpad);
        rest /= Mpad;
        const int q = (int)(rest & 3);
        rest >>= 2;
        const int t = (int)(rest % taps);
        const int r = (int)(rest / taps) * LP_RC + 8 * q + e8;
        float w = 0.f;
        if (r < m.R && mm < Mout) {
            const int pf = bpf ? r * L + mm : mm * m.N + r;
            w = Wt[(size_t)pf * taps + (g.flip ? taps - 1 - t : t)];
        }
        const T hi = (T)w;
        Wp[idx] = hi;
        if (PREC == COGCONV_SPLIT_BF16) Wp[total + idx] = (T)(w - (float)hi);
    }
}

template <int F, int MB, int PREC, bool RING>
__device__ __forceinline__ void igemm_lp_body(const Geom& g, const PlaneMap& m, const IgemmCfg& c,
                                              const float* __restrict__ X,
                                              const typename LpType<PREC>::T* __restrict__ Wp,
                                              float* __restrict__ O, const cogconv_image_io* __restrict__ tab,
                                              const Epi& e) {
    typedef typename LpType<PREC>::T T;
    typedef typename LpType<PREC>::V V;
    constexpr int TAPS = F * F;
    constexpr int MT = 16 * MB;
    constexpr bool SPLIT = PREC == COGCONV_SPLIT_BF16;
    constexpr int TERMS = SPLIT ? 2 : 1;
    extern __shared__ __attribute__((aligned(16))) unsigned char lp_smem[];
    const int TRp = c.tileRows, TCp = c.tileCols;
    // Slots of one 8-plane quarter of the tile, padded to a multiple of 16
    // (quarters start at multiples of 256 bytes): the four quarter groups of
    // lanes of one read then meet no bank conflicts, where strides of 532 or
    // 340 slots measured about three a read.
    const int npix = (TRp * TCp + 15) & ~15;
    // LDS holds both operands quarter-major: 8 planes (16 bytes) of one pixel
    // or output plane per slot, the four 8-plane quarters of a 32-plane chunk
    // in separate planes of slots, so that the 16 lanes of an MFMA row group
    // read 16 consecutive slots (no padding, no bank conflicts).  The filter
    // slice holds TRG tap rows at a time (all F when they fit); RING: one tap
    // row at a time in two buffers, the next row's loads in flight while the
    // current row's MFMAs run.
    const int TRG = RING ? 1 : c.trg;
    const int TAPG = TRG * F;
    const int gElems = TERMS * TAPG * 4 * MT * 8;             // one filter group (buffer)
    T* sB = reinterpret_cast<T*>(lp_smem);                    // [TERMS][4][npix][8], npix padded
    T* sA = sB + (size_t)TERMS * 4 * npix * 8;               // [RING ? 2 : 1][TERMS][TAPG][4][MT][8]
    // Scratch slots (offset from sB) for the stores of lanes beyond the tile
    // width, after the filter slice(s): TERMS x 64 lanes x 8 elements.
    const size_t sJunk = (size_t)TERMS * 4 * npix * 8 + (size_t)(RING ? 2 : 1) * gElems;

    // Output-plane blocks of one tile are adjacent block indices (same XCD,
    // close in time): the tile's later blocks find its input in L2.
    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int mt = bid % c.mTiles;
    bid /= c.mTiles;
    const int tc = bid % c.tilesC;
    bid /= c.tilesC;
    const int tr = bid % c.tilesR;
    const int b = bid / c.tilesR;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g4 = lane >> 4, l16 = lane & 15;
    const int Mout = c.Mout;
    const int m0 = mt * MT;
    const int y0 = tr * 8 * g.dnR - g.hR;
    const int x0 = tc * 32 * g.dnC - g.hC;
    const int HW = g.H * g.W;
    if (tab) {
        // Request table: image b's planes and output planes are entry b's buffers.
        // X and O are re-based so that the image-b offsets below land on them.
        X = reinterpret_cast<const float*>(reinterpret_cast<uintptr_t>(tab[b].in0) - (uintptr_t)b * m.N * HW * 4);
        O = reinterpret_cast<float*>(reinterpret_cast<uintptr_t>(tab[b].out) - (uintptr_t)b * Mout * g.OH * g.OW * 4);
    }
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + (size_t)b * m.N * HW, (uint32_t)m.N * HW * 4u);
    const int Mpad = c.mTiles * MT;
    const int Rpad = c.RC;  // planes padded to a multiple of 32
    const long wTerm = (long)TAPS * Mpad * Rpad;
    const int wChunkBytes = 2 * TAPS * 4 * Mpad * 8;  // one 32-plane chunk of one term
    const __amdgpu_buffer_rsrc_t wr = buffer_rsrc(Wp, (uint32_t)(TERMS * wTerm * 2));  // the plan bounds this
    const bool bordered = g.border == COGCONV_BORDER_CLAMP || g.border == COGCONV_BORDER_CYCLIC;

    // Per-lane column offset of the staged tile (TCp <= 64) with the border applied.
    int colOff;
    bool colIn;
    {
        int x = x0 + lane;
        bool in = (unsigned)x < (unsigned)g.W;
        if (g.border == COGCONV_BORDER_CYCLIC) {
            x %= g.W;
            x += x < 0 ? g.W : 0;
            in = true;
        } else {
            in = in || g.border == COGCONV_BORDER_CLAMP;
            x = min(max(x, 0), g.W - 1);
        }
        colOff = x;
        colIn = in && lane < TCp;
    }
    const int colByte = colIn ? 4 * colOff : LP_OOB;

    // 16-byte staging (c.vec4): the staged columns re-based to xa = x0 rounded
    // down to a multiple of 4; lane (si, gc) loads the 4-column group gc of
    // item slot si (IP items of NG groups per wave pass), eight planes as
    // eight 16-byte loads, and writes the group's (up to) four pixels.  With
    // W % 4 == 0 a group lies wholly inside or wholly outside the field; for
    // Clamp and Cyclic the groups outside are left to a second pass that
    // loads each of their columns from its bordered position.
    const int s4 = ((x0 % 4) + 4) % 4;
    const int NG = (s4 + TCp + 3) >> 2;
    const int IP = 64 / NG;
    const int si = lane / NG, gc = lane - si * NG;
    const int colA = x0 - s4 + 4 * gc;
    const bool gIn = si < IP && (unsigned)colA < (unsigned)g.W;
    const float invTRp = 1.0f / TRp;

    // Filter-slice units (term, tap, 8-plane quarter, output plane), output
    // plane fastest: consecutive threads copy consecutive 16 bytes of the
    // packed bank into consecutive LDS slots.  256 is a multiple of the units
    // of one tap (4 MT), so a thread's quarter and output plane are fixed and
    // its i-th unit lies TPI taps after its first: offsets are affine in i.
    constexpr int UPT = 4 * MT;
    constexpr int TPI = 256 / UPT;
    const int uLow = tid % UPT;                     // quarter * MT + output plane
    const int tg0 = tid / UPT;
    const int unitCol = 2 * (((uLow / MT) * Mpad + m0 + uLow % MT) * 8);  // bytes within a tap
    // Byte offset in the packed bank (the chunk adds its offset as the uniform
    // offset; taps past the filter read 0 through an out-of-range offset) and
    // LDS destination (elements from sB; units past the group go to the
    // scratch slot) of the thread's i-th unit of a group of `tapg` taps
    // starting at tap `tap0`, written into the buffer at `buf` (elements from sB).
    auto group_unit = [&](int i, int tap0, int tapg, int buf, int& off, int& dst) {
        const int L = tg0 + TPI * i;                 // linear tap over [TERMS][tapg]
        const int term = SPLIT && L >= tapg ? 1 : 0;
        const int tg = L - term * tapg;
        const bool live = L < TERMS * tapg;
        off = live && tap0 + tg < TAPS ? (int)(2 * term * wTerm) + (tap0 + tg) * (2 * 4 * Mpad * 8) + unitCol
                                       : LP_OOB;
        dst = live ? buf + (L * UPT + uLow) * 8 : (int)(sJunk + 8 * lane);
    };
    constexpr int UNITS = TERMS * TAPS * MT * 4;
    constexpr int WPT = (UNITS + 255) / 256;
    constexpr bool WPRE = !RING && WPT <= 12;
    const bool onePass = TRG == F;  // block-uniform
    const int sAoff = (int)((size_t)TERMS * 4 * npix * 8);  // elements from sB to the filter buffer(s)
    int wOff[WPRE ? WPT : 1], wDst[WPRE ? WPT : 1];
    if (WPRE && onePass) {
#pragma unroll
        for (int i = 0; i < (WPRE ? WPT : 1); i++) group_unit(i, 0, TAPS, sAoff, wOff[i], wDst[i]);
    }
    // RING: the units of one tap row (per thread), prefetched into registers.
    constexpr int WPG = (TERMS * F + TPI - 1) / TPI;
    int rOff[RING ? WPG : 1], rDst[2][RING ? WPG : 1];
    if (RING) {
#pragma unroll
        for (int i = 0; i < (RING ? WPG : 1); i++) {
            group_unit(i, 0, F, sAoff, rOff[i], rDst[0][i]);
            int o2;
            group_unit(i, 0, F, sAoff + gElems, o2, rDst[1][i]);
        }
    }

    // Per-lane table of the staged tile's row offsets (lane = tile row, at
    // most 64 rows) with the border applied, and the rows' in-field mask.
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

    floatx4 acc[MB][4];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    int boff[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
        const int row = 2 * wave + (pb >> 1);
        const int col = (pb & 1) * 16 + l16;
        boff[pb] = (g4 * npix + (row * g.dnR) * TCp + col * g.dnC) * 8;
    }

    // Input tile: items (8-plane group, row); each lane loads the eight planes
    // of its column, rounds them and writes one 16-byte LDS vector.  The loads
    // of IU items are unconditional and all in flight before the first is
    // used; lanes beyond the tile width store into a scratch slot, so no
    // lane-dependent branch separates the loads.  Masked samples (rows or
    // columns outside the field for Zero/Full/Valid borders) read through an
    // out-of-range lane offset, which the buffer unit answers with 0: one
    // select per item instead of one per sample.  The plane goes into the lane
    // offset and the row into the uniform offset, so planes at or past R (=
    // the image's planes) fall outside the image's buffer and read 0 too; the
    // plan keeps every such offset below 2^32.  A wave owns TRp items of a
    // chunk (7 dn + F); IU covers all of them when dn = 1 and F <= 3, so the
    // chunk's staging costs one memory round trip.
    const int nitems = 4 * TRp;
    // Ring kernels keep fewer items in flight (registers for two blocks a CU).
    constexpr int IU = RING ? (F + 7 < 4 ? F + 7 : 4) : (F + 7 < 10 ? F + 7 : 10);
    const int planeBytes = 4 * HW;
    const bool inTile = lane < TCp;
    auto issue = [&](int r0, int base, float (&v)[IU][8]) {
        int grp = base / TRp;  // wave-uniform; TRp >= 8, so one wrap per step of 4
        int ty = base - grp * TRp;
#pragma unroll
        for (int u = 0; u < IU; u++) {
            const int tyc = min(ty, 63);
            const bool rowOn = base + 4 * u < nitems && ((yMask >> tyc) & 1);
            const int rowByte = 4 * __builtin_amdgcn_readlane(yOffLane, tyc);
            int vo = rowOn ? colByte + planeBytes * (r0 + 8 * grp) : LP_OOB;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                v[u][e] = buffer_load_f32_nt(xr, vo, rowByte);
                vo += planeBytes;
            }
            ty += 4;
            if (ty >= TRp) { ty -= TRp; grp++; }
        }
    };
    auto commit = [&](int base, const float (&v)[IU][8]) {
        int grp = base / TRp;
        int ty = base - grp * TRp;
#pragma unroll
        for (int u = 0; u < IU; u++) {
            if (base + 4 * u < nitems) {  // wave-uniform
                V hv, lv;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const T h = (T)v[u][e];
                    hv[e] = h;
                    if (SPLIT) lv[e] = (T)(v[u][e] - (float)h);
                }
                const size_t o = ((size_t)grp * npix + ty * TCp + lane) * 8;
                *reinterpret_cast<V*>(sB + (inTile ? o : sJunk + 8 * lane)) = hv;
                if (SPLIT) *reinterpret_cast<V*>(sB + (inTile ? (size_t)4 * npix * 8 + o : sJunk + 512 + 8 * lane)) = lv;
            }
            ty += 4;
            if (ty >= TRp) { ty -= TRp; grp++; }
        }
    };
    // Halo columns of a Clamp / Cyclic tile staged with 16-byte loads: the
    // staged columns left of 0 and right of W - 1 (block-uniform counts).
    const int nhl = c.vec4 && bordered ? min(max(-x0, 0), TCp) : 0;
    const int nhr = c.vec4 && bordered ? min(max(x0 + TCp - g.W, 0), TCp - nhl) : 0;
    const int nh = nhl + nhr;
    const int rowBytesW = 2 * F * 4 * Mpad * 8;  // one tap row of one term in the packed bank
    for (int r0 = 0; r0 < m.R; r0 += LP_RC) {
        const int wChunk = (r0 / LP_RC) * wChunkBytes;  // uniform byte offset of the chunk's packed filters
        __syncthreads();
        V wv[WPRE ? WPT : 1];
        if (WPRE && onePass) {
#pragma unroll
            for (int i = 0; i < (WPRE ? WPT : 1); i++)
                wv[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(wr, wOff[i], wChunk, 0));
        }
        V rv[RING ? WPG : 1];
        if (RING) {
#pragma unroll
            for (int i = 0; i < (RING ? WPG : 1); i++)
                rv[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(wr, rOff[i], wChunk, 0));
        }
        if (c.vec4) {  // block-uniform
            constexpr int VP = (F + 12) / 6;  // passes: a wave owns TRp = F + 7 items, IP >= 6 a pass
            constexpr int VG = RING ? 1 : VP;   // passes whose loads are in flight at once
#pragma unroll
            for (int p0 = 0; p0 < VP; p0 += VG) {
            float4 q[VG][8];
            int dBase[VG];
#pragma unroll
            for (int pp = 0; pp < VG; pp++) {
                const int j = (p0 + pp) * IP + si;  // the wave's item slot
                const int it = wave + 4 * j;
                const int grp = fast_div(it, TRp, invTRp), ty = it - grp * TRp;
                const bool itemOk = si < IP && j < TRp;
                int y = y0 + ty;
                bool yIn = (unsigned)y < (unsigned)g.H;
                if (g.border == COGCONV_BORDER_CLAMP) {
                    y = min(max(y, 0), g.H - 1);
                    yIn = true;
                } else if (g.border == COGCONV_BORDER_CYCLIC) {
                    y %= g.H;
                    y += y < 0 ? g.H : 0;
                    yIn = true;
                }
                const bool on = itemOk && gIn && yIn;
                int vo = on ? 4 * ((r0 + 8 * grp) * HW + y * g.W + colA) : LP_OOB;
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    q[pp][e] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(xr, vo, 0, LOAD_NT));
                    vo += planeBytes;
                }
                // Groups outside a Clamp / Cyclic field are written by the halo pass.
                dBase[pp] = itemOk && (gIn || !bordered) ? (grp * npix + ty * TCp) * 8 : -1;
            }
            __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
#pragma unroll
            for (int pp = 0; pp < VG; pp++) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int col = 4 * gc + k - s4;
                    V hv, lv;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const float x = k == 0 ? q[pp][e].x : (k == 1 ? q[pp][e].y : (k == 2 ? q[pp][e].z : q[pp][e].w));
                        const T h = (T)x;
                        hv[e] = h;
                        if (SPLIT) lv[e] = (T)(x - (float)h);
                    }
                    const bool ok = dBase[pp] >= 0 && (unsigned)col < (unsigned)TCp;
                    const size_t o = (size_t)dBase[pp] + (size_t)col * 8;
                    *reinterpret_cast<V*>(sB + (ok ? o : sJunk + 8 * lane)) = hv;
                    if (SPLIT) *reinterpret_cast<V*>(sB + (ok ? (size_t)4 * npix * 8 + o : sJunk + 512 + 8 * lane)) = lv;
                }
            }
            }
            // Halo pass: each slot (quarter, row, halo column) loads its eight
            // planes from the bordered position.
            for (int hq = tid; hq < 4 * TRp * nh; hq += 256) {
                const int k = hq % nh, rest = hq / nh;
                const int ty = rest % TRp, grp = rest / TRp;
                const int cx = k < nhl ? k : TCp - nhr + (k - nhl);
                int x = x0 + cx, y = y0 + ty;
                if (g.border == COGCONV_BORDER_CYCLIC) {
                    x %= g.W;
                    x += x < 0 ? g.W : 0;
                    y %= g.H;
                    y += y < 0 ? g.H : 0;
                } else {
                    x = min(max(x, 0), g.W - 1);
                    y = min(max(y, 0), g.H - 1);
                }
                int vo = 4 * ((r0 + 8 * grp) * HW + y * g.W + x);
                float hvals[8];
#pragma unroll
                for (int e8 = 0; e8 < 8; e8++) {
                    hvals[e8] = buffer_load_f32_nt(xr, vo, 0);
                    vo += planeBytes;
                }
                V hv, lv;
#pragma unroll
                for (int e8 = 0; e8 < 8; e8++) {
                    const T h = (T)hvals[e8];
                    hv[e8] = h;
                    if (SPLIT) lv[e8] = (T)(hvals[e8] - (float)h);
                }
                const size_t o = ((size_t)grp * npix + ty * TCp + cx) * 8;
                *reinterpret_cast<V*>(sB + o) = hv;
                if (SPLIT) *reinterpret_cast<V*>(sB + (size_t)4 * npix * 8 + o) = lv;
            }
        } else {
            for (int base = wave; base < nitems; base += 4 * IU) {
                float v[IU][8];
                issue(r0, base, v);
                __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first store
                commit(base, v);
            }
        }
        for (int tr0 = 0; tr0 < F; tr0 += TRG) {
            if (RING) {
                // Tap row tr0 (prefetched) into buffer tr0 & 1; after the barrier,
                // every wave is past the MFMAs of row tr0 - 2, the buffer's last user.
#pragma unroll
                for (int i = 0; i < (RING ? WPG : 1); i++) *reinterpret_cast<V*>(sB + rDst[tr0 & 1][i]) = rv[i];
                __syncthreads();
                if (tr0 + 1 < F) {
#pragma unroll
                    for (int i = 0; i < (RING ? WPG : 1); i++)
                        rv[i] = __builtin_bit_cast(
                            V, __builtin_amdgcn_raw_buffer_load_b128(wr, rOff[i] + (tr0 + 1) * rowBytesW, wChunk, 0));
                }
            } else if (WPRE && onePass) {
#pragma unroll
                for (int i = 0; i < (WPRE ? WPT : 1); i++) *reinterpret_cast<V*>(sB + wDst[i]) = wv[i];
                __syncthreads();
            } else {
                // Filter slice of tap rows tr0 .. tr0 + TRG - 1 from the pre-packed bank.
                if (tr0 > 0) __syncthreads();  // the previous group's MFMAs are done with the slice
                const int nUnits = TERMS * TAPG;  // taps of the group over both terms; TPI a step
                constexpr int WU = 8;
                for (int i0 = 0; i0 * TPI < nUnits; i0 += WU) {
                    V wb[WU];
                    int dst[WU];
#pragma unroll
                    for (int i = 0; i < WU; i++) {
                        int off;
                        group_unit(i0 + i, tr0 * F, TAPG, sAoff, off, dst[i]);
                        wb[i] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(wr, off, wChunk, 0));
                    }
#pragma unroll
                    for (int i = 0; i < WU; i++) *reinterpret_cast<V*>(sB + dst[i]) = wb[i];
                }
                __syncthreads();
            }
            // Taps of this group: sAg addresses tap t = ti * F + tj of the whole
            // filter, so the per-tap offsets stay compile-time constants.
            const T* sAg = sA + (RING ? (size_t)(tr0 & 1) * gElems : 0) - (size_t)tr0 * F * 4 * MT * 8;
            const T* sAg2 = sAg + (size_t)TAPG * 4 * MT * 8;  // split: the residual term
            const int tEnd = min(F, tr0 + TRG);
            if constexpr (!SPLIT) {
                // The group's rows one at a time (a runtime loop: the row is
                // uniform and the tap offsets within a row are constants, so the
                // LDS reads take immediate offsets from one base per row).
                for (int ti = tr0; ti < tEnd; ti++) {
                    const T* aRow = sAg + ((size_t)(ti * F * 4 + g4) * MT + l16) * 8;
                    const T* bRow = sB + (size_t)ti * TCp * 8;
#pragma unroll
                    for (int tj = 0; tj < F; tj++) {
                        V av[MB], bv[4];
#pragma unroll
                        for (int mb = 0; mb < MB; mb++)
                            av[mb] = *reinterpret_cast<const V*>(aRow + ((size_t)tj * 4 * MT + mb * 16) * 8);
#pragma unroll
                        for (int pb = 0; pb < 4; pb++) bv[pb] = *reinterpret_cast<const V*>(bRow + boff[pb] + tj * 8);
#pragma unroll
                        for (int mb = 0; mb < MB; mb++)
#pragma unroll
                            for (int pb = 0; pb < 4; pb++) acc[mb][pb] = lp_mfma<PREC>(bv[pb], av[mb], acc[mb][pb]);
                    }
                }
                continue;
            }
#pragma unroll
            for (int ti = 0; ti < F; ti++) {
                if (ti < tr0 || ti >= tEnd) continue;  // block-uniform
#pragma unroll
                for (int tj = 0; tj < F; tj++) {
                    const int t = ti * F + tj;
                    const int toff = (ti * TCp + tj) * 8;
                    V av[MB], bv[4];
#pragma unroll
                    for (int mb = 0; mb < MB; mb++)
                        av[mb] = *reinterpret_cast<const V*>(sAg + ((size_t)(t * 4 + g4) * MT + mb * 16 + l16) * 8);
#pragma unroll
                    for (int pb = 0; pb < 4; pb++) bv[pb] = *reinterpret_cast<const V*>(sB + boff[pb] + toff);
                    if (SPLIT) {
                        V al[MB], bl[4];
#pragma unroll
                        for (int mb = 0; mb < MB; mb++)
                            al[mb] = *reinterpret_cast<const V*>(sAg2 + ((size_t)(t * 4 + g4) * MT + mb * 16 + l16) * 8);
#pragma unroll
                        for (int pb = 0; pb < 4; pb++)
                            bl[pb] = *reinterpret_cast<const V*>(sB + (size_t)4 * npix * 8 + boff[pb] + toff);
#pragma unroll
                        for (int mb = 0; mb < MB; mb++)
#pragma unroll
                            for (int pb = 0; pb < 4; pb++) {
                                acc[mb][pb] = lp_mfma<PREC>(bv[pb], al[mb], acc[mb][pb]);
                                acc[mb][pb] = lp_mfma<PREC>(bl[pb], av[mb], acc[mb][pb]);
                            }
                    }
#pragma unroll
                    for (int mb = 0; mb < MB; mb++)
#pragma unroll
                        for (int pb = 0; pb < 4; pb++) acc[mb][pb] = lp_mfma<PREC>(bv[pb], av[mb], acc[mb][pb]);
                }
            }
        }
    }
    // Epilogue: the MFMAs run pixels x planes (A = the staged pixels, B = the
    // filter slice), so lane (g4, l16) of block (mb, pb) holds output plane
    // m0 + 16 mb + l16 at the four adjacent pixels 4 g4 .. 4 g4 + 3 of its
    // 16-pixel run: one 16-byte store each where output rows are 16-byte
    // aligned (a 4-column group is then wholly inside or past the row).
    const bool vec = (g.OW & 3) == 0 && (reinterpret_cast<uintptr_t>(O) & 15) == 0;
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
        const int mm = m0 + mb * 16 + l16;
#pragma unroll
        for (int pb = 0; pb < 4; pb++) {
            const int orow = tr * 8 + 2 * wave + (pb >> 1);
            const int ocol = tc * 32 + (pb & 1) * 16 + 4 * g4;
            if (mm >= Mout || orow >= g.OH || ocol >= g.OW) continue;
            float* dst = O + ((size_t)(b * Mout + mm) * g.OH + orow) * g.OW + ocol;
            if (vec) {
                store_out4(dst, acc[mb][pb][0], acc[mb][pb][1], acc[mb][pb][2], acc[mb][pb][3], e);
            } else {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    if (ocol + v < g.OW) dst[v] = acc[mb][pb][v];
            }
        }
    }
    // The epilogue rewrites the thread's own stores (still in L2), so the
    // accumulators above stay in registers whatever its code size.
    if (e.n) {
        for (int t = 0; t < MB * 4; t++) {
            const int mb = t >> 2, pb = t & 3;
            const int mm = m0 + mb * 16 + l16;
            const int orow = tr * 8 + 2 * wave + (pb >> 1);
            const int ocol = tc * 32 + (pb & 1) * 16 + 4 * g4;
            if (mm >= Mout || orow >= g.OH || ocol >= g.OW) continue;
            float* dst = O + ((size_t)(b * Mout + mm) * g.OH + orow) * g.OW + ocol;
            const int pl = b * Mout + mm, pt = orow * g.OW + ocol;
            if (vec) {
                epi_rewrite4(e, dst, pl, pt);
            } else {
                for (int v = 0; v < 4 && ocol + v < g.OW; v++) epi_rewrite1(e, dst + v, pl, pt + v);
            }
        }
    }
}

template <int F, int MB, int PREC, bool RING>
__global__ __launch_bounds__(256) void k_igemm_lp(Geom g, PlaneMap m, IgemmCfg c, const float* __restrict__ X,
                                                  const typename LpType<PREC>::T* __restrict__ Wp,
                                                  float* __restrict__ O, const cogconv_image_io* __restrict__ tab,
                                                  Epi e) {
    igemm_lp_body<F, MB, PREC, RING>(g, m, c, X, Wp, O, tab, e);
}

// Ring mode with two blocks a CU (LDS within half the CU's): registers held to
// half the SIMD's file so that both blocks' waves are resident.
template <int F, int MB, int PREC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_igemm_lp_ring2(
    Geom g, PlaneMap m, IgemmCfg c, const float* __restrict__ X, const typename LpType<PREC>::T* __restrict__ Wp,
    float* __restrict__ O, const cogconv_image_io* __restrict__ tab, Epi e) {
    igemm_lp_body<F, MB, PREC, true>(g, m, c, X, Wp, O, tab, e);
}

// A ring configuration whose LDS leaves room for a second block on the CU.
bool igemm_lp_two_blocks(const IgemmCfg& c) { return c.ring && c.lds_bytes <= 80 * 1024; }

size_t igemm_lp_lds_bytes(const IgemmCfg& c, int F, int precision) {
    const int terms = precision == COGCONV_SPLIT_BF16 ? 2 : 1;
    // tile [4][npix][8] + filter group [trg * F][4][16 MB][8] (two one-row
    // groups in ring mode), 16-bit elements, + scratch slots of the lanes
    // beyond the tile width
    const size_t group = (size_t)(c.ring ? 1 : c.trg) * F * 16 * c.MB;
    const size_t quarter = ((size_t)c.tileRows * c.tileCols + 15) & ~(size_t)15;  // padded (see the kernel)
    return (size_t)terms * (quarter + (c.ring ? 2 : 1) * group) * 64 +
           (size_t)terms * 64 * 8 * 2;
}

size_t igemm_lp_workspace_bytes(const IgemmCfg& c, int F, int precision) {
    const int terms = precision == COGCONV_SPLIT_BF16 ? 2 : 1;
    return (size_t)terms * F * F * (size_t)c.mTiles * 16 * c.MB * c.RC * 2;
}

// stage: LP_PACK_AND_RUN (both kernels), LP_PACK_ONLY (the filter pre-pack
// into ws), LP_RUN_ONLY (the convolution on a bank packed earlier into ws).
template <int F, int MB, int PREC>
static hipError_t launch_lp3(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x, const float* w,
                             float* out, void* ws, hipStream_t s, int stage, const Epi& ep,
                             const cogconv_image_io* tab) {
    typedef typename LpType<PREC>::T T;
    T* wp = static_cast<T*>(ws);
    if (stage != LP_RUN_ONLY) {
        const long packed = (long)F * F * c.mTiles * 16 * MB * c.RC;
        const long pblocks = std::min((packed + 255) / 256, 4096L);
        hipLaunchKernelGGL((k_pack_filters_lp<PREC>), dim3((unsigned)pblocks), dim3(256), 0, s, g, m, c.Mout,
                           c.mTiles * 16 * MB, c.RC, w, wp);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || stage == LP_PACK_ONLY) return e;
    }
    const long blocks = (long)c.tilesC * c.tilesR * c.mTiles * m.B;
    // Two blocks a CU fit the LDS only for MB <= 2, or MB = 4 with F <= 5 (the
    // planner's sizes); split-BF16 tiles never do.
    constexpr bool TWO = PREC != COGCONV_SPLIT_BF16 && (MB <= 2 || F <= 5);
    if constexpr (TWO) {
        if (igemm_lp_two_blocks(c)) {
            hipLaunchKernelGGL((k_igemm_lp_ring2<F, MB, PREC>), dim3((unsigned)blocks), dim3(256), c.lds_bytes, s, g,
                               m, c, x, wp, out, tab, ep);
            return hipGetLastError();
        }
    }
    if constexpr (PREC != COGCONV_SPLIT_BF16) {  // the planner offers ring mode for BF16 / FP16 only
        if (c.ring) {
            hipLaunchKernelGGL((k_igemm_lp<F, MB, PREC, true>), dim3((unsigned)blocks), dim3(256), c.lds_bytes, s, g,
                               m, c, x, wp, out, tab, ep);
            return hipGetLastError();
        }
    } else if (c.ring) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL((k_igemm_lp<F, MB, PREC, false>), dim3((unsigned)blocks), dim3(256), c.lds_bytes, s, g, m, c,
                       x, wp, out, tab, ep);
    return hipGetLastError();
}

template <int F, int PREC>
static hipError_t launch_lp2(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x, const float* w,
                             float* out, void* ws, hipStream_t s, int stage, const Epi& ep,
                             const cogconv_image_io* tab) {
    switch (c.MB) {
    case 1: return launch_lp3<F, 1, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case 2: return launch_lp3<F, 2, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    default: return launch_lp3<F, 4, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    }
}

template <int PREC>
static hipError_t launch_lp1(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x, const float* w,
                             float* out, void* ws, hipStream_t s, int stage, const Epi& ep,
                             const cogconv_image_io* tab) {
    switch (g.FH) {
    case 1: return launch_lp2<1, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case 3: return launch_lp2<3, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case 5: return launch_lp2<5, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case 7: return launch_lp2<7, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case 9: return launch_lp2<9, PREC>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pack_filters_lp(int precision, const Geom& g, const PlaneMap& m, int Mout, int Mpad, int Rpad,
                                  const float* w, void* wp, hipStream_t s) {
    const long packed = (long)g.FH * g.FW * Mpad * Rpad;
    const long pblocks = std::min((packed + 255) / 256, 4096L);
    if (pblocks <= 0) return hipErrorInvalidConfiguration;
    switch (precision) {
    case COGCONV_BF16:
        hipLaunchKernelGGL((k_pack_filters_lp<COGCONV_BF16>), dim3((unsigned)pblocks), dim3(256), 0, s, g, m, Mout, Mpad,
                           Rpad, w, static_cast<LpType<COGCONV_BF16>::T*>(wp));
        break;
    case COGCONV_FP16:
        hipLaunchKernelGGL((k_pack_filters_lp<COGCONV_FP16>), dim3((unsigned)pblocks), dim3(256), 0, s, g, m, Mout, Mpad,
                           Rpad, w, static_cast<LpType<COGCONV_FP16>::T*>(wp));
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool igemm_lp_supported(int F) { return F == 1 || F == 3 || F == 5 || F == 7 || F == 9; }

hipError_t launch_igemm_lp(int precision, const Geom& g, const PlaneMap& m, const IgemmCfg& cfg, const float* x,
                           const float* w, float* out, void* ws, hipStream_t s, int stage, const Epi& ep,
                           const cogconv_image_io* tab, bool x16) {
    const long blocks = (long)cfg.tilesC * cfg.tilesR * cfg.mTiles * m.B;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    IgemmCfg c = cfg;
    if (tab ? !x16 : (reinterpret_cast<uintptr_t>(x) & 15) != 0) c.vec4 = 0;  // 16-byte loads need 16-byte rows
    switch (precision) {
    case COGCONV_BF16: return launch_lp1<COGCONV_BF16>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case COGCONV_FP16: return launch_lp1<COGCONV_FP16>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    case COGCONV_SPLIT_BF16: return launch_lp1<COGCONV_SPLIT_BF16>(g, m, c, x, w, out, ws, s, stage, ep, tab);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace cogconv
