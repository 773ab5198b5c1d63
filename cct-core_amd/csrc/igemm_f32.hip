// FP32 implicit-GEMM kernel on MFMA for ProjectFrameBlockReduceSum and
// BackProjectFrameBlockReduceSum (the contraction modes of the conv path).
//
//   out[b*Mout + m](y, x) = sum_{r < R} sum_{t < F*F} A[m][r][t] * X_{b*N + r}(y*dn - h + ti, x*dn - h + tj)
//
// ProjectFrameBlockReduceSum:     A[m][r][t] = W[m*N + r][t]   (ConvolveHyperKernel.scala:153-167, 327-328)
// BackProjectFrameBlockReduceSum: A[m][r][t] = W[r*L + m][t]   (ConvolveHyperKernel.scala:330-332)
// with the filter flipped for the convolution orientation (a flip of both axes
// of a row-major F x F filter reverses its flat tap index).  Borders are
// applied while the input tile is staged in LDS (ConvolveHyperKernel.scala:235-317).
//
// gfx950 mapping: v_mfma_f32_16x16x4_f32.  A block of 512 threads (8 wave64)
// owns an 8 x 32 pixel tile of one image and 16*MB output planes.  Waves w and
// w+4 own pixel rows 2(w%4), 2(w%4)+1 (four 16-pixel MFMA column blocks);
// wave w < 4 reduces over the first half of each staged plane chunk, wave
// w >= 4 over the second half, and the two partial sums are added through LDS
// at the end.  The four k-lanes of one MFMA (lane/16) carry four consecutive
// input planes at the same tap; the MFMA takes the pixels as A and the filters
// as B, so a lane's accumulator holds four adjacent pixels of one output plane.  Input planes are staged RC at a time: the
// tile [RC][rows][cols] (one row per wave pass, lanes across columns) and the
// filter slice [RC][F*F][16*MB] (each thread copies the F*F contiguous taps of
// one (plane, filter) pair).  Both staged plane strides are 16 mod 32 words so
// that the four k-lane groups of an MFMA read different LDS banks.
#include "epi.h"
#include "stage.h"

namespace cogconv {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Stages input planes r0 .. r0 + RC - 1 of a tile of TRp rows from stored row
// y0 (border applied) into sB[plane][row][col] (plane stride planeStride, row
// stride TCp): one (plane, row) item per wave pass.  The row of an item is the
// same for all lanes of a wave (scalar work); the lanes' column byte offsets
// vc0 / vc1 (STAGE_OOB: no column or off the field) were computed once per
// block.  Per batch of IU items: offsets and masks first, then all loads
// (unconditional, from in-bounds offsets, no branch between them), then the
// masked LDS stores; items past the end and lanes past the tile width store
// into a scratch slot instead of branching.
template <int IU>
__device__ __forceinline__ void stage_input_f32(const Geom& g, const PlaneMap& m, __amdgpu_buffer_rsrc_t xr, int r0,
                                                int RC, int TRp, int TCp, int planeStride, int y0, int HW, int vc0,
                                                int vc1, bool twoPass, float* sB, float* sJunk, int wave, int lane) {
    const int nrows = RC * TRp;
    for (int base = wave; base < nrows; base += 8 * IU) {
        // Per-lane row table of the batch (lane u = the wave's u-th item):
        // byte offset (border applied), LDS destination (-1: not live) and
        // in-field mask, read back per row with readlane / a ballot bit.
        int rSoff, rDst;
        uint64_t rMask;
        {
            const int item = base + 8 * lane;
            const int rc = item / TRp, ty = item - rc * TRp;
            const int r = r0 + rc;
            int y = y0 + ty;
            bool in = (unsigned)y < (unsigned)g.H;
            if (g.border == COGCONV_BORDER_CYCLIC) {
                y %= g.H;
                y += y < 0 ? g.H : 0;
                in = true;
            } else {
                in = in || g.border == COGCONV_BORDER_CLAMP;
                y = min(max(y, 0), g.H - 1);
            }
            const bool live = lane < IU && item < nrows;
            rSoff = 4 * (min(r, m.R - 1) * HW + y * g.W);
            rDst = live ? rc * planeStride + ty * TCp : -1;
            rMask = __ballot(live && in && r < m.R);
        }
        auto xrow = [&](int u, int& soff, bool& on) {
            soff = __builtin_amdgcn_readlane(rSoff, u);
            on = (rMask >> u) & 1;
        };
        auto xdst = [&](int u) { return __builtin_amdgcn_readlane(rDst, u); };
        if (twoPass)
            stage_rows_fn<true, IU>(xr, xrow, xdst, sB, sJunk, lane, TCp, 1, vc0, vc1);
        else
            stage_rows_fn<false, IU>(xr, xrow, xdst, sB, sJunk, lane, TCp, 1, vc0, vc1);
    }
}


template <int FR, int FC, int MB, bool PH>
__global__ __launch_bounds__(512) void k_igemm_f32(Geom g, PlaneMap m, IgemmCfg c, Phase ph, const float* __restrict__ X,
                                                   const float* __restrict__ Wt, float* __restrict__ O,
                                                   const cogconv_image_io* __restrict__ tab, Epi e) {
    constexpr int TAPS = FR * FC;
    const int TAPSF = ph.FHf * ph.FWf;  // taps of a full filter in the bank
    constexpr int MT = 16 * MB;
    extern __shared__ float smem[];
    float* sB = smem;
    float* sA = smem + (size_t)c.RC * c.planeStride;
    const int aStride = c.aStride;  // floats per staged plane of the filter slice (TAPS*MT padded)
    float* sJunk = sA + (size_t)c.RC * aStride;  // 128 scratch slots for masked-off stores

    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    // Output-plane blocks of one tile are adjacent block indices (the same XCD
    // after the swizzle, and close in time), so the tile's second and later
    // blocks find its input in L2.
    const int mt = bid % c.mTiles;
    bid /= c.mTiles;
    const int tc = bid % c.tilesC;
    bid /= c.tilesC;
    const int tr = bid % c.tilesR;
    const int b = bid / c.tilesR;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rw = wave & 3, pg = wave >> 2;
    const int g4 = lane >> 4, l16 = lane & 15;
    const int Mout = c.Mout;
    const int m0 = mt * MT;
    const int y0 = tr * 8 * g.dnR - g.hR;
    const int x0 = tc * 32 * g.dnC - g.hC;
    const int TRp = c.tileRows, TCp = c.tileCols;
    const bool bpf = m.mode == COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM;
    const int L = m.T / m.N;
    const int HW = g.H * g.W;                       // 32-bit offsets: the plan checks the sizes
    if (tab) {
        // Request table: image b's planes and output planes are entry b's buffers.
        // X and O are re-based so that the image-b offsets below land on them.
        X = reinterpret_cast<const float*>(reinterpret_cast<uintptr_t>(tab[b].in0) - (uintptr_t)b * m.N * HW * 4);
        O = reinterpret_cast<float*>(reinterpret_cast<uintptr_t>(tab[b].out) -
                                     (uintptr_t)b * Mout * ph.OHf * ph.OWf * 4);
    }
    const float* Xb = X + (size_t)b * m.N * HW;
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(Xb, (uint32_t)m.N * HW * 4u);  // the plan bounds this by 2^31
    // Per-lane column offsets of the staged tile (two 64-lane passes at most)
    // with the border applied: the same for every staged row and plane.
    const bool twoPass = TCp > 64;
    const int vc0 = stage_col_offset(g, x0 + lane), vc1 = stage_col_offset(g, x0 + lane + 64);
    const int half = c.RC >> 1;

    floatx4 acc[MB][4];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    int boff[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
        const int row = 2 * rw + (pb >> 1);
        const int col = (pb & 1) * 16 + l16;
        boff[pb] = (row * g.dnR) * TCp + col * g.dnC;
    }

    for (int r0 = 0; r0 < m.R; r0 += c.RC) {
        __syncthreads();
        // Rows a wave stages per batch: 20 (one batch for 16 planes of a
        // 3x3 tile's 10 rows) where the registers allow two blocks per CU;
        // 12 for four row blocks and for the up-sampled phase kernels (2x
        // 7x7 Zero: 700 us at 12, 727 us at 20).
        constexpr int IU = (MB == 4 || PH) ? 12 : 20;
        stage_input_f32<IU>(g, m, xr, r0, c.RC, TRp, TCp, c.planeStride, y0, HW, vc0, vc1, twoPass, sB, sJunk, wave,
                            lane);
        // Filter slice: thread -> (plane rc, filter mm); the F*F taps are
        // contiguous in W.  Loads are unconditional from a clamped pair.
        for (int q = tid; q < c.RC * MT; q += 512) {
            const int mm = q % MT, rc = q / MT;
            const int r = r0 + rc, mg = m0 + mm;
            const bool ok = r < m.R && mg < Mout;
            const int rr = min(r, m.R - 1), mgc = min(mg, Mout - 1);
            const float* src = Wt + (size_t)(bpf ? rr * L + mgc : mgc * m.N + rr) * TAPSF;
            // Tap (a, b) of this launch's (sub-)filter is tap (i0 + step*a,
            // j0 + step*b) of the full filter, flipped for the convolution
            // orientation; contiguous when there is one phase.
            float t[TAPS];
            float* dst = sA + (size_t)rc * aStride + mm;
            if (!PH || ph.step == 1) {  // one phase (PH = false: always), the taps are contiguous
#pragma unroll
                for (int i = 0; i < TAPS; i++) t[i] = src[i];
#pragma unroll
                for (int i = 0; i < TAPS; i++) dst[i * MT] = ok ? (ph.flip ? t[TAPS - 1 - i] : t[i]) : 0.f;
            } else {
#pragma unroll
                for (int i = 0; i < TAPS; i++) {
                    const int a = i / FC, bcol = i - (i / FC) * FC;
                    const int fi = (ph.i0 + ph.step * a) * ph.FWf + ph.j0 + ph.step * bcol;
                    t[i] = src[ph.flip ? TAPSF - 1 - fi : fi];
                }
#pragma unroll
                for (int i = 0; i < TAPS; i++) dst[i * MT] = ok ? t[i] : 0.f;
            }
        }
        __syncthreads();
        for (int p = pg * half; p < (pg + 1) * half; p += 4) {
            const float* bb = sB + (p + g4) * c.planeStride;
            const float* ab = sA + (size_t)(p + g4) * aStride + l16;
#pragma unroll
            for (int ti = 0; ti < FR; ti++) {
#pragma unroll
                for (int tj = 0; tj < FC; tj++) {
                    const int t = ti * FC + tj;
                    float av[MB];
#pragma unroll
                    for (int mb = 0; mb < MB; mb++) av[mb] = ab[t * MT + mb * 16];
                    float bv[4];
#pragma unroll
                    for (int pb = 0; pb < 4; pb++) bv[pb] = bb[boff[pb] + ti * TCp + tj];
#pragma unroll
                    for (int mb = 0; mb < MB; mb++)
#pragma unroll
                        for (int pb = 0; pb < 4; pb++)
                            acc[mb][pb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[pb], av[mb], acc[mb][pb], 0, 0, 0);
                }
            }
        }
    }
    // Epilogue.  The MFMAs run pixels x planes (A = the staged pixels, B = the
    // filter slice), so lane (g4, l16) of block (mb, pb) holds output plane
    // m0 + 16 mb + l16 at the four adjacent pixels 4 g4 .. 4 g4 + 3 of its
    // 16-pixel run.  Dense outputs whose rows start 16-byte aligned: plane
    // group 1 parks its sums in LDS (one 16-byte vector per block and lane),
    // plane group 0 adds them and stores each lane's four pixels as one
    // 16-byte store (a 4-column group is then wholly inside or past the row).
    // Other outputs: the sums meet in an LDS tile [16*MB planes][8 rows][32
    // cols] (plane pitch 260 words, 16-byte writes), written row by row with
    // 8-byte or single-float stores.
    __syncthreads();
    const bool vec = ph.oStride == 1 && ((ph.OWf | ph.oCol0) & 3) == 0 && (reinterpret_cast<uintptr_t>(O) & 15) == 0;
    if (vec) {  // block-uniform
        floatx4* red4 = reinterpret_cast<floatx4*>(smem);  // [rw][MB][4][64]
        if (pg == 1) {
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int pb = 0; pb < 4; pb++) red4[((rw * MB + mb) * 4 + pb) * 64 + lane] = acc[mb][pb];
        }
        __syncthreads();
        if (pg == 0) {
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                const int mm = m0 + mb * 16 + l16;
#pragma unroll
                for (int pb = 0; pb < 4; pb++) {
                    const floatx4 sum = acc[mb][pb] + red4[((rw * MB + mb) * 4 + pb) * 64 + lane];
                    const int orow = tr * 8 + 2 * rw + (pb >> 1);
                    const int ocol = tc * 32 + (pb & 1) * 16 + 4 * g4;
                    if (mm < Mout && orow < g.OH && ocol < g.OW)
                        store_out4(O + ((size_t)(b * Mout + mm) * ph.OHf + ph.oRow0 + orow) * ph.OWf + ph.oCol0 + ocol,
                                   sum[0], sum[1], sum[2], sum[3], e);
                }
            }
            // The epilogue rewrites the thread's own stores (still in L2), so the
            // accumulators above stay in registers whatever its code size.
            if (e.n) {
                for (int t = 0; t < MB * 4; t++) {
                    const int mb = t >> 2, pb = t & 3;
                    const int mm = m0 + mb * 16 + l16;
                    const int orow = tr * 8 + 2 * rw + (pb >> 1);
                    const int ocol = tc * 32 + (pb & 1) * 16 + 4 * g4;
                    if (mm < Mout && orow < g.OH && ocol < g.OW)
                        epi_rewrite4(e, O + ((size_t)(b * Mout + mm) * ph.OHf + ph.oRow0 + orow) * ph.OWf + ph.oCol0 +
                                            ocol,
                                     b * Mout + mm, (ph.oRow0 + orow) * ph.OWf + ph.oCol0 + ocol);
                }
            }
        }
        return;
    }
    constexpr int OP = 260;
    float* sO = smem;
    auto tile = [&](int mb, int pb) {
        return reinterpret_cast<floatx4*>(sO + (mb * 16 + l16) * OP + (2 * rw + (pb >> 1)) * 32 + (pb & 1) * 16 + 4 * g4);
    };
    if (pg == 1) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int pb = 0; pb < 4; pb++) *tile(mb, pb) = acc[mb][pb];
    }
    __syncthreads();
    if (pg == 0) {
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int pb = 0; pb < 4; pb++) *tile(mb, pb) = acc[mb][pb] + *tile(mb, pb);
    }
    __syncthreads();
    if (ph.oStride == 1 && ((ph.OWf | ph.oCol0) & 1) == 0 && (reinterpret_cast<uintptr_t>(O) & 7) == 0) {
        const int c2 = 2 * (tid & 15);  // 8-byte stores, 16 lanes per row
        const int ocol = tc * 32 + c2;
#pragma unroll 4
        for (int item = tid >> 4; item < MT * 8; item += 32) {
            const int mm = item >> 3, row = item & 7;
            const int orow = tr * 8 + row;
            if (m0 + mm < Mout && orow < g.OH && ocol < g.OW) {
                const float2 v = *reinterpret_cast<const float2*>(sO + mm * OP + row * 32 + c2);
                const int pt = (ph.oRow0 + orow) * ph.OWf + ph.oCol0 + ocol, pl = b * Mout + m0 + mm;
                *reinterpret_cast<float2*>(O + ((size_t)(b * Mout + m0 + mm) * ph.OHf + ph.oRow0 + orow) * ph.OWf +
                                           ph.oCol0 + ocol) = make_float2(epi(e, v.x, pl, pt), epi(e, v.y, pl, pt + 1));
            }
        }
    } else {
        const int col = tid & 31;
        const int ocol = tc * 32 + col;
        for (int item = tid >> 5; item < MT * 8; item += 16) {
            const int mm = item >> 3, row = item & 7;
            const int orow = tr * 8 + row;
            if (m0 + mm < Mout && orow < g.OH && ocol < g.OW)
                O[((size_t)(b * Mout + m0 + mm) * ph.OHf + ph.oRow0 + ph.oStride * orow) * ph.OWf + ph.oCol0 +
                  ph.oStride * ocol] = epi(e, sO[mm * OP + row * 32 + col], b * Mout + m0 + mm,
                                           (ph.oRow0 + ph.oStride * orow) * ph.OWf + ph.oCol0 + ph.oStride * ocol);
        }
    }
}

// PH = true instantiations gather the taps of a phase sub-filter; the others
// (one phase) copy contiguous taps and carry no gather code, whose per-tap
// index arithmetic would otherwise be hoisted into (and spill) scalar
// registers in every launch.
template <int FR, int FC, bool PH>
static hipError_t launch_rc(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const Phase& ph, const float* x,
                            const float* w, float* out, hipStream_t s, dim3 grid, const cogconv_image_io* tab,
                            const Epi& e) {
    switch (c.MB) {
    case 1: hipLaunchKernelGGL((k_igemm_f32<FR, FC, 1, PH>), grid, dim3(512), c.lds_bytes, s, g, m, c, ph, x, w, out, tab, e); break;
    case 2: hipLaunchKernelGGL((k_igemm_f32<FR, FC, 2, PH>), grid, dim3(512), c.lds_bytes, s, g, m, c, ph, x, w, out, tab, e); break;
    default: hipLaunchKernelGGL((k_igemm_f32<FR, FC, 4, PH>), grid, dim3(512), c.lds_bytes, s, g, m, c, ph, x, w, out, tab, e); break;
    }
    return hipGetLastError();
}

// Square filters, and the rectangular phase filters of 2x up-sampling (odd F
// up to 9 splits into ceil(F/2) and floor(F/2) taps per dimension).
bool igemm_f32_supported_rc(int FR, int FC) {
    if (FR == FC) return FR == 1 || FR == 3 || FR == 5 || FR == 7 || FR == 9 || FR == 2 || FR == 4;
    const int lo = FR < FC ? FR : FC, hi = FR < FC ? FC : FR;
    return hi == lo + 1 && lo >= 1 && hi <= 5;
}

bool igemm_f32_supported(int F) { return igemm_f32_supported_rc(F, F); }

// LDS floats the output tile of k_igemm_f32's epilogue needs.
size_t igemm_f32_exchange_floats(int MB) { return (size_t)16 * MB * 260; }  // >= the [4 waves][MB][4][64] float4 exchange

hipError_t launch_igemm_f32_phase(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const Phase& ph, const float* x,
                                  const float* w, float* out, hipStream_t s, const Epi& e,
                                  const cogconv_image_io* tab) {
    const long blocks = (long)c.tilesC * c.tilesR * c.mTiles * m.B;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks);
    const int key = g.FH * 10 + g.FW;
    if (ph.step == 1) {
        switch (key) {
        case 11: return launch_rc<1, 1, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 33: return launch_rc<3, 3, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 55: return launch_rc<5, 5, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 77: return launch_rc<7, 7, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 99: return launch_rc<9, 9, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 22: return launch_rc<2, 2, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        case 44: return launch_rc<4, 4, false>(g, m, c, ph, x, w, out, s, grid, tab, e);
        default: return hipErrorInvalidValue;
        }
    }
    switch (key) {
    case 11: return launch_rc<1, 1, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 22: return launch_rc<2, 2, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 33: return launch_rc<3, 3, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 44: return launch_rc<4, 4, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 55: return launch_rc<5, 5, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 12: return launch_rc<1, 2, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 21: return launch_rc<2, 1, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 23: return launch_rc<2, 3, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 32: return launch_rc<3, 2, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 34: return launch_rc<3, 4, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 43: return launch_rc<4, 3, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 45: return launch_rc<4, 5, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    case 54: return launch_rc<5, 4, true>(g, m, c, ph, x, w, out, s, grid, tab, e);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_igemm_f32(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x, const float* w,
                            float* out, hipStream_t s, const Epi& e, const cogconv_image_io* tab) {
    Phase ph;
    ph.i0 = 0;
    ph.j0 = 0;
    ph.step = 1;
    ph.FHf = g.FH;
    ph.FWf = g.FW;
    ph.flip = g.flip;
    ph.oRow0 = 0;
    ph.oCol0 = 0;
    ph.oStride = 1;
    ph.OHf = g.OH;
    ph.OWf = g.OW;
    return launch_igemm_f32_phase(g, m, c, ph, x, w, out, s, e, tab);
}

// ---------------------------------------------------------------------------
// 2x up-sampled ProjectFrame / BackProjectFrame with block reduction, all four
// output phases in one launch.
//
// The zero-inserted input makes output pixel (2R + pr, 2C + pc) a dense
// convolution of the stored samples with the sub-filter of taps
// (i0(pr) + 2a, j0(pc) + 2b), i0(p) = (h - p) mod 2 for a low halo h
// (ConvolveHyperKernel.scala:235-317 reads the same zero-inserted input).  A
// block stages one window of stored samples, TRp x TCp = (8 + UR - 1) x
// (32 + UC - 1) for a tile of 8 x 32 phase pixels, once per plane chunk,
// and every phase reads its taps from that window: no phase re-stages the
// input, and the whole filter is staged once, its taps ordered phase by phase
// (phase (pr, pc) at tap base T(pr, pc), row-major within).
//
// Waves: w % 4 owns phase-pixel rows 2(w%4), 2(w%4)+1 (as in k_igemm_f32); w / 4
// is the row phase pr, and a wave sums every staged plane for both column
// phases, so a lane holds output columns 2C and 2C + 1 of its pixels and its
// stores are 8 adjacent output columns (two 16-byte stores per row) without
// an exchange through LDS.
template <int F, int HP, int PR, int MB, bool TWO>
__device__ __forceinline__ void up2q_chunk(const float* sB, const float* sA, int RC, int planeStride, int aStride,
                                           int TCp, int dr, int dc0, int dc1, const int (&boff)[4], int g4, int l16,
                                           floatx4 (&acc)[2][MB][4]) {
    constexpr int MT = 16 * MB;
    constexpr int FRp = up2_taps(F, HP, PR), FC0 = up2_taps(F, HP, 0), FC1 = up2_taps(F, HP, 1);
    constexpr int T0 = PR ? up2_taps(F, HP, 0) * F : 0;  // taps of row phase 0 (both column phases)
    constexpr int T1 = T0 + FRp * FC0;
    for (int p = 0; p < RC; p += 4) {
        const float* bb = sB + (p + g4) * planeStride + dr * TCp;
        const float* ab = sA + (size_t)(p + g4) * aStride + l16;
        // Two blocks a CU: one sub-filter row at a time, so that the LDS loads
        // the compiler hoists stay within 128 registers.
        constexpr int UNROLL = TWO ? 1 : FRp;
#pragma unroll UNROLL
        for (int a = 0; a < FRp; a++) {
#pragma unroll
            for (int pc = 0; pc < 2; pc++) {
                const int FCp = pc ? FC1 : FC0;
                const int tb = (pc ? T1 : T0) + a * FCp;
                const float* bq = bb + a * TCp + (pc ? dc1 : dc0);
#pragma unroll
                for (int bc = 0; bc < (pc ? FC1 : FC0); bc++) {
                    float av[MB];
#pragma unroll
                    for (int mb = 0; mb < MB; mb++) av[mb] = ab[(tb + bc) * MT + mb * 16];
                    float bv[4];
#pragma unroll
                    for (int pb = 0; pb < 4; pb++) bv[pb] = bq[boff[pb] + bc];
#pragma unroll
                    for (int mb = 0; mb < MB; mb++)
#pragma unroll
                        for (int pb = 0; pb < 4; pb++)
                            acc[pc][mb][pb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[pb], av[mb], acc[pc][mb][pb], 0, 0, 0);
                }
            }
        }
    }
}

template <int F, int HP, int MB, bool TWO>
__device__ __forceinline__ void up2q_body(const Geom& g, const PlaneMap& m, const UpFused& u, const float* __restrict__ X,
                                          const float* __restrict__ Wt, float* __restrict__ O, const Epi& e) {
    constexpr int TAPS = F * F;
    constexpr int MT = 16 * MB;
    const IgemmCfg& c = u.c;
    extern __shared__ float smem[];
    float* sB = smem;
    float* sA = smem + (size_t)c.RC * c.planeStride;
    const int aStride = c.aStride;
    float* sJunk = sA + (size_t)c.RC * aStride;

    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    // Output-plane blocks of one tile are adjacent block indices (the same XCD
    // after the swizzle, and close in time), so the tile's second and later
    // blocks find its input in L2.
    const int mt = bid % c.mTiles;
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
    const int TRp = c.tileRows, TCp = c.tileCols;
    const bool bpf = m.mode == COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM;
    const int L = m.T / m.N;
    const int HW = g.H * g.W;
    const float* Xb = X + (size_t)b * m.N * HW;
    const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(Xb, (uint32_t)m.N * HW * 4u);  // the plan bounds this by 2^31
    const int vc0 = stage_col_offset(g, x0 + lane);  // the window is at most 64 columns wide

    floatx4 acc[2][MB][4];
#pragma unroll
    for (int pc = 0; pc < 2; pc++)
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[pc][i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    int boff[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) boff[pb] = (2 * rw + (pb >> 1)) * TCp + (pb & 1) * 16 + l16;

    for (int r0 = 0; r0 < m.R; r0 += c.RC) {
        __syncthreads();
        stage_input_f32<12>(g, m, xr, r0, c.RC, TRp, TCp, c.planeStride, y0, HW, vc0, STAGE_OOB, false, sB, sJunk,
                            wave, lane);
        // The whole filter of each (plane, filter) pair, its taps reordered
        // phase by phase: tap (i, j) of the filter in the kernel's orientation
        // (flipped for the convolution orientation) belongs to row phase
        // HP ^ (i & 1) at sub-row i / 2 and column phase HP ^ (j & 1) at
        // sub-column j / 2.
        for (int q = tid; q < c.RC * MT; q += 512) {
            const int mm = q % MT, rc = q / MT;
            const int r = r0 + rc, mg = m0 + mm;
            const bool ok = r < m.R && mg < Mout;
            const int rr = min(r, m.R - 1), mgc = min(mg, Mout - 1);
            const float* src = Wt + (size_t)(bpf ? rr * L + mgc : mgc * m.N + rr) * TAPS;
            float* dst = sA + (size_t)rc * aStride + mm;
            // One tap row at a time: row ti of the kernel's orientation is
            // stored row ti (row F - 1 - ti reversed when flipped).
#pragma unroll
            for (int ti = 0; ti < F; ti++) {
                float t[F];
                const float* sr = src + (g.flip ? F - 1 - ti : ti) * F;
#pragma unroll
                for (int j = 0; j < F; j++) t[j] = sr[j];
                const int rp = HP ^ (ti & 1);
#pragma unroll
                for (int tj = 0; tj < F; tj++) {
                    const int cp = HP ^ (tj & 1);
                    const int pos = (rp ? up2_taps(F, HP, 0) * F : 0) +
                                    (cp ? up2_taps(F, HP, rp) * up2_taps(F, HP, 0) : 0) + (ti >> 1) * up2_taps(F, HP, cp) +
                                    (tj >> 1);
                    dst[pos * MT] = ok ? (g.flip ? t[F - 1 - tj] : t[tj]) : 0.f;
                }
            }
        }
        __syncthreads();
        if (pr == 0)
            up2q_chunk<F, HP, 0, MB, TWO>(sB, sA, c.RC, c.planeStride, aStride, TCp, u.dr[0], u.dc[0], u.dc[1], boff, g4, l16,
                                     acc);
        else
            up2q_chunk<F, HP, 1, MB, TWO>(sB, sA, c.RC, c.planeStride, aStride, TCp, u.dr[1], u.dc[0], u.dc[1], boff, g4, l16,
                                     acc);
    }
    // Epilogue: lane (g4, l16) of block (mb, pb) holds output plane
    // m0 + 16 mb + l16 at phase pixels C = 4 g4 .. 4 g4 + 3 of its 16-pixel run,
    // column phase 0 in acc[0] and 1 in acc[1]: output columns 2C .. 2C + 7 of
    // output row 2R + pr.  Two 16-byte stores where the row is 16-byte aligned
    // and the eight columns are inside it; single stores otherwise.
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
    // The epilogue rewrites the thread's own stores (still in L2), so the
    // accumulators above stay in registers whatever its code size.
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

// One block per CU, or (k_igemm_f32_up2q2, one output block) registers held to
// 128 a lane so that two blocks share a CU and one block's stores run under the
// other's MFMAs.
template <int F, int HP, int MB>
__global__ __launch_bounds__(512) void k_igemm_f32_up2q(Geom g, PlaneMap m, UpFused u, const float* __restrict__ X,
                                                        const float* __restrict__ Wt, float* __restrict__ O, Epi e) {
    up2q_body<F, HP, MB, false>(g, m, u, X, Wt, O, e);
}

template <int F, int HP, int MB>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_igemm_f32_up2q2(
    Geom g, PlaneMap m, UpFused u, const float* __restrict__ X, const float* __restrict__ Wt, float* __restrict__ O,
    Epi e) {
    up2q_body<F, HP, MB, true>(g, m, u, X, Wt, O, e);
}

bool igemm_f32_up2q_supported(int F) { return F == 3 || F == 5 || F == 7 || F == 9; }

template <int F, int HP>
static hipError_t launch_up2q_mb(const Geom& g, const PlaneMap& m, const UpFused& u, const float* x, const float* w,
                                 float* out, hipStream_t s, dim3 grid, const Epi& e) {
    if (u.two) {  // one output block only: two spill at 128 registers
        if (u.c.MB != 1) return hipErrorInvalidConfiguration;
        hipLaunchKernelGGL((k_igemm_f32_up2q2<F, HP, 1>), grid, dim3(512), u.c.lds_bytes, s, g, m, u, x, w, out, e);
    } else if (u.c.MB == 1) {
        hipLaunchKernelGGL((k_igemm_f32_up2q<F, HP, 1>), grid, dim3(512), u.c.lds_bytes, s, g, m, u, x, w, out, e);
    } else {
        hipLaunchKernelGGL((k_igemm_f32_up2q<F, HP, 2>), grid, dim3(512), u.c.lds_bytes, s, g, m, u, x, w, out, e);
    }
    return hipGetLastError();
}

hipError_t launch_igemm_f32_up2q(const Geom& g, const PlaneMap& m, const UpFused& u, const float* x, const float* w,
                                 float* out, hipStream_t s, const Epi& e) {
    const long blocks = (long)u.c.tilesC * u.c.tilesR * u.c.mTiles * m.B;
    if (blocks <= 0 || blocks > 0x7fffffffL || u.c.tileCols > 64 || (u.c.MB != 1 && u.c.MB != 2))
        return hipErrorInvalidConfiguration;
    const dim3 grid((unsigned)blocks);
    switch (g.FH * 2 + (u.hp & 1)) {
    case 6: return launch_up2q_mb<3, 0>(g, m, u, x, w, out, s, grid, e);
    case 7: return launch_up2q_mb<3, 1>(g, m, u, x, w, out, s, grid, e);
    case 10: return launch_up2q_mb<5, 0>(g, m, u, x, w, out, s, grid, e);
    case 11: return launch_up2q_mb<5, 1>(g, m, u, x, w, out, s, grid, e);
    case 14: return launch_up2q_mb<7, 0>(g, m, u, x, w, out, s, grid, e);
    case 15: return launch_up2q_mb<7, 1>(g, m, u, x, w, out, s, grid, e);
    case 18: return launch_up2q_mb<9, 0>(g, m, u, x, w, out, s, grid, e);
    case 19: return launch_up2q_mb<9, 1>(g, m, u, x, w, out, s, grid, e);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace cogconv
