// FP32 VALU kernels for every ConvolveOp: the LDS-tiled direct kernel, the
// split-reduction kernel for filter-sized outputs, and blockReduceSum.
//
// Semantics follow ConvolveHyperKernel.scala:205-317 (borders applied to input
// coordinates after the up-sampling phase test), 345-384 (flip for the
// convolution orientation), 404-548 (taps), and the FilterAdjoint handling of
// ConvolveToSmallFieldPipelinedTiledHyperKernel.scala:45-127 (up-sampling of
// the 2nd operand, output of filter size).  The kernels here are written for
// gfx950: 64-wide wavefronts, one wavefront per 64 output columns, the input
// tile and filter chunk staged in LDS, and wavefront shuffles for reductions.
#include "epi.h"
#include "internal.h"
#include "stage.h"

namespace cogconv {

// Operand-1 sample at materialised coordinates (y, x), with the up-sampling
// phase test and the border policy.  Always reads in bounds or returns 0.
__device__ __forceinline__ float read_operand(const float* __restrict__ a, const Geom& g, int y, int x) {
    if (g.upAr > 1) {
        if (y % g.upAr) return 0.0f;
        y /= g.upAr;
    }
    if (g.upAc > 1) {
        if (x % g.upAc) return 0.0f;
        x /= g.upAc;
    }
    switch (g.border) {
    case COGCONV_BORDER_CLAMP:
        y = min(max(y, 0), g.H - 1);
        x = min(max(x, 0), g.W - 1);
        break;
    case COGCONV_BORDER_CYCLIC:
        y %= g.H;
        if (y < 0) y += g.H;
        x %= g.W;
        if (x < 0) x += g.W;
        break;
    default:  // Zero, Full: zero outside.  Valid: only unwritten outputs look outside.
        if ((unsigned)y >= (unsigned)g.H || (unsigned)x >= (unsigned)g.W) return 0.0f;
        break;
    }
    return a[(size_t)y * g.W + x];
}

// Tap (i, j) of operand 2 in cross-correlation form: flipped for the
// convolution orientation, zero off the up-sampling lattice.
__device__ __forceinline__ float read_tap(const float* __restrict__ f, const Geom& g, int i, int j) {
    if (g.flip) {
        i = g.FHe - 1 - i;
        j = g.FWe - 1 - j;
    }
    if (g.upFr > 1) {
        if (i % g.upFr) return 0.0f;
        i /= g.upFr;
    }
    if (g.upFc > 1) {
        if (j % g.upFc) return 0.0f;
        j /= g.upFc;
    }
    return f[(size_t)i * g.FW + j];
}

// ---------------------------------------------------------------------------
// Direct kernel.  Block = 64*RG threads; it owns a (4*RG) x 64 output tile of
// one output plane.  Thread (rg, tx) computes rows 4*rg .. 4*rg+3 of column tx.
// For each plane pair and filter chunk, the input tile and the chunk's taps are
// staged in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_direct(Geom g, PlaneMap m, DirectCfg c, const float* __restrict__ A,
                                                const float* __restrict__ F, float* __restrict__ O, Epi e) {
    extern __shared__ float smem[];
    float* st = smem;
    float* sf = smem + (size_t)c.ldsRows * c.ldsCols;
    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tc = bid % c.tilesC;
    bid /= c.tilesC;
    const int tr = bid % c.tilesR;
    const int o = bid / c.tilesR;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int tx = tid & 63;
    const int rg = tid >> 6;
    const int tileRows = 4 * c.RG;
    const int orow0 = tr * tileRows, ocol0 = tc * 64;
    const int y0 = orow0 * g.dnR - g.hR, x0 = ocol0 * g.dnC - g.hC;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
    const int rbase = rg * 4 * g.dnR;
    for (int r = 0; r < m.R; r++) {
        int pa, pf;
        plane_pair(m, o, r, pa, pf);
        const float* a = A + (size_t)pa * g.H * g.W;
        const float* f = F + (size_t)pf * g.FH * g.FW;
        for (int i0 = 0; i0 < g.FHe; i0 += c.FCR) {
            const int fcr = min(c.FCR, g.FHe - i0);
            const int trows = (tileRows - 1) * g.dnR + fcr;
            for (int j0 = 0; j0 < g.FWe; j0 += c.FCC) {
                const int fcc = min(c.FCC, g.FWe - j0);
                const int tcols = 63 * g.dnC + fcc;
                for (int q = tid; q < fcr * fcc; q += nthr) {
                    const int i = q / fcc, j = q - i * fcc;
                    sf[q] = read_tap(f, g, i0 + i, j0 + j);
                }
                for (int q = tid; q < trows * tcols; q += nthr) {
                    const int ty = q / tcols, txx = q - ty * tcols;
                    st[ty * c.ldsCols + txx] = read_operand(a, g, y0 + i0 + ty, x0 + j0 + txx);
                }
                __syncthreads();
                const int cb = tx * g.dnC;
                for (int i = 0; i < fcr; i++) {
                    const float* row0 = st + (rbase + i) * c.ldsCols + cb;
                    const float* row1 = row0 + g.dnR * c.ldsCols;
                    const float* row2 = row1 + g.dnR * c.ldsCols;
                    const float* row3 = row2 + g.dnR * c.ldsCols;
                    const float* w = sf + i * fcc;
                    for (int j = 0; j < fcc; j++) {
                        const float t = w[j];
                        acc0 = fmaf(t, row0[j], acc0);
                        acc1 = fmaf(t, row1[j], acc1);
                        acc2 = fmaf(t, row2[j], acc2);
                        acc3 = fmaf(t, row3[j], acc3);
                    }
                }
                __syncthreads();
            }
        }
    }
    const int col = ocol0 + tx;
    if (col < g.OW) {
        float* out = O + (size_t)o * g.OH * g.OW + col;
        const int row = orow0 + rg * 4;
        if (row + 0 < g.OH) out[(size_t)(row + 0) * g.OW] = epi(e, acc0, o, (row + 0) * g.OW + col);
        if (row + 1 < g.OH) out[(size_t)(row + 1) * g.OW] = epi(e, acc1, o, (row + 1) * g.OW + col);
        if (row + 2 < g.OH) out[(size_t)(row + 2) * g.OW] = epi(e, acc2, o, (row + 2) * g.OW + col);
        if (row + 3 < g.OH) out[(size_t)(row + 3) * g.OW] = epi(e, acc3, o, (row + 3) * g.OW + col);
    }
}

hipError_t launch_direct(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                         float* out, hipStream_t s, const Epi& e) {
    const long blocks = (long)c.tilesC * c.tilesR * m.nOut;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_direct, dim3((unsigned)blocks), dim3(64 * c.RG), c.lds_bytes, s, g, m, c, a, f, out, e);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Register-tiled direct kernel for ops without sampling and filters of at most
// 16 x 15 (FW a template constant): a block of 256 threads owns an (8 RPT) x 128
// output tile of one output plane; thread (rg, cg) computes rows RPT rg ..
// RPT rg + RPT - 1 and columns 4cg .. 4cg+3.  For each plane pair the input
// tile (border applied) and the taps (flip applied, rows padded to a multiple
// of 4) are staged in LDS.  Each staged input row is read into registers once
// per thread (16-byte reads of 4 + FW - 1 values) and feeds the RPT output rows
// it touches, so a tap row costs 4 RPT FW FMAs for about FW/4 + RPT FW/4 LDS
// reads.  RPT = 2 or 1 shrinks the tile for planes that give too few blocks.
// ---------------------------------------------------------------------------
constexpr int RT_COLS = 128;

// Index v of a plane dimension of n samples after the border policy: clamped,
// wrapped, or (Zero, Valid, Full) false where the sample is outside.
__device__ __forceinline__ bool border_index(int v, int n, int border, int& out) {
    if (border == COGCONV_BORDER_CLAMP) {
        out = min(max(v, 0), n - 1);
        return true;
    }
    if (border == COGCONV_BORDER_CYCLIC) {
        v %= n;
        out = v < 0 ? v + n : v;
        return true;
    }
    out = v;
    return (unsigned)v < (unsigned)n;
}

__host__ __device__ inline int rt_seg_vecs(int fw) { return (fw + 6) / 4; }
__host__ __device__ inline int rt_tap_pitch(int fw) { return (fw + 3) & ~3; }

template <int FW, int RPT>
__global__ __launch_bounds__(256) void k_direct_rt(Geom g, PlaneMap m, DirectCfg c, const float* __restrict__ A,
                                                   const float* __restrict__ F, float* __restrict__ O, Epi e) {
    extern __shared__ float4 smem4[];
    float* smem = reinterpret_cast<float*>(smem4);
    constexpr int ROWS = 8 * RPT;         // output rows per tile
    constexpr int NV = (FW + 6) / 4;      // float4 reads per staged row segment
    constexpr int TP = (FW + 3) & ~3;     // tap row pitch
    constexpr int NT = TP / 4;
    const int TR = ROWS + g.FH - 1;       // staged rows
    const int TCP = c.ldsCols;            // staged row pitch, a multiple of 4, >= 124 + 4 NV
    float* st = smem;
    float* sf = smem + (size_t)TR * TCP;  // FH x TP taps, cross-correlation order
    int bid = xcd_swizzle(blockIdx.x, gridDim.x);
    const int tc = bid % c.tilesC;
    bid /= c.tilesC;
    const int tr = bid % c.tilesR;
    const int o = bid / c.tilesR;
    const int tid = threadIdx.x;
    const int cg = tid & 31, rg = tid >> 5;
    const int orow0 = tr * ROWS, ocol0 = tc * RT_COLS;
    const int y0 = orow0 - g.hR, x0 = ocol0 - g.hC;
    // Staging: half-block h (wave-uniform) stages the first 128 columns of rows
    // h, h+2, ... with lane column xs; the FW - 1 extra columns of every row are
    // spread over the block (up to KE elements a thread).  Every load of a plane
    // pair is issued before the first store.  Row indices after the border
    // policy are wave-uniform; a sample outside the plane reads through the
    // out-of-range offset (0).
    constexpr int SR = 4 * RPT + 8;                                   // rows per half-block, FH <= 16
    constexpr int KE = FW > 1 ? ((8 * RPT + 15) * (FW - 1) + 255) / 256 : 0;
    const int h = __builtin_amdgcn_readfirstlane(tid >> 7), xs = tid & 127;  // 128 | wave: h is uniform
    int xm;
    const bool xmv = border_index(x0 + xs, g.W, g.border, xm);
    const int vcol = xmv ? 4 * xm : STAGE_OOB;
    const int E = TR * (FW - 1);
    // Row table of the half-block: lane l holds the byte offset of row h + 2l
    // after the border policy; bit l of rmask says whether it is in the plane.
    // The load loop reads both with constant lane indices (scalar registers).
    int rowoff;
    unsigned long long rmask;
    {
        const int l = tid & 63, ty = h + 2 * l;
        int yi;
        const bool rv = l < SR && ty < TR && border_index(y0 + ty, g.H, g.border, yi);
        rowoff = rv ? 4 * yi * g.W : 0;
        rmask = __ballot(rv);
    }
    float acc[RPT][4];
#pragma unroll
    for (int a = 0; a < RPT; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[a][q] = 0.f;
    for (int r = 0; r < m.R; r++) {
        int pa, pf;
        plane_pair(m, o, r, pa, pf);
        const __amdgpu_buffer_rsrc_t ar = buffer_rsrc(A + (size_t)pa * g.H * g.W, 4u * (uint32_t)g.H * g.W);
        const float* f = F + (size_t)pf * g.FH * g.FW;
        __syncthreads();
        for (int q = tid; q < g.FH * TP; q += 256) {
            const int i = q / TP, j = q - i * TP;
            sf[q] = j < FW ? read_tap(f, g, i, j) : 0.f;
        }
        float vm[SR], ve[KE > 0 ? KE : 1];
        // Opaque copies: the per-row offsets, selects and store conditions are
        // formed inside the loop instead of being hoisted out of it into
        // registers held (and spilled) across the whole loop.
        int ro = rowoff, vc = vcol, trr = TR;
        unsigned long long rm = rmask;
        asm volatile("" : "+v"(ro), "+v"(vc), "+s"(trr), "+s"(rm));
#pragma unroll
        for (int u = 0; u < SR; u++) {
            const bool yv = (rm >> u) & 1ull;
            vm[u] = buffer_load_f32(ar, yv ? vc : STAGE_OOB, __builtin_amdgcn_readlane(ro, u));
        }
#pragma unroll
        for (int k = 0; k < KE; k++) {
            const int e = tid + 256 * k;
            const int ty = e / (FW > 1 ? FW - 1 : 1), tx = e - ty * (FW > 1 ? FW - 1 : 1);
            int yi, xi;
            const bool v = e < E && border_index(y0 + ty, g.H, g.border, yi) &&
                           border_index(x0 + RT_COLS + tx, g.W, g.border, xi);
            ve[k] = buffer_load_f32(ar, v ? 4 * (yi * g.W + xi) : STAGE_OOB, 0);
        }
        __builtin_amdgcn_sched_barrier(0);  // every load ahead of the first store
#pragma unroll
        for (int u = 0; u < SR; u++) {
            const int ty = h + 2 * u;
            if (ty < trr) st[ty * TCP + xs] = vm[u];
        }
#pragma unroll
        for (int k = 0; k < KE; k++) {
            const int e = tid + 256 * k;
            const int ty = e / (FW > 1 ? FW - 1 : 1), tx = e - ty * (FW > 1 ? FW - 1 : 1);
            if (e < E) st[ty * TCP + RT_COLS + tx] = ve[k];
        }
        __syncthreads();
        // Staged row RPT rg + yy feeds output row RPT rg + a through tap row yy - a.
        for (int yy = 0; yy < g.FH + RPT - 1; yy++) {
            const float4* row = reinterpret_cast<const float4*>(st + (RPT * rg + yy) * TCP) + cg;
            float seg[4 * NV];
#pragma unroll
            for (int k = 0; k < NV; k++) {
                const float4 v = row[k];
                seg[4 * k] = v.x;
                seg[4 * k + 1] = v.y;
                seg[4 * k + 2] = v.z;
                seg[4 * k + 3] = v.w;
            }
#pragma unroll
            for (int a = 0; a < RPT; a++) {
                const int i = yy - a;
                if (i < 0 || i >= g.FH) continue;
                const float4* w4 = reinterpret_cast<const float4*>(sf + i * TP);
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const float4 w = w4[t];
                    const float wt[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int jj = 0; jj < 4; jj++) {
                        const int j = 4 * t + jj;
                        if (j >= FW) break;
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[a][q] = fmaf(wt[jj], seg[q + j], acc[a][q]);
                    }
                }
            }
        }
    }
    float* out = O + (size_t)o * g.OH * g.OW;
    const int col = ocol0 + 4 * cg;
    // 16-byte stores where every row starts 16-byte aligned (then a column
    // group is either wholly inside the row or wholly past it).
    const bool vec = (g.OW & 3) == 0 && (reinterpret_cast<uintptr_t>(O) & 15) == 0;
    if (vec) {
#pragma unroll
        for (int a = 0; a < RPT; a++) {
            const int row = orow0 + RPT * rg + a;
            const int pt = row * g.OW + col;
            if (row < g.OH && col < g.OW)
                *reinterpret_cast<float4*>(out + (size_t)row * g.OW + col) =
                    make_float4(epi(e, acc[a][0], o, pt), epi(e, acc[a][1], o, pt + 1), epi(e, acc[a][2], o, pt + 2),
                                epi(e, acc[a][3], o, pt + 3));
        }
    } else {
#pragma unroll
        for (int a = 0; a < RPT; a++) {
            const int row = orow0 + RPT * rg + a;
            if (row >= g.OH) continue;
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (col + q < g.OW) out[(size_t)row * g.OW + col + q] = epi(e, acc[a][q], o, row * g.OW + col + q);
        }
    }
}

bool direct_rt_supported(const Geom& g) {
    return g.upAr == 1 && g.upAc == 1 && g.dnR == 1 && g.dnC == 1 && g.upFr == 1 && g.upFc == 1 && g.FH >= 1 &&
           g.FH <= 16 && g.FW >= 1 && g.FW <= 15 && (size_t)g.H * g.W * 4 < 0x80000000ull;
}

// Tile geometry and LDS of the register-tiled kernel with `rpt` (1, 2 or 4)
// output rows per thread.  The staged row pitch is a multiple of 4 (16-byte
// row reads) covering the 124 + 4 NV columns the last column group reads.
void direct_rt_config(const Geom& g, int rpt, DirectCfg& c) {
    c.RG = rpt;
    c.FCR = g.FH;
    c.FCC = g.FW;
    c.ldsRows = 8 * rpt + g.FH - 1;
    int p = RT_COLS + g.FW - 1;
    if (p < 124 + 4 * rt_seg_vecs(g.FW)) p = 124 + 4 * rt_seg_vecs(g.FW);
    c.ldsCols = (p + 3) & ~3;
    c.tilesR = (g.OH + 8 * rpt - 1) / (8 * rpt);
    c.tilesC = (g.OW + RT_COLS - 1) / RT_COLS;
    c.lds_bytes = ((size_t)c.ldsRows * c.ldsCols + (size_t)g.FH * rt_tap_pitch(g.FW)) * sizeof(float);
}

template <int RPT>
static hipError_t launch_rt(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                            float* out, hipStream_t s, dim3 grid, const Epi& e) {
    const dim3 block(256);
    switch (g.FW) {
#define RT_CASE(W) \
    case W: hipLaunchKernelGGL((k_direct_rt<W, RPT>), grid, block, c.lds_bytes, s, g, m, c, a, f, out, e); break;
        RT_CASE(1) RT_CASE(2) RT_CASE(3) RT_CASE(4) RT_CASE(5) RT_CASE(6) RT_CASE(7) RT_CASE(8) RT_CASE(9)
        RT_CASE(10) RT_CASE(11) RT_CASE(12) RT_CASE(13) RT_CASE(14) RT_CASE(15)
#undef RT_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_direct_rt(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                            float* out, hipStream_t s, const Epi& e) {
    const long blocks = (long)c.tilesC * c.tilesR * m.nOut;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    if (!direct_rt_supported(g) || c.ldsRows != 8 * c.RG + g.FH - 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)blocks);
    if (c.RG == 4) return launch_rt<4>(g, m, c, a, f, out, s, grid, e);
    if (c.RG == 2) return launch_rt<2>(g, m, c, a, f, out, s, grid, e);
    if (c.RG == 1) return launch_rt<1>(g, m, c, a, f, out, s, grid, e);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Small-output kernel: output planes of at most a few hundred points whose
// every element reduces over a large operand 2 (FilterAdjoint into a filter).
// Block (o, split s, output chunk) sums its share of the (plane pair, tap row,
// tap column) space for 16 outputs; the block reduces with wavefront shuffles
// and writes one partial per output; k_sum_splits adds the partials in split
// order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int SQ = 16;

__global__ __launch_bounds__(256) void k_small_output(Geom g, PlaneMap m, SmallCfg c, const float* __restrict__ A,
                                                      const float* __restrict__ F, float* __restrict__ partial) {
    const int o = blockIdx.x / c.splits;
    const int s = blockIdx.x - o * c.splits;
    const int P = g.OH * g.OW;
    const int q0 = blockIdx.y * SQ;
    const int nq = min(SQ, P - q0);
    int yb[SQ], xb[SQ];
#pragma unroll
    for (int q = 0; q < SQ; q++) {
        const int oq = q0 + min(q, nq - 1);
        const int orow = oq / g.OW, ocol = oq - orow * g.OW;
        yb[q] = orow * g.dnR - g.hR;
        xb[q] = ocol * g.dnC - g.hC;
    }
    float acc[SQ];
#pragma unroll
    for (int q = 0; q < SQ; q++) acc[q] = 0.f;
    const long rowsTotal = (long)m.R * g.FHe;
    const long rb = rowsTotal * s / c.splits, re = rowsTotal * (s + 1) / c.splits;
    const long work = (re - rb) * g.FWe;
    for (long p = threadIdx.x; p < work; p += blockDim.x) {
        const long row = rb + p / g.FWe;
        const int j = (int)(p % g.FWe);
        const int r = (int)(row / g.FHe);
        const int i = (int)(row - (long)r * g.FHe);
        int pa, pf;
        plane_pair(m, o, r, pa, pf);
        const float w = read_tap(F + (size_t)pf * g.FH * g.FW, g, i, j);
        if (w != 0.f) {
            const float* a = A + (size_t)pa * g.H * g.W;
#pragma unroll
            for (int q = 0; q < SQ; q++) acc[q] = fmaf(w, read_operand(a, g, yb[q] + i, xb[q] + j), acc[q]);
        }
    }
    __shared__ float red[4][SQ];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < SQ; q++) {
        float v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wv][q] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nq) {
        const int nw = blockDim.x >> 6;
        float t = 0.f;
        for (int w = 0; w < nw; w++) t += red[w][threadIdx.x];
        partial[((size_t)s * m.nOut + o) * P + q0 + threadIdx.x] = t;
    }
}

__global__ void k_sum_splits(const float* __restrict__ partial, float* __restrict__ out, int splits, long total,
                             long points, Epi e) {
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        float t = 0.f;
        for (int s = 0; s < splits; s++) t += partial[(size_t)s * total + idx];
        const long p = idx / points;
        out[idx] = epi(e, t, (int)p, (int)(idx - p * points));
    }
}

hipError_t launch_small_output(const Geom& g, const PlaneMap& m, const SmallCfg& c, const float* a, const float* f,
                               float* out, float* workspace, hipStream_t s, const Epi& ep) {
    const long gx = (long)m.nOut * c.splits;
    if (gx <= 0 || gx > 0x7fffffffL || c.qchunks > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_small_output, dim3((unsigned)gx, c.qchunks), dim3(256), 0, s, g, m, c, a, f, workspace);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long total = (long)m.nOut * g.OH * g.OW;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)blocks), dim3(256), 0, s, workspace, out, c.splits, total,
                       (long)g.OH * g.OW, ep);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// blockReduceSum: TensorReduceHyperKernel.scala:39-207.  Output plane p sums
// input planes p*factor .. p*factor+factor-1 in order; consecutive threads
// take consecutive pixels, so reads and writes are coalesced.
// ---------------------------------------------------------------------------
__global__ void k_block_reduce_sum(const float* __restrict__ in, float* __restrict__ out, long points, int outPlanes,
                                   int factor, Epi e) {
    const long total = points * outPlanes;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long p = idx / points, q = idx - p * points;
        const float* src = in + (size_t)p * factor * points + q;
        float t = 0.f;
        for (int j = 0; j < factor; j++) t += src[(size_t)j * points];
        out[idx] = epi(e, t, (int)p, (int)q);
    }
}

hipError_t launch_block_reduce_sum(const float* in, float* out, int64_t points, int out_planes, int factor,
                                   hipStream_t s, const Epi& e) {
    const long total = (long)points * out_planes;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_block_reduce_sum, dim3((unsigned)blocks), dim3(256), 0, s, in, out, (long)points, out_planes,
                       factor, e);
    return hipGetLastError();
}

// The elementwise pass of an epilogue over a stored field (in == out allowed).
__global__ void k_elementwise(const float* in, float* out, long points, int planes, Epi e) {
    const long total = points * planes;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long p = idx / points;
        out[idx] = epi(e, in[idx], (int)p, (int)(idx - p * points));
    }
}

hipError_t launch_elementwise(const float* in, float* out, int64_t points, int planes, const Epi& e, hipStream_t s) {
    const long total = (long)points * planes;
    if (total <= 0) return hipSuccess;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_elementwise, dim3((unsigned)blocks), dim3(256), 0, s, in, out, (long)points, planes, e);
    return hipGetLastError();
}

__global__ void k_fill_zero16(uint4* __restrict__ p, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_fill_zero_bytes(unsigned char* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

// Zero-fill by kernels rather than hipMemsetAsync, so that the fill can sit in
// a captured graph.  Unaligned head and tail bytes get a byte kernel.
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
    unsigned char* b = static_cast<unsigned char*>(p);
    size_t head = (16 - reinterpret_cast<uintptr_t>(b) % 16) % 16;
    if (head > bytes) head = bytes;
    const size_t n16 = (bytes - head) / 16;
    const size_t tail = bytes - head - n16 * 16;
    if (head) hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, s, b, head);
    if (n16) {
        const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_fill_zero16, dim3((unsigned)blocks), dim3(256), 0, s,
                           reinterpret_cast<uint4*>(b + head), n16);
    }
    if (tail) hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, s, b + head + n16 * 16, tail);
    return hipGetLastError();
}

// Constant fill of n floats (the autotuner's synthetic inputs).
__global__ void k_fill_value(float* __restrict__ p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill_value(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_value, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace cogconv
