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
#include "internal.h"

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
                                                const float* __restrict__ F, float* __restrict__ O) {
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
        if (row + 0 < g.OH) out[(size_t)(row + 0) * g.OW] = acc0;
        if (row + 1 < g.OH) out[(size_t)(row + 1) * g.OW] = acc1;
        if (row + 2 < g.OH) out[(size_t)(row + 2) * g.OW] = acc2;
        if (row + 3 < g.OH) out[(size_t)(row + 3) * g.OW] = acc3;
    }
}

hipError_t launch_direct(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                         float* out, hipStream_t s) {
    const long blocks = (long)c.tilesC * c.tilesR * m.nOut;
    if (blocks <= 0 || blocks > 0x7fffffffL) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_direct, dim3((unsigned)blocks), dim3(64 * c.RG), c.lds_bytes, s, g, m, c, a, f, out);
    return hipGetLastError();
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

__global__ void k_sum_splits(const float* __restrict__ partial, float* __restrict__ out, int splits, long total) {
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        float t = 0.f;
        for (int s = 0; s < splits; s++) t += partial[(size_t)s * total + idx];
        out[idx] = t;
    }
}

hipError_t launch_small_output(const Geom& g, const PlaneMap& m, const SmallCfg& c, const float* a, const float* f,
                               float* out, float* workspace, hipStream_t s) {
    const long gx = (long)m.nOut * c.splits;
    if (gx <= 0 || gx > 0x7fffffffL || c.qchunks > 65535) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_small_output, dim3((unsigned)gx, c.qchunks), dim3(256), 0, s, g, m, c, a, f, workspace);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long total = (long)m.nOut * g.OH * g.OW;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)blocks), dim3(256), 0, s, workspace, out, c.splits, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// blockReduceSum: TensorReduceHyperKernel.scala:39-207.  Output plane p sums
// input planes p*factor .. p*factor+factor-1 in order; consecutive threads
// take consecutive pixels, so reads and writes are coalesced.
// ---------------------------------------------------------------------------
__global__ void k_block_reduce_sum(const float* __restrict__ in, float* __restrict__ out, long points, int outPlanes,
                                   int factor) {
    const long total = points * outPlanes;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long p = idx / points, q = idx - p * points;
        const float* src = in + (size_t)p * factor * points + q;
        float t = 0.f;
        for (int j = 0; j < factor; j++) t += src[(size_t)j * points];
        out[idx] = t;
    }
}

hipError_t launch_block_reduce_sum(const float* in, float* out, int64_t points, int out_planes, int factor,
                                   hipStream_t s) {
    const long total = (long)points * out_planes;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_block_reduce_sum, dim3((unsigned)blocks), dim3(256), 0, s, in, out, (long)points, out_planes,
                       factor);
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
