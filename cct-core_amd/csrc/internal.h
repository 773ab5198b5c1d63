// Internal declarations of libcogconv (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "cogconv.h"

namespace cogconv {

// Record a message for cogconv_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int check_hip(hipError_t e, const char* what);
void clear_error();

// Plane-level geometry of one ConvolveOp, in "materialised" coordinates:
// operand 1 (in0) up-sampled by (upAr, upAc) is UR x UC, operand 2 (in1)
// up-sampled by (upFr, upFc) is FHe x FWe.  Output element (R, C) sums
//   tap(i, j) * A'(R*dnR - hR + i, C*dnC - hC + j)
// over i < FHe, j < FWe (ConvolveHyperKernel.scala:404-548, 824-830).
struct Geom {
    int H, W;        // stored operand-1 plane
    int FH, FW;      // stored operand-2 plane
    int OH, OW;      // output plane
    int upAr, upAc;  // zero insertion of operand 1 (non-FilterAdjoint upsampling)
    int upFr, upFc;  // zero insertion of operand 2 (FilterAdjoint upsampling)
    int dnR, dnC;    // output step (downsampling)
    int UR, UC;      // H*upAr, W*upAc
    int FHe, FWe;    // FH*upFr, FW*upFc
    int hR, hC;      // low halo: Full F-1, Valid 0, otherwise (F-1)/2
    int border;      // COGCONV_BORDER_*
    int flip;        // 1 for the convolution orientation
};

// Which operand planes feed output plane o (ConvolveHyperKernel.scala:170-199, 319-340, 575-645).
struct PlaneMap {
    int mode;   // COGCONV_* vector mode
    int B;      // batch
    int N;      // operand-1 planes per image
    int T;      // operand-2 planes
    int PA;     // operand-1 planes
    int PFn;    // operand-2 planes
    int nOut;   // output planes
    int R;      // plane pairs summed per output plane
};

__host__ __device__ inline void plane_pair(const PlaneMap& m, int o, int r, int& pa, int& pf) {
    const int B = m.B, N = m.N, T = m.T;
    switch (m.mode) {
    case COGCONV_PLANE_BY_PLANE:
        pa = m.PA == 1 ? 0 : o;
        pf = m.PFn == 1 ? 0 : o;
        break;
    case COGCONV_PROJECT_FRAME: {  // o = b*T + w
        const int b = o / T, w = o - b * T;
        pa = b * N + w % N;
        pf = w;
        break;
    }
    case COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM: {  // o = b*K + k, term r = plane n
        const int K = T / N, b = o / K, k = o - b * K;
        pa = b * N + r;
        pf = k * N + r;
        break;
    }
    case COGCONV_BACK_PROJECT_FRAME: {  // o = b*T + w, L = T/N
        const int L = T / N, b = o / T, w = o - b * T;
        pa = b * N + w % N;
        pf = L * (w % N) + w / N;
        break;
    }
    case COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM: {  // o = b*L + l, term r = plane m
        const int L = T / N, b = o / L, l = o - b * L;
        pa = b * N + r;
        pf = r * L + l;
        break;
    }
    case COGCONV_FILTER_ADJOINT: {  // o = w*B + b (batch-interleaved output)
        const int K = T / B, w = o / B, b = o - w * B;
        pa = b * N + w % N;
        pf = b * K + w / N;
        break;
    }
    default: {  // COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM: o = w, term r = image b
        const int K = T / B;
        pa = r * N + o % N;
        pf = r * K + o / N;
        break;
    }
    }
}

// Block index remapped so that each of the 8 XCDs (which take launched blocks
// round-robin: block i runs on XCD i % 8) works on a contiguous range of tiles,
// whose shared halo rows and filter slices then hit the same L2.  Bijective on
// [0, n): the first 8*floor(n/8) blocks are permuted, the tail keeps its index.
__device__ __forceinline__ int xcd_swizzle(int bid, int n) {
    const int per = n >> 3;
    if (bid >= per * 8) return bid;
    return (bid & 7) * per + (bid >> 3);
}

// n / d for 0 <= n < 2^22 and 1 <= d, from a float reciprocal plus one
// correction step each way (exact in that range).  Used by the staging loops
// in place of an integer division.
__device__ __forceinline__ int fast_div(int n, int d, float inv_d) {
    int q = (int)((float)n * inv_d);
    q -= (q * d > n) ? 1 : 0;
    q += ((q + 1) * d <= n) ? 1 : 0;
    return q;
}

// Elementwise epilogue of a plan (cogconv_plan_set_epilogue): n ops applied in
// order to every output element (plane, point) at its store.  Field operand
// `slot` is read at (srcPlanes == 1 ? 0 : plane) * srcPoints + (srcPoints == 1 ?
// 0 : point): same type as the output, a ScalarField of its field shape, a 0-D
// field with its planes, or a 0-D scalar (ResultType.scala:29-58).  n == 0: none.
struct Epi {
    int n;
    int op[COGCONV_EPI_MAX_OPS];
    float val[COGCONV_EPI_MAX_OPS];
    int slot[COGCONV_EPI_MAX_OPS];
    const float* src[COGCONV_EPI_MAX_OPERANDS];
    int srcPlanes[COGCONV_EPI_MAX_OPERANDS];
    int srcPoints[COGCONV_EPI_MAX_OPERANDS];
};

// Kernel families.
enum KernelKind {
    KK_DIRECT = 1,
    KK_SMALL_OUTPUT = 2,
    KK_IGEMM = 3,
    KK_FA_IGEMM = 4,
    KK_IGEMM_LP = 5,
    KK_FA_IGEMM_LP = 6,
    KK_IGEMM_UP = 7,
    KK_DIRECT_RT = 8,
    KK_FFT = 9,
};

// LDS-tiled direct kernel configuration.
struct DirectCfg {
    int RG;        // row groups of 64 threads; a block is 64*RG threads, tile 4*RG x 64 outputs
    int FCR, FCC;  // filter chunk staged per pass
    int ldsCols;   // row pitch of the staged input tile
    int ldsRows;
    int tilesR, tilesC;
    size_t lds_bytes;
};

// FFT route of a PlaneByPlane plan: expanded plane Er x Ec; PA / PK image and
// filter planes, P output planes.
struct FftCfg {
    int Er, Ec;
    int PA, PK, P;
};

// Split-reduction kernel for small outputs (FilterAdjoint to a filter-sized result).
struct SmallCfg {
    int splits;   // partial sums per output plane
    int qchunks;  // ceil(OH*OW / 16)
};

// Implicit-GEMM configuration (MFMA).
struct IgemmCfg {
    int variant;      // kernel family (IgemmVariant)
    int MB;           // 16-plane output blocks per workgroup
    int RC;           // input planes staged in LDS per pass (multiple of 4)
    int Mout;         // output planes per image
    int mTiles;       // ceil(Mout / (16*MB))
    int tilesR, tilesC;
    int tileRows, tileCols;  // staged input tile (rows x cols)
    int planeStride;  // LDS floats per staged input plane (padded)
    int aStride;      // LDS floats per staged filter-slice plane (padded)
    int splits;       // split-K count (FilterAdjoint)
    int vec4;         // k_igemm_lp: stage 4-column groups with 16-byte loads (see choose_igemm_lp)
    int trg;          // k_igemm_lp: filter tap rows staged in LDS at a time (F: the whole filter)
    int ring;         // k_igemm_lp: one tap row a group, two LDS buffers, the next row prefetched (trg = 1)
    size_t lds_bytes;
};

// FilterAdjointBlockReduceSum implicit GEMM (split reduction).
struct FaCfg {
    int K;          // dy planes per image (GEMM rows)
    int MB;         // 16-row blocks (K <= 16*MB <= 64)
    int CBW;        // 16-column blocks per block (column group width), 3, 6 or 9
    int RR;         // tap rows staged per item
    int CW;         // tap columns per column chunk
    int colChunks;  // ceil(FWe / CW)
    int DWp;        // CW rounded up to 4
    int MTp;        // staged dy row pitch: 16*MB padded to 16 mod 32
    int xRows, xCols;
    int xColsP, xPlaneP;  // LDS pitches of staged x rows and planes (FP32 kernel; bank-spread padding)
    int nCntMax;    // input planes staged per column group
    int colGroups;  // column groups of 16*CBW columns
    int rowChunks;  // ceil(FHe / RR)
    int splits;     // partial results
    int NW;         // waves per block (FP32 kernel: 4, or 8 with the tap columns split between wave pairs)
    // Row-streaming BF16 / FP16 kernel (k_fa_rows_lp, rows = 1): nb x kh waves,
    // kbw dy-plane blocks a wave; tiles of kh*kbw*16 dy planes by nb*16 x
    // planes; 32-pixel column chunks of the dy rows; row splits of rowsPer
    // rows.
    int rows, nb, kh, kbw, nTiles, kTiles, chunks, rowSplits, rowsPer;
    size_t lds_bytes;
};

// One launch of the FP32 implicit GEMM: its (sub-)filter is taps (i0 +
// step*a, j0 + step*b) of the full FHf x FWf filter (flipped for the
// convolution orientation), and its output (row, col) lands at (oRow0 +
// oStride*row, oCol0 + oStride*col) of the full OHf x OWf output.  Up-sampled
// input runs as step^2 such phases, each a dense convolution of the stored
// samples; without up-sampling there is one phase with the identity mapping.
struct Phase {
    int i0, j0, step;
    int FHf, FWf, flip;
    int oRow0, oCol0, oStride;
    int OHf, OWf;
};

// The four phases of 2x up-sampling in one launch (k_igemm_f32_up2q): a block
// stages one window of stored samples and computes all four phases of a tile
// of 8 x 32 phase pixels.  Row phase pr's sub-filter rows start dr[pr] rows
// into the window (hU: the window's low halo in stored rows), column phase
// pc's dc[pc] columns in; phase (pr, pc) covers OHp[pr] x OWp[pc] pixels.
struct UpFused {
    int on;
    int hp;        // parity of the full filter's low halo (selects the tap split)
    int two;       // registers held for two blocks a CU (LDS at most 80 KB)
    IgemmCfg c;    // tilesR / tilesC in phase pixels; tileRows / tileCols: the window
    int hU, wU;
    int dr[2], dc[2];
    int OHp[2], OWp[2];
};

// The phase launches of an up-sampled frame-mode plan (2x: four phases), or
// one fused launch of all four.
struct PhaseSet {
    int n;
    Geom g[4];
    IgemmCfg c[4];
    Phase p[4];
    UpFused f;
};

enum IgemmVariant {
    IV_F32_FRAME = 1,  // FP32 MFMA, ProjectFrame/BackProjectFrame block-reduce modes
};

// The empty epilogue (plain stores).
inline const Epi kNoEpi = {};
// Checks an epilogue against an output of type `t` (cogconv_plan_set_epilogue's
// rules) and builds its device form; `operands` receives its field count.
int epi_build(const cogconv_field_type& t, const cogconv_epilogue* in, Epi* out, int* operands);
// Binds the n field operands of `e` to device buffers.
int epi_bind(Epi* e, int n, const float* const* operands);

// Kernel launchers (device code in *.hip).  `e` is the plan's elementwise
// epilogue, applied by the kernel that stores the final output.
// Register-tiled direct kernel: no sampling, FH <= 16, FW <= 15.
bool direct_rt_supported(const Geom& g);
void direct_rt_config(const Geom& g, int rpt, DirectCfg& c);
hipError_t launch_direct_rt(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                            float* out, hipStream_t s, const Epi& e = kNoEpi);
hipError_t launch_direct(const Geom& g, const PlaneMap& m, const DirectCfg& c, const float* a, const float* f,
                         float* out, hipStream_t s, const Epi& e = kNoEpi);
hipError_t launch_small_output(const Geom& g, const PlaneMap& m, const SmallCfg& c, const float* a, const float* f,
                               float* out, float* workspace, hipStream_t s, const Epi& e = kNoEpi);
bool igemm_f32_supported(int F);
size_t igemm_f32_exchange_floats(int MB);
bool igemm_f32_supported_rc(int FR, int FC);
// `tab` (device memory, one entry per image of the plan's batch): each image's
// operand-1 planes and output planes come from its entry instead of x / out.
// Taps of row (column) phase p of a 2x up-sampled F-tap filter whose low halo
// has parity hp: the taps i with i = (hp ^ p) mod 2.
constexpr int up2_taps(int F, int hp, int p) { return (F - ((hp ^ p) & 1) + 1) / 2; }
bool igemm_f32_up2q_supported(int F);
hipError_t launch_igemm_f32_up2q(const Geom& g, const PlaneMap& m, const UpFused& u, const float* x, const float* w,
                                 float* out, hipStream_t s, const Epi& e);
hipError_t launch_igemm_f32_phase(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const Phase& ph, const float* x,
                                  const float* w, float* out, hipStream_t s, const Epi& e = kNoEpi,
                                  const cogconv_image_io* tab = nullptr);
hipError_t launch_igemm_f32(const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x, const float* w,
                            float* out, hipStream_t s, const Epi& e = kNoEpi, const cogconv_image_io* tab = nullptr);
hipError_t launch_fa_igemm_f32(const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x, const float* dy,
                               float* out, float* ws, hipStream_t s, const Epi& e = kNoEpi);
hipError_t launch_reduce_partials(const float* ws, float* out, int splits, long total, long points, hipStream_t s,
                                  const Epi& e);
size_t fa_igemm_lp_lds_bytes(const FaCfg& c, int OW, int precision);
size_t fa_rows_lp_lds_bytes(const FaCfg& c, int OD);
hipError_t launch_fa_rows_lp(int precision, const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x,
                             const float* dy, float* out, float* ws, hipStream_t s, const Epi& ep);
hipError_t launch_fa_igemm_lp(int precision, const Geom& g, const PlaneMap& m, const FaCfg& c, const float* x,
                              const float* dy, float* out, float* ws, hipStream_t s, const Epi& e = kNoEpi);
bool igemm_lp_supported(int F);
size_t igemm_lp_lds_bytes(const IgemmCfg& c, int F, int precision);
bool igemm_lp_two_blocks(const IgemmCfg& c);
size_t igemm_lp_workspace_bytes(const IgemmCfg& c, int F, int precision);
enum LpStage { LP_PACK_AND_RUN = 0, LP_PACK_ONLY = 1, LP_RUN_ONLY = 2 };
// The BF16 / FP16 filter pack of k_igemm_lp (k_pack_filters_lp) on its own.
hipError_t launch_pack_filters_lp(int precision, const Geom& g, const PlaneMap& m, int Mout, int Mpad, int Rpad,
                                  const float* w, void* wp, hipStream_t s);
// BF16 / FP16 fused four-phase 2x up-sampling (up2q_lp.hip): LDS bytes and
// packed-filter workspace of a configuration; the launch packs the bank into
// ws and runs the kernel.
size_t up2q_lp_lds_bytes(const IgemmCfg& c, int F);
size_t up2q_lp_workspace_bytes(const IgemmCfg& c, int F);
hipError_t launch_up2q_lp(int precision, const Geom& g, const PlaneMap& m, const UpFused& u, const float* x,
                          const float* w, float* out, void* ws, hipStream_t s, const Epi& e);
hipError_t launch_igemm_lp(int precision, const Geom& g, const PlaneMap& m, const IgemmCfg& c, const float* x,
                           const float* w, float* out, void* ws, hipStream_t s, int stage = LP_PACK_AND_RUN,
                           const Epi& e = kNoEpi, const cogconv_image_io* tab = nullptr, bool x16 = false);
hipError_t launch_block_reduce_sum(const float* in, float* out, int64_t points, int out_planes, int factor,
                                   hipStream_t s, const Epi& e = kNoEpi);
// out = epilogue(in) over `planes` planes of `points` (the unmerged elementwise pass).
hipError_t launch_elementwise(const float* in, float* out, int64_t points, int planes, const Epi& e, hipStream_t s);
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);
// FFT route (fft.hip): good transform lengths, workspace, per-plan hipFFT state.
int fft_good_length(int m);
size_t fft_workspace_bytes(const FftCfg& c);
void* fft_state_new();
void fft_release(void* state);
int launch_fft(const Geom& g, const FftCfg& c, void** state, const float* x, const float* f, float* out, void* ws,
               hipStream_t s, const Epi& e);
hipError_t launch_fill_value(float* p, size_t n, float v, hipStream_t s);

// Request-table launches (plan.cpp): plans that take one, and the launch on a
// table already in device memory (x16: every entry's operand 1 is 16-byte aligned).
bool plan_table_capable(const cogconv_plan_s* plan);
int launch_table(cogconv_plan plan, const cogconv_image_io* device_table, bool x16, const float* in1, void* workspace,
                 bool packed, hipStream_t s);

// Device address of the no-allocation entry point's scratch array on `device`
// (current), taken once per device (noalloc.hip).
int noalloc_setup(int device);

// Candidate configurations of a plan (plan.cpp) and the autotuning cache
// (tune.cpp): the tuned variant index for `desc` on the current device and the
// description it had when it was timed.
int plan_variant_count(const cogconv_desc* desc, int* count);
bool tune_cache_lookup(const cogconv_desc* desc, int* variant, std::string* name);

}  // namespace cogconv

// The layout of a plan's packed operand 2 (the low-precision frame kernels'
// filter bank): plans with equal keys pack a bank identically (the field size
// and batch do not enter), so one packed copy serves them all.  Empty for
// plans that do not pack.
std::string cogconv_plan_pack_key(const struct cogconv_plan_s* plan);

struct cogconv_plan_s {
    cogconv_desc desc;
    cogconv_field_type out;
    cogconv::Geom g;
    cogconv::PlaneMap m;
    int kind;
    cogconv::DirectCfg dcfg;
    cogconv::SmallCfg scfg;
    cogconv::IgemmCfg icfg;
    cogconv::FaCfg fcfg;
    cogconv::PhaseSet ps;
    cogconv::FftCfg fft;
    void* fft_state;  // hipFFT handles of an FFT plan (made at its first launch)
    size_t workspace;
    double flops;
    double bytes;
    int variant;  // index into the plan's candidate list (autotuning)
    cogconv::Epi epi;       // elementwise epilogue (epi.n == 0: none); src[] set per launch
    int epi_operands;       // field operands the epilogue reads
};
