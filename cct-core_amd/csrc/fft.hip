// FFT route of PlaneByPlane convolution (DynamicConvolutionGenerator.convolveBy2DFFT,
// src/main/scala/cogx/compiler/codegenerator/opencl/generator/DynamicConvolutionGenerator.scala:322-476,
// with the expanded shape of FFTConvolutionUtils.fftShape2D, FFTConvolutionUtils.scala:50-97).
//
// Per plane: the image is expanded to Er x Ec with its border policy (the rows
// and columns past the image hold its continuation on the high side and, read
// cyclically, on the low side: ExpandBorderHyperKernel.scala:48-130), the filter
// is laid out so that a circular convolution gives the reference's taps (tap
// (i, j) of the cross-correlation form lands at ((h - i) mod Er, (hc - j) mod Ec)),
// both go through a real-to-complex FFT, their spectra multiply with the
// 1 / (Er Ec) scaling folded in, and the inverse FFT is trimmed back to the
// image.  With Er >= H + 2h (Ec >= W + 2hc) the circular convolution equals the
// bordered one on [0, H) x [0, W) for every border; a Cyclic image whose size
// is itself a good FFT length needs no expansion.
//
// The transforms are hipFFT's; the layout, fill, spectrum product and trim are
// the kernels below.  A plan's hipFFT handles are made at its first launch (they
// allocate their own work area then) and reused.
#include <hipfft/hipfft.h>

#include <mutex>

#include "epi.h"
#include "internal.h"

namespace cogconv {

// Smallest n >= m whose factors are all 2, 3, 5 or 7 (lengths hipFFT runs with
// its radix kernels).
int fft_good_length(int m) {
    for (int n = m > 1 ? m : 1;; n++) {
        int r = n;
        for (int p : {2, 3, 5, 7})
            while (r % p == 0) r /= p;
        if (r == 1) return n;
    }
}

// Signed coordinate of expanded index u: the high side continues the image,
// the rest is the wrap of negative coordinates (ExpandBorderHyperKernel.scala:54-118).
__device__ __forceinline__ int fft_signed(int u, int n, int e) { return u < n + (e - n) / 2 ? u : u - e; }

__global__ void k_fft_expand_image(const float* __restrict__ X, float* __restrict__ out, int H, int W, int Er, int Ec,
                                   int border, long total) {
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long plane = idx / ((long)Er * Ec);
        const int q = (int)(idx - plane * Er * Ec);
        const int u = q / Ec, v = q - u * Ec;
        int y = fft_signed(u, H, Er), x = fft_signed(v, W, Ec);
        float val = 0.f;
        if (border == COGCONV_BORDER_CLAMP) {
            y = min(max(y, 0), H - 1);
            x = min(max(x, 0), W - 1);
            val = X[(size_t)plane * H * W + (size_t)y * W + x];
        } else if (border == COGCONV_BORDER_CYCLIC) {
            y %= H;
            y += y < 0 ? H : 0;
            x %= W;
            x += x < 0 ? W : 0;
            val = X[(size_t)plane * H * W + (size_t)y * W + x];
        } else if (y >= 0 && y < H && x >= 0 && x < W) {
            val = X[(size_t)plane * H * W + (size_t)y * W + x];
        }
        out[idx] = val;
    }
}

// G((h - i) mod Er, (hc - j) mod Ec) = t(i, j), t = the taps in cross-correlation
// form (the filter itself, or the filter flipped for the convolution orientation).
__global__ void k_fft_expand_filter(const float* __restrict__ F, float* __restrict__ out, int FH, int FW, int Er,
                                    int Ec, int flip, long total) {
    const int h = (FH - 1) / 2, hc = (FW - 1) / 2;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long plane = idx / ((long)Er * Ec);
        const int q = (int)(idx - plane * Er * Ec);
        const int u = q / Ec, v = q - u * Ec;
        int i = h - u, j = hc - v;
        i += i < 0 ? Er : 0;
        j += j < 0 ? Ec : 0;
        float val = 0.f;
        if (i < FH && j < FW) {
            const int si = flip ? FH - 1 - i : i, sj = flip ? FW - 1 - j : j;
            val = F[(size_t)plane * FH * FW + (size_t)si * FW + sj];
        }
        out[idx] = val;
    }
}

// O[p] = A[pa] * K[pk] * scale over `n` complex bins per plane.
__global__ void k_fft_spectrum_product(const float2* __restrict__ A, const float2* __restrict__ K,
                                       float2* __restrict__ O, long n, int PA, int PK, int P, float scale) {
    const long total = n * P;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long p = idx / n, q = idx - p * n;
        const float2 a = A[(PA == 1 ? 0 : p) * n + q];
        const float2 k = K[(PK == 1 ? 0 : p) * n + q];
        O[idx] = make_float2((a.x * k.x - a.y * k.y) * scale, (a.x * k.y + a.y * k.x) * scale);
    }
}

__global__ void k_fft_trim(const float* __restrict__ R, float* __restrict__ O, int H, int W, int Er, int Ec, long total,
                           Epi e) {
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long plane = idx / ((long)H * W);
        const int q = (int)(idx - plane * H * W);
        const int y = q / W, x = q - y * W;
        O[idx] = epi(e, R[(size_t)plane * Er * Ec + (size_t)y * Ec + x], (int)plane, q);
    }
}

namespace {

unsigned grid_for(long total) { return (unsigned)std::min<long>((total + 255) / 256, 65536L); }

struct FftState {
    std::mutex mu;
    bool made = false;
    hipfftHandle fwdA = 0, fwdK = 0, inv = 0;
};

int fft_check(hipfftResult r, const char* what) {
    if (r == HIPFFT_SUCCESS) return COGCONV_OK;
    return fail(COGCONV_ERR_RUNTIME, "%s failed (hipFFT status %d)", what, (int)r);
}

int make_handle(hipfftHandle* h, int Er, int Ec, hipfftType type, int batch) {
    int n[2] = {Er, Ec};
    const int bins = Er * (Ec / 2 + 1);
    const int reals = Er * Ec;
    const bool r2c = type == HIPFFT_R2C;
    return fft_check(hipfftPlanMany(h, 2, n, nullptr, 1, r2c ? reals : bins, nullptr, 1, r2c ? bins : reals, type,
                                    batch),
                     "hipfftPlanMany");
}

}  // namespace

size_t fft_workspace_bytes(const FftCfg& c) {
    const size_t reals = (size_t)c.Er * c.Ec, bins = (size_t)c.Er * (c.Ec / 2 + 1);
    // realA, realK, realO (floats); specA, specK, specO (complex)
    return 4 * reals * ((size_t)c.PA + c.PK + c.P) + 8 * bins * ((size_t)c.PA + c.PK + c.P) + 256 * 6;
}

void fft_release(void* state) {
    FftState* st = static_cast<FftState*>(state);
    if (!st) return;
    if (st->fwdA) (void)hipfftDestroy(st->fwdA);
    if (st->fwdK && st->fwdK != st->fwdA) (void)hipfftDestroy(st->fwdK);
    if (st->inv) (void)hipfftDestroy(st->inv);
    delete st;
}

int launch_fft(const Geom& g, const FftCfg& c, void** state, const float* x, const float* f, float* out, void* ws,
               hipStream_t s, const Epi& e) {
    if (!*state) return fail(COGCONV_ERR_RUNTIME, "internal: FFT plan state missing");
    FftState* st = static_cast<FftState*>(*state);
    const size_t reals = (size_t)c.Er * c.Ec, bins = (size_t)c.Er * (c.Ec / 2 + 1);
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    char* p = static_cast<char*>(ws);
    float* realA = reinterpret_cast<float*>(p);
    p += align(4 * reals * c.PA);
    float* realK = reinterpret_cast<float*>(p);
    p += align(4 * reals * c.PK);
    float* realO = reinterpret_cast<float*>(p);
    p += align(4 * reals * c.P);
    float2* specA = reinterpret_cast<float2*>(p);
    p += align(8 * bins * c.PA);
    float2* specK = reinterpret_cast<float2*>(p);
    p += align(8 * bins * c.PK);
    float2* specO = reinterpret_cast<float2*>(p);
    std::lock_guard<std::mutex> lk(st->mu);  // hipFFT handles carry their stream: one launch at a time per plan
    if (!st->made) {
        if (int r = make_handle(&st->fwdA, c.Er, c.Ec, HIPFFT_R2C, c.PA)) return r;
        if (c.PK == c.PA) st->fwdK = st->fwdA;
        else if (int r = make_handle(&st->fwdK, c.Er, c.Ec, HIPFFT_R2C, c.PK)) return r;
        if (int r = make_handle(&st->inv, c.Er, c.Ec, HIPFFT_C2R, c.P)) return r;
        st->made = true;
    }
    const long tA = (long)reals * c.PA, tK = (long)reals * c.PK;
    hipLaunchKernelGGL(k_fft_expand_image, dim3(grid_for(tA)), dim3(256), 0, s, x, realA, g.H, g.W, c.Er, c.Ec,
                       g.border, tA);
    hipLaunchKernelGGL(k_fft_expand_filter, dim3(grid_for(tK)), dim3(256), 0, s, f, realK, g.FH, g.FW, c.Er, c.Ec,
                       g.flip, tK);
    if (int r = check_hip(hipGetLastError(), "FFT layout launch")) return r;
    for (hipfftHandle h : {st->fwdA, st->fwdK, st->inv})
        if (int r = fft_check(hipfftSetStream(h, s), "hipfftSetStream")) return r;
    if (int r = fft_check(hipfftExecR2C(st->fwdA, realA, reinterpret_cast<hipfftComplex*>(specA)), "image FFT"))
        return r;
    if (int r = fft_check(hipfftExecR2C(st->fwdK, realK, reinterpret_cast<hipfftComplex*>(specK)), "filter FFT"))
        return r;
    const long tO = (long)bins * c.P;
    hipLaunchKernelGGL(k_fft_spectrum_product, dim3(grid_for(tO)), dim3(256), 0, s, specA, specK, specO, (long)bins,
                       c.PA, c.PK, c.P, 1.0f / ((float)c.Er * (float)c.Ec));
    if (int r = check_hip(hipGetLastError(), "spectrum product launch")) return r;
    if (int r = fft_check(hipfftExecC2R(st->inv, reinterpret_cast<hipfftComplex*>(specO), realO), "inverse FFT"))
        return r;
    const long tOut = (long)g.H * g.W * c.P;
    hipLaunchKernelGGL(k_fft_trim, dim3(grid_for(tOut)), dim3(256), 0, s, realO, out, g.H, g.W, c.Er, c.Ec, tOut, e);
    return check_hip(hipGetLastError(), "FFT trim launch");
}

void* fft_state_new() { return new (std::nothrow) FftState; }

}  // namespace cogconv
