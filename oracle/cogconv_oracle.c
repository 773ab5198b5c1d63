/*
 * cogconv_oracle.c -- CPU restatement of cct-core's ConvolveOp semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * on the product path.  The product (libcogconv) never calls it.
 *
 * It restates, independently of the HIP kernels, what the reference computes:
 *
 *   - output field type and the legality checks of a ConvolveOp
 *       src/main/scala/cogx/compiler/codegenerator/opencl/hyperkernels/ConvolveHyperKernel.scala:1297-1450
 *       src/main/scala/cogx/compiler/parser/syntaxtree/BinaryOperator.scala:65-93
 *       src/main/scala/cogx/compiler/codegenerator/opencl/generator/DynamicConvolutionGenerator.scala:484-539
 *   - the scalar kernel S(x, k; border, orientation, sampling)
 *       ConvolveHyperKernel.scala:205-317 (tile origin, borders), 345-384 (filter flip),
 *       404-548 (multiply-add; column sums accumulated into a row sum), 824-830 (halo)
 *   - the plane maps of every VectorMode
 *       ConvolveHyperKernel.scala:170-199, 319-340, 575-645
 *   - blockReduceSum
 *       src/main/scala/cogx/compiler/codegenerator/opencl/hyperkernels/TensorReduceHyperKernel.scala:39-207
 *
 * The formulation here is deliberately different from the GPU kernels': the
 * up-sampled operand (operand 1, or operand 2 in FilterAdjoint modes) is first
 * materialised with explicit zeros, the filter is materialised flipped for the
 * convolution orientation, and the border is then applied to the materialised
 * operand.  The GPU kernels never materialise anything.
 *
 * Two precisions: FP64 accumulation (the accuracy reference), and FP32
 * accumulation in the reference kernel's order (for each filter row i a column
 * sum over j, added into a row sum; planes of a block reduction added in order).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_output_type  - golden-file shapes (wx/wty), closed-form shapes
 *   oracle_scalar       - 25-case golden set (wx = convolve, wty = crossCorrelate of the
 *                         adjoint), 18 extra golden convolutions, the closed-form
 *                         matrices of StaticConvolutionSpec, torch conv2d on the CPU,
 *                         brute force on tiny inputs, adjoint identities
 *   oracle_conv (plane maps) - the plane-decomposition identities of
 *                         StaticConvolutionSpec and torch conv2d / conv_transpose2d /
 *                         weight-gradient forms
 *   oracle_block_reduce_sum - brute force
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Numeric values shared with include/cogconv.h (the C-ABI uses the same codes). */
enum { OB_ZERO = 0, OB_CLAMP = 1, OB_CYCLIC = 2, OB_VALID = 3, OB_FULL = 4 };
enum { OS_NONE = 0, OS_UP = 1, OS_DOWN = 2 };
enum { OO_XCORR = 0, OO_CONV = 1 };
enum { OM_PBP = 0, OM_PF = 1, OM_PF_BRS = 2, OM_BPF = 3, OM_BPF_BRS = 4, OM_FA = 5, OM_FA_BRS = 6 };

typedef struct {
    int mode, border, orient, sampling, step, batch;
    int in0_dims, in0_rows, in0_cols, in0_planes, in0_order;
    int in1_dims, in1_rows, in1_cols, in1_planes, in1_order;
} oracle_desc;

typedef struct {
    int dims, rows, cols, planes, order;
} oracle_type;

/* Error codes (positive) returned by oracle_output_type. */
enum {
    OE_OK = 0,
    OE_DIMS = 1,          /* field dimension not 1 or 2, or unequal        */
    OE_BATCH = 2,         /* image depth not a multiple of batch           */
    OE_FILTER_DEPTH = 3,  /* filter depth not a multiple of planes/image   */
    OE_EVEN_FILTER = 4,   /* even filter with a centred border             */
    OE_DOWN_DIVIDE = 5,   /* downsample step does not divide the size      */
    OE_UP_CLAMP = 6,      /* integrated upsampling with BorderClamp        */
    OE_PBP_SHAPES = 7,    /* PlaneByPlane tensor shapes incompatible       */
    OE_FA_UP_BORDER = 8,  /* FilterAdjoint upsampling with a non-Valid border */
    OE_ARG = 9            /* out-of-range enum / non-positive size         */
};

static int is_fa(int mode) { return mode == OM_FA || mode == OM_FA_BRS; }

/* Up-sampling factors of operand 1 and operand 2, and the down-sampling step,
 * per axis (a 1-D field has a single column axis; rows stay 1). */
static void sampling_factors(const oracle_desc* d, int* upA_r, int* upA_c, int* upF_r, int* upF_c,
                             int* dn_r, int* dn_c) {
    int s = d->step;
    int two = d->in0_dims == 2;
    *upA_r = *upA_c = *upF_r = *upF_c = *dn_r = *dn_c = 1;
    if (d->sampling == OS_UP && s > 1) {
        if (is_fa(d->mode)) { *upF_c = s; if (two) *upF_r = s; }
        else                { *upA_c = s; if (two) *upA_r = s; }
    } else if (d->sampling == OS_DOWN && s > 1) {
        *dn_c = s; if (two) *dn_r = s;
    }
}

static int border_size(int border, int u, int f) {
    if (border == OB_VALID) return u - f + 1;
    if (border == OB_FULL) return u + f - 1;
    return u;
}

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

int oracle_output_type(const oracle_desc* d, oracle_type* out) {
    if (d->mode < 0 || d->mode > 6 || d->border < 0 || d->border > 4 || d->orient < 0 || d->orient > 1 ||
        d->sampling < 0 || d->sampling > 2 || d->batch < 1 || d->step < 1)
        return OE_ARG;
    if (d->in0_rows < 1 || d->in0_cols < 1 || d->in0_planes < 1 || d->in1_rows < 1 || d->in1_cols < 1 ||
        d->in1_planes < 1)
        return OE_ARG;
    /* BinaryOperator.scala:65-93 */
    if (d->in0_dims < 1 || d->in0_dims > 2 || d->in0_dims != d->in1_dims) return OE_DIMS;
    if (d->in0_dims == 1 && (d->in0_rows != 1 || d->in1_rows != 1)) return OE_ARG;
    int sampling = d->sampling, step = d->step;
    if (step == 1) sampling = OS_NONE; /* CogFunctionAPI.scala:1140-1144 */
    oracle_desc dd = *d;
    dd.sampling = sampling;
    if (d->in0_planes % d->batch) return OE_BATCH;
    int ppi = d->in0_planes / d->batch;
    int T = d->in1_planes;
    switch (d->mode) {
    case OM_PF: case OM_PF_BRS: case OM_BPF: case OM_BPF_BRS:
        if (T % ppi) return OE_FILTER_DEPTH;
        break;
    case OM_FA: case OM_FA_BRS:
        if (T % d->batch) return OE_BATCH;
        break;
    case OM_PBP:
        if (!(d->in0_order == 0 || d->in1_order == 0 || d->in0_planes == d->in1_planes)) return OE_PBP_SHAPES;
        break;
    }
    /* DynamicConvolutionGenerator.scala:500-506: odd filters unless Valid/Full */
    if (d->border != OB_VALID && d->border != OB_FULL) {
        if ((d->in1_cols % 2) == 0 || (d->in0_dims == 2 && (d->in1_rows % 2) == 0)) return OE_EVEN_FILTER;
    }
    /* L512-526: (inSize - contraction) % step == 0 for downsampling */
    if (sampling == OS_DOWN) {
        int cr = 0, cc = 0;
        if (d->border == OB_VALID) { cr = d->in1_rows - 1; cc = d->in1_cols - 1; }
        if (d->border == OB_FULL) { cr = 1 - d->in1_rows; cc = 1 - d->in1_cols; }
        if ((d->in0_cols - cc) % step) return OE_DOWN_DIVIDE;
        if (d->in0_dims == 2 && (d->in0_rows - cr) % step) return OE_DOWN_DIVIDE;
    }
    /* L531-536 */
    if (sampling == OS_UP && d->border == OB_CLAMP) return OE_UP_CLAMP;
    if (sampling == OS_UP && is_fa(d->mode) && d->border != OB_VALID) return OE_FA_UP_BORDER;

    int upA_r, upA_c, upF_r, upF_c, dn_r, dn_c;
    sampling_factors(&dd, &upA_r, &upA_c, &upF_r, &upF_c, &dn_r, &dn_c);
    int U_r = d->in0_rows * upA_r, U_c = d->in0_cols * upA_c;
    int F_r = d->in1_rows * upF_r, F_c = d->in1_cols * upF_c;
    int br = border_size(d->border, U_r, F_r), bc = border_size(d->border, U_c, F_c);
    if (d->in0_dims == 1) br = 1;
    if (br < 1 || bc < 1) return OE_ARG;
    out->dims = d->in0_dims;
    out->rows = ceil_div(br, dn_r);
    out->cols = ceil_div(bc, dn_c);

    int planes = 0, order = 1;
    switch (d->mode) {
    case OM_PF: case OM_BPF: planes = T * d->batch; order = 1; break;
    case OM_PF_BRS: case OM_BPF_BRS:
        planes = d->batch * (T / ppi);
        order = (planes == 1 && d->in0_planes <= 4) ? 0 : 1;
        break;
    case OM_FA: planes = ppi * (T / d->batch) * d->batch; order = planes > 1; break;
    case OM_FA_BRS: planes = ppi * (T / d->batch); order = planes > 1; break;
    case OM_PBP:
        if (d->in0_order == 0) { planes = d->in1_planes; order = d->in1_order; }
        else { planes = d->in0_planes; order = d->in0_order; }
        break;
    }
    out->planes = planes;
    out->order = order;
    return OE_OK;
}

/* ------------------------------------------------------------------------ */
/* Scalar kernel                                                             */
/* ------------------------------------------------------------------------ */

/* Materialise upsample(p, (ur, uc)) with zeros between samples (phase 0). */
static float* upsample_plane(const float* p, int rows, int cols, int ur, int uc) {
    int R = rows * ur, C = cols * uc;
    float* q = (float*)calloc((size_t)R * C, sizeof(float));
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) q[(size_t)(r * ur) * C + c * uc] = p[(size_t)r * cols + c];
    return q;
}

/* Value of the (materialised) operand at (y, x) under the border policy. */
static inline float border_read(const float* a, int U_r, int U_c, int y, int x, int border, int* inside) {
    if (border == OB_CLAMP) {
        if (y < 0) y = 0;
        if (y >= U_r) y = U_r - 1;
        if (x < 0) x = 0;
        if (x >= U_c) x = U_c - 1;
    } else if (border == OB_CYCLIC) {
        y %= U_r; if (y < 0) y += U_r;
        x %= U_c; if (x < 0) x += U_c;
    } else if (y < 0 || y >= U_r || x < 0 || x >= U_c) {
        *inside = 0;
        return 0.0f;
    }
    *inside = 1;
    return a[(size_t)y * U_c + x];
}

/*
 * S(a, f) for one plane pair, output rows R0 <= R < R1.  a is rows x cols, f
 * is frows x fcols (both before sampling).  upA/upF are the zero-insertion
 * factors of a / f, dn the output step.  out receives (R1 - R0) x ow elements,
 * double (prec 0) or float (prec 1).
 */
static void scalar_rows(const float* a, int rows, int cols, const float* f, int frows, int fcols, int border,
                        int orient, int upA_r, int upA_c, int upF_r, int upF_c, int dn_r, int dn_c, int R0, int R1,
                        int ow, int prec, void* out) {
    const float* A = a;
    float* Aup = NULL;
    if (upA_r > 1 || upA_c > 1) { Aup = upsample_plane(a, rows, cols, upA_r, upA_c); A = Aup; }
    int U_r = rows * upA_r, U_c = cols * upA_c;
    const float* Fm = f;
    float* Fup = NULL;
    if (upF_r > 1 || upF_c > 1) { Fup = upsample_plane(f, frows, fcols, upF_r, upF_c); Fm = Fup; }
    int F_r = frows * upF_r, F_c = fcols * upF_c;
    /* taps t[i][j] in cross-correlation form */
    float* t = (float*)malloc((size_t)F_r * F_c * sizeof(float));
    for (int i = 0; i < F_r; i++)
        for (int j = 0; j < F_c; j++)
            t[(size_t)i * F_c + j] =
                orient == OO_CONV ? Fm[(size_t)(F_r - 1 - i) * F_c + (F_c - 1 - j)] : Fm[(size_t)i * F_c + j];
    int h_r, h_c;
    if (border == OB_FULL) { h_r = F_r - 1; h_c = F_c - 1; }
    else if (border == OB_VALID) { h_r = 0; h_c = 0; }
    else { h_r = (F_r - 1) / 2; h_c = (F_c - 1) / 2; }
    for (int R = R0; R < R1; R++) {
        for (int C = 0; C < ow; C++) {
            int y0 = R * dn_r - h_r, x0 = C * dn_c - h_c;
            if (prec == 0) {
                double s = 0.0;
                for (int i = 0; i < F_r; i++)
                    for (int j = 0; j < F_c; j++) {
                        int in;
                        float v = border_read(A, U_r, U_c, y0 + i, x0 + j, border, &in);
                        s += (double)t[(size_t)i * F_c + j] * (double)v;
                    }
                ((double*)out)[(size_t)(R - R0) * ow + C] = s;
            } else {
                float rowSum = 0.0f;
                for (int i = 0; i < F_r; i++) {
                    float colSum = 0.0f;
                    for (int j = 0; j < F_c; j++) {
                        int in;
                        float v = border_read(A, U_r, U_c, y0 + i, x0 + j, border, &in);
                        colSum += v * t[(size_t)i * F_c + j];
                    }
                    rowSum += colSum;
                }
                ((float*)out)[(size_t)(R - R0) * ow + C] = rowSum;
            }
        }
    }
    free(t);
    free(Aup);
    free(Fup);
}

void oracle_scalar(const float* a, int rows, int cols, const float* f, int frows, int fcols, int border, int orient,
                   int upA_r, int upA_c, int upF_r, int upF_c, int dn_r, int dn_c, int oh, int ow, int prec,
                   void* out) {
    scalar_rows(a, rows, cols, f, frows, fcols, border, orient, upA_r, upA_c, upF_r, upF_c, dn_r, dn_c, 0, oh, ow,
                prec, out);
}

/* ------------------------------------------------------------------------ */
/* Plane maps (ConvolveHyperKernel.scala:170-199, 319-340, 575-645)           */
/* ------------------------------------------------------------------------ */

/* Number of plane pairs summed into one output plane. */
static int reduction_length(const oracle_desc* d) {
    int ppi = d->in0_planes / d->batch;
    switch (d->mode) {
    case OM_PF_BRS: return ppi;
    case OM_BPF_BRS: return ppi;
    case OM_FA_BRS: return d->batch;
    default: return 1;
    }
}

/* Operand planes of reduction term r of output plane o. */
void oracle_plane_pair(const oracle_desc* d, int o, int r, int* pa, int* pf) {
    int B = d->batch, N = d->in0_planes / B, T = d->in1_planes;
    switch (d->mode) {
    case OM_PBP:
        *pa = d->in0_planes == 1 ? 0 : o;
        *pf = d->in1_planes == 1 ? 0 : o;
        return;
    case OM_PF: { /* o = b*T + w */
        int b = o / T, w = o % T;
        *pa = b * N + w % N; *pf = w; return;
    }
    case OM_PF_BRS: { /* o = b*K + k; r = n */
        int K = T / N, b = o / K, k = o % K;
        *pa = b * N + r; *pf = k * N + r; return;
    }
    case OM_BPF: { /* o = b*T + w; M = N, L = T/M */
        int L = T / N, b = o / T, w = o % T;
        *pa = b * N + w % N; *pf = L * (w % N) + w / N; return;
    }
    case OM_BPF_BRS: { /* o = b*L + l; r = m */
        int L = T / N, b = o / L, l = o % L;
        *pa = b * N + r; *pf = r * L + l; return;
    }
    case OM_FA: { /* o = w*B + b (batch-interleaved) */
        int K = T / B, w = o / B, b = o % B;
        *pa = b * N + w % N; *pf = b * K + w / N; return;
    }
    case OM_FA_BRS: { /* o = w; r = b */
        int K = T / B, w = o, b = r;
        *pa = b * N + w % N; *pf = b * K + w / N; return;
    }
    }
}

/*
 * Full ConvolveOp on planar fields (addr(p, r, c) = p*rows*cols + r*cols + c).
 * prec 0: out is double, every term accumulated in double.
 * prec 1: out is float, per-plane S in float in the reference order, planes of a
 *         block reduction summed in float in order.
 * Returns an OE_* code.
 */
int oracle_conv(const oracle_desc* d, const float* in0, const float* in1, void* out, int prec, int nthreads) {
    oracle_type ot;
    int e = oracle_output_type(d, &ot);
    if (e) return e;
    oracle_desc dd = *d;
    if (dd.step == 1) dd.sampling = OS_NONE;
    int upA_r, upA_c, upF_r, upF_c, dn_r, dn_c;
    sampling_factors(&dd, &upA_r, &upA_c, &upF_r, &upF_c, &dn_r, &dn_c);
    size_t inA = (size_t)d->in0_rows * d->in0_cols, inF = (size_t)d->in1_rows * d->in1_cols;
    size_t op = (size_t)ot.rows * ot.cols;
    int R = reduction_length(d);
    /* Work items are (output plane, band of up to 8 output rows), so that a
     * few output planes still spread over many threads; every element sums its
     * reduction terms in the same order whatever the banding. */
    const int band = 8, bands = (ot.rows + band - 1) / band;
    const long items = (long)ot.planes * bands;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (long it = 0; it < items; it++) {
        const int o = (int)(it / bands), R0 = (int)(it % bands) * band;
        const int R1 = R0 + band < ot.rows ? R0 + band : ot.rows;
        const size_t n = (size_t)(R1 - R0) * ot.cols, off = (size_t)o * op + (size_t)R0 * ot.cols;
        void* tmp = malloc(n * (prec == 0 ? sizeof(double) : sizeof(float)));
        if (prec == 0) {
            double* acc = (double*)out + off;
            for (size_t q = 0; q < n; q++) acc[q] = 0.0;
            for (int r = 0; r < R; r++) {
                int pa, pf;
                oracle_plane_pair(d, o, r, &pa, &pf);
                scalar_rows(in0 + (size_t)pa * inA, d->in0_rows, d->in0_cols, in1 + (size_t)pf * inF, d->in1_rows,
                            d->in1_cols, d->border, d->orient, upA_r, upA_c, upF_r, upF_c, dn_r, dn_c, R0, R1,
                            ot.cols, 0, tmp);
                for (size_t q = 0; q < n; q++) acc[q] += ((double*)tmp)[q];
            }
        } else {
            float* acc = (float*)out + off;
            for (size_t q = 0; q < n; q++) acc[q] = 0.0f;
            for (int r = 0; r < R; r++) {
                int pa, pf;
                oracle_plane_pair(d, o, r, &pa, &pf);
                scalar_rows(in0 + (size_t)pa * inA, d->in0_rows, d->in0_cols, in1 + (size_t)pf * inF, d->in1_rows,
                            d->in1_cols, d->border, d->orient, upA_r, upA_c, upF_r, upF_c, dn_r, dn_c, R0, R1,
                            ot.cols, 1, tmp);
                for (size_t q = 0; q < n; q++) acc[q] += ((float*)tmp)[q];
            }
        }
        free(tmp);
    }
    return OE_OK;
}

/* blockReduceSum: out[p][q] = sum_{j<factor} in[p*factor + j][q] (planar). */
int oracle_block_reduce_sum(const float* in, int points, int planes, int factor, double* out) {
    if (factor < 1 || planes % factor) return OE_ARG;
    int op = planes / factor;
    for (int p = 0; p < op; p++)
        for (int q = 0; q < points; q++) {
            double s = 0.0;
            for (int j = 0; j < factor; j++) s += in[(size_t)(p * factor + j) * points + q];
            out[(size_t)p * points + q] = s;
        }
    return OE_OK;
}

/* Upsample one plane (zero insertion, phase 0): used by tests to build the
 * expected values of the reference's upsample-based decomposition identities. */
void oracle_upsample(const float* p, int rows, int cols, int ur, int uc, float* out) {
    float* q = upsample_plane(p, rows, cols, ur, uc);
    memcpy(out, q, (size_t)rows * ur * cols * uc * sizeof(float));
    free(q);
}

/*
 * Separable convolution of one 2-D plane (CogFunctionAPI.scala:1277-1316):
 * ConvolveRows2DHyperKernel.scala:73-117 (row pass, border on the column index)
 * then ConvolveColumns2DHyperKernel (column pass over the row-pass result,
 * border on the row index).  The tap i of a filter of length F reads offset
 * i - (F-1)/2; the convolution orientation reverses the filter.  Each pass is
 * materialised in full, in double (prec 0) or in float in tap order (prec 1).
 * border: OB_ZERO, OB_CLAMP or OB_CYCLIC.
 */
static double sep_read(const double* a, int rows, int cols, int y, int x, int border) {
    if (border == OB_CLAMP) {
        y = y < 0 ? 0 : (y >= rows ? rows - 1 : y);
        x = x < 0 ? 0 : (x >= cols ? cols - 1 : x);
    } else if (border == OB_CYCLIC) {
        y %= rows; if (y < 0) y += rows;
        x %= cols; if (x < 0) x += cols;
    } else if (y < 0 || y >= rows || x < 0 || x >= cols) {
        return 0.0;
    }
    return a[(size_t)y * cols + x];
}

int oracle_separable(const float* in, int rows, int cols, const float* rf, int fr, const float* cf, int fc, int border,
                     int orient, int prec, double* out) {
    if (fr < 1 || fc < 1 || fr % 2 == 0 || fc % 2 == 0) return OE_EVEN_FILTER;
    if (border != OB_ZERO && border != OB_CLAMP && border != OB_CYCLIC) return OE_ARG;
    size_t n = (size_t)rows * cols;
    double* a = (double*)malloc(n * sizeof(double));
    double* mid = (double*)malloc(n * sizeof(double));
    for (size_t q = 0; q < n; q++) a[q] = in[q];
    int hr = (fr - 1) / 2, hc = (fc - 1) / 2;
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            double sd = 0.0;
            float sf = 0.0f;
            for (int i = 0; i < fr; i++) {
                float t = rf[orient == OO_CONV ? fr - 1 - i : i];
                double v = sep_read(a, rows, cols, r, c - hr + i, border);
                sd += (double)t * v;
                sf += (float)v * t;
            }
            mid[(size_t)r * cols + c] = prec == 0 ? sd : (double)sf;
        }
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            double sd = 0.0;
            float sf = 0.0f;
            for (int j = 0; j < fc; j++) {
                float t = cf[orient == OO_CONV ? fc - 1 - j : j];
                double v = sep_read(mid, rows, cols, r - hc + j, c, border);
                sd += (double)t * v;
                sf += (float)v * t;
            }
            out[(size_t)r * cols + c] = prec == 0 ? sd : (double)sf;
        }
    free(a);
    free(mid);
    return OE_OK;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
