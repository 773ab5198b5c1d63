// C++ test of the C ABI (include/cogconv.h), linked against libcogconv.so and
// called directly, without the Python binding.  Host-only calls: it runs
// without a GPU.  Built and run by tests/test_capi_cpp.py; prints one line per
// check and exits non-zero on the first failure.
#include <cogconv.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

static int checks = 0;

#define CHECK(cond)                                                                              \
    do {                                                                                         \
        ++checks;                                                                                \
        if (!(cond)) {                                                                           \
            std::fprintf(stderr, "FAIL %s:%d: %s (last error: %s)\n", __FILE__, __LINE__, #cond, \
                         cogconv_last_error());                                                  \
            return 1;                                                                            \
        }                                                                                        \
    } while (0)

static cogconv_field_type ft(int dims, int rows, int cols, int planes) {
    cogconv_field_type t;
    t.dims = dims;
    t.rows = rows;
    t.cols = cols;
    t.planes = planes;
    t.tensor_order = planes == 1 ? 0 : 1;
    return t;
}

static cogconv_desc desc(int mode, int border, int orient, int sampling, int step, int batch, cogconv_field_type a,
                         cogconv_field_type b) {
    cogconv_desc d;
    std::memset(&d, 0, sizeof d);
    d.mode = mode;
    d.border = border;
    d.orientation = orient;
    d.sampling = sampling;
    d.step = step;
    d.batch = batch;
    d.in0 = a;
    d.in1 = b;
    d.precision = COGCONV_FP32;
    d.algo = COGCONV_ALGO_AUTO;
    return d;
}

int main() {
    CHECK(cogconv_abi_version() == COGCONV_ABI_VERSION);
    // Output types (ConvolveHyperKernel.outputFieldType): c2 ProjectFrame+BRS, BorderValid.
    cogconv_field_type out;
    cogconv_desc pf = desc(COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM, COGCONV_BORDER_VALID, COGCONV_CROSS_CORRELATION,
                           COGCONV_NO_SAMPLING, 1, 16, ft(2, 64, 64, 16 * 16), ft(2, 3, 3, 32 * 16));
    CHECK(cogconv_output_type(&pf, &out) == COGCONV_OK);
    CHECK(out.rows == 62 && out.cols == 62 && out.planes == 16 * 32);
    // Down-sampling by 2 with BorderZero: ceil(64 / 2).
    cogconv_desc dn = pf;
    dn.border = COGCONV_BORDER_ZERO;
    dn.sampling = COGCONV_DOWNSAMPLE_OUTPUT;
    dn.step = 2;
    CHECK(cogconv_output_type(&dn, &out) == COGCONV_OK && out.rows == 32 && out.cols == 32);
    // The reference's rejections: even filter with a centred border, up-sampling with BorderClamp.
    cogconv_desc even = desc(COGCONV_PLANE_BY_PLANE, COGCONV_BORDER_ZERO, COGCONV_CONVOLUTION, COGCONV_NO_SAMPLING, 1,
                             1, ft(2, 16, 16, 1), ft(2, 4, 4, 1));
    CHECK(cogconv_output_type(&even, &out) == COGCONV_ERR_ILLEGAL_OP);
    CHECK(std::strlen(cogconv_last_error()) > 0);
    cogconv_desc upclamp = desc(COGCONV_PLANE_BY_PLANE, COGCONV_BORDER_CLAMP, COGCONV_CONVOLUTION,
                                COGCONV_UPSAMPLE_INPUT, 2, 1, ft(2, 16, 16, 1), ft(2, 3, 3, 1));
    CHECK(cogconv_output_type(&upclamp, &out) == COGCONV_ERR_ILLEGAL_OP);
    // Plans are host objects: create, query, describe, destroy.
    cogconv_plan plan = nullptr;
    CHECK(cogconv_plan_create(&pf, &plan) == COGCONV_OK && plan);
    double flops = 0, bytes = 0;
    CHECK(cogconv_plan_flops(plan, &flops) == COGCONV_OK && flops == 566820864.0);
    CHECK(cogconv_plan_bytes(plan, &bytes) == COGCONV_OK && bytes == 12085248.0);
    char name[256];
    CHECK(cogconv_plan_describe(plan, name, sizeof name) == COGCONV_OK && std::strstr(name, "igemm_f32"));
    int32_t tab = 0;
    CHECK(cogconv_plan_table_supported(plan, &tab) == COGCONV_OK && tab == 1);
    int32_t nv = 0;
    CHECK(cogconv_plan_variant_count(&pf, &nv) == COGCONV_OK && nv > 1);
    // Epilogue checks against the output type.
    cogconv_epilogue epi;
    std::memset(&epi, 0, sizeof epi);
    epi.n_ops = 2;
    epi.ops[0].op = COGCONV_EPI_ADD_FIELD;
    epi.ops[0].operand = 0;
    epi.ops[1].op = COGCONV_EPI_MAX_CONST;
    epi.ops[1].value = 0.0f;
    epi.n_operands = 1;
    epi.operands[0] = ft(0, 1, 1, 16 * 32);  // a 0-D field, one value per output plane
    epi.operands[0].tensor_order = 1;
    CHECK(cogconv_plan_set_epilogue(plan, &epi) == COGCONV_OK);
    epi.operands[0] = ft(2, 62, 61, 16 * 32);
    CHECK(cogconv_plan_set_epilogue(plan, &epi) == COGCONV_ERR_ILLEGAL_OP);
    // A launch refuses null buffers before touching a device.
    CHECK(cogconv_launch(plan, nullptr, nullptr, nullptr, nullptr, nullptr) == COGCONV_ERR_INVALID_ARG);
    CHECK(cogconv_plan_destroy(plan) == COGCONV_OK);
    // Separable: its checks come before any device work.
    float dummy[4] = {0, 0, 0, 0};
    cogconv_field_type v = ft(2, 8, 8, 2);
    CHECK(cogconv_separable(dummy, &v, dummy, 3, dummy, 3, COGCONV_BORDER_ZERO, COGCONV_CROSS_CORRELATION, dummy,
                            nullptr, nullptr, nullptr) == COGCONV_ERR_ILLEGAL_OP);
    // A missing graph file fails cleanly.
    cogconv_function fn = nullptr;
    CHECK(cogconv_function_create("/nonexistent/graph.cgraph", 0, &fn) == COGCONV_ERR_INVALID_ARG && !fn);
    std::printf("capi_host_test: %d checks passed\n", checks);
    return 0;
}
