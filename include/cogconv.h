/*
 * cogconv.h -- C ABI of the MI355X-native Field-convolution path (libcogconv.so).
 *
 * Drop-in boundary for cct-core's ConvolveOp path.  The reference reaches the
 * device through JOCL (OpenCL over JNI); this library replaces both
 *   (a) the OpenCL platform layer (device / buffer / command-queue / event
 *       management, src/main/scala/cogx/platform/opencl/), and
 *   (b) the convolution HyperKernels (ConvolveHyperKernel,
 *       ConvolveTiledHyperKernel, ConvolveToSmallField*HyperKernel,
 *       TensorReduceHyperKernel in
 *       src/main/scala/cogx/compiler/codegenerator/opencl/hyperkernels/).
 *
 * Conventions (follow src/main/java/cogjni/CogJNIWrapper.java:29-33,72-85):
 *   - every function returns int status, 0 = COGCONV_OK;
 *   - on failure cogconv_last_error() returns a thread-local message; no C++
 *     exception or abort ever crosses this ABI;
 *   - handles are opaque pointers (a Java caller holds them as `long`);
 *   - the caller owns buffers, streams, events and plans;
 *   - launches are asynchronous (they return after enqueue).
 *
 * Data layout of every field buffer: planar float32,
 *   addr(plane, row, col) = plane*rows*cols + row*cols + col
 * (src/main/scala/cogx/platform/types/FieldMemoryLayoutImpl.scala:250-311,
 *  src/main/scala/cogx/compiler/codegenerator/opencl/fragments/FieldIO.scala:308-343).
 * A batch of B images of N planes stores image b's planes at b*N .. b*N+N-1.
 * A 1-D field of length L is rows = 1, cols = L, dims = 1.
 */
#ifndef COGCONV_H
#define COGCONV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COGCONV_ABI_VERSION 1

/* Status codes. */
enum {
    COGCONV_OK = 0,
    COGCONV_ERR_INVALID_ARG = 1, /* null pointer, bad enum, bad size            */
    COGCONV_ERR_ILLEGAL_OP = 2,  /* the reference rejects this ConvolveOp       */
    COGCONV_ERR_RUNTIME = 3,     /* HIP runtime error                           */
    COGCONV_ERR_OOM = 4,         /* device allocation failed                    */
    COGCONV_ERR_NOT_READY = 5,   /* event / request not complete yet            */
    COGCONV_ERR_UNSUPPORTED = 6, /* legal op, but not offered in this precision */
    COGCONV_ERR_BUDGET = 7       /* serving admission budget exceeded           */
};

/* VectorMode: src/main/scala/cogx/compiler/parser/op/BinaryOpcode.scala:39-48 */
enum {
    COGCONV_PLANE_BY_PLANE = 0,
    COGCONV_PROJECT_FRAME = 1,
    COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM = 2,
    COGCONV_BACK_PROJECT_FRAME = 3,
    COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM = 4,
    COGCONV_FILTER_ADJOINT = 5,
    COGCONV_FILTER_ADJOINT_BLOCK_REDUCE_SUM = 6
};

/* BorderPolicy: src/main/scala/cogx/platform/types/BorderPolicy.scala:30-134 */
enum {
    COGCONV_BORDER_ZERO = 0,
    COGCONV_BORDER_CLAMP = 1,
    COGCONV_BORDER_CYCLIC = 2,
    COGCONV_BORDER_VALID = 3,
    COGCONV_BORDER_FULL = 4
};

/* ConvolutionSamplingPolicy: src/main/scala/cogx/platform/types/ConvolutionSamplingPolicy.scala:27-31 */
enum {
    COGCONV_NO_SAMPLING = 0,
    COGCONV_UPSAMPLE_INPUT = 1,   /* UpsampleInputConvolution(step)   */
    COGCONV_DOWNSAMPLE_OUTPUT = 2 /* DownsampleOutputConvolution(step) */
};

/* FilterOrientation: src/main/scala/cogx/platform/types/FilterOrientation.scala:28-30 */
enum {
    COGCONV_CROSS_CORRELATION = 0,
    COGCONV_CONVOLUTION = 1
};

/* Arithmetic of the plan.  FP32 is the plain form (FP32 products, FP32
 * accumulation).  The others are the precision variants of the
 * contraction modes (ProjectFrame / BackProjectFrame / FilterAdjoint with
 * block reduction): MFMA implicit GEMM with FP32 accumulation. */
enum {
    COGCONV_FP32 = 0,
    COGCONV_BF16 = 1,
    COGCONV_FP16 = 2,
    COGCONV_SPLIT_BF16 = 3 /* three-term x_hi*w_hi + x_hi*w_lo + x_lo*w_hi */
};

/* Kernel-family hint; COGCONV_ALGO_AUTO lets the plan (and its autotuner) pick. */
enum {
    COGCONV_ALGO_AUTO = 0,
    COGCONV_ALGO_DIRECT = 1,        /* LDS-tiled direct kernel, any op         */
    COGCONV_ALGO_SMALL_OUTPUT = 2,  /* split-reduction kernel, small outputs   */
    COGCONV_ALGO_IMPLICIT_GEMM = 3  /* MFMA implicit GEMM, contraction modes   */
};

/* FieldType of one operand / result (src/main/scala/cogx/platform/types/FieldType.scala).
 * tensor_order is 0 for a ScalarField, 1 for a VectorField (planes = vector length). */
typedef struct {
    int32_t dims;   /* 1 or 2 */
    int32_t rows;   /* 1 for dims == 1 */
    int32_t cols;
    int32_t planes; /* tensor points */
    int32_t tensor_order;
} cogconv_field_type;

/* ConvolveOp(border, orientation, sampling, vectorMode, batchSize) plus the two
 * input FieldTypes: BinaryOpcode.scala:64-71, BinaryOperator.scala:59-93. */
typedef struct {
    int32_t mode;        /* COGCONV_PLANE_BY_PLANE .. */
    int32_t border;      /* COGCONV_BORDER_* */
    int32_t orientation; /* COGCONV_CROSS_CORRELATION / COGCONV_CONVOLUTION */
    int32_t sampling;    /* COGCONV_NO_SAMPLING / _UPSAMPLE_INPUT / _DOWNSAMPLE_OUTPUT */
    int32_t step;        /* sampling step; 1 means no sampling */
    int32_t batch;       /* batchSize */
    cogconv_field_type in0; /* image (or dy for BackProjectFrame, x for FilterAdjoint) */
    cogconv_field_type in1; /* filter bank (or dy for FilterAdjoint) */
    int32_t precision;   /* COGCONV_FP32 .. */
    int32_t algo;        /* COGCONV_ALGO_* */
} cogconv_desc;

typedef struct cogconv_plan_s* cogconv_plan;

/* One image of a table launch: its operand-1 planes and its output planes. */
typedef struct {
    const float* in0;
    float* out;
} cogconv_image_io;
typedef struct cogconv_stream_s* cogconv_stream; /* NULL = the device's default stream */
typedef struct cogconv_event_s* cogconv_event;

typedef struct {
    char name[128];
    char arch[64];           /* e.g. "gfx950" */
    int32_t compute_units;
    int32_t wavefront_size;
    int32_t lds_bytes_per_block;
    int32_t max_threads_per_block;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
    int32_t l2_bytes;
    uint64_t hbm_bytes;
} cogconv_device_props;

/* ---- errors and library state (replaces OpenCLPlatform, $OCL/OpenCLPlatform.scala:38-244) ---- */

/* Thread-local message of the last failure on this thread ("" if none). */
const char* cogconv_last_error(void);
int cogconv_abi_version(void);
/* Select `device` as current for this thread; idempotent. */
int cogconv_init(int32_t device);
int cogconv_shutdown(void);
int cogconv_device_count(int32_t* count);
/* OpenCLDevice properties ($OCL/OpenCLDevice.scala:39-402). */
int cogconv_device_props_get(int32_t device, cogconv_device_props* props);
int cogconv_device_synchronize(void);

/* ---- buffers (replaces OpenCLBuffer, $OCL/OpenCLBuffer.scala:53-523) ---- */

int cogconv_buf_alloc(size_t bytes, void** dptr);
int cogconv_buf_free(void* dptr);
/* Blocking copies (OpenCLBuffer.read/write, L207-231). */
int cogconv_h2d(void* dst, const void* src, size_t bytes);
int cogconv_d2h(void* dst, const void* src, size_t bytes);
/* Asynchronous copies on a stream; host memory should come from cogconv_host_alloc. */
int cogconv_h2d_async(void* dst, const void* src, size_t bytes, cogconv_stream stream);
int cogconv_d2h_async(void* dst, const void* src, size_t bytes, cogconv_stream stream);
int cogconv_d2d_async(void* dst, const void* src, size_t bytes, cogconv_stream stream);
/* Zero-fill by a kernel (safe inside captured graphs). */
int cogconv_fill_zero(void* dptr, size_t bytes, cogconv_stream stream);
/* Pinned host memory (the reference's Cog.pinnedBuffers option). */
int cogconv_host_alloc(size_t bytes, void** hptr);
int cogconv_host_free(void* hptr);

/* ---- streams and events (replaces OpenCLParallelCommandQueue, $OCL/OpenCLParallelCommandQueue.scala:45-283) ---- */

int cogconv_stream_create(cogconv_stream* stream);
int cogconv_stream_destroy(cogconv_stream stream);
int cogconv_stream_synchronize(cogconv_stream stream);
int cogconv_event_create(cogconv_event* event);
int cogconv_event_destroy(cogconv_event event);
int cogconv_event_record(cogconv_event event, cogconv_stream stream);
/* Make `stream` wait for `event` (the wait list of a launch, OpenCLDeviceKernel.scala:67-100). */
int cogconv_stream_wait_event(cogconv_stream stream, cogconv_event event);
/* COGCONV_OK when complete, COGCONV_ERR_NOT_READY when pending (phase2Clock's status check). */
int cogconv_event_query(cogconv_event event);
int cogconv_event_synchronize(cogconv_event event);
int cogconv_event_elapsed_ms(cogconv_event start, cogconv_event end, float* ms);

/* ---- the convolution (replaces ConvolveHyperKernel & co.) ---- */

/* Result FieldType of a ConvolveOp, with every legality check the reference
 * applies when the graph is built: ConvolveHyperKernel.outputFieldType
 * ($HK/ConvolveHyperKernel.scala:1297-1450), BinaryOperator.scala:65-93 and
 * DynamicConvolutionGenerator.convolveDirectly (L484-539).  Pure host code:
 * needs no device.  Illegal ops return COGCONV_ERR_ILLEGAL_OP. */
int cogconv_output_type(const cogconv_desc* desc, cogconv_field_type* out);

/* Build an execution plan (kernel choice, tile shape, workspace size).  Replaces
 * ConvolveHyperKernel.apply ($HK/ConvolveHyperKernel.scala:1003-1148).  Needs no
 * device either: kernels are loaded at the first launch. */
int cogconv_plan_create(const cogconv_desc* desc, cogconv_plan* plan);
int cogconv_plan_destroy(cogconv_plan plan);
int cogconv_plan_output_type(cogconv_plan plan, cogconv_field_type* out);
/* Device scratch a launch needs (0 for most plans); the caller allocates it. */
int cogconv_plan_workspace_bytes(cogconv_plan plan, size_t* bytes);
/* Algorithmic FLOPs of one launch: 2 * (multiply-adds with a non-zero tap). */
int cogconv_plan_flops(cogconv_plan plan, double* flops);
/* Algorithmic HBM bytes of one launch: each operand read once, output written once. */
int cogconv_plan_bytes(cogconv_plan plan, double* bytes);
/* Human-readable kernel choice, e.g. "direct tile=16x64 fchunk=16x16". */
int cogconv_plan_describe(cogconv_plan plan, char* buf, size_t len);
/* Enqueue one ConvolveOp: out = op(in0, in1).  in0/in1/out are planar float32
 * device buffers of the plan's FieldTypes; `workspace` holds
 * cogconv_plan_workspace_bytes bytes (may be NULL when that is 0). */
int cogconv_launch(cogconv_plan plan, const float* in0, const float* in1, float* out, void* workspace,
                   cogconv_stream stream);

/* ---- elementwise epilogue (merging of elementwise sinks into the convolution) ----
 * Replaces the kernel merger's forward merge of a convolution into its elementwise
 * sink (HyperKernel.canMergeWithSink, $FR/HyperKernel.scala:478-530, applied by
 * src/main/scala/cogx/compiler/optimizer/HyperKernelMerger.scala under
 * Cog.localMemoryMerging, $M/parameters/Cog.scala:72).  The epilogue's ops run in
 * order on each output element before its store, so a conv -> bias -> nonlinearity
 * chain writes the output once.  Ops follow the reference's elementwise operators:
 * Field +,-,*,/ with a Float or a Field (Field.scala:180-390), max/min with a Float
 * or a Field, abs, exp, log, sq, sqrt, tanh, signum, unary minus
 * (CogFunctionAPI.scala:109-303, 1574-1590).  R* ops take the constant or field
 * as the left operand (c - x, c / x).  Field operands follow the reference's
 * broadcast rule (src/main/scala/cogx/compiler/parser/semantics/ResultType.scala:29-58):
 * the output's own type, a ScalarField of its field shape (broadcast over planes),
 * a 0-D field with its planes (one value per plane), or a 0-D scalar. */
enum {
    COGCONV_EPI_ADD_CONST = 1,
    COGCONV_EPI_SUB_CONST = 2,
    COGCONV_EPI_RSUB_CONST = 3,
    COGCONV_EPI_MUL_CONST = 4,
    COGCONV_EPI_DIV_CONST = 5,
    COGCONV_EPI_RDIV_CONST = 6,
    COGCONV_EPI_MAX_CONST = 7,
    COGCONV_EPI_MIN_CONST = 8,
    COGCONV_EPI_NEG = 16,
    COGCONV_EPI_ABS = 17,
    COGCONV_EPI_SQ = 18,
    COGCONV_EPI_SQRT = 19,
    COGCONV_EPI_EXP = 20,
    COGCONV_EPI_LOG = 21,
    COGCONV_EPI_TANH = 22,
    COGCONV_EPI_SIGNUM = 23,
    COGCONV_EPI_ADD_FIELD = 32,
    COGCONV_EPI_SUB_FIELD = 33,
    COGCONV_EPI_RSUB_FIELD = 34,
    COGCONV_EPI_MUL_FIELD = 35,
    COGCONV_EPI_DIV_FIELD = 36,
    COGCONV_EPI_RDIV_FIELD = 37,
    COGCONV_EPI_MAX_FIELD = 38,
    COGCONV_EPI_MIN_FIELD = 39
};
#define COGCONV_EPI_MAX_OPS 8
#define COGCONV_EPI_MAX_OPERANDS 4

typedef struct {
    int32_t op;      /* COGCONV_EPI_* */
    float value;     /* constant of a *_CONST op */
    int32_t operand; /* operand slot of a *_FIELD op */
} cogconv_epilogue_op;

typedef struct {
    int32_t n_ops;
    cogconv_epilogue_op ops[COGCONV_EPI_MAX_OPS];
    int32_t n_operands;
    cogconv_field_type operands[COGCONV_EPI_MAX_OPERANDS];
} cogconv_epilogue;

/* Attach an epilogue to a plan (NULL or n_ops == 0 removes it).  Checks every op
 * and operand type against the plan's output type (COGCONV_ERR_ILLEGAL_OP when
 * the reference's broadcast rule does not give the output type). */
int cogconv_plan_set_epilogue(cogconv_plan plan, const cogconv_epilogue* epi);
/* cogconv_launch for a plan with an epilogue: operands[i] is the device buffer of
 * operand slot i.  cogconv_launch refuses a plan whose epilogue reads fields. */
int cogconv_launch_epilogue(cogconv_plan plan, const float* in0, const float* in1, float* out, void* workspace,
                            const float* const* operands, cogconv_stream stream);
/* The same ops as their own pass: out = epilogue(in) over a field of `type`
 * (in == out allowed).  The unmerged form of an elementwise sink. */
int cogconv_elementwise(const float* in, float* out, const cogconv_field_type* type, const cogconv_epilogue* epi,
                        const float* const* operands, cogconv_stream stream);

/* Separable convolution (CogFunctionAPI.crossCorrelateSeparable / convolveSeparable,
 * src/main/scala/cogx/api/CogFunctionAPI.scala:1277-1356; Field.scala:1500-1531):
 * a 2-D ScalarField, an odd 1-D row filter, then an odd 1-D column filter, with
 * BorderClamp, BorderZero or BorderCyclic, in one kernel (the reference runs
 * ConvolveRows2DHyperKernel then ConvolveColumns2DHyperKernel).  Filters of up to
 * 63 taps each; `epi` (may be NULL) is applied at the stores. */
int cogconv_separable(const float* in, const cogconv_field_type* type, const float* row_filter, int32_t row_taps,
                      const float* column_filter, int32_t column_taps, int32_t border, int32_t orientation, float* out,
                      const cogconv_epilogue* epi, const float* const* operands, cogconv_stream stream);

/* Resident operand 2 (no counterpart in the reference: for callers that apply one
 * filter bank to many inputs, such as the serving loop).  The BF16 / FP16 /
 * split-BF16 frame plans convert operand 2 into a packed low-precision form in
 * the workspace on every cogconv_launch.  cogconv_plan_pack_operand does that
 * conversion once (*packed = 1; 0 for plans that do not pack, which leave the
 * workspace alone), and cogconv_launch_packed then runs the convolution on the
 * packed form kept in `workspace` (COGCONV_ERR_INVALID_ARG for plans that do
 * not pack).  The workspace must not be written between the two. */
int cogconv_plan_pack_operand(cogconv_plan plan, const float* in1, void* workspace, cogconv_stream stream,
                              int32_t* packed);
int cogconv_launch_packed(cogconv_plan plan, const float* in0, float* out, void* workspace, cogconv_stream stream);

/* No-allocation launch.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Runs cogconv_launch with the plan's workspace
 * taken from a device array the library holds on each device until the process
 * ends (cogconv_noalloc_scratch_bytes gives its size, 64 MiB), so no call
 * allocates or frees memory.  A plan needing more workspace than that fails with
 * COGCONV_ERR_UNSUPPORTED.  All calls on one device share the array: calls that
 * overlap on a device (several threads, or several streams with work in flight)
 * race, so the caller serialises them. */
int cogconv_noalloc_scratch_bytes(size_t* bytes);
int cogconv_launch_noalloc(cogconv_plan plan, const float* in0, const float* in1, float* out, cogconv_stream stream);

/* Request-table launch (no counterpart in the reference: for the serving loop,
 * which stacks requests of one shape whose fields live at unrelated device
 * addresses).  The plan's batch B images read their operand-1 planes from
 * host_table[b].in0 and write their output planes to host_table[b].out instead of
 * consecutive planes of one buffer.  The table is copied to `device_table` (B
 * entries of device memory) on `stream` before the launch; with pinned host memory
 * that copy is asynchronous and host_table must stay unchanged until the stream
 * reaches it.  packed = 1 runs on an operand 2 packed earlier into `workspace`
 * (cogconv_plan_pack_operand; in1 unused).  Offered for the MFMA frame plans
 * without up-sampling (cogconv_plan_table_supported), otherwise
 * COGCONV_ERR_UNSUPPORTED. */
int cogconv_plan_table_supported(cogconv_plan plan, int32_t* supported);
int cogconv_launch_table(cogconv_plan plan, const cogconv_image_io* host_table, cogconv_image_io* device_table,
                         const float* in1, void* workspace, int32_t packed, cogconv_stream stream);

/* ---- autotuning ----
 * Counterpart of Profiler.bestVariant ($M/runtime/execution/Profiler.scala:131-190) and
 * ProfilerCache ($M/runtime/execution/ProfilerCache.scala:39-122).  Every plan has a list of
 * candidate kernel configurations (kernel family, output block, plane chunk, split count),
 * the heuristic choice first.  cogconv_autotune times each candidate on synthetic inputs
 * (input i of n filled with (i+1)/n, as the reference's profiler drives its circuits),
 * writing a flush buffer before every timed step, picks the lowest average step time and
 * records it per device and descriptor; plans built with COGCONV_ALGO_AUTO then use it.
 * The cache persists to a text file (save / load).  Tuning allocates device memory and
 * synchronizes `stream`: call it at graph-build time, not in a serving loop. */
typedef struct cogconv_tune_options {
    int32_t warmup_steps; /* untimed launches per candidate (Cog.profilerWarmupSteps); default 1   */
    int32_t run_steps;    /* timed launches per candidate (Cog.profilerSteps); default 5            */
    int64_t flush_bytes;  /* bytes written before each timed launch (Cog.profilerCacheFlushSizeMB);
                             default 512 MiB, beyond the last-level cache                          */
    int32_t force;        /* time again even if the cache holds a usable sample (Cog.forceProfiling) */
} cogconv_tune_options;

typedef struct cogconv_tune_result {
    int32_t variant;       /* chosen candidate */
    int32_t variants;      /* candidates considered */
    double avg_us;         /* the chosen candidate's average launch time */
    int32_t from_cache;    /* 1: an earlier sample was reused */
    char description[192]; /* cogconv_plan_describe of the chosen candidate */
} cogconv_tune_result;

int cogconv_plan_variant_count(const cogconv_desc* desc, int32_t* count);
/* A plan fixed to candidate `variant` (0 = the heuristic choice). */
int cogconv_plan_create_variant(const cogconv_desc* desc, int32_t variant, cogconv_plan* plan);
int cogconv_plan_variant(cogconv_plan plan, int32_t* variant);
/* `options` may be NULL (defaults); `stream` may be NULL (default stream). */
int cogconv_autotune(const cogconv_desc* desc, const cogconv_tune_options* options, cogconv_stream stream,
                     cogconv_tune_result* result);
int cogconv_tune_cache_clear(void);
int cogconv_tune_cache_save(const char* path);
/* Merges a saved cache; a missing file is not an error. */
int cogconv_tune_cache_load(const char* path);

/* blockReduceSum(field, factor) ($M/api/CogFunctionAPI.scala:755-756,
 * $HK/TensorReduceHyperKernel.scala:39-207): out plane p = sum_{j<factor} in plane p*factor+j.
 * `points` = rows*cols of one plane. */
int cogconv_block_reduce_sum(const float* in, float* out, int64_t points, int32_t in_planes, int32_t factor,
                             cogconv_stream stream);
/* blockReduceSum with an elementwise epilogue on its stores (the reduce kernel
 * merged with its elementwise sinks, as cogconv_plan_set_epilogue does for a
 * convolution); `in_type` is the input field, the result has planes/factor planes. */
int cogconv_block_reduce_sum_epilogue(const float* in, float* out, const cogconv_field_type* in_type, int32_t factor,
                                      const cogconv_epilogue* epi, const float* const* operands,
                                      cogconv_stream stream);

/* ---- compiled graphs: checkpoint / restore and the evaluate() Function Interface ----
 * A compiled graph file holds a convolution graph after its optimizers ran: the
 * fields (types, constant data, which are inputs), and per launch step its
 * descriptor (ConvolveOp, blockReduceSum factor, separable filters, elementwise
 * epilogue), in dependence order.  It replaces the reference's checkpoint of the
 * kernel circuit with OpenCL source (ComputeGraph.write / readFromFile,
 * $M/runtime/ComputeGraph.scala:523-666, OpenCLRestoredDeviceKernel.scala:28-60):
 * restoring re-plans every step for the device (cogconv_plan_create, autotuned
 * choices included) instead of re-running saved kernel code.  The Python
 * ComputeGraph.write() produces the file.
 *
 * The function calls follow the reference's CogFunction interface
 * (src/main/java/cogjni/CogJNIWrapper.java:85-161): cogconv_function_create
 * (cogFunctionCreateComputeFunction), input / output counts, names and
 * FieldTypes, cogconv_function_evaluate (cogFunctionEvaluate2D: one host buffer
 * per input and per output, planar float32), optimal thread count, name, and
 * destroy (cogFunctionShutdown).  evaluate() is reentrant: any number of host
 * threads may call it at once on one function; each call runs on an execution
 * context of its own (stream and device buffers, at most 4 made, later calls
 * wait for a free one), and returns when its outputs are in host memory.  Results
 * are the same whichever threads run beside it. */
typedef struct cogconv_function_s* cogconv_function;

int cogconv_function_create(const char* path, int32_t device, cogconv_function* function);
int cogconv_function_destroy(cogconv_function function);
int cogconv_function_name(cogconv_function function, char* buf, size_t len);
int cogconv_function_optimal_thread_count(cogconv_function function, int32_t* count);
int cogconv_function_input_count(cogconv_function function, int32_t* count);
int cogconv_function_output_count(cogconv_function function, int32_t* count);
/* Name and FieldType of input (output = 0) or output (output = 1) `index`. */
int cogconv_function_io(cogconv_function function, int32_t output, int32_t index, char* name, size_t name_len,
                        cogconv_field_type* type);
int cogconv_function_evaluate(cogconv_function function, const float* const* inputs, float* const* outputs);

/* ---- serving loop ----
 * No counterpart inside the reference's convolution path: the reference specifies a
 * reentrant evaluate() Function Interface (doc/CogFunctionInterfaceERS.docx,
 * src/main/java/cogjni/CogJNIWrapper.java:85-115) but has no request scheduler.  The
 * server keeps filter banks resident in HBM, admits requests against a device-memory
 * budget as request slots free up, stacks requests of one shape bucket (same op, same
 * bank, same field shape) into one launch when the mode allows (batch folded into
 * planes), spreads work over several streams so that transfers and kernels of
 * different requests overlap, and reports latency percentiles. */

typedef struct cogconv_server_s* cogconv_server;

typedef struct {
    int32_t device;
    uint64_t hbm_budget_bytes; /* device memory for request buffers; preallocated at create */
    int32_t max_inflight;      /* launches on the device at once */
    int32_t streams;           /* HIP streams work is spread over */
    int32_t max_stack;         /* requests stacked into one launch (1 = no stacking) */
    int32_t graph_after;       /* launches of one (plan, bank, stream) before its launch sequence is captured
                                  into a hipGraph and replayed (0 = never capture) */
} cogconv_server_config;

/* Request flags. */
enum {
    COGCONV_REQUEST_RESIDENT = 1 /* in0 / in1 / out are device pointers (state resident in HBM): no copies */
};

typedef struct {
    cogconv_desc desc;     /* in0: the request's field; in1: the bank's type when bank >= 0 */
    int32_t bank;          /* resident filter bank, or -1 to send in1 with the request */
    int32_t flags;         /* COGCONV_REQUEST_* */
    const float* host_in0; /* planar fp32; caller keeps it valid until completion */
    const float* host_in1; /* used when bank < 0 */
    float* host_out;       /* planar fp32 result (the op's output FieldType) */
} cogconv_request;

typedef struct {
    uint64_t submitted, completed, rejected;
    uint64_t launches;         /* kernel launches issued */
    uint64_t stacked;          /* requests that shared a launch with another */
    uint64_t graph_launches;   /* launches replayed from a captured hipGraph */
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> result in host memory */
    double first_submit_s, last_complete_s; /* monotonic clock */
    double requests_per_s;     /* completed / (last_complete - first_submit) */
    double gflop;              /* algorithmic GFLOP of completed requests */
    uint64_t peak_bytes;       /* peak device bytes held for requests */
} cogconv_server_stats;

int cogconv_server_create(const cogconv_server_config* config, cogconv_server* server);
/* Waits for in-flight work, then frees everything. */
int cogconv_server_destroy(cogconv_server server);
/* Copy a filter bank into HBM once; requests then name it by id. */
int cogconv_server_register_bank(cogconv_server server, const float* host_weights, const cogconv_field_type* type,
                                 int32_t* bank_id);
/* Non-blocking.  Illegal ops fail here (COGCONV_ERR_ILLEGAL_OP); a request larger than the
 * whole budget fails with COGCONV_ERR_BUDGET; otherwise it queues and gets a ticket. */
int cogconv_server_submit(cogconv_server server, const cogconv_request* request, uint64_t* ticket);
/* COGCONV_OK when the ticket's result is in host memory, COGCONV_ERR_NOT_READY before. */
int cogconv_server_query(cogconv_server server, uint64_t ticket);
/* Submit n requests from the calling thread, request i at (call time + arrival_s[i])
 * (a replayed arrival trace; all zeros = submit back to back).  Blocks until the
 * last one is submitted; tickets[i] receives each ticket (0 if it was refused). */
int cogconv_server_submit_timed(cogconv_server server, const cogconv_request* requests, const double* arrival_s,
                                int64_t n, uint64_t* tickets);
int cogconv_server_wait(cogconv_server server, uint64_t ticket);
int cogconv_server_drain(cogconv_server server);
int cogconv_server_stats_get(cogconv_server server, cogconv_server_stats* stats);
int cogconv_server_stats_reset(cogconv_server server);

#ifdef __cplusplus
}
#endif

#endif /* COGCONV_H */
