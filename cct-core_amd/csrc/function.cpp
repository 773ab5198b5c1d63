// Compiled convolution graphs: the file format that checkpoints them and the
// reentrant evaluate() Function Interface that runs them.
//
// Checkpoint (ComputeGraph.write / restore, src/main/scala/cogx/runtime/ComputeGraph.scala:523-612).
// The reference saves its optimized kernel circuit with each kernel's OpenCL
// source and re-runs that source on restore (OpenCLRestoredDeviceKernel.scala:28-60).
// Here a saved graph keeps each launch step's *descriptor* instead -- the
// ConvolveOp (cogconv_desc), the reduce factor, the separable filters, the
// merged elementwise epilogue -- together with the field types and constant
// field data, and restoring re-plans every step for the device it lands on
// (kernel choice, tile shape and autotuned variant are decided again).
//
// File layout: a text header, then the constant fields' float32 data (little
// endian) in the order of their `field ... const` lines:
//   COGCONV-GRAPH <major> <minor>
//   name <text to end of line>
//   description <text to end of line>                          (optional)
//   field <id> <dims> <rows> <cols> <planes> <order> input|const|computed <name>
//   step conv <result> <in0> <in1> <mode> <border> <orient> <sampling> <step> <batch> <precision>
//   step reduce <result> <src> <factor>
//   step elementwise <result> <src>
//   step separable <result> <x> <row filter> <column filter> <border> <orient>
//   epi <op> <value as %a> <slot>                               (0..8 after a step line)
//   operand <field id>                                         (0..4 after a step line)
//   output <field id> <name>
//   end
//
// Function Interface (the reference's CogFunction JNI API,
// src/main/java/cogjni/CogJNIWrapper.java:85-161): create from a file, list
// inputs and outputs with their FieldTypes, and evaluate(inputs -> outputs)
// from any number of host threads at once.  Each concurrent evaluate runs on
// an execution context of its own (a stream, device buffers for every
// non-constant field, step workspaces); constant fields are uploaded once and
// shared.  A context runs the same plans on the same data in the same order,
// so a result does not depend on which context ran it or what ran beside it.
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

using namespace cogconv;

namespace {

constexpr int kMajor = 1, kMinor = 0;

enum FieldKind { FK_INPUT = 0, FK_CONST = 1, FK_COMPUTED = 2 };
enum StepKind { SK_CONV = 0, SK_REDUCE = 1, SK_ELEMENTWISE = 2, SK_SEPARABLE = 3 };

struct GField {
    cogconv_field_type type;
    int kind;
    std::string name;
    std::vector<float> data;  // constants
    size_t bytes() const { return (size_t)type.rows * type.cols * type.planes * sizeof(float); }
};

struct GStep {
    int kind;
    int result;
    int in[3];
    cogconv_desc desc;  // conv
    int factor;         // reduce
    int border, orient; // separable
    cogconv_epilogue epi;
    int operands[COGCONV_EPI_MAX_OPERANDS];
    cogconv_plan plan = nullptr;
};

struct Context {
    hipStream_t stream = nullptr;
    std::vector<void*> buf;  // per field (constants: shared buffers)
    std::vector<void*> ws;   // per step
    bool busy = false;
};

}  // namespace

struct cogconv_function_s {
    std::string name, description;
    int device = 0;
    std::vector<GField> fields;
    std::vector<GStep> steps;
    std::vector<int> inputs, outputs;
    std::vector<std::string> output_names;
    std::vector<void*> constants;  // device copies of constant fields (by field id; null otherwise)
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<Context>> contexts;
    int max_contexts = 4;
};

namespace {

bool read_type(std::istringstream& ls, cogconv_field_type* t) {
    return static_cast<bool>(ls >> t->dims >> t->rows >> t->cols >> t->planes >> t->tensor_order);
}

int parse(std::istream& in, cogconv_function_s* fn) {
    std::string line;
    if (!std::getline(in, line)) return fail(COGCONV_ERR_INVALID_ARG, "empty graph file");
    {
        std::istringstream ls(line);
        std::string magic;
        int major = 0, minor = 0;
        ls >> magic >> major >> minor;
        if (magic != "COGCONV-GRAPH") return fail(COGCONV_ERR_INVALID_ARG, "not a cogconv graph file");
        if (major != kMajor)
            return fail(COGCONV_ERR_INVALID_ARG, "graph file spec %d.%d, this library reads %d.%d", major, minor,
                        kMajor, kMinor);
    }
    bool ended = false;
    while (!ended && std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        ls >> key;
        if (key == "name" || key == "description") {
            std::string rest;
            std::getline(ls, rest);
            if (!rest.empty() && rest[0] == ' ') rest.erase(0, 1);
            (key == "name" ? fn->name : fn->description) = rest;
        } else if (key == "field") {
            int id = -1;
            GField f;
            std::string kind;
            ls >> id;
            if (!read_type(ls, &f.type) || !(ls >> kind) || id != (int)fn->fields.size())
                return fail(COGCONV_ERR_INVALID_ARG, "bad field line: %s", line.c_str());
            std::getline(ls, f.name);
            if (!f.name.empty() && f.name[0] == ' ') f.name.erase(0, 1);
            f.kind = kind == "input" ? FK_INPUT : kind == "const" ? FK_CONST : kind == "computed" ? FK_COMPUTED : -1;
            if (f.kind < 0 || f.type.rows < 1 || f.type.cols < 1 || f.type.planes < 1)
                return fail(COGCONV_ERR_INVALID_ARG, "bad field line: %s", line.c_str());
            if (f.kind == FK_INPUT) fn->inputs.push_back(id);
            fn->fields.push_back(std::move(f));
        } else if (key == "step") {
            GStep s;
            std::memset(&s.desc, 0, sizeof s.desc);
            std::memset(&s.epi, 0, sizeof s.epi);
            s.in[0] = s.in[1] = s.in[2] = -1;
            s.factor = s.border = s.orient = 0;
            std::string kind;
            ls >> kind >> s.result;
            bool ok = true;
            if (kind == "conv") {
                s.kind = SK_CONV;
                cogconv_desc& d = s.desc;
                ok = static_cast<bool>(ls >> s.in[0] >> s.in[1] >> d.mode >> d.border >> d.orientation >> d.sampling >>
                                       d.step >> d.batch >> d.precision);
            } else if (kind == "reduce") {
                s.kind = SK_REDUCE;
                ok = static_cast<bool>(ls >> s.in[0] >> s.factor);
            } else if (kind == "elementwise") {
                s.kind = SK_ELEMENTWISE;
                ok = static_cast<bool>(ls >> s.in[0]);
            } else if (kind == "separable") {
                s.kind = SK_SEPARABLE;
                ok = static_cast<bool>(ls >> s.in[0] >> s.in[1] >> s.in[2] >> s.border >> s.orient);
            } else {
                ok = false;
            }
            if (!ok) return fail(COGCONV_ERR_INVALID_ARG, "bad step line: %s", line.c_str());
            fn->steps.push_back(s);
        } else if (key == "epi" || key == "operand") {
            if (fn->steps.empty()) return fail(COGCONV_ERR_INVALID_ARG, "%s before any step", key.c_str());
            cogconv_epilogue& e = fn->steps.back().epi;
            if (key == "epi") {
                std::string v;
                cogconv_epilogue_op op;
                if (e.n_ops >= COGCONV_EPI_MAX_OPS || !(ls >> op.op >> v >> op.operand))
                    return fail(COGCONV_ERR_INVALID_ARG, "bad epi line: %s", line.c_str());
                op.value = std::strtof(v.c_str(), nullptr);
                e.ops[e.n_ops++] = op;
            } else {
                int id = -1;
                if (e.n_operands >= COGCONV_EPI_MAX_OPERANDS || !(ls >> id))
                    return fail(COGCONV_ERR_INVALID_ARG, "bad operand line: %s", line.c_str());
                fn->steps.back().operands[e.n_operands++] = id;
            }
        } else if (key == "output") {
            int id = -1;
            ls >> id;
            std::string nm;
            std::getline(ls, nm);
            if (!nm.empty() && nm[0] == ' ') nm.erase(0, 1);
            fn->outputs.push_back(id);
            fn->output_names.push_back(nm);
        } else if (key == "end") {
            ended = true;
        } else if (!key.empty()) {
            return fail(COGCONV_ERR_INVALID_ARG, "unknown graph file line: %s", line.c_str());
        }
    }
    if (!ended) return fail(COGCONV_ERR_INVALID_ARG, "graph file has no end line");
    // Constant data follows, in field order.
    for (GField& f : fn->fields) {
        if (f.kind != FK_CONST) continue;
        f.data.resize(f.bytes() / sizeof(float));
        in.read(reinterpret_cast<char*>(f.data.data()), (std::streamsize)f.bytes());
        if (!in) return fail(COGCONV_ERR_INVALID_ARG, "graph file ends inside the data of field '%s'", f.name.c_str());
    }
    // Reference checks: every id in range, every step's result computed once, steps in dependence order.
    const int nf = (int)fn->fields.size();
    std::vector<char> ready(nf, 0);
    for (int i = 0; i < nf; i++) ready[i] = fn->fields[i].kind != FK_COMPUTED;
    for (const GStep& s : fn->steps) {
        auto bad_id = [&](int id) { return id < 0 || id >= nf; };
        if (bad_id(s.result) || fn->fields[s.result].kind != FK_COMPUTED || ready[s.result])
            return fail(COGCONV_ERR_INVALID_ARG, "step result %d is not a computed field produced once", s.result);
        const int nin = s.kind == SK_CONV ? 2 : s.kind == SK_SEPARABLE ? 3 : 1;
        for (int i = 0; i < nin; i++)
            if (bad_id(s.in[i]) || !ready[s.in[i]])
                return fail(COGCONV_ERR_INVALID_ARG, "step for field %d reads field %d before it is ready", s.result,
                            s.in[i]);
        for (int i = 0; i < s.epi.n_operands; i++)
            if (bad_id(s.operands[i]) || !ready[s.operands[i]])
                return fail(COGCONV_ERR_INVALID_ARG, "epilogue operand %d of field %d is not ready", s.operands[i],
                            s.result);
        ready[s.result] = 1;
    }
    for (int id : fn->outputs)
        if (id < 0 || id >= nf || !ready[id]) return fail(COGCONV_ERR_INVALID_ARG, "output field %d is not produced", id);
    return COGCONV_OK;
}

// Re-plans every step (the rebinding of a restored graph) and checks each
// step's result type against the saved one.
int plan_steps(cogconv_function_s* fn) {
    for (GStep& s : fn->steps) {
        for (int i = 0; i < s.epi.n_operands; i++) s.epi.operands[i] = fn->fields[s.operands[i]].type;
        const cogconv_field_type& rt = fn->fields[s.result].type;
        cogconv_field_type out = rt;
        if (s.kind == SK_CONV) {
            s.desc.in0 = fn->fields[s.in[0]].type;
            s.desc.in1 = fn->fields[s.in[1]].type;
            s.desc.algo = COGCONV_ALGO_AUTO;
            if (int e = cogconv_plan_create(&s.desc, &s.plan)) return e;
            cogconv_plan_output_type(s.plan, &out);
            if (int e = cogconv_plan_set_epilogue(s.plan, &s.epi)) return e;
        } else if (s.kind == SK_REDUCE) {
            const cogconv_field_type& it = fn->fields[s.in[0]].type;
            if (s.factor < 1 || it.planes % s.factor) return fail(COGCONV_ERR_ILLEGAL_OP, "bad reduce factor");
            out = it;
            out.planes = it.planes / s.factor;
        } else {
            out = fn->fields[s.in[0]].type;
        }
        if (out.rows != rt.rows || out.cols != rt.cols || out.planes != rt.planes)
            return fail(COGCONV_ERR_INVALID_ARG, "saved type of field %d (%dx%dx%d) differs from its step's (%dx%dx%d)",
                        s.result, rt.rows, rt.cols, rt.planes, out.rows, out.cols, out.planes);
        if (s.kind != SK_CONV) {
            // Check the epilogue against the result type now, not at the first evaluate.
            Epi tmp;
            int n = 0;
            if (int e = epi_build(rt, &s.epi, &tmp, &n)) return e;
        }
    }
    return COGCONV_OK;
}

void free_context(cogconv_function_s* fn, Context* c) {
    for (size_t i = 0; i < c->buf.size(); i++)
        if (c->buf[i] && !(i < fn->constants.size() && fn->constants[i] == c->buf[i])) (void)hipFree(c->buf[i]);
    for (void* w : c->ws)
        if (w) (void)hipFree(w);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int new_context(cogconv_function_s* fn, std::unique_ptr<Context>* out) {
    std::unique_ptr<Context> c(new (std::nothrow) Context);
    if (!c) return fail(COGCONV_ERR_OOM, "host allocation failed");
    int e = check_hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
    c->buf.assign(fn->fields.size(), nullptr);
    c->ws.assign(fn->steps.size(), nullptr);
    for (size_t i = 0; !e && i < fn->fields.size(); i++) {
        if (fn->constants[i]) c->buf[i] = fn->constants[i];
        else e = check_hip(hipMalloc(&c->buf[i], fn->fields[i].bytes()), "field buffer");
    }
    for (size_t i = 0; !e && i < fn->steps.size(); i++) {
        size_t w = 0;
        if (fn->steps[i].plan) cogconv_plan_workspace_bytes(fn->steps[i].plan, &w);
        if (w) e = check_hip(hipMalloc(&c->ws[i], w), "step workspace");
    }
    if (e) {
        free_context(fn, c.get());
        return e;
    }
    *out = std::move(c);
    return COGCONV_OK;
}

int run_steps(cogconv_function_s* fn, Context* c) {
    cogconv_stream s = reinterpret_cast<cogconv_stream>(c->stream);
    for (size_t i = 0; i < fn->steps.size(); i++) {
        GStep& st = fn->steps[i];
        const float* ops[COGCONV_EPI_MAX_OPERANDS];
        for (int k = 0; k < st.epi.n_operands; k++) ops[k] = static_cast<const float*>(c->buf[st.operands[k]]);
        const float* a = static_cast<const float*>(c->buf[st.in[0]]);
        float* out = static_cast<float*>(c->buf[st.result]);
        int e = COGCONV_OK;
        switch (st.kind) {
        case SK_CONV:
            e = cogconv_launch_epilogue(st.plan, a, static_cast<const float*>(c->buf[st.in[1]]), out, c->ws[i], ops, s);
            break;
        case SK_REDUCE:
            e = cogconv_block_reduce_sum_epilogue(a, out, &fn->fields[st.in[0]].type, st.factor, &st.epi, ops, s);
            break;
        case SK_ELEMENTWISE:
            e = cogconv_elementwise(a, out, &fn->fields[st.result].type, &st.epi, ops, s);
            break;
        default:
            e = cogconv_separable(a, &fn->fields[st.in[0]].type, static_cast<const float*>(c->buf[st.in[1]]),
                                  fn->fields[st.in[1]].type.cols, static_cast<const float*>(c->buf[st.in[2]]),
                                  fn->fields[st.in[2]].type.cols, st.border, st.orient, out, &st.epi, ops, s);
            break;
        }
        if (e) return e;
    }
    return COGCONV_OK;
}

}  // namespace

extern "C" int cogconv_function_create(const char* path, int32_t device, cogconv_function* function) {
    clear_error();
    if (!path || !function) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *function = nullptr;
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(COGCONV_ERR_INVALID_ARG, "cannot open graph file '%s'", path);
    std::unique_ptr<cogconv_function_s> fn(new (std::nothrow) cogconv_function_s);
    if (!fn) return fail(COGCONV_ERR_OOM, "host allocation failed");
    fn->device = device;
    if (int e = parse(in, fn.get())) return e;
    // Planning is host work: a file whose steps do not plan fails before any device call.
    if (int e = plan_steps(fn.get())) {
        for (GStep& s : fn->steps) cogconv_plan_destroy(s.plan);
        return e;
    }
    if (int e = cogconv_init(device)) {
        for (GStep& s : fn->steps) cogconv_plan_destroy(s.plan);
        return e;
    }
    fn->constants.assign(fn->fields.size(), nullptr);
    for (size_t i = 0; i < fn->fields.size(); i++) {
        const GField& f = fn->fields[i];
        if (f.kind != FK_CONST) continue;
        int e = check_hip(hipMalloc(&fn->constants[i], f.bytes()), "constant field");
        if (!e) e = check_hip(hipMemcpy(fn->constants[i], f.data.data(), f.bytes(), hipMemcpyHostToDevice),
                              "constant upload");
        if (e) {
            cogconv_function_destroy(fn.release());
            return e;
        }
    }
    *function = fn.release();
    return COGCONV_OK;
}

extern "C" int cogconv_function_destroy(cogconv_function fn) {
    clear_error();
    if (!fn) return COGCONV_OK;
    {
        std::unique_lock<std::mutex> lk(fn->mu);
        fn->cv.wait(lk, [&] {
            for (auto& c : fn->contexts)
                if (c->busy) return false;
            return true;
        });
    }
    (void)hipSetDevice(fn->device);
    for (auto& c : fn->contexts) free_context(fn, c.get());
    for (void* p : fn->constants)
        if (p) (void)hipFree(p);
    for (GStep& s : fn->steps) cogconv_plan_destroy(s.plan);
    delete fn;
    return COGCONV_OK;
}

extern "C" int cogconv_function_name(cogconv_function fn, char* buf, size_t len) {
    clear_error();
    if (!fn || !buf || !len) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    snprintf(buf, len, "%s", fn->name.c_str());
    return COGCONV_OK;
}

extern "C" int cogconv_function_optimal_thread_count(cogconv_function fn, int32_t* count) {
    clear_error();
    if (!fn || !count) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    // Two evaluates in flight keep the device busy while the other thread's
    // copies and host work run; more contexts only add memory.
    *count = 2;
    return COGCONV_OK;
}

static int io_lookup(cogconv_function fn, int output, int32_t index, const GField** f, const std::string** nm) {
    if (!fn) return fail(COGCONV_ERR_INVALID_ARG, "null function");
    const std::vector<int>& ids = output ? fn->outputs : fn->inputs;
    if (index < 0 || index >= (int32_t)ids.size())
        return fail(COGCONV_ERR_INVALID_ARG, "%s index %d out of range (%d)", output ? "output" : "input", index,
                    (int)ids.size());
    *f = &fn->fields[ids[index]];
    *nm = output ? &fn->output_names[index] : &(*f)->name;
    return COGCONV_OK;
}

extern "C" int cogconv_function_input_count(cogconv_function fn, int32_t* count) {
    clear_error();
    if (!fn || !count) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *count = (int32_t)fn->inputs.size();
    return COGCONV_OK;
}

extern "C" int cogconv_function_output_count(cogconv_function fn, int32_t* count) {
    clear_error();
    if (!fn || !count) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *count = (int32_t)fn->outputs.size();
    return COGCONV_OK;
}

extern "C" int cogconv_function_io(cogconv_function fn, int32_t output, int32_t index, char* name, size_t name_len,
                                   cogconv_field_type* type) {
    clear_error();
    const GField* f = nullptr;
    const std::string* nm = nullptr;
    if (int e = io_lookup(fn, output, index, &f, &nm)) return e;
    if (name && name_len) snprintf(name, name_len, "%s", nm->c_str());
    if (type) *type = f->type;
    return COGCONV_OK;
}

extern "C" int cogconv_function_evaluate(cogconv_function fn, const float* const* inputs, float* const* outputs) {
    clear_error();
    if (!fn) return fail(COGCONV_ERR_INVALID_ARG, "null function");
    if ((!inputs && !fn->inputs.empty()) || (!outputs && !fn->outputs.empty()))
        return fail(COGCONV_ERR_INVALID_ARG, "null input or output array");
    for (size_t i = 0; i < fn->inputs.size(); i++)
        if (!inputs[i]) return fail(COGCONV_ERR_INVALID_ARG, "null input %zu", i);
    for (size_t i = 0; i < fn->outputs.size(); i++)
        if (!outputs[i]) return fail(COGCONV_ERR_INVALID_ARG, "null output %zu", i);
    if (int e = check_hip(hipSetDevice(fn->device), "hipSetDevice")) return e;
    // Take a free context, making one while fewer than max_contexts exist.
    Context* c = nullptr;
    {
        std::unique_lock<std::mutex> lk(fn->mu);
        for (;;) {
            for (auto& x : fn->contexts)
                if (!x->busy) {
                    c = x.get();
                    break;
                }
            if (c || (int)fn->contexts.size() < fn->max_contexts) break;
            fn->cv.wait(lk);
        }
        if (!c) {
            std::unique_ptr<Context> nc;
            lk.unlock();
            const int e = new_context(fn, &nc);
            lk.lock();
            if (e) return e;
            fn->contexts.push_back(std::move(nc));
            c = fn->contexts.back().get();
        }
        c->busy = true;
    }
    int e = COGCONV_OK;
    for (size_t i = 0; !e && i < fn->inputs.size(); i++) {
        const int id = fn->inputs[i];
        e = check_hip(hipMemcpyAsync(c->buf[id], inputs[i], fn->fields[id].bytes(), hipMemcpyHostToDevice, c->stream),
                      "input copy");
    }
    if (!e) e = run_steps(fn, c);
    for (size_t i = 0; !e && i < fn->outputs.size(); i++) {
        const int id = fn->outputs[i];
        e = check_hip(hipMemcpyAsync(outputs[i], c->buf[id], fn->fields[id].bytes(), hipMemcpyDeviceToHost, c->stream),
                      "output copy");
    }
    const int se = check_hip(hipStreamSynchronize(c->stream), "evaluate");
    if (!e) e = se;
    {
        std::lock_guard<std::mutex> lk(fn->mu);
        c->busy = false;
    }
    fn->cv.notify_all();
    return e;
}
