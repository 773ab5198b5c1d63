// Autotuner: times a plan's candidate kernel configurations and remembers the
// fastest, per device and descriptor.
//
// Counterpart of Profiler.bestVariant and ProfilerCache
// (src/main/scala/cogx/runtime/execution/Profiler.scala:131-190, 49-130;
// ProfilerCache.scala:39-122): candidates are timed on synthetic inputs (input
// i of n filled with (i+1)/n, Profiler.scala:84-92), a flush buffer is written
// before every timed step so that each step starts with cold caches
// (Profiler.scala:80-83), the lowest average step time wins (ties keep the
// earlier candidate), and a cached sample is reused unless profiling is forced
// or the sample used fewer steps than asked for (Profiler.scala:100-112).  The
// cache is a text file, one sample per line:
//   device <TAB> descriptor key <TAB> variant <TAB> avg_us <TAB> warmup <TAB> run <TAB> unix_ms <TAB> description
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

using namespace cogconv;

namespace {

struct Sample {
    int variant;
    double avg_us;
    int warmup, run;
    long long created_ms;
    std::string name;
};

std::mutex g_mu;
std::map<std::string, Sample> g_cache;  // key: device + '\t' + descriptor key

std::string desc_key(const cogconv_desc* d) {
    char b[256];
    snprintf(b, sizeof b, "m%d b%d o%d s%d st%d B%d a%d,%d,%d,%d,%d f%d,%d,%d,%d,%d p%d", d->mode, d->border,
             d->orientation, d->sampling, d->step, d->batch, d->in0.dims, d->in0.rows, d->in0.cols, d->in0.planes,
             d->in0.tensor_order, d->in1.dims, d->in1.rows, d->in1.cols, d->in1.planes, d->in1.tensor_order,
             d->precision);
    return b;
}

// Architecture and marketing name of the current device ("" without one).
std::string current_device_name() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return "";
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "";
    std::string n = std::string(p.gcnArchName) + "/" + p.name;
    for (char& c : n)
        if (c == '\t' || c == '\n' || c == ' ') c = '_';
    return n;
}

long long now_ms() {
    return std::chrono::duration_cast<std::chrono::milliseconds>(
               std::chrono::system_clock::now().time_since_epoch())
        .count();
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
    float* f() const { return static_cast<float*>(p); }
};

struct Ev {
    hipEvent_t e = nullptr;
    ~Ev() {
        if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

bool cogconv::tune_cache_lookup(const cogconv_desc* desc, int* variant, std::string* name) {
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_cache.empty()) return false;  // no device query while nothing was tuned
    }
    const std::string key = current_device_name() + "\t" + desc_key(desc);
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key);
    if (it == g_cache.end()) return false;
    *variant = it->second.variant;
    *name = it->second.name;
    return true;
}

extern "C" int cogconv_autotune(const cogconv_desc* desc, const cogconv_tune_options* options, cogconv_stream stream,
                                cogconv_tune_result* result) {
    clear_error();
    if (!desc || !result) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    std::memset(result, 0, sizeof *result);
    cogconv_tune_options o;
    o.warmup_steps = 1;
    o.run_steps = 5;
    o.flush_bytes = 512LL << 20;  // beyond the 256 MiB last-level cache and the L2s of MI355X
    o.force = 0;
    if (options) o = *options;
    if (o.warmup_steps < 0 || o.run_steps < 1 || o.flush_bytes < 0)
        return fail(COGCONV_ERR_INVALID_ARG, "bad tuning options (warmup %d, run %d, flush %lld)", o.warmup_steps,
                    o.run_steps, (long long)o.flush_bytes);
    cogconv_desc d = *desc;
    d.algo = COGCONV_ALGO_AUTO;
    int count = 0;
    int e = plan_variant_count(&d, &count);
    if (e) return e;
    if (count == 0) {
        cogconv_plan p = nullptr;
        return cogconv_plan_create(&d, &p);  // reports why no configuration exists
    }
    result->variants = count;
    const std::string dev = current_device_name();
    if (dev.empty()) return fail(COGCONV_ERR_RUNTIME, "autotuning needs a current device");
    const std::string key = dev + "\t" + desc_key(&d);
    if (!o.force) {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            const Sample& s = it->second;
            if (s.variant < count &&
                (s.run > o.run_steps || (s.run == o.run_steps && s.warmup >= o.warmup_steps))) {
                result->variant = s.variant;
                result->avg_us = s.avg_us;
                result->from_cache = 1;
                snprintf(result->description, sizeof result->description, "%s", s.name.c_str());
                return COGCONV_OK;
            }
        }
    }
    std::vector<cogconv_plan> plans(count, nullptr);
    struct PlanGuard {
        std::vector<cogconv_plan>& v;
        ~PlanGuard() {
            for (cogconv_plan p : v) cogconv_plan_destroy(p);
        }
    } pg{plans};
    size_t ws = 0;
    for (int i = 0; i < count; i++) {
        e = cogconv_plan_create_variant(&d, i, &plans[i]);
        if (e) return e;
        size_t w = 0;
        cogconv_plan_workspace_bytes(plans[i], &w);
        ws = std::max(ws, w);
    }
    const cogconv_field_type& out = plans[0]->out;
    const size_t n0 = (size_t)d.in0.planes * d.in0.rows * d.in0.cols;
    const size_t n1 = (size_t)d.in1.planes * d.in1.rows * d.in1.cols;
    const size_t no = (size_t)out.planes * out.rows * out.cols;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DevBuf b0, b1, bo, bw, bf;
    if ((e = check_hip(b0.alloc(n0 * 4), "hipMalloc (tuning input 0)")) ||
        (e = check_hip(b1.alloc(n1 * 4), "hipMalloc (tuning input 1)")) ||
        (e = check_hip(bo.alloc(no * 4), "hipMalloc (tuning output)")) ||
        (e = check_hip(bw.alloc(ws), "hipMalloc (tuning workspace)")) ||
        (e = check_hip(bf.alloc((size_t)o.flush_bytes), "hipMalloc (cache flush buffer)")))
        return e;
    if ((e = check_hip(launch_fill_value(b0.f(), n0, 0.5f, s), "tuning input fill")) ||
        (e = check_hip(launch_fill_value(b1.f(), n1, 1.0f, s), "tuning input fill")))
        return e;
    Ev e0, e1;
    if ((e = check_hip(hipEventCreate(&e0.e), "hipEventCreate")) ||
        (e = check_hip(hipEventCreate(&e1.e), "hipEventCreate")))
        return e;
    int best = -1;
    double bestUs = 0;
    for (int i = 0; i < count; i++) {
        for (int w = 0; w < o.warmup_steps; w++) {
            e = cogconv_launch(plans[i], b0.f(), b1.f(), bo.f(), bw.p, stream);
            if (e) return e;
        }
        double total_ms = 0;
        for (int r = 0; r < o.run_steps; r++) {
            if (o.flush_bytes && (e = check_hip(launch_fill_zero(bf.p, (size_t)o.flush_bytes, s), "cache flush")))
                return e;
            if ((e = check_hip(hipEventRecord(e0.e, s), "hipEventRecord"))) return e;
            e = cogconv_launch(plans[i], b0.f(), b1.f(), bo.f(), bw.p, stream);
            if (e) return e;
            if ((e = check_hip(hipEventRecord(e1.e, s), "hipEventRecord")) ||
                (e = check_hip(hipEventSynchronize(e1.e), "hipEventSynchronize")))
                return e;
            float ms = 0;
            if ((e = check_hip(hipEventElapsedTime(&ms, e0.e, e1.e), "hipEventElapsedTime"))) return e;
            total_ms += ms;
        }
        const double us = 1000.0 * total_ms / o.run_steps;
        if (best < 0 || us < bestUs) {
            best = i;
            bestUs = us;
        }
    }
    char name[192];
    cogconv_plan_describe(plans[best], name, sizeof name);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = Sample{best, bestUs, o.warmup_steps, o.run_steps, now_ms(), name};
    }
    result->variant = best;
    result->avg_us = bestUs;
    result->from_cache = 0;
    snprintf(result->description, sizeof result->description, "%s", name);
    return COGCONV_OK;
}

extern "C" int cogconv_tune_cache_clear(void) {
    clear_error();
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.clear();
    return COGCONV_OK;
}

extern "C" int cogconv_tune_cache_save(const char* path) {
    clear_error();
    if (!path) return fail(COGCONV_ERR_INVALID_ARG, "null path");
    std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return fail(COGCONV_ERR_RUNTIME, "cannot write %s", tmp.c_str());
    {
        std::lock_guard<std::mutex> lk(g_mu);
        fprintf(f, "# cogconv tune cache 1\n");
        for (const auto& kv : g_cache) {
            const Sample& s = kv.second;
            fprintf(f, "%s\t%d\t%.3f\t%d\t%d\t%lld\t%s\n", kv.first.c_str(), s.variant, s.avg_us, s.warmup, s.run,
                    s.created_ms, s.name.c_str());
        }
    }
    if (fclose(f) != 0 || rename(tmp.c_str(), path) != 0)
        return fail(COGCONV_ERR_RUNTIME, "cannot write %s", path);
    return COGCONV_OK;
}

// Merges the samples of `path` into the cache (a file sample replaces an
// in-memory one only if it used at least as many steps).  A missing file is
// not an error; a malformed line is.
extern "C" int cogconv_tune_cache_load(const char* path) {
    clear_error();
    if (!path) return fail(COGCONV_ERR_INVALID_ARG, "null path");
    FILE* f = fopen(path, "r");
    if (!f) return COGCONV_OK;
    std::vector<std::pair<std::string, Sample>> got;
    char line[1024];
    int lineno = 0;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        size_t n = strlen(line);
        while (n && (line[n - 1] == '\n' || line[n - 1] == '\r')) line[--n] = 0;
        if (n == 0 || line[0] == '#') continue;
        std::vector<std::string> fld;
        std::string cur;
        for (size_t i = 0; i < n; i++) {
            if (line[i] == '\t') {
                fld.push_back(cur);
                cur.clear();
            } else {
                cur += line[i];
            }
        }
        fld.push_back(cur);
        Sample s;
        char* end = nullptr;
        bool ok = fld.size() == 8;
        if (ok) {
            s.variant = (int)strtol(fld[2].c_str(), &end, 10);
            ok = *end == 0 && s.variant >= 0;
        }
        if (ok) {
            s.avg_us = strtod(fld[3].c_str(), &end);
            ok = *end == 0;
        }
        if (ok) {
            s.warmup = (int)strtol(fld[4].c_str(), &end, 10);
            ok = *end == 0;
        }
        if (ok) {
            s.run = (int)strtol(fld[5].c_str(), &end, 10);
            ok = *end == 0;
        }
        if (ok) {
            s.created_ms = strtoll(fld[6].c_str(), &end, 10);
            ok = *end == 0;
        }
        if (!ok) {
            fclose(f);
            return fail(COGCONV_ERR_INVALID_ARG, "%s:%d: malformed tuning sample", path, lineno);
        }
        s.name = fld[7];
        got.emplace_back(fld[0] + "\t" + fld[1], s);
    }
    fclose(f);
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : got) {
        auto it = g_cache.find(kv.first);
        if (it == g_cache.end() || kv.second.run >= it->second.run) g_cache[kv.first] = kv.second;
    }
    return COGCONV_OK;
}
