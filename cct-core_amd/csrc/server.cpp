// Serving loop: continuous admission of convolution requests on one GPU.
//
// The reference specifies a reentrant, thread-safe evaluate() Function Interface
// (doc/CogFunctionInterfaceERS.docx; src/main/java/cogjni/CogJNIWrapper.java:85-115)
// and shares device buffers under a budget (Cog.bufferSharing,
// src/main/scala/cogx/parameters/Cog.scala:83), but has no request scheduler.
// This server:
//   - keeps filter banks resident in HBM (registered once, named by id);
//   - preallocates its whole device budget once and sub-allocates request buffers
//     from it (first fit, 256-byte granules), so the loop never calls hipMalloc;
//   - admits queued requests in arrival order whenever a launch slot and enough
//     budget are free (a request that arrives while others run is admitted as soon
//     as they release memory: continuous admission);
//   - stacks queued requests of one shape bucket (same op, same bank, same field
//     shape) into a single launch when the vector mode keeps images independent
//     (ProjectFrame / BackProjectFrame, with or without block reduction: batch is
//     folded into planes, so stacking is concatenation of planes).  Requests in
//     host memory are copied into consecutive planes; requests whose fields are
//     resident in HBM at unrelated addresses run as one request-table launch
//     (each image's blocks read their buffers from a table the host builds in
//     pinned memory and copies ahead of the kernel on the launch's stream);
//   - captures the kernel of a hot (plan, bank, stream) request-table launch
//     once into a hipGraph and replays it (the table copy stays outside the
//     graph, so a replay runs on that launch's own table); a request alone runs
//     its plain launch;
//   - spreads launches over several streams: copies and kernels of different
//     launches overlap on the device;
//   - records submit -> result-ready latency per request (result in host memory,
//     or in HBM for requests whose state is resident there).
// Kernels and their summation order do not depend on how many images share a
// launch, so a request's result is the same alone and in the mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

using namespace cogconv;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device region.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    size_t in_use = 0, peak = 0;

    static size_t round(size_t b) { return (b + 255) & ~size_t(255); }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                in_use += bytes;
                peak = std::max(peak, in_use);
                return true;
            }
        }
        return false;
    }
    // Last fit from the top of the region: long-lived blocks (packed banks)
    // gather at the end, away from the request blocks that come and go.
    bool alloc_high(size_t bytes, size_t* off) {
        bytes = round(bytes ? bytes : 1);
        for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
            if (it->second >= bytes) {
                const size_t start = it->first, len = it->second;
                *off = start + len - bytes;
                free_.erase(start);
                if (len > bytes) free_[start] = len - bytes;
                in_use += bytes;
                peak = std::max(peak, in_use);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round(bytes ? bytes : 1);
        in_use -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Bank {
    float* dptr;
    cogconv_field_type type;
};

struct Req {
    uint64_t ticket;
    cogconv_request r;
    cogconv_field_type out;
    size_t in0_bytes, in1_bytes, out_bytes;
    double flops;
    cogconv_plan plan = nullptr;  // the server's plan for this request's descriptor
    int bucket = -1;              // shape bucket (per-image descriptor and bank), interned at submit
    double t_submit = 0, t_done = 0;
    bool done = false;
    int waiters = 0;  // threads inside cogconv_server_wait on this request
    int status = COGCONV_OK;
    std::string error;
};

struct Launch {
    std::vector<Req*> reqs;
    int stream;
    hipEvent_t ev;
    std::vector<std::pair<size_t, size_t>> blocks;  // arena (offset, bytes)
    int slot = -1;                                  // pinned request-table slot
};

// A captured launch sequence of one (plan, bank, stream).
struct GraphEntry {
    int uses = 0;
    hipGraphExec_t exec = nullptr;
    bool capturing = false;  // a worker is capturing it (others launch directly meanwhile)
    bool failed = false;     // capture did not work: launch directly
};

bool stackable(int mode) {
    return mode == COGCONV_PROJECT_FRAME || mode == COGCONV_PROJECT_FRAME_BLOCK_REDUCE_SUM ||
           mode == COGCONV_BACK_PROJECT_FRAME || mode == COGCONV_BACK_PROJECT_FRAME_BLOCK_REDUCE_SUM;
}

// Shape bucket: everything but the batch count.
std::string bucket_key(const Req* q) {
    cogconv_desc d = q->r.desc;
    const int per_image = d.in0.planes / d.batch;
    d.in0.planes = per_image;
    d.batch = 1;
    std::string k(reinterpret_cast<const char*>(&d), sizeof d);
    k.append(reinterpret_cast<const char*>(&q->r.bank), sizeof q->r.bank);
    return k;
}

}  // namespace

struct cogconv_server_s {
    cogconv_server_config cfg;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<Bank> banks;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> pending;
    std::vector<Launch*> inflight;
    std::unordered_map<uint64_t, Req*> reqs;
    std::unordered_map<std::string, cogconv_plan> plans;
    // Shape buckets: (plan, bank) of one-image requests and bucket_key of
    // batched ones -> bucket id.
    struct PlanBankHash {
        size_t operator()(const std::pair<cogconv_plan, int>& k) const {
            return std::hash<const void*>()(k.first) * 31 + std::hash<int>()(k.second);
        }
    };
    std::unordered_map<std::pair<cogconv_plan, int>, int, PlanBankHash> plan_buckets;
    std::unordered_map<std::string, int> buckets;
    size_t bucket_count = 0;
    std::vector<hipEvent_t> event_pool;
    // Resident banks packed once per (pack layout, bank) for the plans that
    // pack operand 2 (the low-precision frame kernels): arena offset, kept for
    // the server's life; plans that do not pack map to SIZE_MAX.  Plans of one
    // bank that differ only in field size or batch share one packed copy.
    std::map<std::pair<int, int>, size_t> packed;
    std::unordered_map<cogconv_plan, int> pack_ids;   // plan -> pack layout id
    std::unordered_map<std::string, int> pack_keys;   // cogconv_plan_pack_key -> id
    int pack_id(cogconv_plan p) {
        auto it = pack_ids.find(p);
        if (it != pack_ids.end()) return it->second;
        const int id = pack_keys.emplace(cogconv_plan_pack_key(p), (int)pack_keys.size()).first->second;
        pack_ids.emplace(p, id);
        return id;
    }
    // Request tables: per stream one device table of max_stack entries (a
    // stream runs its launches in order, so the table copy of a launch lands
    // after the previous launch on that stream has read its own), and pinned
    // host slots, one per launch in flight.
    std::vector<cogconv_image_io*> dev_tables;
    // Graphs are captured on a stream of their own per serving stream (no event
    // is ever recorded there): events recorded on a serving stream stay
    // queryable while a launch sequence is being captured.
    std::vector<hipStream_t> capture_streams;
    cogconv_image_io* host_tables = nullptr;
    std::vector<int> free_slots;
    std::map<std::tuple<cogconv_plan, int, int>, GraphEntry> graphs;
    uint64_t graph_launches = 0;
    size_t reserved = 0;  // arena bytes held by packed banks (taken from the top of the arena)
    // Worker threads (two when there are two streams or more): each admits
    // requests, issues launches and retires completions.  All shared state is
    // guarded by mu; issue() drops the lock only around runtime calls
    // (launches, copies, event records), which touch no shared structure, so
    // that the workers' runtime calls and the submitters proceed in parallel.
    std::vector<std::thread> workers;
    int issuing = 0;  // launches being issued with mu dropped (count against max_inflight)
    template <class Fn>
    static auto unlocked(std::unique_lock<std::mutex>& lk, Fn fn) {
        lk.unlock();
        auto r = fn();
        lk.lock();
        return r;
    }
    bool stop = false;
    uint64_t next_ticket = 1;
    int rr = 0;
    // Per stream: a worker is enqueueing on it with mu dropped.  A stream takes
    // one worker's enqueues at a time: its request table is overwritten by the
    // next launch's copy only after the previous launch read it (stream order),
    // which holds only if launches reach the stream one whole launch at a time,
    // and a stream being captured into a graph must receive no other work.
    std::vector<char> stream_issuing;
    // Round robin over the streams no other worker is enqueueing on; -1 if none.
    int take_stream() {
        const int n = (int)streams.size();
        for (int k = 0; k < n; k++) {
            const int si = (rr + k) % n;
            if (!stream_issuing[si]) {
                rr = si + 1;
                stream_issuing[si] = 1;
                return si;
            }
        }
        return -1;
    }
    struct StreamHold {  // gives the stream back when issue() returns (mu held)
        cogconv_server_s* s;
        int si;
        ~StreamHold() {
            if (si >= 0) s->stream_issuing[si] = 0;
        }
    };
    // statistics
    std::vector<double> lat_ms;
    uint64_t submitted = 0, completed = 0, rejected = 0, launches = 0, stacked = 0;
    double first_submit = 0, last_complete = 0, gflop = 0;

    cogconv_plan plan_for(const cogconv_desc& d, int* err) {
        std::string k(reinterpret_cast<const char*>(&d), sizeof d);
        auto it = plans.find(k);
        if (it != plans.end()) return it->second;
        cogconv_plan p = nullptr;
        *err = cogconv_plan_create(&d, &p);
        if (*err) return nullptr;
        plans[k] = p;
        return p;
    }

    hipError_t take_event(hipEvent_t* ev) {
        if (!event_pool.empty()) {
            *ev = event_pool.back();
            event_pool.pop_back();
            return hipSuccess;
        }
        return hipEventCreateWithFlags(ev, hipEventDisableTiming);
    }

    // Requests in host memory stack by copying into consecutive planes;
    // resident ones need a request-table launch, which the MFMA frame plans take.
    bool stack_ok(const Req* h) {
        if (!(h->r.flags & COGCONV_REQUEST_RESIDENT)) return true;
        return h->plan && plan_table_capable(h->plan);
    }

    void fail_reqs(const std::vector<Req*>& g, int status, const std::string& msg) {
        for (Req* q : g) {
            q->status = status;
            q->error = msg;
            q->done = true;
            q->t_done = now_s();
        }
        cv_done.notify_all();
    }

    // Issue one launch for the requests in g (one bucket).  Called with mu held.
    // Returns false when the arena cannot hold it now (nothing issued).
    bool issue(const std::vector<Req*>& g, std::unique_lock<std::mutex>& lk) {
        const Req* h = g.front();
        cogconv_desc d = h->r.desc;
        int B = 0, P0 = 0;
        size_t in0 = 0, outb = 0;
        for (Req* q : g) {
            B += q->r.desc.batch;
            P0 += q->r.desc.in0.planes;
            in0 += q->in0_bytes;
            outb += q->out_bytes;
        }
        d.batch = B;
        d.in0.planes = P0;
        int err = 0;
        cogconv_plan plan = g.size() == 1 ? h->plan : plan_for(d, &err);
        if (!plan) {
            fail_reqs(g, err, cogconv_last_error());
            return true;
        }
        size_t ws = 0;
        cogconv_plan_workspace_bytes(plan, &ws);
        const bool resident = (h->r.flags & COGCONV_REQUEST_RESIDENT) != 0;
        if (resident) {
            // State is already in HBM: only the workspace comes from the budget.
            const int si = take_stream();
            if (si < 0) return false;  // every stream is taken by another worker: retry
            StreamHold hold{this, si};
            hipStream_t s = streams[si];
            const float* in1 = h->r.bank < 0 ? h->r.host_in1 : banks[h->r.bank].dptr;
            // A resident bank a plan packs is packed once (synchronously, the
            // first time) and every later request runs on the packed form.
            size_t pk = SIZE_MAX;
            if (h->r.bank >= 0 && ws) {
                const auto key = std::make_pair(pack_id(plan), h->r.bank);
                auto it = packed.find(key);
                if (it == packed.end()) {
                    size_t o = 0;
                    if (!arena.alloc_high(ws, &o)) return false;
                    int32_t did = 0;
                    const int pst = unlocked(lk, [&] {
                        int r = cogconv_plan_pack_operand(plan, in1, arena.base + o,
                                                          reinterpret_cast<cogconv_stream>(s), &did);
                        if (r == COGCONV_OK && did) r = check_hip(hipStreamSynchronize(s), "bank pack");
                        return r;
                    });
                    if (pst != COGCONV_OK) {
                        arena.release(o, ws);
                        fail_reqs(g, pst, cogconv_last_error());
                        return true;
                    }
                    it = packed.find(key);  // another worker may have packed it meanwhile
                    if (it != packed.end() || !did) arena.release(o, ws);
                    else reserved += Arena::round(ws);
                    if (it == packed.end()) it = packed.emplace(key, did ? o : SIZE_MAX).first;
                }
                pk = it->second;
            }
            std::vector<std::pair<size_t, size_t>> blocks;
            size_t o_ws = 0;
            if (pk == SIZE_MAX && ws) {
                if (!arena.alloc(ws, &o_ws)) return false;
                blocks.push_back({o_ws, ws});
            }
            void* wsp = pk != SIZE_MAX ? arena.base + pk : (ws ? arena.base + o_ws : nullptr);
            // A stack of requests of a table-capable plan on a resident bank runs
            // as one request-table launch (one image per request), which a graph
            // replays once the (plan, bank, stream) is hot.  A request alone runs
            // the plain (packed) launch: the table's host-to-device copy would
            // cost more than it saves (c5 capacity 54 k req/s with tables for
            // single requests, in FP32 and BF16 alike).
            const bool table = g.size() > 1 && h->r.bank >= 0 && plan_table_capable(plan) && !free_slots.empty();
            if (g.size() > 1 && !table) {  // a stack needs a table: the caller issues the head alone
                for (auto& bk : blocks) arena.release(bk.first, bk.second);
                return false;
            }
            int slot = -1;
            bool x16 = true;
            GraphEntry* ge = nullptr;
            bool capture = false;
            if (table) {
                slot = free_slots.back();
                free_slots.pop_back();
                cogconv_image_io* ht = host_tables + (size_t)slot * cfg.max_stack;
                for (size_t i = 0; i < g.size(); i++) {
                    ht[i].in0 = g[i]->r.host_in0;
                    ht[i].out = g[i]->r.host_out;
                    x16 = x16 && (reinterpret_cast<uintptr_t>(g[i]->r.host_in0) & 15) == 0;
                }
                if (cfg.graph_after > 0 && x16) {
                    ge = &graphs[std::make_tuple(plan, h->r.bank, si)];
                    ge->uses++;
                    if (!ge->exec && !ge->failed && !ge->capturing && ge->uses >= cfg.graph_after) {
                        capture = true;
                        ge->capturing = true;
                    }
                }
            }
            hipGraphExec_t exec = ge ? ge->exec : nullptr;
            hipEvent_t ev = nullptr;
            hipError_t e = take_event(&ev);
            bool captured = false, capture_failed = false;
            const int st = e != hipSuccess ? COGCONV_ERR_RUNTIME : unlocked(lk, [&] {
                int r = COGCONV_OK;
                if (table) {
                    r = check_hip(hipMemcpyAsync(dev_tables[si], host_tables + (size_t)slot * cfg.max_stack,
                                                 sizeof(cogconv_image_io) * g.size(), hipMemcpyHostToDevice, s),
                                  "request table copy");
                    if (r == COGCONV_OK && capture) {
                        hipGraph_t graph = nullptr;
                        hipGraphExec_t ex = nullptr;
                        hipStream_t cs = capture_streams[si];
                        bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
                        if (ok) {
                            const int lr = launch_table(plan, dev_tables[si], x16, in1, wsp, pk != SIZE_MAX, cs);
                            ok = hipStreamEndCapture(cs, &graph) == hipSuccess && lr == COGCONV_OK && graph;
                        }
                        ok = ok && hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess;
                        if (graph) (void)hipGraphDestroy(graph);
                        clear_error();
                        if (ok) {
                            exec = ex;
                            captured = true;
                        } else {
                            if (ex) (void)hipGraphExecDestroy(ex);
                            (void)hipGetLastError();
                            capture_failed = true;
                        }
                    }
                    if (r == COGCONV_OK) {
                        if (exec) r = check_hip(hipGraphLaunch(exec, s), "graph launch");
                        else r = launch_table(plan, dev_tables[si], x16, in1, wsp, pk != SIZE_MAX, s);
                    }
                } else if (pk != SIZE_MAX) {
                    r = cogconv_launch_packed(plan, h->r.host_in0, h->r.host_out, wsp,
                                              reinterpret_cast<cogconv_stream>(s));
                } else {
                    r = cogconv_launch(plan, h->r.host_in0, in1, h->r.host_out, wsp,
                                       reinterpret_cast<cogconv_stream>(s));
                }
                if (r == COGCONV_OK) e = hipEventRecord(ev, s);
                return r;
            });
            if (capture) {
                ge->capturing = false;
                if (captured) ge->exec = exec;
                if (capture_failed) ge->failed = true;
            }
            if (st != COGCONV_OK || e != hipSuccess) {
                (void)hipStreamSynchronize(s);
                for (auto& bk : blocks) arena.release(bk.first, bk.second);
                if (slot >= 0) free_slots.push_back(slot);
                if (ev) event_pool.push_back(ev);
                fail_reqs(g, st ? st : COGCONV_ERR_RUNTIME, st ? cogconv_last_error() : hipGetErrorString(e));
                return true;
            }
            if (exec) graph_launches++;
            inflight.push_back(new Launch{g, si, ev, blocks, slot});
            launches++;
            if (g.size() > 1) stacked += g.size();
            return true;
        }
        const bool sendIn1 = h->r.bank < 0;
        std::vector<std::pair<size_t, size_t>> blocks;
        auto take = [&](size_t bytes, size_t* off) {
            if (!arena.alloc(bytes, off)) return false;
            blocks.push_back({*off, bytes});
            return true;
        };
        size_t o_in0 = 0, o_out = 0, o_ws = 0, o_in1 = 0;
        if (!take(in0, &o_in0) || !take(outb, &o_out) || (ws && !take(ws, &o_ws)) ||
            (sendIn1 && !take(h->in1_bytes, &o_in1))) {
            for (auto& b : blocks) arena.release(b.first, b.second);
            return false;
        }
        const int si = take_stream();
        if (si < 0) {
            for (auto& b : blocks) arena.release(b.first, b.second);
            return false;
        }
        StreamHold hold{this, si};
        hipStream_t s = streams[si];
        char* base = arena.base;
        hipEvent_t ev = nullptr;
        hipError_t e = take_event(&ev);
        const float* in1 = sendIn1 ? reinterpret_cast<const float*>(base + o_in1) : banks[h->r.bank].dptr;
        const int st = e != hipSuccess ? COGCONV_ERR_RUNTIME : unlocked(lk, [&] {
            size_t at = 0;
            for (Req* q : g) {
                if (e == hipSuccess)
                    e = hipMemcpyAsync(base + o_in0 + at, q->r.host_in0, q->in0_bytes, hipMemcpyHostToDevice, s);
                at += q->in0_bytes;
            }
            if (e == hipSuccess && sendIn1)
                e = hipMemcpyAsync(base + o_in1, h->r.host_in1, h->in1_bytes, hipMemcpyHostToDevice, s);
            int r = COGCONV_OK;
            if (e == hipSuccess)
                r = cogconv_launch(plan, reinterpret_cast<const float*>(base + o_in0), in1,
                                   reinterpret_cast<float*>(base + o_out), ws ? base + o_ws : nullptr,
                                   reinterpret_cast<cogconv_stream>(s));
            at = 0;
            for (Req* q : g) {
                if (e == hipSuccess && r == COGCONV_OK)
                    e = hipMemcpyAsync(q->r.host_out, base + o_out + at, q->out_bytes, hipMemcpyDeviceToHost, s);
                at += q->out_bytes;
            }
            if (e == hipSuccess && r == COGCONV_OK) e = hipEventRecord(ev, s);
            return r;
        });
        if (e != hipSuccess || st != COGCONV_OK) {
            (void)hipStreamSynchronize(s);
            for (auto& b : blocks) arena.release(b.first, b.second);
            if (ev) event_pool.push_back(ev);
            fail_reqs(g, st ? st : COGCONV_ERR_RUNTIME, st ? cogconv_last_error() : hipGetErrorString(e));
            return true;
        }
        Launch* L = new Launch{g, si, ev, blocks};
        inflight.push_back(L);
        launches++;
        if (g.size() > 1) stacked += g.size();
        return true;
    }

    void run() {
        (void)hipSetDevice(cfg.device);
        std::vector<char> busy;  // per stream: an earlier launch on it is not done
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // Completions.  A stream runs its launches in issue order, so once
            // one is not done the later ones on that stream are skipped: a
            // poll costs one event query per busy stream, not one per launch.
            busy.assign(streams.size(), 0);
            for (size_t i = 0; i < inflight.size();) {
                Launch* L = inflight[i];
                if (busy[L->stream]) {
                    i++;
                    continue;
                }
                const hipError_t q = hipEventQuery(L->ev);
                if (q == hipErrorNotReady) {
                    busy[L->stream] = 1;
                    i++;
                    continue;
                }
                const double t = now_s();
                for (Req* r : L->reqs) {
                    r->done = true;
                    r->t_done = t;
                    if (q != hipSuccess) {
                        r->status = COGCONV_ERR_RUNTIME;
                        r->error = hipGetErrorString(q);
                    } else {
                        completed++;
                        lat_ms.push_back((t - r->t_submit) * 1e3);
                        gflop += r->flops * 1e-9;
                        last_complete = t;
                    }
                }
                for (auto& b : L->blocks) arena.release(b.first, b.second);
                if (L->slot >= 0) free_slots.push_back(L->slot);
                event_pool.push_back(L->ev);
                delete L;
                inflight.erase(inflight.begin() + i);
                cv_done.notify_all();
                cv_work.notify_all();
            }
            // Admissions, in arrival order, while slots and budget allow.
            bool blocked = false;
            // The requests of a launch leave `pending` before it is issued (the
            // other worker must not take them while mu is dropped) and go back
            // to its front if the budget cannot hold them yet.
            while (!pending.empty() && (int)inflight.size() + issuing < cfg.max_inflight) {
                Req* h = pending.front();
                std::vector<Req*> g{h};
                std::vector<size_t> idx{0};
                if (stackable(h->r.desc.mode) && cfg.max_stack > 1 && h->r.bank >= 0 && stack_ok(h)) {
                    // Bucket ids were interned at submit: the scan compares integers,
                    // over a bounded window of the queue.
                    const size_t window = std::min(pending.size(), (size_t)256);
                    for (size_t j = 1; j < window && (int)g.size() < cfg.max_stack; j++) {
                        if (pending[j]->bucket == h->bucket) {
                            g.push_back(pending[j]);
                            idx.push_back(j);
                        }
                    }
                }
                for (size_t k = idx.size(); k-- > 0;) pending.erase(pending.begin() + idx[k]);
                issuing++;
                bool ok = issue(g, lk);
                if (!ok && g.size() > 1) {
                    for (size_t k = g.size(); k-- > 1;) pending.push_front(g[k]);
                    g.resize(1);
                    ok = issue(g, lk);
                }
                issuing--;
                // Waiters on issuing (drain) and on launch slots (other workers).
                cv_done.notify_all();
                cv_work.notify_all();
                if (!ok) {
                    if (inflight.empty() && issuing == 0) {
                        // Nothing holds request memory and the request still does
                        // not fit (the free space is split around packed banks):
                        // it never will, so it fails instead of waiting.
                        fail_reqs(g, COGCONV_ERR_BUDGET, "request does not fit the admission budget");
                        continue;
                    }
                    pending.push_front(g[0]);
                    blocked = true;  // wait for memory
                    break;
                }
            }
            if (!pending.empty() && (int)inflight.size() + issuing >= cfg.max_inflight) blocked = true;
            if (stop && pending.empty() && inflight.empty() && issuing == 0) break;
            if (inflight.empty() && blocked) {
                // Admission is held up by launches another worker is issuing
                // with mu dropped: release mu until that worker notifies (its
                // issuing-- above), or briefly, and try again.
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            } else if (inflight.empty()) {
                cv_work.wait(lk, [&] { return stop || !pending.empty() || !inflight.empty(); });
            } else {
                // Launches in flight: poll again at once (a timed sleep rounds up to
                // the kernel's timer slack, tens of microseconds, which would cap
                // the request rate); yield so submitters can take the lock.
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }
};

extern "C" int cogconv_server_create(const cogconv_server_config* config, cogconv_server* server) {
    clear_error();
    if (!config || !server) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *server = nullptr;
    if (config->hbm_budget_bytes < 4096 || config->max_inflight < 1 || config->streams < 1 || config->max_stack < 1 ||
        config->graph_after < 0)
        return fail(COGCONV_ERR_INVALID_ARG, "bad server configuration");
    int e = cogconv_init(config->device);
    if (e) return e;
    cogconv_server s = new (std::nothrow) cogconv_server_s;
    if (!s) return fail(COGCONV_ERR_OOM, "host allocation failed");
    s->cfg = *config;
    void* base = nullptr;
    e = check_hip(hipMalloc(&base, config->hbm_budget_bytes), "server budget allocation");
    if (e) {
        delete s;
        return e;
    }
    s->arena.base = static_cast<char*>(base);
    s->arena.size = config->hbm_budget_bytes & ~size_t(255);
    s->arena.free_[0] = s->arena.size;
    for (int i = 0; i < config->streams; i++) {
        hipStream_t st;
        e = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream create");
        if (e) {
            (void)cogconv_server_destroy(s);
            return e;
        }
        s->streams.push_back(st);
        e = check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream create");
        if (e) {
            (void)cogconv_server_destroy(s);
            return e;
        }
        s->capture_streams.push_back(st);
    }
    s->stream_issuing.assign(s->streams.size(), 0);
    for (int i = 0; i < config->streams; i++) {
        void* t = nullptr;
        e = check_hip(hipMalloc(&t, sizeof(cogconv_image_io) * config->max_stack), "request table");
        if (e) {
            (void)cogconv_server_destroy(s);
            return e;
        }
        s->dev_tables.push_back(static_cast<cogconv_image_io*>(t));
    }
    {
        void* h = nullptr;
        e = check_hip(hipHostMalloc(&h, sizeof(cogconv_image_io) * config->max_stack * config->max_inflight,
                                    hipHostMallocDefault),
                      "pinned request tables");
        if (e) {
            (void)cogconv_server_destroy(s);
            return e;
        }
        s->host_tables = static_cast<cogconv_image_io*>(h);
        for (int i = config->max_inflight - 1; i >= 0; i--) s->free_slots.push_back(i);
    }
    const int nw = std::min(config->streams, 3);
    for (int i = 0; i < nw; i++) s->workers.emplace_back([s] { s->run(); });
    *server = s;
    return COGCONV_OK;
}

extern "C" int cogconv_server_destroy(cogconv_server s) {
    clear_error();
    if (!s) return COGCONV_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& w : s->workers)
        if (w.joinable()) w.join();
    for (auto& kv : s->graphs)
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto t : s->dev_tables) (void)hipFree(t);
    if (s->host_tables) (void)hipHostFree(s->host_tables);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    for (auto st : s->capture_streams) (void)hipStreamDestroy(st);
    for (auto ev : s->event_pool) (void)hipEventDestroy(ev);
    for (auto& b : s->banks) (void)hipFree(b.dptr);
    for (auto& kv : s->plans) cogconv_plan_destroy(kv.second);
    for (auto& kv : s->reqs) delete kv.second;
    if (s->arena.base) (void)hipFree(s->arena.base);
    delete s;
    return COGCONV_OK;
}

extern "C" int cogconv_server_register_bank(cogconv_server s, const float* host_weights,
                                            const cogconv_field_type* type, int32_t* bank_id) {
    clear_error();
    if (!s || !host_weights || !type || !bank_id) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    const size_t bytes = (size_t)type->rows * type->cols * type->planes * sizeof(float);
    if (!bytes) return fail(COGCONV_ERR_INVALID_ARG, "empty bank");
    float* d = nullptr;
    int e = check_hip(hipMalloc(&d, bytes), "bank allocation");
    if (e) return e;
    e = check_hip(hipMemcpy(d, host_weights, bytes, hipMemcpyHostToDevice), "bank upload");
    if (e) {
        (void)hipFree(d);
        return e;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->banks.push_back({d, *type});
    *bank_id = (int32_t)s->banks.size() - 1;
    return COGCONV_OK;
}

extern "C" int cogconv_server_submit(cogconv_server s, const cogconv_request* request, uint64_t* ticket) {
    clear_error();
    if (!s || !request || !ticket) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    if (!request->host_in0 || !request->host_out) return fail(COGCONV_ERR_INVALID_ARG, "null buffer");
    Req* q = new (std::nothrow) Req;
    if (!q) return fail(COGCONV_ERR_OOM, "host allocation failed");
    q->r = *request;
    cogconv_desc& d = q->r.desc;
    // One critical section on the common path: the request's plan (kept by the
    // server for its life, one per descriptor; made outside the lock the first
    // time) and its shape bucket are resolved here, so that the workers'
    // admission and issue paths compare pointers and integers.
    std::unique_lock<std::mutex> lk(s->mu);
    if (request->bank >= 0) {
        if (request->bank >= (int32_t)s->banks.size()) {
            delete q;
            return fail(COGCONV_ERR_INVALID_ARG, "unknown filter bank %d", request->bank);
        }
        d.in1 = s->banks[request->bank].type;
    } else if (!request->host_in1) {
        delete q;
        return fail(COGCONV_ERR_INVALID_ARG, "request without bank needs host_in1");
    }
    const std::string key(reinterpret_cast<const char*>(&d), sizeof d);
    auto it = s->plans.find(key);
    if (it == s->plans.end()) {
        lk.unlock();
        cogconv_plan p = nullptr;
        const int pe = cogconv_plan_create(&d, &p);  // also rejects an illegal descriptor
        if (pe) {
            delete q;
            return pe;
        }
        lk.lock();
        auto ins = s->plans.emplace(key, p);
        if (!ins.second) cogconv_plan_destroy(p);  // another submitter planned it meanwhile
        it = ins.first;
    }
    q->plan = it->second;
    q->out = q->plan->out;
    q->flops = q->plan->flops;
    q->in0_bytes = (size_t)d.in0.rows * d.in0.cols * d.in0.planes * sizeof(float);
    q->in1_bytes = (size_t)d.in1.rows * d.in1.cols * d.in1.planes * sizeof(float);
    q->out_bytes = (size_t)q->out.rows * q->out.cols * q->out.planes * sizeof(float);
    const size_t ws = q->plan->workspace;
    const bool resident = (request->flags & COGCONV_REQUEST_RESIDENT) != 0;
    const size_t need = resident ? Arena::round(ws)
                                 : Arena::round(q->in0_bytes) + Arena::round(q->out_bytes) + Arena::round(ws) +
                                       (request->bank < 0 ? Arena::round(q->in1_bytes) : 0);
    const size_t room = s->arena.size - s->reserved;
    if (need > room) {
        s->rejected++;
        delete q;
        return fail(COGCONV_ERR_BUDGET, "request needs %zu bytes, budget is %zu (%zu held by packed banks)", need,
                    s->arena.size, s->reserved);
    }
    // Shape bucket: a one-image request's plan and bank name it; a batched
    // request's per-image descriptor does (a string key, the rare case).
    if (d.batch == 1) {
        auto bit = s->plan_buckets.emplace(std::make_pair(q->plan, request->bank), (int)s->bucket_count).first;
        if (bit->second == (int)s->bucket_count) s->bucket_count++;
        q->bucket = bit->second;
    } else {
        auto bit = s->buckets.emplace(bucket_key(q), (int)s->bucket_count).first;
        if (bit->second == (int)s->bucket_count) s->bucket_count++;
        q->bucket = bit->second;
    }
    q->ticket = s->next_ticket++;
    q->t_submit = now_s();
    if (!s->submitted) s->first_submit = q->t_submit;
    s->submitted++;
    s->reqs[q->ticket] = q;
    s->pending.push_back(q);
    *ticket = q->ticket;
    lk.unlock();
    s->cv_work.notify_all();
    return COGCONV_OK;
}

static int ticket_status(cogconv_server s, uint64_t ticket, bool* done) {
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(COGCONV_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second->done;
    if (it->second->done && it->second->status)
        return fail(it->second->status, "request %llu failed: %s", (unsigned long long)ticket,
                    it->second->error.c_str());
    return COGCONV_OK;
}

extern "C" int cogconv_server_query(cogconv_server s, uint64_t ticket) {
    clear_error();
    if (!s) return fail(COGCONV_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    bool done = false;
    int e = ticket_status(s, ticket, &done);
    if (e) return e;
    return done ? COGCONV_OK : COGCONV_ERR_NOT_READY;
}

extern "C" int cogconv_server_wait(cogconv_server s, uint64_t ticket) {
    clear_error();
    if (!s) return fail(COGCONV_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    bool done = false;
    int e = ticket_status(s, ticket, &done);
    if (e) return e;
    // The request stays alive while it has waiters (stats_reset skips it).
    Req* q = s->reqs.find(ticket)->second;
    q->waiters++;
    s->cv_done.wait(lk, [&] { return q->done; });
    q->waiters--;
    return ticket_status(s, ticket, &done);
}

extern "C" int cogconv_server_drain(cogconv_server s) {
    clear_error();
    if (!s) return fail(COGCONV_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight.empty() && s->issuing == 0; });
    return COGCONV_OK;
}

extern "C" int cogconv_server_stats_get(cogconv_server s, cogconv_server_stats* out) {
    clear_error();
    if (!s || !out) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->rejected = s->rejected;
    out->launches = s->launches;
    out->stacked = s->stacked;
    out->graph_launches = s->graph_launches;
    out->first_submit_s = s->first_submit;
    out->last_complete_s = s->last_complete;
    out->gflop = s->gflop;
    out->peak_bytes = s->arena.peak;
    if (!s->lat_ms.empty()) {
        std::vector<double> v = s->lat_ms;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double pos = p * (v.size() - 1);
            const size_t i = (size_t)pos;
            const double f = pos - i;
            return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_ms = sum / v.size();
        out->max_ms = v.back();
        const double span = s->last_complete - s->first_submit;
        out->requests_per_s = span > 0 ? s->completed / span : 0;
    }
    return COGCONV_OK;
}

extern "C" int cogconv_server_stats_reset(cogconv_server s) {
    clear_error();
    if (!s) return fail(COGCONV_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_ms.clear();
    s->submitted = s->completed = s->rejected = s->launches = s->stacked = s->graph_launches = 0;
    s->first_submit = s->last_complete = s->gflop = 0;
    s->arena.peak = s->arena.in_use;
    for (auto it = s->reqs.begin(); it != s->reqs.end();) {
        if (it->second->done && it->second->waiters == 0) {
            delete it->second;
            it = s->reqs.erase(it);
        } else {
            ++it;
        }
    }
    return COGCONV_OK;
}

extern "C" int cogconv_server_submit_timed(cogconv_server s, const cogconv_request* requests, const double* arrival_s,
                                           int64_t n, uint64_t* tickets) {
    clear_error();
    if (!s || (n > 0 && (!requests || !arrival_s || !tickets))) return fail(COGCONV_ERR_INVALID_ARG, "null argument");
    const double t0 = now_s();
    int first_error = COGCONV_OK;
    std::string msg;
    for (int64_t i = 0; i < n; i++) {
        const double due = t0 + arrival_s[i];
        for (;;) {
            const double left = due - now_s();
            if (left <= 0) break;
            if (left > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds((long)((left - 100e-6) * 1e6)));
        }
        const int e = cogconv_server_submit(s, &requests[i], &tickets[i]);
        if (e) {
            tickets[i] = 0;
            if (!first_error) {
                first_error = e;
                msg = cogconv_last_error();
            }
        }
    }
    if (first_error) return fail(first_error, "%s", msg.c_str());
    return COGCONV_OK;
}
