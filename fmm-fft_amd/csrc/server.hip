// Request server over the FMM-FFT plans: continuous batching of same-shape
// requests, plan / workspace cache under an HBM budget, hipGraph replay of
// the execute launch sequence, latency statistics (include/fmmfft.h,
// fmmfft_server_*).  The reference has no counterpart (it runs one transform
// per MPI job, src/fmmfft.c:1192-1288).
//
// One scheduler thread owns every HIP call of the server.  Requests carry
// caller-owned device buffers.  Shapes with N <= kSlotElems are gathered
// into a contiguous slot (one device-to-device copy per request), executed
// as one batched launch sequence captured once per (slot, count, direction)
// as a hipGraph, and scattered back; larger shapes execute directly on the
// caller's buffers.  Each shape owns two slots (two workspaces), so one
// batch can be filled and copied while the previous one computes; different
// shapes run on different streams and overlap on the device.
#include "../../include/fmmfft.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace {

constexpr int64_t kSlotElems = int64_t(1) << 20;   // gather/graph path up to this N
constexpr int64_t kBatchElems = int64_t(1) << 22;  // elements per gathered batch
constexpr int kSlots = 2;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Key {
    int64_t n;
    int32_t p, terms, leaf, dtype, dir;
    bool operator<(const Key &o) const {
        return std::tie(n, p, terms, leaf, dtype, dir) < std::tie(o.n, o.p, o.terms, o.leaf, o.dtype, o.dir);
    }
};

struct Req {
    uint64_t id;
    Key key;
    const void *in;
    void *out;
    double t_submit = 0, t_done = 0;
    int status = FMMFFT_OK;
    bool done = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    void *ws = nullptr, *buf_in = nullptr, *buf_out = nullptr;
    size_t ws_bytes = 0;
    bool busy = false;
    std::vector<std::shared_ptr<Req>> batch;
    std::map<int, hipGraphExec_t> graphs;  // batch count -> executable graph
};

struct Shape {
    Key key;
    fmmfft_plan plan = nullptr;
    int maxb = 1;
    bool gathered = false;
    size_t esize = 8, bytes = 0;
    Slot slots[kSlots];
    double last_used = 0;
};

}  // namespace

struct fmmfft_server_s {
    size_t budget, used = 0, peak = 0;
    int max_batch;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> reqs;
    std::map<Key, std::unique_ptr<Shape>> shapes;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    int device = 0;
    // statistics
    std::vector<double> lat_ms;
    double first_submit = 0, last_done = 0, gflop = 0;
    int64_t batches = 0, plans_created = 0, plans_evicted = 0, graphs_captured = 0, inflight = 0;

    void run();
    bool try_launch(std::unique_lock<std::mutex> &lk);
    void retire(Shape &sh, Slot &sl);
    bool poll_completions();
    Shape *get_shape(const Key &k, int &err);
    void free_shape(Shape &sh);
    bool evict_one(const Shape *keep);
};

namespace {

size_t shape_bytes(const Key &k, int maxb, bool gathered, size_t *ws_bytes) {
    size_t tb = 0, wb = 0;
    if (fmmfft_plan_sizes(k.n, k.p, k.terms, k.leaf, k.dtype, &tb, &wb) != FMMFFT_OK) return 0;
    const size_t esize = k.dtype == FMMFFT_C128 ? 16 : 8;
    *ws_bytes = wb * (size_t)maxb;
    size_t per_slot = *ws_bytes;
    if (gathered) per_slot += 2 * (size_t)maxb * (size_t)k.n * esize;
    return tb + kSlots * per_slot;
}

}  // namespace

void fmmfft_server_s::free_shape(Shape &sh) {
    for (auto &sl : sh.slots) {
        for (auto &g : sl.graphs) (void)hipGraphExecDestroy(g.second);
        sl.graphs.clear();
        if (sl.ws) (void)hipFree(sl.ws);
        if (sl.buf_in) (void)hipFree(sl.buf_in);
        if (sl.buf_out) (void)hipFree(sl.buf_out);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        sl = Slot{};
    }
    if (sh.plan) fmmfft_plan_destroy(sh.plan);
    sh.plan = nullptr;
    used -= sh.bytes;
}

bool fmmfft_server_s::evict_one(const Shape *keep) {
    Shape *victim = nullptr;
    for (auto &kv : shapes) {
        Shape *s = kv.second.get();
        if (s == keep) continue;
        bool idle = true;
        for (auto &sl : s->slots) idle = idle && !sl.busy;
        if (idle && (!victim || s->last_used < victim->last_used)) victim = s;
    }
    if (!victim) return false;
    free_shape(*victim);
    shapes.erase(victim->key);
    ++plans_evicted;
    return true;
}

// Plan + workspaces (+ gather buffers) for a shape, created under the budget.
// err: FMMFFT_OK, FMMFFT_ENOMEM (does not fit even alone: reject) or -1
// (does not fit now: wait for completions).
Shape *fmmfft_server_s::get_shape(const Key &k, int &err) {
    err = FMMFFT_OK;
    auto it = shapes.find(k);
    if (it != shapes.end()) return it->second.get();
    const bool gathered = k.n <= kSlotElems;
    int maxb = gathered ? (int)std::max<int64_t>(1, std::min<int64_t>(max_batch, kBatchElems / k.n)) : 1;
    size_t wsb = 0;
    size_t need = shape_bytes(k, maxb, gathered, &wsb);
    while (need > budget && maxb > 1) {
        maxb /= 2;
        need = shape_bytes(k, maxb, gathered, &wsb);
    }
    if (need == 0 || need > budget) {
        err = FMMFFT_ENOMEM;
        return nullptr;
    }
    while (used + need > budget)
        if (!evict_one(nullptr)) {
            err = -1;
            return nullptr;
        }
    auto sh = std::make_unique<Shape>();
    sh->key = k;
    sh->maxb = maxb;
    sh->gathered = gathered;
    sh->esize = k.dtype == FMMFFT_C128 ? 16 : 8;
    int rc = fmmfft_plan_create(&sh->plan, k.n, k.p, k.terms, k.leaf, k.dtype, 1);
    bool ok = rc == FMMFFT_OK;
    for (auto &sl : sh->slots) {
        if (!ok) break;
        sl.ws_bytes = wsb;
        ok = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) == hipSuccess &&
             hipMalloc(&sl.ws, wsb) == hipSuccess;
        if (ok && gathered) {
            const size_t bb = (size_t)maxb * (size_t)k.n * sh->esize;
            ok = hipMalloc(&sl.buf_in, bb) == hipSuccess && hipMalloc(&sl.buf_out, bb) == hipSuccess;
        }
    }
    sh->bytes = need;
    used += need;
    if (!ok) {
        free_shape(*sh);
        err = rc == FMMFFT_EINVAL || rc == FMMFFT_ENOTSUP ? rc : FMMFFT_ENOMEM;
        return nullptr;
    }
    peak = std::max(peak, used);
    ++plans_created;
    Shape *raw = sh.get();
    shapes[k] = std::move(sh);
    return raw;
}

void fmmfft_server_s::retire(Shape &sh, Slot &sl) {
    const double t = now_s();
    for (auto &r : sl.batch) {
        r->t_done = t;
        r->done = true;
        lat_ms.push_back((r->t_done - r->t_submit) * 1e3);
        gflop += 5.0 * (double)r->key.n * std::log2((double)r->key.n) * 1e-9;
    }
    last_done = t;
    inflight -= (int64_t)sl.batch.size();
    sl.batch.clear();
    sl.busy = false;
    sh.last_used = t;
}

// Retire every finished batch; returns true if any finished.
bool fmmfft_server_s::poll_completions() {
    bool any = false;
    for (auto &kv : shapes)
        for (auto &sl : kv.second->slots)
            if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) {
                retire(*kv.second, sl);
                any = true;
            }
    if (any) cv_done.notify_all();
    return any;
}

// Launch one batch: the oldest pending request's shape, up to maxb requests
// of that shape in arrival order.  Returns false if nothing could start.
bool fmmfft_server_s::try_launch(std::unique_lock<std::mutex> &lk) {
    (void)lk;
    if (pending.empty()) return false;
    const Key key = pending.front()->key;
    int err = FMMFFT_OK;
    Shape *sh = get_shape(key, err);
    if (!sh) {
        if (err == -1) return false;  // wait for memory
        auto r = pending.front();
        pending.pop_front();
        r->status = err;
        r->done = true;
        cv_done.notify_all();
        return true;
    }
    Slot *sl = nullptr;
    for (auto &s : sh->slots)
        if (!s.busy) {
            sl = &s;
            break;
        }
    if (!sl) return false;
    // gather the batch
    std::vector<std::shared_ptr<Req>> batch;
    for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < sh->maxb;) {
        if (!((*it)->key < key) && !(key < (*it)->key)) {
            batch.push_back(*it);
            it = pending.erase(it);
        } else {
            ++it;
        }
    }
    const int cnt = (int)batch.size();
    const size_t sig_bytes = (size_t)key.n * sh->esize;
    int rc = FMMFFT_OK;
    if (sh->gathered) {
        for (int i = 0; i < cnt; ++i)
            if (hipMemcpyAsync((char *)sl->buf_in + i * sig_bytes, batch[i]->in, sig_bytes, hipMemcpyDeviceToDevice,
                               sl->st) != hipSuccess)
                rc = FMMFFT_EHIP;
        auto g = sl->graphs.find(cnt);
        if (g == sl->graphs.end()) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            if (hipStreamBeginCapture(sl->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                int erc = fmmfft_execute_ws(sh->plan, cnt, sl->buf_in, sl->buf_out, key.dir, sl->ws,
                                            sl->ws_bytes, sl->st);
                hipError_t ce = hipStreamEndCapture(sl->st, &graph);
                if (erc == FMMFFT_OK && ce == hipSuccess &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                    sl->graphs[cnt] = exec;
                    ++graphs_captured;
                } else {
                    rc = FMMFFT_EHIP;
                }
                if (graph) (void)hipGraphDestroy(graph);
            } else {
                rc = FMMFFT_EHIP;
            }
            g = sl->graphs.find(cnt);
        }
        if (rc == FMMFFT_OK && g != sl->graphs.end() && hipGraphLaunch(g->second, sl->st) != hipSuccess)
            rc = FMMFFT_EHIP;
        for (int i = 0; i < cnt && rc == FMMFFT_OK; ++i)
            if (hipMemcpyAsync(batch[i]->out, (char *)sl->buf_out + i * sig_bytes, sig_bytes,
                               hipMemcpyDeviceToDevice, sl->st) != hipSuccess)
                rc = FMMFFT_EHIP;
    } else {
        rc = fmmfft_execute_ws(sh->plan, 1, batch[0]->in, batch[0]->out, key.dir, sl->ws, sl->ws_bytes, sl->st);
    }
    if (rc == FMMFFT_OK && hipEventRecord(sl->ev, sl->st) != hipSuccess) rc = FMMFFT_EHIP;
    if (rc != FMMFFT_OK) {
        (void)hipStreamSynchronize(sl->st);
        for (auto &r : batch) {
            r->status = rc;
            r->done = true;
        }
        cv_done.notify_all();
        return true;
    }
    sl->batch = std::move(batch);
    sl->busy = true;
    sh->last_used = now_s();
    inflight += cnt;
    ++batches;
    return true;
}

void fmmfft_server_s::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        poll_completions();
        bool launched = false;
        while (try_launch(lk)) launched = true;
        if (stop && pending.empty() && inflight == 0) break;
        if (launched) continue;
        if (inflight > 0) {
            // nothing can start: wait for the oldest in-flight batch
            hipEvent_t ev = nullptr;
            for (auto &kv : shapes)
                for (auto &sl : kv.second->slots)
                    if (sl.busy && !ev) ev = sl.ev;
            lk.unlock();
            if (ev) (void)hipEventSynchronize(ev);
            lk.lock();
        } else {
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

int fmmfft_server_create(fmmfft_server *out, size_t hbm_budget_bytes, int32_t max_batch) {
    if (!out || hbm_budget_bytes == 0 || max_batch < 1 || max_batch > 4096) return FMMFFT_EINVAL;
    *out = nullptr;
    auto *s = new (std::nothrow) fmmfft_server_s();
    if (!s) return FMMFFT_ENOMEM;
    s->budget = hbm_budget_bytes;
    s->max_batch = max_batch;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return FMMFFT_EHIP;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FMMFFT_OK;
}

int fmmfft_server_submit(fmmfft_server srv, int64_t n, int32_t p_blocks, int32_t terms, int32_t leaf,
                         int32_t dtype, int32_t dir, const void *d_in, void *d_out, uint64_t *ticket) {
    if (!srv || !d_in || !d_out || !ticket) return FMMFFT_EINVAL;
    if (dir != FMMFFT_FORWARD && dir != FMMFFT_INVERSE) return FMMFFT_EINVAL;
    int rc = fmmfft_plan_sizes(n, p_blocks, terms, leaf, dtype, nullptr, nullptr);
    if (rc != FMMFFT_OK) return rc;
    auto r = std::make_shared<Req>();
    r->key = Key{n, p_blocks, terms, leaf, dtype, dir};
    r->in = d_in;
    r->out = d_out;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->stop) return FMMFFT_EINVAL;
        r->id = srv->next_id++;
        r->t_submit = now_s();
        if (srv->first_submit == 0) srv->first_submit = r->t_submit;
        srv->reqs[r->id] = r;
        srv->pending.push_back(r);
        *ticket = r->id;
    }
    srv->cv_work.notify_one();
    return FMMFFT_OK;
}

int fmmfft_server_wait(fmmfft_server srv, uint64_t ticket, double *latency_ms) {
    if (!srv) return FMMFFT_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return FMMFFT_EINVAL;
    auto r = it->second;
    srv->cv_done.wait(lk, [&] { return r->done; });
    srv->reqs.erase(ticket);
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    return r->status;
}

int fmmfft_server_drain(fmmfft_server srv) {
    if (!srv) return FMMFFT_EINVAL;
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->pending.empty() && srv->inflight == 0; });
    return FMMFFT_OK;
}

int fmmfft_server_stats_get(fmmfft_server srv, fmmfft_server_stats *st) {
    if (!srv || !st) return FMMFFT_EINVAL;
    std::lock_guard<std::mutex> g(srv->mu);
    std::vector<double> v = srv->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
        if (v.empty()) return 0.0;
        size_t i = (size_t)std::ceil(q / 100.0 * (double)v.size());
        return v[std::min(v.size() - 1, i == 0 ? 0 : i - 1)];
    };
    double sum = 0;
    for (double x : v) sum += x;
    st->completed = (int64_t)v.size();
    st->batches = srv->batches;
    st->plans_created = srv->plans_created;
    st->plans_evicted = srv->plans_evicted;
    st->graphs_captured = srv->graphs_captured;
    st->p50_ms = pct(50);
    st->p99_ms = pct(99);
    st->mean_ms = v.empty() ? 0 : sum / (double)v.size();
    st->max_ms = v.empty() ? 0 : v.back();
    st->span_s = srv->last_done > srv->first_submit ? srv->last_done - srv->first_submit : 0;
    st->nominal_gflop = srv->gflop;
    st->bytes_in_use = srv->used;
    st->peak_bytes = srv->peak;
    return FMMFFT_OK;
}

int fmmfft_server_reset_stats(fmmfft_server srv) {
    if (!srv) return FMMFFT_EINVAL;
    std::lock_guard<std::mutex> g(srv->mu);
    srv->lat_ms.clear();
    srv->first_submit = srv->last_done = srv->gflop = 0;
    srv->batches = 0;
    return FMMFFT_OK;
}

int fmmfft_server_destroy(fmmfft_server srv) {
    if (!srv) return FMMFFT_EINVAL;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    for (auto &kv : srv->shapes) srv->free_shape(*kv.second);
    srv->shapes.clear();
    delete srv;
    return FMMFFT_OK;
}

}  // extern "C"
