// Request server over the FMM-FFT plans: continuous batching of same-shape
// requests, plan / workspace cache under an HBM budget, hipGraph replay of
// the execute launch sequence, latency statistics (include/fmmfft.h,
// fmmfft_server_*).  The reference has no counterpart (it runs one transform
// per MPI job, src/fmmfft.c:1192-1288).
//
// One scheduler thread owns every shape, slot, plan and graph, and makes
// every HIP call of the server except the submit-time event record.  The
// server mutex guards only what submitters and waiters share with it (the
// pending queue, the tickets, the statistics): the scheduler picks a batch
// under the lock and makes its HIP calls (plan creation, allocation, graph
// capture, launches) with the lock released.
//
// A request carries caller-owned device buffers and the caller's stream: an
// event recorded on that stream at submit time is waited for on the slot
// stream before the request's input is read.  A batch of same-shape
// requests (up to kBatchElems points in all) runs in place on the callers'
// buffers: the slot's device table of per-request input and output
// addresses is refreshed from pinned host memory by one small copy, and the
// batched launch sequence, which reads every signal's addresses from that
// table, is captured once per (slot, count) as a hipGraph and replayed.
// Each shape owns two slots (two workspaces and tables), so one batch is
// set up while the previous one computes, and batches of different shapes
// run on different streams and overlap on the device, up to
// kMaxBatchesInFlight batches at once.  The scheduler launches the oldest
// pending request whose shape has a free slot (a busy shape does not hold
// back the others).
// Completion times come from a timing event recorded after each batch,
// measured against a reference event whose host time is known, so they do
// not depend on when the scheduler notices them.
#include "../../include/fmmfft.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace {

constexpr int64_t kBatchElems = int64_t(1) << 23;  // points per batch (all requests)
constexpr int kSlots = 2;
// Batches on the device at once, all shapes together.  More only queue
// behind the few hardware queues, while a request that waits in the
// server's queue instead joins its shape's next batch: with this cap the
// mixed stream of 52 shapes ran 9.5k req/s at saturation against 8.5k with
// every shape's two slots in use (3.6k against 7.6k launches per 10k
// requests).
constexpr int kMaxBatchesInFlight = 16;

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
    bool operator==(const Key &o) const { return !(*this < o) && !(o < *this); }
};

struct Req {
    uint64_t id;
    Key key;
    const void *in;
    void *out;
    hipEvent_t ready = nullptr;  // recorded on the caller's stream at submit
    double t_submit = 0, t_done = 0;
    int status = FMMFFT_OK;
    bool done = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;  // timing event after the batch's last operation
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // per-request addresses, inputs [0, maxb) then outputs [maxb, 2 maxb):
    // the device table the kernels read and its pinned host staging
    void **tab = nullptr, **htab = nullptr;
    bool busy = false;
    std::vector<std::shared_ptr<Req>> batch;
    std::map<int, hipGraphExec_t> graphs;  // batch count -> executable graph
};

struct Shape {
    Key key;
    fmmfft_plan plan = nullptr;
    int maxb = 1;
    size_t esize = 8, bytes = 0;
    Slot slots[kSlots];
    double last_used = 0;
    Slot *free_slot() {
        for (auto &s : slots)
            if (!s.busy) return &s;
        return nullptr;
    }
};

}  // namespace

struct fmmfft_server_s {
    size_t budget;
    std::atomic<size_t> used{0}, peak{0};
    int max_batch;
    int device = 0;
    // shared with submitters and waiters (guarded by mu)
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Req>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> reqs;
    std::vector<hipEvent_t> event_pool;
    uint64_t next_id = 1;
    bool stop = false;
    int64_t inflight = 0;
    int busy_slots = 0;  // batches on the device (scheduler thread only)
    std::vector<double> lat_ms;
    double first_submit = 0, last_done = 0, gflop = 0;
    int64_t batches = 0, plans_created = 0, plans_evicted = 0, graphs_captured = 0;
    // owned by the scheduler thread
    std::map<Key, std::unique_ptr<Shape>> shapes;
    hipEvent_t ref_ev = nullptr;  // completed reference event ...
    double ref_t = 0;             // ... and the host time it completed at
    std::thread worker;

    void run();
    bool launch_one();
    bool poll_completions();
    void finish(const std::vector<std::shared_ptr<Req>> &batch, int status);
    Shape *get_shape(const Key &k, int &err);
    void free_shape(Shape &sh);
    bool evict_one(const Shape *keep);
};

namespace {

// device memory the server holds for a shape: the plan's tables, and per
// slot the workspace for maxb signals and the address table
size_t shape_bytes(const Key &k, int maxb, size_t *ws_bytes) {
    size_t tb = 0, wb = 0;
    if (fmmfft_plan_sizes(k.n, k.p, k.terms, k.leaf, k.dtype, &tb, &wb) != FMMFFT_OK) return 0;
    *ws_bytes = wb * (size_t)maxb;
    return tb + kSlots * (*ws_bytes + 2 * sizeof(void *) * (size_t)maxb);
}

}  // namespace

void fmmfft_server_s::free_shape(Shape &sh) {
    for (auto &sl : sh.slots) {
        for (auto &g : sl.graphs) (void)hipGraphExecDestroy(g.second);
        sl.graphs.clear();
        if (sl.ws) (void)hipFree(sl.ws);
        if (sl.tab) (void)hipFree(sl.tab);
        if (sl.htab) (void)hipHostFree(sl.htab);
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
    std::lock_guard<std::mutex> g(mu);
    ++plans_evicted;
    return true;
}

// Plan (tables only) + workspaces and address tables for a shape, created
// under the budget.  err: FMMFFT_OK, FMMFFT_ENOMEM / EINVAL / ENOTSUP (cannot
// run even alone: reject) or -1 (does not fit now: wait for completions).
Shape *fmmfft_server_s::get_shape(const Key &k, int &err) {
    err = FMMFFT_OK;
    auto it = shapes.find(k);
    if (it != shapes.end()) return it->second.get();
    int maxb = (int)std::max<int64_t>(1, std::min<int64_t>(max_batch, kBatchElems / k.n));
    size_t wsb = 0;
    size_t need = shape_bytes(k, maxb, &wsb);
    while (need > budget && maxb > 1) {
        maxb /= 2;
        need = shape_bytes(k, maxb, &wsb);
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
    sh->esize = k.dtype == FMMFFT_C128 ? 16 : 8;
    // tables only: every execute passes a slot workspace
    int rc = fmmfft::plan_create_impl(&sh->plan, k.n, k.p, k.terms, k.leaf, k.dtype, maxb, false);
    bool ok = rc == FMMFFT_OK;
    for (auto &sl : sh->slots) {
        if (!ok) break;
        sl.ws_bytes = wsb;
        ok = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreate(&sl.ev) == hipSuccess && hipMalloc(&sl.ws, wsb) == hipSuccess &&
             hipMalloc((void **)&sl.tab, 2 * sizeof(void *) * (size_t)maxb) == hipSuccess &&
             hipHostMalloc((void **)&sl.htab, 2 * sizeof(void *) * (size_t)maxb, hipHostMallocDefault) == hipSuccess;
    }
    sh->bytes = need;
    used += need;
    if (!ok) {
        free_shape(*sh);
        err = rc == FMMFFT_EINVAL || rc == FMMFFT_ENOTSUP ? rc : FMMFFT_ENOMEM;
        return nullptr;
    }
    size_t pk = peak.load();
    while (used > pk && !peak.compare_exchange_weak(pk, used.load())) {
    }
    {
        std::lock_guard<std::mutex> g(mu);
        ++plans_created;
    }
    Shape *raw = sh.get();
    shapes[k] = std::move(sh);
    return raw;
}

// Mark a batch finished (status of every request) and wake the waiters.
// Failed requests complete at the current time; completed batches carry
// their event-measured completion time already.
void fmmfft_server_s::finish(const std::vector<std::shared_ptr<Req>> &batch, int status) {
    const double t = now_s();
    std::lock_guard<std::mutex> g(mu);
    inflight -= (int64_t)batch.size();
    for (auto &r : batch) {
        if (r->ready) event_pool.push_back(r->ready);
        r->ready = nullptr;
        r->status = status;
        if (status != FMMFFT_OK || r->t_done == 0) r->t_done = t;
        r->done = true;
        if (status == FMMFFT_OK) {
            lat_ms.push_back(std::max(0.0, r->t_done - r->t_submit) * 1e3);
            gflop += 5.0 * (double)r->key.n * std::log2((double)r->key.n) * 1e-9;
            last_done = std::max(last_done, r->t_done);
        }
    }
    cv_done.notify_all();
}

// Retire every finished batch (completion time from its timing event);
// returns true if any finished.
bool fmmfft_server_s::poll_completions() {
    bool any = false;
    for (auto &kv : shapes)
        for (auto &sl : kv.second->slots) {
            if (!sl.busy || hipEventQuery(sl.ev) != hipSuccess) continue;
            float ms = 0.f;
            double t = now_s();
            if (hipEventElapsedTime(&ms, ref_ev, sl.ev) == hipSuccess) t = ref_t + 1e-3 * (double)ms;
            for (auto &r : sl.batch) r->t_done = t;
            auto batch = std::move(sl.batch);
            sl.batch.clear();
            sl.busy = false;
            --busy_slots;
            kv.second->last_used = now_s();
            finish(batch, FMMFFT_OK);
            any = true;
        }
    return any;
}

// Launch one batch: the oldest pending request whose shape has a free slot
// (or has no plan yet), with up to maxb requests of that shape in arrival
// order.  Returns false if nothing could start.
bool fmmfft_server_s::launch_one() {
    std::vector<std::shared_ptr<Req>> batch;
    Key key{};
    bool have = false;
    {
        std::lock_guard<std::mutex> g(mu);
        if (busy_slots >= kMaxBatchesInFlight) return false;
        std::vector<Key> busy;  // shapes seen without a free slot
        for (auto it = pending.begin(); it != pending.end() && !have; ++it) {
            const Key &k = (*it)->key;
            if (std::find(busy.begin(), busy.end(), k) != busy.end()) continue;
            auto sit = shapes.find(k);
            if (sit == shapes.end() || sit->second->free_slot()) {
                key = k;
                have = true;
            } else {
                busy.push_back(k);
            }
        }
        if (!have) return false;
        auto sit = shapes.find(key);
        const int maxb = sit == shapes.end() ? 1 : sit->second->maxb;
        // (a shape without a plan yet: its first request now, the rest
        // once the plan exists).  Taken requests count as in flight until
        // they finish or go back to the queue.
        for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < maxb;) {
            if ((*it)->key == key) {
                batch.push_back(*it);
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        inflight += (int64_t)batch.size();
    }
    int err = FMMFFT_OK;
    Shape *sh = get_shape(key, err);
    if (!sh) {
        if (err == -1) {
            // no room now: back to the front of the queue, in order
            std::lock_guard<std::mutex> g(mu);
            for (auto it = batch.rbegin(); it != batch.rend(); ++it) pending.push_front(*it);
            inflight -= (int64_t)batch.size();
            return false;
        }
        finish(batch, err);
        return true;
    }
    if ((int)batch.size() < sh->maxb) {
        // a new shape: fill the batch up to its real size
        std::lock_guard<std::mutex> g(mu);
        const size_t before = batch.size();
        for (auto it = pending.begin(); it != pending.end() && (int)batch.size() < sh->maxb;) {
            if ((*it)->key == key) {
                batch.push_back(*it);
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        inflight += (int64_t)(batch.size() - before);
    }
    Slot *sl = sh->free_slot();
    const int cnt = (int)batch.size();
    int rc = FMMFFT_OK;
    // inputs are complete on the callers' streams
    for (auto &r : batch)
        if (r->ready && hipStreamWaitEvent(sl->st, r->ready, 0) != hipSuccess) rc = FMMFFT_EHIP;
    // the batch's addresses into the slot's table (the pinned staging is
    // rewritten only once the slot's previous batch, and so its copy, ended)
    for (int i = 0; i < cnt; ++i) {
        sl->htab[i] = const_cast<void *>(batch[i]->in);
        sl->htab[sh->maxb + i] = batch[i]->out;
    }
    if (rc == FMMFFT_OK && hipMemcpyAsync(sl->tab, sl->htab, 2 * sizeof(void *) * (size_t)sh->maxb,
                                          hipMemcpyHostToDevice, sl->st) != hipSuccess)
        rc = FMMFFT_EHIP;
    auto gr = sl->graphs.find(cnt);
    if (rc == FMMFFT_OK && gr == sl->graphs.end()) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamBeginCapture(sl->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            int erc = fmmfft::execute_table(sh->plan, cnt, sl->tab, sl->tab + sh->maxb, key.dir, sl->ws,
                                            sl->ws_bytes, sl->st);
            hipError_t ce = hipStreamEndCapture(sl->st, &graph);
            if (erc == FMMFFT_OK && ce == hipSuccess &&
                hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                sl->graphs[cnt] = exec;
                std::lock_guard<std::mutex> lk(mu);
                ++graphs_captured;
            } else {
                rc = FMMFFT_EHIP;
            }
            if (graph) (void)hipGraphDestroy(graph);
        } else {
            rc = FMMFFT_EHIP;
        }
        gr = sl->graphs.find(cnt);
    }
    if (rc == FMMFFT_OK && gr != sl->graphs.end() && hipGraphLaunch(gr->second, sl->st) != hipSuccess)
        rc = FMMFFT_EHIP;
    if (rc == FMMFFT_OK && hipEventRecord(sl->ev, sl->st) != hipSuccess) rc = FMMFFT_EHIP;
    if (rc != FMMFFT_OK) {
        (void)hipStreamSynchronize(sl->st);
        finish(batch, rc);
        return true;
    }
    sl->batch = std::move(batch);
    sl->busy = true;
    ++busy_slots;
    sh->last_used = now_s();
    std::lock_guard<std::mutex> lk(mu);
    ++batches;
    return true;
}

void fmmfft_server_s::run() {
    (void)hipSetDevice(device);
    // reference event for completion times: its host time is taken right
    // after it completes
    if (hipEventCreate(&ref_ev) == hipSuccess && hipEventRecord(ref_ev, nullptr) == hipSuccess &&
        hipEventSynchronize(ref_ev) == hipSuccess)
        ref_t = now_s();
    while (true) {
        const bool done_any = poll_completions();
        bool launched = false;
        while (launch_one()) launched = true;
        std::unique_lock<std::mutex> lk(mu);
        if (stop && pending.empty() && inflight == 0) break;
        if (launched || done_any) continue;
        if (inflight > 0)
            // batches in flight: look again shortly, or as soon as a request
            // arrives
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        else
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
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
                         int32_t dtype, int32_t dir, const void *d_in, void *d_out, void *stream,
                         uint64_t *ticket) {
    if (!srv || !d_in || !d_out || !ticket) return FMMFFT_EINVAL;
    if (dir != FMMFFT_FORWARD && dir != FMMFFT_INVERSE) return FMMFFT_EINVAL;
    int rc = fmmfft_plan_sizes(n, p_blocks, terms, leaf, dtype, nullptr, nullptr);
    if (rc != FMMFFT_OK) return rc;
    // the alignment fmmfft_execute_ws asks of its buffers
    const uintptr_t elem = dtype == FMMFFT_C128 ? 16 : 8;  // bytes of one complex point
    if ((uintptr_t)d_in % 16 != 0 || (uintptr_t)d_out % elem != 0) return FMMFFT_EINVAL;
    auto r = std::make_shared<Req>();
    r->key = Key{n, p_blocks, terms, leaf, dtype, dir};
    r->in = d_in;
    r->out = d_out;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->stop) return FMMFFT_EINVAL;
        if (!srv->event_pool.empty()) {
            r->ready = srv->event_pool.back();
            srv->event_pool.pop_back();
        }
    }
    if (!r->ready && hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) != hipSuccess) return FMMFFT_EHIP;
    if (hipEventRecord(r->ready, static_cast<hipStream_t>(stream)) != hipSuccess) {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->event_pool.push_back(r->ready);
        return FMMFFT_EHIP;
    }
    {
        std::lock_guard<std::mutex> g(srv->mu);
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
    if (latency_ms) *latency_ms = std::max(0.0, r->t_done - r->t_submit) * 1e3;
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
    for (hipEvent_t e : srv->event_pool) (void)hipEventDestroy(e);
    if (srv->ref_ev) (void)hipEventDestroy(srv->ref_ev);
    delete srv;
    return FMMFFT_OK;
}

}  // extern "C"
