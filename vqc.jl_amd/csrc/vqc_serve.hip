// vqc_serve.hip -- continuous-batching request loop for energy + adjoint-gradient
// requests of mixed size (no counterpart in the reference, which evaluates one
// circuit per call, src/circuitdiff.jl:2-8 and src/expecdiff.jl:28).
//
// One scheduler thread owns every HIP call after creation:
//   * requests wait in a FIFO; a free lane (its own HIP stream) takes the head
//     and every later pending request of a similar cost that fits the HBM
//     budget, plans them into one persistent-kernel batch and launches it;
//   * a request of at least `solo_qubits` qubits fills the chip by itself: it
//     runs alone on its lane through the tiled phase kernels (one launch per
//     phase), whose short launches interleave with the other lanes' work;
//   * lanes run concurrently: a batch's persistent grid can be capped to a
//     share of the resident-workgroup capacity (grid_frac), so that a batch
//     launched while another runs finds free CUs;
//   * a finished lane (its event passed) hands results to the waiting callers,
//     returns its memory to the caching allocator and takes the next batch.
// Device memory comes from a caching allocator bounded by the budget (blocks
// are reused, not freed, so no hipFree stalls the other lanes).
#include <algorithm>
#include <cstdio>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vqc_batch_internal.h"
#include "vqc_tile.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    vqc_instance_t inst{};
    std::vector<vqc_op_t> ops;
    std::vector<vqc_term_t> terms;
    std::vector<double> pool, hpool;
    uint64_t bytes = 0;  // device bytes while running (states, program, partials)
    double cost = 0;     // relative work estimate
    double t_submit = 0, t_launch = 0, t_done = 0;
    int status = 0;  // 0 pending, 1 running, 2 done, < 0 error code
    std::string msg;
    double energy[2] = {0, 0};
    std::vector<double> grad;
};

// Caching device allocator: blocks are kept for reuse; a request is served by
// the smallest cached block of at least its size and at most twice it, else by
// hipMalloc. Reserved bytes (live + cached) stay within the budget; cached
// blocks are released only when a new allocation would exceed it.
struct DevCache {
    uint64_t budget = 0;
    uint64_t reserved = 0, live = 0, peak_live = 0;
    std::multimap<size_t, void*> cached;
    std::unordered_map<void*, size_t> sizes;

    static void* alloc_cb(void* ctx, size_t bytes) { return static_cast<DevCache*>(ctx)->alloc(bytes); }
    static void free_cb(void* ctx, void* p) { static_cast<DevCache*>(ctx)->release(p); }

    void* alloc(size_t bytes) {
        bytes = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
        auto it = cached.lower_bound(bytes);
        if (it != cached.end() && it->first <= 2 * bytes) {
            void* p = it->second;
            live += it->first;
            cached.erase(it);
            peak_live = std::max(peak_live, live);
            return p;
        }
        if (reserved + bytes > budget) trim();
        if (reserved + bytes > budget) return nullptr;
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (hipMalloc(&p, bytes) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
        }
        sizes[p] = bytes;
        reserved += bytes;
        live += bytes;
        peak_live = std::max(peak_live, live);
        return p;
    }
    void release(void* p) {
        const size_t b = sizes.at(p);
        live -= b;
        cached.emplace(b, p);
    }
    void trim() {
        for (auto& kv : cached) {
            (void)hipFree(kv.second);
            sizes.erase(kv.second);
            reserved -= kv.first;
        }
        cached.clear();
    }
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    vqc_batch* batch = nullptr;
    vqc_solo* solo = nullptr;  // a solo evaluation instead of a batch
    void* solo_block = nullptr;
    std::vector<Request*> reqs;
    std::vector<double> energies, grads;
    double* h_pool = nullptr;
    size_t h_pool_cap = 0;
    double* h_rows = nullptr;
    size_t h_rows_cap = 0;
};

int grow_pinned(double** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return VQC_OK;
    size_t want = std::max<size_t>(*cap * 2, 1 << 16);
    while (want < bytes) want *= 2;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    VQC_HIP(hipHostMalloc((void**)p, want, hipHostMallocDefault));
    *cap = want;
    return VQC_OK;
}

}  // namespace

struct vqc_server {
    vqc_server_config_t cfg{};
    int device = 0;
    int t_small = 12;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;  // scheduler: new requests or stop
    std::condition_variable cv_done;  // callers: a request finished
    std::deque<Request*> pending;
    std::unordered_map<int64_t, Request*> reqs;
    int64_t next_id = 1;
    uint64_t submits = 0;  // requests queued so far (the scheduler's view of new work)
    bool stop = false;
    int running = 0;  // requests on the device
    DevCache mem;
    std::vector<Lane> lanes;
    std::thread worker;
    // statistics
    int64_t completed = 0, batches = 0, batched_requests = 0;
    int64_t batch_hist[8] = {};  // batches of 1, 2, 3-4, 5-8, 9-16, 17-32, 33-64, 65+ requests
    void count_batch(int ni) {
        int b = 0;
        while (b < 7 && (1 << b) < ni) ++b;
        ++batch_hist[b];
        ++batches;
        batched_requests += ni;
    }

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    bool is_solo(const Request* r) const { return cfg.solo_qubits > 0 && r->inst.nqubits >= cfg.solo_qubits; }
    void run();
    bool launch(Lane& L);
    bool launch_solo(Lane& L, Request* r);
    void finish(Lane& L);
    void fail_requests(std::vector<Request*>& rs, int code, const std::string& msg);
};

void vqc_server::fail_requests(std::vector<Request*>& rs, int code, const std::string& msg) {
    std::lock_guard<std::mutex> g(mu);
    for (Request* r : rs) {
        r->status = code;
        r->msg = msg;
        r->t_done = now();
        --running;
    }
    rs.clear();
    cv_done.notify_all();
}

// Pick the batch for a free lane and launch it. Returns true when launched.
bool vqc_server::launch(Lane& L) {
    std::vector<Request*> sel;
    {
        std::lock_guard<std::mutex> g(mu);
        if (pending.empty()) return false;
        const uint64_t free_bytes = mem.budget > mem.live ? mem.budget - mem.live : 0;
        Request* head = pending.front();
        if (head->bytes > mem.budget) {
            pending.pop_front();
            head->status = VQC_ENOMEM;
            head->msg = "request needs more device memory than the serving budget";
            head->t_done = now();
            cv_done.notify_all();
            return true;
        }
        if (head->bytes > free_bytes) return false;  // admitted once running work frees memory
        if (is_solo(head)) {
            pending.pop_front();
            sel.push_back(head);
            running += 1;
        }
    }
    if (!sel.empty()) return launch_solo(L, sel[0]);
    {
        std::lock_guard<std::mutex> g(mu);
        if (pending.empty()) return false;
        const uint64_t free_bytes = mem.budget > mem.live ? mem.budget - mem.live : 0;
        Request* head = pending.front();
        if (head->bytes > free_bytes) return false;
        uint64_t used = 0;
        const double ratio = cfg.cost_ratio > 1 ? cfg.cost_ratio : 1e300;
        double cmin = head->cost, cmax = head->cost;
        for (auto it = pending.begin(); it != pending.end() && int(sel.size()) < cfg.max_batch;) {
            Request* r = *it;
            const double lo = std::min(cmin, r->cost), hi = std::max(cmax, r->cost);
            if (!is_solo(r) && used + r->bytes <= free_bytes && hi <= ratio * lo) {
                sel.push_back(r);
                used += r->bytes;
                cmin = lo;
                cmax = hi;
                it = pending.erase(it);
            } else {
                ++it;
            }
        }
        running += int(sel.size());
    }
    // pack: shared op / term arrays and pools
    const int ni = int(sel.size());
    std::vector<vqc_instance_t> inst(ni);
    std::vector<vqc_op_t> ops;
    std::vector<vqc_term_t> terms;
    int64_t plen = 0, hlen = 0, glen = 0;
    for (Request* r : sel) {
        plen += int64_t(r->pool.size() / 2);
        hlen += int64_t(r->hpool.size() / 2);
    }
    if (int(grow_pinned(&L.h_pool, &L.h_pool_cap, size_t(std::max<int64_t>(plen + hlen, 1)) * 16)) != VQC_OK) {
        fail_requests(sel, VQC_ENOMEM, "pinned staging allocation failed");
        return true;
    }
    double* pool = L.h_pool;
    double* hpool = L.h_pool + 2 * plen;
    int64_t po = 0, ho = 0;
    for (int i = 0; i < ni; ++i) {
        Request* r = sel[i];
        inst[i] = r->inst;
        inst[i].op_begin = int64_t(ops.size());
        inst[i].term_begin = int64_t(terms.size());
        inst[i].grad_begin = glen;
        glen += r->inst.nparams;
        for (vqc_op_t o : r->ops) {
            o.mat += po;
            if (o.aux >= 0) o.aux += po;
            for (int j = 0; j < 3; ++j)
                if (o.gen[j] >= 0) o.gen[j] += po;
            ops.push_back(o);
        }
        for (vqc_term_t t : r->terms) {
            t.mat += ho;
            terms.push_back(t);
        }
        memcpy(pool + 2 * po, r->pool.data(), r->pool.size() * sizeof(double));
        memcpy(hpool + 2 * ho, r->hpool.data(), r->hpool.size() * sizeof(double));
        po += int64_t(r->pool.size() / 2);
        ho += int64_t(r->hpool.size() / 2);
    }
    BatchOpts o;
    o.stream = L.stream;
    o.dev_alloc = DevCache::alloc_cb;
    o.dev_free = DevCache::free_cb;
    o.alloc_ctx = &mem;
    o.graph = false;
    o.pinned = false;
    int busy = 0;
    for (const Lane& x : lanes) busy += x.batch != nullptr || x.solo != nullptr;
    o.grid_frac = cfg.grid_frac;
    vqc_batch* b = nullptr;
    int rc;
    {
        // the allocator is only touched by this thread; the lock guards the statistics it feeds
        rc = vqc_batch_build(ni, inst.data(), ops.empty() ? nullptr : ops.data(), int64_t(ops.size()),
                             terms.empty() ? nullptr : terms.data(), int64_t(terms.size()), plen, hlen, o, &b);
    }
    if (rc == VQC_ENOMEM && getenv("VQC_SERVE_DEBUG"))
        fprintf(stderr, "[serve] batch build ENOMEM: %s (busy=%d live=%llu)\n", vqc_last_error(), busy,
                (unsigned long long)mem.live);
    if (rc == VQC_ENOMEM && busy > 0) {
        // the estimate admitted it but the allocator could not place it yet:
        // back to the head of the queue, in order, until a running lane finishes
        std::lock_guard<std::mutex> g(mu);
        for (auto it = sel.rbegin(); it != sel.rend(); ++it) pending.push_front(*it);
        running -= int(sel.size());
        return false;
    }
    if (rc != VQC_OK) {
        fail_requests(sel, rc, vqc_last_error());
        return true;
    }
    size_t dev_bytes = 0, pool_bytes = 0, rows_bytes = 0;
    vqc_batch_sizes(b, &dev_bytes, &pool_bytes, &rows_bytes);
    if (grow_pinned(&L.h_rows, &L.h_rows_cap, rows_bytes) != VQC_OK) {
        vqc_batch_release(b);
        fail_requests(sel, VQC_ENOMEM, "pinned staging allocation failed");
        return true;
    }
    // the pools are already in place in the lane's staging (h_pool)
    rc = vqc_batch_enqueue_run(b, pool, hpool, L.h_pool, L.h_rows);
    if (rc == VQC_OK && hipEventRecord(L.done, L.stream) != hipSuccess) rc = vqc_fail(VQC_EHIP, "hipEventRecord failed");
    if (rc != VQC_OK) {
        (void)hipStreamSynchronize(L.stream);
        vqc_batch_release(b);
        fail_requests(sel, rc, vqc_last_error());
        return true;
    }
    const double t = now();
    {
        std::lock_guard<std::mutex> g(mu);
        for (Request* r : sel) {
            r->status = 1;
            r->t_launch = t;
        }
        count_batch(ni);
    }
    L.batch = b;
    L.reqs = std::move(sel);
    L.energies.assign(2 * size_t(ni), 0.0);
    L.grads.assign(size_t(std::max<int64_t>(glen, 1)), 0.0);
    return true;
}

bool vqc_server::launch_solo(Lane& L, Request* r) {
    std::vector<Request*> sel{r};
    const vqc_instance_t& I = r->inst;
    vqc_solo* so = nullptr;
    int rc = vqc_solo_build(I.nqubits, I.nops, r->ops.empty() ? nullptr : r->ops.data(), r->pool.data(),
                            int64_t(r->pool.size() / 2), I.nterms, r->terms.empty() ? nullptr : r->terms.data(),
                            r->hpool.data(), int64_t(r->hpool.size() / 2), I.nparams, &so);
    if (rc != VQC_OK) {
        fail_requests(sel, rc, vqc_last_error());
        return true;
    }
    void* block = mem.alloc(vqc_solo_device_bytes(so));
    if (!block) {
        vqc_solo_free(so);
        int busy = 0;
        for (const Lane& x : lanes) busy += x.batch != nullptr || x.solo != nullptr;
        if (getenv("VQC_SERVE_DEBUG"))
            fprintf(stderr, "[serve] solo alloc failed: bytes=%zu budget=%llu live=%llu reserved=%llu busy=%d\n",
                    vqc_solo_device_bytes(so), (unsigned long long)mem.budget, (unsigned long long)mem.live,
                    (unsigned long long)mem.reserved, busy);
        if (busy) {  // wait for running work to return memory
            std::lock_guard<std::mutex> g(mu);
            pending.push_front(r);
            running -= 1;
            return false;
        }
        fail_requests(sel, VQC_ENOMEM, "device allocation within the serving budget failed");
        return true;
    }
    if (grow_pinned(&L.h_rows, &L.h_rows_cap, vqc_solo_rows_bytes(so)) != VQC_OK) {
        mem.release(block);
        vqc_solo_free(so);
        fail_requests(sel, VQC_ENOMEM, "pinned staging allocation failed");
        return true;
    }
    rc = vqc_solo_enqueue(so, block, L.stream, L.h_rows);
    if (rc == VQC_OK && hipEventRecord(L.done, L.stream) != hipSuccess) rc = vqc_fail(VQC_EHIP, "hipEventRecord failed");
    if (rc != VQC_OK) {
        (void)hipStreamSynchronize(L.stream);
        mem.release(block);
        vqc_solo_free(so);
        fail_requests(sel, rc, vqc_last_error());
        return true;
    }
    {
        std::lock_guard<std::mutex> g(mu);
        r->status = 1;
        r->t_launch = now();
        count_batch(1);
    }
    L.solo = so;
    L.solo_block = block;
    L.reqs = std::move(sel);
    return true;
}

void vqc_server::finish(Lane& L) {
    int rc;
    if (L.solo) {
        L.energies.assign(2, 0.0);
        L.grads.assign(size_t(std::max(L.reqs[0]->inst.nparams, 1)), 0.0);
        rc = vqc_solo_collect(L.solo, L.h_rows, L.energies.data(), L.grads.data());
        mem.release(L.solo_block);
        vqc_solo_free(L.solo);
        L.solo = nullptr;
        L.solo_block = nullptr;
    } else {
        rc = vqc_batch_collect(L.batch, L.h_rows, L.energies.data(), L.grads.data());
        vqc_batch_release(L.batch);
        L.batch = nullptr;
    }
    const std::string msg = rc == VQC_OK ? std::string() : std::string(vqc_last_error());
    const double t = now();
    std::lock_guard<std::mutex> g(mu);
    int64_t goff = 0;
    for (size_t i = 0; i < L.reqs.size(); ++i) {
        Request* r = L.reqs[i];
        const int np = r->inst.nparams;
        if (rc == VQC_OK) {
            r->energy[0] = L.energies[2 * i];
            r->energy[1] = L.energies[2 * i + 1];
            r->grad.assign(L.grads.begin() + goff, L.grads.begin() + goff + np);
            r->status = 2;
        } else {
            r->status = rc;
            r->msg = msg;
        }
        goff += np;
        r->t_done = t;
        ++completed;
        --running;
    }
    L.reqs.clear();
    cv_done.notify_all();
}

void vqc_server::run() {
    (void)hipSetDevice(device);
    for (;;) {
        uint64_t seen;
        {
            std::lock_guard<std::mutex> g(mu);
            seen = submits;
        }
        bool progress = false;
        for (Lane& L : lanes) {
            if (!L.batch && !L.solo) continue;
            const hipError_t q = hipEventQuery(L.done);
            if (q == hipSuccess) {
                finish(L);
                progress = true;
            } else if (q != hipErrorNotReady) {
                (void)hipGetLastError();
                std::vector<Request*> rs = std::move(L.reqs);
                if (L.batch) vqc_batch_release(L.batch);
                if (L.solo) {
                    mem.release(L.solo_block);
                    vqc_solo_free(L.solo);
                }
                L.batch = nullptr;
                L.solo = nullptr;
                L.solo_block = nullptr;
                fail_requests(rs, VQC_EHIP, std::string("device error: ") + hipGetErrorString(q));
                progress = true;
            }
        }
        for (Lane& L : lanes)
            if (!L.batch && !L.solo && launch(L)) progress = true;
        if (progress) continue;
        bool any_running = false;
        for (const Lane& L : lanes) any_running |= L.batch != nullptr || L.solo != nullptr;
        std::unique_lock<std::mutex> g(mu);
        if (stop && pending.empty() && !any_running) break;
        if (submits != seen) continue;  // queued after the launch attempts above: try those first
        if (any_running) cv_work.wait_for(g, std::chrono::microseconds(cfg.poll_us > 0 ? cfg.poll_us : 50));
        else if (pending.empty() && !stop) cv_work.wait(g);
        else if (!pending.empty()) {
            // pending work that does not fit while nothing runs: the head can never be admitted
            Request* head = pending.front();
            pending.pop_front();
            head->status = VQC_ENOMEM;
            char buf[192];
            snprintf(buf, sizeof buf, "request of %llu device bytes does not fit the serving budget (%llu bytes, %llu held)",
                     (unsigned long long)head->bytes, (unsigned long long)mem.budget, (unsigned long long)mem.live);
            head->msg = buf;
            head->t_done = now();
            cv_done.notify_all();
        }
    }
}

extern "C" int vqc_server_default_config(vqc_server_config_t* cfg) {
    if (!cfg) return vqc_fail(VQC_EINVAL, "null config");
    cfg->hbm_budget = 288ull * 1000 * 1000 * 1000;
    cfg->lanes = 3;
    cfg->max_batch = 64;
    cfg->grid_frac = 1.0;
    cfg->cost_ratio = 64.0;
    cfg->poll_us = 50;
    cfg->solo_qubits = 19;
    return VQC_OK;
}

extern "C" int vqc_server_create(const vqc_server_config_t* cfg, vqc_server** out) {
    if (!out) return vqc_fail(VQC_EINVAL, "null output");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    vqc_server_config_t k;
    vqc_server_default_config(&k);
    if (cfg) k = *cfg;
    if (k.lanes < 1 || k.lanes > 8 || k.max_batch < 1 || k.hbm_budget == 0 || !(k.grid_frac > 0) || k.grid_frac > 1 ||
        k.solo_qubits < 0)
        return vqc_fail(VQC_EINVAL, "vqc_server_create: bad configuration");
    vqc_server* s = new vqc_server;
    s->cfg = k;
    s->device = c->device;
    s->t_small = vqc_tile_tuning().t_small;
    s->t0 = Clock::now();
    // the budget is the configured one, capped by what the device has free now
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > (size_t(1) << 30))
        s->mem.budget = std::min<uint64_t>(k.hbm_budget, uint64_t(free_b) - (uint64_t(1) << 30));
    else
        s->mem.budget = k.hbm_budget;
    s->lanes.resize(size_t(k.lanes));
    for (Lane& L : s->lanes) {
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            for (Lane& x : s->lanes) {
                if (x.done) (void)hipEventDestroy(x.done);
                if (x.stream) (void)hipStreamDestroy(x.stream);
            }
            delete s;
            return vqc_fail(VQC_EHIP, "vqc_server_create: stream or event creation failed");
        }
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return VQC_OK;
}

extern "C" int vqc_server_submit(vqc_server* s, const vqc_instance_t* inst, const vqc_op_t* ops, const vqc_term_t* terms,
                                 const double* pool, int64_t pool_len, const double* hpool, int64_t hpool_len,
                                 int64_t* ticket) {
    if (!s || !inst || !ticket || pool_len < 0 || hpool_len < 0) return vqc_fail(VQC_EINVAL, "vqc_server_submit: bad arguments");
    const vqc_instance_t& I = *inst;
    if (I.nqubits < 1 || I.nqubits > 30 || I.nops < 0 || I.nterms < 0 || I.nparams < 0 || (I.nops && !ops) ||
        (I.nterms && !terms) || (pool_len && !pool) || (hpool_len && !hpool))
        return vqc_fail(VQC_EINVAL, "vqc_server_submit: bad instance");
    Request* r = new Request;
    r->inst = I;
    r->inst.op_begin = 0;
    r->inst.term_begin = 0;
    r->inst.grad_begin = 0;
    r->ops.assign(ops + I.op_begin, ops + I.op_begin + I.nops);
    r->terms.assign(terms + I.term_begin, terms + I.term_begin + I.nterms);
    r->pool.assign(pool, pool + 2 * pool_len);
    r->hpool.assign(hpool, hpool + 2 * hpool_len);
    // device bytes: A and B states (tiled instances), program and variants, partials
    const uint64_t L = uint64_t(1) << I.nqubits;
    const uint64_t states = I.nqubits > s->t_small ? 32 * L : 0;
    const uint64_t prog = uint64_t(I.nops * 3 + I.nterms + 8) * 512 + uint64_t(pool_len + hpool_len) * 16 * 4;
    const uint64_t parts = uint64_t(I.nparams + I.nterms + 1) * 1024 * 16;  // partial columns per row
    r->bytes = states + prog + parts + (uint64_t(4) << 20);
    r->cost = double(L) * double(3 * I.nops + 2 * I.nterms + 1);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) {
            delete r;
            return vqc_fail(VQC_EINVAL, "vqc_server_submit: server is shutting down");
        }
        r->id = s->next_id++;
        ++s->submits;
        r->t_submit = s->now();
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        *ticket = r->id;
    }
    s->cv_work.notify_one();
    return VQC_OK;
}

static int take_result(vqc_server* s, Request* r, double* energy2, double* grad, double* times3) {
    // called with s->mu held and r finished
    const int st = r->status;
    if (st == 2) {
        if (energy2) {
            energy2[0] = r->energy[0];
            energy2[1] = r->energy[1];
        }
        if (grad && !r->grad.empty()) memcpy(grad, r->grad.data(), r->grad.size() * sizeof(double));
    }
    if (times3) {
        times3[0] = r->t_submit;
        times3[1] = r->t_launch;
        times3[2] = r->t_done;
    }
    const std::string msg = r->msg;
    s->reqs.erase(r->id);
    delete r;
    if (st != 2) return vqc_fail(st, "vqc_server: %s", msg.c_str());
    return VQC_OK;
}

extern "C" int vqc_server_poll(vqc_server* s, int64_t ticket, double* energy2, double* grad, double* times3) {
    if (!s) return vqc_fail(VQC_EINVAL, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return vqc_fail(VQC_EINVAL, "vqc_server_poll: unknown ticket");
    Request* r = it->second;
    if (r->status == 0 || r->status == 1) return 1;  // not finished
    return take_result(s, r, energy2, grad, times3);
}

extern "C" int vqc_server_wait(vqc_server* s, int64_t ticket, double* energy2, double* grad, double* times3) {
    if (!s) return vqc_fail(VQC_EINVAL, "null server");
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return vqc_fail(VQC_EINVAL, "vqc_server_wait: unknown ticket");
    Request* r = it->second;
    s->cv_done.wait(g, [r] { return r->status != 0 && r->status != 1; });
    return take_result(s, r, energy2, grad, times3);
}

extern "C" int vqc_server_stats(vqc_server* s, vqc_server_stats_t* out) {
    if (!s || !out) return vqc_fail(VQC_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    out->completed = s->completed;
    out->batches = s->batches;
    out->batched_requests = s->batched_requests;
    out->pending = int64_t(s->pending.size());
    out->running = s->running;
    out->budget_bytes = s->mem.budget;
    out->reserved_bytes = s->mem.reserved;
    out->peak_live_bytes = s->mem.peak_live;
    for (int i = 0; i < 8; ++i) out->batch_hist[i] = s->batch_hist[i];
    return VQC_OK;
}

extern "C" int vqc_server_destroy(vqc_server* s) {
    if (!s) return VQC_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();  // drains: every submitted request finishes first
    (void)hipSetDevice(s->device);
    for (Lane& L : s->lanes) {
        (void)hipStreamSynchronize(L.stream);
        if (L.h_pool) (void)hipHostFree(L.h_pool);
        if (L.h_rows) (void)hipHostFree(L.h_rows);
        (void)hipEventDestroy(L.done);
        (void)hipStreamDestroy(L.stream);
    }
    s->mem.trim();
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
    return VQC_OK;
}
