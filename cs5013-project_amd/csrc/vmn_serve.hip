// Continuous-batching request server over one index (host C++, HIP runtime).
//
// Requests (search / insert / delete) are submitted from any thread with
// their arrival time and queued in arrival order. An engine thread:
//  * packs consecutive search requests into the batch of a free slot (a
//    search context with pinned staging for `max_batch` queries), copies the
//    queries in, launches the batched search and the result copy-back on the
//    slot's stream and records an event; up to `n_slots` batches are on the
//    device at once, so requests that arrive while others run are admitted as
//    soon as a slot frees up;
//  * applies an insert or delete when it reaches the head of the queue, after
//    the searches already launched have completed; searches behind it wait.
//    Every search therefore sees exactly the updates that arrived before it,
//    and since each query is traversed on its own wavefront its result does
//    not depend on which other queries share its batch;
//  * consolidates tombstones once `consolidate_every` deletes have
//    accumulated (0: never).
// The reference's request loop is host code that sorts each fixed batch into
// deletes, inserts and searches (parth/src/engine.cu:407-568,
// vamana.cu:1782-1895); this engine admits requests continuously instead.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/vamana_mi355x.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0;
    int32_t kind = 0;
    int64_t n = 0;
    std::vector<float> f;    // search queries / insert vectors
    std::vector<int32_t> i;  // delete ids
    double t_arrival = 0, t_done = 0;
    std::vector<int32_t> ids;  // search results [n x k]; insert: assigned ids
    std::vector<float> dists;
    int status = 0;
    std::string err;
    bool done = false;
};

struct Slot {
    vmn_ctx* ctx = nullptr;
    hipEvent_t ev = nullptr;
    float* h_q = nullptr;
    int32_t* h_ids = nullptr;
    float* h_d = nullptr;
    float* d_q = nullptr;
    int32_t* d_ids = nullptr;
    float* d_d = nullptr;
    std::vector<Request*> reqs;
    int64_t nq = 0;
    bool busy = false;
};

}  // namespace

struct vmn_server {
    vmn_index* idx = nullptr;
    vmn_server_params p{};
    int dim = 0;
    Clock::time_point t0;
    std::vector<Slot> slots;
    std::mutex m;
    std::condition_variable cv;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> all;
    int64_t next_ticket = 1;
    int64_t queued_bytes = 0;
    int64_t deletes_since = 0;
    std::thread th;
    std::atomic<bool> stop{false};
    std::atomic<int64_t> in_flight{0};
    std::string fatal;
    vmn_server_stats counters{};    // engine counters (engine thread writes, under m)
    double drain_from = -1.0;       // when the queue head became an update waiting on slots

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void finish(Request* r, int status, const char* msg) {
        std::lock_guard<std::mutex> g(m);
        r->status = status;
        if (msg) r->err = msg;
        r->t_done = now();
        r->done = true;
        queued_bytes -= (int64_t)(r->f.size() * sizeof(float) + r->i.size() * sizeof(int32_t));
        in_flight--;
    }

    bool any_busy() const {
        for (const Slot& s : slots)
            if (s.busy) return true;
        return false;
    }

    void complete_slots() {
        for (Slot& s : slots) {
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                // a runtime error: the batch's requests fail with it and the
                // slot is freed, so later requests and shutdown still proceed
                const std::string msg = std::string("search batch: ") + hipGetErrorString(q);
                for (Request* r : s.reqs) finish(r, VMN_ERR_HIP, msg.c_str());
                s.reqs.clear();
                s.busy = false;
                std::lock_guard<std::mutex> g(m);
                counters.failed_batches++;
                continue;
            }
            int64_t off = 0;
            for (Request* r : s.reqs) {
                const int64_t cnt = r->n * p.k;
                r->ids.assign(s.h_ids + off * p.k, s.h_ids + off * p.k + cnt);
                r->dists.assign(s.h_d + off * p.k, s.h_d + off * p.k + cnt);
                off += r->n;
                finish(r, VMN_OK, nullptr);
            }
            s.reqs.clear();
            s.busy = false;
        }
    }

    void launch(Slot& s, std::vector<Request*>& batch) {
        int64_t nq = 0;
        for (Request* r : batch) {
            std::memcpy(s.h_q + nq * dim, r->f.data(), r->n * dim * sizeof(float));
            nq += r->n;
        }
        hipStream_t st = (hipStream_t)vmn_ctx_stream(s.ctx);
        hipError_t e = hipMemcpyAsync(s.d_q, s.h_q, nq * dim * sizeof(float), hipMemcpyHostToDevice, st);
        int rc = e == hipSuccess ? vmn_ctx_search(s.ctx, s.d_q, nq, p.k, p.L, p.prec, p.k_rerank, s.d_ids, s.d_d)
                                 : VMN_ERR_HIP;
        if (rc == VMN_OK) {
            e = hipMemcpyAsync(s.h_ids, s.d_ids, nq * p.k * sizeof(int32_t), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s.h_d, s.d_d, nq * p.k * sizeof(float), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(s.ev, st);
            if (e != hipSuccess) rc = VMN_ERR_HIP;
        }
        if (rc != VMN_OK) {
            std::string msg = e != hipSuccess ? hipGetErrorString(e) : vmn_last_error();
            for (Request* r : batch) finish(r, rc, msg.c_str());
            return;
        }
        s.reqs = batch;
        s.nq = nq;
        s.busy = true;
        std::lock_guard<std::mutex> g(m);
        counters.batches++;
        counters.queries += nq;
        if (nq > counters.max_batch_queries) counters.max_batch_queries = nq;
    }

    void apply_update(Request* r) {
        int rc;
        const double t = now();
        bool consolidated = false;
        if (r->kind == VMN_REQ_INSERT) {
            r->ids.resize(r->n);
            rc = vmn_insert(idx, r->f.data(), r->n, r->ids.data(), VMN_MEM_HOST);
        } else {
            rc = vmn_delete_ids(idx, r->i.data(), r->n, VMN_MEM_HOST);
            if (rc == VMN_OK) {
                deletes_since += r->n;
                if (p.consolidate_every > 0 && deletes_since >= p.consolidate_every) {
                    rc = vmn_consolidate(idx);
                    deletes_since = 0;
                    consolidated = true;
                }
            }
        }
        {
            const double dt = now() - t;
            std::lock_guard<std::mutex> g(m);
            if (r->kind == VMN_REQ_INSERT) {
                counters.inserts++;
                counters.insert_s += dt;
            } else {
                counters.deletes++;
                counters.delete_s += dt;
            }
            counters.consolidations += consolidated;
            if (dt > counters.max_update_s) counters.max_update_s = dt;
        }
        finish(r, rc, rc == VMN_OK ? nullptr : vmn_last_error());
    }

    void loop() {
        (void)hipSetDevice(p.device);
        while (true) {
            complete_slots();
            std::unique_lock<std::mutex> lk(m);
            if (queue.empty()) {
                if (stop && !any_busy()) return;
                cv.wait_for(lk, std::chrono::microseconds(any_busy() ? 20 : 500));
                continue;
            }
            Request* head = queue.front();
            if (head->kind != VMN_REQ_SEARCH) {
                if (any_busy()) {  // ordered write: let launched searches finish
                    if (drain_from < 0) drain_from = now();
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(10));
                    continue;
                }
                if (drain_from >= 0) {
                    counters.drain_s += now() - drain_from;
                    drain_from = -1.0;
                }
                queue.pop_front();
                lk.unlock();
                apply_update(head);
                continue;
            }
            Slot* free_slot = nullptr;
            for (Slot& s : slots)
                if (!s.busy) {
                    free_slot = &s;
                    break;
                }
            if (!free_slot) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(10));
                continue;
            }
            std::vector<Request*> batch;
            int64_t nq = 0;
            while (!queue.empty() && queue.front()->kind == VMN_REQ_SEARCH &&
                   nq + queue.front()->n <= p.max_batch) {
                nq += queue.front()->n;
                batch.push_back(queue.front());
                queue.pop_front();
            }
            lk.unlock();
            launch(*free_slot, batch);
        }
    }
};

extern "C" {

void vmn_default_server_params(vmn_server_params* p) {
    std::memset(p, 0, sizeof *p);
    p->k = 10;
    p->L = 64;
    p->prec = VMN_PREC_FP32;
    p->k_rerank = 0;
    p->max_batch = 4096;
    p->n_slots = 3;
    p->max_queued_bytes = (int64_t)1 << 30;
    p->consolidate_every = 0;
    p->device = 0;
}

int vmn_server_create(vmn_index* idx, const vmn_server_params* p, int32_t dim, vmn_server** out) {
    if (!idx || !p || !out) return -1;
    *out = nullptr;
    if (p->n_slots < 1 || p->n_slots > 16 || p->max_batch < 1 || p->k < 1 || p->k > p->L ||
        dim < 1 || dim % 32 != 0)
        return -1;
    auto* s = new vmn_server();
    s->idx = idx;
    s->p = *p;
    s->dim = dim;
    s->t0 = Clock::now();
    if (hipSetDevice(p->device) != hipSuccess) {
        delete s;
        return -2;
    }
    s->slots.resize(p->n_slots);
    for (Slot& sl : s->slots) {
        const size_t qb = (size_t)p->max_batch * dim * sizeof(float);
        const size_t rb = (size_t)p->max_batch * p->k;
        bool ok = vmn_ctx_create(idx, &sl.ctx) == VMN_OK &&
                  hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) == hipSuccess &&
                  hipHostMalloc((void**)&sl.h_q, qb, hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void**)&sl.h_ids, rb * sizeof(int32_t), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc((void**)&sl.h_d, rb * sizeof(float), hipHostMallocDefault) == hipSuccess &&
                  hipMalloc((void**)&sl.d_q, qb) == hipSuccess &&
                  hipMalloc((void**)&sl.d_ids, rb * sizeof(int32_t)) == hipSuccess &&
                  hipMalloc((void**)&sl.d_d, rb * sizeof(float)) == hipSuccess;
        if (!ok) {
            vmn_server_destroy(s);
            return -2;
        }
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return VMN_OK;
}

double vmn_server_now(const vmn_server* s) { return s ? s->now() : -1.0; }

int vmn_server_submit(vmn_server* s, int32_t kind, const void* data, int64_t n, double t_arrival,
                      int64_t* ticket) {
    if (!s || !ticket || (n > 0 && !data) || n < 1) return -1;
    if (kind != VMN_REQ_SEARCH && kind != VMN_REQ_INSERT && kind != VMN_REQ_DELETE) return -1;
    if (kind == VMN_REQ_SEARCH && n > s->p.max_batch) return -1;
    auto* r = new Request();
    r->kind = kind;
    r->n = n;
    r->t_arrival = t_arrival;
    if (kind == VMN_REQ_DELETE) r->i.assign((const int32_t*)data, (const int32_t*)data + n);
    else r->f.assign((const float*)data, (const float*)data + n * s->dim);
    const int64_t bytes = (int64_t)(r->f.size() * sizeof(float) + r->i.size() * sizeof(int32_t));
    {
        std::lock_guard<std::mutex> g(s->m);
        if (s->queued_bytes + bytes > s->p.max_queued_bytes) {  // admission by memory
            delete r;
            return -4;
        }
        r->ticket = s->next_ticket++;
        s->queued_bytes += bytes;
        s->all[r->ticket] = r;
        s->queue.push_back(r);
        s->in_flight++;
        *ticket = r->ticket;
    }
    s->cv.notify_one();
    return VMN_OK;
}

int vmn_server_poll(vmn_server* s, int64_t ticket, int32_t* ids_out, float* dists_out,
                    double* t_arrival, double* t_done) {
    if (!s) return -1;
    Request* r;
    {
        std::lock_guard<std::mutex> g(s->m);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) return -1;
        r = it->second;
        if (!r->done) return 0;
        s->all.erase(it);
    }
    int rc = r->status;
    if (rc == VMN_OK) {
        if (ids_out && !r->ids.empty()) std::memcpy(ids_out, r->ids.data(), r->ids.size() * sizeof(int32_t));
        if (dists_out && !r->dists.empty())
            std::memcpy(dists_out, r->dists.data(), r->dists.size() * sizeof(float));
    }
    if (t_arrival) *t_arrival = r->t_arrival;
    if (t_done) *t_done = r->t_done;
    delete r;
    return rc == VMN_OK ? 1 : rc;
}

int64_t vmn_server_pending(const vmn_server* s) { return s ? s->in_flight.load() : -1; }

int vmn_server_get_stats(const vmn_server* s, vmn_server_stats* out) {
    if (!s || !out) return -1;
    auto* ms = const_cast<vmn_server*>(s);
    std::lock_guard<std::mutex> g(ms->m);
    *out = s->counters;
    return VMN_OK;
}

int vmn_server_reset_stats(vmn_server* s) {
    if (!s) return -1;
    std::lock_guard<std::mutex> g(s->m);
    s->counters = vmn_server_stats{};
    return VMN_OK;
}

int vmn_server_destroy(vmn_server* s) {
    if (!s) return VMN_OK;
    s->stop = true;
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    for (Slot& sl : s->slots) {
        if (sl.ctx) vmn_ctx_destroy(sl.ctx);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.h_q) (void)hipHostFree(sl.h_q);
        if (sl.h_ids) (void)hipHostFree(sl.h_ids);
        if (sl.h_d) (void)hipHostFree(sl.h_d);
        if (sl.d_q) (void)hipFree(sl.d_q);
        if (sl.d_ids) (void)hipFree(sl.d_ids);
        if (sl.d_d) (void)hipFree(sl.d_d);
    }
    for (auto& kv : s->all) delete kv.second;
    delete s;
    return VMN_OK;
}

}  // extern "C"
