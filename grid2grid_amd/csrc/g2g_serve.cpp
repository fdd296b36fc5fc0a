// Serving loop: continuous batching of transform requests under an HBM
// budget (g2g_server_* in include/grid2grid_amd.h).
//
// One host thread owns the device side. Each iteration it (1) retires
// launches whose completion event fired, releasing their requests' footprint,
// and (2) packs the waiting requests that fit the budget (FIFO, so a large
// request is never starved by small ones) into one persistent gather launch on
// a free stream slot. Requests that arrive while launches run join the next
// launch, and launches on different slots overlap on the device.
#include "grid2grid_amd.h"
#include "g2g_internal.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

using namespace g2g;
using clock_type = std::chrono::steady_clock;

namespace {

struct pending_req {
    int64_t ticket;
    g2g_plan plan;
    unsigned char alpha[16], beta[16];
    bool scaled;
    int64_t footprint;
    int64_t bytes;
    clock_type::time_point submitted;
};

struct slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    work_desc *h_descs = nullptr;  // pinned staging of the descriptor table
    work_desc *d_descs = nullptr;
    work_rec *d_work = nullptr;  // per global work index of a launch
    uint32_t work_cap = 0;
    uint32_t *d_queue = nullptr;
    std::vector<pending_req> reqs;
    bool busy = false;
};

struct result {
    double latency = 0;
    int status = G2G_OK;
    std::string error;  // this request's own failure
};

}  // namespace

struct g2g_server_s {
    int64_t budget = 0;
    int max_batch = 1;
    int mode = G2G_SERVE_BATCHED;
    std::vector<slot> slots;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<pending_req> pending;
    std::unordered_map<int64_t, result> finished;  // until poll or wait reports them
    std::unordered_set<int64_t> unreported;         // submitted, not yet reported by poll or wait
    int64_t next_ticket = 1;
    int64_t in_flight = 0;
    int busy_slots = 0;
    bool stop = false;
    std::string error;
    g2g_server_stats st{};
    std::thread th;

    void loop();
    int launch(slot &sl, std::vector<pending_req> &batch);
};

namespace {

int slot_init(slot &s, int max_batch) {
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&s.h_descs, sizeof(work_desc) * max_batch, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&s.d_descs, sizeof(work_desc) * max_batch) != hipSuccess ||
        hipMalloc(&s.d_queue, sizeof(uint32_t) * kQueueWords) != hipSuccess)
        return internal_fail(G2G_ERR_HIP, "server: stream/event/buffer creation failed");
    return G2G_OK;
}

void slot_free(slot &s) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.d_work) (void)hipFree(s.d_work);
    if (s.d_queue) (void)hipFree(s.d_queue);
    if (s.d_descs) (void)hipFree(s.d_descs);
    if (s.h_descs) (void)hipHostFree(s.h_descs);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    s = slot();
}

}  // namespace

int g2g_server_s::launch(slot &sl, std::vector<pending_req> &batch) {
    const int n = (int)batch.size();
    if (mode == G2G_SERVE_PLAIN) {
        for (const pending_req &r : batch) {
            const int rc = g2g_transform_staged(r.plan, r.scaled ? r.alpha : nullptr, r.scaled ? r.beta : nullptr,
                                                sl.stream);
            if (rc != G2G_OK) return rc;
        }
    } else {
        std::vector<g2g_plan_s *> plans(n);
        std::vector<const void *> al(n), be(n);
        for (int k = 0; k < n; ++k) {
            plans[k] = batch[k].plan;
            al[k] = batch[k].scaled ? batch[k].alpha : nullptr;
            be[k] = batch[k].scaled ? batch[k].beta : nullptr;
        }
        std::vector<work_desc> descs;
        uint32_t total_tiles = 0;
        uint64_t total_bytes = 0;
        int rc = internal_descs(n, plans.data(), al.data(), be.data(), descs, total_tiles, total_bytes);
        if (rc != G2G_OK) return rc;
        if (total_tiles > sl.work_cap) {
            if (sl.d_work) (void)hipFree(sl.d_work);
            sl.d_work = nullptr;
            const uint32_t cap = std::max<uint32_t>(total_tiles, 1u << 16);
            if (hipMalloc(&sl.d_work, sizeof(work_rec) * cap) != hipSuccess)
                return internal_fail(G2G_ERR_NOMEM, "server: work table allocation failed");
            sl.work_cap = cap;
        }
        std::memcpy(sl.h_descs, descs.data(), sizeof(work_desc) * n);
        if (hipMemcpyAsync(sl.d_descs, sl.h_descs, sizeof(work_desc) * n, hipMemcpyHostToDevice, sl.stream) !=
            hipSuccess)
            return internal_fail(G2G_ERR_HIP, "server: descriptor upload failed");
        if (launch_work_records(sl.d_descs, n, sl.d_work, sl.stream) != hipSuccess)
            return internal_fail(G2G_ERR_HIP, "server: work table launch failed");
        gather_args a;
        std::memset(&a, 0, sizeof a);
        a.descs = sl.d_descs;
        a.work = sl.d_work;
        a.total_tiles = total_tiles;
        a.queue = sl.d_queue;
        launch_cfg c;
        c.persistent = true;
        c.features = FEAT_COPY;
        for (const work_desc &d : descs) c.features = std::max(c.features, request_features(d.kr.r[0]));
        if (launch_gather(a, c, internal_grid_limit(), sl.stream) != hipSuccess)
            return internal_fail(G2G_ERR_HIP, "server: gather launch failed");
    }
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess)
        return internal_fail(G2G_ERR_HIP, "server: event record failed");
    return G2G_OK;
}

void g2g_server_s::loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progressed = false;
        // retire finished launches
        for (slot &sl : slots) {
            if (!sl.busy) continue;
            lk.unlock();
            const hipError_t q = hipEventQuery(sl.done);
            lk.lock();
            if (q == hipErrorNotReady) continue;
            const auto now = clock_type::now();
            for (const pending_req &r : sl.reqs) {
                result res;
                res.latency = std::chrono::duration<double>(now - r.submitted).count();
                res.status = q == hipSuccess ? G2G_OK : G2G_ERR_HIP;
                if (q != hipSuccess) res.error = std::string("server: launch failed: ") + hipGetErrorString(q);
                finished[r.ticket] = res;
                in_flight -= r.footprint;
                st.completed++;
                st.bytes += r.bytes;
            }
            if (q != hipSuccess && error.empty()) error = std::string("server: ") + hipGetErrorString(q);
            sl.reqs.clear();
            sl.busy = false;
            busy_slots--;
            progressed = true;
            cv_done.notify_all();
        }
        // admit waiting requests into a free slot
        slot *free_slot = nullptr;
        for (slot &sl : slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (free_slot && !pending.empty()) {
            std::vector<pending_req> batch;
            while (!pending.empty() && (int)batch.size() < max_batch &&
                   in_flight + pending.front().footprint <= budget) {
                in_flight += pending.front().footprint;
                batch.push_back(pending.front());
                pending.pop_front();
            }
            if (!batch.empty()) {
                st.peak_in_flight_bytes = std::max(st.peak_in_flight_bytes, in_flight);
                free_slot->busy = true;
                busy_slots++;
                st.launches++;
                lk.unlock();
                const int rc = launch(*free_slot, batch);
                std::string msg = rc == G2G_OK ? "" : g2g_last_error();
                lk.lock();
                free_slot->reqs = std::move(batch);
                if (rc != G2G_OK) {
                    // the event was not recorded: fail these requests now
                    const auto now = clock_type::now();
                    for (const pending_req &r : free_slot->reqs) {
                        result res;
                        res.latency = std::chrono::duration<double>(now - r.submitted).count();
                        res.status = rc;
                        res.error = msg;
                        finished[r.ticket] = res;
                        in_flight -= r.footprint;
                    }
                    if (error.empty()) error = msg;
                    free_slot->reqs.clear();
                    free_slot->busy = false;
                    busy_slots--;
                    cv_done.notify_all();
                }
                progressed = true;
            }
        }
        if (stop && pending.empty() && busy_slots == 0) break;
        if (!progressed) {
            if (busy_slots > 0)
                cv_work.wait_for(lk, std::chrono::microseconds(20));
            else
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

int g2g_server_create(int64_t budget_bytes, int n_streams, int max_batch, int mode, g2g_server *out) {
    if (!out || budget_bytes <= 0 || n_streams < 1 || n_streams > 16 || max_batch < 1 ||
        (mode != G2G_SERVE_BATCHED && mode != G2G_SERVE_PLAIN))
        return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    *out = nullptr;
    g2g_server s = new (std::nothrow) g2g_server_s();
    if (!s) return internal_fail(G2G_ERR_NOMEM, "out of host memory");
    s->budget = budget_bytes;
    s->mode = mode;
    s->max_batch = mode == G2G_SERVE_PLAIN ? 1 : max_batch;
    const int ns = mode == G2G_SERVE_PLAIN ? 1 : n_streams;
    s->slots.resize(ns);
    for (slot &sl : s->slots) {
        const int rc = slot_init(sl, s->max_batch);
        if (rc != G2G_OK) {
            for (slot &x : s->slots) slot_free(x);
            delete s;
            return rc;
        }
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return G2G_OK;
}

int g2g_server_submit(g2g_server s, g2g_plan plan, const void *alpha, const void *beta, int64_t *ticket) {
    if (!s || !plan || !ticket || ((alpha == nullptr) != (beta == nullptr)))
        return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    g2g_plan_info info;
    int rc = g2g_plan_get_info(plan, &info);
    if (rc != G2G_OK) return rc;
    const size_t es = (size_t)info.bytes / (size_t)std::max<int64_t>(1, info.m * info.n);
    pending_req r;
    std::memset(r.alpha, 0, sizeof r.alpha);
    std::memset(r.beta, 0, sizeof r.beta);
    r.plan = plan;
    r.scaled = alpha != nullptr;
    if (r.scaled) {
        std::memcpy(r.alpha, alpha, es);
        std::memcpy(r.beta, beta, es);
    }
    r.bytes = info.bytes;
    r.footprint = 2 * info.bytes;
    std::lock_guard<std::mutex> g(s->mu);
    s->st.submitted++;
    if (r.footprint > s->budget) {
        s->st.rejected++;
        return internal_fail(G2G_ERR_BUDGET, "server: request footprint %lld exceeds the budget %lld",
                             (long long)r.footprint, (long long)s->budget);
    }
    r.ticket = s->next_ticket++;
    r.submitted = clock_type::now();
    *ticket = r.ticket;
    s->unreported.insert(r.ticket);
    s->pending.push_back(r);
    s->cv_work.notify_one();
    return G2G_OK;
}

// Reports a finished ticket once: its entry is erased, so a long-running
// server keeps only results nobody has collected yet.
int g2g_server_poll(g2g_server s, int64_t ticket, int32_t *done, double *latency_s) {
    if (!s || !done) return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) return internal_fail(G2G_ERR_INVALID, "server: unknown ticket");
    if (!s->unreported.count(ticket))
        return internal_fail(G2G_ERR_STATE, "server: ticket %lld was already reported", (long long)ticket);
    auto it = s->finished.find(ticket);
    *done = it != s->finished.end();
    if (*done) {
        const result r = it->second;
        s->finished.erase(it);
        s->unreported.erase(ticket);
        if (latency_s) *latency_s = r.latency;
        if (r.status != G2G_OK) return internal_fail(r.status, "%s", r.error.c_str());
    }
    return G2G_OK;
}

int g2g_server_wait(g2g_server s, int64_t ticket, double *latency_s) {
    if (!s) return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket <= 0 || ticket >= s->next_ticket) return internal_fail(G2G_ERR_INVALID, "server: unknown ticket");
    if (!s->unreported.count(ticket))
        return internal_fail(G2G_ERR_STATE, "server: ticket %lld was already reported", (long long)ticket);
    s->cv_done.wait(lk, [&] { return s->finished.count(ticket) != 0; });
    auto it = s->finished.find(ticket);
    const result r = it->second;
    s->finished.erase(it);
    s->unreported.erase(ticket);
    if (latency_s) *latency_s = r.latency;
    if (r.status != G2G_OK) return internal_fail(r.status, "%s", r.error.c_str());
    return G2G_OK;
}

int g2g_server_drain(g2g_server s) {
    if (!s) return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->busy_slots == 0; });
    if (!s->error.empty()) return internal_fail(G2G_ERR_HIP, "%s", s->error.c_str());
    return G2G_OK;
}

int g2g_server_get_stats(g2g_server s, g2g_server_stats *stats) {
    if (!s || !stats) return internal_fail(G2G_ERR_INVALID, "server: bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    *stats = s->st;
    stats->in_flight_bytes = s->in_flight;
    return G2G_OK;
}

int g2g_server_destroy(g2g_server s) {
    if (!s) return G2G_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    for (slot &sl : s->slots) slot_free(sl);
    delete s;
    return G2G_OK;
}

}  // extern "C"
