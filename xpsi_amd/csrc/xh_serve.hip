// Continuous-batching likelihood server (host runtime).
//
// A server owns `n_slots` pipelines, each with its own HIP stream and device
// buffers for `slot_batch` requests.  Requests are submitted from any thread
// into a FIFO; one dispatcher thread
//   - harvests slots whose batch has finished (event query, no blocking),
//   - admits waiting requests into every free slot, in arrival order, while
//     the in-flight HBM footprint stays within the budget, and launches the
//     batch on the slot's stream (optionally as a replayed hipGraph).
// Batches on different slots overlap on the device; a request arriving while
// others run is admitted as soon as a slot and memory are free.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xpsi_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
    long long id;
    xh_request req;
    Clock::time_point t_submit;
    double bytes;
};

struct Done {
    long long id;
    double lnl;
    double latency;
};

struct Slot {
    xh_pipeline* pl = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Pending> batch;
    std::vector<xh_request> reqs;
    double bytes = 0.0;
};

}  // namespace

struct xh_server {
    std::vector<Slot> slots;
    int slot_batch = 1;
    double budget = 0.0, inflight = 0.0;
    double bytes_per_cell = 0.0, bytes_fixed = 0.0;
    std::deque<Pending> queue;
    std::vector<Done> done;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    std::atomic<bool> stop{false};
    long long next_id = 0;
    long long launched_batches = 0, completed = 0;
    double batch_fill = 0.0;
    std::string error;
};

static double footprint(const xh_server* s, const xh_request& r)
{
    // device bytes the request's work touches: mesh cells, ray tables and
    // per-parallel geometry scale with the mesh size
    double cells = 0.0;
    for (int g = 0; g < 2; g++) {
        int n = r.sqrt_num_cells[g] > 0 ? r.sqrt_num_cells[g] : 32;
        cells += (double)n * n;
    }
    return s->bytes_fixed + s->bytes_per_cell * cells;
}

static void serve_loop(xh_server* s)
{
    std::vector<double> res;
    while (!s->stop.load()) {
        bool progressed = false;
        // harvest finished slots
        for (auto& sl : s->slots) {
            if (!sl.busy) continue;
            if (hipEventQuery(sl.done) != hipSuccess) continue;
            const double* r = xh_pipeline_results(sl.pl);
            auto now = Clock::now();
            std::lock_guard<std::mutex> lk(s->mu);
            for (size_t k = 0; k < sl.batch.size(); k++) {
                double lat = std::chrono::duration<double>(now - sl.batch[k].t_submit).count();
                s->done.push_back({sl.batch[k].id, r[k], lat});
            }
            s->completed += (long long)sl.batch.size();
            s->inflight -= sl.bytes;
            sl.busy = false;
            sl.batch.clear();
            progressed = true;
        }
        // admit into free slots
        for (auto& sl : s->slots) {
            if (sl.busy) continue;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->queue.empty()) break;
                sl.bytes = 0.0;
                while (!s->queue.empty() && (int)sl.batch.size() < s->slot_batch) {
                    const Pending& q = s->queue.front();
                    if (s->inflight + sl.bytes + q.bytes > s->budget && !sl.batch.empty()) break;
                    if (s->inflight + sl.bytes + q.bytes > s->budget && s->inflight > 0.0) break;
                    sl.bytes += q.bytes;
                    sl.batch.push_back(q);
                    s->queue.pop_front();
                }
                if (sl.batch.empty()) continue;
                s->inflight += sl.bytes;
            }
            sl.reqs.resize(sl.batch.size());
            for (size_t k = 0; k < sl.batch.size(); k++) sl.reqs[k] = sl.batch[k].req;
            int rc = xh_pipeline_launch(sl.pl, (int)sl.reqs.size(), sl.reqs.data(), sl.stream);
            if (rc == XH_OK && hipEventRecord(sl.done, sl.stream) == hipSuccess) {
                sl.busy = true;
                s->launched_batches++;
                s->batch_fill += (double)sl.batch.size() / s->slot_batch;
            } else {
                std::lock_guard<std::mutex> lk(s->mu);
                s->error = std::string("launch failed: ") + xh_last_error();
                for (auto& q : sl.batch) s->done.push_back({q.id, 0.0 / 0.0, -1.0});
                s->inflight -= sl.bytes;
                sl.batch.clear();
            }
            progressed = true;
        }
        if (!progressed) {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

extern "C" xh_server* xh_server_create(const xh_model_config* cfg, int n_slots, int slot_batch,
                                       double hbm_budget_bytes, int use_graphs)
{
    if (!cfg || n_slots < 1 || slot_batch < 1) return nullptr;
    xh_server* s = new xh_server();
    s->slot_batch = slot_batch;
    s->budget = hbm_budget_bytes > 0.0 ? hbm_budget_bytes : 1e300;
    const double NR = cfg->num_rays, NL = cfg->num_leaves, NE = cfg->N_E;
    const double NP = cfg->num_phases > 0 ? cfg->num_phases : cfg->num_leaves;
    // per mesh cell: area + per-parallel ray tables and geometry amortised
    s->bytes_per_cell = 8.0 * (1.0 + 3.0 * NR / 64.0 + 5.0 * NL * 2.0 / 64.0);
    s->bytes_fixed = 8.0 * (1000.0 + NE * NP * 2.0 + (double)cfg->N_ch * NP * 2.0);
    xh_model_config c = *cfg;
    c.max_batch = slot_batch;
    s->slots.resize(n_slots);
    for (auto& sl : s->slots) {
        sl.pl = xh_pipeline_create(&c);
        if (!sl.pl || hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            for (auto& t : s->slots) {
                if (t.pl) xh_pipeline_destroy(t.pl);
                if (t.stream) (void)hipStreamDestroy(t.stream);
                if (t.done) (void)hipEventDestroy(t.done);
            }
            delete s;
            return nullptr;
        }
        xh_pipeline_use_graphs(sl.pl, use_graphs);
    }
    s->worker = std::thread(serve_loop, s);
    return s;
}

extern "C" int xh_server_submit(xh_server* s, const xh_request* req, long long* id_out)
{
    Pending p;
    p.req = *req;
    p.t_submit = Clock::now();
    p.bytes = footprint(s, *req);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        p.id = s->next_id++;
        s->queue.push_back(p);
    }
    s->cv.notify_one();
    if (id_out) *id_out = p.id;
    return XH_OK;
}

extern "C" int xh_server_poll(xh_server* s, long long* ids, double* lnl, double* latency, int max_n)
{
    std::lock_guard<std::mutex> lk(s->mu);
    int n = 0;
    while (n < max_n && n < (int)s->done.size()) {
        const Done& d = s->done[n];
        ids[n] = d.id;
        lnl[n] = d.lnl;
        latency[n] = d.latency;
        n++;
    }
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

// stats: queued, in-flight bytes, launched batches, completed, mean batch fill
extern "C" int xh_server_stats(xh_server* s, double* out, int n)
{
    std::lock_guard<std::mutex> lk(s->mu);
    double v[5] = {(double)s->queue.size(), s->inflight, (double)s->launched_batches, (double)s->completed,
                   s->launched_batches ? s->batch_fill / s->launched_batches : 0.0};
    for (int k = 0; k < n && k < 5; k++) out[k] = v[k];
    return XH_OK;
}

extern "C" void xh_server_destroy(xh_server* s)
{
    if (!s) return;
    s->stop.store(true);
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& sl : s->slots) {
        if (sl.busy) (void)hipEventSynchronize(sl.done);
        xh_pipeline_destroy(sl.pl);
        (void)hipStreamDestroy(sl.stream);
        (void)hipEventDestroy(sl.done);
    }
    delete s;
}
