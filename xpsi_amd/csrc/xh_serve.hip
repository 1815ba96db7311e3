// Continuous-batching likelihood server (host runtime).
//
// A server owns `n_slots` pipelines, each with its own HIP stream and device
// buffers for `slot_batch` requests.  Requests are submitted from any thread
// into a FIFO; one dispatcher thread
//   - harvests slots whose batch has finished (event query, no blocking; a
//     failed batch completes its requests with NaN and records the error),
//   - forms a batch for a free slot when a full batch is queued, when the
//     oldest queued request has waited max_wait, or when no slot is busy
//     (an idle device never waits), and
//   - admits queued requests into it in arrival order while the device
//     footprint of all batches in flight stays within the budget, then
//     launches it on the slot's stream (optionally as a replayed hipGraph).
// Batches on different slots overlap on the device; a request arriving while
// others run is admitted as soon as a slot and memory are free.
//
// The footprint of a request is the device memory its work occupies in a
// slot's buffers: per member (super and ceding ring of each region) the mesh
// (S^2 cell areas, S-sized parallel arrays), the ray tables (3 S N_R), the
// per-parallel geometry (5 S N_L per image order, two orders assumed) and the
// flux (N_E N_P), with S the request's sqrt_num_cells; per request the
// energy-integrated and folded signals and the likelihood's per-channel
// arrays.
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
    double budget = 0.0, inflight = 0.0, peak_inflight = 0.0;
    double max_wait = 0.0;       // seconds a partial batch may wait for more requests
    int max_batch_seen = 0;
    bool use_graphs = false;
    // footprint model (bytes), from the model configuration
    int n_regions = 1;
    int members[2] = {1, 1};     // members per region (super + ceding ring)
    int snc[2] = {32, 32};       // configured sqrt_num_cells per region
    double per_cell = 0.0, per_side = 0.0, per_member = 0.0, per_request = 0.0;
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
    double b = s->per_request;
    for (int g = 0; g < s->n_regions; g++) {
        const double n = r.sqrt_num_cells[g] > 0 ? r.sqrt_num_cells[g] : s->snc[g];
        b += s->members[g] * (s->per_member + s->per_side * n + s->per_cell * n * n);
    }
    return b;
}

static void serve_loop(xh_server* s)
{
    std::vector<double> res;
    while (!s->stop.load()) {
        bool progressed = false;
        // harvest finished slots
        int busy = 0;
        for (auto& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) { busy++; continue; }
            const bool ok = q == hipSuccess;
            const double* r = xh_pipeline_results(sl.pl);
            auto now = Clock::now();
            std::lock_guard<std::mutex> lk(s->mu);
            if (!ok) s->error = std::string("batch failed: ") + hipGetErrorString(q);
            for (size_t k = 0; k < sl.batch.size(); k++) {
                double lat = std::chrono::duration<double>(now - sl.batch[k].t_submit).count();
                s->done.push_back({sl.batch[k].id, ok ? r[k] : 0.0 / 0.0, lat});
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
                // batch formation: a full batch, a request at its wait limit,
                // or an idle device
                const double waited =
                    std::chrono::duration<double>(Clock::now() - s->queue.front().t_submit).count();
                if ((int)s->queue.size() < s->slot_batch && waited < s->max_wait && busy > 0) break;
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
                if (s->inflight > s->peak_inflight) s->peak_inflight = s->inflight;
                if ((int)sl.batch.size() > s->max_batch_seen) s->max_batch_seen = (int)sl.batch.size();
            }
            sl.reqs.resize(sl.batch.size());
            for (size_t k = 0; k < sl.batch.size(); k++) sl.reqs[k] = sl.batch[k].req;
            int rc = xh_pipeline_launch(sl.pl, (int)sl.reqs.size(), sl.reqs.data(), sl.stream);
            if (rc == XH_OK && hipEventRecord(sl.done, sl.stream) == hipSuccess) {
                sl.busy = true;
                busy++;
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
    s->n_regions = cfg->n_regions;
    for (int g = 0; g < cfg->n_regions && g < 2; g++) {
        s->members[g] = cfg->region[g].cede ? 2 : 1;
        s->snc[g] = cfg->region[g].sqrt_num_cells;
    }
    s->per_cell = 8.0 + 8.0;                                    // cell areas, boundary-cell list
    s->per_side = 8.0 * (10.0 + 3.0 + 3.0 * NR + 2.0 * 5.0 * NL) + 4.0 * 2.0 * 2.0;
    s->per_member = 8.0 * NE * NP + 256.0;                      // flux, descriptor
    s->per_request = 8.0 * (cfg->n_regions * ((double)cfg->N_in + cfg->N_ch) * NP) +
                     8.0 * (double)cfg->N_ch * (cfg->N_b + 4.0) + 4.0 * cfg->N_ch;
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
        s->use_graphs = use_graphs != 0;
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

// stats: queued, in-flight bytes, launched batches, completed, mean batch
// fill, peak in-flight bytes, largest batch, failed batches recorded (0/1)
extern "C" int xh_server_stats(xh_server* s, double* out, int n)
{
    std::lock_guard<std::mutex> lk(s->mu);
    double v[8] = {(double)s->queue.size(), s->inflight, (double)s->launched_batches, (double)s->completed,
                   s->launched_batches ? s->batch_fill / s->launched_batches : 0.0, s->peak_inflight,
                   (double)s->max_batch_seen, s->error.empty() ? 0.0 : 1.0};
    for (int k = 0; k < n && k < 8; k++) out[k] = v[k];
    return XH_OK;
}

extern "C" int xh_server_warm(xh_server* s)
{
    if (!s) return XH_EINVAL;
    if (!s->use_graphs) return XH_OK;
    std::lock_guard<std::mutex> lk(s->mu);   // the dispatcher launches nothing meanwhile
    for (auto& sl : s->slots) {
        if (sl.busy && hipEventSynchronize(sl.done) != hipSuccess) return XH_EHIP;
        for (int n = 1; n <= s->slot_batch; n++) {
            int rc = xh_pipeline_capture(sl.pl, n, sl.stream);
            if (rc != XH_OK) return rc;
        }
    }
    return XH_OK;
}

extern "C" int xh_server_set_policy(xh_server* s, double max_wait_s)
{
    if (!s || !(max_wait_s >= 0.0)) return XH_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_wait = max_wait_s;
    return XH_OK;
}

extern "C" double xh_server_footprint(xh_server* s, const xh_request* req)
{
    return (s && req) ? footprint(s, *req) : 0.0;
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
