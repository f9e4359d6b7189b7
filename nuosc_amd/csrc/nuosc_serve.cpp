// nuosc_serve.cpp -- continuous-batching run server over the batched RK4
// path (no reference counterpart: the reference runs one simulation per
// process).
//
// Requests (one parameter-scan run each) are queued by any host thread.  A
// dispatcher thread owns the GPU work:
//   * admission: a queued request is admitted when its device footprint
//     (nuosc_estimate_device_bytes) fits beside the resident runs' within
//     the HBM budget.  The queue is scanned in arrival order and any request
//     that fits is taken, so small runs are not held behind a large one;
//     once the oldest queued request has waited longer than the age bound,
//     no younger one is admitted before it (it takes the next memory that
//     frees).  An admitted run is created, initialised and analysed at t = 0;
//   * ticks: all resident runs advance together by k steps through one
//     batch (one launch per RK4 stage for all of them); k stops at the next
//     analysis or completion step of any run, so every run's series is
//     sampled exactly at its own cadence;
//   * retirement: a finished run records its completion time, frees its
//     device memory and leaves the batch, which lets queued runs in.
// A run's numbers do not depend on its batch mates (the batched kernels
// compute each run exactly as the single-run path does).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/nuosc_hip.h"

namespace {

struct Run {
    long long id = 0;
    nuosc_request req{};
    nuosc_sim* sim = nullptr;
    long long bytes = 0;
    int step = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = 0;
    std::string error;
    std::vector<double> series;  // [t, Pee, Peebar, maxdP, |M0|] per sample
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct nuosc_server {
    int device = 0, precision = NUOSC_PREC_FP64;
    long long budget = 0;
    int max_tick = 10;
    double age_bound = 30.0;               // seconds a request may be passed over
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Run*> queue;                // waiting, arrival order
    std::vector<Run*> active;              // resident in the batch (dispatcher only)
    std::map<long long, Run*> done;        // finished, not yet taken
    long long next_id = 1;
    bool stop = false;
    long long resident_bytes = 0;
    long long updates = 0;                 // site*dir RK4 updates done
    double busy_s = 0;                     // dispatcher time spent stepping
    int n_active = 0, n_queued = 0;
    long long n_done = 0;
    std::string err;

    nuosc_batch* batch = nullptr;
    std::thread worker;

    void loop();
    void admit_locked(std::unique_lock<std::mutex>& lk);
    void retire(Run* r, int status, const std::string& e);
    void sample(Run* r);
};

void nuosc_server::sample(Run* r) {
    nuosc_analysis_t a{};
    nuosc_analysis(r->sim, &a);
    r->series.insert(r->series.end(), {a.phy_time, a.surv, a.survb, a.maxdP, a.aM0});
}

void nuosc_server::retire(Run* r, int status, const std::string& e) {
    r->status = status;
    r->error = e;
    r->t_done = now_s(t0);
    if (r->sim) {
        if (batch) nuosc_batch_remove(batch, r->sim);
        nuosc_destroy(r->sim);
        r->sim = nullptr;
    }
    std::lock_guard<std::mutex> g(mu);
    resident_bytes -= r->bytes;
    active.erase(std::remove(active.begin(), active.end(), r), active.end());
    n_active = (int)active.size();
    done[r->id] = r;
    ++n_done;
    cv.notify_all();
}

// admit queued runs that fit (see the admission rule above); called with mu
// held, drops it around the GPU work of each admission (submit may append to
// the queue meanwhile, so each admission rescans it)
void nuosc_server::admit_locked(std::unique_lock<std::mutex>& lk) {
    while (true) {
        Run* r = nullptr;
        const double now = now_s(t0);
        for (size_t k = 0; k < queue.size(); ++k) {
            if (resident_bytes + queue[k]->bytes <= budget) {
                r = queue[k];
                queue.erase(queue.begin() + (long)k);
                break;
            }
            if (k == 0 && now - queue[0]->t_submit >= age_bound) break;   // the oldest goes first
        }
        if (!r) break;
        n_queued = (int)queue.size();
        resident_bytes += r->bytes;
        active.push_back(r);
        n_active = (int)active.size();
        lk.unlock();
        r->t_admit = now_s(t0);
        nuosc_config c = r->req.cfg;
        c.device = device;
        c.precision = precision;
        c.use_graph = 0;
        c.renorm = 0;
        int rc = nuosc_create(&c, &r->sim);
        if (rc == NUOSC_OK)
            rc = nuosc_fill_init(r->sim, r->req.ipt, r->req.alpha, r->req.eps0, r->req.sigma, r->req.lnue,
                                 r->req.lnueb, r->req.lnuex, r->req.lnuebx, r->req.seed);
        if (rc == NUOSC_OK) {
            sample(r);
            rc = nuosc_batch_add(batch, r->sim);
        }
        if (rc != NUOSC_OK) {
            std::string e = r->sim ? nuosc_last_error(r->sim) : nuosc_last_error(nullptr);
            retire(r, rc, e);
        } else if (r->req.nsteps <= 0) {
            retire(r, NUOSC_OK, "");
        }
        lk.lock();
    }
}

void nuosc_server::loop() {
    (void)hipSetDevice(device);  // the device is per host thread
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        admit_locked(lk);
        if (active.empty()) {
            if (stop && queue.empty()) break;
            cv.wait(lk, [&] { return stop || !queue.empty(); });
            continue;
        }
        if (stop) break;  // cancel: resident runs are dropped below
        // tick: up to the next analysis/completion step of any resident run
        int k = max_tick;
        for (Run* r : active) {
            const int ana = std::max(1, r->req.ana_every);
            const int to_ana = ana - (r->step % ana);
            const int to_end = r->req.nsteps - r->step;
            k = std::min(k, std::min(to_ana, to_end));
        }
        std::vector<Run*> act = active;
        lk.unlock();
        const double tb = now_s(t0);
        int rc = nuosc_batch_step(batch, k);
        const double te = now_s(t0);
        long long upd = 0;
        for (Run* r : act) {
            int nx, nz, nv;
            long long lp;
            nuosc_get_dims(r->sim, &nx, &nz, &nv, &lp);
            upd += (long long)nx * nz * nv * k;
        }
        {
            std::lock_guard<std::mutex> g(mu);
            busy_s += te - tb;
            if (rc == NUOSC_OK) updates += upd;
        }
        for (Run* r : act) {
            if (rc != NUOSC_OK) {
                retire(r, rc, nuosc_batch_last_error(batch));
                continue;
            }
            r->step += k;
            const int ana = std::max(1, r->req.ana_every);
            if (r->step % ana == 0 || r->step == r->req.nsteps) sample(r);
            if (r->step >= r->req.nsteps) retire(r, NUOSC_OK, "");
        }
        lk.lock();
    }
    // cancelled runs still resident
    std::vector<Run*> left = active;
    lk.unlock();
    for (Run* r : left) retire(r, NUOSC_EINVAL, "server stopped");
}

extern "C" {

int nuosc_server_create(int device, int precision, long long hbm_budget_bytes, int max_tick_steps,
                        nuosc_server** out) {
    if (!out || hbm_budget_bytes <= 0 || max_tick_steps < 1) return NUOSC_EINVAL;
    *out = nullptr;
    nuosc_server* s = new nuosc_server();
    s->device = device;
    s->precision = precision;
    s->budget = hbm_budget_bytes;
    s->max_tick = max_tick_steps;
    int rc = nuosc_batch_create(device, precision, &s->batch);
    if (rc != NUOSC_OK) {
        delete s;
        return rc;
    }
    s->worker = std::thread([s] { s->loop(); });  // owns every GPU call of the server
    *out = s;
    return NUOSC_OK;
}

int nuosc_server_submit(nuosc_server* s, const nuosc_request* req, long long* id) {
    if (!s || !req) return NUOSC_EINVAL;
    long long bytes = 0;
    nuosc_config c = req->cfg;
    c.precision = s->precision;
    // runs the batch cannot take (nuosc_batch_add) are refused here, not at admission
    if (c.bc != NUOSC_BC_PERIODIC || c.ko_order == 5 || c.adv != NUOSC_ADV_CFD4 || c.nflavor == 3 || c.renorm)
        return NUOSC_EINVAL;
    if (nuosc_estimate_device_bytes(&c, &bytes) != NUOSC_OK) return NUOSC_EINVAL;
    if (bytes > s->budget) return NUOSC_EINVAL;   // could never be admitted
    Run* r = new Run();
    r->req = *req;
    r->bytes = bytes;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) {
        delete r;
        return NUOSC_EINVAL;
    }
    r->id = s->next_id++;
    r->t_submit = now_s(s->t0);
    s->queue.push_back(r);
    s->n_queued = (int)s->queue.size();
    if (id) *id = r->id;
    s->cv.notify_all();
    return NUOSC_OK;
}

int nuosc_server_take_result(nuosc_server* s, long long id, double timeout_s, nuosc_result* info, double* series,
                             int max_samples) {
    if (!s || !info) return NUOSC_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return id < 0 ? !s->done.empty() : s->done.count(id) > 0; };
    if (!ready()) {
        if (timeout_s <= 0) return NUOSC_EAGAIN;
        if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return NUOSC_EAGAIN;
    }
    auto it = id < 0 ? s->done.begin() : s->done.find(id);
    Run* r = it->second;
    s->done.erase(it);
    lk.unlock();
    info->id = r->id;
    info->status = r->status;
    info->nsteps = r->step;
    info->t_submit = r->t_submit;
    info->t_admit = r->t_admit;
    info->t_done = r->t_done;
    info->device_bytes = r->bytes;
    info->n_samples = (int)(r->series.size() / 5);
    if (series && max_samples > 0) {
        const int n = std::min(info->n_samples, max_samples);
        std::copy(r->series.begin(), r->series.begin() + 5 * n, series);
    }
    delete r;
    return NUOSC_OK;
}

int nuosc_server_set_age_bound(nuosc_server* s, double seconds) {
    if (!s || !(seconds >= 0)) return NUOSC_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    s->age_bound = seconds;
    s->cv.notify_all();
    return NUOSC_OK;
}

int nuosc_server_stats(nuosc_server* s, nuosc_server_stats_t* st) {
    if (!s || !st) return NUOSC_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    st->queued = s->n_queued;
    st->active = s->n_active;
    st->finished = s->n_done;
    st->resident_bytes = s->resident_bytes;
    st->updates = (double)s->updates;
    st->busy_s = s->busy_s;
    st->now_s = now_s(s->t0);
    return NUOSC_OK;
}

int nuosc_server_destroy(nuosc_server* s) {
    if (!s) return NUOSC_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (auto& kv : s->done) delete kv.second;
    for (Run* r : s->queue) delete r;
    nuosc_batch_destroy(s->batch);
    delete s;
    return NUOSC_OK;
}

}  // extern "C"
