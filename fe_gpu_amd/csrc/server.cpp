// Request server over a device-resident panel.
//
// Callers submit regression requests (fehip_request) from any thread; worker threads bound to the
// panel's device keep taking the queued requests in arrival order, each admitting as many as fit the
// slot limit and the HBM budget left by the batches already on the device (device bytes the requests'
// work buffers, masks and tables will hold), and solve them as one batch: their columns share the
// within-transform sweeps.  Workers run on their own streams, so a batch that arrives while another
// is on the device is admitted as soon as memory allows and runs alongside it.  Latency is measured
// from submit to completion.
//
// The reference has no serving layer: it estimates one specification per process run
// (src/src_main.f90); each request here follows the reference's estimator exactly as
// fehip_estimate does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fehip.h"
#include "common.hpp"
#include "estimate.hpp"

namespace fehip {
namespace {

using clk = std::chrono::steady_clock;

struct Ticket {
    fehip_request req{};
    std::vector<int32_t> iv_reg, iv_inst;  // owned copies of the option arrays
    double bytes = 0.0;
    clk::time_point t_submit, t_done;
    int status = 0;
    std::string msg;
    bool done = false;
    bool claimed = false;  // a waiter holds this ticket; a second waiter is refused
};

struct Server {
    Panel* panel = nullptr;
    double budget = 0.0;
    double inflight = 0.0;  // bytes held by the batches on the device
    int max_batch = 32;
    clk::duration window{};  // how long a partial batch waits for more requests
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, std::unique_ptr<Ticket>> tickets;
    int64_t next_ticket = 1;
    bool stop = false;
    std::vector<std::thread> workers;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched_requests = 0;
    double busy_s = 0.0;
    std::vector<double> latencies;
    clk::time_point t_first, t_last;
    bool any_done = false;
};

// Device bytes a request holds while its batch is solved: the within-transformed columns (FP32 in the
// mixed form), the panel-order weights and the solved tables of the mixed form.  The temporaries of
// the normal-equation and sandwich step are taken one request at a time, so they are reserved once
// per batch (batch_reserve), not per request.
double request_bytes(const Panel* p, const fehip_request& q) {
    const double n = (double)p->n;
    const double cols = 1.0 + std::max(0, q.K) + std::max(0, q.L);
    const bool f32 = q.opts.demean.precision == 1 && !(q.opts.n_iv_reg > 0 && q.L > 0);
    double groups = 0.0;
    for (int d = 0; d < p->m; ++d) groups += p->d[d].G;
    return (f32 ? 4.0 : 8.0) * n * cols + n + (f32 ? 8.0 * groups * cols : 0.0);
}

double batch_reserve(const Panel* p, const fehip_request& q) {
    const double n = (double)p->n;
    // mask words, residual and score rows of the widest request, refinement buffers
    return 4.0 * n + 8.0 * n * (4.0 + std::max(0, q.K) + std::max(0, q.L));
}

// Each worker runs its batches on its own default stream, so batches of different workers overlap on
// the device.  A worker admits the oldest queued requests that fit the budget left by the batches in
// flight; with nothing in flight the oldest request always goes.
void run_worker(Server* sv, bool capture) {
    (void)hipSetDevice(sv->panel->device);
    set_thread_capture(capture);
    for (;;) {
        std::vector<Ticket*> batch;
        double held = 0.0;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            auto first_fits = [&] {
                if (sv->queue.empty()) return false;
                if (sv->inflight <= 0.0) return true;
                const Ticket* t = sv->tickets[sv->queue.front()].get();
                return sv->inflight + t->bytes + batch_reserve(sv->panel, t->req) <= sv->budget;
            };
            sv->cv_work.wait(lk, [&] { return (sv->stop && sv->queue.empty()) || first_fits(); });
            if (sv->queue.empty()) {  // stop requested and drained
                lk.unlock();
                release_workspaces();  // this thread's workspaces
                return;
            }
            // batching window: a partial batch waits until it is full or its oldest request has waited
            // the window (requests submitted back to back then share one batch)
            if (sv->window.count() > 0 && !sv->stop && (int)sv->queue.size() < sv->max_batch) {
                const auto deadline = sv->tickets[sv->queue.front()]->t_submit + sv->window;
                sv->cv_work.wait_until(lk, deadline, [&] {
                    return sv->stop || sv->queue.empty() || (int)sv->queue.size() >= sv->max_batch;
                });
                if (!first_fits()) continue;  // taken by another worker, or memory in use meanwhile
            }
            // admission: arrival order, up to max_batch requests and the HBM budget
            double bytes = 0.0, reserve = 0.0;
            while (!sv->queue.empty() && (int)batch.size() < sv->max_batch) {
                Ticket* t = sv->tickets[sv->queue.front()].get();
                const double res = std::max(reserve, batch_reserve(sv->panel, t->req));
                if (!(batch.empty() && sv->inflight <= 0.0) && sv->inflight + bytes + t->bytes + res > sv->budget)
                    break;
                bytes += t->bytes;
                reserve = res;
                batch.push_back(t);
                sv->queue.pop_front();
            }
            held = bytes + reserve;
            sv->inflight += held;
        }
        const auto t0 = clk::now();
        std::vector<fehip_request> reqs(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) reqs[i] = batch[i]->req;
        (void)hipGetLastError();
        int e = estimate_requests(sv->panel, reqs.data(), (int)reqs.size());
        std::vector<int> status(batch.size(), e);
        std::vector<std::string> msgs(batch.size(), e ? std::string(last_error()) : std::string());
        if (e && batch.size() > 1) {
            // isolate the failing request(s): run each alone
            for (size_t i = 0; i < batch.size(); ++i) {
                (void)hipGetLastError();
                status[i] = estimate_requests(sv->panel, &reqs[i], 1);
                msgs[i] = status[i] ? std::string(last_error()) : std::string();
            }
        }
        const auto t1 = clk::now();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->inflight -= held;
            sv->busy_s += std::chrono::duration<double>(t1 - t0).count();
            sv->batches += 1;
            sv->batched_requests += (int64_t)batch.size();
            for (size_t i = 0; i < batch.size(); ++i) {
                Ticket* t = batch[i];
                t->status = status[i];
                t->msg = msgs[i];
                t->t_done = t1;
                t->done = true;
                if (status[i]) {
                    sv->failed += 1;
                } else {
                    sv->completed += 1;
                    sv->latencies.push_back(std::chrono::duration<double>(t1 - t->t_submit).count());
                }
            }
            if (!sv->any_done || t0 < sv->t_first) sv->t_first = t0;
            sv->any_done = true;
            if (t1 > sv->t_last) sv->t_last = t1;
        }
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();  // memory freed: another worker may admit
    }
}

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace
}  // namespace fehip

using namespace fehip;

struct fehip_server {
    Server s;
};

extern "C" int fehip_server_create(fehip_panel* panel, const fehip_server_opts* opts, fehip_server** out) {
    if (!panel || !out) return set_error("fehip_server_create: null argument");
    *out = nullptr;
    Panel* p = reinterpret_cast<Panel*>(panel);
    FEHIP_CHECK(hipSetDevice(p->device), "server device");
    size_t free_b = 0, total_b = 0;
    FEHIP_CHECK(hipMemGetInfo(&free_b, &total_b), "memory info");
    auto* srv = new fehip_server();
    Server& sv = srv->s;
    sv.panel = p;
    sv.budget = (opts && opts->hbm_budget > 0) ? (double)opts->hbm_budget
                                                 : 0.8 * (double)(free_b + cache_idle_bytes());
    sv.max_batch = (opts && opts->max_batch > 0) ? std::min(opts->max_batch, 256) : 32;
    const int win_us = !opts ? 2000 : (opts->window_us > 0 ? opts->window_us : (opts->window_us == 0 ? 2000 : 0));
    sv.window = std::chrono::duration_cast<clk::duration>(std::chrono::microseconds(win_us));
    const int nw = (opts && opts->workers > 0) ? std::min(opts->workers, 8) : 2;
    try {
        for (int w = 0; w < nw; ++w) sv.workers.emplace_back(run_worker, &sv, nw == 1);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(sv.mu);
            sv.stop = true;
        }
        sv.cv_work.notify_all();
        for (auto& t : sv.workers)
            if (t.joinable()) t.join();
        delete srv;
        return set_error("fehip_server_create: cannot start the worker threads");
    }
    *out = srv;
    return set_ok();
}

extern "C" int fehip_server_submit(fehip_server* srv, const fehip_request* req, int64_t* ticket) {
    if (!srv || !req || !ticket) return set_error("fehip_server_submit: null argument");
    Server& sv = srv->s;
    if (!req->d_y || !req->result || (req->K > 0 && (!req->d_W || !req->beta || !req->se)) || req->K < 0 ||
        req->L < 0 || (req->L > 0 && !req->d_Z) || req->ld < sv.panel->n)
        return set_error("fehip_server_submit: bad request");
    auto t = std::make_unique<Ticket>();
    t->req = *req;
    if (req->opts.n_iv_reg > 0 && req->opts.iv_reg)
        t->iv_reg.assign(req->opts.iv_reg, req->opts.iv_reg + req->opts.n_iv_reg);
    if (req->opts.n_iv_inst > 0 && req->opts.iv_inst)
        t->iv_inst.assign(req->opts.iv_inst, req->opts.iv_inst + req->opts.n_iv_inst);
    t->req.opts.iv_reg = t->iv_reg.empty() ? nullptr : t->iv_reg.data();
    t->req.opts.n_iv_reg = (int32_t)t->iv_reg.size();
    t->req.opts.iv_inst = t->iv_inst.empty() ? nullptr : t->iv_inst.data();
    t->req.opts.n_iv_inst = (int32_t)t->iv_inst.size();
    t->bytes = request_bytes(sv.panel, *req);
    t->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(sv.mu);
        if (sv.stop) return set_error("fehip_server_submit: server is shutting down");
        const int64_t id = sv.next_ticket++;
        sv.tickets.emplace(id, std::move(t));
        sv.queue.push_back(id);
        sv.submitted += 1;
        *ticket = id;
    }
    sv.cv_work.notify_all();
    return set_ok();
}

extern "C" int fehip_server_wait(fehip_server* srv, int64_t ticket, double* latency_s) {
    if (!srv) return set_error("fehip_server_wait: null server");
    Server& sv = srv->s;
    std::unique_lock<std::mutex> lk(sv.mu);
    auto it = sv.tickets.find(ticket);
    if (it == sv.tickets.end()) return set_error("fehip_server_wait: unknown ticket %lld", (long long)ticket);
    Ticket* t = it->second.get();
    // one waiter per ticket: the ticket is freed when its wait returns, so a second concurrent waiter
    // would read freed memory
    if (t->claimed) return set_error("fehip_server_wait: ticket %lld already has a waiter", (long long)ticket);
    t->claimed = true;
    sv.cv_done.wait(lk, [&] { return t->done; });
    const int status = t->status;
    const std::string msg = t->msg;
    if (latency_s) *latency_s = std::chrono::duration<double>(t->t_done - t->t_submit).count();
    sv.tickets.erase(it);
    lk.unlock();
    if (status) return set_error("%s", msg.c_str());
    return set_ok();
}

extern "C" int fehip_server_get_stats(fehip_server* srv, fehip_server_stats* st) {
    if (!srv || !st) return set_error("fehip_server_get_stats: null argument");
    Server& sv = srv->s;
    std::lock_guard<std::mutex> lk(sv.mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = sv.submitted;
    st->completed = sv.completed;
    st->failed = sv.failed;
    st->batches = sv.batches;
    st->queued = (int64_t)sv.queue.size();
    st->mean_batch = sv.batches ? (double)sv.batched_requests / (double)sv.batches : 0.0;
    st->busy_s = sv.busy_s;
    st->span_s = sv.any_done ? std::chrono::duration<double>(sv.t_last - sv.t_first).count() : 0.0;
    st->p50_s = quantile(sv.latencies, 0.50);
    st->p99_s = quantile(sv.latencies, 0.99);
    st->max_s = sv.latencies.empty() ? 0.0 : *std::max_element(sv.latencies.begin(), sv.latencies.end());
    st->budget_bytes = (int64_t)sv.budget;
    return set_ok();
}

extern "C" int fehip_server_reset_stats(fehip_server* srv) {
    if (!srv) return set_error("fehip_server_reset_stats: null server");
    Server& sv = srv->s;
    std::lock_guard<std::mutex> lk(sv.mu);
    sv.submitted = sv.completed = sv.failed = sv.batches = sv.batched_requests = 0;
    sv.busy_s = 0.0;
    sv.latencies.clear();
    sv.any_done = false;
    return set_ok();
}

extern "C" int fehip_server_destroy(fehip_server* srv) {
    if (!srv) return set_ok();
    {
        std::lock_guard<std::mutex> lk(srv->s.mu);
        srv->s.stop = true;
    }
    srv->s.cv_work.notify_all();
    for (auto& t : srv->s.workers)
        if (t.joinable()) t.join();
    delete srv;
    return set_ok();
}
