// Request loop of libpxh: spectrum requests on many targets served by one
// GPU with continuous batching.
//
// A request is a frequency-response solve (pxh_solve_spectrum arguments) on a
// context the caller prepared (atoms or mesh resident in HBM).  Requests are
// admitted in arrival order as soon as their HBM footprint fits the budget
// and their context is idle; a later, smaller request may pass a waiting
// larger one.  Each admitted request runs on its own context's stream, so the
// work of several requests overlaps on the device; inside a request the
// (frequency, field) systems are batched into solver slots as they free up.
// A request's result is the same alone and in the mix: nothing is shared
// between contexts and every reduction has a fixed order.
//
// The reference runs one structure per process and checks the memory of a
// run before it starts (memory_manager_module.f90:394-397); admission by
// footprint is that check applied per request.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "pxh_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    pxh_ctx *ctx = nullptr;
    std::vector<double> freq;
    double e0 = 0.0, tol = 0.0;
    int rhs_form = PXH_RHS_FIELD, pol_mask = 7, method = PXH_ITERATIVE_OTF, kdim = 0, precision = PXH_FP64;
    size_t bytes = 0;
    // results
    std::vector<double> polar, results, tres;
    std::vector<int32_t> iters;
    int status = PXH_OK;
    std::string error;
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
};

}  // namespace

struct pxh_server {
    int device = 0;
    size_t budget = 0, in_use = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::map<int64_t, Request *> requests;
    std::set<pxh_ctx *> busy;
    std::vector<std::thread> workers;
    bool stop = false;
    int64_t next_id = 1;
    std::string last_error;
};

namespace {

// First queued request (arrival order) that fits the free budget and whose
// context is idle; a request larger than the whole budget is admitted alone.
Request *take_admissible(pxh_server *s) {
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        Request *r = *it;
        if (s->busy.count(r->ctx)) continue;
        const bool fits = s->in_use + r->bytes <= s->budget || s->in_use == 0;
        if (!fits) continue;
        s->queue.erase(it);
        return r;
    }
    return nullptr;
}

void worker_loop(pxh_server *s) {
    (void)hipSetDevice(s->device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || (r = take_admissible(s)) != nullptr; });
            if (!r) return;  // stop requested and nothing admissible
            s->in_use += r->bytes;
            s->busy.insert(r->ctx);
        }
        r->t_start = Clock::now();
        const int nf = (int)r->freq.size();
        r->polar.assign((size_t)18 * nf, 0.0);
        r->results.assign((size_t)20 * nf, 0.0);
        r->iters.assign((size_t)3 * nf, 0);
        r->tres.assign((size_t)3 * nf, 0.0);
        r->status = pxh_solve_spectrum(r->ctx, nf, r->freq.data(), r->e0, r->rhs_form, r->pol_mask, r->method, r->tol,
                                       r->kdim, r->precision, nullptr, r->polar.data(), r->results.data(),
                                       r->iters.data(), r->tres.data());
        if (r->status != PXH_OK) r->error = r->ctx->last_error;
        r->t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->in_use -= r->bytes;
            s->busy.erase(r->ctx);
            r->done = true;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

int pxh_server_create(int device, size_t hbm_budget_bytes, int max_inflight, pxh_server **out) {
    if (!out || max_inflight < 1) return PXH_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PXH_ENODEV;
    if (device < 0 || device >= count) return PXH_EINVAL;
    pxh_server *s = new (std::nothrow) pxh_server();
    if (!s) return PXH_ENOMEM;
    s->device = device;
    size_t free_b = 0, total_b = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete s;
        return PXH_EHIP;
    }
    // default budget: 90 % of the HBM free now
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : free_b / 10 * 9;
    try {
        for (int k = 0; k < max_inflight; ++k) s->workers.emplace_back(worker_loop, s);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto &t : s->workers) t.join();
        delete s;
        return PXH_ENOMEM;
    }
    *out = s;
    return PXH_OK;
}

int pxh_server_submit(pxh_server *s, pxh_ctx *ctx, int n_freq, const double *freq_au, double e0, int rhs_form,
                      int pol_mask, int method, double tol_abs, int krylov_dim, int precision, int64_t *request_id) {
    if (!s || !ctx || !request_id || n_freq <= 0 || !freq_au) return PXH_EINVAL;
    if (ctx->device != s->device) return PXH_EINVAL;
    Request *r = new (std::nothrow) Request();
    if (!r) return PXH_ENOMEM;
    r->ctx = ctx;
    r->freq.assign(freq_au, freq_au + n_freq);
    r->e0 = e0;
    r->rhs_form = rhs_form;
    r->pol_mask = pol_mask;
    r->method = method;
    r->tol = tol_abs;
    r->kdim = krylov_dim;
    r->precision = precision;
    r->bytes = pxh::estimate_solve_bytes(ctx, n_freq, pol_mask, method, krylov_dim, precision);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->requests[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    *request_id = r->id;
    return PXH_OK;
}

int pxh_server_wait(pxh_server *s, int64_t request_id, double *polar, double *results, int32_t *iters,
                    double *true_residual, double *times_ms) {
    if (!s) return PXH_EINVAL;
    Request *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(request_id);
        if (it == s->requests.end()) {
            s->last_error = "unknown request id";
            return PXH_EINVAL;
        }
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->requests.erase(it);
    }
    const size_t nf = r->freq.size();
    if (polar) std::memcpy(polar, r->polar.data(), sizeof(double) * 18 * nf);
    if (results) std::memcpy(results, r->results.data(), sizeof(double) * 20 * nf);
    if (iters) std::memcpy(iters, r->iters.data(), sizeof(int32_t) * 3 * nf);
    if (true_residual) std::memcpy(true_residual, r->tres.data(), sizeof(double) * 3 * nf);
    if (times_ms) {
        times_ms[0] = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
        times_ms[1] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    }
    const int st = r->status;
    if (st != PXH_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = r->error;
    }
    delete r;
    return st;
}

int pxh_server_last_error(pxh_server *s, char *buf, int len) {
    if (!s || !buf || len <= 0) return PXH_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    const int k = std::min<int>(len - 1, (int)s->last_error.size());
    std::memcpy(buf, s->last_error.data(), k);
    buf[k] = '\0';
    return PXH_OK;
}

int pxh_server_destroy(pxh_server *s) {
    if (!s) return PXH_OK;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        // let every queued request finish
        s->cv_done.wait(lk, [&] {
            for (auto &kv : s->requests)
                if (!kv.second->done) return false;
            return true;
        });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto &kv : s->requests) delete kv.second;
    delete s;
    return PXH_OK;
}

}  // extern "C"
