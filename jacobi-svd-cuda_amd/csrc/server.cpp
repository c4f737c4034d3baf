// server.cpp -- request loop with continuous batching and HBM-budget admission (jsvd_server_* in jsvd.h).
//
//   submit()  copies the request into host memory and appends it to the pending FIFO.
//   lanes     `lanes` worker threads, each with its own HIP stream, a batched solver sized for the small
//             class and a cached solver plan.  A free lane admits pending requests while the byte budget
//             allows, then takes the oldest admitted work: one large request (solved by a plan), or every
//             admitted small request at once (one persistent batched launch).  Requests arriving meanwhile
//             wait for the next free lane, so launches of different lanes overlap on the device.
//   wait()    blocks until a request is done and hands its results over.
//
// The batched kernel and the plans compute each request independently of its neighbours, so a request's
// result is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace
{
    using Clock = std::chrono::steady_clock;

    struct Request
    {
        uint64_t id = 0;
        int64_t m = 0;
        int64_t n = 0;
        int precision = JSVD_PRECISION_F64;
        double tol = 1e-14;
        int64_t bytes = 0;
        bool small = false;
        std::vector<double> a;
        std::vector<double> u;
        std::vector<double> sigma;
        std::vector<double> v;
        Clock::time_point submitted;
        Clock::time_point started;
        Clock::time_point finished;
        int sweeps = 0;
        int batch_size = 0;
        int status = JSVD_OK;
        std::string error;
        bool done = false;
    };

    template <typename T>
    struct PinnedBuffer
    {
        T *ptr = nullptr;
        size_t cap = 0;
        ~PinnedBuffer()
        {
            if (ptr)
            {
                (void)hipHostFree(ptr);
            }
        }
        T *ensure(size_t count)
        {
            if (count > cap)
            {
                if (ptr)
                {
                    (void)hipHostFree(ptr);
                    ptr = nullptr;
                }
                size_t c = std::max<size_t>(count, 2 * cap);
                JSVD_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&ptr), c * sizeof(T), hipHostMallocDefault));
                cap = c;
            }
            return ptr;
        }
    };

    double seconds(Clock::duration d) { return std::chrono::duration<double>(d).count(); }
} // namespace

struct jsvd_server
{
    jsvd_server_config cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> pending;
    std::deque<std::shared_ptr<Request>> ready_small;
    std::deque<std::shared_ptr<Request>> ready_large;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
    uint64_t next_id = 1;
    int64_t budget_used = 0;
    int64_t in_flight = 0;
    bool stopping = false;
    jsvd_server_stats stats{};
    std::vector<std::thread> lanes;

    static int64_t footprint(int64_t m, int64_t n)
    {
        // input, U, sigma, V and the solver's working copies of A and V, in doubles
        return 8 * (3 * m * n + 2 * n * n + n);
    }

    // Called with the lock held: moves pending requests into the ready queues while the budget allows (the
    // oldest request is always admitted when nothing else is in flight, so an oversized one cannot starve).
    void admit()
    {
        while (!pending.empty())
        {
            const auto &r = pending.front();
            if (budget_used + r->bytes > cfg.hbm_budget_bytes && in_flight > 0)
            {
                break;
            }
            budget_used += r->bytes;
            ++in_flight;
            stats.budget_peak = std::max(stats.budget_peak, budget_used);
            stats.max_in_flight = std::max(stats.max_in_flight, in_flight);
            (r->small ? ready_small : ready_large).push_back(r);
            pending.pop_front();
        }
    }

    // Called with the lock held.  Returns the next job or an empty vector.
    std::vector<std::shared_ptr<Request>> take_job()
    {
        admit();
        std::vector<std::shared_ptr<Request>> job;
        const bool have_small = !ready_small.empty();
        const bool have_large = !ready_large.empty();
        if (!have_small && !have_large)
        {
            return job;
        }
        const bool large_first =
            have_large && (!have_small || ready_large.front()->submitted <= ready_small.front()->submitted);
        if (large_first)
        {
            job.push_back(ready_large.front());
            ready_large.pop_front();
        }
        else
        {
            while (!ready_small.empty() && static_cast<int>(job.size()) < cfg.max_batch)
            {
                job.push_back(ready_small.front());
                ready_small.pop_front();
            }
        }
        return job;
    }

    void finish(const std::vector<std::shared_ptr<Request>> &job)
    {
        {
            std::lock_guard<std::mutex> lock(mu);
            const auto now = Clock::now();
            for (const auto &r : job)
            {
                r->finished = now;
                r->done = true;
                budget_used -= r->bytes;
                --in_flight;
                ++stats.completed;
            }
            stats.budget_in_use = budget_used;
        }
        cv.notify_all();
    }

    struct LaneState
    {
        hipStream_t stream = nullptr;
        jsvd_batch *batch = nullptr;
        // One plan per path, sized for the largest request the lane has seen and re-targeted at each request's
        // shape and tolerance (jsvd_plan_reshape): no allocation per request.
        jsvd_plan *plan_block = nullptr;
        int64_t block_cap_m = 0;
        int64_t block_cap_n = 0;
        jsvd_plan *plan_plain = nullptr;
        int64_t plain_cap_m = 0;
        int64_t plain_cap_n = 0;
        PinnedBuffer<double> a, u, s, v;
        PinnedBuffer<jsvd_batch_item> items;
        PinnedBuffer<int32_t> sweeps;
        ~LaneState()
        {
            for (jsvd_plan *p : {plan_block, plan_plain})
            {
                if (p)
                {
                    (void)jsvd_plan_destroy(p);
                }
            }
            if (batch)
            {
                (void)jsvd_batch_destroy(batch);
            }
            if (stream)
            {
                (void)hipStreamDestroy(stream);
            }
        }
    };

    static void raise(int status)
    {
        if (status != JSVD_OK)
        {
            throw jsvd::Error(status, jsvd_last_error());
        }
    }

    void run_large(LaneState &ls, Request &r)
    {
        // the block paths need two 32-column blocks
        const bool block = cfg.large_path != JSVD_PATH_PLAIN_F64 && r.n > 32;
        jsvd_plan *&plan = block ? ls.plan_block : ls.plan_plain;
        int64_t &cap_m = block ? ls.block_cap_m : ls.plain_cap_m;
        int64_t &cap_n = block ? ls.block_cap_n : ls.plain_cap_n;
        if (!plan || r.m > cap_m || r.n > cap_n)
        {
            if (plan)
            {
                (void)jsvd_plan_destroy(plan);
                plan = nullptr;
            }
            cap_m = std::max(cap_m, r.m); // cap_m >= r.m >= r.n and every earlier n: a valid m >= n shape
            cap_n = std::max(cap_n, r.n);
            jsvd_config c;
            jsvd_config_default(&c);
            c.epsilon = r.tol;
            c.max_sweeps = cfg.max_sweeps;
            jsvd_mixed_config mx;
            jsvd_mixed_config_default(&mx);
            mx.path = block ? cfg.large_path : JSVD_PATH_PLAIN_F64;
            raise(jsvd_plan_create_ex(cap_m, cap_n, &c, &mx, ls.stream, &plan));
        }
        raise(jsvd_plan_reshape(plan, r.m, r.n, r.tol));
        r.u.resize(r.m * r.n);
        r.sigma.resize(r.n);
        r.v.resize(r.n * r.n);
        int32_t sweeps = 0;
        raise(jsvd_plan_solve_host(plan, r.a.data(), r.u.data(), r.sigma.data(), r.v.data(), &sweeps));
        r.sweeps = sweeps;
        r.batch_size = 1;
    }

    void run_batch(LaneState &ls, std::vector<std::shared_ptr<Request>> &job)
    {
        const size_t k = job.size();
        int64_t na = 0, nu = 0, ns = 0, nv = 0;
        jsvd_batch_item *items = ls.items.ensure(k);
        // most expensive first: better balance across the persistent workgroups
        std::vector<size_t> order(k);
        for (size_t i = 0; i < k; ++i)
        {
            order[i] = i;
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
            return job[x]->m * job[x]->n * job[x]->n > job[y]->m * job[y]->n * job[y]->n;
        });
        for (size_t q = 0; q < k; ++q)
        {
            const Request &r = *job[order[q]];
            jsvd_batch_item &it = items[q];
            it.a_offset = na;
            it.u_offset = nu;
            it.sigma_offset = ns;
            it.v_offset = nv;
            it.rows = static_cast<int32_t>(r.m);
            it.columns = static_cast<int32_t>(r.n);
            it.precision = r.precision;
            it.max_sweeps = cfg.max_sweeps;
            it.tolerance = r.tol;
            na += r.m * r.n;
            nu += r.m * r.n;
            ns += r.n;
            nv += r.n * r.n;
        }
        double *ha = ls.a.ensure(na);
        for (size_t q = 0; q < k; ++q)
        {
            const Request &r = *job[order[q]];
            std::memcpy(ha + items[q].a_offset, r.a.data(), sizeof(double) * r.a.size());
        }
        double *hu = ls.u.ensure(nu);
        double *hs = ls.s.ensure(ns);
        double *hv = ls.v.ensure(nv);
        int32_t *hsw = ls.sweeps.ensure(k);
        void *d_items = nullptr;
        double *d_a = nullptr, *d_u = nullptr, *d_s = nullptr, *d_v = nullptr;
        int32_t *d_sw = nullptr;
        JSVD_HIP_CHECK(hipMallocAsync(&d_items, sizeof(jsvd_batch_item) * k, ls.stream));
        JSVD_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_a), sizeof(double) * na, ls.stream));
        JSVD_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_u), sizeof(double) * nu, ls.stream));
        JSVD_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_s), sizeof(double) * ns, ls.stream));
        JSVD_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_v), sizeof(double) * nv, ls.stream));
        JSVD_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&d_sw), sizeof(int32_t) * k, ls.stream));
        JSVD_HIP_CHECK(hipMemcpyAsync(d_items, items, sizeof(jsvd_batch_item) * k, hipMemcpyHostToDevice, ls.stream));
        JSVD_HIP_CHECK(hipMemcpyAsync(d_a, ha, sizeof(double) * na, hipMemcpyHostToDevice, ls.stream));
        jsvd_batch_buffers buf{d_a, d_u, d_s, d_v, d_sw, na, nu, ns, nv};
        const int st = jsvd_batch_solve_async(ls.batch, static_cast<const jsvd_batch_item *>(d_items),
                                              static_cast<int64_t>(k), &buf, nullptr);
        if (st == JSVD_OK)
        {
            JSVD_HIP_CHECK(hipMemcpyAsync(hu, d_u, sizeof(double) * nu, hipMemcpyDeviceToHost, ls.stream));
            JSVD_HIP_CHECK(hipMemcpyAsync(hs, d_s, sizeof(double) * ns, hipMemcpyDeviceToHost, ls.stream));
            JSVD_HIP_CHECK(hipMemcpyAsync(hv, d_v, sizeof(double) * nv, hipMemcpyDeviceToHost, ls.stream));
            JSVD_HIP_CHECK(hipMemcpyAsync(hsw, d_sw, sizeof(int32_t) * k, hipMemcpyDeviceToHost, ls.stream));
        }
        for (void *p : {static_cast<void *>(d_items), static_cast<void *>(d_a), static_cast<void *>(d_u),
                        static_cast<void *>(d_s), static_cast<void *>(d_v), static_cast<void *>(d_sw)})
        {
            (void)hipFreeAsync(p, ls.stream);
        }
        JSVD_HIP_CHECK(hipStreamSynchronize(ls.stream));
        raise(st);
        for (size_t q = 0; q < k; ++q)
        {
            Request &r = *job[order[q]];
            const jsvd_batch_item &it = items[q];
            r.u.assign(hu + it.u_offset, hu + it.u_offset + r.m * r.n);
            r.sigma.assign(hs + it.sigma_offset, hs + it.sigma_offset + r.n);
            r.v.assign(hv + it.v_offset, hv + it.v_offset + r.n * r.n);
            r.sweeps = hsw[q];
            r.batch_size = static_cast<int>(k);
            if (r.sweeps < 0)
            {
                r.status = JSVD_ERR_INVALID_ARGUMENT;
                r.error = "request rejected by the batched kernel's bounds checks";
            }
        }
    }

    void lane_main()
    {
        (void)hipSetDevice(device);
        LaneState ls;
        std::string init_error; // this lane's own: read by no other thread
        try
        {
            JSVD_HIP_CHECK(hipStreamCreateWithFlags(&ls.stream, hipStreamNonBlocking));
            raise(jsvd_batch_create(cfg.small_max_rows, cfg.small_max_columns, ls.stream, &ls.batch));
        }
        catch (const std::exception &e)
        {
            init_error = e.what();
        }
        for (;;)
        {
            std::vector<std::shared_ptr<Request>> job;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv.wait(lock, [&] {
                    return stopping || !pending.empty() || !ready_small.empty() || !ready_large.empty();
                });
                job = take_job();
                if (job.empty())
                {
                    if (stopping && pending.empty() && ready_small.empty() && ready_large.empty())
                    {
                        return;
                    }
                    // budget exhausted with work in flight: sleep until a completion (or new work) notifies
                    cv.wait(lock);
                    continue;
                }
                const auto now = Clock::now();
                for (auto &r : job)
                {
                    r->started = now;
                }
                if (job.size() == 1 && !job[0]->small)
                {
                    ++stats.plans;
                }
                else
                {
                    ++stats.batches;
                }
            }
            try
            {
                if (!ls.batch)
                {
                    throw jsvd::Error(JSVD_ERR_DEVICE, "lane initialisation failed: " + init_error);
                }
                if (job.size() == 1 && !job[0]->small)
                {
                    run_large(ls, *job[0]);
                }
                else
                {
                    run_batch(ls, job);
                }
            }
            catch (const jsvd::Error &e)
            {
                for (auto &r : job)
                {
                    r->status = e.status();
                    r->error = e.what();
                }
            }
            catch (const std::exception &e)
            {
                for (auto &r : job)
                {
                    r->status = JSVD_ERR_INTERNAL;
                    r->error = e.what();
                }
            }
            finish(job);
        }
    }
};

extern "C" {

void jsvd_server_config_default(jsvd_server_config *cfg)
{
    if (!cfg)
    {
        return;
    }
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->lanes = 8;
    cfg->small_max_rows = 256;
    cfg->small_max_columns = 128;
    cfg->max_batch = 4096;
    cfg->hbm_budget_bytes = int64_t{64} << 30;
    cfg->max_sweeps = 60;
    cfg->large_path = JSVD_PATH_PLAIN_F64;
}

int jsvd_server_create(const jsvd_server_config *cfg, jsvd_server **server)
{
    return jsvd::guarded([&] {
        if (!server)
        {
            jsvd::throw_invalid("server output pointer must not be null.");
        }
        *server = nullptr;
        jsvd_server_config c;
        if (cfg)
        {
            c = *cfg;
        }
        else
        {
            jsvd_server_config_default(&c);
        }
        if (c.lanes < 1 || c.lanes > 32 || c.small_max_columns < 1 || c.small_max_rows < c.small_max_columns ||
            c.small_max_rows > 4096 || c.max_batch < 1 || c.hbm_budget_bytes <= 0 || c.max_sweeps < 1 ||
            c.large_path < JSVD_PATH_PLAIN_F64 || c.large_path > JSVD_PATH_MIXED)
        {
            jsvd::throw_invalid("invalid server configuration.");
        }
        auto s = std::make_unique<jsvd_server>();
        s->cfg = c;
        JSVD_HIP_CHECK(hipGetDevice(&s->device));
        for (int i = 0; i < c.lanes; ++i)
        {
            s->lanes.emplace_back([p = s.get()] { p->lane_main(); });
        }
        *server = s.release();
    });
}

int jsvd_server_destroy(jsvd_server *server)
{
    return jsvd::guarded([&] {
        if (!server)
        {
            return;
        }
        {
            std::lock_guard<std::mutex> lock(server->mu);
            server->stopping = true;
        }
        server->cv.notify_all();
        for (auto &t : server->lanes)
        {
            t.join();
        }
        delete server;
    });
}

int jsvd_server_submit(jsvd_server *server, const double *a, int64_t m, int64_t n, double relerr, uint64_t *id)
{
    return jsvd::guarded([&] {
        if (!server || !a || !id)
        {
            jsvd::throw_invalid("server, matrix and id pointers must not be null.");
        }
        jsvd::validate_shape(m, n);
        if (!(relerr > 0.0))
        {
            jsvd::throw_invalid("relerr must be positive.");
        }
        auto r = std::make_shared<Request>();
        r->m = m;
        r->n = n;
        r->precision = (relerr >= 1e-7) ? JSVD_PRECISION_F32 : JSVD_PRECISION_F64;
        r->tol = (r->precision == JSVD_PRECISION_F32)
                     ? 1e-6
                     : std::max(1e-14, 2.0 * std::sqrt(static_cast<double>(m)) * std::ldexp(1.0, -53));
        r->small = m <= server->cfg.small_max_rows && n <= server->cfg.small_max_columns &&
                   n >= JSVD_BATCH_MIN_COLUMNS;
        r->bytes = jsvd_server::footprint(m, n);
        r->a.assign(a, a + m * n);
        r->submitted = Clock::now();
        {
            std::lock_guard<std::mutex> lock(server->mu);
            if (server->stopping)
            {
                jsvd::throw_invalid("server is shutting down.");
            }
            r->id = server->next_id++;
            server->requests.emplace(r->id, r);
            server->pending.push_back(r);
            ++server->stats.submitted;
        }
        server->cv.notify_all();
        *id = r->id;
    });
}

int jsvd_server_wait(jsvd_server *server, uint64_t id, double *u, double *sigma, double *v, jsvd_request_info *info)
{
    std::shared_ptr<Request> r;
    const int st = jsvd::guarded([&] {
        if (!server)
        {
            jsvd::throw_invalid("server must not be null.");
        }
        std::unique_lock<std::mutex> lock(server->mu);
        auto it = server->requests.find(id);
        if (it == server->requests.end())
        {
            jsvd::throw_invalid("unknown request id.");
        }
        r = it->second;
        server->cv.wait(lock, [&] { return r->done; });
        server->requests.erase(it);
    });
    if (st != JSVD_OK)
    {
        return st;
    }
    if (info)
    {
        info->status = r->status;
        info->sweeps = r->sweeps;
        info->precision = r->precision;
        info->batch_size = r->batch_size;
        info->queued_s = seconds(r->started - r->submitted);
        info->latency_s = seconds(r->finished - r->submitted);
    }
    if (r->status != JSVD_OK)
    {
        jsvd::set_last_error(r->error);
        return r->status;
    }
    if (u)
    {
        std::memcpy(u, r->u.data(), sizeof(double) * r->u.size());
    }
    if (sigma)
    {
        std::memcpy(sigma, r->sigma.data(), sizeof(double) * r->sigma.size());
    }
    if (v)
    {
        std::memcpy(v, r->v.data(), sizeof(double) * r->v.size());
    }
    return JSVD_OK;
}

int jsvd_server_get_stats(jsvd_server *server, jsvd_server_stats *stats)
{
    return jsvd::guarded([&] {
        if (!server || !stats)
        {
            jsvd::throw_invalid("server and stats must not be null.");
        }
        std::lock_guard<std::mutex> lock(server->mu);
        *stats = server->stats;
        stats->budget_in_use = server->budget_used;
    });
}

} // extern "C"
