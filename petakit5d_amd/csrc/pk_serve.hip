// pk_serve.hip -- request loop serving decon jobs on one device.
//
// Continuous batching: a scheduler thread admits queued requests (FIFO) as
// device memory under the context's HBM budget allows, gives each its own
// device state (estimate, spectrum, staging) and pinned OTFs, and advances
// ALL admitted requests by one iteration per step through the batched step
// engine (one launch per pass over every request in flight).  Requests that
// arrive while others run join at the next step; a finished request gets its
// epilogue and output copy, its state is freed, and its latency (submit ->
// output ready) is recorded.  Because the step engine computes every volume
// exactly as the single-volume path does, a request's result is identical
// alone and in any mix.
//
// The reference has no serving loop: its nearest analogues are the dataset
// wrapper's per-file retry loop (XR_decon_data_wrapper.m:332-511) and the
// per-chunk loops of the large-volume drivers (RLdecon_large_in_memory.m:190-260).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "pk_engine.hpp"

namespace pk {

using Clock = std::chrono::steady_clock;

struct ServeReq {
    uint64_t id = 0;
    Chunk c;
    pk_vol in{}, out{};
    std::vector<float> pf, pb;
    int64_t fd[3] = {0, 0, 0}, bd[3] = {0, 0, 0};
    uint64_t hf = 0, hb = 0;  // OTF cache keys of pf / pb, hashed at submission
    DevBuf mem;
    bool pinned = false;
    Clock::time_point t_sub, t_adm, t_done;
    int state = PK_REQ_QUEUED;  // written under Server::mq only
    int iters = 0;
    pk_status st = PK_OK;
    std::string err;
};

struct Server {
    std::thread th;
    std::mutex mq;
    std::condition_variable cv, cv_done;
    std::deque<std::shared_ptr<ServeReq>> queue;
    std::map<uint64_t, std::shared_ptr<ServeReq>> reqs;
    std::vector<std::shared_ptr<ServeReq>> active;  // scheduler thread only
    // finished on the host side, output copy in flight: retired when the
    // event recorded after their last step and output copy completes
    // (scheduler thread only)
    std::vector<std::pair<std::shared_ptr<ServeReq>, hipEvent_t>> draining;
    // admitted, host input still being copied on the copy stream: a request
    // joins the steps once its copy's event completes (scheduler thread only)
    std::vector<std::pair<std::shared_ptr<ServeReq>, hipEvent_t>> staging;
    std::vector<hipEvent_t> ev_free;
    hipEvent_t take_event() {
        hipEvent_t ev = nullptr;
        if (!ev_free.empty()) {
            ev = ev_free.back();
            ev_free.pop_back();
        } else {
            PK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        return ev;
    }
    bool stop = false;
    uint64_t next_id = 1;
    // requests in flight at most.  Capacity of the served mix (144 requests
    // at once, Gvoxel/s, one box, two runs each): 16 -> 3.93-3.97, 24 ->
    // 3.79-3.93, 32 -> 3.89-3.92, 48 -> 3.77-3.81, 64 -> 3.72-3.78, 128 ->
    // 3.61-3.71, with the same p99 at 0.7x load
    int max_active = 32;
    // failure injection for tests of the failure path (PK_SERVE_INJECT_FAILURE="id:k",
    // read at start): request `id` fails on the host before its iteration k + 1
    uint64_t inject_id = 0;
    int inject_k = 0;
    pk_serve_stats stats{};
};

static void unpin(pk_ctx_impl* ctx, ServeReq& r) {
    if (!r.pinned) return;
    if (r.c.Hf) ctx->otf_pin(r.c.Hf, -1);
    if (r.c.Hb) ctx->otf_pin(r.c.Hb, -1);
    r.pinned = false;
}

// Once admitted, a request's OTFs are pinned in the cache: its PSF copies are
// no longer read.
static void drop_psfs(ServeReq& r) {
    std::vector<float>().swap(r.pf);
    std::vector<float>().swap(r.pb);
}

// The input-copy stream (created on first use).  Request arenas come from the
// context's block cache, whose blocks are free of queued work, so a copy on
// this stream need not wait for the steps queued on the engine stream.
static hipStream_t h2d_stream_of(pk_ctx_impl* ctx) {
    if (!ctx->h2d_stream) PK_HIP(hipStreamCreateWithFlags(&ctx->h2d_stream, hipStreamNonBlocking));
    return ctx->h2d_stream;
}

// Before a request's arena goes back to the cache on a failure path: wait for
// whatever may still use it (a no-op when the device is already in error).
static void quiesce(pk_ctx_impl* ctx) {
    for (hipStream_t st : {ctx->h2d_stream, ctx->copy_stream, ctx->stream})
        if (st) (void)hipStreamSynchronize(st);
    (void)hipGetLastError();
}

// Stage an admitted request's input: a host volume is copied on the input-copy
// stream and the request waits in S.staging until that copy completes; a
// device volume is read in place and the request joins the steps at once.
static void place(pk_ctx_impl* ctx, Server& S, const std::shared_ptr<ServeReq>& r) {
    const bool host_in = r->in.mem == PK_MEM_HOST;
    ctx->chunk_place(r->c, r->mem.as<char>(), host_in ? h2d_stream_of(ctx) : nullptr);
    r->t_adm = Clock::now();
    drop_psfs(*r);
    if (host_in) {
        hipEvent_t ev = S.take_event();
        PK_HIP(hipEventRecord(ev, ctx->h2d_stream));
        S.staging.emplace_back(r, ev);
    } else {
        S.active.push_back(r);
    }
}

static void scheduler(pk_ctx_impl* ctx) {
    Server& S = *ctx->srv;
    for (;;) {
        std::vector<std::shared_ptr<ServeReq>> cand;
        {
            std::unique_lock<std::mutex> lk(S.mq);
            S.cv.wait(lk, [&] {
                return S.stop || !S.queue.empty() || !S.active.empty() || !S.draining.empty() || !S.staging.empty();
            });
            if (S.stop && S.active.empty() && S.draining.empty() && S.staging.empty()) {
                for (auto& r : S.queue) {
                    r->state = PK_REQ_CANCELLED;
                    r->st = PK_ERR_CANCELLED;
                }
                S.queue.clear();
                for (hipEvent_t ev : S.ev_free) (void)hipEventDestroy(ev);
                S.ev_free.clear();
                S.cv_done.notify_all();
                return;
            }
            if (!S.stop)
                while (!S.queue.empty() && (int)(S.active.size() + cand.size()) < S.max_active) {
                    cand.push_back(S.queue.front());
                    S.queue.pop_front();
                }
        }
        // outcomes are applied to the shared request records under S.mq below
        std::vector<std::shared_ptr<ServeReq>> admitted, deferred, finished;
        std::vector<std::pair<std::shared_ptr<ServeReq>, Error>> failed;
        bool reject_head = false;
        {
            std::lock_guard<std::recursive_mutex> g(ctx->mu);
            try {
                PK_HIP(hipSetDevice(ctx->device));
            } catch (...) {
            }
            // ---- admission by HBM budget (FIFO: a request that does not fit waits, and so do the ones behind it)
            for (auto& r : cand) {
                if (!deferred.empty()) {
                    deferred.push_back(r);
                    continue;
                }
                try {
                    ctx->chunk_validate(r->in, r->out, r->c.o);
                    const size_t need = ctx->chunk_prepare(r->c, r->in, r->out, r->pf.data(), r->fd, r->hf,
                                                           r->pb.empty() ? nullptr : r->pb.data(), r->bd, r->hb);
                    r->pinned = true;
                    ctx->alloc_pooled(r->mem, need);
                    place(ctx, S, r);
                    admitted.push_back(r);
                } catch (const Error& e) {
                    unpin(ctx, *r);
                    if (r->mem.p) quiesce(ctx);
                    r->mem.release();
                    if (e.st == PK_ERR_BUDGET) {
                        // cached OTFs nobody reads are the first memory to give back; then wait
                        const size_t before = ctx->otf.size();
                        ctx->otf.erase(std::remove_if(ctx->otf.begin(), ctx->otf.end(),
                                                      [](const OtfEntry& x) { return x.pins == 0; }),
                                       ctx->otf.end());
                        bool ok = false;
                        if (ctx->otf.size() != before) {
                            try {
                                const size_t need = ctx->chunk_prepare(r->c, r->in, r->out, r->pf.data(), r->fd, r->hf,
                                                                       r->pb.empty() ? nullptr : r->pb.data(), r->bd,
                                                                       r->hb);
                                r->pinned = true;
                                ctx->alloc_pooled(r->mem, need);
                                place(ctx, S, r);
                                admitted.push_back(r);
                                ok = true;
                            } catch (const Error&) {
                                unpin(ctx, *r);
                                if (r->mem.p) quiesce(ctx);
                                r->mem.release();
                            }
                        }
                        if (!ok) deferred.push_back(r);  // memory frees up as requests finish
                    } else {
                        failed.emplace_back(r, e);
                    }
                }
            }
            if (!deferred.empty() && S.active.empty()) {
                // nothing in flight can free memory: drop unpinned cached OTFs and retry once, then reject
                size_t before = ctx->otf.size();
                ctx->otf.erase(std::remove_if(ctx->otf.begin(), ctx->otf.end(),
                                              [](const OtfEntry& e) { return e.pins == 0; }),
                               ctx->otf.end());
                if (ctx->otf.size() == before) reject_head = true;
            }
            // ---- staged requests whose input copy completed join the steps; with
            // nothing else to run, wait for the oldest copy
            if (!S.staging.empty()) {
                std::vector<std::pair<std::shared_ptr<ServeReq>, hipEvent_t>> still;
                for (size_t i = 0; i < S.staging.size(); ++i) {
                    auto& d = S.staging[i];
                    hipError_t q = hipEventQuery(d.second);
                    if (q == hipErrorNotReady && i == 0 && S.active.empty()) q = hipEventSynchronize(d.second);
                    if (q == hipErrorNotReady) {
                        still.push_back(d);
                        continue;
                    }
                    if (q == hipSuccess) {
                        S.active.push_back(d.first);
                    } else {
                        (void)hipGetLastError();
                        failed.emplace_back(d.first, Error{PK_ERR_HIP, std::string("device error: ") +
                                                                           hipGetErrorString(q)});
                        d.first->mem.release();
                        unpin(ctx, *d.first);
                    }
                    S.ev_free.push_back(d.second);
                }
                S.staging.swap(still);
            }
            // ---- one iteration for every request in flight.  Nothing waits for
            // the device here: the step-table double buffer keeps the host at
            // most two steps ahead, and finished requests retire by event.
            if (S.inject_id) {
                // a request whose host-side work fails mid-flight: it alone fails, its
                // memory and OTF pins are released once its queued steps have run
                std::vector<std::shared_ptr<ServeReq>> keep;
                for (auto& r : S.active) {
                    if (r->id == S.inject_id && r->c.k == S.inject_k) {
                        quiesce(ctx);
                        failed.emplace_back(r, Error{PK_ERR_INVALID_ARG, "injected failure (PK_SERVE_INJECT_FAILURE)"});
                        r->mem.release();
                        unpin(ctx, *r);
                    } else {
                        keep.push_back(r);
                    }
                }
                S.active.swap(keep);
            }
            if (!S.active.empty()) {
                std::vector<std::shared_ptr<ServeReq>> ending;
                try {
                    std::vector<Chunk*> act;
                    for (auto& r : S.active) act.push_back(&r->c);
                    ctx->run_step(act);
                    for (auto& r : S.active)
                        if (r->c.done()) {
                            // host outputs are copied on the copy stream, after the request's last step,
                            // so the next step does not wait for them
                            const bool host_out = r->c.out_host != nullptr;
                            if (host_out && !ctx->copy_stream)
                                PK_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
                            ctx->chunk_finish(r->c, host_out ? ctx->copy_stream : nullptr);
                            ending.push_back(r);
                        }
                    if (!ending.empty()) {
                        hipEvent_t ev = S.take_event();
                        if (ctx->copy_stream) {  // retire after both the steps and the copies
                            ctx->copy_after_stream(ctx->copy_stream);
                            PK_HIP(hipEventRecord(ev, ctx->copy_stream));
                        } else {
                            PK_HIP(hipEventRecord(ev, ctx->stream));
                        }
                        for (auto& r : ending) S.draining.emplace_back(r, ev);
                    }
                    if (ctx->prof_on) {
                        PK_HIP(hipStreamSynchronize(ctx->stream));
                        ctx->prof_flush();
                    }
                } catch (const Error& e) {
                    ending.clear();  // a device error fails every request in flight
                    for (auto& r : S.active) failed.emplace_back(r, e);
                    quiesce(ctx);
                }
                std::vector<std::shared_ptr<ServeReq>> keep;
                for (auto& r : S.active) {
                    bool failed_r = false;
                    for (auto& f : failed) failed_r |= f.first == r;
                    if (failed_r) {
                        r->mem.release();
                        unpin(ctx, *r);
                    } else if (std::find(ending.begin(), ending.end(), r) == ending.end()) {
                        keep.push_back(r);
                    }
                }
                S.active.swap(keep);
            }
            // ---- retire requests whose last step and output copy completed;
            // with nothing else to run, wait for the oldest
            if (!S.draining.empty()) {
                const bool idle = S.active.empty() && S.staging.empty();
                std::vector<std::pair<std::shared_ptr<ServeReq>, hipEvent_t>> still;
                std::vector<hipEvent_t> seen;
                for (size_t i = 0; i < S.draining.size(); ++i) {
                    auto& d = S.draining[i];
                    hipError_t q = hipEventQuery(d.second);
                    if (q == hipErrorNotReady && idle && i == 0) q = hipEventSynchronize(d.second);
                    if (q == hipErrorNotReady) {
                        still.push_back(d);
                        continue;
                    }
                    if (q == hipSuccess) {
                        finished.push_back(d.first);
                    } else {
                        (void)hipGetLastError();
                        failed.emplace_back(d.first, Error{PK_ERR_HIP, std::string("device error: ") +
                                                                           hipGetErrorString(q)});
                    }
                    d.first->mem.release();
                    unpin(ctx, *d.first);
                    if (std::find(seen.begin(), seen.end(), d.second) == seen.end()) seen.push_back(d.second);
                }
                S.draining.swap(still);
                for (hipEvent_t ev : seen) {  // an event shared by requests still draining stays out of the pool
                    bool used = false;
                    for (auto& d : S.draining) used |= d.second == ev;
                    if (!used) S.ev_free.push_back(ev);
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(S.mq);
            const auto now = Clock::now();
            for (auto& r : admitted) r->state = PK_REQ_RUNNING;
            for (auto& r : S.active) r->iters = r->c.k;
            for (auto& r : finished) {
                r->iters = r->c.k;
                r->t_done = now;
                r->state = PK_REQ_DONE;
                S.stats.completed++;
                S.stats.voxels += (double)(r->c.d[0] * r->c.d[1] * r->c.d[2]);
            }
            for (auto& f : failed) {
                f.first->state = PK_REQ_FAILED;
                f.first->st = f.second.st;
                f.first->err = f.second.msg;
                f.first->t_done = now;
                S.stats.failed++;
            }
            for (auto it = deferred.rbegin(); it != deferred.rend(); ++it) S.queue.push_front(*it);
            if (reject_head && !S.queue.empty()) {
                auto r = S.queue.front();
                S.queue.pop_front();
                r->state = PK_REQ_FAILED;
                r->st = PK_ERR_BUDGET;
                r->err = "request exceeds the HBM budget";
                r->t_done = now;
                S.stats.failed++;
            }
            S.stats.active = (int32_t)S.active.size();
            S.stats.queued = (int32_t)S.queue.size();
        }
        S.cv_done.notify_all();
    }
}

void serve_start(pk_ctx_impl* ctx, int max_active) {
    if (ctx->srv) fail(PK_ERR_BUSY, "server already running");
    ctx->srv = std::make_unique<Server>();
    ctx->srv->max_active = max_active > 0 ? max_active : 32;
    if (const char* inj = std::getenv("PK_SERVE_INJECT_FAILURE")) {
        unsigned long long id = 0;
        int k = 0;
        if (std::sscanf(inj, "%llu:%d", &id, &k) == 2 && k >= 0) {
            ctx->srv->inject_id = id;
            ctx->srv->inject_k = k;
        }
    }
    // many request shapes and channels: keep their OTFs (the HBM budget still bounds them)
    ctx->otf_max_entries = std::max<size_t>(ctx->otf_max_entries, 256);
    ctx->srv->th = std::thread(scheduler, ctx);
}

void serve_stop(pk_ctx_impl* ctx) {
    if (!ctx->srv) return;
    {
        std::lock_guard<std::mutex> lk(ctx->srv->mq);
        ctx->srv->stop = true;
    }
    ctx->srv->cv.notify_all();
    if (ctx->srv->th.joinable()) ctx->srv->th.join();
    ctx->srv.reset();
}

uint64_t serve_submit(pk_ctx_impl* ctx, const pk_request& q) {
    if (!ctx->srv) fail(PK_ERR_INVALID_ARG, "server not started");
    if (q.method != 0 && q.method != 1) fail(PK_ERR_INVALID_ARG, "method must be 0 (RL) or 1 (OMW)");
    if (!q.psf_f || (q.method == 1 && !q.psf_b) || q.numit < 0) fail(PK_ERR_INVALID_ARG, "missing PSF or NUMIT < 0");
    auto r = std::make_shared<ServeReq>();
    r->c.method = q.method;
    r->c.numit = q.numit;
    r->c.o = q.opts;
    r->in = q.in;
    r->out = q.out;
    const int64_t nf = q.f_dims[0] * q.f_dims[1] * q.f_dims[2];
    if (nf < 1) fail(PK_ERR_INVALID_ARG, "empty PSF");
    r->pf.assign(q.psf_f, q.psf_f + nf);
    std::memcpy(r->fd, q.f_dims, sizeof(r->fd));
    if (q.method == 1) {
        const int64_t nb = q.b_dims[0] * q.b_dims[1] * q.b_dims[2];
        if (nb < 1) fail(PK_ERR_INVALID_ARG, "empty PSF_b");
        r->pb.assign(q.psf_b, q.psf_b + nb);
        std::memcpy(r->bd, q.b_dims, sizeof(r->bd));
        r->hb = psf_key(r->pb.data(), r->bd);
    }
    r->hf = psf_key(r->pf.data(), r->fd);
    r->t_sub = Clock::now();
    Server& S = *ctx->srv;
    {
        std::lock_guard<std::mutex> lk(S.mq);
        if (S.stop) fail(PK_ERR_CANCELLED, "server stopping");
        r->id = S.next_id++;
        S.reqs[r->id] = r;
        S.queue.push_back(r);
        S.stats.submitted++;
        S.stats.queued = (int32_t)S.queue.size();
    }
    S.cv.notify_all();
    return r->id;
}

static std::shared_ptr<ServeReq> find(Server& S, uint64_t id) {
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) fail(PK_ERR_INVALID_ARG, "unknown request id");
    return it->second;
}

static void report(const ServeReq& r, pk_req_info* info) {
    if (!info) return;
    info->state = r.state;
    info->status = r.st;
    info->iterations_done = r.iters;
    const auto end = r.state == PK_REQ_DONE ? r.t_done : Clock::now();
    info->latency_ms = std::chrono::duration<double, std::milli>(end - r.t_sub).count();
    info->queue_ms = r.state == PK_REQ_QUEUED ? info->latency_ms
                                             : std::chrono::duration<double, std::milli>(r.t_adm - r.t_sub).count();
}

void serve_poll(pk_ctx_impl* ctx, uint64_t id, pk_req_info* info, double wait_ms, std::string* err) {
    if (!ctx->srv) fail(PK_ERR_INVALID_ARG, "server not started");
    Server& S = *ctx->srv;
    std::unique_lock<std::mutex> lk(S.mq);
    auto r = find(S, id);
    if (wait_ms != 0) {
        auto pred = [&] { return r->state >= PK_REQ_DONE; };
        if (wait_ms < 0) S.cv_done.wait(lk, pred);
        else S.cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms), pred);
    }
    report(*r, info);
    if (err) *err = r->err;
}

void serve_cancel(pk_ctx_impl* ctx, uint64_t id) {
    if (!ctx->srv) fail(PK_ERR_INVALID_ARG, "server not started");
    Server& S = *ctx->srv;
    std::lock_guard<std::mutex> lk(S.mq);
    auto r = find(S, id);
    for (auto it = S.queue.begin(); it != S.queue.end(); ++it)
        if ((*it)->id == id) {
            S.queue.erase(it);
            r->state = PK_REQ_CANCELLED;
            r->st = PK_ERR_CANCELLED;
            S.cv_done.notify_all();
            return;
        }
    if (r->state == PK_REQ_RUNNING) fail(PK_ERR_BUSY, "request already running");
}

void serve_release(pk_ctx_impl* ctx, uint64_t id) {
    if (!ctx->srv) fail(PK_ERR_INVALID_ARG, "server not started");
    Server& S = *ctx->srv;
    std::lock_guard<std::mutex> lk(S.mq);
    auto r = find(S, id);
    if (r->state < PK_REQ_DONE) fail(PK_ERR_BUSY, "request not finished");
    S.reqs.erase(id);
}

pk_serve_stats serve_stats(pk_ctx_impl* ctx) {
    if (!ctx->srv) fail(PK_ERR_INVALID_ARG, "server not started");
    std::lock_guard<std::mutex> lk(ctx->srv->mq);
    return ctx->srv->stats;
}

pk_ctx_impl::pk_ctx_impl() = default;
pk_ctx_impl::~pk_ctx_impl() { serve_stop(this); }

}  // namespace pk
