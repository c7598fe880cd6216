// serve.cpp -- asynchronous job interface (tpcf_submit / tpcf_poll /
// tpcf_wait): a request queue with admission by device-memory footprint and
// continuous batching.
//
// The reference has no counterpart; it runs one OS process per (term, voxel)
// job (examples/build_voxel_submission_scripts.py:9-41).  Here a context owns
// a small pool of workers, each with its own HIP stream and workspace.  An
// idle worker takes queued jobs that fit the memory left in the budget, the
// least work first (jobs waiting longer than kStarveSeconds come first, in
// arrival order, so a steady stream of small jobs cannot hold a large one
// back), up to its share of the queued work so that the other workers get
// jobs too and within a band of sizes (a batch completes with its largest
// job; a job larger than the whole budget fails with TPCF_EBUSY), and runs
// them as one batch: pair stages back to back, then one
// persistent triplet launch over all their work items (one for the
// (r1,r2,r3) jobs, one for the triangle parametrizations).  Jobs that arrive
// while a
// batch is on the GPU are picked up by another worker as soon as memory
// allows, so their kernels overlap with the running batch on the device.
// PLAIN-mode jobs (the FP64 baseline) run one per launch.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "tpcf_host.h"

namespace tpcf {

struct Ticket {
    tpcf_job job{};
    tpcf_binspec spec{};
    int mode = TPCF_MODE_FAST;
    uint64_t *hist = nullptr;
    int64_t footprint = 0;
    uint64_t work = 0;  // unique triplets
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int rc = TPCF_OK;
    std::string err;
    tpcf_stats stats{};
    std::chrono::steady_clock::time_point t_submit, t_admit;
};

// Jobs that do not join a batched persistent launch.
static bool solo(const Ticket &t) { return t.mode == TPCF_MODE_PLAIN; }

static constexpr int kWorkers = 3;
static constexpr int kMaxBatch = 64;
static constexpr double kStarveSeconds = 60.0;
// A batch holds jobs of at most kBandFactor times its first job's work (or
// kBandFloor triplets): a batch completes with its largest job, so mixing
// sizes would hold small jobs back.
static constexpr double kBandFactor = 4.0;
static constexpr double kBandFloor = 1e9;

// Sum of the workers' held bytes into in_use, and the peak (under ctx->mu).
static void account(tpcf_ctx *ctx) {
    uint64_t s = 0;
    for (uint64_t h : ctx->held) s += h;
    ctx->in_use = s;
    ctx->peak = std::max(ctx->peak, s);
}

static void worker_main(tpcf_ctx *ctx, int w) {
    (void)hipSetDevice(ctx->device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    std::unique_ptr<Workspace, WorkspaceDeleter> ws(new_workspace());
    for (;;) {
        std::vector<std::shared_ptr<Ticket>> batch;
        std::vector<int64_t> ids;
        bool release_first = false;
        {
            std::unique_lock<std::mutex> lk(ctx->mu);
            for (;;) {
                if (ctx->stopping && ctx->pending.empty()) break;
                // This worker holds `cap` bytes; a batch of total footprint R
                // runs in max(cap, R), or in R once it gives cap back.
                const uint64_t cap = ctx->held[w];
                const uint64_t others = ctx->in_use - cap;
                uint64_t need = 0;
                // This worker's share of the queued work (triplets).
                uint64_t queued = 0, taken = 0;
                for (int64_t id : ctx->pending) queued += ctx->tickets[id]->work;
                const uint64_t share = queued / kWorkers;
                // Visit order: starving jobs by arrival, then the rest by
                // work (stable: ties keep arrival order).
                const auto now = std::chrono::steady_clock::now();
                std::vector<int64_t> order(ctx->pending.begin(), ctx->pending.end());
                auto starving = [&](int64_t id) {
                    return std::chrono::duration<double>(now - ctx->tickets[id]->t_submit).count() > kStarveSeconds;
                };
                std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
                    const bool sa = starving(a), sb = starving(b);
                    if (sa != sb) return sa;
                    if (sa) return false;
                    return ctx->tickets[a]->work < ctx->tickets[b]->work;
                });
                // First fit in that order; PLAIN jobs run alone.
                for (size_t k = 0; k < order.size() && (int)batch.size() < kMaxBatch && (batch.empty() || taken < share);
                     ++k) {
                    const int64_t id = order[k];
                    auto t = ctx->tickets[id];
                    const bool plain = solo(*t);
                    if (plain && !batch.empty()) continue;
                    if (!batch.empty() && solo(*batch[0])) break;
                    if (!batch.empty() && t->mode != batch[0]->mode) continue;
                    if (!batch.empty() &&
                        (double)t->work > kBandFactor * std::max((double)batch[0]->work, kBandFloor))
                        continue;
                    if ((uint64_t)t->footprint > ctx->budget) {
                        t->state = 3;
                        t->rc = TPCF_EBUSY;
                        t->err = "job workspace exceeds the device memory budget";
                        ctx->pending.erase(std::find(ctx->pending.begin(), ctx->pending.end(), id));
                        ctx->cv.notify_all();
                        continue;
                    }
                    if (others + need + (uint64_t)t->footprint <= ctx->budget) {
                        need += t->footprint;
                        taken += t->work;
                        t->state = 1;
                        batch.push_back(t);
                        ids.push_back(id);
                        ctx->pending.erase(std::find(ctx->pending.begin(), ctx->pending.end(), id));
                        if (plain) break;
                    }
                }
                if (!batch.empty()) {
                    const auto now = std::chrono::steady_clock::now();
                    for (auto &t : batch) t->t_admit = now;
                    release_first = others + std::max(cap, need) > ctx->budget;
                    ctx->held[w] = release_first ? need : std::max(cap, need);
                    account(ctx);
                    break;
                }
                if (!ctx->pending.empty() && cap > 0) {
                    // Idle while queued jobs do not fit: give the memory back.
                    workspace_release(ws.get());
                    ctx->held[w] = 0;
                    account(ctx);
                    ctx->cv.notify_all();
                    continue;
                }
                ctx->cv.wait(lk);
            }
            if (batch.empty()) break;
        }
        if (release_first) workspace_release(ws.get());
        std::vector<int> rcs(batch.size(), TPCF_OK);
        std::vector<std::string> errs(batch.size());
        if (solo(*batch[0])) {
            rcs[0] = run_single(ctx, *ws, batch[0]->job, batch[0]->spec, batch[0]->mode, batch[0]->hist,
                                &batch[0]->stats, st);
            if (rcs[0]) errs[0] = tpcf_last_error();
        } else {
            std::vector<tpcf_job> jobs;
            std::vector<tpcf_binspec> specs;
            std::vector<uint64_t *> hists;
            std::vector<tpcf_stats> per(batch.size());
            for (auto &t : batch) {
                jobs.push_back(t->job);
                specs.push_back(t->spec);
                hists.push_back(t->hist);
            }
            const int rc = run_batch(ctx, *ws, jobs.data(), specs.data(), (int)jobs.size(), batch[0]->mode,
                                     hists.data(), per.data(), nullptr, st);
            for (size_t k = 0; k < batch.size(); ++k) {
                rcs[k] = rc;
                if (rc) errs[k] = tpcf_last_error();
                batch[k]->stats = per[k];
            }
        }
        {
            std::lock_guard<std::mutex> g(ctx->mu);
            const auto now = std::chrono::steady_clock::now();
            for (size_t k = 0; k < batch.size(); ++k) {
                batch[k]->rc = rcs[k];
                batch[k]->err = errs[k];
                batch[k]->state = rcs[k] ? 3 : 2;
                batch[k]->stats.wall_ms =
                    std::chrono::duration<double, std::milli>(now - batch[k]->t_submit).count();
                batch[k]->stats.queue_ms =
                    std::chrono::duration<double, std::milli>(batch[k]->t_admit - batch[k]->t_submit).count();
            }
            // what the workspace really holds now (at most what was admitted)
            ctx->held[w] = workspace_held(ws.get());
            account(ctx);
        }
        ctx->cv.notify_all();
    }
    ws.reset();
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->held[w] = 0;
        account(ctx);
    }
    if (st) (void)hipStreamDestroy(st);
}

}  // namespace tpcf

using namespace tpcf;

extern "C" {

int tpcf_submit(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode, uint64_t *hist,
                int64_t *ticket) {
    clear_error();
    if (!ctx || !job || !spec || !hist || !ticket) return set_error(TPCF_EINVAL, "NULL argument");
    uint64_t work = 0;
    const int64_t fp = footprint_bytes(*job, *spec, mode, &work);
    if (fp < 0) return (int)fp;
    auto t = std::make_shared<Ticket>();
    t->job = *job;
    t->spec = *spec;
    t->mode = mode;
    t->hist = hist;
    t->footprint = fp;
    t->work = work;
    t->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (ctx->workers.empty()) {
            ctx->held.assign(kWorkers, 0);
            for (int w = 0; w < kWorkers; ++w) ctx->workers.emplace_back(worker_main, ctx, w);
        }
        const int64_t id = ctx->next_ticket++;
        ctx->tickets[id] = t;
        ctx->pending.push_back(id);
        *ticket = id;
    }
    ctx->cv.notify_all();
    return TPCF_OK;
}

static int finish(tpcf_ctx *ctx, int64_t ticket, tpcf_stats *stats, bool block) {
    std::unique_lock<std::mutex> lk(ctx->mu);
    auto it = ctx->tickets.find(ticket);
    if (it == ctx->tickets.end()) return set_error(TPCF_ENOTFOUND, "unknown ticket %lld", (long long)ticket);
    auto t = it->second;
    if (block) ctx->cv.wait(lk, [&] { return t->state >= 2; });
    if (t->state < 2) return 0;
    if (stats) *stats = t->stats;
    const int rc = t->rc;
    const std::string err = t->err;
    ctx->tickets.erase(it);
    if (rc) return set_error(rc, "%s", err.c_str());
    return 1;
}

int tpcf_poll(tpcf_ctx *ctx, int64_t ticket, tpcf_stats *stats) {
    clear_error();
    if (!ctx) return set_error(TPCF_EINVAL, "ctx is NULL");
    return finish(ctx, ticket, stats, false);
}

int tpcf_wait(tpcf_ctx *ctx, int64_t ticket, tpcf_stats *stats) {
    clear_error();
    if (!ctx) return set_error(TPCF_EINVAL, "ctx is NULL");
    const int r = finish(ctx, ticket, stats, true);
    return r == 1 ? TPCF_OK : r;
}

}  // extern "C"
