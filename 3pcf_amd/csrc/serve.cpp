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
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tpcf_host.h"

namespace tpcf {

// TPCF_SERVE_LOG=1: one stderr line per scheduling event of the serving
// workers (batch taken, share claimed, launch grid, share changes, waits
// for extra launches, batch done), with seconds since the first line.
static bool serve_log_on() {
    static const bool on = [] {
        const char *e = std::getenv("TPCF_SERVE_LOG");
        return e && *e && std::strcmp(e, "0") != 0;
    }();
    return on;
}
__attribute__((format(printf, 1, 2))) static void serve_log(const char *fmt, ...) {
    if (!serve_log_on()) return;
    static const auto t0 = std::chrono::steady_clock::now();
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "[serve %9.3f] %s\n",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), buf);
    std::fflush(stderr);
}

struct Ticket {
    tpcf_job job{};
    tpcf_binspec spec{};
    int mode = TPCF_MODE_FAST;
    uint64_t *hist = nullptr;
    int64_t footprint = 0;
    uint64_t work = 0;  // unique triplets
    bool r123 = false;  // batches hold (r1,r2,r3) jobs or triangle-parametrization jobs
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int rc = TPCF_OK;
    std::string err;
    tpcf_stats stats{};
    std::chrono::steady_clock::time_point t_submit, t_admit;
};

// Jobs that do not join a batched persistent launch.
static bool solo(const Ticket &t) { return t.mode == TPCF_MODE_PLAIN; }

static constexpr int kWorkers = 3;

// A stream with a hardware queue of its own.  HIP maps ordinary streams
// onto a small pool of hardware queues (four by default) shared round robin,
// and kernels of streams that share a queue run one after the other; a
// stream with a CU mask gets a queue of its own, so the serving streams use
// a mask of every CU.  Such streams synchronize with the null stream.
// Falls back to an ordinary non-blocking stream.
static hipError_t queue_stream(int num_cu, hipStream_t *st) {
    std::vector<uint32_t> mask((num_cu + 31) / 32, 0u);
    for (int c = 0; c < num_cu; ++c) mask[c / 32] |= 1u << (c % 32);
    if (hipExtStreamCreateWithCUMask(st, (uint32_t)mask.size(), mask.data()) == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

// Device shares of the workers' persistent launches.  A persistent launch
// fills every CU, so a second worker's kernels would wait for the first
// launch to end; instead each batch holds a share of the CUs.  Per worker
// slot the launch's workgroups read their desired number from pinned host
// memory before every claim and count themselves in device memory
// (ElasticCtl): lowering the desired number makes the surplus workgroups
// leave at their next claim; raising it comes with more workgroups of the
// same launch over the same work queue, enqueued on an idle helper stream
// (one of a small pool of streams with hardware queues of their own).
//
// Shares, shortest work first: every claimed batch keeps a floor (1/16 of
// an equal split, so a long batch keeps progressing), then the batches in
// order of their work (unique triplets) take what they can use of the rest:
// the batch with the least work gets the device and finishes first, and two
// large batches mostly run one after the other instead of dividing the CUs
// and the L2s between them.  A launch can use no more CUs than its items
// need; a batch still staging (host planning, uploads, pair stages) asks for
// an eighth of the device.
struct ElasticScheduler {
    struct Slot {
        bool claimed = false;   // a batch is staging or running
        bool launched = false;  // its persistent launch is enqueued
        double work = 1.0;      // unique triplets of the batch
        int per_cu = 1, nitems = 0, share = 0;
        ElasticHook::Relaunch relaunch;
        hipEvent_t ready = nullptr;   // the launch's inputs are in place
        std::vector<hipEvent_t> ext;  // after each extra launch (reused)
        size_t next_ext = 0;
    };
    static constexpr int kMaxHelpers = 8;
    int num_cu = 0;
    int *desired = nullptr;  // pinned host memory, one per slot
    int *active = nullptr;   // device memory, one per slot
    bool ok = false;         // false: set-up failed, launches keep their full grids
    Slot slots[kWorkers];
    std::vector<hipStream_t> helpers;
    std::mutex mu;

    explicit ElasticScheduler(const tpcf_ctx *ctx) : num_cu(ctx->num_cu) {
        if (hipSetDevice(ctx->device) != hipSuccess) return;
        if (hipHostMalloc(reinterpret_cast<void **>(&desired), sizeof(int) * kWorkers, hipHostMallocCoherent) !=
                hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&active), sizeof(int) * kWorkers) != hipSuccess) {
            (void)hipGetLastError();
            return;
        }
        for (Slot &s : slots)
            if (hipEventCreateWithFlags(&s.ready, hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                return;
            }
        ok = true;
    }
    ~ElasticScheduler() {
        for (hipStream_t h : helpers) (void)hipStreamDestroy(h);
        for (Slot &s : slots) {
            if (s.ready) (void)hipEventDestroy(s.ready);
            for (hipEvent_t e : s.ext) (void)hipEventDestroy(e);
        }
        if (active) (void)hipFree(active);
        if (desired) (void)hipHostFree(desired);
    }
    ElasticCtl ctl(int w) { return ElasticCtl{active + w, desired + w}; }
    void set_desired(int w, int v) { __atomic_store_n(desired + w, v, __ATOMIC_RELEASE); }

    // CUs of every claimed batch (under mu): floor, then by work.
    void plan(double cus[kWorkers]) const {
        int k = 0;
        for (const Slot &s : slots) k += s.claimed;
        double demand[kWorkers], left = num_cu;
        int order[kWorkers], n = 0;
        for (int w = 0; w < kWorkers; ++w) {
            const Slot &s = slots[w];
            cus[w] = 0.0;
            if (!s.claimed) continue;
            demand[w] = s.launched ? std::min<double>(num_cu, std::ceil((double)s.nitems / s.per_cu)) : num_cu / 8.0;
            cus[w] = std::min(demand[w], num_cu / (16.0 * k));
            left -= cus[w];
            order[n++] = w;
        }
        std::sort(order, order + n, [&](int a, int b) { return slots[a].work < slots[b].work; });
        for (int i = 0; i < n && left > 0.0; ++i) {
            const int w = order[i];
            const double give = std::min(demand[w] - cus[w], left);
            cus[w] += give;
            left -= give;
        }
    }
    int target(int w, const double cus[kWorkers], int cap) const {
        const Slot &s = slots[w];
        const int t = (int)(cus[w] * s.per_cu);
        return std::max(1, std::min({t, s.nitems, cap}));
    }
    // A helper stream with nothing enqueued, or a new one (null: none free).
    hipStream_t idle_helper() {
        for (hipStream_t h : helpers)
            if (hipStreamQuery(h) == hipSuccess) return h;
        (void)hipGetLastError();
        if ((int)helpers.size() >= kMaxHelpers) return nullptr;
        hipStream_t h = nullptr;
        if (queue_stream(num_cu, &h) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        helpers.push_back(h);
        return h;
    }
    // Every running launch but skip's to its share (under mu): shrinking
    // only lowers the desired number; growing raises it first, so the new
    // workgroups do not leave at once, then enqueues them after the
    // launch's inputs.
    void rebalance(int skip) {
        double cus[kWorkers];
        plan(cus);
        for (int w = 0; w < kWorkers; ++w) {
            Slot &s = slots[w];
            if (w == skip || !s.launched) continue;
            const int t = target(w, cus, 1 << 30);
            if (t < s.share) {
                serve_log("slot %d shrinks %d -> %d workgroups", w, s.share, t);
                set_desired(w, t);
                s.share = t;
                continue;
            }
            if (t == s.share) continue;
            hipStream_t h = idle_helper();
            if (!h) {
                serve_log("slot %d cannot grow %d -> %d: no idle helper stream (%zu)", w, s.share, t,
                          helpers.size());
                continue;
            }
            if (s.next_ext == s.ext.size()) {
                hipEvent_t e = nullptr;
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    (void)hipGetLastError();
                    continue;
                }
                s.ext.push_back(e);
            }
            set_desired(w, t);
            if (hipStreamWaitEvent(h, s.ready, 0) != hipSuccess || s.relaunch(t - s.share, h, ctl(w)) != TPCF_OK ||
                hipEventRecord(s.ext[s.next_ext], h) != hipSuccess) {
                // whatever was enqueued before the error runs over the same
                // queue; the owner still waits for this helper's work
                (void)hipGetLastError();
                set_desired(w, s.share);
                (void)hipStreamSynchronize(h);
                continue;
            }
            serve_log("slot %d grows %d -> %d workgroups (extra launch %zu)", w, s.share, t, s.next_ext);
            ++s.next_ext;
            s.share = t;
        }
    }
};

// One worker's view of the scheduler, handed to run_batch.
struct WorkerShare final : ElasticHook {
    ElasticScheduler &S;
    int w;
    WorkerShare(ElasticScheduler &s, int w_) : S(s), w(w_) {}
    void claim(double work) override {
        std::lock_guard<std::mutex> g(S.mu);
        S.slots[w].claimed = true;
        S.slots[w].work = std::max(work, 1e6);
        serve_log("slot %d claims (%.3g triplets)", w, work);
        S.rebalance(-1);
    }
    int control(int per_cu, int nitems, Relaunch relaunch, hipStream_t st, ElasticCtl *ctl, int *grid) override {
        std::lock_guard<std::mutex> g(S.mu);
        ElasticScheduler::Slot &s = S.slots[w];
        if (!S.ok || s.launched || nitems < 1) return TPCF_OK;  // not elastic: the full grid
        s.per_cu = std::max(1, per_cu);
        s.nitems = nitems;
        s.launched = true;
        double cus[kWorkers];
        S.plan(cus);
        const int t = S.target(w, cus, *grid);
        if (hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t *>(S.active + w), 1, 0u, st);
            e != hipSuccess || (e = hipEventRecord(s.ready, st)) != hipSuccess) {
            s.launched = false;
            return set_error(TPCF_EDEVICE, "elastic launch set-up failed: %s", hipGetErrorString(e));
        }
        S.set_desired(w, t);
        s.share = t;
        s.relaunch = std::move(relaunch);
        *ctl = S.ctl(w);
        serve_log("slot %d launches %d of %d workgroups over %d items (%d per CU)", w, t, *grid, nitems, s.per_cu);
        *grid = t;
        S.rebalance(w);  // the others, now that this launch's demand is known
        return TPCF_OK;
    }
    void finish() override {
        // The launch's own workgroups have ended, but all of them may have
        // left at a shrink with the rest of the queue to its extra
        // workgroups: the slot keeps its share, and keeps growing, until
        // every extra launch made for it has ended -- then the queue is
        // drained, since the last active workgroup of a launch never leaves.
        size_t waited = 0;
        for (;;) {
            std::vector<hipEvent_t> wait;
            {
                std::lock_guard<std::mutex> g(S.mu);
                ElasticScheduler::Slot &s = S.slots[w];
                if (waited == s.next_ext) {
                    s.claimed = s.launched = false;
                    s.share = 0;
                    s.next_ext = 0;
                    s.relaunch = nullptr;
                    serve_log("slot %d done", w);
                    S.rebalance(-1);
                    return;
                }
                wait.assign(s.ext.begin() + waited, s.ext.begin() + s.next_ext);
                waited = s.next_ext;
            }
            serve_log("slot %d waits for %zu extra launches", w, wait.size());
            for (hipEvent_t e : wait) (void)hipEventSynchronize(e);
        }
    }
};
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
    (void)queue_stream(ctx->num_cu, &st);
    std::unique_ptr<Workspace, WorkspaceDeleter> ws(new_workspace());
    // The worker's own pool (no pool: plain allocations, which wait for
    // the device at a free but never for another worker's stream alone).
    hipMemPool_t pool = nullptr;
    if (worker_pool_create(ctx->device, &pool) == TPCF_OK)
        workspace_order(ws.get(), st, pool);
    else
        clear_error();
    WorkerShare share(*ctx->elastic, w);
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
                    if (!batch.empty() && (t->mode != batch[0]->mode || t->r123 != batch[0]->r123)) continue;
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
                    // Idle while queued jobs do not fit: give the memory back
                    // (the pool hands it to the device at the synchronization,
                    // made without the lock: the stream is idle, but a
                    // caller's null-stream work would hold it).
                    serve_log("worker %d gives its workspace back: %zu queued jobs do not fit", w,
                              ctx->pending.size());
                    workspace_release(ws.get());
                    ctx->held[w] = 0;
                    account(ctx);
                    ctx->cv.notify_all();
                    if (st) {
                        lk.unlock();
                        (void)hipStreamSynchronize(st);
                        lk.lock();
                    }
                    continue;
                }
                ctx->cv.wait(lk);
            }
            if (batch.empty()) break;
        }
        {
            uint64_t work = 0;
            for (auto &t : batch) work += t->work;
            serve_log("worker %d takes %zu jobs (%s, %.3g triplets, first ticket %lld)%s", w, batch.size(),
                      batch[0]->r123 ? "r123" : "sqt/jbj", (double)work, (long long)ids[0],
                      release_first ? ", workspace released first" : "");
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
                                     hists.data(), per.data(), nullptr, st, &share);
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
            serve_log("worker %d batch done (rc %d)", w, rcs[0]);
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
    if (st) {
        (void)hipStreamSynchronize(st);  // the workspace's stream-ordered frees
    }
    // the pool before the stream its frees were ordered on
    if (pool) (void)hipMemPoolDestroy(pool);
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
    t->r123 = spec->param == TPCF_R123;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (ctx->workers.empty()) {
            ctx->held.assign(kWorkers, 0);
            ctx->elastic = std::make_shared<ElasticScheduler>(ctx);
            for (int w = 0; w < kWorkers; ++w) ctx->workers.emplace_back(worker_main, ctx, w);
        }
        t->t_submit = std::chrono::steady_clock::now();  // after the first call's set-up
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
