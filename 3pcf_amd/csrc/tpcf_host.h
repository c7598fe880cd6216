// tpcf_host.h -- host-side internals of libtpcf (not part of the C ABI).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tpcf.h"
#include "tpcf_internal.h"

namespace tpcf {

constexpr int kMaxSqtCells = 36 * 1024;  // LDS histogram limit of the (s,q,theta) kernels
// Catalog coordinates are finite and at most this large in magnitude, so
// every squared distance is finite in FP64.
constexpr double kMaxCoordinate = 1e100;

// Record a message for tpcf_last_error() and return code.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Return the HIP error of expr as TPCF_ENOMEM / TPCF_EDEVICE with a message.
#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return set_error(e_ == hipErrorOutOfMemory ? TPCF_ENOMEM : TPCF_EDEVICE, "%s: %s (%s:%d)", \
                             #expr, hipGetErrorString(e_), __FILE__, __LINE__);                         \
    } while (0)

// Edges e[0..n] of an axis of 1..max_bins bins (triplet axes: kMaxBins).
int axis_edges(const tpcf_axis &ax, double *e, int max_bins = kMaxBins);
void sq_thresholds(const double *e, int n, double *T);
int binspec_check(const tpcf_binspec &s);
int64_t hist_cells(const tpcf_binspec &s);

// Device allocation that remembers its size.  A fixed buffer is a view of
// memory it does not own (the no-allocation entry point's scratch): it never
// grows, reallocates or frees, and reserve() beyond its size is an error.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool fixed = false;
    // Set: allocations are stream-ordered on this stream (hipFreeAsync and
    // hipMallocFromPoolAsync from `pool`), so growing or releasing the
    // buffer does not wait for every kernel on the device, as hipFree does.
    hipStream_t ordered = nullptr;
    hipMemPool_t pool = nullptr;
    ~DevBuf();
    int reserve(size_t need);  // grow-only
    void release();            // free the memory (owned buffers only)
    void view(void *at, size_t n) {
        p = at;
        bytes = n;
        fixed = true;
    }
};

// Union layout of a job: each segment padded to whole 64-point blocks.
struct Layout {
    int nseg = 0;
    int64_t seg_n[3] = {0, 0, 0};
    int64_t seg_off[3] = {0, 0, 0};
    int64_t seg_blk0[3] = {0, 0, 0};
    int64_t seg_nblk[3] = {0, 0, 0};
    int64_t npad = 0, nblk = 0;
};

// Device buffer and events of the pair-count calls (pairs_api.cpp).
struct PairWorkspace {
    DevBuf buf;
    DevBuf cells;  // sorted segment copies, chunk boxes and sort scratch (TPCF_MODE_CELLS)
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PairWorkspace();
};

// Tuned launch shape of the triplet kernels.
struct LaunchShape {
    int waves = 0;     // waves per workgroup (0 = automatic)
    int copies = 1;    // privatized LDS histogram copies
    int per_cu = 0;    // workgroups per CU (0 = occupancy limit)
    // (r1,r2,r3) triplet kernel: -1 chosen per launch from the tiles' bins,
    // 0 run-merging, 1 bit-mask, 2 split; and the scales of the choice (the
    // bit-mask kernel when the tiles' mean bins <= bits_max * sqrt(their
    // fill), else the split form when the mean <= split_mean: items whose
    // (A, B) tile holds at most split_bins bins to the bit-mask kernel, the
    // rest to the run-merging kernel; split_bins < 0: no split form).
    int form = -1;
    float bits_max = 7.0f;
    int split_bins = 5;
    float split_mean = 8.0f;
};

}  // namespace tpcf

struct tpcf_catalog {
    tpcf_ctx *ctx = nullptr;
    double *d = nullptr;   // device SoA [x|y|z]
    int64_t n = 0;
    std::vector<double> h; // host copy (origin / extent computations)
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};  // per-axis bounds (n > 0)
};

namespace tpcf {
struct Ticket;
struct Workspace;
struct ElasticScheduler;

// Per-axis bounds of points [lo_i, hi_i) of a catalog (non-empty range):
// the bounds kept at creation for the whole catalog, else one pass.
inline void seg_bounds(const tpcf_catalog *c, int64_t lo_i, int64_t hi_i, double lo[3], double hi[3]) {
    if (lo_i == 0 && hi_i == c->n) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = c->lo[d];
            hi[d] = c->hi[d];
        }
        return;
    }
    for (int d = 0; d < 3; ++d) {
        const double *v = c->h.data() + (size_t)d * c->n;
        double a = v[lo_i], b = v[lo_i];
        for (int64_t i = lo_i + 1; i < hi_i; ++i) {
            a = std::min(a, v[i]);
            b = std::max(b, v[i]);
        }
        lo[d] = a;
        hi[d] = b;
    }
}

// Run body(lo, hi, worker) over [0, n) split into contiguous ranges on up to
// 16 host threads (one range per thread; small n stays on the caller).
template <class F>
void parallel_ranges(int64_t n, F body) {
    const int64_t nt = std::min<int64_t>(std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency())),
                                         std::max<int64_t>(1, n / 64));
    if (nt <= 1) {
        body((int64_t)0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t w = 0; w < nt; ++w) th.emplace_back(body, n * w / nt, n * (w + 1) / nt, (int)w);
    for (auto &t : th) t.join();
}
}  // namespace tpcf

namespace tpcf {
// A serving batch's share of the device.  claim(work) before staging (work:
// the batch's unique triplets): the running batches' persistent launches
// shrink to make room.  control() when the batch's persistent launch of
// nitems work items is about to be enqueued on st (everything it reads
// already enqueued there): fills *ctl and lowers *grid to the batch's share
// of workgroups (per_cu per CU at full occupancy); relaunch(wgs, stream,
// ctl) enqueues wgs more workgroups of the same launch, which the hook does
// on a stream of its own when the batch's share grows.  finish() once st
// has drained: no more growth, the extra workgroups are waited for and the
// CUs go to the others.
struct ElasticHook {
    using Relaunch = std::function<int(int, hipStream_t, const ElasticCtl &)>;
    virtual ~ElasticHook() = default;
    virtual void claim(double work) = 0;
    virtual int control(int per_cu, int nitems, Relaunch relaunch, hipStream_t st, ElasticCtl *ctl, int *grid) = 0;
    virtual void finish() = 0;
};
// hook: the batch's persistent launch is elastic; the batch then holds jobs
// of one family ((r1,r2,r3), or (s,q,theta) and JBJ).
// Make the workspace's buffers stream-ordered on st (a serving worker's
// stream: a worker growing its workspace then does not wait for the other
// workers' batches), allocated from pool, a pool that only st allocates
// from and frees into.  A pool shared between workers lets an allocation
// reuse a block another worker freed by making the allocating stream wait
// for that worker's stream (the runtime's internal dependencies): the new
// batch then waits for a batch whose elastic launch the scheduler has just
// shrunk to its floor for it.
void workspace_order(Workspace *ws, hipStream_t st, hipMemPool_t pool);
// A stream-ordered device pool for one serving worker: no reuse through
// inserted dependencies, memory handed back at each synchronization.
int worker_pool_create(int device, hipMemPool_t *pool);
int run_batch(tpcf_ctx *ctx, Workspace &ws, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs, int mode,
              uint64_t *const *hist, tpcf_stats *per, tpcf_stats *total, hipStream_t st,
              ElasticHook *hook = nullptr);
int run_single(tpcf_ctx *ctx, Workspace &ws, const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *hist,
               tpcf_stats *stats, hipStream_t st);
// Device workspace bytes of a job (negative: error); *triplets (if given)
// receives its unique triplet count.
int64_t plan_bytes(const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *triplets = nullptr);
Workspace *new_workspace();
void delete_workspace(Workspace *w);
// Device bytes a workspace holds, and giving them back (the workspace must
// be idle: its stream synchronized).
size_t workspace_held(const Workspace *w);
void workspace_release(Workspace *w);
// Admission footprint: the job's region plus an upper bound of its work
// queue (negative: error).
int64_t footprint_bytes(const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *triplets = nullptr);
struct WorkspaceDeleter {
    void operator()(Workspace *w) const { delete_workspace(w); }
};
}  // namespace tpcf

struct tpcf_ctx {
    int device = 0;
    int num_cu = 0;
    size_t lds_per_block = 0;
    hipStream_t stream = nullptr;
    // Catalog uploads (non-blocking: an upload does not wait for work on the
    // serving workers' streams, which synchronize with the null stream).
    hipStream_t upload = nullptr;
    std::mutex call_mu;                    // serializes synchronous calls
    std::unique_ptr<tpcf::Workspace, tpcf::WorkspaceDeleter> ws;  // workspace of synchronous calls
    // tpcf_count_noalloc: the device's fixed scratch array (its address from
    // hipGetSymbolAddress) and a workspace of views into it, events included.
    void *scratch = nullptr;
    std::unique_ptr<tpcf::Workspace, tpcf::WorkspaceDeleter> ws_fixed;
    tpcf::LaunchShape shape;
    tpcf::PairWorkspace pairs;             // workspace of tpcf_count_pairs*
    // serving
    // Device bytes of the serving workers' workspaces: held[w] is what worker
    // w holds, or has been admitted to grow to, in_use their sum, peak the
    // largest in_use so far; admission keeps in_use <= budget.
    uint64_t budget = 0, in_use = 0, peak = 0;
    std::vector<uint64_t> held;
    // Device shares of the workers' persistent launches (serve.cpp).
    std::shared_ptr<tpcf::ElasticScheduler> elastic;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> pending;
    std::map<int64_t, std::shared_ptr<tpcf::Ticket>> tickets;
    std::vector<std::thread> workers;
    int64_t next_ticket = 1;
    bool stopping = false;
};
