// pairs_api.cpp -- pair counts (the two-point function's DD / DR / RR) over
// libtpcf catalogs: tpcf_count_pairs and tpcf_count_pairs_batch.
//
// The reference's equivalent is main() of src/2pcf_C_version.cc:124-411:
// read two catalogs, pick the inner loop's start from the file names
// (:359-369), run the pair loop, write 700 bins.  Here a job names segments
// of device catalogs and 1..4 terms over them; a term (a, a) counts the
// C(n_a, 2) unique pairs of segment a, a term (a, b) the n_a * n_b pairs of
// the product, and every job of a call runs in one persistent launch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "tpcf_host.h"

namespace tpcf {

PairWorkspace::~PairWorkspace() {
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
}

namespace {

constexpr int kPairWaves = 8;          // FAST kernel: waves per workgroup
constexpr int kPlainBlock = 256;       // PLAIN kernel: i points per item
constexpr int kPlainChunk = 4096;      // PLAIN kernel: j points per item
constexpr double kMaxPairHi = 1e100;   // far-point sentinel of the kernel stays beyond every edge
constexpr int kMaxSegPoints = 0x7FFFFFFF - 65536;  // 32-bit point indices in the kernels

int check_pair_job(const tpcf_job &job) {
    if (job.nseg < 1 || job.nseg > 3) return set_error(TPCF_EINVAL, "a job has 1..3 segments, got %d", job.nseg);
    if (job.nterms < 1 || job.nterms > 4) return set_error(TPCF_EINVAL, "a job has 1..4 terms, got %d", job.nterms);
    for (int s = 0; s < job.nseg; ++s) {
        const tpcf_catalog *c = job.seg[s];
        if (!c) return set_error(TPCF_EINVAL, "segment %d has no catalog", s);
        if (job.seg_lo[s] < 0 || job.seg_hi[s] < job.seg_lo[s] || job.seg_hi[s] > c->n)
            return set_error(TPCF_EINVAL, "segment %d range [%lld, %lld) outside catalog of %lld points", s,
                             (long long)job.seg_lo[s], (long long)job.seg_hi[s], (long long)c->n);
        if (job.seg_hi[s] - job.seg_lo[s] > kMaxSegPoints)
            return set_error(TPCF_EINVAL, "segment %d holds more than %d points", s, kMaxSegPoints);
        if (c->ctx != job.seg[0]->ctx) return set_error(TPCF_EINVAL, "segments belong to different contexts");
    }
    for (int t = 0; t < job.nterms; ++t) {
        const int a = job.term[t][0], b = job.term[t][1];
        if (a < 0 || a > b || b >= job.nseg)
            return set_error(TPCF_EINVAL, "pair term %d (%d,%d) must be sorted segment labels < %d", t, a, b,
                             job.nseg);
    }
    return TPCF_OK;
}

// Table D (n+3 doubles: -inf, squared thresholds T[0..n], +inf) and the
// FP32 bin-guess coefficients of a distance axis.
int pair_axis(const tpcf_axis &ax, std::vector<double> &D, float &mul, float &add) {
    std::vector<double> e(ax.n + 1);
    if (axis_edges(ax, e.data(), kMaxPairBins)) return TPCF_EINVAL;
    if (ax.lo < 0) return set_error(TPCF_EINVAL, "distance axis needs lo >= 0");
    if (ax.hi > kMaxPairHi) return set_error(TPCF_EINVAL, "distance axis needs hi <= %g", kMaxPairHi);
    const int n = ax.n;
    D.assign(n + 3, 0.0);
    sq_thresholds(e.data(), n, D.data() + 1);
    D[0] = -HUGE_VAL;
    D[n + 2] = HUGE_VAL;
    if (ax.scale == TPCF_LOG) {
        const double c = n / std::log2(ax.hi / ax.lo);
        mul = (float)(0.5 * c);
        add = (float)(1.0 - std::log2(ax.lo) * c);
    } else {
        const double w = (ax.hi - ax.lo) / n;
        mul = (float)(1.0 / w);
        add = (float)(1.0 - ax.lo / w);
    }
    return TPCF_OK;
}

uint64_t term_pairs(int64_t na, int64_t nb, bool same) {
    return same ? (na < 2 ? 0 : (uint64_t)na * (uint64_t)(na - 1) / 2) : (uint64_t)na * (uint64_t)nb;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct PairShape {
    int copies, stride, hist_words, grid;
};

// FAST launch shape: two workgroups of kPairWaves waves per CU; as many
// privatized histogram copies (up to 16) as fit half the LDS next to the
// threshold table.  Copies sit stride = 1 (mod 64) words apart, so one bin
// of different copies falls in different banks.
PairShape pair_shape(const tpcf_ctx *ctx, int max_n) {
    PairShape s;
    s.stride = ((max_n + 63) / 64) * 64 + 1;
    const size_t budget = 80 * 1024 - 64;  // per workgroup, minus static LDS
    const size_t table = (size_t)(max_n + 3) * 8;
    s.copies = (int)std::max<size_t>(1, std::min<size_t>(16, (budget - table) / ((size_t)s.stride * 4)));
    s.hist_words = (s.copies * s.stride + 1) & ~1;
    const size_t lds = pair_hist_lds_bytes(s.hist_words, max_n);
    const int per_cu = lds * 2 <= 160 * 1024 - 128 ? 2 : 1;
    s.grid = ctx->num_cu * per_cu;
    return s;
}

}  // namespace

// Count the pair terms of njobs jobs (axis axes[j] for job j) in one launch
// on the context's stream; hist[j] receives job j's nterms * n counts.
static int run_pairs(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_axis *axes, int njobs, int mode,
                     uint64_t *const *hist, tpcf_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    if (mode != TPCF_MODE_FAST && mode != TPCF_MODE_PLAIN && mode != TPCF_MODE_FAST64)
        return set_error(TPCF_EINVAL, "unknown mode %d", mode);
    const bool plain = mode == TPCF_MODE_PLAIN;
    // Host tables: per job its threshold table; per (job, term) a term record.
    std::vector<std::vector<double>> tables(njobs);
    std::vector<float> hmul(njobs), hadd(njobs);
    std::vector<size_t> hist_off(njobs), table_off(njobs);
    size_t hist_words = 0, table_words = 0;
    int max_n = 1, nterms_all = 0;
    uint64_t pairs_all = 0;
    for (int j = 0; j < njobs; ++j) {
        if (!hist[j]) return set_error(TPCF_EINVAL, "histogram %d is NULL", j);
        if (int e = check_pair_job(jobs[j])) return e;
        if (int e = pair_axis(axes[j], tables[j], hmul[j], hadd[j])) return e;
        if (jobs[j].seg[0]->ctx != ctx) return set_error(TPCF_EINVAL, "job %d belongs to another context", j);
        hist_off[j] = hist_words;
        hist_words += (size_t)jobs[j].nterms * axes[j].n;
        table_off[j] = table_words;
        table_words += tables[j].size();
        max_n = std::max(max_n, axes[j].n);
        nterms_all += jobs[j].nterms;
        for (int t = 0; t < jobs[j].nterms; ++t) {
            const int a = jobs[j].term[t][0], b = jobs[j].term[t][1];
            pairs_all += term_pairs(jobs[j].seg_hi[a] - jobs[j].seg_lo[a], jobs[j].seg_hi[b] - jobs[j].seg_lo[b],
                                    a == b);
        }
    }
    const PairShape shape = pair_shape(ctx, max_n);
    const int block = plain ? kPlainBlock : kPairWaves * 64 * kPairIpl;
    // j points per item: about 16 items per workgroup of the grid, at least
    // 64 and at most 16384 points.
    int chunk = kPlainChunk;
    if (!plain) {
        const double per_item = (double)pairs_all / (16.0 * shape.grid) / block;
        chunk = (int)std::min(16384.0, std::max(64.0, std::ceil(per_item / 8.0) * 8.0));
    }

    // Layout of the device buffer.
    const size_t o_hist = 0;
    const size_t o_table = align256(o_hist + hist_words * 8);
    const size_t o_terms = align256(o_table + table_words * 8);
    const size_t o_items = align256(o_terms + (size_t)nterms_all * sizeof(PairHistTerm));
    std::vector<PairHistTerm> terms;
    std::vector<PairHistItem> items;
    terms.reserve(nterms_all);
    PairWorkspace &W = ctx->pairs;
    // Items are built after the buffer is sized; term pointers need its base,
    // so size with a first pass over the item counts.
    size_t nitems = 0;
    for (int j = 0; j < njobs; ++j)
        for (int t = 0; t < jobs[j].nterms; ++t) {
            const tpcf_job &J = jobs[j];
            const int a = J.term[t][0], b = J.term[t][1];
            const int64_t na = J.seg_hi[a] - J.seg_lo[a], nb = J.seg_hi[b] - J.seg_lo[b];
            for (int64_t i0 = 0; i0 < na; i0 += block) {
                const int64_t js = a == b ? i0 + 1 : 0;
                if (js < nb) nitems += (size_t)((nb - js + chunk - 1) / chunk);
            }
        }
    if (nitems > 0x7FFFFFFF) return set_error(TPCF_EINVAL, "too many work items (%zu)", nitems);
    const size_t o_counter = align256(o_items + nitems * sizeof(PairHistItem));
    const size_t bytes = o_counter + 256;
    if (int e = W.buf.reserve(bytes)) return e;
    for (auto &e : W.ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    char *base = static_cast<char *>(W.buf.p);
    auto *dhist = reinterpret_cast<unsigned long long *>(base + o_hist);
    auto *dtable = reinterpret_cast<double *>(base + o_table);
    items.reserve(nitems);
    for (int j = 0; j < njobs; ++j) {
        const tpcf_job &J = jobs[j];
        const int n = axes[j].n;
        for (int t = 0; t < J.nterms; ++t) {
            const int a = J.term[t][0], b = J.term[t][1];
            const tpcf_catalog *ca = J.seg[a], *cb = J.seg[b];
            const int64_t na = J.seg_hi[a] - J.seg_lo[a], nb = J.seg_hi[b] - J.seg_lo[b];
            PairHistTerm T{};
            T.ax = ca->d + J.seg_lo[a];
            T.ay = ca->d + ca->n + J.seg_lo[a];
            T.az = ca->d + 2 * ca->n + J.seg_lo[a];
            T.bx = cb->d + J.seg_lo[b];
            T.by = cb->d + cb->n + J.seg_lo[b];
            T.bz = cb->d + 2 * cb->n + J.seg_lo[b];
            T.D = dtable + table_off[j];
            T.hist = dhist + hist_off[j] + (size_t)t * n;
            T.na = (int)na;
            T.nb = (int)nb;
            T.n = n;
            T.same = a == b;
            T.log = axes[j].scale == TPCF_LOG;
            T.h_mul = hmul[j];
            T.h_add = hadd[j];
            const int slot = (int)terms.size();
            terms.push_back(T);
            for (int64_t i0 = 0; i0 < na; i0 += block) {
                const int64_t js = a == b ? i0 + 1 : 0;
                for (int64_t j0 = js; j0 < nb; j0 += chunk)
                    items.push_back({slot, (int)i0, (int)j0, (int)std::min<int64_t>(nb, j0 + chunk)});
            }
        }
    }

    // Upload tables, terms and items; clear histograms and the counter.
    hipStream_t st = ctx->stream;
    std::vector<char> host(o_counter - o_table);
    for (int j = 0; j < njobs; ++j)
        std::memcpy(host.data() + table_off[j] * 8, tables[j].data(), tables[j].size() * 8);
    std::memcpy(host.data() + (o_terms - o_table), terms.data(), terms.size() * sizeof(PairHistTerm));
    if (!items.empty())
        std::memcpy(host.data() + (o_items - o_table), items.data(), items.size() * sizeof(PairHistItem));
    HIP_TRY(hipMemcpyAsync(base + o_table, host.data(), host.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(base + o_hist, 0, hist_words * 8, st));
    HIP_TRY(hipMemsetAsync(base + o_counter, 0, sizeof(int), st));

    PairHistArgs A{};
    A.terms = reinterpret_cast<const PairHistTerm *>(base + o_terms);
    A.items = reinterpret_cast<const PairHistItem *>(base + o_items);
    A.nitems = (int)items.size();
    A.counter = reinterpret_cast<int *>(base + o_counter);
    A.max_n = max_n;
    A.copies = shape.copies;
    A.stride = shape.stride;
    A.hist_words = shape.hist_words;
    HIP_TRY(hipEventRecord(W.ev[0], st));
    if (A.nitems > 0) {
        if (plain) HIP_TRY(launch_pair_plain(A, st));
        else HIP_TRY(launch_pair_hist(A, kPairWaves, shape.grid, st));
    }
    HIP_TRY(hipEventRecord(W.ev[1], st));
    for (int j = 0; j < njobs; ++j)
        HIP_TRY(hipMemcpyAsync(hist[j], dhist + hist_off[j], sizeof(uint64_t) * jobs[j].nterms * axes[j].n,
                               hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, W.ev[0], W.ev[1]));
        *stats = tpcf_stats{};
        stats->pairs = pairs_all;
        stats->device_ms = ms;
        stats->pair_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return TPCF_OK;
}

}  // namespace tpcf

using namespace tpcf;

extern "C" {

int tpcf_count_pairs(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_axis *ax, int mode, uint64_t *hist,
                     tpcf_stats *stats) {
    clear_error();
    if (!ctx || !job || !ax || !hist) return set_error(TPCF_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(ctx->call_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    uint64_t *const h[1] = {hist};
    return run_pairs(ctx, job, ax, 1, mode, h, stats);
}

int tpcf_count_pairs_batch(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_axis *axes, int njobs, int mode,
                           uint64_t *const *hist, tpcf_stats *stats) {
    clear_error();
    if (!ctx || njobs < 0 || (njobs > 0 && (!jobs || !axes || !hist))) return set_error(TPCF_EINVAL, "NULL argument");
    if (njobs == 0) {
        if (stats) *stats = tpcf_stats{};
        return TPCF_OK;
    }
    std::lock_guard<std::mutex> g(ctx->call_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return run_pairs(ctx, jobs, axes, njobs, mode, hist, stats);
}

}  // extern "C"
