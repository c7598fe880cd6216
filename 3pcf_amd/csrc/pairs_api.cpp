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
#include <array>
#include <map>
#include <tuple>
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
constexpr int kCellChunk = 512;        // TPCF_MODE_CELLS: points per box in the host's item culling

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

// A segment as the kernel reads it: SoA base pointers and size.
struct SegView {
    const double *x, *y, *z;
    int n;
    int sorted;          // index of its Hilbert-ordered copy (TPCF_MODE_CELLS), or -1
    const double *jbox;  // device boxes of its kPairBoxPoints runs (TPCF_MODE_CELLS)
};

// Squared FP64 gap between two boxes, in the kernel's order of operations.
// Every pair (p in a, q in b) has |p_c - q_c| >= the gap on each axis, and
// FP64 rounding is monotone, so the kernel's squared distance of every pair
// is at least this value.
double box_gap2(const double *a, const double *b) {
    double g[3];
    for (int c = 0; c < 3; ++c) {
        const double u = b[c] - a[3 + c], v = a[c] - b[3 + c];
        g[c] = std::max(0.0, std::max(u, v));
    }
    return (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
}

// Hilbert-ordered copies of the distinct segments of the jobs (on the
// context's stream) and their chunk boxes (on the host).
int sort_segments(tpcf_ctx *ctx, const tpcf_job *jobs, int njobs, std::vector<std::array<SegView, 3>> &views,
                  std::vector<std::vector<double>> &boxes) {
    std::map<std::tuple<const tpcf_catalog *, int64_t, int64_t>, int> ids;
    std::vector<std::tuple<const tpcf_catalog *, int64_t, int64_t>> segs;
    for (int j = 0; j < njobs; ++j)
        for (int s = 0; s < jobs[j].nseg; ++s) {
            const auto key = std::make_tuple((const tpcf_catalog *)jobs[j].seg[s], jobs[j].seg_lo[s], jobs[j].seg_hi[s]);
            auto it = ids.find(key);
            if (it == ids.end()) {
                it = ids.emplace(key, (int)segs.size()).first;
                segs.push_back(key);
            }
            views[j][s].sorted = it->second;
        }
    const size_t ns = segs.size();
    std::vector<size_t> o_sorted(ns), o_box(ns);
    size_t off = 0, scratch = 0;
    for (size_t i = 0; i < ns; ++i) {
        const int n = (int)(std::get<2>(segs[i]) - std::get<1>(segs[i]));
        o_sorted[i] = off;
        off = align256(off + (size_t)n * 24);
        o_box[i] = off;
        off = align256(off + (size_t)((n + kPairBoxPoints - 1) / kPairBoxPoints) * 48);
        scratch = std::max(scratch, sort_segment_scratch_bytes(n));
    }
    const size_t o_scratch = off;
    if (int e = ctx->pairs.cells.reserve(o_scratch + scratch + 256)) return e;
    char *base = static_cast<char *>(ctx->pairs.cells.p);
    boxes.assign(ns, {});
    for (size_t i = 0; i < ns; ++i) {
        const tpcf_catalog *c = std::get<0>(segs[i]);
        const int64_t lo = std::get<1>(segs[i]);
        const int n = (int)(std::get<2>(segs[i]) - lo);
        if (n == 0) continue;
        double blo[3], bhi[3];
        seg_bounds(c, lo, lo + n, blo, bhi);
        const double span = std::max({bhi[0] - blo[0], bhi[1] - blo[1], bhi[2] - blo[2]});
        double *sorted = reinterpret_cast<double *>(base + o_sorted[i]);
        double *box = reinterpret_cast<double *>(base + o_box[i]);
        HIP_TRY(sort_segment(c->d + lo, c->d + c->n + lo, c->d + 2 * c->n + lo, n, blo, span, base + o_scratch,
                             scratch, sorted, n, kPairBoxPoints, box, ctx->stream));
        boxes[i].resize((size_t)((n + kPairBoxPoints - 1) / kPairBoxPoints) * 6);
        HIP_TRY(hipMemcpyAsync(boxes[i].data(), box, boxes[i].size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // Host boxes of kCellChunk points: unions of the device boxes.
    constexpr int per = kCellChunk / kPairBoxPoints;
    for (auto &b : boxes) {
        const size_t nf = b.size() / 6, nc = (nf + per - 1) / per;
        std::vector<double> cb(nc * 6);
        for (size_t c = 0; c < nc; ++c) {
            for (int a = 0; a < 3; ++a) {
                cb[6 * c + a] = HUGE_VAL;
                cb[6 * c + 3 + a] = -HUGE_VAL;
            }
            for (size_t f = c * per; f < std::min(nf, (c + 1) * per); ++f)
                for (int a = 0; a < 3; ++a) {
                    cb[6 * c + a] = std::min(cb[6 * c + a], b[6 * f + a]);
                    cb[6 * c + 3 + a] = std::max(cb[6 * c + 3 + a], b[6 * f + 3 + a]);
                }
        }
        b.swap(cb);
    }
    for (int j = 0; j < njobs; ++j)
        for (int s = 0; s < jobs[j].nseg; ++s) {
            SegView &v = views[j][s];
            const int i = v.sorted;
            v.n = (int)(std::get<2>(segs[i]) - std::get<1>(segs[i]));
            const double *p = reinterpret_cast<const double *>(base + o_sorted[i]);
            v.jbox = reinterpret_cast<const double *>(base + o_box[i]);
            v.x = p;
            v.y = p + v.n;
            v.z = p + 2 * (size_t)v.n;
        }
    return TPCF_OK;
}

// Items of one term: i blocks of `block` points of segment a against runs
// of j of segment b (at most `chunk` long).  With boxes (TPCF_MODE_CELLS),
// runs skip the chunks whose box gap to the i block's box reaches T[n].
void term_pair_items(int slot, int na, int nb, bool same, int block, int chunk, const std::vector<double> *abox,
                     const std::vector<double> *bbox, double Tn, std::vector<PairHistItem> &out) {
    // i blocks split over host threads, each thread's items appended in
    // block order.
    const int64_t nblocks = ((int64_t)na + block - 1) / block;
    std::vector<std::vector<PairHistItem>> part(16);
    parallel_ranges(nblocks, [&](int64_t k0, int64_t k1, int worker) {
        std::vector<PairHistItem> &mine = part[worker];
        for (int64_t i0 = k0 * block; i0 < std::min<int64_t>(na, k1 * block); i0 += block) {
            const int64_t js = same ? i0 + 1 : 0;
            if (js >= nb) continue;
            if (!abox) {
                for (int64_t j0 = js; j0 < nb; j0 += chunk)
                    mine.push_back({slot, (int)i0, (int)j0, (int)std::min<int64_t>(nb, j0 + chunk)});
                continue;
            }
            double ib[6] = {HUGE_VAL, HUGE_VAL, HUGE_VAL, -HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
            for (int64_t c = i0 / kCellChunk; c * kCellChunk < std::min<int64_t>(na, i0 + block); ++c)
                for (int a = 0; a < 3; ++a) {
                    ib[a] = std::min(ib[a], (*abox)[6 * c + a]);
                    ib[3 + a] = std::max(ib[3 + a], (*abox)[6 * c + 3 + a]);
                }
            int64_t r0 = -1, r1 = -1;
            for (int64_t c = js / kCellChunk; c * kCellChunk < nb; ++c) {
                const int64_t j0 = std::max<int64_t>(js, c * kCellChunk), j1 = std::min<int64_t>(nb, (c + 1) * kCellChunk);
                if (box_gap2(ib, bbox->data() + 6 * c) < Tn) {
                    if (r0 < 0) r0 = j0;
                    r1 = j1;
                    if (r1 - r0 >= chunk) {
                        mine.push_back({slot, (int)i0, (int)r0, (int)r1});
                        r0 = -1;
                    }
                } else if (r0 >= 0) {
                    mine.push_back({slot, (int)i0, (int)r0, (int)r1});
                    r0 = -1;
                }
            }
            if (r0 >= 0) mine.push_back({slot, (int)i0, (int)r0, (int)r1});
        }
    });
    for (auto &p : part) out.insert(out.end(), p.begin(), p.end());
}

}  // namespace

// Count the pair terms of njobs jobs (axis axes[j] for job j) in one launch
// on the context's stream; hist[j] receives job j's nterms * n counts.
static int run_pairs(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_axis *axes, int njobs, int mode,
                     uint64_t *const *hist, tpcf_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    if (mode != TPCF_MODE_FAST && mode != TPCF_MODE_PLAIN && mode != TPCF_MODE_FAST64 && mode != TPCF_MODE_CELLS)
        return set_error(TPCF_EINVAL, "unknown mode %d", mode);
    const bool plain = mode == TPCF_MODE_PLAIN, cells = mode == TPCF_MODE_CELLS;
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
        // Culled runs carry a fraction of the pairs: keep them short enough
        // to balance the persistent workgroups.
        if (cells) chunk = std::min(chunk, 2048);
    }

    // Segments as the kernel reads them: Hilbert-ordered copies with chunk
    // boxes (FAST and CELLS: a wave's pairs are then mostly all in or all out
    // of range), or the catalogs (PLAIN).
    std::vector<std::array<SegView, 3>> views(njobs);
    std::vector<std::vector<double>> boxes;
    if (!plain) {
        if (int e = sort_segments(ctx, jobs, njobs, views, boxes)) return e;
    } else {
        for (int j = 0; j < njobs; ++j)
            for (int s = 0; s < jobs[j].nseg; ++s) {
                const tpcf_catalog *c = jobs[j].seg[s];
                const int64_t lo = jobs[j].seg_lo[s];
                views[j][s] = {c->d + lo, c->d + c->n + lo, c->d + 2 * c->n + lo,
                               (int)(jobs[j].seg_hi[s] - lo), -1, nullptr};
            }
    }
    // Terms and items.
    std::vector<PairHistTerm> terms;
    std::vector<PairHistItem> items;
    terms.reserve(nterms_all);
    for (int j = 0; j < njobs; ++j) {
        const tpcf_job &J = jobs[j];
        const int n = axes[j].n;
        for (int t = 0; t < J.nterms; ++t) {
            const int a = J.term[t][0], b = J.term[t][1];
            const SegView &A = views[j][a], &B = views[j][b];
            PairHistTerm T{};
            T.ax = A.x;
            T.ay = A.y;
            T.az = A.z;
            T.bx = B.x;
            T.by = B.y;
            T.bz = B.z;
            T.na = A.n;
            T.nb = B.n;
            T.n = n;
            T.same = a == b;
            T.log = axes[j].scale == TPCF_LOG;
            T.h_mul = hmul[j];
            T.h_add = hadd[j];
            T.jbox = B.jbox;
            const int slot = (int)terms.size();
            terms.push_back(T);
            term_pair_items(slot, A.n, B.n, a == b, block, chunk, cells ? &boxes[A.sorted] : nullptr,
                            cells ? &boxes[B.sorted] : nullptr, tables[j][n + 1], items);
        }
    }
    if (items.size() > 0x7FFFFFFF) return set_error(TPCF_EINVAL, "too many work items (%zu)", items.size());
    const size_t nitems = items.size();

    // Layout of the device buffer.
    const size_t o_hist = 0;
    const size_t o_table = align256(o_hist + hist_words * 8);
    const size_t o_terms = align256(o_table + table_words * 8);
    const size_t o_items = align256(o_terms + (size_t)nterms_all * sizeof(PairHistTerm));
    const size_t o_counter = align256(o_items + nitems * sizeof(PairHistItem));
    const size_t bytes = o_counter + 256;
    PairWorkspace &W = ctx->pairs;
    if (int e = W.buf.reserve(bytes)) return e;
    for (auto &e : W.ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    char *base = static_cast<char *>(W.buf.p);
    auto *dhist = reinterpret_cast<unsigned long long *>(base + o_hist);
    auto *dtable = reinterpret_cast<double *>(base + o_table);
    for (size_t t = 0, k = 0; t < (size_t)njobs; ++t)
        for (int u = 0; u < jobs[t].nterms; ++u, ++k) {
            terms[k].D = dtable + table_off[t];
            terms[k].hist = dhist + hist_off[t] + (size_t)u * axes[t].n;
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
        else HIP_TRY(launch_pair_hist(A, kPairWaves, cells, shape.grid, st));
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
