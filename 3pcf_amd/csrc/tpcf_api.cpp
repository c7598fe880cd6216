// tpcf_api.cpp -- the C ABI of libtpcf: contexts, device catalogs, job
// planning and the launch sequences of the FAST (FP32 pair stage with FP64
// recheck + integer triplet kernel) and PLAIN (FP64 per-triplet) paths.
//
// The reference's equivalent is main() of each C program:
// src/3pcf_C_version_triangle_parametrization.cc:245-571 (read catalogs,
// pick loop bounds from the file names, run the triple loop, write the
// histogram) and src/3pcf_C_version.cc:147-466.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <thread>

#include "tpcf_host.h"

namespace tpcf {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

static void dev_free(void *p, hipStream_t ordered) {
    if (ordered)
        (void)hipFreeAsync(p, ordered);
    else
        (void)hipFree(p);
}

DevBuf::~DevBuf() {
    if (p && !fixed) dev_free(p, ordered);
}

void DevBuf::release() {
    if (fixed) return;
    if (p) dev_free(p, ordered);
    p = nullptr;
    bytes = 0;
}

int DevBuf::reserve(size_t need) {
    if (need <= bytes) return TPCF_OK;
    if (fixed)
        return set_error(TPCF_ENOMEM, "the job needs %zu bytes in a part of the fixed scratch that holds %zu", need,
                         bytes);
    if (p) {
        dev_free(p, ordered);
        p = nullptr;
        bytes = 0;
    }
    hipError_t e = ordered ? hipMallocFromPoolAsync(&p, need, pool, ordered) : hipMalloc(&p, need);
    if (e != hipSuccess) {
        p = nullptr;
        return set_error(TPCF_ENOMEM, "device allocation of %zu bytes failed: %s", need, hipGetErrorString(e));
    }
    bytes = need;
    return TPCF_OK;
}

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// job planning
// ---------------------------------------------------------------------------

static int check_job(const tpcf_job &job) {
    if (job.nseg < 1 || job.nseg > 3) return set_error(TPCF_EINVAL, "a job has 1..3 segments, got %d", job.nseg);
    if (job.nterms < 1 || job.nterms > 4) return set_error(TPCF_EINVAL, "a job has 1..4 terms, got %d", job.nterms);
    for (int s = 0; s < job.nseg; ++s) {
        const tpcf_catalog *c = job.seg[s];
        if (!c) return set_error(TPCF_EINVAL, "segment %d has no catalog", s);
        if (job.seg_lo[s] < 0 || job.seg_hi[s] < job.seg_lo[s] || job.seg_hi[s] > c->n)
            return set_error(TPCF_EINVAL, "segment %d range [%lld, %lld) outside catalog of %lld points", s,
                             (long long)job.seg_lo[s], (long long)job.seg_hi[s], (long long)c->n);
        if (c->ctx != job.seg[0]->ctx) return set_error(TPCF_EINVAL, "segments belong to different contexts");
    }
    for (int t = 0; t < job.nterms; ++t) {
        const int *tr = job.term[t];
        if (tr[0] < 0 || tr[2] >= job.nseg || tr[0] > tr[1] || tr[1] > tr[2])
            return set_error(TPCF_EINVAL, "term %d (%d,%d,%d) must be sorted segment labels < %d", t, tr[0], tr[1],
                             tr[2], job.nseg);
    }
    return TPCF_OK;
}

static Layout make_layout(const tpcf_job &job) {
    Layout L;
    L.nseg = job.nseg;
    int64_t off = 0;
    for (int s = 0; s < job.nseg; ++s) {
        L.seg_n[s] = job.seg_hi[s] - job.seg_lo[s];
        L.seg_off[s] = off;
        L.seg_blk0[s] = off / kBlock;
        L.seg_nblk[s] = (L.seg_n[s] + kBlock - 1) / kBlock;
        off += L.seg_nblk[s] * kBlock;
    }
    L.npad = off;
    L.nblk = off / kBlock;
    return L;
}

// Exact number of unique triplets of a term.
static uint64_t term_triplets(const Layout &L, const int *tr) {
    auto c2 = [](uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; };
    auto c3 = [](uint64_t n) { return n < 3 ? 0 : n * (n - 1) / 2 * (n - 2) / 3; };
    const uint64_t a = L.seg_n[tr[0]], b = L.seg_n[tr[1]], c = L.seg_n[tr[2]];
    if (tr[0] == tr[1] && tr[1] == tr[2]) return c3(a);
    if (tr[0] == tr[1]) return c2(a) * c;
    if (tr[1] == tr[2]) return a * c2(b);
    return a * b * c;
}

// Squared FP64 gap between the boxes (lo xyz, hi xyz) of two blocks, in the
// order of operations of the squared side: a lower bound of the squared side
// of every pair of their points (each coordinate difference is at least the
// axis gap and FP64 rounding is monotone).
static double block_gap2(const double *a, const double *b) {
    double g[3];
    for (int c = 0; c < 3; ++c) g[c] = std::max(0.0, std::max(b[c] - a[3 + c], a[c] - b[3 + c]));
    return (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
}

// Blocks within reach of each block, as bitsets of nblk bits (block X's
// words at [X * words, (X + 1) * words)): the blocks whose box gap to it is
// below cut2.
struct Reach {
    int64_t words = 0;
    std::vector<uint64_t> bits;
    // First and last word of each row with a bit set (first > last: none).
    std::vector<int32_t> first, last;
    const uint64_t *row(int64_t X) const { return bits.data() + X * words; }
    bool has(int64_t X, int64_t Y) const { return (row(X)[Y >> 6] >> (Y & 63)) & 1u; }
};

static Reach block_reach(const std::vector<double> &boxes, int64_t nblk, double cut2) {
    Reach r;
    r.words = (nblk + 63) / 64;
    r.bits.assign((size_t)(nblk * r.words), 0ull);
    r.first.assign((size_t)nblk, 0);
    r.last.assign((size_t)nblk, -1);
    const double *bx = boxes.data();
    // The union box of the blocks of each word: its gap to a box is at most
    // the gap to any of them (FP64 subtraction, squares and sums are
    // monotone), so a word whose union is out of reach has no bit set.
    std::vector<double> wbox((size_t)6 * r.words);
    for (int64_t w = 0; w < r.words; ++w) {
        double *u = wbox.data() + 6 * w;
        for (int c = 0; c < 3; ++c) {
            u[c] = HUGE_VAL;
            u[3 + c] = -HUGE_VAL;
        }
        for (int64_t Y = w * 64; Y < std::min<int64_t>(nblk, w * 64 + 64); ++Y)
            for (int c = 0; c < 3; ++c) {
                u[c] = std::min(u[c], bx[6 * Y + c]);
                u[3 + c] = std::max(u[3 + c], bx[6 * Y + 3 + c]);
            }
    }
    // Each thread fills whole rows X (every Y), so no word is shared.
    parallel_ranges(nblk, [&](int64_t x0, int64_t x1, int) {
        for (int64_t X = x0; X < x1; ++X) {
            uint64_t *row = r.bits.data() + (size_t)(X * r.words);
            for (int64_t w = 0; w < r.words; ++w) {
                if (!(block_gap2(bx + 6 * X, wbox.data() + 6 * w) < cut2)) continue;
                uint64_t m = 0;
                for (int64_t Y = w * 64; Y < std::min<int64_t>(nblk, w * 64 + 64); ++Y)
                    m |= (uint64_t)(block_gap2(bx + 6 * X, bx + 6 * Y) < cut2) << (Y & 63);
                row[w] = m;
            }
            int32_t f = (int32_t)r.words, l = -1;
            for (int64_t w = 0; w < r.words; ++w)
                if (row[w]) {
                    f = std::min<int32_t>(f, (int32_t)w);
                    l = (int32_t)w;
                }
            r.first[X] = f;
            r.last[X] = l;
        }
    });
    return r;
}

// Longest C-range of one work item, in blocks.  The persistent kernels flush
// their 32-bit LDS counters before an item could take a cell past
// 0xC0000000; an item adds at most 3 * 64^3 per C block to one cell (three
// (s,q,theta) entries per triplet), so one item alone stays far below that.
constexpr int64_t kMaxItemBlocks = 512;
static_assert(3 * kMaxItemBlocks * kBlock * kBlock * kBlock < 0xC0000000ll,
              "one work item could wrap a 32-bit LDS counter");
// Culled items: runs of reachable C blocks at most this far apart share one
// item (the pair stages mark the tiles they skip empty, and the triplet
// kernels skip blocks with an empty tile).
constexpr int64_t kMergeGapBlocks = 32;

// Items [begin, end) of v ordered by C-range length, longest first (the
// persistent kernels take them in that order, so the long ones do not
// finish last), ties in their order: a stable counting sort on the length,
// which is at most kMaxItemBlocks.
static void sort_items_longest_first(std::vector<TripletItem> &v, size_t begin) {
    constexpr int kKeys = (int)kMaxItemBlocks + 1;
    const size_t n = v.size() - begin;
    if (n < 2) return;
    std::vector<size_t> start(kKeys + 1, 0);
    for (size_t i = begin; i < v.size(); ++i) ++start[kKeys - (v[i].c1 - v[i].c0)];
    size_t acc = 0;
    for (int k = 0; k <= kKeys; ++k) {
        const size_t c = start[k];
        start[k] = acc;
        acc += c;
    }
    std::vector<TripletItem> tmp(n);
    for (size_t i = begin; i < v.size(); ++i) tmp[start[kKeys - (v[i].c1 - v[i].c0)]++] = v[i];
    std::copy(tmp.begin(), tmp.end(), v.begin() + (std::ptrdiff_t)begin);
}

// Work items of one term, appended to out longest C-range first (ties in
// (A, B, C) order): blocks A <= B <= C of the term's segments.  With the
// blocks' reach lists (TPCF_MODE_CELLS), only the block triples whose three
// pairwise box gaps are below the cut (no triplet with a squared side at or
// above it can count).
static void term_items(const Layout &L, const int *tr, int term_slot, std::vector<TripletItem> &out,
                       const Reach *reach = nullptr, int64_t merge_gap = 0) {
    const int64_t a0 = L.seg_blk0[tr[0]], a1 = a0 + L.seg_nblk[tr[0]];
    const int64_t b0 = L.seg_blk0[tr[1]], b1 = b0 + L.seg_nblk[tr[1]];
    const int64_t cb0 = L.seg_blk0[tr[2]], cb1 = cb0 + L.seg_nblk[tr[2]];
    if (reach) {
        // Blocks A of segment a split over host threads; each thread's items
        // (in A order) are counted per C-range length, then every thread
        // writes its items at their places in the term's longest-first order
        // (a stable counting sort without the concatenation).
        constexpr int kKeys = (int)kMaxItemBlocks + 1;
        const int64_t last_w = (cb1 - 1) >> 6;
        std::vector<std::vector<TripletItem>> part(16);
        std::vector<std::vector<size_t>> count(16);
        parallel_ranges(a1 - a0, [&](int64_t lo, int64_t hi, int worker) {
            std::vector<TripletItem> &mine = part[worker];
            for (int64_t A = a0 + lo; A < a0 + hi; ++A) {
                const uint64_t *ra = reach->row(A);
                const int64_t bs = tr[0] == tr[1] ? A : b0;
                const int64_t fa = reach->first[A], la = reach->last[A];
                // Each B within reach of A, from the set bits of A's row.
                for (int64_t wb = std::max<int64_t>(bs >> 6, fa); wb <= std::min<int64_t>((b1 - 1) >> 6, la); ++wb) {
                    uint64_t mb = ra[wb];
                    if (wb == bs >> 6) mb &= ~0ull << (bs & 63);
                    if (wb == (b1 - 1) >> 6 && (b1 & 63)) mb &= ~0ull >> (64 - (b1 & 63));
                    for (; mb; mb &= mb - 1) {
                        const int64_t B = wb * 64 + __builtin_ctzll(mb);
                        const uint64_t *rb = reach->row(B);
                        const int64_t cs = std::max<int64_t>((tr[1] == tr[2]) ? B : cb0, cb0);
                        if (cs >= cb1) continue;
                        // Runs of set bits of ra & rb in [cs, cb1), word by
                        // word; runs at most merge_gap blocks apart share an
                        // item (the kernel skips the blocks between them by
                        // their tiles' empty flags), so a run cut at a word
                        // boundary joins up again.
                        int64_t p0 = -1, p1 = -1;
                        auto emit = [&](int64_t r0, int64_t r1) {
                            for (int64_t q = r0; q < r1; q += kMaxItemBlocks) {
                                TripletItem it{};
                                it.term = term_slot;
                                it.A = (int)A;
                                it.B = (int)B;
                                it.c0 = (int)q;
                                it.c1 = (int)std::min<int64_t>(r1, q + kMaxItemBlocks);
                                mine.push_back(it);
                            }
                        };
                        const int64_t w1 = std::min<int64_t>({last_w, la, (int64_t)reach->last[B]});
                        for (int64_t w = std::max<int64_t>({cs >> 6, fa, (int64_t)reach->first[B]}); w <= w1; ++w) {
                            uint64_t m = ra[w] & rb[w];
                            if (w == cs >> 6) m &= ~0ull << (cs & 63);
                            if (w == last_w && (cb1 & 63)) m &= ~0ull >> (64 - (cb1 & 63));
                            while (m) {
                                const int b = __builtin_ctzll(m);
                                const uint64_t ones = ~(m >> b);
                                const int len = ones ? __builtin_ctzll(ones) : 64 - b;
                                const int64_t r0 = w * 64 + b, r1 = r0 + len;
                                m = b + len >= 64 ? 0ull : m & (~0ull << (b + len));
                                if (p1 >= 0 && r0 - p1 <= merge_gap) {
                                    p1 = r1;
                                } else {
                                    if (p1 >= 0) emit(p0, p1);
                                    p0 = r0;
                                    p1 = r1;
                                }
                            }
                        }
                        if (p1 >= 0) emit(p0, p1);
                    }
                }
            }
            std::vector<size_t> &cnt = count[worker];
            cnt.assign(kKeys, 0);
            for (const TripletItem &it : mine) ++cnt[kMaxItemBlocks - (it.c1 - it.c0)];
        });
        // Place of each (key, thread) group: keys in order, threads in A order.
        size_t acc = out.size();
        for (int k = 0; k < kKeys; ++k)
            for (auto &cnt : count)
                if (!cnt.empty()) {
                    const size_t c = cnt[k];
                    cnt[k] = acc;
                    acc += c;
                }
        const size_t total = acc - out.size();
        out.resize(acc);
        const auto place = [&](int w) {
            std::vector<size_t> &at = count[w];
            for (const TripletItem &it : part[w]) out[at[kMaxItemBlocks - (it.c1 - it.c0)]++] = it;
        };
        // A thread each only for large terms (a thread start costs about as
        // much as placing ten thousand items).
        if (total < (size_t(1) << 17)) {
            for (int w = 0; w < 16; ++w) place(w);
        } else {
            std::vector<std::thread> th;
            for (int w = 0; w < 16; ++w)
                if (!part[w].empty()) th.emplace_back(place, w);
            for (auto &t : th) t.join();
        }
        return;
    }
    const size_t before = out.size();
    for (int64_t A = a0; A < a1; ++A)
        for (int64_t B = (tr[0] == tr[1] ? A : b0); B < b1; ++B) {
            const int64_t cs = (tr[1] == tr[2]) ? B : cb0;
            for (int64_t c = cs; c < cb1; c += kMaxItemBlocks) {
                TripletItem it{};
                it.term = term_slot;
                it.A = (int)A;
                it.B = (int)B;
                it.c0 = (int)c;
                it.c1 = (int)std::min(cb1, c + kMaxItemBlocks);
                out.push_back(it);
            }
        }
    sort_items_longest_first(out, before);
}

// Everything one job needs on the device, laid out in one region.
struct JobPlan {
    Layout L;
    int param = TPCF_R123;
    int n = 0;
    int nterms = 0;
    int64_t cells = 0;
    double origin[3] = {0, 0, 0};
    float rmax = 0.f;
    float rel = 0.f;
    std::vector<double> T64;   // squared thresholds of axis 0 (distance / s)
    std::vector<float> T32;
    std::vector<double> Eq;    // SQT/JBJ: q edges
    std::vector<double> Et;    // SQT: cos(pi*edge) of theta/pi; JBJ: (r3-r2)/r1 edges
    int nq = 0, nt = 0;
    std::vector<uint8_t> blk_count;
    uint64_t triplets = 0;
    // byte offsets inside the job's device region
    size_t o_u64 = 0, o_u32 = 0, o_blk = 0, o_T64 = 0, o_T32 = 0, o_Eq = 0, o_Et = 0, o_cnt = 0, o_hist = 0,
           o_tiles = 0, o_any = 0, o_dists = 0, o_masks = 0, o_meta = 0, o_rmask = 0, o_tstat = 0;
    size_t bytes = 0;
    bool need_tiles = true;
    bool bitmask = false;  // tile summaries and row masks for the bit-mask triplet kernel
    double far_s = 0;  // SQT FAST: squared sides above this cannot be in any entry
    double mid_s = 0;  // SQT FAST: squared q sides above this cannot be in any entry
    tpcf_axis axq{}, axt{};  // SQT: q and theta/pi axes
    // TPCF_MODE_CELLS: segments staged in Hilbert order, 64-point block boxes
    // (filled by stage_job), and the squared side no counted triplet reaches.
    bool cull = false;
    bool spatial = false;  // segments staged in Hilbert order (cull, or the (s,q,theta)/JBJ FAST paths)
    double cut2 = HUGE_VAL;
    mutable std::vector<double> boxes;
    mutable Reach reach;  // per block: the blocks within the cut
    size_t o_box = 0, o_sort = 0, sort_bytes = 0, o_need = 0;
};

static int plan_job(const tpcf_job &job, const tpcf_binspec &spec, int mode, JobPlan &P) {
    if (int e = check_job(job)) return e;
    if (int e = binspec_check(spec)) return e;
    if (mode != TPCF_MODE_FAST && mode != TPCF_MODE_PLAIN && mode != TPCF_MODE_FAST64 && mode != TPCF_MODE_CELLS)
        return set_error(TPCF_EINVAL, "unknown mode %d", mode);
    if (spec.param != TPCF_R123 && mode != TPCF_MODE_PLAIN && spec.ax[0].n > kMaxFastSqtBins)
        return set_error(TPCF_EINVAL, "the FAST (s,q,theta) kernel takes up to %d s bins", kMaxFastSqtBins);
    P.L = make_layout(job);
    P.param = spec.param;
    P.n = spec.ax[0].n;
    P.nterms = job.nterms;
    P.cells = hist_cells(spec);
    P.rel = (float)(spec.recheck_rel > 0 ? spec.recheck_rel : 0.0);
    double e[kMaxBins + 1];
    axis_edges(spec.ax[0], e);
    P.T64.resize(P.n + 1);
    sq_thresholds(e, P.n, P.T64.data());
    P.T32.resize(P.n + 1);
    for (int m = 0; m <= P.n; ++m) P.T32[m] = (float)P.T64[m];
    if (spec.param != TPCF_R123) {
        P.nq = spec.ax[1].n;
        P.nt = spec.ax[2].n;
        P.axq = spec.ax[1];
        P.axt = spec.ax[2];
        P.Eq.resize(P.nq + 1);
        P.Et.resize(P.nt + 1);
        axis_edges(spec.ax[1], P.Eq.data());
        axis_edges(spec.ax[2], P.Et.data());
        // theta/pi edges become a descending cosine table: theta >= pi*t
        // <=> cos(theta) <= cos(pi*t) on [0, pi].
        if (spec.param == TPCF_SQT)
            for (double &t : P.Et) t = std::cos(M_PI * t);
    }
    P.blk_count.assign(P.L.nblk, 0);
    for (int s = 0; s < P.L.nseg; ++s)
        for (int64_t b = 0; b < P.L.seg_nblk[s]; ++b)
            P.blk_count[P.L.seg_blk0[s] + b] = (uint8_t)std::min<int64_t>(kBlock, P.L.seg_n[s] - b * kBlock);
    // Origin and extent of the union for the FP32 copy.
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int s = 0; s < job.nseg; ++s) {
        if (job.seg_hi[s] <= job.seg_lo[s]) continue;
        double slo[3], shi[3];
        seg_bounds(job.seg[s], job.seg_lo[s], job.seg_hi[s], slo, shi);
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], slo[d]);
            hi[d] = std::max(hi[d], shi[d]);
        }
    }
    double r = 0;
    for (int d = 0; d < 3; ++d) {
        if (!(hi[d] >= lo[d])) continue;  // empty job
        if (!std::isfinite(lo[d]) || !std::isfinite(hi[d]))
            return set_error(TPCF_EINVAL, "catalog coordinates must be finite");
        P.origin[d] = 0.5 * (lo[d] + hi[d]);
        r = std::max(r, std::max(hi[d] - P.origin[d], P.origin[d] - lo[d]));
    }
    P.rmax = (float)(r * (1.0 + 1e-6)) + 1e-30f;
    P.triplets = 0;
    for (int t = 0; t < job.nterms; ++t) P.triplets += term_triplets(P.L, job.term[t]);
    P.need_tiles = mode != TPCF_MODE_PLAIN;
    size_t o = 0;
    P.o_u64 = o; o = align256(o + sizeof(double) * 3 * P.L.npad);
    P.o_u32 = o; o = align256(o + (P.need_tiles ? sizeof(float4) * P.L.npad : 0));
    P.o_blk = o; o = align256(o + P.L.nblk + 1);
    P.o_T64 = o; o = align256(o + sizeof(double) * (P.n + 1));
    P.o_T32 = o; o = align256(o + sizeof(float) * (P.n + 1));
    P.o_Eq = o; o = align256(o + sizeof(double) * P.Eq.size());
    P.o_Et = o; o = align256(o + sizeof(double) * P.Et.size());
    P.o_cnt = o; o = align256(o + 64);
    P.o_hist = o; o = align256(o + sizeof(uint64_t) * P.cells * P.nterms);
    P.o_tiles = o; o = align256(o + (P.need_tiles ? (size_t)tile_count(P.L.nblk) * kTileBytes : 0));
    P.o_any = o; o = align256(o + (mode != TPCF_MODE_PLAIN ? (size_t)tile_count(P.L.nblk) : 0));
    // (r1,r2,r3) FAST paths with bins for the bit-mask kernel: per tile its
    // summary and row masks.
    const bool bitmask = spec.param == TPCF_R123 && mode != TPCF_MODE_PLAIN && P.n <= kMaxLdsR123Bins;
    P.o_meta = o; o = align256(o + (bitmask ? (size_t)tile_count(P.L.nblk) * sizeof(TileMeta) : 0));
    P.o_rmask = o;
    o = align256(o + (bitmask ? (size_t)tile_count(P.L.nblk) * kRowMaskWords * sizeof(unsigned long long) : 0));
    P.o_tstat = o; o = align256(o + (bitmask ? 192 * sizeof(unsigned long long) : 0));
    P.bitmask = bitmask;
    const bool sqt_fast = spec.param != TPCF_R123 && mode != TPCF_MODE_PLAIN;
    P.o_dists = o; o = align256(o + (sqt_fast ? (size_t)tile_count(P.L.nblk) * kTileBytes * sizeof(float) : 0));
    P.o_masks = o;
    o = align256(o + (sqt_fast ? (size_t)tile_count(P.L.nblk) * kMaskWords * sizeof(unsigned long long) : 0));
    if (sqt_fast && spec.param == TPCF_SQT) {
        // No entry of a triangle with a side >= s_hi*(1+q_hi) can be in range
        // (its s side is < s_hi, the q side < q_hi*s_hi, the third side at
        // most their sum); a relative margin keeps the cut conservative.
        const double dmax = spec.ax[0].hi * (1.0 + std::max(spec.ax[1].hi, 1.0)) * (1.0 + 1e-6);
        P.far_s = dmax * dmax;
        // An entry's q side is below q_hi times its s side, which is below
        // s_hi (and the s side is short itself).
        const double dmid = spec.ax[0].hi * std::max(spec.ax[1].hi, 1.0) * (1.0 + 1e-6);
        P.mid_s = dmid * dmid;
    } else if (sqt_fast) {
        // JBJ: the shortest side l1 < s_hi, l2 <= l1/q_lo and l3 < l2 +
        // v_hi*l1, so no side of an entry reaches s_hi*(1/q_lo + v_hi); with
        // q_lo = 0 there is no such bound and no pair is far.
        const double qlo = spec.ax[1].lo, vhi = std::max(spec.ax[2].hi, 0.0);
        const double dmax = spec.ax[0].hi * (1.0 / qlo + vhi) * (1.0 + 1e-6);
        P.far_s = qlo > 0 ? dmax * dmax : HUGE_VAL;
        // The entry's second side l2 <= l1 / q_lo < s_hi / q_lo.
        const double dmid = spec.ax[0].hi / qlo * (1.0 + 1e-6);
        P.mid_s = qlo > 0 ? dmid * dmid : HUGE_VAL;
    }
    P.cull = mode == TPCF_MODE_CELLS;
    // Spatial order also helps the (s,q,theta) and JBJ kernels unculled: rows
    // and blocks of the masks are empty as a whole more often (10k + 10k,
    // four terms, s [2,20) in a 200 Mpc/h cube: 50.8 -> 42.8 ms).  The
    // (r1,r2,r3) triplet kernel adds runs of equal cells along a tile row,
    // which spatial order makes long.
    P.spatial = mode != TPCF_MODE_PLAIN;
    if (P.spatial) {
        // (r1,r2,r3): every side of a counted triplet is below T[n].
        // (s,q,theta), JBJ: no side of an entry's triangle exceeds far_s.
        if (P.cull) P.cut2 = spec.param == TPCF_R123 ? P.T64[P.n] : std::nextafter(P.far_s, HUGE_VAL);
        P.o_box = o;
        o = align256(o + sizeof(double) * 6 * P.L.nblk);
        int64_t nmax = 0;
        for (int s = 0; s < P.L.nseg; ++s) nmax = std::max(nmax, P.L.seg_n[s]);
        P.sort_bytes = sort_segment_scratch_bytes((int)nmax);
        P.o_sort = o;
        o = align256(o + P.sort_bytes);
        // The reach bitsets on the device: the pair stage skips the tiles
        // no culled item reads.
        P.o_need = o;
        o = align256(o + sizeof(uint64_t) * (size_t)P.L.nblk * (size_t)((P.L.nblk + 63) / 64));
    }
    P.bytes = o;
    return TPCF_OK;
}

// Upload the small tables, gather the segments into the union layout and
// clear histograms/counters.  All on `st`, asynchronously; with culling the
// block boxes are queued for download into P.boxes, and cull_job(), once the
// stream has been synchronized, builds the reach rows from them.
static int stage_job(const tpcf_job &job, const JobPlan &P, char *base, hipStream_t st) {
    double *u64 = reinterpret_cast<double *>(base + P.o_u64);
    // the padding between segments reads as zeros (a fill kernel, like every
    // other clear of a job's region)
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(u64), 6 * P.L.npad, 0u, st));
    for (int s = 0; s < job.nseg; ++s) {
        const tpcf_catalog *c = job.seg[s];
        const int64_t m = P.L.seg_n[s];
        if (!m) continue;
        if (P.spatial) {
            // Hilbert order (the counts do not depend on the order) and the
            // box of every 64-point block.
            double lo[3], hi[3];
            seg_bounds(c, job.seg_lo[s], job.seg_lo[s] + m, lo, hi);
            const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
            const int64_t lo_i = job.seg_lo[s];
            HIP_TRY(sort_segment(c->d + lo_i, c->d + c->n + lo_i, c->d + 2 * c->n + lo_i, (int)m, lo, span,
                                 base + P.o_sort, P.sort_bytes, u64 + P.L.seg_off[s], P.L.npad, kBlock,
                                 reinterpret_cast<double *>(base + P.o_box) + 6 * P.L.seg_blk0[s], st));
            continue;
        }
        for (int d = 0; d < 3; ++d)
            HIP_TRY(hipMemcpyAsync(u64 + d * P.L.npad + P.L.seg_off[s], c->d + d * c->n + job.seg_lo[s],
                                   sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    }
    if (P.cull && P.L.nblk) {
        P.boxes.resize((size_t)6 * P.L.nblk);
        HIP_TRY(hipMemcpyAsync(P.boxes.data(), base + P.o_box, sizeof(double) * P.boxes.size(), hipMemcpyDeviceToHost,
                               st));
    }
    HIP_TRY(hipMemcpyAsync(base + P.o_blk, P.blk_count.data(), P.blk_count.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(base + P.o_T64, P.T64.data(), sizeof(double) * P.T64.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(base + P.o_T32, P.T32.data(), sizeof(float) * P.T32.size(), hipMemcpyHostToDevice, st));
    if (!P.Eq.empty()) {
        HIP_TRY(hipMemcpyAsync(base + P.o_Eq, P.Eq.data(), sizeof(double) * P.Eq.size(), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(base + P.o_Et, P.Et.data(), sizeof(double) * P.Et.size(), hipMemcpyHostToDevice, st));
    }
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(base + P.o_cnt), 16, 0u, st));
    if (P.bitmask) HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(base + P.o_tstat), 384, 0u, st));
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(base + P.o_hist), 2 * P.cells * P.nterms, 0u, st));
    if (P.need_tiles && P.L.npad) HIP_TRY(launch_to_f32(u64, P.L.npad, P.origin, reinterpret_cast<float4 *>(base + P.o_u32), st));
    return TPCF_OK;
}

// The reach rows of a staged culled job (its boxes downloaded), uploaded
// for the pair stage.
static int cull_job(const JobPlan &P, char *base, hipStream_t st) {
    if (!P.cull || !P.L.nblk) return TPCF_OK;
    P.reach = block_reach(P.boxes, P.L.nblk, P.cut2);
    HIP_TRY(hipMemcpyAsync(base + P.o_need, P.reach.bits.data(), sizeof(uint64_t) * P.reach.bits.size(),
                           hipMemcpyHostToDevice, st));
    return TPCF_OK;
}

// Stage jobs[j] at bases[j]: every job's staging queued, one synchronization
// when any job culls (not one per job), then the reach rows.
static int stage_jobs(const tpcf_job *const *jobs, const JobPlan *const *plans, char *const *bases, int njobs,
                      hipStream_t st) {
    bool cull = false;
    for (int j = 0; j < njobs; ++j) {
        if (int e = stage_job(*jobs[j], *plans[j], bases[j], st)) return e;
        cull = cull || (plans[j]->cull && plans[j]->L.nblk);
    }
    if (!cull) return TPCF_OK;
    HIP_TRY(hipStreamSynchronize(st));
    for (int j = 0; j < njobs; ++j)
        if (int e = cull_job(*plans[j], bases[j], st)) return e;
    return TPCF_OK;
}

// shift: 2 when the launch's triplet kernels are the LDS ones (every
// (r1,r2,r3) job of it has at most kMaxLdsR123Bins bins), 0 for the wide one.
static PairJob pair_job(const JobPlan &P, char *base, int shift) {
    PairJob J{};
    J.u64 = reinterpret_cast<const double *>(base + P.o_u64);
    J.u32 = reinterpret_cast<const float4 *>(base + P.o_u32);
    J.blk_count = reinterpret_cast<const uint8_t *>(base + P.o_blk);
    J.tiles = reinterpret_cast<uint8_t *>(base + P.o_tiles);
    J.tile_any = reinterpret_cast<uint8_t *>(base + P.o_any);
    const bool bits = P.bitmask && shift == 2;
    J.meta = bits ? reinterpret_cast<TileMeta *>(base + P.o_meta) : nullptr;
    J.rmask = bits ? reinterpret_cast<unsigned long long *>(base + P.o_rmask) : nullptr;
    J.tstat = bits ? reinterpret_cast<unsigned long long *>(base + P.o_tstat) : nullptr;
    J.shift = shift;
    J.T64 = reinterpret_cast<const double *>(base + P.o_T64);
    J.T32 = reinterpret_cast<const float *>(base + P.o_T32);
    J.rechecked = reinterpret_cast<unsigned long long *>(base + P.o_cnt);
    J.npad = P.L.npad;
    J.nblk = P.L.nblk;
    J.n = P.n;
    J.rmax = P.rmax;
    J.rel = P.rel;
    J.need = P.cull ? reinterpret_cast<const unsigned long long *>(base + P.o_need) : nullptr;
    J.need_words = (P.L.nblk + 63) / 64;
    return J;
}

// ---------------------------------------------------------------------------
// workspaces
// ---------------------------------------------------------------------------

struct Workspace {
    DevBuf region;     // job regions
    DevBuf queue;      // items + terms + counter of the (r1,r2,r3) kernel
    DevBuf sqt_queue;  // items + terms + counter of the (s,q,theta) kernel
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // Control block of the last (r1,r2,r3) LDS-kernel launch on this
    // workspace (device words, kCtl*), or nullptr.
    const int *ctl = nullptr;
    ~Workspace() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    int init() {
        for (auto &e : ev)
            if (!e) HIP_TRY(hipEventCreate(&e));
        return TPCF_OK;
    }
};

void workspace_order(Workspace *ws, hipStream_t st, hipMemPool_t pool) {
    for (DevBuf *b : {&ws->region, &ws->queue, &ws->sqt_queue}) {
        b->ordered = st;
        b->pool = pool;
    }
}

int worker_pool_create(int device, hipMemPool_t *pool) {
    *pool = nullptr;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    HIP_TRY(hipMemPoolCreate(pool, &props));
    int off = 0;
    uint64_t keep = 0;
    if (hipError_t e = hipMemPoolSetAttribute(*pool, hipMemPoolReuseAllowInternalDependencies, &off);
        e != hipSuccess || (e = hipMemPoolSetAttribute(*pool, hipMemPoolAttrReleaseThreshold, &keep)) != hipSuccess) {
        (void)hipMemPoolDestroy(*pool);
        *pool = nullptr;
        return set_error(TPCF_EDEVICE, "cannot configure a worker's memory pool: %s", hipGetErrorString(e));
    }
    return TPCF_OK;
}

// Launch shapes of the LDS triplet kernels for bins up to max_n.  The
// run-merging kernel: as many privatized histogram copies (4, 2 or 1) as fit
// the LDS budget, and enough waves per workgroup for four waves per SIMD at
// the resulting occupancy.  The bit-mask kernel: one histogram copy, 8 waves
// per workgroup, as many workgroups per CU as the occupancy allows.  form:
// which of the two counts the launch (-1: chosen on the device from the
// tiles' bins).  TPCF_R123_SHAPE="waves,copies[,wg_per_cu]" overrides the
// run-merging shape and TPCF_R123_FORM / TPCF_R123_BITS_MAX the choice
// (tuning runs).
struct TripletShape {
    int waves, copies, grid;  // run-merging kernel
    int bwaves, bgrid;        // bit-mask kernel
    int form;
    float bits_max;
    int split_bins;
    float split_mean;
};

static TripletShape triplet_shape(tpcf_ctx *ctx, int max_n) {
    constexpr int kLdsBudget = 155 * 1024;
    TripletShape s{8, 1, 0, 8, 0, ctx->shape.form, ctx->shape.bits_max, ctx->shape.split_bins, ctx->shape.split_mean};
    for (int k : {8, 4, 2, 1})
        if (r123_lds_bytes(max_n, k) <= kLdsBudget) {
            s.copies = k;
            break;
        }
    const int lds = r123_lds_bytes(max_n, s.copies);
    s.waves = lds > kLdsBudget / 2 ? 16 : 8;
    int per_cu_cap = 0;
    if (ctx->shape.waves > 0 && r123_triplet_kernel_ptr(ctx->shape.waves, ctx->shape.copies) &&
        r123_lds_bytes(max_n, ctx->shape.copies) <= kLdsBudget) {
        s.waves = ctx->shape.waves;
        s.copies = ctx->shape.copies;
        per_cu_cap = ctx->shape.per_cu;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, r123_triplet_kernel_ptr(s.waves, s.copies),
                                                     s.waves * 64, r123_lds_bytes(max_n, s.copies)) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (per_cu_cap > 0) per_cu = std::min(per_cu, per_cu_cap);
    s.grid = ctx->num_cu * per_cu;
    int bper_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bper_cu, r123_bits_kernel_ptr(s.bwaves), s.bwaves * 64,
                                                     r123_bits_lds_bytes(max_n)) != hipSuccess ||
        bper_cu < 1)
        bper_cu = 1;
    s.bgrid = ctx->num_cu * bper_cu;
    return s;
}

// The triplet work queue of several jobs: one TripletTerm per (job, term)
// and the items of every term, longest C ranges first within a term
// (dynamic LPT order for the persistent kernel).
struct Queue {
    std::vector<TripletTerm> terms;
    std::vector<TripletItem> items;
    int max_n = 1;
};

static void build_queue(const std::vector<const JobPlan *> &plans, const std::vector<const tpcf_job *> &jobs,
                        const std::vector<char *> &bases, Queue &Q) {
    for (size_t j = 0; j < plans.size(); ++j) {
        const JobPlan &P = *plans[j];
        for (int t = 0; t < P.nterms; ++t) {
            TripletTerm T{};
            T.tiles = reinterpret_cast<const uint8_t *>(bases[j] + P.o_tiles);
            T.tile_any = reinterpret_cast<const uint8_t *>(bases[j] + P.o_any);
            T.meta = P.bitmask ? reinterpret_cast<const TileMeta *>(bases[j] + P.o_meta) : nullptr;
            T.rmask = P.bitmask ? reinterpret_cast<const unsigned long long *>(bases[j] + P.o_rmask) : nullptr;
            T.tstat = P.bitmask ? reinterpret_cast<const unsigned long long *>(bases[j] + P.o_tstat) : nullptr;
            T.nblk = P.L.nblk;
            T.hist = reinterpret_cast<unsigned long long *>(bases[j] + P.o_hist) + (size_t)t * P.cells;
            T.n = P.n;
            term_items(P.L, jobs[j]->term[t], (int)Q.terms.size(), Q.items, P.cull ? &P.reach : nullptr,
                       kMergeGapBlocks);
            Q.terms.push_back(T);
            Q.max_n = std::max(Q.max_n, P.n);
        }
    }
}

// Device copy of a queue.
struct QueueDev {
    const TripletItem *items = nullptr;
    const TripletTerm *terms = nullptr;
    int *counter = nullptr;
    int nitems = 0;
    int nterms = 0;
    int max_n = 1;
};

static int upload_queue(DevBuf &buf, const Queue &Q, hipStream_t st, QueueDev &D) {
    D.nitems = (int)Q.items.size();
    D.nterms = (int)Q.terms.size();
    D.max_n = Q.max_n;
    if (Q.items.empty()) return TPCF_OK;
    const size_t o_terms = align256(sizeof(TripletItem) * Q.items.size());
    const size_t o_ctr = align256(o_terms + sizeof(TripletTerm) * Q.terms.size());
    if (int e = buf.reserve(o_ctr + 256)) return e;
    char *q = reinterpret_cast<char *>(buf.p);
    HIP_TRY(hipMemcpyAsync(q, Q.items.data(), sizeof(TripletItem) * Q.items.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(q + o_terms, Q.terms.data(), sizeof(TripletTerm) * Q.terms.size(), hipMemcpyHostToDevice,
                           st));
    D.items = reinterpret_cast<const TripletItem *>(q);
    D.terms = reinterpret_cast<const TripletTerm *>(q + o_terms);
    D.counter = reinterpret_cast<int *>(q + o_ctr);
    return TPCF_OK;
}

// Reset the queue counter and launch the triplet kernel (kernels only, so
// the sequence can be captured in a graph).
static int launch_queue(tpcf_ctx *ctx, const QueueDev &D, const TripletShape &sh, hipStream_t st,
                        ElasticHook *hook = nullptr) {
    if (!D.nitems) return TPCF_OK;
    // the control block: item counter and the bit-mask kernel's 64-bit op
    // count (the choose kernel writes the form)
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(D.counter), kCtlForm, 0u, st));
    if (D.max_n > kMaxLdsR123Bins) {
        // Wide specs: pair values are unscaled; global-atomic kernel.
        HIP_TRY(launch_r123_triplets_wide(D.items, D.nitems, D.terms, ctx->num_cu * 8, st));
        return TPCF_OK;
    }
    // Both LDS kernels are launched; the one not chosen by the choose kernel
    // returns at once (either one gives the same histograms), and under the
    // split form each counts its own items.  Elastic (serving) launches use
    // one form only: their workgroups leave by one active count per launch,
    // and the last active one stays to drain the queue, which holds for one
    // queue but not for the split form's two.
    const int split_bins = hook ? -1 : sh.split_bins;
    const int form = hook && sh.form == kFormSplit ? kFormRuns : sh.form;
    HIP_TRY(launch_r123_choose(D.terms, D.nterms, D.counter, sh.bits_max, form, split_bins, sh.split_mean, st));
    ElasticCtl ctl{};
    int grid = sh.grid, bgrid = sh.bgrid;
    if (hook) {
        // One share for the launch, sized by the bit-mask kernel's occupancy;
        // the run-merging kernel's workgroups beyond its own occupancy wait
        // to be scheduled, and leave at once when over the share.
        const QueueDev d = D;
        const TripletShape s = sh;
        auto relaunch = [d, s](int wgs, hipStream_t other, const ElasticCtl &c) -> int {
            HIP_TRY(launch_r123_triplets(c, d.items, d.nitems, d.counter, d.terms, d.max_n, s.waves, s.copies, wgs,
                                         other));
            HIP_TRY(launch_r123_bits(c, d.items, d.nitems, d.counter, d.terms, d.max_n, s.bwaves, wgs, other));
            return TPCF_OK;
        };
        if (int e = hook->control(std::max(1, sh.bgrid / ctx->num_cu), D.nitems, relaunch, st, &ctl, &bgrid))
            return e;
        if (ctl.active) grid = bgrid;
    }
    HIP_TRY(launch_r123_triplets(ctl, D.items, D.nitems, D.counter, D.terms, D.max_n, sh.waves, sh.copies, grid,
                                 st));
    HIP_TRY(launch_r123_bits(ctl, D.items, D.nitems, D.counter, D.terms, D.max_n, sh.bwaves, bgrid, st));
    return TPCF_OK;
}

// Launch the triplet stage of several jobs as one persistent kernel.
static int launch_triplets(tpcf_ctx *ctx, Workspace &ws, const std::vector<const JobPlan *> &plans,
                           const std::vector<const tpcf_job *> &jobs, const std::vector<char *> &bases,
                           hipStream_t st, ElasticHook *hook = nullptr) {
    Queue Q;
    build_queue(plans, jobs, bases, Q);
    QueueDev D;
    if (int e = upload_queue(ws.queue, Q, st, D)) return e;
    ws.ctl = D.nitems && D.max_n <= kMaxLdsR123Bins ? D.counter : nullptr;
    HIP_TRY(hipEventRecord(ws.ev[3], st));
    return launch_queue(ctx, D, triplet_shape(ctx, Q.max_n), st, hook);
}

// Triplets a workgroup of the plain (r1,r2,r3) kernel adds to its 32-bit LDS
// counters between flushes: far below 2^32.  TPCF_PLAIN_FLUSH_LIMIT lowers it
// (tests run the flush path many times with a tiny limit).
static unsigned long long plain_flush_limit() {
    unsigned long long lim = 0xC0000000ull;
    if (const char *e = std::getenv("TPCF_PLAIN_FLUSH_LIMIT")) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0 && v < lim) lim = v;
    }
    return lim;
}

static int launch_plain(tpcf_ctx *ctx, Workspace &ws, const JobPlan &P, const tpcf_job &job, char *base,
                        hipStream_t st) {
    for (int t = 0; t < P.nterms; ++t) {
        std::vector<TripletItem> items;
        term_items(P.L, job.term[t], 0, items);
        if (items.empty()) continue;
        if (int e = ws.queue.reserve(sizeof(TripletItem) * items.size())) return e;
        HIP_TRY(hipMemcpyAsync(ws.queue.p, items.data(), sizeof(TripletItem) * items.size(), hipMemcpyHostToDevice, st));
        if (P.param != TPCF_R123) {
            SqtArgs S{};
            S.u64 = reinterpret_cast<const double *>(base + P.o_u64);
            S.blk_count = reinterpret_cast<const uint8_t *>(base + P.o_blk);
            S.Ts = reinterpret_cast<const double *>(base + P.o_T64);
            S.Eq = reinterpret_cast<const double *>(base + P.o_Eq);
            S.Et = reinterpret_cast<const double *>(base + P.o_Et);
            S.items = reinterpret_cast<const TripletItem *>(ws.queue.p);
            S.nitems = (int)items.size();
            S.ns = P.n;
            S.nq = P.nq;
            S.nt = P.nt;
            S.npad = P.L.npad;
            S.hist = reinterpret_cast<unsigned long long *>(base + P.o_hist) + (size_t)t * P.cells;
            const int grid = std::min<int64_t>((int64_t)items.size(), (int64_t)ctx->num_cu * 2);
            HIP_TRY(launch_sqt_plain(S, P.param, grid, st));
            HIP_TRY(hipStreamSynchronize(st));
            continue;
        }
        PlainArgs A{};
        A.u64 = reinterpret_cast<const double *>(base + P.o_u64);
        A.blk_count = reinterpret_cast<const uint8_t *>(base + P.o_blk);
        A.T64 = reinterpret_cast<const double *>(base + P.o_T64);
        A.items = reinterpret_cast<const TripletItem *>(ws.queue.p);
        A.nitems = (int)items.size();
        A.n = P.n;
        A.npad = P.L.npad;
        A.hist = reinterpret_cast<unsigned long long *>(base + P.o_hist) + (size_t)t * P.cells;
        A.flush_limit = plain_flush_limit();
        const int grid = std::min<int64_t>((int64_t)items.size(), (int64_t)ctx->num_cu * 4);
        HIP_TRY(launch_r123_plain(A, grid, st));
        // The item buffer is reused by the next term: finish this one first.
        HIP_TRY(hipStreamSynchronize(st));
    }
    return TPCF_OK;
}

// FAST (s,q,theta) and JBJ for several jobs: every job's pair stage (FP64
// pair codes, FP32 side lengths, candidate masks), then one persistent
// candidate launch over the items of every job and term (FP32 entries, FP64
// recheck near edges; FAST64 sends every candidate through FP64).  A
// workgroup reloads the edge tables and flushes its histogram when the term
// changes.  The host part (SqtLaunch) is built once; sqt_enqueue launches
// kernels only, so a prepared batch can capture it in a graph.
struct SqtLaunch {
    std::vector<SqtPairJob> pairs;
    std::vector<SqtTerm> terms;
    std::vector<TripletItem> items;
    int lds_cells = 1, waves = 8, qcap = 0, per_cu = 0;
    size_t o_terms = 0, o_ctr = 0;
};

// The pair stage of one staged (s,q,theta) or JBJ job.
static SqtPairJob sqt_pair_job(const JobPlan &P, char *base) {
    SqtPairJob J{};
    J.u64 = reinterpret_cast<const double *>(base + P.o_u64);
    J.blk_count = reinterpret_cast<const uint8_t *>(base + P.o_blk);
    J.codes = reinterpret_cast<uint8_t *>(base + P.o_tiles);
    J.dists = reinterpret_cast<float *>(base + P.o_dists);
    J.masks = reinterpret_cast<unsigned long long *>(base + P.o_masks);
    J.tile_any = reinterpret_cast<uint8_t *>(base + P.o_any);
    J.Ts = reinterpret_cast<const double *>(base + P.o_T64);
    J.ns = P.n;
    J.far_s = P.far_s;
    J.mid_s = P.mid_s;
    J.npad = P.L.npad;
    J.nblk = P.L.nblk;
    J.need = P.cull ? reinterpret_cast<const unsigned long long *>(base + P.o_need) : nullptr;
    J.need_words = (P.L.nblk + 63) / 64;
    return J;
}

static int sqt_plan(tpcf_ctx *ctx, const std::vector<const JobPlan *> &plans,
                    const std::vector<const tpcf_job *> &jobs, const std::vector<char *> &bases, int mode,
                    SqtLaunch &L) {
    L = SqtLaunch{};
    for (size_t j = 0; j < plans.size(); ++j) {
        const JobPlan &P = *plans[j];
        char *base = bases[j];
        if (!P.L.nblk) continue;
        const SqtPairJob J = sqt_pair_job(P, base);
        L.pairs.push_back(J);
        SqtTerm T{};
        T.codes = J.codes;
        T.dists = J.dists;
        T.masks = J.masks;
        T.tile_any = J.tile_any;
        T.u64 = J.u64;
        T.Ts = J.Ts;
        T.Eq = reinterpret_cast<const double *>(base + P.o_Eq);
        T.Et = reinterpret_cast<const double *>(base + P.o_Et);
        T.rechecked = reinterpret_cast<unsigned long long *>(base + P.o_cnt);
        T.npad = P.L.npad;
        T.nblk = P.L.nblk;
        T.param = P.param == TPCF_JBJ ? kJbj : kSqt;
        T.ns = P.n;
        T.nq = P.nq;
        T.nt = P.nt;
        T.cells = (int)P.cells;
        T.force_fp64 = mode == TPCF_MODE_FAST64;
        const auto guess = [](const tpcf_axis &ax, float &lo, float &inv, int &lg) {
            lg = ax.scale == TPCF_LOG;
            lo = (float)(lg ? std::log2(ax.lo) : ax.lo);
            inv = (float)(ax.n / (lg ? std::log2(ax.hi / ax.lo) : ax.hi - ax.lo));
        };
        guess(P.axq, T.q_lo, T.q_inv, T.q_log);
        guess(P.axt, T.t_lo, T.t_inv, T.t_log);
        L.lds_cells = std::max(L.lds_cells, T.cells);
        for (int t = 0; t < P.nterms; ++t) {
            T.hist = reinterpret_cast<unsigned long long *>(base + P.o_hist) + (size_t)t * P.cells;
            term_items(P.L, jobs[j]->term[t], (int)L.terms.size(), L.items, P.cull ? &P.reach : nullptr,
                       kMergeGapBlocks);
            L.terms.push_back(T);
        }
    }
    if (L.items.empty()) return TPCF_OK;
    L.o_terms = align256(sizeof(TripletItem) * L.items.size());
    L.o_ctr = align256(L.o_terms + sizeof(SqtTerm) * L.terms.size());
    // Candidate rings of 512 words per wave, fewer where the histogram needs
    // the LDS (the limit on cells keeps 128 possible); 8 waves per workgroup,
    // or 16 when the LDS would leave a CU a single 8-wave workgroup.
    const auto fit = [&](int waves, int &qcap, int &per_cu) -> bool {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, sqt_fast_kernel_ptr(waves)) != hipSuccess) return false;
        size_t lds = 0;
        for (qcap = waves == 8 ? 512 : 256; qcap >= 128; qcap /= 2) {
            lds = sqt_fast_lds_bytes(L.lds_cells, waves, qcap);
            if (fa.sharedSizeBytes + lds <= 160 * 1024) break;
        }
        if (qcap < 128) return false;
        per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sqt_fast_kernel_ptr(waves), waves * 64, lds) !=
            hipSuccess)
            per_cu = 0;
        return per_cu >= 1;
    };
    const bool ok8 = fit(8, L.qcap, L.per_cu);
    if (!ok8 || L.per_cu < 2) {
        int q16 = 0, pc16 = 0;
        if (fit(16, q16, pc16)) {
            L.waves = 16;
            L.qcap = q16;
            L.per_cu = pc16;
        } else if (!ok8) {
            (void)hipGetLastError();
            return set_error(TPCF_EINVAL, "(s,q,theta) histogram of %d cells does not fit the FAST kernel's LDS",
                             L.lds_cells);
        }
    }
    return TPCF_OK;
}

// Copies the items and term descriptors of L into buf (asynchronously on st).
static int sqt_upload(const SqtLaunch &L, DevBuf &buf, hipStream_t st) {
    if (L.items.empty()) return TPCF_OK;
    if (int e = buf.reserve(L.o_ctr + 256)) return e;
    char *q = reinterpret_cast<char *>(buf.p);
    HIP_TRY(hipMemcpyAsync(q, L.items.data(), sizeof(TripletItem) * L.items.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(q + L.o_terms, L.terms.data(), sizeof(SqtTerm) * L.terms.size(), hipMemcpyHostToDevice,
                           st));
    return TPCF_OK;
}

// The kernels of L: pair stages (unless already enqueued), then the counter
// reset and the persistent candidate launch.  When given, ev_pairs is
// recorded after the pair stages and ev_kernel before the candidate launch
// (stage timings).
static int sqt_enqueue(tpcf_ctx *ctx, const SqtLaunch &L, const DevBuf &buf, hipStream_t st, hipEvent_t ev_pairs,
                       hipEvent_t ev_kernel, bool pairs = true, ElasticHook *hook = nullptr) {
    if (pairs)
        for (const SqtPairJob &J : L.pairs) HIP_TRY(launch_sqt_pairs(J, st));
    if (ev_pairs) HIP_TRY(hipEventRecord(ev_pairs, st));
    if (ev_kernel) HIP_TRY(hipEventRecord(ev_kernel, st));
    if (L.items.empty()) return TPCF_OK;
    char *q = reinterpret_cast<char *>(buf.p);
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(q + L.o_ctr), 1, 0u, st));
    SqtFastArgs A{};
    A.terms = reinterpret_cast<const SqtTerm *>(q + L.o_terms);
    A.items = reinterpret_cast<const TripletItem *>(q);
    A.nitems = (int)L.items.size();
    A.counter = reinterpret_cast<int *>(q + L.o_ctr);
    A.lds_cells = L.lds_cells;
    A.qcap = L.qcap;
    int grid = ctx->num_cu * L.per_cu;
    if (hook) {
        const SqtFastArgs a0 = A;
        const int waves = L.waves;
        auto relaunch = [a0, waves](int wgs, hipStream_t other, const ElasticCtl &c) -> int {
            SqtFastArgs a = a0;
            a.ctl = c;
            HIP_TRY(launch_sqt_fast(a, waves, wgs, other));
            return TPCF_OK;
        };
        if (int e = hook->control(L.per_cu, A.nitems, relaunch, st, &A.ctl, &grid)) return e;
    }
    HIP_TRY(launch_sqt_fast(A, L.waves, grid, st));
    return TPCF_OK;
}

static int launch_sqt_fast_jobs(tpcf_ctx *ctx, Workspace &ws, const std::vector<const JobPlan *> &plans,
                                const std::vector<const tpcf_job *> &jobs, const std::vector<char *> &bases, int mode,
                                hipStream_t st, ElasticHook *hook = nullptr) {
    // The pair stages first: the device runs them while the host plans the
    // items.
    for (size_t j = 0; j < plans.size(); ++j)
        if (plans[j]->L.nblk) HIP_TRY(launch_sqt_pairs(sqt_pair_job(*plans[j], bases[j]), st));
    HIP_TRY(hipEventRecord(ws.ev[1], st));
    SqtLaunch L;
    if (int e = sqt_plan(ctx, plans, jobs, bases, mode, L)) return e;
    if (int e = sqt_upload(L, ws.sqt_queue, st)) return e;
    return sqt_enqueue(ctx, L, ws.sqt_queue, st, nullptr, ws.ev[3], false, hook);
}

// The (r1,r2,r3) LDS launch's control block (after the launch has
// finished): the form that counted, the mean bins per tile it was chosen on
// and the bit-mask kernel's modelled VALU lane-ops.  No launch: form -1.
// Read on the launch's own stream: a null-stream hipMemcpy would also wait
// for every kernel on the other blocking streams (the serving workers'
// CU-masked streams are such streams), so a finished batch would wait for
// the longest batch running beside it, and later launches on those streams
// for that copy.
static int read_ctl(const int *ctl, tpcf_stats *s, hipStream_t st) {
    s->mask_ops = 0;
    s->triplet_form = -1;
    s->tile_bins_mean = 0.0;
    s->tile_fill = 0.0;
    if (!ctl) return TPCF_OK;
    int w[8];
    HIP_TRY(hipMemcpyAsync(w, ctl, sizeof w, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    unsigned long long ops;
    std::memcpy(&ops, w + kCtlOps, sizeof ops);
    float mean, fill;
    std::memcpy(&mean, w + kCtlBinsMean, sizeof mean);
    std::memcpy(&fill, w + kCtlFill, sizeof fill);
    s->mask_ops = w[kCtlForm] == kFormBits || w[kCtlForm] == kFormSplit ? ops : 0;
    s->triplet_form = w[kCtlForm];
    s->tile_bins_mean = mean;
    s->tile_fill = fill;
    return TPCF_OK;
}

// Pair-stage and triplet-kernel times from the workspace events.
static int stage_times(Workspace &ws, tpcf_stats *s) {
    float a = 0.f, b = 0.f;
    HIP_TRY(hipEventElapsedTime(&a, ws.ev[0], ws.ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, ws.ev[3], ws.ev[2]));
    s->pair_ms = a;
    s->triplet_ms = b;
    return TPCF_OK;
}

// Run one job synchronously on stream st with workspace ws.
static int run_job(tpcf_ctx *ctx, Workspace &ws, const tpcf_job &job, const tpcf_binspec &spec, int mode,
                   uint64_t *hist, tpcf_stats *stats, hipStream_t st) {
    const auto t0 = std::chrono::steady_clock::now();
    JobPlan P;
    if (int e = plan_job(job, spec, mode, P)) return e;
    if (mode == TPCF_MODE_CELLS) mode = TPCF_MODE_FAST;  // the plan holds the culling
    if (int e = ws.init()) return e;
    ws.ctl = nullptr;
    if (int e = ws.region.reserve(P.bytes)) return e;
    char *base = reinterpret_cast<char *>(ws.region.p);
    {
        const tpcf_job *jp = &job;
        const JobPlan *pp = &P;
        if (int e = stage_jobs(&jp, &pp, &base, 1, st)) return e;
    }
    HIP_TRY(hipEventRecord(ws.ev[0], st));
    HIP_TRY(hipEventRecord(ws.ev[1], st));
    HIP_TRY(hipEventRecord(ws.ev[3], st));
    if (mode == TPCF_MODE_PLAIN) {
        if (P.L.nblk) {
            if (int e = launch_plain(ctx, ws, P, job, base, st)) return e;
        }
    } else if (P.param != TPCF_R123 && P.L.nblk) {
        std::vector<const JobPlan *> plans{&P};
        std::vector<const tpcf_job *> jobs{&job};
        std::vector<char *> bases{base};
        if (int e = launch_sqt_fast_jobs(ctx, ws, plans, jobs, bases, mode, st)) return e;
    } else if (P.L.nblk) {
        HIP_TRY(launch_pair_tiles(pair_job(P, base, P.n <= kMaxLdsR123Bins ? 2 : 0), mode == TPCF_MODE_FAST, st));
        HIP_TRY(hipEventRecord(ws.ev[1], st));
        std::vector<const JobPlan *> plans{&P};
        std::vector<const tpcf_job *> jobs{&job};
        std::vector<char *> bases{base};
        if (int e = launch_triplets(ctx, ws, plans, jobs, bases, st)) return e;
    }
    HIP_TRY(hipEventRecord(ws.ev[2], st));
    HIP_TRY(hipMemcpyAsync(hist, base + P.o_hist, sizeof(uint64_t) * P.cells * P.nterms, hipMemcpyDeviceToHost, st));
    uint64_t rechecked = 0;
    HIP_TRY(hipMemcpyAsync(&rechecked, base + P.o_cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ws.ev[0], ws.ev[2]));
        int64_t nu = 0;
        for (int s = 0; s < P.L.nseg; ++s) nu += P.L.seg_n[s];
        stats->triplets = P.triplets;
        stats->pairs = mode == TPCF_MODE_PLAIN ? 0 : (uint64_t)(nu * (nu - 1) / 2);
        stats->pairs_rechecked = mode == TPCF_MODE_FAST ? rechecked : 0;
        stats->device_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (int e = stage_times(ws, stats)) return e;
        if (int e = read_ctl(ws.ctl, stats, st)) return e;
    }
    return TPCF_OK;
}

// Work items of a planned job: at most one per (A, B) block pair and run of
// up to kMaxItemBlocks C blocks, per term (term_items).
static uint64_t items_bound(const JobPlan &P, const tpcf_job &job) {
    uint64_t items = 0;
    for (int t = 0; t < job.nterms; ++t) {
        const int *tr = job.term[t];
        const uint64_t a = P.L.seg_nblk[tr[0]], b = P.L.seg_nblk[tr[1]], c = P.L.seg_nblk[tr[2]];
        // culled items end where reachable C runs are more than
        // kMergeGapBlocks apart
        items += a * b * (P.cull ? c / kMergeGapBlocks + c / kMaxItemBlocks + 2
                                 : (c + kMaxItemBlocks - 1) / kMaxItemBlocks + 1);
    }
    return items;
}

// Work-queue bytes (items, term table, counter) that upload_queue and
// sqt_upload need at most for these jobs.
static size_t queue_bound(const std::vector<JobPlan> &plans, const tpcf_job *jobs, bool r123) {
    uint64_t items = 0, terms = 0;
    for (size_t j = 0; j < plans.size(); ++j)
        if ((plans[j].param == TPCF_R123) == r123) {
            items += items_bound(plans[j], jobs[j]);
            terms += (uint64_t)jobs[j].nterms;
        }
    if (!items) return 0;
    const size_t tsz = r123 ? sizeof(TripletTerm) : sizeof(SqtTerm);
    return align256(align256(sizeof(TripletItem) * items) + tsz * terms) + 256;
}

// Run several jobs as one batch on stream st: per-job staging and pair
// stage, then one persistent triplet launch over every job's work items.
// per[j] receives job j's counters; total the batch figures (device_ms is
// the batch's kernel time).
int run_batch(tpcf_ctx *ctx, Workspace &ws, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs, int mode,
              uint64_t *const *hist, tpcf_stats *per, tpcf_stats *total, hipStream_t st, ElasticHook *hook) {
    if (mode != TPCF_MODE_FAST && mode != TPCF_MODE_FAST64 && mode != TPCF_MODE_CELLS)
        return set_error(TPCF_EINVAL, "batched counting runs the FAST paths only");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<JobPlan> plans(njobs);
    size_t bytes = 0;
    std::vector<size_t> off(njobs);
    for (int j = 0; j < njobs; ++j) {
        if (int e = plan_job(jobs[j], specs[j], mode, plans[j])) return e;
        off[j] = bytes;
        bytes = align256(bytes + plans[j].bytes);
    }
    if (hook)
        for (int j = 1; j < njobs; ++j)
            if ((plans[j].param == TPCF_R123) != (plans[0].param == TPCF_R123))
                return set_error(TPCF_EINVAL, "an elastic batch holds jobs of one family");
    // The hook's share is given back on every exit, once st has drained.
    struct Share {
        ElasticHook *h;
        hipStream_t st;
        void done() {
            if (!h) return;
            (void)hipStreamSynchronize(st);
            h->finish();
            h = nullptr;
        }
        ~Share() { done(); }
    } share{hook, st};
    if (mode == TPCF_MODE_CELLS) mode = TPCF_MODE_FAST;  // the plans hold the culling
    if (int e = ws.init()) return e;
    ws.ctl = nullptr;
    // An elastic batch reserves its workspace, work-queue buffers included,
    // before it claims its share: nothing after the claim allocates.
    if (int e = ws.region.reserve(std::max<size_t>(bytes, 256))) return e;
    if (hook) {
        // and so do the work-queue buffers, to their bound
        if (int e = ws.queue.reserve(queue_bound(plans, jobs, true))) return e;
        if (int e = ws.sqt_queue.reserve(queue_bound(plans, jobs, false))) return e;
        double work = 0.0;
        for (int j = 0; j < njobs; ++j) work += (double)plans[j].triplets;
        hook->claim(work);
    }
    char *base = reinterpret_cast<char *>(ws.region.p);
    {
        std::vector<const tpcf_job *> jp(njobs);
        std::vector<const JobPlan *> pp(njobs);
        std::vector<char *> bp(njobs);
        for (int j = 0; j < njobs; ++j) {
            jp[j] = &jobs[j];
            pp[j] = &plans[j];
            bp[j] = base + off[j];
        }
        if (int e = stage_jobs(jp.data(), pp.data(), bp.data(), njobs, st)) return e;
    }
    HIP_TRY(hipEventRecord(ws.ev[0], st));
    HIP_TRY(hipEventRecord(ws.ev[3], st));
    // (s,q,theta) and JBJ jobs: one persistent launch of their own.
    std::vector<const JobPlan *> sp;
    std::vector<const tpcf_job *> sj;
    std::vector<char *> sb;
    for (int j = 0; j < njobs; ++j)
        if (plans[j].param != TPCF_R123) {
            sp.push_back(&plans[j]);
            sj.push_back(&jobs[j]);
            sb.push_back(base + off[j]);
        }
    if (!sp.empty())
        if (int e = launch_sqt_fast_jobs(ctx, ws, sp, sj, sb, mode, st, hook)) return e;
    // (r1,r2,r3) jobs: pair stages, then one persistent launch.
    int max_n = 1;
    for (int j = 0; j < njobs; ++j)
        if (plans[j].param == TPCF_R123) max_n = std::max(max_n, plans[j].n);
    const int shift = max_n <= kMaxLdsR123Bins ? 2 : 0;
    std::vector<const JobPlan *> pp;
    std::vector<const tpcf_job *> jj;
    std::vector<char *> bb;
    for (int j = 0; j < njobs; ++j) {
        if (plans[j].param != TPCF_R123 || !plans[j].L.nblk) continue;
        HIP_TRY(launch_pair_tiles(pair_job(plans[j], base + off[j], shift), mode == TPCF_MODE_FAST, st));
        pp.push_back(&plans[j]);
        jj.push_back(&jobs[j]);
        bb.push_back(base + off[j]);
    }
    if (!pp.empty() || sp.empty()) {
        HIP_TRY(hipEventRecord(ws.ev[1], st));
        if (int e = launch_triplets(ctx, ws, pp, jj, bb, st, hook)) return e;
    }
    // The extra workgroups of an elastic launch ran on another stream: the
    // histograms are complete once the hook has waited for them.
    share.done();
    HIP_TRY(hipEventRecord(ws.ev[2], st));
    std::vector<uint64_t> rech(njobs, 0);
    for (int j = 0; j < njobs; ++j) {
        HIP_TRY(hipMemcpyAsync(hist[j], base + off[j] + plans[j].o_hist,
                               sizeof(uint64_t) * plans[j].cells * plans[j].nterms, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&rech[j], base + off[j] + plans[j].o_cnt, 8, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ws.ev[0], ws.ev[2]));
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    tpcf_stats tot{};
    for (int j = 0; j < njobs; ++j) {
        tpcf_stats s{};
        s.triplets = plans[j].triplets;
        int64_t nu = 0;
        for (int q = 0; q < plans[j].L.nseg; ++q) nu += plans[j].L.seg_n[q];
        s.pairs = (uint64_t)(nu * (nu - 1) / 2);
        s.pairs_rechecked = mode == TPCF_MODE_FAST ? rech[j] : 0;
        s.device_ms = ms;
        s.wall_ms = wall;
        if (per) per[j] = s;
        tot.triplets += s.triplets;
        tot.pairs += s.pairs;
        tot.pairs_rechecked += s.pairs_rechecked;
    }
    tot.device_ms = ms;
    tot.wall_ms = wall;
    if (int e = stage_times(ws, &tot)) return e;
    if (int e = read_ctl(ws.ctl, &tot, st)) return e;
    for (int j = 0; j < njobs && per; ++j) {
        per[j].pair_ms = tot.pair_ms;
        per[j].triplet_ms = tot.triplet_ms;
        per[j].mask_ops = tot.mask_ops;
        per[j].triplet_form = tot.triplet_form;
        per[j].tile_bins_mean = tot.tile_bins_mean;
        per[j].tile_fill = tot.tile_fill;
    }
    if (total) *total = tot;
    return TPCF_OK;
}

int run_single(tpcf_ctx *ctx, Workspace &ws, const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *hist,
               tpcf_stats *stats, hipStream_t st) {
    return run_job(ctx, ws, job, spec, mode, hist, stats, st);
}

int64_t plan_bytes(const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *triplets) {
    JobPlan P;
    if (int e = plan_job(job, spec, mode, P)) return e;
    if (triplets) *triplets = P.triplets;
    return (int64_t)P.bytes;
}

Workspace *new_workspace() { return new Workspace(); }
void delete_workspace(Workspace *w) { delete w; }
size_t workspace_held(const Workspace *w) { return w->region.bytes + w->queue.bytes + w->sqt_queue.bytes; }
void workspace_release(Workspace *w) {
    w->region.release();
    w->queue.release();
    w->sqt_queue.release();
}

int64_t footprint_bytes(const tpcf_job &job, const tpcf_binspec &spec, int mode, uint64_t *triplets) {
    JobPlan P;
    if (int e = plan_job(job, spec, mode, P)) return e;
    if (triplets) *triplets = P.triplets;
    // plus the term table and counters
    constexpr size_t kTerm = sizeof(SqtTerm) > sizeof(TripletTerm) ? sizeof(SqtTerm) : sizeof(TripletTerm);
    const size_t queue = align256(sizeof(TripletItem) * items_bound(P, job)) + align256(kTerm * 4) + 256;
    return (int64_t)(P.bytes + queue);
}

// ---------------------------------------------------------------------------
// fixed scratch of tpcf_count_noalloc
// ---------------------------------------------------------------------------

// Device memory of tpcf_count_noalloc: the job's region (staged points, pair
// tiles, tables, histograms) at the start, its work queue after it.  Each GPU
// holds its own copy of the array for the life of the process.  Host code
// never passes `scratch` itself (on the host it is only a stand-in for the
// symbol): it uses the device address hipGetSymbolAddress returns, kept per
// device in g_scratch.
constexpr size_t kScratchBytes = size_t(64) << 20;
// Room kept for the work queue (items, term table, counters) after the region.
constexpr size_t kScratchQueueMin = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[kMaxDevices];

// Device address of the scratch array on `device` (the current device),
// looked up once per device.
static int scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return set_error(TPCF_EDEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(TPCF_EDEVICE, "hipGetSymbolAddress of the fixed scratch failed on device %d: %s", device,
                             hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return TPCF_OK;
}

}  // namespace tpcf

using namespace tpcf;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *tpcf_version(void) { return "libtpcf 0.1 (gfx950)"; }

const char *tpcf_last_error(void) { return g_err.c_str(); }

int tpcf_ctx_create(int device, tpcf_ctx **out) {
    clear_error();
    if (!out) return set_error(TPCF_EINVAL, "out is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(TPCF_EINVAL, "device %d of %d", device, ndev);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(TPCF_EDEVICE, "libtpcf is built for gfx950, device %d is %s", device, prop.gcnArchName);
    auto *c = new tpcf_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->upload, hipStreamNonBlocking) != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return set_error(TPCF_EDEVICE, "stream creation failed");
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) c->budget = (uint64_t)(0.9 * (double)fr);
    c->ws.reset(new Workspace());
    // The no-allocation entry point's scratch address and events, set up
    // here so that the call itself creates nothing.
    c->ws_fixed.reset(new Workspace());
    if (int e = scratch_address(device, &c->scratch)) {
        tpcf_ctx_destroy(c);
        return e;
    }
    if (int e = c->ws_fixed->init()) {
        tpcf_ctx_destroy(c);
        return e;
    }
    // Kernels with an LDS histogram raise their dynamic LDS limit once, to
    // what their static LDS leaves of the 160 KiB.
    std::vector<const void *> lds_kernels{sqt_plain_kernel_ptr(kSqt), sqt_plain_kernel_ptr(kJbj),
                                          pair_hist_kernel_ptr(8, false), pair_hist_kernel_ptr(8, true)};
    for (bool el : {false, true}) {
        lds_kernels.push_back(sqt_fast_kernel_ptr(8, el));
        lds_kernels.push_back(sqt_fast_kernel_ptr(16, el));
        for (int w : {2, 4, 8, 16})
            for (int k : {1, 2, 4, 8})
                if (const void *fn = r123_triplet_kernel_ptr(w, k, el)) lds_kernels.push_back(fn);
        for (int w : {4, 8, 16})
            if (const void *fn = r123_bits_kernel_ptr(w, el)) lds_kernels.push_back(fn);
    }
    for (const void *fn : lds_kernels) {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, fn) != hipSuccess ||
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(160 * 1024 - fa.sharedSizeBytes)) != hipSuccess) {
            (void)hipGetLastError();
            tpcf_ctx_destroy(c);
            return set_error(TPCF_EDEVICE, "cannot raise the dynamic LDS limit of a kernel");
        }
    }
    if (const char *env = std::getenv("TPCF_R123_SHAPE")) {
        int w = 0, k = 0, p = 0;
        if (std::sscanf(env, "%d,%d,%d", &w, &k, &p) >= 2) {
            c->shape.waves = w;
            c->shape.copies = k;
            c->shape.per_cu = p;
        }
    }
    if (const char *env = std::getenv("TPCF_R123_FORM")) {
        if (!std::strcmp(env, "runs")) c->shape.form = kFormRuns;
        if (!std::strcmp(env, "bits")) c->shape.form = kFormBits;
        if (!std::strcmp(env, "split")) c->shape.form = kFormSplit;
    }
    if (const char *env = std::getenv("TPCF_R123_BITS_MAX")) c->shape.bits_max = (float)std::atof(env);
    if (const char *env = std::getenv("TPCF_R123_SPLIT")) {
        int bins = 0;
        float mean = 0.f;
        const int got = std::sscanf(env, "%d,%f", &bins, &mean);
        if (got >= 1) c->shape.split_bins = bins;
        if (got >= 2) c->shape.split_mean = mean;
    }
    *out = c;
    return TPCF_OK;
}

int tpcf_ctx_destroy(tpcf_ctx *ctx) {
    if (!ctx) return TPCF_OK;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->stopping = true;
    }
    ctx->cv.notify_all();
    for (auto &t : ctx->workers) t.join();
    (void)hipSetDevice(ctx->device);
    ctx->ws.reset();
    ctx->ws_fixed.reset();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->upload) (void)hipStreamDestroy(ctx->upload);
    delete ctx;
    return TPCF_OK;
}

int tpcf_ctx_set_memory_budget(tpcf_ctx *ctx, uint64_t bytes) {
    if (!ctx) return set_error(TPCF_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> g(ctx->mu);
    if (bytes) ctx->budget = bytes;
    ctx->cv.notify_all();
    return TPCF_OK;
}

int tpcf_catalog_create(tpcf_ctx *ctx, const double *xyz, int64_t n, tpcf_catalog **out) {
    clear_error();
    if (!ctx || !out || n < 0 || (n > 0 && !xyz)) return set_error(TPCF_EINVAL, "bad catalog arguments");
    double blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
    for (int d = 0; d < 3 && n > 0; ++d) {
        const double *v = xyz + (size_t)d * n;
        blo[d] = bhi[d] = v[0];
        for (int64_t i = 0; i < n; ++i) {
            if (!(std::fabs(v[i]) <= kMaxCoordinate))
                return set_error(TPCF_EINVAL,
                                 "coordinate %lld of the catalog is not finite or exceeds %g in magnitude",
                                 (long long)(d * n + i), kMaxCoordinate);
            blo[d] = std::min(blo[d], v[i]);
            bhi[d] = std::max(bhi[d], v[i]);
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    auto *c = new tpcf_catalog();
    c->ctx = ctx;
    c->n = n;
    c->h.assign(xyz, xyz + 3 * n);
    for (int d = 0; d < 3; ++d) {
        c->lo[d] = blo[d];
        c->hi[d] = bhi[d];
    }
    if (n > 0) {
        if (hipMalloc(&c->d, sizeof(double) * 3 * n) != hipSuccess) {
            delete c;
            return set_error(TPCF_ENOMEM, "device allocation for %lld points failed", (long long)n);
        }
        if (hipMemcpyAsync(c->d, xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->upload) != hipSuccess ||
            hipStreamSynchronize(ctx->upload) != hipSuccess) {
            (void)hipFree(c->d);
            delete c;
            return set_error(TPCF_EDEVICE, "catalog upload failed");
        }
    }
    *out = c;
    return TPCF_OK;
}

int tpcf_catalog_destroy(tpcf_catalog *cat) {
    if (!cat) return TPCF_OK;
    if (cat->d) (void)hipFree(cat->d);
    delete cat;
    return TPCF_OK;
}

int64_t tpcf_catalog_size(const tpcf_catalog *cat) { return cat ? cat->n : -1; }

int64_t tpcf_hist_cells(const tpcf_binspec *spec) {
    if (!spec || binspec_check(*spec)) return -1;
    return hist_cells(*spec);
}

int tpcf_axis_edges(const tpcf_axis *ax, double *edges) {
    clear_error();
    if (!ax || !edges) return set_error(TPCF_EINVAL, "NULL argument");
    return axis_edges(*ax, edges, kMaxPairBins);
}

int tpcf_binspec_check(const tpcf_binspec *spec) {
    clear_error();
    if (!spec) return set_error(TPCF_EINVAL, "spec is NULL");
    return binspec_check(*spec);
}

int64_t tpcf_job_footprint(const tpcf_job *job, const tpcf_binspec *spec, int mode) {
    if (!job || !spec) return set_error(TPCF_EINVAL, "NULL argument");
    return footprint_bytes(*job, *spec, mode);
}

int tpcf_ctx_memory(tpcf_ctx *ctx, uint64_t out[3]) {
    clear_error();
    if (!ctx || !out) return set_error(TPCF_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    out[0] = ctx->in_use;
    out[1] = ctx->peak;
    out[2] = ctx->budget;
    return TPCF_OK;
}

int tpcf_count(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode, uint64_t *hist,
               tpcf_stats *stats) {
    clear_error();
    if (!ctx || !job || !spec || !hist) return set_error(TPCF_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(ctx->call_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return run_job(ctx, *ctx->ws, *job, *spec, mode, hist, stats, ctx->stream);
}

int tpcf_count_noalloc(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode, uint64_t *hist,
                       tpcf_stats *stats) {
    clear_error();
    if (!ctx || !job || !spec || !hist) return set_error(TPCF_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(ctx->call_mu);
    uint64_t triplets = 0;
    const int64_t need = plan_bytes(*job, *spec, mode, &triplets);
    if (need < 0) return (int)need;
    const size_t region = ((size_t)need + 255) & ~size_t(255);
    if (region + kScratchQueueMin > kScratchBytes)
        return set_error(TPCF_ENOMEM,
                         "tpcf_count_noalloc: the job needs %lld bytes of device workspace; the fixed scratch holds "
                         "%zu (%zu kept for the work queue); use tpcf_count",
                         (long long)need, kScratchBytes, kScratchQueueMin);
    HIP_TRY(hipSetDevice(ctx->device));
    // Region first, the queue after it; one job uses only one of the two
    // queues (the (r1,r2,r3)/plain one or the (s,q,theta) one).
    Workspace &ws = *ctx->ws_fixed;
    char *base = reinterpret_cast<char *>(ctx->scratch);
    ws.region.view(base, region);
    ws.queue.view(base + region, kScratchBytes - region);
    ws.sqt_queue.view(base + region, kScratchBytes - region);
    return run_job(ctx, ws, *job, *spec, mode, hist, stats, ctx->stream);
}

int tpcf_count_batch(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs, int mode,
                     uint64_t *const *hist, tpcf_stats *stats) {
    clear_error();
    if (!ctx || !jobs || !specs || !hist || njobs < 0) return set_error(TPCF_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(ctx->call_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<tpcf_stats> per(njobs);
    tpcf_stats total{};
    if (int e = run_batch(ctx, *ctx->ws, jobs, specs, njobs, mode, hist, per.data(), &total, ctx->stream)) return e;
    if (stats) *stats = total;
    return TPCF_OK;
}

void tpcf_free(void *p) { free(p); }

}  // extern "C"

// ---------------------------------------------------------------------------
// prepared batches: stage once, autotune the launch shape, capture the
// counting sequence (histogram reset, pair stages, the persistent candidate
// kernel of the (s,q,theta) and JBJ jobs, the persistent triplet kernel of
// the (r1,r2,r3) jobs) in a HIP graph and replay it
// ---------------------------------------------------------------------------

struct tpcf_batch {
    tpcf_ctx *ctx = nullptr;
    int mode = TPCF_MODE_FAST;
    int njobs = 0;
    std::vector<tpcf_job> jobs;
    std::vector<JobPlan> plans;
    std::vector<size_t> off;
    DevBuf region, queue, sqt_queue;
    SqtLaunch sqt;  // the (s,q,theta) and JBJ jobs' launch
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    QueueDev qd;
    int shift = 2;
    TripletShape shape{8, 1, 0, 8, 0, -1, 7.0f, 5, 8.0f};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string tuning;  // autotuner log
    ~tpcf_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace tpcf {

// The counting sequence of a prepared batch: kernels only.
static int batch_compute(tpcf_batch *b, const TripletShape &sh, hipStream_t st) {
    char *base = reinterpret_cast<char *>(b->region.p);
    for (int j = 0; j < b->njobs; ++j) {
        const JobPlan &P = b->plans[j];
        char *jb = base + b->off[j];
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(jb + P.o_cnt), 16, 0u, st));
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(jb + P.o_hist), 2 * P.cells * P.nterms, 0u, st));
    }
    if (int e = sqt_enqueue(b->ctx, b->sqt, b->sqt_queue, st, nullptr, nullptr)) return e;
    for (int j = 0; j < b->njobs; ++j)
        if (b->plans[j].param == TPCF_R123 && b->plans[j].L.nblk)
            HIP_TRY(launch_pair_tiles(pair_job(b->plans[j], base + b->off[j], b->shift), b->mode == TPCF_MODE_FAST,
                                      st));
    if (!b->qd.nitems) return TPCF_OK;
    return launch_queue(b->ctx, b->qd, sh, st);
}

static int batch_time(tpcf_batch *b, const TripletShape &sh, float &ms) {
    HIP_TRY(hipEventRecord(b->ev[0], b->st));
    if (int e = batch_compute(b, sh, b->st)) return e;
    HIP_TRY(hipEventRecord(b->ev[1], b->st));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    return TPCF_OK;
}

// Candidate launch shapes: 2..16 waves per workgroup (128..1024 threads),
// the largest LDS histogram copy counts that fit, and the occupancy limit or
// half of it as workgroups per CU.
static int batch_autotune(tpcf_batch *b) {
    constexpr int kLdsBudget = 155 * 1024;
    const int n = b->qd.max_n;
    if (n > kMaxLdsR123Bins || !b->qd.nitems) return TPCF_OK;
    // The run-merging kernel's shapes, then the bit-mask kernel's, each
    // timed with its form forced, then the split form with the best shape of
    // each.
    const TripletShape base = b->shape;
    std::vector<TripletShape> cands;
    int kmax = 1;
    for (int k : {8, 4, 2, 1})
        if (r123_lds_bytes(n, k) <= kLdsBudget) {
            kmax = k;
            break;
        }
    for (int k : {kmax, kmax / 2})
        for (int w : {2, 4, 8, 16}) {
            if (k < 1 || !r123_triplet_kernel_ptr(w, k)) continue;
            int occ = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, r123_triplet_kernel_ptr(w, k), w * 64,
                                                             r123_lds_bytes(n, k)) != hipSuccess || occ < 1)
                continue;
            TripletShape c = base;
            c.waves = w;
            c.copies = k;
            c.form = kFormRuns;
            c.grid = b->ctx->num_cu * occ;
            cands.push_back(c);
            if (occ > 1) {
                c.grid = b->ctx->num_cu * std::max(1, occ / 2);
                cands.push_back(c);
            }
        }
    for (int w : {4, 8, 16}) {
        int occ = 0;
        if (!r123_bits_kernel_ptr(w) ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, r123_bits_kernel_ptr(w), w * 64,
                                                         r123_bits_lds_bytes(n)) != hipSuccess || occ < 1)
            continue;
        TripletShape c = base;
        c.bwaves = w;
        c.bgrid = b->ctx->num_cu * occ;
        c.form = kFormBits;
        cands.push_back(c);
    }
    float best = 1e30f, best_runs = 1e30f, best_bits = 1e30f;
    TripletShape runs = base, bits = base;
    char line[200];
    auto consider = [&](const TripletShape &c) -> int {
        float ms = 0.f;
        if (int e = batch_time(b, c, ms)) return e;
        if (c.form == kFormBits)
            std::snprintf(line, sizeof line, "bit-mask waves=%d wg=%d: %.3f ms\n", c.bwaves, c.bgrid, ms);
        else if (c.form == kFormSplit)
            std::snprintf(line, sizeof line,
                          "split (<= %d x bins: bit-mask) waves=%d copies=%d wg=%d, bit-mask waves=%d wg=%d: %.3f ms\n",
                          c.split_bins, c.waves, c.copies, c.grid, c.bwaves, c.bgrid, ms);
        else
            std::snprintf(line, sizeof line, "run-merging waves=%d copies=%d wg=%d: %.3f ms\n", c.waves, c.copies,
                          c.grid, ms);
        b->tuning += line;
        if (c.form == kFormRuns && ms < best_runs) {
            best_runs = ms;
            runs = c;
        }
        if (c.form == kFormBits && ms < best_bits) {
            best_bits = ms;
            bits = c;
        }
        if (ms < best) {
            best = ms;
            b->shape = c;
        }
        return TPCF_OK;
    };
    for (const auto &c : cands)
        if (int e = consider(c)) return e;
    // The split form with the best shape of each kernel.
    if (best_runs < 1e30f && best_bits < 1e30f && base.split_bins >= 0) {
        TripletShape c = runs;
        c.bwaves = bits.bwaves;
        c.bgrid = bits.bgrid;
        c.form = kFormSplit;
        if (int e = consider(c)) return e;
    }
    return TPCF_OK;
}

}  // namespace tpcf

extern "C" {

int tpcf_batch_create(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs, int mode, int flags,
                      tpcf_batch **out) {
    clear_error();
    if (!ctx || !jobs || !specs || njobs < 1 || !out) return set_error(TPCF_EINVAL, "bad batch arguments");
    if (mode != TPCF_MODE_FAST && mode != TPCF_MODE_FAST64 && mode != TPCF_MODE_CELLS)
        return set_error(TPCF_EINVAL, "batches run the FAST paths only");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<tpcf_batch> b(new tpcf_batch());
    b->ctx = ctx;
    b->mode = mode;
    b->njobs = njobs;
    b->jobs.assign(jobs, jobs + njobs);
    b->plans.resize(njobs);
    b->off.resize(njobs);
    size_t bytes = 0;
    int max_n = 1;
    for (int j = 0; j < njobs; ++j) {
        if (int e = plan_job(jobs[j], specs[j], mode, b->plans[j])) return e;
        b->off[j] = bytes;
        bytes = align256(bytes + b->plans[j].bytes);
        if (specs[j].param == TPCF_R123) max_n = std::max(max_n, b->plans[j].n);
    }
    b->shift = max_n <= kMaxLdsR123Bins ? 2 : 0;
    // CELLS: the plans hold the culling (Hilbert staging, reach, culled
    // items, all done once below); the replayed kernels are FAST's.
    if (mode == TPCF_MODE_CELLS) b->mode = TPCF_MODE_FAST;
    HIP_TRY(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    for (auto &e : b->ev) HIP_TRY(hipEventCreate(&e));
    if (int e = b->region.reserve(std::max<size_t>(bytes, 256))) return e;
    char *base = reinterpret_cast<char *>(b->region.p);
    std::vector<const JobPlan *> pp, sp;
    std::vector<const tpcf_job *> jj, sj;
    std::vector<char *> bb, sb;
    {
        std::vector<const tpcf_job *> js(njobs);
        std::vector<const JobPlan *> ps(njobs);
        std::vector<char *> bs(njobs);
        for (int j = 0; j < njobs; ++j) {
            js[j] = &b->jobs[j];
            ps[j] = &b->plans[j];
            bs[j] = base + b->off[j];
        }
        if (int e = stage_jobs(js.data(), ps.data(), bs.data(), njobs, b->st)) return e;
    }
    for (int j = 0; j < njobs; ++j) {
        if (!b->plans[j].L.nblk) continue;
        const bool r123 = b->plans[j].param == TPCF_R123;
        (r123 ? pp : sp).push_back(&b->plans[j]);
        (r123 ? jj : sj).push_back(&b->jobs[j]);
        (r123 ? bb : sb).push_back(base + b->off[j]);
    }
    Queue Q;
    build_queue(pp, jj, bb, Q);
    if (int e = upload_queue(b->queue, Q, b->st, b->qd)) return e;
    if (int e = sqt_plan(ctx, sp, sj, sb, b->mode, b->sqt)) return e;
    if (int e = sqt_upload(b->sqt, b->sqt_queue, b->st)) return e;
    HIP_TRY(hipStreamSynchronize(b->st));
    b->shape = triplet_shape(ctx, b->qd.max_n);
    if (flags & TPCF_BATCH_AUTOTUNE)
        if (int e = batch_autotune(b.get())) return e;
    if (flags & TPCF_BATCH_GRAPH) {
        HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
        const int rc = batch_compute(b.get(), b->shape, b->st);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ce != hipSuccess) return set_error(TPCF_EDEVICE, "graph capture failed: %s", hipGetErrorString(ce));
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    *out = b.release();
    return TPCF_OK;
}

int tpcf_batch_run(tpcf_batch *b, uint64_t *const *hist, tpcf_stats *stats) {
    clear_error();
    if (!b || !hist) return set_error(TPCF_EINVAL, "NULL argument");
    HIP_TRY(hipSetDevice(b->ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(b->ev[0], b->st));
    if (b->exec) HIP_TRY(hipGraphLaunch(b->exec, b->st));
    else if (int e = batch_compute(b, b->shape, b->st)) return e;
    HIP_TRY(hipEventRecord(b->ev[1], b->st));
    char *base = reinterpret_cast<char *>(b->region.p);
    std::vector<uint64_t> rech(b->njobs, 0);
    for (int j = 0; j < b->njobs; ++j) {
        const JobPlan &P = b->plans[j];
        HIP_TRY(hipMemcpyAsync(hist[j], base + b->off[j] + P.o_hist, sizeof(uint64_t) * P.cells * P.nterms,
                               hipMemcpyDeviceToHost, b->st));
        HIP_TRY(hipMemcpyAsync(&rech[j], base + b->off[j] + P.o_cnt, 8, hipMemcpyDeviceToHost, b->st));
    }
    HIP_TRY(hipStreamSynchronize(b->st));
    if (stats) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        tpcf_stats s{};
        for (int j = 0; j < b->njobs; ++j) {
            s.triplets += b->plans[j].triplets;
            int64_t nu = 0;
            for (int q = 0; q < b->plans[j].L.nseg; ++q) nu += b->plans[j].L.seg_n[q];
            s.pairs += (uint64_t)(nu * (nu - 1) / 2);
            s.pairs_rechecked += b->mode == TPCF_MODE_FAST ? rech[j] : 0;
        }
        s.device_ms = ms;
        s.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (int e = read_ctl(b->qd.nitems && b->qd.max_n <= kMaxLdsR123Bins ? b->qd.counter : nullptr, &s, b->st))
            return e;
        *stats = s;
    }
    return TPCF_OK;
}

const char *tpcf_batch_info(const tpcf_batch *b) {
    if (!b) return "";
    static thread_local std::string info;
    char head[160];
    std::snprintf(head, sizeof head,
                  "shape form=%s waves=%d copies=%d wg=%d bits_waves=%d bits_wg=%d graph=%d items=%d sqt_items=%d\n",
                  b->shape.form == kFormBits    ? "bit-mask"
                  : b->shape.form == kFormRuns  ? "run-merging"
                  : b->shape.form == kFormSplit ? "split"
                                                : "auto",
                  b->shape.waves, b->shape.copies, b->shape.grid, b->shape.bwaves, b->shape.bgrid, b->exec ? 1 : 0,
                  b->qd.nitems, (int)b->sqt.items.size());
    info = head + b->tuning;
    return info.c_str();
}

int tpcf_batch_destroy(tpcf_batch *b) {
    if (!b) return TPCF_OK;
    (void)hipSetDevice(b->ctx->device);
    delete b;
    return TPCF_OK;
}

}  // extern "C"
