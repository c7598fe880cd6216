// plan_check.cpp -- CPU checker of libtpcf's job planning, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (3pcf_amd/csrc/Makefile,
// target "san") and run by tests/test_host_sanitized.py.  No GPU is used.
//
// It includes the library's tpcf_api.cpp to reach the planning functions
// and checks them against brute force on seeded random inputs:
//   * term_items (unculled): the items' exact triplet counts add up to the
//     closed-form count of the term, no block triple is covered twice, every
//     item is inside its segments and at most kMaxItemBlocks long, and the
//     items come longest first;
//   * block_reach: every bit equals the box-gap test of its block pair, and
//     the first/last non-empty words of every row are right;
//   * term_items (culled): every block triple whose three pairs are within
//     reach is covered exactly once, none twice;
//   * plan_job: region offsets ascending and 256-byte aligned, per-block
//     point counts summing to the segment sizes, the triplet total equal to
//     the closed forms, for every parametrization and mode.
// Exit status 0 when every check holds; a failed check prints what failed.
#include "tpcf_api.cpp"

#include <random>
#include <set>
#include <tuple>

namespace {

int g_fail = 0;
uint64_t g_terms = 0, g_items = 0, g_reach_bits = 0, g_culled_triples = 0, g_plans = 0;

#define CHECK(cond, ...)                      \
    do {                                      \
        if (!(cond)) {                        \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__); \
            std::fprintf(stderr, "\n");       \
            if (++g_fail > 20) std::exit(1);  \
        }                                     \
    } while (0)

using tpcf::Layout;
using tpcf::TripletItem;

uint64_t c2(uint64_t n) { return n < 2 ? 0 : n * (n - 1) / 2; }
uint64_t c3(uint64_t n) { return n < 3 ? 0 : n * (n - 1) / 2 * (n - 2) / 3; }

// Points of block X of the union layout.
uint64_t block_points(const Layout &L, int64_t X) {
    for (int s = 0; s < L.nseg; ++s)
        if (X >= L.seg_blk0[s] && X < L.seg_blk0[s] + L.seg_nblk[s])
            return (uint64_t)std::min<int64_t>(tpcf::kBlock, L.seg_n[s] - (X - L.seg_blk0[s]) * tpcf::kBlock);
    return 0;
}

// Unique triplets (i in A, j in B, k in C) of term tr: i < j when the first
// two labels are equal and j < k when the last two are.
uint64_t triple_triplets(const Layout &L, const int *tr, int64_t A, int64_t B, int64_t C) {
    const uint64_t a = block_points(L, A), b = block_points(L, B), c = block_points(L, C);
    const bool ab = tr[0] == tr[1] && A == B, bc = tr[1] == tr[2] && B == C;
    if (ab && bc) return c3(a);
    if (ab) return c2(a) * c;
    if (bc) return a * c2(b);
    return a * b * c;
}

tpcf_job make_job(std::vector<tpcf_catalog> &cats, const std::vector<int64_t> &sizes) {
    tpcf_job job{};
    job.nseg = (int)sizes.size();
    for (int s = 0; s < job.nseg; ++s) {
        job.seg[s] = &cats[s];
        job.seg_lo[s] = 0;
        job.seg_hi[s] = sizes[s];
    }
    return job;
}

std::vector<std::array<int, 3>> all_terms(int nseg) {
    std::vector<std::array<int, 3>> t;
    for (int a = 0; a < nseg; ++a)
        for (int b = a; b < nseg; ++b)
            for (int c = b; c < nseg; ++c) t.push_back({a, b, c});
    return t;
}

void check_items_unculled(std::mt19937_64 &g) {
    const int64_t pick[] = {0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 200, 700};
    for (int round = 0; round < 300; ++round) {
        const int nseg = 1 + (int)(g() % 3);
        tpcf_job job{};
        job.nseg = nseg;
        std::vector<int64_t> sizes(nseg);
        for (auto &s : sizes) s = pick[g() % 12];
        // a segment is an index range; the layout only needs the sizes
        for (int s = 0; s < nseg; ++s) {
            job.seg_lo[s] = 0;
            job.seg_hi[s] = sizes[s];
        }
        const Layout L = tpcf::make_layout(job);
        for (const auto &t : all_terms(nseg)) {
            std::vector<TripletItem> items;
            tpcf::term_items(L, t.data(), 3, items);
            uint64_t sum = 0;
            std::set<std::tuple<int, int, int>> seen;
            for (size_t q = 0; q < items.size(); ++q) {
                const TripletItem &it = items[q];
                CHECK(it.term == 3, "term slot %d", it.term);
                CHECK(it.c1 > it.c0 && it.c1 - it.c0 <= tpcf::kMaxItemBlocks, "item length %d", it.c1 - it.c0);
                if (q) CHECK(items[q - 1].c1 - items[q - 1].c0 >= it.c1 - it.c0, "items not longest first");
                CHECK(it.A >= L.seg_blk0[t[0]] && it.A < L.seg_blk0[t[0]] + L.seg_nblk[t[0]], "A outside segment");
                CHECK(it.B >= L.seg_blk0[t[1]] && it.B < L.seg_blk0[t[1]] + L.seg_nblk[t[1]], "B outside segment");
                CHECK(it.c0 >= L.seg_blk0[t[2]] && it.c1 <= L.seg_blk0[t[2]] + L.seg_nblk[t[2]], "C outside segment");
                if (t[0] == t[1]) CHECK(it.A <= it.B, "A > B in one segment");
                if (t[1] == t[2]) CHECK(it.B <= it.c0, "B > C in one segment");
                for (int C = it.c0; C < it.c1; ++C) {
                    CHECK(seen.insert({it.A, it.B, C}).second, "block triple (%d,%d,%d) twice", it.A, it.B, C);
                    sum += triple_triplets(L, t.data(), it.A, it.B, C);
                }
            }
            ++g_terms;
            g_items += items.size();
            CHECK(sum == tpcf::term_triplets(L, t.data()), "term (%d,%d,%d) sizes %lld/%lld/%lld: items %llu, closed "
                  "form %llu", t[0], t[1], t[2], (long long)L.seg_n[0], (long long)L.seg_n[1], (long long)L.seg_n[2],
                  (unsigned long long)sum, (unsigned long long)tpcf::term_triplets(L, t.data()));
        }
    }
}

// Random block boxes (lo xyz, hi xyz) of nblk blocks, in a rough spatial order.
std::vector<double> random_boxes(std::mt19937_64 &g, int64_t nblk, double size) {
    std::uniform_real_distribution<double> u(0.0, size), w(0.0, size / 8);
    std::vector<double> b((size_t)6 * nblk);
    for (int64_t X = 0; X < nblk; ++X) {
        double *p = b.data() + 6 * X;
        const double base = size * (double)X / (double)nblk;
        for (int c = 0; c < 3; ++c) {
            p[c] = c == 0 ? base : u(g);
            p[3 + c] = p[c] + w(g);
        }
    }
    return b;
}

void check_reach_and_culled_items(std::mt19937_64 &g) {
    for (int round = 0; round < 40; ++round) {
        const int nseg = 1 + (int)(g() % 2);
        tpcf_job job{};
        job.nseg = nseg;
        for (int s = 0; s < nseg; ++s) {
            job.seg_lo[s] = 0;
            job.seg_hi[s] = 64 * (int64_t)(1 + g() % 60) - (int64_t)(g() % 64);
        }
        const Layout L = tpcf::make_layout(job);
        const auto boxes = random_boxes(g, L.nblk, 100.0);
        const double cut = 5.0 + (double)(g() % 40);
        const tpcf::Reach R = tpcf::block_reach(boxes, L.nblk, cut * cut);
        for (int64_t X = 0; X < L.nblk; ++X) {
            int32_t f = (int32_t)R.words, l = -1;
            for (int64_t Y = 0; Y < L.nblk; ++Y) {
                const bool want = tpcf::block_gap2(boxes.data() + 6 * X, boxes.data() + 6 * Y) < cut * cut;
                CHECK(R.has(X, Y) == want, "reach bit (%lld,%lld)", (long long)X, (long long)Y);
                ++g_reach_bits;
                if (want) {
                    f = std::min<int32_t>(f, (int32_t)(Y >> 6));
                    l = (int32_t)(Y >> 6);
                }
            }
            if (l >= 0) CHECK(R.first[X] == f && R.last[X] == l, "row %lld first/last words", (long long)X);
        }
        for (const auto &t : all_terms(nseg)) {
            std::vector<TripletItem> items;
            tpcf::term_items(L, t.data(), 0, items, &R, tpcf::kMergeGapBlocks);
            std::set<std::tuple<int, int, int>> seen;
            for (const TripletItem &it : items) {
                CHECK(it.c1 - it.c0 <= tpcf::kMaxItemBlocks, "culled item length");
                for (int C = it.c0; C < it.c1; ++C)
                    CHECK(seen.insert({it.A, it.B, C}).second, "culled block triple (%d,%d,%d) twice", it.A, it.B, C);
            }
            const int64_t a0 = L.seg_blk0[t[0]], b0 = L.seg_blk0[t[1]], c0 = L.seg_blk0[t[2]];
            for (int64_t A = a0; A < a0 + L.seg_nblk[t[0]]; ++A)
                for (int64_t B = t[0] == t[1] ? A : b0; B < b0 + L.seg_nblk[t[1]]; ++B)
                    for (int64_t C = t[1] == t[2] ? B : c0; C < c0 + L.seg_nblk[t[2]]; ++C)
                        if (R.has(A, B) && R.has(A, C) && R.has(B, C) && ++g_culled_triples)
                            CHECK(seen.count({(int)A, (int)B, (int)C}) == 1, "in-reach triple (%lld,%lld,%lld) not "
                                  "covered", (long long)A, (long long)B, (long long)C);
        }
    }
}

void check_plans(std::mt19937_64 &g) {
    tpcf_ctx *ctx = new tpcf_ctx();
    std::uniform_real_distribution<double> u(-50.0, 150.0);
    for (int round = 0; round < 40; ++round) {
        const int nseg = 1 + (int)(g() % 3);
        std::vector<int64_t> sizes(nseg);
        std::vector<tpcf_catalog> cats(nseg);
        for (int s = 0; s < nseg; ++s) {
            sizes[s] = (int64_t)(g() % 400);
            tpcf_catalog &c = cats[s];
            c.ctx = ctx;
            c.n = sizes[s] + (int64_t)(g() % 5);
            c.h.resize((size_t)3 * c.n);
            for (double &v : c.h) v = u(g);
            for (int d = 0; d < 3 && c.n; ++d) {
                c.lo[d] = *std::min_element(c.h.begin() + d * c.n, c.h.begin() + (d + 1) * c.n);
                c.hi[d] = *std::max_element(c.h.begin() + d * c.n, c.h.begin() + (d + 1) * c.n);
            }
        }
        tpcf_job job = make_job(cats, sizes);
        const auto terms = all_terms(nseg);
        job.nterms = std::min<int>(4, (int)terms.size());
        for (int t = 0; t < job.nterms; ++t)
            for (int k = 0; k < 3; ++k) job.term[t][k] = terms[(size_t)(t + round) % terms.size()][k];
        tpcf_binspec spec{};
        spec.param = (int)(g() % 3);
        spec.recheck_rel = 1e-5;
        const int scale = (int)(g() % 2);
        spec.ax[0] = tpcf_axis{1 + (int)(g() % 40), scale ? 1.0 : 0.0, 60.0, scale};
        spec.ax[1] = tpcf_axis{1 + (int)(g() % 20), spec.param == TPCF_JBJ ? 0.1 : 1.0, 3.0, TPCF_LIN};
        spec.ax[2] = tpcf_axis{1 + (int)(g() % 20), 0.0, 1.0, TPCF_LIN};
        for (int mode : {TPCF_MODE_FAST, TPCF_MODE_PLAIN, TPCF_MODE_FAST64, TPCF_MODE_CELLS}) {
            tpcf::JobPlan P;
            const int rc = tpcf::plan_job(job, spec, mode, P);
            CHECK(rc == TPCF_OK, "plan_job rc %d: %s", rc, tpcf_last_error());
            ++g_plans;
            if (rc) continue;
            const size_t offs[] = {P.o_u64, P.o_u32, P.o_blk, P.o_T64, P.o_T32, P.o_Eq, P.o_Et, P.o_cnt,
                                   P.o_hist, P.o_tiles, P.o_any, P.o_dists, P.o_masks};
            for (size_t q = 0; q < sizeof offs / sizeof offs[0]; ++q) {
                CHECK(offs[q] % 256 == 0, "offset %zu not aligned", q);
                if (q) CHECK(offs[q] >= offs[q - 1], "offsets not ascending at %zu", q);
                CHECK(offs[q] <= P.bytes, "offset %zu past the region", q);
            }
            uint64_t tri = 0, pts = 0;
            for (int t = 0; t < job.nterms; ++t) tri += tpcf::term_triplets(P.L, job.term[t]);
            for (uint8_t c : P.blk_count) pts += c;
            uint64_t want_pts = 0;
            for (int64_t s : sizes) want_pts += (uint64_t)s;
            CHECK(P.triplets == tri, "plan triplets");
            CHECK(pts == want_pts, "block counts %llu vs %llu points", (unsigned long long)pts,
                  (unsigned long long)want_pts);
            CHECK(P.bytes >= P.o_hist + sizeof(uint64_t) * (size_t)P.cells * P.nterms, "histograms past the region");
        }
    }
    // Malformed jobs are refused with a message, not planned.
    tpcf_catalog c{};
    c.ctx = ctx;
    c.n = 10;
    c.h.assign(30, 1.0);
    tpcf_job bad{};
    bad.nseg = 1;
    bad.seg[0] = &c;
    bad.seg_hi[0] = 11;  // past the catalog
    bad.nterms = 1;
    tpcf_binspec spec{};
    spec.param = TPCF_R123;
    spec.ax[0] = tpcf_axis{10, 0.0, 10.0, TPCF_LIN};
    tpcf::JobPlan P;
    CHECK(tpcf::plan_job(bad, spec, TPCF_MODE_FAST, P) == TPCF_EINVAL, "range past the catalog accepted");
    bad.seg_hi[0] = 10;
    bad.term[0][0] = 1;  // label beyond the segments
    CHECK(tpcf::plan_job(bad, spec, TPCF_MODE_FAST, P) == TPCF_EINVAL, "bad term label accepted");
    bad.term[0][0] = 0;
    spec.ax[0].n = 0;
    CHECK(tpcf::plan_job(bad, spec, TPCF_MODE_FAST, P) == TPCF_EINVAL, "zero bins accepted");
    delete ctx;
}

}  // namespace

int main() {
    std::mt19937_64 g(20261204);
    check_items_unculled(g);
    check_reach_and_culled_items(g);
    check_plans(g);
    if (g_fail) {
        std::fprintf(stderr, "plan_check: %d checks failed\n", g_fail);
        return 1;
    }
    std::printf("plan_check: all checks passed (%llu terms, %llu items, %llu reach bits, %llu culled triples, "
                "%llu plans)\n", (unsigned long long)g_terms, (unsigned long long)g_items,
                (unsigned long long)g_reach_bits, (unsigned long long)g_culled_triples, (unsigned long long)g_plans);
    return 0;
}
