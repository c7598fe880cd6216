// survey.cpp -- spatial voxelization of a survey into cell jobs (tpcf_survey_*).
//
// The reference splits a survey into voxels by catalog index chunks
// (src/3pcf_C_version_triangle_parametrization.cc:52-80) and runs one process
// per voxel and term (examples/build_voxel_submission_scripts.py:9-41,
// examples/run_CPU_calculation_each_voxels.csh:60-67), whose outputs
// examples/sum_voxels.py adds; index chunks share no triplets across voxels,
// so that sum is not the survey's histogram.  Its other sketch,
// examples/test_predivision_of_distances.py:11-35 (define_ranges) and
// :123-205, divides space into cells of side at least the largest
// separation and pairs each cell with its neighbour cells.  This file builds
// that for triplets, exactly:
//
//   * cells: per axis n = floor(extent / (maxsep (1 + 1e-9))) (at least 1)
//     cells of width extent / n >= maxsep, over the bounding box of data and
//     randoms; a point goes to cell min(n - 1, floor((x - lo) / width)).
//   * every side of a counted triangle is shorter than the spec's reach
//     (r123: the last edge; (s,q,theta): s_hi (1 + max(q_hi, 1)); JBJ:
//     s_hi (1 / q_lo + v_hi)), and maxsep >= reach, so the three points of
//     a counted triangle lie in pairwise adjacent cells (Chebyshev distance
//     at most 1 in cell indices).
//   * a triangle belongs to the lowest linear index c among its points'
//     cells; its other points then lie in c (home, H) or in the 13 cells
//     adjacent to c with a higher index (forward neighbours, U).  The jobs of
//     c count the triangles with at least one point in H and the rest in
//     H u U, which are exactly the triangles c owns plus triangles with a
//     side of at least one cell width (counted by no histogram).  So the
//     cell jobs partition the counted triangles: their sum is the survey's
//     histogram, bit for bit.
//   * per cell and catalog term the terms over the segments (H_D, U_D, H_R,
//     U_R), at most three segments per job:
//       DDD: (H_D, U_D): (0,0,0), (0,0,1), (0,1,1)
//       RRR: (H_R, U_R): the same three
//       DDR: (H_D, U_D, H_R): (0,0,2), (0,1,2), (1,1,2); (H_D, U_D, U_R): (0,0,2), (0,1,2)
//       DRR: (H_R, U_R, H_D): (0,0,2), (0,1,2), (1,1,2); (H_R, U_R, U_D): (0,0,2), (0,1,2)
//     (a DDR triangle has 1 or 2 home data points, or 0 home data points and
//     its random point at home).
// The device holds four catalogs: data and randoms sorted by cell (H ranges)
// and the concatenated forward-neighbour copies of each (U ranges).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

#include "tpcf_host.h"

struct tpcf_survey {
    tpcf_ctx *ctx = nullptr;
    double maxsep = 0;
    int nc[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, width[3] = {1, 1, 1};
    int64_t ncells = 0;
    int64_t npts[2] = {0, 0};
    // 0: data by cell, 1: randoms by cell, 2: data forward copies, 3: randoms forward copies
    tpcf_catalog *cat[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<int64_t> off[4];  // per catalog, ncells + 1 offsets
    std::vector<tpcf_job> jobs;
    std::vector<int32_t> job_cell, job_term;  // cell and catalog term (0 DDD, 1 DDR, 2 DRR, 3 RRR) of each job
};

namespace tpcf {

// Longest side a counted triangle of spec can have (with the margin of the
// kernels' own cuts), or an error for specs with no such bound.
int spec_reach(const tpcf_binspec &spec, double *reach) {
    if (int e = binspec_check(spec)) return e;
    if (spec.param == TPCF_R123) {
        *reach = spec.ax[0].hi;
    } else if (spec.param == TPCF_SQT) {
        *reach = spec.ax[0].hi * (1.0 + std::max(spec.ax[1].hi, 1.0)) * (1.0 + 1e-6);
    } else {
        const double qlo = spec.ax[1].lo, vhi = std::max(spec.ax[2].hi, 0.0);
        if (!(qlo > 0))
            return set_error(TPCF_EINVAL, "a JBJ spec with q_lo = 0 puts no bound on the sides of a counted "
                                          "triangle, so it cannot be split into cells");
        *reach = spec.ax[0].hi * (1.0 / qlo + vhi) * (1.0 + 1e-6);
    }
    return TPCF_OK;
}

}  // namespace tpcf

using namespace tpcf;

namespace {

// Forward neighbour offsets: the 13 of the 26 with a higher linear index.
std::vector<std::array<int, 3>> forward_offsets() {
    std::vector<std::array<int, 3>> v;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c)
                if (a > 0 || (a == 0 && (b > 0 || (b == 0 && c > 0)))) v.push_back({a, b, c});
    return v;
}

void destroy_survey(tpcf_survey *s) {
    if (!s) return;
    for (auto *c : s->cat)
        if (c) tpcf_catalog_destroy(c);
    delete s;
}

}  // namespace

extern "C" {

int tpcf_spec_reach(const tpcf_binspec *spec, double *reach) {
    clear_error();
    if (!spec || !reach) return set_error(TPCF_EINVAL, "NULL argument");
    return spec_reach(*spec, reach);
}

int tpcf_survey_create(tpcf_ctx *ctx, const double *d_xyz, int64_t nd, const double *r_xyz, int64_t nr,
                       double maxsep, tpcf_survey **out) {
    clear_error();
    if (!ctx || !out || nd < 0 || nr < 0 || (nd && !d_xyz) || (nr && !r_xyz))
        return set_error(TPCF_EINVAL, "bad survey arguments");
    if (!(maxsep > 0) || !std::isfinite(maxsep)) return set_error(TPCF_EINVAL, "maxsep must be positive and finite");
    const double *src[2] = {d_xyz, r_xyz};
    const int64_t n[2] = {nd, nr};
    auto *s = new tpcf_survey();
    s->ctx = ctx;
    s->maxsep = maxsep;
    s->npts[0] = nd;
    s->npts[1] = nr;
    // bounding box of both catalogs
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int k = 0; k < 2; ++k)
        for (int d = 0; d < 3; ++d)
            for (int64_t i = 0; i < n[k]; ++i) {
                const double v = src[k][(size_t)d * n[k] + i];
                if (!(std::fabs(v) <= kMaxCoordinate)) {
                    destroy_survey(s);
                    return set_error(TPCF_EINVAL, "coordinate not finite or beyond %g", kMaxCoordinate);
                }
                lo[d] = std::min(lo[d], v);
                hi[d] = std::max(hi[d], v);
            }
    int64_t ncells = 1;
    for (int d = 0; d < 3; ++d) {
        const double ext = nd + nr ? hi[d] - lo[d] : 0.0;
        const double cells = std::floor(ext / (maxsep * (1.0 + 1e-9)));
        s->nc[d] = (int)std::max(1.0, std::min(cells, 1024.0));
        s->lo[d] = nd + nr ? lo[d] : 0.0;
        s->width[d] = ext > 0 ? ext / s->nc[d] : maxsep;
        ncells *= s->nc[d];
    }
    if (ncells > (int64_t)1 << 22) {
        destroy_survey(s);
        return set_error(TPCF_EINVAL, "%lld cells: maxsep is too small for this survey", (long long)ncells);
    }
    s->ncells = ncells;
    const int ny = s->nc[1], nz = s->nc[2];
    auto cell_of = [&](const double *xyz, int64_t m, int64_t i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            const double t = std::floor((xyz[(size_t)d * m + i] - s->lo[d]) / s->width[d]);
            c[d] = (int)std::min<double>(s->nc[d] - 1, std::max(0.0, t));
        }
        return ((int64_t)c[0] * ny + c[1]) * nz + c[2];
    };
    // points sorted by cell (counting sort, stable), per catalog
    std::vector<double> sorted[2], fwd[2];
    const auto offs = forward_offsets();
    for (int k = 0; k < 2; ++k) {
        std::vector<int64_t> cell(n[k]);
        std::vector<int64_t> &o = s->off[k];
        o.assign(ncells + 1, 0);
        for (int64_t i = 0; i < n[k]; ++i) ++o[(cell[i] = cell_of(src[k], n[k], i)) + 1];
        for (int64_t c = 0; c < ncells; ++c) o[c + 1] += o[c];
        std::vector<int64_t> at(o.begin(), o.end() - 1);
        sorted[k].resize((size_t)3 * n[k]);
        for (int64_t i = 0; i < n[k]; ++i) {
            const int64_t p = at[cell[i]]++;
            for (int d = 0; d < 3; ++d) sorted[k][(size_t)d * n[k] + p] = src[k][(size_t)d * n[k] + i];
        }
        // forward-neighbour copies: cell c's block holds the points of its
        // forward neighbours, cells in increasing index
        std::vector<int64_t> &u = s->off[2 + k];
        u.assign(ncells + 1, 0);
        auto neighbours = [&](int64_t c, auto &&visit) {
            const int ci = (int)(c / ((int64_t)ny * nz)), cj = (int)((c / nz) % ny), ck = (int)(c % nz);
            for (const auto &f : offs) {
                const int a = ci + f[0], b = cj + f[1], e = ck + f[2];
                if (a < 0 || a >= s->nc[0] || b < 0 || b >= ny || e < 0 || e >= nz) continue;
                visit(((int64_t)a * ny + b) * nz + e);
            }
        };
        for (int64_t c = 0; c < ncells; ++c) {
            int64_t m = 0;
            neighbours(c, [&](int64_t q) { m += o[q + 1] - o[q]; });
            u[c + 1] = u[c] + m;
        }
        const int64_t total = u[ncells];
        fwd[k].resize((size_t)3 * total);
        for (int64_t c = 0; c < ncells; ++c) {
            int64_t p = u[c];
            neighbours(c, [&](int64_t q) {
                for (int64_t i = o[q]; i < o[q + 1]; ++i, ++p)
                    for (int d = 0; d < 3; ++d) fwd[k][(size_t)d * total + p] = sorted[k][(size_t)d * n[k] + i];
            });
        }
    }
    // device catalogs
    const std::vector<double> *host[4] = {&sorted[0], &sorted[1], &fwd[0], &fwd[1]};
    for (int k = 0; k < 4; ++k) {
        const int64_t m = (int64_t)host[k]->size() / 3;
        if (int e = tpcf_catalog_create(ctx, host[k]->data(), m, &s->cat[k])) {
            destroy_survey(s);
            return e;
        }
    }
    // jobs: per cell, per catalog term (see the header comment)
    enum { HD = 0, HR = 1, UD = 2, UR = 3 };
    for (int64_t c = 0; c < ncells; ++c) {
        auto add = [&](int term, std::initializer_list<int> segs, std::initializer_list<std::array<int, 3>> terms) {
            tpcf_job j{};
            j.nseg = (int)segs.size();
            int q = 0;
            for (int k : segs) {
                j.seg[q] = s->cat[k];
                j.seg_lo[q] = s->off[k][c];
                j.seg_hi[q] = s->off[k][c + 1];
                ++q;
            }
            for (const auto &t : terms) {
                // skip terms with an empty segment: they hold no triangle
                bool empty = false;
                for (int x : t) empty |= j.seg_hi[x] == j.seg_lo[x];
                if (empty) continue;
                std::memcpy(j.term[j.nterms++], t.data(), sizeof(int) * 3);
            }
            if (!j.nterms) return;
            s->jobs.push_back(j);
            s->job_cell.push_back((int32_t)c);
            s->job_term.push_back(term);
        };
        add(0, {HD, UD}, {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}});
        add(1, {HD, UD, HR}, {{0, 0, 2}, {0, 1, 2}, {1, 1, 2}});
        add(1, {HD, UD, UR}, {{0, 0, 2}, {0, 1, 2}});
        add(2, {HR, UR, HD}, {{0, 0, 2}, {0, 1, 2}, {1, 1, 2}});
        add(2, {HR, UR, UD}, {{0, 0, 2}, {0, 1, 2}});
        add(3, {HR, UR}, {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}});
    }
    *out = s;
    return TPCF_OK;
}

int64_t tpcf_survey_cells(const tpcf_survey *s, int dims[3]) {
    clear_error();
    if (!s) return set_error(TPCF_EINVAL, "survey is NULL");
    if (dims)
        for (int d = 0; d < 3; ++d) dims[d] = s->nc[d];
    return s->ncells;
}

int64_t tpcf_survey_jobs(const tpcf_survey *s) {
    clear_error();
    if (!s) return set_error(TPCF_EINVAL, "survey is NULL");
    return (int64_t)s->jobs.size();
}

int tpcf_survey_count(tpcf_ctx *ctx, const tpcf_survey *s, const tpcf_binspec *spec, int mode, uint64_t *hist,
                      uint64_t *per_cell, tpcf_stats *stats) {
    clear_error();
    if (!ctx || !s || !spec || !hist) return set_error(TPCF_EINVAL, "NULL argument");
    if (s->ctx != ctx) return set_error(TPCF_EINVAL, "the survey belongs to another context");
    double reach = 0;
    if (int e = spec_reach(*spec, &reach)) return e;
    if (reach > s->maxsep)
        return set_error(TPCF_EINVAL, "the spec's triangles reach %g, beyond the survey's cell size bound %g", reach,
                         s->maxsep);
    const int64_t cells = hist_cells(*spec);
    const size_t nj = s->jobs.size();
    std::vector<std::vector<uint64_t>> out(nj);
    std::vector<uint64_t *> ptr(nj);
    for (size_t j = 0; j < nj; ++j) {
        out[j].assign((size_t)s->jobs[j].nterms * cells, 0);
        ptr[j] = out[j].data();
    }
    std::memset(hist, 0, sizeof(uint64_t) * 4 * cells);
    if (per_cell) std::memset(per_cell, 0, sizeof(uint64_t) * 4 * cells * s->ncells);
    tpcf_stats total{};
    if (nj) {
        if (mode == TPCF_MODE_PLAIN) {
            for (size_t j = 0; j < nj; ++j) {
                tpcf_stats st{};
                if (int e = tpcf_count(ctx, &s->jobs[j], spec, mode, ptr[j], &st)) return e;
                total.triplets += st.triplets;
                total.device_ms += st.device_ms;
                total.wall_ms += st.wall_ms;
            }
        } else {
            std::vector<tpcf_binspec> specs(nj, *spec);
            if (int e = tpcf_count_batch(ctx, s->jobs.data(), specs.data(), (int)nj, mode, ptr.data(), &total))
                return e;
        }
    }
    for (size_t j = 0; j < nj; ++j) {
        const int t = s->job_term[j];
        uint64_t *h = hist + (size_t)t * cells;
        uint64_t *pc = per_cell ? per_cell + ((size_t)s->job_cell[j] * 4 + t) * cells : nullptr;
        for (int k = 0; k < s->jobs[j].nterms; ++k)
            for (int64_t x = 0; x < cells; ++x) {
                const uint64_t v = out[j][(size_t)k * cells + x];
                h[x] += v;
                if (pc) pc[x] += v;
            }
    }
    if (stats) *stats = total;
    return TPCF_OK;
}

int tpcf_survey_destroy(tpcf_survey *s) {
    destroy_survey(s);
    return TPCF_OK;
}

}  // extern "C"
