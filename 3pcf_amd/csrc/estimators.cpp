// estimators.cpp -- correlation-function estimators over the histograms
// (host arithmetic in FP64, no device work).
//
// Reference: examples/plot_MICE_3pcf.py (normalizations :96-107, the
// estimator :143-156, bin centres :33-40, the triangle sides :43-51, the
// nearest-neighbour 2PCF lookup :9-30 and the reduced amplitude :196-221) and
// examples/plot_MICE_2pcf.py (normalizations :39-50, the estimator :63-69 and
// the table written at :127-131).  The operations and their order follow the
// scripts' numpy expressions, so the FP64 results equal theirs on the same
// inputs.
#include <cmath>
#include <cstdint>

#include "tpcf_host.h"

using namespace tpcf;

namespace {

// float(a * b * c) / d with the product exact (Python integers), rounded
// once to FP64.
double exact_norm(int64_t a, int64_t b, int64_t c, double d) {
    const __int128 p = (__int128)a * (__int128)b * (__int128)c;
    return (double)p / d;
}

// NaN and +inf become 0 (plot_MICE_3pcf.py:155-156); -inf stays.
double clean_ratio(double v) { return (v != v || v == HUGE_VAL) ? 0.0 : v; }

}  // namespace

extern "C" {

int tpcf_triplet_norm(const int64_t ngal[3], int term, double *norm) {
    clear_error();
    if (!ngal || !norm) return set_error(TPCF_EINVAL, "NULL argument");
    const int64_t a = ngal[0], b = ngal[1], c = ngal[2];
    switch (term) {
    case 0: case 3: *norm = exact_norm(a, b - 1, c - 2, 6.0); break;  // DDD, RRR
    case 1: *norm = exact_norm(a, b - 1, c, 2.0); break;              // DDR
    case 2: *norm = exact_norm(a, b, c - 1, 2.0); break;              // DRR
    default: return set_error(TPCF_EINVAL, "term %d is not 0..3 (DDD, DDR, DRR, RRR)", term);
    }
    return TPCF_OK;
}

int tpcf_pair_norm(const int64_t ngal[2], int term, double *norm) {
    clear_error();
    if (!ngal || !norm) return set_error(TPCF_EINVAL, "NULL argument");
    const int64_t a = ngal[0], b = ngal[1];
    switch (term) {
    case 0: case 2: *norm = (double)((__int128)a * a - a) / 2.0; break;  // (N^2 - N) / 2
    case 1: *norm = exact_norm(a, b, 1, 1.0); break;                        // N_0 N_1
    default: return set_error(TPCF_EINVAL, "term %d is not 0..2 (DD, DR, RR)", term);
    }
    return TPCF_OK;
}

int tpcf_zeta(const double *ddd, const double *ddr, const double *drr, const double *rrr, int64_t cells,
              const double norm[4], double *zeta) {
    clear_error();
    if (!ddd || !ddr || !drr || !rrr || !norm || !zeta || cells < 0) return set_error(TPCF_EINVAL, "NULL argument");
    for (int64_t c = 0; c < cells; ++c) {
        const double d3 = ddd[c] / norm[0], d2r = ddr[c] / norm[1];
        const double dr2 = drr[c] / norm[2], r3 = rrr[c] / norm[3];
        zeta[c] = clean_ratio((((d3 - 3.0 * d2r) + 3.0 * dr2) - r3) / r3);
    }
    return TPCF_OK;
}

int tpcf_xi(const double *dd, const double *dr, const double *rr, int64_t n, const double norm[3], double *xi) {
    clear_error();
    if (!dd || !dr || !rr || !norm || !xi || n < 0) return set_error(TPCF_EINVAL, "NULL argument");
    for (int64_t b = 0; b < n; ++b) {
        const double a = dd[b] / norm[0], m = dr[b] / norm[1], r = rr[b] / norm[2];
        xi[b] = clean_ratio(((a - 2.0 * m) + r) / r);
    }
    return TPCF_OK;
}

int tpcf_arange(double lo, double hi, double step, double *out, int64_t cap) {
    clear_error();
    if (!(step != 0.0)) return set_error(TPCF_EINVAL, "step is zero");
    const double len = std::ceil((hi - lo) / step);
    if (!(len > 0)) return 0;
    if (len > (double)cap) return set_error(TPCF_EINVAL, "arange of %g values exceeds the %lld given", len, (long long)cap);
    const int64_t n = (int64_t)len;
    // numpy fills [lo, lo + step, lo + i * ((lo + step) - lo) ...].
    const double second = lo + step, delta = second - lo;
    out[0] = lo;
    if (n > 1) out[1] = second;
    for (int64_t i = 2; i < n; ++i) out[i] = lo + (double)i * delta;
    return (int)n;
}

int tpcf_reduced_q(const double *zeta_theta, int ntheta, double s, double q, double theta_lo, double theta_width,
                   const double *xs, const double *ys, int nxi, double *theta, double *denom, double *qred) {
    clear_error();
    if (!zeta_theta || !xs || !ys || !theta || !denom || !qred || ntheta < 0 || nxi < 1)
        return set_error(TPCF_EINVAL, "bad argument");
    // xi at the table abscissa nearest to v (the first of equal distances).
    auto xi_at = [&](double v) {
        int best = 0;
        double m = std::fabs(v - xs[0]);
        for (int i = 1; i < nxi; ++i) {
            const double d = std::fabs(v - xs[i]);
            if (d < m) {
                m = d;
                best = i;
            }
        }
        return ys[best];
    };
    const double delta = (theta_lo + theta_width) - theta_lo;
    for (int t = 0; t < ntheta; ++t) {
        // np.arange(lo, hi, width)[t]
        const double x = t == 0 ? theta_lo : t == 1 ? theta_lo + theta_width : theta_lo + (double)t * delta;
        const double th = M_PI * (x + theta_width / 2.0);
        const double s12 = s, s23 = s * q, s31 = s * std::sqrt((1.0 + q * q) - (2.0 * q) * std::cos(th));
        const double c12 = xi_at(s12), c23 = xi_at(s23), c31 = xi_at(s31);
        theta[t] = x;
        denom[t] = (c12 * c23 + c23 * c31) + c31 * c12;
        qred[t] = zeta_theta[t] / denom[t];
    }
    return TPCF_OK;
}

}  // extern "C"
