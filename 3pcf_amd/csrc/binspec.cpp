// binspec.cpp -- bin-spec tables shared by every kernel of libtpcf.
//
// The reference hard-codes bin specs as #defines and bins with
// int((d - lo) / width) in float (src/3pcf_C_version_triangle_parametrization.cc:
// 31-42, 84-113; src/3pcf_C_version.cc:9-13, 23-51).  Here an axis is a
// runtime spec whose edges are tabulated once on the host in FP64; kernels
// only compare against the table, so every device and the CPU agree on the
// bin of any value by construction.
#include <cmath>
#include <cstring>

#include "tpcf_host.h"

namespace tpcf {

int axis_edges(const tpcf_axis &ax, double *e, int max_bins) {
    const int n = ax.n;
    if (n < 1 || n > max_bins) return set_error(TPCF_EINVAL, "axis needs 1..%d bins, got %d", max_bins, n);
    if (!(ax.hi > ax.lo) || !std::isfinite(ax.lo) || !std::isfinite(ax.hi))
        return set_error(TPCF_EINVAL, "axis range [%g, %g) is empty", ax.lo, ax.hi);
    if (ax.scale == TPCF_LOG) {
        if (!(ax.lo > 0)) return set_error(TPCF_EINVAL, "log axis needs lo > 0, got %g", ax.lo);
        const double a = std::log(ax.lo), b = std::log(ax.hi);
        for (int m = 0; m <= n; ++m) e[m] = std::exp(a + (b - a) * (double)m / (double)n);
    } else if (ax.scale == TPCF_LIN) {
        for (int m = 0; m <= n; ++m) e[m] = ax.lo + (ax.hi - ax.lo) * (double)m / (double)n;
    } else {
        return set_error(TPCF_EINVAL, "unknown axis scale %d", ax.scale);
    }
    e[0] = ax.lo;
    e[n] = ax.hi;
    for (int m = 0; m < n; ++m)
        if (!(e[m + 1] > e[m])) return set_error(TPCF_EINVAL, "axis edges are not increasing");
    return TPCF_OK;
}

// T[m] = smallest double s >= 0 with sqrt(s) >= e[m].  Because IEEE sqrt is
// monotone, sqrt(s) >= e[m]  <=>  s >= T[m]: comparing squared distances
// against T is exactly comparing distances against the edges.
void sq_thresholds(const double *e, int n, double *T) {
    for (int m = 0; m <= n; ++m) {
        const double em = e[m];
        if (!(em > 0)) {
            T[m] = 0.0;
            continue;
        }
        double t = em * em;
        while (std::sqrt(t) < em) t = std::nextafter(t, INFINITY);
        for (;;) {
            const double p = std::nextafter(t, 0.0);
            if (std::sqrt(p) >= em) t = p;
            else break;
        }
        T[m] = t;
    }
}

int binspec_check(const tpcf_binspec &s) {
    double e[kMaxBins + 1];
    if (s.param == TPCF_R123) {
        if (axis_edges(s.ax[0], e)) return TPCF_EINVAL;
        if (s.ax[0].lo < 0) return set_error(TPCF_EINVAL, "distance axis needs lo >= 0");
        return TPCF_OK;
    }
    if (s.param != TPCF_SQT && s.param != TPCF_JBJ) return set_error(TPCF_EINVAL, "unknown parametrization %d", s.param);
    for (int a = 0; a < 3; ++a)
        if (axis_edges(s.ax[a], e)) return TPCF_EINVAL;
    if (s.ax[0].lo < 0) return set_error(TPCF_EINVAL, "s axis needs lo >= 0");
    if (s.param == TPCF_SQT) {
        axis_edges(s.ax[2], e);
        if (e[0] < 0 || e[s.ax[2].n] > 1) return set_error(TPCF_EINVAL, "theta/pi axis must lie in [0, 1]");
    }
    const int64_t cells = (int64_t)s.ax[0].n * s.ax[1].n * s.ax[2].n;
    if (cells > kMaxSqtCells)
        return set_error(TPCF_EINVAL, "(s,q,theta) histogram of %lld cells exceeds %d", (long long)cells, kMaxSqtCells);
    return TPCF_OK;
}

int64_t hist_cells(const tpcf_binspec &s) {
    if (s.param == TPCF_R123) return (int64_t)s.ax[0].n * s.ax[0].n * s.ax[0].n;
    return (int64_t)s.ax[0].n * s.ax[1].n * s.ax[2].n;
}

}  // namespace tpcf
