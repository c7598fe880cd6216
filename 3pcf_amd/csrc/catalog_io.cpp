// catalog_io.cpp -- catalog readers of libtpcf.
//
// Two text formats exist in the reference:
//   7 columns "%d %f %f %f %f %f %f", x y z in columns 5-7
//     (src/3pcf_C_version_triangle_parametrization.cc:366-394);
//   a "%d" count header followed by "x y z" lines, each value divided by a
//     unit scale (src/3pcf_C_version.cc:311-332, scales :195-208).
// The reference parses into float; here values are parsed as FP64 (strtod,
// correctly rounded), which is what the FP64 definitions are stated on.
// Unlike the reference, a malformed line is an error, not an endless loop.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tpcf_host.h"

using namespace tpcf;

namespace {

bool slurp(const char *path, std::vector<char> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    buf.push_back('\0');
    return ok;
}

// Next whitespace-separated number; false at end of input.
bool next_number(const char *&p, double &v, bool &bad) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
    if (!*p) return false;
    char *end = nullptr;
    errno = 0;
    v = std::strtod(p, &end);
    if (end == p) {
        bad = true;
        return false;
    }
    p = end;
    return true;
}

}  // namespace

extern "C" int64_t tpcf_read_catalog(const char *path, int format, double scale, double **out) {
    clear_error();
    if (!path || !out) return set_error(TPCF_EINVAL, "NULL argument");
    if (format != 7 && format != 3) return set_error(TPCF_EINVAL, "catalog format must be 7 or 3, got %d", format);
    if (!(scale > 0)) return set_error(TPCF_EINVAL, "scale must be positive");
    std::vector<char> buf;
    if (!slurp(path, buf)) return set_error(TPCF_EIO, "cannot read catalog %s", path);
    const char *p = buf.data();
    std::vector<double> xyz;
    bool bad = false;
    double v;
    int64_t n = 0;
    if (format == 7) {
        double row[7];
        for (;;) {
            int k = 0;
            while (k < 7 && next_number(p, row[k], bad)) ++k;
            if (bad) return set_error(TPCF_EIO, "%s: unparsable token after %lld rows", path, (long long)n);
            if (k == 0) break;
            if (k < 7) return set_error(TPCF_EIO, "%s: truncated row %lld (%d of 7 columns)", path, (long long)n, k);
            xyz.insert(xyz.end(), {row[4], row[5], row[6]});
            ++n;
        }
    } else {
        if (!next_number(p, v, bad) || v < 0 || v != (double)(int64_t)v)
            return set_error(TPCF_EIO, "%s: missing point count header", path);
        const int64_t want = (int64_t)v;
        for (int64_t i = 0; i < want; ++i) {
            double r[3];
            for (int c = 0; c < 3; ++c)
                if (!next_number(p, r[c], bad))
                    return set_error(TPCF_EIO, "%s: %lld points announced, %lld read", path, (long long)want,
                                     (long long)i);
            xyz.insert(xyz.end(), {r[0] / scale, r[1] / scale, r[2] / scale});
        }
        n = want;
    }
    double *soa = static_cast<double *>(std::malloc(sizeof(double) * 3 * (n > 0 ? n : 1)));
    if (!soa) return set_error(TPCF_ENOMEM, "host allocation failed");
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[c * n + i] = xyz[3 * i + c];
    *out = soa;
    return n;
}

// RA/Dec/redshift to Cartesian Mpc/h, the conversion of
// sample_data/gen_flat_MICE_data.py: distance D = p0 + p1 z + p2 z^2 (:49-52,
// MICE fit parameters at :82), x = D sin(90 - dec) sin(ra), y = D sin(90 -
// dec) cos(ra), z = D cos(90 - dec) in degrees (:57-72), with the script's
// order of operations.
static const double kMiceFit[3] = {-1.07199554e+00, 3.01313296e+03, -6.38827998e+02};

extern "C" int tpcf_radecz_to_xyz(const double *ra, const double *dec, const double *zred, int64_t n,
                                  const double poly[3], double *xyz) {
    clear_error();
    if (n < 0 || (n > 0 && (!ra || !dec || !zred || !xyz))) return set_error(TPCF_EINVAL, "NULL argument");
    const double *p = poly ? poly : kMiceFit;
    const double rad = M_PI / 180.0;
    for (int64_t i = 0; i < n; ++i) {
        const double z = zred[i];
        const double mpc = (p[0] + p[1] * z) + (p[2] * z) * z;
        const double polar = rad * (90.0 - dec[i]), az = rad * ra[i];
        const double sp = mpc * std::sin(polar);
        xyz[i] = sp * std::sin(az);
        xyz[n + i] = sp * std::cos(az);
        xyz[2 * n + i] = mpc * std::cos(polar);
    }
    return TPCF_OK;
}

// A catalog of "id ra dec z ..." rows (at least four columns; the count of
// the first row is kept for every row, as in the 7-column MICE files of
// sample_data/), converted to Cartesian SoA.
extern "C" int64_t tpcf_read_radecz(const char *path, const double poly[3], double **out) {
    clear_error();
    if (!path || !out) return set_error(TPCF_EINVAL, "NULL argument");
    std::vector<char> buf;
    if (!slurp(path, buf)) return set_error(TPCF_EIO, "cannot read catalog %s", path);
    // Columns of the first row.
    int ncol = 0;
    for (const char *q = buf.data(); *q && *q != '\n';) {
        while (*q == ' ' || *q == '\t' || *q == '\r') ++q;
        if (!*q || *q == '\n') break;
        ++ncol;
        while (*q && *q != ' ' && *q != '\t' && *q != '\r' && *q != '\n') ++q;
    }
    if (ncol < 4) return set_error(TPCF_EIO, "%s: need at least 4 columns (id ra dec z), got %d", path, ncol);
    const char *q = buf.data();
    std::vector<double> ra, dec, zr;
    std::vector<double> row(ncol);
    bool bad = false;
    for (;;) {
        int k = 0;
        while (k < ncol && next_number(q, row[k], bad)) ++k;
        if (bad) return set_error(TPCF_EIO, "%s: unparsable token after %zu rows", path, ra.size());
        if (k == 0) break;
        if (k < ncol) return set_error(TPCF_EIO, "%s: truncated row %zu", path, ra.size());
        ra.push_back(row[1]);
        dec.push_back(row[2]);
        zr.push_back(row[3]);
    }
    const int64_t n = (int64_t)ra.size();
    double *soa = static_cast<double *>(std::malloc(sizeof(double) * 3 * (n > 0 ? n : 1)));
    if (!soa) return set_error(TPCF_ENOMEM, "host allocation failed");
    tpcf_radecz_to_xyz(ra.data(), dec.data(), zr.data(), n, poly, soa);
    *out = soa;
    return n;
}
