// estimate.cpp -- tpcf_estimate: correlation-function estimators over the
// histogram files the counting programs write (the analysis step of
// examples/plot_MICE_3pcf.py and examples/plot_MICE_2pcf.py, without plots).
//
//   tpcf_estimate zeta DDD_file [-o out]
//       reads DDD_file and the files named with DDD replaced by DDR, DRR and
//       RRR (every occurrence, as plot_MICE_3pcf.py:58-61 does); writes the
//       TRI header of DDD_file, then per (s, q) row the theta/pi values of
//       zeta = (DDD - 3 DDR + 3 DRR - RRR) / RRR, each term divided by its
//       triplet count from its own file's catalog sizes.
//   tpcf_estimate reduced DDD_file SBIN QBIN [QBIN ...] [--xi TABLE] [-o out]
//       for s bin SBIN and each q bin: theta/pi (bin low edge), zeta, the
//       denominator xi12 xi23 + xi23 xi31 + xi31 xi12 and the reduced
//       amplitude Q = zeta / denominator, xi from TABLE ("s xi" lines, default
//       master_2pcf_reference.dat) at the nearest s (plot_MICE_3pcf.py:176-221).
//   tpcf_estimate xi DD_file [-o out]
//       reads DD_file and the files named with DD replaced by DR and RR
//       (plot_MICE_2pcf.py:11-13); writes "s xi" lines, s the upper bin edge,
//       xi = (DD - 2 DR + RR) / RR (plot_MICE_2pcf.py:127-131); default out
//       master_2pcf_reference.dat.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tpcf.h"

namespace {

int fail(const char *what, const std::string &arg = "") {
    std::fprintf(stderr, "error: %s%s%s%s%s\n", what, arg.empty() ? "" : " ", arg.c_str(),
                 tpcf_last_error()[0] ? ": " : "", tpcf_last_error());
    return 1;
}

bool read_tokens(const std::string &path, std::vector<std::string> &tok) {
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char buf[512];
    while (std::fscanf(f, "%511s", buf) == 1) tok.emplace_back(buf);
    std::fclose(f);
    return true;
}

std::string replace_all(std::string s, const std::string &from, const std::string &to) {
    for (size_t p = s.find(from); p != std::string::npos; p = s.find(from, p + to.size())) s.replace(p, from.size(), to);
    return s;
}

// A histogram file: catalog sizes, bin counts, range lines (lo, hi, width
// per axis) and the values, parsed as the scripts parse them (whitespace
// tokens, FP64).
struct Hist {
    int64_t ngal[3] = {0, 0, 0};
    int nb[3] = {1, 1, 1};
    double rng[3][3] = {};
    std::vector<double> v;
};

bool read_tri(const std::string &path, Hist &h) {
    std::vector<std::string> t;
    if (!read_tokens(path, t) || t.size() < 15) return false;
    for (int i = 0; i < 3; ++i) h.ngal[i] = std::atoll(t[i].c_str());
    for (int i = 0; i < 3; ++i) h.nb[i] = std::atoi(t[3 + i].c_str());
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) h.rng[a][k] = std::strtod(t[6 + 3 * a + k].c_str(), nullptr);
    const size_t cells = (size_t)h.nb[0] * h.nb[1] * h.nb[2];
    if (t.size() < 15 + cells) return false;
    h.v.resize(cells);
    for (size_t c = 0; c < cells; ++c) h.v[c] = std::strtod(t[15 + c].c_str(), nullptr);
    return true;
}

bool read_two(const std::string &path, Hist &h) {
    std::vector<std::string> t;
    if (!read_tokens(path, t) || t.size() < 6) return false;
    h.ngal[0] = std::atoll(t[0].c_str());
    h.ngal[1] = std::atoll(t[1].c_str());
    h.nb[0] = std::atoi(t[2].c_str());
    for (int k = 0; k < 3; ++k) h.rng[0][k] = std::strtod(t[3 + k].c_str(), nullptr);
    if (t.size() < 6 + (size_t)h.nb[0]) return false;
    h.v.resize(h.nb[0]);
    for (int b = 0; b < h.nb[0]; ++b) h.v[b] = std::strtod(t[6 + b].c_str(), nullptr);
    return true;
}

int zeta_of(const std::string &ddd_path, Hist &H, std::vector<double> &zeta) {
    static const char *names[4] = {"DDD", "DDR", "DRR", "RRR"};
    Hist h[4];
    double norm[4];
    for (int t = 0; t < 4; ++t) {
        const std::string p = replace_all(ddd_path, "DDD", names[t]);
        if (!read_tri(p, h[t])) return fail("cannot read histogram file", p);
        if (tpcf_triplet_norm(h[t].ngal, t, &norm[t]) != TPCF_OK) return fail("bad catalog sizes in", p);
        if (h[t].v.size() != h[0].v.size()) return fail("histogram shapes differ:", p);
    }
    zeta.resize(h[0].v.size());
    if (tpcf_zeta(h[0].v.data(), h[1].v.data(), h[2].v.data(), h[3].v.data(), (int64_t)zeta.size(), norm,
                  zeta.data()) != TPCF_OK)
        return fail("estimator failed");
    H = h[0];
    return 0;
}

int main_zeta(int argc, char **argv, const char *out) {
    if (argc < 1) return fail("usage: tpcf_estimate zeta DDD_file [-o out]");
    Hist H;
    std::vector<double> z;
    if (int e = zeta_of(argv[0], H, z)) return e;
    FILE *f = std::fopen(out ? out : "zeta.dat", "w");
    if (!f) return fail("cannot write", out ? out : "zeta.dat");
    for (int i = 0; i < 3; ++i) std::fprintf(f, "%lld\n", (long long)H.ngal[i]);
    std::fprintf(f, "%d %d %d\n", H.nb[0], H.nb[1], H.nb[2]);
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%-6.3f %-6.3f %-6.3f\n", H.rng[a][0], H.rng[a][1], H.rng[a][2]);
    for (int i = 0; i < H.nb[0] * H.nb[1]; ++i) {
        for (int k = 0; k < H.nb[2]; ++k) std::fprintf(f, "%.17g ", z[(size_t)i * H.nb[2] + k]);
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? 0 : fail("cannot write output");
}

int main_reduced(int argc, char **argv, const char *out, const char *xi_path) {
    if (argc < 3) return fail("usage: tpcf_estimate reduced DDD_file SBIN QBIN [QBIN ...] [--xi TABLE]");
    Hist H;
    std::vector<double> z;
    if (int e = zeta_of(argv[0], H, z)) return e;
    std::vector<std::string> t;
    const std::string xp = xi_path ? xi_path : "master_2pcf_reference.dat";
    if (!read_tokens(xp, t) || t.size() < 2) return fail("cannot read the 2PCF table", xp);
    std::vector<double> xs, ys;
    for (size_t i = 0; i + 1 < t.size(); i += 2) {
        xs.push_back(std::strtod(t[i].c_str(), nullptr));
        ys.push_back(std::strtod(t[i + 1].c_str(), nullptr));
    }
    const int sbin = std::atoi(argv[1]);
    const int ns = H.nb[0], nq = H.nb[1], nt = H.nb[2];
    if (sbin < 0 || sbin >= ns) return fail("s bin out of range:", argv[1]);
    // Bin centre lo + width * i + width / 2 with width = (hi - lo) / n
    // (plot_MICE_3pcf.py:33-40), from the file's range lines.
    auto centre = [&](int axis, int i) {
        const double lo = H.rng[axis][0], hi = H.rng[axis][1];
        const double w = (hi - lo) / (double)H.nb[axis];
        return (lo + w * i) + w / 2.0;
    };
    std::vector<double> th(nt + 1), x(nt), den(nt), q(nt);
    // The theta/pi abscissae: np.arange(lo, hi, width) of the file's range line.
    const int nx = tpcf_arange(H.rng[2][0], H.rng[2][1], H.rng[2][2], th.data(), nt);
    if (nx < 0) return fail("theta range line disagrees with the bin count");
    FILE *f = std::fopen(out ? out : "reduced_q.dat", "w");
    if (!f) return fail("cannot write", out ? out : "reduced_q.dat");
    const double sval = centre(0, sbin);
    std::fprintf(f, "# s bin %d (s = %.17g): q_lo q_hi theta/pi zeta denominator Q\n", sbin, sval);
    for (int a = 2; a < argc; ++a) {
        const int qb = std::atoi(argv[a]);
        if (qb < 0 || qb >= nq) {
            std::fclose(f);
            return fail("q bin out of range:", argv[a]);
        }
        const double qval = centre(1, qb);
        const double *zrow = z.data() + ((size_t)sbin * nq + qb) * nt;
        if (tpcf_reduced_q(zrow, nx, sval, qval, H.rng[2][0], H.rng[2][2], xs.data(), ys.data(), (int)xs.size(),
                           x.data(), den.data(), q.data()) != TPCF_OK) {
            std::fclose(f);
            return fail("reduced amplitude failed");
        }
        const double qw = H.rng[1][2], qlo = H.rng[1][0];
        for (int k = 0; k < nx; ++k)
            std::fprintf(f, "%.6g %.6g %.17g %.17g %.17g %.17g\n", qb * qw + qlo, (qb + 1) * qw + qlo, x[k], zrow[k],
                         den[k], q[k]);
    }
    return std::fclose(f) == 0 ? 0 : fail("cannot write output");
}

int main_xi(int argc, char **argv, const char *out) {
    if (argc < 1) return fail("usage: tpcf_estimate xi DD_file [-o out]");
    static const char *names[3] = {"DD", "DR", "RR"};
    Hist h[3];
    double norm[3];
    const std::string dd = argv[0];
    for (int t = 0; t < 3; ++t) {
        const std::string p = replace_all(dd, "DD", names[t]);
        if (!read_two(p, h[t])) return fail("cannot read pair-count file", p);
        // plot_MICE_2pcf.py:43-50: DD and RR use the file's first catalog size.
        if (tpcf_pair_norm(h[t].ngal, t, &norm[t]) != TPCF_OK) return fail("bad catalog sizes in", p);
        if (h[t].v.size() != h[0].v.size()) return fail("histogram lengths differ:", p);
    }
    const int n = h[0].nb[0];
    std::vector<double> xi(n);
    if (tpcf_xi(h[0].v.data(), h[1].v.data(), h[2].v.data(), n, norm, xi.data()) != TPCF_OK)
        return fail("estimator failed");
    // The range line of the last file read (plot_MICE_2pcf.py:96-101).
    const double lo = h[2].rng[0][0], hi = h[2].rng[0][1], w = h[2].rng[0][2];
    std::vector<double> x(n + 1);
    const int nx = tpcf_arange(lo, hi, w, x.data(), n + 1);
    if (nx < 0) return fail("range line disagrees with the bin count");
    const char *o = out ? out : "master_2pcf_reference.dat";
    FILE *f = std::fopen(o, "w");
    if (!f) return fail("cannot write", o);
    for (int b = 0; b < std::min(nx, n); ++b) std::fprintf(f, "%f %f\n", x[b] + w, xi[b]);
    return std::fclose(f) == 0 ? 0 : fail("cannot write output");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: tpcf_estimate zeta|reduced|xi FILE ... [-o out] [--xi TABLE]\n");
        return 2;
    }
    const char *out = nullptr, *xi = nullptr;
    std::vector<char *> pos;
    for (int i = 2; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!std::strcmp(argv[i], "--xi") && i + 1 < argc) xi = argv[++i];
        else pos.push_back(argv[i]);
    }
    const std::string mode = argv[1];
    if (mode == "zeta") return main_zeta((int)pos.size(), pos.data(), out);
    if (mode == "reduced") return main_reduced((int)pos.size(), pos.data(), out, xi);
    if (mode == "xi") return main_xi((int)pos.size(), pos.data(), out);
    std::fprintf(stderr, "error: unknown mode %s (zeta, reduced, xi)\n", argv[1]);
    return 2;
}
