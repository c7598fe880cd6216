// cli.cpp -- drop-in command-line programs over libtpcf.
//
// One source, four programs (selected at build time with TPCF_PROGRAM):
//   0  3pcf_C_version_triangle_parametrization      (s,q,theta), TRI format
//        argv of src/3pcf_C_version_triangle_parametrization.cc:269-313
//   1  3pcf_C_version_triangle_parametrization_JBJ  JBJ variant, TRI format
//   2  3pcf_C_version                               sorted (r1,r2,r3), cube format
//        argv of src/3pcf_C_version.cc:181-217
//   3  2pcf_C_version                               pair counts (see two_main)
// Usage is the reference's: three catalog paths plus getopt flags, one
// histogram file out (default default_out.dat).  Long options add runtime
// bin specs and the device path:
//   --param r123|sqt|jbj    parametrization
//   --axis I,N,LO,HI[,log]  bin spec of axis I (0..2); r123 uses axis 0
//   --bins N                r123: N bins per side (with --lo/--hi)
//   --lo X --hi X --log     r123 range and log scale
//   --mode fast|plain|fast64|cells
//   --format tri|plain      output layout (default per program)
//   --recheck-rel X         FP64 recheck window (default 1e-5)
//   --device N, --quiet, --stats
//   --all-voxels            with -X DIV: every chunk triple (i,j,k) of DIV in
//                           one batched launch, one file per voxel named by
//                           -o with "{index}" replaced by "%03d%03d%03d"
// Which triplets are counted follows the catalogs given, compared as files
// (realpath), not as strings: "D D R" counts C(N_D,2)*N_R triplets.
#include <getopt.h>
#include <algorithm>
#include <limits.h>
#include <stdlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tpcf.h"

#ifndef TPCF_PROGRAM
#define TPCF_PROGRAM 0
#endif

namespace {

struct Catalog {
    std::string path, real;
    double *xyz = nullptr;
    int64_t n = 0;
};

int die(int code, const char *msg) {
    std::fprintf(stderr, "error: %s%s%s\n", msg, tpcf_last_error()[0] ? ": " : "", tpcf_last_error());
    return code;
}

bool parse_axis(const char *arg, int &idx, tpcf_axis &ax) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s", arg);
    std::vector<char *> tok;
    for (char *p = std::strtok(buf, ","); p; p = std::strtok(nullptr, ",")) tok.push_back(p);
    if (tok.size() < 4 || tok.size() > 5) return false;
    idx = std::atoi(tok[0]);
    ax.n = std::atoi(tok[1]);
    ax.lo = std::atof(tok[2]);
    ax.hi = std::atof(tok[3]);
    ax.scale = (tok.size() == 5 && std::strcmp(tok[4], "log") == 0) ? TPCF_LOG : TPCF_LIN;
    return idx >= 0 && idx < 3;
}

// Reference voxel chunk: src/3pcf_C_version_triangle_parametrization.cc:52-80.
bool vox_chunk(int div, int idx, int64_t n, int64_t &lo, int64_t &hi) {
    const int64_t per = n / div;
    if (per == 0) return false;
    lo = per * idx;
    hi = per * (idx + 1);
    if (idx == div - 1) hi += n % div;
    return true;
}

// TRI-format writer (src/3pcf_C_version_triangle_parametrization.cc:537-569).
bool write_tri(const char *path, const int64_t ng[3], const int nb[3], const double lo[3], const double hi[3],
               const double w[3], const uint64_t *h) {
    FILE *f = std::fopen(path, "w");
    if (!f) return false;
    for (int c = 0; c < 3; ++c) std::fprintf(f, "%lld\n", (long long)ng[c]);
    std::fprintf(f, "%d %d %d\n", nb[0], nb[1], nb[2]);
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%-6.3f %-6.3f %-6.3f\n", lo[a], hi[a], w[a]);
    for (int i = 0; i < nb[0]; ++i)
        for (int j = 0; j < nb[1]; ++j) {
            for (int k = 0; k < nb[2]; ++k)
                std::fprintf(f, "%llu ", (unsigned long long)h[((int64_t)i * nb[1] + j) * nb[2] + k]);
            std::fputc('\n', f);
        }
    return std::fclose(f) == 0;
}

// Cube-format writer (src/3pcf_C_version.cc:440-460): sorted cell
// (r1,r2,r3) lives at n^2*r1 + n*r2 + r3.
bool write_plain(const char *path, int n, const uint64_t *h) {
    FILE *f = std::fopen(path, "w");
    if (!f) return false;
    std::fprintf(f, "%d %d %d\n", n, n, n);
    for (int i = 0; i < n; ++i) {
        std::fprintf(f, "%d\n", i);
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < n; ++k)
                std::fprintf(f, "%llu ", (unsigned long long)h[(int64_t)n * n * k + (int64_t)n * j + i]);
            std::fputc('\n', f);
        }
    }
    return std::fclose(f) == 0;
}

// Pair-count file (src/2pcf_C_version.cc:391-404): both catalog sizes, the
// bin count, one range line, then one count per line.
bool write_two(const char *path, const int64_t ng[2], const tpcf_axis &ax, const uint64_t *h) {
    FILE *f = std::fopen(path, "w");
    if (!f) return false;
    const double w = ax.scale == TPCF_LOG ? std::log(ax.hi / ax.lo) / ax.n : (ax.hi - ax.lo) / ax.n;
    std::fprintf(f, "%lld\n%lld\n%d\n", (long long)ng[0], (long long)ng[1], ax.n);
    std::fprintf(f, "%-6.3f %-6.3f %-6.3f\n", ax.lo, ax.hi, w);
    for (int b = 0; b < ax.n; ++b) std::fprintf(f, "%llu\n", (unsigned long long)h[b]);
    return std::fclose(f) == 0;
}

// Index ranges of the three roles: whole catalogs (div <= 0) or the
// reference's voxel chunks vid[i] of div.  False when a chunk is empty.
bool voxel_ranges(const Catalog cat[3], int div, const int vid[3], int64_t lo[3], int64_t hi[3]) {
    for (int i = 0; i < 3; ++i) {
        lo[i] = 0;
        hi[i] = cat[i].n;
        if (div > 0 && !vox_chunk(div, vid[i], cat[i].n, lo[i], hi[i])) return false;
    }
    return true;
}

// Roles of the same file must come in chunk order; otherwise the voxel
// selects no triplet.
bool selects_nothing(const Catalog cat[3], const int64_t lo[3]) {
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (cat[a].real == cat[b].real && lo[a] > lo[b]) return true;
    return false;
}

// One job: a segment per distinct (file, chunk), one term of the sorted
// segment labels of the three roles.
void voxel_job(const Catalog cat[3], tpcf_catalog *const dev[3], const int64_t lo[3], const int64_t hi[3],
               tpcf_job &job) {
    job = tpcf_job{};
    int label[3] = {-1, -1, -1};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < i; ++j)
            if (cat[j].real == cat[i].real && lo[j] == lo[i]) {
                label[i] = label[j];
                break;
            }
        if (label[i] >= 0) continue;
        label[i] = job.nseg;
        job.seg[job.nseg] = dev[i];
        job.seg_lo[job.nseg] = lo[i];
        job.seg_hi[job.nseg] = hi[i];
        job.nseg++;
    }
    int t3[3] = {label[0], label[1], label[2]};
    std::sort(t3, t3 + 3);
    job.nterms = 1;
    std::memcpy(job.term[0], t3, sizeof t3);
}

// The program's histogram file: cube format, or TRI with the whole
// catalogs' sizes in the header (as the reference writes for a voxel too).
bool write_hist(const char *path, int format, const tpcf_binspec &spec, const Catalog cat[3], const uint64_t *h) {
    if (format == 1) return write_plain(path, spec.ax[0].n, h);
    const int64_t ng[3] = {cat[0].n, cat[1].n, cat[2].n};
    int nb[3];
    double alo[3], ahi[3], w[3];
    for (int a = 0; a < 3; ++a) {
        const tpcf_axis &ax = spec.param == TPCF_R123 ? spec.ax[0] : spec.ax[a];
        nb[a] = ax.n;
        alo[a] = ax.lo;
        ahi[a] = ax.hi;
        w[a] = ax.scale == TPCF_LOG ? std::log(ax.hi / ax.lo) / ax.n : (ax.hi - ax.lo) / ax.n;
    }
    return write_tri(path, ng, nb, alo, ahi, w, h);
}

void print_stats(const tpcf_stats &st) {
    std::printf("{\"triplets\": %llu, \"pairs\": %llu, \"pairs_rechecked\": %llu, \"device_ms\": %.3f, "
                "\"wall_ms\": %.3f}\n",
                (unsigned long long)st.triplets, (unsigned long long)st.pairs,
                (unsigned long long)st.pairs_rechecked, st.device_ms, st.wall_ms);
}

// 2pcf_C_version f0 f1 [-o out] [-l flag] [-X div -x AAABBBCCC] [-a]: the
// argv of src/2pcf_C_version.cc:143-210 (-l is parsed and, as there, has no
// effect).  Two 7-column catalogs; voxel chunk AAA of f0 against chunk BBB of
// f1.  Default axis: 700 linear bins on [0, 70) (src/2pcf_C_version.cc:14-16);
// --bins N --lo X --hi X [--log] change it.  --mode fast|plain|fast64|cells,
// --device N, --quiet, --stats as for the triplet programs.  The catalogs are
// compared as files (realpath): one file gives the C(N,2) unique pairs (i < j,
// no self pairs); a chunk pair (p, q) of one file counts chunk p against
// chunk q for p < q, the unique pairs of chunk p for p == q, and nothing for
// p > q, so the chunk pairs of a voxel run add up to the full run.
int two_main(int argc, char **argv) {
    const char *outfile = "default_out_2pcf.dat";
    int vdiv = -999, vtemp = -999, device = 0, quiet = 0, want_stats = 0, mode = TPCF_MODE_FAST;
    tpcf_axis ax{700, 0.0, 70.0, TPCF_LIN};
    enum { O_BINS = 1000, O_LO, O_HI, O_LOG, O_MODE, O_DEVICE, O_QUIET, O_STATS };
    static option longopts[] = {{"bins", required_argument, nullptr, O_BINS},
                                {"lo", required_argument, nullptr, O_LO},
                                {"hi", required_argument, nullptr, O_HI},
                                {"log", no_argument, nullptr, O_LOG},
                                {"mode", required_argument, nullptr, O_MODE},
                                {"device", required_argument, nullptr, O_DEVICE},
                                {"quiet", no_argument, nullptr, O_QUIET},
                                {"stats", no_argument, nullptr, O_STATS},
                                {nullptr, 0, nullptr, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "ao:l:x:X:", longopts, nullptr)) != -1) {
        switch (c) {
        case 'o': outfile = optarg; break;
        case 'X': vdiv = std::atoi(optarg); break;
        case 'x': vtemp = std::atoi(optarg); break;
        case 'a': case 'l': break;
        case O_BINS: ax.n = std::atoi(optarg); break;
        case O_LO: ax.lo = std::atof(optarg); break;
        case O_HI: ax.hi = std::atof(optarg); break;
        case O_LOG: ax.scale = TPCF_LOG; break;
        case O_MODE:
            if (!std::strcmp(optarg, "fast")) mode = TPCF_MODE_FAST;
            else if (!std::strcmp(optarg, "plain")) mode = TPCF_MODE_PLAIN;
            else if (!std::strcmp(optarg, "fast64")) mode = TPCF_MODE_FAST64;
            else if (!std::strcmp(optarg, "cells")) mode = TPCF_MODE_CELLS;
            else { std::fprintf(stderr, "error: --mode fast|plain|fast64|cells\n"); return 2; }
            break;
        case O_DEVICE: device = std::atoi(optarg); break;
        case O_QUIET: quiet = 1; break;
        case O_STATS: want_stats = 1; break;
        case '?': std::printf("unknown arg %c\n", optopt); break;  // the reference continues
        default: break;
        }
    }
    if (argc - optind < 2) {
        std::printf("\nMust pass in at least two input files on command line!\n");
        return 1;
    }
    std::vector<double> e(ax.n > 0 && ax.n <= 4096 ? ax.n + 1 : 1);
    if (tpcf_axis_edges(&ax, e.data()) != TPCF_OK) return die(2, "bad axis");
    Catalog cat[2];
    for (int i = 0; i < 2; ++i) {
        cat[i].path = argv[optind + i];
        char rp[PATH_MAX];
        if (!realpath(cat[i].path.c_str(), rp)) {
            std::fprintf(stderr, "error: cannot open %s\n", cat[i].path.c_str());
            return 1;
        }
        cat[i].real = rp;
    }
    const bool same_file = cat[0].real == cat[1].real;
    for (int i = 0; i < 2; ++i) {
        if (i == 1 && same_file) {
            cat[1].xyz = cat[0].xyz;
            cat[1].n = cat[0].n;
            break;
        }
        cat[i].n = tpcf_read_catalog(cat[i].path.c_str(), 7, 1.0, &cat[i].xyz);
        if (cat[i].n < 0) return die(1, "cannot read catalog");
        if (!quiet) std::printf("Opening input file %d: %s (%lld points)\n", i, cat[i].path.c_str(), (long long)cat[i].n);
    }
    int64_t lo[2], hi[2];
    int vidx[2] = {0, 0};
    if (vdiv > 0) {
        vidx[0] = vtemp / 1000000;
        vidx[1] = (vtemp / 1000) % 1000;
    }
    for (int i = 0; i < 2; ++i) {
        lo[i] = 0;
        hi[i] = cat[i].n;
        if (vdiv > 0 && vtemp >= 0 && !vox_chunk(vdiv, vidx[i], cat[i].n, lo[i], hi[i])) {
            std::printf("Voxel chunks are bigger than the number of galaxies!");
            return 255;
        }
    }
    const bool one_seg = same_file && lo[0] == lo[1];
    const bool empty = same_file && lo[0] > lo[1];

    tpcf_ctx *ctx = nullptr;
    if (tpcf_ctx_create(device, &ctx) != TPCF_OK) return die(3, "no usable GPU");
    tpcf_catalog *dev[2] = {nullptr, nullptr};
    for (int i = 0; i < 2; ++i) {
        if (i == 1 && same_file) {
            dev[1] = dev[0];
            break;
        }
        if (tpcf_catalog_create(ctx, cat[i].xyz, cat[i].n, &dev[i]) != TPCF_OK) return die(3, "catalog upload failed");
    }
    tpcf_job job{};
    job.nseg = one_seg ? 1 : 2;
    for (int s = 0; s < job.nseg; ++s) {
        job.seg[s] = dev[s];
        job.seg_lo[s] = lo[s];
        job.seg_hi[s] = hi[s];
    }
    job.nterms = 1;
    job.term[0][0] = 0;
    job.term[0][1] = one_seg ? 0 : 1;
    std::vector<uint64_t> h(ax.n, 0);
    tpcf_stats st{};
    if (!empty && tpcf_count_pairs(ctx, &job, &ax, mode, h.data(), &st) != TPCF_OK) return die(4, "counting failed");
    uint64_t total = 0;
    for (auto v : h) total += v;
    const int64_t ng[2] = {cat[0].n, cat[1].n};
    if (!write_two(outfile, ng, ax, h.data())) return die(1, "cannot write the histogram file");
    if (!quiet) {
        std::printf("Output filename is %s\n", outfile);
        std::printf("Total: %llu\n", (unsigned long long)total);
    }
    if (want_stats)
        std::printf("{\"pairs\": %llu, \"device_ms\": %.3f, \"wall_ms\": %.3f}\n", (unsigned long long)st.pairs,
                    st.device_ms, st.wall_ms);
    tpcf_catalog_destroy(dev[0]);
    if (!same_file) tpcf_catalog_destroy(dev[1]);
    tpcf_free(cat[0].xyz);
    if (!same_file) tpcf_free(cat[1].xyz);
    tpcf_ctx_destroy(ctx);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const int program = TPCF_PROGRAM;
    if (program == 3) return two_main(argc, argv);
    const bool triprog = program != 2;
    const char *outfile = "default_out.dat";
    int vdiv = -999, vtemp = -999, device = 0, quiet = 0, want_stats = 0, all_voxels = 0;
    double scale = 1.0;
    tpcf_binspec spec{};
    spec.recheck_rel = 1e-5;
    if (program == 2) {
        // src/3pcf_C_version.cc:169-174: 64 bins on [0, sqrt(3*24*24)).
        spec.param = TPCF_R123;
        spec.ax[0] = {64, 0.0, std::sqrt(3.0 * 24 * 24), TPCF_LIN};
    } else {
        // src/3pcf_C_version_triangle_parametrization.cc:31-42.
        spec.param = program == 1 ? TPCF_JBJ : TPCF_SQT;
        spec.ax[0] = {50, 2.0, 12.0, TPCF_LIN};
        spec.ax[1] = {16, 0.9, 4.1, TPCF_LIN};
        spec.ax[2] = {32, 0.0, 1.0, TPCF_LIN};
    }
    int mode = -1;  // default: FAST
    int format = triprog ? 0 : 1;
    bool have_r123_range = false;
    double r_lo = 0, r_hi = 0;
    int r_n = 0, r_log = 0;

    enum { O_PARAM = 1000, O_AXIS, O_BINS, O_LO, O_HI, O_LOG, O_MODE, O_FORMAT, O_RECHECK, O_DEVICE, O_QUIET, O_STATS, O_ALLVOX };
    static option longopts[] = {{"param", required_argument, nullptr, O_PARAM},
                                {"axis", required_argument, nullptr, O_AXIS},
                                {"bins", required_argument, nullptr, O_BINS},
                                {"lo", required_argument, nullptr, O_LO},
                                {"hi", required_argument, nullptr, O_HI},
                                {"log", no_argument, nullptr, O_LOG},
                                {"mode", required_argument, nullptr, O_MODE},
                                {"format", required_argument, nullptr, O_FORMAT},
                                {"recheck-rel", required_argument, nullptr, O_RECHECK},
                                {"device", required_argument, nullptr, O_DEVICE},
                                {"quiet", no_argument, nullptr, O_QUIET},
                                {"stats", no_argument, nullptr, O_STATS},
                                {"all-voxels", no_argument, nullptr, O_ALLVOX},
                                {nullptr, 0, nullptr, 0}};
    const char *shortopts = triprog ? "ao:l:x:X:" : "ao:L:l:w:sm";
    int c;
    while ((c = getopt_long(argc, argv, shortopts, longopts, nullptr)) != -1) {
        switch (c) {
        case 'o': outfile = optarg; break;
        case 'X': vdiv = std::atoi(optarg); break;
        case 'x': vtemp = std::atoi(optarg); break;
        case 's': scale = 206264.0; break;     // src/3pcf_C_version.cc:195-201
        case 'm': scale = 3437.74677; break;   // src/3pcf_C_version.cc:202-208
        case 'a': case 'l': case 'L': case 'w': break;  // parsed by the reference, no effect on counts
        case O_PARAM:
            if (!std::strcmp(optarg, "r123")) spec.param = TPCF_R123;
            else if (!std::strcmp(optarg, "sqt")) spec.param = TPCF_SQT;
            else if (!std::strcmp(optarg, "jbj")) spec.param = TPCF_JBJ;
            else { std::fprintf(stderr, "error: --param r123|sqt|jbj\n"); return 2; }
            if (spec.param == TPCF_R123 && program != 2) spec.ax[0] = {16, 0.0, 40.0, TPCF_LIN};
            break;
        case O_AXIS: {
            int idx;
            tpcf_axis ax;
            if (!parse_axis(optarg, idx, ax)) { std::fprintf(stderr, "error: --axis I,N,LO,HI[,log]\n"); return 2; }
            spec.ax[idx] = ax;
            break;
        }
        case O_BINS: r_n = std::atoi(optarg); have_r123_range = true; break;
        case O_LO: r_lo = std::atof(optarg); have_r123_range = true; break;
        case O_HI: r_hi = std::atof(optarg); have_r123_range = true; break;
        case O_LOG: r_log = 1; have_r123_range = true; break;
        case O_MODE:
            if (!std::strcmp(optarg, "fast")) mode = TPCF_MODE_FAST;
            else if (!std::strcmp(optarg, "plain")) mode = TPCF_MODE_PLAIN;
            else if (!std::strcmp(optarg, "fast64")) mode = TPCF_MODE_FAST64;
            else if (!std::strcmp(optarg, "cells")) mode = TPCF_MODE_CELLS;
            else { std::fprintf(stderr, "error: --mode fast|plain|fast64|cells\n"); return 2; }
            break;
        case O_FORMAT: format = std::strcmp(optarg, "plain") == 0 ? 1 : 0; break;
        case O_RECHECK: spec.recheck_rel = std::atof(optarg); break;
        case O_DEVICE: device = std::atoi(optarg); break;
        case O_QUIET: quiet = 1; break;
        case O_STATS: want_stats = 1; break;
        case O_ALLVOX: all_voxels = 1; break;
        case '?': std::printf("unknown arg %c\n", optopt); break;  // the reference continues
        default: break;
        }
    }
    if (mode < 0) mode = TPCF_MODE_FAST;
    if (have_r123_range) {
        if (r_n > 0) spec.ax[0].n = r_n;
        if (r_hi > r_lo) {
            spec.ax[0].lo = r_lo;
            spec.ax[0].hi = r_hi;
        }
        spec.ax[0].scale = r_log ? TPCF_LOG : TPCF_LIN;
    }
    if (argc - optind < 3) {
        std::printf("\nMust pass in at least three input files on command line!\n");
        return 1;
    }
    if (format == 1 && spec.param != TPCF_R123) {
        std::fprintf(stderr, "error: the cube format holds (r1,r2,r3) histograms only\n");
        return 2;
    }
    if (tpcf_binspec_check(&spec) != TPCF_OK) return die(2, "bad bin spec");
    if (all_voxels && (vdiv <= 0 || vdiv > 999 || !std::strstr(outfile, "{index}"))) {
        std::fprintf(stderr, "error: --all-voxels needs -X DIV (1..999) and an -o name containing {index}\n");
        return 2;
    }

    // Catalogs, deduplicated by file identity.
    Catalog cat[3];
    for (int i = 0; i < 3; ++i) {
        cat[i].path = argv[optind + i];
        char rp[PATH_MAX];
        if (!realpath(cat[i].path.c_str(), rp)) {
            std::fprintf(stderr, "error: cannot open %s\n", cat[i].path.c_str());
            return 1;
        }
        cat[i].real = rp;
    }
    for (int i = 0; i < 3; ++i) {
        int same = -1;
        for (int j = 0; j < i; ++j)
            if (cat[j].real == cat[i].real) same = j;
        if (same >= 0) {
            cat[i].xyz = cat[same].xyz;
            cat[i].n = cat[same].n;
            continue;
        }
        cat[i].n = tpcf_read_catalog(cat[i].path.c_str(), triprog ? 7 : 3, scale, &cat[i].xyz);
        if (cat[i].n < 0) return die(1, "cannot read catalog");
        if (!quiet) std::printf("Opening input file %d: %s (%lld points)\n", i, cat[i].path.c_str(), (long long)cat[i].n);
    }

    // One device copy per distinct file.
    tpcf_ctx *ctx = nullptr;
    if (tpcf_ctx_create(device, &ctx) != TPCF_OK) return die(3, "no usable GPU");
    tpcf_catalog *dev[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < i; ++j)
            if (cat[j].real == cat[i].real) dev[i] = dev[j];
        if (!dev[i] && tpcf_catalog_create(ctx, cat[i].xyz, cat[i].n, &dev[i]) != TPCF_OK)
            return die(3, "catalog upload failed");
    }
    const int64_t cells = tpcf_hist_cells(&spec);

    if (all_voxels) {
        // Every chunk triple (i,j,k) of -X div, as the per-voxel driver loop
        // (examples/run_CPU_calculation_by_voxels.csh:61-85) runs them, in one
        // batch; one file per voxel, named from the -o template.
        const std::string tmpl = outfile;
        const size_t at = tmpl.find("{index}");
        const int nvox = vdiv * vdiv * vdiv;
        std::vector<tpcf_job> jobs;
        std::vector<int> job_of(nvox, -1);
        for (int v = 0; v < nvox; ++v) {
            const int vid[3] = {v / (vdiv * vdiv), (v / vdiv) % vdiv, v % vdiv};
            int64_t lo[3], hi[3];
            if (!voxel_ranges(cat, vdiv, vid, lo, hi)) {
                std::printf("Voxel chunks are bigger than the number of galaxies!");
                return 255;
            }
            if (selects_nothing(cat, lo)) continue;
            job_of[v] = (int)jobs.size();
            jobs.emplace_back();
            voxel_job(cat, dev, lo, hi, jobs.back());
        }
        std::vector<uint64_t> h((size_t)jobs.size() * cells, 0);
        std::vector<uint64_t *> hp(jobs.size());
        for (size_t j = 0; j < jobs.size(); ++j) hp[j] = h.data() + j * cells;
        const std::vector<tpcf_binspec> specs(jobs.size(), spec);
        tpcf_stats st{};
        if (!jobs.empty()) {
            int rc = TPCF_OK;
            if (mode == TPCF_MODE_PLAIN) {
                for (size_t j = 0; j < jobs.size() && rc == TPCF_OK; ++j) {
                    tpcf_stats sj{};
                    rc = tpcf_count(ctx, &jobs[j], &spec, mode, hp[j], &sj);
                    st.triplets += sj.triplets;
                    st.pairs += sj.pairs;
                    st.pairs_rechecked += sj.pairs_rechecked;
                    st.device_ms += sj.device_ms;
                    st.wall_ms += sj.wall_ms;
                }
            } else {
                rc = tpcf_count_batch(ctx, jobs.data(), specs.data(), (int)jobs.size(), mode, hp.data(), &st);
            }
            if (rc != TPCF_OK) return die(4, "counting failed");
        }
        const std::vector<uint64_t> zero(cells, 0);
        for (int v = 0; v < nvox; ++v) {
            char index[16];
            std::snprintf(index, sizeof index, "%03d%03d%03d", v / (vdiv * vdiv), (v / vdiv) % vdiv, v % vdiv);
            const std::string path = tmpl.substr(0, at) + index + tmpl.substr(at + 7);
            const uint64_t *hv = job_of[v] >= 0 ? hp[job_of[v]] : zero.data();
            if (!write_hist(path.c_str(), format, spec, cat, hv)) return die(1, "cannot write the histogram file");
            if (!quiet) {
                uint64_t total = 0;
                for (int64_t c = 0; c < cells; ++c) total += hv[c];
                std::printf("Output filename is %s\nTotal: %llu\n", path.c_str(), (unsigned long long)total);
            }
        }
        if (!quiet) std::printf("numcalc: %llu\n", (unsigned long long)st.triplets);
        if (want_stats) print_stats(st);
    } else {
        int vid[3] = {0, 0, 0};
        if (vdiv > 0) {
            vid[0] = vtemp / 1000000;
            vid[1] = (vtemp / 1000) % 1000;
            vid[2] = vtemp % 1000;
        }
        int64_t lo[3], hi[3];
        if (!voxel_ranges(cat, vdiv > 0 && vtemp >= 0 ? vdiv : 0, vid, lo, hi)) {
            std::printf("Voxel chunks are bigger than the number of galaxies!");
            return 255;
        }
        tpcf_job job{};
        voxel_job(cat, dev, lo, hi, job);
        std::vector<uint64_t> h(cells, 0);
        tpcf_stats st{};
        if (!selects_nothing(cat, lo) && tpcf_count(ctx, &job, &spec, mode, h.data(), &st) != TPCF_OK)
            return die(4, "counting failed");
        uint64_t total = 0;
        for (auto v : h) total += v;
        if (!write_hist(outfile, format, spec, cat, h.data())) return die(1, "cannot write the histogram file");
        if (!quiet) {
            std::printf("Output filename is %s\n", outfile);
            std::printf("Total: %llu\n", (unsigned long long)total);
            std::printf("numcalc: %llu\n", (unsigned long long)st.triplets);
        }
        if (want_stats) print_stats(st);
    }
    for (int i = 0; i < 3; ++i) {
        bool first = true;
        for (int j = 0; j < i; ++j)
            if (dev[j] == dev[i]) first = false;
        if (first && dev[i]) tpcf_catalog_destroy(dev[i]);
    }
    for (int i = 0; i < 3; ++i) {
        bool first = true;
        for (int j = 0; j < i; ++j)
            if (cat[j].xyz == cat[i].xyz) first = false;
        if (first) tpcf_free(cat[i].xyz);
    }
    tpcf_ctx_destroy(ctx);
    return 0;
}
