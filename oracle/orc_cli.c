/*
 * orc_cli.c -- command-line front end of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see tpcf_oracle.h).  It lets the tests drive the
 * oracle exactly as the reference programs are driven: three catalog paths,
 * getopt flags, one histogram file out.
 *
 *   orc_cli ref-tri   [opts] f0 f1 f2   src/3pcf_C_version_triangle_parametrization.cc:245-571
 *   orc_cli ref-jbj   [opts] f0 f1 f2   src/3pcf_C_version_triangle_parametrization_JBJ.cc
 *   orc_cli ref-plain [opts] f0 f1 f2   src/3pcf_C_version.cc:147-466
 *   orc_cli clean-r123|clean-sqt|clean-jbj [opts] f0 f1 f2   (clean64 semantics)
 *   orc_cli ref-two   [opts] f0 f1      src/2pcf_C_version.cc:124-411
 *   orc_cli clean-two [opts] f0 f1      pair counts, clean64 semantics
 *
 * Options: -o out, -X div, -x AAABBBCCC, -s, -m, --threads N, --kmin-legacy,
 *   --axis IDX,N,LO,HI[,log]   bin spec of axis IDX (0..2),
 *   --count-format              catalogs carry a count header (src/3pcf_C_version.cc)
 */
#define _GNU_SOURCE
#include <getopt.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "tpcf_oracle.h"

static int parse_axis(const char *arg, int *idx, int *n, double *lo, double *hi, int *lg)
{
    char buf[256];
    snprintf(buf, sizeof buf, "%s", arg);
    char *tok[5];
    int nt = 0;
    for (char *p = strtok(buf, ","); p && nt < 5; p = strtok(NULL, ",")) tok[nt++] = p;
    if (nt < 4) return -1;
    *idx = atoi(tok[0]);
    *n = atoi(tok[1]);
    *lo = atof(tok[2]);
    *hi = atof(tok[3]);
    *lg = nt == 5 && strcmp(tok[4], "log") == 0;
    return (*idx < 0 || *idx > 2) ? -1 : 0;
}

/* Pair counts: two catalogs, -X/-x decode two voxel indices
 * (src/2pcf_C_version.cc:187-193). */
static int two_main(const char *mode, int npath, char **paths, const char *out, int vdiv, int vtemp, int nthreads,
                    int an, double alo, double ahi, int alog)
{
    if (npath < 2) {
        fprintf(stderr, "need two catalog paths\n");
        return 2;
    }
    const char *path[2] = {paths[0], paths[1]};
    int vidx[2] = {0, 0};
    if (vdiv > 0) {
        vidx[0] = vtemp / 1000000;
        vidx[1] = (vtemp / 1000) % 1000;
    }
    const int is_ref = strcmp(mode, "ref-two") == 0;
    void *cat[2];
    int64_t ng[2];
    for (int i = 0; i < 2; i++) {
        ng[i] = orc_read_7col(path[i], !is_ref, &cat[i]);
        if (ng[i] < 0) {
            fprintf(stderr, "cannot read %s\n", path[i]);
            return 1;
        }
    }
    int lo[2], hi[2];
    for (int i = 0; i < 2; i++) {
        lo[i] = 0;
        hi[i] = (int)ng[i];
        if (vdiv > 0 && vtemp >= 0) {
            int lh[2];
            if (orc_ref_vox2gal(vdiv, vidx[i], (int)ng[i], lh)) return 1;
            lo[i] = lh[0];
            hi[i] = lh[1];
        }
    }
    if (is_ref) {
        const float *p[2];
        int n[2];
        for (int i = 0; i < 2; i++) {
            const float *P = cat[i];
            int m = hi[i] - lo[i];
            float *Q = malloc(sizeof(float) * 3 * (size_t)(m > 0 ? m : 1));
            for (int t = 0; t < m; t++) {
                Q[t] = P[lo[i] + t];
                Q[m + t] = P[ng[i] + lo[i] + t];
                Q[2 * m + t] = P[2 * ng[i] + lo[i] + t];
            }
            p[i] = Q;
            n[i] = m;
        }
        uint32_t h32[700];
        orc_ref_two(p, n, strcmp(path[0], path[1]) == 0 ? 0 : 1, h32, nthreads);
        uint64_t h[700];
        for (int b = 0; b < 700; b++) h[b] = h32[b];
        return orc_write_two(out, ng, 700, 0.0, 70.0, (70.0 - 0.0) / 700, h) ? 1 : 0;
    }
    orc_axis ax = {700, 0.0, 70.0, ORC_LIN};
    if (an > 0) ax = (orc_axis){an, alo, ahi, alog ? ORC_LOG : ORC_LIN};
    char *rp0 = realpath(path[0], NULL), *rp1 = realpath(path[1], NULL);
    if (!rp0 || !rp1) return 1;
    const int same_file = strcmp(rp0, rp1) == 0;
    const int same = same_file && lo[0] == lo[1];
    const int empty = same_file && lo[0] > lo[1];
    double *seg[2];
    int64_t nseg[2];
    for (int i = 0; i < 2; i++) {
        int m = hi[i] - lo[i];
        const double *P = cat[i];
        seg[i] = malloc(sizeof(double) * 3 * (size_t)(m > 0 ? m : 1));
        for (int t = 0; t < m; t++) {
            seg[i][t] = P[lo[i] + t];
            seg[i][m + t] = P[ng[i] + lo[i] + t];
            seg[i][2 * m + t] = P[2 * ng[i] + lo[i] + t];
        }
        nseg[i] = m;
    }
    uint64_t *h = calloc((size_t)ax.n, 8);
    const int term[1][2] = {{0, same ? 0 : 1}};
    const double *sp[2] = {seg[0], seg[1]};
    if (!empty && orc_clean_pairs(sp, nseg, same ? 1 : 2, term, 1, &ax, h, nthreads)) {
        fprintf(stderr, "bad job\n");
        return 1;
    }
    const double w = ax.scale == ORC_LOG ? log(ax.hi / ax.lo) / ax.n : (ax.hi - ax.lo) / ax.n;
    return orc_write_two(out, ng, ax.n, ax.lo, ax.hi, w, h) ? 1 : 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: orc_cli MODE [opts] f0 f1 f2\n");
        return 2;
    }
    const char *mode = argv[1];
    argv++;
    argc--;
    const char *out = "default_out.dat";
    int vdiv = -999, vtemp = -999, nthreads = 1, kmin_legacy = 0, count_fmt = 0;
    double scale = 1.0;
    int ax_n[3] = {-1, -1, -1}, ax_log[3] = {0, 0, 0};
    double ax_lo[3] = {0, 0, 0}, ax_hi[3] = {0, 0, 0};
    static struct option lo_opts[] = {{"threads", required_argument, 0, 't'},
                                      {"kmin-legacy", no_argument, 0, 'K'},
                                      {"axis", required_argument, 0, 'A'},
                                      {"count-format", no_argument, 0, 'C'},
                                      {0, 0, 0, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "ao:X:x:sm", lo_opts, NULL)) != -1) {
        switch (c) {
        case 'o': out = optarg; break;
        case 'X': vdiv = atoi(optarg); break;
        case 'x': vtemp = atoi(optarg); break;
        case 's': scale = 206264.0; break;    /* src/3pcf_C_version.cc:195-201 */
        case 'm': scale = 3437.74677; break;  /* src/3pcf_C_version.cc:202-208 */
        case 't': nthreads = atoi(optarg); break;
        case 'K': kmin_legacy = 1; break;
        case 'C': count_fmt = 1; break;
        case 'A': {
            int i, n, lg;
            double l, h;
            if (parse_axis(optarg, &i, &n, &l, &h, &lg)) return 2;
            ax_n[i] = n; ax_lo[i] = l; ax_hi[i] = h; ax_log[i] = lg;
            break;
        }
        default: break;
        }
    }
    if (strcmp(mode, "ref-two") == 0 || strcmp(mode, "clean-two") == 0)
        return two_main(mode, argc - optind, argv + optind, out, vdiv, vtemp, nthreads, ax_n[0], ax_lo[0],
                        ax_hi[0], ax_log[0]);
    if (argc - optind < 3) {
        fprintf(stderr, "need three catalog paths\n");
        return 2;
    }
    const char *path[3] = {argv[optind], argv[optind + 1], argv[optind + 2]};
    int vidx[3] = {0, 0, 0};
    if (vdiv > 0) {   /* TRI:306-313 */
        vidx[0] = vtemp / 1000000;
        vidx[1] = (vtemp / 1000) % 1000;
        vidx[2] = vtemp % 1000;
    }

    int is_ref = strncmp(mode, "ref-", 4) == 0;
    int as_double = !is_ref;
    void *cat[3];
    int64_t ng[3];
    for (int i = 0; i < 3; i++) {
        ng[i] = (count_fmt || strcmp(mode, "ref-plain") == 0)
                    ? orc_read_count(path[i], as_double, scale, &cat[i])
                    : orc_read_7col(path[i], as_double, &cat[i]);
        if (ng[i] < 0) {
            fprintf(stderr, "cannot read %s\n", path[i]);
            return 1;
        }
    }

    if (strcmp(mode, "ref-plain") == 0) {
        int nb = ax_n[0] > 0 ? ax_n[0] : 64;
        float hmin = ax_n[0] > 0 ? (float)ax_lo[0] : 0.0f;
        float hmax = ax_n[0] > 0 ? (float)ax_hi[0] : (float)__builtin_sqrt(3 * 24 * 24);
        const float *p[3] = {cat[0], cat[1], cat[2]};
        int n[3] = {(int)ng[0], (int)ng[1], (int)ng[2]};
        uint64_t *h = calloc((size_t)nb * nb * nb, 8), oob = 0;
        orc_ref_plain(p, n, nb, hmin, hmax, h, &oob, nthreads);
        printf("oob: %llu\n", (unsigned long long)oob);
        return orc_write_plain(out, nb, h) ? 1 : 0;
    }

    if (is_ref) {
        /* TRI:31-42 defaults; THETA_NBINS is 32 in the current source. */
        orc_ref_axis ax[3] = {{50, 2.0, 12.0}, {16, 0.9, 4.1}, {32, 0.0, 1.0}};
        for (int a = 0; a < 3; a++)
            if (ax_n[a] > 0) ax[a] = (orc_ref_axis){ax_n[a], ax_lo[a], ax_hi[a]};
        const float *p[3];
        int n[3];
        for (int i = 0; i < 3; i++) {
            const float *P = cat[i];
            int lo = 0, hi = (int)ng[i];
            if (vdiv > 0 && vtemp >= 0) {   /* TRI:408-417 */
                int lh[2];
                if (orc_ref_vox2gal(vdiv, vidx[i], (int)ng[i], lh)) return 1;
                lo = lh[0];
                hi = lh[1];
            }
            float *Q = malloc(sizeof(float) * 3 * (size_t)(hi - lo > 0 ? hi - lo : 1));
            int m = hi - lo;
            for (int t = 0; t < m; t++) {
                Q[t] = P[lo + t];
                Q[m + t] = P[ng[i] + lo + t];
                Q[2 * m + t] = P[2 * ng[i] + lo + t];
            }
            p[i] = Q;
            n[i] = m;
        }
        int which = orc_ref_which(path[0], path[1], path[2]);
        int64_t cells = (int64_t)ax[0].n * ax[1].n * ax[2].n;
        uint32_t *h32 = calloc((size_t)cells, 4);
        uint64_t numcalc = 0;
        orc_ref_tri(p, n, which, strcmp(mode, "ref-jbj") == 0, kmin_legacy, ax, h32, &numcalc,
                    nthreads);
        uint64_t *h = calloc((size_t)cells, 8), total = 0;
        for (int64_t x = 0; x < cells; x++) {
            h[x] = h32[x];
            total += h32[x];
        }
        printf("Total: %llu\nnumcalc: %llu\n", (unsigned long long)total,
               (unsigned long long)numcalc);
        int nb[3] = {ax[0].n, ax[1].n, ax[2].n};
        double lo[3], hi[3], w[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = ax[a].lo;
            hi[a] = ax[a].hi;
            w[a] = (ax[a].hi - ax[a].lo) / ax[a].n;
        }
        return orc_write_tri(out, ng, nb, lo, hi, w, h) ? 1 : 0;
    }

    /* clean64: segments are the distinct (file, chunk) pairs. */
    orc_job job;
    memset(&job, 0, sizeof job);
    if (strcmp(mode, "clean-r123") == 0) job.param = ORC_R123;
    else if (strcmp(mode, "clean-sqt") == 0) job.param = ORC_SQT;
    else if (strcmp(mode, "clean-jbj") == 0) job.param = ORC_JBJ;
    else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 2;
    }
    orc_axis defaults[3] = {{50, 2.0, 12.0, ORC_LIN}, {16, 0.9, 4.1, ORC_LIN}, {32, 0.0, 1.0, ORC_LIN}};
    if (job.param == ORC_R123) defaults[0] = (orc_axis){16, 0.0, 40.0, ORC_LIN};
    for (int a = 0; a < 3; a++) {
        job.ax[a] = defaults[a];
        if (ax_n[a] > 0) job.ax[a] = (orc_axis){ax_n[a], ax_lo[a], ax_hi[a], ax_log[a] ? ORC_LOG : ORC_LIN};
    }
    char *rp[3];
    for (int i = 0; i < 3; i++) {
        rp[i] = realpath(path[i], NULL);
        if (!rp[i]) return 1;
    }
    int chunk_lo[3], chunk_hi[3];
    for (int i = 0; i < 3; i++) {
        chunk_lo[i] = 0;
        chunk_hi[i] = (int)ng[i];
        if (vdiv > 0 && vtemp >= 0) {
            int lh[2];
            if (orc_ref_vox2gal(vdiv, vidx[i], (int)ng[i], lh)) return 1;
            chunk_lo[i] = lh[0];
            chunk_hi[i] = lh[1];
        }
    }
    /* Roles of the same file must be ordered by chunk; a later role in an
     * earlier chunk selects nothing. */
    int empty = 0;
    for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
            if (strcmp(rp[a], rp[b]) == 0 && chunk_lo[a] > chunk_lo[b]) empty = 1;
    /* Segment labels by (file, chunk) identity in role order. */
    int label[3] = {-1, -1, -1};
    int seg_role[3];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < i; j++)
            if (strcmp(rp[i], rp[j]) == 0 && chunk_lo[i] == chunk_lo[j] && label[j] >= 0) {
                label[i] = label[j];
                break;
            }
        if (label[i] < 0) {
            label[i] = job.nseg;
            seg_role[job.nseg] = i;
            job.nseg++;
        }
    }
    double *segbuf[3];
    for (int s = 0; s < job.nseg; s++) {
        int i = seg_role[s];
        int m = chunk_hi[i] - chunk_lo[i];
        const double *P = cat[i];
        segbuf[s] = malloc(sizeof(double) * 3 * (size_t)(m > 0 ? m : 1));
        for (int t = 0; t < m; t++) {
            segbuf[s][t] = P[chunk_lo[i] + t];
            segbuf[s][m + t] = P[ng[i] + chunk_lo[i] + t];
            segbuf[s][2 * m + t] = P[2 * ng[i] + chunk_lo[i] + t];
        }
        job.seg[s] = segbuf[s];
        job.nseg_pts[s] = m;
    }
    int t3[3] = {label[0], label[1], label[2]}, tmp;
    if (t3[0] > t3[1]) { tmp = t3[0]; t3[0] = t3[1]; t3[1] = tmp; }
    if (t3[1] > t3[2]) { tmp = t3[1]; t3[1] = t3[2]; t3[2] = tmp; }
    if (t3[0] > t3[1]) { tmp = t3[0]; t3[0] = t3[1]; t3[1] = tmp; }
    job.nterms = 1;
    memcpy(job.term[0], t3, sizeof t3);
    int64_t cells = orc_hist_cells(&job);
    uint64_t *h = calloc((size_t)cells, 8);
    if (!empty && orc_clean_count(&job, h, nthreads)) {
        fprintf(stderr, "bad job\n");
        return 1;
    }
    int nb[3];
    double lo[3], hi[3], w[3];
    for (int a = 0; a < 3; a++) {
        const orc_axis *A = job.param == ORC_R123 ? &job.ax[0] : &job.ax[a];
        nb[a] = A->n;
        lo[a] = A->lo;
        hi[a] = A->hi;
        w[a] = A->scale == ORC_LOG ? log(A->hi / A->lo) / A->n : (A->hi - A->lo) / A->n;
    }
    return orc_write_tri(out, ng, nb, lo, hi, w, h) ? 1 : 0;
}
