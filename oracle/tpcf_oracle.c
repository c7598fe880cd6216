/*
 * tpcf_oracle.c -- CPU oracle for the brute-force triplet-counting path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product never links it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math; the
 * refcompat32 functions depend on IEEE single-precision SSE arithmetic with no
 * FMA contraction, exactly what the reference's g++ -O1 build on x86-64 does).
 *
 * Reference citations are file:line into the mattbellis/3pcf sources.
 */
#define _GNU_SOURCE
#include "tpcf_oracle.h"

#include <errno.h>
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* Catalog readers                                                          */
/* ======================================================================== */

static int grow(void **buf, int64_t *cap, int64_t need, size_t elem)
{
    if (need <= *cap) return 0;
    int64_t nc = *cap ? *cap * 2 : 4096;
    while (nc < need) nc *= 2;
    void *p = realloc(*buf, (size_t)nc * elem);
    if (!p) return -1;
    *buf = p;
    *cap = nc;
    return 0;
}

/* Repack interleaved xyz triples into SoA [x|y|z]. */
static void *to_soa(const void *xyz, int64_t n, size_t elem)
{
    void *out = malloc((size_t)(3 * (n > 0 ? n : 1)) * elem);
    if (!out) return NULL;
    const char *src = (const char *)xyz;
    char *dst = (char *)out;
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++)
            memcpy(dst + ((size_t)c * n + i) * elem, src + ((size_t)i * 3 + c) * elem, elem);
    return out;
}

/* The reference loops `while (fscanf(...) != EOF)` (TRI:377) and keeps
 * columns 5..7 of every matched line.  A short match that is not EOF would
 * make the reference spin forever; here it is an error. */
int64_t orc_read_7col(const char *path, int as_double, void **out)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    size_t elem = as_double ? sizeof(double) : sizeof(float);
    void *buf = NULL;
    int64_t cap = 0, n = 0;
    for (;;) {
        int id;
        int r;
        if (grow(&buf, &cap, 3 * (n + 1), elem)) { fclose(f); free(buf); return -2; }
        if (as_double) {
            double d0, d1, d2, v[3];
            r = fscanf(f, "%d %lf %lf %lf %lf %lf %lf", &id, &d0, &d1, &d2, &v[0], &v[1], &v[2]);
            if (r == 7) memcpy((double *)buf + 3 * n, v, sizeof v);
        } else {
            float d0, d1, d2, v[3];
            r = fscanf(f, "%d %f %f %f %f %f %f", &id, &d0, &d1, &d2, &v[0], &v[1], &v[2]);
            if (r == 7) memcpy((float *)buf + 3 * n, v, sizeof v);
        }
        if (r == EOF) break;
        if (r != 7) { fclose(f); free(buf); return -3; }
        n++;
    }
    fclose(f);
    *out = to_soa(buf, n, elem);
    free(buf);
    return *out ? n : -2;
}

/* src/3pcf_C_version.cc:311-332: "%d" count, then count lines of "%f %f %f",
 * each divided by the float scale_factor. */
int64_t orc_read_count(const char *path, int as_double, double scale, void **out)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int n;
    if (fscanf(f, "%d", &n) != 1 || n < 0) { fclose(f); return -3; }
    size_t elem = as_double ? sizeof(double) : sizeof(float);
    void *buf = malloc((size_t)(3 * (n > 0 ? n : 1)) * elem);
    if (!buf) { fclose(f); return -2; }
    for (int i = 0; i < n; i++) {
        if (as_double) {
            double v[3];
            if (fscanf(f, "%lf %lf %lf", &v[0], &v[1], &v[2]) != 3) { fclose(f); free(buf); return -3; }
            for (int c = 0; c < 3; c++) ((double *)buf)[3 * i + c] = v[c] / scale;
        } else {
            float v[3];
            float sf = (float)scale;
            if (fscanf(f, "%f %f %f", &v[0], &v[1], &v[2]) != 3) { fclose(f); free(buf); return -3; }
            for (int c = 0; c < 3; c++) ((float *)buf)[3 * i + c] = v[c] / sf;
        }
    }
    fclose(f);
    *out = to_soa(buf, n, elem);
    free(buf);
    return *out ? n : -2;
}

void orc_free(void *p) { free(p); }

/* ======================================================================== */
/* Thread pool helper: dynamic scheduling of an outer index range.          */
/* ======================================================================== */

typedef struct {
    atomic_llong next;
    long long end;
    int chunk;
    void *ctx;
    void (*body)(void *ctx, int tid, long long lo, long long hi);
} orc_pool;

typedef struct {
    orc_pool *pool;
    int tid;
} orc_worker;

static void *pool_main(void *arg)
{
    orc_worker *w = (orc_worker *)arg;
    orc_pool *p = w->pool;
    for (;;) {
        long long lo = atomic_fetch_add(&p->next, p->chunk);
        if (lo >= p->end) break;
        long long hi = lo + p->chunk < p->end ? lo + p->chunk : p->end;
        p->body(p->ctx, w->tid, lo, hi);
    }
    return NULL;
}

static void run_pool(long long begin, long long end, int chunk, int nthreads, void *ctx,
                     void (*body)(void *, int, long long, long long))
{
    if (nthreads < 1) nthreads = 1;
    orc_pool p;
    atomic_init(&p.next, begin);
    p.end = end;
    p.chunk = chunk < 1 ? 1 : chunk;
    p.ctx = ctx;
    p.body = body;
    pthread_t th[256];
    orc_worker w[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; t++) {
        w[t].pool = &p;
        w[t].tid = t;
    }
    for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, pool_main, &w[t]);
    pool_main(&w[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ======================================================================== */
/* refcompat32                                                              */
/* ======================================================================== */

int orc_ref_which(const char *p0, const char *p1, const char *p2)
{
    /* TRI:333-353.  The selector is a bool: assignments of 2 become 1.  The
     * third test (`!=0 && ==0` on the same pair) can never be true. */
    int w = 0;
    if (strcmp(p0, p1) == 0 && strcmp(p0, p2) == 0) w = 0;
    else if (strcmp(p0, p1) != 0 && strcmp(p1, p2) == 0) w = 1;
    else if (strcmp(p0, p1) == 0 && strcmp(p0, p2) != 0) w = 1;
    return w;
}

int orc_ref_vox2gal(int division, int index, int ngals, int lohi[2])
{
    int per = ngals / division;   /* TRI:54 */
    if (per == 0) return -1;      /* TRI:59-63 exits */
    lohi[0] = per * index;
    lohi[1] = per * (index + 1);
    if (index == division - 1) lohi[1] += ngals % division;   /* TRI:70-73 */
    return 0;
}

/* distance_to_bin of TRI:84-113 with flag 0: float width, truncation toward
 * zero, out of range -> -999.  A NaN value passes both range tests and the
 * x86 truncating conversion yields INT_MIN, i.e. a dropped entry. */
static inline int ref_bin(float v, float lo, float hi, int n)
{
    float w = (hi - lo) / (float)n;
    if (v < lo) return -999;
    if (v >= hi) return -999;
    float t = (v - lo) / w;
    if (t != t) return INT_MIN;
    return (int)t;
}

static inline float ref_dist(float ax, float ay, float az, float bx, float by, float bz)
{
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    return sqrtf(dx * dx + dy * dy + dz * dz);   /* TRI:119-122, float throughout */
}

typedef struct {
    const float *p[3];
    int n[3];
    int which, variant, kmin_legacy;
    float lo[3], hi[3];
    int nb[3];
    int64_t cells;
    uint32_t **th_hist;
    uint64_t *th_numcalc;
} ref_tri_ctx;

/* One (s,q,theta) triplet; mirrors distance() of TRI:116-240 (variant 0) or
 * TRIJ:116-229 (variant 1).  Writes up to 3 histogram indices, -1 = none. */
static inline void ref_tri_triplet(const ref_tri_ctx *c, float x0, float y0, float z0,
                                   float x1, float y1, float z1, float x2, float y2, float z2,
                                   int out[3])
{
    float d0 = ref_dist(x0, y0, z0, x1, y1, z1);
    float d1 = ref_dist(x0, y0, z0, x2, y2, z2);
    float d2 = ref_dist(x1, y1, z1, x2, y2, z2);
    out[0] = out[1] = out[2] = -1;
    if (c->variant == 1) {
        /* TRIJ's distance() writes only totbins[0] (TRIJ:192 loops k<1); the
         * caller's bins[1], bins[2] keep their initial 0 (TRIJ main), so every
         * iteration also increments cell 0 twice. */
        out[1] = out[2] = 0;
    }
    if (d0 == 0 || d1 == 0 || d2 == 0) return;   /* TRI:230-234 */

    /* Sort by value (TRI:146-181); only the values matter. */
    float a = d0, b = d1, cc = d2, t;
    if (a > b) { t = a; a = b; b = t; }
    if (b > cc) { t = b; b = cc; cc = t; }
    if (a > b) { t = a; a = b; b = t; }
    const float sh = a, mi = b, lo = cc;
    const float sh2 = sh * sh, mi2 = mi * mi, lo2 = lo * lo;

    const int nq = c->nb[1], nt = c->nb[2];
    if (c->variant == 1) {
        /* TRIJ:192-198: s = shortest, q = shortest/middle,
         * theta = (longest-middle)/shortest. */
        float s = sh, q = sh / mi, th = (lo - mi) / sh;
        int is = ref_bin(s, c->lo[0], c->hi[0], c->nb[0]);
        int iq = ref_bin(q, c->lo[1], c->hi[1], nq);
        int it = ref_bin(th, c->lo[2], c->hi[2], nt);
        if (is >= 0 && iq >= 0 && it >= 0) out[0] = nq * nt * is + nt * iq + it;
        return;
    }
    /* TRI:192-210.  theta/pi is (double)acosf(arg) * 1.0 / 3.14159, stored
     * as float; the third angle is 1.0 - theta0 - theta1 in double. */
    float th0 = (float)(((double)acosf((sh2 + mi2 - lo2) / (2 * sh * mi)) * 1.0) / 3.14159);
    float th1 = (float)(((double)acosf((mi2 + lo2 - sh2) / (2 * mi * lo)) * 1.0) / 3.14159);
    float th2 = (float)((1.0 - (double)th0) - (double)th1);
    const float sv[3] = {sh, mi, sh};
    const float qv[3] = {mi / sh, lo / mi, lo / sh};
    const float tv[3] = {th0, th1, th2};
    for (int k = 0; k < 3; k++) {
        int it = ref_bin(tv[k], c->lo[2], c->hi[2], nt);
        int is = ref_bin(sv[k], c->lo[0], c->hi[0], c->nb[0]);
        int iq = ref_bin(qv[k], c->lo[1], c->hi[1], nq);
        if (is >= 0 && iq >= 0 && it >= 0) out[k] = nq * nt * is + nt * iq + it;
    }
}

static void ref_tri_body(void *vctx, int tid, long long ilo, long long ihi)
{
    ref_tri_ctx *c = (ref_tri_ctx *)vctx;
    uint32_t *h = c->th_hist[tid];
    uint64_t calc = 0;
    const float *P0 = c->p[0], *P1 = c->p[1], *P2 = c->p[2];
    const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
    for (long long i = ilo; i < ihi; i++) {
        /* TRI:484-509 with the bool selector: 0 -> j from i, k from j
         * (legacy: k from j+1); 1 -> j from 0, k from j. */
        int jmin = c->which == 0 ? (int)i : 0;
        for (int j = jmin; j < n1; j++) {
            int kmin = j;
            if (c->which == 0 && c->kmin_legacy) kmin = j + 1;
            for (int k = kmin; k < n2; k++) {
                int idx[3];
                ref_tri_triplet(c, P0[i], P0[n0 + i], P0[2 * n0 + i], P1[j], P1[n1 + j],
                                P1[2 * n1 + j], P2[k], P2[n2 + k], P2[2 * n2 + k], idx);
                calc++;
                for (int b = 0; b < 3; b++)
                    if (idx[b] >= 0) h[idx[b]]++;
            }
        }
    }
    c->th_numcalc[tid] += calc;
}

int orc_ref_tri(const float *const pts[3], const int n[3], int which, int variant,
                int kmin_legacy, const orc_ref_axis ax[3], uint32_t *hist,
                uint64_t *numcalc, int nthreads)
{
    ref_tri_ctx c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < 3; i++) {
        c.p[i] = pts[i];
        c.n[i] = n[i];
        /* The #define values reach distance_to_bin as float parameters. */
        c.lo[i] = (float)ax[i].lo;
        c.hi[i] = (float)ax[i].hi;
        c.nb[i] = ax[i].n;
        if (ax[i].n <= 0) return -1;
    }
    c.which = which;
    c.variant = variant;
    c.kmin_legacy = kmin_legacy;
    c.cells = (int64_t)ax[0].n * ax[1].n * ax[2].n;
    if (nthreads < 1) nthreads = 1;
    c.th_hist = (uint32_t **)calloc((size_t)nthreads, sizeof(uint32_t *));
    c.th_numcalc = (uint64_t *)calloc((size_t)nthreads, sizeof(uint64_t));
    for (int t = 0; t < nthreads; t++) c.th_hist[t] = (uint32_t *)calloc((size_t)c.cells, 4);
    run_pool(0, n[0], 1, nthreads, &c, ref_tri_body);
    memset(hist, 0, (size_t)c.cells * 4);
    uint64_t calc = 0;
    for (int t = 0; t < nthreads; t++) {
        /* unsigned int addition wraps mod 2^32 exactly as the sequential
         * reference counters do. */
        for (int64_t x = 0; x < c.cells; x++) hist[x] += c.th_hist[t][x];
        calc += c.th_numcalc[t];
        free(c.th_hist[t]);
    }
    free(c.th_hist);
    free(c.th_numcalc);
    if (numcalc) *numcalc = calc;
    return 0;
}

/* src/3pcf_C_version.cc:23-51 with flag 0 (the flag is fixed at 0, :175 and
 * :400): under -> 0, over -> nbins+1, else int((d-lo)/w). */
static inline int ref_plain_bin(float d, float lo, float hi, int nbins, float w)
{
    if (d < lo) return 0;
    if (d >= hi) return nbins + 1;
    float t = (d - lo) / w;
    if (t != t) return INT_MIN;
    return (int)t;
}

typedef struct {
    const float *p[3];
    int n[3];
    int nbins;
    float lo, hi, w;
    uint64_t **th_hist;
    uint64_t *th_oob;
} ref_plain_ctx;

static void ref_plain_body(void *vctx, int tid, long long ilo, long long ihi)
{
    ref_plain_ctx *c = (ref_plain_ctx *)vctx;
    uint64_t *h = c->th_hist[tid];
    const int nb = c->nbins;
    const int64_t cells = (int64_t)nb * nb * nb;
    const float *P0 = c->p[0], *P1 = c->p[1], *P2 = c->p[2];
    const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
    uint64_t oob = 0;
    for (long long i = ilo; i < ihi; i++) {
        for (int j = 0; j < n1; j++) {       /* PLAIN:380 ignores jmin */
            for (int k = 0; k < n2; k++) {   /* PLAIN:392 ignores kmin */
                float x0 = P0[i], y0 = P0[n0 + i], z0 = P0[2 * n0 + i];
                float x1 = P1[j], y1 = P1[n1 + j], z1 = P1[2 * n1 + j];
                float x2 = P2[k], y2 = P2[n2 + k], z2 = P2[2 * n2 + k];
                int i0 = ref_plain_bin(ref_dist(x0, y0, z0, x1, y1, z1), c->lo, c->hi, nb, c->w);
                int i1 = ref_plain_bin(ref_dist(x0, y0, z0, x2, y2, z2), c->lo, c->hi, nb, c->w);
                int i2 = ref_plain_bin(ref_dist(x1, y1, z1, x2, y2, z2), c->lo, c->hi, nb, c->w);
                /* PLAIN:137-139: unsorted, stride nbins even for the
                 * nbins+1 overflow index, which therefore aliases other
                 * cells or lands past the array. */
                int64_t idx = (int64_t)nb * nb * i2 + (int64_t)nb * i1 + i0;
                if (idx >= 0 && idx < cells) h[idx]++;
                else oob++;
            }
        }
    }
    c->th_oob[tid] += oob;
}

int orc_ref_plain(const float *const pts[3], const int n[3], int nbins, float hist_min,
                  float hist_max, uint64_t *hist, uint64_t *oob, int nthreads)
{
    if (nbins <= 0) return -1;
    ref_plain_ctx c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < 3; i++) {
        c.p[i] = pts[i];
        c.n[i] = n[i];
    }
    c.nbins = nbins;
    c.lo = hist_min;
    c.hi = hist_max;
    c.w = (hist_max - hist_min) / (float)nbins;   /* PLAIN:174 */
    int64_t cells = (int64_t)nbins * nbins * nbins;
    if (nthreads < 1) nthreads = 1;
    c.th_hist = (uint64_t **)calloc((size_t)nthreads, sizeof(uint64_t *));
    c.th_oob = (uint64_t *)calloc((size_t)nthreads, sizeof(uint64_t));
    for (int t = 0; t < nthreads; t++) c.th_hist[t] = (uint64_t *)calloc((size_t)cells, 8);
    run_pool(0, n[0], 1, nthreads, &c, ref_plain_body);
    memset(hist, 0, (size_t)cells * 8);
    uint64_t o = 0;
    for (int t = 0; t < nthreads; t++) {
        for (int64_t x = 0; x < cells; x++) hist[x] += c.th_hist[t][x];
        o += c.th_oob[t];
        free(c.th_hist[t]);
    }
    free(c.th_hist);
    free(c.th_oob);
    if (oob) *oob = o;
    return 0;
}

/* ---- pair counts: src/2pcf_C_version.cc ------------------------------- */

typedef struct {
    const float *p[2];
    int n[2];
    int which;
    uint32_t **th_hist;
} ref_two_ctx;

/* distance_to_bin (:58-85) with the program's S_LO 0, S_HI 70, S_NBINS 700
 * and flag 0: float width, -999 outside [lo, hi), int((d - lo)/w). */
static int ref_two_bin(float d)
{
    const float lo = 0.0f, hi = 70.0f;
    const int nbins = 700;
    const float w = (hi - lo) / nbins;
    if (d < lo) return -999;
    if (d >= hi) return -999;
    return (int)((d - lo) / w);
}

static void ref_two_body(void *vctx, int tid, long long ilo, long long ihi)
{
    ref_two_ctx *c = (ref_two_ctx *)vctx;
    uint32_t *h = c->th_hist[tid];
    const float *A = c->p[0], *B = c->p[1];
    const int na = c->n[0], nb = c->n[1];
    for (long long i = ilo; i < ihi; i++) {
        /* :358-364: same file -> j from i (self pairs included); else 0 */
        int jmin = c->which == 0 ? (int)i : 0;
        for (int j = jmin; j < nb; j++) {
            float xd = A[i] - B[j];
            float yd = A[na + i] - B[nb + j];
            float zd = A[2 * na + i] - B[2 * nb + j];
            float d = sqrtf(xd * xd + yd * yd + zd * zd);
            int b = ref_two_bin(d);
            if (b >= 0) h[b]++;
        }
    }
}

int orc_ref_two(const float *const pts[2], const int n[2], int which, uint32_t hist[700], int nthreads)
{
    ref_two_ctx c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < 2; i++) {
        c.p[i] = pts[i];
        c.n[i] = n[i];
    }
    c.which = which;
    if (nthreads < 1) nthreads = 1;
    c.th_hist = (uint32_t **)calloc((size_t)nthreads, sizeof(uint32_t *));
    for (int t = 0; t < nthreads; t++) c.th_hist[t] = (uint32_t *)calloc(700, 4);
    run_pool(0, n[0], 1, nthreads, &c, ref_two_body);
    memset(hist, 0, 700 * 4);
    for (int t = 0; t < nthreads; t++) {
        for (int b = 0; b < 700; b++) hist[b] += c.th_hist[t][b];
        free(c.th_hist[t]);
    }
    free(c.th_hist);
    return 0;
}

/* ======================================================================== */
/* clean64                                                                  */
/* ======================================================================== */

int orc_edges(const orc_axis *ax, double *e)
{
    int n = ax->n;
    if (n < 1 || n > 4096 || !(ax->hi > ax->lo)) return -1;
    if (ax->scale == ORC_LOG) {
        if (!(ax->lo > 0)) return -1;
        double a = log(ax->lo), b = log(ax->hi);
        for (int m = 0; m <= n; m++) e[m] = exp(a + (b - a) * (double)m / (double)n);
    } else if (ax->scale == ORC_LIN) {
        for (int m = 0; m <= n; m++) e[m] = ax->lo + (ax->hi - ax->lo) * (double)m / (double)n;
    } else {
        return -1;
    }
    e[0] = ax->lo;
    e[n] = ax->hi;
    for (int m = 0; m < n; m++)
        if (!(e[m + 1] > e[m])) return -1;
    return 0;
}

void orc_sq_thresholds(const double *e, int n, double *T)
{
    for (int m = 0; m <= n; m++) {
        double em = e[m];
        if (!(em > 0)) { T[m] = 0.0; continue; }
        double t = em * em;
        while (sqrt(t) < em) t = nextafter(t, INFINITY);
        for (;;) {
            double p = nextafter(t, 0.0);
            if (p >= 0 && sqrt(p) >= em) t = p;
            else break;
        }
        T[m] = t;
    }
}

int orc_bin_sq(double s, const double *T, int n)
{
    if (!(s >= T[0]) || s >= T[n]) return -1;
    int lo = 0, hi = n;   /* invariant T[lo] <= s < T[hi] */
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s >= T[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* Bin of a value against an ascending edge table, -1 outside [e0, en). */
static inline int bin_val(double v, const double *e, int n)
{
    if (!(v >= e[0]) || v >= e[n]) return -1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (v >= e[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* Bin of a cosine against the descending table C[m] = cos(pi*t[m]):
 * bin m iff C[m+1] < c <= C[m]. */
static inline int bin_cos(double c, const double *C, int n)
{
    if (!(c <= C[0]) || c <= C[n]) return -1;
    int lo = 0, hi = n;   /* invariant c <= C[lo] and c > C[hi] */
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (c <= C[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

static inline double sqdist(const double *A, int64_t na, int64_t i, const double *B, int64_t nb,
                            int64_t j)
{
    double dx = A[i] - B[j];
    double dy = A[na + i] - B[nb + j];
    double dz = A[2 * na + i] - B[2 * nb + j];
    return (dx * dx + dy * dy) + dz * dz;
}

int64_t orc_hist_cells(const orc_job *job)
{
    if (job->param == ORC_R123) return (int64_t)job->ax[0].n * job->ax[0].n * job->ax[0].n;
    return (int64_t)job->ax[0].n * job->ax[1].n * job->ax[2].n;
}

typedef struct {
    const orc_job *job;
    int term;                  /* index into job->term */
    const double *A, *B, *C;   /* the three segments of the term */
    int64_t na, nb, nc;
    int same_ab, same_bc;
    double T0[256], T1[256];   /* squared thresholds of axis 0 (and s axis) */
    double e1[256], e2[256];   /* q edges; theta edges (JBJ) */
    double C2[256];            /* cos table for the theta axis (SQT) */
    int64_t cells;
    uint64_t **th_hist;
} clean_ctx;

static inline void sort3d(double *a, double *b, double *c)
{
    double t;
    if (*a > *b) { t = *a; *a = *b; *b = t; }
    if (*b > *c) { t = *b; *b = *c; *c = t; }
    if (*a > *b) { t = *a; *a = *b; *b = t; }
}

static inline void clean_triplet(const clean_ctx *c, double sab, double sac, double sbc,
                                 uint64_t *h)
{
    const orc_job *job = c->job;
    if (job->param == ORC_R123) {
        const int n = job->ax[0].n;
        int x = orc_bin_sq(sab, c->T0, n);
        if (x < 0) return;
        int y = orc_bin_sq(sac, c->T0, n);
        if (y < 0) return;
        int z = orc_bin_sq(sbc, c->T0, n);
        if (z < 0) return;
        int t;
        if (x > y) { t = x; x = y; y = t; }
        if (y > z) { t = y; y = z; z = t; }
        if (x > y) { t = x; x = y; y = t; }
        h[((int64_t)x * n + y) * n + z]++;
        return;
    }
    if (sab == 0 || sac == 0 || sbc == 0) return;   /* coincident points */
    double s1 = sab, s2 = sac, s3 = sbc;
    sort3d(&s1, &s2, &s3);
    const double l1 = sqrt(s1), l2 = sqrt(s2), l3 = sqrt(s3);
    const int ns = job->ax[0].n, nq = job->ax[1].n, nt = job->ax[2].n;
    if (job->param == ORC_JBJ) {
        int is = orc_bin_sq(s1, c->T0, ns);
        if (is < 0) return;
        int iq = bin_val(l1 / l2, c->e1, nq);
        if (iq < 0) return;
        int it = bin_val((l3 - l2) / l1, c->e2, nt);
        if (it < 0) return;
        h[((int64_t)is * nq + iq) * nt + it]++;
        return;
    }
    /* SQT: three vertex parametrizations per triangle. */
    const int is1 = orc_bin_sq(s1, c->T0, ns);
    const int is2 = orc_bin_sq(s2, c->T0, ns);
    if (is1 >= 0) {
        int iq = bin_val(l2 / l1, c->e1, nq);
        int it = bin_cos(((s1 + s2) - s3) / ((2.0 * l1) * l2), c->C2, nt);
        if (iq >= 0 && it >= 0) h[((int64_t)is1 * nq + iq) * nt + it]++;
        iq = bin_val(l3 / l1, c->e1, nq);
        it = bin_cos(((s1 + s3) - s2) / ((2.0 * l1) * l3), c->C2, nt);
        if (iq >= 0 && it >= 0) h[((int64_t)is1 * nq + iq) * nt + it]++;
    }
    if (is2 >= 0) {
        int iq = bin_val(l3 / l2, c->e1, nq);
        int it = bin_cos(((s2 + s3) - s1) / ((2.0 * l2) * l3), c->C2, nt);
        if (iq >= 0 && it >= 0) h[((int64_t)is2 * nq + iq) * nt + it]++;
    }
}

static void clean_body(void *vctx, int tid, long long ilo, long long ihi)
{
    clean_ctx *c = (clean_ctx *)vctx;
    uint64_t *h = c->th_hist[tid];
    for (long long i = ilo; i < ihi; i++) {
        int64_t j0 = c->same_ab ? i + 1 : 0;
        for (int64_t j = j0; j < c->nb; j++) {
            double sab = sqdist(c->A, c->na, i, c->B, c->nb, j);
            /* (r1,r2,r3): clean_triplet drops every k when the (i,j) side has
             * no bin, so the k loop is skipped (same counts, less time). */
            if (c->job->param == ORC_R123 && orc_bin_sq(sab, c->T0, c->job->ax[0].n) < 0) continue;
            int64_t k0 = c->same_bc ? j + 1 : 0;
            for (int64_t k = k0; k < c->nc; k++) {
                double sac = sqdist(c->A, c->na, i, c->C, c->nc, k);
                double sbc = sqdist(c->B, c->nb, j, c->C, c->nc, k);
                clean_triplet(c, sab, sac, sbc, h);
            }
        }
    }
}

int orc_clean_count(const orc_job *job, uint64_t *hist, int nthreads)
{
    if (job->nseg < 1 || job->nseg > 3 || job->nterms < 1 || job->nterms > 4) return -1;
    const int naxes = job->param == ORC_R123 ? 1 : 3;
    double e[3][256];
    for (int a = 0; a < naxes; a++)
        if (job->ax[a].n > 255 || orc_edges(&job->ax[a], e[a])) return -1;
    int64_t cells = orc_hist_cells(job);
    if (nthreads < 1) nthreads = 1;
    clean_ctx *c = (clean_ctx *)calloc(1, sizeof(clean_ctx));
    c->job = job;
    c->cells = cells;
    orc_sq_thresholds(e[0], job->ax[0].n, c->T0);
    if (job->param != ORC_R123) {
        memcpy(c->e1, e[1], sizeof(double) * (job->ax[1].n + 1));
        memcpy(c->e2, e[2], sizeof(double) * (job->ax[2].n + 1));
        if (job->param == ORC_SQT) {
            /* theta axis is theta/pi in [0,1]; bins compare the cosine rule
             * value against cos(pi * edge). */
            if (e[2][0] < 0 || e[2][job->ax[2].n] > 1) { free(c); return -1; }
            for (int m = 0; m <= job->ax[2].n; m++) c->C2[m] = cos(M_PI * e[2][m]);
        }
    }
    c->th_hist = (uint64_t **)calloc((size_t)nthreads, sizeof(uint64_t *));
    for (int t = 0; t < nthreads; t++) c->th_hist[t] = (uint64_t *)calloc((size_t)cells, 8);
    memset(hist, 0, (size_t)(cells * job->nterms) * 8);
    for (int tm = 0; tm < job->nterms; tm++) {
        const int *tr = job->term[tm];
        if (tr[0] > tr[1] || tr[1] > tr[2] || tr[0] < 0 || tr[2] >= job->nseg) {
            for (int t = 0; t < nthreads; t++) free(c->th_hist[t]);
            free(c->th_hist);
            free(c);
            return -1;
        }
        c->term = tm;
        c->A = job->seg[tr[0]];
        c->B = job->seg[tr[1]];
        c->C = job->seg[tr[2]];
        c->na = job->nseg_pts[tr[0]];
        c->nb = job->nseg_pts[tr[1]];
        c->nc = job->nseg_pts[tr[2]];
        c->same_ab = tr[0] == tr[1];
        c->same_bc = tr[1] == tr[2];
        for (int t = 0; t < nthreads; t++) memset(c->th_hist[t], 0, (size_t)cells * 8);
        run_pool(0, c->na, 1, nthreads, c, clean_body);
        uint64_t *out = hist + (int64_t)tm * cells;
        for (int t = 0; t < nthreads; t++)
            for (int64_t x = 0; x < cells; x++) out[x] += c->th_hist[t][x];
    }
    for (int t = 0; t < nthreads; t++) free(c->th_hist[t]);
    free(c->th_hist);
    free(c);
    return 0;
}

typedef struct {
    const double *A, *B;
    int64_t na, nb;
    int same;
    const double *T;
    int n;
    uint64_t **th_hist;
} clean_pairs_ctx;

static void clean_pairs_body(void *vctx, int tid, long long ilo, long long ihi)
{
    clean_pairs_ctx *c = (clean_pairs_ctx *)vctx;
    uint64_t *h = c->th_hist[tid];
    for (long long i = ilo; i < ihi; i++) {
        int64_t j0 = c->same ? i + 1 : 0;
        for (int64_t j = j0; j < c->nb; j++) {
            int b = orc_bin_sq(sqdist(c->A, c->na, i, c->B, c->nb, j), c->T, c->n);
            if (b >= 0) h[b]++;
        }
    }
}

int orc_clean_pairs(const double *const seg[2], const int64_t nseg_pts[2], int nseg, const int term[][2],
                    int nterms, const orc_axis *ax, uint64_t *hist, int nthreads)
{
    if (nseg < 1 || nseg > 2 || nterms < 1 || nterms > 3) return -1;
    const int n = ax->n;
    double *e = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *T = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    if (!e || !T || orc_edges(ax, e)) {
        free(e);
        free(T);
        return -1;
    }
    orc_sq_thresholds(e, n, T);
    if (nthreads < 1) nthreads = 1;
    clean_pairs_ctx c;
    memset(&c, 0, sizeof c);
    c.T = T;
    c.n = n;
    c.th_hist = (uint64_t **)calloc((size_t)nthreads, sizeof(uint64_t *));
    for (int t = 0; t < nthreads; t++) c.th_hist[t] = (uint64_t *)calloc((size_t)n, 8);
    int rc = 0;
    for (int tm = 0; tm < nterms && rc == 0; tm++) {
        const int a = term[tm][0], b = term[tm][1];
        if (a < 0 || b < a || b >= nseg) {
            rc = -1;
            break;
        }
        c.A = seg[a];
        c.B = seg[b];
        c.na = nseg_pts[a];
        c.nb = nseg_pts[b];
        c.same = a == b;
        for (int t = 0; t < nthreads; t++) memset(c.th_hist[t], 0, (size_t)n * 8);
        run_pool(0, c.na, 1, nthreads, &c, clean_pairs_body);
        uint64_t *out = hist + (int64_t)tm * n;
        memset(out, 0, (size_t)n * 8);
        for (int t = 0; t < nthreads; t++)
            for (int x = 0; x < n; x++) out[x] += c.th_hist[t][x];
    }
    for (int t = 0; t < nthreads; t++) free(c.th_hist[t]);
    free(c.th_hist);
    free(e);
    free(T);
    return rc;
}

/* ======================================================================== */
/* Histogram writers                                                        */
/* ======================================================================== */

/* (s,q,theta) file format of src/3pcf_C_version_triangle_parametrization.cc:
 * 537-569: three catalog sizes, "%d %d %d" bin counts, three "%-6.3f" lines
 * (lo, hi, width) and then n0*n1 lines of n2 "%lu " values. */
int orc_write_tri(const char *path, const int64_t ngal[3], const int nb[3],
                  const double lo[3], const double hi[3], const double width[3],
                  const uint64_t *hist)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    for (int c = 0; c < 3; c++) fprintf(f, "%lld\n", (long long)ngal[c]);
    fprintf(f, "%d %d %d\n", nb[0], nb[1], nb[2]);
    for (int a = 0; a < 3; a++) fprintf(f, "%-6.3f %-6.3f %-6.3f\n", lo[a], hi[a], width[a]);
    for (int i = 0; i < nb[0]; i++)
        for (int j = 0; j < nb[1]; j++) {
            for (int k = 0; k < nb[2]; k++)
                fprintf(f, "%llu ", (unsigned long long)hist[((int64_t)i * nb[1] + j) * nb[2] + k]);
            fprintf(f, "\n");
        }
    return fclose(f) ? -1 : 0;
}

/* (r1,r2,r3) cube format of src/3pcf_C_version.cc:440-460: "n n n", then
 * for every i a line "i" followed by n lines of n values at cell
 * n^2*k + n*j + i. */
int orc_write_plain(const char *path, int n, const uint64_t *hist)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "%d %d %d\n", n, n, n);
    for (int i = 0; i < n; i++) {
        fprintf(f, "%d\n", i);
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < n; k++)
                fprintf(f, "%llu ", (unsigned long long)hist[(int64_t)n * n * k + (int64_t)n * j + i]);
            fprintf(f, "\n");
        }
    }
    return fclose(f) ? -1 : 0;
}

/* Pair-count format of src/2pcf_C_version.cc:383-397. */
int orc_write_two(const char *path, const int64_t ngal[2], int nbins, double lo, double hi, double width,
                  const uint64_t *hist)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "%lld\n%lld\n%d\n", (long long)ngal[0], (long long)ngal[1], nbins);
    fprintf(f, "%-6.3f %-6.3f %-6.3f\n", lo, hi, width);
    for (int b = 0; b < nbins; b++) fprintf(f, "%llu\n", (unsigned long long)hist[b]);
    return fclose(f) ? -1 : 0;
}
