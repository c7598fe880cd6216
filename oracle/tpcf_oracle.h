/*
 * tpcf_oracle.h -- CPU oracle for the brute-force triplet-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libtpcf, the drop-in
 * command-line tools, the kernels) may link, load or call this library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker.
 *
 * Two semantic families live here:
 *
 *  refcompat32  A clean-room restatement of the reference programs' own FP32
 *               arithmetic, loop bounds and quirks, so that the restatement can
 *               be pinned byte-for-byte against the histogram files shipped with
 *               the reference (examples/sandbox/) and against the reference
 *               compiled from source.
 *                 - (s,q,theta)  src/3pcf_C_version_triangle_parametrization.cc
 *                 - JBJ variant  src/3pcf_C_version_triangle_parametrization_JBJ.cc
 *                 - (r1,r2,r3)   src/3pcf_C_version.cc
 *                 - pair counts  src/2pcf_C_version.cc
 *
 *  clean64      The FP64 semantics the GPU kernels must reproduce bit-exactly:
 *               unique triplets, runtime bin specs shared through edge tables,
 *               out-of-range sides dropped, uint64 counters.
 *
 * Parity status of every function is given next to its declaration.
 */
#ifndef TPCF_ORACLE_H
#define TPCF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Catalog readers                                                           */
/* ------------------------------------------------------------------------ */

/* 7-column catalog "%d %f %f %f %f %f %f", columns 5-7 are x y z.
 * Follows src/3pcf_C_version_triangle_parametrization.cc:366-394.
 * as_double=0 parses with %f (float), as_double=1 with %lf.
 * Returns the number of points, or a negative value on error.
 * *out is malloc'ed SoA [x(n) | y(n) | z(n)] of float or double.
 * Pinned: sandbox golden histograms (the parse feeds them) and a token
 * round-trip test in tests/test_oracle_io.py. */
int64_t orc_read_7col(const char *path, int as_double, void **out);

/* Count-header catalog: "%d" N then N x "%f %f %f", each value divided by
 * scale (float division when as_double=0).
 * Follows src/3pcf_C_version.cc:311-332.  Pinned: compiled reference
 * (oracle/_ref) outputs on the same files. */
int64_t orc_read_count(const char *path, int as_double, double scale, void **out);

void orc_free(void *p);

/* ------------------------------------------------------------------------ */
/* refcompat32                                                               */
/* ------------------------------------------------------------------------ */

/* One axis of a compile-time bin spec of the reference (the #define values). */
typedef struct {
    int n;
    double lo, hi;
} orc_ref_axis;

/* Term selector produced by the reference's path-string comparison, which is
 * stored in a C++ bool (src/3pcf_C_version_triangle_parametrization.cc:333-353):
 * 0 = the three paths are equal, or the "D R D" / all-different cases;
 * 1 = every other combination.  Returns 0/1. */
int orc_ref_which(const char *p0, const char *p1, const char *p2);

/* Index-chunk voxel bounds, src/3pcf_C_version_triangle_parametrization.cc:52-80.
 * Returns 0 and fills lohi, or -1 if the chunk is empty (the reference exits). */
int orc_ref_vox2gal(int division, int index, int ngals, int lohi[2]);

/* (s,q,theta) triplet loop, src/3pcf_C_version_triangle_parametrization.cc:477-533
 * with distance() at :116-240 and distance_to_bin() at :84-113.
 *   variant 0 = TRI (three vertex parametrizations per triangle),
 *   variant 1 = JBJ (src/3pcf_C_version_triangle_parametrization_JBJ.cc:192-198).
 *   kmin_legacy=1 selects the older DDD bound kmin=j+1 (the commented line at
 *   :499) that produced examples/sandbox/DDD_voxel4_*.dat.
 * pts[c] points at SoA float [x|y|z] of catalog c (already voxel-sliced),
 * hist has ax[0].n*ax[1].n*ax[2].n uint32 counters (wrapping, as the
 * reference's unsigned int).  numcalc receives the loop iteration count.
 * Pinned: 8/8 DDD_voxel2 + 64/64 DDD_voxel4 + Debbie DDD sandbox goldens,
 * and oracle/_ref builds of the reference. */
int orc_ref_tri(const float *const pts[3], const int n[3], int which, int variant,
                int kmin_legacy, const orc_ref_axis ax[3], uint32_t *hist,
                uint64_t *numcalc, int nthreads);

/* (r1,r2,r3) loop of src/3pcf_C_version.cc:363-432 (every ordered triple,
 * unsorted bin index n^2*i2+n*i1+i0, under -> bin 0, over -> bin n+1, 64
 * linear bins on [0, sqrt(3*24*24)) unless overridden).  Index values at or
 * beyond n^3 are written out of bounds by the reference; here they are
 * counted in *oob instead.  Pinned: oracle/_ref build of the reference. */
int orc_ref_plain(const float *const pts[3], const int n[3], int nbins, float hist_min,
                  float hist_max, uint64_t *hist, uint64_t *oob, int nthreads);

/* Pair counts of src/2pcf_C_version.cc (main loop :346-378): FP32
 * distances sqrt(dx*dx + dy*dy + dz*dz), 700 linear bins on [0,70) with the
 * float width 0.1 and int truncation (:58-85), out of range dropped; same
 * file (which = 0, strcmp of the two paths, :213-224): j from i, so each
 * point's self pair lands in bin 0; different files (which = 1): every
 * pair.  Counters are the reference's unsigned int (wrap).  Pinned: the
 * reference compiled from source (oracle/_ref/two) on sample catalogs,
 * byte for byte through orc_write_two. */
int orc_ref_two(const float *const pts[2], const int n[2], int which, uint32_t hist[700], int nthreads);

/* ------------------------------------------------------------------------ */
/* clean64                                                                   */
/* ------------------------------------------------------------------------ */

enum { ORC_LIN = 0, ORC_LOG = 1 };

typedef struct {
    int n;       /* 1..255 for triplets, 1..4096 for pair counts */
    double lo, hi;
    int scale;   /* ORC_LIN or ORC_LOG */
} orc_axis;

/* Edge table e[0..n]: lin e[m] = lo + (hi-lo)*m/n, log e[m] =
 * exp(log lo + (log hi - log lo)*m/n), with e[0]=lo and e[n]=hi exactly.
 * Returns 0, or -1 for an invalid spec.  Pinned: closed forms in tests. */
int orc_edges(const orc_axis *ax, double *e);

/* Squared thresholds T[m] = min{ s >= 0 : sqrt(s) >= e[m] } (IEEE sqrt),
 * so that sqrt(s) >= e[m]  <=>  s >= T[m].  Pinned: exhaustive neighbour
 * checks in tests (nextafter on both sides). */
void orc_sq_thresholds(const double *e, int n, double *T);

/* Parametrizations. */
enum { ORC_R123 = 0, ORC_SQT = 1, ORC_JBJ = 2 };

/* A triplet-counting job in clean64 semantics.  Up to three distinct
 * catalogs ("segments"); a term is a sorted triple of segment labels
 * (a<=b<=c); its triplets are: (a,a,a) i<j<k in a; (a,a,b) i<j in a, k in b;
 * (a,b,b) i in a, j<k in b; (a,b,c) the full product. */
typedef struct {
    int nseg;
    const double *seg[3];   /* SoA [x|y|z], FP64 */
    int64_t nseg_pts[3];
    int nterms;
    int term[4][3];
    int param;              /* ORC_R123 / ORC_SQT / ORC_JBJ */
    orc_axis ax[3];         /* R123 uses ax[0] for all three sides */
} orc_job;

/* Number of histogram cells of one term: n^3 for R123, n0*n1*n2 otherwise. */
int64_t orc_hist_cells(const orc_job *job);

/* Count every term of the job; hist receives nterms consecutive histograms
 * of orc_hist_cells() uint64 counters.  R123 cell index is
 * (a*n + b)*n + c for sorted side bins a<=b<=c.  SQT/JBJ index is
 * (i_s*nq + i_q)*nt + i_t.  Pinned: closed-form totals, the union identity
 * DDD+DDR+DRR+RRR = C(N_D+N_R,3), an independent numpy brute force on small
 * inputs, exact-geometry triangles (tests/test_oracle_clean.py). */
int orc_clean_count(const orc_job *job, uint64_t *hist, int nthreads);

/* Bin of one FP64 squared distance against thresholds T[0..n]; -1 if out of
 * range. */
int orc_bin_sq(double s, const double *T, int n);

/* Pair counts in clean64 semantics: terms are pairs of segment labels
 * (a <= b); (a,a) counts the unique pairs i<j of segment a, (a,b) every pair
 * of the product.  FP64 squared distances against the squared thresholds of
 * ax (n up to 4096), out of range dropped, uint64 counters; hist receives
 * nterms histograms of ax->n cells.  Pinned: closed-form totals C(N,2) and
 * N_a N_b, the union identity DD+DR+RR = pairs of the union, a numpy brute
 * force, exact lattice distances (tests/test_oracle_pairs.py). */
int orc_clean_pairs(const double *const seg[2], const int64_t nseg_pts[2], int nseg, const int term[][2],
                    int nterms, const orc_axis *ax, uint64_t *hist, int nthreads);

/* ------------------------------------------------------------------------ */
/* Writers                                                                   */
/* ------------------------------------------------------------------------ */

/* (s,q,theta) text format, src/3pcf_C_version_triangle_parametrization.cc:
 * 537-569.  Pinned: byte equality with the examples/sandbox histogram files. */
int orc_write_tri(const char *path, const int64_t ngal[3], const int nb[3],
                  const double lo[3], const double hi[3], const double width[3],
                  const uint64_t *hist);

/* (r1,r2,r3) cube format, src/3pcf_C_version.cc:440-460.  Pinned: byte
 * equality with the oracle/_ref build of the reference. */
int orc_write_plain(const char *path, int n, const uint64_t *hist);

/* Pair-count format of src/2pcf_C_version.cc:383-397: the two catalog
 * sizes, the bin count, one "%-6.3f" line (lo, hi, width) and one "%lu"
 * line per bin.  Pinned: byte equality with the oracle/_ref build. */
int orc_write_two(const char *path, const int64_t ngal[2], int nbins, double lo, double hi, double width,
                  const uint64_t *hist);

#ifdef __cplusplus
}
#endif
#endif
