/*
 * tpcf.h -- C ABI of libtpcf, the MI355X (gfx950) brute-force triplet counter
 * behind the DDD / DDR / DRR / RRR triangle histograms of the three-point
 * correlation function.
 *
 * The reference (mattbellis/3pcf) has no library interface: its boundary is
 * the command line of its C programs plus text histogram files.  Each entry
 * point below names the reference code it replaces.  Everything is plain C:
 * pointers, sizes and status codes; no C++ or torch types cross this line.
 *
 * Conventions
 *   - Every call returns TPCF_OK (0) or a negative TPCF_E* code, and never
 *     exits the process.  tpcf_last_error() gives a thread-local message.
 *   - Coordinates are FP64 structure-of-arrays [x(n) | y(n) | z(n)] owned by
 *     the caller; catalogs are copied to device memory owned by the library
 *     until tpcf_catalog_destroy().
 *   - Histograms are uint64 arrays owned by the caller.
 *   - A context is bound to one device.  tpcf_count() and tpcf_count_batch()
 *     are synchronous; tpcf_submit()/tpcf_poll()/tpcf_wait() are the
 *     asynchronous serving interface and are thread-safe.
 *
 * Semantics ("clean64"): the histograms equal, bit for bit, a CPU evaluation
 * in IEEE FP64 of the following definitions.
 *   - squared side s = ((dx*dx + dy*dy) + dz*dz), d = sqrt(s) (round to
 *     nearest), dx = x_a - x_b in FP64;
 *   - an axis with n bins on [lo, hi) has edges e[m] = lo + (hi-lo)*m/n
 *     (linear) or exp(log lo + (log hi - log lo)*m/n) (log), e[0]=lo,
 *     e[n]=hi; a value v is in bin m iff e[m] <= v < e[m+1], otherwise it is
 *     out of range and its entry is dropped;
 *   - TPCF_R123: sides sorted r1<=r2<=r3, one entry at cell (b1*n+b2)*n+b3,
 *     all three sides use axis 0;
 *   - TPCF_SQT: three entries per triangle (the vertex parametrizations of
 *     src/3pcf_C_version_triangle_parametrization.cc:192-210):
 *       (s=r1, q=r2/r1, angle between r1 and r2),
 *       (s=r2, q=r3/r2, angle between r2 and r3),
 *       (s=r1, q=r3/r1, angle between r1 and r3);
 *     the angle axis is theta/pi in [0,1], binned by comparing the cosine
 *     rule value ((sa+sb)-sc)/((2*da)*db) against cos(pi*edge); triangles
 *     with a zero side are dropped; cell (i_s*nq + i_q)*nt + i_t;
 *   - TPCF_JBJ: one entry (s=r1, q=r1/r2, v=(r3-r2)/r1)
 *     (src/3pcf_C_version_triangle_parametrization_JBJ.cc:192-198);
 *   - triplets are unique: a term is a sorted triple (a<=b<=c) of segment
 *     labels; (a,a,a) counts i<j<k, (a,a,b) i<j in a and k in b, (a,b,b) i in
 *     a and j<k in b, (a,b,c) the full product.  DDD/DDR/DRR/RRR of data
 *     segment 0 and random segment 1 are (0,0,0), (0,0,1), (0,1,1), (1,1,1).
 */
#ifndef TPCF_H
#define TPCF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPCF_VERSION_MAJOR 0
#define TPCF_VERSION_MINOR 1

enum {
    TPCF_OK = 0,
    TPCF_EINVAL = -1,   /* bad argument or bin spec */
    TPCF_EDEVICE = -2,  /* HIP runtime error */
    TPCF_ENOMEM = -3,   /* device or host allocation failed */
    TPCF_EIO = -4,      /* catalog or histogram file error */
    TPCF_EBUSY = -5,    /* admission refused: not enough device memory now */
    TPCF_ENOTFOUND = -6 /* unknown ticket */
};

/* Parametrizations. */
enum { TPCF_R123 = 0, TPCF_SQT = 1, TPCF_JBJ = 2 };
/* Axis scales. */
enum { TPCF_LIN = 0, TPCF_LOG = 1 };
/* Device implementations (identical histograms). */
enum {
    TPCF_MODE_FAST = 0,   /* (r1,r2,r3): FP32 pair stage + FP64 recheck, integer triplet
                             kernel; (s,q,theta) and JBJ: FP64 pair stage, FP32 entry
                             arithmetic with FP64 recheck of triangles near an edge. */
    TPCF_MODE_PLAIN = 1,  /* FP64 distances recomputed per triplet, one job per launch */
    TPCF_MODE_FAST64 = 2, /* as FAST with every decision in FP64 */
    TPCF_MODE_CELLS = 3   /* FAST over Morton-ordered copies of the segments, skipping every
                             group of points that cannot hold a counted pair or triplet: pair
                             runs (pair counts) or block triples ((r1,r2,r3): a side >= the
                             last edge; (s,q,theta), JBJ: a side beyond any entry) whose box
                             gaps rule it out.  Same histograms */
};

typedef struct tpcf_ctx tpcf_ctx;
typedef struct tpcf_catalog tpcf_catalog;
typedef struct tpcf_survey tpcf_survey;

typedef struct {
    int n;        /* bins, 1..255 (pair counts: 1..4096) */
    double lo, hi;
    int scale;    /* TPCF_LIN / TPCF_LOG */
} tpcf_axis;

typedef struct {
    int param;          /* TPCF_R123 / TPCF_SQT / TPCF_JBJ */
    tpcf_axis ax[3];    /* R123 reads ax[0] only */
    double recheck_rel; /* FAST mode: extra relative window around each edge
                           inside which pairs are recomputed in FP64 (0 = the
                           rigorous FP32 error bound alone); default 1e-5 */
} tpcf_binspec;

/* One counting job: up to three segments, each an index range [lo, hi) of a
 * device catalog, and 1..4 terms over them. */
typedef struct {
    int nseg;
    tpcf_catalog *seg[3];
    int64_t seg_lo[3], seg_hi[3];
    int nterms;
    int term[4][3];
} tpcf_job;

typedef struct {
    uint64_t triplets;         /* unique triplets classified (all terms) */
    uint64_t pairs;            /* pairs binned in the pair stage */
    uint64_t pairs_rechecked;  /* FAST: pairs ((r1,r2,r3)) or triangles ((s,q,theta),
                                  JBJ) recomputed in FP64 */
    double device_ms;          /* device time of the counting stages (HIP events) */
    double wall_ms;            /* host wall time of the call (serving: submit to done) */
    double pair_ms;            /* pair stage (FAST modes) */
    double triplet_ms;         /* triplet kernel alone */
    double queue_ms;           /* serving: submit to admission */
    uint64_t mask_ops;         /* (r1,r2,r3) bit-mask triplet kernel: VALU lane-ops of its
                                  cost model (whole launch, its items under the split form;
                                  0 for other kernels) */
    double tile_bins_mean;     /* (r1,r2,r3): mean bins per pair tile with a bin */
    double tile_fill;          /* (r1,r2,r3): fraction of those tiles' pairs with a bin */
    int32_t triplet_form;      /* (r1,r2,r3) triplet kernel that counted: 0 run-merging,
                                  1 bit-mask, 2 split (items whose (A,B) tile holds few bins
                                  by the bit-mask kernel, the rest by the run-merging kernel),
                                  -1 none of these */
    int32_t reserved;
} tpcf_stats;

/* ---- library / context ------------------------------------------------- */

const char *tpcf_version(void);
/* Message of the last failing call on this thread ("" if none). */
const char *tpcf_last_error(void);

/* Bind a context to a HIP device (replaces the implicit CUDA device setup of
 * src/3pcf_CUDA_version_triangle_parametrization.cu:705-731). */
int tpcf_ctx_create(int device, tpcf_ctx **out);
int tpcf_ctx_destroy(tpcf_ctx *ctx);
/* Cap on device memory the context may hold for job workspaces (bytes);
 * 0 = 90% of free memory at creation.  Admission control of tpcf_submit(). */
int tpcf_ctx_set_memory_budget(tpcf_ctx *ctx, uint64_t bytes);

/* ---- catalogs ------------------------------------------------------------ */

/* Upload n points (FP64 SoA) to the device; every coordinate must be finite
 * and at most 1e100 in magnitude (TPCF_EINVAL otherwise).  Replaces the
 * float host arrays filled at
 * src/3pcf_C_version_triangle_parametrization.cc:366-439. */
int tpcf_catalog_create(tpcf_ctx *ctx, const double *xyz_soa, int64_t n, tpcf_catalog **out);
int tpcf_catalog_destroy(tpcf_catalog *cat);
int64_t tpcf_catalog_size(const tpcf_catalog *cat);

/* Read a catalog file: format 7 = "%d %f %f %f %f %f %f" with x y z in
 * columns 5-7 (src/3pcf_C_version_triangle_parametrization.cc:366-394);
 * format 3 = "%d" count header then "x y z" lines, each divided by scale
 * (src/3pcf_C_version.cc:311-332).  Values are parsed as FP64.  *out is
 * malloc'ed SoA; free with tpcf_free().  Returns the point count (>= 0) or a
 * negative error. */
int64_t tpcf_read_catalog(const char *path, int format, double scale, double **out);
void tpcf_free(void *p);

/* RA, Dec (degrees) and redshift to Cartesian Mpc/h (FP64 SoA out, x|y|z):
 * distance D = p0 + p1 z + p2 z^2 with poly = {p0, p1, p2} (NULL: the MICE fit
 * -1.07199554, 3013.13296, -638.827998), x = D sin(90-dec) sin(ra),
 * y = D sin(90-dec) cos(ra), z = D cos(90-dec).  Replaces convert() of
 * sample_data/gen_flat_MICE_data.py:49-99 (same order of operations). */
int tpcf_radecz_to_xyz(const double *ra, const double *dec, const double *zred, int64_t n, const double poly[3],
                       double *xyz_soa);
/* Read "id ra dec z ..." rows (the MICE catalogs of sample_data/) and
 * convert them as tpcf_radecz_to_xyz does; *out as for tpcf_read_catalog. */
int64_t tpcf_read_radecz(const char *path, const double poly[3], double **out);

/* ---- bin specs --------------------------------------------------------- */

/* Cells of one term histogram: n^3 (R123) or n0*n1*n2. */
int64_t tpcf_hist_cells(const tpcf_binspec *spec);
/* Edge table e[0..n] of an axis of 1..4096 bins (the host table every
 * kernel uses). */
int tpcf_axis_edges(const tpcf_axis *ax, double *edges);
/* Validate a spec (returns TPCF_OK or TPCF_EINVAL with a message). */
int tpcf_binspec_check(const tpcf_binspec *spec);

/* ---- counting ---------------------------------------------------------- */

/* Count all terms of one job into hist (nterms * tpcf_hist_cells() uint64,
 * term-major).  Replaces the triple loop + distance() + distance_to_bin()
 * of src/3pcf_C_version_triangle_parametrization.cc:477-533, :116-240,
 * :84-113 and of src/3pcf_C_version.cc:363-432, :54-142, :23-51. */
int tpcf_count(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode,
               uint64_t *hist, tpcf_stats *stats);

/* tpcf_count without any device allocation in the call: the job's device
 * workspace (staged points, pair tiles, tables, histograms, work queue) lives
 * in a fixed 64 MiB scratch array in device memory that each GPU the library
 * runs on holds until the process ends (its address is looked up, and the
 * call's events created, by tpcf_ctx_create).  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, modes and
 * results as tpcf_count.  A job whose workspace does not fit (the limit of
 * tpcf_job_footprint() is 63 MiB, e.g. about 10k points in (r1,r2,r3) FAST
 * mode, far more in PLAIN mode) fails with TPCF_ENOMEM and a message.  Every
 * context on a GPU shares that GPU's scratch: calls that overlap in time on
 * one GPU (from different contexts) race and must be serialized by the
 * caller; calls on one context are serialized by the library. */
int tpcf_count_noalloc(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode,
                       uint64_t *hist, tpcf_stats *stats);

/* Count many jobs (for instance every voxel and term of a survey) in one
 * persistent device launch.  hist[j] receives job j's histograms.  Replaces
 * the per-voxel process fan-out of examples/build_voxel_submission_scripts.py
 * :9-41 and examples/run_CPU_calculation_each_voxels.csh:60-67.  Jobs of any
 * parametrization (one persistent launch per family); mode must be
 * TPCF_MODE_FAST, TPCF_MODE_FAST64 or TPCF_MODE_CELLS. */
int tpcf_count_batch(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs,
                     int mode, uint64_t *const *hist, tpcf_stats *stats);

/* Prepared batch: the jobs are planned and staged on the device once; each
 * tpcf_batch_run() recomputes every histogram (histogram reset, pair stages,
 * one persistent launch for the (s,q,theta) and JBJ jobs and one for the
 * (r1,r2,r3) jobs) -- replayed from a HIP graph with TPCF_BATCH_GRAPH, with
 * the (r1,r2,r3) kernel's launch shape (waves per workgroup, LDS histogram
 * copies, workgroups per CU) picked by timing the candidates on the batch
 * itself with TPCF_BATCH_AUTOTUNE.  Modes FAST, FAST64 and CELLS (culling
 * staged once at creation).  The catalogs must stay alive until
 * tpcf_batch_destroy().  No reference counterpart (the reference runs one
 * process per voxel and term, examples/run_CPU_calculation_each_voxels.csh,
 * with the triangle-parametrization program). */
typedef struct tpcf_batch tpcf_batch;
enum { TPCF_BATCH_AUTOTUNE = 1, TPCF_BATCH_GRAPH = 2 };
int tpcf_batch_create(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_binspec *specs, int njobs, int mode,
                      int flags, tpcf_batch **out);
int tpcf_batch_run(tpcf_batch *batch, uint64_t *const *hist, tpcf_stats *stats);
/* Chosen launch shape and the autotuner's timings, as text. */
const char *tpcf_batch_info(const tpcf_batch *batch);
int tpcf_batch_destroy(tpcf_batch *batch);

/* Serving: enqueue a job (admitted when its device workspace fits the
 * memory budget; otherwise it waits in the queue), returns a ticket.  hist
 * must stay valid until the ticket completes.  Three workers run batches at
 * the same time, their persistent launches sharing the CUs (shortest work
 * first); their streams have hardware queues of their own and synchronize
 * with the legacy null stream, so null-stream work of the caller waits for
 * running batches.  No reference counterpart (the reference runs one
 * process per job). */
int tpcf_submit(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_binspec *spec, int mode,
                uint64_t *hist, int64_t *ticket);
/* Poll: 1 = done (the ticket is then released), 0 = pending, < 0 = the
 * job's error.  Wait: blocks, then TPCF_OK or the job's error. */
int tpcf_poll(tpcf_ctx *ctx, int64_t ticket, tpcf_stats *stats);
int tpcf_wait(tpcf_ctx *ctx, int64_t ticket, tpcf_stats *stats);

/* Device bytes the job's workspace needs (admission footprint): its region
 * plus an upper bound of its work queue. */
int64_t tpcf_job_footprint(const tpcf_job *job, const tpcf_binspec *spec, int mode);
/* Device memory of the serving workers: out[0] bytes held now (the
 * workspaces of running and idle workers, counted at what they were admitted
 * to grow to), out[1] the most held at any time, out[2] the budget.  A worker
 * gives its workspace back when a queued job would not fit beside it. */
int tpcf_ctx_memory(tpcf_ctx *ctx, uint64_t out[3]);

/* ---- surveys split into spatial cells ------------------------------------ */

/* Longest side a counted triangle of spec can have: r123 the last edge;
 * (s,q,theta) s_hi (1 + max(q_hi, 1)); JBJ s_hi (1/q_lo + v_hi) (TPCF_EINVAL
 * when q_lo = 0: no bound), the last two with a 1e-6 relative margin. */
int tpcf_spec_reach(const tpcf_binspec *spec, double *reach);

/* A survey (data and randoms, FP64 SoA [x|y|z]) split into spatial cells of
 * side >= maxsep over their common bounding box, with the jobs that count
 * its four terms cell by cell: each counted triangle (every side shorter
 * than maxsep) belongs to the lowest-index cell among its points' cells and
 * is counted by that cell's jobs, over the cell's own points and copies of
 * the points of its 13 higher-index neighbour cells.  The cell jobs'
 * histograms add up to the survey's, bit for bit.  Replaces the index-chunk
 * voxels of src/3pcf_C_version_triangle_parametrization.cc:52-80 and the
 * per-voxel process fan-out of examples/build_voxel_submission_scripts.py:9-41
 * with the cell division of examples/test_predivision_of_distances.py:11-35
 * (define_ranges) and its neighbour-cell loop (:123-205).  The survey holds
 * four device catalogs until tpcf_survey_destroy(). */
int tpcf_survey_create(tpcf_ctx *ctx, const double *d_xyz, int64_t nd, const double *r_xyz, int64_t nr,
                       double maxsep, tpcf_survey **out);
/* Number of cells (and the cells per axis in dims, if given). */
int64_t tpcf_survey_cells(const tpcf_survey *survey, int dims[3]);
/* Number of cell jobs (at most six per cell). */
int64_t tpcf_survey_jobs(const tpcf_survey *survey);
/* Count every cell job (FAST, FAST64, CELLS: one batched launch; PLAIN: one
 * job at a time).  hist: 4 x cells uint64 (DDD, DDR, DRR, RRR of the whole
 * survey); per_cell (optional): ncells x 4 x cells, the triangles each cell
 * owns (what sum_voxels.py adds up).  spec must reach at most the survey's
 * maxsep (TPCF_EINVAL otherwise). */
int tpcf_survey_count(tpcf_ctx *ctx, const tpcf_survey *survey, const tpcf_binspec *spec, int mode,
                      uint64_t *hist, uint64_t *per_cell, tpcf_stats *stats);
int tpcf_survey_destroy(tpcf_survey *survey);

/* ---- pair counts (two-point function) ---------------------------------- */

/* Pair-count histograms of one job over a distance axis of 1..4096 bins
 * (lin or log, lo >= 0, hi <= 1e100).  A pair term uses term[t][0] <=
 * term[t][1] (term[t][2] is ignored): (a,a) counts the C(n_a,2) pairs i<j of
 * segment a (no self pairs), (a,b) the n_a*n_b pairs of the product, so DD,
 * DR, RR of data segment 0 and random segment 1 are (0,0), (0,1), (1,1).  A
 * pair is in bin m iff e[m] <= d < e[m+1] with d = sqrt((dx*dx + dy*dy) +
 * dz*dz) in FP64, the edges as for the triplet axes; others are dropped.
 * hist: nterms * ax->n uint64, term-major.  Modes: TPCF_MODE_FAST (and
 * TPCF_MODE_FAST64, the same kernel: its FP32 arithmetic only guesses which
 * FP64 comparison to make), TPCF_MODE_PLAIN (binary search per pair) and
 * TPCF_MODE_CELLS (the FAST kernel over Morton-ordered copies of the
 * segments, skipping pairs of point runs whose FP64 box gap already reaches
 * the last squared threshold: same histograms, work roughly proportional to
 * the pairs within the last edge; the idea of
 * examples/test_predivision_of_distances.py:11-227).
 * stats->pairs counts the pairs classified.  Replaces the pair loop,
 * distance() and distance_to_bin() of src/2pcf_C_version.cc:358-387, :89-113
 * and :57-86 (700 FP32 bins on [0, 70), self pairs counted in bin 0). */
int tpcf_count_pairs(tpcf_ctx *ctx, const tpcf_job *job, const tpcf_axis *ax, int mode, uint64_t *hist,
                     tpcf_stats *stats);

/* Many pair-count jobs (for instance every voxel pair of a survey) in one
 * persistent launch: job j over axis axes[j] into hist[j].  Replaces one
 * src/2pcf_C_version.cc process per voxel pair and term
 * (examples/run_CPU_calculation_2pcf.csh:36-53). */
int tpcf_count_pairs_batch(tpcf_ctx *ctx, const tpcf_job *jobs, const tpcf_axis *axes, int njobs, int mode,
                           uint64_t *const *hist, tpcf_stats *stats);

/* ---- estimators (host arithmetic, FP64) ---------------------------------- */

/* Number of triplets a term is normalized by, from the three catalog sizes
 * written in its histogram file: term 0 DDD and 3 RRR: N0 (N1-1) (N2-2) / 6,
 * 1 DDR: N0 (N1-1) N2 / 2, 2 DRR: N0 N1 (N2-1) / 2, the integer product
 * rounded once to FP64 (examples/plot_MICE_3pcf.py:96-107). */
int tpcf_triplet_norm(const int64_t ngal[3], int term, double *norm);
/* Pairs a term is normalized by: term 0 DD and 2 RR: (N0^2 - N0) / 2, 1 DR:
 * N0 N1 (examples/plot_MICE_2pcf.py:43-50). */
int tpcf_pair_norm(const int64_t ngal[2], int term, double *norm);
/* zeta = (DDD/n0 - 3 DDR/n1 + 3 DRR/n2 - RRR/n3) / (RRR/n3) per cell, in
 * that order of operations; NaN and +inf give 0
 * (examples/plot_MICE_3pcf.py:143-156). */
int tpcf_zeta(const double *ddd, const double *ddr, const double *drr, const double *rrr, int64_t cells,
              const double norm[4], double *zeta);
/* xi = (DD/n0 - 2 DR/n1 + RR/n2) / (RR/n2) per bin; NaN and +inf give 0
 * (examples/plot_MICE_2pcf.py:63-69). */
int tpcf_xi(const double *dd, const double *dr, const double *rr, int64_t n, const double norm[3], double *xi);
/* numpy's arange(lo, hi, step) (ceil((hi-lo)/step) values lo, lo+step,
 * lo + i*((lo+step)-lo)) into out[cap]; returns the count or an error. */
int tpcf_arange(double lo, double hi, double step, double *out, int64_t cap);
/* Reduced amplitude along theta for one (s, q): theta[t] = the t-th theta/pi
 * low edge, sides s, s q, s sqrt(1 + q^2 - 2 q cos(pi (theta + width/2))),
 * xi of each side from the table (xs, ys) at the nearest abscissa,
 * denom = xi12 xi23 + xi23 xi31 + xi31 xi12, qred = zeta / denom
 * (examples/plot_MICE_3pcf.py:9-51, :196-221). */
int tpcf_reduced_q(const double *zeta_theta, int ntheta, double s, double q, double theta_lo, double theta_width,
                   const double *xs, const double *ys, int nxi, double *theta, double *denom, double *qred);

#ifdef __cplusplus
}
#endif
#endif
