// tpcf_internal.h -- structures shared by the host runtime (tpcf_api.cpp) and
// the gfx950 kernels (kernels_*.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tpcf {

constexpr int kBlock = 64;          // points per block = one wavefront of lanes
constexpr int kTileBytes = kBlock * kBlock;
constexpr uint8_t kSentinel = 0xFF; // pair bin "out of range / not a pair"
constexpr int kMaxBins = 255;
constexpr int kMaxLdsR123Bins = 32; // LDS-histogram triplet kernel; wider specs use global atomics

// Tile (r, c), r <= c, of a job's pair-bin matrix; tiles are stored
// upper-triangle, row-major over r, 4 KiB each, bytes row-major [i][j].
__host__ __device__ inline int64_t tile_index(int64_t r, int64_t c, int64_t nblk) {
    return r * nblk - (r * (r - 1)) / 2 + (c - r);
}
__host__ __device__ inline int64_t tile_count(int64_t nblk) { return nblk * (nblk + 1) / 2; }

// Per-tile summary the (r1,r2,r3) pair stage writes for the bit-mask
// triplet kernel (bins up to kMaxLdsR123Bins).
struct TileMeta {
    uint32_t bins;            // bit b set: some pair of the tile has bin b
    uint32_t pad;
    unsigned long long rows;  // bit i set: row i of the tile has some pair with a bin
};
// Row masks kept per tile, for its kRowMaskSlots lowest bins present: word
// [slot * kBlock + i] has bit j set where pair (i, j) of the tile has that
// bin.  Bins past the stored slots are built by the triplet kernel.
constexpr int kRowMaskSlots = 12;
constexpr int kRowMaskWords = kRowMaskSlots * kBlock;

// One (job, term) of the triplet stage.
struct TripletTerm {
    const uint8_t *tiles;  // pair-bin tiles of the job
    const uint8_t *tile_any;  // per tile: 1 if some pair of it has a bin (nullptr: not kept)
    const TileMeta *meta;  // per tile (bit-mask kernel; nullptr for wide specs)
    const unsigned long long *rmask;  // kRowMaskWords per tile (bit-mask kernel)
    const unsigned long long *tstat;  // the job's tile counts: 64 slots of (bins, tiles with a bin, their pairs with a bin)
    int64_t nblk;          // blocks in the job's union layout
    unsigned long long *hist;  // nterm histogram, n^3 uint64 (sorted-side cells)
    int n;                 // bins per side
    int pad;
};

// Elastic persistent launches (serving): the workgroups of a launch count
// themselves in *active (device memory) and, before claiming their next
// item, leave while more of them are active than *desired (pinned host
// memory the host lowers to make room for another launch and raises again,
// launching more workgroups over the same queue, when one finishes).  A
// workgroup leaves before it claims anything, so no item is lost.  Both
// pointers null: not elastic.
struct ElasticCtl {
    int *active;
    const int *desired;
};

// One workgroup's step of the protocol, by one thread, before a claim: true
// when the workgroup is to leave (its count already taken off *active).
__device__ inline bool elastic_leave(const ElasticCtl &c) {
    if (!c.active) return false;
    int a = __hip_atomic_load(c.active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const int want = __hip_atomic_load(c.desired, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (a <= want) return false;
        // a failed exchange reloads a: another workgroup came or left
        if (__hip_atomic_compare_exchange_strong(c.active, &a, a - 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return true;
    }
}

// Words of a triplet launch's control block (int *counter of the launchers,
// 256-byte aligned, cleared before each launch up to kCtlForm): the item
// counter, the bit-mask kernel's modelled VALU lane-ops (64 bits), the
// chosen form, and the mean bins per tile and the fill of the tiles the
// choice was made on.
constexpr int kCtlItems = 0, kCtlOps = 2, kCtlForm = 4, kCtlBinsMean = 5, kCtlFill = 6;
// Split form: items whose (A, B) tile holds at most ctlw[kCtlSplit] bins go
// to the bit-mask kernel, which then takes its items from a counter of its
// own (ctlw[kCtlItems2], cleared with the item counter); the others to the
// run-merging kernel.
constexpr int kCtlItems2 = 1, kCtlSplit = 7;
constexpr int kFormRuns = 0, kFormBits = 1, kFormSplit = 2;

// A unit of triplet work: blocks A <= B <= C, C in [c0, c1).
struct TripletItem {
    int term;  // index into the TripletTerm table
    int A, B, c0, c1;
    int pad[3];
};

// Pair stage of one job.
struct PairJob {
    const double *u64;     // union layout SoA: x[npad] y[npad] z[npad]
    const float4 *u32;     // (x-o, y-o, z-o, 0) in FP32
    const uint8_t *blk_count;  // real points in each block (0..64)
    uint8_t *tiles;
    uint8_t *tile_any;     // per tile: 1 if some pair of it has a bin, else 0 (nullptr: none kept)
    // Bit-mask triplet kernel (n <= kMaxLdsR123Bins, shift 2): per tile its
    // TileMeta and kRowMaskWords row masks (nullptr: not written).
    TileMeta *meta;
    unsigned long long *rmask;
    // 64 slots (by tile index) of 3 words: bins summed over the tiles with a
    // bin, their count, and their pairs with a bin.
    unsigned long long *tstat;
    const double *T64;     // squared thresholds [0..n]
    const float *T32;      // T64 rounded to float
    unsigned long long *rechecked;  // FP64 recomputations
    int64_t npad;
    int64_t nblk;
    int n;
    float rmax;            // max |coordinate - origin| over the union (FP32, rounded up)
    float rel;             // extra relative window around each threshold (squared units)
    int shift;             // stored pair value = bin << shift (2 for the LDS kernels, 0 otherwise)
    // Culled form: bit C of word [R * need_words + C / 64] set for the tiles
    // (R, C) the triplet stage reads; the others are skipped.  nullptr: all.
    const unsigned long long *need;
    int64_t need_words;
};

// Plain FP64 baseline of one job.
struct PlainArgs {
    const double *u64;
    const uint8_t *blk_count;
    const double *T64;
    const TripletItem *items;
    int nitems;
    int n;
    int64_t npad;
    unsigned long long *hist;  // sorted cells, n^3
    // The workgroup's 32-bit LDS counters are flushed to hist before the
    // items it has added since the last flush could exceed this many
    // triplets (one item adds at most (c1 - c0) * 64^3 to a cell).
    unsigned long long flush_limit;
};

// Parametrization codes (equal to TPCF_SQT / TPCF_JBJ of the C ABI).
constexpr int kSqt = 1, kJbj = 2;

// (s,q,theta) / JBJ FP64 kernel of one job term.
struct SqtArgs {
    const double *u64;
    const uint8_t *blk_count;
    const double *Ts;  // squared thresholds of the s axis [0..ns]
    const double *Eq;  // q edges [0..nq]
    const double *Et;  // SQT: cos(pi*edge) of the theta/pi axis; JBJ: edges of (r3-r2)/r1
    const TripletItem *items;
    int nitems;
    int ns, nq, nt;
    int64_t npad;
    unsigned long long *hist;  // ns*nq*nt uint64
};
hipError_t launch_sqt_plain(const SqtArgs &P, int param, int grid, hipStream_t st);
const void *sqt_plain_kernel_ptr(int param);

// (s,q,theta) FAST form.  Pair codes: s bin (< kCodeNear), or one of
constexpr uint32_t kCodeFar = 0xFF;   // not a pair, or too long for any entry
constexpr uint32_t kCodeZero = 0xFE;  // coincident points: the triangle is dropped
constexpr uint32_t kCodeNear = 0xFD;  // no s bin, short enough to be a q/theta side
constexpr int kMaxFastSqtBins = 253;

struct SqtPairJob {
    const double *u64;
    const uint8_t *blk_count;
    uint8_t *codes;  // 4 KiB per tile
    float *dists;    // 16 KiB per tile: (float)sqrt(s) of the FP64 squared side
    // 3 KiB per tile, kMaskWords words: bit j of row i near / with an s bin /
    // mid, bit i of column j likewise.
    unsigned long long *masks;
    uint8_t *tile_any;  // per tile: 1 if some pair of it is near, else 0 (also for tiles not computed)
    const double *Ts;
    int ns;
    double far_s;    // squared sides above this are kCodeFar
    double mid_s;    // "mid": near and squared side at most this (a q side bound)
    int64_t npad, nblk;
    const unsigned long long *need;  // as PairJob::need
    int64_t need_words;
};
// Per-tile mask words: row near, row short, column near, column short, row
// mid, column mid.  Every entry has its s side short and its q side mid, and
// a short side is mid too, so a triangle with an entry has at least two mid
// sides.
constexpr int kMaskWords = 6 * 64;
constexpr int kMaskRowNear = 0, kMaskRowShort = 64, kMaskColNear = 128, kMaskColShort = 192;
constexpr int kMaskRowMid = 256, kMaskColMid = 320;
// One (job, term) of a FAST (s,q,theta) launch.
struct SqtTerm {
    const uint8_t *codes;
    const float *dists;
    const unsigned long long *masks;
    const uint8_t *tile_any;  // per tile: 1 if some pair of it is near
    const double *u64;
    const double *Ts, *Eq, *Et;
    unsigned long long *hist;       // this term's ns*nq*nt cells
    unsigned long long *rechecked;  // the job's count of triangles recomputed in FP64
    int64_t npad, nblk;
    int param;  // kSqt or kJbj
    int ns, nq, nt, cells;
    int force_fp64;  // every candidate triangle through the FP64 definition
    // Bin guesses of the q axis and (JBJ) the v axis: lin (v - lo) * inv, log
    // (log2(v) - lo) * inv with lo = log2(axis lo); only a hint, every bin is
    // confirmed against the edge tables.  (The (s,q,theta) theta bin is
    // hinted from a table of the cosine edges the kernel builds in LDS.)
    float q_lo, q_inv, t_lo, t_inv;
    int q_log, t_log;
};
struct SqtFastArgs {
    ElasticCtl ctl;  // serving launches; null pointers otherwise
    const SqtTerm *terms;
    const TripletItem *items;  // sorted by term
    int nitems;
    int *counter;
    int lds_cells;  // histogram words in LDS: the largest term's cells
    // Candidate ring per wave in dynamic LDS after the histogram (words,
    // power of two, >= 128).
    int qcap;
};
hipError_t launch_sqt_pairs(const SqtPairJob &J, hipStream_t st);
hipError_t launch_sqt_fast(const SqtFastArgs &P, int waves, int grid, hipStream_t st);
size_t sqt_fast_lds_bytes(int cells, int waves, int qcap);
const void *sqt_fast_kernel_ptr(int waves, bool elastic = false);  // 8 or 16 waves per workgroup

// Pair counts (kernels_pairs.hip).  Bins per axis up to kMaxPairBins.
constexpr int kMaxPairBins = 4096;
constexpr int kPairIpl = 2;  // i points per lane in the FAST pair kernel
// One (job, term) of a pair-count launch: segment a = (ax, ay, az)[0..na),
// segment b = (bx, by, bz)[0..nb).
struct PairHistTerm {
    const double *ax, *ay, *az, *bx, *by, *bz;
    const double *D;           // n+3 doubles: -inf, squared thresholds T[0..n], +inf
    unsigned long long *hist;  // n uint64
    int na, nb, n;
    int same;                  // a and b are one segment: count i < j
    int log;                   // bin guess from log2 (log axis) or sqrt (linear)
    float h_mul, h_add;        // guess k = floor(g * h_mul + h_add)
    // Culled form: box (lo xyz, hi xyz) of every kPairBoxPoints consecutive
    // points of segment b; each wave skips the boxes whose gap to its own
    // points' box reaches the last squared threshold.  nullptr: no skipping.
    const double *jbox;
};
constexpr int kPairBoxPoints = 64;
// Work item: i in [i0, i0 + block) of segment a against j in [j0, j1) of b.
struct PairHistItem {
    int term, i0, j0, j1;
};
struct PairHistArgs {
    const PairHistTerm *terms;
    const PairHistItem *items;
    int nitems;
    int *counter;
    int max_n;       // largest n of the launch's terms (LDS table size)
    int copies;      // privatized LDS histogram copies
    int stride;      // words between copies
    int hist_words;  // copies * stride, rounded up to an even count
};
hipError_t launch_pair_hist(const PairHistArgs &P, int waves, bool cull, int grid, hipStream_t st);
hipError_t launch_pair_plain(const PairHistArgs &P, hipStream_t st);
size_t pair_hist_lds_bytes(int hist_words, int max_n);
const void *pair_hist_kernel_ptr(int waves, bool cull);  // 8 waves per workgroup
const void *pair_plain_kernel_ptr();

// Hilbert ordering of a segment (kernels_cells.hip): sorted SoA copy
// (x at sorted[0..n), y at sorted[stride..], z at sorted[2*stride..]) and the
// box (lo xyz, hi xyz) of every chunk of consecutive sorted points.
// scratch: sort_segment_scratch_bytes(n) bytes.
size_t sort_segment_scratch_bytes(int n);
hipError_t sort_segment(const double *x, const double *y, const double *z, int n, const double lo[3], double span,
                        void *scratch, size_t scratch_bytes, double *sorted, int64_t stride, int chunk, double *box,
                        hipStream_t st);

// Launchers (kernels_r123.hip).
hipError_t launch_pair_tiles(const PairJob &J, bool f32, hipStream_t st);
hipError_t launch_r123_triplets(const ElasticCtl &ctl, const TripletItem *items, int nitems, int *counter,
                                const TripletTerm *terms,
                                int max_n, int waves, int copies, int grid, hipStream_t st);
hipError_t launch_r123_plain(const PlainArgs &P, int grid, hipStream_t st);
hipError_t launch_r123_triplets_wide(const TripletItem *items, int nitems, const TripletTerm *terms, int grid,
                                     hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_to_f32(const double *u64, int64_t npad, const double o[3], float4 *u32, hipStream_t st);
int r123_lds_bytes(int n, int copies);
const void *r123_triplet_kernel_ptr(int waves, int copies, bool elastic = false);
// The bit-mask triplet kernel, launched beside the run-merging one; a choose
// kernel run first picks which of the two counts the launch.
hipError_t launch_r123_choose(const TripletTerm *terms, int nterms, int *counter, float max_bins, int force,
                              int split_bins, float split_mean, hipStream_t st);
hipError_t launch_r123_bits(const ElasticCtl &ctl, const TripletItem *items, int nitems, int *counter,
                            const TripletTerm *terms, int max_n, int waves, int grid, hipStream_t st);
int r123_bits_lds_bytes(int n);
const void *r123_bits_kernel_ptr(int waves, bool elastic = false);

}  // namespace tpcf
