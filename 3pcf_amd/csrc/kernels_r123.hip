// kernels_r123.hip -- gfx950 kernels of the sorted-side (r1,r2,r3) path.
//
//   pair_tiles_kernel     pair stage: every pair of a job's union layout gets
//                         a u8 bin (or the sentinel) in 64x64 tiles, and every
//                         tile a byte saying whether any pair of it has a bin;
//                         for the bit-mask kernel also the tile's set of bins,
//                         its non-empty rows and the row bit masks of its
//                         lowest bins.  FP32 squared distances against squared
//                         edges, with every pair whose FP32 value lies within
//                         a rigorous error bound (or the requested relative
//                         window) of an edge recomputed in FP64; or FP64
//                         throughout.
//   r123_triplet_kernel   persistent triplet stage over a work queue of
//                         (term, A, B, C-range) block triples: for each row i
//                         of A and point j of B (a lane), the points k of C
//                         with bins y = bin(i,k) and z = bin(j,k) are counted
//                         64 at a time as popcount(row mask of y & row mask of
//                         z), bin pair by bin pair, into per-lane accumulators
//                         of x = bin(i,j); wave sums go to an unsorted-cube LDS
//                         histogram folded into sorted uint64 cells on flush.
//                         Block pairs, C blocks and rows with no pair in range
//                         are skipped.
//   r123_plain_kernel     the FP64 baseline: all three distances of every
//                         triplet recomputed in FP64 and binned, like the
//                         reference's distance() (src/3pcf_C_version.cc:54-142).
//
// Compile with -ffp-contract=off: the FP64 arithmetic must be the unfused
// IEEE sequence ((dx*dx + dy*dy) + dz*dz) that the CPU definition uses.
#include "tpcf_internal.h"

namespace tpcf {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Largest m in [0, n] with T[m] <= s, or -1 when s < T[0]; m == n means the
// value is past the last edge.
template <typename F>
__device__ inline int search_le(F s, const F *T, int n) {
    if (!(s >= T[0])) return -1;
    int lo = 0, hi = n + 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (T[mid] <= s) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Stored pair value: the bin, or the sentinel outside [0, n).
__device__ inline uint32_t bin_of(int m, int n) {
    return (m >= 0 && m < n) ? (uint32_t)m : kSentinel;
}

__device__ inline double sq64(const double *u, int64_t npad, int64_t a, int64_t b) {
    double dx = u[a] - u[b];
    double dy = u[npad + a] - u[npad + b];
    double dz = u[2 * npad + a] - u[2 * npad + b];
    return (dx * dx + dy * dy) + dz * dz;
}

// ---------------------------------------------------------------------------
// pair stage
// ---------------------------------------------------------------------------

// One 64x64 tile (R, C), R <= C, per 256-thread workgroup; thread t owns row
// t/4 and 16 consecutive columns, written as one 16-byte store.
template <bool F32>
__global__ __launch_bounds__(256) void pair_tiles_kernel(PairJob J) {
    const int64_t R = blockIdx.y, C = blockIdx.x;
    if (C < R) return;
    if (J.need && !((J.need[R * J.need_words + (C >> 6)] >> (C & 63)) & 1ull)) {
        // Not computed: marked empty, so work items may span it.
        if (threadIdx.x == 0) {
            if (J.tile_any) J.tile_any[tile_index(R, C, J.nblk)] = 0;
            if (J.meta) J.meta[tile_index(R, C, J.nblk)] = TileMeta{0u, 0u, 0ull};
        }
        return;
    }
    __shared__ float4 sR[kBlock], sC[kBlock];
    __shared__ double sT64[kMaxBins + 1];
    __shared__ float sT32[kMaxBins + 1];
    __shared__ uint32_t s_bins;
    __shared__ unsigned long long s_rows;
    const int t = threadIdx.x;
    const int n = J.n;
    if (t == 0) {
        s_bins = 0u;
        s_rows = 0ull;
    }
    for (int m = t; m <= n; m += 256) {
        sT64[m] = J.T64[m];
        sT32[m] = J.T32[m];
    }
    if (F32) {
        if (t < kBlock) sR[t] = J.u32[R * kBlock + t];
        else if (t < 2 * kBlock) sC[t - kBlock] = J.u32[C * kBlock + t - kBlock];
    }
    __syncthreads();
    const int cntR = J.blk_count[R], cntC = J.blk_count[C];
    const int i = t >> 2;
    const int jb = (t & 3) * 16;
    const int64_t gi = R * kBlock + i;
    uint32_t packed[4] = {0u, 0u, 0u, 0u};
    bool any = false;
    unsigned long long nrecheck = 0;
    const float u = 5.9604644775390625e-8f;  // 2^-24
    const float base = 2.002f * u * J.rmax;
    float4 pi = F32 ? sR[i] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
        const int j = jb + q;
        uint32_t b = kSentinel;
        const bool valid = i < cntR && j < cntC && (R != C || j > i);
        if (valid) {
            const int64_t gj = C * kBlock + j;
            if (F32) {
                const float4 pj = sC[j];
                const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float s = (dx * dx + dy * dy) + dz * dz;
                const int m = search_le(s, sT32, n);
                // |s - s64| <= E: coordinate rounding (2u*rmax per point),
                // the FP32 difference, and three FP32 roundings of the sum.
                const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
                const float ex = base + 1.002f * u * ax;
                const float ey = base + 1.002f * u * ay;
                const float ez = base + 1.002f * u * az;
                const float E = 1.01f * (ex * (2.f * ax + ex) + ey * (2.f * ay + ey) + ez * (2.f * az + ez)) +
                                3.01f * u * s + 1e-30f;
                const double sd = (double)s;
                const double Ed = (double)E + 1e-15 * sd;
                bool nearb = false;
                if (m >= 0) {
                    const double T = sT64[m];
                    nearb |= fabs(sd - T) <= fmax(Ed, 2.0 * (double)J.rel * T);
                }
                if (m < n) {
                    const double T = sT64[m + 1];
                    nearb |= fabs(sd - T) <= fmax(Ed, 2.0 * (double)J.rel * T);
                }
                if (nearb) {
                    ++nrecheck;
                    b = bin_of(search_le(sq64(J.u64, J.npad, gi, gj), sT64, n), n);
                } else {
                    b = bin_of(m, n);
                }
            } else {
                b = bin_of(search_le(sq64(J.u64, J.npad, gi, gj), sT64, n), n);
            }
        }
        packed[q >> 2] |= (b & 0xFFu) << (8 * (q & 3));
        any |= b != kSentinel;
    }
    const int64_t tile = tile_index(R, C, J.nblk);
    uint4 *dst = reinterpret_cast<uint4 *>(J.tiles + tile * kTileBytes + i * kBlock + jb);
    *dst = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    if (F32 && nrecheck) atomicAdd(J.rechecked, nrecheck);
    // Whether any pair of the tile has a bin: a triplet with a side in this
    // tile counts only if that side does, so the triplet stage skips the
    // block pairs (and C blocks) whose tiles have none.
    if (!J.meta) {
        const int tile_has = __syncthreads_or(any ? 1 : 0);
        if (t == 0 && J.tile_any) J.tile_any[tile] = tile_has ? 1 : 0;
        return;
    }
    // Bit-mask kernel (pair values are the bins themselves, n <= 32): the
    // tile's bins and non-empty rows, then the row masks of its lowest
    // kRowMaskSlots bins.
    uint32_t bins = 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t b = (packed[q >> 2] >> (8 * (q & 3))) & 0xFFu;
        if (b != kSentinel) bins |= 1u << b;
    }
    if (bins) {
        atomicOr(&s_bins, bins);
        atomicOr(&s_rows, 1ull << i);
    }
    __syncthreads();
    const uint32_t tb = s_bins;
    if (t == 0) {
        if (J.tile_any) J.tile_any[tile] = tb ? 1 : 0;
        J.meta[tile] = TileMeta{tb, 0u, s_rows};
    }
    uint32_t rest = tb;
    for (int slot = 0; slot < kRowMaskSlots && rest; ++slot) {
        const uint32_t bin = (uint32_t)__builtin_ctz(rest);
        rest &= rest - 1u;
        uint32_t part = 0u;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            part |= (uint32_t)(((packed[q >> 2] >> (8 * (q & 3))) & 0xFFu) == bin) << q;
        // The four threads of row i hold its columns 0-15, 16-31, 32-47, 48-63.
        unsigned long long m = (unsigned long long)part << jb;
        m |= __shfl_xor(m, 1);
        m |= __shfl_xor(m, 2);
        if ((t & 3) == 0) J.rmask[tile * kRowMaskWords + slot * kBlock + i] = m;
    }
}

// ---------------------------------------------------------------------------
// triplet stage (bit-mask form)
// ---------------------------------------------------------------------------

// LDS histogram of one term with n bins: the unsorted cube, word
// (z * n + y) * n + x for x = bin(i,j), y = bin(i,k), z = bin(j,k), folded
// into the term's sorted uint64 cells on flush.
__host__ __device__ inline int r123_lds_words(int n) { return n * n * n; }

// Fold the unsorted cube into the term's sorted uint64 histogram and clear it.
__device__ inline void r123_flush(uint32_t *hist, int n, unsigned long long *g) {
    __syncthreads();
    const int words = r123_lds_words(n);
    for (int c = threadIdx.x; c < words; c += blockDim.x) {
        const uint32_t v = hist[c];
        if (!v) continue;
        hist[c] = 0u;
        int x = c % n, y = (c / n) % n, z = c / (n * n), t;
        if (x > y) { t = x; x = y; y = t; }
        if (y > z) { t = y; y = z; z = t; }
        if (x > y) { t = x; x = y; y = t; }
        atomicAdd(g + (x * n + y) * n + z, (unsigned long long)v);
    }
    __syncthreads();
}

// Tile bytes read per lane: global (not flat) loads, so waiting for them does
// not also wait for LDS operations (flat loads count against both counters).
typedef const uint8_t __attribute__((address_space(1))) *global_u8_ptr;
// Tile summaries read through the constant address space (scalar loads).
typedef const TileMeta __attribute__((address_space(4))) CTileMeta;
struct MetaV {
    uint32_t bins;
    unsigned long long rows;
};
__device__ inline MetaV load_meta(const TileMeta *meta, int64_t t) {
    CTileMeta *p = (CTileMeta *)(meta + t);
    return MetaV{p->bins, p->rows};
}

__device__ inline void lds_add(uint32_t *hist, uint32_t byte_addr, uint32_t v) {
    __hip_atomic_fetch_add(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + byte_addr), v,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Sum of x over the 64 lanes (uniform): inclusive sums within each row of 16
// lanes by DPP shifts, then the four row totals.
__device__ inline uint32_t wave_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 15) + (uint32_t)__builtin_amdgcn_readlane((int)x, 31) +
           (uint32_t)__builtin_amdgcn_readlane((int)x, 47) + (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Row mask of `bin` in a tile for lane r = row r (bit k set where pair
// (r, k) has the bin), from the tile bytes: one ballot per row.  For bins
// past the kRowMaskSlots the pair stage stored.
__device__ inline unsigned long long build_row_mask(global_u8_ptr g, uint32_t bin, int lane) {
    uint32_t lo = 0u, hi = 0u;
    for (int r = 0; r < kBlock; ++r) {
        const unsigned long long m = __ballot((uint32_t)g[r * kBlock + lane] == bin);
        if (lane == r) {
            lo = (uint32_t)m;
            hi = (uint32_t)(m >> 32);
        }
    }
    return (unsigned long long)hi << 32 | lo;
}

// One pass over the rows i of block A (set bits of `rows`) for up to NX x
// bins, NY y bins and NZ z bins of a block triple (A, B, C).  Lane j is
// point j of block B.  Per row, ym[a] is the ballot over the points k of
// block C with bin(i,k) = ys[a]; zm[b] (per lane, from the pair stage) has
// bit k set where bin(j,k) = zs[b]; so popcount(ym[a] & zm[b]) counts the
// triplets (i, j, k) with those two bins, and it goes to the lane's
// accumulator of x = bin(i,j) (a sentinel matches no bin).  The x bins share
// accumulator words as 16-bit fields: bins xs[0] and xs[1] are the low and
// high fields of acc0, xs[2] the high field of acc1, added with one
// multiply-add by the lane's weight 1 or 65536 (0 for a lane whose x is not
// in the pass).  A lane adds at most 64 rows x 64 points = 4096 per field.
// At the end each field is summed over the wave and added to its LDS cell by
// one lane.  A y bin with no point in a row is skipped for that row.
// Returns the pass's VALU lane-ops by the kernel's cost model: per row, NY
// ballot compares, 2 NX for the weights, and for each (y, z) pair of a y bin
// present in the row 2 ANDs, 2 popcounts and one multiply-add per
// accumulator word; per accumulator field 9 for its wave sum.
template <int NX, int NY, int NZ>
__device__ __forceinline__ unsigned long long r123_pass(uint32_t *hist, global_u8_ptr gAB, global_u8_ptr gAC,
                                          unsigned long long rows, const uint32_t (&xs)[3],
                                          const uint32_t (&ys)[3], const uint32_t (&zs)[3],
                                          const unsigned long long (&zm)[3], int lane, uint32_t n) {
    uint32_t acc0[NY][NZ], acc1[NY][NZ];
#pragma unroll
    for (int a = 0; a < NY; ++a)
#pragma unroll
        for (int b = 0; b < NZ; ++b) acc0[a][b] = acc1[a][b] = 0u;
    constexpr int NA = NX > 2 ? 2 : 1;
    const uint32_t nrows = (uint32_t)__builtin_popcountll(rows);
    uint32_t ylive = 0u;  // (row, y bin) pairs with some point
    int i = __builtin_ctzll(rows);
    uint32_t yb = gAC[i * kBlock + lane], xb = gAB[i * kBlock + lane];
    for (;;) {
        rows &= rows - 1ull;
        const int inext = rows ? __builtin_ctzll(rows) : i;
        // the next row's bytes load during this row
        const uint32_t ybn = gAC[inext * kBlock + lane], xbn = gAB[inext * kBlock + lane];
        unsigned long long ym[NY];
#pragma unroll
        for (int a = 0; a < NY; ++a) ym[a] = __ballot(yb == ys[a]);
        const uint32_t w0 = xb == xs[0] ? 1u : (NX > 1 && xb == xs[1] ? 65536u : 0u);
        const uint32_t w1 = NX > 2 && xb == xs[2] ? 65536u : 0u;
#pragma unroll
        for (int a = 0; a < NY; ++a) {
            if (!ym[a]) continue;
            ++ylive;
#pragma unroll
            for (int b = 0; b < NZ; ++b) {
                const uint32_t c = (uint32_t)__builtin_popcount((uint32_t)ym[a] & (uint32_t)zm[b]) +
                                   (uint32_t)__builtin_popcount((uint32_t)(ym[a] >> 32) & (uint32_t)(zm[b] >> 32));
                acc0[a][b] += __umul24(c, w0);
                if (NX > 2) acc1[a][b] += __umul24(c, w1);
            }
        }
        if (!rows) break;
        i = inext;
        yb = ybn;
        xb = xbn;
    }
#pragma unroll
    for (int a = 0; a < NY; ++a)
#pragma unroll
        for (int b = 0; b < NZ; ++b)
#pragma unroll
            for (int s = 0; s < NX; ++s) {
                const uint32_t f = s == 0 ? acc0[a][b] & 0xFFFFu : (s == 1 ? acc0[a][b] >> 16 : acc1[a][b] >> 16);
                const uint32_t tot = wave_sum(f);
                if (tot && lane == 0) lds_add(hist, ((zs[b] * n + ys[a]) * n + xs[s]) * 4u, tot);
            }
    return 64ull * ((unsigned long long)nrows * (NY + 2 * NX) + (unsigned long long)ylive * NZ * (4 + NA) +
                    9ull * NX * NY * NZ);
}

// Up to three bins from the set bits of *rest (lowest first), taken off it.
__device__ __forceinline__ int take3(uint32_t *rest, uint32_t (&v)[3]) {
    int k = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        v[q] = 0u;
        if (*rest) {
            v[q] = (uint32_t)__builtin_ctz(*rest);
            *rest &= *rest - 1u;
            k = q + 1;
        }
    }
    return k;
}

// Persistent: workgroups pull items (term, A, B, C-range) from a global
// counter until the queue is drained; the waves of a workgroup split the C
// blocks of an item (wave w takes C = c0+w, c0+w+WAVES, ...) and share the
// LDS histogram.  A block triple is counted bin by bin: the bins present in
// its three tiles (TileMeta, from the pair stage) are taken up to three per
// side at a time, and each pass (r123_pass) counts every triplet whose three
// bins are in the pass's sets, so the passes together count each triplet of
// the block triple once.  Only rows of block A with some pair in range in
// both tiles (A,B) and (A,C) are visited; block pairs and C blocks whose
// tiles have no pair in range are skipped.
template <int WAVES, bool ELASTIC>
__global__ __launch_bounds__(WAVES * 64) void r123_triplet_kernel(const ElasticCtl ctl,
                                                                  const TripletItem *__restrict__ items,
                                                                  int nitems, int *__restrict__ counter,
                                                                  const TripletTerm *__restrict__ terms,
                                                                  int lds_words, unsigned long long *ops) {
    extern __shared__ uint32_t hist[];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Clear the whole allocation once; every flush clears what its term used,
    // so terms of any bin count start from zero.
    for (int c = threadIdx.x; c < lds_words; c += blockDim.x) hist[c] = 0u;
    bool left = false;
    if constexpr (ELASTIC)
        if (threadIdx.x == 0) atomicAdd(ctl.active, 1);
    int cur = -1, cur_n = 0;
    unsigned long long since_flush = 0, wave_ops = 0;
    unsigned long long *cur_hist = nullptr;
    // One C block adds at most 64^3 counts to a cell; flush the 32-bit LDS
    // counters well before they could wrap.
    const unsigned long long flush_limit = 0xC0000000ull;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            if constexpr (ELASTIC) left = elastic_leave(ctl);
            s_item = left ? nitems : atomicAdd(counter, 1);
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= nitems) break;
        const TripletItem I = items[it];
        const int term = __builtin_amdgcn_readfirstlane(I.term);
        const TripletTerm T = terms[term];
        const int n = __builtin_amdgcn_readfirstlane(T.n);
        const int A = __builtin_amdgcn_readfirstlane(I.A);
        const int B = __builtin_amdgcn_readfirstlane(I.B);
        const int c0 = __builtin_amdgcn_readfirstlane(I.c0);
        const int c1 = __builtin_amdgcn_readfirstlane(I.c1);
        const unsigned long long add = (unsigned long long)(c1 - c0) * (kBlock * kBlock * kBlock);
        if (term != cur || since_flush + add > flush_limit) {
            if (cur >= 0) r123_flush(hist, cur_n, cur_hist);
            cur = term;
            cur_n = n;
            cur_hist = T.hist;
            since_flush = 0;
        }
        since_flush += add;
        const int64_t nblk = T.nblk;
        const MetaV mAB = load_meta(T.meta, tile_index(A, B, nblk));
        // No triplet of the item counts when its (A, B) tile has no pair
        // with a bin.
        if (!mAB.bins) continue;
        const global_u8_ptr gAB = (global_u8_ptr)(T.tiles + tile_index(A, B, nblk) * kTileBytes);
        for (int C = c0 + wave; C < c1; C += WAVES) {
            const int64_t tAC = tile_index(A, C, nblk), tBC = tile_index(B, C, nblk);
            const MetaV mAC = load_meta(T.meta, tAC), mBC = load_meta(T.meta, tBC);
            // Likewise a C block whose (A, C) or (B, C) tile has none, and
            // the rows of A with no pair in range in (A, B) or (A, C).
            if (!mAC.bins || !mBC.bins) continue;
            const unsigned long long rows = mAB.rows & mAC.rows;
            if (!rows) continue;
            const global_u8_ptr gAC = (global_u8_ptr)(T.tiles + tAC * kTileBytes);
            const global_u8_ptr gBC = (global_u8_ptr)(T.tiles + tBC * kTileBytes);
            uint32_t zrest = mBC.bins;
            while (zrest) {
                uint32_t zs[3];
                unsigned long long zm[3];
                const int nz = take3(&zrest, zs);
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    zm[b] = 0ull;
                    if (b < nz) {
                        const int slot = __builtin_popcount(mBC.bins & ((1u << zs[b]) - 1u));
                        zm[b] = slot < kRowMaskSlots ? T.rmask[tBC * kRowMaskWords + slot * kBlock + lane]
                                                     : build_row_mask(gBC, zs[b], lane);
                    }
                }
                uint32_t yrest = mAC.bins;
                while (yrest) {
                    uint32_t ys[3];
                    const int ny = take3(&yrest, ys);
                    uint32_t xrest = mAB.bins;
                    while (xrest) {
                        uint32_t xs[3];
                        const int nx = take3(&xrest, xs);
                        switch ((nx - 1) * 9 + (ny - 1) * 3 + (nz - 1)) {
#define TPCF_PASS(X, Y, Z)                                                                     \
    case (X - 1) * 9 + (Y - 1) * 3 + (Z - 1):                                                  \
        wave_ops += r123_pass<X, Y, Z>(hist, gAB, gAC, rows, xs, ys, zs, zm, lane, (uint32_t)n); \
        break;
                            TPCF_PASS(1, 1, 1) TPCF_PASS(1, 1, 2) TPCF_PASS(1, 1, 3)
                            TPCF_PASS(1, 2, 1) TPCF_PASS(1, 2, 2) TPCF_PASS(1, 2, 3)
                            TPCF_PASS(1, 3, 1) TPCF_PASS(1, 3, 2) TPCF_PASS(1, 3, 3)
                            TPCF_PASS(2, 1, 1) TPCF_PASS(2, 1, 2) TPCF_PASS(2, 1, 3)
                            TPCF_PASS(2, 2, 1) TPCF_PASS(2, 2, 2) TPCF_PASS(2, 2, 3)
                            TPCF_PASS(2, 3, 1) TPCF_PASS(2, 3, 2) TPCF_PASS(2, 3, 3)
                            TPCF_PASS(3, 1, 1) TPCF_PASS(3, 1, 2) TPCF_PASS(3, 1, 3)
                            TPCF_PASS(3, 2, 1) TPCF_PASS(3, 2, 2) TPCF_PASS(3, 2, 3)
                            TPCF_PASS(3, 3, 1) TPCF_PASS(3, 3, 2) TPCF_PASS(3, 3, 3)
#undef TPCF_PASS
                        }
                    }
                }
            }
        }
    }
    if (cur >= 0) r123_flush(hist, cur_n, cur_hist);
    if (ops && lane == 0 && wave_ops) atomicAdd(ops, wave_ops);
    if constexpr (ELASTIC)
        if (threadIdx.x == 0 && !left) atomicSub(ctl.active, 1);
}

// ---------------------------------------------------------------------------
// plain FP64 baseline
// ---------------------------------------------------------------------------

// Add the LDS histogram (n^3 sorted cells) into g and clear it.
__device__ inline void r123_plain_flush(uint32_t *hist, int n3, unsigned long long *g) {
    __syncthreads();
    for (int c = threadIdx.x; c < n3; c += blockDim.x) {
        const uint32_t v = hist[c];
        if (v) {
            atomicAdd(g + c, (unsigned long long)v);
            hist[c] = 0u;
        }
    }
    __syncthreads();
}

// Static round-robin over items (one job per launch).  Lane k owns point k of
// block C; every triplet recomputes its three FP64 squared distances and
// bins each against the squared thresholds, then sorts the three bins.
// The 32-bit LDS counters are flushed to the uint64 histogram before the
// items since the last flush could add more than P.flush_limit to a cell.
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void r123_plain_kernel(PlainArgs P) {
    extern __shared__ uint32_t hist[];
    __shared__ double sT[kMaxBins + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = P.n, n3 = n * n * n;
    const bool lds = n <= kMaxLdsR123Bins;
    for (int m = threadIdx.x; m <= n; m += blockDim.x) sT[m] = P.T64[m];
    if (lds)
        for (int c = threadIdx.x; c < n3; c += blockDim.x) hist[c] = 0u;
    __syncthreads();
    const int64_t np = P.npad;
    unsigned long long since = 0;
    for (int it = blockIdx.x; it < P.nitems; it += gridDim.x) {
        const TripletItem I = P.items[it];
        const unsigned long long add = (unsigned long long)(I.c1 - I.c0) * (kBlock * kBlock * kBlock);
        if (lds && since > 0 && since + add > P.flush_limit) {
            r123_plain_flush(hist, n3, P.hist);
            since = 0;
        }
        since += add;
        const int A = I.A, B = I.B;
        const int cntA = P.blk_count[A], cntB = P.blk_count[B];
        for (int C = I.c0; C < I.c1; ++C) {
            const int cntC = P.blk_count[C];
            const int64_t gk = (int64_t)C * kBlock + lane;
            const bool kval = lane < cntC;
            for (int i = wave; i < cntA; i += WAVES) {
                const int64_t gi = (int64_t)A * kBlock + i;
                if (A == C && lane <= i) continue;
                for (int j = (A == B ? i + 1 : 0); j < cntB; ++j) {
                    if (!kval || (B == C && lane <= j)) continue;
                    const int64_t gj = (int64_t)B * kBlock + j;
                    const double sij = sq64(P.u64, np, gi, gj);
                    const double sik = sq64(P.u64, np, gi, gk);
                    const double sjk = sq64(P.u64, np, gj, gk);
                    int x = search_le(sij, sT, n), y = search_le(sik, sT, n), z = search_le(sjk, sT, n);
                    if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
                    int t;
                    if (x > y) { t = x; x = y; y = t; }
                    if (y > z) { t = y; y = z; z = t; }
                    if (x > y) { t = x; x = y; y = t; }
                    if (lds)
                        __hip_atomic_fetch_add(hist + (x * n + y) * n + z, 1u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    else
                        atomicAdd(P.hist + (x * n + y) * n + z, 1ull);
                }
            }
        }
    }
    if (lds) r123_plain_flush(hist, n3, P.hist);
}

// ---------------------------------------------------------------------------
// wide bin specs (n > 32): pair bins stored unscaled, counts go straight to
// the sorted uint64 histogram in device memory.  Static round-robin.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void r123_triplet_wide_kernel(const TripletItem *__restrict__ items, int nitems,
                                                                const TripletTerm *__restrict__ terms) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const TripletItem I = items[it];
        const TripletTerm T = terms[I.term];
        const int n = T.n;
        // Tiles with no pair in range (or not computed: culled items may
        // span such blocks) are skipped, as in r123_triplet_kernel.
        const uint8_t *tany = T.tile_any;
        if (tany && !tany[tile_index(I.A, I.B, T.nblk)]) continue;
        const uint8_t *tAB = T.tiles + tile_index(I.A, I.B, T.nblk) * kTileBytes;
        for (int C = I.c0 + wave; C < I.c1; C += 4) {
            if (tany && !(tany[tile_index(I.A, C, T.nblk)] & tany[tile_index(I.B, C, T.nblk)])) continue;
            const uint8_t *tAC = T.tiles + tile_index(I.A, C, T.nblk) * kTileBytes;
            const uint8_t *tBC = T.tiles + tile_index(I.B, C, T.nblk) * kTileBytes;
            for (int i = 0; i < kBlock; ++i) {
                const int y = tAC[i * kBlock + lane];
                if (y == kSentinel) continue;
                for (int j = 0; j < kBlock; ++j) {
                    const int x = tAB[i * kBlock + j];
                    const int z = tBC[j * kBlock + lane];
                    if (x == kSentinel || z == kSentinel) continue;
                    int a = x, b = y, c = z, t;
                    if (a > b) { t = a; a = b; b = t; }
                    if (b > c) { t = b; b = c; c = t; }
                    if (a > b) { t = a; a = b; b = t; }
                    atomicAdd(T.hist + ((int64_t)a * n + b) * n + c, 1ull);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// FP32 copy of the union layout relative to the origin o.
__global__ void to_f32_kernel(const double *u64, int64_t npad, double ox, double oy, double oz, float4 *u32) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npad; i += (int64_t)gridDim.x * blockDim.x)
        u32[i] = make_float4((float)(u64[i] - ox), (float)(u64[npad + i] - oy), (float)(u64[2 * npad + i] - oz), 0.f);
}

// ---------------------------------------------------------------------------
// host-side launchers (C++ linkage, used by tpcf_api.cpp)
// ---------------------------------------------------------------------------

hipError_t launch_pair_tiles(const PairJob &J, bool f32, hipStream_t st) {
    dim3 grid((unsigned)J.nblk, (unsigned)J.nblk);
    if (f32) hipLaunchKernelGGL(pair_tiles_kernel<true>, grid, dim3(256), 0, st, J);
    else hipLaunchKernelGGL(pair_tiles_kernel<false>, grid, dim3(256), 0, st, J);
    return hipGetLastError();
}

#define TPCF_R123_SHAPES(X) X(4) X(8) X(16)

// One histogram copy per workgroup: `copies` must be 1.
const void *r123_triplet_kernel_ptr(int waves, int copies, bool elastic) {
    if (copies != 1) return nullptr;
#define TPCF_PTR(W)                                                                                      \
    if (waves == W)                                                                                      \
        return elastic ? reinterpret_cast<const void *>(&r123_triplet_kernel<W, true>)                    \
                       : reinterpret_cast<const void *>(&r123_triplet_kernel<W, false>);
    TPCF_R123_SHAPES(TPCF_PTR)
#undef TPCF_PTR
    return nullptr;
}

int r123_lds_bytes(int n, int copies) { return copies * r123_lds_words(n) * 4; }

hipError_t launch_r123_triplets(const ElasticCtl &ctl, const TripletItem *items, int nitems, int *counter,
                                const TripletTerm *terms,
                                int max_n, int waves, int copies, int grid, hipStream_t st) {
    // The kernel's modelled VALU lane-ops go to the 64-bit word after the
    // item counter (counter[2..3]).
    unsigned long long *ops = reinterpret_cast<unsigned long long *>(counter + 2);
    if (max_n < 1 || max_n > kMaxLdsR123Bins || copies != 1) return hipErrorInvalidValue;
    const int words = r123_lds_words(max_n);
    const size_t lds = (size_t)words * 4;
#define TPCF_LAUNCH(W)                                                                                       \
    if (waves == W) {                                                                                        \
        if (ctl.active)                                                                                      \
            hipLaunchKernelGGL((r123_triplet_kernel<W, true>), dim3(grid), dim3(W * 64), lds, st, ctl, items,  \
                               nitems, counter, terms, words, ops);                                          \
        else                                                                                                 \
            hipLaunchKernelGGL((r123_triplet_kernel<W, false>), dim3(grid), dim3(W * 64), lds, st, ctl, items, \
                               nitems, counter, terms, words, ops);                                          \
        return hipGetLastError();                                                                            \
    }
    TPCF_R123_SHAPES(TPCF_LAUNCH)
#undef TPCF_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_r123_triplets_wide(const TripletItem *items, int nitems, const TripletTerm *terms, int grid,
                                     hipStream_t st) {
    hipLaunchKernelGGL(r123_triplet_wide_kernel, dim3(grid), dim3(256), 0, st, items, nitems, terms);
    return hipGetLastError();
}

hipError_t launch_r123_plain(const PlainArgs &P, int grid, hipStream_t st) {
    const size_t lds = P.n <= kMaxLdsR123Bins ? (size_t)P.n * P.n * P.n * 4 : 0;
    hipLaunchKernelGGL(r123_plain_kernel<4>, dim3(grid), dim3(256), lds, st, P);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(64), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double *u64, int64_t npad, const double o[3], float4 *u32, hipStream_t st) {
    hipLaunchKernelGGL(to_f32_kernel, dim3(256), dim3(256), 0, st, u64, npad, o[0], o[1], o[2], u32);
    return hipGetLastError();
}

}  // namespace tpcf
