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
//   r123_mfma_kernel      persistent triplet stage over a work queue of
//                         (term, A, B, C-range) block triples, bit-mask form
//                         on the matrix cores: per 16 x 16 block (j of B,
//                         i of A) and bin pair (y, z), an int8 product of the
//                         one-hot matrices [bin(j,k) = z] and [bin(i,k) = y]
//                         counts the points k of C with both bins; v_dot4
//                         with the one-hot bytes of x = bin(i,j) routes the
//                         counts to per-lane accumulators; wave sums go to an
//                         unsorted-cube LDS histogram folded into sorted
//                         uint64 cells on flush.  Block pairs, C blocks and
//                         16-row blocks with no pair in range are skipped.
//   r123_bits_kernel      the same passes on the VALU: for each row i of A
//                         and point j of B (a lane), popcount(row mask of y &
//                         row mask of z) counts 64 points k at a time.
//   r123_runs_kernel      run-merging form for tiles with many bins: one LDS
//                         atomic per run of equal cells along a row.
//   r123_choose_kernel    picks the form per launch from the pair stage's
//                         tile statistics.
//   r123_plain_kernel     the FP64 baseline: all three distances of every
//                         triplet recomputed in FP64 and binned, like the
//                         reference's distance() (src/3pcf_C_version.cc:54-142).
//
// Compile with -ffp-contract=off: the FP64 arithmetic must be the unfused
// IEEE sequence ((dx*dx + dy*dy) + dz*dz) that the CPU definition uses.
#include <cstdlib>

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

// Stored pair value: bin << shift, or the sentinel outside [0, n).
__device__ inline uint32_t bin_of(int m, int n, int shift) {
    return (m >= 0 && m < n) ? ((uint32_t)m << shift) : kSentinel;
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
    __shared__ uint32_t s_bins, s_pairs;
    __shared__ unsigned long long s_rows;
    const int t = threadIdx.x;
    const int n = J.n;
    if (t == 0) {
        s_bins = 0u;
        s_pairs = 0u;
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
                    b = bin_of(search_le(sq64(J.u64, J.npad, gi, gj), sT64, n), n, J.shift);
                } else {
                    b = bin_of(m, n, J.shift);
                }
            } else {
                b = bin_of(search_le(sq64(J.u64, J.npad, gi, gj), sT64, n), n, J.shift);
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
    // Bit-mask kernel (pair values 4 * bin, n <= 32): the tile's bins and
    // non-empty rows, then the row masks of its lowest kRowMaskSlots bins,
    // and the job's count of bins over its tiles (64 slots, by tile index).
    uint32_t bins = 0u, npairs = 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t b = (packed[q >> 2] >> (8 * (q & 3))) & 0xFFu;
        if (b != kSentinel) {
            bins |= 1u << (b >> 2);
            ++npairs;
        }
    }
    if (bins) {
        atomicOr(&s_bins, bins);
        atomicOr(&s_rows, 1ull << i);
        atomicAdd(&s_pairs, npairs);
    }
    __syncthreads();
    const uint32_t tb = s_bins;
    if (t == 0) {
        if (J.tile_any) J.tile_any[tile] = tb ? 1 : 0;
        J.meta[tile] = TileMeta{tb, 0u, s_rows};
        if (tb) {
            unsigned long long *st = J.tstat + 3 * (tile & 63);
            atomicAdd(st, (unsigned long long)__builtin_popcount(tb));
            atomicAdd(st + 1, 1ull);
            atomicAdd(st + 2, (unsigned long long)s_pairs);
        }
    }
    uint32_t rest = tb;
    for (int slot = 0; slot < kRowMaskSlots && rest; ++slot) {
        const uint32_t bin = (uint32_t)__builtin_ctz(rest);
        rest &= rest - 1u;
        uint32_t part = 0u;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            part |= (uint32_t)(((packed[q >> 2] >> (8 * (q & 3))) & 0xFFu) == bin << 2) << q;
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
__host__ __device__ inline int bits_lds_words(int n) { return n * n * n; }

// Fold the unsorted cube into the term's sorted uint64 histogram and clear it.
__device__ inline void bits_flush(uint32_t *hist, int n, unsigned long long *g) {
    __syncthreads();
    const int words = bits_lds_words(n);
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
        const unsigned long long m = __ballot((uint32_t)g[r * kBlock + lane] == bin << 2);
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
// high fields of acc0, xs[2] the high and xs[3] the low field of acc1, added
// with one multiply-add by the lane's weight 1 or 65536 (0 for a lane whose
// x is not in the pass).  A lane adds at most 64 rows x 64 points = 4096 per
// field.
// At the end each field is summed over the wave and added to its LDS cell by
// one lane.  A y bin with no point in a row is skipped for that row.
// Returns the pass's VALU lane-ops by the kernel's cost model: per row, NY
// ballot compares, 2 NX for the weights, and for each (y, z) pair of a y bin
// present in the row 2 ANDs, 2 popcounts and one multiply-add per
// accumulator word; per accumulator field 9 for its wave sum.
template <int NX, int NY, int NZ>
__device__ __forceinline__ unsigned long long bits_pass(uint32_t *hist, global_u8_ptr gAB, global_u8_ptr gAC,
                                          unsigned long long rows, const uint32_t (&xs)[4],
                                          const uint32_t (&ys)[4], const uint32_t (&zs)[4],
                                          const unsigned long long (&zm)[4], int lane, uint32_t n) {
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
        for (int a = 0; a < NY; ++a) ym[a] = __ballot(yb == ys[a] << 2);
        const uint32_t w0 = xb == xs[0] << 2 ? 1u : (NX > 1 && xb == xs[1] << 2 ? 65536u : 0u);
        const uint32_t w1 = NX > 2 && xb == xs[2] << 2 ? 65536u : (NX > 3 && xb == xs[3] << 2 ? 1u : 0u);
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
                const uint32_t f = s == 0 ? acc0[a][b] & 0xFFFFu
                                   : (s == 1 ? acc0[a][b] >> 16 : (s == 2 ? acc1[a][b] >> 16 : acc1[a][b] & 0xFFFFu));
                const uint32_t tot = wave_sum(f);
                if (tot && lane == 0) lds_add(hist, ((zs[b] * n + ys[a]) * n + xs[s]) * 4u, tot);
            }
    return 64ull * ((unsigned long long)nrows * (NY + 2 * NX) + (unsigned long long)ylive * NZ * (4 + NA) +
                    9ull * NX * NY * NZ);
}

// Up to N bins from the set bits of *rest (lowest first), taken off it.
template <int N>
__device__ __forceinline__ int take_bins(uint32_t *rest, uint32_t (&v)[N]) {
    int k = 0;
#pragma unroll
    for (int q = 0; q < N; ++q) {
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
// its three tiles (TileMeta, from the pair stage) are taken up to four per
// side at a time (four x bins cost a pass no more than three: both need two
// accumulator words), and each pass (bits_pass)
// counts every triplet whose three
// bins are in the pass's sets, so the passes together count each triplet of
// the block triple once.  Only rows of block A with some pair in range in
// both tiles (A,B) and (A,C) are visited; block pairs and C blocks whose
// tiles have no pair in range are skipped.
template <int WAVES, bool ELASTIC>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(8, 8))) void r123_bits_kernel(const ElasticCtl ctl,
                                                                  const TripletItem *__restrict__ items,
                                                                  int nitems, int *__restrict__ counter,
                                                                  const TripletTerm *__restrict__ terms,
                                                                  int lds_words, int *ctlw) {
    // Launched beside the run-merging kernel; runs only when chosen.
    if (__builtin_amdgcn_readfirstlane(ctlw[kCtlForm]) != kFormBits) return;
    unsigned long long *ops = reinterpret_cast<unsigned long long *>(ctlw + kCtlOps);
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
            if (cur >= 0) bits_flush(hist, cur_n, cur_hist);
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
                uint32_t zs[4];
                unsigned long long zm[4];
                const int nz = take_bins<4>(&zrest, zs);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    zm[b] = 0ull;
                    if (b < nz) {
                        const int slot = __builtin_popcount(mBC.bins & ((1u << zs[b]) - 1u));
                        zm[b] = slot < kRowMaskSlots ? T.rmask[tBC * kRowMaskWords + slot * kBlock + lane]
                                                     : build_row_mask(gBC, zs[b], lane);
                    }
                }
                uint32_t yrest = mAC.bins;
                while (yrest) {
                    uint32_t ys[4];
                    const int ny = take_bins<4>(&yrest, ys);
                    uint32_t xrest = mAB.bins;
                    while (xrest) {
                        uint32_t xs[4];
                        const int nx = take_bins<4>(&xrest, xs);
                        switch ((nx - 1) * 16 + (ny - 1) * 4 + (nz - 1)) {
#define TPCF_PASS(X, Y, Z)                                                                     \
    case (X - 1) * 16 + (Y - 1) * 4 + (Z - 1):                                                 \
        wave_ops += bits_pass<X, Y, Z>(hist, gAB, gAC, rows, xs, ys, zs, zm, lane, (uint32_t)n); \
        break;
                            TPCF_PASS(1, 1, 1) TPCF_PASS(1, 1, 2) TPCF_PASS(1, 1, 3) TPCF_PASS(1, 1, 4)
                            TPCF_PASS(1, 2, 1) TPCF_PASS(1, 2, 2) TPCF_PASS(1, 2, 3) TPCF_PASS(1, 2, 4)
                            TPCF_PASS(1, 3, 1) TPCF_PASS(1, 3, 2) TPCF_PASS(1, 3, 3) TPCF_PASS(1, 3, 4)
                            TPCF_PASS(1, 4, 1) TPCF_PASS(1, 4, 2) TPCF_PASS(1, 4, 3) TPCF_PASS(1, 4, 4)
                            TPCF_PASS(2, 1, 1) TPCF_PASS(2, 1, 2) TPCF_PASS(2, 1, 3) TPCF_PASS(2, 1, 4)
                            TPCF_PASS(2, 2, 1) TPCF_PASS(2, 2, 2) TPCF_PASS(2, 2, 3) TPCF_PASS(2, 2, 4)
                            TPCF_PASS(2, 3, 1) TPCF_PASS(2, 3, 2) TPCF_PASS(2, 3, 3) TPCF_PASS(2, 3, 4)
                            TPCF_PASS(2, 4, 1) TPCF_PASS(2, 4, 2) TPCF_PASS(2, 4, 3) TPCF_PASS(2, 4, 4)
                            TPCF_PASS(3, 1, 1) TPCF_PASS(3, 1, 2) TPCF_PASS(3, 1, 3) TPCF_PASS(3, 1, 4)
                            TPCF_PASS(3, 2, 1) TPCF_PASS(3, 2, 2) TPCF_PASS(3, 2, 3) TPCF_PASS(3, 2, 4)
                            TPCF_PASS(3, 3, 1) TPCF_PASS(3, 3, 2) TPCF_PASS(3, 3, 3) TPCF_PASS(3, 3, 4)
                            TPCF_PASS(3, 4, 1) TPCF_PASS(3, 4, 2) TPCF_PASS(3, 4, 3) TPCF_PASS(3, 4, 4)
                            TPCF_PASS(4, 1, 1) TPCF_PASS(4, 1, 2) TPCF_PASS(4, 1, 3) TPCF_PASS(4, 1, 4)
                            TPCF_PASS(4, 2, 1) TPCF_PASS(4, 2, 2) TPCF_PASS(4, 2, 3) TPCF_PASS(4, 2, 4)
                            TPCF_PASS(4, 3, 1) TPCF_PASS(4, 3, 2) TPCF_PASS(4, 3, 3) TPCF_PASS(4, 3, 4)
                            TPCF_PASS(4, 4, 1) TPCF_PASS(4, 4, 2) TPCF_PASS(4, 4, 3) TPCF_PASS(4, 4, 4)
#undef TPCF_PASS
                        }
                    }
                }
            }
        }
    }
    if (cur >= 0) bits_flush(hist, cur_n, cur_hist);
    if (lane == 0 && wave_ops) atomicAdd(ops, wave_ops);
    if constexpr (ELASTIC)
        if (threadIdx.x == 0 && !left) atomicSub(ctl.active, 1);
}

// ---------------------------------------------------------------------------
// triplet stage, matrix-core form (bit-mask passes as int8 products)
// ---------------------------------------------------------------------------

typedef int mfma_v4i __attribute__((ext_vector_type(4)));

// Bins per side of one matrix-core pass, at most (the registers of the Y
// operands grow with the y bins, the accumulators with all three).
constexpr int kMfmaCX = 4, kMfmaCY = 2, kMfmaCZ = 2;
// LDS after the histogram: the item's x one-hot table, per 16 x 16 block
// (i-block, j-block) and lane one uint4 (four x bins of one pass, each a
// dword of four one-hot bytes).
constexpr int kMfmaXohBytes = 16 * kBlock * 16;
// Then a table of the one-hot bytes of every byte value: entry v holds the
// bits of v as eight bytes 0/1 (byte t of the pair = bit t of v).
constexpr int kMfmaLutBytes = 256 * 8;
__host__ __device__ inline int mfma_hist_words(int n) { return (n * n * n + 3) & ~3; }

// Up to M bins from the set bits of *rest (lowest first), taken off it; the
// other entries of v are 0.
template <int M>
__device__ __forceinline__ int take_upto(uint32_t *rest, uint32_t (&v)[4]) {
    int k = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = 0u;
        if (q < M && *rest) {
            v[q] = (uint32_t)__builtin_ctz(*rest);
            *rest &= *rest - 1u;
            k = q + 1;
        }
    }
    return k;
}

// Sums over the 64 lanes of K values per lane (K a multiple of 4), uniform:
// v_permlane32_swap pairs values so one register holds the 32-lane halves'
// sums of two values (the first in lanes 0-31), v_permlane16_swap pairs
// those so one register holds four values' 16-lane sums (rows 0-3: values
// 4m, 4m + 2, 4m + 1, 4m + 3), then DPP sums within the rows and the four row
// ends are read.  3.5 VALU instructions per value, against 8 for wave_sum.
template <int K>
__device__ __forceinline__ void wave_sums(const uint32_t (&v)[K], uint32_t (&out)[K]) {
    static_assert(K % 4 == 0, "wave_sums takes a multiple of 4 values");
    uint32_t u[K / 2];
#pragma unroll
    for (int m = 0; m < K / 2; ++m) {
        const auto r = __builtin_amdgcn_permlane32_swap(v[2 * m], v[2 * m + 1], false, false);
        u[m] = r[0] + r[1];
    }
#pragma unroll
    for (int m = 0; m < K / 4; ++m) {
        const auto s = __builtin_amdgcn_permlane16_swap(u[2 * m], u[2 * m + 1], false, false);
        uint32_t x = s[0] + s[1];
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
        x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
        out[4 * m] = (uint32_t)__builtin_amdgcn_readlane((int)x, 15);
        out[4 * m + 2] = (uint32_t)__builtin_amdgcn_readlane((int)x, 31);
        out[4 * m + 1] = (uint32_t)__builtin_amdgcn_readlane((int)x, 47);
        out[4 * m + 3] = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    }
}

// The four bytes of w that equal v (v < 0x80) as bytes 1, the others 0.
__device__ inline uint32_t onehot4(uint32_t w, uint32_t v) {
    const uint32_t t = w ^ (v * 0x01010101u);
    const uint32_t u = ((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t;  // bit 7 of a byte set iff the byte is nonzero
    return (~u >> 7) & 0x01010101u;
}

// The int8 operand of bin `bin` of a tile for v_mfma_i32_16x16x64_i8: lane l
// holds row 16 * blk + (l & 15) of the tile and its 16 columns
// k = 16 * (l >> 4) + 4 q + t (byte t of dword q) as one-hot bytes; the A and
// B operands of the instruction both take this k order, so a product sums
// over every column once.  From the tile's stored row mask when the bin has
// a slot, else from the tile bytes.
// A buffer resource over a tile's stored row masks (kRowMaskWords 64-bit
// words): the mask loads of the passes take their base from it and their
// offsets as operands, with no per-lane address arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mfma_mask_rsrc(const unsigned long long *rmask) {
    // dword 3: 32-bit data format, as for untyped dword loads on gfx9
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned long long *>(rmask), (short)0,
                                             kRowMaskWords * 8, 0x00020000);
}
// The dword of the stored row mask an operand needs: bits 16 * (l >> 4) ..
// + 15 of row 16 * blk + (l & 15) are in dword (l >> 5) of that row's mask;
// lane_off = 8 (l & 15) + 4 (l >> 5) is its byte offset from the block's
// first row mask.
__device__ __forceinline__ uint32_t mfma_mask_dword(__amdgpu_buffer_rsrc_t rm, int slot, int blk, uint32_t lane_off) {
    if (slot >= kRowMaskSlots) return 0u;
    return __builtin_amdgcn_raw_buffer_load_b32(rm, lane_off, (slot * kBlock + 16 * blk) * 8, 0);
}

// The operand from its mask dword (mfma_mask_dword) when the bin has a
// slot (its two mask bytes expanded through the LDS table lut), else from
// the tile bytes; *live (uniform) is whether the bin occurs in the block's
// 16 rows at all.
__device__ __forceinline__ mfma_v4i mfma_operand_from(uint32_t mw, const uint2 *lut, global_u8_ptr tile, int slot,
                                                      uint32_t bin, int blk, int lane, bool *live) {
    uint32_t d[4];
    if (slot < kRowMaskSlots) {
        *live = __builtin_amdgcn_ballot_w64(mw != 0u) != 0ull;
        const uint32_t off = 16 * ((lane >> 4) & 1);
        const uint2 lo = lut[__builtin_amdgcn_ubfe(mw, off, 8)];
        const uint2 hi = lut[__builtin_amdgcn_ubfe(mw, off + 8, 8)];
        d[0] = lo.x;
        d[1] = lo.y;
        d[2] = hi.x;
        d[3] = hi.y;
    } else {
        typedef const uint32_t __attribute__((address_space(1))) *global_u32_ptr;
        const global_u32_ptr w = (global_u32_ptr)(tile + (16 * blk + (lane & 15)) * kBlock + 16 * (lane >> 4));
#pragma unroll
        for (int q = 0; q < 4; ++q) d[q] = onehot4(w[q], bin << 2);
        *live = __builtin_amdgcn_ballot_w64((d[0] | d[1] | d[2] | d[3]) != 0u) != 0ull;
    }
    return mfma_v4i{(int)d[0], (int)d[1], (int)d[2], (int)d[3]};
}

// One pass of a block triple (A, B, C) on the matrix cores for up to NX x
// bins, NY y bins and NZ z bins.  For each 16 x 16 block (j of B, i of A)
// and (y, z) pair, v_mfma_i32_16x16x64_i8 of the one-hot operands
// Z[j][k] = [bin(j,k) = z] and Y[i][k] = [bin(i,k) = y] gives
// D[j][i] = the points k of C with both bins; lane l holds
// i = 16 ib + (l & 15) and the four j = 16 jb + 4 (l >> 4) + q, whose
// x = bin(i, j) one-hot bytes the workgroup staged in LDS (xoh).  D (<= 64
// each) is packed into bytes and v_dot4 with the one-hot bytes of each x bin
// adds it to the lane's accumulator of that (x, y, z); a lane adds at most
// 16 blocks x 4 x 64 = 4096 per accumulator.  The i blocks are taken in two
// halves: the Y operands of a half stay in registers while every j block's
// Z operands are built and counted against them (79 VGPRs: 6 waves per
// SIMD, three workgroups of 8 waves per CU).  Blocks of rows with no pair in
// range (i: in (A, B) and (A, C); j: in (B, C)) are skipped, and so is a
// 16 x 16 block when none of the pass's y bins occurs in its i rows or none
// of its z bins in its j rows (one ballot per operand gives a uniform flag;
// within a block every (y, z) pair is counted, which measured faster than a
// test per pair).  At the end the accumulators are summed over the wave
// (wave_sums) and each added to its LDS cell by one lane.  Returns the
// pass's VALU lane-ops by the kernel's cost model: 4 per operand (two mask
// bytes and their table addresses), per (i, j) block and (y, z) pair 3 to
// pack D and NX dot products, per four accumulators 14 for their wave sums.
template <int NX, int NY, int NZ>
__device__ __forceinline__ unsigned long long mfma_pass(uint32_t *hist, const uint4 *xoh, const uint2 *lut,
                                                        global_u8_ptr gAC,
                                                        global_u8_ptr gBC, __amdgpu_buffer_rsrc_t rmAC,
                                                        __amdgpu_buffer_rsrc_t rmBC, unsigned long long rows,
                                                        unsigned long long rowsB, const uint32_t (&xs)[4],
                                                        const uint32_t (&ys)[4], const uint32_t (&zs)[4],
                                                        const int (&sy)[4], const int (&sz)[4], int lane, uint32_t n) {
    uint32_t acc[NX][NY][NZ];
#pragma unroll
    for (int s = 0; s < NX; ++s)
#pragma unroll
        for (int a = 0; a < NY; ++a)
#pragma unroll
            for (int b = 0; b < NZ; ++b) acc[s][a][b] = 0u;
    const uint32_t lane_off = 8u * (uint32_t)(lane & 15) + 4u * (uint32_t)(lane >> 5);
    uint32_t iblk = 0u;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
        if ((rows >> (16 * ib)) & 0xFFFFull) iblk |= 1u << ib;
    uint32_t nops = 0u, nblocks = 0u;
    // The i blocks in two halves of two: the Y operands of a half stay in
    // registers while every j block is counted against them (the Z operands
    // are built once per half and j block).
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const uint32_t hblk = (iblk >> (2 * h)) & 3u;
        if (!hblk) continue;
        // The Y mask dwords of the half and the Z mask dwords of the first
        // j block are loaded before any operand is built, and each j block's
        // loads of the next one's, so the loads' latencies overlap.
        uint32_t ymw[2][NY], zmw[NZ];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < NY; ++a) ymw[u][a] = mfma_mask_dword(rmAC, sy[a], 2 * h + u, lane_off);
#pragma unroll
        for (int b = 0; b < NZ; ++b) zmw[b] = mfma_mask_dword(rmBC, sz[b], 0, lane_off);
        mfma_v4i yop[2][NY];
        uint32_t ylive = 0u;  // bit u: some y bin of the pass occurs in i block 2 h + u
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int a = 0; a < NY; ++a) {
                bool l = false;
                yop[u][a] = ((hblk >> u) & 1u)
                                ? mfma_operand_from(ymw[u][a], lut, gAC, sy[a], ys[a], 2 * h + u, lane, &l)
                                : mfma_v4i{0, 0, 0, 0};
                if (l) ylive |= 1u << u;
            }
        if (!ylive) continue;
#pragma unroll 1
        for (int jb = 0; jb < 4; ++jb) {
            uint32_t zcur[NZ];
#pragma unroll
            for (int b = 0; b < NZ; ++b) {
                zcur[b] = zmw[b];
                zmw[b] = mfma_mask_dword(rmBC, sz[b], (jb + 1) & 3, lane_off);
            }
            if (!((rowsB >> (16 * jb)) & 0xFFFFull)) continue;
            mfma_v4i zop[NZ];
            bool zlive = false;  // some z bin of the pass occurs in j block jb
#pragma unroll
            for (int b = 0; b < NZ; ++b) {
                bool l = false;
                zop[b] = mfma_operand_from(zcur[b], lut, gBC, sz[b], zs[b], jb, lane, &l);
                zlive = zlive || l;
            }
            nops += NZ;
            if (!zlive) continue;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (!((ylive >> u) & 1u)) continue;
                const int ib = 2 * h + u;
                ++nblocks;
                const uint4 x4 = xoh[(ib * 4 + jb) * kBlock + lane];
                const uint32_t xo[4] = {x4.x, x4.y, x4.z, x4.w};
                constexpr int P = NY * NZ;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const int a = p / NZ, b = p % NZ;
                    const mfma_v4i dcur =
                        __builtin_amdgcn_mfma_i32_16x16x64_i8(zop[b], yop[u][a], mfma_v4i{0, 0, 0, 0}, 0, 0, 0);
                    // the four counts (each <= 64) as the bytes of one dword
                    const uint32_t t01 = __builtin_amdgcn_perm((uint32_t)dcur[1], (uint32_t)dcur[0], 0x0c0c0400u);
                    const uint32_t t23 = __builtin_amdgcn_perm((uint32_t)dcur[3], (uint32_t)dcur[2], 0x0c0c0400u);
                    const uint32_t t4 = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
#pragma unroll
                    for (int s = 0; s < NX; ++s)
                        acc[s][a][b] = __builtin_amdgcn_udot4(xo[s], t4, acc[s][a][b], false);
                }
            }
        }
    }
    constexpr int K = NX * NY * NZ, K4 = (K + 3) & ~3;
    uint32_t v[K4], tot[K4];
#pragma unroll
    for (int q = 0; q < K4; ++q) v[q] = q < K ? acc[q / (NY * NZ)][(q / NZ) % NY][q % NZ] : 0u;
    wave_sums<K4>(v, tot);
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int s = q / (NY * NZ), a = (q / NZ) % NY, b = q % NZ;
        if (tot[q] && lane == 0) lds_add(hist, ((zs[b] * n + ys[a]) * n + xs[s]) * 4u, tot[q]);
    }
    return 64ull * (4ull * ((unsigned long long)__builtin_popcount(iblk) * NY + nops) +
                    (unsigned long long)nblocks * NY * NZ * (3 + NX) + 14ull * (K4 / 4));
}

// Persistent, as the bit-mask kernel: workgroups pull items (term, A, B,
// C-range), the waves split the C blocks of an item and share the LDS
// histogram.  Per item and chunk of up to four x bins the workgroup stages
// the x one-hot bytes of the (A, B) tile in LDS, then every C block of the
// item is counted in passes of up to kMfmaCY y and kMfmaCZ z bins
// (mfma_pass).  Items and C blocks whose tiles have no pair in range are
// skipped.
template <int WAVES, bool ELASTIC>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(6, 6))) void r123_mfma_kernel(
    const ElasticCtl ctl, const TripletItem *__restrict__ items, int nitems, int *__restrict__ counter,
    const TripletTerm *__restrict__ terms, int lds_words, int *ctlw) {
    const int form = __builtin_amdgcn_readfirstlane(ctlw[kCtlForm]);
    if (form != kFormBits && form != kFormSplit) return;
    const int split = form == kFormSplit ? __builtin_amdgcn_readfirstlane(ctlw[kCtlSplit]) : 32;
    if (form == kFormSplit) counter = ctlw + kCtlItems2;
    unsigned long long *ops = reinterpret_cast<unsigned long long *>(ctlw + kCtlOps);
    extern __shared__ uint4 mfma_lds[];
    __shared__ int s_item;
    uint32_t *hist = reinterpret_cast<uint32_t *>(mfma_lds);
    uint4 *xoh = reinterpret_cast<uint4 *>(hist + lds_words);
    uint2 *lut = reinterpret_cast<uint2 *>(xoh + 16 * kBlock);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int c = threadIdx.x; c < lds_words; c += blockDim.x) hist[c] = 0u;
    for (int v = threadIdx.x; v < 256; v += blockDim.x)
        lut[v] = make_uint2(__umul24((uint32_t)v & 0xFu, 0x00204081u) & 0x01010101u,
                            __umul24((uint32_t)v >> 4, 0x00204081u) & 0x01010101u);
    bool left = false;
    if constexpr (ELASTIC)
        if (threadIdx.x == 0) atomicAdd(ctl.active, 1);
    int cur = -1, cur_n = 0;
    unsigned long long since_flush = 0, wave_ops = 0;
    unsigned long long *cur_hist = nullptr;
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
            if (cur >= 0) bits_flush(hist, cur_n, cur_hist);
            cur = term;
            cur_n = n;
            cur_hist = T.hist;
            since_flush = 0;
        }
        since_flush += add;
        const int64_t nblk = T.nblk;
        const MetaV mAB = load_meta(T.meta, tile_index(A, B, nblk));
        if (!mAB.bins || __builtin_popcount(mAB.bins) > split) continue;
        const global_u8_ptr gAB = (global_u8_ptr)(T.tiles + tile_index(A, B, nblk) * kTileBytes);
        uint32_t xrest = mAB.bins;
        while (xrest) {
            uint32_t xs[4];
            const int nx = take_upto<kMfmaCX>(&xrest, xs);
            // The x one-hot table of this chunk: entry (block, lane) holds the
            // bytes of row i = 16 ib + (lane & 15), columns
            // j = 16 jb + 4 (lane >> 4) + 0..3 of tile (A, B).
            __syncthreads();
            for (int e = threadIdx.x; e < 16 * kBlock; e += blockDim.x) {
                const int blk = e >> 6, l = e & 63;
                const int i = 16 * (blk >> 2) + (l & 15), j = 16 * (blk & 3) + 4 * (l >> 4);
                typedef const uint32_t __attribute__((address_space(1))) *global_u32_ptr;
                const uint32_t w = *(global_u32_ptr)(gAB + i * kBlock + j);
                uint4 o;
                o.x = onehot4(w, xs[0] << 2);
                o.y = nx > 1 ? onehot4(w, xs[1] << 2) : 0u;
                o.z = nx > 2 ? onehot4(w, xs[2] << 2) : 0u;
                o.w = nx > 3 ? onehot4(w, xs[3] << 2) : 0u;
                xoh[e] = o;
            }
            __syncthreads();
            wave_ops += 64ull * (16 * kBlock / (64 * WAVES)) * 25;
            for (int C = c0 + wave; C < c1; C += WAVES) {
                const int64_t tAC = tile_index(A, C, nblk), tBC = tile_index(B, C, nblk);
                const MetaV mAC = load_meta(T.meta, tAC), mBC = load_meta(T.meta, tBC);
                if (!mAC.bins || !mBC.bins) continue;
                const unsigned long long rows = mAB.rows & mAC.rows;
                if (!rows) continue;
                const global_u8_ptr gAC = (global_u8_ptr)(T.tiles + tAC * kTileBytes);
                const global_u8_ptr gBC = (global_u8_ptr)(T.tiles + tBC * kTileBytes);
                const __amdgpu_buffer_rsrc_t rmAC = mfma_mask_rsrc(T.rmask + tAC * kRowMaskWords);
                const __amdgpu_buffer_rsrc_t rmBC = mfma_mask_rsrc(T.rmask + tBC * kRowMaskWords);
                uint32_t zrest = mBC.bins;
                while (zrest) {
                    uint32_t zs[4];
                    int sz[4];
                    const int nz = take_upto<kMfmaCZ>(&zrest, zs);
#pragma unroll
                    for (int b = 0; b < 4; ++b) sz[b] = __builtin_popcount(mBC.bins & ((1u << zs[b]) - 1u));
                    uint32_t yrest = mAC.bins;
                    while (yrest) {
                        uint32_t ys[4];
                        int sy[4];
                        const int ny = take_upto<kMfmaCY>(&yrest, ys);
#pragma unroll
                        for (int a = 0; a < 4; ++a) sy[a] = __builtin_popcount(mAC.bins & ((1u << ys[a]) - 1u));
                        switch ((nx - 1) * 4 + (ny - 1) * 2 + (nz - 1)) {
#define TPCF_MPASS(X, Y, Z)                                                                                 \
    case (X - 1) * 4 + (Y - 1) * 2 + (Z - 1):                                                               \
        wave_ops += mfma_pass<X, Y, Z>(hist, xoh, lut, gAC, gBC, rmAC, rmBC, rows, mBC.rows, xs, ys, zs, sy, sz, lane, \
                                       (uint32_t)n);                                                       \
        break;
                            TPCF_MPASS(1, 1, 1) TPCF_MPASS(1, 1, 2) TPCF_MPASS(1, 2, 1) TPCF_MPASS(1, 2, 2)
                            TPCF_MPASS(2, 1, 1) TPCF_MPASS(2, 1, 2) TPCF_MPASS(2, 2, 1) TPCF_MPASS(2, 2, 2)
                            TPCF_MPASS(3, 1, 1) TPCF_MPASS(3, 1, 2) TPCF_MPASS(3, 2, 1) TPCF_MPASS(3, 2, 2)
                            TPCF_MPASS(4, 1, 1) TPCF_MPASS(4, 1, 2) TPCF_MPASS(4, 2, 1) TPCF_MPASS(4, 2, 2)
#undef TPCF_MPASS
                        }
                    }
                }
            }
        }
    }
    if (cur >= 0) bits_flush(hist, cur_n, cur_hist);
    if (lane == 0 && wave_ops) atomicAdd(ops, wave_ops);
    if constexpr (ELASTIC)
        if (threadIdx.x == 0 && !left) atomicSub(ctl.active, 1);
}

// ---------------------------------------------------------------------------
// triplet stage, run-merging form (block triples with many bins per tile)
// ---------------------------------------------------------------------------

// Pair bins are stored pre-scaled as 4*bin (a byte offset into a row of
// 32-bit counters); 0xFF is the sentinel, and every valid value is < 128.
//
// LDS histogram of one term with n bins: unsorted cube word
// z*sz + y*sy + x for x = bin(i,j), y = bin(i,k), z = bin(j,k), with odd
// strides sy = n|1 and sz = (n*sy)|1 so that the lanes of one instruction
// (same x, varying y and z) spread over all 32 banks.  A lane whose z is out
// of range adds into a dump word n*sz + lane + y*sy + x instead of branching,
// so the dump needs sz + 64 extra words.  COPIES privatized copies of the cube
// (lane k writes copy k % COPIES) cut same-address serialization of LDS
// atomics when few cells are hot; copies start 32/COPIES banks apart.
__host__ __device__ inline int runs_sy(int n) { return n | 1; }
__host__ __device__ inline int runs_sz(int n) { return (n * runs_sy(n)) | 1; }
__host__ __device__ inline int runs_lds_words(int n) { return (n + 1) * runs_sz(n) + 64; }
// Copy k of the privatized histogram starts k * stride words in; the stride
// is a multiple of 32 words plus 16/copies, so the hot cells of the copies
// fall in different banks (measured on the 50k+50k four-term workload:
// skew 4 for 4 copies is 5 % faster than skew 8).
__host__ __device__ inline int runs_copy_stride(int n, int copies) {
    return ((runs_lds_words(n) + 31) & ~31) + (copies > 1 ? 16 / copies : 0);
}

// Fold the privatized unsorted cubes into the term's sorted uint64 histogram
// and clear them (dump words included).
template <int COPIES>
__device__ inline void runs_flush(uint32_t *hist, int n, unsigned long long *g) {
    __syncthreads();
    const int sy = runs_sy(n), sz = runs_sz(n);
    const int words = runs_lds_words(n);
    const int stride = runs_copy_stride(n, COPIES);
    for (int c = threadIdx.x; c < words; c += blockDim.x) {
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < COPIES; ++k) {
            v += hist[k * stride + c];
            hist[k * stride + c] = 0u;
        }
        int z = c / sz, y = (c % sz) / sy, x = (c % sz) % sy, t;
        if (z < n && y < n && x < n && v) {
            if (x > y) { t = x; x = y; y = t; }
            if (y > z) { t = y; y = z; z = t; }
            if (x > y) { t = x; x = y; y = t; }
            atomicAdd(g + (x * n + y) * n + z, (unsigned long long)v);
        }
    }
    __syncthreads();
}

typedef const uint32_t __attribute__((address_space(4))) *scalar_u32_ptr;

// Column j of a fully valid row: lanes whose run of equal cells ends at j
// (the carry out of doubling w, which holds the run ends of the remaining
// columns from the top bit down) add the run's length to its cell.  Only
// those lanes reach the LDS (an add from every lane, the others into their
// dump slots, measured 10 % slower on the 50k + 50k workload).
__device__ __forceinline__ void runs_column(uint32_t *hist, uint32_t &w, int j, int &last, uint32_t zs,
                                            uint32_t ys, uint32_t x4) {
    uint32_t w2;
    const bool end = __builtin_add_overflow(w, w, &w2);
    w = w2;
    if (end) {
        const uint32_t run = (uint32_t)(j - last);
        last = j;
        lds_add(hist, zs + ys + x4, run);
    }
}

// Persistent: workgroups pull items (term, A, B, C-range) from a global
// counter until the queue is drained; the waves of a workgroup split the C
// blocks of an item (wave w takes C = c0+w, c0+w+WAVES, ...) and share the
// LDS histogram copies.  Lane k owns point k of block C: z = bin(j,k) is
// preloaded for the 64 columns j of block B, y = bin(i,k) is per lane and
// fixed for a row i of block A, and x = bin(i,j) is uniform, read with scalar
// loads.  The blocks hold spatially ordered (Hilbert) points, so along a row a
// lane's cell (x, y, z) often stays the same for several consecutive
// columns: each lane adds one run of equal cells at a time (the run's length
// at its last column).  A run ends where x changes along the row (one ballot
// per row) or z changes along the lane's column (one mask per C block), so
// the add is issued only by the lanes whose run ends there: fewer LDS atomics
// and fewer lanes per atomic, hence fewer bank and same-address conflicts.
// Every triplet is still counted one by one in its run.  A row with no valid
// x is skipped; a row with some invalid x (sparse rows of culled or wide
// geometries, where the run bookkeeping per column costs more than it saves)
// adds one triplet per valid column, the invalid ones skipped by a scalar
// test.
template <int WAVES, int COPIES, bool ELASTIC>
__global__ __launch_bounds__(WAVES * 64) void r123_runs_kernel(const ElasticCtl ctl,
                                                                  const TripletItem *__restrict__ items,
                                                                  int nitems, int *__restrict__ counter,
                                                                  const TripletTerm *__restrict__ terms,
                                                                  int lds_words, const int *ctlw) {
    // Launched beside the bit-mask kernel; runs only when chosen.
    const int form = __builtin_amdgcn_readfirstlane(ctlw[kCtlForm]);
    if (form != kFormRuns && form != kFormSplit) return;
    const int split = form == kFormSplit ? __builtin_amdgcn_readfirstlane(ctlw[kCtlSplit]) : -1;
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
    unsigned long long since_flush = 0;
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
            if (cur >= 0) runs_flush<COPIES>(hist, cur_n, cur_hist);
            cur = term;
            cur_n = n;
            cur_hist = T.hist;
            since_flush = 0;
        }
        since_flush += add;
        const int64_t nblk = T.nblk;
        // No triplet of the item counts when its (A, B) tile has no pair
        // with a bin.
        const uint8_t *tany = T.tile_any;
        if (tany && !tany[tile_index(A, B, nblk)]) continue;
        // split form: the bit-mask kernel counts the items with few x bins
        if (split >= 0 && T.meta && __builtin_popcount(load_meta(T.meta, tile_index(A, B, nblk)).bins) <= split)
            continue;
        const uint32_t sy = (uint32_t)runs_sy(n), sz = (uint32_t)runs_sz(n);
        const uint32_t copy_base = (uint32_t)((lane % COPIES) * runs_copy_stride(n, COPIES)) * 4u;
        const uint32_t dump = copy_base + ((uint32_t)n * sz + (uint32_t)lane) * 4u;
        const uint8_t *tAB = T.tiles + tile_index(A, B, nblk) * kTileBytes;
        const global_u8_ptr gAB = (global_u8_ptr)tAB;
        for (int C = c0 + wave; C < c1; C += WAVES) {
            // Likewise a C block whose (A, C) or (B, C) tile has none.
            if (tany && !(tany[tile_index(A, C, nblk)] & tany[tile_index(B, C, nblk)])) continue;
            const global_u8_ptr gAC = (global_u8_ptr)(T.tiles + tile_index(A, C, nblk) * kTileBytes);
            const global_u8_ptr gBC = (global_u8_ptr)(T.tiles + tile_index(B, C, nblk) * kTileBytes);
            // z of the 64 columns as byte offsets of the z slab in the lane's
            // copy, or the lane's dump slot for the sentinel.
            uint32_t Zs[kBlock];
#pragma unroll
            for (int j = 0; j < kBlock; ++j) Zs[j] = gBC[j * kBlock + lane];
            // Row i's y and x bytes are loaded during row i-1.
            uint32_t y4n = gAC[lane], xvn = gAB[lane];
#pragma unroll
            for (int j = 0; j < kBlock; ++j) Zs[j] = Zs[j] < 128u ? __umul24(Zs[j], sz) + copy_base : dump;
            // Run ends along the lane's column: bit 63-j set where the z slot
            // of column j differs from that of column j+1 (and at j = 63);
            // computed by the first row of the C block that merges runs.
            unsigned long long zend = 0ull;
            bool have_zend = false;
            for (int i = 0; i < kBlock; ++i) {
                const uint32_t y4 = y4n;
                const uint32_t xv = xvn;
                const int inext = i + 1 < kBlock ? i + 1 : i;
                y4n = gAC[inext * kBlock + lane];
                xvn = gAB[inext * kBlock + lane];
                const unsigned long long rowmask = __ballot(xv < 128u);
                if (rowmask == 0ull) continue;
                // Fully valid rows merge runs.  Their run ends are found
                // here, in uniform control flow (every lane active), before
                // the lanes with invalid y drop out.
                const bool runs = rowmask == ~0ull;
                unsigned long long xend = 0ull;
                if (runs) {
                    // Along the row: x of column j differs from x of column
                    // j+1 (lane j+1's value by a wave shift left; lane 63
                    // keeps its own), bit j of the ballot reversed to 63-j.
                    const uint32_t xv1 = (uint32_t)__builtin_amdgcn_update_dpp((int)xv, (int)xv, 0x130, 0xf, 0xf,
                                                                                false);  // wave_shl:1
                    xend = __builtin_bitreverse64(__ballot(xv != xv1));
                    if (!have_zend) {
                        zend = 1ull;
#pragma unroll
                        for (int j = 0; j + 1 < kBlock; ++j)
                            zend |= (unsigned long long)(Zs[j] != Zs[j + 1]) << (kBlock - 1 - j);
                        have_zend = true;
                    }
                }
                if (y4 >= 128u) continue;
                const uint32_t Ys = __umul24(y4, sy);
                // x(i, j) for the 64 columns: one scalar load of the row.
                // (Reading lane j of xv instead would be wrong here: inside
                // the y-divergent region the compiler may rebuild xv in a
                // fresh register for the active lanes only.)
                const scalar_u32_ptr xrow = (scalar_u32_ptr)(tAB + i * kBlock);
                uint32_t xw[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) xw[q] = xrow[q];
                // The row's words are waited for here, once: the columns
                // below may each be skipped, and a wait inside them would
                // also wait for the wave's outstanding LDS adds (scalar loads
                // and LDS share a counter).  Emits no instruction.
                asm volatile("" ::"s"(xw[0]), "s"(xw[1]), "s"(xw[2]), "s"(xw[3]), "s"(xw[4]), "s"(xw[5]),
                             "s"(xw[6]), "s"(xw[7]), "s"(xw[8]), "s"(xw[9]), "s"(xw[10]), "s"(xw[11]),
                             "s"(xw[12]), "s"(xw[13]), "s"(xw[14]), "s"(xw[15]));
                const unsigned long long ends = xend | zend;
                uint32_t whi = (uint32_t)(ends >> 32), wlo = (uint32_t)ends;
                int last = -1;
                if (runs) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = q * 4 + r;
                            const uint32_t x4 = (xw[q] >> (8 * r)) & 0xFFu;
                            runs_column(hist, j < 32 ? whi : wlo, j, last, Zs[j], Ys, x4);
                        }
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const uint32_t x4 = (xw[q] >> (8 * r)) & 0xFFu;
                            if (x4 < 128u) lds_add(hist, Zs[q * 4 + r] + Ys + x4, 1u);
                        }
                    }
                }
            }
        }
    }
    if (cur >= 0) runs_flush<COPIES>(hist, cur_n, cur_hist);
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

// Picks the triplet kernel of a launch from the pair stages' tile counts
// (every term of the launch, 64 slots each): over the tiles with a bin, the
// mean number of bins m and the fraction f of their pairs with one.  The
// bit-mask kernel's work grows with the bin pairs of a block triple (about
// m^2), the run-merging kernel's with its triplets in range (about f), so the
// bit-mask kernel is chosen when m <= max_bins * sqrt(f).  Otherwise, when
// m <= split_mean and split_bins >= 0, the split form: the items whose (A, B)
// tile holds at most split_bins bins go to the bit-mask kernel, the others
// to the run-merging kernel; else the run-merging kernel alone.  force >= 0
// chooses that form.  One 64-thread workgroup; writes ctlw[kCtlForm],
// ctlw[kCtlSplit], and m and f (float bits) to ctlw[kCtlBinsMean],
// ctlw[kCtlFill].
__global__ __launch_bounds__(64) void r123_choose_kernel(const TripletTerm *__restrict__ terms, int nterms,
                                                         int *ctlw, float max_bins, int force, int split_bins,
                                                         float split_mean) {
    const int t = threadIdx.x;
    unsigned long long nb = 0, nt = 0, np = 0;
    for (int q = 0; q < nterms; ++q) {
        const unsigned long long *st = terms[q].tstat;
        if (!st) continue;
        nb += st[3 * t];
        nt += st[3 * t + 1];
        np += st[3 * t + 2];
    }
    for (int o = 32; o > 0; o >>= 1) {
        nb += __shfl_xor(nb, o);
        nt += __shfl_xor(nt, o);
        np += __shfl_xor(np, o);
    }
    if (t == 0) {
        const float mean = nt ? (float)((double)nb / (double)nt) : 0.f;
        const float fill = nt ? (float)((double)np / ((double)nt * kTileBytes)) : 0.f;
        int form = force;
        if (form < 0)
            form = mean <= max_bins * sqrtf(fill)              ? kFormBits
                   : (split_bins >= 0 && mean <= split_mean) ? kFormSplit
                                                             : kFormRuns;
        ctlw[kCtlForm] = form;
        ctlw[kCtlSplit] = split_bins;
        ctlw[kCtlBinsMean] = __float_as_int(mean);
        ctlw[kCtlFill] = __float_as_int(fill);
    }
}

#define TPCF_R123_SHAPES(X)                                                                                    \
    X(2, 1) X(4, 1) X(8, 1) X(16, 1) X(2, 2) X(4, 2) X(8, 2) X(16, 2) X(2, 4) X(4, 4) X(8, 4) X(16, 4) X(2, 8)   \
        X(4, 8) X(8, 8) X(16, 8)
#define TPCF_BITS_SHAPES(X) X(4) X(8) X(16)

const void *r123_triplet_kernel_ptr(int waves, int copies, bool elastic) {
#define TPCF_PTR(W, K)                                                                                   \
    if (waves == W && copies == K)                                                                       \
        return elastic ? reinterpret_cast<const void *>(&r123_runs_kernel<W, K, true>)                    \
                       : reinterpret_cast<const void *>(&r123_runs_kernel<W, K, false>);
    TPCF_R123_SHAPES(TPCF_PTR)
#undef TPCF_PTR
    return nullptr;
}

static bool use_mfma();

const void *r123_bits_kernel_ptr(int waves, bool elastic) {
#define TPCF_PTR(W)                                                                                      \
    if (waves == W && use_mfma())                                                                        \
        return elastic ? reinterpret_cast<const void *>(&r123_mfma_kernel<W, true>)                       \
                       : reinterpret_cast<const void *>(&r123_mfma_kernel<W, false>);                     \
    if (waves == W)                                                                                      \
        return elastic ? reinterpret_cast<const void *>(&r123_bits_kernel<W, true>)                       \
                       : reinterpret_cast<const void *>(&r123_bits_kernel<W, false>);
    TPCF_BITS_SHAPES(TPCF_PTR)
#undef TPCF_PTR
    return nullptr;
}

int r123_lds_bytes(int n, int copies) { return copies * runs_copy_stride(n, copies) * 4; }
// The bit-mask form runs as matrix-core passes (r123_mfma_kernel);
// TPCF_R123_MFMA=0 selects the VALU passes (r123_bits_kernel) instead.
static bool use_mfma() {
    static const bool on = [] {
        const char *e = std::getenv("TPCF_R123_MFMA");
        return !(e && e[0] == '0');
    }();
    return on;
}

int r123_bits_lds_bytes(int n) {
    return use_mfma() ? mfma_hist_words(n) * 4 + kMfmaXohBytes + kMfmaLutBytes : bits_lds_words(n) * 4;
}

hipError_t launch_r123_choose(const TripletTerm *terms, int nterms, int *counter, float max_bins, int force,
                              int split_bins, float split_mean, hipStream_t st) {
    // the split form needs the matrix-core bit-mask kernel's item counter
    if (!use_mfma()) {
        split_bins = -1;
        if (force == kFormSplit) force = kFormRuns;
    }
    hipLaunchKernelGGL(r123_choose_kernel, dim3(1), dim3(64), 0, st, terms, nterms, counter, max_bins, force,
                       split_bins, split_mean);
    return hipGetLastError();
}

hipError_t launch_r123_triplets(const ElasticCtl &ctl, const TripletItem *items, int nitems, int *counter,
                                const TripletTerm *terms,
                                int max_n, int waves, int copies, int grid, hipStream_t st) {
    if (max_n < 1 || max_n > kMaxLdsR123Bins) return hipErrorInvalidValue;
    const int words = copies * runs_copy_stride(max_n, copies);
    const size_t lds = (size_t)words * 4;
#define TPCF_LAUNCH(W, K)                                                                                    \
    if (waves == W && copies == K) {                                                                         \
        if (ctl.active)                                                                                      \
            hipLaunchKernelGGL((r123_runs_kernel<W, K, true>), dim3(grid), dim3(W * 64), lds, st, ctl, items,   \
                               nitems, counter, terms, words, counter);                                      \
        else                                                                                                 \
            hipLaunchKernelGGL((r123_runs_kernel<W, K, false>), dim3(grid), dim3(W * 64), lds, st, ctl, items,  \
                               nitems, counter, terms, words, counter);                                      \
        return hipGetLastError();                                                                            \
    }
    TPCF_R123_SHAPES(TPCF_LAUNCH)
#undef TPCF_LAUNCH
    return hipErrorInvalidValue;
}

hipError_t launch_r123_bits(const ElasticCtl &ctl, const TripletItem *items, int nitems, int *counter,
                            const TripletTerm *terms, int max_n, int waves, int grid, hipStream_t st) {
    if (max_n < 1 || max_n > kMaxLdsR123Bins) return hipErrorInvalidValue;
    if (use_mfma()) {
        const int words = mfma_hist_words(max_n);
        const size_t lds = (size_t)words * 4 + kMfmaXohBytes + kMfmaLutBytes;
#define TPCF_LAUNCH(W)                                                                                       \
    if (waves == W) {                                                                                        \
        if (ctl.active)                                                                                      \
            hipLaunchKernelGGL((r123_mfma_kernel<W, true>), dim3(grid), dim3(W * 64), lds, st, ctl, items,      \
                               nitems, counter, terms, words, counter);                                      \
        else                                                                                                 \
            hipLaunchKernelGGL((r123_mfma_kernel<W, false>), dim3(grid), dim3(W * 64), lds, st, ctl, items,     \
                               nitems, counter, terms, words, counter);                                      \
        return hipGetLastError();                                                                            \
    }
        TPCF_BITS_SHAPES(TPCF_LAUNCH)
#undef TPCF_LAUNCH
        return hipErrorInvalidValue;
    }
    const int words = bits_lds_words(max_n);
    const size_t lds = (size_t)words * 4;
#define TPCF_LAUNCH(W)                                                                                       \
    if (waves == W) {                                                                                        \
        if (ctl.active)                                                                                      \
            hipLaunchKernelGGL((r123_bits_kernel<W, true>), dim3(grid), dim3(W * 64), lds, st, ctl, items,      \
                               nitems, counter, terms, words, counter);                                      \
        else                                                                                                 \
            hipLaunchKernelGGL((r123_bits_kernel<W, false>), dim3(grid), dim3(W * 64), lds, st, ctl, items,     \
                               nitems, counter, terms, words, counter);                                      \
        return hipGetLastError();                                                                            \
    }
    TPCF_BITS_SHAPES(TPCF_LAUNCH)
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
