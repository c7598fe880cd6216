// kernels_sqt_fast.hip -- FAST form of the (s,q,theta) histograms on gfx950.
//
// Pair stage (sqt_pairs_kernel): for every pair of a job's union layout, in
// FP64, the squared side s, its exact s bin against the squared thresholds,
// and a code: the s bin, kCodeNear (no s bin but short enough to be a q or
// angle side), kCodeFar (longer than s_hi*max(q_hi,1): no entry of any
// triangle containing it can be in range) or kCodeZero (coincident points);
// the side length rounded to FP32; and per tile the 64-bit row and column
// masks of "near" and "has an s bin".
//
// Triplet stage (sqt_fast_kernel): persistent over (term, A, B, C-range)
// items like the (r1,r2,r3) kernel.  Candidate triangles (no far or zero
// side, at least one side with an s bin) come from AND/OR of the masks, 64
// columns at a time; they are appended to a per-wave LDS queue and drained 64
// at a time, so lanes run the entry arithmetic together: sides sorted in
// FP32, q and the cosine-rule value in FP32 with rigorous error bounds, bins
// confirmed against the FP32 copies of the host tables.  A triangle whose side
// order is not certain, or any of whose values lies within its error bound of
// an edge, is recomputed from the FP64 coordinates with the same code as the
// plain kernel, so the histograms equal the FP64 definition bit for bit.
#include "tpcf_internal.h"

namespace tpcf {

namespace {

// Term descriptors are read through the constant address space: invariant
// scalar loads the compiler may repeat where needed instead of keeping every
// field in registers.
typedef const SqtTerm __attribute__((address_space(4))) CTerm;

__device__ inline double sq64f(const double *u, int64_t npad, int64_t a, int64_t b) {
    const double dx = u[a] - u[b];
    const double dy = u[npad + a] - u[npad + b];
    const double dz = u[2 * npad + a] - u[2 * npad + b];
    return (dx * dx + dy * dy) + dz * dz;
}

__device__ inline int bin_asc64(double v, const double *e, int n) {
    if (!(v >= e[0]) || v >= e[n]) return -1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v >= e[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline int bin_cos64(double c, const double *C, int n) {
    if (!(c <= C[0]) || c <= C[n]) return -1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (c <= C[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline void inc(uint32_t *h, int cell) {
    __hip_atomic_fetch_add(h + cell, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// FP64 definition of one triangle (the plain kernel's arithmetic).
__device__ inline void entries_fp64(double s1, double s2, double s3, const double *sT, const double *sQ,
                                    const double *sA, int ns, int nq, int nt, uint32_t *hist) {
    if (s1 == 0.0 || s2 == 0.0 || s3 == 0.0) return;
    double t;
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    if (s2 > s3) { t = s2; s2 = s3; s3 = t; }
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    const double l1 = sqrt(s1), l2 = sqrt(s2), l3 = sqrt(s3);
    const int is1 = bin_asc64(s1, sT, ns);
    if (is1 >= 0) {
        int iq = bin_asc64(l2 / l1, sQ, nq);
        int it = bin_cos64(((s1 + s2) - s3) / ((2.0 * l1) * l2), sA, nt);
        if (iq >= 0 && it >= 0) inc(hist, (is1 * nq + iq) * nt + it);
        iq = bin_asc64(l3 / l1, sQ, nq);
        it = bin_cos64(((s1 + s3) - s2) / ((2.0 * l1) * l3), sA, nt);
        if (iq >= 0 && it >= 0) inc(hist, (is1 * nq + iq) * nt + it);
    }
    const int is2 = bin_asc64(s2, sT, ns);
    if (is2 >= 0) {
        const int iq = bin_asc64(l3 / l2, sQ, nq);
        const int it = bin_cos64(((s2 + s3) - s1) / ((2.0 * l2) * l3), sA, nt);
        if (iq >= 0 && it >= 0) inc(hist, (is2 * nq + iq) * nt + it);
    }
}

// FP64 definition of one JBJ triangle (the plain kernel's arithmetic): the
// shortest side's s bin, q = l1/l2 and (l3 - l2)/l1 on ascending axes.
__device__ inline void entry_jbj_fp64(double s1, double s2, double s3, const double *sT, const double *sQ,
                                      const double *sV, int ns, int nq, int nt, uint32_t *hist) {
    if (s1 == 0.0 || s2 == 0.0 || s3 == 0.0) return;
    double t;
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    if (s2 > s3) { t = s2; s2 = s3; s3 = t; }
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    const double l1 = sqrt(s1), l2 = sqrt(s2), l3 = sqrt(s3);
    const int is = bin_asc64(s1, sT, ns);
    if (is < 0) return;
    const int iq = bin_asc64(l1 / l2, sQ, nq);
    if (iq < 0) return;
    const int iv = bin_asc64((l3 - l2) / l1, sV, nt);
    if (iv < 0) return;
    inc(hist, (is * nq + iq) * nt + iv);
}

constexpr float kU = 5.9604644775390625e-8f;  // 2^-24, FP32 unit roundoff

// Bin hint on an axis: lin (v - lo) * inv, log (log2(v) - lo) * inv; -1 below,
// n at or above the range (NaN gives -1).  Branch-free: fmaxf returns -1 for
// a NaN g, the clamp keeps floor(g) in [-1, n].
__device__ inline int guess_bin(float v, float lo, float inv, int lg, int n) {
    const float g = ((lg ? __builtin_amdgcn_logf(v) : v) - lo) * inv;
    return (int)__builtin_floorf(fminf(fmaxf(g, -1.f), (float)n));
}

// Bin m of v against ascending FP32 edges e[0..n] given the hint m in
// [-1, n]: returns m (or -1 out of range) and sets *unsure unless v is more
// than its error d (plus the edge's own FP32 rounding) away from every edge
// that decides it.  A wrong hint always fails the test.
// Branch-free: the hint m is in [-1, n]; the lower edge e[m] decides unless
// m < 0, the upper edge e[m+1] unless m >= n (the tests of the three cases
// m < 0, m >= n and in range, unified).
// Both edges are read whatever m is, and the conditions combined with
// bitwise operators, so the compiler keeps the test free of exec-mask
// branches.
__device__ inline int bracket_asc(float v, float d, const float *e, int n, int m, bool *unsure) {
    const float a = e[max(m, 0)], b = e[min(m + 1, n)];
    const bool lo_ok = (m < 0) | (v - a > d + kU * fabsf(a));
    const bool hi_ok = (m >= n) | (b - v > d + kU * fabsf(b));
    *unsure = *unsure | !(lo_ok & hi_ok);
    return ((m >= 0) & (m < n)) ? m : -1;
}

// Same for the descending cosine table (bin m iff C[m+1] < c <= C[m]); d
// includes the table's rounding.
__device__ inline int bracket_desc(float c, float d, const float *C, int n, int m, bool *unsure) {
    const float a = C[max(m, 0)], b = C[min(m + 1, n)];
    const bool up_ok = (m < 0) | (a - c > d);
    const bool lo_ok = (m >= n) | (c - b > d);
    *unsure = *unsure | !(up_ok & lo_ok);
    return ((m >= 0) & (m < n)) ? m : -1;
}

__device__ inline void sqt_flush64(uint32_t *h, int cells, unsigned long long *g) {
    __syncthreads();
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const uint32_t v = h[c];
        if (v) {
            atomicAdd(g + c, (unsigned long long)v);
            h[c] = 0u;
        }
    }
    __syncthreads();
}

}  // namespace

// One 64x64 tile (R, C), R <= C, per 256-thread workgroup; thread t owns row
// t/4 and 16 consecutive columns.
// One workgroup per tile (R, C), R <= C.  Lane j of each wave owns column j;
// wave w walks rows 16w .. 16w+15 (the row point is uniform: scalar loads).
// A row's masks are wave ballots; a column's mask bits stay in the lane's
// registers and the four waves' parts are OR-ed at the end, so no mask word
// is written by more than one lane (points in spatial order make whole tiles
// near, which would serialize atomics on shared words).
__global__ __launch_bounds__(256) void sqt_pairs_kernel(SqtPairJob J) {
    const int64_t R = blockIdx.y, C = blockIdx.x;
    if (C < R) return;
    if (J.need && !((J.need[R * J.need_words + (C >> 6)] >> (C & 63)) & 1ull)) {
        // Not computed: marked empty, so work items may span it.
        if (threadIdx.x == 0) J.tile_any[tile_index(R, C, J.nblk)] = 0;
        return;
    }
    __shared__ double sT[kMaxBins + 1];
    __shared__ unsigned long long colN[4][kBlock], colS[4][kBlock], colM[4][kBlock];
    const int t = threadIdx.x, j = t & 63, w = t >> 6;
    for (int m = t; m <= J.ns; m += 256) sT[m] = J.Ts[m];
    __syncthreads();
    const int cntR = J.blk_count[R], cntC = J.blk_count[C];
    const int64_t gj = C * kBlock + j;
    const int64_t tile = tile_index(R, C, J.nblk);
    uint8_t *codes = J.codes + tile * kTileBytes;
    float *dists = J.dists + tile * kTileBytes;
    unsigned long long *masks = J.masks + tile * kMaskWords;
    unsigned long long cn = 0ull, cs = 0ull, cm = 0ull;
    for (int r = 0; r < 16; ++r) {
        const int i = w * 16 + r;
        uint32_t code = kCodeFar;
        float d = 0.f;
        bool mid = false;
        if (i < cntR && j < cntC && (R != C || j > i)) {
            const double s = sq64f(J.u64, J.npad, R * kBlock + i, gj);
            d = (float)sqrt(s);
            if (s == 0.0) {
                code = kCodeZero;
            } else if (!(s > J.far_s)) {
                const int m = bin_asc64(s, sT, J.ns);
                code = m >= 0 ? (uint32_t)m : kCodeNear;
                mid = !(s > J.mid_s) || m >= 0;
            }
        }
        codes[i * kBlock + j] = (uint8_t)code;
        dists[i * kBlock + j] = d;
        const bool near = code < kCodeZero, sbin = code < kCodeNear;
        const unsigned long long bn = __ballot(near), bs = __ballot(sbin), bm = __ballot(mid);
        if (j == 0) {
            masks[kMaskRowNear + i] = bn;
            masks[kMaskRowShort + i] = bs;
            masks[kMaskRowMid + i] = bm;
        }
        cn |= (unsigned long long)near << i;
        cs |= (unsigned long long)sbin << i;
        cm |= (unsigned long long)mid << i;
    }
    colN[w][j] = cn;
    colS[w][j] = cs;
    colM[w][j] = cm;
    __syncthreads();
    // Whether any pair of the tile is near (no candidate has a side in a
    // tile without one).
    const int tile_near = __syncthreads_or(cn != 0ull ? 1 : 0);
    if (t == 0) J.tile_any[tile_index(R, C, J.nblk)] = tile_near ? 1 : 0;
    if (t < kBlock) {
        masks[kMaskColNear + t] = (colN[0][t] | colN[1][t]) | (colN[2][t] | colN[3][t]);
        masks[kMaskColShort + t] = (colS[0][t] | colS[1][t]) | (colS[2][t] | colS[3][t]);
        masks[kMaskColMid + t] = (colM[0][t] | colM[1][t]) | (colM[2][t] | colM[3][t]);
    }
}

namespace {

// Makes one wave's LDS stores visible to the other lanes of the same wave.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A queued candidate is k | j << 6 | i << 12 | (C - c0) << 18 for point i of
// block A, j of B and k of C.  With t = e >> 6 the byte offsets of its sides
// in the item's tiles are t & 0xFFF = i*64 + j in (A,B), and two bit-field
// inserts: (C - c0)*4096 + i*64 + k in the tiles of row A from (A, c0) on, and
// (C - c0)*4096 + j*64 + k in those of row B from (B, c0) on (the tiles of one
// row are consecutive in C).
struct SqtItemCtx {
    const uint8_t *cAB, *cAC, *cBC;  // codes of tiles (A,B), (A,c0), (B,c0)
    const float *dAB, *dAC, *dBC;    // FP32 side lengths of the same tiles
    int A, B, c0;
};

// A queued candidate triangle with the codes and FP32 lengths of its sides
// (loads in flight until first used).
struct SqtPending {
    uint32_t e;
    uint32_t xc, yc, zc;
    float xd, yd, zd;
    bool live;
};

// Starts loading the sides of ring entries head .. head+cnt-1, one per lane.
__device__ inline SqtPending sqt_gather(const SqtItemCtx &X, const uint32_t *Q, uint32_t head, uint32_t qmask,
                                        int cnt, int lane) {
    SqtPending p;
    p.live = lane < cnt;
    p.e = p.live ? Q[(head + (uint32_t)lane) & qmask] : 0u;
    p.xc = p.yc = p.zc = kCodeFar;
    p.xd = p.yd = p.zd = 0.f;
    if (p.live) {
        const uint32_t e = p.e, t = e >> 6;
        const uint32_t oAB = t & 0xFFFu;
        const uint32_t oAC = (e & 0x3Fu) | (t & ~0x3Fu);
        const uint32_t oBC = (e & 0xFFFu) | (t & ~0xFFFu);
        p.xc = X.cAB[oAB];
        p.yc = X.cAC[oAC];
        p.zc = X.cBC[oBC];
        p.xd = X.dAB[oAB];
        p.yd = X.dAC[oAC];
        p.zd = X.dBC[oBC];
    }
    return p;
}

// Orders (a, ka), (b, kb) by length without branches.
__device__ __forceinline__ void sort2(float &a, float &b, uint32_t &ka, uint32_t &kb) {
    const bool sw = a > b;
    const float lo = sw ? b : a, hi = sw ? a : b;
    const uint32_t klo = sw ? kb : ka, khi = sw ? ka : kb;
    a = lo;
    b = hi;
    ka = klo;
    kb = khi;
}

// Cells of the theta hint table over the cosine-rule value c in [-1, 1].
constexpr int kCosCells = 2048;

// Bin hint of the cosine-rule value c: the table holds, per cell of
// [-1, 1], 1 + the bin of the cell's upper end (-1 above the range, nt below
// it); with at most one edge in the cell, c is in that bin or the next, and
// one comparison with the next edge decides.  Only a hint: bracket_desc
// confirms it.
__device__ __forceinline__ int cos_hint(float c, const uint8_t *tcos, const float *fA, int nt) {
    const int idx = __builtin_amdgcn_fmed3f(0.f, (c + 1.f) * (0.5f * kCosCells), (float)(kCosCells - 1));
    const int m0 = (int)tcos[idx] - 1;
    const float next = fA[min(m0 + 1, nt)];
    return m0 + (int)((m0 < nt) & (c <= next));
}

// One entry in FP32 from q = b * (1/a) and c = num * rden, with a <= b the
// FP32 lengths (each within u of the FP64 side), 1/a from v_rcp_f32 (1 ulp =
// at most 2u), num = (sa + sb) - sc and rden = (0.5 * (1/a)) * (1/b), two
// v_rcp_f32 values and a rounding: within 5u of 1/(2ab).  Error bounds
// against the FP64 definition:
//   q: within (2u sides + 3u quotient) q, 6u q used;
//   c: the squares are within 3u, so num is within 4u (sa+sb+sc) + u |num|
//   of the FP64 value, and 2ab within 2u: |c - c64| <= 4u (sa+sb+sc)/(2ab)
//   + 9u |c|; 5u (sa+sb+sc) rden, 10u |c| and 2u for the table's rounding
//   used.
// Returns the histogram cell, -1 for "out of range" (certain), and sets
// *unsure when some value lies within its error bound of a deciding edge.
__device__ inline int entry32(int sbin, float q, float num, float rden, float ssum, const float *fQ,
                              const float *fA, const uint8_t *tcos, CTerm &T, bool *unsure) {
    const int nq = T.nq, nt = T.nt;
    const int mq = bracket_asc(q, 6.f * kU * q, fQ, nq, guess_bin(q, T.q_lo, T.q_inv, T.q_log, nq), unsure);
    const float c = num * rden;
    const float dc = 5.f * kU * ssum * rden + 10.f * kU * fabsf(c) + 2.f * kU;
    const int mt = bracket_desc(c, dc, fA, nt, cos_hint(c, tcos, fA, nt), unsure);
    return ((mq < 0) | (mt < 0)) ? -1 : (int)__umul24(__umul24((uint32_t)sbin, nq) + mq, nt) + mt;
}

// Histogram entries of a gathered candidate (FP32 with FP64 recheck), for
// PARAM = kSqt or kJbj.
template <int PARAM>
__device__ inline void sqt_entries(CTerm &T, const SqtItemCtx &X, const SqtPending &p, const double *sT,
                                   const double *sQ, const double *sA, const float *fQ, const float *fA,
                                   const uint8_t *tcos, uint32_t *hist, unsigned long long &nre) {
    if (!p.live) return;
    // Sort the sides by FP32 length, carrying their codes.
    float l1 = p.xd, l2 = p.yd, l3 = p.zd;
    uint32_t k1 = p.xc, k2 = p.yc, k3 = p.zc;
    sort2(l1, l2, k1, k2);
    sort2(l2, l3, k2, k3);
    sort2(l1, l2, k1, k2);
    const int nq = T.nq, nt = T.nt;
    bool unsure = T.force_fp64 != 0;
    // Side order certain only when lengths differ by more than their FP32
    // rounding; lengths far outside [1e-18, 1e18] go to FP64 (no FP32
    // underflow or overflow below).
    if (!(l2 - l1 > 4.f * kU * l2) || !(l3 - l2 > 4.f * kU * l3) || !(l1 > 1e-18f) || !(l3 < 1e18f)) unsure = true;
    int cell[3] = {-1, -1, -1};
    if (!unsure && PARAM == kJbj) {
        // One entry: the shortest side's s bin, q = l1/l2 and v =
        // (l3 - l2)/l1.  With every FP32 side within u of the FP64 one:
        // q = l1 * (1/l2) is within (2u + 3u) q (6u q used); the FP32
        // difference l3 - l2 is within u (l2 + l3) + u |l3 - l2| of the
        // FP64 one, and v within u (l2 + l3)/l1 + 5u |v| (2u (l2 + l3)/l1
        // and 6u |v| used); each edge's own FP32 rounding is added by the
        // bracket test.
        if (k1 < kCodeNear) {
            const float r1 = __builtin_amdgcn_rcpf(l1);
            const float q = l1 * __builtin_amdgcn_rcpf(l2);
            const int mq = bracket_asc(q, 6.f * kU * q, fQ, nq, guess_bin(q, T.q_lo, T.q_inv, T.q_log, nq), &unsure);
            const float v = (l3 - l2) * r1;
            const float dv = 2.f * kU * (l2 + l3) * r1 + 6.f * kU * fabsf(v);
            const int mv = bracket_asc(v, dv, fA, nt, guess_bin(v, T.t_lo, T.t_inv, T.t_log, nt), &unsure);
            if (mq >= 0 && mv >= 0) cell[0] = (int)__umul24(__umul24(k1, nq) + mq, nt) + mv;
        }
    } else if (!unsure) {
        const float s1 = l1 * l1, s2 = l2 * l2, s3 = l3 * l3;
        const float ssum = (s1 + s2) + s3;
        // three reciprocals serve the three entries' q and 1/(2ab)
        const float r1 = __builtin_amdgcn_rcpf(l1), r2 = __builtin_amdgcn_rcpf(l2), r3 = __builtin_amdgcn_rcpf(l3);
        const float h1 = 0.5f * r1, h2 = 0.5f * r2;
        if (k1 < kCodeNear) {
            cell[0] = entry32((int)k1, l2 * r1, (s1 + s2) - s3, h1 * r2, ssum, fQ, fA, tcos, T, &unsure);
            cell[1] = entry32((int)k1, l3 * r1, (s1 + s3) - s2, h1 * r3, ssum, fQ, fA, tcos, T, &unsure);
        }
        if (k2 < kCodeNear)
            cell[2] = entry32((int)k2, l3 * r2, (s2 + s3) - s1, h2 * r3, ssum, fQ, fA, tcos, T, &unsure);
    }
    if (!unsure) {
#pragma unroll
        for (int m = 0; m < 3; ++m)
            if (cell[m] >= 0) inc(hist, cell[m]);
    } else {
        ++nre;
        const uint32_t e = p.e;
        const int k = e & 63, j = (e >> 6) & 63, i = (e >> 12) & 63, C = X.c0 + (int)(e >> 18);
        const int64_t gi = (int64_t)X.A * kBlock + i, gj = (int64_t)X.B * kBlock + j, gk = (int64_t)C * kBlock + k;
        const double sij = sq64f(T.u64, T.npad, gi, gj), sik = sq64f(T.u64, T.npad, gi, gk),
                     sjk = sq64f(T.u64, T.npad, gj, gk);
        if (PARAM == kJbj)
            entry_jbj_fp64(sij, sik, sjk, sT, sQ, sA, T.ns, nq, nt, hist);
        else
            entries_fp64(sij, sik, sjk, sT, sQ, sA, T.ns, nq, nt, hist);
    }
}

}  // namespace

// Inclusive sum of x over the lanes 0..lane of the wave: DPP shifts within
// each row of 16 lanes, then the totals of rows 0..2 (read from lanes 15, 31,
// 47) added to the rows above them.
__device__ inline uint32_t wave_inclusive_sum(uint32_t x, int lane) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    const uint32_t r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31),
                   r2 = __builtin_amdgcn_readlane(x, 47);
    return x + (lane >= 16 ? r0 : 0u) + (lane >= 32 ? r1 : 0u) + (lane >= 48 ? r2 : 0u);
}

// Candidate columns j of row i for this lane: sides (i,j), (i,k), (j,k) all
// near, at least one with an s bin and at least two mid.
__device__ inline unsigned long long row_candidates(unsigned long long xn, unsigned long long xs,
                                                    unsigned long long xm, int i, unsigned long long yn,
                                                    unsigned long long ys, unsigned long long ym,
                                                    unsigned long long zn, unsigned long long zs,
                                                    unsigned long long zm) {
    const unsigned long long yb = 1ull << i;
    if (!(yn & yb)) return 0ull;
    const unsigned long long shortx = (ys & yb) ? ~0ull : (xs | zs);
    const unsigned long long mid2 = (ym & yb) ? (xm | zm) : (xm & zm);
    return xn & zn & shortx & mid2;
}

namespace {

// LDS tables of the current term and the (A,B) row masks of the current item.
struct SqtShared {
    double *sT, *sQ, *sA;
    float *fQ, *fA;
    uint8_t *tcos;
    const unsigned long long *xnear, *xshort, *xmid;
};

// The candidates of one item (term, A, B, C-range) for PARAM, found with the
// pair stage's bit masks and appended to this wave's ring (Q, qcap words),
// drained 64 at a time.
template <int PARAM, int WAVES>
__device__ inline void sqt_item(CTerm &T, const SqtItemCtx &X, int c1, const SqtShared &S, uint32_t *hist,
                                uint32_t *Q, uint32_t qcap, int lane, int wave, unsigned long long &nre) {
    const uint32_t qmask = qcap - 1u;
    const int64_t nblk = T.nblk;
    // This wave's ring of queued candidates: entries [qhead, qtail)
    // (indices mod qcap).
    uint32_t qhead = 0u, qtail = 0u;
    // The last 64 queued entries, their sides in flight: their entries are
    // computed at the next drain (or the item's end), so the loads overlap
    // the scan in between.
    SqtPending pend;
    pend.live = false;
    auto drain_ready = [&]() {
        while (qtail - qhead >= (uint32_t)kBlock) {
            wave_sync();
            sqt_entries<PARAM>(T, X, pend, S.sT, S.sQ, S.sA, S.fQ, S.fA, S.tcos, hist, nre);
            pend = sqt_gather(X, Q, qhead, qmask, kBlock, lane);
            qhead += kBlock;
        }
    };
    for (int C = X.c0 + wave; C < c1; C += WAVES) {
        // Lane k's column of tiles (B,C) and (A,C): bits j of side (j,k)
        // and bits i of side (i,k).
        if (!(T.tile_any[tile_index(X.A, C, nblk)] & T.tile_any[tile_index(X.B, C, nblk)])) continue;
        const unsigned long long *mBC = T.masks + tile_index(X.B, C, nblk) * kMaskWords;
        const unsigned long long zn = mBC[kMaskColNear + lane], zs = mBC[kMaskColShort + lane],
                                 zm = mBC[kMaskColMid + lane];
        if (__ballot(zn != 0ull) == 0ull) continue;
        const unsigned long long *mAC = T.masks + tile_index(X.A, C, nblk) * kMaskWords;
        const unsigned long long yn = mAC[kMaskColNear + lane], ys = mAC[kMaskColShort + lane],
                                 ym = mAC[kMaskColMid + lane];
        const uint32_t lbase = (uint32_t)lane | (uint32_t)(C - X.c0) << 18;
        for (int i = 0; i < kBlock; ++i) {
            const unsigned long long xn = S.xnear[i];
            if (xn == 0ull) continue;
            unsigned long long m = row_candidates(xn, S.xshort[i], S.xmid[i], i, yn, ys, ym, zn, zs, zm);
            const uint32_t cnt = (uint32_t)__popcll(m);
            const uint32_t incl = wave_inclusive_sum(cnt, lane);
            const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
            if (total == 0u) continue;
            const uint32_t base = lbase | (uint32_t)i << 12;
            if (qtail - qhead + total <= qcap) {
                // Compaction: every lane writes its candidates at its offset
                // in the row's block of the ring.
                uint32_t pos = qtail + incl - cnt;
                while (m != 0ull) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1ull;
                    Q[pos & qmask] = base | (uint32_t)j << 6;
                    ++pos;
                }
                qtail += total;
            } else {
                // The row does not fit: one entry per lane per round,
                // draining in between.
                for (;;) {
                    const bool has = m != 0ull;
                    const unsigned long long act = __ballot(has);
                    if (act == 0ull) break;
                    if (has) {
                        const int j = __builtin_ctzll(m);
                        m &= m - 1ull;
                        const uint32_t slot = qtail + __builtin_amdgcn_mbcnt_hi(
                                                          (uint32_t)(act >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
                        Q[slot & qmask] = base | (uint32_t)j << 6;
                    }
                    qtail += (uint32_t)__popcll(act);
                    drain_ready();
                }
            }
            drain_ready();
        }
    }
    sqt_entries<PARAM>(T, X, pend, S.sT, S.sQ, S.sA, S.fQ, S.fA, S.tcos, hist, nre);
    if (qtail != qhead) {
        wave_sync();
        sqt_entries<PARAM>(T, X, sqt_gather(X, Q, qhead, qmask, (int)(qtail - qhead), lane), S.sT, S.sQ, S.sA, S.fQ,
                           S.fA, S.tcos, hist, nre);
        wave_sync();
    }
}

}  // namespace

// Persistent over (term, A, B, C-range) items.  Candidate triangles (no far
// or zero side, at least one side with an s bin) are found with the pair
// stage's bit masks: per item, for every row i of block A, the columns j whose
// side (i,j) is near / has an s bin (shared in LDS); per C block and lane k,
// the rows j of side (j,k) and the rows i of side (i,k) likewise (64-bit
// words in VGPRs).  A row's candidates for lane k are then
//   X_near[i] & Z_near & Y_near(i) & (X_short[i] | Z_short | Y_short(i)).
// Each wave takes the C blocks c0 + wave, c0 + wave + WAVES, ... and appends
// its candidates to an LDS queue, drained 64 at a time with every lane busy.
// The per-item work is compiled once per parametrization ((s,q,theta) and
// JBJ jobs may share a launch).
template <int WAVES, bool ELASTIC>
__global__ __launch_bounds__(WAVES * 64) void sqt_fast_kernel(SqtFastArgs P) {
    extern __shared__ uint32_t hist[];
    __shared__ double sT[kMaxBins + 1], sQ[kMaxBins + 1], sA[kMaxBins + 1];
    __shared__ float fQ[kMaxBins + 1], fA[kMaxBins + 1];
    __shared__ uint8_t tcos[kCosCells];
    __shared__ unsigned long long xnear[kBlock], xshort[kBlock], xmid[kBlock];
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int c = threadIdx.x; c < P.lds_cells; c += blockDim.x) hist[c] = 0u;
    const uint32_t qcap = (uint32_t)P.qcap;
    uint32_t *Q = hist + ((P.lds_cells + 3) & ~3) + wave * qcap;
    const SqtShared S{sT, sQ, sA, fQ, fA, tcos, xnear, xshort, xmid};
    // The current term (its job's tables in LDS); flushing writes and clears
    // the cells the term used, so the next term starts from zero.
    int cur = -1;
    CTerm *Tp = (CTerm *)(P.terms);
    unsigned long long since = 0, nre = 0;
    bool left = false;
    if constexpr (ELASTIC)
        if (threadIdx.x == 0) atomicAdd(P.ctl.active, 1);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            if constexpr (ELASTIC) left = elastic_leave(P.ctl);
            s_item = left ? P.nitems : atomicAdd(P.counter, 1);
        }
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= P.nitems) break;
        const TripletItem I = P.items[it];
        const int term = __builtin_amdgcn_readfirstlane(I.term);
        SqtItemCtx X;
        X.A = __builtin_amdgcn_readfirstlane(I.A);
        X.B = __builtin_amdgcn_readfirstlane(I.B);
        X.c0 = __builtin_amdgcn_readfirstlane(I.c0);
        const int c1 = __builtin_amdgcn_readfirstlane(I.c1);
        const unsigned long long add = 3ull * (unsigned long long)(c1 - X.c0) * (kBlock * kBlock * kBlock);
        if (term != cur || since + add > 0xC0000000ull) {
            if (cur >= 0) sqt_flush64(hist, Tp->cells, Tp->hist);
            if (term != cur) {
                if (nre) atomicAdd(Tp->rechecked, nre);
                nre = 0;
                Tp = (CTerm *)(P.terms) + term;
                CTerm &T = *Tp;
                for (int m = threadIdx.x; m <= kMaxBins; m += blockDim.x) {
                    if (m <= T.ns) sT[m] = T.Ts[m];
                    if (m <= T.nq) {
                        sQ[m] = T.Eq[m];
                        fQ[m] = (float)T.Eq[m];
                    }
                    if (m <= T.nt) {
                        sA[m] = T.Et[m];
                        fA[m] = (float)T.Et[m];
                    }
                }
                __syncthreads();
                // theta hint table: 1 + the bin (descending cosine edges fA)
                // of each cell's upper end
                if (T.param == kSqt) {
                    const int nt = T.nt;
                    for (int c = threadIdx.x; c < kCosCells; c += blockDim.x) {
                        const float v = -1.f + (float)(c + 1) * (2.f / kCosCells);
                        int m;
                        if (v > fA[0]) {
                            m = -1;
                        } else if (v <= fA[nt]) {
                            m = nt;
                        } else {
                            int lo = 0, hi = nt;  // fA[lo] >= v > fA[hi]
                            while (hi - lo > 1) {
                                const int mid = (lo + hi) >> 1;
                                if (v <= fA[mid]) lo = mid;
                                else hi = mid;
                            }
                            m = lo;
                        }
                        tcos[c] = (uint8_t)(m + 1);
                    }
                }
            }
            cur = term;
            since = 0;
        }
        since += add;
        CTerm &T = *Tp;
        const int64_t nblk = T.nblk;
        // No candidate in the item when its (A, B) tile has no near pair
        // (or was not computed).
        if (!T.tile_any[tile_index(X.A, X.B, nblk)]) continue;
        const int64_t tAB = tile_index(X.A, X.B, nblk) * kTileBytes;
        const int64_t tAC = tile_index(X.A, X.c0, nblk) * kTileBytes;
        const int64_t tBC = tile_index(X.B, X.c0, nblk) * kTileBytes;
        X.cAB = T.codes + tAB;
        X.cAC = T.codes + tAC;
        X.cBC = T.codes + tBC;
        X.dAB = T.dists + tAB;
        X.dAC = T.dists + tAC;
        X.dBC = T.dists + tBC;
        const unsigned long long *mAB = T.masks + tile_index(X.A, X.B, nblk) * kMaskWords;
        for (int i = threadIdx.x; i < kBlock; i += blockDim.x) {
            xnear[i] = mAB[kMaskRowNear + i];
            xshort[i] = mAB[kMaskRowShort + i];
            xmid[i] = mAB[kMaskRowMid + i];
        }
        __syncthreads();
        if (T.param == kJbj)
            sqt_item<kJbj, WAVES>(T, X, c1, S, hist, Q, qcap, lane, wave, nre);
        else
            sqt_item<kSqt, WAVES>(T, X, c1, S, hist, Q, qcap, lane, wave, nre);
    }
    if (cur >= 0) {
        sqt_flush64(hist, Tp->cells, Tp->hist);
        if (nre) atomicAdd(Tp->rechecked, nre);
    }
    if constexpr (ELASTIC)
        if (threadIdx.x == 0 && !left) atomicSub(P.ctl.active, 1);
}

hipError_t launch_sqt_pairs(const SqtPairJob &J, hipStream_t st) {
    hipLaunchKernelGGL(sqt_pairs_kernel, dim3((unsigned)J.nblk, (unsigned)J.nblk), dim3(256), 0, st, J);
    return hipGetLastError();
}

const void *sqt_fast_kernel_ptr(int waves, bool elastic) {
    if (waves == 8)
        return elastic ? reinterpret_cast<const void *>(&sqt_fast_kernel<8, true>)
                       : reinterpret_cast<const void *>(&sqt_fast_kernel<8, false>);
    if (waves == 16)
        return elastic ? reinterpret_cast<const void *>(&sqt_fast_kernel<16, true>)
                       : reinterpret_cast<const void *>(&sqt_fast_kernel<16, false>);
    return nullptr;
}

size_t sqt_fast_lds_bytes(int cells, int waves, int qcap) {
    return ((size_t)(cells + 3) & ~(size_t)3) * 4 + (size_t)waves * qcap * 4;
}

hipError_t launch_sqt_fast(const SqtFastArgs &P, int waves, int grid, hipStream_t st) {
    if (P.qcap < 128 || (P.qcap & (P.qcap - 1))) return hipErrorInvalidValue;
    const size_t lds = sqt_fast_lds_bytes(P.lds_cells, waves, P.qcap);
    const bool el = P.ctl.active != nullptr;
    if (waves == 8 && el)
        hipLaunchKernelGGL((sqt_fast_kernel<8, true>), dim3(grid), dim3(512), lds, st, P);
    else if (waves == 8)
        hipLaunchKernelGGL((sqt_fast_kernel<8, false>), dim3(grid), dim3(512), lds, st, P);
    else if (waves == 16 && el)
        hipLaunchKernelGGL((sqt_fast_kernel<16, true>), dim3(grid), dim3(1024), lds, st, P);
    else if (waves == 16)
        hipLaunchKernelGGL((sqt_fast_kernel<16, false>), dim3(grid), dim3(1024), lds, st, P);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace tpcf
