// kernels_sqt.hip -- gfx950 kernels of the triangle parametrizations:
// (s, q, theta) with three vertex entries per triangle
// (src/3pcf_C_version_triangle_parametrization.cc:192-236) and the JBJ
// single-entry variant (src/3pcf_C_version_triangle_parametrization_JBJ.cc:
// 192-198).
//
// sqt_plain_kernel: every triplet recomputes its three FP64 squared sides,
// sorts them, takes IEEE sqrt and divisions in FP64 and bins against the
// host edge tables -- the reference's distance() in FP64.  A workgroup keeps
// a privatized u32 histogram in LDS and folds it into uint64 device memory
// before the counters could wrap.
//
// Compile with -ffp-contract=off: the arithmetic must be the unfused IEEE
// sequence of the CPU definition.
#include "tpcf_internal.h"

namespace tpcf {

namespace {

__device__ inline double sq64(const double *u, int64_t npad, int64_t a, int64_t b) {
    const double dx = u[a] - u[b];
    const double dy = u[npad + a] - u[npad + b];
    const double dz = u[2 * npad + a] - u[2 * npad + b];
    return (dx * dx + dy * dy) + dz * dz;
}

// Bin of v against ascending edges e[0..n]; -1 outside [e0, en).
__device__ inline int bin_asc(double v, const double *e, int n) {
    if (!(v >= e[0]) || v >= e[n]) return -1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (v >= e[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Bin of a cosine against the descending table C[m] = cos(pi*t[m]):
// bin m iff C[m+1] < c <= C[m].
__device__ inline int bin_cos(double c, const double *C, int n) {
    if (!(c <= C[0]) || c <= C[n]) return -1;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (c <= C[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline void lds_inc(uint32_t *h, int cell) {
    __hip_atomic_fetch_add(h + cell, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline void sqt_flush(uint32_t *h, int cells, unsigned long long *g) {
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

// The FP64 definition of one triangle's entries, from its three squared
// sides in any order: sort, IEEE sqrt and divisions, bins against the host
// tables (sT: squared s thresholds, sQ: q edges, sA: cosine table for SQT or
// (r3-r2)/r1 edges for JBJ).  Triangles with a zero side are dropped.
template <int PARAM>
__device__ inline void sqt_entries_fp64(double s1, double s2, double s3, const double *sT, const double *sQ,
                                        const double *sA, int ns, int nq, int nt, uint32_t *hist) {
    if (s1 == 0.0 || s2 == 0.0 || s3 == 0.0) return;
    double t;
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    if (s2 > s3) { t = s2; s2 = s3; s3 = t; }
    if (s1 > s2) { t = s1; s1 = s2; s2 = t; }
    const double l1 = sqrt(s1), l2 = sqrt(s2), l3 = sqrt(s3);
    const int is1 = bin_asc(s1, sT, ns);
    if (PARAM == kJbj) {
        if (is1 < 0) return;
        const int iq = bin_asc(l1 / l2, sQ, nq);
        const int iv = bin_asc((l3 - l2) / l1, sA, nt);
        if (iq >= 0 && iv >= 0) lds_inc(hist, (is1 * nq + iq) * nt + iv);
        return;
    }
    if (is1 >= 0) {
        int iq = bin_asc(l2 / l1, sQ, nq);
        int it = bin_cos(((s1 + s2) - s3) / ((2.0 * l1) * l2), sA, nt);
        if (iq >= 0 && it >= 0) lds_inc(hist, (is1 * nq + iq) * nt + it);
        iq = bin_asc(l3 / l1, sQ, nq);
        it = bin_cos(((s1 + s3) - s2) / ((2.0 * l1) * l3), sA, nt);
        if (iq >= 0 && it >= 0) lds_inc(hist, (is1 * nq + iq) * nt + it);
    }
    const int is2 = bin_asc(s2, sT, ns);
    if (is2 >= 0) {
        const int iq = bin_asc(l3 / l2, sQ, nq);
        const int it = bin_cos(((s2 + s3) - s1) / ((2.0 * l2) * l3), sA, nt);
        if (iq >= 0 && it >= 0) lds_inc(hist, (is2 * nq + iq) * nt + it);
    }
}

}  // namespace

// One entry point for both parametrizations: PARAM == kSqt or kJbj.  Lane k owns point k of block C; triplets are (i in A,
// j in B, k in C) with the uniqueness constraints of equal blocks.
template <int PARAM>
__global__ __launch_bounds__(256) void sqt_plain_kernel(SqtArgs P) {
    extern __shared__ uint32_t hist[];
    __shared__ double sT[kMaxBins + 1], sQ[kMaxBins + 1], sA[kMaxBins + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ns = P.ns, nq = P.nq, nt = P.nt;
    const int cells = ns * nq * nt;
    for (int m = threadIdx.x; m <= kMaxBins; m += blockDim.x) {
        if (m <= ns) sT[m] = P.Ts[m];
        if (m <= nq) sQ[m] = P.Eq[m];
        if (m <= nt) sA[m] = P.Et[m];
    }
    for (int c = threadIdx.x; c < cells; c += blockDim.x) hist[c] = 0u;
    __syncthreads();
    const int64_t np = P.npad;
    // Each item adds at most 3 * 64^3 per C block; flush far below 2^32.
    unsigned long long since = 0;
    for (int it = blockIdx.x; it < P.nitems; it += gridDim.x) {
        const TripletItem I = P.items[it];
        const unsigned long long add = 3ull * (I.c1 - I.c0) * kBlock * kBlock * kBlock;
        if (since + add > 0xC0000000ull) {
            sqt_flush(hist, cells, P.hist);
            since = 0;
        }
        since += add;
        const int A = I.A, B = I.B;
        const int cntA = P.blk_count[A], cntB = P.blk_count[B];
        for (int C = I.c0; C < I.c1; ++C) {
            const int cntC = P.blk_count[C];
            const int64_t gk = (int64_t)C * kBlock + lane;
            const bool kval = lane < cntC;
            for (int i = wave; i < cntA; i += 4) {
                const int64_t gi = (int64_t)A * kBlock + i;
                if (A == C && lane <= i) continue;
                for (int j = (A == B ? i + 1 : 0); j < cntB; ++j) {
                    if (!kval || (B == C && lane <= j)) continue;
                    const int64_t gj = (int64_t)B * kBlock + j;
                    sqt_entries_fp64<PARAM>(sq64(P.u64, np, gi, gj), sq64(P.u64, np, gi, gk), sq64(P.u64, np, gj, gk),
                                            sT, sQ, sA, ns, nq, nt, hist);
                }
            }
        }
    }
    sqt_flush(hist, cells, P.hist);
}

hipError_t launch_sqt_plain(const SqtArgs &P, int param, int grid, hipStream_t st) {
    const size_t lds = (size_t)P.ns * P.nq * P.nt * 4;
    if (param == kSqt)
        hipLaunchKernelGGL(sqt_plain_kernel<kSqt>, dim3(grid), dim3(256), lds, st, P);
    else
        hipLaunchKernelGGL(sqt_plain_kernel<kJbj>, dim3(grid), dim3(256), lds, st, P);
    return hipGetLastError();
}

const void *sqt_plain_kernel_ptr(int param) {
    return param == kSqt ? reinterpret_cast<const void *>(&sqt_plain_kernel<kSqt>)
                                   : reinterpret_cast<const void *>(&sqt_plain_kernel<kJbj>);
}

}  // namespace tpcf
