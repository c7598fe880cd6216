// kernels_pairs.hip -- gfx950 pair-count histograms (the two-point
// function's DD / DR / RR).
//
// Reference: the pair loop of src/2pcf_C_version.cc:358-387 with distance()
// (:89-113) and distance_to_bin() (:57-86): FP32 distances, 700 linear bins
// on [0, 70).  Here every pair is classified in FP64 exactly as the CPU
// definition, ((dx*dx + dy*dy) + dz*dz) against the squared thresholds of the
// axis, for any runtime axis of up to kMaxPairBins bins; (a, a) terms count
// i < j (no self pairs), (a, b) terms the full product.
//
// FAST (pair_hist_kernel<.., CULL = false>): every pair's squared distance
// in FP64 over Hilbert-ordered copies of the segments (the order does not
// change the counts); persistent workgroups pull items -- an i block of
// WAVES*64*IPL points of segment a against a j range of segment b -- from an
// atomic counter.  Each lane keeps IPL i points in registers; the j point is
// the same for the whole wave and comes through scalar loads.  The FP32 sqrt
// (linear axes) or log2 (log axes) of the squared distance gives a bin guess
// k; the guess is confirmed against the FP64 table D[k] <= s < D[k+1] held in
// LDS and walked to the right bin in the rare case FP32 rounding put it off
// by one, so the FP32 arithmetic only ever selects which comparison to make.
// Counts go to privatized 32-bit LDS copies of the histogram (copy = lane %
// copies), flushed with 64-bit atomics on a term change, before any copy
// could wrap, and at exit.
//
// CELLS (pair_hist_kernel<.., CULL = true>): the same over Hilbert-ordered
// copies of the segments; each wave takes the box of its own points and
// runs only the 64-point j runs whose box gap reaches below the last squared
// threshold (a lower bound of every pair's squared distance).
//
// PLAIN (pair_plain_kernel): one item per workgroup of 256 lanes, one i point
// per lane, a binary search of the FP64 thresholds in global memory per pair,
// one LDS histogram flushed at the end of the item.
#include <climits>

#include "tpcf_internal.h"

namespace tpcf {
namespace {

typedef const double __attribute__((address_space(4))) cdouble;
typedef const PairHistTerm __attribute__((address_space(4))) CPairTerm;
typedef const PairHistItem __attribute__((address_space(4))) CPairItem;
typedef const double __attribute__((address_space(1))) gdouble;

// A value every lane of the wave reads at the same address: scalar load.
__device__ __forceinline__ double ld_uniform(const double *p, int j) { return ((cdouble *)p)[j]; }

// Table index guess k in [0, n+1] for squared distance f (FP32).  Linear
// axis: k = floor(d / w + 1 - lo / w); log axis: k = floor(log2(d) * c + 1 -
// log2(lo) * c) with c = ln 2 / (log-bin width); both written in terms of
// f = d^2 (h_mul carries the factor 1/2 of the log form).  Only a hint.
template <bool LOG>
__device__ __forceinline__ int bin_guess(float f, float mul, float add, float top) {
    const float g = LOG ? __builtin_amdgcn_logf(f) : __builtin_amdgcn_sqrtf(f);
    return (int)__builtin_amdgcn_fmed3f(__builtin_fmaf(g, mul, add), 0.0f, top);
}

// Table D: D[0] = -inf, D[1..n+1] = squared thresholds T[0..n], D[n+2] =
// +inf.  The table index of s is the k in [0, n+1] with D[k] <= s < D[k+1];
// bin k-1 when 1 <= k <= n, out of range otherwise.  bin_walk moves a guess
// that fails that test to the right k (NaN ends at 0).
__device__ __noinline__ int bin_walk(double s, int k, const double *D, int n) {
    while (k > 0 && !(s >= D[k])) --k;
    while (k <= n && s >= D[k + 1]) ++k;
    return k;
}

// Pairs of the lane's IPL points with U consecutive j points (j .. j+U-1):
// all squared distances first; if none of the wave's pairs is below the last
// squared threshold Tn (out of range by the FP64 definition itself) the
// batch ends there; otherwise all guesses, then all table reads, so the LDS
// latency is paid once per batch.  (Segments are in Hilbert order, so a
// wave's batch is mostly all in or all out of range.)
template <int IPL, int U, bool LOG, bool DIAG>
__device__ __forceinline__ void pair_batch(const double (&xi)[IPL], const double (&yi)[IPL],
                                           const double (&zi)[IPL], const int (&jmin)[IPL], int j,
                                           const double (&X)[U], const double (&Y)[U], const double (&Z)[U],
                                           int n, float mul, float add, const double *D, double Tn, uint32_t *hk) {
    constexpr int M = IPL * U;
    const float top = (float)(n + 1);
    double s[M];
    int k[M];
    bool near = false;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < IPL; ++q) {
            const double dx = xi[q] - X[u], dy = yi[q] - Y[u], dz = zi[q] - Z[u];
            s[u * IPL + q] = (dx * dx + dy * dy) + dz * dz;
            near = near || s[u * IPL + q] < Tn;
        }
    if (__ballot(near) == 0ull) return;
#pragma unroll
    for (int m = 0; m < M; ++m) k[m] = bin_guess<LOG>((float)s[m], mul, add, top);
    double lo[M], hi[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        lo[m] = D[k[m]];
        hi[m] = D[k[m] + 1];
    }
#pragma unroll
    for (int m = 0; m < M; ++m)
        if (!((s[m] >= lo[m]) & (s[m] < hi[m]))) k[m] = bin_walk(s[m], k[m], D, n);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < IPL; ++q) {
            const int kk = k[u * IPL + q];
            if ((unsigned)(kk - 1) < (unsigned)n && (!DIAG || j + u > jmin[q]))
                __hip_atomic_fetch_add(hk + kk, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
}

constexpr int kJBatch = 4;
// Coordinate of lanes past the end of segment a: every squared distance is
// then about 3e300, finite and above any threshold (axes end below 1e100).
constexpr double kFarPoint = 1e150;

// One item: the lane's IPL points of segment a against j in [j0, j1) of
// segment b.  DIAG: a and b are one segment and the item holds pairs with
// j <= i, which are skipped.
template <int WAVES, int IPL, bool LOG, bool DIAG, bool CULL>
__device__ __forceinline__ void pair_item(CPairTerm *T, int i0, int j0, int j1, const double *D, uint32_t *hk) {
    const int na = T->na, n = T->n;
    const gdouble *ax = (const gdouble *)T->ax, *ay = (const gdouble *)T->ay, *az = (const gdouble *)T->az;
    double xi[IPL], yi[IPL], zi[IPL];
    int jmin[IPL];  // DIAG: the lane counts the pairs with j > jmin
    // The lane's points are consecutive, so a wave's 64 * IPL points are one
    // run of the segment (a compact box in Hilbert order).
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
        const int i = i0 + (int)threadIdx.x * IPL + q;
        if (i < na) {
            xi[q] = ax[i];
            yi[q] = ay[i];
            zi[q] = az[i];
            jmin[q] = i;
        } else {
            xi[q] = yi[q] = zi[q] = kFarPoint;
            jmin[q] = INT_MAX;
        }
    }
    const double *bx = T->bx, *by = T->by, *bz = T->bz;
    const float mul = T->h_mul, add = T->h_add;
    const double Tn = D[n + 1];  // the last squared threshold
    // The pairs of the lane's points with j in [ja, jb).
    auto run = [&](int ja, int jb) {
        int j = ja;
        // kJBatch j points per batch of scalar loads (merged into wide loads).
        for (; j + kJBatch <= jb; j += kJBatch) {
            double X[kJBatch], Y[kJBatch], Z[kJBatch];
#pragma unroll
            for (int u = 0; u < kJBatch; ++u) {
                X[u] = ld_uniform(bx, j + u);
                Y[u] = ld_uniform(by, j + u);
                Z[u] = ld_uniform(bz, j + u);
            }
            pair_batch<IPL, kJBatch, LOG, DIAG>(xi, yi, zi, jmin, j, X, Y, Z, n, mul, add, D, Tn, hk);
        }
        for (; j < jb; ++j) {
            const double X[1] = {ld_uniform(bx, j)}, Y[1] = {ld_uniform(by, j)}, Z[1] = {ld_uniform(bz, j)};
            pair_batch<IPL, 1, LOG, DIAG>(xi, yi, zi, jmin, j, X, Y, Z, n, mul, add, D, Tn, hk);
        }
    };
    if (!CULL) {
        run(j0, j1);
        return;
    }
    const double *jbox = T->jbox;
    // Culled form: the box of the wave's points (uniform after the
    // reductions), then only the j boxes it can reach.
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
#pragma unroll
    for (int q = 0; q < IPL; ++q)
        if (jmin[q] != INT_MAX) {
            lo[0] = fmin(lo[0], xi[q]);
            lo[1] = fmin(lo[1], yi[q]);
            lo[2] = fmin(lo[2], zi[q]);
            hi[0] = fmax(hi[0], xi[q]);
            hi[1] = fmax(hi[1], yi[q]);
            hi[2] = fmax(hi[2], zi[q]);
        }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], m));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], m));
        }
    if (!(lo[0] <= hi[0])) return;  // no point of segment a in this wave
    for (int j = j0; j < j1;) {
        const int c = j / kPairBoxPoints;
        const int je = min(j1, (c + 1) * kPairBoxPoints);
        const double *b = jbox + 6 * c;
        double g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            g[a] = fmax(0.0, fmax(ld_uniform(b, a) - hi[a], lo[a] - ld_uniform(b, 3 + a)));
        // The kernel's squared distance of every pair is at least this
        // (monotone FP64 rounding of the same operations).
        if (__builtin_amdgcn_readfirstlane((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2] < Tn ? 1 : 0)) run(j, je);
        j = je;
    }
}

// Sum the copies of bins [0, n) into the term histogram and clear them.
__device__ void pair_flush(uint32_t *H, int copies, int stride, int n, unsigned long long *hist) {
    __syncthreads();
    for (int b = threadIdx.x; b < n; b += blockDim.x) {
        unsigned long long v = 0;
        for (int c = 0; c < copies; ++c) {
            v += H[c * stride + b];
            H[c * stride + b] = 0u;
        }
        if (v) atomicAdd(hist + b, v);
    }
    __syncthreads();
}

template <int WAVES, int IPL, bool CULL>
__global__ __launch_bounds__(WAVES * 64) void pair_hist_kernel(PairHistArgs P) {
    extern __shared__ __align__(16) uint32_t lds[];
    __shared__ int s_item;
    uint32_t *H = lds;
    double *D = reinterpret_cast<double *>(lds + P.hist_words);
    const int lane = threadIdx.x & 63;
    for (int w = threadIdx.x; w < P.hist_words; w += blockDim.x) H[w] = 0u;
    // hk[k] is bin k-1 of the lane's copy.
    uint32_t *hk = H + (lane % P.copies) * P.stride - 1;
    CPairTerm *terms = (CPairTerm *)P.terms;
    CPairItem *items = (CPairItem *)P.items;
    int cur = -1, cur_n = 0;
    unsigned long long *cur_hist = nullptr;
    unsigned long long since_flush = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(P.counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= P.nitems) break;
        const int term = items[it].term, i0 = items[it].i0, j0 = items[it].j0, j1 = items[it].j1;
        CPairTerm *T = terms + term;
        // A copy gains at most one count per pair the workgroup classifies.
        const unsigned long long add = (unsigned long long)(WAVES * 64 * IPL) * (unsigned long long)(j1 - j0);
        if (term != cur || since_flush + add > 0xFFFFFFFFull) {
            if (cur >= 0) pair_flush(H, P.copies, P.stride, cur_n, cur_hist);
            if (term != cur) {
                const int nd = T->n + 3;
                const gdouble *src = (const gdouble *)T->D;
                for (int w = threadIdx.x; w < nd; w += blockDim.x) D[w] = src[w];
                __syncthreads();
            }
            cur = term;
            cur_n = T->n;
            cur_hist = T->hist;
            since_flush = 0;
        }
        since_flush += add;
        // Pairs with j <= i exist in the item only if j0 <= the last i.
        const bool diag = T->same && j0 < i0 + WAVES * 64 * IPL;
        if (T->log) {
            if (diag) pair_item<WAVES, IPL, true, true, CULL>(T, i0, j0, j1, D, hk);
            else pair_item<WAVES, IPL, true, false, CULL>(T, i0, j0, j1, D, hk);
        } else {
            if (diag) pair_item<WAVES, IPL, false, true, CULL>(T, i0, j0, j1, D, hk);
            else pair_item<WAVES, IPL, false, false, CULL>(T, i0, j0, j1, D, hk);
        }
    }
    if (cur >= 0) pair_flush(H, P.copies, P.stride, cur_n, cur_hist);
}

// FP64 bin of squared distance s by binary search of T[0..n] (global).
__device__ __forceinline__ int plain_bin(double s, const double *T, int n) {
    if (!(s >= T[0]) || !(s < T[n])) return -1;
    int lo = 0, hi = n;  // T[lo] <= s < T[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s >= T[mid]) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void pair_plain_kernel(PairHistArgs P) {
    extern __shared__ __align__(16) uint32_t lds[];
    const PairHistItem I = P.items[blockIdx.x];
    const PairHistTerm &T = P.terms[I.term];
    const int n = T.n;
    for (int b = threadIdx.x; b < n; b += blockDim.x) lds[b] = 0u;
    __syncthreads();
    const double *Tq = T.D + 1;
    const int i = I.i0 + (int)threadIdx.x;
    if (i < T.na) {
        const double x = T.ax[i], y = T.ay[i], z = T.az[i];
        for (int j = (T.same ? max(I.j0, i + 1) : I.j0); j < I.j1; ++j) {
            const double dx = x - T.bx[j], dy = y - T.by[j], dz = z - T.bz[j];
            const double s = (dx * dx + dy * dy) + dz * dz;
            const int b = plain_bin(s, Tq, n);
            if (b >= 0) atomicAdd(lds + b, 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < n; b += blockDim.x)
        if (lds[b]) atomicAdd(T.hist + b, (unsigned long long)lds[b]);
}

}  // namespace

size_t pair_hist_lds_bytes(int hist_words, int max_n) {
    return (size_t)hist_words * 4 + (size_t)(max_n + 3) * 8;
}

const void *pair_hist_kernel_ptr(int waves, bool cull) {
    if (waves == 8)
        return cull ? reinterpret_cast<const void *>(&pair_hist_kernel<8, kPairIpl, true>)
                    : reinterpret_cast<const void *>(&pair_hist_kernel<8, kPairIpl, false>);
    return nullptr;
}

const void *pair_plain_kernel_ptr() { return reinterpret_cast<const void *>(&pair_plain_kernel); }

hipError_t launch_pair_hist(const PairHistArgs &P, int waves, bool cull, int grid, hipStream_t st) {
    const size_t lds = pair_hist_lds_bytes(P.hist_words, P.max_n);
    if (waves != 8) return hipErrorInvalidValue;
    if (cull) hipLaunchKernelGGL((pair_hist_kernel<8, kPairIpl, true>), dim3(grid), dim3(8 * 64), lds, st, P);
    else hipLaunchKernelGGL((pair_hist_kernel<8, kPairIpl, false>), dim3(grid), dim3(8 * 64), lds, st, P);
    return hipGetLastError();
}

hipError_t launch_pair_plain(const PairHistArgs &P, hipStream_t st) {
    hipLaunchKernelGGL(pair_plain_kernel, dim3(P.nitems), dim3(256), (size_t)P.max_n * 4, st, P);
    return hipGetLastError();
}

}  // namespace tpcf
