// kernels_cells.hip -- spatial ordering of a catalog segment: Hilbert keys of
// the coordinates, a device radix sort (rocPRIM), a gather into a sorted SoA
// copy, and the bounding box of every chunk of consecutive sorted points.
//
// Reference idea: examples/test_predivision_of_distances.py:11-227 divides
// the volume into cells no smaller than the maximum separation and pairs each
// cell only with its neighbours.  Here the points are put in Hilbert order, so
// that runs of consecutive points are compact, and the host skips every pair
// of runs whose boxes are farther apart than the last bin edge.  Counts do not
// depend on the order of the points, so the histograms are those of the
// unsorted segment.
#include <rocprim/device/device_radix_sort.hpp>

#include "tpcf_internal.h"

namespace tpcf {
namespace {

constexpr int kKeyBits = 21;

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint32_t quantize(double c, double lo, double inv) {
    const double q = (c - lo) * inv;
    const double top = (double)((1u << kKeyBits) - 1);
    return (uint32_t)(q < 0.0 ? 0.0 : (q > top ? top : q));
}

// Hilbert index of three 21-bit coordinates: Skilling's transform of the axes
// to the "transposed" Hilbert index (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 381 (2004)), whose bits interleaved (axis 0
// most significant at each level) are the key.  Consecutive keys are cells
// one step apart, unlike Morton order, which jumps at every octant boundary:
// along a tile row of Hilbert-ordered points the pair bins change less often,
// so the (r1,r2,r3) kernel's runs of equal cells are longer (the 20k + 20k
// headline geometry: 1995 -> 1895 ms).
__device__ __forceinline__ uint64_t hilbert3(uint32_t x0, uint32_t x1, uint32_t x2) {
    uint32_t X[3] = {x0, x1, x2};
    const uint32_t M = 1u << (kKeyBits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;  // invert the low bits of axis 0
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;  // exchange the low bits of axes 0 and i
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    // Gray encode
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    return spread3(X[2]) | (spread3(X[1]) << 1) | (spread3(X[0]) << 2);
}

__global__ void hilbert_keys_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                    const double *__restrict__ z, int n, double lox, double loy, double loz,
                                    double inv, uint64_t *__restrict__ keys, uint32_t *__restrict__ idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = hilbert3(quantize(x[i], lox, inv), quantize(y[i], loy, inv), quantize(z[i], loz, inv));
        idx[i] = (uint32_t)i;
    }
}

__global__ void gather_kernel(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const uint32_t *__restrict__ idx, int n,
                              int64_t stride, double *__restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = idx[i];
        out[i] = x[s];
        out[stride + i] = y[s];
        out[2 * stride + i] = z[s];
    }
}

// Box of chunk c (points [c*chunk, min(n, (c+1)*chunk))): lo x, y, z then
// hi x, y, z.
__global__ void chunk_bounds_kernel(const double *__restrict__ p, int n, int64_t stride, int chunk, int nchunks,
                                    double *__restrict__ box) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        const int b = c * chunk, e = min(n, b + chunk);
        double lo[3] = {p[b], p[stride + b], p[2 * stride + b]}, hi[3] = {lo[0], lo[1], lo[2]};
        for (int i = b + 1; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                const double v = p[a * stride + i];
                lo[a] = fmin(lo[a], v);
                hi[a] = fmax(hi[a], v);
            }
        for (int a = 0; a < 3; ++a) {
            box[6 * c + a] = lo[a];
            box[6 * c + 3 + a] = hi[a];
        }
    }
}

int grid_for(int n) { return (int)std::min<int64_t>(4096, ((int64_t)n + 255) / 256); }

}  // namespace

size_t sort_segment_scratch_bytes(int n) {
    size_t temp = 0;
    (void)rocprim::radix_sort_pairs(nullptr, temp, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
                                    (uint32_t *)nullptr, (size_t)n, 0, 3 * kKeyBits);
    const auto a = [](size_t v) { return (v + 255) & ~size_t(255); };
    return a(temp) + 2 * a((size_t)n * 8) + 2 * a((size_t)n * 4);
}

hipError_t sort_segment(const double *x, const double *y, const double *z, int n, const double lo[3], double span,
                        void *scratch, size_t scratch_bytes, double *sorted, int64_t stride, int chunk, double *box,
                        hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const auto a = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t temp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                             (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 3 * kKeyBits);
    if (e != hipSuccess) return e;
    char *s = static_cast<char *>(scratch);
    void *tmp = s;
    uint64_t *k0 = reinterpret_cast<uint64_t *>(s + a(temp));
    uint64_t *k1 = reinterpret_cast<uint64_t *>(s + a(temp) + a((size_t)n * 8));
    uint32_t *v0 = reinterpret_cast<uint32_t *>(s + a(temp) + 2 * a((size_t)n * 8));
    uint32_t *v1 = reinterpret_cast<uint32_t *>(s + a(temp) + 2 * a((size_t)n * 8) + a((size_t)n * 4));
    if (a(temp) + 2 * a((size_t)n * 8) + 2 * a((size_t)n * 4) > scratch_bytes) return hipErrorInvalidValue;
    const double inv = span > 0 ? (double)(1u << kKeyBits) / span : 0.0;
    hipLaunchKernelGGL(hilbert_keys_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, y, z, n, lo[0], lo[1], lo[2], inv,
                       k0, v0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = rocprim::radix_sort_pairs(tmp, temp, k0, k1, v0, v1, (size_t)n, 0, 3 * kKeyBits, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, y, z, v1, n, stride, sorted);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int nchunks = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(chunk_bounds_kernel, dim3(grid_for(nchunks)), dim3(256), 0, st, sorted, n, stride, chunk,
                       nchunks, box);
    return hipGetLastError();
}

}  // namespace tpcf
