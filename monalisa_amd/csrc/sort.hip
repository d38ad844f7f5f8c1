// Stable LSD radix sort of (key, value) int32 pairs, 4 bits per pass.
//
// Used to turn point-major (point, neighbour) entries into cell-major
// segments while keeping the entries of each cell in ascending point order
// (the order in which the reference's serial gridders accumulate).
//
// Per pass: (1) each tile of kTile entries counts its 16 digits; (2) the
// digit-major table of tile counts is scanned; (3) each tile ranks its
// entries stably (per-work-item digit counts + a workgroup scan per digit)
// and scatters them.  Integer counts only: the result is exact and
// independent of scheduling.
#include <algorithm>

#include "mona_internal.hpp"

namespace mona {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigits = 16;

__global__ void __launch_bounds__(kSortThreads)
k_tile_hist(const unsigned* __restrict__ keys, int64_t n, int shift, int ntile, int* __restrict__ hist) {
    __shared__ int h[kDigits];
    if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)threadIdx.x * kSortItems;
    for (int i = 0; i < kSortItems; i++)
        if (base + i < n) atomicAdd(&h[(keys[base + i] >> shift) & (kDigits - 1)], 1);
    __syncthreads();
    if (threadIdx.x < kDigits) hist[(int64_t)threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kSortThreads)
k_tile_scatter(const unsigned* __restrict__ keys, const int* __restrict__ vals, int64_t n, int shift, int ntile,
               const int64_t* __restrict__ offs, unsigned* __restrict__ okeys, int* __restrict__ ovals) {
    __shared__ int cnt[kDigits][kSortThreads];
    __shared__ int tot[kDigits];
    const int t = threadIdx.x;
    int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)t * kSortItems;
    unsigned k[kSortItems];
    int v[kSortItems], d[kSortItems];
    int mine[kDigits];
#pragma unroll
    for (int j = 0; j < kDigits; j++) mine[j] = 0;
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        if (base + i < n) {
            k[i] = keys[base + i];
            v[i] = vals[base + i];
            d[i] = (k[i] >> shift) & (kDigits - 1);
        } else {
            d[i] = -1;
        }
    }
#pragma unroll
    for (int i = 0; i < kSortItems; i++)
#pragma unroll
        for (int j = 0; j < kDigits; j++) mine[j] += (d[i] == j) ? 1 : 0;
#pragma unroll
    for (int j = 0; j < kDigits; j++) cnt[j][t] = mine[j];
    __syncthreads();
    // inclusive scan over work-items, one digit row per pass of the loop
    for (int off = 1; off < kSortThreads; off <<= 1) {
        int add[kDigits];
#pragma unroll
        for (int j = 0; j < kDigits; j++) add[j] = (t >= off) ? cnt[j][t - off] : 0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kDigits; j++) cnt[j][t] += add[j];
        __syncthreads();
    }
    if (t < kDigits) tot[t] = cnt[t][kSortThreads - 1];
    __syncthreads();
    int run[kDigits];
#pragma unroll
    for (int j = 0; j < kDigits; j++) run[j] = cnt[j][t] - mine[j];  // exclusive prefix of this work-item
#pragma unroll
    for (int i = 0; i < kSortItems; i++) {
        if (d[i] >= 0) {
            int dd = d[i];
            int r = 0;
#pragma unroll
            for (int j = 0; j < kDigits; j++)
                if (j == dd) { r = run[j]; run[j] = r + 1; }
            int64_t dst = offs[(int64_t)dd * ntile + blockIdx.x] + r;
            okeys[dst] = k[i];
            ovals[dst] = v[i];
        }
    }
}

}  // namespace

// Sorts n pairs stably by the low `bits` bits of the keys.  keys/vals are
// overwritten with the sorted sequence; tmp buffers are allocated here.
void radix_sort_pairs(unsigned* keys, int* vals, int64_t n, int bits, hipStream_t s) {
    if (n <= 1) return;
    MONA_CHECK(n < ((int64_t)1 << 31), MONA_ERR_ARG, "radix sort: too many entries");
    int ntile = (int)((n + kSortTile - 1) / kSortTile);
    DevBuf k2((size_t)n * sizeof(unsigned)), v2((size_t)n * sizeof(int));
    DevBuf hist((size_t)kDigits * ntile * sizeof(int)), offs(((size_t)kDigits * ntile + 1) * sizeof(int64_t));
    unsigned *ka = keys, *kb = k2.as<unsigned>();
    int *va = vals, *vb = v2.as<int>();
    int npass = (bits + 3) / 4;
    for (int p = 0; p < npass; p++) {
        int shift = 4 * p;
        k_tile_hist<<<ntile, kSortThreads, 0, s>>>(ka, n, shift, ntile, hist.as<int>());
        MONA_HIP(hipGetLastError());
        scan_i32_to_i64_exclusive(hist.as<int>(), offs.as<int64_t>(), (int64_t)kDigits * ntile, s);
        k_tile_scatter<<<ntile, kSortThreads, 0, s>>>(ka, va, n, shift, ntile, offs.as<int64_t>(), kb, vb);
        MONA_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        MONA_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToDevice, s));
        MONA_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    MONA_HIP(hipStreamSynchronize(s));
}

}  // namespace mona
